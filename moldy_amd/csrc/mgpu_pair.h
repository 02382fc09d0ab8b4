// The site-site pair term of the reference's kernel() (src/kernel.c:188-460)
// for every potential form, templated on the arithmetic type: double and
// float for the force kernels, Dual<N> for the parameter adjoint (forward-mode
// derivatives of u and forceij with respect to the pair's parameters and the
// charge product, carried alongside the value).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace mgpu {

constexpr double AS_P = 0.3275911;   // Abramowitz and Stegun 7.1.26, src/kernel.c:88-96
constexpr double AS_A1 = 0.254829592, AS_A2 = -0.284496736, AS_A3 = 1.421413741,
                 AS_A4 = -1.453152027, AS_A5 = 1.061405429;

// Value with N first derivatives.
template <int N> struct Dual {
   double v;
   double d[N];
   __host__ __device__ Dual() : v(0) {
      for (int k = 0; k < N; k++) d[k] = 0;
   }
   __host__ __device__ Dual(double x) : v(x) {
      for (int k = 0; k < N; k++) d[k] = 0;
   }
   __host__ __device__ static Dual seed(double x, int k) {
      Dual r(x);
      if (k >= 0 && k < N) r.d[k] = 1.0;
      return r;
   }
   __host__ __device__ Dual &operator+=(const Dual &b) {
      v += b.v;
      for (int k = 0; k < N; k++) d[k] += b.d[k];
      return *this;
   }
   __host__ __device__ Dual &operator*=(const Dual &b) {
      for (int k = 0; k < N; k++) d[k] = d[k] * b.v + v * b.d[k];
      v *= b.v;
      return *this;
   }
};
template <int N> __host__ __device__ inline Dual<N> operator+(Dual<N> a, const Dual<N> &b) { return a += b; }
template <int N> __host__ __device__ inline Dual<N> operator-(Dual<N> a, const Dual<N> &b) {
   a.v -= b.v;
   for (int k = 0; k < N; k++) a.d[k] -= b.d[k];
   return a;
}
template <int N> __host__ __device__ inline Dual<N> operator-(Dual<N> a) {
   a.v = -a.v;
   for (int k = 0; k < N; k++) a.d[k] = -a.d[k];
   return a;
}
template <int N> __host__ __device__ inline Dual<N> operator*(Dual<N> a, const Dual<N> &b) { return a *= b; }
template <int N> __host__ __device__ inline Dual<N> operator+(Dual<N> a, double b) { a.v += b; return a; }
template <int N> __host__ __device__ inline Dual<N> operator+(double b, Dual<N> a) { a.v += b; return a; }
template <int N> __host__ __device__ inline Dual<N> operator-(Dual<N> a, double b) { a.v -= b; return a; }
template <int N> __host__ __device__ inline Dual<N> operator-(double b, Dual<N> a) { return Dual<N>(b) - a; }
template <int N> __host__ __device__ inline Dual<N> operator*(Dual<N> a, double b) {
   a.v *= b;
   for (int k = 0; k < N; k++) a.d[k] *= b;
   return a;
}
template <int N> __host__ __device__ inline Dual<N> operator*(double b, Dual<N> a) { return a * b; }

template <class R> struct Math;
// root(rsq, r, r_r): r = sqrt(rsq) and r_r = 1/r.  FP64 and Dual take the
// correctly rounded square root and divide (the reference's arithmetic);
// FP32 takes the hardware reciprocal square root (v_rsq_f32, about 1 ulp)
// and r = rsq / sqrt(rsq) from it.
template <> struct Math<double> {
   static __host__ __device__ inline double sqrt_(double x) { return sqrt(x); }
   static __host__ __device__ inline double exp_(double x) { return exp(x); }
   static __host__ __device__ inline double rcp(double x) { return 1.0 / x; }
   static __host__ __device__ inline void root(double x, double &r, double &r_r) {
      r = sqrt(x);
      r_r = 1.0 / r;
   }
};
template <> struct Math<float> {
   static __device__ __forceinline__ float sqrt_(float x) { return __builtin_sqrtf(x); }
   static __device__ __forceinline__ float exp_(float x) { return __expf(x); }
   static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
   static __device__ __forceinline__ void root(float x, float &r, float &r_r) {
      r_r = __builtin_amdgcn_rsqf(x);
      r = x * r_r;
   }
};
template <int N> struct Math<Dual<N>> {
   static __host__ __device__ inline Dual<N> sqrt_(Dual<N> x) {
      double s = sqrt(x.v), g = 0.5 / s;
      x.v = s;
      for (int k = 0; k < N; k++) x.d[k] *= g;
      return x;
   }
   static __host__ __device__ inline Dual<N> exp_(Dual<N> x) {
      double e = exp(x.v);
      x.v = e;
      for (int k = 0; k < N; k++) x.d[k] *= e;
      return x;
   }
   static __host__ __device__ inline Dual<N> rcp(Dual<N> x) {
      double r = 1.0 / x.v, g = -r * r;
      x.v = r;
      for (int k = 0; k < N; k++) x.d[k] *= g;
      return x;
   }
   static __host__ __device__ inline void root(Dual<N> x, Dual<N> &r, Dual<N> &r_r) {
      r = sqrt_(x);
      r_r = rcp(r);
   }
};

// One pair: energy u and forceij = -U'(r)/r, restating kernel.c:188-460.
template <class R, int PT, bool COUL>
__host__ __device__ __forceinline__ void pair_term(R rsq, R qq, R norm, R alpha, const R *__restrict__ p,
                                          R &u, R &f) {
   using M = Math<R>;
   R r, r_r, r2r, t = 0, et = 0;
   if (COUL) {
      M::root(rsq, r, r_r);
      R ar = alpha * r;
      t = M::rcp(R(1) + R(AS_P) * ar);
      et = qq * M::exp_(-(ar * ar));
      R poly = t * (R(AS_A1) + t * (R(AS_A2) + t * (R(AS_A3) + t * (R(AS_A4) + t * R(AS_A5)))));
      t = poly * et * r_r;
      et = t + norm * et;
      r2r = r_r * r_r;
   } else if (PT == 0) {
      r2r = M::rcp(rsq);
      r = 0;
      r_r = 0;
   } else {
      M::root(rsq, r, r_r);
      r2r = r_r * r_r;
   }
   if (PT == 0) {  // Lennard-Jones
      R r6 = p[1] * p[1] * r2r;
      r6 = r6 * r6 * r6;
      R r12 = r6 * r6;
      u = t + p[0] * (r12 - r6);
      f = r2r * (R(6) * p[0] * (R(2) * r12 - r6) + et);
   } else if (PT == 1) {  // Buckingham / 6-exp
      R e1 = p[1] * M::exp_(-p[2] * r);
      R r6 = p[0] * r2r * r2r * r2r;
      u = t - r6 + e1;
      f = r2r * (R(-6) * r6 + et) + p[2] * e1 * r_r;
   } else if (PT == 2) {  // MCY
      R e1 = p[0] * M::exp_(-p[1] * r);
      R e2 = -p[2] * M::exp_(-p[3] * r);
      u = t + e1 + e2;
      f = (p[1] * e1 + p[3] * e2) * r_r + et * r2r;
   } else if (PT == 3) {  // generic
      R e1 = p[0] * M::exp_(-p[1] * r);
      R r4 = r2r * r2r, r6 = r2r * r4, r8 = p[5] * r4 * r4, r12 = p[2] * r6 * r6;
      r4 *= p[3];
      r6 *= p[4];
      u = t + e1 + r12 - r4 - r6 - r8;
      f = r2r * (R(12) * r12 - R(4) * r4 - R(6) * r6 - R(8) * r8 + et) + p[1] * e1 * r_r;
   } else if (PT == 4) {  // HIW
      R r4 = r2r * r2r, r6 = r2r * r4, r12 = p[2] * r6 * r6;
      r6 *= p[1];
      r4 *= p[0];
      u = t + r4 + r6 + r12;
      f = r2r * (R(4) * r4 + R(6) * r6 + R(12) * r12 + et);
   } else {  // PT == 6, Morse + BIG
      R e1 = p[0] * M::exp_((p[1] - r) * p[2]);
      R r6 = p[3] * r2r * r2r * r2r;
      R e2 = p[4] * M::exp_(R(-2) * p[5] * (r - p[6]));
      R e3 = -p[4] * R(2) * M::exp_(-p[5] * (r - p[6]));
      u = t + e1 - r6 + e2 + e3;
      f = r2r * (R(-6) * r6 + et) + r_r * (p[2] * e1 + (R(2) * p[5]) * e2 + p[5] * e3);
   }
}

}  // namespace mgpu
