// Reciprocal-space Ewald sum, FP64 path (restates src/ewald.c:280-589).
//
// Per call: per-site trig tables by the reference's recurrence (trig_recur,
// ewald.c:148-193, z table premultiplied by q as at ewald.c:475-481), the
// structure factors S(k) for the reference's half-space k set, the energy
// and stress per k (ewald.c:539-553), and the force back-projection
// (ewald.c:557-579).  Framework sites (the last ns - nsitesxf sites) enter
// S(k) separately so that frame-frame terms are excluded as in the reference.
//
// S(k): a workgroup owns a group of consecutive k vectors (at most 256, from
// at most KROWS (h,k) rows) and a chunk of sites; per tile of sites it stages
// A = q-free exp(i(h a*.r + k b*.r)) for its rows and the q-weighted
// exp(i l c*.r) table in LDS, and each thread accumulates one k vector.
// Chunk partials are reduced in a fixed order (deterministic).
//
// Forces: one thread per site walks the rows; exp(i l c*.r) is advanced along
// a row by multiplication with exp(i c*.r), so the inner loop touches no
// memory except the per-k weights (uniform loads).
#include "mgpu_internal.h"

#include <algorithm>

using namespace mgpu;

namespace {

constexpr int KROWS = 32;   // rows per S(k) workgroup
constexpr int TS = 32;      // sites per LDS tile
constexpr int CHUNK = 8192; // sites per S(k) partial
constexpr double PI_D = 3.14159265358979323846;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
   return make_double2(a.x * b.x - a.y * b.y, a.y * b.x + a.x * b.y);
}

struct TrigArgs {
   double ks[9];     // rows a*, b*, c* (2 pi h^-1)
   int hmax, kmax, lmax, ns, nsp;
};

// trig_recur for the three axes; z tables multiplied by q.
__global__ void k_trig(TrigArgs a, const double *__restrict__ site, const double *__restrict__ chg,
                       double2 *__restrict__ tx, double2 *__restrict__ ty, double2 *__restrict__ tz,
                       double2 *__restrict__ tz1) {
   int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= a.ns) return;
   double x = site[i], y = site[a.nsp + i], z = site[2 * a.nsp + i], q = chg[i];
   const int nmax[3] = {a.hmax, a.kmax, a.lmax};
   double2 *tb[3] = {tx, ty, tz};
#pragma unroll
   for (int ax = 0; ax < 3; ax++) {
      double2 *t = tb[ax];
      double w = ax == 2 ? q : 1.0;
      t[i] = make_double2(w, 0.0);
      if (nmax[ax] >= 1) {
         double kr = a.ks[3 * ax] * x + a.ks[3 * ax + 1] * y + a.ks[3 * ax + 2] * z;
         double s1, c1;
         sincos(kr, &s1, &c1);
         if (ax == 2) tz1[i] = make_double2(c1, s1);
         double cm = c1, sm = s1;
         t[(size_t)a.nsp + i] = make_double2(c1 * w, s1 * w);
         for (int n = 2; n <= nmax[ax]; n++) {
            double cc = cm * c1 - sm * s1;
            double ss = sm * c1 + cm * s1;
            cm = cc;
            sm = ss;
            t[(size_t)n * a.nsp + i] = make_double2(cc * w, ss * w);
         }
      } else if (ax == 2) {
         double s1, c1;
         double kr = a.ks[6] * x + a.ks[7] * y + a.ks[8] * z;
         sincos(kr, &s1, &c1);
         tz1[i] = make_double2(c1, s1);
      }
   }
}

struct SkGroup {
   int r0, nr, k0, nkg;
};

struct SkArgs {
   int nk, ns, nsp, nsxf, lmax, nchunks;
   const KRow *rows;
   const SkGroup *groups;
   const int *krow;
   const double2 *tx, *ty, *tz;
   double *part;    // [nchunks][nk][4]
};

template <bool FW>
__global__ __launch_bounds__(256) void k_sfac(SkArgs a) {
   __shared__ double2 sA[KROWS][TS + 1];
   extern __shared__ double2 sZ[];   // [(lmax+1)][TS+1]
   const SkGroup g = a.groups[blockIdx.x];
   const int chunk = blockIdx.y;
   const int i0 = chunk * CHUNK, i1 = min(a.ns, i0 + CHUNK);
   const int tid = threadIdx.x;
   const int kk = g.k0 + tid;
   const bool active = tid < g.nkg;
   int rloc = 0, l = 0;
   if (active) {
      int r = a.krow[kk];
      rloc = r - g.r0;
      KRow R = a.rows[r];
      l = R.l0 + (kk - R.kstart);
   }
   const int al = l < 0 ? -l : l;
   const double zsgn = l < 0 ? -1.0 : 1.0;
   double sc = 0, ss = 0, scf = 0, ssf = 0;
   for (int t0 = i0; t0 < i1; t0 += TS) {
      const int nt = min(TS, i1 - t0);
      __syncthreads();
      for (int q = tid; q < (a.lmax + 1) * TS; q += 256) {
         int ll = q / TS, t = q - ll * TS;
         sZ[ll * (TS + 1) + t] = t < nt ? a.tz[(size_t)ll * a.nsp + t0 + t] : make_double2(0, 0);
      }
      for (int q = tid; q < g.nr * TS; q += 256) {
         int rr = q / TS, t = q - rr * TS;
         double2 v = make_double2(0, 0);
         if (t < nt) {
            KRow R = a.rows[g.r0 + rr];
            double2 xh = a.tx[(size_t)R.h * a.nsp + t0 + t];
            int ak = R.k < 0 ? -R.k : R.k;
            double2 yk = a.ty[(size_t)ak * a.nsp + t0 + t];
            // qsincos (ewald.c:108-121): k >= 0 uses exp(+i k b*.r), k < 0 the conjugate
            if (R.k < 0) yk.y = -yk.y;
            v = make_double2(xh.x * yk.x - xh.y * yk.y, xh.y * yk.x + xh.x * yk.y);
         }
         sA[rr][t] = v;
      }
      __syncthreads();
      if (active) {
         if (!FW || t0 + nt <= a.nsxf) {
            for (int t = 0; t < nt; t++) {
               double2 A = sA[rloc][t];
               double2 Z = sZ[al * (TS + 1) + t];
               double zi = zsgn * Z.y;
               sc += A.x * Z.x - A.y * zi;
               ss += A.y * Z.x + A.x * zi;
            }
         } else {
            for (int t = 0; t < nt; t++) {
               double2 A = sA[rloc][t];
               double2 Z = sZ[al * (TS + 1) + t];
               double zi = zsgn * Z.y;
               double c = A.x * Z.x - A.y * zi, s = A.y * Z.x + A.x * zi;
               if (t0 + t < a.nsxf) { sc += c; ss += s; }
               else { scf += c; ssf += s; }
            }
         }
      }
   }
   if (active) {
      double *p = a.part + ((size_t)chunk * a.nk + kk) * 4;
      p[0] = sc; p[1] = ss; p[2] = scf; p[3] = ssf;
   }
}

struct EkArgs {
   int nk, nchunks, nrows;
   double vol, r4a, c2;
   const KRow *rows;
   const int *krow;
   const double *part;
   double *w;       // [nk][4]: Wc, Ws (all sites), Wcn, Wsn (framework sites)
   double *bpart;   // [nblocks][8]
};

// Per k: sum the chunk partials, energy, stress, weights (ewald.c:506-553).
__global__ __launch_bounds__(256) void k_ewald_k(EkArgs a) {
   __shared__ double ws[4][8];
   int k = blockIdx.x * 256 + threadIdx.x;
   double pe = 0, s00 = 0, s01 = 0, s02 = 0, s11 = 0, s12 = 0, s22 = 0;
   if (k < a.nk) {
      double scn = 0, ssn = 0, scf = 0, ssf = 0;
      for (int c = 0; c < a.nchunks; c++) {
         const double *p = a.part + ((size_t)c * a.nk + k) * 4;
         scn += p[0]; ssn += p[1]; scf += p[2]; ssf += p[3];
      }
      KRow R = a.rows[a.krow[k]];
      int l = R.l0 + (k - R.kstart);
      double kx = R.kx, ky = R.ky, kz = R.kzt + l * a.c2;
      double ksq = kx * kx + ky * ky + kz * kz;
      double coeff = 2.0 / ((0.25 / PI_D) * a.vol) * exp(ksq * a.r4a) / ksq;
      double coeff2 = 2.0 * (1.0 - ksq * a.r4a) / ksq;
      double pe_k = 0.5 * coeff * (scn * (scn + scf + scf) + ssn * (ssn + ssf + ssf));
      pe = pe_k;
      s00 = pe_k - pe_k * coeff2 * kx * kx;
      s01 = -pe_k * coeff2 * kx * ky;
      s02 = -pe_k * coeff2 * kx * kz;
      s11 = pe_k - pe_k * coeff2 * ky * ky;
      s12 = -pe_k * coeff2 * ky * kz;
      s22 = pe_k - pe_k * coeff2 * kz * kz;
      double *w = a.w + 4 * (size_t)k;
      w[0] = coeff * (scn + scf);
      w[1] = coeff * (ssn + ssf);
      w[2] = coeff * scn;
      w[3] = coeff * ssn;
   }
   double v[7] = {pe, s00, s01, s02, s11, s12, s22};
   int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
   for (int q = 0; q < 7; q++) {
      double t = wave_sum(v[q]);
      if (lane == 0) ws[wv][q] = t;
   }
   __syncthreads();
   if (threadIdx.x < 7) {
      double t = ws[0][threadIdx.x] + ws[1][threadIdx.x] + ws[2][threadIdx.x] + ws[3][threadIdx.x];
      a.bpart[8 * (size_t)blockIdx.x + threadIdx.x] = t;
   }
}

struct EfArgs {
   int ns, nsp, nsxf, nrows;
   double c2;
   const KRow *rows;
   const int *rpart;        // [npart+1] row ranges of the partitions
   const double *w;
   const double2 *tx, *ty, *tz, *tz1;
   double *fpart;           // [npart][3][nsp]
};

// Force back-projection (ewald.c:557-579): one thread per site, the rows
// split over blockIdx.y so that enough waves are resident to hide the
// latency of the l recurrence; partials are summed in a fixed order.
__global__ __launch_bounds__(256) void k_ewald_force(EfArgs a) {
   int i = blockIdx.x * 256 + threadIdx.x;
   if (i >= a.ns) return;
   const int part = blockIdx.y;
   const bool fw = i >= a.nsxf;
   const int wo = fw ? 2 : 0;
   const double2 z1 = a.tz1[i];
   double fx = 0, fy = 0, fz = 0;
   int hcur = -1;
   double2 xh = make_double2(1, 0);
   for (int r = a.rpart[part]; r < a.rpart[part + 1]; r++) {
      const KRow R = a.rows[r];
      if (R.h != hcur) {
         hcur = R.h;
         xh = a.tx[(size_t)R.h * a.nsp + i];
      }
      int ak = R.k < 0 ? -R.k : R.k;
      double2 yk = a.ty[(size_t)ak * a.nsp + i];
      if (R.k < 0) yk.y = -yk.y;
      double2 A = cmul(xh, yk);
      int al0 = R.l0 < 0 ? -R.l0 : R.l0;
      double2 z = a.tz[(size_t)al0 * a.nsp + i];
      if (R.l0 < 0) z.y = -z.y;
      const double *w = a.w + 4 * (size_t)R.kstart + wo;
      double kzl = 0;
      for (int q = 0; q < R.nl; q++) {
         double2 e = cmul(A, z);             // q cos(k.r), q sin(k.r)
         double wc = w[4 * q], wsn = w[4 * q + 1];
         double fc = e.y * wc - e.x * wsn;
         int l = R.l0 + q;
         kzl = R.kzt + l * a.c2;
         fx += R.kx * fc;
         fy += R.ky * fc;
         fz += kzl * fc;
         z = cmul(z, z1);
      }
   }
   double *fp = a.fpart + (size_t)part * 3 * a.nsp;
   fp[i] = fx;
   fp[a.nsp + i] = fy;
   fp[2 * a.nsp + i] = fz;
}

__global__ void k_add_parts(int ns, int nsp, int npart, const double *__restrict__ fpart,
                            double *__restrict__ force) {
   int i = blockIdx.x * 256 + threadIdx.x;
   if (i >= ns) return;
   for (int c = 0; c < 3; c++) {
      double s = 0.0;
      for (int p = 0; p < npart; p++) s += fpart[((size_t)p * 3 + c) * nsp + i];
      force[(size_t)c * nsp + i] += s;
   }
}


}  // namespace

// Shared setup: trig tables (FP64) for the current sites, the z table
// weighted by w (the site charges for the energy and forces).
static void ewald_tables(mgpu_engine *e, const double *w) {
   size_t nsp = e->nsp;
   e->d_tx.reserve((e->hmax + 1) * nsp);
   e->d_ty.reserve((e->kmax + 1) * nsp);
   e->d_tz.reserve((e->lmax + 1) * nsp + nsp);
   TrigArgs t;
   double hi2[9];
   for (int i = 0; i < 9; i++) hi2[i] = 2 * PI_D * e->hinv[i];
   for (int i = 0; i < 9; i++) t.ks[i] = hi2[i];
   t.hmax = e->hmax; t.kmax = e->kmax; t.lmax = e->lmax; t.ns = e->nsites; t.nsp = e->nsp;
   ProfScope ps(e, "ewald_trig");
   k_trig<<<(e->nsites + 255) / 256, 256, 0, e->stream>>>(
       t, e->d_site.p, w, e->d_tx.p, e->d_ty.p, e->d_tz.p, e->d_tz.p + (size_t)(e->lmax + 1) * nsp);
}

// k groups: <= 256 k vectors, <= KROWS rows each, and row partitions of
// the force pass with about equal numbers of k vectors; both are host
// tables uploaded when the k set changes
static void ewald_plan(mgpu_engine *e) {
   if (e->ew_plan_version != e->geom_version) {
      std::vector<SkGroup> groups;
      int r = 0;
      while (r < e->nrows) {
         SkGroup g{r, 0, e->rows[r].kstart, 0};
         while (r < e->nrows && g.nr < KROWS && g.nkg + e->rows[r].nl <= 256) {
            g.nkg += e->rows[r].nl;
            g.nr++;
            r++;
         }
         groups.push_back(g);
      }
      static_assert(sizeof(SkGroup) == sizeof(int4), "SkGroup layout");
      e->d_groups.reserve(groups.size() + 1);
      HIPCHK(hipMemcpyAsync(e->d_groups.p, groups.data(), sizeof(SkGroup) * groups.size(),
                            hipMemcpyHostToDevice, e->stream));
      const int npart = std::max(1, std::min(e->tune_npart, e->nrows / 8));
      std::vector<int> rp(npart + 1, e->nrows);
      rp[0] = 0;
      int p = 1;
      for (int q = 0; q < e->nrows && p < npart; q++)
         if ((long)e->rows[q].kstart * npart >= (long)p * e->nk) rp[p++] = q;
      for (; p < npart; p++) rp[p] = e->nrows;
      e->d_rpart.reserve(npart + 1);
      HIPCHK(hipMemcpyAsync(e->d_rpart.p, rp.data(), sizeof(int) * (npart + 1), hipMemcpyHostToDevice, e->stream));
      HIPCHK(hipStreamSynchronize(e->stream));
      e->ew_ngroups = (int)groups.size();
      e->ew_npart = npart;
      e->ew_plan_version = e->geom_version;
   }
}

void mgpu::run_ewald_fp64(mgpu_engine *e) {
   ewald_plan(e);
   ewald_tables(e, e->d_chg_site.p);
   const size_t nsp = e->nsp;
   const double2 *tz1 = e->d_tz.p + (size_t)(e->lmax + 1) * nsp;
   int nchunks = (e->nsites + CHUNK - 1) / CHUNK;
   e->d_sk.reserve((size_t)nchunks * e->nk * 4 + 4);
   SkArgs s;
   s.nk = e->nk; s.ns = e->nsites; s.nsp = e->nsp; s.nsxf = e->nsitesxf; s.lmax = e->lmax;
   s.nchunks = nchunks; s.rows = e->d_rows.p; s.groups = (const SkGroup *)e->d_groups.p; s.krow = e->d_krow.p;
   s.tx = e->d_tx.p; s.ty = e->d_ty.p; s.tz = e->d_tz.p; s.part = e->d_sk.p;
   size_t shz = sizeof(double2) * (e->lmax + 1) * (TS + 1);
   {
      ProfScope ps(e, "ewald_sfac");
      dim3 g(e->ew_ngroups, nchunks);
      if (e->nsitesxf < e->nsites)
         k_sfac<true><<<g, 256, shz, e->stream>>>(s);
      else
         k_sfac<false><<<g, 256, shz, e->stream>>>(s);
   }
   int nkb = (e->nk + 255) / 256;
   e->d_wk.reserve((size_t)e->nk * 4 + 4);
   e->d_red.reserve((size_t)nkb * 8 + 16);
   EkArgs ek;
   ek.nk = e->nk; ek.nchunks = nchunks; ek.nrows = e->nrows;
   ek.vol = host_det(e->h);
   ek.r4a = -1.0 / (4.0 * e->alpha * e->alpha);
   ek.c2 = 2 * PI_D * e->hinv[8];
   ek.rows = e->d_rows.p; ek.krow = e->d_krow.p; ek.part = e->d_sk.p; ek.w = e->d_wk.p;
   ek.bpart = e->d_red.p;
   {
      ProfScope ps(e, "ewald_kspace");
      k_ewald_k<<<nkb, 256, 0, e->stream>>>(ek);
   }
   e->d_out.reserve(OUT_SIZE);
   reduce_cols(e, e->d_red.p, nkb, 8, e->d_out.p + OUT_RECIP);
   const int npart = e->ew_npart;
   e->d_fpart.reserve((size_t)npart * 3 * nsp);
   EfArgs ef;
   ef.ns = e->nsites; ef.nsp = e->nsp; ef.nsxf = e->nsitesxf; ef.nrows = e->nrows;
   ef.c2 = ek.c2; ef.rows = e->d_rows.p; ef.rpart = e->d_rpart.p; ef.w = e->d_wk.p;
   ef.tx = e->d_tx.p; ef.ty = e->d_ty.p; ef.tz = e->d_tz.p; ef.tz1 = tz1;
   ef.fpart = e->d_fpart.p;
   {
      ProfScope ps(e, "ewald_force");
      dim3 g((e->nsites + 255) / 256, npart);
      k_ewald_force<<<g, 256, 0, e->stream>>>(ef);
      join_side(e);
      k_add_parts<<<(e->nsites + 255) / 256, 256, 0, e->stream>>>(e->nsites, e->nsp, npart, e->d_fpart.p,
                                                                  e->d_force_site.p);
   }
}

// ============================ charge adjoint ============================
// With E = 1/2 sum_k coeff (C^2 + S^2), C + iS = sum_j q_j exp(i k.r_j) and
// F_i = q_i sum_k coeff k (C sin_i - S cos_i) (ewald.c:539-579), the charge
// derivatives by site id a use the id-restricted sums C_a + iS_a =
// sum_{j in a} exp(i k.r_j):
//   dE/dq_a       = sum_k coeff (C C_a + S S_a)
//   dF_i/dq_a     = [id_i = a] phi_i + q_i sum_k coeff k (C_a sin_i - S_a cos_i)
// with phi_i = sum_k coeff k (C sin_i - S cos_i), so that for a cotangent v
//   sum_i v_i.dF_i/dq_a = sum_{i in a} v_i.phi_i + sum_k coeff (C_a T_S - S_a T_C),
//   T_C + i T_S = sum_i q_i (v_i.k) exp(i k.r_i).
namespace {

__global__ void k_sf_sum(int nk, int nchunks, const double *__restrict__ part, double2 *__restrict__ out) {
   int k = blockIdx.x * blockDim.x + threadIdx.x;
   if (k >= nk) return;
   double c = 0, s = 0;
   for (int q = 0; q < nchunks; q++) {
      const double *p = part + ((size_t)q * nk + k) * 4;
      c += p[0];
      s += p[1];
   }
   out[k] = make_double2(c, s);
}

struct KcArgs {
   int nk;
   double vol, r4a, c2;
   const KRow *rows;
   const int *krow;
};

// coeff and k of every k vector (ewald.c:506-531)
__global__ void k_kcoeff(KcArgs a, double4 *__restrict__ out) {
   int k = blockIdx.x * blockDim.x + threadIdx.x;
   if (k >= a.nk) return;
   KRow R = a.rows[a.krow[k]];
   int l = R.l0 + (k - R.kstart);
   double kx = R.kx, ky = R.ky, kz = R.kzt + l * a.c2;
   double ksq = kx * kx + ky * ky + kz * kz;
   out[k] = make_double4(kx, ky, kz, 2.0 / ((0.25 / PI_D) * a.vol) * exp(ksq * a.r4a) / ksq);
}

__global__ void k_weights_w(int nk, const double4 *__restrict__ kc, const double2 *__restrict__ cs,
                            double *__restrict__ w) {
   int k = blockIdx.x * blockDim.x + threadIdx.x;
   if (k >= nk) return;
   double c = kc[k].w;
   w[4 * (size_t)k] = c * cs[k].x;
   w[4 * (size_t)k + 1] = c * cs[k].y;
   w[4 * (size_t)k + 2] = c * cs[k].x;
   w[4 * (size_t)k + 3] = c * cs[k].y;
}

__global__ void k_site_weight(int ns, int mode, int id, const int *__restrict__ stype,
                              const double *__restrict__ chg, const double *__restrict__ v, int nsp,
                              double *__restrict__ w) {
   int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= nsp) return;
   double x = 0.0;
   if (i < ns) {
      if (mode == 0) x = stype[i] == id ? 1.0 : 0.0;       // indicator of site id
      else if (mode == 1) x = 1.0;                         // unit weights
      else x = chg[i] * v[(size_t)(mode - 2) * nsp + i];  // q v_x, q v_y, q v_z
   }
   w[i] = x;
}

// per k: columns 2a (dE/dq_a) and 2a+1 (the k part of the cotangent product)
__global__ __launch_bounds__(256) void k_charge_k(int nk, int nid, const double4 *__restrict__ kc,
                                                  const double2 *__restrict__ csq,
                                                  const double2 *__restrict__ csa,
                                                  const double2 *__restrict__ csv, int have_v,
                                                  double *__restrict__ part) {
   __shared__ double ws[4][2 * MGPU_MAX_ID];
   int k = blockIdx.x * 256 + threadIdx.x;
   const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
   double4 K = k < nk ? kc[k] : make_double4(0, 0, 0, 0);
   double2 cq = k < nk ? csq[k] : make_double2(0, 0);
   double tc = 0, tsn = 0;
   if (k < nk && have_v) {
      double2 vx = csv[k], vy = csv[(size_t)nk + k], vz = csv[2 * (size_t)nk + k];
      tc = K.x * vx.x + K.y * vy.x + K.z * vz.x;
      tsn = K.x * vx.y + K.y * vy.y + K.z * vz.y;
   }
   for (int a = 0; a < nid; a++) {
      double2 ca = k < nk ? csa[(size_t)a * nk + k] : make_double2(0, 0);
      double de = K.w * (cq.x * ca.x + cq.y * ca.y);
      double dv = K.w * (ca.x * tsn - ca.y * tc);
      de = wave_sum(de);
      dv = wave_sum(dv);
      if (lane == 0) { ws[w][2 * a] = de; ws[w][2 * a + 1] = dv; }
   }
   __syncthreads();
   if (threadIdx.x < 2 * nid)
      part[(size_t)blockIdx.x * 2 * nid + threadIdx.x] =
          ws[0][threadIdx.x] + ws[1][threadIdx.x] + ws[2][threadIdx.x] + ws[3][threadIdx.x];
}

// per site id a: sum_{i in a} v_i.phi_i
__global__ __launch_bounds__(256) void k_vphi(int ns, int nsp, int nid, const int *__restrict__ stype,
                                              const double *__restrict__ v, const double *__restrict__ phi,
                                              double *__restrict__ part) {
   __shared__ double ws[4][MGPU_MAX_ID];
   int i = blockIdx.x * 256 + threadIdx.x;
   const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
   double x = 0;
   int t = -1;
   if (i < ns) {
      t = stype[i];
      x = v[i] * phi[i] + v[nsp + i] * phi[nsp + i] + v[2 * (size_t)nsp + i] * phi[2 * (size_t)nsp + i];
   }
   for (int a = 0; a < nid; a++) {
      double y = wave_sum(t == a ? x : 0.0);
      if (lane == 0) ws[w][a] = y;
   }
   __syncthreads();
   if (threadIdx.x < nid)
      part[(size_t)blockIdx.x * nid + threadIdx.x] =
          ws[0][threadIdx.x] + ws[1][threadIdx.x] + ws[2][threadIdx.x] + ws[3][threadIdx.x];
}

__global__ void k_colsum(int nrows, int ncols, const double *__restrict__ part, double *__restrict__ out) {
   int c = blockIdx.x * blockDim.x + threadIdx.x;
   if (c >= ncols) return;
   double t = 0;
   for (int r = 0; r < nrows; r++) t += part[(size_t)r * ncols + c];
   out[c] = t;
}

// dS = sum_a dq_a (C_a, S_a)
__global__ void k_dir_sf(int nk, int nid, const double *__restrict__ dq, const double2 *__restrict__ csa,
                         double2 *__restrict__ out) {
   int k = blockIdx.x * blockDim.x + threadIdx.x;
   if (k >= nk) return;
   double c = 0, s = 0;
   for (int a = 0; a < nid; a++) {
      double2 x = csa[(size_t)a * nk + k];
      c += dq[a] * x.x;
      s += dq[a] * x.y;
   }
   out[k] = make_double2(c, s);
}

// jvp_i += dq_{id_i} phi_i
__global__ void k_jvp_phi(int ns, int nsp, const int *__restrict__ stype, const double *__restrict__ dq,
                          const double *__restrict__ phi, double *__restrict__ out) {
   int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= ns) return;
   double d = dq[stype[i]];
   for (int c = 0; c < 3; c++) out[(size_t)c * nsp + i] += d * phi[(size_t)c * nsp + i];
}

}  // namespace

// sum_i w_i exp(i k.r_i) for every k into cs[nk]
static void weighted_sf(mgpu_engine *e, const double *w, double2 *cs) {
   ewald_tables(e, w);
   int nchunks = (e->nsites + CHUNK - 1) / CHUNK;
   e->d_sk.reserve((size_t)nchunks * e->nk * 4 + 4);
   SkArgs s;
   s.nk = e->nk; s.ns = e->nsites; s.nsp = e->nsp; s.nsxf = e->nsites; s.lmax = e->lmax;
   s.nchunks = nchunks; s.rows = e->d_rows.p; s.groups = (const SkGroup *)e->d_groups.p; s.krow = e->d_krow.p;
   s.tx = e->d_tx.p; s.ty = e->d_ty.p; s.tz = e->d_tz.p; s.part = e->d_sk.p;
   size_t shz = sizeof(double2) * (e->lmax + 1) * (TS + 1);
   dim3 g(e->ew_ngroups, nchunks);
   k_sfac<false><<<g, 256, shz, e->stream>>>(s);
   k_sf_sum<<<(e->nk + 255) / 256, 256, 0, e->stream>>>(e->nk, nchunks, e->d_sk.p, cs);
}

// out[3][nsp] += w_i sum_k coeff k (C sin_i - S cos_i) with (C, S) = cs
static void backproject(mgpu_engine *e, const double *w, const double4 *kc, const double2 *cs, double *out) {
   e->d_wk.reserve((size_t)e->nk * 4 + 4);
   k_weights_w<<<(e->nk + 255) / 256, 256, 0, e->stream>>>(e->nk, kc, cs, e->d_wk.p);
   ewald_tables(e, w);
   const int npart = e->ew_npart;
   e->d_fpart.reserve((size_t)npart * 3 * e->nsp);
   EfArgs ef;
   ef.ns = e->nsites; ef.nsp = e->nsp; ef.nsxf = e->nsites; ef.nrows = e->nrows;
   ef.c2 = 2 * PI_D * e->hinv[8]; ef.rows = e->d_rows.p; ef.rpart = e->d_rpart.p; ef.w = e->d_wk.p;
   ef.tx = e->d_tx.p; ef.ty = e->d_ty.p; ef.tz = e->d_tz.p; ef.tz1 = e->d_tz.p + (size_t)(e->lmax + 1) * e->nsp;
   ef.fpart = e->d_fpart.p;
   dim3 g((e->nsites + 255) / 256, npart);
   k_ewald_force<<<g, 256, 0, e->stream>>>(ef);
   k_add_parts<<<(e->nsites + 255) / 256, 256, 0, e->stream>>>(e->nsites, e->nsp, npart, e->d_fpart.p, out);
}

void mgpu::ewald_charge_adjoint(mgpu_engine *e, const double *v, const double *dq_dev, double *dE, double *vjp,
                                double *jvp) {
   if (e->nsitesxf < e->nsites)
      throw StatusError{MGPU_E_CAPACITY, "charge derivatives of the Ewald sum with a framework"};
   ewald_plan(e);
   const int nk = e->nk, nid = e->max_id, nsp = e->nsp, ns = e->nsites;
   const int nb = (nsp + 255) / 256;
   // scratch: coeff/k, C+iS for q, each id, q v_x..z, direction; phi; weights
   size_t off = 0;
   auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~(size_t)255; return o; };
   const size_t oKc = take(sizeof(double4) * nk), oQ = take(sizeof(double2) * nk),
                oA = take(sizeof(double2) * nk * (size_t)nid), oV = take(sizeof(double2) * nk * 3),
                oD = take(sizeof(double2) * nk), oPhi = take(sizeof(double) * 3 * nsp),
                oW = take(sizeof(double) * nsp), oP = take(sizeof(double) * (size_t)std::max(nk, nsp) / 256 * 2 * nid + 64 * nid);
   e->d_adj.reserve(off);
   uint8_t *B = e->d_adj.p;
   double4 *kc = (double4 *)(B + oKc);
   double2 *csq = (double2 *)(B + oQ), *csa = (double2 *)(B + oA), *csv = (double2 *)(B + oV),
           *csd = (double2 *)(B + oD);
   double *phi = (double *)(B + oPhi), *w = (double *)(B + oW), *part = (double *)(B + oP);
   e->d_out.reserve(OUT_SIZE);
   double *outc = e->d_out.p + OUT_TMP;
   KcArgs ka{nk, host_det(e->h), -1.0 / (4.0 * e->alpha * e->alpha), 2 * PI_D * e->hinv[8], e->d_rows.p, e->d_krow.p};
   k_kcoeff<<<(nk + 255) / 256, 256, 0, e->stream>>>(ka, kc);
   weighted_sf(e, e->d_chg_site.p, csq);
   for (int a = 0; a < nid; a++) {
      k_site_weight<<<nb, 256, 0, e->stream>>>(ns, 0, a, e->d_site_type.p, e->d_chg_site.p, nullptr, nsp, w);
      weighted_sf(e, w, csa + (size_t)a * nk);
   }
   if (v)
      for (int c = 0; c < 3; c++) {
         k_site_weight<<<nb, 256, 0, e->stream>>>(ns, 2 + c, 0, e->d_site_type.p, e->d_chg_site.p, v, nsp, w);
         weighted_sf(e, w, csv + (size_t)c * nk);
      }
   // phi_i: unit site weights against the charges' structure factors
   zero_async(e, phi, sizeof(double) * 3 * nsp);
   k_site_weight<<<nb, 256, 0, e->stream>>>(ns, 1, 0, e->d_site_type.p, e->d_chg_site.p, nullptr, nsp, w);
   backproject(e, w, kc, csq, phi);
   // dE/dq and the k part of the cotangent product
   const int nkb = (nk + 255) / 256;
   k_charge_k<<<nkb, 256, 0, e->stream>>>(nk, nid, kc, csq, csa, csv, v != nullptr, part);
   k_colsum<<<1, 64, 0, e->stream>>>(nkb, 2 * nid, part, outc);
   std::vector<double> h2(2 * nid), hv(nid, 0.0);
   HIPCHK(hipMemcpyAsync(h2.data(), outc, sizeof(double) * 2 * nid, hipMemcpyDeviceToHost, e->stream));
   HIPCHK(hipStreamSynchronize(e->stream));
   if (v) {
      k_vphi<<<nb, 256, 0, e->stream>>>(ns, nsp, nid, e->d_site_type.p, v, phi, part);
      k_colsum<<<1, 64, 0, e->stream>>>(nb, nid, part, outc);
      HIPCHK(hipMemcpyAsync(hv.data(), outc, sizeof(double) * nid, hipMemcpyDeviceToHost, e->stream));
      HIPCHK(hipStreamSynchronize(e->stream));
   }
   for (int a = 0; a < nid; a++) {
      if (dE) dE[a] = h2[2 * a];
      if (vjp) vjp[a] = h2[2 * a + 1] + hv[a];
   }
   if (jvp && dq_dev) {
      k_dir_sf<<<nkb, 256, 0, e->stream>>>(nk, nid, dq_dev, csa, csd);
      backproject(e, e->d_chg_site.p, kc, csd, jvp);
      k_jvp_phi<<<(ns + 255) / 256, 256, 0, e->stream>>>(ns, nsp, e->d_site_type.p, dq_dev, phi, jvp);
   }
}
