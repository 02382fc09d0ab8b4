// Parameter adjoint of eval_forces' potential energy and site forces (FP64):
// derivatives with respect to every pair type's potential parameters and
// every site id's charge.  This has no counterpart in the reference; it
// differentiates exactly what src/force.c:784-1005 (with kernel.c),
// src/ewald.c:280-589 and the constant terms of src/accel.c:293-329 and
// src/force.c:1158-1168 compute.
//
// Real space: one thread per site walks the same full neighbour shell as the
// force kernel and evaluates each pair with Dual arithmetic, which carries
// du/dp_k, dforceij/dp_k (k < n_potpar) and d/d(q_i q_j) beside the value.
// Each unordered pair is seen from both ends with weight 1/2.  Per site the
// sums are kept by the neighbour's site id in a thread-private slice of LDS
// and written out once; a second kernel folds them into the parameter
// vector.  For a site-force cotangent v the pair (i, j) contributes
// dforceij * (v_j - v_i).(r_j - r_i), since F_i = forceij (r_i - r_j)
// (mk_forces, src/force.c:717-743).
//
// Reciprocal space: see ewald_charge_adjoint (mgpu_ewald.hip).
// Constant terms (intramolecular correction, distant potential, self and
// sheet energies): differentiated on the host with the same Dual pair term.
#include "mgpu_internal.h"
#include "mgpu_pair.h"

#include <math.h>
#include <algorithm>

using namespace mgpu;

namespace {

constexpr int ND = MGPU_NPOTP;          // derivative slots: parameters 0..6, charge product 7
constexpr int DQ = MGPU_NPOTP - 1;
using D8 = Dual<ND>;
constexpr double PI_D = 3.14159265358979323846;

struct AdjArgs {
   int nsites, ncells, nx, ny, nz, nfull, max_id, npar, nsp, strict_sub, rs;
   double alpha, norm, cutsq, cut100sq;
   double reloc[81];
   const int *css;
   const int4 *off;
   const double4 *cpos;
   const int2 *cmeta;
   const int *corig;
   const double *potpar;   // [max_id][max_id][8]
   const double *v;        // cotangent [3][nsp] or null
   const double *dpar;     // direction [max_id][max_id][8] or null
   const double *dq;       // direction [max_id] or null
   const double *chg_id;   // [max_id]
   double *rec;            // [nsites][rs]
   double *jvp;            // [3][nsp] accumulated, or null
};

// Slots of a site's record (E half, then V half of rs/2 each):
//   [b*(npar+1) + k]   sum over neighbours of id b of 1/2 du/dp_k (k < npar)
//   [b*(npar+1)+npar]  sum over neighbours of id b of 1/2 du/d(qq) q_i
//   [nid*(npar+1)]     sum over all neighbours of 1/2 du/d(qq) q_j
template <int PT, bool COUL>
__global__ __launch_bounds__(64) void k_adj_real(AdjArgs a) {
   extern __shared__ double sacc[];
   const int p = blockIdx.x * 64 + threadIdx.x;
   double *acc = sacc + (size_t)threadIdx.x * a.rs;
   for (int t = 0; t < a.rs; t++) acc[t] = 0.0;
   if (p >= a.nsites) return;
   const int half = a.rs / 2, nslot = a.npar + 1, self_slot = a.max_id * nslot;
   // cell of position p
   int lo = 0, hi = a.ncells;
   while (hi - lo > 1) {
      int mid = (lo + hi) >> 1;
      if (a.css[mid] <= p) lo = mid; else hi = mid;
   }
   const int cell = lo;
   const int ix = cell / (a.ny * a.nz), iy = cell / a.nz - a.ny * ix, iz = cell - a.nz * (iy + a.ny * ix);
   const double4 Pi = a.cpos[p];
   const int2 mi = a.cmeta[p];
   const int ti = mi.x & 0xff, fwi = mi.x >> 8;
   const int io = a.corig[p];
   double vi[3] = {0, 0, 0};
   if (a.v) { vi[0] = a.v[io]; vi[1] = a.v[a.nsp + io]; vi[2] = a.v[2 * (size_t)a.nsp + io]; }
   double jx = 0, jy = 0, jz = 0;
   const D8 norm(a.norm), alpha(a.alpha);
   for (int o = 0; o < a.nfull; o++) {
      const int4 d = a.off[o];
      int ex = ix + d.x, ey = iy + d.y, ez = iz + d.z;
      int sx = ex < 0 ? -1 : (ex >= a.nx ? 1 : 0);
      int sy = ey < 0 ? -1 : (ey >= a.ny ? 1 : 0);
      int sz = ez < 0 ? -1 : (ez >= a.nz ? 1 : 0);
      int jc = (ez - sz * a.nz) + a.nz * ((ey - sy * a.ny) + a.ny * (ex - sx * a.nx));
      int code = 9 * (sx + 1) + 3 * (sy + 1) + (sz + 1);
      const double shx = a.reloc[3 * code], shy = a.reloc[3 * code + 1], shz = a.reloc[3 * code + 2];
      for (int q = a.css[jc]; q < a.css[jc + 1]; q++) {
         if (o == 0 && q == p) continue;
         const int2 mj = a.cmeta[q];
         const int tj = mj.x & 0xff, fwj = mj.x >> 8;
         if (fwi & fwj) continue;
         const double4 Pj = a.cpos[q];
         double rx = Pj.x - Pi.x + shx, ry = Pj.y - Pi.y + shy, rz = Pj.z - Pi.z + shz;
         double rsq = rx * rx + ry * ry + rz * rz;
         if (a.strict_sub && rsq > a.cutsq) rsq = a.cut100sq;
         const double *pp = a.potpar + (size_t)(ti * a.max_id + tj) * MGPU_NPOTP;
         D8 par[MGPU_NPOTP];
#pragma unroll
         for (int k = 0; k < MGPU_NPOTP; k++) par[k] = D8::seed(pp[k], k < a.npar ? k : -1);
         D8 u, f;
         pair_term<D8, PT, COUL>(D8(rsq), D8::seed(Pi.w * Pj.w, DQ), norm, alpha, par, u, f);
         double *eb = acc + tj * nslot;
         for (int k = 0; k < a.npar; k++) eb[k] += 0.5 * u.d[k];
         eb[a.npar] += 0.5 * u.d[DQ] * Pi.w;
         acc[self_slot] += 0.5 * u.d[DQ] * Pj.w;
         if (a.v) {
            const int jo = a.corig[q];
            const double w = (a.v[jo] - vi[0]) * rx + (a.v[a.nsp + jo] - vi[1]) * ry +
                             (a.v[2 * (size_t)a.nsp + jo] - vi[2]) * rz;
            double *vb = acc + half + tj * nslot;
            for (int k = 0; k < a.npar; k++) vb[k] += 0.5 * f.d[k] * w;
            vb[a.npar] += 0.5 * f.d[DQ] * Pi.w * w;
            acc[half + self_slot] += 0.5 * f.d[DQ] * Pj.w * w;
         }
         if (a.jvp) {
            const double *dp = a.dpar + (size_t)(ti * a.max_id + tj) * MGPU_NPOTP;
            double df = 0.0;
            for (int k = 0; k < a.npar; k++) df += dp[k] * f.d[k];
            df += f.d[DQ] * (a.dq[ti] * a.chg_id[tj] + a.dq[tj] * a.chg_id[ti]);
            jx -= df * rx;
            jy -= df * ry;
            jz -= df * rz;
         }
      }
   }
   double *out = a.rec + (size_t)io * a.rs;
   for (int t = 0; t < a.rs; t++) out[t] = acc[t];
   if (a.jvp) {
      a.jvp[io] += jx;
      a.jvp[a.nsp + io] += jy;
      a.jvp[2 * (size_t)a.nsp + io] += jz;
   }
}

// Fold the per-site records into theta: one thread per theta entry, sites in
// index order (deterministic).  half = 0 for dE, rs/2 for the cotangent part.
__global__ void k_adj_fold(int nsites, int max_id, int npar, int rs, int half, const int *__restrict__ stype,
                           const double *__restrict__ rec, double *__restrict__ out) {
   const int ntypes = max_id * (max_id + 1) / 2, nslot = npar + 1;
   const int col = blockIdx.x * blockDim.x + threadIdx.x;
   if (col >= ntypes * npar + max_id) return;
   double s = 0.0;
   if (col < ntypes * npar) {
      int t = col / npar, k = col - t * npar, a = 0;
      while (t >= max_id - a) { t -= max_id - a; a++; }
      const int b = a + t;
      for (int i = 0; i < nsites; i++) {
         const int ti = stype[i];
         const double *r = rec + (size_t)i * rs + half;
         if (ti == a) s += r[b * nslot + k];
         if (a != b && ti == b) s += r[a * nslot + k];
      }
   } else {
      const int a = col - ntypes * npar;
      for (int i = 0; i < nsites; i++) {
         const double *r = rec + (size_t)i * rs + half;
         if (stype[i] == a) s += r[max_id * nslot];
         s += r[a * nslot + npar];
      }
   }
   out[col] = s;
}

template <int PT, bool COUL> void launch_adj(dim3 g, size_t shm, hipStream_t s, const AdjArgs &a) {
   if (shm > 64 * 1024)
      HIPCHK(hipFuncSetAttribute((const void *)k_adj_real<PT, COUL>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)shm));
   k_adj_real<PT, COUL><<<g, 64, shm, s>>>(a);
}

template <bool COUL> void launch_adj_pt(int pt, dim3 g, size_t shm, hipStream_t s, const AdjArgs &a) {
   switch (pt) {
   case 0: launch_adj<0, COUL>(g, shm, s, a); break;
   case 1: launch_adj<1, COUL>(g, shm, s, a); break;
   case 2: launch_adj<2, COUL>(g, shm, s, a); break;
   case 3: launch_adj<3, COUL>(g, shm, s, a); break;
   case 4: launch_adj<4, COUL>(g, shm, s, a); break;
   case 6: launch_adj<6, COUL>(g, shm, s, a); break;
   default: throw StatusError{MGPU_E_PTYPE, "unknown potential type"};
   }
}

// ---------------------------------------------------------------- host
using D1 = Dual<1>;

template <int PT> D1 pair_u_pt(double rsq, D1 qq, double alpha, const D1 *p) {
   D1 u, f;
   if (alpha > 0.0)
      pair_term<D1, PT, true>(D1(rsq), qq, D1(2.0 * alpha / sqrt(PI_D)), D1(alpha), p, u, f);
   else
      pair_term<D1, PT, false>(D1(rsq), qq, D1(0.0), D1(alpha), p, u, f);
   return u;
}

D1 pair_u(int pt, double rsq, D1 qq, double alpha, const D1 *p) {
   switch (pt) {
   case 0: return pair_u_pt<0>(rsq, qq, alpha, p);
   case 1: return pair_u_pt<1>(rsq, qq, alpha, p);
   case 2: return pair_u_pt<2>(rsq, qq, alpha, p);
   case 3: return pair_u_pt<3>(rsq, qq, alpha, p);
   case 4: return pair_u_pt<4>(rsq, qq, alpha, p);
   case 6: return pair_u_pt<6>(rsq, qq, alpha, p);
   default: throw StatusError{MGPU_E_PTYPE, "unknown potential type"};
   }
}

D1 texp(D1 x) { return Math<D1>::exp_(x); }

// dist_pot, src/kernel.c:103-151, in Dual arithmetic
D1 dist_pot_d(const D1 *p, double rc, int pt) {
   auto sq = [](D1 x) { return x * x; };
   auto cu = [](D1 x) { return x * x * x; };
   const double rc3 = rc * rc * rc;
   switch (pt) {
   case 0: return p[0] * cu(sq(p[1]) * (1.0 / rc)) * (1.0 / 3.0);
   case 1:
      if (p[2].v > 1e-7)
         return p[0] * (1.0 / (3.0 * rc3)) -
                p[1] * texp(-p[2] * rc) *
                    (rc * rc * Math<D1>::rcp(p[2]) + 2 * rc * Math<D1>::rcp(sq(p[2])) + 2.0 * Math<D1>::rcp(cu(p[2])));
      return p[0] * (1.0 / (3.0 * rc3));
   case 2:
      if (p[3].v > 1e-7)
         return p[2] * (rc * rc * Math<D1>::rcp(p[3]) + 2 * rc * Math<D1>::rcp(sq(p[3])) + 2.0 * Math<D1>::rcp(cu(p[3]))) *
                texp(-p[3] * rc);
      return D1(0.0);
   case 3: {
      D1 tail = -p[2] * (1.0 / (9.0 * rc3 * rc3 * rc3)) + p[3] * (1.0 / rc) + p[4] * (1.0 / (3.0 * rc3)) +
                p[5] * (1.0 / (5.0 * rc * rc * rc3));
      if (p[1].v > 1e-7)
         return -p[0] * texp(-p[1] * rc) *
                    (rc * rc * Math<D1>::rcp(p[1]) + 2 * rc * Math<D1>::rcp(sq(p[1])) + 2.0 * Math<D1>::rcp(cu(p[1]))) +
                tail;
      return tail;
   }
   case 6:
      if (p[5].v != 0.0)
         return p[3] * (1.0 / (3.0 * rc3)) +
                2.0 * p[4] *
                    (rc * rc * Math<D1>::rcp(p[5]) + 2 * rc * Math<D1>::rcp(sq(p[5])) + 2.0 * Math<D1>::rcp(cu(p[5]))) *
                    texp(-p[5] * (rc - p[6]));
      return p[3] * (1.0 / (3.0 * rc3));
   case 4: return -p[0] * (1.0 / rc) - p[1] * (1.0 / (3.0 * rc3)) - p[2] * (1.0 / (9.0 * rc3 * rc3 * rc3));
   default: return D1(0.0);
   }
}

double err_fn_host(double x) {   // src/auxil.c:586-593
   double t = 1.0 / (1.0 + AS_P * x);
   if (x >= 0.0) return 1.0 - t * (AS_A1 + t * (AS_A2 + t * (AS_A3 + t * (AS_A4 + t * AS_A5)))) * exp(-x * x);
   return -err_fn_host(-x);
}

// Derivatives of the constant terms of pe[0] + pe[1]: -eintra + dist/V
// (potential parameters and charges) and -self + sheet/V (charges).
void static_grad(mgpu_engine *e, int npar, double *g) {
   const int nid = e->max_id, ntypes = nid * (nid + 1) / 2;
   const double vol = host_det(e->h);
   auto tcol = [&](int a, int b) {
      int lo = std::min(a, b), hi = std::max(a, b), t = 0;
      for (int x = 0; x < lo; x++) t += nid - x;
      return t + (hi - lo);
   };
   std::vector<double> q(e->site_charge);
   // one theta entry at a time: seed it, evaluate the terms in Dual<1>
   const int ntheta = ntypes * npar + nid;
   for (int th = 0; th < ntheta; th++) {
      const bool is_q = th >= ntypes * npar;
      const int qa = is_q ? th - ntypes * npar : -1;
      int ta = -1, tb = -1, tk = -1;
      if (!is_q) {
         int t = th / npar;
         tk = th - t * npar;
         int a = 0;
         while (t >= nid - a) { t -= nid - a; a++; }
         ta = a;
         tb = a + t;
      }
      auto par = [&](int a, int b, D1 *out) {
         const double *pp = e->potpar.data() + MGPU_NPOTP * (a * nid + b);
         bool hit = !is_q && tcol(a, b) == tcol(ta, tb);
         for (int k = 0; k < MGPU_NPOTP; k++) out[k] = D1::seed(pp[k], hit && k == tk ? 0 : -1);
      };
      auto chg = [&](int a) { return D1::seed(q[a], a == qa ? 0 : -1); };
      double total = 0.0;
      // -eintra, src/force.c:654-668 and 1158-1168
      for (int s = 0; s < e->nspecies; s++) {
         if (e->spec_fw[s]) continue;
         const int ns = e->spec_nsites[s];
         const int *id = e->spec_site_id.data() + e->spec_off[s];
         const double *pf = e->spec_pf.data() + 3 * e->spec_off[s];
         double ei = 0.0;
         for (int b = 0; b < ns; b++)
            for (int a = b + 1; a < ns; a++) {
               double d2 = 0;
               for (int c = 0; c < 3; c++) d2 += (pf[3 * a + c] - pf[3 * b + c]) * (pf[3 * a + c] - pf[3 * b + c]);
               double d = sqrt(d2);
               D1 p[MGPU_NPOTP];
               par(id[b], id[a], p);
               ei += pair_u(e->ptype, d * d, chg(id[a]) * chg(id[b]), e->alpha, p).d[0];
            }
         total -= e->spec_nmols[s] * ei;
      }
      // dist / V, src/accel.c:293-329 (potential parameters only)
      if (!is_q) {
         std::vector<long> count(nid, 0);
         for (int sp = 0; sp < e->nspecies; sp++)
            for (int is = 0; is < e->spec_nsites[sp]; is++)
               count[e->spec_site_id[e->spec_off[sp] + is]] += e->spec_nmols[sp];
         for (int a = 1; a < nid; a++)
            for (int b = 1; b < nid; b++) {
               D1 p[MGPU_NPOTP];
               par(a, b, p);
               total -= 2 * PI_D * count[a] * count[b] * dist_pot_d(p, e->cutoff, e->ptype).d[0] / vol;
            }
      }
      // -self + sheet / V, src/ewald.c:367-425 (charges only)
      if (is_q && e->alpha > 0) {
         double dself = 0.0, sq = 0.0;
         int ssite = 0, s = 0;
         std::vector<double> chs;
         for (int sp = 0; sp < e->nspecies; sp++)
            for (int m = 0; m < e->spec_nmols[sp]; m++)
               for (int is = 0; is < e->spec_nsites[sp]; is++) chs.push_back(q[e->spec_site_id[e->spec_off[sp] + is]]);
         for (; s < e->nspecies && !e->spec_fw[s]; s++) {
            const int ns = e->spec_nsites[s];
            const int *id = e->spec_site_id.data() + e->spec_off[s];
            const double *pf = e->spec_pf.data() + 3 * e->spec_off[s];
            double intra = 0.0;
            for (int a = 0; a < ns; a++)
               for (int b = a + 1; b < ns; b++) {
                  double d2 = 0;
                  for (int c = 0; c < 3; c++) d2 += (pf[3 * a + c] - pf[3 * b + c]) * (pf[3 * a + c] - pf[3 * b + c]);
                  double r = sqrt(d2);
                  double dqq = (id[a] == qa ? q[id[b]] : 0.0) + (id[b] == qa ? q[id[a]] : 0.0);
                  intra += dqq * err_fn_host(e->alpha * r) / r;
               }
            dself += e->spec_nmols[s] * intra;
            ssite += ns * e->spec_nmols[s];
         }
         long na = 0;
         for (int i = 0; i < ssite; i++) {
            sq += chs[i];
            if (e->site_type[i] == qa) na++;
         }
         dself += e->alpha / sqrt(PI_D) * 2.0 * q[qa] * na;
         total -= dself;
         double sqall = sq;
         long naall = na;
         for (int i = ssite; i < e->nsites; i++) {
            sqall += chs[i];
            if (e->site_type[i] == qa) naall++;
         }
         const double conv_q = 4.07497263794495e-14 * 1.e-3 * 1.05482230112e-05 / 1.60217733e-19;
         if (fabs(sqall) * conv_q > 1.0e-5)
            total -= PI_D * 2.0 * sqall * naall / (2.0 * e->alpha * e->alpha) / vol;
      }
      g[th] += total;
   }
}

}  // namespace

int mgpu::adjoint_ntheta(const mgpu_engine *e) {
   const int nid = e->max_id;
   return nid * (nid + 1) / 2 * e->npar + nid;
}

void mgpu::run_adjoint(mgpu_engine *e, const double *v_host, const double *dtheta, double *dE, double *vjp,
                       double *jvp_host) {
   if (e->surface_dipole && e->alpha > 1e-7)
      throw StatusError{MGPU_E_CAPACITY, "the parameter adjoint does not include the surface-dipole term"};
   const int nid = e->max_id, npar = e->npar, ntypes = nid * (nid + 1) / 2;
   const int ntheta = ntypes * npar + nid, ns = e->nsites, nsp = e->nsp;
   const int rs = 2 * (nid * (npar + 1) + 1);
   // device inputs
   size_t off = 0;
   auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~(size_t)255; return o; };
   const size_t oRec = take(sizeof(double) * (size_t)ns * rs), oV = take(sizeof(double) * 3 * nsp),
                oJ = take(sizeof(double) * 3 * nsp), oDP = take(sizeof(double) * nid * nid * MGPU_NPOTP),
                oDQ = take(sizeof(double) * nid), oQ = take(sizeof(double) * nid), oT = take(sizeof(double) * 2 * ntheta);
   e->d_adj_real.reserve(off);
   uint8_t *B = e->d_adj_real.p;
   double *rec = (double *)(B + oRec), *dv = (double *)(B + oV), *dj = (double *)(B + oJ),
          *ddp = (double *)(B + oDP), *ddq = (double *)(B + oDQ), *dqid = (double *)(B + oQ),
          *dth = (double *)(B + oT);
   std::vector<double> hdp((size_t)nid * nid * MGPU_NPOTP, 0.0), hdq(nid, 0.0);
   if (dtheta) {
      for (int a = 0; a < nid; a++)
         for (int b = a; b < nid; b++) {
            int col = 0;
            for (int x = 0; x < a; x++) col += nid - x;
            col += b - a;
            for (int k = 0; k < npar; k++) {
               hdp[(size_t)(a * nid + b) * MGPU_NPOTP + k] = dtheta[col * npar + k];
               hdp[(size_t)(b * nid + a) * MGPU_NPOTP + k] = dtheta[col * npar + k];
            }
         }
      for (int a = 0; a < nid; a++) hdq[a] = dtheta[ntypes * npar + a];
   }
   HIPCHK(hipMemcpyAsync(ddp, hdp.data(), sizeof(double) * hdp.size(), hipMemcpyHostToDevice, e->stream));
   HIPCHK(hipMemcpyAsync(ddq, hdq.data(), sizeof(double) * nid, hipMemcpyHostToDevice, e->stream));
   HIPCHK(hipMemcpyAsync(dqid, e->site_charge.data(), sizeof(double) * nid, hipMemcpyHostToDevice, e->stream));
   zero_async(e, dv, sizeof(double) * 3 * nsp);
   if (v_host)
      for (int c = 0; c < 3; c++)
         HIPCHK(hipMemcpyAsync(dv + (size_t)c * nsp, v_host + (size_t)c * ns, sizeof(double) * ns,
                               hipMemcpyHostToDevice, e->stream));
   zero_async(e, dj, sizeof(double) * 3 * nsp);
   // sites and cells of force_calc
   launch_make_sites(e, e->molpbc ? 1 : 0);
   zero_async(e, e->d_iflags.p, sizeof(int) * 48);
   run_binning(e, false);
   AdjArgs a;
   a.nsites = ns; a.ncells = e->ncells; a.nx = e->nx; a.ny = e->ny; a.nz = e->nz; a.nfull = e->n_full;
   a.max_id = nid; a.npar = npar; a.nsp = nsp; a.strict_sub = e->strict && !e->molpbc; a.rs = rs;
   a.alpha = e->alpha; a.norm = 2.0 * e->alpha / sqrt(PI_D);
   a.cutsq = e->cutoff * e->cutoff; a.cut100sq = 10000.0 * a.cutsq;
   int k = 0;
   for (int ii = -1; ii <= 1; ii++)
      for (int jj = -1; jj <= 1; jj++)
         for (int kk = -1; kk <= 1; kk++, k++) {
            a.reloc[3 * k] = e->h[0] * ii + e->h[1] * jj + e->h[2] * kk;
            a.reloc[3 * k + 1] = e->h[3] * ii + e->h[4] * jj + e->h[5] * kk;
            a.reloc[3 * k + 2] = e->h[6] * ii + e->h[7] * jj + e->h[8] * kk;
         }
   a.css = e->d_cell_site_start.p; a.off = e->d_off.p; a.cpos = e->d_cpos.p; a.cmeta = e->d_cmeta.p;
   a.corig = e->d_corig.p; a.potpar = e->d_potpar.p;
   a.v = v_host ? dv : nullptr;
   a.dpar = ddp; a.dq = ddq; a.chg_id = dqid; a.rec = rec;
   a.jvp = (jvp_host && dtheta) ? dj : nullptr;
   const size_t shm = sizeof(double) * 64 * (size_t)rs;
   if (shm > 160 * 1024) throw StatusError{MGPU_E_CAPACITY, "too many site ids for the adjoint"};
   {
      ProfScope ps(e, "adjoint_real");
      dim3 g((ns + 63) / 64);
      if (e->alpha > 0.0) launch_adj_pt<true>(e->ptype, g, shm, e->stream, a);
      else launch_adj_pt<false>(e->ptype, g, shm, e->stream, a);
   }
   k_adj_fold<<<(ntheta + 63) / 64, 64, 0, e->stream>>>(ns, nid, npar, rs, 0, e->d_site_type.p, rec, dth);
   k_adj_fold<<<(ntheta + 63) / 64, 64, 0, e->stream>>>(ns, nid, npar, rs, rs / 2, e->d_site_type.p, rec,
                                                        dth + ntheta);
   std::vector<double> h(2 * (size_t)ntheta);
   HIPCHK(hipMemcpyAsync(h.data(), dth, sizeof(double) * 2 * ntheta, hipMemcpyDeviceToHost, e->stream));
   int fl[41];
   HIPCHK(hipMemcpyAsync(fl, e->d_iflags.p, sizeof fl, hipMemcpyDeviceToHost, e->stream));
   HIPCHK(hipStreamSynchronize(e->stream));
   if (fl[0] & FLAG_BIN_RANGE) throw StatusError{MGPU_E_BIN_RANGE, "Co-ordinate out of range in BIN (fill_cells)"};
   if (fl[0] & FLAG_BIN_ROUND) throw StatusError{MGPU_E_BIN_ROUND, "Rounding problem in BIN (fill_cells)"};
   std::vector<double> ge(h.begin(), h.begin() + ntheta), gv(h.begin() + ntheta, h.end());
   if (e->alpha > 1e-7) {
      std::vector<double> qe(nid, 0.0), qv(nid, 0.0);
      ewald_charge_adjoint(e, v_host ? dv : nullptr, ddq, qe.data(), qv.data(), a.jvp ? dj : nullptr);
      for (int q = 0; q < nid; q++) {
         ge[ntypes * npar + q] += qe[q];
         gv[ntypes * npar + q] += qv[q];
      }
   }
   static_grad(e, npar, ge.data());
   if (dE) std::copy(ge.begin(), ge.end(), dE);
   if (vjp) std::copy(gv.begin(), gv.end(), vjp);
   if (jvp_host) {
      if (!dtheta) {
         std::fill(jvp_host, jvp_host + 3 * (size_t)ns, 0.0);
      } else {
         for (int c = 0; c < 3; c++)
            HIPCHK(hipMemcpyAsync(jvp_host + (size_t)c * ns, dj + (size_t)c * nsp, sizeof(double) * ns,
                                  hipMemcpyDeviceToHost, e->stream));
         HIPCHK(hipStreamSynchronize(e->stream));
      }
   }
}
