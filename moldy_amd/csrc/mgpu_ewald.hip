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

// Shared setup: trig tables (FP64) for the current sites.
static void ewald_tables(mgpu_engine *e) {
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
       t, e->d_site.p, e->d_chg_site.p, e->d_tx.p, e->d_ty.p, e->d_tz.p,
       e->d_tz.p + (size_t)(e->lmax + 1) * nsp);
}

void mgpu::run_ewald_fp64(mgpu_engine *e) {
   ewald_tables(e);
   const size_t nsp = e->nsp;
   const double2 *tz1 = e->d_tz.p + (size_t)(e->lmax + 1) * nsp;
   // k groups: <= 256 k vectors, <= KROWS rows each, and row partitions of
   // the force pass with about equal numbers of k vectors; both are host
   // tables uploaded when the k set changes
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
      const int npart = std::max(1, std::min(16, e->nrows / 8));
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
      k_add_parts<<<(e->nsites + 255) / 256, 256, 0, e->stream>>>(e->nsites, e->nsp, npart, e->d_fpart.p,
                                                                  e->d_force_site.p);
   }
}
