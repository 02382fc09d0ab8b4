// Real-space site-site forces (the work of force_inner + kernel + mk_forces,
// src/force.c:784-1005, src/kernel.c:157-463).
//
// One wavefront owns one link cell and gathers, for each of its sites, the
// interactions with every site in the full neighbour shell (the reference's
// half-shell offsets plus their mirror images), so each unordered pair is
// evaluated from both ends and no atomics are needed.  The pair set is the
// reference's: pair (i,j) is present iff j lies in a cell offset by a
// half-shell vector from i's cell (or the reverse), j != i, and not both are
// framework sites (src/force.c:541-566, 904-912).  Energies and the virial
// are accumulated once per visit with weight 1/2; the virial is the pair form
// sum f r_a r_b, which equals the reference's sum r_i F_i + sum R_n g_n
// (src/force.c:973-997) when site_force starts at zero.
//
// j sites are enumerated virtually: lanes take 64 neighbour cells at a time,
// a wave prefix sum over their site counts numbers the candidate j sites, and
// each lane locates its j by a binary search over that prefix (ds_bpermute).
// Up to MGPU_IB i sites are register-blocked; forces on i are reduced with
// wavefront shuffles at the end of each block.
#include "mgpu_internal.h"
#include "mgpu_pair.h"

using namespace mgpu;

namespace {

struct RSArgs {
   int ncells, nx, ny, nz, nfull, max_id, nsp, strict_sub;
   int skip_intra_u;   // FP32: drop minimum-image intramolecular pair energies
   double alpha, norm, cutsq, cut100sq, half_box_sq;
   double reloc[81];
   const int *css;
   const int4 *off;
   const float4 *offD;
   const double4 *cpos;
   const float4 *cposf;
   const int2 *cmeta;
   const int *corig;
   const double *potpar;
   double *force;
   double *part;
   int *tooclose;
};

template <class R> struct Pos;
template <> struct Pos<double> {
   static __device__ __forceinline__ void load(const RSArgs &a, int j, double &x, double &y,
                                               double &z, double &q) {
      double4 v = a.cpos[j];
      x = v.x; y = v.y; z = v.z; q = v.w;
   }
};
template <> struct Pos<float> {
   static __device__ __forceinline__ void load(const RSArgs &a, int j, float &x, float &y,
                                               float &z, float &q) {
      float4 v = a.cposf[j];
      x = v.x; y = v.y; z = v.z; q = v.w;
   }
};

template <class R, int PT, bool COUL>
__global__ __launch_bounds__(256) void k_realspace(RSArgs a) {
   __shared__ R spot[MGPU_MAX_ID * MGPU_MAX_ID * MGPU_NPOTP];
   __shared__ double sreloc[81];
   const int npp = a.max_id * a.max_id * MGPU_NPOTP;
   for (int t = threadIdx.x; t < npp; t += blockDim.x) spot[t] = (R)a.potpar[t];
   for (int t = threadIdx.x; t < 81; t += blockDim.x) sreloc[t] = a.reloc[t];
   __syncthreads();

   const int lane = threadIdx.x & 63;
   const int cell = blockIdx.x * 4 + (threadIdx.x >> 6);
   if (cell >= a.ncells) return;
   const int ix = cell / (a.ny * a.nz), iy = cell / a.nz - a.ny * ix,
             iz = cell - a.nz * (iy + a.ny * ix);
   const int cs = a.css[cell], ce = a.css[cell + 1];
   double epot = 0, v00 = 0, v01 = 0, v02 = 0, v11 = 0, v12 = 0, v22 = 0;
   const R norm = (R)a.norm, alpha = (R)a.alpha;

   for (int ib = cs; ib < ce; ib += MGPU_IB) {
      const int ni = min(MGPU_IB, ce - ib);
      R xi[MGPU_IB], yi[MGPU_IB], zi[MGPU_IB], qi[MGPU_IB];
      int ti[MGPU_IB], mi[MGPU_IB], fwi[MGPU_IB];
      R fxa[MGPU_IB], fya[MGPU_IB], fza[MGPU_IB];
#pragma unroll
      for (int k = 0; k < MGPU_IB; k++) {
         int ii = ib + (k < ni ? k : 0);
         Pos<R>::load(a, ii, xi[k], yi[k], zi[k], qi[k]);
         int2 m = a.cmeta[ii];
         ti[k] = m.x & 0xff;
         fwi[k] = m.x >> 8;
         mi[k] = m.y;
         fxa[k] = fya[k] = fza[k] = 0;
      }
      for (int o0 = 0; o0 < a.nfull; o0 += 64) {
         const int o = o0 + lane;
         int cnt = 0, jst = 0, code = 13;
         if (o < a.nfull) {
            int4 d = a.off[o];
            int ex = ix + d.x, ey = iy + d.y, ez = iz + d.z;
            int sx = ex < 0 ? -1 : (ex >= a.nx ? 1 : 0);
            int sy = ey < 0 ? -1 : (ey >= a.ny ? 1 : 0);
            int sz = ez < 0 ? -1 : (ez >= a.nz ? 1 : 0);
            int jc = (ez - sz * a.nz) + a.nz * ((ey - sy * a.ny) + a.ny * (ex - sx * a.nx));
            jst = a.css[jc];
            cnt = a.css[jc + 1] - jst;
            code = 9 * (sx + 1) + 3 * (sy + 1) + (sz + 1);
         }
         const int inc = wave_incl_scan(cnt, lane);
         const int tot = __shfl(inc, 63, 64);
         R ep = 0, w00 = 0, w01 = 0, w02 = 0, w11 = 0, w12 = 0, w22 = 0;
         for (int t0 = 0; t0 < tot; t0 += 64) {
            const int kq = t0 + lane;
            const bool active = kq < tot;
            int lo = 0, hi = 63;
#pragma unroll
            for (int s = 0; s < 6; s++) {
               int mid = (lo + hi) >> 1;
               int vm = __shfl(inc, mid, 64);
               if (vm > kq) hi = mid; else lo = mid + 1;
            }
            const int L = lo;
            const int incL = __shfl(inc, L, 64), cntL = __shfl(cnt, L, 64);
            const int jstL = __shfl(jst, L, 64), codeL = __shfl(code, L, 64);
            int j = jstL + (kq - (incL - cntL));
            if (!active) j = cs;
            const int oi = min(o0 + L, a.nfull - 1);   // inactive lanes: any valid offset
            R xj, yj, zj, qj;
            Pos<R>::load(a, j, xj, yj, zj, qj);
            const int2 mj = a.cmeta[j];
            const int tj = mj.x & 0xff, fwj = mj.x >> 8;
            R shx, shy, shz;
            if (sizeof(R) == 8) {
               shx = (R)sreloc[3 * codeL]; shy = (R)sreloc[3 * codeL + 1]; shz = (R)sreloc[3 * codeL + 2];
            } else {
               float4 D = a.offD[oi];
               shx = D.x; shy = D.y; shz = D.z;
            }
#pragma unroll
            for (int k = 0; k < MGPU_IB; k++) {
               if (k >= ni) break;
               const bool skip = !active || (oi == 0 && j == ib + k) || (fwi[k] & fwj);
               R rx = xj - xi[k] + shx, ry = yj - yi[k] + shy, rz = zj - zi[k] + shz;
               R rsq = rx * rx + ry * ry + rz * rz;
               if (!skip && rsq < R(0.25) && mi[k] != mj.y) {
                  int oiorig = a.corig[ib + k], ojorig = a.corig[j];
                  if (oiorig < ojorig) {
                     int slot = atomicAdd(a.tooclose, 1);
                     if (slot < 16) {
                        a.tooclose[1 + 2 * slot] = oiorig;
                        a.tooclose[2 + 2 * slot] = ojorig;
                     }
                  }
               }
               if (a.strict_sub && rsq > (R)a.cutsq) rsq = (R)a.cut100sq;
               if (skip) rsq = R(1);
               R u, f;
               pair_term<R, PT, COUL>(rsq, qi[k] * qj, norm, alpha,
                                      &spot[(ti[k] * a.max_id + tj) * MGPU_NPOTP], u, f);
               if (skip) { u = 0; f = 0; }
               // The minimum-image intramolecular pairs of rigid molecules sum
               // to eintra, which the FP64 path subtracts again; the FP32 path
               // leaves both out so their large cancelling energies never
               // meet FP32 rounding.  Forces are kept.
               if (sizeof(R) == 4 && a.skip_intra_u && mi[k] == mj.y && rsq < (R)a.half_box_sq) u = 0;
               fxa[k] -= f * rx;
               fya[k] -= f * ry;
               fza[k] -= f * rz;
               ep += u;
               w00 += f * rx * rx; w01 += f * rx * ry; w02 += f * rx * rz;
               w11 += f * ry * ry; w12 += f * ry * rz; w22 += f * rz * rz;
            }
         }
         epot += (double)ep;
         v00 += (double)w00; v01 += (double)w01; v02 += (double)w02;
         v11 += (double)w11; v12 += (double)w12; v22 += (double)w22;
      }
#pragma unroll
      for (int k = 0; k < MGPU_IB; k++) {
         if (k >= ni) break;
         double gx = wave_sum((double)fxa[k]);
         double gy = wave_sum((double)fya[k]);
         double gz = wave_sum((double)fza[k]);
         if (lane == 0) {
            int io = a.corig[ib + k];
            a.force[io] += gx;
            a.force[a.nsp + io] += gy;
            a.force[2 * a.nsp + io] += gz;
         }
      }
   }
   epot = wave_sum(epot);
   v00 = wave_sum(v00); v01 = wave_sum(v01); v02 = wave_sum(v02);
   v11 = wave_sum(v11); v12 = wave_sum(v12); v22 = wave_sum(v22);
   if (lane == 0) {
      double *p = a.part + 8 * (size_t)cell;
      p[0] = 0.5 * epot;
      p[1] = 0.5 * v00; p[2] = 0.5 * v01; p[3] = 0.5 * v02;
      p[4] = 0.5 * v11; p[5] = 0.5 * v12; p[6] = 0.5 * v22;
      p[7] = 0;
   }
}

// Deterministic reduction of per-cell partials: out[c] = sum over rows.
__global__ void k_reduce_rows(const double *__restrict__ in, int nrows, int ncols,
                              double *__restrict__ out) {
   // one block per column
   __shared__ double ws[16];
   int col = blockIdx.x;
   double s = 0;
   for (int r = threadIdx.x; r < nrows; r += blockDim.x) s += in[(size_t)r * ncols + col];
   s = wave_sum(s);
   if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
   __syncthreads();
   if (threadIdx.x == 0) {
      double t = 0;
      for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += ws[w];
      out[col] = t;
   }
}

template <class R, bool COUL>
void launch_pt(int pt, dim3 g, dim3 b, hipStream_t s, const RSArgs &a) {
   switch (pt) {
   case 0: k_realspace<R, 0, COUL><<<g, b, 0, s>>>(a); break;
   case 1: k_realspace<R, 1, COUL><<<g, b, 0, s>>>(a); break;
   case 2: k_realspace<R, 2, COUL><<<g, b, 0, s>>>(a); break;
   case 3: k_realspace<R, 3, COUL><<<g, b, 0, s>>>(a); break;
   case 4: k_realspace<R, 4, COUL><<<g, b, 0, s>>>(a); break;
   case 6: k_realspace<R, 6, COUL><<<g, b, 0, s>>>(a); break;
   default: throw StatusError{MGPU_E_PTYPE, "unknown potential type"};
   }
}

}  // namespace

void mgpu::run_realspace(mgpu_engine *e) {
   int nc = e->ncells;
   e->d_cell_part.reserve((size_t)nc * 8 + 8);
   RSArgs a;
   a.ncells = nc; a.nx = e->nx; a.ny = e->ny; a.nz = e->nz; a.nfull = e->n_full;
   a.max_id = e->max_id; a.nsp = e->nsp;
   a.strict_sub = e->strict && !e->molpbc;
   a.alpha = e->alpha;
   a.norm = 2.0 * e->alpha / sqrt(3.14159265358979323846);
   a.cutsq = e->cutoff * e->cutoff;
   a.cut100sq = 10000.0 * a.cutsq;
   a.skip_intra_u = e->precision == MGPU_MIXED && e->intra_exact;
   a.half_box_sq = e->half_box * e->half_box;
   int k = 0;
   for (int ii = -1; ii <= 1; ii++)
      for (int jj = -1; jj <= 1; jj++)
         for (int kk = -1; kk <= 1; kk++, k++) {
            // reloc_v, src/force.c:1284-1293
            a.reloc[3 * k] = e->h[0] * ii + e->h[1] * jj + e->h[2] * kk;
            a.reloc[3 * k + 1] = e->h[3] * ii + e->h[4] * jj + e->h[5] * kk;
            a.reloc[3 * k + 2] = e->h[6] * ii + e->h[7] * jj + e->h[8] * kk;
         }
   a.css = e->d_cell_site_start.p;
   a.off = e->d_off.p;
   a.offD = e->d_offD.p;
   a.cpos = e->d_cpos.p;
   a.cposf = e->d_cposf.p;
   a.cmeta = e->d_cmeta.p;
   a.corig = e->d_corig.p;
   a.potpar = e->d_potpar.p;
   a.force = e->d_force_site.p;
   a.part = e->d_cell_part.p;
   a.tooclose = e->d_iflags.p + 1;
   dim3 g((nc + 3) / 4), b(256);
   bool coul = e->alpha > 0.0;
   {
      ProfScope ps(e, "realspace");
      if (e->precision == MGPU_FP64) {
         if (coul) launch_pt<double, true>(e->ptype, g, b, e->stream, a);
         else launch_pt<double, false>(e->ptype, g, b, e->stream, a);
      } else {
         if (coul) launch_pt<float, true>(e->ptype, g, b, e->stream, a);
         else launch_pt<float, false>(e->ptype, g, b, e->stream, a);
      }
   }
   e->d_out.reserve(OUT_SIZE);
   reduce_cols(e, e->d_cell_part.p, nc, 8, e->d_out.p + OUT_REAL);
}
