/*
 * moldy_oracle.c -- TEST INFRASTRUCTURE ONLY (see moldy_oracle.h).
 *
 * A CPU restatement of the reference algorithm, written for this project and
 * compiled strict-IEEE (-O2 -fno-fast-math -ffp-contract=off) so that the
 * binning arithmetic reproduces the reference's operation order bit for bit.
 * It is the parity checker for the HIP engine and the CPU baseline timed by
 * bench.py.  Nothing in the product imports or links it.
 *
 * Citations are reference-tree paths (src/<file>:<line>).
 */
#include "moldy_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

#define PI       3.14159265358979323846   /* src/defs.h:183 */
#define SQR(x)   ((x) * (x))
#define CUBE(x)  ((x) * (x) * (x))
#define ALPHAMIN 1e-7                     /* src/defs.h:181 */
#define TOO_CLOSE 0.25                    /* src/force.c:89 */
#define NMULT 2.0                         /* src/force.c:88 */
#define EPS0 (0.25 / PI)                  /* src/defs.h:229 */
#define BPAR_TOL 1.0e-7                   /* src/kernel.c:97 */

/* A&S 7.1.26 coefficients, src/kernel.c:88-96 and src/auxil.c:577-585 */
static const double AS_P  = 0.3275911;
static const double AS_A1 = 0.254829592, AS_A2 = -0.284496736,
                    AS_A3 = 1.421413741, AS_A4 = -1.453152027,
                    AS_A5 = 1.061405429;
static inline double as_poly(double t)
{
   return t * (AS_A1 + t * (AS_A2 + t * (AS_A3 + t * (AS_A4 + t * AS_A5))));
}

/* =========================== numeric helpers ============================ */

/* det, src/matrix.c:159-168: sum over cyclic (i,j,k) of a0i(a1j a2k - a1k a2j) */
double orc_det(const double a[9])
{
   double d = 0.0;
   int i, j, k;
   for (i = 0, j = 1, k = 2; i < 3; i++, j = (j + 1) % 3, k = (k + 1) % 3)
      d += a[i] * (a[3 + j] * a[6 + k] - a[3 + k] * a[6 + j]);
   return d;
}

/* invert (adjoint method), src/matrix.c:171-185 */
void orc_invert(const double a[9], double b[9])
{
   double r = 1.0 / orc_det(a);
   int i, j, k, l, m, n;
   for (i = 0, j = 1, k = 2; i < 3; i++, j = (j + 1) % 3, k = (k + 1) % 3)
      for (l = 0, m = 1, n = 2; l < 3; l++, m = (m + 1) % 3, n = (n + 1) % 3)
         b[3 * l + i] = r * (a[3 * j + m] * a[3 * k + n] - a[3 * j + n] * a[3 * k + m]);
}

static void transpose3(const double a[9], double b[9])
{
   double t[9];
   int i, j;
   memcpy(t, a, sizeof t);
   for (i = 0; i < 3; i++)
      for (j = 0; j < 3; j++)
         b[3 * i + j] = t[3 * j + i];
}

/* mat_mul, src/matrix.c:94-115 (each element summed left to right) */
static void matmul3(const double a[9], const double b[9], double c[9])
{
   double r[9];
   int i, j;
   for (i = 0; i < 3; i++)
      for (j = 0; j < 3; j++)
         r[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
   memcpy(c, r, sizeof r);
}

/* mat_vec_mul for one vector, src/matrix.c:46-90 */
static inline void matvec3(const double m[9], const double in[3], double out[3])
{
   double a0 = in[0], a1 = in[1], a2 = in[2];
   out[0] = m[0] * a0 + m[1] * a1 + m[2] * a2;
   out[1] = m[3] * a0 + m[4] * a1 + m[5] * a2;
   out[2] = m[6] * a0 + m[7] * a1 + m[8] * a2;
}

/* q_to_rot, src/quaterns.c:129-154 */
static void q_to_rot(const double *q, double rot[9])
{
   double q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
   double a01 = 2.0 * q0 * q1, a02 = 2.0 * q0 * q2, a03 = 2.0 * q0 * q3;
   double a12 = 2.0 * q1 * q2, a13 = 2.0 * q1 * q3, a23 = 2.0 * q2 * q3;
   rot[1] = a12 - a03;  rot[2] = a13 + a02;
   rot[3] = a12 + a03;  rot[5] = a23 - a01;
   rot[6] = a13 - a02;  rot[7] = a23 + a01;
   q0 = q0 * q0; q1 = q1 * q1; q2 = q2 * q2; q3 = q3 * q3;
   rot[0] = q0 + q1 - q2 - q3;
   rot[4] = q0 - q1 + q2 - q3;
   rot[8] = q0 - q1 - q2 + q3;
}

/* err_fn, src/auxil.c:586-593 (A&S 7.1.26 erf) */
double orc_err_fn(double x)
{
   double t = 1.0 / (1.0 + AS_P * x);
   if (x >= 0.0)
      return 1.0 - as_poly(t) * exp(-x * x);
   return -orc_err_fn(-x);
}

/* precision(), src/auxil.c:339-357: smallest eps with 1+eps > 1 */
double orc_precision(void)
{
   volatile double eps = 0.5, one_plus;
   do {
      eps = eps * 0.5;
      one_plus = 1.0 + eps;
   } while (one_plus > 1.0);
   return eps * 2.0;
}

/* sum(), src/auxil.c:222-256: four interleaved partial sums, delayed by one
 * stride, remainder into the first, combined as ((s1+s2)+s3)+s4. */
double orc_sum(int n, const double *x)
{
   double sa = 0.0, sb = 0.0, sc = 0.0, sd = 0.0;
   double s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0;
   int i;
   for (i = 0; i < n - 3; i += 4) {
      s1 += sa; sa = x[i];
      s2 += sb; sb = x[i + 1];
      s3 += sc; sc = x[i + 2];
      s4 += sd; sd = x[i + 3];
   }
   s1 += sa; s2 += sb; s3 += sc; s4 += sd;
   for (; i < n; i++)
      s1 += x[i];
   return s1 + s2 + s3 + s4;
}

static double vdot_s(int n, const double *x, int ix, const double *y, int iy)
{
   double d = 0.0;
   int i, j;
   for (i = 0, j = 0; i < n; i++, j++)
      d += x[i * ix] * y[j * iy];
   return d;
}

/* ============================== binning ================================= */

/* cellbin, src/force.c:119-137 */
int orc_cellbin(double rc, int nc, double fnc, double eps, int *status)
{
   int ibin;
   if (rc < -0.5 + eps || rc >= 0.5 - eps) {
      if (rc < -0.5 + eps && rc >= -0.5 - eps)
         rc = -0.5;
      else if (rc >= 0.5 - eps && rc <= 0.5 + eps)
         rc = 0.5 - eps;
      else if (status && !*status)
         *status = ORC_ERR_BIN_RANGE;
   }
   ibin = (int)floor((rc + 0.5) * fnc);
   if (ibin >= nc || ibin < 0) {
      if (status && !*status)
         *status = ORC_ERR_BIN_ROUND;
      ibin = ibin < 0 ? 0 : nc - 1;   /* keep the index inside the array */
   }
   return ibin;
}

/* fill_cells, src/force.c:426-486.  The reference pushes each entry onto the
 * head of its cell's linked list, so a cell's list runs in decreasing entry
 * order; the CSR written here keeps exactly that order. */
int orc_fill_cells(const orc_sys *s, const double h[9],
                   const double *x, const double *y, const double *z,
                   const double *c_of_m_frac, int nx, int ny, int nz,
                   int *entry_cell, int *entry_isite, int *entry_num,
                   int *entry_ftype, int *cell_start, int *cell_entries,
                   int *n_frame_types, int *status)
{
   double eps = 8.0 * orc_precision();
   double fnx = nx, fny = ny, fnz = nz;
   double hinv[9], ss[3];
   int ncells = nx * ny * nz;
   int ispec, im, is, isite = 0, imol = 0, ne = 0, frame_type = 1, c, e;
   int *fill;

   orc_invert(h, hinv);
   for (ispec = 0; ispec < s->nspecies; ispec++) {
      int ns = s->spec_nsites[ispec];
      for (im = 0; im < s->spec_nmols[ispec]; im++, imol++) {
         if (s->spec_framework[ispec])
            frame_type++;
         if (!s->molpbc || s->spec_framework[ispec]) {
            for (is = 0; is < ns; is++, isite++) {
               double r[3] = { x[isite], y[isite], z[isite] };
               matvec3(hinv, r, ss);
               c = orc_cellbin(ss[2], nz, fnz, eps, status)
                 + nz * (orc_cellbin(ss[1], ny, fny, eps, status)
                 + ny * orc_cellbin(ss[0], nx, fnx, eps, status));
               entry_cell[ne] = c; entry_isite[ne] = isite;
               entry_num[ne] = 1;  entry_ftype[ne] = frame_type - 1;
               ne++;
            }
         } else {
            const double *cm = c_of_m_frac + 3 * imol;
            c = orc_cellbin(cm[2], nz, fnz, eps, status)
              + nz * (orc_cellbin(cm[1], ny, fny, eps, status)
              + ny * orc_cellbin(cm[0], nx, fnx, eps, status));
            entry_cell[ne] = c; entry_isite[ne] = isite;
            entry_num[ne] = ns; entry_ftype[ne] = 0;
            ne++;
            isite += ns;
         }
      }
   }
   *n_frame_types = frame_type;
   /* CSR in linked-list order (descending entry index within a cell) */
   memset(cell_start, 0, (size_t)(ncells + 1) * sizeof(int));
   for (e = 0; e < ne; e++)
      cell_start[entry_cell[e] + 1]++;
   for (c = 0; c < ncells; c++)
      cell_start[c + 1] += cell_start[c];
   fill = (int *)malloc((size_t)ncells * sizeof(int));
   if (!fill)
      return -ORC_ERR_ALLOC;
   memcpy(fill, cell_start, (size_t)ncells * sizeof(int));
   for (e = ne - 1; e >= 0; e--)
      cell_entries[fill[entry_cell[e]]++] = e;
   free(fill);
   return ne;
}

/* =========================== neighbour cells ============================ */

static double col_norm(const double m[9], int col)
{
   return sqrt(SQR(m[col]) + SQR(m[3 + col]) + SQR(m[6 + col]));
}

/* neighbour_list, src/force.c:167-226 (cell-centre criterion) and
 * strict_neighbour_list, src/force.c:273-421 (corner/face criterion). */
int orc_neighbour_list(const double h[9], double cutoff, int nx, int ny, int nz,
                       int strict, int *offsets, int cap)
{
   double G[9], htr[9], htrinv[9], sv[3], dist;
   int mx, my, mz, ix, iy, iz, i, j, k, n = 0;

   transpose3(h, htr);
   matmul3(htr, h, G);
   orc_invert(htr, htrinv);
   mx = (int)ceil(cutoff * nx * col_norm(htrinv, 0));
   my = (int)ceil(cutoff * ny * col_norm(htrinv, 1));
   mz = (int)ceil(cutoff * nz * col_norm(htrinv, 2));

   if (!strict) {
      for (ix = 0; ix < mx; ix++)
         for (iy = (ix == 0 ? 0 : -my); iy < my; iy++)
            for (iz = (ix == 0 && iy == 0 ? 0 : -mz); iz < mz; iz++) {
               sv[0] = (double)ix / nx; sv[1] = (double)iy / ny; sv[2] = (double)iz / nz;
               dist = 0.0;
               for (i = 0; i < 3; i++)
                  for (j = 0; j < 3; j++)
                     dist += sv[i] * G[3 * i + j] * sv[j];
               if (dist < SQR(cutoff)) {
                  if (ix > nx || ix < -nx || iy > ny || iy < -ny || iz > nz || iz < -nz)
                     return -ORC_ERR_CUTOFF;
                  if (n >= cap)
                     return -ORC_ERR_ALLOC;
                  offsets[3 * n] = ix; offsets[3 * n + 1] = iy; offsets[3 * n + 2] = iz;
                  n++;
               }
            }
      return n;
   } else {
      /* map over ix in [0,mx], iy in [-my-1,my], iz in [-mz-1,mz] */
      int sy = 2 * my + 2, sz = 2 * mz + 2;
      unsigned char *map = (unsigned char *)calloc((size_t)(mx + 1) * sy * sz, 1);
      int nxyz[3], mxyz[3], fc[4][3], a, b, c;
      double proj[3], modabc;
#define MAP(a, b, c) map[((size_t)(a) * sy + ((b) + my + 1)) * sz + ((c) + mz + 1)]
      if (!map)
         return -ORC_ERR_ALLOC;
      for (ix = 0; ix < mx; ix++)
         for (iy = (ix == 0 ? 0 : -my); iy < my; iy++)
            for (iz = (ix == 0 && iy == 0 ? 0 : -mz); iz < mz; iz++) {
               sv[0] = (double)ix / nx; sv[1] = (double)iy / ny; sv[2] = (double)iz / nz;
               dist = 0.0;
               for (i = 0; i < 3; i++)
                  for (j = 0; j < 3; j++)
                     dist += sv[i] * G[3 * i + j] * sv[j];
               if (dist < SQR(cutoff))
                  for (i = 0; i <= 1; i++)
                     for (j = -1; j <= 1; j++)
                        for (k = -1; k <= 1; k++)
                           MAP(ix + i, iy + j, iz + k) = 1;
            }
      nxyz[0] = nx; nxyz[1] = ny; nxyz[2] = nz;
      mxyz[0] = mx; mxyz[1] = my; mxyz[2] = mz;
      for (a = 0; a < 3; a++) {          /* face-face contacts, force.c:338-383 */
         b = (a + 1) % 3; c = (b + 1) % 3;
         proj[0] = proj[1] = proj[2] = 0.0;
         modabc = 0.0;
         for (i = 0; i < 3; i++) {
            modabc += htrinv[3 * i + a];
            proj[i] += htrinv[3 * i + a] * htrinv[3 * i + (a + i) % 3];
         }
         for (i = 0; i < 3; i++)
            proj[i] *= (mxyz[a] - 1) * nxyz[i] / (nxyz[a] * modabc);
         for (i = 0; i < 3; i++)
            fc[0][i] = fc[1][i] = fc[2][i] = fc[3][i] = (int)floor(proj[i]);
         fc[0][a] = fc[1][a] = fc[2][a] = fc[3][a] = mxyz[a];
         fc[1][b] = fc[3][b] = (int)ceil(proj[b]);
         fc[2][c] = fc[3][c] = (int)ceil(proj[c]);
         for (i = 0; i < 4; i++)
            if (fc[i][0] < 0)
               for (j = 0; j < 3; j++)
                  fc[i][j] = -fc[i][j];
         for (i = 0; i < 4; i++) {
            if (fc[i][0] < 0 || fc[i][0] > mx || fc[i][1] < -my - 1 || fc[i][1] > my ||
                fc[i][2] < -mz - 1 || fc[i][2] > mz) {
               free(map);
               return -ORC_ERR_MAP;
            }
            MAP(fc[i][0], fc[i][1], fc[i][2]) = 1;
         }
      }
      for (ix = 0; ix <= mx; ix++)
         for (iy = (ix == 0 ? 0 : -my - 1); iy <= my; iy++)
            for (iz = (ix == 0 && iy == 0 ? 0 : -mz - 1); iz <= mz; iz++)
               if (MAP(ix, iy, iz)) {
                  if (ix > nx || ix < -nx || iy > ny || iy < -ny || iz > nz || iz < -nz) {
                     free(map);
                     return -ORC_ERR_CUTOFF;
                  }
                  if (n >= cap) {
                     free(map);
                     return -ORC_ERR_ALLOC;
                  }
                  offsets[3 * n] = ix; offsets[3 * n + 1] = iy; offsets[3 * n + 2] = iz;
                  n++;
               }
#undef MAP
      free(map);
      return n;
   }
}

/* src/force.c:1153-1157 */
void orc_choose_cells(const orc_sys *s, const double h[9], int *nx, int *ny, int *nz)
{
   double subcell = s->subcell;
   if (subcell <= 0.0)
      subcell = s->cutoff / 5.0;
   *nx = (int)(h[0] / subcell + 0.5);
   *ny = (int)(h[4] / subcell + 0.5);
   *nz = (int)(h[8] / subcell + 0.5);
}

/* ======================== site neighbour lists ========================== */

typedef struct {
   int nx, ny, nz, ncells, ne, n_frame_types, n_nabors;
   int *entry_cell, *entry_isite, *entry_num, *entry_ftype;
   int *cell_start, *cell_entries;
   int *offsets;
} cells_t;

static void cells_free(cells_t *c)
{
   free(c->entry_cell); free(c->entry_isite); free(c->entry_num);
   free(c->entry_ftype); free(c->cell_start); free(c->cell_entries);
   free(c->offsets);
   memset(c, 0, sizeof *c);
}

static int cells_build(const orc_sys *s, const double h[9], const double *x,
                       const double *y, const double *z, const double *cm,
                       int strict, double list_cutoff, cells_t *c, int *status)
{
   int cap;
   memset(c, 0, sizeof *c);
   orc_choose_cells(s, h, &c->nx, &c->ny, &c->nz);
   c->ncells = c->nx * c->ny * c->nz;
   c->entry_cell = (int *)malloc(sizeof(int) * (size_t)(s->nsites + 1));
   c->entry_isite = (int *)malloc(sizeof(int) * (size_t)(s->nsites + 1));
   c->entry_num = (int *)malloc(sizeof(int) * (size_t)(s->nsites + 1));
   c->entry_ftype = (int *)malloc(sizeof(int) * (size_t)(s->nsites + 1));
   c->cell_start = (int *)malloc(sizeof(int) * (size_t)(c->ncells + 1));
   c->cell_entries = (int *)malloc(sizeof(int) * (size_t)(s->nsites + 1));
   if (!c->entry_cell || !c->entry_isite || !c->entry_num || !c->entry_ftype ||
       !c->cell_start || !c->cell_entries)
      return -ORC_ERR_ALLOC;
   c->ne = orc_fill_cells(s, h, x, y, z, cm, c->nx, c->ny, c->nz, c->entry_cell,
                          c->entry_isite, c->entry_num, c->entry_ftype,
                          c->cell_start, c->cell_entries, &c->n_frame_types, status);
   if (c->ne < 0)
      return c->ne;
   if (c->n_frame_types > 2)
      return -ORC_ERR_MULTIFW;
   cap = 4 * 64 * 64 * 64;
   c->offsets = (int *)malloc(sizeof(int) * 3 * (size_t)cap);
   if (!c->offsets)
      return -ORC_ERR_ALLOC;
   c->n_nabors = orc_neighbour_list(h, list_cutoff, c->nx, c->ny, c->nz, strict,
                                    c->offsets, cap);
   if (c->n_nabors < 0)
      return c->n_nabors;
   return 0;
}

/* site_neighbour_list, src/force.c:521-569, for reference cell (ix,iy,iz).
 * Image code k = 9*(a+1)+3*(b+1)+(c+1) for a periodic shift (a,b,c) of the
 * neighbour cell, the numbering of pbclookup/reloc_v (force.c:1269-1293). */
static long cell_site_list(const cells_t *c, int ix, int iy, int iz, int *nab,
                           int *pbctrans, int *nfnab, long cap)
{
   long nn = 0;
   int ftype, jn, e;
   for (ftype = 0; ftype < c->n_frame_types; ftype++) {
      for (jn = (ftype == 0 ? 0 : 1); jn < c->n_nabors; jn++) {
         int ex = ix + c->offsets[3 * jn], ey = iy + c->offsets[3 * jn + 1],
             ez = iz + c->offsets[3 * jn + 2];
         int sx = ex < 0 ? -1 : (ex >= c->nx ? 1 : 0);
         int sy = ey < 0 ? -1 : (ey >= c->ny ? 1 : 0);
         int sz = ez < 0 ? -1 : (ez >= c->nz ? 1 : 0);
         int jc = (ez - sz * c->nz) + c->nz * ((ey - sy * c->ny) + c->ny * (ex - sx * c->nx));
         int k = 9 * (sx + 1) + 3 * (sy + 1) + (sz + 1);
         for (e = c->cell_start[jc]; e < c->cell_start[jc + 1]; e++) {
            int en = c->cell_entries[e], js;
            if (c->entry_ftype[en] != ftype)
               continue;
            for (js = 0; js < c->entry_num[en]; js++) {
               if (nn < cap) {
                  nab[nn] = c->entry_isite[en] + js;
                  pbctrans[nn] = k;
               }
               nn++;
            }
         }
      }
      if (nfnab)
         nfnab[ftype] = (int)nn;
   }
   if (nfnab && c->n_frame_types == 1)
      nfnab[1] = (int)nn;
   return nn;
}

long orc_site_lists(const orc_sys *s, const double h[9], const double *x,
                    const double *y, const double *z, const double *cm,
                    int strict_list, double list_cutoff, long *list_start, int *nab,
                    int *pbctrans, int *nfnab, long cap, int *status)
{
   cells_t c;
   long total = 0;
   int rc = cells_build(s, h, x, y, z, cm, strict_list, list_cutoff, &c, status), ic;
   if (rc < 0) {
      cells_free(&c);
      return rc;
   }
   for (ic = 0; ic < c.ncells; ic++) {
      int ix = ic / (c.ny * c.nz), iy = ic / c.nz - c.ny * ix, iz = ic - c.nz * (iy + c.ny * ix);
      long n;
      list_start[ic] = total;
      n = cell_site_list(&c, ix, iy, iz, nab ? nab + (total < cap ? total : cap) : NULL,
                         pbctrans ? pbctrans + (total < cap ? total : cap) : NULL,
                         nfnab ? nfnab + 2 * ic : NULL, total < cap ? cap - total : 0);
      total += n;
   }
   list_start[c.ncells] = total;
   cells_free(&c);
   return total;
}

/* =============================== kernel ================================= */

/* One pair: energy u and forceij = -U'(r)/r.  Restates the per-jsite bodies of
 * kernel(), src/kernel.c:188-354 (alpha > 0) and 362-460 (alpha <= 0). */
static inline int pair_term(int ptype, int coulomb, double rsq, double qq,
                            double norm, double alpha, const double *p,
                            double *u, double *f)
{
   double r, ar, t, et, r_r, r2r, r4r, r6r, r8r, r12r, e1, e2, e3;
   if (coulomb) {
      r = sqrt(rsq);
      ar = alpha * r;
      t = 1.0 / (1.0 + AS_P * ar);
      et = qq * exp(-SQR(ar));
      r_r = 1.0 / r;
      t = as_poly(t) * et * r_r;       /* erfc(ar) q q / r  (A&S)            */
      et = t + norm * et;             /* + 2 alpha/sqrt(pi) q q exp(-a^2r^2) */
      r2r = SQR(r_r);
      switch (ptype) {
      case ORC_LJ:
         r6r = SQR(p[1]) * r2r; r6r = CUBE(r6r); r12r = SQR(r6r);
         *u = t + p[0] * (r12r - r6r);
         *f = r2r * (6.0 * p[0] * (2 * r12r - r6r) + et);
         return 0;
      case ORC_E6:
         e1 = p[1] * exp(-p[2] * r);
         r6r = p[0] * CUBE(r2r);
         *u = t - r6r + e1;
         *f = r2r * (-6.0 * r6r + et) + p[2] * e1 * r_r;
         return 0;
      case ORC_MCY:
         e1 = p[0] * exp(-p[1] * r);
         e2 = -p[2] * exp(-p[3] * r);
         *u = t + e1 + e2;
         *f = (p[1] * e1 + p[3] * e2) * r_r + et * r2r;
         return 0;
      case ORC_GEN:
         e1 = p[0] * exp(-p[1] * r);
         r4r = SQR(r2r); r6r = r2r * r4r; r8r = p[5] * SQR(r4r);
         r12r = p[2] * SQR(r6r); r4r *= p[3]; r6r *= p[4];
         *u = t + e1 + r12r - r4r - r6r - r8r;
         *f = r2r * (12.0 * r12r - 4.0 * r4r - 6.0 * r6r - 8.0 * r8r + et) + p[1] * e1 * r_r;
         return 0;
      case ORC_MORSE:
         e1 = p[0] * exp((p[1] - r) * p[2]);
         r6r = p[3] * CUBE(r2r);
         e2 = p[4] * exp(-2.0 * p[5] * (r - p[6]));
         e3 = -p[4] * 2.0 * exp(-p[5] * (r - p[6]));
         *u = t + e1 - r6r + e2 + e3;
         *f = r2r * (-6.0 * r6r + et) + r_r * (p[2] * e1 + (2.0 * p[5]) * e2 + p[5] * e3);
         return 0;
      case ORC_HIW:
         r4r = SQR(r2r); r6r = r2r * r4r; r12r = p[2] * SQR(r6r);
         r6r *= p[1]; r4r *= p[0];
         *u = t + r4r + r6r + r12r;
         *f = r2r * (4.0 * r4r + 6.0 * r6r + 12.0 * r12r + et);
         return 0;
      default:
         return ORC_ERR_PTYPE;
      }
   } else {
      switch (ptype) {
      case ORC_LJ:
         r2r = 1.0 / rsq;
         r6r = SQR(p[1]) * r2r; r6r = CUBE(r6r); r12r = SQR(r6r);
         *u = p[0] * (r12r - r6r);
         *f = r2r * 6.0 * p[0] * (2 * r12r - r6r);
         return 0;
      case ORC_E6:
         r = sqrt(rsq); r_r = 1.0 / r;
         e1 = p[1] * exp(-p[2] * r);
         r2r = SQR(r_r);
         r6r = p[0] * r2r * r2r * r2r;
         *u = -r6r + e1;
         *f = -r2r * 6.0 * r6r + p[2] * e1 * r_r;
         return 0;
      case ORC_MCY:
         r = sqrt(rsq); r_r = 1.0 / r;
         e1 = p[0] * exp(-p[1] * r);
         e2 = -p[2] * exp(-p[3] * r);
         *u = e1 + e2;
         *f = (p[1] * e1 + p[3] * e2) * r_r;
         return 0;
      case ORC_GEN:
         r = sqrt(rsq); r_r = 1.0 / r;
         e1 = p[0] * exp(-p[1] * r);
         r2r = SQR(r_r); r4r = SQR(r2r); r6r = r2r * r4r; r8r = p[5] * SQR(r4r);
         r12r = p[2] * SQR(r6r); r4r *= p[3]; r6r *= p[4];
         *u = e1 + r12r - r4r - r6r - r8r;
         *f = r2r * (12.0 * r12r - 4.0 * r4r - 6.0 * r6r - 8.0 * r8r) + p[1] * e1 * r_r;
         return 0;
      case ORC_MORSE:
         r = sqrt(rsq); r_r = 1.0 / r; r2r = SQR(r_r);
         e1 = p[0] * exp((p[1] - r) * p[2]);
         r6r = p[3] * CUBE(r2r);
         e2 = p[4] * exp(-2.0 * p[5] * (r - p[6]));
         e3 = -p[4] * 2.0 * exp(-p[5] * (r - p[6]));
         *u = e1 - r6r + e2 + e3;
         *f = r2r * (-6.0 * r6r) + r_r * (p[2] * e1 + (2.0 * p[5]) * e2 + p[5] * e3);
         return 0;
      case ORC_HIW:
         r = sqrt(rsq); r_r = 1.0 / r; r2r = SQR(r_r);
         r4r = SQR(r2r); r6r = r2r * r4r; r12r = SQR(r6r) * p[2];
         r6r *= p[1]; r4r *= p[0];
         *u = r4r + r6r + r12r;
         *f = r2r * (4.0 * r4r + 6.0 * r6r + 12.0 * r12r);
         return 0;
      default:
         return ORC_ERR_PTYPE;
      }
   }
}

/* kernel(), src/kernel.c:157-463: loop over [jmin, nnab), pe accumulated in a
 * local sum then added once (kernel.c:174, 462). */
int orc_kernel(int jmin, int nnab, double *forceij, double *pe, const double *r_sqr,
               const double *nab_chg, double chg, double norm, double alpha,
               int ptype, double *const *pot)
{
   double ppe = 0.0, u = 0.0, f = 0.0, p[ORC_NPOTP];
   int j, k, rc;
   for (j = jmin; j < nnab; j++) {
      for (k = 0; k < ORC_NPOTP; k++)
         p[k] = pot[k] ? pot[k][j] : 0.0;
      rc = pair_term(ptype, alpha > 0.0, r_sqr[j], nab_chg[j] * chg, norm, alpha, p, &u, &f);
      if (rc)
         return rc;
      ppe += u;
      forceij[j] = f;
   }
   *pe += ppe;
   return 0;
}

/* poteval, src/force.c:632-650 */
double orc_poteval(const double *potpar, double r, int ptype, double chgsq, double alpha)
{
   double u = 0.0, f = 0.0;
   double norm = 2.0 * alpha / sqrt(PI);
   pair_term(ptype, alpha > 0.0, r * r, chgsq * 1.0, norm, alpha, potpar, &u, &f);
   return u;
}

/* dist_pot, src/kernel.c:103-151 */
double orc_dist_pot(const double *p, double rc, int ptype)
{
   switch (ptype) {
   case ORC_LJ:
      return p[0] * CUBE(SQR(p[1]) / rc) / 3.0;
   case ORC_E6:
      if (p[2] > BPAR_TOL)
         return p[0] / (3.0 * CUBE(rc)) - p[1] * exp(-p[2] * rc) *
                (SQR(rc) / p[2] + 2 * rc / SQR(p[2]) + 2.0 / CUBE(p[2]));
      return p[0] / (3.0 * CUBE(rc));
   case ORC_MCY:
      if (p[3] > BPAR_TOL)
         return p[2] * (SQR(rc) / p[3] + 2 * rc / SQR(p[3]) + 2.0 / CUBE(p[3])) *
                exp(-p[3] * rc);
      return 0.0;
   case ORC_GEN:
      if (p[1] > BPAR_TOL)
         return -p[0] * exp(-p[1] * rc) * (SQR(rc) / p[1] + 2 * rc / SQR(p[1]) + 2.0 / CUBE(p[1]))
                - p[2] / (9.0 * CUBE(CUBE(rc))) + p[3] / rc + p[4] / (3.0 * CUBE(rc))
                + p[5] / (5.0 * SQR(rc) * CUBE(rc));
      return -p[2] / (9.0 * CUBE(CUBE(rc))) + p[3] / rc + p[4] / (3.0 * CUBE(rc))
             + p[5] / (5.0 * SQR(rc) * CUBE(rc));
   case ORC_MORSE:
      if (p[5] != 0.0)
         return p[3] / (3.0 * CUBE(rc)) + 2.0 * p[4] * (SQR(rc) / p[5] + 2 * rc / SQR(p[5])
                + 2.0 / CUBE(p[5])) * exp(-p[5] * (rc - p[6]));
      return p[3] / (3.0 * CUBE(rc));
   case ORC_HIW:
      return -p[0] / rc - p[1] / CUBE(rc) / 3.0 - p[2] / CUBE(CUBE(rc)) / 9.0;
   default:
      return 0.0;
   }
}

/* ============================== force_calc ============================== */

static int site_of_species(const orc_sys *s, int ispec)
{
   int i, n = 0;
   for (i = 0; i < ispec; i++)
      n += s->spec_nsites[i];
   return n;
}

/* pe_intra over all non-framework species, src/force.c:654-668, 1158-1168 */
static double eintra_total(const orc_sys *s, const double *chg)
{
   double e = 0.0;
   int ispec, isite = 0, a, b;
   for (ispec = 0; ispec < s->nspecies; ispec++) {
      int ns = s->spec_nsites[ispec], off = site_of_species(s, ispec);
      const int *id = s->spec_site_id + off;
      const double *pf = s->spec_p_f_sites + 3 * off;
      if (!s->spec_framework[ispec]) {
         double ei = 0.0;
         for (b = 0; b < ns; b++)
            for (a = b + 1; a < ns; a++) {
               double d = sqrt(SQR(pf[3 * a] - pf[3 * b]) + SQR(pf[3 * a + 1] - pf[3 * b + 1]) +
                               SQR(pf[3 * a + 2] - pf[3 * b + 2]));
               ei += orc_poteval(s->potpar + ORC_NPOTP * (id[b] * s->max_id + id[a]), d,
                                 s->ptype, chg[isite + a] * chg[isite + b], s->alpha);
            }
         e += s->spec_nmols[ispec] * ei;
      }
      isite += s->spec_nmols[ispec] * ns;
   }
   return e;
}

int orc_force_calc(const orc_sys *s, const double h[9], const double *x,
                   const double *y, const double *z, const double *cm,
                   const double *chg, double *fx, double *fy, double *fz,
                   double *pe, double stress[9], orc_fc_info *info)
{
   cells_t c;
   int status = 0, rc, ic, k, ii, jj, kk, ispec, im, is, isite, mol;
   int *molmap = NULL, *id = NULL, *nab = NULL, *pbc = NULL;
   double *nsx = NULL, *nsy = NULL, *nsz = NULL, *nchg = NULL, *fjx = NULL,
          *fjy = NULL, *fjz = NULL, *rx = NULL, *ry = NULL, *rz = NULL,
          *rsq = NULL, *fij = NULL, *npot = NULL;
   double *potv[ORC_NPOTP];
   double reloc[27][3], gforce[27][3];
   double norm = 2.0 * s->alpha / sqrt(PI);
   double cutsq = SQR(s->cutoff), cut100sq = 10000.0 * cutsq;
   double vol = orc_det(h), s00, s01, s02, s11, s12, s22, subcell;
   long n_nab_sites, maxlist = 0, npairs = 0;
   int nfnab[2];
   orc_fc_info loc;
   if (!info)
      info = &loc;
   memset(info, 0, sizeof *info);

   rc = cells_build(s, h, x, y, z, cm, s->strict_cutoff, s->cutoff, &c, &status);
   if (rc < 0) {
      cells_free(&c);
      return -rc;
   }
   info->nx = c.nx; info->ny = c.ny; info->nz = c.nz;
   info->n_nabors = c.n_nabors; info->n_frame_types = c.n_frame_types;

   info->eintra = eintra_total(s, chg);
   *pe -= info->eintra;

   molmap = (int *)malloc(sizeof(int) * (size_t)s->nsites);
   id = (int *)malloc(sizeof(int) * (size_t)s->nsites);
   for (ispec = 0, isite = 0, mol = 0; ispec < s->nspecies; ispec++) {
      int off = site_of_species(s, ispec);
      for (im = 0; im < s->spec_nmols[ispec]; im++, mol++)
         for (is = 0; is < s->spec_nsites[ispec]; is++, isite++) {
            molmap[isite] = mol;
            id[isite] = s->spec_site_id[off + is];
         }
   }

   /* n_nab_sites as sized by the reference, src/force.c:1232-1235 and
    * max_density, src/force.c:491-507 */
   {
      int nsmax = 0, e;
      double maxd;
      subcell = s->subcell > 0.0 ? s->subcell : s->cutoff / 5.0;
      for (ic = 0; ic < c.ncells; ic++) {
         int nsc = 0;
         for (e = c.cell_start[ic]; e < c.cell_start[ic + 1]; e++)
            nsc += c.entry_num[c.cell_entries[e]];
         if (nsc > nsmax)
            nsmax = nsc;
      }
      maxd = (double)nsmax * c.ncells / vol;
      if (s->strict_cutoff)
         n_nab_sites = (long)(NMULT * 4.19 * CUBE(s->cutoff + subcell) * maxd);
      else
         n_nab_sites = (long)(NMULT * 4.19 * CUBE(s->cutoff) * maxd);
      info->n_nab_sites = n_nab_sites;
   }
   /* list capacity for this restatement: exact maximum over cells */
   for (ic = 0; ic < c.ncells; ic++) {
      int ix = ic / (c.ny * c.nz), iy = ic / c.nz - c.ny * ix, iz = ic - c.nz * (iy + c.ny * ix);
      long n = cell_site_list(&c, ix, iy, iz, NULL, NULL, NULL, 0);
      if (n > maxlist)
         maxlist = n;
   }
   if (maxlist > n_nab_sites)
      status = status ? status : ORC_ERR_TONAB;
   maxlist += 1;
   nab = (int *)malloc(sizeof(int) * (size_t)maxlist);
   pbc = (int *)malloc(sizeof(int) * (size_t)maxlist);
   nsx = (double *)malloc(sizeof(double) * (size_t)maxlist * 14);
   npot = (double *)malloc(sizeof(double) * (size_t)maxlist * ORC_NPOTP);
   if (!molmap || !id || !nab || !pbc || !nsx || !npot) {
      status = ORC_ERR_ALLOC;
      goto done;
   }
   nsy = nsx + maxlist; nsz = nsy + maxlist; nchg = nsz + maxlist;
   fjx = nchg + maxlist; fjy = fjx + maxlist; fjz = fjy + maxlist;
   rx = fjz + maxlist; ry = rx + maxlist; rz = ry + maxlist; rsq = rz + maxlist;
   fij = rsq + maxlist;
   for (k = 0; k < ORC_NPOTP; k++)
      potv[k] = npot + (size_t)k * maxlist;

   /* reloc_v, src/force.c:1284-1293 */
   k = 0;
   for (ii = -1; ii <= 1; ii++)
      for (jj = -1; jj <= 1; jj++)
         for (kk = -1; kk <= 1; kk++) {
            reloc[k][0] = h[0] * ii + h[1] * jj + h[2] * kk;
            reloc[k][1] = h[3] * ii + h[4] * jj + h[5] * kk;
            reloc[k][2] = h[6] * ii + h[7] * jj + h[8] * kk;
            k++;
         }
   memset(gforce, 0, sizeof gforce);

   /* force_inner, src/force.c:784-1005 */
   for (ic = 0; ic < c.ncells; ic++) {
      int ix, iy, iz, jmin, jmax, e, nnab, j, p;
      if (c.cell_start[ic] == c.cell_start[ic + 1])
         continue;
      ix = ic / (c.ny * c.nz); iy = ic / c.nz - c.ny * ix; iz = ic - c.nz * (iy + c.ny * ix);
      nnab = (int)cell_site_list(&c, ix, iy, iz, nab, pbc, nfnab, maxlist);
      for (j = 0; j < nnab; j++) {
         nsx[j] = x[nab[j]]; nsy[j] = y[nab[j]]; nsz[j] = z[nab[j]];
         nchg[j] = chg[nab[j]];
         fjx[j] = fjy[j] = fjz[j] = 0.0;
      }
      jmin = 0;
      for (e = c.cell_start[ic]; e < c.cell_start[ic + 1]; e++) {
         int en = c.cell_entries[e], ft = c.entry_ftype[en], lim;
         if (ft) {
            jmin = 0;
            jmax = nfnab[0];
         } else
            jmax = nnab;
         lim = c.entry_isite[en] + c.entry_num[en];
         for (isite = c.entry_isite[en]; isite < lim; isite++) {
            double s0 = x[isite], s1 = y[isite], s2 = z[isite], fi0 = 0, fi1 = 0, fi2 = 0;
            const double *prow = s->potpar + ORC_NPOTP * (size_t)(id[isite] * s->max_id);
            if (!ft)
               jmin++;
            for (j = 0; j < jmax; j++)
               for (p = 0; p < ORC_NPOTP; p++)
                  potv[p][j] = prow[ORC_NPOTP * id[nab[j]] + p];
            /* mk_r_sqr, src/force.c:692-711 */
            for (j = 0; j < jmax; j++) {
               int kk2 = pbc[j];
               double ax = nsx[j] - s0 + reloc[kk2][0];
               double ay = nsy[j] - s1 + reloc[kk2][1];
               double az = nsz[j] - s2 + reloc[kk2][2];
               rsq[j] = ax * ax + ay * ay + az * az;
               rx[j] = ax; ry[j] = ay; rz[j] = az;
            }
            /* too-close scan, src/force.c:939-949 */
            for (j = jmin; j < jmax; j++)
               if (rsq[j] < TOO_CLOSE && molmap[isite] != molmap[nab[j]]) {
                  if (info->n_too_close < 16) {
                     info->too_close_i[info->n_too_close] = isite;
                     info->too_close_j[info->n_too_close] = nab[j];
                  }
                  info->n_too_close++;
               }
            /* strict cutoff substitution, src/force.c:951-954 */
            if (s->strict_cutoff && !s->molpbc)
               for (j = jmin; j < jmax; j++)
                  if (rsq[j] > cutsq)
                     rsq[j] = cut100sq;
            rc = orc_kernel(jmin, jmax, fij, pe, rsq, nchg, chg[isite], norm, s->alpha,
                            s->ptype, potv);
            if (rc) {
               status = rc;
               goto done;
            }
            npairs += jmax > jmin ? jmax - jmin : 0;
            /* mk_forces, src/force.c:717-743 */
            for (j = jmin; j < jmax; j++) {
               double cx = fij[j] * rx[j], cy = fij[j] * ry[j], cz = fij[j] * rz[j];
               fi0 -= cx; fi1 -= cy; fi2 -= cz;
               fjx[j] += cx; fjy[j] += cy; fjz[j] += cz;
            }
            fx[isite] += fi0; fy[isite] += fi1; fz[isite] += fi2;
         }
      }
      /* scatter_forces, src/force.c:750-778 */
      for (j = 0; j < nnab; j++) {
         int it = nab[j], kk2 = pbc[j];
         fx[it] += fjx[j]; fy[it] += fjy[j]; fz[it] += fjz[j];
         gforce[kk2][0] += fjx[j]; gforce[kk2][1] += fjy[j]; gforce[kk2][2] += fjz[j];
      }
      info->n_list_entries += nnab;
   }
   info->n_pairs = npairs;

   /* virial, src/force.c:973-997 */
   s00 = s01 = s02 = s11 = s12 = s22 = 0.0;
   for (isite = 0; isite < s->nsites; isite++) {
      s00 += x[isite] * fx[isite];
      s01 += y[isite] * fx[isite];
      s02 += z[isite] * fx[isite];
      s11 += y[isite] * fy[isite];
      s12 += z[isite] * fy[isite];
      s22 += z[isite] * fz[isite];
   }
   for (k = 0; k < 27; k++) {
      s00 += reloc[k][0] * gforce[k][0];
      s01 += reloc[k][1] * gforce[k][0];
      s02 += reloc[k][2] * gforce[k][0];
      s11 += reloc[k][1] * gforce[k][1];
      s12 += reloc[k][2] * gforce[k][1];
      s22 += reloc[k][2] * gforce[k][2];
   }
   stress[0] += s00; stress[1] += s01; stress[2] += s02;
   stress[4] += s11; stress[5] += s12; stress[8] += s22;

done:
   free(molmap); free(id); free(nab); free(pbc); free(nsx); free(npot);
   cells_free(&c);
   return status;
}

/* ================================ ewald ================================= */

int orc_kset(const double h[9], double k_cutoff, int *hkl, double *kvec, int cap,
             int *hmax_o, int *kmax_o, int *lmax_o)
{
   double hinvp[9], kcsq = SQR(k_cutoff), kx, ky, kzt, kz, ksq;
   int hmax = (int)floor(k_cutoff / (2 * PI) * h[0]);
   int kmax = (int)floor(k_cutoff / (2 * PI) * sqrt(SQR(h[1]) + SQR(h[4])));
   int lmax = (int)floor(k_cutoff / (2 * PI) * sqrt(SQR(h[2]) + SQR(h[5]) + SQR(h[8])));
   int hh, kk, ll, n = 0, i;
   orc_invert(h, hinvp);
   for (i = 0; i < 9; i++)
      hinvp[i] = 2 * PI * hinvp[i];
   /* src/ewald.c:441-461; astar = row 0, bstar = row 1, cstar = row 2 */
   for (hh = 0; hh <= hmax; hh++)
      for (kk = (hh == 0 ? 0 : -kmax); kk <= kmax; kk++) {
         kx = hh * hinvp[0] + kk * hinvp[3];
         ky = hh * hinvp[1] + kk * hinvp[4];
         kzt = hh * hinvp[2] + kk * hinvp[5];
         ksq = SQR(kx) + SQR(ky);
         for (ll = (hh == 0 && kk == 0 ? 1 : -lmax); ll <= lmax; ll++) {
            kz = kzt + ll * hinvp[8];
            if (SQR(kz) + ksq < kcsq) {
               if (n < cap) {
                  if (hkl) { hkl[3 * n] = hh; hkl[3 * n + 1] = kk; hkl[3 * n + 2] = ll; }
                  if (kvec) { kvec[3 * n] = kx; kvec[3 * n + 1] = ky; kvec[3 * n + 2] = kz; }
               }
               n++;
            }
         }
      }
   if (hmax_o) *hmax_o = hmax;
   if (kmax_o) *kmax_o = kmax;
   if (lmax_o) *lmax_o = lmax;
   return n;
}

/* trig_recur, src/ewald.c:148-193: cos/sin(n k*.r) for n = 0..nmax by the
 * addition formula from cos/sin(k*.r). */
static void trig_recur(double *c, double *s, int nmax, int ns, const double *x,
                       const double *y, const double *z, const double kstar[3])
{
   int n, i;
   for (i = 0; i < ns; i++) {
      c[i] = 1.0; s[i] = 0.0;
   }
   if (nmax >= 1) {
      double *c1 = c + ns, *s1 = s + ns;
      for (i = 0; i < ns; i++) {
         double kr = kstar[0] * x[i] + kstar[1] * y[i] + kstar[2] * z[i];
         c1[i] = cos(kr); s1[i] = sin(kr);
      }
      for (n = 2; n <= nmax; n++) {
         double *cn = c + (size_t)n * ns, *sn = s + (size_t)n * ns;
         double *cm = c + (size_t)(n - 1) * ns, *sm = s + (size_t)(n - 1) * ns;
         for (i = 0; i < ns; i++) {
            double cc = cm[i] * c1[i] - sm[i] * s1[i];
            sn[i] = sm[i] * c1[i] + cm[i] * s1[i];
            cn[i] = cc;
         }
      }
   }
}

int orc_ewald(const orc_sys *s, const double h[9], const double *x,
              const double *y, const double *z, const double *chg, double *fx,
              double *fy, double *fz, double *pe, double stress[9], orc_ew_info *info)
{
   int ns = s->nsites, hmax, kmax, lmax, nk, ik, i, ispec, ssite, nsitesxf;
   double hinvp[9], vol = orc_det(h), r4a = -1.0 / (4.0 * s->alpha * s->alpha);
   double self_energy = 0.0, sheet_energy = 0.0, sq = 0.0, sqsq = 0.0, sqxf;
   int *hkl;
   double *kv, *chx, *shx, *cky, *sky, *clz, *slz, *qc, *qs, *cxy, *sxy;
   int hlast = -1000000, klast = -1000000, frame_flag;
   orc_ew_info loc;
   if (!info)
      info = &loc;
   memset(info, 0, sizeof *info);

   /* self and sheet energy, src/ewald.c:367-425 */
   ssite = 0;
   for (ispec = 0; ispec < s->nspecies && !s->spec_framework[ispec]; ispec++) {
      int nsp = s->spec_nsites[ispec], off = site_of_species(s, ispec), a, b;
      const double *pf = s->spec_p_f_sites + 3 * off;
      double intra = 0.0;
      for (a = 0; a < nsp; a++)
         for (b = a + 1; b < nsp; b++) {
            double r = sqrt(SQR(pf[3 * a] - pf[3 * b]) + SQR(pf[3 * a + 1] - pf[3 * b + 1]) +
                            SQR(pf[3 * a + 2] - pf[3 * b + 2]));
            intra += chg[ssite + a] * chg[ssite + b] * orc_err_fn(s->alpha * r) / r;
         }
      self_energy += s->spec_nmols[ispec] * intra;
      ssite += nsp * s->spec_nmols[ispec];
   }
   nsitesxf = ssite;
   frame_flag = ispec != s->nspecies;
   for (i = 0; i < nsitesxf; i++) {
      sq += chg[i];
      sqsq += SQR(chg[i]);
   }
   self_energy += s->alpha / sqrt(PI) * sqsq;
   sqxf = sq;
   for (; i < ns; i++)
      sq += chg[i];
   if (frame_flag)
      sheet_energy = PI * SQR(sq - sqxf) / (2.0 * SQR(s->alpha));
   {
      /* CONV_Q = QUNIT/_ELCHG, src/defs.h:204,238 */
      const double conv_q = 4.07497263794495e-14 * 1.e-3 * 1.05482230112e-05 / 1.60217733e-19;
      if (fabs(sq) * conv_q > 1.0e-5)
         sheet_energy -= PI * SQR(sq) / (2.0 * SQR(s->alpha));
   }
   info->self_energy = self_energy;
   info->sheet_energy = sheet_energy;
   info->nsitesxf = nsitesxf;
   *pe -= self_energy;
   *pe += sheet_energy / vol;
   for (i = 0; i < 3; i++)
      stress[4 * i] += sheet_energy / vol;

   orc_invert(h, hinvp);
   for (i = 0; i < 9; i++)
      hinvp[i] = 2 * PI * hinvp[i];
   nk = orc_kset(h, s->k_cutoff, NULL, NULL, 0, &hmax, &kmax, &lmax);
   hkl = (int *)malloc(sizeof(int) * 3 * (size_t)(nk + 1));
   kv = (double *)malloc(sizeof(double) * 3 * (size_t)(nk + 1));
   chx = (double *)malloc(sizeof(double) * (size_t)ns *
                          (2 * (hmax + 1) + 2 * (kmax + 1) + 2 * (lmax + 1) + 4) + 64);
   if (!hkl || !kv || !chx) {
      free(hkl); free(kv); free(chx);
      return ORC_ERR_ALLOC;
   }
   orc_kset(h, s->k_cutoff, hkl, kv, nk, NULL, NULL, NULL);
   info->hmax = hmax; info->kmax = kmax; info->lmax = lmax; info->nhkl = nk;
   shx = chx + (size_t)(hmax + 1) * ns;
   cky = shx + (size_t)(hmax + 1) * ns;
   sky = cky + (size_t)(kmax + 1) * ns;
   clz = sky + (size_t)(kmax + 1) * ns;
   slz = clz + (size_t)(lmax + 1) * ns;
   qc = slz + (size_t)(lmax + 1) * ns;
   qs = qc + ns; cxy = qs + ns; sxy = cxy + ns;

   trig_recur(chx, shx, hmax, ns, x, y, z, hinvp);
   trig_recur(cky, sky, kmax, ns, x, y, z, hinvp + 3);
   trig_recur(clz, slz, lmax, ns, x, y, z, hinvp + 6);
   /* q premultiply, src/ewald.c:475-481 */
   for (i = 0; i <= lmax; i++) {
      int is;
      for (is = 0; is < ns; is++) {
         clz[(size_t)i * ns + is] *= chg[is];
         slz[(size_t)i * ns + is] *= chg[is];
      }
   }
   /* k loop, src/ewald.c:496-583 */
   for (ik = 0; ik < nk; ik++) {
      int hh = hkl[3 * ik], kk = hkl[3 * ik + 1], ll = hkl[3 * ik + 2], is;
      double k0 = kv[3 * ik], k1 = kv[3 * ik + 1], k2 = kv[3 * ik + 2];
      double ksq = k0 * k0 + k1 * k1 + k2 * k2;
      double coeff = 2.0 / (EPS0 * vol) * exp(ksq * r4a) / ksq;
      double coeff2 = 2.0 * (1.0 - ksq * r4a) / ksq;
      const double *ch = chx + (size_t)hh * ns, *sh = shx + (size_t)hh * ns;
      const double *ck = cky + (size_t)abs(kk) * ns, *sk = sky + (size_t)abs(kk) * ns;
      const double *cl = clz + (size_t)abs(ll) * ns, *sl = slz + (size_t)abs(ll) * ns;
      double scn, ssn, scf, ssf, sc, ssv, pe_k;
      /* qsincos, src/ewald.c:96-142 */
      if (hh != hlast || kk != klast) {
         if (kk >= 0)
            for (is = 0; is < ns; is++) {
               double t = ch[is] * ck[is] - sh[is] * sk[is];
               sxy[is] = sh[is] * ck[is] + ch[is] * sk[is];
               cxy[is] = t;
            }
         else
            for (is = 0; is < ns; is++) {
               double t = ch[is] * ck[is] + sh[is] * sk[is];
               sxy[is] = sh[is] * ck[is] - ch[is] * sk[is];
               cxy[is] = t;
            }
      }
      if (ll >= 0)
         for (is = 0; is < ns; is++) {
            double t = cxy[is] * cl[is] - sxy[is] * sl[is];
            qs[is] = sxy[is] * cl[is] + cxy[is] * sl[is];
            qc[is] = t;
         }
      else
         for (is = 0; is < ns; is++) {
            double t = cxy[is] * cl[is] + sxy[is] * sl[is];
            qs[is] = sxy[is] * cl[is] - cxy[is] * sl[is];
            qc[is] = t;
         }
      hlast = hh; klast = kk;
      scn = orc_sum(nsitesxf, qc);
      ssn = orc_sum(nsitesxf, qs);
      scf = orc_sum(ns - nsitesxf, qc + nsitesxf);
      ssf = orc_sum(ns - nsitesxf, qs + nsitesxf);
      sc = scn + scf;
      ssv = ssn + ssf;
      pe_k = 0.5 * coeff * (scn * (scn + scf + scf) + ssn * (ssn + ssf + ssf));
      *pe += pe_k;
      ssv *= coeff; sc *= coeff; ssn *= coeff; scn *= coeff;
      stress[0] += pe_k - pe_k * coeff2 * k0 * k0;
      stress[1] -= pe_k * coeff2 * k0 * k1;
      stress[2] -= pe_k * coeff2 * k0 * k2;
      stress[4] += pe_k - pe_k * coeff2 * k1 * k1;
      stress[5] -= pe_k * coeff2 * k1 * k2;
      stress[8] += pe_k - pe_k * coeff2 * k2 * k2;
      for (is = 0; is < nsitesxf; is++) {
         double fc = qs[is] * sc - qc[is] * ssv;
         fx[is] += k0 * fc; fy[is] += k1 * fc; fz[is] += k2 * fc;
      }
      for (is = nsitesxf; is < ns; is++) {
         double fc = qs[is] * scn - qc[is] * ssn;
         fx[is] += k0 * fc; fy[is] += k1 * fc; fz[is] += k2 * fc;
      }
   }
   free(hkl); free(kv); free(chx);
   return 0;
}

/* ============================= eval_forces ============================== */

/* make_sites, src/algorith.c:169-217 (molflag_all: 1 = MOLPBC for every
 * species, 0 = SITEPBC for every species, 2 = SITEPBC for frameworks only). */
void orc_make_sites(const orc_sys *s, const double h[9], const double *cmf,
                    const double *quat, int molflag_all, double *x, double *y, double *z)
{
   double hinv[9], rot[9], cm[3], p[3];
   int ispec, im, is, isite = 0, imol = 0, iq = 0;
   orc_invert(h, hinv);
   for (ispec = 0; ispec < s->nspecies; ispec++) {
      int ns = s->spec_nsites[ispec], off = site_of_species(s, ispec), first = isite;
      const double *pf = s->spec_p_f_sites + 3 * off;
      int molpbc = molflag_all == 1 || (molflag_all == 2 && !s->spec_framework[ispec]);
      for (im = 0; im < s->spec_nmols[ispec]; im++, imol++) {
         matvec3(h, cmf + 3 * imol, cm);
         if (s->spec_rdof[ispec] > 0) {
            q_to_rot(quat + 4 * iq, rot);
            iq++;
            for (is = 0; is < ns; is++) {
               matvec3(rot, pf + 3 * is, p);
               x[isite + is] = p[0] + cm[0];
               y[isite + is] = p[1] + cm[1];
               z[isite + is] = p[2] + cm[2];
            }
         } else
            for (is = 0; is < ns; is++) {
               x[isite + is] = pf[3 * is] + cm[0];
               y[isite + is] = pf[3 * is + 1] + cm[1];
               z[isite + is] = pf[3 * is + 2] + cm[2];
            }
         isite += ns;
      }
      if (!molpbc)
         for (is = first; is < isite; is++) {
            double tx = floor(hinv[0] * x[is] + hinv[1] * y[is] + hinv[2] * z[is] + 0.5);
            double ty = floor(hinv[3] * x[is] + hinv[4] * y[is] + hinv[5] * z[is] + 0.5);
            double tz = floor(hinv[6] * x[is] + hinv[7] * y[is] + hinv[8] * z[is] + 0.5);
            x[is] -= h[0] * tx + h[1] * ty + h[2] * tz;
            y[is] -= h[3] * tx + h[4] * ty + h[5] * tz;
            z[is] -= h[6] * tx + h[7] * ty + h[8] * tz;
         }
   }
}

/* distant_const, src/accel.c:293-329 */
double orc_distant_const(const orc_sys *s, double cutoff, int iflag)
{
   int *count = (int *)calloc((size_t)s->max_id, sizeof(int));
   int ispec, is, a, b;
   double c = 0.0;
   for (ispec = 0; ispec < s->nspecies; ispec++) {
      int off = site_of_species(s, ispec);
      for (is = 0; is < s->spec_nsites[ispec]; is++)
         count[s->spec_site_id[off + is]] += s->spec_nmols[ispec];
   }
   for (a = 1; a < s->max_id; a++)
      for (b = 1; b < s->max_id; b++) {
         const double *p = s->potpar + ORC_NPOTP * (size_t)(a + s->max_id * b);
         c -= 2 * PI * count[a] * count[b] * orc_dist_pot(p, cutoff, s->ptype);
         if (iflag)
            c += 2.0 / 3.0 * PI * count[a] * count[b] * CUBE(cutoff) *
                 orc_poteval(p, cutoff, s->ptype, 0.0, s->alpha);
      }
   free(count);
   return c;
}

int orc_eval_forces(const orc_sys *s, const double h[9], const double *cmf,
                    const double *quat, double pe[2], double dip_mom[3],
                    double stress[9], double *force, double *torque,
                    double *site_xyz_out, double *site_force_out, orc_fc_info *fci,
                    orc_ew_info *ewi)
{
   int ns = s->nsites, ispec, im, is, i, j, isite, imol, iq, rc = 0;
   int nsitesxf = 0, nmolsxf = 0, fspec;
   double vol = orc_det(h), dist, distp;
   double *x = (double *)calloc((size_t)ns * 7 + 8, sizeof(double));
   double *y, *z, *fx, *fy, *fz, *chg, *cmr;
   if (!x)
      return ORC_ERR_ALLOC;
   y = x + ns; z = y + ns; fx = z + ns; fy = fx + ns; fz = fy + ns; chg = fz + ns;
   cmr = (double *)malloc(sizeof(double) * 3 * (size_t)(s->nmols + 1));

   dist = orc_distant_const(s, s->cutoff, 0);
   distp = orc_distant_const(s, s->cutoff, 1);

   for (fspec = 0; fspec < s->nspecies && !s->spec_framework[fspec]; fspec++) {
      nsitesxf += s->spec_nsites[fspec] * s->spec_nmols[fspec];
      nmolsxf += s->spec_nmols[fspec];
   }
   /* charges, src/accel.c:480-484 */
   for (ispec = 0, isite = 0; ispec < s->nspecies; ispec++) {
      int off = site_of_species(s, ispec);
      for (im = 0; im < s->spec_nmols[ispec]; im++)
         for (is = 0; is < s->spec_nsites[ispec]; is++)
            chg[isite++] = s->site_charge[s->spec_site_id[off + is]];
   }
   memset(stress, 0, 9 * sizeof(double));
   dip_mom[0] = dip_mom[1] = dip_mom[2] = 0.0;
   pe[0] = pe[1] = 0.0;

   orc_make_sites(s, h, cmf, quat, s->molpbc ? 1 : 0, x, y, z);
   rc = orc_force_calc(s, h, x, y, z, cmf, chg, fx, fy, fz, &pe[0], stress, fci);
   if (rc)
      goto out;
   if (s->alpha > ALPHAMIN) {
      rc = orc_ewald(s, h, x, y, z, chg, fx, fy, fz, &pe[1], stress, ewi);
      if (rc)
         goto out;
   }
   orc_make_sites(s, h, cmf, quat, 2, x, y, z);
   if (s->alpha > ALPHAMIN) {
      /* dipole and surface term, src/accel.c:542-558 */
      dip_mom[0] = vdot_s(ns, x, 1, chg, 1);
      dip_mom[1] = vdot_s(ns, y, 1, chg, 1);
      dip_mom[2] = vdot_s(ns, z, 1, chg, 1);
      if (s->surface_dipole) {
         double *fa[3] = { fx, fy, fz };
         for (i = 0; i < 3; i++)
            for (is = 0; is < ns; is++)
               fa[i][is] -= 4.0 * PI / (3.0 * vol) * dip_mom[i] * chg[is];
         pe[1] += 2.0 * PI / (3.0 * vol) * (dip_mom[0] * dip_mom[0] + dip_mom[1] * dip_mom[1] +
                                             dip_mom[2] * dip_mom[2]);
      }
   }
   /* mol_force and mol_torque, src/algorith.c:111-163 */
   for (ispec = 0, isite = 0, imol = 0, iq = 0; ispec < s->nspecies; ispec++) {
      int nsp = s->spec_nsites[ispec], off = site_of_species(s, ispec);
      const double *pf = s->spec_p_f_sites + 3 * off;
      for (im = 0; im < s->spec_nmols[ispec]; im++, imol++) {
         double f[3] = { 0, 0, 0 };
         for (is = 0; is < nsp; is++) {
            f[0] += fx[isite + is];
         }
         for (is = 0; is < nsp; is++)
            f[1] += fy[isite + is];
         for (is = 0; is < nsp; is++)
            f[2] += fz[isite + is];
         force[3 * imol] = f[0]; force[3 * imol + 1] = f[1]; force[3 * imol + 2] = f[2];
         if (s->spec_rdof[ispec] > 0) {
            double rot[9], pfs[3], tq[3];
            int a, b, c;
            q_to_rot(quat + 4 * iq, rot);
            for (a = 0, b = 1, c = 2; a < 3; a++, b = (b + 1) % 3, c = (c + 1) % 3) {
               double t = 0.0;
               for (is = 0; is < nsp; is++) {
                  double fs[3] = { fx[isite + is], fy[isite + is], fz[isite + is] };
                  /* inverse rotation: rot^T f */
                  pfs[0] = rot[0] * fs[0] + rot[3] * fs[1] + rot[6] * fs[2];
                  pfs[1] = rot[1] * fs[0] + rot[4] * fs[1] + rot[7] * fs[2];
                  pfs[2] = rot[2] * fs[0] + rot[5] * fs[1] + rot[8] * fs[2];
                  t += pf[3 * is + b] * pfs[c] - pf[3 * is + c] * pfs[b];
               }
               tq[a] = t;
            }
            torque[3 * iq] = tq[0]; torque[3 * iq + 1] = tq[1]; torque[3 * iq + 2] = tq[2];
            iq++;
         }
         isite += nsp;
      }
   }
   /* site -> molecular virial, src/accel.c:575-599 */
   for (imol = 0; imol < s->nmols; imol++)
      matvec3(h, cmf + 3 * imol, cmr + 3 * imol);
   {
      double *fa[3] = { fx, fy, fz }, *sa[3] = { x, y, z };
      for (i = 0; i < 3; i++) {
         for (j = i + 1; j < 3; j++)
            stress[3 * j + i] = stress[3 * i + j];
         for (j = 0; j < 3; j++) {
            stress[3 * i + j] -= vdot_s(nsitesxf, fa[i], 1, sa[j], 1)
                                 - vdot_s(nmolsxf, force + i, 3, cmr + j, 3);
            if (fspec < s->nspecies) {
               int off = site_of_species(s, fspec), nsp = s->spec_nsites[fspec];
               for (im = 0; im < s->spec_nmols[fspec]; im++)
                  stress[3 * i + j] -= vdot_s(nsp, fa[i] + nsitesxf + im * nsp, 1,
                                              s->spec_p_f_sites + 3 * off + j, 3);
            }
         }
      }
   }
   /* distant corrections, src/accel.c:604-606 */
   pe[0] += dist / vol;
   for (i = 0; i < 3; i++)
      stress[4 * i] += distp / vol;
   if (site_xyz_out)
      memcpy(site_xyz_out, x, sizeof(double) * 3 * (size_t)ns);
   if (site_force_out)
      memcpy(site_force_out, fx, sizeof(double) * 3 * (size_t)ns);
out:
   free(x);
   free(cmr);
   return rc;
}

/* ====================================================================== */
/* Molecular dynamics step: the leapfrog splitting of src/accel.c:626-827  */
/* with the symplectic rotation of src/leapfrog.c.  No barostat.           */
/* ====================================================================== */

#define ORC_INERTIA_MIN 1.0e-14
#define ORC_QTOL 1.0e-4
#define ORC_DBL_MIN 2.2250738585072014e-308

/* q_mul / q_conj_mul, src/quaterns.c:38-98 (r may alias p or q) */
static void orc_qmul(const double *p, const double *q, double *r)
{
   double p0 = p[0], p1 = p[1], p2 = p[2], p3 = p[3];
   double q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
   r[0] = p0 * q0 - p1 * q1 - p2 * q2 - p3 * q3;
   r[1] = p1 * q0 + p0 * q1 - p3 * q2 + p2 * q3;
   r[2] = p2 * q0 + p3 * q1 + p0 * q2 - p1 * q3;
   r[3] = p3 * q0 - p2 * q1 + p1 * q2 + p0 * q3;
}
static void orc_qconjmul(const double *p, const double *q, double *r)
{
   double pc[4];
   pc[0] = p[0]; pc[1] = -p[1]; pc[2] = -p[2]; pc[3] = -p[3];
   orc_qmul(pc, q, r);
}

/* rot_substep, src/leapfrog.c:238-244: amom <- rot* amom rot, quat <- quat rot */
static void orc_rot_substep(const double *rot, double *amom, double *quat)
{
   orc_qconjmul(rot, amom, amom);
   orc_qmul(amom, rot, amom);
   orc_qmul(quat, rot, quat);
}

/* make_rot, src/leapfrog.c:193-208 */
static void orc_make_rot(double step, int axis, const double *amom, double rinertia, double *rot)
{
   double angle = 0.5 * step * rinertia * amom[axis + 1];
   rot[0] = cos(angle);
   rot[1] = rot[2] = rot[3] = 0.0;
   rot[axis + 1] = sin(angle);
}

/* make_rot_amom, src/leapfrog.c:213-231 */
static void orc_make_rot_amom(double step, const double *amom, double rinertia, double *rot)
{
   double samom = sqrt(amom[1] * amom[1] + amom[2] * amom[2] + amom[3] * amom[3]);
   double ramom = 1.0 / (samom + (8 * ORC_DBL_MIN));
   double angle = 0.5 * step * rinertia * samom, sa = sin(angle);
   rot[0] = cos(angle);
   rot[1] = sa * ramom * amom[1];
   rot[2] = sa * ramom * amom[2];
   rot[3] = sa * ramom * amom[3];
}

/* rot_ke, src/algorith.c:244-259 */
static double orc_rot_ke(const double *amom, int nmols, double s, const double *inertia)
{
   double ke = 0.0, t;
   int i, m;
   for (i = 0; i < 3; i++)
      if (inertia[i] > ORC_INERTIA_MIN) {
         t = 0.0;
         for (m = 0; m < nmols; m++)
            t += amom[4 * m + i + 1] * amom[4 * m + i + 1];
         ke += t / inertia[i];
      }
   return 0.5 * ke / (s * s);
}

/* trans_ke, src/algorith.c:221-239: real momenta (h^-1)^T mom */
static double orc_trans_ke(const double h[9], const double *mom, int nmols, double s, double mass)
{
   double hi[9], ke = 0.0, p[3];
   int m, i;
   orc_invert(h, hi);
   for (m = 0; m < nmols; m++) {
      const double *q = mom + 3 * m;
      for (i = 0; i < 3; i++)
         p[i] = hi[i] * q[0] + hi[3 + i] * q[1] + hi[6 + i] * q[2];
      ke += p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
   }
   return ke / (2.0 * mass * s * s);
}

static void orc_rinertia(const double *inertia, double *rinertia)
{
   int i;
   for (i = 0; i < 3; i++)
      if (inertia[i] / (inertia[(i + 1) % 3] + inertia[(i + 2) % 3]) < ORC_INERTIA_MIN)
         rinertia[i] = 0.0;
      else
         rinertia[i] = 1.0 / inertia[i];
}

/* Near-symmetry axis of leapf_quat_b, src/leapfrog.c:275-290.  The reference
 * keeps the first call's choice in a static for every later species; the
 * first call is made for the first species with rotational freedom. */
int orc_symmetry_axis(const double *inertia)
{
   double rin[3], idmin = 1.0e37, idiff;
   int i, saxis = 0;
   orc_rinertia(inertia, rin);
   for (i = 0; i < 3; i++) {
      idiff = fabs(rin[(i + 1) % 3] - rin[(i + 2) % 3]);
      if (idiff < idmin) {
         idmin = idiff;
         saxis = i;
      }
   }
   return saxis;
}

/* normalise, src/leapfrog.c:96-113 */
static int orc_normalise(double *quat, int n, int *bad)
{
   int i, j;
   double norm;
   for (i = 0; i < n; i++) {
      norm = 0.0;
      for (j = 0; j < 4; j++)
         norm += quat[4 * i + j] * quat[4 * i + j];
      norm = sqrt(norm);
      if (fabs(norm - 1.0) > ORC_QTOL) {
         *bad = i;
         return ORC_ERR_QNORM;
      }
      for (j = 0; j < 4; j++)
         quat[4 * i + j] /= norm;
   }
   return ORC_OK;
}

/* leapf_quat_b, src/leapfrog.c:255-338 */
static int orc_leapf_quat_b(const orc_dyn *d, double step, double *quat, double *amom,
                            const double *inertia, double *smom, double ts, int nmols,
                            int saxis, int *bad)
{
   double rin[3], rot[4], stepdts = step / ts;
   int o1 = (saxis + 1) % 3, o2 = (saxis + 2) % 3, m;
   orc_rinertia(inertia, rin);
   for (m = 0; m < nmols; m++) {
      orc_make_rot(0.5 * stepdts, o1, amom + 4 * m, rin[o1] - rin[o2], rot);
      orc_rot_substep(rot, amom + 4 * m, quat + 4 * m);
   }
   if (d->const_temp)
      *smom += 0.5 * step * orc_rot_ke(amom, nmols, ts, inertia);
   for (m = 0; m < nmols; m++) {
      orc_make_rot(stepdts, saxis, amom + 4 * m, rin[saxis] - rin[o2], rot);
      orc_rot_substep(rot, amom + 4 * m, quat + 4 * m);
      orc_make_rot_amom(stepdts, amom + 4 * m, rin[o2], rot);
      orc_qmul(quat + 4 * m, rot, quat + 4 * m);
   }
   if (d->const_temp)
      *smom += 0.5 * step * orc_rot_ke(amom, nmols, ts, inertia);
   for (m = 0; m < nmols; m++) {
      orc_make_rot(0.5 * stepdts, o1, amom + 4 * m, rin[o1] - rin[o2], rot);
      orc_rot_substep(rot, amom + 4 * m, quat + 4 * m);
   }
   return orc_normalise(quat, nmols, bad);
}

static double orc_axis_sumsq(const double *amom, int nmols, int axis)
{
   double t = 0.0;
   int m;
   for (m = 0; m < nmols; m++)
      t += amom[4 * m + axis + 1] * amom[4 * m + axis + 1];
   return t;
}

/* leapf_quat_a, src/leapfrog.c:345-392 */
static int orc_leapf_quat_a(const orc_dyn *d, double step, double *quat, double *amom,
                            const double *inertia, double *smom, double ts, int nmols, int *bad)
{
   double rin[3], rot[4], stepdts = step / ts, stepdts2 = step / (ts * ts);
   static const int axis[5] = {0, 1, 2, 1, 0};
   static const double frac[5] = {0.5, 0.5, 1.0, 0.5, 0.5};
   int k, m;
   orc_rinertia(inertia, rin);
   for (k = 0; k < 5; k++) {
      for (m = 0; m < nmols; m++) {
         orc_make_rot(frac[k] * stepdts, axis[k], amom + 4 * m, rin[axis[k]], rot);
         orc_rot_substep(rot, amom + 4 * m, quat + 4 * m);
      }
      if (d->const_temp) {
         if (k == 0) {
            *smom += 0.25 * stepdts2 * rin[0] * orc_axis_sumsq(amom, nmols, 0);
            *smom += 0.25 * stepdts2 * rin[1] * orc_axis_sumsq(amom, nmols, 1);
         } else if (k == 1) {
            *smom += 0.5 * stepdts2 * rin[2] * orc_axis_sumsq(amom, nmols, 2);
         } else if (k == 3) {
            *smom += 0.25 * stepdts2 * rin[1] * orc_axis_sumsq(amom, nmols, 1);
            *smom += 0.25 * stepdts2 * rin[0] * orc_axis_sumsq(amom, nmols, 0);
         }
      }
   }
   return orc_normalise(quat, nmols, bad);
}

/* leapf_all_coords, src/accel.c:360-372, with leapf_com (leapfrog.c:133-146) */
static int orc_leapf_all_coords(const orc_sys *s, const orc_dyn *d, const double h[9],
                                double step, orc_md_state *st, int saxis, int *bad)
{
   double ht[9], G[9], Gi[9], c;
   int i, j, k, sp, m, imol = 0, iq = 0, rc;
   for (i = 0; i < 3; i++)
      for (j = 0; j < 3; j++)
         ht[3 * i + j] = h[3 * j + i];
   for (i = 0; i < 3; i++)
      for (j = 0; j < 3; j++) {
         G[3 * i + j] = 0.0;
         for (k = 0; k < 3; k++)
            G[3 * i + j] += ht[3 * i + k] * h[3 * k + j];
      }
   orc_invert(G, Gi);
   for (sp = 0; sp < s->nspecies; sp++) {
      int nm = s->spec_nmols[sp];
      double A[9];
      c = step / (d->spec_mass[sp] * st->ts);
      for (i = 0; i < 9; i++)
         A[i] = c * Gi[i];
      for (m = imol; m < imol + nm; m++) {
         double *r = st->c_of_m + 3 * m;
         const double *p = st->mom + 3 * m;
         double y0 = r[0] + A[0] * p[0] + A[1] * p[1] + A[2] * p[2];
         double y1 = r[1] + A[3] * p[0] + A[4] * p[1] + A[5] * p[2];
         double y2 = r[2] + A[6] * p[0] + A[7] * p[1] + A[8] * p[2];
         r[0] = y0; r[1] = y1; r[2] = y2;
         /* escape, leapfrog.c:117-128 */
         r[0] -= floor(r[0] + 0.5);
         r[1] -= floor(r[1] + 0.5);
         r[2] -= floor(r[2] + 0.5);
      }
      if (s->spec_rdof[sp] > 0) {
         const double *I = d->spec_inertia + 3 * sp;
         if (d->nosymmetric_rot)
            rc = orc_leapf_quat_a(d, step, st->quat + 4 * iq, st->amom + 4 * iq, I, &st->tsmom,
                                  st->ts, nm, bad);
         else
            rc = orc_leapf_quat_b(d, step, st->quat + 4 * iq, st->amom + 4 * iq, I, &st->tsmom,
                                  st->ts, nm, saxis, bad);
         if (rc) {
            *bad += iq;
            return rc;
         }
         iq += nm;
      }
      imol += nm;
   }
   return ORC_OK;
}

/* leapf_all_momenta, src/accel.c:376-388, with leapf_mom / leapf_amom */
static void orc_leapf_all_momenta(const orc_sys *s, const double h[9], double step,
                                  orc_md_state *st, const double *force, const double *torque)
{
   double A[9], c = step * st->ts;
   int i, j, m, iq = 0, imol = 0, sp;
   for (i = 0; i < 3; i++)
      for (j = 0; j < 3; j++)
         A[3 * i + j] = c * h[3 * j + i];
   for (sp = 0; sp < s->nspecies; sp++) {
      int nm = s->spec_nmols[sp];
      for (m = imol; m < imol + nm; m++) {
         double *p = st->mom + 3 * m;
         const double *f = force + 3 * m;
         double y0 = p[0] + A[0] * f[0] + A[1] * f[1] + A[2] * f[2];
         double y1 = p[1] + A[3] * f[0] + A[4] * f[1] + A[5] * f[2];
         double y2 = p[2] + A[6] * f[0] + A[7] * f[1] + A[8] * f[2];
         p[0] = y0; p[1] = y1; p[2] = y2;
      }
      if (s->spec_rdof[sp] > 0) {
         for (m = iq; m < iq + nm; m++) {
            st->amom[4 * m + 1] += c * torque[3 * m];
            st->amom[4 * m + 2] += c * torque[3 * m + 1];
            st->amom[4 * m + 3] += c * torque[3 * m + 2];
         }
         iq += nm;
      }
      imol += nm;
   }
}

/* tot_ke, src/accel.c:333-345 */
static void orc_tot_ke(const orc_sys *s, const orc_dyn *d, const double h[9],
                       const orc_md_state *st, double *ke_t, double *ke_r)
{
   int sp, imol = 0, iq = 0;
   *ke_t = 0.0;
   *ke_r = 0.0;
   for (sp = 0; sp < s->nspecies; sp++) {
      int nm = s->spec_nmols[sp];
      *ke_t += orc_trans_ke(h, st->mom + 3 * imol, nm, st->ts, d->spec_mass[sp]);
      if (s->spec_rdof[sp] > 0) {
         *ke_r += orc_rot_ke(st->amom + 4 * iq, nm, st->ts, d->spec_inertia + 3 * sp);
         iq += nm;
      }
      imol += nm;
   }
}

/* leapf_nose_therm, src/leapfrog.c:468-474 */
static void orc_nose_therm(double step, double mass, double *s, double *smom)
{
   double scale = (1.0 + *smom * step / (2.0 * mass));
   *smom /= scale;
   *s *= scale * scale;
}

/* ---- barostat pieces, src/leapfrog.c:397-531 and src/algorith.c:261-306 ---- */
static double orc_trace_sqr(const double m[9])
{
   int i;
   double sum = 0.0;
   for (i = 0; i < 9; i++)
      sum += m[i] * m[i];
   return sum;
}

static double orc_ke_cell(const double hmom[9], double w)
{
   double s = 0.0;
   int i;
   for (i = 0; i < 9; i++)
      s += hmom[i] * hmom[i];
   return 0.5 / w * s;
}

/* mk_sigma: sigma = vol * h^T^-1 (cofactors), src/algorith.c:287-294 */
static void orc_mk_sigma(const double h[9], double sg[9])
{
   int i, j, k, l, m, n;
   for (i = 0, j = 1, k = 2; i < 3; i++, j = (j + 1) % 3, k = (k + 1) % 3)
      for (l = 0, m = 1, n = 2; l < 3; l++, m = (m + 1) % 3, n = (n + 1) % 3)
         sg[3 * i + l] = h[3 * j + m] * h[3 * k + n] - h[3 * k + m] * h[3 * j + n];
}

/* leapf_hmom, src/leapfrog.c:415-430 */
static void orc_leapf_hmom(double step, double hmom[9], const double sigma[9], double s,
                           double pressure, int mask)
{
   int i;
   double r = 0.5 * step * s * pressure;
   for (i = 0; i < 9; i++) {
      if (mask & 1)
         hmom[i] = 0;
      else
         hmom[i] -= r * sigma[i];
      mask >>= 1;
   }
}

/* gleap_cell, src/leapfrog.c:479-508 */
static void orc_gleap_cell(const orc_dyn *d, double step, double s, double h[9], double hmom[9],
                           double *smom, int uniform)
{
   double vol = orc_det(h), sigma[9];
   int i;
   if (uniform)
      orc_leapf_hmom(step, hmom, h, 3.0 * vol * s / orc_trace_sqr(h), d->pressure, d->strain_mask);
   else {
      orc_mk_sigma(h, sigma);
      orc_leapf_hmom(step, hmom, sigma, s, d->pressure, d->strain_mask);
   }
   if (d->const_temp)
      *smom -= 0.5 * step * (orc_ke_cell(hmom, d->pmass) + d->pressure * vol);
   for (i = 0; i < 9; i++)   /* leapf_h, src/leapfrog.c:434-442 */
      h[i] += step * s / d->pmass * hmom[i];
   vol = orc_det(h);
   if (d->const_temp)
      *smom -= 0.5 * step * (orc_ke_cell(hmom, d->pmass) + d->pressure * vol);
   if (uniform)
      orc_leapf_hmom(step, hmom, h, 3.0 * vol * s / orc_trace_sqr(h), d->pressure, d->strain_mask);
   else {
      orc_mk_sigma(h, sigma);
      orc_leapf_hmom(step, hmom, sigma, s, d->pressure, d->strain_mask);
   }
}

/* update_hmom, src/leapfrog.c:516-531; stress_part is a virial (times V) */
static void orc_update_hmom(double step, double s, const double h[9], const double sp[9],
                            double hmom[9], int uniform)
{
   double tmp[9], sigma[9], vol, c;
   int i, j, k;
   if (uniform) {
      c = step * s * (sp[0] + sp[4] + sp[8]) / orc_trace_sqr(h);
      for (i = 0; i < 9; i++)
         tmp[i] = c * h[i];
   } else {
      vol = orc_det(h);
      orc_mk_sigma(h, sigma);
      for (i = 0; i < 3; i++)
         for (j = 0; j < 3; j++) {
            tmp[3 * i + j] = 0.0;
            for (k = 0; k < 3; k++)
               tmp[3 * i + j] += sp[3 * i + k] * sigma[3 * k + j];
         }
      c = step * s / vol;
      for (i = 0; i < 9; i++)
         tmp[i] *= c;
   }
   for (i = 0; i < 9; i++)
      hmom[i] += tmp[i];
}

/* stress_kin / energy_dyad, src/accel.c:349-355, src/algorith.c:261-285 */
static void orc_stress_kin(const orc_sys *s, const orc_dyn *d, const orc_md_state *st,
                           double kd[9])
{
   double hi[9], p[3];
   int sp, m, i, j, imol = 0;
   orc_invert(st->h, hi);
   for (i = 0; i < 9; i++)
      kd[i] = 0.0;
   for (sp = 0; sp < s->nspecies; sp++) {
      int nm = s->spec_nmols[sp];
      double dy[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      for (m = imol; m < imol + nm; m++) {
         const double *q = st->mom + 3 * m;
         for (i = 0; i < 3; i++)
            p[i] = hi[i] * q[0] + hi[3 + i] * q[1] + hi[6 + i] * q[2];
         for (i = 0; i < 3; i++)
            for (j = 0; j < 3; j++)
               dy[3 * i + j] += p[i] * p[j];
      }
      for (i = 0; i < 9; i++)
         kd[i] += dy[i] / (d->spec_mass[sp] * st->ts * st->ts);
      imol += nm;
   }
}

int orc_do_step(const orc_sys *s, const orc_dyn *d, orc_md_state *st, int init_H_0,
                double pe[2], double stress[9], double rec[26])
{
   int nm = s->nmols, sp, rc, bad = -1, imol, saxis = 0, m;
   const int press = d->const_pressure != 0;
   const int uni = d->const_pressure > 0 && d->const_pressure % 2 == 0;
   double step = d->step, ke_t, ke_r, dip[3], kd[9];
   double *h = st->h;
   double *force = (double *)calloc((size_t)nm * 6 + 6, sizeof(double)), *torque;
   if (!force)
      return ORC_ERR_ALLOC;
   torque = force + 3 * (size_t)nm + 3;
   for (sp = 0; sp < s->nspecies; sp++)
      if (s->spec_rdof[sp] > 0) {
         saxis = orc_symmetry_axis(d->spec_inertia + 3 * sp);
         break;
      }
   /* H1 half step */
   if (d->const_temp) {
      orc_nose_therm(0.5 * step, d->ttmass, &st->ts, &st->tsmom);
      st->tsmom -= 0.5 * step * d->gkt * (1.0 + log(st->ts));
   }
   /* H2 half step */
   if (press)
      orc_gleap_cell(d, 0.5 * step, st->ts, h, st->hmom, &st->tsmom, uni);
   /* H3 half step */
   if (d->const_temp) {
      orc_tot_ke(s, d, h, st, &ke_t, &ke_r);
      st->tsmom += 0.5 * step * ke_t;
   }
   if (press) {
      orc_stress_kin(s, d, st, kd);
      orc_update_hmom(0.5 * step, st->ts, h, kd, st->hmom, uni);
   }
   rc = orc_leapf_all_coords(s, d, h, 0.5 * step, st, saxis, &bad);
   if (rc)
      goto out;
   /* H4 full step */
   rc = orc_eval_forces(s, h, st->c_of_m, st->quat, pe, dip, stress, force, torque, NULL, NULL,
                        NULL, NULL);
   if (rc)
      goto out;
   orc_leapf_all_momenta(s, h, 0.5 * step, st, force, torque);
   orc_tot_ke(s, d, h, st, &ke_t, &ke_r);
   if (init_H_0) {
      st->H_0 = (ke_t + ke_r) + (pe[0] + pe[1]) + st->tsmom * st->tsmom / (2.0 * d->ttmass) +
                d->gkt * log(st->ts) + orc_ke_cell(st->hmom, d->pmass);
      if (press)
         st->H_0 += d->pressure * orc_det(h);
   }
   if (rec) {
      rec[0] = ke_t;
      rec[1] = ke_r;
      rec[2] = pe[0];
      rec[3] = pe[1];
      rec[4] = st->ts;
      rec[5] = st->tsmom;
      rec[6] = st->H_0;
      rec[7] = orc_det(h);
      for (m = 0; m < 9; m++) {
         rec[8 + m] = st->hmom[m];
         rec[17 + m] = h[m];
      }
   }
   orc_leapf_all_momenta(s, h, 0.5 * step, st, force, torque);
   /* frameworks do not move, src/accel.c:752-757 */
   for (sp = 0, imol = 0; sp < s->nspecies; sp++) {
      if (s->spec_framework[sp])
         for (m = imol; m < imol + s->spec_nmols[sp]; m++)
            st->mom[3 * m] = st->mom[3 * m + 1] = st->mom[3 * m + 2] = 0.0;
      imol += s->spec_nmols[sp];
   }
   if (d->const_temp)
      st->tsmom -= step * (pe[0] + pe[1] - st->H_0);
   if (press)
      orc_update_hmom(step, st->ts, h, stress, st->hmom, uni);
   /* second H3 half step */
   if (d->const_temp) {
      orc_tot_ke(s, d, h, st, &ke_t, &ke_r);
      st->tsmom += 0.5 * step * ke_t;
   }
   if (press) {
      orc_stress_kin(s, d, st, kd);
      orc_update_hmom(0.5 * step, st->ts, h, kd, st->hmom, uni);
   }
   rc = orc_leapf_all_coords(s, d, h, 0.5 * step, st, saxis, &bad);
   if (rc)
      goto out;
   /* second H2 half step */
   if (press)
      orc_gleap_cell(d, 0.5 * step, st->ts, h, st->hmom, &st->tsmom, uni);
   /* final H1 half step */
   if (d->const_temp) {
      st->tsmom -= 0.5 * step * d->gkt * (1.0 + log(st->ts));
      orc_nose_therm(0.5 * step, d->ttmass, &st->ts, &st->tsmom);
   }
out:
   if (rc == ORC_ERR_QNORM)
      st->bad_quat = bad;
   free(force);
   return rc;
}

/* ====================================================================== */
/* Radial distribution functions: rdf_inner (src/force.c:1010-1103, run    */
/* from force_calc at src/force.c:1302-1313) and rdf_accum (src/rdf.c:94-108) */
/* ====================================================================== */

int orc_rdf_block(int max_id, int idi, int idj)
{
   /* init_rdf, src/rdf.c:71-89: blocks for idi = 1.., idj = idi.. */
   int a = idi < idj ? idi : idj, b = idi < idj ? idj : idi, k = 0, i;
   for (i = 1; i < a; i++)
      k += max_id - i;
   return k + (b - a);
}

int orc_rdf(const orc_sys *s, const double h[9], const double *x, const double *y,
            const double *z, const double *cm, double limit, int nbins, float *rdf)
{
   cells_t c;
   int status = 0, rc, ic, isite, is, im, ispec;
   int *id = NULL, *nab = NULL, *pbc = NULL, nfnab[2], ii, jj, kk, k;
   double *rsq = NULL, reloc[27][3];
   double density = s->nsites / orc_det(h), invrho = 1.0 / density, rbin = nbins / limit;
   long maxlist = 0;

   rc = cells_build(s, h, x, y, z, cm, 1, limit, &c, &status);
   if (rc < 0) {
      cells_free(&c);
      return -rc;
   }
   id = (int *)malloc(sizeof(int) * (size_t)s->nsites);
   if (!id) {
      status = ORC_ERR_ALLOC;
      goto done;
   }
   for (ispec = 0, isite = 0; ispec < s->nspecies; ispec++) {
      int off = site_of_species(s, ispec);
      for (im = 0; im < s->spec_nmols[ispec]; im++)
         for (is = 0; is < s->spec_nsites[ispec]; is++, isite++) {
            id[isite] = s->spec_site_id[off + is];
            if (id[isite] < 1) {        /* rdf[0] is never allocated (rdf.c:80) */
               status = ORC_ERR_MAP;
               goto done;
            }
         }
   }
   for (ic = 0; ic < c.ncells; ic++) {
      int ix = ic / (c.ny * c.nz), iy = ic / c.nz - c.ny * ix, iz = ic - c.nz * (iy + c.ny * ix);
      long n = cell_site_list(&c, ix, iy, iz, NULL, NULL, NULL, 0);
      if (n > maxlist)
         maxlist = n;
   }
   maxlist += 1;
   nab = (int *)malloc(sizeof(int) * (size_t)maxlist);
   pbc = (int *)malloc(sizeof(int) * (size_t)maxlist);
   rsq = (double *)malloc(sizeof(double) * (size_t)maxlist);
   if (!nab || !pbc || !rsq) {
      status = ORC_ERR_ALLOC;
      goto done;
   }
   k = 0;
   for (ii = -1; ii <= 1; ii++)
      for (jj = -1; jj <= 1; jj++)
         for (kk = -1; kk <= 1; kk++) {
            reloc[k][0] = h[0] * ii + h[1] * jj + h[2] * kk;
            reloc[k][1] = h[3] * ii + h[4] * jj + h[5] * kk;
            reloc[k][2] = h[6] * ii + h[7] * jj + h[8] * kk;
            k++;
         }
   for (ic = 0; ic < c.ncells; ic++) {
      int ix, iy, iz, jmin, jmax, e, nnab, j;
      if (c.cell_start[ic] == c.cell_start[ic + 1])
         continue;
      ix = ic / (c.ny * c.nz); iy = ic / c.nz - c.ny * ix; iz = ic - c.nz * (iy + c.ny * ix);
      nnab = (int)cell_site_list(&c, ix, iy, iz, nab, pbc, nfnab, maxlist);
      jmin = 0;
      for (e = c.cell_start[ic]; e < c.cell_start[ic + 1]; e++) {
         int en = c.cell_entries[e], ft = c.entry_ftype[en], lim;
         if (ft) {
            jmin = 0;
            jmax = nfnab[0];
         } else
            jmax = nnab;
         lim = c.entry_isite[en] + c.entry_num[en];
         for (isite = c.entry_isite[en]; isite < lim; isite++) {
            double s0 = x[isite], s1 = y[isite], s2 = z[isite];
            if (!ft)
               jmin++;
            for (j = 0; j < jmax; j++) {
               int t = pbc[j];
               double ax = x[nab[j]] - s0 + reloc[t][0];
               double ay = y[nab[j]] - s1 + reloc[t][1];
               double az = z[nab[j]] - s2 + reloc[t][2];
               rsq[j] = ax * ax + ay * ay + az * az;
            }
            for (j = jmin; j < jmax; j++) {
               int bin = (int)(rbin * sqrt(rsq[j]));
               if (bin < nbins) {
                  float *b = rdf + (size_t)orc_rdf_block(s->max_id, id[isite], id[nab[j]]) * nbins;
                  b[bin] += invrho;
               }
            }
         }
      }
   }
done:
   free(id); free(nab); free(pbc); free(rsq);
   cells_free(&c);
   return status;
}
