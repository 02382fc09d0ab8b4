/*
 * moldy_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement (plain C, FP64, single thread) of Moldy's force-evaluation
 * path: link-cell binning, neighbour-cell lists, the site-site kernel, the
 * Ewald reciprocal sum and the eval_forces() glue.  It is the checker for the
 * HIP engine and the CPU timing baseline.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product never links it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference tree, e.g. src/force.c:119).
 */
#ifndef MOLDY_ORACLE_H
#define MOLDY_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NPOTP 8

/* Potential type codes, src/kernel.c:46-59 */
enum { ORC_LJ = 0, ORC_E6 = 1, ORC_MCY = 2, ORC_GEN = 3, ORC_HIW = 4,
       ORC_RSRVD = 5, ORC_MORSE = 6 };

/* Status codes (0 = ok).  Moldy prints these as *E* / *F* messages. */
enum {
   ORC_OK = 0,
   ORC_ERR_BIN_RANGE = 1,     /* "Co-ordinate out of range in BIN"  force.c:130 */
   ORC_ERR_BIN_ROUND = 2,     /* "Rounding problem in BIN"          force.c:134 */
   ORC_ERR_CUTOFF    = 3,     /* CUTOFF fatal                       force.c:206 */
   ORC_ERR_TONAB     = 4,     /* TONAB fatal                        force.c:562 */
   ORC_ERR_MULTIFW   = 5,     /* multiple frameworks                force.c:1219*/
   ORC_ERR_PTYPE     = 6,     /* UNKPTY                             kernel.c:186*/
   ORC_ERR_MAP       = 7,     /* strict map index outside the map (UB in ref) */
   ORC_ERR_ALLOC     = 8,
   ORC_ERR_QNORM     = 9      /* QNORM2 fatal                       leapfrog.c:104 */
};

/*
 * Flat description of the system (the content of system_mt, spec_mt[],
 * pot_mt[] and the contr_mt fields the path reads; src/structs.h:27-161).
 * Species are ordered with frameworks last (src/input.c:95-117).
 */
typedef struct {
   int nsites, nmols, nspecies, max_id, ptype, n_potpar;
   const int    *spec_nmols;      /* [nspecies]                                */
   const int    *spec_nsites;     /* [nspecies]                                */
   const int    *spec_framework;  /* [nspecies]                                */
   const int    *spec_rdof;       /* [nspecies] rotational dof (0 = monatomic) */
   const int    *spec_site_id;    /* concatenated site ids, [sum spec_nsites]  */
   const double *spec_p_f_sites;  /* concatenated principal-frame sites [..][3]*/
   const double *site_charge;     /* [max_id] charge of each site id           */
   const double *potpar;          /* [max_id*max_id][ORC_NPOTP]                */
   /* contr_mt fields read on the path */
   double cutoff, subcell, alpha, k_cutoff;
   int strict_cutoff, molpbc, surface_dipole;
} orc_sys;

typedef struct {
   int nx, ny, nz, n_nabors, n_frame_types;
   long n_list_entries;      /* sum over cells of the site neighbour list length */
   long n_pairs;             /* pairs handed to the kernel                        */
   int  n_too_close;         /* TOOCLS warnings (force.c:944-946)                 */
   int  too_close_i[16], too_close_j[16];
   long n_nab_sites;         /* capacity the reference would allocate (force.c:1232) */
   double eintra;            /* intramolecular correction (force.c:1158-1168)     */
} orc_fc_info;

typedef struct {
   int hmax, kmax, lmax, nhkl, nsitesxf;
   double self_energy, sheet_energy;
} orc_ew_info;

/* ---- numeric helpers (src/matrix.c, src/auxil.c) ---- */
double orc_det(const double a[9]);
void   orc_invert(const double a[9], double b[9]);
double orc_err_fn(double x);
double orc_precision(void);
double orc_sum(int n, const double *x);

/* ---- binning, src/force.c:119-137, 426-486 ---- */
int orc_cellbin(double rc, int nc, double fnc, double eps, int *status);
/* Bins every entry (a site, or a whole molecule when molpbc).  Outputs per
 * entry: cell, first site, number of sites, frame type; and the cell lists in
 * the reference's linked-list order as CSR (cell_start[ncells+1],
 * cell_entries[nentries]).  Returns the number of entries or <0 on error. */
int orc_fill_cells(const orc_sys *s, const double h[9],
                   const double *x, const double *y, const double *z,
                   const double *c_of_m_frac, int nx, int ny, int nz,
                   int *entry_cell, int *entry_isite, int *entry_num,
                   int *entry_ftype, int *cell_start, int *cell_entries,
                   int *n_frame_types, int *status);

/* ---- neighbour-cell offsets, src/force.c:167-226 and 273-421 ----
 * Writes offsets as (dx,dy,dz) triplets; returns the count or <0 on error. */
int orc_neighbour_list(const double h[9], double cutoff, int nx, int ny,
                       int nz, int strict, int *offsets, int cap);

/* nx,ny,nz as force_calc chooses them (src/force.c:1153-1157). */
void orc_choose_cells(const orc_sys *s, const double h[9], int *nx, int *ny,
                      int *nz);

/* ---- site neighbour lists, src/force.c:521-569 ----
 * For every cell, the (nab, pbctrans) list in the reference's order.
 * list_start[ncells+1] (CSR), nab/pbctrans of capacity cap, nfnab[ncells][2]. */
long orc_site_lists(const orc_sys *s, const double h[9],
                    const double *x, const double *y, const double *z,
                    const double *c_of_m_frac, int strict_list, double list_cutoff,
                    long *list_start, int *nab, int *pbctrans, int *nfnab,
                    long cap, int *status);

/* ---- kernel, src/kernel.c:157-463 (one call = one jsite loop) ---- */
int orc_kernel(int jmin, int nnab, double *forceij, double *pe,
               const double *r_sqr, const double *nab_chg, double chg,
               double norm, double alpha, int ptype, double *const *pot);
double orc_poteval(const double *potpar, double r, int ptype, double chgsq,
                   double alpha);
double orc_dist_pot(const double *potpar, double cutoff, int ptype);

/* ---- the boundary: force_calc (src/force.c:1108-1320) ----
 * Accumulates into fx,fy,fz, *pe and the upper triangle of stress[9]. */
int orc_force_calc(const orc_sys *s, const double h[9],
                   const double *x, const double *y, const double *z,
                   const double *c_of_m_frac, const double *chg,
                   double *fx, double *fy, double *fz, double *pe,
                   double stress[9], orc_fc_info *info);

/* ---- the boundary: ewald (src/ewald.c:280-589) ---- */
int orc_ewald(const orc_sys *s, const double h[9],
              const double *x, const double *y, const double *z,
              const double *chg, double *fx, double *fy, double *fz,
              double *pe, double stress[9], orc_ew_info *info);
/* k-set of src/ewald.c:441-461 in the reference's order: hkl[n][3], kvec[n][3] */
int orc_kset(const double h[9], double k_cutoff, int *hkl, double *kvec, int cap,
             int *hmax, int *kmax, int *lmax);

/* ---- the caller: eval_forces (src/accel.c:398-617) ----
 * c_of_m_frac[nmols][3] scaled centre-of-mass coordinates, quat[nmols_r][4].
 * Outputs pe[2], dip_mom[3], stress[9] (full, symmetrised), force[nmols][3],
 * torque[nmols_r][3]; site_force (3 x nsites, optional) receives the
 * site forces after the surface-dipole term. */
int orc_eval_forces(const orc_sys *s, const double h[9],
                    const double *c_of_m_frac, const double *quat,
                    double pe[2], double dip_mom[3], double stress[9],
                    double *force, double *torque, double *site_xyz_out,
                    double *site_force_out, orc_fc_info *fci, orc_ew_info *ewi);

/* make_sites (src/algorith.c:169-217) for the whole system. */
void orc_make_sites(const orc_sys *s, const double h[9], const double *c_of_m_frac,
                    const double *quat, int molflag_all, double *x, double *y,
                    double *z);
/* distant_const (src/accel.c:293-329) */
double orc_distant_const(const orc_sys *s, double cutoff, int iflag);

/* ---- molecular dynamics: do_step (src/accel.c:626-827) with the leapfrog
 * splitting of src/leapfrog.c, at constant h (no barostat) ---- */
typedef struct {
   double step;                  /* control.step                              */
   double ttmass;                /* control.ttmass (thermostat mass)          */
   double pmass;                 /* control.pmass (enters H_0 via ke_cell)    */
   double gkt;                   /* sys->d_of_f * kB * control.temp           */
   int const_temp;               /* 0, or 1 = Nose-Poincare thermostat        */
   int nosymmetric_rot;          /* control.nosymmetric_rot: leapf_quat_a     */
   const double *spec_mass;      /* [nspecies]                                */
   const double *spec_inertia;   /* [nspecies][3] principal moments           */
   int const_pressure;           /* 0; 1 Parrinello-Rahman; 2 uniform dilation */
   double pressure;              /* control.pressure, program units           */
   int strain_mask;              /* control.strain_mask: bit 3i+j freezes h_ij */
} orc_dyn;

typedef struct {
   double *c_of_m;               /* [nmols][3] scaled centre-of-mass coords   */
   double *quat;                 /* [nmols_r][4]                              */
   double *mom;                  /* [nmols][3] scaled momenta                 */
   double *amom;                 /* [nmols_r][4] principal angular momenta    */
   double ts, tsmom, H_0;        /* thermostat s, its momentum, H_0           */
   int bad_quat;                 /* molecule of a QNORM failure               */
   double h[9], hmom[9];         /* cell matrix and its momentum (barostat)   */
} orc_md_state;

/* One do_step on st (its h is the cell; it changes only with a barostat).
 * pe[2] and stress[9] of the force evaluation; rec[26] (optional)
 * = translational KE, rotational KE, pe[0], pe[1], s, p_s, H_0, det(h),
 * hmom[9] and h[9] at the point where do_step evaluates H_0 (after the first
 * momentum half step). */
int orc_do_step(const orc_sys *s, const orc_dyn *d, orc_md_state *st, int init_H_0,
                double pe[2], double stress[9], double rec[26]);
int orc_symmetry_axis(const double *inertia);

/* ---- radial distribution functions: rdf_inner (src/force.c:1010-1103) and
 * rdf_accum (src/rdf.c:94-108).  Adds 1/density per site pair closer than
 * limit (strict neighbour list at limit over force_calc's cells) into
 * rdf[block(id_i, id_j)][nbins], float, blocks in init_rdf's order. ---- */
int orc_rdf(const orc_sys *s, const double h[9], const double *x, const double *y,
            const double *z, const double *c_of_m_frac, double limit, int nbins, float *rdf);
int orc_rdf_block(int max_id, int idi, int idj);

#ifdef __cplusplus
}
#endif
#endif
