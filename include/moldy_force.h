/*
 * moldy_force.h -- Moldy-signature drop-ins implemented by libmoldyforce.so.
 *
 * The three functions below have the exact C signatures of the reference's
 * force-evaluation boundary, so a Moldy build links them in place of
 * force.o and ewald.o:
 *
 *   force_calc  replaces src/force.c:1108  (called from src/accel.c:520)
 *   ewald       replaces src/ewald.c:280   (called from src/accel.c:526)
 *   poteval     replaces src/force.c:632   (called from src/accel.c:323)
 *
 * The struct types restate the memory layouts of src/structs.h (contr_mt
 * :27-85, system_mt :87-118, spec_mt :121-145, pot_mt :156-161) and src/defs.h
 * (real = double, NPOTP = 8); they must stay byte-identical to Moldy's for
 * the link to work, so field order and types follow the reference exactly.
 * The shim reads Moldy's globals `control`, `ithread`, `nthreads` and reports
 * through Moldy's note()/message() (src/output.c:131-187); when those are not
 * provided (a test harness), weak local definitions are used.
 *
 * Precision: environment variable MOLDY_AMD_PRECISION = fp64 (default) or
 * mixed.  Device: MOLDY_AMD_DEVICE (default 0).
 */
#ifndef MOLDY_FORCE_H
#define MOLDY_FORCE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef double real;
typedef real vec_mt[3];
typedef vec_mt *vec_mp;
typedef real quat_mt[4];
typedef quat_mt *quat_mp;
typedef real mat_mt[3][3];
typedef vec_mt *mat_mp;
typedef int boolean;

#define MOLDY_L_NAME 128
#define MOLDY_L_SPEC 32
#define MOLDY_NPOTP 8

typedef struct {
   char title[MOLDY_L_NAME];
   long istep, nsteps;
   double step;
   boolean print_sysdef, new_sysdef, molpbc, reset_averages;
   int scale_options;
   boolean surface_dipole, lattice_start;
   char sysdef[MOLDY_L_NAME], restart_file[MOLDY_L_NAME], save_file[MOLDY_L_NAME],
       dump_file[MOLDY_L_NAME], backup_file[MOLDY_L_NAME], temp_file[MOLDY_L_NAME];
   int spare[20];
   boolean nosymmetric_rot;
   double ewald_accuracy;
   double ttmass, rtmass;
   int const_pressure;
   int const_temp;
   boolean xdr_write, strict_cutoff;
   int strain_mask;
   int nbins;
   unsigned long seed;
   int page_width, page_length;
   long scale_interval, scale_end, begin_average, average_interval, begin_dump, dump_offset,
       dump_interval;
   int dump_level, maxdumps;
   long backup_interval, roll_interval, print_interval, begin_rdf, rdf_interval, rdf_out;
   double temp, pressure, pmass, cutoff, subcell, density, alpha, k_cutoff, limit, cpu_limit;
} contr_mt;

typedef struct {
   int nsites, nmols, nmols_r, nspecies, max_id, d_of_f;
   int ptype, n_potpar;
   vec_mp c_of_m, mom, momp;
   quat_mp quat, amom, amomp;
   mat_mp h, hmom, hmomp;
   real ts, tsmom;
   real H_0;
   real rs, rsmom;
} system_mt, *system_mp;

typedef struct {
   real inertia[3], mass, dipole, charge;
   int nsites, nmols;
   int rdof, framework;
   char name[MOLDY_L_SPEC];
   int *site_id;
   vec_mp p_f_sites;
   vec_mp c_of_m, mom, momp;
   quat_mp quat, amom, amomp;
} spec_mt, *spec_mp;

typedef struct {
   int flag;
   int pad;
   real p[MOLDY_NPOTP];
} pot_mt, *pot_mp;

/* The boundary (signatures of src/accel.c:74-81 and src/ewald.c:280-286). */
void force_calc(real **site, real **site_force, system_mt *system, spec_mt *species, real *chg,
                pot_mt *potpar, double *pe, mat_mt stress);
void ewald(real **site, real **site_force, system_mp system, spec_mt *species, real *chg,
           double *pe, real (*stress)[3]);
double poteval(real *potpar, double r, int ptype, double chgsq);

/* Shim control (no counterpart in the reference): release the engine. */
void moldy_amd_release(void);

#ifdef __cplusplus
}
#endif
#endif
