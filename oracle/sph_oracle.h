/*
 * Test infrastructure only — the oracle. A plain-C FP64 restatement of the reference solver's
 * per-step particle loop (SPH_Calculation, CPP_1-SPHLoops.cpp:478-760), used by tests/, by
 * __graft_entry__.smoke() and by bench.py's CPU-baseline leg as the checker. Nothing in the
 * product links or calls it.
 *
 * Layout follows the reference: an array of particle records, a per-particle neighbour table
 * with 99 kernel-value slots, merge-sort binning, OpenMP static loops. With density mode
 * OR_DENSITY_GS and one thread it reproduces the reference's serial run bit for bit (pinned
 * against the reference link harness, oracle/_ref); OR_DENSITY_JACOBI evaluates the density
 * normalisation from the pre-normalisation densities of all neighbours, the order-free form the
 * GPU computes.
 */
#ifndef SPH_ORACLE_H
#define SPH_ORACLE_H

#include "../include/sphg.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_DENSITY_GS = 0, OR_DENSITY_JACOBI = 1 };

typedef struct or_ctx or_ctx;

or_ctx *or_create(const sphg_problem *pro, const sphg_materials *mat, int n);
void or_destroy(or_ctx *c);
int or_set_field(or_ctx *c, const char *name, const double *v, int n, int ncomp);
int or_get_field(or_ctx *c, const char *name, double *v, int n, int ncomp);
int or_set_field_i(or_ctx *c, const char *name, const int *v, int n);
int or_get_field_i(or_ctx *c, const char *name, int *v, int n);
int or_setting_vis_cons(or_ctx *c);
int or_upload_done(or_ctx *c);
void or_set_mode(or_ctx *c, int density_mode, int nthreads);
void or_set_step(or_ctx *c, int l);
/* Para_Pro::t, the clock the vibration loads read (advanced by dt at the start of every step). */
void or_set_time(or_ctx *c, double t);
/* Para_Vibra and vibra.dat (flv 1, 2): loads[nsteps][3]. */
int or_set_vibration(or_ctx *c, double ttime, int nsteps, double sttime, double edtime, const double *loads);
/* Para_Rain (flr 1), after the particles are set. */
int or_set_rain(or_ctx *c, int rtype, int nrain, int nhor, double press_factor, double drop_velocity);
/* Para_EC (fleob 1): ecp[ecount][3] = start, end, step. */
int or_set_excavation(or_ctx *c, int ecount, const int *ecp);
/* partidel (CPP_15-2-partidel.cpp:16-50) once, outside the step. */
void or_partidel_now(or_ctx *c);
/* Runs nsteps reference steps; returns 0 or the reference error code (3, 5). */
int or_step(or_ctx *c, int nsteps);
/* Phase groups of one step (bit 1 NNPS, 2 density+eos1, 4 strain+eos2, 8 stress, 16 accel+velocity,
 * 32 xsph+boundary, 64 update); bits in increasing order make one step. */
int or_phase(or_ctx *c, int mask);
int or_get_neighbors(or_ctx *c, int *counts, int *ids, int n);
/* Sum over particles of ninflu in the last step (directed pairs). */
long long or_pair_count(or_ctx *c);

/* K-L random-field synthesis (kl_oracle.c, CPP_17-1_K-L_expansion_non.cpp): 200 modes. */
int or_kl_solve_wi(double wi[][3], const double a_x[3], const double dirac_r[3], int ndim);
void or_kl_solve_ramdai(double *eigval, const double wi[][3], const double dirac_r[3], int order_id[][3],
                        int ndim);
int or_kl_synthesize(int ndim, int n, const double *xp, const int *type, const double wi[][3],
                     const double *eigval, const int order_id[][3], const double *uu, const double x_trans[3],
                     const double a[3], int dist, double mean_value, double sd_value, double coe_depth,
                     double soil_max, double water_max, int field, double *value);
/* Correlated pair of SpatialVariables_Generate_2D_corr / _3D_corr (CPP_17-2_K-L_expansion_corr.cpp). */
void or_kl_corr_normals(int dist_a, double mean_a, double sd_a, double mean_b, double sd_b, double coe,
                        const double *u, double *xi_a, double *xi_b, double *coef_out);
/* random_variable_generate (CPP_17-5-Nonstationary-Correlated-RF.cpp:738-863). */
void or_random_variable(const int flags[4], double coe, const int dist_of[6], const double mean_of[6],
                        const double sd_of[6], const double zeta[2], int n, const int *type, double *vals);

#ifdef __cplusplus
}
#endif
/* Correlation-matrix random fields (rf_oracle.c, CPP_17-3-Nonstationary-RF.cpp). */
int or_rf_cells(int ndim, int n, const double *xp, const int *type, double cell_ndx, double dr,
                const double box_nb[6], int nc[3], double base[3], double *cell_dx_out, int *cell_total,
                int *cell_of);
void or_rf_corr(int ndim, int acf_type, const double rl[3], double beta, double beta_1, const int nc[3],
                const double base[3], double cell_dx, int total, double *C);
void or_rf_cholesky(const double *A, double *L, int n);
void or_rf_values(int ndim, int dist, int field, double mean, double std_dev, double coe_depth, double soil_max,
                  double water_max, const int nc[3], const double base[3], double cell_dx, int total, const double *L,
                  const double *zeta, int n, const int *type, const int *cell_of, double fill, double *E_out,
                  double *value);

void or_rf_pair_values(int ndim, int phase, int dist, double coe, const double mean[2], const double sd[2],
                       const double cdep[2], double soil_max, double water_max, const int nc[3], const double base[3],
                       double cell_dx, int total, const double *L, const double *zeta_1, int n, const int *type,
                       const int *cell_of, double fill, double *out_fai, double *out_c);

#endif
