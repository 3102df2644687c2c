/*
 * sphg.h — C-ABI of the MI355X-native per-step particle loop of the multi-phase SPH solver.
 *
 * Every entry point replaces one step function of the reference OpenMP solver; the citation
 * above each names the reference interface (file:line) it stands in for. All reference file
 * citations are into the wodexinxin2015/MultiPhase-SPH-Solver sources.
 *
 * Conventions
 *   - Plain C types only (no HIP or torch types). One context = one simulation (or one batch of
 *     realisations) resident in device memory, driven on one HIP stream the context owns.
 *   - Return value: 0 = OK; the reference's own Err_Deal codes (CPP_0-ErrorDeal.cpp:44-277) where
 *     the reference raises one: 3 = particle outside the cell grid (CPP_3-NNPS.cpp:250-257),
 *     5 = more than 99 neighbours (CPP_3-NNPS.cpp:313-319), 2 = no particles (CPP_1-SPHLoops.cpp:376);
 *     codes >= 100 are this library's own (SPHG_E_*).
 *   - Step functions are asynchronous on the context stream; device-side error codes accumulate
 *     in a sticky error word that sphg_sync() (and every sphg_step()) returns.
 *   - Contexts are independent; a context must not be used from two host threads at once.
 */
#ifndef SPHG_H_INCLUDED
#define SPHG_H_INCLUDED

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPHG_MAX_NB 99       /* neighbour cap of Par_Cell::influ (Class_NNPS.h:72, CPP_3-NNPS.cpp:313) */
#define SPHG_NSOIL 30        /* soil parameter slots: layer*10 + model (CPP_2-initialf.cpp:487-795) */

enum {
    SPHG_OK = 0,
    SPHG_E_NOPARTICLES = 2,
    SPHG_E_CELLRANGE = 3,
    SPHG_E_NEIGHBOURS = 5,
    SPHG_E_ARG = 100,        /* bad argument (null pointer, unknown field, size mismatch) */
    SPHG_E_HIP = 101,        /* HIP runtime error */
    SPHG_E_UNSUPPORTED = 102,/* flag combination this build does not implement */
    SPHG_E_CAPACITY = 103,   /* exceeds the context's capacity */
    SPHG_E_NOCONVERGE = 104, /* K-L root bisection cannot reach |f| <= 1e-6 in FP64 (the reference never returns) */
    SPHG_E_NONFINITE = 105   /* random-field Cholesky factor not finite: the correlation matrix is indefinite beyond
                                what the reference's clamp absorbs (the reference goes on with NaN properties) */
};

/* Arithmetic form of a context. SPHG_FP64 reproduces the reference's FP64 operations in their
 * order. SPHG_MIXED keeps all state and every neighbour sum in FP64 but evaluates distance,
 * kernel value and gradient of each pair in FP32 (separation formed in FP64 first) and the
 * divisions inside pair terms with the hardware reciprocal plus one Newton step. */
enum { SPHG_FP64 = 0, SPHG_MIXED = 1 };

#define SPHG_BUSY (-1)       /* sphg_poll: the context still has work in flight */

/* Problem parameters: Para_FL (Class_Problem.h:10-42), Para_Pro (:44-84), the free-surface
 * thresholds read with Para_Pro (CPP_2-initialf.cpp:75), Para_GF (:150-162), Para_Boundary
 * (:118-133), Para_Inter (:104-116), Para_SSInter (:176-192), Para_Model flags
 * (Class_ParticleProperty.h:10-28) and the Cell_Con domain box (Class_NNPS.h:12-36). */
typedef struct {
    int fla, fls, flr, fli, flv, fleob, flts, flspv, flknl, ntd, flbndy;
    int ndim, fop, inip, loop, ntotal, step_regu, dttype;
    double dr, dt, tt;
    double xsph[3];
    double art_vis[3][2];
    double art_stre_coe;
    double threshold_min, threshold_max;
    double gx, gy, gz;
    int if_cal, if_move;
    double coe_bndy, move_vx[3];
    double interc;
    int permtype;
    double pers, perl, perm;
    int ss_type;
    double ss_myu, ss_zeta;
    int ss_acc_type;
    int flagem, flagsb, flagdp, flagum, flagam, flagbui, flagpl, flagve, flagfocc, flagep;
    double alphats;
    /* Cell_Con: particle bounding box widened by 2*dr (CPP_2-initialf.cpp:976-999) */
    double xmin, xmax, ymin, ymax, zmin, zmax;
    double soil_min[3], soil_max[3], water_min[3], water_max[3];
} sphg_problem;

/* Para_Soil, field for field (Class_ParticleProperty.h:33-101). */
typedef struct {
    double e, poi, eps0, dens, damp1, damp2, cop, ramda, kapa, zmf, ann, bnn, beta;
    double sitas, sitar, epsr, epss, sd, sw, kse, c1, c2, c3, fai, c, ocr, sme0;
    double mzm, azm, r0, bl, bz, zet, vpu, porosity, fs, etao, g0, ds, damp_coe;
} sphg_soil;

/* Para_Fluid (Class_ParticleProperty.h:104-133). */
typedef struct {
    double vis, vpu, dens, paa0, porosity, fai, c;
    int bhtype, press_type, acc_type;
    double pre_coe, strain_min;
} sphg_fluid;

/* Para_Structure (Class_ParticleProperty.h:135-151). */
typedef struct {
    double dens, eps0, e, poi, vpu, porosity, cop;
} sphg_structure;

/* Material tables as initial_mat leaves them (CPP_2-initialf.cpp:325-852). */
typedef struct {
    sphg_soil soil[SPHG_NSOIL];
    sphg_fluid water[3];
    sphg_fluid air;
    sphg_structure structure;
} sphg_materials;

typedef struct sphg_ctx sphg_ctx;

/* Cell_Con domain box of one simulation (Class_NNPS.h:12-36). */
typedef struct {
    double xmin, xmax, ymin, ymax, zmin, zmax;
} sphg_box;

/* Context creation: capacity in particles and the arithmetic form (SPHG_FP64 or SPHG_MIXED).
 * Replaces the allocation block of SPH_Calculation (CPP_1-SPHLoops.cpp:354-378) and the
 * cell-list set-up celllist_ini1/celllist_ini2 (CPP_3-NNPS.cpp:96-171, called at :446-450). */
int sphg_create(sphg_ctx **out, const sphg_problem *pro, const sphg_materials *mat,
                int capacity, int precision);
/* Realisation batch (no counterpart in the reference, which runs one deck per process): nreal
 * simulations of one problem family (same flags, materials, dr, dt) with their own particles and
 * domain boxes in one context. Each keeps its own cell grid and check_domain box, so every
 * realisation computes exactly what it computes alone; one launch per kernel serves all of them.
 * Particles are uploaded realisation-major, counts[r] each. */
int sphg_create_batch(sphg_ctx **out, const sphg_problem *pro, const sphg_materials *mat, int nreal,
                      const int *counts, const sphg_box *boxes, int precision);
void sphg_destroy(sphg_ctx *ctx);
int sphg_device_bytes(const sphg_ctx *ctx, size_t *bytes);

/* Host <-> device state transfer, one named field at a time. Names are the Particle members
 * (Class_particle.h:13-103): x xp vx vxp ax axp ux strep divx interf interfss rho rhop mass hl
 * pre prep e eps deps sig dsig weps adps satu dsatu suct cop vsig veps divq divr meps roue rous
 * prew prea porosity cd poro_cop beta beta_ts type matype permtype etype; the per-particle
 * material copies pParti_ConsPara / pVFluid as "ps_<field>" / "pf_<field>" (ps_fai, ps_c, ps_cop and
 * ps_ds, the properties a random field varies, CPP_17-1_K-L_expansion_non.cpp:575-600, are kept
 * per particle; every other copy must equal its material-table value, or sphg_upload_done returns
 * SPHG_E_UNSUPPORTED); and the
 * boundary records clVar_Boundary vx_water vx_soil vx_air vx_struct (Class_particle.h:190-207)
 * and cl_bndy_pair vel_bndy (:168-177). Values are row-major [n][ncomp] (ncomp 1, 3, 6 or 9);
 * int fields (type matype permtype etype) use the *_i variants. sphg_upload_done() derives the
 * per-particle material slots and must follow the uploads before the first step. */
int sphg_set_field(sphg_ctx *ctx, const char *name, const double *host, int n, int ncomp);
int sphg_get_field(sphg_ctx *ctx, const char *name, double *host, int n, int ncomp);
int sphg_set_field_i(sphg_ctx *ctx, const char *name, const int *host, int n);
int sphg_get_field_i(sphg_ctx *ctx, const char *name, int *host, int n);
int sphg_upload_done(sphg_ctx *ctx, int n);

/* The reference's particle record, member for member and in the same order, so that an array of
 * the reference's Particle (Class_particle.h:13-103, 131 doubles and 4 ints, 1064 bytes) can be
 * passed where an array of sphg_particle is expected. */
typedef struct {
    double x[3], xp[3], vx[3], vxp[3], ax[3], axp[3], ux[3], strep[3];
    double divx[3][3];
    double interf[3], interfss[3];
    double rho, rhop, mass, hl, pre, prep, e;
    double eps[6], deps[6], sig[6], dsig[6], ff_sig[3];
    double weps[3][3];
    double adps[6], tssig[6];
    double tsqc, satu, dsatu, suct, cop, vsig, veps, divq, divr, meps;
    double roue, rous, zeta, rsta, prew, prea, porosity, cd, poro_cop;
    double beta[3][3], beta_ts[3][3];
    int type, matype, permtype, etype;
} sphg_particle;

/* Groups of Particle members for sphg_download_particles. */
enum {
    SPHG_PART_KINEMATICS = 1,   /* x xp vx vxp ax axp ux */
    SPHG_PART_DENSITY = 2,      /* rho rhop mass hl pre prep prew prea e porosity poro_cop */
    SPHG_PART_STRESS = 4,       /* eps deps sig adps strep vsig veps divq divr meps weps beta beta_ts tssig tsqc */
    SPHG_PART_STATE = 8,        /* satu dsatu suct cop roue cd interf interfss */
    SPHG_PART_IDS = 16,         /* type matype permtype etype */
    SPHG_PART_ALL = 31
};

/* Upload of the host particle array as SPH_Calculation holds it (pPar, allocated at
 * CPP_1-SPHLoops.cpp:354-378 and filled by input_par / setting_par): every member the device keeps
 * (all but divx dsig ff_sig weps rous zeta rsta, which the reference recomputes each step or this
 * library does not use; tssig / tsqc only where the TS transform is on), the same as one
 * sphg_set_field / sphg_set_field_i call per member. sphg_upload_done() follows, as after
 * sphg_set_field. The per-particle material copies (pParti_ConsPara, pVFluid) keep their
 * "ps_<field>" / "pf_<field>" fields. */
int sphg_upload_particles(sphg_ctx *ctx, const sphg_particle *p, int n);
/* Download into the host particle array (what output_total, CPP_14-2-output.cpp:15-135, and the
 * reference's other writers read): the members of the groups in `mask` that the device keeps
 * (weps and beta_ts derived as sphg_get_field derives them); members it does not keep are left
 * as they are. */
int sphg_download_particles(sphg_ctx *ctx, sphg_particle *p, int n, unsigned mask);

/* Host-side initial state: setting_vis_cons + setting_par (CPP_2-initialf.cpp:1021-1380) on host
 * arrays (row-major, n particles): fills rho rhop mass hl satu e porosity axp cd roue sig pre prep
 * suct and the per-particle material copies; pro->nboun etc. are not needed. Reads xp, type,
 * matype, etype. Runs on the host only (no device needed). */
int sphg_setting_par(const sphg_problem *pro, const sphg_materials *mat, int n,
                     const double *xp, const int *type, const int *matype, const int *etype,
                     double *rho, double *rhop, double *mass, double *hl, double *satu, double *e,
                     double *porosity, double *axp, double *cd, double *roue, double *sig,
                     double *pre, double *prep, double *suct);

/* ---- step functions: same order and semantics as the reference loop body ---------------- */

/* clNNPS_Fun::parttocell (CPP_3-NNPS.cpp:175-263) incl. Binary_Sort (:597-630). */
int sphg_parttocell(sphg_ctx *ctx);
/* clNNPS_Fun::partisearching with weight() (CPP_3-NNPS.cpp:265-374, :17-48). */
int sphg_partisearching(sphg_ctx *ctx);
/* clNNPS_Fun::kernel_gradient_correction (CPP_3-NNPS.cpp:376-559), called after the search when
 * flknl > 2 (CPP_1-SPHLoops.cpp:544-546; a no-op otherwise): the correction matrix L^-1 of every
 * soil and structure particle, which the later passes apply to those particles' kernel gradients.
 * sphg_step and the graphs include it. */
int sphg_kernel_gradient_correction(sphg_ctx *ctx);
/* clDensity_Fun::density (CPP_6-density.cpp:11-189). The normalisation pass reads every
 * neighbour's pre-normalisation density (Jacobi form), where the reference's serial loop reads
 * the already normalised densities of lower-index neighbours in place (:168-188). */
int sphg_density(sphg_ctx *ctx);
/* clStraStre_Fun::strain_noreg (CPP_7-1-strain.cpp:20-220) fused with the pair sums of eos loop 2
 * (CPP_7-1-eos_CR.cpp:45-177) and of free_surface_check (CPP_7-2-constitutive.cpp:453-480);
 * strain_regu every step_regu steps (CPP_7-2-constitutive.cpp:419-450, Jacobi form),
 * saturation + moisturesuct when flagum == 1 (CPP_7-1-strain.cpp:222-280) and cal_permeability
 * when fli == 1 (CPP_7-2-constitutive.cpp:355-382). eos loop 1 (CPP_7-1-eos_CR.cpp:20-43) runs
 * at the end of sphg_density. */
int sphg_strain(sphg_ctx *ctx);
/* clStraStre_Fun::strain_to_stress (CPP_7-2-constitutive.cpp:13-321; models elastic, SB, DP,
 * UMCC, AMCC, Bui DP, FOCC: CPP_8-1/8-2/8-4/8-5) + stress_regu every step_regu steps (:385-418,
 * Jacobi form) + free-surface stress correction (:481-505) + Stress_eigValue
 * (CPP_7-1-strain.cpp:283-332) + the per-particle tensor of artificial_stress_2d/3d
 * (CPP_9-5-artificialstress.cpp:42-89, :161-227); with flts >= 1 the TS transform
 * TSSMP_before / TSSMP_after around the soil models (CPP_7-3-TS.cpp:12-106), its state readable
 * as the fields "tssig" (6) and "tsqc". */
int sphg_strain_to_stress(sphg_ctx *ctx);
/* clAcce_Fun fluid/soil/structure accelerations (CPP_9-4-derivation_stress.cpp:12-505),
 * artificial_viscosity (CPP_9-3-artificialvis.cpp:12-99), the artificial-stress pair sum
 * (CPP_9-5-artificialstress.cpp:92-140, :230-317), boundary stress effect
 * (CPP_9-4-derivation_stress.cpp:631-721), interact1 / interact2 / interact3 by fli 1 / 2 / 3
 * (CPP_10-interact1.cpp:15-110, :115-267, :270-420; CPP_1-SPHLoops.cpp:649-660),
 * soil_damping_geostress (CPP_9-4-derivation_stress.cpp:850-897) and veocity_update for
 * if_move 0-3 (CPP_11-xsph.cpp:11-142). */
int sphg_accel(sphg_ctx *ctx);
/* clVel_Fun::xsph (CPP_11-xsph.cpp:145-256; skipped when if_move == 2, CPP_1-SPHLoops.cpp:714)
 * + velocity_set_domain (CPP_13-boundary.cpp:71-122) + tm_boundary_2d/3d first loop
 * (CPP_13-boundary.cpp:179-283, :388-520). */
int sphg_xsph(sphg_ctx *ctx);
/* tm_boundary_2d/3d second loop (CPP_13-boundary.cpp:284-386) + check_domain (:20-68) +
 * clUpdate_Fun::update (CPP_12-1-update.cpp:11-55). */
int sphg_update(sphg_ctx *ctx);

/* One whole reference step (CPP_1-SPHLoops.cpp:478-760) nsteps times; advances pro.l. The step
 * sequence is captured once into a hipGraph and replayed. Returns the sticky error word. */
int sphg_step(sphg_ctx *ctx, int nsteps);
int sphg_sync(sphg_ctx *ctx);
/* The time loop of SPH_Calculation (CPP_1-SPHLoops.cpp:478-770) over the context: for
 * l = current step .. loop, while t < tt, t += dt and one step; output after step 1 and after
 * every fop-th step (CPP_1-SPHLoops.cpp:743), and once more, then stop, when t has reached tt
 * (:761-767). Steps between two outputs run as one batch of graph replays. `out` (may be null)
 * gets the step number and time with the state on the device complete; a nonzero return stops
 * the loop and is returned. A step error (the reference's Err_Deal codes 2/3/5, ...) ends the loop
 * and is returned; the state is then that of the end of the batch. *t is the simulated time in
 * and out (pPPro->t). */
typedef int (*sphg_output_fn)(void *user, int step, double t);
int sphg_run(sphg_ctx *ctx, int loop, int fop, double tt, double *t, sphg_output_fn out, void *user);
/* Serving (no counterpart in the reference, which runs one simulation per process): enqueue
 * nsteps steps on the context's stream and return at once; sphg_poll returns SPHG_BUSY while they
 * run, then the sticky error word. Contexts on one device overlap their work. */
int sphg_step_async(sphg_ctx *ctx, int nsteps);
int sphg_poll(sphg_ctx *ctx);
int sphg_current_step(const sphg_ctx *ctx);
/* Sets the step counter pPPro->l for the next step (CPP_1-SPHLoops.cpp:478). */
int sphg_set_step(sphg_ctx *ctx, int l);
/* The clock Para_Pro::t (CPP_1-SPHLoops.cpp:481-483): every step first advances it by dt; the
 * vibration loads read it. Zero after sphg_create; checkpoints keep it. */
int sphg_set_time(sphg_ctx *ctx, double t);
int sphg_get_time(sphg_ctx *ctx, double *t);
/* Time step (timestep, CPP_12-2-dt.cpp:13-20: the constant dt = 0.25 dr / 600 that setting_par
 * stores in Para_Pro::dt). SPHG_DT_CONSTANT (default): every step uses sphg_problem.dt, as the
 * reference does. SPHG_DT_CFL: before every step the device computes
 *   dt = min(c_cour dr / (c_wave + max_i |v_i|), 0.25 sqrt(dr / max_i |a_i|), dt_max)
 * in FP64 as a min-reduction over the particles (v, a: the velocity and acceleration of the last
 * update; type-7 particles left out), which is the reference's constant at rest with c_cour 0.25
 * and c_wave 600. The clock Para_Pro::t advances by each step's dt; sphg_run then takes one step
 * at a time. SPHG_E_ARG for another mode or non-positive coefficients. sphg_get_dt returns the dt
 * of the last step (before any step: sphg_problem.dt). */
#define SPHG_DT_CONSTANT 0
#define SPHG_DT_CFL 1
int sphg_set_dt_mode(sphg_ctx *ctx, int mode, double c_cour, double c_wave, double dt_max);
int sphg_get_dt(sphg_ctx *ctx, double *dt);
/* Vibration loads (flv 1, 2; Para_Vibra, Class_Problem.h:133-149, read by initial_pro's VL section
 * and initial_vib from vibra.dat, CPP_2-initialf.cpp:250-265, :293-322): one period of ttime in
 * nsteps loads loads[nsteps][3], applied while sttime <= t <= edtime. With flv 1 they are the
 * acceleration of the matype-3 / 5 walls (vibration_acceleration), with flv 2 the velocity of the
 * matype-4 / 5 walls (vibration_velocity, CPP_9-1-acceleration_vibration.cpp:11-164); flv > 0 also
 * adds the free-field boundaries (CPP_19-FreeFieldParticles.cpp, CPP_19-FreeFieldParticles_soil.cpp)
 * and soil_damping_vibration. Required before sphg_upload_done when flv > 0 (SPHG_E_ARG otherwise,
 * where the reference stops for a missing vibra.dat). */
int sphg_set_vibration(sphg_ctx *ctx, double ttime, int nsteps, double sttime, double edtime, const double *loads);
/* The VL section of <dir>/input.dat and the loads of <dir>/vibra.dat, as initial_pro and initial_vib
 * read them (CPP_2-initialf.cpp:250-265, :293-322). loads may be null (parameters only) or must hold
 * 3 * cap values (SPHG_E_CAPACITY when cap < nsteps). SPHG_E_ARG for a missing file or section, or
 * nsteps < 1. Host only. */
int sphg_read_vibration(const char *dir, double *ttime, int *nsteps, double *sttime, double *edtime, double *loads,
                        int cap);
/* Particles in the context (pPPro->ntotal, all realisations of a batch), -1 for a null context. */
int sphg_particle_count(const sphg_ctx *ctx);

/* ---- Deck and result files, checkpoints --------------------------------------------------- */

/* initial_pro + initial_mat + input_par (CPP_2-initialf.cpp:22-1018): reads an input.dat deck
 * into the problem and material structs (constructor defaults, then the FL PR XP AT GF IP BP WP
 * AP MS <soil model blocks> SM sections, each found from the top of the file as the reference
 * finds it, then the reference's material post-processing: vpu 600, porosities, default
 * densities, FOCC e) and the particle arrays of the PA section. Host only. Called with any array
 * null it fills pro, mat and *n (pPPro->ntotal) and stops; called again with arrays of cap >= *n
 * particles (xp, vxp: 3 per particle, unused dimensions 0) it reads the particles and fills the
 * Cell_Con box and phase extents of pro. A section that is absent, or that ends (or reaches the
 * end of the file, where the reference would scan forever) before its values, leaves the
 * defaults. Returns 2 when ntotal <= 0 (the reference's error 2), SPHG_E_CAPACITY when
 * cap < ntotal, SPHG_E_ARG for a file that cannot be opened, a missing PA section or a particle
 * row that does not parse (the reference would go on with zeroed particles), or more than three
 * layers of one soil model (beyond the 30 slots). The rain, excavation, vibration and
 * spatial-variability sections belong to flags this library does not run and are not read. */
int sphg_read_input(const char *path, sphg_problem *pro, sphg_materials *mat, int *n, int cap,
                    double *xp, double *vxp, int *type, int *matype, int *permtype, int *etype);

/* Per-particle fields of one .odb table, host arrays of n particles, row-major (3 or 6 values
 * per particle where the Particle member is a vector, Class_particle.h:18-83). */
typedef struct {
    const double *xp, *ux, *ax, *vx, *interf, *interfss, *strep; /* 3 per particle */
    const double *sig, *eps, *adps;                               /* 6 per particle */
    const double *pre, *prea, *prew, *vsig, *veps, *meps, *divq, *divr, *satu, *suct;
    const int *type, *matype;
} sphg_odb_fields;
/* output_total (CPP_14-2-output.cpp:15-135): writes <dir>/data-<l, 10 digits>.odb, the rule
 * line, "l t", one row per particle (records numbered from 1; boundary rows carry interfss and
 * interf where other rows carry ax and vx), the rule line. Host only. */
int sphg_write_odb_fields(const char *dir, int l, double t, int n, const sphg_odb_fields *f);
/* The same table from the context's current state, particles first .. first+count-1 (one
 * realisation of a batch; 0 and the particle count for a single simulation). l and t label the
 * file as the reference's output_total call does (pPPro->l, pPPro->t at the output step). */
int sphg_write_odb(sphg_ctx *ctx, const char *dir, int l, double t, int first, int count);
/* One project directory run the way the reference's executable runs it (main and
 * SPH_Calculation, CPP_0-main.cpp, CPP_1-SPHLoops.cpp:67-782): <dir>/input.dat read with
 * sphg_read_input, the initial state of sphg_setting_par, one context of the given arithmetic
 * form, the time loop of sphg_run, and <dir>/data-<l>.odb written at the reference's output steps
 * (and, after a step error, for the state the run stopped in). Returns the run's error code and
 * the last step written in *last_step (may be null). Random-field decks (flspv >= 1) return
 * SPHG_E_UNSUPPORTED: the K-L entry points serve them through the API. */
int sphg_solve_project(const char *dir, int precision, int *last_step);
/* Checkpoint and resume (no counterpart in the reference, which restarts from input.dat only):
 * sphg_save_state writes every device array of the context except the neighbour lists, which the
 * next step's search rebuilds, together with the step counter and the buffer parity;
 * sphg_load_state restores them into a context created and uploaded from the same deck (same
 * flags, capacity, particle count and arithmetic form; SPHG_E_ARG otherwise). Stepping a
 * restored context gives bit for bit the results of the uninterrupted run. */
int sphg_save_state(sphg_ctx *ctx, const char *path);
int sphg_load_state(sphg_ctx *ctx, const char *path);

/* Parity tap: neighbour lists in reference order (Par_Cell::ninflu / influ[1..], Class_NNPS.h:65-77).
 * counts[n]; ids[n][SPHG_MAX_NB] (entries past counts[i] are -1). Counts above 99 are reported
 * as the true count (the reference keeps counting past the cap before raising error 5). */
int sphg_get_neighbors(sphg_ctx *ctx, int *counts, int *ids, int n);

/* Measurement tap with no counterpart in the reference: reversed = 1 reverses every neighbour list
 * right after the search, so every pair sum of the following steps runs in the opposite order (and
 * sphg_get_neighbors reports the reversed lists); 0 restores the reference order. The distance
 * between the two FP64 runs is the noise floor of the summation order. Meant for decks whose
 * results depend on the list only through its sums (not the soil-structure contact or Takeda-Morris
 * wall searches, which pick the first listed structure or wall particles). */
int sphg_set_list_order(sphg_ctx *ctx, int reversed);

/* Per-kernel timing of the last sphg_step call set (HIP events on the context stream):
 * fills up to cap (name, total ms, launches) rows, returns the row count. */
int sphg_kernel_times(sphg_ctx *ctx, char (*names)[32], double *ms, int *launches, int cap);
int sphg_enable_timing(sphg_ctx *ctx, int on);
/* Autotuner (no counterpart in the reference): the density, normalisation, strain and XSPH
 * passes load 1-6 neighbours per memory round trip, and the best width depends on the deck; the
 * search stages the first 16 or 32 entries of each list in LDS before writing them out, and the
 * better depth depends on the lists' length and the deck's size. Each width set (with a staging
 * depth) is timed over `steps` steps from the current state, the state is restored after each
 * trial, and each pass keeps its fastest choice (chosen[4]: density, normalisation, strain, XSPH;
 * may be null; sphg_get_search_stage reports the depth). Results do not depend on these choices
 * (bit for bit). Allocates a copy of the device state for the duration of the call. */
int sphg_autotune(sphg_ctx *ctx, int steps, int *chosen);
/* Sets the widths directly (density 1/2/4, normalisation 2/4/6, strain 1/2/4, XSPH 1/2/4);
 * SPHG_E_ARG for any other value. */
int sphg_set_batch_widths(sphg_ctx *ctx, const int *widths);
/* The search's list staging depth directly (16 or 32; default 16 in 2D, 32 in 3D); SPHG_E_ARG for
 * any other value. */
int sphg_set_search_stage(sphg_ctx *ctx, int stage);
int sphg_get_search_stage(const sphg_ctx *ctx, int *stage);

/* ---- K-L random fields (CPP_17-1_K-L_expansion_non.cpp) ---------------------------------- */

#define SPHG_KL_TERMS 200    /* klterm (CPP_17-1_K-L_expansion_non.cpp:20) */

/* One spatially variable property (Para_SpatialVariability, Class_spatial_variability.h:44-66)
 * plus the K-L modes that sphg_kl_modes() fills in. */
typedef struct {
    int ndim;
    int dist;                /* 1 normal, 2 lognormal */
    int field;               /* 0 soil fai, 1 soil c, 2 soil cop, 3 soil ds, 4 fluid fai, 5 fluid c */
    double mean_value, sd_value, coe_depth;
    double rl[3];            /* correlation lengths rl_value */
    double a[3], x_trans[3]; /* K-L domain [-a, a] about x_trans */
    double soil_max, water_max;  /* depth-trend references (Cell_Con soil_max / water_max, y or z) */
    double wi[SPHG_KL_TERMS][3];
    double eigval[SPHG_KL_TERMS];
    int order[SPHG_KL_TERMS][3];
} sphg_kl_field;

/* K-L domain of SpatialVariables_Generate_2D/3D (CPP_17-1_K-L_expansion_non.cpp:72-90, :175-191):
 * the extents of the non-boundary particles widened by 20 % per side (input_par's *_nb box,
 * CPP_2-initialf.cpp:915-1003); returns 13 when a half-width is zero, as the reference. */
int sphg_kl_domain(int ndim, int n, const double *xp, const int *type, double a[3], double x_trans[3]);
/* solve_wi + solve_ramdai (CPP_17-1_K-L_expansion_non.cpp:268-438): roots by bisection in FP64
 * on the host, eigenvalue products, the 200 largest in the reference's order. Needs ndim, a, rl. */
int sphg_kl_modes(sphg_kl_field *f);
/* solve_h_2D / solve_h_3D (CPP_17-1_K-L_expansion_non.cpp:499-713) on the device for the
 * context's particles (FP64), with the standard normals xi[200] supplied by the caller (the
 * reference draws them from a clock-seeded generator, :24-34). The result becomes the per-particle
 * property f->field of the field's phase (0 soil fai, 1 c, 2 cop, 3 ds: soil particles; 4 fluid fai,
 * 5 c: water particles), which the step reads in place of the material slot's value; the other
 * particles keep their slot values. */
int sphg_kl_synthesize(sphg_ctx *ctx, const sphg_kl_field *f, const double *xi);
/* The same for the particles [first, first + count) only: one realisation of a batch context
 * (no counterpart in the reference, which holds one realisation per run). */
int sphg_kl_synthesize_part(sphg_ctx *ctx, const sphg_kl_field *f, const double *xi, int first, int count);
/* K-L expansion with correlated friction angle and cohesion (flspv 2,
 * SpatialVariables_Generate_2D_corr / _3D_corr, CPP_17-2_K-L_expansion_corr.cpp:55-473): the
 * standard normals of the two fields are u[200][2] (the reference draws them from a clock-seeded
 * generator, :39-52) mixed by [[1, rho], [0, sqrt(1 - rho^2)]] with rho = coe_frict_cohesion, or for
 * a lognormal f_a the correlation of the underlying normals (:114-140); xi_a / xi_b receive the
 * mixed normals and coefficient (may be null) rho. Host only. */
int sphg_kl_corr_normals(const sphg_kl_field *f_a, const sphg_kl_field *f_b, double coe_frict_cohesion,
                         const double *u, double *xi_a, double *xi_b, double *coefficient);
/* Both fields of the pair into the context's per-particle properties: f_a.field 0 and f_b.field 1
 * (soil fai, c) or 4 and 5 (fluid fai, c), each with its own modes, mean, sd and distribution (the
 * reference takes the mixing coefficient's distribution from f_a). The K-L domain is that of
 * sphg_kl_domain (the same 20 % widening as flspv 1). SPHG_E_ARG for any other field pair. */
int sphg_kl_synthesize_corr(sphg_ctx *ctx, const sphg_kl_field *f_a, const sphg_kl_field *f_b,
                            double coe_frict_cohesion, const double *u);
/* The same for the particles [first, first + count) only (one realisation of a batch). */
int sphg_kl_synthesize_corr_part(sphg_ctx *ctx, const sphg_kl_field *f_a, const sphg_kl_field *f_b,
                                 double coe_frict_cohesion, const double *u, int first, int count);

/* Realisation batch in a common K-L domain: for nreal realisations xi[nreal][200] over the same
 * particles, H = Phi (n x 200) * Xi (200 x nreal) as one FP32 matrix-core GEMM, then the field
 * values value[nreal][n] (FP64 from the FP32 H; particles of other types keep fill). gemm_ms
 * (may be null) receives the GEMM kernel time. */
int sphg_kl_synthesize_batch(const sphg_kl_field *f, int n, const double *xp, const int *type, int nreal,
                             const double *xi, double fill, double *value, double *gemm_ms);

/* The same synthesis as sphg_kl_synthesize_batch with no device allocation: no counterpart in the
 * reference; for callers that cannot allocate (code under stream capture, tight serving loops,
 * harnesses that forbid hipMalloc). All device buffers live in one static array of about 32 MiB per
 * GPU, held until the process ends, through which the particles pass 8192 at a time; nreal <= 256
 * (SPHG_E_CAPACITY beyond). Results equal sphg_kl_synthesize_batch's bit for bit. Runs on the
 * default stream of the current device; calls that overlap on one GPU share the array and race. */
int sphg_kl_synthesize_batch_noalloc(const sphg_kl_field *f, int n, const double *xp, const int *type, int nreal,
                                     const double *xi, double fill, double *value);

/* ---- Correlation-matrix random fields (flspv 3, CPP_17-3-Nonstationary-RF.cpp) ------------- */

/* One spatially variable property generated by the correlation-matrix method
 * (Flag_SpatialVariability and Para_SpatialVariability, Class_spatial_variability.h:10-66, as
 * nonstationary_random_2D/3D use them). sphg_rf_cells fills the cell partition. */
typedef struct {
    int ndim;
    int dist;                 /* 1 normal, 2 lognormal */
    int field;                /* 0 soil fai, 1 soil c, 2 soil cop, 3 soil ds, 4 fluid fai, 5 fluid c */
    int acf;                  /* flag_auto: 1 exponential, 2 Gaussian, otherwise rotated exponential */
    double mean_value, sd_value, coe_depth;
    double rl[3];             /* correlation lengths rl_value */
    double beta, beta_1;      /* rotation angles in degrees (rotated exponential) */
    double cell_ndx, dr;      /* cell edge cell_ndx * dr */
    double box_nb[6];         /* Cell_Con xmin_nb xmax_nb ymin_nb ymax_nb zmin_nb zmax_nb */
    double soil_max, water_max; /* Cell_Con soil_max / water_max on the vertical axis (y in 2D, z in 3D) */
    int nc[3], cell_total;    /* set by sphg_rf_cells */
    double base[3], cell_dx;
} sphg_rf_field;
/* Cell partition (CPP_17-3-Nonstationary-RF.cpp:97-160 2D, :479-565 3D): the non-boundary box
 * widened by 10 % per side, cells of cell_ndx * dr, and the cell of every particle mapped as the
 * reference maps them (2D: water and soil particles; 3D: all particles), -1 otherwise. Host only.
 * Returns 41 when the cell count is not in (0, 15000), 42 when a cell would take a 40th particle
 * (the reference's codes), SPHG_E_ARG for a 3D particle whose cell index falls outside the grid
 * (the reference writes out of bounds there). cell_of may be null. */
int sphg_rf_cells(sphg_rf_field *f, int n, const double *xp, const int *type, int *cell_of);
/* The field for nreal realisations on the device: correlation matrix, Cholesky factor (FP64,
 * blocked, matrix cores for the trailing updates), E = L * zeta (matrix cores), per-particle
 * values value[nreal][n] (the property's phase only; other particles keep fill). zeta holds
 * nreal x cell_total standard normals (the reference draws them from a clock-seeded generator,
 * CPP_17-3-Nonstationary-RF.cpp:32-42). L (may be null) receives the cell_total^2 factor,
 * row-major with zeros above the diagonal; factor_ms (may be null) the Cholesky's device time.
 * Returns SPHG_E_NONFINITE (the outputs are still written) when the factor holds an inf or NaN,
 * where the reference would assign NaN properties without notice. */
int sphg_rf_generate(const sphg_rf_field *f, int n, const double *xp, const int *type, int nreal,
                     const double *zeta, double fill, double *value, double *L, double *factor_ms);

/* Correlated friction angle and cohesion of one phase (flspv 4, nonstationary_correlated_random_2D
 * / _3D, CPP_17-5-Nonstationary-Correlated-RF.cpp:87-737): one correlation matrix (geometry,
 * autocorrelation and lengths of f_fai), its factor, per cell two standard normals zeta
 * [nreal][cell_total][2] mixed by [[1, rho], [0, sqrt(1 - rho^2)]] (rho = coe_frict_cohesion, or
 * the correlation of the underlying normals when f_fai is lognormal), E = L * zeta, and the two
 * properties: f_fai.field 0 (soil) or 4 (fluid), f_c.field one more; both take f_fai's
 * distribution, lognormal values are exp(v), and the fluid pair measures its depth trend on y in
 * 3D too (as the reference does). */
int sphg_rf_generate_pair(const sphg_rf_field *f_fai, const sphg_rf_field *f_c, double coe_frict_cohesion,
                          int n, const double *xp, const int *type, int nreal, const double *zeta, double fill,
                          double *value_fai, double *value_c);

/* Per-particle random-field values from a file (flspv -1, random_field_input,
 * CPP_17-1_K-L_expansion_non.cpp:718-772): <dir>/Parameters.txt with a header line and one line
 * per particle (id, x, y, z, four values, type, matype). values[6][n] (soil fai, c, cop, ds, fluid
 * fai, c) keeps what the caller put there except the values the file gives: fai, c, cop, ds of
 * soil and structure particles, fai and c of water and air particles. Host only; SPHG_E_ARG for a
 * file that cannot be opened (the reference reads from it before the check). The values enter a
 * context with sphg_set_field ("ps_fai" ... "pf_c"). */
int sphg_read_parameters(const char *dir, int n, const int *type, double *values);

/* ---- Dynamic particle sets (flr 1, fleob 1) ---------------------------------------------------- */

/* Para_Rain (Class_Problem.h:83-100) as the RP section gives it (CPP_2-initialf.cpp:180-201):
 * rtype 1 rain, 2 / 3 water-level decrease on the left / right, 4 cement injection (5-7 accepted,
 * nothing happens); nrain reserve particles at the end of the array (type 7); nhor per row (0
 * means 35); initial pressure factor; drop velocity of rain. */
typedef struct {
    int rtype, nrain, nhor;
    double press_factor, drop_velocity;
} sphg_rain;
/* Rain and particle deletion (flr 1; rainini, rainpro, cementpro, waterdel_min / _max, partidel,
 * CPP_15-1-rain.cpp, CPP_15-2-partidel.cpp, called from CPP_1-SPHLoops.cpp:455-511, :757-760).
 * Set before sphg_upload_done, which derives what the reference derives from the initial
 * particles (the water-level limits, rainini's period and velocity, the cement base). Upload
 * returns SPHG_E_ARG for rtype outside 1-7 (the reference's error 6) or a period below one step
 * (the reference divides by zero), SPHG_E_UNSUPPORTED in a batch context. */
int sphg_set_rain(sphg_ctx *ctx, const sphg_rain *rain);
/* Excavation or backfill (fleob 1; exca_or_back, CPP_16-EOB.cpp:13-74): ecp[ecount][3] = first and
 * last particle (1-based, inclusive) and the step at which the stage's particles swap type and
 * etype, one stage per step in order (CPP_1-SPHLoops.cpp:513-520). The reference's reader takes
 * the first two per stage and never assigns the step (CPP_2-initialf.cpp:146), so there the step is
 * indeterminate; here the caller gives it. Set before sphg_upload_done. */
int sphg_set_excavation(sphg_ctx *ctx, int ecount, const int *ecp);
/* The RP section of <dir>/input.dat (CPP_2-initialf.cpp:180-201) after Para_Rain's defaults; nhor 0
 * becomes 35. Host only. */
int sphg_read_rain(const char *dir, sphg_rain *rain);
/* The EC section of <dir>/input.dat (CPP_2-initialf.cpp:125-157): ecount stages of first and last
 * particle into ecp[3 * s], ecp[3 * s + 1]; ecp[3 * s + 2] (the step) is set to 0, at which no stage
 * fires, since the reference never assigns it. ecp may be null (ecount only) or holds 3 * cap ints
 * (SPHG_E_CAPACITY beyond). Returns 8 (the reference's error) when no stage is given. Host only. */
int sphg_read_excavation(const char *dir, int *ecount, int *ecp, int cap);

/* ---- Correlated random variables (flspv 5, CPP_17-5-Nonstationary-Correlated-RF.cpp:738-863) -- */

/* Flag_SpatialVariability members random_variable_generate reads (Class_spatial_variability.h:10-42) */
typedef struct {
    int flag_fai, flag_c, flag_faiw, flag_cw;
    double coe_frict_cohesion;
} sphg_rv_flags;
/* Para_SpatialVariability members it reads (Class_spatial_variability.h:44-66): type 1 normal,
 * 2 lognormal; psv[6] indexed as the reference (0 soil fai, 1 soil c, 4 fluid fai, 5 fluid c) */
typedef struct {
    int type;
    double mean_value, sd_value;
} sphg_rv_param;
/* The homogeneous values of random_variable_generate for the standard normals zeta[2] (the reference
 * draws them from a clock-seeded generator): the distribution is psv[0]'s when flag_fai is set, else
 * psv[4]'s when flag_faiw is set, else normal; the pair is mixed as in flspv 2 and fai takes the
 * mixed normal, c the first. values[k] is set for each flagged property k (0, 1, 4, 5), NaN
 * otherwise. Host only. */
int sphg_random_variable_values(const sphg_rv_flags *fsv, const sphg_rv_param *psv, const double zeta[2],
                                double values[6]);
/* The same values written to every soil (type 2: fai, c) or water (type 1: fai, c) particle in
 * [first, first + count) of the context; nothing is written when the distribution is neither 1 nor 2. */
int sphg_random_variable_generate(sphg_ctx *ctx, const sphg_rv_flags *fsv, const sphg_rv_param *psv,
                                  const double zeta[2], int first, int count);

const char *sphg_error_string(int code);
const char *sphg_build_info(void);

#ifdef __cplusplus
}
#endif
#endif /* SPHG_H_INCLUDED */
