// Device-side data layout and helpers of the SPH step kernels (gfx950).
//
// State lives in HBM as structure-of-arrays: one contiguous FP64 array per scalar component
// (a 3-vector is three arrays), so a wavefront of 64 consecutive particles reads 512 B per
// component, fully coalesced. Neighbour lists are stored column-major, nbr[k * cap + i], for
// the same reason, with the neighbour's type packed beside its id. Kernel values are not stored:
// every pair pass recomputes W and grad W from the gathered positions with the reference's own
// expression, which moves 4 B per pair and pass instead of 36 B.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sphg {

constexpr int kMaxNb = 99;
constexpr int kNSlots = 32;  // 30 soil slots + default row (index 30) + padding

// Per-slot material constants. Loop-invariant transcendental terms (atan/tan/sin/cos of the
// friction angle) are evaluated once on the host with the C library so that the device
// arithmetic that consumes them reproduces the reference's FP64 values exactly.
struct SoilSlot {
    double e, poi, eps0, dens, kapa, zmf, ann, ramda, fs, c, sme0, porosity, vpu, cop, sitas, sitar;
    double damp_coe, beta, ds;
    double bnn, epsr, epss, sd, sw, c1, c2, g0, etao, zet, damp1, damp2;  // UMCC / FOCC parameters
    double exp_c2sw, exp_c1sd;  // exp(c2 * sw), exp(c1 * sd) of moisturesuct (CPP_7-2-constitutive.cpp:333-336)
    double dp_yalf_dil;  // sin(psi') / sqrt(9 + 3 sin^2 psi'), psi' = atan(tan(beta)/fs) (Bui DP dilatancy)
};

struct FluidSlot {
    double vis, vpu, dens, porosity, c, pre_coe, strain_min;
    double tan_fai;   // tan(fai * 0.0174532925199) for the Bingham branch (CPP_7-2-constitutive.cpp:42-43)
    int bhtype, acc_type;
};

// Cell grid of one simulation (celllist_ini1/2, CPP_3-NNPS.cpp:96-171) and its check_domain box.
// A context holds one grid per realisation: their cells are numbered consecutively (realisation r
// owns global cells base + 1 .. base + ncell; global cell 0 parks particles outside their grid),
// so realisations that overlap in space never see each other and each keeps the cell keys,
// neighbour lists and arithmetic it has alone.
struct Grid {
    double xmin[3];          // Cell_Con xmin, ymin, zmin
    double bmin[3], bmax[3]; // Cell_Con box for check_domain
    int cxm, cym, czm, ncell, base, first_moving_bndy, nstencil, pad;
    int nb_off[27];          // stencil offsets in the reference's nncell order
};

// Common neighbour mass of each particle type 0-4.
struct MassTable {
    double t[5];
};

struct Params {
    int n, cap, ndim, flbndy, flknl, fli, if_move, flv, inip, ss_type, flts;
    int ncell;            // cells of all realisations (each grid's own ids run 1..ctotal)
    int water_acc_minor, air_acc_minor;
    double dt, dr, cell_size;
    double gx, gy, gz, gacc;
    double xsph[3];
    double art_vis[3][2];
    double coe_bndy, interc;
    int permtype;
    double pers, perl, perm;
    double tmin, tmax;
    double struct_e, struct_poi;
    int flagum, step_regu;
    double art_stre_coe;
    double alphats;       // alpha of the TS transform (flts >= 1)
    double move_vx[3];
    int hl_uniform;          // every particle has the same smoothing length (hl_j == hl_i bit for bit)
    int nreal;               // realisations (cell grids) in the context
    int fluid_por_uniform;   // every fluid particle's slot has the same porosity
    double ss_myu, ss_zeta;  // Para_SSInter friction coefficient and force factor
    int coupled;             // Para_SSInter acc_type 2 with an interaction type 1-3
    int eig_in_step;         // principal stresses needed within the step (TS transform, artificial stress)
    int mrv_ok;              // mixed form: the momentum kernel writes every particle's mrv (no contact tail)
    double vib_ttime, vib_sttime, vib_edtime;  // Para_Vibra: period, start and end time (flv 1, 2)
    int vib_nsteps;                            // loads per period
    // dynamic particle sets: rain / cement / water-level decrease (flr 1) and excavation or
    // backfill (fleob 1)
    int flr, fleob, rtype, rain_lp, ecount;
    double press_factor, vrain, rx_min, rx_max, cement_base[3];
    double dom_xmin, dom_xmax;                 // Cell_Con xmin / xmax (partidel)
    // time step: 0 constant (Para_Pro::dt, the reference's timestep()), 1 CFL min-reduction:
    // dt = min(dt_cour dr / (dt_cwave + max|v|), 0.25 sqrt(dr / max|a|), dt_max)
    int dt_mode;
    // no water or air particle can exist in the context (none uploaded, no rain, no excavation):
    // every solid's prew / prea is then +0 after the eos pass, so the pore-pressure terms of the
    // momentum sums are +-0 and their gathers are skipped
    int dry;
    double dt_cour, dt_cwave, dt_max;
    // neighbour masses: bit t set when every particle that is, or can become, of type t (0-4)
    // carries the mass mass_t[t], bit for bit; the pair passes then use mass_t[t] (mass_tf[t], the
    // FP64 value of its FP32 rounding, where they read the mixed form's FP32 record) with no gather
    int mass_uniform;
    MassTable mass_t, mass_tf;
    // 2D: every soil / structure particle's out-of-plane components (sig yz, xz; eps zz, yz, xz;
    // adps yz, xz) and every particle's deps zz, yz, xz and spin w02, w12 are +0 (host check at
    // upload). A plane step keeps them +0 (each is +0 plus a zero increment, or the constant the
    // strain pass writes), so the strain and stress passes neither load nor store them; a computed
    // value that is not +0 is stored and raises State::plane_dirty, which ends the skip.
    int plane_zero;
    // vsig / veps / divq / divr / meps are not stored by the step (no consumer inside it: no TS
    // transform, no artificial stress, fixed particle types); k_deq forms them when they are read
    int lazy_deq;
};

// All device arrays of one context (pointers into one HBM allocation).
struct State {
    // kinematics (ping-pong xp / vx by step parity)
    double *xp[2][3], *vx[2][3];
    double *vxp[3], *ax[3], *axp[3], *ux[3], *interf[3], *interfss[3];
    double *rho, *rho_mid, *rhop, *mass, *hl, *pre, *prep, *e;
    double *eps[6], *deps[6], *sig[6], *adps[6];
    double *weps[3];      // w01, w02, w12 of the antisymmetric spin tensor
    double *satu, *dsatu, *suct, *cop, *vsig, *veps, *divq, *divr, *meps, *roue, *prew, *prea;
    double *porosity, *cd, *poro_cop, *color;
    double *strep[3], *beta[9];
    double *vb[4][3];     // clVar_Boundary vx_water/soil/air/struct
    double *vel_bndy[3];  // cl_bndy_pair
    double *ps_c;         // per-particle cohesion (random fields), or null -> slot value
    double *ps_ds;        // per-particle grain size of the drag laws (slot value unless set per particle)
    double *ps_cop;       // per-particle permeability of permtype 0 (slot value unless set per particle)
    double *ps_fai;       // per-particle friction angle (slot value unless set per particle)
    double *pf_fai, *pf_c;  // per-particle fluid friction angle and cohesion (Bingham water; slot values
                            // unless set per particle)
    double *pf_tan;       // from pf_fai on the host: tan(fai * 0.0174532925199)
    double *kgc;          // flknl 3 / 4: kernel-gradient correction L^-1 of particle i, row-major entry r
                          // at kgc[r * cap + i] (soil and structure particles; null otherwise)
    double *tssig[6], *tsqc;  // TS transform (flts >= 1; null otherwise): stress after the transform, qc
    double *ps_dp[3];     // from ps_fai on the host: sin(phi') / sqrt(9 + 3 sin^2 phi'), cos(phi'),
                          // sqrt(9 + 3 sin^2 phi') with phi' = atan(tan(fai) / fs)
    int *type, *matype, *permtype, *etype;
    int8_t *soil_slot, *fluid_slot;
    // cell list
    int *cell, *cell_count, *cell_start, *sorted, *block_sums;
    int *cell_rank;       // particle's arrival rank in its cell at binning (its slot in the cell's run)
    int *nb_cnt, *nbr;    // list entries: id | type << 28, column-major nbr[k * cap + i]
    int *nb_types;        // bit t set when the list holds a neighbour of type t
    double *qsig[6];      // the j factor of the momentum sums: sig / rho / rho (solids),
                          // (porosity * sig) / rho / rho (fluids)
    double *fdens;        // density of the particle's fluid slot (drag partner density)
    // Mixed form only (null in FP64): the neighbour fields the pair passes gather, packed in FP32
    // by the pass that produces them, so one neighbour costs one 16-byte load per record instead
    // of one 8-byte load per field:
    //   mra {m, vxp0, vxp1, vxp2}            update (and the upload)   -> density, strain, accel, XSPH
    //   mrn {rho, pre}                       normalisation             -> strain
    //   mrc {rho, q0, q1, q2} {q3, q4, q5, satu} {prew, prea, porosity, cop}
    //                                         stress (q: the j factor of the stress sums) -> accel
    //   mrv {vx0, vx1, vx2, rho}             momentum (velocity after the update) -> XSPH
    float4 *mra, *mrc, *mrv;
    float2 *mrn;
    double *rij[9];       // clRaij: artificial-stress tensor per particle (persists between steps)
    double *tmp6[6];      // pre-pass copy for the Jacobi strain / stress regularisation
    double *brow[3];      // 2D FP64: boundary-stress rows of the momentum sums, from their own pass
    int *plane_dirty;     // set once a skipped out-of-plane component turned out not to be +0
    int *nw_list;         // ... over the particles the search found next to a wall, and their
    int *nw_count;        // number (cleared before every search)
    // soil-structure interaction (ss_type 1-3): per soil / water particle the structure particles
    // its contact force is passed to (-1: none) and what each receives (f_n[3], f_t[3])
    int *ssi_id[3];
    double *ssi_c[3][6];
    double *cpl[3];       // coupled_Water_Soil_Structure row sums (Para_SSInter acc_type 2)
    // boundary treatment of Tran et al. (flbndy 3/4/13/14): per boundary particle and phase
    // (water, soil, air, structure) the interpolated stress
    double *sigb[4][6];
    // vibration loads and free-field boundaries (flv 1, 2; null otherwise)
    double *vib_loads;    // Para_Vibra loads[nsteps][3]
    double *tnow;         // Para_Pro::t, advanced by dt at the start of every step
    double *vstiff;       // largest diagonal entry (with zero) of the particle's last tangent (depp0)
    double *ffb[3], *ffs[3];  // free-field sums against matype-5 walls / between soil and its columns
    double *vdamp[4];     // soil_damping_vibration sums: damp[3], tempdown
    // dynamic particle sets: dyn = {nrain, cement_flag, nstage}; ecp[ecount][3] = start, end, step
    int *dyn, *ecp;
    int *err, *step;      // sticky error word; current step number l
    int *shear3;          // 2D: set (and kept) once some particle's sig_yz or sig_xz is nonzero
    double *dtp;          // time step of the current step (Para_Pro::dt; the CFL mode rewrites it)
    unsigned long long *dt_red;  // CFL mode: bit patterns of max |v|^2 and max |a|^2 over particles
    const Grid *grids;    // one per realisation
    const int *rid;       // realisation of each particle
    const SoilSlot *soil;
    const FluidSlot *fluid;
};

#define SPHG_ERR_CELL 1
#define SPHG_ERR_NB 2

// weight(): cubic B-spline, CPP_3-NNPS.cpp:17-48 (same operation order; 0.66666667 as written there).
__device__ __forceinline__ void weight_b(double dst, double ra, const double dx[3], double w[4], int ndim) {
    double a1 = 0.0, s = dst / ra;
    if (ndim == 2) {
        a1 = 0.682092613250980 / ra / ra;
        w[2] = 0.0;
    } else {
        a1 = 0.477464829275686 / ra / ra / ra;
    }
    if (s == 0.0) {
        for (int k = 0; k < ndim; k++) w[k] = 0.0;
        w[3] = a1 * (0.66666667 - s * s + 0.5 * s * s * s);
    }
    if (s > 0.0 && s <= 1.0) {
        double c = (-2 * s + 1.5 * s * s), rd = ra * dst;
        for (int k = 0; k < ndim; k++) w[k] = a1 * dx[k] * c / rd;
        w[3] = a1 * (0.66666667 - s * s + 0.5 * s * s * s);
    } else if (s > 1.0 && s <= 2.0) {
        double rd = ra * dst;
        for (int k = 0; k < ndim; k++) w[k] = a1 * (-0.5) * dx[k] * (2.0 - s) * (2.0 - s) / rd;
        w[3] = a1 * (2.0 - s) * (2.0 - s) * (2.0 - s) / 6.0;
    }
}

// weight2(): Yang & Liu kernel, CPP_3-NNPS.cpp:51-82 (flknl 2).
__device__ __forceinline__ void weight_y(double dst, double ra, const double dx[3], double w[4], int ndim) {
    double a1 = 0.0, s = dst / ra;
    if (ndim == 2) {
        a1 = 0.106103295394597 / (ra * ra);
        w[2] = 0.0;
    } else {
        a1 = 0.077010456334788 / (ra * ra * ra);
    }
    if (s == 0.0) {
        for (int k = 0; k < ndim; k++) w[k] = 0.0;
        w[3] = a1 * (6.0 - 6.0 * s + 1.0 * s * s * s);
    }
    if (s > 0.0 && s <= 1.0) {
        double rd = ra * dst;
        for (int k = 0; k < ndim; k++) w[k] = a1 * dx[k] * (-6.0 + 3.0 * s * s) / rd;
        w[3] = a1 * (6.0 - 6.0 * s + 1.0 * s * s * s);
    } else if (s > 1.0 && s <= 2.0) {
        double rd = ra * dst;
        for (int k = 0; k < ndim; k++) w[k] = a1 * (-3.0) * dx[k] * (2.0 - s) * (2.0 - s) / rd;
        w[3] = a1 * (2.0 - s) * (2.0 - s) * (2.0 - s);
    }
}

}  // namespace sphg
