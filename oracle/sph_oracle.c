/*
 * Test infrastructure only — the oracle (see sph_oracle.h). A serial/OpenMP FP64 restatement
 * of the reference step loop; every function names the reference lines it restates. Quirks of
 * the reference are reproduced on purpose and marked "quirk".
 *
 * Parity status of each function is pinned in tests/test_oracle_ref.py against full-step dumps
 * of the reference link harness (oracle/_ref) on seeded decks: bit-equal at one thread in
 * OR_DENSITY_GS mode for every function exercised by the 2D column and 3D soil-water decks.
 */
#include "sph_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NBMAX 100

typedef struct {
    double x[3], xp[3], vx[3], vxp[3], ax[3], axp[3], ux[3], strep[3];
    double divx[3][3], interf[3], interfss[3];
    double rho, rhop, mass, hl, pre, prep, e;
    double eps[6], deps[6], sig[6], dsig[6], weps[3][3], adps[6];
    double satu, dsatu, suct, cop, vsig, veps, divq, divr, meps, roue, rous, prew, prea;
    double porosity, cd, poro_cop;
    double beta[3][3], beta_ts[3][3];
    double Rij[3][3]; /* clRaij (Class_particle.h:181-187) */
    double tssig[6], tsqc; /* TS transform (Class_particle.h:54-55) */
    int type, matype, permtype, etype;
} OPar;

typedef struct {
    int cell_id, ninflu;
    int influ[NBMAX];
    double wij[NBMAX][4];
} OCell;

typedef struct {
    double vx_water[3], vx_soil[3], vx_air[3], vx_struct[3];
    double sig_ph[4][6], rho_ph[4]; /* sig_water/soil/air/struct, rho_water/... (Class_particle.h) */
} OVariB;

struct or_ctx {
    sphg_problem pro;
    sphg_materials mat;
    int n;
    OPar *p;
    OCell *pc;
    sphg_soil *ps;   /* pParti_ConsPara */
    sphg_fluid *pf;  /* pVFluid */
    OVariB *vb;      /* pParti_VariBndy */
    double (*vel_bndy)[3];
    double (*vx_temp)[3];
    double *rho_tmp; /* Jacobi normalisation input */
    /* cell list */
    int cxm, cym, czm, ctotal;
    int *cstart, *cend; /* cell_info[ctotal+4] */
    int (*nncell)[27];  /* cell_link[ctotal+4] */
    int size_m, klog;
    int *s_cell, *s_id, *t_cell, *t_id; /* parti_cellid sorted/temp [size_m] */
    int l;
    double t;
    int density_mode, nthreads;
    /* StiffMat::depp0 (Class_particle.h:105-112): the diagonal of each particle's last tangent */
    double (*depp0)[3];
    /* Para_Vibra (Class_Problem.h:133-149) and the loads of vibra.dat */
    double vib_ttime, vib_sttime, vib_edtime;
    int vib_nsteps;
    double (*vib_loads)[3];
    /* Para_Rain (RP section) with rainini's period lp and velocity, setting_par's x_min / x_max of
     * the water particles of etype 7 and center_cement's base (flr 1) */
    int rain_set, rtype, nrain, nhor, cement_flag, lp;
    double press_factor, drop_velocity, vrain, rx_min, rx_max, cement_base[3];
    /* Para_EC (fleob 1): stages [start, end] and the step each fires at; nstage fired so far */
    int ecount, nstage;
    int (*ecp)[3];
};

/* ------------------------------------------------------------------ kernel function */
/* weight(): cubic B-spline value and gradient (CPP_3-NNPS.cpp:17-48); the truncated
 * 0.66666667 is the reference's constant. */
static void or_weight(double dst, double ra, const double dx[], double *w, int ndim) {
    double a1 = 0.0, s = dst / ra;
    int k;
    if (ndim == 2) {
        a1 = 0.682092613250980 / ra / ra;
        w[ndim] = 0.0;
    }
    if (ndim == 3) a1 = 0.477464829275686 / ra / ra / ra;
    if (s == 0.0) {
        for (k = 0; k < ndim; k++) w[k] = 0.0;
        w[3] = a1 * (0.66666667 - s * s + 0.5 * s * s * s);
    }
    if (s > 0.0 && s <= 1.0) {
        for (k = 0; k < ndim; k++) w[k] = a1 * dx[k] * (-2 * s + 1.5 * s * s) / (ra * dst);
        w[3] = a1 * (0.66666667 - s * s + 0.5 * s * s * s);
    } else if (s > 1.0 && s <= 2.0) {
        for (k = 0; k < ndim; k++) w[k] = a1 * (-0.5) * dx[k] * (2.0 - s) * (2.0 - s) / (ra * dst);
        w[3] = a1 * (2.0 - s) * (2.0 - s) * (2.0 - s) / 6.0;
    }
}

/* weight2(): Yang & Liu kernel (CPP_3-NNPS.cpp:51-82), flknl 2/4. */
static void or_weight2(double dst, double ra, const double dx[], double *w, int ndim) {
    double a1 = 0.0, s = dst / ra;
    int k;
    if (ndim == 2) {
        a1 = 0.106103295394597 / (ra * ra);
        w[ndim] = 0.0;
    }
    if (ndim == 3) a1 = 0.077010456334788 / (ra * ra * ra);
    if (s == 0.0) {
        for (k = 0; k < ndim; k++) w[k] = 0.0;
        w[3] = a1 * (6.0 - 6.0 * s + 1.0 * s * s * s);
    }
    if (s > 0.0 && s <= 1.0) {
        for (k = 0; k < ndim; k++) w[k] = a1 * dx[k] * (-6.0 + 3.0 * s * s) / (ra * dst);
        w[3] = a1 * (6.0 - 6.0 * s + 1.0 * s * s * s);
    } else if (s > 1.0 && s <= 2.0) {
        for (k = 0; k < ndim; k++) w[k] = a1 * (-3.0) * dx[k] * (2.0 - s) * (2.0 - s) / (ra * dst);
        w[3] = a1 * (2.0 - s) * (2.0 - s) * (2.0 - s);
    }
}

/* ------------------------------------------------------------------ context */
static void par_init(OPar *q) { memset(q, 0, sizeof *q); }

static void soil_default(sphg_soil *s) {
    memset(s, 0, sizeof *s);
    s->fs = 1.0;
    s->ds = 1.0;
}
static void fluid_default(sphg_fluid *f) {
    memset(f, 0, sizeof *f);
    f->strain_min = 1.0;
}

or_ctx *or_create(const sphg_problem *pro, const sphg_materials *mat, int n) {
    or_ctx *c = (or_ctx *)calloc(1, sizeof *c);
    int i;
    c->pro = *pro;
    c->mat = *mat;
    c->n = n;
    c->p = (OPar *)calloc(n, sizeof(OPar));
    c->pc = (OCell *)calloc(n, sizeof(OCell));
    c->ps = (sphg_soil *)calloc(n, sizeof(sphg_soil));
    c->pf = (sphg_fluid *)calloc(n, sizeof(sphg_fluid));
    c->vb = (OVariB *)calloc(n, sizeof(OVariB));
    c->vel_bndy = (double(*)[3])calloc(n, sizeof(double[3]));
    c->vx_temp = (double(*)[3])calloc(n, sizeof(double[3]));
    c->rho_tmp = (double *)calloc(n, sizeof(double));
    c->depp0 = (double(*)[3])calloc(n, sizeof(double[3]));
    for (i = 0; i < n; i++) {
        par_init(&c->p[i]);
        soil_default(&c->ps[i]);
        fluid_default(&c->pf[i]);
        c->pc[i].cell_id = -1;
    }
    /* allocation sizes of SPH_Calculation (CPP_1-SPHLoops.cpp:353-356) */
    c->klog = (int)log2((double)n) + 1;
    c->size_m = 1 << c->klog;
    c->s_cell = (int *)malloc(sizeof(int) * c->size_m);
    c->s_id = (int *)malloc(sizeof(int) * c->size_m);
    c->t_cell = (int *)malloc(sizeof(int) * c->size_m);
    c->t_id = (int *)malloc(sizeof(int) * c->size_m);
    for (i = 0; i < c->size_m; i++) { /* parti_cellid() (Class_NNPS.cpp:48-52) */
        c->s_cell[i] = c->t_cell[i] = 10000000;
        c->s_id[i] = c->t_id[i] = 10000000;
    }
    c->l = 1;
    c->t = 0.0;
    c->density_mode = OR_DENSITY_GS;
    c->nthreads = 1;
    return c;
}

void or_destroy(or_ctx *c) {
    if (!c) return;
    free(c->p); free(c->pc); free(c->ps); free(c->pf); free(c->vb); free(c->vel_bndy);
    free(c->vx_temp); free(c->rho_tmp); free(c->cstart); free(c->cend); free(c->nncell);
    free(c->s_cell); free(c->s_id); free(c->t_cell); free(c->t_id);
    free(c->depp0); free(c->vib_loads); free(c->ecp);
    free(c);
}

/* Para_Rain from the RP section (CPP_2-initialf.cpp:180-201, nhor 0 taken as 35), then what the
 * reference derives before its loop from the initial state: x_min / x_max over the water particles
 * of etype 7 (setting_par, :1307-1311), the period lp and velocity of rainini
 * (CPP_15-1-rain.cpp:10-45; nrain / nhor is an integer division there) and the cement base, the
 * centre of the type-7 particles' box (center_cement, :133-163). Call after the particles are set.
 * Returns 6 for a type outside 1-7 (the reference's error), SPHG_E_ARG for lp < 1 where the
 * reference would divide by zero. */
int or_set_rain(or_ctx *c, int rtype, int nrain, int nhor, double press_factor, double drop_velocity) {
    int i, dim = c->pro.ndim;
    double temp, mx[3] = {-1000.0, -1000.0, -1000.0}, mn[3] = {1000.0, 1000.0, 1000.0};
    if (rtype < 1 || rtype > 7) return 6;
    c->rtype = rtype;
    c->nrain = nrain;
    c->nhor = nhor == 0 ? 35 : nhor;
    c->press_factor = press_factor;
    c->drop_velocity = drop_velocity;
    c->cement_flag = 1;
    c->rx_max = -10000.0;
    c->rx_min = 10000.0;
    c->vrain = 0.0;
    c->lp = 0;
    for (i = 0; i < c->n; i++)
        if (c->p[i].type == 1 && c->p[i].etype == 7) {
            c->rx_max = fmax(c->rx_max, c->p[i].xp[dim - 2]);
            c->rx_min = fmin(c->rx_min, c->p[i].xp[dim - 2]);
        }
    if (rtype == 1 || rtype == 4) {
        temp = c->nrain / c->nhor + 1.0;
        c->lp = (int)((c->pro.loop - c->pro.inip) / temp);
        c->vrain = rtype == 1 ? drop_velocity : 2.0;
    } else if (rtype == 2 || rtype == 3) {
        temp = fabs(c->rx_max - c->rx_min) / c->pro.dr + 1.0;
        c->lp = (int)((c->pro.loop - c->pro.inip) / temp);
    }
    for (i = 0; i < c->n; i++)
        if (c->p[i].type == 7) {
            int k;
            for (k = 0; k < 3; k++) {
                mx[k] = fmax(mx[k], c->p[i].xp[k]);
                mn[k] = fmin(mn[k], c->p[i].xp[k]);
            }
        }
    for (i = 0; i < 3; i++) c->cement_base[i] = (mx[i] + mn[i]) * 0.5;
    c->rain_set = 1;
    if (rtype <= 4 && c->lp < 1) return SPHG_E_ARG;
    return 0;
}

/* Para_EC: ecount stages, ecp[i] = {start, end, step} (1-based particle range; the reference's
 * reader fills start and end only, CPP_2-initialf.cpp:146, and never the step). */
int or_set_excavation(or_ctx *c, int ecount, const int *ecp) {
    if (ecount < 1 || !ecp) return 8;  /* the reference's error for no stage */
    free(c->ecp);
    c->ecp = (int(*)[3])malloc(sizeof(int[3]) * ecount);
    memcpy(c->ecp, ecp, sizeof(int[3]) * ecount);
    c->ecount = ecount;
    c->nstage = 0;
    return 0;
}

/* initial_vib (CPP_2-initialf.cpp:293-322) and the VL section (:250-265): one period of ttime
 * in nsteps loads (three components each), applied between sttime and edtime. */
int or_set_vibration(or_ctx *c, double ttime, int nsteps, double sttime, double edtime, const double *loads) {
    if (nsteps < 1 || !loads) return SPHG_E_ARG;
    free(c->vib_loads);
    c->vib_loads = (double(*)[3])malloc(sizeof(double[3]) * nsteps);
    memcpy(c->vib_loads, loads, sizeof(double[3]) * nsteps);
    c->vib_ttime = ttime;
    c->vib_nsteps = nsteps;
    c->vib_sttime = sttime;
    c->vib_edtime = edtime;
    return 0;
}

void or_set_mode(or_ctx *c, int density_mode, int nthreads) {
    c->density_mode = density_mode;
    c->nthreads = nthreads > 0 ? nthreads : 1;
}
void or_set_step(or_ctx *c, int l) { c->l = l; }
void or_set_time(or_ctx *c, double t) { c->t = t; }

/* field access by Particle member name ------------------------------------------------ */
typedef struct { const char *name; size_t off; int ncomp; } FieldDesc;
#define FD(nm, nc) {#nm, offsetof(OPar, nm), nc}
static const FieldDesc PFIELDS[] = {
    FD(x, 3), FD(xp, 3), FD(vx, 3), FD(vxp, 3), FD(ax, 3), FD(axp, 3), FD(ux, 3), FD(strep, 3),
    FD(divx, 9), FD(interf, 3), FD(interfss, 3), FD(rho, 1), FD(rhop, 1), FD(mass, 1), FD(hl, 1),
    FD(pre, 1), FD(prep, 1), FD(e, 1), FD(eps, 6), FD(deps, 6), FD(sig, 6), FD(dsig, 6),
    FD(weps, 9), FD(adps, 6), FD(satu, 1), FD(dsatu, 1), FD(suct, 1), FD(cop, 1), FD(vsig, 1),
    FD(veps, 1), FD(divq, 1), FD(divr, 1), FD(meps, 1), FD(roue, 1), FD(rous, 1), FD(prew, 1),
    FD(prea, 1), FD(porosity, 1), FD(cd, 1), FD(poro_cop, 1), FD(beta, 9), FD(beta_ts, 9), FD(Rij, 9),
    FD(tssig, 6), FD(tsqc, 1),
};
#undef FD
#define SD(nm) {"ps_" #nm, offsetof(sphg_soil, nm), 1}
static const FieldDesc SFIELDS[] = {
    SD(e), SD(poi), SD(eps0), SD(dens), SD(damp1), SD(damp2), SD(cop), SD(ramda), SD(kapa), SD(zmf),
    SD(ann), SD(bnn), SD(beta), SD(sitas), SD(sitar), SD(epsr), SD(epss), SD(sd), SD(sw), SD(kse),
    SD(c1), SD(c2), SD(c3), SD(fai), SD(c), SD(ocr), SD(sme0), SD(mzm), SD(azm), SD(r0), SD(bl),
    SD(bz), SD(zet), SD(vpu), SD(porosity), SD(fs), SD(etao), SD(g0), SD(ds), SD(damp_coe),
};
#undef SD
#define FDF(nm) {"pf_" #nm, offsetof(sphg_fluid, nm), 1}
static const FieldDesc FFIELDS[] = {
    FDF(vis), FDF(vpu), FDF(dens), FDF(paa0), FDF(porosity), FDF(fai), FDF(c), FDF(pre_coe), FDF(strain_min),
};
#undef FDF

static int field_rw(or_ctx *c, const char *name, double *v, int n, int ncomp, int write) {
    size_t i, k;
    if (n != c->n) return SPHG_E_ARG;
    for (i = 0; i < sizeof PFIELDS / sizeof PFIELDS[0]; i++) {
        if (strcmp(PFIELDS[i].name, name)) continue;
        if (ncomp != PFIELDS[i].ncomp) return SPHG_E_ARG;
        for (k = 0; k < (size_t)n; k++) {
            double *f = (double *)((char *)&c->p[k] + PFIELDS[i].off);
            if (write) memcpy(f, v + k * ncomp, sizeof(double) * ncomp);
            else memcpy(v + k * ncomp, f, sizeof(double) * ncomp);
        }
        return 0;
    }
    for (i = 0; i < sizeof SFIELDS / sizeof SFIELDS[0]; i++) {
        if (strcmp(SFIELDS[i].name, name)) continue;
        if (ncomp != 1) return SPHG_E_ARG;
        for (k = 0; k < (size_t)n; k++) {
            double *f = (double *)((char *)&c->ps[k] + SFIELDS[i].off);
            if (write) *f = v[k];
            else v[k] = *f;
        }
        return 0;
    }
    for (i = 0; i < sizeof FFIELDS / sizeof FFIELDS[0]; i++) {
        if (strcmp(FFIELDS[i].name, name)) continue;
        if (ncomp != 1) return SPHG_E_ARG;
        for (k = 0; k < (size_t)n; k++) {
            double *f = (double *)((char *)&c->pf[k] + FFIELDS[i].off);
            if (write) *f = v[k];
            else v[k] = *f;
        }
        return 0;
    }
    {
        const char *inm[3] = {"pf_bhtype", "pf_press_type", "pf_acc_type"};
        for (i = 0; i < 3; i++) {
            if (strcmp(inm[i], name)) continue;
            for (k = 0; k < (size_t)n; k++) {
                int *f = i == 0 ? &c->pf[k].bhtype : i == 1 ? &c->pf[k].press_type : &c->pf[k].acc_type;
                if (write) *f = (int)v[k];
                else v[k] = *f;
            }
            return 0;
        }
    }
    {
        struct { const char *nm; size_t off; } vbf[] = {
            {"vx_water", offsetof(OVariB, vx_water)}, {"vx_soil", offsetof(OVariB, vx_soil)},
            {"vx_air", offsetof(OVariB, vx_air)}, {"vx_struct", offsetof(OVariB, vx_struct)}};
        for (i = 0; i < 4; i++) {
            if (strcmp(vbf[i].nm, name)) continue;
            if (ncomp != 3) return SPHG_E_ARG;
            for (k = 0; k < (size_t)n; k++) {
                double *f = (double *)((char *)&c->vb[k] + vbf[i].off);
                if (write) memcpy(f, v + 3 * k, 24);
                else memcpy(v + 3 * k, f, 24);
            }
            return 0;
        }
    }
    if (!strcmp(name, "vel_bndy")) {
        if (ncomp != 3) return SPHG_E_ARG;
        if (write) memcpy(c->vel_bndy, v, 24 * (size_t)n);
        else memcpy(v, c->vel_bndy, 24 * (size_t)n);
        return 0;
    }
    return SPHG_E_ARG;
}

int or_set_field(or_ctx *c, const char *name, const double *v, int n, int ncomp) {
    return field_rw(c, name, (double *)v, n, ncomp, 1);
}
int or_get_field(or_ctx *c, const char *name, double *v, int n, int ncomp) {
    return field_rw(c, name, v, n, ncomp, 0);
}

static int *ifield(or_ctx *c, const char *name, int k) {
    if (!strcmp(name, "type")) return &c->p[k].type;
    if (!strcmp(name, "matype")) return &c->p[k].matype;
    if (!strcmp(name, "permtype")) return &c->p[k].permtype;
    if (!strcmp(name, "etype")) return &c->p[k].etype;
    return NULL;
}
int or_set_field_i(or_ctx *c, const char *name, const int *v, int n) {
    int k;
    if (n != c->n || !ifield(c, name, 0)) return SPHG_E_ARG;
    for (k = 0; k < n; k++) *ifield(c, name, k) = v[k];
    return 0;
}
int or_get_field_i(or_ctx *c, const char *name, int *v, int n) {
    int k;
    if (n != c->n || !ifield(c, name, 0)) return SPHG_E_ARG;
    for (k = 0; k < n; k++) v[k] = *ifield(c, name, k);
    return 0;
}

/* celllist_ini1 / celllist_ini2 (CPP_3-NNPS.cpp:96-171). */
static void or_celllist_ini(or_ctx *c) {
    const sphg_problem *P = &c->pro;
    double dr = 2.0 * P->dr;
    int i, m, p, kx, ky, kz;
    if (P->ndim == 2) {
        c->cxm = (int)((P->xmax - P->xmin) / dr + 3.0);
        c->cym = (int)((P->ymax - P->ymin) / dr + 3.0);
        c->czm = (int)(1.0);
    } else {
        c->cxm = (int)((P->xmax - P->xmin) / dr + 3.0);
        c->cym = (int)((P->ymax - P->ymin) / dr + 3.0);
        c->czm = (int)((P->zmax - P->zmin) / dr + 3.0);
    }
    c->ctotal = c->cxm * c->cym * c->czm;
    m = c->ctotal;
    free(c->cstart); free(c->cend); free(c->nncell);
    c->cstart = (int *)malloc(sizeof(int) * (m + 4));
    c->cend = (int *)malloc(sizeof(int) * (m + 4));
    c->nncell = (int(*)[27])malloc(sizeof(int[27]) * (m + 4));
    for (i = 0; i < m + 4; i++) { /* cell_info() and cell_link() constructors */
        c->cstart[i] = 2000000;
        c->cend[i] = -2000000;
        for (p = 0; p < 27; p++) c->nncell[i][p] = -1;
    }
    for (i = 0; i <= m; i++) {
        p = 0;
        if (P->ndim == 2) {
            for (ky = 1; ky <= 3; ky++)
                for (kx = 1; kx <= 3; kx++) {
                    int v = (kx - 2) * c->cym + ky - 2 + i;
                    c->nncell[i][p++] = (v < 0 || v > m) ? 0 : v;
                }
        } else {
            for (kz = 1; kz <= 3; kz++)
                for (kx = 1; kx <= 3; kx++)
                    for (ky = 1; ky <= 3; ky++) {
                        int v = (ky - 2) * c->cxm * c->czm + (kx - 2) * c->czm + kz - 2 + i;
                        c->nncell[i][p++] = (v < 0 || v > m) ? 0 : v;
                    }
        }
    }
}

/* setting_vis_cons (CPP_2-initialf.cpp:1021-1058): per-particle copies of the material rows.
 * Reads type/matype/etype; per-particle overrides (random fields) are set afterwards. */
int or_setting_vis_cons(or_ctx *c) {
    int i;
    for (i = 0; i < c->n; i++) {
        const OPar *p = &c->p[i];
        int m = p->matype;
        soil_default(&c->ps[i]);
        fluid_default(&c->pf[i]);
        if (p->type == 1 && m >= 0 && m < 3) c->pf[i] = c->mat.water[m];
        if (p->type == 2) {
            if (m < 30 && m >= 0) c->ps[i] = c->mat.soil[m];
            else if (m >= 50) c->ps[i] = c->mat.soil[m % 10];
        }
        if (p->type == 3) c->pf[i] = c->mat.air;
        if (p->type == 7 && p->etype == 1 && m >= 0 && m < 3) c->pf[i] = c->mat.water[m];
        if (p->type == 7 && p->etype == 2 && m >= 0 && m < 30) c->ps[i] = c->mat.soil[m];
    }
    return 0;
}

int or_upload_done(or_ctx *c) {
    or_celllist_ini(c);
    return 0;
}

/* ------------------------------------------------------------------ NNPS */
/* Merge_CellID (CPP_3-NNPS.cpp:561-595): ties take the right half first. */
static void or_merge(or_ctx *c, int start, int mid, int end) {
    int i, j, k;
    for (i = start; i <= end; i++) {
        c->t_cell[i] = c->s_cell[i];
        c->t_id[i] = c->s_id[i];
    }
    i = start;
    j = mid + 1;
    k = start;
    while (i < mid + 1 && j <= end) {
        if (c->t_cell[i] < c->t_cell[j]) {
            c->s_cell[k] = c->t_cell[i]; c->s_id[k++] = c->t_id[i++];
        } else {
            c->s_cell[k] = c->t_cell[j]; c->s_id[k++] = c->t_id[j++];
        }
    }
    while (i < mid + 1) { c->s_cell[k] = c->t_cell[i]; c->s_id[k++] = c->t_id[i++]; }
    while (j <= end) { c->s_cell[k] = c->t_cell[j]; c->s_id[k++] = c->t_id[j++]; }
}

/* parttocell + Binary_Sort (CPP_3-NNPS.cpp:175-263, :597-630). */
static int or_parttocell(or_ctx *c) {
    const sphg_problem *P = &c->pro;
    int n = c->n, m = c->ctotal, i, j, t = 0;
    double xmin[3] = {P->xmin, P->ymin, P->zmin}, dr = 2.0 * P->dr;
    for (i = 0; i < m; i++) {
        c->cstart[i] = 10000000;
        c->cend[i] = -10000000;
    }
    for (i = 0; i < n; i++) {
        int nt[3] = {2, 2, 2};
        for (j = 0; j < P->ndim; j++) {
            double v = fabs(c->p[i].xp[j] - xmin[j]) / dr + 2.0; /* quirk: fabs mirrors (A6) */
            nt[j] = (int)v;
            if (nt[j] < 2) nt[j] = 2;
        }
        if (P->ndim == 2) c->s_cell[i] = c->cym * (nt[0] - 1) + nt[1];
        else c->s_cell[i] = c->cxm * c->czm * (nt[1] - 1) + c->czm * (nt[0] - 1) + nt[2];
        c->s_id[i] = i;
    }
    for (i = 0; i < n; i++) {
        if (c->s_cell[i] > 0 && c->s_cell[i] <= m) c->pc[i].cell_id = c->s_cell[i];
        else t = 3;
    }
    if (t) return t;
    for (int ik = 1; ik <= c->klog; ik++) {
        int interval = 1 << ik, groups = c->size_m / interval, g;
        for (g = 0; g < groups; g++) {
            int start = g * interval, end = (g + 1) * interval - 1;
            or_merge(c, start, (start + end) >> 1, end);
        }
    }
    for (i = 0; i < n; i++) {
        int cid = c->s_cell[i];
        if (i < c->cstart[cid]) c->cstart[cid] = i;
        if (i > c->cend[cid]) c->cend[cid] = i;
    }
    return 0;
}

/* partisearching (CPP_3-NNPS.cpp:265-374). */
static int or_partisearching(or_ctx *c) {
    const sphg_problem *P = &c->pro;
    int n = c->n, dim = P->ndim, nst = dim == 2 ? 9 : 27, t = 0, i;
    int use_w1 = (P->flknl == 1 || P->flknl == 3);
    for (i = 0; i < n; i++) c->pc[i].ninflu = 0; /* Par_Cell::initial (Class_NNPS.cpp:94-97) */
#pragma omp parallel for schedule(static) num_threads(c->nthreads) reduction(max : t)
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        int nc, nj, np, k;
        if (pi->type == 7) continue;
        nc = c->pc[i].cell_id;
        for (nj = 0; nj < nst; nj++) {
            int nl = c->nncell[nc][nj];
            if (nl <= 0) continue;
            for (np = c->cstart[nl]; np <= c->cend[nl]; np++) {
                int j = c->s_id[np];
                double dst = 0.0, dx[3], ra;
                for (k = 0; k < dim; k++) {
                    dst = dst + (pi->xp[k] - c->p[j].xp[k]) * (pi->xp[k] - c->p[j].xp[k]);
                    dx[k] = pi->xp[k] - c->p[j].xp[k];
                }
                dst = sqrt(dst);
                ra = (pi->hl + c->p[j].hl) / 2;
                if (dst <= 2 * ra && i != j) {
                    int p = ++c->pc[i].ninflu;
                    if (p <= 99) {
                        c->pc[i].influ[p] = j;
                        if (use_w1) or_weight(dst, ra, dx, c->pc[i].wij[p], dim);
                        else or_weight2(dst, ra, dx, c->pc[i].wij[p], dim);
                    } else
                        t = 5;
                }
            }
        }
    }
    return t;
}

/* neighbours actually stored (ninflu may exceed the 99 slots before error 5 is raised) */
static inline int nbcount(const OCell *pc) { return pc->ninflu < 99 ? pc->ninflu : 99; }

/* kernel_gradient_correction (CPP_3-NNPS.cpp:376-496, flknl 3 / 4): for soil and structure
 * particles, L = sum over soil, boundary and structure neighbours of V_j (x_j - x_i) (x) grad W,
 * each listed gradient replaced by L^-1 grad W; L^-1 by cofactors, the identity when
 * |det L| <= 1e-5 (inverse_mat_2D/3D, :498-559); in 2D row and column 2 of L^-1 stay zero. */
static void or_kgc(or_ctx *c) {
    int n = c->n, dim = c->pro.ndim, i;
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        double A[3][3] = {{0}}, B[3][3] = {{0}}, det, x[3], w[3];
        int nc = nbcount(&c->pc[i]), nj, k, l;
        if (pi->type != 2 && pi->type != 4) continue;
        for (nj = 1; nj <= nc; nj++) {
            const OPar *pj = &c->p[c->pc[i].influ[nj]];
            const double *g = c->pc[i].wij[nj];
            if (!(pj->type == 2 || pj->type == 0 || pj->type == 4)) continue;
            for (k = 0; k < dim; k++) x[k] = pj->xp[k] - pi->xp[k];
            for (k = 0; k < dim; k++)
                for (l = 0; l < dim; l++) A[k][l] = A[k][l] + pj->mass / pj->rho * x[k] * g[l];
        }
        if (dim == 2) {
            det = A[0][0] * A[1][1] - A[0][1] * A[1][0];
            if (fabs(det) > 1.0e-5) {
                B[0][0] = A[1][1] / det;
                B[0][1] = -A[0][1] / det;
                B[1][0] = -A[1][0] / det;
                B[1][1] = A[0][0] / det;
            } else {
                B[0][0] = B[1][1] = 1.0;
            }
        } else {
            det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) - A[1][0] * (A[0][1] * A[2][2] - A[0][2] * A[2][1]) +
                  A[2][0] * (A[0][1] * A[1][2] - A[0][2] * A[1][1]);
            if (fabs(det) > 1.0e-5) {
                B[0][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) / det;
                B[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / det;
                B[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / det;
                B[1][0] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) / det;
                B[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / det;
                B[1][2] = (A[1][0] * A[0][2] - A[0][0] * A[1][2]) / det;
                B[2][0] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) / det;
                B[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) / det;
                B[2][2] = (A[0][0] * A[1][1] - A[1][0] * A[0][1]) / det;
            } else {
                B[0][0] = B[1][1] = B[2][2] = 1.0;
            }
        }
        for (nj = 1; nj <= nc; nj++) {
            double *g = c->pc[i].wij[nj];
            for (k = 0; k < 3; k++) w[k] = B[k][0] * g[0] + B[k][1] * g[1] + B[k][2] * g[2];
            for (k = 0; k < 3; k++) g[k] = w[k];
        }
    }
}


/* virtual velocity of boundary neighbour j as seen by particle i (CPP_6-density.cpp:86-120;
 * the same selection is repeated in CPP_7-1-strain.cpp:92-125 and CPP_11-xsph.cpp:187-219). */
static void bndy_vel(const or_ctx *c, int i, int j, int use_vx, double vxj[3]) {
    const OPar *pi = &c->p[i];
    int bt = c->pro.flbndy;
    vxj[0] = vxj[1] = vxj[2] = 0.0;
    if (bt == 1) {
        const double *v = use_vx ? pi->vx : pi->vxp;
        vxj[0] = v[0]; vxj[1] = v[1]; vxj[2] = v[2];
    } else if (bt == 2 || bt == 3 || bt == 4 || bt == 13 || bt == 14) {
        const double *v = NULL;
        if (pi->type == 1) v = c->vb[j].vx_water;
        else if (pi->type == 2) v = c->vb[j].vx_soil;
        else if (pi->type == 3) v = c->vb[j].vx_air;
        else if (pi->type == 4) v = c->vb[j].vx_struct;
        if (v) { vxj[0] = v[0]; vxj[1] = v[1]; vxj[2] = v[2]; }
    }
}

/* ------------------------------------------------------------------ density (CPP_6-density.cpp:11-189) */
static void or_density(or_ctx *c) {
    int n = c->n, ndim = c->pro.ndim, l = c->l, i;
    double dt = c->pro.dt;
    /* loop 1: water porosity from soil neighbours (:22-54) */
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        double up = 0.0, dn = 0.0;
        int nc = nbcount(&c->pc[i]), nj;
        for (nj = 1; nj <= nc; nj++) {
            int j = c->pc[i].influ[nj];
            const OPar *pj = &c->p[j];
            if (pi->type != pj->type && pi->type == 1 && pj->type == 2) {
                up = up + pj->mass * pj->porosity / pj->rho * c->pc[i].wij[nj][3];
                dn = dn + pj->mass / pj->rho * c->pc[i].wij[nj][3];
            }
        }
        if (pi->type == 1 && fabs(dn) > 1.0e-7) {
            pi->poro_cop = pi->porosity;
            pi->porosity = up / dn;
            if (l == 1) pi->poro_cop = pi->porosity;
        } else if (pi->type == 1 && fabs(dn) <= 1.0e-7) {
            pi->poro_cop = pi->porosity;
        }
    }
    /* loop 2: continuity (:56-166) */
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        double drhot = 0.0, up[3] = {0, 0, 0}, dn[3] = {0, 0, 0}, vxj[3];
        int nc = nbcount(&c->pc[i]), nj, k;
        for (nj = 1; nj <= nc; nj++) {
            int j = c->pc[i].influ[nj];
            const OPar *pj = &c->p[j];
            const double *w = c->pc[i].wij[nj];
            if (pi->type == pj->type && pi->type != 7 && pi->type != 0) {
                if (pi->type == 2 || pi->type == 4) {
                    for (k = 0; k < ndim; k++) {
                        up[k] = up[k] + pj->mass * (pi->vxp[k] - pj->vxp[k]) * w[k];
                        dn[k] = dn[k] + pj->mass * (pi->xp[k] - pj->xp[k]) * w[k];
                    }
                }
                if (pi->type == 1 || pi->type == 3)
                    for (k = 0; k < ndim; k++) up[k] = up[k] + pj->mass * (pi->vxp[k] - pj->vxp[k]) * w[k];
            } else if (pi->type >= 1 && pi->type <= 4 && pj->type == 0) {
                bndy_vel(c, i, j, 0, vxj);
                if (pi->type == 2 || pi->type == 4) {
                    for (k = 0; k < ndim; k++) {
                        up[k] = up[k] + pj->mass * (pi->vxp[k] - vxj[k]) * w[k];
                        dn[k] = dn[k] + pj->mass * (pi->xp[k] - pj->xp[k]) * w[k];
                    }
                }
                if (pi->type == 1 || pi->type == 3)
                    for (k = 0; k < ndim; k++) up[k] = up[k] + pj->mass * (pi->vxp[k] - vxj[k]) * w[k];
            }
        }
        if (pi->type != 7) {
            if (pi->type == 2 || pi->type == 4) {
                for (k = 0; k < ndim; k++) {
                    if (fabs(dn[k]) > 1.0e-7) drhot -= up[k] / dn[k];
                    else drhot += up[k];
                }
            }
            if (pi->type == 1 || pi->type == 3)
                for (k = 0; k < ndim; k++) drhot += up[k];
            if (pi->type == 1 && fabs(pi->porosity) > 0.0001)
                pi->rho = pi->rhop + drhot * dt +
                          pi->rhop * (pi->porosity - pi->poro_cop) / (1.0 - pi->poro_cop);
            else
                pi->rho = pi->rhop + drhot * dt;
        }
    }
    /* loop 3: normalisation over same-type neighbours (:168-188). Serial GS reads densities
     * already normalised for lower indices (quirk A1); JACOBI reads all pre-normalisation. */
    if (c->density_mode == OR_DENSITY_JACOBI)
        for (i = 0; i < n; i++) c->rho_tmp[i] = c->p[i].rho;
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        double up = 0.0, dn = 0.0;
        int nc = nbcount(&c->pc[i]), nj;
        for (nj = 1; nj <= nc; nj++) {
            int j = c->pc[i].influ[nj];
            const OPar *pj = &c->p[j];
            if (pi->type == pj->type) {
                double rj = c->density_mode == OR_DENSITY_JACOBI ? c->rho_tmp[j] : pj->rho;
                up = up + pj->mass * c->pc[i].wij[nj][3];
                dn = dn + pj->mass / rj * c->pc[i].wij[nj][3];
            }
        }
        if (pi->type != 7 && fabs(dn) > 1.0e-7) pi->rho = up / dn;
    }
}

/* ------------------------------------------------------------------ strain (CPP_7-1-strain.cpp:20-220) */
static void or_strain(or_ctx *c) {
    int n = c->n, dim = c->pro.ndim, i;
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        double up[9], dn[3], dvx[3], vxj[3];
        int nc = nbcount(&c->pc[i]), nj, k, a, b;
        if (c->l == c->pro.inip)
            for (k = 0; k < 6; k++) pi->eps[k] = 0.0;
        for (k = 0; k < 6; k++) {
            pi->deps[k] = 0.0;
            pi->dsig[k] = 0.0;
        }
        for (k = 0; k < 9; k++) up[k] = 0.0;
        for (a = 0; a < 3; a++) {
            dn[a] = 0.0;
            for (b = 0; b < 3; b++) {
                pi->weps[a][b] = 0.0;
                pi->divx[a][b] = 0.0;
            }
        }
        for (nj = 1; nj <= nc; nj++) {
            int j = c->pc[i].influ[nj], active = 0;
            const OPar *pj = &c->p[j];
            const double *w = c->pc[i].wij[nj];
            if (pi->type == pj->type && pi->type != 7 && pi->type != 0) {
                for (k = 0; k < 3; k++) dvx[k] = pi->vxp[k] - pj->vxp[k];
                active = 1;
            } else if (pi->type >= 1 && pi->type <= 4 && pj->type == 0) {
                bndy_vel(c, i, j, 0, vxj);
                for (k = 0; k < 3; k++) dvx[k] = pi->vxp[k] - vxj[k];
                active = 1;
            }
            if (!active) continue;
            for (a = 0; a < 3; a++)
                for (b = 0; b < 3; b++) up[3 * a + b] = up[3 * a + b] + pj->mass / pj->rho * dvx[a] * w[b];
            if (pi->type == 2 || pi->type == 4)
                for (k = 0; k < 3; k++) dn[k] = dn[k] + pj->mass / pj->rho * (pi->xp[k] - pj->xp[k]) * w[k];
        }
        if (pi->type == 1 || pi->type == 3) {
            for (a = 0; a < 3; a++)
                for (b = 0; b < 3; b++) pi->divx[a][b] = up[3 * a + b];
        } else {
            for (b = 0; b < 3; b++)
                if (fabs(dn[b]) > 1.0e-7)
                    for (a = 0; a < 3; a++) pi->divx[a][b] = up[3 * a + b] / dn[b];
        }
        pi->deps[0] = pi->divx[0][0];
        pi->deps[1] = pi->divx[1][1];
        pi->deps[2] = pi->divx[2][2];
        pi->deps[3] = pi->divx[0][1] + pi->divx[1][0];
        pi->deps[4] = pi->divx[1][2] + pi->divx[2][1];
        pi->deps[5] = pi->divx[2][0] + pi->divx[0][2];
        if (dim == 2) pi->deps[2] = pi->deps[4] = pi->deps[5] = 0.0;
        for (a = 0; a < 3; a++)
            for (b = 0; b < 3; b++)
                pi->weps[a][b] = a == b ? 0.0 : 0.5 * (pi->divx[a][b] - pi->divx[b][a]);
        if (dim == 2) {
            pi->weps[0][2] = pi->weps[1][2] = pi->weps[2][0] = pi->weps[2][1] = pi->weps[2][2] = 0.0;
        }
    }
}

/* cal_permeability (CPP_7-2-constitutive.cpp:355-382) with permeablel/permeabletl
 * (Class_particle.cpp:147-173). */
static void or_permeability(or_ctx *c) {
    int i;
    for (i = 0; i < c->n; i++) {
        OPar *pi = &c->p[i];
        const sphg_problem *P = &c->pro;
        if (pi->type != 2) continue;
        if (P->permtype == 0) pi->cop = c->ps[i].cop;
        else if (P->permtype == 1) pi->cop = P->pers;
        else if (P->permtype == 2) {
            double sitas = c->ps[i].sitas, sitar = c->ps[i].sitar;
            if (pi->satu < sitar) pi->cop = P->pers;
            else if (pi->satu > sitas) pi->cop = P->perl;
            else pi->cop = (P->perl - P->pers) * pi->satu / (sitas - sitar) + P->pers;
        } else if (P->permtype == 3) {
            if (pi->permtype == 1) pi->cop = P->pers;
            if (pi->permtype == 2) pi->cop = P->perm;
            if (pi->permtype == 3) pi->cop = P->perl;
        }
    }
}

/* ------------------------------------------------------------------ eos (CPP_7-1-eos_CR.cpp:12-178) */
/* loop 1 (:20-43): per-particle pressure from density */
static void or_eos1(or_ctx *c) {
    int n = c->n, i;
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        if (pi->type == 1) {
            pi->pre = pi->prep * (pow(pi->rho / c->pf[i].dens / c->pf[i].porosity, 7.0) - 1.0);
            pi->pre = pi->pre + pi->prep;
        }
        if (pi->type == 3) { /* quirk A9: the comma operator discards the density ratio */
            pi->pre = 1.403 * pi->prep * (-1.0);
            pi->pre = pi->pre + pi->prep;
        }
        if (pi->type == 0 && pi->matype == 1) {
            pi->pre = pi->prep * (pow(pi->rho / pi->rhop, 7.0) - 1.0);
            pi->pre = pi->pre + pi->prep;
        }
    }
}

/* loop 2 (:45-177): pore pressures of solids and boundary stress from neighbours */
static void or_eos2(or_ctx *c) {
    int n = c->n, i;
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        double u1 = 0, u2 = 0, u3 = 0, u4 = 0, d1 = 0, d2 = 0, d3 = 0, d4 = 0, us[6] = {0}, ds = 0;
        int nc = nbcount(&c->pc[i]), nj, k;
        for (nj = 1; nj <= nc; nj++) {
            int j = c->pc[i].influ[nj];
            const OPar *pj = &c->p[j];
            double w3 = c->pc[i].wij[nj][3];
            if (pi->type == pj->type) continue;
            if (pi->type == 2 && pj->type == 1) {
                u1 = u1 + pj->mass * pj->pre * w3 / pj->rho;
                d1 = d1 + pj->mass * w3 / pj->rho;
            }
            if (pi->type == 2 && pj->type == 3) {
                u2 = u2 + pj->mass * pj->pre * w3 / pj->rho;
                d2 = d2 + pj->mass * w3 / pj->rho;
            }
            if (pi->type == 4 && pj->type == 1) {
                u3 = u3 + pj->mass * pj->pre * w3 / pj->rho;
                d3 = d3 + pj->mass * w3 / pj->rho;
            }
            if (pi->type == 4 && pj->type == 3) {
                u4 = u4 + pj->mass * pj->pre * w3 / pj->rho;
                d4 = d4 + pj->mass * w3 / pj->rho;
            }
            if (pi->type == 0 && pj->type == 1) {
                u1 = u1 + pj->mass * pj->pre * w3 / pj->rho;
                d1 = d1 + pj->mass * w3 / pj->rho;
            }
            if (pi->type == 0 && pj->type == 3) {
                u2 = u2 + pj->mass * pj->pre * w3 / pj->rho;
                d2 = d2 + pj->mass * w3 / pj->rho;
            }
            if (pi->type == 0 && pj->type == 2) {
                for (k = 0; k < 6; k++) us[k] = us[k] + pj->mass * pj->sig[k] * w3 / pj->rho;
                ds = ds + pj->mass * w3 / pj->rho;
            }
        }
        if (pi->type == 2) {
            pi->prew = d1 != 0.0 ? u1 / d1 : u1;
            pi->prea = d2 != 0.0 ? u2 / d2 : u2;
        } else if (pi->type == 0) {
            pi->prew = d1 != 0.0 ? u1 / d1 : u1;
            pi->prea = d2 != 0.0 ? u2 / d2 : u2;
            for (k = 0; k < 6; k++) pi->sig[k] = ds != 0.0 ? us[k] / ds : us[k];
        } else if (pi->type == 4) {
            pi->prew = d3 != 0.0 ? u3 / d3 : u3;
            pi->prea = d4 != 0.0 ? u4 / d4 : u4;
        }
    }
}

/* ------------------------------------------------------------------ constitutive */
/* deq(), shearstrain(), poro_cal(), stress_adjust() (Class_particle.cpp:121-215) */
static void or_deq(OPar *p) {
    double sxx, syy, szz, j2;
    p->veps = p->eps[0] + p->eps[1] + p->eps[2];
    p->vsig = (p->sig[0] + p->sig[1] + p->sig[2]) / 3;
    p->divq = 0.0;
    sxx = p->sig[0] - p->vsig;
    syy = p->sig[1] - p->vsig;
    szz = p->sig[2] - p->vsig;
    j2 = 0.5 * (sxx * sxx + syy * syy + szz * szz) + p->sig[3] * p->sig[3] + p->sig[4] * p->sig[4] +
         p->sig[5] * p->sig[5];
    p->divq = sqrt(3 * fabs(j2));
    p->divr = p->vsig != 0.0 ? fabs(p->divq / p->vsig) : 0.0;
}
static void or_shearstrain(OPar *p) {
    double sxx = p->eps[0] - p->veps / 3.0, syy = p->eps[1] - p->veps / 3.0, szz = p->eps[2] - p->veps / 3.0;
    double j2 = 0.5 * (sxx * sxx + syy * syy + szz * szz) + p->eps[3] * p->eps[3] + p->eps[4] * p->eps[4] +
                p->eps[5] * p->eps[5];
    p->meps = sqrt(2.0 * fabs(j2));
}
/* TSSMP_before (CPP_7-3-TS.cpp:12-64): transformed stress of Yao's TS method with the
 * Matsuoka-Nakai SMP criterion from last step's principal stresses strep, mean stress vsig and
 * deviatoric stress divq; the original stress is kept in tssig. Quirk: I2 adds strep[2]*strep[1]
 * where the invariant has strep[2]*strep[0]. */
static void or_ts_before(OPar *p, double alpha) {
    double q = p->divq, qc = 0.0, sx[6], I1, I2, I3, temp;
    int k;
    p->tsqc = 0.0;
    for (k = 0; k < 3; k++) sx[k] = -p->sig[k] + p->vsig;
    for (k = 3; k < 6; k++) sx[k] = -p->sig[k];
    I1 = -(p->strep[0] + p->strep[1] + p->strep[2]);
    I2 = p->strep[0] * p->strep[1] + p->strep[1] * p->strep[2] + p->strep[2] * p->strep[1];
    I3 = -p->strep[0] * p->strep[1] * p->strep[2];
    if ((I1 * I1 - 3.0 * I2) >= 1.0e-7) {
        temp = 3.0 * sqrt((I1 * I2 - I3) / (I1 * I2 - 9.0 * I3)) - 1.0;
        if (fabs(temp) >= 1.0e-7)
            qc = alpha * sqrt(I1 * I1 - 3.0 * I2) + 2.0 * (1.0 - alpha) * I1 / (3.0 * sqrt((I1 * I2 - I3) / (I1 * I2 - 9.0 * I3)) - 1.0);
    }
    p->tsqc = qc;
    for (k = 0; k < 6; k++) p->tssig[k] = p->sig[k];
    if (q != 0.0) {
        for (k = 0; k < 3; k++) p->sig[k] = -I1 * 0.33333333 - qc / q * sx[k];
        for (k = 3; k < 6; k++) p->sig[k] = -qc / q * sx[k];
    }
}

/* TSSMP_after (CPP_7-3-TS.cpp:66-106): stress rate back to the real space; sig restored, tssig
 * left holding the transformed stress */
static void or_ts_after(OPar *p) {
    double q = p->divq, qc = p->tsqc, dv = (p->dsig[0] + p->dsig[1] + p->dsig[2]) / 3.0, t[6];
    int k;
    if (qc != 0.0) {
        for (k = 0; k < 3; k++) p->dsig[k] = dv + q / qc * (p->dsig[k] - dv);
        for (k = 3; k < 6; k++) p->dsig[k] = q / qc * p->dsig[k];
    }
    for (k = 0; k < 6; k++) {
        t[k] = p->tssig[k];
        p->tssig[k] = p->sig[k];
        p->sig[k] = t[k];
    }
}

static void or_poro_cal(OPar *p, double dt) {
    double de = (1 + p->e) * (p->deps[0] + p->deps[1] + p->deps[2]) * dt, et = p->e + de;
    if (et < 0.0) et = 0.0;
    p->porosity = et / (1 + et);
    p->e = et;
}
static void or_stress_adjust(OPar *p) {
    if (p->vsig > 0.0) {
        double t = -p->vsig;
        p->sig[0] = p->sig[0] + t;
        p->sig[1] = p->sig[1] + t;
        p->sig[2] = p->sig[2] + t;
        p->sig[3] = p->sig[4] = p->sig[5] = 0.0;
    }
}

static void iso_tangent(double d[6][6], double e0, double poi) {
    double dlam = e0 * poi / (1.0 + poi) / (1.0 - 2.0 * poi), dmu = e0 * 0.5 / (1.0 + poi), dmu2 = dmu * 2.0;
    d[0][0] = d[1][1] = d[2][2] = dlam + dmu2;
    d[0][1] = d[0][2] = d[1][0] = d[1][2] = d[2][0] = d[2][1] = dlam;
    d[3][3] = d[4][4] = d[5][5] = dmu;
}

/* mean-stress dependent Young's modulus shared by the soil models (CPP_8-1-consfactor.cpp:30-41) */
static double soil_modulus(const OPar *p, const sphg_soil *s, double *sm_out) {
    double sm = ((-p->sig[0]) + (-p->sig[1]) + (-p->sig[2])) * 0.3333333333;
    if (sm < 1.0) sm = 1.0;
    if (sm_out) *sm_out = sm;
    if (s->e > 1.0) return s->e;
    return 3 * (1 - 2 * s->poi) * (1 + s->eps0) * fabs(sm) / s->kapa;
}

/* elastic (CPP_8-1-consfactor.cpp:17-59) */
static void or_elastic(double d[6][6], const OPar *p, const sphg_soil *s) {
    iso_tangent(d, soil_modulus(p, s, NULL), s->poi);
}

/* rank-one plastic correction D -= d * w (x) z and plastic strain rate (CPP_8-1-consfactor.cpp:236-273) */
static void plastic_update(double D[6][6], double d, const double w[6], const double z[6]) {
    int a, b;
    for (a = 0; a < 6; a++)
        for (b = 0; b < 6; b++) D[a][b] = D[a][b] - d * w[a] * z[b];
}

/* sub_camclay (CPP_8-1-consfactor.cpp:129-281) */
static void or_sub_camclay(double D[6][6], const OPar *p, const sphg_soil *s, double adps[6]) {
    double sm, e0, poi = s->poi, zmf = s->zmf, ann = s->ann, eps0 = s->eps0;
    double sig[6], ymf, hee, dlam, dmu, dmu2, dx, dy, dz, dxy, dyz, dzx, sjj, tt1, tt2;
    double f11, f22, f33, f12, f23, f31, roue, ttt, hdd, fkk, fij, cc7, d, df, dg, ramda;
    int k;
    for (k = 0; k < 6; k++) sig[k] = -p->sig[k];
    ymf = 3.0 * (zmf - 1.0) / (zmf + 2.0) / s->fs;
    sm = (sig[0] + sig[1] + sig[2]) * 0.3333333333;
    if (sm < 1.0) sm = 1.0;
    e0 = s->e > 1.0 ? s->e : 3 * (1 - 2 * poi) * (1 + eps0) * fabs(sm) / s->kapa;
    hee = (s->ramda - s->kapa) / (1.0 + eps0);
    dlam = e0 * poi / (1.0 + poi) / (1.0 - 2.0 * poi);
    dmu = e0 * 0.5 / (1.0 + poi);
    dmu2 = dmu * 2.0;
    iso_tangent(D, e0, poi);
    dx = sig[0] - sm; dy = sig[1] - sm; dz = sig[2] - sm;
    dxy = sig[3]; dyz = sig[4]; dzx = sig[5];
    sjj = 0.5 * (dx * dx + dy * dy + dz * dz) + dxy * dxy + dyz * dyz + dzx * dzx;
    tt1 = 1.0 / (ymf * ymf * sm * sm + 3.0 * sjj);
    tt2 = 2.0 * ymf * ymf * sm * tt1 - 1.0 / sm;
    f11 = tt2 / 3.0 + dx * tt1 * 3.0;
    f22 = tt2 / 3.0 + dy * tt1 * 3.0;
    f33 = tt2 / 3.0 + dz * tt1 * 3.0;
    f12 = 3.0 * dxy * tt1;
    f23 = 3.0 * dyz * tt1;
    f31 = 3.0 * dzx * tt1;
    roue = (1 + eps0) * log(s->sme0 / sm);
    if (roue < 0.0) roue = 0.0;
    ttt = ann * roue * roue;
    hdd = f11 + f22 + f33 + ttt / sm;
    fkk = f11 + f22 + f33;
    fij = 2.0 * (f11 * f11 + f22 * f22 + f33 * f33) + f12 * f12 + f23 * f23 + f31 * f31;
    cc7 = dlam * (fkk * fkk) + dmu * fij + hdd / hee;
    d = 1.0 / cc7;
    df = dlam * fkk;
    dg = dlam * fkk;
    ramda = (dmu2 * f11 + dlam * fkk) * p->deps[0] + (dmu2 * f22 + dlam * fkk) * p->deps[1] +
            (dmu2 * f33 + dlam * fkk) * p->deps[2] +
            dmu * (f12 * p->deps[3] + f23 * p->deps[4] + f31 * p->deps[5]);
    if (ramda < 0.0) {
        double w[6] = {dg + dmu2 * f11, dg + dmu2 * f22, dg + dmu2 * f33, dmu * f12, dmu * f23, dmu * f31};
        double z[6] = {df + dmu2 * f11, df + dmu2 * f22, df + dmu2 * f33, dmu * f12, dmu * f23, dmu * f31};
        plastic_update(D, d, w, z);
        adps[0] = ramda * d * f11; adps[1] = ramda * d * f22; adps[2] = ramda * d * f33;
        adps[3] = ramda * d * f12; adps[4] = ramda * d * f23; adps[5] = ramda * d * f31;
    }
}

/* dpmodel1 (CPP_8-1-consfactor.cpp:284-448) */
static void or_dpmodel1(double D[6][6], const OPar *p, const sphg_soil *s, double adps[6]) {
    double fai = atan(tan(s->fai * 0.0174532925166667) / s->fs);
    double ci = s->c / s->fs, poi = s->poi, eps0 = s->eps0;
    double sig[6], sm, e0, dlam, dmu, dmu2, dx, dy, dz, dxy, dyz, dzx, sjj, sj, yalf, ykap;
    int k, a, b;
    for (k = 0; k < 6; k++) sig[k] = -p->sig[k];
    for (a = 0; a < 6; a++)
        for (b = 0; b < 6; b++) D[a][b] = 0.0;
    sm = (sig[0] + sig[1] + sig[2]) * 0.3333333333;
    if (sm < 1.0) sm = 1.0;
    e0 = s->e > 1.0 ? s->e : 3 * (1 - 2 * poi) * (1 + eps0) * fabs(sm) / s->kapa;
    dlam = e0 * poi / (1.0 + poi) / (1.0 - 2.0 * poi);
    dmu = e0 * 0.5 / (1.0 + poi);
    dmu2 = dmu * 2.0;
    iso_tangent(D, e0, poi);
    dx = sig[0] - sm; dy = sig[1] - sm; dz = sig[2] - sm;
    dxy = sig[3]; dyz = sig[4]; dzx = sig[5];
    sjj = 0.5 * (dx * dx + dy * dy + dz * dz) + dxy * dxy + dyz * dyz + dzx * dzx;
    sj = sqrt(sjj);
    yalf = sin(fai) / sqrt(9.0 + 3.0 * sin(fai) * sin(fai));
    ykap = 3.0 * ci * cos(fai) / sqrt(9.0 + 3.0 * sin(fai) * sin(fai));
    if ((sj - yalf * sm * 3.0 - ykap) >= 0.0) {
        double f11 = -yalf + dx / sj * 0.5, f22 = -yalf + dy / sj * 0.5, f33 = -yalf + dz / sj * 0.5;
        double f12 = dxy / sj * 0.5, f23 = dyz / sj * 0.5, f31 = dzx / sj * 0.5;
        double fkk = f11 + f22 + f33, yd = 0.0;
        double g11 = -yd + dx / sj * 0.5, g22 = -yd + dy / sj * 0.5, g33 = -yd + dz / sj * 0.5;
        double g12 = dxy / sj * 0.5, g23 = dyz / sj * 0.5, g31 = dzx / sj * 0.5;
        double hdd = g11 + g22 + g33;
        double fij = 2.0 * (f11 * g11 + f22 * g22 + f33 * g33) + f12 * g12 + f23 * g23 + f31 * g31;
        double cc7 = dlam * (fkk * hdd) + dmu * fij + hdd, d = 1.0 / cc7, df = dlam * fkk, dg = dlam * hdd;
        double ramda = (dmu2 * f11 - 3.0 * dlam * yalf) * p->deps[0] +
                       (dmu2 * f22 - 3.0 * dlam * yalf) * p->deps[1] +
                       (dmu2 * f33 - 3.0 * dlam * yalf) * p->deps[2] +
                       dmu * (f12 * p->deps[3] + f23 * p->deps[4] + f31 * p->deps[5]);
        if (ramda * d < 0.0) {
            double w[6] = {dg + dmu2 * g11, dg + dmu2 * g22, dg + dmu2 * g33, dmu * g12, dmu * g23, dmu * g31};
            double z[6] = {df + dmu2 * f11, df + dmu2 * f22, df + dmu2 * f33, dmu * f12, dmu * f23, dmu * f31};
            plastic_update(D, d, w, z);
            adps[0] = ramda * d * f11; adps[1] = ramda * d * f22; adps[2] = ramda * d * f33;
            adps[3] = ramda * d * f12; adps[4] = ramda * d * f23; adps[5] = ramda * d * f31;
        }
    }
}

/* dpmodel_bui: DP with a separate dilatancy angle (CPP_8-1-consfactor.cpp:451-615) */
static void or_dpmodel_bui(double D[6][6], const OPar *p, const sphg_soil *s, double adps[6]) {
    double fai = atan(tan(s->fai * 0.0174532925166667) / s->fs);
    double fai_dil = atan(tan(s->beta * 0.0174532925166667) / s->fs);
    double ci = s->c / s->fs, poi = s->poi, eps0 = s->eps0;
    double sig[6], sm, e0, dlam, dmu, dmu2, dx, dy, dz, dxy, dyz, dzx, sjj, sj, yalf, ykap;
    int k, a, b;
    for (k = 0; k < 6; k++) sig[k] = -p->sig[k];
    for (a = 0; a < 6; a++)
        for (b = 0; b < 6; b++) D[a][b] = 0.0;
    sm = (sig[0] + sig[1] + sig[2]) * 0.3333333333;
    if (sm < 1.0) sm = 1.0;
    e0 = s->e > 1.0 ? s->e : 3 * (1 - 2 * poi) * (1 + eps0) * fabs(sm) / s->kapa;
    dlam = e0 * poi / (1.0 + poi) / (1.0 - 2.0 * poi);
    dmu = e0 * 0.5 / (1.0 + poi);
    dmu2 = dmu * 2.0;
    iso_tangent(D, e0, poi);
    dx = sig[0] - sm; dy = sig[1] - sm; dz = sig[2] - sm;
    dxy = sig[3]; dyz = sig[4]; dzx = sig[5];
    sjj = 0.5 * (dx * dx + dy * dy + dz * dz) + dxy * dxy + dyz * dyz + dzx * dzx;
    sj = sqrt(sjj);
    yalf = sin(fai) / sqrt(9.0 + 3.0 * sin(fai) * sin(fai));
    ykap = 3.0 * ci * cos(fai) / sqrt(9.0 + 3.0 * sin(fai) * sin(fai));
    if ((sj - yalf * sm * 3.0 - ykap) >= 0.0) {
        double f11 = -yalf + dx / sj * 0.5, f22 = -yalf + dy / sj * 0.5, f33 = -yalf + dz / sj * 0.5;
        double f12 = dxy / sj * 0.5, f23 = dyz / sj * 0.5, f31 = dzx / sj * 0.5;
        double fkk = f11 + f22 + f33;
        double yd = sin(fai_dil) / sqrt(9.0 + 3.0 * sin(fai_dil) * sin(fai_dil));
        double g11 = -yd + dx / sj * 0.5, g22 = -yd + dy / sj * 0.5, g33 = -yd + dz / sj * 0.5;
        double g12 = dxy / sj * 0.5, g23 = dyz / sj * 0.5, g31 = dzx / sj * 0.5;
        double hdd = g11 + g22 + g33;
        double fij = 2.0 * (f11 * g11 + f22 * g22 + f33 * g33) + f12 * g12 + f23 * g23 + f31 * g31;
        double cc7 = dlam * (fkk * hdd) + dmu * fij + hdd, d = 1.0 / cc7, df = dlam * fkk, dg = dlam * hdd;
        double ramda = (dmu2 * f11 - 3.0 * dlam * yalf) * p->deps[0] + (dmu2 * f22 - 3.0 * dlam * yalf) * p->deps[1] +
                       (dmu2 * f33 - 3.0 * dlam * yalf) * p->deps[2] +
                       dmu * (f12 * p->deps[3] + f23 * p->deps[4] + f31 * p->deps[5]);
        if (ramda * d < 0.0) {
            double w[6] = {dg + dmu2 * g11, dg + dmu2 * g22, dg + dmu2 * g33, dmu * g12, dmu * g23, dmu * g31};
            double z[6] = {df + dmu2 * f11, df + dmu2 * f22, df + dmu2 * f33, dmu * f12, dmu * f23, dmu * f31};
            plastic_update(D, d, w, z);
            adps[0] = ramda * d * f11; adps[1] = ramda * d * f22; adps[2] = ramda * d * f33;
            adps[3] = ramda * d * f12; adps[4] = ramda * d * f23; adps[5] = ramda * d * f31;
        }
    }
}

/* sub_camunsatu: unsaturated Cam-Clay (CPP_8-4-consfactor-umcc.cpp:17-180); erfs before the
 * tangent update, 0.333333333 as written there */
static void or_sub_camunsatu(double D[6][6], double erfs[6], const OPar *p, const sphg_soil *s, double adps[6]) {
    double poi = s->poi, eps0 = s->eps0, sm, e0, hee, dlam, dmu, dmu2, sig[6];
    double qq = (s->epsr - s->epss) / (s->sitas - s->sitar), ymf, dx, dy, dz, dxy, dyz, dzx, sjj, tt1, tt2;
    double f11, f22, f33, f12, f23, f31, rous, roue, ttt, hdd, fkk, fij, cc7, d, df, dg, ramda, omake;
    int k;
    for (k = 0; k < 6; k++) sig[k] = -p->sig[k];
    ymf = 3.0 * (s->zmf - 1.0) / (s->zmf + 2.0) / s->fs;
    sm = (sig[0] + sig[1] + sig[2]) * 0.333333333;
    if (sm < 1.0) sm = 1.0;
    e0 = s->e > 1.0 ? s->e : 3 * (1 - 2 * poi) * (1 + eps0) * fabs(sm) / s->kapa;
    hee = (s->ramda - s->kapa) / (1.0 + eps0);
    dlam = e0 * poi / (1.0 + poi) / (1.0 - 2.0 * poi);
    dmu = e0 * 0.5 / (1.0 + poi);
    dmu2 = dmu * 2.0;
    iso_tangent(D, e0, poi);
    dx = sig[0] - sm; dy = sig[1] - sm; dz = sig[2] - sm;
    dxy = sig[3]; dyz = sig[4]; dzx = sig[5];
    sjj = 0.5 * (dx * dx + dy * dy + dz * dz) + dxy * dxy + dyz * dyz + dzx * dzx;
    tt1 = 1.0 / (ymf * ymf * sm * sm + 3.0 * sjj);
    tt2 = 2.0 * ymf * ymf * sm * tt1 - 1.0 / sm;
    f11 = tt2 * 0.333333333 + dx * tt1 * 3.0;
    f22 = tt2 * 0.333333333 + dy * tt1 * 3.0;
    f33 = tt2 * 0.333333333 + dz * tt1 * 3.0;
    f12 = 3.0 * dxy * tt1;
    f23 = 3.0 * dyz * tt1;
    f31 = 3.0 * dzx * tt1;
    rous = qq * (s->sitas - p->satu);
    roue = hee * (1.0 + s->eps0) * log(s->sme0 / sm);
    ttt = s->ann * roue + s->bnn * rous;
    ttt = pow(ttt, s->beta);
    hdd = f11 + f22 + f33 + ttt / sm;
    fkk = f11 + f22 + f33;
    fij = 2.0 * (f11 * f11 + f22 * f22 + f33 * f33) + f12 * f12 + f23 * f23 + f31 * f31;
    cc7 = dlam * (fkk * fkk) + dmu * fij + hdd / hee;
    d = 1.0 / cc7;
    df = dlam * fkk;
    dg = dlam * fkk;
    ramda = (dmu2 * f11 + dlam * fkk) * p->deps[0] + (dmu2 * f22 + dlam * fkk) * p->deps[1] +
            (dmu2 * f33 + dlam * fkk) * p->deps[2] + dmu * (f12 * p->deps[3] + f23 * p->deps[4] + f31 * p->deps[5]);
    omake = qq * p->dsatu / (s->ramda - s->kapa);
    ramda = ramda + omake;
    if (ramda < 0.0) {
        erfs[0] = omake * d * (D[0][0] * f11 + D[0][1] * f22 + D[0][2] * f33);
        erfs[1] = omake * d * (D[1][0] * f11 + D[1][1] * f22 + D[1][2] * f33);
        erfs[2] = omake * d * (D[2][0] * f11 + D[2][1] * f22 + D[2][2] * f33);
        erfs[3] = omake * d * (D[3][3] * f12);
        erfs[4] = omake * d * (D[4][4] * f23);
        erfs[5] = omake * d * (D[5][5] * f31);
        {
            double w[6] = {dg + dmu2 * f11, dg + dmu2 * f22, dg + dmu2 * f33, dmu * f12, dmu * f23, dmu * f31};
            double z[6] = {df + dmu2 * f11, df + dmu2 * f22, df + dmu2 * f33, dmu * f12, dmu * f23, dmu * f31};
            plastic_update(D, d, w, z);
        }
        adps[0] = ramda * d * f11; adps[1] = ramda * d * f22; adps[2] = ramda * d * f33;
        adps[3] = ramda * d * f12; adps[4] = ramda * d * f23; adps[5] = ramda * d * f31;
    }
}

/* detA: recursive cofactor determinant (CPP_8-5-consfactor-focc.cpp:275-302) */
static double or_detA(double arr[6][6], int n) {
    double bak[6][6], sum = 0;
    int i, j, k, cc;
    if (n == 1) return arr[0][0];
    for (i = 0; i < n; i++) {
        for (j = 1; j < n; j++) {
            for (cc = 0, k = 0; k < n; k++) {
                if (k == i) continue;
                bak[j - 1][cc++] = arr[j][k];
            }
        }
        sum += (i % 2 == 0 ? 1 : -1) * arr[0][i] * or_detA(bak, n - 1);
    }
    return sum;
}

/* getAStart: adjugate by cofactors (CPP_8-5-consfactor-focc.cpp:305-336) */
static void or_adjugate(double a[6][6], int n, double ans[6][6]) {
    int i, j, k, t;
    double tmp[6][6];
    if (n == 1) {
        ans[0][0] = 1.0;
        return;
    }
    for (i = 0; i < n; i++)
        for (j = 0; j < n; j++) {
            for (k = 0; k < n - 1; k++)
                for (t = 0; t < n - 1; t++) tmp[k][t] = a[k >= i ? k + 1 : k][t >= j ? t + 1 : t];
            ans[j][i] = or_detA(tmp, n - 1);
            if ((i + j) % 2 == 1) ans[j][i] = -ans[j][i];
        }
}

/* fractional_mcc (CPP_8-5-consfactor-focc.cpp:13-240); quirks: sm = sum / 0.33333333 (A13),
 * e approximated by 2.71828, p0 = pa = 98000 */
static void or_fractional_mcc(double D[6][6], double C[6][6], const OPar *p, const sphg_soil *s) {
    double sig[6], sij[6], sm, q, s_val, skl, smn[3][3], sita, poi = s->poi, h1 = s->damp1, h2 = s->damp2;
    double ramda = s->ramda, mc = s->zmf, alpha = s->ann, beta = s->bnn, c = s->c, zeta = s->zet, etao = s->etao;
    double g0 = s->g0, p0 = 98000, pa = 98000, m, ita, df, dg, pc, qc, e = p->e, mf, order, mij[6], nij[6], hp;
    double P[6][6], T[6][6], adj[6][6], g, e0, det_a, dlam, dmu, dmu2, ramda_cc, ymf, tt1, tt2;
    double f11, f22, f33, f12, f23, f31, fkk;
    int a, b, kk;
    for (a = 0; a < 6; a++) {
        sig[a] = -p->sig[a];
        for (b = 0; b < 6; b++) C[a][b] = P[a][b] = 0.0;
    }
    sm = (sig[0] + sig[1] + sig[2]) / 0.33333333;
    for (a = 0; a < 6; a++) sij[a] = a < 3 ? sig[a] - sm : sig[a];
    q = 0.5 * (sij[0] * sij[0] + sij[1] * sij[1] + sij[2] * sij[2]) + sij[3] * sij[3] + sij[4] * sij[4] + sij[5] * sij[5];
    q = sqrt(3.0 * q);
    ita = q / sm;
    if (ita == 0.0) {
        ita = 0.000001;
        q = ita * sm;
    }
    g = g0 * (2.97 - e) * (2.97 - e) / (1.0 + e) * sqrt(sm * pa);
    e0 = 2.0 * g * (1.0 + poi);
    ramda_cc = 0.0;
    ymf = mc;
    dlam = e0 * poi / (1.0 + poi) / (1.0 - 2.0 * poi);
    dmu = e0 * 0.5 / (1.0 + poi);
    dmu2 = dmu * 2.0;
    tt1 = 1.0 / (ymf * ymf * sm * sm + q * q);
    tt2 = 2.0 * ymf * ymf * sm * tt1 - 1.0 / sm;
    f11 = tt2 * 0.33333333 + sij[0] * tt1 * 3.0;
    f22 = tt2 * 0.33333333 + sij[1] * tt1 * 3.0;
    f33 = tt2 * 0.33333333 + sij[2] * tt1 * 3.0;
    f12 = 3.0 * sij[3] * tt1;
    f23 = 3.0 * sij[4] * tt1;
    f31 = 3.0 * sij[5] * tt1;
    fkk = f11 + f22 + f33;
    ramda_cc = (dmu2 * f11 + dlam * fkk) * p->deps[0] + (dmu2 * f22 + dlam * fkk) * p->deps[1] +
               (dmu2 * f33 + dlam * fkk) * p->deps[2] + dmu * (f12 * p->deps[3] + f23 * p->deps[4] + f31 * p->deps[5]);
    smn[0][0] = sij[0]; smn[0][1] = sij[3]; smn[0][2] = sij[5];
    smn[1][0] = sij[3]; smn[1][1] = sij[1]; smn[1][2] = sij[4];
    smn[2][0] = sij[5]; smn[2][1] = sij[4]; smn[2][2] = sij[2];
    s_val = 0.0;
    for (a = 0; a < 3; a++)
        for (b = 0; b < 3; b++)
            for (kk = 0; kk < 3; kk++) s_val = s_val + smn[a][b] * smn[b][kk] * smn[kk][a];
    skl = 0.0;
    for (a = 0; a < 3; a++)
        for (b = 0; b < 3; b++) skl = skl + smn[a][b] * smn[b][a];
    if (skl == 0.0) skl = 0.000001;
    sita = 9.0 * s_val * 0.5 / pow(1.5 * skl, 1.5);
    if (sita < -1.0) sita = -1.0;
    else if (sita > 1.0) sita = 1.0;
    sita = asin(sita) * 0.33333333;
    c = pow(c, 4);
    m = 2 * c / (1 + c + (1 - c) * sin(3.0 * sita));
    m = pow(m, 0.25);
    m = mc * m;
    pc = pa * pow((etao - e) / ramda, 1 / zeta);
    qc = q + m * (sm - pc);
    df = (sm - p0 / 2) * mc * mc / q;
    dg = ((sm - pc) + (2 - alpha) * (pc - sm * 0.5 - q * q / mc / mc / sm / 2)) / ((q - qc) + (2 - alpha) * qc);
    dg = pow(m, alpha + 1) * dg;
    for (a = 0; a < 6; a++) {
        double base = 3.0 * sij[a] * 0.5 / q;
        mij[a] = a < 3 ? 1.0 / sqrt(1.0 + df * df) * (base + df * 0.33333333) : 1.0 / sqrt(1.0 + df * df) * base;
        nij[a] = a < 3 ? 1.0 / sqrt(1.0 + dg * dg) * (base + dg * 0.33333333) : 1.0 / sqrt(1.0 + dg * dg) * base;
    }
    order = beta * (e - etao + ramda * pow(sm / pa, zeta));
    mf = mc * pow(2.71828, -order);
    hp = (h1 - h2 * e) * (1 - e) * g * (mf - ita) / ita * pow(2.71828, order);
    C[0][0] = C[1][1] = C[2][2] = 1.0 / e0;
    C[0][1] = C[0][2] = C[1][0] = C[1][2] = C[2][0] = C[2][1] = -poi / e0;
    C[3][3] = C[4][4] = C[5][5] = 1.0 / g;
    for (a = 0; a < 6; a++)
        for (b = 0; b < 6; b++) P[a][b] = nij[a] * mij[b] / hp;
    for (a = 0; a < 6; a++)
        for (b = 0; b < 6; b++) T[a][b] = ramda_cc < 0.0 ? C[a][b] + P[a][b] : C[a][b];
    det_a = or_detA(T, 6);
    or_adjugate(T, 6, adj);
    for (a = 0; a < 6; a++)
        for (b = 0; b < 6; b++) {
            if (det_a != 0.0)
                D[a][b] = adj[a][b] / det_a;
            else {
                det_a = or_detA(C, 6);
                or_adjugate(C, 6, adj);
                D[a][b] = adj[a][b] / det_a;
            }
        }
}

/* fractional_plastic (CPP_8-5-consfactor-focc.cpp:243-259) */
static void or_fractional_plastic(const double dsig[6], const double deps[6], double adps[6], double C[6][6]) {
    double ee[6];
    int a, b;
    for (a = 0; a < 6; a++) {
        ee[a] = 0.0;
        for (b = 0; b < 6; b++) ee[a] = ee[a] + C[a][b] * dsig[b];
    }
    for (a = 0; a < 6; a++) adps[a] = deps[a] - ee[a];
}

/* DPModel_check: stress-scaling return onto the DP cone (CPP_8-1-consfactor.cpp:779-819); sj is
 * taken from the stress before the increment (as the reference does). */
static void or_dp_check(OPar *p, const sphg_soil *s, double dt) {
    double fai = atan(tan(s->fai * 0.0174532925166667) / s->fs), ci = s->c / s->fs;
    double sm = (p->sig[0] + p->sig[1] + p->sig[2]) / 3.0, ts[6], i1, af, kc, rn, sj;
    double dx = p->sig[0] - sm, dy = p->sig[1] - sm, dz = p->sig[2] - sm;
    int k;
    sj = sqrt(0.5 * (dx * dx + dy * dy + dz * dz) + p->sig[3] * p->sig[3] + p->sig[4] * p->sig[4] +
              p->sig[5] * p->sig[5]);
    for (k = 0; k < 6; k++) ts[k] = p->sig[k] + p->dsig[k] * dt;
    i1 = ts[0] + ts[1] + ts[2];
    af = sin(fai) / sqrt(9.0 + 3.0 * sin(fai) * sin(fai));
    kc = 3.0 * ci * cos(fai) / sqrt(9.0 + 3.0 * sin(fai) * sin(fai));
    rn = (-af * i1 + kc) / sj;
    if (rn < 1.0) {
        for (k = 0; k < 3; k++) ts[k] = rn * (ts[k] - i1 / 3.0) + i1 / 3.0;
        for (k = 3; k < 6; k++) ts[k] = rn * ts[k];
    }
    for (k = 0; k < 6; k++) p->dsig[k] = (ts[k] - p->sig[k]) / dt;
}

/* Jaumann correction of the stress rate (CPP_7-2-constitutive.cpp:236-253) */
static void jaumann(OPar *p) {
    double (*w)[3] = p->weps, *s = p->sig;
    p->dsig[0] = p->dsig[0] + (w[0][0] * s[0] + w[0][1] * s[3] + w[0][2] * s[5]) -
                 (s[0] * w[0][0] + s[3] * w[1][0] + s[5] * w[2][0]);
    p->dsig[1] = p->dsig[1] + (w[1][0] * s[3] + w[1][1] * s[1] + w[1][2] * s[4]) -
                 (s[3] * w[0][1] + s[1] * w[1][1] + s[4] * w[2][1]);
    p->dsig[2] = p->dsig[2] + (w[2][0] * s[5] + w[2][1] * s[4] + w[2][2] * s[2]) -
                 (s[5] * w[0][2] + s[4] * w[1][2] + s[2] * w[2][2]);
    p->dsig[3] = p->dsig[3] + (w[0][0] * s[3] + w[0][1] * s[1] + w[0][2] * s[4]) -
                 (s[0] * w[0][1] + s[3] * w[1][1] + s[5] * w[2][1]);
    p->dsig[4] = p->dsig[4] + (w[1][0] * s[5] + w[1][1] * s[4] + w[1][2] * s[2]) -
                 (s[3] * w[0][2] + s[1] * w[1][2] + s[4] * w[2][2]);
    p->dsig[5] = p->dsig[5] + (w[0][0] * s[5] + w[0][1] * s[4] + w[0][2] * s[2]) -
                 (s[0] * w[0][2] + s[3] * w[1][2] + s[5] * w[2][2]);
}

/* the matype tests of strain_to_stress (CPP_7-2-constitutive.cpp:104-235): model m for
 * matype m, 10+m, 20+m and 50+m; any other matype selects no model (only Jaumann acts). */
static int model_of(int matype) {
    if ((matype >= 0 && matype < 30) || (matype >= 50 && matype < 60)) return matype % 10;
    return -1;
}

static void apply_tangent(OPar *p, double D[6][6]) {
    int a, b;
    for (a = 0; a < 6; a++)
        for (b = 0; b < 6; b++) p->dsig[a] += D[a][b] * p->deps[b];
}

/* pParStiff[i].depp0 = dept (CPP_7-2-constitutive.cpp:112-292): only its diagonal is read later */
static void keep_depp0(or_ctx *c, int i, double D[6][6]) {
    c->depp0[i][0] = D[0][0];
    c->depp0[i][1] = D[1][1];
    c->depp0[i][2] = D[2][2];
}

/* strain_to_stress (CPP_7-2-constitutive.cpp:13-321) for the models the workloads use:
 * Newtonian/Bingham fluid, elastic (0, 6), SB (1), DP (2), AMCC no-op (4, quirk A13), structure. */
static int or_strain_to_stress(or_ctx *c) {
    int n = c->n, i, bad = 0;
    double dt = c->pro.dt;
#pragma omp parallel for schedule(static) num_threads(c->nthreads) reduction(max : bad)
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        /* Cf: the reference's OpenMP-private cijkl is never assigned before fractional_plastic
         * runs for matype 8/18/28 while l < inip (CPP_7-2-constitutive.cpp:17, :24, :255-258);
         * fresh thread stacks hold zeros there in practice, so zero is used (denormal residues
         * of the reference's run-to-run garbage aside) */
        double D[6][6], erfs[6], adps[6], deps[6], Cf[6][6] = {{0}};
        int k, a, b;
        if (pi->type == 1 || pi->type == 3) {
            const sphg_fluid *f = &c->pf[i];
            double vis = f->vis;
            if (pi->type == 1) {
                double dm = (pi->deps[0] + pi->deps[1] + pi->deps[2]) / 3.0, de;
                pi->eps[0] = pi->deps[0] - dm;
                pi->eps[1] = pi->deps[1] - dm;
                pi->eps[2] = pi->deps[2] - dm;
                de = 0.666666666666667 * (pi->eps[0] * pi->eps[0] + pi->eps[1] * pi->eps[1] +
                                          pi->eps[2] * pi->eps[2] + pi->deps[3] * pi->deps[3] * 0.25 +
                                          pi->deps[4] * pi->deps[4] * 0.25 + pi->deps[5] * pi->deps[5] * 0.25);
                de = sqrt(de);
                if (f->bhtype == 1) {
                    double fai = f->fai * 0.0174532925199, tao = pi->pre * tan(fai) + f->c;
                    vis = de < f->strain_min ? f->vis + tao / f->strain_min : f->vis + tao / de;
                }
            }
            for (k = 0; k < 3; k++)
                deps[k] = 2.0 * pi->deps[k] - 2.0 * (pi->deps[0] + pi->deps[1] + pi->deps[2]) / 3.0;
            for (k = 3; k < 6; k++) deps[k] = pi->deps[k];
            for (k = 0; k < 6; k++) pi->sig[k] = k < 3 ? -pi->pre - vis * deps[k] : -vis * deps[k];
        } else if (pi->type == 2) {
            const sphg_soil *s = &c->ps[i];
            int mt = model_of(pi->matype), known = 1;
            for (a = 0; a < 6; a++) {
                erfs[a] = adps[a] = 0.0;
                for (b = 0; b < 6; b++) D[a][b] = 0.0;
            }
            if (c->pro.flts >= 1) or_ts_before(pi, c->pro.alphats);
            if (c->l < c->pro.inip) {
                or_elastic(D, pi, s);
                apply_tangent(pi, D);
                keep_depp0(c, i, D);
            } else if (mt == 0 || mt == 6) {
                or_elastic(D, pi, s);
                apply_tangent(pi, D);
                keep_depp0(c, i, D);
            } else if (mt == 1) {
                or_sub_camclay(D, pi, s, adps);
                apply_tangent(pi, D);
                keep_depp0(c, i, D);
            } else if (mt == 2) {
                or_dpmodel1(D, pi, s, adps);
                apply_tangent(pi, D);
                keep_depp0(c, i, D);
                or_dp_check(pi, s, dt);
            } else if (mt == 3) {
                or_sub_camunsatu(D, erfs, pi, s, adps);
                apply_tangent(pi, D);
                keep_depp0(c, i, D);
            } else if (mt == 4) {
                apply_tangent(pi, D); /* AMCC body is commented out: zero tangent, no erfs */
                keep_depp0(c, i, D);
            } else if (mt == 5) {
                or_dpmodel_bui(D, pi, s, adps);
                apply_tangent(pi, D);
                keep_depp0(c, i, D);
                or_dp_check(pi, s, dt);
            } else if (mt == 8) {
                or_fractional_mcc(D, Cf, pi, s);
                apply_tangent(pi, D);
                keep_depp0(c, i, D);
            } else if (mt >= 0) {
                known = 0; /* 7 viscous-elastic and 9 Mohr-Coulomb are not restated */
            }
            if (!known) bad = SPHG_E_UNSUPPORTED;
            jaumann(pi);
            if (pi->matype == 8 || pi->matype == 18 || pi->matype == 28)
                or_fractional_plastic(pi->dsig, pi->deps, adps, Cf);
            if (c->pro.flts >= 1) or_ts_after(pi);
            for (k = 0; k < 6; k++) {
                pi->adps[k] = pi->adps[k] + adps[k] * dt;
                pi->eps[k] = pi->eps[k] + pi->deps[k] * dt;
                pi->sig[k] = pi->sig[k] + pi->dsig[k] * dt + erfs[k];
            }
            or_deq(pi);
            or_shearstrain(pi);
            or_poro_cal(pi, dt);
        } else if (pi->type == 4) {
            for (a = 0; a < 6; a++)
                for (b = 0; b < 6; b++) D[a][b] = 0.0;
            iso_tangent(D, c->mat.structure.e, c->mat.structure.poi);
            apply_tangent(pi, D);
            keep_depp0(c, i, D);
            jaumann(pi);
            for (k = 0; k < 6; k++) {
                pi->eps[k] = pi->eps[k] + pi->deps[k] * dt;
                pi->sig[k] = pi->sig[k] + pi->dsig[k] * dt;
            }
            or_deq(pi);
            or_shearstrain(pi);
            or_poro_cal(pi, dt);
        }
    }
    return bad;
}

/* free_surface_check (CPP_7-2-constitutive.cpp:453-506) */
static void or_free_surface(or_ctx *c) {
    int n = c->n, i;
    double tmax = c->pro.threshold_max, tmin = c->pro.threshold_min, incl = tmax - tmin;
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        double color = 0.0, xs;
        int nc = nbcount(&c->pc[i]), nj, k;
        for (nj = 1; nj <= nc; nj++) {
            int j = c->pc[i].influ[nj];
            const OPar *pj = &c->p[j];
            if (pi->type == pj->type && pi->type != 0 && pi->type != 7)
                color = color + pj->mass / pj->rho * 1.0 * c->pc[i].wij[nj][3];
            else if (pi->type != pj->type && pi->type > 0 && pi->type < 5 && pj->type == 0)
                color = color + pj->mass / pj->rho * 1.0 * c->pc[i].wij[nj][3];
        }
        if (pi->type == 1) {
            if (color < tmin / 2.0) {
                for (k = 0; k < 3; k++) pi->sig[k] = -1.0;
                for (k = 3; k < 6; k++) pi->sig[k] = 0.0;
            } else if (color >= tmin / 2.0 && color < tmax / 2.0) {
                xs = (2.0 * color - tmin) / incl;
                for (k = 0; k < 6; k++) pi->sig[k] = xs * pi->sig[k];
            }
        } else if (pi->type == 2) {
            if (color < tmin) {
                for (k = 0; k < 3; k++) pi->sig[k] = -1.0;
                for (k = 3; k < 6; k++) pi->sig[k] = 0.0;
            } else if (color >= tmin && color < tmax) {
                xs = (color - tmin) / incl * 1.0;
                for (k = 0; k < 6; k++) pi->sig[k] = xs * pi->sig[k];
            }
            or_stress_adjust(pi);
            or_deq(pi);
        }
    }
}

/* eejcb: cyclic Jacobi eigen-solver (CPP_7-1-strain.cpp:341-437) */
static int or_eejcb(double a[9], double v[9], double eps, int jt) {
    int i, j, p = 0, q = 0, u, w, t, s, l = 1;
    double fm, cn, sn, omega, x, y, d;
    for (i = 0; i < 3; i++)
        for (j = 0; j < 3; j++) v[i * 3 + j] = i == j ? 1.0 : 0.0;
    for (;;) {
        fm = 0.0;
        for (i = 0; i < 3; i++)
            for (j = 0; j < 3; j++) {
                d = fabs(a[i * 3 + j]);
                if (i != j && d > fm) { fm = d; p = i; q = j; }
            }
        if (fm < eps) return 1;
        if (l > jt) return -1;
        l = l + 1;
        u = p * 3 + q; w = p * 3 + p; t = q * 3 + p; s = q * 3 + q;
        x = -a[u];
        y = (a[s] - a[w]) / 2.0;
        omega = x / sqrt(x * x + y * y);
        if (y < 0.0) omega = -omega;
        sn = 1.0 + sqrt(1.0 - omega * omega);
        sn = omega / sqrt(2.0 * sn);
        cn = sqrt(1.0 - sn * sn);
        fm = a[w];
        a[w] = fm * cn * cn + a[s] * sn * sn + a[u] * omega;
        a[s] = fm * sn * sn + a[s] * cn * cn - a[u] * omega;
        a[u] = 0.0;
        a[t] = 0.0;
        for (j = 0; j < 3; j++)
            if (j != p && j != q) {
                u = p * 3 + j; w = q * 3 + j;
                fm = a[u];
                a[u] = fm * cn + a[w] * sn;
                a[w] = -fm * sn + a[w] * cn;
            }
        for (i = 0; i < 3; i++)
            if (i != p && i != q) {
                u = i * 3 + p; w = i * 3 + q;
                fm = a[u];
                a[u] = fm * cn + a[w] * sn;
                a[w] = -fm * sn + a[w] * cn;
            }
        for (i = 0; i < 3; i++) {
            u = i * 3 + p; w = i * 3 + q;
            fm = v[u];
            v[u] = fm * cn + v[w] * sn;
            v[w] = -fm * sn + v[w] * cn;
        }
    }
}

/* Stress_eigValue (CPP_7-1-strain.cpp:283-332) */
static void or_eig(or_ctx *c) {
    int n = c->n, i;
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        double a[9] = {pi->sig[0], pi->sig[3], pi->sig[5], pi->sig[3], pi->sig[1], pi->sig[4],
                       pi->sig[5], pi->sig[4], pi->sig[2]}, v[9];
        int k;
        or_eejcb(a, v, 0.001, 7);
        for (k = 0; k < 3; k++) {
            pi->strep[k] = a[k * 3 + k];
            pi->beta_ts[k][0] = v[k * 3]; pi->beta_ts[k][1] = v[k * 3 + 1]; pi->beta_ts[k][2] = v[k * 3 + 2];
            pi->beta[0][k] = v[k * 3]; pi->beta[1][k] = v[k * 3 + 1]; pi->beta[2][k] = v[k * 3 + 2];
        }
    }
}

/* ------------------------------------------------------------------ accelerations */
/* the nine stress-divergence sums of CPP_9-4-derivation_stress.cpp (e.g. :47-64): component
 * (a,b) pairs stress index SIDX[a][b] with kernel-gradient component b. sgn=+1 adds the i term,
 * sgn=-1 subtracts it (the *_minor* variants). */
static const int SIDX[3][3] = {{0, 3, 5}, {3, 1, 4}, {5, 4, 2}};
static void sdiv(double t[3][3], double mj, const double *aj, double rj, const double *ai, double ri,
                 int sgn, const double *w) {
    int a, b;
    for (a = 0; a < 3; a++)
        for (b = 0; b < 3; b++) {
            int s = SIDX[a][b];
            if (sgn > 0) t[a][b] = t[a][b] + mj * (aj[s] / rj / rj + ai[s] / ri / ri) * w[b];
            else t[a][b] = t[a][b] + mj * (aj[s] / rj / rj - ai[s] / ri / ri) * w[b];
        }
}
static void add_sum(double ax[3], double t[3][3]) {
    ax[0] += t[0][0] + t[0][1] + t[0][2];
    ax[1] += t[1][0] + t[1][1] + t[1][2];
    ax[2] += t[2][0] + t[2][1] + t[2][2];
}

/* fluid_acceleration_water/air and the minor* variants (CPP_9-4-derivation_stress.cpp:12-351) */
static void or_fluid_accel(or_ctx *c, int ftype, int other, int sgn) {
    int n = c->n, i;
    double g[3] = {c->pro.gx, c->pro.gy, c->pro.gz};
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        double t[3][3] = {{0}}, por, ai[6], aj[6];
        int nc = nbcount(&c->pc[i]), nj, k;
        if (pi->type != ftype) continue;
        por = c->pf[i].porosity;
        for (k = 0; k < 3; k++) pi->ax[k] = g[k];
        for (nj = 1; nj <= nc; nj++) {
            int j = c->pc[i].influ[nj];
            const OPar *pj = &c->p[j];
            if (pi->type == pj->type || pj->type == other) {
                for (k = 0; k < 6; k++) {
                    ai[k] = por * pi->sig[k];
                    aj[k] = por * pj->sig[k];
                }
                sdiv(t, pj->mass, aj, pj->rho, ai, pi->rho, sgn, c->pc[i].wij[nj]);
            }
        }
        add_sum(pi->ax, t);
    }
}

/* soil_acceleration_noreg (CPP_9-4-derivation_stress.cpp:354-428) and
 * structure_acceleration_noreg (:431-505) */
static void or_solid_accel(or_ctx *c, int stype) {
    int n = c->n, i;
    double g[3] = {c->pro.gx, c->pro.gy, c->pro.gz};
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        double t[3][3] = {{0}}, por = c->ps[i].porosity;
        int nc = nbcount(&c->pc[i]), nj, k;
        if (pi->type != stype) continue;
        for (k = 0; k < 3; k++) pi->ax[k] = g[k];
        for (nj = 1; nj <= nc; nj++) {
            int j = c->pc[i].influ[nj];
            const OPar *pj = &c->p[j];
            const double *w = c->pc[i].wij[nj];
            double satu, prei, prej;
            if (pi->type != pj->type) continue;
            satu = pj->satu; /* quirk: the neighbour's saturation is used for both terms */
            sdiv(t, pj->mass, pj->sig, pj->rho, pi->sig, pi->rho, 1, w);
            if (stype == 2) {
                double di = c->ps[i].dens; /* quirk A12: parameter density, not rho_i */
                prei = (1 - por) * (-satu * pi->prew + (1 - satu) * pi->prea);
                prej = (1 - por) * (-satu * pj->prew + (1 - satu) * pj->prea);
                for (k = 0; k < 3; k++)
                    t[k][k] = t[k][k] + pj->mass * (prej / pj->rho / pj->rho + prei / di / di) * w[k];
            } else {
                prei = -satu * pi->prew + (1 - satu) * pi->prea;
                prej = -satu * pj->prew + (1 - satu) * pj->prea;
                for (k = 0; k < 3; k++)
                    t[k][k] = t[k][k] + pj->mass * (prej / pj->rho / pj->rho - prei / pi->rho / pi->rho) * w[k];
            }
        }
        add_sum(pi->ax, t);
    }
}

/* artificial_viscosity (CPP_9-3-artificialvis.cpp:12-99) */
static void or_art_visc(or_ctx *c) {
    int n = c->n, ndim = c->pro.ndim, i;
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        double tu[3] = {0, 0, 0};
        int nc = nbcount(&c->pc[i]), nj, k;
        for (nj = 1; nj <= nc; nj++) {
            int j = c->pc[i].influ[nj];
            const OPar *pj = &c->p[j];
            double a = 0, b = 0, cs = 0, inprod = 0, dst = 0, myu, paix, hl, rou, dvx[3], dx[3];
            if (!(pi->type == pj->type && pi->type != 7)) continue;
            for (k = 0; k < ndim; k++) {
                dvx[k] = pi->vxp[k] - pj->vxp[k];
                dx[k] = pi->xp[k] - pj->xp[k];
                dst = dst + dx[k] * dx[k];
                inprod = inprod + dvx[k] * dx[k];
            }
            hl = (pj->hl + pi->hl) * 0.5;
            rou = (pj->rho + pi->rho) * 0.5;
            if (pi->type == 1) { cs = c->pf[i].vpu; a = c->pro.art_vis[0][0]; b = c->pro.art_vis[0][1]; }
            if (pi->type == 2) { cs = c->ps[i].vpu; a = c->pro.art_vis[1][0]; b = c->pro.art_vis[1][1]; }
            if (pi->type == 3) { cs = c->pf[i].vpu; a = c->pro.art_vis[2][0]; b = c->pro.art_vis[2][1]; }
            if (inprod < 0.0) {
                dst = sqrt(dst);
                myu = hl * inprod / (dst * dst + 0.01 * hl * hl);
                paix = (-a * cs * myu + b * myu * myu) / rou;
            } else
                paix = 0.0;
            for (k = 0; k < ndim; k++) tu[k] = tu[k] + pj->mass * paix * c->pc[i].wij[nj][k];
        }
        if (pi->type != 7)
            for (k = 0; k < ndim; k++) pi->ax[k] = pi->ax[k] - tu[k];
    }
}

/* boundary_moving_stress_effect, flbndy 1/2 (CPP_9-4-derivation_stress.cpp:631-721) */
/* boundary_moving_stress_effect (CPP_9-4-derivation_stress.cpp:631-721, flbndy 1/2): the wall
 * mirrors the particle's own stress; boundary_effect_free_struct (:900-990, flbndy 5) the same with
 * the shear components negated for fluids */
static void or_bndy_stress(or_ctx *c) {
    int n = c->n, i, free_struct = c->pro.flbndy == 5;
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        double t[3][3] = {{0}}, as[6], aj[6];
        int nc = nbcount(&c->pc[i]), nj, k;
        if (pi->type == 0 || pi->type == 7) continue;
        for (nj = 1; nj <= nc; nj++) {
            int j = c->pc[i].influ[nj];
            const OPar *pj = &c->p[j];
            if (pj->type != 0) continue;
            for (k = 0; k < 6; k++) as[k] = 0.0;
            if (pi->type == 1 || pi->type == 3) {
                for (k = 0; k < 6; k++) as[k] = pi->sig[k];
            } else if (pi->type == 2 || pi->type == 4) {
                double satu = pi->satu;
                for (k = 0; k < 3; k++) as[k] = pi->sig[k] - satu * pi->prew + (1 - satu) * pi->prea;
                for (k = 3; k < 6; k++) as[k] = pi->sig[k];
            }
            for (k = 0; k < 6; k++) aj[k] = as[k];
            if (free_struct && (pi->type == 1 || pi->type == 3))
                for (k = 3; k < 6; k++) aj[k] = -as[k];
            sdiv(t, pj->mass, aj, pj->rho, as, pi->rho, 1, c->pc[i].wij[nj]);
        }
        add_sum(pi->ax, t);
    }
}

/* boundary_stress_effect_tran (CPP_9-4-derivation_stress.cpp:725-820), flbndy 3/4/13/14: the
 * boundary neighbour contributes the stress its no-slip / free-slip pass interpolated for the
 * particle's phase */
static void or_bndy_stress_tran(or_ctx *c) {
    int n = c->n, i;
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        double t[3][3] = {{0}};
        int nc = nbcount(&c->pc[i]), nj, k;
        if (pi->type == 0 || pi->type == 7) continue;
        for (nj = 1; nj <= nc; nj++) {
            int j = c->pc[i].influ[nj];
            const OPar *pj = &c->p[j];
            double ai[6] = {0, 0, 0, 0, 0, 0}, aj[6] = {0, 0, 0, 0, 0, 0};
            if (pj->type != 0) continue;
            if (pi->type == 1 || pi->type == 3 || pi->type == 4) {
                for (k = 0; k < 6; k++) ai[k] = pi->sig[k];
            } else if (pi->type == 2) {
                double satu = pi->satu;
                for (k = 0; k < 3; k++) ai[k] = pi->sig[k] - satu * pi->prew + (1 - satu) * pi->prea;
                for (k = 3; k < 6; k++) ai[k] = pi->sig[k];
            }
            if (pi->type >= 1 && pi->type <= 4)
                for (k = 0; k < 6; k++) aj[k] = c->vb[j].sig_ph[pi->type - 1][k];
            sdiv(t, pj->mass, aj, pj->rho, ai, pi->rho, 1, c->pc[i].wij[nj]);
        }
        add_sum(pi->ax, t);
    }
}

/* coupled_Water_Soil_Structure (CPP_9-4-derivation_stress.cpp:507-629), Para_SSInter acc_type 2:
 * stress-gradient sums across the structure / soil and structure / water interfaces */
static void or_coupled(or_ctx *c) {
    int n = c->n, i;
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        double t[3][3] = {{0}};
        int nc = nbcount(&c->pc[i]), nj;
        if (!(pi->type == 4 || pi->type == 2 || pi->type == 1)) continue;
        for (nj = 1; nj <= nc; nj++) {
            int j = c->pc[i].influ[nj];
            const OPar *pj = &c->p[j];
            if ((pi->type == 4 && pj->type == 2) || (pi->type == 2 && pj->type == 4) || (pi->type == 4 && pj->type == 1))
                sdiv(t, pj->mass, pj->sig, pj->rho, pi->sig, pi->rho, 1, c->pc[i].wij[nj]);
            if (pi->type == 1 && pj->type == 4) sdiv(t, pj->mass, pj->sig, pj->rho, pi->sig, pi->rho, -1, c->pc[i].wij[nj]);
        }
        add_sum(pi->ax, t);
    }
}

/* interact1: Darcy drag (CPP_10-interact1.cpp:15-110) */
static void or_interact1(or_ctx *c) {
    int n = c->n, ndim = c->pro.ndim, i;
    double xishu = c->pro.interc, dt = c->pro.dt;
    double gacc = sqrt(c->pro.gx * c->pro.gx + c->pro.gy * c->pro.gy + c->pro.gz * c->pro.gz);
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        double tu[3] = {0, 0, 0}, dx[3], dvx[3], inprod, vx1, vx2;
        int nc = nbcount(&c->pc[i]), nj, k;
        for (nj = 1; nj <= nc; nj++) {
            int j = c->pc[i].influ[nj];
            const OPar *pj = &c->p[j];
            double w3 = c->pc[i].wij[nj][3];
            if (pi->type == pj->type) continue;
            inprod = 0.0;
            for (k = 0; k < ndim; k++) {
                dx[k] = pi->xp[k] - pj->xp[k];
                dvx[k] = pi->vxp[k] - pj->vxp[k];
                inprod = inprod + dx[k] * dvx[k];
            }
            if (!(inprod < 0)) continue;
            if (pi->type == 2 && pj->type == 1) {
                for (k = 0; k < ndim; k++)
                    tu[k] = tu[k] + pi->satu * pi->satu * pi->porosity * pi->porosity * c->pf[j].dens * gacc *
                                        dvx[k] / pi->cop * pj->mass / pj->rho * w3;
            } else if (pi->type == 1 && pj->type == 2) {
                for (k = 0; k < ndim; k++)
                    tu[k] = tu[k] + pj->satu * pj->satu * pj->porosity * pj->porosity * c->pf[i].dens * gacc *
                                        dvx[k] / pj->cop * pj->mass / pj->rho * w3;
            } else if (pi->type == 2 && pj->type == 3) {
                for (k = 0; k < ndim; k++)
                    tu[k] = tu[k] + (1 - pi->satu) * (1 - pi->satu) * pi->porosity * pi->porosity * c->pf[j].dens *
                                        gacc * dvx[k] * 2000 * pj->mass / pj->rho * w3;
            } else if (pi->type == 3 && pj->type == 2) {
                for (k = 0; k < ndim; k++)
                    tu[k] = tu[k] + (1 - pj->satu) * (1 - pj->satu) * pj->porosity * pj->porosity * c->pf[i].dens *
                                        gacc * dvx[k] * 2000 * pj->mass / pj->rho * w3;
            }
        }
        if (pi->type != 0 && pi->type != 7) {
            for (k = 0; k < ndim; k++) {
                pi->interf[k] = tu[k];
                vx1 = pi->vxp[k] + 0.5 * (pi->axp[k] + pi->ax[k]) * dt;
                vx2 = pi->vxp[k] + 0.5 * (pi->axp[k] + pi->ax[k] - pi->interf[k] * xishu / pi->rho) * dt;
                if (vx1 * vx2 <= 0.0) {
                    pi->interf[k] = 0.0;
                    pi->ax[k] = -2.0 * pi->vxp[k] / dt - pi->axp[k];
                } else
                    pi->ax[k] -= pi->interf[k] * xishu / pi->rho;
            }
        }
    }
}

/* drag coefficient beta of a soil / water pair: interact2 (fli 2, Ergun below porosity 0.8,
 * Wen-Yu above, CPP_10-interact1.cpp:162-182) or interact3 (fli 3, Kozeny-Carman permeability
 * cd, turbulent term above Reynolds number 60, CPP_10-interact1.cpp:317-333). por and ds are the
 * soil particle's, vis and dens the water's parameter, rhow the water particle's density. */
static double drag_beta(int fli, double mag_v, double por, double ds, double vis, double dens, double rhow) {
    double re, cd, beta;
    if (fli == 2) {
        re = mag_v * dens * por * ds / vis;
        if (re > 1000.0) cd = 0.43;
        else if (fabs(re) > 0.0001) cd = 24.0 * (1.0 + 0.15 * pow(re, 0.687)) / re;
        else cd = 0.0;
        if (por <= 0.8) {
            beta = 150.0 * (1.0 - por) * (1.0 - por) / por / por / ds / ds * vis;
            beta = beta + 1.75 * (1.0 - por) / ds / por * dens * mag_v;
        } else {
            beta = 0.75 * cd * mag_v * dens * (1.0 - por) / ds * pow(por, -2.7);
        }
        return beta;
    }
    re = mag_v * dens * ds / vis;
    cd = por * por * por * ds * ds * 0.0066666667 / (1.0 - por) / (1.0 - por);
    if (re > 60.0) return vis / cd + 1.75 / sqrt(150.0) * rhow * mag_v / sqrt(cd) / pow(por, 1.5);
    return vis / cd;
}

/* interact2 / interact3: drag with laminar and turbulent parts between soil and water, the
 * interact1 air terms, and the interact1 velocity-reversal clamp (CPP_10-interact1.cpp:115-267,
 * :270-420) */
static void or_interact23(or_ctx *c) {
    int n = c->n, ndim = c->pro.ndim, fli = c->pro.fli, i;
    double xishu = c->pro.interc, dt = c->pro.dt;
    double gacc = sqrt(c->pro.gx * c->pro.gx + c->pro.gy * c->pro.gy + c->pro.gz * c->pro.gz);
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        double tu[3] = {0, 0, 0}, dx[3], dvx[3], inprod, mag_v, beta, vx1, vx2;
        int nc = nbcount(&c->pc[i]), nj, k;
        for (nj = 1; nj <= nc; nj++) {
            int j = c->pc[i].influ[nj];
            const OPar *pj = &c->p[j];
            double w3 = c->pc[i].wij[nj][3];
            if (pi->type == pj->type) continue;
            inprod = 0.0;
            mag_v = 0.0;
            for (k = 0; k < ndim; k++) {
                dx[k] = pi->xp[k] - pj->xp[k];
                dvx[k] = pi->vxp[k] - pj->vxp[k];
                mag_v = mag_v + dvx[k] * dvx[k];
                inprod = inprod + dx[k] * dvx[k];
            }
            mag_v = sqrt(mag_v);
            if (!(inprod < 0)) continue;
            if (pi->type == 2 && pj->type == 1) {
                beta = drag_beta(fli, mag_v, pi->porosity, c->ps[i].ds, c->pf[j].vis, c->pf[j].dens, pj->rho);
                for (k = 0; k < ndim; k++) tu[k] = tu[k] + pj->mass / pj->rho * w3 * beta * (pi->vxp[k] - pj->vxp[k]);
            } else if (pi->type == 1 && pj->type == 2) {
                beta = drag_beta(fli, mag_v, pj->porosity, c->ps[j].ds, c->pf[i].vis, c->pf[i].dens, pi->rho);
                for (k = 0; k < ndim; k++) tu[k] = tu[k] + pj->mass / pj->rho * w3 * beta * (pi->vxp[k] - pj->vxp[k]);
            } else if (pi->type == 2 && pj->type == 3) {
                for (k = 0; k < ndim; k++)
                    tu[k] = tu[k] + (1 - pi->satu) * (1 - pi->satu) * pi->porosity * pi->porosity * c->pf[j].dens *
                                        gacc * dvx[k] * 2000 * pj->mass / pj->rho * w3;
            } else if (pi->type == 3 && pj->type == 2) {
                for (k = 0; k < ndim; k++)
                    tu[k] = tu[k] + (1 - pj->satu) * (1 - pj->satu) * pj->porosity * pj->porosity * c->pf[i].dens *
                                        gacc * dvx[k] * 2000 * pj->mass / pj->rho * w3;
            }
        }
        if (pi->type != 0 && pi->type != 7) {
            for (k = 0; k < ndim; k++) {
                pi->interf[k] = tu[k];
                vx1 = pi->vxp[k] + 0.5 * (pi->axp[k] + pi->ax[k]) * dt;
                vx2 = pi->vxp[k] + 0.5 * (pi->axp[k] + pi->ax[k] - pi->interf[k] * xishu / pi->rho) * dt;
                if (vx1 * vx2 <= 0.0) {
                    pi->interf[k] = 0.0;
                    pi->ax[k] = -2.0 * pi->vxp[k] / dt - pi->axp[k];
                } else
                    pi->ax[k] -= pi->interf[k] * xishu / pi->rho;
            }
        }
    }
}


/* ------------------------------------------------------------------ regularisation, saturation */
/* strain_regu / stress_regu: Shepard smoothing of eps or sig over same-type neighbours
 * (CPP_7-2-constitutive.cpp:385-450). In place in the reference (quirk A2, Gauss-Seidel in a
 * serial run); OR_DENSITY_JACOBI smooths from the values before the pass. */
static void or_regu(or_ctx *c, int which_sig) {
    int n = c->n, i;
    double (*prev)[6] = NULL;
    if (c->density_mode == OR_DENSITY_JACOBI) {
        prev = (double(*)[6])malloc(sizeof(double[6]) * (size_t)n);
        for (i = 0; i < n; i++) memcpy(prev[i], which_sig ? c->p[i].sig : c->p[i].eps, 48);
    }
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        double up[6] = {0, 0, 0, 0, 0, 0}, dn = 0.0;
        int nc = nbcount(&c->pc[i]), nj, k;
        for (nj = 1; nj <= nc; nj++) {
            int j = c->pc[i].influ[nj];
            const OPar *pj = &c->p[j];
            const double *v = prev ? prev[j] : (which_sig ? pj->sig : pj->eps);
            if (pi->type != pj->type) continue;
            for (k = 0; k < 6; k++) up[k] = up[k] + pj->mass / pj->rho * c->pc[i].wij[nj][3] * v[k];
            dn = dn + pj->mass / pj->rho * c->pc[i].wij[nj][3];
        }
        if (fabs(dn) > 1.0e-6) {
            double *out = which_sig ? pi->sig : pi->eps;
            for (k = 0; k < 6; k++) out[k] = up[k] / dn;
        }
    }
    free(prev);
}

/* moisturesuct (CPP_7-2-constitutive.cpp:324-338); pi = 3.141592653 (Header_Parameters.h:10) */
static double or_moisturesuct(double satu, double dsatu, const sphg_soil *s) {
    const double pi = 3.141592653;
    double temp;
    if ((satu - s->sitar) < 0.0001) temp = tan(pi * (s->sitas - satu - 0.0001) / (s->sitas - s->sitar) / 2.0);
    temp = tan(pi * (s->sitas - satu) / (s->sitas - s->sitar) / 2.0);
    if (dsatu > 0.0) return log(1 + exp(s->c2 * s->sw) * temp) / s->c2;
    return log(1 + exp(s->c1 * s->sd) * temp) / s->c1;
}

/* saturation (CPP_7-1-strain.cpp:222-280) */
static void or_saturation(or_ctx *c) {
    int n = c->n, i;
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        const sphg_soil *s = &c->ps[i];
        double cof = s->sitas - s->sitar, sitas = s->sitas, sitar = s->sitar, up = 0.0, dn = 0.0, ps;
        int nc = nbcount(&c->pc[i]), nj;
        for (nj = 1; nj <= nc; nj++) {
            int j = c->pc[i].influ[nj];
            const OPar *pj = &c->p[j];
            if (pi->type != pj->type && pi->type == 2 && pj->type == 1) {
                up = up + pj->mass / pj->rho * cof * c->pc[i].wij[nj][3];
                dn = dn + pj->mass / pj->rho * c->pc[i].wij[nj][3];
            }
        }
        if (pi->type != 2) continue;
        ps = pi->satu;
        pi->satu = dn != 0.0 ? up / dn + sitar : up + sitar;
        pi->dsatu = pi->satu - ps;
        if (pi->dsatu < 0) {
            pi->satu = pi->satu - pi->dsatu;
            pi->dsatu = 0.0;
        }
        if (pi->satu > sitas) pi->satu = sitas;
        else if (pi->satu < sitar) pi->satu = sitar;
        pi->suct = or_moisturesuct(pi->satu, pi->dsatu, s);
    }
}

/* ------------------------------------------------------------------ artificial stress */
/* artificial_stress_2d / _3d (CPP_9-5-artificialstress.cpp:28-317); pi*0.78539816325 (quirk A15) */
static void or_inverse_mat(double A[3][3], double B[3][3]) {
    double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) - A[1][0] * (A[0][1] * A[2][2] - A[0][2] * A[2][1]) +
                 A[2][0] * (A[0][1] * A[1][2] - A[0][2] * A[1][1]);
    if (fabs(det) > 1.0e-5) {
        B[0][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) / det;
        B[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / det;
        B[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / det;
        B[1][0] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) / det;
        B[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / det;
        B[1][2] = (A[1][0] * A[0][2] - A[0][0] * A[1][2]) / det;
        B[2][0] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) / det;
        B[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) / det;
        B[2][2] = (A[0][0] * A[1][1] - A[1][0] * A[0][1]) / det;
    }
}
static void mat3mul(double a[3][3], double b[3][3], double c[3][3]) {
    int r, q;
    for (r = 0; r < 3; r++)
        for (q = 0; q < 3; q++) c[r][q] = a[r][0] * b[0][q] + a[r][1] * b[1][q] + a[r][2] * b[2][q];
}

static void or_art_stress(or_ctx *c) {
    const double pi = 3.141592653;
    int n = c->n, ndim = c->pro.ndim, i;
    double xs = c->pro.art_stre_coe;
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
    for (i = 0; i < n; i++) {
        OPar *p = &c->p[i];
        double R[3][3] = {{0}};
        if (p->type != 2) continue;
        if (ndim == 2) {
            double sita, sxx, syy, rxx, ryy;
            if (fabs(p->sig[0] - p->sig[1]) < 1.0e-6) {
                if (p->sig[3] > 0.0) sita = -pi * 0.78539816325;
                else if (p->sig[3] < 0.0) sita = pi * 0.78539816325;
                else sita = 0.0;
            } else
                sita = 0.5 * atan(2.0 * p->sig[3] / (p->sig[0] - p->sig[1]));
            sxx = cos(sita) * cos(sita) * p->sig[0] + 2.0 * sin(sita) * cos(sita) * p->sig[3] + sin(sita) * sin(sita) * p->sig[1];
            syy = sin(sita) * sin(sita) * p->sig[0] - 2.0 * sin(sita) * cos(sita) * p->sig[3] + cos(sita) * cos(sita) * p->sig[1];
            rxx = sxx > 0.0 ? -xs * sxx / p->rho / p->rho : 0.0;
            ryy = syy > 0.0 ? -xs * syy / p->rho / p->rho : 0.0;
            R[0][0] = rxx * cos(sita) * cos(sita) + ryy * sin(sita) * sin(sita);
            R[1][1] = rxx * sin(sita) * sin(sita) + ryy * cos(sita) * cos(sita);
            R[0][1] = (rxx - ryy) * sin(sita) * cos(sita);
            R[1][0] = R[0][1];
        } else {
            double rr[3], bi[3][3] = {{0}}, bti[3][3] = {{0}}, st[3][3] = {{0}}, tmp[3][3];
            int k;
            for (k = 0; k < 3; k++) rr[k] = p->strep[k] > 0.0 ? -xs * p->strep[k] / p->rho / p->rho : 0.0;
            or_inverse_mat(p->beta, bi);
            or_inverse_mat(p->beta_ts, bti);
            st[0][0] = rr[0]; st[1][1] = rr[1]; st[2][2] = rr[2];
            mat3mul(bi, st, tmp);
            mat3mul(tmp, bti, R);
        }
        memcpy(p->Rij, R, sizeof R);
    }
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        double tu[3] = {0, 0, 0};
        int nc = nbcount(&c->pc[i]), nj, k;
        if (pi->type != 2) continue;
        for (nj = 1; nj <= nc; nj++) {
            int j = c->pc[i].influ[nj];
            const OPar *pj = &c->p[j];
            const double *w = c->pc[i].wij[nj];
            double dst = 0.0, hl, q, f;
            if (pi->type != pj->type) continue;
            for (k = 0; k < 3; k++) dst = dst + (pi->xp[k] - pj->xp[k]) * (pi->xp[k] - pj->xp[k]);
            dst = sqrt(dst);
            hl = ndim == 2 ? (pi->hl + pj->hl) * 0.5 : (pi->hl + pj->hl) / 2;
            q = dst / hl;
            if (q >= 0.0 && q < 1.0) f = pow((0.666666666666667 - q * q + 0.5 * q * q * q) * 3.82300884955752, 2.55);
            else if (q >= 1.0 && q < 2.0) f = pow((2.0 - q) * (2.0 - q) * (2.0 - q) * 0.629737609329447, 2.55);
            else f = 0.0;
            if (ndim == 2) {
                tu[0] = tu[0] + pj->mass * f * (pi->Rij[0][0] + pj->Rij[0][0]) * w[0] + pj->mass * f * (pi->Rij[0][1] + pj->Rij[0][1]) * w[1];
                tu[1] = tu[1] + pj->mass * f * (pi->Rij[1][0] + pj->Rij[1][0]) * w[0] + pj->mass * f * (pi->Rij[1][1] + pj->Rij[1][1]) * w[1];
            } else {
                for (k = 0; k < 3; k++)
                    tu[k] = tu[k] + pj->mass * f * (pi->Rij[k][0] + pj->Rij[k][0]) * w[0] +
                            pj->mass * f * (pi->Rij[k][1] + pj->Rij[k][1]) * w[1] +
                            pj->mass * f * (pi->Rij[k][2] + pj->Rij[k][2]) * w[2];
            }
        }
        for (k = 0; k < ndim; k++) pi->ax[k] = pi->ax[k] + tu[k];
    }
}

/* soil_damping_geostress (CPP_9-4-derivation_stress.cpp:850-897) */
static void or_damping_geostress(or_ctx *c) {
    int n = c->n, ndim = c->pro.ndim, i;
    double dt = c->pro.dt;
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        double damp[3] = {0, 0, 0}, cd = pi->cd, vx1, vx2;
        int nc = nbcount(&c->pc[i]), nj, k;
        if (!(pi->type == 2 || pi->type == 4)) continue;
        for (nj = 1; nj <= nc; nj++) {
            int j = c->pc[i].influ[nj];
            const OPar *pj = &c->p[j];
            if (pi->type != pj->type) continue;
            for (k = 0; k < ndim; k++)
                damp[k] = damp[k] + pj->mass / pj->rho * cd * (pi->vxp[k] - pj->vxp[k]) * c->pc[i].wij[nj][3];
        }
        for (k = 0; k < ndim; k++) {
            vx1 = pi->vxp[k] + 0.5 * (pi->axp[k] + pi->ax[k]) * dt;
            vx2 = pi->vxp[k] + 0.5 * (pi->axp[k] + pi->ax[k] - damp[k]) * dt;
            if (vx1 * vx2 <= 0.0) pi->ax[k] = -2.0 * pi->vxp[k] / dt - pi->axp[k];
            else pi->ax[k] -= damp[k];
        }
    }
}

/* soil_damping_vibration (CPP_9-1-acceleration_vibration.cpp:167-224), flv > 0: a velocity damping
 * of soil and structure particles against their own phase, its coefficient from the largest
 * diagonal entry of the particle's last tangent */
static void or_damping_vibration(or_ctx *c) {
    int n = c->n, ndim = c->pro.ndim, i;
    double dt = c->pro.dt;
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        double damp[3] = {0, 0, 0}, cd, stiff, tempdown = 0.0, vx1, vx2;
        int nc = nbcount(&c->pc[i]), nj, k;
        if (!(pi->type == 2 || pi->type == 4)) continue;
        stiff = fmax(0.0, c->depp0[i][0]);
        stiff = fmax(stiff, c->depp0[i][1]);
        stiff = fmax(stiff, c->depp0[i][2]);
        cd = c->ps[i].damp1 + c->ps[i].damp2 * sqrt(stiff / pi->rhop / pi->hl / pi->hl);
        for (nj = 1; nj <= nc; nj++) {
            int j = c->pc[i].influ[nj];
            const OPar *pj = &c->p[j];
            if (pi->type != pj->type) continue;
            for (k = 0; k < ndim; k++)
                damp[k] = damp[k] + pj->mass / pj->rho * cd * (pi->vxp[k] - pj->vxp[k]) * c->pc[i].wij[nj][3];
            tempdown = tempdown + pj->mass / pj->rho * c->pc[i].wij[nj][3];
        }
        for (k = 0; k < ndim; k++) {
            if (fabs(tempdown) < 1.0e-7) tempdown = 1.0e-7;
            vx1 = pi->vxp[k] + 0.5 * (pi->axp[k] + pi->ax[k]) * dt;
            vx2 = pi->vxp[k] + 0.5 * (pi->axp[k] + pi->ax[k] - damp[k] / tempdown) * dt;
            if (vx1 * vx2 <= 0.0) pi->ax[k] = -2.0 * pi->vxp[k] / dt - pi->axp[k];
            else pi->ax[k] -= damp[k] / tempdown;
        }
    }
}

/* ff_stress_2d_solid / ff_stress_3d_solid (CPP_19-FreeFieldParticles.cpp:24-76): viscous stresses
 * of a free-field boundary from the velocity jump, P waves along the boundary normal `type`
 * (1 x, 2 y, 3 z) and S waves along the others */
static void ff_stress(int ndim, double sig[3], const double va[3], const double vb[3], double e, double poi,
                      double dens, int type) {
    double g = e * 0.5 / (1 + poi), k = e * 0.333333333333333 / (1 - 2.0 * poi);
    double c_p = sqrt((k + 4.0 * g * 0.333333333333333) / dens), c_s = sqrt(g / dens);
    if (ndim == 2) {
        if (type == 1) {
            sig[0] = -c_p * dens * (va[0] - vb[0]);
            sig[1] = -c_s * dens * (va[1] - vb[1]);
        } else if (type == 2) {
            sig[0] = -c_s * dens * (va[0] - vb[0]);
            sig[1] = -c_p * dens * (va[1] - vb[1]);
        }
        return;
    }
    if (type == 1) {
        sig[0] = -c_p * dens * (va[0] - vb[0]);
        sig[1] = -c_s * dens * (va[1] - vb[1]);
        sig[2] = -c_s * dens * (va[2] - vb[2]);
    } else if (type == 2) {
        sig[0] = -c_s * dens * (va[0] - vb[0]);
        sig[1] = -c_p * dens * (va[1] - vb[1]);
        sig[2] = -c_s * dens * (va[2] - vb[2]);
    } else if (type == 3) {
        sig[0] = -c_s * dens * (va[0] - vb[0]);
        sig[1] = -c_s * dens * (va[1] - vb[1]);
        sig[2] = -c_p * dens * (va[2] - vb[2]);
    }
}

/* clFreeFieldBndy_Fun::free_field_solid (CPP_19-FreeFieldParticles.cpp:79-141), flv > 0: soil
 * particles against matype-5 boundary particles moving at their interfss, with the elastic
 * constants of a default-constructed Para_Soil (e = poi = 0, SPHLoops.cpp:253) */
static void or_free_field_bndy(or_ctx *c) {
    int n = c->n, ndim = c->pro.ndim, i;
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        double t1[3] = {0, 0, 0};
        int nc = nbcount(&c->pc[i]), nj, k;
        if (pi->type != 2) continue;
        for (nj = 1; nj <= nc; nj++) {
            int j = c->pc[i].influ[nj];
            const OPar *pj = &c->p[j];
            double sig[3] = {0.0, 0.0, 0.0};
            if (!(pj->type == 0 && pj->matype == 5)) continue;
            ff_stress(ndim, sig, pi->vx, pj->interfss, 0.0, 0.0, pj->rho, pj->permtype);
            for (k = 0; k < 3; k++) t1[k] = t1[k] + pj->mass / pj->rho * sig[k] / pi->rho * c->pc[i].wij[nj][k];
        }
        for (k = 0; k < ndim; k++) pi->ax[k] = pi->ax[k] + t1[k];
    }
}

/* clFreeFieldBndy_Fun_soil::free_field_solid (CPP_19-FreeFieldParticles_soil.cpp:79-169), flv > 0:
 * soil of matype < 50 against free-field soil columns (matype >= 50) and back, with the column
 * particle's elastic constants, density and normal */
static void or_free_field_soil(or_ctx *c) {
    int n = c->n, ndim = c->pro.ndim, i;
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        double t1[3] = {0, 0, 0};
        int nc = nbcount(&c->pc[i]), nj, k, col;
        if (pi->type != 2) continue;
        col = pi->matype >= 50;
        for (nj = 1; nj <= nc; nj++) {
            int j = c->pc[i].influ[nj];
            const OPar *pj = &c->p[j];
            double sig[3] = {0.0, 0.0, 0.0};
            if (!(pj->type == 2 && (col ? pj->matype < 50 : pj->matype >= 50))) continue;
            if (!col) ff_stress(ndim, sig, pi->vx, pj->vx, c->ps[j].e, c->ps[j].poi, pj->rho, pj->permtype);
            else ff_stress(ndim, sig, pj->vx, pi->vx, c->ps[i].e, c->ps[i].poi, pi->rho, pi->permtype);
            for (k = 0; k < 3; k++) t1[k] = t1[k] + pj->mass / pj->rho * sig[k] / pi->rho * c->pc[i].wij[nj][k];
        }
        for (k = 0; k < ndim; k++) pi->ax[k] = pi->ax[k] + t1[k];
    }
}

/* vibration_acceleration (flv 1, CPP_9-1-acceleration_vibration.cpp:11-87) and vibration_velocity
 * (flv 2, :90-164): the load of the period step containing t - sttime, while sttime <= t <=
 * edtime, becomes the acceleration (interf; walls of matype 3, 5) or velocity (interfss; matype 4,
 * 5) of the driving walls; a load of magnitude 1000 or more, or a time outside the window, zeroes
 * both instead */
static void or_vibration(or_ctx *c, int velocity) {
    int n = c->n, ndim = c->pro.ndim, i, k, nstep, live = 0;
    double time1 = c->t - c->vib_sttime, time2 = c->t - c->vib_edtime, v[3] = {0, 0, 0}, mx;
    if (time1 >= 0.0 && time2 <= 0.0) {
        nstep = (int)(time1 / c->vib_ttime * (double)c->vib_nsteps);
        if (nstep > c->vib_nsteps - 1) nstep = c->vib_nsteps - 1;
        for (k = 0; k < 3; k++) v[k] = c->vib_loads[nstep][k];
        mx = fmax(0.0, fabs(v[0]));
        mx = fmax(mx, fabs(v[1]));
        mx = fmax(mx, fabs(v[2]));
        live = mx < 1000.0;
    }
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        if (pi->type != 0 || !(pi->matype == (velocity ? 4 : 3) || pi->matype == 5)) continue;
        for (k = 0; k < ndim; k++) {
            if (live && velocity) pi->interfss[k] = v[k];
            else if (live) pi->interf[k] = v[k];
            else pi->interf[k] = pi->interfss[k] = 0.0;
        }
    }
}

/* ------------------------------------------------------------------ dynamic particle sets */
/* rainpro (CPP_15-1-rain.cpp:48-98): the next reserved particles from the end of the array become
 * water, one row at a time (while the next particle's height is within 0.1 dr of this one's
 * before the move); the reference reads one particle past the array at the end of the loop, here
 * the bound is tested first */
static void or_rainpro(or_ctx *c) {
    int k = c->n - c->nrain, ndim = c->pro.ndim;
    double g = ndim == 2 ? c->pro.gy : c->pro.gz;
    if (c->nrain <= 0) return;
    do {
        OPar *q = &c->p[k];
        double temp_xp = q->xp[ndim - 1];
        c->nrain = c->nrain - 1;
        q->xp[ndim - 1] = 0.95 * q->xp[ndim - 1];
        q->type = 1;
        q->matype = 0;
        q->etype = 7;
        q->rhop = c->pf[k].porosity * c->pf[k].dens;
        q->hl = c->pro.dr;
        q->mass = ndim == 2 ? q->rhop * q->hl * q->hl : q->rhop * q->hl * q->hl * q->hl;
        q->vxp[0] = q->vxp[1] = q->vxp[2] = 0.0;
        q->vxp[ndim - 1] = c->vrain;
        q->vx[0] = q->vx[1] = q->vx[2] = 0.0;
        q->prep = q->rhop * c->pro.dr * (-g) * c->press_factor;
        q->satu = 0.99;
        k++;
        if (k > c->n - 1) break;
        if (!(fabs(c->p[k].xp[ndim - 1] - temp_xp) < 0.1 * c->pro.dr)) break;
    } while (1);
}

/* cementpro (CPP_15-1-rain.cpp:101-131): as rainpro, radially from the cement base (vx and etype
 * untouched), while the next particle's matype equals the call's cement flag; the flag advances
 * after every call */
static void or_cementpro(or_ctx *c) {
    int k = c->n - c->nrain, ndim = c->pro.ndim, i, cf = c->cement_flag;
    double g = ndim == 2 ? c->pro.gy : c->pro.gz;
    if (c->nrain > 0) {
        do {
            OPar *q = &c->p[k];
            double drp = 0.0;
            c->nrain = c->nrain - 1;
            for (i = 0; i < ndim; i++) drp = drp + (q->xp[i] - c->cement_base[i]) * (q->xp[i] - c->cement_base[i]);
            drp = sqrt(drp);
            q->type = 1;
            q->matype = 0;
            q->rhop = c->pf[k].porosity * c->pf[k].dens;
            q->hl = c->pro.dr;
            q->mass = ndim == 2 ? q->rhop * q->hl * q->hl : q->rhop * q->hl * q->hl * q->hl;
            for (i = 0; i < 3; i++) q->vxp[i] = c->vrain * (q->xp[i] - c->cement_base[i]) / drp;
            q->prep = q->rhop * c->pro.dr * (-g) * c->press_factor;
            q->satu = 0.99;
            k++;
        } while (k < c->n - 1 && c->p[k].matype == cf);  /* the bound first: the reference reads p[n] */
    }
    c->cement_flag += 1;
}

/* waterdel_min / waterdel_max (CPP_15-2-partidel.cpp:52-130): water particles of etype 7 past
 * x_min + 3 dr (or x_max - 3 dr) on the second-to-last axis return to the reserve */
static void or_waterdel(or_ctx *c, int at_max) {
    int i, ndim = c->pro.ndim;
    double lim = at_max ? c->rx_max - 3.0 * c->pro.dr : c->rx_min + 3.0 * c->pro.dr;
    for (i = 0; i < c->n; i++) {
        OPar *q = &c->p[i];
        double x;
        if (!(q->type == 1 && q->etype == 7)) continue;
        x = q->xp[ndim - 2];
        if (at_max ? x >= lim : x <= lim) {
            q->type = 7;
            q->etype = 1;
            q->vxp[0] = q->vxp[1] = q->vxp[2] = 0.0;
            q->vx[0] = q->vx[1] = q->vx[2] = 0.0;
            q->ax[0] = q->ax[1] = q->ax[2] = 0.0;
            c->nrain = c->nrain + 1;
        }
    }
}

/* partidel (CPP_15-2-partidel.cpp:16-50): water particles 7 dr beyond the domain box in x (within
 * dr) and above y = 0.12 leave the run */
static void or_partidel(or_ctx *c) {
    int i;
    double dr = c->pro.dr;
    for (i = 0; i < c->n; i++) {
        OPar *q = &c->p[i];
        double dx1 = fabs(q->xp[0] - c->pro.xmax - 7.0 * dr), dx2 = fabs(q->xp[0] - c->pro.xmin + 7.0 * dr);
        double dy1 = q->xp[1] - 0.12;
        if (q->type == 1 && (dx1 < dr || dx2 < dr) && dy1 > 0.0) {
            q->type = 7;
            q->vxp[0] = q->vxp[1] = q->vxp[2] = 0.0;
            q->vx[0] = q->vx[1] = q->vx[2] = 0.0;
            q->ax[0] = q->ax[1] = q->ax[2] = 0.0;
        }
    }
}

/* partidel on its own, for the checker's comparison with the reference's (the step calls it after
 * every update when flr is 1) */
void or_partidel_now(or_ctx *c) { or_partidel(c); }

/* exca_or_back (CPP_16-EOB.cpp:13-74): type and etype of the stage's particles swap */
static void or_eob(or_ctx *c) {
    int i, start = c->ecp[c->nstage][0], end = c->ecp[c->nstage][1];
    for (i = start - 1; i < end; i++) {
        int t = c->p[i].type;
        c->p[i].type = c->p[i].etype;
        c->p[i].etype = t;
    }
}

/* ------------------------------------------------------------------ velocity */
/* veocity_update (CPP_11-xsph.cpp:11-142) */
static void or_velocity(or_ctx *c) {
    int n = c->n, ndim = c->pro.ndim, i, k, mv = c->pro.if_move, l = c->l, inip = c->pro.inip;
    double dt = c->pro.dt, ath;
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        if (mv == 0) {
            if (pi->type == 0 && pi->matype < 2) {
                for (k = 0; k < ndim; k++) {
                    pi->interf[k] = 0.0;
                    ath = (pi->interf[k] + pi->interf[k]) * 0.5;
                    pi->interfss[k] = pi->interfss[k] + ath * dt;
                }
            } else if ((pi->type == 0 && pi->matype == 3) || (pi->type == 0 && pi->matype == 5 && c->pro.flv == 1)) {
                for (k = 0; k < ndim; k++) {
                    ath = (pi->interf[k] + pi->interf[k]) * 0.5;
                    pi->interfss[k] = pi->interfss[k] + ath * dt;
                }
            } else if (pi->type != 0) {
                for (k = 0; k < ndim; k++) {
                    ath = (pi->axp[k] + pi->ax[k]) * 0.5;
                    pi->vx[k] = pi->vxp[k] + ath * dt;
                }
            }
        } else if (mv == 1) {
            if (pi->type == 0 && pi->matype == 2 && l > inip) {
                for (k = 0; k < ndim; k++) pi->interfss[k] = c->pro.move_vx[k];
            } else if (pi->type == 0) {
                for (k = 0; k < ndim; k++) pi->interfss[k] = 0.0;
            } else {
                for (k = 0; k < ndim; k++) {
                    ath = (pi->axp[k] + pi->ax[k]) * 0.5;
                    pi->vx[k] = pi->vxp[k] + ath * dt;
                }
            }
        } else if (mv == 2 || mv == 3) {
            int mt = mv == 2 ? 2 : 4;
            if (pi->type == mt && pi->etype == 5 && l > inip) {
                for (k = 0; k < ndim; k++) {
                    pi->axp[k] = 0.0;
                    pi->vxp[k] = c->pro.move_vx[k];
                    pi->vx[k] = c->pro.move_vx[k];
                    pi->ax[k] = 0.0;
                }
            }
            for (k = 0; k < ndim; k++) {
                ath = (pi->axp[k] + pi->ax[k]) * 0.5;
                pi->vx[k] = pi->vxp[k] + ath * dt;
            }
        }
    }
}

/* xsph (CPP_11-xsph.cpp:145-256); boundary neighbours use pPar[j].vx (quirk A16) */
static void or_xsph(or_ctx *c) {
    int n = c->n, ndim = c->pro.ndim, i;
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        double xs = 0.0, tu[3] = {0, 0, 0}, dn = 0.0;
        int nc = nbcount(&c->pc[i]), nj, k;
        for (k = 0; k < ndim; k++) c->vx_temp[i][k] = 0.0;
        for (nj = 1; nj <= nc; nj++) {
            int j = c->pc[i].influ[nj], on = 0;
            const OPar *pj = &c->p[j];
            double w3 = c->pc[i].wij[nj][3];
            if (pi->type == pj->type && pi->type != 0 && pi->type != 7) {
                if (pi->type == 1) xs = c->pro.xsph[0];
                else if (pi->type == 2) xs = c->pro.xsph[1];
                else if (pi->type == 3) xs = c->pro.xsph[2];
                else if (pi->type == 4) xs = c->pro.xsph[1];
                on = 1;
            } else if (pi->type >= 1 && pi->type <= 4 && pj->type == 0) {
                if (pi->type == 1) xs = 0.006;
                else if (pi->type == 2) xs = c->pro.xsph[1] / 5.0;
                else if (pi->type == 3) xs = 0.006;
                else if (pi->type == 4) xs = c->pro.xsph[1] / 5.0;
                on = 1;
            }
            if (!on) continue;
            for (k = 0; k < ndim; k++) tu[k] = tu[k] + xs * pj->mass * (pj->vx[k] - pi->vx[k]) * w3 / pj->rho;
            dn = dn + pj->mass * w3 / pj->rho;
        }
        if (pi->type != 0 && pi->type != 7) {
            if (fabs(dn) > 1.0e-7 && (pi->type == 2 || pi->type == 4))
                for (k = 0; k < ndim; k++) c->vx_temp[i][k] = pi->vx[k] + tu[k] / dn;
            else
                for (k = 0; k < ndim; k++) c->vx_temp[i][k] = pi->vx[k] + tu[k];
        }
    }
    for (i = 0; i < n; i++)
        if (c->p[i].type != 0 && c->p[i].type != 7)
            for (int k = 0; k < ndim; k++) c->p[i].vx[k] = c->vx_temp[i][k];
}

/* ------------------------------------------------------------------ boundary conditions */
/* velocity_set_domain / tmboundary repulsion (CPP_13-boundary.cpp:71-177) */
static void or_vel_repulse(or_ctx *c, double coe, int fluid_only) {
    int n = c->n, ndim = c->pro.ndim, i;
    double dr = c->pro.dr;
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        double tu[3] = {0, 0, 0};
        int nc = nbcount(&c->pc[i]), nj, k;
        if (fluid_only ? !(pi->type == 1 || pi->type == 3) : (pi->type == 0 || pi->type == 7)) continue;
        for (nj = 1; nj <= nc; nj++) {
            int j = c->pc[i].influ[nj];
            const OPar *pj = &c->p[j];
            double inprod = 0, dst = 0, db, beta;
            if (pj->type != 0) continue;
            for (k = 0; k < ndim; k++) {
                inprod = inprod + (pi->vx[k] - pj->interfss[k]) * (pi->xp[k] - pj->xp[k]);
                dst = dst + (pi->xp[k] - pj->xp[k]) * (pi->xp[k] - pj->xp[k]);
            }
            if (inprod < 0.0) {
                dst = sqrt(dst);
                db = coe * dr;
                beta = 1 + db / (dst - coe * dr);
                beta = fmin(1.5, beta);
                for (k = 0; k < ndim; k++) tu[k] = tu[k] + (1 - beta) * (pi->xp[k] - pj->xp[k]) * inprod / dst / dst;
            }
        }
        for (k = 0; k < ndim; k++) pi->vx[k] = pi->vx[k] + tu[k];
    }
}

/* the first three (2D: two) ids of clParti_Pair::dist_sorting (Class_particle.cpp:343-359):
 * quirk A18, only ids are swapped, distances stay in their original slots. */
static void dist_sorting(int total, int *id, const double *dst) {
    int i, j;
    for (i = 0; i < 4; i++)
        for (j = i + 1; j < total; j++)
            if (dst[i] > dst[j]) {
                int t = id[i];
                id[i] = id[j];
                id[j] = t;
            }
}

typedef struct { double x, y, z; } OPt;
typedef struct { double A, B, C, D; } OPlane;

/* cal_plane (CPP_13-boundary.cpp:654-666) */
static void cal_plane(OPt p1, OPt p2, OPt p3, OPlane *o) {
    double dst = -p1.x * (p2.y * p3.z - p3.y * p2.z) - p2.x * (p3.y * p1.z - p1.y * p3.z) -
                 p3.x * (p1.y * p2.z - p2.y * p1.z);
    o->A = p1.y * (p2.z - p3.z) + p2.y * (p3.z - p1.z) + p3.y * (p1.z - p2.z);
    o->B = p1.z * (p2.x - p3.x) + p2.z * (p3.x - p1.x) + p3.z * (p1.x - p2.x);
    o->C = p1.x * (p2.y - p3.y) + p2.x * (p3.y - p1.y) + p3.x * (p1.y - p2.y);
    o->A = o->A / dst;
    o->B = o->B / dst;
    o->C = o->C / dst;
}
/* plane_two_points (CPP_13-boundary.cpp:684-708) */
static void plane_two_points(OPt p1, OPt p2, OPlane *o) {
    double a = p1.x - p2.x, b = p1.y - p2.y, cc = p1.z - p2.z, t = -a * p2.x - b * p2.y - cc * p2.z;
    if (fabs(t) > 1.0e-6) {
        o->A = a / t; o->B = b / t; o->C = cc / t; o->D = 1.0;
    } else {
        o->A = a; o->B = b; o->C = cc; o->D = 0.0;
    }
}

/* tm_boundary_2d / tm_boundary_3d (CPP_13-boundary.cpp:179-614) */
static void or_tm_boundary(or_ctx *c) {
    int n = c->n, ndim = c->pro.ndim, i;
    double dr = c->pro.dr;
#pragma omp parallel num_threads(c->nthreads)
    {
    /* quirk: plane1 is thread-private and default-constructed once per parallel region, and
     * cal_plane never writes D, so D keeps the last plane_two_points value of the same thread */
    double sticky_D = 1.0;
#pragma omp for schedule(static)
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        int ids[NBMAX], total = 0, id = 0, nc = nbcount(&c->pc[i]), nj, k;
        double dsts[NBMAX], dst, da, db, beta;
        if (pi->type == 0 || pi->type == 7) continue;
        for (nj = 1; nj <= nc; nj++) {
            int j = c->pc[i].influ[nj];
            const OPar *pj = &c->p[j];
            if (pj->type != 0) continue;
            ids[total] = j;
            dst = (pi->xp[0] - pj->xp[0]) * (pi->xp[0] - pj->xp[0]) + (pi->xp[1] - pj->xp[1]) * (pi->xp[1] - pj->xp[1]) +
                  (pi->xp[2] - pj->xp[2]) * (pi->xp[2] - pj->xp[2]);
            dsts[total++] = sqrt(dst);
            if (pj->matype > 2) id = j;
        }
        for (k = total; k < NBMAX; k++) { ids[k] = -1; dsts[k] = 500000.0; }
        dist_sorting(total, ids, dsts);
        if ((ndim == 2 && total >= 2) || (ndim == 3 && total >= 3)) {
            OPt p4 = {pi->xp[0], pi->xp[1], pi->xp[2]};
            OPt p1 = {c->p[ids[0]].xp[0], c->p[ids[0]].xp[1], c->p[ids[0]].xp[2]};
            OPt p2 = {c->p[ids[1]].xp[0], c->p[ids[1]].xp[1], c->p[ids[1]].xp[2]};
            if (ndim == 2) {
                /* cal_line (A,B with C=1, quirk A20) + cal_dpline (CPP_13-boundary.cpp:617-652) */
                double den = p2.y * p1.x - p1.y * p2.x, A = (p1.y - p2.y) / den, B = (p2.x - p1.x) / den, tmp;
                dst = fabs(A * p4.x + B * p4.y + 1.0);
                tmp = sqrt(A * A + B * B);
                dst = tmp > 1.0e-3 ? dst / tmp : 0.0;
            } else {
                OPt p3 = {c->p[ids[2]].xp[0], c->p[ids[2]].xp[1], c->p[ids[2]].xp[2]};
                OPlane c1 = {0, 0, 0, 1.0}, c2 = {0, 0, 0, 1.0}, pl = {0, 0, 0, sticky_D};
                double cs, tmp;
                cal_plane(p1, p2, p4, &c1);
                cal_plane(p1, p3, p4, &c2);
                cs = c1.A * c2.A + c1.B * c2.B + c1.C * c2.C;
                cs = cs / sqrt(c1.A * c1.A + c1.B * c1.B + c1.C * c1.C) / sqrt(c2.A * c2.A + c2.B * c2.B + c2.C * c2.C);
                if (fabs(cs) > 0.999999) plane_two_points(p4, p1, &pl);
                else cal_plane(p1, p2, p3, &pl);
                sticky_D = pl.D;
                /* cal_dpplane (CPP_13-boundary.cpp:631-642) */
                dst = fabs(pl.A * p4.x + pl.B * p4.y + pl.C * p4.z + pl.D);
                tmp = sqrt(pl.A * pl.A + pl.B * pl.B + pl.C * pl.C);
                dst = tmp > 1.0e-3 ? dst / tmp : 0.0;
            }
            da = dst > 0.6 * dr ? dst - 0.5 * dr : 0.1 * dr;
            db = 0.5 * dr;
            if (fabs(da) > 1.0e-6) {
                beta = 1.0 + db / da;
                beta = fmin(1.5, beta);
            } else
                beta = 1.5;
            c->vel_bndy[i][0] = (1.0 - beta) * pi->vx[0] + beta * c->p[id].interfss[0];
            c->vel_bndy[i][1] = (1.0 - beta) * pi->vx[1] + beta * c->p[id].interfss[1];
            c->vel_bndy[i][2] = ndim == 2 ? 0.0 : (1.0 - beta) * pi->vx[2] + beta * c->p[id].interfss[2];
        } else {
            for (k = 0; k < 3; k++) c->vel_bndy[i][k] = pi->vx[k];
        }
    }
    }
    /* second loop: Shepard average per phase with the tempup_vel[t][0] quirk (A17) */
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        double tv[4][3] = {{0}}, td[4] = {0, 0, 0, 0};
        int nc = nbcount(&c->pc[i]), nj, t, k;
        double *out[4];
        if (pi->type != 0) continue;
        for (nj = 1; nj <= nc; nj++) {
            int j = c->pc[i].influ[nj];
            const OPar *pj = &c->p[j];
            double w3 = c->pc[i].wij[nj][3];
            if (pj->type < 1 || pj->type > 4) continue;
            t = pj->type - 1;
            tv[t][0] = tv[t][0] + pj->mass * c->vel_bndy[j][0] * w3 / pj->rho;
            tv[t][1] = tv[t][0] + pj->mass * c->vel_bndy[j][1] * w3 / pj->rho;
            tv[t][2] = tv[t][0] + pj->mass * c->vel_bndy[j][2] * w3 / pj->rho;
            td[t] = td[t] + pj->mass * w3 / pj->rho;
        }
        out[0] = c->vb[i].vx_water; out[1] = c->vb[i].vx_soil; out[2] = c->vb[i].vx_air; out[3] = c->vb[i].vx_struct;
        for (t = 0; t < 4; t++)
            for (k = 0; k < 3; k++) out[t][k] = fabs(td[t]) > 1.0e-6 ? tv[t][k] / td[t] : 0.0;
    }
}

/* no_slip_tran / free_slip_tran (CPP_13-boundary.cpp:711-958, :960-1237): every boundary particle
 * takes, per phase (water, soil, air, structure), the Shepard averages of its neighbours' stress
 * (soil: effective stress), density and velocity; its virtual velocity is interfss minus the
 * average velocity (no slip) or interfss plus the average of the velocity with the component
 * along the particle pair removed (free slip, pairs closer than 0.1 dr skipped) */
static void or_slip_tran(or_ctx *c, int free_slip) {
    int n = c->n, i;
    double dr = c->pro.dr;
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        double ts[4][6] = {{0}}, tv[4][3] = {{0}}, tr[4] = {0, 0, 0, 0}, td[4] = {0, 0, 0, 0};
        double *vxo[4];
        int nc = nbcount(&c->pc[i]), nj, k, t;
        if (pi->type != 0) continue;
        for (nj = 1; nj <= nc; nj++) {
            int j = c->pc[i].influ[nj];
            const OPar *pj = &c->p[j];
            double w3 = c->pc[i].wij[nj][3], v1[3], dst, inprod;
            if (pj->type < 1 || pj->type > 4) continue;
            t = pj->type - 1;
            v1[0] = pi->xp[0] - pj->xp[0];
            v1[1] = pi->xp[1] - pj->xp[1];
            v1[2] = pi->xp[2] - pj->xp[2];
            dst = sqrt(v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2]);
            inprod = pj->vx[0] * v1[0] + pj->vx[1] * v1[1] + pj->vx[2] * v1[2];
            if (t == 1) {
                double satu = pj->satu;
                for (k = 0; k < 3; k++)
                    ts[t][k] = ts[t][k] + pj->mass * (pj->sig[k] - satu * pj->prew + (1 - satu) * pj->prea) * w3 / pj->rho;
                for (k = 3; k < 6; k++) ts[t][k] = ts[t][k] + pj->mass * pj->sig[k] * w3 / pj->rho;
            } else {
                for (k = 0; k < 6; k++) ts[t][k] = ts[t][k] + pj->mass * pj->sig[k] * w3 / pj->rho;
            }
            if (!free_slip) {
                for (k = 0; k < 3; k++) tv[t][k] = tv[t][k] + pj->mass * pj->vx[k] * w3 / pj->rho;
            } else if (dst > 0.1 * dr) {
                for (k = 0; k < 3; k++)
                    tv[t][k] = tv[t][k] + pj->mass * (pj->vx[k] - inprod * v1[k] / dst / dst) * w3 / pj->rho;
            }
            tr[t] = tr[t] + pj->mass * pj->rho * w3 / pj->rho;
            td[t] = td[t] + pj->mass * w3 / pj->rho;
        }
        vxo[0] = c->vb[i].vx_water; vxo[1] = c->vb[i].vx_soil; vxo[2] = c->vb[i].vx_air; vxo[3] = c->vb[i].vx_struct;
        for (t = 0; t < 4; t++) {
            if (fabs(td[t]) > 1.0e-6) {
                for (k = 0; k < 6; k++) c->vb[i].sig_ph[t][k] = ts[t][k] / td[t];
                for (k = 0; k < 3; k++)
                    vxo[t][k] = free_slip ? pi->interfss[k] + tv[t][k] / td[t] : pi->interfss[k] - tv[t][k] / td[t];
                c->vb[i].rho_ph[t] = tr[t] / td[t];
            } else {
                for (k = 0; k < 6; k++) c->vb[i].sig_ph[t][k] = 0.0;
                for (k = 0; k < 3; k++) vxo[t][k] = pi->interfss[k];
                c->vb[i].rho_ph[t] = 0.0;
            }
        }
    }
}

/* check_domain (CPP_13-boundary.cpp:20-68), serial semantics of the shared flag/vbx (A3) */
static void or_check_domain(or_ctx *c) {
    const sphg_problem *P = &c->pro;
    double bmax[3] = {P->xmax, P->ymax, P->zmax}, bmin[3] = {P->xmin, P->ymin, P->zmin}, vbx[3] = {0, 0, 0};
    double dr = P->dr;
    int flag = 0, i, k;
    for (i = 0; i < c->n; i++) {
        OPar *pi = &c->p[i];
        if (pi->type != 0 && pi->type != 7) {
            int hit = 0;
            if (pi->xp[0] - bmax[0] >= -0.5 * dr) hit = 1;
            if (pi->xp[0] - bmin[0] <= 0.5 * dr) hit = 1;
            if (pi->xp[1] - bmax[1] >= -0.5 * dr) hit = 1;
            if (pi->xp[1] - bmin[1] <= 0.5 * dr) hit = 1;
            if (pi->xp[2] - bmax[2] >= -0.5 * dr && P->ndim == 3) hit = 1;
            if (pi->xp[2] - bmin[2] <= 0.5 * dr && P->ndim == 3) hit = 1;
            if (hit)
                for (k = 0; k < 3; k++) pi->vx[k] = vbx[k];
        } else if (pi->type == 0 && pi->matype > 2 && flag == 0) {
            flag = 1;
            for (k = 0; k < 3; k++) vbx[k] = pi->interfss[k];
        }
    }
}

/* update (CPP_12-1-update.cpp:11-55) */
static void or_update(or_ctx *c) {
    int n = c->n, ndim = c->pro.ndim, i, k, loop = c->l - c->pro.inip;
    double dt = c->pro.dt;
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        for (k = 0; k < ndim; k++) {
            if (pi->type != 0) {
                pi->x[k] = pi->xp[k] + (pi->vx[k] + pi->vxp[k]) * dt * 0.5;
                if (loop == 0) pi->ux[k] = 0.0;
                else pi->ux[k] = pi->ux[k] + (pi->vx[k] + pi->vxp[k]) * dt * 0.5;
                pi->xp[k] = pi->x[k];
                pi->vxp[k] = pi->vx[k];
                pi->axp[k] = pi->ax[k];
                pi->rhop = pi->rho;
            } else if (pi->matype >= 2 && pi->matype <= 5) {
                pi->x[k] = pi->xp[k] + (pi->interfss[k] + pi->interfss[k]) * dt * 0.5;
                pi->ux[k] = pi->ux[k] + (pi->interfss[k] + pi->interfss[k]) * dt * 0.5;
                pi->xp[k] = pi->x[k];
                pi->rhop = pi->rho;
                pi->prep = pi->pre;
            }
        }
    }
}

/* ------------------------------------------------------------------ soil-structure interaction */
/* interaction_ss_rigid / _deformable / _penalty (CPP_10-1-interact_ss.cpp:17-1038) with the
 * geometry helpers of the same file (:1039-1369), in the serial (one thread) order: one
 * clInterf_SS slot per particle, structure slots accumulated in ascending soil-particle order.
 * The OpenMP-private clLine / clPlane keep their constructor's C = 1 / D = 1 (never assigned)
 * and their last unit normal when set_normal() finds a degenerate one (Class_Problem.cpp:194-258). */
typedef struct { double A, B, C, nx, ny; } OLine;
typedef struct { double x0, y0, z0, a, b, c; } OLine3;
typedef struct { double A, B, C, D, nx, ny, nz; } OPlaneN;
typedef struct { double x, y, z; } OVec;

static double vdot(OVec a, OVec b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static OVec vsub(OVec a, OVec b) { OVec r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }

static double ss_dpline(OPt p, const OLine *l) { /* cal_dpline :1224-1236 */
    double dst = fabs(l->A * p.x + l->B * p.y + l->C), temp = sqrt(l->A * l->A + l->B * l->B);
    return temp > 1.0e-3 ? dst / temp : 0.0;
}
static double ss_dpplane(OPt p, const OPlaneN *q) { /* cal_dpplane :1238-1250 */
    double dst = fabs(q->A * p.x + q->B * p.y + q->C * p.z + q->D), temp = sqrt(q->A * q->A + q->B * q->B + q->C * q->C);
    return temp > 1.0e-3 ? dst / temp : 0.0;
}
static OPt ss_map_line(OPt r, const OLine *l) { /* cal_mappoint_line :1094-1125 (z of the result 0) */
    double dA = l->A, dB = l->B, dp = ss_dpline(r, l), temp = sqrt(dA * dA + dB * dB);
    OPt p1 = {r.x + dp * dA / temp, r.y + dp * dB / temp, 0.0}, p2 = {r.x - dp * dA / temp, r.y - dp * dB / temp, 0.0};
    return ss_dpline(p1, l) < 1.0e-3 ? p1 : p2;
}
static OPt ss_map_plane(OPt r, const OPlaneN *q) { /* cal_mappoint_plane :1127-1160 */
    double dA = q->A, dB = q->B, dC = q->C, dp = ss_dpplane(r, q), temp = sqrt(dA * dA + dB * dB + dC * dC);
    OPt p1 = {r.x + dp * dA / temp, r.y + dp * dB / temp, r.z + dp * dC / temp};
    OPt p2 = {r.x - dp * dA / temp, r.y - dp * dB / temp, r.z - dp * dC / temp};
    return ss_dpplane(p1, q) < 1.0e-3 ? p1 : p2;
}
static void ss_line3(OPt p1, OPt p2, OLine3 *o) { /* cal_line3D :1078-1092 */
    double dst = 0.0;
    o->x0 = p1.x; o->y0 = p1.y; o->z0 = p1.z;
    dst = (p2.x - p1.x) * (p2.x - p1.x) + (p2.y - p1.y) * (p2.y - p1.y) + (p2.z - p1.z) * (p2.z - p1.z);
    dst = sqrt(dst);
    o->a = (p2.x - p1.x) / dst; o->b = (p2.y - p1.y) / dst; o->c = (p2.z - p1.z) / dst;
}
static int ss_intersect(const OLine3 *l1, const OLine3 *l2, OPt *o) { /* cal_intersect_lines3D :1162-1185 */
    double flag = l1->a * l2->b - l2->a * l1->b, t;
    if (!(fabs(flag) > 1.0e-3)) return 1;
    t = (l2->a * (l1->y0 - l2->y0) - l2->b * (l1->x0 - l2->x0)) / flag;
    o->x = l1->x0 + l1->a * t; o->y = l1->y0 + l1->b * t; o->z = l1->z0 + l1->c * t;
    return 0;
}
static double pdist(OPt a, OPt b) {
    return sqrt((b.x - a.x) * (b.x - a.x) + (b.y - a.y) * (b.y - a.y) + (b.z - a.z) * (b.z - a.z));
}
/* linear_interpolate :1187-1202 */
static void ss_interp(OPt p1, OPt p2, OPt r, double m1, double m2, const double v1[3], const double v2[3],
                      double *m, double v[3]) {
    double xk1 = 0.0, xk2 = pdist(p1, p2), xp = pdist(p1, r);
    int k;
    *m = (m1 * fabs(xp - xk2) + m2 * fabs(xp - xk1)) / fabs(xk2 - xk1);
    for (k = 0; k < 3; k++) v[k] = (v1[k] * fabs(xp - xk2) + v2[k] * fabs(xp - xk1)) / fabs(xk2 - xk1);
}
/* linear_interpolate_force :1204-1222 */
static void ss_interp_force(OPt p1, OPt p2, OPt r, const double f[3], double o1[3], double o2[3]) {
    double xk1 = 0.0, xk2 = pdist(p1, p2), xp = pdist(p1, r);
    int k;
    for (k = 0; k < 3; k++) o1[k] = -fabs(xp - xk2) / fabs(xk1 - xk2) * f[k];
    for (k = 0; k < 3; k++) o2[k] = -fabs(xp - xk1) / fabs(xk1 - xk2) * f[k];
}

static void or_ssi(or_ctx *c) {
    int n = c->n, ndim = c->pro.ndim, i, k, var = c->pro.ss_type;
    double dt = c->pro.dt, dr = c->pro.dr, myu = c->pro.ss_myu, zeta = c->pro.ss_zeta;
    double reach = var == 1 ? 0.5 * dr : dr;
    double (*fn)[3] = (double (*)[3])calloc((size_t)n, sizeof *fn), (*ft)[3] = (double (*)[3])calloc((size_t)n, sizeof *ft);
    OLine line = {0.0, 0.0, 1.0, 0.0, 0.0};
    OPlaneN plane = {0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0};
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        int ids[SPHG_MAX_NB + 1], total = 0, nc = nbcount(&c->pc[i]), nj, id1, id2, id3;
        double dsts[SPHG_MAX_NB + 1], vs[3] = {0, 0, 0}, vn[3] = {0, 0, 0}, vt[3] = {0, 0, 0}, ms, dp, mag_fn, mag_ft;
        double f1n[3], f1t[3], f2n[3], f2t[3], f3n[3], f3t[3];
        OPt p1, p2, p3, p4, outp;
        if (!(pi->type == 1 || pi->type == 2)) continue;
        /* structure neighbours in list order with their distances, then dist_sorting */
        for (nj = 1; nj <= nc; nj++) {
            int j = c->pc[i].influ[nj];
            const OPar *pj = &c->p[j];
            double d;
            if (pj->type != 4) continue;
            ids[total] = j;
            d = (pi->xp[0] - pj->xp[0]) * (pi->xp[0] - pj->xp[0]) + (pi->xp[1] - pj->xp[1]) * (pi->xp[1] - pj->xp[1]) +
                (pi->xp[2] - pj->xp[2]) * (pi->xp[2] - pj->xp[2]);
            dsts[total] = sqrt(d);
            total++;
        }
        dist_sorting(total, ids, dsts);
        p4.x = pi->xp[0]; p4.y = pi->xp[1]; p4.z = pi->xp[2];
        if (ndim == 2) {
            OVec a, b, q, v0, v1;
            double u, den;
            if (total < 2) continue;
            id1 = ids[0]; id2 = ids[1];
            p1.x = c->p[id1].xp[0]; p1.y = c->p[id1].xp[1]; p1.z = c->p[id1].xp[2];
            p2.x = c->p[id2].xp[0]; p2.y = c->p[id2].xp[1]; p2.z = c->p[id2].xp[2];
            den = p2.y * p1.x - p1.y * p2.x; /* cal_line :1055-1061 */
            line.A = (p1.y - p2.y) / den;
            line.B = (p2.x - p1.x) / den;
            dp = ss_dpline(p4, &line);
            outp = ss_map_line(p4, &line);
            a.x = p1.x; a.y = p1.y; a.z = p1.z; b.x = p2.x; b.y = p2.y; b.z = p2.z;
            q.x = outp.x; q.y = outp.y; q.z = 0.0;
            v0 = vsub(b, a); v1 = vsub(q, a); /* check_ifinside_line :1296-1315 */
            u = 1.0 / vdot(v0, v0) * vdot(v0, v1);
            if (u < 0 || u > 1) continue;
            if (!(dp < reach)) continue;
            {
                double t = sqrt(line.A * line.A + line.B * line.B), dot;
                OVec e, nn;
                if (t > 1.0e-10) { line.nx = line.A / t; line.ny = line.B / t; }
                e.x = outp.x - p4.x; e.y = outp.y - p4.y; e.z = 0.0;
                nn.x = line.nx; nn.y = line.ny; nn.z = 0.0;
                dot = vdot(e, nn);
                if (!(dot <= 0.0)) { line.nx = -line.nx; line.ny = -line.ny; }
            }
            ss_interp(p1, p2, outp, c->p[id1].mass, c->p[id2].mass, c->p[id1].vxp, c->p[id2].vxp, &ms, vs);
            vs[0] -= pi->vxp[0];
            vs[1] -= pi->vxp[1];
            if (var != 1) {
                vs[0] *= ms / (ms + pi->mass);
                vs[1] *= ms / (ms + pi->mass);
            }
            {
                double v = vs[0] * line.nx + vs[1] * line.ny; /* decomposition_velocity2D */
                vn[0] = v * line.nx; vn[1] = v * line.ny;
                vt[0] = vs[0] - vn[0]; vt[1] = vs[1] - vn[1];
            }
            if (var == 3) {
                mag_fn = (1.0 - zeta) * 2.0 * pi->mass * (dr - dp) / dt / dt;
                fn[i][0] = mag_fn * line.nx; fn[i][1] = mag_fn * line.ny;
            } else {
                fn[i][0] = pi->mass * vn[0] / dt; fn[i][1] = pi->mass * vn[1] / dt;
            }
            ft[i][0] = pi->mass * vt[0] / dt; ft[i][1] = pi->mass * vt[1] / dt;
            mag_fn = sqrt(fn[i][0] * fn[i][0] + fn[i][1] * fn[i][1]);
            mag_ft = sqrt(ft[i][0] * ft[i][0] + ft[i][1] * ft[i][1]);
            if (mag_ft > (mag_fn * myu)) {
                vt[0] = vt[0] * myu * mag_fn / mag_ft; vt[1] = vt[1] * myu * mag_fn / mag_ft;
                ft[i][0] = pi->mass * vt[0] / dt; ft[i][1] = pi->mass * vt[1] / dt;
            }
            ss_interp_force(p1, p2, outp, fn[i], f1n, f2n);
            ss_interp_force(p1, p2, outp, ft[i], f1t, f2t);
            for (k = 0; k < 3; k++) {
                fn[id1][k] += f1n[k]; fn[id2][k] += f2n[k];
                ft[id1][k] += f1t[k]; ft[id2][k] += f2t[k];
            }
        } else {
            OVec A, B, C, P, v0, v1, v2;
            double d00, d01, d02, d11, d12, inv, u, v, tms, tvs[3];
            OLine3 l31, l32;
            OPt sect;
            if (total < 3) continue;
            id1 = ids[0]; id2 = ids[1]; id3 = ids[2];
            p1.x = c->p[id1].xp[0]; p1.y = c->p[id1].xp[1]; p1.z = c->p[id1].xp[2];
            p2.x = c->p[id2].xp[0]; p2.y = c->p[id2].xp[1]; p2.z = c->p[id2].xp[2];
            p3.x = c->p[id3].xp[0]; p3.y = c->p[id3].xp[1]; p3.z = c->p[id3].xp[2];
            {
                OPlane q;
                cal_plane(p1, p2, p3, &q); /* same expressions as CPP_10-1-interact_ss.cpp:1063-1076 */
                plane.A = q.A; plane.B = q.B; plane.C = q.C;
            }
            dp = ss_dpplane(p4, &plane);
            outp = ss_map_plane(p4, &plane);
            A.x = p1.x; A.y = p1.y; A.z = p1.z; B.x = p2.x; B.y = p2.y; B.z = p2.z;
            C.x = p3.x; C.y = p3.y; C.z = p3.z; P.x = outp.x; P.y = outp.y; P.z = outp.z;
            v0 = vsub(C, A); v1 = vsub(B, A); v2 = vsub(P, A); /* check_ifinside_triangle :1317-1344 */
            d00 = vdot(v0, v0); d01 = vdot(v0, v1); d02 = vdot(v0, v2); d11 = vdot(v1, v1); d12 = vdot(v1, v2);
            inv = 1 / (d00 * d11 - d01 * d01);
            u = (d11 * d02 - d01 * d12) * inv;
            if (u < 0 || u > 1) continue;
            v = (d00 * d12 - d01 * d02) * inv;
            if (v < 0 || v > 1) continue;
            if (!(u + v <= 1)) continue;
            if (!(dp < reach)) continue;
            {
                double t = sqrt(plane.A * plane.A + plane.B * plane.B + plane.C * plane.C), dot;
                OVec e, nn;
                if (t > 1.0e-10) { plane.nx = plane.A / t; plane.ny = plane.B / t; plane.nz = plane.C / t; }
                e.x = outp.x - p4.x; e.y = outp.y - p4.y; e.z = outp.z - p4.z;
                nn.x = plane.nx; nn.y = plane.ny; nn.z = plane.nz;
                dot = vdot(e, nn);
                if (!(dot <= 0.0)) { plane.nx = -plane.nx; plane.ny = -plane.ny; plane.nz = -plane.nz; }
            }
            ss_line3(p1, p2, &l31);
            ss_line3(p3, outp, &l32);
            sect.x = sect.y = sect.z = 0.0;
            if (ss_intersect(&l31, &l32, &sect)) continue;
            ss_interp(p1, p2, sect, c->p[id1].mass, c->p[id2].mass, c->p[id1].vxp, c->p[id2].vxp, &tms, tvs);
            ss_interp(p3, sect, outp, c->p[id3].mass, tms, c->p[id3].vxp, tvs, &ms, vs);
            for (k = 0; k < 3; k++) vs[k] -= pi->vxp[k];
            if (var != 1)
                for (k = 0; k < 3; k++) vs[k] *= ms / (ms + pi->mass);
            {
                double w = vs[0] * plane.nx + vs[1] * plane.ny + vs[2] * plane.nz; /* decomposition_velocity3D */
                vn[0] = w * plane.nx; vn[1] = w * plane.ny; vn[2] = w * plane.nz;
                for (k = 0; k < 3; k++) vt[k] = vs[k] - vn[k];
            }
            if (var == 3) {
                mag_fn = (1.0 - zeta) * 2.0 * pi->mass * (dr - dp) / dt / dt;
                fn[i][0] = mag_fn * plane.nx; fn[i][1] = mag_fn * plane.ny; fn[i][2] = mag_fn * plane.nz;
            } else {
                for (k = 0; k < 3; k++) fn[i][k] = pi->mass * vn[k] / dt;
            }
            for (k = 0; k < 3; k++) ft[i][k] = pi->mass * vt[k] / dt;
            mag_fn = sqrt(fn[i][0] * fn[i][0] + fn[i][1] * fn[i][1] + fn[i][2] * fn[i][2]);
            mag_ft = sqrt(ft[i][0] * ft[i][0] + ft[i][1] * ft[i][1] + ft[i][2] * ft[i][2]);
            if (mag_ft > (mag_fn * myu)) {
                for (k = 0; k < 3; k++) vt[k] = vt[k] * myu * mag_fn / mag_ft;
                for (k = 0; k < 3; k++) ft[i][k] = pi->mass * vt[k] / dt;
            }
            ss_interp_force(p3, sect, outp, fn[i], f1n, f3n);
            ss_interp_force(p3, sect, outp, ft[i], f1t, f3t);
            for (k = 0; k < 3; k++) { fn[id3][k] += f1n[k]; ft[id3][k] += f1t[k]; }
            ss_interp_force(p1, p2, sect, f3n, f1n, f2n);
            ss_interp_force(p1, p2, sect, f3t, f1t, f2t);
            for (k = 0; k < 3; k++) {
                fn[id1][k] -= f1n[k]; fn[id2][k] -= f2n[k];
                ft[id1][k] -= f1t[k]; ft[id2][k] -= f2t[k];
            }
        }
    }
    /* thread slots into the particles (one slot here) */
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        if (!(pi->type == 1 || pi->type == 2 || pi->type == 4)) continue;
        for (k = 0; k < 3; k++) {
            pi->interfss[k] = 0.0;
            if (var == 3) pi->interfss[k] += ft[i][k] + fn[i][k];
            else pi->interfss[k] += zeta * (ft[i][k] + fn[i][k]);
            pi->ax[k] += pi->interfss[k] / pi->mass;
        }
    }
    free(fn);
    free(ft);
}

/* boundary_free_solid (CPP_13-boundary.cpp:1239-1515), flbndy 5: contact of a particle with the
 * wall line (2D: types 1-4) or triangle (3D: every particle, as the reference's loop has no type
 * test) of its nearest boundary particles, with the interaction_ss_rigid geometry, the walls'
 * interfss as their velocity, and zeta f / m added to ax directly; 2D adds 0.0 to ax[2] */
static void or_free_solid(or_ctx *c) {
    int n = c->n, ndim = c->pro.ndim, i, k;
    double dt = c->pro.dt, dr = c->pro.dr, myu = c->pro.ss_myu, zeta = c->pro.ss_zeta;
    OLine line = {0.0, 0.0, 1.0, 0.0, 0.0};
    OPlaneN plane = {0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0};
    for (i = 0; i < n; i++) {
        OPar *pi = &c->p[i];
        int ids[SPHG_MAX_NB + 1], total = 0, nc = nbcount(&c->pc[i]), nj, id1, id2, id3;
        double dsts[SPHG_MAX_NB + 1], vs[3] = {0, 0, 0}, vn[3] = {0, 0, 0}, vt[3] = {0, 0, 0}, fn[3] = {0, 0, 0},
                                       ft[3] = {0, 0, 0}, ms, dp, mag_fn, mag_ft;
        OPt p1, p2, p3, p4, outp;
        if (ndim == 2 && !(pi->type >= 1 && pi->type <= 4)) continue;
        for (nj = 1; nj <= nc; nj++) {
            int j = c->pc[i].influ[nj];
            const OPar *pj = &c->p[j];
            double d;
            if (pj->type != 0) continue;
            ids[total] = j;
            d = (pi->xp[0] - pj->xp[0]) * (pi->xp[0] - pj->xp[0]) + (pi->xp[1] - pj->xp[1]) * (pi->xp[1] - pj->xp[1]) +
                (pi->xp[2] - pj->xp[2]) * (pi->xp[2] - pj->xp[2]);
            dsts[total] = sqrt(d);
            total++;
        }
        dist_sorting(total, ids, dsts);
        p4.x = pi->xp[0]; p4.y = pi->xp[1]; p4.z = pi->xp[2];
        if (ndim == 2) {
            OVec a, b, q, v0, v1;
            double u, den;
            if (total < 2) continue;
            id1 = ids[0]; id2 = ids[1];
            p1.x = c->p[id1].xp[0]; p1.y = c->p[id1].xp[1]; p1.z = c->p[id1].xp[2];
            p2.x = c->p[id2].xp[0]; p2.y = c->p[id2].xp[1]; p2.z = c->p[id2].xp[2];
            den = p2.y * p1.x - p1.y * p2.x;
            line.A = (p1.y - p2.y) / den;
            line.B = (p2.x - p1.x) / den;
            dp = ss_dpline(p4, &line);
            outp = ss_map_line(p4, &line);
            a.x = p1.x; a.y = p1.y; a.z = p1.z; b.x = p2.x; b.y = p2.y; b.z = p2.z;
            q.x = outp.x; q.y = outp.y; q.z = 0.0;
            v0 = vsub(b, a); v1 = vsub(q, a);
            u = 1.0 / vdot(v0, v0) * vdot(v0, v1);
            if (u < 0 || u > 1) continue;
            if (!(dp < 0.5 * dr)) continue;
            {
                double t = sqrt(line.A * line.A + line.B * line.B), dot;
                OVec e, nn;
                if (t > 1.0e-10) { line.nx = line.A / t; line.ny = line.B / t; }
                e.x = outp.x - p4.x; e.y = outp.y - p4.y; e.z = 0.0;
                nn.x = line.nx; nn.y = line.ny; nn.z = 0.0;
                dot = vdot(e, nn);
                if (!(dot <= 0.0)) { line.nx = -line.nx; line.ny = -line.ny; }
            }
            ss_interp(p1, p2, outp, c->p[id1].mass, c->p[id2].mass, c->p[id1].interfss, c->p[id2].interfss, &ms, vs);
            vs[0] -= pi->vxp[0];
            vs[1] -= pi->vxp[1];
            {
                double v = vs[0] * line.nx + vs[1] * line.ny;
                vn[0] = v * line.nx; vn[1] = v * line.ny;
                vt[0] = vs[0] - vn[0]; vt[1] = vs[1] - vn[1];
            }
            fn[0] = pi->mass * vn[0] / dt; fn[1] = pi->mass * vn[1] / dt;
            ft[0] = pi->mass * vt[0] / dt; ft[1] = pi->mass * vt[1] / dt;
            mag_fn = sqrt(fn[0] * fn[0] + fn[1] * fn[1]);
            mag_ft = sqrt(ft[0] * ft[0] + ft[1] * ft[1]);
            if (mag_ft > (mag_fn * myu)) {
                vt[0] = vt[0] * myu * mag_fn / mag_ft; vt[1] = vt[1] * myu * mag_fn / mag_ft;
                ft[0] = pi->mass * vt[0] / dt; ft[1] = pi->mass * vt[1] / dt;
            }
            pi->ax[0] += zeta * (ft[0] + fn[0]) / pi->mass;
            pi->ax[1] += zeta * (ft[1] + fn[1]) / pi->mass;
            pi->ax[2] += 0.0;
        } else {
            OVec A, B, C, P, v0, v1, v2;
            double d00, d01, d02, d11, d12, inv, u, v, tms, tvs[3];
            OLine3 l31, l32;
            OPt sect;
            if (total < 3) continue;
            id1 = ids[0]; id2 = ids[1]; id3 = ids[2];
            p1.x = c->p[id1].xp[0]; p1.y = c->p[id1].xp[1]; p1.z = c->p[id1].xp[2];
            p2.x = c->p[id2].xp[0]; p2.y = c->p[id2].xp[1]; p2.z = c->p[id2].xp[2];
            p3.x = c->p[id3].xp[0]; p3.y = c->p[id3].xp[1]; p3.z = c->p[id3].xp[2];
            {
                OPlane q;
                cal_plane(p1, p2, p3, &q);
                plane.A = q.A; plane.B = q.B; plane.C = q.C;
            }
            dp = ss_dpplane(p4, &plane);
            outp = ss_map_plane(p4, &plane);
            A.x = p1.x; A.y = p1.y; A.z = p1.z; B.x = p2.x; B.y = p2.y; B.z = p2.z;
            C.x = p3.x; C.y = p3.y; C.z = p3.z; P.x = outp.x; P.y = outp.y; P.z = outp.z;
            v0 = vsub(C, A); v1 = vsub(B, A); v2 = vsub(P, A);
            d00 = vdot(v0, v0); d01 = vdot(v0, v1); d02 = vdot(v0, v2); d11 = vdot(v1, v1); d12 = vdot(v1, v2);
            inv = 1 / (d00 * d11 - d01 * d01);
            u = (d11 * d02 - d01 * d12) * inv;
            if (u < 0 || u > 1) continue;
            v = (d00 * d12 - d01 * d02) * inv;
            if (v < 0 || v > 1) continue;
            if (!(u + v <= 1)) continue;
            if (!(dp < 0.5 * dr)) continue;
            {
                double t = sqrt(plane.A * plane.A + plane.B * plane.B + plane.C * plane.C), dot;
                OVec e, nn;
                if (t > 1.0e-10) { plane.nx = plane.A / t; plane.ny = plane.B / t; plane.nz = plane.C / t; }
                e.x = outp.x - p4.x; e.y = outp.y - p4.y; e.z = outp.z - p4.z;
                nn.x = plane.nx; nn.y = plane.ny; nn.z = plane.nz;
                dot = vdot(e, nn);
                if (!(dot <= 0.0)) { plane.nx = -plane.nx; plane.ny = -plane.ny; plane.nz = -plane.nz; }
            }
            ss_line3(p1, p2, &l31);
            ss_line3(p3, outp, &l32);
            sect.x = sect.y = sect.z = 0.0;
            if (ss_intersect(&l31, &l32, &sect)) continue;
            ss_interp(p1, p2, sect, c->p[id1].mass, c->p[id2].mass, c->p[id1].interfss, c->p[id2].interfss, &tms, tvs);
            ss_interp(p3, sect, outp, c->p[id3].mass, tms, c->p[id3].interfss, tvs, &ms, vs);
            for (k = 0; k < 3; k++) vs[k] -= pi->vxp[k];
            {
                double w = vs[0] * plane.nx + vs[1] * plane.ny + vs[2] * plane.nz;
                vn[0] = w * plane.nx; vn[1] = w * plane.ny; vn[2] = w * plane.nz;
                for (k = 0; k < 3; k++) vt[k] = vs[k] - vn[k];
            }
            for (k = 0; k < 3; k++) fn[k] = pi->mass * vn[k] / dt;
            for (k = 0; k < 3; k++) ft[k] = pi->mass * vt[k] / dt;
            mag_fn = sqrt(fn[0] * fn[0] + fn[1] * fn[1] + fn[2] * fn[2]);
            mag_ft = sqrt(ft[0] * ft[0] + ft[1] * ft[1] + ft[2] * ft[2]);
            if (mag_ft > (mag_fn * myu)) {
                for (k = 0; k < 3; k++) vt[k] = vt[k] * myu * mag_fn / mag_ft;
                for (k = 0; k < 3; k++) ft[k] = pi->mass * vt[k] / dt;
            }
            for (k = 0; k < 3; k++) pi->ax[k] += zeta * (ft[k] + fn[k]) / pi->mass;
        }
    }
}

/* ------------------------------------------------------------------ one step */
static int unsupported_flags(const or_ctx *c) {
    const sphg_problem *P = &c->pro;
    if (P->flknl < 1 || P->flknl > 4 || P->flr < 0 || P->flr > 1 || P->fleob < 0 || P->fleob > 1 || P->flv < 0 ||
        P->flv > 2)
        return 1;
    if (P->flv > 0 && !c->vib_loads) return 1;
    if (P->flr == 1 && !c->rain_set) return 1;
    if (P->fleob == 1 && !c->ecp) return 1;
    if (P->fli < 0 || P->fli > 3 || P->ss_type < 0 || P->ss_type > 3) return 1;
    if (P->flbndy != 1 && P->flbndy != 2 && P->flbndy != 3 && P->flbndy != 4 && P->flbndy != 5 && P->flbndy != 13 &&
        P->flbndy != 14)
        return 1;
    return 0;
}

/* One reference step as phase groups. The grouping only moves eos loop 1 (which reads the
 * final densities and writes pressures nothing in between reads) ahead of the strain pass, so the
 * arithmetic is the reference's loop body unchanged (CPP_1-SPHLoops.cpp:524-753). */
int or_phase(or_ctx *c, int mask) {
    int e;
    if (unsupported_flags(c)) return SPHG_E_UNSUPPORTED;
    if (mask & 1) {
        int lr = c->l - c->pro.inip;
        c->t = c->t + c->pro.dt;
        /* rain / cement production and water-level decrease, then excavation or backfill
         * (CPP_1-SPHLoops.cpp:495-520) */
        if (c->pro.flr == 1 && lr > 0 && c->lp > 0 && lr % c->lp == 0) {
            if (c->rtype == 1) or_rainpro(c);
            else if (c->rtype == 2) or_waterdel(c, 0);
            else if (c->rtype == 3) or_waterdel(c, 1);
            else if (c->rtype == 4) or_cementpro(c);
        }
        if (c->pro.fleob == 1 && c->nstage < c->ecount && c->l == c->ecp[c->nstage][2]) {
            or_eob(c);
            c->nstage += 1;
        }
        if ((e = or_parttocell(c))) return e;
        if ((e = or_partisearching(c))) return e;
        if (c->pro.flknl > 2) or_kgc(c);
    }
    if (mask & 2) {
        or_density(c);
        or_eos1(c);
    }
    if (mask & 4) {
        or_strain(c);
        if (c->l % c->pro.step_regu == 0) or_regu(c, 0);
        if (c->pro.flagum == 1) or_saturation(c);
        if (c->pro.fli == 1) or_permeability(c);
        or_eos2(c);
    }
    if (mask & 8) {
        if ((e = or_strain_to_stress(c))) return e;
        if (c->l % c->pro.step_regu == 0) or_regu(c, 1);
        or_free_surface(c);
        or_eig(c);
    }
    if (mask & 16) {
        /* CPP_1-SPHLoops.cpp:586-600: the water/air routine by acc_type of water[0] / air */
        or_fluid_accel(c, 1, 3, c->mat.water[0].acc_type == 0 ? 1 : -1);
        or_fluid_accel(c, 3, 1, c->mat.air.acc_type == 0 ? 1 : -1);
        or_solid_accel(c, 2);
        or_solid_accel(c, 4);
        or_art_visc(c);
        if (c->pro.art_stre_coe > 0.0) or_art_stress(c);
        if (c->pro.flbndy == 1 || c->pro.flbndy == 2 || c->pro.flbndy == 5) or_bndy_stress(c);
        else or_bndy_stress_tran(c);
        if (c->pro.flv > 0) {
            or_free_field_bndy(c);
            or_free_field_soil(c);
        }
        if (c->pro.ss_type >= 1 && c->pro.ss_type <= 3 && c->pro.ss_acc_type == 2) or_coupled(c);
        if (c->pro.flv == 1) or_vibration(c, 0);
        if (c->pro.fli == 1) or_interact1(c);
        else if (c->pro.fli == 2 || c->pro.fli == 3) or_interact23(c);
        if (c->pro.ss_type >= 1 && c->pro.ss_type <= 3) or_ssi(c);
        if (c->l <= c->pro.inip) or_damping_geostress(c);
        if (c->pro.flv > 0) or_damping_vibration(c);
        if (c->pro.flbndy == 5) or_free_solid(c);
        or_velocity(c);
        if (c->pro.flv == 2) or_vibration(c, 1);
    }
    if (mask & 32) {
        if (c->pro.if_move != 2) or_xsph(c);
        if (c->pro.flbndy == 1) or_vel_repulse(c, 0.30, 0);
        else if (c->pro.flbndy == 2) {
            or_vel_repulse(c, c->pro.coe_bndy, 1);
            or_tm_boundary(c);
        } else if (c->pro.flbndy == 3 || c->pro.flbndy == 4) {
            or_slip_tran(c, c->pro.flbndy == 4);
        } else if (c->pro.flbndy == 13 || c->pro.flbndy == 14) {
            or_vel_repulse(c, c->pro.coe_bndy, 1);
            or_slip_tran(c, c->pro.flbndy == 14);
        }
    }
    if (mask & 64) {
        or_check_domain(c);
        or_update(c);
        if (c->pro.flr == 1) or_partidel(c);
        c->l++;
    }
    return 0;
}

int or_step(or_ctx *c, int nsteps) {
    int s, e;
    for (s = 0; s < nsteps; s++)
        if ((e = or_phase(c, 127))) return e;
    return 0;
}

int or_get_neighbors(or_ctx *c, int *counts, int *ids, int n) {
    int i, k;
    if (n != c->n) return SPHG_E_ARG;
    for (i = 0; i < n; i++) {
        int nc = nbcount(&c->pc[i]);
        counts[i] = c->pc[i].ninflu;
        for (k = 0; k < SPHG_MAX_NB; k++) ids[(size_t)i * SPHG_MAX_NB + k] = k < nc ? c->pc[i].influ[k + 1] : -1;
    }
    return 0;
}

/* kernel values stored by the last list build for particle i (Par_Cell::wij[1..ninflu]) */
int or_get_wij(or_ctx *c, int i, double *out, int cap) {
    int nc = nbcount(&c->pc[i]), k;
    for (k = 0; k < nc && k < cap; k++) memcpy(out + 4 * k, c->pc[i].wij[k + 1], 32);
    return nc;
}

long long or_pair_count(or_ctx *c) {
    long long s = 0;
    for (int i = 0; i < c->n; i++) s += c->pc[i].ninflu;
    return s;
}
