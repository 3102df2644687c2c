// Host side of the C-ABI (include/sphg.h): context, HBM layout, field transfer, step scheduling.
//
// One context owns one HBM pool holding every particle array (SoA), the cell list and the
// neighbour lists, one HIP stream, and two captured hipGraphs (one per position/velocity buffer
// parity) that replay a whole reference step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstddef>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sphg.h"
#include "sphg_kernels.hip"
#include "sphg_kl.hip"
#include "sphg_ssi.hip"
#include "sphg_rf.hip"

using namespace sphg;

namespace {

constexpr int kBlock = 256;
constexpr int kKlTermsHost = SPHG_KL_TERMS;

struct Timing {
    std::string name;
    hipEvent_t a, b;
};

}  // namespace

struct sphg_ctx {
    sphg_problem pro;
    sphg_materials mat;
    int cap = 0, n = 0, precision = 0;
    Params P{};
    State S{};
    char *pool = nullptr;
    size_t pool_bytes = 0;
    hipStream_t stream = nullptr;
    int cur = 0;        // which xp / vx buffer holds the current state
    int step = 1;       // reference step counter l of the next step
    bool uploaded = false;
    bool graph_ok[2] = {false, false};
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    std::vector<int> h_type, h_matype, h_etype;
    bool timing = false;
    std::vector<Timing> tq;
    std::vector<std::pair<std::string, std::pair<double, int>>> tsum;
    SoilSlot *d_soil = nullptr;
    FluidSlot *d_fluid = nullptr;
    double *d_psc = nullptr;
    bool has_psc = false;
    // ps_ds / ps_cop / ps_fai set by the caller: upload_done keeps them instead of the slot values
    bool has_psds = false, has_pscop = false, has_psfai = false;
    // pf_fai / pf_c (Bingham water strength) set by the caller or by a random-field call
    bool has_pffai = false, has_pfc = false;
    // other per-particle material copies the caller set (ps_e, pf_fai, ...): the step reads these
    // from the material slots, so upload_done refuses them unless they equal the slot values
    std::map<std::string, std::vector<double>> pinned_mat;
    int scan_blocks = 0;
    bool stepped = false;
    bool regu_on = false;  // regularisation passes are in the launch sequence
    bool pre_on = false;   // artificial-stress / geostress-damping pair pass is in the launch sequence
    bool full_models = false;  // some soil particle uses Cam-Clay, unsaturated Cam-Clay or FOCC
    bool rev_lists = false;    // every neighbour list reversed after the search (noise-floor runs)
    // Para_Vibra (flv 1, 2): period, loads per period, window; the table is in S.vib_loads
    double vib_ttime = 0.0, vib_sttime = 0.0, vib_edtime = 0.0;
    int vib_nsteps = 0;
    // Para_Rain (flr 1) and Para_EC (fleob 1) as the caller set them; eob_len: longest stage
    bool rain_set = false;
    sphg_rain rain{};
    int ecount = 0, eob_len = 0;
    bool eig_stale = false;    // strep / beta not written by the last stress pass (k_eig on read)
    // neighbours per load round trip of the density, normalisation, strain and XSPH passes
    // (sphg_autotune); each has the widths kBwChoices[k]
    int bw[4] = {2, 6, 2, 2};
    // list entries the search stages in LDS per thread (16 or 32; sphg_autotune, sphg_set_search_stage)
    int search_stage = 32;
    std::vector<Grid> grids;   // one per realisation (host copy)
    std::vector<int> counts;   // particles per realisation (batch contexts)
    Grid *d_grids = nullptr;
    int *d_rid = nullptr;
};

#define HIPCHK(x)                                                               \
    do {                                                                        \
        hipError_t e_ = (x);                                                    \
        if (e_ != hipSuccess) {                                                 \
            fprintf(stderr, "sphg: %s failed: %s\n", #x, hipGetErrorString(e_)); \
            return SPHG_E_HIP;                                                  \
        }                                                                       \
    } while (0)

static inline int grid_for(int n) { return (n + kBlock - 1) / kBlock; }

// -------------------------------------------------------------------------------- creation
static size_t align256(size_t x) { return (x + 255) & ~size_t(255); }
static constexpr size_t align256c(size_t x) { return (x + 255) & ~size_t(255); }

// Cell grid of one simulation: celllist_ini1 (CPP_3-NNPS.cpp:96-122) and the nncell stencil of
// celllist_ini2 (:125-171) in reference order. Returns false when the grid is empty or too large.
static bool make_grid(const sphg_box &b, int ndim, double dr, Grid &g) {
    const double cs = 2.0 * dr;
    g = Grid{};
    g.xmin[0] = b.xmin; g.xmin[1] = b.ymin; g.xmin[2] = b.zmin;
    g.bmin[0] = b.xmin; g.bmin[1] = b.ymin; g.bmin[2] = b.zmin;
    g.bmax[0] = b.xmax; g.bmax[1] = b.ymax; g.bmax[2] = b.zmax;
    g.cxm = (int)((b.xmax - b.xmin) / cs + 3.0);
    g.cym = (int)((b.ymax - b.ymin) / cs + 3.0);
    g.czm = ndim == 2 ? 1 : (int)((b.zmax - b.zmin) / cs + 3.0);
    long long ct = (long long)g.cxm * g.cym * g.czm;
    if (ct <= 0 || ct > (1LL << 30)) return false;
    g.ncell = (int)ct;
    int p = 0;
    if (ndim == 2) {
        for (int ky = 1; ky <= 3; ky++)
            for (int kx = 1; kx <= 3; kx++) g.nb_off[p++] = (kx - 2) * g.cym + ky - 2;
    } else {
        for (int kz = 1; kz <= 3; kz++)
            for (int kx = 1; kx <= 3; kx++)
                for (int ky = 1; ky <= 3; ky++) g.nb_off[p++] = (ky - 2) * g.cxm * g.czm + (kx - 2) * g.czm + kz - 2;
    }
    g.nstencil = p;
    return true;
}

static int create_common(sphg_ctx **out, const sphg_problem *pro, const sphg_materials *mat, int capacity,
                         int precision, int nreal, const int *counts, const sphg_box *boxes) {
    if (!out || !pro || !mat) return SPHG_E_ARG;
    if (capacity <= 0) return SPHG_E_NOPARTICLES;
    if (capacity > (1 << 28)) return SPHG_E_CAPACITY;  // list entries pack the type above 28 id bits
    if (pro->ndim != 2 && pro->ndim != 3) return SPHG_E_ARG;
    if (precision != SPHG_FP64 && precision != SPHG_MIXED) return SPHG_E_ARG;
    sphg_ctx *c = new sphg_ctx();
    c->pro = *pro;
    c->mat = *mat;
    c->cap = capacity;
    c->search_stage = pro->ndim == 2 ? 16 : 32;  // 2D lists hold about 10-20 entries, 3D about 30
    c->precision = precision;
    Params &P = c->P;
    P.cell_size = 2.0 * pro->dr;
    P.nreal = nreal;
    long long cells = 0;
    c->grids.resize(nreal);
    for (int r = 0; r < nreal; r++) {
        if (!make_grid(boxes[r], pro->ndim, pro->dr, c->grids[r])) {
            delete c;
            return SPHG_E_CAPACITY;
        }
        c->grids[r].base = (int)cells;
        cells += c->grids[r].ncell;
    }
    if (cells > (1LL << 30)) {
        delete c;
        return SPHG_E_CAPACITY;
    }
    if (counts) c->counts.assign(counts, counts + nreal);
    P.ncell = (int)cells;  // all realisations' cells, numbered consecutively

    // HBM layout
    const size_t cap = (size_t)capacity;
    const int ncellp = P.ncell + 2;
    c->scan_blocks = (ncellp + kScanTile - 1) / kScanTile;
    std::vector<std::pair<void **, size_t>> plan;
    State &S = c->S;
    auto D = [&](double *&ptr) { plan.push_back({(void **)&ptr, cap * 8}); };
    auto I = [&](int *&ptr, size_t cnt) { plan.push_back({(void **)&ptr, cnt * 4}); };
    for (int b = 0; b < 2; b++)
        for (int k = 0; k < 3; k++) {
            D(S.xp[b][k]);
            D(S.vx[b][k]);
        }
    for (int k = 0; k < 3; k++) {
        D(S.vxp[k]); D(S.ax[k]); D(S.axp[k]); D(S.ux[k]); D(S.interf[k]); D(S.interfss[k]);
        D(S.strep[k]); D(S.weps[k]); D(S.vel_bndy[k]);
        for (int t = 0; t < 4; t++) D(S.vb[t][k]);
    }
    for (int k = 0; k < 6; k++) {
        D(S.eps[k]); D(S.deps[k]); D(S.sig[k]); D(S.adps[k]);
    }
    for (int k = 0; k < 9; k++) D(S.beta[k]);
    for (double **f : {&S.rho, &S.rho_mid, &S.rhop, &S.mass, &S.hl, &S.pre, &S.prep, &S.e, &S.satu, &S.dsatu,
                       &S.suct, &S.cop, &S.vsig, &S.veps, &S.divq, &S.divr, &S.meps, &S.roue, &S.prew, &S.prea,
                       &S.porosity, &S.cd, &S.poro_cop, &S.color})
        D(*f);
    plan.push_back({(void **)&c->d_psc, cap * 8});
    D(S.ps_ds); D(S.ps_cop); D(S.ps_fai);
    D(S.pf_fai); D(S.pf_c); D(S.pf_tan);
    S.kgc = nullptr;
    if (pro->flknl > 2) plan.push_back({(void **)&S.kgc, 9 * cap * 8});
    for (int k = 0; k < 6; k++) S.tssig[k] = nullptr;
    S.tsqc = nullptr;
    if (pro->flts >= 1) {
        for (int k = 0; k < 6; k++) D(S.tssig[k]);
        D(S.tsqc);
    }
    for (int k = 0; k < 3; k++) D(S.ps_dp[k]);
    I(S.type, cap); I(S.matype, cap); I(S.permtype, cap); I(S.etype, cap);
    plan.push_back({(void **)&S.soil_slot, cap});
    plan.push_back({(void **)&S.fluid_slot, cap});
    I(S.cell, cap); I(S.sorted, cap); I(S.nb_cnt, cap); I(S.nb_types, cap);
    I(S.nbr, cap * kMaxNb);
    for (int k = 0; k < 6; k++) D(S.qsig[k]);
    D(S.fdens);
    S.mra = S.mrc = S.mrv = nullptr;
    S.mrn = nullptr;
    if (precision == SPHG_MIXED) {
        plan.push_back({(void **)&S.mra, cap * sizeof(float4)});
        plan.push_back({(void **)&S.mrc, cap * 3 * sizeof(float4)});
        plan.push_back({(void **)&S.mrv, cap * sizeof(float4)});
        plan.push_back({(void **)&S.mrn, cap * sizeof(float2)});
    }
    // artificial-stress tensor (kept between steps) and the Jacobi regularisation copy
    for (int k = 0; k < 9; k++) {
        if (pro->art_stre_coe > 0.0) D(S.rij[k]);
        else S.rij[k] = nullptr;
    }
    for (int k = 0; k < 6; k++) D(S.tmp6[k]);
    const bool split_brow = pro->ndim == 2 && precision != SPHG_MIXED && pro->flknl <= 2;
    for (int k = 0; k < 3; k++) {
        if (split_brow) D(S.brow[k]);
        else S.brow[k] = nullptr;
    }
    S.nw_list = S.nw_count = nullptr;
    if (split_brow) {
        I(S.nw_list, cap);
        I(S.nw_count, 1);
    }
    for (int q = 0; q < 3; q++) {
        if (pro->ss_type >= 1 && pro->ss_type <= 3) {
            I(S.ssi_id[q], cap);
            for (int k = 0; k < 6; k++) D(S.ssi_c[q][k]);
        } else {
            S.ssi_id[q] = nullptr;
            for (int k = 0; k < 6; k++) S.ssi_c[q][k] = nullptr;
        }
    }
    for (int k = 0; k < 3; k++) {
        if (pro->ss_type >= 1 && pro->ss_type <= 3 && pro->ss_acc_type == 2) D(S.cpl[k]);
        else S.cpl[k] = nullptr;
    }
    const bool tran = pro->flbndy == 3 || pro->flbndy == 4 || pro->flbndy == 13 || pro->flbndy == 14;
    for (int t = 0; t < 4; t++)
        for (int k = 0; k < 6; k++) {
            if (tran) D(S.sigb[t][k]);
            else S.sigb[t][k] = nullptr;
        }
    I(S.cell_count, ncellp); I(S.cell_start, ncellp); I(S.cell_rank, cap);
    I(S.block_sums, (size_t)c->scan_blocks + 1);
    I(S.err, 1); I(S.step, 1); I(S.shear3, 1); I(S.plane_dirty, 1);
    plan.push_back({(void **)&S.dtp, 8});
    plan.push_back({(void **)&S.dt_red, 16});
    plan.push_back({(void **)&S.tnow, 8});
    I(S.dyn, 4);  // rain reserve, cement flag, excavation stage (dynamic particle sets)
    S.ecp = nullptr;
    // vibration loads and free-field boundaries: the clock above, the tangent stiffness and the
    // pair sums (the load table is allocated by sphg_set_vibration)
    S.vib_loads = nullptr;
    S.vstiff = nullptr;
    for (int k = 0; k < 3; k++) S.ffb[k] = S.ffs[k] = nullptr;
    for (int k = 0; k < 4; k++) S.vdamp[k] = nullptr;
    if (pro->flv > 0) {
        D(S.vstiff);
        for (int k = 0; k < 3; k++) { D(S.ffb[k]); D(S.ffs[k]); }
        for (int k = 0; k < 4; k++) D(S.vdamp[k]);
    }
    plan.push_back({(void **)&c->d_grids, sizeof(Grid) * (size_t)nreal});
    I(c->d_rid, cap);
    plan.push_back({(void **)&c->d_soil, sizeof(SoilSlot) * kNSlots});
    plan.push_back({(void **)&c->d_fluid, sizeof(FluidSlot) * 8});
    size_t total = 0;
    for (auto &e : plan) total += align256(e.second);
    if (hipMalloc(&c->pool, total) != hipSuccess) {
        delete c;
        return SPHG_E_CAPACITY;
    }
    c->pool_bytes = total;
    size_t off = 0;
    for (auto &e : plan) {
        *e.first = c->pool + off;
        off += align256(e.second);
    }
    S.soil = c->d_soil;
    S.fluid = c->d_fluid;
    S.grids = c->d_grids;
    S.rid = c->d_rid;
    S.ps_c = nullptr;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        hipFree(c->pool);
        delete c;
        return SPHG_E_HIP;
    }
    // Particle() zero-initialises every member (Class_particle.cpp:12-99); do the same.
    if (hipMemsetAsync(c->pool, 0, total, c->stream) != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess) {
        sphg_destroy(c);
        return SPHG_E_HIP;
    }
    c->h_type.assign(cap, 0);
    c->h_matype.assign(cap, 0);
    c->h_etype.assign(cap, 0);
    *out = c;
    return SPHG_OK;
}

extern "C" int sphg_create(sphg_ctx **out, const sphg_problem *pro, const sphg_materials *mat, int capacity,
                           int precision) {
    if (!pro) return SPHG_E_ARG;
    sphg_box b{pro->xmin, pro->xmax, pro->ymin, pro->ymax, pro->zmin, pro->zmax};
    return create_common(out, pro, mat, capacity, precision, 1, nullptr, &b);
}

extern "C" int sphg_create_batch(sphg_ctx **out, const sphg_problem *pro, const sphg_materials *mat, int nreal,
                                 const int *counts, const sphg_box *boxes, int precision) {
    if (!pro || !counts || !boxes || nreal <= 0) return SPHG_E_ARG;
    long long tot = 0;
    for (int r = 0; r < nreal; r++) {
        if (counts[r] <= 0) return SPHG_E_NOPARTICLES;
        tot += counts[r];
    }
    if (tot > (1 << 28)) return SPHG_E_CAPACITY;
    return create_common(out, pro, mat, (int)tot, precision, nreal, counts, boxes);
}

extern "C" void sphg_destroy(sphg_ctx *c) {
    if (!c) return;
    if (c->stream) hipStreamSynchronize(c->stream);
    for (int b = 0; b < 2; b++)
        if (c->graph[b]) hipGraphExecDestroy(c->graph[b]);
    for (auto &t : c->tq) {
        hipEventDestroy(t.a);
        hipEventDestroy(t.b);
    }
    if (c->pool) hipFree(c->pool);
    if (c->S.vib_loads) hipFree(c->S.vib_loads);
    if (c->S.ecp) hipFree(c->S.ecp);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

extern "C" int sphg_device_bytes(const sphg_ctx *c, size_t *bytes) {
    if (!c || !bytes) return SPHG_E_ARG;
    *bytes = c->pool_bytes;
    return SPHG_OK;
}

// -------------------------------------------------------------------------------- fields
// Returns the device component arrays of a named field (nullptr if unknown).
static int field_ptrs(sphg_ctx *c, const char *name, double **out, int &ncomp, bool for_write,
                      std::vector<double *> *mirror) {
    State &S = c->S;
    const int cur = c->cur;
    std::string nm(name);
    auto v3 = [&](double *const a[3]) { for (int k = 0; k < 3; k++) out[k] = a[k]; ncomp = 3; return 0; };
    auto v6 = [&](double *const a[6]) { for (int k = 0; k < 6; k++) out[k] = a[k]; ncomp = 6; return 0; };
    auto s1 = [&](double *a) { out[0] = a; ncomp = 1; return 0; };
    if (nm == "xp") {
        if (mirror) for (int k = 0; k < 3; k++) mirror->push_back(S.xp[cur ^ 1][k]);
        return v3(S.xp[cur]);
    }
    if (nm == "vx") {
        if (mirror) for (int k = 0; k < 3; k++) mirror->push_back(S.vx[cur ^ 1][k]);
        return v3(S.vx[cur]);
    }
    if (nm == "vxp") return v3(S.vxp);
    if (nm == "ax") return v3(S.ax);
    if (nm == "axp") return v3(S.axp);
    if (nm == "ux") return v3(S.ux);
    if (nm == "interf") return v3(S.interf);
    if (nm == "interfss") return v3(S.interfss);
    if (nm == "strep") return v3(S.strep);
    if (nm == "vel_bndy") return v3(S.vel_bndy);
    if (nm == "vx_water") return v3(S.vb[0]);
    if (nm == "vx_soil") return v3(S.vb[1]);
    if (nm == "vx_air") return v3(S.vb[2]);
    if (nm == "vx_struct") return v3(S.vb[3]);
    if (nm == "eps") return v6(S.eps);
    if (nm == "deps") return v6(S.deps);
    if (nm == "sig") return v6(S.sig);
    if (nm == "adps") return v6(S.adps);
    if (nm == "beta") {
        for (int k = 0; k < 9; k++) out[k] = S.beta[k];
        ncomp = 9;
        return 0;
    }
    if (nm == "tssig") {
        if (!S.tssig[0]) return -1;
        return v6(S.tssig);
    }
    if (nm == "tsqc") {
        if (!S.tsqc) return -1;
        return s1(S.tsqc);
    }
    if (nm == "Rij") {
        if (!S.rij[0]) return -1;
        for (int k = 0; k < 9; k++) out[k] = S.rij[k];
        ncomp = 9;
        return 0;
    }
    struct { const char *n; double *p; } sc[] = {
        {"rho", S.rho}, {"rhop", S.rhop}, {"mass", S.mass}, {"hl", S.hl}, {"pre", S.pre}, {"prep", S.prep},
        {"e", S.e}, {"satu", S.satu}, {"dsatu", S.dsatu}, {"suct", S.suct}, {"cop", S.cop}, {"vsig", S.vsig},
        {"veps", S.veps}, {"divq", S.divq}, {"divr", S.divr}, {"meps", S.meps}, {"roue", S.roue},
        {"prew", S.prew}, {"prea", S.prea}, {"porosity", S.porosity}, {"cd", S.cd}, {"poro_cop", S.poro_cop},
        {"color", S.color}, {"ps_c", c->d_psc}, {"ps_ds", S.ps_ds}, {"ps_cop", S.ps_cop},
        {"ps_fai", S.ps_fai}, {"pf_fai", S.pf_fai}, {"pf_c", S.pf_c}, {"rho_mid", S.rho_mid}};
    for (auto &e : sc)
        if (nm == e.n) return s1(e.p);
    return -1;
}

// strep / beta of the current stress when the last stress pass left them out (no consumer in the
// step): the same Stress_eigValue of the same stress, so the values are those of the pass; with
// them vsig / veps / divq / divr / meps where the step leaves those out too (Params::lazy_deq).
static int ensure_eig(sphg_ctx *c) {
    if (!c->eig_stale) return SPHG_OK;
    hipLaunchKernelGGL(k_eig, dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, c->S, c->P);
    if (c->P.lazy_deq) hipLaunchKernelGGL(k_deq, dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, c->S, c->P);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess) return SPHG_E_HIP;
    c->eig_stale = false;
    return SPHG_OK;
}

// Fields derived on download from what the device keeps (weps: 3 stored components; beta_ts:
// transpose of beta; x: equal to xp for every particle the update moves).
extern "C" int sphg_get_field(sphg_ctx *c, const char *name, double *host, int n, int ncomp) {
    if (!c || !name || !host || n != c->n) return SPHG_E_ARG;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return SPHG_E_HIP;
    std::string nm(name);
    if (nm == "strep" || nm == "beta" || nm == "beta_ts" || nm == "vsig" || nm == "veps" || nm == "divq" ||
        nm == "divr" || nm == "meps") {
        const int rc = ensure_eig(c);
        if (rc) return rc;
    }
    const size_t N = (size_t)n;
    std::vector<double> buf(N);
    auto fetch = [&](double *d, std::vector<double> &h) {
        h.resize(N);
        return hipMemcpy(h.data(), d, N * 8, hipMemcpyDeviceToHost) == hipSuccess;
    };
    if (nm == "weps" || nm == "beta_ts" || nm == "x" || nm == "divx" || nm == "dsig") {
        if (nm == "weps") {
            if (ncomp != 9) return SPHG_E_ARG;
            std::vector<double> w[3];
            for (int k = 0; k < 3; k++)
                if (!fetch(c->S.weps[k], w[k])) return SPHG_E_HIP;
            for (size_t i = 0; i < N; i++) {
                double a = w[0][i], b = w[1][i], d = w[2][i];
                auto ng = [](double x) { return x == 0.0 ? 0.0 : -x; };
                double m[9] = {0.0, a, b, ng(a), 0.0, d, ng(b), ng(d), 0.0};
                memcpy(host + 9 * i, m, 72);
            }
            return SPHG_OK;
        }
        if (nm == "beta_ts") {
            if (ncomp != 9) return SPHG_E_ARG;
            std::vector<double> b[9];
            for (int k = 0; k < 9; k++)
                if (!fetch(c->S.beta[k], b[k])) return SPHG_E_HIP;
            for (size_t i = 0; i < N; i++)
                for (int r = 0; r < 3; r++)
                    for (int q = 0; q < 3; q++) host[9 * i + 3 * r + q] = b[3 * q + r][i];
            return SPHG_OK;
        }
        if (nm == "x") {
            if (ncomp != 3) return SPHG_E_ARG;
            std::vector<double> x[3];
            for (int k = 0; k < 3; k++)
                if (!fetch(c->S.xp[c->cur][k], x[k])) return SPHG_E_HIP;
            for (size_t i = 0; i < N; i++) {
                int t = c->h_type[i], m = c->h_matype[i];
                bool moved = c->stepped && (t != 0 || (m >= 2 && m <= 5));
                for (int k = 0; k < 3; k++)
                    host[3 * i + k] = (moved && k < c->pro.ndim) ? x[k][i] : 0.0;
            }
            return SPHG_OK;
        }
        return SPHG_E_ARG;  // divx / dsig are per-step temporaries not kept in HBM
    }
    double *ptr[9];
    int nc = 0;
    if (field_ptrs(c, name, ptr, nc, false, nullptr) != 0 || nc != ncomp) return SPHG_E_ARG;
    for (int k = 0; k < nc; k++) {
        if (!fetch(ptr[k], buf)) return SPHG_E_HIP;
        for (size_t i = 0; i < N; i++) host[nc * i + k] = buf[i];
    }
    return SPHG_OK;
}

// Para_Soil / Para_Fluid members as per-particle copies "ps_<field>" / "pf_<field>"
// (Class_ParticleProperty.h:10-133) and where each sits in sphg_soil / sphg_fluid
#define SPHG_SOIL_MEMBER(m) {"ps_" #m, offsetof(sphg_soil, m)}
#define SPHG_FLUID_MEMBER(m) {"pf_" #m, offsetof(sphg_fluid, m)}
static const std::pair<const char *, size_t> kSoilMembers[] = {
    SPHG_SOIL_MEMBER(e), SPHG_SOIL_MEMBER(poi), SPHG_SOIL_MEMBER(eps0), SPHG_SOIL_MEMBER(dens),
    SPHG_SOIL_MEMBER(damp1), SPHG_SOIL_MEMBER(damp2), SPHG_SOIL_MEMBER(ramda), SPHG_SOIL_MEMBER(kapa),
    SPHG_SOIL_MEMBER(zmf), SPHG_SOIL_MEMBER(ann), SPHG_SOIL_MEMBER(bnn), SPHG_SOIL_MEMBER(beta),
    SPHG_SOIL_MEMBER(sitas), SPHG_SOIL_MEMBER(sitar), SPHG_SOIL_MEMBER(epsr), SPHG_SOIL_MEMBER(epss),
    SPHG_SOIL_MEMBER(sd), SPHG_SOIL_MEMBER(sw), SPHG_SOIL_MEMBER(kse), SPHG_SOIL_MEMBER(c1),
    SPHG_SOIL_MEMBER(c2), SPHG_SOIL_MEMBER(c3), SPHG_SOIL_MEMBER(ocr), SPHG_SOIL_MEMBER(sme0),
    SPHG_SOIL_MEMBER(mzm), SPHG_SOIL_MEMBER(azm), SPHG_SOIL_MEMBER(r0), SPHG_SOIL_MEMBER(bl),
    SPHG_SOIL_MEMBER(bz), SPHG_SOIL_MEMBER(zet), SPHG_SOIL_MEMBER(vpu), SPHG_SOIL_MEMBER(porosity),
    SPHG_SOIL_MEMBER(fs), SPHG_SOIL_MEMBER(etao), SPHG_SOIL_MEMBER(g0), SPHG_SOIL_MEMBER(damp_coe)};
static const std::pair<const char *, size_t> kFluidMembers[] = {
    SPHG_FLUID_MEMBER(vis), SPHG_FLUID_MEMBER(vpu), SPHG_FLUID_MEMBER(dens), SPHG_FLUID_MEMBER(paa0),
    SPHG_FLUID_MEMBER(porosity), SPHG_FLUID_MEMBER(fai), SPHG_FLUID_MEMBER(c), SPHG_FLUID_MEMBER(pre_coe),
    SPHG_FLUID_MEMBER(strain_min)};
#undef SPHG_SOIL_MEMBER
#undef SPHG_FLUID_MEMBER

// 0 when every pinned per-particle material value equals the value of the particle's slot
static int check_pinned_materials(const sphg_ctx *c, const std::vector<sphg_soil> &soil_of_slot,
                                  const std::vector<sphg_fluid> &fluid_of_slot, const std::vector<int8_t> &ss,
                                  const std::vector<int8_t> &fs) {
    for (const auto &e : c->pinned_mat) {
        for (const auto &m : kSoilMembers)
            if (e.first == m.first)
                for (int i = 0; i < c->n; i++) {
                    const double v = *(const double *)((const char *)&soil_of_slot[ss[i]] + m.second);
                    if (memcmp(&v, &e.second[i], 8) != 0) return SPHG_E_UNSUPPORTED;
                }
        for (const auto &m : kFluidMembers)
            if (e.first == m.first)
                for (int i = 0; i < c->n; i++) {
                    const double v = *(const double *)((const char *)&fluid_of_slot[fs[i]] + m.second);
                    if (memcmp(&v, &e.second[i], 8) != 0) return SPHG_E_UNSUPPORTED;
                }
    }
    return 0;
}

// Constants the step derives from per-particle strength, evaluated with the host's libm as the slot
// constants are: the Drucker-Prager factors of ps_fai (phi' = atan(tan(fai) / fs),
// CPP_8-1-consfactor.cpp:294, :783) and tan of the Bingham friction angle pf_fai
// (CPP_7-2-constitutive.cpp:42-43). Called at upload and whenever either array changes after it.
static int soil_slot_of(int type, int matype, int etype);
static int fluid_slot_of(int type, int matype, int etype);

static int derive_strength_constants(sphg_ctx *c) {
    const int n = c->n;
    std::vector<double> fai(n), dp(3 * (size_t)n);
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(fai.data(), c->S.ps_fai, (size_t)n * 8, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++) {
        const int s = soil_slot_of(c->h_type[i], c->h_matype[i], c->h_etype[i]);
        const double fs = s < SPHG_NSOIL ? c->mat.soil[s].fs : 1.0;
        const double f = atan(tan(fai[i] * 0.0174532925166667) / fs);
        const double sq = sqrt(9.0 + 3.0 * sin(f) * sin(f));
        dp[i] = sin(f) / sq;
        dp[n + i] = cos(f);
        dp[2 * (size_t)n + i] = sq;
    }
    for (int k = 0; k < 3; k++)
        HIPCHK(hipMemcpy(c->S.ps_dp[k], dp.data() + (size_t)k * n, (size_t)n * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(fai.data(), c->S.pf_fai, (size_t)n * 8, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++) fai[i] = tan(fai[i] * 0.0174532925199);
    HIPCHK(hipMemcpy(c->S.pf_tan, fai.data(), (size_t)n * 8, hipMemcpyHostToDevice));
    return SPHG_OK;
}

static int derive_mass_flags(sphg_ctx *c);
static int derive_plane_zero(sphg_ctx *c);
static void drop_graphs(sphg_ctx *c);

extern "C" int sphg_set_field(sphg_ctx *c, const char *name, const double *host, int n, int ncomp) {
    if (!c || !name || !host || n <= 0 || n > c->cap) return SPHG_E_ARG;
    // outputs formed on read from the stress (k_eig, k_deq) are formed from the stress of the last
    // step, before the caller changes anything
    if (ensure_eig(c)) return SPHG_E_HIP;
    if (c->n == 0) c->n = n;
    if (n != c->n) return SPHG_E_ARG;
    std::string nm(name);
    const size_t N = (size_t)n;
    if (nm == "x" || nm == "divx" || nm == "dsig" || nm == "beta_ts" || nm == "weps" || nm == "rous" ||
        (nm.rfind("pf_", 0) == 0 && nm != "pf_fai" && nm != "pf_c") ||
        (nm.rfind("ps_", 0) == 0 && nm != "ps_c" && nm != "ps_ds" && nm != "ps_cop" && nm != "ps_fai")) {
        // derived, per-step temporary, or taken from the material slots
        bool member = false;
        for (const auto &m : kSoilMembers) member = member || nm == m.first;
        for (const auto &m : kFluidMembers) member = member || nm == m.first;
        if (member) {
            if (ncomp != 1) return SPHG_E_ARG;
            c->pinned_mat[nm].assign(host, host + n);
        }
        return SPHG_OK;
    }
    double *ptr[9];
    int nc = 0;
    std::vector<double *> mirror;
    if (field_ptrs(c, name, ptr, nc, true, &mirror) != 0 || nc != ncomp) return SPHG_E_ARG;
    std::vector<double> buf(N);
    for (int k = 0; k < nc; k++) {
        for (size_t i = 0; i < N; i++) buf[i] = host[nc * i + k];
        HIPCHK(hipMemcpy(ptr[k], buf.data(), N * 8, hipMemcpyHostToDevice));
        if (!mirror.empty()) HIPCHK(hipMemcpy(mirror[k], buf.data(), N * 8, hipMemcpyHostToDevice));
    }
    if (nm == "sig" && nc == 6 && c->uploaded && c->P.ndim == 2) {
        // new stresses after sphg_upload_done: out-of-plane shear raises the flag the stress pass keeps
        int sh3 = 0;
        for (size_t i = 0; i < N && !sh3; i++) sh3 = host[6 * i + 4] != 0.0 || host[6 * i + 5] != 0.0;
        if (sh3) HIPCHK(hipMemcpy(c->S.shear3, &sh3, 4, hipMemcpyHostToDevice));
    }
    if (nm == "mass" && c->uploaded)
        if (int rc = derive_mass_flags(c)) return rc;
    if ((nm == "sig" || nm == "eps" || nm == "adps" || nm == "deps") && c->uploaded && c->P.plane_zero) {
        c->P.plane_zero = 0;  // the plane step's zero components are no longer known
        drop_graphs(c);
    }
    if (nm == "ps_c") c->has_psc = true;
    if (nm == "ps_ds") c->has_psds = true;
    if (nm == "ps_cop") c->has_pscop = true;
    if (nm == "ps_fai") c->has_psfai = true;
    if (nm == "pf_fai") c->has_pffai = true;
    if (nm == "pf_c") c->has_pfc = true;
    if (c->uploaded && (nm == "ps_fai" || nm == "pf_fai")) return derive_strength_constants(c);
    return SPHG_OK;
}

extern "C" int sphg_set_field_i(sphg_ctx *c, const char *name, const int *host, int n) {
    if (!c || !name || !host || n <= 0 || n > c->cap) return SPHG_E_ARG;
    if (ensure_eig(c)) return SPHG_E_HIP;
    if (c->n == 0) c->n = n;
    if (n != c->n) return SPHG_E_ARG;
    std::string nm(name);
    int *d = nullptr;
    if (nm == "type") { d = c->S.type; std::copy(host, host + n, c->h_type.begin()); }
    else if (nm == "matype") { d = c->S.matype; std::copy(host, host + n, c->h_matype.begin()); }
    else if (nm == "permtype") d = c->S.permtype;
    else if (nm == "etype") { d = c->S.etype; std::copy(host, host + n, c->h_etype.begin()); }
    else return SPHG_E_ARG;
    HIPCHK(hipMemcpy(d, host, (size_t)n * 4, hipMemcpyHostToDevice));
    c->uploaded = false;
    return SPHG_OK;
}

extern "C" int sphg_get_field_i(sphg_ctx *c, const char *name, int *host, int n) {
    if (!c || !name || !host || n != c->n) return SPHG_E_ARG;
    std::string nm(name);
    int *d = nm == "type" ? c->S.type : nm == "matype" ? c->S.matype : nm == "permtype" ? c->S.permtype
             : nm == "etype" ? c->S.etype : nullptr;
    if (!d) return SPHG_E_ARG;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(host, d, (size_t)n * 4, hipMemcpyDeviceToHost));
    return SPHG_OK;
}

// setting_vis_cons slot selection (CPP_2-initialf.cpp:1021-1058): the per-particle Para_Soil /
// Para_Fluid copies of the reference are rows of the material tables, referenced by index.
static int soil_slot_of(int type, int matype, int etype) {
    if (type == 2) {
        if (matype < 30 && matype >= 0) return matype;
        if (matype >= 50) return matype % 10;
        return 30;
    }
    if (type == 7 && etype == 2 && matype >= 0 && matype < 30) return matype;
    return 30;  // default-constructed Para_Soil
}
static int fluid_slot_of(int type, int matype, int etype) {
    if ((type == 1 || (type == 7 && etype == 1)) && matype >= 0 && matype < 3) return matype;
    if (type == 3) return 3;
    return 4;  // default-constructed Para_Fluid
}

// Rain (flr 1) and excavation (fleob 1) set-up at upload: what the reference derives before its
// loop from the initial particles - x_min / x_max of the water of etype 7 (setting_par,
// CPP_2-initialf.cpp:1307-1311), rainini's period and velocity (CPP_15-1-rain.cpp:10-45; nrain /
// nhor is an integer division there), the cement base (center_cement, :133-163) - and the device
// counters. One realisation per context (the counters are per run).
static int upload_dynamic_sets(sphg_ctx *c) {
    Params &P = c->P;
    const sphg_problem &pro = c->pro;
    P.flr = pro.flr;
    P.fleob = pro.fleob;
    P.ecount = pro.fleob == 1 ? c->ecount : 0;
    P.rtype = 0;
    P.rain_lp = 0;
    if (pro.flr != 1 && pro.fleob != 1) return SPHG_OK;
    if (P.nreal != 1) return SPHG_E_UNSUPPORTED;
    if (pro.fleob == 1 && c->ecount < 1) return SPHG_E_ARG;
    int dyn[4] = {0, 1, 0, 0};
    if (pro.flr == 1) {
        if (!c->rain_set) return SPHG_E_ARG;
        const sphg_rain &r = c->rain;
        if (r.rtype < 1 || r.rtype > 7) return SPHG_E_ARG;  // the reference's error 6
        const int n = c->n, dim = pro.ndim;
        std::vector<double> xp(3 * (size_t)n);
        for (int k = 0; k < 3; k++)
            HIPCHK(hipMemcpy(xp.data() + (size_t)k * n, c->S.xp[c->cur][k], (size_t)n * 8, hipMemcpyDeviceToHost));
        double xmax = -10000.0, xmin = 10000.0, mx[3] = {-1000.0, -1000.0, -1000.0}, mn[3] = {1000.0, 1000.0, 1000.0};
        for (int i = 0; i < n; i++) {
            if (c->h_type[i] == 1 && c->h_etype[i] == 7) {
                xmax = fmax(xmax, xp[(size_t)(dim - 2) * n + i]);
                xmin = fmin(xmin, xp[(size_t)(dim - 2) * n + i]);
            }
            if (c->h_type[i] == 7)
                for (int k = 0; k < 3; k++) {
                    mx[k] = fmax(mx[k], xp[(size_t)k * n + i]);
                    mn[k] = fmin(mn[k], xp[(size_t)k * n + i]);
                }
        }
        const int nhor = r.nhor == 0 ? 35 : r.nhor;
        double temp = 0.0;
        P.vrain = 0.0;
        if (r.rtype == 1 || r.rtype == 4) {
            temp = r.nrain / nhor + 1.0;
            P.vrain = r.rtype == 1 ? r.drop_velocity : 2.0;
        } else if (r.rtype == 2 || r.rtype == 3) {
            temp = fabs(xmax - xmin) / pro.dr + 1.0;
        }
        if (r.rtype <= 4) {
            P.rain_lp = (int)((pro.loop - pro.inip) / temp);
            if (P.rain_lp < 1) return SPHG_E_ARG;  // the reference divides by zero
        }
        P.rtype = r.rtype;
        P.press_factor = r.press_factor;
        P.rx_min = xmin;
        P.rx_max = xmax;
        for (int k = 0; k < 3; k++) P.cement_base[k] = (mx[k] + mn[k]) * 0.5;
        P.dom_xmin = pro.xmin;
        P.dom_xmax = pro.xmax;
        dyn[0] = r.nrain;
    }
    HIPCHK(hipMemcpy(c->S.dyn, dyn, sizeof dyn, hipMemcpyHostToDevice));
    return SPHG_OK;
}

extern "C" int sphg_set_rain(sphg_ctx *c, const sphg_rain *r) {
    if (!c || !r) return SPHG_E_ARG;
    c->rain = *r;
    c->rain_set = true;
    c->uploaded = false;
    return SPHG_OK;
}

extern "C" int sphg_set_excavation(sphg_ctx *c, int ecount, const int *ecp) {
    if (!c || !ecp || ecount < 1) return SPHG_E_ARG;
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->S.ecp) {
        HIPCHK(hipFree(c->S.ecp));
        c->S.ecp = nullptr;
    }
    HIPCHK(hipMalloc(&c->S.ecp, (size_t)ecount * 3 * 4));
    HIPCHK(hipMemcpy(c->S.ecp, ecp, (size_t)ecount * 3 * 4, hipMemcpyHostToDevice));
    c->ecount = ecount;
    c->eob_len = 0;
    for (int s = 0; s < ecount; s++) c->eob_len = std::max(c->eob_len, ecp[3 * s + 1] - ecp[3 * s] + 1);
    c->uploaded = false;
    return SPHG_OK;
}

static int check_supported(const sphg_problem &p) {
    if (p.flknl < 1 || p.flknl > 4) return SPHG_E_UNSUPPORTED;  // 3 / 4: with kernel_gradient_correction
    if (p.flr < 0 || p.flr > 1 || p.fleob < 0 || p.fleob > 1 || p.flv < 0 || p.flv > 2) return SPHG_E_UNSUPPORTED;
    if (p.fli < 0 || p.fli > 3 || p.ss_type < 0 || p.ss_type > 3) return SPHG_E_UNSUPPORTED;
    if (p.flbndy != 1 && p.flbndy != 2 && p.flbndy != 3 && p.flbndy != 4 && p.flbndy != 5 && p.flbndy != 13 &&
        p.flbndy != 14)
        return SPHG_E_UNSUPPORTED;
    if (p.if_move < 0 || p.if_move > 3) return SPHG_E_UNSUPPORTED;
    if (p.step_regu <= 0) return SPHG_E_ARG;  // l % step_regu
    return SPHG_OK;
}

// True when some step l in [first, first + count) runs the regularisation passes (l % step_regu == 0).
static bool regu_due(int step_regu, long long first, int count) {
    if (count <= 0) return false;
    if (first <= 0) return true;
    long long next = ((first + step_regu - 1) / step_regu) * step_regu;
    return next <= first + count - 1;
}

static void drop_graphs(sphg_ctx *c) {
    if (c->graph[0] || c->graph[1]) hipStreamSynchronize(c->stream);  // no launched graph in flight
    for (int b = 0; b < 2; b++) {
        if (c->graph[b]) hipGraphExecDestroy(c->graph[b]);
        c->graph[b] = nullptr;
        c->graph_ok[b] = false;
    }
}

// Adds the regularisation passes to the launch sequence before a batch of steps that needs them
// (they stay in; the kernels test l % step_regu themselves).
static void ensure_regu(sphg_ctx *c, int count) {
    if (c->regu_on || !regu_due(c->P.step_regu, c->step, count)) return;
    c->regu_on = true;
    drop_graphs(c);
}

// The artificial-stress / damping pair pass runs while artificial stress is on or l <= inip.
static void plan_pre(sphg_ctx *c) {
    bool want = c->P.art_stre_coe > 0.0 || c->step <= c->P.inip;
    if (want == c->pre_on) return;
    c->pre_on = want;
    drop_graphs(c);
}

// Neighbour-mass table of the pair passes (Params::mass_uniform / mass_t): per particle type 0-4,
// whether every particle that is of that type, or can become it during the run, carries the same
// mass bit for bit. Excavation / backfill swaps type and etype, so a particle counts for both;
// rain gives reserve particles a water mass on the device, so water never qualifies with rain on.
// Read from the device arrays, so that it also follows sphg_set_field and sphg_load_state.
static int derive_mass_flags(sphg_ctx *c) {
    const int n = c->n;
    Params &P = c->P;
    std::vector<double> m(n);
    std::vector<int> ty(n), et(n);
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(m.data(), c->S.mass, (size_t)n * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(ty.data(), c->S.type, (size_t)n * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(et.data(), c->S.etype, (size_t)n * 4, hipMemcpyDeviceToHost));
    int seen = 0, mixed = 0;
    double mt[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    auto add = [&](int t, double v) {
        if (t < 0 || t > 4) return;
        if (!((seen >> t) & 1)) {
            mt[t] = v;
            seen |= 1 << t;
        } else if (memcmp(&v, &mt[t], 8) != 0) {
            mixed |= 1 << t;
        }
    };
    for (int i = 0; i < n; i++) {
        add(ty[i], m[i]);
        if (P.fleob) add(et[i], m[i]);
    }
    if (P.flr) mixed |= 1 << 1;
    const int mask = seen & ~mixed;
    bool same = mask == P.mass_uniform;
    for (int t = 0; t < 5; t++) {
        const double v = ((mask >> t) & 1) ? mt[t] : 0.0, vf = (double)(float)v;
        same = same && memcmp(&v, &P.mass_t.t[t], 8) == 0;
        P.mass_t.t[t] = v;
        P.mass_tf.t[t] = vf;
    }
    P.mass_uniform = mask;
    if (!same) drop_graphs(c);  // the captured launches hold the old parameters
    return SPHG_OK;
}

// Params::plane_zero: a 2D context whose out-of-plane components are all +0 (bit for bit) and
// stay so under every model of the reduced stress instantiation (elastic, Drucker-Prager, the
// zero tangent of model 4, the elastic structure), with no TS transform and no particle that can
// change type. Read from the device arrays; clears the flag the stress pass raises when a skipped
// component turns out not to be +0.
static int derive_plane_zero(sphg_ctx *c) {
    Params &P = c->P;
    int pz = P.ndim == 2 && !c->full_models && P.flts == 0 && P.fleob == 0 && P.flr == 0;
    if (pz) {
        int sh3 = 0;
        HIPCHK(hipMemcpy(&sh3, c->S.shear3, 4, hipMemcpyDeviceToHost));
        pz = sh3 == 0;
    }
    const size_t n = (size_t)c->n;
    std::vector<unsigned long long> b(n);
    double *const arrays[] = {c->S.sig[4], c->S.sig[5], c->S.eps[2], c->S.eps[4], c->S.eps[5], c->S.adps[4],
                              c->S.adps[5], c->S.deps[2], c->S.deps[4], c->S.deps[5], c->S.weps[1], c->S.weps[2]};
    for (double *a : arrays) {
        if (!pz) break;
        HIPCHK(hipMemcpy(b.data(), a, n * 8, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n && pz; i++) pz = b[i] == 0ull;
    }
    const int zero = 0;
    HIPCHK(hipMemcpy(c->S.plane_dirty, &zero, 4, hipMemcpyHostToDevice));
    if (pz != P.plane_zero) drop_graphs(c);
    P.plane_zero = pz;
    return SPHG_OK;
}

extern "C" int sphg_upload_done(sphg_ctx *c, int n) {
    if (!c || n != c->n) return SPHG_E_ARG;
    int rc = check_supported(c->pro);
    if (rc) return rc;
    Params &P = c->P;
    const sphg_problem &pro = c->pro;
    P.n = n;
    P.cap = c->cap;
    P.ndim = pro.ndim;
    P.flbndy = pro.flbndy;
    P.flknl = pro.flknl;
    P.fli = pro.fli;
    P.if_move = pro.if_move;
    P.flv = pro.flv;
    P.inip = pro.inip;
    P.ss_type = pro.ss_type;
    P.ss_myu = pro.ss_myu;
    P.coupled = (pro.ss_type >= 1 && pro.ss_type <= 3 && pro.ss_acc_type == 2) ? 1 : 0;
    P.ss_zeta = pro.ss_zeta;
    P.flts = pro.flts;
    P.alphats = pro.alphats;
    P.dt = pro.dt;
    P.dr = pro.dr;
    P.gx = pro.gx; P.gy = pro.gy; P.gz = pro.gz;
    P.gacc = sqrt(pro.gx * pro.gx + pro.gy * pro.gy + pro.gz * pro.gz);
    for (int k = 0; k < 3; k++) {
        P.xsph[k] = pro.xsph[k];
        P.art_vis[k][0] = pro.art_vis[k][0];
        P.art_vis[k][1] = pro.art_vis[k][1];
    }
    P.coe_bndy = pro.coe_bndy;
    P.interc = pro.interc;
    P.permtype = pro.permtype;
    P.pers = pro.pers; P.perl = pro.perl; P.perm = pro.perm;
    P.flagum = pro.flagum;
    P.step_regu = pro.step_regu;
    P.art_stre_coe = pro.art_stre_coe;
    for (int k = 0; k < 3; k++) P.move_vx[k] = pro.move_vx[k];
    P.tmin = pro.threshold_min;
    P.tmax = pro.threshold_max;
    P.struct_e = c->mat.structure.e;
    P.struct_poi = c->mat.structure.poi;
    P.eig_in_step = (pro.flts >= 1 || pro.art_stre_coe > 0.0) ? 1 : 0;
    P.lazy_deq = (!P.eig_in_step && pro.fleob == 0) ? 1 : 0;
    P.mrv_ok = (pro.ss_type == 0 && pro.flbndy != 5 && pro.flv == 0) ? 1 : 0;
    if (pro.flv > 0 && !c->S.vib_loads) return SPHG_E_ARG;  // the reference stops when vibra.dat is missing
    P.vib_ttime = c->vib_ttime;
    P.vib_sttime = c->vib_sttime;
    P.vib_edtime = c->vib_edtime;
    P.vib_nsteps = c->vib_nsteps;
    if ((rc = upload_dynamic_sets(c))) return rc;
    P.water_acc_minor = c->mat.water[0].acc_type != 0;
    P.air_acc_minor = c->mat.air.acc_type != 0;
    // realisation of every particle (batch contexts: realisation-major ranges of counts[r]) and
    // each realisation's first type-0 particle with matype > 2 (check_domain's wall velocity)
    {
        const int nr = P.nreal;
        std::vector<int> rid(n, 0), first(nr + 1, 0);
        if (!c->counts.empty()) {
            long long tot = 0;
            for (int r = 0; r < nr; r++) {
                first[r] = (int)tot;
                tot += c->counts[r];
            }
            if (tot != n) return SPHG_E_ARG;
        }
        first[nr] = n;
        for (int r = 0; r < nr; r++) {
            Grid &g = c->grids[r];
            g.first_moving_bndy = n;
            for (int i = first[r]; i < first[r + 1]; i++) {
                rid[i] = r;
                if (g.first_moving_bndy == n && c->h_type[i] == 0 && c->h_matype[i] > 2) g.first_moving_bndy = i;
            }
        }
        HIPCHK(hipMemcpy(c->d_rid, rid.data(), (size_t)n * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->d_grids, c->grids.data(), sizeof(Grid) * nr, hipMemcpyHostToDevice));
        // the search walks particles in the binning's cell order; until the first binning that
        // order is the identity, so a search issued before any binning still visits every particle once
        for (int i = 0; i < n; i++) rid[i] = i;
        HIPCHK(hipMemcpy(c->S.sorted, rid.data(), (size_t)n * 4, hipMemcpyHostToDevice));
    }
    c->full_models = false;
    for (int i = 0; i < n; i++) {
        int m = c->h_matype[i], mt = ((m >= 0 && m < 30) || (m >= 50 && m < 60)) ? m % 10 : -1;
        const bool soil = c->h_type[i] == 2 || (pro.fleob == 1 && c->h_etype[i] == 2);  // backfill
        if (soil && (mt == 1 || mt == 3 || mt == 8)) c->full_models = true;
    }
    // material slots with host-evaluated transcendental constants
    std::vector<SoilSlot> soil(kNSlots);
    std::vector<sphg_soil> soil_src(kNSlots);
    for (int s = 0; s < kNSlots; s++) {
        sphg_soil d{};
        if (s < SPHG_NSOIL) d = c->mat.soil[s];
        else { d.fs = 1.0; d.ds = 1.0; }
        soil_src[s] = d;
        SoilSlot &o = soil[s];
        o.e = d.e; o.poi = d.poi; o.eps0 = d.eps0; o.dens = d.dens; o.kapa = d.kapa; o.zmf = d.zmf; o.ann = d.ann;
        o.ramda = d.ramda; o.fs = d.fs; o.c = d.c; o.sme0 = d.sme0; o.porosity = d.porosity; o.vpu = d.vpu;
        o.cop = d.cop; o.sitas = d.sitas; o.sitar = d.sitar; o.damp_coe = d.damp_coe; o.beta = d.beta; o.ds = d.ds;
        o.bnn = d.bnn; o.epsr = d.epsr; o.epss = d.epss; o.sd = d.sd; o.sw = d.sw; o.c1 = d.c1; o.c2 = d.c2;
        o.g0 = d.g0; o.etao = d.etao; o.zet = d.zet; o.damp1 = d.damp1; o.damp2 = d.damp2;
        o.exp_c2sw = exp(d.c2 * d.sw);  // CPP_7-2-constitutive.cpp:333-336
        o.exp_c1sd = exp(d.c1 * d.sd);
        // dilatancy angle of dpmodel_bui (CPP_8-1-consfactor.cpp:462-463, :511)
        double fai_dil = atan(tan(d.beta * 0.0174532925166667) / d.fs);
        o.dp_yalf_dil = sin(fai_dil) / sqrt(9.0 + 3.0 * sin(fai_dil) * sin(fai_dil));
    }
    std::vector<FluidSlot> fl(8);
    std::vector<sphg_fluid> fluid_src(8);
    for (int s = 0; s < 8; s++) {
        sphg_fluid d{};
        d.strain_min = 1.0;
        if (s < 3) d = c->mat.water[s];
        else if (s == 3) d = c->mat.air;
        fluid_src[s] = d;
        FluidSlot &o = fl[s];
        o.vis = d.vis; o.vpu = d.vpu; o.dens = d.dens; o.porosity = d.porosity; o.c = d.c; o.pre_coe = d.pre_coe;
        o.strain_min = d.strain_min;
        o.tan_fai = tan(d.fai * 0.0174532925199);  // CPP_7-2-constitutive.cpp:42-43
        o.bhtype = d.bhtype;
        o.acc_type = d.acc_type;
    }
    HIPCHK(hipMemcpy(c->d_soil, soil.data(), sizeof(SoilSlot) * kNSlots, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_fluid, fl.data(), sizeof(FluidSlot) * 8, hipMemcpyHostToDevice));
    std::vector<int8_t> ss(n), fs(n);
    for (int i = 0; i < n; i++) {
        ss[i] = (int8_t)soil_slot_of(c->h_type[i], c->h_matype[i], c->h_etype[i]);
        fs[i] = (int8_t)fluid_slot_of(c->h_type[i], c->h_matype[i], c->h_etype[i]);
    }
    if ((rc = check_pinned_materials(c, soil_src, fluid_src, ss, fs))) return rc;
    HIPCHK(hipMemcpy(c->S.soil_slot, ss.data(), n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->S.fluid_slot, fs.data(), n, hipMemcpyHostToDevice));
    // per-particle fluid density (drag partner), and the uniformity flags the pair passes use
    std::vector<double> fd(n), hl(n);
    P.fluid_por_uniform = 1;
    double por0 = 0.0;
    bool seen = false;
    for (int i = 0; i < n; i++) {
        fd[i] = fl[fs[i]].dens;
        int t = c->h_type[i];
        // reserve particles rain or backfill may turn into water count as water
        const bool dyn_water = (pro.flr == 1 || pro.fleob == 1) && t == 7 && c->h_etype[i] == 1;
        if (t == 1 || t == 3 || dyn_water) {
            double pr = fl[fs[i]].porosity;
            if (!seen) { por0 = pr; seen = true; }
            else if (memcmp(&pr, &por0, 8) != 0) P.fluid_por_uniform = 0;
        }
    }
    HIPCHK(hipMemcpy(c->S.fdens, fd.data(), (size_t)n * 8, hipMemcpyHostToDevice));
    // dry context: no water or air particle now, and none can appear (rain or backfill)
    bool fluid = false;
    for (int i = 0; i < n && !fluid; i++) fluid = c->h_type[i] == 1 || c->h_type[i] == 3;
    P.dry = (!fluid && pro.flr == 0 && pro.fleob == 0) ? 1 : 0;
    {
        // per-particle material values pParti_ConsPara[i].ds / .cop / .fai: the soil slot's unless
        // the caller set them (random fields), and the Drucker-Prager constants of each particle's
        // friction angle, evaluated here with the host's libm as the slot constants are
        std::vector<double> v(n), fai(n);
        auto put = [&](double *d, const double *h) { return hipMemcpy(d, h, (size_t)n * 8, hipMemcpyHostToDevice); };
        if (!c->has_psds) {
            for (int i = 0; i < n; i++) v[i] = soil[ss[i]].ds;
            HIPCHK(put(c->S.ps_ds, v.data()));
        }
        if (!c->has_pscop) {
            for (int i = 0; i < n; i++) v[i] = soil[ss[i]].cop;
            HIPCHK(put(c->S.ps_cop, v.data()));
        }
        if (!c->has_psfai) {
            for (int i = 0; i < n; i++) fai[i] = ss[i] < SPHG_NSOIL ? c->mat.soil[ss[i]].fai : 0.0;
            HIPCHK(put(c->S.ps_fai, fai.data()));
        }
        // Para_Fluid fai / c of each particle (Bingham water): its fluid slot's unless set
        if (!c->has_pffai) {
            for (int i = 0; i < n; i++) v[i] = fluid_src[fs[i]].fai;
            HIPCHK(put(c->S.pf_fai, v.data()));
        }
        if (!c->has_pfc) {
            for (int i = 0; i < n; i++) v[i] = fluid_src[fs[i]].c;
            HIPCHK(put(c->S.pf_c, v.data()));
        }
        if ((rc = derive_strength_constants(c))) return rc;
    }
    HIPCHK(hipMemcpy(hl.data(), c->S.hl, (size_t)n * 8, hipMemcpyDeviceToHost));
    P.hl_uniform = 1;
    for (int i = 1; i < n; i++)
        if (memcmp(&hl[i], &hl[0], 8) != 0) P.hl_uniform = 0;
    // the pair passes form the kernel normalisation once per particle from a common smoothing
    // length; setting_par gives every particle hl = dr and the step never changes it
    if (!P.hl_uniform) return SPHG_E_UNSUPPORTED;
    // out-of-plane shear stresses of the uploaded state (2D); the stress pass keeps the flag after
    int sh3 = 0;
    if (P.ndim == 2) {
        for (int k = 4; k < 6 && !sh3; k++) {
            HIPCHK(hipMemcpy(hl.data(), c->S.sig[k], (size_t)n * 8, hipMemcpyDeviceToHost));
            for (int i = 0; i < n && !sh3; i++) sh3 = hl[i] != 0.0;
        }
    }
    HIPCHK(hipMemcpy(c->S.shear3, &sh3, 4, hipMemcpyHostToDevice));
    if ((rc = derive_mass_flags(c))) return rc;
    if ((rc = derive_plane_zero(c))) return rc;
    c->S.ps_c = c->has_psc ? c->d_psc : nullptr;
    int zero = 0;
    HIPCHK(hipMemcpy(c->S.err, &zero, 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->S.step, &c->step, 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->S.dtp, &pro.dt, 8, hipMemcpyHostToDevice));
    if (c->S.mra) {  // mixed form: the packed records from the uploaded state
        hipLaunchKernelGGL(k_mrec_init, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, c->S, c->P, c->cur);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    drop_graphs(c);
    c->regu_on = false;
    c->pre_on = false;
    c->uploaded = true;
    return SPHG_OK;
}

// -------------------------------------------------------------------------------- launches
static void tbegin(sphg_ctx *c, const char *name, Timing &t) {
    if (!c->timing) return;
    t.name = name;
    hipEventCreate(&t.a);
    hipEventCreate(&t.b);
    hipEventRecord(t.a, c->stream);
}
static void tend(sphg_ctx *c, Timing &t) {
    if (!c->timing) return;
    hipEventRecord(t.b, c->stream);
    c->tq.push_back(t);
}

#define LAUNCH(name, kern, grid, ...)                                          \
    do {                                                                       \
        Timing t_;                                                             \
        tbegin(c, name, t_);                                                   \
        hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), 0, c->stream, __VA_ARGS__); \
        tend(c, t_);                                                           \
    } while (0)

// widths each tunable pass is built for (index 0 density, 1 normalisation, 2 strain, 3 XSPH)
static const int kBwChoices[4][3] = {{1, 2, 4}, {2, 4, 6}, {1, 2, 4}, {1, 2, 4}};

// launch of a pass instantiated for the context's width
#define LAUNCH_BW(name, k, slot, W0, W1, W2, grid, ...)                                      \
    do {                                                                                    \
        if (c->bw[slot] == W0) LAUNCH(name, (k<NDIM, M, W0>), grid, __VA_ARGS__);           \
        else if (c->bw[slot] == W1) LAUNCH(name, (k<NDIM, M, W1>), grid, __VA_ARGS__);      \
        else LAUNCH(name, (k<NDIM, M, W2>), grid, __VA_ARGS__);                             \
    } while (0)

template <int NDIM>
static void enqueue_search(sphg_ctx *c) {
    const int cur = c->cur, n = c->n;
    // the search appends the near-wall particles of the split momentum pass to a list it starts empty
    if (c->S.nw_count) LAUNCH("nw_reset", k_fill_int, 1, c->S.nw_count, 1, 0);
    if (c->search_stage == 16) LAUNCH("search", (k_search<NDIM, 16>), grid_for(n), c->S, c->P, cur);
    else LAUNCH("search", (k_search<NDIM, 32>), grid_for(n), c->S, c->P, cur);
}

template <int NDIM>
static void enqueue_nnps(sphg_ctx *c) {
    const int cur = c->cur, n = c->n, M = c->P.ncell + 2;
    LAUNCH("fill_counts", k_fill_int, grid_for(M), c->S.cell_count, M, 0);
    LAUNCH("bin", k_bin<NDIM>, grid_for(n), c->S, c->P, cur);
    LAUNCH("scan_tiles", k_scan_tiles, c->scan_blocks, c->S.cell_count, c->S.cell_start, c->S.block_sums, M);
    {
        Timing t_;
        tbegin(c, "scan_top", t_);
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanBlock), 0, c->stream, c->S.block_sums, c->scan_blocks);
        tend(c, t_);
    }
    LAUNCH("scan_add", k_scan_add, c->scan_blocks, c->S.cell_start, c->S.block_sums, M);
    LAUNCH("scatter", k_scatter, grid_for(n), c->S, c->P);
    LAUNCH("cell_order", k_cell_order, grid_for(M), c->S, M);
    enqueue_search<NDIM>(c);
    if (c->rev_lists) LAUNCH("reverse_lists", k_reverse_lists, grid_for(n), c->S, c->P);
    if (c->P.flknl > 2) LAUNCH("kgc", k_kgc<NDIM>, grid_for(n), c->S, c->P, cur);
}

template <int NDIM, class M>
static void enqueue_step(sphg_ctx *c, int phase_mask) {
    const int cur = c->cur, n = c->n;
    if (phase_mask & 1) {
        if (c->P.dt_mode == SPHG_DT_CFL) {
            LAUNCH("dt_reduce", k_dt_reduce<NDIM>, grid_for(n), c->S, c->P);
            LAUNCH("dt_set", k_dt_set, 1, c->S, c->P);
        }
        if (c->P.flv > 0 || c->P.dt_mode == SPHG_DT_CFL) LAUNCH("time", k_time_step, 1, c->S, c->P);
        if (c->P.flr == 1) {
            LAUNCH("rain", k_rain<NDIM>, 1, c->S, c->P, cur);
            LAUNCH("waterdel", k_waterdel<NDIM>, grid_for(n), c->S, c->P, cur);
        }
        if (c->P.fleob == 1 && c->eob_len > 0) {
            LAUNCH("eob", k_eob, grid_for(c->eob_len), c->S, c->P);
            LAUNCH("eob_next", k_eob_next, 1, c->S, c->P);
        }
        enqueue_nnps<NDIM>(c);
    }
    if (phase_mask & 2) {
        LAUNCH_BW("density", k_density, 0, 1, 2, 4, grid_for(n), c->S, c->P, cur);
        LAUNCH_BW("norm_eos", k_norm, 1, 2, 4, 6, grid_for(n), c->S, c->P, cur);
    }
    if (phase_mask & 4) {
        if (c->P.flagum == 1) LAUNCH("saturation", (k_saturation<NDIM, M>), grid_for(n), c->S, c->P, cur);
        LAUNCH_BW("strain", k_strain, 2, 1, 2, 4, grid_for(n), c->S, c->P, cur);
        if (c->regu_on) {
            LAUNCH("regu_copy", k_regu_copy, grid_for(n), c->S, c->P, 0);
            LAUNCH("regu", (k_regu<NDIM, M>), grid_for(n), c->S, c->P, cur, 0);
        }
    }
    if (phase_mask & 8) {
        auto ks = c->P.flts >= 1 ? (c->full_models ? k_stress<3> : k_stress<2>)
                                 : (c->full_models ? k_stress<1> : k_stress<0>);
        if (c->regu_on) {
            LAUNCH("stress", ks, grid_for(n), c->S, c->P, 1);
            LAUNCH("regu_copy", k_regu_copy, grid_for(n), c->S, c->P, 1);
            LAUNCH("regu", (k_regu<NDIM, M>), grid_for(n), c->S, c->P, cur, 1);
            LAUNCH("stress_post", ks, grid_for(n), c->S, c->P, 2);
        } else {
            LAUNCH("stress", ks, grid_for(n), c->S, c->P, 0);
        }
        if (c->P.eig_in_step) LAUNCH("stress_eig", k_stress_eig, grid_for(n), c->S, c->P);
    }
    if (phase_mask & 16) {
        if (c->pre_on) LAUNCH("accel_pre", (k_accel_pre<NDIM, M>), grid_for(n), c->S, c->P, cur);
        if (c->P.coupled) LAUNCH("coupled", (k_coupled<NDIM, M>), grid_for(n), c->S, c->P, cur);
        if (c->P.fli >= 2) LAUNCH("drag", (k_drag<NDIM, M>), grid_for(n), c->S, c->P, cur);
        if (c->P.flv > 0) LAUNCH("vib", (k_vib<NDIM, M>), grid_for(n), c->S, c->P, cur);
        if constexpr (kSplitBrow<NDIM, M>) LAUNCH("bndy_rows", (k_bndy_rows<NDIM, M>), grid_for(n), c->S, c->P, cur);
        LAUNCH("accel", (k_accel<NDIM, M>), grid_for(n), c->S, c->P, cur);
        if (c->P.flv == 1) LAUNCH("vib_walls_acc", k_vib_walls_acc<NDIM>, grid_for(n), c->S, c->P);
        if (c->P.ss_type != 0) {
            LAUNCH("ssi_soil", k_ssi_soil<NDIM>, grid_for(n), c->S, c->P, cur);
            LAUNCH("ssi_struct", k_ssi_struct, grid_for(n), c->S, c->P);
        }
        if (c->P.ss_type != 0 || c->P.flbndy == 5 || c->P.flv > 0)
            LAUNCH("accel_tail", k_accel_tail<NDIM>, grid_for(n), c->S, c->P, cur);
        if (c->P.flv > 0) LAUNCH("vib_walls_vel", k_vib_walls_vel<NDIM>, grid_for(n), c->S, c->P, cur);
    }
    if (phase_mask & 32) {
        LAUNCH_BW("xsph", k_xsph, 3, 1, 2, 4, grid_for(n), c->S, c->P, cur);
        if (tran_bndy(c->P)) LAUNCH("slip_tran", (k_slip_tran<NDIM, M>), grid_for(n), c->S, c->P, cur);
    }
    if (phase_mask & 64) {
        if (c->P.flbndy == 2) LAUNCH("wall_vb", (k_wall_vb<NDIM, M>), grid_for(n), c->S, c->P, cur);
        LAUNCH("update", (k_update<NDIM, M>), grid_for(n), c->S, c->P, cur);
        if (c->P.flr == 1) LAUNCH("partidel", k_partidel, grid_for(n), c->S, c->P, cur ^ 1);
        LAUNCH("step_inc", k_step_inc, 1, c->S);
    }
}

static void enqueue(sphg_ctx *c, int mask) {
    const bool mixed = c->precision == SPHG_MIXED;
    if (c->P.flknl > 2) {  // passes with the kernel gradient correction
        if (c->P.ndim == 2) mixed ? enqueue_step<2, FastKgc>(c, mask) : enqueue_step<2, ExactKgc>(c, mask);
        else mixed ? enqueue_step<3, FastKgc>(c, mask) : enqueue_step<3, ExactKgc>(c, mask);
        return;
    }
    if (c->P.ndim == 2) mixed ? enqueue_step<2, Fast>(c, mask) : enqueue_step<2, Exact>(c, mask);
    else mixed ? enqueue_step<3, Fast>(c, mask) : enqueue_step<3, Exact>(c, mask);
}

static int read_err(sphg_ctx *c) {
    int e = 0;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(&e, c->S.err, 4, hipMemcpyDeviceToHost));
    if (e & SPHG_ERR_CELL) return SPHG_E_CELLRANGE;
    if (e & SPHG_ERR_NB) return SPHG_E_NEIGHBOURS;
    if (e & 4) return SPHG_E_UNSUPPORTED;
    return SPHG_OK;
}

static int finish_timing(sphg_ctx *c) {
    if (!c->timing || c->tq.empty()) return 0;
    HIPCHK(hipStreamSynchronize(c->stream));
    for (auto &t : c->tq) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, t.a, t.b);
        bool found = false;
        for (auto &s : c->tsum)
            if (s.first == t.name) {
                s.second.first += ms;
                s.second.second += 1;
                found = true;
            }
        if (!found) c->tsum.push_back({t.name, {ms, 1}});
        hipEventDestroy(t.a);
        hipEventDestroy(t.b);
    }
    c->tq.clear();
    return 0;
}

// Individual phases (each advances nothing but its own outputs; sphg_update advances l).
static int phase(sphg_ctx *c, int mask) {
    if (!c || !c->uploaded) return SPHG_E_ARG;
    ensure_regu(c, 1);
    plan_pre(c);
    enqueue(c, mask);
    if ((mask & 8) && !c->P.eig_in_step) c->eig_stale = true;
    if (mask & 64) {
        c->cur ^= 1;
        c->step++;
        c->stepped = true;
    }
    finish_timing(c);
    return read_err(c);
}

extern "C" int sphg_parttocell(sphg_ctx *c) {
    if (!c || !c->uploaded) return SPHG_E_ARG;
    // binning only: the search is issued by sphg_partisearching
    const int n = c->n, M = c->P.ncell + 2, cur = c->cur;
    auto go = [&](auto ndim_tag) {
        constexpr int NDIM = decltype(ndim_tag)::value;
        LAUNCH("fill_counts", k_fill_int, grid_for(M), c->S.cell_count, M, 0);
        LAUNCH("bin", k_bin<NDIM>, grid_for(n), c->S, c->P, cur);
        LAUNCH("scan_tiles", k_scan_tiles, c->scan_blocks, c->S.cell_count, c->S.cell_start, c->S.block_sums, M);
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanBlock), 0, c->stream, c->S.block_sums, c->scan_blocks);
        LAUNCH("scan_add", k_scan_add, c->scan_blocks, c->S.cell_start, c->S.block_sums, M);
        LAUNCH("scatter", k_scatter, grid_for(n), c->S, c->P);
        LAUNCH("cell_order", k_cell_order, grid_for(M), c->S, M);
    };
    if (c->P.ndim == 2) go(std::integral_constant<int, 2>());
    else go(std::integral_constant<int, 3>());
    finish_timing(c);
    return read_err(c);
}
extern "C" int sphg_partisearching(sphg_ctx *c) {
    if (!c || !c->uploaded) return SPHG_E_ARG;
    if (c->P.ndim == 2) enqueue_search<2>(c);
    else enqueue_search<3>(c);
    finish_timing(c);
    return read_err(c);
}
extern "C" int sphg_kernel_gradient_correction(sphg_ctx *c) {
    if (!c || !c->uploaded) return SPHG_E_ARG;
    if (c->P.flknl <= 2) return SPHG_OK;  // the reference calls it only for flknl > 2
    const int n = c->n, cur = c->cur;
    if (c->P.ndim == 2) LAUNCH("kgc", k_kgc<2>, grid_for(n), c->S, c->P, cur);
    else LAUNCH("kgc", k_kgc<3>, grid_for(n), c->S, c->P, cur);
    finish_timing(c);
    return read_err(c);
}
extern "C" int sphg_density(sphg_ctx *c) { return phase(c, 2); }
extern "C" int sphg_strain(sphg_ctx *c) { return phase(c, 4); }
extern "C" int sphg_strain_to_stress(sphg_ctx *c) { return phase(c, 8); }
extern "C" int sphg_accel(sphg_ctx *c) { return phase(c, 16); }
extern "C" int sphg_xsph(sphg_ctx *c) { return phase(c, 32); }
extern "C" int sphg_update(sphg_ctx *c) { return phase(c, 64); }

static int capture(sphg_ctx *c, int parity) {
    hipGraph_t g = nullptr;
    int saved = c->cur;
    c->cur = parity;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    enqueue(c, 127);
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    c->cur = saved;
    if (e != hipSuccess) return SPHG_E_HIP;
    e = hipGraphInstantiate(&c->graph[parity], g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return SPHG_E_HIP;
    c->graph_ok[parity] = true;
    return SPHG_OK;
}

extern "C" int sphg_step(sphg_ctx *c, int nsteps) {
    if (!c || !c->uploaded || nsteps < 0) return SPHG_E_ARG;
    ensure_regu(c, nsteps);
    if (c->timing) {
        for (int s = 0; s < nsteps; s++) {
            plan_pre(c);
            enqueue(c, 127);
            c->cur ^= 1;
            c->step++;
        }
        c->stepped = c->stepped || nsteps > 0;
        if (nsteps > 0 && !c->P.eig_in_step) c->eig_stale = true;
        finish_timing(c);
        return read_err(c);
    }
    for (int s = 0; s < nsteps; s++) {
        int par = c->cur;
        plan_pre(c);
        if (!c->graph_ok[par]) {
            int rc = capture(c, par);
            if (rc) return rc;
        }
        HIPCHK(hipGraphLaunch(c->graph[par], c->stream));
        c->cur ^= 1;
        c->step++;
    }
    c->stepped = c->stepped || nsteps > 0;
    if (nsteps > 0 && !c->P.eig_in_step) c->eig_stale = true;
    return read_err(c);
}

// The time loop with the CFL time step: dt is only known on the device once a step has started,
// so the loop takes one step at a time and reads the device clock Para_Pro::t after each (the
// output steps and the t < tt test are those of sphg_run).
static int run_cfl(sphg_ctx *c, int loop, int fop, double tt, double *t, sphg_output_fn out, void *user) {
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(c->S.tnow, t, 8, hipMemcpyHostToDevice));
    int l = c->step;
    while (l <= loop) {
        if (!(*t < tt)) {
            if (out) {
                int r = out(user, l, *t);
                if (r) return r;
            }
            break;
        }
        int rc = sphg_step(c, 1);
        HIPCHK(hipMemcpy(t, c->S.tnow, 8, hipMemcpyDeviceToHost));
        if (rc) return rc;
        const int last = l++;
        if (out && (last % fop == 0 || last == 1)) {
            int r = out(user, last, *t);
            if (r) return r;
        }
    }
    return SPHG_OK;
}

extern "C" int sphg_set_dt_mode(sphg_ctx *c, int mode, double c_cour, double c_wave, double dt_max) {
    if (!c) return SPHG_E_ARG;
    if (mode == SPHG_DT_CONSTANT) {
        if (c->P.dt_mode != SPHG_DT_CONSTANT) {
            drop_graphs(c);
            c->P.dt_mode = SPHG_DT_CONSTANT;
        }
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipMemcpy(c->S.dtp, &c->pro.dt, 8, hipMemcpyHostToDevice));
        return SPHG_OK;
    }
    if (mode != SPHG_DT_CFL || !(c_cour > 0.0) || !(c_wave > 0.0) || !(dt_max > 0.0) || !std::isfinite(c_cour) ||
        !std::isfinite(c_wave))
        return SPHG_E_ARG;
    drop_graphs(c);
    c->P.dt_mode = SPHG_DT_CFL;
    c->P.dt_cour = c_cour;
    c->P.dt_cwave = c_wave;
    c->P.dt_max = dt_max;
    return SPHG_OK;
}

extern "C" int sphg_get_dt(sphg_ctx *c, double *dt) {
    if (!c || !dt) return SPHG_E_ARG;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(dt, c->S.dtp, 8, hipMemcpyDeviceToHost));
    return SPHG_OK;
}

// Serving: steps are enqueued on the context's stream without waiting, so several contexts'
// work overlaps on the device; sphg_poll reports whether the stream has drained.
extern "C" int sphg_run(sphg_ctx *c, int loop, int fop, double tt, double *t, sphg_output_fn out, void *user) {
    if (!c || !c->uploaded || !t || fop <= 0) return SPHG_E_ARG;
    if (c->P.dt_mode == SPHG_DT_CFL) return run_cfl(c, loop, fop, tt, t, out, user);
    const double dt = c->P.dt;
    int l = c->step;
    // the device clock the vibration loads read starts from the caller's t and then advances
    // step by step exactly as *t does here
    if (c->P.flv > 0) {
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipMemcpy(c->S.tnow, t, 8, hipMemcpyHostToDevice));
    }
    while (l <= loop) {
        if (!(*t < tt)) {  // the reference's else branch: output and leave the loop
            if (out) {
                int r = out(user, l, *t);
                if (r) return r;
            }
            break;
        }
        // steps l .. next, next = the first output step at or after l, all with t < tt at their start
        int next = l == 1 ? 1 : ((l + fop - 1) / fop) * fop;
        if (next > loop) next = loop;
        int k = 0;
        double tk = *t;
        while (l + k <= next && tk < tt) {
            tk = tk + dt;
            k++;
        }
        int rc = sphg_step(c, k);
        *t = tk;
        l += k;
        if (rc) return rc;
        const int last = l - 1;
        if (out && (last % fop == 0 || last == 1)) {
            int r = out(user, last, *t);
            if (r) return r;
        }
    }
    return SPHG_OK;
}

extern "C" int sphg_step_async(sphg_ctx *c, int nsteps) {
    if (!c || !c->uploaded || nsteps < 0 || c->timing) return SPHG_E_ARG;
    ensure_regu(c, nsteps);
    for (int s = 0; s < nsteps; s++) {
        int par = c->cur;
        plan_pre(c);
        if (!c->graph_ok[par]) {
            int rc = capture(c, par);
            if (rc) return rc;
        }
        HIPCHK(hipGraphLaunch(c->graph[par], c->stream));
        c->cur ^= 1;
        c->step++;
    }
    c->stepped = c->stepped || nsteps > 0;
    if (nsteps > 0 && !c->P.eig_in_step) c->eig_stale = true;
    return SPHG_OK;
}

extern "C" int sphg_poll(sphg_ctx *c) {
    if (!c) return SPHG_E_ARG;
    hipError_t q = hipStreamQuery(c->stream);
    if (q == hipErrorNotReady) return SPHG_BUSY;
    if (q != hipSuccess) return SPHG_E_HIP;
    return read_err(c);
}

extern "C" int sphg_sync(sphg_ctx *c) {
    if (!c) return SPHG_E_ARG;
    return read_err(c);
}

extern "C" int sphg_current_step(const sphg_ctx *c) { return c ? c->step : -1; }
extern "C" int sphg_particle_count(const sphg_ctx *c) { return c ? c->n : -1; }

// ------------------------------------------------------------------------------- checkpoints
// File: the header below, then the device pool byte for byte with the neighbour-list region left
// out. Everything a step reads is in the pool (state, material slots, grids, per-particle
// cohesion, sticky error word, device step counter); the host keeps only the step counter, the
// buffer parity and two flags, which the header carries.
namespace {
struct CkptHeader {
    char magic[8];
    int32_t version, n, cap, ndim, precision, nreal, step, cur, stepped, has_psc;
    uint64_t pool_bytes, skip_off, skip_len;
};
constexpr char kCkptMagic[8] = {'S', 'P', 'H', 'G', 'C', 'K', 'P', 'T'};
constexpr size_t kCkptChunk = size_t(64) << 20;

CkptHeader ckpt_header(const sphg_ctx *c) {
    CkptHeader h{};
    memcpy(h.magic, kCkptMagic, 8);
    h.version = 1;
    h.n = c->n; h.cap = c->cap; h.ndim = c->P.ndim; h.precision = c->precision; h.nreal = c->P.nreal;
    h.step = c->step; h.cur = c->cur; h.stepped = c->stepped; h.has_psc = c->has_psc;
    h.pool_bytes = c->pool_bytes;
    h.skip_off = (uint64_t)((char *)c->S.nbr - c->pool);
    h.skip_len = (uint64_t)c->cap * kMaxNb * 4;
    return h;
}

// copies pool[off, off+len) to or from the file through a bounded host buffer
bool ckpt_stream(sphg_ctx *c, FILE *f, size_t off, size_t len, bool save) {
    std::vector<char> buf(std::min(len, kCkptChunk));
    for (size_t done = 0; done < len;) {
        const size_t k = std::min(kCkptChunk, len - done);
        if (save) {
            if (hipMemcpy(buf.data(), c->pool + off + done, k, hipMemcpyDeviceToHost) != hipSuccess) return false;
            if (fwrite(buf.data(), 1, k, f) != k) return false;
        } else {
            if (fread(buf.data(), 1, k, f) != k) return false;
            if (hipMemcpy(c->pool + off + done, buf.data(), k, hipMemcpyHostToDevice) != hipSuccess) return false;
        }
        done += k;
    }
    return true;
}
}  // namespace

extern "C" int sphg_save_state(sphg_ctx *c, const char *path) {
    if (c && ensure_eig(c)) return SPHG_E_HIP;
    if (!c || !path || !c->uploaded) return SPHG_E_ARG;
    HIPCHK(hipStreamSynchronize(c->stream));
    const CkptHeader h = ckpt_header(c);
    FILE *f = fopen(path, "wb");
    if (!f) return SPHG_E_ARG;
    bool ok = fwrite(&h, sizeof h, 1, f) == 1 && ckpt_stream(c, f, 0, h.skip_off, true) &&
              ckpt_stream(c, f, h.skip_off + h.skip_len, h.pool_bytes - h.skip_off - h.skip_len, true);
    ok = (fclose(f) == 0) && ok;
    return ok ? SPHG_OK : SPHG_E_ARG;
}

extern "C" int sphg_load_state(sphg_ctx *c, const char *path) {
    if (!c || !path || !c->uploaded) return SPHG_E_ARG;
    FILE *f = fopen(path, "rb");
    if (!f) return SPHG_E_ARG;
    CkptHeader h{};
    const CkptHeader mine = ckpt_header(c);
    if (fread(&h, sizeof h, 1, f) != 1 || memcmp(h.magic, kCkptMagic, 8) != 0 || h.version != 1 ||
        h.n != mine.n || h.cap != mine.cap || h.ndim != mine.ndim || h.precision != mine.precision ||
        h.nreal != mine.nreal || h.pool_bytes != mine.pool_bytes || h.skip_off != mine.skip_off ||
        h.skip_len != mine.skip_len || (h.cur != 0 && h.cur != 1)) {
        fclose(f);
        return SPHG_E_ARG;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    const bool ok = ckpt_stream(c, f, 0, h.skip_off, false) &&
                    ckpt_stream(c, f, h.skip_off + h.skip_len, h.pool_bytes - h.skip_off - h.skip_len, false);
    fclose(f);
    if (!ok) return SPHG_E_ARG;
    c->step = h.step;
    c->cur = h.cur;
    c->stepped = h.stepped != 0;
    c->eig_stale = false;  // the file was written with strep / beta of its stress
    c->has_psc = h.has_psc != 0;
    if (int rc = derive_mass_flags(c)) return rc;  // the file's masses
    return derive_plane_zero(c);                   // and out-of-plane components
}

extern "C" int sphg_set_vibration(sphg_ctx *c, double ttime, int nsteps, double sttime, double edtime,
                                  const double *loads) {
    if (!c || !loads || nsteps < 1) return SPHG_E_ARG;
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->S.vib_loads) {
        HIPCHK(hipFree(c->S.vib_loads));
        c->S.vib_loads = nullptr;
    }
    HIPCHK(hipMalloc(&c->S.vib_loads, (size_t)nsteps * 3 * 8));
    HIPCHK(hipMemcpy(c->S.vib_loads, loads, (size_t)nsteps * 3 * 8, hipMemcpyHostToDevice));
    c->vib_ttime = ttime;
    c->vib_nsteps = nsteps;
    c->vib_sttime = sttime;
    c->vib_edtime = edtime;
    c->P.vib_ttime = ttime;
    c->P.vib_nsteps = nsteps;
    c->P.vib_sttime = sttime;
    c->P.vib_edtime = edtime;
    drop_graphs(c);  // the captured launches hold the old table pointer and parameters
    return SPHG_OK;
}

extern "C" int sphg_set_time(sphg_ctx *c, double t) {
    if (!c) return SPHG_E_ARG;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(c->S.tnow, &t, 8, hipMemcpyHostToDevice));
    return SPHG_OK;
}

extern "C" int sphg_get_time(sphg_ctx *c, double *t) {
    if (!c || !t) return SPHG_E_ARG;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(t, c->S.tnow, 8, hipMemcpyDeviceToHost));
    return SPHG_OK;
}

extern "C" int sphg_set_step(sphg_ctx *c, int l) {
    if (!c) return SPHG_E_ARG;
    c->step = l;
    HIPCHK(hipMemcpy(c->S.step, &l, 4, hipMemcpyHostToDevice));
    return SPHG_OK;
}

extern "C" int sphg_get_neighbors(sphg_ctx *c, int *counts, int *ids, int n) {
    if (!c || !counts || !ids || n != c->n) return SPHG_E_ARG;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(counts, c->S.nb_cnt, (size_t)n * 4, hipMemcpyDeviceToHost));
    std::vector<int> col((size_t)n);
    for (int k = 0; k < kMaxNb; k++) {
        HIPCHK(hipMemcpy(col.data(), c->S.nbr + (size_t)k * c->cap, (size_t)n * 4, hipMemcpyDeviceToHost));
        for (int i = 0; i < n; i++) {
            int cnt = std::min(counts[i], kMaxNb);
            ids[(size_t)i * kMaxNb + k] = k < cnt ? (col[i] & kIdMask) : -1;
        }
    }
    return SPHG_OK;
}

extern "C" int sphg_enable_timing(sphg_ctx *c, int on) {
    if (!c) return SPHG_E_ARG;
    c->timing = on != 0;
    c->tsum.clear();
    return SPHG_OK;
}

// Autotuner of the neighbour batch widths: each of the three width sets is timed over `steps`
// steps from the context's current state (per-pass HIP events), the state is restored after each
// trial, and every pass keeps its fastest width. The widths change launch shapes only: every
// width gives the same results bit for bit (the summation order is the list order).
extern "C" int sphg_autotune(sphg_ctx *c, int steps, int *chosen) {
    if (!c || !c->uploaded || steps <= 0 || c->timing) return SPHG_E_ARG;
    HIPCHK(hipStreamSynchronize(c->stream));
    // the whole device state but the neighbour lists (the next step's search rebuilds them)
    const size_t skip_off = (size_t)((char *)c->S.nbr - c->pool), skip_len = (size_t)c->cap * kMaxNb * 4;
    const size_t tail = c->pool_bytes - skip_off - skip_len;
    char *bak = nullptr;
    HIPCHK(hipMalloc(&bak, skip_off + tail));
    auto copy_state = [&](bool save) -> hipError_t {
        char *a = save ? bak : c->pool, *b = save ? c->pool : bak;
        const size_t a2 = save ? skip_off : skip_off + skip_len, b2 = save ? skip_off + skip_len : skip_off;
        hipError_t e = hipMemcpyAsync(a, b, skip_off, hipMemcpyDeviceToDevice, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(a + a2, b + b2, tail, hipMemcpyDeviceToDevice, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        return e;
    };
    const int step0 = c->step, cur0 = c->cur;
    const bool stepped0 = c->stepped;
    const int saved_bw[4] = {c->bw[0], c->bw[1], c->bw[2], c->bw[3]}, saved_stage = c->search_stage;
    static const int kStageChoices[3] = {16, 32, 16};
    double best_search = 1e300;
    int pick_stage = saved_stage;
    static const char *kNames[4] = {"density", "norm_eos", "strain", "xsph"};
    double best[4] = {1e300, 1e300, 1e300, 1e300};
    int pick[4] = {saved_bw[0], saved_bw[1], saved_bw[2], saved_bw[3]};
    hipError_t e = copy_state(true);
    int rc = e == hipSuccess ? SPHG_OK : SPHG_E_HIP;
    for (int t = 0; t < 3 && rc == SPHG_OK; t++) {
        for (int k = 0; k < 4; k++) c->bw[k] = kBwChoices[k][t];
        c->search_stage = kStageChoices[t];
        sphg_enable_timing(c, 1);
        rc = sphg_step(c, 1);  // first launches of these instantiations, untimed
        sphg_enable_timing(c, 1);
        if (rc == SPHG_OK) rc = sphg_step(c, steps);
        finish_timing(c);
        for (auto &r : c->tsum) {
            if (r.first == "search" && r.second.second > 0 && r.second.first / r.second.second < best_search) {
                best_search = r.second.first / r.second.second;
                pick_stage = kStageChoices[t];
            }
            for (int k = 0; k < 4; k++)
                if (r.first == kNames[k] && r.second.second > 0) {
                    const double avg = r.second.first / r.second.second;
                    if (avg < best[k]) {
                        best[k] = avg;
                        pick[k] = kBwChoices[k][t];
                    }
                }
        }
        sphg_enable_timing(c, 0);
        // back to the state before the trial
        e = copy_state(false);
        if (e != hipSuccess && rc == SPHG_OK) rc = SPHG_E_HIP;
        c->step = step0;
        c->cur = cur0;
        c->stepped = stepped0;
        if (hipMemcpy(c->S.step, &step0, 4, hipMemcpyHostToDevice) != hipSuccess && rc == SPHG_OK) rc = SPHG_E_HIP;
    }
    hipFree(bak);
    for (int k = 0; k < 4; k++) c->bw[k] = rc == SPHG_OK ? pick[k] : saved_bw[k];
    c->search_stage = rc == SPHG_OK ? pick_stage : saved_stage;
    drop_graphs(c);
    if (chosen)
        for (int k = 0; k < 4; k++) chosen[k] = c->bw[k];
    return rc;
}

extern "C" int sphg_set_search_stage(sphg_ctx *c, int stage) {
    if (!c || (stage != 16 && stage != 32)) return SPHG_E_ARG;
    c->search_stage = stage;
    drop_graphs(c);
    return SPHG_OK;
}

extern "C" int sphg_get_search_stage(const sphg_ctx *c, int *stage) {
    if (!c || !stage) return SPHG_E_ARG;
    *stage = c->search_stage;
    return SPHG_OK;
}

extern "C" int sphg_set_batch_widths(sphg_ctx *c, const int *widths) {
    if (!c || !widths) return SPHG_E_ARG;
    for (int k = 0; k < 4; k++) {
        const int *ch = kBwChoices[k];
        if (widths[k] != ch[0] && widths[k] != ch[1] && widths[k] != ch[2]) return SPHG_E_ARG;
    }
    for (int k = 0; k < 4; k++) c->bw[k] = widths[k];
    drop_graphs(c);
    return SPHG_OK;
}

extern "C" int sphg_set_list_order(sphg_ctx *c, int reversed) {
    if (!c || (reversed != 0 && reversed != 1)) return SPHG_E_ARG;
    if (c->rev_lists != (reversed == 1)) {
        c->rev_lists = reversed == 1;
        drop_graphs(c);
    }
    return SPHG_OK;
}

extern "C" int sphg_kernel_times(sphg_ctx *c, char (*names)[32], double *ms, int *launches, int cap) {
    if (!c) return 0;
    finish_timing(c);
    int k = 0;
    for (auto &s : c->tsum) {
        if (k >= cap) break;
        memset(names[k], 0, 32);
        strncpy(names[k], s.first.c_str(), 31);
        ms[k] = s.second.first;
        launches[k] = s.second.second;
        k++;
    }
    return k;
}

extern "C" const char *sphg_error_string(int code) {
    switch (code) {
    case SPHG_OK: return "ok";
    case SPHG_E_NOPARTICLES: return "no particles (reference error 2)";
    case SPHG_E_CELLRANGE: return "particle outside the cell grid (reference error 3)";
    case SPHG_E_NEIGHBOURS: return "more than 99 neighbours (reference error 5)";
    case SPHG_E_ARG: return "bad argument";
    case SPHG_E_HIP: return "HIP runtime error";
    case SPHG_E_UNSUPPORTED: return "flag combination not implemented";
    case SPHG_E_CAPACITY: return "capacity exceeded";
    case SPHG_E_NOCONVERGE: return "K-L root bisection did not converge";
    case SPHG_E_NONFINITE: return "random-field factor not finite (indefinite correlation matrix)";
    default: return "unknown error";
    }
}

extern "C" const char *sphg_build_info(void) {
    return "sphg gfx950 FP64 step library (contraction off)";
}

// -------------------------------------------------------------------------------- K-L fields
// The per-particle array of spatially variable property `field` (0 soil fai, 1 soil c, 2 soil cop,
// 3 soil ds, 4 fluid fai, 5 fluid c), initialised from the material slots (setting_vis_cons copies,
// CPP_2-initialf.cpp:1021-1058) the first time a random field writes it, so the particles of the
// other phase keep their slot values.
static int property_array(sphg_ctx *c, int field, double **out) {
    bool *has = field == 0 ? &c->has_psfai : field == 1 ? &c->has_psc : field == 2 ? &c->has_pscop
              : field == 3 ? &c->has_psds : field == 4 ? &c->has_pffai : &c->has_pfc;
    double *d = field == 0 ? c->S.ps_fai : field == 1 ? c->d_psc : field == 2 ? c->S.ps_cop
              : field == 3 ? c->S.ps_ds : field == 4 ? c->S.pf_fai : c->S.pf_c;
    if (!*has) {
        std::vector<double> base(c->n);
        for (int i = 0; i < c->n; i++) {
            if (field <= 3) {
                const int s = soil_slot_of(c->h_type[i], c->h_matype[i], c->h_etype[i]);
                sphg_soil m{};
                m.fs = 1.0;
                m.ds = 1.0;
                if (s < SPHG_NSOIL) m = c->mat.soil[s];
                base[i] = field == 0 ? m.fai : field == 1 ? m.c : field == 2 ? m.cop : m.ds;
            } else {
                const int s = fluid_slot_of(c->h_type[i], c->h_matype[i], c->h_etype[i]);
                sphg_fluid m{};
                if (s < 3) m = c->mat.water[s];
                else if (s == 3) m = c->mat.air;
                base[i] = field == 4 ? m.fai : m.c;
            }
        }
        HIPCHK(hipMemcpy(d, base.data(), (size_t)c->n * 8, hipMemcpyHostToDevice));
        *has = true;
    }
    *out = d;
    return SPHG_OK;
}

// After a random field changed a property: the derived constants, and the cohesion pointer the
// captured graphs hold.
static int property_changed(sphg_ctx *c, int field) {
    if (field == 1 && c->S.ps_c != c->d_psc) {
        c->S.ps_c = c->d_psc;
        for (int b = 0; b < 2; b++) c->graph_ok[b] = false;  // the State copy in the graphs changed
    }
    if (c->uploaded && (field == 0 || field == 4)) return derive_strength_constants(c);
    return SPHG_OK;
}

extern "C" int sphg_kl_synthesize_part(sphg_ctx *c, const sphg_kl_field *f, const double *xi, int first,
                                       int count) {
    if (!c || !xi || first < 0 || count <= 0 || first + count > c->n) return SPHG_E_ARG;
    int rc = kl_check(f);
    if (rc) return rc;
    double *dst = nullptr;
    if ((rc = property_array(c, f->field, &dst))) return rc;
    sphg::KlModes m;
    kl_modes_device(*f, xi, m);
    sphg::KlModes *dm = nullptr;
    HIPCHK(hipMalloc(&dm, sizeof(m)));
    hipError_t e = hipMemcpy(dm, &m, sizeof(m), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        sphg::KlParams kp = kl_params(*f, count);
        const int cur = c->cur;
        hipLaunchKernelGGL(sphg::k_kl_fp64, dim3(grid_for(count)), dim3(kBlock), 0, c->stream, c->S.xp[cur][0],
                           c->S.xp[cur][1], c->S.xp[cur][2], c->S.type, dm, kp, first, dst);
        e = hipStreamSynchronize(c->stream);
    }
    hipFree(dm);
    if (e != hipSuccess) return SPHG_E_HIP;
    return property_changed(c, f->field);
}

// Correlation coefficient of the two fields' underlying normals and the mixed standard normals of
// SpatialVariables_Generate_2D_corr / _3D_corr (CPP_17-2_K-L_expansion_corr.cpp:114-140): rho =
// coe_frict_cohesion, or for a lognormal first field log(1 + rho cv_a cv_b) /
// sqrt(log(1 + cv_a^2) log(1 + cv_b^2)); zeta = u [[1, rho], [0, sqrt(1 - rho^2)]] row by row with
// matrix_multi's summation order (:22-37).
extern "C" int sphg_kl_corr_normals(const sphg_kl_field *f_a, const sphg_kl_field *f_b, double coe_frict_cohesion,
                                    const double *u, double *xi_a, double *xi_b, double *coefficient) {
    if (!f_a || !f_b || !u || !xi_a || !xi_b) return SPHG_E_ARG;
    double coe = coe_frict_cohesion;
    if (f_a->dist == 2) {
        coe = log(1.0 + coe_frict_cohesion * (f_a->sd_value / f_a->mean_value) * (f_b->sd_value / f_b->mean_value));
        coe = coe / sqrt(log(1.0 + f_a->sd_value * f_a->sd_value / f_a->mean_value / f_a->mean_value) *
                         log(1.0 + f_b->sd_value * f_b->sd_value / f_b->mean_value / f_b->mean_value));
    }
    const double cov[4] = {1.0, coe, 0.0, sqrt(1.0 - coe * coe)};
    for (int i = 0; i < SPHG_KL_TERMS; i++) {
        double z[2];
        for (int j = 0; j < 2; j++) {
            double t = 0.0;
            for (int k = 0; k < 2; k++) t = t + u[i * 2 + k] * cov[k * 2 + j];
            z[j] = t;
        }
        xi_a[i] = z[0];
        xi_b[i] = z[1];
    }
    if (coefficient) *coefficient = coe;
    return SPHG_OK;
}

extern "C" int sphg_kl_synthesize_corr_part(sphg_ctx *c, const sphg_kl_field *f_a, const sphg_kl_field *f_b,
                                            double coe_frict_cohesion, const double *u, int first, int count) {
    if (!c || !f_a || !f_b || !u) return SPHG_E_ARG;
    if (!((f_a->field == 0 && f_b->field == 1) || (f_a->field == 4 && f_b->field == 5))) return SPHG_E_ARG;
    std::vector<double> xa(SPHG_KL_TERMS), xb(SPHG_KL_TERMS);
    int rc = sphg_kl_corr_normals(f_a, f_b, coe_frict_cohesion, u, xa.data(), xb.data(), nullptr);
    if (!rc) rc = sphg_kl_synthesize_part(c, f_a, xa.data(), first, count);
    if (!rc) rc = sphg_kl_synthesize_part(c, f_b, xb.data(), first, count);
    return rc;
}

extern "C" int sphg_kl_synthesize_corr(sphg_ctx *c, const sphg_kl_field *f_a, const sphg_kl_field *f_b,
                                       double coe_frict_cohesion, const double *u) {
    if (!c) return SPHG_E_ARG;
    return sphg_kl_synthesize_corr_part(c, f_a, f_b, coe_frict_cohesion, u, 0, c->n);
}

// random_variable_generate (CPP_17-5-Nonstationary-Correlated-RF.cpp:738-863): one correlated pair of
// standard normals for the whole domain, mixed as above with the distribution of psv[0] (flag_fai)
// or else psv[4] (flag_faiw), or else the constructor default (normal); fai takes the mixed normal
// and c the first one.
extern "C" int sphg_random_variable_values(const sphg_rv_flags *fsv, const sphg_rv_param *psv,
                                           const double zeta[2], double values[6]) {
    if (!fsv || !psv || !zeta || !values) return SPHG_E_ARG;
    sphg_rv_param t{1, 0.0, 0.0};
    if (fsv->flag_fai == 1) t = psv[0];
    else if (fsv->flag_faiw == 1) t = psv[4];
    const int dist = t.type;
    double coe;
    if (dist == 2 && fsv->flag_fai == 1) {
        coe = log(1.0 + fsv->coe_frict_cohesion * (psv[0].sd_value / psv[0].mean_value) *
                            (psv[1].sd_value / psv[1].mean_value));
        coe = coe / sqrt(log(1.0 + psv[0].sd_value * psv[0].sd_value / psv[0].mean_value / psv[0].mean_value) *
                         log(1.0 + psv[1].sd_value * psv[1].sd_value / psv[1].mean_value / psv[1].mean_value));
    } else if (dist == 2 && fsv->flag_faiw == 1) {
        coe = log(1.0 + fsv->coe_frict_cohesion * (psv[4].sd_value / psv[4].mean_value) *
                            (psv[5].sd_value / psv[5].mean_value));
        coe = coe / sqrt(log(1.0 + psv[4].sd_value * psv[4].sd_value / psv[4].mean_value / psv[4].mean_value) *
                         log(1.0 + psv[5].sd_value * psv[5].sd_value / psv[5].mean_value / psv[5].mean_value));
    } else {
        coe = fsv->coe_frict_cohesion;
    }
    const double cov[4] = {1.0, coe, 0.0, sqrt(1.0 - coe * coe)};
    double z1[2];
    for (int j = 0; j < 2; j++) {
        double tt = 0.0;
        for (int k = 0; k < 2; k++) tt = tt + zeta[k] * cov[k * 2 + j];
        z1[j] = tt;
    }
    // property slot (psv index) and which mixed normal it takes
    const int flag[4] = {fsv->flag_fai, fsv->flag_c, fsv->flag_faiw, fsv->flag_cw};
    const int slot[4] = {0, 1, 4, 5}, zi[4] = {1, 0, 1, 0};
    for (int k = 0; k < 6; k++) values[k] = NAN;
    for (int q = 0; q < 4; q++) {
        if (flag[q] != 1) continue;
        double mean = psv[slot[q]].mean_value, sd = psv[slot[q]].sd_value;
        if (dist == 1) {
            values[slot[q]] = mean + sd * z1[zi[q]];
        } else if (dist == 2) {
            sd = sqrt(log(1.0 + (sd / mean) * (sd / mean)));
            mean = log(mean) - 0.5 * sd * sd;
            values[slot[q]] = exp(mean + sd * z1[zi[q]]);
        }
    }
    return SPHG_OK;
}

extern "C" int sphg_random_variable_generate(sphg_ctx *c, const sphg_rv_flags *fsv, const sphg_rv_param *psv,
                                             const double zeta[2], int first, int count) {
    if (!c || first < 0 || count <= 0 || first + count > c->n) return SPHG_E_ARG;
    double v[6];
    int rc = sphg_random_variable_values(fsv, psv, zeta, v);
    if (rc) return rc;
    const int t = fsv->flag_fai == 1 ? psv[0].type : fsv->flag_faiw == 1 ? psv[4].type : 1;
    if (t != 1 && t != 2) return SPHG_OK;  // the reference assigns nothing for another distribution
    const int flag[4] = {fsv->flag_fai, fsv->flag_c, fsv->flag_faiw, fsv->flag_cw};
    const int slot[4] = {0, 1, 4, 5};
    for (int q = 0; q < 4; q++) {
        if (flag[q] != 1) continue;
        double *dst = nullptr;
        if ((rc = property_array(c, slot[q], &dst))) return rc;
        hipLaunchKernelGGL(k_fill_phase, dim3(grid_for(count)), dim3(kBlock), 0, c->stream, dst, c->S.type,
                           slot[q] <= 1 ? 2 : 1, v[slot[q]], first, count);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
        if ((rc = property_changed(c, slot[q]))) return rc;
    }
    return SPHG_OK;
}

extern "C" int sphg_kl_synthesize(sphg_ctx *c, const sphg_kl_field *f, const double *xi) {
    if (!c) return SPHG_E_ARG;
    return sphg_kl_synthesize_part(c, f, xi, 0, c->n);
}

// one workgroup per 128 x 128 tile of H[npad][rpad]
static dim3 kl_gemm_grid(int npad, int rpad) {
    return dim3((npad + sphg::kGm - 1) / sphg::kGm, (rpad + sphg::kGn - 1) / sphg::kGn);
}

extern "C" int sphg_kl_synthesize_batch(const sphg_kl_field *f, int n, const double *xp, const int *type, int nreal,
                                        const double *xi, double fill, double *value, double *gemm_ms) {
    if (!xp || !type || !xi || !value || n <= 0 || nreal <= 0) return SPHG_E_ARG;
    int rc = kl_check(f);
    if (rc) return rc;
    const int npad = (n + 31) / 32 * 32, rpad = (nreal + 31) / 32 * 32;
    sphg::KlModes m;
    kl_modes_device(*f, nullptr, m);
    std::vector<float> xif((size_t)kKlTermsHost * rpad, 0.0f);
    for (int r = 0; r < nreal; r++)
        for (int j = 0; j < kKlTermsHost; j++) xif[(size_t)j * rpad + r] = (float)xi[(size_t)r * kKlTermsHost + j];
    sphg::KlParams kp = kl_params(*f, n);
    char *pool = nullptr;
    const size_t b_xp = align256((size_t)n * 24), b_ty = align256((size_t)n * 4), b_m = align256(sizeof(m)),
                 b_phi = align256((size_t)npad * kKlTermsHost * 4), b_xi = align256(xif.size() * 4),
                 b_h = align256((size_t)npad * rpad * 4), b_v = align256((size_t)nreal * n * 8);
    HIPCHK(hipMalloc(&pool, b_xp + b_ty + b_m + b_phi + b_xi + b_h + b_v));
    double *dxp = (double *)pool;
    int *dty = (int *)(pool + b_xp);
    sphg::KlModes *dm = (sphg::KlModes *)(pool + b_xp + b_ty);
    float *dphi = (float *)(pool + b_xp + b_ty + b_m);
    float *dxi = (float *)(pool + b_xp + b_ty + b_m + b_phi);
    float *dh = (float *)(pool + b_xp + b_ty + b_m + b_phi + b_xi);
    double *dv = (double *)(pool + b_xp + b_ty + b_m + b_phi + b_xi + b_h);
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipMemcpyAsync(dxp, xp, (size_t)n * 24, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dty, type, (size_t)n * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dm, &m, sizeof(m), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dxi, xif.data(), xif.size() * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        const int tot = npad * kKlTermsHost;
        hipLaunchKernelGGL(sphg::k_kl_phi, dim3((tot + 255) / 256), dim3(256), 0, st, dxp, dty, dm, kp, npad, dphi);
        hipEventRecord(e0, st);
        hipLaunchKernelGGL(sphg::k_kl_gemm, kl_gemm_grid(npad, rpad), dim3(256), 0, st, dphi, dxi, npad, rpad, dh);
        hipEventRecord(e1, st);
        const int totv = n * nreal;
        hipLaunchKernelGGL(sphg::k_kl_values, dim3((totv + 255) / 256), dim3(256), 0, st, dh, dxp, dty, kp, nreal,
                           rpad, fill, dv);
        e = hipMemcpyAsync(value, dv, (size_t)nreal * n * 8, hipMemcpyDeviceToHost, st);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess && gemm_ms) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        *gemm_ms = ms;
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (st) hipStreamDestroy(st);
    hipFree(pool);
    return e == hipSuccess ? SPHG_OK : SPHG_E_HIP;
}

// ------------------------------------------------------ K-L batch synthesis without allocation
// The same synthesis as sphg_kl_synthesize_batch for callers that cannot allocate device memory
// (code under stream capture, tight serving loops): every buffer lives in one file-scope device
// array, `scratch`, of kNaBytes, and the particles pass through it kNaChunk at a time. Each GPU
// holds its own copy of the array; its address is looked up once per device, with that device
// current, and only that address is handed to kernels and copies. Calls that overlap on one GPU
// share the array and race.
namespace {
constexpr int kNaChunk = 8192;  // particles per pass (a multiple of 32)
constexpr int kNaMaxReal = 256; // realisations per call
constexpr size_t kNaXp = 0, kNaTy = kNaXp + align256c((size_t)kNaChunk * 24),
                 kNaModes = kNaTy + align256c((size_t)kNaChunk * 4),
                 kNaPhi = kNaModes + align256c(sizeof(sphg::KlModes)),
                 kNaXiD = kNaPhi + align256c((size_t)kNaChunk * SPHG_KL_TERMS * 4),
                 kNaXi = kNaXiD + align256c((size_t)kNaMaxReal * SPHG_KL_TERMS * 8),
                 kNaH = kNaXi + align256c((size_t)SPHG_KL_TERMS * kNaMaxReal * 4),
                 kNaV = kNaH + align256c((size_t)kNaChunk * kNaMaxReal * 4),
                 kNaBytes = kNaV + align256c((size_t)kNaMaxReal * kNaChunk * 8);
constexpr int kNaMaxDevices = 64;
char *g_na_ptr[kNaMaxDevices];
std::mutex g_na_mutex;
}  // namespace

static __device__ __align__(256) char scratch[kNaBytes];

namespace sphg {
// Xi[200][rpad] (FP32, zero past nreal) from the callers' standard normals xi[nreal][200] (FP64)
__global__ __launch_bounds__(256) void k_kl_xi(const double *xi, int nreal, int rpad, float *out) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= SPHG_KL_TERMS * rpad) return;
    const int j = g / rpad, r = g % rpad;
    out[g] = r < nreal ? (float)xi[(size_t)r * SPHG_KL_TERMS + j] : 0.0f;
}
}  // namespace sphg

// Device address of this GPU's scratch array (looked up on the first call on each device).
static int na_scratch(char **out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNaMaxDevices) return SPHG_E_HIP;
    std::lock_guard<std::mutex> lock(g_na_mutex);
    if (!g_na_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return SPHG_E_HIP;
        g_na_ptr[dev] = (char *)p;
    }
    *out = g_na_ptr[dev];
    return SPHG_OK;
}

extern "C" int sphg_kl_synthesize_batch_noalloc(const sphg_kl_field *f, int n, const double *xp, const int *type,
                                                int nreal, const double *xi, double fill, double *value) {
    if (!xp || !type || !xi || !value || n <= 0 || nreal <= 0) return SPHG_E_ARG;
    if (nreal > kNaMaxReal) return SPHG_E_CAPACITY;
    int rc = kl_check(f);
    if (rc) return rc;
    char *base = nullptr;
    rc = na_scratch(&base);
    if (rc) return rc;
    double *dxp = (double *)(base + kNaXp);
    int *dty = (int *)(base + kNaTy);
    sphg::KlModes *dm = (sphg::KlModes *)(base + kNaModes);
    float *dphi = (float *)(base + kNaPhi), *dxi = (float *)(base + kNaXi), *dh = (float *)(base + kNaH);
    double *dxid = (double *)(base + kNaXiD), *dv = (double *)(base + kNaV);
    const int rpad = (nreal + 31) / 32 * 32;
    sphg::KlModes m;
    kl_modes_device(*f, nullptr, m);
    const hipStream_t st = nullptr;
    hipError_t e = hipMemcpyAsync(dm, &m, sizeof(m), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dxid, xi, (size_t)nreal * SPHG_KL_TERMS * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        hipLaunchKernelGGL(sphg::k_kl_xi, dim3((SPHG_KL_TERMS * rpad + 255) / 256), dim3(256), 0, st,
                           (const double *)dxid, nreal, rpad, dxi);
    for (int first = 0; first < n && e == hipSuccess; first += kNaChunk) {
        const int cn = std::min(kNaChunk, n - first), cpad = (cn + 31) / 32 * 32;
        sphg::KlParams kp = kl_params(*f, cn);
        e = hipMemcpyAsync(dxp, xp + (size_t)first * 3, (size_t)cn * 24, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(dty, type + first, (size_t)cn * 4, hipMemcpyHostToDevice, st);
        if (e != hipSuccess) break;
        const int tot = cpad * kKlTermsHost;
        hipLaunchKernelGGL(sphg::k_kl_phi, dim3((tot + 255) / 256), dim3(256), 0, st, (const double *)dxp,
                           (const int *)dty, (const sphg::KlModes *)dm, kp, cpad, dphi);
        hipLaunchKernelGGL(sphg::k_kl_gemm, kl_gemm_grid(cpad, rpad), dim3(256), 0, st, (const float *)dphi,
                           (const float *)dxi, cpad, rpad, dh);
        hipLaunchKernelGGL(sphg::k_kl_values, dim3((cn * nreal + 255) / 256), dim3(256), 0, st, (const float *)dh,
                           (const double *)dxp, (const int *)dty, kp, nreal, rpad, fill, dv);
        e = hipGetLastError();
        // value[r][first .. first + cn) from the chunk's value[r][0 .. cn)
        if (e == hipSuccess)
            e = hipMemcpy2DAsync(value + first, (size_t)n * 8, dv, (size_t)cn * 8, (size_t)cn * 8, nreal,
                                 hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);  // the next chunk reuses the buffers
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e == hipSuccess ? SPHG_OK : SPHG_E_HIP;
}

// ------------------------------------------------------------- correlation-matrix random fields
static const double kRefPi = 3.141592653;  // Header_Parameters.h:10

extern "C" int sphg_rf_cells(sphg_rf_field *f, int n, const double *xp, const int *type, int *cell_of) {
    if (!f || n <= 0 || !xp || !type || (f->ndim != 2 && f->ndim != 3) || !(f->cell_ndx > 0.0) || !(f->dr > 0.0))
        return SPHG_E_ARG;
    const int nd = f->ndim;
    const double *b = f->box_nb;
    double xmin[3] = {0.0, 0.0, 0.0}, xmax[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < nd; d++) {
        xmin[d] = b[2 * d] - 0.1 * (b[2 * d + 1] - b[2 * d]);
        xmax[d] = b[2 * d + 1] + 0.1 * (b[2 * d + 1] - b[2 * d]);
    }
    const double cell_dx = f->cell_ndx * f->dr;
    long long nc[3] = {1, 1, 1};
    for (int d = 0; d < nd; d++) nc[d] = (long long)((xmax[d] - xmin[d]) / cell_dx) + 1;
    const long long total = nc[0] * nc[1] * nc[2];
    if (!(total > 0 && total < 15000)) return 41;
    f->cell_dx = cell_dx;
    f->cell_total = (int)total;
    for (int d = 0; d < 3; d++) {
        f->nc[d] = (int)nc[d];
        f->base[d] = d < nd ? -(nc[d] * cell_dx - (xmax[d] - xmin[d])) / 2.0 + 0.01 * f->dr + xmin[d] : 0.0;
    }
    std::vector<int> cnt((size_t)total, 0);
    for (int i = 0; i < n; i++) {
        if (cell_of) cell_of[i] = -1;
        if (nd == 2 && type[i] != 1 && type[i] != 2) continue;  // 2D maps water and soil only
        long long np[3] = {0, 0, 0};
        for (int d = 0; d < nd; d++) np[d] = (long long)(fabs(xp[3 * (size_t)i + d] - f->base[d]) / cell_dx);
        const long long c = nd == 2 ? np[1] * nc[0] + np[0] : np[2] * nc[0] * nc[1] + np[1] * nc[0] + np[0];
        if (c < 0 || c >= total) return SPHG_E_ARG;
        if (cnt[(size_t)c] >= 39) return 42;
        cnt[(size_t)c]++;
        if (cell_of) cell_of[i] = (int)c;
    }
    return SPHG_OK;
}

// One random-field generation: cell partition and correlation matrix of the geometry in fin,
// its factor, E = L Z^T for the ncolz columns zcols[ncolz][total], and one value array per spec
// (values[s][specs[s].nreal][n]).
static int rf_run(const sphg_rf_field *fin, int n, const double *xp, const int *type, int ncolz, const double *zcols,
                  int nspec, const sphg::RfValues *specs, double *const *values, double fill, double *L,
                  double *factor_ms) {
    sphg_rf_field f = *fin;
    std::vector<int> cell_of(n > 0 ? n : 0);
    int rc = sphg_rf_cells(&f, n, xp, type, cell_of.data());
    if (rc) return rc;
    using namespace sphg;
    const int total = f.cell_total, npad = (total + kRfB - 1) / kRfB * kRfB, rpad = (ncolz + kRfB - 1) / kRfB * kRfB;
    int vreal = 0;
    for (int q = 0; q < nspec; q++) vreal = std::max(vreal, specs[q].nreal);
    RfParams p{};
    p.ndim = f.ndim;
    p.acf = f.acf;
    p.total = total;
    p.npad = npad;
    p.cell_dx = f.cell_dx;
    p.pi = kRefPi;
    for (int d = 0; d < 3; d++) {
        p.nc[d] = f.nc[d];
        p.base[d] = f.base[d];
        p.rl[d] = f.rl[d];
    }
    const double bt = f.beta * kRefPi / 180.0, bt1 = f.beta_1 * kRefPi / 180.0;  // beta_temp, beta1_temp
    p.cb = cos(bt); p.sb = sin(bt); p.cb1 = cos(bt1); p.sb1 = sin(bt1);
    std::vector<double> zh((size_t)rpad * npad, 0.0);
    for (int r = 0; r < ncolz; r++)
        for (int j = 0; j < total; j++) zh[(size_t)r * npad + j] = zcols[(size_t)r * total + j];
    // E = L Z^T in kRfSlices slices of K (multiples of 64), summed in order afterwards
    constexpr int kRfSlices = 8;
    const int kslice = ((npad / kRfB + kRfSlices - 1) / kRfSlices) * kRfB;
    const size_t bA = align256((size_t)npad * npad * 8), bZ = align256(zh.size() * 8),
                 bE = align256((size_t)npad * rpad * 8), bC = align256((size_t)n * 4), bT = align256((size_t)n * 4),
                 bV = align256((size_t)vreal * n * 8), bP = align256((size_t)kRfSlices * npad * rpad * 8);
    char *pool = nullptr;
    HIPCHK(hipMalloc(&pool, bA + bZ + bE + bC + bT + bV + bP + 256));
    double *dA = (double *)pool, *dZ = (double *)(pool + bA), *dE = (double *)(pool + bA + bZ);
    int *dC = (int *)(pool + bA + bZ + bE), *dT = (int *)(pool + bA + bZ + bE + bC);
    double *dV = (double *)(pool + bA + bZ + bE + bC + bT), *dEp = (double *)(pool + bA + bZ + bE + bC + bT + bV);
    int *dFlag = (int *)(pool + bA + bZ + bE + bC + bT + bV + bP);
    int nonfinite = 0;
    hipStream_t st = nullptr, st2 = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr, ev_f = nullptr, ev_b = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st2, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ev_f, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ev_b, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(ev_b, st2);  // nothing pending on the second stream yet
    if (e == hipSuccess) e = hipMemcpyAsync(dZ, zh.data(), zh.size() * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dC, cell_of.data(), (size_t)n * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dT, type, (size_t)n * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        const size_t tot = (size_t)npad * npad;
        const int gcorr = (int)std::min<size_t>((tot + 255) / 256, 8192);
        hipLaunchKernelGGL(k_rf_corr, dim3(gcorr), dim3(256), 0, st, p, dA);
        hipEventRecord(e0, st);
        // wide panels of 256 columns: inside one, each 64-column step factors its diagonal block,
        // solves the rows below and updates the panel's remaining columns. The trailing update
        // by a finished panel (K = 256, a quarter of the traffic of 64-wide steps) is split: the
        // next panel's columns on the main stream, the rest on a second stream, where it runs
        // while the main stream factors the next panel (one panel of look-ahead). The main
        // stream waits for the second before it touches columns the second still updates.
        constexpr int kWide = 4 * kRfB;
        const double *cA = dA;
        for (int w0 = 0; w0 < npad && e == hipSuccess; w0 += kWide) {
            const int wend = std::min(npad, w0 + kWide);
            for (int k0 = w0; k0 < wend; k0 += kRfB) {
                hipLaunchKernelGGL(k_rf_potrf, dim3(1), dim3(kRfB), 0, st, dA, npad, k0);
                const int below = (npad - k0 - kRfB) / kRfB;
                if (below <= 0) continue;
                hipLaunchKernelGGL(k_rf_trsm, dim3(below), dim3(kRfB), 0, st, dA, npad, k0);
                const int ncol = (wend - k0 - kRfB) / kRfB;
                if (ncol > 0)
                    hipLaunchKernelGGL(k_rf_gemm<kRect>, dim3(below, ncol), dim3(256), 0, st, cA, npad, cA, npad, dA,
                                       npad, k0 + kRfB, k0 + kRfB, k0, kRfB, (size_t)0);
            }
            const int T = (npad - wend) / kRfB;
            if (T <= 0) break;
            const int nnext = std::min(T, kWide / kRfB);  // the next panel's column tiles
            // the rest: rows and columns past the next panel, on the second stream once this
            // panel is factored
            e = hipEventRecord(ev_f, st);
            if (e == hipSuccess) e = hipStreamWaitEvent(st2, ev_f, 0);
            const int Tr = T - nnext;
            if (e == hipSuccess && Tr > 0)
                hipLaunchKernelGGL(k_rf_gemm<kTri>, dim3(Tr * (Tr + 1) / 2), dim3(256), 0, st2, cA, npad, cA, npad, dA,
                                   npad, wend + nnext * kRfB, wend + nnext * kRfB, w0, wend - w0, (size_t)0);
            // the next panel's columns, after the second stream has finished the previous
            // panel's update of those same columns
            if (e == hipSuccess) e = hipStreamWaitEvent(st, ev_b, 0);
            if (e == hipSuccess)
                hipLaunchKernelGGL(k_rf_gemm<kRect>, dim3(T, nnext), dim3(256), 0, st, cA, npad, cA, npad, dA, npad,
                                   wend, wend, w0, wend - w0, (size_t)0);
            if (e == hipSuccess) e = hipEventRecord(ev_b, st2);
        }
        if (e == hipSuccess) e = hipStreamWaitEvent(st, ev_b, 0);
        hipEventRecord(e1, st);
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, st, dFlag, 1, 0);
        hipLaunchKernelGGL(k_rf_nonfinite, dim3((unsigned)(((size_t)total * total + 255) / 256)), dim3(256), 0, st,
                           (const double *)dA, npad, total, dFlag);
        if (e == hipSuccess) e = hipMemcpyAsync(&nonfinite, dFlag, 4, hipMemcpyDeviceToHost, st);
        const size_t es = (size_t)npad * rpad;
        hipLaunchKernelGGL(k_rf_gemm<kProd>, dim3(npad / kRfB, rpad / kRfB, kRfSlices), dim3(256), 0, st, cA, npad,
                           (const double *)dZ, npad, dEp, rpad, 0, 0, 0, kslice, es);
        hipLaunchKernelGGL(k_rf_sum_slices, dim3((unsigned)((es + 255) / 256)), dim3(256), 0, st, (const double *)dEp,
                           kRfSlices, es, dE);
        e = hipGetLastError();
        for (int sidx = 0; sidx < nspec && e == hipSuccess; sidx++) {
            RfValues q = specs[sidx];
            q.total = total;
            q.rpad = rpad;
            q.n = n;
            q.e1 = exp(1.0);
            const long long nv = (long long)n * q.nreal;
            hipLaunchKernelGGL(k_rf_values, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, st, p, q,
                               (const double *)dE, (const int *)dC, (const int *)dT, fill, dV);
            e = hipMemcpyAsync(values[sidx], dV, (size_t)q.nreal * n * 8, hipMemcpyDeviceToHost, st);
        }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess && L) {
        for (int r = 0; r < total && e == hipSuccess; r++) {
            e = hipMemcpy(L + (size_t)r * total, dA + (size_t)r * npad, (size_t)total * 8, hipMemcpyDeviceToHost);
        }
    }
    if (e == hipSuccess && factor_ms) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        *factor_ms = ms;
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (ev_f) hipEventDestroy(ev_f);
    if (ev_b) hipEventDestroy(ev_b);
    if (st2) hipStreamSynchronize(st2);
    if (st) hipStreamDestroy(st);
    if (st2) hipStreamDestroy(st2);
    hipFree(pool);
    if (e != hipSuccess) return SPHG_E_HIP;
    return nonfinite ? SPHG_E_NONFINITE : SPHG_OK;
}

// the spec of one property of the correlation-matrix method (nonstationary_random_2D/3D)
static sphg::RfValues rf_spec_single(const sphg_rf_field &f, int nreal) {
    sphg::RfValues q{};
    q.dist = f.dist;
    q.want_type = f.field <= 3 ? 2 : 1;
    q.nreal = nreal;
    q.ncol = 1;
    q.col = 0;
    q.vaxis = f.ndim == 2 ? 1 : 2;
    q.use_exp = 0;
    q.mean_value = f.mean_value; q.sd_value = f.sd_value; q.coe_depth = f.coe_depth;
    q.top_normal = f.field <= 3 ? f.soil_max : f.water_max;
    q.top_log = f.soil_max;
    return q;
}

extern "C" int sphg_rf_generate(const sphg_rf_field *fin, int n, const double *xp, const int *type, int nreal,
                                const double *zeta, double fill, double *value, double *L, double *factor_ms) {
    if (!fin || !zeta || !value || nreal <= 0 || (fin->dist != 1 && fin->dist != 2) || fin->field < 0 ||
        fin->field > 5)
        return SPHG_E_ARG;
    const sphg::RfValues q = rf_spec_single(*fin, nreal);
    double *vals[1] = {value};
    return rf_run(fin, n, xp, type, nreal, zeta, 1, &q, vals, fill, L, factor_ms);
}

extern "C" int sphg_rf_generate_pair(const sphg_rf_field *f_fai, const sphg_rf_field *f_c, double coe_frict_cohesion,
                                     int n, const double *xp, const int *type, int nreal, const double *zeta,
                                     double fill, double *value_fai, double *value_c) {
    if (!f_fai || !f_c || !zeta || !value_fai || !value_c || nreal <= 0 || (f_fai->dist != 1 && f_fai->dist != 2) ||
        (f_fai->field != 0 && f_fai->field != 4) || f_c->field != f_fai->field + 1)
        return SPHG_E_ARG;
    sphg_rf_field g = *f_fai;
    int rc = sphg_rf_cells(&g, n, xp, type, nullptr);
    if (rc) return rc;
    const int total = g.cell_total;
    // correlation of the two standard normals (CPP_17-5-Nonstationary-Correlated-RF.cpp:215-231): the
    // lognormal pair takes the correlation of the underlying normals
    double coef = coe_frict_cohesion;
    if (f_fai->dist == 2) {
        const double m0 = f_fai->mean_value, s0 = f_fai->sd_value, m1 = f_c->mean_value, s1 = f_c->sd_value;
        coef = log(1.0 + coe_frict_cohesion * (s0 / m0) * (s1 / m1));
        coef = coef / sqrt(log(1.0 + s0 * s0 / m0 / m0) * log(1.0 + s1 * s1 / m1 / m1));
    }
    const double cov[4] = {1.0, coef, 0.0, sqrt(1.0 - coef * coef)};
    // zeta_1 * cov per cell (matrix_multi, :22-36), realisation r's two columns at rows 2r, 2r + 1
    std::vector<double> z((size_t)2 * nreal * total);
    for (int r = 0; r < nreal; r++)
        for (int j = 0; j < total; j++) {
            const double *z1 = zeta + ((size_t)r * total + j) * 2;
            for (int c = 0; c < 2; c++) {
                double t = 0.0;
                t = t + z1[0] * cov[0 * 2 + c];
                t = t + z1[1] * cov[1 * 2 + c];
                z[((size_t)2 * r + c) * total + j] = t;
            }
        }
    sphg::RfValues q[2];
    const sphg_rf_field *fs[2] = {f_fai, f_c};
    for (int k = 0; k < 2; k++) {
        sphg::RfValues &v = q[k];
        v = sphg::RfValues{};
        v.dist = f_fai->dist;  // both properties take the distribution of the friction angle
        v.want_type = f_fai->field == 0 ? 2 : 1;
        v.nreal = nreal;
        v.ncol = 2;
        v.col = k;
        v.use_exp = 1;
        v.mean_value = fs[k]->mean_value; v.sd_value = fs[k]->sd_value; v.coe_depth = fs[k]->coe_depth;
        if (f_fai->field == 0) {  // soil: depth below soil_max on the vertical axis
            v.vaxis = g.ndim == 2 ? 1 : 2;
            v.top_normal = v.top_log = f_fai->soil_max;
        } else {  // fluid: the normal branch below water_max, the lognormal below soil_max, on y in 2D and 3D
            v.vaxis = 1;
            v.top_normal = f_fai->water_max;
            v.top_log = f_fai->soil_max;
        }
    }
    double *vals[2] = {value_fai, value_c};
    return rf_run(f_fai, n, xp, type, 2 * nreal, z.data(), 2, q, vals, fill, nullptr, nullptr);
}
