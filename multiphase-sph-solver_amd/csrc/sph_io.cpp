// Host-side deck and result files: the reference's input.dat reader, and its per-output-step
// particle table (.odb), written from host arrays or from a context's device state.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sphg.h"

namespace {

// output_total's file name (CPP_14-2-output.cpp:27-76): data-<l zero-padded to 10 digits>.odb;
// from l = 1e8 on the reference leaves the number out ("data-.odb") and so does this writer.
std::string odb_name(const char *dir, int l) {
    char num[16] = "";
    if (l < 100000000) snprintf(num, sizeof num, "%010d", l);
    return std::string(dir) + "/data-" + num + ".odb";
}

const char kRule[] =
    "*************************************************************************************************\n";

// One particle row: the counter, 47 values in %14.6e separated by single spaces except a double
// space between the 41st and 42nd value (divq, divr), then type and matype in %4d; the layout of
// CPP_14-2-output.cpp:87-104 (solid and fluid rows) and :111-128 (boundary rows).
void put_row(FILE *f, int rec, const double v[47], int type, int matype) {
    char buf[47 * 16 + 64];
    int p = snprintf(buf, sizeof buf, "%5d", rec);
    for (int k = 0; k < 47; k++) {
        const char *sep = k == 41 ? "  " : " ";
        p += snprintf(buf + p, sizeof buf - p, "%s%14.6e", sep, v[k]);
    }
    p += snprintf(buf + p, sizeof buf - p, " %4d %4d\n", type, matype);
    fwrite(buf, 1, (size_t)p, f);
}

// ---------------------------------------------------------------------------- input.dat
// The reference finds each section afresh from the top of the file: it reads the first two
// characters of every line until they are the section's tag, then the first character of every
// following line until one is 'A' (the values follow, free-form) or 'E' (the section has none).
// Most sections scan lines with "%c%c%*[^\n]%*c"; WP, AP, SF and SV with "%c%c %*[^\n]%*c",
// which differs only on lines shorter than three characters. Both forms are used as they are.
// Where the reference would spin forever (end of file inside a section) this returns false.
bool find_section(FILE *f, const char *tag, bool spaced) {
    rewind(f);
    char c[2] = {0, 0};
    while (!feof(f)) {
        if (spaced) {
            if (fscanf(f, "%c%c %*[^\n]%*c", &c[0], &c[1]) < 0) break;
        } else if (fscanf(f, "%c%c%*[^\n]%*c", &c[0], &c[1]) < 0) {
            break;
        }
        if (c[0] != tag[0] || c[1] != tag[1]) continue;
        do {
            if (fscanf(f, "%c%*[^\n]%*c", &c[0]) < 0) return false;
            if (c[0] == 'A') return true;
        } while (c[0] != 'E');
        return false;
    }
    return false;
}

// Values of a section: each pointer in order, read with %d or %lf like the reference's scanf
// lists; stops at the first value that does not parse (the rest keep their defaults).
struct Val {
    int *i;
    double *d;
};
bool read_vals(FILE *f, std::initializer_list<Val> vals) {
    for (const Val &v : vals) {
        const int r = v.i ? fscanf(f, "%d", v.i) : fscanf(f, "%lf", v.d);
        if (r != 1) return false;
    }
    return true;
}
Val I(int &x) { return {&x, nullptr}; }
Val D(double &x) { return {nullptr, &x}; }

// Para_FL, Para_Pro, Para_Boundary, Para_SSInter and Para_Model constructors
// (Class_Problem.cpp:12-177, Class_ParticleProperty.cpp:9-22): everything zero except these.
void problem_defaults(sphg_problem &p) {
    p = sphg_problem{};
    p.flknl = 1;
    p.flbndy = 2;
    p.ntd = 1;
    p.ndim = 1;
    p.step_regu = 1;
    p.coe_bndy = 0.10;
    p.ss_zeta = 0.5;
    p.ss_acc_type = 1;
}

// Para_Soil / Para_Fluid / Para_Structure constructors (Class_ParticleProperty.cpp:28-124)
void material_defaults(sphg_materials &m) {
    m = sphg_materials{};
    for (sphg_soil &s : m.soil) {
        s.fs = 1.0;
        s.ds = 1.0;
    }
    for (sphg_fluid *w : {&m.water[0], &m.water[1], &m.water[2], &m.air}) w->strain_min = 1.0;
}

// Per-layer value lists of the soil model blocks (initial_mat, CPP_2-initialf.cpp:487-795), in
// model order: elastic, sub-loading Cam-Clay, Drucker-Prager, unsaturated MCC, AMCC, Bui DP,
// perfect plastic, viscoelastic, FOCC, elastoplastic.
struct ModelBlock {
    const char *tag;
    std::vector<double sphg_soil::*> fields;
};
const std::vector<ModelBlock> &model_blocks() {
    using S = sphg_soil;
    static const std::vector<ModelBlock> b = {
        {"EM", {&S::kapa, &S::poi, &S::dens, &S::damp1, &S::damp2, &S::cop, &S::eps0, &S::e, &S::ds}},
        {"SB", {&S::poi, &S::dens, &S::eps0, &S::zmf, &S::sme0, &S::fs, &S::ramda, &S::kapa, &S::damp1, &S::damp2,
                &S::cop, &S::ann, &S::ocr, &S::e, &S::ds}},
        {"DP", {&S::kapa, &S::poi, &S::dens, &S::c, &S::fai, &S::fs, &S::damp1, &S::damp2, &S::cop, &S::eps0, &S::e,
                &S::ds}},
        {"UM", {&S::poi, &S::dens, &S::eps0, &S::zmf, &S::sme0, &S::fs, &S::ramda, &S::kapa, &S::damp1, &S::damp2,
                &S::cop, &S::ann, &S::bnn, &S::beta, &S::epss, &S::epsr, &S::sitas, &S::sitar, &S::sd, &S::sw,
                &S::kse, &S::c1, &S::c2, &S::c3, &S::e, &S::ds}},
        {"AM", {&S::poi, &S::dens, &S::eps0, &S::zmf, &S::sme0, &S::fs, &S::ramda, &S::kapa, &S::damp1, &S::damp2,
                &S::cop, &S::mzm, &S::azm, &S::r0, &S::bl, &S::bz, &S::zet, &S::e, &S::ds}},
        {"BU", {&S::kapa, &S::poi, &S::dens, &S::c, &S::fai, &S::beta, &S::fs, &S::damp1, &S::damp2, &S::cop,
                &S::eps0, &S::e, &S::ds}},
        {"PP", {&S::kapa, &S::poi, &S::dens, &S::damp1, &S::damp2, &S::cop, &S::eps0, &S::e, &S::ds}},
        {"VE", {&S::kapa, &S::poi, &S::dens, &S::c, &S::damp1, &S::damp2, &S::cop, &S::eps0, &S::e, &S::ds}},
        {"FC", {&S::dens, &S::c, &S::eps0, &S::cop, &S::sme0, &S::g0, &S::poi, &S::zmf, &S::ramda, &S::etao, &S::zet,
                &S::ann, &S::bnn, &S::damp1, &S::damp2, &S::ds}},
        {"EP", {&S::kapa, &S::poi, &S::dens, &S::c, &S::fai, &S::fs, &S::damp1, &S::damp2, &S::cop, &S::eps0, &S::e,
                &S::ds}},
    };
    return b;
}

// initial_pro (CPP_2-initialf.cpp:22-290) for the sections this library's flag set uses; the
// vibration section is read by sphg_read_vibration, the rain and excavation sections belong to
// flags it refuses (flr, fleob).
void read_problem(FILE *f, sphg_problem &p) {
    problem_defaults(p);
    if (find_section(f, "FL", false))
        read_vals(f, {I(p.fla), I(p.fls), I(p.flr), I(p.fli), I(p.flv), I(p.fleob), I(p.flts), I(p.flspv),
                      I(p.flknl), I(p.ntd), I(p.flbndy)});
    if (find_section(f, "PR", false))
        read_vals(f, {I(p.ndim), D(p.threshold_min), D(p.threshold_max), I(p.fop), I(p.inip), I(p.loop),
                      I(p.ntotal), I(p.step_regu), I(p.dttype), D(p.dr), D(p.dt)});
    if (find_section(f, "XP", false)) read_vals(f, {D(p.xsph[0]), D(p.xsph[1]), D(p.xsph[2])});
    if (find_section(f, "AT", false))
        read_vals(f, {D(p.art_vis[0][0]), D(p.art_vis[0][1]), D(p.art_vis[1][0]), D(p.art_vis[1][1]),
                      D(p.art_stre_coe), D(p.art_vis[2][0]), D(p.art_vis[2][1])});
    if (find_section(f, "GF", false)) read_vals(f, {D(p.gx), D(p.gy), D(p.gz)});
    if (find_section(f, "IP", false))
        read_vals(f, {D(p.interc), I(p.permtype), D(p.pers), D(p.perl), D(p.perm), I(p.ss_type), D(p.ss_myu),
                      D(p.ss_zeta), I(p.ss_acc_type)});
    if (find_section(f, "BP", false) && read_vals(f, {I(p.if_cal), I(p.if_move), D(p.coe_bndy)}) && p.if_move >= 1)
        read_vals(f, {D(p.move_vx[0]), D(p.move_vx[1]), D(p.move_vx[2])});
    p.tt = p.dt * (p.loop + 1);  // CPP_2-initialf.cpp:277
}

// initial_mat (CPP_2-initialf.cpp:325-852); returns false for more soil layers than slots
bool read_materials(FILE *f, sphg_problem &p, sphg_materials &m) {
    material_defaults(m);
    if (find_section(f, "WP", true)) {
        sphg_fluid *w = m.water;
        read_vals(f, {D(w[0].dens), D(w[0].vis), D(w[0].fai), D(w[0].c), D(w[0].vpu), I(w[0].press_type),
                      I(w[0].acc_type), I(w[0].bhtype), D(w[0].pre_coe), D(w[0].strain_min)});
        for (int k = 1; k < 3; k++) read_vals(f, {D(w[k].dens), D(w[k].vis), D(w[k].fai), D(w[k].c), D(w[k].vpu)});
    }
    for (int k = 1; k < 3; k++) {  // the other two waters share these with the first
        m.water[k].press_type = m.water[0].press_type;
        m.water[k].bhtype = m.water[0].bhtype;
        m.water[k].pre_coe = m.water[0].pre_coe;
        m.water[k].strain_min = m.water[0].strain_min;
    }
    if (p.fla == 1 && find_section(f, "AP", true))
        read_vals(f, {D(m.air.dens), D(m.air.paa0), D(m.air.vis), D(m.air.vpu), I(m.air.press_type),
                      I(m.air.acc_type), D(m.air.pre_coe)});
    double damp_coe = 0.0;
    int *flags[10] = {&p.flagem, &p.flagsb, &p.flagdp, &p.flagum, &p.flagam,
                      &p.flagbui, &p.flagpl, &p.flagve, &p.flagfocc, &p.flagep};
    if (p.fls == 1 && find_section(f, "MS", false)) {
        bool ok = true;
        for (int *fl : flags) ok = ok && read_vals(f, {I(*fl)});
        if (ok && p.flts >= 1) read_vals(f, {D(p.alphats), D(damp_coe)});
    }
    const auto &blocks = model_blocks();
    for (int model = 0; model < 10; model++) {
        const int layers = *flags[model];
        if (p.fls != 1 || layers < 1) continue;
        if (layers * 10 > SPHG_NSOIL) return false;
        if (!find_section(f, blocks[model].tag, false)) continue;
        for (int lay = 0; lay < layers; lay++) {
            sphg_soil &s = m.soil[lay * 10 + model];
            for (double sphg_soil::*fld : blocks[model].fields)
                if (!read_vals(f, {D(s.*fld)})) break;
            if (model == 8) s.e = 5.0e6;  // FOCC (CPP_2-initialf.cpp:738)
        }
    }
    if (find_section(f, "SM", false))
        read_vals(f, {D(m.structure.dens), D(m.structure.eps0), D(m.structure.cop), D(m.structure.e),
                      D(m.structure.poi)});
    for (sphg_soil &s : m.soil) {
        s.vpu = 0600.0;
        s.damp_coe = damp_coe;
    }
    m.structure.vpu = 0600.0;
    double por = 1.0;
    if (p.fls != 0) {
        for (sphg_soil &s : m.soil) s.porosity = s.eps0 / (1 + s.eps0);
        por = 0.0;
        for (const sphg_soil &s : m.soil) por = fmax(por, s.porosity);
    }
    for (sphg_fluid *w : {&m.air, &m.water[0], &m.water[1], &m.water[2]}) w->porosity = por;
    m.structure.porosity = m.structure.eps0 / (1 + m.structure.eps0);
    for (int k = 0; k < 3; k++)
        if (m.water[k].dens <= 0.0) m.water[k].dens = 1000.0;
    if (m.air.dens <= 0.0) m.air.dens = 1.293;
    return true;
}

}  // namespace

extern "C" int sphg_read_input(const char *path, sphg_problem *pro, sphg_materials *mat, int *n, int cap,
                               double *xp, double *vxp, int *type, int *matype, int *permtype, int *etype) {
    if (!path || !pro || !mat || !n) return SPHG_E_ARG;
    FILE *f = fopen(path, "r");
    if (!f) return SPHG_E_ARG;
    sphg_problem p;
    sphg_materials m;
    read_problem(f, p);
    if (!read_materials(f, p, m)) {
        fclose(f);
        return SPHG_E_ARG;
    }
    const int nt = p.ntotal;
    if (nt <= 0) {  // the reference stops with error 2 (CPP_1-SPHLoops.cpp:376)
        fclose(f);
        return SPHG_E_NOPARTICLES;
    }
    *n = nt;
    const bool want = xp && vxp && type && matype && permtype && etype;
    if (!want) {
        fclose(f);
        *pro = p;
        *mat = m;
        return SPHG_OK;
    }
    if (cap < nt) {
        fclose(f);
        return SPHG_E_CAPACITY;
    }
    // input_par (CPP_2-initialf.cpp:854-909): one row per particle, x, v, type, matype,
    // permtype, etype; 2D rows have two coordinates and two velocity components
    std::fill(xp, xp + 3 * (size_t)nt, 0.0);
    std::fill(vxp, vxp + 3 * (size_t)nt, 0.0);
    const int dim = p.ndim == 2 ? 2 : 3;  // as the reference: any ndim but 2 reads 3D rows
    bool ok = find_section(f, "PA", false);
    for (int i = 0; ok && i < nt; i++) {
        double *x = xp + 3 * (size_t)i, *v = vxp + 3 * (size_t)i;
        for (int k = 0; ok && k < dim; k++) ok = read_vals(f, {D(x[k])});
        for (int k = 0; ok && k < dim; k++) ok = read_vals(f, {D(v[k])});
        ok = ok && read_vals(f, {I(type[i]), I(matype[i]), I(permtype[i]), I(etype[i])});
    }
    fclose(f);
    if (!ok) return SPHG_E_ARG;  // the reference would go on with zeroed particles
    // Cell_Con box and phase extents (:911-1017): all particles widened by 2*dr in the first ndim
    // axes; water and soil extents over their particles, the maximum taking fmax(min, x) as the
    // reference does
    double bmax[3], bmin[3];
    for (int j = 0; j < 3; j++) {
        bmax[j] = -10000.0;
        bmin[j] = 10000.0;
        p.water_min[j] = p.soil_min[j] = 100000.000;
        p.water_max[j] = p.soil_max[j] = -100000.000;
    }
    for (int i = 0; i < nt; i++) {
        const double *x = xp + 3 * (size_t)i;
        for (int j = 0; j < 3; j++) {
            bmax[j] = fmax(bmax[j], x[j]);
            bmin[j] = fmin(bmin[j], x[j]);
        }
        double *mn = type[i] == 1 ? p.water_min : type[i] == 2 ? p.soil_min : nullptr;
        double *mx = type[i] == 1 ? p.water_max : type[i] == 2 ? p.soil_max : nullptr;
        if (mn)
            for (int j = 0; j < 3; j++) {
                mn[j] = fmin(mn[j], x[j]);
                mx[j] = fmax(mn[j], x[j]);
            }
    }
    for (int j = 0; j < p.ndim && j < 3; j++) {
        bmax[j] = bmax[j] + 2.0 * p.dr;
        bmin[j] = bmin[j] - 2.0 * p.dr;
    }
    p.xmin = bmin[0]; p.xmax = bmax[0];
    p.ymin = bmin[1]; p.ymax = bmax[1];
    p.zmin = bmin[2]; p.zmax = bmax[2];
    *pro = p;
    *mat = m;
    return SPHG_OK;
}

// random_field_input (CPP_17-1_K-L_expansion_non.cpp:718-772): <dir>/Parameters.txt, a header line
// then one line per particle in particle order; for water and air the 5th and 6th values are the
// fluid fai and c, for soil and structure the 5th to 8th are fai, c, cop and ds; any other
// particle's line is skipped with the reference's two-character scan, whose first character is
// the newline the previous numeric line left (so one line is skipped either way). The reference
// scans before testing the file pointer; a missing file is SPHG_E_ARG here.
extern "C" int sphg_read_parameters(const char *dir, int n, const int *type, double *values) {
    if (!dir || !type || !values || n <= 0) return SPHG_E_ARG;
    std::string path = std::string(dir) + "/Parameters.txt";
    FILE *f = fopen(path.c_str(), "r");
    if (!f) return SPHG_E_ARG;
    char c[2];
    int id, ty, ma;
    double x[3], tmp[2];
    const size_t N = (size_t)n;
    int err = fscanf(f, "%c%c%*[^\n]%*c", &c[0], &c[1]);
    for (int i = 0; i < n; i++) {
        const int t = type[i];
        if (t == 1 || t == 3)
            err = fscanf(f, "%d%lf%lf%lf%lf%lf%lf%lf%d%d", &id, &x[0], &x[1], &x[2], &values[4 * N + i],
                         &values[5 * N + i], &tmp[0], &tmp[1], &ty, &ma);
        else if (t == 2 || t == 4)
            err = fscanf(f, "%d%lf%lf%lf%lf%lf%lf%lf%d%d", &id, &x[0], &x[1], &x[2], &values[i], &values[N + i],
                         &values[2 * N + i], &values[3 * N + i], &ty, &ma);
        else
            err = fscanf(f, "%c%c%*[^\n]%*c", &c[0], &c[1]);
    }
    (void)err;
    fclose(f);
    return SPHG_OK;
}

// The RP section of <dir>/input.dat (CPP_2-initialf.cpp:180-201): rtype, nrain, nhor,
// press_factor, drop_velocity after Para_Rain's defaults (Class_Problem.cpp:74-84); nhor 0 is 35.
extern "C" int sphg_read_rain(const char *dir, sphg_rain *r) {
    if (!dir || !r) return SPHG_E_ARG;
    FILE *f = fopen((std::string(dir) + "/input.dat").c_str(), "r");
    if (!f) return SPHG_E_ARG;
    *r = sphg_rain{};
    if (find_section(f, "RP", false))
        read_vals(f, {I(r->rtype), I(r->nrain), I(r->nhor), D(r->press_factor), D(r->drop_velocity)});
    fclose(f);
    if (r->nhor == 0) r->nhor = 35;
    return SPHG_OK;
}

// The EC section of <dir>/input.dat (CPP_2-initialf.cpp:125-157): ecount, then the first and last
// particle of each stage. The reference never assigns a stage's step (its table is uninitialised
// there); this reader gives step 0, at which no stage fires. ecp holds 3 * cap ints (may be null:
// ecount only). Returns 8 (the reference's error) for no stage.
extern "C" int sphg_read_excavation(const char *dir, int *ecount, int *ecp, int cap) {
    if (!dir || !ecount) return SPHG_E_ARG;
    FILE *f = fopen((std::string(dir) + "/input.dat").c_str(), "r");
    if (!f) return SPHG_E_ARG;
    *ecount = 0;
    int rc = SPHG_OK;
    if (find_section(f, "EC", false) && read_vals(f, {I(*ecount)}) && *ecount >= 1) {
        if (ecp && cap < *ecount) rc = SPHG_E_CAPACITY;
        for (int s = 0; ecp && rc == SPHG_OK && s < *ecount; s++) {
            ecp[3 * s] = ecp[3 * s + 1] = ecp[3 * s + 2] = 0;
            read_vals(f, {I(ecp[3 * s]), I(ecp[3 * s + 1])});
        }
    } else {
        rc = 8;
    }
    fclose(f);
    return rc;
}

// The VL section of <dir>/input.dat (CPP_2-initialf.cpp:250-265: ttime, nsteps, sttime, edtime)
// and the loads of <dir>/vibra.dat (initial_vib, :293-322: nsteps rows of three values). loads may
// be null (then only the parameters are read) or must hold 3 * cap values.
extern "C" int sphg_read_vibration(const char *dir, double *ttime, int *nsteps, double *sttime, double *edtime,
                                   double *loads, int cap) {
    if (!dir || !ttime || !nsteps || !sttime || !edtime) return SPHG_E_ARG;
    FILE *f = fopen((std::string(dir) + "/input.dat").c_str(), "r");
    if (!f) return SPHG_E_ARG;
    *ttime = *sttime = *edtime = 0.0;  // Para_Vibra::Para_Vibra (Class_Problem.cpp:121-129)
    *nsteps = 0;
    const bool ok = find_section(f, "VL", false) && read_vals(f, {D(*ttime), I(*nsteps), D(*sttime), D(*edtime)});
    fclose(f);
    if (!ok || *nsteps < 1) return SPHG_E_ARG;  // the reference reports error 1 without loads
    if (!loads) return SPHG_OK;
    if (cap < *nsteps) return SPHG_E_CAPACITY;
    FILE *v = fopen((std::string(dir) + "/vibra.dat").c_str(), "r");
    if (!v) return SPHG_E_ARG;  // CPP_1-SPHLoops.cpp:306-330 stops with error 1
    for (int i = 0; i < *nsteps; i++) {
        // a short file leaves the rest of the table as it was; the reference's table is
        // uninitialised there, so it is zeroed first
        loads[3 * i] = loads[3 * i + 1] = loads[3 * i + 2] = 0.0;
    }
    for (int i = 0; i < *nsteps; i++)
        if (fscanf(v, "%lf%lf%lf", &loads[3 * i], &loads[3 * i + 1], &loads[3 * i + 2]) != 3) break;
    fclose(v);
    return SPHG_OK;
}

extern "C" int sphg_write_odb_fields(const char *dir, int l, double t, int n, const sphg_odb_fields *f) {
    if (!dir || !f || n < 0) return SPHG_E_ARG;
    const double *v3[] = {f->xp, f->ux, f->ax, f->vx, f->interf, f->interfss, f->strep};
    const double *v6[] = {f->sig, f->eps, f->adps};
    const double *v1[] = {f->pre, f->prea, f->prew, f->vsig, f->veps, f->meps, f->divq, f->divr, f->satu, f->suct};
    for (auto p : v3) if (!p) return SPHG_E_ARG;
    for (auto p : v6) if (!p) return SPHG_E_ARG;
    for (auto p : v1) if (!p) return SPHG_E_ARG;
    if (!f->type || !f->matype) return SPHG_E_ARG;
    FILE *fp = fopen(odb_name(dir, l).c_str(), "w");
    if (!fp) return SPHG_E_ARG;
    fputs(kRule, fp);
    fprintf(fp, "%10d %14.6e\n", l, t);
    for (int i = 0; i < n; i++) {
        const size_t i3 = 3 * (size_t)i, i6 = 6 * (size_t)i;
        const double *ux = f->ux + i3;
        double mag = ux[0] * ux[0] + ux[1] * ux[1] + ux[2] * ux[2];
        mag = sqrt(mag);
        // boundary rows carry the interaction forces where other rows carry ax and vx
        const bool bnd = f->type[i] == 0;
        const double *a = bnd ? f->interfss + i3 : f->ax + i3;
        const double *b = bnd ? f->interf + i3 : f->vx + i3;
        double v[47];
        int k = 0;
        for (int q = 0; q < 3; q++) v[k++] = f->xp[i3 + q];
        for (int q = 0; q < 3; q++) v[k++] = ux[q];
        v[k++] = mag;
        for (int q = 0; q < 3; q++) v[k++] = a[q];
        for (int q = 0; q < 3; q++) v[k++] = b[q];
        v[k++] = f->pre[i];
        v[k++] = f->prea[i];
        v[k++] = f->prew[i];
        for (int q = 0; q < 6; q++) v[k++] = f->sig[i6 + q];
        for (int q = 0; q < 3; q++) v[k++] = f->strep[i3 + q];
        for (int q = 0; q < 6; q++) v[k++] = f->eps[i6 + q];
        for (int q = 0; q < 6; q++) v[k++] = f->adps[i6 + q];
        for (const double *s : {f->vsig, f->veps, f->meps, f->divq, f->divr, f->satu, f->suct}) v[k++] = s[i];
        for (int q = 0; q < 3; q++) v[k++] = f->interfss[i3 + q];
        put_row(fp, i + 1, v, f->type[i], f->matype[i]);
    }
    fputs(kRule, fp);
    return fclose(fp) == 0 ? SPHG_OK : SPHG_E_ARG;
}

extern "C" int sphg_write_odb(sphg_ctx *ctx, const char *dir, int l, double t, int first, int count) {
    if (!ctx || !dir) return SPHG_E_ARG;
    const int n = sphg_particle_count(ctx);
    if (n <= 0 || first < 0 || count < 0 || first > n || count > n - first) return SPHG_E_ARG;
    const size_t N = (size_t)n;
    std::vector<double> d3[7], d6[3], d1[10];
    const char *n3[] = {"xp", "ux", "ax", "vx", "interf", "interfss", "strep"};
    const char *n6[] = {"sig", "eps", "adps"};
    const char *n1[] = {"pre", "prea", "prew", "vsig", "veps", "meps", "divq", "divr", "satu", "suct"};
    for (int k = 0; k < 7; k++) {
        d3[k].resize(3 * N);
        if (int rc = sphg_get_field(ctx, n3[k], d3[k].data(), n, 3)) return rc;
    }
    for (int k = 0; k < 3; k++) {
        d6[k].resize(6 * N);
        if (int rc = sphg_get_field(ctx, n6[k], d6[k].data(), n, 6)) return rc;
    }
    for (int k = 0; k < 10; k++) {
        d1[k].resize(N);
        if (int rc = sphg_get_field(ctx, n1[k], d1[k].data(), n, 1)) return rc;
    }
    std::vector<int> type(N), matype(N);
    if (int rc = sphg_get_field_i(ctx, "type", type.data(), n)) return rc;
    if (int rc = sphg_get_field_i(ctx, "matype", matype.data(), n)) return rc;
    const size_t o = (size_t)first;
    sphg_odb_fields f;
    f.xp = d3[0].data() + 3 * o; f.ux = d3[1].data() + 3 * o; f.ax = d3[2].data() + 3 * o;
    f.vx = d3[3].data() + 3 * o; f.interf = d3[4].data() + 3 * o; f.interfss = d3[5].data() + 3 * o;
    f.strep = d3[6].data() + 3 * o;
    f.sig = d6[0].data() + 6 * o; f.eps = d6[1].data() + 6 * o; f.adps = d6[2].data() + 6 * o;
    f.pre = d1[0].data() + o; f.prea = d1[1].data() + o; f.prew = d1[2].data() + o; f.vsig = d1[3].data() + o;
    f.veps = d1[4].data() + o; f.meps = d1[5].data() + o; f.divq = d1[6].data() + o; f.divr = d1[7].data() + o;
    f.satu = d1[8].data() + o; f.suct = d1[9].data() + o;
    f.type = type.data() + o;
    f.matype = matype.data() + o;
    return sphg_write_odb_fields(dir, l, t, count, &f);
}
