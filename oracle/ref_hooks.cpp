// Test infrastructure only: the oracle's link harness around the reference solver.
//
// This translation unit is linked with the reference's own sources (compiled where they lie,
// under the reference checkout, by oracle/Makefile) in place of the reference's output and
// parameter-report units (CPP_14-1-outmgf.cpp, CPP_14-2-output.cpp, CPP_18-spatialoutput.cpp).
// The reference driver SPH_Calculation (CPP_1-SPHLoops.cpp:67) then runs unmodified; the hooks
// below replace its 7-digit text output (CPP_14-2-output.cpp:15-135) by full-precision binary
// state dumps so the CPU restatement and the GPU path can be compared bit for bit.
//
// Dump format ("SPHD"): magic[4], int32 n, int32 step, double t, int32 nfields, then per field:
// char name[16], int32 ncomp, int32 dtype (0 = float64, 1 = int32), n*ncomp values (row major).
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <vector>
#include <string>
#include "Class_Functions.h"
#include "Class_spatial_variability.h"

namespace {

struct Writer {
    FILE *f;
    int n;
    std::vector<char> body;
    int nfields = 0;
    void put(const void *p, size_t sz) {
        const char *c = (const char *)p;
        body.insert(body.end(), c, c + sz);
    }
    template <class Get>
    void fd(const char *name, int ncomp, Get get) {
        char nm[16] = {0};
        strncpy(nm, name, 15);
        put(nm, 16);
        int dt = 0;
        put(&ncomp, 4);
        put(&dt, 4);
        for (int i = 0; i < n; i++)
            for (int k = 0; k < ncomp; k++) {
                double v = get(i, k);
                put(&v, 8);
            }
        nfields++;
    }
    template <class Get>
    void fi(const char *name, Get get) {
        char nm[16] = {0};
        strncpy(nm, name, 15);
        put(nm, 16);
        int one = 1, dt = 1;
        put(&one, 4);
        put(&dt, 4);
        for (int i = 0; i < n; i++) {
            int v = get(i);
            put(&v, 4);
        }
        nfields++;
    }
    void finish(int step, double t) {
        fwrite("SPHD", 1, 4, f);
        fwrite(&n, 4, 1, f);
        fwrite(&step, 4, 1, f);
        fwrite(&t, 8, 1, f);
        fwrite(&nfields, 4, 1, f);
        fwrite(body.data(), 1, body.size(), f);
        fclose(f);
    }
};

void dump_particles(Writer &w, Particle *p) {
#define V3(nm) w.fd(#nm, 3, [&](int i, int k) { return p[i].nm[k]; })
#define V6(nm) w.fd(#nm, 6, [&](int i, int k) { return p[i].nm[k]; })
#define M33(nm) w.fd(#nm, 9, [&](int i, int k) { return p[i].nm[k / 3][k % 3]; })
#define S1(nm) w.fd(#nm, 1, [&](int i, int) { return p[i].nm; })
#define I1(nm) w.fi(#nm, [&](int i) { return p[i].nm; })
    V3(x); V3(xp); V3(vx); V3(vxp); V3(ax); V3(axp); V3(ux); V3(strep);
    M33(divx); V3(interf); V3(interfss);
    S1(rho); S1(rhop); S1(mass); S1(hl); S1(pre); S1(prep); S1(e);
    V6(eps); V6(deps); V6(sig); V6(dsig); V3(ff_sig); M33(weps); V6(adps);
    V6(tssig); S1(tsqc);
    S1(satu); S1(dsatu); S1(suct); S1(cop); S1(vsig); S1(veps); S1(divq); S1(divr); S1(meps);
    S1(roue); S1(rous); S1(zeta); S1(rsta); S1(prew); S1(prea); S1(porosity); S1(cd); S1(poro_cop);
    M33(beta); M33(beta_ts);
    I1(type); I1(matype); I1(permtype); I1(etype);
#undef V3
#undef V6
#undef M33
#undef S1
#undef I1
}

FILE *open_dump(const char *dir, const char *name) {
    std::string path = std::string(dir) + "/" + name;
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) {
        fprintf(stderr, "ref harness: cannot write %s\n", path.c_str());
        exit(90);
    }
    return f;
}

}  // namespace

// Particle state after each output step (CPP_1-SPHLoops.cpp:762-773 calls this at l % fop == 0 or l == 1).
void clOutput_Fun::output_total(Particle *pVelPar, const Para_Pro &pPPro, char *argv) {
    char name[64];
    snprintf(name, sizeof name, "dump_%06d.bin", pPPro.l);
    Writer w{open_dump(argv, name), pPPro.ntotal};
    dump_particles(w, pVelPar);
    w.finish(pPPro.l, pPPro.t);
}

void clOutput_Fun::mgf(Particle *, const Para_Pro &, const Cell_Con &, FILE *) {}
void clOutput_Fun::output_impact(FILE *, Particle *, const Para_Pro &, char *) {}
void clOutput_Fun::output_flowdist(FILE *, Particle *, const Para_Pro &) {}
void clOutput_Fun::output_status(int, int, int) {}

// Initial state right after setting_par (CPP_1-SPHLoops.cpp:440-443): particles plus the
// per-particle material copies pVFluid / pParti_ConsPara.
void clFParaOutput_Fun::paramters_output(Particle *pPar, Para_Fluid *pVFluid, Para_Soil *pS,
                                         const Para_Pro &pPPro, char *argv) {
    Writer w{open_dump(argv, "dump_init.bin"), pPPro.ntotal};
    dump_particles(w, pPar);
#define PS(nm) w.fd("ps_" #nm, 1, [&](int i, int) { return pS[i].nm; })
    PS(e); PS(poi); PS(eps0); PS(dens); PS(damp1); PS(damp2); PS(cop); PS(ramda); PS(kapa);
    PS(zmf); PS(ann); PS(bnn); PS(beta); PS(sitas); PS(sitar); PS(epsr); PS(epss); PS(sd); PS(sw);
    PS(kse); PS(c1); PS(c2); PS(c3); PS(fai); PS(c); PS(ocr); PS(sme0); PS(mzm); PS(azm); PS(r0);
    PS(bl); PS(bz); PS(zet); PS(vpu); PS(porosity); PS(fs); PS(etao); PS(g0); PS(ds); PS(damp_coe);
#undef PS
#define PF(nm) w.fd("pf_" #nm, 1, [&](int i, int) { return (double)pVFluid[i].nm; })
    PF(vis); PF(vpu); PF(dens); PF(paa0); PF(porosity); PF(fai); PF(c); PF(bhtype);
    PF(press_type); PF(acc_type); PF(pre_coe); PF(strain_min);
#undef PF
    w.finish(0, 0.0);
}

int SPH_Calculation(int argc, char *argv);

// K-L mode: runs the reference's solve_wi, solve_ramdai and solve_h_2D/3D
// (CPP_17-1_K-L_expansion_non.cpp:268-713) on injected standard normals.
// kl_in.bin: int32 ndim, n; double a[3], rl[3], x_trans[3]; int32 dist; double mean, sd,
//            coe_depth, soil_max, water_max; int32 field; double uu[200]; double xp[n][3];
//            int32 type[n]; double value[n] (the field before the synthesis)
// kl_out.bin: double wi[200][3], eigval[200]; int32 order[200][3]; double value[n]; int32 rc
int kl_mode(const char *dir) {
    std::string in = std::string(dir) + "/kl_in.bin", out = std::string(dir) + "/kl_out.bin";
    FILE *f = fopen(in.c_str(), "rb");
    if (!f) return 2;
    int ndim, n, dist, field;
    double a[3], rl[3], xt[3], mean, sd, cdep, smax, wmax, uu[200];
    bool ok = fread(&ndim, 4, 1, f) == 1 && fread(&n, 4, 1, f) == 1 && fread(a, 8, 3, f) == 3 &&
              fread(rl, 8, 3, f) == 3 && fread(xt, 8, 3, f) == 3 && fread(&dist, 4, 1, f) == 1 &&
              fread(&mean, 8, 1, f) == 1 && fread(&sd, 8, 1, f) == 1 && fread(&cdep, 8, 1, f) == 1 &&
              fread(&smax, 8, 1, f) == 1 && fread(&wmax, 8, 1, f) == 1 && fread(&field, 4, 1, f) == 1 &&
              fread(uu, 8, 200, f) == 200;
    std::vector<double> xp((size_t)n * 3), val(n);
    std::vector<int> type(n);
    ok = ok && fread(xp.data(), 8, xp.size(), f) == xp.size() && fread(type.data(), 4, n, f) == (size_t)n &&
         fread(val.data(), 8, n, f) == (size_t)n;
    fclose(f);
    if (!ok) return 2;
    Particle *par = new Particle[n];
    Para_Soil *ps = new Para_Soil[n];
    Para_Fluid *pf = new Para_Fluid[n];
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < 3; k++) par[i].xp[k] = xp[3 * i + k];
        par[i].type = type[i];
        double v = val[i];
        if (field == 0) ps[i].fai = v;
        if (field == 1) ps[i].c = v;
        if (field == 2) ps[i].cop = v;
        if (field == 3) ps[i].ds = v;
        if (field == 4) pf[i].fai = v;
        if (field == 5) pf[i].c = v;
    }
    Para_Pro pro;
    pro.ntotal = n;
    pro.ndim = ndim;
    Para_SpatialVariability psv;
    psv.type = dist;
    psv.mean_value = mean;
    psv.sd_value = sd;
    psv.coe_depth = cdep;
    for (int k = 0; k < 3; k++) {
        psv.rl_value[k] = rl[k];
        psv.a[k] = a[k];
    }
    double wi[200][3] = {{0}}, eig[200];
    int order[200][3];
    clFSpatialVariability_Fun fn;
    fn.solve_wi(wi, psv.a, psv.rl_value, ndim);
    int rc = fn.solve_ramdai(eig, wi, psv.rl_value, order, ndim);
    if (!rc) {
        if (ndim == 2) rc = fn.solve_h_2D(par, pf, ps, pro, wi, eig, uu, xt, order, psv, smax, wmax, field);
        else rc = fn.solve_h_3D(par, pf, ps, pro, wi, eig, uu, xt, order, psv, smax, wmax, field);
    }
    for (int i = 0; i < n; i++) {
        double v = val[i];
        if (field == 0) v = ps[i].fai;
        if (field == 1) v = ps[i].c;
        if (field == 2) v = ps[i].cop;
        if (field == 3) v = ps[i].ds;
        if (field == 4) v = pf[i].fai;
        if (field == 5) v = pf[i].c;
        val[i] = v;
    }
    FILE *o = fopen(out.c_str(), "wb");
    if (!o) return 2;
    fwrite(wi, 8, 600, o);
    fwrite(eig, 8, 200, o);
    fwrite(order, 4, 600, o);
    fwrite(val.data(), 8, n, o);
    fwrite(&rc, 4, 1, o);
    fclose(o);
    delete[] par;
    delete[] ps;
    delete[] pf;
    return 0;
}

// The reference's Cholesky of the correlation-matrix random fields (CPP_17-3-Nonstationary-RF.cpp:45-74),
// an inline function of that unit; oracle/Makefile compiles the unit with -fkeep-inline-functions
// so that its out-of-line copy exists to be called here.
void Cholesky(double *A, double *L, int n);

// Cholesky mode: chol_in.bin: int32 n; double A[n][n] -> chol_out.bin: double L[n][n]
int chol_mode(const char *dir) {
    std::string in = std::string(dir) + "/chol_in.bin", out = std::string(dir) + "/chol_out.bin";
    FILE *f = fopen(in.c_str(), "rb");
    if (!f) return 2;
    int n = 0;
    bool ok = fread(&n, 4, 1, f) == 1 && n > 0;
    std::vector<double> A(ok ? (size_t)n * n : 0), L(A.size(), 0.0);
    ok = ok && fread(A.data(), 8, A.size(), f) == A.size();
    fclose(f);
    if (!ok) return 2;
    Cholesky(A.data(), L.data(), n);
    FILE *o = fopen(out.c_str(), "wb");
    if (!o) return 2;
    fwrite(L.data(), 8, L.size(), o);
    fclose(o);
    return 0;
}

// Random-field modes on given particles, kind 0..4: --rf runs the reference's
// nonstationary_random_2D/3D (CPP_17-3-Nonstationary-RF.cpp:76-849), --rfc its
// nonstationary_correlated_random_2D/3D (CPP_17-5-Nonstationary-Correlated-RF.cpp:87-737), --kln
// SpatialVariables_Generate_2D/3D (CPP_17-1_K-L_expansion_non.cpp:63-265), --klc
// SpatialVariables_Generate_2D_corr/3D_corr (CPP_17-2_K-L_expansion_corr.cpp:55-473), --rv
// random_variable_generate (CPP_17-5-Nonstationary-Correlated-RF.cpp:738-863) and --rfi
// random_field_input (CPP_17-1_K-L_expansion_non.cpp:718-772, <dir>/Parameters.txt); every kind but --rf
// reads the coefficient coe_frict_cohesion after the properties. Their standard normals come from
// a clock-seeded generator, so the outputs are reproducible only where sd_value is 0 (or where
// the values do not depend on them). The K-L kinds append their parameters to
// Para_Spatial_Dist.txt in the working directory.
// rf_in.bin: int32 ndim, n; double dr, box_nb[6] (xmin xmax ymin ymax zmin zmax of the non-boundary
//            particles), soil_max[3], water_max[3]; int32 flags[6]; double cell_ndx; int32 flag_auto;
//            double beta, beta_1; 6 x {int32 type; double mean, sd, rl[3], coe_depth};
//            double xp[n][3]; int32 type[n]; double fai[n], c[n], cop[n], ds[n], fai_w[n], c_w[n]
// rf_out.bin: double fai[n], c[n], cop[n], ds[n], fai_w[n], c_w[n]; int32 rc; double a[6][3]
int rf_mode(const char *dir, int kind) {
    std::string in = std::string(dir) + "/rf_in.bin", out = std::string(dir) + "/rf_out.bin";
    FILE *f = fopen(in.c_str(), "rb");
    if (!f) return 2;
    int ndim, n, flags[6], fauto;
    double dr, box[6], smax[3], wmax[3], cndx, beta, beta1;
    bool ok = fread(&ndim, 4, 1, f) == 1 && fread(&n, 4, 1, f) == 1 && fread(&dr, 8, 1, f) == 1 &&
              fread(box, 8, 6, f) == 6 && fread(smax, 8, 3, f) == 3 && fread(wmax, 8, 3, f) == 3 &&
              fread(flags, 4, 6, f) == 6 && fread(&cndx, 8, 1, f) == 1 && fread(&fauto, 4, 1, f) == 1 &&
              fread(&beta, 8, 1, f) == 1 && fread(&beta1, 8, 1, f) == 1 && n > 0;
    Para_SpatialVariability psv[6];
    for (int k = 0; ok && k < 6; k++) {
        double v[6];
        ok = fread(&psv[k].type, 4, 1, f) == 1 && fread(v, 8, 6, f) == 6;
        psv[k].no = k;
        psv[k].mean_value = v[0];
        psv[k].sd_value = v[1];
        for (int q = 0; q < 3; q++) psv[k].rl_value[q] = v[2 + q];
        psv[k].coe_depth = v[5];
    }
    std::vector<double> xp(ok ? (size_t)n * 3 : 0), vals(ok ? (size_t)n * 6 : 0);
    std::vector<int> type(ok ? n : 0);
    ok = ok && fread(xp.data(), 8, xp.size(), f) == xp.size() && fread(type.data(), 4, n, f) == (size_t)n &&
         fread(vals.data(), 8, vals.size(), f) == vals.size();
    double coe = 0.0;
    if (kind != 0) ok = ok && fread(&coe, 8, 1, f) == 1;
    fclose(f);
    if (!ok) return 2;
    Particle *par = new Particle[n];
    Para_Soil *ps = new Para_Soil[n];
    Para_Fluid *pf = new Para_Fluid[n];
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < 3; k++) par[i].xp[k] = xp[3 * i + k];
        par[i].type = type[i];
        ps[i].fai = vals[i];
        ps[i].c = vals[n + i];
        ps[i].cop = vals[2 * (size_t)n + i];
        ps[i].ds = vals[3 * (size_t)n + i];
        pf[i].fai = vals[4 * (size_t)n + i];
        pf[i].c = vals[5 * (size_t)n + i];
    }
    Para_Pro pro;
    pro.ntotal = n;
    pro.ndim = ndim;
    pro.dr = dr;
    Flag_SpatialVariability fsv;
    fsv.flag_fai = flags[0]; fsv.flag_c = flags[1]; fsv.flag_cop = flags[2]; fsv.flag_ds = flags[3];
    fsv.flag_faiw = flags[4]; fsv.flag_cw = flags[5];
    fsv.cell_ndx = cndx;
    fsv.flag_auto = fauto;
    fsv.beta = beta;
    fsv.beta_1 = beta1;
    fsv.coe_frict_cohesion = coe;
    Cell_Con cc;
    cc.xmin_nb = box[0]; cc.xmax_nb = box[1]; cc.ymin_nb = box[2]; cc.ymax_nb = box[3];
    cc.zmin_nb = box[4]; cc.zmax_nb = box[5];
    for (int k = 0; k < 3; k++) {
        cc.soil_max[k] = smax[k];
        cc.water_max[k] = wmax[k];
    }
    clFSpatialVariability_Fun fn;
    int rc;
    if (kind == 1)
        rc = ndim == 2 ? fn.nonstationary_correlated_random_2D(par, pf, ps, pro, fsv, psv, cc)
                       : fn.nonstationary_correlated_random_3D(par, pf, ps, pro, fsv, psv, cc);
    else if (kind == 2)
        rc = ndim == 2 ? fn.SpatialVariables_Generate_2D_corr(par, pf, ps, pro, fsv, psv, cc)
                       : fn.SpatialVariables_Generate_3D_corr(par, pf, ps, pro, fsv, psv, cc);
    else if (kind == 3)
        rc = ndim == 2 ? fn.SpatialVariables_Generate_2D(par, pf, ps, pro, fsv, psv, cc)
                       : fn.SpatialVariables_Generate_3D(par, pf, ps, pro, fsv, psv, cc);
    else if (kind == 4)
        rc = fn.random_variable_generate(par, pf, ps, pro, fsv, psv);
    else if (kind == 5)
        rc = fn.random_field_input(par, pf, ps, pro, (char *)dir);
    else
        rc = ndim == 2 ? fn.nonstationary_random_2D(par, pf, ps, pro, fsv, psv, cc)
                       : fn.nonstationary_random_3D(par, pf, ps, pro, fsv, psv, cc);
    for (int i = 0; i < n; i++) {
        vals[i] = ps[i].fai;
        vals[n + i] = ps[i].c;
        vals[2 * (size_t)n + i] = ps[i].cop;
        vals[3 * (size_t)n + i] = ps[i].ds;
        vals[4 * (size_t)n + i] = pf[i].fai;
        vals[5 * (size_t)n + i] = pf[i].c;
    }
    FILE *o = fopen(out.c_str(), "wb");
    if (!o) return 2;
    fwrite(vals.data(), 8, vals.size(), o);
    fwrite(&rc, 4, 1, o);
    for (int k = 0; k < 6; k++) fwrite(psv[k].a, 8, 3, o);  // the K-L half-widths the generators set
    fclose(o);
    delete[] par;
    delete[] ps;
    delete[] pf;
    return 0;
}

// Excavation-or-backfill mode: the reference's exca_or_back (CPP_16-EOB.cpp:13-74) for one stage.
// eob_in.bin: int32 n, start, end; int32 type[n], etype[n] -> eob_out.bin: int32 type[n], etype[n]
int eob_mode(const char *dir) {
    std::string in = std::string(dir) + "/eob_in.bin", out = std::string(dir) + "/eob_out.bin";
    FILE *f = fopen(in.c_str(), "rb");
    if (!f) return 2;
    int n = 0, se[2];
    bool ok = fread(&n, 4, 1, f) == 1 && n > 0 && fread(se, 4, 2, f) == 2;
    std::vector<int> ty(ok ? n : 0), et(ok ? n : 0);
    ok = ok && fread(ty.data(), 4, n, f) == (size_t)n && fread(et.data(), 4, n, f) == (size_t)n;
    fclose(f);
    if (!ok) return 2;
    Particle *par = new Particle[n];
    for (int i = 0; i < n; i++) {
        par[i].type = ty[i];
        par[i].etype = et[i];
    }
    Para_Pro pro;
    pro.ntotal = n;
    Para_EC ec;
    ec.ecount = 1;
    ec.memloads();
    ec.ecp[0][0] = se[0];
    ec.ecp[0][1] = se[1];
    ec.ecp[0][2] = 0;
    clExcBac_Fun fn;
    fn.exca_or_back(par, &pro, ec, 1, 0);
    for (int i = 0; i < n; i++) {
        ty[i] = par[i].type;
        et[i] = par[i].etype;
    }
    delete[] par;
    FILE *o = fopen(out.c_str(), "wb");
    if (!o) return 2;
    fwrite(ty.data(), 4, n, o);
    fwrite(et.data(), 4, n, o);
    fclose(o);
    return 0;
}

// Particle-deletion mode: the reference's partidel (CPP_15-2-partidel.cpp:16-50) once.
// del_in.bin: int32 n; double dr, xmin, xmax; double xp[n][3], vxp[n][3], vx[n][3], ax[n][3];
//             int32 type[n] -> del_out.bin: int32 type[n]; double vxp[n][3], vx[n][3], ax[n][3]
int del_mode(const char *dir) {
    std::string in = std::string(dir) + "/del_in.bin", out = std::string(dir) + "/del_out.bin";
    FILE *f = fopen(in.c_str(), "rb");
    if (!f) return 2;
    int n = 0;
    double g[3];
    bool ok = fread(&n, 4, 1, f) == 1 && n > 0 && fread(g, 8, 3, f) == 3;
    std::vector<double> xp(ok ? 3 * (size_t)n : 0), vxp(xp.size()), vx(xp.size()), ax(xp.size());
    std::vector<int> ty(ok ? n : 0);
    ok = ok && fread(xp.data(), 8, xp.size(), f) == xp.size() && fread(vxp.data(), 8, xp.size(), f) == xp.size() &&
         fread(vx.data(), 8, xp.size(), f) == xp.size() && fread(ax.data(), 8, xp.size(), f) == xp.size() &&
         fread(ty.data(), 4, n, f) == (size_t)n;
    fclose(f);
    if (!ok) return 2;
    Particle *par = new Particle[n];
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < 3; k++) {
            par[i].xp[k] = xp[3 * i + k];
            par[i].vxp[k] = vxp[3 * i + k];
            par[i].vx[k] = vx[3 * i + k];
            par[i].ax[k] = ax[3 * i + k];
        }
        par[i].type = ty[i];
    }
    Para_Pro pro;
    pro.ntotal = n;
    pro.dr = g[0];
    Cell_Con cc;
    cc.xmin = g[1];
    cc.xmax = g[2];
    Para_Rain rain;
    clRain_Fun fn;
    fn.partidel(par, &rain, &pro, cc, 1);
    for (int i = 0; i < n; i++) {
        ty[i] = par[i].type;
        for (int k = 0; k < 3; k++) {
            vxp[3 * i + k] = par[i].vxp[k];
            vx[3 * i + k] = par[i].vx[k];
            ax[3 * i + k] = par[i].ax[k];
        }
    }
    delete[] par;
    FILE *o = fopen(out.c_str(), "wb");
    if (!o) return 2;
    fwrite(ty.data(), 4, n, o);
    fwrite(vxp.data(), 8, vxp.size(), o);
    fwrite(vx.data(), 8, vx.size(), o);
    fwrite(ax.data(), 8, ax.size(), o);
    fclose(o);
    return 0;
}

int main(int argc, char **argv) {
    if (argc >= 3 && strcmp(argv[1], "--kl") == 0) return kl_mode(argv[2]);
    if (argc >= 3 && strcmp(argv[1], "--chol") == 0) return chol_mode(argv[2]);
    if (argc >= 3 && strcmp(argv[1], "--rf") == 0) return rf_mode(argv[2], 0);
    if (argc >= 3 && strcmp(argv[1], "--rfc") == 0) return rf_mode(argv[2], 1);
    if (argc >= 3 && strcmp(argv[1], "--klc") == 0) return rf_mode(argv[2], 2);
    if (argc >= 3 && strcmp(argv[1], "--kln") == 0) return rf_mode(argv[2], 3);
    if (argc >= 3 && strcmp(argv[1], "--rv") == 0) return rf_mode(argv[2], 4);
    if (argc >= 3 && strcmp(argv[1], "--rfi") == 0) return rf_mode(argv[2], 5);
    if (argc >= 3 && strcmp(argv[1], "--eob") == 0) return eob_mode(argv[2]);
    if (argc >= 3 && strcmp(argv[1], "--del") == 0) return del_mode(argv[2]);
    if (argc < 2) {
        fprintf(stderr, "usage: %s <project-dir-with-input.dat> | --kl <dir> | --chol <dir> | --rf[c] | --klc | --kln | --rv | --rfi | --eob | --del <dir>\n", argv[0]);
        return 2;
    }
    return SPH_Calculation(3, argv[1]);
}
