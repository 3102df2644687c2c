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

int main(int argc, char **argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s <project-dir-with-input.dat>\n", argv[0]);
        return 2;
    }
    return SPH_Calculation(3, argv[1]);
}
