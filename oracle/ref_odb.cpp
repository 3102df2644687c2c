// Test infrastructure only: runs the reference's own .odb writer, clOutput_Fun::output_total
// (CPP_14-2-output.cpp:15-135, compiled from the reference checkout by oracle/Makefile), on a
// particle state read from a flat binary file, so that the library's writer can be compared with
// it byte for byte (tests/golden/make_odb.py writes the input and keeps the output as a fixture).
//
// Input ("ODBI"): magic[4], int32 n, int32 l, double t, then for each particle 49 doubles
//   xp[3] ux[3] ax[3] vx[3] interf[3] interfss[3] strep[3] sig[6] eps[6] adps[6]
//   pre prea prew vsig veps meps divq divr satu suct
// followed by int32 type, int32 matype.
// Usage: ref_odb <input> <output directory>   (writes <dir>/data-<l>.odb as the reference does)
#include <cstdio>
#include <cstring>
#include <vector>
#include "Class_Functions.h"

int main(int argc, char **argv) {
    if (argc != 3) {
        fprintf(stderr, "usage: %s <input> <outdir>\n", argv[0]);
        return 2;
    }
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 3;
    char magic[4];
    int n = 0, l = 0;
    double t = 0.0;
    if (fread(magic, 1, 4, f) != 4 || memcmp(magic, "ODBI", 4) != 0 || fread(&n, 4, 1, f) != 1 ||
        fread(&l, 4, 1, f) != 1 || fread(&t, 8, 1, f) != 1 || n <= 0)
        return 4;
    std::vector<Particle> par(n);
    for (int i = 0; i < n; i++) {
        double v[49];
        int ti[2];
        if (fread(v, 8, 49, f) != 49 || fread(ti, 4, 2, f) != 2) return 5;
        Particle &p = par[i];
        const double *q = v;
        for (double *dst : {p.xp, p.ux, p.ax, p.vx, p.interf, p.interfss, p.strep})
            for (int k = 0; k < 3; k++) dst[k] = *q++;
        for (double *dst : {p.sig, p.eps, p.adps})
            for (int k = 0; k < 6; k++) dst[k] = *q++;
        for (double *dst : {&p.pre, &p.prea, &p.prew, &p.vsig, &p.veps, &p.meps, &p.divq, &p.divr, &p.satu, &p.suct})
            *dst = *q++;
        p.type = ti[0];
        p.matype = ti[1];
    }
    fclose(f);
    Para_Pro pro;
    pro.l = l;
    pro.t = t;
    pro.ntotal = n;
    clOutput_Fun out;
    out.output_total(par.data(), pro, argv[2]);
    return 0;
}
