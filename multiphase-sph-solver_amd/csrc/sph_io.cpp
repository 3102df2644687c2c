// Host-side result files: the reference's per-output-step particle table (.odb), written from
// host arrays or from a context's device state.
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

}  // namespace

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
