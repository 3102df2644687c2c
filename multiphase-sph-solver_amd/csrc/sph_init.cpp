// Host-side initial particle state: setting_vis_cons + setting_par (CPP_2-initialf.cpp:1021-1380).
// Runs once before the first step; kept on the host because it is a single pass of table
// look-ups, reductions over the phase extents and a few transcendental calls (log, tan) that
// the reference evaluates with the C library.
#include <cmath>
#include <vector>

#include "../../include/sphg.h"

namespace {

// moisturesuct (CPP_7-2-constitutive.cpp:324-338); pi as in Header_Parameters.h:10
double moisturesuct(double satu, double dsatu, const sphg_soil &s) {
    const double pi = 3.141592653;
    double temp = tan(pi * (s.sitas - satu) / (s.sitas - s.sitar) / 2.0);
    if (dsatu > 0.0) return log(1 + exp(s.c2 * s.sw) * temp) / s.c2;  // wetting curve
    return log(1 + exp(s.c1 * s.sd) * temp) / s.c1;
}

}  // namespace

extern "C" int sphg_setting_par(const sphg_problem *pro, const sphg_materials *mat, int n, const double *xp,
                                const int *type, const int *matype, const int *etype, double *rho, double *rhop,
                                double *mass, double *hl, double *satu, double *e, double *porosity, double *axp,
                                double *cd, double *roue, double *sig, double *pre, double *prep, double *suct) {
    if (!pro || !mat || n <= 0 || !xp || !type || !matype || !etype) return SPHG_E_ARG;
    const int dim = pro->ndim;
    const double dr = pro->dr;
    double hmaxw = -100000.0, hmaxs = -100000.0, hminw = 100000.0, hmins = 100000.0;
    double sitas = 0.0, sitar = 0.0;
    for (int i = 0; i < 30; i++) {
        sitas = fmax(sitas, mat->soil[i].sitas);
        sitar = fmax(sitar, mat->soil[i].sitar);
    }
    sphg_soil dsoil{};
    dsoil.fs = 1.0;
    dsoil.ds = 1.0;
    sphg_fluid dfl{};
    dfl.strain_min = 1.0;
    // setting_vis_cons: per-particle copies of the material rows
    auto soil_of = [&](int i) -> const sphg_soil & {
        int t = type[i], m = matype[i];
        if (t == 2) {
            if (m < 30 && m >= 0) return mat->soil[m];
            if (m >= 50) return mat->soil[m % 10];
        }
        if (t == 7 && etype[i] == 2 && m >= 0 && m < 30) return mat->soil[m];
        return dsoil;
    };
    auto fluid_of = [&](int i) -> const sphg_fluid & {
        int t = type[i], m = matype[i];
        if ((t == 1 || (t == 7 && etype[i] == 1)) && m >= 0 && m < 3) return mat->water[m];
        if (t == 3) return mat->air;
        return dfl;
    };
    auto massini = [&](int i) { mass[i] = dim == 2 ? rhop[i] * hl[i] * hl[i] : rhop[i] * hl[i] * hl[i] * hl[i]; };
    const double g3[3] = {pro->gx, pro->gy, pro->gz};
    for (int i = 0; i < n; i++) {
        rho[i] = rhop[i] = mass[i] = hl[i] = satu[i] = e[i] = porosity[i] = cd[i] = roue[i] = 0.0;
        pre[i] = prep[i] = suct[i] = 0.0;
        for (int k = 0; k < 3; k++) axp[3 * i + k] = 0.0;
        for (int k = 0; k < 6; k++) sig[6 * i + k] = 0.0;
    }
    for (int i = 0; i < n; i++) {
        const int t = type[i], m = matype[i], et = etype[i];
        if (t == 0) {
            rhop[i] = 1500.0;
            satu[i] = sitas;
            hl[i] = dr;
            massini(i);
            if (m == 5)
                for (int k = 0; k < 3; k++) axp[3 * i + k] = g3[k];
        } else if (t == 1 || (t == 7 && et == 1)) {
            const sphg_fluid &f = fluid_of(i);
            rhop[i] = f.porosity * f.dens;
            satu[i] = sitas;
            hl[i] = dr;
            massini(i);
            for (int k = 0; k < 3; k++) axp[3 * i + k] = g3[k];
            porosity[i] = f.porosity;
            if (t == 1) {
                hmaxw = fmax(hmaxw, xp[3 * i + dim - 1]);
                hminw = fmin(hminw, xp[3 * i + dim - 1]);
            }
        } else if (t == 2 || (t == 7 && et == 2)) {
            const sphg_soil &s = soil_of(i);
            double por = s.porosity, e0 = s.eps0;
            rhop[i] = (1 - por) * s.dens;
            int md = m % 10;
            bool listed = (m >= 0 && m < 30) || (m >= 50 && m < 60);
            if (listed) {
                satu[i] = md == 3 ? s.sitar : 1.0;
                if (md == 1) roue[i] = (s.ramda - s.kapa) * log(s.ocr);
            }
            if (t == 2) {
                hmaxs = fmax(hmaxs, xp[3 * i + dim - 1]);
                hmins = fmin(hmins, xp[3 * i + dim - 1]);
            }
            for (int k = 0; k < 3; k++) axp[3 * i + k] = g3[k];
            e[i] = e0;
            porosity[i] = por;
            hl[i] = dr;
            massini(i);
            cd[i] = s.damp_coe * sqrt(s.e / rhop[i] / hl[i] / hl[i]);
        } else if (t == 3) {
            const sphg_fluid &f = fluid_of(i);
            rhop[i] = f.porosity * f.dens;
            satu[i] = 0.0;
            hl[i] = dr;
            massini(i);
            for (int k = 0; k < 3; k++) axp[3 * i + k] = g3[k];
        } else if (t == 4) {
            const sphg_structure &st = mat->structure;
            rhop[i] = (1 - st.porosity) * st.dens;
            satu[i] = 1.0;
            e[i] = st.eps0;
            hl[i] = dr;
            massini(i);
            for (int k = 0; k < 3; k++) axp[3 * i + k] = g3[k];
            cd[i] = soil_of(i).damp_coe * sqrt(st.e / rhop[i] / hl[i] / hl[i]);
        }
        rho[i] = rhop[i];
    }
    const double g = dim == 2 ? pro->gy : dim == 3 ? pro->gz : 0.0;
    for (int i = 0; i < n; i++) {
        const int t = type[i], m = matype[i], et = etype[i];
        if (t == 0) {
            prep[i] = 1500.0 * (-g) * fabs(hmaxw - hminw) * 0.5;
            pre[i] = prep[i];
        }
        if (t == 1) {
            const sphg_fluid &f = fluid_of(i);
            if (f.press_type == 0) prep[i] = rhop[i] * (-g) * dr * f.pre_coe;
            else prep[i] = rhop[i] * (-g) * fabs(xp[3 * i + dim - 1] - hmaxw - f.pre_coe * dr);
            pre[i] = prep[i];
        }
        if (t == 7 && et == 1) {
            const sphg_fluid &f = fluid_of(i);
            prep[i] = rhop[i] * (-g) * dr * f.pre_coe;
            pre[i] = prep[i];
        }
        if (t == 2) {
            const sphg_soil &s = soil_of(i);
            if (fabs(s.sme0) > 1.0e-3) {
                for (int k = 0; k < 3; k++) sig[6 * i + k] = -s.sme0;
            } else {
                double v = rho[i] * g * (hmaxs - hmins) * 0.5;
                for (int k = 0; k < 3; k++) sig[6 * i + k] = v;
            }
            if (m == 3 || m == 13 || m == 23) suct[i] = moisturesuct(satu[i], -0.0001, s);
        }
        if (t == 7 && et == 2) {
            const sphg_soil &s = soil_of(i);
            if (fabs(s.sme0) > 1.0e-3)
                for (int k = 0; k < 3; k++) sig[6 * i + k] = -s.sme0;
            if (m == 3 || m == 13 || m == 23) suct[i] = moisturesuct(satu[i], -0.0001, s);
        }
        if (t == 3) {
            const sphg_fluid &f = fluid_of(i);
            if (f.press_type == 0) prep[i] = f.paa0;
            else prep[i] = f.dens * (-g) * fabs(xp[3 * i + dim - 1] - hmaxw - f.pre_coe * dr);
            pre[i] = prep[i];
        }
    }
    return SPHG_OK;
}
