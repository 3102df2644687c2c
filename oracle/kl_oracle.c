/* Test infrastructure only: CPU restatement of the K-L random-field synthesis of the reference
 * (CPP_17-1_K-L_expansion_non.cpp), the checker for the library's K-L entry points.
 *
 * Each function follows the reference function named in its comment, in the same operation
 * order. The standard normals xi are an input here: the reference draws them from a wall-clock
 * seeded generator (generate_SND, :24-34), so parity is defined for injected xi. */
#include <math.h>
#include <string.h>

#include "sph_oracle.h"

#define KL_TERMS 200
#define KL_PRECISION 1.0e-6
static const double KL_PI = 3.141592653; /* Header_Parameters.h:10 */

/* solve_wi (CPP_17-1_K-L_expansion_non.cpp:268-345): bisection for the roots of
 * 2/delta - w tan(w a) (0-based even i) and 2/delta tan(w a) + w (0-based odd i). The reference
 * loops forever where FP64 cannot reach |f| <= 1e-6; the restatement stops after 400 halvings
 * and returns 104 (the checker must not hang). */
int or_kl_solve_wi(double wi[][3], const double a_x[3], const double dirac_r[3], int ndim) {
    int bad = 0;
    double upper[KL_TERMS], lower[KL_TERMS];
    int d, i;
    for (d = 0; d < ndim; d++) {
        double a = a_x[d], dira = dirac_r[d];
        for (i = 0; i < KL_TERMS; i++) {
            if ((i + 1) % 2 == 1) {
                upper[i] = (i + 0.5) * KL_PI / a;
                lower[i] = i * KL_PI / a;
            } else {
                upper[i] = (i + 1.0) * KL_PI / a;
                lower[i] = (i + 0.5) * KL_PI / a;
            }
        }
        for (i = 0; i < KL_TERMS; i++) {
            double x1 = lower[i] + KL_PRECISION, x2 = upper[i] - KL_PRECISION, xm, f1, fm;
            int it = 0;
            if ((i + 1) % 2 == 1) {
                xm = (x1 + x2) * 0.5;
                do {
                    f1 = 2.0 / dira - x1 * tan(x1 * a);
                    fm = 2.0 / dira - xm * tan(xm * a);
                    if ((f1 * fm) < 0.0) {
                        x2 = xm;
                        xm = (x1 + x2) * 0.5;
                    } else {
                        x1 = xm;
                        xm = (x1 + x2) * 0.5;
                    }
                } while (fabs(fm) > KL_PRECISION && ++it < 400);
            } else {
                xm = (x1 + x2) / 2.0;
                do {
                    f1 = 2.0 / dira * tan(x1 * a) + x1;
                    fm = 2.0 / dira * tan(xm * a) + xm;
                    if ((f1 * fm) < 0.0) {
                        x2 = xm;
                        xm = (x1 + x2) * 0.5;
                    } else {
                        x1 = xm;
                        xm = (x1 + x2) * 0.5;
                    }
                } while (fabs(fm) > KL_PRECISION && ++it < 400);
            }
            if (fabs(fm) > KL_PRECISION) bad = 104;
            wi[i][d] = xm;
        }
    }
    return bad;
}

/* solve_ramdai (CPP_17-1_K-L_expansion_non.cpp:348-438): lambda = 4 delta / (4 + w^2 delta^2)
 * per axis, products over axes, the 200 largest by repeated first-maximum search. */
void or_kl_solve_ramdai(double *eigval, const double wi[][3], const double dirac_r[3], int order_id[][3],
                        int ndim) {
    static double prod2[KL_TERMS][KL_TERMS];
    double ramda[KL_TERMS][3];
    int i, d, m, n, k;
    for (i = 0; i < KL_TERMS; i++) order_id[i][0] = order_id[i][1] = order_id[i][2] = i;
    for (d = 0; d < ndim; d++)
        for (i = 0; i < KL_TERMS; i++)
            ramda[i][d] = 4.0 * dirac_r[d] / (4.0 + wi[i][d] * wi[i][d] * dirac_r[d] * dirac_r[d]);
    if (ndim == 2) {
        for (m = 0; m < KL_TERMS; m++)
            for (n = 0; n < KL_TERMS; n++) prod2[m][n] = ramda[m][0] * ramda[n][1];
        for (i = 0; i < KL_TERMS; i++) {
            eigval[i] = -1.0;
            for (m = 0; m < KL_TERMS; m++)
                for (n = 0; n < KL_TERMS; n++)
                    if (eigval[i] < prod2[m][n]) {
                        eigval[i] = prod2[m][n];
                        order_id[i][0] = m;
                        order_id[i][1] = n;
                    }
            prod2[order_id[i][0]][order_id[i][1]] = -1.0;
        }
    } else {
        /* same first-maximum search over the 200^3 products, evaluated on the fly; a product
         * already taken is skipped (the reference overwrites it with -1) */
        for (i = 0; i < KL_TERMS; i++) {
            eigval[i] = -1.0;
            for (m = 0; m < KL_TERMS; m++)
                for (n = 0; n < KL_TERMS; n++)
                    for (k = 0; k < KL_TERMS; k++) {
                        double v = ramda[m][0] * ramda[n][1] * ramda[k][2];
                        int taken = 0, t;
                        if (!(eigval[i] < v)) continue;
                        for (t = 0; t < i; t++)
                            if (order_id[t][0] == m && order_id[t][1] == n && order_id[t][2] == k) {
                                taken = 1;
                                break;
                            }
                        if (taken) continue;
                        eigval[i] = v;
                        order_id[i][0] = m;
                        order_id[i][1] = n;
                        order_id[i][2] = k;
                    }
        }
    }
}

/* solve_faii (CPP_17-1_K-L_expansion_non.cpp:45-60) */
static double kl_faii(double x, double a, double wi, int i) {
    if ((i + 1) % 2 == 1) return 1.0 / sqrt(sin(2.0 * wi * a) / 2.0 / wi + a) * cos(wi * x);
    return 1.0 / sqrt(-sin(2.0 * wi * a) / 2.0 / wi + a) * sin(wi * x);
}

/* solve_h_2D / solve_h_3D (CPP_17-1_K-L_expansion_non.cpp:499-713): for every particle of type
 * other than 0 and 7, H = sum_j sqrt(lambda_j) phi_j(x) xi_j; the value of field `field` (0 soil
 * fai, 1 soil c, 2 soil cop, 3 soil ds, 4 fluid fai, 5 fluid c) of a particle of the matching type
 * becomes mean + sigma H (dist 1) or exp(mu_ln + sigma_ln H) (dist 2), the mean with the depth
 * trend on xp[1] from soil_max / water_max. Other particles keep value[i]. Returns 14 for an
 * unknown distribution, as the reference. */
int or_kl_synthesize(int ndim, int n, const double *xp, const int *type, const double wi[][3],
                     const double *eigval, const int order_id[][3], const double *uu, const double x_trans[3],
                     const double a[3], int dist, double mean_value, double sd_value, double coe_depth,
                     double soil_max, double water_max, int field, double *value) {
    int i, j;
    for (i = 0; i < n; i++) {
        double x_a[3], eigvect[KL_TERMS], temp_h, temp_mean, temp_sigma, v;
        int t = type[i];
        if (t == 0 || t == 7) continue;
        x_a[0] = xp[3 * i] - x_trans[0];
        x_a[1] = xp[3 * i + 1] - x_trans[1];
        x_a[2] = xp[3 * i + 2] - x_trans[2];
        for (j = 0; j < KL_TERMS; j++) {
            int idx = order_id[j][0], idy = order_id[j][1], idz = order_id[j][2];
            double f0 = kl_faii(x_a[0], a[0], wi[idx][0], idx);
            double f1 = kl_faii(x_a[1], a[1], wi[idy][1], idy);
            if (ndim == 2) eigvect[j] = f0 * f1;
            else eigvect[j] = f0 * f1 * kl_faii(x_a[2], a[2], wi[idz][2], idz);
        }
        if (dist != 1 && dist != 2) return 14;
        if (t == 2) temp_mean = mean_value + (soil_max - xp[3 * i + 1]) * coe_depth;
        else if (t == 1) temp_mean = mean_value + (water_max - xp[3 * i + 1]) * coe_depth;
        else temp_mean = mean_value;
        if (dist == 1) {
            temp_sigma = sd_value;
        } else {
            temp_sigma = sqrt(log(1.0 + (sd_value / temp_mean) * (sd_value / temp_mean)));
            temp_mean = log(temp_mean) - 0.5 * temp_sigma * temp_sigma;
        }
        temp_h = 0.0;
        for (j = 0; j < KL_TERMS; j++) temp_h = temp_h + sqrt(eigval[j]) * eigvect[j] * uu[j];
        v = dist == 1 ? temp_mean + temp_h * temp_sigma : exp(temp_mean + temp_h * temp_sigma);
        if ((field <= 3 && t == 2) || (field >= 4 && t == 1)) value[i] = v;
    }
    return 0;
}

/* Standard normals of the correlated pair of SpatialVariables_Generate_2D_corr / _3D_corr
 * (CPP_17-2_K-L_expansion_corr.cpp:114-140 soil, :183-209 fluid): the coefficient is
 * coe_frict_cohesion, or for a lognormal first field log(1 + coe cv_a cv_b) divided by
 * sqrt(log(1 + cv_a^2) log(1 + cv_b^2)); u[200][2] times [[1, coef], [0, sqrt(1 - coef^2)]] in
 * matrix_multi's order (:22-37) gives the fai normals (column 0) and the c normals (column 1). */
void or_kl_corr_normals(int dist_a, double mean_a, double sd_a, double mean_b, double sd_b, double coe,
                        const double *u, double *xi_a, double *xi_b, double *coef_out) {
    double coef = coe, cov[4], zeta[2 * KL_TERMS];
    int i, j, k;
    if (dist_a == 2) {
        coef = log(1.0 + coe * (sd_a / mean_a) * (sd_b / mean_b));
        coef = coef / sqrt(log(1.0 + sd_a * sd_a / mean_a / mean_a) * log(1.0 + sd_b * sd_b / mean_b / mean_b));
    }
    cov[0] = 1.0;
    cov[1] = coef;
    cov[2] = 0.0;
    cov[3] = sqrt(1.0 - coef * coef);
    for (i = 0; i < KL_TERMS; i++)
        for (j = 0; j < 2; j++) {
            double temp = 0.0;
            for (k = 0; k < 2; k++) temp = temp + u[i * 2 + k] * cov[k * 2 + j];
            zeta[i * 2 + j] = temp;
        }
    for (i = 0; i < KL_TERMS; i++) {
        xi_a[i] = zeta[2 * i];
        xi_b[i] = zeta[2 * i + 1];
    }
    if (coef_out) *coef_out = coef;
}

/* random_variable_generate (CPP_17-5-Nonstationary-Correlated-RF.cpp:738-863). flags: flag_fai,
 * flag_c, flag_faiw, flag_cw; dist[6], mean[6], sd[6] indexed like the reference's psv (0 soil fai,
 * 1 soil c, 4 fluid fai, 5 fluid c); zeta[2] the standard normals. The distribution is psv[0]'s
 * when flag_fai is set, else psv[4]'s when flag_faiw is set, else the constructor's 1; the
 * coefficient as above (the lognormal form only with flag_fai or flag_faiw); the pair
 * zeta [[1, coef], [0, sqrt(1 - coef^2)]]; fai takes its second entry, c its first. vals[6][n]:
 * each flagged property of every soil (type 2: 0, 1) or water (type 1: 4, 5) particle. */
void or_random_variable(const int flags[4], double coe, const int dist_of[6], const double mean_of[6],
                        const double sd_of[6], const double zeta[2], int n, const int *type, double *vals) {
    int dist = 1, i, q, j, k;
    double coef, cov[4], z1[2];
    static const int slot[4] = {0, 1, 4, 5}, zi[4] = {1, 0, 1, 0}, ph[4] = {2, 2, 1, 1};
    if (flags[0] == 1) dist = dist_of[0];
    else if (flags[2] == 1) dist = dist_of[4];
    if (dist == 2 && flags[0] == 1) {
        coef = log(1.0 + coe * (sd_of[0] / mean_of[0]) * (sd_of[1] / mean_of[1]));
        coef = coef / sqrt(log(1.0 + sd_of[0] * sd_of[0] / mean_of[0] / mean_of[0]) *
                           log(1.0 + sd_of[1] * sd_of[1] / mean_of[1] / mean_of[1]));
    } else if (dist == 2 && flags[2] == 1) {
        coef = log(1.0 + coe * (sd_of[4] / mean_of[4]) * (sd_of[5] / mean_of[5]));
        coef = coef / sqrt(log(1.0 + sd_of[4] * sd_of[4] / mean_of[4] / mean_of[4]) *
                           log(1.0 + sd_of[5] * sd_of[5] / mean_of[5] / mean_of[5]));
    } else {
        coef = coe;
    }
    cov[0] = 1.0;
    cov[1] = coef;
    cov[2] = 0.0;
    cov[3] = sqrt(1.0 - coef * coef);
    for (j = 0; j < 2; j++) {
        double temp = 0.0;
        for (k = 0; k < 2; k++) temp = temp + zeta[k] * cov[k * 2 + j];
        z1[j] = temp;
    }
    for (q = 0; q < 4; q++) {
        double mean, sd, v;
        if (flags[q] != 1) continue;
        mean = mean_of[slot[q]];
        sd = sd_of[slot[q]];
        if (dist == 1) {
            v = mean + sd * z1[zi[q]];
        } else if (dist == 2) {
            sd = sqrt(log(1.0 + (sd / mean) * (sd / mean)));
            mean = log(mean) - 0.5 * sd * sd;
            v = exp(mean + sd * z1[zi[q]]);
        } else {
            continue;
        }
        for (i = 0; i < n; i++)
            if (type[i] == ph[q]) vals[(size_t)slot[q] * n + i] = v;
    }
}
