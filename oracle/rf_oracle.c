/* Test infrastructure only: CPU restatement of the correlation-matrix random fields of the
 * reference (nonstationary_random_2D / _3D, CPP_17-3-Nonstationary-RF.cpp:76-849, and the
 * correlated pair of nonstationary_correlated_random_2D / _3D,
 * CPP_17-5-Nonstationary-Correlated-RF.cpp:87-737), the checker for the library's sphg_rf_*
 * entry points.
 *
 * Each function follows the reference code named in its comment, in the same operation order.
 * The standard normals zeta are an input: the reference draws them from a wall-clock seeded
 * generator (generate_SND, :32-42), so parity is defined for injected zeta. */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sph_oracle.h"

static const double RF_PI = 3.141592653; /* Header_Parameters.h:10 */

/* cell partition and particle-to-cell map (:97-160 2D, :479-565 3D). Returns 41 (cell count not
 * in (0, 15000)), 42 (a 40th particle in a cell), -1 (a 3D cell index outside the grid, where the
 * reference writes out of bounds), else 0. cell_of[i] = -1 for particles the reference does not
 * map (2D: other than water and soil). */
int or_rf_cells(int ndim, int n, const double *xp, const int *type, double cell_ndx, double dr,
                const double box_nb[6], int nc[3], double base[3], double *cell_dx_out, int *cell_total,
                int *cell_of) {
    double xmin[3], xmax[3], cell_dx;
    int nx, ny, nz, total, i, *cnt;
    xmin[0] = box_nb[0] - 0.1 * (box_nb[1] - box_nb[0]);
    xmax[0] = box_nb[1] + 0.1 * (box_nb[1] - box_nb[0]);
    xmin[1] = box_nb[2] - 0.1 * (box_nb[3] - box_nb[2]);
    xmax[1] = box_nb[3] + 0.1 * (box_nb[3] - box_nb[2]);
    if (ndim == 3) {
        xmin[2] = box_nb[4] - 0.1 * (box_nb[5] - box_nb[4]);
        xmax[2] = box_nb[5] + 0.1 * (box_nb[5] - box_nb[4]);
    } else {
        xmin[2] = 0.0;
        xmax[2] = 0.0;
    }
    cell_dx = cell_ndx * dr;
    nx = (int)((xmax[0] - xmin[0]) / cell_dx) + 1;
    ny = (int)((xmax[1] - xmin[1]) / cell_dx) + 1;
    nz = ndim == 3 ? (int)((xmax[2] - xmin[2]) / cell_dx) + 1 : 1;
    total = nx * ny * nz;
    if (!(total > 0 && total < 15000)) return 41;
    nc[0] = nx;
    nc[1] = ny;
    nc[2] = nz;
    base[0] = -(nx * cell_dx - (xmax[0] - xmin[0])) / 2.0 + 0.01 * dr + xmin[0];
    base[1] = -(ny * cell_dx - (xmax[1] - xmin[1])) / 2.0 + 0.01 * dr + xmin[1];
    base[2] = ndim == 3 ? -(nz * cell_dx - (xmax[2] - xmin[2])) / 2.0 + 0.01 * dr + xmin[2] : 0.0;
    *cell_dx_out = cell_dx;
    *cell_total = total;
    cnt = (int *)calloc((size_t)total, sizeof(int));
    for (i = 0; i < n; i++) {
        int np_x, np_y, np_z, id;
        cell_of[i] = -1;
        if (ndim == 2 && type[i] != 1 && type[i] != 2) continue;
        np_x = (int)(fabs(xp[3 * i] - base[0]) / cell_dx);
        np_y = (int)(fabs(xp[3 * i + 1] - base[1]) / cell_dx);
        if (ndim == 2) {
            id = np_y * nx + np_x;
        } else {
            np_z = (int)(fabs(xp[3 * i + 2] - base[2]) / cell_dx);
            id = np_z * nx * ny + np_y * nx + np_x;
        }
        if (id < 0 || id >= total) {
            free(cnt);
            return -1;
        }
        if (cnt[id] < 39) {
            cnt[id]++;
            cell_of[i] = id;
        } else {
            free(cnt);
            return 42;
        }
    }
    free(cnt);
    return 0;
}

static void rf_cent(int ndim, const int nc[3], const double base[3], double cell_dx, int id, double c[3]) {
    int i, j, k = 0;
    if (ndim == 2) {
        j = id / nc[0];
        i = id % nc[0];
    } else {
        k = id / (nc[0] * nc[1]);
        j = (id % (nc[0] * nc[1])) / nc[0];
        i = id % nc[0];
    }
    c[0] = base[0] + ((double)i + 0.5) * cell_dx;
    c[1] = base[1] + ((double)j + 0.5) * cell_dx;
    c[2] = ndim == 3 ? base[2] + ((double)k + 0.5) * cell_dx : 0.0;
}

/* autocorrelation matrix mat_C (full, row-major) (:203-223 2D, :595-615 3D) */
void or_rf_corr(int ndim, int acf_type, const double rl[3], double beta, double beta_1, const int nc[3],
                const double base[3], double cell_dx, int total, double *C) {
    double beta_temp = beta * RF_PI / 180.0, beta1_temp = beta_1 * RF_PI / 180.0;
    double lx = rl[0], ly = rl[1], lz = rl[2];
    int k, j;
    for (k = 0; k < total; k++) {
        double ck[3];
        rf_cent(ndim, nc, base, cell_dx, k, ck);
        for (j = 0; j < total; j++) {
            double cj[3], dx[3], acf;
            rf_cent(ndim, nc, base, cell_dx, j, cj);
            dx[0] = fabs(ck[0] - cj[0]);
            dx[1] = fabs(ck[1] - cj[1]);
            if (ndim == 2) {
                dx[2] = 0.0;
                if (acf_type == 1) acf = exp(-2.0 * (dx[0] / lx + dx[1] / ly));
                else if (acf_type == 2) acf = exp(-RF_PI * (pow(dx[0] / lx, 2) + pow(dx[1] / ly, 2)));
                else
                    acf = exp(-2.0 * ((dx[0] * cos(beta_temp) + dx[1] * sin(beta_temp)) / lx +
                                      (-dx[0] * sin(beta_temp) + dx[1] * cos(beta_temp)) / ly));
            } else {
                double t[3];
                dx[2] = fabs(ck[2] - cj[2]);
                t[0] = dx[0] * cos(beta1_temp) + dx[1] * sin(beta1_temp);
                t[1] = (-dx[0] * sin(beta1_temp) + dx[1] * cos(beta1_temp)) * cos(beta_temp) + dx[2] * sin(beta_temp);
                t[2] = -(-dx[0] * sin(beta1_temp) + dx[1] * cos(beta1_temp)) * sin(beta_temp) + dx[2] * cos(beta_temp);
                if (acf_type == 1) acf = exp(-2.0 * (dx[0] / lx + dx[1] / ly + dx[2] / lz));
                else if (acf_type == 2)
                    acf = exp(-RF_PI * (pow(dx[0] / lx, 2) + pow(dx[1] / ly, 2) + pow(dx[2] / lz, 2)));
                else acf = exp(-2.0 * (t[0] / lx + t[1] / ly + t[2] / lz));
            }
            C[(size_t)k * total + j] = acf;
        }
    }
}

/* Cholesky (:45-74) serial: column k, diagonal from the ascending sum of squares with the clamp
 * sqrt(max(d, 1e-6)), then the column below; the upper triangle is zero */
void or_rf_cholesky(const double *A, double *L, int n) {
    int k, i, j;
    memset(L, 0, sizeof(double) * (size_t)n * n);
    for (k = 0; k < n; k++) {
        double sum = 0;
        for (i = 0; i < k; i++) sum += L[(size_t)k * n + i] * L[(size_t)k * n + i];
        sum = A[(size_t)k * n + k] - sum;
        L[(size_t)k * n + k] = sqrt(sum > 1.0e-6 ? sum : 1.0e-6);
        for (i = k + 1; i < n; i++) {
            sum = 0;
            for (j = 0; j < k; j++) sum += L[(size_t)i * n + j] * L[(size_t)k * n + j];
            L[(size_t)i * n + k] = (A[(size_t)i * n + k] - sum) / L[(size_t)k * n + k];
        }
    }
}

/* E = L zeta (matrix_multi, :20-30, every column in ascending order) and the per-particle
 * property (:227-450 2D, :620-846 3D): the depth trend below soil_max (soil fields, and the fluid
 * fields' lognormal branch) or water_max (the fluid fields' normal branch), lognormal values as
 * pow(exp(1), v). Particles of the other phase or in no cell keep fill. E (may be null)
 * receives the total cell values. */
void or_rf_values(int ndim, int dist, int field, double mean, double std_dev, double coe_depth, double soil_max,
                  double water_max, const int nc[3], const double base[3], double cell_dx, int total, const double *L,
                  const double *zeta, int n, const int *type, const int *cell_of, double fill, double *E_out,
                  double *value) {
    double *E = (double *)malloc(sizeof(double) * (size_t)total);
    double *cv = (double *)malloc(sizeof(double) * (size_t)total);
    int i, j, want = field <= 3 ? 2 : 1;
    for (i = 0; i < total; i++) {
        double temp = 0.0;
        for (j = 0; j < total; j++) temp = temp + L[(size_t)i * total + j] * zeta[j];
        E[i] = temp;
    }
    for (j = 0; j < total; j++) {
        double c[3], y, top, v;
        rf_cent(ndim, nc, base, cell_dx, j, c);
        y = ndim == 2 ? c[1] : c[2];
        if (dist == 1) {
            top = field <= 3 ? soil_max : water_max;
            if (top > y) v = mean + coe_depth * (top - y) + std_dev * E[j];
            else v = mean + std_dev * E[j];
        } else {
            double temp_mean, std_temp;
            if (soil_max > y) temp_mean = mean + coe_depth * (soil_max - y);
            else temp_mean = mean;
            std_temp = sqrt(log(1.0 + (std_dev / temp_mean) * (std_dev / temp_mean)));
            temp_mean = log(temp_mean) - 0.5 * std_temp * std_temp;
            v = pow(exp(1), temp_mean + std_temp * E[j]);
        }
        cv[j] = v;
    }
    for (i = 0; i < n; i++) value[i] = (cell_of[i] >= 0 && type[i] == want) ? cv[cell_of[i]] : fill;
    if (E_out) memcpy(E_out, E, sizeof(double) * (size_t)total);
    free(E);
    free(cv);
}

/* correlated friction angle and cohesion of one phase (nonstationary_correlated_random_2D/3D,
 * CPP_17-5-Nonstationary-Correlated-RF.cpp:87-737): the mixing of the two normals per cell
 * (:215-239, matrix_multi of :22-36), E = L zeta (:242), and the two properties (:244-395 2D,
 * :574-725 3D). phase 0: soil (fai, c), trend below soil_max on the vertical axis; phase 1:
 * fluid, normal branch below water_max and lognormal below soil_max, both on y. Both properties
 * use dist; lognormal values are exp(v). */
void or_rf_pair_values(int ndim, int phase, int dist, double coe, const double mean[2], const double sd[2],
                       const double cdep[2], double soil_max, double water_max, const int nc[3], const double base[3],
                       double cell_dx, int total, const double *L, const double *zeta_1, int n, const int *type,
                       const int *cell_of, double fill, double *out_fai, double *out_c) {
    double coefficient, cov[4], *zeta, *E, *cv[2];
    int i, j, k, want = phase == 0 ? 2 : 1;
    if (dist == 2) {
        coefficient = log(1.0 + coe * (sd[0] / mean[0]) * (sd[1] / mean[1]));
        coefficient = coefficient / sqrt(log(1.0 + sd[0] * sd[0] / mean[0] / mean[0]) *
                                         log(1.0 + sd[1] * sd[1] / mean[1] / mean[1]));
    } else {
        coefficient = coe;
    }
    cov[0] = 1.0;
    cov[1] = coefficient;
    cov[2] = 0.0;
    cov[3] = sqrt(1.0 - coefficient * coefficient);
    zeta = (double *)malloc(sizeof(double) * (size_t)total * 2);
    E = (double *)malloc(sizeof(double) * (size_t)total * 2);
    for (i = 0; i < total; i++)
        for (j = 0; j < 2; j++) {
            double temp = 0.0;
            for (k = 0; k < 2; k++) temp = temp + zeta_1[i * 2 + k] * cov[k * 2 + j];
            zeta[i * 2 + j] = temp;
        }
    for (i = 0; i < total; i++)
        for (j = 0; j < 2; j++) {
            double temp = 0.0;
            for (k = 0; k < total; k++) temp = temp + L[(size_t)i * total + k] * zeta[k * 2 + j];
            E[i * 2 + j] = temp;
        }
    for (k = 0; k < 2; k++) {
        cv[k] = (double *)malloc(sizeof(double) * (size_t)total);
        for (j = 0; j < total; j++) {
            double c[3], y, v;
            rf_cent(ndim, nc, base, cell_dx, j, c);
            y = (phase == 0 && ndim == 3) ? c[2] : c[1];
            if (dist == 1) {
                double top = phase == 0 ? soil_max : water_max;
                if (top > y) v = mean[k] + cdep[k] * (top - y) + sd[k] * E[j * 2 + k];
                else v = mean[k] + sd[k] * E[j * 2 + k];
            } else {
                double temp_mean, std_temp;
                if (soil_max > y) temp_mean = mean[k] + cdep[k] * (soil_max - y);
                else temp_mean = mean[k];
                std_temp = sqrt(log(1.0 + (sd[k] / temp_mean) * (sd[k] / temp_mean)));
                temp_mean = log(temp_mean) - 0.5 * std_temp * std_temp;
                v = exp(temp_mean + std_temp * E[j * 2 + k]);
            }
            cv[k][j] = v;
        }
    }
    for (i = 0; i < n; i++) {
        int ok = cell_of[i] >= 0 && type[i] == want;
        out_fai[i] = ok ? cv[0][cell_of[i]] : fill;
        out_c[i] = ok ? cv[1][cell_of[i]] : fill;
    }
    free(cv[0]);
    free(cv[1]);
    free(zeta);
    free(E);
}
