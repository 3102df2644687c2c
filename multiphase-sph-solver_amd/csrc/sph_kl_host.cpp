// Host side of the K-L random fields: the domain, the modes (roots and eigenvalue order) in
// FP64 with the C library, as the reference computes them once per property at set-up.
#include <math.h>

#include <algorithm>
#include <vector>

#include "../../include/sphg.h"

namespace {
constexpr double kPi = 3.141592653;  // Header_Parameters.h:10
constexpr double kPrecision = 1.0e-6;  // CPP_17-1_K-L_expansion_non.cpp:21

constexpr int kMaxBisect = 400;  // the interval reaches FP64 resolution in well under 100 halvings

// Root of the characteristic equation of mode i on one axis (solve_wi,
// CPP_17-1_K-L_expansion_non.cpp:280-343): bisection from [lower + 1e-6, upper - 1e-6] until
// |f(xm)| <= 1e-6, same iteration as the reference. Where FP64 cannot resolve |f| <= 1e-6 (very
// large w a), the reference loops forever; here the iteration stops after kMaxBisect halvings and
// *ok is cleared.
double mode_root(int i, double a, double dira, bool *ok) {
    int it = 0;
    double up, lo;
    if ((i + 1) % 2 == 1) {
        up = (i + 0.5) * kPi / a;
        lo = i * kPi / a;
    } else {
        up = (i + 1.0) * kPi / a;
        lo = (i + 0.5) * kPi / a;
    }
    double x1 = lo + kPrecision, x2 = up - kPrecision, xm, f1, fm;
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
        } while (fabs(fm) > kPrecision && ++it < kMaxBisect);
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
        } while (fabs(fm) > kPrecision && ++it < kMaxBisect);
    }
    if (fabs(fm) > kPrecision) *ok = false;
    return xm;
}
}  // namespace

extern "C" int sphg_kl_domain(int ndim, int n, const double *xp, const int *type, double a[3], double x_trans[3]) {
    if (!xp || !type || !a || !x_trans || (ndim != 2 && ndim != 3) || n <= 0) return SPHG_E_ARG;
    double mx[3] = {-10000.0, -10000.0, -10000.0}, mn[3] = {10000.0, 10000.0, 10000.0};
    for (int i = 0; i < n; i++) {
        if (type[i] == 0 || type[i] == 7) continue;
        for (int j = 0; j < 3; j++) {
            mx[j] = fmax(mx[j], xp[3 * i + j]);
            mn[j] = fmin(mn[j], xp[3 * i + j]);
        }
    }
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int j = 0; j < ndim; j++) {
        lo[j] = mn[j] - 0.2 * (mx[j] - mn[j]);
        hi[j] = mx[j] + 0.2 * (mx[j] - mn[j]);
    }
    for (int j = 0; j < 3; j++) {
        x_trans[j] = (hi[j] + lo[j]) / 2.0;
        a[j] = (hi[j] - lo[j]) / 2.0;
    }
    if (a[0] == 0.0 || a[1] == 0.0) return 13;
    return SPHG_OK;
}

extern "C" int sphg_kl_modes(sphg_kl_field *f) {
    if (!f || (f->ndim != 2 && f->ndim != 3)) return SPHG_E_ARG;
    const int nd = f->ndim, T = SPHG_KL_TERMS;
    for (int d = 0; d < 3; d++)
        for (int i = 0; i < T; i++) f->wi[i][d] = 0.0;
    bool ok = true;
    for (int d = 0; d < nd; d++)
        for (int i = 0; i < T && ok; i++) f->wi[i][d] = mode_root(i, f->a[d], f->rl[d], &ok);
    if (!ok) return SPHG_E_NOCONVERGE;
    // equation_ramdai (:37-42)
    double lam[SPHG_KL_TERMS][3];
    for (int d = 0; d < nd; d++)
        for (int i = 0; i < T; i++) {
            const double dr = f->rl[d], w = f->wi[i][d];
            lam[i][d] = 4.0 * dr / (4.0 + w * w * dr * dr);
        }
    // The reference picks the 200 largest products by repeated first-maximum scans over the
    // (m, n[, k]) table in row-major order, overwriting each pick with -1 (:384-433). That is the
    // head of the table sorted by (value descending, row-major index ascending), taken here with
    // a partial sort over the same FP64 products.
    const long long cnt = nd == 2 ? (long long)T * T : (long long)T * T * T;
    std::vector<std::pair<double, int>> v((size_t)cnt);
    long long p = 0;
    for (int m = 0; m < T; m++)
        for (int q = 0; q < T; q++) {
            if (nd == 2) {
                v[p] = {lam[m][0] * lam[q][1], (int)p};
                p++;
            } else {
                for (int k = 0; k < T; k++) {
                    v[p] = {lam[m][0] * lam[q][1] * lam[k][2], (int)p};
                    p++;
                }
            }
        }
    std::partial_sort(v.begin(), v.begin() + T, v.end(), [](const std::pair<double, int> &x, const std::pair<double, int> &y) {
        return x.first > y.first || (x.first == y.first && x.second < y.second);
    });
    for (int i = 0; i < T; i++) {
        const int id = v[i].second;
        f->eigval[i] = v[i].first;
        if (nd == 2) {
            f->order[i][0] = id / T;
            f->order[i][1] = id % T;
            f->order[i][2] = i;  // untouched third index keeps the initial value i (:363-369)
        } else {
            f->order[i][0] = id / (T * T);
            f->order[i][1] = (id / T) % T;
            f->order[i][2] = id % T;
        }
    }
    return SPHG_OK;
}
