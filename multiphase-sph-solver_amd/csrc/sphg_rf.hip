// Correlation-matrix random fields on the device (nonstationary_random_2D / _3D,
// CPP_17-3-Nonstationary-RF.cpp:76-849): the cells' correlation matrix, its Cholesky factor,
// E = L * zeta for one or many realisations, and the per-particle property.
//
// The matrix is dense, symmetric and up to 15,000 cells on a side, so the factor is the one
// cubic-cost step of the set-up. It runs as a blocked right-looking Cholesky in FP64 with 64-wide
// panels: the diagonal block is factored by one wave in LDS (column by column, the reference's
// order within the block, with its clamp sqrt(max(d, 1e-6))), the panel below by a triangular
// solve per row, and the trailing lower triangle is updated by an FP64 matrix-core kernel
// (v_mfma_f64_16x16x4_f64). E = L * Z^T uses the same matrix-core tile, skipping the zero blocks
// above the diagonal. Only the lower triangle of the matrix is stored (zeros above), padded to a
// multiple of 64 with an identity block that leaves the factor of the real cells unchanged.
#include <hip/hip_runtime.h>

namespace sphg {

constexpr int kRfB = 64;  // panel width and tile edge

struct RfParams {
    int ndim, acf, nc[3], total, npad;
    double base[3], cell_dx;
    double rl[3];
    double cb, sb, cb1, sb1;  // cos/sin of beta and beta_1 (radians as the reference forms them)
    double pi;
};

__device__ __forceinline__ void rf_centre(const RfParams &p, int c, double x[3]) {
    const int nx = p.nc[0], ny = p.nc[1];
    int i, j, k = 0;
    if (p.ndim == 2) {
        j = c / nx;
        i = c - j * nx;
    } else {
        k = c / (nx * ny);
        const int r = c - k * nx * ny;
        j = r / nx;
        i = r - j * nx;
    }
    x[0] = p.base[0] + ((double)i + 0.5) * p.cell_dx;
    x[1] = p.base[1] + ((double)j + 0.5) * p.cell_dx;
    x[2] = p.ndim == 2 ? 0.0 : p.base[2] + ((double)k + 0.5) * p.cell_dx;
}

// the autocorrelation between two cell centres (CPP_17-3-Nonstationary-RF.cpp:209-222, :599-614)
__device__ __forceinline__ double rf_acf(const RfParams &p, const double a[3], const double b[3]) {
    const double d0 = fabs(a[0] - b[0]), d1 = fabs(a[1] - b[1]);
    const double lx = p.rl[0], ly = p.rl[1];
    if (p.ndim == 2) {
        if (p.acf == 1) return exp(-2.0 * (d0 / lx + d1 / ly));
        if (p.acf == 2) return exp(-p.pi * (pow(d0 / lx, 2) + pow(d1 / ly, 2)));
        return exp(-2.0 * ((d0 * p.cb + d1 * p.sb) / lx + (-d0 * p.sb + d1 * p.cb) / ly));
    }
    const double d2 = fabs(a[2] - b[2]), lz = p.rl[2];
    if (p.acf == 1) return exp(-2.0 * (d0 / lx + d1 / ly + d2 / lz));
    if (p.acf == 2) return exp(-p.pi * (pow(d0 / lx, 2) + pow(d1 / ly, 2) + pow(d2 / lz, 2)));
    const double t0 = d0 * p.cb1 + d1 * p.sb1;
    const double t1 = (-d0 * p.sb1 + d1 * p.cb1) * p.cb + d2 * p.sb;
    const double t2 = -(-d0 * p.sb1 + d1 * p.cb1) * p.sb + d2 * p.cb;
    return exp(-2.0 * (t0 / lx + t1 / ly + t2 / lz));
}

// A[r][c], lower triangle of the correlation matrix (zeros above), identity on the padding
__global__ __launch_bounds__(256) void k_rf_corr(RfParams p, double *A) {
    const size_t np = (size_t)p.npad, tot = np * np;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(e / np), c = (int)(e - (size_t)r * np);
        double v = 0.0;
        if (c <= r) {
            if (r < p.total) {
                double a[3], b[3];
                rf_centre(p, r, a);
                rf_centre(p, c, b);
                v = rf_acf(p, a, b);
            } else if (r == c) {
                v = 1.0;
            }
        }
        A[e] = v;
    }
}

// Diagonal block A[k0.., k0..] (64 x 64, already updated by the earlier panels), one wave,
// right-looking with row t of the block in lane t's registers: for each column k, l_kk =
// sqrt(max(d, 1e-6)) with d the diagonal left by the earlier columns (the reference's clamp),
// the column below scaled by 1 / l_kk and shared through LDS, and every row below updated by the
// rank-1 product; the strict upper part is set to zero.
__global__ __launch_bounds__(64) void k_rf_potrf(double *A, int ld, int k0) {
    __shared__ double col[kRfB];
    __shared__ double diag;
    const int t = threadIdx.x;
    double *row = A + (size_t)(k0 + t) * ld + k0;
    double x[kRfB];
#pragma unroll
    for (int j = 0; j < kRfB; j++) x[j] = row[j];
#pragma unroll
    for (int k = 0; k < kRfB; k++) {
        if (t == k) {
            const double d = x[k];
            x[k] = sqrt(d > 1.0e-6 ? d : 1.0e-6);
            diag = 1.0 / x[k];
        }
        __syncthreads();
        if (t > k) x[k] = x[k] * diag;
        col[t] = x[k];
        __syncthreads();
        if (t > k) {
            const double lik = x[k];
#pragma unroll
            for (int j = k + 1; j < kRfB; j++) x[j] -= lik * col[j];
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < kRfB; j++) row[j] = j <= t ? x[j] : 0.0;
}

// Panel below the diagonal block: x L_kk^T = a for 64 rows per workgroup, one row per lane held
// in registers, right-looking (x_k = a_k / l_kk, then a_j -= x_k l_jk for j > k).
__global__ __launch_bounds__(64) void k_rf_trsm(double *A, int ld, int k0) {
    __shared__ double Lk[kRfB][kRfB + 1];
    __shared__ double inv[kRfB];
    const int t = threadIdx.x;
    const int r0 = k0 + kRfB + blockIdx.x * kRfB;
    const double *lk = A + (size_t)k0 * ld + k0;
    for (int r = 0; r < kRfB; r++) Lk[r][t] = lk[(size_t)r * ld + t];
    __syncthreads();
    inv[t] = 1.0 / Lk[t][t];
    double *row = A + (size_t)(r0 + t) * ld + k0;
    double x[kRfB];
#pragma unroll
    for (int k = 0; k < kRfB; k++) x[k] = row[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kRfB; k++) {
        x[k] = x[k] * inv[k];
#pragma unroll
        for (int j = k + 1; j < kRfB; j++) x[j] -= x[k] * Lk[j][k];
    }
#pragma unroll
    for (int k = 0; k < kRfB; k++) row[k] = x[k];
}

typedef double f64x4 __attribute__((ext_vector_type(4)));

// C tile (64 x 64) op= A rows x B rows^T over K columns from k0, four waves of 32 x 32, each 2 x 2
// v_mfma_f64_16x16x4_f64 tiles (lane l supplies A[l % 16][l / 16] and B^T[l / 16][l % 16] and
// holds D[l / 16 + 4 v][l % 16] in accumulator v). Modes:
//   kTri   the lower tiles (ti >= tj) of the square block from (r0, c0), walked by a 1D grid;
//          C -= A A^T (the trailing update of the factor);
//   kRect  a 2D grid of tiles from (r0, c0) that skips tiles above the diagonal; C -= A A^T (the
//          update of the columns still inside the current wide panel);
//   kProd  a 3D grid, C_s = A B^T over the K slice [s K, (s + 1) K) of blockIdx.z = s, limited to
//          the nonzero blocks of a lower-triangular A; slice s is written at C + s * cstride and
//          the slices are summed in order by k_rf_sum_slices (E = L Z^T, split so that enough
//          workgroups share the long K of a product with few columns).
enum { kTri = 0, kRect = 1, kProd = 2 };

template <int kMode>
__global__ __launch_bounds__(256) void k_rf_gemm(const double *A, int lda, const double *B, int ldb, double *C,
                                                 int ldc, int r0, int c0, int k0, int K, size_t cstride) {
    constexpr bool kLower = kMode != kProd;
    constexpr int kKc = 32;  // k columns staged per round: 2 x 18 KB of LDS, four workgroups per CU
    // row stride 36: the operand reads (16 rows x 4 k per wave) fall on distinct bank pairs
    __shared__ double As[kRfB][kKc + 4];
    __shared__ double Bs[kRfB][kKc + 4];
    int ti, tj;
    if (kMode == kTri) {
        const int b = blockIdx.x;
        ti = (int)((sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
        while ((ti + 1) * (ti + 2) / 2 <= b) ti++;
        while (ti * (ti + 1) / 2 > b) ti--;
        tj = b - ti * (ti + 1) / 2;
    } else {
        ti = blockIdx.x;
        tj = blockIdx.y;
    }
    const int row0 = r0 + ti * kRfB, col0 = c0 + tj * kRfB;
    if (kMode == kRect && col0 > row0) return;  // above the diagonal: never read, left alone
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    // the update starts from the C tile (read before the first staging, so the two overlap) and
    // subtracts with a negated A operand; the product form starts from zero
    f64x4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int v = 0; v < 4; v++)
                acc[a][b][v] = kLower ? C[(size_t)(row0 + wr + 16 * a + lk + 4 * v) * ldc + col0 + wc + 16 * b + li] : 0.0;
    const double sgn = kLower ? -1.0 : 1.0;
    // the product's slice; A lower triangular: blocks past the diagonal are zero
    const int kbeg = kLower ? 0 : (int)blockIdx.z * K;
    const int kend = kLower ? K : min(kbeg + K, (ti + 1) * kRfB);
    if (!kLower) C += (size_t)blockIdx.z * cstride;
    const int lc = t & (kKc - 1), lr0 = t / kKc;
    for (int kc = kbeg; kc < kend; kc += kKc) {
        // stage the 64 x 32 chunks: in load q each wave reads two row segments of 256 contiguous bytes
#pragma unroll
        for (int q = 0; q < kRfB * kKc / 256; q++) {
            const int lr = q * (256 / kKc) + lr0;
            As[lr][lc] = A[(size_t)(row0 + lr) * lda + k0 + kc + lc];
            Bs[lr][lc] = B[(size_t)(col0 + lr) * ldb + k0 + kc + lc];
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kKc; kk += 4) {
            const double a0 = sgn * As[wr + li][kk + lk], a1 = sgn * As[wr + 16 + li][kk + lk];
            const double b0 = Bs[wc + li][kk + lk], b1 = Bs[wc + 16 + li][kk + lk];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int v = 0; v < 4; v++)
                C[(size_t)(row0 + wr + 16 * a + lk + 4 * v) * ldc + col0 + wc + 16 * b + li] = acc[a][b][v];
}

// E[e] = sum over the slices s = 0, 1, ... of C_s[e], in that order
__global__ __launch_bounds__(256) void k_rf_sum_slices(const double *part, int nslice, size_t stride, double *E) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= stride) return;
    double s = 0.0;
    for (int q = 0; q < nslice; q++) s += part[(size_t)q * stride + e];
    E[e] = s;
}

// Sets *flag when any entry of the n x n lower triangle of the factor is not finite: beyond what the
// clamp sqrt(max(d, 1e-6)) absorbs (an indefinite correlation matrix) the factor degenerates to
// inf / NaN, and every property value drawn from it would be NaN.
__global__ __launch_bounds__(256) void k_rf_nonfinite(const double *A, int ld, int n, int *flag) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * n) return;
    const int r = (int)(e / n), c = (int)(e % n);
    if (c <= r && !isfinite(A[(size_t)r * ld + c])) atomicOr(flag, 1);
}

// How the cell values E become one property: phase, column of E, the vertical axis and the tops
// of the depth trend for the normal and the lognormal branch, and the lognormal map.
struct RfValues {
    int dist, want_type, total, rpad, n, nreal;
    int ncol, col;       // E column of realisation r: r * ncol + col
    int vaxis;           // coordinate of the cell centre the trend is measured on
    int use_exp;         // lognormal value as exp(v) (correlated pair) or pow(exp(1), v)
    double mean_value, sd_value, coe_depth, top_normal, top_log, e1;
};

// per-particle property (CPP_17-3-Nonstationary-RF.cpp:227-450, :620-846;
// CPP_17-5-Nonstationary-Correlated-RF.cpp:244-395, :574-725): the cell's E with the depth trend
// below the branch's top and the lognormal map; particles of the other phase, or in no cell, keep
// fill
__global__ __launch_bounds__(256) void k_rf_values(RfParams p, RfValues q, const double *E, const int *cell_of,
                                                   const int *type, double fill, double *value) {
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long long)q.n * q.nreal) return;
    const int r = (int)(g / q.n), i = (int)(g - (long long)r * q.n);
    const int c = cell_of[i];
    double out = fill;
    if (c >= 0 && type[i] == q.want_type) {
        double x[3];
        rf_centre(p, c, x);
        const double y = x[q.vaxis];
        const double e = E[(size_t)c * q.rpad + (size_t)r * q.ncol + q.col];
        if (q.dist == 1) {
            const double top = q.top_normal;
            out = top > y ? q.mean_value + q.coe_depth * (top - y) + q.sd_value * e : q.mean_value + q.sd_value * e;
        } else {
            double tm = q.top_log > y ? q.mean_value + q.coe_depth * (q.top_log - y) : q.mean_value;
            const double st = sqrt(log(1.0 + (q.sd_value / tm) * (q.sd_value / tm)));
            tm = log(tm) - 0.5 * st * st;
            const double v = tm + st * e;
            out = q.use_exp ? exp(v) : pow(q.e1, v);
        }
    }
    value[g] = out;
}

}  // namespace sphg
