// K-L random-field synthesis on the device (solve_h_2D / solve_h_3D,
// CPP_17-1_K-L_expansion_non.cpp:499-713).
//
// FP64 path (one field of one context): one thread per particle evaluates the 200 products of
// eigenfunctions, sums sqrt(lambda_j) phi_j(x) xi_j in mode order, and maps H to the property.
// Batch path (many realisations in a common domain): Phi(x) scaled by sqrt(lambda) is formed
// once in FP64 and stored as FP32, the realisations' standard normals form Xi, and H = Phi * Xi
// runs on the matrix cores (v_mfma_f32_32x32x2_f32), one wave per 32 x 32 tile of H.
// Mode constants (roots, normalisations 1/sqrt(sin(2wa)/2w +- a), sqrt(lambda)) are evaluated
// on the host with the C library, so only cos/sin(w x) and the lognormal map run on the device.
#include <hip/hip_runtime.h>

namespace sphg {

constexpr int kKlTerms = SPHG_KL_TERMS;

struct KlModes {
    double w[kKlTerms][3];      // root of the mode on each axis
    double alpha[kKlTerms][3];  // normalisation of solve_faii
    int use_cos[kKlTerms][3];   // (idx + 1) % 2 == 1: alpha cos(w x), else alpha sin(w x)
    double seig[kKlTerms];      // sqrt(eigval)
    double xi[kKlTerms];        // standard normals (FP64 path)
};

struct KlParams {
    int ndim, dist, field, n;
    double mean_value, sd_value, coe_depth, soil_max, water_max;
    double x_trans[3];
};

// Value of the property at particle i from H (CPP_17-1_K-L_expansion_non.cpp:552-606)
__device__ __forceinline__ double kl_value(const KlParams &kp, int t, double y, double temp_h) {
    double temp_mean;
    if (t == 2) temp_mean = kp.mean_value + (kp.soil_max - y) * kp.coe_depth;
    else if (t == 1) temp_mean = kp.mean_value + (kp.water_max - y) * kp.coe_depth;
    else temp_mean = kp.mean_value;
    if (kp.dist == 1) return temp_mean + temp_h * kp.sd_value;
    const double sg = sqrt(log(1.0 + (kp.sd_value / temp_mean) * (kp.sd_value / temp_mean)));
    temp_mean = log(temp_mean) - 0.5 * sg * sg;
    return exp(temp_mean + temp_h * sg);
}

__device__ __forceinline__ double kl_eigvect(const KlModes &M, int ndim, int j, const double xa[3]) {
    double f[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double wx = M.w[j][d] * xa[d];
        f[d] = M.alpha[j][d] * (M.use_cos[j][d] ? cos(wx) : sin(wx));
    }
    return ndim == 2 ? f[0] * f[1] : f[0] * f[1] * f[2];
}

// FP64 synthesis into the per-particle cohesion of a context
__global__ __launch_bounds__(256) void k_kl_fp64(const double *x, const double *y, const double *z, const int *type,
                                                 const KlModes *M, KlParams kp, int first, double *out) {
    const int i = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= first + kp.n) return;
    const int t = type[i];
    if (t == 0 || t == 7) return;
    const double xa[3] = {x[i] - kp.x_trans[0], y[i] - kp.x_trans[1], kp.ndim == 3 ? z[i] - kp.x_trans[2] : 0.0};
    double temp_h = 0.0;
    for (int j = 0; j < kKlTerms; j++) temp_h = temp_h + M->seig[j] * kl_eigvect(*M, kp.ndim, j, xa) * M->xi[j];
    const bool match = (kp.field <= 3 && t == 2) || (kp.field >= 4 && t == 1);
    if (match) out[i] = kl_value(kp, t, y[i], temp_h);
}

// Phi[i][j] = sqrt(lambda_j) phi_j(x_i) (FP64, stored FP32); rows past n are zero
__global__ __launch_bounds__(256) void k_kl_phi(const double *xp, const int *type, const KlModes *M, KlParams kp,
                                                int npad, float *phi) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= npad * kKlTerms) return;
    const int i = g / kKlTerms, j = g % kKlTerms;
    float v = 0.0f;
    if (i < kp.n && type[i] != 0 && type[i] != 7) {
        const double xa[3] = {xp[3 * i] - kp.x_trans[0], xp[3 * i + 1] - kp.x_trans[1],
                              kp.ndim == 3 ? xp[3 * i + 2] - kp.x_trans[2] : 0.0};
        v = (float)(M->seig[j] * kl_eigvect(*M, kp.ndim, j, xa));
    }
    phi[(size_t)i * kKlTerms + j] = v;
}

typedef float f32x16 __attribute__((ext_vector_type(16)));

// H[npad][rpad] = Phi[npad][200] * Xi[200][rpad] on the matrix cores (v_mfma_f32_32x32x2_f32:
// lane l supplies A[l % 32][k0 + l / 32] and B[k0 + l / 32][l % 32] and holds
// D[8 (q / 4) + 4 (l / 32) + q % 4][l % 32] in accumulator q). A workgroup of four waves owns 32
// rows of H and kKlCols columns: the 32 x 200 block of Phi is one contiguous 25.6 KB stretch, staged
// into LDS with coalesced loads, and each wave multiplies it with its kKlCols / 4 columns
// of Xi (read from L2, 128 contiguous bytes per half-wave). Phi is read from HBM once and every
// A fragment serves kKlTiles tiles. npad and rpad are multiples of 32; grid.y covers the columns
// in groups of kKlCols.
constexpr int kKlTiles = 2;                 // 32-column tiles per wave
constexpr int kKlCols = 4 * 32 * kKlTiles;  // columns per workgroup

__global__ __launch_bounds__(256) void k_kl_gemm(const float *phi, const float *xi, int rpad, float *h) {
    // rows padded to kKlTerms + 1 floats: the 32 lanes of a half-wave read 32 rows at one k, and an
    // odd row stride puts them in 32 different banks
    constexpr int kLd = kKlTerms + 1;
    __shared__ float a_lds[32 * kLd];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int m0 = blockIdx.x * 32;
    const float *src = phi + (size_t)m0 * kKlTerms;
    for (int q = tid; q < 32 * kKlTerms; q += 256) a_lds[(q / kKlTerms) * kLd + q % kKlTerms] = src[q];
    __syncthreads();
    const int r = lane & 31, kk = lane >> 5;
    const int c0 = blockIdx.y * kKlCols + wv * 32 * kKlTiles;
    f32x16 acc[kKlTiles];
#pragma unroll
    for (int t = 0; t < kKlTiles; t++)
#pragma unroll
        for (int q = 0; q < 16; q++) acc[t][q] = 0.0f;
    bool live[kKlTiles];
#pragma unroll
    for (int t = 0; t < kKlTiles; t++) live[t] = c0 + 32 * t < rpad;
    if (!live[0]) return;  // no barrier follows
#pragma unroll 4
    for (int k0 = 0; k0 < kKlTerms; k0 += 2) {
        const float a = a_lds[r * kLd + k0 + kk];
#pragma unroll
        for (int t = 0; t < kKlTiles; t++) {
            if (!live[t]) continue;
            const float b = xi[(size_t)(k0 + kk) * rpad + c0 + 32 * t + r];
            acc[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < kKlTiles; t++) {
        if (!live[t]) continue;
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int row = 8 * (q / 4) + 4 * kk + (q % 4);
            h[(size_t)(m0 + row) * rpad + c0 + 32 * t + r] = acc[t][q];
        }
    }
}

// value[rr][i] from H[i][rr]; particles of other types keep fill
__global__ __launch_bounds__(256) void k_kl_values(const float *h, const double *xp, const int *type, KlParams kp,
                                                   int nreal, int rpad, double fill, double *value) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= kp.n * nreal) return;
    const int rr = g / kp.n, i = g % kp.n;
    const int t = type[i];
    const bool match = (kp.field <= 3 && t == 2) || (kp.field >= 4 && t == 1);
    double v = fill;
    if (match && t != 0 && t != 7) v = kl_value(kp, t, xp[3 * i + 1], (double)h[(size_t)i * rpad + rr]);
    value[(size_t)rr * kp.n + i] = v;
}

}  // namespace sphg

namespace {

void kl_modes_device(const sphg_kl_field &f, const double *xi, sphg::KlModes &m) {
    for (int j = 0; j < SPHG_KL_TERMS; j++) {
        for (int d = 0; d < 3; d++) {
            const int idx = f.order[j][d];
            const bool use = d < f.ndim;
            const double w = use ? f.wi[idx][d] : 0.0, a = f.a[d];
            m.w[j][d] = w;
            m.use_cos[j][d] = ((idx + 1) % 2 == 1) ? 1 : 0;
            // solve_faii (CPP_17-1_K-L_expansion_non.cpp:45-60)
            if (!use) m.alpha[j][d] = 1.0;
            else if ((idx + 1) % 2 == 1) m.alpha[j][d] = 1.0 / sqrt(sin(2.0 * w * a) / 2.0 / w + a);
            else m.alpha[j][d] = 1.0 / sqrt(-sin(2.0 * w * a) / 2.0 / w + a);
        }
        m.seig[j] = sqrt(f.eigval[j]);
        m.xi[j] = xi ? xi[j] : 0.0;
    }
}

sphg::KlParams kl_params(const sphg_kl_field &f, int n) {
    sphg::KlParams kp{};
    kp.ndim = f.ndim;
    kp.dist = f.dist;
    kp.field = f.field;
    kp.n = n;
    kp.mean_value = f.mean_value;
    kp.sd_value = f.sd_value;
    kp.coe_depth = f.coe_depth;
    kp.soil_max = f.soil_max;
    kp.water_max = f.water_max;
    for (int k = 0; k < 3; k++) kp.x_trans[k] = f.x_trans[k];
    return kp;
}

int kl_check(const sphg_kl_field *f) {
    if (!f || (f->ndim != 2 && f->ndim != 3) || f->field < 0 || f->field > 5) return SPHG_E_ARG;
    if (f->dist != 1 && f->dist != 2) return 14;  // the reference's error for an unknown distribution
    return SPHG_OK;
}

}  // namespace
