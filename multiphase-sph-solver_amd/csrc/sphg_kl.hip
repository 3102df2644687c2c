// K-L random-field synthesis on the device (solve_h_2D / solve_h_3D,
// CPP_17-1_K-L_expansion_non.cpp:499-713).
//
// FP64 path (one field of one context): one thread per particle evaluates the 200 products of
// eigenfunctions, sums sqrt(lambda_j) phi_j(x) xi_j in mode order, and maps H to the property.
// Batch path (many realisations in a common domain): Phi(x) scaled by sqrt(lambda) is formed
// once in FP64 and stored as FP32, the realisations' standard normals form Xi, and H = Phi * Xi
// runs on the matrix cores (v_mfma_f32_32x32x2_f32), one workgroup per 128 x 128 tile of H.
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
// lane l supplies A[l % 32][k + l / 32] and B[k + l / 32][l % 32] and holds
// D[8 (q / 4) + 4 (l / 32) + q % 4][l % 32] in accumulator q). A workgroup of four waves owns a
// 128 x 128 tile of H, each wave a 64 x 64 quarter (2 x 2 MFMA tiles). K runs in chunks of 40:
// the 128 x 40 block of Phi and the 40 x 128 block of Xi are staged in LDS with 16-byte coalesced
// loads, so every element of Phi and Xi is read from memory once per tile it feeds, and each
// fragment read from LDS feeds two MFMAs. The LDS rows are padded (A: 42 floats, B: 160) so that
// the two half-waves' fragment reads, at k and k + 1, fall in 64 distinct banks.
constexpr int kGm = 128, kGn = 128, kGk = 40, kGlda = 42, kGldb = 160;
static_assert(kKlTerms % kGk == 0, "K chunks");

__global__ __launch_bounds__(256) void k_kl_gemm(const float *phi, const float *xi, int npad, int rpad, float *h) {
    __shared__ float a_s[kGm * kGlda];  // [row][k]
    __shared__ float b_s[kGk * kGldb];  // [k][column]
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int m0 = blockIdx.x * kGm, n0 = blockIdx.y * kGn;
    const int wm = (wv >> 1) * 64, wn = (wv & 1) * 64;
    const int r = lane & 31, kk = lane >> 5;
    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int q = 0; q < 16; q++) acc[i][j][q] = 0.0f;
    // each thread stages kPer float4 of A and of B per chunk; the next chunk's loads are issued
    // before the current chunk's MFMAs, so their latency overlaps the matrix work
    constexpr int kPer = kGm * kGk / 4 / 256;
    static_assert(kPer * 256 * 4 == kGm * kGk && kGk * kGn == kGm * kGk, "staging split");
    float4 ra[kPer], rb[kPer];
    auto load = [&](int k0) {
#pragma unroll
        for (int p = 0; p < kPer; p++) {
            const int q = tid + 256 * p;
            const int row = q / (kGk / 4), c4 = q % (kGk / 4), m = m0 + row;
            ra[p] = m < npad ? *(const float4 *)(phi + (size_t)m * kKlTerms + k0 + 4 * c4)
                             : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            const int kr = q / (kGn / 4), d4 = q % (kGn / 4), col = n0 + 4 * d4;
            rb[p] = col < rpad ? *(const float4 *)(xi + (size_t)(k0 + kr) * rpad + col)
                               : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int p = 0; p < kPer; p++) {
            const int q = tid + 256 * p;
            const int row = q / (kGk / 4), c4 = q % (kGk / 4);
            float *d = a_s + row * kGlda + 4 * c4;
            d[0] = ra[p].x;
            d[1] = ra[p].y;
            d[2] = ra[p].z;
            d[3] = ra[p].w;
            const int kr = q / (kGn / 4), d4 = q % (kGn / 4);
            *(float4 *)(b_s + kr * kGldb + 4 * d4) = rb[p];
        }
    };
    load(0);
    store();
    __syncthreads();
    for (int k0 = 0; k0 < kKlTerms; k0 += kGk) {
        const bool more = k0 + kGk < kKlTerms;
        if (more) load(k0 + kGk);
#pragma unroll 4
        for (int k = 0; k < kGk; k += 2) {
            const float a0 = a_s[(wm + r) * kGlda + k + kk], a1 = a_s[(wm + 32 + r) * kGlda + k + kk];
            const float b0 = b_s[(k + kk) * kGldb + wn + r], b1 = b_s[(k + kk) * kGldb + wn + 32 + r];
            acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b1, acc[1][1], 0, 0, 0);
        }
        __syncthreads();
        if (more) {
            store();
            __syncthreads();
        }
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int col = n0 + wn + 32 * j + r;
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const int row = m0 + wm + 32 * i + 8 * (q / 4) + 4 * kk + (q % 4);
                if (row < npad && col < rpad) h[(size_t)row * rpad + col] = acc[i][j][q];
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
