// mlp.hip -- the learned surrogate: BF16 MFMA multilayer perceptron with
// FP32 accumulation, trained and served on the GPU by these kernels.
//
// Layers: n_in -> width x depth (SiLU) -> n_out (linear).  Activations are
// BF16 row-major [rows x features]; weights are BF16 [in x out] copies of FP32
// master weights that Adam updates.  Every matrix product is one GEMM kernel
//   C[M x N] = op(A)[M x K] . op(B)[K x N]
// on v_mfma_f32_16x16x32_bf16: a 256-thread workgroup owns a 64 x 64 tile of
// C (each wavefront a 32 x 32 quarter = 2 x 2 MFMA tiles) and walks K in
// steps of 32 through LDS tiles stored K-contiguous (As[m][k], Bt[n][k]), so
// each lane's 8-element MFMA operand is one 16-byte LDS read.  Epilogues fuse
// bias + SiLU (forward), the SiLU derivative (backward), or store FP32
// (weight gradients, split over K with a deterministic reduction).
//
// MFMA operand layout (16x16x32 bf16, one wavefront): A[i][k] with
// i = lane % 16, k = 8 (lane / 16) + e; B[k][j] with j = lane % 16,
// k = 8 (lane / 16) + e; D[i][j] with j = lane % 16, i = 4 (lane / 16) + e.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mlp.h"

namespace hd {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

constexpr int BM = 64, BN = 64, BK = 32, KPAD = 8;

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z)
{
    const float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

template <bool TA, bool TB, int EPI>
__global__ void __launch_bounds__(256) k_gemm(GemmArgs g)
{
    __shared__ __attribute__((aligned(16))) __bf16 As[BM][BK + KPAD];
    __shared__ __attribute__((aligned(16))) __bf16 Bt[BN][BK + KPAD];
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    // split-K: this workgroup covers K range [kb, ke)
    const int kb = blockIdx.z * g.k_chunk;
    const int ke = min(g.K, kb + g.k_chunk);
    // 16-byte tile loads need 16-byte aligned rows
    const bool a_vec = (g.lda % 8 == 0) && (reinterpret_cast<uintptr_t>(g.A) % 16 == 0);
    const bool b_vec = (g.ldb % 8 == 0) && (reinterpret_cast<uintptr_t>(g.B) % 16 == 0);

    f32x4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

    for (int k0 = kb; k0 < ke; k0 += BK) {
        // ---- A tile (64 x 32) -> As[m][k].  Each thread moves 8 elements that
        // are consecutive in memory: one 16-byte load where all 8 lie inside the
        // matrix (and rows start 16-byte aligned), element loads at the edges.
        if constexpr (!TA) {
            const int r = t >> 2, c = (t & 3) * 8;
            const int m = m0 + r, k = k0 + c;
            bf16x8 v;
            if (a_vec && m < g.M && k + 8 <= ke) {
                v = *reinterpret_cast<const bf16x8*>(g.A + long(m) * g.lda + k);
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) v[e] = (m < g.M && k + e < ke) ? g.A[long(m) * g.lda + k + e] : (__bf16)0.0f;
            }
            *reinterpret_cast<bf16x8*>(&As[r][c]) = v;
        } else {
            const int kk = t >> 3, mm = (t & 7) * 8;
            const int k = k0 + kk, m = m0 + mm;
            bf16x8 v;
            if (a_vec && k < ke && m + 8 <= g.M) {
                v = *reinterpret_cast<const bf16x8*>(g.A + long(k) * g.lda + m);
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) v[e] = (m + e < g.M && k < ke) ? g.A[long(k) * g.lda + m + e] : (__bf16)0.0f;
            }
#pragma unroll
            for (int e = 0; e < 8; ++e) As[mm + e][kk] = v[e];
        }
        // ---- B tile (32 x 64) -> Bt[n][k], the same way
        if constexpr (!TB) {
            const int kk = t >> 3, nn = (t & 7) * 8;
            const int k = k0 + kk, n = n0 + nn;
            bf16x8 v;
            if (b_vec && k < ke && n + 8 <= g.N) {
                v = *reinterpret_cast<const bf16x8*>(g.B + long(k) * g.ldb + n);
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) v[e] = (n + e < g.N && k < ke) ? g.B[long(k) * g.ldb + n + e] : (__bf16)0.0f;
            }
#pragma unroll
            for (int e = 0; e < 8; ++e) Bt[nn + e][kk] = v[e];
        } else {
            const int r = t >> 2, c = (t & 3) * 8;
            const int n = n0 + r, k = k0 + c;
            bf16x8 v;
            if (b_vec && n < g.N && k + 8 <= ke) {
                v = *reinterpret_cast<const bf16x8*>(g.B + long(n) * g.ldb + k);
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) v[e] = (n < g.N && k + e < ke) ? g.B[long(n) * g.ldb + k + e] : (__bf16)0.0f;
            }
            *reinterpret_cast<bf16x8*>(&Bt[r][c]) = v;
        }
        __syncthreads();
        const int ko = 8 * (lane >> 4);
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const bf16x8 av = *reinterpret_cast<const bf16x8*>(&As[wm * 32 + a * 16 + (lane & 15)][ko]);
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const bf16x8 bv = *reinterpret_cast<const bf16x8*>(&Bt[wn * 32 + b * 16 + (lane & 15)][ko]);
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[a][b], 0, 0, 0);
            }
        }
        __syncthreads();
    }

    // ---- epilogue
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int m = m0 + wm * 32 + a * 16 + 4 * (lane >> 4) + e;
                const int n = n0 + wn * 32 + b * 16 + (lane & 15);
                if (m >= g.M || n >= g.N) continue;
                const float v = acc[a][b][e];
                const long o = long(m) * g.ldc + n;
                if constexpr (EPI == EPI_FWD_SILU) {
                    const float z = v + g.bias[n];
                    if (g.Cf) g.Cf[o] = z;
                    g.Cb[o] = (__bf16)silu(z);
                } else if constexpr (EPI == EPI_FWD_LINEAR) {
                    g.Cf[o] = (v + g.bias[n]) * g.scale[n] + g.shift[n];
                } else if constexpr (EPI == EPI_BWD_SILU) {
                    g.Cb[o] = (__bf16)(v * dsilu(g.aux[o]));
                } else {   // EPI_F32: split-K partial sums
                    g.Cf[long(blockIdx.z) * g.M * g.N + long(m) * g.N + n] = v;
                }
            }
}

// sum of `parts` partial matrices (deterministic split-K reduction)
__global__ void k_sum_parts(const float* p, int parts, long n, float* out)
{
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int q = 0; q < parts; ++q) s += p[q * n + i];
    out[i] = s;
}

// The same for few outputs and many parts (the bias gradients' row-chunk
// partials): one workgroup per output, each thread adding a strided share of
// the parts, then a fixed tree over the 256 threads (deterministic).
__global__ void __launch_bounds__(256) k_sum_parts_wide(const float* p, int parts, long n, float* out)
{
    __shared__ float wsum[4];
    const long i = blockIdx.x;
    float s = 0.0f;
    for (int q = threadIdx.x; q < parts; q += 256) s += p[q * n + i];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[i] = (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
}

// column sums of a bf16 [M x N] matrix in row chunks -> partials [chunks x N]
__global__ void k_colsum_bf16(const __bf16* x, int M, int N, int rows_per_chunk, float* part)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int r0 = blockIdx.y * rows_per_chunk;
    const int r1 = min(M, r0 + rows_per_chunk);
    float s = 0.0f;
    for (int r = r0; r < r1; ++r) s += float(x[long(r) * N + n]);
    part[long(blockIdx.y) * N + n] = s;
}

// raw features [M x n_in] -> normalised bf16 [M x ld] (padding columns 0)
__global__ void k_pack_features(const float* x, long M, int n_in, int ld, const float* mean, const float* inv_std,
                                __bf16* out)
{
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= M * ld) return;
    const long m = i / ld;
    const int c = int(i % ld);
    out[i] = (c < n_in) ? (__bf16)((x[m * n_in + c] - mean[c]) * inv_std[c]) : (__bf16)0.0f;
}

// MSE on normalised targets: dY = 2 (y - (t - mean)/std) / (M n_out) as bf16;
// per-block partial sums of the squared error
__global__ void k_mse(const float* y, const float* t, long M, int n_out, const float* mean, const float* inv_std,
                      __bf16* dy, float* part)
{
    __shared__ float red[256];
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    float sq = 0.0f;
    if (i < M * n_out) {
        const int c = int(i % n_out);
        const float d = y[i] - (t[i] - mean[c]) * inv_std[c];
        sq = d * d;
        dy[i] = (__bf16)(2.0f * d / float(M * n_out));
    }
    red[threadIdx.x] = sq;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Adam on FP32 masters; refreshes the BF16 weight copies (w16 may be null
// for biases).  Weight rows beyond `rows_real` of a padded matrix stay zero.
__global__ void k_adam(float* p, const float* g, float* m, float* v, long n, float lr, float b1, float b2, float eps,
                       float c1, float c2, __bf16* w16)
{
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    if (w16) w16[i] = (__bf16)pi;
}

__global__ void k_f32_to_bf16(const float* x, long n, __bf16* y)
{
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (__bf16)x[i];
}

template <bool TA, bool TB, int EPI>
static hipError_t gemm_t(const GemmArgs& g, int splits, hipStream_t s)
{
    dim3 grid((g.N + BN - 1) / BN, (g.M + BM - 1) / BM, splits);
    hipLaunchKernelGGL((k_gemm<TA, TB, EPI>), grid, dim3(256), 0, s, g);
    return hipGetLastError();
}

hipError_t mlp_gemm(const GemmArgs& g0, bool ta, bool tb, int epi, int splits, hipStream_t s)
{
    GemmArgs g = g0;
    if (splits < 1) splits = 1;
    g.k_chunk = ((g.K + splits - 1) / splits + BK - 1) / BK * BK;
    if (!ta && !tb && epi == EPI_FWD_SILU) return gemm_t<false, false, EPI_FWD_SILU>(g, 1, s);
    if (!ta && !tb && epi == EPI_FWD_LINEAR) return gemm_t<false, false, EPI_FWD_LINEAR>(g, 1, s);
    if (!ta && tb && epi == EPI_BWD_SILU) return gemm_t<false, true, EPI_BWD_SILU>(g, 1, s);
    if (ta && !tb && epi == EPI_F32) return gemm_t<true, false, EPI_F32>(g, splits, s);
    if (!ta && !tb && epi == EPI_F32) return gemm_t<false, false, EPI_F32>(g, splits, s);
    return hipErrorInvalidValue;
}

static inline int nblk(long n) { return int((n + 255) / 256); }

hipError_t mlp_sum_parts(const float* p, int parts, long n, float* out, hipStream_t s)
{
    // one thread per output when there are enough outputs to fill the chip;
    // a workgroup per output for the bias gradients (<= 256 outputs, 256 parts),
    // where a thread per output ran a 256-long dependent chain (about 50 us)
    if (n <= 4096 && parts >= 64)
        hipLaunchKernelGGL(k_sum_parts_wide, dim3(unsigned(n)), dim3(256), 0, s, p, parts, n, out);
    else
        hipLaunchKernelGGL(k_sum_parts, dim3(nblk(n)), dim3(256), 0, s, p, parts, n, out);
    return hipGetLastError();
}
hipError_t mlp_colsum(const __bf16* x, int M, int N, int rows_per_chunk, float* part, hipStream_t s)
{
    dim3 grid((N + 255) / 256, (M + rows_per_chunk - 1) / rows_per_chunk);
    hipLaunchKernelGGL(k_colsum_bf16, grid, dim3(256), 0, s, x, M, N, rows_per_chunk, part);
    return hipGetLastError();
}
hipError_t mlp_pack(const float* x, long M, int n_in, int ld, const float* mean, const float* inv_std, __bf16* out,
                    hipStream_t s)
{
    hipLaunchKernelGGL(k_pack_features, dim3(nblk(M * ld)), dim3(256), 0, s, x, M, n_in, ld, mean, inv_std, out);
    return hipGetLastError();
}
hipError_t mlp_mse(const float* y, const float* t, long M, int n_out, const float* mean, const float* inv_std,
                   __bf16* dy, float* part, int* nparts, hipStream_t s)
{
    *nparts = nblk(M * n_out);
    hipLaunchKernelGGL(k_mse, dim3(*nparts), dim3(256), 0, s, y, t, M, n_out, mean, inv_std, dy, part);
    return hipGetLastError();
}
hipError_t mlp_adam(float* p, const float* g, float* m, float* v, long n, float lr, float b1, float b2, float eps,
                    float c1, float c2, __bf16* w16, hipStream_t s)
{
    hipLaunchKernelGGL(k_adam, dim3(nblk(n)), dim3(256), 0, s, p, g, m, v, n, lr, b1, b2, eps, c1, c2, w16);
    return hipGetLastError();
}
hipError_t mlp_to_bf16(const float* x, long n, __bf16* y, hipStream_t s)
{
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(nblk(n)), dim3(256), 0, s, x, n, y);
    return hipGetLastError();
}

}  // namespace hd
