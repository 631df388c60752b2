// mlp_infer.hip -- fused inference of the surrogate network: every layer of
// a 128-row tile in one workgroup, activations resident in LDS between
// layers, weights streamed from L2 in MFMA fragment order.
//
// Shape: n_in <= 16 -> 256 (SiLU) x depth -> n_out <= 16.  A 256-thread
// workgroup (4 wavefronts) owns 128 rows.  Hidden layer l computes
// H_{l+1}^T = W_l^T . H_l^T on v_mfma_f32_16x16x32_bf16 with the weights as
// the MFMA A operand and the activations as the B operand, so each lane's
// result holds 4 consecutive features of one row (one 8-byte LDS store per
// 16 x 16 tile in the epilogue).  Wavefront w owns features [64 w, 64 w + 64)
// for all 128 rows: 4 x 8 tiles = 128 accumulator registers.  Per 32-wide K
// step a wavefront loads 4 weight fragments (1 KiB each, coalesced) from
// global memory and 8 activation fragments from LDS, and issues 32 MFMAs.
// The output layer (256 -> 16) splits the 128 rows over the 4 wavefronts and
// de-standardises in registers; the survival form writes P = exp(-10^y) of
// the first n_out / 2 outputs (log10 tau per energy bin) directly.
//
// Swizzled weight layout, per layer with K padded to a multiple of 32 and N
// to a multiple of 16: fragment (ft, ks) of 64 lanes x 8 bf16 at
// ((ft * KS + ks) * 64 + lane) * 8, lane -> feature n = 16 ft + lane % 16,
// k = 32 ks + 8 (lane / 16) + e.  This is exactly the 16x16x32 A-operand
// register layout (A[i][k], i = lane % 16, k = 8 (lane / 16) + e).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mlp.h"

namespace hd {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) __bf16 bf16x4;
typedef __attribute__((ext_vector_type(4))) float f32x4;

constexpr int FW = 256;          // hidden width of the fused path
constexpr int FTM = 128;         // rows per workgroup
constexpr int FXK = 32;          // first-layer K (n_in zero-padded)
constexpr int FHP = FW + 8;      // LDS row pitch of the activations (bf16)
constexpr int FXP = FXK + 8;     // LDS row pitch of the staged inputs (bf16)

// SiLU with the hardware reciprocal (1 ulp): the result is rounded to BF16
__device__ __forceinline__ float silu_f(float z) { return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }

// one hidden layer's products for this wavefront: acc[f][rt] (features
// 16 (4 wave + f) + 4 (lane / 16) + e, rows 16 rt + lane % 16) starts at the
// bias and accumulates KS K-steps; the next step's weight fragments are
// loaded while the current step's MFMAs issue
template <int KS, int PITCH>
__device__ __forceinline__ void layer_products(const __bf16* __restrict__ src, const bf16x8* __restrict__ W,
                                               const float* __restrict__ b, int wave, int lane, f32x4 (&acc)[4][8])
{
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        const int n0 = 16 * (4 * wave + f) + 4 * (lane >> 4);
        const f32x4 bv = f32x4{b[n0], b[n0 + 1], b[n0 + 2], b[n0 + 3]};
#pragma unroll
        for (int rt = 0; rt < 8; ++rt) acc[f][rt] = bv;
    }
    bf16x8 wf[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) wf[f] = W[((4 * wave + f) * KS) * 64 + lane];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
        bf16x8 wn[4];
        if (ks + 1 < KS) {
#pragma unroll
            for (int f = 0; f < 4; ++f) wn[f] = W[((4 * wave + f) * KS + ks + 1) * 64 + lane];
        }
#pragma unroll
        for (int rt = 0; rt < 8; ++rt) {
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(src + rt * 16 * PITCH + 32 * ks);
#pragma unroll
            for (int f = 0; f < 4; ++f)
                acc[f][rt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf[f], a, acc[f][rt], 0, 0, 0);
        }
        if (ks + 1 < KS) {
#pragma unroll
            for (int f = 0; f < 4; ++f) wf[f] = wn[f];
        }
    }
}

// SiLU of the accumulators -> BF16 activations in LDS (after every
// wavefront has finished reading the layer's input)
__device__ __forceinline__ void layer_store(__bf16* Hs, int wave, int lane, const f32x4 (&acc)[4][8])
{
    __syncthreads();
    const int lr = lane & 15;
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        const int n0 = 16 * (4 * wave + f) + 4 * (lane >> 4);
#pragma unroll
        for (int rt = 0; rt < 8; ++rt) {
            bf16x4 h;
            h[0] = (__bf16)silu_f(acc[f][rt][0]);
            h[1] = (__bf16)silu_f(acc[f][rt][1]);
            h[2] = (__bf16)silu_f(acc[f][rt][2]);
            h[3] = (__bf16)silu_f(acc[f][rt][3]);
            *reinterpret_cast<bf16x4*>(&Hs[(16 * rt + lr) * FHP + n0]) = h;
        }
    }
    __syncthreads();
}

__global__ void k_swizzle(const __bf16* __restrict__ w, int K, int N, int KS, int NT, bf16x8* __restrict__ out)
{
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= long(NT) * KS * 64) return;
    const int lane = int(i & 63);
    const long f = i >> 6;
    const int ks = int(f % KS), ft = int(f / KS);
    const int n = 16 * ft + (lane & 15);
    const int k0 = 32 * ks + 8 * (lane >> 4);
    bf16x8 v;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int k = k0 + e;
        v[e] = (k < K && n < N) ? w[long(k) * N + n] : (__bf16)0.0f;
    }
    out[i] = v;
}

template <int MODE>
__global__ void __launch_bounds__(256, 2) k_mlp_fused(FusedArgs A)
{
    __shared__ __attribute__((aligned(16))) __bf16 Hs[FTM * FHP];
    __bf16* Xs = Hs;   // the staged inputs alias the activations: layer 0 reads them before its epilogue writes
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const long r0 = long(blockIdx.x) * FTM;

    {   // standardised inputs -> Xs[128][32] bf16 (zero padding)
        const int r = t >> 1, c0 = (t & 1) * 16;
        const long row = r0 + r;
        bf16x8 v0, v1;
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const int c = c0 + e;
            float f = 0.0f;
            if (row < A.M && c < A.n_in) f = (A.x[row * A.n_in + c] - A.xm[c]) * A.xi[c];
            if (e < 8) v0[e] = (__bf16)f;
            else v1[e - 8] = (__bf16)f;
        }
        *reinterpret_cast<bf16x8*>(&Xs[r * FXP + c0]) = v0;
        *reinterpret_cast<bf16x8*>(&Xs[r * FXP + c0 + 8]) = v1;
    }
    __syncthreads();

    const int lr = lane & 15, lk = 8 * (lane >> 4);
    const bf16x8* __restrict__ Wall = reinterpret_cast<const bf16x8*>(A.w);
    {
        f32x4 acc[4][8];
        layer_products<1, FXP>(Xs + lr * FXP + lk, Wall + A.wo[0], A.bias + A.bo[0], wave, lane, acc);
        layer_store(Hs, wave, lane, acc);
    }
    for (int l = 1; l < A.L - 1; ++l) {
        f32x4 acc[4][8];
        layer_products<FW / 32, FHP>(Hs + lr * FHP + lk, Wall + A.wo[l], A.bias + A.bo[l], wave, lane, acc);
        layer_store(Hs, wave, lane, acc);
    }

    // output layer: K = 256, one 16-feature tile; wavefront w takes rows [32 w, 32 w + 32)
    const int l = A.L - 1;
    const bf16x8* __restrict__ W = Wall + A.wo[l];
    f32x4 o[2] = {f32x4{0.0f, 0.0f, 0.0f, 0.0f}, f32x4{0.0f, 0.0f, 0.0f, 0.0f}};
#pragma unroll
    for (int ks = 0; ks < FW / 32; ++ks) {
        const bf16x8 wf = W[ks * 64 + lane];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(&Hs[(16 * (2 * wave + q) + lr) * FHP + 32 * ks + lk]);
            o[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf, a, o[q], 0, 0, 0);
        }
    }
    const float* __restrict__ b = A.bias + A.bo[l];
    const int nb = A.n_out / 2;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const long row = r0 + 16 * (2 * wave + q) + lr;
        if (row >= A.M) continue;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int f = 4 * (lane >> 4) + e;
            if (f >= A.n_out) continue;
            const float v = (o[q][e] + b[f]) * A.ys[f] + A.ym[f];
            if constexpr (MODE == 0) {
                A.y[row * A.n_out + f] = v;
            } else {
                if (f < nb) A.y[row * nb + f] = __expf(-exp10f(v));
            }
        }
    }
}

hipError_t mlp_swizzle(const __bf16* w, int K, int N, int KS, int NT, __bf16* out, hipStream_t s)
{
    const long n = long(NT) * KS * 64;
    k_swizzle<<<dim3(unsigned((n + 255) / 256)), dim3(256), 0, s>>>(w, K, N, KS, NT, reinterpret_cast<bf16x8*>(out));
    return hipGetLastError();
}

hipError_t mlp_fused(const FusedArgs& a, int mode, hipStream_t s)
{
    if (a.M <= 0) return hipSuccess;
    const dim3 grid(unsigned((a.M + FTM - 1) / FTM));
    if (mode == 0) k_mlp_fused<0><<<grid, dim3(256), 0, s>>>(a);
    else k_mlp_fused<1><<<grid, dim3(256), 0, s>>>(a);
    return hipGetLastError();
}

}  // namespace hd
