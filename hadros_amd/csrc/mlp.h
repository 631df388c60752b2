// mlp.h -- GEMM argument block and launchers of mlp.hip.
#pragma once

#include <hip/hip_runtime.h>

namespace hd {

enum { EPI_FWD_SILU = 0, EPI_FWD_LINEAR = 1, EPI_BWD_SILU = 2, EPI_F32 = 3 };

struct GemmArgs {
    const __bf16* A;   // op(A) [M x K]; stored [M x lda] or, transposed, [K x lda]
    const __bf16* B;   // op(B) [K x N]; stored [K x ldb] or, transposed, [N x ldb]
    int M, N, K, lda, ldb, ldc;
    int k_chunk;            // split-K chunk (set by mlp_gemm)
    const float* bias;      // [N] (forward epilogues)
    const float* scale;     // [N] output scale (linear epilogue)
    const float* shift;     // [N] output shift (linear epilogue)
    const float* aux;       // [M x ldc] pre-activations for the SiLU derivative
    float* Cf;              // FP32 output (pre-activation / linear output / split-K partials)
    __bf16* Cb;             // BF16 output (activation / back-propagated gradient)
};

// fused inference (mlp_infer.hip): hidden width 256, n_in <= 16, n_out <= 16
constexpr int FUSED_MAX_LAYERS = 8;
struct FusedArgs {
    const float* x;          // [M x n_in] raw features
    long M;
    int n_in, L, n_out, pad;
    const float* xm;         // feature mean
    const float* xi;         // feature 1 / std
    const void* w;           // swizzled bf16 weights (16-byte fragments)
    long wo[FUSED_MAX_LAYERS];   // per layer: first fragment
    long bo[FUSED_MAX_LAYERS];   // per layer: bias offset into `bias`
    const float* bias;
    const float* ys;         // output scale (target std)
    const float* ym;         // output shift (target mean)
    float* y;                // [M x n_out] outputs, or [M x n_out/2] survivals
};

hipError_t mlp_swizzle(const __bf16* w, int K, int N, int KS, int NT, __bf16* out, hipStream_t s);
hipError_t mlp_fused(const FusedArgs& a, int mode, hipStream_t s);

hipError_t mlp_gemm(const GemmArgs& g, bool ta, bool tb, int epi, int splits, hipStream_t s);
hipError_t mlp_sum_parts(const float* p, int parts, long n, float* out, hipStream_t s);
hipError_t mlp_colsum(const __bf16* x, int M, int N, int rows_per_chunk, float* part, hipStream_t s);
hipError_t mlp_pack(const float* x, long M, int n_in, int ld, const float* mean, const float* inv_std, __bf16* out,
                    hipStream_t s);
hipError_t mlp_mse(const float* y, const float* t, long M, int n_out, const float* mean, const float* inv_std,
                   __bf16* dy, float* part, int* nparts, hipStream_t s);
hipError_t mlp_adam(float* p, const float* g, float* m, float* v, long n, float lr, float b1, float b2, float eps,
                    float c1, float c2, __bf16* w16, hipStream_t s);
hipError_t mlp_to_bf16(const float* x, long n, __bf16* y, hipStream_t s);

}  // namespace hd
