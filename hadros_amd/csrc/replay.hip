// replay.hip -- transfer over a resident geodesic cache (image-from-cache).
//
// radiative_transfer::integrate_kerr_ray_spectral (src/radiative_transfer.cpp:
// 457-780) for every ray of a KGEO cache, as in
// apps/compute_kerr_image_from_cache.cpp:605-658.  The reference walks each
// ray sequentially (tau before a point weights that point's emission); here a
// wavefront owns a ray and processes 64 consecutive points per iteration:
// coalesced SoA loads of (r, theta, dl, z), per-bin dtau and emission terms,
// a wavefront prefix sum of dtau gives tau before every point, and a second
// prefix sum adds the 64 emission terms (its lane 63).  The prefix sums are
// DPP lane moves (row shifts 1, 2, 4, 8, then the row-15 and row-31
// broadcasts), run for all bins in lockstep so the bins' dependency chains
// overlap.  HBM traffic is the algorithmic 32 B per point.
#include "kernels.h"

namespace hd {

// DPP move of one 32-bit word; out-of-row and masked-off lanes keep `old`
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ int dpp_word(int old, int v)
{
    return __builtin_amdgcn_update_dpp(old, v, CTRL, ROW_MASK, 0xf, false);
}

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_move(double v)   // identity +0.0
{
    const unsigned long long u = static_cast<unsigned long long>(__double_as_longlong(v));
    const unsigned lo = unsigned(dpp_word<CTRL, ROW_MASK>(0, int(u & 0xffffffffull)));
    const unsigned hi = unsigned(dpp_word<CTRL, ROW_MASK>(0, int(u >> 32)));
    return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_move(float v, float identity)
{
    return __int_as_float(dpp_word<CTRL, ROW_MASK>(__float_as_int(identity), __float_as_int(v)));
}

// inclusive prefix sums over the 64 lanes of N independent values per lane
template <int N>
__device__ __forceinline__ void wave_scan_add(double (&x)[N])
{
#define HD_SCAN_STEP(CTRL, RM)                                            \
    _Pragma("unroll") for (int b = 0; b < N; ++b) x[b] += dpp_move<CTRL, RM>(x[b]);
    HD_SCAN_STEP(0x111, 0xf)   // row_shr:1
    HD_SCAN_STEP(0x112, 0xf)   // row_shr:2
    HD_SCAN_STEP(0x114, 0xf)   // row_shr:4
    HD_SCAN_STEP(0x118, 0xf)   // row_shr:8
    HD_SCAN_STEP(0x142, 0xa)   // row_bcast:15 into rows 1, 3
    HD_SCAN_STEP(0x143, 0xc)   // row_bcast:31 into rows 2, 3
#undef HD_SCAN_STEP
}

template <int N>
__device__ __forceinline__ void wave_scan_add(float (&x)[N])
{
#define HD_SCAN_STEP(CTRL, RM)                                            \
    _Pragma("unroll") for (int b = 0; b < N; ++b) x[b] += dpp_move<CTRL, RM>(x[b], 0.0f);
    HD_SCAN_STEP(0x111, 0xf)
    HD_SCAN_STEP(0x112, 0xf)
    HD_SCAN_STEP(0x114, 0xf)
    HD_SCAN_STEP(0x118, 0xf)
    HD_SCAN_STEP(0x142, 0xa)
    HD_SCAN_STEP(0x143, 0xc)
#undef HD_SCAN_STEP
}

// inclusive prefix maxima (identity -inf)
template <int N>
__device__ __forceinline__ void wave_scan_max(float (&x)[N])
{
#define HD_SCAN_STEP(CTRL, RM)                                            \
    _Pragma("unroll") for (int b = 0; b < N; ++b) x[b] = fmaxf(x[b], dpp_move<CTRL, RM>(x[b], -INFINITY));
    HD_SCAN_STEP(0x111, 0xf)
    HD_SCAN_STEP(0x112, 0xf)
    HD_SCAN_STEP(0x114, 0xf)
    HD_SCAN_STEP(0x118, 0xf)
    HD_SCAN_STEP(0x142, 0xa)
    HD_SCAN_STEP(0x143, 0xc)
#undef HD_SCAN_STEP
}

__device__ __forceinline__ double lane63(double v)
{
    const unsigned long long u = static_cast<unsigned long long>(__double_as_longlong(v));
    const unsigned lo = unsigned(__builtin_amdgcn_readlane(int(u & 0xffffffffull), 63));
    const unsigned hi = unsigned(__builtin_amdgcn_readlane(int(u >> 32), 63));
    return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

__device__ __forceinline__ float lane63(float v) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63)); }

template <int PREC, int NBMAX>
__global__ void __launch_bounds__(256) k_replay(ReplayArgs A)
{
    extern __shared__ double smem[];
    const DevRt& P = A.rt;
    const int n = P.n_sigma;
    const double* lnE = smem;
    const double* lnS = smem + n;
    const double* slope = smem + 2 * n;
    const float* lnEf = reinterpret_cast<const float*>(smem);
    const float* lnSf = lnEf + n;
    const float* slopef = lnEf + 2 * n;
    if constexpr (PREC == HADROS_PREC_FP64) {
        for (int t = threadIdx.x; t < 3 * n; t += blockDim.x) smem[t] = A.sig[t];
    } else {
        float* sf = reinterpret_cast<float*>(smem);
        for (int t = threadIdx.x; t < 3 * n; t += blockDim.x) sf[t] = A.sigf[t];
    }
    __syncthreads();

    const int lane = threadIdx.x & 63;
    const int waves_per_block = blockDim.x >> 6;
    const long nwaves = long(gridDim.x) * waves_per_block;
    unsigned long long pts_done = 0;
    for (long ray = long(blockIdx.x) * waves_per_block + (threadIdx.x >> 6); ray < A.n_rays; ray += nwaves) {
        const long b0 = A.offset[ray], b1 = A.offset[ray + 1];
        double tau[NBMAX], I[NBMAX];
        Kahan ftau[NBMAX];
        ScaledKahan fI[NBMAX];
#pragma unroll
        for (int b = 0; b < NBMAX; ++b) {
            tau[b] = 0.0;
            I[b] = 0.0;
            ftau[b] = {0.0f, 0.0f};
            fI[b] = {{0.0f, 0.0f}, 0.0f, 0};
        }
        for (long base = b0; base < b1; base += 64) {
            const long q = base + lane;
            PointQ pq;
            bool ok = false;
            if (q < b1) ok = point_q(P, A.r[q], A.th[q], A.dl[q], A.z[q], pq);
            if (!ok) {
                pq.z = 1.0;
                pq.lnz = 0.0;
                pq.g3Adl = 0.0;
                pq.nbar_dl = 0.0;
                pq.rho = 0.0;
            }
            if constexpr (PREC == HADROS_PREC_FP64) {
                double inc[NBMAX], em[NBMAX];
#pragma unroll
                for (int b = 0; b < NBMAX; ++b) {
                    inc[b] = 0.0;
                    if (b >= P.nb) continue;
                    const double E = P.emid[b] * pq.z;
                    const double lnEb = P.lnemid[b] + pq.lnz;
                    if (ok && pq.rho > 0.0 && E >= P.e_lo && E <= P.e_hi)
                        inc[b] = pq.nbar_dl * exp(sigma_ln(lnE, lnS, slope, n, lnEb, P.sig_lo, P.sig_inv_d));
                    em[b] = inc[b];   // dtau of this point, until the scan
                }
                wave_scan_add(inc);
#pragma unroll
                for (int b = 0; b < NBMAX; ++b) {
                    const double dtau = em[b];
                    em[b] = 0.0;
                    if (b >= P.nb) continue;
                    const double E = P.emid[b] * pq.z;
                    const double lnEb = P.lnemid[b] + pq.lnz;
                    const double before = tau[b] + (inc[b] - dtau);
                    if (ok && P.spec_ok && pq.g3Adl != 0.0) em[b] = pq.g3Adl * exp(-P.gam * lnEb - E * P.invcut - before);
                }
                wave_scan_add(em);
#pragma unroll
                for (int b = 0; b < NBMAX; ++b) {
                    if (b >= P.nb) continue;
                    I[b] += lane63(em[b]);
                    tau[b] += lane63(inc[b]);
                }
            } else {
                const float lnz = float(pq.lnz);
                const float lP = (pq.g3Adl > 0.0) ? float(log(pq.g3Adl)) : 0.0f;
                const float lN = (pq.nbar_dl > 0.0) ? float(log(pq.nbar_dl)) : 0.0f;
                const float gam = float(P.gam), invcut = float(P.invcut);
                float inc[NBMAX], dt[NBMAX], L[NBMAX], m[NBMAX];
#pragma unroll
                for (int b = 0; b < NBMAX; ++b) {
                    inc[b] = 0.0f;
                    if (b >= P.nb) continue;
                    const double E = P.emid[b] * pq.z;
                    const float lnEb = float(P.lnemid[b]) + lnz;
                    if (ok && pq.rho > 0.0 && pq.nbar_dl > 0.0 && E >= P.e_lo && E <= P.e_hi)
                        inc[b] = __expf(lN + sigma_ln_f(lnEf, lnSf, slopef, n, lnEb, float(P.sig_lo), float(P.sig_inv_d)));
                    dt[b] = inc[b];
                }
                wave_scan_add(inc);
#pragma unroll
                for (int b = 0; b < NBMAX; ++b) {
                    L[b] = -INFINITY;
                    if (b < P.nb) {
                        const double E = P.emid[b] * pq.z;
                        const float lnEb = float(P.lnemid[b]) + lnz;
                        const float before = ftau[b].s + (inc[b] - dt[b]);
                        // emission in log form; the wavefront sum of e^{L - m} uses the chunk maximum m
                        if (ok && P.spec_ok && pq.g3Adl > 0.0) L[b] = lP - gam * lnEb - float(E) * invcut - before;
                    }
                    m[b] = L[b];
                }
                wave_scan_max(m);
#pragma unroll
                for (int b = 0; b < NBMAX; ++b) {
                    m[b] = lane63(m[b]);
                    dt[b] = (m[b] > -INFINITY) ? __expf(L[b] - m[b]) : 0.0f;
                }
                wave_scan_add(dt);
#pragma unroll
                for (int b = 0; b < NBMAX; ++b) {
                    if (b >= P.nb) continue;
                    if (m[b] > -INFINITY) fI[b].add_log(m[b] + __logf(lane63(dt[b])));
                    ftau[b].add(lane63(inc[b]));
                }
            }
        }
        pts_done += (unsigned long long)(b1 - b0);
        if (lane == 0) {
            double bt[NBMAX], bp[NBMAX], bi[NBMAX];
#pragma unroll
            for (int b = 0; b < NBMAX; ++b) {
                if (PREC == HADROS_PREC_FP64) {
                    bt[b] = tau[b];
                    bi[b] = I[b];
                } else {
                    bt[b] = double(ftau[b].s) - double(ftau[b].c);
                    bi[b] = fI[b].set ? (double(fI[b].k.s) - double(fI[b].k.c)) * exp(double(fI[b].off)) : 0.0;
                }
                bp[b] = exp(-bt[b]);
            }
            double agg_t = 0.0, agg_p = 1.0, agg_i = 0.0;
            if (b1 > b0) {
                if (P.mono) {
                    if (P.nb > 0) {
                        agg_t = bt[0];
                        agg_p = bp[0];
                        agg_i = bi[0];
                    }
                } else {
                    double ws = 0.0, ts = 0.0, ps = 0.0, is = 0.0;
#pragma unroll
                    for (int b = 0; b < NBMAX; ++b) {
                        if (b >= P.nb) continue;
                        ws += P.w[b];
                        ts += P.w[b] * bt[b];
                        ps += P.w[b] * bp[b];
                        is += bi[b] * P.de[b];
                    }
                    if (ws > 0.0) {
                        agg_t = ts / ws;
                        agg_p = ps / ws;
                        agg_i = is;
                    }
                }
            }
            if (A.tau) A.tau[ray] = agg_t;
            if (A.p) A.p[ray] = agg_p;
            if (A.I) A.I[ray] = agg_i;
#pragma unroll
            for (int b = 0; b < NBMAX; ++b) {
                if (b >= P.nb) continue;
                if (A.btau) A.btau[ray * P.nb + b] = bt[b];
                if (A.bp) A.bp[ray * P.nb + b] = bp[b];
                if (A.bi) A.bi[ray * P.nb + b] = bi[b];
            }
        }
    }
    if (lane == 0 && A.counters) atomicAdd(A.counters, pts_done);
}

template <int PREC>
static void launch_p(const ReplayArgs& A, int grid, int block, size_t lds, hipStream_t s)
{
    if (A.rt.nb <= 1) hipLaunchKernelGGL((k_replay<PREC, 1>), dim3(grid), dim3(block), lds, s, A);
    else if (A.rt.nb <= 8) hipLaunchKernelGGL((k_replay<PREC, 8>), dim3(grid), dim3(block), lds, s, A);
    else hipLaunchKernelGGL((k_replay<PREC, 16>), dim3(grid), dim3(block), lds, s, A);
}

hipError_t launch_replay(const ReplayArgs& A, int precision, hipStream_t s)
{
    const int block = 256;
    const long waves = A.n_rays;
    long grid = (waves + 3) / 4;
    const long max_grid = 256L * 64;   // persistent-ish: waves loop over rays
    if (grid > max_grid) grid = max_grid;
    if (grid < 1) grid = 1;
    const size_t lds = size_t(3) * A.rt.n_sigma * (precision == HADROS_PREC_FP64 ? sizeof(double) : sizeof(float));
    if (precision == HADROS_PREC_FP64) launch_p<HADROS_PREC_FP64>(A, int(grid), block, lds, s);
    else launch_p<HADROS_PREC_FP32K>(A, int(grid), block, lds, s);
    return hipGetLastError();
}

}  // namespace hd
