// replay.hip -- transfer over a resident geodesic cache (image-from-cache).
//
// radiative_transfer::integrate_kerr_ray_spectral (src/radiative_transfer.cpp:
// 457-780) for every ray of a KGEO cache, as in
// apps/compute_kerr_image_from_cache.cpp:605-658.  The reference walks each
// ray sequentially (tau before a point weights that point's emission); here a
// wavefront owns a ray and processes 64 consecutive points per iteration:
// coalesced SoA loads of (r, theta, dl, z), per-bin dtau and emission terms,
// a wavefront exclusive prefix sum of dtau gives tau before every point, and
// a wavefront reduction adds the 64 emission terms.  HBM traffic is the
// algorithmic 32 B per point.
#include "kernels.h"

namespace hd {

template <typename T>
__device__ __forceinline__ T wave_excl_scan(T x, T& total)
{
    const int lane = threadIdx.x & 63;
    T incl = x;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const T y = __shfl_up(incl, off, 64);
        if (lane >= off) incl += y;
    }
    total = __shfl(incl, 63, 64);
    return incl - x;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T x)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

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
#pragma unroll
                for (int b = 0; b < NBMAX; ++b) {
                    if (b >= P.nb) continue;
                    const double E = P.emid[b] * pq.z;
                    const double lnEb = P.lnemid[b] + pq.lnz;
                    double dtau = 0.0;
                    if (ok && pq.rho > 0.0 && E >= P.e_lo && E <= P.e_hi)
                        dtau = pq.nbar_dl * exp(sigma_ln(lnE, lnS, slope, n, lnEb));
                    double tot;
                    const double before = tau[b] + wave_excl_scan(dtau, tot);
                    double em = 0.0;
                    if (ok && P.spec_ok && pq.g3Adl != 0.0) em = pq.g3Adl * exp(-P.gam * lnEb - E * P.invcut - before);
                    I[b] += wave_sum(em);
                    tau[b] += tot;
                }
            } else {
                const float lnz = float(pq.lnz);
                const float lP = (pq.g3Adl > 0.0) ? float(log(pq.g3Adl)) : 0.0f;
                const float lN = (pq.nbar_dl > 0.0) ? float(log(pq.nbar_dl)) : 0.0f;
                const float gam = float(P.gam), invcut = float(P.invcut);
#pragma unroll
                for (int b = 0; b < NBMAX; ++b) {
                    if (b >= P.nb) continue;
                    const double E = P.emid[b] * pq.z;
                    const float lnEb = float(P.lnemid[b]) + lnz;
                    float dtau = 0.0f;
                    if (ok && pq.rho > 0.0 && pq.nbar_dl > 0.0 && E >= P.e_lo && E <= P.e_hi)
                        dtau = __expf(lN + sigma_ln_f(lnEf, lnSf, slopef, n, lnEb));
                    float tot;
                    const float before = ftau[b].s + wave_excl_scan(dtau, tot);
                    // emission in log form; the wavefront sum of e^{L - m} uses the chunk maximum m
                    float L = -INFINITY;
                    if (ok && P.spec_ok && pq.g3Adl > 0.0) L = lP - gam * lnEb - float(E) * invcut - before;
                    float m = L;
#pragma unroll
                    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
                    if (m > -INFINITY) {
                        const float s = wave_sum(__expf(L - m));
                        fI[b].add_log(m + __logf(s));
                    }
                    ftau[b].add(tot);
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
