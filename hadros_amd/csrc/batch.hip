// batch.hip -- persistent multi-camera tracer with a global ray queue.
//
// Many cameras of uneven resolution, spin, inclination, medium, source and
// spectrum in one launch.  Ray lengths differ by orders of magnitude
// (captured vs escaping vs photon-ring rays), so a fixed grid leaves most
// lanes idle behind the longest ray of their wavefront.  Here a resident grid
// of wavefronts pulls rays from one global counter: every `steps_per_chunk`
// RKF45 attempts each wavefront takes a ballot of its finished lanes, claims
// that many ray indices with one atomic add, and hands them out in lane order
// (mbcnt), so the lanes stay busy until the queue is empty.
//
// Per ray the arithmetic is the plain tracer's (same device functions, same
// operation order), so every image equals the per-camera plain run bit for
// bit.  Camera c owns global ray indices [pix0[c], pix0[c+1]), handed out in
// the tile order of batch_pixel; its outputs are stored at pix0[c] + k with
// pixel k = i*ny + j, as in the plain tracer.
#include "kernels.h"

#ifndef HD_BATCH_WAVES_PER_EU
#define HD_BATCH_WAVES_PER_EU 2
#endif

namespace hd {

struct BatchLane {
    long ray;        // global ray index in queue order, -1 = idle
    int cam;
    int step, npts, captured, exhausted, att_in_step;
    long attempts;
    double pt, pph, inv_e_obs, h;   // inv_e_obs: redshift factor of every point
    Y5 y, yp;
};

// the capture test of src/kerr_camera.cpp: r below the camera's stopping
// radius (read from the camera when tested rather than held in registers)
template <int MODE>
__device__ __forceinline__ bool batch_inside_stop(const DevCam& C, double r)
{
    if constexpr (MODE == HADROS_MODE_CFG) return r < 1.01 * C.r_h;
    else return r <= C.r_h + 1.0e-3;
}

template <int MODE>
__device__ __forceinline__ bool batch_inside_stop_rh(double r_h, double r)
{
    if constexpr (MODE == HADROS_MODE_CFG) return r < 1.01 * r_h;
    else return r <= r_h + 1.0e-3;
}

// global output index of queue ray `ray` of camera `cam` (recomputed when the
// ray finishes rather than held in registers while it is traced)
__device__ __forceinline__ long batch_out_index(const DevCam* __restrict__ cams, const long* __restrict__ pix0, int cam,
                                                long ray)
{
    const DevCam& C = cams[cam];
    int p, q;
    batch_pixel(C, ray - pix0[cam], p, q);
    return pix0[cam] + long(p) * C.sub_ny + q;
}

template <int MODE>
__device__ __forceinline__ void batch_init_ray(const DevCam* __restrict__ cams, int n_cams,
                                               const long* __restrict__ pix0, BatchLane& L, long ray)
{
    const int lo = batch_camera(pix0, n_cams, ray);
    L.ray = ray;
    L.cam = lo;
    const DevCam& C = cams[lo];
    int p, q;   // lattice pixel of the camera's k-th ray in queue order
    batch_pixel(C, ray - pix0[lo], p, q);
    const int i = C.sub_offset + C.sub_stride * p;
    const int j = C.sub_offset + C.sub_stride * q;
    // src/kerr_camera.cpp:110-186 with the observer metric from the host
    const double u = (2.0 * (i + 0.5) / C.nx - 1.0) * C.tanhalf;
    const double v = (2.0 * (j + 0.5) / C.ny - 1.0) * C.tanhalf;
    const double norm = sqrt(1.0 + u * u + v * v);
    const double n_r = -1.0 / norm, n_th = v / norm, n_ph = u / norm;
    const double pt_c = 1.0 / C.alpha;
    const double pr_c = n_r / sqrt(C.g_rr);
    const double pth_c = n_th / sqrt(C.g_thth);
    const double pph_c = n_ph / sqrt(C.g_phph) + C.omega * pt_c;
    L.pt = C.g_tt * pt_c + C.g_tph * pph_c;
    L.pph = C.g_tph * pt_c + C.g_phph * pph_c;
    L.y = Y5{C.r_obs, C.th_obs, 0.0, C.g_rr * pr_c, C.g_thth * pth_c};
    L.yp = L.y;
    L.inv_e_obs = 1.0 / dmax(-(L.pt + C.omega * L.pph) / C.alpha, 1.0e-300);
    L.h = C.h;
    L.step = 0;
    L.npts = 0;
    L.captured = 0;
    L.exhausted = 0;
    L.att_in_step = 0;
    L.attempts = 0;
}

template <int MODE, int PREC, int NBMAX>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(HD_BATCH_WAVES_PER_EU)))
k_trace_batch(BatchArgs A)
{
    extern __shared__ double smem[];
    const int n = A.n_sigma;
    const double* lnE = smem;
    const double* lnS = smem + n;
    const double* slope = smem + 2 * n;
    const float* lnEf = reinterpret_cast<const float*>(smem);
    const float* lnSf = lnEf + n;
    const float* slopef = lnEf + 2 * n;
    // camera start table: in LDS for up to BATCH_LDS_CAMS cameras, else read
    // through the cache hierarchy (one binary search per ray refill)
    long* pix0_lds = reinterpret_cast<long*>(smem + 3 * n);   // [n_cams + 1]
    const bool pix_in_lds = A.n_cams <= BATCH_LDS_CAMS;
    const long* pix0 = pix_in_lds ? pix0_lds : A.pix0;
    if constexpr (PREC == HADROS_PREC_FP64) {
        for (int t = threadIdx.x; t < 3 * n; t += blockDim.x) smem[t] = A.sig[t];
    } else {
        float* sf = reinterpret_cast<float*>(smem);
        for (int t = threadIdx.x; t < 3 * n; t += blockDim.x) sf[t] = A.sigf[t];
    }
    if (pix_in_lds)
        for (int t = threadIdx.x; t <= A.n_cams; t += blockDim.x) pix0_lds[t] = A.pix0[t];
    __syncthreads();

    const int lane = threadIdx.x & 63;
    BatchLane L;
    L.ray = -1;
    double tau[NBMAX], I[NBMAX];
    Kahan ktau[NBMAX];
    ScaledKahan kI[NBMAX];
    bool drained = false;   // wave-uniform: the queue is empty
    bool direct = false;    // redo the lane's attempt in rkf45_cfg's form (rkf45_try)
    unsigned long long c_pts = 0, c_att = 0, c_cap = 0, c_ex = 0;
    int since_refill = A.steps_per_chunk;

    while (true) {
        // ---- refill finished lanes from the global queue (wave-uniform control)
        if (!drained && since_refill >= A.steps_per_chunk) {
            since_refill = 0;
            const unsigned long long idle = __ballot(L.ray < 0);
            if (idle) {
                const int cnt = __popcll(idle);
                const int leader = __ffsll(static_cast<long long>(idle)) - 1;
                unsigned long long base = 0;
                if (lane == leader) base = atomicAdd(A.queue, (unsigned long long)cnt);
                base = __shfl(base, leader, 64);
                if (base + cnt >= (unsigned long long)A.n_rays) drained = true;
                if (L.ray < 0) {
                    const unsigned long long below = idle & ((1ull << lane) - 1ull);
                    const long my = long(base) + __popcll(below);
                    if (my < A.n_rays) {
                        batch_init_ray<MODE>(A.cams, A.n_cams, pix0, L, my);
#pragma unroll
                        for (int b = 0; b < NBMAX; ++b) {
                            tau[b] = 0.0;
                            I[b] = 0.0;
                            ktau[b] = {0.0f, 0.0f};
                            kI[b] = {{0.0f, 0.0f}, 0.0f, 0};
                        }
                        const bool cap0 = batch_inside_stop<MODE>(A.cams[L.cam], L.y.r);
                        if (cap0) L.captured = 1;
                        else L.npts = 1;
                    }
                }
            }
        }
        if (__ballot(L.ray >= 0) == 0ull) {
            if (drained) break;
            since_refill = A.steps_per_chunk;
            continue;
        }
        ++since_refill;

        bool finished = false;
        if (L.ray >= 0 && !(L.captured && L.step == 0)) {
            const DevCam& C = A.cams[L.cam];
            const double a = C.a;
            Y5 y4, y5;
            // a discarded angle-addition attempt is redone on the next pass (rkf45_try)
            if (rkf45_try<MODE>(a, L.y, L.pt, L.pph, L.h, direct, y4, y5)) {
                ++L.attempts;
                bool acc;
                if constexpr (MODE == HADROS_MODE_REF) {
                    const double h_min = C.h * 1.0e-5;
                    const double err = err_ref(y4, y5);
                    acc = (err < C.tol) || (L.h <= h_min);
                    if (!acc) {
                        const double factor = 0.8 * sqrt(sqrt(C.tol / dmax(err, 1.0e-30)));
                        L.h *= dclamp(factor, 0.1, 5.0);
                        if (L.h < h_min) L.h = h_min;
                        if (++L.att_in_step == 50) {
                            y5 = rk4_ref(a, L.y, L.pt, L.pph, C.h);
                            acc = true;
                        }
                    }
                    if (acc) {
                        L.h = C.h;
                        L.att_in_step = 0;
                    }
                } else {
                    const double err = err_cfg(L.y, y4, y5, C.tol);
                    acc = (err <= 1.0) | (L.h <= CFG_H_MIN);
                    L.h = fmax(L.h * cfg_step_factor(err, acc), CFG_H_MIN);
                    if (!acc && L.attempts >= 8L * C.max_steps + 64) {
                        L.exhausted = 1;
                        finished = true;
                    }
                }
                if (acc) {
                    L.yp = L.y;
                    L.y = y5;
                    ++L.step;
                    if (L.step >= C.max_steps) {
                        L.exhausted = 1;
                        finished = true;
                    } else if (batch_inside_stop<MODE>(A.cams[L.cam], L.y.r)) {
                        L.captured = 1;
                        finished = true;
                    } else if (((MODE == HADROS_MODE_CFG) ? (L.y.r > C.r_max) : (L.y.r >= C.r_max)) && L.step > 10) {
                        finished = true;
                    } else {
                        ++L.npts;
                        const double dl = zamo_dl(a, L.y, L.yp);
                        const double z = zamo_energy(a, L.y.r, L.y.th, L.pt, L.pph) * L.inv_e_obs;
                        const DevRt& P = A.rts[C.pad_rt];
                        PointQ q;
                        if (point_q(P, L.y.r, L.y.th, dl, z, q)) {
                            if constexpr (PREC == HADROS_PREC_FP64) {
                                point_rt64<NBMAX>(P, q, lnE, lnS, slope, tau, I);
                            } else {
                                point_rt32<NBMAX>(P, q, lnEf, lnSf, slopef, n, ktau, kI);
                            }
                        }
                    }
                }
            }
        } else if (L.ray >= 0) {
            finished = true;   // captured at the camera: no points
        }

        if (finished) {
            // ---- outputs of ray L.ray (aggregation of src/radiative_transfer.cpp:711-769)
            const DevRt& P = A.rts[A.cams[L.cam].pad_rt];
            const long k = batch_out_index(A.cams, pix0, L.cam, L.ray);
            double bt[NBMAX], bp[NBMAX], bi[NBMAX];
#pragma unroll
            for (int b = 0; b < NBMAX; ++b) {
                if (PREC == HADROS_PREC_FP64) {
                    bt[b] = tau[b];
                    bi[b] = I[b];
                } else {
                    bt[b] = double(ktau[b].s) - double(ktau[b].c);
                    bi[b] = kI[b].set ? (double(kI[b].k.s) - double(kI[b].k.c)) * exp(double(kI[b].off)) : 0.0;
                }
                bp[b] = exp(-bt[b]);
            }
            double agg_t = 0.0, agg_p = 1.0, agg_i = 0.0;
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
            const TraceOut& O = A.out;
            if (O.tau) O.tau[k] = agg_t;
            if (O.p) O.p[k] = agg_p;
            if (O.I) O.I[k] = agg_i;
            if (O.captured) O.captured[k] = L.captured;
            if (O.npoints) O.npoints[k] = L.npts;
            if (O.flags) O.flags[k] = L.exhausted ? HADROS_FLAG_EXHAUSTED : 0;
#pragma unroll
            for (int b = 0; b < NBMAX; ++b) {
                if (b >= A.nb_out) continue;
                const bool on = b < P.nb;
                if (O.btau) O.btau[k * A.nb_out + b] = on ? bt[b] : 0.0;
                if (O.bp) O.bp[k * A.nb_out + b] = on ? bp[b] : 1.0;
                if (O.bi) O.bi[k * A.nb_out + b] = on ? bi[b] : 0.0;
            }
            c_pts += (unsigned long long)L.npts;
            c_att += (unsigned long long)L.attempts;
            c_cap += (unsigned long long)L.captured;
            c_ex += (unsigned long long)L.exhausted;
            L.ray = -1;
        }
    }
    wave_add4(A.out.counters, c_pts, c_att, c_cap, c_ex);
}

#ifndef HD_BATCH_POINTS_WAVES_PER_EU
#define HD_BATCH_POINTS_WAVES_PER_EU 3
#endif

// Geodesic half of the split batch (the transfer runs afterwards in k_replay
// over the chunk pool): the same queue, refill and RKF45 stepping as
// k_trace_batch, with each emitted point's (r, theta, dl, redshift) stored in
// the ray's chunks instead of integrated.  Works on the sub-batch of global
// rays [ray0, ray0 + n_rays); the queue hands out local indices.
template <int MODE>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(MODE == HADROS_MODE_CFG ? HD_BATCH_POINTS_WAVES_PER_EU : 2)))
k_batch_points(BatchPointsArgs A)
{
    extern __shared__ double smem[];
    long* pix0_lds = reinterpret_cast<long*>(smem);   // [n_cams + 1]
    const bool pix_in_lds = A.n_cams <= BATCH_LDS_CAMS;
    const long* pix0 = pix_in_lds ? pix0_lds : A.pix0;
    if (pix_in_lds)
        for (int t = threadIdx.x; t <= A.n_cams; t += blockDim.x) pix0_lds[t] = A.pix0[t];
    __syncthreads();

    const int lane = threadIdx.x & 63;
    BatchLane L;
    L.ray = -1;
    ChunkWriter w;
    w.begin(-1);
    bool drained = false;   // wave-uniform: the queue is empty
    bool direct = false;    // redo the lane's attempt in rkf45_cfg's form (rkf45_try)
    // sin and cos of the lane's L.y.th (CFG mode), kept from the ray's start or
    // its last accepted point for the next attempt's first stage, as the plain
    // split tracer keeps them: the same bits as computing them per attempt
    double sy = 0.0, cy = 1.0;
    // per-thread points / attempts / captured / exhausted totals, in LDS: they
    // change once per ray, and in registers they would cost the 8 VGPRs that
    // keep the kernel at 4 wavefronts per SIMD
    __shared__ unsigned long long cnt[256][4];
    // each lane's redshift factor of its ray (set at the ray's start, read at
    // every path point): in LDS, the registers it held spilled to scratch
    __shared__ double inv_e_lds[256];
    unsigned long long* my_cnt = cnt[threadIdx.x];
    my_cnt[0] = my_cnt[1] = my_cnt[2] = my_cnt[3] = 0ull;
    int since_refill = A.steps_per_chunk;

    while (true) {
        // ---- refill finished lanes from the queue (wave-uniform control)
        if (!drained && since_refill >= A.steps_per_chunk) {
            since_refill = 0;
            const unsigned long long idle = __ballot(L.ray < 0);
            if (idle) {
                const int cnt = __popcll(idle);
                const int leader = __ffsll(static_cast<long long>(idle)) - 1;
                unsigned long long base = 0;
                if (lane == leader) base = atomicAdd(A.queue, (unsigned long long)cnt);
                base = __shfl(base, leader, 64);
                if (base + cnt >= (unsigned long long)A.n_rays) drained = true;
                if (L.ray < 0) {
                    const unsigned long long below = idle & ((1ull << lane) - 1ull);
                    const long my = long(base) + __popcll(below);
                    if (my < A.n_rays) {
                        batch_init_ray<MODE>(A.cams, A.n_cams, pix0, L, A.ray0 + my);
                        if constexpr (MODE == HADROS_MODE_CFG) sincos_m(L.y.th, &sy, &cy);
                        inv_e_lds[threadIdx.x] = L.inv_e_obs;
                        w.begin(my);
                        const bool cap0 = batch_inside_stop<MODE>(A.cams[L.cam], L.y.r);
                        if (cap0) L.captured = 1;
                        else L.npts = 1;
                    }
                }
            }
        }
        if (__ballot(L.ray >= 0) == 0ull) {
            if (drained) break;
            since_refill = A.steps_per_chunk;
            continue;
        }
        ++since_refill;

        bool finished = false;
        if (L.ray >= 0 && !(L.captured && L.step == 0)) {
            const DevCam& C = A.cams[L.cam];
            const double a = C.a;
            Y5 y4, y5;
            // a discarded angle-addition attempt is redone on the next pass (rkf45_try)
            if (rkf45_try_sc<MODE>(a, L.y, sy, cy, L.pt, L.pph, L.h, direct, y4, y5)) {
                // the camera's step-control constants, read together after the
                // attempt (their loads share one wait instead of one each)
                const double tol = C.tol, r_max = C.r_max, r_h = C.r_h;
                const int max_steps = C.max_steps;
                ++L.attempts;
                bool acc;
                if constexpr (MODE == HADROS_MODE_REF) {
                    const double h_min = C.h * 1.0e-5;
                    const double err = err_ref(y4, y5);
                    acc = (err < tol) || (L.h <= h_min);
                    if (!acc) {
                        const double factor = 0.8 * sqrt(sqrt(tol / dmax(err, 1.0e-30)));
                        L.h *= dclamp(factor, 0.1, 5.0);
                        if (L.h < h_min) L.h = h_min;
                        if (++L.att_in_step == 50) {
                            y5 = rk4_ref(a, L.y, L.pt, L.pph, C.h);
                            acc = true;
                        }
                    }
                    if (acc) {
                        L.h = C.h;
                        L.att_in_step = 0;
                    }
                } else {
                    const double err = err_cfg(L.y, y4, y5, tol);
                    acc = (err <= 1.0) | (L.h <= CFG_H_MIN);
                    L.h = fmax(L.h * cfg_step_factor(err, acc), CFG_H_MIN);
                    if (!acc && L.attempts >= 8L * max_steps + 64) {
                        L.exhausted = 1;
                        finished = true;
                    }
                }
                if (acc) {
                    L.yp = L.y;
                    L.y = y5;
                    ++L.step;
                    if (L.step >= max_steps) {
                        L.exhausted = 1;
                        finished = true;
                    } else if (batch_inside_stop_rh<MODE>(r_h, L.y.r)) {
                        L.captured = 1;
                        finished = true;
                    } else if (((MODE == HADROS_MODE_CFG) ? (L.y.r > r_max) : (L.y.r >= r_max)) && L.step > 10) {
                        finished = true;
                    } else {
                        // ---- path point L.step (src/kerr_camera.cpp:236-273)
                        ++L.npts;
                        // the plain split tracer's point forms (sin, cos of theta
                        // by sincos_m, the midpoint's from them), so that the two
                        // round their points alike
                        sincos_m(L.y.th, &sy, &cy);
                        const double dl = (MODE == HADROS_MODE_CFG) ? zamo_dl_sc(a, L.y, L.yp, sy, cy)
                                                                    : zamo_dl(a, L.y, L.yp);
                        w.put(A.cp, L.y.r, L.y.th, dl,
                              zamo_energy_sc(a, L.y.r, sy, cy, L.pt, L.pph) * inv_e_lds[threadIdx.x]);
                    }
                }
            }
        } else if (L.ray >= 0) {
            finished = true;   // captured at the camera: no points
        }

        if (finished) {
            const long g = batch_out_index(A.cams, pix0, L.cam, L.ray);
            w.finish(A.cp, L.npts - 1);   // every point after the observer's was put()
            A.n_st[L.ray - A.ray0] = w.stored;
            const TraceOut& O = A.out;
            if (O.captured) O.captured[g] = L.captured;
            if (O.npoints) O.npoints[g] = L.npts;
            if (O.flags) O.flags[g] = L.exhausted ? HADROS_FLAG_EXHAUSTED : 0;
            my_cnt[0] += (unsigned long long)L.npts;
            my_cnt[1] += (unsigned long long)L.attempts;
            my_cnt[2] += (unsigned long long)L.captured;
            my_cnt[3] += (unsigned long long)L.exhausted;
            L.ray = -1;
        }
    }
    wave_add4(A.out.counters, my_cnt[0], my_cnt[1], my_cnt[2], my_cnt[3]);
}

static size_t batch_points_lds_bytes(int n_cams)
{
    return n_cams <= BATCH_LDS_CAMS ? size_t(n_cams + 1) * sizeof(long) : 0;
}

hipError_t batch_points_resident_blocks(int mode, int block, int n_cams, int* blocks_per_cu)
{
    const void* f = mode == HADROS_MODE_CFG ? reinterpret_cast<const void*>(&k_batch_points<HADROS_MODE_CFG>)
                                            : reinterpret_cast<const void*>(&k_batch_points<HADROS_MODE_REF>);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, f, block, batch_points_lds_bytes(n_cams));
}

hipError_t launch_batch_points(const BatchPointsArgs& A, int mode, int grid, int block, hipStream_t s)
{
    const size_t lds = batch_points_lds_bytes(A.n_cams);
    if (mode == HADROS_MODE_CFG) hipLaunchKernelGGL(k_batch_points<HADROS_MODE_CFG>, dim3(grid), dim3(block), lds, s, A);
    else hipLaunchKernelGGL(k_batch_points<HADROS_MODE_REF>, dim3(grid), dim3(block), lds, s, A);
    return hipGetLastError();
}

template <int MODE, int PREC>
static void launch_b(const BatchArgs& A, int nb, int grid, int block, size_t lds, hipStream_t s)
{
    if (nb <= 1) hipLaunchKernelGGL((k_trace_batch<MODE, PREC, 1>), dim3(grid), dim3(block), lds, s, A);
    else if (nb <= 8) hipLaunchKernelGGL((k_trace_batch<MODE, PREC, 8>), dim3(grid), dim3(block), lds, s, A);
    else hipLaunchKernelGGL((k_trace_batch<MODE, PREC, 16>), dim3(grid), dim3(block), lds, s, A);
}

size_t batch_lds_bytes(int n_sigma, int n_cams, int precision)
{
    (void)precision;   // the FP32 table is stored in the first half of the FP64 one's space
    return size_t(3) * n_sigma * sizeof(double) + (n_cams <= BATCH_LDS_CAMS ? size_t(n_cams + 1) * sizeof(long) : 0);
}

hipError_t batch_resident_blocks(int mode, int precision, int nb, int block, int n_sigma, int n_cams, int* blocks_per_cu)
{
    const size_t lds = batch_lds_bytes(n_sigma, n_cams, precision);
    const void* f = nullptr;
#define HD_PICK(M, P, N) f = reinterpret_cast<const void*>(&k_trace_batch<M, P, N>)
    if (mode == HADROS_MODE_CFG) {
        if (precision == HADROS_PREC_FP64) {
            if (nb <= 1) HD_PICK(HADROS_MODE_CFG, HADROS_PREC_FP64, 1);
            else if (nb <= 8) HD_PICK(HADROS_MODE_CFG, HADROS_PREC_FP64, 8);
            else HD_PICK(HADROS_MODE_CFG, HADROS_PREC_FP64, 16);
        } else {
            if (nb <= 1) HD_PICK(HADROS_MODE_CFG, HADROS_PREC_FP32K, 1);
            else if (nb <= 8) HD_PICK(HADROS_MODE_CFG, HADROS_PREC_FP32K, 8);
            else HD_PICK(HADROS_MODE_CFG, HADROS_PREC_FP32K, 16);
        }
    } else {
        if (precision == HADROS_PREC_FP64) {
            if (nb <= 1) HD_PICK(HADROS_MODE_REF, HADROS_PREC_FP64, 1);
            else if (nb <= 8) HD_PICK(HADROS_MODE_REF, HADROS_PREC_FP64, 8);
            else HD_PICK(HADROS_MODE_REF, HADROS_PREC_FP64, 16);
        } else {
            if (nb <= 1) HD_PICK(HADROS_MODE_REF, HADROS_PREC_FP32K, 1);
            else if (nb <= 8) HD_PICK(HADROS_MODE_REF, HADROS_PREC_FP32K, 8);
            else HD_PICK(HADROS_MODE_REF, HADROS_PREC_FP32K, 16);
        }
    }
#undef HD_PICK
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, f, block, lds);
}

hipError_t launch_trace_batch(const BatchArgs& A, int mode, int precision, int nb, int grid, int block, hipStream_t s)
{
    const size_t lds = batch_lds_bytes(A.n_sigma, A.n_cams, precision);
    if (mode == HADROS_MODE_CFG) {
        if (precision == HADROS_PREC_FP64) launch_b<HADROS_MODE_CFG, HADROS_PREC_FP64>(A, nb, grid, block, lds, s);
        else launch_b<HADROS_MODE_CFG, HADROS_PREC_FP32K>(A, nb, grid, block, lds, s);
    } else {
        if (precision == HADROS_PREC_FP64) launch_b<HADROS_MODE_REF, HADROS_PREC_FP64>(A, nb, grid, block, lds, s);
        else launch_b<HADROS_MODE_REF, HADROS_PREC_FP32K>(A, nb, grid, block, lds, s);
    }
    return hipGetLastError();
}

}  // namespace hd
