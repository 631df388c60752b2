// tracer.hip -- fused per-pixel kernel: Kerr geodesic + tau + per-bin transfer.
//
// One ray per lane, state in registers, sigma table in LDS, nothing
// materialised: the per-pixel loop of KerrCamera::trace_pixel
// (src/kerr_camera.cpp:188-283) with integrate_kerr_ray_spectral
// (src/radiative_transfer.cpp:457-780) folded into every emitted path point.
// A wavefront owns an 8x8 pixel tile so that its 64 rays have similar
// lengths (less tail divergence than a 64-pixel image row).
#include "kernels.h"

// Waves per SIMD the register allocator targets for the fused kernel (the
// FP64 state, six RKF45 stages and the per-bin accumulators compete for VGPRs).
#ifndef HD_TRACE_WAVES_PER_EU
#define HD_TRACE_WAVES_PER_EU 2
#endif

namespace hd {

template <int MODE, int PREC, int NBMAX>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(HD_TRACE_WAVES_PER_EU)))
k_trace_rt(TraceArgs A)
{
    extern __shared__ double smem[];
    const DevRt& P = A.rt;
    const int n = P.n_sigma;
    // ---- stage the sigma table in LDS
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

    const DevCam& C = A.cam;
    const int lane = threadIdx.x & 63;
    const long gw = (long(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const int tiles_y = (C.ny + 7) >> 3;
    const int ti = int(gw / tiles_y), tj = int(gw % tiles_y);
    const int i = ti * 8 + (lane >> 3), j = tj * 8 + (lane & 7);
    const bool valid = (i < C.nx) && (j < C.ny) && (ti * 8 < C.nx);
    const double a = C.a;

    // ---- initial state, src/kerr_camera.cpp:110-186 (observer metric from host)
    const double u = (2.0 * (i + 0.5) / C.nx - 1.0) * C.tanhalf;
    const double v = (2.0 * (j + 0.5) / C.ny - 1.0) * C.tanhalf;
    const double norm = sqrt(1.0 + u * u + v * v);
    const double n_r = -1.0 / norm, n_th = v / norm, n_ph = u / norm;
    const double pt_c = 1.0 / C.alpha;
    const double pr_c = n_r / sqrt(C.g_rr);
    const double pth_c = n_th / sqrt(C.g_thth);
    const double pph_c = n_ph / sqrt(C.g_phph) + C.omega * pt_c;
    const double pt = C.g_tt * pt_c + C.g_tph * pph_c;
    const double pph = C.g_tph * pt_c + C.g_phph * pph_c;
    Y5 y{C.r_obs, C.th_obs, 0.0, C.g_rr * pr_c, C.g_thth * pth_c};
    const double e_obs = -(pt + C.omega * pph) / C.alpha;
    const double inv_e_obs = 1.0 / dmax(e_obs, 1.0e-300);   // redshift factor of every point

    const double r_stop = (MODE == HADROS_MODE_CFG) ? 1.01 * C.r_h : C.r_h + 1.0e-3;

    double tau[NBMAX], I[NBMAX];
    Kahan ktau[NBMAX];
    ScaledKahan kI[NBMAX];
#pragma unroll
    for (int b = 0; b < NBMAX; ++b) {
        tau[b] = 0.0;
        I[b] = 0.0;
        ktau[b] = {0.0f, 0.0f};
        kI[b] = {{0.0f, 0.0f}, 0.0f, 0};
    }

    int step = 0, captured = 0, exhausted = 0, npts = 0;
    long attempts = 0;
    bool done = !valid;
    // step 0: termination checks, then the observer point (dl = 0 adds nothing)
    if (!done) {
        if ((MODE == HADROS_MODE_CFG) ? (y.r < r_stop) : (y.r <= r_stop)) {
            captured = 1;
            done = true;
        } else {
            npts = 1;
        }
    }
    Y5 yp = y;
    double h = C.h;
    const double h_min = C.h * 1.0e-5;
    int att_in_step = 0;
    const long max_attempts = 8L * C.max_steps + 64;

    bool direct = false;   // redo the attempt in rkf45_cfg's form (rkf45_try)
    while (!done) {
        Y5 y4, y5;
        if (!rkf45_try<MODE>(a, y, pt, pph, h, direct, y4, y5)) continue;
        ++attempts;
        bool acc;
        if constexpr (MODE == HADROS_MODE_REF) {
            // KerrGeodesic::step_adaptive (src/kerr_geodesic.cpp:235-328)
            const double err = err_ref(y4, y5);
            acc = (err < C.tol) || (h <= h_min);
            if (!acc) {
                const double factor = 0.8 * sqrt(sqrt(C.tol / dmax(err, 1.0e-30)));   // x^(1/4)
                h *= dclamp(factor, 0.1, 5.0);
                if (h < h_min) h = h_min;
                if (++att_in_step == 50) {
                    y5 = rk4_ref(a, y, pt, pph, C.h);
                    acc = true;
                }
            }
            if (acc) {
                h = C.h;
                att_in_step = 0;
            }
        } else {
            const double err = err_cfg(y, y4, y5, C.tol);
            acc = (err <= 1.0) | (h <= CFG_H_MIN);
            h = fmax(h * cfg_step_factor(err, acc), CFG_H_MIN);
            if (!acc && attempts >= max_attempts) {
                exhausted = 1;
                break;
            }
        }
        if (!acc) continue;
        yp = y;
        y = y5;
        ++step;
        if (step >= C.max_steps) {
            exhausted = 1;
            break;
        }
        if ((MODE == HADROS_MODE_CFG) ? (y.r < r_stop) : (y.r <= r_stop)) {
            captured = 1;
            break;
        }
        if ((MODE == HADROS_MODE_CFG) ? (y.r > C.r_max) : (y.r >= C.r_max)) {
            if (step > 10) break;
        }
        // ---- emit path point `step` (src/kerr_camera.cpp:236-273) and integrate it
        ++npts;
        const double dl = zamo_dl(a, y, yp);
        const double z = zamo_energy(a, y.r, y.th, pt, pph) * inv_e_obs;
        // The scene block is re-read through scalar loads at every point: an
        // empty asm on the pointer keeps the compiler from hoisting ~100 uniform
        // doubles out of the loop into vector registers.
        const DevRt* Pp = A.rtg;
        asm volatile("" : "+s"(Pp));
        const DevRt& Pq = *Pp;
        PointQ q;
        if (point_q(Pq, y.r, y.th, dl, z, q)) {
            if constexpr (PREC == HADROS_PREC_FP64) {
                point_rt64<NBMAX>(Pq, q, lnE, lnS, slope, tau, I);
            } else {
                point_rt32<NBMAX>(Pq, q, lnEf, lnSf, slopef, n, ktau, kI);
            }
        }
    }

    // ---- outputs
    const unsigned long long c_pts = valid ? (unsigned long long)npts : 0ull;
    unsigned long long c_att = valid ? (unsigned long long)attempts : 0ull;
    unsigned long long c_cap = (valid && captured) ? 1ull : 0ull;
    unsigned long long c_ex = (valid && exhausted) ? 1ull : 0ull;
    if (valid) {
        const long k = long(i) * C.ny + j;
        // integrate_kerr_ray_spectral aggregation (src/radiative_transfer.cpp:711-769)
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
        if (O.captured) O.captured[k] = captured;
        if (O.npoints) O.npoints[k] = int(c_pts);
        if (O.flags) O.flags[k] = exhausted ? HADROS_FLAG_EXHAUSTED : 0;
        if (O.alpha) O.alpha[k] = u;
        if (O.beta) O.beta[k] = v;
#pragma unroll
        for (int b = 0; b < NBMAX; ++b) {
            if (b >= P.nb) continue;
            if (O.btau) O.btau[k * P.nb + b] = bt[b];
            if (O.bp) O.bp[k * P.nb + b] = bp[b];
            if (O.bi) O.bi[k * P.nb + b] = bi[b];
        }
    }
    wave_add4(A.out.counters, c_pts, c_att, c_cap, c_ex);
}

// ---- geodesic cache mode: same trace, PathPoints written to [offset[k], ...)
template <int MODE, bool WRITE>
__global__ void __launch_bounds__(256) k_trace_cache(CacheArgs A)
{
    const DevCam& C = A.cam;
    const long kk = long(blockIdx.x) * blockDim.x + threadIdx.x;   // index within the range
    const long k = A.k0 + kk;
    if (k >= A.k1) return;
    const int i = int(k / C.ny), j = int(k % C.ny);
    const double a = C.a;
    const double u = (2.0 * (i + 0.5) / C.nx - 1.0) * C.tanhalf;
    const double v = (2.0 * (j + 0.5) / C.ny - 1.0) * C.tanhalf;
    const double norm = sqrt(1.0 + u * u + v * v);
    const double n_r = -1.0 / norm, n_th = v / norm, n_ph = u / norm;
    const double pt_c = 1.0 / C.alpha;
    const double pr_c = n_r / sqrt(C.g_rr);
    const double pth_c = n_th / sqrt(C.g_thth);
    const double pph_c = n_ph / sqrt(C.g_phph) + C.omega * pt_c;
    const double pt = C.g_tt * pt_c + C.g_tph * pph_c;
    const double pph = C.g_tph * pt_c + C.g_phph * pph_c;
    Y5 y{C.r_obs, C.th_obs, 0.0, C.g_rr * pr_c, C.g_thth * pth_c};
    const double inv_e_obs = 1.0 / dmax(-(pt + C.omega * pph) / C.alpha, 1.0e-300);   // redshift factor of every point
    const double r_stop = (MODE == HADROS_MODE_CFG) ? 1.01 * C.r_h : C.r_h + 1.0e-3;
    double* out = nullptr;
    long cap = 0;
    if constexpr (WRITE) {
        out = A.points + 7 * A.offset[kk];
        cap = A.offset[kk + 1] - A.offset[kk];
    }
    auto emit = [&](long idx, const Y5& yy, double dl, double z) {
        if constexpr (WRITE) {
            if (idx < cap) {
                double st, ct, sp, cp;
                sincos_m(yy.th, &st, &ct);
                sincos_m(yy.ph, &sp, &cp);
                double* o = out + 7 * idx;
                o[0] = yy.r;
                o[1] = yy.th;
                o[2] = yy.r * st * cp;
                o[3] = yy.r * st * sp;
                o[4] = yy.r * ct;
                o[5] = dl;
                o[6] = z;
            }
        }
    };
    int step = 0, captured = 0;
    long npts = 0;
    bool done = false;
    if ((MODE == HADROS_MODE_CFG) ? (y.r < r_stop) : (y.r <= r_stop)) {
        captured = 1;
        done = true;
    } else {
        emit(npts++, y, 0.0, zamo_energy(a, y.r, y.th, pt, pph) * inv_e_obs);
    }
    Y5 yp = y;
    double h = C.h;
    const double h_min = C.h * 1.0e-5;
    int att_in_step = 0;
    long attempts = 0;
    const long max_attempts = 8L * C.max_steps + 64;
    bool direct = false;   // redo the attempt in rkf45_cfg's form (rkf45_try)
    while (!done) {
        Y5 y4, y5;
        if (!rkf45_try<MODE>(a, y, pt, pph, h, direct, y4, y5)) continue;
        ++attempts;
        bool acc;
        if constexpr (MODE == HADROS_MODE_REF) {
            const double err = err_ref(y4, y5);
            acc = (err < C.tol) || (h <= h_min);
            if (!acc) {
                const double factor = 0.8 * sqrt(sqrt(C.tol / dmax(err, 1.0e-30)));   // x^(1/4)
                h *= dclamp(factor, 0.1, 5.0);
                if (h < h_min) h = h_min;
                if (++att_in_step == 50) {
                    y5 = rk4_ref(a, y, pt, pph, C.h);
                    acc = true;
                }
            }
            if (acc) {
                h = C.h;
                att_in_step = 0;
            }
        } else {
            // the controller in its two-branch form: built here with
            // cfg_step_factor, this kernel's points moved by up to 5e-8 relative
            // from those of k_trace_points, k_trace_rt and the batch kernels
            // on the same camera, though the arithmetic is the same; in this
            // form they agree bit for bit
            const double err = err_cfg(y, y4, y5, C.tol);
            acc = (err <= 1.0) || (h <= CFG_H_MIN);
            if (acc) {
                double fac = (err > 0.0) ? 0.9 * inv_fifth_root(err) : 5.0;
                fac = fmin(5.0, fmax(0.2, fac));
                h = fmax(h * fac, CFG_H_MIN);
            } else {
                double fac = (err == err) ? 0.9 * inv_fifth_root(err) : 0.1;
                fac = fmin(0.9, fmax(0.1, fac));
                h = fmax(h * fac, CFG_H_MIN);
                if (attempts >= max_attempts) break;
            }
        }
        if (!acc) continue;
        yp = y;
        y = y5;
        ++step;
        if (step >= C.max_steps) break;
        if ((MODE == HADROS_MODE_CFG) ? (y.r < r_stop) : (y.r <= r_stop)) {
            captured = 1;
            break;
        }
        if ((MODE == HADROS_MODE_CFG) ? (y.r > C.r_max) : (y.r >= C.r_max)) {
            if (step > 10) break;
        }
        emit(npts++, y, zamo_dl(a, y, yp), zamo_energy(a, y.r, y.th, pt, pph) * inv_e_obs);
    }
    if constexpr (!WRITE) {
        A.counts[kk] = npts;
        A.captured[kk] = captured;
    }
}

__global__ void k_zero_u64(unsigned long long* p, int n)
{
    for (int t = threadIdx.x; t < n; t += blockDim.x) p[t] = 0ull;
}

// ---- split form, geodesic half: the same trace as k_trace_rt, with every
// emitted point's (r, theta, dl, redshift) stored in the chunk pool for
// k_replay instead of being integrated here.  Without the per-point transfer
// in the loop the kernel fits 165 VGPRs instead of k_trace_rt's 256 (with
// spills), so 3 wavefronts per SIMD stay resident instead of 2 (the
// angle-addition attempt of rkf45_try costs the registers of a fourth, and
// measured 5 % faster).  Pixel k's first chunk is chunk k;
// further chunks come from a global counter (beyond nx * ny) and are linked
// through next[].  A ray whose chunk lies beyond the pool stops storing
// (the host sees the counter, grows the pool and runs again).
#ifndef HD_POINTS_WAVES_PER_EU
#define HD_POINTS_WAVES_PER_EU 3
#endif

template <int MODE>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(MODE == HADROS_MODE_CFG ? HD_POINTS_WAVES_PER_EU : 2)))
k_trace_points(PointsArgs A)
{
    const DevCam& C = A.cam;
    const int lane = threadIdx.x & 63;
    const long gw = A.tile0 + ((long(blockIdx.x) * blockDim.x + threadIdx.x) >> 6);
    const int tiles_y = (C.ny + 7) >> 3;
    const int ti = int(gw / tiles_y), tj = int(gw % tiles_y);
    const int i = ti * 8 + (lane >> 3), j = tj * 8 + (lane & 7);
    const bool valid = (gw < A.tile1) && (i < C.nx) && (j < C.ny) && (ti * 8 < C.nx);
    const double a = C.a;
    const double u = (2.0 * (i + 0.5) / C.nx - 1.0) * C.tanhalf;
    const double v = (2.0 * (j + 0.5) / C.ny - 1.0) * C.tanhalf;
    const double norm = sqrt(1.0 + u * u + v * v);
    const double n_r = -1.0 / norm, n_th = v / norm, n_ph = u / norm;
    const double pt_c = 1.0 / C.alpha;
    const double pr_c = n_r / sqrt(C.g_rr);
    const double pth_c = n_th / sqrt(C.g_thth);
    const double pph_c = n_ph / sqrt(C.g_phph) + C.omega * pt_c;
    const double pt = C.g_tt * pt_c + C.g_tph * pph_c;
    const double pph = C.g_tph * pt_c + C.g_phph * pph_c;
    Y5 y{C.r_obs, C.th_obs, 0.0, C.g_rr * pr_c, C.g_thth * pth_c};
    const double inv_e_obs = 1.0 / dmax(-(pt + C.omega * pph) / C.alpha, 1.0e-300);   // redshift factor of every point
    const double r_stop = (MODE == HADROS_MODE_CFG) ? 1.01 * C.r_h : C.r_h + 1.0e-3;
    const long k = long(i) * C.ny + j;
    // the camera-plane coordinates are stored now, so they hold no registers through the loop
    if (valid && A.out.alpha) A.out.alpha[k] = u;
    if (valid && A.out.beta) A.out.beta[k] = v;

    ChunkWriter w;
    w.begin(valid ? k : -1);
    int step = 0, captured = 0, exhausted = 0, npts = 0;
    long attempts = 0;
    bool done = !valid;
    // step 0: termination checks, then the observer point (dl = 0: nothing to store)
    if (!done) {
        if ((MODE == HADROS_MODE_CFG) ? (y.r < r_stop) : (y.r <= r_stop)) {
            captured = 1;
            done = true;
        } else {
            npts = 1;
        }
    }
    Y5 yp = y;
    double h = C.h;
    const double h_min = C.h * 1.0e-5;
    int att_in_step = 0;
    const long max_attempts = 8L * C.max_steps + 64;
    bool direct = false;   // redo the attempt in rkf45_cfg's form (rkf45_try)
    // sin and cos of y.th, kept from the last accepted point for the next
    // attempt's first stage (CFG mode; the same bits as computing them there)
    double sy = 0.0, cy = 1.0;
    if constexpr (MODE == HADROS_MODE_CFG) sincos_m(y.th, &sy, &cy);

    while (!done) {
        Y5 y4, y5;
        if (!rkf45_try_sc<MODE>(a, y, sy, cy, pt, pph, h, direct, y4, y5)) continue;
        ++attempts;
        bool acc;
        if constexpr (MODE == HADROS_MODE_REF) {
            const double err = err_ref(y4, y5);
            acc = (err < C.tol) || (h <= h_min);
            if (!acc) {
                const double factor = 0.8 * sqrt(sqrt(C.tol / dmax(err, 1.0e-30)));   // x^(1/4)
                h *= dclamp(factor, 0.1, 5.0);
                if (h < h_min) h = h_min;
                if (++att_in_step == 50) {
                    y5 = rk4_ref(a, y, pt, pph, C.h);
                    acc = true;
                }
            }
            if (acc) {
                h = C.h;
                att_in_step = 0;
            }
        } else {
            const double err = err_cfg(y, y4, y5, C.tol);
            acc = (err <= 1.0) | (h <= CFG_H_MIN);
            h = fmax(h * cfg_step_factor(err, acc), CFG_H_MIN);
            if (!acc && attempts >= max_attempts) {
                exhausted = 1;
                break;
            }
        }
        if (!acc) continue;
        yp = y;
        y = y5;
        ++step;
        if (step >= C.max_steps) {
            exhausted = 1;
            break;
        }
        if ((MODE == HADROS_MODE_CFG) ? (y.r < r_stop) : (y.r <= r_stop)) {
            captured = 1;
            break;
        }
        if ((MODE == HADROS_MODE_CFG) ? (y.r > C.r_max) : (y.r >= C.r_max)) {
            if (step > 10) break;
        }
        // ---- emit path point `step` (src/kerr_camera.cpp:236-273)
        ++npts;
        double z;
        if constexpr (MODE == HADROS_MODE_CFG) {
            sincos_m(y.th, &sy, &cy);
            z = zamo_energy_sc(a, y.r, sy, cy, pt, pph) * inv_e_obs;
        } else {
            z = zamo_energy(a, y.r, y.th, pt, pph) * inv_e_obs;
        }
        if constexpr (MODE == HADROS_MODE_CFG) w.put(A.cp, y.r, y.th, zamo_dl_sc(a, y, yp, sy, cy), z);
        else w.put(A.cp, y.r, y.th, zamo_dl(a, y, yp), z);
    }

    const unsigned long long c_pts = valid ? (unsigned long long)npts : 0ull;
    const unsigned long long c_att = valid ? (unsigned long long)attempts : 0ull;
    const unsigned long long c_cap = (valid && captured) ? 1ull : 0ull;
    const unsigned long long c_ex = (valid && exhausted) ? 1ull : 0ull;
    if (valid) {
        w.finish(A.cp, npts - 1);   // every point after the observer's was put()
        A.n_st[k] = w.stored;
        const TraceOut& O = A.out;
        if (O.captured) O.captured[k] = captured;
        if (O.npoints) O.npoints[k] = npts;
        if (O.flags) O.flags[k] = exhausted ? HADROS_FLAG_EXHAUSTED : 0;
    }
    wave_add4(A.out.counters, c_pts, c_att, c_cap, c_ex);
}

hipError_t launch_trace_points(const PointsArgs& A, int block, hipStream_t s)
{
    const int tiles = A.tile1 - A.tile0;
    if (tiles <= 0) return hipSuccess;
    const long threads = long(tiles) * 64;
    const int grid = int((threads + block - 1) / block);
    if (A.cam.mode == HADROS_MODE_CFG) hipLaunchKernelGGL(k_trace_points<HADROS_MODE_CFG>, dim3(grid), dim3(block), 0, s, A);
    else hipLaunchKernelGGL(k_trace_points<HADROS_MODE_REF>, dim3(grid), dim3(block), 0, s, A);
    return hipGetLastError();
}

hipError_t launch_zero_u64(unsigned long long* p, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, p, n);
    return hipGetLastError();
}

// ------------------------------------------------------------------ launchers
template <int MODE, int PREC>
static hipError_t launch_nb(const TraceArgs& A, int block, size_t lds, hipStream_t s, int grid)
{
    if (A.rt.nb <= 1) {
        hipLaunchKernelGGL((k_trace_rt<MODE, PREC, 1>), dim3(grid), dim3(block), lds, s, A);
    } else if (A.rt.nb <= 8) {
        hipLaunchKernelGGL((k_trace_rt<MODE, PREC, 8>), dim3(grid), dim3(block), lds, s, A);
    } else {
        hipLaunchKernelGGL((k_trace_rt<MODE, PREC, 16>), dim3(grid), dim3(block), lds, s, A);
    }
    return hipGetLastError();
}

hipError_t launch_trace_rt(const TraceArgs& A, int precision, int block, hipStream_t s)
{
    const int tiles = ((A.cam.nx + 7) / 8) * ((A.cam.ny + 7) / 8);
    const long threads = long(tiles) * 64;
    const int grid = int((threads + block - 1) / block);
    const size_t lds = size_t(3) * A.rt.n_sigma * (precision == HADROS_PREC_FP64 ? sizeof(double) : sizeof(float));
    if (A.cam.mode == HADROS_MODE_CFG) {
        return precision == HADROS_PREC_FP64 ? launch_nb<HADROS_MODE_CFG, HADROS_PREC_FP64>(A, block, lds, s, grid)
                                             : launch_nb<HADROS_MODE_CFG, HADROS_PREC_FP32K>(A, block, lds, s, grid);
    }
    return precision == HADROS_PREC_FP64 ? launch_nb<HADROS_MODE_REF, HADROS_PREC_FP64>(A, block, lds, s, grid)
                                         : launch_nb<HADROS_MODE_REF, HADROS_PREC_FP32K>(A, block, lds, s, grid);
}

hipError_t launch_trace_cache(const CacheArgs& A, bool write, hipStream_t s)
{
    const long n = A.k1 - A.k0;
    const int block = 64;
    const int grid = int((n + block - 1) / block);
    if (A.cam.mode == HADROS_MODE_CFG) {
        if (write) hipLaunchKernelGGL((k_trace_cache<HADROS_MODE_CFG, true>), dim3(grid), dim3(block), 0, s, A);
        else hipLaunchKernelGGL((k_trace_cache<HADROS_MODE_CFG, false>), dim3(grid), dim3(block), 0, s, A);
    } else {
        if (write) hipLaunchKernelGGL((k_trace_cache<HADROS_MODE_REF, true>), dim3(grid), dim3(block), 0, s, A);
        else hipLaunchKernelGGL((k_trace_cache<HADROS_MODE_REF, false>), dim3(grid), dim3(block), 0, s, A);
    }
    return hipGetLastError();
}

}  // namespace hd
