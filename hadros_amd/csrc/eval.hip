// eval.hip -- point-wise device evaluations of each physics term, so that
// every device function on the path can be checked against the oracle on its
// own (rho, sigma, emissivity, rhs, metric).
#include "kernels.h"

namespace hd {

__global__ void k_eval_rho(DevMedium m, const double* r, const double* th, double* out, long n)
{
    const long k = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) out[k] = rho(m, r[k], th[k]);
}

__global__ void k_eval_sigma(const double* sig, int ns, double e_lo, double e_hi, double lo0, double inv_d,
                             const double* e, double* out,
                             long n)
{
    const long k = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double x = e[k];
    if (!(x > 0.0) || x < e_lo || x > e_hi) {
        out[k] = __builtin_nan("");
        return;
    }
    out[k] = exp(sigma_ln(sig, sig + ns, sig + 2 * ns, ns, log(x), lo0, inv_d));
}

// emissivity_uhe = (norm * spatial) * S(E) (src/radiative_transfer.cpp:369-395)
__global__ void k_eval_emis(DevRt P, const double* r, const double* th, const double* e, double* out, long n)
{
    const long k = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double E = e[k];
    double j = 0.0;
    if (E > 0.0 && P.spec_ok) {
        const double A = source_factor(P.src, P.med, r[k], th[k], rho(P.med, r[k], th[k]));
        j = A * exp(-P.gam * log(E) - E * P.invcut);
    }
    out[k] = j;
}

// the MeV block of integrate_kerr_ray at one point with redshift factor g:
// T, Ye, kappa, j (src/radiative_transfer.cpp:500-600)
__global__ void k_eval_mev(DevMedium med, DevMev M, const double* r, const double* th, const double* g, double* out,
                           long n)
{
    const long k = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double rh = rho(med, r[k], th[k]);
    const double T = mev_temperature(M, r[k], th[k], rh);
    const double Ye = mev_ye(M, r[k], th[k], rh);
    double kappa = 0.0, j = 0.0;
    mev_kappa_j(M, g[k], rh, T, Ye, kappa, j);
    out[4 * k] = T;
    out[4 * k + 1] = Ye;
    out[4 * k + 2] = kappa;
    out[4 * k + 3] = j;
}

__global__ void k_eval_rhs(double a, int mode, const double* y, double* dy, long n)
{
    const long k = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double* s = y + 8 * k;
    const Y5 yy{s[1], s[2], s[3], s[5], s[6]};
    const Y5 d = (mode == HADROS_MODE_CFG) ? rhs_cfg(a, yy, s[4], s[7]) : rhs_ref(a, yy, s[4], s[7]);
    double* o = dy + 8 * k;
    // dt/dlambda = g^tt p_t + g^tphi p_phi
    double sn, cs;
    sincos_m(s[2], &sn, &cs);
    const Metric5 g = inv_metric(a, s[1], sn, cs);
    o[0] = g.tt * s[4] + g.tph * s[7];
    o[1] = d.r;
    o[2] = d.th;
    o[3] = d.ph;
    o[4] = 0.0;
    o[5] = d.pr;
    o[6] = d.pth;
    o[7] = 0.0;
}

__global__ void k_eval_metric(double a, const double* r, const double* th, double* out, long n)
{
    const long k = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const Metric5 g = metric(a, r[k], th[k]);
    double sn, cs;
    sincos_m(th[k], &sn, &cs);
    const Metric5 gi = inv_metric(a, r[k], sn, cs);
    double* o = out + 10 * k;
    o[0] = g.tt; o[1] = g.tph; o[2] = g.rr; o[3] = g.thth; o[4] = g.phph;
    o[5] = gi.tt; o[6] = gi.tph; o[7] = gi.rr; o[8] = gi.thth; o[9] = gi.phph;
}

// One integrator step of each form the kernels take, on states y (8 doubles
// each): out[11 k ..] = y4 (r, theta, phi, p_r, p_theta), y5 (same), ok.
//   EVAL_STEP_RK4        KerrGeodesic::step_rk4 (rk4_ref): y4 = y5 = result
//   EVAL_STEP_RKF45_REF  a REF-mode Fehlberg attempt (rkf45<REF>)
//   EVAL_STEP_RKF45_CFG  a CFG attempt with sincos_m at every stage (rkf45_cfg)
//   EVAL_STEP_RKF45_CFG_SHIFT  the CFG attempt with the later stages' sin/cos by
//                        angle addition (rkf45_cfg_shift); ok = 0 when a stage
//                        offset leaves sincos_shift's range (the kernels then
//                        redo the attempt with rkf45_cfg)
__global__ void k_eval_step(double a, int kind, double h, const double* y, double* out, long n)
{
    const long k = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double* s = y + 8 * k;
    const Y5 yy{s[1], s[2], s[3], s[5], s[6]};
    Y5 y4{}, y5{};
    bool ok = true;
    if (kind == EVAL_STEP_RK4) {
        y4 = y5 = rk4_ref(a, yy, s[4], s[7], h);
    } else if (kind == EVAL_STEP_RKF45_REF) {
        rkf45<HADROS_MODE_REF>(a, yy, s[4], s[7], h, y4, y5);
    } else if (kind == EVAL_STEP_RKF45_CFG) {
        rkf45_cfg(a, yy, s[4], s[7], h, y4, y5);
    } else {
        ok = rkf45_cfg_shift(a, yy, s[4], s[7], h, y4, y5);
    }
    double* o = out + 11 * k;
    o[0] = y4.r; o[1] = y4.th; o[2] = y4.ph; o[3] = y4.pr; o[4] = y4.pth;
    o[5] = y5.r; o[6] = y5.th; o[7] = y5.ph; o[8] = y5.pr; o[9] = y5.pth;
    o[10] = ok ? 1.0 : 0.0;
}

// The per-point and per-(point, bin) terms of k_replay (the transfer kernel of
// the split tracer, the cache replay and the batch) at points (r, theta, dl,
// z), through the device functions it runs: point_q with exp_tab and log_tab
// (density, source factor, redshift), the sigma lookup the scene's table selects
// (exact-uniform or bracketing) with exp_tab, or for a power-law table n_N dl
// z^s sigma(E_b) (FP32K: e^(ln(n_N dl z^s) + ln sigma(E_b))), and the emission
// exponent.  (The
// fused tracer's library-exp forms are the ones k_eval_rho, k_eval_sigma and
// k_eval_emis run.)  out[EVAL_TRANSFER_COLS k ..] = ok, z, ln z,
// g3Adl, nbar_dl, rho, then per bin dtau_b, then per bin the emission term
// before absorption: FP64 g3Adl e^arg_b, FP32K its log ln(g3Adl) + ln S(E_b)
// (-inf where the point does not emit).  Bins from nb on are 0.
template <int PREC>
__global__ void k_eval_transfer_point(DevRt P, const double* sig, const float* sigf, const double* r,
                                      const double* th, const double* dl, const double* z, double* out, long n)
{
    __shared__ double exp_t[64];
    __shared__ __align__(16) double log_t[2 * LOG_TAB_N];
    if (threadIdx.x < 64) exp_t[threadIdx.x] = EXP_TAB[threadIdx.x];
    if (threadIdx.x < 2 * LOG_TAB_N) log_t[threadIdx.x] = LOG_TAB[threadIdx.x];
    __syncthreads();
    const long k = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    PointQ pq;
    // the transfer kernel's forms: FP64 tables, float-accurate in the FP32 form
    bool ok;
    if constexpr (PREC == HADROS_PREC_FP32K) ok = point_q(P, r[k], th[k], dl[k], z[k], pq, ExpF32{});
    else ok = point_q(P, r[k], th[k], dl[k], z[k], pq, ExpTab{exp_t, log_t});
    if (!ok) {   // as the kernels: a skipped point adds nothing
        pq.z = 1.0;
        pq.lnz = 0.0;
        pq.g3Adl = 0.0;
        pq.nbar_dl = 0.0;
        pq.rho = 0.0;
    }
    double* o = out + long(EVAL_TRANSFER_COLS) * k;
    o[0] = ok ? 1.0 : 0.0;
    o[1] = pq.z;
    o[2] = pq.lnz;
    o[3] = pq.g3Adl;
    o[4] = pq.nbar_dl;
    o[5] = pq.rho;
    const int n_s = P.n_sigma;
    const bool exact = P.sig_exact != 0;
    for (int b = 0; b < HADROS_MAX_BINS; ++b) {
        double dt = 0.0, em = 0.0;
        if (b < P.nb) {
            const double E = P.emid[b] * pq.z;
            if constexpr (PREC == HADROS_PREC_FP64) {
                const double lnEb = P.lnemid[b] + pq.lnz;
                if (P.sig_pow)
                    dt = bin_dtau64_pow(ok && pq.rho > 0.0, E, pq.nbar_dl * exp_tab(P.sig_s * pq.lnz, exp_t),
                                        P.sigb[b], P.e_lo, P.e_hi);
                else
                    dt = bin_dtau64(exact, ok && pq.rho > 0.0, E, lnEb, pq.nbar_dl, sig, sig + n_s, sig + 2 * n_s,
                                    n_s, P.sig_lo, P.sig_inv_d, P.e_lo, P.e_hi, exp_t);
                if (ok && P.spec_ok && pq.g3Adl != 0.0)
                    em = pq.g3Adl * exp_tab(bin_emis_arg(P.gam, P.invcut, E, lnEb), exp_t);
            } else {
                const float lnEb = float(P.lnemid[b]) + float(pq.lnz);
                // the transfer kernel's staged logarithms (ln_f32, rounded to float)
                const float lN = (pq.nbar_dl > 0.0) ? float(ln_f32(pq.nbar_dl)) : 0.0f;
                const float lP = (pq.g3Adl > 0.0) ? float(ln_f32(pq.g3Adl)) : 0.0f;
                const bool ab = ok && pq.rho > 0.0 && pq.nbar_dl > 0.0;
                if (P.sig_pow) {
                    const double lnb = (pq.nbar_dl > 0.0) ? ln_f32(pq.nbar_dl) : 0.0;
                    const float lW = float(lnb + P.sig_s * pq.lnz);
                    dt = (ab && E >= P.e_lo && E <= P.e_hi) ? __expf(lW + logf(float(P.sigb[b]))) : 0.0f;
                } else {
                    dt = bin_dtau32(exact, ab, E, lnEb, lN, sigf, sigf + n_s, sigf + 2 * n_s, n_s, float(P.sig_lo),
                                    float(P.sig_inv_d), P.e_lo, P.e_hi);
                }
                em = (ok && P.spec_ok && pq.g3Adl > 0.0)
                         ? double(bin_emis_log32(lP, float(P.gam), float(P.invcut), E, lnEb))
                         : -HUGE_VAL;
            }
        }
        o[6 + b] = dt;
        o[6 + HADROS_MAX_BINS + b] = em;
    }
}

// Streaming kernels of known HBM traffic, to scale a profiler's byte counters:
// k_stream_read reads x[0, n) once (16-byte loads, grid-stride) and writes one
// partial sum per workgroup; k_stream_fill writes v to x[0, n).
constexpr int STREAM_BLOCK = 256;

__global__ void __launch_bounds__(STREAM_BLOCK) k_stream_read(const double* __restrict__ x, long n,
                                                              double* __restrict__ partial)
{
    const double2* x2 = reinterpret_cast<const double2*>(x);
    const long stride = long(gridDim.x) * STREAM_BLOCK;
    double s = 0.0;
    for (long i = long(blockIdx.x) * STREAM_BLOCK + threadIdx.x; i < n / 2; i += stride) {
        const double2 v = x2[i];
        s += v.x + v.y;
    }
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) s += x[n - 1];
    __shared__ double red[STREAM_BLOCK];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = STREAM_BLOCK / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(STREAM_BLOCK) k_stream_fill(double* __restrict__ x, long n, double v)
{
    double2* x2 = reinterpret_cast<double2*>(x);
    const long stride = long(gridDim.x) * STREAM_BLOCK;
    for (long i = long(blockIdx.x) * STREAM_BLOCK + threadIdx.x; i < n / 2; i += stride) x2[i] = double2{v, v};
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) x[n - 1] = v;
}

hipError_t launch_stream_probe(int kind, double* x, long n, double* partial, double v, int grid, hipStream_t st)
{
    if (kind == 0)
        hipLaunchKernelGGL(k_stream_read, dim3(grid), dim3(STREAM_BLOCK), 0, st, x, n, partial);
    else
        hipLaunchKernelGGL(k_stream_fill, dim3(grid), dim3(STREAM_BLOCK), 0, st, x, n, v);
    return hipGetLastError();
}

static inline int blocks(long n) { return int((n + 255) / 256); }

hipError_t launch_eval_transfer_point(const DevRt& P, int precision, const double* sig, const float* sigf,
                                      const double* r, const double* th, const double* dl, const double* z,
                                      double* out, long n)
{
    if (precision == HADROS_PREC_FP64)
        hipLaunchKernelGGL(k_eval_transfer_point<HADROS_PREC_FP64>, dim3(blocks(n)), dim3(256), 0, 0, P, sig, sigf,
                           r, th, dl, z, out, n);
    else
        hipLaunchKernelGGL(k_eval_transfer_point<HADROS_PREC_FP32K>, dim3(blocks(n)), dim3(256), 0, 0, P, sig, sigf,
                           r, th, dl, z, out, n);
    return hipGetLastError();
}

hipError_t launch_eval_step(double a, int kind, double h, const double* y, double* out, long n)
{
    hipLaunchKernelGGL(k_eval_step, dim3(blocks(n)), dim3(256), 0, 0, a, kind, h, y, out, n);
    return hipGetLastError();
}

hipError_t launch_eval_rho(const DevMedium& m, const double* r, const double* th, double* out, long n)
{
    hipLaunchKernelGGL(k_eval_rho, dim3(blocks(n)), dim3(256), 0, 0, m, r, th, out, n);
    return hipGetLastError();
}
hipError_t launch_eval_sigma(const double* sig, int ns, double e_lo, double e_hi, double lo0, double inv_d,
                             const double* e, double* out, long n)
{
    hipLaunchKernelGGL(k_eval_sigma, dim3(blocks(n)), dim3(256), 0, 0, sig, ns, e_lo, e_hi, lo0, inv_d, e, out, n);
    return hipGetLastError();
}
hipError_t launch_eval_emis(const DevRt& P, const double* r, const double* th, const double* e, double* out, long n)
{
    hipLaunchKernelGGL(k_eval_emis, dim3(blocks(n)), dim3(256), 0, 0, P, r, th, e, out, n);
    return hipGetLastError();
}
hipError_t launch_eval_mev(const DevMedium& med, const DevMev& M, const double* r, const double* th, const double* g,
                           double* out, long n)
{
    hipLaunchKernelGGL(k_eval_mev, dim3(blocks(n)), dim3(256), 0, 0, med, M, r, th, g, out, n);
    return hipGetLastError();
}
hipError_t launch_eval_rhs(double a, int mode, const double* y, double* dy, long n)
{
    hipLaunchKernelGGL(k_eval_rhs, dim3(blocks(n)), dim3(256), 0, 0, a, mode, y, dy, n);
    return hipGetLastError();
}
hipError_t launch_eval_metric(double a, const double* r, const double* th, double* out, long n)
{
    hipLaunchKernelGGL(k_eval_metric, dim3(blocks(n)), dim3(256), 0, 0, a, r, th, out, n);
    return hipGetLastError();
}

}  // namespace hd
