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
// broadcasts).  HBM traffic is the algorithmic 32 B per point.
//
// Two point sources: a cache's contiguous SoA ranges (image-from-cache), or
// the chunk pool written by the split tracer (k_trace_points): 64-point
// chunks of 32-byte (r, theta, dl, z) records, the first at the ray's own
// index, the rest linked through next[].
#include <cstddef>
#include <type_traits>

#include "fastmath.h"
#include "kernels.h"
#include "wave.h"

namespace hd {

#ifndef HD_REPLAY_WAVES_PER_EU
#define HD_REPLAY_WAVES_PER_EU 3
#endif

constexpr int REPLAY_BLOCK = 256;

// A scene block read through the constant address space: wave-uniform scalar loads.
typedef const __attribute__((address_space(4))) DevRt* DevRtC;
constexpr int REPLAY_WAVES = REPLAY_BLOCK / 64;

// FP32 form: per-wavefront, per-bin running sums of the current ray (Kahan
// optical depth, intensity in log form), kept in LDS.  The last lane of each
// bin's lanes (the bin-major layout below) alone updates them; LDS operations
// of one wavefront complete in order, so the next pass reads the update.
static_assert(HADROS_MAX_BINS <= 16, "the ray-end reduction keeps one bin per lane of a 16-lane row");
struct ReplayAcc32 {
    Kahan tau[HADROS_MAX_BINS];
    ScaledKahan I[HADROS_MAX_BINS];
};

// The scene block for one chunk's per-point quantities.  The block is
// invariant over the ray, but held in SGPRs across the chunk loop its medium
// and source fields (every profile kind's) exceed the SGPR file and spill
// into VGPR lanes, one VALU read-back per use.  Hiding the invariance from
// the optimiser (an empty asm on the pointer) leaves them in memory, re-read
// by scalar loads where the chunk uses them.
__device__ __forceinline__ const DevRt& scene_per_chunk(DevRtC Pc)
{
    asm volatile("" : "+s"(Pc));
    return *(const DevRt*)Pc;
}

// Bin-major transfer.  After a chunk's per-point pass (lane = point) the
// points' (z, ln z, n_N dl, g^3 A dl; the FP32 form stages their logarithms)
// go to LDS, and the bins take the lanes:
// with NBP = nb rounded up to a power of two, bin b owns the G = 64 / NBP
// lanes b G .. b G + G - 1, and lane s of them the chunk's points
// s Lc .. s Lc + Lc - 1 (Lc = ceil(points / G), so a partial last chunk is
// spread over every lane; with more than 8 bins the chunk is taken in two
// halves, so that Lc <= RP_SEG).  A lane sums its points' optical depths in path
// order, a prefix sum over the bin's G lanes gives tau before each segment,
// and a second pass over the segment weights each point's emission by
// e^{-tau before it}.  In the FP64 form each lane's running tau and emission
// partial sum live in LDS across the ray (registers held through the
// per-point pass would spill), and the bin's G emission sums are added at the
// ray's end; the FP32 form adds each pass's optical depth to its bin's Kahan
// sum and the pass's log-sum-exp of the emission terms to its bin's
// intensity.
// Staged arrays: 64 points with one pad double after every 8 (a lane's
// segment start is then bank-spread), 72 doubles each.
constexpr int RP_STAGE = 72;
constexpr int RP_SEG = 8;   // points per lane per pass, at most
__host__ __device__ constexpr int replay_stage_index(int q) { return q + (q >> 3); }
__host__ __device__ constexpr int replay_bins_pow2(int nb)
{
    return nb <= 1 ? 1 : nb <= 2 ? 2 : nb <= 4 ? 4 : nb <= 8 ? 8 : 16;
}

// Doubles of dynamic LDS per wavefront: the four staged point arrays, per
// lane one emission exponent per point of its segment, and (FP64 form) per
// lane the running tau and emission sum.
__host__ __device__ constexpr int replay_emission_slots(int precision, int nb)
{
    return 4 * RP_STAGE + 64 * (replay_bins_pow2(nb) < RP_SEG ? replay_bins_pow2(nb) : RP_SEG) +
           (precision == HADROS_PREC_FP64 ? 128 : 0);
}

// The batch's camera start table is staged in LDS when it is short and the
// dynamic LDS stays within 56 KiB (64 KiB per workgroup with the static part).
__host__ __device__ constexpr bool replay_pix0_in_lds(int precision, int nb, int n_sigma, int n_cams)
{
    return n_cams <= BATCH_LDS_CAMS &&
           (long(REPLAY_WAVES) * replay_emission_slots(precision, nb) + 3L * n_sigma + n_cams + 1) * 8 <= 56L * 1024;
}

// LDS tables and per-wavefront scratch of one wavefront of the replay kernel.
struct ReplayLds {
    double* bin_e;
    double* bin_lne;
    double* bin_w;
    double* bin_de;
    double* bin_sig;
    const double* lnE;
    const double* lnS;
    const double* slope;
    const float* lnEf;
    const float* lnSf;
    const float* slopef;
    const double* exp_t;
    const double* log_t;
    double* wsp;
};

// The bin-major layout of a scene with nb bins (replay_emission_slots): lane =
// (bin bl, segment); lanes of bins beyond nb compute on and drop.
struct BinLayout {
    int nbp, lg_g, G, bl;
    bool on;
};

__device__ __forceinline__ BinLayout bin_layout(int nb, int lane)
{
    BinLayout B;
    B.nbp = replay_bins_pow2(nb);
    B.lg_g = 6 - __builtin_ctz(B.nbp);   // log2 G, G = 64 / nbp lanes per bin
    B.G = 1 << B.lg_g;
    B.bl = lane >> B.lg_g;
    B.on = B.bl < nb;
    return B;
}

// FP64 form: each lane's running tau and emission sum, [tau, emission][lane]
__device__ __forceinline__ double* replay_run(const ReplayLds& L, const BinLayout& B)
{
    return L.wsp + 4 * RP_STAGE + 64 * (B.nbp < RP_SEG ? B.nbp : RP_SEG);
}

// A ray's running sums cleared (and, for the batch source, its scene's bin
// constants staged) before its first point.
template <int PREC, int SRC>
__device__ __forceinline__ void ray_begin(const DevRt& P, const ReplayLds& L, ReplayAcc32& a32, const BinLayout& B,
                                          int lane)
{
    if (lane < P.nb) {
        a32.tau[lane] = {0.0f, 0.0f};
        a32.I[lane] = {{0.0f, 0.0f}, 0.0f, 0};
        if constexpr (SRC == REPLAY_SRC_BATCH) {   // this ray's scene
            L.bin_e[lane] = P.emid[lane];
            L.bin_lne[lane] = P.lnemid[lane];
            L.bin_w[lane] = P.w[lane];
            L.bin_de[lane] = P.de[lane];
            L.bin_sig[lane] = P.sigb[lane];
        }
    }
    if constexpr (PREC == HADROS_PREC_FP64) {
        double* run = replay_run(L, B);
        run[lane] = 0.0;
        run[64 + lane] = 0.0;
    }
    __builtin_amdgcn_wave_barrier();
}

// Stage one point (lane = point) at stage index si: z with the absorption flag
// in its sign (z > 0 always), ln z, n_N dl (n_N dl z^s for a power-law sigma
// table, so that n_N sigma(E_b z) dl = n_N dl z^s sigma(E_b)) and g^3 A dl; the
// FP32 form stages their logarithms (-inf where the point does not emit).  pq
// of a point the reference skips, or of a lane without a point, holds the
// neutral values (z 1, ln z 0, nothing absorbed or emitted).  Returns whether
// the point emits.
// The exponential and logarithm of a point's density, source and ln z: the
// FP64 tables, or in the FP32 form (without the MeV channel, which keeps its
// FP64 terms) the float-accurate forms, as every such term of the FP32 form is
// rounded to float (its staged logarithms) before it is used.
template <int PREC, bool MEV>
__device__ __forceinline__ auto point_exp(const ReplayLds& L)
{
    if constexpr (PREC == HADROS_PREC_FP32K && !MEV) return ExpF32{};
    else return ExpTab{L.exp_t, L.log_t};
}

template <int PREC>
__device__ __forceinline__ bool stage_point(const DevRt& P, const ReplayLds& L, const PointQ& pq, bool ok, int si)
{
    double* wsp = L.wsp;
    const bool spow = P.sig_pow != 0;
    if constexpr (PREC == HADROS_PREC_FP64) {
        const bool absorbs = ok && pq.rho > 0.0;
        const bool emits = ok && P.spec_ok && pq.g3Adl != 0.0;
        wsp[si] = absorbs ? pq.z : -pq.z;
        wsp[RP_STAGE + si] = pq.lnz;
        wsp[2 * RP_STAGE + si] = spow ? pq.nbar_dl * exp_tab(P.sig_s * pq.lnz, L.exp_t) : pq.nbar_dl;
        wsp[3 * RP_STAGE + si] = emits ? pq.g3Adl : 0.0;
        return emits;
    } else {
        const bool absorbs = ok && pq.rho > 0.0 && pq.nbar_dl > 0.0;
        const bool emits = ok && P.spec_ok && pq.g3Adl > 0.0;
        // the staged logarithms are rounded to float: ln_f32 is accurate enough
        const double lnb = (pq.nbar_dl > 0.0) ? ln_f32(pq.nbar_dl) : 0.0;
        wsp[si] = absorbs ? pq.z : -pq.z;
        wsp[RP_STAGE + si] = double(float(pq.lnz));
        wsp[2 * RP_STAGE + si] = double(float(spow ? lnb + P.sig_s * pq.lnz : lnb));
        wsp[3 * RP_STAGE + si] = emits ? double(float(ln_f32(pq.g3Adl))) : double(-INFINITY);
        return emits;
    }
}

__device__ __forceinline__ void neutral_point(PointQ& pq)
{
    pq.z = 1.0;
    pq.lnz = 0.0;
    pq.g3Adl = 0.0;
    pq.nbar_dl = 0.0;
    pq.rho = 0.0;
}

// Bin pass over the staged points [s_lo, s_hi) of the current ray, in path
// order (0 <= s_lo < s_hi <= 64).  Lane s of bin b's G lanes takes the points
// s_lo + s Lc .. s_lo + s Lc + Lc - 1 of each pass (Lc <= RP_SEG); points past
// s_hi (another ray's, or none) are masked.  any_emit: some point of the
// range emits.
template <int PREC>
__device__ __forceinline__ void bin_pass(const DevRt& P, const ReplayLds& L, ReplayAcc32& a32, const BinLayout& B,
                                         int lane, int s_lo, int s_hi, bool any_emit)
{
    HD_CONTRACT
    double* wsp = L.wsp;
    double* ey = wsp + 4 * RP_STAGE;   // [point of segment][lane]
    const int G = B.G, lg_g = B.lg_g, bl = B.bl;
    const bool bin_on = B.on;
    const int n = P.n_sigma;
    const int part = RP_SEG << lg_g;   // points per pass: 64, or 32 with 16 bins
    const double e_lo = P.e_lo, e_hi = P.e_hi;
    const bool exact = P.sig_exact != 0;
    const bool spow = P.sig_pow != 0;
    if constexpr (PREC == HADROS_PREC_FP64) {
        double* run = replay_run(L, B);
        // the scene's bin-loop constants
        const double gam = P.gam, invcut = P.invcut;
        const double sig_lo = P.sig_lo, sig_inv_d = P.sig_inv_d;
        const double eb = bin_on ? L.bin_e[bl] : 1.0, lneb = bin_on ? L.bin_lne[bl] : 0.0;
        const double sgb = bin_on ? L.bin_sig[bl] : 0.0;
        for (int s0 = s_lo; s0 < s_hi; s0 += part) {
            const int c2 = (s_hi - s0 < part) ? s_hi - s0 : part;
            const int se = s0 + c2;
            const int lc = (c2 + G - 1) >> lg_g;   // points per lane
            const int q0 = s0 + (lane & (G - 1)) * lc;
            // pass 1: optical depth in path order; each point's emission
            // exponent less tau before it within the segment
            double loc = 0.0;
            if (spow) {
#pragma unroll 2
                for (int k = 0; k < lc; ++k) {
                    const bool in = q0 + k < se;
                    const int qi = replay_stage_index(in ? q0 + k : s0);
                    const double zs = wsp[qi], lnz = wsp[RP_STAGE + qi], wdl = wsp[2 * RP_STAGE + qi];
                    const double E = eb * fabs(zs);
                    const double dt = bin_dtau64_pow(bin_on && in && zs > 0.0, E, wdl, sgb, e_lo, e_hi);
                    ey[k * 64 + lane] = bin_emis_arg(gam, invcut, E, lneb + lnz) - loc;
                    loc += dt;
                }
            } else {
#pragma unroll 2
                for (int k = 0; k < lc; ++k) {
                    const bool in = q0 + k < se;
                    const int qi = replay_stage_index(in ? q0 + k : s0);
                    const double zs = wsp[qi], lnz = wsp[RP_STAGE + qi], nbdl = wsp[2 * RP_STAGE + qi];
                    const double E = eb * fabs(zs);
                    const double lnEb = lneb + lnz;
                    const double dt = bin_dtau64(exact, bin_on && in && zs > 0.0, E, lnEb, nbdl, L.lnE, L.lnS, L.slope,
                                                 n, sig_lo, sig_inv_d, e_lo, e_hi, L.exp_t);
                    ey[k * 64 + lane] = bin_emis_arg(gam, invcut, E, lnEb) - loc;
                    loc += dt;
                }
            }
            const double incl = group_scan_add(loc, lane, G);
            const double tot = __shfl(incl, lane | (G - 1));
            const double tau_run = run[lane];
            const double pre = tau_run + (incl - loc);   // tau before the segment
            // pass 2: emission; exp(arg) is exactly 0 below EXP64_ZERO_BELOW
            // (behind an opaque column), so such terms are skipped
            if (any_emit) {
                double em_acc = run[64 + lane];
#pragma unroll 2
                for (int k = 0; k < lc; ++k) {
                    const bool in = q0 + k < se;
                    const double g3 = in ? wsp[3 * RP_STAGE + replay_stage_index(q0 + k)] : 0.0;
                    const double arg = ey[k * 64 + lane] - pre;
                    if (bin_on && g3 != 0.0 && !(arg < EXP64_ZERO_BELOW)) em_acc += g3 * exp_tab(arg, L.exp_t);
                }
                run[64 + lane] = em_acc;
            }
            run[lane] = tau_run + tot;
            __builtin_amdgcn_wave_barrier();
        }
    } else {
        const float gam = float(P.gam), invcut = float(P.invcut);
        const float sig_lo = float(P.sig_lo), sig_inv_d = float(P.sig_inv_d);
        const double eb = bin_on ? L.bin_e[bl] : 1.0;
        const float lneb = bin_on ? float(L.bin_lne[bl]) : 0.0f;
        const float lsb = (bin_on && spow) ? logf(float(L.bin_sig[bl])) : 0.0f;
        const bool last = (lane & (G - 1)) == G - 1;   // the lane that updates its bin's sums
        for (int s0 = s_lo; s0 < s_hi; s0 += part) {
            const int c2 = (s_hi - s0 < part) ? s_hi - s0 : part;
            const int se = s0 + c2;
            const int lc = (c2 + G - 1) >> lg_g;
            const int q0 = s0 + (lane & (G - 1)) * lc;
            float loc = 0.0f;
#pragma unroll 2
            for (int k = 0; k < lc; ++k) {
                const bool in = q0 + k < se;
                const int qi = replay_stage_index(in ? q0 + k : s0);
                const double zs = wsp[qi];
                const float lnz = float(wsp[RP_STAGE + qi]), lN = float(wsp[2 * RP_STAGE + qi]);
                const double E = eb * fabs(zs);
                const float lnEb = lneb + lnz;
                const bool ab = bin_on && in && zs > 0.0;
                const float dt = spow ? ((ab && E >= e_lo && E <= e_hi) ? __expf(lN + lsb) : 0.0f)
                                      : bin_dtau32(exact, ab, E, lnEb, lN, L.lnEf, L.lnSf, L.slopef, n, sig_lo,
                                                   sig_inv_d, e_lo, e_hi);
                ey[k * 64 + lane] = double(bin_emis_log32(0.0f, gam, invcut, E, lnEb) - loc);
                loc += dt;
            }
            const float incl = group_scan_f<false>(loc, lane, G);
            const float tot = __shfl(incl, lane | (G - 1));
            const float pre = a32.tau[bl].s + (incl - loc);   // tau before the segment
            if (any_emit) {
                // emission in log form: L = ln(g^3 A dl) + ln S(E) - tau before the
                // point; the bin's sum of e^{L - m} uses the pass's maximum m
                float m = -INFINITY;
#pragma unroll 2
                for (int k = 0; k < lc; ++k) {
                    const bool in = q0 + k < se;
                    const float lp = in ? float(wsp[3 * RP_STAGE + replay_stage_index(q0 + k)]) : -INFINITY;
                    const float Lk = (bin_on && lp > -INFINITY) ? lp + float(ey[k * 64 + lane]) - pre : -INFINITY;
                    ey[k * 64 + lane] = double(Lk);
                    m = fmaxf(m, Lk);
                }
                // terms below e^EXP32_LOG_NEGLIGIBLE vanish in the FP64 result or
                // beside any larger term: a pass of only such terms is skipped
                if (__ballot(!(m < EXP32_LOG_NEGLIGIBLE)) != 0ull) {
                    const float mx = __shfl(group_scan_f<true>(m, lane, G), lane | (G - 1));
                    float w = 0.0f;
                    if (mx > -INFINITY) {
#pragma unroll 2
                        for (int k = 0; k < lc; ++k) w += __expf(float(ey[k * 64 + lane]) - mx);
                    }
                    const float ws = group_scan_f<false>(w, lane, G);
                    if (last && bin_on && mx > -INFINITY) a32.I[bl].add_log(mx + __logf(ws));
                }
            }
            __builtin_amdgcn_wave_barrier();
            if (last && bin_on) a32.tau[bl].add(tot);
            __builtin_amdgcn_wave_barrier();
        }
    }
}

// A ray's outputs at global index g (bins nb_out apart): per-bin tau, P, I,
// one lane per bin, and the aggregates of src/radiative_transfer.cpp:763-769;
// a ray without points keeps the defaults.  The MeV sums are the wavefront's
// (wave-uniform) values.
template <int PREC, bool MEV>
__device__ __forceinline__ void ray_end(const ReplayArgs& A, const DevRt& P, const ReplayLds& L, ReplayAcc32& a32,
                                        const BinLayout& B, int lane, long g, int nb_out, bool has_points,
                                        double m_tau, double m_I, double m_leak, double m_rns)
{
    const int nb = P.nb;
    double bi = 0.0, bt64 = 0.0;
    if constexpr (PREC == HADROS_PREC_FP64) {
        // bin b's tau from lane b G, its intensity from the sum over its G lanes
        const double* run = replay_run(L, B);
        const double isum = group_scan_add(run[64 + lane], lane, B.G);
        const int b0l = (lane < B.nbp ? lane : B.nbp - 1) << B.lg_g;
        bi = __shfl(isum, b0l + B.G - 1);
        bt64 = __shfl(run[lane], b0l);
    }
    double bt = 0.0, bp = 1.0;   // bins beyond the scene's own: tau 0, P 1, I 0
    double sums[4] = {0.0, 0.0, 0.0, 0.0};   // sum over bins of w, w tau, w P, I dE
    if (lane < nb) {
        if constexpr (PREC == HADROS_PREC_FP64) {
            bt = bt64;
        } else {
            const Kahan kt = a32.tau[lane];
            const ScaledKahan ki = a32.I[lane];
            bt = double(kt.s) - double(kt.c);
            bi = ki.set ? (double(ki.k.s) - double(ki.k.c)) * exp(double(ki.off)) : 0.0;
        }
        bp = exp(-bt);
        const double w = L.bin_w[lane];
        sums[0] = w;
        sums[1] = w * bt;
        sums[2] = w * bp;
        sums[3] = bi * L.bin_de[lane];
    } else {
        bi = 0.0;
    }
    if (lane < nb_out) {
        // the output index formed here: hoisted out of the ray loop, the three
        // per-lane base addresses stayed live across it and spilled
        long o = g * nb_out + lane;
        asm volatile("" : "+v"(o));
        if (A.btau) A.btau[o] = bt;
        if (A.bp) A.bp[o] = bp;
        if (A.bi) A.bi[o] = bi;
    }
    row_scan_add(sums);
    const double ws = read_lane(sums[0], 15), ts = read_lane(sums[1], 15);
    const double ps = read_lane(sums[2], 15), is = read_lane(sums[3], 15);
    const double t0 = read_lane(bt, 0), p0 = read_lane(bp, 0), i0 = read_lane(bi, 0);
    if (lane == 0) {
        double agg_t = 0.0, agg_p = 1.0, agg_i = 0.0;
        if (has_points) {
            if (P.mono) {
                if (nb > 0) {
                    agg_t = t0;
                    agg_p = p0;
                    agg_i = i0;
                }
            } else if (ws > 0.0) {
                agg_t = ts / ws;
                agg_p = ps / ws;
                agg_i = is;
            }
        }
        if (A.tau) A.tau[g] = agg_t;
        if (A.p) A.p[g] = agg_p;
        if (A.I) A.I[g] = agg_i;
        if constexpr (MEV) {
            if (A.tau_m) A.tau_m[g] = m_tau;
            if (A.p_m) A.p_m[g] = exp(-m_tau);
            if (A.I_m) A.I_m[g] = m_I;
            if (A.rns_m) A.rns_m[g] = m_rns;
            // the reference's leakage factor divides the summed j dl by the same sum
            // (src/radiative_transfer.cpp:630-631, 681-682): 1 wherever it is positive
            if (A.leak_m) A.leak_m[g] = m_leak > 0.0 ? m_leak / m_leak : 1.0;
        }
    }
    __builtin_amdgcn_wave_barrier();
}

// One ray of the replay, one 64-point window at a time: points from the ray's
// source, transfer with the scene's parameters P, outputs at global ray index
// g (bins nb_out apart).  A ray's windows are its own chunk-aligned 64-point
// spans, so its results do not depend on which wavefront takes it (the
// tracer's bands and the batch agree bit for bit with a one-launch run).
template <int PREC, int SRC, bool MEV>
__device__ __forceinline__ unsigned long long replay_ray(const ReplayArgs& A, DevRtC Pc, long ray, long g,
                                                         int nb_out, int lane, ReplayAcc32& a32, const ReplayLds& L)
{
    HD_CONTRACT
    const DevRt& P = *(const DevRt*)Pc;
    const int nb = P.nb;
    // points of this ray: [b0, b1) of the SoA arrays, or n_st points in linked chunks
    long b0 = 0, b1 = 0;
    int chunk = int(ray);
    if constexpr (SRC == REPLAY_SRC_RANGES) {
        b0 = uniform_long(A.offset[ray]);
        b1 = uniform_long(A.offset[ray + 1]);
    } else {
        b1 = uniform_long(A.n_st[ray]);
    }
    const BinLayout B = bin_layout(nb, lane);
    ray_begin<PREC, SRC>(P, L, a32, B, lane);
    // MeV channel of integrate_kerr_ray (src/radiative_transfer.cpp:500-640), kept
    // wave-uniform: tau_mev, I_mev, the summed j dl (leakage) and r_ns.  The
    // spectral reference takes it from its first integrated bin, so a scene
    // with no integrated UHE bin keeps the defaults.
    const bool mev_on = MEV && (P.mono || nb > 0);
    double m_tau = 0.0, m_I = 0.0, m_leak = 0.0, m_rns = -1.0;
    for (long base = b0; base < b1; base += 64) {
        PointQ pq;
        bool ok = false;
        double pr = 0.0, pth = 0.0, pdl = 0.0;
        if constexpr (SRC == REPLAY_SRC_RANGES) {
            const long q = base + lane;
            if (q < b1) {
                pr = A.r[q];
                pth = A.th[q];
                pdl = A.dl[q];
                ok = point_q(scene_per_chunk(Pc), pr, pth, pdl, A.z[q], pq, point_exp<PREC, MEV>(L));
            }
        } else {
            if (base + lane < b1) {
                const double2* c = reinterpret_cast<const double2*>(A.pool + long(chunk) * CHUNK_DOUBLES + 4 * lane);
                const double2 rt = c[0], dz = c[1];
                pr = rt.x;
                pth = rt.y;
                pdl = dz.x;
                ok = point_q(scene_per_chunk(Pc), rt.x, rt.y, dz.x, dz.y, pq, point_exp<PREC, MEV>(L));
            }
            if (base + 64 < b1) chunk = A.next[chunk];
        }
        if constexpr (MEV) {
            if (mev_on) {
                // per point: I -> a I + b, tau += dtau (lane order = path order)
                const DevMev& M = *A.mev;
                double dtau[1] = {0.0}, lk[1] = {0.0}, a = 1.0, b = 0.0;
                if (ok) {
                    const double gz = 1.0 / pq.z;
                    const double dl_cm = pdl * P.rg_cm;
                    const double T = mev_temperature(M, pr, pth, pq.rho);
                    const double Ye = mev_ye(M, pr, pth, pq.rho);
                    double kappa = 0.0, j = 0.0;
                    mev_kappa_j(M, gz, pq.rho, T, Ye, kappa, j);
                    dtau[0] = kappa * dl_cm;
                    const double g3 = pow(gz, 3.0);
                    if (kappa > 0.0 && dtau[0] > 1.0e-12) {
                        a = exp(-dmin(dtau[0], 700.0));
                        b = g3 * (j / kappa) * (1.0 - a);
                    } else {
                        b = g3 * j * dl_cm;
                    }
                    lk[0] = j * dl_cm;
                }
                wave_scan_add(dtau);
                wave_scan_add(lk);
                wave_scan_affine(a, b);
                if (m_rns < 0.0) {
                    const unsigned long long hit = __ballot(ok && m_tau + dtau[0] >= M.ns_tau);
                    if (hit) m_rns = read_lane(pr, __ffsll(static_cast<long long>(hit)) - 1);
                }
                m_tau += lane63(dtau[0]);
                m_leak += lane63(lk[0]);
                m_I = lane63(a) * m_I + lane63(b);
            }
        }
        if (!ok) neutral_point(pq);
        const bool emits = stage_point<PREC>(P, L, pq, ok, replay_stage_index(lane));
        const bool any_emit = __ballot(emits) != 0ull;
        __builtin_amdgcn_wave_barrier();
        const long left = b1 - base;
        bin_pass<PREC>(P, L, a32, B, lane, 0, left < 64 ? int(left) : 64, any_emit);
    }
    ray_end<PREC, MEV>(A, P, L, a32, B, lane, g, nb_out, b1 > b0, m_tau, m_I, m_leak, m_rns);
    return (unsigned long long)(b1 - b0);
}

template <int PREC, int SRC, bool MEV>
__global__ void __launch_bounds__(REPLAY_BLOCK)
// the MeV channel needs the registers; the FP32 form's float-accurate
// per-point terms leave room for 4 waves per SIMD
__attribute__((amdgpu_waves_per_eu(MEV ? 2 : (PREC == HADROS_PREC_FP32K ? 4 : HD_REPLAY_WAVES_PER_EU))))
k_replay(ReplayArgs A)
{
    extern __shared__ double smem[];
    __shared__ ReplayAcc32 acc32[REPLAY_WAVES];
    // per-wavefront bin constants (the batch source changes scene per ray),
    // indexed by the rolled bin loop
    __shared__ double bin_tab[REPLAY_WAVES][5][HADROS_MAX_BINS];
    __shared__ double exp_t[64];   // 2^(j/64) for exp_tab
    __shared__ __align__(16) double log_t[2 * LOG_TAB_N];   // (1/F, ln F) for log_tab
    if (threadIdx.x < 64) exp_t[threadIdx.x] = EXP_TAB[threadIdx.x];
    if (threadIdx.x < 2 * LOG_TAB_N) log_t[threadIdx.x] = LOG_TAB[threadIdx.x];
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    double* bin_e = bin_tab[wv][0];
    double* bin_lne = bin_tab[wv][1];
    double* bin_w = bin_tab[wv][2];
    double* bin_de = bin_tab[wv][3];
    double* bin_sig = bin_tab[wv][4];
    if constexpr (SRC != REPLAY_SRC_BATCH) {
        const DevRt& P = A.rt;
        if (lane < HADROS_MAX_BINS) {
            bin_e[lane] = P.emid[lane];
            bin_lne[lane] = P.lnemid[lane];
            bin_w[lane] = P.w[lane];
            bin_de[lane] = P.de[lane];
            bin_sig[lane] = P.sigb[lane];
        }
    }
    // dynamic LDS: the FP64 form's staged points and emission exponents
    // (replay_emission_slots per wavefront), the sigma table (shared by every scene), then the batch's
    // camera start table
    const int n_em = replay_emission_slots(PREC, SRC == REPLAY_SRC_BATCH ? A.nb_out : A.rt.nb);
    double* wsp = smem + wv * n_em;
    double* sig = smem + REPLAY_WAVES * n_em;
    const int n = A.rt.n_sigma;
    const double* lnE = sig;
    const double* lnS = sig + n;
    const double* slope = sig + 2 * n;
    const float* lnEf = reinterpret_cast<const float*>(sig);
    const float* lnSf = lnEf + n;
    const float* slopef = lnEf + 2 * n;
    if constexpr (PREC == HADROS_PREC_FP64) {
        for (int t = threadIdx.x; t < 3 * n; t += blockDim.x) sig[t] = A.sig[t];
    } else {
        float* sf = reinterpret_cast<float*>(sig);
        for (int t = threadIdx.x; t < 3 * n; t += blockDim.x) sf[t] = A.sigf[t];
    }
    long* pix0_lds = reinterpret_cast<long*>(sig + 3 * n);
    const bool pix_in_lds = SRC == REPLAY_SRC_BATCH && replay_pix0_in_lds(PREC, A.nb_out, n, A.n_cams);
    const long* pix0 = pix_in_lds ? pix0_lds : A.pix0;
    if (pix_in_lds)
        for (int t = threadIdx.x; t <= A.n_cams; t += blockDim.x) pix0_lds[t] = A.pix0[t];
    __syncthreads();

    ReplayAcc32& a32 = acc32[wv];
    // A.rt where it lies in the kernel argument segment (the kernel's only
    // explicit argument starts at offset 0; offsetof needs a standard layout)
    static_assert(std::is_standard_layout<ReplayArgs>::value, "ReplayArgs is read at its kernarg offsets");
    typedef const __attribute__((address_space(4))) char* KernargC;
    const DevRtC rt_c = (DevRtC)((KernargC)__builtin_amdgcn_kernarg_segment_ptr() + offsetof(ReplayArgs, rt));
    const long nwaves = long(gridDim.x) * REPLAY_WAVES;
    const ReplayLds L{bin_e, bin_lne, bin_w, bin_de, bin_sig, lnE, lnS, slope, lnEf, lnSf, slopef, exp_t, log_t, wsp};
    unsigned long long pts_done = 0;
    const long first = long(blockIdx.x) * REPLAY_WAVES + wv;
    if constexpr (SRC == REPLAY_SRC_BATCH) {
        for (long ray = first; ray < A.n_rays; ray += nwaves) {
            // queue index -> camera -> output pixel (batch_pixel's order); the
            // camera and scene tables are read-only here, so they are read
            // through the constant address space (wave-uniform scalar loads)
            const long gq = A.ray0 + ray;
            const int c = __builtin_amdgcn_readfirstlane(batch_camera(pix0, A.n_cams, gq));
            const DevCam& C = *(const DevCam*)(const __attribute__((address_space(4))) DevCam*)(A.cams + c);
            int p, q;
            batch_pixel(C, gq - pix0[c], p, q);
            const long g = pix0[c] + long(p) * C.sub_ny + q;
            pts_done += replay_ray<PREC, SRC, MEV>(A, (DevRtC)(A.rts + C.pad_rt), ray, g, A.nb_out, lane, a32, L);
        }
    } else {
        for (long ray = first; ray < A.n_rays; ray += nwaves) {
            // the chunk source may cover a range of the tracer's rays: ray0 + ray
            const long gr = (SRC == REPLAY_SRC_CHUNKS) ? A.ray0 + ray : ray;
            pts_done += replay_ray<PREC, SRC, MEV>(A, rt_c, gr, gr, A.rt.nb, lane, a32, L);
        }
    }
    if (lane == 0 && A.counters) atomicAdd(A.counters, pts_done);
}

template <int PREC, int SRC>
static void launch_p(const ReplayArgs& A, int grid, size_t lds, hipStream_t s)
{
    // the MeV channel is a separate instantiation, so the UHE-only kernels keep their registers
    if constexpr (SRC != REPLAY_SRC_BATCH) {
        if (A.mev) {
            hipLaunchKernelGGL((k_replay<PREC, SRC, true>), dim3(grid), dim3(REPLAY_BLOCK), lds, s, A);
            return;
        }
    }
    hipLaunchKernelGGL((k_replay<PREC, SRC, false>), dim3(grid), dim3(REPLAY_BLOCK), lds, s, A);
}

hipError_t launch_replay(const ReplayArgs& A, int precision, hipStream_t s)
{
    long grid = (long(A.n_rays) + REPLAY_WAVES - 1) / REPLAY_WAVES;
    const long max_grid = 256L * 64;   // persistent-ish: wavefronts loop over rays
    if (grid > max_grid) grid = max_grid;
    if (grid < 1) grid = 1;
    const bool batch = A.rts != nullptr;
    // the FP32 sigma table takes the first half of the FP64 one's space
    const int n_em = replay_emission_slots(precision, batch ? A.nb_out : A.rt.nb);
    size_t lds = (size_t(REPLAY_WAVES) * n_em + size_t(3) * A.rt.n_sigma) * sizeof(double);
    if (batch && replay_pix0_in_lds(precision, A.nb_out, A.rt.n_sigma, A.n_cams))
        lds += size_t(A.n_cams + 1) * sizeof(long);
    const bool chunks = A.pool != nullptr;
    if (precision == HADROS_PREC_FP64) {
        if (batch) launch_p<HADROS_PREC_FP64, REPLAY_SRC_BATCH>(A, int(grid), lds, s);
        else if (chunks) launch_p<HADROS_PREC_FP64, REPLAY_SRC_CHUNKS>(A, int(grid), lds, s);
        else launch_p<HADROS_PREC_FP64, REPLAY_SRC_RANGES>(A, int(grid), lds, s);
    } else {
        if (batch) launch_p<HADROS_PREC_FP32K, REPLAY_SRC_BATCH>(A, int(grid), lds, s);
        else if (chunks) launch_p<HADROS_PREC_FP32K, REPLAY_SRC_CHUNKS>(A, int(grid), lds, s);
        else launch_p<HADROS_PREC_FP32K, REPLAY_SRC_RANGES>(A, int(grid), lds, s);
    }
    return hipGetLastError();
}

}  // namespace hd
