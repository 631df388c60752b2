// wave.h -- wavefront-wide prefix sums and lane reads on gfx950 (wave64).
//
// Prefix sums use DPP lane moves: row shifts by 1, 2, 4, 8 within each row of
// 16 lanes, then the row-15 broadcast into rows 1 and 3 and the row-31
// broadcast into rows 2 and 3.  Lanes whose source is outside the row, or
// whose row is masked off, keep the identity, so no LDS traffic is needed.
#pragma once

#include <hip/hip_runtime.h>

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

// DPP row shift of a double with every row enabled: lanes whose source is
// outside the row read 0 (bound_ctrl), so no identity needs to be written first
template <int CTRL>
__device__ __forceinline__ double dpp_shift(double v)
{
    const unsigned long long u = static_cast<unsigned long long>(__double_as_longlong(v));
    const unsigned lo = unsigned(__builtin_amdgcn_mov_dpp(int(u & 0xffffffffull), CTRL, 0xf, 0xf, true));
    const unsigned hi = unsigned(__builtin_amdgcn_mov_dpp(int(u >> 32), CTRL, 0xf, 0xf, true));
    return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

// inclusive prefix sums over the 64 lanes of N independent values per lane
template <int N>
__device__ __forceinline__ void wave_scan_add(double (&x)[N])
{
#define HD_SHIFT_STEP(CTRL) _Pragma("unroll") for (int b = 0; b < N; ++b) x[b] += dpp_shift<CTRL>(x[b]);
#define HD_SCAN_STEP(CTRL, RM)                                            \
    _Pragma("unroll") for (int b = 0; b < N; ++b) x[b] += dpp_move<CTRL, RM>(x[b]);
    HD_SHIFT_STEP(0x111)       // row_shr:1
    HD_SHIFT_STEP(0x112)       // row_shr:2
    HD_SHIFT_STEP(0x114)       // row_shr:4
    HD_SHIFT_STEP(0x118)       // row_shr:8
    HD_SCAN_STEP(0x142, 0xa)   // row_bcast:15 into rows 1, 3
    HD_SCAN_STEP(0x143, 0xc)   // row_bcast:31 into rows 2, 3
#undef HD_SCAN_STEP
#undef HD_SHIFT_STEP
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

// inclusive prefix sums within each row of 16 lanes: lane 15 of a row ends
// with the row's total
template <int N>
__device__ __forceinline__ void row_scan_add(double (&x)[N])
{
#define HD_SHIFT_STEP(CTRL) _Pragma("unroll") for (int b = 0; b < N; ++b) x[b] += dpp_shift<CTRL>(x[b]);
    HD_SHIFT_STEP(0x111)
    HD_SHIFT_STEP(0x112)
    HD_SHIFT_STEP(0x114)
    HD_SHIFT_STEP(0x118)
#undef HD_SHIFT_STEP
}

// Inclusive prefix sums within groups of G consecutive lanes (G a power of
// two from 1 to 64, the same on every lane): lane l ends with the sum over
// lanes (l & ~(G-1)) .. l.  Row shifts by 1, 2, 4, 8 masked to the group,
// then for groups of 32 and 64 lanes the row-15 and row-31 broadcasts.
__device__ __forceinline__ double group_scan_add(double x, int lane, int G)
{
    const int gl = lane & (G - 1);
    if (G > 1) { const double d = dpp_shift<0x111>(x); x += (gl >= 1) ? d : 0.0; }
    if (G > 2) { const double d = dpp_shift<0x112>(x); x += (gl >= 2) ? d : 0.0; }
    if (G > 4) { const double d = dpp_shift<0x114>(x); x += (gl >= 4) ? d : 0.0; }
    if (G > 8) { const double d = dpp_shift<0x118>(x); x += (gl >= 8) ? d : 0.0; }
    if (G > 16) x += dpp_move<0x142, 0xa>(x);   // row_bcast:15 into rows 1, 3
    if (G > 32) x += dpp_move<0x143, 0xc>(x);   // row_bcast:31 into rows 2, 3
    return x;
}

// The same for floats, with the sum's or the maximum's identity in lanes
// that have no source
template <bool MAX>
__device__ __forceinline__ float group_scan_f(float x, int lane, int G)
{
    const float id = MAX ? -INFINITY : 0.0f;
    const int gl = lane & (G - 1);
#define HD_GSTEP(CTRL, RM, S)                                         \
    {                                                                 \
        const float d = dpp_move<CTRL, RM>(x, id);                    \
        const float v = (gl >= (S)) ? d : id;                         \
        x = MAX ? fmaxf(x, v) : x + v;                                \
    }
    if (G > 1) HD_GSTEP(0x111, 0xf, 1)
    if (G > 2) HD_GSTEP(0x112, 0xf, 2)
    if (G > 4) HD_GSTEP(0x114, 0xf, 4)
    if (G > 8) HD_GSTEP(0x118, 0xf, 8)
    if (G > 16) HD_GSTEP(0x142, 0xa, 0)   // row_bcast:15 into rows 1, 3
    if (G > 32) HD_GSTEP(0x143, 0xc, 0)   // row_bcast:31 into rows 2, 3
#undef HD_GSTEP
    return x;
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

// DPP move of a double whose lanes without a source keep `identity`
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_move_id(double v, double identity)
{
    const unsigned long long u = static_cast<unsigned long long>(__double_as_longlong(v));
    const unsigned long long id = static_cast<unsigned long long>(__double_as_longlong(identity));
    const unsigned lo = unsigned(dpp_word<CTRL, ROW_MASK>(int(id & 0xffffffffull), int(u & 0xffffffffull)));
    const unsigned hi = unsigned(dpp_word<CTRL, ROW_MASK>(int(id >> 32), int(u >> 32)));
    return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

// Inclusive scan of affine maps x -> a x + b in lane order: lane l ends with
// the composition of the maps of lanes 0..l (lane 0's applied first), i.e.
// the pair (a_0 ... a_l, sum_k b_k a_{k+1} ... a_l).  Identity (1, 0).
__device__ __forceinline__ void wave_scan_affine(double& a, double& b)
{
#define HD_AFFINE_STEP(CTRL, RM)                                   \
    {                                                               \
        const double pa = dpp_move_id<CTRL, RM>(a, 1.0);            \
        const double pb = dpp_move_id<CTRL, RM>(b, 0.0);            \
        b = pb * a + b;                                             \
        a = pa * a;                                                 \
    }
    HD_AFFINE_STEP(0x111, 0xf)
    HD_AFFINE_STEP(0x112, 0xf)
    HD_AFFINE_STEP(0x114, 0xf)
    HD_AFFINE_STEP(0x118, 0xf)
    HD_AFFINE_STEP(0x142, 0xa)
    HD_AFFINE_STEP(0x143, 0xc)
#undef HD_AFFINE_STEP
}

__device__ __forceinline__ double lane63(double v)
{
    const unsigned long long u = static_cast<unsigned long long>(__double_as_longlong(v));
    const unsigned lo = unsigned(__builtin_amdgcn_readlane(int(u & 0xffffffffull), 63));
    const unsigned hi = unsigned(__builtin_amdgcn_readlane(int(u >> 32), 63));
    return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

__device__ __forceinline__ float lane63(float v) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63)); }

// a value the same on every lane (a wave-uniform load), moved to scalar
// registers so that loops bounded by it are counted in scalar code
__device__ __forceinline__ long uniform_long(long v)
{
    const unsigned long long u = static_cast<unsigned long long>(v);
    const unsigned lo = unsigned(__builtin_amdgcn_readfirstlane(int(u & 0xffffffffull)));
    const unsigned hi = unsigned(__builtin_amdgcn_readfirstlane(int(u >> 32)));
    return static_cast<long>((static_cast<unsigned long long>(hi) << 32) | lo);
}

// value of lane `l` (wave-uniform) for every lane
__device__ __forceinline__ double read_lane(double v, int l)
{
    const unsigned long long u = static_cast<unsigned long long>(__double_as_longlong(v));
    const unsigned lo = unsigned(__builtin_amdgcn_readlane(int(u & 0xffffffffull), l));
    const unsigned hi = unsigned(__builtin_amdgcn_readlane(int(u >> 32), l));
    return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

}  // namespace hd
