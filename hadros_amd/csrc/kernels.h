// kernels.h -- kernel argument blocks and launchers shared by the .hip units.
#pragma once

#include "physics.h"

namespace hd {

struct TraceOut {
    double *tau, *p, *I, *alpha, *beta, *btau, *bp, *bi;
    int *captured, *npoints, *flags;
    unsigned long long* counters;   // [0] points, [1] attempts, [2] captured, [3] exhausted
};

struct TraceArgs {
    DevCam cam;
    DevRt rt;
    const DevRt* rtg;     // device copy of rt, read per path point (see k_trace_rt)
    const double* sig;    // 3*n_sigma: ln E, ln sigma, slope (FP64)
    const float* sigf;    // same, FP32
    TraceOut out;
};

// chunk pool of the split tracer: chunk c holds up to 64 points as 32-byte
// records (r, theta, dl, z) at pool + c * CHUNK_DOUBLES + 4 p, so each lane's
// stores fill whole 128-byte lines every 4 points
constexpr int CHUNK_DOUBLES = 256;
enum { REPLAY_SRC_RANGES = 0, REPLAY_SRC_CHUNKS = 1, REPLAY_SRC_BATCH = 2 };

// The pool as the geodesic kernels see it: chunks [0, base) are the rays'
// first chunks (ray r starts in chunk r), further chunks are handed out by
// the counter from `base` on.
struct ChunkPool {
    double* pool;        // capacity * CHUNK_DOUBLES
    int* next;           // [capacity] next chunk of the same ray
    unsigned* ctr;       // extra chunks requested
    long base, capacity;
};

// One lane's writer of its ray's path points.  Two registers of state: the
// chunk being filled (-1 once the pool has run out) and the points stored.
// After running out the ray keeps integrating; finish() then adds the chunks
// it still lacked to the counter, so the counter ends at the total demand.
struct ChunkWriter {
    int cur, stored;
    __device__ __forceinline__ void begin(long first) { cur = int(first); stored = 0; }
    __device__ __forceinline__ void put(const ChunkPool& P, double r, double th, double dl, double z)
    {
        if (cur < 0) return;
        const int fill = stored & 63;
        if (fill == 0 && stored > 0) {
            const long c = P.base + long(atomicAdd(P.ctr, 1u));
            if (c >= P.capacity) {
                cur = -1;
                return;
            }
            P.next[cur] = int(c);
            cur = int(c);
        }
        // one 32-byte record: (r, theta) and (dl, z) as two 16-byte stores
        double2* o = reinterpret_cast<double2*>(P.pool + long(cur) * CHUNK_DOUBLES + 4 * fill);
        o[0] = make_double2(r, th);
        o[1] = make_double2(dl, z);
        ++stored;
    }
    // emitted: points put() was called with; only for a ray that began
    __device__ __forceinline__ void finish(const ChunkPool& P, int emitted) const
    {
        if (cur >= 0) return;
        // the failed request was the ray's stored / 64-th; it needed ceil(emitted / 64) - 1
        const int need = (emitted + 63) / 64 - 1, requested = stored / 64;
        if (need > requested) atomicAdd(P.ctr, unsigned(need - requested));
    }
};

struct PointsArgs {
    DevCam cam;
    ChunkPool cp;           // base = nx * ny: chunk k < base is pixel k's first chunk
    int* n_st;              // [nx * ny] points stored per pixel
    TraceOut out;           // captured, npoints, flags, alpha, beta, counters
    int tile0, tile1;       // the launch's 8x8 tiles [tile0, tile1), row-major over tile rows
};

struct CacheArgs {
    DevCam cam;
    long k0, k1;          // pixel range [k0, k1), pixel k = i*ny + j
    const long* offset;   // [npix+1] (write pass)
    double* points;       // 7 doubles per point (write pass)
    long* counts;         // [npix] (count pass)
    int* captured;        // [npix] (count pass)
};

struct ReplayArgs {
    DevRt rt;
    const double* sig;
    const float* sigf;
    int n_rays;
    const long* offset;                 // ranges source: [n_rays+1]
    const double *r, *th, *dl, *z;
    const double* pool;                 // chunks source (non-null selects it)
    const int* next;                    // [capacity] next chunk of the same ray
    const int* n_st;                    // [n_rays] points stored per ray
    // batch source (non-null rts selects it, with the pool): local ray r is
    // global ray ray0 + r, whose camera (last c with pix0[c] <= ray) has the
    // transfer parameters rts[cams[c].pad_rt]; bins are written nb_out apart
    const DevRt* rts;
    const DevCam* cams;
    const long* pix0;
    int n_cams, nb_out;
    long ray0;
    double *tau, *p, *I, *btau, *bp, *bi;
    unsigned long long* counters;
    // MeV channel (non-null mev selects it; not with the batch source)
    const DevMev* mev;
    double *tau_m, *p_m, *I_m, *rns_m, *leak_m;
};

struct OpacityArgs {
    DevMedium m;
    const double* sig;        // 3*n_sigma: ln E, ln sigma, slope
    int n_sigma, nr, ntheta, pad;
    double sig_lo, sig_inv_d; // index guess of the sigma lookup
    double ln_enu, rg_cm, r_min, dr, tau_s;
    double *theta, *r_tau, *tau_total, *cls;
    int* crossing;
};

// batches of up to this many cameras keep their camera start table in LDS
constexpr int BATCH_LDS_CAMS = 1024;

struct BatchArgs {
    const DevCam* cams;   // [n_cams], cams[c].pad_rt indexes rts
    const long* pix0;     // [n_cams+1] first global ray of each camera
    const DevRt* rts;     // per-scene transfer parameters
    const double* sig;    // shared sigma table (3*n_sigma, FP64)
    const float* sigf;    // same, FP32
    int n_sigma, n_cams, nb_out, steps_per_chunk;
    long n_rays;
    unsigned long long* queue;   // next unclaimed ray
    TraceOut out;                // indexed by global ray (bins: ray*nb_out + b)
};

// Geodesic half of the split batch: the rays [ray0, ray0 + n_rays) of the
// batch's global ray order (a sub-batch), each storing its path points in the
// chunk pool (local ray r starts in chunk r, so cp.base = n_rays).
struct BatchPointsArgs {
    const DevCam* cams;   // [n_cams]
    const long* pix0;     // [n_cams+1] first global ray of each camera
    int n_cams, steps_per_chunk;
    long ray0, n_rays;
    unsigned long long* queue;   // next unclaimed local ray
    ChunkPool cp;
    int* n_st;                   // [n_rays] points stored per local ray
    TraceOut out;                // captured, npoints, flags by global ray; counters
};

// camera of global ray `ray`: the last c with pix0[c] <= ray
__device__ __forceinline__ int batch_camera(const long* __restrict__ pix0, int n_cams, long ray)
{
    int lo = 0, hi = n_cams - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pix0[mid] <= ray) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Queue order of a batch camera's (sub-)lattice pixels: first the 8x8 tiles
// that fit whole, tile by tile (64 consecutive rays are one tile, so the rays
// a wavefront takes are neighbours on the sky and step alike), then the
// pixels right of them row by row, then those below them row by row.  Maps
// the index k in that order to the lattice pixel (p, q); the ray's outputs go
// to pixel p * sub_ny + q of the camera.
__device__ __forceinline__ void batch_pixel(const DevCam& C, long k, int& p, int& q)
{
    const int TX = C.sub_nx >> 3, TY = C.sub_ny >> 3;
    const long full = 64L * TX * TY;
    if (k < full) {
        const int t = int(k >> 6), l = int(k & 63);
        p = (t / TY) * 8 + (l >> 3);
        q = (t % TY) * 8 + (l & 7);
        return;
    }
    long r = k - full;
    const int wr = C.sub_ny - 8 * TY;   // right strip: every row, columns 8 TY ..
    const long right = long(C.sub_nx) * wr;
    if (r < right) {
        p = int(r / wr);
        q = 8 * TY + int(r % wr);
        return;
    }
    r -= right;   // bottom strip: rows 8 TX .., columns < 8 TY (TY > 0 here)
    p = 8 * TX + int(r / (8 * TY));
    q = int(r % (8 * TY));
}

// Sum four 64-bit per-lane values over the wavefront and add them to
// counters[0..3] from lane 0 (one atomic per wave, not per lane).
__device__ __forceinline__ void wave_add4(unsigned long long* c, unsigned long long a, unsigned long long b,
                                          unsigned long long d, unsigned long long e)
{
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
        d += __shfl_xor(d, off, 64);
        e += __shfl_xor(e, off, 64);
    }
    if ((threadIdx.x & 63) == 0 && c) {
        atomicAdd(c + 0, a);
        atomicAdd(c + 1, b);
        atomicAdd(c + 2, d);
        atomicAdd(c + 3, e);
    }
}

// host helpers of hadros_abi.hip: set the thread's last error (returns code),
// check for a usable gfx950 device (HADROS_OK or HADROS_ERR_DEVICE)
int set_error(int code, const char* msg);
int device_ready();

hipError_t launch_zero_u64(unsigned long long* p, int n, hipStream_t s);
hipError_t launch_trace_rt(const TraceArgs& A, int precision, int block, hipStream_t s);
hipError_t launch_trace_points(const PointsArgs& A, int block, hipStream_t s);
hipError_t launch_trace_cache(const CacheArgs& A, bool write, hipStream_t s);
hipError_t launch_replay(const ReplayArgs& A, int precision, hipStream_t s);
hipError_t launch_opacity_surface(const OpacityArgs& A, hipStream_t s);
hipError_t launch_trace_batch(const BatchArgs& A, int mode, int precision, int nb, int grid, int block, hipStream_t s);
hipError_t batch_resident_blocks(int mode, int precision, int nb, int block, int n_sigma, int n_cams, int* blocks_per_cu);
hipError_t launch_batch_points(const BatchPointsArgs& A, int mode, int grid, int block, hipStream_t s);
hipError_t batch_points_resident_blocks(int mode, int block, int n_cams, int* blocks_per_cu);

// point-wise hooks (eval.hip)
hipError_t launch_eval_rho(const DevMedium& m, const double* r, const double* th, double* out, long n);
hipError_t launch_eval_sigma(const double* sig, int n_sigma, double e_lo, double e_hi, double lo0, double inv_d,
                             const double* e, double* out, long n);
hipError_t launch_eval_emis(const DevRt& P, const double* r, const double* th, const double* e, double* out,
                            long n);
hipError_t launch_eval_rhs(double a, int mode, const double* y, double* dy, long n);
hipError_t launch_eval_mev(const DevMedium& med, const DevMev& M, const double* r, const double* th, const double* g,
                           double* out, long n);
hipError_t launch_eval_metric(double a, const double* r, const double* th, double* out, long n);
enum { EVAL_STEP_RK4 = 0, EVAL_STEP_RKF45_REF = 1, EVAL_STEP_RKF45_CFG = 2, EVAL_STEP_RKF45_CFG_SHIFT = 3 };
hipError_t launch_eval_step(double a, int kind, double h, const double* y, double* out, long n);
hipError_t launch_stream_probe(int kind, double* x, long n, double* partial, double v, int grid, hipStream_t st);
constexpr int EVAL_TRANSFER_COLS = 6 + 2 * HADROS_MAX_BINS;
hipError_t launch_eval_transfer_point(const DevRt& P, int precision, const double* sig, const float* sigf,
                                      const double* r, const double* th, const double* dl, const double* z,
                                      double* out, long n);

}  // namespace hd
