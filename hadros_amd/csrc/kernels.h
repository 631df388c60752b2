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

struct PointsArgs {
    DevCam cam;
    long n_rays;            // nx * ny: chunk k < n_rays is pixel k's first chunk
    long capacity;          // chunks in the pool
    double* pool;           // capacity * CHUNK_DOUBLES
    int* next;              // [capacity]
    int* n_st;              // [n_rays] points stored per pixel
    unsigned* chunk_ctr;    // extra chunks requested
    TraceOut out;           // captured, npoints, flags, alpha, beta, counters
};

struct CacheArgs {
    DevCam cam;
    long k0, k1;          // pixel range [k0, k1), pixel k = i*ny + j
    const long* offset;   // [npix+1] (write pass)
    double* points;       // 7 doubles per point (write pass)
    long* counts;         // [npix] (count pass)
    int* captured;        // [npix] (count pass)
};

// chunk pool of the split tracer: chunk c holds up to 64 points as 32-byte
// records (r, theta, dl, z) at pool + c * CHUNK_DOUBLES + 4 p, so each lane's
// stores fill whole 128-byte lines every 4 points
constexpr int CHUNK_DOUBLES = 256;
enum { REPLAY_SRC_RANGES = 0, REPLAY_SRC_CHUNKS = 1 };

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
    double *tau, *p, *I, *btau, *bp, *bi;
    unsigned long long* counters;
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

// point-wise hooks (eval.hip)
hipError_t launch_eval_rho(const DevMedium& m, const double* r, const double* th, double* out, long n);
hipError_t launch_eval_sigma(const double* sig, int n_sigma, double e_lo, double e_hi, double lo0, double inv_d,
                             const double* e, double* out, long n);
hipError_t launch_eval_emis(const DevRt& P, const double* r, const double* th, const double* e, double* out,
                            long n);
hipError_t launch_eval_rhs(double a, int mode, const double* y, double* dy, long n);
hipError_t launch_eval_metric(double a, const double* r, const double* th, double* out, long n);

}  // namespace hd
