/*
 * hadros.h -- C-ABI of the MI355X (gfx950) per-pixel Kerr tracer:
 * null-geodesic integration, DIS optical depth tau = int n_N sigma_CC(E) dl,
 * and per-energy-bin transfer I = int j e^{-tau} dl.
 *
 * Plain C: structs of doubles and int32 pairs, caller-owned buffers, no C++
 * types, no exceptions across the boundary.  Every entry point returns an
 * int status (HADROS_OK == 0); the message of the last failure on the calling
 * thread is available from hadros_last_error().
 *
 * The reference exposes this path as C++ value classes plus positional-argv
 * programs; each declaration below names the reference interface it replaces.
 * All compute entry points run on the current HIP device (gfx950).  There is
 * no CPU fallback: without a usable device they fail with HADROS_ERR_DEVICE.
 */
#ifndef HADROS_H
#define HADROS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HADROS_ABI_VERSION 1

/* ---- status codes (reference failure modes, see each entry point) ------- */
enum {
    HADROS_OK = 0,
    HADROS_ERR_ARG = 1,              /* invalid argument / shape               */
    HADROS_ERR_TABLE = 2,            /* sigma table: <2 rows or E not increasing
                                        (src/sigma_table.cpp:42-52)            */
    HADROS_ERR_RANGE = 3,            /* energy outside sigma table
                                        (src/sigma_table.cpp:86-94)            */
    HADROS_ERR_KIND = 4,             /* unknown profile/source/spectrum name
                                        (src/torus_profile.cpp:85-90,
                                         src/radiative_transfer.cpp:108-111,152-155) */
    HADROS_ERR_CACHE_MAGIC = 5,      /* KGEO bad magic
                                        (apps/compute_kerr_image_from_cache.cpp:388-391) */
    HADROS_ERR_CACHE_VERSION = 6,    /* KGEO version != 1 (:393-397)            */
    HADROS_ERR_CACHE_TRUNCATED = 7,  /* EOF mid-ray or npoints < 0 (:672-700)   */
    HADROS_ERR_DEVICE = 8,           /* HIP runtime failure / no gfx950 device  */
    HADROS_ERR_IO = 9,               /* file open/write failure                 */
    HADROS_ERR_MAX_POINTS = 10       /* a traced ray exceeded the point buffer
                                        (apps/compute_kerr_image_cuda.cu:1787-1823) */
};

/* ---- integrator semantics ---------------------------------------------- */
enum {
    /* Reference semantics (src/kerr_geodesic.cpp:194-329, src/kerr_camera.cpp:257-351):
       finite-difference d g^{mu nu}, RKF45 restarted from h every step,
       absolute max-norm error < tol, 50 attempts then RK4, stop at
       r <= r_+ + 1e-3 or (r >= r_max and step > 10), max_steps points. */
    HADROS_MODE_REF = 0,
    /* Error-controlled semantics: analytic Hamiltonian derivatives, RKF45
       with step growth, mixed error |y5-y4| <= tol*(1+|y|), stop at
       r < 1.01 r_+ or (r > r_max and step > 10), max_steps points. */
    HADROS_MODE_CFG = 1
};

enum {
    HADROS_PREC_FP64 = 0,            /* everything FP64                        */
    HADROS_PREC_FP32K = 1            /* FP64 geodesic + FP32 Kahan tau / I     */
};

/* density morphologies, src/torus_profile.cpp:44-91 (include/torus_profile.hpp:6-15) */
enum {
    HADROS_PROFILE_GAUSSIAN = 0,
    HADROS_PROFILE_POWERLAW = 1,
    HADROS_PROFILE_GAUSSIAN_FUNNEL = 2,
    HADROS_PROFILE_POWERLAW_FUNNEL = 3,
    HADROS_PROFILE_GAUSSIAN_ENVELOPE = 4,
    HADROS_PROFILE_POWERLAW_ENVELOPE = 5,
    HADROS_PROFILE_POWERLAW_FUNNEL_ENVELOPE = 6,
    HADROS_PROFILE_COLLAPSAR_NDAF_LIKE = 7
};

/* UHE source morphologies, include/radiative_transfer.hpp:39-45 */
enum {
    HADROS_SOURCE_INNER_RING = 0,
    HADROS_SOURCE_FUNNEL_WALL = 1,
    HADROS_SOURCE_JET_BASE = 2,
    HADROS_SOURCE_SHOCK_LAYER = 3,
    HADROS_SOURCE_DENSITY_WEIGHTED = 4
};

/* UHE spectral models, include/radiative_transfer.hpp:23-27 */
enum {
    HADROS_SPECTRUM_MONOCHROMATIC = 0,
    HADROS_SPECTRUM_POWERLAW = 1,
    HADROS_SPECTRUM_POWERLAW_CUTOFF = 2
};

#define HADROS_MAX_BINS 16
#define HADROS_MAX_SIGMA 1024

/* Camera: replaces KerrCamera(a, r_obs, theta_obs, fov_deg, nx, ny, r_max, h)
   (include/kerr_camera.hpp:10-19) plus the integrator knobs that the
   reference hard-codes (tolerance include/kerr_geodesic.hpp:12,
   max steps src/kerr_camera.cpp:222). */
typedef struct hadros_camera {
    double a_spin;          /* dimensionless spin a                          */
    double r_obs;           /* observer radius [r_g]                         */
    double theta_obs_rad;   /* observer inclination [rad]                    */
    double fov_deg;         /* full field of view [deg]                      */
    int32_t nx, ny;         /* pixels                                        */
    double r_max;           /* escape radius [r_g]                           */
    double h;               /* (initial) affine step                         */
    double tol;             /* REF: absolute error tol; CFG: relative tol    */
    int32_t mode;           /* HADROS_MODE_*                                 */
    int32_t max_steps;      /* max emitted points per ray                    */
} hadros_camera;

/* Density: replaces TorusProfile(...) (include/torus_profile.hpp:19-33). */
typedef struct hadros_medium {
    int32_t kind;           /* HADROS_PROFILE_*                              */
    int32_t reserved;
    double rho0;            /* g/cm^3                                        */
    double r0;              /* r_g                                           */
    double sigma_r;         /* r_g                                           */
    double h_over_r;
    double radial_power;
    double funnel_depletion;
    double funnel_theta_rad;
    double envelope_rho0;   /* g/cm^3                                        */
    double envelope_alpha;
    double r_min, r_max;    /* r_g                                           */
    double rho_floor;       /* g/cm^3                                        */
} hadros_medium;

/* UHE source: replaces UHESourceParams (include/radiative_transfer.hpp:47-64). */
typedef struct hadros_source {
    int32_t kind;           /* HADROS_SOURCE_*                               */
    int32_t reserved;
    double r_center;
    double sigma_r;
    double theta_width_rad;
    double powerlaw;
    double emax_GeV;
    double norm;
    double funnel_theta_rad;
    double density_q;
    double radial_s;
    double rho_ref;
    double cutoff_min, cutoff_max;
    double gradient_dr;
    double gradient_dtheta_rad;
} hadros_source;

/* UHE spectrum: replaces UHESpectralParams (include/radiative_transfer.hpp:29-37). */
typedef struct hadros_spectrum {
    int32_t kind;           /* HADROS_SPECTRUM_*                             */
    int32_t n_bins;
    double gamma;
    double ecut_GeV;
    double e_min_GeV, e_max_GeV;
} hadros_spectrum;

/* MeV thermal channel: replaces MeVThermalParams (include/radiative_transfer.hpp:66-106)
   with mev_neutrino::MeVPhysicsParams (include/mev_neutrino_physics.hpp:42-80).  With
   enabled = 0 the MeV outputs keep the reference's "not computed" values
   (tau 0, P 1, I 0, r_ns -1, leakage 1). */
enum {
    HADROS_MEV_MODEL_TOY = 0, HADROS_MEV_MODEL_PHYSICAL = 1
};
enum {
    HADROS_MEV_NU_E = 0, HADROS_MEV_ANTI_NU_E = 1, HADROS_MEV_NU_X = 2
};
enum {
    HADROS_MEV_T_CONSTANT = 0, HADROS_MEV_T_INNER_HOT_TORUS = 1, HADROS_MEV_T_RADIAL_POWERLAW = 2,
    HADROS_MEV_T_TORUS_PLUS_COOL_ENVELOPE = 3, HADROS_MEV_T_COLLAPSAR_INNER_HOT = 4
};
enum {
    HADROS_MEV_YE_CONSTANT = 0, HADROS_MEV_YE_NEUTRON_RICH_TORUS = 1, HADROS_MEV_YE_FUNNEL_PROTON_RICH = 2,
    HADROS_MEV_YE_TORUS_ENVELOPE_CONTRAST = 3, HADROS_MEV_YE_COLLAPSAR_NEUTRON_RICH = 4
};
enum {
    HADROS_MEV_SPECTRAL_MONOCHROMATIC = 0, HADROS_MEV_SPECTRAL_FERMI_DIRAC_BAND = 1
};
/* bits of hadros_mev.processes (all set by default, except DEGENERACY) */
enum {
    HADROS_MEV_URCA = 1 << 0, HADROS_MEV_PAIR = 1 << 1, HADROS_MEV_BREMS = 1 << 2,
    HADROS_MEV_ABSORPTION = 1 << 3, HADROS_MEV_SCATTERING = 1 << 4, HADROS_MEV_DEGENERACY = 1 << 5,
    HADROS_MEV_ABS_N = 1 << 6, HADROS_MEV_ABS_P = 1 << 7, HADROS_MEV_SCAT_N = 1 << 8,
    HADROS_MEV_SCAT_P = 1 << 9, HADROS_MEV_SCAT_E = 1 << 10
};
typedef struct hadros_mev {
    int32_t enabled;
    int32_t model;            /* HADROS_MEV_MODEL_*                             */
    int32_t flavor;           /* HADROS_MEV_NU_E / _ANTI_NU_E / _NU_X           */
    int32_t thermal_profile;  /* HADROS_MEV_T_*                                 */
    int32_t ye_profile;       /* HADROS_MEV_YE_*                                */
    int32_t spectral_mode;    /* HADROS_MEV_SPECTRAL_*                          */
    int32_t n_bins;           /* Fermi-Dirac band bins                          */
    uint32_t processes;       /* HADROS_MEV_* bits                              */
    double enu_obs_MeV;
    double norm;
    double sigma_abs0_cm2, sigma_scat0_cm2, sigma_scat_e0_cm2;
    double neutrinosphere_tau;
    double T0_MeV, T_floor_MeV, T_power;
    double Ye_torus, Ye_funnel, Ye_envelope, Ye_floor, Ye_ceil;
    double E_min_MeV, E_max_MeV;
} hadros_mev;

/* Per-ray MeV outputs (the .dat columns tau_mev P_surv_mev I_obs_mev
   r_neutrinosphere_rg leakage_mev, src/radiative_transfer.cpp:763-769).
   Caller-owned host arrays, any may be NULL. */
typedef struct hadros_mev_out {
    double* tau;
    double* p_surv;
    double* intensity;
    double* r_neutrinosphere;
    double* leakage;
} hadros_mev_out;

/* Everything the transfer integral needs besides the ray and the sigma table:
   the argv block of apps/compute_kerr_image_from_cache.cpp:57-205 that is on
   the UHE path. */
typedef struct hadros_scene {
    hadros_medium medium;
    hadros_source source;
    hadros_spectrum spectrum;
    double enu_obs_GeV;     /* observed energy (monochromatic)               */
    double m_bh_msun;       /* black-hole mass [M_sun] -> r_g in cm           */
} hadros_scene;

/* Per-pixel outputs (caller-owned, host memory unless stated otherwise).
   Pixel index k = i*ny + j (ray_id of apps/compute_kerr_geodesics.cpp:110).
   tau/p_surv/intensity are the aggregates written to the image rows
   (src/radiative_transfer.cpp:763-769); bin_* hold the per-bin values
   (k*n_bins + b), bins skipped by the reference (w <= 0) hold tau=0, P=1, I=0.
   Any pointer may be NULL to skip that output. */
typedef struct hadros_image_out {
    double* tau;
    double* p_surv;
    double* intensity;
    int32_t* captured;
    int32_t* npoints;       /* points emitted along the ray                  */
    int32_t* flags;         /* bit0: ray exhausted max_steps                 */
    double* alpha;          /* tan-plane u (alpha_rg of the reference)       */
    double* beta;           /* tan-plane v (beta_rg)                         */
    double* bin_tau;
    double* bin_p;
    double* bin_i;
} hadros_image_out;

#define HADROS_FLAG_EXHAUSTED 1

typedef struct hadros_stats {
    double kernel_ms;       /* device time of the compute kernels (HIP events) */
    double total_ms;        /* wall time of the call                          */
    int64_t total_points;   /* accepted steps over all rays                   */
    int64_t total_attempts; /* RKF45 attempts over all rays                   */
    int32_t n_captured;
    int32_t n_exhausted;
    int32_t n_bins;         /* bins actually integrated                       */
    int32_t block_size;     /* workgroup size used                            */
} hadros_stats;

/* Geodesic cache (KGEO content) as structure-of-arrays.
   Replaces RayPath/PathPoint (include/ray.hpp:6-29) streamed from the KGEO
   file (apps/compute_kerr_image_from_cache.cpp:660-711).  Ray k owns points
   [offset[k], offset[k+1]).  Only r, theta, dl, redshift are read by the
   transfer; x, y, z are not needed. */
typedef struct hadros_cache_view {
    int32_t n_rays;
    int32_t reserved;
    double a_bh;            /* spin stored in the cache header               */
    const int64_t* offset;  /* [n_rays+1]                                    */
    const double* r;        /* [n_points]                                    */
    const double* theta;
    const double* dl;
    const double* redshift;
} hadros_cache_view;

/* ---- library / errors --------------------------------------------------- */
const char* hadros_last_error(void);
int hadros_abi_version(void);
/* Name and compute-unit count of the current device. */
int hadros_device_info(char* name, int name_cap, int* cu_count, int* gcn_arch);

/* ---- defaults and name parsing ------------------------------------------ */
/* argv defaults of apps/compute_kerr_geodesics.cpp:26-34 (ref mode, tol 1e-6
   as KerrCamera passes no tolerance: include/kerr_geodesic.hpp:12). */
void hadros_camera_defaults(hadros_camera* cam);
/* argv defaults of apps/compute_kerr_image_from_cache.cpp:33-55. */
void hadros_scene_defaults(hadros_scene* scene);
/* TorusProfile::parse_profile_kind (src/torus_profile.cpp:44-91). */
int hadros_parse_profile(const char* name, int32_t* kind);
/* parse_uhe_source_model (src/radiative_transfer.cpp:80-112). */
int hadros_parse_source(const char* name, int32_t* kind);
/* parse_uhe_spectral_model (src/radiative_transfer.cpp:132-156). */
int hadros_parse_spectrum(const char* name, int32_t* kind);
/* MeVThermalParams defaults (include/radiative_transfer.hpp:66-106), enabled = 1. */
void hadros_mev_defaults(hadros_mev* mev);
/* parse_mev_model / _flavor / _thermal_profile / _ye_profile / _spectral_mode
   (src/mev_neutrino_physics.cpp:36-233): the reference's names and aliases;
   HADROS_ERR_KIND with the reference's message otherwise. */
int hadros_parse_mev_model(const char* name, int32_t* kind);
int hadros_parse_mev_flavor(const char* name, int32_t* kind);
int hadros_parse_mev_thermal(const char* name, int32_t* kind);
int hadros_parse_mev_ye(const char* name, int32_t* kind);
int hadros_parse_mev_spectral(const char* name, int32_t* kind);
const char* hadros_profile_name(int32_t kind);
const char* hadros_source_name(int32_t kind);
const char* hadros_spectrum_name(int32_t kind);

/* ---- sigma table -------------------------------------------------------- */
/* SigmaTable::SigmaTable (src/sigma_table.cpp:9-53): parse a 3-column
   "E_GeV sigma_GeV^-2 sigma_cm2" file, skip '#', blank, malformed and
   non-positive rows; require >= 2 rows and strictly increasing E.
   Writes up to cap rows; *n receives the row count. */
int hadros_sigma_load(const char* path, double* e_gev, double* sigma_gev2,
                      double* sigma_cm2, int cap, int* n);
/* Validation part of the same constructor for in-memory tables. */
int hadros_sigma_check(const double* e_gev, const double* sigma_cm2, int n);

/* ---- the hot path ------------------------------------------------------- */
/* Fused trace + tau + per-bin transfer for a whole camera, one ray per lane:
   KerrCamera::trace_pixel (src/kerr_camera.cpp:257-351) followed by
   radiative_transfer::integrate_kerr_ray_spectral
   (src/radiative_transfer.cpp:687-780) without materialising the path --
   the loop of apps/compute_kerr_geodesics.cpp:106-140 +
   apps/compute_kerr_image_from_cache.cpp:605-658 in one launch, and the
   direct mode of apps/compute_kerr_image_cuda.cu:2239-2606. */
int hadros_trace_rt(const hadros_camera* cam, const hadros_scene* scene,
                    const double* sigma_e_gev, const double* sigma_cm2, int n_sigma,
                    int precision, hadros_image_out* out, hadros_stats* stats);

/* Allocation-free variant of hadros_trace_rt.  It has no counterpart in the
   reference: it is for callers that cannot allocate device memory (code under
   HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
   It runs the fused kernel (the kernels of hadros_tracer_run with
   HADROS_PATH_FUSED, so its results equal that path's bit for bit) and calls
   no hipMalloc, hipFree or host-to-device copy: the scene block, the sigma
   table and the counters live in a static device array of
   HADROS_NOALLOC_SCRATCH_BYTES per GPU, kept until the process ends, and
   reach it as kernel arguments.
   - Every non-NULL member of `out` is DEVICE memory of the current device,
     sized as for hadros_trace_rt (nx*ny per pixel column, nx*ny*n_bins per
     bin column, where n_bins is the number of bins the scene integrates).
   - `counters`: NULL, or 4 int64 of device memory that receive the points,
     attempts, captured and exhausted counts (cleared by the call).
   - Asynchronous on `stream` (a hipStream_t or NULL); no host synchronisation.
   - Calls that overlap on one GPU (on different streams or threads) share
     the scratch and race: order them on one stream.
   - Host inputs are read before the call returns.
   - The scratch holds tables of up to HADROS_MAX_SIGMA rows; a larger
     layout is rejected with HADROS_ERR_ARG. */
#define HADROS_NOALLOC_SCRATCH_BYTES 32768
int hadros_trace_rt_noalloc(const hadros_camera* cam, const hadros_scene* scene,
                            const double* sigma_e_gev, const double* sigma_cm2, int n_sigma,
                            int precision, const hadros_image_out* out, int64_t* counters,
                            void* stream);

/* Re-usable plan of hadros_trace_rt: device buffers are allocated once,
   run() only launches (asynchronously on `stream`, a hipStream_t or NULL),
   fetch() synchronises and copies results to host. */
typedef struct hadros_tracer hadros_tracer;
int hadros_tracer_create(const hadros_camera* cam, const hadros_scene* scene,
                         const double* sigma_e_gev, const double* sigma_cm2, int n_sigma,
                         int precision, hadros_tracer** out);
/* workgroup size of the tracer kernels: 0 = the default of the running form
   (128 fused, 64 split), else 64, 128, 192 or 256. */
int hadros_tracer_set_block(hadros_tracer* t, int block_size);
/* How a plan runs (no reference counterpart; results agree to rounding):
   HADROS_PATH_FUSED: one kernel integrates the transfer at every path point
   (the form the batch kernel shares bit for bit);
   HADROS_PATH_SPLIT: a geodesic kernel stores the path points in a device
   chunk pool, then the cache-replay kernel integrates them (fewer registers
   per kernel, more resident wavefronts);
   HADROS_PATH_AUTO (default): split in CFG mode while the pool stays under
   8 GB, fused otherwise.  The split form's pool grows on demand. */
enum { HADROS_PATH_AUTO = 0, HADROS_PATH_FUSED = 1, HADROS_PATH_SPLIT = 2 };
int hadros_tracer_set_path(hadros_tracer* t, int path);
/* Segments of the split form (no reference counterpart; scheduling only, the
   results are bit for bit the same for every count).  The camera's rows of
   8x8 tiles are cut into n_segments contiguous bands; band 0 runs on the
   caller's stream, band i > 0 on a stream of the plan's own (default
   priority, created on first use) that waits for the caller's stream at the
   start of the run and that the caller's stream waits for at its end.  All
   bands' geodesic launches are queued first, then each band's transfer
   launch behind its own geodesic launch: the transfer of the bands whose
   rays finish early runs while the longest rays are still traced.
   0 (default) = automatic: on cameras of at least 16 tile rows, four bands
   at 0, 3/8, 1/2 and 5/8 of the tile rows, so that the two middle bands,
   where a camera aimed at the hole has its longest rays (near the photon
   orbit), hold a quarter of the rows and the outer bands' transfer runs
   under their tail; on smaller cameras up to HADROS_TRACER_MAX_SEGMENTS
   equal bands of at least four tile rows each.  n >= 1 = n equal bands
   (1 = one geodesic launch, then one transfer launch); counts above the
   camera's tile rows are cut to them. */
#define HADROS_TRACER_MAX_SEGMENTS 4
int hadros_tracer_set_segments(hadros_tracer* t, int n_segments);
int hadros_tracer_run(hadros_tracer* t, void* stream);
/* Device time (ms) of the last run's kernel, measured with HIP events on the
   launch stream; waits for the run to finish. */
int hadros_tracer_last_ms(hadros_tracer* t, double* ms);
/* The last run's time split by kernel (HIP events on the launch streams).
   Split form: geodesic_ms spans the run's start to the end of the last
   band's geodesic launch, transfer_ms the first band's transfer start to the
   run's end; with one segment these are the two launches' own durations,
   with more they overlap.  The fused form puts its one kernel in
   geodesic_ms and 0 in transfer_ms.  No reference counterpart (measurement
   hook). */
int hadros_tracer_kernel_ms(hadros_tracer* t, double* geodesic_ms, double* transfer_ms);
int hadros_tracer_fetch(hadros_tracer* t, hadros_image_out* out, hadros_stats* stats);
/* MeV channel of integrate_kerr_ray (src/radiative_transfer.cpp:500-640) for the
   plan's next runs (NULL or enabled = 0 turns it off).  It runs in the split
   form's transfer kernel, so a plan with MeV runs split (HADROS_PATH_FUSED
   is then an error).  hadros_tracer_fetch_mev copies the last run's columns. */
int hadros_tracer_set_mev(hadros_tracer* t, const hadros_mev* mev);
int hadros_tracer_fetch_mev(hadros_tracer* t, hadros_mev_out* out);
int hadros_tracer_destroy(hadros_tracer* t);

/* optical_depth::tau_along_ray and survival_probability
   (src/optical_depth.cpp:50-91) for every ray of a cache: tau = sum over the
   ray's points of rho/m_u sigma(E_inf z) dl r_g, with r_g of 3 M_sun whatever
   the black hole (:83-85), every point counted (no horizon cut) and points
   whose local energy E_inf z leaves the table skipped; P = exp(-tau).  tau
   and p_surv receive n_rays values each (either may be NULL). */
int hadros_tau_along_ray(const hadros_cache_view* cache, const hadros_medium* medium,
                         const double* sigma_e_gev, const double* sigma_cm2, int n_sigma,
                         double enu_inf_GeV, double* tau, double* p_surv);

/* Cache replay: radiative_transfer::integrate_kerr_ray_spectral over every ray
   of a KGEO cache (apps/compute_kerr_image_from_cache.cpp:605-658;
   apps/compute_kerr_image_cuda.cu:1293-1335).  Host arrays in, host results
   out; rays with zero points give tau=0, P=1, I=0 (:611-618). */
int hadros_rt_from_cache(const hadros_cache_view* cache, const hadros_scene* scene,
                         const double* sigma_e_gev, const double* sigma_cm2, int n_sigma,
                         int precision, hadros_image_out* out, hadros_stats* stats);

/* Re-usable plan of the replay with the cache resident in HBM. */
typedef struct hadros_replay hadros_replay;
int hadros_replay_create(const hadros_cache_view* cache, const hadros_scene* scene,
                         const double* sigma_e_gev, const double* sigma_cm2, int n_sigma,
                         int precision, hadros_replay** out);
int hadros_replay_run(hadros_replay* p, void* stream);
int hadros_replay_last_ms(hadros_replay* p, double* ms);
int hadros_replay_fetch(hadros_replay* p, hadros_image_out* out, hadros_stats* stats);
/* MeV channel of the replay (as hadros_tracer_set_mev / _fetch_mev). */
int hadros_replay_set_mev(hadros_replay* p, const hadros_mev* mev);
int hadros_replay_fetch_mev(hadros_replay* p, hadros_mev_out* out);
int hadros_replay_destroy(hadros_replay* p);

/* Geodesic cache on the GPU: the --geodesic-cache mode of
   apps/compute_kerr_image_cuda.cu:1478-1535,1624-1826 and the per-pixel
   output of apps/compute_kerr_geodesics.cpp:106-140.  Traces every pixel and
   writes its PathPoints {r, theta, x, y, z, dl, redshift} (7 doubles each,
   include/ray.hpp:6-14) into `points` at row offset[k], where offset is an
   exclusive prefix sum of the per-ray point counts computed on the device.
   Call first with points == NULL to obtain counts/offsets only (offset has
   nx*ny+1 entries; offset[nx*ny] is the total), then with a buffer of
   offset[nx*ny]*7 doubles.  captured[k] as in BinaryRayHeader. */
int hadros_trace_cache(const hadros_camera* cam, int64_t* offset, int32_t* captured,
                       double* points, int64_t points_cap);
/* Same for pixels [k0, k1) only (offset has k1-k0+1 entries, relative to
   the first pixel of the range), so that large images stream in batches
   as the reference's per-ray file writes do. */
int hadros_trace_cache_range(const hadros_camera* cam, int64_t k0, int64_t k1, int64_t* offset,
                             int32_t* captured, double* points, int64_t points_cap);

/* Batched tracer: many cameras (each with its own spin, inclination,
   resolution, medium, source and spectrum; one integrator mode and one sigma
   table for all) traced by one persistent kernel that pulls rays from a
   global queue and refills finished lanes by wavefront ballot.  Each image
   equals the plain hadros_trace_rt of its camera, run in the same form
   (hadros_batch_set_path / hadros_tracer_set_path), bit for bit.  There is no
   single reference interface: the reference runs one camera per process
   (apps/compute_kerr_image_cuda.cu:2418-2483 tiles one image; the parameter
   scans of scripts/run_robustness_scans.py:103-165 launch one process per
   camera). */
typedef struct hadros_batch hadros_batch;
int hadros_batch_create(const hadros_camera* cams, const hadros_scene* scenes, int n_cams,
                        const double* sigma_e_gev, const double* sigma_cm2, int n_sigma,
                        int precision, hadros_batch** out);
/* Same, tracing only the pixel sub-lattice i, j = offset + stride * (0, 1, ...)
   of every camera (e.g. stride 10, offset 5: 1 % of a 256x256 image, for
   spot checks); outputs are indexed by the sub-lattice pixel p * sub_ny + q. */
int hadros_batch_create_sub(const hadros_camera* cams, const hadros_scene* scenes, int n_cams,
                            const double* sigma_e_gev, const double* sigma_cm2, int n_sigma,
                            int precision, int stride, int offset, hadros_batch** out);
/* all rays' outputs in global ray order; bins are [ray x n_bins_max] with
   n_bins_max the largest bin count of the batch's scenes */
int hadros_batch_fetch_all(hadros_batch* b, hadros_image_out* out);
/* as hadros_batch_fetch_all, but every non-NULL pointer of `out` is DEVICE
   memory of the current device: asynchronous device-to-device copies on
   `stream`, ordered after the batch's last run (the library's own product
   surface for consumers that keep the outputs on the GPU, such as the
   surrogate scan's spot checks; no reference counterpart). */
int hadros_batch_fetch_all_device(hadros_batch* b, const hadros_image_out* out, void* stream);
/* workgroup size (64..256, multiple of 64) and RKF45 attempts between refills */
int hadros_batch_set_tuning(hadros_batch* b, int block, int steps_per_chunk);
/* times every (block in {64,128,256}) x (steps_per_chunk in {1,4,16,64}) on
   the first eighth of the rays, re-times the three fastest twice more and
   keeps the pair with the least total time */
int hadros_batch_autotune(hadros_batch* b, int* block, int* steps_per_chunk);
/* How the batch runs (no reference counterpart): HADROS_PATH_FUSED, the
   persistent kernel integrating the transfer at every point; HADROS_PATH_SPLIT,
   sub-batches of rays whose persistent geodesic kernel stores the path points
   in a device chunk pool, each followed by the cache-replay kernel over those
   rays; HADROS_PATH_AUTO (default): split in CFG mode, fused in REF mode.
   The split form makes
   hadros_batch_run's results final only once a fetch, stats or last_ms call
   has checked the pool (it grows the pool and runs again when rays ran out of
   chunks). */
int hadros_batch_set_path(hadros_batch* b, int path);
/* Bound (bytes) of the split form's chunk pool; the rays per sub-batch follow
   from it.  0 = the default, 64 GB. */
int hadros_batch_set_pool(hadros_batch* b, double max_bytes);
int hadros_batch_run(hadros_batch* b, void* stream);
int hadros_batch_last_ms(hadros_batch* b, double* ms);
/* outputs of camera `cam` (pixel k = i*ny + j of that camera) */
int hadros_batch_fetch(hadros_batch* b, int cam, hadros_image_out* out);
int hadros_batch_stats(hadros_batch* b, hadros_stats* stats);
int hadros_batch_info(hadros_batch* b, int64_t* n_rays, int* block, int* steps_per_chunk, int* grid);
int hadros_batch_destroy(hadros_batch* b);

/* ---- learned surrogate --------------------------------------------------
   A multilayer perceptron n_in -> width x depth (SiLU) -> n_out with BF16
   weights and activations on MFMA and FP32 accumulation, trained with Adam
   by the library's own forward/backward kernels.  It maps per-ray features
   (alpha, beta, spin, cos inclination, profile one-hot, log10 rho0, slope)
   to log10 tau and survival per energy bin.  No reference counterpart: the
   reference has no learned component; its parameter scans re-run the tracer
   for every point (scripts/run_robustness_scans.py:103-165).
   x, y are DEVICE pointers to row-major FP32 [n x n_in] / [n x n_out]. */
typedef struct hadros_mlp hadros_mlp;
/* n_in <= 16, width a multiple of 16; train_batch / infer_batch bound the
   rows of one training step / one inference chunk.  He-uniform init from seed. */
int hadros_mlp_create(int n_in, int width, int depth, int n_out, int train_batch,
                      int infer_batch, uint64_t seed, hadros_mlp** out);
int hadros_mlp_n_params(hadros_mlp* m, int64_t* n);
/* parameters as FP32 in the order W1 (n_in x width, row-major), b1, ..., W_L, b_L */
int hadros_mlp_get_params(hadros_mlp* m, float* host, int64_t n);
int hadros_mlp_set_params(hadros_mlp* m, const float* host, int64_t n);
/* feature and target standardisation (host arrays of n_in / n_out) */
int hadros_mlp_set_norm(hadros_mlp* m, const float* x_mean, const float* x_std,
                        const float* y_mean, const float* y_std);
/* y = net(x) for any n, asynchronous on stream.  Width 256 with n_out <= 16
   runs the fused kernel (all layers of a 128-row tile in one workgroup,
   activations in LDS); other shapes run one GEMM per layer in infer_batch
   chunks. */
int hadros_mlp_forward(hadros_mlp* m, const float* x, int64_t n, float* y, void* stream);
/* survival form of the fused forward: outputs [0, n_out/2) are log10 tau per
   energy bin; p[row * n_out/2 + b] = exp(-10^y_b).  Needs the fused shape. */
int hadros_mlp_survival(hadros_mlp* m, const float* x, int64_t n, float* p, void* stream);
/* 0: fused forward where the shape allows (default); 1: per-layer GEMMs */
int hadros_mlp_set_infer_path(hadros_mlp* m, int path);
/* gradient of the mean squared error on standardised targets w.r.t. every
   parameter (host FP32, get_params order) and the loss, without updating */
int hadros_mlp_grad(hadros_mlp* m, const float* x, const float* y, int64_t n, float* grad, float* loss);
/* one Adam step (beta 0.9 / 0.999, eps 1e-8) on n <= train_batch rows */
int hadros_mlp_train_step(hadros_mlp* m, const float* x, const float* y, int64_t n, float lr, float* loss);
int hadros_mlp_destroy(hadros_mlp* m);

/* ---- point-wise device evaluations (parity hooks for each physics term) -- */
/* Opacity surface of a density model for one neutrino energy: replaces the
   loop of apps/extract_opacity_surface.cpp:127-206.  For theta_k =
   pi k / (ntheta - 1), k < ntheta: tau(r) = integral of rho/m_u sigma(E) dl
   from r to m->r_max on the uniform grid of nr nodes over [m->r_min,
   m->r_max] (trapezoidal density, :165-174); tau_total = tau(r_min); r_tau =
   the innermost radius where tau falls through tau_surface, linear in tau
   between nodes, or -1 (:176-190); classification 0/1/2/3 for tau_total
   < 0.1 / < 1 / < 3 / else (:55-68); crossing 1 if r_tau was found.  Host
   arrays of ntheta.  HADROS_ERR_ARG for ntheta < 2, nr < 4 or r_max <= r_min
   (:95-98), HADROS_ERR_RANGE for E outside the table (:101-104). */
int hadros_opacity_surface(const hadros_medium* m, const double* sigma_e_gev, const double* sigma_cm2,
                           int n_sigma, double enu_GeV, double m_bh_msun, double tau_surface, int nr,
                           int ntheta, double* theta, double* r_tau, double* tau_total,
                           double* classification, int* crossing);

/* TorusProfile::rho (src/torus_profile.cpp:261-264), n points. */
int hadros_eval_rho(const hadros_medium* m, const double* r, const double* th,
                    double* rho, int64_t n);
/* SigmaTable::sigma_cm2 (src/sigma_table.cpp:55-58,80-124); out-of-range
   energies give NaN (the reference throws HADROS_ERR_RANGE-equivalent). */
int hadros_eval_sigma(const double* sigma_e_gev, const double* sigma_cm2, int n_sigma,
                      const double* e, double* out, int64_t n);
/* radiative_transfer::emissivity_uhe (src/radiative_transfer.cpp:369-395)
   with the spectral model of `scene`. */
int hadros_eval_emissivity(const hadros_scene* scene, const double* r, const double* th,
                           const double* e_local, double* out, int64_t n);
/* The MeV channel at points (r, theta) of the scene's medium, for a path
   point whose frequency ratio is g = 1 / redshift: out is n x 4 =
   (T_MeV, Ye, kappa_cm^-1, j) as mev_temperature_profile_MeV, mev_ye_profile,
   and the opacity / emissivity of integrate_kerr_ray's MeV block at
   E_local = E_obs / g, or the Fermi-Dirac band's kappa average and j sum
   (src/mev_neutrino_physics.cpp:235-771, src/radiative_transfer.cpp:500-600). */
int hadros_eval_mev(const hadros_scene* scene, const hadros_mev* mev, const double* r, const double* th,
                    const double* g, double* out, int64_t n);
/* KerrGeodesic::rhs (src/kerr_geodesic.cpp:83-135) for mode REF (finite
   differences) or the analytic derivative for mode CFG; y and dy are n x 8
   (t, r, theta, phi, p_t, p_r, p_theta, p_phi). */
int hadros_eval_rhs(double a_spin, int mode, const double* y, double* dy, int64_t n);
/* One step of the integrator forms on n states y (n x 8): out is n x 11, the
   4th- and 5th-order results (r, theta, phi, p_r, p_theta each) and a flag.
   kind 0: KerrGeodesic::step_rk4 (src/kerr_geodesic.cpp:137-192), both
   results equal; 1: one Fehlberg attempt of step_adaptive
   (src/kerr_geodesic.cpp:239-308) on the REF right-hand side; 2: the same
   attempt on the CFG right-hand side; 3: the CFG attempt with the later
   stages' sin/cos by angle addition, as the tracing kernels take it (flag 0
   when a stage's theta offset leaves that form's range; the kernels then
   redo the attempt as kind 2). */
int hadros_eval_step(double a_spin, int kind, double h, const double* y, double* out, int64_t n);
/* The per-point and per-(point, bin) terms of integrate_kerr_ray_spectral
   (src/radiative_transfer.cpp:640-685) as the transfer kernel of the split
   tracer, the cache replay and the batch computes them, at n points
   (r, theta, dl, redshift) of a ray of spin a_bh: out is n x
   HADROS_EVAL_TRANSFER_COLS = ok (0 inside r_h + 1e-3), z, ln z, g^3 A dl_cm,
   n_N dl_cm, rho, then per bin dtau_b = n_N sigma(E_b) dl_cm (0 where rho <= 0
   or E_b = E_mid,b z is outside the table), then per bin the emission term
   before absorption g^3 j(E_b) dl_cm -- for HADROS_PREC_FP32K its natural
   log, -inf where the point does not emit.  Bins past the scene's count are 0. */
#define HADROS_EVAL_TRANSFER_COLS (6 + 2 * HADROS_MAX_BINS)
int hadros_eval_transfer_point(const hadros_scene* scene, const double* sigma_e_gev, const double* sigma_cm2,
                               int n_sigma, double a_bh, int precision, const double* r, const double* th,
                               const double* dl, const double* redshift, double* out, int64_t n);
/* Measurement helper; it has no counterpart in the reference.  Streaming
   kernels of known HBM traffic, for scaling a profiler's per-kernel byte
   counters (FETCH_SIZE, WRITE_SIZE): kind 0 (hd::k_stream_read) reads the n
   doubles of buf once and writes HADROS_STREAM_GRID partial sums (their total
   is the sum of buf) into partial; kind 1 (hd::k_stream_fill) writes value to
   the n doubles of buf.  Device pointers, buf 16-byte aligned; asynchronous
   on `stream` (a hipStream_t or NULL). */
#define HADROS_STREAM_GRID 2048
int hadros_stream_probe(int kind, double* buf, int64_t n, double* partial, double value, void* stream);
/* KerrMetric::metric / inverse_metric (src/kerr_metric.cpp:34-116):
   5 distinct components each (tt, tphi, rr, thth, phph) per point, n x 10. */
int hadros_eval_metric(double a_spin, const double* r, const double* th,
                       double* out, int64_t n);

#ifdef __cplusplus
}
#endif

#endif /* HADROS_H */
