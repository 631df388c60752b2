/*
 * hadros_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU FP64 restatement of the HADROS per-pixel path (Kerr geodesics, DIS
 * optical depth, UHE transfer) used as the correctness oracle for the HIP
 * library.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may load it, and only as the checker or the timed CPU baseline; the
 * product (libhadros, the CLIs) never links or calls it.
 *
 * Each function follows the reference source cited beside it, with the same
 * operation order so that REF mode reproduces the reference bit for bit when
 * both are built with g++ -O3 -ffp-contract=off on x86-64.
 *
 * Pinning (what checks each part; details in tests/):
 *   metric / inverse metric / lapse / omega / horizon:
 *       known answers printed by the reference's apps/test_kerr_metric.cpp
 *       (tests/golden/known_answers.json, pointwise.npz) + g.g^-1 = 1.
 *   rhs (REF, finite differences): reference harness vectors (bit-exact) and
 *       the analytic derivative (agreement to FD truncation error).
 *   rhs (CFG, analytic): central differences of the Hamiltonian, H = 0 and
 *       conserved p_t, p_phi, Carter Q along traced rays.
 *   step / trace_pixel (REF): KGEO caches written by the compiled reference
 *       (per-ray point counts, captured flags, sampled points bit-exact).
 *   trace (CFG): invariants above, a = 0 shadow edge b = sqrt(27), and
 *       convergence to the REF path as both tolerances shrink.
 *   sigma table: sigma = A E^B tables reproduced to 1e-13 (exact log-log),
 *       sigma_GBW(1e9) known answer, reference harness vectors.
 *   rho (8 kinds), source weights (5), spectral weights (3), emissivity:
 *       reference harness vectors on random (r, theta, E) grids, bit-exact.
 *   integrate_kerr_ray(_spectral): image rows written by the compiled
 *       reference's compute_kerr_image_from_cache on reference caches.
 */
#ifndef HADROS_ORACLE_H
#define HADROS_ORACLE_H

#include "../include/hadros.h"

#ifdef __cplusplus
extern "C" {
#endif

/* KerrMetric (src/kerr_metric.cpp). out: tt, tphi, rr, thth, phph */
void or_metric(double a, double r, double th, double out[5]);
void or_inverse_metric(double a, double r, double th, double out[5]);
double or_lapse(double a, double r, double th);
double or_omega(double a, double r, double th);
double or_horizon(double a);

/* KerrGeodesic::rhs (src/kerr_geodesic.cpp:83-135) in REF mode, or the
   analytic Hamiltonian derivative in CFG mode.  y, dy: 8 doubles. */
void or_rhs(double a, int mode, const double y[8], double dy[8]);
double or_hamiltonian(double a, const double y[8]);

/* One accepted step in REF mode: KerrGeodesic::step_adaptive
   (src/kerr_geodesic.cpp:194-329). */
void or_step_ref(double a, double h, double tol, double y[8]);

/* KerrGeodesic::step_rk4 (src/kerr_geodesic.cpp:137-192): one classic RK4
   step of size h on the REF right-hand side (the fallback of step_adaptive
   after 50 rejected attempts, :328). */
void or_step_rk4(double a, double h, double y[8]);

/* One Fehlberg 4(5) attempt of size h (tableau of src/kerr_geodesic.cpp:239-308)
   on the REF or CFG right-hand side: the 4th- and 5th-order results. */
void or_rkf45_attempt(double a, int mode, double h, const double y[8], double y4[8], double y5[8]);

/* optical_depth::tau_along_ray (src/optical_depth.cpp:56-91) over every ray of
   a cache, with survival_probability (:50-53): tau_out, p_out per ray. */
int or_tau_along_ray(const hadros_cache_view* c, const hadros_medium* m, const double* sig_e, const double* sig_s,
                     int n_sig, double enu_inf, double* tau_out, double* p_out);

/* The state (t, r, theta, phi, p_t, p_r, p_theta, p_phi) at every emitted
   point of or_trace_pixel's ray; returns the number of points. */
int64_t or_trace_states(const hadros_camera* cam, int i, int j, double* states, int64_t cap);

/* KerrCamera::initial_state (src/kerr_camera.cpp:110-186): y0 of pixel (i,j). */
void or_initial_state(const hadros_camera* cam, int i, int j, double y[8]);

/* KerrCamera::trace_pixel (src/kerr_camera.cpp:188-283) in cam->mode.
   Writes up to cap points of 7 doubles {r, th, x, y, z, dl, redshift};
   returns the number of points emitted (may exceed cap: then only cap are
   stored).  *captured, *exhausted, *attempts receive the ray's status. */
int64_t or_trace_pixel(const hadros_camera* cam, int i, int j, double* pts, int64_t cap,
                       int* captured, int* exhausted, int64_t* attempts);

/* sigma table log-log interpolation (src/sigma_table.cpp:80-124).
   Returns NaN for energies the reference would throw on. */
double or_sigma(const double* e, const double* s, int n, double x);

/* TorusProfile::rho (src/torus_profile.cpp:261-264). */
double or_rho(const hadros_medium* m, double r, double th);
/* radiative_transfer::uhe_spectral_weight (src/radiative_transfer.cpp:172-207). */
double or_spectral_weight(const hadros_scene* s, double e);
/* radiative_transfer::uhe_source_spatial_weight (:209-333). */
double or_source_weight(const hadros_scene* s, double r, double th);
/* radiative_transfer::emissivity_uhe (:369-395). */
double or_emissivity(const hadros_scene* s, double r, double th, double e);

/* Energy bins of integrate_kerr_ray_spectral (:687-780): writes up to
   HADROS_MAX_BINS (E_mid, dE, w) triples of bins that are integrated
   (w > 0 and finite) and returns their count; for a monochromatic spectrum
   one bin (Enu_obs, 1, 1).  Returns -1 if the spectral range is empty. */
int or_bins(const hadros_scene* s, const double* sig_e, int n_sig,
            double* e_mid, double* de, double* w);

/* integrate_kerr_ray_spectral over one ray given as 7-double points.
   out_agg: tau, P, I; bin_*: per-bin tau/P/I for the or_bins() bins. */
void or_integrate_ray(const hadros_scene* s, const double* sig_e, const double* sig_s,
                      int n_sig, double a_bh, const double* pts, int64_t n,
                      double out_agg[3], double* bin_tau, double* bin_p, double* bin_i);

/* Whole camera: trace + transfer per pixel with OpenMP schedule(dynamic)
   over pixels (apps/compute_kerr_geodesics.cpp:106).  Output layout as
   hadros_image_out (pixel k = i*ny + j); per-bin arrays sized
   npix * HADROS_MAX_BINS are indexed k*n_bins + b.
   nthreads <= 0 uses the OpenMP default. */
int or_trace_image(const hadros_camera* cam, const hadros_scene* s,
                   const double* sig_e, const double* sig_s, int n_sig,
                   int nthreads, hadros_image_out* out, hadros_stats* st);

/* Same, for pixels [k0, k1) only (bounded CPU samples). */
int or_trace_image_range(const hadros_camera* cam, const hadros_scene* s,
                         const double* sig_e, const double* sig_s, int n_sig,
                         int nthreads, int64_t k0, int64_t k1,
                         hadros_image_out* out, hadros_stats* st);

/* Transfer over a cache (apps/compute_kerr_image_from_cache.cpp:605-658). */
int or_rt_from_cache(const hadros_cache_view* c, const hadros_scene* s,
                     const double* sig_e, const double* sig_s, int n_sig,
                     int nthreads, hadros_image_out* out);

/* The MeV channel (src/mev_neutrino_physics.cpp, src/radiative_transfer.cpp:500-640).
   or_mev_point: T_MeV, Ye, kappa (cm^-1) and j at (r, theta) for a path point
   with g = 1 / redshift.  or_rt_from_cache_mev: per ray of the cache the MeV
   columns tau_mev, P_surv_mev, I_obs_mev, r_neutrinosphere_rg, leakage_mev
   (mev_out[5 k + c]) as integrate_kerr_ray_spectral reports them. */
void or_mev_point(const hadros_scene* s, const hadros_mev* m, double r, double th, double g, double out[4]);
int or_rt_from_cache_mev(const hadros_cache_view* c, const hadros_scene* s, const double* sig_e, const double* sig_s,
                         int n_sig, const hadros_mev* m, int nthreads, double* mev_out);

/* apps/extract_opacity_surface.cpp:127-206: per polar angle theta_k =
   pi k / (ntheta - 1), the radial optical depth from r to R_max of a nu of
   energy E on the uniform grid r_i = R_min + i dr (nr points, trapezoidal
   density, n = rho / m_u, sigma(E) of the table), its total, the radius where
   it crosses tau_surface (linear in tau between samples, -1 if none), the
   classification 0/1/2/3 (tau < 0.1 / < 1 / < 3 / else) and the crossing flag.
   Returns 0, or 1 for an invalid grid / 2 for E outside the table. */
int or_opacity_surface(const hadros_medium* m, const double* sig_e, const double* sig_s, int n_sig,
                       double enu_GeV, double m_bh_msun, double tau_surface, int nr, int ntheta,
                       double* theta, double* r_tau, double* tau_total, double* cls, int* crossing);

#ifdef __cplusplus
}
#endif

#endif
