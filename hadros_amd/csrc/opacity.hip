// opacity.hip -- axisymmetric opacity surfaces r_tau(theta) of a density
// model for one neutrino energy (apps/extract_opacity_surface.cpp:127-206).
//
// One lane per polar angle: the lane walks the uniform radial grid inwards
// from R_max, evaluating the density once per node and accumulating the
// trapezoidal optical depth in the reference's order, so no density or tau
// array is stored.  Of the nodes where tau crosses tau_surface going outwards
// (tau[i] >= tau_s >= tau[i+1]) the innermost is kept, which is the first
// match of the reference's outward search.
#include "kernels.h"

namespace hd {

__global__ void __launch_bounds__(256) k_opacity_surface(OpacityArgs A)
{
    const int it = int(blockIdx.x) * blockDim.x + threadIdx.x;
    if (it >= A.ntheta) return;
    const double th = PI * double(it) / double(A.ntheta - 1);
    // sigma(E) of the table (src/sigma_table.cpp:80-124), E inside the table
    const int n = A.n_sigma;
    const double sigma_cm2 = exp(sigma_ln(A.sig, A.sig + n, A.sig + 2 * n, n, A.ln_enu, A.sig_lo, A.sig_inv_d));
    const double dr = A.dr;
    double rho_next = rho(A.m, A.r_min + dr * double(A.nr - 1), th);
    double tau_next = 0.0;
    int cand = -1;
    double c_t0 = 0.0, c_t1 = 0.0;
    for (int ir = A.nr - 2; ir >= 0; --ir) {
        const double rho_here = rho(A.m, A.r_min + dr * double(ir), th);
        const double rho_mid = 0.5 * (rho_here + rho_next);
        const double nb = rho_mid / M_U_G;
        const double dtau = nb * sigma_cm2 * dr * A.rg_cm;
        const double tau_here = tau_next + dtau;
        if (tau_here >= A.tau_s && tau_next <= A.tau_s) {
            cand = ir;
            c_t0 = tau_here;
            c_t1 = tau_next;
        }
        rho_next = rho_here;
        tau_next = tau_here;
    }
    const double tt = tau_next;   // tau at R_min
    double rt = -1.0;
    int cf = 0;
    if (tt >= A.tau_s && cand >= 0) {
        const double r0 = A.r_min + dr * double(cand);
        const double frac = (A.tau_s - c_t0) / dmax(c_t1 - c_t0, -1.0e-300);
        rt = r0 + dclamp(frac, 0.0, 1.0) * dr;
        cf = 1;
    }
    A.theta[it] = th;
    A.r_tau[it] = rt;
    A.tau_total[it] = tt;
    A.cls[it] = tt < 0.1 ? 0.0 : tt < 1.0 ? 1.0 : tt < 3.0 ? 2.0 : 3.0;
    A.crossing[it] = cf;
}

hipError_t launch_opacity_surface(const OpacityArgs& A, hipStream_t s)
{
    hipLaunchKernelGGL(k_opacity_surface, dim3(unsigned((A.ntheta + 255) / 256)), dim3(256), 0, s, A);
    return hipGetLastError();
}

}  // namespace hd
