// opacity.hip -- axisymmetric opacity surfaces r_tau(theta) of a density
// model for one neutrino energy (apps/extract_opacity_surface.cpp:127-206).
//
// One wavefront per polar angle.  It walks the uniform radial grid inwards
// from R_max in chunks of 64 nodes (lane l of chunk c holds node
// nr - 1 - 64 c - l), evaluates the density once per node, forms the
// trapezoidal dtau with the outer neighbour's density (one lane up, or the
// previous chunk's last node), and accumulates tau with a wavefront prefix sum
// on top of the previous chunk's innermost tau.  Of the nodes where tau falls
// through tau_surface going outwards (tau[i] >= tau_s >= tau[i+1]) the
// innermost is kept, which is the first match of the reference's outward
// search.  The reference sums tau node by node; the prefix sum associates the
// same terms differently (relative differences of order 1e-16).
#include "kernels.h"
#include "wave.h"

namespace hd {

__global__ void __launch_bounds__(256) k_opacity_surface(OpacityArgs A)
{
    const int it = int((long(blockIdx.x) * blockDim.x + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    if (it >= A.ntheta) return;   // wave-uniform
    const double th = PI * double(it) / double(A.ntheta - 1);
    // sigma(E) of the table (src/sigma_table.cpp:80-124), E inside the table
    const int n = A.n_sigma;
    const double sigma_cm2 = exp(sigma_ln(A.sig, A.sig + n, A.sig + 2 * n, n, A.ln_enu, A.sig_lo, A.sig_inv_d));
    const double dr = A.dr;
    const int nr = A.nr;
    double carry_rho = 0.0, carry_tau = 0.0;   // density and tau at the previous chunk's innermost node
    int cand = -1;
    double c_t0 = 0.0, c_t1 = 0.0;
    const int nchunks = (nr + 63) / 64;
    for (int c = 0; c < nchunks; ++c) {
        const int ir = nr - 1 - (64 * c + lane);
        const bool valid = ir >= 0;
        const double rho_here = valid ? rho(A.m, A.r_min + dr * double(ir), th) : 0.0;
        double rho_out = __shfl_up(rho_here, 1, 64);   // node ir + 1
        if (lane == 0) rho_out = carry_rho;
        double d[1] = {0.0};
        if (valid && ir <= nr - 2) {
            const double rho_mid = 0.5 * (rho_here + rho_out);
            const double nb = rho_mid / M_U_G;
            d[0] = nb * sigma_cm2 * dr * A.rg_cm;
        }
        wave_scan_add(d);
        const double tau_here = carry_tau + d[0];
        double tau_out = __shfl_up(tau_here, 1, 64);
        if (lane == 0) tau_out = carry_tau;
        const unsigned long long hit =
            __ballot(valid && ir <= nr - 2 && tau_here >= A.tau_s && tau_out <= A.tau_s);
        if (hit) {
            const int L = 63 - __clzll(hit);   // innermost matching node of this chunk
            cand = nr - 1 - (64 * c + L);
            c_t0 = read_lane(tau_here, L);
            c_t1 = read_lane(tau_out, L);
        }
        const int last = min(63, nr - 1 - 64 * c);
        carry_rho = read_lane(rho_here, last);
        carry_tau = read_lane(tau_here, last);
    }
    if (lane != 0) return;
    const double tt = carry_tau;   // tau at R_min
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
    hipLaunchKernelGGL(k_opacity_surface, dim3(unsigned((A.ntheta + 3) / 4)), dim3(256), 0, s, A);   // wave per angle
    return hipGetLastError();
}

}  // namespace hd
