// states_dev.h -- device face-state reconstruction (reconstruction.pyx).
#pragma once
#include <hip/hip_runtime.h>

// callable from kernels and from host code (the host build runs the same
// functions for the CPU checks of the adjoint)
#ifndef PHD_HD
#define PHD_HD __host__ __device__
#endif

namespace phd {

// face-frame copies of the cell states (reconstruction.pyx:838-852; the
// constant reconstruction 330-352 is the same copy)
template <int D>
PHD_HD __forceinline__ void boost_copy(const double* pi, const double* pj, const double* fv,
                                           int boost, double* L, double* R)
{
#pragma unroll
    for (int n = 0; n < D + 2; n++) { L[n] = pi[n]; R[n] = pj[n]; }
    if (boost) {
#pragma unroll
        for (int k = 0; k < D; k++) { L[1 + k] = pi[1 + k] - fv[k]; R[1 + k] = pj[1 + k] - fv[k]; }
    }
}

// GIZMO pairwise limiter, Hopkins (2015) appendix B4 (reconstruction.pyx:
// 877-921), applied to the extrapolated states L, R.
//   sepi_mag, sepj_mag : SQUARED distances face centroid - cell centroid
//                        (the square roots are taken here, as in the reference)
//   sl, sr             : the face-frame cell states before extrapolation
template <int D>
PHD_HD __forceinline__ void gizmo_limit(double sepi_mag, double sepj_mag, double diff,
                                            const double* pi, const double* pj, const double* sl,
                                            const double* sr, double* L, double* R)
{
    constexpr int NF = D + 2;
    const double psi1 = 0.5, psi2 = 0.25;
    sepi_mag = sqrt(sepi_mag);
    sepj_mag = sqrt(sepj_mag);
#pragma unroll
    for (int n = 0; n < NF; n++) {
        double delta1 = psi1 * fabs(sl[n] - sr[n]);
        double delta2 = psi2 * fabs(sl[n] - sr[n]);
        double phi_min = fmin(sl[n], sr[n]);
        double phi_max = fmax(sl[n], sr[n]);
        double phibar_l = sl[n] + sepi_mag / diff * (sr[n] - sl[n]);
        double phibar_r = sr[n] + sepj_mag / diff * (sl[n] - sr[n]);
        // the two fallbacks (892-900) exclude each other for finite values:
        // the first needs phi_max < 0, the second phi_min > 0, and phi_min <=
        // phi_max; so one evaluation serves the lane's case
        const bool fp = !((phi_max + delta1) * phi_max >= 0.);
        const bool fm = !((phi_min - delta1) * phi_min >= 0.);
        double phi_plus = phi_max + delta1, phi_minus = phi_min - delta1;
        if (fp || fm) {
            const double x = fp ? phi_max : phi_min;
            const double y = x / (1 + delta1 / fabs(x));
            if (fp) phi_plus = y;
            else phi_minus = y;
        }
        if (pi[n] < pj[n]) {
            L[n] = fmax(phi_minus, fmin(phibar_l + delta2, L[n]));
            R[n] = fmin(phi_plus, fmax(phibar_r - delta2, R[n]));
        } else if (pi[n] > pj[n]) {
            L[n] = fmin(phi_plus, fmax(phibar_l - delta2, L[n]));
            R[n] = fmax(phi_minus, fmin(phibar_r + delta2, R[n]));
        } else {
            L[n] = sl[n];
            R[n] = sr[n];
        }
    }
    // negativity fallback (914-921), inside the gizmo branch as in the reference
    if (L[0] < 0.0 || L[D + 1] < 0.0) {
#pragma unroll
        for (int n = 0; n < NF; n++) L[n] = sl[n];
    }
    if (R[0] < 0.0 || R[D + 1] < 0.0) {
#pragma unroll
        for (int n = 0; n < NF; n++) R[n] = sr[n];
    }
}

// PieceWiseLinear.add_spatial body (reconstruction.pyx:849-921).
//   com   : face centroid;  ci, cj : cell centroids x + dcom (861-862)
//   diff  : |x_j - x_i| (866, 875)
//   pi,pj : lab-frame primitives (the GIZMO branch test at 902/906)
template <int D>
PHD_HD __forceinline__ void spatial_states(const double* com, double diff, const double* ci,
                                               const double* cj, const double* pi,
                                               const double* pj, const double* gi,
                                               const double* gj, int gizmo, double* L, double* R)
{
    constexpr int NF = D + 2;
    double sl[NF], sr[NF];
#pragma unroll
    for (int n = 0; n < NF; n++) { sl[n] = L[n]; sr[n] = R[n]; }
    double sepi_mag = 0.0, sepj_mag = 0.0;
#pragma unroll
    for (int k = 0; k < D; k++) {
        double sepi = com[k] - ci[k];
        double sepj = com[k] - cj[k];
        sepi_mag += sepi * sepi;
        sepj_mag += sepj * sepj;
#pragma unroll
        for (int n = 0; n < NF; n++) {
            L[n] += gi[n * D + k] * sepi;
            R[n] += gj[n * D + k] * sepj;
        }
    }
    if (gizmo) gizmo_limit<D>(sepi_mag, sepj_mag, diff, pi, pj, sl, sr, L, R);
}

// one side of PieceWiseLinear.add_temporal (reconstruction.pyx:1018-1072):
// half-step primitive Euler predictor in the face frame.  Gradient groups:
// d rho/dx_k = g[k], d v_q/dx_k = g[(1+q)D+k], d p/dx_k = g[(D+1)D+k].
template <int D>
PHD_HD __forceinline__ void temporal_side(const double* p, const double* g, const double* fv,
                                              int boost, double gamma, double dt, double* S)
{
    constexpr int NF = D + 2;
    double s0[NF], v[D];
#pragma unroll
    for (int n = 0; n < NF; n++) s0[n] = S[n];
#pragma unroll
    for (int k = 0; k < D; k++) {
        v[k] = boost ? p[1 + k] - fv[k] : p[1 + k];
        S[1 + k] -= dt * g[(D + 1) * D + k] / p[0];
    }
#pragma unroll
    for (int k = 0; k < D; k++) {
        double div = g[(1 + k) * D + k];
        S[0] -= dt * (p[0] * div + v[k] * g[k]);
        S[D + 1] -= dt * (gamma * p[D + 1] * div + v[k] * g[(D + 1) * D + k]);
#pragma unroll
        for (int q = 0; q < D; q++) S[1 + q] -= dt * v[k] * g[(1 + q) * D + k];
    }
    if (S[0] < 0.0 || S[D + 1] < 0.0) {
#pragma unroll
        for (int n = 0; n < NF; n++) S[n] = s0[n];
    }
}

} // namespace phd
