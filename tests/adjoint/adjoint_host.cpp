// Host build of the device adjoint functions (csrc/adjoint_dev.h) and of
// the forward functions they differentiate, for the CPU derivative checks
// in tests/test_adjoint_cpu.py.  Test infrastructure: nothing in the
// library loads it.
#include <hip/hip_runtime.h>
#include "adjoint_dev.h"

using namespace phd;

namespace {

// forward limited gradient of one cell, same arithmetic as k_grad_sell
// with the AREPO limiter (reconstruction.pyx:563-737)
template <int D>
void grad_cell_fwd(int deg, const double* rows, const double* com, const double* area,
                   const double* pv, const double* xi, const double* cx, double vol, double* grad)
{
    constexpr int NF = D + 2;
    double df[NF * D], pmax[NF], pmin[NF];
    for (int n = 0; n < NF; n++) {
        pmax[n] = pmin[n] = pv[n];
        for (int k = 0; k < D; k++) df[n * D + k] = 0.0;
    }
    for (int e = 0; e < deg; e++) {
        const double* r = rows + 8 * e;
        double cfx[D], dr[D], rr = 0.0;
        for (int k = 0; k < D; k++) {
            cfx[k] = com[3 * e + k] - 0.5 * (xi[k] + r[k]);
            dr[k] = xi[k] - r[k];
            rr += dr[k] * dr[k];
        }
        rr = sqrt(rr);
        const double rv = rr * vol;
        for (int n = 0; n < NF; n++) {
            const double pj = r[D + n];
            pmax[n] = fmax(pmax[n], pj);
            pmin[n] = fmin(pmin[n], pj);
            const double d_dif = pj - pv[n], d_sum = pj + pv[n];
            for (int k = 0; k < D; k++)
                df[n * D + k] += area[e] * (d_dif * cfx[k] - 0.5 * d_sum * dr[k]) / rv;
        }
    }
    for (int n = 0; n < NF; n++) {
        double dpos = 0.0, dneg = 0.0, alpha = 1.0;
        for (int e = 0; e < deg; e++) {
            double dphi = 0.0;
            for (int k = 0; k < D; k++) dphi += df[n * D + k] * (com[3 * e + k] - cx[k]);
            if (dphi > 0) dpos = fmax(dpos, dphi);
            else if (dphi < 0) dneg = fmin(dneg, dphi);
        }
        if (dpos > 0) alpha = fmin(alpha, fmax((pmax[n] - pv[n]) / dpos, 0.));
        if (dneg < 0) alpha = fmin(alpha, fmax((pmin[n] - pv[n]) / dneg, 0.));
        for (int k = 0; k < D; k++) grad[n * D + k] = alpha * df[n * D + k];
    }
}

template <int D>
struct ArrNbr {
    int n;
    const double *rows, *com, *area;
    double* lnb;
    PHD_HD int deg() const { return n; }
    PHD_HD bool valid(int) const { return true; }
    PHD_HD void row(int e, double* r) const { for (int q = 0; q < 8; q++) r[q] = rows[8 * e + q]; }
    PHD_HD void geo(int e, double* c, double& a) const {
        for (int k = 0; k < 3; k++) c[k] = com[3 * e + k];
        a = area[e];
    }
    PHD_HD void out(int e, const double* lam) const { for (int q = 0; q < D + 2; q++) lnb[(D + 2) * e + q] = lam[q]; }
};

} // namespace

extern "C" {

int ah_hllc(int dim, const double* L, const double* R, const double* n, const double* fv,
            double gamma, const double* lF, double* F, double* lL, double* lR)
{
    const Gas g = make_gas(gamma);
    if (dim == 2) {
        if (F) { riemann_hll<2, PHD_SOLVER_HLLC>(L, R, n, fv, g, 1, F); deboost<2>(fv, F); }
        if (lF) hllc_vjp<2>(L, R, n, fv, g, lF, lL, lR);
    } else {
        if (F) { riemann_hll<3, PHD_SOLVER_HLLC>(L, R, n, fv, g, 1, F); deboost<3>(fv, F); }
        if (lF) hllc_vjp<3>(L, R, n, fv, g, lF, lL, lR);
    }
    return 0;
}

int ah_temporal(int dim, const double* p, const double* g, const double* fv, int boost,
                double gamma, double dt, const double* S_in, const double* lS, double* S_out,
                double* lS_in, double* lp, double* lg)
{
    const int nf = dim + 2;
    if (S_out) {
        for (int q = 0; q < nf; q++) S_out[q] = S_in[q];
        if (dim == 2) temporal_side<2>(p, g, fv, boost, gamma, dt, S_out);
        else temporal_side<3>(p, g, fv, boost, gamma, dt, S_out);
    }
    if (lS) {
        if (dim == 2) temporal_side_vjp<2>(p, g, fv, boost, gamma, dt, S_in, lS, lS_in, lp, lg);
        else temporal_side_vjp<3>(p, g, fv, boost, gamma, dt, S_in, lS, lS_in, lp, lg);
    }
    return 0;
}

int ah_spatial(int dim, const double* com, double diff, const double* ci, const double* cj,
               const double* pi, const double* pj, const double* gi, const double* gj,
               const double* fv, int boost, int gizmo, const double* lL, const double* lR,
               double* L, double* R, double* lpi, double* lpj, double* lgi, double* lgj)
{
    if (L) {
        if (dim == 2) { boost_copy<2>(pi, pj, fv, boost, L, R); spatial_states<2>(com, diff, ci, cj, pi, pj, gi, gj, gizmo, L, R); }
        else { boost_copy<3>(pi, pj, fv, boost, L, R); spatial_states<3>(com, diff, ci, cj, pi, pj, gi, gj, gizmo, L, R); }
    }
    if (lL) {
        if (dim == 2) spatial_states_vjp<2>(com, diff, ci, cj, pi, pj, gi, gj, fv, boost, gizmo, lL, lR, lpi, lpj, lgi, lgj);
        else spatial_states_vjp<3>(com, diff, ci, cj, pi, pj, gi, gj, fv, boost, gizmo, lL, lR, lpi, lpj, lgi, lgj);
    }
    return 0;
}

int ah_eos(int dim, int to_prim, const double* x, double vol, double gamma, const double* lam_out,
           double* y, double* lam_in)
{
    const int nf = dim + 2;
    if (y) {
        if (to_prim) {
            y[0] = x[0] / vol;
            double vsq = 0.;
            for (int k = 0; k < dim; k++) { y[1 + k] = x[1 + k] / x[0]; vsq += y[1 + k] * y[1 + k]; }
            y[nf - 1] = (x[nf - 1] / vol - .5 * y[0] * vsq) * (gamma - 1.);
        } else {
            double vsq = 0.;
            for (int k = 0; k < dim; k++) vsq += x[1 + k] * x[1 + k];
            y[0] = x[0] * vol;
            for (int k = 0; k < dim; k++) y[1 + k] = x[0] * x[1 + k] * vol;
            y[nf - 1] = (.5 * x[0] * vsq + x[nf - 1] / (gamma - 1.)) * vol;
        }
    }
    if (lam_out) {
        if (to_prim) { if (dim == 2) prim_from_cons_vjp<2>(x, vol, gamma, lam_out, lam_in); else prim_from_cons_vjp<3>(x, vol, gamma, lam_out, lam_in); }
        else { if (dim == 2) cons_from_prim_vjp<2>(x, vol, gamma, lam_out, lam_in); else cons_from_prim_vjp<3>(x, vol, gamma, lam_out, lam_in); }
    }
    return 0;
}

int ah_grad_cell(int dim, int deg, const double* rows, const double* com, const double* area,
                 const double* pv, const double* xi, const double* cx, double vol,
                 const double* lgrad, double* grad, double* lself, double* lnb)
{
    if (grad) {
        if (dim == 2) grad_cell_fwd<2>(deg, rows, com, area, pv, xi, cx, vol, grad);
        else grad_cell_fwd<3>(deg, rows, com, area, pv, xi, cx, vol, grad);
    }
    if (lgrad) {
        if (dim == 2) { ArrNbr<2> nb{deg, rows, com, area, lnb}; grad_cell_vjp<2>(nb, pv, xi, cx, vol, lgrad, lself); }
        else { ArrNbr<3> nb{deg, rows, com, area, lnb}; grad_cell_vjp<3>(nb, pv, xi, cx, vol, lgrad, lself); }
    }
    return 0;
}

// pow_b1 over n (base, exponent) pairs, for its check against libm pow
void ah_pow_b1(long n, const double* x, const double* y, double* out)
{
    for (long i = 0; i < n; i++) out[i] = pow_b1(x[i], y[i]);
}

}
