// riemann_dev.h -- device Riemann solvers in the moving face frame.
//
// Each function evaluates the reference's expressions in the reference's
// order (riemann/riemann.pyx); with FMA contraction off (the library's
// compile flags) every intermediate is the same IEEE double as in the
// reference's x86-64 build.
#pragma once
#include <hip/hip_runtime.h>
#include "phd_flux.h"

// callable from kernels and from host code (the host build runs the same
// functions for the CPU checks of the adjoint)
#ifndef PHD_HD
#define PHD_HD __host__ __device__
#endif

namespace phd {

// Gamma-only subexpressions of the HLL-family solvers, evaluated once on the
// host with the reference's expressions.  Division and subtraction are
// correctly rounded on both sides, so each constant has the same bits as the
// per-face evaluation in riemann.pyx.
struct Gas {
    double gamma;
    double z;        // (gamma - 1.)/(2.*gamma)        riemann.pyx:452
    double inv_z;    // 1./z                            riemann.pyx:460-461
    double gp1_2g;   // (gamma+1.)/(2.*gamma)           riemann.pyx:488,497
    double gm1_gp1;  // (gamma - 1.)/(gamma + 1.)       riemann.pyx:470-471
};

inline Gas make_gas(double gamma)
{
    Gas g;
    g.gamma = gamma;
    g.z = (gamma - 1.) / (2. * gamma);
    g.inv_z = 1. / g.z;
    g.gp1_2g = (gamma + 1.) / (2. * gamma);
    g.gm1_gp1 = (gamma - 1.) / (gamma + 1.);
    return g;
}

// pow for bases at or near one.  In a uniform region every TRRS base is one
// or, when the flow moves, one plus an ulp or two (u* rounds differently from
// u), so most faces land here.  For |x - 1| (|y| + 3) <= 2^-20 the binomial
// series (1 + d)^y = 1 + y d + y(y-1)/2 d^2 + ... is cut after the quadratic
// term: the rest is below (|d|(|y| + 3))^3 / 6 < 2^-62, under a thousandth of
// an ulp of 1, and the Horner sum rounds once at the final 1 + (...), so the
// result is within an ulp of the exact power (libm's pow is within one ulp of
// it too).  x - 1 is exact there (Sterbenz), and x == 1 gives exactly 1, as
// pow(1, y) does (C99 F.9.4.4).  A wave whose lanes all take the series never
// executes the pow sequence.
PHD_HD __forceinline__ double pow_b1(double x, double y)
{
    const double d = x - 1.0;
    if (__builtin_expect(fabs(d) * (fabs(y) + 3.0) <= 0x1p-20, 1))
        return 1.0 + d * (y + d * (0.5 * y * (y - 1.0)));
    return pow(x, y);
}

// HLL.get_waves (riemann.pyx:375-504): Toro PVRS / TRRS / TSRS p* estimate
// and Toro eq. 10.68-10.70 wave speeds.
PHD_HD __forceinline__ void get_waves(double dl, double ul, double pl, double dr, double ur,
                                          double pr, const Gas& g, double& sl, double& sc,
                                          double& sr)
{
    const double gamma = g.gamma;
    const double Q = 2., c_floor = 1.0E-10;
    double cl = fmax(sqrt(gamma * pl / dl), c_floor);
    double cr = fmax(sqrt(gamma * pr / dr), c_floor);
    double d_avg = .5 * (dl + dr);
    double c_avg = .5 * (cl + cr);
    double p_star = fmax(0., .5 * (pl + pr) + .5 * (ul - ur) * d_avg * c_avg);
    double p_min = fmin(pl, pr);
    double p_max = fmax(pl, pr);
    if (((p_max / p_min) < Q) && ((p_min < p_star) && (p_star < p_max))) {
        // PVRS: p_star kept (u_star at 445 is not used further)
    } else if (p_star <= p_min) {
        double plr = pow_b1(pl / pr, g.z);
        double u_star = (plr * ul / cl + ur / cr + 2. * (plr - 1.) / (gamma - 1.)) /
                        (plr / cl + 1. / cr);
        p_star = .5 * pl * pow_b1(1. + (gamma - 1.) * (ul - u_star) / (2. * cl), g.inv_z);
        p_star += .5 * pr * pow_b1(1. + (gamma - 1.) * (u_star - ur) / (2. * cr), g.inv_z);
    } else {
        double Al = 2. / ((gamma + 1.) * dl);
        double Ar = 2. / ((gamma + 1.) * dr);
        double Bl = pl * g.gm1_gp1;
        double Br = pr * g.gm1_gp1;
        double gl = sqrt(Al / (p_star + Bl));
        double gr = sqrt(Ar / (p_star + Br));
        p_star = (gl * pl + gr * pr - (ur - ul)) / (gl + gr);
    }
    if (p_star <= pl) sl = ul - cl;
    else sl = ul - cl * sqrt(1. + g.gp1_2g * (p_star / pl - 1.));
    if (p_star <= pr) sr = ur + cr;
    else sr = ur + cr * sqrt(1. + g.gp1_2g * (p_star / pr - 1.));
    sc = (pr - pl + dl * ul * (sl - ul) - dr * ur * (sr - ur)) / (dl * (sl - ul) - dr * (sr - ur));
}

// pure upwind state flux in the moving frame (riemann.pyx:624-631, 672-679)
template <int D>
PHD_HD __forceinline__ void side_flux(const double* S, const double* n, double Vn, double vsq,
                                          double wn, double gamma, double* F)
{
    double d = S[0], p = S[D + 1];
    F[0] = d * (Vn - wn);
    F[D + 1] = (0.5 * d * vsq + p / (gamma - 1.0)) * (Vn - wn) + p * Vn;
#pragma unroll
    for (int k = 0; k < D; k++) F[1 + k] = d * S[1 + k] * (Vn - wn) + p * n[k];
}

// HLL (riemann.pyx:243-373) and HLLC (riemann.pyx:521-682) interface flux
template <int D, int SOLVER>
PHD_HD __forceinline__ void riemann_hll(const double* L, const double* R, const double* n,
                                            const double* fv, const Gas& g, int boost, double* F)
{
    const double gamma = g.gamma;
    double _dl = L[0], _pl = L[D + 1], _dr = R[0], _pr = R[D + 1];
    double Vnl = 0.0, Vnr = 0.0, vl_sq = 0.0, vr_sq = 0.0, wn = 0.0;
#pragma unroll
    for (int k = 0; k < D; k++) {
        vl_sq += L[1 + k] * L[1 + k];
        vr_sq += R[1 + k] * R[1 + k];
        Vnl += L[1 + k] * n[k];
        Vnr += R[1 + k] * n[k];
        wn += fv[k] * n[k];
    }
    if (boost) wn = 0.;
    double sl, sc, sr;
    get_waves(_dl, Vnl, _pl, _dr, Vnr, _pr, g, sl, sc, sr);
    if (SOLVER == PHD_SOLVER_HLL && (sl < wn) && (wn <= sr)) {
        // Toro eq. 10.20-10.21 (riemann.pyx:346-361)
        double fac1 = sr - wn, fac2 = sl - wn, fac3 = sr - sl;
        F[0] = (_dl * Vnl * fac1 - _dr * Vnr * fac2 - sl * _dl * fac1 + sr * _dr * fac2) / fac3;
#pragma unroll
        for (int k = 0; k < D; k++)
            F[1 + k] = ((_dl * L[1 + k] * Vnl + _pl * n[k]) * fac1 -
                        (_dr * R[1 + k] * Vnr + _pr * n[k]) * fac2 -
                        sl * (_dl * L[1 + k]) * fac1 + sr * (_dr * R[1 + k]) * fac2) / fac3;
        double el = 0.5 * _dl * vl_sq + _pl / (gamma - 1.0);
        double er = 0.5 * _dr * vr_sq + _pr / (gamma - 1.0);
        F[D + 1] = ((el + _pl) * Vnl * fac1 - (er + _pr) * Vnr * fac2 - sl * el * fac1 +
                    sr * er * fac2) / fac3;
    } else {
        // The two upwind fluxes (riemann.pyx:624-631, 672-679) and the two
        // HLLC star fluxes (Toro eq. 10.38-10.39, 636-652 and 654-670) are
        // each one expression applied to the left or the right side's data:
        // each lane selects its side and the expression is evaluated once,
        // so a wave with faces on both sides runs one evaluation, and every
        // lane computes the operations of its own branch.
        const bool mid = !(wn <= sl) && (sl < wn) && (wn <= sr);
        const bool left = mid ? (wn <= sc) : (wn <= sl);
        double S[D + 2];
#pragma unroll
        for (int q = 0; q < D + 2; q++) S[q] = left ? L[q] : R[q];
        const double d = S[0], p = S[D + 1];
        const double Vn = left ? Vnl : Vnr, vsq = left ? vl_sq : vr_sq, s = left ? sl : sr;
        if (SOLVER == PHD_SOLVER_HLLC && mid) {
            double factor_1 = d * (s - Vn) / (s - sc);
            double factor_2 = factor_1 * (s - wn) * (sc - Vn) + p;
            double frho = d * (Vn - s) + factor_1 * (s - wn);
            double e = 0.5 * d * vsq + p / (gamma - 1.0);
            F[0] = frho;
            F[D + 1] = (e + p) * Vn - s * e +
                       (s - wn) * factor_1 * (e / d + (sc - Vn) * (sc + p / (d * (s - Vn))));
#pragma unroll
            for (int k = 0; k < D; k++) F[1 + k] = frho * S[1 + k] + factor_2 * n[k];
        } else {
            side_flux<D>(S, n, Vn, vsq, wn, gamma, F);
        }
    }
}

// RiemannBase.deboost (riemann.pyx:190-222), Pakmor & Springel (2011) eq. 17
template <int D>
PHD_HD __forceinline__ void deboost(const double* fv, double* F)
{
#pragma unroll
    for (int k = 0; k < D; k++) {
        F[D + 1] += fv[k] * (0.5 * fv[k] * F[0] + F[1 + k]);
        F[1 + k] += fv[k] * F[0];
    }
}

// ---- Exact solver (riemann.pyx:684-1090), Toro chapter 4 -------------------
PHD_HD inline double ex_p_func(double d, double p, double c, double gamma, double p_old)
{
    if (p_old <= p) return 2. * c / (gamma - 1.) * (pow_b1(p_old / p, (gamma - 1.) / (2. * gamma)) - 1.);
    double Ak = 2. / (d * (gamma + 1.));
    double Bk = p * (gamma - 1.) / (gamma + 1.);
    return (p_old - p) * sqrt(Ak / (p_old + Bk));
}

PHD_HD inline double ex_p_deriv(double d, double p, double c, double gamma, double p_old)
{
    if (p_old <= p) return pow_b1(p_old / p, -(gamma + 1.) / (2. * gamma)) / (c * d);
    double Ak = 2. / (d * (gamma + 1.));
    double Bk = p * (gamma - 1.) / (gamma + 1.);
    return sqrt(Ak / (p_old + Bk)) * (1. - .5 * (p_old - p) / (Bk + p_old));
}

// p_guess + get_pstar (riemann.pyx:975-1090): Newton, TOL 1e-6, 1000 steps
PHD_HD inline bool ex_pstar(double dl, double ul, double pl, double cl, double dr, double ur,
                                double pr, double cr, double gamma, double& out)
{
    double ppv = .5 * (pl + pr) - .125 * (ur - ul) * (dl + dr) * (cl + cr);
    double p_max = (pl > pr) ? pl : pr;
    double p_min = (pl < pr) ? pl : pr;
    double q_max = p_max / p_min;
    double p_old;
    if ((q_max <= 2.) && (p_min <= ppv && ppv <= p_max)) {
        p_old = ppv;
    } else if (ppv <= p_min) {
        double p_lr = pow_b1(pl / pr, (gamma - 1.) / (2. * gamma));
        double u_star = (p_lr * ul / cl + ur / cr + 2. * (p_lr - 1.) / (gamma - 1.));
        u_star = u_star / (p_lr / cl + 1. / cr);
        double p_tl = pow_b1(1. + (gamma - 1.) * (ul - u_star) / (2. * cl), 2. * gamma / (gamma - 1.));
        double p_tr = pow_b1(1. + (gamma - 1.) * (u_star - ur) / (2. * cr), 2. * gamma / (gamma - 1.));
        p_old = .5 * (pl * p_tl + pr * p_tr);
    } else {
        double gl = sqrt((2. / (dl * (gamma + 1.))) / ((gamma - 1.) * pl / (gamma + 1.) + ppv));
        double gr = sqrt((2. / (dr * (gamma + 1.))) / ((gamma - 1.) * pr / (gamma + 1.) + ppv));
        p_old = (gl * pl + gr * pr - (ur - ul)) / (gr + gl);
    }
    for (int it = 0; it < 1000; it++) {
        double fl = ex_p_func(dl, pl, cl, gamma, p_old);
        double fr = ex_p_func(dr, pr, cr, gamma, p_old);
        double dfl = ex_p_deriv(dl, pl, cl, gamma, p_old);
        double dfr = ex_p_deriv(dr, pr, cr, gamma, p_old);
        double p_new = p_old - (fl + fr + ur - ul) / (dfl + dfr);
        if ((2. * fabs((p_new - p_old) / (p_new + p_old))) <= 1.0e-6) { out = p_new; return true; }
        if (p_new < 0.) p_new = 1.0e-6;
        p_old = p_new;
    }
    return false;
}

template <int D>
PHD_HD __forceinline__ void ex_state_flux(double d, double p, const double* v, const double* n,
                                              double gamma, double* F)
{
    double v_sq = 0., vn = 0.;
#pragma unroll
    for (int k = 0; k < D; k++) { vn += v[k] * n[k]; v_sq += v[k] * v[k]; }
    F[0] = d * vn;
    F[D + 1] = (0.5 * d * v_sq + gamma * p / (gamma - 1.0)) * vn;
#pragma unroll
    for (int k = 0; k < D; k++) F[1 + k] = d * v[k] * vn + p * n[k];
}

// Exact.riemann_solver sampling at x/t = 0 (riemann.pyx:763-970).  Returns
// false on Newton non-convergence.
template <int D>
PHD_HD bool riemann_exact(const double* L, const double* R, const double* n, double gamma,
                              double* F)
{
    double _dl = L[0], _pl = L[D + 1], _dr = R[0], _pr = R[D + 1];
    double cl = sqrt(gamma * _pl / _dl), cr = sqrt(gamma * _pr / _dr);
    double vnl = 0.0, vnr = 0.0;
#pragma unroll
    for (int k = 0; k < D; k++) { vnl += L[1 + k] * n[k]; vnr += R[1 + k] * n[k]; }
    double p_star;
    if (!ex_pstar(_dl, vnl, _pl, cl, _dr, vnr, _pr, cr, gamma, p_star)) return false;
    double fl = ex_p_func(_dl, _pl, cl, gamma, p_star);
    double fr = ex_p_func(_dr, _pr, cr, gamma, p_star);
    double u_star = 0.5 * (vnl + vnr + fr - fl);
    double v[D];
    const double* S;
    double d = 0, p = 0, shift = 0, vn0;
    bool data = false;
    if (0.0 <= u_star) {
        S = L; vn0 = vnl;
        if (p_star <= _pl) {
            if (0.0 <= vnl - cl) data = true;
            else {
                double c_star_l = cl * pow_b1(p_star / _pl, (gamma - 1.0) / (2.0 * gamma));
                if (0.0 >= u_star - c_star_l) { d = _dl * pow_b1(p_star / _pl, 1.0 / gamma); p = p_star; shift = u_star; }
                else {
                    double c = (2.0 / (gamma + 1.0)) * (cl + 0.5 * (gamma - 1.0) * vnl);
                    d = _dl * pow_b1(c / cl, 2.0 / (gamma - 1.0));
                    p = _pl * pow_b1(c / cl, 2.0 * gamma / (gamma - 1.0));
                    shift = c;
                }
            }
        } else {
            double sl = vnl - cl * sqrt((gamma + 1.0) * p_star / (2.0 * gamma * _pl) +
                                        (gamma - 1.0) / (2.0 * gamma));
            if (0.0 <= sl) data = true;
            else {
                d = _dl * (p_star / _pl + (gamma - 1.0) / (gamma + 1.0)) /
                    (p_star * (gamma - 1.0) / ((gamma + 1.0) * _pl) + 1.0);
                p = p_star; shift = u_star;
            }
        }
    } else {
        S = R; vn0 = vnr;
        if (p_star >= _pr) {
            double sr = vnr + cr * sqrt((gamma + 1.0) * p_star / (2.0 * gamma * _pr) +
                                        (gamma - 1.0) / (2.0 * gamma));
            if (0.0 >= sr) data = true;
            else {
                d = _dr * (p_star / _pr + (gamma - 1.0) / (gamma + 1.0)) /
                    (p_star * (gamma - 1.0) / ((gamma + 1.0) * _pr) + 1.0);
                p = p_star; shift = u_star;
            }
        } else {
            if (0.0 >= vnr + cr) data = true;
            else {
                double c_star_r = cr * pow_b1(p_star / _pr, (gamma - 1.0) / (2.0 * gamma));
                if (0.0 <= u_star + c_star_r) { d = _dr * pow_b1(p_star / _pr, 1.0 / gamma); p = p_star; shift = u_star; }
                else {
                    double c = (2.0 / (gamma + 1.0)) * (cr - 0.5 * (gamma - 1.0) * vnr);
                    double u_tmp = (2.0 / (gamma + 1.0)) * (-cr + 0.5 * (gamma - 1.0) * vnr);
                    d = _dr * pow_b1(c / cr, 2.0 / (gamma - 1.0));
                    p = _pr * pow_b1(c / cr, 2.0 * gamma / (gamma - 1.0));
                    shift = u_tmp;
                }
            }
        }
    }
    if (data) {
        // unperturbed data state (riemann.pyx:809-813 and mirrors)
        double dd = S[0], pp = S[D + 1], vsq = 0.0;
#pragma unroll
        for (int k = 0; k < D; k++) vsq += S[1 + k] * S[1 + k];
        F[0] = dd * vn0;
        F[D + 1] = (0.5 * dd * vsq + gamma * pp / (gamma - 1.0)) * vn0;
#pragma unroll
        for (int k = 0; k < D; k++) F[1 + k] = dd * S[1 + k] * vn0 + pp * n[k];
        return true;
    }
#pragma unroll
    for (int k = 0; k < D; k++) v[k] = S[1 + k] + (shift - vn0) * n[k];
    ex_state_flux<D>(d, p, v, n, gamma, F);
    return true;
}

// Riemann solve + deboost as RiemannBase.compute_fluxes does (riemann.pyx:74-98);
// Exact always deboosts (972).
template <int D, int SOLVER>
PHD_HD __forceinline__ bool solve_face(const double* L, const double* R, const double* n,
                                           const double* fv, const Gas& g, int boost, double* F)
{
    if (SOLVER == PHD_SOLVER_EXACT) {
        if (!riemann_exact<D>(L, R, n, g.gamma, F)) return false;
        deboost<D>(fv, F);
        return true;
    }
    riemann_hll<D, SOLVER>(L, R, n, fv, g, boost, F);
    if (boost) deboost<D>(fv, F);
    return true;
}

} // namespace phd
