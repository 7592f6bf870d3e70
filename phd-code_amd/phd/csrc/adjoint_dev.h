// adjoint_dev.h -- hand-written reverse mode (vector-Jacobian products) of
// the per-face and per-cell pieces of the static-mesh core step:
//
//   EOS both ways                 equation_state.pyx:43-117
//   HLLC + wave estimates         riemann.pyx:375-504, 521-682
//   MUSCL-Hancock predictor       reconstruction.pyx:1018-1072
//   extrapolation + GIZMO limiter reconstruction.pyx:849-921
//   eq. 21 gradient + AREPO       reconstruction.pyx:563-737
//
// Each *_vjp function recomputes the forward values it needs with the same
// expressions and branch tests as the forward functions (riemann_dev.h,
// states_dev.h), so it differentiates exactly the branch the forward took,
// then accumulates lambda_in += J^T lambda_out.  The functions are host +
// device: the CPU tests check every one against central differences.
#pragma once
#include "riemann_dev.h"
#include "states_dev.h"

namespace phd {

// ---- equation of state ----------------------------------------------------
// prim = P(cons, vol): d = m/V, v = mv/m, p = (E/V - .5 d v^2)(gamma-1)
template <int D>
PHD_HD inline void prim_from_cons_vjp(const double* u, double vol, double gamma,
                                      const double* lp, double* lu)
{
    const double d = u[0] / vol;
    double v[D], vsq = 0.;
    for (int k = 0; k < D; k++) {
        v[k] = u[1 + k] / u[0];
        vsq += v[k] * v[k];
    }
    // p = (E/V - .5 d vsq)(g-1)
    const double lpp = lp[D + 1];
    lu[D + 1] += lpp * (gamma - 1.) / vol;
    const double ld = lp[0] - lpp * (gamma - 1.) * 0.5 * vsq;
    double lv[D];
    for (int k = 0; k < D; k++) lv[k] = lp[1 + k] - lpp * (gamma - 1.) * d * v[k];
    lu[0] += ld / vol;
    for (int k = 0; k < D; k++) {
        lu[1 + k] += lv[k] / u[0];
        lu[0] -= lv[k] * v[k] / u[0];
    }
}

// cons = C(prim, vol): m = dV, mv = d v V, E = (.5 d v^2 + p/(gamma-1)) V
template <int D>
PHD_HD inline void cons_from_prim_vjp(const double* p, double vol, double gamma,
                                      const double* lu, double* lp)
{
    const double d = p[0];
    double vsq = 0.;
    for (int k = 0; k < D; k++) vsq += p[1 + k] * p[1 + k];
    lp[0] += lu[0] * vol + lu[D + 1] * 0.5 * vsq * vol;
    for (int k = 0; k < D; k++) {
        lp[0] += lu[1 + k] * p[1 + k] * vol;
        lp[1 + k] += lu[1 + k] * d * vol + lu[D + 1] * d * p[1 + k] * vol;
    }
    lp[D + 1] += lu[D + 1] * vol / (gamma - 1.);
}

// ---- wave estimates ----------------------------------------------------------
// reverse of get_waves; lam_* accumulate
PHD_HD inline void get_waves_vjp(double dl, double ul, double pl, double dr, double ur, double pr,
                                 const Gas& g, double lsl, double lsc, double lsr, double& ldl,
                                 double& lul, double& lpl, double& ldr, double& lur, double& lpr)
{
    const double gamma = g.gamma;
    const double Q = 2., c_floor = 1.0E-10;
    const double cl0 = sqrt(gamma * pl / dl), cr0 = sqrt(gamma * pr / dr);
    const double cl = fmax(cl0, c_floor), cr = fmax(cr0, c_floor);
    const double d_avg = .5 * (dl + dr);
    const double c_avg = .5 * (cl + cr);
    const double ps0 = .5 * (pl + pr) + .5 * (ul - ur) * d_avg * c_avg;
    const double ps = fmax(0., ps0);
    const double p_min = fmin(pl, pr), p_max = fmax(pl, pr);
    int branch;   // 0 PVRS, 1 TRRS, 2 TSRS
    double p_star = ps;
    // TRRS intermediates
    double plr = 0, u_star = 0, Nu = 0, Du = 0, bl = 0, br = 0, powl = 0, powr = 0;
    // TSRS intermediates
    double Al = 0, Ar = 0, Bl = 0, Br = 0, gl = 0, gr = 0, Np = 0, Dp = 0;
    if (((p_max / p_min) < Q) && ((p_min < ps) && (ps < p_max))) {
        branch = 0;
    } else if (ps <= p_min) {
        branch = 1;
        plr = pow_b1(pl / pr, g.z);
        Nu = plr * ul / cl + ur / cr + 2. * (plr - 1.) / (gamma - 1.);
        Du = plr / cl + 1. / cr;
        u_star = Nu / Du;
        bl = 1. + (gamma - 1.) * (ul - u_star) / (2. * cl);
        br = 1. + (gamma - 1.) * (u_star - ur) / (2. * cr);
        powl = pow_b1(bl, g.inv_z);
        powr = pow_b1(br, g.inv_z);
        p_star = .5 * pl * powl;
        p_star += .5 * pr * powr;
    } else {
        branch = 2;
        Al = 2. / ((gamma + 1.) * dl);
        Ar = 2. / ((gamma + 1.) * dr);
        Bl = pl * g.gm1_gp1;
        Br = pr * g.gm1_gp1;
        gl = sqrt(Al / (ps + Bl));
        gr = sqrt(Ar / (ps + Br));
        Np = gl * pl + gr * pr - (ur - ul);
        Dp = gl + gr;
        p_star = Np / Dp;
    }
    double sl, sr, ql = 1., qr = 1.;
    const bool shock_l = !(p_star <= pl), shock_r = !(p_star <= pr);
    if (!shock_l) sl = ul - cl;
    else { ql = sqrt(1. + g.gp1_2g * (p_star / pl - 1.)); sl = ul - cl * ql; }
    if (!shock_r) sr = ur + cr;
    else { qr = sqrt(1. + g.gp1_2g * (p_star / pr - 1.)); sr = ur + cr * qr; }
    const double N = pr - pl + dl * ul * (sl - ul) - dr * ur * (sr - ur);
    const double Dn = dl * (sl - ul) - dr * (sr - ur);
    const double sc = N / Dn;

    // ---- reverse
    double lcl = 0., lcr = 0., lp_star = 0.;
    const double lN = lsc / Dn, lDn = -lsc * sc / Dn;
    lpr += lN;
    lpl -= lN;
    ldl += lN * ul * (sl - ul) + lDn * (sl - ul);
    lul += lN * dl * (sl - 2. * ul) - lDn * dl;
    lsl += lN * dl * ul + lDn * dl;
    ldr -= lN * ur * (sr - ur) + lDn * (sr - ur);
    lur -= lN * dr * (sr - 2. * ur) - lDn * dr;
    lsr -= lN * dr * ur + lDn * dr;
    // sr
    lur += lsr;
    if (!shock_r) lcr += lsr;
    else {
        lcr += lsr * qr;
        const double lq = lsr * cr, dq = g.gp1_2g / (2. * qr) * lq;
        lp_star += dq / pr;
        lpr -= dq * p_star / (pr * pr);
    }
    // sl
    lul += lsl;
    if (!shock_l) lcl -= lsl;
    else {
        lcl -= lsl * ql;
        const double lq = -lsl * cl, dq = g.gp1_2g / (2. * ql) * lq;
        lp_star += dq / pl;
        lpl -= dq * p_star / (pl * pl);
    }
    double lps = 0.;
    if (branch == 0) {
        lps += lp_star;
    } else if (branch == 1) {
        lpl += lp_star * .5 * powl;
        lpr += lp_star * .5 * powr;
        const double lbl = bl != 0. ? lp_star * .5 * pl * g.inv_z * powl / bl : 0.;
        const double lbr = br != 0. ? lp_star * .5 * pr * g.inv_z * powr / br : 0.;
        double lus = 0.;
        lul += lbl * (gamma - 1.) / (2. * cl);
        lus -= lbl * (gamma - 1.) / (2. * cl);
        lcl -= lbl * (gamma - 1.) * (ul - u_star) / (2. * cl * cl);
        lus += lbr * (gamma - 1.) / (2. * cr);
        lur -= lbr * (gamma - 1.) / (2. * cr);
        lcr -= lbr * (gamma - 1.) * (u_star - ur) / (2. * cr * cr);
        const double lNu = lus / Du, lDu = -lus * u_star / Du;
        double lplr = lNu * (ul / cl + 2. / (gamma - 1.)) + lDu / cl;
        lul += lNu * plr / cl;
        lcl -= lNu * plr * ul / (cl * cl) + lDu * plr / (cl * cl);
        lur += lNu / cr;
        lcr -= lNu * ur / (cr * cr) + lDu / (cr * cr);
        const double x = pl / pr;
        const double lx = x != 0. ? lplr * g.z * plr / x : 0.;
        lpl += lx / pr;
        lpr -= lx * pl / (pr * pr);
    } else {
        const double lNp = lp_star / Dp, lDp = -lp_star * p_star / Dp;
        const double lgl = lNp * pl + lDp, lgr = lNp * pr + lDp;
        lpl += lNp * gl;
        lpr += lNp * gr;
        lur -= lNp;
        lul += lNp;
        // gl = sqrt(Al/(ps+Bl))
        {
            const double s = ps + Bl, t = Al / s, lt = lgl / (2. * gl);
            const double lAl = lt / s, ls = -lt * t / s;
            lps += ls;
            lpl += ls * g.gm1_gp1;
            ldl -= lAl * Al / dl;
        }
        {
            const double s = ps + Br, t = Ar / s, lt = lgr / (2. * gr);
            const double lAr = lt / s, ls = -lt * t / s;
            lps += ls;
            lpr += ls * g.gm1_gp1;
            ldr -= lAr * Ar / dr;
        }
    }
    // ps = max(0, ps0)
    const double lps0 = ps0 > 0. ? lps : 0.;
    lpl += .5 * lps0;
    lpr += .5 * lps0;
    lul += .5 * lps0 * d_avg * c_avg;
    lur -= .5 * lps0 * d_avg * c_avg;
    const double ld_avg = .5 * lps0 * (ul - ur) * c_avg;
    const double lc_avg = .5 * lps0 * (ul - ur) * d_avg;
    ldl += .5 * ld_avg;
    ldr += .5 * ld_avg;
    lcl += .5 * lc_avg;
    lcr += .5 * lc_avg;
    // cl = max(sqrt(gamma pl/dl), c_floor)
    if (cl0 > c_floor) {
        lpl += lcl * gamma / (2. * dl * cl0);
        ldl -= lcl * cl0 / (2. * dl);
    }
    if (cr0 > c_floor) {
        lpr += lcr * gamma / (2. * dr * cr0);
        ldr -= lcr * cr0 / (2. * dr);
    }
}

// ---- HLLC flux (boosted frame: face speed wn = 0) ---------------------------
// reverse of side_flux for state S (wn = 0): F0 = d Vn, F_k = d v_k Vn + p n_k,
// F_E = (.5 d vsq + p/(g-1)) Vn + p Vn
template <int D>
PHD_HD inline void side_flux_vjp(const double* S, const double* n, double Vn, double vsq,
                                 double gamma, const double* lF, double* lS, double& lVn,
                                 double& lvsq)
{
    const double d = S[0], p = S[D + 1];
    const double E = .5 * d * vsq + p / (gamma - 1.);
    lS[0] += lF[0] * Vn + lF[D + 1] * .5 * vsq * Vn;
    lVn += lF[0] * d + lF[D + 1] * (E + p);
    lS[D + 1] += lF[D + 1] * (Vn / (gamma - 1.) + Vn);
    lvsq += lF[D + 1] * .5 * d * Vn;
    for (int k = 0; k < D; k++) {
        lS[0] += lF[1 + k] * S[1 + k] * Vn;
        lVn += lF[1 + k] * d * S[1 + k];
        lS[1 + k] += lF[1 + k] * d * Vn;
        lS[D + 1] += lF[1 + k] * n[k];
    }
}

// reverse of the HLLC star-state flux of side S with wave s (sl or sr) and
// contact sc (riemann.pyx:636-670, wn = 0)
template <int D>
PHD_HD inline void star_flux_vjp(const double* S, const double* n, double Vn, double vsq,
                                 double s, double sc, double gamma, const double* lF, double* lS,
                                 double& lVn, double& lvsq, double& ls, double& lsc)
{
    const double d = S[0], p = S[D + 1];
    const double a1 = s - Vn, b1 = s - sc, c1 = sc - Vn;
    const double f1 = d * a1 / b1;
    const double frho = d * (Vn - s) + f1 * s;
    const double e = .5 * d * vsq + p / (gamma - 1.);
    const double Y = sc + p / (d * a1);
    const double X = e / d + c1 * Y;
    double lfrho = lF[0], lf2 = 0.;
    for (int k = 0; k < D; k++) {
        lfrho += lF[1 + k] * S[1 + k];
        lS[1 + k] += lF[1 + k] * frho;
        lf2 += lF[1 + k] * n[k];
    }
    const double lE = lF[D + 1];
    double le = lE * (Vn - s);
    lS[D + 1] += lE * Vn;
    lVn += lE * (e + p);
    ls += -lE * e + lE * f1 * X;
    double lf1 = lE * s * X;
    const double lX = lE * s * f1;
    le += lX / d;
    lS[0] -= lX * e / (d * d);
    double lc1 = lX * Y;
    const double lY = lX * c1;
    lsc += lY;
    {
        const double da = d * a1, W = p / da, lda = -lY * W / da;
        lS[D + 1] += lY / da;
        lS[0] += lda * a1;
        // a1 part handled below through la1
        double la1 = lda * d;
        // factor_2 = f1 s c1 + p
        lf1 += lf2 * s * c1;
        ls += lf2 * f1 * c1;
        lc1 += lf2 * f1 * s;
        lS[D + 1] += lf2;
        // frho = d (Vn - s) + f1 s
        lS[0] += lfrho * (Vn - s);
        lVn += lfrho * d;
        ls += -lfrho * d + lfrho * f1;
        lf1 += lfrho * s;
        // e
        lS[0] += le * .5 * vsq;
        lvsq += le * .5 * d;
        lS[D + 1] += le / (gamma - 1.);
        // f1 = d a1 / b1
        lS[0] += lf1 * a1 / b1;
        la1 += lf1 * d / b1;
        const double lb1 = -lf1 * f1 / b1;
        // c1 = sc - Vn, b1 = s - sc, a1 = s - Vn
        lsc += lc1 - lb1;
        lVn -= lc1 + la1;
        ls += lb1 + la1;
    }
}

// lambda_L, lambda_R += J^T lambda_F for the boosted HLLC flux (wn = 0)
// followed by the deboost with face velocity fv (riemann.pyx:190-222)
template <int D>
PHD_HD inline void hllc_vjp(const double* L, const double* R, const double* n, const double* fv,
                            const Gas& g, const double* lF_out, double* lL, double* lR)
{
    constexpr int NF = D + 2;
    const double gamma = g.gamma;
    // deboost reverse: forward is, for k in order,
    //   F_E += fv_k (.5 fv_k F0 + F_k);  F_k += fv_k F0
    double lF[NF];
    for (int q = 0; q < NF; q++) lF[q] = lF_out[q];
    for (int k = D - 1; k >= 0; k--) {
        lF[0] += lF[1 + k] * fv[k];
        lF[0] += lF[D + 1] * fv[k] * .5 * fv[k];
        lF[1 + k] += lF[D + 1] * fv[k];
    }
    double Vnl = 0., Vnr = 0., vl_sq = 0., vr_sq = 0.;
    for (int k = 0; k < D; k++) {
        vl_sq += L[1 + k] * L[1 + k];
        vr_sq += R[1 + k] * R[1 + k];
        Vnl += L[1 + k] * n[k];
        Vnr += R[1 + k] * n[k];
    }
    const double wn = 0.;
    double sl, sc, sr;
    get_waves(L[0], Vnl, L[D + 1], R[0], Vnr, R[D + 1], g, sl, sc, sr);
    double lVnl = 0., lVnr = 0., lvl = 0., lvr = 0., lsl = 0., lsc = 0., lsr = 0.;
    if (wn <= sl) {
        side_flux_vjp<D>(L, n, Vnl, vl_sq, gamma, lF, lL, lVnl, lvl);
    } else if ((sl < wn) && (wn <= sr)) {
        if (wn <= sc) star_flux_vjp<D>(L, n, Vnl, vl_sq, sl, sc, gamma, lF, lL, lVnl, lvl, lsl, lsc);
        else star_flux_vjp<D>(R, n, Vnr, vr_sq, sr, sc, gamma, lF, lR, lVnr, lvr, lsr, lsc);
    } else {
        side_flux_vjp<D>(R, n, Vnr, vr_sq, gamma, lF, lR, lVnr, lvr);
    }
    // the upwind side fluxes do not read the wave estimates; skipping their
    // reverse also keeps a NaN estimate (vacuum-forming states, where the
    // forward falls through to an upwind flux) out of the lambdas
    if (lsl != 0. || lsc != 0. || lsr != 0.)
        get_waves_vjp(L[0], Vnl, L[D + 1], R[0], Vnr, R[D + 1], g, lsl, lsc, lsr, lL[0], lVnl,
                      lL[D + 1], lR[0], lVnr, lR[D + 1]);
    for (int k = 0; k < D; k++) {
        lL[1 + k] += lVnl * n[k] + 2. * lvl * L[1 + k];
        lR[1 + k] += lVnr * n[k] + 2. * lvr * R[1 + k];
    }
}

// ---- MUSCL-Hancock predictor -------------------------------------------------
// reverse of temporal_side: S_out = S_in - dt*terms(p, g, fv), or S_in when
// the predicted density or pressure is negative
template <int D>
PHD_HD inline void temporal_side_vjp(const double* p, const double* g, const double* fv,
                                     int boost, double gamma, double dt, const double* S_in,
                                     const double* lS, double* lS_in, double* lp, double* lg)
{
    constexpr int NF = D + 2;
    // the forward reverts to S_in on a negative predicted density or
    // pressure; recompute the prediction for that test
    double P[NF];
    for (int q = 0; q < NF; q++) P[q] = S_in[q];
    double v[D];
    for (int k = 0; k < D; k++) {
        v[k] = boost ? p[1 + k] - fv[k] : p[1 + k];
        P[1 + k] -= dt * g[(D + 1) * D + k] / p[0];
    }
    for (int k = 0; k < D; k++) {
        double div = g[(1 + k) * D + k];
        P[0] -= dt * (p[0] * div + v[k] * g[k]);
        P[D + 1] -= dt * (gamma * p[D + 1] * div + v[k] * g[(D + 1) * D + k]);
        for (int q = 0; q < D; q++) P[1 + q] -= dt * v[k] * g[(1 + q) * D + k];
    }
    for (int q = 0; q < NF; q++) lS_in[q] += lS[q];
    if (P[0] < 0.0 || P[D + 1] < 0.0) return;
    double lv[D];
    for (int k = 0; k < D; k++) lv[k] = 0.;
    for (int k = 0; k < D; k++) {
        const double gP = g[(D + 1) * D + k];
        const double lt = -lS[1 + k];           // S_k -= t, t = dt gP / p0
        lg[(D + 1) * D + k] += lt * dt / p[0];
        lp[0] -= lt * dt * gP / (p[0] * p[0]);
    }
    for (int k = 0; k < D; k++) {
        const double div = g[(1 + k) * D + k];
        const double gr = g[k], gP = g[(D + 1) * D + k];
        double ldiv = 0.;
        // S0 -= dt (p0 div + v_k g_rho_k)
        lp[0] -= dt * lS[0] * div;
        ldiv -= dt * lS[0] * p[0];
        lv[k] -= dt * lS[0] * gr;
        lg[k] -= dt * lS[0] * v[k];
        // SE -= dt (gamma pE div + v_k gP_k)
        lp[D + 1] -= dt * lS[D + 1] * gamma * div;
        ldiv -= dt * lS[D + 1] * gamma * p[D + 1];
        lv[k] -= dt * lS[D + 1] * gP;
        lg[(D + 1) * D + k] -= dt * lS[D + 1] * v[k];
        // S_q -= dt v_k g_(1+q)k
        for (int q = 0; q < D; q++) {
            lv[k] -= dt * lS[1 + q] * g[(1 + q) * D + k];
            lg[(1 + q) * D + k] -= dt * lS[1 + q] * v[k];
        }
        lg[(1 + k) * D + k] += ldiv;
    }
    for (int k = 0; k < D; k++) lp[1 + k] += lv[k];
}

// ---- spatial extrapolation + GIZMO limiter -----------------------------------
// derivative of phi_plus / phi_minus (reconstruction.pyx:889-899) w.r.t.
// phi (= phi_max or phi_min) and delta1; sgn = +1 for plus, -1 for minus
PHD_HD inline double gizmo_bound(double phi, double delta1, int sgn, double& dphi, double& ddelta)
{
    const double shifted = phi + sgn * delta1;
    if (shifted * phi >= 0.) {
        dphi = 1.;
        ddelta = sgn;
        return shifted;
    }
    const double s = fabs(phi), h = 1 + delta1 / s;
    const double sg = phi > 0. ? 1. : -1.;
    dphi = 1. / h + (delta1 / s) / (h * h) * (phi * sg / s);
    ddelta = -phi / (s * h * h);
    return phi / h;
}

// reverse of boost_copy + spatial_states (+ gizmo_limit) for one face.
// lL, lR: lambda of the limited states; accumulates into the lambdas of the
// cell primitives pi, pj and gradients gi, gj.
template <int D>
PHD_HD inline void spatial_states_vjp(const double* com, double diff, const double* ci,
                                      const double* cj, const double* pi, const double* pj,
                                      const double* gi, const double* gj, const double* fv,
                                      int boost, int gizmo, const double* lL, const double* lR,
                                      double* lpi, double* lpj, double* lgi, double* lgj)
{
    constexpr int NF = D + 2;
    double sl[NF], sr[NF], Lx[NF], Rx[NF];
    boost_copy<D>(pi, pj, fv, boost, sl, sr);
    double sepi[D], sepj[D], si2 = 0., sj2 = 0.;
    for (int n = 0; n < NF; n++) { Lx[n] = sl[n]; Rx[n] = sr[n]; }
    for (int k = 0; k < D; k++) {
        sepi[k] = com[k] - ci[k];
        sepj[k] = com[k] - cj[k];
        si2 += sepi[k] * sepi[k];
        sj2 += sepj[k] * sepj[k];
        for (int n = 0; n < NF; n++) {
            Lx[n] += gi[n * D + k] * sepi[k];
            Rx[n] += gj[n * D + k] * sepj[k];
        }
    }
    // lambdas of sl, sr (cell states) and Lx, Rx (extrapolations)
    double lsl[NF], lsr[NF], lLx[NF], lRx[NF];
    for (int n = 0; n < NF; n++) { lsl[n] = 0.; lsr[n] = 0.; lLx[n] = 0.; lRx[n] = 0.; }
    if (!gizmo) {
        for (int n = 0; n < NF; n++) { lLx[n] = lL[n]; lRx[n] = lR[n]; }
    } else {
        // the negativity fallback replaces a whole side by its cell state;
        // recompute the limited values before it (gizmo_limit's loop)
        bool fallL = false, fallR = false;
        {
            double Lt[NF], Rt[NF];
            for (int n = 0; n < NF; n++) { Lt[n] = Lx[n]; Rt[n] = Rx[n]; }
            const double psi1 = 0.5, psi2 = 0.25;
            const double smi = sqrt(si2), smj = sqrt(sj2);
            for (int n = 0; n < NF; n++) {
                double delta1 = psi1 * fabs(sl[n] - sr[n]);
                double delta2 = psi2 * fabs(sl[n] - sr[n]);
                double phi_min = fmin(sl[n], sr[n]);
                double phi_max = fmax(sl[n], sr[n]);
                double phibar_l = sl[n] + smi / diff * (sr[n] - sl[n]);
                double phibar_r = sr[n] + smj / diff * (sl[n] - sr[n]);
                double phi_plus, phi_minus;
                if ((phi_max + delta1) * phi_max >= 0.) phi_plus = phi_max + delta1;
                else phi_plus = phi_max / (1 + delta1 / fabs(phi_max));
                if ((phi_min - delta1) * phi_min >= 0.) phi_minus = phi_min - delta1;
                else phi_minus = phi_min / (1 + delta1 / fabs(phi_min));
                if (pi[n] < pj[n]) {
                    Lt[n] = fmax(phi_minus, fmin(phibar_l + delta2, Lt[n]));
                    Rt[n] = fmin(phi_plus, fmax(phibar_r - delta2, Rt[n]));
                } else if (pi[n] > pj[n]) {
                    Lt[n] = fmin(phi_plus, fmax(phibar_l - delta2, Lt[n]));
                    Rt[n] = fmax(phi_minus, fmin(phibar_r + delta2, Rt[n]));
                } else {
                    Lt[n] = sl[n];
                    Rt[n] = sr[n];
                }
            }
            fallL = Lt[0] < 0.0 || Lt[D + 1] < 0.0;
            fallR = Rt[0] < 0.0 || Rt[D + 1] < 0.0;
        }
        const double smi = sqrt(si2), smj = sqrt(sj2);
        const double ai = smi / diff, aj = smj / diff;
        for (int n = 0; n < NF; n++) {
            const double a = sl[n], b = sr[n];
            const double dif = a - b, sgd = dif > 0. ? 1. : (dif < 0. ? -1. : 0.);
            const double delta1 = 0.5 * fabs(dif), delta2 = 0.25 * fabs(dif);
            // d delta / d a = c sgd, d delta / d b = -c sgd
            const double phi_min = fmin(a, b), phi_max = fmax(a, b);
            const bool min_is_a = !(b < a), max_is_a = !(a < b);
            double dpp, dpd, dmp, dmd;
            const double phi_plus = gizmo_bound(phi_max, delta1, 1, dpp, dpd);
            const double phi_minus = gizmo_bound(phi_min, delta1, -1, dmp, dmd);
            const double pbl = a + ai * (b - a), pbr = b + aj * (a - b);
            // accumulators for d/da, d/db of the selected expressions
            auto add_plus = [&](double lam) {   // phi_plus(phi_max(a,b), delta1(a,b))
                if (max_is_a) lsl[n] += lam * dpp; else lsr[n] += lam * dpp;
                lsl[n] += lam * dpd * 0.5 * sgd;
                lsr[n] -= lam * dpd * 0.5 * sgd;
            };
            auto add_minus = [&](double lam) {
                if (min_is_a) lsl[n] += lam * dmp; else lsr[n] += lam * dmp;
                lsl[n] += lam * dmd * 0.5 * sgd;
                lsr[n] -= lam * dmd * 0.5 * sgd;
            };
            auto add_pbl = [&](double lam, double sd2) {   // phibar_l + sd2*delta2
                lsl[n] += lam * (1. - ai) + lam * sd2 * 0.25 * sgd;
                lsr[n] += lam * ai - lam * sd2 * 0.25 * sgd;
            };
            auto add_pbr = [&](double lam, double sd2) {   // phibar_r + sd2*delta2
                lsr[n] += lam * (1. - aj) - lam * sd2 * 0.25 * sgd;
                lsl[n] += lam * aj + lam * sd2 * 0.25 * sgd;
            };
            const double lLn = fallL ? 0. : lL[n], lRn = fallR ? 0. : lR[n];
            if (fallL) lsl[n] += lL[n];
            if (fallR) lsr[n] += lR[n];
            if (pi[n] < pj[n]) {
                // L = max(phi_minus, min(pbl + delta2, Lx))
                const double in = fmin(pbl + delta2, Lx[n]);
                if (phi_minus > in) add_minus(lLn);
                else if (pbl + delta2 < Lx[n]) add_pbl(lLn, 1.);
                else lLx[n] += lLn;
                // R = min(phi_plus, max(pbr - delta2, Rx))
                const double inr = fmax(pbr - delta2, Rx[n]);
                if (phi_plus < inr) add_plus(lRn);
                else if (pbr - delta2 > Rx[n]) add_pbr(lRn, -1.);
                else lRx[n] += lRn;
            } else if (pi[n] > pj[n]) {
                // L = min(phi_plus, max(pbl - delta2, Lx))
                const double in = fmax(pbl - delta2, Lx[n]);
                if (phi_plus < in) add_plus(lLn);
                else if (pbl - delta2 > Lx[n]) add_pbl(lLn, -1.);
                else lLx[n] += lLn;
                // R = max(phi_minus, min(pbr + delta2, Rx))
                const double inr = fmin(pbr + delta2, Rx[n]);
                if (phi_minus > inr) add_minus(lRn);
                else if (pbr + delta2 < Rx[n]) add_pbr(lRn, 1.);
                else lRx[n] += lRn;
            } else {
                lsl[n] += lLn;
                lsr[n] += lRn;
            }
        }
    }
    // Lx = sl + gi . sepi, Rx = sr + gj . sepj; sl, sr = boosted copies of pi, pj
    for (int n = 0; n < NF; n++) {
        lpi[n] += lsl[n] + lLx[n];
        lpj[n] += lsr[n] + lRx[n];
        for (int k = 0; k < D; k++) {
            lgi[n * D + k] += lLx[n] * sepi[k];
            lgj[n * D + k] += lRx[n] * sepj[k];
        }
    }
}

// ---- eq. 21 gradient + AREPO limiter, one real cell ----------------------------
// Nbr supplies the cell's face list in order: deg(), and for entry e
// valid(e), row(e, r) = [x_j | prim_j], geo(e, com, area).  Given lgrad =
// lambda of the limited gradient (component n*D+k), accumulates lambda of
// the cell's own primitives into lself and writes, per entry, the lambda of
// the neighbour's primitives through out(e, lam).
template <int D, class Nbr>
PHD_HD inline void grad_cell_vjp(const Nbr& nb, const double* pv, const double* xi,
                                 const double* cx, double vol, const double* lgrad,
                                 double* lself)
{
    constexpr int NF = D + 2;
    const int deg = nb.deg();
    double df[NF * D], pmax[NF], pmin[NF];
    int amax[NF], amin[NF];
    for (int n = 0; n < NF; n++) {
        pmax[n] = pmin[n] = pv[n];
        amax[n] = amin[n] = -1;
        for (int k = 0; k < D; k++) df[n * D + k] = 0.0;
    }
    for (int e = 0; e < deg; e++) {
        if (!nb.valid(e)) continue;
        double r[8], com[3], area;
        nb.row(e, r);
        nb.geo(e, com, area);
        double cfx[D], dr[D], rr = 0.0;
        for (int k = 0; k < D; k++) {
            cfx[k] = com[k] - 0.5 * (xi[k] + r[k]);
            dr[k] = xi[k] - r[k];
            rr += dr[k] * dr[k];
        }
        rr = sqrt(rr);
        const double rv = rr * vol;
        for (int n = 0; n < NF; n++) {
            const double pj = r[D + n];
            if (pj > pmax[n]) { pmax[n] = pj; amax[n] = e; }
            if (pj < pmin[n]) { pmin[n] = pj; amin[n] = e; }
            const double d_dif = pj - pv[n], d_sum = pj + pv[n];
            for (int k = 0; k < D; k++) df[n * D + k] += area * (d_dif * cfx[k] - 0.5 * d_sum * dr[k]) / rv;
        }
    }
    // limiter: alpha_n = min(1, (pmax-pv)/dpos, (pmin-pv)/dneg), each >= 0
    double dpos[NF], dneg[NF];
    int epos[NF], eneg[NF];
    for (int n = 0; n < NF; n++) { dpos[n] = 0.0; dneg[n] = 0.0; epos[n] = eneg[n] = -1; }
    for (int e = 0; e < deg; e++) {
        double com[3], area;
        nb.geo(e, com, area);
        double rel[D];
        for (int k = 0; k < D; k++) rel[k] = com[k] - cx[k];
        for (int n = 0; n < NF; n++) {
            double dphi = 0.0;
            for (int k = 0; k < D; k++) dphi += df[n * D + k] * rel[k];
            if (dphi > 0 && dphi > dpos[n]) { dpos[n] = dphi; epos[n] = e; }
            else if (dphi < 0 && dphi < dneg[n]) { dneg[n] = dphi; eneg[n] = e; }
        }
    }
    double ldf[NF * D], lnb_max[NF], lnb_min[NF];
    for (int n = 0; n < NF; n++) {
        double alpha = 1.0;
        int which = 0;
        double qp = 0.0, qn = 0.0;
        if (dpos[n] > 0) {
            qp = fmax((pmax[n] - pv[n]) / dpos[n], 0.);
            if (qp < alpha) { alpha = qp; which = 1; }
        }
        if (dneg[n] < 0) {
            qn = fmax((pmin[n] - pv[n]) / dneg[n], 0.);
            if (qn < alpha) { alpha = qn; which = 2; }
        }
        double lalpha = 0.0;
        for (int k = 0; k < D; k++) {
            lalpha += lgrad[n * D + k] * df[n * D + k];
            ldf[n * D + k] = alpha * lgrad[n * D + k];
        }
        lnb_max[n] = lnb_min[n] = 0.0;
        if (which != 0 && alpha > 0.) {
            const double dd = which == 1 ? dpos[n] : dneg[n];
            const int ee = which == 1 ? epos[n] : eneg[n];
            const double l = lalpha / dd;
            // alpha = (phi_ext - pv)/dd
            if (which == 1) lnb_max[n] = l; else lnb_min[n] = l;
            lself[n] -= l;
            const double ldd = -lalpha * alpha / dd;
            double com[3], area;
            nb.geo(ee, com, area);
            for (int k = 0; k < D; k++) ldf[n * D + k] += ldd * (com[k] - cx[k]);
        }
        // phi_max / phi_min may be the cell's own value
        if (amax[n] < 0) lself[n] += lnb_max[n];
        if (amin[n] < 0) lself[n] += lnb_min[n];
    }
    for (int e = 0; e < deg; e++) {
        double lam[NF];
        for (int n = 0; n < NF; n++) lam[n] = 0.0;
        if (nb.valid(e)) {
            double r[8], com[3], area;
            nb.row(e, r);
            nb.geo(e, com, area);
            double cfx[D], dr[D], rr = 0.0;
            for (int k = 0; k < D; k++) {
                cfx[k] = com[k] - 0.5 * (xi[k] + r[k]);
                dr[k] = xi[k] - r[k];
                rr += dr[k] * dr[k];
            }
            rr = sqrt(rr);
            const double rv = rr * vol;
            for (int n = 0; n < NF; n++) {
                for (int k = 0; k < D; k++) {
                    lam[n] += ldf[n * D + k] * area * (cfx[k] - 0.5 * dr[k]) / rv;
                    lself[n] += ldf[n * D + k] * area * (-cfx[k] - 0.5 * dr[k]) / rv;
                }
                if (amax[n] == e) lam[n] += lnb_max[n];
                if (amin[n] == e) lam[n] += lnb_min[n];
            }
        }
        nb.out(e, lam);
    }
}

} // namespace phd
