/*
 * oracle_dual.cpp -- forward-mode automatic differentiation of the CPU
 * oracle, TEST INFRASTRUCTURE ONLY (same rules as phd_oracle.c: only tests/
 * load it).
 *
 * phd_oracle.c is compiled a second time, as C++, with every `double`
 * replaced by a dual number (value, tangent).  Each arithmetic operation and
 * each math function propagates the tangent by its derivative; comparisons,
 * fmin/fmax and branches act on the values only, so the tangent follows the
 * branch the values take.  One pass therefore gives the exact directional
 * derivative J v of the oracle's computation (to rounding), for the branch
 * pattern of the given state: a reference gradient with no truncation error,
 * independent of the hand-written reverse mode in adjoint_dev.h.
 *
 * ord_rollout runs the static-mesh step the adjoint differentiates
 * (StaticMeshMUSCLHancock.evolve_timestep, integrate.py:349-378, face and
 * generator velocities zero): conserved fields from the initial primitives,
 * ghost refresh + EOS, then per step orc_core_step (AREPO + HLLC) followed by
 * the ghost refresh (domain_manager.pyx:759-774, boundary.pyx:336-348) and
 * IdealGas.primitive_from_conservative.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <stdlib.h>
#include <vector>

namespace ad {

struct Dual {
    double v, d;
    Dual() : v(0.0), d(0.0) {}
    Dual(double x) : v(x), d(0.0) {}
    Dual(double x, double t) : v(x), d(t) {}
    Dual& operator+=(const Dual& b) { d += b.d; v += b.v; return *this; }
    Dual& operator-=(const Dual& b) { d -= b.d; v -= b.v; return *this; }
    Dual& operator*=(const Dual& b) { d = d * b.v + v * b.d; v *= b.v; return *this; }
    Dual& operator/=(const Dual& b)
    {
        const double q = v / b.v;
        d = (d - q * b.d) / b.v;
        v = q;
        return *this;
    }
};

inline Dual operator+(Dual a, const Dual& b) { return a += b; }
inline Dual operator-(Dual a, const Dual& b) { return a -= b; }
inline Dual operator*(Dual a, const Dual& b) { return a *= b; }
inline Dual operator/(Dual a, const Dual& b) { return a /= b; }
inline Dual operator-(const Dual& a) { return Dual(-a.v, -a.d); }
inline Dual operator+(const Dual& a) { return a; }
inline bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
inline bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
inline bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
inline bool operator>=(const Dual& a, const Dual& b) { return a.v >= b.v; }
inline bool operator==(const Dual& a, const Dual& b) { return a.v == b.v; }
inline bool operator!=(const Dual& a, const Dual& b) { return a.v != b.v; }

inline Dual sqrt(const Dual& a)
{
    const double r = ::sqrt(a.v);
    return Dual(r, a.d / (2.0 * r));
}
inline Dual fabs(const Dual& a) { return a.v < 0.0 ? -a : (a.v > 0.0 ? a : Dual(a.v, 0.0)); }
// C99 fmax/fmin: a NaN operand loses; otherwise the larger / smaller value
inline Dual fmax(const Dual& a, const Dual& b)
{
    if (a.v != a.v) return b;
    if (b.v != b.v) return a;
    return a.v >= b.v ? a : b;
}
inline Dual fmin(const Dual& a, const Dual& b)
{
    if (a.v != a.v) return b;
    if (b.v != b.v) return a;
    return a.v <= b.v ? a : b;
}
inline Dual pow(const Dual& a, const Dual& b)
{
    const double r = ::pow(a.v, b.v);
    double t = 0.0;
    if (a.d != 0.0) t += b.v * ::pow(a.v, b.v - 1.0) * a.d;
    if (b.d != 0.0) t += ::log(a.v) * r * b.d;
    return Dual(r, t);
}
inline Dual exp(const Dual& a)
{
    const double r = ::exp(a.v);
    return Dual(r, r * a.d);
}
inline Dual log(const Dual& a) { return Dual(::log(a.v), a.d / a.v); }

} // namespace ad

namespace dual {
using ad::Dual;
#define double Dual
#include "phd_oracle.c"
#undef double
} // namespace dual

using ad::Dual;

namespace {

std::vector<Dual> to_dual(const double* v, const double* t, long n)
{
    std::vector<Dual> out(n);
    for (long i = 0; i < n; i++) out[i] = Dual(v[i], t ? t[i] : 0.0);
    return out;
}

struct Fields {
    std::vector<std::vector<Dual>> f;
    std::vector<Dual*> p;
    Fields(int m, long n) : f(m, std::vector<Dual>(n)), p(m)
    {
        for (int q = 0; q < m; q++) p[q] = f[q].data();
    }
    Dual* const* ptr() { return p.data(); }
};

// ghost conserved fields from their images, momenta negated on each axis
// where a reflective exterior ghost lies outside the box; then the EOS
void refresh(int dim, long n, Dual gamma, int boundary, const int8_t* tag, const int8_t* type,
             const int32_t* map, const std::vector<std::vector<Dual>>& x, const double* lo,
             const double* hi, Dual* vol, Fields& cons, Fields& prim)
{
    for (long g = 0; g < n; g++) {
        if (tag[g] != 1) continue;
        const int s = map[g];
        for (int q = 0; q < dim + 2; q++) cons.f[q][g] = cons.f[q][s];
        if (boundary == 0 && type[g] == 2)
            for (int k = 0; k < dim; k++) {
                if (x[k][g].v < lo[k]) cons.f[1 + k][g] = -cons.f[1 + k][g];
                if (x[k][g].v > hi[k]) cons.f[1 + k][g] = -cons.f[1 + k][g];
            }
    }
    dual::orc_prim_from_cons(dim, (int)n, gamma, cons.ptr(), vol, prim.ptr());
}

} // namespace

extern "C" int ord_rollout(int dim, int n, int nf, double gamma, int n_steps, const double* dts,
                           int boost, int gizmo, int boundary, const int8_t* tag,
                           const int8_t* type, const int32_t* map, double* const* x,
                           double* const* dcx, const double* vol, const double* area,
                           double* const* fij, double* const* nx, const int32_t* pair_i,
                           const int32_t* pair_j, const int32_t* nbr_off,
                           const int32_t* nbr_face, const double* lo, const double* hi,
                           double* const* prim0, double* const* dprim0, double* const* prim_out,
                           double* const* dprim_out)
{
    const int NF = dim + 2;
    std::vector<std::vector<Dual>> X(dim), DC(dim), FIJ(dim), NX(dim), FV(dim);
    std::vector<Dual*> px(dim), pdc(dim), pf(dim), pn(dim), pv(dim);
    for (int k = 0; k < dim; k++) {
        X[k] = to_dual(x[k], nullptr, n);
        DC[k] = to_dual(dcx[k], nullptr, n);
        FIJ[k] = to_dual(fij[k], nullptr, nf);
        NX[k] = to_dual(nx[k], nullptr, nf);
        FV[k] = std::vector<Dual>(nf);
        px[k] = X[k].data();
        pdc[k] = DC[k].data();
        pf[k] = FIJ[k].data();
        pn[k] = NX[k].data();
        pv[k] = FV[k].data();
    }
    std::vector<Dual> VOL = to_dual(vol, nullptr, n), AREA = to_dual(area, nullptr, nf);
    Dual LO[3], HI[3];
    for (int k = 0; k < 3; k++) { LO[k] = k < dim ? lo[k] : 0.0; HI[k] = k < dim ? hi[k] : 0.0; }
    Fields prim(NF, n), cons(NF, n), grad(dim * NF, n), pl(NF, nf), pr(NF, nf), flux(NF, nf);
    for (int q = 0; q < NF; q++)
        for (long i = 0; i < n; i++) prim.f[q][i] = Dual(prim0[q][i], dprim0 ? dprim0[q][i] : 0.0);
    const Dual G = gamma;
    dual::orc_cons_from_prim(dim, n, G, prim.ptr(), VOL.data(), cons.ptr());
    refresh(dim, n, G, boundary, tag, type, map, X, lo, hi, VOL.data(), cons, prim);
    for (int s = 0; s < n_steps; s++) {
        int32_t bad = -1;
        const int rc = dual::orc_core_step(dim, n, nf, G, Dual(dts[s]), 1, boost, gizmo, 0,
                                           boundary, tag, type, map, px.data(), pdc.data(),
                                           VOL.data(), prim.ptr(), cons.ptr(), AREA.data(),
                                           pf.data(), pn.data(), pv.data(), pair_i, pair_j,
                                           nbr_off, nbr_face, LO, HI, grad.ptr(), pl.ptr(),
                                           pr.ptr(), flux.ptr(), &bad);
        if (rc) return rc;
        refresh(dim, n, G, boundary, tag, type, map, X, lo, hi, VOL.data(), cons, prim);
    }
    for (int q = 0; q < NF; q++)
        for (long i = 0; i < n; i++) {
            prim_out[q][i] = prim.f[q][i].v;
            if (dprim_out) dprim_out[q][i] = prim.f[q][i].d;
        }
    return 0;
}
