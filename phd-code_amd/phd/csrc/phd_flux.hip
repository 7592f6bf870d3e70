// phd_flux.hip -- MI355X (gfx950) kernels for the Voronoi face-flux pipeline
// and the C ABI declared in include/phd_flux.h.
//
// Data layout in HBM
//   cells  : structure-of-arrays, one double array per named field, reals
//            first then ghosts (the reference's particle container order).
//   grad   : cell-major rows, grad[c*ld + n*D + k] = d prim_n / d x_k.  With
//            ld = 16 (3D) / 8 (2D) a row is one 128 B / 64 B line, so the face
//            kernel fetches a neighbour's whole gradient in one line.
//   faces  : structure-of-arrays, sorted by pair_i (emission order of
//            Tess2d/Tess3d::extract_geometry, tess.cpp:363, tess3.cpp:509).
//   lists  : CSR of face ids per cell, ascending (tess.cpp:379-382).
//
// Arithmetic: FP64 VALU only.  The library is compiled with
// -ffp-contract=off; BITWISE kernels evaluate each expression in the
// reference's order (correctly rounded +,-,*,/,sqrt), so results are
// bit-identical to the reference's x86-64 build except where a libm pow()
// result differs in the last place.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>
#include "phd_flux.h"

#define PHD_BLOCK 256

struct phd_handle {
    int device;
    hipStream_t stream;
    int mode;
    unsigned long long* d_err;   // [0] update key, [1] neighbor cell, [2] exact face, [3] dt bits
    unsigned long long* h_err;   // pinned host mirror
};

namespace {

constexpr unsigned long long NOERR = ~0ull;
constexpr unsigned long long INF_BITS = 0x7ff0000000000000ull;

struct FaceArrays { double* q[5]; };

__device__ __forceinline__ void record_min(unsigned long long* p, unsigned long long v)
{
    atomicMin(p, v);
}

__global__ void k_reset_err(unsigned long long* e)
{
    if (threadIdx.x < 3) e[threadIdx.x] = NOERR;
    if (threadIdx.x == 3) e[3] = INF_BITS;
}

inline unsigned nblocks(long long n) { return (unsigned)((n + PHD_BLOCK - 1) / PHD_BLOCK); }

// ---------------------------------------------------------------------------
// equation of state  (equation_state.pyx:43-121)
template <int D>
__global__ __launch_bounds__(PHD_BLOCK) void k_cons_from_prim(phd_cells c, double gamma)
{
    int i = blockIdx.x * PHD_BLOCK + threadIdx.x;
    if (i >= c.n_cells) return;
    double d = c.prim[0][i], vol = c.volume[i];
    double m = d * vol;
    c.cons[0][i] = m;
    double v_sq = 0.;
#pragma unroll
    for (int k = 0; k < D; k++) {
        double v = c.prim[1 + k][i];
        c.cons[1 + k][i] = v * m;
        v_sq += v * v;
    }
    c.cons[D + 1][i] = (.5 * d * v_sq + c.prim[D + 1][i] / (gamma - 1.)) * vol;
}

template <int D>
__global__ __launch_bounds__(PHD_BLOCK) void k_prim_from_cons(phd_cells c, double gamma)
{
    int i = blockIdx.x * PHD_BLOCK + threadIdx.x;
    if (i >= c.n_cells) return;
    double m = c.cons[0][i], vol = c.volume[i];
    double d = m / vol;
    c.prim[0][i] = d;
    double v_sq = 0.;
#pragma unroll
    for (int k = 0; k < D; k++) {
        double v = c.cons[1 + k][i] / m;
        c.prim[1 + k][i] = v;
        v_sq += v * v;
    }
    c.prim[D + 1][i] = (c.cons[D + 1][i] / vol - .5 * d * v_sq) * (gamma - 1.);
}

// cell radius (riemann.pyx:155-158, mesh.pyx:414-417); pi is numpy.pi
template <int D>
__device__ __forceinline__ double cell_radius(double vol)
{
    if (D == 2) return sqrt(vol / 3.141592653589793);
    return pow(3.0 * vol / (4.0 * 3.141592653589793), 1.0 / 3.0);
}

// ---------------------------------------------------------------------------
// CFL time step  (riemann.pyx:120-188): min over real cells of R/c (boost)
// or R/(c+|v|).  min is order independent, so a wave-level min followed by
// an atomic min on the IEEE bit pattern (positive doubles order as unsigned
// integers) is exact and deterministic.
template <int D>
__global__ __launch_bounds__(PHD_BLOCK) void k_cfl(phd_cells c, double gamma, int boost,
                                                   unsigned long long* err)
{
    int i = blockIdx.x * PHD_BLOCK + threadIdx.x;
    double v = __longlong_as_double((long long)INF_BITS);
    if (i < c.n_cells && c.tag[i] == 0) {
        double d = c.prim[0][i], p = c.prim[D + 1][i];
        double cs = sqrt(gamma * p / d);
        double R = cell_radius<D>(c.volume[i]);
        if (boost) {
            v = R / cs;
        } else {
            double vsq = 0.0;
#pragma unroll
            for (int k = 0; k < D; k++) { double u = c.prim[1 + k][i]; vsq += u * u; }
            v = R / (cs + sqrt(vsq));
        }
        if (!(v >= 0.0)) v = __longlong_as_double((long long)INF_BITS); // fmin ignores NaN
    }
    // wavefront min (64 lanes) via cross-lane shuffles
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
    __shared__ double wmin[PHD_BLOCK / 64];
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) wmin[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = wmin[0];
        for (int w = 1; w < PHD_BLOCK / 64; w++) b = fmin(b, wmin[w]);
        record_min(&err[3], (unsigned long long)__double_as_longlong(b));
    }
}

// ---------------------------------------------------------------------------
// generator velocities (mesh.pyx:363-426), all cells
template <int D>
__global__ __launch_bounds__(PHD_BLOCK) void k_gen_vel(phd_cells c, double gamma, double eta,
                                                       int regularize)
{
    int i = blockIdx.x * PHD_BLOCK + threadIdx.x;
    if (i >= c.n_cells) return;
    double w[D];
#pragma unroll
    for (int k = 0; k < D; k++) w[k] = c.prim[1 + k][i];
    if (regularize) {
        double cs = sqrt(gamma * c.prim[D + 1][i] / c.prim[0][i]);
        double dc[D], d = 0.0;
#pragma unroll
        for (int k = 0; k < D; k++) { dc[k] = c.dcom[k][i]; d += dc[k] * dc[k]; }
        d = sqrt(d);
        double R = cell_radius<D>(c.volume[i]);
        if ((0.9 <= d / (eta * R)) && (d / (eta * R) < 1.1)) {
#pragma unroll
            for (int k = 0; k < D; k++)
                w[k] += cs * dc[k] * (d - 0.9 * eta * R) / (d * 0.2 * eta * R);
        } else if (1.1 <= d / (eta * R)) {
#pragma unroll
            for (int k = 0; k < D; k++) w[k] += cs * dc[k] / d;
        }
    }
#pragma unroll
    for (int k = 0; k < D; k++) c.w[k][i] = w[k];
}

// face velocities (mesh.pyx:428-477), all faces
template <int D>
__global__ __launch_bounds__(PHD_BLOCK) void k_face_vel(phd_cells c, phd_faces f)
{
    int m = blockIdx.x * PHD_BLOCK + threadIdx.x;
    if (m >= f.n_faces) return;
    int i = f.pair_i[m], j = f.pair_j[m];
    double factor = 0.0, denom = 0.0, wi[D], wj[D], dx[D];
#pragma unroll
    for (int k = 0; k < D; k++) {
        double xi = c.x[k][i], xj = c.x[k][j];
        wi[k] = c.w[k][i];
        wj[k] = c.w[k][j];
        dx[k] = xj - xi;
        factor += (wi[k] - wj[k]) * (f.com[k][m] - 0.5 * (xi + xj));
        denom += dx[k] * dx[k];
    }
    factor /= denom;
#pragma unroll
    for (int k = 0; k < D; k++) f.velocity[k][m] = 0.5 * (wi[k] + wj[k]) + factor * dx[k];
}

// ---------------------------------------------------------------------------
// gradients + slope limiter (reconstruction.pyx:563-737), one thread per cell.
//
// The AREPO limiter (683-703) takes alpha_n = min(1, min_faces fmax(psi,0))
// with psi = (phi_max - phi_i)/dphi for dphi > 0 and (phi_min - phi_i)/dphi
// for dphi < 0.  The numerators are fixed per cell (>= 0 and <= 0) and a
// correctly rounded quotient is monotone in its divisor, so the minimum over
// faces is reached at the largest positive dphi and at the most negative
// dphi: two divisions per field instead of one per face, same bits.
template <int D, int LIM>
__global__ __launch_bounds__(PHD_BLOCK) void k_gradients(phd_cells c, phd_faces f,
                                                         double* __restrict__ grad, int ld,
                                                         unsigned long long* err)
{
    constexpr int NF = D + 2;
    int i = blockIdx.x * PHD_BLOCK + threadIdx.x;
    if (i >= c.n_cells || c.tag[i] != 0) return;

    double xi[D], cx[D], pv[NF], pmax[NF], pmin[NF], df[NF * D];
#pragma unroll
    for (int k = 0; k < D; k++) { xi[k] = c.x[k][i]; cx[k] = xi[k] + c.dcom[k][i]; }
    const double vol = c.volume[i];
#pragma unroll
    for (int n = 0; n < NF; n++) {
        pv[n] = c.prim[n][i];
        pmax[n] = pmin[n] = pv[n];
#pragma unroll
        for (int k = 0; k < D; k++) df[n * D + k] = 0.0;
    }
    const int beg = f.nbr_off[i], end = f.nbr_off[i + 1];
    for (int m = beg; m < end; m++) {
        const int fid = f.nbr_face[m];
        const double area = f.area[fid];
        const int a = f.pair_i[fid], b = f.pair_j[fid];
        int j;
        if (i == a) j = b;
        else if (i == b) j = a;
        else { record_min(&err[1], (unsigned long long)i); return; }
        double cfx[D], dr[D], r = 0.0;
#pragma unroll
        for (int k = 0; k < D; k++) {
            double xj = c.x[k][j];
            cfx[k] = f.com[k][fid] - 0.5 * (xi[k] + xj);
            dr[k] = xi[k] - xj;
            r += dr[k] * dr[k];
        }
        r = sqrt(r);
        const double rv = r * vol;
#pragma unroll
        for (int n = 0; n < NF; n++) {
            double pj = c.prim[n][j];
            pmax[n] = fmax(pmax[n], pj);
            pmin[n] = fmin(pmin[n], pj);
            double d_dif = pj - pv[n];
            double d_sum = pj + pv[n];
#pragma unroll
            for (int k = 0; k < D; k++)
                df[n * D + k] += area * (d_dif * cfx[k] - 0.5 * d_sum * dr[k]) / rv;
        }
    }

    double alpha[NF];
#pragma unroll
    for (int n = 0; n < NF; n++) alpha[n] = 1.0;
    if (LIM == PHD_LIMITER_AREPO) {
        double dpos[NF], dneg[NF];
#pragma unroll
        for (int n = 0; n < NF; n++) { dpos[n] = 0.0; dneg[n] = 0.0; }
        for (int m = beg; m < end; m++) {
            const int fid = f.nbr_face[m];
            double rel[D];
#pragma unroll
            for (int k = 0; k < D; k++) rel[k] = f.com[k][fid] - cx[k];
#pragma unroll
            for (int n = 0; n < NF; n++) {
                double dphi = 0.0;
#pragma unroll
                for (int k = 0; k < D; k++) dphi += df[n * D + k] * rel[k];
                if (dphi > 0) dpos[n] = fmax(dpos[n], dphi);
                else if (dphi < 0) dneg[n] = fmin(dneg[n], dphi);
            }
        }
#pragma unroll
        for (int n = 0; n < NF; n++) {
            if (dpos[n] > 0) alpha[n] = fmin(alpha[n], fmax((pmax[n] - pv[n]) / dpos[n], 0.));
            if (dneg[n] < 0) alpha[n] = fmin(alpha[n], fmax((pmin[n] - pv[n]) / dneg[n], 0.));
        }
    } else {
        // TESS limiter (705-731); as in the reference only alpha[0] changes.
        for (int n = 0; n < NF; n++) {
            int j = i;
            for (int m = beg; m < end; m++) {
                const int fid = f.nbr_face[m];
                const int a = f.pair_i[fid], b = f.pair_j[fid];
                if (i == a) j = b;
                else if (i == b) j = a;
                double dphi = 0.0;
#pragma unroll
                for (int k = 0; k < D; k++) dphi += df[n * D + k] * (f.com[k][fid] - cx[k]);
                double psi;
                if (dphi > 0.0 || dphi < 0.0) {
                    double q = (c.prim[n][j] - pv[n]) / dphi;
                    psi = (0. > q) ? 0. : q;
                } else {
                    psi = 1.0;
                }
                alpha[0] = fmin(alpha[0], fmax(psi, 0.));
            }
        }
    }
    double* row = grad + (size_t)i * ld;
#pragma unroll
    for (int n = 0; n < NF; n++)
#pragma unroll
        for (int k = 0; k < D; k++) row[n * D + k] = alpha[n] * df[n * D + k];
}

// ghost gradients (domain_manager.pyx:819-837 + boundary.pyx:265-307):
// one thread per (cell, gradient component)
template <int D>
__global__ __launch_bounds__(PHD_BLOCK) void k_ghost_grad(phd_cells c, int reflect, double lo0,
                                                          double lo1, double lo2, double hi0,
                                                          double hi1, double hi2,
                                                          double* __restrict__ grad, int ld)
{
    constexpr int G = D * (D + 2);
    long long t = (long long)blockIdx.x * PHD_BLOCK + threadIdx.x;
    int i = (int)(t / G), g = (int)(t % G);
    if (i >= c.n_cells || c.tag[i] != 1) return;
    double v = grad[(size_t)c.map[i] * ld + g];
    // velocity group entry dv[dim*q + q] = grad[(1+q)*D + q]
    if (reflect && c.type[i] == 2 && g >= D && g < (D + 1) * D) {
        int q = g / D - 1, k = g % D;
        if (q == k) {
            double lo = q == 0 ? lo0 : (q == 1 ? lo1 : lo2);
            double hi = q == 0 ? hi0 : (q == 1 ? hi1 : hi2);
            double xq = c.x[q][i];
            if (xq < lo) v *= -1;
            if (xq > hi) v *= -1;
        }
    }
    grad[(size_t)i * ld + g] = v;
}

// ---------------------------------------------------------------------------
// face states.  Spatial: PieceWiseLinear.add_spatial (741-921); temporal:
// add_temporal (923-1072); constant: PieceWiseConstant.compute_states
// (254-352).  L/R hold the primitive group in the face frame when boosted.
template <int D>
__device__ __forceinline__ void load_grad_row(const double* __restrict__ row, double* g)
{
    constexpr int G = D * (D + 2);
    const double2* r2 = reinterpret_cast<const double2*>(row);
#pragma unroll
    for (int q = 0; q < G / 2; q++) { double2 v = r2[q]; g[2 * q] = v.x; g[2 * q + 1] = v.y; }
    if (G & 1) g[G - 1] = row[G - 1];
}

template <int D>
__device__ __forceinline__ void copy_states(const phd_cells& c, const double* fv, int boost,
                                            int i, int j, double* pi, double* pj, double* L,
                                            double* R)
{
    constexpr int NF = D + 2;
#pragma unroll
    for (int n = 0; n < NF; n++) {
        pi[n] = c.prim[n][i];
        pj[n] = c.prim[n][j];
        L[n] = pi[n];
        R[n] = pj[n];
    }
    if (boost) {
#pragma unroll
        for (int k = 0; k < D; k++) { L[1 + k] = pi[1 + k] - fv[k]; R[1 + k] = pj[1 + k] - fv[k]; }
    }
}

template <int D>
__device__ __forceinline__ void spatial_states(const phd_cells& c, const phd_faces& f, int m,
                                               int i, int j, const double* pi, const double* pj,
                                               const double* gi, const double* gj, int gizmo,
                                               double* L, double* R)
{
    constexpr int NF = D + 2;
    const double psi1 = 0.5, psi2 = 0.25;
    double sl[NF], sr[NF];
#pragma unroll
    for (int n = 0; n < NF; n++) { sl[n] = L[n]; sr[n] = R[n]; }
    double diff_mag = 0.0, sepi_mag = 0.0, sepj_mag = 0.0;
#pragma unroll
    for (int k = 0; k < D; k++) {
        double com = f.com[k][m];
        double xik = c.x[k][i], xjk = c.x[k][j];
        double sepi = com - (xik + c.dcom[k][i]);
        double sepj = com - (xjk + c.dcom[k][j]);
        sepi_mag += sepi * sepi;
        sepj_mag += sepj * sepj;
        diff_mag += (xjk - xik) * (xjk - xik);
#pragma unroll
        for (int n = 0; n < NF; n++) {
            L[n] += gi[n * D + k] * sepi;
            R[n] += gj[n * D + k] * sepj;
        }
    }
    sepi_mag = sqrt(sepi_mag);
    sepj_mag = sqrt(sepj_mag);
    diff_mag = sqrt(diff_mag);
    if (!gizmo) return;
#pragma unroll
    for (int n = 0; n < NF; n++) {
        double delta1 = psi1 * fabs(sl[n] - sr[n]);
        double delta2 = psi2 * fabs(sl[n] - sr[n]);
        double phi_min = fmin(sl[n], sr[n]);
        double phi_max = fmax(sl[n], sr[n]);
        double phibar_l = sl[n] + sepi_mag / diff_mag * (sr[n] - sl[n]);
        double phibar_r = sr[n] + sepj_mag / diff_mag * (sl[n] - sr[n]);
        double phi_plus, phi_minus;
        if ((phi_max + delta1) * phi_max >= 0.) phi_plus = phi_max + delta1;
        else phi_plus = phi_max / (1 + delta1 / fabs(phi_max));
        if ((phi_min - delta1) * phi_min >= 0.) phi_minus = phi_min - delta1;
        else phi_minus = phi_min / (1 + delta1 / fabs(phi_min));
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
    if (L[0] < 0.0 || L[D + 1] < 0.0) {
#pragma unroll
        for (int n = 0; n < NF; n++) L[n] = sl[n];
    }
    if (R[0] < 0.0 || R[D + 1] < 0.0) {
#pragma unroll
        for (int n = 0; n < NF; n++) R[n] = sr[n];
    }
}

// one side of add_temporal (1018-1072)
template <int D>
__device__ __forceinline__ void temporal_side(const double* p, const double* g, const double* fv,
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

// ---------------------------------------------------------------------------
// Riemann solvers in the face frame (riemann.pyx)
__device__ __forceinline__ void get_waves(double dl, double ul, double pl, double dr, double ur,
                                          double pr, double gamma, double& sl, double& sc,
                                          double& sr)
{
    // HLL.get_waves (riemann.pyx:375-504)
    const double Q = 2., c_floor = 1.0E-10;
    double cl = fmax(sqrt(gamma * pl / dl), c_floor);
    double cr = fmax(sqrt(gamma * pr / dr), c_floor);
    double d_avg = .5 * (dl + dr);
    double c_avg = .5 * (cl + cr);
    double p_star = fmax(0., .5 * (pl + pr) + .5 * (ul - ur) * d_avg * c_avg);
    double p_min = fmin(pl, pr);
    double p_max = fmax(pl, pr);
    if (((p_max / p_min) < Q) && ((p_min < p_star) && (p_star < p_max))) {
        // PVRS: p_star kept
    } else if (p_star <= p_min) {
        double z = (gamma - 1.) / (2. * gamma);
        double plr = pow(pl / pr, z);
        double u_star = (plr * ul / cl + ur / cr + 2. * (plr - 1.) / (gamma - 1.)) /
                        (plr / cl + 1. / cr);
        p_star = .5 * pl * pow(1. + (gamma - 1.) * (ul - u_star) / (2. * cl), 1. / z);
        p_star += .5 * pr * pow(1. + (gamma - 1.) * (u_star - ur) / (2. * cr), 1. / z);
    } else {
        double Al = 2. / ((gamma + 1.) * dl);
        double Ar = 2. / ((gamma + 1.) * dr);
        double Bl = pl * ((gamma - 1.) / (gamma + 1.));
        double Br = pr * ((gamma - 1.) / (gamma + 1.));
        double gl = sqrt(Al / (p_star + Bl));
        double gr = sqrt(Ar / (p_star + Br));
        p_star = (gl * pl + gr * pr - (ur - ul)) / (gl + gr);
    }
    if (p_star <= pl) sl = ul - cl;
    else sl = ul - cl * sqrt(1. + ((gamma + 1.) / (2. * gamma)) * (p_star / pl - 1.));
    if (p_star <= pr) sr = ur + cr;
    else sr = ur + cr * sqrt(1. + ((gamma + 1.) / (2. * gamma)) * (p_star / pr - 1.));
    sc = (pr - pl + dl * ul * (sl - ul) - dr * ur * (sr - ur)) / (dl * (sl - ul) - dr * (sr - ur));
}

template <int D>
__device__ __forceinline__ void side_flux(const double* S, const double* n, double Vn, double vsq,
                                          double wn, double gamma, double* F)
{
    // pure upwind state flux in the moving frame (riemann.pyx:624-631, 672-679)
    double d = S[0], p = S[D + 1];
    F[0] = d * (Vn - wn);
    F[D + 1] = (0.5 * d * vsq + p / (gamma - 1.0)) * (Vn - wn) + p * Vn;
#pragma unroll
    for (int k = 0; k < D; k++) F[1 + k] = d * S[1 + k] * (Vn - wn) + p * n[k];
}

template <int D, int SOLVER>
__device__ __forceinline__ void riemann(const double* L, const double* R, const double* n,
                                        const double* fv, double gamma, int boost, double* F)
{
    constexpr int NF = D + 2;
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
    get_waves(_dl, Vnl, _pl, _dr, Vnr, _pr, gamma, sl, sc, sr);
    if (wn <= sl) {
        side_flux<D>(L, n, Vnl, vl_sq, wn, gamma, F);
    } else if ((sl < wn) && (wn <= sr)) {
        if (SOLVER == PHD_SOLVER_HLL) {
            // HLL.riemann_solver (riemann.pyx:346-361)
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
            // HLLC star states (riemann.pyx:633-670), Toro eq. 10.38-10.39
            if (wn <= sc) {
                double factor_1 = _dl * (sl - Vnl) / (sl - sc);
                double factor_2 = factor_1 * (sl - wn) * (sc - Vnl) + _pl;
                double frho = _dl * (Vnl - sl) + factor_1 * (sl - wn);
                double el = 0.5 * _dl * vl_sq + _pl / (gamma - 1.0);
                F[0] = frho;
                F[D + 1] = (el + _pl) * Vnl - sl * el +
                           (sl - wn) * factor_1 *
                               (el / _dl + (sc - Vnl) * (sc + _pl / (_dl * (sl - Vnl))));
#pragma unroll
                for (int k = 0; k < D; k++) F[1 + k] = frho * L[1 + k] + factor_2 * n[k];
            } else {
                double factor_1 = _dr * (sr - Vnr) / (sr - sc);
                double factor_2 = factor_1 * (sr - wn) * (sc - Vnr) + _pr;
                double frho = _dr * (Vnr - sr) + factor_1 * (sr - wn);
                double er = 0.5 * _dr * vr_sq + _pr / (gamma - 1.0);
                F[0] = frho;
                F[D + 1] = (er + _pr) * Vnr - sr * er +
                           (sr - wn) * factor_1 *
                               (er / _dr + (sc - Vnr) * (sc + _pr / (_dr * (sr - Vnr))));
#pragma unroll
                for (int k = 0; k < D; k++) F[1 + k] = frho * R[1 + k] + factor_2 * n[k];
            }
        }
    } else {
        side_flux<D>(R, n, Vnr, vr_sq, wn, gamma, F);
    }
    (void)NF;
}

// RiemannBase.deboost (riemann.pyx:190-222), Pakmor & Springel 2011 eq. 17
template <int D>
__device__ __forceinline__ void deboost(const double* fv, double* F)
{
#pragma unroll
    for (int k = 0; k < D; k++) {
        F[D + 1] += fv[k] * (0.5 * fv[k] * F[0] + F[1 + k]);
        F[1 + k] += fv[k] * F[0];
    }
}

// ---- Exact solver (riemann.pyx:684-1090), Toro chapter 4 ------------------
__device__ double ex_p_func(double d, double p, double c, double gamma, double p_old)
{
    if (p_old <= p) return 2. * c / (gamma - 1.) * (pow(p_old / p, (gamma - 1.) / (2. * gamma)) - 1.);
    double Ak = 2. / (d * (gamma + 1.));
    double Bk = p * (gamma - 1.) / (gamma + 1.);
    return (p_old - p) * sqrt(Ak / (p_old + Bk));
}

__device__ double ex_p_deriv(double d, double p, double c, double gamma, double p_old)
{
    if (p_old <= p) return pow(p_old / p, -(gamma + 1.) / (2. * gamma)) / (c * d);
    double Ak = 2. / (d * (gamma + 1.));
    double Bk = p * (gamma - 1.) / (gamma + 1.);
    return sqrt(Ak / (p_old + Bk)) * (1. - .5 * (p_old - p) / (Bk + p_old));
}

__device__ bool ex_pstar(double dl, double ul, double pl, double cl, double dr, double ur,
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
        double p_lr = pow(pl / pr, (gamma - 1.) / (2. * gamma));
        double u_star = (p_lr * ul / cl + ur / cr + 2. * (p_lr - 1.) / (gamma - 1.));
        u_star = u_star / (p_lr / cl + 1. / cr);
        double p_tl = pow(1. + (gamma - 1.) * (ul - u_star) / (2. * cl), 2. * gamma / (gamma - 1.));
        double p_tr = pow(1. + (gamma - 1.) * (u_star - ur) / (2. * cr), 2. * gamma / (gamma - 1.));
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
__device__ __forceinline__ void ex_state_flux(double d, double p, const double* v, const double* n,
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

// Exact.riemann_solver (riemann.pyx:700-972).  Returns false on Newton
// non-convergence.
template <int D>
__device__ bool riemann_exact(const double* L, const double* R, const double* n, double gamma,
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
    double d, p, shift;
    bool fan = false;
    if (0.0 <= u_star) {
        S = L;
        if (p_star <= _pl) {
            double s_hl = vnl - cl;
            if (0.0 <= s_hl) { d = _dl; p = _pl; shift = vnl; }
            else {
                double c_star_l = cl * pow(p_star / _pl, (gamma - 1.0) / (2.0 * gamma));
                double s_tl = u_star - c_star_l;
                if (0.0 >= s_tl) { d = _dl * pow(p_star / _pl, 1.0 / gamma); p = p_star; shift = u_star; }
                else {
                    double c = (2.0 / (gamma + 1.0)) * (cl + 0.5 * (gamma - 1.0) * vnl);
                    d = _dl * pow(c / cl, 2.0 / (gamma - 1.0));
                    p = _pl * pow(c / cl, 2.0 * gamma / (gamma - 1.0));
                    shift = c;
                }
                fan = true;
            }
        } else {
            double sl = vnl - cl * sqrt((gamma + 1.0) * p_star / (2.0 * gamma * _pl) +
                                        (gamma - 1.0) / (2.0 * gamma));
            if (0.0 <= sl) { d = _dl; p = _pl; shift = vnl; }
            else {
                d = _dl * (p_star / _pl + (gamma - 1.0) / (gamma + 1.0)) /
                    (p_star * (gamma - 1.0) / ((gamma + 1.0) * _pl) + 1.0);
                p = p_star; shift = u_star; fan = true;
            }
        }
        if (!fan) {
            // unperturbed left data (riemann.pyx:809-813)
            double vsq = 0.0;
#pragma unroll
            for (int k = 0; k < D; k++) vsq += L[1 + k] * L[1 + k];
            F[0] = _dl * vnl;
            F[D + 1] = (0.5 * _dl * vsq + gamma * _pl / (gamma - 1.0)) * vnl;
#pragma unroll
            for (int k = 0; k < D; k++) F[1 + k] = _dl * L[1 + k] * vnl + _pl * n[k];
            return true;
        }
#pragma unroll
        for (int k = 0; k < D; k++) v[k] = S[1 + k] + (shift - vnl) * n[k];
    } else {
        S = R;
        if (p_star >= _pr) {
            double sr = vnr + cr * sqrt((gamma + 1.0) * p_star / (2.0 * gamma * _pr) +
                                        (gamma - 1.0) / (2.0 * gamma));
            if (0.0 >= sr) { d = _dr; p = _pr; shift = vnr; }
            else {
                d = _dr * (p_star / _pr + (gamma - 1.0) / (gamma + 1.0)) /
                    (p_star * (gamma - 1.0) / ((gamma + 1.0) * _pr) + 1.0);
                p = p_star; shift = u_star; fan = true;
            }
        } else {
            double s_hr = vnr + cr;
            if (0.0 >= s_hr) { d = _dr; p = _pr; shift = vnr; }
            else {
                double c_star_r = cr * pow(p_star / _pr, (gamma - 1.0) / (2.0 * gamma));
                double s_tr = u_star + c_star_r;
                if (0.0 <= s_tr) { d = _dr * pow(p_star / _pr, 1.0 / gamma); p = p_star; shift = u_star; }
                else {
                    double c = (2.0 / (gamma + 1.0)) * (cr - 0.5 * (gamma - 1.0) * vnr);
                    double u_tmp = (2.0 / (gamma + 1.0)) * (-cr + 0.5 * (gamma - 1.0) * vnr);
                    d = _dr * pow(c / cr, 2.0 / (gamma - 1.0));
                    p = _pr * pow(c / cr, 2.0 * gamma / (gamma - 1.0));
                    shift = u_tmp;
                }
                fan = true;
            }
        }
        if (!fan) {
            double vsq = 0.0;
#pragma unroll
            for (int k = 0; k < D; k++) vsq += R[1 + k] * R[1 + k];
            F[0] = _dr * vnr;
            F[D + 1] = (0.5 * _dr * vsq + gamma * _pr / (gamma - 1.0)) * vnr;
#pragma unroll
            for (int k = 0; k < D; k++) F[1 + k] = _dr * R[1 + k] * vnr + _pr * n[k];
            return true;
        }
#pragma unroll
        for (int k = 0; k < D; k++) v[k] = S[1 + k] + (shift - vnr) * n[k];
    }
    ex_state_flux<D>(d, p, v, n, gamma, F);
    return true;
}

// ---------------------------------------------------------------------------
// Face kernels
template <int D>
__device__ __forceinline__ void load_face_vec(const double* const* a, int m, double* v)
{
#pragma unroll
    for (int k = 0; k < D; k++) v[k] = a[k][m];
}

// Materialised states: PieceWiseConstant.compute_states or
// PieceWiseLinear.add_spatial / add_temporal.
template <int D, int ORDER>
__global__ __launch_bounds__(PHD_BLOCK) void k_states(phd_cells c, phd_faces f,
                                                      const double* __restrict__ grad, int ld,
                                                      double gamma, double dt, int boost,
                                                      int gizmo, int stages, FaceArrays left,
                                                      FaceArrays right)
{
    constexpr int NF = D + 2;
    int m = blockIdx.x * PHD_BLOCK + threadIdx.x;
    if (m >= f.n_faces) return;
    int i = f.pair_i[m], j = f.pair_j[m];
    double fv[D], pi[NF], pj[NF], L[NF], R[NF];
    load_face_vec<D>(f.velocity, m, fv);
    if (ORDER == PHD_ORDER_CONSTANT || (stages & PHD_STAGE_SPATIAL)) {
        copy_states<D>(c, fv, boost, i, j, pi, pj, L, R);
    } else {
#pragma unroll
        for (int n = 0; n < NF; n++) {
            pi[n] = c.prim[n][i];
            pj[n] = c.prim[n][j];
            L[n] = left.q[n][m];
            R[n] = right.q[n][m];
        }
    }
    if (ORDER == PHD_ORDER_LINEAR) {
        double gi[D * (D + 2)], gj[D * (D + 2)];
        load_grad_row<D>(grad + (size_t)i * ld, gi);
        load_grad_row<D>(grad + (size_t)j * ld, gj);
        if (stages & PHD_STAGE_SPATIAL) spatial_states<D>(c, f, m, i, j, pi, pj, gi, gj, gizmo, L, R);
        if (stages & PHD_STAGE_TEMPORAL) {
            temporal_side<D>(pi, gi, fv, boost, gamma, dt, L);
            temporal_side<D>(pj, gj, fv, boost, gamma, dt, R);
        }
    }
#pragma unroll
    for (int n = 0; n < NF; n++) { left.q[n][m] = L[n]; right.q[n][m] = R[n]; }
}

template <int D, int SOLVER>
__device__ __forceinline__ bool solve_face(const double* L, const double* R, const double* n,
                                           const double* fv, double gamma, int boost, double* F)
{
    if (SOLVER == PHD_SOLVER_EXACT) {
        if (!riemann_exact<D>(L, R, n, gamma, F)) return false;
        deboost<D>(fv, F);
        return true;
    }
    riemann<D, SOLVER>(L, R, n, fv, gamma, boost, F);
    if (boost) deboost<D>(fv, F);
    return true;
}

// Riemann fluxes from materialised states (riemann.pyx:74-98)
template <int D, int SOLVER>
__global__ __launch_bounds__(PHD_BLOCK) void k_riemann(phd_faces f, FaceArrays left,
                                                       FaceArrays right, double gamma, int boost,
                                                       FaceArrays flux, unsigned long long* err)
{
    constexpr int NF = D + 2;
    int m = blockIdx.x * PHD_BLOCK + threadIdx.x;
    if (m >= f.n_faces) return;
    double L[NF], R[NF], n[D], fv[D], F[NF];
#pragma unroll
    for (int q = 0; q < NF; q++) { L[q] = left.q[q][m]; R[q] = right.q[q][m]; }
    load_face_vec<D>(f.normal, m, n);
    load_face_vec<D>(f.velocity, m, fv);
    if (!solve_face<D, SOLVER>(L, R, n, fv, gamma, boost, F)) {
        record_min(&err[2], (unsigned long long)m);
        return;
    }
#pragma unroll
    for (int q = 0; q < NF; q++) flux.q[q][m] = F[q];
}

// Fused states + Riemann + deboost: states stay in registers.
template <int D, int ORDER, int SOLVER>
__global__ __launch_bounds__(PHD_BLOCK) void k_face_flux(phd_cells c, phd_faces f,
                                                         const double* __restrict__ grad, int ld,
                                                         double gamma, double dt, int boost,
                                                         int gizmo, FaceArrays flux,
                                                         unsigned long long* err)
{
    constexpr int NF = D + 2;
    int m = blockIdx.x * PHD_BLOCK + threadIdx.x;
    if (m >= f.n_faces) return;
    int i = f.pair_i[m], j = f.pair_j[m];
    double fv[D], n[D], pi[NF], pj[NF], L[NF], R[NF], F[NF];
    load_face_vec<D>(f.velocity, m, fv);
    copy_states<D>(c, fv, boost, i, j, pi, pj, L, R);
    if (ORDER == PHD_ORDER_LINEAR) {
        double gi[D * (D + 2)], gj[D * (D + 2)];
        load_grad_row<D>(grad + (size_t)i * ld, gi);
        load_grad_row<D>(grad + (size_t)j * ld, gj);
        spatial_states<D>(c, f, m, i, j, pi, pj, gi, gj, gizmo, L, R);
        temporal_side<D>(pi, gi, fv, boost, gamma, dt, L);
        temporal_side<D>(pj, gj, fv, boost, gamma, dt, R);
    }
    load_face_vec<D>(f.normal, m, n);
    if (!solve_face<D, SOLVER>(L, R, n, fv, gamma, boost, F)) {
        record_min(&err[2], (unsigned long long)m);
        return;
    }
#pragma unroll
    for (int q = 0; q < NF; q++) flux.q[q][m] = F[q];
}

// ---------------------------------------------------------------------------
// conserved update (mesh.pyx:479-551).  One thread per real cell walks the
// cell's face list in ascending face id, which is the order in which the
// reference's face loop touches that cell, so every running sum is
// bit-identical and no atomics are needed.  The positivity check of the
// reference (546-551) fires at the first face n at which a running mass or
// energy of either endpoint is <= 0; each endpoint reports its first
// failure as key 2n (mass) or 2n+1 (energy) and the minimum key wins.
template <int D>
__global__ __launch_bounds__(PHD_BLOCK) void k_update(phd_cells c, phd_faces f, FaceArrays flux,
                                                      double dt, unsigned long long* err)
{
    constexpr int NF = D + 2;
    int i = blockIdx.x * PHD_BLOCK + threadIdx.x;
    if (i >= c.n_cells || c.tag[i] != 0) return;
    double u[NF];
#pragma unroll
    for (int q = 0; q < NF; q++) u[q] = c.cons[q][i];
    const int beg = f.nbr_off[i], end = f.nbr_off[i + 1];
    for (int m = beg; m < end; m++) {
        const int fid = f.nbr_face[m];
        const double t = dt * f.area[fid];
        const int a = f.pair_i[fid], b = f.pair_j[fid];
        int other;
        if (a == i) {
            other = b;
#pragma unroll
            for (int q = 0; q < NF; q++) u[q] -= t * flux.q[q][fid];
        } else {
            other = a;
#pragma unroll
            for (int q = 0; q < NF; q++) u[q] += t * flux.q[q][fid];
        }
        bool bad_m = u[0] <= 0.0, bad_e = u[D + 1] <= 0.0;
        if (c.tag[other] != 0) {
            bad_m = bad_m || c.cons[0][other] <= 0.0;
            bad_e = bad_e || c.cons[D + 1][other] <= 0.0;
        }
        if (bad_m || bad_e) {
            record_min(&err[0], 2ull * (unsigned long long)fid + (bad_m ? 0ull : 1ull));
            break;
        }
    }
#pragma unroll
    for (int q = 0; q < NF; q++) c.cons[q][i] = u[q];
}

// Static-mesh epilogue: DomainManager.update_ghost_fields(particles,
// conservative, True) (domain_manager.pyx:759-774) with the boundary rule
// Reflective.update_fields (boundary.pyx:336-348: momentum and velocity of
// exterior ghosts negated on each axis where they lie outside the box),
// then IdealGas.primitive_from_conservative (equation_state.pyx:104-117)
// over all cells.  Ghosts read only their image's conserved values, which
// this kernel does not write.
template <int D>
__global__ __launch_bounds__(PHD_BLOCK) void k_ghost_refresh_eos(phd_cells c, int reflect,
                                                                 double lo0, double lo1,
                                                                 double lo2, double hi0,
                                                                 double hi1, double hi2,
                                                                 double gamma)
{
    constexpr int NF = D + 2;
    int i = blockIdx.x * PHD_BLOCK + threadIdx.x;
    if (i >= c.n_cells) return;
    double u[NF];
    if (c.tag[i] == 1) {
        const int s = c.map[i];
#pragma unroll
        for (int q = 0; q < NF; q++) u[q] = c.cons[q][s];
        if (reflect && c.type[i] == 2) {
            const double lo[3] = {lo0, lo1, lo2}, hi[3] = {hi0, hi1, hi2};
#pragma unroll
            for (int k = 0; k < D; k++) {
                double xk = c.x[k][i];
                if (xk < lo[k]) u[1 + k] *= -1;
                if (xk > hi[k]) u[1 + k] *= -1;
            }
        }
#pragma unroll
        for (int q = 0; q < NF; q++) c.cons[q][i] = u[q];
    } else {
#pragma unroll
        for (int q = 0; q < NF; q++) u[q] = c.cons[q][i];
    }
    const double vol = c.volume[i];
    const double d = u[0] / vol;
    c.prim[0][i] = d;
    double v_sq = 0.;
#pragma unroll
    for (int k = 0; k < D; k++) {
        double v = u[1 + k] / u[0];
        c.prim[1 + k][i] = v;
        v_sq += v * v;
    }
    c.prim[D + 1][i] = (u[D + 1] / vol - .5 * d * v_sq) * (gamma - 1.);
}

// ---------------------------------------------------------------------------
// host helpers
int check_cells(const phd_cells* c)
{
    if (!c || (c->dim != 2 && c->dim != 3) || c->n_cells < 0) return PHD_E_ARG;
    return 0;
}
int check_faces(const phd_faces* f)
{
    if (!f || f->n_faces < 0) return PHD_E_ARG;
    return 0;
}

FaceArrays fa(double* const* p, int nf)
{
    FaceArrays a;
    for (int q = 0; q < 5; q++) a.q[q] = q < nf ? p[q] : nullptr;
    return a;
}

#define HIPCHK(x) do { if ((x) != hipSuccess) return PHD_E_HIP; } while (0)
#define LAUNCHCHK() do { if (hipPeekAtLastError() != hipSuccess) { (void)hipGetLastError(); return PHD_E_HIP; } } while (0)

template <int D>
int launch_gradients(phd_handle* h, const phd_cells* c, const phd_faces* f, int limiter,
                     double* grad, int ld)
{
    if (c->n_cells == 0) return 0;
    if (limiter == PHD_LIMITER_AREPO)
        k_gradients<D, PHD_LIMITER_AREPO><<<nblocks(c->n_cells), PHD_BLOCK, 0, h->stream>>>(*c, *f, grad, ld, h->d_err);
    else if (limiter == PHD_LIMITER_TESS)
        k_gradients<D, PHD_LIMITER_TESS><<<nblocks(c->n_cells), PHD_BLOCK, 0, h->stream>>>(*c, *f, grad, ld, h->d_err);
    else return PHD_E_MODE;
    LAUNCHCHK();
    return 0;
}

template <int D>
int launch_ghost(phd_handle* h, const phd_cells* c, int boundary, const double* lo,
                 const double* hi, double* grad, int ld)
{
    if (c->n_cells == 0) return 0;
    constexpr int G = D * (D + 2);
    double l[3] = {0, 0, 0}, u[3] = {0, 0, 0};
    if (boundary == PHD_BOUNDARY_REFLECTIVE) {
        if (!lo || !hi) return PHD_E_ARG;
        for (int k = 0; k < D; k++) { l[k] = lo[k]; u[k] = hi[k]; }
    } else if (boundary != PHD_BOUNDARY_PERIODIC) return PHD_E_MODE;
    if (!c->map) return 0;
    long long n = (long long)c->n_cells * G;
    k_ghost_grad<D><<<nblocks(n), PHD_BLOCK, 0, h->stream>>>(*c, boundary == PHD_BOUNDARY_REFLECTIVE, l[0], l[1], l[2], u[0], u[1], u[2], grad, ld);
    LAUNCHCHK();
    return 0;
}

template <int D, int ORDER>
int launch_face_flux_o(phd_handle* h, const phd_cells* c, const phd_faces* f, const double* grad,
                       int ld, double gamma, double dt, int boost, int gizmo, int solver,
                       FaceArrays flux)
{
    unsigned nb = nblocks(f->n_faces);
    switch (solver) {
    case PHD_SOLVER_HLL:
        k_face_flux<D, ORDER, PHD_SOLVER_HLL><<<nb, PHD_BLOCK, 0, h->stream>>>(*c, *f, grad, ld, gamma, dt, boost, gizmo, flux, h->d_err); break;
    case PHD_SOLVER_HLLC:
        k_face_flux<D, ORDER, PHD_SOLVER_HLLC><<<nb, PHD_BLOCK, 0, h->stream>>>(*c, *f, grad, ld, gamma, dt, boost, gizmo, flux, h->d_err); break;
    case PHD_SOLVER_EXACT:
        k_face_flux<D, ORDER, PHD_SOLVER_EXACT><<<nb, PHD_BLOCK, 0, h->stream>>>(*c, *f, grad, ld, gamma, dt, boost, gizmo, flux, h->d_err); break;
    default: return PHD_E_MODE;
    }
    LAUNCHCHK();
    return 0;
}

template <int D>
int launch_face_flux(phd_handle* h, const phd_cells* c, const phd_faces* f, const double* grad,
                     int ld, double gamma, double dt, int boost, int gizmo, int order, int solver,
                     FaceArrays flux)
{
    if (f->n_faces == 0) return 0;
    if (order == PHD_ORDER_LINEAR)
        return launch_face_flux_o<D, PHD_ORDER_LINEAR>(h, c, f, grad, ld, gamma, dt, boost, gizmo, solver, flux);
    if (order == PHD_ORDER_CONSTANT)
        return launch_face_flux_o<D, PHD_ORDER_CONSTANT>(h, c, f, grad, ld, gamma, dt, boost, gizmo, solver, flux);
    return PHD_E_MODE;
}

template <int D>
int launch_update(phd_handle* h, const phd_cells* c, const phd_faces* f, FaceArrays flux, double dt)
{
    if (c->n_cells == 0) return 0;
    k_update<D><<<nblocks(c->n_cells), PHD_BLOCK, 0, h->stream>>>(*c, *f, flux, dt, h->d_err);
    LAUNCHCHK();
    return 0;
}

int grad_ok(const phd_cells* c, const double* grad, int ld)
{
    int D = c->dim;
    if (!grad || ld < D * (D + 2) || (ld & 1) || (((uintptr_t)grad) & 15)) return PHD_E_ARG;
    return 0;
}

} // namespace

// ===========================================================================
// C ABI
extern "C" {

int phd_abi_version(void) { return PHD_ABI_VERSION; }

const char* phd_error_string(int code)
{
    switch (code) {
    case 0: return "ok";
    case PHD_E_ARG: return "invalid argument";
    case PHD_E_HIP: return "HIP runtime error";
    case PHD_E_MODE: return "unsupported mode";
    case PHD_F_NEIGHBORS: return "ERROR: incorrect neighbors!";
    case PHD_F_MASS: return "Mass less than zero in flux update";
    case PHD_F_ENERGY: return "Energy less than zero in flux update";
    case PHD_F_EXACT: return "No convergence in Exact Riemann Solver";
    default: return "unknown error";
    }
}

int phd_handle_create(phd_handle** out, int device)
{
    if (!out) return PHD_E_ARG;
    *out = nullptr;
    HIPCHK(hipSetDevice(device));
    phd_handle* h = new phd_handle();
    h->device = device;
    h->stream = nullptr;
    h->mode = PHD_MODE_BITWISE;
    if (hipMalloc(&h->d_err, 4 * sizeof(unsigned long long)) != hipSuccess) { delete h; return PHD_E_HIP; }
    if (hipHostMalloc(&h->h_err, 4 * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess) {
        (void)hipFree(h->d_err); delete h; return PHD_E_HIP;
    }
    k_reset_err<<<1, 64, 0, h->stream>>>(h->d_err);
    if (hipStreamSynchronize(h->stream) != hipSuccess) {
        (void)hipFree(h->d_err); (void)hipHostFree(h->h_err); delete h; return PHD_E_HIP;
    }
    *out = h;
    return 0;
}

int phd_handle_destroy(phd_handle* h)
{
    if (!h) return PHD_E_ARG;
    (void)hipStreamSynchronize(h->stream);
    (void)hipFree(h->d_err);
    (void)hipHostFree(h->h_err);
    delete h;
    return 0;
}

int phd_set_stream(phd_handle* h, void* stream)
{
    if (!h) return PHD_E_ARG;
    h->stream = (hipStream_t)stream;
    return 0;
}

int phd_set_mode(phd_handle* h, int mode)
{
    if (!h) return PHD_E_ARG;
    if (mode != PHD_MODE_BITWISE && mode != PHD_MODE_FAST) return PHD_E_MODE;
    h->mode = mode;
    return 0;
}

int phd_get_error(phd_handle* h, int32_t* code, int64_t* index)
{
    if (!h || !code || !index) return PHD_E_ARG;
    HIPCHK(hipMemcpyAsync(h->h_err, h->d_err, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    unsigned long long e0 = h->h_err[0], e1 = h->h_err[1], e2 = h->h_err[2];
    *code = 0;
    *index = -1;
    if (e1 != NOERR) { *code = PHD_F_NEIGHBORS; *index = (int64_t)e1; }
    else if (e2 != NOERR) { *code = PHD_F_EXACT; *index = (int64_t)e2; }
    else if (e0 != NOERR) { *code = (e0 & 1ull) ? PHD_F_ENERGY : PHD_F_MASS; *index = (int64_t)(e0 >> 1); }
    if (*code) {
        k_reset_err<<<1, 64, 0, h->stream>>>(h->d_err);
        LAUNCHCHK();
    }
    return 0;
}

int phd_conservative_from_primitive(phd_handle* h, const phd_cells* c, double gamma)
{
    if (!h || check_cells(c)) return PHD_E_ARG;
    if (c->n_cells == 0) return 0;
    if (c->dim == 2) k_cons_from_prim<2><<<nblocks(c->n_cells), PHD_BLOCK, 0, h->stream>>>(*c, gamma);
    else k_cons_from_prim<3><<<nblocks(c->n_cells), PHD_BLOCK, 0, h->stream>>>(*c, gamma);
    LAUNCHCHK();
    return 0;
}

int phd_primitive_from_conservative(phd_handle* h, const phd_cells* c, double gamma)
{
    if (!h || check_cells(c)) return PHD_E_ARG;
    if (c->n_cells == 0) return 0;
    if (c->dim == 2) k_prim_from_cons<2><<<nblocks(c->n_cells), PHD_BLOCK, 0, h->stream>>>(*c, gamma);
    else k_prim_from_cons<3><<<nblocks(c->n_cells), PHD_BLOCK, 0, h->stream>>>(*c, gamma);
    LAUNCHCHK();
    return 0;
}

int phd_compute_time_step(phd_handle* h, const phd_cells* c, double gamma, double cfl, int boost,
                          double* dt_out)
{
    if (!h || check_cells(c) || !dt_out || c->n_cells == 0) return PHD_E_ARG;
    k_reset_err<<<1, 64, 0, h->stream>>>(h->d_err);  // also resets the fault words
    LAUNCHCHK();
    if (c->dim == 2) k_cfl<2><<<nblocks(c->n_cells), PHD_BLOCK, 0, h->stream>>>(*c, gamma, boost, h->d_err);
    else k_cfl<3><<<nblocks(c->n_cells), PHD_BLOCK, 0, h->stream>>>(*c, gamma, boost, h->d_err);
    LAUNCHCHK();
    HIPCHK(hipMemcpyAsync(h->h_err + 3, h->d_err + 3, sizeof(unsigned long long), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    double m;
    memcpy(&m, &h->h_err[3], sizeof(double));
    *dt_out = cfl * m;
    return 0;
}

int phd_assign_generator_velocities(phd_handle* h, const phd_cells* c, double gamma, double eta,
                                    int regularize)
{
    if (!h || check_cells(c)) return PHD_E_ARG;
    if (c->n_cells == 0) return 0;
    if (c->dim == 2) k_gen_vel<2><<<nblocks(c->n_cells), PHD_BLOCK, 0, h->stream>>>(*c, gamma, eta, regularize);
    else k_gen_vel<3><<<nblocks(c->n_cells), PHD_BLOCK, 0, h->stream>>>(*c, gamma, eta, regularize);
    LAUNCHCHK();
    return 0;
}

int phd_assign_face_velocities(phd_handle* h, const phd_cells* c, const phd_faces* f)
{
    if (!h || check_cells(c) || check_faces(f)) return PHD_E_ARG;
    if (f->n_faces == 0) return 0;
    if (c->dim == 2) k_face_vel<2><<<nblocks(f->n_faces), PHD_BLOCK, 0, h->stream>>>(*c, *f);
    else k_face_vel<3><<<nblocks(f->n_faces), PHD_BLOCK, 0, h->stream>>>(*c, *f);
    LAUNCHCHK();
    return 0;
}

int phd_compute_gradients(phd_handle* h, const phd_cells* c, const phd_faces* f, int limiter,
                          double* grad, int32_t grad_ld)
{
    if (!h || check_cells(c) || check_faces(f) || grad_ok(c, grad, grad_ld)) return PHD_E_ARG;
    return c->dim == 2 ? launch_gradients<2>(h, c, f, limiter, grad, grad_ld)
                       : launch_gradients<3>(h, c, f, limiter, grad, grad_ld);
}

int phd_update_ghost_gradients(phd_handle* h, const phd_cells* c, int boundary, const double* lo,
                               const double* hi, double* grad, int32_t grad_ld)
{
    if (!h || check_cells(c) || grad_ok(c, grad, grad_ld)) return PHD_E_ARG;
    return c->dim == 2 ? launch_ghost<2>(h, c, boundary, lo, hi, grad, grad_ld)
                       : launch_ghost<3>(h, c, boundary, lo, hi, grad, grad_ld);
}

int phd_compute_states(phd_handle* h, const phd_cells* c, const phd_faces* f, const double* grad,
                       int32_t grad_ld, double gamma, double dt, int boost, int gizmo, int order,
                       int stages, double* const* left, double* const* right)
{
    if (!h || check_cells(c) || check_faces(f) || !left || !right) return PHD_E_ARG;
    if (order == PHD_ORDER_LINEAR && grad_ok(c, grad, grad_ld)) return PHD_E_ARG;
    if (order != PHD_ORDER_LINEAR && order != PHD_ORDER_CONSTANT) return PHD_E_MODE;
    if (f->n_faces == 0) return 0;
    int nf = c->dim + 2;
    FaceArrays L = fa(left, nf), R = fa(right, nf);
    unsigned nb = nblocks(f->n_faces);
    if (c->dim == 2) {
        if (order == PHD_ORDER_LINEAR) k_states<2, PHD_ORDER_LINEAR><<<nb, PHD_BLOCK, 0, h->stream>>>(*c, *f, grad, grad_ld, gamma, dt, boost, gizmo, stages, L, R);
        else k_states<2, PHD_ORDER_CONSTANT><<<nb, PHD_BLOCK, 0, h->stream>>>(*c, *f, grad, grad_ld, gamma, dt, boost, gizmo, stages, L, R);
    } else {
        if (order == PHD_ORDER_LINEAR) k_states<3, PHD_ORDER_LINEAR><<<nb, PHD_BLOCK, 0, h->stream>>>(*c, *f, grad, grad_ld, gamma, dt, boost, gizmo, stages, L, R);
        else k_states<3, PHD_ORDER_CONSTANT><<<nb, PHD_BLOCK, 0, h->stream>>>(*c, *f, grad, grad_ld, gamma, dt, boost, gizmo, stages, L, R);
    }
    LAUNCHCHK();
    return 0;
}

int phd_compute_fluxes(phd_handle* h, const phd_faces* f, int dim, double* const* left,
                       double* const* right, double gamma, int boost, int solver,
                       double* const* flux)
{
    if (!h || check_faces(f) || (dim != 2 && dim != 3) || !left || !right || !flux) return PHD_E_ARG;
    if (f->n_faces == 0) return 0;
    int nf = dim + 2;
    FaceArrays L = fa(left, nf), R = fa(right, nf), F = fa(flux, nf);
    unsigned nb = nblocks(f->n_faces);
#define RIE(DD, SS) k_riemann<DD, SS><<<nb, PHD_BLOCK, 0, h->stream>>>(*f, L, R, gamma, boost, F, h->d_err)
    if (dim == 2) {
        if (solver == PHD_SOLVER_HLL) RIE(2, PHD_SOLVER_HLL);
        else if (solver == PHD_SOLVER_HLLC) RIE(2, PHD_SOLVER_HLLC);
        else if (solver == PHD_SOLVER_EXACT) RIE(2, PHD_SOLVER_EXACT);
        else return PHD_E_MODE;
    } else {
        if (solver == PHD_SOLVER_HLL) RIE(3, PHD_SOLVER_HLL);
        else if (solver == PHD_SOLVER_HLLC) RIE(3, PHD_SOLVER_HLLC);
        else if (solver == PHD_SOLVER_EXACT) RIE(3, PHD_SOLVER_EXACT);
        else return PHD_E_MODE;
    }
#undef RIE
    LAUNCHCHK();
    return 0;
}

int phd_face_flux(phd_handle* h, const phd_cells* c, const phd_faces* f, const double* grad,
                  int32_t grad_ld, double gamma, double dt, int boost, int gizmo, int order,
                  int solver, double* const* flux)
{
    if (!h || check_cells(c) || check_faces(f) || !flux) return PHD_E_ARG;
    if (order == PHD_ORDER_LINEAR && grad_ok(c, grad, grad_ld)) return PHD_E_ARG;
    FaceArrays F = fa(flux, c->dim + 2);
    return c->dim == 2
               ? launch_face_flux<2>(h, c, f, grad, grad_ld, gamma, dt, boost, gizmo, order, solver, F)
               : launch_face_flux<3>(h, c, f, grad, grad_ld, gamma, dt, boost, gizmo, order, solver, F);
}

int phd_update_from_fluxes(phd_handle* h, const phd_cells* c, const phd_faces* f,
                           double* const* flux, double dt)
{
    if (!h || check_cells(c) || check_faces(f) || !flux) return PHD_E_ARG;
    FaceArrays F = fa(flux, c->dim + 2);
    return c->dim == 2 ? launch_update<2>(h, c, f, F, dt) : launch_update<3>(h, c, f, F, dt);
}

int phd_core_step(phd_handle* h, const phd_cells* c, const phd_faces* f, double* grad,
                  int32_t grad_ld, int limiter, int boundary, const double* lo, const double* hi,
                  double gamma, double dt, int boost, int gizmo, int solver, double* const* flux)
{
    int rc = phd_compute_gradients(h, c, f, limiter, grad, grad_ld);
    if (rc) return rc;
    rc = phd_update_ghost_gradients(h, c, boundary, lo, hi, grad, grad_ld);
    if (rc) return rc;
    rc = phd_face_flux(h, c, f, grad, grad_ld, gamma, 0.5 * dt, boost, gizmo, PHD_ORDER_LINEAR,
                       solver, flux);
    if (rc) return rc;
    return phd_update_from_fluxes(h, c, f, flux, dt);
}

int phd_ghost_refresh_eos(phd_handle* h, const phd_cells* c, int boundary, const double* lo,
                          const double* hi, double gamma)
{
    if (!h || check_cells(c)) return PHD_E_ARG;
    if (c->n_cells == 0) return 0;
    double l[3] = {0, 0, 0}, u[3] = {0, 0, 0};
    if (boundary == PHD_BOUNDARY_REFLECTIVE) {
        if (!lo || !hi) return PHD_E_ARG;
        for (int k = 0; k < c->dim; k++) { l[k] = lo[k]; u[k] = hi[k]; }
    } else if (boundary != PHD_BOUNDARY_PERIODIC) return PHD_E_MODE;
    const int refl = boundary == PHD_BOUNDARY_REFLECTIVE;
    if (!c->map) {
        // no ghosts: plain EOS
        return phd_primitive_from_conservative(h, c, gamma);
    }
    if (c->dim == 2)
        k_ghost_refresh_eos<2><<<nblocks(c->n_cells), PHD_BLOCK, 0, h->stream>>>(*c, refl, l[0], l[1], l[2], u[0], u[1], u[2], gamma);
    else
        k_ghost_refresh_eos<3><<<nblocks(c->n_cells), PHD_BLOCK, 0, h->stream>>>(*c, refl, l[0], l[1], l[2], u[0], u[1], u[2], gamma);
    LAUNCHCHK();
    return 0;
}

} // extern "C"
