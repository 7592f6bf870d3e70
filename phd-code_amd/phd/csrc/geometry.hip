// geometry.hip -- Voronoi mesh geometry on the device from the host's
// Delaunay topology: the device half of Tess2d::extract_geometry
// (tess.cpp:238-398) and Tess3d::extract_geometry (tess3.cpp:334-562).
//
// The host keeps the triangulation only (faces, ghost images and each
// face's ring: the generators whose bisectors bound it, in polygon order;
// include/phd_mesh.h).  Four kernels:
//   k_geo_ghost_pos   ghost positions, the reference's image formulas
//                     (boundary.pyx:103-104, 125-126 reflective; 378-380
//                     periodic), one thread per ghost;
//   k_geo_faces       one thread per face: Voronoi vertices as
//                     circumcentres of the ring's Delaunay simplices, in the
//                     owner generator's frame; area, centroid and normal by
//                     the reference's formulas; each side's triangle (2D) /
//                     tetrahedron (3D) volume and centroid moment into a
//                     workspace row of 2 x (1 + D) doubles;
//   k_geo_cells       one thread per real cell: its faces' contributions in
//                     ascending face id (the order of its face list), then
//                     volume and centroid offset (tess.cpp:389-393,
//                     tess3.cpp:551-559);
//   k_geo_ghost_cells ghost volume and dcom copied from the image
//                     (mesh.pyx:141-142, 323).
// Every expression is evaluated in the order of the CPU oracle's
// orc_extract_geometry (no FMA contraction, correctly rounded division and
// square root), so the results are bit-identical to it.
//
// The 3D face kernel makes three passes over the face's vertices (vertex
// mean, fan about the mean, tetrahedra about the face centroid) and
// recomputes each circumcentre per pass instead of holding a variable-length
// vertex list: the ring entries and generator rows it re-reads stay in L1/L2,
// and the extra arithmetic is cheaper than per-lane scratch arrays.
#include <stdlib.h>
#include "flux_common.h"

#define PHD_ERR_SLOT_TOPO 6   // first invalid face (or cell | 1 << 62) of a topology

namespace {

#define GHIPCHK(x) do { if ((x) != hipSuccess) return PHD_E_HIP; } while (0)
#define GLAUNCHCHK() do { if (hipPeekAtLastError() != hipSuccess) { (void)hipGetLastError(); return PHD_E_HIP; } } while (0)

constexpr unsigned long long TOPO_CELL = 1ull << 62;

// image `code` of generator coordinate xk on axis a (orc_image)
__device__ __forceinline__ double image_coord(const phd_topology& t, double xk, int a, int code)
{
    if (code == 0) return xk;
    if (t.boundary == 1) {
        int idx = code - 1;
        for (int b = 0; b < a; b++) idx /= 3;
        return xk + (double)(idx % 3 - 1) * (t.hi[a] - t.lo[a]);
    }
    if ((code - 1) >> 1 != a) return xk;
    const double edge = ((code - 1) & 1) ? t.hi[a] : t.lo[a];
    return xk - 2 * (xk - edge);
}

// a valid image code: reflective 0..2D, periodic 0..3^D with the centre
// (no shift) excluded
template <int D>
__device__ __forceinline__ bool code_ok(const phd_topology& t, int code)
{
    if (t.boundary == 1) {
        const int n3 = D == 2 ? 9 : 27, centre = D == 2 ? 4 : 13;
        return code == 0 || (code >= 1 && code <= n3 && code - 1 != centre);
    }
    return code >= 0 && code <= 2 * D;
}

// ring entry r relative to xo (orc_ring_pos)
template <int D>
__device__ __forceinline__ void ring_pos(const phd_topology& t, const phd_geometry& g, int r,
                                         const double* xo, double* p)
{
    const int k = t.ring_k[r], code = t.ring_code[r];
#pragma unroll
    for (int a = 0; a < D; a++) p[a] = image_coord(t, g.x[a][k], a, code) - xo[a];
}

__device__ __forceinline__ void cross3(const double* u, const double* v, double* w)
{
    w[0] = u[1] * v[2] - u[2] * v[1];
    w[1] = u[2] * v[0] - u[0] * v[2];
    w[2] = u[0] * v[1] - u[1] * v[0];
}

// circumcentre of 0, a, b, c (orc_circ3)
__device__ __forceinline__ void circ3(const double* a, const double* b, const double* c, double* p)
{
    double bc[3], ca[3], ab[3];
    const double aa = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
    const double bb = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
    const double cc = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
    cross3(b, c, bc);
    cross3(c, a, ca);
    cross3(a, b, ab);
    const double det = a[0] * bc[0] + a[1] * bc[1] + a[2] * bc[2];
    const double r = 1.0 / (det + det);
#pragma unroll
    for (int k = 0; k < 3; k++) p[k] = (aa * bc[k] + bb * ca[k] + cc * ab[k]) * r;
}

// circumcentre of 0, a, b (orc_circ2)
__device__ __forceinline__ void circ2(const double* a, const double* b, double* p)
{
    const double aa = a[0] * a[0] + a[1] * a[1], bb = b[0] * b[0] + b[1] * b[1];
    const double d = a[0] * b[1] - a[1] * b[0];
    const double r = 1.0 / (d + d);
    p[0] = (aa * b[1] - bb * a[1]) * r;
    p[1] = (bb * a[0] - aa * b[0]) * r;
}

// A face's two workspace rows: its contribution to the owner (plane 0) and
// to the other cell (plane 1), 1 + D doubles each: volume term, centroid
// moment.  Plane 1 starts n_faces rows after plane 0.
struct Rows {
    double* o;
    double* j;
};

template <int D>
__device__ __forceinline__ void store_side(double* r, double v, const double* m)
{
    if constexpr (D == 3) {
        *reinterpret_cast<double2*>(r) = make_double2(v, m[0]);
        *reinterpret_cast<double2*>(r + 2) = make_double2(m[1], m[2]);
    } else {
        r[0] = v;
        r[1] = m[0];
        r[2] = m[1];
    }
}

template <int D>
__global__ __launch_bounds__(PHD_BLOCK) void k_geo_ghost_pos(phd_topology t, phd_geometry g,
                                                             unsigned long long* err)
{
    const long long q = xcd_index(t.n_ghost);
    if (q >= t.n_ghost) return;
    const int k = t.ghost_map[q], code = t.ghost_code[q];
    if (k < 0 || k >= t.n_real || code == 0 || !code_ok<D>(t, code)) {
        atomicMin(err, TOPO_CELL | (unsigned long long)(t.n_real + q));
        return;
    }
#pragma unroll
    for (int a = 0; a < D; a++) g.x[a][t.n_real + q] = image_coord(t, g.x[a][k], a, code);
}

// The block's ring entries staged in LDS: entry r (rb <= r < rb + staged)
// holds the absolute image position of its generator and a validity flag.
// Entries past the staged range (a block with more than RING_CAP entries)
// are read from global memory.
constexpr int RING_CAP = 1408;   // ~ 5.5 entries per face; 37 KB of LDS per block

struct RingStage {
    double* px;
    double* py;
    double* pz;
    const unsigned char* ok;
    int rb, staged;
};

// ring entry r relative to xo, from LDS when staged (same bits as ring_pos)
__device__ __forceinline__ void ring_get(const phd_topology& t, const phd_geometry& g,
                                         const RingStage& L, int r, const double* xo, double* p)
{
    const int q = r - L.rb;
    if (q >= 0 && q < L.staged) {
        p[0] = L.px[q] - xo[0];
        p[1] = L.py[q] - xo[1];
        p[2] = L.pz[q] - xo[2];
    } else {
        ring_pos<3>(t, g, r, xo, p);
    }
}

__device__ __forceinline__ bool ring_ok(const phd_topology& t, const RingStage& L, int r)
{
    const int q = r - L.rb;
    if (q >= 0 && q < L.staged) return L.ok[q] != 0;
    const int k = t.ring_k[r];
    return k >= 0 && k < t.n_real && code_ok<3>(t, t.ring_code[r]);
}

// 3D face whose ring is staged whole: pass 1 computes vertex q (the
// circumcentre of a, P_q, P_{q+1 mod m}) and stores it over P_q, which no
// later vertex of this face needs (P_0 stays in registers for the last);
// each ring entry belongs to one face only.  The fan and the tetrahedra then
// read the vertices back.  Same arithmetic as face3's recomputing passes.
__device__ __forceinline__ void face3_staged(const phd_geometry& g, double* px, double* py,
                                             double* pz, long long f, int m, const double* xo,
                                             const double* A, Rows row)
{
    const double third = 1.0 / 3.0, fourth = 0.25;
    double P0[3], Pc[3], Pn[3], w[3], c[3] = {0.0, 0.0, 0.0};
    P0[0] = px[0] - xo[0];
    P0[1] = py[0] - xo[1];
    P0[2] = pz[0] - xo[2];
#pragma unroll
    for (int k = 0; k < 3; k++) Pc[k] = P0[k];
    for (int q = 0; q < m; q++) {
        if (q + 1 < m) {
            Pn[0] = px[q + 1] - xo[0];
            Pn[1] = py[q + 1] - xo[1];
            Pn[2] = pz[q + 1] - xo[2];
        } else {
#pragma unroll
            for (int k = 0; k < 3; k++) Pn[k] = P0[k];
        }
        circ3(A, Pc, Pn, w);
        px[q] = w[0];
        py[q] = w[1];
        pz[q] = w[2];
#pragma unroll
        for (int k = 0; k < 3; k++) { c[k] += w[k]; Pc[k] = Pn[k]; }
    }
#pragma unroll
    for (int k = 0; k < 3; k++) c[k] *= 1.0 / m;
    double v1[3], v2[3], n3[3], s[3] = {0.0, 0.0, 0.0};
    v1[0] = px[m - 1] - c[0];
    v1[1] = py[m - 1] - c[1];
    v1[2] = pz[m - 1] - c[2];
    double area1 = 0.0;
    for (int q = 0; q < m; q++) {
        v2[0] = px[q] - c[0];
        v2[1] = py[q] - c[1];
        v2[2] = pz[q] - c[2];
        cross3(v1, v2, n3);
        const double a3 = sqrt(n3[0] * n3[0] + n3[1] * n3[1] + n3[2] * n3[2]);
#pragma unroll
        for (int k = 0; k < 3; k++) s[k] += (c[k] + (v1[k] + v2[k]) * third) * a3;
        area1 += a3;
#pragma unroll
        for (int k = 0; k < 3; k++) v1[k] = v2[k];
    }
    const double l1 = sqrt(area1);
    const double l2 = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    if (area1 == 0.0 || l1 < 1.0e-10 * l2) {
        g.area[f] = 0.0;
#pragma unroll
        for (int k = 0; k < 3; k++) g.com[k][f] = xo[k] + c[k];
        const double z[3] = {0.0, 0.0, 0.0};
        store_side<3>(row.o, 0.0, z);
        store_side<3>(row.j, 0.0, z);
        return;
    }
    double fc[3], cvo[3], cvj[3];
#pragma unroll
    for (int k = 0; k < 3; k++) fc[k] = s[k] * (1.0 / area1);
    g.area[f] = 0.5 * area1;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        g.com[k][f] = xo[k] + fc[k];
        cvo[k] = 0.0 - fc[k];
        cvj[k] = A[k] - fc[k];
    }
    double vo = 0.0, vj = 0.0, mo[3] = {0.0, 0.0, 0.0}, mj[3] = {0.0, 0.0, 0.0};
    v1[0] = px[m - 1] - fc[0];
    v1[1] = py[m - 1] - fc[1];
    v1[2] = pz[m - 1] - fc[2];
    for (int q = 0; q < m; q++) {
        double cr[3];
        v2[0] = px[q] - fc[0];
        v2[1] = py[q] - fc[1];
        v2[2] = pz[q] - fc[2];
        cross3(v1, v2, cr);
        double v4 = fabs(cr[0] * cvo[0] + cr[1] * cvo[1] + cr[2] * cvo[2]);
        vo += v4;
#pragma unroll
        for (int k = 0; k < 3; k++) mo[k] += (fc[k] + (v1[k] + v2[k] + cvo[k]) * fourth) * v4;
        v4 = fabs(cr[0] * cvj[0] + cr[1] * cvj[1] + cr[2] * cvj[2]);
        vj += v4;
#pragma unroll
        for (int k = 0; k < 3; k++)
            mj[k] += ((fc[k] + (v1[k] + v2[k] + cvj[k]) * fourth) - A[k]) * v4;
#pragma unroll
        for (int k = 0; k < 3; k++) v1[k] = v2[k];
    }
    store_side<3>(row.o, vo, mo);
    store_side<3>(row.j, vj, mj);
}

// 3D face f: vertex mean, fan about it, tetrahedra about the face centroid
// (orc_face_geometry).  Vertex t is the circumcentre of a, P_t, P_{t+1 mod m}.
__device__ __forceinline__ void face3(const phd_topology& t, const phd_geometry& g,
                                      const RingStage& L, long long f, int r0, int m,
                                      const double* xo, const double* A, Rows row)
{
    constexpr int D = 3;
    const double hn = 1.0 / sqrt(A[0] * A[0] + A[1] * A[1] + A[2] * A[2]);
#pragma unroll
    for (int k = 0; k < D; k++) g.normal[k][f] = A[k] * hn;
    const int q0 = r0 - L.rb;
    if (q0 >= 0 && q0 + m <= L.staged) {
        face3_staged(g, L.px + q0, L.py + q0, L.pz + q0, f, m, xo, A, row);
        return;
    }
    const double third = 1.0 / 3.0, fourth = 0.25;
    double P0[3], Pl[3], Pc[3], Pn[3], w[3], c[3] = {0.0, 0.0, 0.0};
    ring_get(t, g, L, r0, xo, P0);
    ring_get(t, g, L, r0 + m - 1, xo, Pl);
    // vertex mean
#pragma unroll
    for (int k = 0; k < 3; k++) Pc[k] = P0[k];
    for (int q = 0; q < m; q++) {
        if (q + 1 < m) ring_get(t, g, L, r0 + q + 1, xo, Pn);
        else {
#pragma unroll
            for (int k = 0; k < 3; k++) Pn[k] = P0[k];
        }
        circ3(A, Pc, Pn, w);
#pragma unroll
        for (int k = 0; k < 3; k++) { c[k] += w[k]; Pc[k] = Pn[k]; }
    }
#pragma unroll
    for (int k = 0; k < 3; k++) c[k] *= 1.0 / m;
    // fan about c: area and unnormalised centroid
    double v1[3], v2[3], n3[3], s[3] = {0.0, 0.0, 0.0};
    circ3(A, Pl, P0, w);
#pragma unroll
    for (int k = 0; k < 3; k++) { v1[k] = w[k] - c[k]; Pc[k] = P0[k]; }
    double area1 = 0.0;
    for (int q = 0; q < m; q++) {
        if (q + 1 < m) ring_get(t, g, L, r0 + q + 1, xo, Pn);
        else {
#pragma unroll
            for (int k = 0; k < 3; k++) Pn[k] = P0[k];
        }
        circ3(A, Pc, Pn, w);
#pragma unroll
        for (int k = 0; k < 3; k++) { v2[k] = w[k] - c[k]; Pc[k] = Pn[k]; }
        cross3(v1, v2, n3);
        const double a3 = sqrt(n3[0] * n3[0] + n3[1] * n3[1] + n3[2] * n3[2]);
#pragma unroll
        for (int k = 0; k < 3; k++) s[k] += (c[k] + (v1[k] + v2[k]) * third) * a3;
        area1 += a3;
#pragma unroll
        for (int k = 0; k < 3; k++) v1[k] = v2[k];
    }
    const double l1 = sqrt(area1);
    const double l2 = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    if (area1 == 0.0 || l1 < 1.0e-10 * l2) {
        g.area[f] = 0.0;
#pragma unroll
        for (int k = 0; k < D; k++) g.com[k][f] = xo[k] + c[k];
        const double z[3] = {0.0, 0.0, 0.0};
        store_side<3>(row.o, 0.0, z);
        store_side<3>(row.j, 0.0, z);
        return;
    }
    double fc[3], cvo[3], cvj[3];
#pragma unroll
    for (int k = 0; k < 3; k++) fc[k] = s[k] * (1.0 / area1);
    g.area[f] = 0.5 * area1;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        g.com[k][f] = xo[k] + fc[k];
        cvo[k] = 0.0 - fc[k];
        cvj[k] = A[k] - fc[k];
    }
    // tetrahedra (face centroid, v1, v2, apex) for both cells
    double vo = 0.0, vj = 0.0, mo[3] = {0.0, 0.0, 0.0}, mj[3] = {0.0, 0.0, 0.0};
    circ3(A, Pl, P0, w);
#pragma unroll
    for (int k = 0; k < 3; k++) { v1[k] = w[k] - fc[k]; Pc[k] = P0[k]; }
    for (int q = 0; q < m; q++) {
        if (q + 1 < m) ring_get(t, g, L, r0 + q + 1, xo, Pn);
        else {
#pragma unroll
            for (int k = 0; k < 3; k++) Pn[k] = P0[k];
        }
        circ3(A, Pc, Pn, w);
        double cr[3];
#pragma unroll
        for (int k = 0; k < 3; k++) { v2[k] = w[k] - fc[k]; Pc[k] = Pn[k]; }
        cross3(v1, v2, cr);
        double v4 = fabs(cr[0] * cvo[0] + cr[1] * cvo[1] + cr[2] * cvo[2]);
        vo += v4;
#pragma unroll
        for (int k = 0; k < 3; k++) mo[k] += (fc[k] + (v1[k] + v2[k] + cvo[k]) * fourth) * v4;
        v4 = fabs(cr[0] * cvj[0] + cr[1] * cvj[1] + cr[2] * cvj[2]);
        vj += v4;
#pragma unroll
        for (int k = 0; k < 3; k++)
            mj[k] += ((fc[k] + (v1[k] + v2[k] + cvj[k]) * fourth) - A[k]) * v4;
#pragma unroll
        for (int k = 0; k < 3; k++) v1[k] = v2[k];
    }
    store_side<3>(row.o, vo, mo);
    store_side<3>(row.j, vj, mj);
}

// 2D face f (orc_face_geometry): the segment between the circumcentres of
// a, P_0 and a, P_1.
__device__ __forceinline__ void face2(const phd_topology& t, const phd_geometry& g, long long f,
                                      int r0, const double* xo, const double* a, Rows row)
{
    constexpr int W = 3;
    double p[2], w0[2], w1[2], fm[2];
    ring_pos<2>(t, g, r0, xo, p);
    circ2(a, p, w0);
    ring_pos<2>(t, g, r0 + 1, xo, p);
    circ2(a, p, w1);
    const double h = sqrt(a[0] * a[0] + a[1] * a[1]);
    const double xe = w1[0] - w0[0];
    const double ye = w1[1] - w0[1];
    const double ar = sqrt(xe * xe + ye * ye);
    fm[0] = 0.5 * (w0[0] + w1[0]);
    fm[1] = 0.5 * (w0[1] + w1[1]);
    const double wt = 0.25 * ar * h;
    g.area[f] = ar;
    double out[2 * W];
    out[0] = wt;
    out[W] = wt;
#pragma unroll
    for (int k = 0; k < 2; k++) {
        g.normal[k][f] = a[k] / h;
        g.com[k][f] = xo[k] + fm[k];
        out[1 + k] = wt * (2.0 * fm[k] / 3.0);
        out[W + 1 + k] = wt * (2.0 * (fm[k] - a[k]) / 3.0);
    }
    store_side<2>(row.o, out[0], out + 1);
    store_side<2>(row.j, out[W], out + W + 1);
}

// One face per thread; its two workspace rows (struct Rows).  In 3D
// the block first sorts its faces by ring length (lanes then take them in
// descending order, so a wave's loops run about as long as its faces' rings
// rather than its longest one; every face's arithmetic is its own, so the
// results do not depend on the lane) and stages its ring entries' generator
// images in LDS (coalesced index loads, all gathers in flight at once), so
// the three vertex passes read LDS instead of repeating dependent gathers.
template <int D>
__global__ __launch_bounds__(PHD_BLOCK) void k_geo_faces(phd_topology t, phd_geometry g,
                                                         double* __restrict__ ws,
                                                         int32_t* __restrict__ owner_start,
                                                         int cap, unsigned long long* err)
{
    constexpr int W = 1 + D;
    long long f = xcd_index(t.n_faces);
    RingStage L{nullptr, nullptr, nullptr, nullptr, 0, 0};
    if constexpr (D == 3) {
        constexpr int NB = 32;
        __shared__ int s_cur[NB];
        __shared__ long long s_face[PHD_BLOCK];
        __shared__ double s_px[RING_CAP], s_py[RING_CAP], s_pz[RING_CAP];
        __shared__ unsigned char s_ok[RING_CAP];
        const int tid = threadIdx.x;
        const long long fb = f - tid;
        const long long fe = fb + PHD_BLOCK < t.n_faces ? fb + PHD_BLOCK : t.n_faces;
        if (tid < NB) s_cur[tid] = 0;
        __syncthreads();
        int key = NB - 1;   // past the end: last
        if (f < t.n_faces) {
            const int mm = t.ring_off[f + 1] - t.ring_off[f];
            key = NB - 1 - (mm < 0 ? 0 : (mm > NB - 2 ? NB - 2 : mm));
        }
        atomicAdd(&s_cur[key], 1);
        __syncthreads();
        if (tid < 64) {
            // exclusive prefix over the NB bins, one wave
            const int c = tid < NB ? s_cur[tid] : 0;
            int incl = c;
#pragma unroll
            for (int d = 1; d < NB; d <<= 1) {
                const int u = __shfl_up(incl, d, 64);
                if (tid >= d) incl += u;
            }
            if (tid < NB) s_cur[tid] = incl - c;
        }
        __syncthreads();
        s_face[atomicAdd(&s_cur[key], 1)] = f;
        // stage the block's ring range [ring_off[fb], ring_off[fe])
        int rb = 0, staged = 0;
        if (fb < fe) {
            const int n_ring = t.ring_off[t.n_faces];
            rb = t.ring_off[fb];
            const int re = t.ring_off[fe];
            if (rb >= 0 && re >= rb && re <= n_ring) staged = re - rb < cap ? re - rb : cap;
        }
        for (int q = tid; q < staged; q += PHD_BLOCK) {
            const int k = t.ring_k[rb + q], code = t.ring_code[rb + q];
            const bool ok = k >= 0 && k < t.n_real && code_ok<3>(t, code);
            s_ok[q] = ok ? 1 : 0;
            s_px[q] = ok ? image_coord(t, g.x[0][k], 0, code) : 0.0;
            s_py[q] = ok ? image_coord(t, g.x[1][k], 1, code) : 0.0;
            s_pz[q] = ok ? image_coord(t, g.x[2][k], 2, code) : 0.0;
        }
        __syncthreads();
        f = s_face[tid];
        L = RingStage{s_px, s_py, s_pz, s_ok, rb, staged};
    }
    if (f >= t.n_faces) return;
    const int na = t.n_real + t.n_ghost;
    const int o = t.pair_i[f], j = t.pair_j[f];
    const int r0 = t.ring_off[f], m = t.ring_off[f + 1] - r0;
    const int n_ring = t.ring_off[t.n_faces];
    const int prev = f > 0 ? t.pair_i[f - 1] : -1;
    bool ok = o >= 0 && o < t.n_real && j >= 0 && j < na && j != o && r0 >= 0 &&
              r0 + m <= n_ring && (D == 2 ? m == 2 : m >= 3) && prev <= o;
    // first and one-past-last face of its owner's run (faces are in owner
    // order); owner_start holds 2 n_real entries: starts, then ends
    if (ok && prev != o) owner_start[o] = (int32_t)f;
    if (ok && (f + 1 == t.n_faces || t.pair_i[f + 1] != o))
        owner_start[t.n_real + o] = (int32_t)(f + 1);
    for (int r = 0; ok && r < m; r++) {
        if constexpr (D == 3) {
            ok = ring_ok(t, L, r0 + r);
        } else {
            const int k = t.ring_k[r0 + r];
            ok = k >= 0 && k < t.n_real && code_ok<D>(t, t.ring_code[r0 + r]);
        }
    }
    const Rows row{ws + (size_t)f * W, ws + ((size_t)t.n_faces + (size_t)f) * W};
    if (!ok) {
        atomicMin(err, (unsigned long long)f);
        const double z[3] = {0.0, 0.0, 0.0};
        store_side<D>(row.o, 0.0, z);
        store_side<D>(row.j, 0.0, z);
        return;
    }
    double xo[D], a[D];
#pragma unroll
    for (int k = 0; k < D; k++) {
        xo[k] = g.x[k][o];
        a[k] = g.x[k][j] - xo[k];
    }
    if constexpr (D == 2) face2(t, g, f, r0, xo, a, row);
    else face3(t, g, L, f, r0, m, xo, a, row);
}

// One real cell: sum of its faces' side contributions in ascending face id
// (the order of its face list).  The faces before its own run belong to
// lower cells (plane-1 rows, gathered); its own run [start, end) is one
// contiguous id range (plane-0 rows, streamed) and is the tail of its list.
template <int D>
__global__ __launch_bounds__(PHD_BLOCK) void k_geo_cells(phd_topology t, phd_geometry g,
                                                         const double* __restrict__ ws,
                                                         const int32_t* __restrict__ owner_start,
                                                         unsigned long long* err)
{
    constexpr int W = 1 + D;
    const long long i = xcd_index(t.n_real);
    if (i >= t.n_real) return;
    double v = 0.0, mm[3] = {0.0, 0.0, 0.0};
    const int e0 = t.nbr_off[i], e1 = t.nbr_off[i + 1];
    const int s0 = owner_start[i], s1 = owner_start[t.n_real + i];
    const int nown = s0 == INT32_MAX ? 0 : s1 - s0;
    const int ninc = e1 - e0 - nown;
    if (e0 < 0 || e1 < e0 || e1 > t.nbr_off[t.n_real + t.n_ghost] || nown < 0 || ninc < 0) {
        atomicMin(err, TOPO_CELL | (unsigned long long)i);
        return;
    }
    const double* plane1 = ws + (size_t)t.n_faces * W;
    for (int e = e0; e < e0 + ninc; e++) {
        const int fe = t.nbr_face[e];
        if (fe < 0 || fe >= t.n_faces || fe >= s0) {
            atomicMin(err, TOPO_CELL | (unsigned long long)i);
            return;
        }
        const double* q = plane1 + (size_t)fe * W;
        v += q[0];
#pragma unroll
        for (int k = 0; k < D; k++) mm[k] += q[1 + k];
    }
    for (int u = 0; u < nown; u++) {
        const double* q = ws + (size_t)(s0 + u) * W;
        v += q[0];
#pragma unroll
        for (int k = 0; k < D; k++) mm[k] += q[1 + k];
    }
    if (D == 2) {
        g.volume[i] = v;
#pragma unroll
        for (int k = 0; k < D; k++) g.dcom[k][i] = mm[k] / v;
    } else {
#pragma unroll
        for (int k = 0; k < D; k++) g.dcom[k][i] = mm[k] * (1.0 / v);
        g.volume[i] = v * (1.0 / 6.0);
    }
}

template <int D>
__global__ __launch_bounds__(PHD_BLOCK) void k_geo_ghost_cells(phd_topology t, phd_geometry g)
{
    const long long q = xcd_index(t.n_ghost);
    if (q >= t.n_ghost) return;
    const int k = t.ghost_map[q];
    if (k < 0 || k >= t.n_real) return;   // reported by k_geo_ghost_pos
    g.volume[t.n_real + q] = g.volume[k];
#pragma unroll
    for (int a = 0; a < D; a++) g.dcom[a][t.n_real + q] = g.dcom[a][k];
}

__global__ void k_geo_fill_i32(int32_t* a, long long n, int32_t v)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// workspace: the two planes of per-face rows, then each real cell's first
// owned face and one past its last
inline size_t ws_rows_bytes(int dim, long long n_faces)
{
    const size_t b = (size_t)n_faces * 2 * (size_t)(1 + dim) * sizeof(double);
    return (b + 255) & ~(size_t)255;
}

// LDS entries staged per block: RING_CAP, or less when PHD_GEOMETRY_STAGE_CAP
// says so (0 sends every 3D face through the recomputing path; for tests)
int stage_cap()
{
    const char* e = getenv("PHD_GEOMETRY_STAGE_CAP");
    if (!e || !*e) return RING_CAP;
    const long v = strtol(e, nullptr, 10);
    return v < 0 ? 0 : (v > RING_CAP ? RING_CAP : (int)v);
}

template <int D>
int launch_geometry(phd_handle* h, const phd_topology& t, const phd_geometry& g, void* work,
                    unsigned long long* err)
{
    hipStream_t s = h->stream;
    double* ws = (double*)work;
    int32_t* owner_start = (int32_t*)((char*)work + ws_rows_bytes(D, t.n_faces));
    k_geo_fill_i32<<<nblocks(t.n_real), PHD_BLOCK, 0, s>>>(owner_start, t.n_real, INT32_MAX);
    k_geo_fill_i32<<<nblocks(t.n_real), PHD_BLOCK, 0, s>>>(owner_start + t.n_real, t.n_real, 0);
    if (t.n_ghost > 0)
        k_geo_ghost_pos<D><<<xcd_grid(t.n_ghost), PHD_BLOCK, 0, s>>>(t, g, err);
    if (t.n_faces > 0)
        k_geo_faces<D><<<xcd_grid(t.n_faces), PHD_BLOCK, 0, s>>>(t, g, ws, owner_start,
                                                                 stage_cap(), err);
    k_geo_cells<D><<<xcd_grid(t.n_real), PHD_BLOCK, 0, s>>>(t, g, ws, owner_start, err);
    if (t.n_ghost > 0)
        k_geo_ghost_cells<D><<<xcd_grid(t.n_ghost), PHD_BLOCK, 0, s>>>(t, g);
    GLAUNCHCHK();
    return 0;
}

} // namespace

extern "C" {

size_t phd_geometry_workspace_bytes(int dim, int64_t n_real, int64_t n_faces)
{
    if ((dim != 2 && dim != 3) || n_faces < 0 || n_real < 0) return 0;
    return ws_rows_bytes(dim, n_faces) + 2 * (size_t)n_real * sizeof(int32_t);
}

int phd_extract_geometry(phd_handle* h, const phd_topology* t, const phd_geometry* g,
                         void* workspace, size_t workspace_bytes, int64_t* bad)
{
    if (bad) *bad = -1;
    if (!h || !t || !g || (t->dim != 2 && t->dim != 3) || (t->boundary != 0 && t->boundary != 1))
        return PHD_E_ARG;
    const int D = t->dim;
    if (t->n_real <= 0 || t->n_ghost < 0 || t->n_faces < 0) return PHD_E_ARG;
    if (!t->pair_i || !t->pair_j || !t->ring_off || !t->nbr_off || !t->nbr_face) return PHD_E_ARG;
    if (t->n_faces > 0 && (!t->ring_k || !t->ring_code || !g->area)) return PHD_E_ARG;
    if (t->n_ghost > 0 && (!t->ghost_map || !t->ghost_code)) return PHD_E_ARG;
    if (!g->volume) return PHD_E_ARG;
    for (int k = 0; k < D; k++) {
        if (!g->x[k] || !g->dcom[k]) return PHD_E_ARG;
        if (t->n_faces > 0 && (!g->com[k] || !g->normal[k])) return PHD_E_ARG;
        if (!(t->hi[k] > t->lo[k])) return PHD_E_ARG;
    }
    if (workspace_bytes < phd_geometry_workspace_bytes(D, t->n_real, t->n_faces) || !workspace)
        return PHD_E_ARG;
    unsigned long long* err = h->d_err + PHD_ERR_SLOT_TOPO;
    k_fill_u64<<<1, 64, 0, h->stream>>>(err, 1, ~0ull);
    const int rc = D == 2 ? launch_geometry<2>(h, *t, *g, workspace, err)
                          : launch_geometry<3>(h, *t, *g, workspace, err);
    if (rc) return rc;
    GHIPCHK(hipMemcpyAsync(h->h_err + PHD_ERR_SLOT_TOPO, err, sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, h->stream));
    GHIPCHK(hipStreamSynchronize(h->stream));
    const unsigned long long v = h->h_err[PHD_ERR_SLOT_TOPO];
    if (v != ~0ull) {
        // a face index, or TOPO_CELL | cell (a face-list entry or ghost map
        // out of range): report the face, or -2 - cell
        if (bad) *bad = (v & TOPO_CELL) ? -2 - (int64_t)(v & ~TOPO_CELL) : (int64_t)v;
        return PHD_E_ARG;
    }
    return 0;
}

} // extern "C"
