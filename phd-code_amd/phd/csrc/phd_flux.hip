// phd_flux.hip -- MI355X (gfx950) kernels of the Voronoi face-flux pipeline
// and the C ABI declared in include/phd_flux.h.
//
// Data layout in HBM
//   API arrays (caller-owned, the reference's named fields): structure of
//   arrays, one double array per field; reals first, then ghosts.
//   Plan (library-owned, one per frozen mesh; phd_plan_create):
//     face_rec[F][W]  com(D) | area | normal(D) | |x_j - x_i|, W = 2D + 2:
//                     48 B/face in 2D, 64 B/face in 3D
//     cell_geo[Na][4] centroid x+dcom (3) | volume                  32 B/cell
//     csr[2F]         {face id, other cell | incoming bit} per entry of each
//                     cell's ascending face list (the order of
//                     Tess2d/3d::extract_geometry, tess.cpp:379-382)
//     xp[Na][8]       position(3) | primitive(5) in 3D, primitive(4) |
//                     position(2) | 0 0 in 2D, packed per step       64 B/cell
//     perm[Na]        visiting order of the sliced lists (degree-sorted
//                     windows of 256 cells)
//     sliced lists    per slot: other cell, face id | incoming bit, face
//                     centroid, area
//     g[F][D+2]       (dt*area)*flux per face, one unpadded row of D+2
//                     doubles (32 B in 2D, 40 B in 3D), written by the face kernel
//                     (coalesced) and applied by the update as -g to the
//                     owner and +g to the neighbour
//   Cell record rows rec[Na][ld] (the gradient container): primitive(D+2) |
//   gradient(D(D+2)) | centroid(D); written by the gradient kernels, read
//   whole by the face kernel (one 192 B row per 3D cell).
//   Every gather of the hot kernels therefore touches one packed row
//   instead of one cache segment per field.
//
// Scheduling: blocks are remapped so that each of the 8 XCDs processes one
// contiguous chunk of cells / faces (hardware dispatches workgroup b to XCD
// b mod 8); with Morton-ordered cells, neighbours then share that XCD's L2.
//
// Arithmetic: FP64 VALU only, compiled with -ffp-contract=off; every
// expression follows the reference's order, so results are bit-identical to
// its x86-64 build except where libm pow() and the GPU's pow() (or, for bases
// within 2^-20 of one, its binomial series) round differently.
#include <string.h>
#include <algorithm>
#include <vector>
#include "flux_common.h"

using namespace phd;

namespace {

__global__ void k_reset_err(unsigned long long* e)
{
    if (threadIdx.x < 3) e[threadIdx.x] = NOERR;
    if (threadIdx.x == 3) e[3] = INF_BITS;
}

__global__ void k_set_err(unsigned long long* e, int slot, unsigned long long v)
{
    if (threadIdx.x == 0) atomicMin(&e[slot], v);
}

// first cell with tag t into *first (atomic min)
__global__ __launch_bounds__(PHD_BLOCK) void k_first_tag(int n, const int8_t* __restrict__ tag,
                                                         int t, unsigned long long* first)
{
    const long long i = (long long)blockIdx.x * PHD_BLOCK + threadIdx.x;
    if (i < n && tag[i] == t) atomicMin(first, (unsigned long long)i);
}


// ---------------------------------------------------------------------------
// equation of state (equation_state.pyx:43-117), all cells
template <int D>
__global__ __launch_bounds__(PHD_BLOCK) void k_cons_from_prim(phd_cells c, double gamma)
{
    long long i = xcd_index(c.n_cells);
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
    long long i = xcd_index(c.n_cells);
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

template <int D>
__global__ __launch_bounds__(PHD_BLOCK) void k_cfl(phd_cells c, double gamma, int boost,
                                                   unsigned long long* err)
{
    long long i = (long long)blockIdx.x * PHD_BLOCK + threadIdx.x;
    double v = __longlong_as_double((long long)INF_BITS);
    if (i < c.n_cells && c.tag[i] == 0) v = cfl_value<D>(c, i, gamma, boost);
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
    __shared__ double wmin[PHD_BLOCK / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) wmin[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = wmin[0];
        for (int w = 1; w < PHD_BLOCK / 64; w++) b = fmin(b, wmin[w]);
        atomicMin(&err[3], (unsigned long long)__double_as_longlong(b));
    }
}

// generator velocities (mesh.pyx:363-426), all cells
template <int D>
__global__ __launch_bounds__(PHD_BLOCK) void k_gen_vel(phd_cells c, double gamma, double eta,
                                                       int regularize)
{
    long long i = xcd_index(c.n_cells);
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
    long long m = xcd_index(f.n_faces);
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
// plan construction (once per frozen mesh)
template <int D>
__global__ __launch_bounds__(PHD_BLOCK) void k_plan_faces(phd_cells c, phd_faces f, double* rec)
{
    long long m = xcd_index(f.n_faces);
    if (m >= f.n_faces) return;
    using F = FrLay<D>;
    int i = f.pair_i[m], j = f.pair_j[m];
    double r[F::W];
    double diff = 0.0;
#pragma unroll
    for (int k = 0; k < D; k++) {
        r[F::C + k] = f.com[k][m];
        r[F::N + k] = f.normal[k][m];
        double dx = c.x[k][j] - c.x[k][i];
        diff += dx * dx;          // reconstruction.pyx:866
    }
    r[F::A] = f.area[m];
    r[F::DIFF] = sqrt(diff);      // reconstruction.pyx:875
    double2* o = reinterpret_cast<double2*>(rec + (size_t)m * F::W);
#pragma unroll
    for (int q = 0; q < F::W / 2; q++) o[q] = make_double2(r[2 * q], r[2 * q + 1]);
}

template <int D>
__global__ __launch_bounds__(PHD_BLOCK) void k_plan_cells(phd_cells c, double* geo, double* xp)
{
    long long i = xcd_index(c.n_cells);
    if (i >= c.n_cells) return;
    double r[GEO] = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < D; k++) r[k] = c.x[k][i] + c.dcom[k][i];   // reconstruction.pyx:626, 861
    r[3] = c.volume[i];
    double2* o = reinterpret_cast<double2*>(geo + (size_t)i * GEO);
    o[0] = make_double2(r[0], r[1]);
    o[1] = make_double2(r[2], r[3]);
    if (D == 2)   // the packed row's centroid pair
        reinterpret_cast<double2*>(xp + (size_t)i * XPW + XpLay<D>::C)[0] = make_double2(r[0], r[1]);
}

// per entry {face id, other cell | incoming bit}; the neighbor check of
// reconstruction.pyx:647-652 is evaluated here once per mesh
__global__ __launch_bounds__(PHD_BLOCK) void k_plan_csr(phd_cells c, phd_faces f, int2* csr,
                                                        unsigned long long* bad)
{
    long long i = xcd_index(c.n_cells);
    if (i >= c.n_cells) return;
    const int beg = f.nbr_off[i], end = f.nbr_off[i + 1];
    for (int e = beg; e < end; e++) {
        int fid = f.nbr_face[e];
        int a = f.pair_i[fid], b = f.pair_j[fid];
        int other;
        if (a == (int)i) other = b;
        else if (b == (int)i) other = a | INCOMING;
        else {
            other = INVALID;
            if (c.tag[i] == 0) atomicMin(bad, (unsigned long long)i);
        }
        csr[e] = make_int2(fid, other);
    }
}

// owner_runs check: in every real cell's list, the owned faces (not
// incoming) are consecutive ids after all incoming faces; clears flag[0]
// (NOERR) otherwise
__global__ __launch_bounds__(PHD_BLOCK) void k_plan_runs(phd_cells c, const int2* csr,
                                                         const int32_t* off,
                                                         unsigned long long* flag)
{
    long long i = xcd_index(c.n_cells);
    if (i >= c.n_cells || c.tag[i] != 0) return;
    bool owned = false, ok = true;
    int prev = 0;
    for (int e = off[i]; e < off[i + 1]; e++) {
        const int2 en = csr[e];
        if ((en.y & 0x7fffffff) == INVALID) { ok = false; break; }
        if (en.y & INCOMING) {
            if (owned) { ok = false; break; }
        } else {
            if (owned && en.x != prev + 1) { ok = false; break; }
            owned = true;
            prev = en.x;
        }
    }
    if (!ok) atomicMin(flag, 0ull);
}

constexpr int SLICE = 16;
constexpr int SORT_WINDOW = PHD_BLOCK;   // cells per degree-sorted window

// width of each 16-cell slice = the largest face-list length of its real cells
// position -> cell order: inside each window of SORT_WINDOW consecutive
// cells, real cells by descending face count (ties by id), then ghosts
__global__ __launch_bounds__(PHD_BLOCK) void k_plan_perm(phd_cells c, const int32_t* off,
                                                         int32_t* perm)
{
    __shared__ int key[SORT_WINDOW];
    const long long w0 = (long long)blockIdx.x * SORT_WINDOW;
    const int l = threadIdx.x;
    const long long i = w0 + l;
    int k = -2;                                   // past the end
    if (i < c.n_cells) k = c.tag[i] == 0 ? off[i + 1] - off[i] : -1;
    key[l] = k;
    __syncthreads();
    int rank = 0;
    for (int m = 0; m < SORT_WINDOW; m++) {
        const int km = key[m];
        rank += (km > k) | ((km == k) & (m < l));
    }
    if (i < c.n_cells) perm[w0 + rank] = (int32_t)i;
}

__global__ __launch_bounds__(PHD_BLOCK) void k_plan_slice_width(phd_cells c, const int32_t* off,
                                                                const int32_t* perm,
                                                                int32_t* width, int n_slices)
{
    long long s = (long long)blockIdx.x * PHD_BLOCK + threadIdx.x;
    if (s >= n_slices) return;
    int w = 0;
    for (int l = 0; l < SLICE; l++) {
        long long t = s * SLICE + l;
        if (t >= c.n_cells) break;
        const int i = perm[t];
        if (c.tag[i] == 0) w = max(w, off[i + 1] - off[i]);
    }
    width[s] = w;
}

__global__ __launch_bounds__(PHD_BLOCK) void k_fill_i32(int32_t* a, long long n, int32_t v)
{
    long long i = (long long)blockIdx.x * PHD_BLOCK + threadIdx.x;
    if (i < n) a[i] = v;
}

// slot contents for real cells; per face the slots it feeds
template <int D>
__global__ __launch_bounds__(PHD_BLOCK) void k_plan_slots(phd_cells c, phd_faces f,
                                                          const int2* csr, const int32_t* off,
                                                          const int32_t* perm,
                                                          const int32_t* sbase, int32_t* sj,
                                                          double* sx, double* sy, double* sz,
                                                          double* sa, int32_t* sfi, int2* cinfo)
{
    long long t = xcd_index(c.n_cells);
    if (t >= c.n_cells) return;
    const int i = perm[t];
    if (c.tag[i] != 0) return;
    const int beg = off[i], end = off[i + 1];
    const long long base = sbase[t / SLICE] + (t % SLICE);
    int n_in = 0, own0 = -1;
    for (int e = beg; e < end; e++) {
        const int2 en = csr[e];
        const long long slot = base + (long long)(e - beg) * SLICE;
        if (en.y & INCOMING) n_in++;
        else if (own0 < 0) own0 = en.x;
        sj[slot] = en.y;
        const int fid = en.x;
        sx[slot] = f.com[0][fid];
        sy[slot] = f.com[1][fid];
        sz[slot] = D == 3 ? f.com[D - 1][fid] : 0.0;
        sa[slot] = f.area[fid];
        sfi[slot] = (en.y & 0x7fffffff) == INVALID ? INVALID : (fid | (en.y & INCOMING));
    }
    cinfo[t] = make_int2(n_in, own0);
}

// Exclusive prefix sum of one int per thread over a block of SORT_WINDOW
// threads; total receives the block's sum.  part: SORT_WINDOW ints of LDS.
__device__ __forceinline__ int block_scan_excl(int v, int* part, int& total)
{
    const int tid = threadIdx.x;
    part[tid] = v;
    __syncthreads();
    for (int d = 1; d < SORT_WINDOW; d <<= 1) {
        const int a = tid >= d ? part[tid - d] : 0;
        __syncthreads();
        part[tid] += a;
        __syncthreads();
    }
    total = part[SORT_WINDOW - 1];
    const int ex = part[tid] - v;
    __syncthreads();
    return ex;
}

// The 3D gradient kernel's halo of window w (cells [w0, w1) = positions
// [w0, w1), since the degree sort permutes within windows), one block per
// window: the list entries of its real cells that name a cell outside the
// window are gathered into LDS, sorted (bitonic) and made distinct; the
// first hm are the halo, halo[w*hm + k] in ascending order, hcnt[w] of them.
// Each list slot then gets its staging row: j - w0 inside the window,
// SORT_WINDOW + k for the k-th halo cell, LOC_GLOBAL for a distinct cell past
// hm, LOC_NONE for an invalid neighbour (padding slots keep the LOC_NONE the
// host fills in).  A window whose lists name more than HALO_CAP outside
// entries keeps no halo: all of them are LOC_GLOBAL (read from HBM).
constexpr int HALO_CAP = 4096;

__global__ __launch_bounds__(SORT_WINDOW) void k_plan_halo(long long na,
                                                           const int8_t* __restrict__ tag,
                                                           const int32_t* __restrict__ off,
                                                           const int32_t* __restrict__ perm,
                                                           const int32_t* __restrict__ sbase,
                                                           const int32_t* __restrict__ sj,
                                                           int hm, int32_t* __restrict__ hcnt,
                                                           int32_t* __restrict__ halo,
                                                           uint16_t* __restrict__ sloc)
{
    __shared__ int32_t key[HALO_CAP];
    __shared__ int part[SORT_WINDOW];
    const long long w = blockIdx.x;
    const long long w0 = w * SORT_WINDOW, w1 = min(na, w0 + SORT_WINDOW);
    const long long t = w0 + threadIdx.x;
    int deg = 0;
    long long base = 0;
    if (t < w1) {
        const int i = perm[t];
        if (tag[i] == 0) {
            deg = off[i + 1] - off[i];
            base = sbase[t / SLICE] + (t % SLICE);
        }
    }
    auto outside = [&](int j) { return j != INVALID && (j < w0 || j >= w1); };
    int mine = 0;
    for (int r = 0; r < deg; r++) mine += outside(sj[base + (long long)r * SLICE] & 0x7fffffff);
    int n = 0;
    int at = block_scan_excl(mine, part, n);
    const bool fit = n <= HALO_CAP;               // block-uniform
    int u = 0;
    if (fit) {
        for (int r = 0; r < deg && mine > 0; r++) {
            const int j = sj[base + (long long)r * SLICE] & 0x7fffffff;
            if (outside(j)) key[at++] = j;
        }
        int np = 1;
        while (np < n) np <<= 1;
        for (int k = n + threadIdx.x; k < np; k += SORT_WINDOW) key[k] = INT_MAX;
        __syncthreads();
        for (int k = 2; k <= np; k <<= 1) {
            for (int d = k >> 1; d > 0; d >>= 1) {
                for (int a = threadIdx.x; a < np; a += SORT_WINDOW) {
                    const int b = a ^ d;
                    if (b > a) {
                        const int x = key[a], y = key[b];
                        if ((x > y) == ((a & k) == 0)) { key[a] = y; key[b] = x; }
                    }
                }
                __syncthreads();
            }
        }
        // distinct values of key[0, n): each thread reads a run of at most
        // 16, counts the first of each value, and after the block's reads
        // and scan writes them packed from the front
        constexpr int CH = HALO_CAP / SORT_WINDOW;
        const int c0 = threadIdx.x * CH, c1 = min(n, c0 + CH);
        int v[CH];
        unsigned first = 0;
        int nu = 0;
        int last = (c0 > 0 && c0 < n) ? key[c0 - 1] : -1;
#pragma unroll
        for (int q = 0; q < CH; q++) {
            v[q] = 0;
            if (c0 + q < c1) {
                v[q] = key[c0 + q];
                if (v[q] != last) { first |= 1u << q; nu++; }
                last = v[q];
            }
        }
        int o = block_scan_excl(nu, part, u);     // its barriers follow every read above
#pragma unroll
        for (int q = 0; q < CH; q++)
            if ((first >> q) & 1u) key[o++] = v[q];
        __syncthreads();
    }
    const int kept = min(u, hm);
    if (threadIdx.x == 0) hcnt[w] = kept;
    for (int k = threadIdx.x; k < kept; k += SORT_WINDOW) halo[w * hm + k] = key[k];
    for (int r = 0; r < deg; r++) {
        const long long slot = base + (long long)r * SLICE;
        const int j = sj[slot] & 0x7fffffff;
        uint16_t loc = LOC_NONE;
        if (j != INVALID) {
            if (j >= w0 && j < w1) {
                loc = (uint16_t)(j - w0);
            } else if (!fit) {
                loc = LOC_GLOBAL;
            } else {
                int lo = 0, hi = u;                   // first key >= j (present)
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (key[mid] < j) lo = mid + 1;
                    else hi = mid;
                }
                loc = lo < hm ? (uint16_t)(SORT_WINDOW + lo) : LOC_GLOBAL;
            }
        }
        sloc[slot] = loc;
    }
}

// ---------------------------------------------------------------------------
// packed [position | primitive] rows for the neighbour gathers
template <int D>
__global__ __launch_bounds__(PHD_BLOCK) void k_pack_xp(phd_cells c, double* xp)
{
    long long i = xcd_index(c.n_cells);
    if (i >= c.n_cells) return;
    double r[XPW] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < D; k++) r[XpLay<D>::X + k] = c.x[k][i];
#pragma unroll
    for (int n = 0; n < D + 2; n++) r[XpLay<D>::P + n] = c.prim[n][i];
    double2* o = reinterpret_cast<double2*>(xp + (size_t)i * XPW);
#pragma unroll
    for (int q = 0; q < XpLay<D>::REFRESH_PAIRS; q++) o[q] = make_double2(r[2 * q], r[2 * q + 1]);
}

// ---------------------------------------------------------------------------
// Division by one divisor shared by many numerators.  hipcc lowers a / b on
// gfx950 to div_scale, rcp, two Newton steps on the reciprocal, q0 = a*r,
// e = fma(-b, q0, a), div_fmas, div_fixup.  For |a|, |b| in [2^-200, 2^200]
// div_scale returns its operand unchanged with VCC clear (no operand or
// quotient is near the overflow/underflow limits), div_fmas reduces to
// fma(e, r, q0) and div_fixup to the identity, so the steps below give the
// bits of a / b while the reciprocal and its two Newton steps are computed
// once per divisor.  The residual is formed as b*q0 - a and subtracted, which
// is the same exact value with the opposite sign; for b > 0 that makes a zero
// numerator come out as the correctly signed zero.  Callers check the range
// and b > 0 and use the ordinary division otherwise.
__device__ __forceinline__ double rcp_newton(double b)
{
    double r = __builtin_amdgcn_rcp(b);
    double e = fma(-b, r, 1.0);
    r = fma(r, e, r);
    e = fma(-b, r, 1.0);
    return fma(r, e, r);
}

__device__ __forceinline__ double div_shared(double a, double b, double r)
{
    const double q = a * r;
    const double e = fma(b, q, -a);
    return fma(-e, r, q);
}

// |x| in [2^-200, 2^200] or x == 0; false for inf and NaN (non-short-circuit
// operators: the test compiles to compares and mask logic, no branches)
__device__ __forceinline__ bool div_operand_ok(double x)
{
    const double ax = fabs(x);
    return (ax <= 0x1p200) & ((ax >= 0x1p-200) | (x == 0.0));
}

// one entry of a sliced list: neighbour id (INCOMING bit cleared), face
// centroid, face area
template <int D> struct SlotEnt {
    int j;
    double com[3];
    double area;
};

template <int D, bool J = true>
__device__ __forceinline__ SlotEnt<D> load_slot(const int32_t* sj, const double* sx,
                                                const double* sy, const double* sz,
                                                const double* sa, long long slot)
{
    SlotEnt<D> s;
    s.j = J ? __builtin_nontemporal_load(sj + slot) & 0x7fffffff : INVALID;   // !J: j unread
    s.com[0] = sx[slot];               // cached: the limiter pass reads them again
    s.com[1] = sy[slot];
    s.com[2] = D == 3 ? sz[slot] : 0.0;
    s.area = __builtin_nontemporal_load(sa + slot);
    return s;
}

// packed [position | primitive] row of cell j
__device__ __forceinline__ void load_row(const double* xp, int j, double* r)
{
#pragma unroll
    for (int q = 0; q < XPW / 2; q++) ld2(xp + (size_t)j * XPW + 2 * q, r[2 * q], r[2 * q + 1]);
}

// gradients + slope limiter (reconstruction.pyx:563-737), one thread per
// real cell; writes the cell record row [prim | alpha*df | centroid].  The
// neighbour loop walks the cell's sliced list (sbase/sj/sgeo, see phd_plan),
// so the per-entry face data is streamed with one coalesced load per field
// per 16 lanes and only the packed [position | primitive] row of each
// neighbour is gathered.
//
// The AREPO limiter (683-703) takes alpha_n = min(1, min_faces fmax(psi,0))
// with psi = (phi_max - phi_i)/dphi for dphi > 0 and (phi_min - phi_i)/dphi
// for dphi < 0.  The numerators are fixed per cell (>= 0 and <= 0) and a
// correctly rounded quotient is monotone in its divisor, so the minimum over
// faces is reached at the largest positive dphi and at the most negative
// dphi: two divisions per field instead of one per face, same bits.
template <int D, int LIM, int BS = PHD_BLOCK, int XCD = 1>
__global__ __launch_bounds__(BS) void k_grad_sell(int n_cells, const int8_t* __restrict__ tag,
                                                         const double* __restrict__ geo,
                                                         const double* __restrict__ vols,
                                                         const double* __restrict__ xp,
                                                         const int32_t* __restrict__ off,
                                                         const int32_t* __restrict__ perm,
                                                         const int32_t* __restrict__ sbase,
                                                         const int32_t* __restrict__ sj,
                                                         const double* __restrict__ sx,
                                                         const double* __restrict__ sy,
                                                         const double* __restrict__ sz,
                                                         const double* __restrict__ sa,
                                                         const int32_t* __restrict__ hcnt,
                                                         const int32_t* __restrict__ halo,
                                                         const uint16_t* __restrict__ sloc,
                                                         double* __restrict__ rec, int ld)
{
    using Y = Lay<D>;
    constexpr int NF = Y::NF;
    // A block of SORT_WINDOW threads owns one whole window of cells (the
    // degree sort permutes within windows): every thread stages its cell's
    // packed row in LDS, and the block then gathers the window's halo (the
    // distinct outside cells its lists name, up to HM of them) into the rows
    // after it, all at once.  Each slot's sloc names the staged row, so the
    // entry loop reads every neighbour from LDS: no entry waits on a
    // per-lane HBM gather (about 1/4 of the entries in 3D, 1/10 in 2D
    // lie outside the window; a wave would wait for its slowest lane).
    constexpr bool STAGE = BS == SORT_WINDOW;
    constexpr int HM = STAGE ? halo_max<D>() : 0;
    constexpr bool HALO = HM > 0;
    __shared__ double2 rows[STAGE ? (SORT_WINDOW + HM) * XPW / 2 : 1];
    // 2D AREPO: the first KC entries of each list keep their face centroid
    // in LDS from the gradient pass for the limiter pass ([entry][thread],
    // thread-private), so the limiter streams only the later slots again.
    // 32 KB with the rows: four blocks still share a CU, as the 128 VGPRs
    // allow (4096-problem step -0.8 %).  In 3D the 6 entries that would fit
    // beside three blocks changed nothing measurable.
    constexpr int KC = STAGE && LIM == PHD_LIMITER_AREPO && D == 2 ? 4 : 0;
    __shared__ double2 ccache[KC > 0 ? KC * SORT_WINDOW : 1];
    const long long t = sched_index<BS, XCD>(n_cells);
    const bool live = t < n_cells;
    const int i = live ? perm[t] : 0;
    const long long w0 = t - threadIdx.x;          // the window's first cell
    double own[XPW];
    if (live) {
#pragma unroll
        for (int q = 0; q < XPW / 2; q++) ld2(xp + (size_t)i * XPW + 2 * q, own[2 * q], own[2 * q + 1]);
    }
    if (STAGE) {
        if (live) {
#pragma unroll
            for (int q = 0; q < XPW / 2; q++)
                rows[(i - w0) * (XPW / 2) + q] = make_double2(own[2 * q], own[2 * q + 1]);
        }
        if (HALO && w0 < n_cells) {
            const long long w = w0 / SORT_WINDOW;
            const int hn = hcnt[w];                // at most HM
            for (int k = threadIdx.x; k < hn; k += BS) {
                const int j = halo[w * HM + k];
#pragma unroll
                for (int q = 0; q < XPW / 2; q++) {
                    double a, b;
                    ld2(xp + (size_t)j * XPW + 2 * q, a, b);
                    rows[(SORT_WINDOW + k) * (XPW / 2) + q] = make_double2(a, b);
                }
            }
        }
        __syncthreads();
    }
    if (!live || tag[i] != 0) return;

    double cx[D], vol, xi[D], pv[NF];
    {
        if (D == 2) {
            // the packed row's centroid pair and the dense volume array: 24
            // bytes per cell less than the 32-byte geometry row
#pragma unroll
            for (int k = 0; k < D; k++) cx[k] = own[XpLay<D>::C + k];
            vol = vols[i];
        } else {
            double g[GEO];
            ld2(geo + (size_t)i * GEO, g[0], g[1]);
            ld2(geo + (size_t)i * GEO + 2, g[2], g[3]);
#pragma unroll
            for (int k = 0; k < D; k++) cx[k] = g[k];
            vol = g[3];
        }
#pragma unroll
        for (int k = 0; k < D; k++) xi[k] = own[XpLay<D>::X + k];
#pragma unroll
        for (int n = 0; n < NF; n++) pv[n] = own[XpLay<D>::P + n];
    }
    double pmax[NF], pmin[NF], df[NF * D];
#pragma unroll
    for (int n = 0; n < NF; n++) {
        pmax[n] = pmin[n] = pv[n];
#pragma unroll
        for (int k = 0; k < D; k++) df[n * D + k] = 0.0;
    }
    const int deg = off[i + 1] - off[i];
    const long long base = sbase[t / SLICE] + (t % SLICE);
    // GB entries in flight: their slot loads, then their row reads, then the
    // entries applied one by one in list order (the same sums).  Two pay in
    // 2D; in 3D the extra registers cost a wave per SIMD.  LB: face centroids
    // in flight in the limiter pass (two measured best in 2D and 3D)
    constexpr int GB = D == 2 ? 2 : 1, LB = 2;
    for (int e0 = 0; e0 < deg; e0 += GB) {
        SlotEnt<D> cur[GB];
        double rb[GB][XPW];
#pragma unroll
        for (int b = 0; b < GB; b++) {
            cur[b].j = INVALID;
            if (e0 + b < deg) {
                const long long slot = base + (long long)(e0 + b) * SLICE;
                if (HALO) {
                    // j carries the staged row; LOC_NONE marks an invalid
                    // neighbour, LOC_GLOBAL one past the staged halo
                    cur[b] = load_slot<D, false>(sj, sx, sy, sz, sa, slot);
                    const int loc = __builtin_nontemporal_load(sloc + slot);
                    cur[b].j = loc == LOC_NONE ? INVALID : loc;
                } else {
                    cur[b] = load_slot<D>(sj, sx, sy, sz, sa, slot);
                }
            }
        }
        if (KC > 0 && e0 < KC) {
#pragma unroll
            for (int b = 0; b < GB; b++)
                if (e0 + b < KC && e0 + b < deg)
                    ccache[(e0 + b) * SORT_WINDOW + threadIdx.x] = make_double2(cur[b].com[0], cur[b].com[1]);
        }
#pragma unroll
        for (int b = 0; b < GB; b++) {
            const int j = cur[b].j;
            if (j == INVALID) continue;
            const long long lj = (long long)j - w0;
            if (HALO && j != LOC_GLOBAL) {
#pragma unroll
                for (int q = 0; q < XPW / 2; q++) {
                    const double2 v = rows[j * (XPW / 2) + q];
                    rb[b][2 * q] = v.x;
                    rb[b][2 * q + 1] = v.y;
                }
            } else if (HALO) {
                const long long slot = base + (long long)(e0 + b) * SLICE;
                load_row(xp, sj[slot] & 0x7fffffff, rb[b]);
            } else if (STAGE && lj >= 0 && lj < SORT_WINDOW) {
#pragma unroll
                for (int q = 0; q < XPW / 2; q++) {
                    const double2 v = rows[lj * (XPW / 2) + q];
                    rb[b][2 * q] = v.x;
                    rb[b][2 * q + 1] = v.y;
                }
            } else {
                load_row(xp, j, rb[b]);
            }
        }
#pragma unroll
        for (int b = 0; b < GB; b++) {
            if (cur[b].j == INVALID) continue;
            const double* com = cur[b].com;
            const double area = cur[b].area;
            const double* r = rb[b];
            double cfx[D], dr[D], rr = 0.0;
#pragma unroll
            for (int k = 0; k < D; k++) {
                const double xj = r[XpLay<D>::X + k];
                cfx[k] = com[k] - 0.5 * (xi[k] + xj);
                dr[k] = xi[k] - xj;
                rr += dr[k] * dr[k];
            }
            rr = sqrt(rr);
            const double rv = rr * vol;
            // Springel (2010) eq. 21 summand, area*(...)/(r*vol) (677-681)
            double num[NF * D];
            bool shared = div_operand_ok(rv) & (rv > 0.0);
#pragma unroll
            for (int n = 0; n < NF; n++) {
                const double pj = r[XpLay<D>::P + n];
                pmax[n] = fmax(pmax[n], pj);
                pmin[n] = fmin(pmin[n], pj);
                const double d_dif = pj - pv[n];
                const double d_sum = pj + pv[n];
#pragma unroll
                for (int k = 0; k < D; k++) {
                    num[n * D + k] = area * (d_dif * cfx[k] - 0.5 * d_sum * dr[k]);
                    shared &= div_operand_ok(num[n * D + k]);
                }
            }
            if (shared) {
                const double rcp = rcp_newton(rv);
#pragma unroll
                for (int q = 0; q < NF * D; q++) df[q] += div_shared(num[q], rv, rcp);
            } else {
#pragma unroll
                for (int q = 0; q < NF * D; q++) df[q] += num[q] / rv;
            }
        }
    }

    double alpha[NF];
#pragma unroll
    for (int n = 0; n < NF; n++) alpha[n] = 1.0;
    if (LIM == PHD_LIMITER_AREPO) {
        double dpos[NF], dneg[NF];
#pragma unroll
        for (int n = 0; n < NF; n++) { dpos[n] = 0.0; dneg[n] = 0.0; }
        for (int e0 = 0; e0 < deg; e0 += LB) {
            double relb[LB][D];
#pragma unroll
            for (int b = 0; b < LB; b++) {
                if (e0 + b >= deg) continue;
                if (KC > 0 && e0 + b < KC) {
                    const double2 v = ccache[(e0 + b) * SORT_WINDOW + threadIdx.x];
                    relb[b][0] = v.x - cx[0];
                    relb[b][1] = v.y - cx[1];
                    continue;
                }
                const long long slot = base + (long long)(e0 + b) * SLICE;
                relb[b][0] = __builtin_nontemporal_load(sx + slot) - cx[0];
                relb[b][1] = __builtin_nontemporal_load(sy + slot) - cx[1];
                if (D == 3) relb[b][D - 1] = __builtin_nontemporal_load(sz + slot) - cx[D - 1];
            }
#pragma unroll
            for (int b = 0; b < LB; b++) {
                if (e0 + b >= deg) continue;
#pragma unroll
                for (int n = 0; n < NF; n++) {
                    double dphi = 0.0;
#pragma unroll
                    for (int k = 0; k < D; k++) dphi += df[n * D + k] * relb[b][k];
                    // dpos starts at 0 and only grows, dneg starts at 0 and
                    // only shrinks, and fmax/fmin skip NaN: the unconditional
                    // forms leave each one exactly as the reference's
                    // "dphi > 0" / "dphi < 0" branches do (a -0 is never
                    // taken as > 0 or < 0 below)
                    dpos[n] = fmax(dpos[n], dphi);
                    dneg[n] = fmin(dneg[n], dphi);
                }
            }
        }
#pragma unroll
        for (int n = 0; n < NF; n++) {
            if (dpos[n] > 0) alpha[n] = fmin(alpha[n], fmax((pmax[n] - pv[n]) / dpos[n], 0.));
            if (dneg[n] < 0) alpha[n] = fmin(alpha[n], fmax((pmin[n] - pv[n]) / dneg[n], 0.));
        }
    } else {
        // TESS limiter (reconstruction.pyx:705-731); only alpha[0] changes
        for (int n = 0; n < NF; n++) {
            for (int e = 0; e < deg; e++) {
                const long long slot = base + (long long)e * SLICE;
                const int j = sj[slot] & 0x7fffffff;
                if (j == INVALID) continue;
                double dphi = 0.0;
                const double cm[3] = {sx[slot], sy[slot], D == 3 ? sz[slot] : 0.0};
#pragma unroll
                for (int k = 0; k < D; k++) dphi += df[n * D + k] * (cm[k] - cx[k]);
                double psi;
                if (dphi > 0.0 || dphi < 0.0) {
                    double q = (xp[(size_t)j * XPW + XpLay<D>::P + n] - pv[n]) / dphi;
                    psi = (0. > q) ? 0. : q;
                } else {
                    psi = 1.0;
                }
                alpha[0] = fmin(alpha[0], fmax(psi, 0.));
            }
        }
    }
    double* row = rec + (size_t)i * ld;
    double out[Y::LD];
#pragma unroll
    for (int n = 0; n < NF; n++) out[Y::P0 + n] = pv[n];
#pragma unroll
    for (int n = 0; n < NF; n++)
#pragma unroll
        for (int k = 0; k < D; k++) out[Y::G0 + n * D + k] = alpha[n] * df[n * D + k];
#pragma unroll
    for (int k = 0; k < D; k++) out[Y::C0 + k] = cx[k];
    if (Y::LD > Y::REC) out[Y::LD - 1] = 0.0;
#pragma unroll
    for (int q = 0; q < Y::LD / 2; q++)
        reinterpret_cast<double2*>(row)[q] = make_double2(out[2 * q], out[2 * q + 1]);
}

// ghost cell records: gradient copied from the image (domain_manager.pyx:
// 829-834) with the reflective diagonal flip (boundary.pyx:291-307);
// primitive and centroid are the ghost's own
template <int D>
__global__ __launch_bounds__(PHD_BLOCK) void k_ghost_rec(phd_cells c,
                                                         const int32_t* __restrict__ gl,
                                                         int n_ghosts,
                                                         const double* __restrict__ geo,
                                                         int reflect, double lo0, double lo1,
                                                         double lo2, double hi0, double hi1,
                                                         double hi2, double* __restrict__ rec,
                                                         int ld)
{
    using Y = Lay<D>;
    const long long k = xcd_index(n_ghosts);     // one thread per listed ghost
    if (k >= n_ghosts) return;
    const int i = gl[k];
    const double* src = rec + (size_t)c.map[i] * ld;
    double out[Y::LD];
#pragma unroll
    for (int q = 0; q < Y::LD / 2; q++) ld2(src + 2 * q, out[2 * q], out[2 * q + 1]);
#pragma unroll
    for (int n = 0; n < Y::NF; n++) out[Y::P0 + n] = c.prim[n][i];
    if (reflect && c.type[i] == 2) {
        const double lo[3] = {lo0, lo1, lo2}, hi[3] = {hi0, hi1, hi2};
#pragma unroll
        for (int q = 0; q < D; q++) {
            const double xq = c.x[q][i];
            // velocity group dv[D*q + q] = d v_q / d x_q
            if (xq < lo[q]) out[Y::G0 + (1 + q) * D + q] *= -1;
            if (xq > hi[q]) out[Y::G0 + (1 + q) * D + q] *= -1;
        }
    }
#pragma unroll
    for (int k = 0; k < D; k++) out[Y::C0 + k] = geo[(size_t)i * GEO + k];
    double* row = rec + (size_t)i * ld;
#pragma unroll
    for (int q = 0; q < Y::LD / 2; q++)
        reinterpret_cast<double2*>(row)[q] = make_double2(out[2 * q], out[2 * q + 1]);
}

// ---------------------------------------------------------------------------
// face kernels
// Materialised states: PieceWiseConstant.compute_states or
// PieceWiseLinear.add_spatial / add_temporal.
template <int D, int ORDER>
__global__ __launch_bounds__(PHD_BLOCK) void k_states(phd_cells c, phd_faces f,
                                                      const double* __restrict__ frec,
                                                      const double* __restrict__ rec, int ld,
                                                      double gamma, double dt, int boost,
                                                      int gizmo, int stages, FaceArrays left,
                                                      FaceArrays right)
{
    constexpr int NF = D + 2;
    long long m = xcd_index(f.n_faces);
    if (m >= f.n_faces) return;
    const int i = f.pair_i[m], j = f.pair_j[m];
    double com[D], n[D], area, diff, fv[D], L[NF], R[NF];
    load_face_rec<D>(frec, m, com, area, n, diff);
#pragma unroll
    for (int k = 0; k < D; k++) fv[k] = f.velocity[k][m];
    if (ORDER == PHD_ORDER_LINEAR && !(stages & PHD_STAGE_SPATIAL)) {
        // temporal only: start from the stored (spatial) states
#pragma unroll
        for (int q = 0; q < NF; q++) { L[q] = left.q[q][m]; R[q] = right.q[q][m]; }
    }
    face_states<D, ORDER>(c, rec, ld, i, j, com, diff, fv, gamma, dt, boost, gizmo, stages, L, R);
#pragma unroll
    for (int q = 0; q < NF; q++) { left.q[q][m] = L[q]; right.q[q][m] = R[q]; }
}

// Riemann fluxes from materialised states (riemann.pyx:74-98)
template <int D, int SOLVER>
__global__ __launch_bounds__(PHD_BLOCK) void k_riemann(phd_faces f, FaceArrays left,
                                                       FaceArrays right, Gas gas, int boost,
                                                       FaceArrays flux, unsigned long long* err)
{
    constexpr int NF = D + 2;
    long long m = xcd_index(f.n_faces);
    if (m >= f.n_faces) return;
    double L[NF], R[NF], n[D], fv[D], F[NF];
#pragma unroll
    for (int q = 0; q < NF; q++) { L[q] = left.q[q][m]; R[q] = right.q[q][m]; }
#pragma unroll
    for (int k = 0; k < D; k++) { n[k] = f.normal[k][m]; fv[k] = f.velocity[k][m]; }
    if (!solve_face<D, SOLVER>(L, R, n, fv, gas, boost, F)) {
        atomicMin(&err[2], (unsigned long long)m);
        return;
    }
#pragma unroll
    for (int q = 0; q < NF; q++) flux.q[q][m] = F[q];
}

// Fused states + Riemann + deboost; the states never leave registers.
// OUT bit 1: write the flux arrays (riemann.fluxes); bit 2: write
// g[m] = (dt_full*area)*flux, the update's summand (mesh.pyx:530-544 forms
// dt*a*flux for each endpoint; the product is the same for both).
// BATCH: the time step is per problem, dt_full = bdt[fprob[m]] and dt_half =
// 0.5 * dt_full (the arguments are ignored), and the mesh is static (face
// velocities zero, not read).
// BG: the boost and gizmo switches as template constants (bit 0 boost, bit 1
// gizmo), so each launch runs a kernel compiled for its case: with boost the
// face-frame velocity wn is the constant +0.0 and every x - wn folds to x
// (an exact IEEE identity, so the bits are unchanged), and the dead branches
// leave the register allocation (3D Sedov face kernel 0.81 -> 0.68 ms).
template <int D, int ORDER, int SOLVER, int OUT, int BATCH = 0, int BS = PHD_BLOCK, int XCD = 1,
          int BG = 0>
__global__ __launch_bounds__(BS) void k_face(phd_cells c, phd_faces f,
                                             const double* __restrict__ frec,
                                             const double* __restrict__ rec, int ld,
                                             Gas gas, double dt_half, double dt_full,
                                             int boost, int gizmo, FaceArrays flux,
                                             double* __restrict__ g,
                                             unsigned long long* err,
                                             const double* __restrict__ bdt,
                                             const int32_t* __restrict__ fprob)
{
    constexpr int NF = D + 2;
    boost = BG & 1;
    gizmo = (BG >> 1) & 1;
    long long m = sched_index<BS, XCD>(f.n_faces);
    if (m >= f.n_faces) return;
    const int i = f.pair_i[m], j = f.pair_j[m];
    if (BATCH) {
        dt_full = bdt[fprob[m]];
        dt_half = 0.5 * dt_full;
        if (dt_full == 0.0) {              // a frozen problem: no change
            const double zero[NF] = {};
            store_summand<D>(g, m, zero);
            return;
        }
    }
    double com[D], n[D], area, diff, fv[D], L[NF], R[NF], F[NF];
    load_face_rec<D>(frec, m, com, area, n, diff);
#pragma unroll
    for (int k = 0; k < D; k++) fv[k] = BATCH ? 0.0 : f.velocity[k][m];   // batches: static meshes
    face_states<D, ORDER>(c, rec, ld, i, j, com, diff, fv, gas.gamma, dt_half, boost, gizmo,
                          PHD_STAGE_SPATIAL | PHD_STAGE_TEMPORAL, L, R);
    if (!solve_face<D, SOLVER>(L, R, n, fv, gas, boost, F)) {
        atomicMin(&err[2], (unsigned long long)m);
        return;
    }
    if (OUT & 1) {
#pragma unroll
        for (int q = 0; q < NF; q++) flux.q[q][m] = F[q];
    }
    if (OUT & 2) {
        const double t = dt_full * area;
        double row[NF];
#pragma unroll
        for (int q = 0; q < NF; q++) row[q] = t * F[q];
        store_summand<D>(g, m, row);
    }
}

// ---------------------------------------------------------------------------
// conserved update (mesh.pyx:479-551).  One thread per real cell walks the
// cell's face list in ascending face id, the order in which the reference's
// face loop reaches that cell, so every running sum is bit-identical and no
// atomics are needed: the owner (pair_i) subtracts (dt*area)*flux, the
// neighbour adds it.  The positivity check (546-551) fires at the first face
// n where a running mass or energy of either endpoint is <= 0; each endpoint
// reports its first failure as key 2n (mass) or 2n+1 (energy) and the
// minimum key wins.  Ghost endpoints never change: a ghost fails at the
// first face of its list.
template <int D>
__device__ __forceinline__ bool ghost_fault(const phd_cells& c, int i, int beg, int end,
                                            const int2* csr, unsigned long long* err)
{
    if (c.tag[i] == 0) return false;
    if (beg < end) {
        const bool bm = c.cons[0][i] <= 0.0, be = c.cons[D + 1][i] <= 0.0;
        if (bm || be) atomicMin(&err[0], 2ull * (unsigned long long)csr[beg].x + (bm ? 0ull : 1ull));
    }
    return true;
}

template <int D>
__device__ __forceinline__ bool apply_summand(double* u, const double* t, bool incoming,
                                              int fid, unsigned long long* err)
{
    constexpr int NF = D + 2;
    if (incoming) {
#pragma unroll
        for (int q = 0; q < NF; q++) u[q] += t[q];
    } else {
#pragma unroll
        for (int q = 0; q < NF; q++) u[q] -= t[q];
    }
    const bool bm = u[0] <= 0.0, be = u[D + 1] <= 0.0;
    if (bm || be) {
        atomicMin(&err[0], 2ull * (unsigned long long)fid + (bm ? 0ull : 1ull));
        return true;
    }
    return false;
}

// API path: riemann.fluxes + face areas, per-cell lists from the CSR
template <int D>
__global__ __launch_bounds__(PHD_BLOCK) void k_update(phd_cells c, const double* __restrict__ area,
                                                      FaceArrays flux,
                                                      const int2* __restrict__ csr,
                                                      const int32_t* __restrict__ off, double dt,
                                                      unsigned long long* err)
{
    constexpr int NF = D + 2;
    long long ii = xcd_index(c.n_cells);
    if (ii >= c.n_cells) return;
    const int i = (int)ii;
    const int beg = off[i], end = off[i + 1];
    if (ghost_fault<D>(c, i, beg, end, csr, err)) return;
    double u[NF];
#pragma unroll
    for (int q = 0; q < NF; q++) u[q] = c.cons[q][i];
    for (int e = beg; e < end; e++) {
        const int2 en = csr[e];
        if ((en.y & 0x7fffffff) == INVALID) continue;
        const int fid = en.x;
        double t[NF];
        const double s = dt * area[fid];
#pragma unroll
        for (int q = 0; q < NF; q++) t[q] = s * flux.q[q][fid];
        if (apply_summand<D>(u, t, en.y & INCOMING, fid, err)) break;
    }
#pragma unroll
    for (int q = 0; q < NF; q++) c.cons[q][i] = u[q];
}

// fused-step path: the per-face summand rows g of the face kernel, the
// per-cell lists as sliced face ids (sfi, coalesced) in the plan's visiting
// order.  Faces are numbered in owner order, so a cell's list is its
// incoming faces followed by one run of consecutive owned face ids
// (owner_runs): the incoming rows are gathered, the owned run is read as
// consecutive rows without touching the list again.
// Conserved update of the cell at sliced position t (cell i = perm[t]):
// u = its conserved values plus its summand rows in list order.  False for a
// ghost (nothing to write; a ghost with faces reports a fault).
template <int D, int OR>
__device__ __forceinline__ bool update_cell(const phd_cells& c, long long t, int i,
                                            const double* __restrict__ g,
                                            const int2* __restrict__ csr,
                                            const int32_t* __restrict__ off,
                                            const int32_t* __restrict__ sbase,
                                            const int32_t* __restrict__ sfi,
                                            const int2* __restrict__ cinfo,
                                            unsigned long long* err, double* u)
{
    // B summand rows in flight per batch (3D cells have ~16 faces, 2D ~6):
    // eight in both; in 2D eight took the 4096-problem ensemble step 1.4 %
    // below four, though the fused update then holds 102 VGPRs (4 waves per
    // SIMD instead of 8)
    constexpr int NF = D + 2, GW = summand_width<D>(), B = 8;
    const int beg = off[i], end = off[i + 1];
    if (ghost_fault<D>(c, i, beg, end, csr, err)) return false;
#pragma unroll
    for (int q = 0; q < NF; q++) u[q] = c.cons[q][i];
    const int deg = end - beg;
    const long long base = sbase[t / SLICE] + (t % SLICE);
    if (OR) {
        // incoming rows (gathered) then the owned run, B rows in flight at a
        // time; applied one by one in list order
        const int2 ci = cinfo[t];
        const int n_in = ci.x, n_own = deg - ci.x;
        bool stop = false;
        for (int e0 = 0; e0 < n_in && !stop; e0 += B) {
            int fb[B];
            double sb[B][GW];
#pragma unroll
            for (int b = 0; b < B; b++)
                fb[b] = e0 + b < n_in
                            ? (__builtin_nontemporal_load(sfi + base + (long long)(e0 + b) * SLICE) & 0x7fffffff)
                            : -1;
#pragma unroll
            for (int b = 0; b < B; b++)
                if (fb[b] >= 0) load_summand<D>(g, fb[b], sb[b]);
#pragma unroll
            for (int b = 0; b < B; b++)
                if (!stop && fb[b] >= 0) stop = apply_summand<D>(u, sb[b], true, fb[b], err);
        }
        for (int k0 = 0; k0 < n_own && !stop; k0 += B) {
            double sb[B][GW];
#pragma unroll
            for (int b = 0; b < B; b++)
                if (k0 + b < n_own) load_summand<D>(g, ci.y + k0 + b, sb[b]);
#pragma unroll
            for (int b = 0; b < B; b++)
                if (!stop && k0 + b < n_own) stop = apply_summand<D>(u, sb[b], false, ci.y + k0 + b, err);
        }
    } else {
        for (int e = 0; e < deg; e++) {
            const int en = __builtin_nontemporal_load(sfi + base + (long long)e * SLICE);
            if (en == INVALID) continue;
            const int fid = en & 0x7fffffff;
            double s[GW];
            load_summand<D>(g, fid, s);
            if (apply_summand<D>(u, s, en & INCOMING, fid, err)) break;
        }
    }
    return true;
}

// OR: the plan's owner_runs flag as a template constant (the kernel holds
// only the path it takes)
template <int D, int BS = PHD_BLOCK, int XCD = 1, int OR = 1>
__global__ __launch_bounds__(BS) void k_update_sell(phd_cells c,
                                                           const double* __restrict__ g,
                                                           const int2* __restrict__ csr,
                                                           const int32_t* __restrict__ off,
                                                           const int32_t* __restrict__ perm,
                                                           const int32_t* __restrict__ sbase,
                                                           const int32_t* __restrict__ sfi,
                                                           const int2* __restrict__ cinfo,
                                                           unsigned long long* err)
{
    // At SORT_WINDOW threads a block owns one whole window (the degree sort
    // permutes within windows): the updated values pass through LDS and are
    // stored in natural cell order, whole cache lines per wave.
    constexpr int NF = D + 2;
    constexpr bool STAGE = BS == SORT_WINDOW;
    __shared__ double su[STAGE ? NF * SORT_WINDOW : 1];
    const long long n = c.n_cells;
    const long long t = sched_index<BS, XCD>(n);
    const long long w0 = t - threadIdx.x;          // the window's first cell
    const int i = t < n ? perm[t] : 0;
    double u[NF];
    const bool real = t < n && update_cell<D, OR>(c, t, i, g, csr, off, sbase, sfi, cinfo, err, u);
    if (!STAGE) {
        if (real) {
#pragma unroll
            for (int q = 0; q < NF; q++) c.cons[q][i] = u[q];
        }
        return;
    }
    if (real) {
#pragma unroll
        for (int q = 0; q < NF; q++) su[q * SORT_WINDOW + (i - w0)] = u[q];
    }
    __syncthreads();
    const long long k = w0 + threadIdx.x;
    if (k < n && c.tag[k] == 0) {
#pragma unroll
        for (int q = 0; q < NF; q++) c.cons[q][k] = su[q * SORT_WINDOW + threadIdx.x];
    }
}

// Static-mesh epilogue: DomainManager.update_ghost_fields(particles,
// conservative, True) (domain_manager.pyx:759-774) with the boundary rule
// Reflective.update_fields (boundary.pyx:336-348: momentum and velocity of
// exterior ghosts negated on each axis where they lie outside the box),
// then IdealGas.primitive_from_conservative (equation_state.pyx:104-117)
// over all cells; optionally refreshes the packed [position | primitive]
// rows.  Ghosts read only their image's conserved values, which this kernel
// does not write.
//
// BATCH (a union of problems, problem p owning cells [cs[p], cs[p+1])): also
// folds each real cell's CFL bound into dtbits[p] (atomic min on the bit
// pattern: exact) and flags bad[p] when the cell's mass, energy, density or
// pressure is not positive; a wave reduces per problem first, one atomic
// per problem it touches.
// IdealGas.primitive_from_conservative (equation_state.pyx:104-117) for one
// cell, in the reference's operation order
template <int D>
__device__ __forceinline__ void eos_prim(const double* u, double vol, double gamma, double* p)
{
    p[0] = u[0] / vol;
    double v_sq = 0.;
#pragma unroll
    for (int k = 0; k < D; k++) {
        p[1 + k] = u[1 + k] / u[0];
        v_sq += p[1 + k] * p[1 + k];
    }
    p[D + 1] = (u[D + 1] / vol - .5 * p[0] * v_sq) * (gamma - 1.);
}

// the primitive part of cell i's packed [position | primitive] row (the
// positions of a static mesh stay as written)
template <int D>
__device__ __forceinline__ void store_xp_prim(double* xp, long long i, const double* p)
{
    double* r = xp + (size_t)i * XPW;
    if (D == 2) {
        reinterpret_cast<double2*>(r)[0] = make_double2(p[0], p[1]);
        reinterpret_cast<double2*>(r)[1] = make_double2(p[2], p[3]);
    } else {
        r[3] = p[0];
        reinterpret_cast<double2*>(r)[2] = make_double2(p[1], p[2]);
        reinterpret_cast<double2*>(r)[3] = make_double2(p[3], p[D + 1]);
    }
}

// problem of cell i, searching forward from p0 (a problem at or before i's)
__device__ __forceinline__ int problem_from(int p0, int i, const int32_t* __restrict__ cs,
                                            int n_problems)
{
    int p = p0;
    while (p + 1 < n_problems && cs[p + 1] <= i) p++;
    return p;
}

// problem of cell i by binary search
__device__ __forceinline__ int problem_of(int i, const int32_t* __restrict__ cs, int n_problems)
{
    int lo = 0, hi = n_problems - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cs[mid] <= i) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// p0: the problem of the block's first cell (problems own contiguous cell
// ranges, so a lane's problem is found by stepping forward from it)
template <int D>
__device__ __forceinline__ void fold_cfl(bool real, int i, const double* u, const double* p,
                                         double vol, double gamma, int boost,
                                         const int32_t* __restrict__ cs, int n_problems,
                                         int p0, unsigned long long* dtbits, int32_t* bad)
{
    double v = __longlong_as_double((long long)INF_BITS);
    int b = 0, pid = -1;
    if (real) {
        const double cs_ = sqrt(gamma * p[D + 1] / p[0]);
        const double R = cell_radius<D>(vol);
        if (boost) {
            v = R / cs_;
        } else {
            double vsq = 0.0;
#pragma unroll
            for (int k = 0; k < D; k++) vsq += p[1 + k] * p[1 + k];
            v = R / (cs_ + sqrt(vsq));
        }
        if (!(v >= 0.0)) v = __longlong_as_double((long long)INF_BITS);
        b = !(u[0] > 0.0) || !(u[D + 1] > 0.0) || !(p[0] > 0.0) || !(p[D + 1] > 0.0);
        pid = problem_from(p0, i, cs, n_problems);
    }
    // distinct problems of the wave, smallest first
    int todo = pid;
    for (;;) {
        int q = todo < 0 ? 0x7fffffff : todo;
        for (int off = 32; off > 0; off >>= 1) q = min(q, __shfl_xor(q, off, 64));
        if (q == 0x7fffffff) break;
        double m = pid == q ? v : __longlong_as_double((long long)INF_BITS);
        int bb = pid == q ? b : 0;
        for (int off = 32; off > 0; off >>= 1) {
            m = fmin(m, __shfl_xor(m, off, 64));
            bb |= __shfl_xor(bb, off, 64);
        }
        if ((threadIdx.x & 63) == 0) {
            if (m < __longlong_as_double((long long)INF_BITS))
                atomicMin(&dtbits[q], (unsigned long long)__double_as_longlong(m));
            if (bb) atomicOr(&bad[q], 1);
        }
        if (pid == q) todo = -1;
    }
}

// conserved values of ghost i from its image, momenta negated on each axis
// where a reflective exterior ghost lies outside the box (domain_manager.pyx:
// 759-774, boundary.pyx:336-348); stored into the ghost and returned in u
template <int D>
__device__ __forceinline__ void ghost_cons(const phd_cells& c, int i, int reflect, double lo0,
                                           double lo1, double lo2, double hi0, double hi1,
                                           double hi2, double* u)
{
    constexpr int NF = D + 2;
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
}

template <int D, int BATCH = 0>
__global__ __launch_bounds__(PHD_BLOCK) void k_ghost_refresh_eos(phd_cells c, int reflect,
                                                                 double lo0, double lo1,
                                                                 double lo2, double hi0,
                                                                 double hi1, double hi2,
                                                                 double gamma, double* xp,
                                                                 const int32_t* __restrict__ bcs,
                                                                 int n_problems, int boost,
                                                                 unsigned long long* dtbits,
                                                                 int32_t* bad)
{
    constexpr int NF = D + 2;
    const long long ii = xcd_index(c.n_cells);
    const bool in = ii < c.n_cells;      // no early exit: the BATCH fold is wave-wide
    const int i = in ? (int)ii : 0;
    __shared__ int s_p0;
    if (BATCH) {
        const long long w0 = ii - threadIdx.x;     // the block's first cell
        if (threadIdx.x == 0) s_p0 = w0 < c.n_cells ? problem_of((int)w0, bcs, n_problems) : 0;
        __syncthreads();
    }
    double u[NF], p[NF], vol = 0.0;
    bool real = false;
    if (in) {
        if (c.tag[i] == 1) {
            ghost_cons<D>(c, i, reflect, lo0, lo1, lo2, hi0, hi1, hi2, u);
        } else {
            real = true;
#pragma unroll
            for (int q = 0; q < NF; q++) u[q] = c.cons[q][i];
        }
        vol = c.volume[i];
        eos_prim<D>(u, vol, gamma, p);
#pragma unroll
        for (int q = 0; q < NF; q++) c.prim[q][i] = p[q];
        if (xp) {
            double r[XPW] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int k = 0; k < D; k++) r[XpLay<D>::X + k] = c.x[k][i];
#pragma unroll
            for (int n = 0; n < NF; n++) r[XpLay<D>::P + n] = p[n];
            double2* o = reinterpret_cast<double2*>(xp + (size_t)i * XPW);
#pragma unroll
            for (int q = 0; q < XpLay<D>::REFRESH_PAIRS; q++) o[q] = make_double2(r[2 * q], r[2 * q + 1]);
        }
    }
    if (BATCH) fold_cfl<D>(real, i, u, p, vol, gamma, boost, bcs, n_problems, s_p0, dtbits, bad);
}

// BATCH conserved update fused with the epilogue of its real cells: the
// update of k_update_sell, then the primitives (equation_state.pyx:104-117),
// the primitive part of the packed row and the per-problem CFL fold, as
// k_ghost_refresh_eos<D, 1> does for real cells, with the same expressions.
// A block of SORT_WINDOW threads owns a whole window (the degree sort
// permutes within windows): the updated conserved values pass through LDS
// and every field is then written in natural cell order, whole cache lines
// per wave.  Ghosts are left to k_batch_ghosts, launched after this kernel
// (their images are the updated real cells).  PRIM = 0 leaves the real
// cells' primitive arrays alone: no kernel of a batched step reads them (the
// gradients read the packed rows), so replayed steps defer them to one
// k_batch_prim after the last step.
template <int D, int BS = PHD_BLOCK, int XCD = 1, int PRIM = 1, int OR = 1>
__global__ __launch_bounds__(BS) void k_update_eos(phd_cells c, const double* __restrict__ g,
                                                   const int2* __restrict__ csr,
                                                   const int32_t* __restrict__ off,
                                                   const int32_t* __restrict__ perm,
                                                   const int32_t* __restrict__ sbase,
                                                   const int32_t* __restrict__ sfi,
                                                   const int2* __restrict__ cinfo,
                                                   unsigned long long* err, double gamma,
                                                   double* __restrict__ xp,
                                                   const int32_t* __restrict__ bcs,
                                                   int n_problems, int boost,
                                                   unsigned long long* dtbits, int32_t* bad)
{
    constexpr int NF = D + 2;
    constexpr bool STAGE = BS == SORT_WINDOW;
    __shared__ double su[STAGE ? NF * SORT_WINDOW : 1];
    __shared__ int s_p0;
    const long long n = c.n_cells;
    const long long t = sched_index<BS, XCD>(n);
    const long long w0 = t - threadIdx.x;          // the block's first position
    const long long ws = w0 - w0 % SORT_WINDOW;    // its window's first cell
    if (threadIdx.x == 0) s_p0 = ws < n ? problem_of((int)ws, bcs, n_problems) : 0;
    double u[NF], p[NF];
    int i = t < n ? perm[t] : 0;
    bool real = t < n && update_cell<D, OR>(c, t, i, g, csr, off, sbase, sfi, cinfo, err, u);
    if (STAGE) {
        if (real) {
#pragma unroll
            for (int q = 0; q < NF; q++) su[q * SORT_WINDOW + (i - w0)] = u[q];
        }
        __syncthreads();
        const long long k = w0 + threadIdx.x;
        i = k < n ? (int)k : 0;
        real = k < n && c.tag[i] == 0;
        if (real) {
#pragma unroll
            for (int q = 0; q < NF; q++) u[q] = su[q * SORT_WINDOW + threadIdx.x];
        }
    } else {
        __syncthreads();                           // s_p0
    }
    double vol = 0.0;
    if (real) {
        vol = c.volume[i];
        eos_prim<D>(u, vol, gamma, p);
#pragma unroll
        for (int q = 0; q < NF; q++) {
            c.cons[q][i] = u[q];
            if (PRIM) c.prim[q][i] = p[q];
        }
        store_xp_prim<D>(xp, i, p);
    }
    fold_cfl<D>(real, i, u, p, vol, gamma, boost, bcs, n_problems, s_p0, dtbits, bad);
}

// BATCH ghosts after k_update_eos: conserved values from their images
// (ghost_cons), primitives and the primitive part of the packed row; gl
// lists the ghost cells
template <int D>
__global__ __launch_bounds__(PHD_BLOCK) void k_batch_ghosts(phd_cells c,
                                                            const int32_t* __restrict__ gl,
                                                            int n_ghosts, int reflect,
                                                            double lo0, double lo1, double lo2,
                                                            double hi0, double hi1, double hi2,
                                                            double gamma, double* __restrict__ xp)
{
    constexpr int NF = D + 2;
    const long long k = xcd_index(n_ghosts);
    if (k >= n_ghosts) return;
    const int i = gl[k];
    double u[NF], p[NF];
    ghost_cons<D>(c, i, reflect, lo0, lo1, lo2, hi0, hi1, hi2, u);
    eos_prim<D>(u, c.volume[i], gamma, p);
#pragma unroll
    for (int q = 0; q < NF; q++) c.prim[q][i] = p[q];
    store_xp_prim<D>(xp, i, p);
}

// primitive arrays from the packed rows' primitive part (the same values the
// step's EOS wrote there), after steps whose update deferred them
template <int D>
__global__ __launch_bounds__(PHD_BLOCK) void k_batch_prim(phd_cells c, const double* __restrict__ xp)
{
    const long long i = xcd_index(c.n_cells);
    if (i >= c.n_cells) return;
    const double* r = xp + (size_t)i * XPW + XpLay<D>::P;
#pragma unroll
    for (int q = 0; q < D + 2; q++) c.prim[q][i] = r[q];
}

// ---------------------------------------------------------------------------
// host helpers
#define HIPCHK(x) do { if ((x) != hipSuccess) return PHD_E_HIP; } while (0)
#define LAUNCHCHK() do { if (hipPeekAtLastError() != hipSuccess) { (void)hipGetLastError(); return PHD_E_HIP; } } while (0)

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

int check_plan(const phd_plan* p, const phd_cells* c)
{
    if (!p || !c || p->dim != c->dim || p->n_cells != c->n_cells) return PHD_E_ARG;
    return 0;
}

int check_plan_faces(const phd_plan* p, const phd_faces* f)
{
    if (!p || !f || p->n_faces != f->n_faces) return PHD_E_ARG;
    return 0;
}

// cell-record rows: ld >= the layout's width, even, 16-byte aligned; the
// array may be null only when there are no cells
int check_rec(const phd_cells* c, const double* rec, int ld)
{
    int need = c->dim == 3 ? Lay<3>::LD : Lay<2>::LD;
    if ((!rec && c->n_cells > 0) || ld < need || (ld & 1) || (((uintptr_t)rec) & 15))
        return PHD_E_ARG;
    return 0;
}

FaceArrays fa(double* const* p, int nf)
{
    FaceArrays a;
    for (int q = 0; q < 5; q++) a.q[q] = (p && q < nf) ? p[q] : nullptr;
    return a;
}

void bounds3(int dim, int boundary, const double* lo, const double* hi, double* l, double* u)
{
    for (int k = 0; k < 3; k++) { l[k] = 0; u[k] = 0; }
    if (boundary == PHD_BOUNDARY_REFLECTIVE && lo && hi)
        for (int k = 0; k < dim; k++) { l[k] = lo[k]; u[k] = hi[k]; }
}



template <int D>
int launch_pack(phd_handle* h, phd_plan* p, const phd_cells* c)
{
    if (c->n_cells == 0) return 0;
    k_pack_xp<D><<<xcd_grid(c->n_cells), PHD_BLOCK, 0, h->stream>>>(*c, p->xp);
    LAUNCHCHK();
    return 0;
}

template <int D>
int launch_grad(phd_handle* h, phd_plan* p, const phd_cells* c, int limiter, double* rec, int ld)
{
    if (c->n_cells == 0) return 0;
    if (p->bad_cell >= 0) {
        k_set_err<<<1, 64, 0, h->stream>>>(h->d_err, 1, (unsigned long long)p->bad_cell);
        LAUNCHCHK();
    }
    const unsigned nb = xcd_grid(c->n_cells);
#define GS(L) k_grad_sell<D, L><<<nb, PHD_BLOCK, 0, h->stream>>>(c->n_cells, c->tag, p->cell_geo, c->volume, p->xp, p->off, p->perm, p->sbase, p->sj, p->sgeo[0], p->sgeo[1], p->sgeo[2], p->sgeo[3], p->hcnt, p->halo, p->sloc, rec, ld)
    if (limiter == PHD_LIMITER_AREPO) GS(PHD_LIMITER_AREPO);
    else if (limiter == PHD_LIMITER_TESS) GS(PHD_LIMITER_TESS);
    else return PHD_E_MODE;
#undef GS
    LAUNCHCHK();
    return 0;
}

template <int D>
int launch_ghost(phd_handle* h, phd_plan* p, const phd_cells* c, int boundary, const double* lo,
                 const double* hi, double* rec, int ld)
{
    if (c->n_cells == 0) return 0;
    if (boundary != PHD_BOUNDARY_REFLECTIVE && boundary != PHD_BOUNDARY_PERIODIC) return PHD_E_MODE;
    if (boundary == PHD_BOUNDARY_REFLECTIVE && (!lo || !hi)) return PHD_E_ARG;
    if (p->n_ghosts == 0) return 0;
    if (!c->map) return PHD_E_ARG;          // ghosts need their images
    double l[3], u[3];
    bounds3(D, boundary, lo, hi, l, u);
    k_ghost_rec<D><<<xcd_grid(p->n_ghosts), PHD_BLOCK, 0, h->stream>>>(
        *c, p->ghosts, p->n_ghosts, p->cell_geo, boundary == PHD_BOUNDARY_REFLECTIVE, l[0], l[1],
        l[2], u[0], u[1], u[2], rec, ld);
    LAUNCHCHK();
    return 0;
}

template <int D, int ORDER, int OUT>
int launch_face_o(phd_handle* h, phd_plan* p, const phd_cells* c, const phd_faces* f,
                  const double* rec, int ld, double gamma, double dt_half, double dt_full,
                  int boost, int gizmo, int solver, FaceArrays flux)
{
    const unsigned nb = xcd_grid(f->n_faces);
    // the gizmo switch only acts in the linear reconstruction
    const int bg = (boost ? 1 : 0) | (ORDER == PHD_ORDER_LINEAR && gizmo ? 2 : 0);
#define FACE(S, B) k_face<D, ORDER, S, OUT, 0, PHD_BLOCK, 1, B><<<nb, PHD_BLOCK, 0, h->stream>>>(*c, *f, p->face_rec, rec, ld, make_gas(gamma), dt_half, dt_full, boost, gizmo, flux, p->g, h->d_err, nullptr, nullptr)
#define FACE_BG(S)                                                           \
    do {                                                                     \
        if (bg == 0) FACE(S, 0);                                             \
        else if (bg == 1) FACE(S, 1);                                        \
        else if constexpr (ORDER == PHD_ORDER_LINEAR) {                      \
            if (bg == 2) FACE(S, 2);                                         \
            else FACE(S, 3);                                                 \
        }                                                                    \
    } while (0)
    switch (solver) {
    case PHD_SOLVER_HLL: FACE_BG(PHD_SOLVER_HLL); break;
    case PHD_SOLVER_HLLC: FACE_BG(PHD_SOLVER_HLLC); break;
    case PHD_SOLVER_EXACT: FACE_BG(PHD_SOLVER_EXACT); break;
    default: return PHD_E_MODE;
    }
#undef FACE_BG
#undef FACE
    LAUNCHCHK();
    return 0;
}

template <int D, int OUT>
int launch_face(phd_handle* h, phd_plan* p, const phd_cells* c, const phd_faces* f,
                const double* rec, int ld, double gamma, double dt_half, double dt_full,
                int boost, int gizmo, int order, int solver, FaceArrays flux)
{
    if (f->n_faces == 0) return 0;
    if (order == PHD_ORDER_LINEAR)
        return launch_face_o<D, PHD_ORDER_LINEAR, OUT>(h, p, c, f, rec, ld, gamma, dt_half, dt_full, boost, gizmo, solver, flux);
    // constant order only through the flux entry point (the fused step
    // reconstructs linearly), so only its output mode is compiled
    if constexpr (OUT == 1) {
        if (order == PHD_ORDER_CONSTANT)
            return launch_face_o<D, PHD_ORDER_CONSTANT, OUT>(h, p, c, f, rec, ld, gamma, dt_half, dt_full, boost, gizmo, solver, flux);
    }
    return PHD_E_MODE;
}

template <int D, int SRC>
int launch_update(phd_handle* h, phd_plan* p, const phd_cells* c, const phd_faces* f,
                  FaceArrays flux, double dt)
{
    if (c->n_cells == 0) return 0;
    if (SRC == 1 && p->owner_runs)
        k_update_sell<D, PHD_BLOCK, 1, 1><<<xcd_grid(c->n_cells), PHD_BLOCK, 0, h->stream>>>(
            *c, p->g, p->csr, p->off, p->perm, p->sbase, p->sfi, p->cinfo, h->d_err);
    else if (SRC == 1)
        k_update_sell<D, PHD_BLOCK, 1, 0><<<xcd_grid(c->n_cells), PHD_BLOCK, 0, h->stream>>>(
            *c, p->g, p->csr, p->off, p->perm, p->sbase, p->sfi, p->cinfo, h->d_err);
    else
        k_update<D><<<xcd_grid(c->n_cells), PHD_BLOCK, 0, h->stream>>>(
            *c, f ? f->area : nullptr, flux, p->csr, p->off, dt, h->d_err);
    LAUNCHCHK();
    return 0;
}

// sliced face lists of a plan: widths on the device, bases by a host prefix
// sum (plan creation is the synchronising, allocating call), then the slots
// the ghost cells (tag 1) in ascending order, for the kernels that touch
// ghosts only
int build_ghost_list(phd_handle* h, phd_plan* p, const phd_cells* c)
{
    const long long na = c->n_cells;
    if (na == 0) return 0;
    std::vector<int8_t> tag((size_t)na);
    HIPCHK(hipMemcpyAsync(tag.data(), c->tag, (size_t)na, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    std::vector<int32_t> gl;
    for (long long i = 0; i < na; i++)
        if (tag[i] == 1) gl.push_back((int32_t)i);
    p->n_ghosts = (int)gl.size();
    if (gl.empty()) return 0;
    HIPCHK(hipMalloc(&p->ghosts, 4 * gl.size()));
    p->bytes += 4 * gl.size();
    HIPCHK(hipMemcpyAsync(p->ghosts, gl.data(), 4 * gl.size(), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

int build_slots(phd_handle* h, phd_plan* p, const phd_cells* c, const phd_faces* f)
{
    const long long na = c->n_cells, nf = f->n_faces;
    const int D = c->dim, NF = D + 2;
    const long long ns = (na + SLICE - 1) / SLICE;
    int32_t* d_w = nullptr;
    HIPCHK(hipMalloc(&d_w, sizeof(int32_t) * (size_t)(ns + 1)));
    int32_t* hw = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ns + 1));
    if (!hw) { (void)hipFree(d_w); return PHD_E_ARG; }
    int rc = 0;
    long long total = 0;
    do {
        if (na > 0) k_plan_perm<<<(unsigned)((na + SORT_WINDOW - 1) / SORT_WINDOW), SORT_WINDOW, 0, h->stream>>>(*c, p->off, p->perm);
        if (ns > 0) k_plan_slice_width<<<nblocks(ns), PHD_BLOCK, 0, h->stream>>>(*c, p->off, p->perm, d_w, (int)ns);
        if (hipPeekAtLastError() != hipSuccess) { (void)hipGetLastError(); rc = PHD_E_HIP; break; }
        if (hipMemcpyAsync(hw, d_w, sizeof(int32_t) * (size_t)ns, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
            hipStreamSynchronize(h->stream) != hipSuccess) { rc = PHD_E_HIP; break; }
        for (long long s = 0; s < ns; s++) {
            long long w = hw[s];
            hw[s] = (int32_t)total;
            total += w * SLICE;
        }
        hw[ns] = (int32_t)total;
        if (total >= INT32_MAX) { rc = PHD_E_ARG; break; }
    } while (0);
    (void)hipFree(d_w);
    if (rc) { free(hw); return rc; }
    p->n_slots = total;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t s_base = al(sizeof(int32_t) * (size_t)(ns + 1));
    const size_t s_j = al(sizeof(int32_t) * (size_t)total);
    const size_t s_d = al(sizeof(double) * (size_t)total);
    const size_t s_g = al(sizeof(double) * (size_t)NF * (size_t)nf);   // summand_width
    const size_t s_ci = al(sizeof(int2) * (size_t)na);
    const size_t bytes = s_base + 2 * s_j + 4 * s_d + s_g + s_ci;
    if (hipMalloc(&p->slot_base, bytes) != hipSuccess) { free(hw); p->slot_base = nullptr; return PHD_E_HIP; }
    p->bytes += bytes;
    char* b = (char*)p->slot_base;
    p->sbase = (int32_t*)b; b += s_base;
    p->sj = (int32_t*)b; b += s_j;
    p->sfi = (int32_t*)b; b += s_j;
    for (int q = 0; q < 4; q++) { p->sgeo[q] = (double*)b; b += s_d; }
    p->g = (double*)b; b += s_g;
    p->cinfo = (int2*)b;
    do {
        if (hipMemcpyAsync(p->sbase, hw, sizeof(int32_t) * (size_t)(ns + 1), hipMemcpyHostToDevice, h->stream) != hipSuccess) { rc = PHD_E_HIP; break; }
        if (total > 0) k_fill_i32<<<nblocks(total), PHD_BLOCK, 0, h->stream>>>(p->sj, total, INVALID);
        if (total > 0) k_fill_i32<<<nblocks(total), PHD_BLOCK, 0, h->stream>>>(p->sfi, total, INVALID);
        if (na > 0) {
            if (D == 2) k_plan_slots<2><<<xcd_grid(na), PHD_BLOCK, 0, h->stream>>>(*c, *f, p->csr, p->off, p->perm, p->sbase, p->sj, p->sgeo[0], p->sgeo[1], p->sgeo[2], p->sgeo[3], p->sfi, p->cinfo);
            else k_plan_slots<3><<<xcd_grid(na), PHD_BLOCK, 0, h->stream>>>(*c, *f, p->csr, p->off, p->perm, p->sbase, p->sj, p->sgeo[0], p->sgeo[1], p->sgeo[2], p->sgeo[3], p->sfi, p->cinfo);
        }
        if (hipPeekAtLastError() != hipSuccess) { (void)hipGetLastError(); rc = PHD_E_HIP; break; }
        if (hipStreamSynchronize(h->stream) != hipSuccess) rc = PHD_E_HIP;
    } while (0);
    free(hw);
    return rc;
}

// Per window of SORT_WINDOW cells, the halo the 3D gradient kernel stages
// in LDS beside the window's own rows (at most halo_max rows, fixed stride
// per window), and each slot's row in that staging: k_plan_halo, one block
// per window, after the sliced lists are built.
int build_halo(phd_handle* h, phd_plan* p, const phd_cells* c)
{
    const long long na = c->n_cells, nw = (na + SORT_WINDOW - 1) / SORT_WINDOW;
    const long long nsl = p->n_slots;
    const int HM = c->dim == 2 ? halo_max<2>() : halo_max<3>();
    p->hcnt = p->halo = nullptr;
    p->sloc = nullptr;
    if (HM == 0) return 0;            // no halo staging in this dimension
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t s_hc = al(4 * (size_t)nw + 4), s_ha = al(4 * (size_t)nw * HM + 4);
    const size_t s_sl = al(2 * (size_t)nsl + 2);
    if (hipMalloc(&p->halo_base, s_hc + s_ha + s_sl) != hipSuccess) { p->halo_base = nullptr; return PHD_E_HIP; }
    p->bytes += s_hc + s_ha + s_sl;
    char* b = (char*)p->halo_base;
    p->hcnt = (int32_t*)b; b += s_hc;
    p->halo = (int32_t*)b; b += s_ha;
    p->sloc = (uint16_t*)b;
    if (nsl > 0) HIPCHK(hipMemsetAsync(p->sloc, 0xFF, 2 * (size_t)nsl, h->stream));   // LOC_NONE
    if (nw > 0) {
        k_plan_halo<<<(unsigned)nw, SORT_WINDOW, 0, h->stream>>>(na, c->tag, p->off, p->perm, p->sbase,
                                                                 p->sj, HM, p->hcnt, p->halo, p->sloc);
        LAUNCHCHK();
    }
    return 0;
}

} // namespace

// ===========================================================================
// C ABI
extern "C" {

int phd_abi_version(void) { return PHD_ABI_VERSION; }

int phd_cell_record_ld(int dim)
{
    return dim == 3 ? Lay<3>::LD : (dim == 2 ? Lay<2>::LD : PHD_E_ARG);
}

const char* phd_error_string(int code)
{
    switch (code) {
    case 0: return "ok";
    case PHD_E_ARG: return "invalid argument";
    case PHD_E_HIP: return "HIP runtime error";
    case PHD_E_MODE: return "unsupported mode";
    case PHD_E_SIZE: return "input exceeds the fixed scratch size of an allocation-free entry point";
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
    if (hipMalloc(&h->d_err, PHD_ERR_SLOTS * sizeof(unsigned long long)) != hipSuccess) { delete h; return PHD_E_HIP; }
    if (hipHostMalloc(&h->h_err, PHD_ERR_SLOTS * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess) {
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
    if (mode != PHD_MODE_BITWISE) return PHD_E_MODE;
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

// ---- plan ----------------------------------------------------------------
int phd_plan_create(phd_handle* h, const phd_cells* c, const phd_faces* f, phd_plan** out)
{
    if (!h || !out || check_cells(c) || check_faces(f) || !f->nbr_off) return PHD_E_ARG;
    *out = nullptr;
    const int D = c->dim, NF = D + 2;
    const long long na = c->n_cells, nf = f->n_faces;
    int32_t n_entries = 0;
    if (na > 0) {
        HIPCHK(hipMemcpyAsync(&n_entries, f->nbr_off + na, sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    if (n_entries < 0 || (n_entries > 0 && !f->nbr_face)) return PHD_E_ARG;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t s_face = al(sizeof(double) * FREC * (size_t)nf);
    const size_t s_geo = al(sizeof(double) * GEO * (size_t)na);
    const size_t s_csr = al(sizeof(int2) * (size_t)n_entries);
    const size_t s_off = al(sizeof(int32_t) * (size_t)(na + 1));
    const size_t s_xp = al(sizeof(double) * XPW * (size_t)na);
    const size_t s_perm = al(sizeof(int32_t) * (size_t)na);
    const size_t s_bad = al(4 * sizeof(unsigned long long));
    const size_t total = s_face + s_geo + s_csr + s_off + s_xp + s_perm + s_bad;
    phd_plan* p = new phd_plan();
    p->dim = D;
    p->n_cells = (int)na;
    p->n_faces = (int)nf;
    p->n_entries = n_entries;
    p->device = h->device;
    p->bytes = total;
    p->bad_cell = -1;
    if (hipMalloc(&p->base, total) != hipSuccess) { delete p; return PHD_E_HIP; }
    char* b = (char*)p->base;
    p->face_rec = (double*)b; b += s_face;
    p->cell_geo = (double*)b; b += s_geo;
    p->csr = (int2*)b; b += s_csr;
    p->off = (int32_t*)b; b += s_off;
    p->xp = (double*)b; b += s_xp;
    p->perm = (int32_t*)b; b += s_perm;
    p->d_bad = (unsigned long long*)b;
    p->slot_base = nullptr;
    p->ghosts = nullptr;
    p->n_ghosts = 0;
    p->halo_base = nullptr;
    int rc = 0;
    do {
        if (hipMemcpyAsync(p->off, f->nbr_off, sizeof(int32_t) * (size_t)(na + 1), hipMemcpyDeviceToDevice, h->stream) != hipSuccess) { rc = PHD_E_HIP; break; }
        k_reset_err<<<1, 64, 0, h->stream>>>(p->d_bad);   // slot 0 = NOERR
        if (nf > 0) {
            if (D == 2) k_plan_faces<2><<<xcd_grid(nf), PHD_BLOCK, 0, h->stream>>>(*c, *f, p->face_rec);
            else k_plan_faces<3><<<xcd_grid(nf), PHD_BLOCK, 0, h->stream>>>(*c, *f, p->face_rec);
        }
        if (na > 0) {
            if (D == 2) k_plan_cells<2><<<xcd_grid(na), PHD_BLOCK, 0, h->stream>>>(*c, p->cell_geo, p->xp);
            else k_plan_cells<3><<<xcd_grid(na), PHD_BLOCK, 0, h->stream>>>(*c, p->cell_geo, p->xp);
            k_plan_csr<<<xcd_grid(na), PHD_BLOCK, 0, h->stream>>>(*c, *f, p->csr, p->d_bad);
            k_plan_runs<<<xcd_grid(na), PHD_BLOCK, 0, h->stream>>>(*c, p->csr, p->off, p->d_bad + 1);
        }
        if (hipPeekAtLastError() != hipSuccess) { (void)hipGetLastError(); rc = PHD_E_HIP; break; }
        unsigned long long bad[2] = {NOERR, NOERR};
        if (hipMemcpyAsync(bad, p->d_bad, sizeof(bad), hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
            hipStreamSynchronize(h->stream) != hipSuccess) { rc = PHD_E_HIP; break; }
        p->bad_cell = bad[0] == NOERR ? -1 : (long long)bad[0];
        p->owner_runs = bad[1] == NOERR;
        rc = build_slots(h, p, c, f);
        if (!rc) rc = build_halo(h, p, c);
        if (!rc) rc = build_ghost_list(h, p, c);
    } while (0);
    if (rc) {
        if (p->ghosts) (void)hipFree(p->ghosts);
        if (p->halo_base) (void)hipFree(p->halo_base);
        if (p->slot_base) (void)hipFree(p->slot_base);
        (void)hipFree(p->base);
        delete p;
        return rc;
    }
    *out = p;
    return 0;
}

int phd_plan_destroy(phd_plan* p)
{
    if (!p) return PHD_E_ARG;
    (void)hipDeviceSynchronize();
    if (p->ghosts) (void)hipFree(p->ghosts);
    if (p->halo_base) (void)hipFree(p->halo_base);
    if (p->slot_base) (void)hipFree(p->slot_base);
    (void)hipFree(p->base);
    delete p;
    return 0;
}

int64_t phd_plan_bytes(const phd_plan* p) { return p ? (int64_t)p->bytes : -1; }

// ---- equation of state ----------------------------------------------------
int phd_conservative_from_primitive(phd_handle* h, const phd_cells* c, double gamma)
{
    if (!h || check_cells(c)) return PHD_E_ARG;
    if (c->n_cells == 0) return 0;
    if (c->dim == 2) k_cons_from_prim<2><<<xcd_grid(c->n_cells), PHD_BLOCK, 0, h->stream>>>(*c, gamma);
    else k_cons_from_prim<3><<<xcd_grid(c->n_cells), PHD_BLOCK, 0, h->stream>>>(*c, gamma);
    LAUNCHCHK();
    return 0;
}

int phd_primitive_from_conservative(phd_handle* h, const phd_cells* c, double gamma)
{
    if (!h || check_cells(c)) return PHD_E_ARG;
    if (c->n_cells == 0) return 0;
    if (c->dim == 2) k_prim_from_cons<2><<<xcd_grid(c->n_cells), PHD_BLOCK, 0, h->stream>>>(*c, gamma);
    else k_prim_from_cons<3><<<xcd_grid(c->n_cells), PHD_BLOCK, 0, h->stream>>>(*c, gamma);
    LAUNCHCHK();
    return 0;
}

int phd_compute_time_step(phd_handle* h, const phd_cells* c, double gamma, double cfl, int boost,
                          double* dt_out)
{
    if (!h || check_cells(c) || !dt_out || c->n_cells == 0) return PHD_E_ARG;
    k_reset_err<<<1, 64, 0, h->stream>>>(h->d_err);   // also clears the fault words
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
    if (c->dim == 2) k_gen_vel<2><<<xcd_grid(c->n_cells), PHD_BLOCK, 0, h->stream>>>(*c, gamma, eta, regularize);
    else k_gen_vel<3><<<xcd_grid(c->n_cells), PHD_BLOCK, 0, h->stream>>>(*c, gamma, eta, regularize);
    LAUNCHCHK();
    return 0;
}

int phd_assign_face_velocities(phd_handle* h, const phd_cells* c, const phd_faces* f)
{
    if (!h || check_cells(c) || check_faces(f)) return PHD_E_ARG;
    if (f->n_faces == 0) return 0;
    if (c->dim == 2) k_face_vel<2><<<xcd_grid(f->n_faces), PHD_BLOCK, 0, h->stream>>>(*c, *f);
    else k_face_vel<3><<<xcd_grid(f->n_faces), PHD_BLOCK, 0, h->stream>>>(*c, *f);
    LAUNCHCHK();
    return 0;
}

// ---- reconstruction -------------------------------------------------------
int phd_pack_cells(phd_handle* h, phd_plan* p, const phd_cells* c)
{
    if (!h || check_cells(c) || check_plan(p, c)) return PHD_E_ARG;
    return c->dim == 2 ? launch_pack<2>(h, p, c) : launch_pack<3>(h, p, c);
}

int phd_compute_gradients(phd_handle* h, phd_plan* p, const phd_cells* c, int limiter,
                          double* rec, int32_t rec_ld)
{
    if (!h || check_cells(c) || check_plan(p, c) || check_rec(c, rec, rec_ld)) return PHD_E_ARG;
    int rc = phd_pack_cells(h, p, c);
    if (rc) return rc;
    return c->dim == 2 ? launch_grad<2>(h, p, c, limiter, rec, rec_ld)
                       : launch_grad<3>(h, p, c, limiter, rec, rec_ld);
}

int phd_update_ghost_gradients(phd_handle* h, phd_plan* p, const phd_cells* c, int boundary,
                               const double* lo, const double* hi, double* rec, int32_t rec_ld)
{
    if (!h || check_cells(c) || check_plan(p, c) || check_rec(c, rec, rec_ld)) return PHD_E_ARG;
    return c->dim == 2 ? launch_ghost<2>(h, p, c, boundary, lo, hi, rec, rec_ld)
                       : launch_ghost<3>(h, p, c, boundary, lo, hi, rec, rec_ld);
}

int phd_compute_states(phd_handle* h, phd_plan* p, const phd_cells* c, const phd_faces* f,
                       const double* rec, int32_t rec_ld, double gamma, double dt, int boost,
                       int gizmo, int order, int stages, double* const* left,
                       double* const* right)
{
    if (!h || check_cells(c) || check_faces(f) || check_plan(p, c) || check_plan_faces(p, f) ||
        !left || !right)
        return PHD_E_ARG;
    if (order == PHD_ORDER_LINEAR && check_rec(c, rec, rec_ld)) return PHD_E_ARG;
    if (order != PHD_ORDER_LINEAR && order != PHD_ORDER_CONSTANT) return PHD_E_MODE;
    if (f->n_faces == 0) return 0;
    const int nf = c->dim + 2;
    FaceArrays L = fa(left, nf), R = fa(right, nf);
    const unsigned nb = xcd_grid(f->n_faces);
#define ST(DD, OO) k_states<DD, OO><<<nb, PHD_BLOCK, 0, h->stream>>>(*c, *f, p->face_rec, rec, rec_ld, gamma, dt, boost, gizmo, stages, L, R)
    if (c->dim == 2) {
        if (order == PHD_ORDER_LINEAR) ST(2, PHD_ORDER_LINEAR); else ST(2, PHD_ORDER_CONSTANT);
    } else {
        if (order == PHD_ORDER_LINEAR) ST(3, PHD_ORDER_LINEAR); else ST(3, PHD_ORDER_CONSTANT);
    }
#undef ST
    LAUNCHCHK();
    return 0;
}

// ---- Riemann ---------------------------------------------------------------
int phd_compute_fluxes(phd_handle* h, const phd_faces* f, int dim, double* const* left,
                       double* const* right, double gamma, int boost, int solver,
                       double* const* flux)
{
    if (!h || check_faces(f) || (dim != 2 && dim != 3) || !left || !right || !flux) return PHD_E_ARG;
    if (f->n_faces == 0) return 0;
    const int nf = dim + 2;
    FaceArrays L = fa(left, nf), R = fa(right, nf), F = fa(flux, nf);
    const unsigned nb = xcd_grid(f->n_faces);
#define RIE(DD, SS) k_riemann<DD, SS><<<nb, PHD_BLOCK, 0, h->stream>>>(*f, L, R, make_gas(gamma), boost, F, h->d_err)
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

int phd_face_flux(phd_handle* h, phd_plan* p, const phd_cells* c, const phd_faces* f,
                  const double* rec, int32_t rec_ld, double gamma, double dt, int boost, int gizmo,
                  int order, int solver, double* const* flux)
{
    if (!h || check_cells(c) || check_faces(f) || check_plan(p, c) || check_plan_faces(p, f) ||
        !flux)
        return PHD_E_ARG;
    if (order == PHD_ORDER_LINEAR && check_rec(c, rec, rec_ld)) return PHD_E_ARG;
    FaceArrays F = fa(flux, c->dim + 2);
    return c->dim == 2
               ? launch_face<2, 1>(h, p, c, f, rec, rec_ld, gamma, dt, 0.0, boost, gizmo, order, solver, F)
               : launch_face<3, 1>(h, p, c, f, rec, rec_ld, gamma, dt, 0.0, boost, gizmo, order, solver, F);
}

// ---- update ------------------------------------------------------------------
int phd_update_from_fluxes(phd_handle* h, phd_plan* p, const phd_cells* c, const phd_faces* f,
                           double* const* flux, double dt)
{
    if (!h || check_cells(c) || check_faces(f) || check_plan(p, c) || check_plan_faces(p, f) ||
        !flux)
        return PHD_E_ARG;
    FaceArrays F = fa(flux, c->dim + 2);
    return c->dim == 2 ? launch_update<2, 0>(h, p, c, f, F, dt)
                       : launch_update<3, 0>(h, p, c, f, F, dt);
}

// ---- core-step stages -------------------------------------------------------
int phd_gradients_packed(phd_handle* h, phd_plan* p, const phd_cells* c, int limiter,
                         double* rec, int32_t rec_ld)
{
    if (!h || check_cells(c) || check_plan(p, c) || check_rec(c, rec, rec_ld)) return PHD_E_ARG;
    return c->dim == 2 ? launch_grad<2>(h, p, c, limiter, rec, rec_ld)
                       : launch_grad<3>(h, p, c, limiter, rec, rec_ld);
}

int phd_face_summands(phd_handle* h, phd_plan* p, const phd_cells* c, const phd_faces* f,
                      const double* rec, int32_t rec_ld, double gamma, double dt, int boost,
                      int gizmo, int solver, double* const* flux)
{
    if (!h || check_cells(c) || check_faces(f) || check_plan(p, c) || check_plan_faces(p, f) ||
        check_rec(c, rec, rec_ld))
        return PHD_E_ARG;
    const int D = c->dim;
    FaceArrays F = fa(flux, D + 2);
    if (flux)
        return D == 2 ? launch_face<2, 3>(h, p, c, f, rec, rec_ld, gamma, 0.5 * dt, dt, boost, gizmo, PHD_ORDER_LINEAR, solver, F)
                      : launch_face<3, 3>(h, p, c, f, rec, rec_ld, gamma, 0.5 * dt, dt, boost, gizmo, PHD_ORDER_LINEAR, solver, F);
    return D == 2 ? launch_face<2, 2>(h, p, c, f, rec, rec_ld, gamma, 0.5 * dt, dt, boost, gizmo, PHD_ORDER_LINEAR, solver, F)
                  : launch_face<3, 2>(h, p, c, f, rec, rec_ld, gamma, 0.5 * dt, dt, boost, gizmo, PHD_ORDER_LINEAR, solver, F);
}

int phd_update_from_summands(phd_handle* h, phd_plan* p, const phd_cells* c)
{
    if (!h || check_cells(c) || check_plan(p, c)) return PHD_E_ARG;
    FaceArrays F = fa(nullptr, 0);
    return c->dim == 2 ? launch_update<2, 1>(h, p, c, nullptr, F, 0.0)
                       : launch_update<3, 1>(h, p, c, nullptr, F, 0.0);
}

// ---- fused core step -----------------------------------------------------------
int phd_core_step(phd_handle* h, phd_plan* p, const phd_cells* c, const phd_faces* f,
                  double* rec, int32_t rec_ld, int limiter, int boundary, const double* lo,
                  const double* hi, double gamma, double dt, int boost, int gizmo, int solver,
                  double* const* flux, int flags)
{
    if (!h || check_cells(c) || check_faces(f) || check_plan(p, c) || check_plan_faces(p, f) ||
        check_rec(c, rec, rec_ld))
        return PHD_E_ARG;
    int rc = 0;
    if (!(flags & PHD_CORE_XP_FRESH)) rc = phd_pack_cells(h, p, c);
    if (!rc) rc = phd_gradients_packed(h, p, c, limiter, rec, rec_ld);
    if (!rc) rc = phd_update_ghost_gradients(h, p, c, boundary, lo, hi, rec, rec_ld);
    if (!rc) rc = phd_face_summands(h, p, c, f, rec, rec_ld, gamma, dt, boost, gizmo, solver, flux);
    if (!rc) rc = phd_update_from_summands(h, p, c);
    return rc;
}

int phd_ghost_refresh_eos(phd_handle* h, phd_plan* p, const phd_cells* c, int boundary,
                          const double* lo, const double* hi, double gamma)
{
    if (!h || check_cells(c)) return PHD_E_ARG;
    if (p && check_plan(p, c)) return PHD_E_ARG;
    if (c->n_cells == 0) return 0;
    if (boundary != PHD_BOUNDARY_REFLECTIVE && boundary != PHD_BOUNDARY_PERIODIC) return PHD_E_MODE;
    if (boundary == PHD_BOUNDARY_REFLECTIVE && (!lo || !hi)) return PHD_E_ARG;
    double l[3], u[3];
    bounds3(c->dim, boundary, lo, hi, l, u);
    const int refl = boundary == PHD_BOUNDARY_REFLECTIVE;
    double* xp = p ? p->xp : nullptr;
    if (!c->map) {
        // no images: valid only without ghosts (the plan lists them; without
        // a plan the tags are scanned, synchronously)
        if (p && p->n_ghosts > 0) return PHD_E_ARG;
        if (!p) {
            unsigned long long* slot = h->d_err + PHD_ERR_SLOT_GHOST;
            k_fill_u64<<<1, 64, 0, h->stream>>>(slot, 1, NOERR);
            k_first_tag<<<nblocks(c->n_cells), PHD_BLOCK, 0, h->stream>>>(c->n_cells, c->tag, 1, slot);
            LAUNCHCHK();
            HIPCHK(hipMemcpyAsync(h->h_err + PHD_ERR_SLOT_GHOST, slot, sizeof(unsigned long long),
                                  hipMemcpyDeviceToHost, h->stream));
            HIPCHK(hipStreamSynchronize(h->stream));
            if (h->h_err[PHD_ERR_SLOT_GHOST] != NOERR) return PHD_E_ARG;
        }
        int rc = phd_primitive_from_conservative(h, c, gamma);
        if (rc || !p) return rc;
        return phd_pack_cells(h, p, c);
    }
    if (c->dim == 2)
        k_ghost_refresh_eos<2><<<xcd_grid(c->n_cells), PHD_BLOCK, 0, h->stream>>>(*c, refl, l[0], l[1], l[2], u[0], u[1], u[2], gamma, xp, nullptr, 0, 0, nullptr, nullptr);
    else
        k_ghost_refresh_eos<3><<<xcd_grid(c->n_cells), PHD_BLOCK, 0, h->stream>>>(*c, refl, l[0], l[1], l[2], u[0], u[1], u[2], gamma, xp, nullptr, 0, 0, nullptr, nullptr);
    LAUNCHCHK();
    return 0;
}

} // extern "C"

// ---------------------------------------------------------------------------
// batched ensembles: many independent problems in one union mesh (problem p
// owns cells [cell_start[p], cell_start[p+1]) and a run of faces).  Every
// kernel is flat over the union; the time step is per problem: one block per
// problem takes the CFL minimum over its real cells, and the face kernel
// reads its problem's dt through a per-face problem id.

struct phd_batch {
    int dim, n_problems, n_cells, n_faces;
    int32_t* cell_start;    // n_problems + 1
    int32_t* fprob;         // per face
    double* dt;             // per problem
    double* hist;           // hist_cap x n_problems
    long long hist_cap;
    int32_t* counter;       // steps recorded since the last reset
    int32_t* fault;         // per problem: 0, or 1 + the step count when it was frozen
    unsigned long long* dtbits;  // per problem: CFL minimum folded by the refresh (bits)
    int32_t* badp;          // per problem: a non-positive state seen by the refresh
    int32_t* bad;           // creation check: faces joining two problems
    int32_t* ghosts;        // the ghost cells (tag 1), ascending
    int n_ghosts;
    hipStream_t stream;
    hipEvent_t ev_in, ev_out;
    hipGraph_t graph;
    hipGraphExec_t exec;
    phd_cells cap_cells;    // the captured step's cells and packed rows, for the
    double* cap_xp;         // primitive arrays written after the replays
    phd_schedule sched;
    void* base;
    size_t bytes;
};

namespace {

// One thread per problem, at the start of a step: dt[p] = cfl * the CFL
// minimum the last refresh folded (riemann.pyx:120-188), or 0 for a frozen
// problem: one whose state the refresh found with a non-positive mass,
// energy, density or pressure (the previous update failed, where the
// reference raises for its single problem) is frozen from then on and
// recorded in fault[p]; the other problems go on.  Clears the accumulators.
__global__ __launch_bounds__(PHD_BLOCK) void k_batch_dt(int n_problems, double cfl,
                                                        unsigned long long* __restrict__ dtbits,
                                                        int32_t* __restrict__ badp,
                                                        int32_t* __restrict__ fault,
                                                        double* __restrict__ dt,
                                                        double* __restrict__ hist,
                                                        long long hist_cap,
                                                        const int32_t* __restrict__ counter)
{
    const int p = blockIdx.x * PHD_BLOCK + threadIdx.x;
    if (p >= n_problems) return;
    const int k = counter[0];
    if (badp[p] && fault[p] == 0) fault[p] = k + 1;
    const double d = fault[p] ? 0.0 : cfl * __longlong_as_double((long long)dtbits[p]);
    dt[p] = d;
    if (hist && k < hist_cap) hist[(long long)k * n_problems + p] = d;
    dtbits[p] = INF_BITS;
    badp[p] = 0;
}

__global__ __launch_bounds__(PHD_BLOCK) void k_batch_acc_clear(int n_problems,
                                                               unsigned long long* dtbits,
                                                               int32_t* badp)
{
    const int p = blockIdx.x * PHD_BLOCK + threadIdx.x;
    if (p >= n_problems) return;
    dtbits[p] = INF_BITS;
    badp[p] = 0;
}

__global__ void k_batch_tick(int32_t* counter)
{
    if (threadIdx.x == 0) counter[0] += 1;
}

__global__ void k_batch_zero(int32_t* counter)
{
    if (threadIdx.x == 0) counter[0] = 0;
}

__global__ void k_batch_clear(int32_t* a, int n)
{
    const int i = blockIdx.x * PHD_BLOCK + threadIdx.x;
    if (i < n) a[i] = 0;
}

// problem of each face (binary search of pair_i in cell_start); flags a
// face whose pair_j lies in another problem
__global__ __launch_bounds__(PHD_BLOCK) void k_batch_fprob(phd_faces f, int n_problems,
                                                           const int32_t* __restrict__ cell_start,
                                                           int32_t* __restrict__ fprob,
                                                           int32_t* bad)
{
    long long m = xcd_index(f.n_faces);
    if (m >= f.n_faces) return;
    const int i = f.pair_i[m], j = f.pair_j[m];
    int lo = 0, hi = n_problems - 1;
    while (lo < hi) {                    // largest p with cell_start[p] <= i
        const int mid = (lo + hi + 1) >> 1;
        if (cell_start[mid] <= i) lo = mid; else hi = mid - 1;
    }
    fprob[m] = lo;
    if (j < cell_start[lo] || j >= cell_start[lo + 1]) bad[0] = 1;
}

// the step's three main kernels under a schedule (block size, XCD remap)
#define PHD_SCHED_CASES(X)                                    \
    X(64, 0) X(64, 1) X(128, 0) X(128, 1) X(256, 0) X(256, 1)

template <int D>
int batch_grad(phd_handle* h, phd_plan* p, const phd_cells* c, const phd_schedule& s, double* rec,
               int ld)
{
    if (p->bad_cell >= 0) {
        k_set_err<<<1, 64, 0, h->stream>>>(h->d_err, 1, (unsigned long long)p->bad_cell);
        LAUNCHCHK();
    }
    const long long n = c->n_cells;
#define X(BS, XC)                                                                              \
    if (s.grad_block == BS && s.xcd == XC) {                                                   \
        k_grad_sell<D, PHD_LIMITER_AREPO, BS, XC><<<sched_grid<BS, XC>(n), BS, 0, h->stream>>>( \
            c->n_cells, c->tag, p->cell_geo, c->volume, p->xp, p->off, p->perm, p->sbase,      \
            p->sj, p->sgeo[0], p->sgeo[1], p->sgeo[2], p->sgeo[3], p->hcnt, p->halo, p->sloc,  \
            rec, ld);                                                                          \
        LAUNCHCHK();                                                                           \
        return 0;                                                                              \
    }
    PHD_SCHED_CASES(X)
#undef X
    return PHD_E_ARG;
}

template <int D>
int batch_face(phd_handle* h, phd_batch* b, phd_plan* p, const phd_cells* c, const phd_faces* f,
               const double* rec, int ld, double gamma, int boost, int gizmo)
{
    const phd_schedule& s = b->sched;
    const long long n = f->n_faces;
    if (n == 0) return 0;
    const FaceArrays none = fa(nullptr, 0);
    const int bg = (boost ? 1 : 0) | (gizmo ? 2 : 0);
#define XB(BS, XC, B)                                                                          \
    k_face<D, PHD_ORDER_LINEAR, PHD_SOLVER_HLLC, 2, 1, BS, XC, B>                              \
        <<<sched_grid<BS, XC>(n), BS, 0, h->stream>>>(*c, *f, p->face_rec, rec, ld,             \
                                                      make_gas(gamma), 0.0, 0.0, boost, gizmo,  \
                                                      none, p->g, h->d_err, b->dt, b->fprob)
#define X(BS, XC)                                                                              \
    if (s.face_block == BS && s.xcd == XC) {                                                   \
        if (bg == 0) XB(BS, XC, 0);                                                            \
        else if (bg == 1) XB(BS, XC, 1);                                                       \
        else if (bg == 2) XB(BS, XC, 2);                                                       \
        else XB(BS, XC, 3);                                                                    \
        LAUNCHCHK();                                                                           \
        return 0;                                                                              \
    }
    PHD_SCHED_CASES(X)
#undef X
#undef XB
    return PHD_E_ARG;
}

// the update fused with the real cells' epilogue, then the ghosts
template <int D>
int batch_update_eos(phd_handle* h, phd_batch* b, phd_plan* p, const phd_cells* c, int boundary,
                     const double* lo, const double* hi, double gamma, int boost, bool prim)
{
    const phd_schedule& s = b->sched;
    const long long n = c->n_cells;
    bool done = false;
#define XU(BS, XC, PR, OR)                                                                     \
    k_update_eos<D, BS, XC, PR, OR><<<sched_grid<BS, XC>(n), BS, 0, h->stream>>>(              \
        *c, p->g, p->csr, p->off, p->perm, p->sbase, p->sfi, p->cinfo, h->d_err, gamma, p->xp,  \
        b->cell_start, b->n_problems, boost, b->dtbits, b->badp)
#define X(BS, XC)                                                                              \
    if (!done && s.update_block == BS && s.xcd == XC) {                                        \
        if (prim && p->owner_runs) XU(BS, XC, 1, 1);                                           \
        else if (prim) XU(BS, XC, 1, 0);                                                       \
        else if (p->owner_runs) XU(BS, XC, 0, 1);                                              \
        else XU(BS, XC, 0, 0);                                                                 \
        LAUNCHCHK();                                                                           \
        done = true;                                                                           \
    }
    PHD_SCHED_CASES(X)
#undef X
#undef XU
    if (!done) return PHD_E_ARG;
    if (b->n_ghosts > 0) {
        double l[3], u[3];
        bounds3(D, boundary, lo, hi, l, u);
        k_batch_ghosts<D><<<xcd_grid(b->n_ghosts), PHD_BLOCK, 0, h->stream>>>(
            *c, b->ghosts, b->n_ghosts, boundary == PHD_BOUNDARY_REFLECTIVE, l[0], l[1], l[2],
            u[0], u[1], u[2], gamma, p->xp);
        LAUNCHCHK();
    }
    return 0;
}
#undef PHD_SCHED_CASES

bool sched_ok(const phd_schedule& s)
{
    auto ok = [](int v) { return v == 64 || v == 128 || v == 256; };
    return ok(s.grad_block) && ok(s.face_block) && ok(s.update_block) && (s.xcd == 0 || s.xcd == 1);
}

void drop_graph(phd_batch* b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// ghost refresh + EOS of the union, folding the per-problem CFL minimum and
// fault flags for the next step
template <int D>
int batch_refresh(phd_handle* h, phd_batch* b, phd_plan* p, const phd_cells* c, int boundary,
                  const double* lo, const double* hi, double gamma, int boost)
{
    double l[3], u[3];
    bounds3(D, boundary, lo, hi, l, u);
    k_ghost_refresh_eos<D, 1><<<xcd_grid(c->n_cells), PHD_BLOCK, 0, h->stream>>>(
        *c, boundary == PHD_BOUNDARY_REFLECTIVE, l[0], l[1], l[2], u[0], u[1], u[2], gamma,
        p->xp, b->cell_start, b->n_problems, boost, b->dtbits, b->badp);
    LAUNCHCHK();
    return 0;
}

unsigned pblocks(int n) { return (unsigned)((n + PHD_BLOCK - 1) / PHD_BLOCK); }

// batch_step flag (internal, above the public PHD_CORE_* bits): the update
// leaves the real cells' primitive arrays to a k_batch_prim after the steps
constexpr int BATCH_DEFER_PRIM = 1 << 16;

template <int D>
int batch_step(phd_handle* h, phd_batch* b, phd_plan* p, const phd_cells* c, const phd_faces* f,
               double* rec, int ld, int boundary, const double* lo, const double* hi,
               double gamma, double cfl, int boost, int gizmo, int flags)
{
    k_batch_dt<<<pblocks(b->n_problems), PHD_BLOCK, 0, h->stream>>>(
        b->n_problems, cfl, b->dtbits, b->badp, b->fault, b->dt, b->hist, b->hist_cap, b->counter);
    LAUNCHCHK();
    int rc = 0;
    if (!(flags & PHD_CORE_XP_FRESH)) rc = phd_pack_cells(h, p, c);
    if (!rc) rc = batch_grad<D>(h, p, c, b->sched, rec, ld);
    if (!rc) rc = phd_update_ghost_gradients(h, p, c, boundary, lo, hi, rec, ld);
    if (!rc) rc = batch_face<D>(h, b, p, c, f, rec, ld, gamma, boost, gizmo);
    if (!rc) rc = batch_update_eos<D>(h, b, p, c, boundary, lo, hi, gamma, boost,
                                      !(flags & BATCH_DEFER_PRIM));
    if (rc) return rc;
    k_batch_tick<<<1, 64, 0, h->stream>>>(b->counter);
    LAUNCHCHK();
    return 0;
}

int batch_args_ok(phd_handle* h, phd_batch* b, phd_plan* p, const phd_cells* c, const phd_faces* f,
                  const double* rec, int ld, int boundary, const double* lo, const double* hi)
{
    if (!h || !b || check_cells(c) || check_faces(f) || check_plan(p, c) ||
        check_plan_faces(p, f) || check_rec(c, rec, ld))
        return PHD_E_ARG;
    if (b->dim != c->dim || b->n_cells != c->n_cells || b->n_faces != f->n_faces) return PHD_E_ARG;
    if (boundary != PHD_BOUNDARY_REFLECTIVE && boundary != PHD_BOUNDARY_PERIODIC) return PHD_E_MODE;
    if (boundary == PHD_BOUNDARY_REFLECTIVE && (!lo || !hi)) return PHD_E_ARG;
    if (!c->map) return PHD_E_ARG;       // ghosts are refreshed from their images
    return 0;
}

} // namespace

extern "C" {

int phd_batch_create(phd_handle* h, phd_plan* p, const phd_cells* c, const phd_faces* f,
                     int32_t n_problems, const int32_t* cell_start, int64_t hist_steps,
                     phd_batch** out)
{
    if (!h || !out || check_cells(c) || check_faces(f) || check_plan(p, c) ||
        check_plan_faces(p, f) || n_problems < 1 || !cell_start || hist_steps < 0)
        return PHD_E_ARG;
    *out = nullptr;
    if (cell_start[0] != 0 || cell_start[n_problems] != c->n_cells) return PHD_E_ARG;
    for (int q = 0; q < n_problems; q++)
        if (cell_start[q + 1] <= cell_start[q]) return PHD_E_ARG;
    // the ghost list (tags are 0 real or 1 ghost)
    std::vector<int8_t> tag((size_t)c->n_cells);
    if (c->n_cells > 0 &&
        (hipMemcpyAsync(tag.data(), c->tag, (size_t)c->n_cells, hipMemcpyDeviceToHost, h->stream) !=
             hipSuccess ||
         hipStreamSynchronize(h->stream) != hipSuccess))
        return PHD_E_HIP;
    std::vector<int32_t> gl;
    for (int i = 0; i < c->n_cells; i++) {
        if (tag[i] == 1) gl.push_back(i);
        else if (tag[i] != 0) return PHD_E_ARG;
    }
    auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
    const size_t s_gl = al(4 * gl.size() + 4);
    const size_t s_cs = al(4 * (size_t)(n_problems + 1)), s_fp = al(4 * (size_t)f->n_faces + 4);
    const size_t s_dt = al(8 * (size_t)n_problems), s_h = al(8 * (size_t)hist_steps * (size_t)n_problems);
    const size_t s_ct = al(8), s_fa = al(4 * (size_t)n_problems), s_db = al(8 * (size_t)n_problems);
    phd_batch* b = new phd_batch();
    b->dim = c->dim;
    b->n_problems = n_problems;
    b->n_cells = c->n_cells;
    b->n_faces = f->n_faces;
    b->hist_cap = hist_steps;
    b->sched = phd_schedule{256, 256, 256, 1};
    b->n_ghosts = (int)gl.size();
    b->bytes = s_cs + s_fp + s_dt + s_h + 2 * s_ct + 2 * s_fa + s_db + s_gl;
    if (hipMalloc(&b->base, b->bytes) != hipSuccess) { delete b; return PHD_E_HIP; }
    char* q = (char*)b->base;
    b->cell_start = (int32_t*)q; q += s_cs;
    b->fprob = (int32_t*)q; q += s_fp;
    b->dt = (double*)q; q += s_dt;
    b->hist = hist_steps > 0 ? (double*)q : nullptr; q += s_h;
    b->counter = (int32_t*)q; q += s_ct;
    b->fault = (int32_t*)q; q += s_fa;
    b->dtbits = (unsigned long long*)q; q += s_db;
    b->badp = (int32_t*)q; q += s_fa;
    b->bad = (int32_t*)q; q += s_ct;
    b->ghosts = (int32_t*)q;
    int32_t hbad = 0;
    bool ok = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess;
    if (ok) ok = hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming) == hipSuccess;
    if (ok) ok = hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming) == hipSuccess;
    ok = ok && hipMemcpyAsync(b->cell_start, cell_start, 4 * (size_t)(n_problems + 1),
                              hipMemcpyHostToDevice, h->stream) == hipSuccess;
    if (!gl.empty())
        ok = ok && hipMemcpyAsync(b->ghosts, gl.data(), 4 * gl.size(), hipMemcpyHostToDevice,
                                  h->stream) == hipSuccess;
    if (ok) {
        k_batch_zero<<<1, 64, 0, h->stream>>>(b->counter);
        k_batch_zero<<<1, 64, 0, h->stream>>>(b->bad);
        k_batch_clear<<<(n_problems + PHD_BLOCK - 1) / PHD_BLOCK, PHD_BLOCK, 0, h->stream>>>(b->fault, n_problems);
        k_batch_acc_clear<<<pblocks(n_problems), PHD_BLOCK, 0, h->stream>>>(n_problems, b->dtbits, b->badp);
        if (f->n_faces > 0)
            k_batch_fprob<<<xcd_grid(f->n_faces), PHD_BLOCK, 0, h->stream>>>(*f, n_problems, b->cell_start,
                                                                           b->fprob, b->bad);
        ok = hipPeekAtLastError() == hipSuccess;
    }
    ok = ok && hipMemcpyAsync(&hbad, b->bad, 4, hipMemcpyDeviceToHost, h->stream) == hipSuccess;
    ok = ok && hipStreamSynchronize(h->stream) == hipSuccess;
    if (!ok || hbad) {
        (void)hipGetLastError();
        phd_batch_destroy(b);
        return ok ? PHD_E_ARG : PHD_E_HIP;
    }
    *out = b;
    return 0;
}

int phd_batch_destroy(phd_batch* b)
{
    if (!b) return PHD_E_ARG;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    if (b->ev_in) (void)hipEventDestroy(b->ev_in);
    if (b->ev_out) (void)hipEventDestroy(b->ev_out);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    (void)hipFree(b->base);
    delete b;
    return 0;
}

int phd_batch_set_schedule(phd_batch* b, const phd_schedule* s)
{
    if (!b || !s || !sched_ok(*s)) return PHD_E_ARG;
    drop_graph(b);
    b->sched = *s;
    return 0;
}

int phd_batch_step(phd_handle* h, phd_batch* b, phd_plan* p, const phd_cells* c,
                   const phd_faces* f, double* rec, int32_t rec_ld, int boundary,
                   const double* lo, const double* hi, double gamma, double cfl, int boost,
                   int gizmo, int flags)
{
    const int rc = batch_args_ok(h, b, p, c, f, rec, rec_ld, boundary, lo, hi);
    if (rc) return rc;
    if (flags & ~PHD_CORE_XP_FRESH) return PHD_E_ARG;
    return c->dim == 2 ? batch_step<2>(h, b, p, c, f, rec, rec_ld, boundary, lo, hi, gamma, cfl, boost, gizmo, flags)
                       : batch_step<3>(h, b, p, c, f, rec, rec_ld, boundary, lo, hi, gamma, cfl, boost, gizmo, flags);
}

int phd_batch_capture(phd_handle* h, phd_batch* b, phd_plan* p, const phd_cells* c,
                      const phd_faces* f, double* rec, int32_t rec_ld, int boundary,
                      const double* lo, const double* hi, double gamma, double cfl, int boost,
                      int gizmo)
{
    int rc = batch_args_ok(h, b, p, c, f, rec, rec_ld, boundary, lo, hi);
    if (rc) return rc;
    drop_graph(b);
    phd_handle hc = *h;            // same error words, the batch's stream
    hc.stream = b->stream;
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int fl = PHD_CORE_XP_FRESH | BATCH_DEFER_PRIM;
    rc = c->dim == 2 ? batch_step<2>(&hc, b, p, c, f, rec, rec_ld, boundary, lo, hi, gamma, cfl, boost, gizmo, fl)
                     : batch_step<3>(&hc, b, p, c, f, rec, rec_ld, boundary, lo, hi, gamma, cfl, boost, gizmo, fl);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc || e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return rc ? rc : PHD_E_HIP;
    }
    b->graph = g;
    b->cap_cells = *c;
    b->cap_xp = p->xp;
    if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGetLastError();
        drop_graph(b);
        return PHD_E_HIP;
    }
    return 0;
}

int phd_batch_launch(phd_handle* h, phd_batch* b, int32_t n_steps)
{
    if (!h || !b || !b->exec || n_steps < 0) return PHD_E_ARG;
    HIPCHK(hipEventRecord(b->ev_in, h->stream));
    HIPCHK(hipStreamWaitEvent(b->stream, b->ev_in, 0));
    for (int s = 0; s < n_steps; s++) HIPCHK(hipGraphLaunch(b->exec, b->stream));
    if (n_steps > 0) {
        // the captured steps deferred the real cells' primitive arrays
        const phd_cells& cc = b->cap_cells;
        if (cc.dim == 2) k_batch_prim<2><<<xcd_grid(cc.n_cells), PHD_BLOCK, 0, b->stream>>>(cc, b->cap_xp);
        else k_batch_prim<3><<<xcd_grid(cc.n_cells), PHD_BLOCK, 0, b->stream>>>(cc, b->cap_xp);
        LAUNCHCHK();
    }
    HIPCHK(hipEventRecord(b->ev_out, b->stream));
    HIPCHK(hipStreamWaitEvent(h->stream, b->ev_out, 0));
    return 0;
}

int phd_batch_refresh(phd_handle* h, phd_batch* b, phd_plan* p, const phd_cells* c,
                      int boundary, const double* lo, const double* hi, double gamma, int boost)
{
    if (!h || !b || check_cells(c) || check_plan(p, c)) return PHD_E_ARG;
    if (b->dim != c->dim || b->n_cells != c->n_cells || !c->map) return PHD_E_ARG;
    if (boundary != PHD_BOUNDARY_REFLECTIVE && boundary != PHD_BOUNDARY_PERIODIC) return PHD_E_MODE;
    if (boundary == PHD_BOUNDARY_REFLECTIVE && (!lo || !hi)) return PHD_E_ARG;
    k_batch_acc_clear<<<pblocks(b->n_problems), PHD_BLOCK, 0, h->stream>>>(b->n_problems, b->dtbits, b->badp);
    LAUNCHCHK();
    return c->dim == 2 ? batch_refresh<2>(h, b, p, c, boundary, lo, hi, gamma, boost)
                       : batch_refresh<3>(h, b, p, c, boundary, lo, hi, gamma, boost);
}

int phd_batch_reset_history(phd_handle* h, phd_batch* b)
{
    if (!h || !b) return PHD_E_ARG;
    k_batch_zero<<<1, 64, 0, h->stream>>>(b->counter);
    k_batch_clear<<<(b->n_problems + PHD_BLOCK - 1) / PHD_BLOCK, PHD_BLOCK, 0, h->stream>>>(b->fault, b->n_problems);
    LAUNCHCHK();
    return 0;
}

int phd_batch_read(phd_handle* h, phd_batch* b, double* dt, int64_t* steps, double* history,
                   int32_t* faults)
{
    if (!h || !b) return PHD_E_ARG;
    int32_t k = 0;
    HIPCHK(hipMemcpyAsync(&k, b->counter, 4, hipMemcpyDeviceToHost, h->stream));
    if (dt) HIPCHK(hipMemcpyAsync(dt, b->dt, 8 * (size_t)b->n_problems, hipMemcpyDeviceToHost, h->stream));
    if (faults) HIPCHK(hipMemcpyAsync(faults, b->fault, 4 * (size_t)b->n_problems, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (steps) *steps = k;
    if (history && b->hist && k > 0) {
        const long long n = (k < b->hist_cap ? k : b->hist_cap) * (long long)b->n_problems;
        HIPCHK(hipMemcpyAsync(history, b->hist, 8 * (size_t)n, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    return 0;
}

} // extern "C"
