// flux_common.h -- definitions shared by the translation units of
// libphd_flux.so (forward kernels in phd_flux.hip, adjoint kernels in
// adjoint.hip): the handle and plan structures, record layouts, launch
// helpers and the per-face state builders.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "phd_flux.h"
#include "riemann_dev.h"
#include "states_dev.h"

#define PHD_BLOCK 256

// handle scratch words: [0] update key, [1] neighbor cell, [2] exact face,
// [3] dt bits, [4] first real cell outside the box, [5] first ghost cell;
// [6..7] spare
#define PHD_ERR_SLOTS 8
#define PHD_ERR_SLOT_OUTSIDE 4
#define PHD_ERR_SLOT_GHOST 5

struct phd_handle {
    int device;
    hipStream_t stream;
    int mode;
    unsigned long long* d_err;   // PHD_ERR_SLOTS words
    unsigned long long* h_err;   // pinned host mirror
};

struct phd_plan {
    int dim, n_cells, n_faces;
    long long n_entries;
    int device;
    void* base;                  // single allocation holding every buffer below
    size_t bytes;
    double* face_rec;
    double* cell_geo;
    int2* csr;
    int32_t* off;
    double* xp;
    unsigned long long* d_bad;
    long long bad_cell;          // first real cell with an inconsistent face list, or -1
    // sliced lists: cells are visited in the order perm (within each window
    // of 256 consecutive cells, by descending face count; reals before
    // ghosts), and the cell at position t keeps entry r of its face list in
    // slot sbase[t/16] + 16 r + t%16.  A wavefront reads 4 x 16 consecutive
    // slots per entry, and its 64 lanes have nearly equal list lengths.
    int32_t* perm;               // position -> cell, n_cells
    long long n_slots;
    void* slot_base;             // second allocation holding the slot arrays
    int32_t* sbase;              // n_slices + 1
    int32_t* sj;                 // other cell | incoming bit (INVALID for padding)
    int32_t* sfi;                // face id | incoming bit (INVALID for padding)
    int2* cinfo;                 // per position: incoming count, first owned face
    double* sgeo[4];             // face centroid x, y, z, area per slot
    double* g;                   // per-face update summand rows (dt*area)*flux, F x GW
    int owner_runs;              // every real cell's owned faces are one id run that
                                 // follows all its incoming faces (faces in owner order)
    int32_t* ghosts;             // third allocation: the ghost cells (tag 1), ascending
    int n_ghosts;
    // fourth allocation (3D): each window's halo, the first halo_max of the
    // distinct cells outside the window that its real cells' lists name
    // (ascending, halo[w * halo_max + k] for k < hcnt[w]), and per slot the
    // row of the gradient kernel's LDS staging that holds the neighbour
    // (sloc: window rows 0..255, then halo rows; LOC_GLOBAL past the
    // staging, LOC_NONE for an invalid neighbour)
    void* halo_base;
    int32_t* hcnt;
    int32_t* halo;
    uint16_t* sloc;
};

namespace {

using namespace phd;

constexpr unsigned long long NOERR = ~0ull;
constexpr unsigned long long INF_BITS = 0x7ff0000000000000ull;
constexpr int INCOMING = (int)0x80000000u;
constexpr int INVALID = 0x7fffffff;
constexpr int FREC = 8;      // doubles per face record, at most
constexpr uint16_t LOC_GLOBAL = 0xFFFE, LOC_NONE = 0xFFFF;   // sloc markers
// halo rows the gradient kernel stages beside its window's 256 rows: 576 in
// 3D (53 KB of LDS per block, three blocks per CU as its registers allow);
// none in 2D, where a tenth of the entries lie outside the window and the
// staging code would cost the kernel its fourth wave per SIMD
template <int D> constexpr int halo_max() { return D == 2 ? 0 : 576; }

// Face record: com(D) | area | normal(D) | |x_j - x_i|, 6 doubles in 2D and
// 8 in 3D (rows stay 16-byte aligned for the paired loads)
template <int D> struct FrLay {
    static constexpr int W = 2 * D + 2;
    static constexpr int C = 0, A = D, N = D + 1, DIFF = 2 * D + 1;
};
constexpr int GEO = 4;       // doubles per cell geometry row
constexpr int XPW = 8;       // doubles per packed cell row

// Packed cell row: [position | primitive] in 3D, [primitive | position |
// centroid] in 2D, where the four primitives fill the row's first aligned
// 32-byte sector, so the per-step refresh of a static mesh rewrites one whole
// sector.  The 2D centroid x + dcom is written with the plan (it is the
// plan's cell_geo centroid) and left alone by the per-step refreshes.
template <int D> struct XpLay {
    static constexpr int X = D == 2 ? 4 : 0;     // first position component
    static constexpr int P = D == 2 ? 0 : 3;     // first primitive
    static constexpr int C = D == 2 ? 6 : -1;    // first centroid component (2D)
    // double2 pairs a refresh rewrites: all but the 2D centroid pair
    static constexpr int REFRESH_PAIRS = D == 2 ? 3 : XPW / 2;
};

struct FaceArrays { double* q[5]; };

// doubles per summand row: the D+2 conserved components, unpadded (32 B in
// 2D, moved as 16-byte pairs; 40 B in 3D, moved as single doubles)
template <int D> constexpr int summand_width() { return D + 2; }

template <int D> struct Lay {
    static constexpr int NF = D + 2;                 // primitive fields
    static constexpr int G = D * (D + 2);            // gradient components
    static constexpr int REC = NF + G + D;           // used doubles per record row
    static constexpr int LD = (REC + 1) & ~1;        // 24 (3D) / 14 (2D)
    static constexpr int P0 = 0, G0 = NF, C0 = NF + G;
};

inline unsigned nblocks(long long n) { return (unsigned)((n + PHD_BLOCK - 1) / PHD_BLOCK); }
// grid for an XCD-remapped launch over n items: a multiple of 8 blocks
inline unsigned xcd_grid(long long n) { unsigned nb = nblocks(n); return 8u * ((nb + 7u) / 8u); }

// logical item index of this thread under a launch schedule: BS items per
// block; XCD = 1 remaps blocks so that each of the 8 XCDs (which take
// dispatched blocks round-robin) works on one contiguous eighth of the items
// and keeps its neighbours' rows in its own L2.  n items in total.
template <int BS, int XCD>
__device__ __forceinline__ long long sched_index(long long n)
{
    const long long b = blockIdx.x;
    if (!XCD) return b * BS + threadIdx.x;
    const long long nb = (n + BS - 1) / BS;
    const long long per = (nb + 7) / 8;
    const long long L = (b & 7) * per + (b >> 3);
    return L * BS + threadIdx.x;
}

template <int BS, int XCD>
inline unsigned sched_grid(long long n)
{
    const unsigned nb = (unsigned)((n + BS - 1) / BS);
    return XCD ? 8u * ((nb + 7u) / 8u) : nb;
}

__device__ __forceinline__ long long xcd_index(long long n) { return sched_index<PHD_BLOCK, 1>(n); }

// cell radius (riemann.pyx:155-158, mesh.pyx:414-417); pi is numpy.pi
template <int D>
__device__ __forceinline__ double cell_radius(double vol)
{
    if (D == 2) return sqrt(vol / 3.141592653589793);
    return pow(3.0 * vol / (4.0 * 3.141592653589793), 1.0 / 3.0);
}

// CFL time step (riemann.pyx:120-188).  min is order independent: a
// wavefront min and an atomic min on the bit pattern (non-negative doubles
// order as unsigned integers) give the exact, deterministic minimum.
// cfl_value: the bound of one real cell; +inf for NaN, which fmin would
// otherwise skip
template <int D>
__device__ __forceinline__ double cfl_value(const phd_cells& c, long long i, double gamma, int boost)
{
    double v;
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
    if (!(v >= 0.0)) v = __longlong_as_double((long long)INF_BITS);
    return v;
}

__global__ void k_fill_u64(unsigned long long* a, int n, unsigned long long v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

template <int D>
__device__ __forceinline__ void store_summand(double* g, long long m, const double* row)
{
    constexpr int GW = summand_width<D>();
    double* o = g + (size_t)m * GW;
    if (GW % 2 == 0) {
#pragma unroll
        for (int q = 0; q < GW / 2; q++)
            reinterpret_cast<double2*>(o)[q] = make_double2(row[2 * q], row[2 * q + 1]);
    } else {
#pragma unroll
        for (int q = 0; q < GW; q++) o[q] = row[q];
    }
}

template <int D>
__device__ __forceinline__ void load_summand(const double* g, long long m, double* s)
{
    constexpr int GW = summand_width<D>();
    const double* r = g + (size_t)m * GW;
    if (GW % 2 == 0) {
#pragma unroll
        for (int q = 0; q < GW / 2; q++) {
            const double2 v = reinterpret_cast<const double2*>(r)[q];
            s[2 * q] = v.x;
            s[2 * q + 1] = v.y;
        }
    } else {
#pragma unroll
        for (int q = 0; q < GW; q++) s[q] = r[q];
    }
}

__device__ __forceinline__ void ld2(const double* p, double& a, double& b)
{
    double2 v = *reinterpret_cast<const double2*>(p);
    a = v.x;
    b = v.y;
}

template <int D>
__device__ __forceinline__ void load_face_rec(const double* frec, long long m, double* com,
                                              double& area, double* n, double& diff)
{
    using F = FrLay<D>;
    double r[F::W];
#pragma unroll
    for (int q = 0; q < F::W / 2; q++) ld2(frec + (size_t)m * F::W + 2 * q, r[2 * q], r[2 * q + 1]);
#pragma unroll
    for (int k = 0; k < D; k++) { com[k] = r[F::C + k]; n[k] = r[F::N + k]; }
    area = r[F::A];
    diff = r[F::DIFF];
}

template <int D>
__device__ __forceinline__ void load_rec(const double* rec, int ld, int i, double* row)
{
    using Y = Lay<D>;
    const double* p = rec + (size_t)i * ld;
#pragma unroll
    for (int q = 0; q < Y::LD / 2; q++) ld2(p + 2 * q, row[2 * q], row[2 * q + 1]);
}

// left/right states of a face from its two cell-record rows ri, rj (in
// registers, or ri in LDS): spatial extrapolation + GIZMO, then the
// half-step predictor
template <int D>
__device__ __forceinline__ void face_states_rows(const double* ri, const double* rj,
                                                 const double* com, double diff, const double* fv,
                                                 double gamma, double dt, int boost, int gizmo,
                                                 int stages, double* L, double* R)
{
    using Y = Lay<D>;
    const double* pi = ri + Y::P0;
    const double* pj = rj + Y::P0;
    if (stages & PHD_STAGE_SPATIAL) {
        boost_copy<D>(pi, pj, fv, boost, L, R);
        spatial_states<D>(com, diff, ri + Y::C0, rj + Y::C0, pi, pj, ri + Y::G0, rj + Y::G0,
                          gizmo, L, R);
    }
    if (stages & PHD_STAGE_TEMPORAL) {
        temporal_side<D>(pi, ri + Y::G0, fv, boost, gamma, dt, L);
        temporal_side<D>(pj, rj + Y::G0, fv, boost, gamma, dt, R);
    }
}

// left/right states of face m into L, R (primitive group, face frame)
template <int D, int ORDER>
__device__ __forceinline__ void face_states(const phd_cells& c, const double* rec, int ld, int i,
                                            int j, const double* com, double diff,
                                            const double* fv, double gamma, double dt, int boost,
                                            int gizmo, int stages, double* L, double* R)
{
    using Y = Lay<D>;
    if (ORDER == PHD_ORDER_CONSTANT) {
        double pi[Y::NF], pj[Y::NF];
#pragma unroll
        for (int n = 0; n < Y::NF; n++) { pi[n] = c.prim[n][i]; pj[n] = c.prim[n][j]; }
        boost_copy<D>(pi, pj, fv, boost, L, R);
        return;
    }
    double ri[Y::LD], rj[Y::LD];
    load_rec<D>(rec, ld, i, ri);
    load_rec<D>(rec, ld, j, rj);
    face_states_rows<D>(ri, rj, com, diff, fv, gamma, dt, boost, gizmo, stages, L, R);
}

} // namespace
