// adjoint.hip -- reverse of one static-mesh hydro step on the GPU.
//
// The step being differentiated is what phd_core_step + phd_ghost_refresh_eos
// do with face velocities held fixed (a static or prescribed mesh):
//
//   prim = EOS(cons) (ghosts: images, boundary flips)
//   rec  = limited gradients (eq. 21 + AREPO), ghost gradients
//   L, R = extrapolation + GIZMO, MUSCL-Hancock predictor at dt/2
//   F    = HLLC + deboost
//   cons' = cons - sum (dt*area) F over owned faces, + over incoming
//
// phd_core_step_vjp maps lambda' = dL/dcons' (real cells) to
// lambda = dL/dcons, all reductions deterministic:
//   K1a face, Riemann   one thread per face: forward states, HLLC reverse;
//                       lambda of the predicted states L, R into the row
//   K1b face, states    one thread per face: predictor and extrapolation
//                       reverse; lambda of (p_i, g_i, p_j, g_j) as the row
//                       (two kernels: one would hold 272 registers, one
//                       wave per SIMD)
//   K2 cell             one thread per real cell: its face rows and its
//                       ghost copies' (with the boundary flips), in list
//                       order, then the gradient VJP: lambda of its own
//                       primitives, and one vector per list entry for the
//                       entry's other cell
//   K3 combine          one thread per real cell: own + neighbour entries
//                       (through the mirror-entry table), EOS VJP, + lambda'
#include "flux_common.h"
#include "adjoint_dev.h"

#include <vector>

struct phd_adjoint {
    int dim, n_cells, n_faces, n_real;
    long long n_entries;
    int32_t* mirror;        // per list entry: position of the same face in the other cell's list
    int32_t* gofs;          // n_real + 1: ghost copies of each real cell
    int32_t* gids;
    double* frow;           // F x FW face adjoint rows
    double* contrib;        // n_entries x NF: gradient-VJP lambda for the entry's other cell
    double* lself;          // Na x NF: gradient-VJP lambda for the cell's own primitives
    void* base;
    size_t bytes;
};

namespace {

template <int D> constexpr int pg_width() { return Lay<D>::NF + Lay<D>::G; }
template <int D> constexpr int frow_width() { return 2 * pg_width<D>(); }

__global__ __launch_bounds__(PHD_BLOCK) void k_adj_mirror(int n_cells, const int2* csr,
                                                          const int32_t* off, int32_t* mirror)
{
    long long c = xcd_index(n_cells);
    if (c >= n_cells) return;
    for (int e = off[c]; e < off[c + 1]; e++) {
        const int2 en = csr[e];
        const int o = en.y & 0x7fffffff;
        int pos = -1;
        if (o != INVALID) {
            int lo = off[o], hi = off[o + 1] - 1;
            while (lo <= hi) {       // lists are ascending in face id
                const int mid = (lo + hi) >> 1;
                const int f = csr[mid].x;
                if (f == en.x) { pos = mid; break; }
                if (f < en.x) lo = mid + 1; else hi = mid - 1;
            }
        }
        mirror[e] = pos;
    }
}

// flips of an exterior ghost: bit q set when it lies outside the box on axis q
template <int D>
__device__ __forceinline__ int ghost_flips(const phd_cells& c, int g, int reflect, const double* lo,
                                           const double* hi)
{
    int m = 0;
    if (!reflect || c.type[g] != 2) return 0;
#pragma unroll
    for (int q = 0; q < D; q++) {
        const double xq = c.x[q][g];
        if (xq < lo[q] || xq > hi[q]) m |= 1 << q;
    }
    return m;
}

// K1a: forward face states + HLLC reverse -> lambda of the predicted states.
// BG: boost (bit 0) and gizmo (bit 1) as template constants, as in k_face
template <int D, int BG>
__global__ __launch_bounds__(PHD_BLOCK) void k_adj_face_a(phd_cells c, phd_faces f,
                                                          const double* __restrict__ frec,
                                                          const double* __restrict__ rec, int ld,
                                                          Gas gas, double dt, int boost, int gizmo,
                                                          FaceArrays lam_out,
                                                          double* __restrict__ frow)
{
    using Y = Lay<D>;
    constexpr int NF = Y::NF, W = frow_width<D>();
    boost = BG & 1;
    gizmo = (BG >> 1) & 1;
    long long m = xcd_index(f.n_faces);
    if (m >= f.n_faces) return;
    const int i = f.pair_i[m], j = f.pair_j[m];
    double com[D], n[D], area, diff, fv[D];
    load_face_rec<D>(frec, m, com, area, n, diff);
#pragma unroll
    for (int k = 0; k < D; k++) fv[k] = f.velocity[k][m];
    double ri[Y::LD], rj[Y::LD];
    load_rec<D>(rec, ld, i, ri);
    load_rec<D>(rec, ld, j, rj);
    const double* pi = ri + Y::P0;
    const double* pj = rj + Y::P0;
    const double* gi = ri + Y::G0;
    const double* gj = rj + Y::G0;
    double L[NF], R[NF];
    boost_copy<D>(pi, pj, fv, boost, L, R);
    spatial_states<D>(com, diff, ri + Y::C0, rj + Y::C0, pi, pj, gi, gj, gizmo, L, R);
    const double dt_half = 0.5 * dt;
    temporal_side<D>(pi, gi, fv, boost, gas.gamma, dt_half, L);
    temporal_side<D>(pj, gj, fv, boost, gas.gamma, dt_half, R);
    const double t = dt * area;
    const bool jreal = c.tag[j] == 0;
    double lF[NF], lL[NF], lR[NF];
#pragma unroll
    for (int q = 0; q < NF; q++) {
        lF[q] = t * ((jreal ? lam_out.q[q][j] : 0.0) - lam_out.q[q][i]);
        lL[q] = lR[q] = 0.0;
    }
    hllc_vjp<D>(L, R, n, fv, gas, lF, lL, lR);
    double* o = frow + (size_t)m * W;
#pragma unroll
    for (int q = 0; q < NF; q++) { o[q] = lL[q]; o[NF + q] = lR[q]; }
}

// K1b: predictor and extrapolation reverse -> the face row (BG as in K1a)
template <int D, int BG>
__global__ __launch_bounds__(PHD_BLOCK) void k_adj_face_b(phd_cells c, phd_faces f,
                                                          const double* __restrict__ frec,
                                                          const double* __restrict__ rec, int ld,
                                                          Gas gas, double dt, int boost, int gizmo,
                                                          double* __restrict__ frow)
{
    using Y = Lay<D>;
    constexpr int NF = Y::NF, G = Y::G, W = frow_width<D>();
    boost = BG & 1;
    gizmo = (BG >> 1) & 1;
    long long m = xcd_index(f.n_faces);
    if (m >= f.n_faces) return;
    const int i = f.pair_i[m], j = f.pair_j[m];
    double com[D], n[D], area, diff, fv[D];
    load_face_rec<D>(frec, m, com, area, n, diff);
#pragma unroll
    for (int k = 0; k < D; k++) fv[k] = f.velocity[k][m];
    double ri[Y::LD], rj[Y::LD];
    load_rec<D>(rec, ld, i, ri);
    load_rec<D>(rec, ld, j, rj);
    const double* pi = ri + Y::P0;
    const double* pj = rj + Y::P0;
    const double* gi = ri + Y::G0;
    const double* gj = rj + Y::G0;
    double Ls[NF], Rs[NF];
    boost_copy<D>(pi, pj, fv, boost, Ls, Rs);
    spatial_states<D>(com, diff, ri + Y::C0, rj + Y::C0, pi, pj, gi, gj, gizmo, Ls, Rs);
    double lL[NF], lR[NF];
    const double* r = frow + (size_t)m * W;
#pragma unroll
    for (int q = 0; q < NF; q++) { lL[q] = r[q]; lR[q] = r[NF + q]; }
    const double dt_half = 0.5 * dt;
    double lLs[NF], lRs[NF], lpi[NF], lpj[NF], lgi[G], lgj[G];
#pragma unroll
    for (int q = 0; q < NF; q++) { lLs[q] = lRs[q] = lpi[q] = lpj[q] = 0.0; }
#pragma unroll
    for (int q = 0; q < G; q++) { lgi[q] = lgj[q] = 0.0; }
    temporal_side_vjp<D>(pi, gi, fv, boost, gas.gamma, dt_half, Ls, lL, lLs, lpi, lgi);
    temporal_side_vjp<D>(pj, gj, fv, boost, gas.gamma, dt_half, Rs, lR, lRs, lpj, lgj);
    spatial_states_vjp<D>(com, diff, ri + Y::C0, rj + Y::C0, pi, pj, gi, gj, fv, boost, gizmo, lLs,
                          lRs, lpi, lpj, lgi, lgj);
    double row[W];
#pragma unroll
    for (int q = 0; q < NF; q++) { row[q] = lpi[q]; row[pg_width<D>() + q] = lpj[q]; }
#pragma unroll
    for (int q = 0; q < G; q++) { row[NF + q] = lgi[q]; row[pg_width<D>() + NF + q] = lgj[q]; }
    double2* o = reinterpret_cast<double2*>(frow + (size_t)m * W);
#pragma unroll
    for (int q = 0; q < W / 2; q++) o[q] = make_double2(row[2 * q], row[2 * q + 1]);
}

template <int D>
struct CsrNbr {
    const int2* csr;
    int beg, n;
    const double* xp;
    const double* frec;
    double* contrib;
    PHD_HD int deg() const { return n; }
    PHD_HD bool valid(int e) const { return (csr[beg + e].y & 0x7fffffff) != INVALID; }
    PHD_HD void row(int e, double* r) const
    {
        const int o = csr[beg + e].y & 0x7fffffff;
        // as [position | primitive], the order grad_cell_vjp reads
        const double* s = xp + (size_t)o * XPW;
#pragma unroll
        for (int k = 0; k < D; k++) r[k] = s[XpLay<D>::X + k];
#pragma unroll
        for (int q = 0; q < D + 2; q++) r[D + q] = s[XpLay<D>::P + q];
    }
    PHD_HD void geo(int e, double* cm, double& a) const
    {
        using F = FrLay<D>;
        const double* fr = frec + (size_t)csr[beg + e].x * F::W;
        cm[0] = fr[F::C];
        cm[1] = fr[F::C + 1];
        cm[2] = D == 3 ? fr[F::C + 2] : 0.0;
        a = fr[F::A];
    }
    PHD_HD void out(int e, const double* lam) const
    {
#pragma unroll
        for (int q = 0; q < D + 2; q++) contrib[(size_t)(beg + e) * (D + 2) + q] = lam[q];
    }
};

// sum of the face rows of cell cc's list, in list order (the cell's side of
// each face row)
template <int D>
__device__ __forceinline__ void gather_rows(int cc, const int2* __restrict__ csr,
                                            const int32_t* __restrict__ off,
                                            const double* __restrict__ frow, double* s)
{
    constexpr int PW = pg_width<D>(), W = frow_width<D>();
    for (int e = off[cc]; e < off[cc + 1]; e++) {
        const int2 en = csr[e];
        if ((en.y & 0x7fffffff) == INVALID) continue;
        const double* r = frow + (size_t)en.x * W + ((en.y & INCOMING) ? PW : 0);
#pragma unroll
        for (int q = 0; q < PW; q++) s[q] += r[q];
    }
}

// K2: one thread per real cell.  lambda of its primitives and limited
// gradient = its face rows + those of its ghost copies (a ghost g of k has
// prim_g = prim_k with velocity q negated for each flip axis q, and grad_g =
// grad_k with d v_q / d x_q negated); then the gradient VJP: lself gets the
// primitives' total (faces + own gradient), contrib one vector per list
// entry for the entry's other cell
template <int D>
__global__ __launch_bounds__(PHD_BLOCK) void k_adj_cell(phd_cells c, int n_real,
                                                        const int2* __restrict__ csr,
                                                        const int32_t* __restrict__ off,
                                                        const int32_t* __restrict__ gofs,
                                                        const int32_t* __restrict__ gids,
                                                        int reflect, double lo0, double lo1,
                                                        double lo2, double hi0, double hi1,
                                                        double hi2, const double* __restrict__ xp,
                                                        const double* __restrict__ geo,
                                                        const double* __restrict__ frec,
                                                        const double* __restrict__ frow,
                                                        double* __restrict__ contrib,
                                                        double* __restrict__ lself)
{
    using Y = Lay<D>;
    constexpr int NF = Y::NF, PW = pg_width<D>();
    long long ii = xcd_index(n_real);
    if (ii >= n_real) return;
    const int i = (int)ii;
    const double lo[3] = {lo0, lo1, lo2}, hi[3] = {hi0, hi1, hi2};
    double s[PW];
#pragma unroll
    for (int q = 0; q < PW; q++) s[q] = 0.0;
    gather_rows<D>(i, csr, off, frow, s);
    for (int a = gofs[i]; a < gofs[i + 1]; a++) {
        const int g = gids[a];
        const int fl = ghost_flips<D>(c, g, reflect, lo, hi);
        double t[PW];
#pragma unroll
        for (int q = 0; q < PW; q++) t[q] = 0.0;
        gather_rows<D>(g, csr, off, frow, t);
#pragma unroll
        for (int q = 0; q < D; q++) {
            if (fl & (1 << q)) {
                t[1 + q] = -t[1 + q];
                t[NF + (1 + q) * D + q] = -t[NF + (1 + q) * D + q];
            }
        }
#pragma unroll
        for (int q = 0; q < PW; q++) s[q] += t[q];
    }
    double pv[NF], xi[D], cx[D], vol;
#pragma unroll
    for (int q = 0; q < NF; q++) pv[q] = xp[(size_t)i * XPW + XpLay<D>::P + q];
#pragma unroll
    for (int k = 0; k < D; k++) { xi[k] = xp[(size_t)i * XPW + XpLay<D>::X + k]; cx[k] = geo[(size_t)i * GEO + k]; }
    vol = geo[(size_t)i * GEO + 3];
    CsrNbr<D> nb{csr, off[i], off[i + 1] - off[i], xp, frec, contrib};
    grad_cell_vjp<D>(nb, pv, xi, cx, vol, s + NF, s);
#pragma unroll
    for (int q = 0; q < NF; q++) lself[(size_t)i * NF + q] = s[q];
}

// K3
template <int D>
__global__ __launch_bounds__(PHD_BLOCK) void k_adj_combine(phd_cells c, int n_real,
                                                           const int2* __restrict__ csr,
                                                           const int32_t* __restrict__ off,
                                                           const int32_t* __restrict__ mirror,
                                                           const int32_t* __restrict__ gofs,
                                                           const int32_t* __restrict__ gids,
                                                           int reflect, double lo0, double lo1,
                                                           double lo2, double hi0, double hi1,
                                                           double hi2,
                                                           const double* __restrict__ lself,
                                                           const double* __restrict__ contrib,
                                                           double gamma, FaceArrays lam_out,
                                                           FaceArrays lam_in)
{
    constexpr int NF = D + 2, PW = pg_width<D>();
    long long kk = xcd_index(n_real);
    if (kk >= n_real) return;
    const int k = (int)kk;
    const double lo[3] = {lo0, lo1, lo2}, hi[3] = {hi0, hi1, hi2};
    double lp[NF];
#pragma unroll
    for (int q = 0; q < NF; q++) lp[q] = lself[(size_t)k * NF + q];
    // neighbours' gradient VJPs that read this cell's primitives
    for (int e = off[k]; e < off[k + 1]; e++) {
        const int o = csr[e].y & 0x7fffffff;
        const int src = mirror[e];
        if (o == INVALID || src < 0 || c.tag[o] != 0) continue;
#pragma unroll
        for (int q = 0; q < NF; q++) lp[q] += contrib[(size_t)src * NF + q];
    }
    // ... and those that read a ghost copy of it
    for (int a = gofs[k]; a < gofs[k + 1]; a++) {
        const int g = gids[a];
        const int fl = ghost_flips<D>(c, g, reflect, lo, hi);
        for (int e = off[g]; e < off[g + 1]; e++) {
            const int o = csr[e].y & 0x7fffffff;
            const int src = mirror[e];
            if (o == INVALID || src < 0 || c.tag[o] != 0) continue;
            double t[NF];
#pragma unroll
            for (int q = 0; q < NF; q++) t[q] = contrib[(size_t)src * NF + q];
#pragma unroll
            for (int q = 0; q < D; q++)
                if (fl & (1 << q)) t[1 + q] = -t[1 + q];
#pragma unroll
            for (int q = 0; q < NF; q++) lp[q] += t[q];
        }
    }
    double u[NF], lu[NF];
#pragma unroll
    for (int q = 0; q < NF; q++) { u[q] = c.cons[q][k]; lu[q] = lam_out.q[q][k]; }
    prim_from_cons_vjp<D>(u, c.volume[k], gamma, lp, lu);
#pragma unroll
    for (int q = 0; q < NF; q++) lam_in.q[q][k] = lu[q];
}

template <int D, int TO_PRIM>
__global__ __launch_bounds__(PHD_BLOCK) void k_adj_eos(phd_cells c, double gamma, FaceArrays lam_y,
                                                       FaceArrays lam_x)
{
    constexpr int NF = D + 2;
    long long ii = xcd_index(c.n_cells);
    if (ii >= c.n_cells) return;
    const int i = (int)ii;
    if (c.tag[i] != 0) return;
    double x[NF], ly[NF], lx[NF];
#pragma unroll
    for (int q = 0; q < NF; q++) {
        x[q] = TO_PRIM ? c.cons[q][i] : c.prim[q][i];
        ly[q] = lam_y.q[q][i];
        lx[q] = lam_x.q[q][i];
    }
    if (TO_PRIM) prim_from_cons_vjp<D>(x, c.volume[i], gamma, ly, lx);
    else cons_from_prim_vjp<D>(x, c.volume[i], gamma, ly, lx);
#pragma unroll
    for (int q = 0; q < NF; q++) lam_x.q[q][i] = lx[q];
}

FaceArrays fa5(double* const* a, int n)
{
    FaceArrays r;
    for (int q = 0; q < 5; q++) r.q[q] = (a && q < n) ? a[q] : nullptr;
    return r;
}

#define ADJ_CHK()                                                                    \
    do {                                                                             \
        if (hipPeekAtLastError() != hipSuccess) { (void)hipGetLastError(); return PHD_E_HIP; } \
    } while (0)

template <int D>
int step_vjp(phd_handle* h, phd_plan* p, phd_adjoint* a, const phd_cells* c, const phd_faces* f,
             double* rec, int ld, int boundary, const double* lo, const double* hi, double gamma,
             double dt, int boost, int gizmo, double* const* lam_out, double* const* lam_in,
             int flags)
{
    double l[3] = {0, 0, 0}, u[3] = {0, 0, 0};
    const int refl = boundary == PHD_BOUNDARY_REFLECTIVE;
    if (refl) for (int k = 0; k < D; k++) { l[k] = lo[k]; u[k] = hi[k]; }
    // forward records of this step: packed rows, gradients, ghost gradients
    if (!(flags & PHD_VJP_REC_FRESH)) {
        int rc = phd_pack_cells(h, p, c);
        if (!rc) rc = phd_gradients_packed(h, p, c, PHD_LIMITER_AREPO, rec, ld);
        if (!rc) rc = phd_update_ghost_gradients(h, p, c, boundary, lo, hi, rec, ld);
        if (rc) return rc;
    }
    const FaceArrays Lo = fa5(lam_out, D + 2), Li = fa5(lam_in, D + 2);
    if (f->n_faces > 0) {
        const Gas gas = make_gas(gamma);
        const unsigned nb = xcd_grid(f->n_faces);
        const int bg = (boost ? 1 : 0) | (gizmo ? 2 : 0);
#define ADJ_FACES(B)                                                                           \
        do {                                                                                   \
            k_adj_face_a<D, B><<<nb, PHD_BLOCK, 0, h->stream>>>(*c, *f, p->face_rec, rec, ld,  \
                                                                gas, dt, boost, gizmo, Lo,     \
                                                                a->frow);                      \
            k_adj_face_b<D, B><<<nb, PHD_BLOCK, 0, h->stream>>>(*c, *f, p->face_rec, rec, ld,  \
                                                                gas, dt, boost, gizmo,         \
                                                                a->frow);                      \
        } while (0)
        if (bg == 0) ADJ_FACES(0);
        else if (bg == 1) ADJ_FACES(1);
        else if (bg == 2) ADJ_FACES(2);
        else ADJ_FACES(3);
#undef ADJ_FACES
    }
    k_adj_cell<D><<<xcd_grid(a->n_real), PHD_BLOCK, 0, h->stream>>>(
        *c, a->n_real, p->csr, p->off, a->gofs, a->gids, refl, l[0], l[1], l[2], u[0], u[1], u[2],
        p->xp, p->cell_geo, p->face_rec, a->frow, a->contrib, a->lself);
    k_adj_combine<D><<<xcd_grid(a->n_real), PHD_BLOCK, 0, h->stream>>>(
        *c, a->n_real, p->csr, p->off, a->mirror, a->gofs, a->gids, refl, l[0], l[1], l[2], u[0],
        u[1], u[2], a->lself, a->contrib, gamma, Lo, Li);
    ADJ_CHK();
    return 0;
}

} // namespace

extern "C" {

int phd_adjoint_create(phd_handle* h, phd_plan* p, const phd_cells* c, const phd_faces* f,
                       phd_adjoint** out)
{
    if (!h || !p || !c || !f || !out || (c->dim != 2 && c->dim != 3)) return PHD_E_ARG;
    if (p->n_cells != c->n_cells || p->n_faces != f->n_faces || p->dim != c->dim) return PHD_E_ARG;
    *out = nullptr;
    const int D = c->dim, NF = D + 2, G = D * (D + 2);
    const long long na = c->n_cells, nf = f->n_faces, ne = p->n_entries;
    // real count and ghost -> image lists on the host
    std::vector<int8_t> tag(na);
    std::vector<int32_t> map(na);
    if (na > 0) {
        if (hipMemcpyAsync(tag.data(), c->tag, na, hipMemcpyDeviceToHost, h->stream) != hipSuccess) return PHD_E_HIP;
        if (c->map && hipMemcpyAsync(map.data(), c->map, 4 * na, hipMemcpyDeviceToHost, h->stream) != hipSuccess) return PHD_E_HIP;
        if (hipStreamSynchronize(h->stream) != hipSuccess) return PHD_E_HIP;
    }
    long long nr = 0;
    while (nr < na && tag[nr] == 0) nr++;
    std::vector<int32_t> gofs(nr + 1, 0), gids;
    for (long long g = nr; g < na; g++) {
        const int k = map[g];
        if (k < 0 || k >= nr) return PHD_E_ARG;
        gofs[k + 1]++;
    }
    for (long long k = 0; k < nr; k++) gofs[k + 1] += gofs[k];
    gids.resize(gofs[nr]);
    {
        std::vector<int32_t> fill(gofs.begin(), gofs.end() - 1);
        for (long long g = nr; g < na; g++) gids[fill[map[g]]++] = (int32_t)g;
    }
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t s_m = al(4 * (size_t)ne), s_go = al(4 * (size_t)(nr + 1)), s_gi = al(4 * gids.size() + 4);
    const size_t s_fr = al(8 * (size_t)(2 * (NF + G)) * (size_t)nf);
    const size_t s_ct = al(8 * (size_t)NF * (size_t)ne), s_ls = al(8 * (size_t)NF * (size_t)na);
    phd_adjoint* a = new phd_adjoint();
    a->dim = D;
    a->n_cells = (int)na;
    a->n_faces = (int)nf;
    a->n_real = (int)nr;
    a->n_entries = ne;
    a->bytes = s_m + s_go + s_gi + s_fr + s_ct + s_ls;
    if (hipMalloc(&a->base, a->bytes) != hipSuccess) { delete a; return PHD_E_HIP; }
    char* b = (char*)a->base;
    a->mirror = (int32_t*)b; b += s_m;
    a->gofs = (int32_t*)b; b += s_go;
    a->gids = (int32_t*)b; b += s_gi;
    a->frow = (double*)b; b += s_fr;
    a->contrib = (double*)b; b += s_ct;
    a->lself = (double*)b;
    bool ok = hipMemcpyAsync(a->gofs, gofs.data(), 4 * gofs.size(), hipMemcpyHostToDevice, h->stream) == hipSuccess;
    if (ok && !gids.empty())
        ok = hipMemcpyAsync(a->gids, gids.data(), 4 * gids.size(), hipMemcpyHostToDevice, h->stream) == hipSuccess;
    if (ok && na > 0) {
        k_adj_mirror<<<xcd_grid(na), PHD_BLOCK, 0, h->stream>>>((int)na, p->csr, p->off, a->mirror);
        ok = hipPeekAtLastError() == hipSuccess;
    }
    ok = ok && hipStreamSynchronize(h->stream) == hipSuccess;
    if (!ok) {
        (void)hipGetLastError();
        (void)hipFree(a->base);
        delete a;
        return PHD_E_HIP;
    }
    *out = a;
    return 0;
}

int phd_adjoint_destroy(phd_adjoint* a)
{
    if (!a) return PHD_E_ARG;
    (void)hipFree(a->base);
    delete a;
    return 0;
}

int phd_core_step_vjp(phd_handle* h, phd_plan* p, phd_adjoint* a, const phd_cells* c,
                      const phd_faces* f, double* rec, int32_t rec_ld, int boundary,
                      const double* lo, const double* hi, double gamma, double dt, int boost,
                      int gizmo, double* const* lam_out, double* const* lam_in, int flags)
{
    if (!h || !p || !a || !c || !f || !rec || !lam_out || !lam_in) return PHD_E_ARG;
    if (flags & ~PHD_VJP_REC_FRESH) return PHD_E_ARG;
    if (a->n_cells != c->n_cells || a->n_faces != f->n_faces || a->dim != c->dim) return PHD_E_ARG;
    if (boundary != PHD_BOUNDARY_REFLECTIVE && boundary != PHD_BOUNDARY_PERIODIC) return PHD_E_MODE;
    if (boundary == PHD_BOUNDARY_REFLECTIVE && (!lo || !hi)) return PHD_E_ARG;
    if (rec_ld < phd_cell_record_ld(c->dim) || (rec_ld & 1) || (((uintptr_t)rec) & 15)) return PHD_E_ARG;
    return c->dim == 2 ? step_vjp<2>(h, p, a, c, f, rec, rec_ld, boundary, lo, hi, gamma, dt, boost, gizmo, lam_out, lam_in, flags)
                       : step_vjp<3>(h, p, a, c, f, rec, rec_ld, boundary, lo, hi, gamma, dt, boost, gizmo, lam_out, lam_in, flags);
}

int phd_eos_vjp(phd_handle* h, const phd_cells* c, double gamma, int to_prim,
                double* const* lam_y, double* const* lam_x)
{
    if (!h || !c || !lam_y || !lam_x || (c->dim != 2 && c->dim != 3)) return PHD_E_ARG;
    if (c->n_cells == 0) return 0;
    const int D = c->dim;
    const FaceArrays Y = fa5(lam_y, D + 2), X = fa5(lam_x, D + 2);
    const unsigned nb = xcd_grid(c->n_cells);
    if (D == 2) {
        if (to_prim) k_adj_eos<2, 1><<<nb, PHD_BLOCK, 0, h->stream>>>(*c, gamma, Y, X);
        else k_adj_eos<2, 0><<<nb, PHD_BLOCK, 0, h->stream>>>(*c, gamma, Y, X);
    } else {
        if (to_prim) k_adj_eos<3, 1><<<nb, PHD_BLOCK, 0, h->stream>>>(*c, gamma, Y, X);
        else k_adj_eos<3, 0><<<nb, PHD_BLOCK, 0, h->stream>>>(*c, gamma, Y, X);
    }
    ADJ_CHK();
    return 0;
}

} // extern "C"
