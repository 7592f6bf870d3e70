// domain.hip -- HIP kernels and C ABI of the serial domain-manager and
// boundary-condition loops that touch cell fields between the flux steps of
// the moving-mesh integrator (domain/domain_manager.pyx, domain/boundary.pyx).
// One thread per cell; every write is the reference's own expression, so the
// fields are bit-identical to its loops.
#include "flux_common.h"

namespace {

// DomainManager.move_generators (domain_manager.pyx:566-590): x += dt*w for
// real cells.  The positions are the cells' own position arrays.
template <int D>
__global__ __launch_bounds__(PHD_BLOCK) void k_move_generators(phd_cells c, double dt)
{
    const long long i = xcd_index(c.n_cells);
    if (i >= c.n_cells || c.tag[i] != 0) return;
#pragma unroll
    for (int k = 0; k < D; k++) const_cast<double*>(c.x[k])[i] += dt * c.w[k][i];
}

// Periodic.migrate_particles (boundary.pyx:502-538): a real cell outside the
// box (intersect_bounds with r = 0, boundary.pyx:12-37) is wrapped on each
// axis, both tests made on the position before any shift (xs), so each axis
// moves at most once.
template <int D>
__global__ __launch_bounds__(PHD_BLOCK) void k_migrate_periodic(phd_cells c, double lo0,
                                                                double lo1, double lo2,
                                                                double hi0, double hi1,
                                                                double hi2, double L0,
                                                                double L1, double L2)
{
    const long long i = xcd_index(c.n_cells);
    if (i >= c.n_cells || c.tag[i] != 0) return;
    const double lo[3] = {lo0, lo1, lo2}, hi[3] = {hi0, hi1, hi2}, L[3] = {L0, L1, L2};
    double xs[D];
    bool inside = true;
#pragma unroll
    for (int k = 0; k < D; k++) {
        xs[k] = c.x[k][i];
        if (xs[k] + 0.0 < lo[k]) inside = false;
        if (xs[k] - 0.0 > hi[k]) inside = false;
    }
    if (inside) return;
#pragma unroll
    for (int k = 0; k < D; k++) {
        double x = xs[k];
        if (xs[k] <= lo[k]) x += L[k];
        if (xs[k] >= hi[k]) x -= L[k];
        const_cast<double*>(c.x[k])[i] = x;
    }
}

// Reflective.migrate_particles (boundary.pyx:236-263): the first real cell
// outside the box, by index, into *first (atomic min: the reference raises
// at the first one its loop meets).
template <int D>
__global__ __launch_bounds__(PHD_BLOCK) void k_first_outside(phd_cells c, double lo0, double lo1,
                                                             double lo2, double hi0, double hi1,
                                                             double hi2,
                                                             unsigned long long* first)
{
    const long long i = xcd_index(c.n_cells);
    if (i >= c.n_cells || c.tag[i] != 0) return;
    const double lo[3] = {lo0, lo1, lo2}, hi[3] = {hi0, hi1, hi2};
    bool inside = true;
#pragma unroll
    for (int k = 0; k < D; k++) {
        const double x = c.x[k][i];
        if (x + 0.0 < lo[k]) inside = false;
        if (x - 0.0 > hi[k]) inside = false;
    }
    if (!inside) atomicMin(first, (unsigned long long)i);
}

// Reflective.update_fields (boundary.pyx:309-348): velocity and momentum of
// every exterior cell negated on each axis where it lies outside the box.
template <int D>
__global__ __launch_bounds__(PHD_BLOCK) void k_reflect_fields(phd_cells c, double lo0, double lo1,
                                                              double lo2, double hi0, double hi1,
                                                              double hi2)
{
    const long long i = xcd_index(c.n_cells);
    if (i >= c.n_cells || c.type[i] != 2) return;
    const double lo[3] = {lo0, lo1, lo2}, hi[3] = {hi0, hi1, hi2};
#pragma unroll
    for (int k = 0; k < D; k++) {
        const double x = c.x[k][i];
        if (x < lo[k]) {
            c.prim[1 + k][i] *= -1;
            c.cons[1 + k][i] *= -1;
        }
        if (x > hi[k]) {
            c.prim[1 + k][i] *= -1;
            c.cons[1 + k][i] *= -1;
        }
    }
}

constexpr int GF_MAX = 16;   // fields per launch
struct GhostFields {
    void* p[GF_MAX];
    int32_t size[GF_MAX];    // bytes per element: 1, 4 or 8
    int32_t n;
};

// DomainManager.update_ghost_fields, serial branch (domain_manager.pyx:
// 757-770): each ghost takes every listed field from its image
__global__ __launch_bounds__(PHD_BLOCK) void k_copy_ghost_fields(int n_cells,
                                                                 const int8_t* __restrict__ tag,
                                                                 const int32_t* __restrict__ map,
                                                                 GhostFields gf)
{
    const long long i = xcd_index(n_cells);
    if (i >= n_cells || tag[i] != 1) return;
    const long long s = map[i];
    for (int f = 0; f < gf.n; f++) {
        char* a = static_cast<char*>(gf.p[f]);
        switch (gf.size[f]) {
        case 8: reinterpret_cast<int64_t*>(a)[i] = reinterpret_cast<const int64_t*>(a)[s]; break;
        case 4: reinterpret_cast<int32_t*>(a)[i] = reinterpret_cast<const int32_t*>(a)[s]; break;
        default: a[i] = a[s]; break;
        }
    }
}

void box3(int dim, const double* lo, const double* hi, double* l, double* u)
{
    for (int k = 0; k < 3; k++) {
        l[k] = k < dim ? lo[k] : 0.0;
        u[k] = k < dim ? hi[k] : 0.0;
    }
}

#define DHIPCHK(x) do { if ((x) != hipSuccess) return PHD_E_HIP; } while (0)
#define DLAUNCHCHK() do { if (hipPeekAtLastError() != hipSuccess) { (void)hipGetLastError(); return PHD_E_HIP; } } while (0)

int dom_check(const phd_handle* h, const phd_cells* c)
{
    if (!h || !c || (c->dim != 2 && c->dim != 3) || c->n_cells < 0) return PHD_E_ARG;
    if (c->n_cells > 0 && !c->tag) return PHD_E_ARG;
    return 0;
}

} // namespace

extern "C" {

int phd_move_generators(phd_handle* h, const phd_cells* c, double dt)
{
    if (dom_check(h, c)) return PHD_E_ARG;
    if (c->n_cells == 0) return 0;
    for (int k = 0; k < c->dim; k++)
        if (!c->x[k] || !c->w[k]) return PHD_E_ARG;
    if (c->dim == 2) k_move_generators<2><<<xcd_grid(c->n_cells), PHD_BLOCK, 0, h->stream>>>(*c, dt);
    else k_move_generators<3><<<xcd_grid(c->n_cells), PHD_BLOCK, 0, h->stream>>>(*c, dt);
    DLAUNCHCHK();
    return 0;
}

int phd_migrate_periodic(phd_handle* h, const phd_cells* c, const double* lo, const double* hi)
{
    if (dom_check(h, c) || !lo || !hi) return PHD_E_ARG;
    if (c->n_cells == 0) return 0;
    for (int k = 0; k < c->dim; k++)
        if (!c->x[k]) return PHD_E_ARG;
    double l[3], u[3], L[3];
    box3(c->dim, lo, hi, l, u);
    for (int k = 0; k < 3; k++) L[k] = u[k] - l[k];     // DomainManager.translate
    if (c->dim == 2)
        k_migrate_periodic<2><<<xcd_grid(c->n_cells), PHD_BLOCK, 0, h->stream>>>(*c, l[0], l[1], l[2], u[0], u[1], u[2], L[0], L[1], L[2]);
    else
        k_migrate_periodic<3><<<xcd_grid(c->n_cells), PHD_BLOCK, 0, h->stream>>>(*c, l[0], l[1], l[2], u[0], u[1], u[2], L[0], L[1], L[2]);
    DLAUNCHCHK();
    return 0;
}

int phd_first_outside(phd_handle* h, const phd_cells* c, const double* lo, const double* hi,
                      int64_t* first)
{
    if (dom_check(h, c) || !lo || !hi || !first) return PHD_E_ARG;
    *first = -1;
    if (c->n_cells == 0) return 0;
    for (int k = 0; k < c->dim; k++)
        if (!c->x[k]) return PHD_E_ARG;
    double l[3], u[3];
    box3(c->dim, lo, hi, l, u);
    unsigned long long* slot = h->d_err + PHD_ERR_SLOT_OUTSIDE;
    k_fill_u64<<<1, 64, 0, h->stream>>>(slot, 1, ~0ull);
    if (c->dim == 2)
        k_first_outside<2><<<xcd_grid(c->n_cells), PHD_BLOCK, 0, h->stream>>>(*c, l[0], l[1], l[2], u[0], u[1], u[2], slot);
    else
        k_first_outside<3><<<xcd_grid(c->n_cells), PHD_BLOCK, 0, h->stream>>>(*c, l[0], l[1], l[2], u[0], u[1], u[2], slot);
    DLAUNCHCHK();
    DHIPCHK(hipMemcpyAsync(h->h_err + PHD_ERR_SLOT_OUTSIDE, slot, sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, h->stream));
    DHIPCHK(hipStreamSynchronize(h->stream));
    const unsigned long long v = h->h_err[PHD_ERR_SLOT_OUTSIDE];
    *first = v == ~0ull ? -1 : (int64_t)v;
    return 0;
}

int phd_reflect_fields(phd_handle* h, const phd_cells* c, const double* lo, const double* hi)
{
    if (dom_check(h, c) || !lo || !hi) return PHD_E_ARG;
    if (c->n_cells == 0) return 0;
    if (!c->type) return PHD_E_ARG;
    for (int k = 0; k < c->dim; k++)
        if (!c->x[k] || !c->prim[1 + k] || !c->cons[1 + k]) return PHD_E_ARG;
    double l[3], u[3];
    box3(c->dim, lo, hi, l, u);
    if (c->dim == 2)
        k_reflect_fields<2><<<xcd_grid(c->n_cells), PHD_BLOCK, 0, h->stream>>>(*c, l[0], l[1], l[2], u[0], u[1], u[2]);
    else
        k_reflect_fields<3><<<xcd_grid(c->n_cells), PHD_BLOCK, 0, h->stream>>>(*c, l[0], l[1], l[2], u[0], u[1], u[2]);
    DLAUNCHCHK();
    return 0;
}

int phd_update_ghost_fields(phd_handle* h, const phd_cells* c, void* const* fields,
                            const int32_t* elem_size, int32_t n_fields)
{
    if (dom_check(h, c) || n_fields < 0 || (n_fields > 0 && (!fields || !elem_size)))
        return PHD_E_ARG;
    if (c->n_cells == 0 || n_fields == 0) return 0;
    for (int f = 0; f < n_fields; f++) {
        if (!fields[f]) return PHD_E_ARG;
        if (elem_size[f] != 1 && elem_size[f] != 4 && elem_size[f] != 8) return PHD_E_ARG;
    }
    if (!c->map) return PHD_E_ARG;
    for (int f0 = 0; f0 < n_fields; f0 += GF_MAX) {
        GhostFields gf;
        gf.n = n_fields - f0 < GF_MAX ? n_fields - f0 : GF_MAX;
        for (int f = 0; f < GF_MAX; f++) {
            gf.p[f] = f < gf.n ? fields[f0 + f] : nullptr;
            gf.size[f] = f < gf.n ? elem_size[f0 + f] : 0;
        }
        k_copy_ghost_fields<<<xcd_grid(c->n_cells), PHD_BLOCK, 0, h->stream>>>(c->n_cells, c->tag, c->map, gf);
        DLAUNCHCHK();
    }
    return 0;
}

} // extern "C"
