// voronoi.cpp -- host Voronoi mesh builder (see include/phd_mesh.h).
//
// Each real cell starts as the domain box (reflective) or as the cube of
// half-width L/2 about its generator (periodic) and is clipped by the
// bisector planes of nearby generators, taken in order of distance from a
// uniform bin grid, until the security criterion holds: every generator not
// yet used is farther than twice the cell's largest vertex distance.
//
// Reflective walls need no ghost candidates: for a point inside the box the
// mirror image of any generator is never closer than the generator itself,
// so the box-clipped cell of the real generators is exact, and a wall facet
// of cell i is its face with i's own mirror image (boundary.pyx:96-140
// creates one mirror per wall crossing).  Periodic images are found by
// wrapping bin indices (boundary.pyx:378-398).
//
// Cells are built in parallel (OpenMP); faces are then numbered serially in
// ascending pair_i, so the output does not depend on the thread count.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <vector>
#include <omp.h>
#include "phd_mesh.h"

namespace {

// neighbor key: real index plus image code.  code 0 = the real generator
// itself; reflective: 1 + 2*axis + side (mirror across that wall); periodic:
// 1 + index of the shift in {-1,0,1}^dim (center excluded).
struct NKey {
    int64_t k;
    int32_t code;
};

struct FaceRec {
    int64_t k;        // neighbor real index
    int32_t code;     // image code (0 = real neighbor)
    double area;
    double com[3];    // absolute
    double q[3];      // neighbor position - generator (absolute offset)
    int32_t r0, rn;   // the face's ring: Work::ring[r0 .. r0 + rn)
};

struct CellOut {
    double volume;
    double dcom[3];
};

struct Grid {
    int dim;
    int nb[3];
    double lo[3], hi[3], L[3], bs[3];
    std::vector<int64_t> start;   // bin -> first index into order
    std::vector<int64_t> order;   // point ids sorted by bin
    int bin_of(const double* p, int a) const
    {
        int b = (int)std::floor((p[a] - lo[a]) / bs[a]);
        return std::min(std::max(b, 0), nb[a] - 1);
    }
};

struct Cand {
    double d2;
    double q[3];
    int64_t k;
    int32_t code;
};

// ---- 3D convex polyhedron ----------------------------------------------
struct Poly3 {
    std::vector<double> v;          // 3*nv, relative to the generator
    std::vector<int> fstart, fcnt;  // face cycles (outward CCW)
    std::vector<int> fv;            // flat vertex ids
    std::vector<NKey> ftag;
    std::vector<double> fq;         // 3 per face: plane vector q
    int nv() const { return (int)v.size() / 3; }
    int nf() const { return (int)fcnt.size(); }
};

void box3(Poly3& P, const double* a, const double* b, const NKey* tags, const double* qs)
{
    // a = lower corner, b = upper corner (relative).  Faces: -x,+x,-y,+y,-z,+z
    P.v.clear(); P.fstart.clear(); P.fcnt.clear(); P.fv.clear(); P.ftag.clear(); P.fq.clear();
    for (int c = 0; c < 8; c++) {
        P.v.push_back((c & 1) ? b[0] : a[0]);
        P.v.push_back((c & 2) ? b[1] : a[1]);
        P.v.push_back((c & 4) ? b[2] : a[2]);
    }
    // outward CCW cycles
    static const int F[6][4] = {{0, 4, 6, 2}, {1, 3, 7, 5}, {0, 1, 5, 4},
                                {2, 6, 7, 3}, {0, 2, 3, 1}, {4, 5, 7, 6}};
    for (int f = 0; f < 6; f++) {
        P.fstart.push_back((int)P.fv.size());
        P.fcnt.push_back(4);
        for (int t = 0; t < 4; t++) P.fv.push_back(F[f][t]);
        P.ftag.push_back(tags[f]);
        for (int t = 0; t < 3; t++) P.fq.push_back(qs[3 * f + t]);
    }
}

// Clip by {p : p.q <= |q|^2/2}.  Returns false if the plane does not cut.
bool clip3(Poly3& P, const double* q, NKey tag, Poly3& out, std::vector<double>& s,
           std::vector<int>& remap, std::vector<int>& ekey_a, std::vector<int>& ekey_b,
           std::vector<int>& ekey_v, std::vector<int>& nxt)
{
    const double qq = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
    const double d = 0.5 * qq, eps = 1e-13 * qq;
    const int nv = P.nv();
    s.resize(nv);
    bool any_out = false;
    for (int a = 0; a < nv; a++) {
        const double* p = &P.v[3 * a];
        s[a] = p[0] * q[0] + p[1] * q[1] + p[2] * q[2] - d;
        if (s[a] > eps) any_out = true;
    }
    if (!any_out) return false;
    out.v.clear(); out.fstart.clear(); out.fcnt.clear(); out.fv.clear(); out.ftag.clear(); out.fq.clear();
    remap.assign(nv, -1);
    for (int a = 0; a < nv; a++)
        if (s[a] <= eps) {
            remap[a] = out.nv();
            out.v.insert(out.v.end(), &P.v[3 * a], &P.v[3 * a] + 3);
        }
    ekey_a.clear(); ekey_b.clear(); ekey_v.clear();
    auto crossing = [&](int a, int b) -> int {
        int lo_ = std::min(a, b), hi_ = std::max(a, b);
        for (size_t t = 0; t < ekey_a.size(); t++)
            if (ekey_a[t] == lo_ && ekey_b[t] == hi_) return ekey_v[t];
        // point on the edge where s = 0 (s[lo_] and s[hi_] differ in sign)
        double sa = s[lo_], sb = s[hi_];
        double t = sa / (sa - sb);
        const double* pa = &P.v[3 * lo_];
        const double* pb = &P.v[3 * hi_];
        int id = out.nv();
        out.v.push_back(pa[0] + t * (pb[0] - pa[0]));
        out.v.push_back(pa[1] + t * (pb[1] - pa[1]));
        out.v.push_back(pa[2] + t * (pb[2] - pa[2]));
        ekey_a.push_back(lo_); ekey_b.push_back(hi_); ekey_v.push_back(id);
        return id;
    };
    // next-vertex map of the cap cycle (entry -> exit), keyed by new vertex id
    nxt.clear();
    for (int f = 0; f < P.nf(); f++) {
        const int st = P.fstart[f], cnt = P.fcnt[f];
        const int start_out = (int)out.fv.size();
        int exitp = -1, entryp = -1;
        for (int t = 0; t < cnt; t++) {
            int u = P.fv[st + t], w = P.fv[st + (t + 1) % cnt];
            bool iu = s[u] <= eps, iw = s[w] <= eps;
            if (iu) out.fv.push_back(remap[u]);
            if (iu != iw) {
                int x = crossing(u, w);
                out.fv.push_back(x);
                if (iu) exitp = x; else entryp = x;
            }
        }
        int c = (int)out.fv.size() - start_out;
        if (c >= 3) {
            out.fstart.push_back(start_out);
            out.fcnt.push_back(c);
            out.ftag.push_back(P.ftag[f]);
            out.fq.insert(out.fq.end(), &P.fq[3 * f], &P.fq[3 * f] + 3);
        } else {
            out.fv.resize(start_out);
        }
        if (exitp >= 0 && entryp >= 0) {
            if ((int)nxt.size() < out.nv()) nxt.resize(out.nv() + 64, -1);
            nxt[entryp] = exitp;
        }
    }
    // cap face: chain entry -> exit
    if ((int)nxt.size() < out.nv()) nxt.resize(out.nv(), -1);
    int startv = -1;
    for (int a = 0; a < out.nv(); a++) if (nxt[a] >= 0) { startv = a; break; }
    if (startv >= 0) {
        const int st = (int)out.fv.size();
        int v = startv, guard = 0;
        do {
            out.fv.push_back(v);
            v = nxt[v];
            if (++guard > 4096 || v < 0) { out.fv.resize(st); v = -2; break; }
        } while (v != startv);
        int c = (int)out.fv.size() - st;
        if (v != -2 && c >= 3) {
            out.fstart.push_back(st);
            out.fcnt.push_back(c);
            out.ftag.push_back(tag);
            out.fq.insert(out.fq.end(), q, q + 3);
        } else {
            out.fv.resize(st);
        }
    }
    return true;
}

double rmax2_3(const Poly3& P)
{
    double r = 0;
    for (int a = 0; a < P.nv(); a++) {
        const double* p = &P.v[3 * a];
        r = std::max(r, p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    }
    return r;
}

// ---- 2D convex polygon ---------------------------------------------------
struct Poly2 {
    std::vector<double> v;   // 2*nv, CCW
    std::vector<NKey> etag;  // tag of edge starting at vertex a
    std::vector<double> eq;  // 2 per edge
    int nv() const { return (int)v.size() / 2; }
};

bool clip2(Poly2& P, const double* q, NKey tag, Poly2& out, std::vector<double>& s)
{
    const double qq = q[0] * q[0] + q[1] * q[1];
    const double d = 0.5 * qq, eps = 1e-13 * qq;
    const int nv = P.nv();
    s.resize(nv);
    bool any_out = false;
    for (int a = 0; a < nv; a++) {
        s[a] = P.v[2 * a] * q[0] + P.v[2 * a + 1] * q[1] - d;
        if (s[a] > eps) any_out = true;
    }
    if (!any_out) return false;
    out.v.clear(); out.etag.clear(); out.eq.clear();
    for (int a = 0; a < nv; a++) {
        int b = (a + 1) % nv;
        bool ia = s[a] <= eps, ib = s[b] <= eps;
        if (ia) {
            out.v.push_back(P.v[2 * a]); out.v.push_back(P.v[2 * a + 1]);
            out.etag.push_back(P.etag[a]);
            out.eq.push_back(P.eq[2 * a]); out.eq.push_back(P.eq[2 * a + 1]);
        }
        if (ia != ib) {
            double t = s[a] / (s[a] - s[b]);
            double x = P.v[2 * a] + t * (P.v[2 * b] - P.v[2 * a]);
            double y = P.v[2 * a + 1] + t * (P.v[2 * b + 1] - P.v[2 * a + 1]);
            out.v.push_back(x); out.v.push_back(y);
            if (ia) { out.etag.push_back(tag); out.eq.push_back(q[0]); out.eq.push_back(q[1]); }
            else { out.etag.push_back(P.etag[a]); out.eq.push_back(P.eq[2 * a]); out.eq.push_back(P.eq[2 * a + 1]); }
        }
    }
    return true;
}

double rmax2_2(const Poly2& P)
{
    double r = 0;
    for (int a = 0; a < P.nv(); a++) {
        double x = P.v[2 * a], y = P.v[2 * a + 1];
        r = std::max(r, x * x + y * y);
    }
    return r;
}

} // namespace

struct phdm_mesh {
    int dim;
    bool has_geometry;
    int64_t n_real, n_ghost, n_faces;
    std::vector<double> gx[3];
    std::vector<int32_t> gmap;
    std::vector<int8_t> gcode;
    std::vector<int32_t> ring_off, ring_k;
    std::vector<int8_t> ring_code;
    std::vector<double> volume, dcom[3];
    std::vector<double> area, com[3], normal[3];
    std::vector<int32_t> pi, pj;
    std::vector<int32_t> nbr_off, nbr_face;
};

namespace {

// gather candidates of shell r (Chebyshev distance exactly r in bins) around
// generator i; returns false when the shell lies entirely outside a
// non-periodic grid.
void shell_candidates(const Grid& g, const double* const* x, int64_t i, const int* cb, int r,
                      bool periodic, std::vector<Cand>& out)
{
    const int dim = g.dim;
    int lo_[3] = {0, 0, 0}, hi_[3] = {0, 0, 0};
    for (int a = 0; a < dim; a++) { lo_[a] = -r; hi_[a] = r; }
    double xi[3] = {x[0][i], x[1][i], dim == 3 ? x[2][i] : 0.0};
    for (int oz = lo_[2]; oz <= hi_[2]; oz++)
        for (int oy = lo_[1]; oy <= hi_[1]; oy++)
            for (int ox = lo_[0]; ox <= hi_[0]; ox++) {
                int o[3] = {ox, oy, oz};
                int cheb = std::max(std::abs(ox), std::max(std::abs(oy), std::abs(oz)));
                if (cheb != r) continue;
                int bb[3] = {0, 0, 0}, sh[3] = {0, 0, 0};
                bool skip = false;
                for (int a = 0; a < dim; a++) {
                    int b = cb[a] + o[a];
                    if (periodic) {
                        // wrap, remembering the image shift
                        int w = (int)std::floor((double)b / g.nb[a]);
                        sh[a] = w;
                        b -= w * g.nb[a];
                    } else if (b < 0 || b >= g.nb[a]) {
                        skip = true;
                    }
                    bb[a] = b;
                }
                if (skip) continue;
                int64_t bin = bb[0] + (int64_t)g.nb[0] * (bb[1] + (int64_t)g.nb[1] * bb[2]);
                for (int64_t t = g.start[bin]; t < g.start[bin + 1]; t++) {
                    int64_t k = g.order[t];
                    int32_t code = 0;
                    if (periodic) {
                        int idx = 0, mul = 1;
                        bool zero = true;
                        for (int a = 0; a < dim; a++) {
                            idx += (sh[a] + 1) * mul;
                            mul *= 3;
                            if (sh[a] != 0) zero = false;
                        }
                        if (!zero) {
                            if (std::abs(sh[0]) > 1 || std::abs(sh[1]) > 1 || std::abs(sh[2]) > 1) continue;
                            code = 1 + idx;
                        }
                    }
                    if (k == i && code == 0) continue;
                    Cand c;
                    c.k = k;
                    c.code = code;
                    c.d2 = 0;
                    for (int a = 0; a < 3; a++) c.q[a] = 0;
                    for (int a = 0; a < dim; a++) {
                        c.q[a] = x[a][k] + sh[a] * g.L[a] - xi[a];
                        c.d2 += c.q[a] * c.q[a];
                    }
                    out.push_back(c);
                }
            }
}

// distance from xi to the boundary of its (2r+1)^dim bin block: every
// generator outside the block is at least this far away
double covered_radius(const Grid& g, const double* xi, const int* cb, int r, bool periodic)
{
    double rc = 1e300;
    for (int a = 0; a < g.dim; a++) {
        double blo = g.lo[a] + (cb[a] - r) * g.bs[a];
        double bhi = g.lo[a] + (cb[a] + r + 1) * g.bs[a];
        double dlo = xi[a] - blo, dhi = bhi - xi[a];
        if (!periodic) {
            // beyond the box there is nothing to find
            if (cb[a] - r <= 0) dlo = 1e300;
            if (cb[a] + r + 1 >= g.nb[a]) dhi = 1e300;
        }
        rc = std::min(rc, std::min(dlo, dhi));
    }
    return rc;
}

struct Work {
    Poly3 p3a, p3b;
    Poly2 p2a, p2b;
    std::vector<double> s;
    std::vector<int> remap, ea, eb, ev, nxt;
    std::vector<Cand> cands;
    std::vector<NKey> ring;                            // rings of the emitted faces
    std::vector<std::pair<int64_t, int>> edges;        // directed edge -> face
};

// The face across edge (u -> v) of a face of P: the face holding (v -> u).
// W.edges holds every directed edge of P, sorted.  -1 if there is none.
int twin_face(const Work& W, int u, int v)
{
    const int64_t key = ((int64_t)v << 32) | (uint32_t)u;
    auto it = std::lower_bound(W.edges.begin(), W.edges.end(), std::make_pair(key, -1));
    if (it == W.edges.end() || it->first != key) return -1;
    return it->second;
}

// Build real cell i; appends its emitted faces (neighbor real index > i, or
// any image) to `faces` and their rings to W.ring.  The ring of a face is the
// Delaunay star of the edge (i, j): the keys of the generators whose bisector
// planes bound the face, in the order of the face's edges.  In 2D it is the
// two neighbours meeting the face at its end points (vertex t is the
// circumcentre of i, j, ring[t]); in 3D ring[t] is the face across the
// face's edge t, so vertex t is the circumcentre of i, j, ring[t],
// ring[t + 1 mod m].  geometry = false skips the cell's volume and centroid.
// Returns false if the cell could not be closed.
bool build_cell(const Grid& g, const double* const* x, int64_t i, bool periodic, Work& W,
                std::vector<FaceRec>& faces, CellOut& co, bool geometry)
{
    const int dim = g.dim;
    double xi[3] = {x[0][i], x[1][i], dim == 3 ? x[2][i] : 0.0};
    int cb[3] = {0, 0, 0};
    for (int a = 0; a < dim; a++) cb[a] = g.bin_of(xi, a);

    // initial cell and its tags
    NKey tags[6];
    double qs[18];
    double a_[3], b_[3];
    for (int a = 0; a < dim; a++) {
        if (periodic) {
            a_[a] = -0.5 * g.L[a];
            b_[a] = 0.5 * g.L[a];
        } else {
            a_[a] = g.lo[a] - xi[a];
            b_[a] = g.hi[a] - xi[a];
        }
        for (int side = 0; side < 2; side++) {
            int t = 2 * a + side;
            double* q = &qs[3 * t];
            q[0] = q[1] = q[2] = 0;
            if (periodic) {
                // image of i itself shifted by -/+ L along a
                q[a] = side ? g.L[a] : -g.L[a];
                int idx = 0, mul = 1;
                for (int c = 0; c < dim; c++) { idx += ((c == a ? (side ? 1 : -1) : 0) + 1) * mul; mul *= 3; }
                tags[t] = NKey{i, 1 + idx};
            } else {
                // mirror image of i across the wall
                q[a] = side ? 2.0 * (g.hi[a] - xi[a]) : 2.0 * (g.lo[a] - xi[a]);
                tags[t] = NKey{i, 1 + 2 * a + side};
            }
        }
    }

    bool use_a = true;
    double r2;
    if (dim == 3) {
        box3(W.p3a, a_, b_, tags, qs);
        r2 = rmax2_3(W.p3a);
    } else {
        Poly2& P = W.p2a;
        P.v = {a_[0], a_[1], b_[0], a_[1], b_[0], b_[1], a_[0], b_[1]};
        // edges: bottom (-y), right (+x), top (+y), left (-x)
        P.etag = {tags[2], tags[1], tags[3], tags[0]};
        P.eq = {qs[6], qs[7], qs[3], qs[4], qs[9], qs[10], qs[0], qs[1]};
        r2 = rmax2_2(P);
    }

    for (int r = 0;; r++) {
        W.cands.clear();
        shell_candidates(g, x, i, cb, r, periodic, W.cands);
        std::sort(W.cands.begin(), W.cands.end(),
                  [](const Cand& A, const Cand& B) {
                      if (A.d2 != B.d2) return A.d2 < B.d2;
                      if (A.k != B.k) return A.k < B.k;
                      return A.code < B.code;
                  });
        for (const Cand& c : W.cands) {
            if (c.d2 >= 4.0 * r2) break;   // plane at |q|/2 >= rmax cannot cut
            bool cut;
            NKey tag{c.k, c.code};
            if (dim == 3) {
                Poly3& src = use_a ? W.p3a : W.p3b;
                Poly3& dst = use_a ? W.p3b : W.p3a;
                cut = clip3(src, c.q, tag, dst, W.s, W.remap, W.ea, W.eb, W.ev, W.nxt);
                if (cut) { use_a = !use_a; r2 = rmax2_3(use_a ? W.p3a : W.p3b); }
            } else {
                Poly2& src = use_a ? W.p2a : W.p2b;
                Poly2& dst = use_a ? W.p2b : W.p2a;
                cut = clip2(src, c.q, tag, dst, W.s);
                if (cut) { use_a = !use_a; r2 = rmax2_2(use_a ? W.p2a : W.p2b); }
            }
        }
        double rc = covered_radius(g, xi, cb, r, periodic);
        if (4.0 * r2 <= rc * rc) break;
        bool all_out = true;
        for (int a = 0; a < dim; a++)
            if (periodic || cb[a] - r > 0 || cb[a] + r + 1 < g.nb[a]) all_out = false;
        if (all_out) break;
        if (r > 64) return false;
    }

    // extract faces, volume and centroid (reference formulas)
    double vol = 0, cen[3] = {0, 0, 0};
    if (dim == 2) {
        const Poly2& P = use_a ? W.p2a : W.p2b;
        const int nv = P.nv();
        for (int a = 0; a < nv; a++) {
            int b = (a + 1) % nv;
            double x1 = P.v[2 * a], y1 = P.v[2 * a + 1], x2 = P.v[2 * b], y2 = P.v[2 * b + 1];
            const double* q = &P.eq[2 * a];
            double h = std::sqrt(q[0] * q[0] + q[1] * q[1]);
            double xe = x2 - x1, ye = y2 - y1;
            double area = std::sqrt(xe * xe + ye * ye);
            if (area <= 1e-14 * h) continue;
            double fx = 0.5 * (x1 + x2), fy = 0.5 * (y1 + y2);
            if (geometry) {
                vol += 0.25 * area * h;
                cen[0] += 0.25 * area * h * (2.0 * fx / 3.0);
                cen[1] += 0.25 * area * h * (2.0 * fy / 3.0);
            }
            const NKey& t = P.etag[a];
            if (t.code != 0 || t.k > i) {
                FaceRec fr;
                fr.k = t.k; fr.code = t.code; fr.area = area;
                fr.com[0] = xi[0] + fx; fr.com[1] = xi[1] + fy; fr.com[2] = 0;
                fr.q[0] = q[0]; fr.q[1] = q[1]; fr.q[2] = 0;
                // end points: a on edge a-1, a+1 on edge a+1
                fr.r0 = (int32_t)W.ring.size(); fr.rn = 2;
                W.ring.push_back(P.etag[(a + nv - 1) % nv]);
                W.ring.push_back(P.etag[b]);
                faces.push_back(fr);
            }
        }
        co.volume = vol;
        co.dcom[0] = cen[0] / vol;
        co.dcom[1] = cen[1] / vol;
        co.dcom[2] = 0;
    } else {
        const Poly3& P = use_a ? W.p3a : W.p3b;
        W.edges.clear();
        for (int f = 0; f < P.nf(); f++)
            for (int t = 0; t < P.fcnt[f]; t++) {
                const int u = P.fv[P.fstart[f] + t], v = P.fv[P.fstart[f] + (t + 1) % P.fcnt[f]];
                W.edges.push_back(std::make_pair(((int64_t)u << 32) | (uint32_t)v, f));
            }
        std::sort(W.edges.begin(), W.edges.end());
        for (int f = 0; f < P.nf(); f++) {
            const int st = P.fstart[f], nvtx = P.fcnt[f];
            double c[3] = {0, 0, 0};
            for (int t = 0; t < nvtx; t++)
                for (int a = 0; a < 3; a++) c[a] += P.v[3 * P.fv[st + t] + a];
            for (int a = 0; a < 3; a++) c[a] *= 1.0 / nvtx;
            double fc[3] = {0, 0, 0}, area1 = 0;
            const double third = 1.0 / 3.0;
            double v1[3];
            for (int a = 0; a < 3; a++) v1[a] = P.v[3 * P.fv[st + nvtx - 1] + a] - c[a];
            for (int t = 0; t < nvtx; t++) {
                double v2[3];
                for (int a = 0; a < 3; a++) v2[a] = P.v[3 * P.fv[st + t] + a] - c[a];
                double n3[3] = {v1[1] * v2[2] - v1[2] * v2[1], v1[2] * v2[0] - v1[0] * v2[2],
                                v1[0] * v2[1] - v1[1] * v2[0]};
                double a3 = std::sqrt(n3[0] * n3[0] + n3[1] * n3[1] + n3[2] * n3[2]);
                for (int a = 0; a < 3; a++) fc[a] += (c[a] + (v1[a] + v2[a]) * third) * a3;
                area1 += a3;
                for (int a = 0; a < 3; a++) v1[a] = v2[a];
            }
            const double* q = &P.fq[3 * f];
            double qn = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
            if (area1 <= 1e-20 * qn * qn) continue;   // degenerate facet
            for (int a = 0; a < 3; a++) fc[a] *= 1.0 / area1;
            const NKey& tg = P.ftag[f];
            if (tg.code != 0 || tg.k > i) {
                FaceRec fr;
                fr.k = tg.k; fr.code = tg.code; fr.area = 0.5 * area1;
                for (int a = 0; a < 3; a++) { fr.com[a] = xi[a] + fc[a]; fr.q[a] = q[a]; }
                fr.r0 = (int32_t)W.ring.size(); fr.rn = nvtx;
                for (int t = 0; t < nvtx; t++) {
                    const int g_ = twin_face(W, P.fv[st + t], P.fv[st + (t + 1) % nvtx]);
                    if (g_ < 0) return false;   // open polyhedron
                    W.ring.push_back(P.ftag[g_]);
                }
                faces.push_back(fr);
            }
            if (!geometry) continue;
            // tetrahedra (facet centroid, v1, v2, generator) -- tess3.cpp:537-549
            const double fourth = 0.25;
            double cv[3] = {-fc[0], -fc[1], -fc[2]};
            for (int a = 0; a < 3; a++) v1[a] = P.v[3 * P.fv[st + nvtx - 1] + a] - fc[a];
            for (int t = 0; t < nvtx; t++) {
                double v2[3];
                for (int a = 0; a < 3; a++) v2[a] = P.v[3 * P.fv[st + t] + a] - fc[a];
                double cr[3] = {v1[1] * v2[2] - v1[2] * v2[1], v1[2] * v2[0] - v1[0] * v2[2],
                                v1[0] * v2[1] - v1[1] * v2[0]};
                double v4 = std::fabs(cr[0] * cv[0] + cr[1] * cv[1] + cr[2] * cv[2]);
                vol += v4;
                for (int a = 0; a < 3; a++) cen[a] += (fc[a] + (v1[a] + v2[a] + cv[a]) * fourth) * v4;
                for (int a = 0; a < 3; a++) v1[a] = v2[a];
            }
        }
        for (int a = 0; a < 3; a++) co.dcom[a] = cen[a] / vol;
        co.volume = vol / 6.0;
    }
    return true;
}

// Position of the image `code` of generator k along axis a, as the
// reference's ghost creation places it: a periodic shift by (index - 1) *
// translate (boundary.pyx:378-380), a reflective mirror x - 2 (x - edge)
// (boundary.pyx:103-104, 125-126).
double image_coord(const Grid& g, bool periodic, double xk, int a, int32_t code)
{
    if (code == 0) return xk;
    if (periodic) {
        int idx = code - 1;
        for (int b = 0; b < a; b++) idx /= 3;
        return xk + (double)(idx % 3 - 1) * g.L[a];
    }
    const int axis = (code - 1) >> 1, side = (code - 1) & 1;
    if (axis != a) return xk;
    const double edge = side ? g.hi[a] : g.lo[a];
    return xk - 2 * (xk - edge);
}

phdm_mesh* build_mesh(int dim, int64_t n, const double* const* x, const double* lo,
                      const double* hi, int boundary, int n_threads, int* status, bool geometry)
{
    if (status) *status = 0;
    if ((dim != 2 && dim != 3) || n < 1 || !x || !lo || !hi ||
        (boundary != PHDM_REFLECTIVE && boundary != PHDM_PERIODIC)) {
        if (status) *status = -1;
        return nullptr;
    }
    const bool periodic = boundary == PHDM_PERIODIC;
    for (int64_t i = 0; i < n; i++)
        for (int a = 0; a < dim; a++)
            if (!(periodic ? (x[a][i] >= lo[a] && x[a][i] < hi[a])
                           : (x[a][i] > lo[a] && x[a][i] < hi[a]))) {
                if (status) *status = -2;   // generator outside the box
                return nullptr;
            }
    Grid g;
    g.dim = dim;
    double vol = 1;
    for (int a = 0; a < dim; a++) {
        g.lo[a] = lo[a]; g.hi[a] = hi[a]; g.L[a] = hi[a] - lo[a];
        vol *= g.L[a];
    }
    double spacing = std::pow(vol / (double)n, 1.0 / dim);
    for (int a = 0; a < 3; a++) {
        if (a < dim) {
            g.nb[a] = std::max(1, (int)std::floor(g.L[a] / spacing));
            if (periodic) g.nb[a] = std::max(g.nb[a], 3);
            g.bs[a] = g.L[a] / g.nb[a];
        } else {
            g.nb[a] = 1; g.bs[a] = 1; g.lo[a] = 0; g.hi[a] = 1; g.L[a] = 1;
        }
    }
    int64_t nbins = (int64_t)g.nb[0] * g.nb[1] * g.nb[2];
    std::vector<int64_t> binid(n);
    g.start.assign(nbins + 1, 0);
    for (int64_t i = 0; i < n; i++) {
        double p[3] = {x[0][i], x[1][i], dim == 3 ? x[2][i] : 0.0};
        int64_t b = 0, mul = 1;
        for (int a = 0; a < dim; a++) { b += g.bin_of(p, a) * mul; mul *= g.nb[a]; }
        binid[i] = b;
        g.start[b + 1]++;
    }
    for (int64_t b = 0; b < nbins; b++) g.start[b + 1] += g.start[b];
    g.order.resize(n);
    {
        std::vector<int64_t> fill(g.start.begin(), g.start.end() - 1);
        for (int64_t i = 0; i < n; i++) g.order[fill[binid[i]]++] = i;
    }

    if (n_threads > 0) omp_set_num_threads(n_threads);
    const int64_t CH = 2048;
    const int64_t nch = (n + CH - 1) / CH;
    std::vector<std::vector<FaceRec>> chunk_faces(nch);
    std::vector<std::vector<NKey>> chunk_ring(nch);
    std::vector<std::vector<int32_t>> chunk_cnt(nch);
    std::vector<CellOut> cells(n);
    int fail = 0;
#pragma omp parallel
    {
        Work W;
        std::vector<FaceRec> tmp;
#pragma omp for schedule(dynamic, 1)
        for (int64_t ch = 0; ch < nch; ch++) {
            int64_t b = ch * CH, e = std::min(n, b + CH);
            std::vector<FaceRec>& out = chunk_faces[ch];
            std::vector<NKey>& rout = chunk_ring[ch];
            chunk_cnt[ch].resize(e - b);
            for (int64_t i = b; i < e; i++) {
                tmp.clear();
                W.ring.clear();
                if (!build_cell(g, x, i, periodic, W, tmp, cells[i], geometry)) {
#pragma omp atomic write
                    fail = 1;
                }
                // reals first (ascending id), then images (by image id, then code)
                std::sort(tmp.begin(), tmp.end(), [](const FaceRec& A, const FaceRec& B) {
                    bool ga = A.code != 0, gb = B.code != 0;
                    if (ga != gb) return !ga;
                    if (!ga) return A.k < B.k;
                    if (A.k != B.k) return A.k < B.k;
                    return A.code < B.code;
                });
                chunk_cnt[ch][i - b] = (int32_t)tmp.size();
                for (FaceRec& fr : tmp) {
                    const int32_t r0 = (int32_t)rout.size();
                    rout.insert(rout.end(), W.ring.begin() + fr.r0, W.ring.begin() + fr.r0 + fr.rn);
                    fr.r0 = r0;
                }
                out.insert(out.end(), tmp.begin(), tmp.end());
            }
        }
    }
    if (fail) {
        if (status) *status = -3;
        return nullptr;
    }

    phdm_mesh* M = new phdm_mesh();
    M->dim = dim;
    M->has_geometry = geometry;
    M->n_real = n;
    int64_t nf = 0, nring = 0;
    for (auto& v : chunk_faces) nf += (int64_t)v.size();
    for (auto& v : chunk_ring) nring += (int64_t)v.size();
    if (nf > INT32_MAX / 2 || nring > INT32_MAX) {
        delete M;
        if (status) *status = -4;
        return nullptr;
    }
    M->n_faces = nf;
    M->pi.resize(nf);
    M->pj.resize(nf);
    M->ring_off.resize(nf + 1);
    M->ring_k.resize(nring);
    M->ring_code.resize(nring);
    M->ring_off[0] = 0;
    if (geometry) {
        M->area.resize(nf);
        for (int a = 0; a < dim; a++) { M->com[a].resize(nf); M->normal[a].resize(nf); }
    }
    // ghost numbering in order of first appearance along the face list
    std::unordered_map<uint64_t, int32_t> gid;
    gid.reserve(1024);
    int64_t fcount = 0;
    for (int64_t ch = 0; ch < nch; ch++) {
        int64_t b = ch * CH;
        size_t t = 0;
        for (size_t c = 0; c < chunk_cnt[ch].size(); c++) {
            int64_t i = b + (int64_t)c;
            for (int32_t u = 0; u < chunk_cnt[ch][c]; u++, t++) {
                const FaceRec& fr = chunk_faces[ch][t];
                int64_t j;
                if (fr.code == 0) {
                    j = fr.k;
                } else {
                    uint64_t key = (uint64_t)fr.k * 64u + (uint64_t)fr.code;
                    auto it = gid.find(key);
                    if (it == gid.end()) {
                        int32_t id = (int32_t)M->gmap.size();
                        gid.emplace(key, id);
                        M->gmap.push_back((int32_t)fr.k);
                        M->gcode.push_back((int8_t)fr.code);
                        for (int a = 0; a < dim; a++)
                            M->gx[a].push_back(image_coord(g, periodic, x[a][fr.k], a, fr.code));
                        j = n + id;
                    } else {
                        j = n + it->second;
                    }
                }
                M->pi[fcount] = (int32_t)i;
                M->pj[fcount] = (int32_t)j;
                const int32_t r0 = M->ring_off[fcount];
                for (int32_t r = 0; r < fr.rn; r++) {
                    M->ring_k[r0 + r] = (int32_t)chunk_ring[ch][fr.r0 + r].k;
                    M->ring_code[r0 + r] = (int8_t)chunk_ring[ch][fr.r0 + r].code;
                }
                M->ring_off[fcount + 1] = r0 + fr.rn;
                if (geometry) {
                    M->area[fcount] = fr.area;
                    double qn = 0;
                    for (int a = 0; a < dim; a++) qn += fr.q[a] * fr.q[a];
                    qn = std::sqrt(qn);
                    for (int a = 0; a < dim; a++) {
                        M->com[a][fcount] = fr.com[a];
                        M->normal[a][fcount] = fr.q[a] / qn;
                    }
                }
                fcount++;
            }
        }
        std::vector<FaceRec>().swap(chunk_faces[ch]);
        std::vector<NKey>().swap(chunk_ring[ch]);
    }
    M->n_ghost = (int64_t)M->gmap.size();
    const int64_t na = n + M->n_ghost;
    if (geometry) {
        M->volume.resize(na);
        for (int a = 0; a < dim; a++) M->dcom[a].resize(na);
        for (int64_t i = 0; i < n; i++) {
            M->volume[i] = cells[i].volume;
            for (int a = 0; a < dim; a++) M->dcom[a][i] = cells[i].dcom[a];
        }
        for (int64_t gg = 0; gg < M->n_ghost; gg++) {
            int64_t k = M->gmap[gg];
            M->volume[n + gg] = cells[k].volume;
            for (int a = 0; a < dim; a++) M->dcom[a][n + gg] = cells[k].dcom[a];
        }
    }
    // per-cell face lists, ascending face id (tess.cpp:379-382)
    M->nbr_off.assign(na + 1, 0);
    for (int64_t f = 0; f < nf; f++) { M->nbr_off[M->pi[f] + 1]++; M->nbr_off[M->pj[f] + 1]++; }
    for (int64_t c = 0; c < na; c++) M->nbr_off[c + 1] += M->nbr_off[c];
    M->nbr_face.resize(2 * nf);
    {
        std::vector<int32_t> fill(M->nbr_off.begin(), M->nbr_off.end() - 1);
        for (int64_t f = 0; f < nf; f++) {
            M->nbr_face[fill[M->pi[f]]++] = (int32_t)f;
            M->nbr_face[fill[M->pj[f]]++] = (int32_t)f;
        }
    }
    return M;
}

} // namespace

extern "C" {

phdm_mesh* phdm_build(int dim, int64_t n, const double* const* x, const double* lo,
                      const double* hi, int boundary, int n_threads, int* status)
{
    return build_mesh(dim, n, x, lo, hi, boundary, n_threads, status, true);
}

phdm_mesh* phdm_build_topology(int dim, int64_t n, const double* const* x, const double* lo,
                               const double* hi, int boundary, int n_threads, int* status)
{
    return build_mesh(dim, n, x, lo, hi, boundary, n_threads, status, false);
}

int phdm_counts(const phdm_mesh* m, int64_t* n_ghost, int64_t* n_faces, int64_t* n_list)
{
    if (!m) return -1;
    if (n_ghost) *n_ghost = m->n_ghost;
    if (n_faces) *n_faces = m->n_faces;
    if (n_list) *n_list = 2 * m->n_faces;
    return 0;
}

int phdm_fill(const phdm_mesh* m, double* const* ghost_x, int32_t* ghost_map, double* volume,
              double* const* dcom, double* area, double* const* com, double* const* normal,
              int32_t* pair_i, int32_t* pair_j, int32_t* nbr_off, int32_t* nbr_face)
{
    if (!m) return -1;
    const int dim = m->dim;
    const int64_t ng = m->n_ghost, na = m->n_real + m->n_ghost, nf = m->n_faces;
    if (!m->has_geometry) {
        bool asked = volume || area;
        for (int a = 0; a < dim; a++)
            asked = asked || (dcom && dcom[a]) || (com && com[a]) || (normal && normal[a]);
        if (asked) return -2;   // a topology-only mesh holds no geometry
    }
    for (int a = 0; a < dim; a++) {
        if (ghost_x && ghost_x[a] && ng) memcpy(ghost_x[a], m->gx[a].data(), ng * sizeof(double));
        if (dcom && dcom[a]) memcpy(dcom[a], m->dcom[a].data(), na * sizeof(double));
        if (com && com[a] && nf) memcpy(com[a], m->com[a].data(), nf * sizeof(double));
        if (normal && normal[a] && nf) memcpy(normal[a], m->normal[a].data(), nf * sizeof(double));
    }
    if (ghost_map && ng) memcpy(ghost_map, m->gmap.data(), ng * sizeof(int32_t));
    if (volume) memcpy(volume, m->volume.data(), na * sizeof(double));
    if (area && nf) memcpy(area, m->area.data(), nf * sizeof(double));
    if (pair_i && nf) memcpy(pair_i, m->pi.data(), nf * sizeof(int32_t));
    if (pair_j && nf) memcpy(pair_j, m->pj.data(), nf * sizeof(int32_t));
    if (nbr_off) memcpy(nbr_off, m->nbr_off.data(), (na + 1) * sizeof(int32_t));
    if (nbr_face && nf) memcpy(nbr_face, m->nbr_face.data(), 2 * nf * sizeof(int32_t));
    return 0;
}

int phdm_ring_count(const phdm_mesh* m, int64_t* n_ring)
{
    if (!m || !n_ring) return -1;
    *n_ring = (int64_t)m->ring_k.size();
    return 0;
}

int phdm_fill_topology(const phdm_mesh* m, int8_t* ghost_code, int32_t* ring_off,
                       int32_t* ring_k, int8_t* ring_code)
{
    if (!m) return -1;
    const int64_t ng = m->n_ghost, nf = m->n_faces, nr = (int64_t)m->ring_k.size();
    if (ghost_code && ng) memcpy(ghost_code, m->gcode.data(), ng);
    if (ring_off) memcpy(ring_off, m->ring_off.data(), (nf + 1) * sizeof(int32_t));
    if (ring_k && nr) memcpy(ring_k, m->ring_k.data(), nr * sizeof(int32_t));
    if (ring_code && nr) memcpy(ring_code, m->ring_code.data(), nr);
    return 0;
}

void phdm_free(phdm_mesh* m) { delete m; }

} // extern "C"
