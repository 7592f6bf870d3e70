/*
 * phd_mesh.h -- C ABI of the host Voronoi mesh builder (libphd_mesh.so).
 *
 * Host-side replacement for the reference's CGAL tessellation +
 * geometry extraction (mesh/mesh.pyx:168-323, Tess2d::extract_geometry
 * tess.cpp:238-398, Tess3d::extract_geometry tess3.cpp:334-562) and for
 * serial ghost creation (domain/boundary.pyx:70-142 reflective,
 * 350-400 periodic).  It produces the arrays the flux pipeline consumes, in
 * the reference's conventions:
 *   - reals are cells 0..n_real-1, ghosts follow;
 *   - a face is emitted once, from its lower-id cell, so pair_i < pair_j and
 *     pair_i is real; faces are numbered in emission order (ascending pair_i);
 *   - normal = (x_j - x_i)/|x_j - x_i|;
 *   - 2D area = Voronoi edge length, com = edge midpoint (tess.cpp:326-338);
 *     3D area = 1/2 sum |v1 x v2| over a fan about the vertex mean, com =
 *     fan-triangle-area-weighted centroid (tess3.cpp:452-483, 515);
 *   - volume / dcom from the triangle (2D, tess.cpp:350-393) or tetrahedron
 *     (3D, tess3.cpp:537-559) decomposition;
 *   - per-cell face lists in ascending face id (tess.cpp:379-382);
 *   - ghost volume and dcom are copies of the image's (no reflection),
 *     as update_ghost_fields does (mesh.pyx:141-142, 323).
 * The builder clips each real cell's polyhedron by bisector planes of nearby
 * generators (a host-side algorithm; the reference uses a Delaunay
 * triangulation), so for the same generators the face set can differ from
 * CGAL's only on degenerate (co-spherical) configurations.
 */
#ifndef PHD_MESH_H
#define PHD_MESH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PHDM_REFLECTIVE 0
#define PHDM_PERIODIC   1

typedef struct phdm_mesh phdm_mesh;

/* Build the mesh of n_real generators x[k][i] in the box [lo, hi].
 * n_threads <= 0 uses every OpenMP thread.  On failure returns NULL and
 * sets *status (< 0). */
phdm_mesh* phdm_build(int dim, int64_t n_real, const double* const* x,
                      const double* lo, const double* hi, int boundary,
                      int n_threads, int* status);
int phdm_counts(const phdm_mesh* m, int64_t* n_ghost, int64_t* n_faces,
                int64_t* n_list);
/* Copy results into caller arrays (host memory):
 *   ghost_x[k][g], ghost_map[g]                  (g < n_ghost; every ghost is
 *                                                an exterior image)
 *   volume[c], dcom[k][c]                        (c < n_real + n_ghost)
 *   area[f], com[k][f], normal[k][f], pair_i[f], pair_j[f]
 *   nbr_off[c] (n_real + n_ghost + 1), nbr_face[n_list]                  */
int phdm_fill(const phdm_mesh* m, double* const* ghost_x, int32_t* ghost_map,
              double* volume, double* const* dcom, double* area,
              double* const* com, double* const* normal, int32_t* pair_i,
              int32_t* pair_j, int32_t* nbr_off, int32_t* nbr_face);

/* Topology-only build (the host half of GPU geometry extraction): the same
 * faces, pairs, ghosts, ghost positions and per-cell face lists as
 * phdm_build, and no face or cell geometry -- phdm_fill returns -2 if asked
 * for volume, dcom, area, com or normal of such a mesh.  The geometry is then
 * computed on the device by phd_extract_geometry (phd_flux.h) from the
 * rings below.  Replaces the CGAL triangulation that
 * Tess2d/Tess3d::extract_geometry (tess.cpp:238-398, tess3.cpp:334-562)
 * walk. */
phdm_mesh* phdm_build_topology(int dim, int64_t n_real, const double* const* x,
                               const double* lo, const double* hi, int boundary,
                               int n_threads, int* status);
/* Number of ring entries (sum of ring lengths over faces); any mesh. */
int phdm_ring_count(const phdm_mesh* m, int64_t* n_ring);
/* The Delaunay star of each face (i, j) = (pair_i, pair_j), any mesh:
 *   ring_off[f] .. ring_off[f+1]   (n_faces + 1 offsets)
 *   ring_k[r], ring_code[r]        the generator: real index and image code
 *   ghost_code[g]                  image code of each ghost (g < n_ghost)
 * Image code 0 is the generator itself; reflective 1 + 2*axis + side
 * (side 1 = upper wall), the mirror x - 2 (x - wall) (boundary.pyx:103-104);
 * periodic 1 + sum_a (s_a + 1) 3^a for the shift s in {-1, 0, 1}^dim, the
 * copy x + s translate (boundary.pyx:368-380).  ghost_x is that image of
 * ghost_map's generator.
 * In 2D a ring has 2 entries and vertex t (t = 0, 1) of the face is the
 * circumcentre of generators i, j and ring[t]; in 3D vertex t of the
 * face polygon (t < m, the ring length) is the circumcentre of i, j, ring[t]
 * and ring[(t + 1) mod m] -- the Delaunay tetrahedra around edge (i, j),
 * whose duals CGAL's cell circulator visits (tess3.cpp:423-440). */
int phdm_fill_topology(const phdm_mesh* m, int8_t* ghost_code, int32_t* ring_off,
                       int32_t* ring_k, int8_t* ring_code);
void phdm_free(phdm_mesh* m);

#ifdef __cplusplus
}
#endif
#endif /* PHD_MESH_H */
