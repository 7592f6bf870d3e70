/*
 * phd_flux.h -- C ABI of the MI355X (gfx950) Voronoi face-flux library
 * (libphd_flux.so).
 *
 * The library replaces the compiled inner loops of the reference's
 * reconstruction / riemann / mesh / equation_state plugins.  Each entry
 * point below names the reference method whose loop it replaces
 * (path:line in rickyfernandez/phd-code).
 *
 * Conventions
 *   - Every pointer in phd_cells / phd_faces and every array argument is a
 *     DEVICE pointer, borrowed for the duration of the call.  The library
 *     never frees or keeps them.
 *   - Work is enqueued on the handle's HIP stream and returns immediately,
 *     except the calls documented as synchronising (time step, error read).
 *   - Fields follow the reference's named groups:
 *       primitive    = density, velocity-x.., pressure
 *                      (utils/particle_creator.pyx:50-52)
 *       conservative = mass, momentum-x.., energy
 *                      (utils/particle_creator.pyx:53-55)
 *       gradient component g = n*dim + k is d prim_n / d x_k
 *                      (reconstruction/reconstruction.pyx:470-480); gradients
 *                      live in cell-record rows (phd_cell_record_ld).
 *   - Tags: 0 Real, 1 Ghost.  Types: 2 Exterior (utils/particle_tags.py).
 *   - Faces carry pair_i < pair_j with pair_i real, and per-cell face lists
 *     (nbr_off/nbr_face) in ascending face id (mesh/tess.cpp:363-382,
 *     mesh/tess3.cpp:509-534).
 *   - Return value: 0 = OK, < 0 = API misuse (PHD_E_*), > 0 = a numeric fault
 *     detected on the device (PHD_F_*), reported by phd_get_error().
 *   - Arithmetic is IEEE-754 double.  PHD_MODE_BITWISE evaluates every
 *     expression in the reference's operation order without FMA contraction.
 */
#ifndef PHD_FLUX_H
#define PHD_FLUX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PHD_ABI_VERSION 1

/* API misuse */
#define PHD_E_ARG      (-1)   /* bad argument (dim, sizes, NULL pointer) */
#define PHD_E_HIP      (-2)   /* a HIP runtime call failed */
#define PHD_E_MODE     (-3)   /* unsupported solver / limiter / mode */
#define PHD_E_SIZE     (-4)   /* input larger than a fixed-size scratch allows */
/* device-detected numeric faults (phd_get_error) */
#define PHD_F_NEIGHBORS 1     /* "ERROR: incorrect neighbors!"  reconstruction.pyx:650-652 */
#define PHD_F_MASS      2     /* "Mass less than zero in flux update"  mesh.pyx:546-548 */
#define PHD_F_ENERGY    3     /* "Energy less than zero in flux update" mesh.pyx:549-551 */
#define PHD_F_EXACT     4     /* "No convergence in Exact Riemann Solver" riemann.pyx:1088-1090 */

/* solvers (riemann/riemann.pyx: HLL 225, HLLC 507, Exact 684) */
#define PHD_SOLVER_HLL   0
#define PHD_SOLVER_HLLC  1
#define PHD_SOLVER_EXACT 2
/* limiters (reconstruction.pyx:406-411) */
#define PHD_LIMITER_AREPO 0
#define PHD_LIMITER_TESS  1
/* boundary conditions (domain/boundary.pyx: Reflective 70, Periodic 350) */
#define PHD_BOUNDARY_REFLECTIVE 0
#define PHD_BOUNDARY_PERIODIC   1
/* arithmetic modes */
#define PHD_MODE_BITWISE 0
/* reconstruction stages for phd_compute_states */
#define PHD_STAGE_SPATIAL  1
#define PHD_STAGE_TEMPORAL 2
/* reconstruction order */
#define PHD_ORDER_CONSTANT 0
#define PHD_ORDER_LINEAR   1

typedef struct phd_handle phd_handle;
typedef struct phd_plan phd_plan;

/* Cell (particle) arrays, n_cells = real + ghost, told apart by tag.  The
 * reference's containers put the ghosts after the reals, which the adjoint
 * requires; a batched union puts each problem's ghosts after its reals. */
typedef struct phd_cells {
    int32_t        dim;        /* 2 or 3 */
    int32_t        n_cells;    /* real + ghost */
    const int8_t*  tag;        /* "tag"  */
    const int8_t*  type;       /* "type" */
    const int32_t* map;        /* "map": image index of each ghost (may be NULL if no ghosts) */
    const double*  x[3];       /* "position-x".. */
    const double*  dcom[3];    /* "dcom-x".. */
    const double*  volume;     /* "volume" */
    double*        prim[5];    /* primitive group */
    double*        cons[5];    /* conservative group */
    double*        w[3];       /* "w-x".. generator velocity */
} phd_cells;

/* Face arrays (mesh.pyx:16-46) plus the per-cell face lists (mesh.pxd:34). */
typedef struct phd_faces {
    int32_t        n_faces;
    const double*  area;
    const double*  com[3];
    const double*  normal[3];
    double*        velocity[3];
    const int32_t* pair_i;
    const int32_t* pair_j;
    const int32_t* nbr_off;    /* n_cells + 1 offsets into nbr_face */
    const int32_t* nbr_face;   /* face ids, ascending within each cell */
} phd_faces;

/* Flags of phd_core_step */
#define PHD_CORE_XP_FRESH 1   /* the plan's packed cell rows already hold the
                                 current primitives (written by the previous
                                 phd_ghost_refresh_eos with this plan) */

/* ---- library / handle ---------------------------------------------------- */
int phd_abi_version(void);
const char* phd_error_string(int code);
/* Create a handle on `device`; allocates the handle's small device scratch
 * (error word, reduction cell) once. */
int phd_handle_create(phd_handle** out, int device);
int phd_handle_destroy(phd_handle* h);
/* Stream for all later work (a hipStream_t; NULL = default stream). */
int phd_set_stream(phd_handle* h, void* stream);
/* Arithmetic mode; PHD_MODE_BITWISE is the only mode of this build. */
int phd_set_mode(phd_handle* h, int mode);
/* Synchronising: read and clear the device error word.  *code is 0 or a
 * PHD_F_* value, *index the first failing face (flux update, Exact solver)
 * or cell (neighbors) in the reference's loop order. */
int phd_get_error(phd_handle* h, int32_t* code, int64_t* index);

/* ---- mesh plan ----------------------------------------------------------- */
/* A plan is the device-side acceleration structure of one frozen mesh: packed
 * face records (centroid, area, normal, generator distance), cell centroids
 * and volumes, per-cell face lists as {face, other cell, incoming bit}, the
 * list of ghost cells (the tags at creation), in 3D each 256-cell window's
 * halo of outside neighbours, and the per-step scratch of the fused kernels.
 * It is built on the device (6 ms at 2.1 M cells) from the mesh
 * arrays the reference's Mesh.build_geometry produces (mesh.pyx:255-323) and
 * must be recreated whenever the mesh is rebuilt.  The only allocating and
 * synchronising calls of the library besides the handle's. */
int phd_plan_create(phd_handle* h, const phd_cells* c, const phd_faces* f, phd_plan** out);
int phd_plan_destroy(phd_plan* p);
int64_t phd_plan_bytes(const phd_plan* p);
/* Doubles per cell-record row: primitive(dim+2) | gradient(dim(dim+2)) |
 * centroid(dim), padded to even (24 in 3D, 14 in 2D).  The gradient of cell
 * c, component g = n*dim+k, is rec[c*ld + dim + 2 + g]. */
int phd_cell_record_ld(int dim);
/* Refresh the plan's packed rows (position and primitive of each cell) from
 * the cells. */
int phd_pack_cells(phd_handle* h, phd_plan* p, const phd_cells* c);

/* ---- equation of state (equation_state/equation_state.pyx) --------------- */
/* IdealGas.conservative_from_primitive, equation_state.pyx:43-79 */
int phd_conservative_from_primitive(phd_handle* h, const phd_cells* c, double gamma);
/* IdealGas.primitive_from_conservative, equation_state.pyx:81-117 */
int phd_primitive_from_conservative(phd_handle* h, const phd_cells* c, double gamma);

/* ---- time step ---------------------------------------------------------- */
/* RiemannBase.compute_time_step, riemann.pyx:120-188.  Synchronising:
 * writes cfl * min over real cells to *dt_out (host). */
int phd_compute_time_step(phd_handle* h, const phd_cells* c, double gamma,
                          double cfl, int boost, double* dt_out);

/* Allocation-free variant of phd_compute_time_step.  No counterpart in the
 * reference: for callers that cannot allocate device memory (harnesses that
 * forbid hipMalloc, serving loops).  Needs no handle and allocates nothing:
 * its per-block partial minima live in a fixed device array of 512 KiB per
 * GPU (resolved with hipGetSymbolAddress on the first call on each device
 * and held until the process ends).  Runs on `stream` (a hipStream_t of the
 * current device; NULL = default stream) and synchronises it.  Same bits as
 * phd_compute_time_step.  At most phd_noalloc_max_cells() cells (16,777,216):
 * larger inputs return PHD_E_SIZE.  Calls that overlap on one GPU share its
 * array and race. */
int phd_compute_time_step_noalloc(const phd_cells* c, double gamma, double cfl,
                                  int boost, void* stream, double* dt_out);
int phd_noalloc_max_cells(void);

/* ---- mesh motion (mesh/mesh.pyx) ---------------------------------------- */
/* Mesh.assign_generator_velocities, mesh.pyx:363-426 (all cells) */
int phd_assign_generator_velocities(phd_handle* h, const phd_cells* c,
                                    double gamma, double eta, int regularize);
/* Mesh.assign_face_velocities, mesh.pyx:428-477 (all faces) */
int phd_assign_face_velocities(phd_handle* h, const phd_cells* c,
                               const phd_faces* f);

/* ---- reconstruction (reconstruction/reconstruction.pyx) ----------------- */
/* PieceWiseLinear.compute_gradients, reconstruction.pyx:563-737 (real
 * cells): Springel (2010) eq. 21 gradients and the AREPO (limiter 0) or TESS
 * (1) limiter, written into the cell-record rows.  The ghost transfer at 739
 * is phd_update_ghost_gradients. */
int phd_compute_gradients(phd_handle* h, phd_plan* p, const phd_cells* c,
                          int limiter, double* rec, int32_t rec_ld);
/* DomainManager.update_ghost_gradients (serial), domain_manager.pyx:819-837,
 * with Reflective.update_gradients (boundary.pyx:265-307) or the periodic
 * no-op (boundary.pyx:457-478).  lo/hi: host arrays of dim box bounds.
 * PHD_E_ARG when the plan lists ghosts and c->map is NULL. */
int phd_update_ghost_gradients(phd_handle* h, phd_plan* p, const phd_cells* c,
                               int boundary, const double* lo, const double* hi,
                               double* rec, int32_t rec_ld);
/* PieceWiseConstant.compute_states (reconstruction.pyx:254-352) when
 * order == PHD_ORDER_CONSTANT; PieceWiseLinear.add_spatial (741-921) and/or
 * add_temporal (923-1072) when order == PHD_ORDER_LINEAR, selected by the
 * `stages` bit mask.  left/right: dim+2 face arrays each (primitive group).
 * With stages == PHD_STAGE_TEMPORAL the states are read and updated in place. */
int phd_compute_states(phd_handle* h, phd_plan* p, const phd_cells* c,
                       const phd_faces* f, const double* rec, int32_t rec_ld,
                       double gamma, double dt, int boost, int gizmo, int order,
                       int stages, double* const* left, double* const* right);

/* ---- Riemann solvers (riemann/riemann.pyx) ------------------------------ */
/* RiemannBase.compute_fluxes (74-98) with HLL (243-373), HLLC (521-682) or
 * Exact (700-972) and RiemannBase.deboost (190-222), from materialised
 * states.  flux: dim+2 face arrays (conservative group). */
int phd_compute_fluxes(phd_handle* h, const phd_faces* f, int dim,
                       double* const* left, double* const* right, double gamma,
                       int boost, int solver, double* const* flux);
/* Fused PieceWise{Constant,Linear}.compute_states + compute_fluxes: the face
 * states are built in registers and never written.  Bit-identical to
 * phd_compute_states followed by phd_compute_fluxes. */
int phd_face_flux(phd_handle* h, phd_plan* p, const phd_cells* c,
                  const phd_faces* f, const double* rec, int32_t rec_ld,
                  double gamma, double dt, int boost, int gizmo, int order,
                  int solver, double* const* flux);

/* ---- conserved update (mesh/mesh.pyx) ----------------------------------- */
/* Mesh.update_from_fluxes, mesh.pyx:479-551.  Deterministic: each real cell
 * applies its faces in ascending face id, as the reference's face loop does,
 * so the result is bit-identical.  Positivity faults are recorded in the
 * error word (PHD_F_MASS / PHD_F_ENERGY at the first failing face). */
int phd_update_from_fluxes(phd_handle* h, phd_plan* p, const phd_cells* c,
                           const phd_faces* f, double* const* flux, double dt);

/* ---- core-step stages (what phd_core_step runs, callable one by one) ----- */
/* phd_compute_gradients without the packing of the cell rows: the plan's
 * packed rows must be current (phd_pack_cells or phd_ghost_refresh_eos). */
int phd_gradients_packed(phd_handle* h, phd_plan* p, const phd_cells* c,
                         int limiter, double* rec, int32_t rec_ld);
/* Fused states (at dt/2) + Riemann + deboost that writes, per face, the
 * update summands (dt*area)*flux into the plan (mesh.pyx:530-544 forms the
 * same product for each endpoint), and the fluxes when `flux` is not NULL. */
int phd_face_summands(phd_handle* h, phd_plan* p, const phd_cells* c,
                      const phd_faces* f, const double* rec, int32_t rec_ld,
                      double gamma, double dt, int boost, int gizmo, int solver,
                      double* const* flux);
/* Mesh.update_from_fluxes (mesh.pyx:479-551) from the plan's summands. */
int phd_update_from_summands(phd_handle* h, phd_plan* p, const phd_cells* c);

/* ---- fused core step ---------------------------------------------------- */
/* One MovingMeshMUSCLHancock hydro step on a frozen mesh with precomputed
 * face velocities (integrate.py:396-408): gradients + limiter, ghost
 * gradients, fused states (at dt/2) + Riemann + deboost, conserved update at
 * dt.  The update's summands (dt*area)*flux are formed once per face by the
 * face kernel, the same product the reference forms for each endpoint.
 * `flux` (dim+2 face arrays) may be NULL when the caller does not need the
 * fluxes themselves. */
int phd_core_step(phd_handle* h, phd_plan* p, const phd_cells* c,
                  const phd_faces* f, double* rec, int32_t rec_ld, int limiter,
                  int boundary, const double* lo, const double* hi,
                  double gamma, double dt, int boost, int gizmo, int solver,
                  double* const* flux, int flags);

/* ---- static-mesh epilogue ---------------------------------------------- */
/* DomainManager.update_ghost_fields(particles, conservative, True)
 * (domain_manager.pyx:759-774, Reflective.update_fields boundary.pyx:309-348)
 * followed by IdealGas.primitive_from_conservative (equation_state.pyx:81-117)
 * over all cells, fused (StaticMeshMUSCLHancock.evolve_timestep,
 * integrate.py:371-377).  With a plan, also refreshes its packed cell rows
 * (then phd_core_step may be called with PHD_CORE_XP_FRESH).  c->map may be
 * NULL only when there are no ghosts: PHD_E_ARG when the plan lists ghosts,
 * and without a plan the tags are scanned first (synchronising). */
int phd_ghost_refresh_eos(phd_handle* h, phd_plan* p, const phd_cells* c,
                          int boundary, const double* lo, const double* hi,
                          double gamma);

/* ---- domain and boundary loops between flux steps ------------------------
 * (domain/domain_manager.pyx, domain/boundary.pyx; serial branches) */
/* DomainManager.move_generators, domain_manager.pyx:566-590: x += dt*w for
 * real cells.  Writes the cells' position arrays (c->x). */
int phd_move_generators(phd_handle* h, const phd_cells* c, double dt);
/* Periodic.migrate_particles, boundary.pyx:502-538: each real cell outside
 * the box [lo, hi] (host arrays of dim bounds) is wrapped on every axis
 * where its position before the wrap is <= lo (+ (hi - lo)) or >= hi
 * (- (hi - lo)).  Writes c->x. */
int phd_migrate_periodic(phd_handle* h, const phd_cells* c, const double* lo,
                         const double* hi);
/* Reflective.migrate_particles, boundary.pyx:236-263 (the check that raises
 * "particle left domain in reflective boundary condition!!"): *first = the
 * smallest index of a real cell outside the box, or -1.  Synchronising. */
int phd_first_outside(phd_handle* h, const phd_cells* c, const double* lo,
                      const double* hi, int64_t* first);
/* Reflective.update_fields, boundary.pyx:309-348: velocity (primitive
 * group) and momentum (conservative group) of every exterior cell (type 2)
 * negated on each axis where it lies outside the box. */
int phd_reflect_fields(phd_handle* h, const phd_cells* c, const double* lo,
                       const double* hi);
/* DomainManager.update_ghost_fields, serial branch, domain_manager.pyx:
 * 757-770: every ghost (tag 1) takes each of the n_fields arrays' value at
 * its image c->map[i].  fields: device arrays of n_cells elements of
 * elem_size bytes (1, 4 or 8) each; the boundary rule is the caller's
 * separate phd_reflect_fields, as in the reference. */
int phd_update_ghost_fields(phd_handle* h, const phd_cells* c,
                            void* const* fields, const int32_t* elem_size,
                            int32_t n_fields);

/* ---- mesh geometry from Delaunay topology --------------------------------
 * The device half of Tess2d::extract_geometry (tess.cpp:238-398) and
 * Tess3d::extract_geometry (tess3.cpp:334-562), called by Mesh.build_geometry
 * (mesh.pyx:255-323) after the host's topology-only build
 * (phdm_build_topology, phd_mesh.h).  The topology is the host's triangulation
 * in the builder's layout: faces in owner order (pair_i real and non-decreasing,
 * pair_j > pair_i or a ghost), ghosts n_real .. n_real + n_ghost - 1 with their image's real
 * index and image code, each face's ring (the generators whose bisectors
 * bound it, in polygon order) and the per-cell face lists.  All arrays are
 * device arrays. */
typedef struct phd_topology {
    int32_t        dim;        /* 2 or 3 */
    int32_t        boundary;   /* 0 reflective, 1 periodic */
    int32_t        n_real, n_ghost, n_faces;
    double         lo[3], hi[3];
    const int32_t* ghost_map;  /* n_ghost: real index of each ghost's image */
    const int8_t*  ghost_code; /* n_ghost: image code */
    const int32_t* pair_i;     /* n_faces */
    const int32_t* pair_j;
    const int32_t* ring_off;   /* n_faces + 1 */
    const int32_t* ring_k;     /* ring entries: real index ... */
    const int8_t*  ring_code;  /* ... and image code */
    const int32_t* nbr_off;    /* n_real + n_ghost + 1 */
    const int32_t* nbr_face;   /* face ids, ascending within each cell */
} phd_topology;

/* Outputs (and the generator positions) of phd_extract_geometry. */
typedef struct phd_geometry {
    double* x[3];       /* n_real + n_ghost: real rows read, ghost rows written */
    double* volume;     /* n_real + n_ghost */
    double* dcom[3];    /* n_real + n_ghost, centroid - generator */
    double* area;       /* n_faces */
    double* com[3];
    double* normal[3];
} phd_geometry;

/* Bytes of the workspace phd_extract_geometry needs: 2 (1 + dim) doubles
 * per face (each side's volume term and centroid moment), rounded up to 256
 * bytes, then two int32 per real cell (its run of owned faces). */
size_t phd_geometry_workspace_bytes(int dim, int64_t n_real, int64_t n_faces);
/* Ghost positions (the reference's image formulas, boundary.pyx:103-104,
 * 378-380), then face area, centroid and normal and cell volume and dcom,
 * with the reference's formulas (Voronoi vertices are circumcentres of the
 * ring's Delaunay simplices; 3D faces by a fan about the vertex mean and the
 * zero-area rule of tess3.cpp:494-502; volumes by triangles / tetrahedra on
 * the faces; ghost volume and dcom copied from the image).  Each face is
 * computed once, in its owner generator's frame, and each cell sums its
 * faces in ascending face id.  Synchronising.  A topology entry out of range
 * returns PHD_E_ARG and *bad = the first such face, or -2 - c for cell (or
 * ghost) c. */
int phd_extract_geometry(phd_handle* h, const phd_topology* t, const phd_geometry* g,
                         void* workspace, size_t workspace_bytes, int64_t* bad);

/* ---- batched ensembles ---------------------------------------------------
 * No counterpart in the reference, which runs one problem per process.
 * Many independent problems are laid out as one union mesh: problem p owns
 * cells [cell_start[p], cell_start[p+1]) (its reals, then its ghosts) and a
 * contiguous run of faces in owner order; face, neighbour-list and ghost-map
 * indices are global, and no face joins two problems.  All problems share
 * the boundary kind and box.  Every kernel of a step is flat over the union,
 * so one launch serves all problems; the time step is per problem (a
 * per-problem CFL minimum folded on the device).  A step can be captured
 * once in a HIP graph and replayed with one graph launch per step, with no
 * host synchronisation.  Static meshes (face velocities taken as zero, as
 * StaticMeshMUSCLHancock sets them, integrate.py:340-347), limiter AREPO and
 * solver HLLC only. */
typedef struct {
    int32_t grad_block;    /* threads per block of the gradient kernel: 64, 128, 256 */
    int32_t face_block;    /* ... of the face kernel */
    int32_t update_block;  /* ... of the update kernel */
    int32_t xcd;           /* 1: blocks remapped so each XCD takes a contiguous
                              eighth of the cells / faces; 0: dispatch order */
} phd_schedule;
typedef struct phd_batch phd_batch;
/* cell_start: host array of n_problems + 1 offsets; hist_steps: capacity of
 * the per-step time-step history (steps x problems, 0 for none).
 * Allocating and synchronising (it also reads the tags, to list the ghost
 * cells); fails with PHD_E_ARG if a face joins two problems or a tag is
 * neither 0 nor 1. */
int phd_batch_create(phd_handle* h, phd_plan* p, const phd_cells* c,
                     const phd_faces* f, int32_t n_problems,
                     const int32_t* cell_start, int64_t hist_steps,
                     phd_batch** out);
int phd_batch_destroy(phd_batch* b);
/* Launch schedule of the step's three main kernels (drops a captured
 * graph).  Default {256, 256, 256, 1}. */
int phd_batch_set_schedule(phd_batch* b, const phd_schedule* s);
/* Ghost refresh + EOS of phd_ghost_refresh_eos over the union, also folding
 * each problem's CFL minimum and fault flag for the next step, and writing
 * the packed rows whole (positions included).  Call after any change of the
 * conserved fields made outside the batch's steps. */
int phd_batch_refresh(phd_handle* h, phd_batch* b, phd_plan* p, const phd_cells* c,
                      int boundary, const double* lo, const double* hi, double gamma,
                      int boost);
/* One step of every problem on h's stream, no host synchronisation:
 * dt[p] = cfl * min over p's real cells of the CFL bound (riemann.pyx:
 * 120-188) as the last step or refresh folded it, the core step of
 * phd_core_step with that dt per problem, whose update also evaluates the
 * EOS of the real cells, refreshes their packed rows' primitives and folds
 * the next CFL minimum, then the ghosts' refresh (the fields equal those of
 * phd_core_step + phd_ghost_refresh_eos bit for bit), and the history
 * record.  The packed rows' positions are not rewritten (static meshes).
 * flags: PHD_CORE_XP_FRESH as for phd_core_step. */
int phd_batch_step(phd_handle* h, phd_batch* b, phd_plan* p, const phd_cells* c,
                   const phd_faces* f, double* rec, int32_t rec_ld, int boundary,
                   const double* lo, const double* hi, double gamma, double cfl,
                   int boost, int gizmo, int flags);
/* Capture one phd_batch_step (packed rows fresh, as every step leaves them)
 * into a graph on the batch's own stream.  The captured update writes the
 * real cells' primitive arrays only once per phd_batch_launch, after its
 * last step (no kernel of a step reads them; the values are the same).  The
 * arrays named by the arguments must stay allocated while the graph is
 * used. */
int phd_batch_capture(phd_handle* h, phd_batch* b, phd_plan* p, const phd_cells* c,
                      const phd_faces* f, double* rec, int32_t rec_ld, int boundary,
                      const double* lo, const double* hi, double gamma, double cfl,
                      int boost, int gizmo);
/* n_steps graph launches, then the real cells' primitive arrays from the
 * packed rows, ordered after the work already on h's stream and before the
 * work enqueued there afterwards; no host synchronisation.  The cells must
 * hold a refreshed state (packed rows fresh). */
int phd_batch_launch(phd_handle* h, phd_batch* b, int32_t n_steps);
/* Reset the step counter of the history and the fault marks (on h's
 * stream). */
int phd_batch_reset_history(phd_handle* h, phd_batch* b);
/* Synchronise and copy out: dt (n_problems, the last step's), the number of
 * steps recorded since the last reset, the history (steps x problems) and
 * the fault marks (n_problems: 0, or 1 + the step count at which the
 * problem was frozen); NULL skips an output.  A problem is frozen (dt = 0,
 * state unchanged) from the step after one that left a real cell with
 * non-positive mass, energy, density or pressure, where the reference
 * raises for its single problem; the other problems go on, and the fault is
 * also recorded as for phd_core_step. */
int phd_batch_read(phd_handle* h, phd_batch* b, double* dt, int64_t* steps,
                   double* history, int32_t* faults);

/* ---- reverse mode (adjoint) ----------------------------------------------
 * No counterpart in the reference, which has no derivatives; these serve
 * gradient-based inversion and optimisation through the step above.  The
 * differentiated step is phd_core_step (AREPO limiter, HLLC, face
 * velocities held fixed: a static or prescribed mesh) followed by
 * phd_ghost_refresh_eos. */
typedef struct phd_adjoint phd_adjoint;
/* Per-mesh adjoint workspace (mirror-entry table, ghost lists, per-face and
 * per-cell adjoint buffers).  Allocating and synchronising; recreate with
 * the plan. */
int phd_adjoint_create(phd_handle* h, phd_plan* p, const phd_cells* c,
                       const phd_faces* f, phd_adjoint** out);
int phd_adjoint_destroy(phd_adjoint* a);
/* Flags of phd_core_step_vjp */
#define PHD_VJP_REC_FRESH 1   /* `rec` already holds this state's limited and
                                 ghost gradients (what phd_core_step wrote into
                                 it from this state) and the plan's packed rows
                                 are current (phd_ghost_refresh_eos ran with
                                 this plan since the fields last changed) */
/* Given lam_out = dL/dcons after the step (real cells; dim+2 arrays of
 * n_cells), writes lam_in = dL/dcons before it (real entries; the ghost
 * entries are not touched).  `c` holds the state the step started from
 * (conserved and primitive fields consistent, ghosts refreshed); `rec` is a
 * cell-record buffer (phd_cell_record_ld), rebuilt here unless flags has
 * PHD_VJP_REC_FRESH.  Deterministic: every sum runs in a fixed order. */
int phd_core_step_vjp(phd_handle* h, phd_plan* p, phd_adjoint* a, const phd_cells* c,
                      const phd_faces* f, double* rec, int32_t rec_ld, int boundary,
                      const double* lo, const double* hi, double gamma, double dt,
                      int boost, int gizmo, double* const* lam_out,
                      double* const* lam_in, int flags);
/* lam_x += J^T lam_y for the equation of state of the real cells:
 * to_prim = 1: y = prim(cons), x = cons;  to_prim = 0: y = cons(prim),
 * x = prim (equation_state.pyx:43-117). */
int phd_eos_vjp(phd_handle* h, const phd_cells* c, double gamma, int to_prim,
                double* const* lam_y, double* const* lam_x);

#ifdef __cplusplus
}
#endif
#endif /* PHD_FLUX_H */
