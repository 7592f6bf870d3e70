/*
 * phd_oracle.c -- CPU ORACLE for the Voronoi face-flux pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a serial C restatement of the
 * reference's Cython loops, used as the parity checker for the HIP library
 * and as the timed CPU baseline.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product path never does.
 *
 * Every function follows the loop nest and floating-point operation order of
 * the reference source it cites, so that a build with
 * `gcc -O2 -ffp-contract=off` (no FMA contraction, no fast-math) performs the
 * same IEEE-754 double operations as the reference's own x86-64 build.
 *
 * Conventions (all as in the reference):
 *   - primitive fields  prim[n], n = 0..D+1 : density, velocity-x.., pressure
 *       (particle_creator.pyx:50-52)
 *   - conservative      cons[n]             : mass, momentum-x.., energy
 *       (particle_creator.pyx:53-55)
 *   - gradients         grad[n*D+k] = d prim_n / d x_k
 *       (reconstruction.pyx:470-480)
 *   - tags: 0 real, 1 ghost; types: 2 exterior (particle_tags.py:1-6)
 *   - per-cell face lists are CSR (nbr_off, nbr_face) in ascending face id,
 *     the order Tess2d/Tess3d::extract_geometry pushes them
 *     (tess.cpp:379-382, tess3.cpp:531-534)
 *
 * Squares written `x**2` / `pow(x, 2.0)` in the reference are written x*x
 * here: gcc folds pow(x, 2.0) to x*x at -O2, so the reference binary
 * computes the same correctly rounded product.
 *
 * Parity pinning of each function is in the test suite
 * (tests/test_oracle_riemann.py, tests/test_oracle_cells.py): known answers
 * from the reference tests, closed forms, conservation identities (the
 * Rankine-Hugoniot conditions of the HLLC star states), the exact solution
 * where an estimate is exact, hand-worked cases and brute force on tiny
 * inputs.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <stdlib.h>

#define REAL 0
#define GHOST 1
#define EXTERIOR 2

/* error codes shared with the HIP library's error word */
#define ORC_OK 0
#define ORC_ERR_NEIGHBORS 1     /* reconstruction.pyx:650-652 */
#define ORC_ERR_MASS 2          /* mesh.pyx:546-548 */
#define ORC_ERR_ENERGY 3        /* mesh.pyx:549-551 */
#define ORC_ERR_EXACT 4         /* riemann.pyx:1088-1090 */

static const double PI = 3.141592653589793; /* numpy.pi */

/* Loops whose iterations are independent (per cell or per face) carry
 * ORC_PFOR: with -fopenmp (libphd_oracle_omp.so, the all-cores CPU line)
 * they run in parallel, each iteration doing exactly the serial work, so
 * both builds give the same bits; without it the macro is empty. */
#define ORC_STR(x) #x
#ifdef _OPENMP
#include <omp.h>
#define ORC_PFOR(clauses) _Pragma(ORC_STR(omp parallel for schedule(static) clauses))
#else
#define ORC_PFOR(clauses)
#endif

int orc_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* equation of state  (equation_state.pyx:119-121)                          */
static double sound_speed(double gamma, double d, double p)
{
    return sqrt(gamma*p/d);
}

/* conservative_from_primitive (equation_state.pyx:43-79), all cells */
void orc_cons_from_prim(int dim, int n, double gamma, double* const* prim,
                        const double* vol, double* const* cons)
{
    for (int i = 0; i < n; i++) {
        cons[0][i] = prim[0][i]*vol[i];
        double v_sq = 0.;
        for (int k = 0; k < dim; k++) {
            cons[1+k][i] = prim[1+k][i]*cons[0][i];
            v_sq += prim[1+k][i]*prim[1+k][i];
        }
        cons[dim+1][i] = (.5*prim[0][i]*v_sq + prim[dim+1][i]/(gamma-1.))*vol[i];
    }
}

/* primitive_from_conservative (equation_state.pyx:81-117), all cells */
void orc_prim_from_cons(int dim, int n, double gamma, double* const* cons,
                        const double* vol, double* const* prim)
{
    for (int i = 0; i < n; i++) {
        prim[0][i] = cons[0][i]/vol[i];
        double v_sq = 0.;
        for (int k = 0; k < dim; k++) {
            prim[1+k][i] = cons[1+k][i]/cons[0][i];
            v_sq += prim[1+k][i]*prim[1+k][i];
        }
        prim[dim+1][i] = (cons[dim+1][i]/vol[i] - .5*prim[0][i]*v_sq)*(gamma-1.);
    }
}

/* ------------------------------------------------------------------------ */
/* cell radius used by the CFL and regularisation
 * (riemann.pyx:155-158, mesh.pyx:414-417)                                  */
static double cell_radius(int dim, double vol)
{
    if (dim == 2)
        return sqrt(vol/PI);
    return pow(3.0*vol/(4.0*PI), 1.0/3.0);
}

/* RiemannBase.compute_time_step (riemann.pyx:120-188) */
double orc_compute_time_step(int dim, int n, double gamma, double cfl, int boost,
                             const int8_t* tag, double* const* prim,
                             const double* vol)
{
    double c = sound_speed(gamma, prim[0][0], prim[dim+1][0]);
    double R = cell_radius(dim, vol[0]);
    double dt, vsq;
    if (boost) {
        dt = R/c;
    } else {
        vsq = 0.0;
        for (int k = 0; k < dim; k++)
            vsq += prim[1+k][0]*prim[1+k][0];
        dt = R/(c + sqrt(vsq));
    }
    for (int i = 0; i < n; i++) {
        if (tag[i] != REAL) continue;
        c = sound_speed(gamma, prim[0][i], prim[dim+1][i]);
        R = cell_radius(dim, vol[i]);
        if (boost) {
            dt = fmin(R/c, dt);
        } else {
            vsq = 0.0;
            for (int k = 0; k < dim; k++)
                vsq += prim[1+k][i]*prim[1+k][i];
            dt = fmin(R/(c + sqrt(vsq)), dt);
        }
    }
    return cfl*dt;
}

/* ------------------------------------------------------------------------ */
/* Mesh.assign_generator_velocities (mesh.pyx:363-426), all cells */
void orc_generator_velocities(int dim, int n, double gamma, double eta,
                              int regularize, double* const* prim,
                              double* const* dcx, const double* vol,
                              double* const* wx)
{
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < dim; k++)
            wx[k][i] = prim[1+k][i];
        if (!regularize) continue;
        double c = sound_speed(gamma, prim[0][i], prim[dim+1][i]);
        double d = 0.0;
        for (int k = 0; k < dim; k++)
            d += dcx[k][i]*dcx[k][i];
        d = sqrt(d);
        double R = cell_radius(dim, vol[i]);
        if ((0.9 <= d/(eta*R)) && (d/(eta*R) < 1.1)) {
            for (int k = 0; k < dim; k++)
                wx[k][i] += c*dcx[k][i]*(d - 0.9*eta*R)/(d*0.2*eta*R);
        } else if (1.1 <= d/(eta*R)) {
            for (int k = 0; k < dim; k++)
                wx[k][i] += c*dcx[k][i]/d;
        }
    }
}

/* Mesh.assign_face_velocities (mesh.pyx:428-477), all faces */
void orc_face_velocities(int dim, int nf, double* const* x, double* const* wx,
                         double* const* fij, const int32_t* pair_i,
                         const int32_t* pair_j, double* const* fv)
{
    for (int m = 0; m < nf; m++) {
        int i = pair_i[m], j = pair_j[m];
        double factor = 0.0, denom = 0.0;
        for (int k = 0; k < dim; k++) {
            factor += (wx[k][i] - wx[k][j])*(fij[k][m] - 0.5*(x[k][i] + x[k][j]));
            denom  += (x[k][j] - x[k][i])*(x[k][j] - x[k][i]);
        }
        factor /= denom;
        for (int k = 0; k < dim; k++)
            fv[k][m] = 0.5*(wx[k][i] + wx[k][j]) + factor*(x[k][j] - x[k][i]);
    }
}

/* ------------------------------------------------------------------------ */
/* PieceWiseLinear.compute_gradients (reconstruction.pyx:563-737), real
 * cells only.  limiter: 0 = arepo (683-703), 1 = tess (705-731).
 * Returns ORC_ERR_NEIGHBORS (and *bad_cell) on an inconsistent face list. */
int orc_compute_gradients(int dim, int n, const int8_t* tag, double* const* x,
                          double* const* dcx, const double* vol,
                          double* const* prim, const double* face_area,
                          double* const* fij, const int32_t* pair_i,
                          const int32_t* pair_j, const int32_t* nbr_off,
                          const int32_t* nbr_face, int limiter,
                          double* const* grad, int32_t* bad_cell)
{
    const int nfld = dim + 2;
    double phi_max[5], phi_min[5], alpha[5], df[15];
    double xi[3], xj[3], cx[3], cfx[3], dr[3];

    /* the neighbour check (reconstruction.pyx:647-652): the first real cell
     * in loop order with a face that is not its own raises */
    for (int i = 0; i < n; i++) {
        if (tag[i] != REAL) continue;
        for (int m = nbr_off[i]; m < nbr_off[i+1]; m++) {
            int fid = nbr_face[m];
            if (i != pair_i[fid] && i != pair_j[fid]) {
                if (bad_cell) *bad_cell = i;
                return ORC_ERR_NEIGHBORS;
            }
        }
    }

    ORC_PFOR(private(phi_max, phi_min, alpha, df, xi, xj, cx, cfx, dr))
    for (int i = 0; i < n; i++) {
        if (tag[i] != REAL) continue;
        for (int k = 0; k < dim; k++) {
            xi[k] = x[k][i];
            cx[k] = xi[k] + dcx[k][i];
        }
        double _vol = vol[i];
        for (int f = 0; f < nfld; f++) {
            phi_max[f] = phi_min[f] = prim[f][i];
            alpha[f] = 1.0;
            for (int k = 0; k < dim; k++)
                df[dim*f+k] = 0;
        }

        /* Eq. 21 accumulation over the cell's faces (640-681) */
        int j = -1;
        for (int m = nbr_off[i]; m < nbr_off[i+1]; m++) {
            int fid = nbr_face[m];
            double area = face_area[fid];
            if (i == pair_i[fid]) j = pair_j[fid];
            else j = pair_i[fid];      /* i == pair_j[fid]: checked above */

            double r = 0.0;
            for (int k = 0; k < dim; k++) {
                xj[k] = x[k][j];
                cfx[k] = fij[k][fid] - 0.5*(xi[k] + xj[k]);
                dr[k] = xi[k] - xj[k];
                r += dr[k]*dr[k];
            }
            r = sqrt(r);
            for (int f = 0; f < nfld; f++) {
                phi_max[f] = fmax(phi_max[f], prim[f][j]);
                phi_min[f] = fmin(phi_min[f], prim[f][j]);
                double d_dif = prim[f][j] - prim[f][i];
                double d_sum = prim[f][j] + prim[f][i];
                for (int k = 0; k < dim; k++)
                    df[dim*f+k] += area*(d_dif*cfx[k] - 0.5*d_sum*dr[k])/(r*_vol);
            }
        }

        if (limiter == 0) {
            /* AREPO limiter, Eq. 30 (683-703).  dphi is a Python float in the
             * reference; its arithmetic is the same IEEE double arithmetic. */
            for (int f = 0; f < nfld; f++) {
                for (int m = nbr_off[i]; m < nbr_off[i+1]; m++) {
                    int fid = nbr_face[m];
                    double dphi = 0.0;
                    for (int k = 0; k < dim; k++)
                        dphi += df[dim*f+k]*(fij[k][fid] - cx[k]);
                    double psi;
                    if (dphi > 0) psi = (phi_max[f] - prim[f][i])/dphi;
                    else if (dphi < 0) psi = (phi_min[f] - prim[f][i])/dphi;
                    else psi = 1.0;
                    alpha[f] = fmin(alpha[f], fmax(psi, 0.));
                }
            }
        } else if (limiter == 1) {
            /* TESS limiter (705-731); the reference updates alpha[0] only. */
            for (int f = 0; f < nfld; f++) {
                for (int m = nbr_off[i]; m < nbr_off[i+1]; m++) {
                    int fid = nbr_face[m];
                    if (i == pair_i[fid]) j = pair_j[fid];
                    else if (i == pair_j[fid]) j = pair_i[fid];
                    double dphi = 0.0;
                    for (int k = 0; k < dim; k++)
                        dphi += df[dim*f+k]*(fij[k][fid] - cx[k]);
                    double psi;
                    if (dphi > 0.0 || dphi < 0.0) {
                        double q = (prim[f][j] - prim[f][i])/dphi;
                        psi = (0. > q) ? 0. : q;      /* Python max(q, 0.) */
                    } else {
                        psi = 1.0;
                    }
                    alpha[0] = fmin(alpha[0], fmax(psi, 0.));
                }
            }
        }

        for (int f = 0; f < nfld; f++)
            for (int k = 0; k < dim; k++)
                grad[dim*f+k][i] = alpha[f]*df[dim*f+k];
    }
    return ORC_OK;
}

/* DomainManager.update_ghost_gradients, serial branch
 * (domain_manager.pyx:819-837) followed by the boundary rule:
 *   boundary 0 = Reflective.update_gradients (boundary.pyx:265-307),
 *   boundary 1 = Periodic.update_gradients (no-op, boundary.pyx:457-478). */
void orc_ghost_gradients(int dim, int n, const int8_t* tag, const int8_t* type,
                         const int32_t* map, double* const* x,
                         const double* lo, const double* hi, int boundary,
                         double* const* grad)
{
    const int ng = dim*(dim + 2);
    ORC_PFOR()
    for (int i = 0; i < n; i++)
        if (tag[i] == GHOST)
            for (int g = 0; g < ng; g++)
                grad[g][i] = grad[g][map[i]];
    if (boundary != 0) return;
    for (int i = 0; i < n; i++) {
        if (type[i] != EXTERIOR) continue;
        for (int j = 0; j < dim; j++) {
            /* velocity group dv[dim*j + j] = grad[(1+j)*dim + j] */
            if (x[j][i] < lo[j]) grad[(1+j)*dim + j][i] *= -1;
            if (x[j][i] > hi[j]) grad[(1+j)*dim + j][i] *= -1;
        }
    }
}

/* Static-mesh epilogue of StaticMeshMUSCLHancock.evolve_timestep
 * (integrate.py:371-377): DomainManager.update_ghost_fields(particles,
 * conservative, True), serial branch (domain_manager.pyx:757-774): every
 * ghost takes its image's conserved fields, then the boundary rule
 * (Reflective.update_fields, boundary.pyx:309-348: momentum and velocity of
 * exterior cells negated on each axis where they lie outside the box;
 * Periodic: none), then IdealGas.primitive_from_conservative over all cells
 * (equation_state.pyx:81-117). */
void orc_ghost_refresh_eos(int dim, int n, double gamma, const int8_t* tag,
                           const int8_t* type, const int32_t* map,
                           double* const* x, const double* lo,
                           const double* hi, int boundary, const double* vol,
                           double* const* cons, double* const* prim)
{
    for (int i = 0; i < n; i++)
        if (tag[i] == GHOST)
            for (int q = 0; q < dim + 2; q++)
                cons[q][i] = cons[q][map[i]];
    if (boundary == 0) {
        for (int i = 0; i < n; i++) {
            if (type[i] != EXTERIOR) continue;
            for (int k = 0; k < dim; k++) {
                if (x[k][i] < lo[k]) { prim[1+k][i] *= -1; cons[1+k][i] *= -1; }
                if (x[k][i] > hi[k]) { prim[1+k][i] *= -1; cons[1+k][i] *= -1; }
            }
        }
    }
    orc_prim_from_cons(dim, n, gamma, cons, vol, prim);
}

/* ------------------------------------------------------------------------ */
/* PieceWiseConstant.compute_states (reconstruction.pyx:254-352) */
void orc_states_constant(int dim, int nf, double* const* prim, double* const* fv,
                         const int32_t* pair_i, const int32_t* pair_j, int boost,
                         double* const* pl, double* const* pr)
{
    for (int m = 0; m < nf; m++) {
        int i = pair_i[m], j = pair_j[m];
        pl[0][m] = prim[0][i];
        pr[0][m] = prim[0][j];
        pl[dim+1][m] = prim[dim+1][i];
        pr[dim+1][m] = prim[dim+1][j];
        for (int k = 0; k < dim; k++) {
            if (boost) {
                pl[1+k][m] = prim[1+k][i] - fv[k][m];
                pr[1+k][m] = prim[1+k][j] - fv[k][m];
            } else {
                pl[1+k][m] = prim[1+k][i];
                pr[1+k][m] = prim[1+k][j];
            }
        }
    }
}

/* PieceWiseLinear.add_spatial (reconstruction.pyx:741-921) */
void orc_add_spatial(int dim, int nf, double* const* prim, double* const* x,
                     double* const* dcx, double* const* grad,
                     double* const* fij, double* const* fv,
                     const int32_t* pair_i, const int32_t* pair_j, int boost,
                     int gizmo, double* const* pl, double* const* pr)
{
    const int nfld = dim + 2;
    const double psi1 = 0.5, psi2 = 0.25;
    double state_l[5], state_r[5];

    ORC_PFOR(private(state_l, state_r))
    for (int m = 0; m < nf; m++) {
        int i = pair_i[m], j = pair_j[m];
        for (int f = 0; f < nfld; f++) {
            pl[f][m] = prim[f][i];
            pr[f][m] = prim[f][j];
        }
        for (int k = 0; k < dim; k++) {
            if (boost) {
                pl[1+k][m] = prim[1+k][i] - fv[k][m];
                pr[1+k][m] = prim[1+k][j] - fv[k][m];
            }
        }
        for (int f = 0; f < nfld; f++) {
            state_l[f] = pl[f][m];
            state_r[f] = pr[f][m];
        }

        double diff_mag = 0.0, sepi_mag = 0.0, sepj_mag = 0.0;
        for (int k = 0; k < dim; k++) {
            double sepi = fij[k][m] - (x[k][i] + dcx[k][i]);
            double sepj = fij[k][m] - (x[k][j] + dcx[k][j]);
            sepi_mag += sepi*sepi;
            sepj_mag += sepj*sepj;
            diff_mag += (x[k][j] - x[k][i])*(x[k][j] - x[k][i]);
            for (int f = 0; f < nfld; f++) {
                pl[f][m] += grad[f*dim+k][i]*sepi;
                pr[f][m] += grad[f*dim+k][j]*sepj;
            }
        }
        sepi_mag = sqrt(sepi_mag);
        sepj_mag = sqrt(sepj_mag);
        diff_mag = sqrt(diff_mag);

        if (!gizmo) continue;
        /* GIZMO pairwise limiter, Hopkins (2015) appendix B4 (877-912) */
        for (int f = 0; f < nfld; f++) {
            double delta1 = psi1*fabs(state_l[f] - state_r[f]);
            double delta2 = psi2*fabs(state_l[f] - state_r[f]);
            double phi_min = fmin(state_l[f], state_r[f]);
            double phi_max = fmax(state_l[f], state_r[f]);
            double phibar_l = state_l[f] + sepi_mag/diff_mag*(state_r[f] - state_l[f]);
            double phibar_r = state_r[f] + sepj_mag/diff_mag*(state_l[f] - state_r[f]);
            double phi_plus, phi_minus;
            if ((phi_max + delta1)*phi_max >= 0.) phi_plus = phi_max + delta1;
            else phi_plus = phi_max/(1 + delta1/fabs(phi_max));
            if ((phi_min - delta1)*phi_min >= 0.) phi_minus = phi_min - delta1;
            else phi_minus = phi_min/(1 + delta1/fabs(phi_min));

            if (prim[f][i] < prim[f][j]) {
                pl[f][m] = fmax(phi_minus, fmin(phibar_l + delta2, pl[f][m]));
                pr[f][m] = fmin(phi_plus, fmax(phibar_r - delta2, pr[f][m]));
            } else if (prim[f][i] > prim[f][j]) {
                pl[f][m] = fmin(phi_plus, fmax(phibar_l - delta2, pl[f][m]));
                pr[f][m] = fmax(phi_minus, fmin(phibar_r + delta2, pr[f][m]));
            } else {
                pl[f][m] = state_l[f];
                pr[f][m] = state_r[f];
            }
        }
        /* negativity fallback (914-921) */
        if (pl[0][m] < 0.0 || pl[dim+1][m] < 0.0)
            for (int f = 0; f < nfld; f++) pl[f][m] = state_l[f];
        if (pr[0][m] < 0.0 || pr[dim+1][m] < 0.0)
            for (int f = 0; f < nfld; f++) pr[f][m] = state_r[f];
    }
}

/* PieceWiseLinear.add_temporal (reconstruction.pyx:923-1072); dt is the
 * half step the integrator passes (integrate.py:398-400). */
void orc_add_temporal(int dim, int nf, double gamma, double dt,
                      double* const* prim, double* const* grad,
                      double* const* fv, const int32_t* pair_i,
                      const int32_t* pair_j, int boost,
                      double* const* pl, double* const* pr)
{
    const int nfld = dim + 2;
    double state_l[5], state_r[5], vi[3], vj[3];
    /* gradient groups: dd[k] = grad[k], dv[n*dim+k] = grad[(1+n)*dim+k],
     * dp[k] = grad[(dim+1)*dim+k]  (reconstruction.pyx:462-484) */
    double* const* dd = grad;
    double* const* dv = grad + dim;
    double* const* dp = grad + (dim+1)*dim;

    ORC_PFOR(private(state_l, state_r, vi, vj))
    for (int m = 0; m < nf; m++) {
        int i = pair_i[m], j = pair_j[m];
        for (int f = 0; f < nfld; f++) {
            state_l[f] = pl[f][m];
            state_r[f] = pr[f][m];
        }
        for (int k = 0; k < dim; k++) {
            if (boost) {
                vi[k] = prim[1+k][i] - fv[k][m];
                vj[k] = prim[1+k][j] - fv[k][m];
            } else {
                vi[k] = prim[1+k][i];
                vj[k] = prim[1+k][j];
            }
            pl[1+k][m] -= dt*dp[k][i]/prim[0][i];
            pr[1+k][m] -= dt*dp[k][j]/prim[0][j];
        }
        for (int k = 0; k < dim; k++) {
            pl[0][m] -= dt*(prim[0][i]*dv[(dim+1)*k][i] + vi[k]*dd[k][i]);
            pr[0][m] -= dt*(prim[0][j]*dv[(dim+1)*k][j] + vj[k]*dd[k][j]);
            pl[dim+1][m] -= dt*(gamma*prim[dim+1][i]*dv[(dim+1)*k][i] + vi[k]*dp[k][i]);
            pr[dim+1][m] -= dt*(gamma*prim[dim+1][j]*dv[(dim+1)*k][j] + vj[k]*dp[k][j]);
            for (int q = 0; q < dim; q++) {
                pl[1+q][m] -= dt*vi[k]*dv[q*dim+k][i];
                pr[1+q][m] -= dt*vj[k]*dv[q*dim+k][j];
            }
        }
        if (pl[0][m] < 0.0 || pl[dim+1][m] < 0.0)
            for (int f = 0; f < nfld; f++) pl[f][m] = state_l[f];
        if (pr[0][m] < 0.0 || pr[dim+1][m] < 0.0)
            for (int f = 0; f < nfld; f++) pr[f][m] = state_r[f];
    }
}

/* ------------------------------------------------------------------------ */
/* HLL.get_waves (riemann.pyx:375-504): Toro PVRS / TRRS / TSRS estimates.
 * get_waves_est also returns the star-pressure estimate the wave speeds are
 * built from and the branch that produced it (0 PVRS, 1 TRRS, 2 TSRS), which
 * the reference keeps local; orc_get_waves_pstar exposes them to the tests. */
static void get_waves_est(double dl, double ul, double pl, double dr, double ur,
                          double pr, double gamma, double* sl, double* sc,
                          double* sr, double* p_est, int* branch)
{
    const double Q = 2., c_floor = 1.0E-10;
    double cl = fmax(sqrt(gamma*pl/dl), c_floor);
    double cr = fmax(sqrt(gamma*pr/dr), c_floor);
    double d_avg = .5*(dl + dr);
    double c_avg = .5*(cl + cr);
    double p_star = fmax(0., .5*(pl + pr) + .5*(ul - ur)*d_avg*c_avg);
    double p_min = fmin(pl, pr);
    double p_max = fmax(pl, pr);

    if (((p_max/p_min) < Q) && ((p_min < p_star) && (p_star < p_max))) {
        /* PVRS; u_star (445) is not used further */
        *branch = 0;
    } else if (p_star <= p_min) {
        *branch = 1;
        double z = (gamma - 1.)/(2.*gamma);
        double plr = pow(pl/pr, z);
        double u_star = (plr*ul/cl + ur/cr + 2.*(plr - 1.)/(gamma - 1.))/
                        (plr/cl + 1./cr);
        p_star  = .5*pl*pow(1. + (gamma - 1.)*(ul - u_star)/(2.*cl), 1./z);
        p_star += .5*pr*pow(1. + (gamma - 1.)*(u_star - ur)/(2.*cr), 1./z);
    } else {
        *branch = 2;
        double Al = 2./((gamma + 1.)*dl);
        double Ar = 2./((gamma + 1.)*dr);
        double Bl = pl*((gamma - 1.)/(gamma + 1.));
        double Br = pr*((gamma - 1.)/(gamma + 1.));
        double gl = sqrt(Al/(p_star + Bl));
        double gr = sqrt(Ar/(p_star + Br));
        p_star = (gl*pl + gr*pr - (ur - ul))/(gl + gr);
        /* u_star (479) is not used further */
    }
    double _sl, _sr;
    if (p_star <= pl) _sl = ul - cl;
    else _sl = ul - cl*sqrt(1.+((gamma+1.)/(2.*gamma))*(p_star/pl - 1.));
    if (p_star <= pr) _sr = ur + cr;
    else _sr = ur + cr*sqrt(1. + ((gamma+1.)/(2.*gamma))*(p_star/pr - 1.));
    *sc = (pr - pl + dl*ul*(_sl - ul) - dr*ur*(_sr - ur))/(dl*(_sl - ul) - dr*(_sr - ur));
    *sl = _sl;
    *sr = _sr;
    *p_est = p_star;
}

void orc_get_waves(double dl, double ul, double pl, double dr, double ur,
                   double pr, double gamma, double* sl, double* sc, double* sr)
{
    double p_est;
    int branch;
    get_waves_est(dl, ul, pl, dr, ur, pr, gamma, sl, sc, sr, &p_est, &branch);
}

void orc_get_waves_pstar(double dl, double ul, double pl, double dr, double ur,
                         double pr, double gamma, double* sl, double* sc,
                         double* sr, double* p_est, int* branch)
{
    get_waves_est(dl, ul, pl, dr, ur, pr, gamma, sl, sc, sr, p_est, branch);
}

/* RiemannBase.deboost (riemann.pyx:190-222), Pakmor & Springel (2011) eq 17 */
static void deboost(int dim, int nf, double* const* fv, double* const* flux)
{
    for (int m = 0; m < nf; m++)
        for (int k = 0; k < dim; k++) {
            flux[dim+1][m] += fv[k][m]*(0.5*fv[k][m]*flux[0][m] + flux[1+k][m]);
            flux[1+k][m] += fv[k][m]*flux[0][m];
        }
}

/* HLL.riemann_solver (riemann.pyx:243-373) */
void orc_hll(int dim, int nf, double gamma, int boost, double* const* sL,
             double* const* sR, double* const* nx, double* const* fv,
             double* const* flux)
{
    for (int i = 0; i < nf; i++) {
        double _dl = sL[0][i], _pl = sL[dim+1][i];
        double _dr = sR[0][i], _pr = sR[dim+1][i];
        double Vnl = 0.0, Vnr = 0.0, vl_sq = 0.0, vr_sq = 0.0, wn = 0.0;
        for (int k = 0; k < dim; k++) {
            double vl_tmp = sL[1+k][i], vr_tmp = sR[1+k][i], nx_tmp = nx[k][i];
            vl_sq += vl_tmp*vl_tmp;
            vr_sq += vr_tmp*vr_tmp;
            Vnl += vl_tmp*nx_tmp;
            Vnr += vr_tmp*nx_tmp;
            wn += fv[k][i]*nx_tmp;
        }
        if (boost) wn = 0.;
        double sl, sc, sr;
        orc_get_waves(_dl, Vnl, _pl, _dr, Vnr, _pr, gamma, &sl, &sc, &sr);
        if (wn <= sl) {
            flux[0][i] = _dl*(Vnl - wn);
            flux[dim+1][i] = (0.5*_dl*vl_sq + _pl/(gamma - 1.0))*(Vnl - wn) + _pl*Vnl;
            for (int k = 0; k < dim; k++)
                flux[1+k][i] = _dl*sL[1+k][i]*(Vnl - wn) + _pl*nx[k][i];
        } else if ((sl < wn) && (wn <= sr)) {
            double fac1 = sr - wn, fac2 = sl - wn, fac3 = sr - sl;
            flux[0][i] = (_dl*Vnl*fac1 - _dr*Vnr*fac2 - sl*_dl*fac1 + sr*_dr*fac2)/fac3;
            for (int k = 0; k < dim; k++)
                flux[1+k][i] = ((_dl*sL[1+k][i]*Vnl + _pl*nx[k][i])*fac1
                                - (_dr*sR[1+k][i]*Vnr + _pr*nx[k][i])*fac2
                                - sl*(_dl*sL[1+k][i])*fac1 + sr*(_dr*sR[1+k][i])*fac2)/fac3;
            double el = 0.5*_dl*vl_sq + _pl/(gamma - 1.0);
            double er = 0.5*_dr*vr_sq + _pr/(gamma - 1.0);
            flux[dim+1][i] = ((el + _pl)*Vnl*fac1 - (er + _pr)*Vnr*fac2
                              - sl*el*fac1 + sr*er*fac2)/fac3;
        } else {
            flux[0][i] = _dr*(Vnr - wn);
            flux[dim+1][i] = (0.5*_dr*vr_sq + _pr/(gamma - 1.0))*(Vnr - wn) + _pr*Vnr;
            for (int k = 0; k < dim; k++)
                flux[1+k][i] = _dr*sR[1+k][i]*(Vnr - wn) + _pr*nx[k][i];
        }
    }
    if (boost) deboost(dim, nf, fv, flux);
}

/* HLLC.riemann_solver (riemann.pyx:521-682) */
void orc_hllc(int dim, int nf, double gamma, int boost, double* const* sL,
              double* const* sR, double* const* nx, double* const* fv,
              double* const* flux)
{
    ORC_PFOR()
    for (int i = 0; i < nf; i++) {
        double _dl = sL[0][i], _pl = sL[dim+1][i];
        double _dr = sR[0][i], _pr = sR[dim+1][i];
        /* cl, cr at 593-594 are never read */
        double Vnl = 0.0, Vnr = 0.0, vl_sq = 0.0, vr_sq = 0.0, wn = 0.0;
        for (int k = 0; k < dim; k++) {
            double vlk = sL[1+k][i], vrk = sR[1+k][i], nk = nx[k][i];
            vl_sq += vlk*vlk;
            vr_sq += vrk*vrk;
            Vnl += vlk*nk;
            Vnr += vrk*nk;
            wn += fv[k][i]*nk;
        }
        if (boost == 1) wn = 0.;
        double sl, s_contact, sr;
        orc_get_waves(_dl, Vnl, _pl, _dr, Vnr, _pr, gamma, &sl, &s_contact, &sr);

        if (wn <= sl) {
            flux[0][i] = _dl*(Vnl - wn);
            flux[dim+1][i] = (0.5*_dl*vl_sq + _pl/(gamma - 1.0))*(Vnl - wn) + _pl*Vnl;
            for (int k = 0; k < dim; k++)
                flux[1+k][i] = _dl*sL[1+k][i]*(Vnl - wn) + _pl*nx[k][i];
        } else if ((sl < wn) && (wn <= sr)) {
            if (wn <= s_contact) {
                double factor_1 = _dl*(sl - Vnl)/(sl - s_contact);
                double factor_2 = factor_1*(sl - wn)*(s_contact - Vnl) + _pl;
                double frho = _dl*(Vnl - sl) + factor_1*(sl - wn);
                double el = 0.5*_dl*vl_sq + _pl/(gamma-1.0);
                flux[0][i] = frho;
                flux[dim+1][i] = (el + _pl)*Vnl - sl*el +
                    (sl - wn)*factor_1*(el/_dl + (s_contact - Vnl)*
                    (s_contact + _pl/(_dl*(sl - Vnl))));
                for (int k = 0; k < dim; k++)
                    flux[1+k][i] = frho*sL[1+k][i] + factor_2*nx[k][i];
            } else {
                double factor_1 = _dr*(sr - Vnr)/(sr - s_contact);
                double factor_2 = factor_1*(sr - wn)*(s_contact - Vnr) + _pr;
                double frho = _dr*(Vnr - sr) + factor_1*(sr - wn);
                double er = 0.5*_dr*vr_sq + _pr/(gamma-1.0);
                flux[0][i] = frho;
                flux[dim+1][i] = (er + _pr)*Vnr - sr*er +
                    (sr - wn)*factor_1*(er/_dr + (s_contact - Vnr)*
                    (s_contact + _pr/(_dr*(sr - Vnr))));
                for (int k = 0; k < dim; k++)
                    flux[1+k][i] = frho*sR[1+k][i] + factor_2*nx[k][i];
            }
        } else {
            flux[0][i] = _dr*(Vnr - wn);
            flux[dim+1][i] = (0.5*_dr*vr_sq + _pr/(gamma - 1.0))*(Vnr - wn) + _pr*Vnr;
            for (int k = 0; k < dim; k++)
                flux[1+k][i] = _dr*sR[1+k][i]*(Vnr - wn) + _pr*nx[k][i];
        }
    }
    if (boost) deboost(dim, nf, fv, flux);
}

/* ------------------------------------------------------------------------ */
/* Exact Riemann solver (riemann.pyx:684-1090), Toro chapter 4.             */
static double ex_p_guess(double dl, double ul, double pl, double cl, double dr,
                         double ur, double pr, double cr, double gamma)
{
    /* p_guess (riemann.pyx:975-1011) */
    double ppv = .5*(pl + pr) - .125*(ur - ul)*(dl + dr)*(cl + cr);
    double p_max = (pl > pr) ? pl : pr;        /* max(pl, pr) */
    double p_min = (pl < pr) ? pl : pr;        /* min(pl, pr) */
    double q_max = p_max/p_min;
    double p0;
    if ((q_max <= 2.) && (p_min <= ppv && ppv <= p_max)) {
        p0 = ppv;
    } else if (ppv <= p_min) {
        double p_lr = pow(pl/pr, (gamma - 1.)/(2.*gamma));
        double u_star = (p_lr*ul/cl + ur/cr + 2.*(p_lr - 1.)/(gamma - 1.));
        u_star = u_star/(p_lr/cl + 1./cr);
        double p_tl = pow(1. + (gamma - 1.)*(ul - u_star)/(2.*cl), 2.*gamma/(gamma - 1.));
        double p_tr = pow(1. + (gamma - 1.)*(u_star - ur)/(2.*cr), 2.*gamma/(gamma - 1.));
        p0 = .5*(pl*p_tl + pr*p_tr);
    } else {
        double gl = sqrt((2./(dl*(gamma + 1.)))/((gamma - 1.)*pl/(gamma + 1.) + ppv));
        double gr = sqrt((2./(dr*(gamma + 1.)))/((gamma - 1.)*pr/(gamma + 1.) + ppv));
        p0 = (gl*pl + gr*pr - (ur - ul))/(gr + gl);
    }
    return p0;
}

static double ex_p_func(double d, double p, double c, double gamma, double p_old)
{
    /* p_func (riemann.pyx:1014-1032) */
    if (p_old <= p)
        return 2.*c/(gamma - 1.)*(pow(p_old/p, (gamma - 1.)/(2.*gamma)) - 1.);
    double Ak = 2./(d*(gamma + 1.));
    double Bk = p*(gamma - 1.)/(gamma + 1.);
    return (p_old - p)*sqrt(Ak/(p_old + Bk));
}

static double ex_p_func_deriv(double d, double p, double c, double gamma, double p_old)
{
    /* p_func_deriv (riemann.pyx:1035-1054) */
    if (p_old <= p)
        return pow(p_old/p, -(gamma + 1.)/(2.*gamma))/(c*d);
    double Ak = 2./(d*(gamma + 1.));
    double Bk = p*(gamma - 1.)/(gamma + 1.);
    return sqrt(Ak/(p_old + Bk))*(1. - .5*(p_old - p)/(Bk + p_old));
}

static int ex_get_pstar(double dl, double ul, double pl, double cl, double dr,
                        double ur, double pr, double cr, double gamma, double* out)
{
    /* get_pstar (riemann.pyx:1057-1090): Newton iteration, TOL 1e-6 */
    const double TOL = 1.0e-6;
    double p_old = ex_p_guess(dl, ul, pl, cl, dr, ur, pr, cr, gamma);
    for (int it = 0; it < 1000; it++) {
        double fl = ex_p_func(dl, pl, cl, gamma, p_old);
        double fr = ex_p_func(dr, pr, cr, gamma, p_old);
        double dfl = ex_p_func_deriv(dl, pl, cl, gamma, p_old);
        double dfr = ex_p_func_deriv(dr, pr, cr, gamma, p_old);
        double p_new = p_old - (fl + fr + ur - ul)/(dfl + dfr);
        if ((2.*fabs((p_new - p_old)/(p_new + p_old))) <= TOL) { *out = p_new; return 0; }
        if (p_new < 0.) p_new = TOL;
        p_old = p_new;
    }
    return ORC_ERR_EXACT;
}

static void ex_state_flux(int dim, int i, double gamma, double _d, double _p,
                          const double* v, double* const* nx, double* const* flux)
{
    double v_sq = 0., vn = 0.;
    for (int k = 0; k < dim; k++) {
        vn += v[k]*nx[k][i];
        v_sq += v[k]*v[k];
    }
    flux[0][i] = _d*vn;
    flux[dim+1][i] = (0.5*_d*v_sq + gamma*_p/(gamma - 1.0))*vn;
    for (int k = 0; k < dim; k++)
        flux[1+k][i] = _d*v[k]*vn + _p*nx[k][i];
}

/* Exact.riemann_solver (riemann.pyx:700-972); always deboosts (972). */
int orc_exact(int dim, int nf, double gamma, double* const* sL, double* const* sR,
              double* const* nx, double* const* fv, double* const* flux,
              int32_t* bad_face)
{
    double v[3];
    for (int i = 0; i < nf; i++) {
        double _dl = sL[0][i], _pl = sL[dim+1][i];
        double _dr = sR[0][i], _pr = sR[dim+1][i];
        double cl = sqrt(gamma*_pl/_dl);
        double cr = sqrt(gamma*_pr/_dr);
        double vnl = 0.0, vnr = 0.0, vl_sq = 0.0, vr_sq = 0.0;
        for (int k = 0; k < dim; k++) {
            vnl += sL[1+k][i]*nx[k][i];
            vnr += sR[1+k][i]*nx[k][i];
            vl_sq += sL[1+k][i]*sL[1+k][i];
            vr_sq += sR[1+k][i]*sR[1+k][i];
        }
        double p_star;
        if (ex_get_pstar(_dl, vnl, _pl, cl, _dr, vnr, _pr, cr, gamma, &p_star)) {
            if (bad_face) *bad_face = i;
            return ORC_ERR_EXACT;
        }
        double fl = ex_p_func(_dl, _pl, cl, gamma, p_star);
        double fr = ex_p_func(_dr, _pr, cr, gamma, p_star);
        double u_star = 0.5*(vnl + vnr + fr - fl);

        if (0.0 <= u_star) {
            if (p_star <= _pl) {
                double s_hl = vnl - cl;
                if (0.0 <= s_hl) {
                    flux[0][i] = _dl*vnl;
                    flux[dim+1][i] = (0.5*_dl*vl_sq + gamma*_pl/(gamma - 1.0))*vnl;
                    for (int k = 0; k < dim; k++)
                        flux[1+k][i] = _dl*sL[1+k][i]*vnl + _pl*nx[k][i];
                } else {
                    double c_star_l = cl*pow(p_star/_pl, (gamma - 1.0)/(2.0*gamma));
                    double s_tl = u_star - c_star_l;
                    if (0.0 >= s_tl) {
                        double _d = _dl*pow(p_star/_pl, 1.0/gamma);
                        for (int k = 0; k < dim; k++)
                            v[k] = sL[1+k][i] + (u_star - vnl)*nx[k][i];
                        ex_state_flux(dim, i, gamma, _d, p_star, v, nx, flux);
                    } else {
                        double c = (2.0/(gamma + 1.0))*(cl + 0.5*(gamma - 1.0)*vnl);
                        double _d = _dl*pow(c/cl, 2.0/(gamma - 1.0));
                        double _p = _pl*pow(c/cl, 2.0*gamma/(gamma - 1.0));
                        for (int k = 0; k < dim; k++)
                            v[k] = sL[1+k][i] + (c - vnl)*nx[k][i];
                        ex_state_flux(dim, i, gamma, _d, _p, v, nx, flux);
                    }
                }
            } else {
                double sl = vnl - cl*sqrt((gamma + 1.0)*p_star/(2.0*gamma*_pl)
                                          + (gamma - 1.0)/(2.0*gamma));
                if (0.0 <= sl) {
                    flux[0][i] = _dl*vnl;
                    flux[dim+1][i] = (0.5*_dl*vl_sq + gamma*_pl/(gamma - 1.0))*vnl;
                    for (int k = 0; k < dim; k++)
                        flux[1+k][i] = _dl*sL[1+k][i]*vnl + _pl*nx[k][i];
                } else {
                    double _d = _dl*(p_star/_pl + (gamma - 1.0)/(gamma + 1.0))
                                /(p_star*(gamma - 1.0)/((gamma + 1.0)*_pl) + 1.0);
                    for (int k = 0; k < dim; k++)
                        v[k] = sL[1+k][i] + (u_star - vnl)*nx[k][i];
                    ex_state_flux(dim, i, gamma, _d, p_star, v, nx, flux);
                }
            }
        } else {
            if (p_star >= _pr) {
                double sr = vnr + cr*sqrt((gamma + 1.0)*p_star/(2.0*gamma*_pr)
                                          + (gamma-1.0)/(2.0*gamma));
                if (0.0 >= sr) {
                    flux[0][i] = _dr*vnr;
                    flux[dim+1][i] = (0.5*_dr*vr_sq + gamma*_pr/(gamma - 1.0))*vnr;
                    for (int k = 0; k < dim; k++)
                        flux[1+k][i] = _dr*sR[1+k][i]*vnr + _pr*nx[k][i];
                } else {
                    double _d = _dr*(p_star/_pr + (gamma - 1.0)/(gamma + 1.0))
                                /(p_star*(gamma - 1.0)/((gamma + 1.0)*_pr) + 1.0);
                    for (int k = 0; k < dim; k++)
                        v[k] = sR[1+k][i] + (u_star - vnr)*nx[k][i];
                    ex_state_flux(dim, i, gamma, _d, p_star, v, nx, flux);
                }
            } else {
                double s_hr = vnr + cr;
                if (0.0 >= s_hr) {
                    flux[0][i] = _dr*vnr;
                    flux[dim+1][i] = (0.5*_dr*vr_sq + gamma*_pr/(gamma - 1.0))*vnr;
                    for (int k = 0; k < dim; k++)
                        flux[1+k][i] = _dr*sR[1+k][i]*vnr + _pr*nx[k][i];
                } else {
                    double c_star_r = cr*pow(p_star/_pr, (gamma-1.0)/(2.0*gamma));
                    double s_tr = u_star + c_star_r;
                    if (0.0 <= s_tr) {
                        double _d = _dr*pow(p_star/_pr, 1.0/gamma);
                        for (int k = 0; k < dim; k++)
                            v[k] = sR[1+k][i] + (u_star - vnr)*nx[k][i];
                        ex_state_flux(dim, i, gamma, _d, p_star, v, nx, flux);
                    } else {
                        double c = (2.0/(gamma + 1.0))*(cr - 0.5*(gamma - 1.0)*vnr);
                        double u_tmp = (2.0/(gamma + 1.0))*(-cr + 0.5*(gamma-1.0)*vnr);
                        double _d = _dr*pow(c/cr, 2.0/(gamma - 1.0));
                        double _p = _pr*pow(c/cr, 2.0*gamma/(gamma - 1.0));
                        for (int k = 0; k < dim; k++)
                            v[k] = sR[1+k][i] + (u_tmp - vnr)*nx[k][i];
                        ex_state_flux(dim, i, gamma, _d, _p, v, nx, flux);
                    }
                }
            }
        }
    }
    deboost(dim, nf, fv, flux);
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Mesh.update_from_fluxes (mesh.pyx:479-551).  Returns ORC_ERR_MASS /
 * ORC_ERR_ENERGY with *bad_face at the first face whose running endpoint
 * values fail the positivity check, leaving the arrays as the reference
 * leaves them at its raise. */
int orc_update_from_fluxes(int dim, int nf, double dt, const int8_t* tag,
                           const double* area, const int32_t* pair_i,
                           const int32_t* pair_j, double* const* flux,
                           double* const* cons, int32_t* bad_face)
{
    double* mm = cons[0];
    double* ee = cons[dim+1];
    for (int n = 0; n < nf; n++) {
        int i = pair_i[n], j = pair_j[n];
        double a = area[n];
        if (tag[i] == REAL) {
            mm[i] -= dt*a*flux[0][n];
            ee[i] -= dt*a*flux[dim+1][n];
            for (int k = 0; k < dim; k++)
                cons[1+k][i] -= dt*a*flux[1+k][n];
        }
        if (tag[j] == REAL) {
            mm[j] += dt*a*flux[0][n];
            ee[j] += dt*a*flux[dim+1][n];
            for (int k = 0; k < dim; k++)
                cons[1+k][j] += dt*a*flux[1+k][n];
        }
        if (mm[i] <= 0.0 || mm[j] <= 0.0) { if (bad_face) *bad_face = n; return ORC_ERR_MASS; }
        if (ee[i] <= 0.0 || ee[j] <= 0.0) { if (bad_face) *bad_face = n; return ORC_ERR_ENERGY; }
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Core face-flux step on a frozen mesh, in the order of
 * MovingMeshMUSCLHancock.evolve_timestep (integrate.py:396-408):
 * gradients + limiter, ghost gradients, spatial + temporal states at dt/2,
 * HLL, HLLC or Exact + deboost, conserved update.  Used as the timed CPU
 * baseline; states and fluxes live in caller-provided scratch. */
int orc_core_step(int dim, int n, int nf, double gamma, double dt, int solver,
                  int boost, int gizmo, int limiter, int boundary,
                  const int8_t* tag, const int8_t* type, const int32_t* map,
                  double* const* x, double* const* dcx, const double* vol,
                  double* const* prim, double* const* cons,
                  const double* area, double* const* fij, double* const* nx,
                  double* const* fv, const int32_t* pair_i,
                  const int32_t* pair_j, const int32_t* nbr_off,
                  const int32_t* nbr_face, const double* lo, const double* hi,
                  double* const* grad, double* const* pl, double* const* pr,
                  double* const* flux, int32_t* bad)
{
    int rc = orc_compute_gradients(dim, n, tag, x, dcx, vol, prim, area, fij,
                                   pair_i, pair_j, nbr_off, nbr_face, limiter,
                                   grad, bad);
    if (rc) return rc;
    orc_ghost_gradients(dim, n, tag, type, map, x, lo, hi, boundary, grad);
    orc_add_spatial(dim, nf, prim, x, dcx, grad, fij, fv, pair_i, pair_j, boost,
                    gizmo, pl, pr);
    orc_add_temporal(dim, nf, gamma, 0.5*dt, prim, grad, fv, pair_i, pair_j,
                     boost, pl, pr);
    /* riemann.compute_fluxes (riemann.pyx:74-98): HLL 0, HLLC 1, Exact 2;
     * the Exact solver raises before the update when Newton fails */
    if (solver == 0) orc_hll(dim, nf, gamma, boost, pl, pr, nx, fv, flux);
    else if (solver == 1) orc_hllc(dim, nf, gamma, boost, pl, pr, nx, fv, flux);
    else {
        rc = orc_exact(dim, nf, gamma, pl, pr, nx, fv, flux, bad);
        if (rc) return rc;
    }
    return orc_update_from_fluxes(dim, nf, dt, tag, area, pair_i, pair_j, flux,
                                  cons, bad);
}

/* ------------------------------------------------------------------------ */
/* Geometry of a Voronoi mesh from its Delaunay topology: the face set
 * (pair_i, pair_j), ghost images (ghost_map, ghost_code) and the ring of
 * each face -- the generators whose bisectors bound it, in polygon order
 * (include/phd_mesh.h, phdm_fill_topology).  Follows
 * Tess2d::extract_geometry (tess.cpp:238-398) and
 * Tess3d::extract_geometry (tess3.cpp:334-562):
 *   - Voronoi vertices are circumcentres of Delaunay triangles/tetrahedra
 *     (CGAL's tess.dual(cc), tess3.cpp:435-437);
 *   - 2D: area = |p2 - p1|, com = midpoint, volume += area h / 4, centroid
 *     += (area h / 4)(2 f / 3 + x / 3) (tess.cpp:313-355), dcom = c / vol -
 *     x (tess.cpp:389-393), normal = (x_j - x_i) / h;
 *   - 3D: fan about the vertex mean c (tess3.cpp:452-483), the zero-area
 *     rule (tess3.cpp:494-502), com = sum(c3 a3) / area1, area = area1 / 2,
 *     normal = (x_j - x_i) * (1 / |x_j - x_i|) (tess3.cpp:504-517),
 *     volume and centroid from the tetrahedra (face centroid, v1, v2,
 *     generator) (tess3.cpp:537-559);
 *   - ghost positions: periodic x + (index - 1) translate
 *     (boundary.pyx:378-380), reflective x - 2 (x - edge)
 *     (boundary.pyx:103-104, 125-126); ghost volume and dcom are copies of
 *     the image's (mesh.pyx:141-142, 323).
 * Two choices the reference leaves to CGAL are fixed here: every face is
 * computed once, in the frame of its owner generator x_o (o = pair_i), and
 * feeds both cells -- cell o with the apex at 0, cell j with the apex at
 * x_j - x_o; and a cell's volume and centroid sum its faces' contributions
 * in ascending face id (the order of its face list), the centroid relative
 * to its own generator.  The reference's zero-area test compares
 * sqrt(area1) with the distance of the unnormalised centroid sum from the
 * generator, in the owner frame |sum c3 a3|; a face with area1 == 0 is
 * zero as well (the reference reaches the same via its absolute frame).  A
 * zero face keeps its place in the face list with area 0 and com = x_o + c
 * and adds nothing to the volumes.
 * A circumcentre of 0, a, b (2D) / 0, a, b, c (3D) solves p.q = |q|^2 / 2:
 *   2D: p = (|a|^2 (b_y, -b_x) + |b|^2 (-a_y, a_x)) / (2 (a_x b_y - a_y b_x))
 *   3D: p = (|a|^2 (b x c) + |b|^2 (c x a) + |c|^2 (a x b)) / (2 a.(b x c))
 * The per-face contributions live in ws: 2 sides x (1 + dim) doubles per
 * face (volume term, then the centroid moments). */

/* image `code` of a generator coordinate xk on axis a */
static double orc_image(int boundary, const double* lo, const double* hi,
                        double xk, int a, int code)
{
    if (code == 0) return xk;
    if (boundary == 1) {
        int idx = code - 1, b;
        for (b = 0; b < a; b++) idx /= 3;
        return xk + (double)(idx % 3 - 1)*(hi[a] - lo[a]);
    }
    if ((code - 1) >> 1 != a) return xk;
    {
        const double edge = ((code - 1) & 1) ? hi[a] : lo[a];
        return xk - 2*(xk - edge);
    }
}

/* ring member r relative to xo */
static void orc_ring_pos(int dim, int boundary, const double* lo, const double* hi,
                         double* const* x, int32_t k, int code, const double* xo,
                         double* p)
{
    int a;
    for (a = 0; a < dim; a++)
        p[a] = orc_image(boundary, lo, hi, x[a][k], a, code) - xo[a];
}

static void orc_circ2(const double* a, const double* b, double* p)
{
    const double aa = a[0]*a[0] + a[1]*a[1], bb = b[0]*b[0] + b[1]*b[1];
    const double d = a[0]*b[1] - a[1]*b[0];
    const double r = 1.0/(d + d);
    p[0] = (aa*b[1] - bb*a[1])*r;
    p[1] = (bb*a[0] - aa*b[0])*r;
}

static void orc_cross(const double* u, const double* v, double* w)
{
    w[0] = u[1]*v[2] - u[2]*v[1];
    w[1] = u[2]*v[0] - u[0]*v[2];
    w[2] = u[0]*v[1] - u[1]*v[0];
}

static void orc_circ3(const double* a, const double* b, const double* c, double* p)
{
    double bc[3], ca[3], ab[3];
    int k;
    const double aa = a[0]*a[0] + a[1]*a[1] + a[2]*a[2];
    const double bb = b[0]*b[0] + b[1]*b[1] + b[2]*b[2];
    const double cc = c[0]*c[0] + c[1]*c[1] + c[2]*c[2];
    orc_cross(b, c, bc);
    orc_cross(c, a, ca);
    orc_cross(a, b, ab);
    {
        const double det = a[0]*bc[0] + a[1]*bc[1] + a[2]*bc[2];
        const double r = 1.0/(det + det);
        for (k = 0; k < 3; k++) p[k] = (aa*bc[k] + bb*ca[k] + cc*ab[k])*r;
    }
}

/* vertex t of 3D face f (ring r0 .. r0+m), relative to xo */
static void orc_vertex3(int boundary, const double* lo, const double* hi,
                        double* const* x, const int32_t* ring_k,
                        const int8_t* ring_code, int r0, int m, int t,
                        const double* xo, const double* a, double* w)
{
    double p[3], q[3];
    const int u = r0 + t, v = r0 + (t + 1 == m ? 0 : t + 1);
    orc_ring_pos(3, boundary, lo, hi, x, ring_k[u], ring_code[u], xo, p);
    orc_ring_pos(3, boundary, lo, hi, x, ring_k[v], ring_code[v], xo, q);
    orc_circ3(a, p, q, w);
}

static void orc_face_geometry(int dim, int boundary, const double* lo,
                              const double* hi, double* const* x,
                              const int32_t* pair_i, const int32_t* pair_j,
                              const int32_t* ring_off, const int32_t* ring_k,
                              const int8_t* ring_code, int f, double* area,
                              double* const* com, double* const* normal,
                              double* ws)
{
    const int o = pair_i[f], j = pair_j[f];
    const int r0 = ring_off[f], m = ring_off[f + 1] - ring_off[f];
    const int W = 1 + dim;
    double xo[3], a[3];
    int k, t;
    for (k = 0; k < dim; k++) {
        xo[k] = x[k][o];
        a[k] = x[k][j] - xo[k];
    }
    if (dim == 2) {
        double p[2], w0[2], w1[2], fm[2];
        double xe, ye, ar, h, wt;
        orc_ring_pos(2, boundary, lo, hi, x, ring_k[r0], ring_code[r0], xo, p);
        orc_circ2(a, p, w0);
        orc_ring_pos(2, boundary, lo, hi, x, ring_k[r0 + 1], ring_code[r0 + 1], xo, p);
        orc_circ2(a, p, w1);
        h = sqrt(a[0]*a[0] + a[1]*a[1]);
        xe = w1[0] - w0[0];
        ye = w1[1] - w0[1];
        ar = sqrt(xe*xe + ye*ye);
        fm[0] = 0.5*(w0[0] + w1[0]);
        fm[1] = 0.5*(w0[1] + w1[1]);
        wt = 0.25*ar*h;
        area[f] = ar;
        for (k = 0; k < 2; k++) {
            normal[k][f] = a[k]/h;
            com[k][f] = xo[k] + fm[k];
            ws[(size_t)f*2*W + 1 + k] = wt*(2.0*fm[k]/3.0);
            ws[(size_t)f*2*W + W + 1 + k] = wt*(2.0*(fm[k] - a[k])/3.0);
        }
        ws[(size_t)f*2*W] = wt;
        ws[(size_t)f*2*W + W] = wt;
        return;
    }
    {
        const double third = 1.0/3.0, fourth = 0.25;
        double c[3] = {0.0, 0.0, 0.0}, s[3] = {0.0, 0.0, 0.0};
        double w[3], v1[3], v2[3], n3[3], fc[3], cvo[3], cvj[3];
        double area1 = 0.0, vo = 0.0, vj = 0.0;
        double mo[3] = {0.0, 0.0, 0.0}, mj[3] = {0.0, 0.0, 0.0};
        const double hn = 1.0/sqrt(a[0]*a[0] + a[1]*a[1] + a[2]*a[2]);
        for (k = 0; k < 3; k++) normal[k][f] = a[k]*hn;
        /* vertex mean */
        for (t = 0; t < m; t++) {
            orc_vertex3(boundary, lo, hi, x, ring_k, ring_code, r0, m, t, xo, a, w);
            for (k = 0; k < 3; k++) c[k] += w[k];
        }
        for (k = 0; k < 3; k++) c[k] *= 1.0/m;
        /* fan about c: area and unnormalised centroid */
        orc_vertex3(boundary, lo, hi, x, ring_k, ring_code, r0, m, m - 1, xo, a, w);
        for (k = 0; k < 3; k++) v1[k] = w[k] - c[k];
        for (t = 0; t < m; t++) {
            double a3;
            orc_vertex3(boundary, lo, hi, x, ring_k, ring_code, r0, m, t, xo, a, w);
            for (k = 0; k < 3; k++) v2[k] = w[k] - c[k];
            orc_cross(v1, v2, n3);
            a3 = sqrt(n3[0]*n3[0] + n3[1]*n3[1] + n3[2]*n3[2]);
            for (k = 0; k < 3; k++) s[k] += (c[k] + (v1[k] + v2[k])*third)*a3;
            area1 += a3;
            for (k = 0; k < 3; k++) v1[k] = v2[k];
        }
        {
            const double l1 = sqrt(area1);
            const double l2 = sqrt(s[0]*s[0] + s[1]*s[1] + s[2]*s[2]);
            if (area1 == 0.0 || l1 < 1.0e-10*l2) {
                area[f] = 0.0;
                for (k = 0; k < 3; k++) com[k][f] = xo[k] + c[k];
                for (k = 0; k < 2*W; k++) ws[(size_t)f*2*W + k] = 0.0;
                return;
            }
        }
        for (k = 0; k < 3; k++) fc[k] = s[k]*(1.0/area1);
        area[f] = 0.5*area1;
        for (k = 0; k < 3; k++) {
            com[k][f] = xo[k] + fc[k];
            cvo[k] = 0.0 - fc[k];
            cvj[k] = a[k] - fc[k];
        }
        /* tetrahedra (face centroid, v1, v2, apex) for both cells */
        orc_vertex3(boundary, lo, hi, x, ring_k, ring_code, r0, m, m - 1, xo, a, w);
        for (k = 0; k < 3; k++) v1[k] = w[k] - fc[k];
        for (t = 0; t < m; t++) {
            double cr[3], v4;
            orc_vertex3(boundary, lo, hi, x, ring_k, ring_code, r0, m, t, xo, a, w);
            for (k = 0; k < 3; k++) v2[k] = w[k] - fc[k];
            orc_cross(v1, v2, cr);
            v4 = fabs(cr[0]*cvo[0] + cr[1]*cvo[1] + cr[2]*cvo[2]);
            vo += v4;
            for (k = 0; k < 3; k++) mo[k] += (fc[k] + (v1[k] + v2[k] + cvo[k])*fourth)*v4;
            v4 = fabs(cr[0]*cvj[0] + cr[1]*cvj[1] + cr[2]*cvj[2]);
            vj += v4;
            for (k = 0; k < 3; k++)
                mj[k] += ((fc[k] + (v1[k] + v2[k] + cvj[k])*fourth) - a[k])*v4;
            for (k = 0; k < 3; k++) v1[k] = v2[k];
        }
        ws[(size_t)f*2*W] = vo;
        ws[(size_t)f*2*W + W] = vj;
        for (k = 0; k < 3; k++) {
            ws[(size_t)f*2*W + 1 + k] = mo[k];
            ws[(size_t)f*2*W + W + 1 + k] = mj[k];
        }
    }
}

void orc_extract_geometry(int dim, int boundary, const double* lo, const double* hi,
                          int n, int ng, int nf, double* const* x,
                          const int32_t* ghost_map, const int8_t* ghost_code,
                          const int32_t* pair_i, const int32_t* pair_j,
                          const int32_t* ring_off, const int32_t* ring_k,
                          const int8_t* ring_code, const int32_t* nbr_off,
                          const int32_t* nbr_face, double* volume,
                          double* const* dcom, double* area, double* const* com,
                          double* const* normal, double* ws)
{
    const int W = 1 + dim;
    int g, f, i;
    ORC_PFOR()
    for (g = 0; g < ng; g++) {
        int a;
        for (a = 0; a < dim; a++)
            x[a][n + g] = orc_image(boundary, lo, hi, x[a][ghost_map[g]], a, ghost_code[g]);
    }
    ORC_PFOR()
    for (f = 0; f < nf; f++)
        orc_face_geometry(dim, boundary, lo, hi, x, pair_i, pair_j, ring_off,
                          ring_k, ring_code, f, area, com, normal, ws);
    ORC_PFOR()
    for (i = 0; i < n; i++) {
        double v = 0.0, mm[3] = {0.0, 0.0, 0.0};
        int e, k;
        for (e = nbr_off[i]; e < nbr_off[i + 1]; e++) {
            const int fe = nbr_face[e];
            const double* q = ws + (size_t)fe*2*W + (pair_i[fe] == i ? 0 : W);
            v += q[0];
            for (k = 0; k < dim; k++) mm[k] += q[1 + k];
        }
        if (dim == 2) {
            volume[i] = v;
            for (k = 0; k < 2; k++) dcom[k][i] = mm[k]/v;
        } else {
            for (k = 0; k < 3; k++) dcom[k][i] = mm[k]*(1.0/v);
            volume[i] = v*(1.0/6.0);
        }
    }
    ORC_PFOR()
    for (g = 0; g < ng; g++) {
        int k;
        volume[n + g] = volume[ghost_map[g]];
        for (k = 0; k < dim; k++) dcom[k][n + g] = dcom[k][ghost_map[g]];
    }
}
