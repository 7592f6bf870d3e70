"""GPU entry points on empty inputs, through the C ABI.

A mesh with no cells and no faces is a legal input: the reference's loops
run zero times, so every stage returns 0, writes nothing and launches no
kernel (a launch of zero blocks would be a HIP error).  Zero-length arrays
have null data pointers, and a null cell-record array is accepted when there
are no cells.  The time step of an empty mesh is an argument error: the
reference seeds its minimum from cell 0 (riemann.pyx:120-188).  Afterwards
the device error word is clear.
"""
import ctypes

import pytest
import torch

pytestmark = pytest.mark.gpu


def _empty_mesh(dim, dev):
    from phd import _native as nat
    keep = []

    def ptr(dtype=torch.float64):
        t = torch.empty(0, dtype=dtype, device=dev)
        keep.append(t)
        return t.data_ptr() or None

    c = nat.Cells()
    c.dim, c.n_cells = dim, 0
    c.tag, c.type, c.map = ptr(torch.int8), ptr(torch.int8), ptr(torch.int32)
    for k in range(dim):
        c.x[k], c.dcom[k], c.w[k] = ptr(), ptr(), ptr()
    c.volume = ptr()
    for q in range(dim + 2):
        c.prim[q], c.cons[q] = ptr(), ptr()
    f = nat.Faces()
    f.n_faces = 0
    f.area = ptr()
    for k in range(dim):
        f.com[k], f.normal[k], f.velocity[k] = ptr(), ptr(), ptr()
    f.pair_i, f.pair_j = ptr(torch.int32), ptr(torch.int32)
    off = torch.zeros(1, dtype=torch.int32, device=dev)
    keep.append(off)
    f.nbr_off, f.nbr_face = off.data_ptr(), ptr(torch.int32)
    return c, f, keep


@pytest.mark.parametrize("dim", [2, 3])
def test_empty_mesh_every_stage(gpu, dim):
    from phd import _native as nat
    print("case empty mesh dim=%d" % dim, flush=True)
    lib, h = nat.flux(), nat.handle()
    c, f, keep = _empty_mesh(dim, gpu)
    cr, fr = ctypes.byref(c), ctypes.byref(f)
    plan = nat.Plan(c, f)
    p = plan.ptr
    ld = nat.record_ld(dim)
    lo, hi = nat.DblArray3(0.0, 0.0, 0.0), nat.DblArray3(1.0, 1.0, 1.0)
    lo_, hi_ = ctypes.cast(lo, ctypes.c_void_p), ctypes.cast(hi, ctypes.c_void_p)
    none5 = nat.PtrArray5()
    g = 1.4
    calls = {
        "phd_conservative_from_primitive": lambda: lib.phd_conservative_from_primitive(h, cr, g),
        "phd_primitive_from_conservative": lambda: lib.phd_primitive_from_conservative(h, cr, g),
        "phd_assign_generator_velocities":
            lambda: lib.phd_assign_generator_velocities(h, cr, g, 0.25, 1),
        "phd_assign_face_velocities": lambda: lib.phd_assign_face_velocities(h, cr, fr),
        "phd_pack_cells": lambda: lib.phd_pack_cells(h, p, cr),
        "phd_compute_gradients": lambda: lib.phd_compute_gradients(h, p, cr, 0, None, ld),
        "phd_update_ghost_gradients":
            lambda: lib.phd_update_ghost_gradients(h, p, cr, nat.BOUNDARY_REFLECTIVE, lo_, hi_,
                                                   None, ld),
        "phd_compute_states":
            lambda: lib.phd_compute_states(h, p, cr, fr, None, ld, g, 0.01, 1, 1,
                                           nat.ORDER_LINEAR,
                                           nat.STAGE_SPATIAL | nat.STAGE_TEMPORAL, none5, none5),
        "phd_compute_fluxes":
            lambda: lib.phd_compute_fluxes(h, fr, dim, none5, none5, g, 1, nat.SOLVER_HLLC, none5),
        "phd_face_flux":
            lambda: lib.phd_face_flux(h, p, cr, fr, None, ld, g, 0.01, 1, 1, nat.ORDER_LINEAR,
                                      nat.SOLVER_HLLC, none5),
        "phd_update_from_fluxes": lambda: lib.phd_update_from_fluxes(h, p, cr, fr, none5, 0.01),
        "phd_gradients_packed": lambda: lib.phd_gradients_packed(h, p, cr, 0, None, ld),
        "phd_face_summands":
            lambda: lib.phd_face_summands(h, p, cr, fr, None, ld, g, 0.01, 1, 1, nat.SOLVER_HLLC,
                                          none5),
        "phd_update_from_summands": lambda: lib.phd_update_from_summands(h, p, cr),
        "phd_core_step":
            lambda: lib.phd_core_step(h, p, cr, fr, None, ld, nat.LIMITER_AREPO,
                                      nat.BOUNDARY_REFLECTIVE, lo_, hi_, g, 0.01, 1, 1,
                                      nat.SOLVER_HLLC, None, 0),
        "phd_ghost_refresh_eos":
            lambda: lib.phd_ghost_refresh_eos(h, p, cr, nat.BOUNDARY_REFLECTIVE, lo_, hi_, g),
        "phd_move_generators": lambda: lib.phd_move_generators(h, cr, 0.01),
        "phd_migrate_periodic": lambda: lib.phd_migrate_periodic(h, cr, lo_, hi_),
        "phd_reflect_fields": lambda: lib.phd_reflect_fields(h, cr, lo_, hi_),
        "phd_update_ghost_fields":
            lambda: lib.phd_update_ghost_fields(h, cr, (ctypes.c_void_p * 1)(c.volume or 1),
                                                (ctypes.c_int32 * 1)(8), 1),
    }
    failed = {name: rc for name, rc in ((n, call()) for n, call in calls.items()) if rc != 0}
    assert not failed, failed
    first = ctypes.c_int64(-2)
    assert lib.phd_first_outside(h, cr, lo_, hi_, ctypes.byref(first)) == 0 and first.value == -1
    dt = ctypes.c_double(-1.0)
    assert lib.phd_compute_time_step(h, cr, g, 0.3, 1, ctypes.byref(dt)) == nat.PHD_E_ARG
    torch.cuda.synchronize()
    nat.raise_device_fault(h)
    plan.close()
    del keep
