"""ctypes binding of the two native libraries of this package.

libphd_flux.so  -- the HIP (gfx950) face-flux library, C ABI in
                   include/phd_flux.h.  Every hot-path plugin method calls it;
                   there is no CPU fallback: a missing library or a machine
                   without a GPU raises NativeUnavailable.
libphd_mesh.so  -- the host Voronoi mesh builder, C ABI in include/phd_mesh.h.

Both are built in-tree by phd.build (hipcc / g++).
"""
import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
FLUX_LIB_PATH = os.path.join(_HERE, "libphd_flux.so")
MESH_LIB_PATH = os.path.join(_HERE, "libphd_mesh.so")

c_vp = ctypes.c_void_p
c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_dbl = ctypes.c_double

# constants mirrored from include/phd_flux.h
PHD_ABI_VERSION = 1
PHD_E_ARG, PHD_E_HIP, PHD_E_MODE, PHD_E_SIZE = -1, -2, -3, -4
PHD_F_NEIGHBORS, PHD_F_MASS, PHD_F_ENERGY, PHD_F_EXACT = 1, 2, 3, 4
SOLVER_HLL, SOLVER_HLLC, SOLVER_EXACT = 0, 1, 2
LIMITER_AREPO, LIMITER_TESS = 0, 1
BOUNDARY_REFLECTIVE, BOUNDARY_PERIODIC = 0, 1
MODE_BITWISE = 0
CORE_XP_FRESH = 1
VJP_REC_FRESH = 1
STAGE_SPATIAL, STAGE_TEMPORAL = 1, 2
ORDER_CONSTANT, ORDER_LINEAR = 0, 1

# fault messages of the reference (mesh.pyx:548,551; reconstruction.pyx:652;
# riemann.pyx:1090)
FAULT_MESSAGES = {
    PHD_F_NEIGHBORS: "ERROR: incorrect neighbors!",
    PHD_F_MASS: "Mass less than zero in flux update",
    PHD_F_ENERGY: "Energy less than zero in flux update",
    PHD_F_EXACT: "No convergence in Exact Riemann Solver",
}


class NativeUnavailable(RuntimeError):
    """The HIP library is not built or no GPU is visible."""


class Cells(ctypes.Structure):
    _fields_ = [("dim", c_i32), ("n_cells", c_i32), ("tag", c_vp), ("type", c_vp),
                ("map", c_vp), ("x", c_vp * 3), ("dcom", c_vp * 3), ("volume", c_vp),
                ("prim", c_vp * 5), ("cons", c_vp * 5), ("w", c_vp * 3)]


class Schedule(ctypes.Structure):
    _fields_ = [("grad_block", c_i32), ("face_block", c_i32), ("update_block", c_i32),
                ("xcd", c_i32)]


class Faces(ctypes.Structure):
    _fields_ = [("n_faces", c_i32), ("area", c_vp), ("com", c_vp * 3), ("normal", c_vp * 3),
                ("velocity", c_vp * 3), ("pair_i", c_vp), ("pair_j", c_vp),
                ("nbr_off", c_vp), ("nbr_face", c_vp)]


class Topology(ctypes.Structure):
    _fields_ = [("dim", c_i32), ("boundary", c_i32), ("n_real", c_i32), ("n_ghost", c_i32),
                ("n_faces", c_i32), ("lo", c_dbl * 3), ("hi", c_dbl * 3), ("ghost_map", c_vp),
                ("ghost_code", c_vp), ("pair_i", c_vp), ("pair_j", c_vp), ("ring_off", c_vp),
                ("ring_k", c_vp), ("ring_code", c_vp), ("nbr_off", c_vp), ("nbr_face", c_vp)]


class Geometry(ctypes.Structure):
    _fields_ = [("x", c_vp * 3), ("volume", c_vp), ("dcom", c_vp * 3), ("area", c_vp),
                ("com", c_vp * 3), ("normal", c_vp * 3)]


PtrArray5 = c_vp * 5
DblArray3 = c_dbl * 3

_FLUX_SIGS = {
    "phd_abi_version": ([], c_i32),
    "phd_error_string": ([c_i32], ctypes.c_char_p),
    "phd_handle_create": ([ctypes.POINTER(c_vp), c_i32], c_i32),
    "phd_handle_destroy": ([c_vp], c_i32),
    "phd_set_stream": ([c_vp, c_vp], c_i32),
    "phd_set_mode": ([c_vp, c_i32], c_i32),
    "phd_get_error": ([c_vp, ctypes.POINTER(c_i32), ctypes.POINTER(c_i64)], c_i32),
    "phd_plan_create": ([c_vp, c_vp, c_vp, ctypes.POINTER(c_vp)], c_i32),
    "phd_plan_destroy": ([c_vp], c_i32),
    "phd_plan_bytes": ([c_vp], c_i64),
    "phd_cell_record_ld": ([c_i32], c_i32),
    "phd_pack_cells": ([c_vp, c_vp, c_vp], c_i32),
    "phd_conservative_from_primitive": ([c_vp, c_vp, c_dbl], c_i32),
    "phd_primitive_from_conservative": ([c_vp, c_vp, c_dbl], c_i32),
    "phd_compute_time_step": ([c_vp, c_vp, c_dbl, c_dbl, c_i32, ctypes.POINTER(c_dbl)], c_i32),
    "phd_assign_generator_velocities": ([c_vp, c_vp, c_dbl, c_dbl, c_i32], c_i32),
    "phd_assign_face_velocities": ([c_vp, c_vp, c_vp], c_i32),
    "phd_compute_gradients": ([c_vp, c_vp, c_vp, c_i32, c_vp, c_i32], c_i32),
    "phd_update_ghost_gradients": ([c_vp, c_vp, c_vp, c_i32, c_vp, c_vp, c_vp, c_i32], c_i32),
    "phd_compute_states": ([c_vp, c_vp, c_vp, c_vp, c_vp, c_i32, c_dbl, c_dbl, c_i32, c_i32,
                            c_i32, c_i32, c_vp, c_vp], c_i32),
    "phd_compute_fluxes": ([c_vp, c_vp, c_i32, c_vp, c_vp, c_dbl, c_i32, c_i32, c_vp], c_i32),
    "phd_face_flux": ([c_vp, c_vp, c_vp, c_vp, c_vp, c_i32, c_dbl, c_dbl, c_i32, c_i32, c_i32,
                       c_i32, c_vp], c_i32),
    "phd_update_from_fluxes": ([c_vp, c_vp, c_vp, c_vp, c_vp, c_dbl], c_i32),
    "phd_core_step": ([c_vp, c_vp, c_vp, c_vp, c_vp, c_i32, c_i32, c_i32, c_vp, c_vp, c_dbl,
                       c_dbl, c_i32, c_i32, c_i32, c_vp, c_i32], c_i32),
    "phd_ghost_refresh_eos": ([c_vp, c_vp, c_vp, c_i32, c_vp, c_vp, c_dbl], c_i32),
    "phd_adjoint_create": ([c_vp, c_vp, c_vp, c_vp, ctypes.POINTER(c_vp)], c_i32),
    "phd_adjoint_destroy": ([c_vp], c_i32),
    "phd_core_step_vjp": ([c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_i32, c_i32, c_vp, c_vp, c_dbl,
                           c_dbl, c_i32, c_i32, c_vp, c_vp, c_i32], c_i32),
    "phd_eos_vjp": ([c_vp, c_vp, c_dbl, c_i32, c_vp, c_vp], c_i32),
    "phd_batch_create": ([c_vp, c_vp, c_vp, c_vp, c_i32, c_vp, c_i64, ctypes.POINTER(c_vp)],
                         c_i32),
    "phd_batch_destroy": ([c_vp], c_i32),
    "phd_batch_set_schedule": ([c_vp, c_vp], c_i32),
    "phd_batch_step": ([c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_i32, c_i32, c_vp, c_vp, c_dbl, c_dbl,
                        c_i32, c_i32, c_i32], c_i32),
    "phd_batch_capture": ([c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_i32, c_i32, c_vp, c_vp, c_dbl,
                           c_dbl, c_i32, c_i32], c_i32),
    "phd_batch_launch": ([c_vp, c_vp, c_i32], c_i32),
    "phd_batch_refresh": ([c_vp, c_vp, c_vp, c_vp, c_i32, c_vp, c_vp, c_dbl, c_i32], c_i32),
    "phd_batch_reset_history": ([c_vp, c_vp], c_i32),
    "phd_batch_read": ([c_vp, c_vp, c_vp, c_vp, c_vp, c_vp], c_i32),
    "phd_gradients_packed": ([c_vp, c_vp, c_vp, c_i32, c_vp, c_i32], c_i32),
    "phd_face_summands": ([c_vp, c_vp, c_vp, c_vp, c_vp, c_i32, c_dbl, c_dbl, c_i32, c_i32,
                           c_i32, c_vp], c_i32),
    "phd_update_from_summands": ([c_vp, c_vp, c_vp], c_i32),
    "phd_move_generators": ([c_vp, c_vp, c_dbl], c_i32),
    "phd_migrate_periodic": ([c_vp, c_vp, c_vp, c_vp], c_i32),
    "phd_first_outside": ([c_vp, c_vp, c_vp, c_vp, ctypes.POINTER(c_i64)], c_i32),
    "phd_reflect_fields": ([c_vp, c_vp, c_vp, c_vp], c_i32),
    "phd_update_ghost_fields": ([c_vp, c_vp, c_vp, c_vp, c_i32], c_i32),
    "phd_geometry_workspace_bytes": ([c_i32, c_i64, c_i64], ctypes.c_size_t),
    "phd_extract_geometry": ([c_vp, c_vp, c_vp, c_vp, ctypes.c_size_t, ctypes.POINTER(c_i64)],
                             c_i32),
    "phd_compute_time_step_noalloc": ([c_vp, c_dbl, c_dbl, c_i32, c_vp, ctypes.POINTER(c_dbl)],
                                      c_i32),
    "phd_noalloc_max_cells": ([], c_i32),
}

# symbols added by later parts of the library register here too
EXTRA_FLUX_SIGS = {}

_MESH_SIGS = {
    "phdm_build": ([c_i32, c_i64, c_vp, c_vp, c_vp, c_i32, c_i32, ctypes.POINTER(c_i32)], c_vp),
    "phdm_counts": ([c_vp, ctypes.POINTER(c_i64), ctypes.POINTER(c_i64),
                     ctypes.POINTER(c_i64)], c_i32),
    "phdm_fill": ([c_vp] + [c_vp] * 11, c_i32),
    "phdm_build_topology": ([c_i32, c_i64, c_vp, c_vp, c_vp, c_i32, c_i32, ctypes.POINTER(c_i32)],
                            c_vp),
    "phdm_ring_count": ([c_vp, ctypes.POINTER(c_i64)], c_i32),
    "phdm_fill_topology": ([c_vp] + [c_vp] * 4, c_i32),
    "phdm_free": ([c_vp], None),
}

_flux = None
_mesh = None
_handles = {}


def _bind(lib, sigs):
    for name, (args, res) in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res


def flux_symbols():
    """Names of every entry point declared in include/phd_flux.h."""
    return sorted(list(_FLUX_SIGS) + list(EXTRA_FLUX_SIGS))


def load_flux_library(path=FLUX_LIB_PATH):
    """Load libphd_flux.so (no GPU needed to load it)."""
    global _flux
    if _flux is None:
        if not os.path.exists(path):
            raise NativeUnavailable(
                "libphd_flux.so is not built (%s); run __graft_entry__.build()" % path)
        lib = ctypes.CDLL(path)
        _bind(lib, _FLUX_SIGS)
        _bind(lib, EXTRA_FLUX_SIGS)
        if lib.phd_abi_version() != PHD_ABI_VERSION:
            raise NativeUnavailable("libphd_flux.so ABI version mismatch")
        _flux = lib
    return _flux


def mesh_lib():
    global _mesh
    if _mesh is None:
        if not os.path.exists(MESH_LIB_PATH):
            raise NativeUnavailable(
                "libphd_mesh.so is not built (%s); run __graft_entry__.build()" % MESH_LIB_PATH)
        lib = ctypes.CDLL(MESH_LIB_PATH)
        _bind(lib, _MESH_SIGS)
        _mesh = lib
    return _mesh


def flux():
    """The HIP library, for compute calls: needs a visible GPU."""
    if not torch.cuda.is_available():
        raise NativeUnavailable(
            "phd: the face-flux pipeline runs only on the GPU through libphd_flux.so "
            "and no GPU is visible (torch.cuda.is_available() is False)")
    return load_flux_library()


def handle(device=None):
    """Per-device library handle, with its stream set to torch's current one."""
    lib = flux()
    dev = torch.cuda.current_device() if device is None else int(device)
    h = _handles.get(dev)
    if h is None:
        hp = c_vp()
        rc = lib.phd_handle_create(ctypes.byref(hp), dev)
        if rc != 0:
            raise RuntimeError("phd_handle_create failed: %s" % lib.phd_error_string(rc).decode())
        h = hp.value
        _handles[dev] = h
    lib.phd_set_stream(h, c_vp(torch.cuda.current_stream(dev).cuda_stream))
    return h


def check(rc, what):
    if rc != 0:
        lib = load_flux_library()
        raise RuntimeError("%s failed: %s" % (what, lib.phd_error_string(rc).decode()))


def raise_device_fault(h):
    """Synchronise, read the device error word and raise the reference's
    RuntimeError for the first fault in the reference's loop order."""
    lib = flux()
    code = c_i32()
    idx = c_i64()
    check(lib.phd_get_error(h, ctypes.byref(code), ctypes.byref(idx)), "phd_get_error")
    if code.value:
        err = RuntimeError(FAULT_MESSAGES[code.value])
        err.index = idx.value
        raise err


def record_ld(dim):
    """Doubles per cell-record row, as the library reports it
    (include/phd_flux.h phd_cell_record_ld; loading needs no GPU)."""
    ld = load_flux_library().phd_cell_record_ld(int(dim))
    if ld < 0:
        raise ValueError("no cell-record layout for dim %r" % (dim,))
    return ld


class Plan(object):
    """Owner of one phd_plan (device acceleration structure of a frozen mesh)."""

    def __init__(self, cells, faces):
        lib, h = flux(), handle()
        p = c_vp()
        check(lib.phd_plan_create(h, ctypes.byref(cells), ctypes.byref(faces), ctypes.byref(p)),
              "phd_plan_create")
        self.ptr = p.value
        self._lib = lib

    @property
    def nbytes(self):
        return int(self._lib.phd_plan_bytes(self.ptr))

    def close(self):
        if self.ptr:
            self._lib.phd_plan_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ptr(t):
    return c_vp(t.data_ptr()) if t is not None else c_vp()


def ptr_array(tensors, n=5):
    arr = PtrArray5()
    for q, t in enumerate(tensors):
        arr[q] = t.data_ptr()
    return arr
