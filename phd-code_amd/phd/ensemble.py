"""Batched ensembles of independent 2D (or 3D) Riemann problems on the GPU.

Each problem is a static-mesh run (StaticMeshMUSCLHancock.evolve_timestep,
integrate.py:349-378, with generator and face velocities zero) of a
Riemann problem in the reflective unit square (cube): left state for
x < 0.5, right state otherwise, on its own jittered-lattice Voronoi mesh.  The ensemble is
held as one union mesh (libphd_flux.so phd_batch_*): problem p owns cells
[cell_start[p], cell_start[p+1]) -- its reals, then its ghosts -- and a
contiguous run of faces, with global indices.  One step of every problem is
a handful of kernels flat over the union plus a per-problem CFL minimum on
the device, captured once in a HIP graph and replayed with one graph launch
per step.

The per-problem form runs the same problems through the ordinary per-problem
entry points (phd_core_step + phd_ghost_refresh_eos on each problem's own
plan), over views into the same field arrays, for the comparison and the
parity checks.

Seeded draws (numpy default_rng(seed)): cells per problem log-uniform in
[n_min, n_max] (the lattice side is the rounded square or cube root),
left/right density and pressure log-uniform in [0.1, 10], left/right
x-velocity uniform in [-1, 1], the other components 0; gamma 1.4; HLLC in the lab frame (boost off, so
the CFL bound includes |v|), PieceWiseLinear with the AREPO and GIZMO
limiters, CFL 0.3.
"""
import ctypes
import math
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch

from . import _native as nat
from .mesh import build_voronoi
from .problems import jittered_lattice

GAMMA = 1.4
CFL = 0.3
BOOST = 0
GIZMO = 1
LO = (0.0, 0.0, 0.0)
HI = (1.0, 1.0, 1.0)


def draw_problems(n_problems, seed=0, n_min=1000, n_max=64000, dim=2):
    """Seeded problem parameters (dict of numpy arrays, one entry per
    problem, and the dimension)."""
    if dim not in (2, 3):
        raise ValueError("dim must be 2 or 3")
    rng = np.random.default_rng(seed)
    cells = np.exp(rng.uniform(math.log(n_min), math.log(n_max), n_problems))
    root = np.sqrt(cells) if dim == 2 else np.cbrt(cells)
    side = np.maximum(2, np.rint(root)).astype(np.int64)

    def logu(n):
        return np.exp(rng.uniform(math.log(0.1), math.log(10.0), n))
    return {"side": side, "rho_l": logu(n_problems), "rho_r": logu(n_problems),
            "p_l": logu(n_problems), "p_r": logu(n_problems),
            "v_l": rng.uniform(-1.0, 1.0, n_problems), "v_r": rng.uniform(-1.0, 1.0, n_problems),
            "mesh_seed": rng.integers(0, 2**31 - 1, n_problems), "dim": dim}


def _mesh(args):
    side, seed, dim = args
    x = jittered_lattice(int(side), dim, int(seed))
    return x, build_voronoi(dim, x, LO[:dim], HI[:dim], nat.BOUNDARY_REFLECTIVE, 1)


class HostEnsemble(object):
    """The union mesh on the host (numpy), with the per-problem local index
    arrays the per-problem views need."""

    def __init__(self, spec, n_workers=16):
        self.spec = spec
        P = len(spec["side"])
        D = int(spec.get("dim", 2))
        self.n_problems, self.dim = P, D
        with ThreadPoolExecutor(max_workers=max(1, n_workers)) as ex:
            parts = list(ex.map(_mesh, zip(spec["side"], spec["mesh_seed"], [D] * P)))
        n_real = np.array([x[0].size for x, _ in parts], dtype=np.int64)
        n_ghost = np.array([g["ghost_map"].size for _, g in parts], dtype=np.int64)
        n_face = np.array([g["area"].size for _, g in parts], dtype=np.int64)
        n_ent = np.array([g["nbr_face"].size for _, g in parts], dtype=np.int64)
        na_p = n_real + n_ghost
        cs = np.zeros(P + 1, dtype=np.int64)
        np.cumsum(na_p, out=cs[1:])
        fs = np.zeros(P + 1, dtype=np.int64)
        np.cumsum(n_face, out=fs[1:])
        es = np.zeros(P + 1, dtype=np.int64)
        np.cumsum(n_ent, out=es[1:])
        if cs[-1] >= 2**31 or fs[-1] >= 2**31 or es[-1] >= 2**31:
            raise ValueError("ensemble too large for 32-bit indices")
        self.n_real, self.n_ghost, self.n_face = n_real, n_ghost, n_face
        self.cell_start, self.face_start, self.entry_start = cs, fs, es
        na, nf = int(cs[-1]), int(fs[-1])
        self.n_cells, self.n_faces = na, nf

        def cat(fn, dtype):
            return np.concatenate([fn(p, x, g) for p, (x, g) in enumerate(parts)]).astype(dtype)

        self.x = [cat(lambda p, x, g, k=k: np.concatenate([x[k], g["ghost_x"][k]]), np.float64)
                  for k in range(D)]
        self.tag = cat(lambda p, x, g: np.concatenate(
            [np.zeros(n_real[p]), np.ones(n_ghost[p])]), np.int8)
        self.type = cat(lambda p, x, g: np.concatenate(
            [np.full(n_real[p], 3), np.full(n_ghost[p], 2)]), np.int8)
        self.map_local = cat(lambda p, x, g: np.concatenate(
            [np.arange(n_real[p]), g["ghost_map"]]), np.int32)
        self.volume = cat(lambda p, x, g: g["volume"], np.float64)
        self.dcom = [cat(lambda p, x, g, k=k: g["dcom"][k], np.float64) for k in range(D)]
        self.area = cat(lambda p, x, g: g["area"], np.float64)
        self.com = [cat(lambda p, x, g, k=k: g["com"][k], np.float64) for k in range(D)]
        self.normal = [cat(lambda p, x, g, k=k: g["normal"][k], np.float64) for k in range(D)]
        self.pair_i_local = cat(lambda p, x, g: g["pair_i"], np.int32)
        self.pair_j_local = cat(lambda p, x, g: g["pair_j"], np.int32)
        self.nbr_face_local = cat(lambda p, x, g: g["nbr_face"], np.int32)
        # per problem n_cells + 1 offsets: block p starts at cell_start[p] + p
        self.nbr_off_local = cat(lambda p, x, g: g["nbr_off"], np.int32)
        # global indices: cell local + cell_start[p], face local + face_start[p]
        cell_of_face = np.repeat(cs[:-1], n_face)
        self.pair_i = (self.pair_i_local + cell_of_face).astype(np.int32)
        self.pair_j = (self.pair_j_local + cell_of_face).astype(np.int32)
        self.map = (self.map_local + np.repeat(cs[:-1], na_p)).astype(np.int32)
        self.nbr_face = (self.nbr_face_local + np.repeat(fs[:-1], n_ent)).astype(np.int32)
        off = np.empty(na + 1, dtype=np.int64)
        keep = np.ones(self.nbr_off_local.size, dtype=bool)
        keep[(cs[1:] + np.arange(1, P + 1)) - 1] = False      # drop each block's last offset
        off[:na] = self.nbr_off_local[keep].astype(np.int64) + np.repeat(es[:-1], na_p)
        off[na] = es[-1]
        self.nbr_off = off.astype(np.int32)
        # initial primitive state
        pid = np.repeat(np.arange(P), na_p)
        left = self.x[0] < 0.5
        self.density = np.where(left, spec["rho_l"][pid], spec["rho_r"][pid])
        self.pressure = np.where(left, spec["p_l"][pid], spec["p_r"][pid])
        self.vx = np.where(left, spec["v_l"][pid], spec["v_r"][pid])
        self.v = [self.vx] + [np.zeros(na) for _ in range(D - 1)]
        self.vy = self.v[1]


class Ensemble(object):
    """The union mesh on the device, its plan and batch object; per-problem
    views on demand."""

    def __init__(self, host, device=None, hist_steps=0, boost=BOOST, gizmo=GIZMO):
        dev = torch.device("cuda", torch.cuda.current_device()) if device is None else device
        self.host, self.device = host, dev
        self.boost, self.gizmo = int(boost), int(gizmo)
        self.lib, self.h = nat.flux(), nat.handle()

        def t(a, dtype=None):
            return torch.as_tensor(np.ascontiguousarray(a), device=dev, dtype=dtype)
        H = host
        na, nf = H.n_cells, H.n_faces
        D = self.dim = H.dim
        self.n_cells, self.n_faces, self.n_problems = na, nf, H.n_problems
        self.x = [t(a) for a in H.x]
        self.tag, self.type = t(H.tag), t(H.type)
        self.map, self.map_local = t(H.map), t(H.map_local)
        self.volume = t(H.volume)
        self.dcom = [t(a) for a in H.dcom]
        self.prim = [torch.zeros(na, dtype=torch.float64, device=dev) for _ in range(D + 2)]
        self.cons = [torch.zeros(na, dtype=torch.float64, device=dev) for _ in range(D + 2)]
        self.w = [torch.zeros(na, dtype=torch.float64, device=dev) for _ in range(D)]
        self.area = t(H.area)
        self.com = [t(a) for a in H.com]
        self.normal = [t(a) for a in H.normal]
        self.fvel = [torch.zeros(nf, dtype=torch.float64, device=dev) for _ in range(D)]
        self.pair_i, self.pair_j = t(H.pair_i), t(H.pair_j)
        self.nbr_off, self.nbr_face = t(H.nbr_off), t(H.nbr_face)
        self.pair_i_local, self.pair_j_local = t(H.pair_i_local), t(H.pair_j_local)
        self.nbr_off_local, self.nbr_face_local = t(H.nbr_off_local), t(H.nbr_face_local)
        self.ld = nat.record_ld(D)
        self.rec = torch.zeros((na, self.ld), dtype=torch.float64, device=dev)
        self.cells = self._cells(0, na, self.map)
        self.faces = self._faces(0, nf, self.pair_i, self.pair_j, self.nbr_off, self.nbr_face)
        self.plan = nat.Plan(self.cells, self.faces)
        self.lo = (ctypes.c_double * 3)(*[LO[k] if k < D else 0.0 for k in range(3)])
        self.hi = (ctypes.c_double * 3)(*[HI[k] if k < D else 0.0 for k in range(3)])
        cs = (ctypes.c_int32 * (H.n_problems + 1))(*[int(v) for v in H.cell_start])
        b = ctypes.c_void_p()
        nat.check(self.lib.phd_batch_create(self.h, self.plan.ptr, ctypes.byref(self.cells),
                                            ctypes.byref(self.faces), H.n_problems, cs,
                                            int(hist_steps), ctypes.byref(b)),
                  "phd_batch_create")
        self.batch = b.value
        self.hist_steps = int(hist_steps)
        self.schedule = None
        self.captured = False
        self._views = None

    # -- views -------------------------------------------------------------------
    def _cells(self, c0, n, mapt):
        c = nat.Cells()
        c.dim, c.n_cells = self.dim, n
        e = 8 * c0
        c.tag = self.tag.data_ptr() + c0
        c.type = self.type.data_ptr() + c0
        c.map = mapt.data_ptr() + 4 * c0
        for k in range(self.dim):
            c.x[k] = self.x[k].data_ptr() + e
            c.dcom[k] = self.dcom[k].data_ptr() + e
            c.w[k] = self.w[k].data_ptr() + e
        c.volume = self.volume.data_ptr() + e
        for q in range(self.dim + 2):
            c.prim[q] = self.prim[q].data_ptr() + e
            c.cons[q] = self.cons[q].data_ptr() + e
        return c

    def _faces(self, f0, n, pi, pj, off_ptr_t, nbr_t, off0=0, e0=0):
        f = nat.Faces()
        f.n_faces = n
        e = 8 * f0
        f.area = self.area.data_ptr() + e
        for k in range(self.dim):
            f.com[k] = self.com[k].data_ptr() + e
            f.normal[k] = self.normal[k].data_ptr() + e
            f.velocity[k] = self.fvel[k].data_ptr() + e
        f.pair_i = pi.data_ptr() + 4 * f0
        f.pair_j = pj.data_ptr() + 4 * f0
        f.nbr_off = off_ptr_t.data_ptr() + 4 * off0
        f.nbr_face = nbr_t.data_ptr() + 4 * e0 if nbr_t.numel() else 0
        return f

    def problem_views(self):
        """Per problem (cells, faces, plan, rec pointer): views into the
        union's field arrays with local indices, each with its own plan."""
        if self._views is None:
            H = self.host
            views = []
            for p in range(H.n_problems):
                c0, f0 = int(H.cell_start[p]), int(H.face_start[p])
                n = int(H.cell_start[p + 1]) - c0
                c = self._cells(c0, n, self.map_local)
                f = self._faces(f0, int(H.n_face[p]), self.pair_i_local, self.pair_j_local,
                                self.nbr_off_local, self.nbr_face_local, off0=c0 + p,
                                e0=int(H.entry_start[p]))
                views.append((c, f, nat.Plan(c, f), self.rec.data_ptr() + 8 * self.ld * c0))
            self._views = views
        return self._views

    # -- state -------------------------------------------------------------------
    def reset(self):
        """Initial primitive state, conserved fields from it, ghosts refreshed."""
        H = self.host
        for q, a in enumerate([H.density] + H.v + [H.pressure]):
            self.prim[q].copy_(torch.as_tensor(a, device=self.device))
        nat.check(self.lib.phd_conservative_from_primitive(self.h, ctypes.byref(self.cells), GAMMA),
                  "phd_conservative_from_primitive")
        self.refresh()

    def refresh(self):
        """Ghost refresh + EOS (phd_ghost_refresh_eos's), folding each
        problem's CFL minimum for the next batched step."""
        nat.check(self.lib.phd_batch_refresh(self.h, self.batch, self.plan.ptr,
                                             ctypes.byref(self.cells), nat.BOUNDARY_REFLECTIVE,
                                             ctypes.cast(self.lo, ctypes.c_void_p),
                                             ctypes.cast(self.hi, ctypes.c_void_p), GAMMA, self.boost),
                  "phd_batch_refresh")

    def state(self):
        return [t.clone() for t in self.cons + self.prim]

    def set_state(self, s):
        for t, v in zip(self.cons + self.prim, s):
            t.copy_(v)
        self.refresh()

    def restart(self, s):
        """set_state(s) and a fresh history (step count and fault marks)."""
        self.set_state(s)
        self.reset_history()

    # -- batched steps ------------------------------------------------------------
    def set_schedule(self, grad_block=256, face_block=256, update_block=256, xcd=1):
        s = nat.Schedule(grad_block, face_block, update_block, xcd)
        nat.check(self.lib.phd_batch_set_schedule(self.batch, ctypes.byref(s)),
                  "phd_batch_set_schedule")
        self.schedule = (grad_block, face_block, update_block, xcd)
        self.captured = False

    def _args(self):
        return (self.h, self.batch, self.plan.ptr, ctypes.byref(self.cells),
                ctypes.byref(self.faces), self.rec.data_ptr(), self.ld, nat.BOUNDARY_REFLECTIVE,
                ctypes.cast(self.lo, ctypes.c_void_p), ctypes.cast(self.hi, ctypes.c_void_p),
                GAMMA, CFL, self.boost, self.gizmo)

    def step_eager(self):
        """One batched step by direct launches (the packed rows are fresh
        after reset(), set_state() and every step)."""
        nat.check(self.lib.phd_batch_step(*self._args(), nat.CORE_XP_FRESH), "phd_batch_step")

    def capture(self):
        nat.check(self.lib.phd_batch_capture(*self._args()), "phd_batch_capture")
        self.captured = True

    def run(self, n_steps):
        """n_steps graph launches (captures first if needed); asynchronous."""
        if not self.captured:
            self.capture()
        nat.check(self.lib.phd_batch_launch(self.h, self.batch, int(n_steps)), "phd_batch_launch")

    def reset_history(self):
        nat.check(self.lib.phd_batch_reset_history(self.h, self.batch), "phd_batch_reset_history")

    def read(self, history=False):
        """(dt of the last step per problem, steps recorded, history or None,
        fault marks: 0, or 1 + the step count at which the problem was
        frozen)."""
        P = self.n_problems
        dt = np.zeros(P)
        faults = np.zeros(P, dtype=np.int32)
        k = ctypes.c_int64()
        nat.check(self.lib.phd_batch_read(self.h, self.batch, dt.ctypes.data, ctypes.byref(k),
                                          None, faults.ctypes.data), "phd_batch_read")
        hist = None
        if history and k.value > 0:
            hist = np.zeros((min(k.value, self.hist_steps), P))
            nat.check(self.lib.phd_batch_read(self.h, self.batch, dt.ctypes.data,
                                              ctypes.byref(k), hist.ctypes.data, None),
                      "phd_batch_read")
        return dt, k.value, hist, faults

    def check(self):
        """Raise the reference's RuntimeError for a device fault, if any."""
        nat.raise_device_fault(self.h)

    def clear_errors(self):
        """Read and clear the handle's device error word (a frozen problem's
        failed update also lands there); returns (code, index)."""
        code, idx = ctypes.c_int32(), ctypes.c_int64()
        nat.check(self.lib.phd_get_error(self.h, ctypes.byref(code), ctypes.byref(idx)),
                  "phd_get_error")
        return code.value, idx.value

    def problem_cells(self, problems):
        """Boolean mask over the union's cells of the given problems."""
        H = self.host
        pid = np.repeat(np.arange(H.n_problems), H.cell_start[1:] - H.cell_start[:-1])
        return torch.as_tensor(np.isin(pid, problems), device=self.device)

    # -- per-problem form -----------------------------------------------------------
    def time_step_problem(self, p):
        c, f, plan, rec = self.problem_views()[p]
        dt = ctypes.c_double()
        nat.check(self.lib.phd_compute_time_step(self.h, ctypes.byref(c), GAMMA, CFL, self.boost,
                                                 ctypes.byref(dt)), "phd_compute_time_step")
        return dt.value

    def step_problems(self, dts, first=False):
        """One step of every problem through the per-problem entry points,
        problem p with time step dts[p]; a problem with dts[p] == 0 (frozen
        by the batched run) is skipped.  first: the per-problem packed rows
        are stale (after a batched run or a state change)."""
        lib, h = self.lib, self.h
        lo, hi = ctypes.cast(self.lo, ctypes.c_void_p), ctypes.cast(self.hi, ctypes.c_void_p)
        flags = 0 if first else nat.CORE_XP_FRESH
        for p, (c, f, plan, rec) in enumerate(self.problem_views()):
            if dts[p] == 0.0:
                continue
            cr, fr = ctypes.byref(c), ctypes.byref(f)
            rc = lib.phd_core_step(h, plan.ptr, cr, fr, rec, self.ld, nat.LIMITER_AREPO,
                                   nat.BOUNDARY_REFLECTIVE, lo, hi, GAMMA, float(dts[p]),
                                   self.boost, self.gizmo, nat.SOLVER_HLLC, None, flags)
            if rc:
                nat.check(rc, "phd_core_step")
            rc = lib.phd_ghost_refresh_eos(h, plan.ptr, cr, nat.BOUNDARY_REFLECTIVE, lo, hi, GAMMA)
            if rc:
                nat.check(rc, "phd_ghost_refresh_eos")

    def close(self):
        if self.batch:
            self.lib.phd_batch_destroy(self.batch)
            self.batch = None
        if self._views:
            for _, _, plan, _ in self._views:
                plan.close()
            self._views = None
        self.plan.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(ens, steps=3, candidates=None, log=None):
    """Pick the launch schedule of the batched step by timing: every
    candidate runs `steps` graph launches from the same saved state (median
    of two trials, HIP events on the caller's stream); the fastest is kept
    and set.  Returns (best schedule, {schedule: ms per step})."""
    if candidates is None:
        candidates = [(g, fb, u, x) for x in (1, 0) for g in (256, 128, 64)
                      for fb in (256, 128, 64) for u in (256, 128)]
    s0 = ens.state()
    times = {}
    for cand in candidates:
        ens.set_schedule(*cand)
        trial = []
        for _ in range(2):
            ens.restart(s0)
            ens.run(1)                     # capture + warm
            torch.cuda.synchronize()
            ens.restart(s0)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            ens.run(steps)
            e1.record()
            torch.cuda.synchronize()
            trial.append(e0.elapsed_time(e1) / steps)
        times[cand] = sorted(trial)[0]
        if log:
            log("schedule %s: %.3f ms/step" % (cand, times[cand]))
    best = min(times, key=times.get)
    ens.set_schedule(*best)
    ens.restart(s0)
    ens.clear_errors()
    return best, times
