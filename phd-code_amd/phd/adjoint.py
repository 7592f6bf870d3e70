"""Reverse mode of the static-mesh hydro step, on the GPU.

StaticStepper runs StaticMeshMUSCLHancock.evolve_timestep (integrate.py:
349-378) on a frozen mesh through the C ABI: phd_core_step (gradients +
AREPO limiter, ghost gradients, states + HLLC, update) and
phd_ghost_refresh_eos.  Its vjp() is the hand-written reverse of one such
step (libphd_flux.so: phd_core_step_vjp), deterministic like the forward.

Rollout is a torch.autograd.Function over n steps: forward keeps a
checkpoint of the full cell state every `every` steps; backward replays each
segment from its checkpoint, keeping the segment's step states, and runs the
step VJPs in reverse.  Gradients are taken with respect to the initial
primitive fields of the real cells; the time steps are those the forward
took (CFL on each step's state), held fixed.
"""
import ctypes

import torch

from . import _native as nat
from .views import cells_view, faces_view


class StaticStepper(object):
    def __init__(self, problem):
        self.p = problem
        self.lib, self.h = nat.flux(), nat.handle()
        pc, mesh = problem.particles, problem.mesh
        dim = problem.dim
        self.dim, self.nf = dim, dim + 2
        # static mesh: generator and face velocities are zero (integrate.py:339-347)
        for f in pc.carray_named_groups["w"]:
            pc[f][:] = 0.0
        for f in mesh.faces.carray_named_groups["velocity"]:
            mesh.faces[f][:] = 0.0
        self.cells, self.faces = cells_view(pc), faces_view(mesh)
        recon = problem.recon
        if recon.slope_limiter != nat.LIMITER_AREPO:
            raise RuntimeError("the adjoint supports the AREPO limiter")
        if problem.riemann.solver != nat.SOLVER_HLLC:
            raise RuntimeError("the adjoint supports the HLLC solver")
        recon.grad.resize(pc.get_carray_size())
        self.grad = recon.grad
        problem.riemann.fluxes.resize(mesh.faces.get_carray_size())
        dm = problem.domain_manager
        self.lo = (ctypes.c_double * 3)(*(dm.xmin + [0.0] * (3 - dim)))
        self.hi = (ctypes.c_double * 3)(*(dm.xmax + [0.0] * (3 - dim)))
        self.boundary = problem.boundary.native_code
        self.gamma = problem.eos.get_gamma()
        self.gizmo = int(recon.gizmo_limiter)
        self.boost = int(problem.riemann.boost)
        self.plan = mesh.plan(pc)
        adj = ctypes.c_void_p()
        nat.check(self.lib.phd_adjoint_create(self.h, self.plan, ctypes.byref(self.cells),
                                              ctypes.byref(self.faces), ctypes.byref(adj)),
                  "phd_adjoint_create")
        self.adj = adj.value
        pc_ = pc
        self.real = pc_["tag"] == 0
        self.cons_names = pc_.carray_named_groups["conservative"]
        self.prim_names = pc_.carray_named_groups["primitive"]
        self.xp_of = None           # the cells view the plan's packed rows were built from
        self.capture_initial()

    def capture_initial(self):
        """Take the current real-cell primitives as the initial state that
        rollouts start from (the fields they do not differentiate)."""
        pc = self.p.particles
        self.prim_init = [pc[f][self.real].clone() for f in self.prim_names]

    def lane(self):
        """The stepper's second execution lane (created on first use)."""
        if getattr(self, "_lane", None) is None:
            self._lane = Lane(self)
        return self._lane

    def close(self):
        if getattr(self, "_lane", None) is not None:
            self._lane.close()
            self._lane = None
        if self.adj:
            self.lib.phd_adjoint_destroy(self.adj)
            self.adj = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state ------------------------------------------------------------------
    def state(self):
        pc = self.p.particles
        return [pc[f].clone() for f in self.cons_names + self.prim_names]

    def set_state(self, s):
        pc = self.p.particles
        for f, t in zip(self.cons_names + self.prim_names, s):
            pc[f].copy_(t)
        self.xp_of = None

    def save_cons(self, dst):
        """Copy the conserved fields (all cells) into `dst`."""
        pc = self.p.particles
        for t, f in zip(dst, self.cons_names):
            t.copy_(pc[f])

    def view(self, cons, prim):
        """A cells view of the same mesh whose conserved and primitive
        fields live in the given tensors (dim+2 each, all cells)."""
        v = nat.Cells.from_buffer_copy(self.cells)
        for q, t in enumerate(cons):
            v.cons[q] = t.data_ptr()
        for q, t in enumerate(prim):
            v.prim[q] = t.data_ptr()
        v._keep = (cons, prim)
        return v

    def refresh(self, cells=None):
        """Ghost conserved fields from the reals + EOS, all cells of `cells`
        (default: the container); also rewrites the plan's packed rows."""
        c = self.cells if cells is None else cells
        nat.check(self.lib.phd_ghost_refresh_eos(self.h, self.plan, ctypes.byref(c),
                                                 self.boundary,
                                                 ctypes.cast(self.lo, ctypes.c_void_p),
                                                 ctypes.cast(self.hi, ctypes.c_void_p),
                                                 self.gamma), "phd_ghost_refresh_eos")
        self.xp_of = c

    def time_step(self):
        return self.p.riemann.compute_time_step(self.p.particles, self.p.eos)

    # -- one step and its reverse ----------------------------------------------
    def new_rec(self):
        """A cell-record buffer shaped like the stepper's own."""
        return torch.empty_like(self.grad.data)

    def step(self, dt, rec=None):
        """One step + refresh.  rec: a buffer (new_rec()) that keeps this
        state's limited and ghost gradients for a later vjp(rec=...)."""
        r = self.grad.data if rec is None else rec
        flags = nat.CORE_XP_FRESH if self.xp_of is self.cells else 0
        rc = self.lib.phd_core_step(self.h, self.plan, ctypes.byref(self.cells),
                                    ctypes.byref(self.faces), r.data_ptr(),
                                    self.grad.ld, nat.LIMITER_AREPO, self.boundary,
                                    ctypes.cast(self.lo, ctypes.c_void_p),
                                    ctypes.cast(self.hi, ctypes.c_void_p), self.gamma, float(dt),
                                    self.boost, self.gizmo, nat.SOLVER_HLLC, None, flags)
        nat.check(rc, "phd_core_step")
        self.refresh()

    def vjp(self, dt, lam_out, rec=None, out=None, cells=None):
        """lam_out: dL/dcons after the step (dim+2 tensors over all cells,
        real entries used); returns dL/dcons before it (real entries written;
        the ghost entries of `out` are left as they are, zeros when None).
        `cells` (default: the container) must hold the state the step
        started from; rec: the gradients step(rec=...) kept for this state,
        after which the state must have been refreshed (refresh(cells))."""
        c = self.cells if cells is None else cells
        lam_in = [torch.zeros_like(t) for t in lam_out] if out is None else out
        flags = 0
        if rec is not None:
            if self.xp_of is not c:
                raise RuntimeError("vjp(rec=...) needs refresh() of these cells first")
            flags = nat.VJP_REC_FRESH
        r = self.grad.data if rec is None else rec
        rc = self.lib.phd_core_step_vjp(self.h, self.plan, self.adj, ctypes.byref(c),
                                        ctypes.byref(self.faces), r.data_ptr(),
                                        self.grad.ld, self.boundary,
                                        ctypes.cast(self.lo, ctypes.c_void_p),
                                        ctypes.cast(self.hi, ctypes.c_void_p), self.gamma,
                                        float(dt), self.boost, self.gizmo,
                                        nat.ptr_array(lam_out), nat.ptr_array(lam_in), flags)
        nat.check(rc, "phd_core_step_vjp")
        return lam_in

    def eos_vjp(self, to_prim, lam_y):
        lam_x = [torch.zeros_like(t) for t in lam_y]
        nat.check(self.lib.phd_eos_vjp(self.h, ctypes.byref(self.cells), self.gamma,
                                       int(to_prim), nat.ptr_array(lam_y),
                                       nat.ptr_array(lam_x)), "phd_eos_vjp")
        return lam_x

    def set_initial(self, prim_real):
        """Primitive fields of the real cells (dim+2 tensors of n_real),
        conserved fields from them, ghosts refreshed."""
        pc = self.p.particles
        for f, t in zip(self.prim_names, prim_real):
            pc[f][self.real] = t
        self.p.eos.conservative_from_primitive(pc)
        self.xp_of = None
        self.refresh()


class Lane(object):
    """A second execution lane of a StaticStepper: its own stream (default
    priority), library handle, plan and cell buffers over the same mesh, so
    that forward steps can run concurrently with the stepper's own work."""

    def __init__(self, st):
        self.st = st
        pc = st.p.particles
        dev = pc[st.cons_names[0]].device
        self.stream = torch.cuda.Stream(device=dev)
        hp = ctypes.c_void_p()
        nat.check(st.lib.phd_handle_create(ctypes.byref(hp), dev.index), "phd_handle_create")
        self.h = hp.value
        st.lib.phd_set_stream(self.h, ctypes.c_void_p(self.stream.cuda_stream))
        # the conserved fields as rows of one buffer: a snapshot is one copy
        ref = pc[st.cons_names[0]]
        self.cons_buf = torch.empty((len(st.cons_names), ref.numel()), dtype=ref.dtype,
                                    device=dev)
        self.cons = list(self.cons_buf.unbind(0))
        self.prim = [torch.empty_like(pc[f]) for f in st.prim_names]
        self.cells = st.view(self.cons, self.prim)
        self.plan = nat.Plan(self.cells, st.faces)

    def load(self, state):
        """Cells of this lane = a full state (conserved + primitive)."""
        for t, v in zip(self.cons + self.prim, state):
            t.copy_(v)

    def step(self, dt, rec, packed):
        st = self.st
        lo, hi = ctypes.cast(st.lo, ctypes.c_void_p), ctypes.cast(st.hi, ctypes.c_void_p)
        rc = st.lib.phd_core_step(self.h, self.plan.ptr, ctypes.byref(self.cells),
                                  ctypes.byref(st.faces), rec.data_ptr(), st.grad.ld,
                                  nat.LIMITER_AREPO, st.boundary, lo, hi, st.gamma, float(dt),
                                  st.boost, st.gizmo, nat.SOLVER_HLLC, None,
                                  nat.CORE_XP_FRESH if packed else 0)
        nat.check(rc, "phd_core_step")
        nat.check(st.lib.phd_ghost_refresh_eos(self.h, self.plan.ptr, ctypes.byref(self.cells),
                                               st.boundary, lo, hi, st.gamma),
                  "phd_ghost_refresh_eos")

    def close(self):
        if self.h:
            torch.cuda.current_stream().wait_stream(self.stream)
            self.stream.synchronize()
            self.plan.close()
            self.st.lib.phd_handle_destroy(self.h)
            self.h = None


class Rollout(torch.autograd.Function):
    """Final real-cell density after n static-mesh steps, as a function of
    the initial real-cell density and one initial velocity component.

    Backward, per segment [s0, s1), last segment first: a second lane
    (stream, handle, plan, cells) restores the segment's checkpoint and
    replays it, keeping each step's start conserved fields and the gradient
    records steps s0..s1-2 wrote; then, on the caller's stream, each step in
    reverse gets a refresh of a cells view over its kept conserved fields
    (primitives, ghosts and packed rows are functions of them) and its VJP,
    on the kept gradients (step s1-1: recomputed).  The replay of the next
    segment runs on the lane while this segment's VJPs run: two buffer sets,
    ordered by events."""

    @staticmethod
    def forward(ctx, rho0, v0, stepper, n_steps, every, axis, dts=None):
        st = stepper
        pc = st.p.particles
        prim0 = [t.clone() for t in st.prim_init]
        prim0[0] = rho0.detach()
        prim0[1 + axis] = v0.detach()
        st.set_initial(prim0)
        checkpoints, dts, dts_in = {}, [], dts
        for n in range(n_steps):
            if n % every == 0:
                checkpoints[n] = st.state()
            dt = st.time_step() if dts_in is None else dts_in[n]
            dts.append(dt)
            st.step(dt)
        ctx.st, ctx.n, ctx.every, ctx.axis = st, n_steps, every, axis
        ctx.checkpoints, ctx.dts, ctx.prim0 = checkpoints, dts, prim0
        ctx.final = st.state()
        return pc["density"][st.real].clone()

    @staticmethod
    def backward(ctx, g_rho):
        st = ctx.st
        pc = st.p.particles
        nfld = st.nf
        # final density = prim(cons_N)[0]
        st.set_state(ctx.final)
        lam_prim = [torch.zeros_like(pc[f]) for f in st.prim_names]
        lam_prim[0][st.real] = g_rho
        lam = st.eos_vjp(1, lam_prim)
        bufs = (lam, [torch.zeros_like(t) for t in lam])   # ghost entries stay zero
        nseg = min(ctx.every, ctx.n)
        lane = st.lane()
        main = torch.cuda.current_stream()
        prim_tmp = [torch.empty_like(pc[f]) for f in st.prim_names]
        sets = []
        for _ in range(2):
            snap_bufs = [torch.empty_like(lane.cons_buf) for _ in range(nseg)]
            snaps = [list(sb.unbind(0)) for sb in snap_bufs]
            sets.append({"snap_bufs": snap_bufs,
                         "views": [st.view(sn, prim_tmp) for sn in snaps],
                         "recs": [st.new_rec() for _ in range(nseg)],
                         "ready": torch.cuda.Event(), "free": torch.cuda.Event()})
        for b in sets:
            b["free"].record(main)
        segs = [(s0, min(s0 + ctx.every, ctx.n)) for s0 in reversed(sorted(ctx.checkpoints))]

        def replay(k):
            s0, s1 = segs[k]
            b = sets[k % 2]
            with torch.cuda.stream(lane.stream):
                lane.stream.wait_event(b["free"])
                lane.load(ctx.checkpoints[s0])
                for n in range(s0, s1):
                    b["snap_bufs"][n - s0].copy_(lane.cons_buf)
                    if n < s1 - 1:
                        lane.step(ctx.dts[n], b["recs"][n - s0], packed=n > s0)
                b["ready"].record(lane.stream)

        replay(0)
        for k, (s0, s1) in enumerate(segs):
            if k + 1 < len(segs):
                replay(k + 1)
            b = sets[k % 2]
            main.wait_event(b["ready"])
            for n in reversed(range(s0, s1)):
                out = bufs[1] if lam is bufs[0] else bufs[0]
                view = b["views"][n - s0]
                st.refresh(view)
                lam = st.vjp(ctx.dts[n], lam, rec=b["recs"][n - s0] if n < s1 - 1 else None,
                             out=out, cells=view)
            b["free"].record(main)
        main.wait_stream(lane.stream)
        # initial conserved fields = cons(prim_0)
        st.set_state(ctx.checkpoints[0])
        lam_p0 = st.eos_vjp(0, lam)
        st.set_state(ctx.final)
        return (lam_p0[0][st.real].clone(), lam_p0[1 + ctx.axis][st.real].clone(), None, None,
                None, None, None)


def rollout(stepper, rho0, v0, n_steps, every=10, axis=1, dts=None):
    """Differentiable final density after n_steps (see Rollout).  dts: the
    time steps to take (default: the CFL step of each state)."""
    return Rollout.apply(rho0, v0, stepper, n_steps, every, axis, dts)


def forward_dts(stepper, rho0, v0, n_steps, axis=1, dts=None):
    """Plain forward run; returns (final real density, time steps taken)."""
    st = stepper
    pc = st.p.particles
    prim0 = [t.clone() for t in st.prim_init]
    prim0[0] = rho0
    prim0[1 + axis] = v0
    st.set_initial(prim0)
    out = []
    for n in range(n_steps):
        dt = st.time_step() if dts is None else dts[n]
        out.append(dt)
        st.step(dt)
    return pc["density"][st.real].clone(), out
