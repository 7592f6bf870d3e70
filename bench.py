"""Benchmark: Voronoi face-flux pipeline, 3D Sedov blast, FP64, one MI355X.

Workload: 128^3 jittered-lattice generators (about 2.1 M cells, 16.6 M faces),
reflective unit box, gamma 1.4, rho 1, p 1e-5, E = 1 in the 32 cells nearest
the centre; PieceWiseLinear (AREPO limiter, GIZMO pairwise limiter) + HLLC,
boosted face frame.  The mesh is built once on the host and stays frozen.

A step is one GPU hydro step on that mesh, all inputs resident in HBM:
  gradients + limiter, ghost gradients, fused spatial/temporal states + HLLC
  + deboost, deterministic conserved update (the core face-flux step of
  MovingMeshMUSCLHancock.evolve_timestep, integrate.py:396-408), then the
  ghost conserved refresh + EOS epilogue (integrate.py:371-377) so the state
  stays physical from step to step.
value = faces / (wall time per step).  Per-kernel HIP-event times give the
roofline line for the dominant kernel.  The CPU baseline is the oracle
(serial C restatement of the reference loops) timed on one host core on one
full core step of the same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 128]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "phd-code_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(dim, nr, ng, nf):
    """Bytes each kernel must move at least once (every array touched once at
    its natural size; 8-byte doubles, 4-byte indices, 1-byte tags)."""
    na = nr + ng
    nfld = dim + 2
    ngrad = dim * nfld
    face_geom = 8 * (1 + 3 * dim)             # area, com, normal, face velocity
    b = {}
    # gradients: cell prim/x/dcom/volume/tag, face area/com/pairs, CSR, grad out
    b["gradients"] = (na * (8 * (nfld + 2 * dim + 1) + 1) + nf * (8 * (1 + dim) + 8)
                      + 4 * (na + 1) + 4 * 2 * nf + nr * 8 * ngrad)
    # ghost gradients: read image rows, write ghost rows, map/tag/type/x
    b["ghost_gradients"] = ng * (2 * 8 * ngrad + 4 + 2 + 8 * dim) + na
    # fused face kernel: face geometry + pairs, cell prim/x/dcom/grad, flux out
    b["face_flux"] = (nf * (8 * (3 * dim) + 8 + 8 * nfld)
                      + na * 8 * (nfld + 2 * dim + ngrad))
    # update: flux + area + pairs via the CSR, cons in/out, tags
    b["update"] = (nf * (8 * nfld + 8 + 8) + 4 * 2 * nf + 4 * (na + 1)
                   + nr * 2 * 8 * nfld + na)
    # epilogue: cons (ghost image gather), prim out, volume, tag/type/map/x
    b["epilogue"] = (na * (8 * nfld * 2 + 8 + 2) + ng * (4 + 8 * dim + 8 * nfld))
    # the byte model of the core step (gradients .. update)
    b["core_model"] = (8 * (na * (3 * dim + 3) + 2 * na * dim * (dim + 2) + 2 * nr * (dim + 2)
                            + nf * (1 + 3 * dim) + 2 * nf * (dim + 2))
                       + 4 * (4 * nf + nr + 1 + ng))
    return b


def cpu_baseline(problem, dt):
    """One full core step of the same workload through the oracle, one core."""
    from phd import problems
    from oracle import oracle
    a = problems.host_arrays(problem)
    oracle.lib()
    t0 = time.perf_counter()
    r = oracle.core_step(a, dt)
    t = time.perf_counter() - t0
    if r["rc"] != 0:
        raise RuntimeError("oracle core step reported fault %d" % r["rc"])
    return {"value": a["nf"] / t, "unit": "face-updates/s", "cores": 1, "kind": "port",
            "sample": "one full core step (gradients+limiter, ghost gradients, spatial+temporal "
                      "states, HLLC+deboost, update) of the same 128^3 Sedov mesh, %d faces, "
                      "%.2f s on one core" % (a["nf"], t)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the face-flux pipeline here is the single-GPU "
              "kernel at its real shape (multi-GPU sharding is a design, not part of this "
              "build); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from phd import problems
    from phd import _native as nat

    torch.cuda.set_device(0)
    t0 = time.time()
    prob = problems.sedov_3d(n=args.n, seed=0)
    step = problems.FrozenStep(prob)
    step.prologue()
    dt = step.time_step()
    nr, nf = prob.n_real, prob.n_faces
    na = prob.particles.get_carray_size()
    ng = na - nr
    print("setup %.1fs: %d real cells, %d ghosts, %d faces, dt %.3e" %
          (time.time() - t0, nr, ng, nf, dt), file=sys.stderr, flush=True)

    h = nat.handle()
    lib = nat.flux()
    import ctypes
    c, f = ctypes.byref(step.cells), ctypes.byref(step.faces)
    lo = ctypes.cast(step.lo, ctypes.c_void_p)
    hi = ctypes.cast(step.hi, ctypes.c_void_p)
    g_ptr, ld = step.grad.data.data_ptr(), step.grad.ld
    gamma = step.gamma

    plan = step.plan

    def launches():
        """The kernels of one step, in order (what phd_core_step runs, with
        the packed cell rows refreshed by the previous step's epilogue)."""
        return [
            ("gradients", lambda: lib.phd_gradients_packed(h, plan, c, step.limiter, g_ptr, ld)),
            ("ghost_gradients", lambda: lib.phd_update_ghost_gradients(
                h, plan, c, step.boundary, lo, hi, g_ptr, ld)),
            ("face_flux", lambda: lib.phd_face_summands(
                h, plan, c, f, g_ptr, ld, gamma, dt, step.boost, step.gizmo, step.solver, None)),
            ("update", lambda: lib.phd_update_from_summands(h, plan, c)),
            ("epilogue", lambda: lib.phd_ghost_refresh_eos(h, plan, c, step.boundary, lo, hi,
                                                           gamma)),
        ]

    L = launches()

    def one_step():
        for name, fn in L:
            nat.check(fn(), name)

    nat.check(lib.phd_pack_cells(h, plan, c), "pack")
    for _ in range(args.warmup):
        one_step()
    torch.cuda.synchronize()
    step.check()

    # timed region: K whole steps, wall clock, synchronised both sides
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        one_step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t1
    step.check()
    ms = wall * 1e3 / args.steps

    # per-kernel HIP events on the library's stream (torch's current stream)
    names = [n for n, _ in L]
    ev = {n: [] for n in names}
    reps = max(5, min(args.steps, 20))
    stream = torch.cuda.current_stream()
    marks = []
    for _ in range(reps):
        row = [torch.cuda.Event(enable_timing=True)]
        row[0].record(stream)
        for name, fn in L:
            nat.check(fn(), name)
            e = torch.cuda.Event(enable_timing=True)
            e.record(stream)
            row.append(e)
        marks.append(row)
    torch.cuda.synchronize()
    step.check()
    for row in marks:
        for q, name in enumerate(names):
            ev[name].append(row[q].elapsed_time(row[q + 1]))
    kern_ms = {n: sorted(v)[len(v) // 2] for n, v in ev.items()}
    core_ms = sum(kern_ms[n] for n in ("gradients", "ghost_gradients", "face_flux", "update"))

    B = algorithmic_bytes(prob.dim, nr, ng, nf)
    dom = max(("gradients", "face_flux", "update"), key=lambda n: kern_ms[n])
    achieved = B[dom] / (kern_ms[dom] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": dom}

    result = {
        "metric": "Voronoi face-flux updates/s FP64 (3D Sedov 2M cells) + %HBM roofline; "
                  "adjoint/fwd",
        "value": nf / (ms * 1e-3),
        "unit": "face-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "3D Sedov blast, %d^3 jittered-lattice generators, frozen "
                               "Voronoi mesh, HLLC + PieceWiseLinear(arepo, gizmo), FP64 "
                               "BITWISE kernels; step = core face-flux step + ghost/EOS "
                               "epilogue" % args.n,
                   "cells": nr, "ghosts": ng, "faces": nf, "parallelism": "single GPU"},
        "kernel_ms": {k: round(v, 4) for k, v in kern_ms.items()},
        "core_step_ms": round(core_ms, 4),
        "core_face_updates_per_s": nf / (core_ms * 1e-3),
        "core_model_frac_hbm": round(B["core_model"] / (core_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "roofline": roofline,
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(prob, dt)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
