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
value = faces / (wall time per step).  roofline is the core step (the four
kernels of gradients .. update, HIP-event medians on the library's stream)
against the algorithmic byte model B_core of that step; "kernels" carries
each kernel's own line.  The adjoint block times the forward and reverse of
a checkpointed KH 256^2 rollout (50 steps, a checkpoint every 10) and
reports reverse/forward.  The ensemble block times 4096 independent 2D
Riemann problems (1k-64k cells each) as one batched graph launch per step
against the per-problem launches, over 100 steps.  The CPU baseline is the
oracle (serial C restatement of the reference loops) on one host core, 20
core steps of the same Sedov set-up at half the lattice side per axis, with
one full-size step beside it; the OpenMP line runs 20 full-size steps.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 128]
                       [--no-cpu-baseline] [--no-adjoint] [--no-ensemble]
       python bench.py --workload ensemble [--problems 4096] [--steps 100]
                       [--schedule G,F,U,X] [--batched-only]   (batched ensemble, its own line)
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
# FP64 VALU issue rate: a wave64 FP64 VALU instruction occupies a SIMD for 4
# cycles; 256 CUs x 4 SIMDs at the 2.4 GHz peak engine clock
VALU_WAVE_INSTR_PER_S = 256 * 4 * 2.4e9 / 4
CSRC = os.path.join(ROOT, "phd-code_amd", "phd", "csrc")
PROFILES = os.path.join(ROOT, "profiles")
# bench stage -> rocprofv3 kernel name of the 3D kernel it launches
KERNEL_NAMES = {"gradients": "k_grad_sell<3, 0, 256, 1>",
                "face_flux": "k_face<3, 1, 1, 2, 0, 256, 1, 3>",
                "update": "k_update_sell<3, 256, 1, 1>", "ghost_gradients": "k_ghost_rec<3>",
                "epilogue": "k_ghost_refresh_eos<3, 0>"}
COPY_KERNEL = "k_pack_xp<3>"
CORE_STAGES = ("gradients", "ghost_gradients", "face_flux", "update")


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
    # ghost gradients: read image rows, write ghost rows, ghost list/map/type/x
    b["ghost_gradients"] = ng * (2 * 8 * ngrad + 4 + 4 + 1 + 8 * dim)
    # fused face kernel: face geometry (centroid, normal, velocity, area),
    # pair_i/pair_j, cell prim/x/dcom/grad, flux out
    b["face_flux"] = (nf * (8 * (3 * dim) + 8 + 4 * 2 + 8 * nfld)
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


def code_text(text):
    """C/C++ source without its comments, whitespace runs collapsed (string
    and character literals kept as written)."""
    out, i, n = [], 0, len(text)
    while i < n:
        c = text[i]
        if c in "\"'":
            j = i + 1
            while j < n and text[j] != c:
                j += 2 if text[j] == "\\" else 1
            out.append(text[i:j + 1])
            i = j + 1
        elif text.startswith("//", i):
            j = text.find("\n", i)
            i = n if j < 0 else j
        elif text.startswith("/*", i):
            j = text.find("*/", i + 2)
            out.append(" ")
            i = n if j < 0 else j + 2
        else:
            out.append(c)
            i += 1
    return " ".join("".join(out).split())


def source_hash():
    """sha256 over the code of the HIP sources and the C header (comments
    and layout excluded): a traffic record is used only for the kernels it
    was measured on."""
    import hashlib
    hsh = hashlib.sha256()
    files = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                   if f.endswith((".hip", ".h")))
    files.append(os.path.join(ROOT, "include", "phd_flux.h"))
    for fn in files:
        with open(fn, encoding="utf-8") as fh:
            code = code_text(fh.read())
        hsh.update(os.path.basename(fn).encode() + b"\0" + code.encode())
    return hsh.hexdigest()


def traffic_path(n):
    return os.path.join(PROFILES, "traffic_sedov3d_%d.json" % n)


def _pmc_means(csv_path, counter):
    """Mean per-dispatch value of one counter for each kernel in a
    rocprofv3 --pmc counter_collection.csv."""
    import csv
    import re
    vals = {}
    with open(csv_path) as fh:
        for r in csv.DictReader(fh):
            if r["Counter_Name"] != counter:
                continue
            m = re.search(r"(k_\w+<[^>]*>)", r["Kernel_Name"])
            if m:
                vals.setdefault(m.group(1), []).append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in vals.items()}


def write_traffic(fetch_csv, write_csv, n, command, sq_csv=None):
    """Turn two rocprofv3 --pmc runs of this benchmark (FETCH_SIZE, then
    WRITE_SIZE; KiB per dispatch) into HBM bytes per launch per kernel.
    Each counter is scaled by what it reads for k_pack_xp, a copy kernel of
    known size: it reads (D+2)+D doubles per cell and writes one 64-byte row
    per cell, all cells."""
    fetch = _pmc_means(fetch_csv, "FETCH_SIZE")
    write = _pmc_means(write_csv, "WRITE_SIZE")
    valu = _pmc_means(sq_csv, "SQ_INSTS_VALU") if sq_csv else {}
    from phd.mesh import build_voronoi
    from phd.problems import jittered_lattice
    g = build_voronoi(3, jittered_lattice(n, 3, 0), [0.0] * 3, [1.0] * 3, 0)
    na_bytes = (g["n_real"] + g["n_ghost"]) * 64
    f_cal = na_bytes / (fetch[COPY_KERNEL] * 1024.0)
    w_cal = na_bytes / (write[COPY_KERNEL] * 1024.0)
    kernels = {}
    for stage, kn in KERNEL_NAMES.items():
        if kn in fetch and kn in write:
            rd = fetch[kn] * 1024.0 * f_cal
            wr = write[kn] * 1024.0 * w_cal
            kernels[stage] = {"kernel": kn, "fetch_kib": fetch[kn], "write_kib": write[kn],
                              "read_bytes": rd, "write_bytes": wr, "traffic_bytes": rd + wr}
            if kn in valu:
                kernels[stage]["valu_wave_instr"] = valu[kn]
    rec = {"workload": "sedov3d n=%d" % n, "source_sha256": source_hash(),
           "copy_kernel": COPY_KERNEL, "copy_known_bytes": na_bytes,
           "fetch_scale": f_cal, "write_scale": w_cal, "kernels": kernels,
           "command": command}
    os.makedirs(PROFILES, exist_ok=True)
    with open(traffic_path(n), "w") as fh:
        json.dump(rec, fh, indent=1)
    return rec


def measured_traffic(n, stage):
    """HBM bytes per launch of a stage's kernel from the committed traffic
    record, or None when there is none for the current sources."""
    try:
        with open(traffic_path(n)) as fh:
            rec = json.load(fh)
    except (OSError, ValueError):
        return None
    if rec.get("source_sha256") != source_hash() or stage not in rec.get("kernels", {}):
        return None
    return rec["kernels"][stage]["traffic_bytes"]


def measured_valu(n, stage):
    """Wave-level FP64 VALU instructions per launch of a stage's kernel from
    the committed counter record (SQ_INSTS_VALU), or None."""
    try:
        with open(traffic_path(n)) as fh:
            rec = json.load(fh)
    except (OSError, ValueError):
        return None
    if rec.get("source_sha256") != source_hash():
        return None
    return rec.get("kernels", {}).get(stage, {}).get("valu_wave_instr")


def stream_copy_gbs(nbytes=2 << 30, reps=10):
    """Device-to-device copy bandwidth measured on this GPU (read + write
    bytes / time, HIP events), the practical HBM ceiling beside the 8 TB/s
    peak."""
    import torch
    a = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    return 2.0 * nbytes / (ms * 1e-3) / 1e9


CPU_SAMPLES = 20


def _timed_core_steps(a, dt, n, omp):
    from oracle import oracle
    oracle.lib(omp)
    times = []
    for _ in range(n):
        t0 = time.perf_counter()
        r = oracle.core_step(a, dt, omp=omp)
        times.append(time.perf_counter() - t0)
        if r["rc"] != 0:
            raise RuntimeError("oracle core step reported fault %d" % r["rc"])
    return times


def cpu_baseline(a_sample, dt_sample, n_sample, a_full, dt_full):
    """The oracle's core step (gradients + limiter, ghost gradients, spatial +
    temporal states, HLLC + deboost, update) on one host core.  The timed
    sample is CPU_SAMPLES steps of the same Sedov set-up at n_sample^3 (a
    128^3 step takes about 7 s on one core); one step of the full mesh is
    timed beside it to show the per-face rate carries over."""
    times = _timed_core_steps(a_sample, dt_sample, CPU_SAMPLES, omp=False)
    t = sorted(times)[len(times) // 2]
    full = _timed_core_steps(a_full, dt_full, 1, omp=False)[0] if a_full is not None else None
    out = {"value": a_sample["nf"] / t, "unit": "face-updates/s", "cores": 1, "kind": "port",
           "sample": "median of %d core steps of the Sedov set-up at %d^3 (%d faces), one core "
                     "of %s: %.3f s per step (min %.3f, max %.3f)" %
                     (len(times), n_sample, a_sample["nf"], cpu_model(), t, min(times),
                      max(times))}
    if full is not None:
        out["full_size_check"] = {"faces": a_full["nf"], "seconds": round(full, 3),
                                  "face_updates_per_s": a_full["nf"] / full}
    return out


def cpu_baseline_omp(a, dt, n):
    """The same core step through the OpenMP build of the oracle (its
    independent per-cell and per-face loops on all threads OpenMP is given;
    the update stays serial, in face order), CPU_SAMPLES steps of the full
    mesh: the all-cores CPU line."""
    from oracle import oracle
    threads = oracle.omp_threads()
    times = _timed_core_steps(a, dt, CPU_SAMPLES, omp=True)
    t = sorted(times)[len(times) // 2]
    return {"value": a["nf"] / t, "unit": "face-updates/s", "cores": threads, "kind": "port",
            "sample": "median of %d core steps of the %d^3 Sedov mesh (%d faces) through the "
                      "OpenMP build of the oracle, %d threads on %s (%d logical CPUs): %.3f s "
                      "per step (min %.3f, max %.3f)" %
                      (len(times), n, a["nf"], threads, cpu_model(), os.cpu_count() or 0, t,
                       min(times), max(times))}


def adjoint_timing(n=256, n_steps=50, every=10, reps=3):
    """Forward and reverse of a checkpointed KH rollout on a static mesh:
    n^2 cells, n_steps steps, a checkpoint every `every` steps, gradient of
    an L2 density misfit with respect to the initial density and v_y.  The
    forward is the rollout's own (steps + checkpoints, time steps fixed to
    the CFL steps of a first run); the backward replays each segment from
    its checkpoint and runs the step VJPs in reverse."""
    import torch
    from phd import problems
    from phd.adjoint import StaticStepper, forward_dts, rollout
    st = StaticStepper(problems.kh_2d(n=n, seed=0))
    real = st.real
    pc = st.p.particles
    rho0 = pc["density"][real].clone()
    vy0 = pc["velocity-y"][real].clone()
    _, dts = forward_dts(st, rho0, vy0, n_steps)
    target = rho0.clone()
    fwd, bwd = [], []
    for _ in range(reps + 1):
        r0 = rho0.clone().requires_grad_(True)
        v0 = vy0.clone().requires_grad_(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rho = rollout(st, r0, v0, n_steps, every=every, dts=dts)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ((rho - target) ** 2).sum().backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        fwd.append(t1 - t0)
        bwd.append(t2 - t1)
    fwd, bwd = sorted(fwd[1:])[reps // 2], sorted(bwd[1:])[reps // 2]
    st.close()
    return {"workload": "KH %d^2 static mesh, %d steps, checkpoint every %d, HLLC + "
                        "PieceWiseLinear(arepo, gizmo)" % (n, n_steps, every),
            "forward_ms": round(fwd * 1e3, 3), "backward_ms": round(bwd * 1e3, 3),
            "ratio": round(bwd / fwd, 3)}


def ensemble_timing(n_problems=4096, steps=100, seed=0, log=None, schedule=None,
                    batched_only=False):
    """Batched ensemble of independent 2D Riemann problems (cells per problem
    log-uniform in [1k, 64k]): autotuned schedule, one graph launch per step.
    Both forms run the same `steps` steps from the same initial state and are
    timed over steps 1 .. steps-1 (step 0 warms each form): the batched run,
    then the per-problem form (each problem through phd_core_step +
    phd_ghost_refresh_eos on its own plan, with the batched run's per-problem
    time steps, so no host synchronisation).  The two final states are
    compared bit for bit.  schedule: a fixed (grad, face, update block, xcd)
    schedule instead of the autotuner's search.  batched_only: time the
    batched run alone (no per-problem form; for profiling the batched
    kernels)."""
    import torch
    from phd import ensemble as E
    log = log or (lambda m: None)
    t0 = time.time()
    spec = E.draw_problems(n_problems, seed=seed)
    H = E.HostEnsemble(spec, n_workers=16)
    log("ensemble host build %.1fs: %d problems, %d cells, %d faces" %
        (time.time() - t0, n_problems, H.n_cells, H.n_faces))
    t1 = time.time()
    ens = E.Ensemble(H, hist_steps=steps + 8)
    del H.x, H.dcom, H.area, H.com, H.normal
    ens.reset()
    torch.cuda.synchronize()
    log("ensemble device set-up %.1fs" % (time.time() - t1))
    best, tried = E.autotune(ens, steps=3, log=log,
                             candidates=[tuple(schedule)] if schedule else None)
    s0 = ens.state()
    # batched run
    ens.restart(s0)
    ens.run(2)                                  # warm
    ens.restart(s0)
    ens.run(1)
    torch.cuda.synchronize()
    tb = time.perf_counter()
    ens.run(steps - 1)
    torch.cuda.synchronize()
    wall_b = time.perf_counter() - tb
    _, k, hist, faults = ens.read(history=True)
    ens.clear_errors()            # frozen problems' failed updates land there too
    assert k == steps
    if batched_only:
        ms_b = wall_b * 1e3 / (steps - 1)
        out = {"workload": "%d independent 2D Riemann problems, batched run only" % n_problems,
               "cells": ens.n_cells, "faces": ens.n_faces, "steps": steps,
               "batched_ms_per_step": round(ms_b, 3),
               "batched_face_updates_per_s": ens.n_faces / (ms_b * 1e-3),
               "schedule": dict(zip(("grad_block", "face_block", "update_block", "xcd"), best))}
        ens.close()
        return out
    after_b = ens.state()
    # per-problem form over the same steps
    t2 = time.time()
    ens.set_state(s0)
    views = ens.problem_views()
    log("per-problem plans %.1fs" % (time.time() - t2))
    ens.step_problems(hist[0], first=True)
    torch.cuda.synchronize()
    tp = time.perf_counter()
    for s in range(1, steps):
        ens.step_problems(hist[s])
    torch.cuda.synchronize()
    wall_p = time.perf_counter() - tp
    ens.clear_errors()
    same = all(torch.equal(a, b) for a, b in zip(after_b, ens.state()))
    del after_b
    nf = ens.n_faces
    ms_b = wall_b * 1e3 / (steps - 1)
    ms_p = wall_p * 1e3 / (steps - 1)
    out = {"workload": "%d independent 2D Riemann problems (static jittered-lattice Voronoi "
                       "meshes, cells log-uniform in [1k, 64k], reflective unit box, HLLC + "
                       "PieceWiseLinear(arepo, gizmo), CFL 0.3 per problem), FP64" % n_problems,
           "cells": ens.n_cells, "faces": nf, "steps": steps, "steps_timed": steps - 1,
           "batched_ms_per_step": round(ms_b, 3),
           "batched_face_updates_per_s": nf / (ms_b * 1e-3),
           "per_problem_ms_per_step": round(ms_p, 3),
           "speedup": round(ms_p / ms_b, 2),
           "schedule": dict(zip(("grad_block", "face_block", "update_block", "xcd"), best)),
           "autotune_ms": {"%d/%d/%d/%d" % c: round(v, 3) for c, v in sorted(tried.items())},
           "bitwise_equal_to_per_problem": bool(same),
           "frozen_problems": int((faults > 0).sum()),
           "frozen_note": "a problem whose update left a non-positive mass/energy/density/"
                          "pressure is frozen from the next step (the reference raises for "
                          "its single problem); the others go on",
           "launches_per_step": 1}
    views = None
    ens.close()
    return out


def geometry_bytes(dim, nr, ng, nf, n_ring):
    """Minimum HBM bytes of phd_extract_geometry: every input read once and
    every output written once.  Reads: pair_i, pair_j, ring_off (12 B per
    face), ring entries (index + code, 5 B each), face lists (2 x 4 B per
    face + 4 B per cell offset), generator positions (8 D B per real cell),
    ghost map + code (5 B per ghost).  Writes: face area, centroid, normal
    (8 (1 + 2 D) B per face), volume + centroid offset (8 (1 + D) B per cell,
    ghosts too), ghost positions (8 D B per ghost).  The per-face workspace
    (2 (1 + D) doubles written, half of it read back per side) is this
    design's and counted apart."""
    na = nr + ng
    reads = 12 * nf + 5 * n_ring + 8 * nf + 4 * (na + 1) + 8 * dim * nr + 5 * ng
    writes = 8 * (1 + 2 * dim) * nf + 8 * (1 + dim) * na + 8 * dim * ng
    workspace = 8 * 2 * (1 + dim) * nf * 2
    return {"model": reads + writes, "workspace": workspace}


def remesh_timing(prob, reps=10):
    """Re-meshing of the benchmark mesh, split as Mesh.build_geometry does:
    device -> host positions, the host topology-only build (all host
    threads), topology upload, phd_extract_geometry on the device (HIP events
    on the library's stream; repeated on the live mesh, same values).  The
    host builder's full build (its own geometry on the host) and its upload
    size are measured beside it."""
    import numpy as np
    import torch
    from phd.mesh import build_voronoi, extract_geometry, upload_topology
    pc, mesh, dm = prob.particles, prob.mesh, prob.domain_manager
    dim, nr = prob.dim, prob.n_real
    pos = pc.carray_named_groups["position"]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x = [pc[f][:nr].cpu().numpy() for f in pos]
    t1 = time.perf_counter()
    bnd = dm.boundary_condition.native_code
    topo = build_voronoi(dim, x, dm.xmin, dm.xmax, bnd, geometry=False)
    t2 = time.perf_counter()
    dev = pc.device
    tdev = upload_topology(topo, dev)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    full = build_voronoi(dim, x, dm.xmin, dm.xmax, bnd, geometry=True)
    t4 = time.perf_counter()
    keys_full = ("ghost_x", "ghost_map", "volume", "dcom", "area", "com", "normal", "pair_i",
                 "pair_j", "nbr_off", "nbr_face")
    keys_topo = ("ghost_map", "ghost_code", "pair_i", "pair_j", "ring_off", "ring_k",
                 "ring_code", "nbr_off", "nbr_face")

    def nbytes(d, keys):
        return int(sum(sum(a.nbytes for a in d[k]) if isinstance(d[k], list) else d[k].nbytes
                       for k in keys))
    same = all(np.array_equal(topo[k], full[k]) for k in keys_topo if k in full)
    full_bytes = nbytes(full, keys_full)
    del full
    fc = mesh.faces
    stream = torch.cuda.current_stream()
    times = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        extract_geometry(tdev, dim, dm.xmin, dm.xmax, bnd, [pc[f] for f in pos], pc["volume"],
                         [pc[f] for f in pc.carray_named_groups["dcom"]], fc)
        e1.record(stream)
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1))
    ms = sorted(times)[len(times) // 2]
    nf, ng = topo["pair_i"].size, topo["ghost_map"].size
    B = geometry_bytes(dim, nr, ng, nf, topo["ring_k"].size)
    return {
        "workload": "re-mesh of the benchmark mesh: %d cells, %d ghosts, %d faces, %d ring "
                    "entries" % (nr, ng, nf, topo["ring_k"].size),
        "positions_d2h_ms": round((t1 - t0) * 1e3, 2),
        "host_topology_ms": round((t2 - t1) * 1e3, 1),
        "topology_h2d_ms": round((t3 - t2) * 1e3, 2),
        "topology_h2d_bytes": nbytes(topo, keys_topo),
        "host_full_build_ms": round((t4 - t3) * 1e3, 1),
        "host_full_upload_bytes": full_bytes,
        "topology_equal_to_full_build": bool(same),
        "device_geometry_ms": round(ms, 4),
        "device_geometry_algorithmic_bytes": B["model"],
        "device_geometry_workspace_bytes": B["workspace"],
        "device_geometry_achieved_gbs": round(B["model"] / (ms * 1e-3) / 1e9, 1),
        "device_geometry_frac_hbm": round(B["model"] / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "host_threads": int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count()),
        "_topology": topo,
        "_x": x,
        "_box": (list(dm.xmin), list(dm.xmax), bnd),
    }


def cpu_geometry_omp(x, topo, lo, hi, boundary, reps=3):
    """The oracle's geometry extraction (OpenMP build, all host threads) on
    the same topology: the CPU line beside phd_extract_geometry."""
    from oracle import oracle
    dim = len(x)
    best = None
    for _ in range(reps):
        t0 = time.perf_counter()
        oracle.extract_geometry(dim, x, topo, lo, hi, boundary, omp=True)
        t = time.perf_counter() - t0
        best = t if best is None else min(best, t)
    return round(best * 1e3, 1)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the batched-ensemble block of the default run")
    ap.add_argument("--workload", choices=("sedov", "ensemble"), default="sedov",
                    help="ensemble: only the batched-ensemble measurement, as its own line")
    ap.add_argument("--problems", type=int, default=4096)
    ap.add_argument("--batched-only", action="store_true",
                    help="ensemble: time the batched run only (no per-problem form)")
    ap.add_argument("--schedule", default=None, metavar="G,F,U,X",
                    help="ensemble: a fixed schedule (threads per block of the gradient, "
                         "face and update kernels, XCD remap 0/1) instead of the autotuner")
    ap.add_argument("--sq-from", default=None, metavar="SQ_CSV",
                    help="with --traffic-from: a third rocprofv3 --pmc run with SQ_INSTS_VALU")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="write the traffic record from two rocprofv3 --pmc runs of this "
                         "benchmark (FETCH_SIZE, WRITE_SIZE) and exit; no GPU needed")
    ap.add_argument("--traffic-command", default="",
                    help="the profiling command, stored in the traffic record")
    args = ap.parse_args()
    if args.traffic_from:
        rec = write_traffic(args.traffic_from[0], args.traffic_from[1], args.n,
                            args.traffic_command, args.sq_from)
        print(json.dumps(rec, indent=1))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the face-flux pipeline here is the single-GPU "
              "kernel at its real shape (multi-GPU sharding is a design, not part of this "
              "build); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from phd import problems
    from phd import _native as nat

    torch.cuda.set_device(0)
    if args.workload == "ensemble":
        def log(m):
            print(m, file=sys.stderr, flush=True)
        sched = [int(v) for v in args.schedule.split(",")] if args.schedule else None
        r = ensemble_timing(args.problems, steps=args.steps, log=log, schedule=sched,
                            batched_only=args.batched_only)
        r.update({"metric": "batched ensemble face-updates/s FP64 (one graph launch per step) "
                            "and speed-up over per-problem launches",
                  "value": r["batched_face_updates_per_s"], "unit": "face-updates/s",
                  "n_gpus": 1, "higher_is_better": True, "dtype": "f64", "data": "synthetic"})
        print(json.dumps(r), flush=True)
        return
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
    # headline roofline: the core step (gradients, ghost gradients, face,
    # update) against the algorithmic byte model B_core of the whole step
    core_traffic = [measured_traffic(args.n, k) for k in CORE_STAGES]
    achieved = B["core_model"] / (core_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": None if None in core_traffic else round(sum(core_traffic)),
                "kernel": "core step: " + ", ".join(KERNEL_NAMES[k] for k in CORE_STAGES),
                "algorithmic_bytes": B["core_model"], "time_ms": round(core_ms, 4),
                "bytes_per_face": round(B["core_model"] / nf, 2)}
    # per-kernel lines: each kernel's own minimum bytes, time and measured
    # traffic / FP64 VALU issue utilisation from the committed counter record
    kernels = {}
    for k in ("gradients", "ghost_gradients", "face_flux", "update", "epilogue"):
        line = {"kernel": KERNEL_NAMES[k], "ms": round(kern_ms[k], 4),
                "algorithmic_bytes": B[k],
                "achieved_gbs": round(B[k] / (kern_ms[k] * 1e-3) / 1e9, 1),
                "frac_hbm": round(B[k] / (kern_ms[k] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        tr = measured_traffic(args.n, k)
        line["traffic"] = None if tr is None else round(tr)
        vi = measured_valu(args.n, k)
        if vi is not None:
            floor_ms = vi / VALU_WAVE_INSTR_PER_S * 1e3
            line["valu_floor_ms"] = round(floor_ms, 4)
            line["valu_frac"] = round(floor_ms / kern_ms[k], 4)
        kernels[k] = line
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
        "roofline": roofline,
        "kernels": kernels,
        "stream_copy_gbs": round(stream_copy_gbs(), 1),
        "cpu_baseline": None,
        "cpu_baseline_omp": None,
        "adjoint": None,
        "ensemble": None,
    }
    remesh = remesh_timing(prob)
    topo_full, x_full, box = remesh.pop("_topology"), remesh.pop("_x"), remesh.pop("_box")
    result["remesh"] = remesh
    # host copies for the CPU lines, then free the device problem
    a_full = problems.host_arrays(prob) if not args.no_cpu_baseline else None
    del L, step, prob, plan
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    if not args.no_adjoint:
        result["adjoint"] = adjoint_timing()
        result["adjoint_over_fwd"] = result["adjoint"]["ratio"]
        torch.cuda.empty_cache()
    if not args.no_ensemble:
        def log(m):
            print(m, file=sys.stderr, flush=True)
        result["ensemble"] = ensemble_timing(args.problems, steps=100, log=log)
        torch.cuda.empty_cache()
    if not args.no_cpu_baseline:
        n_s = max(8, args.n // 2)
        ps = problems.sedov_3d(n=n_s, seed=0)
        st = problems.FrozenStep(ps)
        st.prologue()
        dt_s = st.time_step()
        a_s = problems.host_arrays(ps)
        del st, ps
        result["cpu_baseline"] = cpu_baseline(a_s, dt_s, n_s, a_full, dt)
        result["cpu_baseline_omp"] = cpu_baseline_omp(a_full, dt, args.n)
        result["remesh"]["cpu_geometry_omp_ms"] = cpu_geometry_omp(x_full, topo_full, *box)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
