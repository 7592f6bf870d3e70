"""Meshes whose faces are not numbered by owner, through the core step.

The builder numbers faces in owner order, so every cell's list is its
incoming faces followed by one contiguous run of owned faces, and the
update kernels take the path that streams that run.  A mesh from elsewhere
need not be ordered that way (the reference accepts any face numbering and
sums each cell's faces in ascending id, mesh.pyx:479-551).  Here the faces
of a builder mesh are renumbered by a seeded permutation -- every per-face
array moved, every cell's list remapped and sorted ascending -- so the plan
finds no owner runs and the update walks each list through its face ids.
The step is compared with the oracle's on the same renumbered arrays, bit
for bit, and with the step on the builder's numbering, from which it
differs only by the order of each cell's sums (to the 1e-12 bar).
"""
import numpy as np
import pytest
import torch

from parity import check_fields

pytestmark = pytest.mark.gpu

CASES = [("sod_2d", 16), ("kh_2d", 16), ("sedov_3d", 7)]


def _renumber(p, seed):
    """Permute the face ids of problem p's mesh in place; returns the
    permutation (old id -> new id)."""
    mesh = p.mesh
    nf = mesh.faces.get_carray_size()
    perm = np.random.default_rng(seed).permutation(nf)
    inv = np.empty(nf, dtype=np.int64)
    inv[perm] = np.arange(nf)
    for name in list(mesh.faces.carrays):
        t = mesh.faces[name]
        t.copy_(t[torch.as_tensor(inv, device=t.device)].clone())
    off = mesh.neighbors.offsets.cpu().numpy()
    lst = mesh.neighbors.faces.cpu().numpy().astype(np.int64)
    new = perm[lst]
    for i in range(off.size - 1):
        new[off[i]:off[i + 1]] = np.sort(new[off[i]:off[i + 1]])
    mesh.neighbors.faces.copy_(torch.as_tensor(new, dtype=mesh.neighbors.faces.dtype))
    mesh.invalidate_plan()
    return perm


def _owner_runs(a):
    """The mesh property the plan checks: in every real cell's list the
    owned faces (the cell is pair_i) form one run of consecutive ids after
    all its incoming faces."""
    off, lst = a["nbr_off"], a["nbr_face"]
    for i in np.nonzero(a["tag"] == 0)[0]:
        ids = lst[off[i]:off[i + 1]]
        own = a["pair_i"][ids] == i
        k = int(np.argmax(own)) if own.any() else ids.size
        if own[:k].any() or not own[k:].all():
            return False
        if own.any() and np.any(np.diff(ids[k:]) != 1):
            return False
    return True


@pytest.mark.parametrize("name,n", CASES)
def test_core_step_on_renumbered_faces(gpu, name, n):
    from phd import problems
    from oracle import oracle

    print("case %s n=%d renumbered faces" % (name, n), flush=True)
    p = problems.PROBLEMS[name](n=n, seed=3)
    perm = _renumber(p, seed=11)
    step = problems.FrozenStep(p)
    assert not _owner_runs(problems.host_arrays(p)), "the renumbered mesh still has owner runs"
    step.prologue()
    dt = step.time_step()
    a = problems.host_arrays(p)
    ref = oracle.core_step(a, dt, solver=step.solver, boost=bool(step.boost),
                           gizmo=bool(step.gizmo), limiter=step.limiter)
    assert ref["rc"] == 0 and ref["bad"] < 0
    step.run(dt)
    step.check()
    torch.cuda.synchronize()
    real = a["tag"] == 0
    flux = [t.cpu().numpy() for t in p.riemann.fluxes.group("conservative")]
    cons = [t.cpu().numpy() for t in p.particles.group("conservative")]
    check_fields(flux, ref["flux"], "flux", bitwise=True)
    check_fields(cons, ref["cons"], "cons", mask=real, bitwise=True)
    # the same faces in the builder's numbering give the same fluxes
    q = problems.PROBLEMS[name](n=n, seed=3)
    sq = problems.FrozenStep(q)
    assert _owner_runs(problems.host_arrays(q))
    sq.prologue()
    assert sq.time_step() == dt
    sq.run(dt)
    torch.cuda.synchronize()
    flux0 = [t.cpu().numpy() for t in q.riemann.fluxes.group("conservative")]
    check_fields([f[perm] for f in flux], flux0, "flux vs builder numbering")
