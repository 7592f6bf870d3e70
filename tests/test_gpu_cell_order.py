"""Meshes whose real cells are not in a space-filling-curve order.

The problem builders number cells along a Morton curve, so a window of 256
consecutive cells is a compact region: the gradient kernel finds most
neighbours among the window's own rows and, in 3D, the rest in the window's
staged halo.  The reference accepts any numbering (its loops run over
indices).  Here the real cells of a builder mesh are renumbered by a seeded
permutation -- every per-cell array moved, the face pairs, ghost images and
per-cell face lists remapped -- so a window's neighbours are scattered over
the whole mesh: in 3D its halo overflows the staging and most entries take
the gathered path, and the faces are no longer numbered by owner.  A third
case puts the 256 real cells of highest degree first, so that the first
window's lists name more outside entries than the plan's halo build sorts
(4096): that window keeps no halo and reads every outside row from HBM.
The core step is compared with the oracle's on the same renumbered arrays,
bit for bit.
"""
import numpy as np
import pytest
import torch

from parity import check_fields

pytestmark = pytest.mark.gpu

CASES = [("sod_2d", 20), ("sedov_3d", 9)]


HALO_SORT_CAP = 4096     # outside entries a window's halo build sorts


def _renumber_cells(p, seed=None, by_degree=False):
    """Permute the real cells of problem p in place (ghosts keep their
    indices): at random, or by descending degree (ties by index); returns
    sigma (old real index -> new)."""
    pc, mesh = p.particles, p.mesh
    nr = int((pc["tag"] == 0).sum())
    n = pc.get_carray_size()
    assert bool((pc["tag"][:nr] == 0).all()), "reals must come first"
    sigma = np.arange(n)
    if by_degree:
        deg = np.diff(mesh.neighbors.offsets.cpu().numpy())[:nr]
        order = np.lexsort((np.arange(nr), -deg))      # old index of each new position
        sigma[order] = np.arange(nr)
    else:
        sigma[:nr] = np.random.default_rng(seed).permutation(nr)
    inv = np.empty(n, dtype=np.int64)
    inv[sigma] = np.arange(n)
    dev = pc["tag"].device
    inv_t = torch.as_tensor(inv, device=dev)
    sig_t = torch.as_tensor(sigma, device=dev)
    for name in list(pc.carrays):
        t = pc[name]
        t.copy_(t[inv_t].clone())
    m = pc["map"]
    ghost = pc["tag"] != 0
    m[ghost] = sig_t[m[ghost].long()].to(m.dtype)
    for name in ("pair-i", "pair-j"):
        t = mesh.faces[name]
        t.copy_(sig_t[t.long()].to(t.dtype))
    off = mesh.neighbors.offsets.cpu().numpy()
    lst = mesh.neighbors.faces.cpu().numpy()
    cnt = np.diff(off)
    new_cnt = cnt[inv]
    new_off = np.concatenate([[0], np.cumsum(new_cnt)]).astype(off.dtype)
    new_lst = np.concatenate([lst[off[inv[k]]:off[inv[k] + 1]] for k in range(n)]).astype(lst.dtype)
    mesh.neighbors.offsets.copy_(torch.as_tensor(new_off))
    mesh.neighbors.faces.copy_(torch.as_tensor(new_lst))
    mesh.invalidate_plan()
    return sigma


def _outside_entries(a, w0, w1):
    """List entries of the real cells in [w0, w1) that name a cell outside
    that range."""
    off, lst = a["nbr_off"], a["nbr_face"]
    k = 0
    for i in range(w0, w1):
        if a["tag"][i] != 0:
            continue
        ids = lst[off[i]:off[i + 1]]
        other = np.where(a["pair_i"][ids] == i, a["pair_j"][ids], a["pair_i"][ids])
        k += int(((other < w0) | (other >= w1)).sum())
    return k


@pytest.mark.parametrize("name,n,how", [c + ("random",) for c in CASES] + [("sedov_3d", 16, "degree")])
def test_core_step_on_renumbered_cells(gpu, name, n, how):
    from phd import problems
    from oracle import oracle

    print("case %s n=%d renumbered cells (%s)" % (name, n, how), flush=True)
    p = problems.PROBLEMS[name](n=n, seed=5)
    _renumber_cells(p, seed=13, by_degree=how == "degree")
    if how == "degree":
        assert _outside_entries(problems.host_arrays(p), 0, 256) > HALO_SORT_CAP
    step = problems.FrozenStep(p)
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
    g_gpu = p.recon.grad.rows().cpu().numpy()
    check_fields([g_gpu[:, q] for q in range(len(ref["grad"]))], ref["grad"], "grad",
                 mask=real, bitwise=True)
    flux = [t.cpu().numpy() for t in p.riemann.fluxes.group("conservative")]
    cons = [t.cpu().numpy() for t in p.particles.group("conservative")]
    check_fields(flux, ref["flux"], "flux", bitwise=True)
    check_fields(cons, ref["cons"], "cons", mask=real, bitwise=True)
