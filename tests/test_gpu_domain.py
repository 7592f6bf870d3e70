"""GPU domain and boundary loops through the C ABI, against the reference's
loops restated per element on the host (domain_manager.pyx:566-590, 757-770;
boundary.pyx:236-263, 309-348, 502-538).  Every result is bit-identical: the
device evaluates the same single IEEE operation per element.
"""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

LO, HI = [0.0, 0.0, 0.0], [1.0, 1.0, 1.0]


def _cells(dim, n_real, n_ghost, dev, seed=0):
    """A cell set with reals first, then ghosts (map -> a real image); the
    positions of some reals sit on or outside the box faces."""
    from phd import _native as nat
    rng = np.random.default_rng(seed)
    n = n_real + n_ghost
    a = {
        "tag": np.r_[np.zeros(n_real, np.int8), np.ones(n_ghost, np.int8)],
        "type": np.r_[np.full(n_real, 3, np.int8), np.full(n_ghost, 2, np.int8)],
        "map": np.r_[np.zeros(n_real, np.int32), rng.integers(0, n_real, n_ghost).astype(np.int32)],
        "x": [rng.uniform(-0.2, 1.2, n) for _ in range(dim)],
        "w": [rng.normal(size=n) for _ in range(dim)],
        "prim": [rng.normal(size=n) for _ in range(dim + 2)],
        "cons": [rng.normal(size=n) for _ in range(dim + 2)],
    }
    d = {k: (torch.as_tensor(v, device=dev) if not isinstance(v, list)
             else [torch.as_tensor(x, device=dev) for x in v]) for k, v in a.items()}
    c = nat.Cells()
    c.dim, c.n_cells = dim, n
    c.tag, c.type, c.map = d["tag"].data_ptr(), d["type"].data_ptr(), d["map"].data_ptr()
    for k in range(dim):
        c.x[k], c.w[k] = d["x"][k].data_ptr(), d["w"][k].data_ptr()
    for q in range(dim + 2):
        c.prim[q], c.cons[q] = d["prim"][q].data_ptr(), d["cons"][q].data_ptr()
    return a, d, c


def _box(dim):
    from phd import _native as nat
    return nat.DblArray3(*LO), nat.DblArray3(*HI)


@pytest.mark.parametrize("dim", [2, 3])
def test_move_generators(gpu, dim):
    from phd import _native as nat
    print("case move_generators dim=%d" % dim, flush=True)
    lib, h = nat.flux(), nat.handle()
    a, d, c = _cells(dim, 5000, 700, gpu, seed=dim)
    dt = 0.0137
    assert lib.phd_move_generators(h, ctypes.byref(c), dt) == 0
    torch.cuda.synchronize()
    real = a["tag"] == 0
    for k in range(dim):
        want = a["x"][k].copy()
        want[real] = want[real] + dt * a["w"][k][real]        # domain_manager.pyx:588-590
        np.testing.assert_array_equal(d["x"][k].cpu().numpy(), want)


@pytest.mark.parametrize("dim", [2, 3])
def test_migrate_periodic(gpu, dim):
    from phd import _native as nat
    print("case migrate_periodic dim=%d" % dim, flush=True)
    lib, h = nat.flux(), nat.handle()
    a, d, c = _cells(dim, 4000, 300, gpu, seed=10 + dim)
    # edge cases on cells 0..3: exactly on xmin while outside on the next
    # axis; just below xmin where x + L rounds to xmax; exactly on xmax while
    # outside on another axis; on xmin but inside on every axis (not moved)
    edits = [(0, 0, 0.0, 1, -0.05), (1, 0, -2.0 ** -60, 1, 0.5), (2, 0, 1.0, 1, 1.3),
             (3, 0, 0.0, 1, 0.5)]
    for i, k0, v0, k1, v1 in edits:
        for k in range(dim):
            v = v0 if k == k0 else (v1 if k == k1 else 0.5)
            d["x"][k][i] = v
            a["x"][k][i] = v
    lo, hi = _box(dim)
    assert lib.phd_migrate_periodic(h, ctypes.byref(c), lo, hi) == 0
    torch.cuda.synchronize()
    want = [x.copy() for x in a["x"]]
    for i in range(a["tag"].size):                        # boundary.pyx:523-538
        if a["tag"][i] != 0:
            continue
        xs = [a["x"][k][i] for k in range(dim)]
        if all(LO[k] <= xs[k] <= HI[k] for k in range(dim)):
            continue
        for k in range(dim):
            if xs[k] <= LO[k]:
                want[k][i] += HI[k] - LO[k]
            if xs[k] >= HI[k]:
                want[k][i] -= HI[k] - LO[k]
    got = [x.cpu().numpy() for x in d["x"]]
    for k in range(dim):
        np.testing.assert_array_equal(got[k], want[k])
    # the hand cases: moved by exactly +L / -L once, or not at all
    assert got[0][0] == 1.0 and got[0][1] == -2.0 ** -60 + 1.0 and got[0][2] == 0.0
    assert got[0][3] == 0.0


@pytest.mark.parametrize("dim", [2, 3])
def test_first_outside_and_reflect_fields(gpu, dim):
    from phd import _native as nat
    print("case first_outside/reflect_fields dim=%d" % dim, flush=True)
    lib, h = nat.flux(), nat.handle()
    a, d, c = _cells(dim, 3000, 500, gpu, seed=20 + dim)
    lo, hi = _box(dim)
    first = ctypes.c_int64(-2)
    assert lib.phd_first_outside(h, ctypes.byref(c), lo, hi, ctypes.byref(first)) == 0
    real = a["tag"] == 0
    out = np.zeros(real.size, bool)
    for k in range(dim):
        out |= (a["x"][k] < LO[k]) | (a["x"][k] > HI[k])
    idx = np.nonzero(out & real)[0]
    assert first.value == (idx[0] if idx.size else -1)
    # all reals inside: -1
    for k in range(dim):
        d["x"][k][:3000].clamp_(0.0, 1.0)
    assert lib.phd_first_outside(h, ctypes.byref(c), lo, hi, ctypes.byref(first)) == 0
    assert first.value == -1
    xs = [x.cpu().numpy() for x in d["x"]]
    assert lib.phd_reflect_fields(h, ctypes.byref(c), lo, hi) == 0
    torch.cuda.synchronize()
    ext = a["type"] == 2
    for k in range(dim):                                   # boundary.pyx:336-348
        flip = np.where(ext & (xs[k] < LO[k]), -1.0, 1.0) * np.where(ext & (xs[k] > HI[k]), -1.0, 1.0)
        np.testing.assert_array_equal(d["prim"][1 + k].cpu().numpy(), a["prim"][1 + k] * flip)
        np.testing.assert_array_equal(d["cons"][1 + k].cpu().numpy(), a["cons"][1 + k] * flip)
    for q in (0, dim + 1):
        np.testing.assert_array_equal(d["prim"][q].cpu().numpy(), a["prim"][q])
        np.testing.assert_array_equal(d["cons"][q].cpu().numpy(), a["cons"][q])


def test_update_ghost_fields_every_width(gpu):
    from phd import _native as nat
    print("case update_ghost_fields", flush=True)
    lib, h = nat.flux(), nat.handle()
    a, d, c = _cells(3, 6000, 900, gpu, seed=30)
    n = a["tag"].size
    rng = np.random.default_rng(31)
    fields = [torch.as_tensor(rng.normal(size=n), device=gpu),
              torch.as_tensor(rng.integers(-2 ** 40, 2 ** 40, n), device=gpu),
              torch.as_tensor(rng.integers(-2 ** 30, 2 ** 30, n).astype(np.int32), device=gpu),
              torch.as_tensor(rng.integers(-100, 100, n).astype(np.int8), device=gpu)]
    fields = fields * 5                                  # 20 fields: two launches
    before = [f.cpu().numpy().copy() for f in fields]
    ptrs = (ctypes.c_void_p * len(fields))(*[f.data_ptr() for f in fields])
    size = (ctypes.c_int32 * len(fields))(*[f.element_size() for f in fields])
    assert lib.phd_update_ghost_fields(h, ctypes.byref(c), ptrs, size, len(fields)) == 0
    torch.cuda.synchronize()
    g = np.nonzero(a["tag"] == 1)[0]
    for f, b in zip(fields, before[:4]):                 # domain_manager.pyx:765-770
        want = b.copy()
        want[g] = b[a["map"][g]]
        np.testing.assert_array_equal(f.cpu().numpy(), want)
    bad = (ctypes.c_int32 * 1)(3)
    assert lib.phd_update_ghost_fields(h, ctypes.byref(c), ptrs, bad, 1) == nat.PHD_E_ARG
    c.map = None
    assert lib.phd_update_ghost_fields(h, ctypes.byref(c), ptrs, size, 1) == nat.PHD_E_ARG


def test_null_map_with_ghosts_is_an_argument_error(gpu):
    """A plan that lists ghosts needs c->map (the ghost gradients and the
    ghost refresh read the images); without a plan the refresh scans the
    tags."""
    from phd import _native as nat
    from phd import problems
    print("case null map with ghosts", flush=True)
    lib, h = nat.flux(), nat.handle()
    prob = problems.sod_2d(n=12, seed=0)
    step = problems.FrozenStep(prob)
    c = step.cells
    assert int((prob.particles["tag"] == 1).sum()) > 0
    c.map = None
    lo, hi = nat.DblArray3(*LO), nat.DblArray3(*HI)
    lo_, hi_ = ctypes.cast(lo, ctypes.c_void_p), ctypes.cast(hi, ctypes.c_void_p)
    rec = step.grad
    assert lib.phd_update_ghost_gradients(h, step.plan, ctypes.byref(c), nat.BOUNDARY_REFLECTIVE,
                                          lo_, hi_, rec.data.data_ptr(), rec.ld) == nat.PHD_E_ARG
    assert lib.phd_ghost_refresh_eos(h, step.plan, ctypes.byref(c), nat.BOUNDARY_REFLECTIVE,
                                     lo_, hi_, 1.4) == nat.PHD_E_ARG
    assert lib.phd_ghost_refresh_eos(h, None, ctypes.byref(c), nat.BOUNDARY_REFLECTIVE,
                                     lo_, hi_, 1.4) == nat.PHD_E_ARG
    torch.cuda.synchronize()
    nat.raise_device_fault(h)
