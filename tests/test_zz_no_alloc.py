"""The allocation-free time step (phd_compute_time_step_noalloc) through the
C ABI equals phd_compute_time_step bit for bit, on two calls in a row (the
second on a smaller input, so that stale partial minima from the first call
would show), in 2D and 3D, boosted and not; inputs beyond its fixed scratch
are rejected with PHD_E_SIZE before any launch."""
import ctypes

import pytest
import torch

pytestmark = pytest.mark.gpu


def _dt(lib, h, c, gamma, cfl, boost):
    dt = ctypes.c_double(-1.0)
    assert lib.phd_compute_time_step(h, ctypes.byref(c), gamma, cfl, boost, ctypes.byref(dt)) == 0
    return dt.value


def _dt_noalloc(lib, c, gamma, cfl, boost):
    from phd import _native as nat
    dt = ctypes.c_double(-1.0)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = lib.phd_compute_time_step_noalloc(ctypes.byref(c), gamma, cfl, boost, stream,
                                           ctypes.byref(dt))
    assert rc == 0, nat.load_flux_library().phd_error_string(rc)
    return dt.value


@pytest.mark.parametrize("name,n", [("sedov_3d", 24), ("kh_2d", 160), ("sod_2d", 96)])
def test_noalloc_time_step_equals_handle_time_step(gpu, name, n):
    from phd import _native as nat
    from phd import problems
    from phd.views import cells_view
    print("case noalloc time step %s n=%d" % (name, n), flush=True)
    lib, h = nat.flux(), nat.handle()
    p = problems.PROBLEMS[name](n=n, seed=3)
    pc = p.particles
    # a real cell in the second half of the arrays sets the first call's
    # minimum (a sound speed a thousand times the largest); the second call,
    # over the first half, must not see it
    k = 3 * pc.get_carray_size() // 4
    assert int(pc["tag"][k]) == 0
    pc["pressure"][k] = 1e6 * float(pc["pressure"].max())
    c = cells_view(pc)
    gamma = p.eos.get_gamma()
    for boost in (1, 0):
        first = _dt_noalloc(lib, c, gamma, 0.3, boost)
        assert first == _dt(lib, h, c, gamma, 0.3, boost)
        # second call, smaller input: the cells of the first half only
        half = nat.Cells.from_buffer_copy(c)
        half.n_cells = c.n_cells // 2
        second = _dt_noalloc(lib, half, gamma, 0.3, boost)
        assert second == _dt(lib, h, half, gamma, 0.3, boost)
        assert second > first
    torch.cuda.synchronize()


def test_noalloc_rejects_inputs_beyond_its_scratch(gpu):
    from phd import _native as nat
    from phd import problems
    from phd.views import cells_view
    print("case noalloc size limit", flush=True)
    lib = nat.flux()
    p = problems.sod_2d(n=8, seed=0)
    c = cells_view(p.particles)
    big = nat.Cells.from_buffer_copy(c)
    big.n_cells = lib.phd_noalloc_max_cells() + 1      # rejected on the host, nothing launched
    dt = ctypes.c_double(-1.0)
    rc = lib.phd_compute_time_step_noalloc(ctypes.byref(big), 1.4, 0.3, 1, None, ctypes.byref(dt))
    assert rc == nat.PHD_E_SIZE and dt.value == -1.0
    assert lib.phd_noalloc_max_cells() == 65536 * 256
    ok = _dt_noalloc(lib, c, 1.4, 0.3, 1)
    assert ok > 0.0
