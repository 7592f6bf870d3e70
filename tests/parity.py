"""Shared comparison helpers for the parity tests.

Tolerance (from the parity target of the port): |got - want| <=
1e-12 * max(|got|, |want|, s) with s = 1e-6 * max|want| over the array, a
per-array floor for entries that are round-off-sized relative to the array's
scale (static regions give exact zeros in one code and 1e-20-sized values in
the other).  Bitwise agreement is reported separately.
"""
import numpy as np

RTOL = 1e-12


def rel_check(got, want, rtol=RTOL, floor_frac=1e-6, name="", scale=None):
    """scale: a per-entry floor (the magnitude of the terms each entry is
    formed from) in place of floor_frac of the array maximum."""
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    assert got.shape == want.shape, (name, got.shape, want.shape)
    if want.size == 0:
        return 1.0
    if scale is None:
        s = floor_frac * float(np.max(np.abs(want)))
    else:
        s = np.asarray(scale, dtype=np.float64)
        assert s.shape == want.shape and np.all(s > 0), name
    scale = np.maximum(np.maximum(np.abs(got), np.abs(want)), s)
    scale = np.where(scale == 0, 1.0, scale)
    err = np.abs(got - want) / scale
    worst = float(np.max(err))
    assert np.all(np.isfinite(got) == np.isfinite(want)), name
    assert worst <= rtol, "%s: worst relative difference %.3e (index %d: got %r want %r)" % (
        name, worst, int(np.argmax(err)), got.flat[int(np.argmax(err))],
        want.flat[int(np.argmax(err))])
    return float(np.mean(got == want))


def check_fields(got_list, want_list, name, mask=None, bitwise=False, scales=None, **kw):
    """rel_check per field; with bitwise=True every entry must also be
    bit-identical (checked after the tolerance, so a failure reports the
    worst relative difference first); scales: per-field per-entry floors
    (see rel_check).  Returns the smallest bit-identical fraction."""
    fr = []
    for q, (g, w) in enumerate(zip(got_list, want_list)):
        g = np.asarray(g)
        w = np.asarray(w)
        sc = None if scales is None else np.asarray(scales[q])
        if mask is not None:
            g, w = g[mask], w[mask]
            sc = None if sc is None else sc[mask]
        fr.append(rel_check(g, w, name="%s[%d]" % (name, q), scale=sc, **kw))
        if bitwise:
            assert_bitwise(g, w, "%s[%d]" % (name, q))
    return min(fr) if fr else 1.0


def assert_bitwise(got, want, name=""):
    """Exact equality of two float64 arrays (NaNs equal when both are NaN);
    reports the first differing entry and the count."""
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    assert got.shape == want.shape, (name, got.shape, want.shape)
    same = (got == want) | (np.isnan(got) & np.isnan(want))
    if not np.all(same):
        bad = np.nonzero(~same)[0]
        i = int(bad[0])
        raise AssertionError("%s: %d of %d entries differ; first at %d: got %r want %r" % (
            name, bad.size, got.size, i, got[i], want[i]))


def assert_time_step(dt, dt_ref, dim, name=""):
    """The CFL time step of RiemannBase.compute_time_step: identical in 2D,
    where the cell radius is sqrt(V / pi) (correctly rounded on both sides);
    in 3D the radius is pow(3 V / (4 pi), 1/3) (riemann.pyx:155-158), and
    the GPU's pow and the host libm's are each within an ulp of the true
    power without always giving the same bits (glibc's own pow is not
    correctly rounded on about 1 in 1300 operands), so the minimum may differ
    by one ulp."""
    if dim == 2:
        assert dt == dt_ref, (name, dt, dt_ref)
    else:
        lo, hi = np.nextafter(dt_ref, 0.0), np.nextafter(dt_ref, np.inf)
        assert lo <= dt <= hi, (name, dt, dt_ref)


def flux_scales(left, right, gamma, fv=None):
    """Per face and flux component, the magnitude of the terms the flux is
    formed from, from the two primitive states of the face (density,
    velocity components, pressure; lists of arrays) and the face velocity:
    with s = |v| + |w| + c, mass rho s, momentum rho s^2 + p, energy
    (rho s^2 / 2 + gamma p / (gamma - 1)) s, the larger side's.  An exact
    Riemann solution's sampled state (a star state included) stays within a
    small factor of these, so round-off in the flux is about 1e-16 of them:
    the parity bound 1e-12 * max(|flux|, scale) is tight per face, however
    small the face's flux is against the array's largest."""
    dim = len(left) - 2
    wmag = 0.0 if fv is None else np.sqrt(sum(np.asarray(f, dtype=np.float64) ** 2 for f in fv))
    out = None
    for st in (left, right):
        rho = np.abs(np.asarray(st[0], dtype=np.float64))
        p = np.abs(np.asarray(st[dim + 1], dtype=np.float64))
        v = np.sqrt(sum(np.asarray(st[1 + k], dtype=np.float64) ** 2 for k in range(dim)))
        s = v + wmag + np.sqrt(gamma * p / rho)
        sc = [rho * s] + [rho * s * s + p] * dim + [(0.5 * rho * s * s + gamma * p / (gamma - 1.0)) * s]
        out = sc if out is None else [np.maximum(a, b) for a, b in zip(out, sc)]
    return out
