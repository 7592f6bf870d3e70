"""Union-mesh assembly of a batched ensemble (host only): global indices,
owner order, problem isolation, face lists and ghost maps."""
import numpy as np
import pytest


def _host(n=8, seed=3, dim=2):
    from phd import ensemble as E
    return E.HostEnsemble(E.draw_problems(n, seed=seed, n_min=100, n_max=900, dim=dim),
                          n_workers=4)


def test_draws_are_seeded_and_in_range():
    from phd import ensemble as E
    a, b = E.draw_problems(64, seed=7), E.draw_problems(64, seed=7)
    for k in a:
        assert np.array_equal(a[k], b[k])
    cells = a["side"] ** 2
    assert cells.min() >= 900 and cells.max() <= 64516      # rounded sqrt of [1k, 64k]
    for k in ("rho_l", "rho_r", "p_l", "p_r"):
        assert a[k].min() >= 0.1 and a[k].max() <= 10.0
    assert np.abs(a["v_l"]).max() <= 1.0 and np.abs(a["v_r"]).max() <= 1.0
    c3 = E.draw_problems(64, seed=7, dim=3)
    assert np.array_equal(c3["rho_l"], a["rho_l"])            # same draws, cube-root sides
    assert (c3["side"] ** 3).min() >= 729 and (c3["side"] ** 3).max() <= 68921


@pytest.mark.parametrize("dim", [2, 3])
def test_union_layout(dim):
    H = _host(dim=dim)
    cs, fs = H.cell_start, H.face_start
    assert cs[0] == 0 and cs[-1] == H.n_cells and fs[-1] == H.n_faces
    # each block: reals then ghosts
    for p in range(H.n_problems):
        t = H.tag[cs[p]:cs[p + 1]]
        assert np.all(t[:H.n_real[p]] == 0) and np.all(t[H.n_real[p]:] == 1)
    # faces in owner order, owner real, both ends in the face's problem
    assert np.all(np.diff(H.pair_i) >= 0) and np.all(H.pair_i < H.pair_j)
    assert np.all(H.tag[H.pair_i] == 0)
    pi = np.searchsorted(cs, H.pair_i, side="right") - 1
    pj = np.searchsorted(cs, H.pair_j, side="right") - 1
    assert np.array_equal(pi, pj)
    assert np.array_equal(pi, np.repeat(np.arange(H.n_problems), H.n_face))
    # global = local + block start
    assert np.array_equal(H.pair_i - np.repeat(cs[:-1], H.n_face), H.pair_i_local)


@pytest.mark.parametrize("dim", [2, 3])
def test_face_lists_and_ghost_maps(dim):
    H = _host(dim=dim)
    assert H.nbr_off[-1] == H.nbr_face.size and np.all(np.diff(H.nbr_off) >= 0)
    for c in range(H.n_cells):
        fl = H.nbr_face[H.nbr_off[c]:H.nbr_off[c + 1]]
        assert np.all(np.diff(fl) > 0)                       # ascending face ids
        assert np.all((H.pair_i[fl] == c) | (H.pair_j[fl] == c))
    # every face is listed by both of its cells
    cnt = np.bincount(H.nbr_face, minlength=H.n_faces)
    assert np.all(cnt == 2)
    g = np.where(H.tag == 1)[0]
    cs = H.cell_start
    assert np.array_equal(np.searchsorted(cs, H.map[g], side="right") - 1,
                          np.searchsorted(cs, g, side="right") - 1)
    assert np.all(H.tag[H.map[g]] == 0)
    # local per-problem offsets describe the same lists
    for p in range(H.n_problems):
        n = cs[p + 1] - cs[p]
        lo = H.nbr_off_local[cs[p] + p: cs[p] + p + n + 1]
        assert np.array_equal(lo - lo[0], H.nbr_off[cs[p]:cs[p + 1] + 1] - H.nbr_off[cs[p]])


@pytest.mark.parametrize("dim", [2, 3])
def test_initial_state_is_the_riemann_split(dim):
    H = _host(dim=dim)
    s = H.spec
    pid = np.repeat(np.arange(H.n_problems), H.cell_start[1:] - H.cell_start[:-1])
    left = H.x[0] < 0.5
    assert np.array_equal(H.density[left], s["rho_l"][pid][left])
    assert np.array_equal(H.pressure[~left], s["p_r"][pid][~left])
    assert np.array_equal(H.vx[left], s["v_l"][pid][left])
    assert len(H.x) == len(H.v) == dim and not any(v.any() for v in H.v[1:])
