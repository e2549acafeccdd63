"""Index tables of the tied-RBM kernel (pygfmc_amd.torus): the torus is
recognised from the translation table, and every table read the kernel
makes (theta: block of w_f at code(k) + off(h); factors: the class block at
code(i) + off(h)) names W[h][k], W[h][i] and W[h][j] for every hidden unit
and slot, on square lattices and on a torus with a three-site basis."""
import numpy as np
import pytest

from pygfmc_amd import RBMSpinHalf, SpinHalf, chain, square
from pygfmc_amd.lattice import square_translations
from pygfmc_amd.torus import _torus_perm, find_torus, tied_tables


@pytest.mark.parametrize("Lx,Ly", [(4, 4), (10, 10), (6, 4), (3, 5)])
def test_find_torus_square(Lx, Ly):
    assert find_torus(square_translations(Lx, Ly)) == (1, Lx, Ly)


def test_find_torus_rejects_other_tables():
    T = square_translations(4, 4)
    assert find_torus(T[::-1]) is None          # rows in another order
    assert find_torus(T[:, ::-1]) is None
    assert find_torus(_torus_perm(3, 2, 2)) == (3, 2, 2)


def emulate(t, W, H, edge_i, edge_j):
    """Every read of the kernel, from the tables, against W: entry idx of the
    kernel's table holds (T[idx], T[idx - pair_code]) for a pair chunk (the
    lane's units a and b) and T[idx] for a single chunk."""
    F, Q, nT, D = t["F"], t["Q"], t["n_T"], t["pair_code"]
    n_pc, n_sc = t["n_pc"], t["n_sc"]
    lanes = 64
    N = W.shape[1]
    ncls = t["n_cls"]
    FQ = F * Q
    # the raw table as (filter, first site, partner site) per entry; the
    # weight block as (filter, site, -1); the zero block as (-1, -1, -1)
    raw = np.full(((2 * ncls + 1) * FQ + Q, 3), -1, dtype=np.int64)
    qp = t["q_pair"].reshape(ncls, Q)
    for c in range(ncls):
        for sg in range(2):
            for f in range(F):
                o = (2 * c + sg) * FQ + f * Q
                raw[o:o + Q, 0] = f
                raw[o:o + Q, 1] = t["q_site"]
                raw[o:o + Q, 2] = qp[c]
    for f in range(F):
        o = 2 * ncls * FQ + f * Q
        raw[o:o + Q, 0] = f
        raw[o:o + Q, 1] = t["q_site"]

    def w_of(e):          # the weight an entry names: w_f at its site
        return np.where(e[..., 0] >= 0, W[np.maximum(e[..., 0], 0) * nT, np.maximum(e[..., 1], 0)],
                        0.0)

    def wj_of(e):
        return W[e[..., 0] * nT, e[..., 2]]

    # register u of a lane reads entry idx(lane chunk) (x) or idx - D (y)
    regs = []
    for c in range(n_pc):
        regs.append((c, 0, t["lane_unit"][2 * c * lanes:(2 * c + 1) * lanes]))
        regs.append((c, D, t["lane_unit"][(2 * c + 1) * lanes:(2 * c + 2) * lanes]))
    for c in range(n_sc):
        regs.append((n_pc + c, 0,
                     t["lane_unit"][(2 * n_pc + c) * lanes:(2 * n_pc + c + 1) * lanes]))
    seen = np.concatenate([u for _, _, u in regs])
    assert np.array_equal(np.sort(seen[seen >= 0]), np.arange(H))
    code = t["slot_code"] & 0xfffff
    cls = (t["slot_code"] >> 20) & 0x3ff
    swap = (t["slot_code"] >> 30) & 1
    first = np.where(swap == 1, edge_j, edge_i)
    second = np.where(swap == 1, edge_i, edge_j)
    assert np.array_equal(code, t["site_code"][first])
    for c, shift, units in regs:
        off = t["lane_off"][c * lanes:(c + 1) * lanes]
        real = units >= 0
        # theta reads: the weight block at code(k)
        idx = 2 * ncls * FQ + t["site_code"][None, :] + off[:, None] - shift
        assert idx.min() >= 0 and idx.max() < len(raw)
        got = w_of(raw[idx])
        np.testing.assert_array_equal(got[real], W[units[real]])
        # factor reads of both signs: entry names w_f at first - r and
        # second - r
        for sg in range(2):
            idx = (2 * cls + sg)[None, :] * FQ + code[None, :] + off[:, None] - shift
            assert idx.min() >= 0 and idx.max() < len(raw)
            e = raw[idx[real]]
            np.testing.assert_array_equal(w_of(e), W[units[real]][:, first])
            np.testing.assert_array_equal(wj_of(e), W[units[real]][:, second])


@pytest.mark.parametrize("Lx,Ly,F", [(4, 4, 3), (10, 10, 4), (6, 4, 2), (6, 6, 2), (5, 5, 2)])
def test_tables_reproduce_the_weights_square(Lx, Ly, F):
    L = square(Lx, Ly)
    g = RBMSpinHalf.symmetric(L.extras["translations"], F, seed=Lx + Ly, scale=0.3)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float32)
    t = tied_tables(H.edge_i, H.edge_j, g)
    assert t is not None and t["torus"] == (1, Lx, Ly)
    assert t["n_cls"] == 2           # x and y bonds, each from its canonical end
    emulate(t, g.W, g.hidden, H.edge_i, H.edge_j)


def test_tables_reproduce_the_weights_chain():
    L = chain(12)
    T = square_translations(12, 1)
    g = RBMSpinHalf.symmetric(T, 3, seed=3, scale=0.3)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float32)
    t = tied_tables(H.edge_i, H.edge_j, g)
    assert t is not None and t["torus"] in ((1, 12, 1), (1, 1, 12))
    emulate(t, g.W, g.hidden, H.edge_i, H.edge_j)


def test_tables_reproduce_the_weights_with_a_basis():
    # a 3 x 2 torus of three-site cells, bonds inside the cell and to the
    # neighbouring cells (every pair of basis sites and displacement kinds)
    B, Lx, Ly = 3, 3, 2
    T = _torus_perm(B, Lx, Ly)
    site = lambda b, cx, cy: B * ((cx % Lx) + Lx * (cy % Ly)) + b  # noqa: E731
    edges = set()
    for cy in range(Ly):
        for cx in range(Lx):
            for a, b, dx, dy in ((0, 1, 0, 0), (0, 2, 0, 0), (1, 2, 0, 0), (1, 0, 1, 0),
                                 (2, 0, 0, 1), (1, 2, 1, -1)):
                i, j = site(a, cx, cy), site(b, cx + dx, cy + dy)
                if i != j:
                    edges.add((min(i, j), max(i, j)))
    edges = sorted(edges)
    g = RBMSpinHalf.symmetric(T, 2, seed=7, scale=0.3)
    H = SpinHalf.create(edges, -1.0, -1.0, 1.0, dtype=np.float32)
    t = tied_tables(H.edge_i, H.edge_j, g)
    assert t is not None and t["torus"] == (B, Lx, Ly)
    emulate(t, g.W, g.hidden, H.edge_i, H.edge_j)


def test_untied_guides_have_no_tables():
    L = square(4, 4)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float32)
    g = RBMSpinHalf(16, alpha=2, seed=1, scale=0.3)
    assert tied_tables(H.edge_i, H.edge_j, g) is None          # no translations
    t = RBMSpinHalf.symmetric(L.extras["translations"], 2, seed=1, scale=0.3)
    W = t.W.copy()
    W[5, 3] += 0.5                                               # tying broken
    broken = t.with_parameters(W, t.a, t.b)
    assert tied_tables(H.edge_i, H.edge_j, broken) is None
    assert tied_tables(H.edge_i, H.edge_j, t.with_parameters(t.W, t.a, t.b)) is not None
