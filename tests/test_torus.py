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
    """Every read of the kernel, from the tables, against W."""
    F, Q, nT = t["F"], t["Q"], t["n_T"]
    Hp = len(t["unit_off"])
    wt = np.zeros((F + 1) * Q)               # the w block, then the zero block
    for f in range(F):
        wt[f * Q:(f + 1) * Q] = W[f * nT][t["q_site"]]
    N = W.shape[1]
    # theta reads
    got = wt[t["site_code"][None, :] + t["unit_off"][:, None]]
    ref = np.zeros((Hp, N))
    ref[:H] = W
    np.testing.assert_array_equal(got, ref)
    # factor reads: entry q of block f names i - r (w_f there is W[h][i]) and,
    # through the class's pair table, j - r (W[h][j])
    code = t["slot_code"] & 0xfffff
    cls = t["slot_code"] >> 20
    h = np.arange(H)
    f = h // nT
    q = code[None, :] + t["unit_off"][:H, None] - (f * Q)[:, None]
    assert q.min() >= 0 and q.max() < Q
    wi = W[(f * nT)[:, None], t["q_site"][q]]
    pj = t["q_pair"].reshape(t["n_cls"], Q)[cls[None, :], q]
    assert pj.min() >= 0
    wj = W[(f * nT)[:, None], pj]
    np.testing.assert_array_equal(wi, W[:, edge_i])
    np.testing.assert_array_equal(wj, W[:, edge_j])


@pytest.mark.parametrize("Lx,Ly,F", [(4, 4, 3), (10, 10, 4), (6, 4, 2)])
def test_tables_reproduce_the_weights_square(Lx, Ly, F):
    L = square(Lx, Ly)
    g = RBMSpinHalf.symmetric(L.extras["translations"], F, seed=Lx + Ly, scale=0.3)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float32)
    t = tied_tables(H.edge_i, H.edge_j, g)
    assert t is not None and t["torus"] == (1, Lx, Ly)
    assert t["n_cls"] == 4
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
