"""Lattice invariants and the keyword contract of run / run_gfmc (no GPU).

The kagome edge set defines the 432-site workload model, so it is checked
against the geometry itself, independently of the factory's bond template:
N = 3 Lx Ly sites, 2N edges, coordination 4, every edge of unit length under
the minimum image of the torus and every unit-distance pair an edge, every
up triangle and every hexagon side a bond (the reference factory asserts the
hexagon sides, codes/pygfmc/hamiltonian/lattice.py:247-250).
"""
import numpy as np
import pytest

from pygfmc_amd import kagome, run, run_gfmc, square


def min_image_dist(pos, torus, i, j):
    d = pos[j] - pos[i]
    best = np.inf
    for m in (-1, 0, 1):
        for n in (-1, 0, 1):
            best = min(best, np.linalg.norm(d + m * torus[0] + n * torus[1]))
    return best


def site_at(pos, torus, point):
    """Index of the site at ``point`` modulo the torus (None if absent)."""
    for m in (-1, 0, 1, 2):
        for n in (-1, 0, 1, 2):
            q = point - m * torus[0] - n * torus[1]
            k = np.flatnonzero(np.linalg.norm(pos - q, axis=1) < 1e-9)
            if k.size:
                return int(k[0])
    return None


@pytest.mark.parametrize("Lx,Ly", [(2, 2), (3, 3), (2, 5), (12, 12)])
def test_kagome_edges_are_the_geometric_nearest_neighbours(Lx, Ly):
    lat = kagome(Lx, Ly)
    a1, a2 = np.array([2.0, 0.0]), np.array([1.0, np.sqrt(3.0)])
    torus = np.array([Lx * a1, Ly * a2])
    N = 3 * Lx * Ly
    assert lat.N == N
    pos = np.array([x * a1 + y * a2 + b for y in range(Ly) for x in range(Lx)
                    for b in (np.zeros(2), a1 / 2, a2 / 2)])
    np.testing.assert_allclose(lat.extras["positions"], pos)
    edges = set(map(tuple, lat.edges))
    assert len(edges) == len(lat.edges) == 2 * N
    assert all(i < j for i, j in edges)
    deg = np.zeros(N, dtype=int)
    for i, j in edges:
        deg[i] += 1
        deg[j] += 1
        assert min_image_dist(pos, torus, i, j) == pytest.approx(1.0, abs=1e-12)
    assert np.all(deg == 4)
    # completeness: every pair at unit distance under the minimum image is a bond
    if N <= 108:
        unit = {(i, j) for i in range(N) for j in range(i + 1, N)
                if abs(min_image_dist(pos, torus, i, j) - 1.0) < 1e-9}
        assert unit == edges
    # up triangles inside each cell and every side of each cell's hexagon
    offs = [np.array([np.cos(k * np.pi / 3), np.sin(k * np.pi / 3)]) for k in range(6)]
    for y in range(Ly):
        for x in range(Lx):
            b = 3 * (x + Lx * y)
            for i, j in ((b, b + 1), (b, b + 2), (b + 1, b + 2)):
                assert (i, j) in edges
            c = x * a1 + y * a2 + (a1 + a2) / 2
            hexa = [site_at(pos, torus, c + o) for o in offs]
            assert None not in hexa and len(set(hexa)) == 6
            for k in range(6):
                i, j = hexa[k], hexa[(k + 1) % 6]
                assert (min(i, j), max(i, j)) in edges


def test_kagome_rejects_small_tori():
    with pytest.raises(ValueError, match="Lx, Ly >= 2"):
        kagome(1, 3)


@pytest.mark.parametrize("L", [3, 4, 10, 32])
def test_square_edges_are_the_nearest_neighbours(L):
    lat = square(L, L)
    N = L * L
    assert lat.N == N and len(lat.edges) == 2 * N == len(set(map(tuple, lat.edges)))
    deg = np.zeros(N, dtype=int)
    for i, j in lat.edges:
        xi, yi, xj, yj = i % L, i // L, j % L, j // L
        dx, dy = min(abs(xi - xj), L - abs(xi - xj)), min(abs(yi - yj), L - abs(yi - yj))
        assert dx + dy == 1
        deg[i] += 1
        deg[j] += 1
    assert np.all(deg == 4)


def test_unknown_keyword_is_a_type_error():
    """run_gfmc has no catch-all (driver.py:1148-1158): a misspelt keyword
    fails at the call, before anything runs."""
    with pytest.raises(TypeError):
        run_gfmc(None, 10, None, 0.2, None, [2], warmup_step=5)
    with pytest.raises(TypeError):
        run(None, 10, None, 0.2, None, [2], number_of_output=3)


@pytest.mark.parametrize("bad", [0, -1, 1.5, True, "two"])
def test_n_devices_is_validated(bad):
    with pytest.raises(ValueError, match="n_devices"):
        run(None, 10, None, 0.2, None, [2], n_devices=bad)


def test_more_than_one_device_is_refused():
    with pytest.raises(ValueError, match="one GPU"):
        run(None, 10, None, 0.2, None, [2], n_devices=8)


def test_step_offset_must_be_zero():
    with pytest.raises(ValueError, match="rng_key_next"):
        run_gfmc(None, 10, None, 0.2, None, [2], _step_offset=5)


def test_square_translations_are_lattice_automorphisms():
    L = square(5, 4)
    T = L.extras["translations"]
    assert T.shape == (20, 20)
    edges = set(map(tuple, L.edges))
    for r in range(20):
        assert sorted(T[r]) == list(range(20))
        for i, j in L.edges:
            a, b = int(T[r, i]), int(T[r, j])
            assert (min(a, b), max(a, b)) in edges


def test_translation_invariant_rbm_ties_its_parameters():
    from pygfmc_amd import RBMSpinHalf
    from pygfmc_amd.training import flat_params
    L = square(4, 4)
    g = RBMSpinHalf.symmetric(L.extras["translations"], 3, seed=2, scale=0.3)
    assert g.hidden == 48
    flat = flat_params(g)
    for o in np.unique(g.orbits):
        vals = flat[g.orbits == o]
        assert np.all(vals == vals[0])
    assert len(np.unique(g.orbits)) == 3 * 16 + 3 + 1


def test_rbm_filter_support_freezes_the_entries_outside_it():
    """A filter support keeps the entries outside it at 0 in every hidden
    unit (orbit -1, in no orbit table), and the orbit tables hold exactly the
    live parameters."""
    import torch

    from pygfmc_amd import RBMSpinHalf
    from pygfmc_amd.training import flat_params, orbit_tables
    L = square(6, 6)
    x, y = np.arange(36) % 6, np.arange(36) // 6
    dx, dy = np.minimum(x, 6 - x), np.minimum(y, 6 - y)
    sup = np.maximum(dx, dy) <= 1
    g = RBMSpinHalf.from_jastrow(L.edges, L.extras["translations"], 0.2, 3, seed=4, scale=0.3,
                                 support=sup)
    T = L.extras["translations"]
    for f in range(3):
        for r in range(36):
            row = g.W[f * 36 + r]
            assert np.all(row[T[r][~sup]] == 0.0)
    flat = flat_params(g)
    frozen = g.orbits < 0
    assert frozen.sum() == 3 * 36 * (36 - sup.sum())
    assert np.all(flat[frozen] == 0.0)
    off, mem, m = orbit_tables(g.orbits, "cpu")
    assert m == 3 * int(sup.sum()) + 1 + 3
    assert sorted(mem.numpy().tolist()) == np.flatnonzero(~frozen).tolist()
    assert int(off[-1]) == int((~frozen).sum())
    with pytest.raises(ValueError, match="support must hold"):
        RBMSpinHalf.from_jastrow(L.edges, T, 0.2, 3, support=(dx + dy) == 0)


def test_rbm_from_jastrow_is_the_jastrow_guide():
    """log psi_RBM - log psi_Jastrow is the same constant on every
    configuration (the BF16 rounding of the filter weight moves beta to
    lncosh(2 bf16(w)) / 2)."""
    from helpers import basis, rbm_log_psi
    from pygfmc_amd import JastrowSpinHalf, RBMSpinHalf
    L = square(4, 4)
    g = RBMSpinHalf.from_jastrow(L.edges, L.extras["translations"], 0.15, 4, scale=0.0)
    w = float(g.W[0, 0])
    beta_eff = np.log(np.cosh(2 * w)) / 2
    assert abs(beta_eff - 0.15) < 2e-3
    J = JastrowSpinHalf(16, L.edges, beta=beta_eff)
    x = basis(16)[::97]
    lj = -beta_eff * np.sum([x[:, i] * x[:, j] for i, j in L.edges], axis=0)
    d = rbm_log_psi(g, x) - lj
    np.testing.assert_allclose(d, d[0], atol=1e-9)
    assert J.N == 16
