"""Pin the oracle's MLP guide on the CPU: the tabulated tanh against tanh,
log psi against an independent vectorised NumPy evaluation of the same
definition, and the local energies against the Kronecker-product matrix."""
import numpy as np
import pytest

from helpers import basis, index_of, spinhalf_matrix
from pygfmc_amd import MLPSpinHalf, SpinHalf, chain, kagome
from pygfmc_amd.guides import bf16_flush, mlp_tanh_table
from pygfmc_amd.problem import lower


def test_tabulated_tanh(orc):
    z = np.linspace(-12, 12, 48001).astype(np.float32)
    t = orc.mlp_tanh(z)
    assert np.max(np.abs(t - np.tanh(z.astype(np.float64)))) < 2e-6
    np.testing.assert_array_equal(orc.mlp_tanh(-z), -t)
    T, D = orc.mlp_table()
    nodes = (np.arange(2048) / 256.0).astype(np.float32)
    np.testing.assert_array_equal(orc.mlp_tanh(nodes), T)
    assert orc.mlp_tanh(np.float32(50.0)) == np.float32(np.tanh(8.0))
    T2, D2 = mlp_tanh_table()
    np.testing.assert_array_equal(T, T2)
    np.testing.assert_array_equal(D, D2)


def _bf16(a):
    f = np.ascontiguousarray(a, dtype=np.float32)
    b = f.view(np.uint32).astype(np.uint64)
    return ((((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16) & 0xFFFFFFFF).astype(
        np.uint32).view(np.float32)


def _tanh_tab(z, T, D):
    z = np.asarray(z, dtype=np.float32)
    a = np.abs(z)
    t = np.minimum(a * np.float32(256.0), np.float32(2048.0))
    j = np.minimum(t.astype(np.int64), 2047)
    f = (t - j.astype(np.float32)).astype(np.float32)
    h = (f.astype(np.float64) * D[j].astype(np.float64) + T[j].astype(np.float64)).astype(
        np.float32)
    return np.copysign(h, z).astype(np.float32)


def numpy_log_psi(g, configs):
    """The MLP guide's definition in matrix form: exact FP64 layer 1, BF16
    activations, layer 2 in FP64 rounded to FP32, the output in the order
    (c mod 32 over c / 32, then a butterfly over the 32 columns)."""
    T, D = mlp_tanh_table()
    x = np.atleast_2d(np.asarray(configs, dtype=np.float64))
    W1 = bf16_flush(g.W1).astype(np.float64)
    b1 = bf16_flush(g.b1).astype(np.float64)
    W2 = bf16_flush(g.W2).astype(np.float64)
    z1 = x @ W1.T + b1
    h = _bf16(_tanh_tab(z1.astype(np.float32), T, D))
    z2 = (h.astype(np.float64) @ W2.T).astype(np.float32)
    v = (g.w3 * _tanh_tab(z2 + g.b2, T, D)).astype(np.float32)
    NB = g.hidden[1] // 32
    s = np.zeros((x.shape[0], 32), dtype=np.float32)
    for t in range(NB):
        s = (s + v[:, t * 32:(t + 1) * 32]).astype(np.float32)
    for d in (16, 8, 4, 2, 1):
        s = (s + s[:, np.arange(32) ^ d]).astype(np.float32)
    return (s[:, 0] + np.float32(g.b3)).astype(np.float32)


NETS = [((32, 32), 12, 0), ((64, 64), 12, 1), ((128, 256), 24, 2), ((96, 128), 8, 3)]


@pytest.mark.parametrize("hidden,N,seed", NETS, ids=[f"{h[0]}x{h[1]}" for h, _, _ in NETS])
def test_log_psi_equals_the_vectorised_definition(orc, hidden, N, seed):
    H = SpinHalf.create(chain(N).edges, -1.0, -1.0, 1.0, dtype=np.float32)
    g = MLPSpinHalf(N, hidden, seed=seed, scale=1.5, out_scale=1.0)
    cfg = np.random.default_rng(seed).choice([-1.0, 1.0], size=(40, N))
    got = orc.Problem(lower(H, g)).log_psi(orc.pack(cfg))
    np.testing.assert_array_equal(got.astype(np.float32), numpy_log_psi(g, cfg))
    assert np.std(got) > 0.05


@pytest.mark.parametrize("name", ["kagome2x2", "chain8_hx"])
def test_local_energies_equal_the_matrix_column(orc, name):
    """e_L(x) = sum_x' H_x'x psi(x') / psi(x) over the Kronecker matrix, with
    psi from the oracle's log psi, for every configuration."""
    if name == "kagome2x2":
        L = kagome(2, 2)
        H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float32)
    else:
        H = SpinHalf.create(chain(8).edges, -1.0, -0.4, 0.7, h_x=0.6, dtype=np.float32)
    g = MLPSpinHalf(H.N, (64, 32), seed=5, scale=1.5, out_scale=1.0)
    P = orc.Problem(lower(H, g))
    cfg = basis(H.N)
    sp = orc.pack(cfg)
    lp = P.log_psi(sp).astype(np.float32)
    _, e_loc, _ = P.rates(sp)
    M = spinhalf_matrix(H).tocsc()
    idx = index_of(cfg)
    pos = np.empty_like(idx)
    pos[idx] = np.arange(len(idx))
    rng = np.random.default_rng(0)
    for c in rng.choice(len(cfg), size=60, replace=False):
        col = M[:, idx[c]]
        rows, vals = col.indices, col.data
        lr = (lp[pos[rows]] - lp[c]).astype(np.float32).astype(np.float64)
        want = float(np.sum(vals * np.exp(lr)))
        assert abs(e_loc[c] - want) <= 1e-9 * (1 + abs(want))
