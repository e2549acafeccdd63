"""Pin the oracle's MLP guide on the CPU: the tabulated tanh against tanh,
log psi against an independent vectorised NumPy evaluation of the same
definition, and the local energies against the Kronecker-product matrix."""
import numpy as np
import pytest

from helpers import basis, index_of, spinhalf_matrix
from pygfmc_amd import MLPSpinHalf, SpinHalf, chain, kagome
from pygfmc_amd.guides import bf16_flush, mlp_l1_grid, mlp_tanh_table
from pygfmc_amd.problem import lower


def test_tabulated_tanh(orc):
    z = np.linspace(-12, 12, 48001).astype(np.float32)
    t = orc.mlp_tanh(z)
    assert np.max(np.abs(t - np.tanh(z.astype(np.float64)))) < 2e-6
    np.testing.assert_array_equal(orc.mlp_tanh(-z), -t)
    T, D = orc.mlp_table()
    nodes = (np.arange(2048) / 256.0).astype(np.float32)
    np.testing.assert_array_equal(orc.mlp_tanh(nodes), T)
    sat = orc.mlp_tanh(np.float32(50.0))
    assert sat == orc.mlp_tanh(np.float32(8.5)) and abs(sat - np.tanh(8.0)) < 1e-6
    T2, D2 = mlp_tanh_table()
    np.testing.assert_array_equal(T, T2)
    np.testing.assert_array_equal(D, D2)


def test_layer1_grid_is_the_oracles_and_integral_at_scale_2_16(orc):
    """W1 and b1 live on the layer-1 grid: BF16 values with magnitudes below
    2^-9 flushed and above 64 clamped; every grid value times 2^16 is an
    integer of magnitude at most 2^22 (so the layer-1 sums of up to 510 sites
    are exact 32-bit integers)."""
    edge = np.array([0.0, 2.0 ** -9, -(2.0 ** -9), 2.0 ** -9 * (1 - 2.0 ** -9),
                     2.0 ** -9 * (1 - 2.0 ** -7), 2.0 ** -10, 2.0 ** -24, 1e-30, 63.9, 64.0,
                     64.2, 64.4, -65.0, 1e6, -1e6, 0.3, -0.77], dtype=np.float32)
    rng = np.random.default_rng(5)
    v = np.concatenate([edge] + [rng.normal(scale=s, size=4000).astype(np.float32)
                                 for s in (1e-3, 1e-2, 0.1, 1.0, 30.0, 300.0)])
    g = mlp_l1_grid(v)
    np.testing.assert_array_equal(g.view(np.uint32), orc.mlp_l1_grid(v).view(np.uint32))
    np.testing.assert_array_equal(g[np.abs(g) > 0], bf16_flush(v)[np.abs(g) > 0].clip(-64, 64))
    assert np.all((np.abs(g) >= 2.0 ** -9) | (g == 0)) and np.abs(g).max() == 64.0
    q = g.astype(np.float64) * 65536.0
    assert np.all(q == np.round(q)) and np.abs(q).max() <= 2.0 ** 22
    assert g[1] == 2.0 ** -9 and g[2] == -(2.0 ** -9) and g[5] == 0.0 and g[10] == 64.0
    # rounding to BF16 happens before the flush: 2^-9 (1 - 2^-9) rounds up to 2^-9
    assert g[3] == 2.0 ** -9 and g[4] == 0.0


def test_integer_route_of_the_activation_is_the_tabulated_tanh(orc):
    """The GEMM kernels build s(z) from a = min(|z| 2^16, 2^19): table entry
    j = a >> 8 of (T_j, D_j / 256) (entry 2048: the saturated value, slope 0)
    and s = fma(a & 255, D_j / 256, T_j) with the sign of z.  Every a up to
    and past saturation, both signs, against the oracle's FP32 route
    (t = min(256 |z|, 2048 - 2^-13), j = floor(t), fma(fract t, D_j, T_j))."""
    T, D = orc.mlp_table()
    s_sat = np.float32(np.longdouble(np.float32(2047.9998779296875) - np.float32(2047.0))
                       * np.longdouble(D[2047]) + np.longdouble(T[2047]))
    Ti = np.append(T, s_sat).astype(np.float32)
    Di = np.append(D * np.float32(1.0 / 256.0), np.float32(0.0)).astype(np.float32)
    a = np.arange(0, 2 ** 19 + 4096, dtype=np.int64)
    ac = np.minimum(a, 2 ** 19)
    j = ac >> 8
    f = (ac & 255).astype(np.longdouble)
    # |T| < 1 and f D_j / 256 a multiple of 2^-32 T_j: the longdouble sum is exact,
    # one rounding to FP32 as the fma
    v = (f * Di[j].astype(np.longdouble) + Ti[j].astype(np.longdouble)).astype(np.float32)
    z = (a.astype(np.float64) / 65536.0).astype(np.float32)
    assert np.all(z.astype(np.float64) * 65536.0 == a)
    ref = orc.mlp_tanh(z)
    np.testing.assert_array_equal(v.view(np.uint32), ref.view(np.uint32))
    np.testing.assert_array_equal((-v).view(np.uint32), orc.mlp_tanh(-z).view(np.uint32))
    assert s_sat == orc.mlp_tanh(np.float32(8.0)) == orc.mlp_tanh(np.float32(100.0))


def _bf16(a):
    f = np.ascontiguousarray(a, dtype=np.float32)
    b = f.view(np.uint32).astype(np.uint64)
    return ((((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16) & 0xFFFFFFFF).astype(
        np.uint32).view(np.float32)


def _tanh_tab(z, T, D):
    z = np.asarray(z, dtype=np.float32)
    a = np.abs(z)
    t = np.minimum(a * np.float32(256.0), np.float32(2047.9998779296875))
    j = t.astype(np.int64)
    f = (t - np.floor(t)).astype(np.float32)
    h = (f.astype(np.float64) * D[j].astype(np.float64) + T[j].astype(np.float64)).astype(
        np.float32)
    return np.copysign(h, z).astype(np.float32)


def numpy_log_psi(g, configs):
    """The MLP guide's definition in matrix form: exact FP64 layer 1, BF16
    activations, layer 2 in FP64 rounded to FP32, the output in the order
    (c mod 16 over c / 16, then a butterfly over the 16 columns)."""
    T, D = mlp_tanh_table()
    x = np.atleast_2d(np.asarray(configs, dtype=np.float64))
    W1 = mlp_l1_grid(g.W1).astype(np.float64)
    b1 = mlp_l1_grid(g.b1).astype(np.float64)
    W2 = bf16_flush(g.W2).astype(np.float64)
    z1 = x @ W1.T + b1
    h = _bf16(_tanh_tab(z1.astype(np.float32), T, D))
    z2 = (h.astype(np.float64) @ W2.T).astype(np.float32)
    v = (g.w3 * _tanh_tab(z2 + g.b2, T, D)).astype(np.float32)
    NC = g.hidden[1] // 16
    s = np.zeros((x.shape[0], 16), dtype=np.float32)
    for t in range(NC):
        s = (s + v[:, t * 16:(t + 1) * 16]).astype(np.float32)
    for d in (8, 4, 2, 1):
        s = (s + s[:, np.arange(16) ^ d]).astype(np.float32)
    return (s[:, 0] + np.float32(g.b3)).astype(np.float32)


NETS = [((64, 32), 12, 0), ((64, 64), 12, 1), ((128, 256), 24, 2), ((192, 128), 8, 3)]


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


def _torch_tab(z, T, D):
    """sign(z) (T_j + f D_j) written as sign(z) T_j + D_j (256 z - j sign(z))
    below saturation, so that autograd gives the slope 256 D_j at z = 0 as
    well (the derivative of |z| there would be taken as 0)."""
    import torch
    tmax = 2047.9998779296875
    t = torch.clamp(torch.abs(z) * 256.0, max=tmax).detach()
    j = torch.floor(t).long()
    sg = torch.sign(z).detach()
    free = (torch.abs(z) * 256.0 < tmax).detach()
    lin = sg * T[j] + D[j] * (256.0 * z - j.double() * sg)
    sat = sg * (T[j] + (tmax - j.double()) * D[j])
    return torch.where(free, lin, sat)


def _torch_bf16(v):
    import torch
    b = torch.as_tensor(_bf16(v.detach().numpy().astype(np.float32)), dtype=torch.float64)
    return v + (b - v).detach()


@pytest.mark.parametrize("name", ["kagome2x2", "chain8_hx"])
def test_variance_gradient_equals_torch_autograd(orc, name):
    """The oracle's gradient of the local-energy variance against FP64
    autograd of an independent torch evaluation: straight-through BF16
    rounding, the table interpolation differentiated piecewise."""
    import torch
    if name == "kagome2x2":
        L = kagome(2, 2)
        H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float32)
    else:
        H = SpinHalf.create(chain(8).edges, -1.0, -0.4, 0.7, h_x=0.6, dtype=np.float32)
    g = MLPSpinHalf(H.N, (64, 32), seed=7, scale=1.5, out_scale=1.0)
    N, (h1, h2) = H.N, g.hidden
    rng = np.random.default_rng(1)
    cfg = rng.choice([-1.0, 1.0], size=(30, N))
    P = orc.Problem(lower(H, g))
    grad, mean, var = P.mlp_vmin_grad(orc.pack(cfg))
    T, D = (torch.as_tensor(a, dtype=torch.float64) for a in mlp_tanh_table())
    flat = np.concatenate([g.W1.ravel(), g.b1, g.W2.ravel(), g.b2, g.w3, [g.b3]])
    p = torch.tensor(flat, dtype=torch.float64, requires_grad=True)
    o = np.cumsum([0, h1 * N, h1, h2 * h1, h2, h2, 1])
    W1, b1 = p[o[0]:o[1]].reshape(h1, N), p[o[1]:o[2]]
    W2, b2 = p[o[2]:o[3]].reshape(h2, h1), p[o[3]:o[4]]
    w3, b3 = p[o[4]:o[5]], p[o[5]]

    def log_psi(x):
        z1 = x @ W1.T + b1
        h = _torch_bf16(_torch_tab(z1, T, D))
        u = h @ W2.T + b2
        return _torch_tab(u, T, D) @ w3 + b3

    J_pm = float(np.asarray(H.J_pm).ravel()[0])
    J_pmpm = float(np.asarray(H.J_pmpm).ravel()[0])
    hx = float(np.asarray(H.h_x_half).ravel()[0]) if H.n_hx else 0.0
    Es = []
    for x in cfg:
        xs, bare = [], []
        for s in range(H.n_off):
            y = x.copy()
            if s < H.n_hx:
                y[s] = -y[s]
                b = hx
            else:
                i, j = H.edge_i[s - H.n_hx], H.edge_j[s - H.n_hx]
                b = J_pm if x[i] != x[j] else J_pmpm
                y[i], y[j] = -y[i], -y[j]
            if b != 0.0:
                xs.append(y)
                bare.append(b)
        jz = float(np.asarray(H.J_z).ravel()[0])
        diag = sum(jz / 4 * (1 if x[i] == x[j] else -1) for i, j in zip(H.edge_i, H.edge_j))
        lp0 = log_psi(torch.as_tensor(x[None], dtype=torch.float64))[0]
        lps = log_psi(torch.as_tensor(np.array(xs), dtype=torch.float64))
        Es.append(diag + torch.sum(torch.as_tensor(bare, dtype=torch.float64) *
                                   torch.exp(lps - lp0)))
    E = torch.stack(Es)
    Lvar = torch.mean((E - E.mean()) ** 2)
    Lvar.backward()
    want = p.grad.numpy()
    assert abs(var - float(Lvar.detach())) < 1e-5 * (1 + float(Lvar.detach()))
    assert np.max(np.abs(grad - want)) < 2e-5 * np.max(np.abs(want))
    assert np.max(np.abs(want)) > 1e-3
