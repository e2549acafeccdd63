"""RBM training on the GPU: theta by MFMA and the variance-loss gradient
against the FP64 oracle, Adam against torch.optim.Adam, parameter upload,
and a training run that must beat the plain method's error bar."""
import sys
import warnings

import numpy as np
import pytest
import torch

from helpers import ground_energy, random_rbm
from pygfmc_amd import (JastrowSpinHalf, RBMSpinHalf, SpinHalf, chain, evaluate_observables,
                        prepare_walkers_sz, run, square, train_variance)
from pygfmc_amd import _lib
from pygfmc_amd.problem import Problem, lower
from pygfmc_amd.training import RBMTrainer, flat_params, rbm_from_flat

pytestmark = pytest.mark.gpu

# FP32 tolerances.  theta: sums of N exact BF16 x (+-1) products in FP32.
# Gradient: FP32 rates (1e-5 relative each) and FP32 accumulation over
# samples and moves; compared relative to the largest gradient entry.
THETA_ATOL = 2e-6
GRAD_TOL = 2e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def to_dev(spins):
    return torch.as_tensor(np.ascontiguousarray(spins).view(np.int32), device="cuda")


def train_systems():
    out = []
    H = SpinHalf.create(chain(12).edges, -1.0, -0.6, 0.7, h_x=0.4, dtype=np.float64)
    out.append(("chain12_hx", H, random_rbm(12, alpha=2, seed=1, scale=0.3), 100, False))
    H = SpinHalf.create(chain(13).edges, -1.0, -1.0, 1.0, h_x=0.3, dtype=np.float64)
    out.append(("chain13_odd_hx", H, random_rbm(13, alpha=3, seed=2, scale=0.3), 70, False))
    H = SpinHalf.create(square(4, 4).edges, -1.0, -1.0, 1.0, dtype=np.float64)
    out.append(("square4_heis", H, random_rbm(16, alpha=4, seed=3, scale=0.2), 200, True))
    H = SpinHalf.create(square(10, 10).edges, -1.0, -1.0, 1.0, dtype=np.float64)
    out.append(("square10_alpha4", H, random_rbm(100, alpha=4, seed=4, scale=0.05), 40, True))
    return out


SYSTEMS = train_systems()
IDS = [s[0] for s in SYSTEMS]


def samples(orc, H, n, sz):
    return orc.init_sz(H.N, n, H.N // 2, 11) if sz else orc.init_random(H.N, n, 11)


@pytest.mark.parametrize("name,H,G,n,sz", SYSTEMS, ids=IDS)
def test_theta_mfma_matches_oracle(orc, name, H, G, n, sz):
    say("theta mfma", name)
    prob = Problem(H, G, _lib.GFMC_FP32)
    spins = samples(orc, H, n, sz)
    ref = orc.Problem(lower(H, G)).rbm_theta(spins)
    th = torch.empty((n, G.hidden), dtype=torch.float32, device="cuda")
    _lib.check(_lib.load().gfmc_rbm_theta(prob.handle, n, _lib.ptr(to_dev(spins)),
                                          _lib.ptr(th), _lib.stream()), "gfmc_rbm_theta")
    th = th.cpu().numpy()
    err = np.abs(th - ref).max()
    say("  max |theta - oracle|", err, "max |theta|", np.abs(ref).max())
    assert err <= THETA_ATOL * max(1.0, np.abs(ref).max())


def device_grad(H, G, spins, clip=0.0):
    prob = Problem(H, G, _lib.GFMC_FP32)
    tr = RBMTrainer(prob, G, clip=clip)
    out2 = torch.empty(2, dtype=torch.float64, device="cuda")
    tr.gradient(to_dev(spins), out2)
    o = out2.cpu().numpy()
    return tr.grad.cpu().numpy(), o[0], o[1]


@pytest.mark.parametrize("name,H,G,n,sz", SYSTEMS, ids=IDS)
def test_vmin_gradient_matches_oracle(orc, name, H, G, n, sz):
    say("vmin gradient", name)
    spins = samples(orc, H, n, sz)
    g_ref, m_ref, v_ref = orc.Problem(lower(H, G)).rbm_vmin_grad(spins)
    g, m, v = device_grad(H, G, spins)
    scale = np.abs(g_ref).max()
    err = np.abs(g - g_ref).max() / scale
    nW = G.hidden * G.N
    parts = [np.abs(g[s] - g_ref[s]).max() / scale for s in
             (slice(0, nW), slice(nW, nW + G.N), slice(nW + G.N, None))]
    say(f"  grad rel err {err:.2e} (W {parts[0]:.2e} a {parts[1]:.2e} b {parts[2]:.2e}); "
        f"mean {m} vs {m_ref}; var {v} vs {v_ref}")
    assert m == pytest.approx(m_ref, rel=1e-5)
    assert v == pytest.approx(v_ref, rel=2e-5)
    assert err < GRAD_TOL


def test_vmin_gradient_with_clip_matches_oracle(orc):
    name, H, G, n, sz = SYSTEMS[2]
    spins = samples(orc, H, n, sz)
    g_ref, _, _ = orc.Problem(lower(H, G)).rbm_vmin_grad(spins, clip=1.0)
    g, _, _ = device_grad(H, G, spins, clip=1.0)
    err = np.abs(g - g_ref).max() / np.abs(g_ref).max()
    say("  clipped grad rel err", err)
    assert err < GRAD_TOL


def test_vmin_gradient_is_deterministic(orc):
    name, H, G, n, sz = SYSTEMS[3]
    spins = samples(orc, H, 300, sz)
    g1, m1, v1 = device_grad(H, G, spins)
    g2, m2, v2 = device_grad(H, G, spins)
    np.testing.assert_array_equal(g1, g2)
    assert (m1, v1) == (m2, v2)


def test_adam_matches_torch_optim():
    say("adam")
    rng = np.random.default_rng(0)
    n = 1000
    p0 = rng.normal(size=n).astype(np.float32)
    ref = torch.tensor(p0.copy(), requires_grad=True)
    opt = torch.optim.Adam([ref], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    p = torch.tensor(p0, device="cuda")
    m = torch.zeros_like(p)
    v = torch.zeros_like(p)
    for t in range(1, 6):
        g = rng.normal(size=n).astype(np.float32)
        ref.grad = torch.tensor(g)
        opt.step()
        gd = torch.tensor(g, device="cuda")
        _lib.check(_lib.load().gfmc_adam_step(n, _lib.ptr(p), _lib.ptr(gd), _lib.ptr(m),
                                              _lib.ptr(v), t, 1e-2, 0.9, 0.999, 1e-8,
                                              _lib.stream()), "gfmc_adam_step")
    np.testing.assert_allclose(p.cpu().numpy(), ref.detach().numpy(), rtol=1e-6, atol=1e-7)


def test_set_params_equals_a_fresh_problem(orc):
    """Parameters pushed from FP32 device memory give bit-identical rates to
    a problem created from the same (BF16-rounded) parameters on the host."""
    name, H, G, n, sz = SYSTEMS[0]
    rng = np.random.default_rng(3)
    flat = rng.normal(scale=0.2, size=G.hidden * G.N + G.N + G.hidden).astype(np.float32)
    prob = Problem(H, G, _lib.GFMC_FP32)
    tr = RBMTrainer(prob, G)
    tr.params.copy_(torch.as_tensor(flat, device="cuda"))
    tr.push()
    spins = to_dev(samples(orc, H, 64, sz))
    r1, e1, _ = prob.rates(spins)
    G2 = rbm_from_flat(G, flat)
    r2, e2, _ = Problem(H, G2, _lib.GFMC_FP32).rates(spins)
    np.testing.assert_array_equal(r1.cpu().numpy(), r2.cpu().numpy())
    np.testing.assert_array_equal(e1.cpu().numpy(), e2.cpu().numpy())
    np.testing.assert_array_equal(flat_params(tr.guide()), flat_params(G2))


def _sigma_run(H, G, nw, n_steps, seed):
    w = prepare_walkers_sz(nw, H.N, seed)
    _, _, _, stats, raw = run(w, n_steps, H, 0.2, G, [4, 32], rng_key=seed, warmup_steps=200,
                              verbosity=0)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        E, _, err, _ = evaluate_observables(raw, n_blocks=20)
    return E[32], err[32]


def test_training_shrinks_variance_and_beats_the_plain_error_bar():
    """4x4 Heisenberg: variance minimisation on GFMC samples from a nearly
    flat RBM cuts the local-energy variance several-fold, and GFMC with the
    trained guide has a smaller error bar than with the plain Jastrow guide
    (beta 0.3) or the untrained RBM at equal walker-steps, with both
    energies at the exact value."""
    say("training square4")
    L = square(4, 4)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float32)
    e0 = ground_energy(H)
    rbm0 = RBMSpinHalf(16, alpha=4, seed=0, scale=0.01)
    w = prepare_walkers_sz(1024, 16, 3)
    trained, hist = train_variance(H, rbm0, w, n_iter=300, steps_per_iter=2, lr=0.02,
                                   rng_key=5)
    v0, v1 = hist["variance"][:5].mean(), hist["variance"][-20:].mean()
    say(f"  variance {v0:.4f} -> {v1:.4f} in {hist['seconds']:.2f} s; E {hist['mean_e'][-1]}")
    assert v1 < 0.2 * v0
    jas = JastrowSpinHalf(16, L.edges, beta=0.3, dtype=np.float32)
    res = {}
    for name, G in (("jastrow", jas), ("rbm0", rbm0), ("trained", trained)):
        res[name] = _sigma_run(H, G, 1024, 1500, 7)
        say(f"  {name}: E = {res[name][0]:.5f} +- {res[name][1]:.5f} (exact {e0:.5f})")
    for name, (E, err) in res.items():
        assert abs(E - e0) < 4 * err + 2e-3, name
    assert res["trained"][1] < 0.5 * res["jastrow"][1]
    assert res["trained"][1] < 0.5 * res["rbm0"][1]


def test_training_a_guide_whose_ratios_overflow_is_an_error():
    """An RBM with weights of magnitude 50 on 4x4: its ratios overflow FP32,
    the sampling walkers' weights become non-finite (their steps end without
    a hop) and train_variance reports the divergence instead of returning a
    guide trained on garbage (or stepping forever)."""
    say("training a diverged guide")
    L = square(4, 4)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float32)
    rng = np.random.default_rng(2)
    rbm = RBMSpinHalf(16, alpha=1, W=50.0 * rng.choice([-1.0, 1.0], size=(16, 16)))
    w = prepare_walkers_sz(256, 16, 3)
    with pytest.raises(ValueError, match="not finite from iteration"):
        train_variance(H, rbm, w, n_iter=3, steps_per_iter=2, lr=0.02, rng_key=5,
                       warmup_steps=2)


def test_filter_support_stays_zero_through_natural_gradient_training():
    """A tied RBM whose filters exist on a 3 x 3 window (from_jastrow with
    a support): natural-gradient training moves the window's entries and
    leaves every entry outside it at exactly 0 in every hidden unit."""
    say("filter support training")
    from pygfmc_amd import RBMSpinHalf
    L = square(4, 4)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float32)
    x, y = np.arange(16) % 4, np.arange(16) // 4
    dx, dy = np.minimum(x, 4 - x), np.minimum(y, 4 - y)
    sup = np.maximum(dx, dy) <= 1
    rbm0 = RBMSpinHalf.from_jastrow(L.edges, L.extras["translations"], 0.15, 2, seed=1,
                                    scale=0.01, support=sup)
    w = prepare_walkers_sz(1024, 16, 3)
    G, hist = train_variance(H, rbm0, w, n_iter=40, steps_per_iter=2, lr=1.0, rng_key=3,
                             optimizer="sr", sr_max_norm=0.05, warmup_steps=20)
    assert np.all(np.isfinite(hist["variance"]))
    T = L.extras["translations"]
    moved = 0
    for f in range(2):
        for r in range(16):
            row0, row = rbm0.W[f * 16 + r], G.W[f * 16 + r]
            assert np.all(row[T[r][~sup]] == 0.0)
            moved += int(np.any(row[T[r][sup]] != row0[T[r][sup]]))
    assert moved == 32


def test_tied_rbm_natural_gradient_pieces_match_numpy(orc):
    """The orbit-summed log-derivatives of a translation-invariant RBM
    against the oracle's FP64 theta, the orbit sum and update, and the
    natural-gradient solve against a dense FP64 solve of the same system."""
    from pygfmc_amd import RBMSpinHalf
    from pygfmc_amd.training import RBMTrainer, flat_params
    from pygfmc_amd.problem import Problem, lower
    say("tied rbm sr pieces")
    L = square(4, 4)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float32)
    G = RBMSpinHalf.symmetric(L.extras["translations"], 3, seed=4, scale=0.3)
    p = Problem(H, G, _lib.GFMC_FP32)
    tr = RBMTrainer(p, G, optimizer="sr", sr_shift=1e-2, sr_abs=1e-4)
    off, mem, m = tr._orbits
    assert m == 3 * 16 + 3 + 1
    n = 700
    spins = orc.init_sz(16, n, 8, 3)
    sp = torch.as_tensor(spins.view(np.int32), device="cuda")
    # O against the oracle's theta
    theta = orc.Problem(lower(H, G)).rbm_theta(spins)
    x = orc.unpack(spins, 16)
    t = np.tanh(theta)
    HN = G.hidden * 16
    d_full = np.concatenate([(t[:, :, None] * x[:, None, :]).reshape(n, HN), x, t], axis=1)
    O_ref = np.zeros((n, m))
    uniq, inv = np.unique(G.orbits, return_inverse=True)
    for q in range(len(uniq)):
        O_ref[:, q] = d_full[:, inv == q].sum(axis=1)
    O = torch.empty((n, m), dtype=torch.float32, device="cuda")
    th = torch.empty((n, G.hidden), dtype=torch.float32, device="cuda")
    L_ = _lib.load()
    _lib.check(L_.gfmc_rbm_orbit_O(p.handle, n, _lib.ptr(sp), m, _lib.ptr(off), _lib.ptr(mem),
                                   _lib.ptr(O), _lib.ptr(th), _lib.stream()), "orbit_O")
    np.testing.assert_allclose(O.cpu().numpy(), O_ref, atol=2e-4)
    # orbit sum of a vector and the update
    rng = np.random.default_rng(1)
    v = rng.normal(size=tr.n_params).astype(np.float32)
    vd = torch.as_tensor(v, device="cuda")
    out = torch.empty(m, dtype=torch.float32, device="cuda")
    _lib.check(L_.gfmc_orbit_sum(tr.n_params, _lib.ptr(vd), m, _lib.ptr(off), _lib.ptr(mem),
                                 _lib.ptr(out), _lib.stream()), "orbit_sum")
    want = np.array([v[inv == q].astype(np.float64).sum() for q in range(m)])
    np.testing.assert_allclose(out.cpu().numpy(), want, rtol=1e-5, atol=1e-5)
    params = torch.as_tensor(flat_params(G), device="cuda")
    step = torch.as_tensor(rng.normal(size=m).astype(np.float32), device="cuda")
    _lib.check(L_.gfmc_orbit_axpy(tr.n_params, _lib.ptr(params), m, _lib.ptr(off), _lib.ptr(mem),
                                  _lib.ptr(step), -0.5, _lib.stream()), "orbit_axpy")
    np.testing.assert_allclose(params.cpu().numpy(),
                               flat_params(G) - 0.5 * step.cpu().numpy()[inv], rtol=1e-6,
                               atol=1e-6)
    # the SR solve
    g = torch.as_tensor(rng.normal(size=m).astype(np.float32), device="cuda")
    delta = torch.empty(m, dtype=torch.float32, device="cuda")
    work = torch.empty(int(L_.gfmc_sr_work_doubles(m)), dtype=torch.float64, device="cuda")
    _lib.check(L_.gfmc_sr_solve(n, m, _lib.ptr(O), _lib.ptr(g), 1e-2, 1e-3, 500, 1e-10, 0.0, 0,
                                _lib.ptr(delta), _lib.ptr(work), _lib.stream()), "sr_solve")
    Oh = O.cpu().numpy().astype(np.float64)
    Oc = Oh - Oh.mean(axis=0)
    S = Oc.T @ Oc / n
    A = S + np.diag(1e-2 * np.diag(S) + 1e-3)
    d_ref = np.linalg.solve(A, g.cpu().numpy().astype(np.float64))
    np.testing.assert_allclose(delta.cpu().numpy(), d_ref, rtol=1e-4,
                               atol=1e-4 * np.abs(d_ref).max())
    # the trust region scales the same direction to sqrt(d^T S d) = max_norm
    # (warm-started from the first call's solution)
    _lib.check(L_.gfmc_sr_solve(n, m, _lib.ptr(O), _lib.ptr(g), 1e-2, 1e-3, 500, 1e-10, 0.01, 1,
                                _lib.ptr(delta), _lib.ptr(work), _lib.stream()), "sr_solve")
    d2 = delta.cpu().numpy().astype(np.float64)
    assert np.sqrt(d2 @ S @ d2) == pytest.approx(min(0.01, np.sqrt(d_ref @ S @ d_ref)), rel=1e-3)
    np.testing.assert_allclose(d2 / np.linalg.norm(d2), d_ref / np.linalg.norm(d_ref), atol=1e-3)
    # a warm start from that solution for another right-hand side converges
    # to the other solution
    g2 = torch.as_tensor(rng.normal(size=m).astype(np.float32), device="cuda")
    _lib.check(L_.gfmc_sr_solve(n, m, _lib.ptr(O), _lib.ptr(g2), 1e-2, 1e-3, 500, 1e-10, 0.0, 1,
                                _lib.ptr(delta), _lib.ptr(work), _lib.stream()), "sr_solve")
    d_ref2 = np.linalg.solve(A, g2.cpu().numpy().astype(np.float64))
    np.testing.assert_allclose(delta.cpu().numpy(), d_ref2, rtol=1e-4,
                               atol=1e-4 * np.abs(d_ref2).max())
