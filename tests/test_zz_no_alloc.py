"""The reconfiguration entry point that allocates nothing
(gfmc_reconfigure_static, scratch in a static device array) against the one
with a caller workspace (gfmc_reconfigure): identical parents and {log G -
log n, E_n} bit for bit, on two calls in a row with the second population
smaller than the first, and the size limit rejected with an error."""
import ctypes
import sys

import numpy as np
import pytest
import torch

from pygfmc_amd import _lib

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def inputs(n, seed):
    rng = np.random.default_rng(seed)
    lw = rng.normal(scale=1.5, size=n)
    lw[rng.uniform(size=n) < 0.05] = -np.inf
    lw[0] = 0.3
    e = rng.normal(size=n) - 4.0
    return (torch.as_tensor(lw, device="cuda"), torch.as_tensor(e, device="cuda"))


def with_workspace(lw, e, seed, step, psi):
    L = _lib.load()
    n = lw.numel()
    par = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    out2 = torch.full((2,), np.nan, dtype=torch.float64, device="cuda")
    work = torch.empty(_lib.work_doubles(n), dtype=torch.float64, device="cuda")
    p = None if psi is None else ctypes.byref(ctypes.c_double(psi))
    _lib.check(L.gfmc_reconfigure(n, _lib.ptr(lw), _lib.ptr(e), seed, step, _lib.TAG_RECONFIG,
                                  p, _lib.ptr(par), _lib.ptr(out2), _lib.ptr(work),
                                  _lib.stream()), "gfmc_reconfigure")
    torch.cuda.synchronize()
    return par.cpu().numpy(), out2.cpu().numpy()


def without_allocation(lw, e, seed, step, psi):
    L = _lib.load()
    n = lw.numel()
    par = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    out2 = torch.full((2,), np.nan, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    p = None if psi is None else ctypes.byref(ctypes.c_double(psi))
    _lib.check(L.gfmc_reconfigure_static(n, _lib.ptr(lw), _lib.ptr(e), seed, step,
                                         _lib.TAG_RECONFIG, p, _lib.ptr(par), _lib.ptr(out2),
                                         _lib.stream()), "gfmc_reconfigure_static")
    torch.cuda.synchronize()
    return par.cpu().numpy(), out2.cpu().numpy()


@pytest.mark.parametrize("psi", [None, 0.8125], ids=["philox", "given_psi"])
def test_static_scratch_reconfiguration_equals_workspace_one(orc, psi):
    for k, n in enumerate((70001, 2500)):      # the second call is smaller
        say("reconfigure_static", n, psi)
        lw, e = inputs(n, 40 + k)
        seed, step = 99 + k, 1234 + k
        par_a, out_a = with_workspace(lw, e, seed, step, psi)
        par_b, out_b = without_allocation(lw, e, seed, step, psi)
        np.testing.assert_array_equal(par_b, par_a)
        np.testing.assert_array_equal(out_b, out_a)
        # and the oracle's systematic resampling of the same weights
        ps = orc.reconfig_psi(seed, step, 1) if psi is None else psi
        par_ref, lg_ref, E_ref, _ = orc.reconfigure(lw.cpu().numpy(), e.cpu().numpy(), ps)
        np.testing.assert_array_equal(par_b, par_ref)
        assert out_b[0] == pytest.approx(lg_ref, rel=1e-13, abs=1e-13)
        assert out_b[1] == pytest.approx(E_ref, rel=1e-12, abs=1e-13)


def test_static_scratch_limit_is_an_error():
    L = _lib.load()
    dummy = torch.zeros(4, dtype=torch.float64, device="cuda")
    n = _lib.RECONFIGURE_STATIC_MAX_WALKERS + 1
    rc = L.gfmc_reconfigure_static(n, _lib.ptr(dummy), _lib.ptr(dummy), 1, 0, 1, None,
                                   _lib.ptr(dummy), _lib.ptr(dummy), _lib.stream())
    assert rc == -1
    assert b"static scratch" in L.gfmc_last_error()
    # a valid call right after still works
    lw, e = inputs(1000, 7)
    par_a, out_a = with_workspace(lw, e, 5, 6, None)
    par_b, out_b = without_allocation(lw, e, 5, 6, None)
    np.testing.assert_array_equal(par_b, par_a)
    np.testing.assert_array_equal(out_b, out_a)
