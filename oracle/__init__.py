"""CPU oracle of the GFMC hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg use this package, and only as the checker / baseline; the product
(pygfmc_amd) never imports it.  ``liboracle.so`` is built from
``gfmc_oracle.c`` by ``python -m oracle.build``.

Functions take the descriptor arrays of a lowered problem (a mapping of
numpy arrays and scalars, as produced by ``pygfmc_amd.problem.lower``) and
host numpy buffers.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# PYGFMC_ORACLE_LIB selects another build of the same source (the
# sanitizer build of build.build_sanitized, run by the tests)
LIB_PATH = os.environ.get("PYGFMC_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")

_i32p = ctypes.POINTER(ctypes.c_int32)
_f64p = ctypes.POINTER(ctypes.c_double)
_f32p = ctypes.POINTER(ctypes.c_float)
_u32p = ctypes.POINTER(ctypes.c_uint32)


class Desc(ctypes.Structure):
    """Layout of gfmc_problem_desc (include/gfmc.h)."""
    _fields_ = [
        ("n_sites", ctypes.c_int32), ("n_edges", ctypes.c_int32),
        ("n_hx", ctypes.c_int32), ("n_jpm", ctypes.c_int32),
        ("has_jz", ctypes.c_int32), ("uniform_jz", ctypes.c_int32),
        ("has_hz", ctypes.c_int32), ("guide", ctypes.c_int32),
        ("edge_i", _i32p), ("edge_j", _i32p),
        ("jz", _f64p), ("jpm", _f64p), ("jpmpm", _f64p), ("hz", _f64p), ("hx_half", _f64p),
        ("jas_max_deg", ctypes.c_int32),
        ("jas_site_other", _i32p), ("jas_site_J", _f64p), ("jas_h", _f64p),
        ("jas_edge_other", _i32p), ("jas_edge_J", _f64p),
        ("rbm_hidden", ctypes.c_int32),
        ("rbm_W", _f32p), ("rbm_a", _f32p), ("rbm_b", _f32p),
        ("dep_off", _i32p), ("dep_slots", _i32p), ("edge_off", _i32p), ("dep_edges", _i32p),
        ("mlp_h1", ctypes.c_int32), ("mlp_h2", ctypes.c_int32),
        ("mlp_W1", _f32p), ("mlp_b1", _f32p), ("mlp_W2", _f32p), ("mlp_b2", _f32p),
        ("mlp_w3", _f32p), ("mlp_b3", _f32p),
    ]


class DiagDesc(ctypes.Structure):
    """Layout of gfmc_diag_desc (include/gfmc.h)."""
    _fields_ = [
        ("n_sites", ctypes.c_int32), ("n_cols", ctypes.c_int32),
        ("kind", _i32p), ("edge_off", _i32p), ("edge_i", _i32p), ("edge_j", _i32p),
        ("edge_c", _f64p), ("a", _f64p), ("b", _f64p), ("c", _f64p),
        ("n_p", ctypes.c_int32), ("m", _i32p),
    ]


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            from . import build
            build.build()
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        D = ctypes.POINTER(Desc)
        L.orc_philox.argtypes = [_u32p, _u32p, _u32p]
        L.orc_u53.restype = ctypes.c_double
        L.orc_u53.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        L.orc_wave_tree.argtypes = [_f64p]
        L.orc_set_lanes.argtypes = [ctypes.c_int]
        L.orc_blocked_scan.restype = ctypes.c_double
        L.orc_blocked_scan.argtypes = [_f64p, ctypes.c_int, _f64p]
        L.orc_sample_many.argtypes = [_f64p, ctypes.c_int, ctypes.c_int, _f64p, _i32p]
        L.orc_rates_batch.argtypes = [D, ctypes.c_int64, _u32p, _f64p, _f64p, _f64p]
        L.orc_log_psi_batch.argtypes = [D, ctypes.c_int64, _u32p, _f64p]
        L.orc_propagate.restype = ctypes.c_int
        L.orc_propagate.argtypes = [D, ctypes.c_int64, _u32p, _f64p, _f64p, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_uint64, ctypes.c_uint32,
                                    ctypes.c_uint32, _i32p, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_double]
        L.orc_propagate_ids.restype = ctypes.c_int
        L.orc_propagate_ids.argtypes = [D, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64), _u32p,
                                        _f64p, _f64p, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, _i32p,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.orc_reconfigure.restype = ctypes.c_int
        L.orc_reconfigure.argtypes = [ctypes.c_int64, _f64p, _f64p, ctypes.c_double, _i32p,
                                      _f64p, _f64p]
        L.orc_reconfig_psi.restype = ctypes.c_double
        L.orc_reconfig_psi.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32]
        L.orc_gather.argtypes = [ctypes.c_int, ctypes.c_int64, _u32p, _f64p, _i32p, _u32p, _f64p]
        L.orc_init_sz.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int, ctypes.c_uint64,
                                  _u32p]
        L.orc_init_random.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_uint64, _u32p]
        L.orc_pack.argtypes = [ctypes.c_int, ctypes.c_int64, _f64p, _u32p]
        L.orc_unpack.argtypes = [ctypes.c_int, ctypes.c_int64, _u32p, _f64p]
        L.orc_rbm_vmin_grad.argtypes = [D, ctypes.c_int64, _u32p, ctypes.c_double, _f64p,
                                        _f64p]
        L.orc_rbm_theta_batch.argtypes = [D, ctypes.c_int64, _u32p, _f64p]
        L.orc_diag_values.argtypes = [ctypes.POINTER(DiagDesc), ctypes.c_int64, _u32p, _f64p]
        L.orc_mlp_table.argtypes = [_f32p, _f32p]
        L.orc_mlp_tanh.restype = ctypes.c_float
        L.orc_mlp_tanh.argtypes = [ctypes.c_float]
        L.orc_bf16_flush.restype = ctypes.c_float
        L.orc_bf16_flush.argtypes = [ctypes.c_float]
        L.orc_mlp_l1_grid.restype = ctypes.c_float
        L.orc_mlp_l1_grid.argtypes = [ctypes.c_float]
        L.orc_mlp_hidden.argtypes = [D, ctypes.POINTER(ctypes.c_int8), _f32p]
        L.orc_mlp_O.argtypes = [D, ctypes.POINTER(ctypes.c_int8), _f64p]
        L.orc_mlp_vmin_grad.argtypes = [D, ctypes.c_int64, _u32p, ctypes.c_double, _f64p, _f64p]
        L.orc_neg_log_unit.restype = ctypes.c_double
        L.orc_neg_log_unit.argtypes = [ctypes.c_double]
        L.orc_apply_operator.argtypes = [D, ctypes.c_int64, _u32p, ctypes.c_uint64,
                                         ctypes.c_uint32, ctypes.c_uint32, _u32p, _f64p]
        del vp
        _LIB = L
    return _LIB


def _p(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


class Problem:
    """A descriptor built from lowered arrays (keeps them alive)."""

    def __init__(self, lowered):
        self.lowered = lowered
        self.arrays = {k: (None if v is None else np.ascontiguousarray(v))
                       for k, v in lowered.arrays.items()}
        d = Desc()
        for k, v in lowered.scalars.items():
            setattr(d, k, int(v))
        for k, v in self.arrays.items():
            if v is None:
                continue
            ct = {np.dtype(np.int32): ctypes.c_int32, np.dtype(np.float32): ctypes.c_float,
                  np.dtype(np.float64): ctypes.c_double}[v.dtype]
            setattr(d, k, _p(v, ct))
        self.desc = d
        self.N = int(lowered.scalars["n_sites"])
        self.n_off = int(lowered.scalars["n_hx"]) + int(lowered.scalars["n_jpm"])
        self.W = (self.N + 31) // 32

    def rates(self, spins):
        spins = np.ascontiguousarray(spins, dtype=np.uint32)
        nw = spins.shape[0]
        M = self.n_off + 1
        r = np.empty((nw, M))
        e = np.empty(nw)
        ld = np.empty((nw, M))
        lib().orc_rates_batch(ctypes.byref(self.desc), nw, _p(spins, ctypes.c_uint32),
                              _p(r, ctypes.c_double), _p(e, ctypes.c_double),
                              _p(ld, ctypes.c_double))
        return r, e, ld

    def mlp_n_params(self):
        sc = self.lowered.scalars
        h1, h2 = int(sc["mlp_h1"]), int(sc["mlp_h2"])
        return h1 * self.N + h1 + h2 * h1 + 2 * h2 + 1

    def mlp_O(self, config):
        """d log psi / dp of one +-1 configuration (flat parameter order)."""
        c = np.ascontiguousarray(np.asarray(config), dtype=np.int8)
        out = np.empty(self.mlp_n_params())
        lib().orc_mlp_O(ctypes.byref(self.desc), c.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)),
                        _p(out, ctypes.c_double))
        return out

    def mlp_vmin_grad(self, spins, clip=0.0):
        """(flat gradient, mean E, variance) of the local-energy variance."""
        s = np.ascontiguousarray(spins, dtype=np.uint32)
        g = np.empty(self.mlp_n_params())
        out2 = np.empty(2)
        lib().orc_mlp_vmin_grad(ctypes.byref(self.desc), s.shape[0], _p(s, ctypes.c_uint32),
                                float(clip), _p(g, ctypes.c_double), _p(out2, ctypes.c_double))
        return g, float(out2[0]), float(out2[1])

    def mlp_hidden(self, config):
        """Layer-1 activations (BF16 values) of +-1 configurations."""
        c = np.atleast_2d(np.asarray(config)).astype(np.int8)
        H1 = int(self.lowered.scalars["mlp_h1"])
        out = np.empty((c.shape[0], H1), dtype=np.float32)
        for i in range(c.shape[0]):
            row = np.ascontiguousarray(c[i])
            lib().orc_mlp_hidden(ctypes.byref(self.desc), row.ctypes.data_as(
                ctypes.POINTER(ctypes.c_int8)), _p(out[i], ctypes.c_float))
        return out

    def log_psi(self, spins):
        spins = np.ascontiguousarray(spins, dtype=np.uint32)
        out = np.empty(spins.shape[0])
        lib().orc_log_psi_batch(ctypes.byref(self.desc), spins.shape[0],
                                _p(spins, ctypes.c_uint32), _p(out, ctypes.c_double))
        return out

    def rbm_vmin_grad(self, spins, clip=0.0):
        """(flat gradient [W, a, b], mean E, variance) of the local-energy
        variance over the samples ``spins`` (RBM guide)."""
        spins = np.ascontiguousarray(spins, dtype=np.uint32)
        Hh = int(self.lowered.scalars["rbm_hidden"])
        g = np.empty(Hh * self.N + self.N + Hh)
        out2 = np.empty(2)
        lib().orc_rbm_vmin_grad(ctypes.byref(self.desc), spins.shape[0],
                                _p(spins, ctypes.c_uint32), float(clip),
                                _p(g, ctypes.c_double), _p(out2, ctypes.c_double))
        return g, out2[0], out2[1]

    def rbm_theta(self, spins):
        spins = np.ascontiguousarray(spins, dtype=np.uint32)
        Hh = int(self.lowered.scalars["rbm_hidden"])
        out = np.empty((spins.shape[0], Hh))
        lib().orc_rbm_theta_batch(ctypes.byref(self.desc), spins.shape[0],
                                  _p(spins, ctypes.c_uint32), _p(out, ctypes.c_double))
        return out

    def apply_operator(self, spins, seed, step, tag):
        """This problem as a measurement operator (orc_apply_operator):
        returns (successor spins, values)."""
        spins = np.ascontiguousarray(spins, dtype=np.uint32)
        out = np.empty_like(spins)
        vals = np.empty(spins.shape[0])
        lib().orc_apply_operator(ctypes.byref(self.desc), spins.shape[0],
                                 _p(spins, ctypes.c_uint32), int(seed), int(step) & 0xFFFFFFFF,
                                 int(tag), _p(out, ctypes.c_uint32), _p(vals, ctypes.c_double))
        return out, vals

    def propagate(self, spins, log_w, e_loc, beta, alw, seed, step, tag, n_threads=0,
                  lanes=64, method=0, lam=0.0):
        """In place on copies; returns (spins, log_w, e_loc, evals).  lanes:
        lanes per walker of the kernel being mirrored (blocked order);
        method 0 / 1 / 2 = cont_time / single_step / multi_step (beta is then
        the sub-step count, lam the shift)."""
        spins = np.array(spins, dtype=np.uint32, copy=True)
        log_w = np.array(log_w, dtype=np.float64, copy=True)
        e_loc = np.array(e_loc, dtype=np.float64, copy=True)
        nw = spins.shape[0]
        ev = np.zeros(nw, dtype=np.int32)
        lib().orc_set_lanes(int(lanes))
        try:
            lib().orc_propagate(ctypes.byref(self.desc), nw, _p(spins, ctypes.c_uint32),
                                _p(log_w, ctypes.c_double), _p(e_loc, ctypes.c_double),
                                float(beta), float(alw), int(seed), int(step) & 0xFFFFFFFF,
                                int(tag), _p(ev, ctypes.c_int32), int(n_threads), int(method),
                                float(lam))
        finally:
            lib().orc_set_lanes(64)
        return spins, log_w, e_loc, ev

    def propagate_ids(self, ids, spins, log_w, e_loc, beta, alw, seed, step, tag, n_threads=0,
                      lanes=64, method=0, lam=0.0):
        """:meth:`propagate` of a sample of a larger population: row i is
        walker ``ids[i]`` (its Philox counter index)."""
        ids = np.ascontiguousarray(ids, dtype=np.int64)
        spins = np.array(spins, dtype=np.uint32, copy=True)
        log_w = np.array(log_w, dtype=np.float64, copy=True)
        e_loc = np.array(e_loc, dtype=np.float64, copy=True)
        nw = spins.shape[0]
        assert ids.shape == (nw,)
        ev = np.zeros(nw, dtype=np.int32)
        lib().orc_set_lanes(int(lanes))
        try:
            lib().orc_propagate_ids(ctypes.byref(self.desc), nw, _p(ids, ctypes.c_int64),
                                    _p(spins, ctypes.c_uint32), _p(log_w, ctypes.c_double),
                                    _p(e_loc, ctypes.c_double), float(beta), float(alw),
                                    int(seed), int(step) & 0xFFFFFFFF, int(tag),
                                    _p(ev, ctypes.c_int32), int(n_threads), int(method),
                                    float(lam))
        finally:
            lib().orc_set_lanes(64)
        return spins, log_w, e_loc, ev


def philox(ctr, key):
    ctr = np.ascontiguousarray(ctr, dtype=np.uint32)
    key = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.empty(4, dtype=np.uint32)
    lib().orc_philox(_p(ctr, ctypes.c_uint32), _p(key, ctypes.c_uint32),
                     _p(out, ctypes.c_uint32))
    return out


def neg_log_unit(q):
    """-log(q) as the continuous-time hold computes it (orc_neg_log_unit)."""
    return np.array([lib().orc_neg_log_unit(float(x)) for x in np.atleast_1d(q)])


def wave_tree(v):
    """Inclusive scan of 64 values in the kernels' wave-tree order."""
    s = np.array(v, dtype=np.float64, copy=True)
    assert s.shape == (64,)
    lib().orc_wave_tree(_p(s, ctypes.c_double))
    return s


def blocked_scan(v):
    v = np.ascontiguousarray(v, dtype=np.float64)
    incl = np.empty(64)
    tot = lib().orc_blocked_scan(_p(v, ctypes.c_double), len(v), _p(incl, ctypes.c_double))
    return tot, incl


def sample_many(weights, uniforms):
    w = np.ascontiguousarray(weights, dtype=np.float64)
    u = np.ascontiguousarray(np.atleast_1d(uniforms), dtype=np.float64)
    out = np.empty(len(u), dtype=np.int32)
    lib().orc_sample_many(_p(w, ctypes.c_double), len(w), len(u), _p(u, ctypes.c_double),
                          _p(out, ctypes.c_int32))
    return out


def reconfigure(log_w, e_loc, psi):
    lw = np.ascontiguousarray(log_w, dtype=np.float64)
    el = np.ascontiguousarray(e_loc, dtype=np.float64)
    n = len(lw)
    parents = np.empty(n, dtype=np.int32)
    out2 = np.empty(2)
    cdf = np.empty(n)
    lib().orc_reconfigure(n, _p(lw, ctypes.c_double), _p(el, ctypes.c_double), float(psi),
                          _p(parents, ctypes.c_int32), _p(out2, ctypes.c_double),
                          _p(cdf, ctypes.c_double))
    return parents, out2[0], out2[1], cdf


def reconfig_psi(seed, step, tag=1):
    return lib().orc_reconfig_psi(int(seed), int(step) & 0xFFFFFFFF, int(tag))


def gather(N, spins, e_loc, parents):
    spins = np.ascontiguousarray(spins, dtype=np.uint32)
    e = np.ascontiguousarray(e_loc, dtype=np.float64)
    par = np.ascontiguousarray(parents, dtype=np.int32)
    so = np.empty_like(spins)
    eo = np.empty_like(e)
    lib().orc_gather(int(N), len(par), _p(spins, ctypes.c_uint32), _p(e, ctypes.c_double),
                     _p(par, ctypes.c_int32), _p(so, ctypes.c_uint32), _p(eo, ctypes.c_double))
    return so, eo


def init_sz(N, nw, n_up, seed):
    out = np.empty((nw, (N + 31) // 32), dtype=np.uint32)
    lib().orc_init_sz(int(N), int(nw), int(n_up), int(seed), _p(out, ctypes.c_uint32))
    return out


def init_random(N, nw, seed):
    out = np.empty((nw, (N + 31) // 32), dtype=np.uint32)
    lib().orc_init_random(int(N), int(nw), int(seed), _p(out, ctypes.c_uint32))
    return out


def pack(config):
    c = np.ascontiguousarray(config, dtype=np.float64)
    nw, N = c.shape
    out = np.empty((nw, (N + 31) // 32), dtype=np.uint32)
    lib().orc_pack(N, nw, _p(c, ctypes.c_double), _p(out, ctypes.c_uint32))
    return out


def unpack(spins, N):
    s = np.ascontiguousarray(spins, dtype=np.uint32)
    out = np.empty((s.shape[0], N))
    lib().orc_unpack(int(N), s.shape[0], _p(s, ctypes.c_uint32), _p(out, ctypes.c_double))
    return out


def run_chain(prob, spins, n_steps, beta, alw, seed, step0, warmup=False, n_threads=0,
              lanes=64, method=0, lam=0.0):
    """n_steps branching steps (propagate, reconfigure, gather) with the
    same counters as the GPU step loop; returns (spins, e_loc, series)."""
    tp, tr = (3, 4) if warmup else (0, 1)
    nw = spins.shape[0]
    lw = np.zeros(nw)
    e = np.zeros(nw)
    series = np.empty((n_steps, 2))
    for t in range(n_steps):
        spins, lw, e, _ = prob.propagate(spins, lw, e, beta, alw, seed, step0 + t, tp,
                                         n_threads=n_threads, lanes=lanes, method=method,
                                         lam=lam)
        psi = reconfig_psi(seed, step0 + t, tr)
        par, lg, En, _ = reconfigure(lw, e, psi)
        series[t] = (lg, En)
        spins, e = gather(prob.N, spins, e, par)
        lw = np.zeros(nw)
    return spins, e, series


def mlp_table():
    """(T, D) of the tabulated tanh the MLP guide uses."""
    T = np.empty(2048, dtype=np.float32)
    D = np.empty(2048, dtype=np.float32)
    lib().orc_mlp_table(_p(T, ctypes.c_float), _p(D, ctypes.c_float))
    return T, D


def mlp_tanh(z):
    return np.array([lib().orc_mlp_tanh(float(v)) for v in np.ravel(z)],
                    dtype=np.float32).reshape(np.shape(z))


def bf16_flush(v):
    return np.array([lib().orc_bf16_flush(float(a)) for a in np.ravel(v)],
                    dtype=np.float32).reshape(np.shape(v))


def mlp_l1_grid(v):
    return np.array([lib().orc_mlp_l1_grid(float(a)) for a in np.ravel(v)],
                    dtype=np.float32).reshape(np.shape(v))


class Diag:
    """Diagonal operator columns (the arrays of a gfmc_diag_desc: kind,
    edge_off, edge_i, edge_j, edge_c, a, b, c, N), as a descriptor."""

    def __init__(self, cols):
        self.cols = cols
        self.keep = [np.ascontiguousarray(getattr(cols, k), dtype=dt) for k, dt in
                     (("kind", np.int32), ("edge_off", np.int32), ("edge_i", np.int32),
                      ("edge_j", np.int32), ("edge_c", np.float64), ("a", np.float64),
                      ("b", np.float64), ("c", np.float64))]
        kind, eoff, ei, ej, ec, a, b, c = self.keep
        d = DiagDesc()
        d.n_sites = int(cols.N)
        d.n_cols = len(kind)
        d.kind, d.edge_off = _p(kind, ctypes.c_int32), _p(eoff, ctypes.c_int32)
        d.edge_i, d.edge_j = _p(ei, ctypes.c_int32), _p(ej, ctypes.c_int32)
        d.edge_c, d.a = _p(ec, ctypes.c_double), _p(a, ctypes.c_double)
        d.b, d.c = _p(b, ctypes.c_double), _p(c, ctypes.c_double)
        self.desc = d
        self.N = int(cols.N)
        self.V = len(kind)

    def values(self, spins):
        """(nw, n_cols) column values of packed walkers."""
        s = np.ascontiguousarray(spins, dtype=np.uint32)
        out = np.empty((s.shape[0], self.V))
        lib().orc_diag_values(ctypes.byref(self.desc), s.shape[0], _p(s, ctypes.c_uint32),
                              _p(out, ctypes.c_double))
        return out


def diag_chain(prob, diag, m_list, spins, n_steps, beta, alw, seed, step0, lanes=64, method=0,
               lam=0.0, history=False):
    """run_chain with the forward-walking record of driver.py:240-288 after
    every resampling: a ring of R = max(m) + 1 slots of parents (initially
    the identity) and column values (initially zero); per step the ring slot
    buf gets the parents and the values of the resampled walkers, each
    walker's ancestor m_j steps back is followed through slots buf, buf-1,
    ..., and the walker mean of the ancestors' values at slot buf - m_j is
    row[:, j].  Returns (spins, e_loc, series, rows[n_steps, n_cols, n_p])
    and, with ``history``, the per-step parents and values."""
    m_list = [int(m) for m in m_list]
    R = max(m_list) + 1
    nw = spins.shape[0]
    ring_par = np.tile(np.arange(nw, dtype=np.int64), (R, 1))
    ring_val = np.zeros((R, nw, diag.V))
    buf = 0
    lw = np.zeros(nw)
    e = np.zeros(nw)
    series = np.empty((n_steps, 2))
    rows = np.empty((n_steps, diag.V, len(m_list)))
    pars, vals = [], []
    for t in range(n_steps):
        spins, lw, e, _ = prob.propagate(spins, lw, e, beta, alw, seed, step0 + t, 0,
                                         lanes=lanes, method=method, lam=lam)
        psi = reconfig_psi(seed, step0 + t, 1)
        par, lg, En, _ = reconfigure(lw, e, psi)
        series[t] = (lg, En)
        spins, e = gather(prob.N, spins, e, par)
        lw = np.zeros(nw)
        v = diag.values(spins)
        if history:
            pars.append(par.copy())
            vals.append(v)
        ring_par[buf] = par
        ring_val[buf] = v
        anc = np.arange(nw)
        k = 0
        for j, m in enumerate(m_list):
            while k < m:
                anc = ring_par[(buf - k) % R][anc]
                k += 1
            rows[t, :, j] = ring_val[(buf - m) % R][anc].mean(axis=0)
        buf = (buf + 1) % R
    if history:
        return spins, e, series, rows, pars, vals
    return spins, e, series, rows


def forward_walk_operator(h, op, spins, m_list, beta, alw, seed, step, tag_op, values=None,
                          lanes=64, method=0, lam=0.0, history=False):
    """forward_walk_operator (evolution.py:892-1101) over the oracle's
    pieces: apply ``op`` (an oracle Problem) at position ``step`` (or start
    from ``values`` with the configurations unchanged when ``op`` is None),
    then max(m_list) branching steps of ``h`` at positions step + 1, ...
    with the production tags, the values following their walkers through
    every resampling.  Returns (fw, norm) at the lags m_list: the walker
    mean of the values after m steps and 0.0 + log G_1 + ... + log G_m; with
    ``history`` also the per-step log G and means."""
    m_list = [int(m) for m in m_list]
    if op is None:
        s = np.array(spins, dtype=np.uint32, copy=True)
        v = np.array(values, dtype=np.float64, copy=True)
    else:
        s, v = op.apply_operator(spins, seed, step, tag_op)
    nw = s.shape[0]
    means, logG = [float(np.mean(v))], [0.0]
    e = np.zeros(nw)
    for k in range(1, max(m_list) + 1):
        pos = (int(step) + k) & 0xFFFFFFFF
        s, lw, e, _ = h.propagate(s, np.zeros(nw), e, beta, alw, seed, pos, 0, lanes=lanes,
                                  method=method, lam=lam)
        par, lg, _, _ = reconfigure(lw, e, reconfig_psi(seed, pos, 1))
        s, e = gather(h.N, s, e, par)
        v = v[par]
        means.append(float(np.mean(v)))
        logG.append(float(lg))
    fw, norm = [], []
    acc, k = 0.0, 0
    for m in m_list:
        while k < m:
            k += 1
            acc = acc + logG[k]
        fw.append(means[m])
        norm.append(acc)
    if history:
        return np.array(fw), np.array(norm), np.array(logG), np.array(means)
    return np.array(fw), np.array(norm)


class lanes:
    """Context: the oracle's blocked order over ``G`` lanes per walker."""

    def __init__(self, G):
        self.G = int(G)

    def __enter__(self):
        lib().orc_set_lanes(self.G)
        return self

    def __exit__(self, *exc):
        lib().orc_set_lanes(64)
        return False
