"""Lowering of a (SpinHalf, guide) pair to the flat descriptor of the C-ABI
and the device-side handle the engine drives.

The lowering is the job of the reference's ``guide_cache`` /
``JastrowSpinHalf.build_cache`` (evolution.py:193-202,
jastrow_spin_half.py:400-459) plus the host-built dependency lists the
propagation kernel uses to re-evaluate only the slots a move touches.
"""
from __future__ import annotations

import ctypes
import hashlib
from collections import OrderedDict

import numpy as np
import torch

from . import _lib
from .guides import JastrowSpinHalf, MLPSpinHalf, RBMSpinHalf
from .hamiltonian import SpinHalf

__all__ = ["lower", "Problem", "get_problem"]

# A move touching more than this many slots on average switches the kernel to
# a full re-evaluation after every move.
_DEP_LIMIT = 256


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def dependency_lists(H, guide, cache=None):
    """For every off-diagonal slot (move) s: the unique slots whose rate reads
    a spin that s flips, and the unique edges whose parallel/anti-parallel
    state s can change.  Returns (dep_off, dep_slots, edge_off, dep_edges)."""
    N, n_hx, n_jpm = H.N, H.n_hx, H.n_jpm
    ei, ej = np.asarray(H.edge_i), np.asarray(H.edge_j)
    jas = isinstance(guide, JastrowSpinHalf)
    readers = [[] for _ in range(N)]            # site -> slots reading it
    if n_hx:
        site_J = guide.site_J if jas else None
        for k in range(N):
            readers[k].append(k)
            if jas:
                for q in range(guide.max_deg):
                    if site_J[k, q] != 0.0:
                        readers[int(guide.site_other[k, q])].append(k)
    if n_jpm:
        for e in range(len(ei)):
            s = n_hx + e
            readers[int(ei[e])].append(s)
            readers[int(ej[e])].append(s)
            if jas:
                for side in range(2):
                    for q in range(guide.max_deg):
                        if cache.edge_J[e, side, q] != 0.0:
                            readers[int(cache.edge_other[e, side, q])].append(s)
    incident = [[] for _ in range(N)]
    for e in range(len(ei)):
        incident[int(ei[e])].append(e)
        incident[int(ej[e])].append(e)
    readers = [np.unique(np.asarray(r, dtype=np.int64)) for r in readers]
    incident = [np.unique(np.asarray(r, dtype=np.int64)) for r in incident]
    n_off = n_hx + n_jpm
    dep_off = np.zeros(n_off + 1, dtype=np.int64)
    edge_off = np.zeros(n_off + 1, dtype=np.int64)
    slots, edges = [], []
    for s in range(n_off):
        if s < n_hx:
            sites = (s,)
        else:
            e = s - n_hx
            sites = (int(ei[e]), int(ej[e]))
        d = np.unique(np.concatenate([readers[v] for v in sites]))
        g = np.unique(np.concatenate([incident[v] for v in sites]))
        slots.append(d)
        edges.append(g)
        dep_off[s + 1] = dep_off[s] + len(d)
        edge_off[s + 1] = edge_off[s] + len(g)
    cat = lambda L: np.concatenate(L) if L else np.zeros(0, dtype=np.int64)  # noqa: E731
    return _i32(dep_off), _i32(cat(slots)), _i32(edge_off), _i32(cat(edges))


class Lowered:
    """Host arrays of a descriptor (kept alive while the descriptor is used)."""

    def __init__(self, H, guide, full=None):
        if not isinstance(H, SpinHalf):
            raise TypeError(f"pygfmc_amd propagates SpinHalf Hamiltonians; got "
                            f"{type(H).__name__}.")
        if guide is not None and not isinstance(guide, (JastrowSpinHalf, RBMSpinHalf,
                                                        MLPSpinHalf)):
            raise TypeError(f"guide must be None, JastrowSpinHalf, RBMSpinHalf or MLPSpinHalf, "
                            f"got {type(guide).__name__}")
        if guide is not None and guide.N != H.N:
            raise ValueError(f"guide has N={guide.N} sites, the Hamiltonian {H.N}")
        self.H, self.guide = H, guide
        E = len(H.edge_i)
        a = {}
        a["edge_i"] = _i32(H.edge_i)
        a["edge_j"] = _i32(H.edge_j)
        a["jz"] = _f64(H.per_edge(H.J_z))
        a["jpm"] = _f64(H.per_edge(H.J_pm))
        a["jpmpm"] = _f64(H.per_edge(H.J_pmpm))
        a["hz"] = _f64(H.per_site(H.h_z)) if H.has_hz else None
        a["hx_half"] = _f64(H.per_site(H.h_x_half)) if H.n_hx else None
        cache = None
        rbm = isinstance(guide, RBMSpinHalf)
        mlp = isinstance(guide, MLPSpinHalf)
        if mlp:
            a["mlp_W1"] = np.ascontiguousarray(guide.W1.reshape(-1), dtype=np.float32)
            a["mlp_b1"] = np.ascontiguousarray(guide.b1, dtype=np.float32)
            a["mlp_W2"] = np.ascontiguousarray(guide.W2.reshape(-1), dtype=np.float32)
            a["mlp_b2"] = np.ascontiguousarray(guide.b2, dtype=np.float32)
            a["mlp_w3"] = np.ascontiguousarray(guide.w3, dtype=np.float32)
            a["mlp_b3"] = np.asarray([guide.b3], dtype=np.float32)
        elif rbm:
            a["rbm_W"] = np.ascontiguousarray(guide.W.reshape(-1), dtype=np.float32)
            a["rbm_a"] = np.ascontiguousarray(guide.a, dtype=np.float32)
            a["rbm_b"] = np.ascontiguousarray(guide.b, dtype=np.float32)
        elif guide is not None:
            cache = guide.build_cache(H)
            a["jas_site_other"] = _i32(guide.site_other.reshape(-1))
            a["jas_site_J"] = _f64(guide.site_J.reshape(-1))
            a["jas_h"] = _f64(guide.h)
            a["jas_edge_other"] = _i32(cache.edge_other.reshape(-1))
            a["jas_edge_J"] = _f64(cache.edge_J.reshape(-1))
        n_off = H.n_off
        if rbm or mlp:
            full = True         # the network changes everywhere: every slot after every move
        else:
            dep = dependency_lists(H, guide, cache)
            if full is None:
                full = n_off > 0 and dep[0][-1] > n_off * max(_DEP_LIMIT, n_off // 4)
            if not full:
                a["dep_off"], a["dep_slots"], a["edge_off"], a["dep_edges"] = dep
        self.arrays = a
        self.scalars = dict(n_sites=H.N, n_edges=E, n_hx=H.n_hx, n_jpm=H.n_jpm,
                            has_jz=int(H.has_jz), uniform_jz=int(H.uniform_jz),
                            has_hz=int(H.has_hz),
                            guide=(_lib.GUIDE_RBM if rbm else _lib.GUIDE_MLP if mlp
                                   else _lib.GUIDE_JASTROW if guide is not None
                                   else _lib.GUIDE_NONE),
                            jas_max_deg=(guide.max_deg if isinstance(guide, JastrowSpinHalf)
                                         else 0),
                            rbm_hidden=guide.hidden if rbm else 0,
                            mlp_h1=guide.hidden[0] if mlp else 0,
                            mlp_h2=guide.hidden[1] if mlp else 0)
        self.full = bool(full)

    def desc(self):
        d = _lib.ProblemDesc()
        for k, v in self.scalars.items():
            setattr(d, k, v)
        for k, v in self.arrays.items():
            if v is None:
                continue
            ct = {np.dtype(np.int32): ctypes.c_int32, np.dtype(np.float32): ctypes.c_float,
                  np.dtype(np.float64): ctypes.c_double}[v.dtype]
            setattr(d, k, v.ctypes.data_as(ctypes.POINTER(ct)))
        return d


def lower(H, guide=None, full=None):
    """The host descriptor arrays of ``H`` with ``guide``."""
    return Lowered(H, guide, full)


class Problem:
    """Device tables of one (SpinHalf, guide, precision) and the calls on them."""

    def __init__(self, H, guide, precision, full=None, tied=True):
        lib = _lib.load()
        _lib.device()
        self.H, self.guide = H, guide
        self.precision = precision
        self.lowered = lower(H, guide, full)
        d = self.lowered.desc()
        h = ctypes.c_void_p()
        _lib.check(lib.gfmc_problem_create(ctypes.byref(d), precision, ctypes.byref(h)),
                   "gfmc_problem_create")
        self.handle = h
        # a translation-tied RBM on a torus propagates with table factors
        # (gfmc_rbm_set_tied) where the problem allows it (pair moves with
        # J_pmpm = 0, at most 512 hidden units, tables within the LDS)
        self.tied = None
        if tied and isinstance(guide, RBMSpinHalf) and H.n_hx == 0:
            from .torus import tied_tables
            t = tied_tables(H.edge_i, H.edge_j, guide)
            if t is not None:
                rc = lib.gfmc_rbm_set_tied(
                    h, t["F"], t["n_T"], t["Q"], t["n_cls"], t["pair_code"], t["n_pc"],
                    t["n_sc"], t["lane_off"].ctypes.data, t["lane_unit"].ctypes.data,
                    t["slot_code"].ctypes.data, t["site_code"].ctypes.data,
                    t["q_site"].ctypes.data, t["q_pair"].ctypes.data, _lib.stream())
                if rc == 0:
                    self.tied = t
                elif rc != _lib.GFMC_EUNSUP:
                    _lib.check(rc, "gfmc_rbm_set_tied")
        info = _lib.ProblemInfo()
        _lib.check(lib.gfmc_problem_get_info(h, ctypes.byref(info)), "gfmc_problem_get_info")
        self.info = info
        self.n_sites = info.n_sites
        self.n_words = info.n_words
        self.m_total = info.m_total
        self.guide_kind = int(info.guide)
        self._mlp_buf = None
        self._mlp_cap = 0
        self.launch = (1, 0)
        self._wpb_ok = None

    def reserve(self, n_walkers):
        """Attach (or grow) the MLP guide's propagation workspace, a device
        buffer this object owns (gfmc_mlp_work_bytes / gfmc_mlp_set_work),
        for calls on up to ``n_walkers`` walkers; a no-op for the other
        guides, whose kernels need no workspace."""
        n = int(n_walkers)
        if self.guide_kind != _lib.GUIDE_MLP or n <= self._mlp_cap:
            return
        lib = _lib.load()
        nb = int(lib.gfmc_mlp_work_bytes(self.handle, n))
        if nb < 0:
            raise _lib.GfmcError(f"MLP guide: no workspace size for {n} walkers")
        buf = torch.empty(nb + 256, dtype=torch.uint8, device=_lib.device())
        off = (-buf.data_ptr()) % 256
        _lib.check(lib.gfmc_mlp_set_work(self.handle, buf.data_ptr() + off, nb, n,
                                         _lib.stream()), "gfmc_mlp_set_work")
        self._mlp_buf = buf
        self._mlp_cap = n

    def mlp_step_stats(self):
        """(steps of the MLP guide captured into a graph that ended with a
        walker unfinished, the largest hop-iteration count of one step)
        since the workspace was attached (gfmc_mlp_unfinished; zeros for
        the other guides)."""
        if self.guide_kind != _lib.GUIDE_MLP:
            return 0, 0
        out = (ctypes.c_int64 * 2)()
        _lib.check(_lib.load().gfmc_mlp_unfinished(self.handle, out), "gfmc_mlp_unfinished")
        return int(out[0]), int(out[1])

    def set_mlp_hop_cap(self, hop_cap):
        """Hop iterations of an MLP-guided step captured into a graph."""
        _lib.check(_lib.load().gfmc_mlp_set_hop_cap(self.handle, int(hop_cap)),
                   "gfmc_mlp_set_hop_cap")

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                _lib.load().gfmc_problem_destroy(self.handle)
        except Exception:
            pass

    # -- lanes per walker of the propagation kernel ----------------------------
    @property
    def lanes_ok(self):
        """Usable lanes-per-walker values (64 always; 32 / 16 on the record
        path when every move's dependent slots fit the group)."""
        return [g for g in (64, 32, 16) if self.info.lanes_ok & g]

    @property
    def lanes(self):
        return int(self.info.lanes)

    def lowered_fingerprint(self):
        h = hashlib.sha1(self.H.fingerprint())
        h.update(b"none" if self.guide is None else self.guide.fingerprint())
        h.update(str(self.precision).encode())
        return h.digest()

    def set_lanes(self, lanes):
        _lib.check(_lib.load().gfmc_problem_set_lanes(self.handle, int(lanes)),
                   "gfmc_problem_set_lanes")
        _lib.check(_lib.load().gfmc_problem_get_info(self.handle, ctypes.byref(self.info)),
                   "gfmc_problem_get_info")

    @property
    def wpb_ok(self):
        """Usable wavefronts per workgroup of the propagation kernel: 1, 2,
        4 on the record paths, 1 elsewhere (the library's answer, probed
        once)."""
        if self._wpb_ok is None:
            lib = _lib.load()
            ok = lib.gfmc_problem_set_launch(self.handle, 2, self.launch[1]) == 0
            lib.gfmc_problem_set_launch(self.handle, self.launch[0], self.launch[1])
            self._wpb_ok = (1, 2, 4) if ok else (1,)
        return self._wpb_ok

    def set_launch(self, waves_per_block, xcd_remap):
        """Launch shape of the propagation kernel (gfmc_problem_set_launch):
        1, 2 or 4 wavefronts per workgroup, XCD-contiguous walker ranges."""
        _lib.check(_lib.load().gfmc_problem_set_launch(self.handle, int(waves_per_block),
                                                       int(bool(xcd_remap))),
                   "gfmc_problem_set_launch")
        self.launch = (int(waves_per_block), int(bool(xcd_remap)))

    # -- calls ---------------------------------------------------------------
    def rates(self, spins, want_rates=True, want_eloc=True, want_logdiff=False):
        nw = spins.shape[0]
        self.reserve(nw)
        dev = spins.device
        r = torch.empty((nw, self.m_total), dtype=torch.float64, device=dev) if want_rates else None
        e = torch.empty(nw, dtype=torch.float64, device=dev) if want_eloc else None
        ld = (torch.empty((nw, self.m_total), dtype=torch.float64, device=dev)
              if want_logdiff else None)
        _lib.check(_lib.load().gfmc_rates(self.handle, nw, _lib.ptr(spins), _lib.ptr(r),
                                          _lib.ptr(e), _lib.ptr(ld), _lib.stream()),
                   "gfmc_rates")
        return r, e, ld

    def log_psi(self, spins):
        self.reserve(spins.shape[0])
        out = torch.empty(spins.shape[0], dtype=torch.float64, device=spins.device)
        _lib.check(_lib.load().gfmc_log_psi(self.handle, spins.shape[0], _lib.ptr(spins),
                                            _lib.ptr(out), _lib.stream()), "gfmc_log_psi")
        return out

    def propagate(self, spins, log_w, e_loc, beta, alw, seed, step, tag, evals=None):
        self.reserve(spins.shape[0])
        _lib.check(_lib.load().gfmc_propagate(
            self.handle, spins.shape[0], _lib.ptr(spins), _lib.ptr(log_w), _lib.ptr(e_loc),
            float(beta), float(alw), int(seed) & 0xFFFFFFFFFFFFFFFF, int(step) & 0xFFFFFFFF,
            int(tag), _lib.ptr(evals), _lib.stream()), "gfmc_propagate")


_PROBLEMS: "OrderedDict[bytes, Problem]" = OrderedDict()
_MAX_PROBLEMS = 16


def precision_of(H, precision=None):
    if precision is not None:
        if precision in ("fp64", "float64", _lib.GFMC_FP64):
            return _lib.GFMC_FP64
        if precision in ("fp32", "float32", _lib.GFMC_FP32):
            return _lib.GFMC_FP32
        raise ValueError(f"precision must be 'fp64' or 'fp32', got {precision!r}")
    return _lib.GFMC_FP64 if np.dtype(H.dtype) == np.float64 else _lib.GFMC_FP32


def get_problem(H, guide=None, precision=None, full=None):
    """Cached :class:`Problem` for (H, guide, precision).  The RBM guide runs
    in FP32 on BF16 weights whatever the Hamiltonian's dtype."""
    prec = precision_of(H, precision)
    if precision is None and isinstance(guide, (RBMSpinHalf, MLPSpinHalf)):
        prec = _lib.GFMC_FP32
    h = hashlib.sha1(H.fingerprint())
    h.update(b"none" if guide is None else guide.fingerprint())
    h.update(str((prec, full)).encode())
    key = h.digest()
    p = _PROBLEMS.get(key)
    if p is None:
        p = Problem(H, guide, prec, full)
        _PROBLEMS[key] = p
        while len(_PROBLEMS) > _MAX_PROBLEMS:
            _PROBLEMS.popitem(last=False)
    else:
        _PROBLEMS.move_to_end(key)
    return p


# -- helpers behind SpinHalf.get_rates / Jastrow.log_psi ---------------------

def _configs_to_spins(config, n_sites):
    from .walkers import pack_configs
    c = np.asarray(config, dtype=np.float64)
    single = c.ndim == 1
    c = c.reshape(-1, n_sites)
    return pack_configs(torch.as_tensor(c, device=_lib.device())), single


def device_rates(H, guide, config):
    p = get_problem(H, guide)
    spins, single = _configs_to_spins(config, H.N)
    r, _, _ = p.rates(spins, want_eloc=False)
    r = r.cpu().numpy()
    return r[0] if single else r


def device_log_diff(H, guide, config):
    p = get_problem(H, guide)
    spins, single = _configs_to_spins(config, H.N)
    _, _, ld = p.rates(spins, want_rates=False, want_eloc=False, want_logdiff=True)
    ld = ld.cpu().numpy()
    return ld[0] if single else ld


def device_log_psi(guide, config):
    # a field-free Hamiltonian on no edges carries the guide's tables
    H = SpinHalf.create([], N=guide.N, dtype=np.float64)
    p = get_problem(H, guide, precision=(_lib.GFMC_FP32
                                         if isinstance(guide, (RBMSpinHalf, MLPSpinHalf))
                                         else _lib.GFMC_FP64))
    spins, single = _configs_to_spins(config, guide.N)
    out = p.log_psi(spins).cpu().numpy()
    return out[0] if single else out
