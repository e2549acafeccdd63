"""Guiding wavefunctions.

``JastrowSpinHalf`` (jastrow_spin_half.py:42-499): psi_T(x) =
exp(sum_pairs J_ij x_i x_j + sum_i h_i x_i) with J = -beta.  The host builds
the per-site partner tables (__init__:79-157) and the per-edge
external-partner cache (build_cache:400-459); every evaluation runs on the
GPU through the rate / log-psi kernels.
"""
from __future__ import annotations

import numpy as np

from .hamiltonian import edges_array

__all__ = ["JastrowSpinHalf", "JastrowOpCache"]


class JastrowOpCache:
    """External-partner tables of a (Jastrow, SpinHalf) pair
    (JastrowOpCache, jastrow_spin_half.py:31-39)."""

    def __init__(self, edge_other, edge_J):
        self.edge_other = edge_other    # (n_edges, 2, max_deg) int32
        self.edge_J = edge_J            # (n_edges, 2, max_deg) float64, 0 on excluded/padding


class JastrowSpinHalf:
    """Two-body Jastrow guide with sparse couplings.

    ``JastrowSpinHalf(N, pairs, beta, h=None, dtype=np.float32)``: ``beta``
    scalar (tied) or one per pair, ``h`` scalar, ``(N,)`` or None.  Pairs are
    canonicalised to (min, max); each site lists its partners in pair order
    and both endpoints read the same coupling.
    """

    def __init__(self, N, pairs, beta, h=None, dtype=np.float32):
        N = int(N)
        if pairs is None:
            raise ValueError("pairs must be provided (n_pairs, 2) int array)")
        pairs_np, _, n_min = edges_array(pairs)
        if n_min > N:
            raise ValueError(f"pair indices must lie in [0, {N}).")
        pairs_np = np.sort(pairs_np.astype(np.int32), axis=1)
        n_pairs = pairs_np.shape[0]
        beta_arr = np.asarray(beta, dtype=float)
        if beta_arr.ndim > 0 and beta_arr.shape != (n_pairs,):
            raise ValueError(f"beta must be a scalar or have shape ({n_pairs},), "
                             f"got {beta_arr.shape}")
        J_np = -np.broadcast_to(beta_arr, (n_pairs,)).astype(float)
        if h is None:
            h_np = np.zeros(N)
        else:
            h_in = np.asarray(h, dtype=float)
            if h_in.ndim > 1 or (h_in.ndim == 1 and h_in.shape != (N,)):
                raise ValueError(f"h must be a scalar or have shape ({N},), got {h_in.shape}")
            h_np = np.broadcast_to(h_in, (N,)).copy()
        site_partners = [[] for _ in range(N)]
        for p_idx, (a, b) in enumerate(pairs_np):
            site_partners[a].append((int(b), p_idx))
            site_partners[b].append((int(a), p_idx))
        max_deg = max((len(s) for s in site_partners), default=0) or 1
        site_other = -np.ones((N, max_deg), dtype=np.int32)
        site_pair = np.zeros((N, max_deg), dtype=np.int32)
        for k in range(N):
            for slot, (other, p_idx) in enumerate(site_partners[k]):
                site_other[k, slot] = other
                site_pair[k, slot] = p_idx
        self.dtype = np.dtype(dtype)
        self.N = N
        self.pairs = pairs_np
        self.site_valid_mask = site_other >= 0
        self.site_other = np.clip(site_other, 0, None).astype(np.int32)
        self._site_pair = site_pair
        self._J = J_np
        self._tied = beta_arr.ndim == 0
        self._h = None if h is None else h_np
        self._cache = {}

    # -- parameters --------------------------------------------------------
    @property
    def site_J(self):
        """J per (site, partner slot), 0 on padding (site_J, :159-163)."""
        return np.where(self.site_valid_mask, self._J[self._site_pair], 0.0)

    @property
    def h(self):
        return np.zeros(self.N) if self._h is None else self._h.copy()

    @property
    def beta(self):
        return -self._J.copy()

    @property
    def tied_beta(self):
        return self._tied

    @property
    def max_deg(self):
        return self.site_other.shape[1]

    def fingerprint(self):
        return b"jastrow|" + np.int64(self.N).tobytes() + self.pairs.tobytes() + \
            self._J.tobytes() + self.h.tobytes()

    # -- operator cache (build_cache, :400-459) -----------------------------
    def build_cache(self, op):
        if not hasattr(op, "n_hx_static") or not hasattr(op, "edge_i"):
            raise TypeError(f"{type(self).__name__} needs an operator with the SpinHalf "
                            f"transition layout; got {type(op).__name__}.")
        ei = np.asarray(op.edge_i, dtype=np.int64)
        ej = np.asarray(op.edge_j, dtype=np.int64)
        key = (ei.tobytes(), ej.tobytes(), self._J.tobytes())
        hit = self._cache.get(key)
        if hit is not None:
            return hit
        site_J = self.site_J
        i_part, j_part = self.site_other[ei], self.site_other[ej]
        i_J, j_J = site_J[ei], site_J[ej]
        i_mask = self.site_valid_mask[ei] & ~(i_part == ej[:, None])
        j_mask = self.site_valid_mask[ej] & ~(j_part == ei[:, None])
        edge_other = np.stack([i_part, j_part], axis=1).astype(np.int32)
        edge_J = np.where(np.stack([i_mask, j_mask], axis=1), np.stack([i_J, j_J], axis=1), 0.0)
        out = JastrowOpCache(edge_other, edge_J)
        self._cache[key] = out
        return out

    # -- evaluation on the device --------------------------------------------
    def log_psi(self, config, cache=None):
        """log psi_T of one ``(N,)`` or a batch ``(B, N)`` of configurations
        (log_psi, :299-324), by the log-psi kernel."""
        from .problem import device_log_psi
        return device_log_psi(self, config)

    def log_diff_from_operator(self, config, op, cache=None):
        """log psi(x'_k) - log psi(x) in the SpinHalf layout of ``op``
        (log_diff_from_operator, :369-398), by the rate kernel."""
        from .problem import device_log_diff
        return device_log_diff(op, self, config)

    def log_diff_from_configs(self, config, new_configs, cache=None):
        new = np.asarray(new_configs, dtype=float)
        return self.log_psi(new) - self.log_psi(np.asarray(config, dtype=float)[None])[0]


def bf16_round(a):
    """float32 values rounded to the nearest BF16 value (ties to even),
    returned as float32."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    bits = f.view(np.uint32).astype(np.uint64)
    r = ((bits + 0x7FFF + ((bits >> 16) & 1)) >> 16) << 16
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32).reshape(f.shape)


class RBMSpinHalf:
    """Restricted-Boltzmann-machine guide (no counterpart in the reference,
    whose network guide is ClusterMLP, wavefunction/cluster_mlp.py:133-439):

        log psi(x) = sum_k a_k x_k + sum_h lncosh(b_h + sum_k W_hk x_k)

    with ``hidden = alpha * N`` units.  The weights W are held on the BF16
    grid (the kernels read them as BF16 and compute in FP32); a and b are
    FP32.  Trained by :func:`pygfmc_amd.training.train_variance`.
    """

    def __init__(self, N, alpha=4, hidden=None, W=None, a=None, b=None, seed=0, scale=0.01,
                 dtype=np.float32, orbits=None, translations=None):
        self.N = int(N)
        H = int(hidden) if hidden is not None else int(alpha) * self.N
        rng = np.random.default_rng(seed)
        if W is None:
            W = rng.normal(scale=scale, size=(H, self.N))
        W = np.asarray(W, dtype=np.float32)
        if W.shape != (H, self.N):
            raise ValueError(f"W must have shape ({H}, {self.N}), got {W.shape}")
        self.hidden = H
        self.dtype = np.dtype(dtype)
        self.W = bf16_round(W)
        self.a = np.zeros(self.N, np.float32) if a is None else np.asarray(a, np.float32).copy()
        self.b = np.zeros(H, np.float32) if b is None else np.asarray(b, np.float32).copy()
        if self.a.shape != (self.N,) or self.b.shape != (H,):
            raise ValueError("a must have shape (N,) and b shape (hidden,)")
        # tied parameters: orbits[q] = the orbit of flat parameter q (W row-major,
        # then a, then b); parameters of one orbit are equal and train together;
        # orbit -1 holds the frozen parameters (kept at their value by training)
        self.orbits = None
        if orbits is not None:
            o = np.asarray(orbits, dtype=np.int64).reshape(-1)
            if o.shape != (H * self.N + self.N + H,) or o.min() < -1:
                raise ValueError("orbits must give every flat parameter an orbit index >= 0, "
                                 "or -1 for a frozen parameter")
            self.orbits = o
        # the site permutations of a translation-invariant RBM (hidden unit
        # f * n_T + r is filter f moved by translation r), when it is one
        self.translations = (None if translations is None
                             else np.asarray(translations, dtype=np.int64))

    @classmethod
    def symmetric(cls, translations, n_filters, filters=None, b=None, a=0.0, seed=0,
                  scale=0.01, support=None):
        """A translation-invariant RBM: hidden unit (f, r) = f * n_T + r is
        filter f moved by translation r, W[(f, r), T_r(k)] = filters[f, k],
        b[(f, r)] = b[f], a_k = a (one value).  ``translations``: (n_T, N)
        site permutations T_r (row r maps site k to T_r(k)); n_filters * n_T
        hidden units; the tied parameters are recorded in ``orbits``.
        ``support`` (boolean, N or (n_filters, N)): the filter entries that
        exist; the others are 0 and frozen (orbit -1), so every hidden unit
        sees only the sites its filter's support covers."""
        T = np.asarray(translations, dtype=np.int64)
        n_T, N = T.shape
        F = int(n_filters)
        rng = np.random.default_rng(seed)
        filt = (rng.normal(scale=scale, size=(F, N)) if filters is None
                else np.asarray(filters, dtype=np.float64).reshape(F, N))
        sup = (np.ones((F, N), dtype=bool) if support is None
               else np.broadcast_to(np.asarray(support, dtype=bool), (F, N)))
        filt = np.where(sup, filt, 0.0)
        bf = np.zeros(F) if b is None else np.asarray(b, dtype=np.float64).reshape(F)
        H = F * n_T
        W = np.zeros((H, N))
        oW = np.zeros((H, N), dtype=np.int64)
        for f in range(F):
            for r in range(n_T):
                W[f * n_T + r, T[r]] = filt[f]
                oW[f * n_T + r, T[r]] = np.where(sup[f], f * N + np.arange(N), -1)
        orbits = np.concatenate([oW.reshape(-1), np.full(N, F * N + F),
                                 F * N + np.repeat(np.arange(F), n_T)])
        return cls(N, hidden=H, W=W, a=np.full(N, a), b=np.repeat(bf, n_T), orbits=orbits,
                   translations=T)

    @classmethod
    def from_jastrow(cls, edges, translations, beta, n_filters, seed=0, scale=0.01,
                     support=None):
        """A translation-invariant RBM that starts as the nearest-neighbour
        Jastrow guide log psi = -beta sum_edges x_i x_j (+ const): one filter
        per class of translation-equivalent edges, w at site 0 and -w at the
        edge's other end, lncosh(w (x_i - x_j)) = lncosh(2 w) on anti-parallel
        bonds and 0 on parallel ones, so lncosh(2 w) / 2 = beta,
        w = acosh(e^{2 beta}) / 2 (before BF16 rounding); the remaining
        filters start small and random.  ``support``: as in
        :meth:`symmetric` (must hold site 0 and each class's other end)."""
        T = np.asarray(translations, dtype=np.int64)
        n_T, N = T.shape
        inv = np.empty_like(T)
        for r in range(n_T):
            inv[r, T[r]] = np.arange(N)
        at = {int(T[r, 0]): r for r in range(n_T)}
        classes = set()
        for i, j in edges:
            d1, d2 = int(inv[at[int(i)], int(j)]), int(inv[at[int(j)], int(i)])
            classes.add(min(d1, d2))
        classes = sorted(classes)
        if len(classes) > n_filters:
            raise ValueError(f"{len(classes)} edge classes need at least as many filters, "
                             f"got n_filters={n_filters}")
        w = float(np.arccosh(np.exp(2.0 * float(beta)))) / 2.0
        rng = np.random.default_rng(seed)
        filt = rng.normal(scale=scale, size=(int(n_filters), N))
        for f, d in enumerate(classes):
            filt[f] = 0.0
            filt[f, 0] = w
            filt[f, d] = -w
            if support is not None and not (np.broadcast_to(support, (len(filt), N))[f, [0, d]]
                                            .all()):
                raise ValueError(f"support must hold site 0 and site {d} of filter {f}")
        return cls.symmetric(T, n_filters, filters=filt, support=support)

    def parameters(self):
        return dict(W=self.W.copy(), a=self.a.copy(), b=self.b.copy())

    def with_parameters(self, W, a, b):
        return RBMSpinHalf(self.N, hidden=self.hidden, W=W, a=a, b=b, dtype=self.dtype,
                           orbits=self.orbits, translations=self.translations)

    def fingerprint(self):
        return (b"rbm|" + np.int64(self.N).tobytes() + np.int64(self.hidden).tobytes()
                + self.W.tobytes() + self.a.tobytes() + self.b.tobytes())

    def log_psi(self, config, cache=None):
        from .problem import device_log_psi
        return device_log_psi(self, config)

    def log_diff_from_operator(self, config, op, cache=None):
        from .problem import device_log_diff
        return device_log_diff(op, self, config)

    def log_diff_from_configs(self, config, new_configs, cache=None):
        new = np.asarray(new_configs, dtype=float)
        return self.log_psi(new) - self.log_psi(np.asarray(config, dtype=float)[None])[0]


def bf16_flush(a):
    """BF16 rounding (nearest even) with magnitudes below 2^-24 flushed to 0:
    the grid the MLP guide's W1, b1 and W2 live on."""
    r = bf16_round(a)
    return np.where(np.abs(r) < np.float32(2.0 ** -24), np.float32(0.0), r).astype(np.float32)


def mlp_l1_grid(a):
    """The MLP's layer-1 grid (W1, b1): BF16 rounding (nearest even) with
    magnitudes below 2^-9 flushed to 0 and above 64 clamped to +-64, so every
    value is an integer multiple of 2^-16 of magnitude at most 2^6 and the
    layer-1 sums of up to 510 sites are exact 32-bit integers at scale 2^16."""
    r = bf16_round(a)
    r = np.where(np.abs(r) < np.float32(2.0 ** -9), np.float32(0.0), r)
    return np.clip(r, np.float32(-64.0), np.float32(64.0)).astype(np.float32)


def mlp_tanh_table():
    """(T, D) of the tabulated tanh: T_j = tanh(j / 256) and D_j = T_{j+1} - T_j
    in FP32, j = 0..2047."""
    j = np.arange(2049, dtype=np.float64)
    T = np.tanh(j / 256.0).astype(np.float32)
    return T[:-1].copy(), (T[1:] - T[:-1]).astype(np.float32)


class MLPSpinHalf:
    """Two-layer perceptron guide (no counterpart in the reference, whose
    network guide is ClusterMLP, wavefunction/cluster_mlp.py:133-439):

        log psi(x) = b3 + sum_c w3_c s(b2_c + sum_k W2_ck h_k),
        h_k = bf16(s(b1_k + sum_n W1_kn x_n)),

    ``s`` the tabulated tanh (2048 linear pieces on [0, 8), odd, saturating
    at tanh(8)), ``bf16`` rounding to nearest even.  W1 and b1 are held on the
    layer-1 grid (:func:`mlp_l1_grid`: BF16, magnitudes below 2^-9 flushed to
    0, above 64 clamped), W2 on the BF16 grid (magnitudes below 2^-24
    flushed); b2, w3 and b3 are FP32.  ``hidden = (h1, h2)``: h1 a multiple of
    64, h2 one of 32, 64, 128, 256; at most 510 sites.
    """

    def __init__(self, N, hidden=(512, 256), W1=None, b1=None, W2=None, b2=None, w3=None,
                 b3=0.0, seed=0, scale=1.0, out_scale=0.1, dtype=np.float32):
        self.N = int(N)
        if not 1 <= self.N <= 510:
            raise ValueError(f"the MLP guide takes 1 to 510 sites, got N={self.N}")
        h1, h2 = (int(hidden[0]), int(hidden[1]))
        if h1 <= 0 or h1 % 64 or h2 not in (32, 64, 128, 256):
            raise ValueError(f"hidden must be (multiple of 64, one of 32/64/128/256); got "
                             f"{hidden!r}")
        self.hidden = (h1, h2)
        rng = np.random.default_rng(seed)
        if W1 is None:
            W1 = rng.normal(scale=scale / np.sqrt(self.N), size=(h1, self.N))
        if W2 is None:
            W2 = rng.normal(scale=scale / np.sqrt(h1), size=(h2, h1))
        if w3 is None:
            w3 = rng.normal(scale=out_scale / np.sqrt(h2), size=h2)
        self.W1 = mlp_l1_grid(np.asarray(W1, dtype=np.float32).reshape(h1, self.N))
        self.W2 = bf16_flush(np.asarray(W2, dtype=np.float32).reshape(h2, h1))
        self.b1 = mlp_l1_grid(np.zeros(h1) if b1 is None else np.asarray(b1, dtype=np.float32))
        self.b2 = np.zeros(h2, np.float32) if b2 is None else np.asarray(b2, np.float32).copy()
        self.w3 = np.asarray(w3, dtype=np.float32).reshape(h2).copy()
        self.b3 = np.float32(b3)
        self.dtype = np.dtype(dtype)
        if self.b1.shape != (h1,) or self.b2.shape != (h2,):
            raise ValueError("b1 must have shape (h1,) and b2 shape (h2,)")

    def parameters(self):
        return dict(W1=self.W1.copy(), b1=self.b1.copy(), W2=self.W2.copy(), b2=self.b2.copy(),
                    w3=self.w3.copy(), b3=np.float32(self.b3))

    def with_parameters(self, W1, b1, W2, b2, w3, b3):
        return MLPSpinHalf(self.N, self.hidden, W1=W1, b1=b1, W2=W2, b2=b2, w3=w3, b3=b3,
                           dtype=self.dtype)

    def fingerprint(self):
        return (b"mlp|" + np.asarray([self.N, *self.hidden], np.int64).tobytes()
                + self.W1.tobytes() + self.b1.tobytes() + self.W2.tobytes()
                + self.b2.tobytes() + self.w3.tobytes() + np.float32(self.b3).tobytes())

    def log_psi(self, config, cache=None):
        from .problem import device_log_psi
        return device_log_psi(self, config)

    def log_diff_from_operator(self, config, op, cache=None):
        from .problem import device_log_diff
        return device_log_diff(op, self, config)

    def log_diff_from_configs(self, config, new_configs, cache=None):
        new = np.asarray(new_configs, dtype=float)
        return self.log_psi(new) - self.log_psi(np.asarray(config, dtype=float)[None])[0]
