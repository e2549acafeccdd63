"""The spin-1/2 XYZ Hamiltonian in fields (SpinHalf.py:70-405).

    H = sum_<ij> [J_x S^x_i S^x_j + J_y S^y_i S^y_j + J_z S^z_i S^z_j]
        - sum_i [h_z S^z_i + h_x S^x_i]

Host side only: the constructor validates and stores the couplings in the
reference's stored-coefficient convention (fields negated, J_pm = (J_x+J_y)/4,
J_pmpm = (J_x-J_y)/4); the rates are evaluated by the HIP kernels
(``get_rates`` / ``diagonal_weight`` below call the device).
"""
from __future__ import annotations

import numpy as np

__all__ = ["SpinHalf", "edges_array", "transverse_field_ising", "tfim_chain_ground_energy"]


def edges_array(edges):
    """Validated ``(n_edges, 2)`` int64 array, edge count, implied N
    (_common.edges_array, codes/pygfmc/hamiltonian/_common.py:7-31)."""
    raw = np.asarray(edges)
    if raw.shape == (0,):
        return np.zeros((0, 2), dtype=np.int64), 0, 0
    if raw.ndim != 2 or raw.shape[1] != 2:
        raise ValueError("edges must have shape (n_edges, 2).")
    if (raw.dtype.kind not in "iuf" or not np.all(np.isfinite(raw)) or np.any(raw < 0)
            or np.any(raw >= 2**63) or np.any(raw != np.floor(raw))):
        raise ValueError("edge indices must be nonnegative integers fitting in int64.")
    arr = raw.astype(np.int64)
    if np.any(arr[:, 0] == arr[:, 1]):
        raise ValueError("self-edges are not supported; edge endpoints must differ.")
    return arr, len(arr), (int(arr.max()) + 1 if arr.size else 0)


def _jpmpm_message(value):
    return (f"J_pmpm = (J_x - J_y) / 4 = {value:.6g} > 0 violates the "
            "sign-problem-free convention (J_pmpm <= 0).\n"
            "  Fix: swap S^x <-> S^y, i.e. pass J_x and J_y the other way round "
            "(a symmetry only when h_x = 0), or pass sign_check=False for a "
            "measurement operator.")


def _jpm_message(value):
    return (f"J_pm = (J_x + J_y) / 4 = {value:.6g} > 0 violates the "
            "sign-problem-free convention (J_pm <= 0): the off-diagonal elements "
            "of the propagator change sign and the walk cannot sample it.\n"
            "  Fix (Hamiltonian): on a BIPARTITE lattice apply the sublattice "
            "rotation, i.e. pass -J_x and -J_y.  For an off-diagonal MEASUREMENT "
            "operator build it with sign_check=False.")


def _is_array(v):
    return hasattr(v, "__len__")


class SpinHalf:
    """Spin-1/2 XYZ model (SpinHalf.py:70-142); build with :meth:`create`.

    Rate layout (as the reference): ``[diag, n_hx site flips (-h_x/2 each),
    n_jpm edge slots (J_pm if anti-parallel else J_pmpm)]``.
    """

    def __init__(self, *, dtype, edge_i, edge_j, N, J_z, J_pm, J_pmpm, h_z, h_x_half,
                 n_hx_static, n_jpm_static, has_jz, uniform_jz, has_hz):
        self.dtype = np.dtype(dtype)
        self.edge_i = edge_i
        self.edge_j = edge_j
        self.N = int(N)
        self.J_z = J_z
        self.J_pm = J_pm
        self.J_pmpm = J_pmpm
        self.h_z = h_z
        self.h_x_half = h_x_half
        self.n_hx_static = int(n_hx_static)
        self.n_jpm_static = int(n_jpm_static)
        self.has_jz = bool(has_jz)
        self.uniform_jz = bool(uniform_jz)
        self.has_hz = bool(has_hz)

    # -- structure -------------------------------------------------------
    @property
    def edges(self):
        return np.stack([np.asarray(self.edge_i), np.asarray(self.edge_j)], axis=1)

    @property
    def n_hx(self):
        return self.n_hx_static

    @property
    def n_jpm(self):
        return self.n_jpm_static

    @property
    def n_off(self):
        return self.n_hx_static + self.n_jpm_static

    @property
    def M_total(self):
        return self.n_hx_static + self.n_jpm_static + 1

    @property
    def local_hilbert_space(self):
        return (1.0, -1.0)

    def fingerprint(self):
        """Bytes identifying the model (used to cache device tables)."""
        parts = [str((self.N, self.n_hx_static, self.n_jpm_static, self.has_jz,
                      self.uniform_jz, self.has_hz, self.dtype.str)).encode()]
        for a in (self.edge_i, self.edge_j, self.J_z, self.J_pm, self.J_pmpm,
                  self.h_z, self.h_x_half):
            parts.append(b"-" if a is None else np.ascontiguousarray(a, dtype=np.float64).tobytes())
        return b"|".join(parts)

    # -- construction (SpinHalf.create, SpinHalf.py:144-288) -------------
    @classmethod
    def create(cls, edges, J_x=0, J_y=0, J_z=0, h_z=None, h_x=None, N=None, *,
               dtype=np.float32, sign_check=True):
        """Build a spin-1/2 XYZ Hamiltonian.

        ``dtype`` selects the rate arithmetic of the kernels: float64 runs the
        FP64 kernels, float32 the FP32 ones (weights and accumulators are
        FP64 either way).  Everything else as SpinHalf.create.
        """
        # fields enter with a minus sign; store the coefficients of +sum
        h_z = None if h_z is None else -np.asarray(h_z, dtype=float)
        h_x = None if h_x is None else -np.asarray(h_x, dtype=float)
        edges_arr, _, N_min = edges_array(edges)
        for site_arr in (h_z, h_x):
            if site_arr is not None and np.ndim(site_arr) > 0:
                N_min = max(N_min, np.shape(site_arr)[0])
        if N is None:
            N = N_min
        else:
            N = int(N)
            if N < N_min:
                raise ValueError(f"N={N} is smaller than required by the given edges/site "
                                 f"arrays (need at least {N_min})")
        n_edges = len(edges_arr)
        uniform = not (_is_array(J_x) or _is_array(J_y) or _is_array(J_z))
        if uniform:
            J_pmpm = (float(J_x) - float(J_y)) / 4.0
            if J_pmpm > 0 and sign_check:
                raise ValueError(_jpmpm_message(J_pmpm))
            J_z_arr = np.array([float(J_z)])
            J_pm_arr = np.array([(float(J_x) + float(J_y)) / 4.0])
            J_pmpm_arr = np.array([J_pmpm])
        else:
            def per_edge(v):
                a = np.asarray(v, dtype=float)
                if a.ndim == 0 or a.shape[0] == 1:
                    return np.broadcast_to(a.reshape(1), (n_edges,)).copy()
                return a
            jx, jy, J_z_arr = per_edge(J_x), per_edge(J_y), per_edge(J_z)
            if jx.shape[0] != n_edges or jy.shape[0] != n_edges or J_z_arr.shape[0] != n_edges:
                raise ValueError("Non-uniform couplings must match the number of edges.")
            J_pm_arr = (jx + jy) / 4.0
            J_pmpm_arr = (jx - jy) / 4.0
            if sign_check and np.any(J_pmpm_arr > 0):
                bad = np.flatnonzero(J_pmpm_arr > 0)
                raise ValueError(_jpmpm_message(float(J_pmpm_arr[bad[0]]))
                                 + f"\n  Offending edges: {bad.tolist()[:10]}"
                                 + (" ..." if bad.size > 10 else "")
                                 + f" ({bad.size} of {n_edges}).")
        if h_z is None:
            h_z_arr = None
        elif np.ndim(h_z) > 0:
            h_z_arr = np.asarray(h_z, dtype=float)
            if h_z_arr.shape[0] != N:
                raise ValueError(f"Non-uniform h_z must have length N={N}, got {h_z_arr.shape[0]}.")
        else:
            h_z_arr = np.array([float(h_z)])
        if h_x is None:
            h_x_half_arr = None
        elif np.ndim(h_x) > 0:
            h_x_arr = np.asarray(h_x, dtype=float)
            if h_x_arr.shape[0] != N:
                raise ValueError(f"Non-uniform h_x must have length N={N}, got {h_x_arr.shape[0]}.")
            if sign_check and np.any(h_x_arr > 0):
                raise ValueError("h_x must be >= 0 for all sites (sign-problem-free "
                                 "convention; H contains -h_x sum S^x).")
            h_x_half_arr = h_x_arr / 2
        else:
            if sign_check and float(h_x) > 0:
                raise ValueError("h_x must be >= 0 (sign-problem-free convention; H "
                                 "contains -h_x sum S^x).")
            h_x_half_arr = np.array([float(h_x) / 2])
        if sign_check and np.any(J_pm_arr > 0):
            raise ValueError(_jpm_message(float(np.max(J_pm_arr))))
        n_hx = N if (h_x_half_arr is not None and np.linalg.norm(h_x_half_arr) > 0) else 0
        n_jpm = n_edges if (np.linalg.norm(J_pm_arr) > 0 or np.linalg.norm(J_pmpm_arr) > 0) else 0
        return cls(dtype=dtype,
                   edge_i=edges_arr[:, 0].astype(np.int32),
                   edge_j=edges_arr[:, 1].astype(np.int32),
                   N=N, J_z=J_z_arr, J_pm=J_pm_arr, J_pmpm=J_pmpm_arr,
                   h_z=h_z_arr, h_x_half=h_x_half_arr,
                   n_hx_static=n_hx, n_jpm_static=n_jpm,
                   has_jz=bool(np.linalg.norm(J_z_arr) > 0),
                   uniform_jz=len(J_z_arr) == 1,
                   has_hz=h_z_arr is not None and bool(np.linalg.norm(h_z_arr) > 0))

    # -- per-edge / per-site expansions ----------------------------------
    def per_edge(self, a):
        return np.broadcast_to(np.asarray(a, dtype=float), (len(self.edge_i),)).copy()

    def per_site(self, a):
        return np.broadcast_to(np.asarray(a, dtype=float), (self.N,)).copy()

    # -- evaluation on the device ----------------------------------------
    def get_rates(self, config, return_configs=False, guide=None):
        """Rates of one ``(N,)`` or a batch ``(B, N)`` of +-1 configurations,
        evaluated by the rate kernel (SpinHalf.get_rates:316-369).  With
        ``guide`` the guide-transformed rates (transition_rates)."""
        from .problem import device_rates
        rates = device_rates(self, guide, config)
        if not return_configs:
            return rates, None
        return rates, self.successors(config)

    def diagonal_weight(self, config):
        """H_xx (SpinHalf.diagonal_weight:291-314), from the rate kernel."""
        rates, _ = self.get_rates(config)
        return rates[..., 0]

    def generate_state(self, config, idx):
        """Config after transition ``idx`` (idx 0: unchanged), SpinHalf.py:371-405."""
        c = np.array(config, dtype=float, copy=True)
        idx = int(idx)
        if idx == 0:
            return c
        off = idx - 1
        if off < self.n_hx:
            c[off] = -c[off]
        else:
            e = off - self.n_hx
            c[self.edge_i[e]] = -c[self.edge_i[e]]
            c[self.edge_j[e]] = -c[self.edge_j[e]]
        return c

    def successors(self, config):
        """(M_total, N) successor configurations of one configuration."""
        c = np.asarray(config, dtype=float)
        return np.stack([self.generate_state(c, k) for k in range(self.M_total)])


def transverse_field_ising(edges, J=1.0, h=1.0, N=None, *, dtype=np.float64):
    """H = -J sum_<ij> sigma^z_i sigma^z_j - h sum_i sigma^x_i in Pauli
    matrices, as the SpinHalf operator with J_z = -4 J and h_x = 2 h
    (sigma = 2 S); sign-free for h >= 0."""
    return SpinHalf.create(edges, J_x=0.0, J_y=0.0, J_z=-4.0 * np.asarray(J, dtype=float),
                           h_x=2.0 * np.asarray(h, dtype=float), N=N, dtype=dtype)


def tfim_chain_ground_energy(L, J=1.0, h=1.0):
    """Ground energy of the periodic transverse-field Ising chain of L sites
    (Pauli convention above) from its free-fermion solution (Pfeuty, Ann.
    Phys. 57, 79 (1970)): E_0 = -sum_k sqrt(J^2 + h^2 - 2 J h cos k) over the
    antiperiodic momenta k = (2n + 1) pi / L, n = 0..L-1."""
    k = (2.0 * np.arange(int(L)) + 1.0) * np.pi / int(L)
    return float(-np.sum(np.sqrt(J * J + h * h - 2.0 * J * h * np.cos(k))))
