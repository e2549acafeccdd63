"""Diagonal measurement operators and their lowering to the device's value
columns (hamiltonian/observables.py, SpinHalf.diagonal_weight).

A diagonal operator needs only ``diagonal_weight(config)``.  Two families are
measured on the GPU:

* :class:`~pygfmc_amd.hamiltonian.SpinHalf` -- its diagonal part
  ``sum_e J_z/4 s_i s_j - sum_k h_z s_k / 2`` (SpinHalf.py:291-314), one
  value per walker;
* :class:`SzSzStructureFactor` -- ``(|sum_k e^{i q.r_k} s_k|^2 - N) / 4``
  per wavevector q (observables.py:68-165), one value per wavevector;
* :class:`Identity` -- a constant, one value per walker.

``lower_diag_operators`` turns a ``{name: operator}`` mapping into the
column tables of ``gfmc_diag_desc``: one ``GFMC_DIAG_PAIR`` column per
SpinHalf, one ``GFMC_DIAG_SF`` column per wavevector.
"""
from __future__ import annotations

import hashlib

import numpy as np

from .hamiltonian import SpinHalf

__all__ = ["SzSzStructureFactor", "Identity", "is_valid_k", "lower_diag_operators",
           "DiagLowering"]


def is_valid_k(k_vectors, lattice_vectors, atol=1e-5):
    """Whether each wavevector lies in the reciprocal lattice of the periodic
    cell whose vectors are the rows of ``lattice_vectors``: k . L_j / (2 pi)
    an integer (within ``atol``) for every row (observables.py:13-65).
    ``bool`` for one ``(d,)`` wavevector, a boolean ``(n_k,)`` array
    otherwise."""
    lat = np.asarray(lattice_vectors, dtype=np.float64)
    if lat.ndim != 2 or lat.shape[0] != lat.shape[1]:
        raise ValueError(f"lattice_vectors must be (d, d), rows the periodic vectors; got "
                         f"{lat.shape}.")
    k = np.asarray(k_vectors, dtype=np.float64)
    single = k.ndim == 1
    k = np.atleast_2d(k)
    if k.ndim != 2 or k.shape[1] != lat.shape[1]:
        raise ValueError(f"k_vectors must be (d,) or (n_k, d) with d = {lat.shape[1]}; got "
                         f"{np.shape(k_vectors)}.")
    m = k @ lat.T / (2 * np.pi)
    ok = np.all(np.abs(m - np.round(m)) <= atol, axis=1)
    return bool(ok[0]) if single else ok


class SzSzStructureFactor:
    """S^zz(q) = sum_{i<j} 2 cos(q.(r_i - r_j)) S^z_i S^z_j for one or many
    wavevectors q, evaluated as (|rho_q|^2 - N) / 4 with rho_q = sum_k
    e^{i q.r_k} s_k (s = +-1): one value per wavevector.  Exact for
    wavevectors of the cell's reciprocal lattice (:func:`is_valid_k`), which
    the caller supplies.

    ``dtype`` rounds the stored cos / sin tables (built in float64), as the
    reference's table dtype does (default float32); the device evaluates in
    FP64 with the rounded tables.
    """

    def __init__(self, positions, k_vectors, dtype=np.float32):
        r = np.atleast_2d(np.asarray(positions, dtype=np.float64))
        k = np.atleast_2d(np.asarray(k_vectors, dtype=np.float64))
        if r.ndim != 2 or k.ndim != 2:
            raise ValueError(f"positions must be (N, d) and k_vectors (n_k, d); got "
                             f"{r.shape} and {k.shape}.")
        if k.shape[1] != r.shape[1]:
            raise ValueError(f"k_vectors have dimension {k.shape[1]} but positions have "
                             f"{r.shape[1]}; a wavevector needs one component per spatial "
                             "dimension of the lattice.")
        kr = k @ r.T
        self.dtype = np.dtype(dtype)
        self.cos = np.cos(kr).astype(self.dtype).astype(np.float64)
        self.sin = np.sin(kr).astype(self.dtype).astype(np.float64)
        self.offset = float(np.asarray(0.25 * r.shape[0], dtype=self.dtype))
        self.k_vectors = k
        self.N = int(r.shape[0])

    @property
    def n_values(self):
        return int(self.cos.shape[0])

    def diagonal_weight(self, config):
        """``(n_k,)`` for one ``(N,)`` configuration of +-1, ``(B, n_k)`` for
        a batch (host NumPy)."""
        c = np.asarray(config, dtype=np.float64)
        return 0.25 * ((c @ self.cos.T) ** 2 + (c @ self.sin.T) ** 2) - self.offset

    def fingerprint(self):
        return b"|".join([b"sf", np.ascontiguousarray(self.cos).tobytes(),
                          np.ascontiguousarray(self.sin).tobytes(),
                          np.float64(self.offset).tobytes()])


class Identity:
    """Measurement-only constant operator on N sites: one diagonal element
    ``value`` (1 by default) and no off-diagonal slot (M_total = 1).  As an
    off-diagonal measurement its forward-walking value is the shadow chain's
    bare weight product, which replays the main chain exactly when the
    shadow consumes the main chain's own future draws, so its estimate is 1
    for every window (the identity check of test_identity_forward_walk.py)."""

    M_total = 1
    n_off = 0

    def __init__(self, N, value=1.0):
        self.N = int(N)
        self.value = float(value)

    def diagonal_weight(self, config):
        c = np.asarray(config, dtype=np.float64)
        return np.full(c.shape[:-1], self.value)

    def fingerprint(self):
        return b"identity|" + np.asarray([self.N], dtype=np.int64).tobytes() + \
            np.float64(self.value).tobytes()


class DiagLowering:
    """Column tables of a ``{name: operator}`` mapping (``gfmc_diag_desc``
    layout) and the map back from columns to named series."""

    def __init__(self, names, widths, kind, edge_off, edge_i, edge_j, edge_c, a, b, c, N):
        self.names = list(names)
        self.widths = list(widths)
        self.kind = np.asarray(kind, dtype=np.int32)
        self.edge_off = np.asarray(edge_off, dtype=np.int32)
        self.edge_i = np.asarray(edge_i, dtype=np.int32)
        self.edge_j = np.asarray(edge_j, dtype=np.int32)
        self.edge_c = np.asarray(edge_c, dtype=np.float64)
        self.a = np.ascontiguousarray(a, dtype=np.float64)
        self.b = np.ascontiguousarray(b, dtype=np.float64)
        self.c = np.asarray(c, dtype=np.float64)
        self.N = int(N)

    @property
    def n_cols(self):
        return int(len(self.kind))

    def split(self, rows):
        """``rows`` (..., n_cols, n_p) -> {name: (..., n_p)} for a one-value
        operator, {name: (..., n_vals, n_p)} otherwise (driver.py:841-860)."""
        out = {}
        v = 0
        for name, w in zip(self.names, self.widths):
            out[name] = rows[..., v, :] if w == 1 else rows[..., v:v + w, :]
            v += w
        return out

    def split_values(self, vals):
        """Per-walker values (nw, n_cols) -> {name: (nw,) or (nw, n_vals)}."""
        out = {}
        v = 0
        for name, w in zip(self.names, self.widths):
            out[name] = vals[:, v] if w == 1 else vals[:, v:v + w]
            v += w
        return out

    def sha1(self):
        h = hashlib.sha1()
        for arr in (self.kind, self.edge_off, self.edge_i, self.edge_j, self.edge_c, self.a,
                    self.b, self.c):
            h.update(np.ascontiguousarray(arr).tobytes())
        h.update(repr((self.names, self.widths)).encode())
        return h.hexdigest()


def lower_diag_operators(ops, N):
    """Lower ``{name: operator}`` (SpinHalf or SzSzStructureFactor) on N
    sites to value columns, in the mapping's order."""
    if not ops:
        raise ValueError("no diagonal operators")
    names, widths, kind = [], [], []
    edge_off, ei, ej, ec = [0], [], [], []
    a_rows, b_rows, c = [], [], []
    for name, op in ops.items():
        if not isinstance(name, str):
            raise TypeError(f"operator names must be strings; got {name!r}")
        if isinstance(op, SpinHalf):
            if op.N != N:
                raise ValueError(f"diagonal operator {name!r} has {op.N} sites, the "
                                 f"Hamiltonian {N}")
            names.append(name)
            widths.append(1)
            kind.append(0)
            if op.has_jz:
                jz = np.broadcast_to(np.asarray(op.J_z, dtype=np.float64),
                                     (len(op.edge_i),))
                ei.extend(int(x) for x in op.edge_i)
                ej.extend(int(x) for x in op.edge_j)
                ec.extend(float(x) / 4.0 for x in jz)
            edge_off.append(len(ei))
            a = np.zeros(N)
            if op.has_hz:
                a[:] = np.broadcast_to(np.asarray(op.h_z, dtype=np.float64), (N,)) / 2.0
            a_rows.append(a)
            b_rows.append(np.zeros(N))
            c.append(0.0)
        elif isinstance(op, SzSzStructureFactor):
            if op.N != N:
                raise ValueError(f"diagonal operator {name!r} has {op.N} sites, the "
                                 f"Hamiltonian {N}")
            names.append(name)
            widths.append(op.n_values)
            for q in range(op.n_values):
                kind.append(1)
                edge_off.append(len(ei))
                a_rows.append(op.cos[q])
                b_rows.append(op.sin[q])
                c.append(op.offset)
        elif isinstance(op, Identity):
            if op.N != N:
                raise ValueError(f"diagonal operator {name!r} has {op.N} sites, the "
                                 f"Hamiltonian {N}")
            names.append(name)
            widths.append(1)
            kind.append(0)
            edge_off.append(len(ei))
            a_rows.append(np.zeros(N))
            b_rows.append(np.zeros(N))
            c.append(op.value)
        else:
            raise TypeError(f"diagonal operator {name!r} of type {type(op).__name__} is not "
                            "supported on the GPU (SpinHalf, SzSzStructureFactor and "
                            "Identity are)")
    if len(kind) > 256:
        raise ValueError(f"{len(kind)} value columns; at most 256 are supported")
    return DiagLowering(names, widths, kind, edge_off, ei, ej, ec, np.stack(a_rows),
                        np.stack(b_rows), c, N)
