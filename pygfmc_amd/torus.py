"""Index tables of a translation-tied RBM guide on a torus, for the
propagation kernel that reads guide factors from per-filter tables
(gfmc_rbm_set_tied).  No counterpart in the reference (its network guide is
a JAX module; this is the device layout of the RBM guide of this package).

A torus of Lx x Ly cells with B sites each numbers site k as
B (cx + Lx cy) + b; translation r = rx + Lx ry moves cell (cx, cy) to
((cx + rx) % Lx, (cy + ry) % Ly).  An RBM built by
``RBMSpinHalf.symmetric`` from these translations has
W[f n_T + r, k] = w_f(k - r), w_f = row f n_T (r = 0, the identity).

Every site gets a code on a doubled grid of row stride S = 2 Lx - 1,
code(k) = (cy S + cx) B + b, and a hidden unit (f, r) the offset
off(f, r) = f Q + O - (ry S + rx) B with O = ((Ly - 1) S + Lx - 1) B and
Q = (2 Ly - 1) S B, so that table entry q = code(k) + off(f, r) - f Q of
filter block f holds w_f at the site k - r: one integer addition per
(slot, hidden unit) replaces the periodic subtraction.  For a pair slot
(i, j) of displacement class c (the cell displacement and basis sites of
j relative to i), entry q of class c's table also names the site j - r.
"""
from __future__ import annotations

import numpy as np

__all__ = ["find_torus", "tied_tables"]


def _torus_perm(B, Lx, Ly):
    """The translation table of a (B, Lx, Ly) torus: row rx + Lx ry."""
    k = np.arange(B * Lx * Ly)
    b, cell = k % B, k // B
    cx, cy = cell % Lx, cell // Lx
    return np.array([B * (((cx + rx) % Lx) + Lx * ((cy + ry) % Ly)) + b
                     for ry in range(Ly) for rx in range(Lx)], dtype=np.int64)


def find_torus(translations):
    """(B, Lx, Ly) whose torus translations equal ``translations`` row for
    row, or None."""
    T = np.asarray(translations, dtype=np.int64)
    if T.ndim != 2:
        return None
    n_T, N = T.shape
    if n_T == 0 or N % n_T:
        return None
    B = N // n_T
    for Lx in range(1, n_T + 1):
        if n_T % Lx:
            continue
        Ly = n_T // Lx
        if np.array_equal(_torus_perm(B, Lx, Ly), T):
            return B, Lx, Ly
    return None


def tied_tables(edge_i, edge_j, guide, lanes=64):
    """The index tables of ``guide`` (RBMSpinHalf) for the pair slots
    (edge_i[s], edge_j[s]), or None when the guide is not a tied RBM on a
    torus.  Returns dict(F, n_T, Q, n_cls, unit_off [Hp], slot_code
    [n_slots] = code(i) | class << 20, site_code [N], q_site [Q], q_pair
    [n_cls Q] (-1 where the entry's basis site is not class c's first
    site)), all int32.  Checks that W is tied as the tables assume."""
    T = getattr(guide, "translations", None)
    if T is None:
        return None
    tor = find_torus(T)
    if tor is None:
        return None
    B, Lx, Ly = tor
    n_T, N = T.shape
    H = int(guide.hidden)
    if H % n_T or N != guide.N:
        return None
    F = H // n_T
    W = np.asarray(guide.W)
    # W[f n_T + r, T_r(k)] == W[f n_T, k] for all f, r, k
    for f in range(F):
        base = W[f * n_T]
        rows = W[f * n_T:(f + 1) * n_T]
        if not np.array_equal(np.take_along_axis(rows, T, axis=1), np.broadcast_to(base, T.shape)):
            return None
    S = 2 * Lx - 1
    Q = (2 * Ly - 1) * S * B
    O = ((Ly - 1) * S + Lx - 1) * B
    k = np.arange(N)
    kb, kcell = k % B, k // B
    kcx, kcy = kcell % Lx, kcell // Lx
    site_code = (kcy * S + kcx) * B + kb
    q = np.arange(Q)
    qb, qc = q % B, q // B
    qx, qy = qc % S - (Lx - 1), qc // S - (Ly - 1)
    q_site = B * ((qx % Lx) + Lx * (qy % Ly)) + qb
    Hp = (H + lanes - 1) // lanes * lanes
    unit_off = np.empty(Hp, dtype=np.int64)
    h = np.arange(H)
    f, r = h // n_T, h % n_T
    rx, ry = r % Lx, r // Lx
    unit_off[:H] = f * Q + O - (ry * S + rx) * B
    # padded units read the zero block after the weight block (filters x Q
    # entries further on from filter 0's)
    unit_off[H:] = F * Q + O
    # check: entry code(k) + off(f, r) - f Q of block f names the site k - r
    Tinv = np.empty_like(T)
    for rr in range(n_T):
        Tinv[rr, T[rr]] = np.arange(N)
    qq = site_code[None, :] + (O - (ry[:n_T] * S + rx[:n_T]) * B)[:, None]
    if qq.min() < 0 or qq.max() >= Q or not np.array_equal(q_site[qq], Tinv):
        raise AssertionError("torus tables: site codes do not invert the translations")
    # displacement classes of the slots
    ei = np.asarray(edge_i, dtype=np.int64)
    ej = np.asarray(edge_j, dtype=np.int64)
    key = np.stack([kb[ei], kb[ej], (kcx[ej] - kcx[ei]) % Lx, (kcy[ej] - kcy[ei]) % Ly], axis=1)
    cls_keys, cls_of = np.unique(key, axis=0, return_inverse=True)
    cls_of = cls_of.reshape(-1)
    n_cls = len(cls_keys)
    q_pair = np.full((n_cls, Q), -1, dtype=np.int64)
    for c, (bi, bj, dx, dy) in enumerate(cls_keys):
        sel = qb == bi
        q_pair[c, sel] = B * (((qx[sel] + dx) % Lx) + Lx * ((qy[sel] + dy) % Ly)) + bj
    # check: for slot s and translation r, the class entry names j - r
    qs = site_code[ei][None, :] + (O - (ry[:n_T] * S + rx[:n_T]) * B)[:, None]
    if not np.array_equal(q_pair[cls_of[None, :], qs], Tinv[:, ej]):
        raise AssertionError("torus tables: pair entries do not name the partner sites")
    if Q >= (1 << 20):
        return None
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)  # noqa: E731
    return dict(F=F, n_T=n_T, Q=Q, n_cls=n_cls, unit_off=i32(unit_off),
                slot_code=i32(site_code[ei] | (cls_of << 20)), site_code=i32(site_code),
                q_site=i32(q_site), q_pair=i32(q_pair.reshape(-1)), torus=(B, Lx, Ly))
