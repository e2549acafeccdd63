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


def _pairing(Lx, Ly, B, S):
    """The translation delta that pairs hidden units (f, r) and (f, r + delta)
    without a wrap (r in the lower half along a doubled direction) and its
    code (delta_y S + delta_x) B, or (None, 0) when both sides are odd."""
    if Ly % 2 == 0:
        return (0, Ly // 2), (Ly // 2) * S * B
    if Lx % 2 == 0:
        return (Lx // 2, 0), (Lx // 2) * B
    return None, 0


def tied_tables(edge_i, edge_j, guide, lanes=64):
    """The index tables of ``guide`` (RBMSpinHalf) for the pair slots
    (edge_i[s], edge_j[s]), or None when the guide is not a tied RBM on a
    torus.  All int32:

    unit_off [Hp]      f Q + O - code(r) per hidden unit (padding: F Q + O)
    slot_code [n_slots] code(first site) | class << 20 | swapped << 30: a
                       slot is read from its canonical end (the one whose
                       displacement key to the other is the smaller), so a
                       class and its reverse share one table block
    site_code [N], q_site [Q], q_pair [n_cls Q] (-1 where the entry's basis
                       site is not class c's first site)
    pair_code          code(delta) of the unit pairing (0: no pairing)
    n_pc, n_sc         pair and single lane chunks
    packed             the single chunk holds at most 16 units (four slots
                       share it)
    lane_off [(n_pc + n_sc) lanes]   the offset of the chunk's (first) unit
    lane_unit [(2 n_pc + n_sc) lanes] hidden unit of register u = 2c, 2c + 1
                       (pair chunk c) or 2 n_pc + c (single chunk c); -1 pads

    Unit b = (f, r + delta) of a pair has off(b) = off(a) - pair_code, so the
    kernel's entry (T[idx], T[idx - pair_code]) at idx = block + code(i) +
    off(a) holds both factors of the pair.  Checks that W is tied as the
    tables assume and that every read names the right weight."""
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
    pad_off = F * Q + O
    unit_off[H:] = pad_off
    # check: entry code(k) + off(f, r) - f Q of block f names the site k - r
    Tinv = np.empty_like(T)
    for rr in range(n_T):
        Tinv[rr, T[rr]] = np.arange(N)
    qq = site_code[None, :] + (O - (ry[:n_T] * S + rx[:n_T]) * B)[:, None]
    if qq.min() < 0 or qq.max() >= Q or not np.array_equal(q_site[qq], Tinv):
        raise AssertionError("torus tables: site codes do not invert the translations")
    # displacement classes of the slots, each read from its canonical end
    ei0 = np.asarray(edge_i, dtype=np.int64)
    ej0 = np.asarray(edge_j, dtype=np.int64)
    fwd = np.stack([kb[ei0], kb[ej0], (kcx[ej0] - kcx[ei0]) % Lx, (kcy[ej0] - kcy[ei0]) % Ly], 1)
    rev = np.stack([kb[ej0], kb[ei0], (kcx[ei0] - kcx[ej0]) % Lx, (kcy[ei0] - kcy[ej0]) % Ly], 1)
    swap = np.array([tuple(b) < tuple(a) for a, b in zip(fwd, rev)], dtype=bool)
    ei = np.where(swap, ej0, ei0)
    ej = np.where(swap, ei0, ej0)
    key = np.where(swap[:, None], rev, fwd)
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
    if Q >= (1 << 20) or n_cls >= (1 << 10):
        return None
    # unit pairs (f, r), (f, r + delta) and their lane chunks: full pair
    # chunks; a remainder of more than half a chunk of pairs takes one more
    # pair chunk, a smaller one goes to single chunks as separate units
    delta, pair_code = _pairing(Lx, Ly, B, S)
    if delta is not None:
        low = (ry < Ly // 2) if delta[1] else (rx < Lx // 2)
        ua = h[low]
        ub = ua + (delta[1] * Lx + delta[0])
        pairs = np.stack([ua, ub], 1)
        if not np.array_equal(np.sort(pairs.reshape(-1)), h) or \
                not np.array_equal(unit_off[ub], unit_off[ua] - pair_code):
            raise AssertionError("torus tables: unit pairs do not share a table entry")
        n_pc, rem = divmod(len(pairs), lanes)
        if rem > lanes // 2:
            n_pc, rem = n_pc + 1, 0
        singles = pairs[len(pairs) - rem:].reshape(-1) if rem else np.zeros(0, np.int64)
        pairs = pairs[:len(pairs) - rem]
    else:
        pair_code, n_pc, pairs, singles = 0, 0, np.zeros((0, 2), np.int64), h
    n_sc = (len(singles) + lanes - 1) // lanes
    lane_off = np.full((n_pc + n_sc) * lanes, pad_off, dtype=np.int64)
    lane_unit = np.full((2 * n_pc + n_sc) * lanes, -1, dtype=np.int64)
    for t, (a, b) in enumerate(pairs):
        c, l = divmod(t, lanes)
        lane_off[c * lanes + l] = unit_off[a]
        lane_unit[2 * c * lanes + l] = a
        lane_unit[(2 * c + 1) * lanes + l] = b
    for t, u in enumerate(singles):
        c, l = divmod(t, lanes)
        lane_off[(n_pc + c) * lanes + l] = unit_off[u]
        lane_unit[(2 * n_pc + c) * lanes + l] = u
    # a single chunk of at most 16 units is evaluated four slots at a time
    packed = n_sc == 1 and bool(np.all(lane_unit[(2 * n_pc) * lanes + 16:] < 0))
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)  # noqa: E731
    return dict(F=F, n_T=n_T, Q=Q, n_cls=n_cls, unit_off=i32(unit_off), packed=packed,
                slot_code=i32(site_code[ei] | (cls_of << 20) | (swap.astype(np.int64) << 30)),
                site_code=i32(site_code), q_site=i32(q_site), q_pair=i32(q_pair.reshape(-1)),
                pair_code=int(pair_code), n_pc=int(n_pc), n_sc=int(n_sc),
                lane_off=i32(lane_off), lane_unit=i32(lane_unit), torus=(B, Lx, Ly))
