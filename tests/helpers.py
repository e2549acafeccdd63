"""Independent constructions the tests check against.

* ``spinhalf_matrix``: the SpinHalf Hamiltonian from Kronecker products of
  single-site operators, sharing no code with the rate kernels (the same
  construction as the reference's tests/core/reference.py:19-54).
* ``jastrow_log_psi``: log psi_T by the explicit double sum over pairs.
* ``basis``: every +-1 configuration of N sites, site 0 the most significant
  digit, digit (1 - sigma) / 2.
"""
import itertools

import numpy as np
from scipy import sparse

from pygfmc_amd import JastrowSpinHalf, SpinHalf, chain, kagome, square


def basis(N):
    digits = np.array(list(itertools.product([0, 1], repeat=N)), dtype=np.int64)
    return 1.0 - 2.0 * digits


def index_of(configs):
    c = np.atleast_2d(configs)
    N = c.shape[1]
    digits = ((1 - c) / 2).astype(np.int64)
    return digits @ (2 ** np.arange(N - 1, -1, -1))


def spinhalf_matrix(H):
    sz = sparse.csr_matrix(np.array([[0.5, 0.0], [0.0, -0.5]]))
    sp = sparse.csr_matrix(np.array([[0.0, 1.0], [0.0, 0.0]]))
    sm = sparse.csr_matrix(np.array([[0.0, 0.0], [1.0, 0.0]]))
    L = H.N

    def site(op, i):
        return sparse.kron(sparse.kron(sparse.eye(2 ** i), op), sparse.eye(2 ** (L - i - 1)),
                           "csr")

    Z = [site(sz, i) for i in range(L)]
    Sp = [site(sp, i) for i in range(L)]
    Sm = [site(sm, i) for i in range(L)]
    M = sparse.csr_matrix((2 ** L, 2 ** L))
    E = len(H.edge_i)
    jpm, jpmpm, jz = H.per_edge(H.J_pm), H.per_edge(H.J_pmpm), H.per_edge(H.J_z)
    for e in range(E):
        a, b = int(H.edge_i[e]), int(H.edge_j[e])
        M = M + jpm[e] * (Sp[a] @ Sm[b] + Sm[a] @ Sp[b])
        M = M + jpmpm[e] * (Sp[a] @ Sp[b] + Sm[a] @ Sm[b])
        M = M + jz[e] * (Z[a] @ Z[b])
    if H.h_z is not None:
        for i, hz in enumerate(H.per_site(H.h_z)):
            M = M + hz * Z[i]
    if H.n_hx:
        for i, hx in enumerate(H.per_site(H.h_x_half)):
            M = M + hx * (Sp[i] + Sm[i])
    return M.tocsr()


def jastrow_log_psi(G, configs):
    c = np.atleast_2d(np.asarray(configs, dtype=float))
    beta = G.beta
    out = np.zeros(c.shape[0])
    for (a, b), bt in zip(G.pairs, beta):
        out += -bt * c[:, a] * c[:, b]
    out += c @ G.h
    return out


def ground_energy(H):
    from scipy.sparse.linalg import eigsh
    M = spinhalf_matrix(H)
    if M.shape[0] <= 512:
        return float(np.linalg.eigvalsh(M.toarray())[0])
    return float(eigsh(M, k=1, which="SA")[0][0])


# -- systems used across tests --------------------------------------------------

def golden_system(dtype=np.float64, guide=True):
    """chain(8), J_x = J_y = -1, J_z = 1, h_x = 0.5; Jastrow beta 0.3, h 0.05
    (the reference's golden SpinHalf system, test_golden_outputs.py:191-215)."""
    e = chain(8).edges
    H = SpinHalf.create(e, J_x=-1.0, J_y=-1.0, J_z=1.0, h_x=0.5, dtype=dtype)
    G = JastrowSpinHalf(8, e, beta=0.3, h=0.05, dtype=dtype) if guide else None
    return H, G


def random_model(N=8, seed=0, dtype=np.float64, lattice="chain", fields=True,
                 uniform=False, guide_pairs="edges"):
    """A sign-problem-free XYZ model with random per-edge couplings and
    fields and a Jastrow guide with random per-pair couplings."""
    rng = np.random.default_rng(seed)
    if lattice == "chain":
        lat = chain(N)
    elif lattice == "square":
        lat = square(int(np.sqrt(N)), int(np.sqrt(N)))
    else:
        lat = kagome(2, 2)
    edges = lat.edges
    E = len(edges)
    if uniform:
        jx, jy, jz = -1.0, -0.6, 0.7
    else:
        jx = -rng.uniform(0.5, 1.5, E)
        jy = jx + rng.uniform(0.0, 0.5, E)   # J_x - J_y <= 0
        jz = rng.uniform(-1.0, 2.0, E)
    kw = {}
    if fields:
        kw = dict(h_z=rng.uniform(-0.5, 0.5, lat.N), h_x=rng.uniform(0.1, 0.8, lat.N))
    H = SpinHalf.create(edges, J_x=jx, J_y=jy, J_z=jz, dtype=dtype, **kw)
    if guide_pairs == "edges":
        pairs = edges
    else:   # next-nearest neighbours on top: partners beyond the bond
        pairs = list(edges) + [(i, (i + 2) % lat.N) for i in range(lat.N)]
    G = JastrowSpinHalf(lat.N, pairs, beta=rng.uniform(0.05, 0.4, len(pairs)),
                        h=rng.uniform(-0.2, 0.2, lat.N), dtype=dtype)
    return H, G


def random_rbm(N, alpha=2, seed=0, scale=0.4):
    """An RBM guide with non-trivial random parameters (BF16 weights)."""
    from pygfmc_amd.guides import RBMSpinHalf
    rng = np.random.default_rng(seed)
    H = alpha * N
    return RBMSpinHalf(N, hidden=H, W=rng.normal(scale=scale, size=(H, N)),
                       a=rng.normal(scale=0.2, size=N), b=rng.normal(scale=0.2, size=H))


def rbm_log_psi(G, configs):
    """a.x + sum_h log cosh(b_h + W_h . x), explicitly, in float64."""
    c = np.atleast_2d(np.asarray(configs, dtype=np.float64))
    th = c @ G.W.astype(np.float64).T + G.b.astype(np.float64)
    out = c @ G.a.astype(np.float64) + np.sum(np.log(np.cosh(th)), axis=1)
    return out if np.ndim(configs) > 1 else out[0]
