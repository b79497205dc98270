"""Brute-force Fock-space exact diagonalisation of the Bose-Hubbard models,
used to pin the MPO and the DMRG ground energies (textbook construction, no
reference code involved).

H = sum_{bonds (p,q)} J (b+_p b_q + b_p b+_q) + sum_p (mu n_p + U n_p (n_p-1))
with occupations 0..d-1, fixed total particle number (the reference's
spinless convention: no 1/2 on the interaction).
"""
import itertools

import numpy as np
import scipy.sparse as sps
import scipy.sparse.linalg as spla


def basis(L, d, N):
    states = [s for s in itertools.product(range(d), repeat=L) if sum(s) == N]
    return states


def hamiltonian(L, d, N, bonds, U, mu):
    states = basis(L, d, N)
    index = {s: i for i, s in enumerate(states)}
    rows, cols, vals = [], [], []
    for a, s in enumerate(states):
        diag = sum(mu * n + U * n * (n - 1) for n in s)
        rows.append(a)
        cols.append(a)
        vals.append(diag)
        for p, q, J in bonds:
            for src, dst in ((q, p), (p, q)):
                # b+_dst b_src
                if s[src] > 0 and s[dst] < d - 1:
                    t = list(s)
                    amp = np.sqrt(t[src]) * np.sqrt(t[dst] + 1)
                    t[src] -= 1
                    t[dst] += 1
                    rows.append(index[tuple(t)])
                    cols.append(a)
                    vals.append(J * amp)
    n = len(states)
    return sps.csr_matrix((vals, (rows, cols)), shape=(n, n)), states


def ground_energy(L, d, N, bonds, U, mu=0.0):
    H, states = hamiltonian(L, d, N, bonds, U, mu)
    if H.shape[0] < 400:
        return float(np.linalg.eigvalsh(H.toarray())[0])
    w = spla.eigsh(H, k=1, which='SA', tol=1e-13)[0]
    return float(w[0])


def full_hamiltonian(L, d, bonds, U, mu=0.0):
    """Dense H on the full product space (all particle numbers), basis index
    sum_p n_p d^(L-1-p) (site 0 most significant)."""
    dim = d ** L
    H = np.zeros((dim, dim))
    for a in range(dim):
        s = np.unravel_index(a, [d] * L)
        H[a, a] += sum(mu * n + U * n * (n - 1) for n in s)
        for p, q, J in bonds:
            for src, dst in ((q, p), (p, q)):
                if s[src] > 0 and s[dst] < d - 1:
                    t = list(s)
                    amp = np.sqrt(t[src]) * np.sqrt(t[dst] + 1)
                    t[src] -= 1
                    t[dst] += 1
                    H[np.ravel_multi_index(t, [d] * L), a] += J * amp
    return H
