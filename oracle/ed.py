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


# ---------------------------------------------------------------------------
# General Fock-space form (complex hopping amplitudes), used for the
# two-component model: MPS site 2p holds species 0 of lattice site p, site
# 2p+1 species 1.
# ---------------------------------------------------------------------------
def _apply_hop(s, dst, src, d):
    """b+_dst b_src on occupation tuple s: (new state, amplitude) or None."""
    if dst == src:
        return tuple(s), float(s[src])
    if s[src] == 0 or s[dst] >= d - 1:
        return None
    t = list(s)
    amp = np.sqrt(t[src]) * np.sqrt(t[dst] + 1)
    t[src] -= 1
    t[dst] += 1
    return tuple(t), amp


def fock_full(L, d, hops, diag):
    """Dense H on the full product space (basis index sum_p n_p d^(L-1-p)).
    hops: (dst, src, a) meaning a b+_dst b_src; diag(s): diagonal energy."""
    dim = d ** L
    H = np.zeros((dim, dim), dtype=complex)
    for k in range(dim):
        s = tuple(int(v) for v in np.unravel_index(k, [d] * L))
        H[k, k] += diag(s)
        for dst, src, a in hops:
            r = _apply_hop(s, dst, src, d)
            if r is not None:
                H[np.ravel_multi_index(r[0], [d] * L), k] += a * r[1]
    return H


def fock_ground_energy(L, d, hops, diag, keep):
    """Lowest eigenvalue of H restricted to the states with keep(s) True."""
    states = [s for s in itertools.product(range(d), repeat=L) if keep(s)]
    index = {s: i for i, s in enumerate(states)}
    rows, cols, vals = [], [], []
    for k, s in enumerate(states):
        rows.append(k)
        cols.append(k)
        vals.append(diag(s))
        for dst, src, a in hops:
            r = _apply_hop(s, dst, src, d)
            if r is not None:
                rows.append(index[r[0]])
                cols.append(k)
                vals.append(a * r[1])
    n = len(states)
    H = sps.csr_matrix((np.asarray(vals, dtype=complex), (rows, cols)), shape=(n, n))
    if n < 400:
        return float(np.linalg.eigvalsh(H.toarray())[0])
    return float(spla.eigsh(H, k=1, which='SA', tol=1e-13)[0][0])


def spinful_diag(m):
    """On-site energy of the two-component model m (SpinfulBoseSquare):
    mu n + U_tt/2 n(n-1) per species and U_ud n_up n_down per lattice site
    (mu taken equal for both species)."""
    mu = m.mu[0]
    uu, ud_, ud = m.inter

    def diag(s):
        e = 0.0
        for p in range(len(s) // 2):
            nu, nd = s[2 * p], s[2 * p + 1]
            e += mu * nu + 0.5 * uu * nu * (nu - 1) + mu * nd + 0.5 * ud_ * nd * (nd - 1) + ud * nu * nd
        return e
    return diag


def spinful_ground_energy(m):
    keep = lambda s: sum(s[0::2]) == m.N[0] and sum(s[1::2]) == m.N[1]
    return fock_ground_energy(m.L, m.d, m.hops(), spinful_diag(m), keep)
