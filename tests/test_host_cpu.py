"""The C-ABI library loads and exports every declared symbol, and its host
restatements (rand stream, QuickSort, truncation rule, QN tables, MPO) agree
bit-exactly with the oracle; the host eigensolvers (RealSymMatrixDiag and the
lowest-eigenpair solve of the dense edge sites) agree with a 2 x 2 closed form
and with numpy's eigh.  No GPU is touched."""
import os
import re

import numpy as np
import pytest

import dmrgcc_amd as dm
from oracle.dmrg import truncate_dims as o_truncate
from oracle.model import SpinfulBoseSquare, SpinlessBoseSquare
from oracle.rng import GlibcRand
from oracle.sort import quicksort as o_quicksort

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "dmrgcc_hip.h")).read()
    return sorted(set(re.findall(r"\b(dmrgcc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = dm.lib()
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(dm.exported_symbols())


def test_error_path_reports_status():
    with pytest.raises(dm.DmrgError):
        dm.quicksort([1.0, 2.0], flag=7)


def test_rand_stream_matches_oracle():
    g = GlibcRand(1)
    ref = g.draws(5000)
    assert np.array_equal(dm.rand_draws(5000), ref)
    g.skip(12345)
    assert np.array_equal(dm.rand_draws(100, skip=17345), g.draws(100))


@pytest.mark.parametrize("flag", [0, 1])
def test_quicksort_matches_oracle(flag):
    rng = np.random.default_rng(7)
    for n in (0, 1, 2, 9, 64, 500):
        v = rng.integers(0, 8, n).astype(float) * 0.5
        idx = np.arange(n)
        a, b = list(v), list(idx)
        o_quicksort(a, b, flag=flag)
        pv, pi = dm.quicksort(v, idx, flag)
        assert np.array_equal(pv, a)
        assert np.array_equal(pi, b)


def test_truncate_rule_matches_oracle():
    rng = np.random.default_rng(11)
    for trial in range(200):
        ns = int(rng.integers(1, 8))
        sv = []
        for _ in range(ns):
            k = int(rng.integers(1, 12))
            s = np.sort(rng.choice([1e-14, 1e-9, 1e-3, 0.01, 0.1, 0.25, 0.5], k))[::-1].copy()
            sv.append(s)
        max_dim = int(rng.choice([-1, 1, 3, 8, 40]))
        canpre = float(rng.choice([1e-12, 1e-10, 1e-8, 2.0]))
        a = [s.copy() for s in sv]
        t_ref = o_truncate(max_dim, canpre, a)
        t, b = dm.truncate_dims([s.copy() for s in sv], max_dim, canpre)
        assert list(t) == t_ref
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


MODELS = [
    (2, 2, 4, 3, -1.0, -1.2, 2.0, 0.3, False, False),
    (4, 4, 16, 4, -1.0, -1.0, 2.0, 0.0, False, False),   # 4x4 open, unit filling
    (8, 4, 32, 4, -1.0, -1.0, 3.0, 0.0, False, True),    # 8x4 cylinder, unit filling
    (5, 3, 7, 3, -0.5, -1.5, 1.0, -0.2, True, True),
    (6, 1, 6, 4, -1.0, 0.0, 4.0, 0.0, False, False),
]


@pytest.mark.parametrize("args", MODELS)
def test_qn_tables_and_mpo_match_oracle(args):
    nx, ny, N, d, jx, jy, U, mu, px, py = args
    om = SpinlessBoseSquare(nx, ny, N, d, jx=jx, jy=jy, U=U, mu=mu, period_x=px, period_y=py)
    pm = dm.SpinlessModel(nx, ny, N, d, hop_x=jx, hop_y=jy, U=U, mu=mu, period_x=px, period_y=py)
    tabs = om.space_tables()
    for b in range(om.L + 1):
        assert list(pm.space_table(b)) == [q[0] for q in tabs[b]]
    W, qt = om.mpo()
    for s in range(om.L):
        assert np.array_equal(pm.mpo_site(s), W[s]), s
    for b in range(om.L + 1):
        assert list(pm.mpo_table(b)) == [q[0] for q in qt[b]]


SPINFUL_MODELS = [
    # nx, ny, n_up, n_down, d, flux, jx, jy, (U_uu, U_dd, U_ud), mu, px, py
    (2, 2, 2, 2, 3, 0.25, -1.0, -1.0, (4.0, 4.0, 2.0), 0.0, False, False),
    (4, 4, 8, 8, 3, 0.25, -1.0, -1.0, (4.0, 4.0, 2.0), 0.0, False, False),  # 4x4, local dim 9
    (3, 3, 2, 4, 3, 0.2, -0.7, -1.1, (1.0, 2.0, 1.5), 0.3, True, True),
]


@pytest.mark.parametrize("args", SPINFUL_MODELS)
def test_spinful_tables_and_mpo_match_oracle(args):
    nx, ny, nu, nd, d, flux, jx, jy, (uu, udd, ud), mu, px, py = args
    om = SpinfulBoseSquare(nx, ny, nu, nd, d, flux=flux, jx=jx, jy=jy, U_uu=uu, U_dd=udd, U_ud=ud,
                           mu=mu, period_x=px, period_y=py)
    pm = dm.SpinfulModel(nx, ny, nu, nd, d, flux=flux, hop_x=jx, hop_y=jy, U_uu=uu, U_dd=udd,
                         U_ud=ud, mu=mu, period_x=px, period_y=py)
    assert pm.complex and pm.nq == 2 and pm.L == 2 * nx * ny
    tabs = om.space_tables()
    for b in range(om.L + 1):
        assert [tuple(q) for q in pm.space_table(b)] == list(tabs[b])
    W, qt = om.mpo()
    for s in range(om.L):
        assert np.array_equal(pm.mpo_site(s), W[s]), s
    for b in range(om.L + 1):
        assert [tuple(q) for q in pm.mpo_table(b)] == list(qt[b])


def test_complex_spinless_mpo_is_real_mpo():
    pr = dm.SpinlessModel(4, 2, 8, 3, U=2.0)
    pz = dm.SpinlessModel(4, 2, 8, 3, U=2.0, complex=True)
    om = SpinlessBoseSquare(4, 2, 8, 3, U=2.0, complex=True)
    W, _ = om.mpo()
    assert pz.complex and not pr.complex
    for s in range(8):
        assert np.array_equal(pz.mpo_site(s), W[s])
        assert np.array_equal(pz.mpo_site(s).real, pr.mpo_site(s))
        assert not np.any(pz.mpo_site(s).imag)


def test_random_matrices_match_oracle_stream():
    # the threaded jump-ahead fill reproduces one sequential pass of the
    # stream: real raw draws, then complex with the expansion's diagonal
    shapes = [(900, 25)] * 30 + [(3, 7), (0, 5), (17, 1), (25, 40)]
    M = dm.random_matrices(shapes, skip=12345)
    g = GlibcRand(1)
    g.skip(12345)
    for m, (r, c) in zip(M, shapes):
        assert np.array_equal(m, g.random_matrix(r, c))
    w = [complex(0.3 * k, -0.2) for k in range(len(shapes))]
    Mz = dm.random_matrices(shapes, skip=777, complex=True, w=w)
    g = GlibcRand(1)
    g.skip(777)
    for k, (m, (r, c)) in enumerate(zip(Mz, shapes)):
        z = g.complex_matrix(r, c)
        for q in range(min(r, c)):
            z[q, q] = 0.1 * w[k]
        assert np.array_equal(m, z), k


@pytest.mark.parametrize("skip", [4095, 4096, 100003, 300000])
def test_rand_jump_ahead(skip):
    g = GlibcRand(1)
    g.skip(skip)
    assert list(dm.rand_draws(40, skip=skip)) == list(g.draws(40))


def test_sym_eig_closed_form_2x2():
    # eigenvalues of [[a, b], [b, c]]: (a + c)/2 -+ sqrt(((a - c)/2)^2 + b^2)
    a, b, c = 1.5, -0.75, -2.0
    r = np.hypot((a - c) / 2, b)
    lo, hi = (a + c) / 2 - r, (a + c) / 2 + r
    w, v = dm.sym_eig(np.array([[a, b], [b, c]]))
    assert abs(w[0] - lo) < 1e-15 and abs(w[1] - hi) < 1e-15
    # column 0 is proportional to (b, lo - a), largest |component| positive
    u = np.array([b, lo - a]) / np.hypot(b, lo - a)
    u *= np.sign(u[np.argmax(np.abs(u))])
    assert np.abs(v[:, 0] - u).max() < 1e-15
    e0, v0 = dm.sym_eig_lowest(np.array([[a, b], [b, c]]))
    assert abs(e0 - lo) < 1e-15 and np.abs(v0 - u).max() < 1e-15


@pytest.mark.parametrize("n", [1, 3, 17, 64, 99])
def test_sym_eig_matches_numpy(n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n))
    a = a + a.T
    w, v = dm.sym_eig(a)
    ref = np.linalg.eigvalsh(a)
    scale = np.abs(a).max() * n
    assert np.abs(w - ref).max() < 1e-14 * scale
    assert np.abs(a @ v - v * w).max() < 1e-13 * scale
    assert np.abs(v.T @ v - np.eye(n)).max() < 1e-13
    e0, v0 = dm.sym_eig_lowest(a)
    assert abs(e0 - ref[0]) < 1e-14 * scale
    assert abs(np.linalg.norm(v0) - 1.0) < 1e-14
    assert np.abs(a @ v0 - e0 * v0).max() < 1e-13 * scale
    # same vector as the Jacobi solver (non-degenerate lowest eigenvalue, same sign rule)
    assert np.abs(v0 - v[:, 0]).max() < 1e-10
    assert v0[np.argmax(np.abs(v0))] > 0


@pytest.mark.parametrize("m", [5, 40, 99])
def test_sym_eig_lowest_complex_embedding(m):
    # the dense complex edge solve diagonalises [[A, -B], [B, A]] for H = A + iB;
    # its lowest eigenvalue is H's (twice) and (x, y) gives the ground state x + iy
    rng = np.random.default_rng(100 + m)
    h = rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))
    h = h + h.conj().T
    emb = np.block([[h.real, -h.imag], [h.imag, h.real]])
    e0, v0 = dm.sym_eig_lowest(emb)
    ev, evec = np.linalg.eigh(h)
    assert abs(e0 - ev[0]) < 1e-13 * np.abs(h).max() * m
    g = v0[:m] + 1j * v0[m:]
    assert abs(np.linalg.norm(g) - 1.0) < 1e-13
    assert abs(abs(np.vdot(evec[:, 0], g)) - 1.0) < 1e-10


def test_sym_eig_lowest_tridiagonal_and_split():
    # a Lanczos-type tridiagonal matrix, and one that splits into two blocks
    d = np.array([2.0, -1.0, 0.5, 3.0, -2.5, 1.0])
    e = np.array([0.3, -0.7, 0.2, 1e-3, 0.9])
    for ee in (e, np.where(np.arange(5) == 2, 0.0, e)):
        t = np.diag(d) + np.diag(ee, 1) + np.diag(ee, -1)
        e0, v0 = dm.sym_eig_lowest(t)
        ref, rv = np.linalg.eigh(t)
        assert abs(e0 - ref[0]) < 1e-14
        assert abs(abs(v0 @ rv[:, 0]) - 1.0) < 1e-12
