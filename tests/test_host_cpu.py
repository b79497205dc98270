"""The C-ABI library loads and exports every declared symbol, and its host
restatements (rand stream, QuickSort, truncation rule, QN tables, MPO) agree
bit-exactly with the oracle.  No GPU is touched."""
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
