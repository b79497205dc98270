"""Pins the CPU oracle against known answers, closed forms and brute force
(no GPU).  See oracle/__init__.py for the list of what pins what."""
import math

import numpy as np
import pytest

from oracle import ed
from oracle.dmrg import DMRG, heff_apply, heff_dense, heff_tasks, svd, truncate_dims
from oracle.model import SpinfulBoseSquare, SpinlessBoseSquare, parallel_matrix
from oracle.rng import GlibcRand
from oracle.sort import quicksort
from tests._support import oracle_state, sweep_rows


def test_rand_known_answers():
    g = GlibcRand(1)
    assert [g.rand() for _ in range(4)] == [1804289383, 846930886, 1681692777, 1714636915]


def test_random_matrix_is_column_major():
    # RealMatrix(3,3)->RandomMatrix() fills column by column
    m = GlibcRand(1).random_matrix(3, 3)
    assert np.allclose(m[:, 0], [0.83, 0.86, 0.77])
    assert np.allclose(m[0, 1], 0.15)


def test_quicksort_reference_example():
    # test/test_real_matrix.cc:9-20: ascending with index
    v = [0.2, 0.4, 2.6, 0.8, 1, 3.2]
    idx = list(range(6))
    quicksort(v, idx, flag=1)
    assert v == [0.2, 0.4, 0.8, 1, 2.6, 3.2]
    assert idx == [0, 1, 3, 4, 2, 5]


@pytest.mark.parametrize("flag", [0, 1])
def test_quicksort_is_sorted_permutation(flag):
    rng = np.random.default_rng(3)
    for n in (1, 2, 7, 50, 300):
        v = list(rng.integers(0, 6, n).astype(float))  # many ties
        idx = list(range(n))
        orig = list(v)
        quicksort(v, idx, flag=flag)
        assert sorted(orig, reverse=(flag == 0)) == v
        assert [orig[i] for i in idx] == v


def test_svd_closed_form():
    U, s, V = svd(np.array([[1.0, 2, 3], [4, 5, 6]]))
    exact = [math.sqrt((91 + math.sqrt(8065)) / 2), math.sqrt((91 - math.sqrt(8065)) / 2)]
    assert np.allclose(s, exact, rtol=1e-14)
    assert V.shape == (3, 3) and np.all(V[2] == 0)


def test_parallel_matrix_example():
    # [[1,3,1],[2,4,2]]: column 2 = column 0
    pos, T = parallel_matrix(np.array([[1.0, 3, 1], [2, 4, 2]]), 0)
    assert pos == [0, 1]
    assert np.allclose(T, [[1, 0, 1], [0, 1, 0], [0, 0, 0]])
    # rows: row 1 = 2 * row 0 in [[1,2],[2,4],[0,1]]
    pos, T = parallel_matrix(np.array([[1.0, 2], [2, 4], [0, 1]]), 1)
    assert pos == [0, 2]
    assert np.allclose(T[1], [2, 0, 0])


def test_truncate_rule():
    # canpre cut, then global top-max_dim, then forced single state
    sv = [np.array([0.9, 0.3, 1e-12]), np.array([0.5, 0.2]), np.array([1e-13])]
    t = truncate_dims(-1, 1e-10, [s.copy() for s in sv])
    assert t == [2, 2, 1]
    s2 = [s.copy() for s in sv]
    t = truncate_dims(3, 1e-10, s2)
    assert t == [2, 1, 1] and s2[2][0] == 0.0
    s3 = [s.copy() for s in sv]
    t = truncate_dims(2, 1e-10, s3)
    assert t == [1, 1, 1]


def contract_mpo(W):
    M = W[0][0]
    for s in range(1, len(W)):
        M = np.einsum('aXY,abpq->bXpYq', M, W[s])
        b, X, p, Y, q = M.shape
        M = M.reshape(b, X * p, Y * q)
    return M[0]


@pytest.mark.parametrize("nx,ny,d,py", [(2, 2, 3, False), (3, 2, 3, False), (2, 3, 2, True),
                                        (4, 1, 3, False), (3, 3, 2, True)])
def test_mpo_equals_bruteforce(nx, ny, d, py):
    m = SpinlessBoseSquare(nx, ny, nx * ny // 2, d, jx=-1.0, jy=-0.7, U=1.3, mu=0.2, period_y=py)
    W, _ = m.mpo()
    Hb = ed.full_hamiltonian(m.L, d, m.bonds(), m.U, m.mu)
    assert np.abs(contract_mpo(W) - Hb).max() < 1e-12


def test_ed_closed_forms():
    # two sites, one boson, d=2: E0 = -|J|
    assert abs(ed.ground_energy(2, 2, 1, [(0, 1, -1.0)], 0.0) + 1.0) < 1e-12
    # open chain of L sites, one boson, U irrelevant: E0 = 2J cos(pi/(L+1)) (J<0)
    L = 6
    e = ed.ground_energy(L, 3, 1, [(i, i + 1, -1.0) for i in range(L - 1)], 3.0)
    assert abs(e + 2 * math.cos(math.pi / (L + 1))) < 1e-12


def test_heff_block_equals_dense_and_symmetric():
    dm = oracle_state(3, 2, 6, 3, maxm=30, half=2)
    site = 2
    lat, W = dm.space.lat[site], dm.W[site]
    H = heff_dense(lat, W, dm.Lenv[site], dm.Renv[site])
    n = H.shape[0]
    rng = np.random.default_rng(0)
    x = rng.standard_normal(n)
    y = heff_apply(lat, W, dm.Lenv[site], dm.Renv[site], x, np.zeros(n))
    assert np.allclose(y, H @ x, atol=1e-12 * np.abs(y).max())
    assert np.abs(H - H.T).max() < 1e-12 * np.abs(H).max()


@pytest.mark.parametrize("nx,ny,N,d,U,py", [(3, 2, 6, 3, 2.0, False), (2, 2, 4, 4, 3.0, False),
                                            (4, 1, 4, 3, 1.0, False), (2, 3, 6, 3, 2.0, True)])
def test_dmrg_reaches_exact_ground_energy(nx, ny, N, d, U, py):
    m = SpinlessBoseSquare(nx, ny, N, d, jx=-1.0, jy=-1.0, U=U, period_y=py)
    dm = DMRG(m, sweep_rows(80, 6, niter=10, canpre=1e-12), 3, first_block=20, first_dim=4)
    E = dm.run()
    e0 = ed.ground_energy(m.L, d, N, m.bonds(), U)
    assert abs(E[-1] - e0) < 1e-10 * abs(e0)
    # variational: no half sweep goes below the exact energy
    assert min(E) > e0 - 1e-10 * abs(e0)


@pytest.mark.parametrize("nx,ny,d,flux,py", [(2, 2, 2, 0.25, False), (1, 3, 2, 0.2, True),
                                             (3, 1, 3, 0.1, False), (2, 2, 3, 0.125, False)])
def test_spinful_mpo_equals_bruteforce(nx, ny, d, flux, py):
    # complex MPO with flux phases against the Fock-space H built from the
    # lattice geometry (Landau gauge on y bonds)
    m = SpinfulBoseSquare(nx, ny, 1, 1, d, flux=flux, jx=-1.0, jy=-0.8, U_uu=4.0, U_dd=3.0,
                          U_ud=2.0, mu=0.3, period_y=py)
    W, _ = m.mpo()
    Hb = ed.fock_full(m.L, d, m.hops(), ed.spinful_diag(m))
    H = contract_mpo(W)
    assert np.abs(H - Hb).max() < 1e-12
    assert np.abs(H - H.conj().T).max() < 1e-12
    if ny > 1:  # y bonds carry the flux: H is genuinely complex
        assert np.abs(Hb.imag).max() > 0.1


def test_spinful_qn_rule_and_tables():
    m = SpinfulBoseSquare(2, 2, 2, 3, 3)
    tabs = m.space_tables()
    assert tabs[0] == [(0, 0)] and tabs[-1] == [(2, 3)]
    for i in range(1, m.L + 1):
        up, dn = (i + 1) // 2, i // 2
        full = {(a, b) for a in range(0, 2 * up + 1) for b in range(0, 2 * dn + 1)
                if 2 - a <= 2 * (4 - up) and 3 - b <= 2 * (4 - dn) and a <= 2 and b <= 3}
        assert set(tabs[i]) == full
    # even sites carry species 0, odd sites species 1
    assert m.check(0, (0, 0), (2, 0)) == 2 and m.check(0, (0, 0), (0, 1)) == -1
    assert m.check(1, (1, 0), (1, 2)) == 2 and m.check(1, (0, 0), (1, 0)) == -1


def test_complex_random_matrix_draw_order():
    # two draws per entry, column-major; the first draw of a pair is the
    # imaginary part (g++ evaluates the second constructor argument first)
    a, b = GlibcRand(1), GlibcRand(1)
    z = a.complex_matrix(2, 3)
    v = 0.01 * (b.draws(12) % 100)
    for j in range(3):
        for i in range(2):
            k = 2 * (j * 2 + i)
            assert z[i, j] == complex(v[k + 1], v[k])


@pytest.mark.parametrize("which", ["spinful", "zspinless"])
def test_complex_dmrg_reaches_exact_ground_energy(which):
    if which == "spinful":
        m = SpinfulBoseSquare(2, 2, 2, 2, 3, flux=0.25, U_uu=4.0, U_dd=4.0, U_ud=2.0)
        e0 = ed.spinful_ground_energy(m)
        rows = sweep_rows(100, 6, niter=10, canpre=1e-12)
    else:
        m = SpinlessBoseSquare(3, 2, 6, 3, U=2.0, complex=True)
        e0 = ed.ground_energy(m.L, 3, 6, m.bonds(), 2.0)
        rows = sweep_rows(80, 6, niter=10, canpre=1e-12)
    dmr = DMRG(m, rows, 3, first_block=20, first_dim=4)
    E = dmr.run()
    assert abs(E[-1] - e0) < 1e-10 * abs(e0)
    assert min(E) > e0 - 1e-10 * abs(e0)
