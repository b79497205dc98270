"""MPO files (RealTensorHamiltonian::Write/ReadTensorHamiltonian,
real_tensor_hamiltonian.cc:74-150) and environment files
(RealTensorContraction::Write/ReadTensorContraction,
real_tensor_contraction.cc:73-109): a byte fixture written field by field in
the reference's order is read by the library and by the oracle; files go
library -> oracle and oracle -> library unchanged.  The MPO tests need no
GPU; the environment files of a run are checked in tests/test_gpu_files.py."""
import struct

import numpy as np
import pytest

import dmrgcc_amd as dm
from oracle.io import read_contraction, read_hamiltonian, write_hamiltonian
from oracle.model import SpinfulBoseSquare, SpinlessBoseSquare


def _i(*v):
    return struct.pack("<%di" % len(v), *v)


def _mat(m, cplx=False):
    m = np.asarray(m, dtype=np.complex128 if cplx else np.float64)
    return _i(m.shape[0], m.shape[1], m.size) + m.tobytes()


def mpo_fixture(cplx=False):
    """Two sites, d = 2, basic operators I and n, bond tables of sizes 1, 2,
    1 (one quantum number; two for the complex two-component model),
    W_0 = [I, n] (1 x 2), W_1 = [n; 2 I] (2 x 1)."""
    I = np.eye(2)
    n = np.diag([0.0, 1.0])
    if cplx:
        n = n * (1 - 0.5j)
    W0 = np.stack([I, n])[None]          # (1, 2, 2, 2)
    W1 = np.stack([n, 2 * I])[:, None]   # (2, 1, 2, 2)
    if cplx:
        tables = [[(0, 0)], [(0, 0), (1, 0)], [(0, 0)]]
        qt = _i(2) + _i(1, 2, 1) + _i(0, 0) + _i(0, 0, 1, 0) + _i(0, 0)
    else:
        tables = [[(0,)], [(0,), (1,)], [(0,)]]
        qt = _i(1) + _i(1, 2, 1) + _i(0) + _i(0, 1) + _i(0)
    b = _i(2, 3, 1) + _i(2) + _mat(I, cplx) + _mat(n, cplx) + qt
    for w in (W0, W1):
        b += _i(2, w.shape[0], w.shape[1])
        for a in range(w.shape[0]):
            for c in range(w.shape[1]):
                b += _mat(w[a, c], cplx)
    return b, [W0, W1], [I, n], tables


@pytest.mark.parametrize("cplx", [False, True])
def test_mpo_byte_fixture_read_by_both(tmp_path, cplx):
    raw, W, basic, tables = mpo_fixture(cplx)
    p = tmp_path / "mpo.dat"
    p.write_bytes(raw)
    oW, obasic, otab = read_hamiltonian(str(p), cplx)
    assert all(np.array_equal(a, b) for a, b in zip(oW, W))
    assert all(np.array_equal(a, b) for a, b in zip(obasic, basic))
    assert otab == tables
    # the library reads it into a two-site, d = 2 model of the same kind
    m = dm.SpinfulModel(1, 1, 1, 1, 2, U_ud=1.0) if cplx else dm.SpinlessModel(2, 1, 1, 2, U=1.0)
    m.read_mpo(str(p))
    for s in range(2):
        assert np.array_equal(m.mpo_site(s), W[s])
    assert np.array_equal(np.asarray(m.mpo_table(1)).reshape(2, -1), np.array(tables[1]))
    # and writes the same bytes back
    q = tmp_path / "mpo2.dat"
    m.write_mpo(str(q))
    assert q.read_bytes() == raw


@pytest.mark.parametrize("which", ["spinless", "spinful"])
def test_mpo_roundtrips_library_and_oracle(tmp_path, which):
    if which == "spinless":
        m = dm.SpinlessModel(3, 2, 6, 3, hop_x=-1.0, hop_y=-0.75, U=2.0, mu=0.1, period_y=True)
        o = SpinlessBoseSquare(3, 2, 6, 3, jx=-1.0, jy=-0.75, U=2.0, mu=0.1, period_y=True)
        other = dm.SpinlessModel(3, 2, 6, 3, hop_x=-0.3, hop_y=-0.2, U=7.0, period_y=False)
    else:
        m = dm.SpinfulModel(2, 2, 2, 2, 3, flux=0.25, U_uu=4.0, U_dd=4.0, U_ud=2.0)
        o = SpinfulBoseSquare(2, 2, 2, 2, 3, flux=0.25, U_uu=4.0, U_dd=4.0, U_ud=2.0)
        other = dm.SpinfulModel(2, 2, 2, 2, 3, flux=0.1, U_uu=1.0, U_dd=2.0, U_ud=0.5)
    cplx = m.complex
    W_ref, tabs_ref = o.mpo()
    p = tmp_path / "lib.dat"
    m.write_mpo(str(p))
    oW, obasic, otab = read_hamiltonian(str(p), cplx)
    assert len(oW) == m.L
    for s in range(m.L):
        assert np.array_equal(oW[s], W_ref[s])
    assert all(np.array_equal(a, b) for a, b in zip(obasic, o.basic_ops()))
    assert [list(map(tuple, t)) for t in otab] == [list(map(tuple, t)) for t in tabs_ref]
    # the oracle's file read into a model with other couplings: its MPO
    # becomes the file's
    q = tmp_path / "oracle.dat"
    write_hamiltonian(str(q), W_ref, o.basic_ops(), tabs_ref, cplx)
    assert q.read_bytes() == p.read_bytes()
    other.read_mpo(str(q))
    for s in range(m.L):
        assert np.array_equal(other.mpo_site(s), W_ref[s])


def test_model_without_terms_is_refused():
    # one lattice site, no bond and no on-site term: an empty MPO
    with pytest.raises(dm.DmrgError, match="no term"):
        dm.SpinfulModel(1, 1, 1, 1, 2)


def test_mpo_read_refuses_other_models(tmp_path):
    p = tmp_path / "mpo.dat"
    dm.SpinlessModel(3, 2, 6, 3, U=2.0).write_mpo(str(p))
    with pytest.raises(dm.DmrgError, match="number of sites"):
        dm.SpinlessModel(2, 2, 4, 3, U=2.0).read_mpo(str(p))
    with pytest.raises(dm.DmrgError):
        dm.SpinlessModel(3, 2, 6, 4, U=2.0).read_mpo(str(p))   # another physical dimension
    p.write_bytes(p.read_bytes()[:-8])
    with pytest.raises(dm.DmrgError, match="ends early"):
        dm.SpinlessModel(3, 2, 6, 3, U=2.0).read_mpo(str(p))


def env_fixture(cplx=False):
    """Two channels: channel 0 holds blocks (0, 1) 2 x 3 and (1, 0) 1 x 1,
    channel 1 holds none."""
    a = np.arange(6, dtype=float).reshape(2, 3) - 2.5
    b = np.array([[7.25]])
    if cplx:
        a = a + 1j * a[::-1]
        b = b * (1 + 2j)
    raw = _i(2) + _i(0, 1) + _mat(a, cplx) + _i(1, 0) + _mat(b, cplx) + _i(0)
    return raw, a, b


@pytest.mark.parametrize("cplx", [False, True])
def test_env_byte_fixture_oracle_reader(tmp_path, cplx):
    raw, a, b = env_fixture(cplx)
    p = tmp_path / "env.dat"
    p.write_bytes(raw)
    env = read_contraction(str(p), 2, cplx)
    assert (env[0].left, env[0].right) == ([0, 1], [1, 0])
    assert np.array_equal(env[0].mats[0], a) and np.array_equal(env[0].mats[1], b)
    assert env[1].n == 0
    with pytest.raises(ValueError, match="trailing"):
        read_contraction(str(p), 1, cplx)
