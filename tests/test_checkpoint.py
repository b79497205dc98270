"""Tensor-space files (RealTensorSpace::Write/ReadTensorSpace,
real_tensor_space.cc:171-277): the oracle's restatement of the layout, round
trip and byte layout on the CPU, then on the device against the library's
writer and reader (checkpoint after every sweep, resume from a file)."""
import struct

import numpy as np
import pytest

import dmrgcc_amd as dm
from oracle.dmrg import DMRG, Space
from oracle.io import read_space, write_space
from oracle.model import SpinfulBoseSquare, SpinlessBoseSquare
from oracle.rng import GlibcRand
from tests._support import sweep_rows

ROWS = sweep_rows(40, 6, niter=8, canpre=1e-8)


def _oracle(model, rows, resume=None):
    return DMRG(model, rows, len(rows) // 2, first_block=20, first_dim=4, resume=resume)


def _models():
    return [("real", SpinlessBoseSquare(3, 2, 6, 3, U=2.0),
             lambda: dm.SpinlessModel(3, 2, 6, 3, U=2.0)),
            # no up/down exchange symmetry: an exactly degenerate bond spectrum
            # leaves the singular vectors inside it to each SVD's choice, and
            # the expansion noise (L K random) then differs between any two
            # implementations, so trajectories are compared away from such
            # symmetric states
            ("spinful", SpinfulBoseSquare(3, 2, 2, 3, 3, flux=0.2, U_uu=1.0, U_dd=2.0, U_ud=1.5),
             lambda: dm.SpinfulModel(3, 2, 2, 3, 3, flux=0.2, U_uu=1.0, U_dd=2.0, U_ud=1.5))]


def _same_space(a, b, exact=False):
    """Tables, sectors, dims, physical indices and block shapes bit-exact;
    the entries either bit-exact, or through the Frobenius norm of every
    block, which no gauge change of a bond (a unitary mixing of a sector's
    singular vectors, e.g. the phase of a vector whose largest-modulus
    entry is tied) alters."""
    assert a.tables == b.tables
    assert (a.max_block, a.max_dim) == (b.max_block, b.max_dim)
    assert a.canpre == b.canpre and a.noise == b.noise
    for la, lb in zip(a.lat, b.lat):
        assert (la.d, la.left_block, la.left_dim, la.right_block, la.right_dim) == \
               (lb.d, lb.left_block, lb.left_dim, lb.right_block, lb.right_dim)
        assert la.phys == lb.phys
        assert (la.ket.left, la.ket.right) == (lb.ket.left, lb.ket.right)
        for ma, mb in zip(la.ket.mats, lb.ket.mats):
            assert ma.shape == mb.shape
            if exact:
                assert np.array_equal(ma, mb)
            elif ma.size:
                assert abs(np.linalg.norm(ma) - np.linalg.norm(mb)) < 1e-8


@pytest.mark.parametrize("which", [0, 1])
def test_oracle_space_file_round_trip(tmp_path, which):
    name, om, _ = _models()[which]
    o = _oracle(om, ROWS[:2])
    o.run()
    p = tmp_path / "space.dat"
    write_space(str(p), o.space)
    raw = p.read_bytes()
    L = om.L
    assert struct.unpack_from("<4i", raw, 0) == (L, L + 1, L - 1, om.nq)
    # trailing: max_block, max_dim, canonical precision, noise factor
    mb, md = struct.unpack_from("<2i", raw, len(raw) - 24)
    cp, nf = struct.unpack_from("<2d", raw, len(raw) - 16)
    assert (mb, md, cp, nf) == (o.space.max_block, o.space.max_dim, o.space.canpre, o.space.noise)
    fresh = Space(om, GlibcRand(1))
    read_space(str(p), fresh)
    _same_space(o.space, fresh, exact=True)


@pytest.mark.gpu
@pytest.mark.parametrize("which", [0, 1])
def test_record_file_matches_oracle_state(tmp_path, which):
    # the library's checkpoint after the last sweep, read by the oracle's
    # reader, equals the oracle's own state after the same sweeps
    name, om, mk = _models()[which]
    p = tmp_path / "record.dat"
    r = mk().run(ROWS[:4], first_block=20, first_dim=4, record=str(p))
    o = _oracle(om, ROWS[:4])
    e = o.run()
    assert np.abs(r["energies"] - np.array(e)).max() < 1e-10 * abs(e[-1])
    got = Space(om, GlibcRand(1))
    read_space(str(p), got)
    _same_space(got, o.space)


@pytest.mark.gpu
@pytest.mark.parametrize("which", [0, 1])
def test_resume_from_oracle_file(tmp_path, which):
    # cache_resume: the library starts from a file the oracle wrote and runs
    # the remaining sweeps as the oracle does from the same file
    name, om, mk = _models()[which]
    o1 = _oracle(om, ROWS[:2])
    o1.run()
    p = tmp_path / "resume.dat"
    write_space(str(p), o1.space)
    # A saved space is right-canonical on every bond but the first, so the
    # SVDs of the resumed Initialize see all-equal singular values and each
    # implementation picks its own vectors inside them (MKL's choice is just
    # as arbitrary).  The expansion noise L K random is not gauge-covariant,
    # so the comparison runs the resumed sweeps without it: energies, kept
    # dimensions and the rand() count are then gauge-independent.
    rest = sweep_rows(40, 4, niter=8, canpre=1e-8, noise=[-1] * 4)
    r = mk().run(rest, first_block=20, first_dim=4, resume=str(p))
    o2 = _oracle(om, rest, resume=str(p))
    e = np.array(o2.run())
    assert r["rand_draws"] == o2.rng.count
    assert np.abs(r["energies"] - e).max() < 1e-10 * np.abs(e).max()
    # and the library's own writer reproduces the file it read
    run = mk().start(rest, first_block=20, first_dim=4, resume=str(p))
    q = tmp_path / "again.dat"
    run.write_space(str(q))
    a, b = Space(om, GlibcRand(1)), Space(om, GlibcRand(1))
    read_space(str(q), a)
    o3 = _oracle(om, rest, resume=str(p))
    o3.initialize()
    _same_space(a, o3.space)
