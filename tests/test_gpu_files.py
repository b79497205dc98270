"""Environment files of a run (RealTensorContraction::WriteTensorContraction /
ReadTensorContraction, real_tensor_contraction.cc:73-109): the library's files
read by the oracle, the oracle's and a byte fixture read by the library."""
import numpy as np
import pytest

import dmrgcc_amd as dm
from oracle.io import read_contraction, write_contraction
from tests._support import oracle_state, sweep_rows
from tests.test_files import env_fixture

pytestmark = pytest.mark.gpu


def _export(run, site, which, nch):
    blocks, data = run.site_blocks(site, which)
    out = [[] for _ in range(nch)]
    for (o, l, r, rows, cols, off) in blocks:
        out[o].append((l, r, data[off:off + rows * cols].reshape(rows, cols)))
    return out


def _same(env, ref):
    assert len(env) == len(ref)
    for bm, blocks in zip(env, ref):
        assert bm.n == len(blocks)
        for k, (l, r, m) in enumerate(blocks):
            assert (bm.left[k], bm.right[k]) == (l, r)
            assert np.array_equal(bm.mats[k], m)


def _lib_env(res, nch):
    host = res.to_host()
    out = [[] for _ in range(nch)]
    for (o, l, r, rows, cols, off) in res.blocks:
        out[o].append((l, r, host[off:off + rows * cols].reshape(rows, cols)))
    return out


@pytest.mark.parametrize("which", ["spinless", "spinful"])
def test_run_environment_files(tmp_path, which):
    if which == "spinless":
        model = dm.SpinlessModel(3, 2, 6, 3, U=2.0)
    else:
        model = dm.SpinfulModel(2, 2, 2, 2, 3, flux=0.25, U_uu=4.0, U_dd=4.0, U_ud=2.0)
    run = model.start(sweep_rows(30, 4, niter=6), first_dim=4)
    site = 2
    assert run.advance(2, site)
    W = model.mpo_site(site)
    for leigh, nch in ((0, W.shape[0]), (1, W.shape[1])):
        p = tmp_path / ("env%d.dat" % leigh)
        run.write_env(site, leigh, str(p))
        ref = _export(run, site, 1 + leigh, nch)
        _same(read_contraction(str(p), nch, model.complex), ref)
        got = dm.env_file_read(str(p), nch, model.complex)
        for a, b in zip(_lib_env(got, nch), ref):
            assert [(l, r) for l, r, _ in a] == [(l, r) for l, r, _ in b]
            for (_, _, x), (_, _, y) in zip(a, b):
                assert np.array_equal(x, y)


def test_oracle_and_fixture_environment_files_read_by_library(tmp_path):
    dmo = oracle_state(3, 2, 6, 3, maxm=30, half=2, site=2, niter=6)
    env = dmo.Lenv[2]
    p = tmp_path / "oracle_env.dat"
    write_contraction(str(p), env)
    got = _lib_env(dm.env_file_read(str(p), len(env)), len(env))
    for bm, blocks in zip(env, got):
        assert [(l, r) for l, r, _ in blocks] == list(zip(bm.left, bm.right))
        for (_, _, x), m in zip(blocks, bm.mats):
            assert np.array_equal(x, m)
    for cplx in (False, True):
        raw, a, b = env_fixture(cplx)
        q = tmp_path / ("fixture%d.dat" % cplx)
        q.write_bytes(raw)
        got = _lib_env(dm.env_file_read(str(q), 2, cplx), 2)
        assert [(l, r) for l, r, _ in got[0]] == [(0, 1), (1, 0)] and got[1] == []
        assert np.array_equal(got[0][0][2], a) and np.array_equal(got[0][1][2], b)
    with pytest.raises(dm.DmrgError, match="trailing"):
        dm.env_file_read(str(q), 1, True)


def test_library_resumes_the_hand_written_space_file(tmp_path):
    # cache_resume from the byte fixture: the library and the oracle read the
    # same file and run the same sweeps (no noise) to the same energies
    from oracle.dmrg import DMRG
    from tests.test_oracle_sweep import space_fixture
    _, raw, _ = space_fixture()
    p = tmp_path / "fixture_space.dat"
    p.write_bytes(raw)
    rows = sweep_rows(20, 4, niter=6, canpre=1e-10, noise=[-1] * 4)
    r = dm.SpinlessModel(2, 1, 2, 3, hop_x=-1.0, U=1.0).run(rows, first_block=20, first_dim=4, resume=str(p))
    from oracle.model import SpinlessBoseSquare
    o = DMRG(SpinlessBoseSquare(2, 1, 2, 3, jx=-1.0, U=1.0), rows, 2, resume=str(p))
    e = np.array(o.run())
    assert np.abs(r["energies"] - e).max() < 1e-12 * np.abs(e).max()
