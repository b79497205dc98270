"""Timing entry points of the HBM-side kernels: environment updates on a
run's current state and the Lanczos vector kernels.  They report sane
numbers and leave the run untouched (the sweep continues bit-identically)."""
import numpy as np
import pytest

import dmrgcc_amd as dm
from tests._support import sweep_rows

pytestmark = pytest.mark.gpu


def _run():
    return dm.SpinlessModel(3, 2, 6, 3, U=2.0).start(sweep_rows(30, 4, niter=6), first_dim=4)


def test_env_time_leaves_run_untouched():
    ref = _run()
    ref.advance()
    e_ref = np.array(ref.result()["energies"])

    run = _run()
    assert run.advance(2, 3)
    for leigh in (1, 2):
        t = run.env_time(3, leigh, reps=2)
        assert t["f_ms"] > 0 and t["out_ms"] > 0 and t["flops"] > 0
        assert 0 < t["bytes_min"] <= t["bytes_exec"]
    run.advance()
    e = np.array(run.result()["energies"])
    assert np.array_equal(e, e_ref)


def test_env_time_rejects_bad_arguments():
    run = _run()
    assert run.advance(2, 3)
    with pytest.raises(dm.DmrgError):
        run.env_time(99, 1)
    with pytest.raises(dm.DmrgError):
        run.env_time(3, 3)


def test_vector_time():
    t = dm.vector_time(1 << 16, reps=5)
    for k in ("dot_ms", "axpy_ms", "lanczos_step_ms"):
        assert t[k] > 0
    assert t["bytes"]["axpy"] == 24.0 * (1 << 16)
