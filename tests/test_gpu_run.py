"""Full sweeps on the device against the oracle's sweeps on the same inputs
(same seeded rand() stream) and against exact diagonalisation.

With a truncation cut-off that FP64 singular values resolve (canpre 1e-8)
the whole trajectory is reproduced: every rand() draw, every truncation and
the per-half-sweep energies.  At canpre 1e-12 singular values sit at the
round-off floor of any SVD, single truncation decisions may flip, and only
the energies are compared."""
import numpy as np
import pytest

import dmrgcc_amd as dm
from oracle import ed
from oracle.dmrg import DMRG
from oracle.model import SpinlessBoseSquare
from tests._support import sweep_rows

pytestmark = pytest.mark.gpu

CASES = [
    # nx, ny, N, d, U, period_y, maxm, nhalf, first_dim
    (3, 2, 6, 3, 2.0, False, 60, 6, 4),
    (2, 3, 6, 4, 3.0, True, 80, 6, 6),
    (4, 3, 12, 3, 2.0, False, 40, 6, 6),
]


def _run_both(case, canpre):
    nx, ny, N, d, U, py, maxm, nhalf, fd = case
    rows = sweep_rows(maxm, nhalf, niter=10, canpre=canpre)
    om = SpinlessBoseSquare(nx, ny, N, d, jx=-1.0, jy=-1.0, U=U, period_y=py)
    o = DMRG(om, rows, nhalf // 2, first_block=20, first_dim=fd)
    e_ref = np.array(o.run())
    pm = dm.SpinlessModel(nx, ny, N, d, hop_x=-1.0, hop_y=-1.0, U=U, period_y=py)
    r = pm.run(rows, first_block=20, first_dim=fd, precision=dm.FP64)
    return om, o, e_ref, r


@pytest.mark.parametrize("case", CASES)
def test_run_trajectory_matches_oracle(case):
    om, o, e_ref, r = _run_both(case, 1e-8)
    assert r["rand_draws"] == o.rng.count
    assert np.abs(r["energies"] - e_ref).max() < 1e-10 * np.abs(e_ref).max()
    assert np.allclose(r["trunc_err"], o.trunc_err, rtol=1e-6, atol=1e-14)


@pytest.mark.parametrize("case", CASES[:2])
def test_run_energies_match_oracle_and_ed(case):
    om, o, e_ref, r = _run_both(case, 1e-12)
    assert np.abs(r["energies"] - e_ref).max() < 1e-9 * np.abs(e_ref).max()
    e0 = ed.ground_energy(om.L, om.d, om.N, om.bonds(), om.U)
    assert abs(r["energies"][-1] - e0) < 1e-9 * abs(e0)


@pytest.mark.parametrize("prec", [dm.FP32, dm.BF16X3])
def test_run_mixed_precision_energy(prec):
    nx, ny, N, d, U, py, maxm, nhalf, fd = CASES[0]
    rows = sweep_rows(maxm, nhalf, niter=10, canpre=1e-12)
    pm = dm.SpinlessModel(nx, ny, N, d, hop_x=-1.0, hop_y=-1.0, U=U, period_y=py)
    om = SpinlessBoseSquare(nx, ny, N, d, jx=-1.0, jy=-1.0, U=U, period_y=py)
    e0 = ed.ground_energy(om.L, d, N, om.bonds(), U)
    rmx = pm.run(rows, first_dim=fd, precision=prec, refine=True)
    assert abs(rmx["energies"][-1] - e0) < 1e-9 * abs(e0)


@pytest.mark.parametrize("prec", [dm.FP64, dm.BF16X3])
def test_pipelined_lanczos_is_bit_identical(monkeypatch, prec):
    # the Lanczos loop decides each iteration's stop while the device runs the
    # next one; the arithmetic is that of the blocking loop, bit for bit
    nx, ny, N, d, U, py, maxm, nhalf, fd = CASES[2]
    rows = sweep_rows(maxm, nhalf, niter=10, canpre=1e-8)
    pm = dm.SpinlessModel(nx, ny, N, d, hop_x=-1.0, hop_y=-1.0, U=U, period_y=py)
    fast = pm.run(rows, first_block=20, first_dim=fd, precision=prec)
    monkeypatch.setenv("DMRGCC_LANCZOS_BLOCKING", "1")
    slow = pm.run(rows, first_block=20, first_dim=fd, precision=prec)
    assert np.array_equal(fast["energies"], slow["energies"])
    assert np.array_equal(fast["trunc_err"], slow["trunc_err"])
    assert (fast["matvecs"], fast["rand_draws"]) == (slow["matvecs"], slow["rand_draws"])
