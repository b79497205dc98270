"""Complex sweeps (ComplexTensorRundmrg) on the device against the oracle's
complex sweeps on the same inputs (same seeded rand() stream, two draws per
complex entry) and against exact diagonalisation: the complex spinless model
(the real Hamiltonian on complex storage) and two-component bosons with flux.

As for the real sweeps, a truncation cut-off that FP64 singular values
resolve (canpre 1e-8) reproduces the trajectory: every rand() draw, every
truncation and the per-half-sweep energies."""
import numpy as np
import pytest

import dmrgcc_amd as dm
from oracle import ed
from oracle.dmrg import DMRG
from oracle.model import SpinfulBoseSquare, SpinlessBoseSquare
from tests._support import sweep_rows

pytestmark = pytest.mark.gpu

SPINFUL = [
    # nx, ny, n_up, n_down, d, flux, (U_uu, U_dd, U_ud), period_y, maxm, nhalf, first_dim
    (2, 2, 2, 2, 3, 0.25, (4.0, 4.0, 2.0), False, 100, 6, 4),
    (3, 2, 2, 3, 3, 0.2, (1.0, 2.0, 1.5), False, 60, 6, 4),
]


def _spinful_pair(case):
    nx, ny, nu, nd, d, flux, (uu, udd, ud), py, maxm, nhalf, fd = case
    om = SpinfulBoseSquare(nx, ny, nu, nd, d, flux=flux, U_uu=uu, U_dd=udd, U_ud=ud, period_y=py)
    pm = dm.SpinfulModel(nx, ny, nu, nd, d, flux=flux, U_uu=uu, U_dd=udd, U_ud=ud, period_y=py)
    return om, pm


@pytest.mark.parametrize("case", SPINFUL)
def test_spinful_trajectory_matches_oracle(case):
    *_, maxm, nhalf, fd = case
    om, pm = _spinful_pair(case)
    assert pm.complex and pm.nq == 2
    rows = sweep_rows(maxm, nhalf, niter=10, canpre=1e-8)
    o = DMRG(om, rows, nhalf // 2, first_block=20, first_dim=fd)
    e_ref = np.array(o.run())
    r = pm.run(rows, first_block=20, first_dim=fd, precision=dm.FP64)
    assert r["rand_draws"] == o.rng.count
    assert np.abs(r["energies"] - e_ref).max() < 1e-10 * np.abs(e_ref).max()
    assert np.allclose(r["trunc_err"], o.trunc_err, rtol=1e-6, atol=1e-14)


def test_spinful_reaches_exact_ground_energy():
    case = SPINFUL[0]
    *_, maxm, nhalf, fd = case
    om, pm = _spinful_pair(case)
    e0 = ed.spinful_ground_energy(om)
    r = pm.run(sweep_rows(maxm, nhalf, niter=10, canpre=1e-12), first_block=20, first_dim=fd)
    assert abs(r["energies"][-1] - e0) < 1e-9 * abs(e0)
    assert min(r["energies"]) > e0 - 1e-9 * abs(e0)


@pytest.mark.parametrize("prec", [dm.FP32, dm.BF16X3])
def test_spinful_mixed_precision_energy(prec):
    case = SPINFUL[0]
    *_, maxm, nhalf, fd = case
    om, pm = _spinful_pair(case)
    e0 = ed.spinful_ground_energy(om)
    r = pm.run(sweep_rows(maxm, nhalf, niter=10, canpre=1e-12), first_block=20, first_dim=fd,
               precision=prec, refine=True)
    assert abs(r["energies"][-1] - e0) < 1e-9 * abs(e0)


def test_complex_spinless_matches_oracle_and_real_ground_state():
    nx, ny, N, d, U = 3, 2, 6, 3, 2.0
    rows = sweep_rows(60, 6, niter=10, canpre=1e-8)
    om = SpinlessBoseSquare(nx, ny, N, d, U=U, complex=True)
    o = DMRG(om, rows, 3, first_block=20, first_dim=4)
    e_ref = np.array(o.run())
    pm = dm.SpinlessModel(nx, ny, N, d, U=U, complex=True)
    r = pm.run(rows, first_block=20, first_dim=4)
    assert r["rand_draws"] == o.rng.count
    assert np.abs(r["energies"] - e_ref).max() < 1e-10 * np.abs(e_ref).max()
    e0 = ed.ground_energy(om.L, d, N, om.bonds(), U)
    rr = dm.SpinlessModel(nx, ny, N, d, U=U).run(rows, first_block=20, first_dim=4)
    assert abs(r["energies"][-1] - e0) < 1e-8 * abs(e0)
    assert abs(rr["energies"][-1] - e0) < 1e-8 * abs(e0)


def test_complex_run_exports_complex_blocks():
    case = SPINFUL[0]
    *_, maxm, nhalf, fd = case
    om, pm = _spinful_pair(case)
    run = pm.start(sweep_rows(maxm, nhalf, niter=10, canpre=1e-8), first_block=20, first_dim=fd)
    assert run.advance(2, 3)
    blocks, data = run.site_blocks(3, 0)
    assert data.dtype == np.complex128
    assert abs(np.vdot(data, data).real - 1.0) < 1e-12
    h = run.heff(3)
    assert h.complex
