"""Two-site effective Hamiltonian (L_j x W_j W_j+1 x Theta x R_j+1) against
an independent numpy contraction over the exported blocks, real and complex;
its Lanczos energy against the one-site energy of the same state and the
exact ground energy."""
import numpy as np
import pytest
import torch

import dmrgcc_amd as dm
from oracle import ed
from oracle.model import SpinfulBoseSquare, SpinlessBoseSquare
from tests._support import sweep_rows

pytestmark = pytest.mark.gpu


def _blocks(run, site, which, cplx):
    blocks, data = run.site_blocks(site, which)
    out = []
    for (o, l, r, rows, cols, off) in blocks:
        out.append((o, l, r, data[off:off + rows * cols].reshape(rows, cols)))
    return out


def _two_site_reference(run, model, j, x_blocks):
    """y = H_eff2 x on the block list of Theta (numpy, block by block)."""
    cplx = model.complex
    A = _blocks(run, j, 0, cplx)
    B = _blocks(run, j + 1, 0, cplx)
    Lb = _blocks(run, j, 1, cplx)
    Rb = _blocks(run, j + 1, 2, cplx)
    PA, PB = run.phys(j), run.phys(j + 1)
    Wj, Wk = model.mpo_site(j), model.mpo_site(j + 1)
    d = model.d
    # Theta block list in the library's order: A blocks, then B blocks sharing m
    tb = []
    for (_, l, m, a) in A:
        for (_, m2, r, b) in B:
            if m2 == m:
                tb.append((l, m, r, a @ b))
    if not x_blocks:
        return tb, []
    L = {(o, l1, l2): m for (o, l1, l2, m) in Lb}
    R = {(o, r1, r2): m for (o, r1, r2, m) in Rb}
    ys = [np.zeros_like(x) for x in x_blocks]
    for bi, (l2, m2, r2, _) in enumerate(tb):       # output block
        p1o, p2o = PA[l2][m2], PB[m2][r2]
        for bj, (l, m, r, _) in enumerate(tb):     # input block
            p1, p2 = PA[l][m], PB[m][r]
            w = np.einsum('ab,bc->ac', Wj[:, :, p1o, p1], Wk[:, :, p2o, p2])  # (o1, o3)
            for o1, o3 in zip(*np.nonzero(w)):
                lm = L.get((o1, l2, l))
                rm = R.get((o3, r, r2))
                if lm is None or rm is None:
                    continue
                ys[bi] = ys[bi] + w[o1, o3] * (lm @ x_blocks[bj] @ rm)
    return tb, ys


CASES = [("real", lambda: dm.SpinlessModel(3, 2, 6, 3, U=2.0),
          SpinlessBoseSquare(3, 2, 6, 3, U=2.0)),
         ("spinful", lambda: dm.SpinfulModel(2, 2, 2, 2, 3, flux=0.25, U_uu=4.0, U_dd=4.0, U_ud=2.0),
          SpinfulBoseSquare(2, 2, 2, 2, 3, flux=0.25, U_uu=4.0, U_dd=4.0, U_ud=2.0))]


@pytest.mark.parametrize("name,mk,om", CASES)
def test_two_site_heff_matches_numpy(name, mk, om):
    model = mk()
    run = model.start(sweep_rows(30, 6, niter=8, canpre=1e-10), first_block=20, first_dim=4)
    j = 2
    assert run.advance(2, j)
    plan, theta = run.heff2(j)
    cplx = model.complex
    th = theta.cpu().numpy()
    if cplx:
        th = th.view(np.complex128)
    # Theta = A B
    tb, _ = _two_site_reference(run, model, j, [])
    flat = np.concatenate([t[3].ravel() for t in tb])
    assert plan.info()["psi_dim"] == flat.size
    assert np.abs(th - flat).max() < 1e-13
    # matvec on a random vector, y = y0 + H x
    rng = np.random.default_rng(3)
    x = rng.standard_normal(flat.size) + (1j * rng.standard_normal(flat.size) if cplx else 0)
    y0 = rng.standard_normal(flat.size) + (1j * rng.standard_normal(flat.size) if cplx else 0)
    xb, off = [], 0
    for t in tb:
        xb.append(x[off:off + t[3].size].reshape(t[3].shape))
        off += t[3].size
    _, ys = _two_site_reference(run, model, j, xb)
    ref = y0 + np.concatenate([y.ravel() for y in ys])
    for prec, tol in ((dm.FP64, 1e-12), (dm.FP32, 1e-5), (dm.BF16X3, 2e-4)):
        xd = torch.tensor(x.view(np.float64) if cplx else x, device="cuda")
        yd = torch.tensor(y0.view(np.float64) if cplx else y0, device="cuda")
        plan.apply(xd, yd, prec)
        got = yd.cpu().numpy()
        got = got.view(np.complex128) if cplx else got
        assert np.abs(got - ref).max() < tol * np.abs(ref).max(), prec


@pytest.mark.parametrize("name,mk,om", CASES)
def test_two_site_lanczos_energy(name, mk, om):
    model = mk()
    run = model.start(sweep_rows(60, 6, niter=10, canpre=1e-12), first_block=20, first_dim=4)
    j = 2
    assert run.advance(4, j)
    e1, _ = run.heff(j).lanczos(torch.tensor(_ket(run, j, model.complex), device="cuda"), 20)
    plan, theta = run.heff2(j)
    e2, _ = plan.lanczos(theta.clone(), 20)
    e0 = ed.spinful_ground_energy(om) if name == "spinful" else ed.ground_energy(om.L, 3, 6, om.bonds(), 2.0)
    # the two-site space contains the one-site one; neither beats the exact energy
    assert e2 <= e1 + 1e-10 * abs(e1)
    assert e2 >= e0 - 1e-9 * abs(e0)
    assert abs(e2 - e0) < 1e-8 * abs(e0)


def _ket(run, site, cplx):
    _, data = run.site_blocks(site, 0)
    return data.view(np.float64).copy() if cplx else data.copy()


@pytest.mark.parametrize("name,mk,om", CASES)
def test_two_site_sweeps_reach_exact_energy(name, mk, om):
    model = mk()
    e0 = ed.spinful_ground_energy(om) if name == "spinful" else ed.ground_energy(om.L, 3, 6, om.bonds(), 2.0)
    rows = sweep_rows(100, 6, niter=10, canpre=1e-12)
    r = model.run(rows, first_block=20, first_dim=4, two_site=True)
    assert abs(r["energies"][-1] - e0) < 1e-9 * abs(e0)
    assert min(r["energies"]) > e0 - 1e-9 * abs(e0)
    # mixed-precision H_eff with the FP64 refinement reaches it too
    rm = mk().run(rows, first_block=20, first_dim=4, two_site=True, precision=dm.BF16X3)
    assert abs(rm["energies"][-1] - e0) < 1e-9 * abs(e0)


def test_two_site_truncated_sweeps_are_variational():
    om = SpinlessBoseSquare(4, 2, 8, 3, U=2.0)
    e0 = ed.ground_energy(om.L, 3, 8, om.bonds(), 2.0)
    r = dm.SpinlessModel(4, 2, 8, 3, U=2.0).run(sweep_rows(8, 8, niter=10, canpre=1e-12), first_block=20,
                                                 first_dim=4, two_site=True)
    e = np.array(r["energies"])
    assert (e > e0 - 1e-10).all()
    assert max(r["trunc_err"][2:]) > 0          # m = 8 truncates
    assert e[-1] <= e[1] + 1e-10                # sweeps do not lose energy
    assert abs(e[-1] - e0) < 1e-2 * abs(e0)
