"""Batched H_eff of independent instances: one grouped launch per stage for
plans of different models, sizes and arithmetic kinds (real and complex),
bit-identical to applying the plans one by one; the HIP-graph form replayed
several times with eager work in between."""
import numpy as np
import pytest
import torch

import dmrgcc_amd as dm
from tests._support import sweep_rows

pytestmark = pytest.mark.gpu


def _instances():
    runs = [dm.SpinlessModel(3, 2, 6, 3, U=2.0).start(sweep_rows(30, 4, niter=6), first_dim=4),
            dm.SpinlessModel(6, 1, 6, 4, U=5.0).start(sweep_rows(40, 4, niter=6), first_dim=4),
            dm.SpinfulModel(2, 2, 2, 2, 3, flux=0.25, U_uu=4.0, U_dd=4.0, U_ud=2.0)
              .start(sweep_rows(30, 4, niter=6), first_dim=4)]
    plans = []
    for run in runs:
        assert run.advance(2, 2)
        plans.append(run.heff(2))
    return runs, plans


@pytest.mark.parametrize("prec", [dm.FP64, dm.BF16X3])
def test_batch_equals_plan_by_plan(monkeypatch, prec):
    # the batch runs the on-the-fly split BF16x3 form: the plans are built
    # without the pre-split stages to compare bit for bit
    monkeypatch.setenv("DMRGCC_PRESPLIT", "0")
    runs, plans = _instances()
    rng = np.random.default_rng(11)
    xs, ys, refs = [], [], []
    for p in plans:
        n = p.info()["psi_dim"] * (2 if p.complex else 1)
        x = torch.tensor(rng.standard_normal(n), device="cuda")
        y0 = torch.tensor(rng.standard_normal(n), device="cuda")
        yr = y0.clone()
        p.apply(x, yr, prec)
        xs.append(x)
        ys.append(y0)
        refs.append(yr.cpu().numpy())
    y_init = [y.clone() for y in ys]
    b = dm.HeffBatch(plans, xs, ys)
    assert b.info()["plans"] == 3
    b.apply(prec, graph=False)
    for y, r in zip(ys, refs):
        assert np.array_equal(y.cpu().numpy(), r)
    # graph replays: reset y eagerly, replay, compare, several times
    for rep in range(3):
        for y, y0 in zip(ys, y_init):
            y.copy_(y0)
        torch.cuda.synchronize()
        b.apply(prec, graph=True)
        for y, r in zip(ys, refs):
            assert np.array_equal(y.cpu().numpy(), r), rep


def test_batch_refuses_retiled_plan():
    # re-tiling a plan replaces the descriptors and buffers a batch built from
    # it points at: the batch then refuses to launch, graph or not
    runs, plans = _instances()
    xs, ys = [], []
    for p in plans:
        n = p.info()["psi_dim"] * (2 if p.complex else 1)
        xs.append(torch.zeros(n, dtype=torch.float64, device="cuda"))
        ys.append(torch.zeros(n, dtype=torch.float64, device="cuda"))
    b = dm.HeffBatch(plans, xs, ys)
    b.apply(dm.FP64, graph=False)
    plans[1].tune(dm.FP64, reps=1)
    for graph in (False, True):
        with pytest.raises(dm.DmrgError, match="re-tiled"):
            b.apply(dm.FP64, graph=graph)
    b2 = dm.HeffBatch(plans, xs, ys)  # a new batch is fine
    b2.apply(dm.FP64, graph=True)
    torch.cuda.synchronize()


def test_batch_refuses_repeated_plan_and_stale_plans():
    runs, plans = _instances()
    n = plans[0].info()["psi_dim"]
    x = [torch.zeros(n, dtype=torch.float64, device="cuda") for _ in range(2)]
    y = [torch.zeros(n, dtype=torch.float64, device="cuda") for _ in range(2)]
    with pytest.raises(dm.DmrgError, match="more than once"):
        dm.HeffBatch([plans[0], plans[0]], x, y)
    # advancing the run that built a plan makes the plan stale: every call
    # that would read its buffers is refused, info() still answers
    b = dm.HeffBatch([plans[0]], x[:1], y[:1])
    b.apply(dm.FP64, graph=False)
    assert runs[0].advance(3, 2)
    with pytest.raises(dm.DmrgError, match="stale"):
        plans[0].apply(x[0], y[0], dm.FP64)
    with pytest.raises(dm.DmrgError, match="stale"):
        plans[0].lanczos(x[0], 4)
    with pytest.raises(dm.DmrgError, match="stale"):
        b.apply(dm.FP64, graph=True)
    assert plans[0].info()["psi_dim"] == n
    fresh = runs[0].heff(2)
    assert fresh.info()["psi_dim"] > 0
    torch.cuda.synchronize()
