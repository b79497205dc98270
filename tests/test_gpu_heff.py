"""H_eff matvec, environment updates and Lanczos on the device against the
oracle, on states taken from real oracle sweeps."""
import numpy as np
import pytest

import dmrgcc_amd as dm
from oracle.dmrg import heff_apply, heff_dense, heff_flops, lanczos, left_env, right_env
from tests._support import env_blocks, oracle_state, site_description

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

STATES = {
    # (nx, ny, N, d, maxm, half, site)
    "3x2_d3": (3, 2, 6, 3, 30, 2, 2),
    "4x3_d4_left": (4, 3, 12, 4, 120, 2, 5),
    "4x3_d4_right": (4, 3, 12, 4, 120, 3, 7),
}
_cache = {}


def state(name):
    if name not in _cache:
        nx, ny, N, d, maxm, half, site = STATES[name]
        _cache[name] = (oracle_state(nx, ny, N, d, maxm=maxm, half=half, site=site, niter=6), site)
    return _cache[name]


def make_plan(dmo, site):
    lat = dmo.space.lat[site]
    phys, kl, kr, krows, kcols = site_description(lat)
    (Lo, Ll, Lr, Lrows, Lcols, Loff), Ld = env_blocks(dmo.Lenv[site])
    (Ro, Rl, Rr, Rrows, Rcols, Roff), Rd = env_blocks(dmo.Renv[site])
    Ldev = torch.tensor(Ld if Ld.size else np.zeros(1), device="cuda")
    Rdev = torch.tensor(Rd if Rd.size else np.zeros(1), device="cuda")
    plan = dm.HeffPlan(dmo.W[site], phys, kl, kr, krows, kcols,
                       dm.BlockList(Lo, Ll, Lr, Lrows, Lcols, Loff), Ldev,
                       dm.BlockList(Ro, Rl, Rr, Rrows, Rcols, Roff), Rdev)
    plan._dev = (Ldev, Rdev)
    return plan, lat


@pytest.mark.parametrize("name", list(STATES))
def test_heff_fp64_matches_oracle(name):
    dmo, site = state(name)
    plan, lat = make_plan(dmo, site)
    n = lat.ket.dim()
    assert plan.info()["psi_dim"] == n
    rng = np.random.default_rng(2)
    x = rng.standard_normal(n)
    y0 = rng.standard_normal(n)
    ref = heff_apply(lat, dmo.W[site], dmo.Lenv[site], dmo.Renv[site], x, y0.copy())
    xd = torch.tensor(x, device="cuda")
    yd = torch.tensor(y0, device="cuda")
    plan.apply(xd, yd, dm.FP64)          # accumulate: y0 + H x
    got = yd.cpu().numpy()
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()
    info = plan.info()
    assert info["flops_ref"] == heff_flops(lat, dmo.W[site], dmo.Lenv[site], dmo.Renv[site])


@pytest.mark.parametrize("tile", ["32", "64"])
@pytest.mark.parametrize("split_k", [0, 1, 40])
@pytest.mark.parametrize("prec", [dm.FP64, dm.BF16X3])
def test_heff_split_k(monkeypatch, split_k, prec, tile):
    # split-K cuts heavy output blocks into contribution chunks that are summed
    # by a second kernel; any threshold must give the unsplit result, with
    # 32x32 wave tiles or 64x64 workgroup tiles
    dmo, site = state("4x3_d4_left")
    monkeypatch.setenv("DMRGCC_SPLIT_K", str(split_k))
    monkeypatch.setenv("DMRGCC_TILE", tile)
    plan, lat = make_plan(dmo, site)
    n = lat.ket.dim()
    x = np.random.default_rng(6).standard_normal(n)
    y0 = np.random.default_rng(7).standard_normal(n)
    ref = heff_apply(lat, dmo.W[site], dmo.Lenv[site], dmo.Renv[site], x, y0.copy())
    yd = torch.tensor(y0, device="cuda")
    plan.apply(torch.tensor(x, device="cuda"), yd, prec)
    err = np.abs(yd.cpu().numpy() - ref).max() / np.abs(ref).max()
    assert err < (1e-12 if prec == dm.FP64 else 2e-4), err


@pytest.mark.parametrize("tile", ["32", "64"])
@pytest.mark.parametrize("prec", [dm.FP64, dm.FP32, dm.BF16X3])
def test_heff_split_k_repeats_bit_identical(monkeypatch, prec, tile):
    # heavy output blocks cut into split-K chunks: back-to-back matvecs, and
    # graph replays with eager matvecs between them, give the same bits
    dmo, site = state("4x3_d4_left")
    monkeypatch.setenv("DMRGCC_SPLIT_K", "40")
    monkeypatch.setenv("DMRGCC_TILE", tile)
    plan, lat = make_plan(dmo, site)
    n = lat.ket.dim()
    x = torch.tensor(np.random.default_rng(9).standard_normal(n), device="cuda")
    outs = []
    for _ in range(4):
        y = torch.zeros(n, dtype=torch.float64, device="cuda")
        plan.apply(x, y, prec)
        outs.append(y.cpu().numpy())
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    yg = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=s):
        plan.apply_async(x, yg, 1.0, 0.0, prec, stream=s)
    for rep in range(3):
        ye = torch.zeros(n, dtype=torch.float64, device="cuda")
        plan.apply(x, ye, prec)   # eager launches between the replays
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(yg.cpu().numpy(), outs[0]), rep
        assert np.array_equal(ye.cpu().numpy(), outs[0]), rep
    ref = heff_apply(lat, dmo.W[site], dmo.Lenv[site], dmo.Renv[site], x.cpu().numpy(), np.zeros(n))
    tol = {dm.FP64: 1e-12, dm.FP32: 1e-5, dm.BF16X3: 2e-4}[prec]
    assert np.abs(outs[0] - ref).max() < tol * np.abs(ref).max()


@pytest.mark.parametrize("prec", [dm.FP64, dm.BF16X3])
def test_heff_tune_keeps_results(prec):
    # the autotuner re-tiles the plan with every candidate and keeps the
    # fastest; the tuned plan still gives the oracle's matvec, and the chosen
    # tiling can be made the default of later plans
    dmo, site = state("4x3_d4_left")
    plan, lat = make_plan(dmo, site)
    t = plan.tune(prec, reps=3)
    cands = t["candidates"]
    assert len(cands) >= 4 and {c[0] for c in cands} == {32, 64}
    best = min(cands, key=lambda c: c[2])
    dflt = [c for c in cands if c[:2] == (64, 256)][0]
    if best[2] > 0.97 * dflt[2]:
        best = dflt
    assert t["best"] == (best[0], best[1]) == plan.config()
    n = lat.ket.dim()
    x = np.random.default_rng(8).standard_normal(n)
    ref = heff_apply(lat, dmo.W[site], dmo.Lenv[site], dmo.Renv[site], x, np.zeros(n))
    for tile, sk, _ in cands[:2] + [best]:
        dm.gemm_config(tile, sk)
        try:
            p2, _ = make_plan(dmo, site)
            assert p2.config() == (tile, sk)
        finally:
            dm.gemm_config(-1, -1)
        yd = torch.zeros(n, dtype=torch.float64, device="cuda")
        p2.apply(torch.tensor(x, device="cuda"), yd, prec)
        err = np.abs(yd.cpu().numpy() - ref).max() / np.abs(ref).max()
        assert err < (1e-12 if prec == dm.FP64 else 2e-4), (tile, sk, err)
    yd = torch.zeros(n, dtype=torch.float64, device="cuda")
    plan.apply(torch.tensor(x, device="cuda"), yd, prec)
    err = np.abs(yd.cpu().numpy() - ref).max() / np.abs(ref).max()
    assert err < (1e-12 if prec == dm.FP64 else 2e-4), err
    p3, _ = make_plan(dmo, site)
    assert p3.config() == (64, 256)  # defaults restored


@pytest.mark.parametrize("prec,tol", [(dm.FP32, 1e-5), (dm.BF16X3, 2e-4)])
def test_heff_mixed_within_tolerance(prec, tol):
    dmo, site = state("4x3_d4_left")
    plan, lat = make_plan(dmo, site)
    n = lat.ket.dim()
    x = np.random.default_rng(4).standard_normal(n)
    ref = heff_apply(lat, dmo.W[site], dmo.Lenv[site], dmo.Renv[site], x, np.zeros(n))
    yd = torch.zeros(n, dtype=torch.float64, device="cuda")
    plan.apply(torch.tensor(x, device="cuda"), yd, prec)
    err = np.abs(yd.cpu().numpy() - ref).max() / np.abs(ref).max()
    assert err < tol, err


@pytest.mark.parametrize("split_k", ["0", "40"])
def test_heff_presplit_matches_unpacked_bf16x3(monkeypatch, split_k):
    # the pre-split BF16x3 stages (operands packed once into bf16 pair
    # groups, T^T written as pairs) against the on-the-fly split: the same
    # products, each term scaled in FP64 instead of before the split
    dmo, site = state("4x3_d4_left")
    monkeypatch.setenv("DMRGCC_SPLIT_K", split_k)
    monkeypatch.setenv("DMRGCC_TILE", "64")
    pre, lat = make_plan(dmo, site)
    monkeypatch.setenv("DMRGCC_PRESPLIT", "0")
    old, _ = make_plan(dmo, site)
    n = lat.ket.dim()
    x = np.random.default_rng(11).standard_normal(n)
    y0 = np.random.default_rng(12).standard_normal(n)
    ref = heff_apply(lat, dmo.W[site], dmo.Lenv[site], dmo.Renv[site], x, y0.copy())
    got = {}
    for key, plan in (("pre", pre), ("old", old)):
        yd = torch.tensor(y0, device="cuda")
        plan.apply(torch.tensor(x, device="cuda"), yd, dm.BF16X3)
        got[key] = yd.cpu().numpy()
        err = np.abs(got[key] - ref).max() / np.abs(ref).max()
        assert err < 2e-4, (key, err)
    diff = np.abs(got["pre"] - got["old"]).max() / np.abs(ref).max()
    assert diff < 2e-5, diff
    # the FP64 and FP32 forms of the pre-split plan keep the unpacked T
    for prec, tol in ((dm.FP64, 1e-12), (dm.FP32, 1e-5)):
        yd = torch.tensor(y0, device="cuda")
        pre.apply(torch.tensor(x, device="cuda"), yd, prec)
        assert np.abs(yd.cpu().numpy() - ref).max() < tol * np.abs(ref).max()


def test_heff_is_symmetric_and_equals_dense():
    dmo, site = state("3x2_d3")
    plan, lat = make_plan(dmo, site)
    H = heff_dense(lat, dmo.W[site], dmo.Lenv[site], dmo.Renv[site])
    n = H.shape[0]
    X = torch.eye(n, dtype=torch.float64, device="cuda")
    cols = []
    for c in range(n):
        y = torch.zeros(n, dtype=torch.float64, device="cuda")
        plan.apply(X[c].contiguous(), y, dm.FP64)
        cols.append(y.cpu().numpy())
    Hg = np.stack(cols, axis=1)
    assert np.abs(Hg - H).max() < 1e-12 * np.abs(H).max()
    assert np.abs(Hg - Hg.T).max() < 1e-12 * np.abs(H).max()


@pytest.mark.parametrize("name", list(STATES))
@pytest.mark.parametrize("leigh", [0, 1])
def test_env_update_matches_oracle(name, leigh):
    dmo, site = state(name)
    lat = dmo.space.lat[site]
    phys, kl, kr, krows, kcols = site_description(lat)
    env = dmo.Lenv[site] if leigh == 0 else dmo.Renv[site]
    (o, l, r, rows, cols, off), data = env_blocks(env)
    ket = torch.tensor(lat.ket.to_vector(), device="cuda")
    E = torch.tensor(data if data.size else np.zeros(1), device="cuda")
    res = dm.env_update(leigh, dmo.W[site], phys, kl, kr, krows, kcols, ket,
                        dm.BlockList(o, l, r, rows, cols, off), E)
    ref = (left_env if leigh == 0 else right_env)(lat, dmo.W[site], env)
    host = res.to_host()
    got = {}
    for (ch, bl, br, nr_, nc, of) in res.blocks:
        got[(ch, bl, br)] = host[of:of + nr_ * nc].reshape(nr_, nc)
    nref = sum(b.n for b in ref)
    assert len(res.blocks) == nref
    for ch, bm in enumerate(ref):
        for k in range(bm.n):
            m = bm.mats[k]
            g = got[(ch, bm.left[k], bm.right[k])]
            assert g.shape == m.shape
            if m.size:
                assert np.abs(g - m).max() <= 1e-12 * max(1.0, np.abs(m).max())


@pytest.mark.parametrize("name", ["4x3_d4_left", "4x3_d4_right"])
def test_lanczos_matches_oracle(name):
    dmo, site = state(name)
    plan, lat = make_plan(dmo, site)
    ket = torch.tensor(lat.ket.to_vector(), device="cuda")
    lat2 = lat.copy()
    e_ref = lanczos(lat2, dmo.W[site], dmo.Lenv[site], dmo.Renv[site], 10)
    e, iters = plan.lanczos(ket, 10, dm.FP64)
    assert abs(e - e_ref) < 1e-10 * abs(e_ref)
    v = ket.cpu().numpy()
    vr = lat2.ket.to_vector()
    assert abs(abs(np.dot(v, vr)) - 1.0) < 1e-8
    # mixed precision with FP64 Rayleigh-quotient refinement
    ket2 = torch.tensor(lat.ket.to_vector(), device="cuda")
    e2, _ = plan.lanczos(ket2, 10, dm.BF16X3, refine=True)
    assert abs(e2 - e_ref) < 1e-8 * abs(e_ref)


def test_apply_async_alpha_beta_and_graph_capture():
    # y = alpha H x + beta y on a given stream without host waits; captured
    # into a CUDA graph and replayed with eager work between the replays
    dmo, site = state("4x3_d4_left")
    plan, lat = make_plan(dmo, site)
    n = lat.ket.dim()
    rng = np.random.default_rng(12)
    x = rng.standard_normal(n)
    y0 = rng.standard_normal(n)
    hx = heff_apply(lat, dmo.W[site], dmo.Lenv[site], dmo.Renv[site], x, np.zeros(n))
    xd = torch.tensor(x, device="cuda")
    for alpha, beta in ((1.0, 1.0), (0.5, -2.0), (1.0, 0.0), (0.0, 3.0), (-1.5, 1.0)):
        yd = torch.tensor(y0, device="cuda")
        plan.apply_async(xd, yd, alpha, beta, dm.FP64)
        torch.cuda.synchronize()
        ref = alpha * hx + beta * y0
        assert np.abs(yd.cpu().numpy() - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()), (alpha, beta)
    # graph: y = 2 H x (beta = 0 never reads y)
    yg = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=s):
        plan.apply_async(xd, yg, 2.0, 0.0, dm.FP64, stream=s)
    for rep in range(3):
        xd.mul_(-0.5).add_(0.25)   # eager work between the replays
        ye = torch.zeros(n, dtype=torch.float64, device="cuda")
        plan.apply(xd, ye, dm.FP64)
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(yg, 2.0 * ye), rep
