"""Grouped block GEMM (all three arithmetic forms) and the batched sector SVD
against numpy FP64, on ragged shapes that exercise every tile edge."""
import numpy as np
import pytest

import dmrgcc_amd as dm

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

SHAPES = [(1, 1, 1), (3, 5, 2), (16, 16, 4), (17, 33, 65), (64, 64, 64), (65, 1, 130),
          (100, 200, 37), (200, 150, 300), (7, 129, 9), (257, 3, 5)]

# relative (to |A||B| scale) tolerances per arithmetic form
TOL = {dm.FP64: 1e-13, dm.FP32: 1e-5, dm.BF16X3: 5e-5}


@pytest.mark.parametrize("tile", ["32", "64"])
@pytest.mark.parametrize("prec", [dm.FP64, dm.FP32, dm.BF16X3])
def test_block_gemm_matches_numpy(monkeypatch, prec, tile):
    # tile: 32x32 wave tiles or 64x64 workgroup tiles (DMRGCC_TILE)
    monkeypatch.setenv("DMRGCC_TILE", tile)
    rng = np.random.default_rng(5)
    As, Bs, ref = [], [], []
    for (M, N, K) in SHAPES:
        a = rng.standard_normal((M, K))
        b = rng.standard_normal((K, N))
        As.append(torch.tensor(a, device="cuda"))
        Bs.append(torch.tensor(b, device="cuda"))
        ref.append(a @ b)
    Cs = [torch.full((M, N), np.nan, dtype=torch.float64, device="cuda") for (M, N, K) in SHAPES]
    dm.block_gemm(As, Bs, Cs, prec)
    torch.cuda.synchronize()
    for (M, N, K), c, r in zip(SHAPES, Cs, ref):
        c = c.cpu().numpy()
        scale = np.sqrt(K)  # entries of A@B are O(sqrt(K))
        err = np.abs(c - r).max() / scale
        assert np.isfinite(c).all()
        assert err < TOL[prec], (M, N, K, err)


@pytest.mark.parametrize("tile", ["32", "64"])
def test_block_gemm_fp64_is_exact_on_integers(monkeypatch, tile):
    # small integers: every product and partial sum is exact in FP64, so the
    # MFMA result must equal numpy bit for bit (catches any lane/layout slip)
    monkeypatch.setenv("DMRGCC_TILE", tile)
    rng = np.random.default_rng(1)
    for (M, N, K) in SHAPES:
        a = rng.integers(-3, 4, (M, K)).astype(float)
        b = rng.integers(-3, 4, (K, N)).astype(float)
        c = torch.zeros((M, N), dtype=torch.float64, device="cuda")
        dm.block_gemm([torch.tensor(a, device="cuda")], [torch.tensor(b, device="cuda")], [c], dm.FP64)
        assert np.array_equal(c.cpu().numpy(), a @ b), (M, N, K)


@pytest.mark.parametrize("shape", [(1, 1), (1, 7), (7, 1), (5, 5), (40, 12), (12, 40), (120, 90),
                                   (300, 64), (33, 257)])
def test_sector_svd(shape):
    rng = np.random.default_rng(9)
    M, N = shape
    a = rng.standard_normal((M, N))
    # graded spectrum down to 1e-12 to check small singular values
    u, s, vt = np.linalg.svd(a, full_matrices=False)
    s = np.logspace(0, -12, len(s))
    a = (u * s) @ vt
    U, S, Vt = dm.sector_svd([torch.tensor(a, device="cuda")])
    U, S, Vt = U[0].cpu().numpy(), S[0].cpu().numpy(), Vt[0].cpu().numpy()
    k = min(M, N)
    assert np.all(np.diff(S) <= 0)
    assert np.abs(S - s).max() < 1e-14 * max(M, N)
    assert np.abs((U * S) @ Vt - a).max() < 1e-14 * max(M, N)
    assert np.abs(U.T @ U - np.eye(k)).max() < 1e-12
    assert np.abs(Vt @ Vt.T - np.eye(k)).max() < 1e-12


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("shapes", [[(120, 90)], [(200, 150), (150, 200), (64, 64), (90, 20)],
                                    [(600, 560)], [(700, 650), (650, 700)], [(1024, 300)],
                                    [(1100, 90)], [(1800, 150)], [(2048, 64)], [(2100, 100)],
                                    [(3000, 60)]])
def test_sector_svd_preconditioned(shapes, mode):
    # QR-preconditioned path: register panels of 32 columns (up to 576 rows),
    # 16 columns (up to 1024 rows) and 8 columns (up to 2048 rows), LDS panels
    # of 16 / 8 columns beyond those, and the global panel (M too large for
    # LDS); mixed with problems below the QR size
    rng = np.random.default_rng(11)
    mats, spec = [], []
    for (M, N) in shapes:
        a = rng.standard_normal((M, N))
        u, s, vt = np.linalg.svd(a, full_matrices=False)
        s = np.logspace(0, -12, len(s))
        mats.append((u * s) @ vt)
        spec.append(s)
    U, S, Vt = dm.sector_svd([torch.tensor(a, device="cuda") for a in mats], precondition=mode)
    for a, s, u, sg, vt in zip(mats, spec, U, S, Vt):
        u, sg, vt = u.cpu().numpy(), sg.cpu().numpy(), vt.cpu().numpy()
        M, N = a.shape
        k = min(M, N)
        assert np.all(np.diff(sg) <= 0)
        assert np.abs(sg - s).max() < 1e-14 * max(M, N)
        assert np.abs((u * sg) @ vt - a).max() < 1e-14 * max(M, N)
        big = sg > 1e-4
        if (M >= N) == (mode == 1) or k < 48:  # U exact, V for the well-resolved singular values
            assert np.abs(u.T @ u - np.eye(k)).max() < 1e-12
            vb = vt[big]
            assert np.abs(vb @ vb.T - np.eye(vb.shape[0])).max() < 1e-9
        else:
            assert np.abs(vt @ vt.T - np.eye(k)).max() < 1e-12
            ub = u[:, big]
            assert np.abs(ub.T @ ub - np.eye(ub.shape[1])).max() < 1e-9
        # same singular values and sign convention as the plain Jacobi
        U0, S0, Vt0 = dm.sector_svd([torch.tensor(a, device="cuda")])
        assert np.abs(S0[0].cpu().numpy() - sg).max() < 1e-14 * max(M, N)
        u0 = U0[0].cpu().numpy()
        ok = np.abs(u0[:, big] - u[:, big]).max()
        assert ok < 1e-8, ok
