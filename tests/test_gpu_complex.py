"""Complex (complex128) twins of the hot path on the device against the
oracle: block GEMM (ComplexMatrix::MultiplyToMatrix), H_eff matvec
(ComplexTensorContraction::MultiplyEffectHamilton), environment updates with
the Hermitian-conjugate bra, and complex Lanczos.

Hermitian complex operators are made from real oracle states by a phase gauge
on the site's physical index, W'(p, q) = exp(-i t_p) W(p, q) exp(i t_q): H'
is unitarily equivalent to the real H, so its spectrum is known exactly."""
import copy

import numpy as np
import pytest

import dmrgcc_amd as dm
from oracle.blocks import BlockMat
from oracle.dmrg import heff_apply, heff_dense, heff_flops, lanczos, left_env, right_env
from tests._support import env_blocks, oracle_state, site_description

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

SHAPES = [(1, 1, 1), (3, 5, 2), (17, 33, 65), (64, 64, 64), (100, 200, 37), (7, 129, 9)]
TOL = {dm.FP64: 1e-13, dm.FP32: 1e-5, dm.BF16X3: 5e-5}
_cache = {}


def state(name):
    if name not in _cache:
        if name == "3x2_d3":
            _cache[name] = (oracle_state(3, 2, 6, 3, maxm=30, half=2, site=2, niter=6), 2)
        else:
            _cache[name] = (oracle_state(4, 3, 12, 4, maxm=120, half=2, site=5, niter=6), 5)
    return _cache[name]


def gauge_W(W, seed):
    d = W.shape[2]
    t = np.random.default_rng(seed).uniform(0, 2 * np.pi, d)
    ph = np.exp(1j * t)
    return W * np.conj(ph)[None, None, :, None] * ph[None, None, None, :]


def complex_env(env, rng, amount):
    out = []
    for bm in env:
        mats = [m + amount * 1j * rng.standard_normal(m.shape) for m in bm.mats]
        out.append(BlockMat(list(bm.left), list(bm.right), mats))
    return out


def cdev(a):
    a = np.asarray(a, dtype=np.complex128)
    return torch.tensor(a if a.size else np.zeros(1, dtype=np.complex128), device="cuda")


def zplan(W, lat, Lenv, Renv):
    phys, kl, kr, krows, kcols = site_description(lat)
    (Lo, Ll, Lr, Lrows, Lcols, Loff), Ld = env_blocks(Lenv)
    (Ro, Rl, Rr, Rrows, Rcols, Roff), Rd = env_blocks(Renv)
    Ldev, Rdev = cdev(Ld), cdev(Rd)
    plan = dm.HeffPlan(W, phys, kl, kr, krows, kcols, dm.BlockList(Lo, Ll, Lr, Lrows, Lcols, Loff), Ldev,
                       dm.BlockList(Ro, Rl, Rr, Rrows, Rcols, Roff), Rdev)
    plan._dev = (Ldev, Rdev)
    assert plan.complex
    return plan


@pytest.mark.parametrize("prec", [dm.FP64, dm.FP32, dm.BF16X3])
def test_zblock_gemm_matches_numpy(prec):
    rng = np.random.default_rng(8)
    As, Bs, Cs, ref = [], [], [], []
    for (M, N, K) in SHAPES:
        a = rng.standard_normal((M, K)) + 1j * rng.standard_normal((M, K))
        b = rng.standard_normal((K, N)) + 1j * rng.standard_normal((K, N))
        As.append(cdev(a))
        Bs.append(cdev(b))
        Cs.append(torch.full((M, N), complex(np.nan, np.nan), dtype=torch.complex128, device="cuda"))
        ref.append(a @ b)
    dm.block_gemm(As, Bs, Cs, prec)
    for (M, N, K), c, r in zip(SHAPES, Cs, ref):
        c = c.cpu().numpy()
        assert np.isfinite(c).all()
        assert np.abs(c - r).max() / np.sqrt(2 * K) < TOL[prec], (M, N, K)


def test_zblock_gemm_fp64_exact_on_gaussian_integers():
    rng = np.random.default_rng(2)
    for (M, N, K) in SHAPES:
        a = rng.integers(-3, 4, (M, K)) + 1j * rng.integers(-3, 4, (M, K))
        b = rng.integers(-3, 4, (K, N)) + 1j * rng.integers(-3, 4, (K, N))
        c = torch.zeros((M, N), dtype=torch.complex128, device="cuda")
        dm.block_gemm([cdev(a)], [cdev(b)], [c], dm.FP64)
        assert np.array_equal(c.cpu().numpy(), a @ b), (M, N, K)


@pytest.mark.parametrize("name", ["3x2_d3", "4x3_d4"])
def test_zheff_matches_oracle(name):
    # fully complex data (non-Hermitian is fine for a contraction)
    dmo, site = state(name)
    rng = np.random.default_rng(5)
    lat = dmo.space.lat[site]
    W = gauge_W(dmo.W[site], 1) * (1 + 0.2j)
    Lc = complex_env(dmo.Lenv[site], rng, 0.3)
    Rc = complex_env(dmo.Renv[site], rng, 0.3)
    plan = zplan(W, lat, Lc, Rc)
    n = lat.ket.dim()
    info = plan.info()
    assert info["psi_dim"] == n
    assert info["flops_ref"] == heff_flops(lat, W, Lc, Rc)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y0 = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = heff_apply(lat, W, Lc, Rc, x, y0.copy())
    for prec, tol in ((dm.FP64, 1e-12), (dm.FP32, 1e-5), (dm.BF16X3, 2e-4)):
        yd = cdev(y0)
        plan.apply(cdev(x), yd, prec)
        err = np.abs(yd.cpu().numpy() - ref).max() / np.abs(ref).max()
        assert err < tol, (prec, err)


def test_zheff_gauge_is_hermitian_and_unitarily_equivalent():
    dmo, site = state("3x2_d3")
    lat = dmo.space.lat[site]
    W = gauge_W(dmo.W[site], 3)
    plan = zplan(W, lat, dmo.Lenv[site], dmo.Renv[site])
    H = heff_dense(lat, W, dmo.Lenv[site], dmo.Renv[site])
    n = H.shape[0]
    cols = []
    for c in range(n):
        e = np.zeros(n, dtype=complex)
        e[c] = 1.0
        y = cdev(np.zeros(n))
        plan.apply(cdev(e), y, dm.FP64)
        cols.append(y.cpu().numpy())
    Hg = np.stack(cols, axis=1)
    scale = np.abs(H).max()
    assert np.abs(Hg - H).max() < 1e-12 * scale
    assert np.abs(Hg - Hg.conj().T).max() < 1e-12 * scale
    Hr = heff_dense(lat, dmo.W[site], dmo.Lenv[site], dmo.Renv[site])
    assert np.abs(np.linalg.eigvalsh(Hg) - np.linalg.eigvalsh(Hr)).max() < 1e-11 * scale


@pytest.mark.parametrize("leigh", [0, 1])
def test_zenv_update_matches_oracle(leigh):
    dmo, site = state("4x3_d4")
    rng = np.random.default_rng(11 + leigh)
    lat = copy.deepcopy(dmo.space.lat[site])
    lat.ket.mats = [m + 0.5j * rng.standard_normal(m.shape) for m in lat.ket.mats]
    W = gauge_W(dmo.W[site], 4)
    env = complex_env(dmo.Lenv[site] if leigh == 0 else dmo.Renv[site], rng, 0.2)
    phys, kl, kr, krows, kcols = site_description(lat)
    (o, l, r, rows, cols, off), data = env_blocks(env)
    res = dm.env_update(leigh, W, phys, kl, kr, krows, kcols, cdev(lat.ket.to_vector()),
                        dm.BlockList(o, l, r, rows, cols, off), cdev(data))
    ref = (left_env if leigh == 0 else right_env)(lat, W, env)
    host = res.to_host()
    got = {(ch, bl, br): host[of:of + nr_ * nc].reshape(nr_, nc)
           for (ch, bl, br, nr_, nc, of) in res.blocks}
    assert len(res.blocks) == sum(b.n for b in ref)
    for ch, bm in enumerate(ref):
        for k in range(bm.n):
            m = bm.mats[k]
            g = got[(ch, bm.left[k], bm.right[k])]
            assert g.shape == m.shape
            if m.size:
                assert np.abs(g - m).max() <= 1e-12 * max(1.0, np.abs(m).max())


@pytest.mark.parametrize("name", ["3x2_d3", "4x3_d4"])
def test_zlanczos_matches_oracle(name):
    # Hermitian gauge-transformed operator, complex start vector: device and
    # oracle complex Lanczos agree; the dense edge path (3x2) also reaches the
    # exact ground energy of the real operator
    dmo, site = state(name)
    rng = np.random.default_rng(13)
    lat = copy.deepcopy(dmo.space.lat[site])
    lat.ket.mats = [m * np.exp(1j * rng.uniform(0, 2 * np.pi)) for m in lat.ket.mats]
    W = gauge_W(dmo.W[site], 5)
    plan = zplan(W, lat, dmo.Lenv[site], dmo.Renv[site])
    ket = cdev(lat.ket.to_vector())
    lat2 = copy.deepcopy(lat)
    e_ref = lanczos(lat2, W, dmo.Lenv[site], dmo.Renv[site], 10)
    e, _ = plan.lanczos(ket, 10, dm.FP64)
    assert abs(e - e_ref) < 1e-10 * abs(e_ref)
    v, vr = ket.cpu().numpy(), lat2.ket.to_vector()
    assert abs(abs(np.vdot(vr, v)) - 1.0) < 1e-8
    if lat.ket.dim() < 100:
        Hr = heff_dense(dmo.space.lat[site], dmo.W[site], dmo.Lenv[site], dmo.Renv[site])
        assert abs(e - np.linalg.eigvalsh(Hr)[0]) < 1e-10 * abs(e)
    else:
        # mixed precision: after 10 unconverged iterations the BF16x3 Krylov
        # space differs from the FP64 one (matvec error ~1e-5), so the refined
        # FP64 Rayleigh quotient agrees to ~1e-7 and, being variational, is
        # never below the exact ground energy, which the longer FP64 run bounds
        ket2 = cdev(lat.ket.to_vector())
        e2, _ = plan.lanczos(ket2, 10, dm.BF16X3, refine=True)
        assert abs(e2 - e_ref) < 1e-6 * abs(e_ref)
        ket3 = cdev(lat.ket.to_vector())
        e_long, _ = plan.lanczos(ket3, 60, dm.FP64)
        assert e2 >= e_long - 1e-9 * abs(e_long)


@pytest.mark.parametrize("shapes", [[(1, 1), (5, 5), (40, 12), (12, 40)], [(120, 90), (90, 120)],
                                    [(300, 64), (33, 257)],
                                    # QR panels of 16 columns in LDS (many panels, square), of 8
                                    # columns in LDS (M > 622), in global memory (M > 1267)
                                    [(600, 500), (470, 470)], [(1300, 64)], [(1400, 100), (100, 1500)]])
def test_zsector_svd(shapes):
    # complex sector SVD (ComplexMatrix::SVDMatrix) against numpy on graded
    # spectra down to 1e-12, with the phase convention of the oracle
    from oracle.dmrg import fix_signs
    rng = np.random.default_rng(21)
    mats, spec = [], []
    for (M, N) in shapes:
        a = rng.standard_normal((M, N)) + 1j * rng.standard_normal((M, N))
        u, s, vh = np.linalg.svd(a, full_matrices=False)
        s = np.logspace(0, -12, len(s))
        mats.append((u * s) @ vh)
        spec.append(s)
    U, S, Vt = dm.sector_svd([cdev(a) for a in mats])
    for a, s, u, sg, vt in zip(mats, spec, U, S, Vt):
        u, sg, vt = u.cpu().numpy(), sg.cpu().numpy(), vt.cpu().numpy()
        M, N = a.shape
        k = min(M, N)
        assert np.all(np.diff(sg) <= 0)
        assert np.abs(sg - s).max() < 1e-14 * max(M, N)
        assert np.abs((u * sg) @ vt - a).max() < 1e-14 * max(M, N)
        assert np.abs(u.conj().T @ u - np.eye(k)).max() < 1e-12
        assert np.abs(vt @ vt.conj().T - np.eye(k)).max() < 1e-12
        # phase convention: largest-modulus entry of each left vector real > 0
        top = u[np.argmax(np.abs(u), axis=0), np.arange(k)]
        assert np.abs(top.imag).max() < 1e-12 and (top.real > 0).all()
        ur, _, _ = np.linalg.svd(a, full_matrices=False)
        ur = ur * fix_signs(ur)[None, :]
        big = sg > 1e-4
        assert np.abs(ur[:, big] - u[:, big]).max() < 1e-8


@pytest.mark.parametrize("qr", ["1", "0"])
def test_zsector_svd_degenerate(qr, monkeypatch):
    # rank-deficient sectors with exact zero columns, a real-valued complex
    # matrix (reflectors with real alpha and no tail) and repeated singular
    # values, with and without the QR preconditioning
    monkeypatch.setenv("DMRGCC_ZSVD_QR", qr)
    rng = np.random.default_rng(5)
    a1 = rng.standard_normal((200, 40)) + 1j * rng.standard_normal((200, 40))
    a1 = a1 @ (rng.standard_normal((40, 90)) + 1j * rng.standard_normal((40, 90)))  # rank 40
    a1[:, 7] = 0
    a1[:, 50:54] = 0
    a2 = rng.standard_normal((150, 80)).astype(np.complex128)
    a2[:, 0] = 0
    q, _ = np.linalg.qr(rng.standard_normal((120, 60)) + 1j * rng.standard_normal((120, 60)))
    p, _ = np.linalg.qr(rng.standard_normal((60, 60)) + 1j * rng.standard_normal((60, 60)))
    s3 = np.repeat([2.0, 1.0, 0.5], 20)
    a3 = (q * s3) @ p.conj().T
    mats = [a1, a2, a3]
    U, S, Vt = dm.sector_svd([cdev(a) for a in mats])
    for a, u, sg, vt in zip(mats, U, S, Vt):
        u, sg, vt = u.cpu().numpy(), sg.cpu().numpy(), vt.cpu().numpy()
        M, N = a.shape
        k = min(M, N)
        ref = np.linalg.svd(a, compute_uv=False)
        scale = ref[0]
        assert np.abs(sg - ref).max() < 1e-13 * max(M, N) * scale
        assert np.abs(vt @ vt.conj().T - np.eye(k)).max() < 1e-12
        # left vectors of (numerically) zero singular values are not defined
        # (an exactly zero one gives a NaN column): the live part only
        live = sg > 1e-10 * scale
        uu = u[:, live]
        assert np.isfinite(uu).all()
        assert np.abs((uu * sg[live]) @ vt[live] - a).max() < 1e-13 * max(M, N) * scale
        assert np.abs(uu.conj().T @ uu - np.eye(live.sum())).max() < 1e-11
