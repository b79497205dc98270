"""dmrgcc_block_gemm_noalloc: the allocation-free form of dmrgcc_block_gemm
(GEMM descriptors in a fixed per-GPU device arena) gives the same products
as dmrgcc_block_gemm, bit for bit, on two calls in a row (the second smaller,
so stale arena contents from the first call would show), in each arithmetic;
sizes beyond the arena are refused with status 3."""
import ctypes

import numpy as np
import pytest

import dmrgcc_amd as dm

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

vp = ctypes.c_void_p


def _call(fn, As, Bs, Cs, precision):
    n = len(As)
    arr = lambda ts: (vp * max(n, 1))(*[vp(t.data_ptr()) for t in ts])
    i32 = lambda v: (ctypes.c_int * max(n, 1))(*v)
    M = i32([a.shape[0] for a in As])
    K = i32([a.shape[1] for a in As])
    N = i32([b.shape[1] for b in Bs])
    torch.cuda.synchronize()
    return fn(n, arr(As), arr(Bs), arr(Cs), M, N, K, precision, vp())


def _case(rng, shapes):
    As = [torch.tensor(rng.standard_normal((m, k)), device="cuda") for m, n, k in shapes]
    Bs = [torch.tensor(rng.standard_normal((k, n)), device="cuda") for m, n, k in shapes]
    return As, Bs


@pytest.mark.parametrize("precision", [dm.FP64, dm.FP32, dm.BF16X3])
def test_noalloc_equals_block_gemm(precision):
    lib = dm.lib()
    fn = lib.dmrgcc_block_gemm_noalloc
    rng = np.random.default_rng(7)
    big = [(300, 257, 700), (64, 64, 64), (1, 129, 33), (97, 5, 1000), (0, 3, 4), (130, 130, 0)]
    small = [(17, 40, 9), (33, 1, 65)]
    for shapes in (big, small):
        As, Bs = _case(rng, shapes)
        ref = [torch.full((m, n), 7.0, dtype=torch.float64, device="cuda") for m, n, k in shapes]
        got = [torch.full((m, n), -3.0, dtype=torch.float64, device="cuda") for m, n, k in shapes]
        assert _call(lib.dmrgcc_block_gemm, As, Bs, ref, precision) == 0
        assert _call(fn, As, Bs, got, precision) == 0, lib.dmrgcc_last_error().decode()
        torch.cuda.synchronize()
        for g, r, a, b in zip(got, ref, As, Bs):
            assert torch.equal(g, r)
            if g.numel() and precision == dm.FP64:
                assert torch.allclose(g, a @ b, rtol=1e-12, atol=1e-12 * max(1.0, a.shape[1]))


def test_noalloc_refuses_sizes_beyond_the_arena():
    lib = dm.lib()
    # 40 GEMMs of 1024 x 1024 outputs: 40 * 256 tiles of 16 bytes fit, but
    # 4000 such GEMMs' tile lists (64 MB) do not; the check comes before any
    # device work, so the operands need not exist beyond one small matrix
    a = torch.zeros((1024, 1), dtype=torch.float64, device="cuda")
    b = torch.zeros((1, 1024), dtype=torch.float64, device="cuda")
    c = torch.zeros((1024, 1024), dtype=torch.float64, device="cuda")
    n = 4000
    arr = lambda t: (vp * n)(*([vp(t.data_ptr())] * n))
    i32 = lambda v: (ctypes.c_int * n)(*([v] * n))
    rc = lib.dmrgcc_block_gemm_noalloc(n, arr(a), arr(b), arr(c), i32(1024), i32(1024), i32(1), dm.FP64, vp())
    assert rc == 3
    assert "arena" in lib.dmrgcc_last_error().decode()
