"""The drop-in class API (include/dmrgcc/): a C++ harness written against the
reference's class names and member signatures (tests/cxx/dropin_test.cpp,
built by dmrgcc_amd/build.py into tests/cxx/libdropin_test.so and loaded here
with ctypes) drives one site of an oracle state through
MultiplyEffectHamilton, Left/RightComputeTensorContraction, LanczosMethod and
Left/RightCanonicalTensorLattice, and runs the reference's real spinless
driver on a 3x2 input file through RealTensorRundmrg::Run; each result is
compared with the oracle."""
import ctypes
import os

import numpy as np
import pytest

import dmrgcc_amd as dm
from oracle.dmrg import DMRG, heff_apply, lanczos, left_canonical, left_env, right_canonical, right_env
from oracle.model import SpinlessBoseSquare
from tests._support import env_blocks, oracle_state

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
HARNESS = os.path.join(HERE, "cxx", "libdropin_test.so")
GOLDEN = os.path.join(HERE, "golden")

ip = lambda a: np.ascontiguousarray(a, dtype=np.int32).ctypes.data_as(ctypes.POINTER(ctypes.c_int))
lp = lambda a: np.ascontiguousarray(a, dtype=np.int64).ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def harness():
    dm.lib()  # the library (and its HIP runtime) first
    h = ctypes.CDLL(HARNESS)
    h.dropin_site_create.restype = ctypes.c_void_p
    h.dropin_site_lanczos.restype = ctypes.c_double
    for f in ("dropin_site_env", "dropin_site_heff", "dropin_site_lanczos", "dropin_site_env_update",
              "dropin_site_canonical", "dropin_site_destroy"):
        getattr(h, f).argtypes = None
    return h


class Site:
    def __init__(self, h, dmo, site):
        self.h, self.lat = h, dmo.space.lat[site]
        self.W, self.L, self.R = dmo.W[site], dmo.Lenv[site], dmo.Renv[site]
        lat = self.lat
        ket = lat.ket
        self.phys = np.array(lat.phys, dtype=np.int32)
        self.keep = [np.ascontiguousarray(self.W, dtype=np.float64), ket.to_vector()]
        rows = [m.shape[0] for m in ket.mats]
        cols = [m.shape[1] for m in ket.mats]
        self.ptr = ctypes.c_void_p(h.dropin_site_create(
            ctypes.c_int(lat.d), ctypes.c_int(self.W.shape[0]), ctypes.c_int(self.W.shape[1]), dp(self.keep[0]),
            ctypes.c_int(lat.nl), ctypes.c_int(lat.nr), ip(lat.left_dim), ip(lat.right_dim), ip(self.phys),
            ctypes.c_int(ket.n), ip(ket.left), ip(ket.right), ip(rows), ip(cols), dp(self.keep[1])))
        for leigh, env in ((0, self.L), (1, self.R)):
            (o, l, r, er, ec, off), data = env_blocks(env)
            data = np.ascontiguousarray(data if data.size else np.zeros(1))
            self.keep.append(data)
            h.dropin_site_env(self.ptr, ctypes.c_int(leigh), ctypes.c_int(len(o)), ip(o), ip(l), ip(r), ip(er),
                              ip(ec), lp(off), dp(data))

    def __del__(self):
        self.h.dropin_site_destroy(self.ptr)


@pytest.fixture(scope="module")
def site():
    dmo = oracle_state(3, 2, 6, 3, maxm=40, half=2, site=2, niter=8)
    return Site(harness(), dmo, 2)


def test_multiply_effect_hamilton(site):
    n = site.lat.ket.dim()
    y0 = np.random.default_rng(1).standard_normal(n)
    ref = heff_apply(site.lat, site.W, site.L, site.R, site.lat.ket.to_vector(), y0.copy())
    for prec, tol in ((dm.FP64, 1e-12), (dm.BF16X3, 2e-4)):
        y = y0.copy()
        site.h.dropin_site_heff(site.ptr, dp(y), ctypes.c_int(prec))
        assert np.abs(y - ref).max() <= tol * np.abs(ref).max(), prec


@pytest.mark.parametrize("leigh", [0, 1])
def test_compute_tensor_contraction(site, leigh):
    ref = (left_env if leigh == 0 else right_env)(site.lat, site.W, site.L if leigh == 0 else site.R)
    cap = 4096
    o, l, r, rows, cols = (np.zeros(cap, dtype=np.int32) for _ in range(5))
    off = np.zeros(cap, dtype=np.int64)
    data = np.zeros(1 << 20)
    n = site.h.dropin_site_env_update(site.ptr, ctypes.c_int(leigh), ctypes.c_int(cap), ip(o), ip(l), ip(r),
                                      ip(rows), ip(cols), lp(off), dp(data), ctypes.c_int64(data.size))
    # ip() copies only non-int32 input; these are int32 already, so the
    # harness wrote into the arrays themselves
    assert n == sum(b.n for b in ref)
    got = {(int(o[i]), int(l[i]), int(r[i])): data[off[i]:off[i] + rows[i] * cols[i]].reshape(rows[i], cols[i])
           for i in range(n)}
    for ch, bm in enumerate(ref):
        for k in range(bm.n):
            g = got[(ch, bm.left[k], bm.right[k])]
            assert g.shape == bm.mats[k].shape
            if g.size:
                assert np.abs(g - bm.mats[k]).max() <= 1e-12 * max(1.0, np.abs(bm.mats[k]).max())


def test_lanczos_method(site):
    lat2 = site.lat.copy()
    e_ref = lanczos(lat2, site.W, site.L, site.R, 10)
    v = np.zeros(site.lat.ket.dim())
    e = site.h.dropin_site_lanczos(site.ptr, ctypes.c_int(10), dp(v))
    assert abs(e - e_ref) <= 1e-10 * abs(e_ref)
    assert abs(abs(np.dot(v, lat2.ket.to_vector())) - 1.0) < 1e-8


@pytest.mark.parametrize("leigh", [0, 1])
def test_canonical_tensor_lattice(site, leigh):
    max_dim, canpre = 12, 1e-10
    lat2 = site.lat.copy()
    (left_canonical if leigh == 0 else right_canonical)(lat2, max_dim, canpre)
    nb = site.lat.ket.n
    ns = site.lat.nr if leigh == 0 else site.lat.nl
    ket = np.zeros(1 << 18)
    can = np.zeros(1 << 18)
    rows, cols = np.zeros(nb, dtype=np.int32), np.zeros(nb, dtype=np.int32)
    dims, match = np.zeros(ns, dtype=np.int32), np.zeros(ns, dtype=np.int32)
    disc = ctypes.c_double()
    got_ns = site.h.dropin_site_canonical(site.ptr, ctypes.c_int(leigh), ctypes.c_int(max_dim), ctypes.c_double(canpre),
                                          dp(ket), ctypes.c_int64(ket.size), ip(rows), ip(cols), ip(dims), ip(match),
                                          dp(can), ctypes.c_int64(can.size), ctypes.byref(disc))
    assert got_ns == ns
    assert list(dims) == list(lat2.right_dim if leigh == 0 else lat2.left_dim)
    assert list(match) == list(lat2.match_dim)
    assert [tuple(m.shape) for m in lat2.ket.mats] == list(zip(rows, cols))
    ref_ket = lat2.ket.to_vector()
    assert np.abs(ket[:ref_ket.size] - ref_ket).max() < 1e-10
    ref_can = np.concatenate([m.ravel() for m in lat2.canonical.mats])
    assert np.abs(can[:ref_can.size] - ref_can).max() < 1e-10 * max(1.0, np.abs(ref_can).max())


def test_rundmrg_run_matches_oracle():
    h = harness()
    path = os.path.join(GOLDEN, "input_spinless_small")
    c = dm.input_run_config(path)
    e = np.zeros(64)
    bonds = np.zeros(64, dtype=np.int32)
    n = h.dropin_run(os.fsencode(path), dp(e), ctypes.c_int(64), ip(bonds))
    assert n == len(c["table"])
    om = SpinlessBoseSquare(3, 2, 6, 3, jx=-1.0, jy=-0.75, U=2.0, mu=0.1, period_y=True)
    o = DMRG(om, c["table"], n // 2, first_block=c["first_block"], first_dim=c["first_dim"])
    e_ref = np.array(o.run())
    assert np.abs(e[:n] - e_ref).max() < 1e-10 * np.abs(e_ref).max()
    ref_bonds = []
    for lat in o.space.lat:
        used = {}
        for r, m in zip(lat.ket.right, lat.ket.mats):
            used[r] = m.shape[1]
        ref_bonds.append(sum(used.values()))
    assert list(bonds[:om.L]) == ref_bonds
