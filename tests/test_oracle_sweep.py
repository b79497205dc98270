"""Pins of the oracle's sweep-level path by mathematics (no GPU): the
subspace expansion (LeftExpanTensorContraction, real_tensor_contraction.cc:
169-459; ComputeExpanTensorLattice and ExpanTensorSpace, real_tensor_space.cc:
319-387, :552-706; LeftExpanTensorLattice, real_tensor_lattice.cc:979-1063),
the Lanczos recurrence (real_tensor_lanczos.cc:14-201) and the tensor-space
file layout (real_tensor_space.cc:182-208, real_tensor_lattice.cc:218-239),
each checked against what it must satisfy rather than against a second
restatement of the same loops."""
import copy
import struct

import numpy as np
import pytest

from oracle.blocks import BlockMat
from oracle.dmrg import (DMRG, heff_apply, heff_dense, lanczos, left_canonical, left_env,
                         left_expan_contraction, left_merge)
from oracle.io import read_space, write_space
from oracle.model import SpinfulBoseSquare, SpinlessBoseSquare

NOISE_OFF = -1


def stopped(model, maxm, half, site, noise, niter=8, canpre=1e-10, first_dim=4, sweeps=2):
    rows = [(maxm, 400, niter, canpre, noise[i] if i < len(noise) else NOISE_OFF) for i in range(2 * sweeps)]
    dmo = DMRG(model, rows, sweeps, first_block=20, first_dim=first_dim)
    dmo.run(stop=(half, site))
    return dmo


MODELS = {
    "2x2": lambda: SpinlessBoseSquare(2, 2, 4, 3, jx=-1.0, jy=-1.0, U=2.0),
    "3x2": lambda: SpinlessBoseSquare(3, 2, 6, 3, jx=-1.0, jy=-0.7, U=1.5, period_y=False),
}


def lanczos_then_expand(dmo, j, niter=8):
    """The left-to-right site step up to the expansion contraction
    (RealTensorRundmrg::Run, real_tensor_rundmrg.cc:147-156)."""
    sp = dmo.space
    e = lanczos(sp.lat[j], dmo.W[j], dmo.Lenv[j], dmo.Renv[j], niter)
    mapping = sp.expan_lattice(1, j, dmo.op_tables[j + 1])
    before = dmo.rng.count
    left_expan_contraction(sp.lat[j], dmo.W[j], dmo.Lenv[j], sp.expan, mapping, sp.noise, dmo.rng)
    return e, mapping, dmo.rng.count - before


@pytest.mark.parametrize("name", list(MODELS))
def test_expansion_keeps_the_original_blocks(name):
    dmo = stopped(MODELS[name](), 40, 2, 1, [2, 2, 2, 2])
    sp = dmo.space
    j = 1
    lanczos_then_expand(dmo, j)
    origin = sp.lat[j].copy()
    E = copy.deepcopy(sp.expan)
    sp.expand(0, j)
    lat = sp.lat[j]
    assert lat.left_block == origin.left_block            # the left bond is untouched
    assert set(lat.right_block) == set(origin.right_block) | set(E.right_block)
    assert lat.right_block == sorted(lat.right_block)     # sorted union (CombineTensorLattice)
    for k in range(lat.ket.n):
        l, r = lat.ket.left[k], lat.ket.right[k]
        q = lat.right_block[r]
        m = lat.ket.mats[k]
        c0 = 0
        if q in origin.right_block:
            p0 = origin.ket.find(l, origin.right_block.index(q))
            if p0 != -1:
                old = origin.ket.mats[p0]
                assert np.array_equal(m[:old.shape[0], :old.shape[1]], old)   # unchanged, leading
                c0 = old.shape[1]
        if q in E.right_block:
            p1 = E.ket.find(l, E.right_block.index(q))
            if p1 != -1:
                x = E.ket.mats[p1]
                assert np.array_equal(m[:, c0:c0 + x.shape[1]], x)           # the expansion, appended
                c0 += x.shape[1]
        assert not np.any(m[:, c0:])                                          # the rest is zero padding
        # rows: the sector dimension; columns: at most the sector dimension
        # (the real stack leaves a block that took expansion columns as wide
        # as its own terms, real_tensor_lattice.cc:1018-1040), exactly it
        # where only zero padding was added
        assert m.shape[0] == lat.left_dim[l] and m.shape[1] <= lat.right_dim[r]
        if q not in E.right_block:
            assert m.shape[1] == lat.right_dim[r]


@pytest.mark.parametrize("name", list(MODELS))
def test_expansion_columns_and_sectors(name):
    base = stopped(MODELS[name](), 40, 2, 1, [2, 2, 2, 2])
    j = 1
    runs = {}
    for eps in (1e-2, 1e-5):
        dmo = copy.deepcopy(base)
        dmo.space.noise = eps
        e, mapping, draws = lanczos_then_expand(dmo, j)
        runs[eps] = (dmo, mapping, draws)
    (d1, mp1, n1), (d2, mp2, n2) = runs[1e-2], runs[1e-5]
    E1, E2 = d1.space.expan, d2.space.expan
    # the noise factor is a pure scale: same draws, same blocks up to 1e3
    assert n1 == n2 > 0 and mp1 == mp2
    assert (E1.ket.left, E1.ket.right) == (E2.ket.left, E2.ket.right)
    for a, b in zip(E1.ket.mats, E2.ket.mats):
        assert a.shape == b.shape
        assert np.allclose(a, 1e3 * b, rtol=1e-13, atol=0)
    # new right sectors: QN merges of the MPO channels on the old right
    # sectors that the table holds, in table order, at most max_block
    sp, lat = d1.space, d1.space.lat[j]
    tab = sp.tables[j + 1]
    closure = {sp.model.merge(oq, tab[ob]) for oq in d1.op_tables[j + 1] for ob in lat.right_block}
    assert all(tab[b] in closure for b in E1.right_block)
    assert E1.right_block == sorted(E1.right_block) and len(E1.right_block) <= sp.max_block
    # every column of an expansion block lies in the span of the products
    # L[o1](i, k) K(k, j) that may feed it (any channel, any k, any j whose
    # merged sector is this one): a wrong operand or channel leaves the span
    W, Lenv, ket = d1.W[j], d1.Lenv[j], lat.ket
    for pos in range(E1.ket.n):
        i, r = E1.ket.left[pos], E1.ket.right[pos]
        q = tab[E1.right_block[r]]
        cols = []
        for o1 in range(W.shape[0]):
            Lb = Lenv[o1]
            for a in range(Lb.n):
                if Lb.left[a] != i or Lb.mats[a].shape[0] == 0:
                    continue
                for kp in ket.same(0, Lb.right[a]):
                    jq = tab[lat.right_block[ket.right[kp]]]
                    if any(sp.model.merge(oq, jq) == q for oq in d1.op_tables[j + 1]):
                        cols.append(Lb.mats[a] @ ket.mats[kp])
        X = E1.ket.mats[pos]
        if X.size == 0 or not np.any(X):
            continue
        B = np.hstack(cols)
        Q, s, _ = np.linalg.svd(B, full_matrices=False)
        Q = Q[:, s > 1e-12 * s[0]]
        resid = X - Q @ (Q.T @ X)
        assert np.abs(resid).max() <= 1e-11 * np.abs(X).max()


@pytest.mark.parametrize("name", list(MODELS))
@pytest.mark.parametrize("noise", [2, NOISE_OFF])
def test_expansion_preserves_the_physical_state(name, noise):
    # Expansion adds basis states with zero amplitude (the next site gets
    # zero rows for them), so with no truncation the state that reaches the
    # next site has the Rayleigh quotient the Lanczos solve at this site found.
    half, j = 2, 1
    model = MODELS[name]()
    a = stopped(model, 400, half, j, [2, 2, noise, noise])
    e_site = lanczos(a.space.lat[j].copy(), a.W[j], a.Lenv[j], a.Renv[j], 8)
    b = stopped(model, 400, half, j + 1, [2, 2, noise, noise])
    lat = b.space.lat[j + 1]
    x = lat.ket.to_vector()
    hx = heff_apply(lat, b.W[j + 1], b.Lenv[j + 1], b.Renv[j + 1], x, np.zeros_like(x))
    rq = float(np.dot(x, hx) / np.dot(x, x))
    assert abs(rq - e_site) <= 1e-10 * abs(e_site)


def test_lanczos_basis_and_tridiagonal():
    # modified Gram-Schmidt keeps the Krylov basis orthonormal, and the
    # recurrence's alpha / beta are the matrix elements of H in that basis
    dmo = stopped(MODELS["3x2"](), 40, 2, 2, [2, 2, 2, 2])
    lat = dmo.space.lat[2]
    W, L, R = dmo.W[2], dmo.Lenv[2], dmo.Renv[2]
    assert lat.ket.dim() >= 100   # the iterative path, not the dense one
    H = heff_dense(lat, W, L, R)
    tr = {}
    e = lanczos(lat.copy(), W, L, R, 12, trace=tr)
    V = np.stack(tr["psi"][:len(tr["alpha"])], axis=1)
    k = V.shape[1]
    assert np.abs(V.T @ V - np.eye(k)).max() < 1e-12
    T = np.diag(tr["alpha"]) + np.diag(tr["beta"][:k - 1], 1) + np.diag(tr["beta"][:k - 1], -1)
    assert np.abs(V.T @ H @ V - T).max() < 1e-10 * np.abs(H).max()
    assert abs(e - np.linalg.eigvalsh(T)[0]) < 1e-12 * abs(e)
    assert e >= np.linalg.eigvalsh(H)[0] - 1e-12 * abs(e)   # variational


def _i(*v):
    return struct.pack("<%di" % len(v), *v)


def space_fixture():
    """A two-site tensor-space file of the 2x1 spinless model (N = 2, d = 3)
    written field by field in the order of WriteTensorSpace /
    WriteTensorLattice / WriteMatrixBlock / WriteMatrix
    (real_tensor_space.cc:182-208, real_tensor_lattice.cc:218-239,
    real_matrix_block.cc:227-240, real_matrix.cc:139-151), with the blocks
    the model's quantum-number rule allows.  Returns (model, bytes, blocks)."""
    model = SpinlessBoseSquare(2, 1, 2, 3, jx=-1.0, U=1.0)
    tabs = model.space_tables()
    q = [tabs[1][0][0], tabs[1][1][0]]           # the two middle sectors used
    a0, a1 = np.array([[0.5]]), np.array([[-1.25, 0.375]])
    b0, b1 = np.array([[2.0]]), np.array([[0.75], [-0.5]])
    raw = _i(2, 3, 1, 1) + _i(*[len(t) for t in tabs])
    for t in tabs:
        for qq in t:
            raw += _i(*qq)
    # site 0: d, nl, (left_block, left_dim), nr, (right_block, right_dim) x 2,
    # physics_index[1][2], blocks (0, 0) 1 x 1 and (0, 1) 1 x 2
    raw += _i(3, 1, 0, 1, 2, 0, 1, 1, 2, q[0], q[1])
    raw += _i(2) + _i(0, 0, 1, 1, 1) + a0.tobytes() + _i(0, 1, 1, 2, 2) + a1.tobytes()
    # site 1: nl = 2 (the same sectors), nr = 1, physics_index[2][1],
    # blocks (0, 0) 1 x 1 and (1, 0) 2 x 1
    n2 = tabs[2][0][0]
    raw += _i(3, 2, 0, 1, 1, 2, 1, 0, 1, n2 - q[0], n2 - q[1])
    raw += _i(2) + _i(0, 0, 1, 1, 1) + b0.tobytes() + _i(1, 0, 2, 1, 2) + b1.tobytes()
    raw += _i(7, 11) + struct.pack("<dd", 1e-9, 0.125)
    return model, raw, (a0, a1, b0, b1, q, n2)


def test_tensor_space_byte_fixture(tmp_path):
    # the oracle reads exactly the written values and writes the same bytes
    model, raw, (a0, a1, b0, b1, q, n2) = space_fixture()
    p = tmp_path / "space.dat"
    p.write_bytes(raw)
    dmo = DMRG(model, [(10, 10, 4, 1e-8, 2)] * 2, 1)
    sp = dmo.space
    read_space(str(p), sp)
    assert [list(map(tuple, t)) for t in sp.tables] == [list(map(tuple, t)) for t in model.space_tables()]
    l0, l1 = sp.lat
    assert (l0.left_block, l0.left_dim, l0.right_block, l0.right_dim) == ([0], [1], [0, 1], [1, 2])
    assert l0.phys == [[q[0], q[1]]] and (l0.ket.left, l0.ket.right) == ([0, 0], [0, 1])
    assert np.array_equal(l0.ket.mats[0], a0) and np.array_equal(l0.ket.mats[1], a1)
    assert (l1.left_block, l1.left_dim, l1.right_block, l1.right_dim) == ([0, 1], [1, 2], [0], [1])
    assert l1.phys == [[n2 - q[0]], [n2 - q[1]]]
    assert np.array_equal(l1.ket.mats[0], b0) and np.array_equal(l1.ket.mats[1], b1)
    assert (sp.max_block, sp.max_dim, sp.canpre, sp.noise) == (7, 11, 1e-9, 0.125)
    again = tmp_path / "again.dat"
    write_space(str(again), sp)
    assert again.read_bytes() == raw
