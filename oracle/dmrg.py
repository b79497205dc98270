"""Single-site DMRG with subspace expansion, oracle copy (real FP64 and
complex128: the complex twins in src/dmrg/complexdmrg/ run the same steps on
Complex entries, with the Hermitian conjugate bra in the environment updates,
two rand() draws per random entry, and complex-specific expansion padding).

Restates src/dmrg/realdmrg/:
* real_tensor_contraction.cc  -- H_eff matvec (:1223-1291), dense H_eff
  (:1136-1205), environment updates (:782-1119), expansion (:169-765);
* real_tensor_lanczos.cc      -- Lanczos + Gram-Schmidt (:14-201);
* real_tensor_lattice.cc      -- truncation rule (:501-555), left/right
  canonicalisation (:558-676, :794-911), merge (:914-977), combine
  (:361-484), expansion fill (:979-1149);
* real_tensor_space.cc        -- initial MPS (:68-134), sector lists
  (:442-550), expansion lattice (:552-706), expand/match (:319-440);
* real_tensor_network.cc      -- environment bookkeeping (:3-154);
* real_tensor_rundmrg.cc      -- sweep table (:70-113), Run (:115-245),
  Initialize (:247-287).

Environment blocks: L[o1] has blocks (i, k) = (bra left sector, ket left
sector) of shape D^L_i x D^L_k; R[o2] has blocks (l, j) = (ket right sector,
bra right sector) of shape D^R_l x D^R_j.
"""
import math
import time

import numpy as np

from .blocks import BlockMat, Lattice, empty
from .model import left_check_zero, right_check_zero, mixed_check_zero
from .rng import GlibcRand
from .sort import quicksort


# --------------------------------------------------------------------------
# H_eff
# --------------------------------------------------------------------------
def heff_tasks(lat, W, Lenv, Renv):
    """Task enumeration of MultiplyEffectHamilton (real_tensor_contraction.cc
    :1223-1291), in the reference's loop order.

    Returns a list of (o2, r_pos, k_pos, [(out_pos, o1, l_pos, w), ...]):
    T = K[k_pos] @ R[o2][r_pos]; out[out_pos] += w * L[o1][l_pos] @ T.
    A complex W (ComplexTensorContraction::MultiplyEffectHamilton,
    complex_tensor_contraction.cc:1244-1312) keeps complex w; an entry is used
    when its real or imaginary part is nonzero (:1288)."""
    lb, rb = W.shape[0], W.shape[1]
    ket = lat.ket
    same_r = [ket.same(1, r) for r in range(lat.nr)]
    tasks = []
    for o2 in range(rb):
        R = Renv[o2]
        for ri in range(R.n):
            ldx, jdx = R.left[ri], R.right[ri]
            if R.mats[ri].shape[0] == 0:
                continue
            for kp in same_r[ldx]:
                kdx = ket.left[kp]
                p2 = lat.phys[kdx][ldx]
                if not right_check_zero(W, o2, p2):
                    continue
                outs = []
                for op in same_r[jdx]:
                    idx = ket.left[op]
                    p1 = lat.phys[idx][jdx]
                    for o1 in range(lb):
                        w = W[o1, o2, p1, p2]
                        if w != 0.0:
                            lp = Lenv[o1].find(idx, kdx)
                            if lp != -1:
                                outs.append((op, o1, lp, w.item()))
                tasks.append((o2, ri, kp, outs))
    return tasks


def heff_flops(lat, W, Lenv, Renv, tasks=None):
    """F_ref: real flops of one H_eff application in the reference order
    (complex: 8 real flops per multiply-add)."""
    if tasks is None:
        tasks = heff_tasks(lat, W, Lenv, Renv)
    f = 0
    ket = lat.ket
    for o2, ri, kp, outs in tasks:
        R = Renv[o2].mats[ri]
        K = ket.mats[kp]
        f += 2 * K.shape[0] * K.shape[1] * R.shape[1]
        for op, o1, lp, w in outs:
            Lm = Lenv[o1].mats[lp]
            f += 2 * Lm.shape[0] * Lm.shape[1] * R.shape[1]
    return 4 * f if np.iscomplexobj(W) else f


def heff_apply(lat, W, Lenv, Renv, x, y, tasks=None):
    """MultiplyEffectHamilton: y += H_eff x (x in ket-vector layout)."""
    if tasks is None:
        tasks = heff_tasks(lat, W, Lenv, Renv)
    ket = lat.ket
    offs = ket.offsets()
    shapes = [m.shape for m in ket.mats]
    for o2, ri, kp, outs in tasks:
        K = x[offs[kp]:offs[kp] + shapes[kp][0] * shapes[kp][1]].reshape(shapes[kp])
        T = K @ Renv[o2].mats[ri]
        for op, o1, lp, w in outs:
            S = (Lenv[o1].mats[lp] @ T) * w
            n = S.size
            y[offs[op]:offs[op] + n] += S.ravel()
    return y


def heff_dense(lat, W, Lenv, Renv):
    """ComputeEffectHamilton (real_tensor_contraction.cc:1136-1205):
    H[(i,j),(k,l)] += w * L[o1](i,k) (x) R[o2](l,j)^T."""
    ket = lat.ket
    dim = ket.dim()
    offs = ket.offsets()
    cplx = np.iscomplexobj(W) or any(np.iscomplexobj(m) for m in ket.mats)
    H = np.zeros((dim, dim), dtype=complex if cplx else float)
    lb, rb = W.shape[0], W.shape[1]
    for o1 in range(lb):
        Lb = Lenv[o1]
        for a in range(Lb.n):
            idx, kdx = Lb.left[a], Lb.right[a]
            Lm = Lb.mats[a]
            if Lm.shape[0] == 0:
                continue
            for o2 in range(rb):
                Rb = Renv[o2]
                for b in range(Rb.n):
                    ldx, jdx = Rb.left[b], Rb.right[b]
                    Rm = Rb.mats[b]
                    if Rm.shape[0] == 0:
                        continue
                    p0 = ket.find(idx, jdx)
                    p1 = ket.find(kdx, ldx)
                    if p0 != -1 and p1 != -1:
                        w = W[o1, o2, lat.phys[idx][jdx], lat.phys[kdx][ldx]]
                        if w != 0:
                            kr = np.kron(Lm, Rm.T) * w
                            if kr.size == 0:
                                continue
                            H[offs[p0]:offs[p0] + kr.shape[0], offs[p1]:offs[p1] + kr.shape[1]] += kr
    return H


# --------------------------------------------------------------------------
# Environment updates
# --------------------------------------------------------------------------
def left_env(lat, W, Lenv):
    """LeftComputeTensorContraction (real_tensor_contraction.cc:782-940):
    Lnew[o2](i,j) += w(p1,p2) * B(k,i)^T L[o1](k,l) K(l,j); complex
    (complex_tensor_contraction.cc:803-978) with the bra B^H (:929)."""
    lb, rb = W.shape[0], W.shape[1]
    ket = lat.ket
    nr = lat.nr
    same_l = [ket.same(0, l) for l in range(lat.nl)]
    flag = np.zeros((rb, nr, nr), dtype=bool)
    for o1 in range(lb):
        Lb = Lenv[o1]
        for a in range(Lb.n):
            kdx, ldx = Lb.left[a], Lb.right[a]
            if Lb.mats[a].shape[0] == 0:
                continue
            for kp in same_l[kdx]:
                idx = ket.right[kp]
                p1 = lat.phys[kdx][idx]
                if not left_check_zero(W, o1, p1):
                    continue
                for lp in same_l[ldx]:
                    jdx = ket.right[lp]
                    p2 = lat.phys[ldx][jdx]
                    for o2 in range(rb):
                        if W[o1, o2, p1, p2] != 0.0:
                            flag[o2, idx, jdx] = True
    out = []
    for o2 in range(rb):
        li, ri = np.nonzero(flag[o2])
        out.append(BlockMat(li.tolist(), ri.tolist()))
    for o1 in range(lb):
        Lb = Lenv[o1]
        for a in range(Lb.n):
            kdx, ldx = Lb.left[a], Lb.right[a]
            Lm = Lb.mats[a]
            if Lm.shape[0] == 0:
                continue
            for kp in same_l[kdx]:
                idx = ket.right[kp]
                p1 = lat.phys[kdx][idx]
                if not left_check_zero(W, o1, p1):
                    continue
                F = ket.mats[kp].conj().T @ Lm
                for lp in same_l[ldx]:
                    jdx = ket.right[lp]
                    p2 = lat.phys[ldx][jdx]
                    S = F @ ket.mats[lp]
                    for o2 in range(rb):
                        w = W[o1, o2, p1, p2]
                        if w != 0.0:
                            pos = out[o2].find(idx, jdx)
                            _add_block(out[o2], pos, w, S)
    return out


def right_env(lat, W, Renv):
    """RightComputeTensorContraction (real_tensor_contraction.cc:959-1119):
    Rnew[o1](j,i) += w(p1,p2) * K(j,l) R[o2](l,k) B(i,k)^T; complex
    (complex_tensor_contraction.cc:980-1158) with B^H (:1113)."""
    lb, rb = W.shape[0], W.shape[1]
    ket = lat.ket
    nl = lat.nl
    same_r = [ket.same(1, r) for r in range(lat.nr)]
    flag = np.zeros((lb, nl, nl), dtype=bool)
    for o2 in range(rb):
        Rb = Renv[o2]
        for a in range(Rb.n):
            ldx, kdx = Rb.left[a], Rb.right[a]
            if Rb.mats[a].shape[0] == 0:
                continue
            for lp in same_r[ldx]:
                jdx = ket.left[lp]
                p2 = lat.phys[jdx][ldx]
                if not right_check_zero(W, o2, p2):
                    continue
                for kp in same_r[kdx]:
                    idx = ket.left[kp]
                    p1 = lat.phys[idx][kdx]
                    for o1 in range(lb):
                        if W[o1, o2, p1, p2] != 0.0:
                            flag[o1, jdx, idx] = True
    out = []
    for o1 in range(lb):
        li, ri = np.nonzero(flag[o1])
        out.append(BlockMat(li.tolist(), ri.tolist()))
    for o2 in range(rb):
        Rb = Renv[o2]
        for a in range(Rb.n):
            ldx, kdx = Rb.left[a], Rb.right[a]
            Rm = Rb.mats[a]
            if Rm.shape[0] == 0:
                continue
            for lp in same_r[ldx]:
                jdx = ket.left[lp]
                p2 = lat.phys[jdx][ldx]
                if not right_check_zero(W, o2, p2):
                    continue
                F = ket.mats[lp] @ Rm
                for kp in same_r[kdx]:
                    idx = ket.left[kp]
                    p1 = lat.phys[idx][kdx]
                    S = F @ ket.mats[kp].conj().T
                    for o1 in range(lb):
                        w = W[o1, o2, p1, p2]
                        if w != 0.0:
                            pos = out[o1].find(jdx, idx)
                            _add_block(out[o1], pos, w, S)
    return out


def _add_block(bm, pos, w, S):
    """RealMatrixBlock::AddToMatrixBlock -> RealMatrix::AddToMatrix(factor, m)
    (real_matrix.cc:248-276): an empty block takes the (scaled) copy."""
    if S.size == 0:
        return
    if bm.mats[pos].size == 0:
        bm.mats[pos] = w * S
    else:
        if bm.mats[pos].shape != S.shape:
            raise ValueError("Matrix indeces do not match in AddToMatrix")
        bm.mats[pos] = bm.mats[pos] + w * S


def boundary_env(nbond):
    """DefineTensorContraction (real_tensor_contraction.cc:111-137)."""
    return [BlockMat([0], [0], [np.ones((1, 1))]) for _ in range(nbond)]


# --------------------------------------------------------------------------
# Lanczos
# --------------------------------------------------------------------------
def fix_signs(vecs):
    """Sign convention for eigen/singular vectors (columns of `vecs`): the
    largest-magnitude component (first on ties) is made positive.  The
    reference leaves the sign to LAPACK; fixing it makes the trajectory
    reproducible, because the expansion noise contracts a fixed random matrix
    with the gauge-carrying bond index."""
    idx = np.argmax(np.abs(vecs), axis=0)
    top = vecs[idx, np.arange(vecs.shape[1])]
    if np.iscomplexobj(vecs):
        # complex vectors: the phase that makes that component real positive
        mag = np.abs(top)
        return np.where(mag > 0, np.conj(top) / np.where(mag > 0, mag, 1.0), 1.0)
    sg = np.where(top < 0, -1.0, 1.0)
    return sg


def sym_eig(A):
    """RealSymMatrixDiag (src/util/general.cc:47-64; complex Hermitian:
    ComplexSymMatrixDiag, :66-83): ascending eigenvalues, eigenvectors in
    columns (sign / phase convention of fix_signs)."""
    w, v = np.linalg.eigh(A)
    return w, v * fix_signs(v)[None, :]


def lanczos(lat, W, Lenv, Renv, num_iter, stats=None, trace=None):
    """LanczosMethod (real_tensor_lanczos.cc:14-169; complex
    complex_tensor_lanczos.cc:14-160: complex alpha subtracted, beta real,
    the tridiagonal's real diagonal diagonalised).  Writes the Ritz vector
    into lat.ket and returns the energy.  With a dict `trace`, the Krylov
    vectors and the alpha / beta of every iteration are stored there."""
    psi_dim = lat.ket.dim()
    if psi_dim < 100:
        H = heff_dense(lat, W, Lenv, Renv)
        ev, vec = sym_eig(H)
        lat.ket.from_vector(vec[:, 0])
        if stats is not None:
            stats['dense'] = stats.get('dense', 0) + 1
        return float(ev[0])
    tasks = heff_tasks(lat, W, Lenv, Renv)
    psi = [lat.ket.to_vector()]
    cplx = np.iscomplexobj(psi[0]) or np.iscomplexobj(W)
    if cplx:
        psi[0] = psi[0].astype(complex)
    h = np.zeros(psi_dim, dtype=complex if cplx else float)
    alpha, beta = [], []
    old = 10000.0
    energy = None
    for i in range(num_iter):
        heff_apply(lat, W, Lenv, Renv, psi[i], h, tasks)
        if stats is not None:
            stats['matvec'] = stats.get('matvec', 0) + 1
        a = np.vdot(psi[i], h) if cplx else float(np.dot(psi[i], h))
        h -= a * psi[i]
        b = math.sqrt(float(np.vdot(h, h).real))
        a = float(np.real(a))
        alpha.append(a)
        beta.append(b)
        k = i + 1
        T = np.diag(alpha) + np.diag(beta[:k - 1], 1) + np.diag(beta[:k - 1], -1)
        ev, vec = sym_eig(T)
        new = float(ev[0])
        if trace is not None:
            trace["psi"], trace["alpha"], trace["beta"] = [v.copy() for v in psi], list(alpha), list(beta)
        if b < 1e-15 or abs(new - old) < 1e-15 or k == num_iter:
            energy = new
            res = np.zeros(psi_dim, dtype=h.dtype)
            for j in range(k):
                res += psi[j] * vec[j, 0]
            lat.ket.from_vector(res)
            break
        old = new
        if k < num_iter:
            nxt = h / b
            h = -b * psi[i]
            psi.append(nxt)
            gram_schmidt(psi)
    return energy


def gram_schmidt(vecs):
    """GramSchmidtMethod (real_tensor_lanczos.cc:188-201): the newest vector
    is orthogonalised against all earlier ones, each coefficient taken from
    the already-updated vector, then normalised."""
    v = vecs[-1]
    for u in vecs[:-1]:
        c = np.vdot(u, v) if np.iscomplexobj(v) else float(np.dot(u, v))
        v -= c * u
    v /= math.sqrt(float(np.vdot(v, v).real))


# --------------------------------------------------------------------------
# Truncation / canonical forms
# --------------------------------------------------------------------------
def truncate_dims(max_dim, canpre, svals):
    """ComputeTruncateDim (real_tensor_lattice.cc:501-555).  svals is a list
    of per-sector descending singular-value arrays (modified in place when a
    sector is forced to keep one state with sigma := 0)."""
    if canpre > 1:
        canpre = 0.0
    t = [int(np.sum(s > canpre)) for s in svals]
    total = sum(t)
    if total > 0 and max_dim > 0 and max_dim < total:
        vals, blk = [], []
        for i, s in enumerate(svals):
            for j in range(t[i]):
                vals.append(float(s[j]))
                blk.append(i)
        quicksort(vals, blk, flag=0)
        t = [0] * len(svals)
        for i in range(max_dim):
            t[blk[i]] += 1
    for i in range(len(svals)):
        if t[i] == 0:
            t[i] = 1
            svals[i][0] = 0.0
    return t


def svd(A):
    """RealMatrix::SVDMatrix (real_matrix.cc:588-620; complex
    ComplexMatrix::SVDMatrix, complex_matrix.cc:601-637, right factor V^H):
    thin SVD; the right factor is returned padded to n rows (rows >=
    min(m,n) are zero)."""
    m, n = A.shape
    U, s, Vt = np.linalg.svd(A, full_matrices=False)
    sg = fix_signs(U)
    U = U * sg[None, :]
    Vt = Vt * np.conj(sg)[:, None]
    Vfull = np.zeros((n, n), dtype=A.dtype)
    Vfull[:Vt.shape[0]] = Vt
    return U, s.copy(), Vfull


def _dtype(bm):
    return np.complex128 if any(np.iscomplexobj(m) for m in bm.mats) else np.float64


def left_canonical(lat, max_dim, canpre, info=None):
    """LeftCanonicalTensorLattice (real_tensor_lattice.cc:558-676)."""
    ket = lat.ket
    nr = lat.nr
    Us, Vs, S = [None] * nr, [None] * nr, [None] * nr
    for r in range(nr):
        pos = ket.same(1, r)
        rows = sum(ket.mats[p].shape[0] for p in pos)
        cols = lat.right_dim[r]
        if rows == 0:
            Vs[r] = np.zeros((cols, cols))
            S[r] = np.zeros(max(cols, 1))
            S[r] = S[r][:1]
        else:
            A = np.zeros((rows, cols), dtype=_dtype(ket))
            acc = 0
            for p in pos:
                m = ket.mats[p]
                A[acc:acc + m.shape[0], :m.shape[1]] = m
                acc += m.shape[0]
            Us[r], S[r], Vs[r] = svd(A)
    t = truncate_dims(max_dim, canpre, S)
    if info is not None:
        info['svals'] = [s.copy() for s in S]
        info['trunc'] = list(t)
    lat.match_dim = [0] * nr
    can = []
    for r in range(nr):
        if Us[r] is not None:
            acc = 0
            for p in ket.same(1, r):
                m = ket.mats[p]
                rows = m.shape[0]
                ket.mats[p] = Us[r][acc:acc + rows, :t[r]].copy()
                acc += rows
        lat.right_dim[r] = t[r]
        ncol = Vs[r].shape[1]
        lat.match_dim[r] = ncol
        can.append(S[r][:t[r], None] * Vs[r][:t[r], :])
    lat.canonical = BlockMat(list(range(nr)), list(range(nr)), can)


def right_canonical(lat, max_dim, canpre, info=None):
    """RightCanonicalTensorLattice (real_tensor_lattice.cc:794-911)."""
    ket = lat.ket
    nl = lat.nl
    Us, Vs, S = [None] * nl, [None] * nl, [None] * nl
    for l in range(nl):
        pos = ket.same(0, l)
        rows = lat.left_dim[l]
        cols = sum(ket.mats[p].shape[1] for p in pos)
        if cols == 0:
            Us[l] = np.zeros((rows, rows))
            S[l] = np.zeros(1)
        else:
            A = np.zeros((rows, cols), dtype=_dtype(ket))
            acc = 0
            for p in pos:
                m = ket.mats[p]
                A[:m.shape[0], acc:acc + m.shape[1]] = m
                acc += m.shape[1]
            Us[l], S[l], Vs[l] = svd(A)
    t = truncate_dims(max_dim, canpre, S)
    if info is not None:
        info['svals'] = [s.copy() for s in S]
        info['trunc'] = list(t)
    lat.match_dim = [0] * nl
    can = []
    for l in range(nl):
        if Vs[l] is not None:
            acc = 0
            for p in ket.same(0, l):
                m = ket.mats[p]
                cols = m.shape[1]
                ket.mats[p] = Vs[l][:t[l], acc:acc + cols].copy()
                acc += cols
        lat.left_dim[l] = t[l]
        lat.match_dim[l] = Us[l].shape[0]
        Ut = np.zeros((Us[l].shape[0], t[l]), dtype=Us[l].dtype)
        kk = min(t[l], Us[l].shape[1])
        Ut[:, :kk] = Us[l][:, :kk] * S[l][:kk][None, :]
        can.append(Ut)
    lat.canonical = BlockMat(list(range(nl)), list(range(nl)), can)


def left_merge(lat, prev):
    """LeftMergeTensorLattice (real_tensor_lattice.cc:914-945)."""
    if prev.canonical is None:
        raise RuntimeError("Canonical tensor is nullptr in LeftMergeTensorLattice!")
    ket = lat.ket
    for l in range(lat.nl):
        C = prev.canonical.mats[l]
        for p in ket.same(0, l):
            ket.mats[p] = _mm(C, ket.mats[p])
        lat.left_dim[l] = C.shape[0]
    ket.normalize()
    prev.canonical = None
    prev.match_dim = None


def right_merge(lat, nxt):
    """RightMergeTensorLattice (real_tensor_lattice.cc:947-977)."""
    if nxt.canonical is None:
        raise RuntimeError("Canonical tensor is nullptr in RightMergeTensorLattice!")
    ket = lat.ket
    for r in range(lat.nr):
        C = nxt.canonical.mats[r]
        for p in ket.same(1, r):
            ket.mats[p] = _mm(ket.mats[p], C)
        lat.right_dim[r] = C.shape[1]
    ket.normalize()
    nxt.canonical = None
    nxt.match_dim = None


def _mm(A, B):
    """RealMatrix::MultiplyToMatrix (real_matrix.cc:291-315)."""
    if A.shape[1] != B.shape[0]:
        raise ValueError("Matrix indeces do not match in MultiplyToMatrix")
    return A @ B


# --------------------------------------------------------------------------
# Space (MPS) bookkeeping
# --------------------------------------------------------------------------
class Space:
    """RealTensorSpace + model space subclass state."""

    def __init__(self, model, rng):
        self.model = model
        self.L = model.L
        self.d = model.d
        self.tables = model.space_tables()
        self.lat = [Lattice(self.d) for _ in range(self.L)]
        self.rng = rng
        self.max_block = 0
        self.max_dim = -1
        self.canpre = 1e-30
        self.noise = 10.0
        self.expan = None
        self.tracer = None
        self.cplx = bool(getattr(model, "complex", False))

    def set_params(self, max_block, max_dim, canpre, noise):
        """DefineTensorSpaceParameter (real_tensor_space.cc:59-66)."""
        self.max_block, self.max_dim, self.canpre, self.noise = max_block, max_dim, canpre, noise

    def check(self, site, li, ri):
        return self.model.check(site, self.tables[site][li], self.tables[site + 1][ri])

    def lattice_block(self, site):
        """ComputeLatticeBlock (real_tensor_space.cc:442-504)."""
        if site == 0:
            left = [0]
        else:
            left = list(self.lat[site - 1].right_block)
        right = []
        for r in range(len(self.tables[site + 1])):
            if any(self.check(site, lb, r) != -1 for lb in left):
                right.append(r)
            if len(right) == self.max_block:
                break
        return left, right

    def tensor_index(self, site, lat=None):
        """ComputeTensorIndex (real_tensor_space.cc:506-550): row-major over
        (left position, right position) pairs passing the QN rule."""
        lat = self.lat[site] if lat is None else lat
        li, ri, pi = [], [], []
        for i, lbq in enumerate(lat.left_block):
            for j, rbq in enumerate(lat.right_block):
                p = self.check(site, lbq, rbq)
                if p != -1:
                    li.append(i)
                    ri.append(j)
                    pi.append(p)
        return li, ri, pi

    def define(self, first_block, first_dim):
        """DefineTensorSpace (real_tensor_space.cc:68-134)."""
        self.max_block = first_block
        self.max_dim = -1
        self.noise = 10.0
        self.canpre = 1e-30
        L = self.L
        for i in range(L):
            left, right = self.lattice_block(i)
            ld = [1 if i == 0 else first_dim] * len(left)
            rd = [1 if i == L - 1 else first_dim] * len(right)
            self.lat[i].define_sectors(left, right, ld, rd)
        for i in range(L):
            li, ri, pi = self.tensor_index(i)
            rows = 1 if i == 0 else first_dim
            cols = 1 if i == L - 1 else first_dim
            self.lat[i].define_blocks(li, ri, pi, rows, cols, self.rng, self.cplx)
            if self.tracer:
                self.tracer(-2, i, 0.0, "define_random")
            self.canonical(0, i)
            if self.tracer:
                self.tracer(-2, i, 0.0, "define_canon")
            if i > 0:
                self.merge(0, i)
            if self.tracer:
                self.tracer(-2, i, 0.0, "define_merge")

    def canonical(self, leigh, site, info=None):
        if leigh == 0:
            left_canonical(self.lat[site], self.max_dim, self.canpre, info)
        else:
            right_canonical(self.lat[site], self.max_dim, self.canpre, info)

    def merge(self, leigh, site):
        """MergeTensorSpace (real_tensor_space.cc:310-316)."""
        if leigh == 0 and site > 0:
            left_merge(self.lat[site], self.lat[site - 1])
        elif leigh == 1 and site < self.L - 1:
            right_merge(self.lat[site], self.lat[site + 1])

    def expansion_active(self):
        return not (self.max_block < 0 or self.max_dim < 0 or self.noise > 1)

    def expan_lattice(self, leigh, site, op_table):
        """ComputeExpanTensorLattice (real_tensor_space.cc:552-706).  leigh=1
        expands the right bond (left-to-right sweep), leigh=0 the left bond.
        Returns the mapping table [op channel][old sector position]."""
        lat = self.lat[site]
        tab = self.tables[site + leigh]
        old = lat.left_block if leigh == 0 else lat.right_block
        merge = self.model.merge
        flag = [False] * len(tab)
        merged = [[merge(oq, tab[ob]) for ob in old] for oq in op_table]
        for row in merged:
            for mq in row:
                for k in range(len(tab)):
                    if tab[k] == mq:
                        flag[k] = True
        blocks = []
        for k in range(len(tab)):
            if flag[k]:
                blocks.append(k)
                if len(blocks) == self.max_block:
                    break
        mapping = []
        for row in merged:
            mr = []
            for mq in row:
                hit = -1
                for k, b in enumerate(blocks):
                    if tab[b] == mq:
                        hit = k
                        break
                mr.append(hit)
            mapping.append(mr)
        E = Lattice(self.d)
        if leigh == 0:
            left, right = blocks, list(lat.right_block)
        else:
            left, right = list(lat.left_block), blocks
        E.define_sectors(left, right, [1] * len(left), [1] * len(right))
        li, ri, pi = self.tensor_index(site, E)
        E.define_blocks(li, ri, pi, 1, 1, self.rng, self.cplx)
        E.ket.reset()
        self.expan = E
        return mapping

    def expand(self, leigh, site):
        """ExpanTensorSpace (real_tensor_space.cc:319-387); leigh=0 means the
        left-to-right sweep (right bond grows)."""
        if not self.expansion_active():
            return
        if self.expan is None:
            raise RuntimeError("expan_tensor_lattice is nullptr!")
        origin = self.lat[site].copy()
        E = self.expan
        lat = self.lat[site]
        if leigh == 0:
            left, ld = list(origin.left_block), list(origin.left_dim)
            right, rd = combine(origin.right_block, origin.right_dim, E.right_block, E.right_dim)
        else:
            right, rd = list(origin.right_block), list(origin.right_dim)
            left, ld = combine(origin.left_block, origin.left_dim, E.left_block, E.left_dim)
        lat.define_sectors(left, right, ld, rd)
        li, ri, pi = self.tensor_index(site)
        lat.define_blocks(li, ri, pi, 1, 1, self.rng, self.cplx)
        lat.ket.reset()
        if leigh == 0:
            left_expan_fill(lat, origin, E)
        else:
            right_expan_fill(lat, origin, E)
        self.expan = None

    def match(self, leigh, site):
        """MatchTensorSpace (real_tensor_space.cc:389-440)."""
        if not self.expansion_active():
            return
        origin = self.lat[site].copy()
        lat = self.lat[site]
        if leigh == 0 and site > 0:
            prev = self.lat[site - 1]
            left, ld = list(prev.right_block), list(prev.match_dim)
            right, rd = list(origin.right_block), list(origin.right_dim)
        elif leigh == 1 and site < self.L - 1:
            nxt = self.lat[site + 1]
            left, ld = list(origin.left_block), list(origin.left_dim)
            right, rd = list(nxt.left_block), list(nxt.match_dim)
        else:
            left, ld, right, rd = [], [], [], []
        lat.define_sectors(left, right, ld, rd)
        li, ri, pi = self.tensor_index(site)
        lat.define_blocks(li, ri, pi, 1, 1, self.rng, self.cplx)
        lat.ket.reset()
        if leigh == 0 and site > 0:
            right_expan_fill(lat, origin, None)
        elif leigh == 1 and site < self.L - 1:
            left_expan_fill(lat, origin, None)


def combine(blk0, dim0, blk1, dim1):
    """CombineTensorLattice (real_tensor_lattice.cc:361-484): sorted union of
    the sector lists, dims added.  The reference's first counting loop bounds
    its inner index by the second list's length (:426); this restates the
    intended union (the two agree whenever the expansion list is a prefix of
    the flagged sectors, which ComputeExpanTensorLattice guarantees)."""
    union = list(blk0) + [b for b in blk1 if b not in blk0]
    quicksort(union, None, flag=1)
    dims = []
    for b in union:
        v = 0
        if b in blk0:
            v += dim0[blk0.index(b)]
        if b in blk1:
            v += dim1[blk1.index(b)]
        dims.append(v)
    return union, dims


def _concat(A, B, axis):
    """RealMatrix::ExpanMatrix (real_matrix.cc:533-583)."""
    if axis == 1:
        if A.shape[0] != B.shape[0]:
            raise ValueError("row or column dimension is not match in ExpanMatrix")
        return np.concatenate([A, B], axis=1)
    if A.shape[1] != B.shape[1]:
        raise ValueError("row or column dimension is not match in ExpanMatrix")
    return np.concatenate([A, B], axis=0)


def _add(A, B):
    """RealMatrix::AddToMatrix (real_matrix.cc:220-246)."""
    if B.size == 0:
        return A
    if A.size == 0:
        return B.copy()
    if A.shape != B.shape:
        raise ValueError("Matrix indeces do not match in AddToMatrix")
    return A + B


def left_expan_fill(lat, origin, E):
    """LeftExpanTensorLattice (real_tensor_lattice.cc:979-1063)."""
    ket = lat.ket
    for r in range(lat.nr):
        e0 = origin.right_block.index(lat.right_block[r]) if lat.right_block[r] in origin.right_block else -1
        e1 = -1
        if E is not None and lat.right_block[r] in E.right_block:
            e1 = E.right_block.index(lat.right_block[r])
        for l in range(lat.nl):
            pos = ket.find(l, r)
            if pos == -1:
                continue
            m = ket.mats[pos]
            if e0 != -1:
                p0 = origin.ket.find(l, e0)
                if p0 != -1:
                    m = _add(m, origin.ket.mats[p0])
            if e1 != -1:
                p1 = E.ket.find(l, e1)
                if p1 != -1:
                    x = E.ket.mats[p1]
                    m = _add(m, x) if m.shape[0] == 0 else _concat(m, x, 1)
            if e0 != -1 and e1 == -1:
                extra = lat.right_dim[r] - m.shape[1]
                if extra < 0:
                    raise ValueError("expan_dim < 0 in LeftExpanTensorLattice")
                m = _concat(m, np.zeros((lat.left_dim[l], extra)), 1)
            if e0 == -1 and e1 == -1:
                m = _add(m, np.zeros((lat.left_dim[l], lat.right_dim[r])))
            ket.mats[pos] = m


def right_expan_fill(lat, origin, E):
    """RightExpanTensorLattice (real_tensor_lattice.cc:1066-1149)."""
    ket = lat.ket
    for l in range(lat.nl):
        e0 = origin.left_block.index(lat.left_block[l]) if lat.left_block[l] in origin.left_block else -1
        e1 = -1
        if E is not None and lat.left_block[l] in E.left_block:
            e1 = E.left_block.index(lat.left_block[l])
        for r in range(lat.nr):
            pos = ket.find(l, r)
            if pos == -1:
                continue
            m = ket.mats[pos]
            if e0 != -1:
                p0 = origin.ket.find(e0, r)
                if p0 != -1:
                    m = _add(m, origin.ket.mats[p0])
            if e1 != -1:
                p1 = E.ket.find(e1, r)
                if p1 != -1:
                    x = E.ket.mats[p1]
                    m = _add(m, x) if m.shape[0] == 0 else _concat(m, x, 0)
            if e0 != -1 and e1 == -1:
                extra = lat.left_dim[l] - m.shape[0]
                if extra < 0:
                    raise ValueError("expan_dim < 0 in RightExpanTensorLattice")
                m = _concat(m, np.zeros((extra, lat.right_dim[r])), 0)
            if e0 == -1 and e1 == -1:
                m = _add(m, np.zeros((lat.left_dim[l], lat.right_dim[r])))
            ket.mats[pos] = m


# --------------------------------------------------------------------------
# Expansion contractions
# --------------------------------------------------------------------------
def _expan_width(D, d):
    """expan_dim clamp (real_tensor_contraction.cc:203-206, :383-386)."""
    w = int(math.sqrt(D)) * d * d
    return min(max(w, 1), 25)


def left_expan_contraction(lat, W, Lenv, E, mapping, noise, rng):
    """LeftExpanTensorContraction (real_tensor_contraction.cc:169-459;
    complex_tensor_contraction.cc:169-470)."""
    cplx = np.iscomplexobj(W)
    lb, rb = W.shape[0], W.shape[1]
    ket = lat.ket
    same_l = [ket.same(0, l) for l in range(lat.nl)]
    exist = [False] * E.nr
    for o1 in range(lb):
        Lb = Lenv[o1]
        for a in range(Lb.n):
            idx, kdx = Lb.left[a], Lb.right[a]
            if Lb.mats[a].shape[0] == 0:
                continue
            for kp in same_l[kdx]:
                jdx = ket.right[kp]
                p2 = lat.phys[kdx][jdx]
                if not mixed_check_zero(W, o1, p2):
                    continue
                for o2 in range(rb):
                    ldx = mapping[o2][jdx]
                    pos = E.ket.find(idx, ldx)
                    if ldx != -1 and pos != -1:
                        p1 = E.phys[idx][ldx]
                        if W[o1, o2, p1, p2] != 0.0:
                            exist[ldx] = True
    keep = [i for i in range(E.nr) if exist[i]]
    leigh_block = [E.right_block[i] for i in keep]
    phys = [[E.phys[j][i] for i in keep] for j in range(E.nl)]
    etab = []
    for o2 in range(rb):
        row = []
        for r in range(lat.nr):
            ldx = mapping[o2][r]
            hit = -1
            if ldx != -1:
                for i, b in enumerate(leigh_block):
                    if E.right_block[ldx] == b:
                        hit = i
                        break
            row.append(hit)
        etab.append(row)
    E.right_block = leigh_block
    E.left_dim = list(lat.left_dim)
    E.right_dim = [0] * len(leigh_block)
    E.phys = phys
    li, ri = [], []
    for i in range(E.nl):
        for j in range(E.nr):
            if E.phys[i][j] != -1:
                li.append(i)
                ri.append(j)
    E.ket = BlockMat(li, ri)
    d = lat.d
    for o1 in range(lb):
        Lb = Lenv[o1]
        for a in range(Lb.n):
            idx, kdx = Lb.left[a], Lb.right[a]
            Lm = Lb.mats[a]
            if Lm.shape[0] == 0:
                continue
            for kp in same_l[kdx]:
                jdx = ket.right[kp]
                p2 = lat.phys[kdx][jdx]
                if not mixed_check_zero(W, o1, p2):
                    continue
                first = _mm(Lm, ket.mats[kp])
                e0 = first.shape[1]
                e1 = _expan_width(e0, d)
                for o2 in range(rb):
                    ldx = etab[o2][jdx]
                    pos = E.ket.find(idx, ldx)
                    if ldx != -1 and pos != -1:
                        p1 = E.phys[idx][ldx]
                        w = W[o1, o2, p1, p2]
                        if w != 0.0:
                            rnd = rng.complex_matrix(e0, e1) if cplx else rng.random_matrix(e0, e1)
                            for t in range(min(e0, e1)):
                                rnd[t, t] = 0.1 * w
                            sec = _mm(first, rnd)
                            cur = E.ket.mats[pos]
                            E.ket.mats[pos] = _add(cur, sec) if cur.shape[0] == 0 else _concat(cur, sec, 1)
    for r in range(E.nr):
        E.right_dim[r] = 0
        for l in range(E.nl):
            pos = E.ket.find(l, r)
            if pos != -1:
                E.right_dim[r] = max(E.right_dim[r], E.ket.mats[pos].shape[1])
        for l in range(E.nl):
            pos = E.ket.find(l, r)
            if pos != -1 and E.ket.mats[pos].shape[0] == 0:
                if cplx:  # complex_tensor_contraction.cc:446-451
                    E.ket.mats[pos] = np.zeros((E.left_dim[l], E.right_dim[r]), dtype=complex)
                    continue
                # the reference sizes this zero block by left_dim_[r] (:448)
                if r >= len(E.left_dim):
                    raise IndexError("zero expansion block with r >= num_left_block (reference reads out of range)")
                E.ket.mats[pos] = _add(E.ket.mats[pos], np.zeros((E.left_dim[r], E.right_dim[r])))
            elif pos != -1 and cplx:
                # complex pads short blocks with zero columns (:453-462)
                m = E.ket.mats[pos]
                extra = E.right_dim[r] - m.shape[1]
                if extra > 0:
                    E.ket.mats[pos] = _concat(m, np.zeros((m.shape[0], extra), dtype=complex), 1)
    E.ket.mats = [m * noise for m in E.ket.mats]


def right_expan_contraction(lat, W, Renv, E, mapping, noise, rng):
    """RightExpanTensorContraction (real_tensor_contraction.cc:477-765;
    complex_tensor_contraction.cc:488-786)."""
    cplx = np.iscomplexobj(W)
    lb, rb = W.shape[0], W.shape[1]
    ket = lat.ket
    same_r = [ket.same(1, r) for r in range(lat.nr)]
    exist = [False] * E.nl
    for o2 in range(rb):
        Rb = Renv[o2]
        for a in range(Rb.n):
            kdx, idx = Rb.left[a], Rb.right[a]
            if Rb.mats[a].shape[0] == 0:
                continue
            for kp in same_r[kdx]:
                jdx = ket.left[kp]
                p2 = lat.phys[jdx][kdx]
                if not right_check_zero(W, o2, p2):
                    continue
                for o1 in range(lb):
                    ldx = mapping[o1][jdx]
                    pos = E.ket.find(ldx, idx)
                    if ldx != -1 and pos != -1:
                        p1 = E.phys[ldx][idx]
                        if W[o1, o2, p1, p2] != 0.0:
                            exist[ldx] = True
    keep = [i for i in range(E.nl) if exist[i]]
    leigh_block = [E.left_block[i] for i in keep]
    phys = [list(E.phys[i]) for i in keep]
    etab = []
    for o1 in range(lb):
        row = []
        for l in range(lat.nl):
            ldx = mapping[o1][l]
            hit = -1
            if ldx != -1:
                for i, b in enumerate(leigh_block):
                    if E.left_block[ldx] == b:
                        hit = i
                        break
            row.append(hit)
        etab.append(row)
    E.left_block = leigh_block
    E.right_dim = list(lat.right_dim)
    E.left_dim = [0] * len(leigh_block)
    E.phys = phys
    li, ri = [], []
    for i in range(E.nl):
        for j in range(E.nr):
            if E.phys[i][j] != -1:
                li.append(i)
                ri.append(j)
    E.ket = BlockMat(li, ri)
    d = lat.d
    for o2 in range(rb):
        Rb = Renv[o2]
        for a in range(Rb.n):
            kdx, idx = Rb.left[a], Rb.right[a]
            Rm = Rb.mats[a]
            if Rm.shape[0] == 0:
                continue
            for kp in same_r[kdx]:
                jdx = ket.left[kp]
                p2 = lat.phys[jdx][kdx]
                if not right_check_zero(W, o2, p2):
                    continue
                first = _mm(ket.mats[kp], Rm)
                e1 = first.shape[0]
                e0 = _expan_width(e1, d)
                for o1 in range(lb):
                    ldx = etab[o1][jdx]
                    pos = E.ket.find(ldx, idx)
                    if ldx != -1 and pos != -1:
                        p1 = E.phys[ldx][idx]
                        w = W[o1, o2, p1, p2]
                        if w != 0.0:
                            rnd = rng.complex_matrix(e0, e1) if cplx else rng.random_matrix(e0, e1)
                            for t in range(min(e0, e1)):
                                rnd[t, t] = 0.1 * w
                            sec = _mm(rnd, first)
                            cur = E.ket.mats[pos]
                            E.ket.mats[pos] = _add(cur, sec) if cur.shape[0] == 0 else _concat(cur, sec, 0)
    for l in range(E.nl):
        E.left_dim[l] = 0
        for r in range(E.nr):
            pos = E.ket.find(l, r)
            if pos != -1:
                E.left_dim[l] = max(E.left_dim[l], E.ket.mats[pos].shape[0])
        for r in range(E.nr):
            pos = E.ket.find(l, r)
            if pos != -1 and E.ket.mats[pos].shape[0] == 0:
                if cplx:  # complex_tensor_contraction.cc:762-767
                    E.ket.mats[pos] = np.zeros((E.left_dim[l], E.right_dim[r]), dtype=complex)
                    continue
                # the reference sizes this zero block by left_dim_[r] (:754)
                if r >= len(E.left_dim):
                    raise IndexError("zero expansion block with r >= num_left_block (reference reads out of range)")
                E.ket.mats[pos] = _add(E.ket.mats[pos], np.zeros((E.left_dim[r], E.right_dim[r])))
            elif pos != -1 and cplx:
                # complex pads short blocks with zero rows (:769-778)
                m = E.ket.mats[pos]
                extra = E.left_dim[l] - m.shape[0]
                if extra > 0:
                    E.ket.mats[pos] = _concat(m, np.zeros((extra, m.shape[1]), dtype=complex), 0)
    E.ket.mats = [m * noise for m in E.ket.mats]


# --------------------------------------------------------------------------
# Sweep driver
# --------------------------------------------------------------------------
def sweep_table(rows, num_sweep):
    """RealTensorRundmrg::Sweep (real_tensor_rundmrg.cc:70-113).  rows are
    (maxm, maxb, niter, canpre, noise) per half sweep; a maxm of 0 ends the
    table and later rows repeat the last one, with the reference's fill-in
    (max_dim := max_block of the last row, :94).  Noise becomes 0.1**noise."""
    n = 2 * num_sweep
    tab = [list(r) for r in rows[:n]]
    nlast = n
    for i, r in enumerate(tab):
        if r[0] == 0:
            nlast = i - 1
            break
    if len(tab) < n and nlast == n:
        nlast = len(tab) - 1
    out = tab[:nlast + 1]
    while len(out) < n:
        last = tab[nlast]
        out.append([last[1], last[1], last[2], last[3], last[4]])
    return [(int(a), int(b), int(c), float(e), 0.1 ** float(f)) for a, b, c, e, f in out]


class DMRG:
    """RealTensorRundmrg + RealTensorNetwork + RealTensorLanczos."""

    def __init__(self, model, sweeps, num_sweep, first_block=20, first_dim=10, seed=1, trace=False,
                 resume=None):
        self.model = model
        self.rng = GlibcRand(seed)
        self.space = Space(model, self.rng)
        self.W, self.op_tables = model.mpo()
        self.L = model.L
        self.table = sweep_table(sweeps, num_sweep)
        self.trace = trace
        if trace:
            self.space.tracer = self._trace
        if resume is not None:  # cache_resume (real_tensor_rundmrg.cc:44-52)
            from .io import read_space
            read_space(resume, self.space)
        else:
            self.space.define(first_block, first_dim)
        # tensor_contraction_[site] holds L (left_bond of W[site]) and R
        # (right_bond of W[site]) (real_tensor_network.cc:3-33)
        self.Lenv = [None] * (self.L + 1)
        self.Renv = [None] * (self.L + 1)
        self.energies = []
        self.trunc_err = []
        self.stats = {}

    def initialize(self):
        """Initialize (real_tensor_rundmrg.cc:247-287)."""
        L = self.L
        self.Lenv[0] = boundary_env(self.W[0].shape[0])
        self.Renv[L - 1] = boundary_env(self.W[L - 1].shape[1])
        sp = self.space
        for i in range(L - 1, 0, -1):
            sp.canonical(1, i)
            if self.trace:
                self._trace(-1, i, 0.0, "init_canon")
            self.Renv[i - 1] = right_env(sp.lat[i], self.W[i], self.Renv[i])
            sp.merge(1, i - 1)
            if self.trace:
                self._trace(-1, i - 1, 0.0, "init_merge")
        sp.canonical(1, 0)

    def _trace(self, h, j, e, what):
        lat = self.space.lat[j]
        v = lat.ket.to_vector()
        print("TRACE %d %d %-17s E=%.12f nl=%d nr=%d ld=%d rd=%d nblk=%d size=%d rng=%d nrm2=%.12e" % (
            h, j, what, e if e is not None else 0.0, lat.nl, lat.nr, sum(lat.left_dim), sum(lat.right_dim),
            lat.ket.n, lat.ket.dim(), self.rng.count, float(np.vdot(v, v).real)), flush=True)

    def _discarded(self, info):
        s2 = 0.0
        for s, t in zip(info['svals'], info['trunc']):
            s2 += float(np.sum(np.asarray(s[t:]) ** 2))
        return s2

    def run(self, verbose=False, stop=None):
        """Run (real_tensor_rundmrg.cc:115-245), without disk cache.

        stop=(h, j) returns just before the Lanczos solve of site j in half
        sweep h (the MPS, environments and RNG are then consistent for j)."""
        t0 = time.time()
        self.initialize()
        sp = self.space
        L = self.L
        energy = None
        for s in range(len(self.table) // 2):
            maxm, maxb, niter, canpre, noise = self.table[2 * s]
            sp.set_params(maxb, maxm, canpre, noise)
            terr = 0.0
            for j in range(L - 1):
                if stop == (2 * s, j):
                    return self.energies
                energy = lanczos(sp.lat[j], self.W[j], self.Lenv[j], self.Renv[j], niter, self.stats)
                tr = self._trace if self.trace else (lambda *a: None)
                tr(2 * s, j, energy, "lanczos")
                self.Renv[j] = [BlockMat(b.left, b.right) for b in self.Renv[j]]
                mapping = sp.expan_lattice(1, j, self.op_tables[j + 1])
                left_expan_contraction(sp.lat[j], self.W[j], self.Lenv[j], sp.expan, mapping, sp.noise, self.rng)
                tr(2 * s, j, energy, "expan_contraction")
                sp.expand(0, j)
                tr(2 * s, j, energy, "expand")
                info = {}
                sp.canonical(0, j, info)
                tr(2 * s, j, energy, "canonical")
                terr = max(terr, self._discarded(info))
                self.Lenv[j + 1] = left_env(sp.lat[j], self.W[j], self.Lenv[j])
                sp.match(0, j + 1)
                tr(2 * s, j + 1, energy, "match")
                sp.merge(0, j + 1)
                tr(2 * s, j + 1, energy, "merge")
            self.energies.append(energy)
            self.trunc_err.append(terr)
            if verbose:
                print("right energy = %.12f" % energy, flush=True)
            maxm, maxb, niter, canpre, noise = self.table[2 * s + 1]
            sp.set_params(maxb, maxm, canpre, noise)
            terr = 0.0
            for j in range(L - 1, 0, -1):
                if stop == (2 * s + 1, j):
                    return self.energies
                energy = lanczos(sp.lat[j], self.W[j], self.Lenv[j], self.Renv[j], niter, self.stats)
                tr = self._trace if self.trace else (lambda *a: None)
                tr(2 * s + 1, j, energy, "lanczos")
                self.Lenv[j] = [BlockMat(b.left, b.right) for b in self.Lenv[j]]
                mapping = sp.expan_lattice(0, j, self.op_tables[j])
                right_expan_contraction(sp.lat[j], self.W[j], self.Renv[j], sp.expan, mapping, sp.noise, self.rng)
                tr(2 * s + 1, j, energy, "expan_contraction")
                sp.expand(1, j)
                tr(2 * s + 1, j, energy, "expand")
                info = {}
                sp.canonical(1, j, info)
                tr(2 * s + 1, j, energy, "canonical")
                terr = max(terr, self._discarded(info))
                self.Renv[j - 1] = right_env(sp.lat[j], self.W[j], self.Renv[j])
                sp.match(1, j - 1)
                tr(2 * s + 1, j - 1, energy, "match")
                sp.merge(1, j - 1)
                tr(2 * s + 1, j - 1, energy, "merge")
            self.energies.append(energy)
            self.trunc_err.append(terr)
            if verbose:
                print("left energy = %.12f" % energy, flush=True)
        self.total_time = time.time() - t0
        return self.energies
