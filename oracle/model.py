"""Real spinless Bose-Hubbard model on a square lattice, oracle copy.

Restates
* model/space/real_spinless_bose_square_space.cc:26-83 (QN tables, QN rule),
* src/dmrg/realdmrg/real_tensor_space.cc:716-736 (ReorderQuantumTable),
* model/hamiltonian/real_spinless_bose_square_hamiltonian.cc:41-228 (terms),
* src/dmrg/realdmrg/real_tensor_hamiltonian.cc:151-254 (incremental MPO),
* src/dmrg/realdmrg/real_tensor_operator.cc:141-409 (expansion,
  deparallelisation, zero checks),
* src/tensor/real_matrix.cc:438-531 + src/util/general.h:146-180
  (ParallelMatrix / isParallelElement).

The MPO site tensor W has shape (left_bond, right_bond, d, d) with
W[o1, o2, p1, p2], p1 = bra (output) index, p2 = ket (input) index.
"""
import math

import numpy as np

from .sort import quicksort


def reorder_table(table, mean):
    """ReorderQuantumTable (real_tensor_space.cc:716-736): ascending QuickSort
    of sum_j (q_j - mean_j)^2 with an index permutation."""
    dev = [sum((q[j] - mean[j]) ** 2 for j in range(len(q))) for q in table]
    idx = list(range(len(table)))
    quicksort(dev, idx, flag=1)
    return [table[i] for i in idx]


def parallel_matrix(mat, leigh):
    """RealMatrix::ParallelMatrix (real_matrix.cc:438-531).

    leigh=0 works on columns, leigh=1 on rows.  Returns (positions,
    transfer) with positions the kept (unparallel) column/row indices and
    transfer the (n x n) prefactor matrix: for leigh=0,
    mat[:, c] = sum_k mat[:, positions[k]] * transfer[k, c]."""
    a = mat if leigh == 0 else mat.T
    ncol = a.shape[1]
    pos = [0]
    transfer = np.zeros((ncol, ncol))
    transfer[0, 0] = 1.0
    for c in range(ncol):
        col = a[:, c]
        found = -1
        pref = 0.0
        for k, pk in enumerate(pos):
            ref = a[:, pk]
            refnz = np.abs(ref) > 1e-10
            # isParallelElement: a row with |ref|<=1e-10 and |col|>1e-10 breaks it
            if np.any(~refnz & (np.abs(col) > 1e-10)):
                continue
            if np.any(refnz):
                fac = col[refnz] / ref[refnz]
                p0 = fac[0]
                if np.any(np.abs(fac - p0) > 1e-8):
                    continue
                pref = p0
            else:
                pref = 0.0
            found = k
            break
        if found >= 0:
            transfer[found, c] = pref
        else:
            pos.append(c)
            transfer[len(pos) - 1, c] = 1.0
    if leigh == 1:
        transfer = transfer.T.copy()
    return pos, transfer


def expand_operator(W, op, leigh, d):
    """RealTensorOperator::ExpanTensorOperator (real_tensor_operator.cc:141-197)."""
    lb, rb = W.shape[0], W.shape[1]
    if lb == 0 and rb == 0:
        out = np.zeros((1, 1, d, d))
        out[0, 0] = op
        return out
    e0 = 0 if leigh == 0 else 1
    e1 = 0 if leigh == 1 else 1
    out = np.zeros((lb + e0, rb + e1, d, d))
    out[:lb, :rb] = W
    out[e0 * lb, e1 * rb] = op
    return out


def left_parallel(W, d):
    """LeftParallelTensorOperator (real_tensor_operator.cc:199-236).
    Rows of the reshape are (l, p1, p2) row-major, columns r."""
    lb, rb = W.shape[:2]
    resh = W.transpose(0, 2, 3, 1).reshape(lb * d * d, rb)
    pos, T = parallel_matrix(resh, 0)
    Wn = W[:, pos, :, :].copy()
    return Wn, pos, T


def right_parallel(W, d):
    """RightParallelTensorOperator (real_tensor_operator.cc:238-275).
    Columns of the reshape are (r, p1, p2) row-major, rows l."""
    lb, rb = W.shape[:2]
    resh = W.reshape(lb, rb * d * d)
    pos, T = parallel_matrix(resh, 1)
    Wn = W[pos, :, :, :].copy()
    return Wn, pos, T


def left_merge(W, npos, T):
    """LeftMergeTensorOperator (real_tensor_operator.cc:277-318):
    W'[l] = sum_m T[l, m] W[m], l < npos."""
    if T.shape[1] != W.shape[0]:
        raise ValueError("merge dimension is not match in LeftMergeTensorOperator")
    return np.einsum('lm,mrab->lrab', T[:npos, :], W)


def right_merge(W, npos, T):
    """RightMergeTensorOperator (real_tensor_operator.cc:320-361):
    W'[:, r] = sum_m T[m, r] W[:, m], r < npos."""
    if T.shape[0] != W.shape[1]:
        raise ValueError("merge dimension is not match in RightMergeTensorOperator")
    return np.einsum('mr,lmab->lrab', T[:, :npos], W)


class MPOBuilder:
    """RealTensorHamiltonian's incremental MPO construction
    (real_tensor_hamiltonian.cc:151-254)."""

    def __init__(self, num_site, d, num_quantum, basic):
        self.L = num_site
        self.d = d
        self.nq = num_quantum
        self.basic = basic
        self.W = [np.zeros((0, 0, d, d)) for _ in range(num_site)]
        self.qt = [[] for _ in range(num_site + 1)]
        self.qt[0] = [tuple([0] * num_quantum)]
        self.qt[num_site] = [tuple([0] * num_quantum)]

    def expan(self, site, op_index, coef, table):
        """ExpanTensorHamiltonian (real_tensor_hamiltonian.cc:151-187)."""
        leigh = -1
        if site == 0:
            leigh = 0
        if site == self.L - 1:
            leigh = 1
        self.W[site] = expand_operator(self.W[site], coef * self.basic[op_index], leigh, self.d)
        if site < self.L - 1:
            self.qt[site + 1].append(tuple(table))

    def parallel(self):
        """ParallelTensorHamiltonian (real_tensor_hamiltonian.cc:189-214)."""
        L, d = self.L, self.d
        for i in range(L - 1):
            self.W[i], pos, T = left_parallel(self.W[i], d)
            self.W[i + 1] = left_merge(self.W[i + 1], len(pos), T)
            self.qt[i + 1] = [self.qt[i + 1][p] for p in pos]
        for i in range(L - 1, 0, -1):
            self.W[i], pos, T = right_parallel(self.W[i], d)
            self.W[i - 1] = right_merge(self.W[i - 1], len(pos), T)
            self.qt[i] = [self.qt[i][p] for p in pos]


class SpinlessBoseSquare:
    """RealSpinlessBoseSquareSpace + RealSpinlessBoseSquareHamiltonian."""

    def __init__(self, nx, ny, num_boson, d, jx=-1.0, jy=-1.0, U=0.0, mu=0.0,
                 period_x=False, period_y=False):
        self.nx, self.ny, self.N, self.d = nx, ny, num_boson, d
        self.jx, self.jy, self.U, self.mu = jx, jy, U, mu
        self.period_x, self.period_y = period_x, period_y
        self.L = nx * ny
        self.nq = 1

    # ---- space (model/space/real_spinless_bose_square_space.cc) ----
    def space_tables(self):
        """DefineQuantumTable (real_spinless_bose_square_space.cc:26-62)."""
        L, N, d = self.L, self.N, self.d
        tables = []
        for i in range(L + 1):
            if i == 0:
                lo, hi, mean = 0, 0, 0.0
            else:
                lo = max(0, N - (L - i) * (d - 1))
                hi = min((d - 1) * i, N)
                mean = float(N) / L * i
            t = [(q,) for q in range(lo, hi + 1)]
            tables.append(reorder_table(t, [mean]))
        return tables

    def check(self, site, lq, rq):
        """CheckQuantumTable (real_spinless_bose_square_space.cc:72-83):
        physical index = right QN - left QN if in [0, d)."""
        for i in range(self.d):
            if lq[0] + i == rq[0]:
                return i
        return -1

    @staticmethod
    def merge(opq, spq):
        """MergeQuantumTable (real_spinless_bose_square_space.cc:64-70)."""
        return (opq[0] + spq[0],)

    # ---- Hamiltonian (model/hamiltonian/real_spinless_bose_square_hamiltonian.cc) ----
    def basic_ops(self):
        """DefineBasicTensor (:41-64): 0=I, 1=a^+, 2=a^-, 3=n,
        4=mu*n + U*n(n-1) (no factor 1/2)."""
        d = self.d
        ops = [np.zeros((d, d)) for _ in range(5)]
        for i in range(d):
            ops[0][i, i] = 1.0
            ops[3][i, i] = i
            ops[4][i, i] = self.mu * i + self.U * i * (i - 1)
        for i in range(1, d):
            ops[1][i, i - 1] = math.sqrt(i)
            ops[2][i - 1, i] = math.sqrt(i)
        return ops

    def _hop1(self, b, p1, p2, coef):
        """DefineHamiltonian1 (:67-106): a^+_{p1} a^-_{p2}, channel QN +1."""
        if p1 >= p2:
            raise ValueError("position1 >= position2 in DefineHamiltonian1")
        for i in range(self.L):
            op, c, t = 0, 1.0, 0
            if i == p1:
                op, c, t = 1, 1.0, 1
            if p1 < i < p2:
                op, c, t = 0, 1.0, 1
            if i == p2:
                op, c, t = 2, coef, 0
            b.expan(i, op, c, [t])

    def _hop2(self, b, p1, p2, coef):
        """DefineHamiltonian2 (:108-147): a^-_{p1} a^+_{p2}, channel QN -1."""
        if p1 >= p2:
            raise ValueError("position1 >= position2 in DefineHamiltonian2")
        for i in range(self.L):
            op, c, t = 0, 1.0, 0
            if i == p1:
                op, c, t = 2, 1.0, -1
            if p1 < i < p2:
                op, c, t = 0, 1.0, -1
            if i == p2:
                op, c, t = 1, coef, 0
            b.expan(i, op, c, [t])

    def _onsite(self, b, p, coef):
        """DefineHamiltonian3 (:149-169)."""
        for i in range(self.L):
            op, c = (4, coef) if i == p else (0, 1.0)
            b.expan(i, op, c, [0])

    def mpo(self):
        """DefineTensorHamiltonian (:171-228): returns (W list, MPO bond
        QN tables).  Site p = j + i*ny (row by row)."""
        nx, ny = self.nx, self.ny
        b = MPOBuilder(self.L, self.d, 1, self.basic_ops())
        for i in range(nx):
            for j in range(ny):
                p0 = j + i * ny
                if self.jx != 0:
                    p1 = j + ((i + 1) % nx) * ny
                    if i + 1 < nx:
                        self._hop1(b, p0, p1, self.jx)
                        self._hop2(b, p0, p1, self.jx)
                    elif nx > 2 and self.period_x:
                        self._hop1(b, p1, p0, self.jx)
                        self._hop2(b, p1, p0, self.jx)
                if self.jy != 0:
                    p1 = (j + 1) % ny + i * ny
                    if j + 1 < ny:
                        self._hop1(b, p0, p1, self.jy)
                        self._hop2(b, p0, p1, self.jy)
                    elif ny > 2 and self.period_y:
                        self._hop1(b, p1, p0, self.jy)
                        self._hop2(b, p1, p0, self.jy)
                if self.mu != 0 or self.U != 0:
                    self._onsite(b, p0, 1.0)
                b.parallel()
        return b.W, b.qt

    def bonds(self):
        """List of hopping bonds (p, q, amplitude) as the MPO encodes them."""
        out = []
        nx, ny = self.nx, self.ny
        for i in range(nx):
            for j in range(ny):
                p0 = j + i * ny
                if self.jx != 0:
                    p1 = j + ((i + 1) % nx) * ny
                    if i + 1 < nx or (nx > 2 and self.period_x):
                        out.append((min(p0, p1), max(p0, p1), self.jx))
                if self.jy != 0:
                    p1 = (j + 1) % ny + i * ny
                    if j + 1 < ny or (ny > 2 and self.period_y):
                        out.append((min(p0, p1), max(p0, p1), self.jy))
        return out


def left_check_zero(W, l, p1):
    """LeftCheckZero (real_tensor_operator.cc:363-377): True if any
    |W[l, r, p1, p2]| > 1e-15."""
    return bool(np.any(np.abs(W[l, :, p1, :]) > 1e-15))


def right_check_zero(W, r, p2):
    """RightCheckZero (real_tensor_operator.cc:379-393)."""
    return bool(np.any(np.abs(W[:, r, :, p2]) > 1e-15))


def mixed_check_zero(W, l, p2):
    """MixedCheckZero (real_tensor_operator.cc:395-409)."""
    return bool(np.any(np.abs(W[l, :, :, p2]) > 1e-15))
