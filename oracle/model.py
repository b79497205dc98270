"""Bose-Hubbard models on a square lattice, oracle copy: real spinless, its
complex twin, and complex two-component (spinful) bosons with flux.

Restates
* model/space/real_spinless_bose_square_space.cc:26-83 (QN tables, QN rule),
* src/dmrg/realdmrg/real_tensor_space.cc:716-736 (ReorderQuantumTable),
* model/hamiltonian/real_spinless_bose_square_hamiltonian.cc:41-228 (terms),
* src/dmrg/realdmrg/real_tensor_hamiltonian.cc:151-254 (incremental MPO),
* src/dmrg/realdmrg/real_tensor_operator.cc:141-409 (expansion,
  deparallelisation, zero checks),
* src/tensor/real_matrix.cc:438-531 + src/util/general.h:146-180
  (ParallelMatrix / isParallelElement; complex: complex_matrix.cc:451-540,
  general.h:182-207),
* model/space/complex_spinful_bose_square_space.cc:28-125 and
  model/hamiltonian/complex_spinful_bose_square_hamiltonian.cc:45-350.

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


def _norm(z):
    """Norm (src/tensor/complex.h:112-115): sqrt(re^2 + im^2); |x| for reals."""
    if np.iscomplexobj(z):
        return np.sqrt(z.real * z.real + z.imag * z.imag)
    return np.abs(z)


def _quot(a, b):
    """operator/ of src/tensor/complex.h (a * conj(b) / |b|^2, written out);
    plain division for reals."""
    if not (np.iscomplexobj(a) or np.iscomplexobj(b)):
        return a / b
    a, b = np.asarray(a, dtype=complex), np.asarray(b, dtype=complex)
    k = 1.0 / (b.real * b.real + b.imag * b.imag)
    return ((a.real * b.real + a.imag * b.imag) * k) + 1j * ((a.imag * b.real - a.real * b.imag) * k)


def parallel_matrix(mat, leigh):
    """RealMatrix::ParallelMatrix (real_matrix.cc:438-531).

    leigh=0 works on columns, leigh=1 on rows.  Returns (positions,
    transfer) with positions the kept (unparallel) column/row indices and
    transfer the (n x n) prefactor matrix: for leigh=0,
    mat[:, c] = sum_k mat[:, positions[k]] * transfer[k, c]."""
    a = mat if leigh == 0 else mat.T
    ncol = a.shape[1]
    pos = [0]
    transfer = np.zeros((ncol, ncol), dtype=a.dtype)
    transfer[0, 0] = 1.0
    for c in range(ncol):
        col = a[:, c]
        found = -1
        pref = 0.0
        for k, pk in enumerate(pos):
            ref = a[:, pk]
            refnz = _norm(ref) > 1e-10
            # isParallelElement: a row with |ref|<=1e-10 and |col|>1e-10 breaks it
            if np.any(~refnz & (_norm(col) > 1e-10)):
                continue
            if np.any(refnz):
                fac = _quot(col[refnz], ref[refnz])
                p0 = fac[0]
                if np.any(_norm(fac - p0) > 1e-8):
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
    dt = np.result_type(W.dtype, op.dtype)
    if lb == 0 and rb == 0:
        out = np.zeros((1, 1, d, d), dtype=dt)
        out[0, 0] = op
        return out
    e0 = 0 if leigh == 0 else 1
    e1 = 0 if leigh == 1 else 1
    out = np.zeros((lb + e0, rb + e1, d, d), dtype=dt)
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

    def __init__(self, num_site, d, num_quantum, basic, cplx=False):
        self.L = num_site
        self.d = d
        self.nq = num_quantum
        self.basic = [np.asarray(b, dtype=complex) for b in basic] if cplx else basic
        self.W = [np.zeros((0, 0, d, d), dtype=complex if cplx else float) for _ in range(num_site)]
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
                 period_x=False, period_y=False, complex=False):
        self.nx, self.ny, self.N, self.d = nx, ny, num_boson, d
        self.jx, self.jy, self.U, self.mu = jx, jy, U, mu
        self.period_x, self.period_y = period_x, period_y
        self.L = nx * ny
        self.nq = 1
        # complex=True: ComplexSpinlessBoseSquare{Space,Hamiltonian}, the same
        # model on Complex storage (model/*/complex_spinless_bose_square_*)
        self.complex = complex

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
        b = MPOBuilder(self.L, self.d, 1, self.basic_ops(), self.complex)
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


class SpinfulBoseSquare:
    """ComplexSpinfulBoseSquareSpace + ComplexSpinfulBoseSquareHamiltonian.

    Lattice site p = j + i*ny becomes MPS sites 2p (species 0, up) and 2p+1
    (species 1, down), each of local dimension d; QNs are (N_up, N_down).
    Scalars given for jx, jy, flux, mu apply to both species."""

    def __init__(self, nx, ny, n_up, n_down, d, flux=0.0, jx=-1.0, jy=-1.0, U_uu=0.0, U_dd=0.0,
                 U_ud=0.0, mu=0.0, period_x=False, period_y=False):
        pair = lambda v: (float(v[0]), float(v[1])) if np.ndim(v) else (float(v), float(v))
        self.nx, self.ny, self.d = nx, ny, d
        self.N = (n_up, n_down)
        self.flux, self.jx, self.jy, self.mu = pair(flux), pair(jx), pair(jy), pair(mu)
        self.inter = (float(U_uu), float(U_dd), float(U_ud))
        self.period_x, self.period_y = period_x, period_y
        self.L = 2 * nx * ny
        self.nq = 2
        self.complex = True

    # ---- space (model/space/complex_spinful_bose_square_space.cc) ----
    def space_tables(self):
        """DefineQuantumTable (:28-96): bond i has seen (i+1)//2 up and i//2
        down sites; the reachable rectangle of (N_up, N_down), N_up fastest,
        then ReorderQuantumTable on both QNs."""
        L, d = self.L, self.d
        half = L // 2
        tables = []
        for i in range(L + 1):
            if i == 0:
                lo, hi, mean = [0, 0], [0, 0], [0.0, 0.0]
            else:
                pos = [(i + 1) // 2, i // 2]
                lo = [max(0, self.N[t] - (half - pos[t]) * (d - 1)) for t in range(2)]
                hi = [min((d - 1) * pos[t], self.N[t]) for t in range(2)]
                mean = [float(self.N[t]) / half * pos[t] for t in range(2)]
            t = [(q0, q1) for q1 in range(lo[1], hi[1] + 1) for q0 in range(lo[0], hi[0] + 1)]
            tables.append(reorder_table(t, mean))
        return tables

    def check(self, site, lq, rq):
        """CheckQuantumTable (:105-125): the species of the site (parity)
        takes the physical occupation."""
        shift = site % 2
        for i in range(self.d):
            a = (i, 0) if shift == 0 else (0, i)
            if lq[0] + a[0] == rq[0] and lq[1] + a[1] == rq[1]:
                return i
        return -1

    @staticmethod
    def merge(opq, spq):
        """MergeQuantumTable (:98-103)."""
        return (opq[0] + spq[0], opq[1] + spq[1])

    # ---- Hamiltonian (model/hamiltonian/complex_spinful_bose_square_hamiltonian.cc) ----
    def basic_ops(self):
        """DefineBasicTensor (:45-70): 0=I, 1=a^+, 2=a^-, 3=n,
        4=mu n + U_uu/2 n(n-1), 5=mu n + U_dd/2 n(n-1), both with the
        species-0 chemical potential as in the reference (:61)."""
        d = self.d
        ops = [np.zeros((d, d)) for _ in range(6)]
        for i in range(d):
            ops[0][i, i] = 1.0
            ops[3][i, i] = i
            ops[4][i, i] = self.mu[0] * i + 0.5 * self.inter[0] * i * (i - 1)
            ops[5][i, i] = self.mu[0] * i + 0.5 * self.inter[1] * i * (i - 1)
        for i in range(1, d):
            ops[1][i, i - 1] = math.sqrt(i)
            ops[2][i - 1, i] = math.sqrt(i)
        return ops

    def _string(self, b, p1, p2, ops, coef, q):
        for i in range(self.L):
            if i == p1:
                b.expan(i, ops[0], 1.0, q)
            elif p1 < i < p2:
                b.expan(i, 0, 1.0, q)
            elif i == p2:
                b.expan(i, ops[1], coef, (0, 0))
            else:
                b.expan(i, 0, 1.0, (0, 0))

    def _hop1(self, b, p1, p2, coef):
        """DefineHamiltonian1 (:73-126): a^+_{p1} a^-_{p2} coef."""
        if p1 >= p2:
            raise ValueError("position1 >= position2 in DefineHamiltonian1")
        self._string(b, p1, p2, (1, 2), coef, (1, 0) if p1 % 2 == 0 else (0, 1))

    def _hop2(self, b, p1, p2, coef):
        """DefineHamiltonian2 (:129-181): a^-_{p1} a^+_{p2} coef."""
        if p1 >= p2:
            raise ValueError("position1 >= position2 in DefineHamiltonian2")
        self._string(b, p1, p2, (2, 1), coef, (-1, 0) if p1 % 2 == 0 else (0, -1))

    def _onsite(self, b, p, coef):
        """DefineHamiltonian3 (:184-210)."""
        for i in range(self.L):
            if i == p:
                b.expan(i, 4 if p % 2 == 0 else 5, coef, (0, 0))
            else:
                b.expan(i, 0, 1.0, (0, 0))

    def _inter(self, b, p1, p2, coef):
        """DefineHamiltonian4 (:213-240): n_{p1} n_{p2} coef."""
        for i in range(self.L):
            if i == p1:
                b.expan(i, 3, 1.0, (0, 0))
            elif i == p2:
                b.expan(i, 3, coef, (0, 0))
            else:
                b.expan(i, 0, 1.0, (0, 0))

    @staticmethod
    def phase(i, flux):
        """Phase(2 pi (x+1) flux) (src/tensor/complex.h:127-130)."""
        th = 2 * math.pi * (i + 1) * flux
        return complex(math.cos(th), math.sin(th))

    def mpo(self):
        """DefineTensorHamiltonian (:242-350)."""
        nx, ny = self.nx, self.ny
        b = MPOBuilder(self.L, self.d, 2, self.basic_ops(), cplx=True)
        for i in range(nx):
            for j in range(ny):
                p0 = j + i * ny
                for t in range(2):
                    if self.jx[t] == 0.0:
                        continue
                    p1 = j + ((i + 1) % nx) * ny
                    if i + 1 < nx:
                        self._hop1(b, 2 * p0 + t, 2 * p1 + t, self.jx[t])
                        self._hop2(b, 2 * p0 + t, 2 * p1 + t, self.jx[t])
                    elif nx > 2 and self.period_x:
                        self._hop1(b, 2 * p1 + t, 2 * p0 + t, self.jx[t])
                        self._hop2(b, 2 * p1 + t, 2 * p0 + t, self.jx[t])
                for t in range(2):
                    if self.jy[t] == 0.0:
                        continue
                    p1 = (j + 1) % ny + i * ny
                    ph = self.phase(i, self.flux[t])
                    if j + 1 < ny:
                        self._hop1(b, 2 * p0 + t, 2 * p1 + t, ph.conjugate() * self.jy[t])
                        self._hop2(b, 2 * p0 + t, 2 * p1 + t, ph * self.jy[t])
                    elif ny > 2 and self.period_y:
                        self._hop2(b, 2 * p1 + t, 2 * p0 + t, ph.conjugate() * self.jy[t])
                        self._hop1(b, 2 * p1 + t, 2 * p0 + t, ph * self.jy[t])
                if self.mu[0] != 0.0 or self.inter[0] != 0.0:
                    self._onsite(b, 2 * p0, 1.0)
                if self.mu[1] != 0.0 or self.inter[1] != 0.0:
                    self._onsite(b, 2 * p0 + 1, 1.0)
                if self.inter[2] != 0.0:
                    self._inter(b, 2 * p0, 2 * p0 + 1, self.inter[2])
                b.parallel()
        return b.W, b.qt

    def hops(self):
        """Hopping terms (dst, src, amplitude): amplitude * b+_dst b_src on
        MPS sites, from the lattice geometry: x bonds J_x (b+_p b_q + h.c.),
        y bonds J_y (e^{-i theta} b+_{(x,y)} b_{(x,y+1)} + h.c.) with
        theta = 2 pi (x+1) flux (Landau gauge)."""
        out = []
        nx, ny = self.nx, self.ny
        for i in range(nx):
            for j in range(ny):
                p = j + i * ny
                for t in range(2):
                    if i + 1 < nx or (nx > 2 and self.period_x):
                        q = j + ((i + 1) % nx) * ny
                        out.append((2 * p + t, 2 * q + t, self.jx[t]))
                        out.append((2 * q + t, 2 * p + t, self.jx[t]))
                    if j + 1 < ny or (ny > 2 and self.period_y):
                        q = (j + 1) % ny + i * ny
                        th = 2 * math.pi * (i + 1) * self.flux[t]
                        a = self.jy[t] * complex(math.cos(th), -math.sin(th))
                        out.append((2 * p + t, 2 * q + t, a))
                        out.append((2 * q + t, 2 * p + t, a.conjugate()))
        return [h for h in out if h[2] != 0]


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
