"""Tensor-space, environment and MPO files, oracle copy (test infrastructure
only).

Restates RealTensorSpace::Write/ReadTensorSpace (src/dmrg/realdmrg/
real_tensor_space.cc:171-277), RealTensorLattice::Write/ReadTensorLattice
(real_tensor_lattice.cc:207-298), RealMatrixBlock::Write/ReadMatrixBlock
(src/tensor/real_matrix_block.cc:216-283) and RealMatrix::Write/ReadMatrix
(src/tensor/real_matrix.cc:128-175); the complex twins write each entry as
the 16-byte Complex {real, imag} (src/tensor/complex_matrix.cc:139-175).
int is 4 bytes, little-endian (x86-64 native order).

Layout:
  int num_site, num_site+1, num_site-1, num_quantum
  int num_table[num_site+1]; int quantum_table[i][j][k]
  per site: int physics_dim; int nl; (int left_block, int left_dim) x nl;
            int nr; (int right_block, int right_dim) x nr;
            int physics_index[nl][nr];
            int num_block; per block: int left, int right,
                int rows, int cols, int rows*cols, entries (row-major)
  int max_block, int max_dim, double canonical_precision, double noise_factor
"""
import struct

import numpy as np

from .blocks import BlockMat, Lattice


def _ints(f, *v):
    f.write(struct.pack("<%di" % len(v), *v))


def write_space(path, space):
    L = space.L
    tabs = space.tables
    nq = len(tabs[0][0])
    with open(path, "wb") as f:
        _ints(f, L, L + 1, L - 1, nq)
        _ints(f, *[len(t) for t in tabs])
        for t in tabs:
            for q in t:
                _ints(f, *q)
        for lat in space.lat:
            _ints(f, lat.d, lat.nl)
            for b, dd in zip(lat.left_block, lat.left_dim):
                _ints(f, b, dd)
            _ints(f, lat.nr)
            for b, dd in zip(lat.right_block, lat.right_dim):
                _ints(f, b, dd)
            for row in lat.phys:
                if row:
                    _ints(f, *row)
            ket = lat.ket
            _ints(f, ket.n)
            for k in range(ket.n):
                m = ket.mats[k]
                _ints(f, ket.left[k], ket.right[k], m.shape[0], m.shape[1], m.size)
                dt = np.complex128 if space.cplx else np.float64
                f.write(np.ascontiguousarray(m, dtype=dt).tobytes())
        _ints(f, space.max_block, space.max_dim)
        f.write(struct.pack("<dd", space.canpre, space.noise))


class _Reader:
    def __init__(self, data):
        self.b = data
        self.p = 0

    def ints(self, n):
        v = struct.unpack_from("<%di" % n, self.b, self.p)
        self.p += 4 * n
        return list(v)

    def arr(self, n, dt):
        a = np.frombuffer(self.b, dtype=dt, count=n, offset=self.p).copy()
        self.p += a.nbytes
        return a


def read_space(path, space):
    """Fill `space` (oracle.dmrg.Space) from a tensor-space file."""
    with open(path, "rb") as f:
        r = _Reader(f.read())
    L, Lpp, Lmm, nq = r.ints(4)
    if L != space.L or Lpp != L + 1 or Lmm != L - 1:
        raise ValueError("tensor space file has another number of sites")
    nt = r.ints(Lpp)
    tabs = [[tuple(r.ints(nq)) for _ in range(n)] for n in nt]
    dt = np.complex128 if space.cplx else np.float64
    lats = []
    for _ in range(L):
        lat = Lattice(r.ints(1)[0])
        nl = r.ints(1)[0]
        lv = r.ints(2 * nl)
        nr = r.ints(1)[0]
        rv = r.ints(2 * nr)
        lat.define_sectors(lv[0::2], rv[0::2], lv[1::2], rv[1::2])
        for i in range(nl):
            lat.phys[i] = r.ints(nr) if nr else []
        nb = r.ints(1)[0]
        left, right, mats = [], [], []
        for _ in range(nb):
            lidx, ridx, rows, cols, tot = r.ints(5)
            if tot != rows * cols:
                raise ValueError("bad matrix header in tensor space file")
            left.append(lidx)
            right.append(ridx)
            mats.append(r.arr(tot, dt).reshape(rows, cols))
        lat.ket = BlockMat(left, right, mats)
        lats.append(lat)
    space.max_block, space.max_dim = r.ints(2)
    space.canpre, space.noise = struct.unpack_from("<dd", r.b, r.p)
    r.p += 16
    if r.p != len(r.b):
        raise ValueError("trailing bytes in tensor space file")
    space.tables = tabs
    space.lat = lats


# --------------------------------------------------------------------------
# Environment files: RealTensorContraction::Write/ReadTensorContraction
# (src/dmrg/realdmrg/real_tensor_contraction.cc:73-109): per channel
#   int num_block; per block: int left, int right, int rows, int cols,
#   int rows*cols, entries (row-major)
# --------------------------------------------------------------------------
def write_contraction(path, env, cplx=False):
    dt = np.complex128 if cplx else np.float64
    with open(path, "wb") as f:
        for bm in env:
            _ints(f, bm.n)
            for k in range(bm.n):
                m = bm.mats[k]
                _ints(f, bm.left[k], bm.right[k], m.shape[0], m.shape[1], m.size)
                f.write(np.ascontiguousarray(m, dtype=dt).tobytes())


def read_contraction(path, nbond, cplx=False):
    dt = np.complex128 if cplx else np.float64
    with open(path, "rb") as f:
        r = _Reader(f.read())
    env = []
    for _ in range(nbond):
        nb = r.ints(1)[0]
        left, right, mats = [], [], []
        for _ in range(nb):
            lidx, ridx, rows, cols, tot = r.ints(5)
            if tot != rows * cols:
                raise ValueError("bad matrix header in environment file")
            left.append(lidx)
            right.append(ridx)
            mats.append(r.arr(tot, dt).reshape(rows, cols))
        env.append(BlockMat(left, right, mats))
    if r.p != len(r.b):
        raise ValueError("trailing bytes in environment file")
    return env


# --------------------------------------------------------------------------
# MPO files: RealTensorHamiltonian::Write/ReadTensorHamiltonian
# (src/dmrg/realdmrg/real_tensor_hamiltonian.cc:74-150) and
# RealTensorOperator::Write/ReadTensorOperator (real_tensor_operator.cc:96-139):
#   int num_site, num_site+1, num_site-1
#   int num_operator; per basic operator: int rows, cols, rows*cols, entries
#   int num_quantum; int num_table[num_site+1]; int quantum_table[i][j][k]
#   per site: int physics_dim, left_bond, right_bond;
#             W[o1][o2] for o1 < left_bond, o2 < right_bond: int d, d, d*d, entries
# The reference's reader stops after the quantum tables and reads the basic
# operators into unallocated matrices (:132); this reader reads everything.
# --------------------------------------------------------------------------
def write_hamiltonian(path, W, basic, tables, cplx=False):
    dt = np.complex128 if cplx else np.float64
    L = len(W)
    nq = len(tables[0][0])
    with open(path, "wb") as f:
        _ints(f, L, L + 1, L - 1, len(basic))
        for b in basic:
            _ints(f, b.shape[0], b.shape[1], b.size)
            f.write(np.ascontiguousarray(b, dtype=dt).tobytes())
        _ints(f, nq)
        _ints(f, *[len(t) for t in tables])
        for t in tables:
            for q in t:
                _ints(f, *q)
        for w in W:
            lb, rb, d, _ = w.shape
            _ints(f, d, lb, rb)
            for a in range(lb):
                for b in range(rb):
                    _ints(f, d, d, d * d)
                    f.write(np.ascontiguousarray(w[a, b], dtype=dt).tobytes())


def read_hamiltonian(path, cplx=False):
    """Returns (W list of (lb, rb, d, d) arrays, basic operators, bond tables)."""
    dt = np.complex128 if cplx else np.float64
    with open(path, "rb") as f:
        r = _Reader(f.read())
    L, Lpp, Lmm, nop = r.ints(4)
    if Lpp != L + 1 or Lmm != L - 1:
        raise ValueError("bad site counts in MPO file")
    basic = []
    for _ in range(nop):
        rows, cols, tot = r.ints(3)
        basic.append(r.arr(tot, dt).reshape(rows, cols))
    nq = r.ints(1)[0]
    nt = r.ints(Lpp)
    tables = [[tuple(r.ints(nq)) for _ in range(n)] for n in nt]
    W = []
    for _ in range(L):
        d, lb, rb = r.ints(3)
        w = np.zeros((lb, rb, d, d), dtype=dt)
        for a in range(lb):
            for b in range(rb):
                rows, cols, tot = r.ints(3)
                if (rows, cols, tot) != (d, d, d * d):
                    raise ValueError("bad operator matrix header in MPO file")
                w[a, b] = r.arr(tot, dt).reshape(d, d)
        W.append(w)
    if r.p != len(r.b):
        raise ValueError("trailing bytes in MPO file")
    return W, basic, tables
