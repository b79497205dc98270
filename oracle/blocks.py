"""Block-sparse matrix and MPS site tensor containers, oracle copy.

Restates src/tensor/real_matrix_block.cc (FindMatrixBlock :314-380,
ComputePartMatrixBlockDim :154-162, VectorizeMatrixBlock :173-194,
NormalizeMatrixBlock :164-171) and the RealTensorLattice fields
(src/dmrg/realdmrg/real_tensor_lattice.h:10-42).  A block is a 2-D float64
numpy array; an empty (reset) block has shape (0, 0).
"""
import numpy as np


def empty():
    return np.zeros((0, 0))


class BlockMat:
    def __init__(self, left=(), right=(), mats=None):
        self.left = list(left)
        self.right = list(right)
        if mats is None:
            mats = [empty() for _ in self.left]
        self.mats = list(mats)

    @property
    def n(self):
        return len(self.left)

    def copy(self):
        return BlockMat(self.left, self.right, [m.copy() for m in self.mats])

    def find(self, l, r):
        """FindMatrixBlock(left, right): first position or -1."""
        for i in range(self.n):
            if self.left[i] == l and self.right[i] == r:
                return i
        return -1

    def same(self, leigh, target):
        """FindMatrixBlock(num, positions, leigh, target)."""
        key = self.left if leigh == 0 else self.right
        return [i for i in range(self.n) if key[i] == target]

    def dim(self):
        return sum(m.size for m in self.mats)

    def offsets(self):
        out, acc = [], 0
        for m in self.mats:
            out.append(acc)
            acc += m.size
        return out

    def to_vector(self):
        if not self.mats:
            return np.zeros(0)
        return np.concatenate([m.ravel() for m in self.mats])

    def from_vector(self, v):
        acc = 0
        for i, m in enumerate(self.mats):
            dt = np.result_type(m.dtype, np.asarray(v).dtype)
            self.mats[i] = np.asarray(v[acc:acc + m.size], dtype=dt).reshape(m.shape).copy()
            acc += m.size

    def normalize(self):
        """NormalizeMatrixBlock (real_matrix_block.cc:164-171; complex:
        sum of SquareNorm, complex_matrix_block.cc:163-170)."""
        if any(np.iscomplexobj(m) for m in self.mats):
            s = sum(float(np.sum(m.real * m.real + m.imag * m.imag)) for m in self.mats)
        else:
            s = sum(float(np.sum(m * m)) for m in self.mats)
        f = np.sqrt(1.0 / s)
        self.mats = [m * f for m in self.mats]

    def reset(self):
        self.mats = [empty() for _ in self.mats]


class Lattice:
    """RealTensorLattice: one MPS site."""

    def __init__(self, d):
        self.d = d
        self.left_block = []
        self.left_dim = []
        self.right_block = []
        self.right_dim = []
        self.phys = []  # [nl][nr], -1 where no block
        self.ket = BlockMat()
        self.match_dim = None
        self.canonical = None

    def copy(self):
        o = Lattice(self.d)
        o.left_block = list(self.left_block)
        o.left_dim = list(self.left_dim)
        o.right_block = list(self.right_block)
        o.right_dim = list(self.right_dim)
        o.phys = [list(r) for r in self.phys]
        o.ket = self.ket.copy()
        return o

    @property
    def nl(self):
        return len(self.left_block)

    @property
    def nr(self):
        return len(self.right_block)

    def define_sectors(self, left_block, right_block, left_dim, right_dim):
        """DefineTensorLattice (real_tensor_lattice.cc:300-340)."""
        self.left_block = list(left_block)
        self.left_dim = list(left_dim)
        self.right_block = list(right_block)
        self.right_dim = list(right_dim)
        self.phys = [[-1] * len(right_block) for _ in left_block]

    def define_blocks(self, li, ri, pi, rows, cols, rng, cplx=False):
        """DefineTensorLattice (real_tensor_lattice.cc:343-359): every block
        gets a fresh RandomMatrix(rows, cols) (consumes rand(); complex
        blocks, complex_tensor_lattice.cc:343-359, two draws per entry)."""
        mats = []
        for k in range(len(li)):
            self.phys[li[k]][ri[k]] = pi[k]
            mats.append(rng.complex_matrix(rows, cols) if cplx else rng.random_matrix(rows, cols))
        self.ket = BlockMat(li, ri, mats)


def lattice_from_export(d, phys, blocks, data):
    """A Lattice from a site exported as flat block lists: the physical-index
    table phys[nl][nr] and the ket blocks (o, l, r, rows, cols, off) over the
    flat `data` (the VectorizeMatrixBlock layout).  Sector positions stand
    for the QN-table indices; sector dimensions are read off the blocks."""
    phys = np.asarray(phys)
    lat = Lattice(d)
    nl, nr = phys.shape
    lat.left_block = list(range(nl))
    lat.right_block = list(range(nr))
    lat.left_dim = [0] * nl
    lat.right_dim = [0] * nr
    lat.phys = phys.tolist()
    mats = []
    for (_, l, r, rows, cols, off) in blocks:
        mats.append(np.array(data[off:off + rows * cols]).reshape(rows, cols))
        lat.left_dim[l] = rows
        lat.right_dim[r] = cols
    lat.ket = BlockMat([b[1] for b in blocks], [b[2] for b in blocks], mats)
    return lat


def env_from_export(blocks, data, nch):
    """An environment (one BlockMat per MPO channel) from flat block lists
    (o, l, r, rows, cols, off) over `data`."""
    out = [BlockMat() for _ in range(nch)]
    for (o, l, r, rows, cols, off) in blocks:
        out[o].left.append(l)
        out[o].right.append(r)
        out[o].mats.append(np.array(data[off:off + rows * cols]).reshape(rows, cols))
    return out
