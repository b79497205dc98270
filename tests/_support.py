"""Shared helpers for the tests: oracle states and their conversion into the
C-ABI block descriptions."""
import numpy as np

from oracle.dmrg import DMRG, lanczos, left_env, right_env
from oracle.model import SpinlessBoseSquare


def sweep_rows(maxm, nhalf, niter=10, canpre=1e-10, maxb=400, noise=None):
    noise = noise or ([2, 2, 3, 3] + [-1] * 100)
    return [(maxm, maxb, niter, canpre, noise[i]) for i in range(nhalf)]


def oracle_state(nx=3, ny=2, N=6, d=3, U=2.0, maxm=40, half=2, site=2, niter=8, first_dim=6,
                 period_y=False):
    """Run the oracle DMRG up to the Lanczos solve of `site` in half sweep
    `half` and return the engine: lattice `site` and its L/R environments are
    then consistent (the state the reference's LanczosMethod sees)."""
    m = SpinlessBoseSquare(nx, ny, N, d, jx=-1.0, jy=-1.0, U=U, period_y=period_y)
    rows = sweep_rows(maxm, half + 2, niter)
    dm = DMRG(m, rows, (half + 2) // 2, first_block=20, first_dim=first_dim)
    dm.run(stop=(half, site))
    return dm


def site_description(lat):
    """phys table and ket block list of an oracle Lattice."""
    phys = np.array(lat.phys, dtype=np.int32).reshape(lat.nl, lat.nr)
    ket = lat.ket
    rows = [m.shape[0] for m in ket.mats]
    cols = [m.shape[1] for m in ket.mats]
    return phys, ket.left, ket.right, rows, cols


def env_blocks(env):
    """Flatten an oracle environment (list over channels of BlockMat) into
    (o, l, r, rows, cols, off) lists and the concatenated data."""
    o, l, r, rows, cols, off, data = [], [], [], [], [], [], []
    acc = 0
    for ch, bm in enumerate(env):
        for k in range(bm.n):
            m = bm.mats[k]
            o.append(ch)
            l.append(bm.left[k])
            r.append(bm.right[k])
            rows.append(m.shape[0])
            cols.append(m.shape[1])
            off.append(acc)
            data.append(m.ravel())
            acc += m.size
    flat = np.concatenate(data) if data else np.zeros(0)
    return (o, l, r, rows, cols, off), flat
