"""Write the sweep-trajectory fixtures the GPU tests compare the library with.

Each fixture is the oracle's run (oracle/dmrg.py, a numpy restatement of
RealTensorRundmrg::Run, src/dmrg/realdmrg/real_tensor_rundmrg.cc:115-245,
and of its complex twin) of one workload: per half sweep the energy and the
largest discarded weight, the number of rand() draws consumed, and the bond
dimension of every MPS bond at the end.

    python tests/golden/make_trajectories.py [c1] [c4_m64]

  c1      4x4 open square, L = 16, N = 16, d = 4, U = 4 (U_ref = 2), t = 1,
          m = 200, 6 sweeps, niter 10, canpre 1e-10 (about 30 s)
  c4_m64  4x4 two-component bosons, flux 1/4, N_up = N_down = 8, d = 3,
          U = 4, U' = 2, t = 1, the m = 2000 workload's model at m = 64,
          2 sweeps (a few minutes)
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.dmrg import DMRG  # noqa: E402
from oracle.model import SpinfulBoseSquare, SpinlessBoseSquare  # noqa: E402

NOISE = [2, 2, 3, 3, 4, 4] + [-1] * 200


def rows(maxm, sweeps, niter=10, canpre=1e-10, maxb=400):
    return [(maxm, maxb, niter, canpre, NOISE[i]) for i in range(2 * sweeps)]


def bond_dims(dm):
    """Bond dimension to the right of every site: the sum of its right
    sector dimensions over the sectors its ket blocks use."""
    out = []
    for lat in dm.space.lat:
        used = {}
        for r, m in zip(lat.ket.right, lat.ket.mats):
            used[r] = m.shape[1]
        out.append(int(sum(used.values())))
    return out


def record(name, model, table, sweeps):
    t0 = time.time()
    dm = DMRG(model, table, sweeps, first_block=20, first_dim=10)
    dm.run()
    out = {"workload": name, "table": [list(r) for r in table], "energies": [float(e) for e in dm.energies],
           "trunc_err": [float(t) for t in dm.trunc_err], "rand_draws": int(dm.rng.count),
           "bond_dims": bond_dims(dm), "oracle_seconds": time.time() - t0}
    path = os.path.join(HERE, "%s_trajectory.json" % name)
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("%s: %.1f s, E = %.12f, %d draws -> %s" % (name, out["oracle_seconds"], out["energies"][-1],
                                                      out["rand_draws"], path))


def main():
    which = sys.argv[1:] or ["c1"]
    if "c1" in which:
        record("c1", SpinlessBoseSquare(4, 4, 16, 4, jx=-1.0, jy=-1.0, U=2.0), rows(200, 6), 6)
    if "c4_m64" in which:
        record("c4_m64", SpinfulBoseSquare(4, 4, 8, 8, 3, flux=0.25, jx=-1.0, jy=-1.0, U_uu=4.0, U_dd=4.0,
                                           U_ud=2.0), rows(64, 2), 2)


if __name__ == "__main__":
    main()
