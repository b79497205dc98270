"""The five workloads the hot path is measured and tested on.

Each is a model of the reference's drivers (model/real_spinless_bose_square_model.cc,
model/complex_spinful_bose_square_model.cc) with its sweep table
(RealTensorRundmrg::Sweep, src/dmrg/realdmrg/real_tensor_rundmrg.cc:70-113).

Conventions: the reference's spinless on-site term is U_ref n(n-1) with no
1/2 (model/hamiltonian/real_spinless_bose_square_hamiltonian.cc), so the
standard Bose-Hubbard U enters as U_ref = U / 2.  The two-component model
already carries the 1/2 and takes U, U' as they are.  Hopping t = 1 is the
reference's J = -1.

  cylinder-c1   4x4 open square, L = 16, N = 16, d = 4, U = 4, m = 200, 6 sweeps
  cylinder-c2   8x4 cylinder (periodic y), L = 32, N = 32, d = 4, U = 6, m = 1000, 8 sweeps
  chains-c3     64 chains, L from {16,24,32,48,64}, U uniform in [2, 10],
                m from {64,128,256}, unit filling, d = 4 (SplitMix64 seed 20261203)
  spinful-c4    4x4 two-component bosons, flux 1/4, N_up = N_down = 8, d = 3,
                U = 4, U' = 2, m = 2000, 8 sweeps (32 MPS sites)
  cylinder-c5   10x6 cylinder (periodic y), L = 60, N = 60, d = 4, U = 8, m = 4000, 2 sweeps
"""

NOISE = [2, 2, 3, 3, 4, 4] + [-1] * 200

# (nx, ny, N, d, U (standard), period_y, maxm, sweeps)
SPINLESS = {
    "c1": (4, 4, 16, 4, 4.0, False, 200, 6),
    "c2": (8, 4, 32, 4, 6.0, True, 1000, 8),
    "c5": (10, 6, 60, 4, 8.0, True, 4000, 2),
}
# (nx, ny, n_up, n_down, d, flux, U_uu, U_dd, U_ud, maxm, sweeps)
SPINFUL = {"c4": (4, 4, 8, 8, 3, 0.25, 4.0, 4.0, 2.0, 2000, 8)}

C3_SEED = 20261203
C3_LENGTHS = (16, 24, 32, 48, 64)
C3_MAXM = (64, 128, 256)


def table(maxm, sweeps, niter=10, canpre=1e-10, maxb=400, noise=None):
    """Sweep table rows (maxm, maxb, niter, canpre, noise exponent), one per
    half sweep: the noise exponents 2,2,3,3,4,4 then -1 (noise 0.1^-1 = 10
    switches the expansion off, real_tensor_rundmrg.cc:96-113)."""
    noise = NOISE if noise is None else noise
    return [(maxm, maxb, niter, canpre, noise[i]) for i in range(2 * sweeps)]


def spinless(name, dm=None):
    """(library model, sweep table) of a spinless workload."""
    if dm is None:
        import dmrgcc_amd as dm
    nx, ny, N, d, U, py, maxm, sweeps = SPINLESS[name]
    model = dm.SpinlessModel(nx, ny, N, d, hop_x=-1.0, hop_y=-1.0, U=U / 2, period_y=py)
    return model, table(maxm, sweeps)


def spinful(name="c4", maxm=None, sweeps=None, dm=None):
    """(library model, sweep table) of the two-component workload."""
    if dm is None:
        import dmrgcc_amd as dm
    nx, ny, nu, nd, d, flux, uu, udd, ud, m0, s0 = SPINFUL[name]
    model = dm.SpinfulModel(nx, ny, nu, nd, d, flux=flux, hop_x=-1.0, hop_y=-1.0, U_uu=uu, U_dd=udd,
                            U_ud=ud)
    return model, table(maxm or m0, sweeps or s0)


def splitmix64(seed):
    """SplitMix64 (Steele, Lea, Flood 2014): the generator the chain scan
    draws its instances from."""
    state = seed & 0xFFFFFFFFFFFFFFFF
    while True:
        state = (state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        yield z ^ (z >> 31)


def chain_draws(n=64, seed=C3_SEED):
    """The chain scan's instances in draw order: (L, U, m) with L from
    {16,24,32,48,64}, U uniform in [2, 10) (53-bit mantissa) and m from
    {64,128,256}, three draws per instance."""
    g = splitmix64(seed)
    out = []
    for _ in range(n):
        L = C3_LENGTHS[next(g) % 5]
        U = 2.0 + 8.0 * (next(g) >> 11) * 2.0 ** -53
        m = C3_MAXM[next(g) % 3]
        out.append((L, U, m))
    return out
