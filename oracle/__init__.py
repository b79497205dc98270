"""CPU oracle for the dmrgcc DMRG hot path -- TEST INFRASTRUCTURE ONLY.

This package is a plain numpy restatement of the reference algorithms
(jiandong94/dmrgcc, real spinless Bose-Hubbard stack).  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker.  The product (``dmrgcc_amd``) never routes
through it.

Each function cites the reference file:line it restates.

How each part is pinned (checks that would catch a dropped term, a wrong sign
or index, or a transposed operand):

* ``rng``        -- glibc TYPE_3 ``rand()``; known-answer test seed 1 ->
                    1804289383, 846930886, 1681692777, 1714636915 and the
                    column-major ``RandomMatrix`` fill (0.83, 0.86, 0.77, ...).
* ``sort``       -- the reference QuickSort; known answer from
                    test/test_real_matrix.cc:9-20 and brute-force checks that
                    the result is a sorted permutation.
* ``model``      -- QN tables against the closed-form ranges; the MPO
                    (incremental build + deparallelisation) contracted to a
                    dense matrix equals a brute-force Fock-space Hamiltonian;
                    ``parallel_matrix`` against the closed-form example
                    [[1,3,1],[2,4,2]].
* ``dmrg``       -- block matvec equals the reference's independent dense
                    path (L (x) R^T * w) and the brute-force Hamiltonian
                    projected on the MPS basis; H_eff symmetry; full sweeps
                    reach the exact-diagonalisation ground energy of small
                    lattices; SVD truncation rule on hand-made spectra;
                    sigma([[1,2,3],[4,5,6]]) closed form.
* ``ed``         -- textbook Fock-space exact diagonalisation; pinned by
                    closed forms (two sites, U=0 tight-binding spectrum).
* ``dmrg``, the sweep-level path (tests/test_oracle_sweep.py) -- the
                    expansion keeps every original block unchanged as the
                    leading part of the enlarged one, appends the expansion
                    block and pads the rest with zeros; its new sectors are
                    QN merges of the MPO channels with the old sectors, in
                    table order, at most max_block; the noise factor is a pure
                    scale (same draws); every expansion column lies in the span
                    of the L K products that may feed it; with expansion on or
                    off and no truncation, the state reaching the next site has
                    the Rayleigh quotient the Lanczos solve found (expansion,
                    canonical form, match and merge preserve the physical
                    state).  Lanczos: the Krylov basis is orthonormal to
                    1e-12 and alpha / beta are H's matrix elements in it.
* ``io``         -- hand-written byte fixtures in the reference's field order
                    (tensor-space, environment and MPO files) read back
                    value by value and rewritten byte for byte.

Parity unpinned: the order in which the reference pairs two rand() draws in
one complex entry (C++ leaves it unspecified; g++ is followed) and
truncation decisions at near-ties of singular values, since the reference
cannot be built here (it needs ``mkl.h``, which this image lacks).
"""
