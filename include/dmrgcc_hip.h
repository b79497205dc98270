/*
 * dmrgcc_hip.h -- C ABI of the MI355X DMRG hot path.
 *
 * The reference (jiandong94/dmrgcc) has no FFI: its hot path sits behind a
 * C++ class API.  Each entry point below replaces one member function of that
 * API (cited as file:line); plain pointers and sizes only.  Device pointers
 * are HIP device addresses (HBM); host pointers are ordinary memory.
 *
 * Conventions
 *  - Every function returns 0 on success or a nonzero status:
 *      1 HIP runtime error, 2 shape / index mismatch (the reference prints and
 *      calls exit(-1) there), 3 invalid argument, 4 no convergence,
 *      5 internal error.  dmrgcc_last_error() returns the message.
 *  - Matrices are row-major FP64.  A block-sparse matrix is a list of blocks
 *    (sector indices l, r; rows; cols) stored back to back in block order in
 *    one buffer: exactly the reference's VectorizeMatrixBlock layout
 *    (src/tensor/real_matrix_block.cc:173-194).
 *  - MPO site tensors are dense host arrays W[o1][o2][p1][p2] (bra p1, ket
 *    p2), lb*rb*d*d doubles.
 *  - precision: 0 = FP64 (MFMA f64), 1 = FP32 (MFMA f32),
 *               2 = BF16x3 (split bf16 MFMA, FP32 accumulate).
 *  - stream: a hipStream_t, or NULL for the library's own stream (a
 *    non-blocking stream: the caller finishes the work that produces the
 *    inputs before the call).  Every entry point except
 *    dmrgcc_heff_apply_async has completed its device work when it returns.
 *  - Not thread-safe per handle.
 */
#ifndef DMRGCC_HIP_H_
#define DMRGCC_HIP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- library ---------------------------------------------------------- */
const char* dmrgcc_last_error(void);
int dmrgcc_version(void);

/* ---- host-side restatements (no GPU needed) ---------------------------- */

/* glibc rand() stream from `seed` (the reference never seeds: seed 1),
 * src/tensor/real_matrix.cc:201-208.  Writes n successive values. */
int dmrgcc_rand_draws(unsigned seed, int64_t skip, int n, int* out);
/* n random matrices in draw order from the stream after `skip` draws:
 * RealMatrix::RandomMatrix (src/tensor/real_matrix.cc:201-208) or, with
 * is_complex, ComplexMatrix::RandomMatrix (src/tensor/complex_matrix.cc:
 * 201-208, two draws per entry, the first one the imaginary part as g++
 * evaluates the constructor arguments); row-major, concatenated in out
 * (interleaved re, im when complex).  With w != NULL the diagonal of matrix
 * k is set to 0.1 * w[k] (complex: w holds re, im pairs), as the subspace
 * expansion does (real_tensor_contraction.cc:397-399).  Long skips jump
 * ahead in O(log skip); large sets are filled by several host threads. */
int dmrgcc_random_matrices(unsigned seed, int64_t skip, int n, const int* rows, const int* cols,
                           int is_complex, const double* w, double* out);

/* QuickSort (src/util/general.ih:6-65), flag 0 descending / 1 ascending,
 * sorting `vals` in place together with `index`. */
int dmrgcc_quicksort(int n, double* vals, int* index, int flag);

/* ComputeTruncateDim (src/dmrg/realdmrg/real_tensor_lattice.cc:501-555).
 * Sector s has counts[s] descending singular values in `svals` (concatenated;
 * modified in place when a sector is forced to keep one state with sigma 0).
 * Writes the kept dimension per sector to `trunc`. */
int dmrgcc_truncate_dims(int nsec, const int* counts, double* svals, int max_dim,
                         double canpre, int* trunc);

/* RealSymMatrixDiag (src/util/general.cc:47-64): all eigenpairs of the dense
 * symmetric n x n row-major matrix `a`, eigenvalues ascending in `w`, `a`
 * overwritten by the eigenvectors as columns (a[k*n + j] = component k of
 * vector j), as the reference's dsyevd call leaves them; each vector's
 * largest |component| is positive.  Host cyclic Jacobi (the Lanczos
 * tridiagonal eigensolve). */
int dmrgcc_sym_eig(int n, double* a, double* w);
/* The lowest eigenpair only (what RealSymMatrixDiag's callers read: w[0] and
 * column 0), by Householder tridiagonalisation, Sturm bisection and inverse
 * iteration (the dense edge-site solve of ComputeEffectHamilton,
 * real_tensor_contraction.cc:1136-1205).  `a` is not modified. */
int dmrgcc_sym_eig_lowest(int n, const double* a, double* e0, double* v0);

/* Real spinless Bose-Hubbard on an nx x ny square lattice (site j + i*ny):
 * RealSpinlessBoseSquareSpace / RealSpinlessBoseSquareHamiltonian
 * (model/space/real_spinless_bose_square_space.cc:26-83,
 *  model/hamiltonian/real_spinless_bose_square_hamiltonian.cc:41-228).
 * A model handle owns the MPS QN tables and the deparallelised MPO. */
typedef struct dmrgcc_model dmrgcc_model;
int dmrgcc_model_spinless_create(int nx, int ny, int num_boson, int physics_dim, double hop_x,
                                 double hop_y, double U, double mu, int period_x, int period_y,
                                 dmrgcc_model** out);
/* The complex twin (ComplexSpinlessBoseSquareSpace /
 * ComplexSpinlessBoseSquareHamiltonian, model/space/complex_spinless_bose_square_space.cc,
 * model/hamiltonian/complex_spinless_bose_square_hamiltonian.cc): the same
 * MPO on complex storage; runs on it use the complex sweep engine
 * (ComplexTensorRundmrg). */
int dmrgcc_model_zspinless_create(int nx, int ny, int num_boson, int physics_dim, double hop_x,
                                  double hop_y, double U, double mu, int period_x, int period_y,
                                  dmrgcc_model** out);
/* Complex two-component bosons (ComplexSpinfulBoseSquareSpace /
 * ComplexSpinfulBoseSquareHamiltonian, model/space/complex_spinful_bose_square_space.cc:3-125,
 * model/hamiltonian/complex_spinful_bose_square_hamiltonian.cc:3-350):
 * L = 2*nx*ny MPS sites (up, down per lattice site), QNs (N_up, N_down).
 * flux, hop_x, hop_y, mu: two entries (per species); inter: U_uu, U_dd, U_ud.
 * y bonds carry the phase exp(i*2*pi*(x+1)*flux). */
int dmrgcc_model_spinful_create(int nx, int ny, int n_up, int n_down, int physics_dim,
                                const double* flux, const double* hop_x, const double* hop_y,
                                const double* inter, const double* mu, int period_x, int period_y,
                                dmrgcc_model** out);
/* Model from a driver's input file: InputGroup(path, "input")
 * (src/util/input.cc:98-331) read as the reference's model drivers read it.
 * driver: "real_spinless_bose_square" (model/real_spinless_bose_square_model.cc:3-43),
 * "complex_spinless_bose_square" (model/complex_spinless_bose_square_model.cc:3-43) or
 * "complex_spinful_bose_square" (model/complex_spinful_bose_square_model.cc:3-56).
 * echo != 0 prints the reference's "Got/Def input.<key> = value" lines.
 * A missing required key fails with "there is no item: <key>". */
int dmrgcc_input_model_create(const char* path, const char* driver, int echo, dmrgcc_model** out);
/* The run section of the same file, as the RealTensorRundmrg constructor
 * reads it (real_tensor_rundmrg.cc:3-42) with Sweep (:70-113): first_block,
 * first_dim and the sweep table, one row per half sweep (nhalf =
 * 2*num_sweep) after the reference's fill-in rule; noise holds the table's
 * exponents (dmrgcc_run applies 0.1^noise).  Pass maxm == NULL to get
 * nhalf first.  flags (4 ints, may be NULL): disk_cache, cache_record,
 * cache_resume, record_process. */
int dmrgcc_input_run_config(const char* path, int echo, int* first_block, int* first_dim, int* nhalf,
                            int* maxm, int* maxb, int* niter, double* canpre, double* noise, int* flags);
/* One int (kind 0), double (1) or bool (2) value of `group` as
 * InputGroup::GetInt / GetDouble / GetBool read it (src/util/input.cc:
 * 203-262): *found = 0/1, *value = the value (bools as 0 / 1), or
 * default_value when absent; echo prints the reference's Got / Def line
 * (Def only with has_default). */
int dmrgcc_input_get(const char* path, const char* group, const char* key, int kind, int echo,
                     int has_default, double default_value, double* value, int* found);
/* One string value of the "input" group (InputGroup::GetString,
 * input.cc:223-244; a key with nothing after '=' is found and empty):
 * the value into buf (NUL-terminated, buflen bytes), *found = 0/1. */
int dmrgcc_input_string(const char* path, const char* key, char* buf, int buflen, int* found);
int dmrgcc_model_destroy(dmrgcc_model* m);
/* MPO files: RealTensorHamiltonian::WriteTensorHamiltonian /
 * ReadTensorHamiltonian (real_tensor_hamiltonian.cc:74-150, the site
 * operators as RealTensorOperator::WriteTensorOperator,
 * real_tensor_operator.cc:96-139): int num_site, num_site+1, num_site-1;
 * int num_operator, the basic operators (rows, cols, rows*cols, entries);
 * int num_quantum, the MPO bond table sizes and entries; per site int
 * physics_dim, left_bond, right_bond and the W[o1][o2] d x d matrices.
 * Entries are doubles, 16-byte Complex for a complex model; native order.
 * read replaces the model's MPO (same sites, quantum numbers and kind).
 * The reference's reader never allocates the basic operators it reads
 * into (real_tensor_hamiltonian.cc:132) and does not read the site
 * operators back; this one reads both. */
int dmrgcc_model_write_mpo(const dmrgcc_model* m, const char* path);
int dmrgcc_model_read_mpo(dmrgcc_model* m, const char* path);
int dmrgcc_model_num_site(const dmrgcc_model* m);
/* Number of sites, local dimension, QNs per sector and whether the MPO is
 * complex (any pointer may be NULL). */
int dmrgcc_model_info(const dmrgcc_model* m, int* num_site, int* physics_dim, int* num_quantum,
                      int* is_complex);
/* MPS bond QN table `bond` (0..L): count, then num_quantum ints per entry. */
int dmrgcc_model_space_table(const dmrgcc_model* m, int bond, int* count, int* qn);
/* MPO site tensor: bond dims, then W (lb*rb*d*d entries; doubles, or
 * interleaved (re, im) pairs for a complex model) if W != NULL. */
int dmrgcc_model_mpo_site(const dmrgcc_model* m, int site, int* lb, int* rb, double* W);
/* MPO bond QN table (0..L): count, then num_quantum ints per channel. */
int dmrgcc_model_mpo_table(const dmrgcc_model* m, int bond, int* count, int* qn);

/* ---- block GEMM (RealMatrix::MultiplyToMatrix, real_matrix.cc:291-315) -- */
/* Batched C_i = A_i * B_i, row-major, device pointers, one launch. */
int dmrgcc_block_gemm(int n, const double* const* A, const double* const* B, double* const* C,
                      const int* M, const int* N, const int* K, int precision, void* stream);

/* No counterpart in the reference: the allocation-free form of
 * dmrgcc_block_gemm, for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and the same products, bit for bit; the GEMM
 * descriptors (items, terms, the tile queue) and any split-K partials go
 * into a fixed 8 MiB device arena that each GPU holds until the process
 * ends, so no hipMalloc / hipFree is called.  A call whose descriptors do
 * not fit (about 500,000 64x64 output tiles) fails with status 3.  Every
 * call on one GPU shares that GPU's arena: calls that overlap on one GPU
 * (two host threads, or two streams before the first call returns) race. */
int dmrgcc_block_gemm_noalloc(int n, const double* const* A, const double* const* B, double* const* C,
                              const int* M, const int* N, const int* K, int precision, void* stream);

/* Complex (ComplexMatrix::MultiplyToMatrix, complex_matrix.cc:294-318):
 * C_i = A_i * B_i with interleaved complex entries (re, im), so every
 * matrix of M x N entries is 2*M*N doubles. */
int dmrgcc_zblock_gemm(int n, const double* const* A, const double* const* B, double* const* C,
                       const int* M, const int* N, const int* K, int precision, void* stream);

/* ---- effective Hamiltonian (MultiplyEffectHamilton,
 *      src/dmrg/realdmrg/real_tensor_contraction.cc:1223-1291) ---------- */
typedef struct dmrgcc_heff dmrgcc_heff;
/* Site description:
 *   d, lb, rb, W (host)          MPO site tensor
 *   nl, nr, phys (host nl*nr)    physical index per sector pair, -1 if none
 *   nket, ket_l/r/rows/cols      ket block list (defines the vector layout)
 *   nL, L_o/l/r/rows/cols/off    left environment blocks: channel o1, bra
 *                                sector l, ket sector r, offset into L_data
 *   nR, R_o/l/r/rows/cols/off    right environment blocks: channel o2, ket
 *                                sector l, bra sector r, offset into R_data
 * L_data / R_data are device buffers that must outlive the plan. */
int dmrgcc_heff_create(int d, int lb, int rb, const double* W, int nl, int nr, const int* phys,
                       int nket, const int* ket_l, const int* ket_r, const int* ket_rows,
                       const int* ket_cols, int nL, const int* L_o, const int* L_l,
                       const int* L_r, const int* L_rows, const int* L_cols,
                       const int64_t* L_off, const double* L_data, int nR, const int* R_o,
                       const int* R_l, const int* R_r, const int* R_rows, const int* R_cols,
                       const int64_t* R_off, const double* R_data, void* stream,
                       dmrgcc_heff** out);
/* Complex plan (ComplexTensorContraction::MultiplyEffectHamilton,
 * src/dmrg/complexdmrg/complex_tensor_contraction.cc:1244-1312): the same
 * description with W (lb*rb*d*d entries), L_data and R_data holding
 * interleaved complex entries; offsets and shapes count entries.
 * dmrgcc_heff_apply / _info / _time / _destroy and dmrgcc_lanczos take the
 * handle as for a real plan, with vectors of 2*psi_dim doubles; F_ref counts
 * 8 real flops per complex multiply-add. */
int dmrgcc_zheff_create(int d, int lb, int rb, const double* W, int nl, int nr, const int* phys,
                        int nket, const int* ket_l, const int* ket_r, const int* ket_rows,
                        const int* ket_cols, int nL, const int* L_o, const int* L_l,
                        const int* L_r, const int* L_rows, const int* L_cols,
                        const int64_t* L_off, const double* L_data, int nR, const int* R_o,
                        const int* R_l, const int* R_r, const int* R_rows, const int* R_cols,
                        const int64_t* R_off, const double* R_data, void* stream,
                        dmrgcc_heff** out);
/* y += H_eff x  (the reference's accumulate-into-state semantics).
 * x, y: device vectors of psi_dim doubles.  BF16x3 on a real plan tiled
 * 64x64 runs the pre-split form: the plan's L and R blocks are packed once
 * into bf16 hi/lo pairs, x is packed per apply, and every term is scaled in
 * FP64 after its product.  A batch (dmrgcc_heff_batch_*) and the 32x32
 * tiling split the FP64 operands on the fly, with the scale folded into the
 * split; the two forms agree to the BF16x3 accuracy, not bit for bit.  The
 * environment variable DMRGCC_PRESPLIT=0, read when a plan is built, keeps
 * that plan on the on-the-fly form. */
int dmrgcc_heff_apply(dmrgcc_heff* h, const double* x, double* y, int precision);
/* y = alpha * H_eff x + beta * y on `stream` (NULL: the plan's own stream),
 * returning without waiting: only kernel launches (no copies, memsets or
 * host synchronisation), so the call can be captured into a HIP graph.
 * alpha = beta = 1 is the reference's accumulate (state += H psi); beta = 0
 * does not read y.  The caller orders the work that produces x and y before
 * it on `stream`; x and y must not overlap.  No counterpart in the
 * reference (MultiplyEffectHamilton is synchronous, on host memory). */
int dmrgcc_heff_apply_async(dmrgcc_heff* h, const double* x, double* y, double alpha, double beta,
                            int precision, void* stream);
/* info[0] = psi_dim, [1] = F_ref (reference-order flops), [2] = F_exec,
 * [3] = minimum HBM bytes per matvec, [4] = stage-1 tasks, [5] = tiles,
 * [6] = stage-1 flops, [7] = stage-2 flops, [8] = stage-1 tiles,
 * [9] = stage-2 tiles, [10] = 1 when a BF16x3 matvec runs on the pre-split
 * bf16 pair operands (gemm_bf16x3_pre_kernel), else 0.  info holds 11
 * doubles. */
int dmrgcc_heff_info(const dmrgcc_heff* h, double* info);
int dmrgcc_heff_destroy(dmrgcc_heff* h);

/* Tile autotuner (no counterpart in the reference, whose MKL calls have no
 * tiles): re-tiles the plan with every candidate (tile mode 64 = 64x64
 * workgroup tiles or 32 = 32x32 wave tiles; split-K threshold), times `reps`
 * back-to-back matvecs of `precision` each with HIP events and keeps the
 * fastest (the default 64 / 256 unless another candidate is more than 3 %
 * faster).  Writes *ncand entries (cap >= 8) of tile_mode / split_k / ms.
 * Re-tiling invalidates H_eff batches built from this plan. */
int dmrgcc_heff_tune(dmrgcc_heff* h, int precision, int reps, int cap, int* ncand, int* tile_mode,
                     int* split_k, double* ms);
/* The plan's current tiling. */
int dmrgcc_heff_config(const dmrgcc_heff* h, int* tile_mode, int* split_k);
/* Process-wide tiling for GEMM batches built from now on (the autotuner's
 * choice applied to later sites and sweeps); -1 restores the default of
 * that field (environment DMRGCC_TILE / DMRGCC_SPLIT_K, else 64 / 256). */
int dmrgcc_gemm_config(int tile_mode, int split_k);

/* Lanczos ground state (RealTensorLanczos::LanczosMethod,
 * src/dmrg/realdmrg/real_tensor_lanczos.cc:14-169) on the plan's site:
 * `ket` (device, psi_dim) is the start vector and receives the Ritz vector.
 * precision != 0 runs the iterations mixed and, if refine, returns the FP64
 * Rayleigh quotient of the Ritz vector.  iters_out may be NULL. */
int dmrgcc_lanczos(dmrgcc_heff* h, double* ket, int num_iter, int precision, int refine,
                   double* energy, int* iters_out);

/* ---- device buffers (the boundary's upload / download of blocks) ------- */
/* Device memory from the library's caching pool; upload / download copy
 * `bytes` between host and device memory on `stream` (NULL: the null
 * stream) and return when the copy is done. */
int dmrgcc_alloc(int64_t bytes, void** dev);
int dmrgcc_free(void* dev);
int dmrgcc_upload(void* dev, const void* host, int64_t bytes, void* stream);
int dmrgcc_download(void* host, const void* dev, int64_t bytes, void* stream);

/* ---- block-sparse result handle (environment updates) ------------------ */
typedef struct dmrgcc_blocks dmrgcc_blocks;
int dmrgcc_blocks_count(const dmrgcc_blocks* b);
/* Block i: channel, l, r, rows, cols, element offset into the data buffer. */
int dmrgcc_blocks_get(const dmrgcc_blocks* b, int i, int* o, int* l, int* r, int* rows, int* cols,
                      int64_t* off);
const double* dmrgcc_blocks_data(const dmrgcc_blocks* b);
/* Copy all block data (sum rows*cols entries; 2 doubles each for a complex
 * result) to host memory. */
int dmrgcc_blocks_download(const dmrgcc_blocks* b, double* host);
int dmrgcc_blocks_destroy(dmrgcc_blocks* b);

/* Environment update (LeftComputeTensorContraction :782-940 for leigh 0,
 * RightComputeTensorContraction :959-1119 for leigh 1): the site
 * description as for dmrgcc_heff_create plus the ket data (device) and one
 * environment (L for leigh 0, R for leigh 1).  Output: the next environment,
 * one block-sparse matrix per channel (o2 for leigh 0, o1 for leigh 1). */
int dmrgcc_env_update(int leigh, int d, int lb, int rb, const double* W, int nl, int nr,
                      const int* phys, int nket, const int* ket_l, const int* ket_r,
                      const int* ket_rows, const int* ket_cols, const double* ket_data, int nE,
                      const int* E_o, const int* E_l, const int* E_r, const int* E_rows,
                      const int* E_cols, const int64_t* E_off, const double* E_data, void* stream,
                      dmrgcc_blocks** out);

/* Complex environment update (ComplexTensorContraction::
 * {Left,Right}ComputeTensorContraction, complex_tensor_contraction.cc:803,
 * :980; the bra enters as the Hermitian conjugate of the ket): arguments as
 * dmrgcc_env_update with interleaved complex W, ket and environment data;
 * the result handle then holds 2 doubles per entry. */
int dmrgcc_zenv_update(int leigh, int d, int lb, int rb, const double* W, int nl, int nr,
                       const int* phys, int nket, const int* ket_l, const int* ket_r,
                       const int* ket_rows, const int* ket_cols, const double* ket_data, int nE,
                       const int* E_o, const int* E_l, const int* E_r, const int* E_rows,
                       const int* E_cols, const int64_t* E_off, const double* E_data, void* stream,
                       dmrgcc_blocks** out);

/* ---- canonical form of one site ----------------------------------------
 * RealTensorLattice::LeftCanonicalTensorLattice (leigh 0,
 * real_tensor_lattice.cc:558-676) / RightCanonicalTensorLattice (leigh 1,
 * :794-911) with (max_dim, canonical_precision) on a site given as its
 * sector dimensions (ldim: nl, rdim: nr), ket block list and device ket
 * data (interleaved complex when is_complex).  Per sector the stacked
 * blocks are decomposed by the sector SVD, ComputeTruncateDim
 * (:501-555) picks the kept states, the orthonormal factor replaces the ket
 * blocks and Sigma V^T (left) / U Sigma (right) forms the canonical tensor,
 * one block per sector (l = r = sector).  Outputs: the new ket and
 * canonical tensor (channel 0 of the handles), the new sector dimensions
 * (right dims for leigh 0: nr entries; left dims for leigh 1: nl), the match
 * dimensions and the discarded weight (may be NULL).  Signs: the largest
 * |entry| of every kept orthonormal vector is positive (real positive). */
int dmrgcc_lattice_canonical(int leigh, int is_complex, int nl, int nr, const int* ldim, const int* rdim,
                             int nket, const int* ket_l, const int* ket_r, const int* ket_rows,
                             const int* ket_cols, const double* ket_data, int max_dim, double canpre,
                             void* stream, dmrgcc_blocks** ket_out, dmrgcc_blocks** can_out, int* dims_out,
                             int* match_dim, double* discarded);

/* ---- sector SVD (RealMatrix::SVDMatrix, real_matrix.cc:588-620) -------- */
/* Batched thin SVD of n row-major device matrices A_i (M_i x N_i), one-sided
 * Jacobi on the device.  Outputs (device): U_i (M_i x k_i), S_i (k_i,
 * descending), Vt_i (k_i x N_i), k_i = min(M_i, N_i).  precondition selects
 * the QR preconditioning (matrices with k_i >= 48; about a third of the
 * Jacobi sweeps): 0 = none, plain one-sided Jacobi: the short-side factor
 * is orthonormal to working precision, the long-side one (the rotated
 * columns divided by sigma) only on singular values well above eps * |A|;
 * 1 = A = Q R and Jacobi on R^T: the long-side factor (U when M_i >= N_i,
 * else Vt) is orthonormal to working precision, the short-side one only for
 * singular values well above eps * |A| (sigma = 0 columns come back zero);
 * 2 = a second QR of R^T and Jacobi on its R^T: the short-side factor is
 * orthonormal to working precision, the long-side one as in 0.  The sweep's
 * canonical step uses the form whose exact factor becomes the new site
 * tensor. */
int dmrgcc_sector_svd(int n, const double* const* A, const int* M, const int* N,
                      double* const* U, double* const* S, double* const* Vt, int precondition,
                      void* stream);

/* Complex (ComplexMatrix::SVDMatrix, complex_matrix.cc:601-637; zgesdd in
 * the reference): A_i, U_i, Vt_i interleaved complex row-major (A = U S Vt,
 * Vt = V^H), S_i real descending.  Complex one-sided block Jacobi; sectors
 * with min(M, N) >= 48 are QR-preconditioned as in the sweep (environment
 * DMRGCC_ZSVD_QR=0 turns that off); the largest-modulus entry of every column
 * of U is real positive. */
int dmrgcc_zsector_svd(int n, const double* const* A, const int* M, const int* N,
                       double* const* U, double* const* S, double* const* Vt, void* stream);

/* ---- full sweep (RealTensorRundmrg::Run, real_tensor_rundmrg.cc:115-245;
 *      ComplexTensorRundmrg::Run, complex_tensor_rundmrg.cc:115-245) ---- */
/* Runs the model from DefineTensorSpace through nhalf half sweeps; a complex
 * model runs the complex engine (complex MPS, environments, SVD and random
 * draws; complex_matrix.cc:201-208 draws two rand() per entry).
 * Per half sweep h: maxm[h], maxb[h], niter[h], canpre[h], noise[h] (the
 * table exponent; 0.1^noise is applied as in the reference).  Outputs per
 * half sweep (host, nhalf entries): energies, discarded weight, seconds.
 * stats (host, 24): [0] total seconds, [1] matvecs, [2] dense solves,
 * [3] matvec seconds, [4] matvec flops (F_ref), [5] rand() draws, [6] Lanczos
 * solves, [7] max bond dim, [8..14] seconds in Lanczos, expansion
 * contraction, expansion fill, canonicalisation, of which SVD, environment
 * update, match+merge, [15..16] largest SVD M and N, [17] SVD count,
 * [18] sum of M*N*N over SVDs, [19] seconds building H_eff plans (part of [8]),
 * [20..22] sweep autotuner runs, last tile mode, last split-K. */
int dmrgcc_run(dmrgcc_model* m, int first_block, int first_dim, int nhalf, const int* maxm,
               const int* maxb, const int* niter, const double* canpre, const double* noise,
               int precision, int refine, int verbose, double* energies, double* trunc,
               double* seconds, double* stats);
/* Two-site sweeps (the reference sweeps one site with subspace expansion;
 * this is the two-site form): per pair, Lanczos on the two-site H_eff
 * (dmrgcc_dmrg_heff2), then the SVD split of Theta per middle sector with
 * ComputeTruncateDim (maxm, canpre of the table row) moves the centre; no
 * expansion noise (the table's noise column is ignored).  Same arguments
 * and outputs as dmrgcc_run; the initial MPS is the reference's. */
int dmrgcc_run_two_site(dmrgcc_model* m, int first_block, int first_dim, int nhalf, const int* maxm,
                        const int* maxb, const int* niter, const double* canpre, const double* noise,
                        int precision, int refine, int verbose, double* energies, double* trunc,
                        double* seconds, double* stats);
/* dmrgcc_run with the reference's checkpoint files
 * (real_tensor_rundmrg.cc:44-52, :231-235): resume_path != NULL starts from
 * that tensor-space file instead of DefineTensorSpace (cache_resume);
 * record_path != NULL writes the tensor space there after every sweep
 * (cache_record).  The file format is RealTensorSpace::WriteTensorSpace's
 * (real_tensor_space.cc:182-208; see dmrgcc_dmrg_write_space). */
int dmrgcc_run_files(dmrgcc_model* m, int first_block, int first_dim, int nhalf, const int* maxm,
                     const int* maxb, const int* niter, const double* canpre, const double* noise,
                     int precision, int refine, int verbose, const char* resume_path,
                     const char* record_path, double* energies, double* trunc, double* seconds,
                     double* stats);

/* ---- resumable sweep (RealTensorRundmrg, real_tensor_rundmrg.cc:3-287) -- */
/* A handle owns the MPS, environments and rand() stream of one run.
 * create: DefineTensorSpace + Initialize (same sweep-table arguments as
 * dmrgcc_run).  advance: run site steps until the Lanczos solve of `site`
 * in half sweep `half` is next (*reached = 1) or the table ends
 * (*reached = 0; pass half = -1 to run to the end).  heff: the H_eff plan of
 * `site` on the current state (valid until the next advance); ket: device
 * pointer and length of that site's ket vector.  result: as dmrgcc_run.
 * set_precision: arithmetic of the following Lanczos solves.
 * Plans from dmrgcc_dmrg_heff / dmrgcc_dmrg_heff2 point into the run's
 * environment and site buffers: every dmrgcc_dmrg_advance and
 * dmrgcc_dmrg_destroy makes them stale, and apply / lanczos / tune / time /
 * state / batch calls on a stale plan fail with status 3 (info still
 * works).  Destroy the plan, then build a new one. */
typedef struct dmrgcc_dmrg dmrgcc_dmrg;
int dmrgcc_dmrg_create(dmrgcc_model* m, int first_block, int first_dim, int nhalf, const int* maxm,
                       const int* maxb, const int* niter, const double* canpre, const double* noise,
                       int precision, int refine, int verbose, dmrgcc_dmrg** out);
int dmrgcc_dmrg_create_files(dmrgcc_model* m, int first_block, int first_dim, int nhalf, const int* maxm,
                             const int* maxb, const int* niter, const double* canpre, const double* noise,
                             int precision, int refine, int verbose, const char* resume_path,
                             const char* record_path, dmrgcc_dmrg** out);
/* RealTensorSpace::WriteTensorSpace (real_tensor_space.cc:171-208) of the
 * current state: num_site, num_site+1, num_site-1, num_quantum, the table
 * sizes and quantum tables, then per site WriteTensorLattice
 * (real_tensor_lattice.cc:218-239: physics_dim, left sectors (index, dim),
 * right sectors, physics_index, ket blocks as WriteMatrixBlock /
 * WriteMatrix), then max_block, max_dim (int) and canonical precision,
 * noise factor (double).  int is 4 bytes, entries 8-byte doubles or 16-byte
 * Complex, native byte order. */
int dmrgcc_dmrg_write_space(dmrgcc_dmrg* h, const char* path);
int dmrgcc_dmrg_advance(dmrgcc_dmrg* h, int half, int site, int* reached);
/* Environment files: RealTensorContraction::WriteTensorContraction(leigh,
 * name) (real_tensor_contraction.cc:73-90) of `site` on the current state
 * (leigh 0: its left environment L_site, 1: its right environment R_site):
 * per channel the block count and per block left, right, rows, cols,
 * rows*cols and the entries.  env_file_read: ReadTensorContraction
 * (:92-109) of nbond channels into a block handle (channel = o). */
int dmrgcc_dmrg_write_env(dmrgcc_dmrg* h, int site, int leigh, const char* path);
int dmrgcc_env_file_read(const char* path, int nbond, int is_complex, dmrgcc_blocks** out);
int dmrgcc_dmrg_set_precision(dmrgcc_dmrg* h, int precision, int refine);
/* The tile autotuner inside the sweep (no reference counterpart; off by
 * default): the H_eff plan of the first site with psi_dim >= min_dim
 * (20000 when min_dim <= 0), and of
 * every later site whose psi_dim is 4x the last tuned one, is re-tiled with
 * every candidate (as dmrgcc_heff_tune, 3 timed matvecs each) and the fastest
 * tiling is used for the plans that follow.  stats[20..22] of the result:
 * tunes run, the last tile mode and split-K threshold chosen. */
int dmrgcc_dmrg_set_autotune(dmrgcc_dmrg* h, int on, int64_t min_dim);
int dmrgcc_dmrg_heff(dmrgcc_dmrg* h, int site, dmrgcc_heff** out);
int dmrgcc_dmrg_ket(dmrgcc_dmrg* h, int site, double** ket, int64_t* n);
/* Two-site effective Hamiltonian of (site, site+1) on the current state:
 * L_site x W_site W_site+1 x Theta x R_site+1, the matvec the reference's
 * single-site sweep does not have (it is the two-site form the hot path is
 * specified by).  Theta = A_site * B_site+1 is built on the device, one block
 * (l, r) per middle sector m, in A's block order then B's, physical index
 * p1 * d + p2; the fused MPO contracts W_site W_site+1 over their shared
 * bond.  The plan takes dmrgcc_heff_apply / _info / _time / dmrgcc_lanczos
 * like a one-site plan; dmrgcc_heff_state copies Theta (psi_dim entries)
 * into dst (device).  Valid when the run is paused before the Lanczos solve
 * of `site` (site holds the centre, site+1 is right-canonical, L_site and
 * R_site+1 are current). */
/* Batched H_eff of independent instances (no reference counterpart: the
 * reference runs one instance per process): y_i += H_i x_i for n plans in
 * one grouped launch per stage.  x[i], y[i] are device vectors of plan i,
 * bound at creation; the plans must outlive the batch.  apply with
 * use_graph != 0 captures the launches once per arithmetic into a HIP
 * graph and replays it.  time: ms[0] mean graph replay of the whole batch,
 * ms[1] mean of the same products plan by plan.  info: [0] sum of F_ref,
 * [1] tiles, [2] plans.  Each plan appears at most once (its stage buffer
 * is written by the launch): a repeated plan pointer, or a stale run plan,
 * fails with status 3.  One operator applied to several vectors takes one
 * plan per vector. */
typedef struct dmrgcc_heff_batch dmrgcc_heff_batch;
int dmrgcc_heff_batch_create(int n, dmrgcc_heff* const* plans, const double* const* x, double* const* y,
                             dmrgcc_heff_batch** out);
int dmrgcc_heff_batch_apply(dmrgcc_heff_batch* b, int precision, int use_graph);
int dmrgcc_heff_batch_time(dmrgcc_heff_batch* b, int precision, int reps, double* ms);
int dmrgcc_heff_batch_info(const dmrgcc_heff_batch* b, double* info);
int dmrgcc_heff_batch_destroy(dmrgcc_heff_batch* b);
int dmrgcc_dmrg_heff2(dmrgcc_dmrg* h, int site, dmrgcc_heff** out);
int dmrgcc_heff_state(const dmrgcc_heff* h, double* dst);
/* Export the current blocks of `site` to host memory (tests, CPU baseline):
 * which 0 = ket, 1 = left environment, 2 = right environment.  With
 * o == NULL only *count and *total (elements) are written.  Offsets count
 * entries; on a complex model data receives 2 * total doubles (re, im)
 * and the ket vector holds interleaved pairs. */
int dmrgcc_dmrg_site_blocks(dmrgcc_dmrg* h, int site, int which, int* count, int64_t* total, int* o,
                            int* l, int* r, int* rows, int* cols, int64_t* off, double* data);
/* Physical-index table of `site` (nl x nr, -1 where no block). */
int dmrgcc_dmrg_phys(dmrgcc_dmrg* h, int site, int* nl, int* nr, int* phys);
int dmrgcc_dmrg_result(dmrgcc_dmrg* h, int* nhalf_done, double* energies, double* trunc,
                       double* seconds, double* stats);
int dmrgcc_dmrg_destroy(dmrgcc_dmrg* h);

/* Time `reps` H_eff applications (y += H x on internal buffers) with HIP
 * events on the plan's stream.  ms (5 entries): [0] mean stage-1 (T = K R)
 * launch, [1] mean stage-2 (out += w L T) launch including its split-K
 * reduction, [2] mean whole matvec, back to back, [3] / [4] mean stage-1 /
 * stage-2 GEMM kernel alone, back to back. */
int dmrgcc_heff_time(dmrgcc_heff* h, int precision, int reps, double* ms);

/* Time `reps` environment updates of `site` on the run's current state:
 * leigh 1 = LeftComputeTensorContraction (real_tensor_contraction.cc:
 * 782-940) from L_site, leigh 2 = RightComputeTensorContraction (:959-1119)
 * from R_site; the results are discarded.  out (5 entries): [0] / [1] mean
 * GPU ms of the F launch and of the output launch (HIP events), [2] bytes
 * with every used old-environment block, ket block and new block read or
 * written once, [3] the same plus the intermediate F written and read,
 * [4] executed flops. */
int dmrgcc_dmrg_env_time(dmrgcc_dmrg* h, int site, int leigh, int reps, double* out);
/* Lanczos vector kernels on n-double device vectors, mean ms over `reps`
 * back-to-back launches (HIP events): [0] the two-pass dot (reads 16 n B),
 * [1] axpy with a device scalar (24 n B), [2] the Lanczos step
 * psi_next = h / beta, h = -beta psi (32 n B).  No reference counterpart (its MKL BLAS-1 calls). */
int dmrgcc_vector_time(int64_t n, int reps, double* ms);

#ifdef __cplusplus
}
#endif

#endif /* DMRGCC_HIP_H_ */
