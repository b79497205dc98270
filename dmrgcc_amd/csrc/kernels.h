// Device primitives of the DMRG hot path: descriptor structs shared by host
// and device code, and the host-side launch wrappers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "common.h"

namespace dmrg {

// Size of the fixed per-GPU device arena of the allocation-free entry point.
constexpr size_t kNoAllocArenaBytes = size_t(8) << 20;
// Device address of the current GPU's arena (looked up with
// hipGetSymbolAddress on the first call per device); *bytes = its size.
void* noalloc_arena(size_t* bytes);

// Arithmetic form of a contraction.
enum Precision : int {
  kFP64 = 0,    // FP64 MFMA (v_mfma_f64_16x16x4f64), FP64 accumulate
  kFP32 = 1,    // FP32 MFMA (v_mfma_f32_16x16x4f32), FP32 accumulate
  kBF16x3 = 2,  // split bf16 hi/lo, three v_mfma_f32_16x16x32_bf16 products,
                // FP32 accumulate
};

// One product term  scale * op(A) * op(B)  accumulated into an output item.
// A is M x K (row-major with row stride lda) or, with kTransA, stored K x M;
// B is K x N (row stride ldb) or, with kTransB, stored N x K.
// kRelA/kRelB: the operand pointer is an element offset from the launch's
// `in` base; kRelC on an item: C is an offset from the launch's `out` base.
// One uploaded plan then serves any input/output vector pair.
// kStrided2 (term) / kStrided2C (item): consecutive elements along the
// contiguous dimension sit two doubles apart (one part of an interleaved
// complex matrix); lda / ldb / ldc and offsets still count doubles.
enum : int { kTransA = 1, kTransB = 2, kRelA = 4, kRelB = 8, kStrided2 = 16 };
enum : int { kRelC = 1, kStrided2C = 2, kPairC = 4 };
// Pre-split operands (GemmBatch::presplit, BF16x3 only): every operand is
// stored k-contiguous as bf16 pair groups, 32 bytes per 8 consecutive k of
// one row: the 8 hi halves, then the 8 lo halves (x = hi + lo, hi = bf16(x),
// lo = bf16(x - hi), from x rounded to FP32), rows padded with zero k to a
// whole group.  A is M x K by rows (m), B is stored N x K by rows (n); lda /
// ldb count groups per row and K is the real k extent.  kPairC on an item:
// C is written the same way, transposed (row n, groups over m; ldc in
// groups), so that it can be the B operand of a later pre-split batch.
// Pointers and offsets of pair data count doubles (4 per group).
struct GemmContrib {
  const double* A;
  const double* B;
  int64_t lda, ldb;
  int K;
  int flags;
  double scale;
};

// One row-major output matrix C (M x N, row stride ldc) and its product terms
// [c0, c1).  accumulate=1: C += sum; accumulate=0: C = sum.
struct GemmItem {
  double* C;
  int64_t ldc;
  int M, N;
  int c0, c1;
  int accumulate;
  int flags;
};

// One workgroup: output tile (m0, n0) of item `item`.
struct GemmTile {
  int item, m0, n0, pad;
};

// Summed-K threshold above which finalize() splits an item: 256, or the
// value of the environment variable DMRGCC_SPLIT_K (0 disables splitting),
// or the value set by set_gemm_config.
int default_split_k();
// Tile shape of finalize(): 32 or 64 (0 means 64), from set_gemm_config, else
// the environment variable DMRGCC_TILE.
int default_tile_mode();
// Workgroup cap of a persistent launch (0 = one workgroup per tile): from
// set_gemm_grid_cap, else the environment variable DMRGCC_GRID_CAP, else 0.
int default_grid_cap();
void set_gemm_grid_cap(int cap);
// Process-wide GEMM configuration for batches built from now on (the tile
// autotuner's choice); a negative value restores the environment / built-in
// default of that field.
void set_gemm_config(int tile_mode, int split_k);

// Split-K of one heavy output item: partial sums of chunks 1..n-1 land in
// `part` (n-1 consecutive M x N buffers) and are added into C afterwards in
// chunk order, so results stay deterministic.
struct SplitTask {
  double* C;
  int64_t ldc;
  int M, N;
  int nparts, flags;
  const double* part;
};

// A grouped GEMM launch: items with their contribution lists.  Built on the
// host, uploaded once, launched many times (the H_eff plan reuses it for every
// Lanczos iteration).  finalize() cuts items whose summed K exceeds
// `split_k` into contribution chunks (split-K) so that no single workgroup
// carries the whole launch.
struct GemmBatch {
  std::vector<GemmItem> items;
  std::vector<GemmContrib> contribs;
  std::vector<GemmTile> tiles;
  std::vector<SplitTask> splits;
  DevBuf<double> d_desc;       // finalize: items, contribs, tiles, splits in one buffer
  DevBuf<SplitTask> d_splits;  // merge: the merged batch's split tasks
  DevBuf<double> part;
  // where launch() reads the descriptors: d_desc (finalize) or a caller's
  // arena (place)
  GemmItem* dev_items = nullptr;
  GemmContrib* dev_contribs = nullptr;
  GemmTile* dev_tiles = nullptr;
  SplitTask* dev_splits = nullptr;
  std::vector<int> part_items;  // split-K chunk items awaiting bind_parts()
  // Optional placement key per item (begin_item): items with one key have
  // their tiles on one XCD, so operands they share are fetched into one L2.
  std::vector<int> item_key;
  int split_k = default_split_k();
  // 32: 32x32 wave tiles (four independent tiles per workgroup); otherwise
  // 64x64 workgroup tiles (environment DMRGCC_TILE)
  int tile_mode = default_tile_mode();
  bool wave_tiles = false;  // set by finalize: four 32x32 tiles per workgroup
  bool presplit = false;    // operands in pair groups (64x64 tiles, BF16x3 only)
  static constexpr int kPreTerms = 128;  // terms per pre-split item (layout cuts longer lists)
  // persistent grid: at most grid_cap workgroups walk the tiles with a
  // static stride (0: one workgroup per tile, or per four wave tiles)
  int grid_cap = default_grid_cap();
  unsigned launch_grid(int units) const;
  double flops = 0;  // executed real flops (2*M*N*K summed)
  // Groups: end_group() closes the items added since the previous call into
  // one group that launch_group() runs on its own (a chain of dependent
  // GEMMs described by one upload).  Batches that use groups are not split.
  std::vector<int> group_items;  // item index where each group ends
  std::vector<int> group_tiles;  // tile index where each group ends (after finalize)

  void clear() {
    items.clear(); contribs.clear(); tiles.clear(); splits.clear();
    group_items.clear(); group_tiles.clear(); part_items.clear(); item_key.clear(); flops = 0;
  }
  // Start an output item; subsequent add() calls attach terms to it.  key >= 0:
  // the item's placement key (items sharing a key go to one XCD).
  void begin_item(double* C, int64_t ldc, int M, int N, int accumulate, int flags = 0, int key = -1);
  void add(const double* A, int64_t lda, const double* B, int64_t ldb, int K,
           int flags, double scale);
  int end_group() { group_items.push_back((int)items.size()); return (int)group_items.size() - 1; }
  // Finish construction: tile every item and upload the descriptors.
  void finalize(hipStream_t stream, int tile_m = 64, int tile_n = 64);
  // The host half of finalize(): split-K and tiling only.  Returns the
  // number of split-K partial doubles the batch needs (their items hold
  // offsets until bind_parts()).
  int64_t layout(int tile_m = 64, int tile_n = 64);
  void bind_parts(double* base);
  // Device bytes place() needs for the descriptors and `psize` partials.
  size_t placed_bytes(int64_t psize) const;
  // Instead of finalize()'s allocation and upload: put the descriptors and
  // split-K partials of a laid-out batch into `arena` (device memory of
  // `cap` bytes owned by the caller) with plain copies on `stream`.
  // Allocates nothing; launch() then reads the arena.
  void place(void* arena, size_t cap, int64_t psize, hipStream_t stream);
  // reduce = false launches the GEMM kernel only (timing of that kernel)
  void launch(hipStream_t stream, int precision, const double* in = nullptr,
              double* out = nullptr, bool reduce = true) const;
  void launch_group(hipStream_t stream, int group, int precision = 0) const;
  // Concatenate finalized batches into this one (one launch for all): the
  // relative operands and outputs of batch i are bound to ins[i] / outs[i]
  // (made absolute), split tasks kept; the source batches' buffers (stage
  // buffers, split-K partials) must outlive this batch.
  void merge(const std::vector<const GemmBatch*>& bs, const std::vector<const double*>& ins,
             const std::vector<double*>& outs, hipStream_t stream);
  bool empty() const { return tiles.empty(); }
};

// Complex (interleaved re, im) products on a real batch.  One complex term
// s * op(A) * op(B), with op = transpose (kTransA/kTransB), optional complex
// conjugation of each operand, s = sr + i si, expands into up to four real
// terms per part; zgemm_item adds the real-part and the imaginary-part items
// of one complex output C (M x N).  Leading dimensions count complex
// elements; pointers address element (0, 0) (or carry complex offsets with
// kRelA / kRelB / kRelC).
struct ZTerm {
  const double* A;
  int64_t lda;
  const double* B;
  int64_t ldb;
  int K;
  int flags;  // kTransA, kTransB, kRelA, kRelB
  bool conjA, conjB;
  double sr, si;
};
void zgemm_item(GemmBatch& g, double* C, int64_t ldc, int M, int N, int accumulate, int cflags,
                const std::vector<ZTerm>& terms);

// Pair-group packing of one matrix view: element (r, k) = src[r*srs + k*sks]
// for r < rows, k < K goes to row r, group k/8 of dst (G = ceil(K/8) groups
// per row, 4 doubles each), the padding k of the last group as zeros.
struct PackTask {
  const double* src;
  int64_t srs, sks;
  int rows, K;
  double* dst;
};
// One launch for all tasks (grid y over a task's rows x groups, max_groups
// the largest of those counts).  A non-null `base`
// makes every src an element offset from it (the input vector of an H_eff
// apply).
void launch_pack(const PackTask* dev_tasks, int ntasks, int64_t max_groups, const double* base,
                 hipStream_t stream);

// Strided 2-D copy task: dst[i*dsr + j*dsc] = alpha * src[i*ssr + j*ssc]
// for i < rows, j < cols.  src == nullptr writes zeros.  The tasks of one
// launch run concurrently: their destinations must not overlap.
struct CopyTask {
  const double* src;
  double* dst;
  int rows, cols;
  int64_t ssr, ssc, dsr, dsc;
  double alpha;
};
void launch_copies(const std::vector<CopyTask>& tasks, DevBuf<CopyTask>& scratch,
                   hipStream_t stream);

// ---- vector primitives (Lanczos) ----
// Deterministic two-pass dot product: result written to *out (device).
void dot(const double* x, const double* y, int64_t n, double* out,
         DevBuf<double>& partial, hipStream_t stream);
// Complex forms on n interleaved entries (ComplexTensorLanczos::VectorMultiply
// / VectorSubtraction, complex_tensor_lanczos.cc:163-176): out[0..1] =
// <x|y> = sum conj(x) y (deterministic); y += sign * alpha * x with alpha
// complex on the device.
void zdot(const double* x, const double* y, int64_t n, double* out,
          DevBuf<double>& partial, hipStream_t stream);
void zaxpy_dev(double* y, const double* x, int64_t n, const double* alpha,
               double sign, hipStream_t stream);
// y += sign * (*alpha) * x   (alpha on device)
void axpy_dev(double* y, const double* x, int64_t n, const double* alpha,
              double sign, hipStream_t stream);
// y = a*y + b*x (host scalars)
void axpby(double* y, const double* x, int64_t n, double a, double b,
           hipStream_t stream);
// mode 0: x /= sqrt(*nrm2)   (GramSchmidtMethod, real_tensor_lanczos.cc:198-200)
// mode 1: x *= sqrt(1/(*nrm2)) (NormalizeMatrixBlock, real_matrix_block.cc:164-171)
void scale_inv_sqrt(double* x, int64_t n, const double* nrm2, hipStream_t stream, int mode = 0);
// psi_next = h / beta; h = -beta * psi   (real_tensor_lanczos.cc:140-151)
void lanczos_step(double* psi_next, double* h, const double* psi, int64_t n, double beta,
                  hipStream_t stream);
// y = sum_j coef[j] * V[j*stride + :]   (coef on device; Ritz vector,
// real_tensor_lanczos.cc:116-123, summed in increasing j)
void combine(double* y, const double* V, int64_t stride, const double* coef, int k,
             int64_t n, hipStream_t stream);
void fill(double* x, int64_t n, double v, hipStream_t stream);

// ---- fused Lanczos vector steps ----
// Each keeps the decomposition of dot(): per-block partial sums over the
// grid of dot_grid(n) blocks (complex: (re, im) pairs of sum conj(x) y over
// entries), and every block that needs a coefficient sums the partials of
// the previous step itself, in dot()'s final order.  A projection step of
// Gram-Schmidt is then one launch, with results equal to dot + axpy.
// Partial buffers hold 2 * kDotPartials doubles; in and out must differ.
constexpr int kDotPartials = 512;
int dot_grid(int64_t n);
// part <- partials of <x|y>
void lz_dot_partial(const double* x, const double* y, int64_t n, bool cplx, double* part,
                    hipStream_t stream);
// out[0..1] = final(part_a) (alpha; real: out[1] = 0), out[2] = Re final(part_b) (beta^2);
// out_host (device-visible pinned memory, optional) receives the same three,
// then out_host[3] = 1 (the ready flag the host polls)
void lz_finals(const double* part_a, const double* part_b, int64_t n, bool cplx, double* out,
               hipStream_t stream, double* out_host = nullptr);
// The Lanczos steps after the alpha dot (basis x_j = x + j ld, j < nj, each
// of n entries; partial buffers of dot_grid(n) partials per value; part_m
// holds (cplx ? 2 : 1) nj groups).  lz_alpha_multi: alpha = final(part_a),
// h -= alpha psi, part_b <- partials of |h|^2, part_m <- partials of every
// <x_j|h>.  lz_next_all: beta from part_b (or sqrt(*beta2) on the device when
// part_b is null), c_j = <x_j|h> / beta, nxt = (h / beta - sum_j c_j x_j) / r
// with r^2 = 1 - sum_j |c_j|^2, h = -beta psi.  With slot, block 0 also forms
// alpha from part_a and writes (alpha, beta^2) to slot[0..2] and slot_host[0..2]
// (device-visible pinned memory), then slot_host[3] = 1 (the ready flag the
// host polls).
void lz_alpha_multi(double* h, const double* psi, const double* x, int64_t ld, int nj, int64_t n, bool cplx,
                    const double* part_a, double* part_b, double* part_m, hipStream_t stream);
void lz_next_all(double* nxt, double* h, const double* psi, const double* x, int64_t ld, int nj, int64_t n,
                 bool cplx, const double* part_m, hipStream_t stream, const double* beta2 = nullptr,
                 const double* part_a = nullptr, const double* part_b = nullptr, double* slot = nullptr,
                 double* slot_host = nullptr);

// Column gather for the SVD scatter: for i < rows, jj < ncol:
//   dst[i*dsr + jj*dsc] = scale[jj] * src[(r0 + i) + perm[jj]*ld]
// (src column-major).  A NaN scale writes the unit vector e_jj instead
// (completion of a zero singular vector).  perm/scale are device arrays.
// flags: kGatherComplex = interleaved complex entries (indices count
// entries), kGatherConj = conjugate what is written.
enum : int { kGatherComplex = 1, kGatherConj = 2 };
struct ColGatherTask {
  const double* src;
  int64_t ld;
  int r0, rows, ncol, flags;
  const int* perm;
  const double* scale;
  double* dst;
  int64_t dsr, dsc;
};
// Device rand() stream (the glibc TYPE_3 generator of GlibcRand): draws
// 0 .. total-1 after the state whose generation window is w0 (61 words,
// GlibcRand::window), as bytes (draw % 100).  Lane c makes draws
// [c kRngChunk, (c + 1) kRngChunk) from the window coef[c] (31 words,
// GlibcRand::chunk_powers with step kRngChunk) gives, so `out` holds
// ceil(total / kRngChunk) kRngChunk bytes and coef that many chunks.
constexpr int kRngChunk = 992;  // 8 runs of 124 = 4 x 31 draws
void rng_stream(const uint32_t* w0, const uint32_t* coef, int64_t total, uint8_t* out, hipStream_t stream);
// One random matrix of the expansion: RandomMatrix(rows, cols) from draws
// first .. (row-major at R[roff], filled column by column; complex: two
// draws per entry, the imaginary part first), then its diagonal set to
// 0.1 (w + i wi) when diag.
struct RngDraw {
  int64_t first, roff;
  int rows, cols;
  double w, wi;
  int diag, pad;
};
void rng_layout(const RngDraw* draws, int ndraws, int64_t max_entries, const uint8_t* raw, double* R, int cs,
                hipStream_t stream);
void launch_col_gather(const std::vector<ColGatherTask>& tasks, DevBuf<ColGatherTask>& scratch,
                       hipStream_t stream);

// ---- batched one-sided Jacobi SVD ----
// Each problem: column-major G (M x N, M >= N, leading dim M) is rotated in
// place into G*V with mutually orthogonal columns; V (N x N column-major) is
// initialised to identity and accumulates the rotations; sigma[j] = |G e_j|.
// Sign convention (fixes the gauge the reference leaves to LAPACK): the
// largest-magnitude entry (first one on ties) of every left singular vector
// is made positive, flipping the matching right singular vector with it.  The
// left vector is column j of G (u_in_v = 0) or of V (u_in_v = 1, when G holds
// the transposed matrix).
// Columns whose squared norm falls below (eps |A|_F)^2 are numerically zero
// and are not rotated (their singular values sit at the round-off floor); the
// squared Frobenius norm is computed on the device into fro2.
//
// precond asks for QR preconditioning when N is large enough.  With
// B_1 = A = Q_1 R_1 and B_k = R_{k-1}^T = Q_k R_k (blocked Householder), the
// Jacobi runs on X = R_K^T = W J^T, which is closer to diagonal with every
// QR (a third of the sweeps or less); both forms return G = U*Sigma and V as
// above:
//   1 (K = 3): U = Q_1 Q_3 J, G <- U*Sigma, V <- Q_2 W / sigma.  U is
//      orthonormal to working precision for every singular value; V only for
//      singular values well above eps |A| (zero sigma gives a zero column).
//      For callers that take the orthonormal factor from G.
//   2 (K = 2): V <- Q_2 J (orthonormal to working precision, as in the plain
//      Jacobi) and G <- Q_1 W.  For callers that take it from V.
struct SvdProblem {
  double* G;
  double* V;
  double* sigma;
  double* fro2;
  int M, N;
  int u_in_v;
  int precond;
};
void jacobi_svd(const std::vector<SvdProblem>& probs, DevBuf<SvdProblem>& scratch,
                DevBuf<int>& sweeps_out, hipStream_t stream, int max_sweeps = 60);

// Complex form (the SVD of ComplexMatrix::SVDMatrix, complex_matrix.cc:601-637):
// G and V hold interleaved complex entries (G: M x N column-major, leading
// dimension M entries); one-sided block Jacobi with complex rotations, G <-
// G V, V unitary, sigma_j = |G e_j|.  Phase convention: the largest-modulus
// entry of every left vector is made real positive, the right vector taking
// the same phase.  precond != 0 with N >= 48: QR-preconditioned (A P = Q_1
// R_1 with the columns presorted by norm, R_1^H = Q_2 R_2, Jacobi on R_2^H);
// the left factor is then formed from Q_1 and V = P Q_2 J.
void zjacobi_svd(const std::vector<SvdProblem>& probs, DevBuf<SvdProblem>& scratch,
                 DevBuf<int>& sweeps_out, hipStream_t stream, int max_sweeps = 80);

}  // namespace dmrg
