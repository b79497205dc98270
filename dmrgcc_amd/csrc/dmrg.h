// Host-side DMRG engine: block-sparse site tensors and environments resident
// in HBM, host bookkeeping of quantum-number sectors, device kernels for all
// numerics.  Mirrors the reference classes RealTensorLattice,
// RealTensorContraction, RealTensorLanczos, RealTensorSpace,
// RealTensorNetwork and RealTensorRundmrg (src/dmrg/realdmrg/).
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "common.h"
#include "kernels.h"
#include "model.h"
#include "rng.h"

namespace dmrg {

// One block of a block-sparse matrix: sector positions (l, r), shape, element
// offset into the tensor's device buffer (row-major).
struct Block {
  int l, r, rows, cols;
  int64_t off;
};

// RealMatrixBlock: blocks concatenated in block order in one HBM buffer; the
// flat buffer is exactly the reference's VectorizeMatrixBlock layout.
// cs = 2 holds complex entries interleaved (re, im) (ComplexMatrixBlock);
// offsets and `size` count entries, the buffer holds cs * size doubles.
struct BlockTensor {
  std::vector<Block> blk;
  int64_t size = 0;
  int cs = 1;
  DevBuf<double> data;
  const double* ext = nullptr;  // non-owning view of caller memory (C ABI)

  int n() const { return (int)blk.size(); }
  void define(const std::vector<int>& l, const std::vector<int>& r);
  void allocate();  // offsets from shapes, fresh (uninitialised) buffer
  int find(int l, int r) const;
  std::vector<int> same(int leigh, int target) const;
  double* base() const { return ext ? const_cast<double*>(ext) : data.data(); }
  double* ptr(int i) const { return base() + cs * blk[i].off; }
  int64_t elems(int i) const { return (int64_t)blk[i].rows * blk[i].cols; }
  int64_t doubles() const { return cs * size; }
};

using Env = std::vector<BlockTensor>;  // one block-sparse matrix per MPO channel

// Hashed (l, r) -> block index of a block list as it is when built; find
// returns the first match, as BlockTensor::find (FindMatrixBlock,
// real_matrix_block.cc:314-380) does, in O(1) for the contraction loops.
class BlockIndex {
 public:
  explicit BlockIndex(const BlockTensor& t);
  int find(int l, int r) const {
    const auto it = m_.find(key(l, r));
    return it == m_.end() ? -1 : it->second;
  }

 private:
  static int64_t key(int l, int r) { return ((int64_t)l << 32) | (uint32_t)r; }
  std::unordered_map<int64_t, int> m_;
};

// RealTensorLattice: one MPS site.
struct Lattice {
  int d = 0;
  std::vector<int> lb, ld, rb, rd;   // sector QN-table indices and dims
  std::vector<std::vector<int>> phys; // [nl][nr], -1 where no block
  BlockTensor ket;
  std::vector<int> match_dim;
  bool has_can = false;
  BlockTensor can;  // canonical_tensor_: one block per sector
  int nl() const { return (int)lb.size(); }
  int nr() const { return (int)rb.size(); }
  void define_sectors(const std::vector<int>& l, const std::vector<int>& r,
                      const std::vector<int>& ldim, const std::vector<int>& rdim);
};

// Per-instance device resources.
struct Context {
  hipStream_t stream = nullptr;
  bool own_stream = false;
  GlibcRand rng;
  DevBuf<double> partial, scal;
  DevBuf<CopyTask> copy_scratch;
  DevBuf<ColGatherTask> gather_scratch;
  DevBuf<SvdProblem> svd_scratch;
  DevBuf<int> svd_sweeps;
  DevBuf<int> ivec;
  DevBuf<double> dvec;
  DevBuf<double> work;
  explicit Context(hipStream_t s = nullptr);
  ~Context();
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  // waits for the stream, then drops the buffers retired since the last sync
  void sync() {
    DMRG_HIP(hipStreamSynchronize(stream));
    retired.clear();
  }
  // keeps a device buffer alive until the next sync(): for buffers that work
  // already queued on the stream may still use when their owner lets go
  template <typename T>
  void retire(DevBuf<T>&& b) {
    if (b.data()) retired.push_back(std::make_shared<DevBuf<T>>(std::move(b)));
  }
  void retire(GemmBatch& g) {
    retire(std::move(g.d_desc));
    retire(std::move(g.d_splits));
    retire(std::move(g.part));
  }
  std::vector<std::shared_ptr<void>> retired;
  // pinned host staging of at least n doubles, reused between calls (grown
  // when needed; growing discards the contents)
  double* pinned(int64_t n);
  // pinned host memory of the Lanczos loop (4 * 200 + 200 doubles: per
  // iteration alpha and beta^2, then the Ritz coefficients), separate from
  // the staging above so that it never waits for another copy
  double* lanczos_host();
  double* lanczos_host_dev();  // the device's address of lanczos_host()
  // at least n events (timing disabled), reused between calls
  std::vector<hipEvent_t> events(int n);
  // Device rand() stream state (random_matrices_device, dmrg.cpp): the chunk
  // coefficients made so far (host copy and device table), the window, the
  // raw draws and the matrix descriptors
  std::vector<uint32_t> rng_coef_host;
  DevBuf<uint32_t> rng_coef, rng_w0;
  DevBuf<uint8_t> rng_raw;
  DevBuf<RngDraw> rng_draws;

 private:
  double* pinned_ = nullptr;
  int64_t pinned_n_ = 0;
  double* lz_host_ = nullptr;
  double* lz_host_dev_ = nullptr;
  std::vector<hipEvent_t> events_;
};

// H_eff matvec plan of one site (MultiplyEffectHamilton,
// real_tensor_contraction.cc:1223-1291): stage 1 T = K(k,l) R[o2](l,j) per
// (o2, R block, ket block); stage 2 out(i,j) += sum w L[o1](i,k) T per output
// block.  Built once per Lanczos solve, launched once per iteration.
class HeffPlan {
 public:
  void build(const Lattice& lat, const SiteOperator& W, const Env& L, const Env& R,
             hipStream_t stream);
  // General form: state blocks (l, r) with physical index bphys[k] (nr right
  // sectors); blocks may share a sector pair.  The two-site plan uses it
  // with the fused MPO (fuse_sites) and p = p1 * d + p2.
  void build_blocks(const BlockTensor& ket, const std::vector<int>& bphys, int nr, const SiteOperator& W,
                    const Env& L, const Env& R, hipStream_t stream);
  // y (+)= H x ; x, y in ket-vector layout (psi_dim elements).
  void apply(const double* x, double* y, int precision, hipStream_t stream);
  // launch one stage only (timing); gemm_only skips the split-K reduction
  void apply_stage(int stage, const double* x, double* y, int precision, hipStream_t stream,
                   bool gemm_only = false);
  double flops() const { return s1_.flops + s2_.flops; }  // executed
  double fref() const { return fref_; }                      // reference order
  int64_t psi_dim() const { return psi_dim_; }               // entries (complex: pairs)
  bool complex() const { return cplx_; }
  int64_t doubles() const { return (cplx_ ? 2 : 1) * psi_dim_; }
  int64_t tasks() const { return (int64_t)s1_.items.size(); }
  int64_t bytes_min() const { return bytes_min_; }
  const GemmBatch& stage1() const { return s1_; }
  const GemmBatch& stage2() const { return s2_; }

  // Tile autotuner: re-tile both stages with every (tile mode, split-K)
  // candidate, time `reps` back-to-back matvecs of `precision` each, keep
  // the fastest (the default 64x64 / split-K 256 unless another is more
  // than 3 % faster).  Returns the timing of every candidate (in order); the
  // plan ends up tiled with the chosen one.
  struct Tuning { int tile_mode, split_k; double ms; };
  std::vector<Tuning> tune(int precision, int reps, hipStream_t stream);
  // re-tile both stages (tile_mode 32 or 64, split-K threshold)
  void retile(int tile_mode, int split_k, hipStream_t stream);
  int tile_mode() const { return s1_.tile_mode; }
  int split_k() const { return s1_.split_k; }
  // bumped by every build and retile: batches built from the plan check it
  int generation() const { return gen_; }
  // tiling of the next build (-1: the process-wide default of that field)
  void set_tiling(int tile_mode, int split_k) { want_tile_ = tile_mode; want_split_ = split_k; }
  // Pre-split BF16x3 form of the next build (default on; the environment
  // variable DMRGCC_PRESPLIT=0 turns it off): the used L and R blocks are
  // packed once into bf16 pair groups, x is packed per apply, stage 1 writes
  // T^T as pair groups, and the BF16x3 apply on 64x64 tiles runs both
  // stages on the pre-split kernel.  FP64 and FP32 keep the FP64 T.
  void set_presplit(bool on) { want_pre_ = on; }
  bool presplit() const { return pre_; }
  // whether an apply in `precision` runs on the pre-split operands
  bool use_pre(int precision) const { return pre_ && precision == kBF16x3 && !s1_.wave_tiles; }

 private:
  GemmBatch s1_, s2_;
  int want_tile_ = -1, want_split_ = -1;
  // pre-split BF16x3 stages: their batches, items and terms (for retile), and
  // the pair-group buffers of x, T^T and the used L and R blocks
  bool want_pre_ = true, pre_ = false;
  GemmBatch p1_, p2_;
  std::vector<GemmItem> pitems1_, pitems2_;
  std::vector<GemmContrib> pterms1_, pterms2_;
  DevBuf<double> Xs_, Ts_, LRs_;
  DevBuf<PackTask> d_xpack_;
  DevBuf<PackTask> d_lrpack_;  // the one-off L/R packing's descriptors (kept: no wait at build)
  int nxpack_ = 0;
  int64_t xpack_groups_ = 0;
  // items and terms of both stages before tiling, kept for retile()
  std::vector<GemmItem> items1_, items2_;
  std::vector<GemmContrib> terms1_, terms2_;
  int gen_ = 0;
  DevBuf<double> T_;
  bool cplx_ = false;
  int64_t psi_dim_ = 0;
  int64_t bytes_min_ = 0;
  double fref_ = 0;
};

// Many H_eff plans (independent instances, any mix of sizes) applied in one
// grouped launch per stage: y_i += H_i x_i for every plan, with x_i / y_i
// bound at build time.  apply_graph captures those launches once per
// arithmetic into a HIP graph and replays it.  The plans must outlive the
// batch.
class HeffBatch {
 public:
  ~HeffBatch();
  void build(const std::vector<HeffPlan*>& plans, const std::vector<const double*>& x,
             const std::vector<double*>& y, hipStream_t stream);
  void apply(int precision, hipStream_t stream);
  void apply_graph(int precision, hipStream_t stream);
  // the same products plan by plan, two launches each (reference timing)
  void apply_each(int precision, hipStream_t stream);
  double fref() const { return fref_; }
  int64_t tiles() const { return (int64_t)(s1_.tiles.size() + s2_.tiles.size()); }

 private:
  GemmBatch s1_, s2_;
  std::vector<HeffPlan*> plans_;
  std::vector<int> gens_;  // the plans' generations at build time
  void check_plans() const;
  std::vector<const double*> x_;
  std::vector<double*> y_;
  double fref_ = 0;
  hipGraph_t graph_[3] = {nullptr, nullptr, nullptr};
  hipGraphExec_t exec_[3] = {nullptr, nullptr, nullptr};
};

// Two-site MPO W_j W_{j+1} contracted over the shared bond:
// W[o1][o3](p1 d + p2, q1 d + q2) = sum_o2 W_j[o1][o2](p1, q1) W_{j+1}[o2][o3](p2, q2)
// (physical dimension d^2; complex when either factor is).
SiteOperator fuse_sites(const SiteOperator& a, const SiteOperator& b);

// Two-site state of sites (j, j+1): Theta = A_j * B_{j+1} per middle sector,
// its blocks' physical indices and the fused MPO.
struct TwoSite {
  BlockTensor theta;
  std::vector<int> bphys;
  std::vector<std::pair<int, int>> src;  // (A block, B block) of each Theta block
  SiteOperator W2;
};

// Environment updates (LeftComputeTensorContraction :782-940,
// RightComputeTensorContraction :959-1119).  With `tm`, the two GEMM
// launches (F = B^T L / F = W R B-part, then the new environment) are timed
// with HIP events and their algorithmic bytes and flops are counted.
struct EnvTiming {
  double ms_f = 0, ms_out = 0;  // GPU time of the F launch and of the output launch
  double bytes_min = 0;         // each used old-environment block, ket block and new block once
  double bytes_exec = 0;        // bytes_min plus the intermediate F written and read back
  double flops = 0;             // executed real flops of both launches
};
Env left_env_update(const Lattice& lat, const SiteOperator& W, const Env& L, Context& cx,
                    EnvTiming* tm = nullptr);
Env right_env_update(const Lattice& lat, const SiteOperator& W, const Env& R, Context& cx,
                     EnvTiming* tm = nullptr);
// Environment files (RealTensorContraction::Write/ReadTensorContraction,
// real_tensor_contraction.cc:73-109): per channel RealMatrixBlock::
// WriteMatrixBlock (block count, then per block left, right and
// RealMatrix::WriteMatrix: rows, cols, rows*cols, entries; complex entries
// as the 16-byte Complex).  The reader takes the channel count, as the
// reference's object does.
void WriteTensorContraction(const Env& env, const std::string& path, hipStream_t stream);
Env ReadTensorContraction(const std::string& path, int nbond, int cs, hipStream_t stream);

// Random matrices of the expansion in draw order (RandomMatrix,
// real_matrix.cc:201-208; complex_matrix.cc:201-208 with cs = 2): matrix k
// is rows[k] x cols[k], row-major (complex: interleaved), concatenated in
// `out`; with w != nullptr its diagonal is then set to 0.1 * w[k] (complex:
// w holds re, im pairs), as the expansion does (real_tensor_contraction.cc:
// 397-399).  Large sets are generated by several host threads from
// jump-ahead copies of `rng`, which ends after the last draw.
void random_matrices(GlibcRand& rng, int n, const int* rows, const int* cols, int cs, const double* w,
                     double* out);
// cs = 2: complex (1 + 0i) boundary blocks
Env boundary_env(int nbond, Context& cx, int cs = 1);

// Dense symmetric eigen-decomposition (cyclic Jacobi), ascending eigenvalues,
// eigenvectors in columns of the row-major n x n `vec`.  Used for the Lanczos
// tridiagonal and the dense edge-site H_eff (RealSymMatrixDiag,
// src/util/general.cc:47-64), both FP64 on the host.
void sym_eig(int n, std::vector<double> a, std::vector<double>& eval, std::vector<double>& vec);
// Lowest eigenvalue and its unit eigenvector (largest |component| positive).
void sym_eig_lowest(int n, std::vector<double> a, double& e0, std::vector<double>& v0);

// ComputeTruncateDim (real_tensor_lattice.cc:501-555): kept states per
// sector from per-sector descending singular values (modified in place when a
// sector is forced to keep one state with sigma 0).
std::vector<int> truncate_dims(int max_dim, double canpre, std::vector<std::vector<double>>& sv);

struct LanczosStats {
  int matvecs = 0, dense = 0, solves = 0;
  int tunes = 0, tile_mode = 0, split_k = 0;  // sweep autotuner: runs and the last choice
  double matvec_seconds = 0, flops = 0;
};

// Lanczos on one site's H_eff (LanczosMethod, real_tensor_lanczos.cc:14-169):
// psi_dim < 100 takes the dense path (dense H_eff + host eigensolver), else
// device Lanczos with modified Gram-Schmidt.  `ket` (device, n doubles) is
// the start vector and receives the ground vector.  Returns the energy.
double lanczos_solve(HeffPlan& plan, double* ket, int num_iter, int precision, int refine,
                     Context& cx, LanczosStats& st, int* iters = nullptr);

struct RunParams {
  int first_block = 20, first_dim = 10;
  std::vector<int> maxm, maxb, niter;          // per half sweep
  std::vector<double> canpre, noise;           // noise already 0.1^x
  int precision = kFP64;  // H_eff arithmetic inside Lanczos
  // tile autotuner inside the sweep: the H_eff plan of the first site whose
  // psi_dim reaches tune_min_dim, and of every later site whose psi_dim is 4x
  // the last tuned one, is re-tiled with each candidate and timed; later
  // plans are built with the fastest tiling
  bool autotune = false;
  int64_t tune_min_dim = 20000;
  int refine = 1;         // mixed precision: final FP64 Rayleigh quotient
  int verbose = 0;
  // cache_resume: start from this tensor-space file instead of
  // DefineTensorSpace (real_tensor_rundmrg.cc:44-52); cache_record: write
  // the tensor space to this file after every sweep (:231-235)
  std::string resume_path, record_path;
  // Two-site sweeps: Lanczos on the two-site H_eff of each pair, then the
  // SVD split of Theta moves the centre (no subspace expansion).
  bool two_site = false;
};

// Wall time per sweep phase (seconds, host clock around device-synchronised
// work) and SVD sizes seen.
struct PhaseTimes {
  double lanczos = 0, expan = 0, expand = 0, canonical = 0, svd = 0, env = 0, merge = 0;
  double plan = 0;  // H_eff plan construction, part of lanczos
  int64_t svd_max_m = 0, svd_max_n = 0, svd_calls = 0;
  double svd_work = 0;  // sum of M*N*N over sector SVDs
};

// Left (leigh 0) / right (leigh 1) canonical form of one site
// (Left/RightCanonicalTensorLattice, real_tensor_lattice.cc:558-676,
// :794-911): per sector the stacked blocks' SVD on the device,
// ComputeTruncateDim with (max_dim, canpre), the orthonormal factor into the
// ket blocks and Sigma V^T (U Sigma) into lat.can; lat.match_dim and the
// truncated sector dimensions updated.  Returns the discarded weight; with
// `ph` the SVD time and sizes are added there.
double canonicalize(Lattice& lat, int leigh, int max_dim, double canpre, Context& cx,
                    PhaseTimes* ph = nullptr);

struct RunResult {
  PhaseTimes phase;
  std::vector<double> energies;   // per half sweep (right, left, right, ...)
  std::vector<double> trunc_err;  // max discarded weight per half sweep
  std::vector<double> half_seconds;
  double total_seconds = 0;
  LanczosStats stats;
  uint64_t rng_draws = 0;
  std::vector<int> final_bond_dims;
};

// RealTensorRundmrg: owns the MPS, the environments and drives the sweeps.
// On a complex Hamiltonian it is ComplexTensorRundmrg
// (src/dmrg/complexdmrg/complex_tensor_rundmrg.cc): complex MPS and
// environments, complex random draws, complex SVD and expansion padding.
class RealTensorRundmrg {
 public:
  RealTensorRundmrg(const RealTensorSpace& space, const RealTensorHamiltonian& ham,
                    const RunParams& p, hipStream_t stream = nullptr);
  RunResult Run();
  // Resumable driving: Start() builds and right-canonicalises the initial MPS
  // (DefineTensorSpace + Initialize); Advance(h, j) runs site steps until the
  // Lanczos solve of site j in half sweep h is next (returns true) or the
  // sweep table is exhausted (returns false, pass h = -1 to run to the end).
  // Wall time spent between Advance calls is excluded from the sweep times.
  void Start();
  bool Advance(int stop_h, int stop_j);
  RunResult Result();
  int num_site() const { return L_sites_; }
  bool complex() const { return cs_ == 2; }
  void set_precision(int precision, int refine) { p_.precision = precision; p_.refine = refine; }
  void set_autotune(bool on, int64_t min_dim) {
    p_.autotune = on;
    p_.tune_min_dim = min_dim > 0 ? min_dim : 20000;
  }
  // H_eff plan on the current state of `site` (valid when the site is next).
  void BuildHeff(int site, HeffPlan& plan) {
    plan.build(lat_[site], ham_.sites()[site], L_[site], R_[site], cx_.stream);
  }
  // Two-site H_eff of (site, site+1) on the current state: L_site x W_site
  // W_{site+1} x Theta x R_{site+1}, with Theta = A_site B_{site+1} built on
  // the device (valid when `site` is the next Lanczos site: site holds the
  // centre, site+1 is right-canonical, L_site and R_site+1 are current).
  void BuildTwoSite(int site, TwoSite& ts, HeffPlan& plan);

  // Tensor-space files (RealTensorSpace::Write/ReadTensorSpace,
  // real_tensor_space.cc:171-277): the bond QN tables, every site lattice
  // (real_tensor_lattice.cc:207-298) with its ket blocks
  // (real_matrix_block.cc:216-283, real_matrix.cc:128-175; complex entries
  // are the 16-byte Complex, complex_matrix.cc:139-175) and the current
  // max_block, max_dim, canonical precision and noise factor.  Native
  // byte order, as the reference writes it.
  void WriteTensorSpace(const std::string& path);
  void ReadTensorSpace(const std::string& path);

  // pieces exposed for tests
  void DefineTensorSpace(int first_block, int first_dim);
  void Initialize();
  double LanczosMethod(int num_iter, int site);
  // the H_eff plan of `site` as LanczosMethod builds it (tiling choice and
  // autotune included)
  std::unique_ptr<HeffPlan> make_plan(int site);
  Lattice& lattice(int i) { return lat_[i]; }
  Env& left_env(int i) { return L_[i]; }
  Env& right_env(int i) { return R_[i]; }
  const SiteOperator& op(int i) const { return ham_.sites()[i]; }
  Context& context() { return cx_; }

 private:
  void set_params(int maxb, int maxm, double canpre, double noise) {
    max_block_ = maxb; max_dim_ = maxm; canpre_ = canpre; noise_ = noise;
  }
  bool expansion_active() const { return !(max_block_ < 0 || max_dim_ < 0 || noise_ > 1); }
  int check(int site, int lq, int rq) const;
  void lattice_block(int site, std::vector<int>& left, std::vector<int>& right) const;
  void tensor_index(int site, const Lattice& lat, std::vector<int>& li, std::vector<int>& ri,
                    std::vector<int>& pi) const;
  void random_blocks(Lattice& lat, const std::vector<int>& li, const std::vector<int>& ri,
                     const std::vector<int>& pi, int rows, int cols);
  double canonical(int leigh, int site);  // returns discarded weight
  void merge(int leigh, int site);
  std::vector<std::vector<int>> expan_lattice(int leigh, int site, const QNTable& op_table,
                                              Lattice& E);
  // The expansion's GEMMs and buffers when their launch is deferred (the
  // host work done ahead, while the Lanczos iterations run)
  struct ExpanWork {
    Lattice E;
    std::vector<std::vector<int>> mapping;
    GemmBatch s1, s2;
    DevBuf<double> F, Rd;
    bool pending = false;
  };
  // defer: build everything and queue what does not read the ket (the zero
  // fill of E, the random matrices, the descriptors) but leave the two
  // GEMMs to launch_expansion
  void expan_contraction(int leigh, int site, Lattice& E, const std::vector<std::vector<int>>& mapping,
                         ExpanWork* defer = nullptr);
  void launch_expansion(ExpanWork& w);
  void expand(int leigh, int site, Lattice& E);
  void match(int leigh, int site);
  void trace(int h, int j, double e, const char* what);
  void Step(int h, int j);
  void Step2(int h, int j);
  double split_two_site(int a, const TwoSite& ts, bool l2r);  // returns discarded weight
  bool started_ = false;
  int cur_h_ = 0, cur_j_ = 0;
  double half_e_ = 0, half_terr_ = 0, half_t0_ = 0, t0_ = 0;
  double paused_at_ = -1.0;  // set while stopped between Advance calls
  RunResult res_;

  const RealTensorSpace& space_;
  const RealTensorHamiltonian& ham_;
  RunParams p_;
  Context cx_;
  int L_sites_ = 0, d_ = 0;
  int cs_ = 1;  // 2: complex MPS and environments (ComplexTensorRundmrg)
  std::vector<QNTable> tables_;
  std::vector<Lattice> lat_;
  std::vector<Env> L_, R_;
  int max_block_ = 0, max_dim_ = -1;
  double canpre_ = 1e-30, noise_ = 10;
  LanczosStats stats_;
  PhaseTimes ph_;
  int64_t tuned_psi_ = 0;                  // psi_dim of the last autotuned plan
  int tune_tile_ = -1, tune_split_ = -1;   // its choice, used by later plans
  // the next step's plan, built at the end of a step while the GPU still runs
  // that step's environment update and merge
  std::unique_ptr<HeffPlan> next_plan_;
  int next_plan_site_ = -1;
};

}  // namespace dmrg
