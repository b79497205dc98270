// HIP kernels of the DMRG hot path for gfx950 (MI355X, CDNA4).
//
// * grouped block GEMM: every sector product of the H_eff matvec, the
//   environment updates, the merges and the expansion runs through one
//   grouped launch described by uploaded descriptors (one workgroup per
//   output tile, in a cost-sorted, XCD-aware order).  A workgroup owns one
//   64x64 output tile of one output block and accumulates all of that
//   block's product terms
//   (K-concatenation), so there are no atomics and results are deterministic.
//   Four waves per workgroup, each a 32x32 sub-tile = 2x2 MFMA 16x16 blocks.
//   FP64 uses v_mfma_f64_16x16x4f64, FP32 v_mfma_f32_16x16x4f32, BF16x3
//   splits both operands into bf16 hi+lo and issues three
//   v_mfma_f32_16x16x32_bf16 per block (hi*hi + hi*lo + lo*hi).
// * strided block copies (stacking for the SVD, padding, scatter)
// * Lanczos vector primitives with deterministic two-pass reductions
// * batched one-sided (Hestenes) Jacobi SVD, one workgroup per sector.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <unordered_map>

#include "kernels.h"

// Fixed device arena of the allocation-free entry point
// (dmrgcc_block_gemm_noalloc): GEMM descriptors and split-K partials.  Every
// GPU the library runs on holds its own copy until the process ends; calls
// on one GPU share it.
static __device__ __align__(256) unsigned char scratch[dmrg::kNoAllocArenaBytes];

namespace dmrg {

namespace {

constexpr int kTile = 64;     // output tile edge
constexpr int kXcds = 8;      // MI355X: 8 XCDs, workgroups dispatched round-robin
constexpr int kThreads = 256; // 4 waves

typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));

// ---------------------------------------------------------------------------
// FP64 / FP32 MFMA tile kernel.  LDS holds A and B k-major (As[k][m],
// Bs[k][n]) so that an MFMA operand fetch (lane l reads row k = l/16, column
// l%16) is a contiguous 16-element run per quarter wave.
// ---------------------------------------------------------------------------
template <typename T>
struct MfmaAcc;
// Output element v of a lane sits in column lane%16 and a row that differs
// between the two instructions (measured on gfx950): the f64 form interleaves
// rows (lane/16 + 4v), the f32 form blocks them (4*(lane/16) + v).
template <>
struct MfmaAcc<double> {
  typedef d4v V;
  __device__ static V mma(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <>
struct MfmaAcc<float> {
  typedef f4v V;
  __device__ static V mma(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

// Operand staging shared by the GEMM kernels.  One K-step covers KT = 32
// columns of A (64 x 32 tile) and rows of B (32 x 64 tile); each of the 256
// threads holds 8 elements of each, prefetched into registers while the MFMAs
// of the previous step run.  Element q of thread tid:
//   A, not transposed: m = tid/16 + 16*(q/2), k = 2*(tid%16) + q%2
//   A, transposed:     k = tid/8,             m = 8*(tid%8) + q   (8 consecutive m)
//   B, not transposed: k = tid/8,             n = 8*(tid%8) + q
//   B, transposed:     n = tid/16 + 16*(q/2), k = 2*(tid%16) + q%2
// For the operands stored with k contiguous (A not transposed, B transposed)
// sixteen consecutive lanes read one row's 32 k, so one load instruction
// covers four whole 256-byte row segments; a thread-per-row run of 8 k would
// touch 32 cache lines per instruction for 512 bytes.
constexpr int kKT = 32;

struct StepRegs {
  double a[8], b[8];
};

// Operands live in device (global) memory.  The descriptors carry generic
// pointers, which the compiler would load through FLAT instructions; FLAT
// loads count against the LDS wait counter too, so every LDS read of a K-step
// would also wait for the next step's prefetch.  Loads go through global
// pointers instead.
typedef const double __attribute__((address_space(1)))* gcdptr;
typedef double __attribute__((address_space(1)))* gdptr;
__device__ inline gcdptr gload_ptr(const double* p) { return (gcdptr)p; }
__device__ inline gdptr gstore_ptr(double* p) { return (gdptr)p; }

// Workgroup barrier that orders LDS traffic only: __syncthreads() also waits
// for every global load in flight (its fence covers all address spaces),
// which would drain the operand loads issued for later K-steps.
__device__ inline void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}


// v[q] = sc * p[q * step] for q < nv, 0 for q >= nv (nv may be <= 0).  A
// partial run reads clamped addresses (its last valid element) without
// per-element branches; a run with no valid element issues no loads.  p must
// address a valid element whenever nv > 0.
__device__ inline void load_run8(gcdptr p, int64_t step, int nv, double sc, double (&v)[8]) {
  if (nv >= 8) {
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = p[q * step] * sc;
  } else if (nv > 0) {
    const int last = nv - 1;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const double x = p[(q < last ? q : last) * step];
      v[q] = q < nv ? x * sc : 0.0;
    }
  } else {  // the whole run lies outside the operand: no loads
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = 0.0;
  }
}

// Row-major operand X (rows r, k contiguous with element stride es): v[2i],
// v[2i+1] = sc * X[r0 + tid/16 + 16 i][k0 + 2*(tid%16) + {0, 1}], zero outside
// R rows and K columns; rows outside R issue no loads.
__device__ inline void load_rows_k(gcdptr X, int64_t ld, int es, int r0, int R, int k0, int K,
                                   int tid, double sc, double (&v)[8]) {
  const int kk = k0 + 2 * (tid & 15);
  const int ka = kk < K ? kk : K - 1, kb = kk + 1 < K ? kk + 1 : K - 1;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = r0 + (tid >> 4) + 16 * i;
    double x0 = 0.0, x1 = 0.0;
    if (r < R) {
      const gcdptr row = X + (int64_t)r * ld;
      x0 = row[(int64_t)ka * es];
      x1 = row[(int64_t)kb * es];
    }
    v[2 * i] = (r < R && kk < K) ? x0 * sc : 0.0;
    v[2 * i + 1] = (r < R && kk + 1 < K) ? x1 * sc : 0.0;
  }
}

// Element (r, c) runs of the K-step: the run starts at row r (clamped into
// [0, R)) and column c (clamped into [0, C)); it holds nv valid elements.
__device__ inline int run_len(int r, int R, int c, int C) { return r < R ? min(C - c, 8) : 0; }
__device__ inline int clamp_hi(int x, int n) { return x < n ? x : n - 1; }

__device__ inline void load_step(const GemmContrib& cb, const double* __restrict__ in_base, int m0,
                                 int n0, int M, int N, int k0, int tid, StepRegs& r) {
  gcdptr A = gload_ptr((cb.flags & kRelA) ? in_base + (intptr_t)cb.A : cb.A);
  gcdptr B = gload_ptr((cb.flags & kRelB) ? in_base + (intptr_t)cb.B : cb.B);
  const int K = cb.K;
  const double sc = cb.scale;
  const int es = (cb.flags & kStrided2) ? 2 : 1;
  if (!(cb.flags & kTransA)) {
    load_rows_k(A, cb.lda, es, m0, M, k0, K, tid, sc, r.a);
  } else {
    const int gk = k0 + (tid >> 3), mb = m0 + (tid & 7) * 8;
    load_run8(A + (int64_t)clamp_hi(gk, K) * cb.lda + (int64_t)clamp_hi(mb, M) * es, es,
              run_len(gk, K, mb, M), sc, r.a);
  }
  if (!(cb.flags & kTransB)) {
    const int gk = k0 + (tid >> 3), nb = n0 + (tid & 7) * 8;
    load_run8(B + (int64_t)clamp_hi(gk, K) * cb.ldb + (int64_t)clamp_hi(nb, N) * es, es,
              run_len(gk, K, nb, N), 1.0, r.b);
  } else {
    load_rows_k(B, cb.ldb, es, n0, N, k0, K, tid, 1.0, r.b);
  }
}

// (contribution, k0) walk of one item: advance to the next K-step
template <int KT = kKT>
__device__ inline bool next_step(const GemmContrib* __restrict__ cons, int c1, int& c, int& k0,
                                 GemmContrib& cb) {
  k0 += KT;
  if (k0 < cb.K) return true;
  k0 = 0;
  while (++c < c1) {
    cb = cons[c];
    if (cb.K > 0) return true;
  }
  return false;
}

template <typename T>
__device__ inline void store_step(const GemmContrib& cb, int tid, const StepRegs& r,
                                  T (*As)[kTile + 2], T (*Bs)[kTile + 2]) {
  if (!(cb.flags & kTransA)) {
    const int m = tid >> 4, k = 2 * (tid & 15);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      As[k][m + 16 * i] = (T)r.a[2 * i];
      As[k + 1][m + 16 * i] = (T)r.a[2 * i + 1];
    }
  } else {
    const int k = tid >> 3, mb = (tid & 7) * 8;
#pragma unroll
    for (int q = 0; q < 8; ++q) As[k][mb + q] = (T)r.a[q];
  }
  if (!(cb.flags & kTransB)) {
    const int k = tid >> 3, nb = (tid & 7) * 8;
#pragma unroll
    for (int q = 0; q < 8; ++q) Bs[k][nb + q] = (T)r.b[q];
  } else {
    const int n = tid >> 4, k = 2 * (tid & 15);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      Bs[k][n + 16 * i] = (T)r.b[2 * i];
      Bs[k + 1][n + 16 * i] = (T)r.b[2 * i + 1];
    }
  }
}

template <typename T, bool kLoop = false>
__global__ __launch_bounds__(kThreads) void gemm_mfma_kernel(
    const GemmTile* __restrict__ tiles, const GemmItem* __restrict__ items,
    const GemmContrib* __restrict__ cons, const double* __restrict__ in_base,
    double* __restrict__ out_base, int ntiles) {
  constexpr int LD = kTile + 2;  // padded LDS row
  __shared__ T As[kKT][LD];
  __shared__ T Bs[kKT][LD];
  typedef typename MfmaAcc<T>::V V;

  // one tile per workgroup, or (persistent grid) tiles blockIdx.x, +gridDim.x, ...
  for (int tix = blockIdx.x; tix < ntiles; tix += gridDim.x) {
  const GemmTile tl = tiles[tix];
  if (tl.item < 0) {  // padding of the XCD-interleaved order
    if constexpr (kLoop) continue;
    else return;
  }
  const GemmItem it = items[tl.item];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int M = it.M, N = it.N;
  const int m0 = tl.m0, n0 = tl.n0;

  V acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = V{0, 0, 0, 0};

  // first non-empty contribution
  int c = it.c0, k0 = 0;
  GemmContrib cb{};
  bool live = false;
  for (; c < it.c1; ++c) {
    cb = cons[c];
    if (cb.K > 0) { live = true; break; }
  }
  StepRegs r;
  if (live) load_step(cb, in_base, m0, n0, M, N, 0, tid, r);
  while (live) {
    store_step<T>(cb, tid, r, As, Bs);
    __syncthreads();
    GemmContrib nc = cb;
    int c2 = c, k2 = k0;
    live = next_step(cons, it.c1, c2, k2, nc);
    if (live) load_step(nc, in_base, m0, n0, M, N, k2, tid, r);  // in flight during the MFMAs
#pragma unroll
    for (int ks = 0; ks < kKT; ks += 4) {
      const int kr = ks + (lane >> 4);
      T af[2], bf[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) af[a] = As[kr][wm * 32 + a * 16 + (lane & 15)];
#pragma unroll
      for (int b = 0; b < 2; ++b) bf[b] = Bs[kr][wn * 32 + b * 16 + (lane & 15)];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = MfmaAcc<T>::mma(af[a], bf[b], acc[a][b]);
    }
    __syncthreads();
    c = c2; k0 = k2; cb = nc;
  }
  // epilogue: lane holds D[row(lane, v)][lane%16] of each 16x16 block
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int gm = m0 + wm * 32 + a * 16 + MfmaAcc<T>::row(lane, v);
        const int gn = n0 + wn * 32 + b * 16 + (lane & 15);
        if (gm < M && gn < N) {
          gdptr p = gstore_ptr((it.flags & kRelC) ? out_base + (intptr_t)it.C : it.C) +
                    (int64_t)gm * it.ldc + (int64_t)gn * ((it.flags & kStrided2C) ? 2 : 1);
          const double r2 = (double)acc[a][b][v];
          *p = it.accumulate ? (*p + r2) : r2;
        }
      }
  if constexpr (!kLoop) break;  // one tile per workgroup: no loop at all
  }
}

// ---------------------------------------------------------------------------
// BF16x3 kernel.  K-tile 32; LDS keeps hi/lo bf16 planes m-major with k
// contiguous (Ah[m][k]) so lane l fetches 8 consecutive k of row l%16 in one
// 16-byte read: exactly the v_mfma_f32_16x16x32_bf16 operand layout (lane l
// supplies A[l%16][8*(l/16) .. +7] and B[8*(l/16) .. +7][l%16]).
// ---------------------------------------------------------------------------
// BF16x3 staging: every thread holds 8 consecutive k of one row of A (m) and
// of one column of B (n), so each hi/lo plane is written to LDS with one
// 16-byte store.  Element q of thread tid:
//   A, not transposed: m = tid/16 + 16*(q/2), k = 2*(tid%16) + q%2 (pairs; as above)
//   A, transposed:     m = tid%64, k = 8*(tid/64) + q  (64 threads read 64 consecutive m)
//   B, not transposed: n = tid%64, k = 8*(tid/64) + q
//   B, transposed:     n = tid/16 + 16*(q/2), k = 2*(tid%16) + q%2
typedef float f8v __attribute__((ext_vector_type(8)));

// KT = 32 * NR: every thread loads NR runs of 8 k per operand, the runs 32 k
// apart (run j covers k0 + 32 j + [the offsets above]).
template <int NR>
struct PackedRegs {
  double a[NR][8], b[NR][8];
};

template <int NR>
__device__ inline void load_step_packed(const GemmContrib& cb, const double* __restrict__ in_base,
                                        int m0, int n0, int M, int N, int k0, int tid,
                                        PackedRegs<NR>& r) {
  gcdptr A = gload_ptr((cb.flags & kRelA) ? in_base + (intptr_t)cb.A : cb.A);
  gcdptr B = gload_ptr((cb.flags & kRelB) ? in_base + (intptr_t)cb.B : cb.B);
  const int K = cb.K;
  const double sc = cb.scale;
  const int es = (cb.flags & kStrided2) ? 2 : 1;
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    const int kj = k0 + 32 * j;
    if (!(cb.flags & kTransA)) {
      load_rows_k(A, cb.lda, es, m0, M, kj, K, tid, sc, r.a[j]);
    } else {
      const int gm = m0 + (tid & 63), kb = kj + (tid >> 6) * 8;
      load_run8(A + (int64_t)clamp_hi(kb, K) * cb.lda + (int64_t)clamp_hi(gm, M) * es, cb.lda,
                run_len(gm, M, kb, K), sc, r.a[j]);
    }
    if (!(cb.flags & kTransB)) {
      const int gn = n0 + (tid & 63), kb = kj + (tid >> 6) * 8;
      load_run8(B + (int64_t)clamp_hi(kb, K) * cb.ldb + (int64_t)clamp_hi(gn, N) * es, cb.ldb,
                run_len(gn, N, kb, K), 1.0, r.b[j]);
    } else {
      load_rows_k(B, cb.ldb, es, n0, N, kj, K, tid, 1.0, r.b[j]);
    }
  }
}

// x = hi + lo with hi = bf16(x), lo = bf16(x - hi) (round to nearest even)
__device__ inline void split8(const double* v, bf8v& h, bf8v& l) {
  f8v f;
#pragma unroll
  for (int q = 0; q < 8; ++q) f[q] = (float)v[q];
  h = __builtin_convertvector(f, bf8v);
  const f8v hf = __builtin_convertvector(h, f8v);
  l = __builtin_convertvector(f - hf, bf8v);
}

// K-step of the BF16x3 kernel: 32 or 64 (compile-time)
#ifndef DMRG_BF16X3_KT
#define DMRG_BF16X3_KT 32
#endif
constexpr int kBKT = DMRG_BF16X3_KT;
static_assert(kBKT == 32 || kBKT == 64, "BF16x3 K-step of 32 or 64");

typedef float f2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf2v __attribute__((ext_vector_type(2)));

// the same split for a pair
__device__ inline void split2(double x0, double x1, bf2v& h, bf2v& l) {
  const f2v f{(float)x0, (float)x1};
  h = __builtin_convertvector(f, bf2v);
  const f2v hf = __builtin_convertvector(h, f2v);
  l = __builtin_convertvector(f - hf, bf2v);
}

// one operand of a K-step into its hi/lo planes X[row][k] (row = m for A, n
// for B); `kcont` selects the pair layout of the k-contiguous operands
template <int LD>
__device__ inline void store_planes(bool kcont, int tid, int kofs, const double (&v)[8],
                                    __bf16 (*Xh)[LD], __bf16 (*Xl)[LD], int row_run, int k_run) {
  if (kcont) {
    const int row = tid >> 4, k = kofs + 2 * (tid & 15);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      bf2v h, l;
      split2(v[2 * i], v[2 * i + 1], h, l);
      *reinterpret_cast<bf2v*>(&Xh[row + 16 * i][k]) = h;
      *reinterpret_cast<bf2v*>(&Xl[row + 16 * i][k]) = l;
    }
  } else {
    bf8v h, l;
    split8(v, h, l);
    *reinterpret_cast<bf8v*>(&Xh[row_run][kofs + k_run]) = h;
    *reinterpret_cast<bf8v*>(&Xl[row_run][kofs + k_run]) = l;
  }
}

// acc[a][b] += lo.hi, hi.lo, hi.hi of the 2 x 2 blocks of a wave, issued as
// three passes over the four independent accumulators
__device__ inline void bf16x3_products(const bf8v (&ah)[2], const bf8v (&al)[2], const bf8v (&bh)[2],
                                       const bf8v (&bl)[2], f4v (&acc)[2][2]) {
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], bh[b], acc[a][b], 0, 0, 0);
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bl[b], acc[a][b], 0, 0, 0);
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bh[b], acc[a][b], 0, 0, 0);
}

template <bool kLoop = false>
__global__ __launch_bounds__(kThreads) void gemm_bf16x3_kernel(
    const GemmTile* __restrict__ tiles, const GemmItem* __restrict__ items,
    const GemmContrib* __restrict__ cons, const double* __restrict__ in_base,
    double* __restrict__ out_base, int ntiles) {
  constexpr int KT = kBKT, NR = KT / 32;
  constexpr int LD = KT + 8;  // rows of 80 or 144 bytes keep 16-byte alignment
  __shared__ __attribute__((aligned(16))) __bf16 Ah[kTile][LD];
  __shared__ __attribute__((aligned(16))) __bf16 Al[kTile][LD];
  __shared__ __attribute__((aligned(16))) __bf16 Bh[kTile][LD];
  __shared__ __attribute__((aligned(16))) __bf16 Bl[kTile][LD];

  for (int tix = blockIdx.x; tix < ntiles; tix += gridDim.x) {
  const GemmTile tl = tiles[tix];
  if (tl.item < 0) {  // padding of the XCD-interleaved order
    if constexpr (kLoop) continue;
    else return;
  }
  const GemmItem it = items[tl.item];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int M = it.M, N = it.N;
  const int m0 = tl.m0, n0 = tl.n0;

  f4v acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f4v{0, 0, 0, 0};

  int c = it.c0, k0 = 0;
  GemmContrib cb{};
  bool live = false;
  for (; c < it.c1; ++c) {
    cb = cons[c];
    if (cb.K > 0) { live = true; break; }
  }
  PackedRegs<NR> r;
  if (live) load_step_packed<NR>(cb, in_base, m0, n0, M, N, 0, tid, r);
  while (live) {
#pragma unroll
    for (int j = 0; j < NR; ++j) {
      // runs of the transposed A and the plain B: row tid%64, k 8*(tid/64)
      store_planes<LD>(!(cb.flags & kTransA), tid, 32 * j, r.a[j], Ah, Al, tid & 63, (tid >> 6) * 8);
      store_planes<LD>((cb.flags & kTransB) != 0, tid, 32 * j, r.b[j], Bh, Bl, tid & 63, (tid >> 6) * 8);
    }
    __syncthreads();
    GemmContrib nc = cb;
    int c2 = c, k2 = k0;
    live = next_step<KT>(cons, it.c1, c2, k2, nc);
    if (live) load_step_packed<NR>(nc, in_base, m0, n0, M, N, k2, tid, r);  // in flight during the MFMAs
#pragma unroll
    for (int j = 0; j < NR; ++j) {
      const int kc = 32 * j + 8 * (lane >> 4);
      bf8v ah[2], al[2], bh[2], bl[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const int rr = wm * 32 + a * 16 + (lane & 15);
        ah[a] = *reinterpret_cast<const bf8v*>(&Ah[rr][kc]);
        al[a] = *reinterpret_cast<const bf8v*>(&Al[rr][kc]);
      }
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const int rr = wn * 32 + b * 16 + (lane & 15);
        bh[b] = *reinterpret_cast<const bf8v*>(&Bh[rr][kc]);
        bl[b] = *reinterpret_cast<const bf8v*>(&Bl[rr][kc]);
      }
      // each product over all four blocks before the next, so no MFMA waits
      // for the one before it (the per-block order lo.hi, hi.lo, hi.hi stays)
      bf16x3_products(ah, al, bh, bl, acc);
    }
    __syncthreads();
    c = c2; k0 = k2; cb = nc;
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int gm = m0 + wm * 32 + a * 16 + 4 * (lane >> 4) + v;
        const int gn = n0 + wn * 32 + b * 16 + (lane & 15);
        if (gm < M && gn < N) {
          gdptr p = gstore_ptr((it.flags & kRelC) ? out_base + (intptr_t)it.C : it.C) +
                    (int64_t)gm * it.ldc + (int64_t)gn * ((it.flags & kStrided2C) ? 2 : 1);
          const double r2 = (double)acc[a][b][v];
          *p = it.accumulate ? (*p + r2) : r2;
        }
      }
  if constexpr (!kLoop) break;  // one tile per workgroup: no loop at all
  }
}

// ---------------------------------------------------------------------------
// Pre-split BF16x3 GEMM (GemmBatch::presplit).  The operands arrive as bf16
// pair groups, so a lane's MFMA operand (8 consecutive k of one row, hi and
// lo) is one 32-byte group: a K-step of the 64x64 tile is one 16-byte load
// pair per thread and operand, copied to LDS as it is (no FP64 -> bf16
// conversion, half the bytes of FP64 operands).  Four waves of 32x32 (2 x 2
// blocks of 16x16); three K-steps (in the same or the following terms) are
// in flight in registers while the current one is multiplied.
// Each contribution accumulates in FP32 and is added, times its scale, into
// an FP64 tile total.
// ---------------------------------------------------------------------------
typedef unsigned int u4v __attribute__((ext_vector_type(4)));
typedef const u4v __attribute__((address_space(1)))* gcu4ptr;

// one thread's share of a K-step: group 4 s + tid % 4 of row tid / 4 of the
// A tile and of the B tile (zeros past the item or the term's k)
struct PairRegs {
  u4v ah, al, bh, bl;
  bool oka, okb;  // in range; applied when the registers are stored, so that
                  // nothing waits for the loads before then
};

__device__ inline void load_pair_regs(const GemmContrib& cb, const double* __restrict__ in_base, int s,
                                      int m0, int M, int n0, int N, int tid, PairRegs& r) {
  const double* A = (cb.flags & kRelA) ? in_base + (intptr_t)cb.A : cb.A;
  const double* B = (cb.flags & kRelB) ? in_base + (intptr_t)cb.B : cb.B;
  const int G = (cb.K + 7) >> 3;
  const int g = 4 * s + (tid & 3), row = tid >> 2;
  const bool gok = g < G;
  const gcu4ptr pa = (gcu4ptr)A, pb = (gcu4ptr)B;
  const int64_t ia = ((int64_t)min(m0 + row, M - 1) * cb.lda + min(g, G - 1)) * 2;
  const int64_t ib = ((int64_t)min(n0 + row, N - 1) * cb.ldb + min(g, G - 1)) * 2;
  r.ah = pa[ia];
  r.al = pa[ia + 1];
  r.bh = pb[ib];
  r.bl = pb[ib + 1];
  r.oka = gok && m0 + row < M;
  r.okb = gok && n0 + row < N;
}

__global__ __launch_bounds__(kThreads) void gemm_bf16x3_pre_kernel(
    const GemmTile* __restrict__ tiles, const GemmItem* __restrict__ items,
    const GemmContrib* __restrict__ cons, const double* __restrict__ in_base,
    double* __restrict__ out_base) {
  constexpr int LD = 40;  // bf16 per LDS row: 32 k + 8 (16-byte aligned rows, as the split kernel)
  __shared__ GemmContrib scon[GemmBatch::kPreTerms];  // the item's terms
  __shared__ __attribute__((aligned(16))) __bf16 Ah[kTile][LD];
  __shared__ __attribute__((aligned(16))) __bf16 Al[kTile][LD];
  __shared__ __attribute__((aligned(16))) __bf16 Bh[kTile][LD];
  __shared__ __attribute__((aligned(16))) __bf16 Bl[kTile][LD];
  const GemmTile tl = tiles[blockIdx.x];
  if (tl.item < 0) return;  // padding of the XCD-interleaved order
  const GemmItem it = items[tl.item];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int M = it.M, N = it.N;
  const int m0 = tl.m0, n0 = tl.n0;
  const int am0 = m0 + wm * 32, bn0 = n0 + wn * 32;
  for (int e = tid; e < it.c1 - it.c0; e += blockDim.x) scon[e] = cons[it.c0 + e];
  __syncthreads();
  // Terms are read from LDS only (the layout caps a pre-split item at
  // GemmBatch::kPreTerms terms): a read that may be LDS or global would be a FLAT load,
  // whose wait also drains the operand loads in flight.
  auto con = [&](int c) -> GemmContrib { return scon[c - it.c0]; };
  f4v acc[2][2];
  double tot[2][2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      acc[a][b] = f4v{0, 0, 0, 0};
#pragma unroll
      for (int v = 0; v < 4; ++v) tot[a][b][v] = 0.0;
    }
  // Steps in order (term, 32-k step).  The load cursor runs three steps
  // ahead of the compute cursor: a thread's share of a step is 64 bytes, so
  // a ring of three register slots keeps three steps in flight per workgroup
  // while one LDS buffer serves the step being multiplied.  Every term has
  // K > 0 (the plan drops empty ones), and the loads are issued without
  // branches (past the last step they re-read a valid step, marked out of
  // range), so the compiler's waits stay per slot.
  int pc = it.c0, ps = 0;
  int cc = it.c0, cs = 0;
  GemmContrib ccb = con(cc);
  auto load_next = [&](PairRegs& r) -> bool {
    const bool valid = pc < it.c1;
    const GemmContrib cb = con(valid ? pc : it.c1 - 1);
    load_pair_regs(cb, in_base, valid ? ps : 0, m0, M, n0, N, tid, r);
    r.oka = r.oka && valid;
    r.okb = r.okb && valid;
    if (valid && 32 * (++ps) >= cb.K) {
      ps = 0;
      ++pc;
    }
    return valid;
  };
  auto step = [&](PairRegs& r) {
    lds_barrier();  // the previous step's LDS reads are done (the loads in flight stay so)
    {
      const int row = tid >> 2, k = 8 * (tid & 3);
      const u4v z{0, 0, 0, 0};
      *reinterpret_cast<u4v*>(&Ah[row][k]) = r.oka ? r.ah : z;
      *reinterpret_cast<u4v*>(&Al[row][k]) = r.oka ? r.al : z;
      *reinterpret_cast<u4v*>(&Bh[row][k]) = r.okb ? r.bh : z;
      *reinterpret_cast<u4v*>(&Bl[row][k]) = r.okb ? r.bl : z;
    }
    lds_barrier();
    const int kc = 8 * (lane >> 4);
    bf8v ah[2], al[2], bh[2], bl[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const int rr = wm * 32 + a * 16 + (lane & 15);
      ah[a] = *reinterpret_cast<const bf8v*>(&Ah[rr][kc]);
      al[a] = *reinterpret_cast<const bf8v*>(&Al[rr][kc]);
    }
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int rr = wn * 32 + b * 16 + (lane & 15);
      bh[b] = *reinterpret_cast<const bf8v*>(&Bh[rr][kc]);
      bl[b] = *reinterpret_cast<const bf8v*>(&Bl[rr][kc]);
    }
    bf16x3_products(ah, al, bh, bl, acc);
    if (32 * (++cs) >= ccb.K) {  // term done: scale into the FP64 total
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
#pragma unroll
          for (int v = 0; v < 4; ++v) tot[a][b][v] += ccb.scale * (double)acc[a][b][v];
          acc[a][b] = f4v{0, 0, 0, 0};
        }
      cs = 0;
      if (++cc < it.c1) ccb = con(cc);
    }
  };
  PairRegs r0, r1, r2;
  if (it.c1 <= it.c0) return;  // (not tiled by the layout)
  bool v0 = load_next(r0);
  bool v1 = load_next(r1);
  bool v2 = load_next(r2);
  while (v0) {
    step(r0);
    v0 = load_next(r0);
    if (!v1) break;
    step(r1);
    v1 = load_next(r1);
    if (!v2) break;
    step(r2);
    v2 = load_next(r2);
  }
  double* Cb = (it.flags & kRelC) ? out_base + (intptr_t)it.C : it.C;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int gm = am0 + a * 16 + 4 * (lane >> 4);  // rows gm .. gm + 3
      const int gn = bn0 + b * 16 + (lane & 15);
      if (gn >= N) continue;
      if (it.flags & kPairC) {
        // row gn of the transposed pair matrix, group gm / 8, slots gm % 8 ..
        __bf16 hv[4], lv[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const float x = (float)tot[a][b][v];
          hv[v] = (__bf16)x;
          lv[v] = (__bf16)(x - (float)hv[v]);
        }
        __bf16* g = reinterpret_cast<__bf16*>(Cb + ((int64_t)gn * it.ldc + (gm >> 3)) * 4) + (gm & 7);
        if (gm + 3 < M) {
          typedef __bf16 bf4v __attribute__((ext_vector_type(4)));
          *reinterpret_cast<bf4v*>(g) = bf4v{hv[0], hv[1], hv[2], hv[3]};
          *reinterpret_cast<bf4v*>(g + 8) = bf4v{lv[0], lv[1], lv[2], lv[3]};
        } else {
#pragma unroll
          for (int v = 0; v < 4; ++v)
            if (gm + v < M) {
              g[v] = hv[v];
              g[v + 8] = lv[v];
            }
        }
      } else {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          if (gm + v >= M) continue;
          gdptr p = gstore_ptr(Cb) + (int64_t)(gm + v) * it.ldc + gn;
          *p = it.accumulate ? (*p + tot[a][b][v]) : tot[a][b][v];
        }
      }
    }
}

// Pair-group packing (PackTask), one thread per (row, group)
__global__ __launch_bounds__(256) void pack_pairs_kernel(const PackTask* __restrict__ tasks,
                                                         const double* __restrict__ base) {
  const PackTask t = tasks[blockIdx.x];
  const int G = (t.K + 7) >> 3;
  const int64_t n = (int64_t)t.rows * G;
  const double* src = base ? base + (intptr_t)t.src : t.src;
  for (int64_t e = (int64_t)blockIdx.y * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.y * blockDim.x) {
    const int r = (int)(e / G), g = (int)(e % G);
    double v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int k = 8 * g + q;
      v[q] = k < t.K ? src[(int64_t)r * t.srs + (int64_t)k * t.sks] : 0.0;
    }
    bf8v h, l;
    split8(v, h, l);
    double* d = t.dst + 4 * e;
    *reinterpret_cast<bf8v*>(d) = h;
    *reinterpret_cast<bf8v*>(d + 2) = l;
  }
}

// ---------------------------------------------------------------------------
// Wave-tile GEMM kernels: every wave of the 256-thread workgroup owns one
// 32x32 output tile (tiles[4*blockIdx.x + wave]) with its own K walk and its
// own LDS slice, so the four tiles of a workgroup may belong to different
// items and run different K lengths.  No block-level barrier: a wave stores
// its staged operands, fences its own LDS traffic and reads them back.  For
// block-sparse operands with 50-600 rows this wastes less of every tile than
// 64x64 tiles and gives a launch four times as many independent tiles.
// One K-step stages 32 x 32 of A and 32 x 32 of B, 16 elements of each per
// lane (mapping at wave_load).
// ---------------------------------------------------------------------------
constexpr int kWT = 32;  // wave tile edge

struct WaveRegs {
  double a[16], b[16];
};

// The wave-tile staging of a K-step, element q of lane l:
//   k contiguous (A not transposed, B transposed): row l/16 + 4*(q/2),
//     k = 2*(l%16) + q%2 (four 256-byte row segments per load instruction);
//   k-major (A transposed, B not transposed): row-run index l%32,
//     k = 16*(l/32) + q (32 lanes read 32 consecutive elements).
__device__ inline void wave_rows_k(gcdptr X, int64_t ld, int es, int r0, int R, int k0, int K, int lane,
                                   double sc, double (&v)[16]) {
  const int kk = k0 + 2 * (lane & 15);
  const int ka = kk < K ? kk : K - 1, kb = kk + 1 < K ? kk + 1 : K - 1;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int r = r0 + (lane >> 4) + 4 * i;
    double x0 = 0.0, x1 = 0.0;
    if (r < R) {
      const gcdptr row = X + (int64_t)r * ld;
      x0 = row[(int64_t)ka * es];
      x1 = row[(int64_t)kb * es];
    }
    v[2 * i] = (r < R && kk < K) ? x0 * sc : 0.0;
    v[2 * i + 1] = (r < R && kk + 1 < K) ? x1 * sc : 0.0;
  }
}

__device__ inline void wave_k_major(gcdptr X, int64_t ld, int es, int c0, int C, int k0, int K, int lane,
                                    double sc, double (&v)[16]) {
  const int gc = c0 + (lane & 31);
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int kb = k0 + (lane >> 5) * 16 + 8 * j;
    double t[8];
    load_run8(X + (int64_t)clamp_hi(kb, K) * ld + (int64_t)clamp_hi(gc, C) * es, ld, run_len(gc, C, kb, K),
              sc, t);
#pragma unroll
    for (int q = 0; q < 8; ++q) v[8 * j + q] = t[q];
  }
}

__device__ inline void wave_load(const GemmContrib& cb, const double* __restrict__ in_base, int m0,
                                 int n0, int M, int N, int k0, int lane, WaveRegs& r) {
  gcdptr A = gload_ptr((cb.flags & kRelA) ? in_base + (intptr_t)cb.A : cb.A);
  gcdptr B = gload_ptr((cb.flags & kRelB) ? in_base + (intptr_t)cb.B : cb.B);
  const int K = cb.K;
  const int es = (cb.flags & kStrided2) ? 2 : 1;
  if (!(cb.flags & kTransA)) wave_rows_k(A, cb.lda, es, m0, M, k0, K, lane, cb.scale, r.a);
  else wave_k_major(A, cb.lda, es, m0, M, k0, K, lane, cb.scale, r.a);
  if (!(cb.flags & kTransB)) wave_k_major(B, cb.ldb, es, n0, N, k0, K, lane, 1.0, r.b);
  else wave_rows_k(B, cb.ldb, es, n0, N, k0, K, lane, 1.0, r.b);
}

// one operand of a wave K-step into a k-major tile X[k][row] (FP64 / FP32)
template <typename T, int LD>
__device__ inline void wave_store_kmajor(bool kcont, int lane, const double (&v)[16], T (*X)[LD]) {
  if (kcont) {
    const int row = lane >> 4, k = 2 * (lane & 15);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      X[k][row + 4 * i] = (T)v[2 * i];
      X[k + 1][row + 4 * i] = (T)v[2 * i + 1];
    }
  } else {
    const int row = lane & 31, kb = (lane >> 5) * 16;
#pragma unroll
    for (int q = 0; q < 16; ++q) X[kb + q][row] = (T)v[q];
  }
}

// one operand of a wave K-step into hi/lo planes X[row][k] (BF16x3)
template <int LD>
__device__ inline void wave_store_planes(bool kcont, int lane, const double (&v)[16], __bf16 (*Xh)[LD],
                                         __bf16 (*Xl)[LD]) {
  if (kcont) {
    const int row = lane >> 4, k = 2 * (lane & 15);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      bf2v h, l;
      split2(v[2 * i], v[2 * i + 1], h, l);
      *reinterpret_cast<bf2v*>(&Xh[row + 4 * i][k]) = h;
      *reinterpret_cast<bf2v*>(&Xl[row + 4 * i][k]) = l;
    }
  } else {
    const int row = lane & 31, kb = (lane >> 5) * 16;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      bf8v h, l;
      split8(v + 8 * j, h, l);
      *reinterpret_cast<bf8v*>(&Xh[row][kb + 8 * j]) = h;
      *reinterpret_cast<bf8v*>(&Xl[row][kb + 8 * j]) = l;
    }
  }
}

// LDS traffic of this wave ordered before what follows (other lanes' stores
// visible to this lane's loads, and loads done before the next stores)
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T>
__device__ inline void wave_store_out(const GemmItem& it, double* __restrict__ out_base, int m0, int n0,
                                      int lane, const typename MfmaAcc<T>::V (&acc)[2][2]) {
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int gm = m0 + a * 16 + MfmaAcc<T>::row(lane, v);
        const int gn = n0 + b * 16 + (lane & 15);
        if (gm < it.M && gn < it.N) {
          gdptr p = gstore_ptr((it.flags & kRelC) ? out_base + (intptr_t)it.C : it.C) +
                    (int64_t)gm * it.ldc + (int64_t)gn * ((it.flags & kStrided2C) ? 2 : 1);
          const double r2 = (double)acc[a][b][v];
          *p = it.accumulate ? (*p + r2) : r2;
        }
      }
}

template <typename T, bool kLoop = false>
__global__ __launch_bounds__(kThreads) void gemm_wave_kernel(
    const GemmTile* __restrict__ tiles, const GemmItem* __restrict__ items,
    const GemmContrib* __restrict__ cons, const double* __restrict__ in_base,
    double* __restrict__ out_base, int ntiles) {
  constexpr int LD = kWT + 2;
  __shared__ T As[4][kKT][LD];  // k-major, per wave
  __shared__ T Bs[4][kKT][LD];
  typedef typename MfmaAcc<T>::V V;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  // ntiles counts groups of four tiles (one per wave)
  for (int gix = blockIdx.x; gix < ntiles; gix += gridDim.x) {
  const GemmTile tl = tiles[gix * 4 + wave];
  if (tl.item < 0) {  // no block barriers: a wave without a tile skips it
    if constexpr (kLoop) continue;
    else return;
  }
  const GemmItem it = items[tl.item];
  const int M = it.M, N = it.N, m0 = tl.m0, n0 = tl.n0;
  T (*as)[LD] = As[wave];
  T (*bs)[LD] = Bs[wave];

  V acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = V{0, 0, 0, 0};
  int c = it.c0, k0 = 0;
  GemmContrib cb{};
  bool live = false;
  for (; c < it.c1; ++c) {
    cb = cons[c];
    if (cb.K > 0) { live = true; break; }
  }
  WaveRegs r;
  if (live) wave_load(cb, in_base, m0, n0, M, N, 0, lane, r);
  while (live) {
    wave_store_kmajor<T, LD>(!(cb.flags & kTransA), lane, r.a, as);
    wave_store_kmajor<T, LD>((cb.flags & kTransB) != 0, lane, r.b, bs);
    wave_lds_sync();
    GemmContrib nc = cb;
    int c2 = c, k2 = k0;
    live = next_step(cons, it.c1, c2, k2, nc);
    if (live) wave_load(nc, in_base, m0, n0, M, N, k2, lane, r);  // in flight during the MFMAs
#pragma unroll
    for (int ks = 0; ks < kKT; ks += 4) {
      const int kr = ks + (lane >> 4);
      T af[2], bf[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) af[a] = as[kr][a * 16 + (lane & 15)];
#pragma unroll
      for (int b = 0; b < 2; ++b) bf[b] = bs[kr][b * 16 + (lane & 15)];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = MfmaAcc<T>::mma(af[a], bf[b], acc[a][b]);
    }
    wave_lds_sync();
    c = c2; k0 = k2; cb = nc;
  }
  wave_store_out<T>(it, out_base, m0, n0, lane, acc);
  if constexpr (!kLoop) break;
  }
}

// BF16x3 wave-tile kernel: per-wave hi/lo planes m-major with k contiguous,
// as in gemm_bf16x3_kernel.
template <bool kLoop = false>
__global__ __launch_bounds__(kThreads) void gemm_bf16x3_wave_kernel(
    const GemmTile* __restrict__ tiles, const GemmItem* __restrict__ items,
    const GemmContrib* __restrict__ cons, const double* __restrict__ in_base,
    double* __restrict__ out_base, int ntiles) {
  constexpr int LD = kKT + 8;
  __shared__ __attribute__((aligned(16))) __bf16 Ah[4][kWT][LD];
  __shared__ __attribute__((aligned(16))) __bf16 Al[4][kWT][LD];
  __shared__ __attribute__((aligned(16))) __bf16 Bh[4][kWT][LD];
  __shared__ __attribute__((aligned(16))) __bf16 Bl[4][kWT][LD];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  for (int gix = blockIdx.x; gix < ntiles; gix += gridDim.x) {
  const GemmTile tl = tiles[gix * 4 + wave];
  if (tl.item < 0) {
    if constexpr (kLoop) continue;
    else return;
  }
  const GemmItem it = items[tl.item];
  const int M = it.M, N = it.N, m0 = tl.m0, n0 = tl.n0;
  __bf16 (*ah_)[LD] = Ah[wave];
  __bf16 (*al_)[LD] = Al[wave];
  __bf16 (*bh_)[LD] = Bh[wave];
  __bf16 (*bl_)[LD] = Bl[wave];

  f4v acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f4v{0, 0, 0, 0};
  int c = it.c0, k0 = 0;
  GemmContrib cb{};
  bool live = false;
  for (; c < it.c1; ++c) {
    cb = cons[c];
    if (cb.K > 0) { live = true; break; }
  }
  WaveRegs r;
  if (live) wave_load(cb, in_base, m0, n0, M, N, 0, lane, r);
  while (live) {
    wave_store_planes<LD>(!(cb.flags & kTransA), lane, r.a, ah_, al_);
    wave_store_planes<LD>((cb.flags & kTransB) != 0, lane, r.b, bh_, bl_);
    wave_lds_sync();
    GemmContrib nc = cb;
    int c2 = c, k2 = k0;
    live = next_step(cons, it.c1, c2, k2, nc);
    if (live) wave_load(nc, in_base, m0, n0, M, N, k2, lane, r);  // in flight during the MFMAs
    const int kc = 8 * (lane >> 4);
    bf8v ah[2], al[2], bh[2], bl[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const int rr = a * 16 + (lane & 15);
      ah[a] = *reinterpret_cast<const bf8v*>(&ah_[rr][kc]);
      al[a] = *reinterpret_cast<const bf8v*>(&al_[rr][kc]);
    }
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int rr = b * 16 + (lane & 15);
      bh[b] = *reinterpret_cast<const bf8v*>(&bh_[rr][kc]);
      bl[b] = *reinterpret_cast<const bf8v*>(&bl_[rr][kc]);
    }
    bf16x3_products(ah, al, bh, bl, acc);
    wave_lds_sync();
    c = c2; k0 = k2; cb = nc;
  }
  wave_store_out<float>(it, out_base, m0, n0, lane, acc);
  if constexpr (!kLoop) break;
  }
}

// split-K reduction: C (+)= sum of the partial buffers, in chunk order
// (s = ((0 + p_1) + p_2) + ..., then C += s).  Thread t of block x owns
// elements 4 (256 x + t) .. + 3 of task y's M x N output, read and written
// row-major; the loads of four partials for all four elements are issued
// together, so a thread waits for memory once per four partials (the rows-
// by-workgroup form waited once per element row).
constexpr int kSplitEl = 4;
__global__ __launch_bounds__(256) void split_reduce_kernel(const SplitTask* __restrict__ tasks,
                                                           double* __restrict__ out_base) {
  const SplitTask t = tasks[blockIdx.y];
  const int64_t mn = (int64_t)t.M * t.N;
  const int64_t e0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * kSplitEl;
  if (e0 >= mn) return;
  gdptr C = gstore_ptr((t.flags & kRelC) ? out_base + (intptr_t)t.C : t.C);
  gcdptr part = gload_ptr(t.part);
  const int es = (t.flags & kStrided2C) ? 2 : 1;
  double sum[kSplitEl];
#pragma unroll
  for (int q = 0; q < kSplitEl; ++q) sum[q] = 0.0;
  for (int c0 = 0; c0 < t.nparts; c0 += 4) {
    double v[4][kSplitEl];
#pragma unroll
    for (int cc = 0; cc < 4; ++cc)
#pragma unroll
      for (int q = 0; q < kSplitEl; ++q)  // clamped addresses, always in bounds
        v[cc][q] = part[(int64_t)min(c0 + cc, t.nparts - 1) * mn + min(e0 + q, mn - 1)];
#pragma unroll
    for (int cc = 0; cc < 4; ++cc)
      if (c0 + cc < t.nparts)
#pragma unroll
        for (int q = 0; q < kSplitEl; ++q) sum[q] += v[cc][q];
  }
#pragma unroll
  for (int q = 0; q < kSplitEl; ++q) {
    const int64_t e = e0 + q;
    if (e < mn) C[(e / t.N) * t.ldc + (e % t.N) * es] += sum[q];
  }
}

// ---------------------------------------------------------------------------
// strided copies
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void copy_kernel(const CopyTask* __restrict__ tasks) {
  const CopyTask t = tasks[blockIdx.x];
  const int total = t.rows * t.cols;  // one row chunk (launch_copies)
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    const int i = e / t.cols, j = e - i * t.cols;
    const double v = t.src ? t.alpha * t.src[i * t.ssr + j * t.ssc] : 0.0;
    t.dst[i * t.dsr + j * t.dsc] = v;
  }
}

// ---------------------------------------------------------------------------
// vector primitives
// ---------------------------------------------------------------------------
constexpr int kDotBlocks = 512;

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// DPP lane move of a double within rows of 16 lanes (kCtrl: a DPP control
// code: 0xB1 / 0x4E quad permutations (lane ^ 1, lane ^ 2), 0x141 half-row
// mirror, 0x140 row mirror)
template <int kCtrl>
__device__ inline double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, kCtrl, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), kCtrl, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Sum over each row of 16 lanes, the same value in all 16 (a + b and b + a
// round alike), on DPP moves: no LDS round trips
__device__ inline double row16_sum(double v) {
  v += dpp_move<0xB1>(v);
  v += dpp_move<0x4E>(v);
  v += dpp_move<0x141>(v);
  v += dpp_move<0x140>(v);
  return v;
}

__device__ inline double readlane_d(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// v of lane src (ds_bpermute, both halves)
__device__ inline double bperm_d(double v, int src) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_ds_bpermute(src << 2, (int)b);
  const int hi = __builtin_amdgcn_ds_bpermute(src << 2, (int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Sum over the wave, wave-uniform: row sums, then rows (0 + 1) + (2 + 3)
__device__ inline double wave_sum_dpp(double v) {
  v = row16_sum(v);
  return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
}

__global__ __launch_bounds__(256) void dot_partial_kernel(const double* __restrict__ x,
                                                          const double* __restrict__ y,
                                                          int64_t n, double* __restrict__ part) {
  __shared__ double red[4];
  double s = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    s += x[i] * y[i];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ __launch_bounds__(256) void dot_final_kernel(const double* __restrict__ part, int np,
                                                        double* __restrict__ out) {
  __shared__ double red[4];
  double s = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) s += part[i];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) *out = (red[0] + red[1]) + (red[2] + red[3]);
}

// <x|y> = sum conj(x_i) y_i over n interleaved complex entries: partial
// (re, im) per block, then a fixed-order final sum (deterministic)
__global__ __launch_bounds__(256) void zdot_partial_kernel(const double* __restrict__ x,
                                                           const double* __restrict__ y,
                                                           int64_t n, double* __restrict__ part) {
  __shared__ double red[2][4];
  double sr = 0.0, si = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double xr = x[2 * i], xi = x[2 * i + 1], yr = y[2 * i], yi = y[2 * i + 1];
    sr += xr * yr + xi * yi;
    si += xr * yi - xi * yr;
  }
  sr = wave_sum(sr);
  si = wave_sum(si);
  if ((threadIdx.x & 63) == 0) { red[0][threadIdx.x >> 6] = sr; red[1][threadIdx.x >> 6] = si; }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
    part[2 * blockIdx.x + 1] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
  }
}

__global__ __launch_bounds__(256) void zdot_final_kernel(const double* __restrict__ part, int np,
                                                         double* __restrict__ out) {
  __shared__ double red[2][4];
  double sr = 0.0, si = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) { sr += part[2 * i]; si += part[2 * i + 1]; }
  sr = wave_sum(sr);
  si = wave_sum(si);
  if ((threadIdx.x & 63) == 0) { red[0][threadIdx.x >> 6] = sr; red[1][threadIdx.x >> 6] = si; }
  __syncthreads();
  if (threadIdx.x == 0) {
    out[0] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
    out[1] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
  }
}

// y += sign * alpha * x, alpha = (alpha[0], alpha[1]) complex on the device
__global__ void zaxpy_dev_kernel(double* __restrict__ y, const double* __restrict__ x, int64_t n,
                                 const double* __restrict__ alpha, double sign) {
  const double ar = sign * alpha[0], ai = sign * alpha[1];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double xr = x[2 * i], xi = x[2 * i + 1];
    y[2 * i] += ar * xr - ai * xi;
    y[2 * i + 1] += ar * xi + ai * xr;
  }
}

__global__ void axpy_dev_kernel(double* __restrict__ y, const double* __restrict__ x, int64_t n,
                                const double* __restrict__ alpha, double sign) {
  const double a = sign * (*alpha);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] += a * x[i];
}

__global__ void axpby_kernel(double* __restrict__ y, const double* __restrict__ x, int64_t n,
                             double a, double b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = a * y[i] + b * x[i];
}

__global__ void scale_inv_sqrt_kernel(double* __restrict__ x, int64_t n,
                                      const double* __restrict__ nrm2, int mode) {
  if (mode == 0) {
    const double r = sqrt(*nrm2);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
      x[i] /= r;
  } else {
    const double r = sqrt(1.0 / (*nrm2));
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
      x[i] *= r;
  }
}

__global__ void lanczos_step_kernel(double* __restrict__ nxt, double* __restrict__ h,
                                    const double* __restrict__ psi, int64_t n, double beta) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    nxt[i] = h[i] / beta;
    h[i] = -beta * psi[i];
  }
}

__global__ __launch_bounds__(256) void col_gather_kernel(const ColGatherTask* __restrict__ tasks) {
  // workgroup (task, y) takes elements y*256 + tid + k*256*gridDim.y
  const ColGatherTask t = tasks[blockIdx.x];
  const int total = t.rows * t.ncol;  // < 2^31 (checked on the host)
  const bool cz = t.flags & kGatherComplex;
  const double cj = (t.flags & kGatherConj) ? -1.0 : 1.0;
  for (int e = blockIdx.y * blockDim.x + threadIdx.x; e < total; e += gridDim.y * blockDim.x) {
    const int jj = e / t.rows, i = e - jj * t.rows;
    const double s = t.scale[jj];
    const int64_t di = (int64_t)i * t.dsr + (int64_t)jj * t.dsc;
    if (!cz) {
      double v;
      if (s != s) v = (t.r0 + i == jj) ? 1.0 : 0.0;
      else v = s * t.src[(int64_t)(t.r0 + i) + (int64_t)t.perm[jj] * t.ld];
      t.dst[di] = v;
    } else {
      double vr, vi;
      if (s != s) {
        vr = (t.r0 + i == jj) ? 1.0 : 0.0;
        vi = 0.0;
      } else {
        const int64_t si = (int64_t)(t.r0 + i) + (int64_t)t.perm[jj] * t.ld;
        vr = s * t.src[2 * si];
        vi = s * t.src[2 * si + 1];
      }
      t.dst[2 * di] = vr;
      t.dst[2 * di + 1] = cj * vi;
    }
  }
}

__global__ void combine_kernel(double* __restrict__ y, const double* __restrict__ V, int64_t stride,
                               const double* __restrict__ coef, int k, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int j = 0; j < k; ++j) s += V[j * stride + i] * coef[j];
    y[i] = s;
  }
}

__global__ void fill_kernel(double* __restrict__ x, int64_t n, double v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] = v;
}

// ---------------------------------------------------------------------------
// Fused Lanczos vector steps.  Partial sums follow dot_partial_kernel's
// decomposition (grid g = dot_grid(n) blocks of 256 threads, element i to
// thread i mod (256 g), each thread summing its elements in increasing i, a
// wave then a fixed four-wave combine); a coefficient is the sum of the
// previous step's partials in dot_final_kernel's order, formed by every
// block that needs it.  The element updates equal axpy_dev / zaxpy_dev /
// lanczos_step, so a fused sequence gives the same doubles as the unfused
// one (real vectors; complex norms sum |y_i|^2 over entries).
// ---------------------------------------------------------------------------
template <bool kC>
__device__ inline void lz_block_store(double sr, double si, double* __restrict__ part) {
  __shared__ double red[2][4];
  sr = wave_sum(sr);
  if (kC) si = wave_sum(si);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { red[0][w] = sr; red[1][w] = si; }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (kC) {
      part[2 * blockIdx.x] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
      part[2 * blockIdx.x + 1] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
    } else {
      part[blockIdx.x] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
    }
  }
}

// sum of np partials (complex: pairs) in dot_final_kernel's order, returned
// to every thread of the block
template <bool kC>
__device__ inline void lz_final(const double* __restrict__ part, int np, double& cr, double& ci) {
  __shared__ double red[2][4];
  __shared__ double res[2];
  double sr = 0.0, si = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) {
    if (kC) { sr += part[2 * i]; si += part[2 * i + 1]; }
    else sr += part[i];
  }
  sr = wave_sum(sr);
  if (kC) si = wave_sum(si);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { red[0][w] = sr; red[1][w] = si; }
  __syncthreads();
  if (threadIdx.x == 0) {
    res[0] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
    res[1] = kC ? (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]) : 0.0;
  }
  __syncthreads();
  cr = res[0];
  ci = res[1];
}

template <bool kC>
__global__ __launch_bounds__(256) void lz_dot_partial_kernel(const double* __restrict__ x,
                                                             const double* __restrict__ y, int64_t n,
                                                             double* __restrict__ part) {
  double sr = 0.0, si = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (kC) {
      const double xr = x[2 * i], xi = x[2 * i + 1], yr = y[2 * i], yi = y[2 * i + 1];
      sr += xr * yr + xi * yi;
      si += xr * yi - xi * yr;
    } else {
      sr += x[i] * y[i];
    }
  }
  lz_block_store<kC>(sr, si, part);
}

template <bool kC>
__global__ __launch_bounds__(256) void lz_finals_kernel(const double* __restrict__ part_a,
                                                        const double* __restrict__ part_b, int np,
                                                        double* __restrict__ out, double* __restrict__ out_host) {
  double ar, ai, br, bi;
  lz_final<kC>(part_a, np, ar, ai);
  lz_final<false>(part_b, np, br, bi);
  if (threadIdx.x == 0) {
    out[0] = ar;
    out[1] = ai;
    out[2] = br;
    if (out_host) {  // pinned host memory: the values, then the ready flag
      out_host[0] = ar;
      out_host[1] = ai;
      out_host[2] = br;
      __threadfence_system();
      out_host[3] = 1.0;
      __threadfence_system();
    }
  }
}

// ---------------------------------------------------------------------------
// The rest of a Lanczos iteration after the matvec and the alpha dot, in two
// launches.  The reference (real_tensor_lanczos.cc:80-151, GramSchmidtMethod
// :188-201) subtracts alpha psi_i, forms beta = |h|, y = h / beta, projects y
// against psi_0 .. psi_i one after the other and normalises.  Here
// lz_alpha_multi subtracts alpha psi_i and forms the partials of |h|^2 and of
// every <psi_j|h>; lz_next_all forms beta, c_j = <psi_j|h> / beta and writes
// (h / beta - sum_j c_j psi_j) / r with r^2 = 1 - sum_j |c_j|^2 (|h / beta|
// = 1 by the definition of beta).  For an orthonormal basis this is the
// reference's vector: its sequential coefficients differ from c_j by terms of
// second order in the overlaps, which the three-term recurrence leaves at the
// rounding level, and |y - sum c_j psi_j|^2 = |y|^2 - sum |c_j|^2.  Partial
// sums: grid g = dot_grid(n), element e to thread e mod (256 g), a fixed block
// combine; every block that needs a coefficient sums the partials in one
// fixed order.
// ---------------------------------------------------------------------------
constexpr int kMultiChunk = 4;  // overlaps accumulated per pass over the elements

// block sums of NV per-thread values; value v < nvalid goes to part[(base +
// v) g + blockIdx]
template <int NV>
__device__ inline void lz_block_store_multi(const double (&val)[NV], int nvalid, int base,
                                            double* __restrict__ part, int g) {
  __shared__ double red[2 * kMultiChunk][4];
  static_assert(NV <= 2 * kMultiChunk, "values per block store");
  const int w = threadIdx.x >> 6;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const double t = wave_sum(val[v]);
    if ((threadIdx.x & 63) == 0) red[v][w] = t;
  }
  __syncthreads();
  if ((int)threadIdx.x < nvalid) {
    const int v = threadIdx.x;
    part[(int64_t)(base + v) * g + blockIdx.x] = (red[v][0] + red[v][1]) + (red[v][2] + red[v][3]);
  }
  __syncthreads();
}

// alpha = final(part_a); h -= alpha psi; part_b <- partials of |h|^2;
// part_m <- partials of <x_j|h> for j < nj (x_j = x + j ld; complex: value
// pairs, groups 2 j and 2 j + 1)
template <bool kC>
__global__ __launch_bounds__(256) void lz_alpha_multi_kernel(double* __restrict__ h,
                                                             const double* __restrict__ psi,
                                                             const double* __restrict__ x, int64_t ld, int nj,
                                                             int64_t n, const double* __restrict__ part_a,
                                                             double* __restrict__ part_b,
                                                             double* __restrict__ part_m) {
  const int g = gridDim.x;
  double cr, ci;
  lz_final<kC>(part_a, g, cr, ci);
  const double ar = -1.0 * cr, ai = -1.0 * ci;
  const int64_t stride = (int64_t)g * blockDim.x;
  const int64_t e0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  double nrm = 0.0;
  for (int64_t e = e0; e < n; e += stride) {
    if (kC) {
      const double xr = psi[2 * e], xi = psi[2 * e + 1];
      const double yr = h[2 * e] + (ar * xr - ai * xi);
      const double yi = h[2 * e + 1] + (ar * xi + ai * xr);
      h[2 * e] = yr;
      h[2 * e + 1] = yi;
      nrm += yr * yr + yi * yi;
    } else {
      const double y = h[e] + ar * psi[e];
      h[e] = y;
      nrm += y * y;
    }
  }
  {
    const double v[1] = {nrm};
    lz_block_store_multi(v, 1, 0, part_b, g);
  }
  // overlaps, kMultiChunk vectors per pass (this thread's own elements of h)
  for (int j0 = 0; j0 < nj; j0 += kMultiChunk) {
    double sr[kMultiChunk], si[kMultiChunk];
#pragma unroll
    for (int u = 0; u < kMultiChunk; ++u) sr[u] = si[u] = 0.0;
    for (int64_t e = e0; e < n; e += stride) {
      if (kC) {
        const double yr = h[2 * e], yi = h[2 * e + 1];
#pragma unroll
        for (int u = 0; u < kMultiChunk; ++u)
          if (j0 + u < nj) {
            const double* xj = x + (int64_t)(j0 + u) * ld;
            const double xr = xj[2 * e], xi = xj[2 * e + 1];
            sr[u] += xr * yr + xi * yi;
            si[u] += xr * yi - xi * yr;
          }
      } else {
        const double y = h[e];
#pragma unroll
        for (int u = 0; u < kMultiChunk; ++u)
          if (j0 + u < nj) sr[u] += x[(int64_t)(j0 + u) * ld + e] * y;
      }
    }
    const int nu = min(kMultiChunk, nj - j0);
    if (kC) {
      double v[2 * kMultiChunk];
#pragma unroll
      for (int u = 0; u < kMultiChunk; ++u) {
        v[2 * u] = sr[u];
        v[2 * u + 1] = si[u];
      }
      lz_block_store_multi(v, 2 * nu, 2 * j0, part_m, g);
    } else {
      lz_block_store_multi(sr, nu, j0, part_m, g);
    }
  }
}

// beta: formed from part_b (block 0 then also forms alpha from part_a and
// publishes (alpha, beta^2) to slot and slot_host, then the ready flag
// slot_host[3]), or sqrt of the device value beta2.  c_j = final(part_m group)
// / beta; nxt = (h / beta - sum_j c_j x_j) / r, r^2 = 1 - sum_j |c_j|^2 (j
// ascending; a non-positive r^2, a breakdown the solve's beta test stops at,
// divides by 1); h = -beta psi.
template <bool kC>
__global__ __launch_bounds__(256) void lz_next_all_kernel(
    double* __restrict__ nxt, double* __restrict__ h, const double* __restrict__ psi,
    const double* __restrict__ x, int64_t ld, int nj, int64_t n, const double* __restrict__ part_m,
    const double* __restrict__ part_a, const double* __restrict__ part_b, const double* __restrict__ beta2,
    double* __restrict__ slot, double* __restrict__ slot_host) {
  extern __shared__ double coef[];  // (kC ? 2 : 1) nj values
  const int g = gridDim.x;
  double beta;
  if (part_b) {
    double br, bi;
    lz_final<false>(part_b, g, br, bi);
    beta = sqrt(br);
    if (slot && blockIdx.x == 0) {
      double ar, ai;
      lz_final<kC>(part_a, g, ar, ai);
      if (threadIdx.x == 0) {
        slot[0] = ar;
        slot[1] = ai;
        slot[2] = br;
        slot_host[0] = ar;
        slot_host[1] = ai;
        slot_host[2] = br;
        // the ready flag after the values, both pushed out to host memory
        __threadfence_system();
        slot_host[3] = 1.0;
        __threadfence_system();
      }
    }
  } else {
    beta = sqrt(beta2[0]);
  }
  const int nvals = (kC ? 2 : 1) * nj;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int v = w; v < nvals; v += 4) {
    double s = 0.0;
    for (int i = lane; i < g; i += 64) s += part_m[(int64_t)v * g + i];
    s = wave_sum(s);
    if (lane == 0) coef[v] = s / beta;
  }
  __syncthreads();
  double r2 = 1.0;
  for (int v = 0; v < nvals; ++v) r2 -= coef[v] * coef[v];
  const double r = r2 > 0.0 ? sqrt(r2) : 1.0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    if (kC) {
      double vr = h[2 * e] / beta, vi = h[2 * e + 1] / beta;
      for (int j = 0; j < nj; ++j) {
        const double* xj = x + (int64_t)j * ld;
        const double xr = xj[2 * e], xi = xj[2 * e + 1], cr = coef[2 * j], ci = coef[2 * j + 1];
        vr -= cr * xr - ci * xi;
        vi -= cr * xi + ci * xr;
      }
      nxt[2 * e] = vr / r;
      nxt[2 * e + 1] = vi / r;
      h[2 * e] = -beta * psi[2 * e];
      h[2 * e + 1] = -beta * psi[2 * e + 1];
    } else {
      double v = h[e] / beta;
      for (int j = 0; j < nj; ++j) v -= coef[j] * x[(int64_t)j * ld + e];
      nxt[e] = v / r;
      h[e] = -beta * psi[e];
    }
  }
}

inline int grid_for(int64_t n) {
  int64_t g = (n + 255) / 256;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, 2048));
}

// ---------------------------------------------------------------------------
// One-sided (Hestenes) Jacobi SVD, batched over sectors.  Columns are paired
// by the round-robin (circle) ordering; one launch performs one round for
// every active sector, one wave per column pair, so a round costs one short
// kernel however many sectors there are.  The host runs the rounds of a
// sweep and retires sectors whose sweep made no rotation.
// ---------------------------------------------------------------------------
// Columns per workgroup of the per-column SVD epilogues (grid y splits a
// problem's columns; every column is computed as by a single workgroup).
constexpr int kFinCols = 32;

__global__ __launch_bounds__(256) void jacobi_init_kernel(const SvdProblem* __restrict__ probs,
                                                          double* __restrict__ part) {
  // workgroup (problem, y): columns [c0, c1) of V set to the identity's, and
  // the partial |G|^2 of the same columns of G into part[problem * gy + y]
  const SvdProblem p = probs[blockIdx.x];
  const int c0 = blockIdx.y * kFinCols, c1 = min(p.N, c0 + kFinCols);
  const int64_t nv = (int64_t)max(c1 - c0, 0) * p.N, ng = (int64_t)max(c1 - c0, 0) * p.M;
  for (int64_t e = threadIdx.x; e < nv; e += blockDim.x) {
    const int64_t idx = (int64_t)c0 * p.N + e;
    p.V[idx] = (idx % (p.N + 1) == 0) ? 1.0 : 0.0;
  }
  __shared__ double red[4];
  double s = 0.0;
  const double* g = p.G + (int64_t)c0 * p.M;
  for (int64_t e = threadIdx.x; e < ng; e += blockDim.x) s += g[e] * g[e];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0)
    part[(int64_t)blockIdx.x * gridDim.y + blockIdx.y] = (red[0] + red[1]) + (red[2] + red[3]);
}

// |A|_F^2 of every problem from its column-chunk partials, in chunk order
__global__ __launch_bounds__(64) void svd_fro2_kernel(const SvdProblem* __restrict__ probs,
                                                      const double* __restrict__ part, int gy) {
  double s = 0.0;
  for (int i = threadIdx.x; i < gy; i += 64) s += part[(int64_t)blockIdx.x * gy + i];
  s = wave_sum(s);
  if (threadIdx.x == 0) *probs[blockIdx.x].fro2 = s;
}

// Block one-sided Jacobi: the N columns are cut into blocks of kJb; one
// launch performs one round of the block round-robin, one workgroup of kJw
// waves per block pair ("visit", up to kJc = 2 kJb columns).  A visit:
//  1. forms the kJc x kJc Gram matrix of its columns with FP64 MFMA: the waves
//     take interleaved 16-row slabs and keep partial 16 x 16 blocks, summed in
//     a fixed order in LDS (deterministic);
//  2. runs one sweep of two-sided Jacobi on it in registers (wave 0,
//     inner_sweep16; the rotation angles are those of the one-sided method,
//     computed from the Gram entries), accumulating the rotation Q;
//  3. applies Q to its columns of G, and the previous round's rotation to its
//     columns of V, with MFMA (slabs over the waves).
// Pairs whose columns are already orthogonal to tol (or numerically zero) are
// not rotated; a visit that rotates nothing leaves its columns untouched.
// The round time is one visit's critical path, so the visit width and the
// number of waves per visit are compile-time tunables.  DMRG_JACOBI_JB=16
// gives 32-column visits (three local rounds of two 16-column sweeps each,
// every column pair of the two blocks rotated once per visit): half the
// rounds per sweep at the same sweep count, but on the 8x4 m = 1000 sweep a
// round took 41 us against 16 us (half the visits, twice the bytes each), so
// 16-column visits stay the default.
#ifndef DMRG_JACOBI_JB
#define DMRG_JACOBI_JB 8
#endif
#ifndef DMRG_JACOBI_WAVES
#define DMRG_JACOBI_WAVES 8
#endif
constexpr int kJb = DMRG_JACOBI_JB;
constexpr int kJc = 2 * kJb;
constexpr int kJw = DMRG_JACOBI_WAVES;
constexpr int kJn = kJc / 16;  // 16 x 16 MFMA blocks per visit edge
static_assert(kJc == 16 || kJc == 32, "visits of 16 or 32 columns");
static_assert(kJw >= 2 && kJw <= 16, "2 to 16 waves per visit");
static_assert(kJc == 16 || kJw >= 8, "32-column visits update their 32 x 32 Gram with eight waves");

// Slabs of 16 rows handled together by one wave (their loads in flight at once)
#ifndef DMRG_JACOBI_SLABS
#define DMRG_JACOBI_SLABS 2
#endif
constexpr int kSlabs = DMRG_JACOBI_SLABS;

// rows [0, rows) of the visit's columns: X[:, col] <- X[:, col] * Q, by
// waves wave, wave + nwaves, ... of the workgroup (16-row slabs)
__device__ inline void apply_rotation(double* __restrict__ X, int64_t ld, int rows,
                                      const int* col, const double (*Q)[kJc + 1], int lane,
                                      int wave, int nwaves) {
  const int i = lane & 15, kk = lane >> 4;
  constexpr int kT = kJc / 4;  // MFMA k-steps
  int cs[kT];
#pragma unroll
  for (int t = 0; t < kT; ++t) cs[t] = col[4 * t + kk];
  int cout[kJn];
#pragma unroll
  for (int h = 0; h < kJn; ++h) cout[h] = col[16 * h + i];
  double qb[kJn][kT];
#pragma unroll
  for (int h = 0; h < kJn; ++h)
#pragma unroll
    for (int t = 0; t < kT; ++t) qb[h][t] = Q[4 * t + kk][16 * h + i];
  for (int base = wave * 16; base < rows; base += 16 * nwaves * kSlabs) {
    double a[kSlabs][kT];
#pragma unroll
    for (int u = 0; u < kSlabs; ++u)
#pragma unroll
      for (int t = 0; t < kT; ++t) {
        const int r = base + 16 * nwaves * u + i;
        a[u][t] = (cs[t] >= 0 && r < rows) ? X[(int64_t)cs[t] * ld + r] : 0.0;
      }
#pragma unroll
    for (int u = 0; u < kSlabs; ++u) {
      d4v acc[kJn];
#pragma unroll
      for (int h = 0; h < kJn; ++h) {
        acc[h] = d4v{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int t = 0; t < kT; ++t)
          acc[h] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u][t], qb[h][t], acc[h], 0, 0, 0);
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = base + 16 * nwaves * u + kk + 4 * v;
        if (r < rows) {
#pragma unroll
          for (int h = 0; h < kJn; ++h)
            if (cout[h] >= 0) X[(int64_t)cout[h] * ld + r] = acc[h][v];
        }
      }
    }
  }
}

// One sweep of two-sided Jacobi in registers on the 16 x 16 Gram
// G[sel[r]][sel[c]] (r, c < 16) of an LDS array with row length LD, by one
// wave; its rotation goes to Q[sel[r]][sel[c]].  Returns whether any pair
// rotated (the same value in every lane).  Pairs (a, b) rotate when
// G_ab^2 > tol2 G_aa G_bb and both norms exceed zero2.
template <int LD>
__device__ inline int inner_sweep16(const double (*G)[LD], double (*Q)[LD], const int* sel, int lane,
                                    double tol2, double zero2) {
  // The inner sweep in registers, XOR order: step d rotates the pairs
  // (x, x ^ d), d = 1 .. 15 (every pair once).  Lane (c, q) = c + 16 q
  // holds G[c ^ (4 q + s)][c] and Q[c ^ (4 q + s)][c] in slot s < 4 and the
  // diagonal D = G[c][c]; every lane forms its column's rotation from the
  // same three values as its partner (bit-identical angles).  Per step:
  // column update G J, Q J (the partner column sits at lane ^ (d | (d >> 2)
  // << 4), slot s ^ (d & 3)), then the row update through the symmetry of
  // J^T G J: G'[r][c] = G'[c][r] = (G J)[c][r] cs + (G J)[c ^ d][r] s_c.
  const int c = lane & 15, q = lane >> 4;
  double g[4], qv[4];
#pragma unroll
  for (int s2 = 0; s2 < 4; ++s2) {
    g[s2] = G[sel[c ^ (4 * q + s2)]][sel[c]];
    qv[s2] = (4 * q + s2) == 0 ? 1.0 : 0.0;
  }
  double D = G[sel[c]][sel[c]];
  int any = 0;
#pragma unroll
  for (int d = 1; d < 16; ++d) {
    const int hb = d >= 8 ? 8 : d >= 4 ? 4 : d >= 2 ? 2 : 1;  // top bit of d
    const bool lo = (c & hb) == 0;
    const int a = lo ? c : (c ^ d);
    const double Dp = bperm_d(D, lane ^ d);
    // G[a ^ d][a] sits in column a at offset d: quarter d >> 2, slot d & 3
    const double gm = bperm_d(g[d & 3], a + 16 * (d >> 2));
    const double al = lo ? D : Dp, be = lo ? Dp : D;
    double cs = 1.0, sn = 0.0;
    const bool doit = al > zero2 && be > zero2 && gm * gm > tol2 * al * be;
    if (doit) {
      // tan(theta) = sgn(dd) gg / (|dd| + hypot(dd, gg)), dd = be - al,
      // gg = 2 gm (the one-sided angle); rsq/rcp with one Newton step each
      const double dd = be - al, gg = 2.0 * gm;
      const double w = dd * dd + gg * gg;
      double ri = __builtin_amdgcn_rsq(w);
      ri = ri * (1.5 - 0.5 * w * ri * ri);
      const double den = fabs(dd) + w * ri;
      double rc = __builtin_amdgcn_rcp(den);
      rc = rc * (2.0 - den * rc);
      const double t = (dd >= 0.0 ? gg : -gg) * rc;
      const double u = 1.0 + t * t;
      double ci = __builtin_amdgcn_rsq(u);
      ci = ci * (1.5 - 0.5 * u * ci * ci);
      cs = ci;
      sn = ci * t;
      any = 1;
    }
    // a step that rotates no pair leaves G and Q as they are: skip its
    // exchanges (wave-uniform test)
    if (!__any(doit)) continue;
    // new column a = cs a - sn b, new column b = sn a + cs b
    const double sg = lo ? -sn : sn;
    const int mask = d | ((d >> 2) << 4);
    double h[4], pq[4];
#pragma unroll
    for (int s2 = 0; s2 < 4; ++s2) {
      const double pg = bperm_d(g[s2 ^ (d & 3)], lane ^ mask);
      pq[s2] = bperm_d(qv[s2 ^ (d & 3)], lane ^ mask);  // all read before any update
      h[s2] = cs * g[s2] + sg * pg;
    }
#pragma unroll
    for (int s2 = 0; s2 < 4; ++s2) qv[s2] = cs * qv[s2] + sg * pq[s2];
#pragma unroll
    for (int s2 = 0; s2 < 4; ++s2) {
      const int r = c ^ (4 * q + s2);
      const double hc = bperm_d(h[s2], r + 16 * q);                      // (G J)[c][r]
      const double hd = bperm_d(h[s2 ^ (d & 3)], r + 16 * (q ^ (d >> 2)));  // (G J)[c ^ d][r]
      g[s2] = cs * hc + sg * hd;
    }
    D = lo ? cs * cs * al - 2.0 * cs * sn * gm + sn * sn * be
           : sn * sn * al + 2.0 * cs * sn * gm + cs * cs * be;
  }
#pragma unroll
  for (int s2 = 0; s2 < 4; ++s2) Q[sel[c ^ (4 * q + s2)]][sel[c]] = qv[s2];
  return __any(any);
}

// column c of visit slot `slot` (< kJc) of block pair pr in round `round`
// of the block round-robin over nb2 (even) blocks; -1 past the last column
__device__ inline int visit_column(int round, int pr, int nb, int nb2, int N, int slot) {
  int bi, bj;
  if (pr == 0) { bi = round; bj = nb2 - 1; }
  else { bi = (round + pr) % (nb2 - 1); bj = (round + nb2 - 1 - pr) % (nb2 - 1); }
  const int blk = slot < kJb ? bi : bj;
  const int c = blk * kJb + (slot & (kJb - 1));
  return (blk < nb && c < N) ? c : -1;
}

// Sweep state of the round launches, per problem: rotated (set by the visits
// that rotate), active (problems still sweeping; the visits of the others
// exit at once), and host, the device address of pinned host memory that
// receives rotated at the end of each sweep.
struct SweepCtl {
  int* rotated;
  int* active;
  int* host;
  int np;
};

// The V half of a round launch (workgroups nvisits .. 2 nvisits - 1): visit
// slot v applies the previous round's rotation of its (problem, block pair)
// to V with the whole workgroup, on CUs of its own, while the G half of the
// same launch runs the current round (V is read by nothing else during the
// sweep, and every V column still receives its rotations in round order).
// The flush launch of a sweep has one more workgroup, which ends the sweep:
// rotated goes to the host and becomes the next sweep's active, and is
// cleared (the visits that wrote it ran in earlier launches).
template <int kC, int kW, typename QT>
__device__ inline bool round_v_part(const SvdProblem* __restrict__ probs, const int* __restrict__ vprob,
                                    const int* __restrict__ vpair, int round, const double* __restrict__ qbuf,
                                    const int* __restrict__ qflag, int nvisits, int kb, int cs, SweepCtl ctl,
                                    QT apply) {
  if ((int)blockIdx.x < nvisits) return false;
  if ((int)blockIdx.x >= 2 * nvisits) {
    for (int i = threadIdx.x; i < ctl.np; i += blockDim.x) {
      const int r = ctl.rotated[i];
      ctl.host[i] = r;
      ctl.active[i] = r;
      ctl.rotated[i] = 0;
    }
    return true;
  }
  const int v = (int)blockIdx.x - nvisits;
  if (round < 1 || !qflag[((round - 1) & 1) * nvisits + v]) return true;
  __shared__ double Qp[2][kC][kC + 1];  // the previous round's rotation (complex: re, im)
  __shared__ int colp[kC];
  const int pi = vprob[v], pr = vpair[v];
  const SvdProblem p = probs[pi];
  const int N = p.N;
  const int nb = (N + kb - 1) / kb, nb2 = nb + (nb & 1);
  if (round - 1 >= nb2 - 1) return true;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const double* qs = qbuf + ((int64_t)((round - 1) & 1) * nvisits + v) * (cs * kC * kC);
  for (int e = tid; e < cs * kC * kC; e += 64 * kW) {
    const int h = e / (kC * kC), r = (e / kC) % kC, c = e % kC;
    Qp[h][r][c] = qs[e];
  }
  if (tid < kC) {
    int bi, bj;
    if (pr == 0) { bi = round - 1; bj = nb2 - 1; }
    else { bi = (round - 1 + pr) % (nb2 - 1); bj = (round - 1 + nb2 - 1 - pr) % (nb2 - 1); }
    const int blk = tid < kb ? bi : bj;
    const int c = blk * kb + (tid % kb);
    colp[tid] = (blk < nb && c < N) ? c : -1;
  }
  __syncthreads();
  apply(p.V, N, colp, Qp, lane, wave);
  return true;
}

// One round.  Launched with 2 nvisits workgroups: the first nvisits run the
// round's visits (Gram, inner sweep, rotation of G), the others apply the
// previous round's rotations to V (round_v_part).  The rotation and a flag
// go through qbuf / qflag, double-buffered by round parity; a launch with
// round == the sweep's round count only flushes the last rotations and ends
// the sweep.
__global__ __launch_bounds__(64 * kJw) void jacobi_gram_round_kernel(
    const SvdProblem* __restrict__ probs, const int* __restrict__ vprob,
    const int* __restrict__ vpair, int round, SweepCtl ctl,
    double* __restrict__ qbuf, int* __restrict__ qflag, int nvisits) {
  int* __restrict__ rotated = ctl.rotated;
  if (round_v_part<kJc, kJw>(probs, vprob, vpair, round, qbuf, qflag, nvisits, kJb, 1, ctl,
                             [](double* V, int N, const int* colp, const double (*Qp)[kJc][kJc + 1], int lane,
                                int wave) { apply_rotation(V, N, N, colp, Qp[0], lane, wave, kJw); }))
    return;
  __shared__ double Gp[kJw][kJc][kJc + 1];  // partial Grams, one per wave
  __shared__ double Q[kJc][kJc + 1];
  __shared__ int col[kJc];
  __shared__ int flag;
  double (*G)[kJc + 1] = Gp[0];  // the summed Gram overwrites wave 0's partial
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // visit blockIdx.x is column-block pair pr of problem pi (host-built map)
  const int pi = vprob[blockIdx.x], pr = vpair[blockIdx.x];
  const SvdProblem p = probs[pi];
  const int M = p.M, N = p.N;
  const int nb = (N + kJb - 1) / kJb, nb2 = nb + (nb & 1);
  if (round >= nb2 - 1) return;
  if (!ctl.active[pi]) {  // converged in an earlier sweep: nothing for V either
    if (tid == 0) qflag[(round & 1) * nvisits + blockIdx.x] = 0;
    return;
  }
  if (tid < kJc) col[tid] = visit_column(round, pr, nb, nb2, N, tid);
  __syncthreads();
  // partial Gram of this wave's slabs: lane (i, kk) feeds X[row][col_{16h+i}]
  {
    const int i = lane & 15, kk = lane >> 4;
    const double* __restrict__ xc[kJn];
    bool ok[kJn];
#pragma unroll
    for (int h = 0; h < kJn; ++h) {
      const int c = col[16 * h + i];
      ok[h] = c >= 0;
      xc[h] = p.G + (int64_t)(c >= 0 ? c : 0) * M;
    }
    d4v acc[kJn][kJn];
#pragma unroll
    for (int a = 0; a < kJn; ++a)
#pragma unroll
      for (int b = 0; b < kJn; ++b) acc[a][b] = d4v{0.0, 0.0, 0.0, 0.0};
    for (int base = wave * 16; base < M; base += 16 * kJw * kSlabs) {
      double v[kSlabs][4][kJn];
#pragma unroll
      for (int u = 0; u < kSlabs; ++u)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int r = base + 16 * kJw * u + 4 * kk + t;
#pragma unroll
          for (int h = 0; h < kJn; ++h) v[u][t][h] = (ok[h] && r < M) ? xc[h][r] : 0.0;
        }
#pragma unroll
      for (int u = 0; u < kSlabs; ++u)
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
          for (int a = 0; a < kJn; ++a)
#pragma unroll
            for (int b = 0; b < kJn; ++b)
              acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(v[u][t][a], v[u][t][b], acc[a][b], 0, 0, 0);
    }
#pragma unroll
    for (int a = 0; a < kJn; ++a)
#pragma unroll
      for (int b = 0; b < kJn; ++b)
#pragma unroll
        for (int v = 0; v < 4; ++v) Gp[wave][16 * a + kk + 4 * v][16 * b + i] = acc[a][b][v];
  }
  __syncthreads();
  for (int e = tid; e < kJc * kJc; e += blockDim.x) {
    const int r = e / kJc, c = e % kJc;
    double sum = Gp[0][r][c];
    for (int w = 1; w < kJw; ++w) sum += Gp[w][r][c];
    G[r][c] = sum;
    Q[r][c] = r == c ? 1.0 : 0.0;
  }
  if (kJc == 32 && tid == 0) flag = 0;  // (32-column visits: set by the sweeping waves)
  __syncthreads();
  if constexpr (kJc == 16) {
  if (wave == 0) {
    const double tol = 2.220446049250313e-16 * (double)M;  // above the dot-product round-off
    const double zero2 = 4.930380657631324e-32 * (*p.fro2);  // (eps |A|_F)^2
    const int ident[16] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15};
    const int any = inner_sweep16<kJc + 1>(G, Q, ident, lane, tol * tol, zero2);
    if (lane == 0) flag = any;
  }
  } else {
  // 32-column visits: three local rounds over the four 8-column sub-blocks
  // s0 = slots 0-7, s1 = 8-15 (block bi), s2 = 16-23, s3 = 24-31 (block bj):
  // (s0 s1)(s2 s3), (s0 s2)(s1 s3), (s0 s3)(s1 s2).  In each, waves 0 and 1
  // sweep one 16-column pair each (inner_sweep16 on its part of the 32 x 32
  // Gram), their rotations form Qr, and the workgroup updates G <- Qr^T G Qr
  // and Q <- Q Qr on MFMA.  Every column pair of the two blocks is rotated once per visit, so a sweep needs
  // half the rounds of 16-column visits.
  const double tol = 2.220446049250313e-16 * (double)M;  // above the dot-product round-off
  const double zero2 = 4.930380657631324e-32 * (*p.fro2);  // (eps |A|_F)^2
  double (*Qr)[kJc + 1] = Gp[1];  // (partial Grams no longer needed)
  double (*Tm)[kJc + 1] = Gp[2];
  double (*Qn)[kJc + 1] = Gp[3];
  int any = 0;
  for (int lr = 0; lr < 3; ++lr) {
    if (wave < 2) {
      // sub-block pairs of this local round: (0 1)(2 3), (0 2)(1 3), (0 3)(1 2)
      const int sa = wave == 0 ? 0 : (lr == 0 ? 2 : 1);
      const int sb = wave == 0 ? lr + 1 : (lr == 0 ? 3 : (lr == 1 ? 3 : 2));
      int sel[16];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        sel[k] = 8 * sa + k;
        sel[8 + k] = 8 * sb + k;
      }
      any |= inner_sweep16<kJc + 1>(G, Qr, sel, lane, tol * tol, zero2);
      // zero this pair's rows in the other pair's columns (the two pairs
      // partition the slots, so Qr is written whole)
      const int oa = wave == 1 ? 0 : (lr == 0 ? 2 : 1);
      const int ob = wave == 1 ? lr + 1 : (lr == 0 ? 3 : (lr == 1 ? 3 : 2));
      for (int e = lane; e < 16 * 16; e += 64) {
        const int r = e >> 4, c = e & 15;
        Qr[sel[r]][c < 8 ? 8 * oa + c : 8 * ob + c - 8] = 0.0;
      }
    }
    __syncthreads();
    // Tm = G Qr (waves 0-3), Qn = Q Qr (waves 4-7): one 16 x 16 block each
    {
      const int i = lane & 15, kk = lane >> 4;
      const int w4 = wave & 3, bi = w4 >> 1, bj = w4 & 1;
      const double (*L)[kJc + 1] = wave < 4 ? G : Q;
      double (*O)[kJc + 1] = wave < 4 ? Tm : Qn;
      if (wave < 8) {
        d4v acc = d4v{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int t = 0; t < kJc / 4; ++t)
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(L[16 * bi + i][4 * t + kk], Qr[4 * t + kk][16 * bj + i],
                                                     acc, 0, 0, 0);
#pragma unroll
        for (int v = 0; v < 4; ++v) O[16 * bi + kk + 4 * v][16 * bj + i] = acc[v];
      }
    }
    __syncthreads();
    // G = Qr^T Tm (waves 0-3); Q = Qn (waves 4-7)
    {
      const int i = lane & 15, kk = lane >> 4;
      const int w4 = wave & 3, bi = w4 >> 1, bj = w4 & 1;
      if (wave < 4) {
        d4v acc = d4v{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int t = 0; t < kJc / 4; ++t)
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Qr[4 * t + kk][16 * bi + i], Tm[4 * t + kk][16 * bj + i],
                                                     acc, 0, 0, 0);
#pragma unroll
        for (int v = 0; v < 4; ++v) G[16 * bi + kk + 4 * v][16 * bj + i] = acc[v];
      } else if (wave < 8) {
#pragma unroll
        for (int v = 0; v < 4; ++v) Q[16 * bi + kk + 4 * v][16 * bj + i] = Qn[16 * bi + kk + 4 * v][16 * bj + i];
      }
    }
    __syncthreads();
  }
  if (wave < 2 && lane == 0 && any) flag = 1;
  }
  __syncthreads();
  int* qf = qflag + (round & 1) * nvisits + blockIdx.x;
  if (!flag) {
    if (tid == 0) *qf = 0;
    return;
  }
  apply_rotation(p.G, M, M, col, Q, lane, wave, kJw);
  double* qd = qbuf + ((int64_t)(round & 1) * nvisits + blockIdx.x) * (kJc * kJc);
  for (int e = tid; e < kJc * kJc; e += blockDim.x) qd[e] = Q[e / kJc][e % kJc];
  if (tid == 0) {
    *qf = 1;
    rotated[pi] = 1;
  }
}

// singular values and the sign convention, one workgroup per problem
__global__ __launch_bounds__(256) void jacobi_finish_kernel(const SvdProblem* __restrict__ probs) {
  const SvdProblem p = probs[blockIdx.x];
  if (p.precond) return;  // finished by qr_finish_kernel
  const int M = p.M, N = p.N;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  const int jend = min(N, (int)(blockIdx.y + 1) * kFinCols);
  for (int j = blockIdx.y * kFinCols + wave; j < jend; j += nwaves) {
    double* g = p.G + (int64_t)j * M;
    double* v = p.V + (int64_t)j * N;
    double s = 0.0;
    for (int i = lane; i < M; i += 64) s += g[i] * g[i];
    s = wave_sum(s);
    if (lane == 0) p.sigma[j] = sqrt(s);
    // sign convention: largest |u_i| (first on ties) positive
    const double* u = p.u_in_v ? v : g;
    const int n = p.u_in_v ? N : M;
    double best = -1.0, val = 0.0;
    int bi = 0x7fffffff;
    for (int i = lane; i < n; i += 64) {
      const double x = fabs(u[i]);
      if (x > best) { best = x; bi = i; val = u[i]; }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double ob = __shfl_xor(best, o, 64);
      const int oi = __shfl_xor(bi, o, 64);
      const double ov = __shfl_xor(val, o, 64);
      if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; val = ov; }
    }
    if (val < 0.0) {
      for (int i = lane; i < M; i += 64) g[i] = -g[i];
      for (int i = lane; i < N; i += 64) v[i] = -v[i];
    }
  }
}

// ---------------------------------------------------------------------------
// Complex one-sided block Jacobi (interleaved complex G, V).  Visits of 16
// columns (blocks of 8), 8 waves: Hermitian Gram from real MFMA products
// (Re = Xr^T Xr + Xi^T Xi, Im = Xr^T Xi - Xi^T Xr), one sweep of two-sided
// complex Jacobi on it in LDS with the unitary
//   J = [[c, s e^{i phi}], [-s e^{-i phi}, c]],  e^{i phi} = G_ab / |G_ab|
// (the real rotation on (a, e^{-i phi} b), phase restored), and X <- X Q with
// four real MFMA chains per slab.
// ---------------------------------------------------------------------------
constexpr int kZw = 8;
constexpr int kZc = 16;

__global__ __launch_bounds__(256) void zjacobi_init_kernel(const SvdProblem* __restrict__ probs,
                                                           double* __restrict__ part) {
  // as jacobi_init_kernel, on interleaved complex G and V
  const SvdProblem p = probs[blockIdx.x];
  const int c0 = blockIdx.y * kFinCols, c1 = min(p.N, c0 + kFinCols);
  const int64_t nv = (int64_t)max(c1 - c0, 0) * p.N, ng = 2 * (int64_t)max(c1 - c0, 0) * p.M;
  for (int64_t e = threadIdx.x; e < nv; e += blockDim.x) {
    const int64_t idx = (int64_t)c0 * p.N + e;
    p.V[2 * idx] = (idx % (p.N + 1) == 0) ? 1.0 : 0.0;
    p.V[2 * idx + 1] = 0.0;
  }
  __shared__ double red[4];
  double s = 0.0;
  const double* g = p.G + 2 * (int64_t)c0 * p.M;
  for (int64_t e = threadIdx.x; e < ng; e += blockDim.x) s += g[e] * g[e];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0)
    part[(int64_t)blockIdx.x * gridDim.y + blockIdx.y] = (red[0] + red[1]) + (red[2] + red[3]);
}

// rows [0, rows) of the visit's complex columns: X[:, col] <- X[:, col] * Q
__device__ inline void zapply_rotation(double* __restrict__ X, int64_t ld, int rows, const int* col,
                                       const double (*Qr)[kZc + 1], const double (*Qi)[kZc + 1],
                                       int lane, int wave, int nwaves) {
  const int i = lane & 15, kk = lane >> 4;
  int cs[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) cs[t] = col[4 * t + kk];
  const int cout = col[i];
  double qr[4], qi[4], nqi[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    qr[t] = Qr[4 * t + kk][i];
    qi[t] = Qi[4 * t + kk][i];
    nqi[t] = -qi[t];
  }
  for (int base = wave * 16; base < rows; base += 16 * nwaves) {
    double ar[4], ai[4];
    const int r = base + i;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const bool ok = cs[t] >= 0 && r < rows;
      const int64_t ix = 2 * ((int64_t)(cs[t] >= 0 ? cs[t] : 0) * ld + r);
      ar[t] = ok ? X[ix] : 0.0;
      ai[t] = ok ? X[ix + 1] : 0.0;
    }
    d4v re = {0.0, 0.0, 0.0, 0.0}, im = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      re = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[t], qr[t], re, 0, 0, 0);
      re = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[t], nqi[t], re, 0, 0, 0);
      im = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[t], qi[t], im, 0, 0, 0);
      im = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[t], qr[t], im, 0, 0, 0);
    }
    if (cout >= 0) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int rr = base + kk + 4 * v;
        if (rr < rows) {
          const int64_t ix = 2 * ((int64_t)cout * ld + rr);
          X[ix] = re[v];
          X[ix + 1] = im[v];
        }
      }
    }
  }
}

// One round, as jacobi_gram_round_kernel: launched with 2 nvisits
// workgroups, the first nvisits running the round's visits and the others
// applying the previous round's rotations to V (round_v_part).  Rotations and
// flags go through qbuf (Qr then Qi, 2 kZc^2 doubles per slot) and qflag,
// double-buffered by round parity; a launch with round == the sweep's round
// count only flushes the last rotations and ends the sweep.
__global__ __launch_bounds__(64 * kZw) void zjacobi_round_kernel(
    const SvdProblem* __restrict__ probs, const int* __restrict__ vprob,
    const int* __restrict__ vpair, int round, SweepCtl ctl,
    double* __restrict__ qbuf, int* __restrict__ qflag, int nvisits) {
  constexpr int kB = kZc / 2;
  int* __restrict__ rotated = ctl.rotated;
  if (round_v_part<kZc, kZw>(probs, vprob, vpair, round, qbuf, qflag, nvisits, kB, 2, ctl,
                             [](double* V, int N, const int* colp, const double (*Qp)[kZc][kZc + 1], int lane,
                                int wave) { zapply_rotation(V, N, N, colp, Qp[0], Qp[1], lane, wave, kZw); }))
    return;
  __shared__ double Pr[kZw][kZc][kZc + 1];
  __shared__ double Pi[kZw][kZc][kZc + 1];
  __shared__ double Qr[kZc][kZc + 1];
  __shared__ double Qi[kZc][kZc + 1];
  __shared__ int col[kZc];
  __shared__ int flag;
  double (*Gr)[kZc + 1] = Pr[0];
  double (*Gi)[kZc + 1] = Pi[0];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // visit blockIdx.x is column-block pair pr of problem pi (host-built map)
  const int pi = vprob[blockIdx.x], pr = vpair[blockIdx.x];
  const SvdProblem p = probs[pi];
  const int M = p.M, N = p.N;
  const int nb = (N + kB - 1) / kB, nb2 = nb + (nb & 1);
  if (round >= nb2 - 1) return;
  if (!ctl.active[pi]) {  // converged in an earlier sweep: nothing for V either
    if (tid == 0) qflag[(round & 1) * nvisits + blockIdx.x] = 0;
    return;
  }
  if (tid < kZc) {
    int bi, bj;
    if (pr == 0) { bi = round; bj = nb2 - 1; }
    else { bi = (round + pr) % (nb2 - 1); bj = (round + nb2 - 1 - pr) % (nb2 - 1); }
    const int blk = tid < kB ? bi : bj;
    const int c = blk * kB + (tid & (kB - 1));
    col[tid] = (blk < nb && c < N) ? c : -1;
  }
  __syncthreads();
  {
    const int i = lane & 15, kk = lane >> 4;
    const int ci = col[i];
    const double* __restrict__ xc = p.G + 2 * (int64_t)(ci >= 0 ? ci : 0) * M;
    d4v rr = {0.0, 0.0, 0.0, 0.0}, ii = rr, ri = rr, ir = rr;
    for (int base = wave * 16; base < M; base += 16 * kZw) {
      double vr[4], vi[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int r = base + 4 * kk + t;
        const bool ok = ci >= 0 && r < M;
        vr[t] = ok ? xc[2 * r] : 0.0;
        vi[t] = ok ? xc[2 * r + 1] : 0.0;
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        rr = __builtin_amdgcn_mfma_f64_16x16x4f64(vr[t], vr[t], rr, 0, 0, 0);
        ii = __builtin_amdgcn_mfma_f64_16x16x4f64(vi[t], vi[t], ii, 0, 0, 0);
        ri = __builtin_amdgcn_mfma_f64_16x16x4f64(vr[t], vi[t], ri, 0, 0, 0);
        ir = __builtin_amdgcn_mfma_f64_16x16x4f64(vi[t], vr[t], ir, 0, 0, 0);
      }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      Pr[wave][kk + 4 * v][i] = rr[v] + ii[v];
      Pi[wave][kk + 4 * v][i] = ri[v] - ir[v];
    }
  }
  __syncthreads();
  {
    const int r = tid >> 4, c = tid & 15;
    if (tid < kZc * kZc) {
      double sr = Pr[0][r][c], si = Pi[0][r][c];
      for (int w = 1; w < kZw; ++w) { sr += Pr[w][r][c]; si += Pi[w][r][c]; }
      Gr[r][c] = sr;
      Gi[r][c] = si;
      Qr[r][c] = r == c ? 1.0 : 0.0;
      Qi[r][c] = 0.0;
    }
  }
  __syncthreads();
  if (wave == 0) {
    const double tol = 2.220446049250313e-16 * (double)M;
    const double tol2 = tol * tol;
    const double zero2 = 4.930380657631324e-32 * (*p.fro2);
    const int pp = lane >> 3, i0 = (lane & 7) * 2;
    int any = 0;
    for (int r = 0; r < kZc - 1; ++r) {
      int a, b;
      if (pp == 0) { a = r; b = kZc - 1; }
      else { a = (r + pp) % (kZc - 1); b = (r + kZc - 1 - pp) % (kZc - 1); }
      if (a > b) { const int t = a; a = b; b = t; }
      const double al = Gr[a][a], be = Gr[b][b], gr = Gr[a][b], gi = Gi[a][b];
      const double ag2 = gr * gr + gi * gi;
      double c = 1.0, s = 0.0, ph_r = 1.0, ph_i = 0.0;
      const bool doit = al > zero2 && be > zero2 && ag2 > tol2 * al * be;
      if (doit) {
        double ri = __builtin_amdgcn_rsq(ag2);
        ri = ri * (1.5 - 0.5 * ag2 * ri * ri);
        const double ag = ag2 * ri;
        ph_r = gr * ri;
        ph_i = gi * ri;
        const double d = be - al, g = 2.0 * ag;
        const double w = d * d + g * g;
        double rw = __builtin_amdgcn_rsq(w);
        rw = rw * (1.5 - 0.5 * w * rw * rw);
        const double den = fabs(d) + w * rw;
        double rc = __builtin_amdgcn_rcp(den);
        rc = rc * (2.0 - den * rc);
        const double t = (d >= 0.0 ? g : -g) * rc;
        const double u = 1.0 + t * t;
        double ci = __builtin_amdgcn_rsq(u);
        ci = ci * (1.5 - 0.5 * u * ci * ci);
        c = ci;
        s = ci * t;
        any = 1;
      }
      // a step that rotates no pair leaves G and Q as they are (wave-uniform)
      if (!__any(doit)) continue;
      // J = [[c, s e], [-s conj(e)]...]: x_a' = c x_a - s conj(e) x_b,
      // x_b' = s e x_a + c x_b, with e = (ph_r, ph_i)
      const double mr = s * ph_r, mi = s * ph_i;  // s e
      wave_lds_sync();
      if (doit) {
        for (int ii = i0; ii < i0 + 2; ++ii) {
          double ar = Gr[ii][a], ai = Gi[ii][a], br = Gr[ii][b], bi = Gi[ii][b];
          Gr[ii][a] = c * ar - (mr * br + mi * bi);
          Gi[ii][a] = c * ai - (mr * bi - mi * br);
          Gr[ii][b] = (mr * ar - mi * ai) + c * br;
          Gi[ii][b] = (mr * ai + mi * ar) + c * bi;
          ar = Qr[ii][a]; ai = Qi[ii][a]; br = Qr[ii][b]; bi = Qi[ii][b];
          Qr[ii][a] = c * ar - (mr * br + mi * bi);
          Qi[ii][a] = c * ai - (mr * bi - mi * br);
          Qr[ii][b] = (mr * ar - mi * ai) + c * br;
          Qi[ii][b] = (mr * ai + mi * ar) + c * bi;
        }
      }
      wave_lds_sync();
      if (doit) {  // G <- J^H G: row a' = c a - s e b, row b' = s conj(e) a + c b
        for (int ii = i0; ii < i0 + 2; ++ii) {
          const double ar = Gr[a][ii], ai = Gi[a][ii], br = Gr[b][ii], bi = Gi[b][ii];
          Gr[a][ii] = c * ar - (mr * br - mi * bi);
          Gi[a][ii] = c * ai - (mr * bi + mi * br);
          Gr[b][ii] = (mr * ar + mi * ai) + c * br;
          Gi[b][ii] = (mr * ai - mi * ar) + c * bi;
        }
      }
      wave_lds_sync();
    }
    any = __any(any);
    if (lane == 0) flag = any;
  }
  __syncthreads();
  int* qf = qflag + (round & 1) * nvisits + blockIdx.x;
  if (!flag) {
    if (tid == 0) *qf = 0;
    return;
  }
  zapply_rotation(p.G, M, M, col, Qr, Qi, lane, wave, kZw);
  double* qd = qbuf + ((int64_t)(round & 1) * nvisits + blockIdx.x) * (2 * kZc * kZc);
  {
    const int e = tid & (kZc * kZc - 1);
    qd[tid] = tid < kZc * kZc ? Qr[e / kZc][e % kZc] : Qi[e / kZc][e % kZc];
  }
  if (tid == 0) {
    *qf = 1;
    rotated[pi] = 1;
  }
}

// sigma and the phase convention, one workgroup per problem, wave per column
__global__ __launch_bounds__(256) void zjacobi_finish_kernel(const SvdProblem* __restrict__ probs) {
  const SvdProblem p = probs[blockIdx.x];
  const int M = p.M, N = p.N;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  const int jend = min(N, (int)(blockIdx.y + 1) * kFinCols);
  for (int j = blockIdx.y * kFinCols + wave; j < jend; j += nwaves) {
    double* g = p.G + 2 * (int64_t)j * M;
    double* v = p.V + 2 * (int64_t)j * N;
    double s = 0.0;
    for (int i = lane; i < 2 * M; i += 64) s += g[i] * g[i];
    s = wave_sum(s);
    if (lane == 0) p.sigma[j] = sqrt(s);
    const double* u = p.u_in_v ? v : g;
    const int n = p.u_in_v ? N : M;
    double best = -1.0, vr = 0.0, vi = 0.0;
    int bi = 0x7fffffff;
    for (int i = lane; i < n; i += 64) {
      const double x = u[2 * i] * u[2 * i] + u[2 * i + 1] * u[2 * i + 1];
      if (x > best) { best = x; bi = i; vr = u[2 * i]; vi = u[2 * i + 1]; }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double ob = __shfl_xor(best, o, 64);
      const int oi = __shfl_xor(bi, o, 64);
      const double ovr = __shfl_xor(vr, o, 64);
      const double ovi = __shfl_xor(vi, o, 64);
      if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; vr = ovr; vi = ovi; }
    }
    if (best <= 0.0) continue;
    const double m = sqrt(best);
    const double pr = vr / m, pim = -vi / m;  // conj(top) / |top|
    for (int i = lane; i < M; i += 64) {
      const double a = g[2 * i], b = g[2 * i + 1];
      g[2 * i] = a * pr - b * pim;
      g[2 * i + 1] = a * pim + b * pr;
    }
    for (int i = lane; i < N; i += 64) {
      const double a = v[2 * i], b = v[2 * i + 1];
      v[2 * i] = a * pr - b * pim;
      v[2 * i + 1] = a * pim + b * pr;
    }
  }
}

// ---------------------------------------------------------------------------
// QR preconditioning of the sector SVD: blocked Householder QR A = Q R
// (panels of b columns, compact WY form Q_k = I - Y_k T_k Y_k^T), X = R^T,
// and later C <- Q [S; 0] (back-transformation of a seed S).
// Per job, panel k covers columns c0 = k*b .. c0+bw-1 and rows c0 .. M-1
// (m = M - c0 of them); Y_k and P_k = Y_k T_k are m x bw column-major (ld m)
// at offset k*b*M of the job's Y and P buffers.
// ---------------------------------------------------------------------------
struct QrJob {
  double* A;        // M x N column-major (ld M), factorised in place
  double* Y;        // Householder vectors, panel k at Y + k*b*M
  double* P;        // Y_k T_k, same layout
  double* X;        // N x N column-major: R^T
  double* C;        // M x N (ld M): target of the back-transformation
  const double* S;  // N x N seed, C = [S; 0] first (null: C already holds it)
  double* W;        // finish: N x N rotated R^T (sigma_j = |W e_j|)
  double* V;        // finish: N x N output W / sigma
  double* sigma;
  int M, N;
  int u_in_v, finish;
};

// Factorise panel k of every problem (one workgroup each).  kLds: the panel is
// staged in LDS (dynamic, m*b + b*b + b doubles); otherwise it is worked on in
// place in global memory and LDS holds only T and tau (b*b + b doubles).
// kU > 0 (b = kB, m <= 64 kU, LDS sized for 64 kU rows): the column loop keeps the panel in
// registers instead, wave w of kWv holding columns w, w + kWv, ... and lane
// l rows l, l + 64, ...; a column is written to LDS once, by its wave, when
// its last update is done, and each reflector value read from there serves
// the kB / kWv columns of a lane.  A column step is bound by the dependent
// FP64 chain of one wave (reductions, square root, divisions), not by the
// barrier: kU = 9 panels take 71 us with 16 waves against 75 us with 8.
template <bool kLds, int kU = 0, int kWv = 8, int kB = 32>
__global__ __launch_bounds__(64 * kWv) void qr_panel_kernel(const QrJob* __restrict__ probs, int k,
                                                       int b) {
  extern __shared__ double smem[];
  __shared__ double red[kWv];
  const QrJob p = probs[blockIdx.x];
  const int M = p.M, N = p.N;
  const int c0 = k * b;
  if (c0 >= N) return;
  const int bw = min(b, N - c0);
  const int m = M - c0;  // >= bw because M >= N
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = (int)(blockDim.x >> 6);
  double* Ag = p.A + c0 + (int64_t)c0 * M;  // panel origin, ld M
  double* pan;
  int64_t ld;
  double* T;
  if (kLds && kU > 0) {
    pan = smem;
    ld = 64 * kU;
    T = smem + (int64_t)ld * b;
  } else if (kLds) {
    pan = smem;
    ld = m;
    T = smem + (int64_t)m * bw;
    // one column per pass of each wave group, four loads in flight per thread
    for (int j = tid / 128; j < bw; j += (int)blockDim.x / 128) {
      const double* src = Ag + (int64_t)j * M;
      double* dst = pan + (int64_t)j * m;
      for (int i0 = tid % 128; i0 < m; i0 += 512) {
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = i0 + 128 * u < m ? src[i0 + 128 * u] : 0.0;
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (i0 + 128 * u < m) dst[i0 + 128 * u] = v[u];
      }
    }
  } else {
    pan = Ag;
    ld = M;
    T = smem;
  }
  double* tau = T + b * b;
  __syncthreads();
  // Column j: every thread forms the reflector H = I - tau v v^T with
  // v = (1, x[1:] * scal) from |x[1:]|^2 (the same value for every thread),
  // the trailing panel columns are updated in parallel (16 threads per
  // column, shuffle-reduced dot), and x is scaled into v after the barrier.
  // The 16 threads that update column j+1 also form its |x[1:]|^2 for the
  // next step (column 0's comes from a block reduction), so a column costs
  // one barrier.
  __shared__ double colss[2];
  if constexpr (kU > 0) {
    // wave w holds columns w, w + kWv, ... of the panel, lane l rows l + 64 u
    // (u < kU): a reflector value read from LDS serves kQ columns.
    // The LDS columns have 64 kU rows (ld), zero past m, so only register 0
    // (rows 0 .. 63, the ones that can be <= j < kB) needs a row test.
    constexpr int nwv = kWv;
    static_assert(kB % kWv == 0 && kB <= 64, "whole columns per wave, panel rows j < 64 in register 0");
    constexpr int kQ = kB / kWv;  // columns per wave
    double y[kQ][kU];
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
      const int c = wave + nwv * q;
      const double* src = Ag + (int64_t)min(c, bw - 1) * M;
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int i = lane + 64 * u;
        const double v = src[min(i, m - 1)];  // clamped, in bounds
        y[q][u] = (c < bw && i < m) ? v : 0.0;
      }
    }
    // column 0 to LDS with |x[1:]|^2
    if (wave == 0) {
      double ss = 0.0;
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        pan[lane + 64 * u] = y[0][u];
        ss += (u > 0 || lane >= 1) ? y[0][u] * y[0][u] : 0.0;
      }
      ss = wave_sum_dpp(ss);
      if (lane == 0) colss[0] = ss;
    }
    __syncthreads();
    double scal_prev = 0.0, tj_prev = 0.0, beta_prev = 0.0;
    for (int j = 0; j < bw; ++j) {
      // column j - 1 (read by nobody from here on) as v = (beta, scal x[j:]),
      // from its owner's registers, which hold its final values
      if (j > 0 && (j - 1) % nwv == wave) {
        const int q = (j - 1) / nwv;
        double* xp = pan + (int64_t)(j - 1) * ld;
        const double sc = tj_prev != 0.0 ? scal_prev : 1.0;
#pragma unroll
        for (int qq = 0; qq < kQ; ++qq)
          if (qq == q) {
            if (lane >= j - 1) xp[lane] = lane == j - 1 ? beta_prev : y[qq][0] * sc;
#pragma unroll
            for (int u = 1; u < kU; ++u) xp[lane + 64 * u] = y[qq][u] * sc;
          }
      }
      const double* x = pan + (int64_t)j * ld;  // column j
      // reflector values first (they do not wait for the scalars)
      double xv[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) xv[u] = x[lane + 64 * u];
      xv[0] = lane > j ? xv[0] : 0.0;
      const double t = colss[j & 1];
      const double alpha = x[j];
      // x^T y over rows > j for the wave's kQ columns, reduced together
      // (a column that is done or absent gives an unused value)
      double dot[kQ];
#pragma unroll
      for (int q = 0; q < kQ; ++q) {
        double dp[2] = {0.0, 0.0};
#pragma unroll
        for (int u = 0; u < kU; ++u) dp[u & 1] += xv[u] * y[q][u];
        dot[q] = row16_sum(dp[0] + dp[1]);
      }
#pragma unroll
      for (int q = 0; q < kQ; ++q)
        dot[q] = (readlane_d(dot[q], 0) + readlane_d(dot[q], 16)) +
                 (readlane_d(dot[q], 32) + readlane_d(dot[q], 48));
      double tj = 0.0, beta = alpha, scal = 0.0;
      if (t > 0.0) {
        beta = -copysign(sqrt(alpha * alpha + t), alpha);
        tj = (beta - alpha) / beta;
        scal = 1.0 / (alpha - beta);
      }
      if (tid == 0) tau[j] = tj;
#pragma unroll
      for (int q = 0; q < kQ; ++q) {
        const int c = wave + nwv * q;
        if (c <= j || c >= bw) continue;  // uniform over the wave
        // d = tau (y_j + scal x^T y); y -= d v, v = (1, scal x[j+1:]); row
        // j < kB sits in lane j, register 0
        if (tj != 0.0) {
          const double yj = readlane_d(y[q][0], j);
          const double d = (yj + scal * dot[q]) * tj;
          const double ds = d * scal;
          y[q][0] -= lane == j ? d : ds * xv[0];
#pragma unroll
          for (int u = 1; u < kU; ++u) y[q][u] -= ds * xv[u];
        }
        if (c == j + 1) {  // final: to LDS with |x[1:]|^2 for the next step
          double* xo = pan + (int64_t)c * ld;
          double ss = 0.0;
#pragma unroll
          for (int u = 0; u < kU; ++u) {
            xo[lane + 64 * u] = y[q][u];
            ss += (u > 0 || lane > c) ? y[q][u] * y[q][u] : 0.0;
          }
          ss = wave_sum_dpp(ss);
          if (lane == 0) colss[(j + 1) & 1] = ss;
        }
      }
      scal_prev = scal;
      tj_prev = tj;
      beta_prev = beta;
      __syncthreads();
    }
    if ((bw - 1) % nwv == wave) {
      const int q = (bw - 1) / nwv;
      double* xp = pan + (int64_t)(bw - 1) * ld;
      const double sc = tj_prev != 0.0 ? scal_prev : 1.0;
#pragma unroll
      for (int qq = 0; qq < kQ; ++qq)
        if (qq == q) {
          if (lane >= bw - 1) xp[lane] = lane == bw - 1 ? beta_prev : y[qq][0] * sc;
#pragma unroll
          for (int u = 1; u < kU; ++u) xp[lane + 64 * u] = y[qq][u] * sc;
        }
    }
  } else {
  {
    double ss = 0.0;
    for (int i = 1 + tid; i < m; i += blockDim.x) ss += pan[i] * pan[i];
    ss = wave_sum(ss);
    if (lane == 0) red[wave] = ss;
    __syncthreads();
    if (tid == 0) {
      double t = 0.0;
      for (int w = 0; w < nw; ++w) t += red[w];
      colss[0] = t;
    }
    __syncthreads();
  }
  for (int j = 0; j < bw; ++j) {
    double* x = pan + (int64_t)j * ld + j;  // x[0 .. len-1], len = m - j
    const int len = m - j;
    const double t = colss[j & 1];
    const double alpha = x[0];
    double tj = 0.0, beta = alpha, scal = 0.0;
    if (t > 0.0) {
      beta = -copysign(sqrt(alpha * alpha + t), alpha);
      tj = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
    if (tid == 0) tau[j] = tj;
    const int c = j + 1 + (tid >> 4), sub = tid & 15;
    if (c < bw) {
      double* y = pan + (int64_t)c * ld + j;
      const bool next = c == j + 1;  // this group also forms column j+1's norm
      double acc = 0.0;
      if (tj != 0.0) {
        double d = 0.0;
        // unrolled so that the LDS reads of four iterations are in flight
        // together (same accumulation order)
#pragma unroll 4
        for (int i = 1 + sub; i < len; i += 16) d += x[i] * y[i];
        d += __shfl_xor(d, 8, 64);
        d += __shfl_xor(d, 4, 64);
        d += __shfl_xor(d, 2, 64);
        d += __shfl_xor(d, 1, 64);
        d = (y[0] + scal * d) * tj;  // tau v^T y
        const double ds = d * scal;
#pragma unroll 4
        for (int i = 1 + sub; i < len; i += 16) {
          const double yn = y[i] - ds * x[i];
          y[i] = yn;
          if (next && i >= 2) acc += yn * yn;
        }
        if (sub == 0) y[0] -= d;
      } else if (next) {
        for (int i = 2 + sub; i < len; i += 16) acc += y[i] * y[i];
      }
      if (next) {
        acc += __shfl_xor(acc, 8, 64);
        acc += __shfl_xor(acc, 4, 64);
        acc += __shfl_xor(acc, 2, 64);
        acc += __shfl_xor(acc, 1, 64);
        if (sub == 0) colss[(j + 1) & 1] = acc;
      }
    }
    __syncthreads();
    if (tj != 0.0)
      for (int i = 1 + tid; i < len; i += blockDim.x) x[i] *= scal;
    if (tid == 0) x[0] = beta;
  }
  }
  __syncthreads();
  // S = Y^T Y with FP64 MFMA, one wave per upper 16 x 16 block (a <= b);
  // Y[r][c] = 0 (r < c), 1 (r = c), v_c[r] (r > c).  The strictly upper part
  // of S goes to the strictly lower triangle of T (T[j][l] = S[l][j], l < j).
  {
    const int nbk = (bw + 15) >> 4;
    int blk = wave, ba = 0, bb = 0;
    for (int a = 0; a < nbk; ++a)
      for (int c2 = a; c2 < nbk; ++c2) {
        if (blk == 0) { ba = a; bb = c2; }
        --blk;
      }
    if (blk < 0) {  // this wave owns block (ba, bb)
      const int i = lane & 15, kk = lane >> 4;
      const int cl = ba * 16 + i, cj = bb * 16 + i;
      d4v acc = {0.0, 0.0, 0.0, 0.0};
      // loads from clamped (in-bounds) addresses first, the triangle and edge
      // selects after them, so a slab's eight loads are in flight together
      const double* pa = pan + (int64_t)min(cl, bw - 1) * ld;
      const double* pb = pan + (int64_t)min(cj, bw - 1) * ld;
      for (int base = 0; base < m; base += 16) {
        double va[4], vb[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int rc = min(base + 4 * kk + t, m - 1);
          va[t] = pa[rc];
          vb[t] = pb[rc];
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int r = base + 4 * kk + t;
          const double ya = (r < m && cl < bw) ? (r < cl ? 0.0 : (r == cl ? 1.0 : va[t])) : 0.0;
          const double yb = (r < m && cj < bw) ? (r < cj ? 0.0 : (r == cj ? 1.0 : vb[t])) : 0.0;
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ya, yb, acc, 0, 0, 0);
        }
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int l = ba * 16 + kk + 4 * v, j = bb * 16 + i;
        if (l < j && j < bw) T[j * b + l] = acc[v];
      }
    }
  }
  __syncthreads();
  // T upper triangular (dlarft): T[j][j] = tau_j, T[0:j, j] = -tau_j T[0:j, 0:j] S[0:j, j]
  // (S[l][j] = y_l^T y_j sits at T[j][l]).  Row r of T depends only on its own
  // earlier entries, T[r][j] = -tau_j sum_{r <= l < j} T[r][l] S[l][j], so the
  // 16 lanes of row r build it left to right with no barrier: lane sub keeps
  // T[r][r + sub] and T[r][r + sub + 16] in registers and the partial sums are
  // reduced within the 16 lanes by DPP.
  {
    const int r = tid >> 4, sub = tid & 15;
    double t0 = 0.0, t1 = 0.0;
    if (r < bw) {
      for (int j = r; j < bw; ++j) {
        double sum = 0.0;
        const int l0 = r + sub, l1 = r + sub + 16;
        if (l0 < j) sum += t0 * T[j * b + l0];
        if (l1 < j) sum += t1 * T[j * b + l1];
        sum = row16_sum(sum);  // the 16 lanes of row r form one DPP row
        const double v = j == r ? tau[j] : -tau[j] * sum;
        if (j - r == sub) t0 = v;
        if (j - r == sub + 16) t1 = v;
      }
      if (r + sub < bw) T[r * b + r + sub] = t0;
      if (r + sub + 16 < bw) T[r * b + r + sub + 16] = t1;
    }
  }
  __syncthreads();
  // Y_k (unit lower trapezoidal) and P_k = Y_k T_k
  double* Yk = p.Y + (int64_t)k * b * M;
  double* Pk = p.P + (int64_t)k * b * M;
  for (int64_t e = tid; e < (int64_t)m * bw; e += blockDim.x) {
    const int j = (int)(e / m), i = (int)(e % m);
    Yk[e] = i < j ? 0.0 : (i == j ? 1.0 : pan[(int64_t)j * ld + i]);
  }
  {
    // P = Y T with FP64 MFMA: 16 x 16 output tiles over the waves, K = the
    // panel's columns in steps of 4 (Y[i][l] = 0 for i < l, 1 for i = l; T
    // upper triangular, so K stops at the tile's last column; its strict lower
    // triangle still holds Y^T Y and is read as zero)
    const int i16 = lane & 15, kk = lane >> 4;
    const int nib = (m + 15) >> 4, njb = (bw + 15) >> 4;
    for (int tile = wave; tile < nib * njb; tile += nw) {
      const int ib = tile / njb, jb = tile % njb;
      const int ia = ib * 16 + i16;       // row of the A operand (Y)
      const int jc = jb * 16 + i16;       // column of the B operand (T)
      const int kend = min(bw, jb * 16 + 16);
      d4v acc = {0.0, 0.0, 0.0, 0.0};
      for (int l0 = 0; l0 < kend; l0 += 4) {
        const int l = l0 + kk;
        const double pv = pan[(int64_t)min(l, bw - 1) * ld + min(ia, m - 1)];  // clamped, in bounds
        const double tv = T[min(l, bw - 1) * b + min(jc, bw - 1)];
        const double ya = (l < bw && ia < m) ? (ia < l ? 0.0 : (ia == l ? 1.0 : pv)) : 0.0;
        const double tb = (l <= jc && jc < bw) ? tv : 0.0;  // upper triangle only (S below)
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ya, tb, acc, 0, 0, 0);
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = ib * 16 + kk + 4 * v, col = jb * 16 + i16;
        if (row < m && col < bw) Pk[(int64_t)col * m + row] = acc[v];
      }
    }
  }
  if (kLds) {  // R block of the panel back to A (upper triangle of its top bw rows)
    for (int64_t e = tid; e < (int64_t)bw * bw; e += blockDim.x) {
      const int j = (int)(e / bw), i = (int)(e % bw);
      if (i <= j) Ag[i + (int64_t)j * M] = pan[(int64_t)j * ld + i];
    }
  }
}

// Block reflector of panel k applied to a matrix, X <- X - L (R^T X), with
// X the trailing columns of A and (L, R) = (Y_k, P_k) while factorising
// (Q_k^T A), or X = C and (L, R) = (P_k, Y_k) in the back-transformation
// (Q_k C).  Workgroup (column block, job) takes kQrUpCols columns of X
// (m rows, from row c0): Z = R^T X_blk (b x kQrUpCols, K = m) on FP64 MFMA
// with K split over the waves and the partial Z summed in a fixed order in
// LDS, then X_blk -= L Z in 16-row slabs.  One launch per panel replaces the
// two grouped GEMMs of the same product and spreads it over every column
// block of every job.
constexpr int kQrUpCols = 16;
template <bool kBack>
__global__ __launch_bounds__(512) void qr_update_kernel(const QrJob* __restrict__ probs, int k, int b) {
  constexpr int kW = 8;
  __shared__ double Zp[kW][32][kQrUpCols + 1];
  __shared__ double Z[32][kQrUpCols + 1];
  const QrJob p = probs[blockIdx.y];
  const int M = p.M, N = p.N;
  const int c0 = k * b;
  if (c0 >= N) return;
  const int bw = min(b, N - c0), m = M - c0;
  const int x0 = kBack ? 0 : c0 + bw;  // first column of X
  const int ncol = N - x0;
  const int j0 = blockIdx.x * kQrUpCols;
  if (j0 >= ncol) return;
  double* X = (kBack ? p.C : p.A) + c0 + (int64_t)(x0 + j0) * M;  // ld M, m rows
  const double* Yk = p.Y + (int64_t)k * b * M;                   // ld m
  const double* Pk = p.P + (int64_t)k * b * M;
  const double* L = kBack ? Pk : Yk;
  const double* R = kBack ? Yk : Pk;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int i = lane & 15, kk = lane >> 4;
  const bool colok = j0 + i < ncol;
  const double* xc = X + (int64_t)(colok ? i : 0) * M;
  // 1. partial Z over this wave's row chunks of 16: lane (i, kk) feeds rows
  //    r0 + 4 kk .. r0 + 4 kk + 3 (one MFMA each) of Z rows 16 h + i
  {
    d4v acc[2] = {d4v{0.0, 0.0, 0.0, 0.0}, d4v{0.0, 0.0, 0.0, 0.0}};
    const double* rc[2];
    bool rok[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      rok[h] = 16 * h + i < bw;
      rc[h] = R + (int64_t)(rok[h] ? 16 * h + i : 0) * m;
    }
    for (int r0 = 16 * wave; r0 < m; r0 += 16 * kW) {
      double xa[4], ra[2][4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int r = r0 + 4 * kk + t;
        const int rr = min(r, m - 1);
        xa[t] = (colok && r < m) ? xc[rr] : 0.0;
#pragma unroll
        for (int h = 0; h < 2; ++h) ra[h][t] = (rok[h] && r < m) ? rc[h][rr] : 0.0;
      }
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int h = 0; h < 2; ++h)
          acc[h] = __builtin_amdgcn_mfma_f64_16x16x4f64(ra[h][t], xa[t], acc[h], 0, 0, 0);
    }
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int v = 0; v < 4; ++v) Zp[wave][16 * h + kk + 4 * v][i] = acc[h][v];
  }
  __syncthreads();
  for (int e = tid; e < 32 * kQrUpCols; e += blockDim.x) {
    const int r = e / kQrUpCols, c = e % kQrUpCols;
    double s = Zp[0][r][c];
#pragma unroll
    for (int w = 1; w < kW; ++w) s += Zp[w][r][c];
    Z[r][c] = s;
  }
  __syncthreads();
  // 2. X_blk -= L Z, 16-row slabs over the waves (K = the panel's columns)
  double zb[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) zb[t] = Z[4 * t + kk][i];  // rows >= bw are zero
  for (int r0 = 16 * wave; r0 < m; r0 += 16 * kW) {
    const int ra = min(r0 + i, m - 1);
    double la[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int l = 4 * t + kk;
      const double lv = L[(int64_t)min(l, bw - 1) * m + ra];  // clamped, in bounds
      la[t] = (l < bw && r0 + i < m) ? -lv : 0.0;
    }
    d4v acc;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = r0 + kk + 4 * v;
      const double xv = xc[min(r, m - 1)];
      acc[v] = (colok && r < m) ? xv : 0.0;
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(la[t], zb[t], acc, 0, 0, 0);
    if (colok) {
      double* xw = X + (int64_t)i * M;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = r0 + kk + 4 * v;
        if (r < m) xw[r] = acc[v];
      }
    }
  }
}

// X = R^T (N x N column-major) from the upper triangle of the factorised A
__global__ __launch_bounds__(256) void qr_rt_kernel(const QrJob* __restrict__ probs) {
  const QrJob p = probs[blockIdx.y];
  const int N = p.N, M = p.M;
  const int64_t nn = (int64_t)N * N;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nn;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / N), j = (int)(e % N);  // X[j + i N] = R[i][j]
    p.X[e] = j >= i ? p.A[i + (int64_t)j * M] : 0.0;
  }
}

// C = [S; 0] (M x N) ahead of the back-transformation C <- Q C
__global__ __launch_bounds__(256) void qr_seed_kernel(const QrJob* __restrict__ probs) {
  const QrJob p = probs[blockIdx.y];
  if (!p.S) return;
  const int N = p.N, M = p.M;
  const int64_t mn = (int64_t)M * N;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < mn;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int j = (int)(e / M), i = (int)(e % M);
    p.C[e] = i < N ? p.S[i + (int64_t)j * N] : 0.0;
  }
}

// Jobs with finish set (C = Q J, the left factor, exactly orthonormal):
// sigma_j = |W e_j|, sign convention on the left vector (column j of C, or of
// W when the problem holds the transposed matrix), then C_j <- sigma_j C_j and
// V_j <- W_j / sigma_j (zero when sigma_j = 0).  One wave per column.
__global__ __launch_bounds__(256) void qr_finish_kernel(const QrJob* __restrict__ probs) {
  const QrJob p = probs[blockIdx.x];
  if (!p.finish) return;
  const int M = p.M, N = p.N;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  const int jend = min(N, (int)(blockIdx.y + 1) * kFinCols);
  for (int j = blockIdx.y * kFinCols + wave; j < jend; j += nwaves) {
    double* a = p.C + (int64_t)j * M;
    double* w = p.W + (int64_t)j * N;
    double* v = p.V + (int64_t)j * N;
    double s = 0.0;
    for (int i = lane; i < N; i += 64) s += w[i] * w[i];
    s = sqrt(wave_sum(s));
    if (lane == 0) p.sigma[j] = s;
    const double* u = p.u_in_v ? w : a;
    const int n = p.u_in_v ? N : M;
    double best = -1.0, val = 0.0;
    int bi = 0x7fffffff;
    for (int i = lane; i < n; i += 64) {
      const double x = fabs(u[i]);
      if (x > best) { best = x; bi = i; val = u[i]; }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double ob = __shfl_xor(best, o, 64);
      const int oi = __shfl_xor(bi, o, 64);
      const double ov = __shfl_xor(val, o, 64);
      if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; val = ov; }
    }
    const double sg = val < 0.0 ? -1.0 : 1.0;
    const double sa = sg * s, sv = s > 0.0 ? sg / s : 0.0;
    for (int i = lane; i < M; i += 64) a[i] *= sa;
    for (int i = lane; i < N; i += 64) v[i] = w[i] * sv;
  }
}

// ---------------------------------------------------------------------------
// Complex QR preconditioning.  A, Y, P, X and C hold interleaved complex
// entries; leading dimensions count complex entries.  Reflectors follow
// LAPACK's zlarfg: H = I - tau v v^H with H^H x = (beta, 0, ..., 0), beta
// real, v = (1, ...).  Panel k has Q_k = H_1 ... H_bw = I - Y T Y^H (T upper
// triangular, zlarft's forward columnwise recursion) and P_k = Y_k T_k, so the
// factorisation applies Q_k^H = I - Y P^H to the trailing columns and the
// back-transformation applies Q_k = I - P Y^H: both are X <- X - L (R^H X),
// as in the real form.  Panels are at most kZqrB columns wide.
// ---------------------------------------------------------------------------
constexpr int kZqrB = 16;

// Panel k of every problem, one workgroup each.  kLds: the panel is staged in
// LDS (2 (m b + b b + b) doubles of dynamic LDS); otherwise it is worked on in
// place in global memory and LDS holds T and tau (2 (b b + b) doubles).
template <bool kLds>
__global__ __launch_bounds__(512) void zqr_panel_kernel(const QrJob* __restrict__ probs, int k, int b) {
  extern __shared__ double smem[];
  __shared__ double red[8];
  __shared__ double colss[2];
  const QrJob p = probs[blockIdx.x];
  const int M = p.M, N = p.N;
  const int c0 = k * b;
  if (c0 >= N) return;
  const int bw = min(b, N - c0);
  const int m = M - c0;  // >= bw because M >= N
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = (int)(blockDim.x >> 6);
  double* Ag = p.A + 2 * (c0 + (int64_t)c0 * M);  // panel origin, ld M
  double* pan;
  int64_t ld;
  double* T;
  if (kLds) {
    pan = smem;
    ld = m;
    T = smem + 2 * (int64_t)m * bw;
    const int64_t n2 = 2 * (int64_t)m;
    for (int64_t e = tid; e < n2 * bw; e += blockDim.x) {
      const int j = (int)(e / n2);
      pan[e] = Ag[e - j * n2 + 2 * (int64_t)j * M];
    }
  } else {
    pan = Ag;
    ld = M;
    T = smem;
  }
  double* tau = T + 2 * b * b;
  __syncthreads();
  {
    double ss = 0.0;  // |x[1:]|^2 of column 0
    for (int i = 2 + tid; i < 2 * m; i += blockDim.x) ss += pan[i] * pan[i];
    ss = wave_sum(ss);
    if (lane == 0) red[wave] = ss;
    __syncthreads();
    if (tid == 0) {
      double t = 0.0;
      for (int w = 0; w < nw; ++w) t += red[w];
      colss[0] = t;
    }
    __syncthreads();
  }
  // Column j: every thread forms the reflector from alpha = x[0] and
  // |x[1:]|^2 (the same values for every thread); 32 threads per trailing
  // column apply H^H = I - conj(tau) v v^H (shuffle-reduced dot), the group of
  // column j + 1 also forming that column's |x[1:]|^2 for the next step, and x
  // is scaled into v after the barrier.  No reflector (tau = 0) when x[1:] = 0
  // and alpha is real.
  for (int j = 0; j < bw; ++j) {
    double* x = pan + 2 * ((int64_t)j * ld + j);  // x[0 .. len-1], len = m - j
    const int len = m - j;
    const double t = colss[j & 1];
    const double ar = x[0], ai = x[1];
    const bool refl = t > 0.0 || ai != 0.0;
    double tr = 0.0, ti = 0.0, beta = ar, sr = 0.0, si = 0.0;  // tau, beta, scal = 1 / (alpha - beta)
    if (refl) {
      beta = -copysign(sqrt(ar * ar + ai * ai + t), ar);
      tr = (beta - ar) / beta;
      ti = -ai / beta;
      const double dr = ar - beta, den = dr * dr + ai * ai;
      sr = dr / den;
      si = -ai / den;
    }
    if (tid == 0) {
      tau[2 * j] = tr;
      tau[2 * j + 1] = ti;
    }
    const int c = j + 1 + (tid >> 5), sub = tid & 31;
    if (c < bw) {
      double* y = pan + 2 * ((int64_t)c * ld + j);
      const bool next = c == j + 1;
      double acc = 0.0;
      if (refl) {
        double dr = 0.0, di = 0.0;  // sum over i >= 1 of conj(x_i) y_i
#pragma unroll 4
        for (int i = 1 + sub; i < len; i += 32) {
          const double xr = x[2 * i], xi = x[2 * i + 1], yr = y[2 * i], yi = y[2 * i + 1];
          dr += xr * yr + xi * yi;
          di += xr * yi - xi * yr;
        }
#pragma unroll
        for (int o = 16; o > 0; o >>= 1) {
          dr += __shfl_xor(dr, o, 64);
          di += __shfl_xor(di, o, 64);
        }
        // w = v^H y = y_0 + conj(scal) dot; d = conj(tau) w; y -= d v
        const double wr = y[0] + (sr * dr + si * di), wi = y[1] + (sr * di - si * dr);
        const double Dr = tr * wr + ti * wi, Di = tr * wi - ti * wr;
        const double Sr = Dr * sr - Di * si, Si = Dr * si + Di * sr;  // d scal
#pragma unroll 4
        for (int i = 1 + sub; i < len; i += 32) {
          const double xr = x[2 * i], xi = x[2 * i + 1];
          const double yr = y[2 * i] - (Sr * xr - Si * xi), yi = y[2 * i + 1] - (Sr * xi + Si * xr);
          y[2 * i] = yr;
          y[2 * i + 1] = yi;
          if (next && i >= 2) acc += yr * yr + yi * yi;
        }
        if (sub == 0) {
          y[0] -= Dr;
          y[1] -= Di;
        }
      } else if (next) {
        for (int i = 2 + sub; i < len; i += 32) acc += y[2 * i] * y[2 * i] + y[2 * i + 1] * y[2 * i + 1];
      }
      if (next) {
#pragma unroll
        for (int o = 16; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
        if (sub == 0) colss[(j + 1) & 1] = acc;
      }
    }
    __syncthreads();
    if (refl)
      for (int i = 1 + tid; i < len; i += blockDim.x) {
        const double xr = x[2 * i], xi = x[2 * i + 1];
        x[2 * i] = xr * sr - xi * si;
        x[2 * i + 1] = xr * si + xi * sr;
      }
    if (tid == 0) {
      x[0] = beta;
      x[1] = 0.0;
    }
  }
  __syncthreads();
  // S[l][j] = y_l^H y_j (l < j) into the strict lower triangle of T (T[j][l]),
  // four threads per pair; Y[r][c] = 0 (r < c), 1 (r = c), v_c[r] (r > c)
  {
    const int npair = bw * (bw - 1) / 2;
    const int pr = tid >> 2, sub = tid & 3;
    if (pr < npair) {  // uniform over the pair's four lanes
      int j = 1;
      while ((j + 1) * j / 2 <= pr) ++j;
      const int l = pr - j * (j - 1) / 2;
      const double* yl = pan + 2 * (int64_t)l * ld;
      const double* yj = pan + 2 * (int64_t)j * ld;
      double s_r = 0.0, s_i = 0.0;
      for (int r = j + 1 + sub; r < m; r += 4) {
        const double a = yl[2 * r], bq = yl[2 * r + 1], cq = yj[2 * r], dq = yj[2 * r + 1];
        s_r += a * cq + bq * dq;
        s_i += a * dq - bq * cq;
      }
      s_r += __shfl_xor(s_r, 2, 64);
      s_i += __shfl_xor(s_i, 2, 64);
      s_r += __shfl_xor(s_r, 1, 64);
      s_i += __shfl_xor(s_i, 1, 64);
      if (sub == 0) {  // row j: conj(v_l[j]) * 1
        T[2 * (j * b + l)] = s_r + yl[2 * j];
        T[2 * (j * b + l) + 1] = s_i - yl[2 * j + 1];
      }
    }
  }
  __syncthreads();
  // T upper triangular: T[j][j] = tau_j, T[0:j, j] = -tau_j T[0:j, 0:j] S[0:j, j]
  // (S[l][j] at T[j][l]); as in the real panel, the 16 lanes of row r build
  // it left to right with no barrier, lane sub keeping T[r][r + sub]
  {
    const int r = tid >> 4, sub = tid & 15;  // bw <= 16: one entry per lane
    double tr0 = 0.0, ti0 = 0.0;
    if (r < bw) {
      for (int j = r; j < bw; ++j) {
        double s_r = 0.0, s_i = 0.0;
        const int l = r + sub;
        if (l < j) {
          const double cq = T[2 * (j * b + l)], dq = T[2 * (j * b + l) + 1];
          s_r += tr0 * cq - ti0 * dq;
          s_i += tr0 * dq + ti0 * cq;
        }
        s_r = row16_sum(s_r);
        s_i = row16_sum(s_i);
        const double tjr = tau[2 * j], tji = tau[2 * j + 1];
        const double vr = j == r ? tjr : -(tjr * s_r - tji * s_i);
        const double vi = j == r ? tji : -(tjr * s_i + tji * s_r);
        if (j - r == sub) {
          tr0 = vr;
          ti0 = vi;
        }
      }
      if (r + sub < bw) {
        T[2 * (r * b + r + sub)] = tr0;
        T[2 * (r * b + r + sub) + 1] = ti0;
      }
    }
  }
  __syncthreads();
  // Y_k (unit lower trapezoidal) and P_k = Y_k T_k, m x bw each (ld m)
  double* Yk = p.Y + 2 * (int64_t)k * b * M;
  double* Pk = p.P + 2 * (int64_t)k * b * M;
  for (int64_t e = tid; e < (int64_t)m * bw; e += blockDim.x) {
    const int j = (int)(e / m), i = (int)(e % m);
    const double* yi = pan + 2 * ((int64_t)j * ld + i);
    Yk[2 * e] = i < j ? 0.0 : (i == j ? 1.0 : yi[0]);
    Yk[2 * e + 1] = i <= j ? 0.0 : yi[1];
    double s_r = 0.0, s_i = 0.0;
    const int lmax = min(j, i);
    for (int l = 0; l <= lmax; ++l) {
      const double* yl = pan + 2 * ((int64_t)l * ld + i);
      const double a = l == i ? 1.0 : yl[0], bq = l == i ? 0.0 : yl[1];
      const double cq = T[2 * (l * b + j)], dq = T[2 * (l * b + j) + 1];
      s_r += a * cq - bq * dq;
      s_i += a * dq + bq * cq;
    }
    Pk[2 * e] = s_r;
    Pk[2 * e + 1] = s_i;
  }
  if (kLds) {  // R block of the panel back to A (upper triangle of its top bw rows)
    for (int64_t e = tid; e < (int64_t)bw * bw; e += blockDim.x) {
      const int j = (int)(e / bw), i = (int)(e % bw);
      if (i <= j) {
        Ag[2 * (i + (int64_t)j * M)] = pan[2 * ((int64_t)j * ld + i)];
        Ag[2 * (i + (int64_t)j * M) + 1] = pan[2 * ((int64_t)j * ld + i) + 1];
      }
    }
  }
}

// Complex block reflector of panel k, X <- X - L (R^H X) with (L, R) = (Y_k,
// P_k) while factorising and (P_k, Y_k) in the back-transformation.  As
// qr_update_kernel: workgroup (column block, job) takes kQrUpCols columns of
// X; Z = R^H X_blk on FP64 MFMA (four real chains: Re Z = Rr^T Xr + Ri^T Xi,
// Im Z = Rr^T Xi - Ri^T Xr) with the rows split over the waves and the partial
// Z summed in a fixed order in LDS, then X_blk -= L Z in 16-row slabs.
template <bool kBack>
__global__ __launch_bounds__(512) void zqr_update_kernel(const QrJob* __restrict__ probs, int k, int b) {
  constexpr int kW = 8;
  __shared__ double Zp[kW][2][16][kQrUpCols + 1];
  __shared__ double Z[2][16][kQrUpCols + 1];
  const QrJob p = probs[blockIdx.y];
  const int M = p.M, N = p.N;
  const int c0 = k * b;
  if (c0 >= N) return;
  const int bw = min(b, N - c0), m = M - c0;
  const int x0 = kBack ? 0 : c0 + bw;  // first column of X
  const int ncol = N - x0;
  const int j0 = blockIdx.x * kQrUpCols;
  if (j0 >= ncol) return;
  double* X = (kBack ? p.C : p.A) + 2 * (c0 + (int64_t)(x0 + j0) * M);  // ld M, m rows
  const double* Yk = p.Y + 2 * (int64_t)k * b * M;                       // ld m
  const double* Pk = p.P + 2 * (int64_t)k * b * M;
  const double* L = kBack ? Pk : Yk;
  const double* R = kBack ? Yk : Pk;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int i = lane & 15, kk = lane >> 4;
  const bool colok = j0 + i < ncol;
  const double* xc = X + 2 * (int64_t)(colok ? i : 0) * M;
  const bool rok = i < bw;
  const double* rc = R + 2 * (int64_t)(rok ? i : 0) * m;
  {
    d4v zr = {0.0, 0.0, 0.0, 0.0}, zi = {0.0, 0.0, 0.0, 0.0};
    for (int r0 = 16 * wave; r0 < m; r0 += 16 * kW) {
      double xr[4], xi[4], rr[4], ri[4], nri[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int r = r0 + 4 * kk + t;
        const int rq = min(r, m - 1);
        const double a = xc[2 * rq], bq = xc[2 * rq + 1], cq = rc[2 * rq], dq = rc[2 * rq + 1];
        const bool okx = colok && r < m, okr = rok && r < m;
        xr[t] = okx ? a : 0.0;
        xi[t] = okx ? bq : 0.0;
        rr[t] = okr ? cq : 0.0;
        ri[t] = okr ? dq : 0.0;
        nri[t] = -ri[t];
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        zr = __builtin_amdgcn_mfma_f64_16x16x4f64(rr[t], xr[t], zr, 0, 0, 0);
        zr = __builtin_amdgcn_mfma_f64_16x16x4f64(ri[t], xi[t], zr, 0, 0, 0);
        zi = __builtin_amdgcn_mfma_f64_16x16x4f64(rr[t], xi[t], zi, 0, 0, 0);
        zi = __builtin_amdgcn_mfma_f64_16x16x4f64(nri[t], xr[t], zi, 0, 0, 0);
      }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      Zp[wave][0][kk + 4 * v][i] = zr[v];
      Zp[wave][1][kk + 4 * v][i] = zi[v];
    }
  }
  __syncthreads();
  for (int e = tid; e < 2 * 16 * kQrUpCols; e += blockDim.x) {
    const int h = e / (16 * kQrUpCols), r = (e / kQrUpCols) % 16, c = e % kQrUpCols;
    double s = Zp[0][h][r][c];
#pragma unroll
    for (int w = 1; w < kW; ++w) s += Zp[w][h][r][c];
    Z[h][r][c] = s;
  }
  __syncthreads();
  // Re X -= Lr Zr - Li Zi, Im X -= Lr Zi + Li Zr (rows of Z >= bw are zero)
  double zbr[4], zbi[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    zbr[t] = Z[0][4 * t + kk][i];
    zbi[t] = Z[1][4 * t + kk][i];
  }
  for (int r0 = 16 * wave; r0 < m; r0 += 16 * kW) {
    const int ra = min(r0 + i, m - 1);
    double nlr[4], li[4], nli[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int l = 4 * t + kk;
      const double* lp = L + 2 * ((int64_t)min(l, bw - 1) * m + ra);  // clamped, in bounds
      const double a = lp[0], bq = lp[1];
      const bool ok = l < bw && r0 + i < m;
      nlr[t] = ok ? -a : 0.0;
      li[t] = ok ? bq : 0.0;
      nli[t] = -li[t];
    }
    d4v ar, ai;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = r0 + kk + 4 * v;
      const int rq = min(r, m - 1);
      const double a = xc[2 * rq], bq = xc[2 * rq + 1];
      const bool ok = colok && r < m;
      ar[v] = ok ? a : 0.0;
      ai[v] = ok ? bq : 0.0;
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      ar = __builtin_amdgcn_mfma_f64_16x16x4f64(nlr[t], zbr[t], ar, 0, 0, 0);
      ar = __builtin_amdgcn_mfma_f64_16x16x4f64(li[t], zbi[t], ar, 0, 0, 0);
      ai = __builtin_amdgcn_mfma_f64_16x16x4f64(nlr[t], zbi[t], ai, 0, 0, 0);
      ai = __builtin_amdgcn_mfma_f64_16x16x4f64(nli[t], zbr[t], ai, 0, 0, 0);
    }
    if (colok) {
      double* xw = X + 2 * (int64_t)i * M;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = r0 + kk + 4 * v;
        if (r < m) {
          xw[2 * r] = ar[v];
          xw[2 * r + 1] = ai[v];
        }
      }
    }
  }
}

// X = R^H (N x N complex) from the upper triangle of the factorised A
__global__ __launch_bounds__(256) void zqr_rt_kernel(const QrJob* __restrict__ probs) {
  const QrJob p = probs[blockIdx.y];
  const int N = p.N, M = p.M;
  const int64_t nn = (int64_t)N * N;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nn;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / N), j = (int)(e % N);  // X[j + i N] = conj(R[i][j])
    const double* a = p.A + 2 * (i + (int64_t)j * M);
    const bool up = j >= i;
    p.X[2 * e] = up ? a[0] : 0.0;
    p.X[2 * e + 1] = up ? -a[1] : 0.0;
  }
}

// C = [S; 0] (M x N complex) ahead of the back-transformation C <- Q C
__global__ __launch_bounds__(256) void zqr_seed_kernel(const QrJob* __restrict__ probs) {
  const QrJob p = probs[blockIdx.y];
  if (!p.S) return;
  const int N = p.N, M = p.M;
  const int64_t mn = (int64_t)M * N;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < mn;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int j = (int)(e / M), i = (int)(e % M);
    const bool in = i < N;
    const double* s = p.S + 2 * ((int64_t)(in ? i : 0) + (int64_t)j * N);
    p.C[2 * e] = in ? s[0] : 0.0;
    p.C[2 * e + 1] = in ? s[1] : 0.0;
  }
}

// ---------------------------------------------------------------------------
// Column presort of the QR-preconditioned problems.  The first QR factors
// A P, P ordering the columns by decreasing norm (ties: lower index first);
// the Jacobi on the resulting triangle then needs fewer sweeps.  A P = U S W^T
// gives A = U S (P W)^T, so the right factor comes back with its rows
// scattered through P (row perm[j] of the result is row j of the computed
// factor).
// ---------------------------------------------------------------------------
struct SortJob {
  const double* A;  // M x N (ld M)
  double* dst;      // M x N: A P; later the N x N scratch of the row scatter
  double* key;      // N column norms^2
  int* perm;        // column j of A P is column perm[j] of A
  double* C;        // N x N (ld N) whose rows are scattered back through P
  int M, N;
  int cs;           // doubles per entry: 1 real, 2 interleaved complex
};
constexpr int kSortMaxN = 8192;  // keys staged in LDS

// key[j] = |A e_j|^2, one wave per column
__global__ __launch_bounds__(256) void sort_colnorm_kernel(const SortJob* __restrict__ jobs) {
  const SortJob p = jobs[blockIdx.x];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int jend = min(p.N, (int)(blockIdx.y + 1) * kFinCols);
  const int mc = p.cs * p.M;
  for (int j = blockIdx.y * kFinCols + wave; j < jend; j += 4) {
    const double* a = p.A + (int64_t)j * mc;
    double s = 0.0;
    for (int i = lane; i < mc; i += 64) s += a[i] * a[i];
    s = wave_sum(s);
    if (lane == 0) p.key[j] = s;
  }
}

// perm from ranks: column i goes to position #{k : key_k > key_i, or equal
// with k < i} (a total order, so the ranks are a permutation)
__global__ __launch_bounds__(1024) void sort_rank_kernel(const SortJob* __restrict__ jobs) {
  extern __shared__ double keys[];
  const SortJob p = jobs[blockIdx.x];
  for (int i = threadIdx.x; i < p.N; i += blockDim.x) keys[i] = p.key[i];
  __syncthreads();
  for (int i = threadIdx.x; i < p.N; i += blockDim.x) {
    const double ki = keys[i];
    int r = 0;
    for (int k = 0; k < p.N; ++k) {
      const double kk = keys[k];
      r += (kk > ki || (kk == ki && k < i)) ? 1 : 0;
    }
    p.perm[r] = i;
  }
}

// dst[:, j] = A[:, perm[j]]
__global__ __launch_bounds__(256) void sort_gather_kernel(const SortJob* __restrict__ jobs) {
  const SortJob p = jobs[blockIdx.x];
  const int c0 = blockIdx.y * kFinCols, c1 = min(p.N, c0 + kFinCols);
  const int mc = p.cs * p.M;
  const int64_t n = (int64_t)max(c1 - c0, 0) * mc;
  for (int64_t e = threadIdx.x; e < n; e += blockDim.x) {
    const int j = c0 + (int)(e / mc), i = (int)(e % mc);
    p.dst[e + (int64_t)c0 * mc] = p.A[i + (int64_t)p.perm[j] * mc];
  }
}

// dst[perm[i], j] = C[i, j]  (mode 0), then C = dst  (mode 1)
__global__ __launch_bounds__(256) void sort_scatter_rows_kernel(const SortJob* __restrict__ jobs, int mode) {
  const SortJob p = jobs[blockIdx.x];
  const int N = p.N, cs = p.cs;
  const int c0 = blockIdx.y * kFinCols, c1 = min(N, c0 + kFinCols);
  const int64_t n = (int64_t)max(c1 - c0, 0) * N * cs;
  for (int64_t e = threadIdx.x; e < n; e += blockDim.x) {
    const int64_t at = e + (int64_t)c0 * N * cs;  // double index
    if (mode == 0) {
      const int64_t en = at / cs;  // entry index
      const int part = (int)(at - en * cs);
      const int j = (int)(en / N), i = (int)(en % N);
      p.dst[(p.perm[i] + (int64_t)j * N) * cs + part] = p.C[at];
    } else {
      p.C[at] = p.dst[at];
    }
  }
}

}  // namespace

// ===========================================================================
// host wrappers
// ===========================================================================
void GemmBatch::begin_item(double* C, int64_t ldc, int M, int N, int accumulate, int flags, int key) {
  if (key >= 0 || !item_key.empty()) {
    item_key.resize(items.size(), -1);
    item_key.push_back(key);
  }
  GemmItem it{};
  it.flags = flags;
  it.C = C;
  it.ldc = ldc;
  it.M = M;
  it.N = N;
  it.c0 = it.c1 = (int)contribs.size();
  it.accumulate = accumulate;
  items.push_back(it);
}

void GemmBatch::add(const double* A, int64_t lda, const double* B, int64_t ldb, int K, int flags,
                    double scale) {
  require(!items.empty(), "GemmBatch::add without begin_item", kErrInternal);
  GemmContrib c{A, B, lda, ldb, K, flags, scale};
  contribs.push_back(c);
  items.back().c1 = (int)contribs.size();
  flops += 2.0 * items.back().M * (double)items.back().N * K;
}

namespace {
std::atomic<int> g_split_k{-1}, g_tile_mode{-1};
}  // namespace

int default_split_k() {
  const int g = g_split_k.load();
  if (g >= 0) return g;
  const char* e = std::getenv("DMRGCC_SPLIT_K");
  return e ? std::max(0, std::atoi(e)) : 256;
}

namespace {
std::atomic<int> g_grid_cap{-1};
}  // namespace

int default_grid_cap() {
  const int g = g_grid_cap.load();
  if (g >= 0) return g;
  const char* e = std::getenv("DMRGCC_GRID_CAP");
  return e ? std::max(0, std::atoi(e)) : 0;
}

void set_gemm_grid_cap(int cap) { g_grid_cap.store(cap < 0 ? -1 : cap); }

int default_tile_mode() {
  const int g = g_tile_mode.load();
  if (g >= 0) return g;
  const char* e = std::getenv("DMRGCC_TILE");
  if (!e) return 0;
  const int v = std::atoi(e);
  return (v == 32 || v == 64) ? v : 0;
}

void set_gemm_config(int tile_mode, int split_k) {
  g_tile_mode = tile_mode < 0 ? -1 : (tile_mode == 32 ? 32 : 64);
  g_split_k = split_k < 0 ? -1 : split_k;
}

void GemmBatch::finalize(hipStream_t stream, int tile_m, int tile_n) {
  const int64_t psize = layout(tile_m, tile_n);
  if (psize > 0) {
    part.ensure((size_t)psize);
    bind_parts(part.data());
  }
  // the four descriptor arrays in one device buffer and one upload (each
  // array 256-byte aligned)
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t bi = al(items.size() * sizeof(GemmItem)), bc = al(contribs.size() * sizeof(GemmContrib));
  const size_t bt = al(tiles.size() * sizeof(GemmTile)), bs = al(splits.size() * sizeof(SplitTask));
  const size_t total = bi + bc + bt + bs;
  std::vector<char> img(std::max<size_t>(total, 1));
  if (!items.empty()) std::memcpy(img.data(), items.data(), items.size() * sizeof(GemmItem));
  if (!contribs.empty()) std::memcpy(img.data() + bi, contribs.data(), contribs.size() * sizeof(GemmContrib));
  if (!tiles.empty()) std::memcpy(img.data() + bi + bc, tiles.data(), tiles.size() * sizeof(GemmTile));
  if (!splits.empty()) std::memcpy(img.data() + bi + bc + bt, splits.data(), splits.size() * sizeof(SplitTask));
  d_desc.ensure(std::max<size_t>(total, 8) / sizeof(double));
  if (total) upload_bytes(d_desc.data(), img.data(), total, stream);
  char* base = reinterpret_cast<char*>(d_desc.data());
  dev_items = reinterpret_cast<GemmItem*>(base);
  dev_contribs = reinterpret_cast<GemmContrib*>(base + bi);
  dev_tiles = reinterpret_cast<GemmTile*>(base + bi + bc);
  dev_splits = reinterpret_cast<SplitTask*>(base + bi + bc + bt);
}

void GemmBatch::bind_parts(double* base) {
  for (int i : part_items) items[i].C = base + (intptr_t)items[i].C;
  for (SplitTask& t : splits) t.part = base + (intptr_t)t.part;
  part_items.clear();
}

namespace {
size_t align256(size_t b) { return (b + 255) & ~size_t(255); }
}  // namespace

size_t GemmBatch::placed_bytes(int64_t psize) const {
  return align256(items.size() * sizeof(GemmItem)) + align256(contribs.size() * sizeof(GemmContrib)) +
         align256(tiles.size() * sizeof(GemmTile)) + align256(splits.size() * sizeof(SplitTask)) +
         align256((size_t)psize * sizeof(double));
}

void GemmBatch::place(void* arena, size_t cap, int64_t psize, hipStream_t stream) {
  const size_t need = placed_bytes(psize);
  require(need <= cap,
          "GEMM descriptors and split-K partials need " + std::to_string(need) + " bytes, the arena holds " +
              std::to_string(cap),
          kErrArgument);
  char* p = static_cast<char*>(arena);
  auto take = [&](size_t bytes) { char* q = p; p += align256(bytes); return q; };
  GemmItem* di = reinterpret_cast<GemmItem*>(take(items.size() * sizeof(GemmItem)));
  GemmContrib* dc = reinterpret_cast<GemmContrib*>(take(contribs.size() * sizeof(GemmContrib)));
  GemmTile* dt = reinterpret_cast<GemmTile*>(take(tiles.size() * sizeof(GemmTile)));
  SplitTask* ds = reinterpret_cast<SplitTask*>(take(splits.size() * sizeof(SplitTask)));
  double* dp = reinterpret_cast<double*>(take((size_t)psize * sizeof(double)));
  if (psize > 0) bind_parts(dp);
  // plain copies from the host vectors: no staging buffer, nothing allocated
  auto put = [&](void* dst, const void* src, size_t bytes) {
    if (bytes) DMRG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
  };
  put(di, items.data(), items.size() * sizeof(GemmItem));
  put(dc, contribs.data(), contribs.size() * sizeof(GemmContrib));
  put(dt, tiles.data(), tiles.size() * sizeof(GemmTile));
  put(ds, splits.data(), splits.size() * sizeof(SplitTask));
  dev_items = di;
  dev_contribs = dc;
  dev_tiles = dt;
  dev_splits = ds;
}

int64_t GemmBatch::layout(int tile_m, int tile_n) {
  tiles.clear();
  splits.clear();
  group_tiles.clear();
  part_items.clear();
  int64_t psize = 0;
  // placement keys of the items (split-K chunks take their item's)
  std::vector<int> key(items.size(), -1);
  for (size_t i = 0; i < item_key.size() && i < items.size(); ++i) key[i] = item_key[i];
  if (!group_items.empty() && group_items.back() != (int)items.size()) end_group();
  // split-K: chunk the contribution lists of heavy items (pre-split batches:
  // also every item with more than kPreTerms terms, whatever split_k)
  if ((split_k > 0 || presplit) && group_items.empty()) {
    std::vector<GemmItem> extra;
    struct Pending { size_t item; int first_extra, n; };
    std::vector<Pending> pend;
    std::vector<int> cuts;
    const size_t nitems = items.size();
    for (size_t i = 0; i < nitems; ++i) {
      GemmItem& it = items[i];
      int64_t ktot = 0;
      for (int c = it.c0; c < it.c1; ++c) ktot += contribs[c].K;
      const bool heavy = split_k > 0 && ktot > split_k;
      const bool many = presplit && it.c1 - it.c0 > kPreTerms;
      if (!(heavy || many) || it.c1 - it.c0 < 2) continue;
      // chunk boundaries by accumulated K (and term count)
      cuts.assign(1, it.c0);
      int64_t acc = 0;
      for (int c = it.c0; c < it.c1; ++c) {
        const bool full = (split_k > 0 && acc >= split_k) || (presplit && c - cuts.back() >= kPreTerms);
        if (full && c > cuts.back()) { cuts.push_back(c); acc = 0; }
        acc += contribs[c].K;
      }
      cuts.push_back(it.c1);
      const int nch = (int)cuts.size() - 1;
      if (nch < 2) continue;
      Pending pd{i, (int)extra.size(), nch - 1};
      for (int ch = 1; ch < nch; ++ch) {
        GemmItem e = it;
        e.c0 = cuts[ch];
        e.c1 = cuts[ch + 1];
        e.accumulate = 0;
        e.flags = 0;
        e.C = reinterpret_cast<double*>((intptr_t)psize);  // fixed up after allocation
        e.ldc = it.N;
        psize += (int64_t)it.M * it.N;
        extra.push_back(e);
      }
      it.c1 = cuts[1];
      pend.push_back(pd);
    }
    if (!extra.empty()) {
      // chunk outputs hold offsets into the partial buffer until
      // bind_parts() gives it an address
      const size_t base = items.size();
      items.reserve(base + extra.size());
      part_items.reserve(part_items.size() + extra.size());
      for (GemmItem& e : extra) {
        part_items.push_back((int)items.size());
        items.push_back(e);
      }
      key.resize(items.size(), -1);
      for (const Pending& pd : pend)
        for (int e = 0; e < pd.n; ++e) key[base + pd.first_extra + e] = key[pd.item];
      for (const Pending& pd : pend) {
        const GemmItem& it = items[pd.item];
        splits.push_back(SplitTask{it.C, it.ldc, it.M, it.N, pd.n, it.flags,
                                   items[base + pd.first_extra].C});
      }
    }
  }
  // Largest work first (within each group) so the long items start early.
  std::vector<double> cost(items.size(), 0.0);
  for (size_t i = 0; i < items.size(); ++i) {
    double k = 0;
    for (int c = items[i].c0; c < items[i].c1; ++c) k += contribs[c].K;
    cost[i] = k;
  }
  // XCD-aware order: workgroup w of a launch runs on XCD w % kXcds, and each
  // XCD has its own L2.  Every item's tiles go to one XCD queue (greedy by
  // cost, largest first) so the tiles that share A rows and B columns share
  // an L2; the queues are interleaved into the launch order, padded with
  // empty tiles (item -1) where a queue runs short.
  // Tile shape: 64x64 workgroup tiles, or 32x32 wave tiles (four per
  // workgroup, DMRGCC_TILE = 32).  Wave tiles waste less of each padded tile
  // on small blocks but read twice the operand bytes per flop; on the
  // H_eff launches measured here the 64x64 form is the faster (BF16x3:
  // 0.105 vs 0.158 ms per C2 matvec), so it is the default.
  wave_tiles = group_items.empty() && tile_m == kTile && tile_n == kTile && tile_mode == 32;
  // With placement keys: the items of one key form a group whose tiles all go
  // to one XCD queue; groups (and unkeyed items) to the least-loaded queue,
  // costliest first.
  const bool keyed = tile_m == tile_n && std::any_of(key.begin(), key.end(), [](int k) { return k >= 0; });
  // costliest first, ties in item order (the order a stable sort gives)
  auto by_cost = [&](int a, int b) { return cost[a] > cost[b] || (cost[a] == cost[b] && a < b); };
  auto live = [&](const GemmItem& it) { return it.M > 0 && it.N > 0 && !(it.c0 == it.c1 && it.accumulate); };
  auto ntiles = [](const GemmItem& it, int ts) { return ((it.M + ts - 1) / ts) * ((it.N + ts - 1) / ts); };
  // The items of one launch group in queue order: qx[n] is the XCD queue of
  // the n-th item of `seq`; the tiles are then laid out queue by queue from
  // counts, without growing vectors.
  std::vector<int> seq, qx, qoff(kXcds + 1), qfill(kXcds);
  std::vector<GemmTile> qt;
  auto assign = [&](const std::vector<int>& order, int ts) {
    seq.clear();
    qx.clear();
    std::vector<double> load(kXcds, 0.0);
    auto least = [&] { return (int)(std::min_element(load.begin(), load.end()) - load.begin()); };
    if (keyed) {
      // items of one key form a group whose tiles all go to one queue; groups
      // (and unkeyed items) to the least-loaded queue, costliest first
      std::unordered_map<int, int> kg;
      std::vector<int> gfirst, gnext, glast;
      std::vector<double> gcost;
      std::vector<int> nextof(items.size(), -1);
      for (int i : order) {
        const GemmItem& it = items[i];
        if (!live(it)) continue;
        int g = (int)gfirst.size();
        if (key[i] >= 0) {
          auto f = kg.find(key[i]);
          if (f != kg.end()) g = f->second;
          else kg[key[i]] = g;
        }
        if (g == (int)gfirst.size()) {
          gfirst.push_back(i);
          glast.push_back(i);
          gcost.push_back(0.0);
        } else {
          nextof[glast[g]] = i;
          glast[g] = i;
        }
        gcost[g] += (cost[i] + 32.0) * (double)ntiles(it, ts);
      }
      std::vector<int> gorder(gfirst.size());
      std::iota(gorder.begin(), gorder.end(), 0);
      std::sort(gorder.begin(), gorder.end(),
                [&](int x, int y) { return gcost[x] > gcost[y] || (gcost[x] == gcost[y] && x < y); });
      for (int g : gorder) {
        const int x = least();
        for (int i = gfirst[g]; i >= 0; i = nextof[i]) {
          seq.push_back(i);
          qx.push_back(x);
        }
        load[x] += gcost[g];
      }
    } else {
      for (int i : order) {
        const GemmItem& it = items[i];
        if (!live(it)) continue;
        const int x = least();
        seq.push_back(i);
        qx.push_back(x);
        load[x] += (cost[i] + 32.0) * ntiles(it, ts);
      }
    }
    // tiles per queue, then each queue's tiles in item order
    std::fill(qoff.begin(), qoff.end(), 0);
    for (size_t n = 0; n < seq.size(); ++n) qoff[qx[n] + 1] += ntiles(items[seq[n]], ts);
    for (int x = 0; x < kXcds; ++x) qoff[x + 1] += qoff[x];
    qt.resize((size_t)qoff[kXcds]);
    for (int x = 0; x < kXcds; ++x) qfill[x] = qoff[x];
    for (size_t n = 0; n < seq.size(); ++n) {
      const GemmItem& it = items[seq[n]];
      GemmTile* d = qt.data() + qfill[qx[n]];
      int c = 0;
      for (int m0 = 0; m0 < it.M; m0 += ts)
        for (int n0 = 0; n0 < it.N; n0 += ts) d[c++] = GemmTile{seq[n], m0, n0, 0};
      qfill[qx[n]] += c;
    }
  };
  auto qsize = [&](int x) { return (size_t)(qoff[x + 1] - qoff[x]); };
  auto qat = [&](int x, size_t k) { return qt[(size_t)qoff[x] + k]; };
  std::vector<int> ends = group_items;
  if (ends.empty()) ends.push_back((int)items.size());
  int first = 0;
  if (wave_tiles) {
    // 32x32 tiles in per-XCD queues as below, cut into runs of four (one
    // workgroup); workgroup w runs on XCD w % kXcds
    std::vector<int> order(items.size());
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), by_cost);
    assign(order, kWT);
    size_t depth = 0;
    for (int x = 0; x < kXcds; ++x) depth = std::max(depth, (qsize(x) + 3) / 4);
    tiles.resize(depth * kXcds * 4);
    size_t o = 0;
    for (size_t k = 0; k < depth; ++k)
      for (int x = 0; x < kXcds; ++x)
        for (int w = 0; w < 4; ++w) tiles[o++] = 4 * k + w < qsize(x) ? qat(x, 4 * k + w) : GemmTile{-1, 0, 0, 0};
    while (tiles.size() >= 4 && tiles[tiles.size() - 1].item < 0 && tiles[tiles.size() - 2].item < 0 &&
           tiles[tiles.size() - 3].item < 0 && tiles[tiles.size() - 4].item < 0)
      tiles.resize(tiles.size() - 4);
    group_tiles.push_back((int)tiles.size());
    ends.clear();
  }
  for (int end : ends) {
    std::vector<int> order;
    order.reserve((size_t)std::max(end - first, 0));
    for (int i = first; i < end; ++i) order.push_back(i);
    std::sort(order.begin(), order.end(), by_cost);
    assign(order, tile_m);
    size_t depth = 0;
    for (int x = 0; x < kXcds; ++x) depth = std::max(depth, qsize(x));
    const size_t t0 = tiles.size();
    tiles.resize(t0 + depth * kXcds);
    size_t o = t0;
    for (size_t k = 0; k < depth; ++k)
      for (int x = 0; x < kXcds; ++x) tiles[o++] = k < qsize(x) ? qat(x, k) : GemmTile{-1, 0, 0, 0};
    // trailing empty tiles of the last row are not launched
    while (!tiles.empty() && tiles.back().item < 0 && (int)tiles.size() > (group_tiles.empty() ? 0 : group_tiles.back()))
      tiles.pop_back();
    group_tiles.push_back((int)tiles.size());
    first = end;
  }
  return psize;
}

namespace {
// one launch of the 64x64 (launch_tile) or wave-tile (launch_wave) kernels;
// kLoop: the persistent form (each workgroup walks tiles with a stride of the
// grid)
template <bool kLoop>
void launch_tile(int precision, dim3 grid, dim3 block, hipStream_t stream, const GemmTile* t, const GemmItem* it,
                 const GemmContrib* c, const double* in, double* out, int n) {
  if (precision == kFP64)
    hipLaunchKernelGGL(HIP_KERNEL_NAME(gemm_mfma_kernel<double, kLoop>), grid, block, 0, stream, t, it, c, in, out,
                       n);
  else if (precision == kFP32)
    hipLaunchKernelGGL(HIP_KERNEL_NAME(gemm_mfma_kernel<float, kLoop>), grid, block, 0, stream, t, it, c, in, out,
                       n);
  else
    hipLaunchKernelGGL(HIP_KERNEL_NAME(gemm_bf16x3_kernel<kLoop>), grid, block, 0, stream, t, it, c, in, out, n);
}
template <bool kLoop>
void launch_wave(int precision, dim3 grid, dim3 block, hipStream_t stream, const GemmTile* t, const GemmItem* it,
                 const GemmContrib* c, const double* in, double* out, int n) {
  if (precision == kFP64)
    hipLaunchKernelGGL(HIP_KERNEL_NAME(gemm_wave_kernel<double, kLoop>), grid, block, 0, stream, t, it, c, in, out,
                       n);
  else if (precision == kFP32)
    hipLaunchKernelGGL(HIP_KERNEL_NAME(gemm_wave_kernel<float, kLoop>), grid, block, 0, stream, t, it, c, in, out,
                       n);
  else
    hipLaunchKernelGGL(HIP_KERNEL_NAME(gemm_bf16x3_wave_kernel<kLoop>), grid, block, 0, stream, t, it, c, in, out,
                       n);
}
}  // namespace

unsigned GemmBatch::launch_grid(int units) const {
  // a persistent grid of at most grid_cap workgroups (a multiple of the XCD
  // count, so that tile t stays on XCD t % 8) walks the tile list with a
  // stride of the grid; grid_cap 0 launches one workgroup per tile
  if (grid_cap <= 0 || units <= grid_cap) return (unsigned)std::max(units, 1);
  return (unsigned)std::max(kXcds, grid_cap - grid_cap % kXcds);
}

void GemmBatch::launch(hipStream_t stream, int precision, const double* in, double* out,
                       bool reduce) const {
  if (tiles.empty()) return;
  require(precision >= kFP64 && precision <= kBF16x3, "unknown precision", kErrArgument);
  const dim3 block(kThreads);
  if (presplit) {
    require(precision == kBF16x3 && !wave_tiles, "pre-split batches are BF16x3 on 64x64 tiles", kErrInternal);
    hipLaunchKernelGGL(gemm_bf16x3_pre_kernel, dim3((unsigned)tiles.size()), block, 0, stream, dev_tiles,
                       dev_items, dev_contribs, in, out);
  } else if (wave_tiles) {
    const int ng = (int)(tiles.size() / 4);
    const dim3 grid(launch_grid(ng));
    if ((int)grid.x < ng)
      launch_wave<true>(precision, grid, block, stream, dev_tiles, dev_items, dev_contribs, in, out, ng);
    else
      launch_wave<false>(precision, grid, block, stream, dev_tiles, dev_items, dev_contribs, in, out, ng);
  } else {
    const int nt = (int)tiles.size();
    const dim3 grid(launch_grid(nt));
    if ((int)grid.x < nt)
      launch_tile<true>(precision, grid, block, stream, dev_tiles, dev_items, dev_contribs, in, out, nt);
    else
      launch_tile<false>(precision, grid, block, stream, dev_tiles, dev_items, dev_contribs, in, out, nt);
  }
  DMRG_HIP(hipGetLastError());
  if (reduce && !splits.empty()) {
    int64_t mx = 0;
    for (const SplitTask& t : splits) mx = std::max(mx, (int64_t)t.M * t.N);
    const unsigned gx = (unsigned)std::max<int64_t>(1, (mx + 256 * kSplitEl - 1) / (256 * kSplitEl));
    hipLaunchKernelGGL(split_reduce_kernel, dim3(gx, (unsigned)splits.size()), dim3(256), 0, stream,
                       dev_splits, out);
    DMRG_HIP(hipGetLastError());
  }
}

void GemmBatch::merge(const std::vector<const GemmBatch*>& bs, const std::vector<const double*>& ins,
                      const std::vector<double*>& outs, hipStream_t stream) {
  require(bs.size() == ins.size() && bs.size() == outs.size(), "merge: one input and output per batch",
          kErrArgument);
  clear();
  std::vector<SplitTask> sp;
  for (size_t i = 0; i < bs.size(); ++i) {
    const GemmBatch& b = *bs[i];
    require(b.group_items.empty(), "merge: grouped batches cannot be merged", kErrArgument);
    const int cbase = (int)contribs.size();
    for (GemmContrib c : b.contribs) {
      if (c.flags & kRelA) { c.A = ins[i] + (intptr_t)c.A; c.flags &= ~kRelA; }
      if (c.flags & kRelB) { c.B = ins[i] + (intptr_t)c.B; c.flags &= ~kRelB; }
      contribs.push_back(c);
    }
    for (GemmItem it : b.items) {
      if (it.flags & kRelC) { it.C = outs[i] + (intptr_t)it.C; it.flags &= ~kRelC; }
      it.c0 += cbase;
      it.c1 += cbase;
      items.push_back(it);
    }
    for (SplitTask t : b.splits) {
      if (t.flags & kRelC) { t.C = outs[i] + (intptr_t)t.C; t.flags &= ~kRelC; }
      sp.push_back(t);
    }
    flops += b.flops;
  }
  const int sk = split_k;
  split_k = 0;  // the sources are split already
  finalize(stream);
  split_k = sk;
  splits = sp;
  if (!splits.empty()) upload(d_splits, splits, stream);
  dev_splits = d_splits.data();
}

void GemmBatch::launch_group(hipStream_t stream, int group, int precision) const {
  require(group >= 0 && group < (int)group_tiles.size(), "GemmBatch: no such group", kErrInternal);
  const int t0 = group == 0 ? 0 : group_tiles[group - 1], t1 = group_tiles[group];
  if (t1 <= t0) return;
  const dim3 grid((unsigned)(t1 - t0)), block(kThreads);
  const int nt = t1 - t0;
  const GemmTile* tl = dev_tiles + t0;
  if (precision == kFP64)
    hipLaunchKernelGGL(gemm_mfma_kernel<double>, grid, block, 0, stream, tl, dev_items,
                       dev_contribs, nullptr, nullptr, nt);
  else if (precision == kFP32)
    hipLaunchKernelGGL(gemm_mfma_kernel<float>, grid, block, 0, stream, tl, dev_items,
                       dev_contribs, nullptr, nullptr, nt);
  else
    hipLaunchKernelGGL(gemm_bf16x3_kernel<false>, grid, block, 0, stream, tl, dev_items,
                       dev_contribs, nullptr, nullptr, nt);
  DMRG_HIP(hipGetLastError());
}

namespace {
// complex element offset -> double offset of part p (0 re, 1 im), for real
// or relative (offset-encoded) pointers alike
template <typename P>
P* part_ptr(P* ptr, int p, bool relative) {
  if (relative) return reinterpret_cast<P*>(2 * (intptr_t)ptr + p);
  return ptr + p;
}
}  // namespace

void zgemm_item(GemmBatch& g, double* C, int64_t ldc, int M, int N, int accumulate, int cflags,
                const std::vector<ZTerm>& terms) {
  const bool relc = cflags & kRelC;
  for (int part = 0; part < 2; ++part) {
    g.begin_item(part_ptr(C, part, relc), 2 * ldc, M, N, accumulate, cflags | kStrided2C);
    for (const ZTerm& t : terms) {
      const double sa = t.conjA ? -1.0 : 1.0, sb = t.conjB ? -1.0 : 1.0;
      // (A part, B part, scale): P = A'B' with A' = Ar + i sa Ai, B' = Br + i sb Bi
      struct R { int pa, pb; double s; };
      R rs[4];
      if (part == 0)
        rs[0] = {0, 0, t.sr}, rs[1] = {1, 1, -t.sr * sa * sb}, rs[2] = {0, 1, -t.si * sb},
        rs[3] = {1, 0, -t.si * sa};
      else
        rs[0] = {0, 1, t.sr * sb}, rs[1] = {1, 0, t.sr * sa}, rs[2] = {0, 0, t.si},
        rs[3] = {1, 1, -t.si * sa * sb};
      for (const R& r : rs) {
        if (r.s == 0.0) continue;
        g.add(part_ptr(t.A, r.pa, t.flags & kRelA), 2 * t.lda, part_ptr(t.B, r.pb, t.flags & kRelB),
              2 * t.ldb, t.K, t.flags | kStrided2, r.s);
      }
    }
  }
}

void launch_copies(const std::vector<CopyTask>& tasks, DevBuf<CopyTask>& scratch,
                   hipStream_t stream) {
  if (tasks.empty()) return;
  // split tasks into row chunks of <= 8K elements (a single row may be
  // longer), one workgroup each
  std::vector<CopyTask> split;
  split.reserve(tasks.size());
  for (const CopyTask& t : tasks) {
    if (t.rows <= 0 || t.cols <= 0) continue;
    require(t.cols < (1 << 30), "copy: row too long", kErrArgument);
    int rows_per = std::max(1, (int)(8192 / std::max(1, t.cols)));
    for (int r0 = 0; r0 < t.rows; r0 += rows_per) {
      CopyTask c = t;
      c.rows = std::min(rows_per, t.rows - r0);
      if (c.src) c.src = t.src + r0 * t.ssr;
      c.dst = t.dst + r0 * t.dsr;
      split.push_back(c);
    }
  }
  if (split.empty()) return;
  upload(scratch, split, stream);
  hipLaunchKernelGGL(copy_kernel, dim3((unsigned)split.size()), dim3(256), 0, stream,
                     scratch.data());
  DMRG_HIP(hipGetLastError());
  // the descriptor buffer may be reused by the next call: wait for the kernel
  DMRG_HIP(hipStreamSynchronize(stream));
}

void launch_pack(const PackTask* dev_tasks, int ntasks, int64_t max_groups, const double* base,
                 hipStream_t stream) {
  if (ntasks <= 0) return;
  const unsigned gy = (unsigned)std::max<int64_t>(1, std::min<int64_t>(64, (max_groups + 255) / 256));
  hipLaunchKernelGGL(pack_pairs_kernel, dim3((unsigned)ntasks, gy), dim3(256), 0, stream, dev_tasks, base);
  DMRG_HIP(hipGetLastError());
}

void dot(const double* x, const double* y, int64_t n, double* out, DevBuf<double>& partial,
         hipStream_t stream) {
  partial.ensure(kDotBlocks);
  const int g = (int)std::max<int64_t>(1, std::min<int64_t>(kDotBlocks, (n + 255) / 256));
  hipLaunchKernelGGL(dot_partial_kernel, dim3(g), dim3(256), 0, stream, x, y, n, partial.data());
  hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(256), 0, stream, partial.data(), g, out);
  DMRG_HIP(hipGetLastError());
}

void zdot(const double* x, const double* y, int64_t n, double* out, DevBuf<double>& partial,
          hipStream_t stream) {
  partial.ensure(2 * kDotBlocks);
  const int g = (int)std::max<int64_t>(1, std::min<int64_t>(kDotBlocks, (n + 255) / 256));
  hipLaunchKernelGGL(zdot_partial_kernel, dim3(g), dim3(256), 0, stream, x, y, n, partial.data());
  hipLaunchKernelGGL(zdot_final_kernel, dim3(1), dim3(256), 0, stream, partial.data(), g, out);
  DMRG_HIP(hipGetLastError());
}

void zaxpy_dev(double* y, const double* x, int64_t n, const double* alpha, double sign,
               hipStream_t stream) {
  hipLaunchKernelGGL(zaxpy_dev_kernel, dim3(grid_for(n)), dim3(256), 0, stream, y, x, n, alpha, sign);
  DMRG_HIP(hipGetLastError());
}

void axpy_dev(double* y, const double* x, int64_t n, const double* alpha, double sign,
              hipStream_t stream) {
  hipLaunchKernelGGL(axpy_dev_kernel, dim3(grid_for(n)), dim3(256), 0, stream, y, x, n, alpha, sign);
  DMRG_HIP(hipGetLastError());
}

void axpby(double* y, const double* x, int64_t n, double a, double b, hipStream_t stream) {
  hipLaunchKernelGGL(axpby_kernel, dim3(grid_for(n)), dim3(256), 0, stream, y, x, n, a, b);
  DMRG_HIP(hipGetLastError());
}

void scale_inv_sqrt(double* x, int64_t n, const double* nrm2, hipStream_t stream, int mode) {
  if (n <= 0) return;
  hipLaunchKernelGGL(scale_inv_sqrt_kernel, dim3(grid_for(n)), dim3(256), 0, stream, x, n, nrm2, mode);
  DMRG_HIP(hipGetLastError());
}

int dot_grid(int64_t n) {
  return (int)std::max<int64_t>(1, std::min<int64_t>(kDotBlocks, (n + 255) / 256));
}

void lz_dot_partial(const double* x, const double* y, int64_t n, bool cplx, double* part,
                    hipStream_t stream) {
  const int g = dot_grid(n);
  if (cplx) hipLaunchKernelGGL(lz_dot_partial_kernel<true>, dim3(g), dim3(256), 0, stream, x, y, n, part);
  else hipLaunchKernelGGL(lz_dot_partial_kernel<false>, dim3(g), dim3(256), 0, stream, x, y, n, part);
  DMRG_HIP(hipGetLastError());
}

void lz_finals(const double* part_a, const double* part_b, int64_t n, bool cplx, double* out,
               hipStream_t stream, double* out_host) {
  const int g = dot_grid(n);
  if (cplx)
    hipLaunchKernelGGL(lz_finals_kernel<true>, dim3(1), dim3(256), 0, stream, part_a, part_b, g, out, out_host);
  else
    hipLaunchKernelGGL(lz_finals_kernel<false>, dim3(1), dim3(256), 0, stream, part_a, part_b, g, out, out_host);
  DMRG_HIP(hipGetLastError());
}

void lz_alpha_multi(double* h, const double* psi, const double* x, int64_t ld, int nj, int64_t n, bool cplx,
                    const double* part_a, double* part_b, double* part_m, hipStream_t stream) {
  if (n <= 0) return;
  const int g = dot_grid(n);
  if (cplx)
    hipLaunchKernelGGL(lz_alpha_multi_kernel<true>, dim3(g), dim3(256), 0, stream, h, psi, x, ld, nj, n, part_a,
                       part_b, part_m);
  else
    hipLaunchKernelGGL(lz_alpha_multi_kernel<false>, dim3(g), dim3(256), 0, stream, h, psi, x, ld, nj, n, part_a,
                       part_b, part_m);
  DMRG_HIP(hipGetLastError());
}

void lz_next_all(double* nxt, double* h, const double* psi, const double* x, int64_t ld, int nj, int64_t n,
                 bool cplx, const double* part_m, hipStream_t stream, const double* beta2, const double* part_a,
                 const double* part_b, double* slot, double* slot_host) {
  if (n <= 0) return;
  require(nj >= 1, "lz_next_all: needs at least one basis vector", kErrInternal);
  require((part_b && (!slot || (part_a && slot_host))) || (!part_b && beta2),
          "lz_next_all: beta needs the beta^2 partials or the device value; publishing needs alpha's partials "
          "and both slots",
          kErrInternal);
  const int g = dot_grid(n);
  const size_t lds = (size_t)(cplx ? 2 : 1) * nj * sizeof(double);
  if (cplx)
    hipLaunchKernelGGL(lz_next_all_kernel<true>, dim3(g), dim3(256), lds, stream, nxt, h, psi, x, ld, nj, n, part_m,
                       part_a, part_b, beta2, slot, slot_host);
  else
    hipLaunchKernelGGL(lz_next_all_kernel<false>, dim3(g), dim3(256), lds, stream, nxt, h, psi, x, ld, nj, n,
                       part_m, part_a, part_b, beta2, slot, slot_host);
  DMRG_HIP(hipGetLastError());
}

void lanczos_step(double* psi_next, double* h, const double* psi, int64_t n, double beta,
                  hipStream_t stream) {
  hipLaunchKernelGGL(lanczos_step_kernel, dim3(grid_for(n)), dim3(256), 0, stream, psi_next, h, psi,
                     n, beta);
  DMRG_HIP(hipGetLastError());
}

namespace {
// Lane c: its chunk's window st[t] = sum_j coef[c][j] w[j + t] (the 31
// latest raw values, oldest first), then the recurrence r = r[-31] + r[-3]
// unrolled over 124 steps (st[u % 31] += st[(u + 28) % 31]), four draws per
// 32-bit store.
__global__ __launch_bounds__(256) void rng_stream_kernel(const uint32_t* __restrict__ w0,
                                                          const uint32_t* __restrict__ coef, int64_t nchunks,
                                                          uint32_t* __restrict__ out) {
  __shared__ uint32_t w[61];
  if (threadIdx.x < 61) w[threadIdx.x] = w0[threadIdx.x];
  __syncthreads();
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nchunks) return;
  uint32_t st[31];
  {
    uint32_t cf[31];
#pragma unroll
    for (int j = 0; j < 31; ++j) cf[j] = coef[31 * c + j];
#pragma unroll
    for (int t = 0; t < 31; ++t) {
      uint32_t v = 0;
#pragma unroll
      for (int j = 0; j < 31; ++j) v += cf[j] * w[j + t];
      st[t] = v;
    }
  }
  uint32_t* o = out + c * (kRngChunk / 4);
  for (int run = 0; run < kRngChunk / 124; ++run) {
#pragma unroll
    for (int q = 0; q < 31; ++q) {
      uint32_t word = 0;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int u = (4 * q + b) % 31;
        st[u] += st[(u + 28) % 31];
        word |= ((st[u] >> 1) % 100u) << (8 * b);
      }
      o[31 * run + q] = word;
    }
  }
}

__global__ __launch_bounds__(256) void rng_layout_kernel(const RngDraw* __restrict__ draws,
                                                          const uint8_t* __restrict__ raw,
                                                          double* __restrict__ R, int cs) {
  const RngDraw d = draws[blockIdx.y];
  const int64_t n = (int64_t)d.rows * d.cols;
  double* out = R + d.roff;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / d.cols), j = (int)(e % d.cols);
    const int64_t k = (int64_t)j * d.rows + i;  // its draw: column-major fill
    const bool dg = d.diag && i == j;
    if (cs == 1) {
      out[e] = dg ? 0.1 * d.w : 0.01 * (double)raw[d.first + k];
    } else {
      out[2 * e] = dg ? 0.1 * d.w : 0.01 * (double)raw[d.first + 2 * k + 1];
      out[2 * e + 1] = dg ? 0.1 * d.wi : 0.01 * (double)raw[d.first + 2 * k];
    }
  }
}
}  // namespace

void rng_stream(const uint32_t* w0, const uint32_t* coef, int64_t total, uint8_t* out, hipStream_t stream) {
  if (total <= 0) return;
  const int64_t nchunks = (total + kRngChunk - 1) / kRngChunk;
  hipLaunchKernelGGL(rng_stream_kernel, dim3((unsigned)((nchunks + 255) / 256)), dim3(256), 0, stream, w0, coef,
                     nchunks, reinterpret_cast<uint32_t*>(out));
  DMRG_HIP(hipGetLastError());
}

void rng_layout(const RngDraw* draws, int ndraws, int64_t max_entries, const uint8_t* raw, double* R, int cs,
                hipStream_t stream) {
  if (ndraws <= 0 || max_entries <= 0) return;
  const unsigned gx = (unsigned)std::min<int64_t>(64, (max_entries + 255) / 256);
  for (int d0 = 0; d0 < ndraws; d0 += 65535) {
    const unsigned gy = (unsigned)std::min(65535, ndraws - d0);
    hipLaunchKernelGGL(rng_layout_kernel, dim3(gx, gy), dim3(256), 0, stream, draws + d0, raw, R, cs);
    DMRG_HIP(hipGetLastError());
  }
}

void launch_col_gather(const std::vector<ColGatherTask>& tasks, DevBuf<ColGatherTask>& scratch,
                       hipStream_t stream) {
  std::vector<ColGatherTask> live;
  for (const ColGatherTask& t : tasks)
    if (t.rows > 0 && t.ncol > 0) live.push_back(t);
  if (live.empty()) return;
  int64_t most = 0;
  for (const ColGatherTask& t : live) {
    const int64_t tot = (int64_t)t.rows * t.ncol;
    require(tot < (int64_t)1 << 31, "col_gather: block too large", kErrArgument);
    most = std::max(most, tot);
  }
  upload(scratch, live, stream);
  const unsigned gy = (unsigned)std::max<int64_t>(1, std::min<int64_t>(64, (most + 4095) / 4096));
  hipLaunchKernelGGL(col_gather_kernel, dim3((unsigned)live.size(), gy), dim3(256), 0, stream,
                     scratch.data());
  DMRG_HIP(hipGetLastError());
  DMRG_HIP(hipStreamSynchronize(stream));
}

void combine(double* y, const double* V, int64_t stride, const double* coef, int k, int64_t n,
             hipStream_t stream) {
  hipLaunchKernelGGL(combine_kernel, dim3(grid_for(n)), dim3(256), 0, stream, y, V, stride, coef, k, n);
  DMRG_HIP(hipGetLastError());
}

void fill(double* x, int64_t n, double v, hipStream_t stream) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n)), dim3(256), 0, stream, x, n, v);
  DMRG_HIP(hipGetLastError());
}

namespace {
constexpr int kQrMinN = 48;  // smaller problems are not worth the QR launches

// Pinned host memory for the per-sweep flags, one buffer per host thread (a
// driver call runs on one thread and waits for each of its sweeps).
int* sweep_host_flags(int np, int** dev) {
  struct Buf {
    int* host = nullptr;
    int* dev = nullptr;
    int cap = 0;
  };
  thread_local Buf b;
  if (np > b.cap) {
    if (b.host) DMRG_HIP(hipHostFree(b.host));
    b.cap = std::max(np, 256);
    DMRG_HIP(hipHostMalloc((void**)&b.host, (size_t)b.cap * sizeof(int), hipHostMallocDefault));
    DMRG_HIP(hipHostGetDevicePointer((void**)&b.dev, b.host, 0));
  }
  *dev = b.dev;
  return b.host;
}

// The sweeps of a batched Jacobi.  One visit map serves the whole call (every
// problem with N > 1, kb-column blocks, a visit per block pair); a sweep is
// rounds 0 .. R-1 (R: the most rounds of a still active problem) and a flush
// launch, whose extra workgroup leaves the per-problem rotation flags in
// pinned host memory and makes them the next sweep's active set, so the host
// only waits for the sweep and reads them.  launch(grid, round, ctl, vprob,
// vpair, qflag, visits) issues one round launch.  Returns the sweeps run;
// last[i] = the last sweep problem i took part in.
template <typename Launch>
int run_sweeps(const std::vector<SvdProblem>& probs, int kb, DevBuf<int>& ctl_buf, hipStream_t stream,
               int max_sweeps, std::vector<int>& last, int& visits_out, Launch launch) {
  const int np = (int)probs.size();
  std::vector<int> active, nrounds(np, 0), map;
  std::vector<int> pairs;
  for (int i = 0; i < np; ++i)
    if (probs[i].N > 1) {
      active.push_back(i);
      const int nb = (probs[i].N + kb - 1) / kb, nb2 = nb + (nb & 1);
      nrounds[i] = nb2 - 1;
      for (int pr = 0; pr < nb2 / 2; ++pr) {
        map.push_back(i);
        pairs.push_back(pr);
      }
    }
  const int visits = (int)pairs.size();
  visits_out = visits;
  last.assign(np, 0);
  if (active.empty()) return 0;
  // device ints: active[np] | rotated[np] | vprob[visits] | vpair[visits] |
  // qflag[2 visits]
  ctl_buf.ensure(2 * (size_t)np + 4 * (size_t)visits);
  std::vector<int> h(2 * (size_t)np, 0);
  for (int i : active) h[i] = 1;
  h.insert(h.end(), map.begin(), map.end());
  h.insert(h.end(), pairs.begin(), pairs.end());
  upload_bytes(ctl_buf.data(), h.data(), h.size() * sizeof(int), stream);
  SweepCtl ctl{ctl_buf.data() + np, ctl_buf.data(), nullptr, np};
  int* hflags = sweep_host_flags(np, &ctl.host);
  const int* vprob = ctl_buf.data() + 2 * np;
  const int* vpair = vprob + visits;
  int* qflag = ctl_buf.data() + 2 * np + 2 * visits;
  int sweeps = 0;
  for (int sweep = 0; sweep < max_sweeps && !active.empty(); ++sweep) {
    ++sweeps;
    int rounds = 0;
    for (int i : active) {
      rounds = std::max(rounds, nrounds[i]);
      last[i] = sweep + 1;
    }
    for (int r = 0; r < rounds; ++r) launch(dim3(2 * visits), r, ctl, vprob, vpair, qflag, visits);
    launch(dim3(2 * visits + 1), rounds, ctl, vprob, vpair, qflag, visits);  // flush, sweep end
    DMRG_HIP(hipGetLastError());
    DMRG_HIP(hipStreamSynchronize(stream));
    std::vector<int> still;
    for (int i : active)
      if (hflags[i]) still.push_back(i);
    active.swap(still);
  }
  return sweeps;
}

// One batched blocked-Householder QR over a set of jobs: per panel k the
// panel kernel, then the trailing update Q_k^H A (qr_update_kernel or its
// complex twin); the back-transformation C <- Q_1 ... Q_npan C applies the
// panels in reverse.
struct QrPhase {
  std::vector<QrJob> jobs;
  DevBuf<QrJob> d_jobs;
  DevBuf<double> work;
  int b = 32, npan = 0, maxN = 0, maxM = 0, reg_u = 0, cs = 1;
  bool lds = true;
  size_t lds_bytes = 0;

  // jobs: every field but Y and P set by the caller (X included); the plan
  // adds the Y and P buffers.  cs = 2: interleaved complex jobs.
  void plan(std::vector<QrJob> in, hipStream_t stream, int cs_ = 1) {
    jobs = std::move(in);
    cs = cs_;
    maxM = maxN = 0;
    for (const QrJob& j : jobs) { maxM = std::max(maxM, j.M); maxN = std::max(maxN, j.N); }
    // widest panel whose LDS copy fits (cs (m*b + b*b + b) doubles plus the static bytes)
    const size_t budget = 160 * 1024 - 256;
    lds = false;
    reg_u = 0;
    if (cs == 2) {
      for (int bb : {kZqrB, 8}) {
        const size_t need = 2 * ((size_t)maxM * bb + (size_t)bb * bb + bb) * sizeof(double);
        if (need <= budget) { b = bb; lds = true; lds_bytes = need; break; }
      }
      if (!lds) { b = kZqrB; lds_bytes = 2 * ((size_t)b * b + b) * sizeof(double); }
    } else {
      for (int bb : {32, 16, 8}) {
        const size_t need = ((size_t)maxM * bb + (size_t)bb * bb + bb) * sizeof(double);
        if (need <= budget) { b = bb; lds = true; lds_bytes = need; break; }
      }
      if (!lds) { b = 32; lds_bytes = ((size_t)b * b + b) * sizeof(double); }
      // register column loop (panel columns padded to 64 kU rows): 32-column
      // panels up to 576 rows, 16-column panels up to 1024 rows (16 waves, a
      // column each), 8-column panels up to 2048 rows (8 waves)
      if (lds && b == 32 && maxM <= 64 * 9) reg_u = maxM <= 64 * 4 ? 4 : 9;
      else if (lds && b == 16 && maxM <= 64 * 16) reg_u = 16;
      else if (lds && b == 8 && maxM <= 64 * 32) reg_u = 32;
      if (reg_u) lds_bytes = ((size_t)64 * reg_u * b + (size_t)b * b + b) * sizeof(double);
    }
    npan = (maxN + b - 1) / b;
    int64_t total = 0;
    std::vector<int64_t> off;
    for (const QrJob& j : jobs) {
      const int np = (j.N + b - 1) / b;
      off.push_back(total);
      total += 2 * (int64_t)cs * np * b * j.M;
    }
    work.ensure((size_t)std::max<int64_t>(total, 1));
    for (size_t t = 0; t < jobs.size(); ++t) {
      QrJob& r = jobs[t];
      const int np = (r.N + b - 1) / b;
      r.Y = work.data() + off[t];
      r.P = r.Y + (int64_t)cs * np * b * r.M;
    }
    upload(d_jobs, jobs, stream);
  }

  // A = Q R in place, X = R^H
  void factor(hipStream_t stream) const {
    const unsigned nj = (unsigned)jobs.size();
    if (cs == 2) {
      for (int k = 0; k < npan; ++k) {
        if (lds)
          hipLaunchKernelGGL(zqr_panel_kernel<true>, dim3(nj), dim3(512), lds_bytes, stream, d_jobs.data(), k, b);
        else
          hipLaunchKernelGGL(zqr_panel_kernel<false>, dim3(nj), dim3(512), lds_bytes, stream, d_jobs.data(), k, b);
        DMRG_HIP(hipGetLastError());
        const unsigned gx = (unsigned)((maxN - std::min(maxN, k * b) + kQrUpCols - 1) / kQrUpCols);
        if (gx > 0)
          hipLaunchKernelGGL(zqr_update_kernel<false>, dim3(gx, nj), dim3(512), 0, stream, d_jobs.data(), k, b);
        DMRG_HIP(hipGetLastError());
      }
      const unsigned gx = (unsigned)std::min<int64_t>(256, ((int64_t)maxN * maxN + 255) / 256);
      hipLaunchKernelGGL(zqr_rt_kernel, dim3(gx, nj), dim3(256), 0, stream, d_jobs.data());
      DMRG_HIP(hipGetLastError());
      return;
    }
    for (int k = 0; k < npan; ++k) {
      if (reg_u == 4)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(qr_panel_kernel<true, 4>), dim3(nj), dim3(512), lds_bytes, stream,
                           d_jobs.data(), k, b);
      else if (reg_u == 9)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(qr_panel_kernel<true, 9, 16>), dim3(nj), dim3(1024), lds_bytes, stream,
                           d_jobs.data(), k, b);
      else if (reg_u == 16)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(qr_panel_kernel<true, 16, 16, 16>), dim3(nj), dim3(1024), lds_bytes,
                           stream, d_jobs.data(), k, b);
      else if (reg_u == 32)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(qr_panel_kernel<true, 32, 8, 8>), dim3(nj), dim3(512), lds_bytes,
                           stream, d_jobs.data(), k, b);
      else if (lds)
        hipLaunchKernelGGL(qr_panel_kernel<true>, dim3(nj), dim3(512), lds_bytes, stream,
                           d_jobs.data(), k, b);
      else
        hipLaunchKernelGGL(qr_panel_kernel<false>, dim3(nj), dim3(512), lds_bytes, stream,
                           d_jobs.data(), k, b);
      DMRG_HIP(hipGetLastError());
      const unsigned gx = (unsigned)((maxN - std::min(maxN, k * b) + kQrUpCols - 1) / kQrUpCols);
      if (gx > 0) hipLaunchKernelGGL(qr_update_kernel<false>, dim3(gx, nj), dim3(512), 0, stream, d_jobs.data(), k, b);
      DMRG_HIP(hipGetLastError());
    }
    const unsigned gx = (unsigned)std::min<int64_t>(256, ((int64_t)maxN * maxN + 255) / 256);
    hipLaunchKernelGGL(qr_rt_kernel, dim3(gx, nj), dim3(256), 0, stream, d_jobs.data());
    DMRG_HIP(hipGetLastError());
  }

  // C = Q [S; 0] (or Q C when S is null), then the finish of flagged jobs
  // (real jobs only)
  void back(hipStream_t stream) const {
    const unsigned nj = (unsigned)jobs.size();
    const unsigned gx = (unsigned)std::min<int64_t>(256, ((int64_t)maxM * maxN + 255) / 256);
    const unsigned gu = (unsigned)((maxN + kQrUpCols - 1) / kQrUpCols);
    if (cs == 2) {
      hipLaunchKernelGGL(zqr_seed_kernel, dim3(gx, nj), dim3(256), 0, stream, d_jobs.data());
      DMRG_HIP(hipGetLastError());
      for (int k = npan - 1; k >= 0; --k)
        hipLaunchKernelGGL(zqr_update_kernel<true>, dim3(gu, nj), dim3(512), 0, stream, d_jobs.data(), k, b);
      DMRG_HIP(hipGetLastError());
      return;
    }
    hipLaunchKernelGGL(qr_seed_kernel, dim3(gx, nj), dim3(256), 0, stream, d_jobs.data());
    DMRG_HIP(hipGetLastError());
    for (int k = npan - 1; k >= 0; --k)
      hipLaunchKernelGGL(qr_update_kernel<true>, dim3(gu, nj), dim3(512), 0, stream, d_jobs.data(), k, b);
    DMRG_HIP(hipGetLastError());
    const unsigned gy = (unsigned)std::max(1, (maxN + kFinCols - 1) / kFinCols);
    hipLaunchKernelGGL(qr_finish_kernel, dim3(nj, gy), dim3(256), 0, stream, d_jobs.data());
    DMRG_HIP(hipGetLastError());
  }
};
}  // namespace

void jacobi_svd(const std::vector<SvdProblem>& probs_in, DevBuf<SvdProblem>& scratch,
                DevBuf<int>& sweeps_out, hipStream_t stream, int max_sweeps) {
  if (probs_in.empty()) return;
  const int np = (int)probs_in.size();
  // Jacobi view of every problem.  With B_1 = A = Q_1 R_1 and
  // B_k = R_{k-1}^T = Q_k R_k, the Jacobi runs on X = R_K^T (X J = W):
  //   precond 1 (K = 3): A = (Q_1 Q_3 J) (Q_2 W)^T, G <- Q_1 Q_3 J exact;
  //   precond 2 (K = 2): A = (Q_1 W) (Q_2 J)^T,    V <- Q_2 J exact.
  // Each extra QR of the transposed triangle makes X more nearly diagonal
  // (fewer sweeps).  B_1 is A P with the columns in decreasing norm order
  // (up to kSortMaxN columns); the right factor is then P (Q_2 W) or P (Q_2 J)
  // (rows scattered back before the finish).  On the sectors of an m = 1000
  // sweep with expansion noise this takes the K = 2 form from 13 - 14 sweeps
  // to 8 - 9; the K = 3 form stays at 8.
  std::vector<SvdProblem> probs = probs_in;
  std::vector<int> mode(np, 0), pq;
  for (int i = 0; i < np; ++i) {
    const SvdProblem& p = probs_in[i];
    require(p.M >= p.N && p.N >= 1, "jacobi_svd: needs M >= N >= 1", kErrInternal);
    if (p.N >= kQrMinN && (p.precond == 1 || p.precond == 2)) mode[i] = p.precond;
    if (mode[i]) pq.push_back(i);
  }
  QrPhase ph1, ph2, ph3;
  DevBuf<double> xbuf, sbuf;
  DevBuf<int> sperm;
  DevBuf<SortJob> d_sort;
  std::vector<SortJob> sorts;
  int nq1 = 0, sort_n = 0;
  if (!pq.empty()) {
    // X_k = R_k^T, N x N each: three per precond-1 problem, two per precond 2
    int64_t total = 0, stotal = 0, ptotal = 0;
    for (int i : pq) {
      const SvdProblem& p = probs_in[i];
      total += (int64_t)(mode[i] == 1 ? 3 : 2) * p.N * p.N;
      if (p.N <= kSortMaxN) {
        stotal += (int64_t)p.M * p.N + p.N;
        ptotal += p.N;
      }
    }
    xbuf.ensure((size_t)total);
    sbuf.ensure((size_t)std::max<int64_t>(stotal, 1));
    sperm.ensure((size_t)std::max<int64_t>(ptotal, 1));
    std::vector<QrJob> j1, j2, j3;
    double* x = xbuf.data();
    double* sb = sbuf.data();
    int* sp = sperm.data();
    for (int i : pq) {
      const SvdProblem& p = probs_in[i];
      const int64_t nn = (int64_t)p.N * p.N;
      double* X1 = x;
      double* X2 = x + nn;
      double* X3 = mode[i] == 1 ? x + 2 * nn : nullptr;
      x += (mode[i] == 1 ? 3 : 2) * nn;
      // the first QR factors A P (a presorted copy) when the keys fit in LDS
      const double* A0 = p.G;
      if (p.N <= kSortMaxN) {
        SortJob sj{p.G, sb, sb + (int64_t)p.M * p.N, sp, mode[i] == 1 ? X3 : p.V, p.M, p.N, 1};
        sb += (int64_t)p.M * p.N + p.N;
        sp += p.N;
        sorts.push_back(sj);
        sort_n = std::max(sort_n, p.N);
        A0 = sj.dst;
      }
      QrJob a{};
      a.A = const_cast<double*>(A0); a.X = X1; a.C = p.G; a.V = p.V; a.sigma = p.sigma;
      a.M = p.M; a.N = p.N; a.u_in_v = p.u_in_v;
      if (mode[i] == 1) { a.S = p.V; a.W = X3; a.finish = 1; }  // [Q_3 J; 0], Q_2 W
      else a.S = X2;                                           // [W; 0]
      j1.push_back(a);
      QrJob b{};
      b.A = X1; b.X = X2; b.C = mode[i] == 1 ? X3 : p.V;  // Q_2 W or Q_2 J
      b.M = p.N; b.N = p.N;
      j2.push_back(b);
      if (mode[i] == 1) {
        QrJob c{};
        c.A = X2; c.X = X3; c.C = p.V;  // Q_3 J
        c.M = p.N; c.N = p.N;
        j3.push_back(c);
        probs[i].G = X3;
        ++nq1;
      } else {
        probs[i].G = X2;
      }
      probs[i].M = p.N;
    }
    ph1.plan(j1, stream);
    ph2.plan(j2, stream);
    if (!j3.empty()) ph3.plan(j3, stream);
    if (!sorts.empty()) {
      upload(d_sort, sorts, stream);
      const unsigned ns = (unsigned)sorts.size(), gy = (unsigned)((sort_n + kFinCols - 1) / kFinCols);
      hipLaunchKernelGGL(sort_colnorm_kernel, dim3(ns, gy), dim3(256), 0, stream, d_sort.data());
      hipLaunchKernelGGL(sort_rank_kernel, dim3(ns), dim3(1024), sort_n * sizeof(double), stream,
                         d_sort.data());
      hipLaunchKernelGGL(sort_gather_kernel, dim3(ns, gy), dim3(256), 0, stream, d_sort.data());
      DMRG_HIP(hipGetLastError());
    }
    ph1.factor(stream);
    ph2.factor(stream);
    if (!j3.empty()) ph3.factor(stream);
  }
  upload(scratch, probs, stream);
  int init_n = 1;
  for (const SvdProblem& q : probs) init_n = std::max(init_n, q.N);
  const int init_gy = (init_n + kFinCols - 1) / kFinCols;
  DevBuf<double> fro_part((size_t)np * init_gy);
  hipLaunchKernelGGL(jacobi_init_kernel, dim3(np, init_gy), dim3(256), 0, stream, scratch.data(),
                     fro_part.data());
  hipLaunchKernelGGL(svd_fro2_kernel, dim3(np), dim3(64), 0, stream, scratch.data(), fro_part.data(), init_gy);
  DMRG_HIP(hipGetLastError());
  // the rotations of two rounds in qbuf
  DevBuf<double> qbuf;
  std::vector<int> last;
  int visits = 0;
  {
    int nv = 0;
    for (const SvdProblem& q : probs)
      if (q.N > 1) {
        const int nb = (q.N + kJb - 1) / kJb;
        nv += (nb + (nb & 1)) / 2;
      }
    qbuf.ensure((size_t)std::max(2 * nv, 1) * kJc * kJc);
  }
  const int sweeps_done = run_sweeps(
      probs, kJb, sweeps_out, stream, max_sweeps, last, visits,
      [&](dim3 grid, int r, SweepCtl ctl, const int* vprob, const int* vpair, int* qflag, int nv) {
        hipLaunchKernelGGL(jacobi_gram_round_kernel, grid, dim3(64 * kJw), 0, stream, scratch.data(), vprob,
                           vpair, r, ctl, qbuf.data(), qflag, nv);
      });
  if (nq1 > 0) ph3.back(stream);        // V = Q_3 J             (precond 1)
  if (!pq.empty()) {
    ph2.back(stream);                   // X3 = Q_2 W or V = Q_2 J
    if (!sorts.empty()) {               // rows of X3 or V back through P
      const unsigned ns = (unsigned)sorts.size(), gy = (unsigned)((sort_n + kFinCols - 1) / kFinCols);
      hipLaunchKernelGGL(sort_scatter_rows_kernel, dim3(ns, gy), dim3(256), 0, stream, d_sort.data(), 0);
      hipLaunchKernelGGL(sort_scatter_rows_kernel, dim3(ns, gy), dim3(256), 0, stream, d_sort.data(), 1);
      DMRG_HIP(hipGetLastError());
    }
    ph1.back(stream);                   // G = Q_1 [Q_3 J; 0] (+ finish) or Q_1 [W; 0]
  }
  // sigma and sign convention from G = U Sigma and V (plain and precond 2)
  for (int i = 0; i < np; ++i) {
    probs[i] = probs_in[i];
    probs[i].precond = mode[i] == 1 ? 1 : 0;
  }
  upload(scratch, probs, stream);
  int fin_n = 1;
  for (const SvdProblem& q : probs_in) fin_n = std::max(fin_n, q.N);
  hipLaunchKernelGGL(jacobi_finish_kernel, dim3(np, (fin_n + kFinCols - 1) / kFinCols), dim3(256), 0, stream,
                     scratch.data());
  DMRG_HIP(hipGetLastError());
  // the phases' buffers are released when this function returns
  if (!pq.empty()) DMRG_HIP(hipStreamSynchronize(stream));
  if (std::getenv("DMRG_SVD_STATS")) {
    int maxN = 0; int64_t mn2 = 0;
    for (const SvdProblem& p : probs_in) { maxN = std::max(maxN, p.N); mn2 += (int64_t)p.M * p.N * p.N; }
    std::printf("SVDSTAT np=%d qr=%d qr3=%d maxN=%d sweeps=%d MN2=%lld |", np, (int)pq.size(),
                nq1, maxN,
                sweeps_done, (long long)mn2);
    for (int i = 0; i < np; ++i)
      if ((int64_t)probs_in[i].M * probs_in[i].N > 20000)
        std::printf(" %dx%d:%d:%d", probs_in[i].M, probs_in[i].N, mode[i], last[i]);
    std::printf("\n");
  }
}

void zjacobi_svd(const std::vector<SvdProblem>& probs_in, DevBuf<SvdProblem>& scratch,
                 DevBuf<int>& sweeps_out, hipStream_t stream, int max_sweeps) {
  if (probs_in.empty()) return;
  const int np = (int)probs_in.size();
  for (const SvdProblem& p : probs_in)
    require(p.M >= p.N && p.N >= 1, "zjacobi_svd: needs M >= N >= 1", kErrInternal);
  // QR preconditioning (problems with precond set and N >= kQrMinN): with
  // A P = Q_1 R_1 (P: the columns in decreasing norm order) and R_1^H = Q_2 R_2
  // the Jacobi runs on X = R_2^H (X J = W), and A = (Q_1 [W; 0]) (P Q_2 J)^H:
  // the left factor and sigma come from G = Q_1 [W; 0], V = P Q_2 J is
  // unitary to rounding.  DMRGCC_ZSVD_QR=0 turns it off.
  const char* qenv = std::getenv("DMRGCC_ZSVD_QR");
  const bool qr_on = !qenv || std::atoi(qenv) != 0;
  std::vector<SvdProblem> probs = probs_in;
  std::vector<int> pq;
  for (int i = 0; i < np; ++i)
    if (qr_on && probs_in[i].precond != 0 && probs_in[i].N >= kQrMinN) pq.push_back(i);
  QrPhase ph1, ph2;
  DevBuf<double> xbuf, sbuf;
  DevBuf<int> sperm;
  DevBuf<SortJob> d_sort;
  std::vector<SortJob> sorts;
  int sort_n = 0;
  if (!pq.empty()) {
    int64_t total = 0, stotal = 0, ptotal = 0;
    for (int i : pq) {
      const SvdProblem& p = probs_in[i];
      total += 4 * (int64_t)p.N * p.N;  // X_1, X_2 (complex)
      if (p.N <= kSortMaxN) {
        stotal += 2 * (int64_t)p.M * p.N + p.N;
        ptotal += p.N;
      }
    }
    xbuf.ensure((size_t)total);
    sbuf.ensure((size_t)std::max<int64_t>(stotal, 1));
    sperm.ensure((size_t)std::max<int64_t>(ptotal, 1));
    std::vector<QrJob> j1, j2;
    double* x = xbuf.data();
    double* sb = sbuf.data();
    int* sp = sperm.data();
    for (int i : pq) {
      const SvdProblem& p = probs_in[i];
      const int64_t nn = (int64_t)p.N * p.N;
      double* X1 = x;
      double* X2 = x + 2 * nn;
      x += 4 * nn;
      const double* A0 = p.G;
      if (p.N <= kSortMaxN) {  // the first QR factors a presorted copy
        SortJob sj{p.G, sb, sb + 2 * (int64_t)p.M * p.N, sp, p.V, p.M, p.N, 2};
        sb += 2 * (int64_t)p.M * p.N + p.N;
        sp += p.N;
        sorts.push_back(sj);
        sort_n = std::max(sort_n, p.N);
        A0 = sj.dst;
      }
      QrJob a{};
      a.A = const_cast<double*>(A0); a.X = X1; a.C = p.G; a.S = X2;  // G = Q_1 [W; 0]
      a.M = p.M; a.N = p.N;
      j1.push_back(a);
      QrJob c{};
      c.A = X1; c.X = X2; c.C = p.V;  // V = Q_2 J
      c.M = p.N; c.N = p.N;
      j2.push_back(c);
      probs[i].G = X2;
      probs[i].M = p.N;
    }
    ph1.plan(j1, stream, 2);
    ph2.plan(j2, stream, 2);
    if (!sorts.empty()) {
      upload(d_sort, sorts, stream);
      const unsigned ns = (unsigned)sorts.size(), gy = (unsigned)((sort_n + kFinCols - 1) / kFinCols);
      hipLaunchKernelGGL(sort_colnorm_kernel, dim3(ns, gy), dim3(256), 0, stream, d_sort.data());
      hipLaunchKernelGGL(sort_rank_kernel, dim3(ns), dim3(1024), sort_n * sizeof(double), stream,
                         d_sort.data());
      hipLaunchKernelGGL(sort_gather_kernel, dim3(ns, gy), dim3(256), 0, stream, d_sort.data());
      DMRG_HIP(hipGetLastError());
    }
    ph1.factor(stream);
    ph2.factor(stream);
  }
  upload(scratch, probs, stream);
  int init_n = 1;
  for (const SvdProblem& q : probs) init_n = std::max(init_n, q.N);
  const int init_gy = (init_n + kFinCols - 1) / kFinCols;
  DevBuf<double> fro_part((size_t)np * init_gy);
  hipLaunchKernelGGL(zjacobi_init_kernel, dim3(np, init_gy), dim3(256), 0, stream, scratch.data(),
                     fro_part.data());
  hipLaunchKernelGGL(svd_fro2_kernel, dim3(np), dim3(64), 0, stream, scratch.data(), fro_part.data(), init_gy);
  DMRG_HIP(hipGetLastError());
  constexpr int kB = kZc / 2;
  // the rotations of two rounds in qbuf (Qr then Qi per slot)
  DevBuf<double> qbuf;
  std::vector<int> last;
  int visits = 0;
  {
    int nv = 0;
    for (const SvdProblem& q : probs)
      if (q.N > 1) {
        const int nb = (q.N + kB - 1) / kB;
        nv += (nb + (nb & 1)) / 2;
      }
    qbuf.ensure((size_t)std::max(2 * nv, 1) * 2 * kZc * kZc);
  }
  const int sweeps_done = run_sweeps(
      probs, kB, sweeps_out, stream, max_sweeps, last, visits,
      [&](dim3 grid, int r, SweepCtl ctl, const int* vprob, const int* vpair, int* qflag, int nv) {
        hipLaunchKernelGGL(zjacobi_round_kernel, grid, dim3(64 * kZw), 0, stream, scratch.data(), vprob, vpair,
                           r, ctl, qbuf.data(), qflag, nv);
      });
  if (!pq.empty()) {
    ph2.back(stream);  // V = Q_2 J
    if (!sorts.empty()) {  // rows of V back through P
      const unsigned ns = (unsigned)sorts.size(), gy = (unsigned)((sort_n + kFinCols - 1) / kFinCols);
      hipLaunchKernelGGL(sort_scatter_rows_kernel, dim3(ns, gy), dim3(256), 0, stream, d_sort.data(), 0);
      hipLaunchKernelGGL(sort_scatter_rows_kernel, dim3(ns, gy), dim3(256), 0, stream, d_sort.data(), 1);
      DMRG_HIP(hipGetLastError());
    }
    ph1.back(stream);  // G = Q_1 [W; 0]
    upload(scratch, probs_in, stream);
  }
  int fin_n = 1;
  for (const SvdProblem& q : probs_in) fin_n = std::max(fin_n, q.N);
  hipLaunchKernelGGL(zjacobi_finish_kernel, dim3(np, (fin_n + kFinCols - 1) / kFinCols), dim3(256), 0, stream,
                     scratch.data());
  DMRG_HIP(hipGetLastError());
  // the phases' buffers are released when this function returns
  if (!pq.empty()) DMRG_HIP(hipStreamSynchronize(stream));
  if (std::getenv("DMRG_SVD_STATS")) {
    int maxN = 0;
    for (const SvdProblem& q : probs_in) maxN = std::max(maxN, q.N);
    std::printf("ZSVDSTAT np=%d qr=%d maxN=%d sweeps=%d |", np, (int)pq.size(), maxN, sweeps_done);
    for (int i = 0; i < np; ++i)
      if ((int64_t)probs_in[i].M * probs_in[i].N > 20000)
        std::printf(" %dx%d:%d", probs_in[i].M, probs_in[i].N, last[i]);
    std::printf("\n");
  }
}

// ---------------------------------------------------------------------------
// The allocation-free arena: its device address, looked up once per device
// with that device current (the array's name in host code is only a
// host-side stand-in, never a device address).
// ---------------------------------------------------------------------------
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_arena_mu;
void* g_arena[kMaxDevices] = {};
}  // namespace

void* noalloc_arena(size_t* bytes) {
  int dev = 0;
  DMRG_HIP(hipGetDevice(&dev));
  require(dev >= 0 && dev < kMaxDevices, "device index out of range", kErrArgument);
  std::lock_guard<std::mutex> lk(g_arena_mu);
  if (!g_arena[dev]) {
    void* p = nullptr;
    DMRG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    require(p != nullptr, "no device address for the arena", kErrHip);
    g_arena[dev] = p;
  }
  if (bytes) *bytes = kNoAllocArenaBytes;
  return g_arena[dev];
}

}  // namespace dmrg
