// DMRG engine: host bookkeeping of the block-sparse MPS and environments,
// device numerics through the grouped-GEMM / copy / SVD / vector kernels.
//
// Every routine restates the reference algorithm named in its comment
// (src/dmrg/realdmrg/*.cc); loop orders that decide block order, sector
// order or the order of rand() draws follow the reference exactly.
#include "dmrg.h"

#include <algorithm>
#include <future>
#include <unordered_map>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <numeric>
#include <set>
#include <thread>

#include "quicksort.h"

namespace dmrg {

namespace {

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// CombineTensorLattice (real_tensor_lattice.cc:361-484): ascending union of
// two sector lists with dims added.  The reference's first counting loop
// bounds its inner index by the second list's length (:426); the intended
// union is restated (identical whenever the expansion list is a prefix of
// the flagged sectors, which the expansion lattice construction guarantees).
void combine_sectors(const std::vector<int>& b0, const std::vector<int>& d0,
                     const std::vector<int>& b1, const std::vector<int>& d1, std::vector<int>& out,
                     std::vector<int>& dims) {
  out = b0;
  for (int b : b1)
    if (std::find(b0.begin(), b0.end(), b) == b0.end()) out.push_back(b);
  quicksort(out, nullptr, 1);
  dims.assign(out.size(), 0);
  for (size_t i = 0; i < out.size(); ++i) {
    auto a = std::find(b0.begin(), b0.end(), out[i]);
    if (a != b0.end()) dims[i] += d0[a - b0.begin()];
    auto c = std::find(b1.begin(), b1.end(), out[i]);
    if (c != b1.end()) dims[i] += d1[c - b1.begin()];
  }
}

int index_of(const std::vector<int>& v, int x) {
  auto it = std::find(v.begin(), v.end(), x);
  return it == v.end() ? -1 : (int)(it - v.begin());
}

// expan_dim clamp (real_tensor_contraction.cc:203-206, :383-386)
int expan_width(int D, int d) {
  int w = (int)std::sqrt((double)D) * d * d;
  return std::min(std::max(w, 1), 25);
}

struct Shape {
  int rows = 0, cols = 0;
  int64_t size() const { return (int64_t)rows * cols; }
};

// Push a copy given in entry units (pointers at entry starts, strides in
// entries) as real copies: one, or the re and im parts of interleaved
// complex entries (cs = 2).
void push_copy(std::vector<CopyTask>& v, CopyTask c, int cs) {
  if (cs == 1) {
    v.push_back(c);
    return;
  }
  c.ssr *= 2; c.ssc *= 2; c.dsr *= 2; c.dsc *= 2;
  for (int part = 0; part < 2; ++part) {
    CopyTask t = c;
    if (t.src) t.src += part;
    t.dst += part;
    v.push_back(t);
  }
}

// One expansion random matrix: RandomMatrix(rows, cols) (complex: two draws
// per entry) at rnd[roff], then its diagonal set to 0.1 * (w + i wi)
// (real_tensor_contraction.cc:397-399, complex_tensor_contraction.cc:397-399).
struct RandomDraw {
  int64_t roff;
  int rows, cols;
  double w, wi;
  bool diag = true;  // false: leave the raw draws on the diagonal
};

// Fill the random matrices in draw order from `rng`, which then stands after
// the last draw.  Large sets are split over host threads, each starting from
// a jump-ahead copy of the stream, so the values are those of one sequential
// pass.
void fill_random(const std::vector<RandomDraw>& ds, double* rnd, int cs, GlibcRand& rng) {
  std::vector<int64_t> first(ds.size() + 1, 0);
  for (size_t k = 0; k < ds.size(); ++k) first[k + 1] = first[k] + (int64_t)cs * ds[k].rows * ds[k].cols;
  const int64_t total = first.back();
  auto work = [&](size_t k0, size_t k1) {
    GlibcRand g = rng;
    g.skip(first[k0]);
    for (size_t k = k0; k < k1; ++k) {
      const RandomDraw& d = ds[k];
      double* R = rnd + d.roff;
      if (cs == 2) {
        g.complex_matrix(d.rows, d.cols, R);
        for (int q = 0; q < (d.diag ? std::min(d.rows, d.cols) : 0); ++q) {
          R[2 * ((int64_t)q * d.cols + q)] = 0.1 * d.w;
          R[2 * ((int64_t)q * d.cols + q) + 1] = 0.1 * d.wi;
        }
      } else {
        g.random_matrix(d.rows, d.cols, R);
        for (int q = 0; q < (d.diag ? std::min(d.rows, d.cols) : 0); ++q) R[(int64_t)q * d.cols + q] = 0.1 * d.w;
      }
    }
  };
  const int64_t hw = std::max(1u, std::thread::hardware_concurrency());
  const int nt = (int)std::min<int64_t>(std::min<int64_t>(hw, 16), total / (1 << 18));
  if (nt <= 1) {
    work(0, ds.size());
  } else {
    std::vector<std::thread> th;
    size_t k0 = 0;
    for (int t = 0; t < nt && k0 < ds.size(); ++t) {
      const int64_t target = total * (t + 1) / nt;
      size_t k1 = k0 + 1;
      while (k1 < ds.size() && first[k1] < target) ++k1;
      if (t == nt - 1) k1 = ds.size();
      th.emplace_back(work, k0, k1);
      k0 = k1;
    }
    for (auto& x : th) x.join();
  }
  rng.skip(total);
}

// The random matrices of `ds` made on the device into R (device memory): the
// draws fill_random makes on the host, from the stream's current window,
// and `rng` then advanced past them as fill_random leaves it.  Lane c of
// rng_stream starts chunk c of the draws from the window that the
// coefficients of x^(c kRngChunk) mod P give (GlibcRand::chunk_powers),
// made once per context and kept on the device.
void random_matrices_device(const std::vector<RandomDraw>& ds, int cs, double* R, Context& cx) {
  int64_t total = 0, most = 0;
  std::vector<RngDraw> dd;
  dd.reserve(ds.size());
  for (const RandomDraw& d : ds) {
    const int64_t n = (int64_t)d.rows * d.cols;
    dd.push_back(RngDraw{total, d.roff, d.rows, d.cols, d.w, d.wi, d.diag ? 1 : 0, 0});
    total += (int64_t)cs * n;
    most = std::max(most, n);
  }
  if (total == 0) return;
  const int64_t nchunks = (total + kRngChunk - 1) / kRngChunk;
  const int64_t have = (int64_t)cx.rng_coef_host.size() / 31;
  if (nchunks > have) {
    const int64_t want = std::max(nchunks, 2 * have);
    cx.rng_coef_host.resize((size_t)want * 31);
    GlibcRand::chunk_powers(kRngChunk, have, want - have,
                            have ? cx.rng_coef_host.data() + 31 * (have - 1) : nullptr,
                            cx.rng_coef_host.data() + 31 * have);
    upload(cx.rng_coef, cx.rng_coef_host, cx.stream);
  }
  std::vector<uint32_t> w(61);
  cx.rng.window(w.data());
  upload(cx.rng_w0, w, cx.stream);
  cx.rng_raw.ensure((size_t)(nchunks * kRngChunk));
  rng_stream(cx.rng_w0.data(), cx.rng_coef.data(), total, cx.rng_raw.data(), cx.stream);
  upload(cx.rng_draws, dd, cx.stream);
  rng_layout(cx.rng_draws.data(), (int)dd.size(), most, cx.rng_raw.data(), R, cs, cx.stream);
  cx.rng.skip(total);
}

// One piece of a block assembled by the expansion / match fills.
struct Piece {
  const double* src;  // nullptr: zeros
  int rows, cols;
  int64_t ssr;
  int r0, c0;
};

}  // namespace

void random_matrices(GlibcRand& rng, int n, const int* rows, const int* cols, int cs, const double* w,
                     double* out) {
  require(n >= 0 && (cs == 1 || cs == 2) && (n == 0 || (rows && cols && out)), "bad arguments",
          kErrArgument);
  std::vector<RandomDraw> ds;
  int64_t off = 0;
  for (int k = 0; k < n; ++k) {
    require(rows[k] >= 0 && cols[k] >= 0, "bad matrix shape", kErrArgument);
    ds.push_back(RandomDraw{off, rows[k], cols[k], w ? w[cs * k] : 0.0, (w && cs == 2) ? w[cs * k + 1] : 0.0,
                            w != nullptr});
    off += (int64_t)cs * rows[k] * cols[k];
  }
  std::vector<double> buf((size_t)off);
  fill_random(ds, buf.data(), cs, rng);
  std::copy(buf.begin(), buf.end(), out);
}

// ComputeTruncateDim (real_tensor_lattice.cc:501-555).  `sv` holds each
// sector's singular values in descending order; a sector left with no state
// keeps one with its first singular value set to 0.
std::vector<int> truncate_dims(int max_dim, double canpre, std::vector<std::vector<double>>& sv) {
  if (canpre > 1) canpre = 0.0;
  const int ns = (int)sv.size();
  std::vector<int> t(ns, 0);
  int total = 0;
  for (int i = 0; i < ns; ++i) {
    for (double s : sv[i])
      if (s > canpre) ++t[i];
    total += t[i];
  }
  if (total > 0 && max_dim > 0 && max_dim < total) {
    std::vector<double> vals;
    std::vector<int> blk;
    vals.reserve(total);
    blk.reserve(total);
    for (int i = 0; i < ns; ++i)
      for (int j = 0; j < t[i]; ++j) {
        vals.push_back(sv[i][j]);
        blk.push_back(i);
      }
    std::fill(t.begin(), t.end(), 0);
    quicksort(vals, &blk, 0);
    for (int i = 0; i < max_dim; ++i) ++t[blk[i]];
  }
  for (int i = 0; i < ns; ++i)
    if (t[i] == 0) {
      t[i] = 1;
      sv[i][0] = 0.0;
    }
  return t;
}

// ===========================================================================
// BlockTensor / Lattice
// ===========================================================================
void BlockTensor::define(const std::vector<int>& l, const std::vector<int>& r) {
  blk.clear();
  for (size_t i = 0; i < l.size(); ++i) blk.push_back(Block{l[i], r[i], 0, 0, 0});
  size = 0;
  data.reset();
}

void BlockTensor::allocate() {
  int64_t acc = 0;
  for (Block& b : blk) {
    b.off = acc;
    acc += (int64_t)b.rows * b.cols;
  }
  size = acc;
  data.resize((size_t)(cs * acc));
}

int BlockTensor::find(int l, int r) const {
  for (int i = 0; i < n(); ++i)
    if (blk[i].l == l && blk[i].r == r) return i;
  return -1;
}

BlockIndex::BlockIndex(const BlockTensor& t) {
  m_.reserve(2 * (size_t)t.n());
  for (int i = 0; i < t.n(); ++i) m_.emplace(key(t.blk[i].l, t.blk[i].r), i);
}

std::vector<int> BlockTensor::same(int leigh, int target) const {
  std::vector<int> out;
  for (int i = 0; i < n(); ++i)
    if ((leigh == 0 ? blk[i].l : blk[i].r) == target) out.push_back(i);
  return out;
}

void Lattice::define_sectors(const std::vector<int>& l, const std::vector<int>& r,
                             const std::vector<int>& ldim, const std::vector<int>& rdim) {
  lb = l;
  rb = r;
  ld = ldim;
  rd = rdim;
  phys.assign(l.size(), std::vector<int>(r.size(), -1));
}

Context::Context(hipStream_t s) : stream(s) {
  if (!stream) {
    // A non-blocking stream (default priority): launches on a stream that
    // synchronises with the null stream cost tens of microseconds of host
    // time each.  Callers finish the work producing the inputs before a call;
    // every entry point waits for its own work before returning.
    DMRG_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    own_stream = true;
  }
  scal.resize(64);
}

Context::~Context() {
  upload_forget_stream(stream);
  if (!events_.empty()) (void)hipStreamSynchronize(stream);
  for (hipEvent_t e : events_) (void)hipEventDestroy(e);
  if (lz_host_) (void)hipHostFree(lz_host_);
  if (pinned_) (void)hipHostFree(pinned_);
  if (own_stream && stream) (void)hipStreamDestroy(stream);
}

double* Context::lanczos_host() {
  if (!lz_host_) DMRG_HIP(hipHostMalloc((void**)&lz_host_, sizeof(double) * (4 * 200 + 200 + 8)));
  return lz_host_;
}

double* Context::lanczos_host_dev() {
  // the device's address of the same pinned buffer, for kernels that write it
  if (!lz_host_dev_) DMRG_HIP(hipHostGetDevicePointer((void**)&lz_host_dev_, lanczos_host(), 0));
  return lz_host_dev_;
}

std::vector<hipEvent_t> Context::events(int n) {
  while ((int)events_.size() < n) {
    hipEvent_t e;
    DMRG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    events_.push_back(e);
  }
  return std::vector<hipEvent_t>(events_.begin(), events_.begin() + n);
}

double* Context::pinned(int64_t n) {
  if (n > pinned_n_) {
    // the previous contents are not kept; callers finish their copies
    // (synchronise the stream) before asking again
    DMRG_HIP(hipStreamSynchronize(stream));
    if (pinned_) DMRG_HIP(hipHostFree(pinned_));
    pinned_ = nullptr;
    pinned_n_ = 0;
    const int64_t cap = std::max<int64_t>(n, 1 << 20);
    DMRG_HIP(hipHostMalloc((void**)&pinned_, (size_t)cap * sizeof(double), hipHostMallocDefault));
    pinned_n_ = cap;
  }
  return pinned_;
}

// ===========================================================================
// dense symmetric eigensolver (cyclic Jacobi)
// ===========================================================================
void sym_eig(int n, std::vector<double> a, std::vector<double>& eval, std::vector<double>& vec) {
  std::vector<double> v((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) v[(size_t)i * n + i] = 1.0;
  auto A = [&](int i, int j) -> double& { return a[(size_t)i * n + j]; };
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, dia = 0.0;
    for (int i = 0; i < n; ++i) {
      dia += A(i, i) * A(i, i);
      for (int j = i + 1; j < n; ++j) off += A(i, j) * A(i, j);
    }
    if (off <= 1e-34 * dia || off == 0.0) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = A(p, q);
        if (apq == 0.0) continue;
        const double app = A(p, p), aqq = A(q, q);
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::hypot(1.0, theta));
        const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
        for (int k = 0; k < n; ++k) {
          const double akp = A(k, p), akq = A(k, q);
          A(k, p) = c * akp - s * akq;
          A(k, q) = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = A(p, k), aqk = A(q, k);
          A(p, k) = c * apk - s * aqk;
          A(q, k) = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double vkp = v[(size_t)k * n + p], vkq = v[(size_t)k * n + q];
          v[(size_t)k * n + p] = c * vkp - s * vkq;
          v[(size_t)k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return A(x, x) < A(y, y); });
  eval.resize(n);
  vec.assign((size_t)n * n, 0.0);
  for (int j = 0; j < n; ++j) {
    eval[j] = A(order[j], order[j]);
    // sign convention: largest |component| (first on ties) positive
    int bi = 0;
    for (int i = 1; i < n; ++i)
      if (std::fabs(v[(size_t)i * n + order[j]]) > std::fabs(v[(size_t)bi * n + order[j]])) bi = i;
    const double sg = v[(size_t)bi * n + order[j]] < 0 ? -1.0 : 1.0;
    for (int i = 0; i < n; ++i) vec[(size_t)i * n + j] = sg * v[(size_t)i * n + order[j]];
  }
}

// Lowest eigenpair of a dense symmetric matrix (n x n, row-major): the part of
// RealSymMatrixDiag (general.cc:47-64) the dense edge-site solve uses.
//  1. Householder reduction to tridiagonal form T = Q^T A Q (Golub & Van Loan,
//     Matrix Computations, Alg. 8.3.1, with the cancellation-free reflector of
//     Alg. 5.1.1);
//  2. the smallest eigenvalue of T by bisection on the Sturm count of
//     T - x I (Golub & Van Loan, sec. 8.4.1), to the spacing of doubles;
//  3. its eigenvector by inverse iteration on T - lambda I (tridiagonal
//     elimination with partial pivoting, three steps);
//  4. x = Q y, normalised, largest |component| (first on ties) positive.
// For a degenerate lowest eigenvalue the vector lies in its eigenspace, as
// with any eigensolver.
void sym_eig_lowest(int n, std::vector<double> a, double& e0, std::vector<double>& v0) {
  auto A = [&](int i, int j) -> double& { return a[(size_t)i * n + j]; };
  std::vector<std::vector<double>> hv(std::max(n - 2, 0));
  std::vector<double> hb(std::max(n - 2, 0), 0.0);
  for (int k = 0; k + 2 < n; ++k) {
    const int m = n - k - 1;
    std::vector<double> v(m);
    for (int i = 0; i < m; ++i) v[i] = A(k + 1 + i, k);
    double sigma = 0.0;
    for (int i = 1; i < m; ++i) sigma += v[i] * v[i];
    double beta = 0.0;
    if (sigma > 0.0) {
      const double x0 = v[0];
      const double mu = std::sqrt(x0 * x0 + sigma);
      const double w0 = x0 <= 0.0 ? x0 - mu : -sigma / (x0 + mu);
      beta = 2.0 * w0 * w0 / (sigma + w0 * w0);
      for (int i = 1; i < m; ++i) v[i] /= w0;
      v[0] = 1.0;
      // A22 <- P A22 P with P = I - beta v v^T: p = beta A22 v,
      // w = p - (beta/2)(p^T v) v, A22 -= v w^T + w v^T
      std::vector<double> pw(m, 0.0);
      for (int i = 0; i < m; ++i) {
        double t = 0.0;
        for (int j = 0; j < m; ++j) t += A(k + 1 + i, k + 1 + j) * v[j];
        pw[i] = beta * t;
      }
      double pv = 0.0;
      for (int i = 0; i < m; ++i) pv += pw[i] * v[i];
      const double c = 0.5 * beta * pv;
      for (int i = 0; i < m; ++i) pw[i] -= c * v[i];
      for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) A(k + 1 + i, k + 1 + j) -= v[i] * pw[j] + pw[i] * v[j];
      A(k + 1, k) = A(k, k + 1) = mu;  // P x = |x| e_1
    }
    hv[k] = std::move(v);
    hb[k] = beta;
  }
  std::vector<double> d(n), e(std::max(n - 1, 0));
  for (int i = 0; i < n; ++i) d[i] = A(i, i);
  for (int i = 0; i + 1 < n; ++i) e[i] = A(i + 1, i);
  // smallest eigenvalue: Sturm count of eigenvalues below x
  double tnorm = 0.0, lo = 0.0, hi = 0.0;
  for (int i = 0; i < n; ++i) {
    const double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < n ? std::fabs(e[i]) : 0.0);
    lo = i == 0 ? d[i] - r : std::min(lo, d[i] - r);
    hi = i == 0 ? d[i] + r : std::max(hi, d[i] + r);
    tnorm = std::max(tnorm, std::fabs(d[i]) + r);
  }
  const double tiny = std::numeric_limits<double>::min() / std::numeric_limits<double>::epsilon();
  auto below = [&](double x) {
    int cnt = 0;
    double q = d[0] - x;
    if (q == 0.0) q = -tiny;
    if (q < 0.0) ++cnt;
    for (int i = 1; i < n; ++i) {
      q = d[i] - x - e[i - 1] * e[i - 1] / q;
      if (q == 0.0) q = -tiny;
      if (q < 0.0) ++cnt;
    }
    return cnt;
  };
  lo -= std::numeric_limits<double>::epsilon() * tnorm + tiny;
  hi += std::numeric_limits<double>::epsilon() * tnorm + tiny;
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (below(mid) >= 1) hi = mid;
    else lo = mid;
  }
  const double lam = 0.5 * (lo + hi);
  // inverse iteration: (T - lam I) y = b, elimination with row interchanges
  std::vector<double> y(n, 1.0);
  const double pivmin = std::numeric_limits<double>::epsilon() * std::max(tnorm, tiny);
  for (int step = 0; step < 3 && n > 1; ++step) {
    // rows of the banded system: r0 (diag), r1 (super), r2 (second super) after pivoting
    std::vector<double> dd(n), du(n, 0.0), du2(n, 0.0), rhs = y;
    std::vector<double> dl(n, 0.0);
    for (int i = 0; i < n; ++i) dd[i] = d[i] - lam;
    for (int i = 0; i + 1 < n; ++i) { du[i] = e[i]; dl[i] = e[i]; }
    for (int i = 0; i + 1 < n; ++i) {
      if (std::fabs(dd[i]) >= std::fabs(dl[i])) {
        if (std::fabs(dd[i]) < pivmin) dd[i] = dd[i] < 0 ? -pivmin : pivmin;
        const double f = dl[i] / dd[i];
        dd[i + 1] -= f * du[i];
        if (i + 2 < n) du2[i] = 0.0;
        rhs[i + 1] -= f * rhs[i];
      } else {  // swap rows i and i+1
        const double f = dd[i] / dl[i];
        dd[i] = dl[i];
        const double t = dd[i + 1];
        dd[i + 1] = du[i] - f * t;
        du[i] = t;
        if (i + 2 < n) {
          du2[i] = du[i + 1];
          du[i + 1] = -f * du2[i];
        }
        std::swap(rhs[i], rhs[i + 1]);
        rhs[i + 1] -= f * rhs[i];
      }
    }
    if (std::fabs(dd[n - 1]) < pivmin) dd[n - 1] = dd[n - 1] < 0 ? -pivmin : pivmin;
    // back substitution
    y[n - 1] = rhs[n - 1] / dd[n - 1];
    if (n >= 2) y[n - 2] = (rhs[n - 2] - du[n - 2] * y[n - 1]) / dd[n - 2];
    for (int i = n - 3; i >= 0; --i) y[i] = (rhs[i] - du[i] * y[i + 1] - du2[i] * y[i + 2]) / dd[i];
    double nrm = 0.0;
    for (double x : y) nrm += x * x;
    nrm = std::sqrt(nrm);
    for (double& x : y) x /= nrm;
  }
  // x = Q y = P_0 P_1 ... P_{n-3} y
  for (int k = n - 3; k >= 0; --k) {
    if (hb[k] == 0.0) continue;
    const std::vector<double>& v = hv[k];
    double t = 0.0;
    for (size_t i = 0; i < v.size(); ++i) t += v[i] * y[k + 1 + i];
    t *= hb[k];
    for (size_t i = 0; i < v.size(); ++i) y[k + 1 + i] -= t * v[i];
  }
  double nrm = 0.0;
  for (double x : y) nrm += x * x;
  nrm = std::sqrt(nrm);
  int bi = 0;
  for (int i = 0; i < n; ++i) {
    y[i] /= nrm;
    if (std::fabs(y[i]) > std::fabs(y[bi])) bi = i;
  }
  if (y[bi] < 0.0)
    for (double& x : y) x = -x;
  e0 = n == 1 ? d[0] : lam;
  v0 = std::move(y);
}

// ===========================================================================
// H_eff plan (MultiplyEffectHamilton, real_tensor_contraction.cc:1223-1291)
// ===========================================================================
void HeffPlan::build(const Lattice& lat, const SiteOperator& W, const Env& L, const Env& R,
                     hipStream_t stream) {
  std::vector<int> bphys(lat.ket.n());
  for (int k = 0; k < lat.ket.n(); ++k) bphys[k] = lat.phys[lat.ket.blk[k].l][lat.ket.blk[k].r];
  build_blocks(lat.ket, bphys, lat.nr(), W, L, R, stream);
}

void HeffPlan::build_blocks(const BlockTensor& ket, const std::vector<int>& bphys, int nr,
                            const SiteOperator& W, const Env& L, const Env& R, hipStream_t stream) {
  // Blocks (l, r) of the state with physical index bphys[k]; several blocks
  // may share a sector pair (the two-site state, one block per middle sector).
  s1_.clear();
  s2_.clear();
  s1_.split_k = s2_.split_k = want_split_ >= 0 ? want_split_ : default_split_k();
  s1_.tile_mode = s2_.tile_mode = want_tile_ >= 0 ? want_tile_ : default_tile_mode();
  psi_dim_ = ket.size;
  // complex plans (ComplexTensorContraction::MultiplyEffectHamilton,
  // complex_tensor_contraction.cc:1244-1312): same task structure, every
  // complex product expanded into real MFMA terms on interleaved views
  cplx_ = ket.cs == 2;
  std::vector<std::vector<int>> same_r(nr);
  for (int r = 0; r < nr; ++r) same_r[r] = ket.same(1, r);
  // block (l, r) -> index per L channel, the first match as FindMatrixBlock
  // returns it (real_matrix_block.cc:314-380), hashed once per plan instead
  // of a scan per term
  std::vector<BlockIndex> lidx;
  lidx.reserve(L.size());
  for (const BlockTensor& b : L) lidx.emplace_back(b);

  struct Out { int o1, lp; double w, wi; };
  struct Task { int o2, ri, kp, jdx; int64_t toff; std::vector<Out> outs; std::vector<int> ops; };
  std::vector<Task> tasks;
  int64_t tsize = 0;
  std::vector<char> usedL, usedR;
  for (int o2 = 0; o2 < W.rb; ++o2) {
    const BlockTensor& Rb = R[o2];
    for (int ri = 0; ri < Rb.n(); ++ri) {
      const Block& rbk = Rb.blk[ri];
      if (rbk.rows == 0) continue;
      const int ldx = rbk.l, jdx = rbk.r;
      for (int kp : same_r[ldx]) {
        const int kdx = ket.blk[kp].l;
        const int p2 = bphys[kp];
        if (!W.right_nonzero(o2, p2)) continue;
        Task t;
        t.o2 = o2; t.ri = ri; t.kp = kp; t.jdx = jdx;
        for (int op : same_r[jdx]) {
          const int idx = ket.blk[op].l;
          const int p1 = bphys[op];
          for (int o1 = 0; o1 < W.lb; ++o1) {
            if (W.used(o1, o2, p1, p2)) {
              const int lp = lidx[o1].find(idx, kdx);
              if (lp != -1) {
                t.outs.push_back(Out{o1, lp, W.at(o1, o2, p1, p2), W.im(o1, o2, p1, p2)});
                t.ops.push_back(op);
              }
            }
          }
        }
        const Block& kb = ket.blk[kp];
        require(kb.cols == rbk.rows, "Matrix indeces do not match in MultiplyToMatrix (K*R)");
        t.toff = tsize;
        if (!t.outs.empty()) tsize += (int64_t)kb.rows * rbk.cols;
        tasks.push_back(std::move(t));
      }
    }
  }
  const int cs = cplx_ ? 2 : 1;
  T_.resize((size_t)std::max<int64_t>(cs * tsize, 1));
  // stage-1 items that read one R block share a placement key (their tiles on
  // one XCD, the block fetched into one L2); DMRGCC_XCD_KEYS=0 turns it off
  std::vector<int> rbase(R.size() + 1, 0);
  for (size_t i = 0; i < R.size(); ++i) rbase[i + 1] = rbase[i] + R[i].n();
  const char* kev = std::getenv("DMRGCC_XCD_KEYS");
  const bool rkeys = !(kev && kev[0] == '0');
  auto rkey = [&](const Task& t) { return rkeys ? rbase[t.o2] + t.ri : -1; };
  // stage 1: T = K * R  (K relative to the input vector)
  for (const Task& t : tasks) {
    if (t.outs.empty()) continue;
    const Block& kb = ket.blk[t.kp];
    const BlockTensor& Rb = R[t.o2];
    const Block& rbk = Rb.blk[t.ri];
    const double* krel = reinterpret_cast<const double*>((intptr_t)kb.off);
    if (cplx_) {
      zgemm_item(s1_, T_.data() + cs * t.toff, rbk.cols, kb.rows, rbk.cols, 0, 0,
                 {ZTerm{krel, kb.cols, Rb.ptr(t.ri), rbk.cols, kb.cols, kRelA, false, false, 1.0, 0.0}});
      continue;
    }
    s1_.begin_item(T_.data() + t.toff, rbk.cols, kb.rows, rbk.cols, 0, 0, rkey(t));
    s1_.add(krel, kb.cols, Rb.ptr(t.ri), rbk.cols, kb.cols, kRelA, 1.0);
  }
  // stage 2: group terms by output block, keeping the reference's term order
  std::vector<std::vector<std::pair<int, int>>> by_out(ket.n());  // (task, term)
  for (size_t ti = 0; ti < tasks.size(); ++ti)
    for (size_t k = 0; k < tasks[ti].outs.size(); ++k) by_out[tasks[ti].ops[k]].push_back({(int)ti, (int)k});
  for (int op = 0; op < ket.n(); ++op) {
    if (by_out[op].empty()) continue;
    const Block& ob = ket.blk[op];
    double* orel = reinterpret_cast<double*>((intptr_t)ob.off);
    std::vector<ZTerm> zt;
    if (!cplx_) s2_.begin_item(orel, ob.cols, ob.rows, ob.cols, 1, kRelC);
    for (auto [ti, k] : by_out[op]) {
      const Task& t = tasks[ti];
      const Out& o = t.outs[k];
      const Block& lbk = L[o.o1].blk[o.lp];
      const Block& kb = ket.blk[t.kp];
      const Block& rbk = R[t.o2].blk[t.ri];
      require(lbk.rows == ob.rows && lbk.cols == kb.rows && rbk.cols == ob.cols,
              "Matrix indeces do not match in MultiplyToMatrix (L*T)");
      if (cplx_)
        zt.push_back(ZTerm{L[o.o1].ptr(o.lp), lbk.cols, T_.data() + cs * t.toff, rbk.cols, lbk.cols, 0,
                           false, false, o.w, o.wi});
      else
        s2_.add(L[o.o1].ptr(o.lp), lbk.cols, T_.data() + t.toff, rbk.cols, lbk.cols, 0, o.w);
    }
    if (cplx_) zgemm_item(s2_, orel, ob.cols, ob.rows, ob.cols, 1, kRelC, zt);
  }
  // Pre-split BF16x3 stages (real plans): the same items and terms on bf16
  // pair groups.  Offsets first (every used L and R block packed once, x per
  // ket block, T^T per task), then the buffers, then the descriptors.
  pre_ = false;
  p1_.clear();
  p2_.clear();
  {
    const char* ev = std::getenv("DMRGCC_PRESPLIT");
    const bool on = want_pre_ && !cplx_ && !(ev && ev[0] == '0');
    if (on) {
      auto groups = [](int K) { return (int64_t)((K + 7) / 8); };
      std::vector<std::vector<int64_t>> loff(L.size()), roff(R.size());
      for (size_t i = 0; i < L.size(); ++i) loff[i].assign(L[i].n(), -1);
      for (size_t i = 0; i < R.size(); ++i) roff[i].assign(R[i].n(), -1);
      int64_t lr = 0, xs = 0, ts = 0, lrmax = 0;
      std::vector<PackTask> lrpack, xpack;
      std::vector<int64_t> xoff(ket.n(), -1), toff_s(tasks.size(), -1);
      for (int kp = 0; kp < ket.n(); ++kp) {
        const Block& b = ket.blk[kp];
        if ((int64_t)b.rows * b.cols == 0) continue;
        xoff[kp] = xs;
        xs += (int64_t)b.rows * groups(b.cols) * 4;
      }
      for (size_t ti = 0; ti < tasks.size(); ++ti) {
        const Task& t = tasks[ti];
        if (t.outs.empty()) continue;
        const Block& kb = ket.blk[t.kp];
        const Block& rbk = R[t.o2].blk[t.ri];
        toff_s[ti] = ts;
        ts += (int64_t)rbk.cols * groups(kb.rows) * 4;
        if (roff[t.o2][t.ri] < 0) {
          roff[t.o2][t.ri] = lr;
          lr += (int64_t)rbk.cols * groups(rbk.rows) * 4;
        }
        for (const Out& o : t.outs)
          if (loff[o.o1][o.lp] < 0) {
            const Block& lbk = L[o.o1].blk[o.lp];
            loff[o.o1][o.lp] = lr;
            lr += (int64_t)lbk.rows * groups(lbk.cols) * 4;
          }
      }
      Xs_.ensure((size_t)std::max<int64_t>(xs, 1));
      Ts_.ensure((size_t)std::max<int64_t>(ts, 1));
      LRs_.ensure((size_t)std::max<int64_t>(lr, 1));
      fill(Ts_.data(), std::max<int64_t>(ts, 1), 0.0, stream);  // the padding k stay zero
      // pack tasks: the used R blocks as B^T (rows n, k = R rows), L as A
      for (size_t i = 0; i < R.size(); ++i)
        for (int b = 0; b < R[i].n(); ++b)
          if (roff[i][b] >= 0) {
            const Block& bk = R[i].blk[b];
            lrpack.push_back(PackTask{R[i].ptr(b), 1, bk.cols, bk.cols, bk.rows, LRs_.data() + roff[i][b]});
            lrmax = std::max<int64_t>(lrmax, (int64_t)bk.cols * groups(bk.rows));
          }
      for (size_t i = 0; i < L.size(); ++i)
        for (int b = 0; b < L[i].n(); ++b)
          if (loff[i][b] >= 0) {
            const Block& bk = L[i].blk[b];
            lrpack.push_back(PackTask{L[i].ptr(b), bk.cols, 1, bk.rows, bk.cols, LRs_.data() + loff[i][b]});
            lrmax = std::max<int64_t>(lrmax, (int64_t)bk.rows * groups(bk.cols));
          }
      int64_t xmax = 0;
      for (int kp = 0; kp < ket.n(); ++kp) {
        if (xoff[kp] < 0) continue;
        const Block& b = ket.blk[kp];
        xpack.push_back(PackTask{reinterpret_cast<const double*>((intptr_t)b.off), b.cols, 1, b.rows, b.cols,
                                 Xs_.data() + xoff[kp]});
        xmax = std::max<int64_t>(xmax, (int64_t)b.rows * groups(b.cols));
      }
      if (!lrpack.empty()) {
        upload(d_lrpack_, lrpack, stream);
        launch_pack(d_lrpack_.data(), (int)lrpack.size(), lrmax, nullptr, stream);
      }
      upload(d_xpack_, xpack, stream);
      nxpack_ = (int)xpack.size();
      xpack_groups_ = xmax;
      // stage 1: T^T (pairs) = K * R; stage 2: y += w L T, the same term order
      for (size_t ti = 0; ti < tasks.size(); ++ti) {
        const Task& t = tasks[ti];
        if (t.outs.empty()) continue;
        const Block& kb = ket.blk[t.kp];
        const Block& rbk = R[t.o2].blk[t.ri];
        if (kb.cols <= 0) continue;  // (no K: the FP64 T of such a task is zero as well)
        p1_.begin_item(Ts_.data() + toff_s[ti], groups(kb.rows), kb.rows, rbk.cols, 0, kPairC, rkey(t));
        p1_.add(Xs_.data() + xoff[t.kp], groups(kb.cols), LRs_.data() + roff[t.o2][t.ri], groups(rbk.rows),
                kb.cols, 0, 1.0);
      }
      for (int op = 0; op < ket.n(); ++op) {
        if (by_out[op].empty()) continue;
        const Block& ob = ket.blk[op];
        p2_.begin_item(reinterpret_cast<double*>((intptr_t)ob.off), ob.cols, ob.rows, ob.cols, 1, kRelC);
        for (auto [ti, k] : by_out[op]) {
          const Task& t = tasks[ti];
          const Out& o = t.outs[k];
          const Block& lbk = L[o.o1].blk[o.lp];
          const Block& kb = ket.blk[t.kp];
          if (lbk.cols <= 0) continue;  // empty terms are dropped (the kernel assumes K > 0)
          p2_.add(LRs_.data() + loff[o.o1][o.lp], groups(lbk.cols), Ts_.data() + toff_s[ti], groups(kb.rows),
                  lbk.cols, 0, o.w);
        }
      }
      pitems1_ = p1_.items;
      pterms1_ = p1_.contribs;
      pitems2_ = p2_.items;
      pterms2_ = p2_.contribs;
      for (GemmBatch* g : {&p1_, &p2_}) {
        g->presplit = true;
        g->tile_mode = 64;
        g->split_k = s2_.split_k;
        g->finalize(stream);
      }
      pre_ = true;
    }
  }
  // F_ref counts the reference's K*R products even when no L term uses them.
  double fref = 0.0;
  int64_t lbytes = 0, rbytes = 0;
  std::vector<std::vector<char>> uL(L.size()), uR(R.size());
  for (size_t i = 0; i < L.size(); ++i) uL[i].assign(L[i].n(), 0);
  for (size_t i = 0; i < R.size(); ++i) uR[i].assign(R[i].n(), 0);
  for (const Task& t : tasks) {
    const Block& kb = ket.blk[t.kp];
    const Block& rbk = R[t.o2].blk[t.ri];
    fref += 2.0 * kb.rows * (double)kb.cols * rbk.cols;
    if (!uR[t.o2][t.ri]) { uR[t.o2][t.ri] = 1; rbytes += (int64_t)rbk.rows * rbk.cols; }
    for (const Out& o : t.outs) {
      const Block& lbk = L[o.o1].blk[o.lp];
      fref += 2.0 * lbk.rows * (double)lbk.cols * rbk.cols;
      if (!uL[o.o1][o.lp]) { uL[o.o1][o.lp] = 1; lbytes += (int64_t)lbk.rows * lbk.cols; }
    }
  }
  // complex: 8 real flops per multiply-add and 16 bytes per entry
  fref_ = cplx_ ? 4.0 * fref : fref;
  bytes_min_ = 8 * cs * (2 * psi_dim_ + lbytes + rbytes);
  ++gen_;
  items1_ = s1_.items;
  terms1_ = s1_.contribs;
  items2_ = s2_.items;
  terms2_ = s2_.contribs;
  s1_.finalize(stream);
  s2_.finalize(stream);
}

void HeffPlan::retile(int tile_mode, int split_k, hipStream_t stream) {
  // the previous launches may still read the descriptors being replaced
  DMRG_HIP(hipStreamSynchronize(stream));
  ++gen_;
  GemmBatch* st[4] = {&s1_, &s2_, &p1_, &p2_};
  const std::vector<GemmItem>* it[4] = {&items1_, &items2_, &pitems1_, &pitems2_};
  const std::vector<GemmContrib>* tm[4] = {&terms1_, &terms2_, &pterms1_, &pterms2_};
  for (int k = 0; k < (pre_ ? 4 : 2); ++k) {
    GemmBatch& g = *st[k];
    g.items = *it[k];
    g.contribs = *tm[k];
    g.tile_mode = k >= 2 || tile_mode != 32 ? 64 : 32;  // the pre-split stages tile 64x64
    g.split_k = std::max(0, split_k);
    g.finalize(stream);
  }
}

std::vector<HeffPlan::Tuning> HeffPlan::tune(int precision, int reps, hipStream_t stream) {
  require(reps > 0, "tune: reps must be positive", kErrArgument);
  static const int kCand[][2] = {{64, 128}, {64, 256}, {64, 512}, {64, 1024}, {32, 256}, {32, 512}};
  const int64_t n = std::max<int64_t>(doubles(), 1);
  DevBuf<double> x((size_t)n), y((size_t)n);
  fill(x.data(), n, 1e-3, stream);
  fill(y.data(), n, 0.0, stream);
  hipEvent_t e0, e1;
  DMRG_HIP(hipEventCreate(&e0));
  DMRG_HIP(hipEventCreate(&e1));
  std::vector<Tuning> out;
  for (const auto& c : kCand) {
    retile(c[0], c[1], stream);
    apply(x.data(), y.data(), precision, stream);  // warm-up
    apply(x.data(), y.data(), precision, stream);
    DMRG_HIP(hipEventRecord(e0, stream));
    for (int r = 0; r < reps; ++r) apply(x.data(), y.data(), precision, stream);
    DMRG_HIP(hipEventRecord(e1, stream));
    DMRG_HIP(hipEventSynchronize(e1));
    float ms = 0;
    DMRG_HIP(hipEventElapsedTime(&ms, e0, e1));
    out.push_back(Tuning{c[0], c[1], ms / reps});
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  // the default tiling (64x64, split-K 256) stays unless a candidate beats it
  // by more than 3 %, so that close calls do not flip between runs
  size_t best = 1;
  for (size_t i = 0; i < out.size(); ++i)
    if (out[i].ms < out[best].ms) best = i;
  if (out[best].ms > 0.97 * out[1].ms) best = 1;
  retile(out[best].tile_mode, out[best].split_k, stream);
  return out;
}

void HeffPlan::apply(const double* x, double* y, int precision, hipStream_t stream) {
  if (use_pre(precision)) {
    launch_pack(d_xpack_.data(), nxpack_, xpack_groups_, x, stream);
    p1_.launch(stream, precision, nullptr, nullptr);
    p2_.launch(stream, precision, nullptr, y);
    return;
  }
  s1_.launch(stream, precision, x, nullptr);
  s2_.launch(stream, precision, nullptr, y);
}

void HeffPlan::apply_stage(int stage, const double* x, double* y, int precision,
                           hipStream_t stream, bool gemm_only) {
  if (use_pre(precision)) {
    if (stage == 1) {
      launch_pack(d_xpack_.data(), nxpack_, xpack_groups_, x, stream);
      p1_.launch(stream, precision, nullptr, nullptr, !gemm_only);
    } else {
      p2_.launch(stream, precision, nullptr, y, !gemm_only);
    }
    return;
  }
  if (stage == 1) s1_.launch(stream, precision, x, nullptr, !gemm_only);
  else s2_.launch(stream, precision, nullptr, y, !gemm_only);
}

// ===========================================================================
// Environment updates
// ===========================================================================
Env boundary_env(int nbond, Context& cx, int cs) {
  // DefineTensorContraction (real_tensor_contraction.cc:111-137; complex
  // complex_tensor_contraction.cc:111-137: the 1x1 block holds 1 + 0i)
  Env e(nbond);
  for (int i = 0; i < nbond; ++i) {
    e[i].define({0}, {0});
    e[i].blk[0].rows = e[i].blk[0].cols = 1;
    e[i].cs = cs;
    e[i].allocate();
    fill(e[i].data.data(), cs, 0.0, cx.stream);
    fill(e[i].data.data(), 1, 1.0, cx.stream);
  }
  return e;
}

namespace {

// Output block assembly shared by both environment updates: every block of
// out[o] collects (F index, ket block, w) terms; its shape comes from the
// terms (AddToMatrixBlock into an empty block takes the first term's shape).
struct EnvTerm { int f; int kp; double w, wi; };

void finish_env(Env& out, std::vector<std::vector<std::vector<EnvTerm>>>& terms,
                const std::vector<Shape>& fshape, const std::vector<int64_t>& foff,
                DevBuf<double>& F, const BlockTensor& ket, bool right, GemmBatch& s1, Context& cx,
                EnvTiming* tm) {
  GemmBatch s2;
  for (size_t o = 0; o < out.size(); ++o) {
    BlockTensor& bt = out[o];
    for (int b = 0; b < bt.n(); ++b) {
      Shape sh;
      for (const EnvTerm& t : terms[o][b]) {
        const Block& kb = ket.blk[t.kp];
        Shape s{fshape[t.f].rows, right ? kb.rows : kb.cols};
        if (s.size() == 0) continue;  // AddToMatrix ignores an empty term
        if (sh.size() == 0) sh = s;
        else require(sh.rows == s.rows && sh.cols == s.cols, "Matrix indeces do not match in AddToMatrix");
      }
      bt.blk[b].rows = sh.rows;
      bt.blk[b].cols = sh.cols;
    }
    bt.cs = ket.cs;
    bt.allocate();
    const bool cplx = ket.cs == 2;
    for (int b = 0; b < bt.n(); ++b) {
      const Block& ob = bt.blk[b];
      if ((int64_t)ob.rows * ob.cols == 0) continue;
      std::vector<ZTerm> zt;
      if (!cplx) s2.begin_item(bt.ptr(b), ob.cols, ob.rows, ob.cols, 0);
      for (const EnvTerm& t : terms[o][b]) {
        const Block& kb = ket.blk[t.kp];
        const Shape fs = fshape[t.f];
        double* Ff = F.data() + ket.cs * foff[t.f];
        if (!right) {
          // Lnew += w * F(i,l) * K(l,j)
          require(fs.cols == kb.rows, "Matrix indeces do not match in MultiplyToMatrix (F*K)");
          if (cplx) zt.push_back(ZTerm{Ff, fs.cols, ket.ptr(t.kp), kb.cols, fs.cols, 0, false, false, t.w, t.wi});
          else s2.add(Ff, fs.cols, ket.ptr(t.kp), kb.cols, fs.cols, 0, t.w);
        } else {
          // Rnew += w * F(j,k) * B(i,k)^T  (complex: B^H)
          require(fs.cols == kb.cols, "Matrix indeces do not match in MultiplyToMatrix (F*B^T)");
          if (cplx)
            zt.push_back(ZTerm{Ff, fs.cols, ket.ptr(t.kp), kb.cols, fs.cols, kTransB, false, true, t.w, t.wi});
          else s2.add(Ff, fs.cols, ket.ptr(t.kp), kb.cols, fs.cols, kTransB, t.w);
        }
      }
      if (cplx) zgemm_item(s2, bt.ptr(b), ob.cols, ob.rows, ob.cols, 0, 0, zt);
    }
  }
  s1.finalize(cx.stream);
  s2.finalize(cx.stream);
  if (!tm) {
    s1.launch(cx.stream, kFP64);
    s2.launch(cx.stream, kFP64);
    // no wait: F and the descriptors stay allocated until the next sync, so
    // the host goes on while the GPU computes the environment
    cx.retire(std::move(F));
    cx.retire(s1);
    cx.retire(s2);
    return;
  }
  hipEvent_t e[3];
  for (auto& ev : e) DMRG_HIP(hipEventCreate(&ev));
  DMRG_HIP(hipEventRecord(e[0], cx.stream));
  s1.launch(cx.stream, kFP64);
  DMRG_HIP(hipEventRecord(e[1], cx.stream));
  s2.launch(cx.stream, kFP64);
  DMRG_HIP(hipEventRecord(e[2], cx.stream));
  DMRG_HIP(hipEventSynchronize(e[2]));
  float a = 0, b = 0;
  DMRG_HIP(hipEventElapsedTime(&a, e[0], e[1]));
  DMRG_HIP(hipEventElapsedTime(&b, e[1], e[2]));
  for (auto& ev : e) (void)hipEventDestroy(ev);
  tm->ms_f = a;
  tm->ms_out = b;
  tm->flops = s1.flops + s2.flops;
  // bytes: the caller adds the old environment and ket blocks
  int64_t fsz = 0, osz = 0;
  for (const Shape& sh : fshape) fsz += sh.size();
  for (const BlockTensor& bt : out) osz += bt.size;
  tm->bytes_min = 8.0 * ket.cs * (double)osz;
  tm->bytes_exec = tm->bytes_min + 2 * 8.0 * ket.cs * (double)fsz;
  cx.sync();
}

// the used blocks of an old environment and of the ket, once each, in bytes
void env_input_bytes(EnvTiming* tm, const Env& E, const std::vector<std::pair<int, int>>& used_e,
                     const BlockTensor& ket, const std::vector<int>& used_k) {
  if (!tm) return;
  double b = 0;
  std::set<std::pair<int, int>> se(used_e.begin(), used_e.end());
  for (const auto& [o, a] : se) b += (double)E[o].blk[a].rows * E[o].blk[a].cols;
  std::set<int> sk(used_k.begin(), used_k.end());
  for (int k : sk) b += (double)ket.blk[k].rows * ket.blk[k].cols;
  tm->bytes_min += 8.0 * ket.cs * b;
  tm->bytes_exec += 8.0 * ket.cs * b;
}

}  // namespace

Env left_env_update(const Lattice& lat, const SiteOperator& W, const Env& L, Context& cx, EnvTiming* tm) {
  // LeftComputeTensorContraction (real_tensor_contraction.cc:782-940)
  const BlockTensor& ket = lat.ket;
  const int nr = lat.nr();
  std::vector<std::vector<int>> same_l(lat.nl());
  for (int l = 0; l < lat.nl(); ++l) same_l[l] = ket.same(0, l);
  std::vector<char> flag((size_t)W.rb * nr * nr, 0);
  for (int o1 = 0; o1 < W.lb; ++o1) {
    const BlockTensor& Lb = L[o1];
    for (int a = 0; a < Lb.n(); ++a) {
      if (Lb.blk[a].rows == 0) continue;
      const int kdx = Lb.blk[a].l, ldx = Lb.blk[a].r;
      for (int kp : same_l[kdx]) {
        const int idx = ket.blk[kp].r;
        const int p1 = lat.phys[kdx][idx];
        if (!W.left_nonzero(o1, p1)) continue;
        for (int lp : same_l[ldx]) {
          const int jdx = ket.blk[lp].r;
          const int p2 = lat.phys[ldx][jdx];
          for (int o2 = 0; o2 < W.rb; ++o2)
            if (W.used(o1, o2, p1, p2)) flag[((size_t)o2 * nr + idx) * nr + jdx] = 1;
        }
      }
    }
  }
  Env out(W.rb);
  std::vector<std::vector<std::vector<EnvTerm>>> terms(W.rb);
  for (int o2 = 0; o2 < W.rb; ++o2) {
    std::vector<int> li, ri;
    for (int i = 0; i < nr; ++i)
      for (int j = 0; j < nr; ++j)
        if (flag[((size_t)o2 * nr + i) * nr + j]) { li.push_back(i); ri.push_back(j); }
    out[o2].define(li, ri);
    terms[o2].resize(li.size());
  }
  std::vector<BlockIndex> oidx;
  oidx.reserve(out.size());
  for (const BlockTensor& b : out) oidx.emplace_back(b);
  // F = B(k,i)^T * L[o1](k,l) for each (o1, L block, ket block)
  std::vector<Shape> fshape;
  std::vector<int64_t> foff;
  struct FTask { int o1, a, kp; };
  std::vector<FTask> ft;
  int64_t fsize = 0;
  for (int o1 = 0; o1 < W.lb; ++o1) {
    const BlockTensor& Lb = L[o1];
    for (int a = 0; a < Lb.n(); ++a) {
      if (Lb.blk[a].rows == 0) continue;
      const int kdx = Lb.blk[a].l, ldx = Lb.blk[a].r;
      for (int kp : same_l[kdx]) {
        const int idx = ket.blk[kp].r;
        const int p1 = lat.phys[kdx][idx];
        if (!W.left_nonzero(o1, p1)) continue;
        require(ket.blk[kp].rows == Lb.blk[a].rows, "Matrix indeces do not match in MultiplyToMatrix (B^T*L)");
        const int f = (int)ft.size();
        ft.push_back(FTask{o1, a, kp});
        fshape.push_back(Shape{ket.blk[kp].cols, Lb.blk[a].cols});
        foff.push_back(fsize);
        fsize += fshape.back().size();
        for (int lp : same_l[ldx]) {
          const int jdx = ket.blk[lp].r;
          const int p2 = lat.phys[ldx][jdx];
          for (int o2 = 0; o2 < W.rb; ++o2) {
            if (W.used(o1, o2, p1, p2)) {
              const int pos = oidx[o2].find(idx, jdx);
              terms[o2][pos].push_back(EnvTerm{f, lp, W.at(o1, o2, p1, p2), W.im(o1, o2, p1, p2)});
            }
          }
        }
      }
    }
  }
  DevBuf<double> F((size_t)std::max<int64_t>(ket.cs * fsize, 1));
  GemmBatch s1;
  for (size_t f = 0; f < ft.size(); ++f) {
    const Block& kb = ket.blk[ft[f].kp];
    const BlockTensor& Lb = L[ft[f].o1];
    const Block& lbk = Lb.blk[ft[f].a];
    if (fshape[f].size() == 0) continue;
    double* Ff = F.data() + ket.cs * foff[f];
    if (ket.cs == 2) {  // F = B^H * L  (complex_tensor_contraction.cc:929)
      zgemm_item(s1, Ff, fshape[f].cols, fshape[f].rows, fshape[f].cols, 0, 0,
                 {ZTerm{ket.ptr(ft[f].kp), kb.cols, Lb.ptr(ft[f].a), lbk.cols, kb.rows, kTransA, true,
                        false, 1.0, 0.0}});
      continue;
    }
    s1.begin_item(Ff, fshape[f].cols, fshape[f].rows, fshape[f].cols, 0);
    s1.add(ket.ptr(ft[f].kp), kb.cols, Lb.ptr(ft[f].a), lbk.cols, kb.rows, kTransA, 1.0);
  }
  finish_env(out, terms, fshape, foff, F, ket, false, s1, cx, tm);
  if (tm) {
    std::vector<std::pair<int, int>> ue;
    std::vector<int> uk;
    for (const FTask& t : ft) { ue.push_back({t.o1, t.a}); uk.push_back(t.kp); }
    for (const auto& to : terms)
      for (const auto& tb : to)
        for (const EnvTerm& t : tb) uk.push_back(t.kp);
    env_input_bytes(tm, L, ue, ket, uk);
  }
  return out;
}

Env right_env_update(const Lattice& lat, const SiteOperator& W, const Env& R, Context& cx, EnvTiming* tm) {
  // RightComputeTensorContraction (real_tensor_contraction.cc:959-1119)
  const BlockTensor& ket = lat.ket;
  const int nl = lat.nl();
  std::vector<std::vector<int>> same_r(lat.nr());
  for (int r = 0; r < lat.nr(); ++r) same_r[r] = ket.same(1, r);
  std::vector<char> flag((size_t)W.lb * nl * nl, 0);
  for (int o2 = 0; o2 < W.rb; ++o2) {
    const BlockTensor& Rb = R[o2];
    for (int a = 0; a < Rb.n(); ++a) {
      if (Rb.blk[a].rows == 0) continue;
      const int ldx = Rb.blk[a].l, kdx = Rb.blk[a].r;
      for (int lp : same_r[ldx]) {
        const int jdx = ket.blk[lp].l;
        const int p2 = lat.phys[jdx][ldx];
        if (!W.right_nonzero(o2, p2)) continue;
        for (int kp : same_r[kdx]) {
          const int idx = ket.blk[kp].l;
          const int p1 = lat.phys[idx][kdx];
          for (int o1 = 0; o1 < W.lb; ++o1)
            if (W.used(o1, o2, p1, p2)) flag[((size_t)o1 * nl + jdx) * nl + idx] = 1;
        }
      }
    }
  }
  Env out(W.lb);
  std::vector<std::vector<std::vector<EnvTerm>>> terms(W.lb);
  for (int o1 = 0; o1 < W.lb; ++o1) {
    std::vector<int> li, ri;
    for (int j = 0; j < nl; ++j)
      for (int i = 0; i < nl; ++i)
        if (flag[((size_t)o1 * nl + j) * nl + i]) { li.push_back(j); ri.push_back(i); }
    out[o1].define(li, ri);
    terms[o1].resize(li.size());
  }
  std::vector<BlockIndex> oidx;
  oidx.reserve(out.size());
  for (const BlockTensor& b : out) oidx.emplace_back(b);
  std::vector<Shape> fshape;
  std::vector<int64_t> foff;
  struct FTask { int o2, a, lp; };
  std::vector<FTask> ft;
  int64_t fsize = 0;
  for (int o2 = 0; o2 < W.rb; ++o2) {
    const BlockTensor& Rb = R[o2];
    for (int a = 0; a < Rb.n(); ++a) {
      if (Rb.blk[a].rows == 0) continue;
      const int ldx = Rb.blk[a].l, kdx = Rb.blk[a].r;
      for (int lp : same_r[ldx]) {
        const int jdx = ket.blk[lp].l;
        const int p2 = lat.phys[jdx][ldx];
        if (!W.right_nonzero(o2, p2)) continue;
        require(ket.blk[lp].cols == Rb.blk[a].rows, "Matrix indeces do not match in MultiplyToMatrix (K*R)");
        const int f = (int)ft.size();
        ft.push_back(FTask{o2, a, lp});
        fshape.push_back(Shape{ket.blk[lp].rows, Rb.blk[a].cols});
        foff.push_back(fsize);
        fsize += fshape.back().size();
        for (int kp : same_r[kdx]) {
          const int idx = ket.blk[kp].l;
          const int p1 = lat.phys[idx][kdx];
          for (int o1 = 0; o1 < W.lb; ++o1) {
            if (W.used(o1, o2, p1, p2)) {
              const int pos = oidx[o1].find(jdx, idx);
              terms[o1][pos].push_back(EnvTerm{f, kp, W.at(o1, o2, p1, p2), W.im(o1, o2, p1, p2)});
            }
          }
        }
      }
    }
  }
  DevBuf<double> F((size_t)std::max<int64_t>(ket.cs * fsize, 1));
  GemmBatch s1;
  for (size_t f = 0; f < ft.size(); ++f) {
    const Block& kb = ket.blk[ft[f].lp];
    const BlockTensor& Rb = R[ft[f].o2];
    const Block& rbk = Rb.blk[ft[f].a];
    if (fshape[f].size() == 0) continue;
    double* Ff = F.data() + ket.cs * foff[f];
    if (ket.cs == 2) {
      zgemm_item(s1, Ff, fshape[f].cols, fshape[f].rows, fshape[f].cols, 0, 0,
                 {ZTerm{ket.ptr(ft[f].lp), kb.cols, Rb.ptr(ft[f].a), rbk.cols, kb.cols, 0, false, false,
                        1.0, 0.0}});
      continue;
    }
    s1.begin_item(Ff, fshape[f].cols, fshape[f].rows, fshape[f].cols, 0);
    s1.add(ket.ptr(ft[f].lp), kb.cols, Rb.ptr(ft[f].a), rbk.cols, kb.cols, 0, 1.0);
  }
  finish_env(out, terms, fshape, foff, F, ket, true, s1, cx, tm);
  if (tm) {
    std::vector<std::pair<int, int>> ue;
    std::vector<int> uk;
    for (const FTask& t : ft) { ue.push_back({t.o2, t.a}); uk.push_back(t.lp); }
    for (const auto& to : terms)
      for (const auto& tb : to)
        for (const EnvTerm& t : tb) uk.push_back(t.kp);
    env_input_bytes(tm, R, ue, ket, uk);
  }
  return out;
}

// ===========================================================================
// RealTensorRundmrg
// ===========================================================================
RealTensorRundmrg::RealTensorRundmrg(const RealTensorSpace& space, const RealTensorHamiltonian& ham,
                                     const RunParams& p, hipStream_t stream)
    : space_(space), ham_(ham), p_(p), cx_(stream) {
  L_sites_ = space.num_site();
  d_ = space.physics_dim();
  cs_ = ham.complex() ? 2 : 1;
  require(ham.num_site() == L_sites_, "space and Hamiltonian disagree on the number of sites",
          kErrArgument);
  tables_ = space.DefineQuantumTable();
  lat_.clear();
  lat_.resize(L_sites_);
  for (auto& l : lat_) l.d = d_;
  L_.clear();
  L_.resize(L_sites_ + 1);
  R_.clear();
  R_.resize(L_sites_ + 1);
}

int RealTensorRundmrg::check(int site, int lq, int rq) const {
  return space_.CheckQuantumTable(site, tables_[site][lq], tables_[site + 1][rq]);
}

void RealTensorRundmrg::lattice_block(int site, std::vector<int>& left, std::vector<int>& right) const {
  // ComputeLatticeBlock (real_tensor_space.cc:442-504)
  if (site == 0) left = {0};
  else left = lat_[site - 1].rb;
  right.clear();
  const int nt = (int)tables_[site + 1].size();
  for (int r = 0; r < nt; ++r) {
    for (int l : left)
      if (check(site, l, r) != -1) {
        right.push_back(r);
        break;
      }
    if ((int)right.size() == max_block_) break;
  }
}

void RealTensorRundmrg::tensor_index(int site, const Lattice& lat, std::vector<int>& li,
                                     std::vector<int>& ri, std::vector<int>& pi) const {
  // ComputeTensorIndex (real_tensor_space.cc:506-550)
  li.clear(); ri.clear(); pi.clear();
  for (int i = 0; i < lat.nl(); ++i)
    for (int j = 0; j < lat.nr(); ++j) {
      const int p = check(site, lat.lb[i], lat.rb[j]);
      if (p != -1) { li.push_back(i); ri.push_back(j); pi.push_back(p); }
    }
}

void RealTensorRundmrg::random_blocks(Lattice& lat, const std::vector<int>& li,
                                      const std::vector<int>& ri, const std::vector<int>& pi,
                                      int rows, int cols) {
  // DefineTensorLattice (real_tensor_lattice.cc:343-359; complex
  // complex_tensor_lattice.cc:343-359)
  lat.ket.define(li, ri);
  lat.ket.cs = cs_;
  std::vector<double> host;
  for (size_t k = 0; k < li.size(); ++k) {
    lat.phys[li[k]][ri[k]] = pi[k];
    lat.ket.blk[k].rows = rows;
    lat.ket.blk[k].cols = cols;
    const size_t o = host.size();
    host.resize(o + (size_t)cs_ * rows * cols);
    if (cs_ == 2) cx_.rng.complex_matrix(rows, cols, host.data() + o);
    else cx_.rng.random_matrix(rows, cols, host.data() + o);
  }
  lat.ket.allocate();
  if (!host.empty())
    DMRG_HIP(hipMemcpyAsync(lat.ket.data.data(), host.data(), host.size() * sizeof(double),
                            hipMemcpyHostToDevice, cx_.stream));
  cx_.sync();
}

void RealTensorRundmrg::DefineTensorSpace(int first_block, int first_dim) {
  // DefineTensorSpace (real_tensor_space.cc:68-134)
  max_block_ = first_block;
  max_dim_ = -1;
  noise_ = 10;
  canpre_ = 1e-30;
  for (int i = 0; i < L_sites_; ++i) {
    std::vector<int> left, right;
    lattice_block(i, left, right);
    std::vector<int> ldim(left.size(), i == 0 ? 1 : first_dim);
    std::vector<int> rdim(right.size(), i == L_sites_ - 1 ? 1 : first_dim);
    lat_[i].define_sectors(left, right, ldim, rdim);
  }
  for (int i = 0; i < L_sites_; ++i) {
    std::vector<int> li, ri, pi;
    tensor_index(i, lat_[i], li, ri, pi);
    random_blocks(lat_[i], li, ri, pi, i == 0 ? 1 : first_dim, i == L_sites_ - 1 ? 1 : first_dim);
    if (p_.verbose >= 2) trace(-2, i, 0.0, "define_random");
    canonical(0, i);
    if (p_.verbose >= 2) trace(-2, i, 0.0, "define_canon");
    if (i > 0) merge(0, i);
    if (p_.verbose >= 2) trace(-2, i, 0.0, "define_merge");
  }
}

// --------------------------------------------------------------------------
// canonical forms: Left/RightCanonicalTensorLattice
// (real_tensor_lattice.cc:558-676, :794-911) with the per-sector SVD on the
// device (one-sided Jacobi) and the truncation rule on the host.
// --------------------------------------------------------------------------
double RealTensorRundmrg::canonical(int leigh, int site) {
  return canonicalize(lat_[site], leigh, max_dim_, canpre_, cx_, &ph_);
}

double canonicalize(Lattice& lat, int leigh, int max_dim, double canpre, Context& cx, PhaseTimes* ph) {
  PhaseTimes scratch_times;
  PhaseTimes& ph_ = ph ? *ph : scratch_times;
  Context& cx_ = cx;
  const int max_dim_ = max_dim;
  const double canpre_ = canpre;
  BlockTensor& ket = lat.ket;
  const bool left = leigh == 0;
  const int cs = ket.cs;
  const int ns = left ? lat.nr() : lat.nl();
  struct Sec {
    std::vector<int> pos;
    int rows = 0, cols = 0;  // stacked matrix A
    bool svd = false, trans = false;
    int M = 0, N = 0;
    int64_t g = 0, v = 0, s = 0;
  };
  std::vector<Sec> sec(ns);
  int64_t wsize = 0;
  std::vector<CopyTask> copies;
  for (int s = 0; s < ns; ++s) {
    Sec& S = sec[s];
    S.pos = ket.same(left ? 1 : 0, s);
    if (left) {
      S.cols = lat.rd[s];
      for (int p : S.pos) {
        S.rows += ket.blk[p].rows;
        require(ket.blk[p].cols <= S.cols, "Matrix indices are out of range in set_matrix_element");
      }
      S.svd = S.rows > 0;
    } else {
      S.rows = lat.ld[s];
      for (int p : S.pos) {
        S.cols += ket.blk[p].cols;
        require(ket.blk[p].rows <= S.rows, "Matrix indices are out of range in set_matrix_element");
      }
      S.svd = S.cols > 0;
    }
    if (!S.svd) continue;
    require(S.rows > 0 && S.cols > 0, "empty sector matrix in SVD");
    S.trans = S.rows < S.cols;
    S.M = std::max(S.rows, S.cols);
    S.N = std::min(S.rows, S.cols);
    S.g = wsize; wsize += (int64_t)cs * S.M * S.N;  // doubles
    S.v = wsize; wsize += (int64_t)cs * S.N * S.N;
  }
  // sigma[N] and |A|_F^2 of every sector, contiguous after the matrices so
  // that one copy brings all singular values to the host
  const int64_t sig0 = wsize;
  for (int s = 0; s < ns; ++s) {
    Sec& S = sec[s];
    if (!S.svd) continue;
    S.s = wsize;
    wsize += S.N + 1;
  }
  cx_.work.ensure((size_t)std::max<int64_t>(wsize, 1));
  double* w = cx_.work.data();
  std::vector<SvdProblem> probs;
  // zero the stacked matrices first (ragged blocks leave gaps); a separate
  // launch, since copies and fills of one buffer must not race
  fill(w, wsize, 0.0, cx_.stream);
  for (int s = 0; s < ns; ++s) {
    Sec& S = sec[s];
    if (!S.svd) continue;
    int acc = 0;
    for (int p : S.pos) {
      const Block& b = ket.blk[p];
      if ((int64_t)b.rows * b.cols == 0) {
        acc += left ? b.rows : b.cols;
        continue;
      }
      CopyTask c{ket.ptr(p), nullptr, b.rows, b.cols, b.cols, 1, 0, 0, 1.0};
      // element (i, j) of block p sits at A(acc+i, j) (left) or A(i, acc+j) (right)
      const int64_t M = S.M;
      if (left) {
        if (!S.trans) { c.dst = w + S.g + (int64_t)cs * acc; c.dsr = 1; c.dsc = M; }
        else { c.dst = w + S.g + (int64_t)cs * acc * M; c.dsr = M; c.dsc = 1; }
        acc += b.rows;
      } else {
        if (!S.trans) { c.dst = w + S.g + (int64_t)cs * acc * M; c.dsr = 1; c.dsc = M; }
        else { c.dst = w + S.g + (int64_t)cs * acc; c.dsr = M; c.dsc = 1; }
        acc += b.cols;
      }
      push_copy(copies, c, cs);
    }
    // QR preconditioning that keeps exact the factor read as orthonormal:
    // G (left canonical of a tall sector, right canonical of a wide one) or V
    const int precond = (left && !S.trans) || (!left && S.trans) ? 1 : 2;
    probs.push_back(SvdProblem{w + S.g, w + S.v, w + S.s, w + S.s + S.N, S.M, S.N, S.trans ? 1 : 0,
                               precond});
  }
  launch_copies(copies, cx_.copy_scratch, cx_.stream);
  const double tsvd = now();
  // complex: ComplexMatrix::SVDMatrix (complex_matrix.cc:601-637)
  if (cs == 2) zjacobi_svd(probs, cx_.svd_scratch, cx_.svd_sweeps, cx_.stream);
  else jacobi_svd(probs, cx_.svd_scratch, cx_.svd_sweeps, cx_.stream);
  cx_.sync();
  ph_.svd += now() - tsvd;
  for (const SvdProblem& q : probs) {
    ph_.svd_max_m = std::max<int64_t>(ph_.svd_max_m, q.M);
    ph_.svd_max_n = std::max<int64_t>(ph_.svd_max_n, q.N);
    ph_.svd_work += (double)q.M * q.N * q.N;
  }
  ph_.svd_calls += (int64_t)probs.size();
  // singular values to the host, sorted descending per sector
  std::vector<std::vector<double>> sv(ns), strue(ns);
  std::vector<std::vector<int>> perm(ns);
  {
    std::vector<double> sig((size_t)std::max<int64_t>(wsize - sig0, 1));
    if (wsize > sig0)
      DMRG_HIP(hipMemcpyAsync(sig.data(), w + sig0, (wsize - sig0) * sizeof(double), hipMemcpyDeviceToHost,
                              cx_.stream));
    cx_.sync();
    for (int s = 0; s < ns; ++s) {
      Sec& S = sec[s];
      if (!S.svd) {
        sv[s] = {0.0};
        strue[s] = {0.0};
        perm[s] = {0};
        continue;
      }
      strue[s].assign(sig.begin() + (S.s - sig0), sig.begin() + (S.s - sig0) + S.N);
    }
    for (int s = 0; s < ns; ++s) {
      if (!sec[s].svd) continue;
      std::vector<double>& v = strue[s];
      std::vector<int> ord(v.size());
      std::iota(ord.begin(), ord.end(), 0);
      std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return v[a] > v[b]; });
      std::vector<double> sorted(v.size());
      for (size_t k = 0; k < ord.size(); ++k) sorted[k] = v[ord[k]];
      perm[s] = ord;
      strue[s] = sorted;
      sv[s] = sorted;
    }
  }
  std::vector<int> t = truncate_dims(max_dim_, canpre_, sv);
  double discarded = 0.0;
  for (int s = 0; s < ns; ++s)
    for (size_t j = t[s]; j < strue[s].size(); ++j) discarded += strue[s][j] * strue[s][j];

  // device arrays: permutations and scales
  std::vector<int> hperm;
  std::vector<double> hscale;
  struct Off { int64_t perm, su, sc; };
  std::vector<Off> off(ns);
  for (int s = 0; s < ns; ++s) {
    if (!sec[s].svd) continue;
    const Sec& S = sec[s];
    off[s].perm = hperm.size();
    for (int j = 0; j < t[s]; ++j) hperm.push_back(perm[s][j]);
    // scale for the orthonormal factor taken from G*V (1/sigma, NaN -> unit vector)
    off[s].su = hscale.size();
    for (int j = 0; j < t[s]; ++j) {
      const double st = strue[s][j];
      hscale.push_back(st > 0 ? 1.0 / st : std::numeric_limits<double>::quiet_NaN());
    }
    // scale for the canonical factor: sigma (from V) or sigma/sigma_true (from G*V)
    off[s].sc = hscale.size();
    const bool from_gv = (left && S.trans) || (!left && !S.trans);
    for (int j = 0; j < t[s]; ++j) {
      const double st = strue[s][j], sf = sv[s][j];
      if (from_gv) hscale.push_back(st > 0 ? sf / st : 0.0);
      else hscale.push_back(sf);
    }
    // unit-scale for factors taken straight from V
    (void)S;
  }
  const int64_t ones_off = hscale.size();
  int tmax = 1;
  for (int s = 0; s < ns; ++s) tmax = std::max(tmax, t[s]);
  for (int j = 0; j < tmax; ++j) hscale.push_back(1.0);
  upload(cx_.ivec, hperm.empty() ? std::vector<int>{0} : hperm, cx_.stream);
  upload(cx_.dvec, hscale, cx_.stream);
  const int* dperm = cx_.ivec.data();
  const double* dscale = cx_.dvec.data();

  // new ket and canonical tensors
  BlockTensor nk;
  nk.blk = ket.blk;
  nk.cs = cs;
  BlockTensor can;
  can.cs = cs;
  {
    std::vector<int> idx(ns);
    std::iota(idx.begin(), idx.end(), 0);
    can.define(idx, idx);
  }
  lat.match_dim.assign(ns, 0);
  for (int s = 0; s < ns; ++s) {
    const Sec& S = sec[s];
    if (S.svd)
      for (int p : S.pos) {
        if (left) nk.blk[p].cols = t[s];
        else nk.blk[p].rows = t[s];
      }
    if (left) {
      can.blk[s].rows = t[s];
      can.blk[s].cols = S.cols;  // right_singular_tensor column count
      lat.match_dim[s] = S.cols;
      lat.rd[s] = t[s];
    } else {
      can.blk[s].rows = S.rows;
      can.blk[s].cols = t[s];
      lat.match_dim[s] = S.rows;
      lat.ld[s] = t[s];
    }
  }
  nk.allocate();
  can.allocate();
  std::vector<CopyTask> cps;
  std::vector<ColGatherTask> gts;
  // sectors without an SVD keep their (empty) ket blocks; canonical is zero
  for (int s = 0; s < ns; ++s) {
    const Sec& S = sec[s];
    if (!S.svd) {
      for (int p : S.pos)
        if (nk.elems(p) > 0)
          push_copy(cps, CopyTask{ket.ptr(p), nk.ptr(p), nk.blk[p].rows, nk.blk[p].cols,
                                  nk.blk[p].cols, 1, nk.blk[p].cols, 1, 1.0}, cs);
      if (can.elems(s) > 0)
        push_copy(cps, CopyTask{nullptr, can.ptr(s), 1, (int)can.elems(s), 0, 0, 0, 1, 0.0}, cs);
      continue;
    }
    const double* G = w + S.g;
    const double* V = w + S.v;
    const int* pp = dperm + off[s].perm;
    const double* su = dscale + off[s].su;
    const double* sc = dscale + off[s].sc;
    const double* one = dscale + ones_off;
    int acc = 0;
    for (int p : S.pos) {
      const Block& b = nk.blk[p];
      const int len = left ? ket.blk[p].rows : ket.blk[p].cols;
      ColGatherTask g{};
      g.perm = pp;
      g.ncol = t[s];
      g.r0 = acc;
      g.rows = len;
      g.dst = nk.ptr(p);
      if (left) {  // ket block (len x t): row i, column jj
        g.dsr = b.cols; g.dsc = 1;
        if (!S.trans) { g.src = G; g.ld = S.M; g.scale = su; }
        else { g.src = V; g.ld = S.N; g.scale = one; }
      } else {  // ket block (t x len): column c = i, row jj
        g.dsr = 1; g.dsc = b.cols;
        if (!S.trans) { g.src = V; g.ld = S.N; g.scale = one; }
        else { g.src = G; g.ld = S.M; g.scale = su; }
      }
      // complex: A = U Sigma W^H (G = A) or A = conj(W) Sigma U'^T (G = A^T),
      // so every factor read from V = W is conjugated
      if (cs == 2) g.flags = kGatherComplex | (g.src == V ? kGatherConj : 0);
      gts.push_back(g);
      acc += len;
    }
    ColGatherTask c{};
    c.perm = pp;
    c.ncol = t[s];
    c.r0 = 0;
    c.dst = can.ptr(s);
    c.scale = sc;
    if (left) {  // canonical t x cols: row jj = sigma * v^T
      c.dsr = 1; c.dsc = S.cols; c.rows = S.cols;
      if (!S.trans) { c.src = V; c.ld = S.N; }
      else { c.src = G; c.ld = S.M; }
    } else {  // canonical rows x t: column jj = u * sigma
      c.dsr = t[s]; c.dsc = 1; c.rows = S.rows;
      if (!S.trans) { c.src = G; c.ld = S.M; }
      else { c.src = V; c.ld = S.N; }
    }
    if (cs == 2) c.flags = kGatherComplex | (c.src == V ? kGatherConj : 0);
    gts.push_back(c);
  }
  launch_copies(cps, cx_.copy_scratch, cx_.stream);
  launch_col_gather(gts, cx_.gather_scratch, cx_.stream);
  cx_.retire(std::move(ket.data));  // read by the copies just queued
  ket = std::move(nk);
  lat.can = std::move(can);
  lat.has_can = true;
  return discarded;
}

// Left/RightMergeTensorLattice (real_tensor_lattice.cc:914-977)
void RealTensorRundmrg::merge(int leigh, int site) {
  Lattice* lat;
  Lattice* nb;
  if (leigh == 0 && site > 0) { lat = &lat_[site]; nb = &lat_[site - 1]; }
  else if (leigh == 1 && site < L_sites_ - 1) { lat = &lat_[site]; nb = &lat_[site + 1]; }
  else return;
  require(nb->has_can, leigh == 0 ? "Canonical tensor is nullptr in LeftMergeTensorLattice!"
                                  : "Canonical tensor is nullptr in RightMergeTensorLattice!");
  BlockTensor& ket = lat->ket;
  const BlockTensor& can = nb->can;
  BlockTensor nk;
  nk.blk = ket.blk;
  nk.cs = ket.cs;
  const int ns = leigh == 0 ? lat->nl() : lat->nr();
  require(can.n() >= ns, "canonical tensor has fewer sectors than the merged site");
  for (int p = 0; p < ket.n(); ++p) {
    const Block& b = ket.blk[p];
    if (leigh == 0) {
      const Block& c = can.blk[b.l];
      require(c.cols == b.rows, "Matrix indeces do not match in MultiplyToMatrix (merge)");
      nk.blk[p].rows = c.rows;
    } else {
      const Block& c = can.blk[b.r];
      require(b.cols == c.rows, "Matrix indeces do not match in MultiplyToMatrix (merge)");
      nk.blk[p].cols = c.cols;
    }
  }
  nk.allocate();
  GemmBatch gb;
  for (int p = 0; p < ket.n(); ++p) {
    const Block& b = ket.blk[p];
    const Block& o = nk.blk[p];
    if ((int64_t)o.rows * o.cols == 0) continue;
    if (nk.cs == 2) {  // ComplexTensorLattice merge: complex products
      if (leigh == 0) {
        const Block& c = can.blk[b.l];
        zgemm_item(gb, nk.ptr(p), o.cols, o.rows, o.cols, 0, 0,
                   {ZTerm{can.ptr(b.l), c.cols, ket.ptr(p), b.cols, c.cols, 0, false, false, 1.0, 0.0}});
      } else {
        const Block& c = can.blk[b.r];
        zgemm_item(gb, nk.ptr(p), o.cols, o.rows, o.cols, 0, 0,
                   {ZTerm{ket.ptr(p), b.cols, can.ptr(b.r), c.cols, b.cols, 0, false, false, 1.0, 0.0}});
      }
      continue;
    }
    gb.begin_item(nk.ptr(p), o.cols, o.rows, o.cols, 0);
    if (leigh == 0) {
      const Block& c = can.blk[b.l];
      gb.add(can.ptr(b.l), c.cols, ket.ptr(p), b.cols, c.cols, 0, 1.0);
    } else {
      const Block& c = can.blk[b.r];
      gb.add(ket.ptr(p), b.cols, can.ptr(b.r), c.cols, b.cols, 0, 1.0);
    }
  }
  gb.finalize(cx_.stream);
  gb.launch(cx_.stream, kFP64);
  for (int s = 0; s < ns; ++s) {
    if (leigh == 0) lat->ld[s] = can.blk[s].rows;
    else lat->rd[s] = can.blk[s].cols;
  }
  // NormalizeTensorLattice (real_matrix_block.cc:164-171; complex: sum of
  // |z|^2 over the interleaved doubles, complex_matrix_block.cc:163-170)
  double* sc = cx_.scal.data();
  dot(nk.data.data(), nk.data.data(), nk.doubles(), sc, cx_.partial, cx_.stream);
  scale_inv_sqrt(nk.data.data(), nk.doubles(), sc, cx_.stream, 1);
  // no wait here: the replaced ket and canonical tensor stay allocated until
  // the next sync, for the queued GEMMs that read them
  cx_.retire(std::move(ket.data));
  ket = std::move(nk);
  nb->has_can = false;
  cx_.retire(std::move(nb->can.data));
  cx_.retire(gb);
  nb->can = BlockTensor{};
  nb->match_dim.clear();
}

void RealTensorRundmrg::Initialize() {
  // Initialize (real_tensor_rundmrg.cc:247-287)
  L_[0] = boundary_env(ham_.sites()[0].lb, cx_, cs_);
  R_[L_sites_ - 1] = boundary_env(ham_.sites()[L_sites_ - 1].rb, cx_, cs_);
  for (int i = L_sites_ - 1; i > 0; --i) {
    canonical(1, i);
    if (p_.verbose >= 2) trace(-1, i, 0.0, "init_canon");
    R_[i - 1] = right_env_update(lat_[i], ham_.sites()[i], R_[i], cx_);
    merge(1, i - 1);
    if (p_.verbose >= 2) trace(-1, i - 1, 0.0, "init_merge");
  }
  canonical(1, 0);
}

// --------------------------------------------------------------------------
// Lanczos (RealTensorLanczos::LanczosMethod, real_tensor_lanczos.cc:14-169)
// --------------------------------------------------------------------------
std::unique_ptr<HeffPlan> RealTensorRundmrg::make_plan(int site) {
  auto plan = std::make_unique<HeffPlan>();
  const double t0 = now();
  const int64_t n = lat_[site].ket.size;
  const bool tune = p_.autotune && n >= p_.tune_min_dim && n >= 4 * tuned_psi_;
  if (!tune) plan->set_tiling(tune_tile_, tune_split_);
  plan->set_presplit(p_.precision == kBF16x3);  // the packed operands serve BF16x3 only
  plan->build(lat_[site], ham_.sites()[site], L_[site], R_[site], cx_.stream);
  if (tune) {
    // the plan ends up tiled with the fastest candidate; the trial matvecs
    // run on scratch vectors, the state is untouched
    plan->tune(p_.precision, 3, cx_.stream);
    tune_tile_ = plan->tile_mode() == 32 ? 32 : 64;
    tune_split_ = plan->split_k();
    tuned_psi_ = n;
    ++stats_.tunes;
    stats_.tile_mode = tune_tile_;
    stats_.split_k = tune_split_;
  }
  ph_.plan += now() - t0;
  return plan;
}

double RealTensorRundmrg::LanczosMethod(int num_iter, int site) {
  // a plan built ahead for this site by the previous step (same lattice,
  // operator and environments: nothing changes them in between)
  std::unique_ptr<HeffPlan> plan;
  if (next_plan_ && next_plan_site_ == site) plan = std::move(next_plan_);
  else plan = make_plan(site);
  next_plan_.reset();
  next_plan_site_ = -1;
  return lanczos_solve(*plan, lat_[site].ket.data.data(), num_iter, p_.precision, p_.refine, cx_,
                       stats_);
}

double lanczos_solve(HeffPlan& plan, double* ket, int num_iter, int precision, int refine,
                     Context& cx, LanczosStats& st, int* iters) {
  // Real (RealTensorLanczos::LanczosMethod, real_tensor_lanczos.cc:14-169)
  // and complex (ComplexTensorLanczos, complex_tensor_lanczos.cc:14-160)
  // forms.  Complex vectors are n interleaved pairs (nd = 2n doubles): alpha
  // and the Gram-Schmidt overlaps are complex, beta and the tridiagonal
  // eigenproblem real (zheevd reads only the real diagonal).
  const int64_t n = plan.psi_dim();
  const bool cz = plan.complex();
  const int64_t nd = plan.doubles();
  ++st.solves;
  if (n < 100) {
    // dense H_eff (ComputeEffectHamilton, real_tensor_contraction.cc:1136-1205)
    // built column by column from the FP64 device matvec, then the FP64 host
    // eigensolver (RealSymMatrixDiag; a complex Hermitian H through its real
    // embedding [[Re, -Im], [Im, Re]], whose lowest eigenvector (x, y) gives
    // the ground state x + i y)
    ++st.dense;
    if (iters) *iters = 0;
    if (n == 0) return 0.0;
    DevBuf<double> X((size_t)n * nd), Y((size_t)n * nd);
    std::vector<double> eye((size_t)n * nd, 0.0);
    for (int64_t i = 0; i < n; ++i) eye[(size_t)i * nd + (cz ? 2 * i : i)] = 1.0;
    DMRG_HIP(hipMemcpyAsync(X.data(), eye.data(), eye.size() * sizeof(double), hipMemcpyHostToDevice,
                            cx.stream));
    fill(Y.data(), n * nd, 0.0, cx.stream);
    for (int64_t c = 0; c < n; ++c) plan.apply(X.data() + c * nd, Y.data() + c * nd, kFP64, cx.stream);
    std::vector<double> y((size_t)n * nd);
    DMRG_HIP(hipMemcpyAsync(y.data(), Y.data(), y.size() * sizeof(double), hipMemcpyDeviceToHost, cx.stream));
    cx.sync();
    const int64_t m = cz ? 2 * n : n;
    std::vector<double> H((size_t)m * m);
    for (int64_t i = 0; i < n; ++i)
      for (int64_t c = 0; c < n; ++c) {
        if (!cz) {
          H[(size_t)i * n + c] = y[(size_t)c * n + i];
        } else {
          const double re = y[(size_t)c * nd + 2 * i], im = y[(size_t)c * nd + 2 * i + 1];
          H[(size_t)i * m + c] = re;
          H[(size_t)(i + n) * m + c + n] = re;
          H[(size_t)(i + n) * m + c] = im;
          H[(size_t)i * m + c + n] = -im;
        }
      }
    double e0 = 0.0;
    std::vector<double> v0;
    sym_eig_lowest((int)m, H, e0, v0);
    std::vector<double> g(nd);
    for (int64_t i = 0; i < n; ++i) {
      if (!cz) {
        g[i] = v0[i];
      } else {
        g[2 * i] = v0[i];
        g[2 * i + 1] = v0[i + n];
      }
    }
    DMRG_HIP(hipMemcpyAsync(ket, g.data(), nd * sizeof(double), hipMemcpyHostToDevice, cx.stream));
    cx.sync();
    return e0;
  }
  require(num_iter >= 1 && num_iter <= 200, "num_iter must be in [1, 200]", kErrArgument);
  // Device scalars: per iteration i, slot sc[4 i ..] = (alpha re, im, beta^2);
  // then the Ritz coefficients and the refinement's two dots.  The host reads
  // each slot through the context's pinned Lanczos buffer.
  constexpr int kSlot = 4;
  cx.scal.ensure((size_t)kSlot * num_iter + num_iter + 8);
  double* sc = cx.scal.data();
  double* coef_dev = sc + kSlot * num_iter;
  double* rq_dev = coef_dev + num_iter;
  double* host = cx.lanczos_host();
  double* host_dev = cx.lanczos_host_dev();
  DevBuf<double> psi((size_t)num_iter * nd), h((size_t)nd);
  double* P = psi.data();
  DMRG_HIP(hipMemcpyAsync(P, ket, nd * sizeof(double), hipMemcpyDeviceToDevice, cx.stream));
  fill(h.data(), nd, 0.0, cx.stream);
  std::vector<double> alpha, beta;
  double old_energy = 10000.0, energy = 0.0;
  // partial-sum buffers of the fused vector steps: part_a holds the alpha
  // partials, part_b the beta^2 partials, part_m the overlaps of h with the
  // basis for the next vector's Gram-Schmidt
  cx.partial.ensure(2 * 2 * kDotPartials + (size_t)(2 * num_iter) * kDotPartials);
  double* part_a = cx.partial.data();
  double* part_b = part_a + 2 * kDotPartials;
  double* part_m = part_b + 2 * kDotPartials;
  // Slot i's ready flag (host[kSlot i + 3]) is set by the launch that
  // publishes iteration i's coefficients, after them; the host polls it
  // instead of recording an event in the stream (an event between two
  // launches idles the GPU for several microseconds).
  for (int i = 0; i < num_iter; ++i) host[kSlot * i + 3] = 0.0;
  // psi_{i+1} = h / beta_i, h = -beta_i psi_i, then GramSchmidtMethod
  // (real_tensor_lanczos.cc:140-151, :188-201; complex_tensor_lanczos.cc:
  // 178-192), which projects out psi_0 .. psi_i and normalises: one launch,
  // lz_next_all, from the overlaps <psi_j|h> that iteration i's
  // lz_alpha_multi formed (c_j = <psi_j|h> / beta_i, r^2 = 1 - sum |c_j|^2).
  // For the orthonormal basis this is the reference's vector: its sequential
  // coefficients differ by terms of second order in the overlaps, which the
  // recurrence leaves at the rounding level.  beta_i is formed on the device.
  // With `publish`, the launch also forms alpha_i and beta_i^2 from iteration
  // i's partials and writes them to slot i and to pinned host memory with the
  // ready flag: iteration i needs no launch of its own for its coefficients,
  // and no copy.
  auto next_vector = [&](int i, bool publish) {
    double* pn = P + (int64_t)(i + 1) * nd;
    lz_next_all(pn, h.data(), P + (int64_t)i * nd, P, nd, i + 1, n, cz, part_m, cx.stream, nullptr, part_a, part_b,
                publish ? sc + kSlot * i : nullptr, publish ? host_dev + kSlot * i : nullptr);
  };
  // the coefficients of iteration i by a launch of their own: (alpha,
  // beta^2) to slot i and to pinned host memory, with the ready flag
  auto publish_alone = [&](int i) {
    lz_finals(part_a, part_b, n, cz, sc + kSlot * i, cx.stream, host_dev + kSlot * i);
  };
  // wait for slot i's flag; a stream that fails or drains without setting it
  // ends the wait with an error
  auto wait_published = [&](int i) {
    const volatile double* flag = host + kSlot * i + 3;
    for (uint64_t spin = 1; *flag != 1.0; ++spin) {
      if ((spin & 0xfff) == 0) {
        const hipError_t q = hipStreamQuery(cx.stream);
        if (q == hipSuccess) {
          require(*flag == 1.0, "lanczos: the stream drained without publishing the coefficients",
                  kErrInternal);
          break;
        }
        if (q != hipErrorNotReady) DMRG_HIP(q);
      }
      __builtin_ia32_pause();
    }
  };
  // iteration i on the stream: psi_i (when i > 0, publishing iteration i-1's
  // coefficients unless `blocking`), h += H psi_i, alpha = <psi_i|h>,
  // h -= alpha psi_i, beta^2 = <h|h> (as partials)
  bool blocking = false;
  auto launch_iteration = [&](int i) {
    if (i > 0) next_vector(i - 1, !blocking);
    double* pi = P + (int64_t)i * nd;
    plan.apply(pi, h.data(), precision, cx.stream);
    lz_dot_partial(pi, h.data(), n, cz, part_a, cx.stream);   // alpha = <psi|H|psi>
    // h -= alpha psi; partials of <h|h> and of <psi_j|h>, j <= i
    lz_alpha_multi(h.data(), pi, P, nd, i + 1, n, cz, part_a, part_b, part_m, cx.stream);
    if (blocking) publish_alone(i);
  };
  blocking = std::getenv("DMRGCC_LANCZOS_BLOCKING") != nullptr;
  launch_iteration(0);
  for (int i = 0; i < num_iter; ++i) {
    const int k = i + 1;
    const double t0 = now();
    if (!blocking) {
      if (k < num_iter) launch_iteration(k);  // publishes iteration i first
      else publish_alone(i);
    }
    wait_published(i);
    st.matvec_seconds += now() - t0;
    st.flops += plan.fref();
    ++st.matvecs;
    const double* ab = host + kSlot * i;
    alpha.push_back(ab[0]);
    beta.push_back(std::sqrt(ab[2]));
    std::vector<double> T((size_t)k * k, 0.0), ev, vec;
    for (int j = 0; j < k; ++j) {
      T[(size_t)j * k + j] = alpha[j];
      if (j > 0) T[(size_t)j * k + j - 1] = T[(size_t)(j - 1) * k + j] = beta[j - 1];
    }
    sym_eig(k, T, ev, vec);
    const double e_new = ev[0];
    // FP64 keeps the reference's stop (beta < 1e-15, |dE| < 1e-15, niter;
    // real_tensor_lanczos.cc:110-112).  A mixed-precision run converges to
    // the eigenvector of the rounded operator, whose FP64 Rayleigh quotient
    // has an error floor of order (operand precision)^2 |H|; Ritz-value
    // changes below eta |E| cannot lower the refined energy, so the loop
    // stops there (eta: 1e-13 FP32, 1e-10 split BF16x3 with its ~16-bit
    // operands) instead of running to niter against an unreachable 1e-15.
    const double eta = precision == kFP32 ? 1e-13 : precision == kFP64 ? 0.0 : 1e-10;
    const double de_tol = std::max(1e-15, eta * std::fabs(e_new));
    if (beta[i] < 1e-15 || std::fabs(e_new - old_energy) < de_tol || k == num_iter) {
      energy = e_new;
      if (iters) *iters = k;
      double* coef = host + kSlot * num_iter;
      for (int j = 0; j < k; ++j) coef[j] = vec[(size_t)j * k];
      DMRG_HIP(hipMemcpyAsync(coef_dev, coef, k * sizeof(double), hipMemcpyHostToDevice, cx.stream));
      combine(ket, P, nd, coef_dev, k, nd, cx.stream);
      if (precision != kFP64 && refine) {
        // FP64 Rayleigh-quotient refinement of the mixed-precision Ritz pair
        fill(h.data(), nd, 0.0, cx.stream);
        plan.apply(ket, h.data(), kFP64, cx.stream);
        dot(ket, h.data(), nd, rq_dev, cx.partial, cx.stream);   // Re <ket|H ket>
        dot(ket, ket, nd, rq_dev + 1, cx.partial, cx.stream);
        DMRG_HIP(hipMemcpyAsync(host, rq_dev, 2 * sizeof(double), hipMemcpyDeviceToHost, cx.stream));
        cx.sync();
        st.flops += plan.fref();
        ++st.matvecs;
        energy = host[0] / host[1];
      }
      cx.sync();
      break;
    }
    old_energy = e_new;
    if (blocking && k < num_iter) launch_iteration(k);
  }
  return energy;
}

// --------------------------------------------------------------------------
// Subspace expansion
// --------------------------------------------------------------------------
std::vector<std::vector<int>> RealTensorRundmrg::expan_lattice(int leigh, int site,
                                                               const QNTable& op_table, Lattice& E) {
  // ComputeExpanTensorLattice (real_tensor_space.cc:552-706); leigh 1 grows
  // the right bond, leigh 0 the left bond.
  const Lattice& lat = lat_[site];
  const QNTable& tab = tables_[site + leigh];
  const std::vector<int>& old = leigh == 0 ? lat.lb : lat.rb;
  std::vector<char> flag(tab.size(), 0);
  std::vector<std::vector<QN>> merged(op_table.size());
  for (size_t o = 0; o < op_table.size(); ++o)
    for (int ob : old) {
      QN m = space_.MergeQuantumTable(op_table[o], tab[ob]);
      merged[o].push_back(m);
      for (size_t k = 0; k < tab.size(); ++k)
        if (tab[k] == m) flag[k] = 1;
    }
  std::vector<int> blocks;
  for (size_t k = 0; k < tab.size(); ++k)
    if (flag[k]) {
      blocks.push_back((int)k);
      if ((int)blocks.size() == max_block_) break;
    }
  std::vector<std::vector<int>> mapping(op_table.size());
  for (size_t o = 0; o < op_table.size(); ++o)
    for (const QN& m : merged[o]) {
      int hit = -1;
      for (size_t k = 0; k < blocks.size(); ++k)
        if (tab[blocks[k]] == m) { hit = (int)k; break; }
      mapping[o].push_back(hit);
    }
  E = Lattice{};
  E.d = d_;
  if (leigh == 0)
    E.define_sectors(blocks, lat.rb, std::vector<int>(blocks.size(), 1), std::vector<int>(lat.rb.size(), 1));
  else
    E.define_sectors(lat.lb, blocks, std::vector<int>(lat.lb.size(), 1), std::vector<int>(blocks.size(), 1));
  std::vector<int> li, ri, pi;
  tensor_index(site, E, li, ri, pi);
  // DefineTensorLattice(..., 1, 1) draws one rand() per block, then the
  // blocks are reset to empty
  for (size_t k = 0; k < li.size(); ++k) E.phys[li[k]][ri[k]] = pi[k];
  cx_.rng.skip((int64_t)cs_ * li.size());
  E.ket.define(li, ri);
  E.ket.cs = cs_;
  return mapping;
}

void RealTensorRundmrg::expan_contraction(int leigh, int site, Lattice& E,
                                          const std::vector<std::vector<int>>& mapping, ExpanWork* defer) {
  // leigh 0: LeftExpanTensorContraction (real_tensor_contraction.cc:169-459)
  // leigh 1: RightExpanTensorContraction (real_tensor_contraction.cc:477-765)
  const Lattice& lat = lat_[site];
  const BlockTensor& ket = lat.ket;
  const SiteOperator& W = ham_.sites()[site];
  const bool left = leigh == 0;
  const bool active = expansion_active();
  const int d = lat.d;
  // ---- which expansion sectors receive a term ----
  const int ne = left ? E.nr() : E.nl();
  std::vector<char> exist(ne, 0);
  std::vector<std::vector<int>> same(left ? lat.nl() : lat.nr());
  for (size_t s = 0; s < same.size(); ++s) same[s] = ket.same(left ? 0 : 1, (int)s);
  const Env& env = left ? L_[site] : R_[site];
  const int nouter = left ? W.lb : W.rb;
  const int ninner = left ? W.rb : W.lb;
  const BlockIndex eidx0(E.ket);
  auto wat = [&](int outer, int inner, int p1, int p2) {
    return left ? W.at(outer, inner, p1, p2) : W.at(inner, outer, p1, p2);
  };
  auto wim = [&](int outer, int inner, int p1, int p2) {
    return left ? W.im(outer, inner, p1, p2) : W.im(inner, outer, p1, p2);
  };
  const int cs = cs_;
  for (int ob = 0; ob < nouter; ++ob) {
    const BlockTensor& Eb = env[ob];
    for (int a = 0; a < Eb.n(); ++a) {
      if (Eb.blk[a].rows == 0) continue;
      const int idx = left ? Eb.blk[a].l : Eb.blk[a].r;
      const int kdx = left ? Eb.blk[a].r : Eb.blk[a].l;
      for (int kp : same[kdx]) {
        const int jdx = left ? ket.blk[kp].r : ket.blk[kp].l;
        const int p2 = left ? lat.phys[kdx][jdx] : lat.phys[jdx][kdx];
        if (left ? !W.mixed_nonzero(ob, p2) : !W.right_nonzero(ob, p2)) continue;
        for (int ib = 0; ib < ninner; ++ib) {
          const int ldx = mapping[ib][jdx];
          const int pos = left ? eidx0.find(idx, ldx) : eidx0.find(ldx, idx);
          if (ldx != -1 && pos != -1) {
            const int p1 = left ? E.phys[idx][ldx] : E.phys[ldx][idx];
            if (wat(ob, ib, p1, p2) != 0.0 || wim(ob, ib, p1, p2) != 0.0) exist[ldx] = 1;
          }
        }
      }
    }
  }
  // ---- rebuild the expansion lattice with the existing sectors only ----
  std::vector<int> keep, leigh_block;
  for (int i = 0; i < ne; ++i)
    if (exist[i]) { keep.push_back(i); leigh_block.push_back(left ? E.rb[i] : E.lb[i]); }
  std::vector<std::vector<int>> etab(ninner);
  for (int ib = 0; ib < ninner; ++ib)
    for (size_t r = 0; r < mapping[ib].size(); ++r) {
      const int ldx = mapping[ib][r];
      int hit = -1;
      if (ldx != -1)
        for (size_t i = 0; i < leigh_block.size(); ++i)
          if ((left ? E.rb[ldx] : E.lb[ldx]) == leigh_block[i]) { hit = (int)i; break; }
      etab[ib].push_back(hit);
    }
  {
    std::vector<std::vector<int>> ph;
    if (left) {
      for (int j = 0; j < E.nl(); ++j) {
        std::vector<int> row;
        for (int i : keep) row.push_back(E.phys[j][i]);
        ph.push_back(row);
      }
      E.rb = leigh_block;
      E.ld = lat.ld;
      E.rd.assign(leigh_block.size(), 0);
    } else {
      for (int i : keep) ph.push_back(E.phys[i]);
      E.lb = leigh_block;
      E.rd = lat.rd;
      E.ld.assign(leigh_block.size(), 0);
    }
    E.phys = ph;
    std::vector<int> li, ri;
    for (int i = 0; i < E.nl(); ++i)
      for (int j = 0; j < E.nr(); ++j)
        if (E.phys[i][j] != -1) { li.push_back(i); ri.push_back(j); }
    E.ket.define(li, ri);
  }
  const BlockIndex eidx(E.ket);
  // ---- enumerate the expansion terms, drawing the random matrices ----
  struct First { int ob, a, kp; Shape sh; int64_t off; };
  struct Term { int f; int pos; int64_t roff; int e0, e1; int64_t at; };
  std::vector<First> firsts;
  std::vector<Term> terms;
  int64_t rnd_size = 0;           // doubles of the random matrices
  std::vector<RandomDraw> draws;  // the random matrices, in draw order
  int64_t pending = 0;            // draws of an inactive expansion
  std::vector<Shape> eshape(E.ket.n());
  int64_t fsize = 0;
  for (int ob = 0; ob < nouter; ++ob) {
    const BlockTensor& Eb = env[ob];
    for (int a = 0; a < Eb.n(); ++a) {
      if (Eb.blk[a].rows == 0) continue;
      const int idx = left ? Eb.blk[a].l : Eb.blk[a].r;
      const int kdx = left ? Eb.blk[a].r : Eb.blk[a].l;
      for (int kp : same[kdx]) {
        const int jdx = left ? ket.blk[kp].r : ket.blk[kp].l;
        const int p2 = left ? lat.phys[kdx][jdx] : lat.phys[jdx][kdx];
        if (left ? !W.mixed_nonzero(ob, p2) : !W.right_nonzero(ob, p2)) continue;
        First f{ob, a, kp, Shape{}, fsize};
        if (left) {
          require(Eb.blk[a].cols == ket.blk[kp].rows, "Matrix indeces do not match in MultiplyToMatrix (L*K)");
          f.sh = Shape{Eb.blk[a].rows, ket.blk[kp].cols};
        } else {
          require(ket.blk[kp].cols == Eb.blk[a].rows, "Matrix indeces do not match in MultiplyToMatrix (K*R)");
          f.sh = Shape{ket.blk[kp].rows, Eb.blk[a].cols};
        }
        const int fi = (int)firsts.size();
        firsts.push_back(f);
        fsize += f.sh.size();
        // left: random (D_jdx x w), w from D_jdx = first cols;
        // right: random (w x D_jdx), w from D_jdx = first rows
        const int Dj = left ? f.sh.cols : f.sh.rows;
        const int wdt = expan_width(Dj, d);
        const int r_rows = left ? Dj : wdt, r_cols = left ? wdt : Dj;
        for (int ib = 0; ib < ninner; ++ib) {
          const int ldx = etab[ib][jdx];
          const int pos = left ? eidx.find(idx, ldx) : eidx.find(ldx, idx);
          if (ldx != -1 && pos != -1) {
            const int p1 = left ? E.phys[idx][ldx] : E.phys[ldx][idx];
            const double w = wat(ob, ib, p1, p2), wi = wim(ob, ib, p1, p2);
            if (w != 0.0 || wi != 0.0) {
              if (!active) {
                pending += (int64_t)cs * r_rows * r_cols;
                continue;
              }
              Term t;
              t.f = fi;
              t.pos = pos;
              t.e0 = r_rows;
              t.e1 = r_cols;
              t.roff = rnd_size;  // doubles
              rnd_size += (int64_t)cs * r_rows * r_cols;
              draws.push_back(RandomDraw{t.roff, r_rows, r_cols, w, wi});
              // product shape and its place in the (concatenated) block
              Shape ps = left ? Shape{f.sh.rows, r_cols} : Shape{r_rows, f.sh.cols};
              Shape& es = eshape[pos];
              if (es.rows == 0) {  // AddToMatrix into an empty block
                t.at = 0;
                if (ps.size() > 0) es = ps;
              } else if (left) {   // ExpanMatrix(1, .): append columns
                require(es.rows == ps.rows, "row or column dimension is not match in ExpanMatrix");
                t.at = es.cols;
                es.cols += ps.cols;
              } else {             // ExpanMatrix(0, .): append rows
                require(es.cols == ps.cols, "row or column dimension is not match in ExpanMatrix");
                t.at = es.rows;
                es.rows += ps.rows;
              }
              terms.push_back(t);
            }
          }
        }
      }
    }
  }
  if (!active) {  // ExpanTensorSpace returns early: E is never used
    cx_.rng.skip(pending);
    return;
  }
  // ---- dims of the expansion sectors; zero blocks for empty ones ----
  if (left) {
    for (int r = 0; r < E.nr(); ++r) {
      E.rd[r] = 0;
      for (int l = 0; l < E.nl(); ++l) {
        const int pos = eidx.find(l, r);
        if (pos != -1) E.rd[r] = std::max(E.rd[r], eshape[pos].cols);
      }
      for (int l = 0; l < E.nl(); ++l) {
        const int pos = eidx.find(l, r);
        if (pos != -1 && eshape[pos].rows == 0) {
          std::printf("row of expan_tensor is zero in LeftExpanTensorContraction\n");
          if (cs == 2) {  // complex_tensor_contraction.cc:446-451
            eshape[pos] = Shape{E.ld[l], E.rd[r]};
            continue;
          }
          // the reference sizes this zero block by left_dim_[r] (:448)
          require(r < E.nl(), "zero expansion block with r >= num_left_block (reference reads out of range)");
          eshape[pos] = Shape{E.ld[r], E.rd[r]};
        } else if (pos != -1 && cs == 2) {
          // complex pads short blocks with zero columns (complex_tensor_contraction.cc:453-462)
          eshape[pos].cols = std::max(eshape[pos].cols, E.rd[r]);
        }
      }
    }
  } else {
    for (int l = 0; l < E.nl(); ++l) {
      E.ld[l] = 0;
      for (int r = 0; r < E.nr(); ++r) {
        const int pos = eidx.find(l, r);
        if (pos != -1) E.ld[l] = std::max(E.ld[l], eshape[pos].rows);
      }
      for (int r = 0; r < E.nr(); ++r) {
        const int pos = eidx.find(l, r);
        if (pos != -1 && eshape[pos].rows == 0) {
          std::printf("row of expan_tensor is zero in RightExpanTensorContraction\n");
          if (cs == 2) {  // complex_tensor_contraction.cc:762-767
            eshape[pos] = Shape{E.ld[l], E.rd[r]};
            continue;
          }
          require(r < E.nl(), "zero expansion block with r >= num_left_block (reference reads out of range)");
          eshape[pos] = Shape{E.ld[r], E.rd[r]};
        } else if (pos != -1 && cs == 2) {
          // complex pads short blocks with zero rows (complex_tensor_contraction.cc:769-778)
          eshape[pos].rows = std::max(eshape[pos].rows, E.ld[l]);
        }
      }
    }
  }
  for (int p = 0; p < E.ket.n(); ++p) {
    E.ket.blk[p].rows = eshape[p].rows;
    E.ket.blk[p].cols = eshape[p].cols;
  }
  E.ket.cs = cs;
  E.ket.allocate();
  fill(E.ket.data.data(), E.ket.doubles(), 0.0, cx_.stream);
  // ---- device: first = L*K (or K*R), then noise * first * random ----
  DevBuf<double> F((size_t)std::max<int64_t>(cs * fsize, 1)), Rd((size_t)std::max<int64_t>(rnd_size, 1));
  // the random matrices, drawn on the device (the draws of fill_random)
  random_matrices_device(draws, cs, Rd.data(), cx_);
  GemmBatch s1, s2;
  std::vector<char> fused(firsts.size(), 0);
  for (const Term& t : terms) fused[t.f] = 1;
  for (size_t fi = 0; fi < firsts.size(); ++fi) {
    if (!fused[fi]) continue;
    const First& f = firsts[fi];
    if (f.sh.size() == 0) continue;
    const BlockTensor& Eb = env[f.ob];
    const Block& eb = Eb.blk[f.a];
    const Block& kb = ket.blk[f.kp];
    if (cs == 2) {
      const ZTerm zt = left ? ZTerm{Eb.ptr(f.a), eb.cols, ket.ptr(f.kp), kb.cols, eb.cols, 0, false, false, 1.0, 0.0}
                            : ZTerm{ket.ptr(f.kp), kb.cols, Eb.ptr(f.a), eb.cols, kb.cols, 0, false, false, 1.0, 0.0};
      zgemm_item(s1, F.data() + cs * f.off, f.sh.cols, f.sh.rows, f.sh.cols, 0, 0, {zt});
      continue;
    }
    s1.begin_item(F.data() + f.off, f.sh.cols, f.sh.rows, f.sh.cols, 0);
    if (left) s1.add(Eb.ptr(f.a), eb.cols, ket.ptr(f.kp), kb.cols, eb.cols, 0, 1.0);
    else s1.add(ket.ptr(f.kp), kb.cols, Eb.ptr(f.a), eb.cols, kb.cols, 0, 1.0);
  }
  for (const Term& t : terms) {
    const First& f = firsts[t.f];
    const Block& eb = E.ket.blk[t.pos];
    if (left) {  // cols [at, at+e1) of the block: first (Di x Dj) * rnd (Dj x e1)
      if ((int64_t)f.sh.rows * t.e1 == 0) continue;
      if (cs == 2) {
        zgemm_item(s2, E.ket.ptr(t.pos) + cs * t.at, eb.cols, f.sh.rows, t.e1, 0, 0,
                   {ZTerm{F.data() + cs * f.off, f.sh.cols, Rd.data() + t.roff, t.e1, f.sh.cols, 0, false,
                          false, noise_, 0.0}});
        continue;
      }
      s2.begin_item(E.ket.ptr(t.pos) + t.at, eb.cols, f.sh.rows, t.e1, 0);
      s2.add(F.data() + f.off, f.sh.cols, Rd.data() + t.roff, t.e1, f.sh.cols, 0, noise_);
    } else {     // rows [at, at+e0): rnd (e0 x Dj) * first (Dj x Di)
      if ((int64_t)t.e0 * f.sh.cols == 0) continue;
      if (cs == 2) {
        zgemm_item(s2, E.ket.ptr(t.pos) + cs * t.at * eb.cols, eb.cols, t.e0, f.sh.cols, 0, 0,
                   {ZTerm{Rd.data() + t.roff, t.e1, F.data() + cs * f.off, f.sh.cols, t.e1, 0, false, false,
                          noise_, 0.0}});
        continue;
      }
      s2.begin_item(E.ket.ptr(t.pos) + t.at * eb.cols, eb.cols, t.e0, f.sh.cols, 0);
      s2.add(Rd.data() + t.roff, t.e1, F.data() + f.off, f.sh.cols, t.e1, 0, noise_);
    }
  }
  s1.finalize(cx_.stream);
  s2.finalize(cx_.stream);
  if (defer) {
    defer->s1 = std::move(s1);
    defer->s2 = std::move(s2);
    defer->F = std::move(F);
    defer->Rd = std::move(Rd);
    defer->pending = true;
    return;
  }
  s1.launch(cx_.stream, kFP64);
  s2.launch(cx_.stream, kFP64);
  // no wait: the buffers stay allocated until the next sync (the SVD's)
  cx_.retire(std::move(F));
  cx_.retire(std::move(Rd));
  cx_.retire(s1);
  cx_.retire(s2);
}

void RealTensorRundmrg::launch_expansion(ExpanWork& w) {
  if (!w.pending) return;
  w.s1.launch(cx_.stream, kFP64);
  w.s2.launch(cx_.stream, kFP64);
  cx_.retire(std::move(w.F));
  cx_.retire(std::move(w.Rd));
  cx_.retire(w.s1);
  cx_.retire(w.s2);
  w.pending = false;
}

namespace {

// Assemble a new ket from pieces (expansion / match fills).
void assemble(BlockTensor& nk, const std::vector<std::vector<Piece>>& pieces, Context& cx) {
  nk.allocate();
  fill(nk.data.data(), nk.doubles(), 0.0, cx.stream);
  std::vector<CopyTask> cps;
  for (int p = 0; p < nk.n(); ++p)
    for (const Piece& pc : pieces[p]) {
      if (!pc.src || (int64_t)pc.rows * pc.cols == 0) continue;
      push_copy(cps, CopyTask{pc.src, nk.ptr(p) + nk.cs * ((int64_t)pc.r0 * nk.blk[p].cols + pc.c0), pc.rows,
                              pc.cols, pc.ssr, 1, nk.blk[p].cols, 1, 1.0}, nk.cs);
    }
  launch_copies(cps, cx.copy_scratch, cx.stream);
}

// LeftExpanTensorLattice / RightExpanTensorLattice
// (real_tensor_lattice.cc:979-1149): block (l, r) of the new lattice is the
// origin block, extended by the expansion block (columns for the left fill,
// rows for the right fill) or zero-padded to the new sector dimension.
void fill_expanded(Lattice& lat, const Lattice& origin, const Lattice* E, bool left_fill, Context& cx) {
  BlockTensor nk;
  nk.blk = lat.ket.blk;
  nk.cs = origin.ket.cs;
  std::vector<std::vector<Piece>> pieces(nk.n());
  for (int p = 0; p < nk.n(); ++p) {
    const int l = nk.blk[p].l, r = nk.blk[p].r;
    int e0, e1 = -1;
    if (left_fill) {
      e0 = index_of(origin.rb, lat.rb[r]);
      if (E) e1 = index_of(E->rb, lat.rb[r]);
    } else {
      e0 = index_of(origin.lb, lat.lb[l]);
      if (E) e1 = index_of(E->lb, lat.lb[l]);
    }
    Shape m;
    auto add_piece = [&](const BlockTensor& bt, int q) {
      const Block& b = bt.blk[q];
      if ((int64_t)b.rows * b.cols == 0) return;  // AddToMatrix of an empty matrix
      if (m.size() == 0) {
        m = Shape{b.rows, b.cols};
        pieces[p].push_back(Piece{bt.ptr(q), b.rows, b.cols, b.cols, 0, 0});
      } else {
        require(m.rows == b.rows && m.cols == b.cols, "Matrix indeces do not match in AddToMatrix");
        throw Error(kErrInternal, "expansion fill: AddToMatrix onto a non-empty block");
      }
    };
    if (e0 != -1) {
      const int q = left_fill ? origin.ket.find(l, e0) : origin.ket.find(e0, r);
      if (q != -1) add_piece(origin.ket, q);
    }
    if (e1 != -1) {
      const int q = left_fill ? E->ket.find(l, e1) : E->ket.find(e1, r);
      if (q != -1) {
        const Block& b = E->ket.blk[q];
        if (m.rows == 0) {
          add_piece(E->ket, q);
        } else if (left_fill) {
          require(m.rows == b.rows, "row or column dimension is not match in ExpanMatrix");
          if ((int64_t)b.rows * b.cols > 0) pieces[p].push_back(Piece{E->ket.ptr(q), b.rows, b.cols, b.cols, 0, m.cols});
          m.cols += b.cols;
        } else {
          require(m.cols == b.cols, "row or column dimension is not match in ExpanMatrix");
          if ((int64_t)b.rows * b.cols > 0) pieces[p].push_back(Piece{E->ket.ptr(q), b.rows, b.cols, b.cols, m.rows, 0});
          m.rows += b.rows;
        }
      }
    }
    if (e0 != -1 && e1 == -1) {
      if (left_fill) {
        const int extra = lat.rd[r] - m.cols;
        require(extra >= 0, "expan_dim < 0 in LeftExpanTensorLattice");
        require(m.rows == lat.ld[l], "row or column dimension is not match in ExpanMatrix");
        m.cols += extra;
      } else {
        const int extra = lat.ld[l] - m.rows;
        require(extra >= 0, "expan_dim < 0 in RightExpanTensorLattice");
        require(m.cols == lat.rd[r], "row or column dimension is not match in ExpanMatrix");
        m.rows += extra;
      }
    }
    if (e0 == -1 && e1 == -1) {
      if (m.size() == 0 && (int64_t)lat.ld[l] * lat.rd[r] > 0) m = Shape{lat.ld[l], lat.rd[r]};
      else if ((int64_t)lat.ld[l] * lat.rd[r] > 0)
        require(m.rows == lat.ld[l] && m.cols == lat.rd[r], "Matrix indeces do not match in AddToMatrix");
    }
    nk.blk[p].rows = m.rows;
    nk.blk[p].cols = m.cols;
  }
  assemble(nk, pieces, cx);
  lat.ket = std::move(nk);
}

}  // namespace

void RealTensorRundmrg::expand(int leigh, int site, Lattice& E) {
  // ExpanTensorSpace (real_tensor_space.cc:319-387); leigh 0 = left-to-right
  if (!expansion_active()) return;
  Lattice& lat = lat_[site];
  Lattice origin;
  origin.d = lat.d;
  origin.lb = lat.lb; origin.ld = lat.ld; origin.rb = lat.rb; origin.rd = lat.rd;
  origin.phys = lat.phys;
  origin.ket = std::move(lat.ket);
  std::vector<int> nb, nd;
  if (leigh == 0) {
    combine_sectors(origin.rb, origin.rd, E.rb, E.rd, nb, nd);
    lat.define_sectors(origin.lb, nb, origin.ld, nd);
  } else {
    combine_sectors(origin.lb, origin.ld, E.lb, E.ld, nb, nd);
    lat.define_sectors(nb, origin.rb, nd, origin.rd);
  }
  std::vector<int> li, ri, pi;
  tensor_index(site, lat, li, ri, pi);
  for (size_t k = 0; k < li.size(); ++k) lat.phys[li[k]][ri[k]] = pi[k];
  cx_.rng.skip((int64_t)cs_ * li.size());  // DefineTensorLattice(..., 1, 1)
  lat.ket.define(li, ri);
  fill_expanded(lat, origin, &E, leigh == 0, cx_);
}

void RealTensorRundmrg::match(int leigh, int site) {
  // MatchTensorSpace (real_tensor_space.cc:389-440)
  if (!expansion_active()) return;
  Lattice& lat = lat_[site];
  Lattice origin;
  origin.d = lat.d;
  origin.lb = lat.lb; origin.ld = lat.ld; origin.rb = lat.rb; origin.rd = lat.rd;
  origin.phys = lat.phys;
  origin.ket = std::move(lat.ket);
  if (leigh == 0 && site > 0) {
    const Lattice& prev = lat_[site - 1];
    lat.define_sectors(prev.rb, origin.rb, prev.match_dim, origin.rd);
  } else if (leigh == 1 && site < L_sites_ - 1) {
    const Lattice& nxt = lat_[site + 1];
    lat.define_sectors(origin.lb, nxt.lb, origin.ld, nxt.match_dim);
  } else {
    lat.define_sectors({}, {}, {}, {});
  }
  std::vector<int> li, ri, pi;
  tensor_index(site, lat, li, ri, pi);
  for (size_t k = 0; k < li.size(); ++k) lat.phys[li[k]][ri[k]] = pi[k];
  cx_.rng.skip((int64_t)cs_ * li.size());
  lat.ket.define(li, ri);
  if (leigh == 0 && site > 0) fill_expanded(lat, origin, nullptr, false, cx_);
  else if (leigh == 1 && site < L_sites_ - 1) fill_expanded(lat, origin, nullptr, true, cx_);
}

// --------------------------------------------------------------------------
// Run (real_tensor_rundmrg.cc:115-245), no disk cache
// --------------------------------------------------------------------------
void RealTensorRundmrg::trace(int h, int j, double e, const char* what) {
  const Lattice& lat = lat_[j];
  long ld = 0, rd = 0;
  for (int v : lat.ld) ld += v;
  for (int v : lat.rd) rd += v;
  double nrm = 0.0;
  if (lat.ket.size > 0) {
    dot(lat.ket.data.data(), lat.ket.data.data(), lat.ket.doubles(), cx_.scal.data() + 30, cx_.partial,
        cx_.stream);
    DMRG_HIP(hipMemcpyAsync(&nrm, cx_.scal.data() + 30, sizeof(double), hipMemcpyDeviceToHost, cx_.stream));
    cx_.sync();
  }
  std::printf("TRACE %d %d %-17s E=%.12f nl=%d nr=%d ld=%ld rd=%ld nblk=%d size=%ld rng=%llu nrm2=%.12e\n", h, j,
              what, e, lat.nl(), lat.nr(), ld, rd, lat.ket.n(), (long)lat.ket.size,
              (unsigned long long)cx_.rng.count(), nrm);
  std::fflush(stdout);
}

void RealTensorRundmrg::Start() {
  t0_ = now();
  stats_ = LanczosStats{};
  ph_ = PhaseTimes{};
  res_ = RunResult{};
  if (!p_.resume_path.empty()) ReadTensorSpace(p_.resume_path);
  else DefineTensorSpace(p_.first_block, p_.first_dim);
  Initialize();
  cur_h_ = 0;
  cur_j_ = 0;
  half_e_ = 0;
  half_terr_ = 0;
  half_t0_ = now();
  started_ = true;
}

// One site step of half sweep h at site j: Lanczos, expansion, canonical
// form, environment update, match and merge (real_tensor_rundmrg.cc:145-175
// and :183-214).
void RealTensorRundmrg::Step(int h, int j) {
  const auto& W = ham_.sites();
  const bool l2r = (h % 2) == 0;
  if ((l2r && j == 0) || (!l2r && j == L_sites_ - 1)) {
    set_params(p_.maxb[h], p_.maxm[h], p_.canpre[h], p_.noise[h]);
    half_terr_ = 0;
    half_t0_ = now();
  }
  double tp = now();
  // The expansion's host work (sector bookkeeping, term lists, the random
  // draws, the GEMM descriptors) reads only the block structure of the site,
  // its environment and the rand() state, none of which the Lanczos solve
  // touches: it runs on a worker thread while the Lanczos iterations run, and
  // only the two GEMMs that read the new ket are launched afterwards.  The
  // rand() draws keep their order (the solve draws none).
  ExpanWork xw;
  auto prep = std::async(std::launch::async, [&] {
    xw.mapping = l2r ? expan_lattice(1, j, ham_.bond_tables()[j + 1], xw.E)
                     : expan_lattice(0, j, ham_.bond_tables()[j], xw.E);
    expan_contraction(l2r ? 0 : 1, j, xw.E, xw.mapping, &xw);
  });
  const double energy = LanczosMethod(p_.niter[h], j);
  half_e_ = energy;
  ph_.lanczos += now() - tp; tp = now();
  if (p_.verbose >= 2) trace(h, j, energy, "lanczos");
  prep.get();
  // ResetTensorNetwork keeps the block list but empties the matrices; the
  // reset environment is not read again before it is recomputed.
  for (auto& b : (l2r ? R_[j] : L_[j])) b.define(std::vector<int>(), std::vector<int>());
  Lattice& E = xw.E;
  launch_expansion(xw);
  ph_.expan += now() - tp; tp = now();
  if (p_.verbose >= 2) trace(h, j, energy, "expan_contraction");
  expand(l2r ? 0 : 1, j, E);
  ph_.expand += now() - tp; tp = now();
  if (p_.verbose >= 2) trace(h, j, energy, "expand");
  half_terr_ = std::max(half_terr_, canonical(l2r ? 0 : 1, j));
  ph_.canonical += now() - tp; tp = now();
  if (p_.verbose >= 2) trace(h, j, energy, "canonical");
  if (l2r) L_[j + 1] = left_env_update(lat_[j], W[j], L_[j], cx_);
  else R_[j - 1] = right_env_update(lat_[j], W[j], R_[j], cx_);
  ph_.env += now() - tp; tp = now();
  const int nb = l2r ? j + 1 : j - 1;
  match(l2r ? 0 : 1, nb);
  if (p_.verbose >= 2) trace(h, nb, energy, "match");
  merge(l2r ? 0 : 1, nb);
  ph_.merge += now() - tp;
  if (p_.verbose >= 2) trace(h, nb, energy, "merge");
  // the next step works on site nb (in this half sweep, or first in the next
  // one): its plan is built on the host now, while the GPU runs the
  // environment update and merge queued above
  next_plan_ = make_plan(nb);
  next_plan_site_ = nb;
  if ((l2r && j == L_sites_ - 2) || (!l2r && j == 1)) {
    res_.energies.push_back(half_e_);
    res_.trunc_err.push_back(half_terr_);
    res_.half_seconds.push_back(now() - half_t0_);
    // cout fixed, precision 12 (real_tensor_rundmrg.cc:122-123) or 14 (complex_tensor_rundmrg.cc:123)
    if (p_.verbose)
      std::printf("%s energy = %.*f\n", l2r ? "right" : "left", cs_ == 2 ? 14 : 12, half_e_), std::fflush(stdout);
    if (!l2r && !p_.record_path.empty()) WriteTensorSpace(p_.record_path);
  }
}

// One two-site step of half sweep h: the pair (j, j+1) left to right, the
// pair (j-1, j) right to left, with the centre on j; the SVD split moves the
// centre to the pair's other site.
void RealTensorRundmrg::Step2(int h, int j) {
  const auto& W = ham_.sites();
  const bool l2r = (h % 2) == 0;
  if ((l2r && j == 0) || (!l2r && j == L_sites_ - 1)) {
    set_params(p_.maxb[h], p_.maxm[h], p_.canpre[h], p_.noise[h]);
    half_terr_ = 0;
    half_t0_ = now();
  }
  const int a = l2r ? j : j - 1;
  double tp = now();
  TwoSite ts;
  HeffPlan plan;
  BuildTwoSite(a, ts, plan);
  ph_.plan += now() - tp;
  const double energy = lanczos_solve(plan, ts.theta.data.data(), p_.niter[h], p_.precision, p_.refine, cx_,
                                      stats_);
  half_e_ = energy;
  ph_.lanczos += now() - tp; tp = now();
  if (p_.verbose >= 2) trace(h, a, energy, "lanczos2");
  half_terr_ = std::max(half_terr_, split_two_site(a, ts, l2r));
  ph_.canonical += now() - tp; tp = now();
  if (p_.verbose >= 2) trace(h, a, energy, "split");
  if (l2r) L_[a + 1] = left_env_update(lat_[a], W[a], L_[a], cx_);
  else R_[a] = right_env_update(lat_[a + 1], W[a + 1], R_[a + 1], cx_);
  ph_.env += now() - tp;
  if ((l2r && j == L_sites_ - 2) || (!l2r && j == 1)) {
    res_.energies.push_back(half_e_);
    res_.trunc_err.push_back(half_terr_);
    res_.half_seconds.push_back(now() - half_t0_);
    if (p_.verbose)
      std::printf("%s energy = %.*f\n", l2r ? "right" : "left", cs_ == 2 ? 14 : 12, half_e_), std::fflush(stdout);
    if (!l2r && !p_.record_path.empty()) WriteTensorSpace(p_.record_path);
  }
}

double RealTensorRundmrg::split_two_site(int a, const TwoSite& ts, bool l2r) {
  // Theta's blocks regrouped into one site tensor whose other bond is the far
  // bond of the pair (left to right: blocks (l, m) holding [Theta(l,m,r) ...]
  // side by side over B's blocks r; right to left: blocks (m, r) holding
  // Theta(l,m,r) stacked over A's blocks l), then the canonical form of that
  // tensor (per-sector SVD, ComputeTruncateDim) keeps the orthonormal factor
  // on one site and its canonical tensor (Sigma V^H or U Sigma) is cut back
  // into the blocks of the other.
  Lattice& A = lat_[a];
  Lattice& B = lat_[a + 1];
  const int cs = cs_;
  const BlockTensor& th = ts.theta;
  const int site = l2r ? a : a + 1;
  Lattice& keep = l2r ? A : B;  // receives the orthonormal factor
  Lattice& other = l2r ? B : A;
  const BlockTensor& kk = keep.ket;  // block list of the kept site (its order)
  // fat blocks: one per block of the kept site
  std::vector<int> fl, fr, frows, fcols;
  std::vector<std::vector<int>> parts(kk.n());  // Theta blocks of each fat block, in order
  for (size_t t = 0; t < ts.src.size(); ++t) parts[l2r ? ts.src[t].first : ts.src[t].second].push_back((int)t);
  if (!l2r) {  // stacked over A's blocks in A's order
    for (auto& v : parts)
      std::stable_sort(v.begin(), v.end(), [&](int x, int y) { return ts.src[x].first < ts.src[y].first; });
  }
  BlockTensor F;
  F.cs = cs;
  for (int k = 0; k < kk.n(); ++k) { fl.push_back(kk.blk[k].l); fr.push_back(kk.blk[k].r); }
  F.define(fl, fr);
  std::vector<int> fdim(l2r ? keep.nr() : keep.nl(), 0);  // far-side extent per sector
  for (int k = 0; k < kk.n(); ++k) {
    int ext = 0, other_ext = -1;
    for (int t : parts[k]) {
      const Block& b = th.blk[t];
      ext += l2r ? b.cols : b.rows;
      const int o = l2r ? b.rows : b.cols;
      require(other_ext < 0 || other_ext == o, "two-site blocks do not line up", kErrInternal);
      other_ext = o;
    }
    if (other_ext < 0) other_ext = l2r ? kk.blk[k].rows : kk.blk[k].cols;
    F.blk[k].rows = l2r ? other_ext : ext;
    F.blk[k].cols = l2r ? ext : other_ext;
    const int sec = l2r ? kk.blk[k].r : kk.blk[k].l;
    fdim[sec] = std::max(fdim[sec], ext);
  }
  F.allocate();
  fill(F.data.data(), F.doubles(), 0.0, cx_.stream);
  std::vector<CopyTask> cps;
  for (int k = 0; k < kk.n(); ++k) {
    int at = 0;
    for (int t : parts[k]) {
      const Block& b = th.blk[t];
      if ((int64_t)b.rows * b.cols > 0) {
        double* dst = F.ptr(k) + (int64_t)cs * (l2r ? at : (int64_t)at * F.blk[k].cols);
        push_copy(cps, CopyTask{th.ptr(t), dst, b.rows, b.cols, b.cols, 1, F.blk[k].cols, 1, 1.0}, cs);
      }
      at += l2r ? b.cols : b.rows;
    }
  }
  launch_copies(cps, cx_.copy_scratch, cx_.stream);
  // canonical form of the fat tensor on the kept site
  std::vector<int> far_keep = l2r ? keep.rd : keep.ld;
  keep.ket = std::move(F);
  if (l2r) keep.rd = fdim; else keep.ld = fdim;
  (void)far_keep;
  const double discarded = canonical(l2r ? 0 : 1, site);
  // the canonical tensor, per sector of the shared bond, cut into the other
  // site's blocks (in its own order); the shared bond's dims follow
  const BlockTensor& can = keep.can;
  BlockTensor nk;
  nk.cs = cs;
  nk.blk = other.ket.blk;
  const int nsec = l2r ? keep.nr() : keep.nl();
  std::vector<int> tdim(nsec);
  for (int m = 0; m < nsec; ++m) tdim[m] = l2r ? can.blk[m].rows : can.blk[m].cols;
  // offset of each of the other site's blocks inside its sector's fat extent
  std::vector<int> off(other.ket.n(), 0);
  {
    std::vector<int> acc(nsec, 0);
    // the extent order is the order of the parts of any fat block of that
    // sector: B's block order (left to right) or A's block order (right to left)
    for (int k = 0; k < other.ket.n(); ++k) {
      const Block& ob = other.ket.blk[k];
      const int m = l2r ? ob.l : ob.r;
      off[k] = acc[m];
      acc[m] += l2r ? ob.cols : ob.rows;
    }
  }
  for (int k = 0; k < nk.n(); ++k) {
    const int m = l2r ? nk.blk[k].l : nk.blk[k].r;
    if (l2r) nk.blk[k].rows = tdim[m];
    else nk.blk[k].cols = tdim[m];
  }
  nk.allocate();
  cps.clear();
  for (int k = 0; k < nk.n(); ++k) {
    const Block& b = nk.blk[k];
    if ((int64_t)b.rows * b.cols == 0) continue;
    const int m = l2r ? b.l : b.r;
    const Block& cb = can.blk[m];
    // left to right: can[m] is t_m x extent, columns [off, off + cols);
    // right to left: can[m] is extent x t_m, rows [off, off + rows)
    const double* src = can.ptr(m) + (int64_t)cs * (l2r ? off[k] : (int64_t)off[k] * cb.cols);
    push_copy(cps, CopyTask{src, nk.ptr(k), b.rows, b.cols, cb.cols, 1, b.cols, 1, 1.0}, cs);
  }
  launch_copies(cps, cx_.copy_scratch, cx_.stream);
  cx_.sync();
  other.ket = std::move(nk);
  if (l2r) other.ld = tdim; else other.rd = tdim;
  keep.has_can = false;
  keep.can = BlockTensor{};
  keep.match_dim.clear();
  return discarded;
}

bool RealTensorRundmrg::Advance(int stop_h, int stop_j) {
  if (!started_) Start();
  if (paused_at_ >= 0.0) {  // wall time between Advance calls is not sweep time
    const double gap = now() - paused_at_;
    half_t0_ += gap;
    t0_ += gap;
    paused_at_ = -1.0;
  }
  const int nhalf = (int)p_.maxm.size();
  while (cur_h_ < nhalf) {
    const bool l2r = (cur_h_ % 2) == 0;
    if (!l2r && cur_j_ == 0) cur_j_ = L_sites_ - 1;
    if (cur_h_ == stop_h && cur_j_ == stop_j) {
      cx_.sync();
      paused_at_ = now();
      return true;
    }
    if (p_.two_site) Step2(cur_h_, cur_j_);
    else Step(cur_h_, cur_j_);
    if (l2r) {
      if (++cur_j_ == L_sites_ - 1) { ++cur_h_; cur_j_ = 0; }
    } else {
      if (--cur_j_ == 0) { ++cur_h_; cur_j_ = 0; }
    }
  }
  return false;
}

RunResult RealTensorRundmrg::Result() {
  cx_.sync();
  RunResult res = res_;
  res.total_seconds = (paused_at_ >= 0.0 ? paused_at_ : now()) - t0_;
  res.stats = stats_;
  res.phase = ph_;
  res.rng_draws = cx_.rng.count();
  res.final_bond_dims.clear();
  for (int i = 0; i < L_sites_; ++i) {
    int s = 0;
    for (int v : lat_[i].rd) s += v;
    res.final_bond_dims.push_back(s);
  }
  return res;
}

// --------------------------------------------------------------------------
// Batched H_eff of independent instances
// --------------------------------------------------------------------------
HeffBatch::~HeffBatch() {
  for (int p = 0; p < 3; ++p) {
    if (exec_[p]) (void)hipGraphExecDestroy(exec_[p]);
    if (graph_[p]) (void)hipGraphDestroy(graph_[p]);
  }
}

void HeffBatch::build(const std::vector<HeffPlan*>& plans, const std::vector<const double*>& x,
                      const std::vector<double*>& y, hipStream_t stream) {
  require(plans.size() == x.size() && plans.size() == y.size(), "one input and output vector per plan",
          kErrArgument);
  std::vector<const GemmBatch*> b1, b2;
  std::vector<const double*> none_in(plans.size(), nullptr);
  std::vector<double*> none_out(plans.size(), nullptr);
  fref_ = 0;
  for (HeffPlan* p : plans) {
    b1.push_back(&p->stage1());
    b2.push_back(&p->stage2());
    fref_ += p->fref();
  }
  // stage 1 reads x (relative A), writes each plan's own stage buffer;
  // stage 2 reads that buffer, accumulates into y (relative C)
  s1_.merge(b1, x, none_out, stream);
  s2_.merge(b2, none_in, y, stream);
  plans_ = plans;
  gens_.clear();
  for (HeffPlan* p : plans) gens_.push_back(p->generation());
  x_ = x;
  y_ = y;
  DMRG_HIP(hipStreamSynchronize(stream));
}

void HeffBatch::check_plans() const {
  // a plan rebuilt or re-tiled since build() no longer owns the buffers and
  // descriptors this batch points at
  for (size_t i = 0; i < plans_.size(); ++i)
    require(plans_[i]->generation() == gens_[i], "H_eff batch: a plan was rebuilt or re-tiled after the batch",
            kErrArgument);
}

void HeffBatch::apply_each(int precision, hipStream_t stream) {
  check_plans();
  for (size_t i = 0; i < plans_.size(); ++i) plans_[i]->apply(x_[i], y_[i], precision, stream);
}

void HeffBatch::apply(int precision, hipStream_t stream) {
  check_plans();
  s1_.launch(stream, precision);
  s2_.launch(stream, precision);
}

void HeffBatch::apply_graph(int precision, hipStream_t stream) {
  require(precision >= 0 && precision <= 2, "unknown precision", kErrArgument);
  check_plans();  // before any capture starts
  if (!exec_[precision]) {
    // kernels only (no memset nodes); captured on this thread's stream
    DMRG_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    try {
      apply(precision, stream);
    } catch (...) {
      // never leave the stream capturing
      hipGraph_t partial = nullptr;
      (void)hipStreamEndCapture(stream, &partial);
      if (partial) (void)hipGraphDestroy(partial);
      throw;
    }
    DMRG_HIP(hipStreamEndCapture(stream, &graph_[precision]));
    DMRG_HIP(hipGraphInstantiate(&exec_[precision], graph_[precision], nullptr, nullptr, 0));
  }
  DMRG_HIP(hipGraphLaunch(exec_[precision], stream));
}

// --------------------------------------------------------------------------
// Two-site effective Hamiltonian
// --------------------------------------------------------------------------
SiteOperator fuse_sites(const SiteOperator& a, const SiteOperator& b) {
  require(a.rb == b.lb && a.d == b.d, "MPO bond dimensions do not match in fuse_sites", kErrShape);
  const bool cz = a.complex() || b.complex();
  const int d = a.d, D = d * d;
  SiteOperator w;
  w.lb = a.lb;
  w.rb = b.rb;
  w.d = D;
  w.w.assign((size_t)w.lb * w.rb * D * D, 0.0);
  if (cz) w.wi.assign(w.w.size(), 0.0);
  for (int o1 = 0; o1 < a.lb; ++o1)
    for (int o3 = 0; o3 < b.rb; ++o3)
      for (int p1 = 0; p1 < d; ++p1)
        for (int q1 = 0; q1 < d; ++q1)
          for (int p2 = 0; p2 < d; ++p2)
            for (int q2 = 0; q2 < d; ++q2) {
              double re = 0.0, im = 0.0;
              for (int o2 = 0; o2 < a.rb; ++o2) {
                const double ar = a.at(o1, o2, p1, q1), ai = a.im(o1, o2, p1, q1);
                const double br = b.at(o2, o3, p2, q2), bi = b.im(o2, o3, p2, q2);
                re += ar * br - ai * bi;
                im += ar * bi + ai * br;
              }
              const size_t k = (((size_t)o1 * w.rb + o3) * D + (p1 * d + p2)) * D + (q1 * d + q2);
              w.w[k] = re;
              if (cz) w.wi[k] = im;
            }
  return w;
}

void RealTensorRundmrg::BuildTwoSite(int site, TwoSite& ts, HeffPlan& plan) {
  require(site >= 0 && site + 1 < L_sites_, "two-site pair out of range", kErrArgument);
  const Lattice& A = lat_[site];
  const Lattice& B = lat_[site + 1];
  const int d = d_;
  require(A.nr() == B.nl(), "middle bond sectors differ between the two sites", kErrShape);
  // Theta blocks: for each block (l, m) of A in its order, the blocks (m, r)
  // of B in B's order; physical index p1 * d + p2
  BlockTensor& th = ts.theta;
  th = BlockTensor{};
  th.cs = cs_;
  std::vector<int> li, ri;
  std::vector<std::pair<int, int>> src;
  ts.bphys.clear();
  ts.src.clear();
  for (int a = 0; a < A.ket.n(); ++a) {
    const Block& ab = A.ket.blk[a];
    for (int b : B.ket.same(0, ab.r)) {
      const Block& bb = B.ket.blk[b];
      require(ab.cols == bb.rows, "Matrix indeces do not match in MultiplyToMatrix (A*B)");
      li.push_back(ab.l);
      ri.push_back(bb.r);
      src.push_back({a, b});
      ts.src.push_back({a, b});
      ts.bphys.push_back(A.phys[ab.l][ab.r] * d + B.phys[bb.l][bb.r]);
    }
  }
  th.define(li, ri);
  for (size_t k = 0; k < src.size(); ++k) {
    th.blk[k].rows = A.ket.blk[src[k].first].rows;
    th.blk[k].cols = B.ket.blk[src[k].second].cols;
  }
  th.allocate();
  GemmBatch g;
  for (size_t k = 0; k < src.size(); ++k) {
    const Block& ab = A.ket.blk[src[k].first];
    const Block& bb = B.ket.blk[src[k].second];
    const Block& tb = th.blk[k];
    if ((int64_t)tb.rows * tb.cols == 0) continue;
    if (cs_ == 2) {
      zgemm_item(g, th.ptr((int)k), tb.cols, tb.rows, tb.cols, 0, 0,
                 {ZTerm{A.ket.ptr(src[k].first), ab.cols, B.ket.ptr(src[k].second), bb.cols, ab.cols, 0, false,
                        false, 1.0, 0.0}});
    } else {
      g.begin_item(th.ptr((int)k), tb.cols, tb.rows, tb.cols, 0);
      g.add(A.ket.ptr(src[k].first), ab.cols, B.ket.ptr(src[k].second), bb.cols, ab.cols, 0, 1.0);
    }
  }
  g.finalize(cx_.stream);
  g.launch(cx_.stream, kFP64);
  ts.W2 = fuse_sites(ham_.sites()[site], ham_.sites()[site + 1]);
  plan.set_presplit(p_.precision == kBF16x3);
  plan.build_blocks(th, ts.bphys, B.nr(), ts.W2, L_[site], R_[site + 1], cx_.stream);
  cx_.sync();
}

// --------------------------------------------------------------------------
// Tensor-space files
// --------------------------------------------------------------------------
namespace {

template <typename T>
void put(std::ostream& os, const T& v) {
  os.write(reinterpret_cast<const char*>(&v), sizeof(T));
}
template <typename T>
T get(std::istream& is) {
  T v{};
  is.read(reinterpret_cast<char*>(&v), sizeof(T));
  require(bool(is), "tensor space file ends early", kErrArgument);
  return v;
}

// RealMatrixBlock::WriteMatrixBlock (real_matrix_block.cc:227-240): block
// count, then per block (left, right, RealMatrix::WriteMatrix: rows, cols,
// rows*cols, the entries row-major)
void write_blocks(std::ostream& os, const BlockTensor& t, hipStream_t stream) {
  std::vector<double> h((size_t)t.doubles());
  if (!h.empty()) {
    DMRG_HIP(hipMemcpyAsync(h.data(), t.base(), h.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
    DMRG_HIP(hipStreamSynchronize(stream));
  }
  put<int>(os, t.n());
  for (const Block& b : t.blk) {
    put<int>(os, b.l);
    put<int>(os, b.r);
    put<int>(os, b.rows);
    put<int>(os, b.cols);
    put<int>(os, b.rows * b.cols);
    os.write(reinterpret_cast<const char*>(h.data() + t.cs * b.off),
             (std::streamsize)(sizeof(double) * t.cs * b.rows * b.cols));
  }
}

// RealMatrixBlock::ReadMatrixBlock (real_matrix_block.cc:254-283)
void read_blocks(std::istream& is, BlockTensor& t, int cs, hipStream_t stream) {
  const int n = get<int>(is);
  require(n >= 0, "bad block count in tensor space file", kErrArgument);
  std::vector<int> l(n), r(n);
  std::vector<std::vector<double>> data(n);
  std::vector<std::pair<int, int>> shape(n);
  for (int k = 0; k < n; ++k) {
    l[k] = get<int>(is);
    r[k] = get<int>(is);
    const int rows = get<int>(is), cols = get<int>(is), total = get<int>(is);
    require(rows >= 0 && cols >= 0 && total == rows * cols, "bad matrix header in tensor space file",
            kErrArgument);
    shape[k] = {rows, cols};
    data[k].resize((size_t)cs * total);
    is.read(reinterpret_cast<char*>(data[k].data()), (std::streamsize)(sizeof(double) * data[k].size()));
    require(bool(is), "tensor space file ends early", kErrArgument);
  }
  t = BlockTensor{};
  t.cs = cs;
  t.define(l, r);
  for (int k = 0; k < n; ++k) {
    t.blk[k].rows = shape[k].first;
    t.blk[k].cols = shape[k].second;
  }
  t.allocate();
  std::vector<double> h((size_t)t.doubles());
  for (int k = 0; k < n; ++k) std::copy(data[k].begin(), data[k].end(), h.begin() + cs * t.blk[k].off);
  if (!h.empty()) {
    DMRG_HIP(hipMemcpyAsync(t.data.data(), h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, stream));
    DMRG_HIP(hipStreamSynchronize(stream));
  }
}

// RealTensorLattice::WriteTensorLattice (real_tensor_lattice.cc:218-239)
void write_lattice(std::ostream& os, const Lattice& lat, hipStream_t stream) {
  put<int>(os, lat.d);
  put<int>(os, lat.nl());
  for (int i = 0; i < lat.nl(); ++i) {
    put<int>(os, lat.lb[i]);
    put<int>(os, lat.ld[i]);
  }
  put<int>(os, lat.nr());
  for (int j = 0; j < lat.nr(); ++j) {
    put<int>(os, lat.rb[j]);
    put<int>(os, lat.rd[j]);
  }
  for (int i = 0; i < lat.nl(); ++i)
    for (int j = 0; j < lat.nr(); ++j) put<int>(os, lat.phys[i][j]);
  write_blocks(os, lat.ket, stream);
}

// RealTensorLattice::ReadTensorLattice (real_tensor_lattice.cc:252-298);
// the canonical tensor and match dims are not stored
void read_lattice(std::istream& is, Lattice& lat, int cs, hipStream_t stream) {
  lat = Lattice{};
  lat.d = get<int>(is);
  const int nl = get<int>(is);
  require(nl >= 0, "bad sector count in tensor space file", kErrArgument);
  std::vector<int> lb(nl), ld(nl);
  for (int i = 0; i < nl; ++i) {
    lb[i] = get<int>(is);
    ld[i] = get<int>(is);
  }
  const int nr = get<int>(is);
  require(nr >= 0, "bad sector count in tensor space file", kErrArgument);
  std::vector<int> rb(nr), rd(nr);
  for (int j = 0; j < nr; ++j) {
    rb[j] = get<int>(is);
    rd[j] = get<int>(is);
  }
  lat.define_sectors(lb, rb, ld, rd);
  for (int i = 0; i < nl; ++i)
    for (int j = 0; j < nr; ++j) lat.phys[i][j] = get<int>(is);
  read_blocks(is, lat.ket, cs, stream);
}

}  // namespace

void WriteTensorContraction(const Env& env, const std::string& path, hipStream_t stream) {
  // RealTensorContraction::WriteTensorContraction
  // (real_tensor_contraction.cc:73-90): one WriteMatrixBlock per channel
  std::ofstream os(path, std::ios::binary | std::ios::out);
  require(os.good(), "can't open environment file for writing: " + path, kErrArgument);
  for (const BlockTensor& t : env) write_blocks(os, t, stream);
  os.flush();
  require(os.good(), "write error on environment file: " + path, kErrInternal);
}

Env ReadTensorContraction(const std::string& path, int nbond, int cs, hipStream_t stream) {
  // ReadTensorContraction (real_tensor_contraction.cc:92-109): nbond
  // ReadMatrixBlock in channel order
  std::ifstream is(path, std::ios::binary | std::ios::in);
  require(is.good(), "can't open environment file: " + path, kErrArgument);
  require(nbond > 0, "an environment has at least one channel", kErrArgument);
  Env env(nbond);
  for (BlockTensor& t : env) read_blocks(is, t, cs, stream);
  is.peek();
  require(is.eof(), "trailing bytes in environment file: " + path, kErrArgument);
  return env;
}

void RealTensorRundmrg::WriteTensorSpace(const std::string& path) {
  // RealTensorSpace::WriteTensorSpace (real_tensor_space.cc:182-208)
  cx_.sync();
  std::ofstream os(path, std::ios::binary | std::ios::out);
  require(os.good(), "can't open tensor space file for writing: " + path, kErrArgument);
  const int nq = space_.num_quantum();
  put<int>(os, L_sites_);
  put<int>(os, L_sites_ + 1);
  put<int>(os, L_sites_ - 1);
  put<int>(os, nq);
  for (int i = 0; i <= L_sites_; ++i) put<int>(os, (int)tables_[i].size());
  for (int i = 0; i <= L_sites_; ++i)
    for (const QN& q : tables_[i])
      for (int k = 0; k < nq; ++k) put<int>(os, q[k]);
  for (int i = 0; i < L_sites_; ++i) write_lattice(os, lat_[i], cx_.stream);
  put<int>(os, max_block_);
  put<int>(os, max_dim_);
  put<double>(os, canpre_);
  put<double>(os, noise_);
  os.flush();
  require(os.good(), "write error on tensor space file: " + path, kErrInternal);
}

void RealTensorRundmrg::ReadTensorSpace(const std::string& path) {
  // RealTensorSpace::ReadTensorSpace (real_tensor_space.cc:221-277).  The
  // reference sizes two of its arrays by num_quantum where num_site+1 is
  // meant (:245, :248); this reads the layout its writer produces.
  std::ifstream is(path, std::ios::binary | std::ios::in);
  require(is.good(), "can't open tensor space file: " + path, kErrArgument);
  const int L = get<int>(is), Lpp = get<int>(is), Lmm = get<int>(is), nq = get<int>(is);
  require(L == L_sites_ && Lpp == L + 1 && Lmm == L - 1, "tensor space file has another number of sites",
          kErrArgument);
  require(nq == space_.num_quantum(), "tensor space file has another number of quantum numbers",
          kErrArgument);
  std::vector<int> nt(Lpp);
  for (int i = 0; i < Lpp; ++i) {
    nt[i] = get<int>(is);
    require(nt[i] >= 0, "bad quantum table size in tensor space file", kErrArgument);
  }
  std::vector<QNTable> tabs(Lpp);
  for (int i = 0; i < Lpp; ++i)
    for (int j = 0; j < nt[i]; ++j) {
      QN q(nq);
      for (int k = 0; k < nq; ++k) q[k] = get<int>(is);
      tabs[i].push_back(q);
    }
  std::vector<Lattice> lats(L);
  for (int i = 0; i < L; ++i) {
    read_lattice(is, lats[i], cs_, cx_.stream);
    require(lats[i].d == d_, "tensor space file has another physical dimension", kErrArgument);
  }
  max_block_ = get<int>(is);
  max_dim_ = get<int>(is);
  canpre_ = get<double>(is);
  noise_ = get<double>(is);
  tables_ = tabs;
  lat_ = std::move(lats);
}

RunResult RealTensorRundmrg::Run() {
  Start();
  Advance(-1, -1);
  RunResult r = Result();
  if (p_.verbose) std::printf("Total Time : %.*f s\n", cs_ == 2 ? 14 : 12, r.total_seconds), std::fflush(stdout);
  return r;
}

}  // namespace dmrg
