// Model layer: QN tables, QN rule, MPO construction and deparallelisation.
#include "model.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <fstream>

#include "common.h"
#include "quicksort.h"

namespace dmrg {

bool SiteOperator::left_nonzero(int o1, int p1) const {
  for (int r = 0; r < rb; ++r)
    for (int p2 = 0; p2 < d; ++p2)
      if (std::hypot(at(o1, r, p1, p2), im(o1, r, p1, p2)) > 1e-15) return true;
  return false;
}
bool SiteOperator::right_nonzero(int o2, int p2) const {
  for (int l = 0; l < lb; ++l)
    for (int p1 = 0; p1 < d; ++p1)
      if (std::hypot(at(l, o2, p1, p2), im(l, o2, p1, p2)) > 1e-15) return true;
  return false;
}
bool SiteOperator::mixed_nonzero(int o1, int p2) const {
  for (int r = 0; r < rb; ++r)
    for (int p1 = 0; p1 < d; ++p1)
      if (std::hypot(at(o1, r, p1, p2), im(o1, r, p1, p2)) > 1e-15) return true;
  return false;
}

namespace {

using zd = std::complex<double>;

// Norm of an entry (complex: Norm, src/tensor/complex.h:112-115) and the
// quotient of isParallelElement (complex: operator/ of complex.h, written
// out so that the prefactors, which enter the merged MPO, are the same bits).
inline double mag(double x) { return std::fabs(x); }
inline double mag(const zd& z) { return std::sqrt(z.real() * z.real() + z.imag() * z.imag()); }
inline double quot(double a, double b) { return a / b; }
inline zd quot(const zd& a, const zd& b) {
  const double k = 1.0 / (b.real() * b.real() + b.imag() * b.imag());
  return zd((a.real() * b.real() + a.imag() * b.imag()) * k, (a.imag() * b.real() - a.real() * b.imag()) * k);
}
inline zd cmul(const zd& a, const zd& b) {
  return zd(a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real());
}
inline double cmul(double a, double b) { return a * b; }

// isParallelElement (src/util/general.h:146-180, complex :182-207) applied
// along a whole column: returns true (and the prefactor) when `col` is
// parallel to `ref`.
template <typename T>
bool parallel_to(const T* col, const T* ref, int n, int stride, T& prefactor) {
  bool zero = true;
  prefactor = T(0);
  for (int i = 0; i < n; ++i) {
    const T e0 = col[(size_t)i * stride], e1 = ref[(size_t)i * stride];
    if (mag(e1) <= 1e-10) {
      if (mag(e0) > 1e-10) return false;
    } else {
      const T f = quot(e0, e1);
      if (zero) {
        prefactor = f;
        zero = false;
      } else if (mag(f - prefactor) > 1e-8) {
        return false;
      }
    }
  }
  return true;
}

template <typename T>
std::vector<int> parallel_t(const std::vector<T>& a, int rows, int cols, int leigh, std::vector<T>& transfer) {
  // Work on "vectors": columns (leigh 0) or rows (leigh 1).
  const int nv = leigh == 0 ? cols : rows;
  const int len = leigh == 0 ? rows : cols;
  const int stride = leigh == 0 ? cols : 1;
  auto vec = [&](int v) { return leigh == 0 ? &a[v] : &a[(size_t)v * cols]; };
  std::vector<int> pos{0};
  transfer.assign((size_t)nv * nv, T(0));
  auto T_ = [&](int k, int v) -> T& {
    // leigh 0: transfer(k, v); leigh 1: transfer(v, k)
    return leigh == 0 ? transfer[(size_t)k * nv + v] : transfer[(size_t)v * nv + k];
  };
  T_(0, 0) = T(1);
  for (int v = 0; v < nv; ++v) {
    bool found = false;
    for (size_t k = 0; k < pos.size(); ++k) {
      T pref;
      if (parallel_to(vec(v), vec(pos[k]), len, stride, pref)) {
        T_((int)k, v) = pref;
        found = true;
        break;
      }
    }
    if (!found) {
      pos.push_back(v);
      T_((int)pos.size() - 1, v) = T(1);
    }
  }
  return pos;
}

std::vector<zd> to_z(const std::vector<double>& a) {
  std::vector<zd> z(a.size() / 2);
  for (size_t i = 0; i < z.size(); ++i) z[i] = zd(a[2 * i], a[2 * i + 1]);
  return z;
}
std::vector<double> from_z(const std::vector<zd>& z) {
  std::vector<double> a(2 * z.size());
  for (size_t i = 0; i < z.size(); ++i) { a[2 * i] = z[i].real(); a[2 * i + 1] = z[i].imag(); }
  return a;
}

// Site tensor entries as T (complex: w + i wi).
template <typename T> std::vector<T> load(const SiteOperator& W);
template <> std::vector<double> load<double>(const SiteOperator& W) { return W.w; }
template <> std::vector<zd> load<zd>(const SiteOperator& W) {
  std::vector<zd> z(W.w.size());
  for (size_t i = 0; i < z.size(); ++i) z[i] = zd(W.w[i], W.wi[i]);
  return z;
}
void store(SiteOperator& W, const std::vector<double>& v) { W.w = v; W.wi.clear(); }
void store(SiteOperator& W, const std::vector<zd>& v) {
  W.w.resize(v.size());
  W.wi.resize(v.size());
  for (size_t i = 0; i < v.size(); ++i) { W.w[i] = v[i].real(); W.wi[i] = v[i].imag(); }
}

// ParallelTensorHamiltonian on entries of type T.
template <typename T>
void parallel_sweep(std::vector<SiteOperator>& Ws, std::vector<QNTable>& qt, int num_site, int d) {
  const int dd = d * d;
  for (int i = 0; i < num_site - 1; ++i) {
    // LeftParallelTensorOperator (real_tensor_operator.cc:199-236): rows (l,p1,p2)
    SiteOperator& W = Ws[i];
    const std::vector<T> w = load<T>(W);
    const int rows = W.lb * dd, cols = W.rb;
    std::vector<T> resh((size_t)rows * cols);
    for (int l = 0; l < W.lb; ++l)
      for (int r = 0; r < W.rb; ++r)
        for (int p = 0; p < dd; ++p) resh[((size_t)l * dd + p) * cols + r] = w[((size_t)l * W.rb + r) * dd + p];
    std::vector<T> Tm;
    std::vector<int> pos = parallel_t(resh, rows, cols, 0, Tm);
    const int np = (int)pos.size();
    std::vector<T> nw((size_t)W.lb * np * dd, T(0));
    for (int l = 0; l < W.lb; ++l)
      for (int r = 0; r < np; ++r)
        for (int p = 0; p < dd; ++p) nw[((size_t)l * np + r) * dd + p] = resh[((size_t)l * dd + p) * cols + pos[r]];
    W.rb = np;
    store(W, nw);
    // LeftMergeTensorOperator (real_tensor_operator.cc:277-318): W'[l] = sum_m T[l][m] W[m]
    SiteOperator& X = Ws[i + 1];
    require(cols == X.lb, "merge dimension is not match in LeftMergeTensorOperator");
    const std::vector<T> x = load<T>(X);
    std::vector<T> m((size_t)np * X.rb * dd, T(0));
    for (int l = 0; l < np; ++l)
      for (int r = 0; r < X.rb; ++r)
        for (int p = 0; p < dd; ++p) {
          T e = T(0);
          for (int k = 0; k < cols; ++k) e += cmul(Tm[(size_t)l * cols + k], x[((size_t)k * X.rb + r) * dd + p]);
          m[((size_t)l * X.rb + r) * dd + p] = e;
        }
    X.lb = np;
    store(X, m);
    QNTable q;
    for (int p : pos) q.push_back(qt[i + 1][p]);
    qt[i + 1] = q;
  }
  for (int i = num_site - 1; i > 0; --i) {
    // RightParallelTensorOperator (real_tensor_operator.cc:238-275): columns (r,p1,p2)
    SiteOperator& W = Ws[i];
    const int rows = W.lb, cols = W.rb * dd;
    const std::vector<T> resh = load<T>(W);  // (l, r, p1, p2) row-major == rows x cols
    std::vector<T> Tm;
    std::vector<int> pos = parallel_t(resh, rows, cols, 1, Tm);
    const int np = (int)pos.size();
    std::vector<T> nw((size_t)np * W.rb * dd, T(0));
    for (int l = 0; l < np; ++l)
      for (int c = 0; c < cols; ++c) nw[(size_t)l * cols + c] = resh[(size_t)pos[l] * cols + c];
    W.lb = np;
    store(W, nw);
    // RightMergeTensorOperator (real_tensor_operator.cc:320-361): W'[:, r] = sum_m T[m][r] W[:, m]
    SiteOperator& X = Ws[i - 1];
    require(rows == X.rb, "merge dimension is not match in RightMergeTensorOperator");
    const std::vector<T> x = load<T>(X);
    std::vector<T> m((size_t)X.lb * np * dd, T(0));
    for (int l = 0; l < X.lb; ++l)
      for (int r = 0; r < np; ++r)
        for (int p = 0; p < dd; ++p) {
          T e = T(0);
          for (int k = 0; k < rows; ++k) e += cmul(Tm[(size_t)k * rows + r], x[((size_t)l * X.rb + k) * dd + p]);
          m[((size_t)l * np + r) * dd + p] = e;
        }
    X.rb = np;
    store(X, m);
    QNTable q;
    for (int p : pos) q.push_back(qt[i][p]);
    qt[i] = q;
  }
}

}  // namespace

std::vector<int> parallel_matrix(const std::vector<double>& a, int rows, int cols, int leigh,
                                 std::vector<double>& transfer) {
  return parallel_t(a, rows, cols, leigh, transfer);
}

std::vector<int> zparallel_matrix(const std::vector<double>& a, int rows, int cols, int leigh,
                                  std::vector<double>& transfer) {
  std::vector<zd> t;
  std::vector<int> pos = parallel_t(to_z(a), rows, cols, leigh, t);
  transfer = from_z(t);
  return pos;
}

void RealTensorHamiltonian::Init(int num_site, int physics_dim, int num_quantum,
                                 std::vector<std::vector<double>> basic) {
  num_site_ = num_site;
  d_ = physics_dim;
  nq_ = num_quantum;
  basic_ = std::move(basic);
  W_.assign(num_site, SiteOperator{});
  for (auto& w : W_) w.d = d_;
  qt_.assign(num_site + 1, QNTable{});
  qt_[0] = QNTable{QN(nq_, 0)};
  qt_[num_site] = QNTable{QN(nq_, 0)};
}

namespace {
template <typename T>
void put(std::ostream& os, const T& v) {
  os.write(reinterpret_cast<const char*>(&v), sizeof(T));
}
template <typename T>
T get(std::istream& is) {
  T v{};
  is.read(reinterpret_cast<char*>(&v), sizeof(T));
  require(bool(is), "MPO file ends early", kErrArgument);
  return v;
}
// RealMatrix::WriteMatrix (real_matrix.cc:139-151): rows, cols, rows*cols,
// the entries (ComplexMatrix: re, im pairs, complex_matrix.cc:139-175)
void put_matrix(std::ostream& os, int rows, int cols, const double* re, const double* im, bool cplx) {
  put<int>(os, rows);
  put<int>(os, cols);
  put<int>(os, rows * cols);
  for (int e = 0; e < rows * cols; ++e) {
    put<double>(os, re[e]);
    if (cplx) put<double>(os, im ? im[e] : 0.0);
  }
}
void get_matrix(std::istream& is, int rows, int cols, double* re, double* im, bool cplx) {
  const int r = get<int>(is), c = get<int>(is), n = get<int>(is);
  require(r == rows && c == cols && n == rows * cols, "bad matrix header in MPO file", kErrArgument);
  for (int e = 0; e < n; ++e) {
    re[e] = get<double>(is);
    const double i = cplx ? get<double>(is) : 0.0;
    if (im) im[e] = i;
  }
}
}  // namespace

void RealTensorHamiltonian::WriteTensorHamiltonian(const std::string& path) const {
  std::ofstream os(path, std::ios::binary | std::ios::out);
  require(os.good(), "can't open MPO file for writing: " + path, kErrArgument);
  put<int>(os, num_site_);
  put<int>(os, num_site_ + 1);
  put<int>(os, num_site_ - 1);
  put<int>(os, (int)basic_.size());
  for (size_t k = 0; k < basic_.size(); ++k)
    put_matrix(os, d_, d_, basic_[k].data(), k < basic_im_.size() ? basic_im_[k].data() : nullptr, cplx_);
  put<int>(os, nq_);
  for (int i = 0; i <= num_site_; ++i) put<int>(os, (int)qt_[i].size());
  for (int i = 0; i <= num_site_; ++i)
    for (const QN& q : qt_[i])
      for (int k = 0; k < nq_; ++k) put<int>(os, q[k]);
  const int dd = d_ * d_;
  for (const SiteOperator& W : W_) {
    put<int>(os, W.d);
    put<int>(os, W.lb);
    put<int>(os, W.rb);
    for (int a = 0; a < W.lb; ++a)
      for (int b = 0; b < W.rb; ++b) {
        const size_t o = ((size_t)a * W.rb + b) * dd;
        put_matrix(os, W.d, W.d, W.w.data() + o, W.wi.empty() ? nullptr : W.wi.data() + o, cplx_);
      }
  }
  os.flush();
  require(os.good(), "write error on MPO file: " + path, kErrInternal);
}

void RealTensorHamiltonian::ReadTensorHamiltonian(const std::string& path) {
  // The reference's reader reads the basic operators into matrices it never
  // allocated (real_tensor_hamiltonian.cc:132: basic_operator_[i] is an
  // uninitialised pointer) and never reads the site operators back (it
  // deletes tensor_hamiltonian_ and stops after the quantum tables).  Fixed
  // here: both are read, in the order the writer puts them.
  std::ifstream is(path, std::ios::binary | std::ios::in);
  require(is.good(), "can't open MPO file: " + path, kErrArgument);
  const int L = get<int>(is), Lpp = get<int>(is), Lmm = get<int>(is);
  require(L == num_site_ && Lpp == L + 1 && Lmm == L - 1, "MPO file has another number of sites", kErrArgument);
  const int nop = get<int>(is);
  require(nop >= 0, "bad operator count in MPO file", kErrArgument);
  std::vector<std::vector<double>> basic(nop, std::vector<double>((size_t)d_ * d_));
  std::vector<std::vector<double>> basic_im(cplx_ ? nop : 0, std::vector<double>((size_t)d_ * d_));
  for (int k = 0; k < nop; ++k) get_matrix(is, d_, d_, basic[k].data(), cplx_ ? basic_im[k].data() : nullptr, cplx_);
  const int nq = get<int>(is);
  require(nq == nq_, "MPO file has another number of quantum numbers", kErrArgument);
  std::vector<int> nt(Lpp);
  for (int& n : nt) {
    n = get<int>(is);
    require(n >= 0, "bad quantum table size in MPO file", kErrArgument);
  }
  std::vector<QNTable> qt(Lpp);
  for (int i = 0; i < Lpp; ++i)
    for (int j = 0; j < nt[i]; ++j) {
      QN q(nq);
      for (int& v : q) v = get<int>(is);
      qt[i].push_back(q);
    }
  std::vector<SiteOperator> W(L);
  for (SiteOperator& w : W) {
    w.d = get<int>(is);
    w.lb = get<int>(is);
    w.rb = get<int>(is);
    require(w.d == d_ && w.lb > 0 && w.rb > 0, "bad operator header in MPO file", kErrArgument);
    const size_t n = (size_t)w.lb * w.rb * d_ * d_;
    w.w.assign(n, 0.0);
    if (cplx_) w.wi.assign(n, 0.0);
    for (int a = 0; a < w.lb; ++a)
      for (int b = 0; b < w.rb; ++b) {
        const size_t o = ((size_t)a * w.rb + b) * d_ * d_;
        get_matrix(is, d_, d_, w.w.data() + o, cplx_ ? w.wi.data() + o : nullptr, cplx_);
      }
  }
  for (int i = 0; i < L; ++i)
    require((i == 0 || W[i].lb == W[i - 1].rb) && (int)qt[i].size() == W[i].lb && (int)qt[i + 1].size() == W[i].rb,
            "MPO file: bond dimensions and quantum tables do not match", kErrArgument);
  basic_ = std::move(basic);
  basic_im_ = std::move(basic_im);
  qt_ = std::move(qt);
  W_ = std::move(W);
}

void RealTensorHamiltonian::ExpanTensorHamiltonian(int site, int op_index, double cr, double ci,
                                                   const QN& table) {
  // RealTensorOperator::ExpanTensorOperator (real_tensor_operator.cc:141-197;
  // complex: complex_tensor_operator.cc:141-197, coefficient * basic)
  int leigh = -1;
  if (site == 0) leigh = 0;
  if (site == num_site_ - 1) leigh = 1;
  SiteOperator& W = W_[site];
  const int d = d_, dd = d_ * d_;
  SiteOperator n;
  n.d = d;
  int e0 = 1, e1 = 1;
  if (W.lb == 0 && W.rb == 0) {
    n.lb = n.rb = 1;
    e0 = e1 = 0;
  } else {
    e0 = leigh == 0 ? 0 : 1;
    e1 = leigh == 1 ? 0 : 1;
    n.lb = W.lb + e0;
    n.rb = W.rb + e1;
  }
  n.w.assign((size_t)n.lb * n.rb * dd, 0.0);
  if (cplx_) n.wi.assign(n.w.size(), 0.0);
  for (int l = 0; l < W.lb; ++l)
    for (int r = 0; r < W.rb; ++r)
      for (int p = 0; p < dd; ++p) {
        const size_t to = ((size_t)l * n.rb + r) * dd + p, from = ((size_t)l * W.rb + r) * dd + p;
        n.w[to] = W.w[from];
        if (cplx_) n.wi[to] = W.wi[from];
      }
  const int il = e0 * W.lb, ir = e1 * W.rb;
  for (int p = 0; p < dd; ++p) {
    const size_t to = ((size_t)il * n.rb + ir) * dd + p;
    n.w[to] = cr * basic_[op_index][p];
    if (cplx_) n.wi[to] = ci * basic_[op_index][p];
  }
  W = std::move(n);
  if (site < num_site_ - 1) qt_[site + 1].push_back(table);
}

void RealTensorHamiltonian::ParallelTensorHamiltonian() {
  // nothing to deparallelise while a site still has no term (a lattice
  // without any bond or on-site term leaves the MPO empty; the model
  // constructors refuse that)
  for (const SiteOperator& W : W_)
    if (W.lb == 0 || W.rb == 0) return;
  if (cplx_) parallel_sweep<zd>(W_, qt_, num_site_, d_);
  else parallel_sweep<double>(W_, qt_, num_site_, d_);
}

std::vector<QNTable> RealSpinlessBoseSquareSpace::DefineQuantumTable() const {
  // real_spinless_bose_square_space.cc:26-62 with ReorderQuantumTable
  // (real_tensor_space.cc:716-736)
  const int L = num_site();
  std::vector<QNTable> tabs;
  for (int i = 0; i <= L; ++i) {
    int lo = 0, hi = 0;
    double mean = 0.0;
    if (i > 0) {
      lo = std::max(0, n_ - (L - i) * (d_ - 1));
      hi = std::min((d_ - 1) * i, n_);
      mean = (double)n_ / L * i;
    }
    std::vector<double> dev;
    std::vector<int> idx;
    QNTable t;
    for (int q = lo; q <= hi; ++q) {
      t.push_back(QN{q});
      dev.push_back(std::pow(q - mean, 2));
      idx.push_back((int)idx.size());
    }
    quicksort(dev, &idx, 1);
    QNTable r;
    for (int k : idx) r.push_back(t[k]);
    tabs.push_back(r);
  }
  return tabs;
}

int RealSpinlessBoseSquareSpace::CheckQuantumTable(int, const QN& left, const QN& right) const {
  for (int i = 0; i < d_; ++i)
    if (left[0] + i == right[0]) return i;
  return -1;
}

RealSpinlessBoseSquareHamiltonian::RealSpinlessBoseSquareHamiltonian(
    int nx, int ny, int d, double /*flux: stored, never used by the reference*/, double hop_x,
    double hop_y, double U, double mu, bool period_x, bool period_y)
    : nx_(nx), ny_(ny), hop_x_(hop_x), hop_y_(hop_y), U_(U), mu_(mu), px_(period_x), py_(period_y) {
  // DefineBasicTensor (real_spinless_bose_square_hamiltonian.cc:41-64):
  // 0=I, 1=a^+, 2=a^-, 3=n, 4=mu*n + U*n(n-1)
  std::vector<std::vector<double>> ops(5, std::vector<double>((size_t)d * d, 0.0));
  for (int i = 0; i < d; ++i) {
    ops[0][i * d + i] = 1.0;
    ops[3][i * d + i] = i;
    ops[4][i * d + i] = mu * i + U * i * (i - 1);
  }
  for (int i = 1; i < d; ++i) {
    ops[1][i * d + (i - 1)] = std::sqrt((double)i);
    ops[2][(i - 1) * d + i] = std::sqrt((double)i);
  }
  Init(nx * ny, d, 1, ops);
}

void RealSpinlessBoseSquareHamiltonian::Hop1(int p1, int p2, double c) {
  require(p1 < p2, "position1 >= position2 in DefineHamiltonian1", kErrArgument);
  for (int i = 0; i < num_site_; ++i) {
    int op = 0, t = 0;
    double k = 1.0;
    if (i == p1) { op = 1; k = 1.0; t = 1; }
    if (i > p1 && i < p2) { op = 0; k = 1.0; t = 1; }
    if (i == p2) { op = 2; k = c; t = 0; }
    ExpanTensorHamiltonian(i, op, k, QN{t});
  }
}

void RealSpinlessBoseSquareHamiltonian::Hop2(int p1, int p2, double c) {
  require(p1 < p2, "position1 >= position2 in DefineHamiltonian2", kErrArgument);
  for (int i = 0; i < num_site_; ++i) {
    int op = 0, t = 0;
    double k = 1.0;
    if (i == p1) { op = 2; k = 1.0; t = -1; }
    if (i > p1 && i < p2) { op = 0; k = 1.0; t = -1; }
    if (i == p2) { op = 1; k = c; t = 0; }
    ExpanTensorHamiltonian(i, op, k, QN{t});
  }
}

void RealSpinlessBoseSquareHamiltonian::OnSite(int p, double c) {
  for (int i = 0; i < num_site_; ++i) {
    if (i == p) ExpanTensorHamiltonian(i, 4, c, QN{0});
    else ExpanTensorHamiltonian(i, 0, 1.0, QN{0});
  }
}

void RealSpinlessBoseSquareHamiltonian::DefineTensorHamiltonian() {
  // real_spinless_bose_square_hamiltonian.cc:171-228; site p = j + i*ny
  for (int i = 0; i < nx_; ++i)
    for (int j = 0; j < ny_; ++j) {
      const int p0 = j + i * ny_;
      if (hop_x_ != 0) {
        const int p1 = j + ((i + 1) % nx_) * ny_;
        if (i + 1 < nx_) { Hop1(p0, p1, hop_x_); Hop2(p0, p1, hop_x_); }
        else if (nx_ > 2 && px_) { Hop1(p1, p0, hop_x_); Hop2(p1, p0, hop_x_); }
      }
      if (hop_y_ != 0) {
        const int p1 = (j + 1) % ny_ + i * ny_;
        if (j + 1 < ny_) { Hop1(p0, p1, hop_y_); Hop2(p0, p1, hop_y_); }
        else if (ny_ > 2 && py_) { Hop1(p1, p0, hop_y_); Hop2(p1, p0, hop_y_); }
      }
      if (mu_ != 0 || U_ != 0) OnSite(p0, 1.0);
      ParallelTensorHamiltonian();
    }
}

// ---------------------------------------------------------------------------
// Complex spinful Bose-Hubbard
// ---------------------------------------------------------------------------
std::vector<QNTable> ComplexSpinfulBoseSquareSpace::DefineQuantumTable() const {
  // complex_spinful_bose_square_space.cc:28-96: species 0 (up) sits on even
  // MPS sites, so bond i has seen (i+1)/2 up and i/2 down sites; the table is
  // the rectangle of reachable (N_up, N_down), N_up fastest, ordered by
  // ReorderQuantumTable (complex_tensor_space.cc:719-741).
  const int L = num_site(), half = L / 2;
  std::vector<QNTable> tabs;
  for (int i = 0; i <= L; ++i) {
    int lo[2] = {0, 0}, hi[2] = {0, 0};
    double mean[2] = {0.0, 0.0};
    if (i > 0) {
      const int pos[2] = {(i + 1) / 2, i / 2};
      for (int t = 0; t < 2; ++t) {
        lo[t] = std::max(0, n_[t] - (half - pos[t]) * (d_ - 1));
        hi[t] = std::min((d_ - 1) * pos[t], n_[t]);
        mean[t] = (double)n_[t] / half * pos[t];
      }
    }
    const int n0 = hi[0] - lo[0] + 1, n1 = hi[1] - lo[1] + 1;
    require(n0 > 0 && n1 > 0, "empty quantum table (particle numbers out of range)", kErrArgument);
    QNTable t((size_t)n0 * n1);
    for (int t0 = lo[0]; t0 <= hi[0]; ++t0)
      for (int t1 = lo[1]; t1 <= hi[1]; ++t1) t[(t0 - lo[0]) + (size_t)(t1 - lo[1]) * n0] = QN{t0, t1};
    std::vector<double> dev;
    std::vector<int> idx;
    for (const QN& q : t) {
      dev.push_back(std::pow(q[0] - mean[0], 2) + std::pow(q[1] - mean[1], 2));
      idx.push_back((int)idx.size());
    }
    quicksort(dev, &idx, 1);
    QNTable r;
    for (int k : idx) r.push_back(t[k]);
    tabs.push_back(r);
  }
  return tabs;
}

int ComplexSpinfulBoseSquareSpace::CheckQuantumTable(int site, const QN& left, const QN& right) const {
  // complex_spinful_bose_square_space.cc:105-125: site parity picks the species
  const int shift = site % 2;
  for (int i = 0; i < d_; ++i) {
    const int a0 = shift == 0 ? i : 0, a1 = shift == 1 ? i : 0;
    if (left[0] + a0 == right[0] && left[1] + a1 == right[1]) return i;
  }
  return -1;
}

ComplexSpinfulBoseSquareHamiltonian::ComplexSpinfulBoseSquareHamiltonian(
    int nx, int ny, int d, const double flux[2], const double hop_x[2], const double hop_y[2],
    const double inter[3], const double mu[2], bool period_x, bool period_y)
    : nx_(nx), ny_(ny), px_(period_x), py_(period_y) {
  for (int t = 0; t < 2; ++t) {
    flux_[t] = flux[t];
    hop_x_[t] = hop_x[t];
    hop_y_[t] = hop_y[t];
    mu_[t] = mu[t];
  }
  for (int t = 0; t < 3; ++t) inter_[t] = inter[t];
  // DefineBasicTensor (complex_spinful_bose_square_hamiltonian.cc:45-70):
  // 0=I, 1=a^+, 2=a^-, 3=n, 4=mu n + U_uu/2 n(n-1), 5=mu n + U_dd/2 n(n-1).
  // Both on-site operators take the species-0 chemical potential, as the
  // reference does (:61).
  std::vector<std::vector<double>> ops(6, std::vector<double>((size_t)d * d, 0.0));
  for (int i = 0; i < d; ++i) {
    ops[0][i * d + i] = 1.0;
    ops[3][i * d + i] = i;
    ops[4][i * d + i] = mu_[0] * i + 0.5 * inter_[0] * i * (i - 1);
    ops[5][i * d + i] = mu_[0] * i + 0.5 * inter_[1] * i * (i - 1);
  }
  for (int i = 1; i < d; ++i) {
    ops[1][i * d + (i - 1)] = std::sqrt((double)i);
    ops[2][(i - 1) * d + i] = std::sqrt((double)i);
  }
  cplx_ = true;
  Init(2 * nx * ny, d, 2, ops);
}

void ComplexSpinfulBoseSquareHamiltonian::Hop1(int p1, int p2, double cr, double ci) {
  // DefineHamiltonian1 (:73-126): a^+_{p1} a^-_{p2}, channel QN +1 of the species of p1
  require(p1 < p2, "position1 >= position2 in DefineHamiltonian1", kErrArgument);
  const QN q = p1 % 2 == 0 ? QN{1, 0} : QN{0, 1};
  for (int i = 0; i < num_site_; ++i) {
    if (i == p1) ExpanTensorHamiltonian(i, 1, 1.0, 0.0, q);
    else if (i > p1 && i < p2) ExpanTensorHamiltonian(i, 0, 1.0, 0.0, q);
    else if (i == p2) ExpanTensorHamiltonian(i, 2, cr, ci, QN{0, 0});
    else ExpanTensorHamiltonian(i, 0, 1.0, 0.0, QN{0, 0});
  }
}

void ComplexSpinfulBoseSquareHamiltonian::Hop2(int p1, int p2, double cr, double ci) {
  // DefineHamiltonian2 (:129-181): a^-_{p1} a^+_{p2}, channel QN -1
  require(p1 < p2, "position1 >= position2 in DefineHamiltonian2", kErrArgument);
  const QN q = p1 % 2 == 0 ? QN{-1, 0} : QN{0, -1};
  for (int i = 0; i < num_site_; ++i) {
    if (i == p1) ExpanTensorHamiltonian(i, 2, 1.0, 0.0, q);
    else if (i > p1 && i < p2) ExpanTensorHamiltonian(i, 0, 1.0, 0.0, q);
    else if (i == p2) ExpanTensorHamiltonian(i, 1, cr, ci, QN{0, 0});
    else ExpanTensorHamiltonian(i, 0, 1.0, 0.0, QN{0, 0});
  }
}

void ComplexSpinfulBoseSquareHamiltonian::OnSite(int p, double c) {
  // DefineHamiltonian3 (:184-210): operator 4 on up sites, 5 on down sites
  for (int i = 0; i < num_site_; ++i) {
    if (i == p) ExpanTensorHamiltonian(i, p % 2 == 0 ? 4 : 5, c, 0.0, QN{0, 0});
    else ExpanTensorHamiltonian(i, 0, 1.0, 0.0, QN{0, 0});
  }
}

void ComplexSpinfulBoseSquareHamiltonian::Inter(int p1, int p2, double c) {
  // DefineHamiltonian4 (:213-240): n_{p1} n_{p2} c
  for (int i = 0; i < num_site_; ++i) {
    if (i == p1) ExpanTensorHamiltonian(i, 3, 1.0, 0.0, QN{0, 0});
    else if (i == p2) ExpanTensorHamiltonian(i, 3, c, 0.0, QN{0, 0});
    else ExpanTensorHamiltonian(i, 0, 1.0, 0.0, QN{0, 0});
  }
}

void ComplexSpinfulBoseSquareHamiltonian::DefineTensorHamiltonian() {
  // complex_spinful_bose_square_hamiltonian.cc:242-350; physical site
  // p = j + i*ny, MPS sites 2p (up) and 2p+1 (down).  y bonds carry
  // phase = exp(i 2 pi (i+1) flux): a^+_p a_q conj(phase) + a_p a^+_q phase.
  for (int i = 0; i < nx_; ++i)
    for (int j = 0; j < ny_; ++j) {
      const int p0 = j + i * ny_;
      for (int t = 0; t < 2; ++t) {
        if (hop_x_[t] == 0.0) continue;
        const int p1 = j + ((i + 1) % nx_) * ny_;
        if (i + 1 < nx_) {
          Hop1(2 * p0 + t, 2 * p1 + t, hop_x_[t], 0.0);
          Hop2(2 * p0 + t, 2 * p1 + t, hop_x_[t], 0.0);
        } else if (nx_ > 2 && px_) {
          Hop1(2 * p1 + t, 2 * p0 + t, hop_x_[t], 0.0);
          Hop2(2 * p1 + t, 2 * p0 + t, hop_x_[t], 0.0);
        }
      }
      for (int t = 0; t < 2; ++t) {
        if (hop_y_[t] == 0.0) continue;
        const int p1 = (j + 1) % ny_ + i * ny_;
        const double theta = 2 * M_PI * (i + 1) * flux_[t];
        const double c = std::cos(theta), s = std::sin(theta);
        if (j + 1 < ny_) {
          Hop1(2 * p0 + t, 2 * p1 + t, c * hop_y_[t], -s * hop_y_[t]);
          Hop2(2 * p0 + t, 2 * p1 + t, c * hop_y_[t], s * hop_y_[t]);
        } else if (ny_ > 2 && py_) {
          Hop2(2 * p1 + t, 2 * p0 + t, c * hop_y_[t], -s * hop_y_[t]);
          Hop1(2 * p1 + t, 2 * p0 + t, c * hop_y_[t], s * hop_y_[t]);
        }
      }
      if (mu_[0] != 0.0 || inter_[0] != 0.0) OnSite(2 * p0, 1.0);
      if (mu_[1] != 0.0 || inter_[1] != 0.0) OnSite(2 * p0 + 1, 1.0);
      if (inter_[2] != 0.0) Inter(2 * p0, 2 * p0 + 1, inter_[2]);
      ParallelTensorHamiltonian();
    }
}

}  // namespace dmrg
