// Model layer: QN tables, QN rule, MPO construction and deparallelisation.
#include "model.h"

#include <algorithm>
#include <cmath>

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

// isParallelElement (src/util/general.h:146-180) applied along a whole
// column: returns true (and the prefactor) when `col` is parallel to `ref`.
static bool parallel_to(const double* col, const double* ref, int n, int stride,
                        double& prefactor) {
  bool zero = true;
  prefactor = 0.0;
  for (int i = 0; i < n; ++i) {
    const double e0 = col[(size_t)i * stride], e1 = ref[(size_t)i * stride];
    if (std::fabs(e1) <= 1e-10) {
      if (std::fabs(e0) > 1e-10) return false;
    } else {
      const double f = e0 / e1;
      if (zero) {
        prefactor = f;
        zero = false;
      } else if (std::fabs(f - prefactor) > 1e-8) {
        return false;
      }
    }
  }
  return true;
}

std::vector<int> parallel_matrix(const std::vector<double>& a, int rows, int cols, int leigh,
                                 std::vector<double>& transfer) {
  // Work on "vectors": columns (leigh 0) or rows (leigh 1).
  const int nv = leigh == 0 ? cols : rows;
  const int len = leigh == 0 ? rows : cols;
  const int stride = leigh == 0 ? cols : 1;
  auto vec = [&](int v) { return leigh == 0 ? &a[v] : &a[(size_t)v * cols]; };
  std::vector<int> pos{0};
  transfer.assign((size_t)nv * nv, 0.0);
  auto T = [&](int k, int v) -> double& {
    // leigh 0: transfer(k, v); leigh 1: transfer(v, k)
    return leigh == 0 ? transfer[(size_t)k * nv + v] : transfer[(size_t)v * nv + k];
  };
  T(0, 0) = 1.0;
  for (int v = 0; v < nv; ++v) {
    bool found = false;
    for (size_t k = 0; k < pos.size(); ++k) {
      double pref;
      if (parallel_to(vec(v), vec(pos[k]), len, stride, pref)) {
        T((int)k, v) = pref;
        found = true;
        break;
      }
    }
    if (!found) {
      pos.push_back(v);
      T((int)pos.size() - 1, v) = 1.0;
    }
  }
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

void RealTensorHamiltonian::ExpanTensorHamiltonian(int site, int op_index, double coefficient,
                                                   const QN& table) {
  // RealTensorOperator::ExpanTensorOperator (real_tensor_operator.cc:141-197)
  int leigh = -1;
  if (site == 0) leigh = 0;
  if (site == num_site_ - 1) leigh = 1;
  SiteOperator& W = W_[site];
  const int d = d_;
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
  n.w.assign((size_t)n.lb * n.rb * d * d, 0.0);
  for (int l = 0; l < W.lb; ++l)
    for (int r = 0; r < W.rb; ++r)
      for (int p = 0; p < d * d; ++p) n.w[((size_t)l * n.rb + r) * d * d + p] = W.w[((size_t)l * W.rb + r) * d * d + p];
  const int il = e0 * W.lb, ir = e1 * W.rb;
  for (int p = 0; p < d * d; ++p)
    n.w[((size_t)il * n.rb + ir) * d * d + p] = coefficient * basic_[op_index][p];
  W = std::move(n);
  if (site < num_site_ - 1) qt_[site + 1].push_back(table);
}

void RealTensorHamiltonian::ParallelTensorHamiltonian() {
  const int d = d_, dd = d_ * d_;
  for (int i = 0; i < num_site_ - 1; ++i) {
    // LeftParallelTensorOperator (real_tensor_operator.cc:199-236): rows (l,p1,p2)
    SiteOperator& W = W_[i];
    const int rows = W.lb * dd, cols = W.rb;
    std::vector<double> resh((size_t)rows * cols);
    for (int l = 0; l < W.lb; ++l)
      for (int r = 0; r < W.rb; ++r)
        for (int p = 0; p < dd; ++p) resh[((size_t)l * dd + p) * cols + r] = W.w[((size_t)l * W.rb + r) * dd + p];
    std::vector<double> T;
    std::vector<int> pos = parallel_matrix(resh, rows, cols, 0, T);
    const int np = (int)pos.size();
    SiteOperator n;
    n.d = d; n.lb = W.lb; n.rb = np;
    n.w.assign((size_t)n.lb * np * dd, 0.0);
    for (int l = 0; l < W.lb; ++l)
      for (int r = 0; r < np; ++r)
        for (int p = 0; p < dd; ++p) n.w[((size_t)l * np + r) * dd + p] = resh[((size_t)l * dd + p) * cols + pos[r]];
    W = std::move(n);
    // LeftMergeTensorOperator (real_tensor_operator.cc:277-318): W'[l] = sum_m T[l][m] W[m]
    SiteOperator& X = W_[i + 1];
    require(cols == X.lb, "merge dimension is not match in LeftMergeTensorOperator");
    SiteOperator m;
    m.d = d; m.lb = np; m.rb = X.rb;
    m.w.assign((size_t)np * X.rb * dd, 0.0);
    for (int l = 0; l < np; ++l)
      for (int r = 0; r < X.rb; ++r)
        for (int p = 0; p < dd; ++p) {
          double e = 0.0;
          for (int k = 0; k < cols; ++k) e += T[(size_t)l * cols + k] * X.w[((size_t)k * X.rb + r) * dd + p];
          m.w[((size_t)l * X.rb + r) * dd + p] = e;
        }
    X = std::move(m);
    QNTable q;
    for (int p : pos) q.push_back(qt_[i + 1][p]);
    qt_[i + 1] = q;
  }
  for (int i = num_site_ - 1; i > 0; --i) {
    // RightParallelTensorOperator (real_tensor_operator.cc:238-275): columns (r,p1,p2)
    SiteOperator& W = W_[i];
    const int rows = W.lb, cols = W.rb * dd;
    std::vector<double> resh(W.w);  // (l, r, p1, p2) row-major == rows x cols
    std::vector<double> T;
    std::vector<int> pos = parallel_matrix(resh, rows, cols, 1, T);
    const int np = (int)pos.size();
    SiteOperator n;
    n.d = d; n.lb = np; n.rb = W.rb;
    n.w.assign((size_t)np * W.rb * dd, 0.0);
    for (int l = 0; l < np; ++l)
      for (int c = 0; c < cols; ++c) n.w[(size_t)l * cols + c] = resh[(size_t)pos[l] * cols + c];
    W = std::move(n);
    // RightMergeTensorOperator (real_tensor_operator.cc:320-361): W'[:, r] = sum_m T[m][r] W[:, m]
    SiteOperator& X = W_[i - 1];
    require(rows == X.rb, "merge dimension is not match in RightMergeTensorOperator");
    SiteOperator m;
    m.d = d; m.lb = X.lb; m.rb = np;
    m.w.assign((size_t)X.lb * np * dd, 0.0);
    for (int l = 0; l < X.lb; ++l)
      for (int r = 0; r < np; ++r)
        for (int p = 0; p < dd; ++p) {
          double e = 0.0;
          for (int k = 0; k < rows; ++k) e += T[(size_t)k * rows + r] * X.w[((size_t)l * X.rb + k) * dd + p];
          m.w[((size_t)l * np + r) * dd + p] = e;
        }
    X = std::move(m);
    QNTable q;
    for (int p : pos) q.push_back(qt_[i][p]);
    qt_[i] = q;
  }
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

}  // namespace dmrg
