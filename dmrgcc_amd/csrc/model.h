// Model layer of the host API: quantum-number tables, the QN rule and the
// incremental MPO builder, mirroring the reference's abstract
// RealTensorSpace / RealTensorHamiltonian interface
// (src/dmrg/realdmrg/real_tensor_space.h:137-150,
//  src/dmrg/realdmrg/real_tensor_hamiltonian.h:6-78) and its Bose-Hubbard
// subclasses: real spinless (model/space/real_spinless_bose_square_space.*,
// model/hamiltonian/real_spinless_bose_square_hamiltonian.*), complex
// spinless and complex spinful (model/*/complex_spin*_bose_square_*).  The
// complex twins (ComplexTensorSpace / ComplexTensorHamiltonian) have the same
// interface; a Hamiltonian built complex() stores the imaginary parts of its
// MPO in SiteOperator::wi.
#pragma once

#include <cstddef>
#include <string>
#include <vector>

namespace dmrg {

using QN = std::vector<int>;
using QNTable = std::vector<QN>;

// Dense operator of one MPO site: W[o1][o2](p1, p2), p1 = bra, p2 = ket,
// stored as ((o1*rb + o2)*d + p1)*d + p2.
// Complex operators (ComplexTensorOperator) keep the imaginary parts in wi
// (same layout); wi is empty for a real operator.
struct SiteOperator {
  int lb = 0, rb = 0, d = 0;
  std::vector<double> w, wi;
  double at(int o1, int o2, int p1, int p2) const {
    return w[(((size_t)o1 * rb + o2) * d + p1) * d + p2];
  }
  double& at(int o1, int o2, int p1, int p2) {
    return w[(((size_t)o1 * rb + o2) * d + p1) * d + p2];
  }
  double im(int o1, int o2, int p1, int p2) const {
    return wi.empty() ? 0.0 : wi[(((size_t)o1 * rb + o2) * d + p1) * d + p2];
  }
  bool complex() const { return !wi.empty(); }
  // an entry the contractions use (real: w != 0; complex: re != 0 or im != 0,
  // complex_tensor_contraction.cc:1288)
  bool used(int o1, int o2, int p1, int p2) const {
    return at(o1, o2, p1, p2) != 0.0 || im(o1, o2, p1, p2) != 0.0;
  }
  // RealTensorOperator::{Left,Right,Mixed}CheckZero
  // (real_tensor_operator.cc:363-409; complex: Norm > 1e-15,
  // complex_tensor_operator.cc:363-409): any |entry| > 1e-15 along the
  // remaining indices.
  bool left_nonzero(int o1, int p1) const;
  bool right_nonzero(int o2, int p2) const;
  bool mixed_nonzero(int o1, int p2) const;
};

// RealTensorSpace: the abstract model-space interface.
class RealTensorSpace {
 public:
  virtual ~RealTensorSpace() = default;
  virtual int num_site() const = 0;
  virtual int physics_dim() const = 0;
  virtual int num_quantum() const = 0;
  // DefineQuantumTable: one QN table per MPS bond (num_site + 1 of them).
  virtual std::vector<QNTable> DefineQuantumTable() const = 0;
  // MergeQuantumTable: QN of an MPO channel applied to an MPS sector.
  virtual QN MergeQuantumTable(const QN& op, const QN& space) const = 0;
  // CheckQuantumTable: physical index connecting two sectors, or -1.
  virtual int CheckQuantumTable(int site, const QN& left, const QN& right) const = 0;
};

// RealTensorHamiltonian: owns the MPO and its bond QN tables and provides
// the incremental construction used by model subclasses.
class RealTensorHamiltonian {
 public:
  virtual ~RealTensorHamiltonian() = default;
  virtual void DefineTensorHamiltonian() = 0;

  int num_site() const { return num_site_; }
  int num_quantum() const { return nq_; }
  bool complex() const { return cplx_; }
  const std::vector<SiteOperator>& sites() const { return W_; }
  const std::vector<QNTable>& bond_tables() const { return qt_; }

  // MPO files: RealTensorHamiltonian::Write/ReadTensorHamiltonian
  // (real_tensor_hamiltonian.cc:74-150) with RealTensorOperator::Write/
  // ReadTensorOperator (real_tensor_operator.cc:96-139).  Layout: int
  // num_site, num_site+1, num_site-1; int num_operator and the basic
  // operators (RealMatrix::WriteMatrix: rows, cols, rows*cols, entries);
  // int num_quantum, the bond table sizes and entries; per site int
  // physics_dim, left_bond, right_bond and the left_bond*right_bond matrices
  // W[o1][o2] (d x d).  Entries are 8-byte doubles, or 16-byte Complex on a
  // complex Hamiltonian; native byte order.  Reading replaces this
  // Hamiltonian's MPO (same number of sites, quantum numbers and arithmetic
  // kind required).
  void WriteTensorHamiltonian(const std::string& path) const;
  void ReadTensorHamiltonian(const std::string& path);

 protected:
  void Init(int num_site, int physics_dim, int num_quantum, std::vector<std::vector<double>> basic);
  // ExpanTensorHamiltonian (real_tensor_hamiltonian.cc:151-187; complex
  // coefficient cr + i ci: complex_tensor_hamiltonian.cc:151-187)
  void ExpanTensorHamiltonian(int site, int op_index, double coefficient, const QN& table) {
    ExpanTensorHamiltonian(site, op_index, coefficient, 0.0, table);
  }
  void ExpanTensorHamiltonian(int site, int op_index, double cr, double ci, const QN& table);
  // ParallelTensorHamiltonian (real_tensor_hamiltonian.cc:189-214,
  // complex_tensor_hamiltonian.cc:189-214)
  void ParallelTensorHamiltonian();

  bool cplx_ = false;  // set by complex models before DefineTensorHamiltonian
  int num_site_ = 0, d_ = 0, nq_ = 0;
  std::vector<std::vector<double>> basic_;  // d*d row-major operators
  // imaginary parts of the basic operators of a complex MPO read from a
  // file (the models' own basic operators are real); empty otherwise
  std::vector<std::vector<double>> basic_im_;
  std::vector<SiteOperator> W_;
  std::vector<QNTable> qt_;
};

// RealMatrix::ParallelMatrix on a row-major rows x cols matrix
// (real_matrix.cc:438-531).  leigh 0 = columns, 1 = rows.  Returns kept
// positions; transfer is n x n row-major (n = cols or rows).
std::vector<int> parallel_matrix(const std::vector<double>& a, int rows, int cols, int leigh,
                                 std::vector<double>& transfer);
// ComplexMatrix::ParallelMatrix (complex_matrix.cc:451-540) on interleaved
// (re, im) entries; transfer is n x n interleaved complex.
std::vector<int> zparallel_matrix(const std::vector<double>& a, int rows, int cols, int leigh,
                                  std::vector<double>& transfer);

class RealSpinlessBoseSquareSpace : public RealTensorSpace {
 public:
  RealSpinlessBoseSquareSpace(int nx, int ny, int num_boson, int physics_dim)
      : nx_(nx), ny_(ny), n_(num_boson), d_(physics_dim) {}
  int num_site() const override { return nx_ * ny_; }
  int physics_dim() const override { return d_; }
  int num_quantum() const override { return 1; }
  std::vector<QNTable> DefineQuantumTable() const override;
  QN MergeQuantumTable(const QN& op, const QN& space) const override { return QN{op[0] + space[0]}; }
  int CheckQuantumTable(int site, const QN& left, const QN& right) const override;

 private:
  int nx_, ny_, n_, d_;
};

class RealSpinlessBoseSquareHamiltonian : public RealTensorHamiltonian {
 public:
  RealSpinlessBoseSquareHamiltonian(int nx, int ny, int physics_dim, double flux, double hop_x,
                                    double hop_y, double U, double mu, bool period_x,
                                    bool period_y);
  void DefineTensorHamiltonian() override;

 private:
  void Hop1(int p1, int p2, double c);
  void Hop2(int p1, int p2, double c);
  void OnSite(int p, double c);
  int nx_, ny_;
  double hop_x_, hop_y_, U_, mu_;
  bool px_, py_;
};

// The complex spinless model is the real one on Complex storage
// (model/space/complex_spinless_bose_square_space.cc and
// model/hamiltonian/complex_spinless_bose_square_hamiltonian.cc are the real
// files with Complex coefficients; the flux argument is stored and unused).
using ComplexSpinlessBoseSquareSpace = RealSpinlessBoseSquareSpace;
class ComplexSpinlessBoseSquareHamiltonian : public RealSpinlessBoseSquareHamiltonian {
 public:
  ComplexSpinlessBoseSquareHamiltonian(int nx, int ny, int physics_dim, double flux, double hop_x,
                                       double hop_y, double U, double mu, bool period_x,
                                       bool period_y)
      : RealSpinlessBoseSquareHamiltonian(nx, ny, physics_dim, flux, hop_x, hop_y, U, mu, period_x,
                                          period_y) {
    cplx_ = true;
  }
};

// Two-component bosons (model/space/complex_spinful_bose_square_space.cc):
// physical site p becomes MPS sites 2p (up) and 2p+1 (down), each of local
// dimension physics_dim; U(1) x U(1) quantum numbers (N_up, N_down).
class ComplexSpinfulBoseSquareSpace : public RealTensorSpace {
 public:
  ComplexSpinfulBoseSquareSpace(int nx, int ny, int n_up, int n_down, int physics_dim)
      : nx_(nx), ny_(ny), d_(physics_dim) {
    n_[0] = n_up;
    n_[1] = n_down;
  }
  int num_site() const override { return 2 * nx_ * ny_; }
  int physics_dim() const override { return d_; }
  int num_quantum() const override { return 2; }
  std::vector<QNTable> DefineQuantumTable() const override;
  QN MergeQuantumTable(const QN& op, const QN& space) const override {
    return QN{op[0] + space[0], op[1] + space[1]};
  }
  int CheckQuantumTable(int site, const QN& left, const QN& right) const override;

 private:
  int nx_, ny_, d_;
  int n_[2];
};

// model/hamiltonian/complex_spinful_bose_square_hamiltonian.cc: per species
// t (0 = up, 1 = down) hopping hop_x[t], hop_y[t] with the Landau-gauge phase
// exp(i 2 pi (x+1) flux[t]) on y bonds, on-site mu n + U_tt/2 n(n-1), and
// the inter-species U_ud n_up n_down.  inter = {U_uu, U_dd, U_ud}.
class ComplexSpinfulBoseSquareHamiltonian : public RealTensorHamiltonian {
 public:
  ComplexSpinfulBoseSquareHamiltonian(int nx, int ny, int physics_dim, const double flux[2],
                                      const double hop_x[2], const double hop_y[2],
                                      const double inter[3], const double mu[2], bool period_x,
                                      bool period_y);
  void DefineTensorHamiltonian() override;

 private:
  void Hop1(int p1, int p2, double cr, double ci);
  void Hop2(int p1, int p2, double cr, double ci);
  void OnSite(int p, double c);
  void Inter(int p1, int p2, double c);
  int nx_, ny_;
  double flux_[2], hop_x_[2], hop_y_[2], inter_[3], mu_[2];
  bool px_, py_;
};

}  // namespace dmrg
