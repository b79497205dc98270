// Drop-in RealMatrix / RealMatrixBlock (src/tensor/real_matrix.h,
// src/tensor/real_matrix_block.h): host matrices in the reference's layout
// (row-major doubles) holding the data the hot-path members exchange with the
// MI355X library.  Only the members the hot path uses are provided.
#ifndef DMRGCC_DROPIN_REAL_MATRIX_H_
#define DMRGCC_DROPIN_REAL_MATRIX_H_

#include <cmath>
#include <cstdlib>
#include <memory>
#include <vector>

class RealMatrix {
 public:
  RealMatrix() = default;
  // zero-filled row x column matrix (real_matrix.cc:15-25)
  RealMatrix(int row, int column) : row_(row), column_(column), e_((size_t)row * column, 0.0) {}
  explicit RealMatrix(const RealMatrix* m) : row_(m->row_), column_(m->column_), e_(m->e_) {}
  int get_row() const { return row_; }
  int get_column() const { return column_; }
  int get_total_element_num() const { return row_ * column_; }
  double* get_matrix_element() { return e_.data(); }
  const double* get_matrix_element() const { return e_.data(); }
  double get_matrix_element(int i, int j) const { return e_[(size_t)i * column_ + j]; }
  void set_matrix_element(int i, int j, double v) { e_[(size_t)i * column_ + j] = v; }
  // RealMatrix::RandomMatrix (real_matrix.cc:201-208): 0.01 * (rand() % 100),
  // filled column by column from the process's rand() stream
  void RandomMatrix() {
    for (int j = 0; j < column_; ++j)
      for (int i = 0; i < row_; ++i) set_matrix_element(i, j, 0.01 * (std::rand() % 100));
  }
  double SumSquareMatrix() const {
    double s = 0.0;
    for (double x : e_) s += x * x;
    return s;
  }
  void MultiplyToScalar(double f) {
    for (double& x : e_) x *= f;
  }

 private:
  int row_ = 0, column_ = 0;
  std::vector<double> e_;
};

class RealMatrixBlock {
 public:
  RealMatrixBlock() = default;
  explicit RealMatrixBlock(int num_block) : left_(num_block, 0), right_(num_block, 0) { fill_empty(); }
  RealMatrixBlock(int num_block, const int* left_block, const int* right_block)
      : left_(left_block, left_block + num_block), right_(right_block, right_block + num_block) {
    fill_empty();
  }
  int get_num_block() const { return (int)left_.size(); }
  int* get_left_block() { return left_.data(); }
  int* get_right_block() { return right_.data(); }
  RealMatrix* get_matrix_block(int i) { return m_[i].get(); }
  const RealMatrix* get_matrix_block(int i) const { return m_[i].get(); }
  // a copy of `m` becomes block i (RealMatrixBlock::set_matrix_block)
  void set_matrix_block(int i, const RealMatrix* m) { m_[i].reset(new RealMatrix(m)); }
  // FindMatrixBlock (real_matrix_block.cc:314-380): first block (left,
  // right), or -1
  int FindMatrixBlock(int left, int right) const {
    for (int i = 0; i < get_num_block(); ++i)
      if (left_[i] == left && right_[i] == right) return i;
    return -1;
  }
  // all blocks whose left (leigh 0) or right (leigh 1) sector is target;
  // the caller deletes[] positions, as in the reference
  void FindMatrixBlock(int& num, int*& positions, int leigh, int target) const {
    std::vector<int> p;
    for (int i = 0; i < get_num_block(); ++i)
      if ((leigh == 0 ? left_[i] : right_[i]) == target) p.push_back(i);
    num = (int)p.size();
    positions = new int[p.size() + 1];
    for (size_t k = 0; k < p.size(); ++k) positions[k] = p[k];
  }
  int ComputeMatrixBlockDim() const { return ComputePartMatrixBlockDim(get_num_block()); }
  int ComputePartMatrixBlockDim(int position) const {
    int d = 0;
    for (int i = 0; i < position; ++i) d += m_[i]->get_total_element_num();
    return d;
  }
  // VectorizeMatrixBlock (real_matrix_block.cc:173-194): blocks to state
  // (direction true) or state to blocks (false), blocks back to back
  void VectorizeMatrixBlock(bool direction, double* state) {
    int off = 0;
    for (auto& m : m_) {
      double* e = m->get_matrix_element();
      const int n = m->get_total_element_num();
      for (int j = 0; j < n; ++j) {
        if (direction) state[off + j] = e[j];
        else e[j] = state[off + j];
      }
      off += n;
    }
  }
  // NormalizeMatrixBlock (real_matrix_block.cc:164-171)
  void NormalizeMatrixBlock() {
    double s = 0.0;
    for (auto& m : m_) s += m->SumSquareMatrix();
    const double f = std::sqrt(1.0 / s);
    for (auto& m : m_) m->MultiplyToScalar(f);
  }

 private:
  void fill_empty() {
    m_.clear();
    for (size_t i = 0; i < left_.size(); ++i) m_.emplace_back(new RealMatrix());
  }
  std::vector<int> left_, right_;
  std::vector<std::unique_ptr<RealMatrix>> m_;
};

#endif  // DMRGCC_DROPIN_REAL_MATRIX_H_
