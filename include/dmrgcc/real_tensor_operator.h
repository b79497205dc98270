// Drop-in RealTensorOperator (src/dmrg/realdmrg/real_tensor_operator.h:6-95):
// one MPO site as a left_bond x right_bond grid of physics_dim x
// physics_dim matrices, W[o1][o2](p1 = bra, p2 = ket).
#ifndef DMRGCC_DROPIN_REAL_TENSOR_OPERATOR_H_
#define DMRGCC_DROPIN_REAL_TENSOR_OPERATOR_H_

#include <vector>

#include "real_matrix.h"

class RealTensorOperator {
 public:
  RealTensorOperator() = default;
  explicit RealTensorOperator(int physics_dim) : physics_dim_(physics_dim) {}
  int get_physics_dim() const { return physics_dim_; }
  int get_left_bond() const { return left_bond_; }
  int get_right_bond() const { return right_bond_; }
  // zero operator of the given bond dimensions (DefineTensorOperator with
  // the bonds a Hamiltonian's expansion would reach)
  void DefineTensorOperator(int left_bond, int right_bond) {
    left_bond_ = left_bond;
    right_bond_ = right_bond;
    w_.assign((size_t)left_bond * right_bond, RealMatrix(physics_dim_, physics_dim_));
  }
  // tensor_operator_[o1][o2]
  RealMatrix* get_tensor_operator(int o1, int o2) { return &w_[(size_t)o1 * right_bond_ + o2]; }
  // the dense lb x rb x d x d array the C ABI takes
  std::vector<double> Dense() const {
    const int d = physics_dim_;
    std::vector<double> out((size_t)left_bond_ * right_bond_ * d * d);
    for (size_t k = 0; k < w_.size(); ++k)
      for (int e = 0; e < d * d; ++e) out[k * d * d + e] = w_[k].get_matrix_element()[e];
    return out;
  }

 private:
  int physics_dim_ = 0, left_bond_ = 0, right_bond_ = 0;
  std::vector<RealMatrix> w_;
};

#endif  // DMRGCC_DROPIN_REAL_TENSOR_OPERATOR_H_
