// Drop-in RealTensorLattice (src/dmrg/realdmrg/real_tensor_lattice.h:8-252):
// one MPS site on the host in the reference's layout (sector lists, physics
// index table, ket blocks, canonical tensor).  The canonical forms run on
// the MI355X through dmrgcc_lattice_canonical.
#ifndef DMRGCC_DROPIN_REAL_TENSOR_LATTICE_H_
#define DMRGCC_DROPIN_REAL_TENSOR_LATTICE_H_

#include <memory>
#include <vector>

#include "device.h"
#include "real_matrix.h"

class RealTensorLattice {
 public:
  RealTensorLattice() = default;
  explicit RealTensorLattice(int physics_dim) : physics_dim_(physics_dim) {}

  int get_physics_dim() const { return physics_dim_; }
  int get_num_left_block() const { return (int)left_block_.size(); }
  int* get_left_block() { return left_block_.data(); }
  int* get_left_dim() { return left_dim_.data(); }
  int get_num_right_block() const { return (int)right_block_.size(); }
  int* get_right_block() { return right_block_.data(); }
  int* get_right_dim() { return right_dim_.data(); }
  int** get_physics_index() { return phys_rows_.data(); }
  RealMatrixBlock* get_ket_tensor() { return ket_tensor_.get(); }
  RealMatrixBlock* get_canonical_tensor() { return canonical_tensor_.get(); }
  int* get_match_dim() { return match_dim_.empty() ? nullptr : match_dim_.data(); }

  // DefineTensorLattice (real_tensor_lattice.cc:300-340): sector lists, all
  // physics indices -1
  void DefineTensorLattice(int num_left_block, int num_right_block, const int* left_block,
                           const int* right_block, const int* left_dim, const int* right_dim) {
    left_block_.assign(left_block, left_block + num_left_block);
    left_dim_.assign(left_dim, left_dim + num_left_block);
    right_block_.assign(right_block, right_block + num_right_block);
    right_dim_.assign(right_dim, right_dim + num_right_block);
    phys_.assign((size_t)num_left_block * num_right_block, -1);
    phys_rows_.resize(num_left_block);
    for (int i = 0; i < num_left_block; ++i) phys_rows_[i] = phys_.data() + (size_t)i * num_right_block;
  }
  // DefineTensorLattice (real_tensor_lattice.cc:343-359): blocks at sector
  // positions (left_index[i], right_index[i]) with their physics index, each
  // a fresh RandomMatrix(row, column) from the process's rand()
  void DefineTensorLattice(int num_block, const int* left_index, const int* right_index,
                           const int* physics_index, int row, int column) {
    ket_tensor_.reset(new RealMatrixBlock(num_block, left_index, right_index));
    RealMatrix m(row, column);
    for (int i = 0; i < num_block; ++i) {
      phys_rows_[left_index[i]][right_index[i]] = physics_index[i];
      m.RandomMatrix();
      ket_tensor_->set_matrix_block(i, &m);
    }
  }

  int ComputeKetTensorDim() const { return ket_tensor_ ? ket_tensor_->ComputeMatrixBlockDim() : 0; }
  int ComputePartKetTensorDim(int position) const { return ket_tensor_->ComputePartMatrixBlockDim(position); }
  void NormalizeTensorLattice() { ket_tensor_->NormalizeMatrixBlock(); }
  void VectorizeTensorLattice(bool direction, double* state) { ket_tensor_->VectorizeMatrixBlock(direction, state); }

  // LeftCanonicalTensorLattice (real_tensor_lattice.cc:558-676) and
  // RightCanonicalTensorLattice (:794-911) on the device: the ket blocks
  // become the orthonormal factor (truncated by ComputeTruncateDim with
  // max_dim, canonical_precision), canonical_tensor_ receives Sigma V^T /
  // U Sigma per sector and match_dim_ the sector dimensions before the cut.
  void LeftCanonicalTensorLattice(int max_dim, double canonical_precision) {
    canonical(0, max_dim, canonical_precision);
  }
  void RightCanonicalTensorLattice(int max_dim, double canonical_precision) {
    canonical(1, max_dim, canonical_precision);
  }
  // discarded weight (sum of dropped sigma^2) of the last canonical form (no
  // counterpart in the reference, which computes it and drops it)
  double get_discarded_weight() const { return discarded_; }

 private:
  void canonical(int leigh, int max_dim, double canpre) {
    using namespace dmrgcc_dropin;
    const int nb = ket_tensor_->get_num_block();
    std::vector<int> kl(nb), kr(nb), rows(nb), cols(nb);
    std::vector<double> data(ComputeKetTensorDim());
    for (int b = 0; b < nb; ++b) {
      kl[b] = ket_tensor_->get_left_block()[b];
      kr[b] = ket_tensor_->get_right_block()[b];
      rows[b] = ket_tensor_->get_matrix_block(b)->get_row();
      cols[b] = ket_tensor_->get_matrix_block(b)->get_column();
    }
    ket_tensor_->VectorizeMatrixBlock(true, data.data());
    DeviceVector dk;
    dk.upload(data.data(), (int64_t)data.size());
    const int ns = leigh == 0 ? get_num_right_block() : get_num_left_block();
    std::vector<int> dims(ns), match(ns);
    dmrgcc_blocks *ko = nullptr, *co = nullptr;
    check(dmrgcc_lattice_canonical(leigh, 0, get_num_left_block(), get_num_right_block(), left_dim_.data(),
                                   right_dim_.data(), nb, kl.data(), kr.data(), rows.data(), cols.data(),
                                   dk.data(), max_dim, canpre, nullptr, &ko, &co, dims.data(), match.data(),
                                   &discarded_));
    // the new ket blocks, same block list
    std::unique_ptr<RealMatrixBlock> nk(new RealMatrixBlock(nb, kl.data(), kr.data()));
    fill_from(ko, *nk);
    // the canonical tensor: one block per sector
    std::vector<int> idx(ns);
    for (int s = 0; s < ns; ++s) idx[s] = s;
    std::unique_ptr<RealMatrixBlock> nc(new RealMatrixBlock(ns, idx.data(), idx.data()));
    fill_from(co, *nc);
    dmrgcc_blocks_destroy(ko);
    dmrgcc_blocks_destroy(co);
    ket_tensor_ = std::move(nk);
    canonical_tensor_ = std::move(nc);
    match_dim_ = match;
    (leigh == 0 ? right_dim_ : left_dim_) = dims;
  }
  // blocks of a library result handle, in order, into `dst`
  static void fill_from(dmrgcc_blocks* h, RealMatrixBlock& dst) {
    using namespace dmrgcc_dropin;
    const int n = dmrgcc_blocks_count(h);
    int64_t total = 0;
    std::vector<int64_t> offs(n);
    std::vector<int> rr(n), cc(n);
    for (int i = 0; i < n; ++i) {
      int o, l, r;
      check(dmrgcc_blocks_get(h, i, &o, &l, &r, &rr[i], &cc[i], &offs[i]));
      total += (int64_t)rr[i] * cc[i];
    }
    std::vector<double> host(total > 0 ? total : 1);
    check(dmrgcc_blocks_download(h, host.data()));
    for (int i = 0; i < n && i < dst.get_num_block(); ++i) {
      RealMatrix m(rr[i], cc[i]);
      for (int64_t e = 0; e < (int64_t)rr[i] * cc[i]; ++e) m.get_matrix_element()[e] = host[offs[i] + e];
      dst.set_matrix_block(i, &m);
    }
  }

  int physics_dim_ = 0;
  std::vector<int> left_block_, left_dim_, right_block_, right_dim_;
  std::vector<int> phys_;
  std::vector<int*> phys_rows_;
  std::unique_ptr<RealMatrixBlock> ket_tensor_;
  std::vector<int> match_dim_;
  std::unique_ptr<RealMatrixBlock> canonical_tensor_;
  double discarded_ = 0.0;
};

#endif  // DMRGCC_DROPIN_REAL_TENSOR_LATTICE_H_
