// Drop-in RealTensorContraction (src/dmrg/realdmrg/real_tensor_contraction.h:
// 8-96): the left and right environments of one site on the host, in the
// reference's layout (one RealMatrixBlock per MPO channel), with the hot-path
// members running on the MI355X:
//   MultiplyEffectHamilton   -> dmrgcc_heff_create / dmrgcc_heff_apply
//   Left/RightComputeTensorContraction -> dmrgcc_env_update
// The H_eff plan and the device copies of L and R are kept between calls
// and rebuilt when the site's block structure, the operator or the
// environments change (the environments through this class's own members,
// or after MarkModified() when a caller writes their blocks directly).
#ifndef DMRGCC_DROPIN_REAL_TENSOR_CONTRACTION_H_
#define DMRGCC_DROPIN_REAL_TENSOR_CONTRACTION_H_

#include <memory>
#include <vector>

#include "device.h"
#include "real_tensor_lattice.h"
#include "real_tensor_operator.h"

class RealTensorContraction {
 public:
  RealTensorContraction(int left_bond, int right_bond) : left_bond_(left_bond), right_bond_(right_bond) {
    for (int i = 0; i < left_bond; ++i) left_.emplace_back(new RealMatrixBlock(0));
    for (int i = 0; i < right_bond; ++i) right_.emplace_back(new RealMatrixBlock(0));
    refresh_views();
  }
  ~RealTensorContraction() { drop_plan(); }
  RealTensorContraction(const RealTensorContraction&) = delete;
  RealTensorContraction& operator=(const RealTensorContraction&) = delete;

  int get_left_bond() const { return left_bond_; }
  int get_right_bond() const { return right_bond_; }
  RealMatrixBlock** get_left_contraction_tensor() { return left_view_.data(); }
  RealMatrixBlock** get_right_contraction_tensor() { return right_view_.data(); }
  // environments edited through the getters above: rebuild the device copy
  void MarkModified() { ++version_; }

  // DefineTensorContraction (real_tensor_contraction.cc:111-137): the
  // boundary environment of leigh 0 (left) / 1 (right), one 1x1 block of
  // value 1 per channel
  void DefineTensorContraction(int leigh) {
    auto& side = leigh == 0 ? left_ : right_;
    const int zero = 0;
    RealMatrix one(1, 1);
    one.set_matrix_element(0, 0, 1.0);
    for (auto& b : side) {
      b.reset(new RealMatrixBlock(1, &zero, &zero));
      b->set_matrix_block(0, &one);
    }
    refresh_views();
    ++version_;
  }

  // MultiplyEffectHamilton (real_tensor_contraction.cc:1223-1291):
  // state += H_eff * ket, the ket read from the lattice, `state` in the
  // ket-vector layout (ComputeKetTensorDim doubles)
  void MultiplyEffectHamilton(RealTensorLattice* tensor_lattice, RealTensorOperator* tensor_operator,
                              double* state) {
    using namespace dmrgcc_dropin;
    dmrgcc_heff* h = plan(tensor_lattice, tensor_operator);
    const int64_t n = tensor_lattice->ComputeKetTensorDim();
    std::vector<double> ket(n);
    tensor_lattice->VectorizeTensorLattice(true, ket.data());
    x_.upload(ket.data(), n);
    y_.upload(state, n);
    check(dmrgcc_heff_apply(h, x_.data(), y_.data(), precision_));
    y_.download(state, n);
  }

  // The Lanczos solve of RealTensorLanczos::LanczosMethod on this site's
  // plan: the lattice's ket is the start vector and receives the ground
  // vector; returns the energy.
  double SolveGroundState(RealTensorLattice* tensor_lattice, RealTensorOperator* tensor_operator,
                          int num_iter) {
    using namespace dmrgcc_dropin;
    dmrgcc_heff* h = plan(tensor_lattice, tensor_operator);
    const int64_t n = tensor_lattice->ComputeKetTensorDim();
    std::vector<double> ket(n);
    tensor_lattice->VectorizeTensorLattice(true, ket.data());
    x_.upload(ket.data(), n);
    double e = 0.0;
    int it = 0;
    check(dmrgcc_lanczos(h, x_.data(), num_iter, precision_, 1, &e, &it));
    x_.download(ket.data(), n);
    tensor_lattice->VectorizeTensorLattice(false, ket.data());
    return e;
  }

  // LeftComputeTensorContraction (real_tensor_contraction.cc:782-940): the
  // left environment of the next site, from this one's and the lattice,
  // into tensor_contraction's left side
  void LeftComputeTensorContraction(RealTensorLattice* tensor_lattice, RealTensorOperator* tensor_operator,
                                    RealTensorContraction* tensor_contraction) {
    env_update(0, tensor_lattice, tensor_operator, left_, tensor_contraction->left_);
    tensor_contraction->refresh_views();
    ++tensor_contraction->version_;
  }
  // RightComputeTensorContraction (:959-1119): the right environment of the
  // previous site into tensor_contraction's right side
  void RightComputeTensorContraction(RealTensorLattice* tensor_lattice, RealTensorOperator* tensor_operator,
                                     RealTensorContraction* tensor_contraction) {
    env_update(1, tensor_lattice, tensor_operator, right_, tensor_contraction->right_);
    tensor_contraction->refresh_views();
    ++tensor_contraction->version_;
  }

  // arithmetic of the H_eff products (0 FP64, the reference's; 1 FP32;
  // 2 BF16x3); no counterpart in the reference
  void set_precision(int precision) { precision_ = precision; }

 private:
  typedef std::vector<std::unique_ptr<RealMatrixBlock>> Side;

  static dmrgcc_dropin::FlatBlocks flatten(Side& side) {
    dmrgcc_dropin::FlatBlocks f;
    int64_t acc = 0;
    for (size_t ch = 0; ch < side.size(); ++ch) {
      RealMatrixBlock& b = *side[ch];
      for (int i = 0; i < b.get_num_block(); ++i) {
        const RealMatrix* m = b.get_matrix_block(i);
        f.o.push_back((int)ch);
        f.l.push_back(b.get_left_block()[i]);
        f.r.push_back(b.get_right_block()[i]);
        f.rows.push_back(m->get_row());
        f.cols.push_back(m->get_column());
        f.off.push_back(acc);
        const int64_t n = (int64_t)m->get_row() * m->get_column();
        f.data.insert(f.data.end(), m->get_matrix_element(), m->get_matrix_element() + n);
        acc += n;
      }
    }
    if (f.data.empty()) f.data.push_back(0.0);
    return f;
  }

  struct SiteShape {
    std::vector<int> v;
    bool operator==(const SiteShape& o) const { return v == o.v; }
  };
  static SiteShape shape_of(RealTensorLattice* lat) {
    SiteShape s;
    const int nl = lat->get_num_left_block(), nr = lat->get_num_right_block();
    s.v = {nl, nr};
    for (int i = 0; i < nl; ++i)
      for (int j = 0; j < nr; ++j) s.v.push_back(lat->get_physics_index()[i][j]);
    RealMatrixBlock* k = lat->get_ket_tensor();
    for (int b = 0; b < k->get_num_block(); ++b) {
      s.v.push_back(k->get_left_block()[b]);
      s.v.push_back(k->get_right_block()[b]);
      s.v.push_back(k->get_matrix_block(b)->get_row());
      s.v.push_back(k->get_matrix_block(b)->get_column());
    }
    return s;
  }

  void site_args(RealTensorLattice* lat, std::vector<int>& phys, std::vector<int>& kl, std::vector<int>& kr,
                 std::vector<int>& rows, std::vector<int>& cols) {
    const int nl = lat->get_num_left_block(), nr = lat->get_num_right_block();
    phys.resize((size_t)nl * nr);
    for (int i = 0; i < nl; ++i)
      for (int j = 0; j < nr; ++j) phys[(size_t)i * nr + j] = lat->get_physics_index()[i][j];
    RealMatrixBlock* k = lat->get_ket_tensor();
    for (int b = 0; b < k->get_num_block(); ++b) {
      kl.push_back(k->get_left_block()[b]);
      kr.push_back(k->get_right_block()[b]);
      rows.push_back(k->get_matrix_block(b)->get_row());
      cols.push_back(k->get_matrix_block(b)->get_column());
    }
  }

  dmrgcc_heff* plan(RealTensorLattice* lat, RealTensorOperator* op) {
    using namespace dmrgcc_dropin;
    SiteShape shape = shape_of(lat);
    if (h_ && lat == plan_lat_ && op == plan_op_ && version_ == plan_version_ && shape == plan_shape_) return h_;
    drop_plan();
    FlatBlocks L = flatten(left_), R = flatten(right_);
    Ld_.upload(L.data.data(), (int64_t)L.data.size());
    Rd_.upload(R.data.data(), (int64_t)R.data.size());
    std::vector<int> phys, kl, kr, rows, cols;
    site_args(lat, phys, kl, kr, rows, cols);
    const std::vector<double> W = op->Dense();
    check(dmrgcc_heff_create(op->get_physics_dim(), op->get_left_bond(), op->get_right_bond(), W.data(),
                             lat->get_num_left_block(), lat->get_num_right_block(), phys.data(), (int)kl.size(),
                             kl.data(), kr.data(), rows.data(), cols.data(), L.n(), L.o.data(), L.l.data(),
                             L.r.data(), L.rows.data(), L.cols.data(), L.off.data(), Ld_.data(), R.n(),
                             R.o.data(), R.l.data(), R.r.data(), R.rows.data(), R.cols.data(), R.off.data(),
                             Rd_.data(), nullptr, &h_));
    plan_lat_ = lat;
    plan_op_ = op;
    plan_version_ = version_;
    plan_shape_ = shape;
    return h_;
  }
  void drop_plan() {
    if (h_) dmrgcc_heff_destroy(h_);
    h_ = nullptr;
  }

  void env_update(int leigh, RealTensorLattice* lat, RealTensorOperator* op, Side& from, Side& into) {
    using namespace dmrgcc_dropin;
    FlatBlocks E = flatten(from);
    DeviceVector Ed, Kd;
    Ed.upload(E.data.data(), (int64_t)E.data.size());
    std::vector<double> ket(lat->ComputeKetTensorDim() > 0 ? lat->ComputeKetTensorDim() : 1);
    lat->VectorizeTensorLattice(true, ket.data());
    Kd.upload(ket.data(), (int64_t)ket.size());
    std::vector<int> phys, kl, kr, rows, cols;
    site_args(lat, phys, kl, kr, rows, cols);
    const std::vector<double> W = op->Dense();
    dmrgcc_blocks* out = nullptr;
    check(dmrgcc_env_update(leigh, op->get_physics_dim(), op->get_left_bond(), op->get_right_bond(), W.data(),
                            lat->get_num_left_block(), lat->get_num_right_block(), phys.data(), (int)kl.size(),
                            kl.data(), kr.data(), rows.data(), cols.data(), Kd.data(), E.n(), E.o.data(),
                            E.l.data(), E.r.data(), E.rows.data(), E.cols.data(), E.off.data(), Ed.data(),
                            nullptr, &out));
    const int nb = dmrgcc_blocks_count(out);
    const int nch = leigh == 0 ? op->get_right_bond() : op->get_left_bond();
    std::vector<std::vector<int>> bl(nch), br(nch), bi(nch);
    std::vector<int> o(nb), l(nb), r(nb), rr(nb), cc(nb);
    std::vector<int64_t> off(nb);
    int64_t total = 0;
    for (int i = 0; i < nb; ++i) {
      check(dmrgcc_blocks_get(out, i, &o[i], &l[i], &r[i], &rr[i], &cc[i], &off[i]));
      bl[o[i]].push_back(l[i]);
      br[o[i]].push_back(r[i]);
      bi[o[i]].push_back(i);
      total += (int64_t)rr[i] * cc[i];
    }
    std::vector<double> host(total > 0 ? total : 1);
    check(dmrgcc_blocks_download(out, host.data()));
    dmrgcc_blocks_destroy(out);
    into.clear();
    for (int ch = 0; ch < nch; ++ch) {
      std::unique_ptr<RealMatrixBlock> b(new RealMatrixBlock((int)bl[ch].size(), bl[ch].data(), br[ch].data()));
      for (size_t k = 0; k < bi[ch].size(); ++k) {
        const int i = bi[ch][k];
        RealMatrix m(rr[i], cc[i]);
        for (int64_t e = 0; e < (int64_t)rr[i] * cc[i]; ++e) m.get_matrix_element()[e] = host[off[i] + e];
        b->set_matrix_block((int)k, &m);
      }
      into.push_back(std::move(b));
    }
  }

  void refresh_views() {
    left_view_.clear();
    right_view_.clear();
    for (auto& b : left_) left_view_.push_back(b.get());
    for (auto& b : right_) right_view_.push_back(b.get());
  }

  int left_bond_, right_bond_;
  Side left_, right_;
  std::vector<RealMatrixBlock*> left_view_, right_view_;
  int64_t version_ = 0;
  int precision_ = 0;
  // the cached H_eff plan and the device data it reads
  dmrgcc_heff* h_ = nullptr;
  RealTensorLattice* plan_lat_ = nullptr;
  RealTensorOperator* plan_op_ = nullptr;
  int64_t plan_version_ = -1;
  SiteShape plan_shape_;
  dmrgcc_dropin::DeviceVector Ld_, Rd_, x_, y_;
};

#endif  // DMRGCC_DROPIN_REAL_TENSOR_CONTRACTION_H_
