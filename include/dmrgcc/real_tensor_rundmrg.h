// Drop-in RealTensorRundmrg (src/dmrg/realdmrg/real_tensor_rundmrg.h:6-59)
// and the spinless Bose-Hubbard model classes its drivers build
// (model/space/real_spinless_bose_square_space.h,
// model/hamiltonian/real_spinless_bose_square_hamiltonian.h).  The whole
// sweep runs on the MI355X through the library's resumable run
// (dmrgcc_dmrg_create_files / _advance / _result); the constructor reads the
// run section of the input file as the reference's does (cache settings,
// first_block, first_dim, num_sweep, the sweeps table with its fill-in).
// disk_cache is read and ignored: environments stay in device memory.
#ifndef DMRGCC_DROPIN_REAL_TENSOR_RUNDMRG_H_
#define DMRGCC_DROPIN_REAL_TENSOR_RUNDMRG_H_

#include <chrono>
#include <cstdio>
#include <string>
#include <vector>

#include "input.h"
#include "real_tensor_network.h"

// The library model behind a space + Hamiltonian pair.
class DeviceModelSource {
 public:
  virtual ~DeviceModelSource() = default;
  virtual dmrgcc_model* CreateModel(const RealTensorSpace& space) const = 0;
};

// RealSpinlessBoseSquareSpace (model/space/real_spinless_bose_square_space.cc)
class RealSpinlessBoseSquareSpace : public RealTensorSpace {
 public:
  RealSpinlessBoseSquareSpace(int num_site_x, int num_site_y, int num_boson, int physics_dim)
      : RealTensorSpace(num_site_x * num_site_y, physics_dim), nx_(num_site_x), ny_(num_site_y),
        num_boson_(num_boson) {}
  int get_num_boson() const { return num_boson_; }

 private:
  int nx_, ny_, num_boson_;
};

// RealSpinlessBoseSquareHamiltonian
// (model/hamiltonian/real_spinless_bose_square_hamiltonian.cc): flux is
// accepted and unused, as in the reference.
class RealSpinlessBoseSquareHamiltonian : public RealTensorHamiltonian, public DeviceModelSource {
 public:
  RealSpinlessBoseSquareHamiltonian(int num_site_x, int num_site_y, int physics_dim, double flux, double Jx,
                                    double Jy, double U, double mu, bool period_x, bool period_y)
      : RealTensorHamiltonian(num_site_x * num_site_y, physics_dim), nx_(num_site_x), ny_(num_site_y), d_(physics_dim),
        flux_(flux), jx_(Jx), jy_(Jy), U_(U), mu_(mu), px_(period_x), py_(period_y) {}
  // DefineTensorHamiltonian (:171-228): the deparallelised MPO, one operator
  // per site (the particle number does not enter the MPO)
  void DefineTensorHamiltonian() override {
    dmrgcc_model* m = make(nx_ * ny_);
    for (int s = 0; s < num_site_; ++s) {
      int lb = 0, rb = 0;
      dmrgcc_dropin::check(dmrgcc_model_mpo_site(m, s, &lb, &rb, nullptr));
      std::vector<double> W((size_t)lb * rb * d_ * d_);
      dmrgcc_dropin::check(dmrgcc_model_mpo_site(m, s, &lb, &rb, W.data()));
      RealTensorOperator* op = op_[s].get();
      op->DefineTensorOperator(lb, rb);
      for (int a = 0; a < lb; ++a)
        for (int b = 0; b < rb; ++b)
          for (int e = 0; e < d_ * d_; ++e)
            op->get_tensor_operator(a, b)->get_matrix_element()[e] = W[((size_t)a * rb + b) * d_ * d_ + e];
    }
    dmrgcc_model_destroy(m);
  }
  dmrgcc_model* CreateModel(const RealTensorSpace& space) const override {
    const auto* sp = dynamic_cast<const RealSpinlessBoseSquareSpace*>(&space);
    if (!sp) {
      std::printf("RealSpinlessBoseSquareHamiltonian needs a RealSpinlessBoseSquareSpace\n");
      std::exit(-1);
    }
    return make(sp->get_num_boson());
  }

 private:
  dmrgcc_model* make(int num_boson) const {
    dmrgcc_model* m = nullptr;
    dmrgcc_dropin::check(dmrgcc_model_spinless_create(nx_, ny_, num_boson, d_, jx_, jy_, U_, mu_, px_ ? 1 : 0,
                                                      py_ ? 1 : 0, &m));
    return m;
  }
  int nx_, ny_, d_;
  double flux_, jx_, jy_, U_, mu_;
  bool px_, py_;
};

class RealTensorRundmrg {
 public:
  // the run section of `input` (real_tensor_rundmrg.cc:3-42, Sweep :70-113)
  RealTensorRundmrg(RealTensorSpace* space, RealTensorHamiltonian* hamiltonian, InputGroup& input)
      : space_(space), hamiltonian_(hamiltonian), path_(input.file_name()) {
    using namespace dmrgcc_dropin;
    int nhalf = 0;
    check(dmrgcc_input_run_config(path_.c_str(), 1, &first_block_, &first_dim_, &nhalf, nullptr, nullptr, nullptr,
                                  nullptr, nullptr, nullptr));
    maxm_.resize(nhalf);
    maxb_.resize(nhalf);
    niter_.resize(nhalf);
    canpre_.resize(nhalf);
    noise_.resize(nhalf);
    int flags[4] = {0, 0, 0, 0};
    check(dmrgcc_input_run_config(path_.c_str(), 0, &first_block_, &first_dim_, &nhalf, maxm_.data(),
                                  maxb_.data(), niter_.data(), canpre_.data(), noise_.data(), flags));
    cache_record_ = flags[1] != 0;
    cache_resume_ = flags[2] != 0;
    cache_name_ = read_string("cache_name");
    record_name_ = read_string("record_name");
    resume_name_ = read_string("resume_name");
  }

  // Run (real_tensor_rundmrg.cc:115-245): the sweeps, printing the
  // reference's sweep headers, "right energy" / "left energy" after every
  // half sweep and the total time; afterwards the space's lattices hold the
  // final MPS.
  void Run() {
    using namespace dmrgcc_dropin;
    const auto t0 = std::chrono::steady_clock::now();
    const auto* src = dynamic_cast<const DeviceModelSource*>(hamiltonian_);
    if (!src) {
      std::printf("RealTensorRundmrg: the Hamiltonian has no device model\n");
      std::exit(-1);
    }
    dmrgcc_model* model = src->CreateModel(*space_);
    const int nhalf = (int)maxm_.size();
    const std::string record = cache_name_ + "/" + record_name_;
    dmrgcc_dmrg* run = nullptr;
    check(dmrgcc_dmrg_create_files(model, first_block_, first_dim_, nhalf, maxm_.data(), maxb_.data(),
                                   niter_.data(), canpre_.data(), noise_.data(), 0, 1, 0,
                                   cache_resume_ ? resume_name_.c_str() : nullptr,
                                   cache_record_ ? record.c_str() : nullptr, &run));
    const int L = space_->get_num_site();
    energies_.assign(nhalf, 0.0);
    for (int h = 0; h < nhalf; ++h) {
      if (h % 2 == 0) std::printf("------------------------ sweep=%d ------------------------\n", h / 2);
      int reached = 0;
      // run to the first Lanczos solve of the next half sweep (or the end)
      if (h + 1 < nhalf) check(dmrgcc_dmrg_advance(run, h + 1, (h + 1) % 2 == 0 ? 0 : L - 1, &reached));
      else check(dmrgcc_dmrg_advance(run, -1, -1, &reached));
      int done = 0;
      check(dmrgcc_dmrg_result(run, &done, energies_.data(), nullptr, nullptr, nullptr));
      std::printf("%s energy = %.12f\n", h % 2 == 0 ? "right" : "left", energies_[h]);
    }
    for (int s = 0; s < L; ++s) download_site(run, s);
    dmrgcc_dmrg_destroy(run);
    dmrgcc_model_destroy(model);
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("Total Time : %.12f s\n", sec);
  }

  // energy after every half sweep of the last Run (no counterpart in the
  // reference, which only prints them)
  const std::vector<double>& get_energies() const { return energies_; }

 private:
  std::string read_string(const char* key) {
    std::vector<char> buf(4096);
    int found = 0;
    dmrgcc_dropin::check(dmrgcc_input_string(path_.c_str(), key, buf.data(), (int)buf.size(), &found));
    return found ? std::string(buf.data()) : std::string();
  }
  // site s of the run into the space's lattice: sector lists (positions),
  // physics indices, ket blocks
  void download_site(dmrgcc_dmrg* run, int s) {
    using namespace dmrgcc_dropin;
    int nl = 0, nr = 0, count = 0;
    int64_t total = 0;
    check(dmrgcc_dmrg_phys(run, s, &nl, &nr, nullptr));
    std::vector<int> phys((size_t)nl * nr);
    check(dmrgcc_dmrg_phys(run, s, &nl, &nr, phys.data()));
    check(dmrgcc_dmrg_site_blocks(run, s, 0, &count, &total, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                  nullptr));
    std::vector<int> o(count), l(count), r(count), rows(count), cols(count);
    std::vector<int64_t> off(count);
    std::vector<double> data(total > 0 ? total : 1);
    check(dmrgcc_dmrg_site_blocks(run, s, 0, &count, &total, o.data(), l.data(), r.data(), rows.data(),
                                  cols.data(), off.data(), data.data()));
    std::vector<int> lb(nl), rb(nr), ld(nl, 0), rd(nr, 0);
    for (int i = 0; i < nl; ++i) lb[i] = i;
    for (int j = 0; j < nr; ++j) rb[j] = j;
    for (int b = 0; b < count; ++b) {
      ld[l[b]] = rows[b];
      rd[r[b]] = cols[b];
    }
    RealTensorLattice* lat = space_->get_tensor_lattice(s);
    lat->DefineTensorLattice(nl, nr, lb.data(), rb.data(), ld.data(), rd.data());
    std::vector<int> pi(count);
    for (int b = 0; b < count; ++b) pi[b] = phys[(size_t)l[b] * nr + r[b]];
    lat->DefineTensorLattice(count, l.data(), r.data(), pi.data(), 0, 0);
    for (int b = 0; b < count; ++b) {
      RealMatrix m(rows[b], cols[b]);
      for (int64_t e = 0; e < (int64_t)rows[b] * cols[b]; ++e) m.get_matrix_element()[e] = data[off[b] + e];
      lat->get_ket_tensor()->set_matrix_block(b, &m);
    }
  }

  RealTensorSpace* space_;
  RealTensorHamiltonian* hamiltonian_;
  std::string path_, cache_name_, record_name_, resume_name_;
  bool cache_record_ = false, cache_resume_ = false;
  int first_block_ = 0, first_dim_ = 0;
  std::vector<int> maxm_, maxb_, niter_;
  std::vector<double> canpre_, noise_;
  std::vector<double> energies_;
};

#endif  // DMRGCC_DROPIN_REAL_TENSOR_RUNDMRG_H_
