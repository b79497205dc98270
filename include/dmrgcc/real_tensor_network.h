// Drop-in RealTensorSpace, RealTensorHamiltonian, RealTensorNetwork and
// RealTensorLanczos (src/dmrg/realdmrg/real_tensor_{space,hamiltonian,
// network,lanczos}.h): the containers the reference's Lanczos reads its site
// from, and LanczosMethod on the MI355X (dmrgcc_lanczos).
#ifndef DMRGCC_DROPIN_REAL_TENSOR_NETWORK_H_
#define DMRGCC_DROPIN_REAL_TENSOR_NETWORK_H_

#include <memory>
#include <vector>

#include "real_tensor_contraction.h"
#include "real_tensor_lattice.h"
#include "real_tensor_operator.h"

// RealTensorSpace (real_tensor_space.h:6-140): the MPS, one lattice per site.
class RealTensorSpace {
 public:
  RealTensorSpace(int num_site, int physics_dim) : num_site_(num_site), physics_dim_(physics_dim) {
    for (int i = 0; i < num_site; ++i) lat_.emplace_back(new RealTensorLattice(physics_dim));
  }
  virtual ~RealTensorSpace() = default;
  int get_num_site() const { return num_site_; }
  int get_num_site_pp() const { return num_site_ + 1; }
  int get_num_site_mm() const { return num_site_ - 1; }
  int get_physics_dim() const { return physics_dim_; }
  RealTensorLattice* get_tensor_lattice(int site) { return lat_[site].get(); }

 protected:
  int num_site_, physics_dim_;
  std::vector<std::unique_ptr<RealTensorLattice>> lat_;
};

// RealTensorHamiltonian (real_tensor_hamiltonian.h:6-83): the MPO, one
// operator per site.
class RealTensorHamiltonian {
 public:
  RealTensorHamiltonian(int num_site, int physics_dim) : num_site_(num_site), physics_dim_(physics_dim) {
    for (int i = 0; i < num_site; ++i) op_.emplace_back(new RealTensorOperator(physics_dim));
  }
  virtual ~RealTensorHamiltonian() = default;
  int get_num_site() const { return num_site_; }
  RealTensorOperator* get_tensor_hamiltonian(int site) { return op_[site].get(); }
  virtual void DefineTensorHamiltonian() {}

 protected:
  int num_site_, physics_dim_;
  std::vector<std::unique_ptr<RealTensorOperator>> op_;
};

// RealTensorNetwork (real_tensor_network.h:8-60): tensor_contraction_[site]
// holds L (left bond of the site's operator) and R (its right bond).
class RealTensorNetwork {
 public:
  RealTensorNetwork(RealTensorSpace* space, RealTensorHamiltonian* hamiltonian)
      : space_(space), hamiltonian_(hamiltonian) {
    for (int i = 0; i < space->get_num_site(); ++i) {
      RealTensorOperator* op = hamiltonian->get_tensor_hamiltonian(i);
      con_.emplace_back(new RealTensorContraction(op->get_left_bond(), op->get_right_bond()));
    }
  }
  RealTensorContraction* get_tensor_contraction(int site) { return con_[site].get(); }

 private:
  RealTensorSpace* space_;
  RealTensorHamiltonian* hamiltonian_;
  std::vector<std::unique_ptr<RealTensorContraction>> con_;
};

// RealTensorLanczos (real_tensor_lanczos.h:6-31)
class RealTensorLanczos {
 public:
  RealTensorLanczos(RealTensorSpace* space, RealTensorHamiltonian* hamiltonian, RealTensorNetwork* network)
      : space_(space), hamiltonian_(hamiltonian), network_(network) {}
  // LanczosMethod (real_tensor_lanczos.cc:14-169): the ground state of the
  // site's effective Hamiltonian from the lattice's ket (written back into
  // it), its energy in result; psi_dim < 100 takes the dense solve, as in
  // the reference
  void LanczosMethod(int num_iter, int site, double& result) {
    result = network_->get_tensor_contraction(site)->SolveGroundState(
        space_->get_tensor_lattice(site), hamiltonian_->get_tensor_hamiltonian(site), num_iter);
  }

 private:
  RealTensorSpace* space_;
  RealTensorHamiltonian* hamiltonian_;
  RealTensorNetwork* network_;
};

#endif  // DMRGCC_DROPIN_REAL_TENSOR_NETWORK_H_
