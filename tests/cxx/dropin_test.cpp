// Test harness of the drop-in class API (include/dmrgcc/): one site built
// from given blocks and driven through the reference's member functions,
// and a whole run from a driver input file.  Compiled into a shared library
// that the GPU tests load with ctypes; every call ends up in
// libdmrgcc_amd.so.
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#include "../../include/dmrgcc/dmrgcc.h"

namespace {

struct Site {
  int d, lb, rb, nl, nr;
  std::vector<int> ldim, rdim, phys, kl, kr, krows, kcols;
  std::vector<double> W, ket;
  std::unique_ptr<RealTensorSpace> space;
  std::unique_ptr<RealTensorHamiltonian> ham;
  std::unique_ptr<RealTensorNetwork> net;
  std::unique_ptr<RealTensorContraction> next;  // receives environment updates

  // the site's lattice from the stored blocks
  void fill_lattice(RealTensorLattice* lat) const {
    std::vector<int> li(nl), ri(nr);
    std::iota(li.begin(), li.end(), 0);
    std::iota(ri.begin(), ri.end(), 0);
    lat->DefineTensorLattice(nl, nr, li.data(), ri.data(), ldim.data(), rdim.data());
    const int nb = (int)kl.size();
    std::vector<int> pi(nb);
    for (int b = 0; b < nb; ++b) pi[b] = phys[(size_t)kl[b] * nr + kr[b]];
    lat->DefineTensorLattice(nb, kl.data(), kr.data(), pi.data(), 0, 0);
    size_t off = 0;
    for (int b = 0; b < nb; ++b) {
      RealMatrix m(krows[b], kcols[b]);
      std::memcpy(m.get_matrix_element(), ket.data() + off, sizeof(double) * krows[b] * kcols[b]);
      off += (size_t)krows[b] * kcols[b];
      lat->get_ket_tensor()->set_matrix_block(b, &m);
    }
  }
};

void set_side(RealMatrixBlock** side, int nch, int n, const int* o, const int* l, const int* r, const int* rows,
              const int* cols, const int64_t* off, const double* data) {
  for (int ch = 0; ch < nch; ++ch) {
    std::vector<int> bl, br, idx;
    for (int i = 0; i < n; ++i)
      if (o[i] == ch) { bl.push_back(l[i]); br.push_back(r[i]); idx.push_back(i); }
    *side[ch] = RealMatrixBlock((int)bl.size(), bl.data(), br.data());
    for (size_t k = 0; k < idx.size(); ++k) {
      const int i = idx[k];
      RealMatrix m(rows[i], cols[i]);
      std::memcpy(m.get_matrix_element(), data + off[i], sizeof(double) * rows[i] * cols[i]);
      side[ch]->set_matrix_block((int)k, &m);
    }
  }
}

}  // namespace

extern "C" {

void* dropin_site_create(int d, int lb, int rb, const double* W, int nl, int nr, const int* ldim, const int* rdim,
                         const int* phys, int nket, const int* kl, const int* kr, const int* krows,
                         const int* kcols, const double* ket) {
  auto* s = new Site;
  s->d = d; s->lb = lb; s->rb = rb; s->nl = nl; s->nr = nr;
  s->ldim.assign(ldim, ldim + nl);
  s->rdim.assign(rdim, rdim + nr);
  s->phys.assign(phys, phys + (size_t)nl * nr);
  s->kl.assign(kl, kl + nket);
  s->kr.assign(kr, kr + nket);
  s->krows.assign(krows, krows + nket);
  s->kcols.assign(kcols, kcols + nket);
  size_t n = 0;
  for (int b = 0; b < nket; ++b) n += (size_t)krows[b] * kcols[b];
  s->ket.assign(ket, ket + n);
  s->W.assign(W, W + (size_t)lb * rb * d * d);
  s->space.reset(new RealTensorSpace(1, d));
  s->ham.reset(new RealTensorHamiltonian(1, d));
  RealTensorOperator* op = s->ham->get_tensor_hamiltonian(0);
  op->DefineTensorOperator(lb, rb);
  for (int a = 0; a < lb; ++a)
    for (int b = 0; b < rb; ++b)
      std::memcpy(op->get_tensor_operator(a, b)->get_matrix_element(), W + ((size_t)a * rb + b) * d * d,
                  sizeof(double) * d * d);
  s->fill_lattice(s->space->get_tensor_lattice(0));
  s->net.reset(new RealTensorNetwork(s->space.get(), s->ham.get()));
  s->next.reset(new RealTensorContraction(lb, rb));
  return s;
}

// leigh 0: L of the site (channels = lb), 1: R (channels = rb)
void dropin_site_env(void* h, int leigh, int n, const int* o, const int* l, const int* r, const int* rows,
                     const int* cols, const int64_t* off, const double* data) {
  auto* s = static_cast<Site*>(h);
  RealTensorContraction* c = s->net->get_tensor_contraction(0);
  if (leigh == 0) set_side(c->get_left_contraction_tensor(), s->lb, n, o, l, r, rows, cols, off, data);
  else set_side(c->get_right_contraction_tensor(), s->rb, n, o, l, r, rows, cols, off, data);
  c->MarkModified();
}

// state += H_eff ket  (MultiplyEffectHamilton), in the given arithmetic
void dropin_site_heff(void* h, double* state, int precision) {
  auto* s = static_cast<Site*>(h);
  RealTensorContraction* c = s->net->get_tensor_contraction(0);
  c->set_precision(precision);
  c->MultiplyEffectHamilton(s->space->get_tensor_lattice(0), s->ham->get_tensor_hamiltonian(0), state);
  c->set_precision(0);
}

// LanczosMethod from the stored ket; the Ritz vector into ket_out
double dropin_site_lanczos(void* h, int niter, double* ket_out) {
  auto* s = static_cast<Site*>(h);
  RealTensorLanczos lz(s->space.get(), s->ham.get(), s->net.get());
  RealTensorLattice* lat = s->space->get_tensor_lattice(0);
  s->fill_lattice(lat);
  double e = 0.0;
  lz.LanczosMethod(niter, 0, e);
  lat->VectorizeTensorLattice(true, ket_out);
  s->fill_lattice(lat);  // restore the stored ket
  return e;
}

// Left (leigh 0) / RightComputeTensorContraction into a fresh contraction;
// returns the block count, blocks (channel, l, r, rows, cols, offset) and
// data written up to the capacities
int dropin_site_env_update(void* h, int leigh, int cap, int* o, int* l, int* r, int* rows, int* cols,
                           int64_t* off, double* data, int64_t dcap) {
  auto* s = static_cast<Site*>(h);
  RealTensorContraction* c = s->net->get_tensor_contraction(0);
  RealTensorLattice* lat = s->space->get_tensor_lattice(0);
  RealTensorOperator* op = s->ham->get_tensor_hamiltonian(0);
  if (leigh == 0) c->LeftComputeTensorContraction(lat, op, s->next.get());
  else c->RightComputeTensorContraction(lat, op, s->next.get());
  RealMatrixBlock** side = leigh == 0 ? s->next->get_left_contraction_tensor()
                                      : s->next->get_right_contraction_tensor();
  const int nch = leigh == 0 ? s->next->get_left_bond() : s->next->get_right_bond();
  int k = 0;
  int64_t acc = 0;
  for (int ch = 0; ch < nch; ++ch)
    for (int i = 0; i < side[ch]->get_num_block(); ++i, ++k) {
      const RealMatrix* m = side[ch]->get_matrix_block(i);
      const int64_t ne = (int64_t)m->get_row() * m->get_column();
      if (k < cap) {
        o[k] = ch; l[k] = side[ch]->get_left_block()[i]; r[k] = side[ch]->get_right_block()[i];
        rows[k] = m->get_row(); cols[k] = m->get_column(); off[k] = acc;
        if (acc + ne <= dcap) std::memcpy(data + acc, m->get_matrix_element(), sizeof(double) * ne);
      }
      acc += ne;
    }
  return k;
}

// Left (leigh 0) / RightCanonicalTensorLattice on a copy of the site:
// the new ket blocks (in block order), sector dims, match dims and the
// canonical tensor's blocks (in sector order), up to the capacities
int dropin_site_canonical(void* h, int leigh, int max_dim, double canpre, double* ket_out, int64_t kcap,
                          int* rows_out, int* cols_out, int* dims_out, int* match_out, double* can_out,
                          int64_t ccap, double* discarded) {
  auto* s = static_cast<Site*>(h);
  RealTensorLattice lat(s->d);
  s->fill_lattice(&lat);
  if (leigh == 0) lat.LeftCanonicalTensorLattice(max_dim, canpre);
  else lat.RightCanonicalTensorLattice(max_dim, canpre);
  RealMatrixBlock* k = lat.get_ket_tensor();
  int64_t acc = 0;
  for (int b = 0; b < k->get_num_block(); ++b) {
    const RealMatrix* m = k->get_matrix_block(b);
    rows_out[b] = m->get_row();
    cols_out[b] = m->get_column();
    const int64_t ne = (int64_t)m->get_row() * m->get_column();
    if (acc + ne <= kcap) std::memcpy(ket_out + acc, m->get_matrix_element(), sizeof(double) * ne);
    acc += ne;
  }
  const int ns = leigh == 0 ? lat.get_num_right_block() : lat.get_num_left_block();
  const int* dims = leigh == 0 ? lat.get_right_dim() : lat.get_left_dim();
  acc = 0;
  for (int i = 0; i < ns; ++i) {
    dims_out[i] = dims[i];
    match_out[i] = lat.get_match_dim()[i];
    const RealMatrix* m = lat.get_canonical_tensor()->get_matrix_block(i);
    const int64_t ne = (int64_t)m->get_row() * m->get_column();
    if (acc + ne <= ccap) std::memcpy(can_out + acc, m->get_matrix_element(), sizeof(double) * ne);
    acc += ne;
  }
  *discarded = lat.get_discarded_weight();
  return ns;
}

void dropin_site_destroy(void* h) { delete static_cast<Site*>(h); }

// The reference's real spinless driver (model/real_spinless_bose_square_model.cc)
// on `path`, written against the drop-in classes; energies of every half
// sweep and the final bond dimension after every site.
int dropin_run(const char* path, double* energies, int cap, int* bond_dims) {
  InputGroup input(path, "input");
  int num_site_x = input.getInt("num_site_x");
  int num_site_y = input.getInt("num_site_y");
  int num_boson = input.getInt("num_boson");
  int physics_dim = input.getInt("physics_dim", 2);
  double flux = input.getDouble("flux");
  double flux_value = flux * PI;
  double Jx = input.getDouble("Jx");
  double Jy = input.getDouble("Jy");
  double U = input.getDouble("U");
  double mu = input.getDouble("mu");
  bool period_x = input.getBool("period_x");
  bool period_y = input.getBool("period_y");
  RealTensorSpace* space = new RealSpinlessBoseSquareSpace(num_site_x, num_site_y, num_boson, physics_dim);
  RealTensorHamiltonian* hamiltonian = new RealSpinlessBoseSquareHamiltonian(
      num_site_x, num_site_y, physics_dim, flux_value, Jx, Jy, U, mu, period_x, period_y);
  RealTensorRundmrg* rundmrg = new RealTensorRundmrg(space, hamiltonian, input);
  rundmrg->Run();
  const std::vector<double>& e = rundmrg->get_energies();
  int n = (int)e.size();
  for (int i = 0; i < n && i < cap; ++i) energies[i] = e[i];
  for (int s = 0; s < space->get_num_site(); ++s) {
    RealTensorLattice* lat = space->get_tensor_lattice(s);
    int dim = 0;
    std::vector<int> seen(lat->get_num_right_block(), 0);
    RealMatrixBlock* k = lat->get_ket_tensor();
    for (int b = 0; b < k->get_num_block(); ++b) {
      const int r = k->get_right_block()[b];
      if (!seen[r]) { seen[r] = 1; dim += k->get_matrix_block(b)->get_column(); }
    }
    bond_dims[s] = dim;
  }
  delete rundmrg;
  delete hamiltonian;
  delete space;
  return n;
}

}  // extern "C"
