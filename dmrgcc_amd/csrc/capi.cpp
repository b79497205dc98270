// extern "C" entry points (declared in include/dmrgcc_hip.h).  Each wraps one
// piece of the engine, converting exceptions into status codes.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>

#include "../../include/dmrgcc_hip.h"
#include "dmrg.h"
#include "input.h"
#include "quicksort.h"
#include "rng.h"

using namespace dmrg;

struct dmrgcc_model {
  std::unique_ptr<RealTensorSpace> space;
  std::unique_ptr<RealTensorHamiltonian> ham;
  std::vector<QNTable> tables;
};

struct dmrgcc_heff {
  std::unique_ptr<Context> cx;
  Lattice lat;
  SiteOperator W;
  Env L, R;
  HeffPlan plan;
  LanczosStats st;
  TwoSite two;  // the two-site state of a dmrgcc_dmrg_heff2 plan
  // A plan built from a run (dmrgcc_dmrg_heff / _heff2) holds device
  // pointers into that run's environments and site tensors.  It is valid
  // while the run's epoch equals the value recorded here; advancing or
  // destroying the run bumps the epoch (the buffers go back to the pool).
  std::shared_ptr<const int64_t> run_epoch;
  int64_t epoch = 0;
};

struct dmrgcc_dmrg {
  dmrgcc_model* model;
  std::unique_ptr<RealTensorRundmrg> run;
  std::shared_ptr<int64_t> epoch = std::make_shared<int64_t>(0);
  ~dmrgcc_dmrg() { ++*epoch; }
};

struct dmrgcc_heff_batch {
  std::unique_ptr<Context> cx;
  HeffBatch batch;
  std::vector<dmrgcc_heff*> plans;
};

struct dmrgcc_blocks {
  std::vector<int> o, l, r, rows, cols;
  std::vector<int64_t> off;
  int cs = 1;  // 2: complex entries, interleaved
  DevBuf<double> data;
};

namespace {

void require_live(const dmrgcc_heff* h) {
  require(!h->run_epoch || *h->run_epoch == h->epoch,
          "stale H_eff plan: the run it was built from has advanced or been destroyed", kErrArgument);
}

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return kErrInternal;
  }
}

// cs = 2: W holds interleaved complex entries
SiteOperator make_op(int d, int lb, int rb, const double* W, int cs = 1) {
  require(W != nullptr && d > 0 && lb > 0 && rb > 0, "invalid MPO site tensor", kErrArgument);
  SiteOperator op;
  op.d = d; op.lb = lb; op.rb = rb;
  const size_t ne = (size_t)lb * rb * d * d;
  if (cs == 1) {
    op.w.assign(W, W + ne);
  } else {
    op.w.resize(ne);
    op.wi.resize(ne);
    for (size_t e = 0; e < ne; ++e) { op.w[e] = W[2 * e]; op.wi[e] = W[2 * e + 1]; }
  }
  return op;
}

void make_lattice(Lattice& lat, int d, int nl, int nr, const int* phys, int nket, const int* kl,
                  const int* kr, const int* krows, const int* kcols, const double* data,
                  int cs = 1) {
  lat.ket.cs = cs;
  require(nl >= 0 && nr >= 0 && nket >= 0, "negative sizes", kErrArgument);
  lat.d = d;
  lat.lb.resize(nl);
  lat.rb.resize(nr);
  std::iota(lat.lb.begin(), lat.lb.end(), 0);
  std::iota(lat.rb.begin(), lat.rb.end(), 0);
  lat.ld.assign(nl, 0);
  lat.rd.assign(nr, 0);
  lat.phys.assign(nl, std::vector<int>(nr, -1));
  for (int i = 0; i < nl; ++i)
    for (int j = 0; j < nr; ++j) lat.phys[i][j] = phys[(size_t)i * nr + j];
  int64_t acc = 0;
  for (int b = 0; b < nket; ++b) {
    require(kl[b] >= 0 && kl[b] < nl && kr[b] >= 0 && kr[b] < nr, "ket block sector out of range");
    require(lat.phys[kl[b]][kr[b]] != -1, "ket block without a physical index");
    lat.ket.blk.push_back(Block{kl[b], kr[b], krows[b], kcols[b], acc});
    acc += (int64_t)krows[b] * kcols[b];
    lat.ld[kl[b]] = krows[b];
    lat.rd[kr[b]] = kcols[b];
  }
  lat.ket.size = acc;
  lat.ket.ext = data;
}

Env make_env(int nbond, int n, const int* o, const int* l, const int* r, const int* rows,
             const int* cols, const int64_t* off, const double* data, int cs = 1) {
  Env e(nbond);
  for (int i = 0; i < n; ++i) {
    require(o[i] >= 0 && o[i] < nbond, "environment channel out of range");
    e[o[i]].blk.push_back(Block{l[i], r[i], rows[i], cols[i], off[i]});
  }
  for (auto& b : e) { b.ext = data; b.cs = cs; }
  return e;
}

int heff_create(int cs, int d, int lb, int rb, const double* W, int nl, int nr, const int* phys,
                int nket, const int* ket_l, const int* ket_r, const int* ket_rows,
                const int* ket_cols, int nL, const int* L_o, const int* L_l, const int* L_r,
                const int* L_rows, const int* L_cols, const int64_t* L_off, const double* L_data,
                int nR, const int* R_o, const int* R_l, const int* R_r, const int* R_rows,
                const int* R_cols, const int64_t* R_off, const double* R_data, void* stream,
                dmrgcc_heff** out) {
  return guarded([&] {
    require(out != nullptr, "null output", kErrArgument);
    auto h = std::make_unique<dmrgcc_heff>();
    h->cx = std::make_unique<Context>(static_cast<hipStream_t>(stream));
    h->W = make_op(d, lb, rb, W, cs);
    make_lattice(h->lat, d, nl, nr, phys, nket, ket_l, ket_r, ket_rows, ket_cols, nullptr, cs);
    h->L = make_env(lb, nL, L_o, L_l, L_r, L_rows, L_cols, L_off, L_data, cs);
    h->R = make_env(rb, nR, R_o, R_l, R_r, R_rows, R_cols, R_off, R_data, cs);
    h->plan.build(h->lat, h->W, h->L, h->R, h->cx->stream);
    h->cx->sync();  // descriptors in place: the plan may run on any stream
    *out = h.release();
  });
}

int env_update(int cs, int leigh, int d, int lb, int rb, const double* W, int nl, int nr,
               const int* phys, int nket, const int* ket_l, const int* ket_r, const int* ket_rows,
               const int* ket_cols, const double* ket_data, int nE, const int* E_o, const int* E_l,
               const int* E_r, const int* E_rows, const int* E_cols, const int64_t* E_off,
               const double* E_data, void* stream, dmrgcc_blocks** out) {
  return guarded([&] {
    require(out && (leigh == 0 || leigh == 1), "bad arguments", kErrArgument);
    Context cx(static_cast<hipStream_t>(stream));
    SiteOperator op = make_op(d, lb, rb, W, cs);
    Lattice lat;
    make_lattice(lat, d, nl, nr, phys, nket, ket_l, ket_r, ket_rows, ket_cols, ket_data, cs);
    Env e = make_env(leigh == 0 ? lb : rb, nE, E_o, E_l, E_r, E_rows, E_cols, E_off, E_data, cs);
    Env res = leigh == 0 ? left_env_update(lat, op, e, cx) : right_env_update(lat, op, e, cx);
    auto b = std::make_unique<dmrgcc_blocks>();
    b->cs = cs;
    int64_t total = 0;
    for (const auto& t : res) total += t.size;
    b->data.resize((size_t)std::max<int64_t>(cs * total, 1));
    int64_t acc = 0;
    for (size_t o = 0; o < res.size(); ++o) {
      for (const Block& k : res[o].blk) {
        b->o.push_back((int)o);
        b->l.push_back(k.l);
        b->r.push_back(k.r);
        b->rows.push_back(k.rows);
        b->cols.push_back(k.cols);
        b->off.push_back(acc + k.off);
      }
      if (res[o].size > 0)
        DMRG_HIP(hipMemcpyAsync(b->data.data() + cs * acc, res[o].data.data(),
                                cs * res[o].size * sizeof(double), hipMemcpyDeviceToDevice, cx.stream));
      acc += res[o].size;
    }
    cx.sync();
    *out = b.release();
  });
}

}  // namespace

namespace {
// The items of RealMatrix::MultiplyToMatrix calls: C_i = A_i B_i, row-major
void block_gemm_items(GemmBatch& gb, int n, const double* const* A, const double* const* B,
                      double* const* C, const int* M, const int* N, const int* K, int precision) {
  require(n >= 0 && (n == 0 || (A && B && C && M && N && K)), "bad arguments", kErrArgument);
  require(precision >= 0 && precision <= 2, "unknown precision", kErrArgument);
  for (int i = 0; i < n; ++i) {
    require(M[i] >= 0 && N[i] >= 0 && K[i] >= 0, "negative GEMM size");
    gb.begin_item(C[i], N[i], M[i], N[i], 0);
    gb.add(A[i], K[i], B[i], N[i], K[i], 0, 1.0);
  }
}

RunParams make_params(int first_block, int first_dim, int nhalf, const int* maxm, const int* maxb,
                      const int* niter, const double* canpre, const double* noise, int precision,
                      int refine, int verbose) {
  require(nhalf > 0 && maxm && maxb && niter && canpre && noise, "bad sweep table", kErrArgument);
  require(precision >= 0 && precision <= 2, "unknown precision", kErrArgument);
  RunParams p;
  p.first_block = first_block;
  p.first_dim = first_dim;
  for (int h = 0; h < nhalf; ++h) {
    p.maxm.push_back(maxm[h]);
    p.maxb.push_back(maxb[h]);
    p.niter.push_back(niter[h]);
    p.canpre.push_back(canpre[h]);
    p.noise.push_back(std::pow(0.1, noise[h]));
  }
  p.precision = precision;
  p.refine = refine;
  p.verbose = verbose;
  return p;
}

void fill_result(const RunResult& r, int* nhalf_done, double* energies, double* trunc, double* seconds,
                 double* stats) {
  if (nhalf_done) *nhalf_done = (int)r.energies.size();
  for (size_t h = 0; h < r.energies.size(); ++h) {
    if (energies) energies[h] = r.energies[h];
    if (trunc) trunc[h] = r.trunc_err[h];
    if (seconds) seconds[h] = r.half_seconds[h];
  }
  if (stats) {
    stats[0] = r.total_seconds;
    stats[1] = r.stats.matvecs;
    stats[2] = r.stats.dense;
    stats[3] = r.stats.matvec_seconds;
    stats[4] = r.stats.flops;
    stats[5] = (double)r.rng_draws;
    stats[6] = r.stats.solves;
    int mx = 0;
    for (int v : r.final_bond_dims) mx = std::max(mx, v);
    stats[7] = mx;
    const PhaseTimes& ph = r.phase;
    stats[8] = ph.lanczos; stats[9] = ph.expan; stats[10] = ph.expand; stats[11] = ph.canonical;
    stats[12] = ph.svd; stats[13] = ph.env; stats[14] = ph.merge;
    stats[15] = (double)ph.svd_max_m; stats[16] = (double)ph.svd_max_n;
    stats[17] = (double)ph.svd_calls; stats[18] = ph.svd_work;
    stats[19] = ph.plan;
    stats[20] = r.stats.tunes;
    stats[21] = r.stats.tile_mode;
    stats[22] = r.stats.split_k;
  }
}
}  // namespace

extern "C" {

const char* dmrgcc_last_error(void) { return last_error(); }
int dmrgcc_version(void) { return 1; }

int dmrgcc_rand_draws(unsigned seed, int64_t skip, int n, int* out) {
  return guarded([&] {
    require(out != nullptr && n >= 0 && skip >= 0, "bad arguments", kErrArgument);
    GlibcRand g(seed);
    g.skip(skip);
    for (int i = 0; i < n; ++i) out[i] = g.next();
  });
}

int dmrgcc_random_matrices(unsigned seed, int64_t skip, int n, const int* rows, const int* cols,
                           int is_complex, const double* w, double* out) {
  return guarded([&] {
    require(skip >= 0, "bad arguments", kErrArgument);
    GlibcRand g(seed);
    g.skip(skip);
    random_matrices(g, n, rows, cols, is_complex ? 2 : 1, w, out);
  });
}

int dmrgcc_quicksort(int n, double* vals, int* index, int flag) {
  return guarded([&] {
    require(vals != nullptr && n >= 0 && (flag == 0 || flag == 1), "bad arguments", kErrArgument);
    std::vector<double> v(vals, vals + n);
    std::vector<int> idx;
    if (index) idx.assign(index, index + n);
    quicksort(v, index ? &idx : nullptr, flag);
    std::memcpy(vals, v.data(), n * sizeof(double));
    if (index) std::memcpy(index, idx.data(), n * sizeof(int));
  });
}

int dmrgcc_sym_eig(int n, double* a, double* w) {
  return guarded([&] {
    require(n >= 1 && a && w, "bad arguments", kErrArgument);
    std::vector<double> ev, vec;
    sym_eig(n, std::vector<double>(a, a + (size_t)n * n), ev, vec);
    std::copy(ev.begin(), ev.end(), w);
    std::copy(vec.begin(), vec.end(), a);
  });
}

int dmrgcc_sym_eig_lowest(int n, const double* a, double* e0, double* v0) {
  return guarded([&] {
    require(n >= 1 && a && e0 && v0, "bad arguments", kErrArgument);
    std::vector<double> v;
    sym_eig_lowest(n, std::vector<double>(a, a + (size_t)n * n), *e0, v);
    std::copy(v.begin(), v.end(), v0);
  });
}

int dmrgcc_truncate_dims(int nsec, const int* counts, double* svals, int max_dim, double canpre,
                         int* trunc) {
  return guarded([&] {
    require(nsec >= 0 && counts && svals && trunc, "bad arguments", kErrArgument);
    std::vector<std::vector<double>> sv(nsec);
    int64_t acc = 0;
    for (int s = 0; s < nsec; ++s) {
      require(counts[s] >= 1, "every sector needs at least one singular value", kErrArgument);
      sv[s].assign(svals + acc, svals + acc + counts[s]);
      acc += counts[s];
    }
    std::vector<int> t = truncate_dims(max_dim, canpre, sv);
    acc = 0;
    for (int s = 0; s < nsec; ++s) {
      trunc[s] = t[s];
      svals[acc] = sv[s][0];
      acc += counts[s];
    }
  });
}

int dmrgcc_model_spinless_create(int nx, int ny, int num_boson, int physics_dim, double hop_x,
                                 double hop_y, double U, double mu, int period_x, int period_y,
                                 dmrgcc_model** out) {
  return guarded([&] {
    require(out && nx > 0 && ny > 0 && physics_dim > 1 && num_boson >= 0, "bad model arguments",
            kErrArgument);
    auto m = std::make_unique<dmrgcc_model>();
    m->space = std::make_unique<RealSpinlessBoseSquareSpace>(nx, ny, num_boson, physics_dim);
    m->ham = std::make_unique<RealSpinlessBoseSquareHamiltonian>(nx, ny, physics_dim, 0.0, hop_x, hop_y,
                                                                 U, mu, period_x != 0, period_y != 0);
    m->ham->DefineTensorHamiltonian();
    for (const SiteOperator& w : m->ham->sites())
      require(w.lb > 0 && w.rb > 0, "the Hamiltonian has no term on some site of this lattice", kErrArgument);
    m->tables = m->space->DefineQuantumTable();
    *out = m.release();
  });
}

int dmrgcc_model_zspinless_create(int nx, int ny, int num_boson, int physics_dim, double hop_x,
                                  double hop_y, double U, double mu, int period_x, int period_y,
                                  dmrgcc_model** out) {
  return guarded([&] {
    require(out && nx > 0 && ny > 0 && physics_dim > 1 && num_boson >= 0, "bad model arguments",
            kErrArgument);
    auto m = std::make_unique<dmrgcc_model>();
    m->space = std::make_unique<ComplexSpinlessBoseSquareSpace>(nx, ny, num_boson, physics_dim);
    m->ham = std::make_unique<ComplexSpinlessBoseSquareHamiltonian>(nx, ny, physics_dim, 0.0, hop_x, hop_y,
                                                                    U, mu, period_x != 0, period_y != 0);
    m->ham->DefineTensorHamiltonian();
    for (const SiteOperator& w : m->ham->sites())
      require(w.lb > 0 && w.rb > 0, "the Hamiltonian has no term on some site of this lattice", kErrArgument);
    m->tables = m->space->DefineQuantumTable();
    *out = m.release();
  });
}

int dmrgcc_model_spinful_create(int nx, int ny, int n_up, int n_down, int physics_dim,
                                const double* flux, const double* hop_x, const double* hop_y,
                                const double* inter, const double* mu, int period_x, int period_y,
                                dmrgcc_model** out) {
  return guarded([&] {
    require(out && flux && hop_x && hop_y && inter && mu, "bad model arguments", kErrArgument);
    require(nx > 0 && ny > 0 && physics_dim > 1 && n_up >= 0 && n_down >= 0, "bad model arguments",
            kErrArgument);
    auto m = std::make_unique<dmrgcc_model>();
    m->space = std::make_unique<ComplexSpinfulBoseSquareSpace>(nx, ny, n_up, n_down, physics_dim);
    m->ham = std::make_unique<ComplexSpinfulBoseSquareHamiltonian>(nx, ny, physics_dim, flux, hop_x, hop_y,
                                                                   inter, mu, period_x != 0, period_y != 0);
    m->ham->DefineTensorHamiltonian();
    for (const SiteOperator& w : m->ham->sites())
      require(w.lb > 0 && w.rb > 0, "the Hamiltonian has no term on some site of this lattice", kErrArgument);
    m->tables = m->space->DefineQuantumTable();
    *out = m.release();
  });
}

int dmrgcc_model_info(const dmrgcc_model* m, int* num_site, int* physics_dim, int* num_quantum,
                      int* is_complex) {
  return guarded([&] {
    require(m, "bad arguments", kErrArgument);
    if (num_site) *num_site = m->space->num_site();
    if (physics_dim) *physics_dim = m->space->physics_dim();
    if (num_quantum) *num_quantum = m->space->num_quantum();
    if (is_complex) *is_complex = m->ham->complex() ? 1 : 0;
  });
}

int dmrgcc_input_model_create(const char* path, const char* driver, int echo, dmrgcc_model** out) {
  return guarded([&] {
    require(path && driver && out, "bad arguments", kErrArgument);
    InputGroup::echo = echo != 0;
    InputGroup input(path, "input");
    auto m = std::make_unique<dmrgcc_model>();
    read_model(input, driver, m->space, m->ham);
    m->ham->DefineTensorHamiltonian();
    for (const SiteOperator& w : m->ham->sites())
      require(w.lb > 0 && w.rb > 0, "the Hamiltonian has no term on some site of this lattice", kErrArgument);
    m->tables = m->space->DefineQuantumTable();
    *out = m.release();
    std::fflush(stdout);
  });
}

int dmrgcc_input_string(const char* path, const char* key, char* buf, int buflen, int* found) {
  return guarded([&] {
    require(path && key && buf && buflen > 0, "bad arguments", kErrArgument);
    const bool echo = InputGroup::echo;
    InputGroup::echo = false;
    InputGroup input(path, "input");
    std::string v;
    const int got = input.GetString(key, v);
    InputGroup::echo = echo;
    if (found) *found = got;
    require((int)v.size() < buflen, "buffer too small for input value", kErrArgument);
    std::memcpy(buf, v.c_str(), v.size() + 1);
  });
}

int dmrgcc_input_get(const char* path, const char* group, const char* key, int kind, int echo,
                     int has_default, double default_value, double* value, int* found) {
  return guarded([&] {
    require(path && group && key && value && found && kind >= 0 && kind <= 2, "bad arguments", kErrArgument);
    const bool saved = InputGroup::echo;
    InputGroup::echo = echo != 0;
    InputGroup g(path, group);
    int got = 0;
    if (kind == 0) {
      int v = 0;
      got = g.GetInt(key, v, has_default != 0, (int)default_value);
      *value = got ? v : default_value;
    } else if (kind == 1) {
      double v = 0.0;
      got = g.GetDouble(key, v, has_default != 0, default_value);
      *value = got ? v : default_value;
    } else {
      bool v = false;
      got = g.GetBool(key, v, has_default != 0, default_value != 0.0);
      *value = got ? (v ? 1.0 : 0.0) : default_value;
    }
    InputGroup::echo = saved;
    *found = got;
  });
}

int dmrgcc_input_run_config(const char* path, int echo, int* first_block, int* first_dim, int* nhalf,
                            int* maxm, int* maxb, int* niter, double* canpre, double* noise, int* flags) {
  return guarded([&] {
    require(path && nhalf, "bad arguments", kErrArgument);
    InputGroup::echo = echo != 0;
    InputGroup input(path, "input");
    const RunConfig c = read_run_config(input);
    *nhalf = (int)c.maxm.size();
    if (first_block) *first_block = c.first_block;
    if (first_dim) *first_dim = c.first_dim;
    for (size_t i = 0; i < c.maxm.size(); ++i) {
      if (maxm) maxm[i] = c.maxm[i];
      if (maxb) maxb[i] = c.maxb[i];
      if (niter) niter[i] = c.niter[i];
      if (canpre) canpre[i] = c.canpre[i];
      if (noise) noise[i] = c.noise[i];
    }
    if (flags) {
      flags[0] = c.disk_cache;
      flags[1] = c.cache_record;
      flags[2] = c.cache_resume;
      flags[3] = c.record_process;
    }
    std::fflush(stdout);
  });
}

int dmrgcc_model_destroy(dmrgcc_model* m) {
  delete m;
  return 0;
}

int dmrgcc_model_num_site(const dmrgcc_model* m) { return m ? m->space->num_site() : -1; }

int dmrgcc_model_space_table(const dmrgcc_model* m, int bond, int* count, int* qn) {
  return guarded([&] {
    require(m && count && bond >= 0 && bond < (int)m->tables.size(), "bad arguments", kErrArgument);
    *count = (int)m->tables[bond].size();
    const int nq = m->space->num_quantum();
    if (qn)
      for (int i = 0; i < *count; ++i)
        for (int q = 0; q < nq; ++q) qn[(size_t)i * nq + q] = m->tables[bond][i][q];
  });
}

int dmrgcc_model_mpo_site(const dmrgcc_model* m, int site, int* lb, int* rb, double* W) {
  return guarded([&] {
    require(m && lb && rb && site >= 0 && site < m->ham->num_site(), "bad arguments", kErrArgument);
    const SiteOperator& op = m->ham->sites()[site];
    *lb = op.lb;
    *rb = op.rb;
    if (W && !op.complex()) std::memcpy(W, op.w.data(), op.w.size() * sizeof(double));
    if (W && op.complex())
      for (size_t i = 0; i < op.w.size(); ++i) {
        W[2 * i] = op.w[i];
        W[2 * i + 1] = op.wi[i];
      }
  });
}

int dmrgcc_model_mpo_table(const dmrgcc_model* m, int bond, int* count, int* qn) {
  return guarded([&] {
    require(m && count, "bad arguments", kErrArgument);
    const auto& t = m->ham->bond_tables();
    require(bond >= 0 && bond < (int)t.size(), "bond out of range", kErrArgument);
    *count = (int)t[bond].size();
    const int nq = m->ham->num_quantum();
    if (qn)
      for (int i = 0; i < *count; ++i)
        for (int q = 0; q < nq; ++q) qn[(size_t)i * nq + q] = t[bond][i][q];
  });
}

int dmrgcc_block_gemm(int n, const double* const* A, const double* const* B, double* const* C,
                      const int* M, const int* N, const int* K, int precision, void* stream) {
  return guarded([&] {
    hipStream_t s = static_cast<hipStream_t>(stream);
    GemmBatch gb;
    block_gemm_items(gb, n, A, B, C, M, N, K, precision);
    gb.finalize(s);
    gb.launch(s, precision);
    DMRG_HIP(hipStreamSynchronize(s));
  });
}

int dmrgcc_block_gemm_noalloc(int n, const double* const* A, const double* const* B, double* const* C,
                              const int* M, const int* N, const int* K, int precision, void* stream) {
  return guarded([&] {
    hipStream_t s = static_cast<hipStream_t>(stream);
    GemmBatch gb;
    block_gemm_items(gb, n, A, B, C, M, N, K, precision);
    // the same tiling as dmrgcc_block_gemm, placed in the fixed arena
    const int64_t psize = gb.layout();
    size_t cap = 0;
    void* arena = noalloc_arena(&cap);
    gb.place(arena, cap, psize, s);
    gb.launch(s, precision);
    DMRG_HIP(hipStreamSynchronize(s));
  });
}

int dmrgcc_zblock_gemm(int n, const double* const* A, const double* const* B, double* const* C,
                       const int* M, const int* N, const int* K, int precision, void* stream) {
  return guarded([&] {
    require(n >= 0 && (n == 0 || (A && B && C && M && N && K)), "bad arguments", kErrArgument);
    require(precision >= 0 && precision <= 2, "unknown precision", kErrArgument);
    hipStream_t s = static_cast<hipStream_t>(stream);
    GemmBatch gb;
    for (int i = 0; i < n; ++i) {
      require(M[i] >= 0 && N[i] >= 0 && K[i] >= 0, "negative GEMM size");
      zgemm_item(gb, C[i], N[i], M[i], N[i], 0, 0,
                 {ZTerm{A[i], K[i], B[i], N[i], K[i], 0, false, false, 1.0, 0.0}});
    }
    gb.finalize(s);
    gb.launch(s, precision);
    DMRG_HIP(hipStreamSynchronize(s));
  });
}

int dmrgcc_heff_create(int d, int lb, int rb, const double* W, int nl, int nr, const int* phys,
                       int nket, const int* ket_l, const int* ket_r, const int* ket_rows,
                       const int* ket_cols, int nL, const int* L_o, const int* L_l,
                       const int* L_r, const int* L_rows, const int* L_cols,
                       const int64_t* L_off, const double* L_data, int nR, const int* R_o,
                       const int* R_l, const int* R_r, const int* R_rows, const int* R_cols,
                       const int64_t* R_off, const double* R_data, void* stream,
                       dmrgcc_heff** out) {
  return heff_create(1, d, lb, rb, W, nl, nr, phys, nket, ket_l, ket_r, ket_rows, ket_cols, nL, L_o,
                     L_l, L_r, L_rows, L_cols, L_off, L_data, nR, R_o, R_l, R_r, R_rows, R_cols,
                     R_off, R_data, stream, out);
}

int dmrgcc_zheff_create(int d, int lb, int rb, const double* W, int nl, int nr, const int* phys,
                        int nket, const int* ket_l, const int* ket_r, const int* ket_rows,
                        const int* ket_cols, int nL, const int* L_o, const int* L_l,
                        const int* L_r, const int* L_rows, const int* L_cols,
                        const int64_t* L_off, const double* L_data, int nR, const int* R_o,
                        const int* R_l, const int* R_r, const int* R_rows, const int* R_cols,
                        const int64_t* R_off, const double* R_data, void* stream,
                        dmrgcc_heff** out) {
  return heff_create(2, d, lb, rb, W, nl, nr, phys, nket, ket_l, ket_r, ket_rows, ket_cols, nL, L_o,
                     L_l, L_r, L_rows, L_cols, L_off, L_data, nR, R_o, R_l, R_r, R_rows, R_cols,
                     R_off, R_data, stream, out);
}

int dmrgcc_heff_apply(dmrgcc_heff* h, const double* x, double* y, int precision) {
  return guarded([&] {
    require(h && x && y, "null argument", kErrArgument);
    require(precision >= 0 && precision <= 2, "unknown precision", kErrArgument);
    require_live(h);
    h->plan.apply(x, y, precision, h->cx->stream);
    h->cx->sync();
  });
}

int dmrgcc_heff_apply_async(dmrgcc_heff* h, const double* x, double* y, double alpha, double beta,
                            int precision, void* stream) {
  return guarded([&] {
    require(h && x && y, "null argument", kErrArgument);
    require(precision >= 0 && precision <= 2, "unknown precision", kErrArgument);
    require(x != y, "x and y must not overlap", kErrArgument);
    require_live(h);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : h->cx->stream;
    const int64_t n = h->plan.doubles();
    // y = alpha H x + beta y as y <- (beta / alpha) y, y += H x, y <- alpha y:
    // kernels only (no memset, no host wait), so the call can be captured
    if (alpha == 0.0) {
      if (beta == 0.0) fill(y, n, 0.0, s);
      else if (beta != 1.0) axpby(y, y, n, beta, 0.0, s);
      return;
    }
    if (beta == 0.0) fill(y, n, 0.0, s);
    else if (beta != alpha) axpby(y, y, n, beta / alpha, 0.0, s);
    h->plan.apply(x, y, precision, s);
    if (alpha != 1.0) axpby(y, y, n, alpha, 0.0, s);
  });
}

int dmrgcc_heff_info(const dmrgcc_heff* h, double* info) {
  return guarded([&] {
    require(h && info, "null argument", kErrArgument);
    info[0] = (double)h->plan.psi_dim();
    info[1] = h->plan.fref();
    info[2] = h->plan.flops();
    info[3] = (double)h->plan.bytes_min();
    info[4] = (double)h->plan.tasks();
    info[5] = (double)(h->plan.stage1().tiles.size() + h->plan.stage2().tiles.size());
    info[6] = h->plan.stage1().flops;
    info[7] = h->plan.stage2().flops;
    info[8] = (double)h->plan.stage1().tiles.size();
    info[9] = (double)h->plan.stage2().tiles.size();
    info[10] = h->plan.use_pre(kBF16x3) ? 1.0 : 0.0;
  });
}

int dmrgcc_heff_destroy(dmrgcc_heff* h) {
  delete h;
  return 0;
}

int dmrgcc_heff_tune(dmrgcc_heff* h, int precision, int reps, int cap, int* ncand, int* tile_mode,
                     int* split_k, double* ms) {
  return guarded([&] {
    require(h && ncand && tile_mode && split_k && ms, "null argument", kErrArgument);
    require(precision >= 0 && precision <= 2 && reps > 0, "bad arguments", kErrArgument);
    require_live(h);
    const std::vector<HeffPlan::Tuning> t = h->plan.tune(precision, reps, h->cx->stream);
    require(cap >= (int)t.size(), "tune: output arrays too short", kErrArgument);
    *ncand = (int)t.size();
    for (size_t i = 0; i < t.size(); ++i) {
      tile_mode[i] = t[i].tile_mode;
      split_k[i] = t[i].split_k;
      ms[i] = t[i].ms;
    }
  });
}

int dmrgcc_heff_config(const dmrgcc_heff* h, int* tile_mode, int* split_k) {
  return guarded([&] {
    require(h && tile_mode && split_k, "null argument", kErrArgument);
    *tile_mode = h->plan.tile_mode() == 32 ? 32 : 64;
    *split_k = h->plan.split_k();
  });
}

int dmrgcc_gemm_config(int tile_mode, int split_k) {
  return guarded([&] {
    require(tile_mode < 0 || tile_mode == 0 || tile_mode == 32 || tile_mode == 64, "tile mode: 32 or 64",
            kErrArgument);
    set_gemm_config(tile_mode, split_k);
  });
}

int dmrgcc_lanczos(dmrgcc_heff* h, double* ket, int num_iter, int precision, int refine,
                   double* energy, int* iters_out) {
  return guarded([&] {
    require(h && ket && energy, "null argument", kErrArgument);
    require(precision >= 0 && precision <= 2, "unknown precision", kErrArgument);
    require_live(h);
    int it = 0;
    *energy = lanczos_solve(h->plan, ket, num_iter, precision, refine, *h->cx, h->st, &it);
    if (iters_out) *iters_out = it;
  });
}

int dmrgcc_alloc(int64_t bytes, void** dev) {
  return guarded([&] {
    require(dev && bytes >= 0, "bad arguments", kErrArgument);
    *dev = bytes > 0 ? DevicePool::get().alloc((size_t)bytes) : nullptr;
  });
}

int dmrgcc_free(void* dev) {
  return guarded([&] {
    if (dev) DevicePool::get().release(dev);
  });
}

int dmrgcc_upload(void* dev, const void* host, int64_t bytes, void* stream) {
  return guarded([&] {
    require(bytes >= 0 && (bytes == 0 || (dev && host)), "bad arguments", kErrArgument);
    if (bytes == 0) return;
    hipStream_t s = static_cast<hipStream_t>(stream);
    DMRG_HIP(hipMemcpyAsync(dev, host, (size_t)bytes, hipMemcpyHostToDevice, s));
    DMRG_HIP(hipStreamSynchronize(s));
  });
}

int dmrgcc_download(void* host, const void* dev, int64_t bytes, void* stream) {
  return guarded([&] {
    require(bytes >= 0 && (bytes == 0 || (dev && host)), "bad arguments", kErrArgument);
    if (bytes == 0) return;
    hipStream_t s = static_cast<hipStream_t>(stream);
    DMRG_HIP(hipMemcpyAsync(host, dev, (size_t)bytes, hipMemcpyDeviceToHost, s));
    DMRG_HIP(hipStreamSynchronize(s));
  });
}

int dmrgcc_blocks_count(const dmrgcc_blocks* b) { return b ? (int)b->o.size() : -1; }

int dmrgcc_blocks_get(const dmrgcc_blocks* b, int i, int* o, int* l, int* r, int* rows, int* cols,
                      int64_t* off) {
  return guarded([&] {
    require(b && i >= 0 && i < (int)b->o.size(), "block index out of range", kErrArgument);
    if (o) *o = b->o[i];
    if (l) *l = b->l[i];
    if (r) *r = b->r[i];
    if (rows) *rows = b->rows[i];
    if (cols) *cols = b->cols[i];
    if (off) *off = b->off[i];
  });
}

const double* dmrgcc_blocks_data(const dmrgcc_blocks* b) { return b ? b->data.data() : nullptr; }

int dmrgcc_blocks_download(const dmrgcc_blocks* b, double* host) {
  return guarded([&] {
    require(b && host, "null argument", kErrArgument);
    int64_t total = 0;
    for (size_t i = 0; i < b->rows.size(); ++i) total += (int64_t)b->rows[i] * b->cols[i];
    if (total > 0)
      DMRG_HIP(hipMemcpy(host, b->data.data(), b->cs * total * sizeof(double), hipMemcpyDeviceToHost));
  });
}

int dmrgcc_blocks_destroy(dmrgcc_blocks* b) {
  delete b;
  return 0;
}

int dmrgcc_lattice_canonical(int leigh, int is_complex, int nl, int nr, const int* ldim, const int* rdim,
                             int nket, const int* ket_l, const int* ket_r, const int* ket_rows,
                             const int* ket_cols, const double* ket_data, int max_dim, double canpre,
                             void* stream, dmrgcc_blocks** ket_out, dmrgcc_blocks** can_out, int* dims_out,
                             int* match_dim, double* discarded) {
  return guarded([&] {
    require(leigh == 0 || leigh == 1, "leigh must be 0 or 1", kErrArgument);
    require(nl > 0 && nr > 0 && ldim && rdim && nket >= 0 && ket_out && can_out && dims_out && match_dim,
            "bad arguments", kErrArgument);
    require(nket == 0 || (ket_l && ket_r && ket_rows && ket_cols && ket_data), "bad block list", kErrArgument);
    const int cs = is_complex ? 2 : 1;
    Context cx(static_cast<hipStream_t>(stream));
    Lattice lat;
    std::vector<int> li(nl), ri(nr);
    std::iota(li.begin(), li.end(), 0);
    std::iota(ri.begin(), ri.end(), 0);
    lat.define_sectors(li, ri, std::vector<int>(ldim, ldim + nl), std::vector<int>(rdim, rdim + nr));
    lat.ket.cs = cs;
    int64_t acc = 0;
    for (int b = 0; b < nket; ++b) {
      require(ket_l[b] >= 0 && ket_l[b] < nl && ket_r[b] >= 0 && ket_r[b] < nr, "ket block sector out of range");
      require(ket_rows[b] >= 0 && ket_cols[b] >= 0, "negative block shape");
      lat.ket.blk.push_back(Block{ket_l[b], ket_r[b], ket_rows[b], ket_cols[b], acc});
      acc += (int64_t)ket_rows[b] * ket_cols[b];
    }
    lat.ket.size = acc;
    lat.ket.ext = ket_data;
    const double d = canonicalize(lat, leigh, max_dim, canpre, cx);
    auto export_blocks = [&](const BlockTensor& t) {
      auto b = std::make_unique<dmrgcc_blocks>();
      b->cs = cs;
      for (const Block& k : t.blk) {
        b->o.push_back(0);
        b->l.push_back(k.l);
        b->r.push_back(k.r);
        b->rows.push_back(k.rows);
        b->cols.push_back(k.cols);
        b->off.push_back(k.off);
      }
      b->data.resize((size_t)std::max<int64_t>(t.doubles(), 1));
      if (t.size > 0)
        DMRG_HIP(hipMemcpyAsync(b->data.data(), t.base(), t.doubles() * sizeof(double),
                                hipMemcpyDeviceToDevice, cx.stream));
      return b;
    };
    auto k = export_blocks(lat.ket);
    auto c = export_blocks(lat.can);
    cx.sync();
    const std::vector<int>& dims = leigh == 0 ? lat.rd : lat.ld;
    std::copy(dims.begin(), dims.end(), dims_out);
    std::copy(lat.match_dim.begin(), lat.match_dim.end(), match_dim);
    if (discarded) *discarded = d;
    *ket_out = k.release();
    *can_out = c.release();
  });
}

int dmrgcc_env_update(int leigh, int d, int lb, int rb, const double* W, int nl, int nr,
                      const int* phys, int nket, const int* ket_l, const int* ket_r,
                      const int* ket_rows, const int* ket_cols, const double* ket_data, int nE,
                      const int* E_o, const int* E_l, const int* E_r, const int* E_rows,
                      const int* E_cols, const int64_t* E_off, const double* E_data, void* stream,
                      dmrgcc_blocks** out) {
  return env_update(1, leigh, d, lb, rb, W, nl, nr, phys, nket, ket_l, ket_r, ket_rows, ket_cols,
                    ket_data, nE, E_o, E_l, E_r, E_rows, E_cols, E_off, E_data, stream, out);
}

int dmrgcc_zenv_update(int leigh, int d, int lb, int rb, const double* W, int nl, int nr,
                       const int* phys, int nket, const int* ket_l, const int* ket_r,
                       const int* ket_rows, const int* ket_cols, const double* ket_data, int nE,
                       const int* E_o, const int* E_l, const int* E_r, const int* E_rows,
                       const int* E_cols, const int64_t* E_off, const double* E_data, void* stream,
                       dmrgcc_blocks** out) {
  return env_update(2, leigh, d, lb, rb, W, nl, nr, phys, nket, ket_l, ket_r, ket_rows, ket_cols,
                    ket_data, nE, E_o, E_l, E_r, E_rows, E_cols, E_off, E_data, stream, out);
}

int dmrgcc_sector_svd(int n, const double* const* A, const int* M, const int* N, double* const* U,
                      double* const* S, double* const* Vt, int precondition, void* stream) {
  return guarded([&] {
    require(n >= 0 && (n == 0 || (A && M && N && U && S && Vt)), "bad arguments", kErrArgument);
    require(precondition >= 0 && precondition <= 2, "precondition must be 0, 1 or 2", kErrArgument);
    Context cx(static_cast<hipStream_t>(stream));
    struct P { int M, N, m, k; bool trans; int64_t g, v, s; };
    std::vector<P> ps(n);
    int64_t w = 0;
    for (int i = 0; i < n; ++i) {
      require(M[i] > 0 && N[i] > 0, "empty matrix in sector SVD");
      P& p = ps[i];
      p.trans = M[i] < N[i];
      p.m = std::max(M[i], N[i]);
      p.k = std::min(M[i], N[i]);
      p.M = M[i]; p.N = N[i];
      p.g = w; w += (int64_t)p.m * p.k;
      p.v = w; w += (int64_t)p.k * p.k;
      p.s = w; w += p.k + 1;  // sigma[k] and |A|_F^2
    }
    DevBuf<double> work((size_t)std::max<int64_t>(w, 1));
    std::vector<CopyTask> cps;
    std::vector<SvdProblem> probs;
    for (int i = 0; i < n; ++i) {
      const P& p = ps[i];
      CopyTask c{A[i], work.data() + p.g, p.M, p.N, p.N, 1, 0, 0, 1.0};
      if (!p.trans) { c.dsr = 1; c.dsc = p.m; }  // G = A, column-major
      else { c.dsr = p.m; c.dsc = 1; }           // G = A^T
      cps.push_back(c);
      probs.push_back(SvdProblem{work.data() + p.g, work.data() + p.v, work.data() + p.s,
                                 work.data() + p.s + p.k, p.m, p.k, p.trans ? 1 : 0,
                                 precondition});
    }
    launch_copies(cps, cx.copy_scratch, cx.stream);
    jacobi_svd(probs, cx.svd_scratch, cx.svd_sweeps, cx.stream);
    std::vector<int> hperm;
    std::vector<double> hsc, hone;
    std::vector<int64_t> poff(n);
    std::vector<std::vector<double>> sorted(n);
    for (int i = 0; i < n; ++i) {
      const P& p = ps[i];
      std::vector<double> s(p.k);
      DMRG_HIP(hipMemcpyAsync(s.data(), work.data() + p.s, p.k * sizeof(double), hipMemcpyDeviceToHost,
                              cx.stream));
      cx.sync();
      std::vector<int> ord(p.k);
      std::iota(ord.begin(), ord.end(), 0);
      std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return s[a] > s[b]; });
      poff[i] = hperm.size();
      for (int j = 0; j < p.k; ++j) {
        hperm.push_back(ord[j]);
        sorted[i].push_back(s[ord[j]]);
        hsc.push_back(s[ord[j]] > 0 ? 1.0 / s[ord[j]] : std::numeric_limits<double>::quiet_NaN());
      }
    }
    const int64_t ones = hsc.size();
    int kmax = 1;
    for (const P& p : ps) kmax = std::max(kmax, p.k);
    for (int j = 0; j < kmax; ++j) hsc.push_back(1.0);
    upload(cx.ivec, hperm.empty() ? std::vector<int>{0} : hperm, cx.stream);
    upload(cx.dvec, hsc, cx.stream);
    std::vector<ColGatherTask> gts;
    for (int i = 0; i < n; ++i) {
      const P& p = ps[i];
      const int* pp = cx.ivec.data() + poff[i];
      const double* inv = cx.dvec.data() + poff[i];
      const double* one = cx.dvec.data() + ones;
      const double* G = work.data() + p.g;
      const double* V = work.data() + p.v;
      // U (M x k) row-major, Vt (k x N) row-major
      ColGatherTask u{}, v{};
      u.perm = v.perm = pp;
      u.ncol = v.ncol = p.k;
      u.dst = U[i]; u.dsr = p.k; u.dsc = 1; u.rows = p.M;
      v.dst = Vt[i]; v.dsr = 1; v.dsc = p.N; v.rows = p.N;
      if (!p.trans) {  // A = (G V^-1) -> U = G/sigma, V = V
        u.src = G; u.ld = p.m; u.scale = inv;
        v.src = V; v.ld = p.k; v.scale = one;
      } else {         // A^T = G -> U = V, V = G/sigma
        u.src = V; u.ld = p.k; u.scale = one;
        v.src = G; v.ld = p.m; v.scale = inv;
      }
      gts.push_back(u);
      gts.push_back(v);
      DMRG_HIP(hipMemcpyAsync(S[i], sorted[i].data(), p.k * sizeof(double), hipMemcpyHostToDevice,
                              cx.stream));
    }
    launch_col_gather(gts, cx.gather_scratch, cx.stream);
    cx.sync();
  });
}


int dmrgcc_zsector_svd(int n, const double* const* A, const int* M, const int* N, double* const* U,
                       double* const* S, double* const* Vt, void* stream) {
  return guarded([&] {
    require(n >= 0 && (n == 0 || (A && M && N && U && S && Vt)), "bad arguments", kErrArgument);
    Context cx(static_cast<hipStream_t>(stream));
    // G = A (M >= N) or A^T (M < N), column-major complex; G W = U' Sigma
    struct P { int M, N, m, k; bool trans; int64_t g, v, s; };
    std::vector<P> ps(n);
    int64_t w = 0;
    for (int i = 0; i < n; ++i) {
      require(M[i] > 0 && N[i] > 0, "empty matrix in sector SVD");
      P& p = ps[i];
      p.trans = M[i] < N[i];
      p.m = std::max(M[i], N[i]);
      p.k = std::min(M[i], N[i]);
      p.M = M[i]; p.N = N[i];
      p.g = w; w += 2 * (int64_t)p.m * p.k;
      p.v = w; w += 2 * (int64_t)p.k * p.k;
      p.s = w; w += p.k + 1;  // sigma[k] and |A|_F^2
    }
    DevBuf<double> work((size_t)std::max<int64_t>(w, 1));
    std::vector<CopyTask> cps;
    std::vector<SvdProblem> probs;
    for (int i = 0; i < n; ++i) {
      const P& p = ps[i];
      for (int part = 0; part < 2; ++part) {
        CopyTask c{A[i] + part, work.data() + p.g + part, p.M, p.N, 2 * (int64_t)p.N, 2, 0, 0, 1.0};
        if (!p.trans) { c.dsr = 2; c.dsc = 2 * (int64_t)p.m; }
        else { c.dsr = 2 * (int64_t)p.m; c.dsc = 2; }
        cps.push_back(c);
      }
      probs.push_back(SvdProblem{work.data() + p.g, work.data() + p.v, work.data() + p.s,
                                 work.data() + p.s + p.k, p.m, p.k, p.trans ? 1 : 0, 2});
    }
    launch_copies(cps, cx.copy_scratch, cx.stream);
    zjacobi_svd(probs, cx.svd_scratch, cx.svd_sweeps, cx.stream);
    std::vector<int> hperm;
    std::vector<double> hsc;
    std::vector<int64_t> poff(n);
    std::vector<std::vector<double>> sorted(n);
    for (int i = 0; i < n; ++i) {
      const P& p = ps[i];
      std::vector<double> sv(p.k);
      DMRG_HIP(hipMemcpyAsync(sv.data(), work.data() + p.s, p.k * sizeof(double), hipMemcpyDeviceToHost,
                              cx.stream));
      cx.sync();
      std::vector<int> ord(p.k);
      std::iota(ord.begin(), ord.end(), 0);
      std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return sv[a] > sv[b]; });
      poff[i] = hperm.size();
      for (int j = 0; j < p.k; ++j) {
        hperm.push_back(ord[j]);
        sorted[i].push_back(sv[ord[j]]);
        hsc.push_back(sv[ord[j]] > 0 ? 1.0 / sv[ord[j]] : std::numeric_limits<double>::quiet_NaN());
      }
    }
    const int64_t ones = hsc.size();
    int kmax = 1;
    for (const P& p : ps) kmax = std::max(kmax, p.k);
    for (int j = 0; j < kmax; ++j) hsc.push_back(1.0);
    upload(cx.ivec, hperm.empty() ? std::vector<int>{0} : hperm, cx.stream);
    upload(cx.dvec, hsc, cx.stream);
    std::vector<ColGatherTask> gts;
    for (int i = 0; i < n; ++i) {
      const P& p = ps[i];
      const int* pp = cx.ivec.data() + poff[i];
      const double* inv = cx.dvec.data() + poff[i];
      const double* one = cx.dvec.data() + ones;
      const double* G = work.data() + p.g;
      const double* V = work.data() + p.v;
      // U (M x k) and Vt (k x N), row-major complex
      ColGatherTask u{}, v{};
      u.perm = v.perm = pp;
      u.ncol = v.ncol = p.k;
      u.dst = U[i]; u.dsr = p.k; u.dsc = 1; u.rows = p.M;
      v.dst = Vt[i]; v.dsr = 1; v.dsc = p.N; v.rows = p.N;
      if (!p.trans) {  // A = U' Sigma W^H: U = G / sigma, Vt = W^H
        u.src = G; u.ld = p.m; u.scale = inv; u.flags = kGatherComplex;
        v.src = V; v.ld = p.k; v.scale = one; v.flags = kGatherComplex | kGatherConj;
      } else {         // A^T = U' Sigma W^H: U = conj(W), Vt = U'^T = (G / sigma)^T
        u.src = V; u.ld = p.k; u.scale = one; u.flags = kGatherComplex | kGatherConj;
        v.src = G; v.ld = p.m; v.scale = inv; v.flags = kGatherComplex;
      }
      gts.push_back(u);
      gts.push_back(v);
      DMRG_HIP(hipMemcpyAsync(S[i], sorted[i].data(), p.k * sizeof(double), hipMemcpyHostToDevice,
                              cx.stream));
    }
    launch_col_gather(gts, cx.gather_scratch, cx.stream);
    cx.sync();
  });
}

int dmrgcc_dmrg_create(dmrgcc_model* m, int first_block, int first_dim, int nhalf, const int* maxm,
                       const int* maxb, const int* niter, const double* canpre, const double* noise,
                       int precision, int refine, int verbose, dmrgcc_dmrg** out) {
  return dmrgcc_dmrg_create_files(m, first_block, first_dim, nhalf, maxm, maxb, niter, canpre, noise,
                                  precision, refine, verbose, nullptr, nullptr, out);
}

int dmrgcc_dmrg_create_files(dmrgcc_model* m, int first_block, int first_dim, int nhalf, const int* maxm,
                             const int* maxb, const int* niter, const double* canpre, const double* noise,
                             int precision, int refine, int verbose, const char* resume_path,
                             const char* record_path, dmrgcc_dmrg** out) {
  return guarded([&] {
    require(m && out, "null argument", kErrArgument);
    RunParams p = make_params(first_block, first_dim, nhalf, maxm, maxb, niter, canpre, noise, precision,
                              refine, verbose);
    if (resume_path) p.resume_path = resume_path;
    if (record_path) p.record_path = record_path;
    auto h = std::make_unique<dmrgcc_dmrg>();
    h->model = m;
    h->run = std::make_unique<RealTensorRundmrg>(*m->space, *m->ham, p);
    h->run->Start();
    *out = h.release();
  });
}

int dmrgcc_dmrg_write_space(dmrgcc_dmrg* h, const char* path) {
  return guarded([&] {
    require(h && path, "null argument", kErrArgument);
    h->run->WriteTensorSpace(path);
  });
}

int dmrgcc_dmrg_write_env(dmrgcc_dmrg* h, int site, int leigh, const char* path) {
  return guarded([&] {
    require(h && path && site >= 0 && site < h->run->num_site() && (leigh == 0 || leigh == 1), "bad arguments",
            kErrArgument);
    const Env& e = leigh == 0 ? h->run->left_env(site) : h->run->right_env(site);
    require(!e.empty(), "no environment at this site on the current state", kErrArgument);
    h->run->context().sync();
    WriteTensorContraction(e, path, h->run->context().stream);
  });
}

int dmrgcc_env_file_read(const char* path, int nbond, int is_complex, dmrgcc_blocks** out) {
  return guarded([&] {
    require(path && out && nbond > 0, "bad arguments", kErrArgument);
    const int cs = is_complex ? 2 : 1;
    Context cx(nullptr);
    Env e = ReadTensorContraction(path, nbond, cs, cx.stream);
    auto b = std::make_unique<dmrgcc_blocks>();
    b->cs = cs;
    int64_t total = 0;
    for (const auto& t : e) total += t.size;
    b->data.resize((size_t)std::max<int64_t>(cs * total, 1));
    int64_t acc = 0;
    for (size_t o = 0; o < e.size(); ++o) {
      for (const Block& k : e[o].blk) {
        b->o.push_back((int)o);
        b->l.push_back(k.l);
        b->r.push_back(k.r);
        b->rows.push_back(k.rows);
        b->cols.push_back(k.cols);
        b->off.push_back(acc + k.off);
      }
      if (e[o].size > 0)
        DMRG_HIP(hipMemcpyAsync(b->data.data() + cs * acc, e[o].data.data(), cs * e[o].size * sizeof(double),
                                hipMemcpyDeviceToDevice, cx.stream));
      acc += e[o].size;
    }
    cx.sync();
    *out = b.release();
  });
}

int dmrgcc_model_write_mpo(const dmrgcc_model* m, const char* path) {
  return guarded([&] {
    require(m && path, "null argument", kErrArgument);
    m->ham->WriteTensorHamiltonian(path);
  });
}

int dmrgcc_model_read_mpo(dmrgcc_model* m, const char* path) {
  return guarded([&] {
    require(m && path, "null argument", kErrArgument);
    m->ham->ReadTensorHamiltonian(path);
  });
}

int dmrgcc_dmrg_advance(dmrgcc_dmrg* h, int half, int site, int* reached) {
  return guarded([&] {
    require(h != nullptr, "null handle", kErrArgument);
    ++*h->epoch;  // plans built from the run point at buffers this may replace
    const bool r = h->run->Advance(half, site);
    if (reached) *reached = r ? 1 : 0;
  });
}

int dmrgcc_dmrg_set_precision(dmrgcc_dmrg* h, int precision, int refine) {
  return guarded([&] {
    require(h != nullptr && precision >= 0 && precision <= 2, "bad arguments", kErrArgument);
    h->run->set_precision(precision, refine);
  });
}

int dmrgcc_dmrg_set_autotune(dmrgcc_dmrg* h, int on, int64_t min_dim) {
  return guarded([&] {
    require(h != nullptr, "null handle", kErrArgument);
    h->run->set_autotune(on != 0, min_dim);
  });
}

int dmrgcc_dmrg_heff(dmrgcc_dmrg* h, int site, dmrgcc_heff** out) {
  return guarded([&] {
    require(h && out && site >= 0 && site < h->run->num_site(), "bad arguments", kErrArgument);
    auto e = std::make_unique<dmrgcc_heff>();
    e->cx = std::make_unique<Context>(nullptr);
    h->run->BuildHeff(site, e->plan);
    h->run->context().sync();
    e->run_epoch = h->epoch;
    e->epoch = *h->epoch;
    *out = e.release();
  });
}

int dmrgcc_heff_batch_create(int n, dmrgcc_heff* const* plans, const double* const* x, double* const* y,
                             dmrgcc_heff_batch** out) {
  return guarded([&] {
    require(n > 0 && plans && x && y && out, "bad arguments", kErrArgument);
    // Each plan owns one stage buffer; the same plan twice in one launch
    // would have both instances write it at once.  Checked before anything
    // is dereferenced or allocated.
    {
      std::vector<const dmrgcc_heff*> seen(plans, plans + n);
      std::sort(seen.begin(), seen.end());
      require(std::adjacent_find(seen.begin(), seen.end()) == seen.end(),
              "H_eff batch: the same plan appears more than once (one stage buffer per plan)",
              kErrArgument);
    }
    for (int i = 0; i < n; ++i) {
      require(plans[i] != nullptr, "null plan in batch", kErrArgument);
      require_live(plans[i]);
    }
    auto b = std::make_unique<dmrgcc_heff_batch>();
    b->cx = std::make_unique<Context>(nullptr);
    std::vector<HeffPlan*> ps;
    std::vector<const double*> xs;
    std::vector<double*> ys;
    for (int i = 0; i < n; ++i) {
      require(plans[i] && x[i] && y[i], "null plan or vector in batch", kErrArgument);
      ps.push_back(&plans[i]->plan);
      xs.push_back(x[i]);
      ys.push_back(y[i]);
      b->plans.push_back(plans[i]);
    }
    b->batch.build(ps, xs, ys, b->cx->stream);
    *out = b.release();
  });
}

int dmrgcc_heff_batch_apply(dmrgcc_heff_batch* b, int precision, int use_graph) {
  return guarded([&] {
    require(b && precision >= 0 && precision <= 2, "bad arguments", kErrArgument);
    for (const dmrgcc_heff* p : b->plans) require_live(p);
    if (use_graph) b->batch.apply_graph(precision, b->cx->stream);
    else b->batch.apply(precision, b->cx->stream);
    DMRG_HIP(hipStreamSynchronize(b->cx->stream));
  });
}

int dmrgcc_heff_batch_time(dmrgcc_heff_batch* b, int precision, int reps, double* ms) {
  return guarded([&] {
    require(b && ms && reps > 0 && precision >= 0 && precision <= 2, "bad arguments", kErrArgument);
    for (const dmrgcc_heff* p : b->plans) require_live(p);
    hipStream_t s = b->cx->stream;
    hipEvent_t e0, e1;
    DMRG_HIP(hipEventCreate(&e0));
    DMRG_HIP(hipEventCreate(&e1));
    float t = 0;
    // [0] graph replay of the whole batch, [1] the plans one after another
    b->batch.apply_graph(precision, s);
    DMRG_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) b->batch.apply_graph(precision, s);
    DMRG_HIP(hipEventRecord(e1, s));
    DMRG_HIP(hipEventSynchronize(e1));
    DMRG_HIP(hipEventElapsedTime(&t, e0, e1));
    ms[0] = t / reps;
    DMRG_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) b->batch.apply_each(precision, s);
    DMRG_HIP(hipEventRecord(e1, s));
    DMRG_HIP(hipEventSynchronize(e1));
    DMRG_HIP(hipEventElapsedTime(&t, e0, e1));
    ms[1] = t / reps;
    DMRG_HIP(hipEventDestroy(e0));
    DMRG_HIP(hipEventDestroy(e1));
  });
}

int dmrgcc_heff_batch_info(const dmrgcc_heff_batch* b, double* info) {
  return guarded([&] {
    require(b && info, "null argument", kErrArgument);
    info[0] = b->batch.fref();
    info[1] = (double)b->batch.tiles();
    info[2] = (double)b->plans.size();
  });
}

int dmrgcc_heff_batch_destroy(dmrgcc_heff_batch* b) {
  delete b;
  return 0;
}

int dmrgcc_dmrg_heff2(dmrgcc_dmrg* h, int site, dmrgcc_heff** out) {
  return guarded([&] {
    require(h && out && site >= 0 && site + 1 < h->run->num_site(), "bad arguments", kErrArgument);
    auto e = std::make_unique<dmrgcc_heff>();
    e->cx = std::make_unique<Context>(nullptr);
    h->run->BuildTwoSite(site, e->two, e->plan);
    h->run->context().sync();
    e->run_epoch = h->epoch;
    e->epoch = *h->epoch;
    *out = e.release();
  });
}

int dmrgcc_heff_state(const dmrgcc_heff* h, double* dst) {
  return guarded([&] {
    require(h && dst, "null argument", kErrArgument);
    require(h->two.theta.size > 0, "no two-site state on this plan", kErrArgument);
    require_live(h);
    DMRG_HIP(hipMemcpyAsync(dst, h->two.theta.base(), h->two.theta.doubles() * sizeof(double),
                            hipMemcpyDeviceToDevice, h->cx->stream));
    DMRG_HIP(hipStreamSynchronize(h->cx->stream));
  });
}

int dmrgcc_dmrg_ket(dmrgcc_dmrg* h, int site, double** ket, int64_t* n) {
  return guarded([&] {
    require(h && ket && n && site >= 0 && site < h->run->num_site(), "bad arguments", kErrArgument);
    Lattice& lat = h->run->lattice(site);
    *ket = lat.ket.data.data();
    *n = lat.ket.size;
  });
}

int dmrgcc_dmrg_site_blocks(dmrgcc_dmrg* h, int site, int which, int* count, int64_t* total, int* o,
                            int* l, int* r, int* rows, int* cols, int64_t* off, double* data) {
  return guarded([&] {
    require(h && count && total && site >= 0 && site < h->run->num_site() && which >= 0 && which <= 2,
            "bad arguments", kErrArgument);
    std::vector<const BlockTensor*> ts;
    if (which == 0) ts.push_back(&h->run->lattice(site).ket);
    else for (const auto& b : (which == 1 ? h->run->left_env(site) : h->run->right_env(site))) ts.push_back(&b);
    int n = 0;
    int64_t tot = 0;
    for (const BlockTensor* t : ts) { n += t->n(); tot += t->size; }
    *count = n;
    *total = tot;
    if (!o) return;
    int k = 0;
    int64_t acc = 0;
    for (size_t c = 0; c < ts.size(); ++c) {
      const BlockTensor* t = ts[c];
      for (const Block& b : t->blk) {
        o[k] = (int)c; l[k] = b.l; r[k] = b.r; rows[k] = b.rows; cols[k] = b.cols; off[k] = acc + b.off;
        ++k;
      }
      if (data && t->size > 0)
        DMRG_HIP(hipMemcpy(data + t->cs * acc, t->base(), t->doubles() * sizeof(double),
                           hipMemcpyDeviceToHost));
      acc += t->size;
    }
  });
}

int dmrgcc_dmrg_phys(dmrgcc_dmrg* h, int site, int* nl, int* nr, int* phys) {
  return guarded([&] {
    require(h && nl && nr && site >= 0 && site < h->run->num_site(), "bad arguments", kErrArgument);
    const Lattice& lat = h->run->lattice(site);
    *nl = lat.nl();
    *nr = lat.nr();
    if (phys)
      for (int i = 0; i < lat.nl(); ++i)
        for (int j = 0; j < lat.nr(); ++j) phys[(size_t)i * lat.nr() + j] = lat.phys[i][j];
  });
}

int dmrgcc_dmrg_result(dmrgcc_dmrg* h, int* nhalf_done, double* energies, double* trunc,
                       double* seconds, double* stats) {
  return guarded([&] {
    require(h != nullptr, "null handle", kErrArgument);
    fill_result(h->run->Result(), nhalf_done, energies, trunc, seconds, stats);
  });
}

int dmrgcc_dmrg_destroy(dmrgcc_dmrg* h) {
  delete h;
  return 0;
}

int dmrgcc_dmrg_env_time(dmrgcc_dmrg* h, int site, int leigh, int reps, double* out) {
  return guarded([&] {
    require(h && out && reps > 0 && (leigh == 1 || leigh == 2), "bad arguments", kErrArgument);
    RealTensorRundmrg& r = *h->run;
    require(site >= 0 && site < r.num_site(), "site out of range", kErrArgument);
    double f = 0, o = 0;
    EnvTiming tm;
    for (int i = 0; i < reps; ++i) {
      tm = EnvTiming{};
      Env e = leigh == 1 ? left_env_update(r.lattice(site), r.op(site), r.left_env(site), r.context(), &tm)
                         : right_env_update(r.lattice(site), r.op(site), r.right_env(site), r.context(), &tm);
      f += tm.ms_f;
      o += tm.ms_out;
    }
    out[0] = f / reps;
    out[1] = o / reps;
    out[2] = tm.bytes_min;
    out[3] = tm.bytes_exec;
    out[4] = tm.flops;
  });
}

int dmrgcc_vector_time(int64_t n, int reps, double* ms) {
  return guarded([&] {
    require(ms && n > 0 && reps > 0, "bad arguments", kErrArgument);
    Context cx(nullptr);
    DevBuf<double> x((size_t)n), y((size_t)n), h((size_t)n), s(2);
    fill(x.data(), n, 1e-3, cx.stream);
    fill(y.data(), n, 2e-3, cx.stream);
    fill(h.data(), n, 0.0, cx.stream);
    fill(s.data(), 2, 0.5, cx.stream);
    hipEvent_t e[2];
    for (auto& ev : e) DMRG_HIP(hipEventCreate(&ev));
    float t[3] = {0, 0, 0};
    for (int k = 0; k < 3; ++k) {
      for (int rep = -1; rep < reps; ++rep) {  // rep -1: warm-up
        if (rep == 0) DMRG_HIP(hipEventRecord(e[0], cx.stream));
        if (k == 0) dot(x.data(), y.data(), n, s.data(), cx.partial, cx.stream);
        else if (k == 1) axpy_dev(y.data(), x.data(), n, s.data() + 1, -1.0, cx.stream);
        else lanczos_step(h.data(), y.data(), x.data(), n, 0.25, cx.stream);
      }
      DMRG_HIP(hipEventRecord(e[1], cx.stream));
      DMRG_HIP(hipEventSynchronize(e[1]));
      DMRG_HIP(hipEventElapsedTime(&t[k], e[0], e[1]));
    }
    for (auto& ev : e) (void)hipEventDestroy(ev);
    for (int k = 0; k < 3; ++k) ms[k] = t[k] / reps;
  });
}

int dmrgcc_heff_time(dmrgcc_heff* h, int precision, int reps, double* ms) {
  return guarded([&] {
    require(h && ms && reps > 0 && precision >= 0 && precision <= 2, "bad arguments", kErrArgument);
    require_live(h);
    const int64_t n = std::max<int64_t>(h->plan.doubles(), 1);
    DevBuf<double> x((size_t)n), y((size_t)n);
    fill(x.data(), n, 1e-3, h->cx->stream);
    fill(y.data(), n, 0.0, h->cx->stream);
    hipStream_t s = h->cx->stream;
    hipEvent_t e[3];
    for (auto& ev : e) DMRG_HIP(hipEventCreate(&ev));
    // warm-up
    h->plan.apply(x.data(), y.data(), precision, s);
    float t1 = 0, t2 = 0, tm = 0;
    for (int r = 0; r < reps; ++r) {
      DMRG_HIP(hipEventRecord(e[0], s));
      h->plan.apply_stage(1, x.data(), y.data(), precision, s);
      DMRG_HIP(hipEventRecord(e[1], s));
      h->plan.apply_stage(2, x.data(), y.data(), precision, s);
      DMRG_HIP(hipEventRecord(e[2], s));
      DMRG_HIP(hipEventSynchronize(e[2]));
      float a, b;
      DMRG_HIP(hipEventElapsedTime(&a, e[0], e[1]));
      DMRG_HIP(hipEventElapsedTime(&b, e[1], e[2]));
      t1 += a; t2 += b;
    }
    // back-to-back matvecs, no host synchronisation in between
    DMRG_HIP(hipEventRecord(e[0], s));
    for (int r = 0; r < reps; ++r) h->plan.apply(x.data(), y.data(), precision, s);
    DMRG_HIP(hipEventRecord(e[1], s));
    DMRG_HIP(hipEventSynchronize(e[1]));
    DMRG_HIP(hipEventElapsedTime(&tm, e[0], e[1]));
    // each stage's GEMM kernel alone, back to back (what a kernel trace
    // reports per dispatch)
    float g[2] = {0, 0};
    for (int st = 1; st <= 2; ++st) {
      DMRG_HIP(hipEventRecord(e[0], s));
      for (int r = 0; r < reps; ++r) h->plan.apply_stage(st, x.data(), y.data(), precision, s, true);
      DMRG_HIP(hipEventRecord(e[1], s));
      DMRG_HIP(hipEventSynchronize(e[1]));
      DMRG_HIP(hipEventElapsedTime(&g[st - 1], e[0], e[1]));
    }
    for (auto& ev : e) (void)hipEventDestroy(ev);
    ms[0] = t1 / reps;
    ms[1] = t2 / reps;
    ms[2] = tm / reps;
    ms[3] = g[0] / reps;
    ms[4] = g[1] / reps;
  });
}

int dmrgcc_run(dmrgcc_model* m, int first_block, int first_dim, int nhalf, const int* maxm,
               const int* maxb, const int* niter, const double* canpre, const double* noise,
               int precision, int refine, int verbose, double* energies, double* trunc,
               double* seconds, double* stats) {
  return dmrgcc_run_files(m, first_block, first_dim, nhalf, maxm, maxb, niter, canpre, noise, precision,
                          refine, verbose, nullptr, nullptr, energies, trunc, seconds, stats);
}

int dmrgcc_run_two_site(dmrgcc_model* m, int first_block, int first_dim, int nhalf, const int* maxm,
                        const int* maxb, const int* niter, const double* canpre, const double* noise,
                        int precision, int refine, int verbose, double* energies, double* trunc,
                        double* seconds, double* stats) {
  return guarded([&] {
    require(m != nullptr, "null model", kErrArgument);
    RunParams p = make_params(first_block, first_dim, nhalf, maxm, maxb, niter, canpre, noise, precision,
                              refine, verbose);
    p.two_site = true;
    RealTensorRundmrg run(*m->space, *m->ham, p);
    fill_result(run.Run(), nullptr, energies, trunc, seconds, stats);
  });
}

int dmrgcc_run_files(dmrgcc_model* m, int first_block, int first_dim, int nhalf, const int* maxm,
                     const int* maxb, const int* niter, const double* canpre, const double* noise,
                     int precision, int refine, int verbose, const char* resume_path,
                     const char* record_path, double* energies, double* trunc, double* seconds,
                     double* stats) {
  return guarded([&] {
    require(m != nullptr, "null model", kErrArgument);
    RunParams p = make_params(first_block, first_dim, nhalf, maxm, maxb, niter, canpre, noise, precision,
                              refine, verbose);
    if (resume_path) p.resume_path = resume_path;
    if (record_path) p.record_path = record_path;
    RealTensorRundmrg run(*m->space, *m->ham, p);
    fill_result(run.Run(), nullptr, energies, trunc, seconds, stats);
  });
}

}  // extern "C"
