// Input-file parsing of the reference's drivers (src/util/input.cc:1-331,
// real_tensor_rundmrg.cc:3-113, model/*_model.cc), restated on a
// std::istringstream over the file's text so that reads, put-backs and the
// stream extractors behave as on the reference's ifstream.
#include "input.h"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <fstream>

#include "common.h"

namespace dmrg {

bool InputGroup::echo = true;

void InputFile::open() {
  // InputFile::open (input.cc:3-10): every lookup starts at the top
  if (!loaded_) {
    std::ifstream f(name_, std::ios::binary);
    require(f.good(), "can't open input file", kErrArgument);
    std::ostringstream ss;
    ss << f.rdbuf();
    text_ = ss.str();
    loaded_ = true;
  }
  is_.clear();
  is_.str(text_);
  is_.seekg(0);
}

namespace {

// eatwhite / eatblank (input.cc:24-50)
void eatwhite(std::istream& is) {
  char c;
  while (is.get(c))
    if (!std::isspace((unsigned char)c)) {
      is.putback(c);
      break;
    }
}
void eatblank(std::istream& is) {
  char c;
  while (is.get(c))
    if (!std::isblank((unsigned char)c)) {
      is.putback(c);
      break;
    }
}

// matchbracket (input.cc:52-65): skip to the bracket closing an open one
void matchbracket(std::istream& is) {
  int nest = 1;
  char c;
  while (is >> c) {
    if (c == '{') nest++;
    else if (c == '}') nest--;
    if (nest == 0) break;
  }
  require(nest == 0, "unterminated bracket", kErrArgument);
}

// gettoken (input.cc:67-93): an identifier, or one non-identifier character
int gettoken(std::istream& is, std::string& s) {
  eatwhite(is);
  std::string t;
  for (int i = 0; i < 511; ++i) {
    char ch;
    if (!is.get(ch)) return 0;
    if (std::isalpha((unsigned char)ch) || ch == '_' || (i != 0 && std::isdigit((unsigned char)ch))) {
      t.push_back(ch);
    } else {
      if (i == 0) {
        s = std::string(1, ch);
        return 1;
      }
      is.putback(ch);
      break;
    }
  }
  s = t;
  return 1;
}

// the rest of a line that did not match, with any bracketed group on it
// (input.cc:138-151, :179-191); false at the end of the file
bool skip_rest(std::istream& is) {
  char c;
  while (is.get(c)) {
    if (c == '{') matchbracket(is);
    if (c == '\n') {
      eatwhite(is);
      return true;
    }
  }
  return false;
}

}  // namespace

InputGroup::InputGroup(const std::string& file_name, const std::string& group_name)
    : file_(std::make_shared<InputFile>(file_name)), name_(group_name) {}

InputGroup::InputGroup(InputGroup& parent, const std::string& name)
    : file_(parent.file_), parent_(&parent), name_(name) {}

int InputGroup::GotoGroup() {
  // input.cc:112-154
  if (parent_) {
    if (!parent_->GotoGroup()) return 0;
  } else {
    file_->open();
  }
  std::istream& is = file();
  eatwhite(is);
  while (true) {
    std::string s;
    if (!gettoken(is, s)) return 0;
    if (s == name_) {
      eatwhite(is);
      char c = 0;
      is.get(c);
      require(c == '{', "bracket does not follow group name", kErrArgument);
      eatwhite(is);
      return 1;
    }
    if (s[0] == '{') matchbracket(is);
    if (!skip_rest(is)) return 0;
  }
}

int InputGroup::GotoToken(const std::string& s) {
  // input.cc:157-193
  if (!GotoGroup()) return 0;
  std::istream& is = file();
  while (true) {
    std::string t;
    if (!gettoken(is, t)) return 0;
    if (t[0] == '}') return 0;
    if (t == s) {
      eatwhite(is);
      char c;
      if (!is.get(c)) return 0;
      if (c != '=') return 0;
      eatblank(is);
      return 1;
    }
    if (t[0] == '{') matchbracket(is);
    if (!skip_rest(is)) return 0;
  }
}

void InputGroup::SkipLine() {
  // input.cc:195-201
  char c = '\0';
  while (c != '\n')
    if (!file().get(c)) break;
  eatwhite(file());
}

int InputGroup::GetInt(const std::string& s, int& res, bool hasdf, int df) {
  if (!GotoToken(s) || !(file() >> res)) {
    if (hasdf && echo) std::printf("Def %s.%s = %d\n", name_.c_str(), s.c_str(), df);
    return 0;
  }
  if (echo) std::printf("Got %s.%s = %d\n", name_.c_str(), s.c_str(), res);
  return 1;
}

int InputGroup::GetDouble(const std::string& s, double& res, bool hasdf, double df) {
  if (!GotoToken(s) || !(file() >> res)) {
    if (hasdf && echo) std::printf("Def %s.%s = %.2f\n", name_.c_str(), s.c_str(), df);
    return 0;
  }
  if (echo) std::printf("Got %s.%s = %.2f\n", name_.c_str(), s.c_str(), res);
  return 1;
}

int InputGroup::GetString(const std::string& s, std::string& res, bool hasdf, const std::string& df) {
  // input.cc:223-244: a key with nothing after '=' counts as found, empty
  if (GotoToken(s)) {
    char c;
    if (file().get(c)) {
      if (std::isspace((unsigned char)c)) {
        if (echo) {
          if (hasdf) std::printf("Def %s.%s = %s\n", name_.c_str(), s.c_str(), df.c_str());
          else std::printf("No  %s.%s = %s\n", name_.c_str(), s.c_str(), "None");
        }
        return 1;
      }
      file().putback(c);
    }
  }
  if (!GotoToken(s) || !(file() >> res)) {
    if (hasdf && echo) std::printf("Def %s.%s = %s\n", name_.c_str(), s.c_str(), df.c_str());
    return 0;
  }
  if (echo) std::printf("Got %s.%s = %s\n", name_.c_str(), s.c_str(), res.c_str());
  return 1;
}

int InputGroup::GetBool(const std::string& s, bool& res, bool hasdf, bool df) {
  // input.cc:246-262: "true"/"1" and "false"/"0"; other words leave res
  std::string t;
  if (!GotoToken(s) || !(file() >> t)) {
    if (hasdf && echo) std::printf("Def %s.%s = %d\n", name_.c_str(), s.c_str(), (int)df);
    return 0;
  }
  if (t == "true" || t == "1") res = true;
  else if (t == "false" || t == "0") res = false;
  if (echo) std::printf("Got %s.%s = %d\n", name_.c_str(), s.c_str(), (int)res);
  return 1;
}

int InputGroup::getInt(const std::string& s, int def) {
  int res = 0;
  return GetInt(s, res, true, def) ? res : def;
}
double InputGroup::getDouble(const std::string& s, double def) {
  double res = 0.0;
  return GetDouble(s, res, true, def) ? res : def;
}
std::string InputGroup::getString(const std::string& s, const std::string& def) {
  std::string res;
  return GetString(s, res, true, def) ? res : def;
}
bool InputGroup::getBool(const std::string& s, bool def) {
  bool res = false;
  return GetBool(s, res, true, def) ? res : def;
}
int InputGroup::getInt(const std::string& s) {
  int res = 0;
  require(GetInt(s, res), "there is no item: " + s, kErrArgument);
  return res;
}
double InputGroup::getDouble(const std::string& s) {
  double res = 0.0;
  require(GetDouble(s, res), "there is no item: " + s, kErrArgument);
  return res;
}
std::string InputGroup::getString(const std::string& s) {
  std::string res;
  require(GetString(s, res), "there is no item: " + s, kErrArgument);
  return res;
}
bool InputGroup::getBool(const std::string& s) {
  bool res = false;
  require(GetBool(s, res), "there is no item: " + s, kErrArgument);
  return res;
}

RunConfig read_run_config(InputGroup& input) {
  // RealTensorRundmrg::RealTensorRundmrg (real_tensor_rundmrg.cc:3-42), same
  // keys in the complex twin (complex_tensor_rundmrg.cc:3-42)
  RunConfig c;
  c.disk_cache = input.getBool("disk_cache");
  c.cache_name = input.getString("cache_name");
  c.cache_record = input.getBool("cache_record");
  c.record_name = input.getString("record_name");
  c.cache_resume = input.getBool("cache_resume");
  c.resume_name = input.getString("resume_name");
  c.record_process = input.getBool("record_process");
  c.process_name = input.getString("process_name");
  c.first_block = input.getInt("first_block");
  c.first_dim = input.getInt("first_dim");
  c.num_sweep = input.getInt("num_sweep");
  const int n = 2 * c.num_sweep;
  require(n > 0, "num_sweep must be positive", kErrArgument);
  // Sweep (real_tensor_rundmrg.cc:70-113): rows after the header line; a
  // row with maxm 0, or the first one the extractor cannot read (which then
  // stores 0), ends the table, and the rest repeats the last row with maxm
  // taken from its maxb column, as the reference does
  InputGroup table(input, "sweeps");
  require(table.GotoGroup(), "Couldn't find table " + table.name(), kErrArgument);
  table.SkipLine();
  std::vector<int> maxm(n, 0), maxb(n, 0), niter(n, 0);
  std::vector<double> canpre(n, 0.0), noise(n, 0.0);
  int nlast = n;
  std::istream& is = table.file();
  for (int i = 0; i < n; ++i) {
    is >> maxm[i] >> maxb[i] >> niter[i] >> canpre[i] >> noise[i];
    if (maxm[i] == 0) {
      nlast = i - 1;
      break;
    }
  }
  require(nlast >= 0, "empty sweep table", kErrArgument);
  for (int i = nlast + 1; i < n; ++i) {
    maxm[i] = maxb[nlast];
    maxb[i] = maxb[nlast];
    niter[i] = niter[nlast];
    canpre[i] = canpre[nlast];
    noise[i] = noise[nlast];
  }
  if (InputGroup::echo) {
    std::printf("Sweeps:\n");
    for (int i = 0; i < n; ++i)
      std::printf("%d Maxm=%d, Maxb=%d, Niter=%d, Canon=%g, Noise=%g\n", i + 1, maxm[i], maxb[i], niter[i],
                  canpre[i], noise[i]);
  }
  c.maxm = maxm;
  c.maxb = maxb;
  c.niter = niter;
  c.canpre = canpre;
  c.noise = noise;
  return c;
}

void read_model(InputGroup& input, const std::string& driver, std::unique_ptr<RealTensorSpace>& space,
                std::unique_ptr<RealTensorHamiltonian>& ham) {
  if (driver == "real_spinless_bose_square" || driver == "complex_spinless_bose_square") {
    // model/real_spinless_bose_square_model.cc:11-22 (the complex driver
    // reads the same keys); flux is read and passed on, and the Hamiltonian
    // does not use it
    const int nx = input.getInt("num_site_x");
    const int ny = input.getInt("num_site_y");
    const int nb = input.getInt("num_boson");
    const int d = input.getInt("physics_dim", 2);
    const double flux = input.getDouble("flux") * M_PI;
    const double jx = input.getDouble("Jx");
    const double jy = input.getDouble("Jy");
    const double U = input.getDouble("U");
    const double mu = input.getDouble("mu");
    const bool px = input.getBool("period_x");
    const bool py = input.getBool("period_y");
    require(nx > 0 && ny > 0 && d > 1 && nb >= 0, "bad model parameters in input file", kErrArgument);
    space = std::make_unique<RealSpinlessBoseSquareSpace>(nx, ny, nb, d);
    if (driver == "real_spinless_bose_square")
      ham = std::make_unique<RealSpinlessBoseSquareHamiltonian>(nx, ny, d, flux, jx, jy, U, mu, px, py);
    else
      ham = std::make_unique<ComplexSpinlessBoseSquareHamiltonian>(nx, ny, d, flux, jx, jy, U, mu, px, py);
  } else if (driver == "complex_spinful_bose_square") {
    // model/complex_spinful_bose_square_model.cc:11-36
    const int nx = input.getInt("num_site_x");
    const int ny = input.getInt("num_site_y");
    const int d = input.getInt("physics_dim", 2);
    const bool px = input.getBool("period_x");
    const bool py = input.getBool("period_y");
    const int nup = input.getInt("Nup");
    const int ndn = input.getInt("Ndown");
    double flux[2], jx[2], jy[2], mu[2], U[3];
    for (int t = 0; t < 2; ++t) {
      flux[t] = input.getDouble("flux");
      jx[t] = input.getDouble("Jx");
      jy[t] = input.getDouble("Jy");
      mu[t] = input.getDouble("mu");
    }
    U[0] = input.getDouble("U_uu");
    U[1] = input.getDouble("U_dd");
    U[2] = input.getDouble("U_ud");
    require(nx > 0 && ny > 0 && d > 1 && nup >= 0 && ndn >= 0, "bad model parameters in input file",
            kErrArgument);
    space = std::make_unique<ComplexSpinfulBoseSquareSpace>(nx, ny, nup, ndn, d);
    ham = std::make_unique<ComplexSpinfulBoseSquareHamiltonian>(nx, ny, d, flux, jx, jy, U, mu, px, py);
  } else {
    throw Error(kErrArgument, "unknown model driver: " + driver);
  }
}

}  // namespace dmrg
