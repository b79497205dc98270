// Input files of the reference's drivers: a one-pass index of the text
// (line heads, bracket depths) and lookups as walks over that index.  The
// grammar and the reference behaviours kept are listed with each function;
// values go through the C++ stream extractors, so "2.5" read as an int is 2
// and an empty value makes the extractor read on into the next line, as in
// the reference (src/util/input.cc, real_tensor_rundmrg.cc:3-113).
#include "input.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <functional>

#include "common.h"

namespace dmrg {

bool InputGroup::echo = true;

namespace {
constexpr size_t kNone = std::string::npos;
bool is_ws(char c) { return std::isspace((unsigned char)c) != 0; }
bool ident_start(char c) { return std::isalpha((unsigned char)c) || c == '_'; }
bool ident_char(char c) { return ident_start(c) || std::isdigit((unsigned char)c); }
}  // namespace

// ---------------------------------------------------------------------------
// The index
// ---------------------------------------------------------------------------
void InputFile::load() {
  std::ifstream f(name_, std::ios::binary);
  require(f.good(), "can't open input file", kErrArgument);
  std::ostringstream ss;
  ss << f.rdbuf();
  text_ = ss.str();
  const size_t n = text_.size();
  // depth_[i]: '{' minus '}' strictly before character i
  depth_.assign(n + 1, 0);
  for (size_t i = 0; i < n; ++i)
    depth_[i + 1] = depth_[i] + (text_[i] == '{') - (text_[i] == '}');
  // a line head is the first non-blank character of the file and the first
  // non-whitespace character after each newline (blank lines add none)
  heads_.clear();
  bool line_start = true;
  for (size_t i = 0; i < n; ++i) {
    if (text_[i] == '\n') { line_start = true; continue; }
    if (is_ws(text_[i])) continue;
    if (line_start) heads_.push_back(i);
    line_start = false;
  }
  is_.str(text_);
  loaded_ = true;
}

const std::string& InputFile::text() {
  if (!loaded_) load();
  return text_;
}

std::istream& InputFile::at(size_t pos) {
  text();
  is_.clear();
  is_.seekg((std::streamoff)std::min(pos, text_.size()));
  return is_;
}

size_t InputFile::next_char(size_t pos) const {
  while (pos < text_.size() && is_ws(text_[pos])) ++pos;
  return pos < text_.size() ? pos : kNone;
}

size_t InputFile::token_start(size_t from) const { return next_char(from); }

std::string InputFile::token(size_t pos) const {
  if (pos >= text_.size()) return std::string();
  if (!ident_start(text_[pos])) return std::string(1, text_[pos]);
  size_t e = pos;
  while (e < text_.size() && ident_char(text_[e])) ++e;
  return text_.substr(pos, e - pos);
}

std::vector<size_t> InputFile::visible_heads(size_t from) const {
  // A scan reads the token at a line head; when it is not the one sought
  // it skips to the end of the line, jumping over every bracketed region
  // that opens on the way (a region's own lines are never heads of the
  // scan).  So a later line head is met exactly when its bracket depth is
  // the lowest depth seen since `from`: a higher one lies inside a region
  // opened after the scan started.  Regions the scan starts inside do not
  // stop it: past their closing bracket it goes on in the enclosing text.
  std::vector<size_t> out;
  if (from == kNone || from >= text_.size()) return out;
  out.push_back(from);
  int floor = depth_[from];
  size_t prev = from;
  auto it = std::upper_bound(heads_.begin(), heads_.end(), from);
  for (; it != heads_.end(); ++it) {
    const size_t h = *it;
    for (size_t i = prev + 1; i <= h; ++i) floor = std::min(floor, depth_[i]);
    if (depth_[h] == floor) out.push_back(h);
    prev = h;
  }
  return out;
}

// ---------------------------------------------------------------------------
// Group and key lookup
// ---------------------------------------------------------------------------
InputGroup::InputGroup(const std::string& file_name, const std::string& group_name)
    : file_(std::make_shared<InputFile>(file_name)), name_(group_name) {}

InputGroup::InputGroup(InputGroup& parent, const std::string& name)
    : file_(parent.file_), parent_(&parent), name_(name) {}

size_t InputGroup::body() {
  // Grammar: a group is its name as a line head, then (after any
  // whitespace, newlines included) '{'; the body starts at the first token
  // after the bracket.  The top-level group is sought from the start of
  // the file, a nested one from its parent's body start: the first line
  // head equal to the name wins.  A name head not followed by '{' is an
  // error ("bracket does not follow group name").  The search is not bounded
  // by the parent's closing bracket (the reference's scan is line-based and
  // runs on).  End of file: not found (the reference's scan does not end
  // at an end of file that has no final newline; here it does).
  const std::string& t = file_->text();
  size_t from = 0;
  if (parent_) {
    from = parent_->body();
    if (from == kNone) return kNone;
  }
  for (size_t h : file_->visible_heads(file_->token_start(from))) {
    const std::string tok = file_->token(h);
    if (tok != name_) continue;
    const size_t b = file_->next_char(h + tok.size());
    require(b != kNone && t[b] == '{', "bracket does not follow group name", kErrArgument);
    const size_t first = file_->next_char(b + 1);
    return first == kNone ? t.size() : first;
  }
  return kNone;
}

size_t InputGroup::value(const std::string& key) {
  // Grammar: an entry is `key = value` with the key as a line head of the
  // group body.  The first head equal to the key decides: when the next
  // non-whitespace character is not '=', the key counts as absent.  A head
  // '}' (the group's closing bracket at the start of a line) ends the
  // search; nested groups are jumped over, so their keys are not seen.  The
  // value starts after '=' and the blanks (spaces, tabs) that follow it.
  const size_t b = body();
  if (b == kNone) return kNone;
  const std::string& t = file_->text();
  for (size_t h : file_->visible_heads(b)) {
    const std::string tok = file_->token(h);
    if (tok == "}") return kNone;
    if (tok != key) continue;
    const size_t eq = file_->next_char(h + tok.size());
    if (eq == kNone || t[eq] != '=') return kNone;
    size_t v = eq + 1;
    while (v < t.size() && (t[v] == ' ' || t[v] == '\t')) ++v;
    return v;
  }
  return kNone;
}

int InputGroup::GotoGroup() {
  const size_t b = body();
  if (b == kNone) return 0;
  file_->at(b);
  return 1;
}

int InputGroup::GotoToken(const std::string& s) {
  const size_t v = value(s);
  if (v == kNone) return 0;
  file_->at(v);
  return 1;
}

void InputGroup::SkipLine() {
  std::istream& is = file();
  std::string rest;
  std::getline(is, rest);
  is >> std::ws;
}

// ---------------------------------------------------------------------------
// Typed values and the reference's echo lines
// ---------------------------------------------------------------------------
namespace {
// Extract a T at the value of `key`: 1 when the key is there and the
// extractor succeeds.
template <typename T>
int extract(InputGroup& g, const std::string& key, T& out) {
  if (!g.GotoToken(key)) return 0;
  T v{};
  if (!(g.file() >> v)) return 0;
  out = v;
  return 1;
}

void say(const char* tag, const std::string& group, const std::string& key, const std::string& val) {
  if (InputGroup::echo) std::printf("%s %s.%s = %s\n", tag, group.c_str(), key.c_str(), val.c_str());
}
std::string fmt_int(long v) { return std::to_string(v); }
std::string fmt_dbl(double v) {
  char b[64];
  std::snprintf(b, sizeof b, "%.2f", v);
  return b;
}
}  // namespace

int InputGroup::GetInt(const std::string& s, int& res, bool hasdf, int df) {
  if (!extract(*this, s, res)) {
    if (hasdf) say("Def", name_, s, fmt_int(df));
    return 0;
  }
  say("Got", name_, s, fmt_int(res));
  return 1;
}

int InputGroup::GetDouble(const std::string& s, double& res, bool hasdf, double df) {
  if (!extract(*this, s, res)) {
    if (hasdf) say("Def", name_, s, fmt_dbl(df));
    return 0;
  }
  say("Got", name_, s, fmt_dbl(res));
  return 1;
}

int InputGroup::GetString(const std::string& s, std::string& res, bool hasdf, const std::string& df) {
  // a key whose value is empty (whitespace right after the blanks that
  // follow '=') is found, with res left as it is
  const size_t v = value(s);
  const std::string& t = file_->text();
  if (v != kNone && v < t.size() && is_ws(t[v])) {
    if (hasdf) say("Def", name_, s, df);
    else say("No ", name_, s, "None");
    return 1;
  }
  if (!extract(*this, s, res)) {
    if (hasdf) say("Def", name_, s, df);
    return 0;
  }
  say("Got", name_, s, res);
  return 1;
}

int InputGroup::GetBool(const std::string& s, bool& res, bool hasdf, bool df) {
  // "true" / "1" and "false" / "0"; any other word is found and leaves res
  std::string w;
  if (!extract(*this, s, w)) {
    if (hasdf) say("Def", name_, s, fmt_int(df));
    return 0;
  }
  if (w == "true" || w == "1") res = true;
  if (w == "false" || w == "0") res = false;
  say("Got", name_, s, fmt_int(res));
  return 1;
}

int InputGroup::getInt(const std::string& s, int def) {
  int r = 0;
  return GetInt(s, r, true, def) ? r : def;
}
double InputGroup::getDouble(const std::string& s, double def) {
  double r = 0.0;
  return GetDouble(s, r, true, def) ? r : def;
}
std::string InputGroup::getString(const std::string& s, const std::string& def) {
  std::string r;
  return GetString(s, r, true, def) ? r : def;
}
bool InputGroup::getBool(const std::string& s, bool def) {
  bool r = false;
  return GetBool(s, r, true, def) ? r : def;
}

namespace {
template <typename T, typename F>
T required(const std::string& key, F&& get) {
  T r{};
  require(get(r) != 0, "there is no item: " + key, kErrArgument);
  return r;
}
}  // namespace

int InputGroup::getInt(const std::string& s) {
  return required<int>(s, [&](int& r) { return GetInt(s, r); });
}
double InputGroup::getDouble(const std::string& s) {
  return required<double>(s, [&](double& r) { return GetDouble(s, r); });
}
std::string InputGroup::getString(const std::string& s) {
  return required<std::string>(s, [&](std::string& r) { return GetString(s, r); });
}
bool InputGroup::getBool(const std::string& s) {
  return required<bool>(s, [&](bool& r) { return GetBool(s, r); });
}

// ---------------------------------------------------------------------------
// The run section and the sweep table
// ---------------------------------------------------------------------------
namespace {

struct SweepRow {
  int maxm = 0, maxb = 0, niter = 0;
  double canpre = 0.0, noise = 0.0;
};

// The table rows after its header line (RealTensorRundmrg::Sweep,
// real_tensor_rundmrg.cc:70-113), five stream-extracted fields per row.  A
// row whose maxm is 0, or is where the extractor fails (it stores 0 there,
// and every later field of a failed stream stays 0 here, where the
// reference leaves it uninitialised), ends the table.  Rows past the last
// one read repeat it, with maxm taken from its maxb column (the reference's
// fill-in).  An empty table is an error here (the reference reads index -1).
std::vector<SweepRow> read_sweeps(InputGroup& table, int nhalf) {
  require(table.GotoGroup() != 0, "Couldn't find table " + table.name(), kErrArgument);
  table.SkipLine();  // the column names
  std::istream& is = table.file();
  std::vector<SweepRow> rows;
  while ((int)rows.size() < nhalf) {
    SweepRow r;
    is >> r.maxm >> r.maxb >> r.niter >> r.canpre >> r.noise;
    if (r.maxm == 0) break;
    rows.push_back(r);
  }
  require(!rows.empty(), "empty sweep table", kErrArgument);
  const SweepRow last = rows.back();
  while ((int)rows.size() < nhalf) {
    SweepRow f = last;
    f.maxm = last.maxb;
    rows.push_back(f);
  }
  return rows;
}

}  // namespace

RunConfig read_run_config(InputGroup& input) {
  // The keys of the RealTensorRundmrg / ComplexTensorRundmrg constructors
  // (real_tensor_rundmrg.cc:3-42, complex_tensor_rundmrg.cc:3-42), in the
  // order they read (and echo) them.
  RunConfig c;
  struct Key { const char* name; std::function<void()> read; };
  const Key keys[] = {
      {"disk_cache", [&] { c.disk_cache = input.getBool("disk_cache"); }},
      {"cache_name", [&] { c.cache_name = input.getString("cache_name"); }},
      {"cache_record", [&] { c.cache_record = input.getBool("cache_record"); }},
      {"record_name", [&] { c.record_name = input.getString("record_name"); }},
      {"cache_resume", [&] { c.cache_resume = input.getBool("cache_resume"); }},
      {"resume_name", [&] { c.resume_name = input.getString("resume_name"); }},
      {"record_process", [&] { c.record_process = input.getBool("record_process"); }},
      {"process_name", [&] { c.process_name = input.getString("process_name"); }},
      {"first_block", [&] { c.first_block = input.getInt("first_block"); }},
      {"first_dim", [&] { c.first_dim = input.getInt("first_dim"); }},
      {"num_sweep", [&] { c.num_sweep = input.getInt("num_sweep"); }},
  };
  for (const Key& k : keys) k.read();
  const int nhalf = 2 * c.num_sweep;
  require(nhalf > 0, "num_sweep must be positive", kErrArgument);
  InputGroup table(input, "sweeps");
  for (const SweepRow& r : read_sweeps(table, nhalf)) {
    c.maxm.push_back(r.maxm);
    c.maxb.push_back(r.maxb);
    c.niter.push_back(r.niter);
    c.canpre.push_back(r.canpre);
    c.noise.push_back(r.noise);
  }
  if (InputGroup::echo) {
    std::printf("Sweeps:\n");
    for (int i = 0; i < nhalf; ++i)
      std::printf("%d Maxm=%d, Maxb=%d, Niter=%d, Canon=%g, Noise=%g\n", i + 1, c.maxm[i], c.maxb[i],
                  c.niter[i], c.canpre[i], c.noise[i]);
  }
  return c;
}

// ---------------------------------------------------------------------------
// The model drivers
// ---------------------------------------------------------------------------
void read_model(InputGroup& input, const std::string& driver, std::unique_ptr<RealTensorSpace>& space,
                std::unique_ptr<RealTensorHamiltonian>& ham) {
  const bool spinless = driver == "real_spinless_bose_square" || driver == "complex_spinless_bose_square";
  const bool spinful = driver == "complex_spinful_bose_square";
  require(spinless || spinful, "unknown model driver: " + driver, kErrArgument);
  // keys in the order each driver's main() reads them
  const int nx = input.getInt("num_site_x");
  const int ny = input.getInt("num_site_y");
  if (spinless) {
    // model/{real,complex}_spinless_bose_square_model.cc:11-22; flux (in
    // units of pi) is passed on and the Hamiltonian does not use it
    const int nb = input.getInt("num_boson");
    const int d = input.getInt("physics_dim", 2);
    const double flux = M_PI * input.getDouble("flux");
    const double jx = input.getDouble("Jx"), jy = input.getDouble("Jy");
    const double U = input.getDouble("U"), mu = input.getDouble("mu");
    const bool px = input.getBool("period_x"), py = input.getBool("period_y");
    require(nx > 0 && ny > 0 && d > 1 && nb >= 0, "bad model parameters in input file", kErrArgument);
    space = std::make_unique<RealSpinlessBoseSquareSpace>(nx, ny, nb, d);
    if (driver[0] == 'r')
      ham = std::make_unique<RealSpinlessBoseSquareHamiltonian>(nx, ny, d, flux, jx, jy, U, mu, px, py);
    else
      ham = std::make_unique<ComplexSpinlessBoseSquareHamiltonian>(nx, ny, d, flux, jx, jy, U, mu, px, py);
    return;
  }
  // model/complex_spinful_bose_square_model.cc:11-36: per-species values
  // are read twice from the same keys (both species alike)
  const int d = input.getInt("physics_dim", 2);
  const bool px = input.getBool("period_x"), py = input.getBool("period_y");
  const int nup = input.getInt("Nup"), ndn = input.getInt("Ndown");
  double flux[2], jx[2], jy[2], mu[2], U[3];
  for (int sp = 0; sp < 2; ++sp) {
    flux[sp] = input.getDouble("flux");
    jx[sp] = input.getDouble("Jx");
    jy[sp] = input.getDouble("Jy");
    mu[sp] = input.getDouble("mu");
  }
  const char* inter[3] = {"U_uu", "U_dd", "U_ud"};
  for (int k = 0; k < 3; ++k) U[k] = input.getDouble(inter[k]);
  require(nx > 0 && ny > 0 && d > 1 && nup >= 0 && ndn >= 0, "bad model parameters in input file",
          kErrArgument);
  space = std::make_unique<ComplexSpinfulBoseSquareSpace>(nx, ny, nup, ndn, d);
  ham = std::make_unique<ComplexSpinfulBoseSquareHamiltonian>(nx, ny, d, flux, jx, jy, U, mu, px, py);
}

}  // namespace dmrg
