// Input files of the reference's model drivers.  The public classes keep the
// reference's interface (InputFile / InputGroup with GotoGroup, GotoToken,
// SkipLine, Get* / get*, src/util/input.h:7-89), its key names and its
// "Got / Def / No" echo lines; the sweep table and the drivers' keys are
// those of RealTensorRundmrg (real_tensor_rundmrg.cc:3-113) and
// model/*_model.cc.
//
// How a file is read here: the text is loaded once and indexed (the
// position of every line head, the bracket depth before every character).
// A lookup is then a walk over the line heads a scan from the group body
// can see, and the value is extracted with the C++ stream extractors from
// the position the walk returns.
#pragma once

#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "model.h"

namespace dmrg {

class InputFile {
 public:
  explicit InputFile(std::string name) : name_(std::move(name)) {}
  const std::string& filename() const { return name_; }
  // the text and its index, loaded on first use ("can't open input file")
  const std::string& text();
  // the stream the Get* extractors read, positioned at `pos`
  std::istream& at(size_t pos);
  std::istream& file() { return is_; }
  // First token position at or after `from` (whitespace skipped), or npos.
  size_t token_start(size_t from) const;
  // The token at `pos`: an identifier [A-Za-z_][A-Za-z0-9_]*, or one other
  // character.
  std::string token(size_t pos) const;
  // The line heads a line-by-line scan starting at token position `from`
  // meets, in order: `from` itself, then the first token of every later
  // line that is not inside a bracket opened after `from`.
  std::vector<size_t> visible_heads(size_t from) const;
  // first non-whitespace position at or after `pos` (npos at the end)
  size_t next_char(size_t pos) const;

 private:
  void load();
  std::string name_, text_;
  bool loaded_ = false;
  std::vector<int> depth_;     // bracket depth before each character
  std::vector<size_t> heads_;  // first token of each line, in order
  std::istringstream is_;
};

class InputGroup {
 public:
  InputGroup(const std::string& file_name, const std::string& group_name);
  InputGroup(InputGroup& parent, const std::string& name);
  const std::string& name() const { return name_; }
  std::istream& file() { return file_->file(); }

  // Position the stream at the first token of the group body (1), or 0
  // when the group is not found.
  int GotoGroup();
  // Position the stream at the value of `key = value` in the group (1), or 0.
  int GotoToken(const std::string& s);
  // Skip the rest of the current line and the whitespace after it.
  void SkipLine();
  int GetInt(const std::string& s, int& res, bool hasdf = false, int df = 0);
  int GetDouble(const std::string& s, double& res, bool hasdf = false, double df = 0.0);
  int GetString(const std::string& s, std::string& res, bool hasdf = false, const std::string& df = "");
  int GetBool(const std::string& s, bool& res, bool hasdf = false, bool df = false);
  // with a default: the default when the key is absent
  int getInt(const std::string& s, int def);
  double getDouble(const std::string& s, double def);
  std::string getString(const std::string& s, const std::string& def);
  bool getBool(const std::string& s, bool def);
  // required: "there is no item: <key>" when absent
  int getInt(const std::string& s);
  double getDouble(const std::string& s);
  std::string getString(const std::string& s);
  bool getBool(const std::string& s);

  // print the reference's "Got/Def <group>.<key> = value" lines
  static bool echo;

 private:
  // body start of this group (npos when absent)
  size_t body();
  // value position of `key` in this group (npos when absent)
  size_t value(const std::string& key);
  std::shared_ptr<InputFile> file_;
  InputGroup* parent_ = nullptr;
  std::string name_;
};

// The run section read by the RealTensorRundmrg / ComplexTensorRundmrg
// constructors (real_tensor_rundmrg.cc:3-42, :70-113): cache settings,
// first_block / first_dim and the sweep table, one row per half sweep after
// the reference's fill-in rule; `noise` keeps the table's exponents.
struct RunConfig {
  bool disk_cache = false, cache_record = false, cache_resume = false, record_process = false;
  std::string cache_name, record_name, resume_name, process_name;
  int first_block = 0, first_dim = 0, num_sweep = 0;
  std::vector<int> maxm, maxb, niter;
  std::vector<double> canpre, noise;
};
RunConfig read_run_config(InputGroup& input);

// The model a reference driver builds from its input group:
// "real_spinless_bose_square" (model/real_spinless_bose_square_model.cc:11-22),
// "complex_spinless_bose_square" (model/complex_spinless_bose_square_model.cc:11-22),
// "complex_spinful_bose_square" (model/complex_spinful_bose_square_model.cc:11-36).
void read_model(InputGroup& input, const std::string& driver, std::unique_ptr<RealTensorSpace>& space,
                std::unique_ptr<RealTensorHamiltonian>& ham);

}  // namespace dmrg
