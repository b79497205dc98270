// Input files of the reference's model drivers: InputFile / InputGroup
// (src/util/input.h:7-89, src/util/input.cc:1-331), the sweep table of
// RealTensorRundmrg (real_tensor_rundmrg.cc:3-113) and the model drivers'
// keys (model/*_model.cc).
//
// Format: named groups `name { ... }` holding `key = value` lines; a nested
// group (the sweep table) is a header line followed by whitespace-separated
// rows.  Every lookup rescans from the top of the file, skipping lines and
// bracketed sub-groups that do not match, as the reference does; values are
// read with the C++ stream extractors, so their parsing is the same.
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
  void open();  // (re)start at the top of the file
  std::istream& file() { return is_; }
  const std::string& filename() const { return name_; }

 private:
  std::string name_, text_;
  bool loaded_ = false;
  std::istringstream is_;
};

class InputGroup {
 public:
  InputGroup(const std::string& file_name, const std::string& group_name);
  InputGroup(InputGroup& parent, const std::string& name);
  const std::string& name() const { return name_; }
  std::istream& file() { return file_->file(); }

  int GotoGroup();
  int GotoToken(const std::string& s);
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
