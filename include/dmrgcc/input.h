// Drop-in InputGroup (src/util/input.h:41-89) over the library's reader:
// values are read, echoed ("Got / Def input.<key> = value") and defaulted
// exactly as the reference's InputGroup does (dmrgcc_input_get,
// dmrgcc_input_string); a missing required key prints "there is no item:
// <key>" and exits.
#ifndef DMRGCC_DROPIN_INPUT_H_
#define DMRGCC_DROPIN_INPUT_H_

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "device.h"

class InputGroup {
 public:
  InputGroup(std::string file_name, std::string group_name)
      : file_(std::move(file_name)), name_(std::move(group_name)) {}
  const std::string& name() const { return name_; }
  const std::string& file_name() const { return file_; }

  int getInt(const std::string& s, int def) { return (int)get(s, 0, true, def); }
  double getDouble(const std::string& s, double def) { return get(s, 1, true, def); }
  bool getBool(const std::string& s, bool def) { return get(s, 2, true, def ? 1.0 : 0.0) != 0.0; }
  int getInt(const std::string& s) { return (int)get(s, 0, false, 0.0); }
  double getDouble(const std::string& s) { return get(s, 1, false, 0.0); }
  bool getBool(const std::string& s) { return get(s, 2, false, 0.0) != 0.0; }
  std::string getString(const std::string& s, const std::string& def) {
    std::string v;
    return string_value(s, v) ? v : def;
  }
  std::string getString(const std::string& s) {
    std::string v;
    if (!string_value(s, v)) fail(s);
    return v;
  }

 private:
  double get(const std::string& key, int kind, bool has_default, double def) {
    double v = 0.0;
    int found = 0;
    dmrgcc_dropin::check(dmrgcc_input_get(file_.c_str(), name_.c_str(), key.c_str(), kind, 1, has_default, def,
                                           &v, &found));
    if (!found && !has_default) fail(key);
    return v;
  }
  bool string_value(const std::string& key, std::string& v) {
    std::vector<char> buf(4096);
    int found = 0;
    dmrgcc_dropin::check(dmrgcc_input_string(file_.c_str(), key.c_str(), buf.data(), (int)buf.size(), &found));
    if (found) {
      v = buf.data();
      std::printf("Got %s.%s = %s\n", name_.c_str(), key.c_str(), v.c_str());
    }
    return found != 0;
  }
  [[noreturn]] static void fail(const std::string& key) {
    std::printf("there is no item: %s\n", key.c_str());
    std::exit(-1);
  }
  std::string file_, name_;
};

#endif  // DMRGCC_DROPIN_INPUT_H_
