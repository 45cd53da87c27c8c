// options.hpp -- command-line and runcard options.
//
// The reference parses its options with boost::program_options
// (src/options.cpp:38-204); boost is not available here, so this is a small
// parser with the semantics the reference relies on:
//   * "--name value", "--name=value", "-n value", "-nvalue"; the first
//     positional token is the runcard ("input");
//   * runcard lines "name = value", '#' starts a comment;
//   * a value stored first wins (the command line is stored before the
//     runcard, so it overrides it); defaults apply to options never given;
//   * switches take no value on the command line and "true/false/1/0/yes/no/
//     on/off" in a runcard; implicit values are used when an option is given
//     without a value;
//   * a pass may allow unknown options (the first pass, before the model's
//     options exist) or reject them (the second pass).
#pragma once

#include <map>
#include <string>
#include <vector>

struct OptionSpec {
  enum Kind { UINT, INT, DOUBLE, STRING, SWITCH };
  std::string name;
  char short_name = 0;
  Kind kind = STRING;
  void* target = nullptr;
  bool has_default = false;
  std::string default_value;
  bool has_implicit = false;
  std::string implicit_value;
  std::string help;
};

struct OptionGroup {
  std::string title;
  std::vector<OptionSpec> opts;

  OptionGroup& add_uint(const std::string& name, unsigned* t, const std::string& help);
  OptionGroup& add_int(const std::string& name, int* t, const std::string& help);
  OptionGroup& add_double(const std::string& name, double* t, const std::string& help);
  OptionGroup& add_string(const std::string& name, std::string* t, const std::string& help);
  OptionGroup& add_switch(const std::string& name, bool* t, const std::string& help);
  // modifiers of the last added option
  OptionGroup& short_name(char c);
  OptionGroup& default_value(const std::string& v);
  OptionGroup& implicit_value(const std::string& v);
};

using option_list = std::vector<OptionGroup>;

class VariablesMap {
 public:
  // stores name=value unless name is already stored (first value wins)
  void store(const std::string& name, const std::string& value, const std::string& source);
  bool count(const std::string& name) const { return values_.count(name) != 0; }
  const std::string& get(const std::string& name) const { return values_.at(name).first; }
  const std::map<std::string, std::pair<std::string, std::string>>& all() const { return values_; }
  void clear() { values_.clear(); }

 private:
  std::map<std::string, std::pair<std::string, std::string>> values_;
};

// Parses argv with the given groups; unknown options throw unless
// allow_unknown.  Positional tokens are stored under "input".
void parse_command_line(int argc, char** argv, const std::vector<const OptionGroup*>& groups,
                        bool allow_unknown, VariablesMap& vm);
// Parses a runcard file.
void parse_config_file(const std::string& path, const std::vector<const OptionGroup*>& groups,
                       bool allow_unknown, VariablesMap& vm);
// Writes stored values (or defaults) into the bound variables; throws on
// values that do not convert.
void notify(const std::vector<const OptionGroup*>& groups, const VariablesMap& vm);
// Help text of the groups.
std::string describe(const std::vector<const OptionGroup*>& groups);
