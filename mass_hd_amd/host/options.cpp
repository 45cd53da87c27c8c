// options.cpp -- see options.hpp.
#include "options.hpp"

#include <cerrno>
#include <cstdlib>
#include <fstream>
#include <set>
#include <sstream>

#include "error_msg.hpp"

namespace {

OptionSpec& last(OptionGroup& g) {
  if (g.opts.empty()) throw error_msg("option modifier without an option");
  return g.opts.back();
}

const OptionSpec* find_long(const std::vector<const OptionGroup*>& gs, const std::string& n) {
  for (auto* g : gs)
    for (auto& o : g->opts)
      if (o.name == n) return &o;
  return nullptr;
}

const OptionSpec* find_short(const std::vector<const OptionGroup*>& gs, char c) {
  for (auto* g : gs)
    for (auto& o : g->opts)
      if (o.short_name && o.short_name == c) return &o;
  return nullptr;
}

std::string trim(const std::string& s) {
  const auto b = s.find_first_not_of(" \t\r\n");
  if (b == std::string::npos) return "";
  const auto e = s.find_last_not_of(" \t\r\n");
  return s.substr(b, e - b + 1);
}

bool parse_bool(const std::string& v, bool& out) {
  static const char* t[] = {"true", "1", "yes", "on"};
  static const char* f[] = {"false", "0", "no", "off"};
  for (auto* s : t)
    if (v == s) return out = true, true;
  for (auto* s : f)
    if (v == s) return out = false, true;
  return false;
}

}  // namespace

OptionGroup& OptionGroup::add_uint(const std::string& n, unsigned* t, const std::string& h) {
  OptionSpec o;
  o.name = n;
  o.kind = OptionSpec::UINT;
  o.target = t;
  o.help = h;
  opts.push_back(o);
  return *this;
}
OptionGroup& OptionGroup::add_int(const std::string& n, int* t, const std::string& h) {
  OptionSpec o;
  o.name = n;
  o.kind = OptionSpec::INT;
  o.target = t;
  o.help = h;
  opts.push_back(o);
  return *this;
}
OptionGroup& OptionGroup::add_double(const std::string& n, double* t, const std::string& h) {
  OptionSpec o;
  o.name = n;
  o.kind = OptionSpec::DOUBLE;
  o.target = t;
  o.help = h;
  opts.push_back(o);
  return *this;
}
OptionGroup& OptionGroup::add_string(const std::string& n, std::string* t, const std::string& h) {
  OptionSpec o;
  o.name = n;
  o.kind = OptionSpec::STRING;
  o.target = t;
  o.help = h;
  opts.push_back(o);
  return *this;
}
OptionGroup& OptionGroup::add_switch(const std::string& n, bool* t, const std::string& h) {
  OptionSpec o;
  o.name = n;
  o.kind = OptionSpec::SWITCH;
  o.target = t;
  o.help = h;
  o.has_default = true;
  o.default_value = "false";
  o.has_implicit = true;
  o.implicit_value = "true";
  opts.push_back(o);
  return *this;
}
OptionGroup& OptionGroup::short_name(char c) {
  last(*this).short_name = c;
  return *this;
}
OptionGroup& OptionGroup::default_value(const std::string& v) {
  last(*this).has_default = true;
  last(*this).default_value = v;
  return *this;
}
OptionGroup& OptionGroup::implicit_value(const std::string& v) {
  last(*this).has_implicit = true;
  last(*this).implicit_value = v;
  return *this;
}

void VariablesMap::store(const std::string& name, const std::string& value,
                         const std::string& source) {
  // the first stored value wins; later sources are ignored
  if (values_.find(name) == values_.end()) values_[name] = {value, source};
}

void parse_command_line(int argc, char** argv, const std::vector<const OptionGroup*>& gs,
                        bool allow_unknown, VariablesMap& vm) {
  const std::string src = "command line";
  std::set<std::string> seen;  // one source may give an option only once
  bool have_input = false;
  for (int i = 1; i < argc; ++i) {
    std::string tok = argv[i];
    const OptionSpec* o = nullptr;
    std::string value;
    bool has_value = false;
    if (tok.size() > 2 && tok[0] == '-' && tok[1] == '-') {
      std::string name = tok.substr(2);
      const auto eq = name.find('=');
      if (eq != std::string::npos) {
        value = name.substr(eq + 1);
        name = name.substr(0, eq);
        has_value = true;
      }
      o = find_long(gs, name);
      if (!o) {
        if (allow_unknown) continue;
        throw error_msg("unrecognised option '--", name, "'");
      }
    } else if (tok.size() > 1 && tok[0] == '-' && !(tok[1] >= '0' && tok[1] <= '9')) {
      o = find_short(gs, tok[1]);
      if (!o) {
        if (allow_unknown) continue;
        throw error_msg("unrecognised option '", tok, "'");
      }
      if (tok.size() > 2) {
        value = tok.substr(2);
        has_value = true;
      }
    } else {
      if (!have_input) {
        vm.store("input", tok, src);
        have_input = true;
        continue;
      }
      if (allow_unknown) continue;
      throw error_msg("too many positional options have been specified on the command line");
    }
    if (!has_value) {
      if (o->kind == OptionSpec::SWITCH || o->has_implicit) {
        value = o->kind == OptionSpec::SWITCH ? "true" : o->implicit_value;
      } else if (i + 1 < argc) {
        value = argv[++i];
      } else {
        throw error_msg("the required argument for option '--", o->name, "' is missing");
      }
    } else if (o->kind == OptionSpec::SWITCH) {
      throw error_msg("option '--", o->name, "' does not take any arguments");
    }
    if (!seen.insert(o->name).second)
      throw error_msg("option '--", o->name, "' cannot be specified more than once");
    vm.store(o->name, value, src);
  }
}

void parse_config_file(const std::string& path, const std::vector<const OptionGroup*>& gs,
                       bool allow_unknown, VariablesMap& vm) {
  std::ifstream in(path);
  if (!in.good()) throw error_msg("can not open runcard file ", path);
  const std::string src = path;
  std::set<std::string> seen;
  std::string line;
  unsigned lineno = 0;
  while (std::getline(in, line)) {
    ++lineno;
    const auto hash = line.find('#');
    if (hash != std::string::npos) line = line.substr(0, hash);
    line = trim(line);
    if (line.empty()) continue;
    const auto eq = line.find('=');
    if (eq == std::string::npos)
      throw error_msg("invalid line ", lineno, " in runcard ", path, ": '", line, "'");
    const std::string name = trim(line.substr(0, eq));
    const std::string value = trim(line.substr(eq + 1));
    const OptionSpec* o = find_long(gs, name);
    if (!o) {
      if (allow_unknown) continue;
      throw error_msg("unrecognised option '", name, "'");
    }
    if (!seen.insert(name).second)
      throw error_msg("option '--", name, "' cannot be specified more than once");
    vm.store(name, value, src);
  }
}

void notify(const std::vector<const OptionGroup*>& gs, const VariablesMap& vm) {
  for (auto* g : gs)
    for (auto& o : g->opts) {
      std::string v;
      if (vm.count(o.name))
        v = vm.get(o.name);
      else if (o.has_default)
        v = o.default_value;
      else
        continue;
      const auto bad = [&]() {
        return error_msg("the argument ('", v, "') for option '--", o.name, "' is invalid");
      };
      switch (o.kind) {
        case OptionSpec::UINT: {
          if (v.empty() || v[0] == '-') throw bad();
          errno = 0;
          char* end = nullptr;
          const unsigned long x = std::strtoul(v.c_str(), &end, 10);
          if (errno || *end || x > 0xffffffffUL) throw bad();
          *static_cast<unsigned*>(o.target) = (unsigned)x;
          break;
        }
        case OptionSpec::INT: {
          errno = 0;
          char* end = nullptr;
          const long x = std::strtol(v.c_str(), &end, 10);
          if (v.empty() || errno || *end || x < -2147483647L - 1 || x > 2147483647L) throw bad();
          *static_cast<int*>(o.target) = (int)x;
          break;
        }
        case OptionSpec::DOUBLE: {
          errno = 0;
          char* end = nullptr;
          const double x = std::strtod(v.c_str(), &end);
          if (v.empty() || errno || *end) throw bad();
          *static_cast<double*>(o.target) = x;
          break;
        }
        case OptionSpec::STRING:
          *static_cast<std::string*>(o.target) = v;
          break;
        case OptionSpec::SWITCH: {
          bool b = false;
          if (!parse_bool(v, b)) throw bad();
          *static_cast<bool*>(o.target) = b;
          break;
        }
      }
    }
}

std::string describe(const std::vector<const OptionGroup*>& gs) {
  std::ostringstream s;
  for (auto* g : gs) {
    s << g->title << ":\n";
    for (auto& o : g->opts) {
      std::string lhs = "  ";
      if (o.short_name) lhs += std::string("-") + o.short_name + " [ --" + o.name + " ]";
      else lhs += "--" + o.name;
      if (o.kind != OptionSpec::SWITCH) lhs += " arg";
      if (o.has_default && o.kind != OptionSpec::SWITCH) lhs += " (=" + o.default_value + ")";
      if (lhs.size() < 34) lhs += std::string(34 - lhs.size(), ' ');
      s << lhs << " " << o.help << "\n";
    }
    s << "\n";
  }
  return s.str();
}
