// serialization.hpp -- JSON archive in the reference's on-disk format.
//
// Reproduces the output of the reference's oarchive
// (src/serialization.hpp:187-361, src/serialization.cpp:9-104) so that the
// frame and parameter files stay readable by its post-processing
// (POST/massPy/archive.py):
//
//   {
//     "id" : "frame",
//     "version" : 1,
//     "data" : {
//       "phi" : {
//         "type" : "array(double)",
//         "value" : [ 0.5, 0.51, ... ]
//       }, ...
//     }
//   }
//
// Numbers go through operator<< of a default std::stringstream (six
// significant digits), and a NaN sets the bad-value flag that the writers turn
// into "nan found while writing file.".
#pragma once

#include <array>
#include <ostream>
#include <sstream>
#include <string>
#include <vector>

// type names of the archive (src/serialization.hpp:158-177)
template <class T>
struct archive_type_name;
template <>
struct archive_type_name<double> {
  static std::string name() { return "double"; }
};
template <>
struct archive_type_name<unsigned> {
  static std::string name() { return "unsigned"; }
};
template <>
struct archive_type_name<int> {
  static std::string name() { return "int"; }
};
template <>
struct archive_type_name<bool> {
  static std::string name() { return "bool"; }
};
template <>
struct archive_type_name<std::string> {
  static std::string name() { return "string"; }
};
template <class T>
struct archive_type_name<std::vector<T>> {
  static std::string name() { return "array(" + archive_type_name<T>::name() + ")"; }
};
template <class T, size_t M>
struct archive_type_name<std::array<T, M>> {
  static std::string name() { return "array(" + archive_type_name<T>::name() + ")"; }
};

class oarchive {
  std::ostream& stream_;
  unsigned level_ = 0;
  bool first_ = true;
  bool bad_ = false;
  static constexpr unsigned kPad = 2;

  void open_group() {
    stream_ << '{';
    ++level_;
    first_ = true;
  }
  void close_group() {
    --level_;
    stream_ << '\n' << std::string(kPad * level_, ' ') << '}';
    first_ = false;
  }
  void new_line() {
    if (!first_) stream_ << ',';
    first_ = false;
    stream_ << '\n' << std::string(kPad * level_, ' ');
  }
  void add_key(const std::string& key) {
    new_line();
    stream_ << '"' << key << "\" : ";
  }

  template <class T>
  void element(const T& v) {
    if (v != v) bad_ = true;
    std::stringstream ss;
    ss << v;
    stream_ << ss.str();
  }
  void element(const std::string& v) { stream_ << '"' << v << '"'; }
  template <class T>
  void element(const std::vector<T>& v) {
    stream_ << "[ ";
    bool f = true;
    for (const auto& x : v) {
      if (!f) stream_ << ", ";
      f = false;
      element(x);
    }
    stream_ << " ]";
  }
  template <class T, size_t M>
  void element(const std::array<T, M>& v) {
    stream_ << "[ ";
    for (size_t i = 0; i < M; ++i) {
      if (i) stream_ << ", ";
      element(v[i]);
    }
    stream_ << " ]";
  }

 public:
  oarchive(std::ostream& s, const std::string& id, unsigned version) : stream_(s) {
    open_group();
    add_key("id");
    element(id);
    add_key("version");
    element(version);
    add_key("data");
    open_group();
  }
  ~oarchive() {
    close_group();
    stream_ << '\n' << '}';
  }
  oarchive(const oarchive&) = delete;
  oarchive& operator=(const oarchive&) = delete;

  // ar.add("name", value) is the reference's  ar & auto_name(value)
  template <class T>
  oarchive& add(const std::string& name, const T& value) {
    add_key(name);
    open_group();
    add_key("type");
    element(archive_type_name<T>::name());
    add_key("value");
    element(value);
    close_group();
    return *this;
  }

  bool bad_value() const { return bad_; }
};
