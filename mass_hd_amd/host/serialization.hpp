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
// Numbers are written as operator<< of a default std::ostream writes them
// (the reference's archive streams them): doubles as "%.6g" in the C locale
// (std::num_put with precision 6 and no floatfield), integers in decimal,
// bools as 1/0.  std::to_chars with chars_format::general and precision 6 is
// specified to give printf's "%.6g" character for character, at a fraction
// of a stream's cost, and large arrays are formatted in parallel chunks
// joined in order, so the bytes are those of the one-number-at-a-time
// stream.  A NaN sets the bad-value flag that the writers turn into "nan
// found while writing file.".
#pragma once

#include <algorithm>
#include <array>
#include <charconv>
#include <ostream>
#include <sstream>
#include <string>
#include <thread>
#include <type_traits>
#include <utility>
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

namespace archive_detail {

// one value as operator<< of a default std::ostream writes it; returns
// whether it is a NaN
inline bool put(std::string& s, double v) {
  char b[40];
  const auto r = std::to_chars(b, b + sizeof b, v, std::chars_format::general, 6);
  s.append(b, r.ptr);
  return v != v;
}
// (character types would stream as characters: none is archived)
template <class T>
inline typename std::enable_if<std::is_integral<T>::value && !std::is_same<T, bool>::value &&
                                   sizeof(T) >= sizeof(short),
                               bool>::type
put(std::string& s, T v) {
  char b[24];
  const auto r = std::to_chars(b, b + sizeof b, v);
  s.append(b, r.ptr);
  return false;
}
inline bool put(std::string& s, bool v) {
  s += v ? '1' : '0';
  return false;
}
template <class T, size_t M>
inline bool put(std::string& s, const std::array<T, M>& a) {
  bool nan = false;
  s += "[ ";
  for (size_t i = 0; i < M; ++i) {
    if (i) s += ", ";
    nan |= put(s, a[i]);
  }
  s += " ]";
  return nan;
}

// threads for a list of n values: one below 32k values, else up to 16, at
// least 16k values each
inline unsigned list_workers(size_t n) {
  if (n < (size_t(1) << 15)) return 1;
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  return (unsigned)std::min<size_t>({size_t(16), size_t(hw), n >> 14});
}

// "v0, v1, ..." of a vector (std::vector<bool> included) as consecutive
// pieces, formatted in parallel for long vectors; sets nan if any value is
// a NaN
template <class V>
std::vector<std::string> format_items(const V& v, bool& nan) {
  const size_t n = v.size();
  const unsigned nt = list_workers(n);
  std::vector<std::string> part(nt);
  std::vector<char> bad(nt, 0);
  auto work = [&](unsigned t) {
    const size_t a = n * t / nt, b = n * (t + 1) / nt;
    // built in the thread's own string and moved out at the end: the
    // headers of neighbouring strings share cache lines
    std::string s;
    s.reserve((b - a) * 14 + 8);
    bool any = false;
    for (size_t i = a; i < b; ++i) {
      if (i) s += ", ";
      any |= put(s, static_cast<typename V::value_type>(v[i]));
    }
    part[t] = std::move(s);
    bad[t] = any;
  };
  if (nt == 1) {
    work(0);
  } else {
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < nt; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto& th : pool) th.join();
  }
  for (unsigned t = 0; t < nt; ++t) nan = nan || bad[t];
  return part;
}

// "[ v0, v1, ... ]" in one string
template <class V>
std::string format_list(const V& v, bool& nan) {
  std::string out = "[ ";
  for (const auto& p : format_items(v, nan)) out += p;
  out += " ]";
  return out;
}

}  // namespace archive_detail

// An output stream that keeps what is written as a list of pieces: the long
// lists of a frame are handed over as they were formatted, without copying
// them into one growing buffer, and written out piece by piece.
class piece_buffer : public std::streambuf {
  std::vector<std::string> pieces_;
  std::string cur_;

  void seal() {
    if (!cur_.empty()) {
      pieces_.push_back(std::move(cur_));
      cur_.clear();
    }
  }

 protected:
  int_type overflow(int_type c) override {
    if (!traits_type::eq_int_type(c, traits_type::eof())) cur_ += traits_type::to_char_type(c);
    return traits_type::not_eof(c);
  }
  std::streamsize xsputn(const char* p, std::streamsize n) override {
    cur_.append(p, (size_t)n);
    return n;
  }

 public:
  void take(std::string&& piece) {
    seal();
    pieces_.push_back(std::move(piece));
  }
  void write_to(std::ostream& o) {
    seal();
    for (const auto& p : pieces_) o.write(p.data(), (std::streamsize)p.size());
  }
};

class piece_stream : public std::ostream {
  piece_buffer buf_;

 public:
  piece_stream() : std::ostream(nullptr) { rdbuf(&buf_); }
  piece_buffer& pieces() { return buf_; }
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
    std::string s;
    if (archive_detail::put(s, v)) bad_ = true;
    stream_ << s;
  }
  void element(const std::string& v) { stream_ << '"' << v << '"'; }
  template <class T>
  void element(const std::vector<T>& v) {
    std::vector<std::string> items = archive_detail::format_items(v, bad_);
    stream_ << "[ ";
    if (auto* ps = dynamic_cast<piece_stream*>(&stream_)) {
      for (auto& p : items) ps->pieces().take(std::move(p));
    } else {
      for (const auto& p : items) stream_ << p;
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
