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
#include <condition_variable>
#include <fstream>
#include <functional>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <ostream>
#include <sstream>
#include <string>
#include <string_view>
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
// 10^k for |k| <= 300, each the double nearest the exact power (strtod
// rounds correctly)
struct Pow10Table {
  static constexpr int kSpan = 300;
  double p[2 * kSpan + 1];
  Pow10Table() {
    char b[16];
    for (int k = -kSpan; k <= kSpan; ++k) {
      std::snprintf(b, sizeof b, "1e%d", k);
      p[k + kSpan] = std::strtod(b, nullptr);
    }
  }
  double operator()(int k) const { return p[k + kSpan]; }
};
inline const Pow10Table kPow10;

// "%.6g" of v without a library call where that is safe: q = v * 10^(5-e)
// with e = floor(log10|v|) carries a relative error below 2.3e-16 (one
// rounding in the power, one in the product), i.e. below 2.3e-10 at q < 1e6,
// so when the fraction of q is more than 1e-9 away from one half, rounding q
// to the nearest integer gives the six correctly rounded digits.  Writes at
// most 13 characters at b and returns their end, or nullptr (having written
// nothing) for zeros, infinities, NaNs, magnitudes outside [1e-290, 1e290]
// and near-ties, which take std::to_chars.
inline char* put_g6(char* b, double v) {
  const double a = std::fabs(v);
  if (!(a >= 1e-290 && a <= 1e290)) return nullptr;
  uint64_t bits;
  std::memcpy(&bits, &a, sizeof bits);
  const int e2 = (int)(bits >> 52) - 1023;  // a in [2^e2, 2^(e2+1)): a is normal here
  // floor(e2 log10 2) = (e2 * 78913) >> 18 for |e2| <= 1100: floor(log10 a)
  // or one less
  int e = (e2 * 78913) >> 18;
  double x = a * kPow10(5 - e);
  if (x >= 1e6) {
    ++e;
    x = a * kPow10(5 - e);
  }
  const long t = (long)x;  // x > 0: truncation is floor
  const double f = x - (double)t;
  if (std::fabs(f - 0.5) < 1e-9) return nullptr;
  int q = (int)t + (f > 0.5 ? 1 : 0);
  if (q == 1000000) {
    q = 100000;
    ++e;
  }
  if (q < 100000 || q > 999999) return nullptr;
  // the six digits as three independent pairs
  static constexpr char kPairs[] =
      "00010203040506070809101112131415161718192021222324252627282930313233343536373839"
      "40414243444546474849505152535455565758596061626364656667686970717273747576777879"
      "8081828384858687888990919293949596979899";
  const int p0 = q / 10000, p1 = (q / 100) % 100, p2 = q % 100;
  const char d[6] = {kPairs[2 * p0], kPairs[2 * p0 + 1], kPairs[2 * p1],
                     kPairs[2 * p1 + 1], kPairs[2 * p2], kPairs[2 * p2 + 1]};
  int n = 6;  // significant digits once trailing zeros are dropped
  while (n > 1 && d[n - 1] == '0') --n;
  char* w = b;
  if (v < 0) *w++ = '-';
  if (e < -4 || e >= 6) {
    *w++ = d[0];
    if (n > 1) {
      *w++ = '.';
      for (int i = 1; i < n; ++i) *w++ = d[i];
    }
    *w++ = 'e';
    int ae = e;
    if (e < 0) {
      *w++ = '-';
      ae = -e;
    } else {
      *w++ = '+';
    }
    if (ae >= 100) *w++ = (char)('0' + ae / 100);
    *w++ = (char)('0' + (ae / 10) % 10);
    *w++ = (char)('0' + ae % 10);
  } else if (e >= 0) {
    for (int i = 0; i <= e; ++i) *w++ = d[i];
    if (n > e + 1) {
      *w++ = '.';
      for (int i = e + 1; i < n; ++i) *w++ = d[i];
    }
  } else {
    *w++ = '0';
    *w++ = '.';
    for (int i = 0; i < -e - 1; ++i) *w++ = '0';
    for (int i = 0; i < n; ++i) *w++ = d[i];
  }
  return w;
}

// one double as operator<< writes it, at most 16 characters; sets nan
inline char* put_raw(char* b, double v, bool& nan) {
  if (char* w = put_g6(b, v)) return w;
  nan = nan || v != v;
  return std::to_chars(b, b + 16, v, std::chars_format::general, 6).ptr;
}

inline bool put(std::string& s, double v) {
  char b[16];
  bool nan = false;
  s.append(b, put_raw(b, v, nan));
  return nan;
}

// "[ v0, ..., v(M-1) ]" of a node's values, at most 20 M + 4 characters
template <size_t M>
inline char* put_raw(char* w, const std::array<double, M>& a, bool& nan) {
  *w++ = '[';
  *w++ = ' ';
  for (size_t i = 0; i < M; ++i) {
    if (i) {
      *w++ = ',';
      *w++ = ' ';
    }
    w = put_raw(w, a[i], nan);
  }
  *w++ = ' ';
  *w++ = ']';
  return w;
}

// value types written straight into a buffer, with a bound on their length
template <class T>
struct raw_width {
  static constexpr size_t value = 0;
};
template <>
struct raw_width<double> {
  static constexpr size_t value = 16;
};
template <size_t M>
struct raw_width<std::array<double, M>> {
  static constexpr size_t value = 20 * M + 4;
};
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

// Buffers of the long lists, kept for the next frame: a frame's lists have
// the sizes of the last frame's, and memory the process has already touched
// costs no page faults (at ~70 MB per 500^2 frame those took as long as the
// formatting).  At most 1 GiB is kept.
class buffer_pool {
  std::mutex mu_;
  std::vector<std::pair<size_t, std::unique_ptr<char[]>>> free_;
  size_t bytes_ = 0;

 public:
  // a buffer of at least n bytes and its capacity
  std::unique_ptr<char[]> get(size_t n, size_t& cap) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      size_t best = free_.size();
      for (size_t i = 0; i < free_.size(); ++i)
        if (free_[i].first >= n && (best == free_.size() || free_[i].first < free_[best].first))
          best = i;
      if (best < free_.size()) {
        cap = free_[best].first;
        std::unique_ptr<char[]> b = std::move(free_[best].second);
        free_.erase(free_.begin() + (std::ptrdiff_t)best);
        bytes_ -= cap;
        return b;
      }
    }
    cap = n;
    return std::unique_ptr<char[]>(new char[n]);
  }
  void put(std::unique_ptr<char[]> b, size_t cap) {
    if (!b) return;
    std::lock_guard<std::mutex> lk(mu_);
    if (bytes_ + cap > (size_t(1) << 30)) return;  // freed here
    bytes_ += cap;
    free_.emplace_back(cap, std::move(b));
  }
};
inline buffer_pool& list_buffers() {
  static buffer_pool pool;
  return pool;
}

// A run of formatted text in storage of its own (not zero-filled first:
// the long lists are written straight into it), handed back to the buffer
// pool when it is dropped.
struct piece {
  std::unique_ptr<char[]> data;
  size_t size = 0, cap = 0;
  piece() = default;
  piece(const char* p, size_t n) : data(new char[n ? n : 1]), size(n), cap(n ? n : 1) {
    if (n) std::memcpy(data.get(), p, n);
  }
  piece(piece&&) = default;
  piece& operator=(piece&& o) {
    if (this != &o) {
      release();
      data = std::move(o.data);
      size = o.size;
      cap = o.cap;
    }
    return *this;
  }
  ~piece() { release(); }
  void release() {
    if (data && cap >= (size_t(1) << 16)) list_buffers().put(std::move(data), cap);
    data.reset();
  }
  std::string_view view() const { return std::string_view(data.get(), size); }
};

// Worker threads for the long lists, started on first use and kept until
// the process ends: a frame formats ~20 lists, and starting and joining a
// thread per chunk of each took longer than the formatting itself.
class worker_pool {
  std::mutex mu_;
  std::condition_variable wake_, done_;
  std::vector<std::thread> threads_;
  const std::function<void(unsigned)>* job_ = nullptr;
  unsigned njobs_ = 0, next_ = 0, busy_ = 0;
  uint64_t generation_ = 0;
  bool stop_ = false;

  // claims and runs chunks of the current job until none is left; mu_ held
  // on entry and exit
  void drain(std::unique_lock<std::mutex>& lk) {
    while (next_ < njobs_) {
      const unsigned i = next_++;
      ++busy_;
      const std::function<void(unsigned)>* job = job_;
      lk.unlock();
      (*job)(i);
      lk.lock();
      --busy_;
    }
    if (busy_ == 0) done_.notify_all();
  }
  void loop() {
    std::unique_lock<std::mutex> lk(mu_);
    uint64_t seen = generation_;
    for (;;) {
      wake_.wait(lk, [&] { return stop_ || generation_ != seen; });
      if (stop_) return;
      seen = generation_;
      drain(lk);
    }
  }

 public:
  ~worker_pool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    wake_.notify_all();
    for (auto& t : threads_) t.join();
  }
  // runs f(0) ... f(n-1), the calling thread taking part; returns when all
  // have finished.  One job at a time (the archive is written by one thread).
  void run(unsigned n, const std::function<void(unsigned)>& f) {
    std::unique_lock<std::mutex> lk(mu_);
    while (threads_.size() + 1 < n) threads_.emplace_back([this] { loop(); });
    job_ = &f;
    njobs_ = n;
    next_ = 0;
    busy_ = 0;
    ++generation_;
    wake_.notify_all();
    drain(lk);
    done_.wait(lk, [&] { return next_ >= njobs_ && busy_ == 0; });
    job_ = nullptr;
    njobs_ = 0;
  }
};
inline worker_pool& list_workers_pool() {
  static worker_pool pool;
  return pool;
}

// "v0, v1, ..." of a vector (std::vector<bool> included) as consecutive
// pieces, formatted in parallel for long vectors; sets nan if any value is
// a NaN
template <class V>
std::vector<piece> format_items(const V& v, bool& nan) {
  const size_t n = v.size();
  const unsigned nt = list_workers(n);
  std::vector<piece> part(nt);
  std::vector<char> bad(nt, 0);
  auto work = [&](unsigned t) {
    const size_t a = n * t / nt, b = n * (t + 1) / nt;
    bool any = false;
    constexpr size_t width = raw_width<typename V::value_type>::value;
    piece out;
    if constexpr (width > 0) {
      // doubles (and nodes of them) straight into the piece: ", " and at
      // most `width` characters each
      out.data = list_buffers().get((b - a) * (width + 2) + 2, out.cap);
      char* w = out.data.get();
      for (size_t i = a; i < b; ++i) {
        if (i) {
          *w++ = ',';
          *w++ = ' ';
        }
        w = put_raw(w, v[i], any);
      }
      out.size = (size_t)(w - out.data.get());
    } else {
      std::string s;
      s.reserve((b - a) * 14 + 8);
      for (size_t i = a; i < b; ++i) {
        if (i) s += ", ";
        any |= put(s, static_cast<typename V::value_type>(v[i]));
      }
      out = piece(s.data(), s.size());
    }
    // moved out at the end: the headers of neighbouring pieces share cache
    // lines
    part[t] = std::move(out);
    bad[t] = any;
  };
  if (nt == 1) {
    work(0);
  } else {
    const std::function<void(unsigned)> job = work;
    list_workers_pool().run(nt, job);
  }
  for (unsigned t = 0; t < nt; ++t) nan = nan || bad[t];
  return part;
}

// "[ v0, v1, ... ]" in one string
template <class V>
std::string format_list(const V& v, bool& nan) {
  std::string out = "[ ";
  for (const auto& p : format_items(v, nan)) out += p.view();
  out += " ]";
  return out;
}

}  // namespace archive_detail

// An output stream that keeps what is written as a list of pieces: the long
// lists of a frame are handed over as they were formatted, without copying
// them into one growing buffer, and written out piece by piece.
class piece_buffer : public std::streambuf {
  std::vector<archive_detail::piece> pieces_;
  std::string cur_;
  // streaming to a file: an I/O thread writes the pieces in order as they
  // come, while the lists after them are formatted
  std::ofstream* file_ = nullptr;
  std::thread io_;
  std::mutex mu_;
  std::condition_variable cv_;
  size_t written_ = 0;  // pieces the I/O thread has taken
  bool closing_ = false;

  void push(archive_detail::piece&& p) {
    if (!file_) {
      pieces_.push_back(std::move(p));
      return;
    }
    {
      std::lock_guard<std::mutex> lk(mu_);
      pieces_.push_back(std::move(p));
    }
    cv_.notify_one();
  }
  void seal() {
    if (!cur_.empty()) {
      push(archive_detail::piece(cur_.data(), cur_.size()));
      cur_.clear();
    }
  }
  void io_loop() {
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
      cv_.wait(lk, [&] { return closing_ || written_ < pieces_.size(); });
      if (written_ == pieces_.size()) return;  // closing, all written
      archive_detail::piece p = std::move(pieces_[written_++]);
      lk.unlock();
      file_->write(p.data.get(), (std::streamsize)p.size);
      p.release();  // back to the buffer pool
      lk.lock();
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
  ~piece_buffer() { finish(); }
  void take(archive_detail::piece&& p) {
    seal();
    push(std::move(p));
  }
  void write_to(std::ostream& o) {
    seal();
    for (const auto& p : pieces_) o.write(p.data.get(), (std::streamsize)p.size);
  }
  // from now on the pieces go to the file, in order, from an I/O thread
  // (call before anything is written)
  void stream_to(std::ofstream& f) {
    file_ = &f;
    io_ = std::thread([this] { io_loop(); });
  }
  // everything written: the I/O thread has written every piece
  void finish() {
    if (!file_) return;
    seal();
    {
      std::lock_guard<std::mutex> lk(mu_);
      closing_ = true;
    }
    cv_.notify_one();
    if (io_.joinable()) io_.join();
    file_ = nullptr;
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
    std::vector<archive_detail::piece> items = archive_detail::format_items(v, bad_);
    stream_ << "[ ";
    if (auto* ps = dynamic_cast<piece_stream*>(&stream_)) {
      for (auto& p : items) ps->pieces().take(std::move(p));
    } else {
      for (const auto& p : items) stream_ << p.view();
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

namespace archive_detail {
inline bool any_nan(double v) { return v != v; }
template <size_t M>
inline bool any_nan(const std::array<double, M>& a) {
  for (double v : a)
    if (v != v) return true;
  return false;
}
template <class T>
inline bool any_nan(const std::vector<T>& v) {
  if constexpr (std::is_same<T, double>::value || std::is_same<T, std::array<double, 9>>::value) {
    for (const auto& x : v)
      if (any_nan(x)) return true;
  }
  return false;
}
template <class T>
inline bool any_nan(const T&) {
  return false;
}
}  // namespace archive_detail

// The values of a frame, kept as they are handed over (downloaded fields
// are moved in) and written out later, in the same order, through an
// oarchive: the host program formats and writes a frame in a thread of its
// own while the GPU runs the next steps.
class frame_values {
  struct entry_base {
    std::string name;
    virtual ~entry_base() {}
    virtual void emit(oarchive& ar) const = 0;
    virtual bool has_nan() const = 0;
  };
  template <class T>
  struct entry : entry_base {
    T value;
    void emit(oarchive& ar) const override { ar.add(name, value); }
    bool has_nan() const override { return archive_detail::any_nan(value); }
  };
  std::vector<std::unique_ptr<entry_base>> entries_;

 public:
  template <class T>
  frame_values& add(const std::string& name, T&& value) {
    auto e = std::make_unique<entry<typename std::decay<T>::type>>();
    e->name = name;
    e->value = std::forward<T>(value);
    entries_.push_back(std::move(e));
    return *this;
  }
  // what the archive's bad-value flag would say of these values
  bool has_nan() const {
    for (const auto& e : entries_)
      if (e->has_nan()) return true;
    return false;
  }
  void emit(oarchive& ar) const {
    for (const auto& e : entries_) e->emit(ar);
  }
};
