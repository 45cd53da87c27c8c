// Checks the host program's JSON writer (mass_hd_amd/host/serialization.hpp)
// against the stream it replaces, byte for byte:
//  - every double as printf("%.6g") and as operator<< of a default
//    std::stringstream writes it: random bit patterns (all exponents,
//    subnormals, infinities, NaNs), scaled normal values, values on and next
//    to the rounding boundaries of six significant digits, exact halfway
//    cases, signed zeros, integers of every size;
//  - whole archives of vectors (sizes around the parallel split, empty and
//    one-element lists; double, unsigned, int, bool and 9-value nodes)
//    against the same archive written one value at a time through a
//    stringstream, and the NaN flag.
// Prints "format ok <values> (short path <n>)" and exits 0 -- the short exact
// path must have written most of the doubles -- or lists the first
// mismatches.
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include <unistd.h>

#include "serialization.hpp"

static int bad = 0;
static long checked = 0;
static long doubles = 0, fast = 0;  // doubles checked, and those the short exact path wrote

static std::string streamed(double v) {
  std::stringstream ss;
  ss << v;
  return ss.str();
}

static void check_double(double v) {
  char probe[16];
  ++doubles;
  if (archive_detail::put_g6(probe, v)) ++::fast;
  std::string fast;
  archive_detail::put(fast, v);
  char b[64];
  std::snprintf(b, sizeof b, "%.6g", v);
  ++checked;
  if (fast != b || (v == v && fast != streamed(v))) {
    if (bad++ < 20)
      std::fprintf(stderr, "double %.17g (bits %016llx): writer '%s', printf '%s', stream '%s'\n",
                   v, (unsigned long long)[&] { uint64_t u; std::memcpy(&u, &v, 8); return u; }(),
                   fast.c_str(), b, streamed(v).c_str());
  }
}

template <class T>
static void check_int(T v) {
  std::string fast;
  archive_detail::put(fast, v);
  std::stringstream ss;
  ss << v;
  ++checked;
  if (fast != ss.str() && bad++ < 20)
    std::fprintf(stderr, "integer: writer '%s', stream '%s'\n", fast.c_str(), ss.str().c_str());
}

// the archive as the one-value-at-a-time stream writes it
template <class T>
static void slow_value(std::ostream& o, const T& v) {
  std::stringstream ss;
  ss << v;
  o << ss.str();
}
template <class T, size_t M>
static void slow_value(std::ostream& o, const std::array<T, M>& a) {
  o << "[ ";
  for (size_t i = 0; i < M; ++i) {
    if (i) o << ", ";
    slow_value(o, a[i]);
  }
  o << " ]";
}
template <class V>
static std::string slow_list(const V& v) {
  std::stringstream o;
  o << "[ ";
  bool first = true;
  for (const auto& x : v) {
    if (!first) o << ", ";
    first = false;
    slow_value(o, static_cast<typename V::value_type>(x));
  }
  o << " ]";
  return o.str();
}

template <class V>
static void check_list(const V& v, const char* what) {
  bool nan = false;
  const std::string fast = archive_detail::format_list(v, nan);
  const std::string slow = slow_list(v);
  checked += (long)v.size();
  if (fast != slow) {
    size_t i = 0;
    while (i < fast.size() && i < slow.size() && fast[i] == slow[i]) ++i;
    if (bad++ < 20)
      std::fprintf(stderr, "%s list of %zu: differs at byte %zu ('%.40s' vs '%.40s')\n", what,
                   v.size(), i, fast.c_str() + std::min(i, fast.size()),
                   slow.c_str() + std::min(i, slow.size()));
  }
  if (nan) {
    if (bad++ < 20) std::fprintf(stderr, "%s list of %zu: NaN flag without a NaN\n", what, v.size());
  }
}

int main(int argc, char** argv) {
  // --lists: only the lists and frames (the parallel parts), for a run under
  // ThreadSanitizer
  const bool lists_only = argc > 1 && std::strcmp(argv[1], "--lists") == 0;
  std::mt19937_64 gen(20261204);
  std::normal_distribution<double> nd(0., 1.);
  std::uniform_real_distribution<double> ud(0., 1.);
  if (!lists_only) {
  // random bit patterns: every exponent, subnormals, inf and NaN payloads
  for (int i = 0; i < 2000000; ++i) {
    const uint64_t u = gen();
    double v;
    std::memcpy(&v, &u, 8);
    check_double(v);
  }
  // the values a frame holds: scaled normal and uniform draws
  for (int e = -40; e <= 40; ++e) {
    const double s = std::pow(10., e);
    for (int i = 0; i < 20000; ++i) {
      check_double(nd(gen) * s);
      check_double(ud(gen) * s);
    }
  }
  // on and next to the boundaries of six significant digits: d.ddddd5 x 10^e
  // and the doubles around them, integers with exact halves
  for (int e = -30; e <= 30; ++e) {
    for (int i = 0; i < 4000; ++i) {
      const long m = 100000 + (long)(gen() % 900000);  // six digits
      const double v = (m + 0.5) * std::pow(10., e - 5);
      check_double(v);
      check_double(std::nextafter(v, 0.));
      check_double(std::nextafter(v, 1e300));
      check_double(-v);
    }
  }
  // exact ties at the sixth digit: seven-digit integers ending in 5 and
  // six-digit integers plus one half
  for (long m = 100000; m < 300000; ++m) {
    check_double((double)(m * 10 + 5));
    check_double((double)m + 0.5);
  }
  for (long m = 1; m < 2000000; m += 7) check_double((double)m);
  const double special[] = {0.,
                            -0.,
                            1.,
                            -1.,
                            0.1,
                            0.5,
                            999999.5,
                            9999995.,
                            1e-5,
                            1e-4,
                            0.0001,
                            123456.,
                            1234565.,
                            1e21,
                            1e-300,
                            std::numeric_limits<double>::min(),
                            std::numeric_limits<double>::denorm_min(),
                            std::numeric_limits<double>::max(),
                            std::numeric_limits<double>::infinity(),
                            -std::numeric_limits<double>::infinity(),
                            std::numeric_limits<double>::quiet_NaN()};
  for (double v : special) check_double(v);
  for (int i = 0; i < 200000; ++i) {
    check_int((unsigned)gen());
    check_int((int)gen());
    check_int((long long)gen());
  }
  check_int(0u);
  check_int(std::numeric_limits<int>::min());
  }  // !lists_only

  // whole lists, serial and parallel
  const size_t sizes[] = {0, 1, 2, 1000, 32767, 32768, 32769, 100003, 262144, 1 << 20};
  for (size_t n : sizes) {
    std::vector<double> d(n);
    for (auto& x : d) x = nd(gen) * std::pow(10., (int)(gen() % 17) - 8);
    check_list(d, "double");
    std::vector<unsigned> u(n);
    for (auto& x : u) x = (unsigned)gen();
    check_list(u, "unsigned");
    std::vector<int> s(n);
    for (auto& x : s) x = (int)gen();
    check_list(s, "int");
    std::vector<bool> b(n);
    for (size_t i = 0; i < n; ++i) b[i] = gen() & 1;
    check_list(b, "bool");
    std::vector<std::array<double, 9>> ff(n / 9 + 1);
    for (auto& a : ff)
      for (auto& x : a) x = 40. / 9. + nd(gen) * 1e-3;
    check_list(ff, "node");
  }
  // a NaN anywhere in a long list sets the flag
  {
    std::vector<double> d(300000, 0.25);
    d[271828] = std::numeric_limits<double>::quiet_NaN();
    bool nan = false;
    archive_detail::format_list(d, nan);
    if (!nan && bad++ < 20) std::fprintf(stderr, "NaN in a long list not flagged\n");
  }
  // a whole frame through the piece stream (the frame writer's) and through
  // a stringstream: the same bytes
  {
    std::vector<std::array<double, 9>> ff(70000);
    for (auto& a : ff)
      for (auto& x : a) x = 40. / 9. + nd(gen) * 1e-3;
    std::vector<double> phi(250000);
    for (auto& x : phi) x = ud(gen);
    std::vector<bool> mask(250000);
    for (size_t i = 0; i < mask.size(); ++i) mask[i] = (gen() % 7) == 0;
    auto fill = [&](std::ostream& o) {
      oarchive ar(o, "frame", 1);
      ar.add("ff", ff).add("phi", phi).add("division_mask", mask).add("LX", 500u).add("name",
                                                                                     std::string("x"));
    };
    std::stringstream a;
    fill(a);
    piece_stream b;
    fill(b);
    std::stringstream bb;
    b.pieces().write_to(bb);
    // and streamed to a file from the I/O thread as the lists are formatted
    const std::string path = "/tmp/format_check_" + std::to_string((long)getpid()) + ".json";
    {
      std::ofstream f(path);
      piece_stream c;
      c.pieces().stream_to(f);
      fill(c);
      c.pieces().finish();
    }
    std::string from_file;
    {
      std::ifstream in(path, std::ios::binary);
      std::stringstream ss;
      ss << in.rdbuf();
      from_file = ss.str();
    }
    std::remove(path.c_str());
    if (from_file != a.str() && bad++ < 20)
      std::fprintf(stderr, "frame streamed to a file differs (%zu vs %zu bytes)\n",
                   from_file.size(), a.str().size());
    checked += (long)(9 * ff.size() + phi.size() + mask.size());
    if (a.str() != bb.str() && bad++ < 20)
      std::fprintf(stderr, "frame through the piece stream differs (%zu vs %zu bytes)\n",
                   bb.str().size(), a.str().size());
  }
  // a frame collected first and written later (the host program's frame
  // writer thread): the same bytes as the archive written directly, and the
  // NaN check of the collected values says what the archive's flag says
  {
    std::vector<std::array<double, 9>> ff(20000);
    for (auto& a : ff)
      for (auto& x : a) x = 40. / 9. + nd(gen) * 1e-3;
    std::vector<double> phi(70000);
    for (auto& x : phi) x = ud(gen);
    std::vector<bool> mask(70000);
    for (size_t i = 0; i < mask.size(); ++i) mask[i] = (gen() % 5) == 0;
    std::stringstream direct;
    {
      oarchive ar(direct, "frame", 1);
      ar.add("ff", ff).add("phi", phi).add("division_mask", mask);
    }
    frame_values fv;
    fv.add("ff", ff).add("phi", std::vector<double>(phi)).add("division_mask", mask);
    std::stringstream later;
    {
      oarchive ar(later, "frame", 1);
      fv.emit(ar);
    }
    if (direct.str() != later.str() && bad++ < 20)
      std::fprintf(stderr, "a collected frame writes other bytes\n");
    if (fv.has_nan() && bad++ < 20) std::fprintf(stderr, "NaN reported in a frame without one\n");
    for (int where = 0; where < 2; ++where) {
      frame_values f2;
      auto ff2 = ff;
      auto phi2 = phi;
      if (where == 0)
        ff2[12345][7] = std::numeric_limits<double>::quiet_NaN();
      else
        phi2[54321] = -std::numeric_limits<double>::quiet_NaN();
      f2.add("ff", std::move(ff2)).add("phi", std::move(phi2)).add("division_mask", mask);
      std::stringstream o;
      oarchive ar(o, "frame", 1);
      f2.emit(ar);
      if ((!f2.has_nan() || !ar.bad_value()) && bad++ < 20)
        std::fprintf(stderr, "NaN in a collected frame not reported (case %d)\n", where);
    }
  }
  if (bad) {
    std::fprintf(stderr, "%d mismatches in %ld values\n", bad, checked);
    return 1;
  }
  if (!lists_only && ::fast < doubles / 2) {
    std::fprintf(stderr, "the short path wrote only %ld of %ld doubles\n", ::fast, doubles);
    return 1;
  }
  std::printf("format ok %ld (short path %ld of %ld doubles)\n", checked, ::fast, doubles);
  return 0;
}
