"""The host program's JSON writer (mass_hd_amd/host/serialization.hpp)
formats doubles by an exact six-digit path (std::to_chars near ties) and
long lists in parallel chunks; the
reference's archive streams each number through a std::stringstream
(src/serialization.hpp:187-361).  tests/native/format_check.cpp holds the
writer to the stream byte for byte: 14.8 million doubles (random bit
patterns, values on and next to the sixth-digit rounding boundaries, exact
ties, subnormals, signed zeros, infinities, NaNs) against printf("%.6g") and
operator<<, integer and bool lists, lists on both sides of the parallel
split, and a whole frame through the writer's piece stream against the same
frame through a stringstream."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "mass_hd_amd", "bin", "format_check")


def test_writer_matches_stream_byte_for_byte():
    assert os.path.exists(CHECK), "format_check is not built (run __graft_entry__.build())"
    r = subprocess.run([CHECK], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    assert r.stdout.startswith("format ok")


def test_parallel_writer_has_no_data_races():
    """The list and frame parts of the same check (pooled worker threads,
    the buffer pool, the piece stream) under ThreadSanitizer."""
    tsan = CHECK + "_tsan"
    assert os.path.exists(tsan), "format_check_tsan is not built (run __graft_entry__.build())"
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66")
    r = subprocess.run([tsan, "--lists"], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, timeout=600, env=env)
    assert r.returncode == 0 and "ThreadSanitizer" not in r.stderr, r.stderr[-3000:]
    assert r.stdout.startswith("format ok")
