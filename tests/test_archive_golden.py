"""The host program's archive stays readable by the reference's post-processing.

tests/golden/make_archive_golden.py had the reference's reader
(POST/massPy/archive.py) load a small archive written by this build's
serializer (tests/golden/archive_harness.cpp) and recorded what it returned
(tests/golden/archive_read.json) next to the archive files it read
(tests/golden/archive_sample/).  Here the serializer writes that archive again:
it must be byte for byte the files the reader parsed, and the recorded values
must be the written ones (frame count from nsteps/nstart/ninfo, integer and
boolean parameters as such, numbers to six significant digits, ff as nine
values per node).
"""
import json
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
HOST = os.path.join(os.path.dirname(HERE), "mass_hd_amd", "host")


def test_serializer_writes_the_archive_the_reference_reader_parsed(tmp_path):
    exe = str(tmp_path / "h")
    subprocess.check_call(["g++", "-std=c++17", "-I" + HOST, "-o", exe,
                           os.path.join(GOLDEN, "archive_harness.cpp")])
    out = tmp_path / "ar"
    out.mkdir()
    subprocess.check_call([exe, str(out)])
    sample = os.path.join(GOLDEN, "archive_sample")
    names = sorted(os.listdir(sample))
    assert names == sorted(os.listdir(out)) == ["frame0.json", "frame10.json", "frame5.json",
                                                "parameters.json"]
    for n in names:
        with open(os.path.join(sample, n), "rb") as a, open(out / n, "rb") as b:
            assert a.read() == b.read(), n


def test_reference_reader_values():
    with open(os.path.join(GOLDEN, "archive_read.json")) as f:
        g = json.load(f)
    p = g["parameters"]
    assert g["num_frames"] == 3 == (p["nsteps"] - p["nstart"]) // p["ninfo"] + 1
    assert p["LX"] == 3 and p["LY"] == 2 and p["bc"] == 0 and p["conserve_phi"] is False
    assert p["model_name"] == "lyotropic" and p["init_config"] == "random"
    assert p["tauNem"] == 2.0 and p["zeta"] == 0.02 and p["xi"] == 0.7
    assert [fr["time"] for fr in g["frames"]] == [0, 5, 10]
    for fr in g["frames"]:
        t = fr["time"]
        k = np.arange(6)
        ff = np.array(fr["ff"])
        assert ff.shape == (6, 9)
        exp = 40.0 / 9.0 + 0.01 * (k[:, None] + np.arange(9)[None, :]) - 1e-7 * t
        assert np.allclose(ff, exp, rtol=1e-5, atol=0)  # six significant digits
        assert np.allclose(fr["QQxx"], 0.5 * np.cos(0.3 * k + t), rtol=1e-5, atol=1e-6)
        assert np.allclose(fr["phi"], 0.5 + 1e-3 * k - 2.5e-5 * t, rtol=1e-5, atol=0)
