"""Writes tests/golden/archive_read.json: what the reference's own reader
(POST/massPy/archive.py, the post-processing of MASS) returns for a small
archive that the host program's serializer writes (tests/golden/
archive_harness.cpp).  The archive files themselves are kept beside it
(tests/golden/archive_sample/), so the CPU test can check that the serializer
still writes exactly what the reader was seen to parse.

Run in the build container, where the reference is available:
    python tests/golden/make_archive_golden.py
"""
import json
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_POST = "/root/reference/POST/massPy"


def build_and_write(outdir):
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "h")
        subprocess.check_call(["g++", "-std=c++17", "-I" + os.path.join(ROOT, "mass_hd_amd", "host"),
                               "-o", exe, os.path.join(HERE, "archive_harness.cpp")])
        os.makedirs(outdir, exist_ok=True)
        subprocess.check_call([exe, outdir])


def main():
    sample = os.path.join(HERE, "archive_sample")
    if os.path.isdir(sample):
        shutil.rmtree(sample)
    build_and_write(sample)
    sys.path.insert(0, REF_POST)
    import archive  # the reference's reader
    ar = archive.loadarchive(sample)
    params = {k: v for k, v in ar.__dict__.items() if not k.startswith("_")}
    frames = []
    for fr in ar.read_frames():
        frames.append({"time": fr.time,
                       "ff": fr.ff.tolist(), "QQxx": fr.QQxx.tolist(), "phi": fr.phi.tolist()})
    out = {"reader": "POST/massPy/archive.py Archive (loadarchive, read_frames)",
           "num_frames": ar.num_frames, "parameters": params, "frames": frames}
    with open(os.path.join(HERE, "archive_read.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", os.path.join(HERE, "archive_read.json"))


if __name__ == "__main__":
    main()
