"""The C++ host program (mass_hd_amd/bin/mass_hip): option passes, runcard
semantics, error paths and the JSON archive format.  These run on the CPU:
every case here stops before the model needs the GPU, except the ones marked
gpu at the end.
"""
import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "mass_hd_amd", "bin", "mass_hip")
HOST = os.path.join(ROOT, "mass_hd_amd", "host")

RUNCARD = """# plain lyotropic model, 32^2
model = lyotropic
LX = 32
LY = 32
nsteps = 20
ninfo = 10
seed = 1
config = random
angle = 90
noise = 1
Qamp = 0.5
phinoise = 0.05
tauNem = 2.0
tauIso = 2.0
GammaQ = 0.1
GammaP = 0.1
xi = 0.7
zeta = 0.02
LL = 0.04
AA = 0.01
KK = 0.04
CC = 0
zetaI = 0
conc = 0.5
friction = 0
npc = 1
"""


# the reference's proliferation runcard (go-or-grow), at 64^2 and 20 steps
GROWTH_RUNCARD = """model    = go-or-grow
config   = circle
nsteps   = 20
ninfo    = 10
LX       = 64
LY       = 64
bc       = 0
seed     = 1001
GammaQ   = 0.3
GammaP   = 5e-2
zeta     = 0
zetaI    = 0
friction = 0.0
tauNem   = 1.0
tauIso   = 1.0
AA       = 0.03
CC       = 0.1
LL       = 0.002
KK       = 0.06
rho      = 40
xi       = 0.3
B        = 0.1
Snem     = 1
level    = 32
radius   = 12
conc     = 1
noise    = 1
angle    = 0
relax-steps = 100
relax-dt    = 1.0
relax-phi   = 1
relax-Q     = 1
chi-config = noise
chi0       = 1
chi-noise  = 0
chi-length = 0
Dchi       = 0
Achi      = 0
Ochi      = 0
phiswitch = 0
growTogether = 0
beta         = 0
death-rate   = 0
death-time   = 1
death-radius = 0
division-rate   = 1e-1
division-time   = 2
division-radius = 3
alpha           = 0.001
phi-critical    = 1.2
"""


def run(args, cwd=None):
    r = subprocess.run([EXE] + args, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       cwd=cwd, timeout=120)
    return r.returncode, r.stdout, r.stderr


@pytest.fixture
def card(tmp_path):
    def make(text=RUNCARD):
        p = tmp_path / "run.dat"
        p.write_text(text)
        return str(p)
    return make


def test_no_argument():
    rc, out, err = run([])
    assert rc == 1 and "no argument provided" in err


def test_list_models():
    rc, out, err = run(["--list"])
    assert rc == 0 and "lyotropic" in out


def test_help_shows_model_options():
    rc, out, _ = run(["-h", "-m", "lyotropic"])
    assert rc == 0
    for key in ("GammaQ", "npc", "config", "precision", "--LX", "--nsteps"):
        assert key in out


def test_go_or_grow_options():
    rc, out, _ = run(["-h", "-m", "go-or-grow"])
    assert rc == 0
    for key in ("division-rate", "phi-critical", "chi-config", "relax-steps", "growTogether",
                "Snem", "GammaQ"):
        assert key in out
    rc, out, _ = run(["--list"])
    assert "go-or-grow" in out


def test_division_model_options():
    rc, out, _ = run(["-h", "-m", "lyotropic-with-division"])
    assert rc == 0
    for key in ("division-rate", "death-radius", "phi-critical", "GammaQ"):
        assert key in out
    assert "chi-config" not in out


def test_go_or_grow_bad_values(card):
    rc, _, err = run([card(GROWTH_RUNCARD.replace("relax-phi   = 1", "relax-phi   = x")),
                      "--no-write"])
    assert rc == 1 and "relax-phi" in err
    rc, _, err = run([card(GROWTH_RUNCARD + "precision = fp32\n"), "--no-write"])
    assert rc == 1


def test_missing_runcard(tmp_path):
    rc, _, err = run([str(tmp_path / "nope.dat")])
    assert rc == 1 and "can not open runcard file" in err


def test_unknown_model(card):
    rc, _, err = run([card(RUNCARD.replace("lyotropic", "nematic"))])
    assert rc == 1 and "can not be found" in err


def test_unknown_key_rejected_in_second_pass(card):
    rc, _, err = run([card(RUNCARD + "foo = 3\n")])
    assert rc == 1 and "unrecognised option 'foo'" in err


def test_duplicate_key_rejected(card):
    rc, _, err = run([card(RUNCARD + "LX = 64\n")])
    assert rc == 1 and "more than once" in err


def test_bad_value(card):
    rc, _, err = run([card(RUNCARD.replace("LX = 32", "LX = abc"))])
    assert rc == 1 and "is invalid" in err


def test_nsubsteps_rejected(card):
    # lyotropic.cpp:64-66, raised before any device work
    rc, _, err = run([card(RUNCARD + "nsubsteps = 2\n"), "--no-write"])
    assert rc == 1 and "nsubsteps=1" in err


def test_boundary_conditions_rejected(card):
    # bc 0-4 (periodic, channels, boxes) are offered; the other values are not
    rc, _, err = run([card(RUNCARD + "bc = 5\n"), "--no-write"])
    assert rc == 1 and "bc=5" in err and "channel" in err


def test_command_line_overrides_runcard(card):
    # the command line is stored first, so it wins; bc on the command line
    rc, _, err = run([card(RUNCARD + "bc = 6\n"), "--bc", "7", "--no-write"])
    assert rc == 1 and "bc=7" in err
    # and a runcard value alone is used
    rc, _, err = run([card(RUNCARD + "bc = 6\n"), "--no-write"])
    assert rc == 1 and "bc=6" in err


def test_bad_precision(card):
    rc, _, err = run([card(RUNCARD + "precision = fp16\n"), "--no-write"])
    assert rc == 1 and "precision" in err


_SER = r"""
#include <array>
#include <iostream>
#include <string>
#include <vector>
#include "serialization.hpp"
int main() {
  unsigned u = 7; double d = 0.1; std::string s = "lyotropic";
  std::vector<double> v = {1.0, 2.5, -3e-7, 123456789.0};
  std::vector<std::array<double, 9>> a(2);
  for (int i = 0; i < 9; ++i) { a[0][i] = i; a[1][i] = 40.0 / (i + 1); }
  {
    oarchive ar(std::cout, "frame", 1);
    ar.add("u", u).add("d", d).add("s", s).add("v", v).add("ff", a);
  }
  return 0;
}
"""


def test_archive_format(tmp_path):
    """The archive writes the reference's JSON layout (id/version/data, each
    entry a {type, value} group) with six-significant-digit numbers."""
    src = tmp_path / "t.cpp"
    src.write_text(_SER)
    exe = str(tmp_path / "t")
    subprocess.check_call(["g++", "-std=c++17", "-I" + HOST, "-o", exe, str(src)])
    out = subprocess.check_output([exe], text=True)
    assert out.startswith('{\n  "id" : "frame",\n  "version" : 1,\n  "data" : {\n    "u" : {\n'
                          '      "type" : "unsigned",\n      "value" : 7\n    },')
    j = json.loads(out)
    assert j["id"] == "frame" and j["version"] == 1
    d = j["data"]
    assert d["u"] == {"type": "unsigned", "value": 7}
    assert d["d"] == {"type": "double", "value": 0.1}
    assert d["s"] == {"type": "string", "value": "lyotropic"}
    assert d["v"]["type"] == "array(double)"
    assert d["v"]["value"] == [1.0, 2.5, -3e-07, 123457000.0]  # %g, 6 digits
    assert d["ff"]["type"] == "array(array(double))"
    assert np.allclose(d["ff"]["value"][1], [40.0 / (i + 1) for i in range(9)], rtol=1e-5)
    assert '"value" : [ 1, 2.5, -3e-07, 1.23457e+08 ]' in out


# ------------------------------------------------------------------ on the GPU

@pytest.mark.gpu
def test_host_run_matches_oracle(card, tmp_path, require_gpu):
    """End to end through the Model interface: runcard -> frames on disk; the
    frame fields equal the oracle's to the archive's six digits."""
    from oracle import pyoracle as po
    outdir = tmp_path / "out"
    rc, out, err = run([card(), "-o", str(outdir), "-f"])
    assert rc == 0, err
    assert "fcheck:" in out
    frames = sorted(os.listdir(outdir))
    assert frames == ["frame0.json", "frame10.json", "frame20.json", "parameters.json"]
    params = json.load(open(outdir / "parameters.json"))["data"]
    assert params["LX"]["value"] == 32 and params["model_name"]["value"] == "lyotropic"
    assert params["init_config"]["value"] == "random"
    o = po.Oracle(32, 32, npc=1, GammaQ=0.1, GammaP=0.1, xi=0.7, tauNem=2.0, tauIso=2.0,
                  rho=40.0, friction=0.0, AA=0.01, CC=0.0, LL=0.04, KK=0.04, zeta=0.02,
                  zetaI=0.0, conc=0.5)
    o.configure(config="random", seed=1, angle=90, noise=1, qamp=0.5, phinoise=0.05)
    f0 = json.load(open(outdir / "frame0.json"))["data"]
    assert not np.any(f0["sigmaXX"]["value"])  # zero before the first step
    assert np.allclose(f0["QQxx"]["value"], o.get("QQxx"), rtol=1e-5, atol=1e-6)
    o.step(20)
    f20 = json.load(open(outdir / "frame20.json"))["data"]
    for k in ("QQxx", "QQyx", "phi", "sigmaXY", "sigma_active_xx"):
        assert np.allclose(f20[k]["value"], o.get(k), rtol=1e-5, atol=1e-6), k
    assert np.allclose(np.array(f20["ff"]["value"]), o.get("ff"), rtol=1e-5)


@pytest.mark.gpu
@pytest.mark.parametrize("bc", [2, 4])
def test_host_walled_run_matches_oracle(card, tmp_path, require_gpu, bc):
    """A runcard with walls (bc 2: no-slip channel, bc 4: no-slip box) gives
    the oracle's frames to the archive's six digits."""
    from oracle import pyoracle as po
    outdir = tmp_path / "out"
    rc, out, err = run([card(RUNCARD + "bc = %d\n" % bc), "-o", str(outdir), "-f"])
    assert rc == 0, err
    o = po.Oracle(32, 32, npc=1, bc=bc, GammaQ=0.1, GammaP=0.1, xi=0.7, tauNem=2.0, tauIso=2.0,
                  rho=40.0, friction=0.0, AA=0.01, CC=0.0, LL=0.04, KK=0.04, zeta=0.02,
                  zetaI=0.0, conc=0.5)
    o.configure(config="random", seed=1, angle=90, noise=1, qamp=0.5, phinoise=0.05)
    o.step(20)
    f20 = json.load(open(outdir / "frame20.json"))["data"]
    for k in ("QQxx", "QQyx", "phi", "sigmaXX", "sigmaXY"):
        assert np.allclose(f20[k]["value"], o.get(k), rtol=1e-5, atol=1e-6), k
    assert np.allclose(np.array(f20["ff"]["value"]), o.get("ff"), rtol=1e-5)


@pytest.mark.gpu
def test_host_go_or_grow_run_matches_oracle(card, tmp_path, require_gpu):
    """The reference's proliferation runcard (64^2): frames with the masks,
    m and chi equal the oracle's to the archive's six digits."""
    from oracle import pyoracle as po
    outdir = tmp_path / "out"
    rc, out, err = run([card(GROWTH_RUNCARD), "-o", str(outdir), "-f"])
    assert rc == 0, err
    params = json.load(open(outdir / "parameters.json"))["data"]
    assert params["model_name"]["value"] == "go-or-grow"
    assert params["division_rate"]["value"] == 0.1 and params["relax_steps"]["value"] == 100
    o = po.Oracle(64, 64, npc=1, GammaQ=0.3, GammaP=5e-2, xi=0.3, tauNem=1.0, tauIso=1.0,
                  rho=40.0, friction=0.0, AA=0.03, CC=0.1, LL=0.002, KK=0.06, zeta=0.0,
                  zetaI=0.0, conc=1.0)
    o.set_growth(B=0.1, Snem=1.0, relax_steps=100, relax_dt=1.0, relax_phi=1, relax_Q=1,
                 chi_config="noise", chi0=1.0, alpha=0.001, beta=0.0, division_rate=0.1,
                 division_time=2.0, division_radius=3.0, death_rate=0.0, death_time=1.0,
                 death_radius=0.0, phi_critical=1.2)
    o.configure(config="circle", seed=1001, radius=12.0, level=32.0, noise=1.0, angle=0.0)
    o.step(20)
    f = json.load(open(outdir / "frame20.json"))["data"]
    for k in ("QQxx", "QQyx", "phi", "m", "chi", "sigmaXX"):
        assert np.allclose(f[k]["value"], o.get(k), rtol=1e-5, atol=1e-6), k
    div, death = o.masks()
    assert np.array_equal(np.array(f["division_mask"]["value"]) != 0, div != 0)
    assert not np.any(f["death_mask"]["value"])


@pytest.mark.gpu
def test_host_division_run_matches_oracle(card, tmp_path, require_gpu):
    """model = lyotropic-with-division: frames with the masks equal the
    oracle's to the archive's six digits."""
    from oracle import pyoracle as po
    text = RUNCARD + """model = lyotropic-with-division
alpha = 0.002
beta = 0.001
division-rate = 0.2
division-time = 2
division-radius = 2
death-rate = 0.05
death-time = 1
death-radius = 1
phi-critical = 1.2
"""
    outdir = tmp_path / "out"
    rc, out, err = run([card(text.replace("model = lyotropic\n", "")), "-o", str(outdir), "-f"])
    assert rc == 0, err
    params = json.load(open(outdir / "parameters.json"))["data"]
    assert params["model_name"]["value"] == "lyotropic-with-division"
    o = po.Oracle(32, 32, npc=1, GammaQ=0.1, GammaP=0.1, xi=0.7, tauNem=2.0, tauIso=2.0,
                  rho=40.0, friction=0.0, AA=0.01, CC=0.0, LL=0.04, KK=0.04, zeta=0.02,
                  zetaI=0.0, conc=0.5)
    o.set_growth(division_only=1, alpha=0.002, beta=0.001, division_rate=0.2, division_time=2.0,
                 division_radius=2.0, death_rate=0.05, death_time=1.0, death_radius=1.0,
                 phi_critical=1.2, chi0=0.0)
    o.configure(config="random", seed=1, angle=90, noise=1, qamp=0.5, phinoise=0.05)
    o.step(20)
    f = json.load(open(outdir / "frame20.json"))["data"]
    for k in ("QQxx", "QQyx", "phi", "sigmaXX"):
        assert np.allclose(f[k]["value"], o.get(k), rtol=1e-5, atol=1e-6), k
    div, death = o.masks()
    assert np.array_equal(np.array(f["division_mask"]["value"]) != 0, div != 0)
    assert np.array_equal(np.array(f["death_mask"]["value"]) != 0, death != 0)
    assert "m" not in f


@pytest.mark.gpu
def test_host_restart_from_checkpoint(card, tmp_path, require_gpu):
    """20 steps + checkpoint, then a run restarted from it for 10 steps, ends
    in the frame of one 30-step run (the same state: identical text)."""
    full = tmp_path / "full"
    rc, _, err = run([card(RUNCARD.replace("nsteps = 20", "nsteps = 30")), "-o", str(full), "-f"])
    assert rc == 0, err
    ckp = tmp_path / "state.ckp"
    rc, _, err = run([card(RUNCARD + "checkpoint = %s\n" % ckp), "--no-write"])
    assert rc == 0, err
    assert ckp.exists() and ckp.read_bytes()[:8] == b"MASSLYO1"
    part = tmp_path / "part"
    rc, _, err = run([card(RUNCARD.replace("nsteps = 20", "nsteps = 10") +
                           "restart = %s\n" % ckp), "-o", str(part), "-f"])
    assert rc == 0, err
    a = json.load(open(full / "frame30.json"))["data"]
    b = json.load(open(part / "frame10.json"))["data"]
    for k in ("ff", "QQxx", "QQyx", "phi"):
        assert a[k]["value"] == b[k]["value"], k
    rc, _, err = run([card(RUNCARD + "restart = %s\n" % (tmp_path / "missing")), "--no-write"])
    assert rc == 1 and "restart file" in err



@pytest.mark.gpu
@pytest.mark.parametrize("flag", ["-c", "--compress-full"])
def test_host_compressed_output(card, tmp_path, require_gpu, flag):
    """The reference's zip output modes (write.cpp:48-59): -c zips every
    file on its own, --compress-full adds them all to one archive; either
    way the archives hold exactly the files an uncompressed run writes."""
    import shutil
    import zipfile
    if shutil.which("zip") is None:
        pytest.skip("zip is not installed")
    plain = tmp_path / "plain"
    rc, out, err = run([card(), "-o", str(plain) + "/", "-f"])
    assert rc == 0, err
    if flag == "-c":
        outdir = tmp_path / "zipped"
        rc, out, err = run([card(), "-o", str(outdir) + "/", "-f", "-c"])
        assert rc == 0, err
        names = sorted(os.listdir(outdir))
        assert names == sorted(f + ".zip" for f in os.listdir(plain))
        for z in names:
            with zipfile.ZipFile(outdir / z) as zf:
                (inner,) = zf.namelist()
                assert inner == z[:-4]
                assert zf.read(inner) == (plain / inner).read_bytes(), inner
    else:
        archive = tmp_path / "run"
        rc, out, err = run([card(), "-o", str(archive), "-f", "--compress-full"])
        assert rc == 0, err
        with zipfile.ZipFile(str(archive) + ".zip") as zf:
            assert sorted(zf.namelist()) == sorted(os.listdir(plain))
            for inner in zf.namelist():
                assert zf.read(inner) == (plain / inner).read_bytes(), inner


@pytest.mark.gpu
def test_host_nstart(card, tmp_path, require_gpu):
    """Frames from step nstart on, nstart rounded up to a multiple of ninfo
    (options.cpp:206-216): nstart = 5 with ninfo = 10 writes frames 10 and
    20 only; their contents are those of a run that writes every frame."""
    full = tmp_path / "full"
    rc, out, err = run([card(), "-o", str(full) + "/", "-f"])
    assert rc == 0, err
    late = tmp_path / "late"
    rc, out, err = run([card(RUNCARD.replace("ninfo = 10", "ninfo = 10\nnstart = 5")), "-o",
                        str(late) + "/", "-f"])
    assert rc == 0, err
    assert sorted(os.listdir(late)) == ["frame10.json", "frame20.json", "parameters.json"]
    for f in ("frame10.json", "frame20.json"):
        assert (late / f).read_bytes() == (full / f).read_bytes(), f


@pytest.mark.gpu
def test_host_nan_stops_before_the_frame(card, tmp_path, require_gpu):
    """A NaN in a frame's fields stops the run with the reference's message
    (write.cpp:41) and exit status 1 before that frame's file exists; the
    earlier frames are on disk.  tauNem = tauIso = 0 makes the first
    collision 0/0 (f starts at its equilibrium)."""
    text = RUNCARD.replace("tauNem = 2.0", "tauNem = 0").replace("tauIso = 2.0", "tauIso = 0")
    outdir = tmp_path / "out"
    rc, out, err = run([card(text), "-o", str(outdir) + "/", "-f"])
    assert rc == 1
    assert "nan found while writing file." in err
    files = sorted(os.listdir(outdir))
    assert "frame0.json" in files and "frame10.json" not in files, files
