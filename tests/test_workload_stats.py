"""The long-run comparison statistics of the configuration workloads
(mass_hd_amd.workloads.relative_stats): relative differences from the plain
run per checkpoint and as window means, on series with known answers."""
import numpy as np
import pytest

from mass_hd_amd.workloads import relative_stats


def _series(vals_u, vals_d, every=100):
    return [(every * (k + 1), float(u), float(d)) for k, (u, d) in enumerate(zip(vals_u, vals_d))]


def test_identical_runs_have_zero_differences():
    u = np.linspace(1.0, 2.0, 40)
    d = np.linspace(3e-3, 1e-3, 40)
    s = {"plain": _series(u, d), "twin": _series(u, d)}
    r = relative_stats(s, windows=4)["twin"]
    for name in ("u_rms", "defect_density"):
        assert r[name]["checkpoint_max_abs"] == 0.0
        assert r[name]["window_means"] == [0.0] * 4
    assert r["checkpoints"] == 40


def test_constant_bias_and_windows():
    """A run 2% above the plain one at every checkpoint: every checkpoint and
    window difference is +2%; windows of 10 checkpoints of 100 steps span
    1,000 steps."""
    u = np.full(40, 5.0)
    d = np.linspace(1.0, 2.0, 40)
    s = {"plain": _series(u, d), "surrogate": _series(u * 1.02, d * 1.02)}
    r = relative_stats(s, windows=4)["surrogate"]
    for name in ("u_rms", "defect_density"):
        assert r[name]["checkpoint_mean"] == pytest.approx(0.02, rel=1e-12)
        assert r[name]["window_means"] == pytest.approx([0.02] * 4, rel=1e-12)
        assert r[name]["window_max_abs"] == pytest.approx(0.02, rel=1e-12)
        assert r[name]["window_steps"] == 1000


def test_window_means_are_ratios_of_window_averages():
    """Alternating +10% / -10% checkpoints: the checkpoint extremes are 10%,
    while a window mean is the ratio of the window's averages (not the mean
    of the ratios)."""
    base = np.array([1.0, 3.0] * 10)
    run = base * np.array([1.1, 0.9] * 10)
    s = {"plain": _series(base, base), "surrogate": _series(run, run)}
    r = relative_stats(s, windows=2)["surrogate"]["defect_density"]
    assert r["checkpoint_max_abs"] == pytest.approx(0.1)
    w = (1.1 * 1.0 + 0.9 * 3.0) / 4.0 - 1.0   # (sum run) / (sum base) - 1 per window
    assert r["window_means"] == pytest.approx([w, w])


def test_shorter_run_and_fp64_single_window():
    """An FP64 run over the first checkpoints only is compared over those,
    as one window."""
    base = np.linspace(1.0, 2.0, 30)
    s = {"plain": _series(base, base), "fp64": _series(base[:10] * 1.001, base[:10])}
    r = relative_stats(s, windows=10)["fp64"]
    assert r["checkpoints"] == 10
    assert len(r["u_rms"]["window_means"]) == 1
    assert r["u_rms"]["window_means"][0] == pytest.approx(0.001, rel=1e-9)
    assert r["defect_density"]["checkpoint_max_abs"] == 0.0
