"""numpy restatement of the reference's post-processing definitions used by the
run diagnostics.

TEST INFRASTRUCTURE ONLY (the checker of mass_lyo_diagnostics).

  velocity     POST/massPy/base_modules/flow.py:4-11
  director     POST/massPy/nematic/nematicPy.py:4-15
  charge       POST/massPy/base_modules/defects.py:15-50
  defects      POST/massPy/base_modules/defects.py:58-110 with the threshold
               0.4 of nematicPy.py:74-95 (breadth-first collapse of 4-connected
               same-sign regions, periodic)
"""
import numpy as np

RING = [(1, 0), (1, 1), (0, 1), (-1, 1), (-1, 0), (-1, -1), (0, -1), (1, -1), (1, 0)]


def velocity(ff, LX, LY):
    d = ff.sum(axis=1)
    vx = (ff[:, 1] - ff[:, 2] + ff[:, 5] - ff[:, 6] - ff[:, 7] + ff[:, 8]) / d
    vy = (ff[:, 3] - ff[:, 4] + ff[:, 5] - ff[:, 6] + ff[:, 7] - ff[:, 8]) / d
    return vx.reshape(LX, LY), vy.reshape(LX, LY)


def director(Qxx, Qyx):
    with np.errstate(invalid="ignore", divide="ignore"):
        S = np.sqrt(Qxx ** 2 + Qyx ** 2)
        nx = np.sqrt(.5 * (1 + Qxx / S))
        ny = np.sign(Qyx) * np.sqrt(.5 * (1 - Qxx / S))
    return S, nx, ny


def _wang(a, b):
    ang = np.arctan2(a[0] * b[1] - a[1] * b[0], a[0] * b[0] + a[1] * b[1])
    return np.where(np.abs(ang) > .5 * np.pi, ang - np.sign(ang) * np.pi, ang)


def charge(Qxx, Qyx, LX, LY):
    _, nx, ny = director(np.asarray(Qxx, float).reshape(-1), np.asarray(Qyx, float).reshape(-1))
    v = np.stack([nx, ny]).reshape(2, LX, LY)
    w = np.zeros((LX, LY))
    with np.errstate(invalid="ignore"):
        for s0, s1 in zip(RING[:-1], RING[1:]):
            v0 = np.roll(v, shift=s0, axis=(1, 2))
            v1 = np.roll(v, shift=s1, axis=(1, 2))
            w += _wang(v0, v1)
    return w / (2 * np.pi)


def defects(Qxx, Qyx, LX, LY, threshold=.4):
    """List of (x, y, charge) in the order the reference's scan finds them."""
    w = charge(Qxx, Qyx, LX, LY)
    out = []
    for idx in np.ndindex(w.shape):
        if abs(w[idx]) > threshold:
            s = np.sign(w[idx])
            w[idx] = 0.
            queue, detected = [idx], [idx]
            while queue:
                i, j = queue.pop(0)
                for nbr in [(i + 1, j), (i - 1, j), (i, j + 1), (i, j - 1)]:
                    ni, nj = nbr[0] % LX, nbr[1] % LY
                    if s * w[ni, nj] > threshold and nbr not in detected:
                        detected.append(nbr)
                        queue.append(nbr)
                        w[ni, nj] = 0.
            x, y = np.mean(detected, 0)
            out.append((x % LX, y % LY, .5 * s))
    return out


def phi_histogram(phi, bins=64):
    p = np.asarray(phi, float)
    h = np.zeros(bins + 2, dtype=np.int64)
    under = ~(p >= 0)
    over = p >= 1
    mid = ~under & ~over
    b = np.minimum((p[mid] * bins).astype(np.int64), bins - 1)
    h[0] = under.sum()
    h[bins + 1] = over.sum()
    np.add.at(h, 1 + b, 1)
    return h
