"""FP64 numpy restatement of the surrogate corrector network.

TEST INFRASTRUCTURE ONLY: the checker of the GPU network kernels.

Network: four 3x3 periodic convolutions, 8 -> 32 -> 32 -> 32 -> 8 channels
(7 input and 5 output channels used, the rest zero), tanh-GELU after the first
three.  Layout: activations [LX, LY, C] with node (x, y); tap t = 3*(dx+1) +
(dy+1) reads the input at (x+dx, y+dy) (periodic); weights W[t, co, ci].
"""
import numpy as np

TAPS = [(dx, dy) for dx in (-1, 0, 1) for dy in (-1, 0, 1)]
CH = [8, 32, 32, 32, 8]
C_GELU = np.sqrt(2.0 / np.pi)


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(C_GELU * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    u = C_GELU * (x + 0.044715 * x ** 3)
    t = np.tanh(u)
    return 0.5 * (1 + t) + 0.5 * x * (1 - t * t) * C_GELU * (1 + 3 * 0.044715 * x * x)


def conv(a, W, b):
    """a [LX, LY, Ci], W [9, Co, Ci], b [Co] -> [LX, LY, Co]"""
    out = np.broadcast_to(b, a.shape[:2] + (W.shape[1],)).copy()
    for t, (dx, dy) in enumerate(TAPS):
        s = np.roll(a, shift=(-dx, -dy), axis=(0, 1))  # s[x, y] = a[x+dx, y+dy]
        out += s @ W[t].T
    return out


def conv_dgrad(d, W):
    """gradient wrt the conv input given d = dL/dout [LX, LY, Co]"""
    g = np.zeros(d.shape[:2] + (W.shape[2],))
    for t, (dx, dy) in enumerate(TAPS):
        g += np.roll(d @ W[t], shift=(dx, dy), axis=(0, 1))
    return g


def conv_wgrad(a, d):
    """dL/dW [9, Co, Ci] and dL/db [Co]"""
    dW = np.zeros((9, d.shape[2], a.shape[2]))
    for t, (dx, dy) in enumerate(TAPS):
        s = np.roll(a, shift=(-dx, -dy), axis=(0, 1))
        dW[t] = np.einsum("xyo,xyi->oi", d, s)
    return dW, d.sum(axis=(0, 1))


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    ps = []
    for l in range(4):
        ci, co = CH[l], CH[l + 1]
        W = rng.normal(0, np.sqrt(2.0 / (9 * ci)), (9, co, ci))
        ps.append([W, np.zeros(co)])
    return ps


def forward(x, ps):
    cache = [x]
    a = x
    for l, (W, b) in enumerate(ps):
        z = conv(a, W, b)
        cache.append(z)
        a = gelu(z) if l < 3 else z
    return a, cache


def loss_and_grads(x, y, ps):
    """L = mean over nodes of sum over channels of (out - y)^2 / 2"""
    out, cache = forward(x, ps)
    n = x.shape[0] * x.shape[1]
    d = (out - y) / n
    L = 0.5 * np.sum((out - y) ** 2) / n
    grads = [None] * 4
    for l in range(3, -1, -1):
        a_in = cache[0] if l == 0 else gelu(cache[l])
        dW, db = conv_wgrad(a_in, d)
        grads[l] = [dW, db]
        if l > 0:
            d = conv_dgrad(d, ps[l][0]) * gelu_grad(cache[l])
    return L, grads
