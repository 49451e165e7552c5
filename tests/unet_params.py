"""Parameter layout of the U-Net preconditioner and hand-built weight sets
whose output is known in closed form (shifts, block averages), used to pin
both the FP64 oracle and the GPU network.

Layout: for conv k (cin_k -> cout_k) the weights [27][cout][cin] then the bias
[cout]; then the 1x1 head weights [16] and its bias. Tap t is the offset
(t//9 - 1, (t//3)%3 - 1, t%3 - 1) along (x, y, z).
"""
import numpy as np

CIN = [8, 16, 16, 32, 32, 64, 64, 32, 32, 16]
COUT = [16, 16, 32, 32, 64, 64, 32, 32, 16, 16]
CENTER = 13


def offsets():
    off_w, off_b, o = [], [], 0
    for ci, co in zip(CIN, COUT):
        off_w.append(o)
        o += 27 * ci * co
        off_b.append(o)
        o += co
    return off_w, off_b, o, o + 17


OFF_W, OFF_B, OFF_HEAD, COUNT = offsets()


def conv_weights(p, k):
    return p[OFF_W[k]:OFF_W[k] + 27 * CIN[k] * COUT[k]].reshape(27, COUT[k], CIN[k])


def conv_bias(p, k):
    return p[OFF_B[k]:OFF_B[k] + COUT[k]]


def head(p):
    return p[OFF_HEAD:OFF_HEAD + 16], p[OFF_HEAD + 16:OFF_HEAD + 17]


def tap_of(dx, dy, dz):
    return (dx + 1) * 9 + (dy + 1) * 3 + (dz + 1)


def random_params(seed, scale=1.0):
    """He-scaled normal weights, small random biases."""
    rng = np.random.default_rng(seed)
    p = np.zeros(COUNT)
    for k in range(10):
        fan = 27 * (4 if k == 0 else CIN[k])
        w = conv_weights(p, k)
        w[:] = rng.normal(0.0, scale * np.sqrt(2.0 / fan), w.shape)
        if k == 0:
            w[:, :, 4:] = 0.0
        conv_bias(p, k)[:] = rng.normal(0.0, 0.05, COUT[k])
    hw, hb = head(p)
    hw[:] = rng.normal(0.0, 0.25, 16)
    hb[:] = 0.01
    return p


def shift_params(dx, dy, dz, channel=0):
    """z = relu(x[channel] shifted by (dx,dy,dz)) through c0 and c1's channel 0,
    the level-0 skip (c8 zero) and c9: z[v] = relu(x[v + d]) with zero padding."""
    p = np.zeros(COUNT)
    conv_weights(p, 0)[tap_of(dx, dy, dz), 0, channel] = 1.0
    conv_weights(p, 1)[CENTER, 0, 0] = 1.0
    conv_weights(p, 9)[CENTER, 0, 0] = 1.0
    head(p)[0][0] = 1.0
    return p


def coarse_path_params():
    """Routes input channel 0 through both pools, the bottom level and both
    up-samplings on channel 0/1, with every skip carrying nothing the head
    reads: for x >= 0, z = nearest-up4(4x4x4 block means of x0)."""
    p = np.zeros(COUNT)
    c = CENTER
    conv_weights(p, 0)[c, 0, 0] = 1.0   # a0.0 = x0
    conv_weights(p, 1)[c, 0, 0] = 1.0   # e0.0 = a0.0
    conv_weights(p, 2)[c, 0, 0] = 1.0   # a1.0 = pool(e0).0
    conv_weights(p, 3)[c, 0, 0] = 1.0   # e1.0
    conv_weights(p, 4)[c, 0, 0] = 1.0   # a2.0 = pool(e1).0
    conv_weights(p, 5)[c, 0, 0] = 1.0   # e2.0
    conv_weights(p, 6)[c, 1, 0] = 1.0   # s1.1 = up(e2).0 (e1.1 = 0)
    conv_weights(p, 7)[c, 1, 1] = 1.0   # d1.1
    conv_weights(p, 8)[c, 1, 1] = 1.0   # s0.1 = up(d1).1 (e0.1 = 0)
    conv_weights(p, 9)[c, 1, 1] = 1.0   # d0.1
    head(p)[0][1] = 1.0
    return p


def block_mean_up(x, f):
    """Nearest up-sampling by f of the f^3 block means of x [B,X,Y,Z]."""
    B, X, Y, Z = x.shape
    m = x.reshape(B, X // f, f, Y // f, f, Z // f, f).mean(axis=(2, 4, 6))
    return m.repeat(f, 1).repeat(f, 2).repeat(f, 3)


def shifted(x, dx, dy, dz):
    """out[v] = x[v + d] with zeros outside, x [B,X,Y,Z]."""
    out = np.zeros_like(x)
    B, X, Y, Z = x.shape

    def rng(n, d):
        return slice(max(0, -d), min(n, n - d)), slice(max(0, d), min(n, n + d))

    (ox, ix), (oy, iy), (oz, iz) = rng(X, dx), rng(Y, dy), rng(Z, dz)
    out[:, ox, oy, oz] = x[:, ix, iy, iz]
    return out
