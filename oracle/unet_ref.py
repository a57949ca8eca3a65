"""CPU oracle of the learned LSRTM preconditioner (U-Net) — TEST INFRASTRUCTURE ONLY.

The reference has no network; this file is the FP64 restatement of the U-Net
that libds2d.so trains and runs with BF16 MFMA kernels, written so that
tests/, __graft_entry__.smoke() and bench.py can check the GPU against it.
The product path never imports it.

Network (levels L, base channels C0, c_l = C0 * 2**l):
  encoder l = 0..L-1 : conv3x3(cin -> c_l) + ReLU, conv3x3(c_l -> c_l) + ReLU,
                       cin = 1 at l = 0, else c_{l-1}; 2x2 average pool between levels
  decoder l = L-2..0 : input concat(skip e_l, nearest-2x upsample of d_{l+1}),
                       conv3x3(c_l + c_{l+1} -> c_l) + ReLU, conv3x3(c_l -> c_l) + ReLU
  head               : 1x1 conv c_0 -> 1 (+ bias), no activation
  loss               : mean over pixels of (y - t)^2
Tensors are NHWC (H = x, W = z, C innermost); 3x3 weights are [tap][Cout][Cin]
with tap = (dx + 1) * 3 + (dz + 1), zero padding 1.  Parameters are flattened
in the order enc0a, enc0b, ..., enc{L-1}b, dec{L-2}a, dec{L-2}b, ..., dec0b,
head, each weight followed by its bias.

With bf16=True the forward/backward round to bfloat16 (round-to-nearest-even)
exactly where the GPU stores BF16: 3x3 weights of every layer but the first,
every post-ReLU activation, pooled maps and the gradient maps fed to the
MFMA kernels; accumulation stays FP64 here (FP32 on the GPU).

Parity pins (tests/test_unet_oracle.py): central finite differences of the
FP64 loss against every parameter block; conv3x3 against scipy.signal
correlate; pool/upsample adjointness <Ax, y> = <x, A^T y>.
"""
import numpy as np

TAPS = [(dx, dz) for dx in (-1, 0, 1) for dz in (-1, 0, 1)]


def bf16(a):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def layers(levels=4, base=32):
    """[(name, cin, cout, k)] in parameter order; k = 3 (3x3) or 1 (head)."""
    c = [base << l for l in range(levels)]
    out = []
    for l in range(levels):
        out.append((f"enc{l}a", 1 if l == 0 else c[l - 1], c[l], 3))
        out.append((f"enc{l}b", c[l], c[l], 3))
    for l in range(levels - 2, -1, -1):
        out.append((f"dec{l}a", c[l] + c[l + 1], c[l], 3))
        out.append((f"dec{l}b", c[l], c[l], 3))
    out.append(("head", c[0], 1, 1))
    return out


def num_params(levels=4, base=32):
    return sum((9 if k == 3 else 1) * ci * co + co for _, ci, co, k in layers(levels, base))


def unflatten(theta, levels=4, base=32):
    p, o = {}, 0
    for name, ci, co, k in layers(levels, base):
        nw = (9 if k == 3 else 1) * co * ci
        w = theta[o:o + nw]
        p[name + ".w"] = w.reshape(9, co, ci) if k == 3 else w.reshape(co, ci)
        o += nw
        p[name + ".b"] = theta[o:o + co]
        o += co
    assert o == theta.size
    return p


def flatten(p, levels=4, base=32):
    return np.concatenate([np.concatenate([p[n + ".w"].ravel(), p[n + ".b"].ravel()])
                           for n, _, _, _ in layers(levels, base)])


def _pad(x):
    return np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))


def conv3x3(x, w, b=None):
    """y[n,h,w,co] = sum_tap,ci x[n,h+dx,w+dz,ci] * w[tap,co,ci] (+ b[co])."""
    n, h, wd, _ = x.shape
    xp = _pad(x)
    y = np.zeros((n, h, wd, w.shape[1]))
    for t, (dx, dz) in enumerate(TAPS):
        y += xp[:, 1 + dx:1 + dx + h, 1 + dz:1 + dz + wd, :] @ w[t].T
    return y if b is None else y + b


def conv3x3_dgrad(dy, w):
    """Adjoint of conv3x3 in x: dx[p] = sum_tap dy[p - off(tap)] @ w[tap]."""
    n, h, wd, _ = dy.shape
    dp = _pad(dy)
    dx = np.zeros((n, h, wd, w.shape[2]))
    for t, (ox, oz) in enumerate(TAPS):
        dx += dp[:, 1 - ox:1 - ox + h, 1 - oz:1 - oz + wd, :] @ w[t]
    return dx


def conv3x3_wgrad(x, dy):
    """dw[tap,co,ci] = sum_p dy[p,co] * x[p + off(tap), ci]."""
    n, h, wd, ci = x.shape
    xp = _pad(x)
    dy2 = dy.reshape(-1, dy.shape[-1])
    dw = np.zeros((9, dy.shape[-1], ci))
    for t, (dx, dz) in enumerate(TAPS):
        dw[t] = dy2.T @ xp[:, 1 + dx:1 + dx + h, 1 + dz:1 + dz + wd, :].reshape(-1, ci)
    return dw


def pool2(x):
    n, h, w, c = x.shape
    return x.reshape(n, h // 2, 2, w // 2, 2, c).mean(axis=(2, 4))


def pool2_bwd(dy):
    return np.repeat(np.repeat(dy, 2, axis=1), 2, axis=2) / 4.0


def up2(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up2_bwd(dy):
    n, h, w, c = dy.shape
    return dy.reshape(n, h // 2, 2, w // 2, 2, c).sum(axis=(2, 4))


def forward(theta, x, levels=4, base=32, bf=False):
    """x: [N,H,W] float.  Returns (y [N,H,W], cache)."""
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    p = unflatten(np.asarray(theta, dtype=np.float64), levels, base)
    W = {n: (p[n + ".w"] if (not bf or n in ("enc0a", "head")) else bf16(p[n + ".w"]))
         for n, _, _, _ in layers(levels, base)}
    a = {}
    inp = np.asarray(x, dtype=np.float64)[..., None]
    cur = inp
    for l in range(levels):
        a[f"in{l}a"] = cur
        a[f"enc{l}a"] = r(np.maximum(conv3x3(cur, W[f"enc{l}a"], p[f"enc{l}a.b"]), 0))
        a[f"enc{l}b"] = r(np.maximum(conv3x3(a[f"enc{l}a"], W[f"enc{l}b"], p[f"enc{l}b.b"]), 0))
        if l < levels - 1:
            cur = r(pool2(a[f"enc{l}b"]))
            a[f"pool{l}"] = cur
    low = a[f"enc{levels - 1}b"]
    for l in range(levels - 2, -1, -1):
        cat = np.concatenate([a[f"enc{l}b"], up2(low)], axis=-1)
        a[f"cat{l}"] = cat
        a[f"dec{l}a"] = r(np.maximum(conv3x3(cat, W[f"dec{l}a"], p[f"dec{l}a.b"]), 0))
        a[f"dec{l}b"] = r(np.maximum(conv3x3(a[f"dec{l}a"], W[f"dec{l}b"], p[f"dec{l}b.b"]), 0))
        low = a[f"dec{l}b"]
    y = low @ p["head.w"][0] + p["head.b"][0]
    return y, (p, W, a, levels, base, bf)


def loss_grad(theta, x, t, levels=4, base=32, bf=False):
    """MSE loss and its gradient with respect to the flattened parameters."""
    y, (p, W, a, L, C0, bfl) = forward(theta, x, levels, base, bf)
    r = bf16 if bfl else (lambda v: v)
    P = y.size
    diff = y - np.asarray(t, dtype=np.float64)
    loss = float((diff * diff).sum() / P)
    dy = 2.0 * diff / P
    g = {}
    top = a["dec0b"] if L > 1 else a["enc0b"]
    g["head.w"] = (dy[..., None] * top).sum(axis=(0, 1, 2))[None, :]
    g["head.b"] = np.array([dy.sum()])
    dz = r(dy[..., None] * p["head.w"][0] * (top > 0))

    def conv_back(name, inp, dzv):
        g[name + ".w"] = conv3x3_wgrad(inp, dzv)
        g[name + ".b"] = dzv.sum(axis=(0, 1, 2))
        return conv3x3_dgrad(dzv, W[name])

    dskip = {}
    for l in range(0, L - 1):
        # decoder level l, walked from the top (l = 0) down
        d_in = conv_back(f"dec{l}b", a[f"dec{l}a"], dz)
        dz = r(d_in * (a[f"dec{l}a"] > 0))
        d_cat = conv_back(f"dec{l}a", a[f"cat{l}"], dz)
        cl = a[f"enc{l}b"].shape[-1]
        dskip[l] = d_cat[..., :cl]
        low = a[f"dec{l + 1}b"] if l + 1 < L - 1 else a[f"enc{L - 1}b"]
        dz = r(up2_bwd(d_cat[..., cl:]) * (low > 0))
    for l in range(L - 1, -1, -1):
        d_in = conv_back(f"enc{l}b", a[f"enc{l}a"], dz)
        dz = r(d_in * (a[f"enc{l}a"] > 0))
        if l == 0:
            g["enc0a.w"] = conv3x3_wgrad(a["in0a"], dz)
            g["enc0a.b"] = dz.sum(axis=(0, 1, 2))
            break
        dp = conv_back(f"enc{l}a", a[f"in{l}a"], dz)
        dz = r((pool2_bwd(dp) + dskip[l - 1]) * (a[f"enc{l - 1}b"] > 0))
    return loss, flatten(g, L, C0), y


def init_params(levels=4, base=32, seed=0):
    """He-uniform weights, zero biases (host-side initialisation of the tests)."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, ci, co, k in layers(levels, base):
        fan = ci * (9 if k == 3 else 1)
        lim = np.sqrt(6.0 / fan)
        p[name + ".w"] = rng.uniform(-lim, lim, size=(9, co, ci) if k == 3 else (co, ci))
        p[name + ".b"] = np.zeros(co)
    return flatten(p, levels, base)
