"""CPU FP64 restatement of the learned warm start's network (match-series_amd/csrc/ms_unet.hip):
forward pass, loss and backward pass in numpy. Test infrastructure only: tests/ compare the GPU's
gradients with it; nothing in the product imports it.

The network has no counterpart in the reference. This restatement is pinned by the mathematics
(tests/test_unet_oracle.py): its convolution against a brute-force loop, its backward pass against
central finite differences of its forward pass (exact arithmetic, bf16=False), and the gradient
of the loss with respect to the output against the closed form.

With bf16=True the values are rounded to BF16 (round to nearest even, through FP32) where the GPU
stores them in BF16: the packed input, every convolution's weights and output, every pooled tensor,
and the pre-activation gradient a convolution's backward pass multiplies with. Everything else is
FP64 (the GPU accumulates in FP32), so the comparison isolates accumulation-order effects.

Layout: activations (B, H, W, C) like the GPU's NHWC; the flat parameter vector as the C ABI's
(per convolution [9][Cout][Cin] weights with tap = 3 (dy + 1) + dx + 1, then Cout biases; then the
head's [2][C] weights and 2 biases).
"""
import numpy as np


def round_bf16(x):
    """FP64 -> FP32 -> BF16 (round to nearest even) -> FP64."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def graph(scales, base):
    """Activations [(scale, channels)], convolutions [dict], pools [(from, to)], head input."""
    acts = [(0, 16)]
    convs, pools, skip = [], [], []
    cur, ch = 0, 16
    for s in range(scales):
        co = base << s
        if s > 0:
            acts.append((s, ch))
            pools.append((cur, len(acts) - 1))
            cur = len(acts) - 1
        acts.append((s, co))
        convs.append(dict(in1=cur, in2=-1, up1=0, out=len(acts) - 1, cin1=ch, cin2=0, cout=co, cin_real=2 if s == 0 else ch))
        acts.append((s, co))
        convs.append(dict(in1=len(acts) - 2, in2=-1, up1=0, out=len(acts) - 1, cin1=co, cin2=0, cout=co, cin_real=co))
        skip.append(len(acts) - 1)
        cur, ch = len(acts) - 1, co
    for s in range(scales - 2, -1, -1):
        co = base << s
        acts.append((s, co))
        convs.append(dict(in1=cur, in2=skip[s], up1=1, out=len(acts) - 1, cin1=ch, cin2=co, cout=co, cin_real=ch + co))
        acts.append((s, co))
        convs.append(dict(in1=len(acts) - 2, in2=-1, up1=0, out=len(acts) - 1, cin1=co, cin2=0, cout=co, cin_real=co))
        cur, ch = len(acts) - 1, co
    off = 0
    for c in convs:
        cin = c["cin1"] + c["cin2"]
        c["woff"], off = off, off + 9 * c["cout"] * cin
        c["boff"], off = off, off + c["cout"]
    head = dict(input=cur, woff=off, boff=off + 2 * base)
    return acts, convs, pools, head, off + 2 * base + 2


def param_count(scales, base):
    return graph(scales, base)[4]


def _virtual(acts_y, c):
    x1 = acts_y[c["in1"]]
    if c["up1"]:
        x1 = np.repeat(np.repeat(x1, 2, axis=1), 2, axis=2)
    return x1 if c["in2"] < 0 else np.concatenate([x1, acts_y[c["in2"]]], axis=3)


def conv3x3(x, w, acc32=False):
    """y[b, y, x, o] = sum_tap sum_i w[tap, o, i] x[b, y + dy, x + dx, i] (zero outside), accumulated
    in FP64, or in FP32 with acc32 (another summation that is as valid for BF16 operands: the two
    differ by accumulation rounding only, which sets the noise floor of any comparison)."""
    B, H, W, _ = x.shape
    ft = np.float32 if acc32 else np.float64
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0))).astype(ft)
    y = np.zeros((B, H, W, w.shape[1]), dtype=ft)
    for t in range(9):
        dy, dx = t // 3 - 1, t % 3 - 1
        y += xp[:, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W, :] @ w[t].T.astype(ft)
    return y.astype(np.float64)


def forward(params, x_in, scales, base, bf16=True, acc32=False):
    """x_in: (B, 2, H, W) pairs. Returns (out (B, 2, H, W), cache)."""
    acts, convs, pools, head, _ = graph(scales, base)
    rnd = round_bf16 if bf16 else (lambda a: a)
    B, _, H, W = x_in.shape
    ys = [None] * len(acts)
    x0 = np.zeros((B, H, W, 16))
    x0[..., 0] = x_in[:, 0]
    x0[..., 1] = x_in[:, 1]
    ys[0] = rnd(x0)
    pi = 0
    for c in convs:
        while pi < len(pools) and pools[pi][1] == c["in1"]:
            src = ys[pools[pi][0]]
            ys[pools[pi][1]] = rnd(0.25 * ((src[:, 0::2, 0::2] + src[:, 0::2, 1::2]) + (src[:, 1::2, 0::2] + src[:, 1::2, 1::2])))
            pi += 1
        cin = c["cin1"] + c["cin2"]
        w = rnd(params[c["woff"]:c["woff"] + 9 * c["cout"] * cin].reshape(9, c["cout"], cin))
        b = params[c["boff"]:c["boff"] + c["cout"]]
        ys[c["out"]] = rnd(np.maximum(conv3x3(_virtual(ys, c), w, acc32) + b, 0.0))
    xh = ys[head["input"]]
    wh = params[head["woff"]:head["woff"] + 2 * base].reshape(2, base)
    bh = params[head["boff"]:head["boff"] + 2]
    out = np.einsum("bhwc,oc->bohw", xh, wh) + bh[None, :, None, None]
    return out, ys


def loss_and_grad(params, x_in, target, scales, base, bf16=True, acc32=False):
    """Loss 1/(2 n) sum (out - target)^2 and its gradient with respect to the flat parameters
    (acc32: the forward convolutions accumulated in FP32)."""
    acts, convs, pools, head, npar = graph(scales, base)
    rnd = round_bf16 if bf16 else (lambda a: a)
    out, ys = forward(params, x_in, scales, base, bf16, acc32)
    n = out.size
    e = out - target
    loss = 0.5 * np.sum(e * e) / n
    dout = e / n
    g = np.zeros(npar)
    dys = [np.zeros_like(y) for y in ys]
    xh = ys[head["input"]]
    wh = params[head["woff"]:head["woff"] + 2 * base].reshape(2, base)
    dys[head["input"]] += np.einsum("bohw,oc->bhwc", dout, wh)
    g[head["woff"]:head["woff"] + 2 * base] = np.einsum("bohw,bhwc->oc", dout, xh).ravel()
    g[head["boff"]:head["boff"] + 2] = dout.sum(axis=(0, 2, 3))
    pi = len(pools)
    for k in range(len(convs) - 1, -1, -1):
        c = convs[k]
        cin = c["cin1"] + c["cin2"]
        y = ys[c["out"]]
        dz_exact = dys[c["out"]] * (y > 0)
        g[c["boff"]:c["boff"] + c["cout"]] = dz_exact.sum(axis=(0, 1, 2))
        dz = rnd(dz_exact)
        xv = _virtual(ys, c)
        B, H, W, _ = xv.shape
        xp = np.pad(xv, ((0, 0), (1, 1), (1, 1), (0, 0)))
        dw = np.zeros((9, c["cout"], cin))
        for t in range(9):
            dy, dx = t // 3 - 1, t % 3 - 1
            dw[t] = np.einsum("bhwo,bhwi->oi", dz, xp[:, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W, :])
        g[c["woff"]:c["woff"] + dw.size] = dw.ravel()
        if k == 0:
            continue
        w = rnd(params[c["woff"]:c["woff"] + 9 * c["cout"] * cin].reshape(9, c["cout"], cin))
        dxp = np.zeros_like(xp)
        for t in range(9):
            dy, dx = t // 3 - 1, t % 3 - 1
            dxp[:, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W, :] += dz @ w[t]
        dxv = dxp[:, 1:H + 1, 1:W + 1, :]
        d1 = dxv[..., :c["cin1"]]
        if c["up1"]:
            d1 = (d1[:, 0::2, 0::2] + d1[:, 0::2, 1::2]) + (d1[:, 1::2, 0::2] + d1[:, 1::2, 1::2])
        dys[c["in1"]] += d1
        if c["in2"] >= 0:
            dys[c["in2"]] += dxv[..., c["cin1"]:]
        while pi > 0 and pools[pi - 1][1] == c["in1"]:
            pi -= 1
            dsrc = dys[pools[pi][1]]
            up = np.repeat(np.repeat(dsrc, 2, axis=1), 2, axis=2)
            dys[pools[pi][0]] += 0.25 * up
    return loss, g, out
