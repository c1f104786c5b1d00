"""Pins the CPU restatement of the warm-start network (oracle/unet_ref.py) by the mathematics: its
3x3 convolution against a brute-force loop, its network graph against the parameter count of the
stated architecture, and its backward pass (exact arithmetic) against central finite differences of
its forward pass for parameters of every layer."""
import numpy as np

import unet_ref as U


def test_conv3x3_brute_force():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 5, 6, 3))
    w = rng.standard_normal((9, 4, 3))
    y = U.conv3x3(x, w)
    ref = np.zeros((2, 5, 6, 4))
    for b in range(2):
        for yy in range(5):
            for xx in range(6):
                for o in range(4):
                    a = 0.0
                    for t in range(9):
                        sy, sx = yy + t // 3 - 1, xx + t % 3 - 1
                        if 0 <= sy < 5 and 0 <= sx < 6:
                            a += np.dot(w[t, o], x[b, sy, sx])
                    ref[b, yy, xx, o] = a
    np.testing.assert_allclose(y, ref, rtol=1e-12, atol=1e-12)


def test_graph_parameter_count():
    # 4 scales, 32 base channels: 3x3 convolutions 2(pad 16)->32->32, 32->64->64, 64->128->128,
    # 128->256->256, (256+128)->128->128, (128+64)->64->64, (64+32)->32->32, 1x1 head 32->2
    pairs = [(16, 32), (32, 32), (32, 64), (64, 64), (64, 128), (128, 128), (128, 256), (256, 256),
             (384, 128), (128, 128), (192, 64), (64, 64), (96, 32), (32, 32)]
    n = sum(9 * ci * co + co for ci, co in pairs) + 2 * 32 + 2
    assert U.param_count(4, 32) == n
    real = n - 9 * 14 * 32  # without the zero weights of the padded input channels
    assert 1.9e6 < real < 2.0e6  # "about 2M parameters"


def test_backward_matches_finite_differences():
    scales, base = 2, 32
    rng = np.random.default_rng(7)
    acts, convs, pools, head, npar = U.graph(scales, base)
    p = np.zeros(npar)
    for c in convs:
        cin = c["cin1"] + c["cin2"]
        w = rng.standard_normal((9, c["cout"], cin)) * np.sqrt(2.0 / (9 * c["cin_real"]))
        if c["in1"] == 0:
            w[:, :, c["cin_real"]:] = 0.0
        p[c["woff"]:c["woff"] + w.size] = w.ravel()
        p[c["boff"]:c["boff"] + c["cout"]] = 0.05 * rng.standard_normal(c["cout"])
    p[head["woff"]:head["woff"] + 2 * base] = rng.standard_normal(2 * base) / np.sqrt(base)
    p[head["boff"]:head["boff"] + 2] = [0.1, -0.2]
    x = rng.random((1, 2, 64, 64))
    tgt = rng.standard_normal((1, 2, 64, 64))
    loss, g, _ = U.loss_and_grad(p, x, tgt, scales, base, bf16=False)
    # parameters of every convolution (weights of each tap region and biases) and of the head
    idx = []
    for c in convs:
        cin = c["cin1"] + c["cin2"]
        for t in (0, 4, 8):
            idx.append(c["woff"] + (t * c["cout"] + rng.integers(c["cout"])) * cin + rng.integers(c["cin_real"]))
        idx.append(c["boff"] + rng.integers(c["cout"]))
    idx += [head["woff"] + 3, head["woff"] + base + 5, head["boff"], head["boff"] + 1]
    eps = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += eps
        lp = U.loss_and_grad(q, x, tgt, scales, base, bf16=False)[0]
        q[i] -= 2 * eps
        lm = U.loss_and_grad(q, x, tgt, scales, base, bf16=False)[0]
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(abs(fd), 1e-3), (i, fd, g[i])
