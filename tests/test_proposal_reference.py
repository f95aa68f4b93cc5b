"""Pins of the FP64 restatement of the learned band proposal (the checker
of its GPU kernels): central finite differences of the loss for the
backward pass, Adam's first step in closed form, and the network input's
2x2 means."""
import numpy as np
import pytest

from chiaro_amd.proposal import (ReferenceNet, adam_reference, from_bf16, layer_specs,
                                 quarter_input, quarter_means, to_bf16)


def _params(width, height, rng, scale=0.2):
    out = []
    for (cin, cout, *_r) in layer_specs(width, height):
        cp_in, cp_out = (cin + 15) // 16 * 16, (cout + 15) // 16 * 16
        w = np.zeros((9 * cp_in, cp_out))
        for t in range(9):
            w[t * cp_in:t * cp_in + cin, :cout] = rng.normal(size=(cin, cout)) * scale
        b = np.zeros(cp_out)
        b[:cout] = rng.normal(size=cout) * 0.1
        out += [w.reshape(-1), b]
    return np.concatenate(out)


def _loss(net, x, labels, frozen=None, label_count=129):
    """The training loss per pixel; the scale term's |t - c| (capped at 32
    labels) is the value `frozen` (the unperturbed error), as the gradient
    treats it."""
    acts = net.forward(x)
    out = acts[-1].reshape(-1, acts[-1].shape[-1])
    t = labels.reshape(-1).astype(np.float64)
    m = t > 0
    c = (label_count - 1) * 0.5 + (label_count - 1) * 0.125 * out[:, 0]
    s = np.clip(out[:, 1], -3.0, 5.0)
    z = out[:, 2]
    ad = np.abs(c - t)
    a = ad if frozen is None else frozen
    y = (a <= 6.0).astype(np.float64)
    loss = (ad * 0.125 + np.minimum(a, 32.0) * np.exp(-s) + s
            + (np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z)))))
    return float(loss[m].sum()) / labels.size, ad


@pytest.mark.parametrize("label_count", [129, 257])
def test_backward_matches_finite_differences(label_count):
    rng = np.random.default_rng(0)
    w, h = 7, 6
    params = _params(w, h, rng)
    x = np.zeros((2, h, w, 16))
    x[..., :3] = rng.uniform(0, 1, size=(2, h, w, 3))
    labels = rng.integers(1, label_count, size=(2, h, w)).astype(np.uint16)
    labels[0, 0, :3] = 0                      # unlabelled pixels contribute nothing
    net = ReferenceNet(w, h, params, bf16=False)
    acts = net.forward(x)
    _, dout = net.loss_grad(acts[-1], labels, label_count)
    grads = net.backward(acts, dout)
    base, frozen = _loss(net, x, labels, label_count=label_count)
    total, _ = net.loss_grad(acts[-1], labels, label_count)
    assert abs(total / labels.size - base) <= 1e-9 * abs(base)
    live = np.flatnonzero(np.abs(params) > 0)
    for i in rng.choice(live, size=40, replace=False):
        eps = 1e-5
        p1, p2 = params.copy(), params.copy()
        p1[i] += eps
        p2[i] -= eps
        fd = (_loss(ReferenceNet(w, h, p1, bf16=False), x, labels, frozen, label_count)[0] -
              _loss(ReferenceNet(w, h, p2, bf16=False), x, labels, frozen, label_count)[0]) / (2 * eps)
        assert abs(fd - grads[i]) <= 1e-5 + 1e-3 * abs(fd), (i, fd, grads[i])


def test_adam_first_step_closed_form():
    """With m = v = 0, step 1 moves each weight by lr * g / (|g| + eps)."""
    rng = np.random.default_rng(1)
    w = rng.normal(size=64).astype(np.float32)
    g = rng.normal(size=64).astype(np.float32)
    z = np.zeros(64, np.float32)
    w1, m1, v1 = adam_reference(w, g, z, z, 1, 1e-3)
    assert np.allclose(m1, 0.1 * g, rtol=1e-6) and np.allclose(v1, 0.001 * g * g, rtol=1e-5)
    assert np.allclose(w - w1, 1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-4)


def test_cell_means_of_four():
    """Network input on an 8 px grid: 4x4 cell means, summed row by row."""
    rng = np.random.default_rng(3)
    p = rng.random((17, 22)).astype(np.float32)
    got = quarter_means([p], 4)[..., 0]
    want = np.zeros((4, 5), np.float32)
    for j in range(4):
        for i in range(5):
            acc = np.float32(0.0)
            for dy in range(4):
                for dx in range(4):
                    acc = np.float32(acc + p[min(4 * j + dy, 16), min(4 * i + dx, 21)])
            want[j, i] = np.float32(1.0 / 16) * acc
    assert np.array_equal(got, want)


def test_bf16_round_trip_and_quarter_input():
    x = np.array([1.0, 1.00390625, 1.005859375, -2.5, 3.0e-3], np.float32)
    r = from_bf16(to_bf16(x))
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties go to the even 1.0
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125 and r[3] == -2.5
    p = np.arange(24, dtype=np.float32).reshape(4, 6)
    q = from_bf16(quarter_input([p, 2 * p]))
    assert q.shape == (2, 3, 16) and not q[..., 2:].any()
    want = (p[0, 0] + p[0, 1] + p[1, 0] + p[1, 1]) / 4 / 8
    assert q[0, 0, 0] == from_bf16(to_bf16(np.float32([want])))[0]
    assert q[1, 2, 1] == from_bf16(to_bf16(np.float32([2 * (p[2, 4] + p[2, 5] + p[3, 4] + p[3, 5]) / 32])))[0]
