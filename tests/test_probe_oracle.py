"""The CPU oracle of the learned probe model, pinned before it judges the GPU:
its fp64 backpropagation against central differences of its own loss, its
bf16 rounding against torch's bfloat16 conversion, and its Adam step against
the published update written out for one parameter."""
import numpy as np

import probe_oracle as po


def _model(seed):
    r = np.random.default_rng(seed)
    w = np.zeros(po.N_PARAMS, np.float32)
    w[po.O_W1:po.O_B1] = r.uniform(-0.25, 0.25, po.NH * po.NI)
    w[po.O_B1:po.O_W2] = r.uniform(-0.1, 0.1, po.NH)
    w[po.O_W2:po.O_B2] = r.uniform(-0.15, 0.15, po.NH * po.NH)
    w[po.O_B2:po.O_W3] = r.uniform(-0.1, 0.1, po.NH)
    w[po.O_W3:po.O_B3] = r.uniform(-0.15, 0.15, po.NH)
    w[po.O_B3] = 0.3
    X = r.standard_normal((64, po.NI)).astype(np.float32)
    t = r.uniform(0, 8, 64).astype(np.float32)
    return w, X, t


def test_backprop_matches_central_differences():
    w, X, t = _model(1)
    g = po.grads(w, X, t, bf16_points=False)
    r = np.random.default_rng(2)
    picks = np.concatenate([r.choice(np.arange(a, b), 6, replace=False) for a, b in
                            ((po.O_W1, po.O_B1), (po.O_B1, po.O_W2), (po.O_W2, po.O_B2), (po.O_B2, po.O_W3),
                             (po.O_W3, po.O_B3))] + [[po.O_B3]])
    w64 = w.astype(np.float64)
    for i in picks:
        h = 1e-6
        wp, wm = w64.copy(), w64.copy()
        wp[i] += h
        wm[i] -= h
        fd = (po.loss(wp, X, t) - po.loss(wm, X, t)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


def test_bf16_rounding_matches_torch():
    import torch

    r = np.random.default_rng(3)
    a = np.concatenate([r.standard_normal(20000).astype(np.float32) * 10.0 ** r.integers(-6, 6, 20000),
                        np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -9, -1 - 2 ** -8, 0.0, -0.0], np.float32)])
    want = torch.from_numpy(a).to(torch.bfloat16).to(torch.float32).numpy()
    got = po.bf16(a)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert got[-6] == 1.0 and got[-5] == 1.0 and got[-4] == 1 + 2 ** -7  # ties to even; above the tie rounds up


def test_bf16_points_change_gradients_by_little():
    w, X, t = _model(4)
    a = po.grads(w, X, t, bf16_points=False)
    b = po.grads(w, X, t, bf16_points=True)
    for lo, hi in ((po.O_W1, po.O_B1), (po.O_W2, po.O_B2), (po.O_W3, po.O_B3)):
        rel = np.linalg.norm(a[lo:hi] - b[lo:hi]) / np.linalg.norm(a[lo:hi])
        assert 0 < rel < 0.15, (lo, rel)


def test_adam_step_written_out():
    g = np.array([0.5, -2.0, 1e-3])
    w = np.array([1.0, 1.0, 1.0])
    w1, m1, v1 = po.adam_step(w, np.zeros(3), np.zeros(3), g, 1, 0.01)
    # first step: m_hat = g, v_hat = g^2, so each weight moves by lr * sign(g) (eps aside)
    assert np.allclose(w1, w - 0.01 * g / (np.abs(g) + 1e-8))
    w2, m2, v2 = po.adam_step(w1, m1, v1, g, 2, 0.01)
    mh = (0.9 * 0.1 * g + 0.1 * g) / (1 - 0.81)
    vh = (0.999 * 0.001 * g * g + 0.001 * g * g) / (1 - 0.999 ** 2)
    assert np.allclose(w2, w1 - 0.01 * mh / (np.sqrt(vh) + 1e-8))
