"""Pins of the MLP oracle: finite differences fix the FP64 backward, torch's
bfloat16 cast fixes the rounding, torch.optim.Adam fixes the Adam step."""
import numpy as np
import pytest

import mlp_oracle as M


def _problem(D=5, O=2, n=7, seed=0):
    rng = np.random.default_rng(seed)
    lay = M.layout(D, O)
    # weights ~ N(0, 1/fan_in) keep outputs and the loss O(1), so the
    # differences' round-off stays far below the tolerance
    p = rng.normal(0, 0.1, lay["total"])
    for name, fan in (("W1", D), ("W2", M.H), ("W3", M.H)):
        o, shape = lay[name]
        size = int(np.prod(shape))
        p[o:o + size] = rng.normal(0, 1.0 / np.sqrt(fan), size)
    x = rng.normal(0, 1, (n, D)).astype(np.float32)
    t = rng.normal(0, 2, (n, O))
    xmean = x.mean(0)
    xinv = 1.0 / x.std(0)
    return p, x, t, xmean, xinv, t.mean(0), float(t.std())


def test_fp64_gradient_matches_central_differences():
    p, x, t, xm, xi, ym, ys = _problem()
    loss, g = M.forward_backward(p, x, t, xm, xi, ym, ys)
    rng = np.random.default_rng(1)
    lay = M.layout(5, 2)
    idx = np.concatenate([rng.choice(lay["total"], 60, replace=False),
                          np.arange(lay["b3"][0], lay["total"]), lay["W1"][0] + np.arange(5)])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        num = (M.forward_backward(pp, x, t, xm, xi, ym, ys)[0] - M.forward_backward(pm, x, t, xm, xi, ym, ys)[0]) / (2 * h)
        assert num == pytest.approx(g[i], rel=1e-5, abs=1e-8), i
    assert loss > 0


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(2)
    v = np.concatenate([rng.normal(0, 1, 5000), rng.normal(0, 1e-30, 100), [0.0, -0.0, 1.0, 1.0 + 2 ** -8,
                        1.0 + 2 ** -8 + 2 ** -20, 1.0 + 3 * 2 ** -8, 3.4e38, -3.4e38, 1e-40]]).astype(np.float32)
    want = torch.from_numpy(v).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(M.bf16(v).view(np.uint32), want.view(np.uint32))
    assert M.bf16(np.float32([1.0 + 2 ** -8]))[0] == 1.0  # tie goes to even


def test_bf16_points_are_a_small_perturbation():
    p, x, t, xm, xi, ym, ys = _problem(D=40, n=300, seed=3)
    l64, g64 = M.forward_backward(p, x, t, xm, xi, ym, ys)
    l16, g16 = M.forward_backward(p, x, t, xm, xi, ym, ys, bf16_points=True)
    assert l16 == pytest.approx(l64, rel=2e-2)
    assert np.linalg.norm(g16 - g64) / np.linalg.norm(g64) < 3e-2


def test_adam_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(4)
    w0 = rng.normal(size=1000).astype(np.float32)
    grads = [rng.normal(size=1000).astype(np.float32) for _ in range(3)]
    tw = torch.nn.Parameter(torch.from_numpy(w0.copy()))
    opt = torch.optim.Adam([tw], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    w, m, v = w0, np.zeros_like(w0), np.zeros_like(w0)
    for t, g in enumerate(grads, start=1):
        tw.grad = torch.from_numpy(g.copy())
        opt.step()
        w, m, v = M.adam(w, m, v, g, t)
    np.testing.assert_allclose(w, tw.detach().numpy(), rtol=1e-6, atol=5e-8)  # torch orders the terms differently
