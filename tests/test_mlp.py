"""The MLP continuation regressor: its CPU oracle pinned against torch
(bf16 conversion, tanh GELU, autograd, Adam) and central differences; the
GPU forward, gradient and Adam kernels checked against the oracle; and the
config-5 regressor checked against the polynomial LSMC and the FD judge."""

import ctypes as C

import numpy as np
import pytest
import torch

from oracle import mlp as om

F = C.POINTER(C.c_float)


def random_params(seed=0, scale=1.0):
    g = np.random.default_rng(seed)
    p = np.zeros(om.COUNT, np.float32)
    p[om.W1:om.B1] = g.uniform(-1, 1, 192) * 0.6 * scale
    p[om.B1:om.W2] = g.uniform(-0.3, 0.3, 64)
    p[om.W2:om.B2] = g.uniform(-1, 1, 4096) * 0.22 * scale
    p[om.B2:om.W3] = g.uniform(-0.3, 0.3, 64)
    p[om.W3:om.B3] = g.uniform(-1, 1, 64) * 0.3 * scale
    p[om.B3] = 0.1
    return p


def random_samples(n, seed=1):
    g = np.random.default_rng(seed)
    x = np.empty((n, 4), np.float32)
    x[:, 0] = g.uniform(-2.0, 0.0, n)
    x[:, 1] = g.uniform(-1.0, 1.0, n)
    x[:, 2] = g.uniform(-1.0, 1.0, n)
    x[:, 3] = np.maximum(-x[:, 0], 0) + g.normal(0, 0.3, n)
    return x


def torch_net(p):
    w1, b1, w2, b2, w3, b3 = (torch.tensor(np.asarray(a, np.float64)) for a in om.unpack(p))
    b3 = torch.tensor([float(b3)], dtype=torch.float64)
    params = [t.clone().requires_grad_(True) for t in (w1, b1, w2, b2, w3, b3)]

    def run(x):
        gelu = lambda z: torch.nn.functional.gelu(z, approximate="tanh")
        h1 = gelu(x[:, :3] @ params[0].T + params[1])
        h2 = gelu(h1 @ params[2].T + params[3])
        return h2 @ params[4] + params[5]
    return params, run


# ------------------------------------------------------------ oracle pins
def test_bf16_rounding_is_torch_bfloat16():
    v = np.concatenate([np.random.default_rng(3).normal(0, 10, 20000).astype(np.float32),
                        np.float32([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8, 3.0e-38, 65504.0])])
    ref = torch.tensor(v).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(om.bf16(v), ref)


def test_gelu_is_torch_tanh_gelu_and_its_derivative():
    x = np.linspace(-8, 8, 4001)
    g, d = om.gelu(x)
    t = torch.tensor(x, requires_grad=True)
    y = torch.nn.functional.gelu(t, approximate="tanh")
    y.sum().backward()
    assert np.allclose(g, y.detach().numpy(), rtol=0, atol=1e-14)
    assert np.allclose(d, t.grad.numpy(), rtol=0, atol=1e-12)


def test_oracle_gradient_is_torch_autograd():
    p, x = random_params(), random_samples(300).astype(np.float64)
    grad, loss = om.forward_backward(p, x, bf16_ops=False)
    params, run = torch_net(p)
    xt = torch.tensor(x)
    l = ((run(xt) - xt[:, 3]) ** 2).sum() / len(x)
    l.backward()
    ref = np.concatenate([params[0].grad.numpy().ravel(), params[1].grad.numpy(), params[2].grad.numpy().ravel(),
                          params[3].grad.numpy(), params[4].grad.numpy(), params[5].grad.numpy()])
    assert loss == pytest.approx(l.item(), rel=1e-12)
    assert np.allclose(grad, ref, rtol=1e-10, atol=1e-13)


def test_oracle_gradient_by_central_differences():
    p, x = random_params(5).astype(np.float64), random_samples(64, 6).astype(np.float64)
    grad, _ = om.forward_backward(p, x, bf16_ops=False)
    idx = np.random.default_rng(7).choice(om.COUNT, 60, replace=False).tolist() + [om.B3, om.W1, om.W2 + 65]
    for k in idx:
        e = np.zeros(om.COUNT)
        e[k] = 1e-6
        up = om.forward_backward(p + e, x, bf16_ops=False)[1]
        down = om.forward_backward(p - e, x, bf16_ops=False)[1]
        assert grad[k] == pytest.approx((up - down) / 2e-6, rel=1e-5, abs=1e-9), k


def test_adam_is_torch_adam():
    p0, x = random_params(9), random_samples(200, 10)
    p, m, v = p0.copy(), np.zeros(om.COUNT, np.float32), np.zeros(om.COUNT, np.float32)
    tp = torch.tensor(p0, requires_grad=True)
    opt = torch.optim.Adam([tp], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    for step in range(1, 6):
        grad, _ = om.forward_backward(p, x)
        p, m, v = om.adam(p, m, v, grad, step, 1e-3)
        tp.grad = torch.tensor(grad.astype(np.float32))
        opt.step()
    assert np.allclose(p, tp.detach().numpy(), rtol=0, atol=2e-7)


def test_first_adam_step_moves_each_weight_by_the_learning_rate():
    g = np.random.default_rng(2).normal(0, 1, om.COUNT).astype(np.float32)
    p, _, _ = om.adam(np.zeros(om.COUNT, np.float32), np.zeros(om.COUNT, np.float32),
                      np.zeros(om.COUNT, np.float32), g, 1, 1e-3)
    # Bias correction makes the first step -lr g / (|g| + eps): +-lr unless |g| is near eps.
    g64 = g.astype(np.float64)
    assert np.allclose(p, -1e-3 * g64 / (np.abs(g64) + 1e-8), rtol=1e-5, atol=0)
    assert np.mean(np.abs(np.abs(p) - 1e-3) < 1e-6) > 0.99


# ---------------------------------------------------------------- GPU
def as_ptr(a):
    return a.ctypes.data_as(F)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 37, 1000, 4096])
def test_gpu_forward_matches_the_oracle(gpu, n):
    from kiyosi_amd import _native
    p, x = random_params(11), random_samples(n, 12)
    out = np.zeros(n, np.float32)
    gpu.check(_native.lib().kg_mlp_forward(gpu.handle, as_ptr(p), as_ptr(x), n, as_ptr(out)))
    ref, _ = om.forward(p, x.astype(np.float64))
    assert np.allclose(out, ref, rtol=2e-3, atol=2e-3), np.abs(out - ref).max()


@pytest.mark.gpu
@pytest.mark.parametrize("n", [37, 1000, 4096])
def test_gpu_gradient_matches_the_oracle(gpu, n):
    from kiyosi_amd import _native
    p, x = random_params(13), random_samples(n, 14)
    grad = np.zeros(om.COUNT, np.float32)
    loss = C.c_double()
    gpu.check(_native.lib().kg_mlp_gradient(gpu.handle, as_ptr(p), as_ptr(x), n, as_ptr(grad), C.byref(loss)))
    ref, ref_loss = om.forward_backward(p, x.astype(np.float64))
    assert loss.value == pytest.approx(ref_loss, rel=2e-3)
    for name, a, b in (("W1", om.W1, om.B1), ("b1", om.B1, om.W2), ("W2", om.W2, om.B2), ("b2", om.B2, om.W3),
                       ("W3", om.W3, om.B3), ("b3", om.B3, om.COUNT)):
        err = np.linalg.norm(grad[a:b] - ref[a:b]) / max(np.linalg.norm(ref[a:b]), 1e-12)
        assert err < 5e-3, (name, err)


@pytest.mark.gpu
def test_gpu_adam_steps_match_the_oracle(gpu):
    from kiyosi_amd import _native
    p0, x = random_params(15), random_samples(1000, 16)
    p = p0.copy()
    gpu.check(_native.lib().kg_mlp_train(gpu.handle, as_ptr(p), as_ptr(x), len(x), 256, 6, 1e-3))
    ref, _, _ = om.train(p0, x, 256, 6, 1e-3)
    moved = np.abs(ref - p0)
    diff = np.abs(p - ref)
    assert moved.max() > 1e-3
    # A parameter whose gradient is close to zero can take an Adam step of the
    # other sign, so a small fraction may differ by up to 2 lr per step.
    assert np.mean(diff < 1e-4) > 0.97, np.sort(diff)[-10:]
    assert diff.max() <= 2 * 6 * 1e-3 + 1e-6


@pytest.mark.gpu
def test_gpu_training_lowers_the_loss(gpu):
    from kiyosi_amd import _native
    p, x = random_params(17), random_samples(8192, 18)
    before = C.c_double()
    g = np.zeros(om.COUNT, np.float32)
    gpu.check(_native.lib().kg_mlp_gradient(gpu.handle, as_ptr(p), as_ptr(x), len(x), as_ptr(g), C.byref(before)))
    gpu.check(_native.lib().kg_mlp_train(gpu.handle, as_ptr(p), as_ptr(x), len(x), 1024, 400, 1e-3))
    after = C.c_double()
    gpu.check(_native.lib().kg_mlp_gradient(gpu.handle, as_ptr(p), as_ptr(x), len(x), as_ptr(g), C.byref(after)))
    assert after.value < 0.5 * before.value


@pytest.mark.gpu
def test_config5_mlp_matches_poly4_beats_quadratic_and_meets_the_fd_judge(gpu, oracle):
    """American put under the config-5 local-volatility surface, 2^20 training
    and 2^20 independent pricing paths, FP32 stepping. The MLP exercise rule
    priced out of sample is a lower bound of the 50-date Bermudan price that
    (a) does not fall below the degree-4 polynomial rule by more than half a
    standard error, (b) gains over the reference's quadratic regression, and
    (c) lies within one standard error, plus the 0.01 gap between 50 exercise
    dates and continuous exercise, of the FD American price."""
    from conftest import fd_richardson
    from kiyosi_amd import workloads
    from kiyosi_amd.engines import MonteCarloVanillaEngine
    option, context, _, training = workloads.american_put_local_vol()
    engine = MonteCarloVanillaEngine(path_count=1 << 20, step_count=50, seed=1729, precision="fp32")

    def report(basis):
        return engine.american_report(option, context, training_path_count=training, basis=basis)
    mlp, poly, quad = report("mlp"), report("poly4"), report("quadratic")
    fd = fd_richardson(oracle, context.model_parameters.surface._struct(), dividend=0.0)
    se = mlp.standard_error
    print(f"mlp {mlp.price:.5f} poly4 {poly.price:.5f} quadratic {quad.price:.5f} fd {fd:.5f} se {se:.5f} "
          f"gain over quadratic {mlp.price - quad.price:+.5f} over poly4 {mlp.price - poly.price:+.5f}")
    assert mlp.dates_fitted == 49
    assert mlp.price >= poly.price - 0.5 * se
    assert mlp.price > quad.price
    assert mlp.price <= fd + 2 * se
    assert fd - mlp.price <= se + 0.01


@pytest.mark.gpu
def test_config5_mlp_against_poly4_with_a_paired_error(gpu):
    """MLP minus poly-4 out-of-sample prices over six seeds: each seed prices
    both rules on the same fresh pricing paths, so the seed-to-seed spread
    of the difference is its paired (common-random-number) error, far below
    either price's own error. The MLP must not be significantly below
    poly-4."""
    import math

    import numpy as np
    from kiyosi_amd import workloads
    from kiyosi_amd.engines import MonteCarloVanillaEngine
    option, context, _, training = workloads.american_put_local_vol()
    diffs, marginal = [], []
    for seed in range(101, 107):
        engine = MonteCarloVanillaEngine(path_count=1 << 20, step_count=50, seed=seed, precision="fp32")
        mlp = engine.american_report(option, context, training_path_count=training, basis="mlp")
        poly = engine.american_report(option, context, training_path_count=training, basis="poly4")
        diffs.append(mlp.price - poly.price)
        marginal.append(mlp.standard_error)
    d = np.array(diffs)
    paired = d.std(ddof=1) / math.sqrt(len(d))
    print(f"MLP - poly4 over {len(d)} seeds: {d.mean():+.5f} +- {paired:.5f} (paired); "
          f"one price's error {np.mean(marginal):.5f}")
    assert paired < 0.25 * np.mean(marginal)
    assert d.mean() > -3.0 * paired
