"""Learned cost model: the FP64 reference is pinned by central finite
differences; the HIP forward/backward kernels are checked against it on the
GPU; features are checked on a small design."""
import numpy as np
import pytest

import mlp_ref

F = 6


def _params(seed=0):
    rng = np.random.default_rng(seed)
    n = 64 * F + 64 + 64 * 64 + 64 + 64 + 1
    return rng.normal(0, 0.3, n)


def test_reference_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    p = _params(1)
    x = rng.normal(0, 1, (7, F))
    t = rng.normal(0, 1, 7)
    g = mlp_ref.gradients(p, x, t, F)
    eps = 1e-6
    for i in rng.choice(len(p), 60, replace=False):
        d = np.zeros_like(p)
        d[i] = eps
        num = (mlp_ref.loss(p + d, x, t, F) - mlp_ref.loss(p - d, x, t, F)) / (2 * eps)
        assert abs(num - g[i]) <= 1e-6 + 1e-5 * abs(num), (i, num, g[i])


def test_reference_forward_closed_form():
    # with W2 = I, b = 0 and W3 = 1, the network computes sum(relu(relu(W1 x)))
    p = np.zeros(64 * F + 64 + 64 * 64 + 64 + 64 + 1)
    W1 = np.arange(64 * F).reshape(64, F) / 1000.0 - 0.1
    p[:64 * F] = W1.ravel()
    o = 64 * F + 64
    p[o:o + 64 * 64] = np.eye(64).ravel()
    p[o + 64 * 64 + 64:o + 64 * 64 + 128] = 1.0
    x = np.ones((2, F))
    want = np.maximum(x @ W1.T, 0).sum(1)
    assert np.allclose(mlp_ref.forward(p, x, F), want)


def test_features_of_a_small_design():
    from rtlflow_amd.compiler import Compiled, Options
    from rtlflow_amd.costmodel import FeatureExtractor
    from rtlflow_amd.verilog import elaborate
    src = """
    module f(input clk, input [7:0] a, input [127:0] w, output [7:0] o, output [127:0] q);
        reg [7:0] r; reg [127:0] rw;
        wire [7:0] t = a[0] ? a + r : a - r;
        always @(posedge clk) begin r <= t; rw <= rw ^ w; end
        assign o = r;
        assign q = rw + w;
    endmodule"""
    comp = Compiled(elaborate(src), Options(threads=1))
    fx = FeatureExtractor(comp)
    allv = list(range(len(comp.sch.vertices)))
    instr, fan_in, fan_out, state, br, wide = fx.raw(allv)
    assert instr > 10 and br >= 1 and wide >= 2
    assert fan_in >= 3          # a, w and clk (plus the clock latch) come from outside
    assert fan_out == 0         # nothing outside the whole design reads its writes
    assert state >= 1 + 16 + 1 + 16


@pytest.mark.gpu
def test_hip_mlp_matches_fp64_reference(gpu):
    from rtlflow_amd.costmodel import MLP
    rng = np.random.default_rng(3)
    p = _params(3).astype(np.float32)
    x = rng.normal(0, 1, (777, F)).astype(np.float32)
    t = rng.normal(0, 1, 777).astype(np.float32)
    net = MLP(p)
    y = net.forward(x)
    want = mlp_ref.forward(p, x, F)
    assert np.allclose(y, want, rtol=1e-4, atol=1e-4)
    loss, g = net.gradients(x, t)
    gw = mlp_ref.gradients(p, x, t, F)
    assert abs(loss - mlp_ref.loss(p, x, t, F)) <= 1e-4 * max(1.0, abs(loss))
    # FP32 kernels vs FP64 restatement: relative 1e-3 of the gradient scale
    assert np.max(np.abs(g - gw)) <= 1e-3 * max(1.0, np.max(np.abs(gw)))
    # deterministic reduction: identical on a second call
    _, g2 = net.gradients(x, t)
    assert np.array_equal(g, g2)
    net.close()


@pytest.mark.gpu
def test_hip_mlp_training_fits_a_known_function(gpu):
    from rtlflow_amd.costmodel import train
    rng = np.random.default_rng(4)
    x = rng.uniform(0, 3, (2000, F)).astype(np.float32)
    y = (0.7 * x[:, 0] + np.log1p(x[:, 3]) - 0.2 * x[:, 4]).astype(np.float32)
    m = train(x, y, epochs=120, lr=3e-3, batch=128, seed=2)
    ep, tr, va = m["history"][-1]
    assert va < 0.02 * float(np.var(y))
