"""Learned warm start: the HIP MLP kernels (bf16 MFMA GEMM with FP32 accumulation, GELU, MSE,
Adam) against a float64 numpy model that applies the same bf16 roundings (inputs, weights,
activations and the output gradients fed to the matrix cores), and the iterations the trained
network saves over cold starts on a family of RAOCPs.
"""
import numpy as np
import pytest
from scipy.special import erf

from oracle.oracle import Oracle
from spock_amd import runtime as rt
from spock_amd.api import configs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return rt.Context(0)


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a32 = np.ascontiguousarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def unpack(p, dims):
    Ws, bs, off = [], [], 0
    for a, b in zip(dims[:-1], dims[1:]):
        Ws.append(p[off:off + a * b].reshape(a, b).astype(np.float64))
        off += a * b
        bs.append(p[off:off + b].astype(np.float64))
        off += b
    return Ws, bs


def reference(p, dims, x, t):
    """Forward output, loss and parameter gradients with the kernels' bf16 roundings."""
    Ws, bs = unpack(p, dims)
    acts, zs = [bf16(x)], []
    h = acts[0]
    for l in range(len(Ws)):
        z = h @ bf16(Ws[l]) + bs[l]
        if l + 1 < len(Ws):
            zs.append(z.astype(np.float32).astype(np.float64))
            h = bf16(gelu(zs[-1]))
            acts.append(h)
        else:
            y = z
    n = y.size
    loss = float(np.mean((y - t) ** 2))
    d = (2.0 * (y.astype(np.float32) - t) / n).astype(np.float32).astype(np.float64)
    grads = [None] * (2 * len(Ws))
    for l in range(len(Ws) - 1, -1, -1):
        grads[2 * l] = acts[l].T @ bf16(d)
        grads[2 * l + 1] = d.sum(0)
        if l == 0:
            break
        da = bf16(d) @ bf16(Ws[l]).T
        d = (da * gelu_grad(zs[l - 1])).astype(np.float32).astype(np.float64)
    return y, loss, np.concatenate([g.ravel() for g in grads])


def relnorm(a, b):
    return float(np.linalg.norm(a - b) / max(1e-30, np.linalg.norm(b)))


# the last case is the warm start's own network (50 -> 3 x 512 -> 52,360, a minibatch of 256): the
# output layer's GEMMs run on the 128 x 128 CDNA4 kernel, its input gradient with split K
@pytest.mark.parametrize("dims,rows", [((5, 24, 24, 37), 19), ((50, 64, 64, 64, 300), 70),
                                       ((50, 200, 200, 200, 1000), 130), ((50, 512, 512, 512, 52360), 256)])
def test_mlp_forward_and_gradients_match_cpu(ctx, dims, rows):
    m = rt.WarmStartMLP(ctx, dims[0], dims[-1], hidden=dims[1], layers=len(dims) - 2, seed=5)
    rng = np.random.default_rng(1)
    x = rng.standard_normal((rows, dims[0])).astype(np.float32)
    t = rng.standard_normal((rows, dims[-1])).astype(np.float32)
    p = m.params()
    p[np.arange(p.size) % 7 == 3] *= -1.0  # break the symmetry of the default zero biases
    p += 0.01 * rng.standard_normal(p.size).astype(np.float32)
    m.set_params(p)
    y = m.forward(x)
    loss = m.loss_backward(t)
    g = m.grads()
    y_ref, loss_ref, g_ref = reference(p, list(dims), x, t)
    assert relnorm(y, y_ref) < 5e-3
    assert abs(loss - loss_ref) <= 1e-2 * loss_ref
    assert relnorm(g, g_ref) < 2e-2
    # every layer's block of the gradient, separately (a transposed operand would fail here)
    Ws, _ = unpack(p, list(dims))
    off = 0
    for W in Ws:
        n = W.size + W.shape[1]
        assert relnorm(g[off:off + n], g_ref[off:off + n]) < 3e-2
        off += n


def test_mlp_adam_fits_a_smooth_map(ctx):
    rng = np.random.default_rng(2)
    A = rng.standard_normal((8, 20)) / np.sqrt(8)
    x = rng.standard_normal((512, 8)).astype(np.float32)
    t = np.tanh(x @ A).astype(np.float32)
    m = rt.WarmStartMLP(ctx, 8, 20, hidden=64, layers=2, seed=3)
    first = None
    for it in range(300):
        m.forward(x)
        loss = m.loss_backward(t)
        first = loss if first is None else first
        m.adam(lr=3e-3)
    assert loss < 0.05 * first


def test_warm_start_saves_iterations(ctx, problem_dirs):
    """Train on solved instances of the 31-node problem, warm-start held-out ones: fewer
    iterations than cold starts, accepted against the cold solves' own thresholds, same cost."""
    d = problem_dirs["small"]
    o = Oracle(d)
    p = rt.Problem(ctx, d)
    rng = np.random.default_rng(40)
    n_train, n_test = 1024, 64
    x_train = rng.uniform(-1.0, 1.0, (n_train, o.nx))
    x_test = rng.uniform(-1.0, 1.0, (n_test, o.nx))
    opts = dict(tol_abs=1e-3, tol_rel=1e-3, max_iters=4000)
    train = rt.Batch(ctx, [p] * n_train, rt.F64, shared=True)
    train.set_x0(list(x_train))
    train.solve(rt.options(**opts))
    res_train, _ = train.results()
    assert sum(r["status"] == 0 for r in res_train) >= 0.9 * n_train
    m = rt.WarmStartMLP(ctx, o.nx, o.n_iter, hidden=256, layers=3, seed=7)
    m.dataset(train, x_train)
    losses = m.train(epochs=60, minibatch=128, lr=1e-3, seed=1)
    assert losses[-1] < 0.5 * losses[0]

    cold = rt.Batch(ctx, [p] * n_test, rt.F64, shared=True)
    cold.set_x0(list(x_test))
    cold.solve(rt.options(**opts))
    res_cold, _ = cold.results()
    warm = rt.Batch(ctx, [p] * n_test, rt.F64, shared=True)
    warm.set_x0(list(x_test))
    m.predict(x_test, warm)
    warm.set_tolerances([r["tol"] for r in res_cold])
    warm.solve(rt.options(**opts), warm=True)
    res_warm, _ = warm.results()
    ok = [k for k in range(n_test) if res_cold[k]["status"] == 0]
    assert len(ok) >= 0.9 * n_test
    assert all(res_warm[k]["status"] == 0 for k in ok)
    it_cold = np.mean([res_cold[k]["iters"] for k in ok])
    it_warm = np.mean([res_warm[k]["iters"] for k in ok])
    print(f"warm start: mean iterations {it_warm:.1f} vs cold {it_cold:.1f}")
    assert it_warm < 0.8 * it_cold
    # cost against a tightly converged FP64 solve: the warm start is no less accurate than cold starts
    tight = rt.Batch(ctx, [p] * n_test, rt.F64, shared=True)
    tight.set_x0(list(x_test))
    tight.solve(rt.options(tol_abs=1e-4, tol_rel=1e-4, max_iters=10000))
    res_tight, _ = tight.results()
    s = o.n_prim - o.sS
    zt = tight.iterates().reshape(n_test, -1)[:, s]
    zc = cold.iterates().reshape(n_test, -1)[:, s]
    zw = warm.iterates().reshape(n_test, -1)[:, s]
    good = [k for k in ok if res_tight[k]["status"] == 0]
    assert len(good) >= 0.8 * n_test
    err_c = np.array([abs(zc[k] - zt[k]) / max(1.0, abs(zt[k])) for k in good])
    err_w = np.array([abs(zw[k] - zt[k]) / max(1.0, abs(zt[k])) for k in good])
    print(f"cost error vs tight solve: warm mean {err_w.mean():.2e} max {err_w.max():.2e}, "
          f"cold mean {err_c.mean():.2e} max {err_c.max():.2e}")
    assert err_w.mean() <= 2.0 * err_c.mean() + 1e-3


def test_warm_start_example_runs():
    """spock.cu_amd/examples/warm_start_batch.py end to end on the 31-node problem (small sizes)."""
    import importlib.util
    import os
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "spock.cu_amd", "examples",
                        "warm_start_batch.py")
    spec = importlib.util.spec_from_file_location("warm_start_batch", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    ok, states = mod.main(states=256, train=256)
    assert ok >= 0.8 * states
