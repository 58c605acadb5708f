"""Cost-model MLP kernels on the GPU vs the FP64 oracle (gradients, inference,
Adam) and training behaviour."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(n_in, rows, seed):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-1, 1, (rows, n_in)).astype(np.float32)
    t = rng.normal(0, 1, rows).astype(np.float32)
    return X, t


@pytest.mark.parametrize("n_in,rows", [(1, 1), (5, 100), (17, 128), (17, 129), (32, 1000),
                                       (12, 10007), (34, 500), (48, 257)])
def test_gradient_matches_fp64_oracle(clops, oracle, ctx, queue, n_in, rows):
    m = clops.Mlp(ctx, n_in, seed=n_in * 31 + rows)
    X, t = _case(n_in, rows, rows)
    loss, g = m.gradient(queue, X, t)
    p = m.params().astype(np.float64)
    rl, rg = oracle.mlp_grad(p, X.astype(np.float64), t.astype(np.float64))
    # FP32 accumulation over up to 10^4 rows vs FP64: relative 1e-4 of the
    # gradient's scale.
    scale = np.max(np.abs(rg)) + 1e-12
    assert np.max(np.abs(g - rg)) <= 1e-4 * scale
    assert loss == pytest.approx(rl, rel=1e-4)


def test_gradient_is_reproducible(clops, ctx, queue):
    m = clops.Mlp(ctx, 9, seed=3)
    X, t = _case(9, 5000, 1)
    _, g1 = m.gradient(queue, X, t)
    _, g2 = m.gradient(queue, X, t)
    assert np.array_equal(g1, g2)


def test_predict_matches_oracle(clops, oracle, ctx, queue):
    m = clops.Mlp(ctx, 17, seed=11)
    X, _ = _case(17, 3001, 2)
    y = m.predict(queue, X)
    ry = oracle.mlp_forward(m.params().astype(np.float64), X.astype(np.float64))
    assert np.allclose(y, ry, rtol=1e-5, atol=1e-5)


def test_adam_steps_match_oracle(clops, oracle, ctx, queue):
    n_in, rows = 6, 300
    m = clops.Mlp(ctx, n_in, seed=7)
    X, t = _case(n_in, rows, 4)
    p = m.params().astype(np.float64)
    mom, vel = np.zeros_like(p), np.zeros_like(p)
    for step in range(1, 4):
        _, g = oracle.mlp_grad(p, X.astype(np.float64), t.astype(np.float64))
        p, mom, vel = oracle.adam_step(p, mom, vel, g, step, 1e-2)
    m.train(queue, X, t, 3, 1e-2)
    assert np.allclose(m.params(), p, rtol=1e-4, atol=1e-5)


def test_training_fits_a_known_function(clops, ctx, queue):
    rng = np.random.default_rng(9)
    X = rng.uniform(0, 1, (4096, 4)).astype(np.float32)
    t = (3 * X[:, 0] - 2 * X[:, 1] * X[:, 2] + np.abs(X[:, 3] - 0.5)).astype(np.float32)
    m = clops.Mlp(ctx, 4, seed=1)
    l0, _ = m.gradient(queue, X, t)
    l1 = m.train(queue, X, t, 1500, 3e-3)
    assert l1 < 0.01 * l0 and l1 < 2e-3


def test_save_load_roundtrip(clops, ctx, queue, tmp_path):
    m = clops.Mlp(ctx, 13, seed=5)
    path = str(tmp_path / "w.clomlp")
    m.save(path)
    m2 = clops.Mlp(ctx, path=path)
    assert np.array_equal(m.params(), m2.params())
    X, _ = _case(13, 50, 8)
    assert np.array_equal(m.predict(queue, X), m2.predict(queue, X))


def test_mlp_argument_errors(clops, ctx, queue):
    with pytest.raises(clops.ClopsError):
        clops.Mlp(ctx, 0)
    with pytest.raises(clops.ClopsError):
        clops.Mlp(ctx, 49)  # inputs 1..48
    with pytest.raises(clops.ClopsError):
        clops.Mlp(ctx, path="/nonexistent/w.clomlp")
