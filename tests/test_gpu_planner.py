"""Learned fusion planner on the GPU: the MLP kernels' gradients against the
FP64 restatement (oracle/mlp.py), training convergence, engine-recorded
training data, and learned plans keeping the <=1e-5 amplitude bound."""

import numpy as np
import pytest

import dqsim
from dqsim import planner, workloads
from oracle import mlp, oracle

pytestmark = pytest.mark.gpu


def test_mlp_gradient_matches_fp64_oracle():
    rng = np.random.default_rng(3)
    w = planner.init_weights(5)
    F = planner.N_FEATURES
    w[F * 64:F * 64 + 64] = rng.normal(0, 0.1, 64)  # non-zero biases
    X = rng.normal(0, 1, size=(37, planner.N_FEATURES)).astype(np.float32)
    y = rng.normal(0, 1, size=37).astype(np.float32)
    g, loss = planner.grad(w, X, y)
    gref, lref = mlp.grad(w.astype(np.float64), X.astype(np.float64), y.astype(np.float64))
    assert abs(loss - lref) <= 1e-5 * max(1.0, lref)
    scale = np.abs(gref).max()
    assert np.abs(g - gref).max() <= 1e-4 * scale
    out = planner.predict(w, X)
    ref, _ = mlp.forward(w, X)
    assert np.abs(out - ref).max() <= 1e-4 * max(1.0, np.abs(ref).max())


def test_training_reduces_loss_and_fits_a_known_function():
    rng = np.random.default_rng(8)
    X = rng.uniform(0, 1, size=(2048, planner.N_FEATURES)).astype(np.float32)
    y = (2.0 * X[:, 0] - X[:, 3] + 0.5 * np.sin(3 * X[:, 8])).astype(np.float32)
    w0 = planner.init_weights(1)
    _, l0 = planner.grad(w0, X, y)
    w, curve = planner.train(X, y, epochs=120, batch=256, lr=3e-3, seed=2, w0=w0)
    _, l1 = planner.grad(w, X, y)
    assert l1 < 0.05 * l0 and l1 < 0.01
    assert curve[-1] < curve[0]
    w2, _ = planner.train(X, y, epochs=120, batch=256, lr=3e-3, seed=2, w0=w0)
    assert np.array_equal(w, w2)  # deterministic training


def test_engine_records_training_data_and_learned_plan_is_accurate():
    circs = list(planner.training_circuits(4, nmin=14, nmax=16, dmin=4, dmax=10, seed=1))
    X, y = planner.dataset(circs, seed=1)
    assert X.shape[1] == planner.N_FEATURES and len(X) == len(y) > 20 and np.isfinite(y).all()
    w, _ = planner.train(X, y, epochs=60, batch=64, lr=3e-3, seed=0)
    planner.install(w)
    try:
        c = workloads.random_su2(n=12, depth=10, seed=9)
        want, _ = oracle.simulate(c.data, 1)
        sim = dqsim.StatevectorSimulator(seed=1, precision="complex64", fusion="learned", max_fused_qubits=5)
        assert np.abs(sim.simulate(c).statevector - want).max() < 1e-5
    finally:
        planner.install(None)
