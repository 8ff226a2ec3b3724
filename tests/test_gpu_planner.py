"""Learned fusion planner on the GPU: the MLP kernels' gradients against the
FP64 restatement (oracle/mlp.py), training convergence, engine-recorded
training data, learned plans keeping the <=1e-5 amplitude bound, and the
shipped weights against the greedy planner on held-out circuits."""

import time

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


# ---- learned vs greedy on the shipping path (out-of-place tile sweeps) ----
# Held-out circuits: seeds 20_000 + n never occur in training (training
# circuits use seed * 1_000_003 + i with seed 11). The shipped weights were
# trained on engine timings of other circuits of the same families. Both
# planners get blocks of up to 5 qubits (tile sweeps take up to 4): greedy
# fuses as far as that allows, the learned planner picks the block width and
# the tile rule by predicted time.
HELD_OUT = [("su2", 28, 40), ("brickwork", 29, 40), ("qft", 30, 0), ("mixed", 31, 40), ("su2", 32, 40),
            ("brickwork", 32, 60), ("qft", 33, 0)]
_AB = {}


def _wall_ms(sim, c):
    t = time.perf_counter()
    sim.evolve_timed(c, record=False)
    return (time.perf_counter() - t) * 1e3


@pytest.mark.parametrize("family,n,depth", HELD_OUT)
def test_learned_plan_not_slower_than_greedy_on_held_out_circuits(family, n, depth):
    import dqsim._core as core
    assert planner.install_default(), "shipped planner weights missing"
    try:
        c = planner.family_circuit(family, n, depth, 20_000 + n)
        mem = (1 << n) * 8 * 3 // (1 << 20) + 4096  # two complex64 states: out-of-place sweeps
        sims = {f: dqsim.StatevectorSimulator(seed=1, precision="complex64", fusion=f, max_fused_qubits=5,
                                              max_memory_mb=mem) for f in ("greedy", "learned")}
        choice = core.plan_json(c, precision="complex64", fusion="learned", max_fused_qubits=5,
                                out_of_place=True)["learned_choices"]
        print(f"[planner-ab] {family} n={n} depth={depth} learned choice {choice}", flush=True)
        for f in sims:
            _wall_ms(sims[f], c)  # warm-up
        ms = {f: [] for f in sims}
        for _ in range(3):
            for f in sims:
                ms[f].append(_wall_ms(sims[f], c))
        med = {f: float(np.median(v)) for f, v in ms.items()}
        # Accuracy: the same bound for both plans, against the complex128
        # path (pinned to the FP64 oracle by the parity tests) up to 32
        # qubits; at 33 qubits (no room for a complex128 state beside two
        # complex64 ones) the two plans against each other.
        res = {f: sims[f].evolve_timed(c, record=False, keep=True)[4] for f in ("greedy",)}
        if n <= 32:
            exact = dqsim.StatevectorSimulator(seed=1, precision="complex128", max_memory_mb=mem * 2)
            ref = exact.evolve_timed(c, record=False, keep=True)[4]
            err_g, _ = res["greedy"].compare(ref)
            del res
            lres = sims["learned"].evolve_timed(c, record=False, keep=True)[4]
            err_l, _ = lres.compare(ref)
            del lres, ref
        else:
            lres = sims["learned"].evolve_timed(c, record=False, keep=True)[4]
            err_l, _ = lres.compare(res["greedy"])
            err_g = 0.0
            del lres, res
        gain = med["greedy"] / med["learned"]
        _AB[(family, n)] = (med["greedy"], med["learned"])
        print(f"[planner-ab] {family} n={n}: greedy {med['greedy']:.1f} ms, learned {med['learned']:.1f} ms, "
              f"gain {gain:.3f}, max err greedy {err_g:.2e} learned {err_l:.2e}", flush=True)
        assert err_g <= 1e-5 and err_l <= 1e-5
        # The learned plan is never slower beyond timing noise (3 %) and the
        # host time of scoring its candidates (dry emissions of six plans and
        # the model on their launches: 1-3 ms; where the model keeps greedy's
        # plan, as for su2-28, that time is the whole difference).
        assert med["learned"] <= 1.03 * med["greedy"] + 3.0
    finally:
        planner.install(None)


def test_learned_total_not_slower_than_greedy():
    if len(_AB) < len(HELD_OUT):
        pytest.skip("needs every held-out circuit timed in this session")
    g = sum(v[0] for v in _AB.values())
    l = sum(v[1] for v in _AB.values())
    print(f"[planner-ab] total greedy {g:.1f} ms, learned {l:.1f} ms, gain {g / l:.3f}", flush=True)
    assert l <= g
