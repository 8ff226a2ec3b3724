"""Pins the CPU oracle before it is trusted as the checker.

Each check is chosen so that a plausible mistake in the oracle (a dropped
term, a wrong sign, a transposed operand, a wrong bit order, a sequential
instead of pairwise sum) fails it:
  - committed golden vectors: rocRAND's Philox engine and the independent
    numpy brute force (tests/bruteforce.py) on every gate kind;
  - the reference's own known-answer procedures (src/engine.rs:568-754,
    src/gates.rs:302-377, shots.rs:268-301, tests/test_statevector.py);
  - closed forms (QFT of a product state), invariants (mirror circuits,
    norm) and the Random123 Philox known-answer vectors.
"""

import json
import math
import os

import numpy as np
import pytest

import bruteforce as bf
from dqsim import workloads
from oracle import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


# ---------------------------------------------------------------------------
# Philox
# ---------------------------------------------------------------------------
def test_philox_random123_known_answers():
    # Known-answer vectors published with Random123 (philox4x32, 10 rounds).
    assert oracle.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert oracle.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert oracle.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_philox_matches_rocrand_engine():
    for row in load("philox_rocrand.json"):
        s, st, d = row["seed"], row["stream"], row["draw"]
        ctr = [d & 0xFFFFFFFF, d >> 32, st & 0xFFFFFFFF, st >> 32]
        assert oracle.philox(ctr, [s & 0xFFFFFFFF, s >> 32]) == row["words"], row
        w = row["words"]
        u = ((w[1] << 32 | w[0]) >> 11) * 2.0 ** -53
        assert oracle.uniform53(s, st, d) == u


# ---------------------------------------------------------------------------
# Gates and whole circuits against the independent brute force
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("case", load("small_circuits.json"), ids=lambda c: c["name"])
def test_oracle_matches_bruteforce_golden(case):
    want = np.array([complex(a, b) for a, b in case["expected"]])
    got, _ = oracle.simulate({"qregs": {"q": {"name": "q", "base": 0, "size": case["n"]}},
                              "cregs": {}, "instructions": case["instructions"]}, 1)
    tol = 1e-13 if case["name"] == "tiny_rx" else 4e-12
    np.testing.assert_allclose(got, want, atol=tol, rtol=tol)


def test_tiny_rotation_survives_fusion():
    theta = 1e-11
    got, _ = oracle.simulate(workloads.Circuit(1, [{"kind": "rx", "qubit": 0, "theta": theta}]).data, 1)
    np.testing.assert_allclose(got, [math.cos(theta / 2), -1j * math.sin(theta / 2)], rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("kind,expected", [
    ("x", np.array([[0, 1], [1, 0]])), ("h", bf.H), ("s", bf.S), ("t", bf.T), ("sx", bf.SX), ("sxdg", bf.SXDG),
    ("y", bf.Y), ("z", bf.Z), ("sdg", bf.SDG), ("tdg", bf.TDG)])
def test_fixed_gates_are_textbook(kind, expected):
    np.testing.assert_allclose(oracle.gate_1q(kind), expected, atol=1e-15)


def test_parametric_gates_are_textbook_and_unitary():
    th, ph, la = 0.731, -0.412, 1.137
    np.testing.assert_allclose(oracle.gate_1q("u3", th, ph, la), bf.u3(th, ph, la), atol=1e-15)
    np.testing.assert_allclose(oracle.gate_1q("u2", 0, ph, la), bf.u3(math.pi / 2, ph, la), atol=1e-15)
    np.testing.assert_allclose(oracle.gate_1q("rx", th), bf.rx(th), atol=1e-15)
    np.testing.assert_allclose(oracle.gate_1q("ry", th), bf.ry(th), atol=1e-15)
    np.testing.assert_allclose(oracle.gate_1q("rz", 0, ph), bf.rz(ph), atol=1e-15)
    np.testing.assert_allclose(oracle.gate_1q("p", 0, 0, la), bf.p(la), atol=1e-15)
    for k in ("u3", "u2", "rx", "ry", "rz", "p"):
        m = oracle.gate_1q(k, th, ph, la)
        np.testing.assert_allclose(m @ m.conj().T, np.eye(2), atol=1e-15)


def test_rccx_and_rc3x_match_elementary_decompositions():
    # gates.rs:302-377: RCCX = H T CX Tdg CX T CX Tdg H; RC3X = CS . C(RCCX).
    for kind, k, mat in (("rccx", 3, bf.rccx_matrix()), ("rc3x", 4, bf.rc3x_matrix())):
        names = ["control1", "control2", "control3"][: k - 1] + ["target"]
        for col in range(1 << k):
            insts = [{"kind": "x", "qubit": k - 1 - i} for i in range(k) if (col >> (k - 1 - i)) & 1]
            insts.append({"kind": kind, **{nm: k - 1 - i for i, nm in enumerate(names)}})
            got, _ = oracle.simulate(workloads.Circuit(k, insts).data, 1)
            got_op = np.empty(1 << k, dtype=complex)
            for b in range(1 << k):
                got_op[b] = got[sum(((b >> (k - 1 - i)) & 1) << (k - 1 - i) for i in range(k))]
            np.testing.assert_allclose(got_op, mat[:, col], atol=1e-13)


# ---------------------------------------------------------------------------
# Kernel procedures of the reference's engine tests (src/engine.rs:568-687)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n", [5, 12])
def test_dense_kernel_matches_dense_reference(n):
    for qs in ([0], [n - 1], [n - 1, 0], [0, 3, 1], [4, 2, 0, 3, 1]):
        mat = bf.fourier(1 << len(qs))
        st = bf.test_input(n)
        want = bf.dense_reference(st, mat, qs)
        oracle.apply_dense(st, qs, mat)
        assert np.abs(st - want).max() < 1e-12


@pytest.mark.parametrize("n", [4, 12])
def test_controlled_one_and_x_match_dense_reference(n):
    u = bf.fourier(2)
    for target in (0, n - 1):
        control = n - 1 if target == 0 else 0
        for value in (False, True):
            for is_x in (False, True):
                blk = bf.X if is_x else u
                full = np.eye(4, dtype=complex)
                r = 2 if value else 0
                full[r:r + 2, r:r + 2] = blk
                st = bf.test_input(n)
                want = bf.dense_reference(st, full, [control, target])
                if is_x:
                    oracle.apply_x(st, target, [(control, value)])
                else:
                    oracle.apply_one(st, target, u, [(control, value)])
                assert np.abs(st - want).max() < 1e-12


@pytest.mark.parametrize("n", [4, 12, 13])
def test_diagonal_and_controlled_swap_match_dense_reference(n):
    for qs in ([0, n - 1], [n - 1, 0]):
        d = np.exp(1j * 0.3 * np.arange(4))
        st = bf.test_input(n)
        want = bf.dense_reference(st, np.diag(d), qs)
        oracle.apply_diagonal(st, qs, d)
        assert np.abs(st - want).max() < 1e-12
    for a, b in ((0, n - 1), (n - 1, 0), (n - 1, n - 2)):
        c = next(q for q in range(n) if q not in (a, b))
        for ctl in ([], [(c, True)], [(c, False)]):
            qubits = [a, b] if not ctl else [c, a, b]
            m = bf.SWAP if not ctl else (np.kron(np.diag([1, 0]), np.eye(4)) + np.kron(np.diag([0, 1]), bf.SWAP)
                                         if ctl[0][1] else
                                         np.kron(np.diag([1, 0]), bf.SWAP) + np.kron(np.diag([0, 1]), np.eye(4)))
            st = bf.test_input(n)
            want = bf.dense_reference(st, m, qubits)
            oracle.apply_swap(st, a, b, ctl)
            assert np.abs(st - want).max() < 1e-12


def test_marginal_known_answers():
    kat = load("reference_kats.json")["marginal"]
    st = np.array([complex(a, b) for a, b in kat["amplitudes"]])
    for case in kat["cases"]:
        got = oracle.marginal(st, case["qubits"])
        np.testing.assert_allclose(got, np.array(case["expected36"]) / 36.0, atol=kat["tolerance"], rtol=0)


def test_sampling_edge_cases():
    # engine.rs:707-731: zero-probability outcomes are never drawn, even for
    # the smallest draws; empty selections give the empty key.
    st = np.array([0, 2], dtype=complex)
    for seed in range(20):
        assert set(oracle.sample_marginal(st, [0], 10, seed).tolist()) == {1}
    assert oracle.counts_from_indices(oracle.sample_marginal(st, [], 7, 3), 0) == {"": 7}
    assert oracle.sample_marginal(st, [0], 0, 3).size == 0
    counts = oracle.counts_from_indices(oracle.sample_marginal(np.array([2, 2], dtype=complex), [0], 10000, 123), 1)
    assert sum(counts.values()) == 10000 and abs(counts["0"] - 5000) < 250


@pytest.mark.parametrize("n", [1, 12])
def test_measurement_normalizes_both_branches(n):
    # engine.rs:733-754: r = 0 selects outcome 1, r = 1 - 2^-53 outcome 0.
    for target in (0, n - 1):
        for outcome, r in ((1, 0.0), (0, 1.0 - 2.0 ** -53)):
            st = np.zeros(1 << n, dtype=complex)
            st[0] = 2
            st[1 << target] = 2
            got, p0, p1 = oracle.measure(st, target, r)
            assert got == outcome
            assert abs(np.vdot(st, st).real - 1.0) < 1e-14
            assert st[outcome << target] == 1.0


def test_pairwise_tree_order():
    # A sequential sum gives 1.0 here; the pairwise tree (1e16+1) + (1-1e16)
    # loses both ones.
    assert oracle.ptree(np.array([1e16, 1.0, -1e16, 1.0])) == 0.0
    x = np.arange(1, 1025, dtype=float)
    assert oracle.ptree(x) == 1024 * 1025 / 2
    assert oracle.ptree(np.array([0.1] * 8)) == ((0.1 + 0.1) + (0.1 + 0.1)) + ((0.1 + 0.1) + (0.1 + 0.1))


# ---------------------------------------------------------------------------
# Closed forms and invariants
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n", [3, 5])
def test_qft_closed_form_matches_bruteforce(n):
    circ, angles = workloads.qft(n, seed=11, measure=False)
    want = bf.statevector(n, circ.data["instructions"])
    np.testing.assert_allclose(workloads.qft_closed_form(angles, np.arange(1 << n)), want, atol=1e-13)


@pytest.mark.parametrize("n", [6, 10])
def test_oracle_qft_matches_closed_form(n):
    circ, angles = workloads.qft(n, seed=n, measure=False)
    got, _ = oracle.simulate(circ.data, 1)
    np.testing.assert_allclose(got, workloads.qft_closed_form(angles, np.arange(1 << n)), atol=1e-12)


def test_mirror_circuit_returns_to_zero():
    circ = workloads.random_su2(n=12, depth=6, seed=3)
    got, _ = oracle.simulate(workloads.mirror(circ).data, 1)
    want = np.zeros(1 << 12, dtype=complex)
    want[0] = 1
    assert np.abs(got - want).max() < 1e-12


def test_brickwork_norm_and_bruteforce_small():
    circ = workloads.brickwork(n=6, layers=4, seed=42, measure=False)
    got, _ = oracle.simulate(circ.data, 1)
    np.testing.assert_allclose(got, bf.statevector(6, circ.data["instructions"]), atol=1e-12)


# ---------------------------------------------------------------------------
# Compile and shot driver (execution.rs:368-416, shots.rs:268-301)
# ---------------------------------------------------------------------------
def test_compile_lowering_kinds():
    ins = [{"kind": "id", "qubit": 0}, {"kind": "h", "qubit": 0}, {"kind": "h", "qubit": 0},
           {"kind": "cx", "control": 0, "target": 1}, {"kind": "cz", "control": 0, "target": 1},
           {"kind": "rzz", "a": 0, "b": 1, "theta": 0.3}, {"kind": "rxx", "a": 0, "b": 1, "theta": 0.3},
           {"kind": "x", "qubit": 1}, {"kind": "measure", "qubit": 1, "cbit": 0}, {"kind": "rz", "qubit": 0,
                                                                                   "phi": 0.2}]
    # H.H is diag(2s^2, 2s^2) in FP64 (off-diagonals cancel exactly, the
    # diagonal is 1 + 2^-52), so it is kept and lowered to a diagonal op.
    assert oracle.compile_kinds(ins) == ["diagonal", "x", "diagonal", "diagonal", "dense", "x", "measure",
                                         "diagonal"]


def test_terminal_sampling_keeps_correlations_and_last_write():
    c = workloads.Circuit(2, [{"kind": "h", "qubit": 0}, {"kind": "cx", "control": 0, "target": 1},
                              {"kind": "measure", "qubit": 0, "cbit": 0}, {"kind": "measure", "qubit": 1, "cbit": 1},
                              {"kind": "measure", "qubit": 1, "cbit": 2}], 3)
    counts, info = oracle.run_shots(c.data, 1024, 5, budget_mb=1)
    assert info["execution_strategy"] == "terminal_sampling"
    assert sum(counts.values()) == 1024 and set(counts) == {"000", "111"}


def test_budget_selects_one_worker_and_sampling_fallback():
    kat = load("reference_kats.json")["budget"]
    data = workloads.Circuit(kat["n"], kat["instructions"], kat["cbits"]).data
    import ctypes

    arr, cnt = oracle.to_insts(data["instructions"])
    bits = np.zeros((kat["shots"], 1), dtype=np.uint8)
    info = oracle.ShotInfo()
    rc = oracle.lib().or_run_shots(arr, cnt, kat["n"], kat["cbits"], kat["shots"], 5, kat["budget_bytes_ok"], 1,
                                   kat["max_parallel_shots"], 1, 1024, 16, bits.ctypes.data, ctypes.byref(info))
    assert rc == 0
    assert oracle.STRATEGIES[info.strategy] == kat["strategy"]
    assert info.parallel_shots == kat["parallel_shots"] and info.working_bytes == kat["working_bytes"]
    rc = oracle.lib().or_run_shots(arr, cnt, kat["n"], kat["cbits"], kat["shots"], 5, kat["budget_bytes_fail"], 0, 0,
                                   1, 1024, 16, bits.ctypes.data, ctypes.byref(info))
    assert rc == 2


def test_reference_deterministic_results():
    for case in load("reference_kats.json")["deterministic"]:
        data = workloads.Circuit(case["n"], case["instructions"], case["cbits"]).data
        if "shot_counts" in case:
            for terminal in (True, False):
                counts, _ = oracle.run_shots(data, case["shots"], 22, sample_terminal=terminal)
                assert counts == case["shot_counts"]
        st, bits = oracle.simulate(data, 1729)
        if "classical_bits" in case:
            assert bits == {int(k): v for k, v in case["classical_bits"].items()}
            np.testing.assert_allclose(st, [complex(a, b) for a, b in case["statevector"]], atol=1e-14)
        if "counts" in case:
            for c in case["counts"]:
                idx = oracle.sample_marginal(st, c["qubits"], c["shots"], 2)
                assert oracle.counts_from_indices(idx, len(c["qubits"])) == c["expected"]


# ---------------------------------------------------------------------------
# Planner MLP oracle (oracle/mlp.py): central finite differences
# ---------------------------------------------------------------------------
def test_mlp_oracle_gradient_matches_finite_differences():
    from oracle import mlp

    rng = np.random.default_rng(4)
    w = rng.normal(0, 0.3, size=mlp.F * 64 + 64 + 64 * 64 + 64 + 64 + 1)
    X = rng.normal(0, 1, size=(9, mlp.F))
    y = rng.normal(0, 1, size=9)
    g, _ = mlp.grad(w, X, y)
    idx = rng.choice(w.size, size=60, replace=False).tolist() + [w.size - 1, 0, mlp.F * 64]
    for i in idx:
        e = np.zeros_like(w)
        e[i] = 1e-6
        fd = (mlp.loss(w + e, X, y) - mlp.loss(w - e, X, y)) / 2e-6
        assert abs(fd - g[i]) < 1e-6 + 1e-5 * abs(g[i]), i


# ---------------------------------------------------------------------------
# Result queries and samplers that the reference's tests do not cover
# ---------------------------------------------------------------------------
def _random_state(n, rng):
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return v / np.linalg.norm(v)


def test_fidelity_matches_vdot_and_closed_forms():
    # |<a|b>|^2 (simulator.rs:178-190): numpy's vdot conjugates its first
    # argument; a dropped conjugate or a swapped real/imaginary part fails.
    rng = np.random.default_rng(2024)
    for n in (1, 3, 8, 11):
        a, b = _random_state(n, rng), _random_state(n, rng)
        want = abs(np.vdot(a, b)) ** 2
        assert abs(oracle.fidelity(a, b) - want) <= 1e-14 + 1e-12 * want
        assert abs(oracle.fidelity(a, a) - 1.0) < 1e-13
        assert abs(oracle.fidelity(a, np.exp(0.731j) * a) - 1.0) < 1e-13  # global phase
        assert abs(oracle.fidelity(a, b) - oracle.fidelity(b, a)) < 1e-15
    zero, plus = np.array([1, 0], dtype=complex), np.array([1, 1], dtype=complex) / math.sqrt(2)
    assert oracle.fidelity(zero, np.array([0, 1], dtype=complex)) == 0.0  # orthogonal
    assert abs(oracle.fidelity(zero, plus) - 0.5) < 1e-15
    # (1, i)/sqrt2 with itself: 1 with the conjugate, 0 without it.
    y = np.array([1, 1j]) / math.sqrt(2)
    assert abs(oracle.fidelity(y, y) - 1.0) < 1e-15
    assert abs(oracle.fidelity(y, y.conj())) < 1e-15


def _inverse_cdf(p, shots, seed):
    # Textbook inverse-CDF sampling over the full sequential CDF.
    cdf = np.cumsum(p) / p.sum()
    cdf[-1] = 1.0
    r = np.array([oracle.uniform53(seed, s, 0) for s in range(shots)])
    return np.searchsorted(cdf, r, side="right")


def test_batch_sample_matches_inverse_cdf_on_dyadic_states():
    # Probabilities that are multiples of 2^-16 make every partial sum exact,
    # so the chunked search (chunk totals, binary search, in-chunk scan) must
    # pick exactly the outcome of plain inverse-CDF sampling, for any chunking.
    rng = np.random.default_rng(7)
    for n in (1, 4, 7, 10):
        amp = rng.integers(0, 3, size=1 << n).astype(float)
        amp[rng.integers(1 << n)] = 2.0  # at least one non-zero outcome
        w = amp * amp  # 0, 1 or 4: |a|^2 of amp * 2^-8 is w * 2^-16 exactly
        amp = amp * 2.0 ** -8
        want = _inverse_cdf(w, 600, 31 + n)
        for block in (1, 2, 16, 256, 1024):
            got = oracle.batch_sample(amp.astype(complex), 600, 31 + n, block)
            assert np.array_equal(got, want), (n, block)
    # Zero-probability outcomes are never drawn, and a basis state always
    # gives its own index.
    st = np.zeros(64, dtype=complex)
    st[37] = 1.0
    assert set(oracle.batch_sample(st, 100, 3, 16).tolist()) == {37}


def test_measure_f32_rounds_the_collapse_like_complex64():
    # Kept amplitudes are scaled in FP64 by 1/sqrt(p) and rounded to FP32;
    # restated here with numpy's float32 conversion.
    rng = np.random.default_rng(12)
    for n in (2, 6, 12):
        st = _random_state(n, rng).astype(np.complex64).astype(np.complex128)
        for q, r in ((0, 0.3), (n - 1, 0.8), (n // 2, 0.5)):
            ref = st.copy()
            got = st.copy()
            out, p0, p1 = oracle.measure(got, q, r, f32=True)
            out64, q0, q1 = oracle.measure(ref, q, r)
            assert (out, p0, p1) == (out64, q0, q1)
            keep = ((np.arange(1 << n) >> q) & 1) == out
            scale = 1.0 / math.sqrt(p1 if out else p0)
            want = np.where(keep, st, 0)
            want = (want.real * scale).astype(np.float32) + 1j * (want.imag * scale).astype(np.float32)
            assert np.array_equal(got, want.astype(np.complex128))
            assert np.array_equal(ref[keep] * 0 + (st[keep].real * scale + 1j * st[keep].imag * scale), ref[keep])
            st = got


def test_terminal_trajectories_match_oracle_run_shots():
    # The trajectory helper (prefix + ordered measurements, stream = shot,
    # draw = measurement index) against the oracle's own trajectory driver.
    n = 7
    unitary = workloads.random_su2(n=n, depth=4, seed=3).data["instructions"]
    measures = [(3, 0), (0, 1), (6, 2), (3, 3), (1, 1)]
    ins = unitary + [{"kind": "measure", "qubit": q, "cbit": c} for q, c in measures]
    circ = {"qregs": {"q": {"name": "q", "base": 0, "size": n}}, "cregs": {"c": {"name": "c", "base": 0, "size": 4}},
            "instructions": ins}
    prefix, _ = oracle.simulate({**circ, "instructions": unitary, "cregs": {}}, 1)
    want, _ = oracle.run_shots(circ, 300, 19, sample_terminal=False)
    assert oracle.terminal_trajectories(prefix, measures, 4, 300, 19) == want


def test_canonical_cdf_exact_on_dyadic_bins_at_every_level():
    # Sums of multiples of 2^-20 below 2^33 are exact in FP64, so the block
    # scan (any addition order) must equal the exact integer prefix, across
    # one block, two levels (2048 < N <= 2048^2) and a partial last block.
    rng = np.random.default_rng(5)
    for count in (1, 7, 2048, 2049, 6000, (1 << 13) + 5, 1 << 16):
        k = rng.integers(0, 1 << 10, size=count)
        x = k.astype(np.float64) * 2.0 ** -20
        want = np.cumsum(k).astype(np.float64) * 2.0 ** -20
        got = oracle.cdf(x)
        assert np.array_equal(got, want), count


def test_canonical_cdf_is_the_block_scan_not_a_sequential_sum():
    # Bin 0 = 1 and 2047 bins of 2^-53: a sequential sum never moves off 1.0
    # (each 1 + 2^-53 rounds back to 1); in the block scan thread 0 loses its
    # seven small bins the same way, but the other 255 threads sum theirs to
    # 2^-50 each before they reach 1.0, so the total is exactly 1 + 255 * 2^-50.
    x = np.full(2048, 2.0 ** -53)
    x[0] = 1.0
    c = oracle.cdf(x)
    assert math.fsum(x[:1]) == 1.0 and sum(x.tolist()) == 1.0
    assert c[-1] == 1.0 + 255 * 2.0 ** -50
    assert c[7] == 1.0 and c[15] == 1.0 + 2.0 ** -50
    # Non-decreasing, last entry = total, within the FP64 error bound of the
    # exact sum on random bins at three levels.
    rng = np.random.default_rng(8)
    for count in (3000, 5 << 20):
        x = rng.random(count) ** 3
        c = oracle.cdf(x)
        assert np.all(np.diff(c) >= 0)
        assert abs(c[-1] - math.fsum(x)) <= 64 * 2.0 ** -53 * math.fsum(x)
