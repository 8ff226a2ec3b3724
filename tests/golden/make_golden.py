"""Writes the committed golden fixtures in this directory.

  philox_rocrand.json ... Philox4x32-10 blocks from rocRAND's header-only
                          engine (rocrand_philox4x32_10.h, a library routine),
                          compiled host-only with hipcc; pins the Philox draw
                          shared by the engine and the oracle.
  small_circuits.json ... small circuits (every standard gate kind on an
                          entangled input, random circuits, remote gates) with
                          statevectors from the independent numpy brute force
                          (tests/bruteforce.py).
  reference_kats.json ... known answers stated in the reference's own tests
                          (src/engine.rs:689-754, tests/test_statevector.py),
                          as data.

Run from the repository root: python tests/golden/make_golden.py
"""

import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import bruteforce  # noqa: E402

PHILOX_SRC = r"""
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
int main() {
  const unsigned long long seeds[] = {0ull, 1ull, 42ull, 1729ull, 0x123456789abcdefull, ~0ull};
  const unsigned long long streams[] = {0ull, 1ull, 7ull, 4095ull, 0xfffffffff0ull};
  const unsigned long long draws[] = {0ull, 1ull, 2ull, 123456789ull};
  for (auto s : seeds) for (auto st : streams) for (auto d : draws) {
    rocrand_device::philox4x32_10_engine e(s, st, 4ull * d);
    uint4 r = e.next4();
    printf("%llu %llu %llu %u %u %u %u\n", s, st, d, r.x, r.y, r.z, r.w);
  }
  return 0;
}
"""


def philox():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "p.cpp")
        exe = os.path.join(d, "p")
        with open(src, "w") as f:
            f.write(PHILOX_SRC)
        subprocess.run(["hipcc", "-O1", src, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    rows = []
    for line in out.strip().splitlines():
        s, st, d, *w = line.split()
        rows.append({"seed": int(s), "stream": int(st), "draw": int(d), "words": [int(x) for x in w]})
    return rows


def inst(kind, **f):
    return {"kind": kind, **f}


def gate_cases():
    theta, phi, lam, gamma = 0.731, -0.412, 1.137, 0.289
    single, ctl, pair = ("qubit",), ("control", "target"), ("a", "b")
    tri = ("control1", "control2", "target")
    quad = ("control1", "control2", "control3", "target")
    return [
        ("id", single, {}), ("u0", single, {}), ("x", single, {}), ("y", single, {}), ("z", single, {}),
        ("h", single, {}), ("s", single, {}), ("sdg", single, {}), ("t", single, {}), ("tdg", single, {}),
        ("sx", single, {}), ("sxdg", single, {}), ("u3", single, dict(theta=theta, phi=phi, lam=lam)),
        ("u2", single, dict(phi=phi, lam=lam)), ("u1", single, dict(lam=lam)),
        ("u", single, dict(theta=theta, phi=phi, lam=lam)), ("p", single, dict(lam=lam)),
        ("rx", single, dict(theta=theta)), ("ry", single, dict(theta=theta)), ("rz", single, dict(phi=phi)),
        ("cx", ctl, {}), ("cz", ctl, {}), ("cy", ctl, {}), ("ch", ctl, {}), ("swap", pair, {}), ("csx", ctl, {}),
        ("crx", ctl, dict(theta=theta)), ("cry", ctl, dict(theta=theta)), ("crz", ctl, dict(lam=lam)),
        ("cu1", ctl, dict(lam=lam)), ("cp", ctl, dict(lam=lam)), ("cu3", ctl, dict(theta=theta, phi=phi, lam=lam)),
        ("cu", ctl, dict(theta=theta, phi=phi, lam=lam, gamma=gamma)), ("rxx", pair, dict(theta=theta)),
        ("rzz", pair, dict(theta=theta)), ("ccx", tri, {}), ("cswap", ("control", "target1", "target2"), {}),
        ("rccx", tri, {}), ("rc3x", quad, {}), ("c3x", quad, {}), ("c3sqrtx", quad, {}),
        ("c4x", ("control1", "control2", "control3", "control4", "target"), {}),
    ]


def prefix(n):
    out = [inst("u", qubit=q, theta=0.21 + 0.37 * q, phi=-0.13 + 0.31 * q, lam=0.7 - 0.09 * q) for q in range(n)]
    out += [inst("cx", control=q, target=q + 1) for q in range(n - 1)]
    return out


def small_circuits():
    cases = []
    for kind, names, params in gate_cases():
        n = len(names) + 1
        for qubits in (list(range(len(names))), list(range(n - 1, 0, -1))):
            insts = prefix(n) + [inst(kind, **dict(zip(names, qubits)), **params)]
            cases.append({"name": f"{kind}@{qubits}", "n": n, "instructions": insts})
    for name, qs, params in [("remote_link_phi_plus", [0, 2], []), ("remote_epr", [2, 1], []),
                             ("remote_link_psi_plus", [1, 0], []), ("remote_link_psi_minus", [0, 1], []),
                             ("remote_cz", [2, 0], []), ("REMOTE_CX", [1, 2], []), ("remote_rzz", [0, 2], [0.4]),
                             ("remote_cu1", [1, 0], [0.82]), ("remote_barrier", [0, 1, 2], [])]:
        insts = prefix(3) + [inst("gate", name=name, qubits=qs, params=params)]
        cases.append({"name": f"gate:{name}", "n": 3, "instructions": insts})
    rng = np.random.default_rng(1907)
    gc = gate_cases()
    for run in range(4):
        n = 5
        insts = []
        for _ in range(40):
            kind, names, params = gc[int(rng.integers(len(gc)))]
            qubits = rng.choice(n, size=len(names), replace=False).tolist()
            insts.append(inst(kind, **dict(zip(names, qubits)), **params))
        cases.append({"name": f"random5q#{run}", "n": n, "instructions": insts})
    cases.append({"name": "tiny_rx", "n": 1, "instructions": [inst("rx", qubit=0, theta=1e-11)]})
    for c in cases:
        psi = bruteforce.statevector(c["n"], c["instructions"])
        c["expected"] = [[float(z.real), float(z.imag)] for z in psi]
    return cases


def reference_kats():
    """Known answers from the reference's tests, restated as data."""
    return {
        # src/engine.rs:689-705: amplitudes sqrt(i/36), i = 1..8
        "marginal": {"amplitudes": [[(i / 36) ** 0.5, 0.0] for i in range(1, 9)],
                     "cases": [{"qubits": [0, 2], "expected36": [4, 12, 6, 14]},
                               {"qubits": [2, 0], "expected36": [4, 6, 12, 14]},
                               {"qubits": [], "expected36": [36]},
                               {"qubits": [2, 1, 0], "expected36": list(range(1, 9))}],
                     "tolerance": 1e-14},
        # tests/test_statevector.py deterministic results
        "deterministic": [
            {"test": "little_endian x0,x2 (test_statevector.py:65-75)", "n": 4, "cbits": 0,
             "instructions": [inst("x", qubit=0), inst("x", qubit=2)],
             "probabilities": {"all": {"5": 1.0}, "[0, 1]": {"2": 1.0}, "[1, 0]": {"1": 1.0}, "[]": {"0": 1.0}},
             "counts": [{"shots": 17, "qubits": [0, 1], "expected": {"10": 17}},
                        {"shots": 17, "qubits": [], "expected": {"": 17}}]},
            {"test": "reset + nested conditionals (test_statevector.py:173-186)", "n": 2, "cbits": 3,
             "instructions": [inst("x", qubit=0), inst("measure", qubit=0, cbit=2), inst("reset", qubit=0),
                              inst("measure", qubit=0, cbit=0),
                              {"kind": "conditional", "condition": {"creg_base": 1, "creg_size": 2, "creg_value": 2},
                               "op": {"kind": "conditional",
                                      "condition": {"creg_base": 0, "creg_size": 1, "creg_value": 0},
                                      "op": inst("x", qubit=1)}},
                              inst("measure", qubit=1, cbit=1)],
             "classical_bits": {"0": 0, "1": 1, "2": 1}, "statevector": [[0, 0], [0, 0], [1, 0], [0, 0]],
             "shots": 64, "shot_counts": {"110": 64}},
            {"test": "wrapper x+measure (test_statevector.py:281-286)", "n": 1, "cbits": 1,
             "instructions": [inst("x", qubit=0), inst("measure", qubit=0, cbit=0)],
             "shots": 19, "shot_counts": {"1": 19}},
        ],
        # src/monolithic/statevector/shots.rs:289-301
        "budget": {"instructions": [inst("h", qubit=0), inst("measure", qubit=0, cbit=0)], "n": 2, "cbits": 1,
                   "shots": 10, "budget_bytes_ok": 65, "max_parallel_shots": 8, "strategy": "trajectories",
                   "parallel_shots": 1, "working_bytes": 65, "budget_bytes_fail": 64},
    }


def main():
    with open(os.path.join(HERE, "philox_rocrand.json"), "w") as f:
        json.dump(philox(), f)
    with open(os.path.join(HERE, "small_circuits.json"), "w") as f:
        json.dump(small_circuits(), f)
    with open(os.path.join(HERE, "reference_kats.json"), "w") as f:
        json.dump(reference_kats(), f, indent=1)


if __name__ == "__main__":
    main()
