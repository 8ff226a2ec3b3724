"""Independent brute-force statevector simulator for small circuits (numpy).

Used only to pin the oracle and the engine on small inputs. Every gate is
written from its textbook definition (Qiskit conventions, matrix index MSB =
first operand, as the reference's gates.rs:101 states), embedded into the full
2^n x 2^n operator and applied by a dense matrix-vector product. It shares no
code with the oracle (oracle/sv_oracle.c) or the engine.
"""

from __future__ import annotations

import math
from typing import List, Sequence

import numpy as np

S2 = 1.0 / math.sqrt(2.0)
I2 = np.eye(2, dtype=complex)
X = np.array([[0, 1], [1, 0]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
Z = np.array([[1, 0], [0, -1]], dtype=complex)
H = np.array([[1, 1], [1, -1]], dtype=complex) * S2
S = np.diag([1, 1j])
SDG = np.diag([1, -1j])
T = np.diag([1, np.exp(1j * math.pi / 4)])
TDG = np.diag([1, np.exp(-1j * math.pi / 4)])
SX = 0.5 * np.array([[1 + 1j, 1 - 1j], [1 - 1j, 1 + 1j]])
SXDG = SX.conj().T


def u3(theta, phi, lam):
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    return np.array([[c, -np.exp(1j * lam) * s], [np.exp(1j * phi) * s, np.exp(1j * (phi + lam)) * c]])


def rx(t):
    return math.cos(t / 2) * I2 - 1j * math.sin(t / 2) * X


def ry(t):
    return math.cos(t / 2) * I2 - 1j * math.sin(t / 2) * Y


def rz(t):
    return np.diag([np.exp(-0.5j * t), np.exp(0.5j * t)])


def p(lam):
    return np.diag([1, np.exp(1j * lam)])


def controlled(u, nctl=1):
    d = u.shape[0]
    m = np.eye(d << nctl, dtype=complex)
    m[-d:, -d:] = u
    return m


SWAP = np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=complex)


def embed(m: np.ndarray, qubits: Sequence[int], n: int) -> np.ndarray:
    """Full operator of `m` acting on `qubits` (qubits[0] = MSB of m's index);
    basis index bit q is qubit q (little-endian)."""
    k = len(qubits)
    N = 1 << n
    full = np.zeros((N, N), dtype=complex)
    mask = sum(1 << q for q in qubits)
    for b in range(N):
        col = 0
        for q in qubits:
            col = (col << 1) | ((b >> q) & 1)
        rest = b & ~mask
        for row in range(1 << k):
            out = rest
            for i, q in enumerate(qubits):
                if (row >> (k - 1 - i)) & 1:
                    out |= 1 << q
            full[out, b] += m[row, col]
    return full


def rccx_matrix() -> np.ndarray:
    """RCCX from its elementary definition H, T, CX(c2), Tdg, CX(c1), T,
    CX(c2), Tdg, H on the target (operands [c1, c2, t])."""
    n = 3
    seq = [(H, [2]), (T, [2]), (controlled(X), [1, 2]), (TDG, [2]), (controlled(X), [0, 2]), (T, [2]),
           (controlled(X), [1, 2]), (TDG, [2]), (H, [2])]
    full = np.eye(8, dtype=complex)
    for m, qs in seq:
        full = embed(m, [n - 1 - q for q in qs], n) @ full
    # operands [c1, c2, t] map to matrix bits MSB..LSB: qubit index n-1-i
    return full


def rc3x_matrix() -> np.ndarray:
    """RC3X(c, a, b; t) = CS(c, a) . C_c(RCCX(a, b; t)) (identity stated in the
    reference's gates.rs:362-376 tests)."""
    n = 4
    crccx = controlled(rccx_matrix(), 1)
    full = embed(crccx, [3, 2, 1, 0], n)
    cs = embed(controlled(S), [3, 2], n)
    return cs @ full


PHI_PLUS = np.array([[S2, 0, S2, 0], [0, S2, 0, S2], [0, S2, 0, -S2], [S2, 0, -S2, 0]], dtype=complex)
PSI_MINUS = np.array([[0, S2, 0, S2], [S2, 0, S2, 0], [-S2, 0, S2, 0], [0, -S2, 0, S2]], dtype=complex)
PSI_PLUS = np.kron(np.diag([1, -1]), I2) @ PSI_MINUS


def gate(inst: dict):
    """(matrix, operand list) of a transport instruction."""
    k = inst["kind"]
    th, ph, la, ga = (inst.get(x, 0.0) for x in ("theta", "phi", "lam", "gamma"))
    single = {"id": I2, "u0": I2, "x": X, "y": Y, "z": Z, "h": H, "s": S, "sdg": SDG, "t": T, "tdg": TDG,
              "sx": SX, "sxdg": SXDG}
    if k in single:
        return single[k], [inst["qubit"]]
    if k in ("u3", "u"):
        return u3(th, ph, la), [inst["qubit"]]
    if k == "u2":
        return u3(math.pi / 2, ph, la), [inst["qubit"]]
    if k in ("u1", "p"):
        return p(la), [inst["qubit"]]
    if k == "rx":
        return rx(th), [inst["qubit"]]
    if k == "ry":
        return ry(th), [inst["qubit"]]
    if k == "rz":
        return rz(ph), [inst["qubit"]]
    ct = [inst.get("control"), inst.get("target")]
    two = {"cx": X, "cy": Y, "cz": Z, "ch": H, "csx": SX}
    if k in two:
        return controlled(two[k]), ct
    if k == "crx":
        return controlled(rx(th)), ct
    if k == "cry":
        return controlled(ry(th)), ct
    if k == "crz":
        return controlled(rz(la)), ct
    if k in ("cu1", "cp"):
        return controlled(p(la)), ct
    if k == "cu3":
        return controlled(u3(th, ph, la)), ct
    if k == "cu":
        return controlled(np.exp(1j * ga) * u3(th, ph, la)), ct
    if k == "swap":
        return SWAP, [inst["a"], inst["b"]]
    if k == "rxx":
        return math.cos(th / 2) * np.eye(4) - 1j * math.sin(th / 2) * np.kron(X, X), [inst["a"], inst["b"]]
    if k == "rzz":
        return np.diag(np.exp(-0.5j * th * np.array([1, -1, -1, 1]))), [inst["a"], inst["b"]]
    if k == "ccx":
        return controlled(X, 2), [inst["control1"], inst["control2"], inst["target"]]
    if k == "cswap":
        return controlled(SWAP, 1), [inst["control"], inst["target1"], inst["target2"]]
    if k == "rccx":
        return rccx_matrix(), [inst["control1"], inst["control2"], inst["target"]]
    if k == "rc3x":
        return rc3x_matrix(), [inst["control1"], inst["control2"], inst["control3"], inst["target"]]
    if k == "c3x":
        return controlled(X, 3), [inst["control1"], inst["control2"], inst["control3"], inst["target"]]
    if k == "c3sqrtx":
        return controlled(SX, 3), [inst["control1"], inst["control2"], inst["control3"], inst["target"]]
    if k == "c4x":
        return controlled(X, 4), [inst[f"control{i}"] for i in range(1, 5)] + [inst["target"]]
    if k == "gate":
        name = inst["name"].lower()
        qs = list(inst["qubits"])
        if name in ("remote_link_phi_plus", "remote_epr", "epr"):
            return PHI_PLUS, qs
        if name == "remote_link_psi_plus":
            return PSI_PLUS, qs
        if name == "remote_link_psi_minus":
            return PSI_MINUS, qs
        if name in ("nonlocal_cz", "remote_cz"):
            return controlled(Z), qs
        if name == "remote_cx":
            return controlled(X), qs
        if name == "remote_rzz":
            return gate({"kind": "rzz", "a": qs[0], "b": qs[1], "theta": inst["params"][0]})
        if name == "remote_cu1":
            return controlled(p(inst["params"][0])), qs
    raise ValueError(f"brute force: unsupported {k}")


def statevector(n: int, instructions: List[dict]) -> np.ndarray:
    psi = np.zeros(1 << n, dtype=complex)
    psi[0] = 1.0
    for inst in instructions:
        if inst["kind"] == "barrier" or (inst["kind"] == "gate" and inst["name"].lower() == "remote_barrier"):
            continue
        m, qs = gate(inst)
        psi = embed(m, qs, n) @ psi
    return psi


def dense_reference(state: np.ndarray, matrix: np.ndarray, qubits: Sequence[int]) -> np.ndarray:
    """Out-of-place dense application: each output basis index sums over the
    input indices that differ on the gate's qubits (qubits[0] = MSB)."""
    n = state.size.bit_length() - 1
    return embed(np.asarray(matrix, dtype=complex), list(qubits), n) @ state


def fourier(dim: int) -> np.ndarray:
    j = np.arange(dim)
    return np.exp(2j * np.pi * np.outer(j, j) / dim) / math.sqrt(dim)


def test_input(n: int) -> np.ndarray:
    """Deterministic normalised input sin(i + 0.25) + i cos(i + 0.5)."""
    i = np.arange(1 << n, dtype=float)
    s = np.sin(i + 0.25) + 1j * np.cos(i + 0.5)
    return s / np.linalg.norm(s)
