// Circuit transport IR, validation, gate matrices and the statevector
// compile step (1-qubit fusion and lowering to kernel operations).
//
// Semantics follow the reference crate:
//   transport schema ........ src/types.rs:7-296
//   operand order ........... src/types.rs:346-464 (Instruction::qubits)
//   validation .............. src/validation.rs:5-182,
//                             src/monolithic/statevector/validation.rs:3-12
//   gate matrices ........... src/gates.rs:23-274, src/types.rs:303-344
//   compile / lowering ...... src/monolithic/statevector/execution.rs:21-274,368-416
#pragma once

#include <array>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace dq {

// Complex arithmetic is written out by hand everywhere (no std::complex
// operators) so that the operation order is explicit and identical to the
// device kernels and to num-complex: mul = (ar*br - ai*bi, ar*bi + ai*br).
struct cplx {
  double re = 0.0, im = 0.0;
  cplx() = default;
  constexpr cplx(double r, double i) : re(r), im(i) {}
};
inline cplx cmul(cplx a, cplx b) { return cplx(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
inline cplx cadd(cplx a, cplx b) { return cplx(a.re + b.re, a.im + b.im); }
inline bool ceq(cplx a, cplx b) { return a.re == b.re && a.im == b.im; }

using Mat2 = std::array<cplx, 4>;  // row-major [[m0,m1],[m2,m3]]

enum class Kind : uint8_t {
  Id, X, Y, Z, H, S, Sdg, T, Tdg, Sx, Sxdg,
  U3, U2, U1, U, P, Rx, Ry, Rz, U0,
  Cx, Cz, Cy, Ch, Swap, Csx,
  Crx, Cry, Crz, Cu1, Cp, Cu3, Cu, Rxx, Rzz,
  Ccx, Cswap, Rccx, Rc3x, C3x, C3sqrtx, C4x,
  Gate, Measure, Reset, Conditional, Barrier, Classical,
};

struct Condition {
  uint64_t creg_base = 0, creg_size = 0, creg_value = 0;
};

struct Instruction {
  Kind kind = Kind::Barrier;
  // Operands in the order of the reference's Instruction::qubits().
  std::vector<uint64_t> qubits;
  double theta = 0, phi = 0, lam = 0, gamma = 0;
  std::string name;            // Gate / Classical
  std::vector<double> params;  // Gate
  uint64_t cbit = 0;           // Measure
  Condition cond;              // Conditional
  std::shared_ptr<Instruction> op;  // Conditional body
};

struct Register {
  std::string key, name;
  uint64_t base = 0, size = 0;
};

struct Circuit {
  std::vector<Register> qregs, cregs;
  std::vector<Instruction> instructions;
  uint64_t num_qubits() const;
  uint64_t num_cbits() const;
};

// Parse the JSON transport. Throws ValueErr on malformed input.
Circuit parse_circuit(const char* json, size_t len);
// Reference validation, plus the statevector n < 64 rule. Throws ValueErr.
void validate_circuit(const Circuit& c);

// ---- compiled operations (execution.rs:21-56) ----
enum class OpKind : uint8_t { One, X, Swap, Diagonal, Dense, Measure, Reset, Conditional, Noop };

struct Op {
  OpKind kind = OpKind::One;
  uint32_t target = 0, a = 0, b = 0, qubit = 0;
  uint64_t cbit = 0;
  std::vector<std::pair<uint32_t, bool>> controls;
  std::vector<uint32_t> qubits;  // Diagonal / Dense, qubits[0] = MSB
  std::vector<cplx> mat;         // One: 4; Diagonal: 2^k; Dense: 4^k row-major
  Condition cond;
  std::shared_ptr<Op> inner;
  bool is_unitary() const {
    return kind != OpKind::Measure && kind != OpKind::Reset && kind != OpKind::Conditional;
  }
};

std::vector<Op> compile(const std::vector<Instruction>& insts);

// Pieces of the above for the P-block backend: sorted quantum register
// ranges, the reference validation without the statevector qubit-count
// limit, gate_matrix_1q (src/types.rs:303-344) and one instruction compiled
// alone (execution.rs:66-274 per-instruction rules; false for no-ops).
std::vector<std::pair<uint64_t, uint64_t>> quantum_ranges(const Circuit& c);
void validate_circuit_rules(const Circuit& c);
bool one_qubit_matrix(const Instruction& in, uint32_t& q, Mat2& m);
bool compile_instruction(const Instruction& in, Op& op);

// Gate library.
namespace gates {
Mat2 X();
Mat2 Y();
Mat2 Z();
Mat2 h();
Mat2 s_gate();
Mat2 sdg();
Mat2 t_gate();
Mat2 tdg();
Mat2 sx();
Mat2 sxdg();
Mat2 u3(double theta, double phi, double lam);
Mat2 u2(double phi, double lam);
Mat2 p(double lam);
Mat2 rx(double theta);
Mat2 ry(double theta);
Mat2 rz(double phi);
std::vector<cplx> rxx(double theta);
std::vector<cplx> rzz(double theta);
std::vector<cplx> rccx();
std::vector<cplx> rc3x();
std::vector<cplx> psi_minus();
std::vector<cplx> psi_plus();
std::vector<cplx> phi_plus();
}  // namespace gates

struct ValueErr : std::exception {
  std::string msg;
  explicit ValueErr(std::string m) : msg(std::move(m)) {}
  const char* what() const noexcept override { return msg.c_str(); }
};

}  // namespace dq
