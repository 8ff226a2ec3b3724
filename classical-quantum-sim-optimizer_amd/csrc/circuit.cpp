// Transport parsing, validation, gate matrices and compile.
// See circuit.hpp for the reference files each part follows.
#include "circuit.hpp"

#include <algorithm>
#include <cmath>
#include <map>

#include "json.hpp"

namespace dq {

namespace {

[[noreturn]] void parse_fail(const std::string& what) {
  throw ValueErr("Circuit JSON parse error: " + what);
}

const JVal& need(const JVal& obj, const char* key) {
  const JVal* v = obj.get(key);
  if (!v) parse_fail(std::string("missing field `") + key + "`");
  return *v;
}

void check_duplicates(const JVal& obj) {
  for (size_t i = 0; i < obj.obj.size(); ++i)
    for (size_t j = 0; j < i; ++j)
      if (obj.obj[i].first == obj.obj[j].first)
        parse_fail("duplicate field `" + obj.obj[i].first + "`");
}

uint64_t as_u64(const JVal& v, const char* key) {
  if (v.type != JVal::NUM) parse_fail(std::string("invalid type for `") + key + "`, expected an unsigned integer");
  if (!v.is_int) parse_fail(std::string("invalid type: floating point for `") + key + "`, expected an unsigned integer");
  if (v.overflow) parse_fail(std::string("integer out of range for `") + key + "`");
  if (v.neg && v.u != 0) parse_fail(std::string("invalid value: negative integer for `") + key + "`");
  return v.u;
}
uint64_t get_u64(const JVal& obj, const char* key) { return as_u64(need(obj, key), key); }

double get_f64(const JVal& obj, const char* key) {
  const JVal& v = need(obj, key);
  if (v.type != JVal::NUM) parse_fail(std::string("invalid type for `") + key + "`, expected f64");
  return v.d;
}

std::string get_str(const JVal& obj, const char* key) {
  const JVal& v = need(obj, key);
  if (v.type != JVal::STR) parse_fail(std::string("invalid type for `") + key + "`, expected a string");
  return v.s;
}

struct KindInfo {
  const char* tag;
  Kind kind;
  const char* operands[5];
  const char* params[4];
};

// Field names per instruction kind (src/types.rs:60-296) with operands in
// Instruction::qubits() order (src/types.rs:346-464).
const KindInfo kKinds[] = {
    {"id", Kind::Id, {"qubit"}, {}},
    {"x", Kind::X, {"qubit"}, {}},
    {"y", Kind::Y, {"qubit"}, {}},
    {"z", Kind::Z, {"qubit"}, {}},
    {"h", Kind::H, {"qubit"}, {}},
    {"s", Kind::S, {"qubit"}, {}},
    {"sdg", Kind::Sdg, {"qubit"}, {}},
    {"t", Kind::T, {"qubit"}, {}},
    {"tdg", Kind::Tdg, {"qubit"}, {}},
    {"sx", Kind::Sx, {"qubit"}, {}},
    {"sxdg", Kind::Sxdg, {"qubit"}, {}},
    {"u3", Kind::U3, {"qubit"}, {"theta", "phi", "lam"}},
    {"u2", Kind::U2, {"qubit"}, {"phi", "lam"}},
    {"u1", Kind::U1, {"qubit"}, {"lam"}},
    {"u", Kind::U, {"qubit"}, {"theta", "phi", "lam"}},
    {"p", Kind::P, {"qubit"}, {"lam"}},
    {"rx", Kind::Rx, {"qubit"}, {"theta"}},
    {"ry", Kind::Ry, {"qubit"}, {"theta"}},
    {"rz", Kind::Rz, {"qubit"}, {"phi"}},
    {"u0", Kind::U0, {"qubit"}, {}},
    {"cx", Kind::Cx, {"control", "target"}, {}},
    {"cz", Kind::Cz, {"control", "target"}, {}},
    {"cy", Kind::Cy, {"control", "target"}, {}},
    {"ch", Kind::Ch, {"control", "target"}, {}},
    {"swap", Kind::Swap, {"a", "b"}, {}},
    {"csx", Kind::Csx, {"control", "target"}, {}},
    {"crx", Kind::Crx, {"control", "target"}, {"theta"}},
    {"cry", Kind::Cry, {"control", "target"}, {"theta"}},
    {"crz", Kind::Crz, {"control", "target"}, {"lam"}},
    {"cu1", Kind::Cu1, {"control", "target"}, {"lam"}},
    {"cp", Kind::Cp, {"control", "target"}, {"lam"}},
    {"cu3", Kind::Cu3, {"control", "target"}, {"theta", "phi", "lam"}},
    {"cu", Kind::Cu, {"control", "target"}, {"theta", "phi", "lam", "gamma"}},
    {"rxx", Kind::Rxx, {"a", "b"}, {"theta"}},
    {"rzz", Kind::Rzz, {"a", "b"}, {"theta"}},
    {"ccx", Kind::Ccx, {"control1", "control2", "target"}, {}},
    {"cswap", Kind::Cswap, {"control", "target1", "target2"}, {}},
    {"rccx", Kind::Rccx, {"control1", "control2", "target"}, {}},
    {"rc3x", Kind::Rc3x, {"control1", "control2", "control3", "target"}, {}},
    {"c3x", Kind::C3x, {"control1", "control2", "control3", "target"}, {}},
    {"c3sqrtx", Kind::C3sqrtx, {"control1", "control2", "control3", "target"}, {}},
    {"c4x", Kind::C4x, {"control1", "control2", "control3", "control4", "target"}, {}},
};

Instruction parse_instruction(const JVal& v, int depth) {
  if (v.type != JVal::OBJ) parse_fail("instruction must be an object");
  if (depth > 128) parse_fail("conditional nesting too deep");
  check_duplicates(v);
  std::string tag = get_str(v, "kind");
  Instruction in;
  for (const KindInfo& k : kKinds) {
    if (tag != k.tag) continue;
    in.kind = k.kind;
    for (const char* o : k.operands)
      if (o) in.qubits.push_back(get_u64(v, o));
    for (const char* p : k.params) {
      if (!p) continue;
      double x = get_f64(v, p);
      if (!std::strcmp(p, "theta")) in.theta = x;
      else if (!std::strcmp(p, "phi")) in.phi = x;
      else if (!std::strcmp(p, "lam")) in.lam = x;
      else in.gamma = x;
    }
    return in;
  }
  if (tag == "gate") {
    in.kind = Kind::Gate;
    in.name = get_str(v, "name");
    const JVal& ps = need(v, "params");
    if (ps.type != JVal::ARR) parse_fail("invalid type for `params`, expected a sequence");
    for (const JVal& x : ps.arr) {
      if (x.type != JVal::NUM) parse_fail("invalid type in `params`, expected f64");
      in.params.push_back(x.d);
    }
    const JVal& qs = need(v, "qubits");
    if (qs.type != JVal::ARR) parse_fail("invalid type for `qubits`, expected a sequence");
    for (const JVal& x : qs.arr) in.qubits.push_back(as_u64(x, "qubits"));
    return in;
  }
  if (tag == "measure") {
    in.kind = Kind::Measure;
    in.qubits.push_back(get_u64(v, "qubit"));
    in.cbit = get_u64(v, "cbit");
    return in;
  }
  if (tag == "reset") {
    in.kind = Kind::Reset;
    in.qubits.push_back(get_u64(v, "qubit"));
    return in;
  }
  if (tag == "conditional") {
    in.kind = Kind::Conditional;
    const JVal& c = need(v, "condition");
    if (c.type != JVal::OBJ) parse_fail("invalid type for `condition`");
    check_duplicates(c);
    in.cond.creg_base = get_u64(c, "creg_base");
    in.cond.creg_size = get_u64(c, "creg_size");
    in.cond.creg_value = get_u64(c, "creg_value");
    in.op = std::make_shared<Instruction>(parse_instruction(need(v, "op"), depth + 1));
    in.qubits = in.op->qubits;
    return in;
  }
  if (tag == "barrier") {
    in.kind = Kind::Barrier;
    return in;
  }
  if (tag == "classical") {
    in.kind = Kind::Classical;
    in.name = get_str(v, "name");
    return in;
  }
  parse_fail("unknown variant `" + tag + "`");
}

std::vector<Register> parse_registers(const JVal& v, const char* what) {
  if (v.type != JVal::OBJ) parse_fail(std::string("invalid type for `") + what + "`, expected a map");
  std::vector<Register> regs;
  for (auto& kv : v.obj) {
    const JVal& r = kv.second;
    if (r.type != JVal::OBJ) parse_fail("register must be an object");
    check_duplicates(r);
    Register reg;
    reg.key = kv.first;
    reg.name = get_str(r, "name");
    reg.size = get_u64(r, "size");
    reg.base = get_u64(r, "base");
    // A map with a repeated key keeps the last value.
    auto it = std::find_if(regs.begin(), regs.end(), [&](const Register& x) { return x.key == reg.key; });
    if (it != regs.end()) *it = reg;
    else regs.push_back(reg);
  }
  return regs;
}

using Ranges = std::vector<std::pair<uint64_t, uint64_t>>;

Ranges register_ranges(const std::vector<Register>& regs, const char* kind) {
  Ranges ranges;
  for (const Register& r : regs) {
    if (r.base > UINT64_MAX - r.size)
      throw ValueErr(std::string(kind) + " register \"" + r.key + "\" has an overflowing range");
    if (r.size != 0) ranges.emplace_back(r.base, r.base + r.size);
  }
  std::sort(ranges.begin(), ranges.end());
  for (size_t i = 1; i < ranges.size(); ++i)
    if (ranges[i].first < ranges[i - 1].second)
      throw ValueErr(std::string("Overlapping ") + kind + " registers");
  return ranges;
}

bool declared(uint64_t index, const Ranges& r) {
  for (auto& x : r)
    if (x.first <= index && index < x.second) return true;
  return false;
}

std::string lower(std::string s) {
  for (char& c : s)
    if (c >= 'A' && c <= 'Z') c = (char)(c - 'A' + 'a');
  return s;
}

// src/validation.rs:44-182
void validate_instruction(const Instruction& in, const Ranges& qubits, const Ranges& cbits) {
  if (in.kind == Kind::Conditional) {
    const Condition& c = in.cond;
    if (c.creg_size == 0 || c.creg_size > 64) throw ValueErr("Condition width must be between 1 and 64 bits");
    if (c.creg_base > UINT64_MAX - c.creg_size) throw ValueErr("Condition has an overflowing classical range");
    for (uint64_t bit = c.creg_base; bit < c.creg_base + c.creg_size; ++bit)
      if (!declared(bit, cbits)) throw ValueErr("Condition refers to undeclared classical bits");
    if (c.creg_size < 64 && c.creg_value >= (1ull << c.creg_size))
      throw ValueErr("Condition value does not fit its classical width");
    validate_instruction(*in.op, qubits, cbits);
    return;
  }
  for (size_t i = 0; i < in.qubits.size(); ++i) {
    uint64_t q = in.qubits[i];
    if (!declared(q, qubits)) throw ValueErr("Qubit " + std::to_string(q) + " is not declared by a quantum register");
    for (size_t j = 0; j < i; ++j)
      if (in.qubits[j] == q)
        throw ValueErr("Gate operands must be distinct; qubit " + std::to_string(q) + " is repeated");
  }
  if (in.kind == Kind::Measure && !declared(in.cbit, cbits))
    throw ValueErr("Classical bit " + std::to_string(in.cbit) + " is not declared by a classical register");
  if (in.kind == Kind::Classical) throw ValueErr("Unsupported classical operation: \"" + in.name + "\"");
  if (in.kind == Kind::Gate) {
    std::string n = lower(in.name);
    int arity = -1;
    size_t nparams = 0;
    if (n == "remote_link_phi_plus" || n == "remote_link_psi_minus" || n == "remote_link_psi_plus" ||
        n == "nonlocal_cz" || n == "remote_cz" || n == "remote_cx" || n == "remote_epr" || n == "epr") {
      arity = 2;
    } else if (n == "remote_rzz" || n == "remote_cu1") {
      arity = 2;
      nparams = 1;
    } else if (n == "remote_barrier") {
      arity = -1;
    } else {
      throw ValueErr("Unsupported generic gate: \"" + in.name + "\". Decompose it before simulating.");
    }
    if (arity >= 0 && in.qubits.size() != (size_t)arity)
      throw ValueErr("Gate \"" + in.name + "\" requires " + std::to_string(arity) + " qubits, received " +
                     std::to_string(in.qubits.size()));
    if (in.params.size() != nparams)
      throw ValueErr("Gate \"" + in.name + "\" requires " + std::to_string(nparams) + " parameters, received " +
                     std::to_string(in.params.size()));
  }
  bool finite = true;
  switch (in.kind) {
    case Kind::U3: case Kind::U: case Kind::Cu3:
      finite = std::isfinite(in.theta) && std::isfinite(in.phi) && std::isfinite(in.lam);
      break;
    case Kind::Cu:
      finite = std::isfinite(in.theta) && std::isfinite(in.phi) && std::isfinite(in.lam) && std::isfinite(in.gamma);
      break;
    case Kind::U2: finite = std::isfinite(in.phi) && std::isfinite(in.lam); break;
    case Kind::U1: case Kind::P: case Kind::Crz: case Kind::Cu1: case Kind::Cp: finite = std::isfinite(in.lam); break;
    case Kind::Rx: case Kind::Ry: case Kind::Crx: case Kind::Cry: case Kind::Rxx: case Kind::Rzz:
      finite = std::isfinite(in.theta);
      break;
    case Kind::Rz: finite = std::isfinite(in.phi); break;
    case Kind::Gate:
      for (double x : in.params) finite = finite && std::isfinite(x);
      break;
    default: break;
  }
  if (!finite) throw ValueErr("Gate parameters must be finite numbers");
  switch (in.kind) {
    case Kind::U3: case Kind::U2: case Kind::U: case Kind::Cu3: case Kind::Cu:
      if (!std::isfinite(in.phi + in.lam)) throw ValueErr("The combined phase phi + lam must be finite");
      break;
    default: break;
  }
}

}  // namespace

uint64_t Circuit::num_qubits() const {
  uint64_t m = 0;
  for (auto& r : qregs) m = std::max(m, r.base + r.size);
  return m;
}
uint64_t Circuit::num_cbits() const {
  uint64_t m = 0;
  for (auto& r : cregs) m = std::max(m, r.base + r.size);
  return m;
}

Circuit parse_circuit(const char* json, size_t len) {
  JVal root;
  try {
    root = JsonParser(json, len).parse();
  } catch (const JsonError& e) {
    throw ValueErr(e.what());
  }
  if (root.type != JVal::OBJ) parse_fail("circuit must be an object");
  check_duplicates(root);
  Circuit c;
  c.qregs = parse_registers(need(root, "qregs"), "qregs");
  if (const JVal* cr = root.get("cregs")) c.cregs = parse_registers(*cr, "cregs");
  const JVal& insts = need(root, "instructions");
  if (insts.type != JVal::ARR) parse_fail("invalid type for `instructions`, expected a sequence");
  c.instructions.reserve(insts.arr.size());
  for (const JVal& v : insts.arr) c.instructions.push_back(parse_instruction(v, 0));
  return c;
}

// src/validation.rs:5-13 then src/monolithic/statevector/validation.rs:3-12
void validate_circuit(const Circuit& c) {
  Ranges q = register_ranges(c.qregs, "quantum");
  Ranges b = register_ranges(c.cregs, "classical");
  for (size_t i = 0; i < c.instructions.size(); ++i) {
    try {
      validate_instruction(c.instructions[i], q, b);
    } catch (const ValueErr& e) {
      throw ValueErr("Instruction " + std::to_string(i) + ": " + e.msg);
    }
  }
  if (c.num_qubits() >= 64) throw ValueErr("Statevector qubit count must be less than 64");
}

// ---------------------------------------------------------------------------
// Gate matrices (src/gates.rs:23-274). Expressions mirror the reference's
// arithmetic so the FP64 values are bit-identical.
// ---------------------------------------------------------------------------
namespace gates {
namespace {
constexpr double kS = 0.70710678118654752440;  // FRAC_1_SQRT_2
constexpr double kPi = 3.14159265358979323846;
cplx r(double x) { return cplx(x, 0.0); }
cplx neg(cplx a) { return cplx(-a.re, -a.im); }
cplx ei(double x) { return cplx(std::cos(x), std::sin(x)); }
}  // namespace

Mat2 X() { return {r(0), r(1), r(1), r(0)}; }
Mat2 Y() { return {r(0), cplx(0, -1), cplx(0, 1), r(0)}; }
Mat2 Z() { return {r(1), r(0), r(0), r(-1)}; }
Mat2 h() { return {r(kS), r(kS), r(kS), r(-kS)}; }
Mat2 s_gate() { return {r(1), r(0), r(0), cplx(0, 1)}; }
Mat2 sdg() { return {r(1), r(0), r(0), cplx(0, -1)}; }
Mat2 t_gate() { return {r(1), r(0), r(0), cplx(kS, kS)}; }
Mat2 tdg() { return {r(1), r(0), r(0), cplx(kS, -kS)}; }
Mat2 sx() { return {cplx(0.5, 0.5), cplx(0.5, -0.5), cplx(0.5, -0.5), cplx(0.5, 0.5)}; }
Mat2 sxdg() { return {cplx(0.5, -0.5), cplx(0.5, 0.5), cplx(0.5, 0.5), cplx(0.5, -0.5)}; }
Mat2 u3(double theta, double phi, double lam) {
  double ct = std::cos(theta / 2.0), st = std::sin(theta / 2.0);
  return {r(ct), neg(cmul(ei(lam), r(st))), cmul(ei(phi), r(st)), cmul(ei(phi + lam), r(ct))};
}
Mat2 u2(double phi, double lam) { return u3(kPi / 2.0, phi, lam); }
Mat2 p(double lam) { return {r(1), r(0), r(0), ei(lam)}; }
Mat2 rx(double theta) {
  double c2 = std::cos(theta / 2.0), s2 = std::sin(theta / 2.0);
  return {r(c2), cplx(0, -s2), cplx(0, -s2), r(c2)};
}
Mat2 ry(double theta) {
  double c2 = std::cos(theta / 2.0), s2 = std::sin(theta / 2.0);
  return {r(c2), r(-s2), r(s2), r(c2)};
}
Mat2 rz(double phi) { return {ei(-phi / 2.0), r(0), r(0), ei(phi / 2.0)}; }

static std::vector<cplx> identity(int dim) {
  std::vector<cplx> m((size_t)dim * dim, r(0));
  for (int i = 0; i < dim; ++i) m[(size_t)i * dim + i] = r(1);
  return m;
}
std::vector<cplx> rxx(double theta) {
  double cv = std::cos(theta / 2.0), sv = std::sin(theta / 2.0);
  cplx o = r(0), c = r(cv), s = cplx(0, -sv);
  return {c, o, o, s, o, c, s, o, o, s, c, o, s, o, o, c};
}
std::vector<cplx> rzz(double theta) {
  cplx ep(std::cos(theta / 2.0), std::sin(theta / 2.0));
  cplx em(std::cos(theta / 2.0), -std::sin(theta / 2.0));
  cplx o = r(0);
  return {em, o, o, o, o, ep, o, o, o, o, ep, o, o, o, o, em};
}
std::vector<cplx> rccx() {
  auto m = identity(8);
  m[5 * 8 + 5] = r(-1);
  m[6 * 8 + 6] = r(0);
  m[7 * 8 + 7] = r(0);
  m[6 * 8 + 7] = cplx(0, -1);
  m[7 * 8 + 6] = cplx(0, 1);
  return m;
}
std::vector<cplx> rc3x() {
  auto m = identity(16);
  m[12 * 16 + 12] = cplx(0, 1);
  m[13 * 16 + 13] = cplx(0, -1);
  m[14 * 16 + 14] = r(0);
  m[15 * 16 + 15] = r(0);
  m[14 * 16 + 15] = r(1);
  m[15 * 16 + 14] = r(-1);
  return m;
}
std::vector<cplx> psi_minus() {
  cplx o = r(0), s = r(kS), n = r(-kS);
  return {o, s, o, s, s, o, s, o, n, o, s, o, o, n, o, s};
}
std::vector<cplx> psi_plus() {
  auto m = psi_minus();
  for (int j = 8; j < 16; ++j) m[j] = neg(m[j]);
  return m;
}
std::vector<cplx> phi_plus() {
  cplx o = r(0), s = r(kS), n = r(-kS);
  return {s, o, s, o, o, s, o, s, o, s, o, n, s, o, n, o};
}
}  // namespace gates

// ---------------------------------------------------------------------------
// Compile (src/monolithic/statevector/execution.rs:66-274, 368-416)
// ---------------------------------------------------------------------------
namespace {

bool mat_is_x(const Mat2& m) {
  Mat2 x = gates::X();
  for (int i = 0; i < 4; ++i)
    if (!ceq(m[i], x[i])) return false;
  return true;
}

// Operation::one (execution.rs:66-87)
Op make_one(uint32_t target, const Mat2& m, std::vector<std::pair<uint32_t, bool>> controls) {
  Op op;
  const cplx zero(0, 0);
  if (mat_is_x(m)) {
    op.kind = OpKind::X;
    op.target = target;
    op.controls = std::move(controls);
  } else if (ceq(m[1], zero) && ceq(m[2], zero)) {
    op.kind = OpKind::Diagonal;
    for (auto& c : controls) op.qubits.push_back(c.first);
    op.qubits.push_back(target);
    op.mat.assign((size_t)1 << op.qubits.size(), cplx(1, 0));
    size_t index = 0;
    for (auto& c : controls) index = (index << 1) | (c.second ? 1 : 0);
    index <<= 1;
    op.mat[index] = m[0];
    op.mat[index + 1] = m[3];
  } else {
    op.kind = OpKind::One;
    op.target = target;
    op.mat.assign(m.begin(), m.end());
    op.controls = std::move(controls);
  }
  return op;
}

// Operation::dense (execution.rs:89-106)
Op make_dense(std::vector<uint32_t> qubits, const std::vector<cplx>& m) {
  size_t dim = (size_t)1 << qubits.size();
  bool diag = true;
  for (size_t i = 0; i < dim && diag; ++i)
    for (size_t j = 0; j < dim; ++j)
      if (i != j && !ceq(m[i * dim + j], cplx(0, 0))) {
        diag = false;
        break;
      }
  Op op;
  op.qubits = std::move(qubits);
  if (diag) {
    op.kind = OpKind::Diagonal;
    for (size_t i = 0; i < dim; ++i) op.mat.push_back(m[i * dim + i]);
  } else {
    op.kind = OpKind::Dense;
    op.mat = m;
  }
  return op;
}

// gate_matrix_1q (src/types.rs:303-344)
bool matrix_1q(const Instruction& in, uint32_t& q, Mat2& m) {
  switch (in.kind) {
    case Kind::Id: case Kind::U0: m = {cplx(1, 0), cplx(0, 0), cplx(0, 0), cplx(1, 0)}; break;
    case Kind::X: m = gates::X(); break;
    case Kind::Y: m = gates::Y(); break;
    case Kind::Z: m = gates::Z(); break;
    case Kind::H: m = gates::h(); break;
    case Kind::S: m = gates::s_gate(); break;
    case Kind::Sdg: m = gates::sdg(); break;
    case Kind::T: m = gates::t_gate(); break;
    case Kind::Tdg: m = gates::tdg(); break;
    case Kind::Sx: m = gates::sx(); break;
    case Kind::Sxdg: m = gates::sxdg(); break;
    case Kind::U3: case Kind::U: m = gates::u3(in.theta, in.phi, in.lam); break;
    case Kind::U2: m = gates::u2(in.phi, in.lam); break;
    case Kind::U1: case Kind::P: m = gates::p(in.lam); break;
    case Kind::Rx: m = gates::rx(in.theta); break;
    case Kind::Ry: m = gates::ry(in.theta); break;
    case Kind::Rz: m = gates::rz(in.phi); break;
    default: return false;
  }
  q = (uint32_t)in.qubits[0];
  return true;
}

// Returns false for a no-op (barrier, remote_barrier).
bool compile_one(const Instruction& in, Op& out) {
  uint32_t q;
  Mat2 m;
  if (matrix_1q(in, q, m)) {
    out = make_one(q, m, {});
    return true;
  }
  auto Q = [&](size_t i) { return (uint32_t)in.qubits[i]; };
  auto controlled = [&](const Mat2& u) { return make_one(Q(1), u, {{Q(0), true}}); };
  switch (in.kind) {
    case Kind::Cx: out = controlled(gates::X()); return true;
    case Kind::Cy: out = controlled(gates::Y()); return true;
    case Kind::Cz: out = controlled(gates::Z()); return true;
    case Kind::Ch: out = controlled(gates::h()); return true;
    case Kind::Csx: out = controlled(gates::sx()); return true;
    case Kind::Crx: out = controlled(gates::rx(in.theta)); return true;
    case Kind::Cry: out = controlled(gates::ry(in.theta)); return true;
    case Kind::Crz: out = controlled(gates::rz(in.lam)); return true;
    case Kind::Cu1: case Kind::Cp: out = controlled(gates::p(in.lam)); return true;
    case Kind::Cu3: out = controlled(gates::u3(in.theta, in.phi, in.lam)); return true;
    case Kind::Cu: {
      Mat2 u = gates::u3(in.theta, in.phi, in.lam);
      cplx phase(std::cos(in.gamma), std::sin(in.gamma));
      for (auto& v : u) v = cmul(v, phase);
      out = controlled(u);
      return true;
    }
    case Kind::Swap:
      out = Op();
      out.kind = OpKind::Swap;
      out.a = Q(0);
      out.b = Q(1);
      return true;
    case Kind::Cswap:
      out = Op();
      out.kind = OpKind::Swap;
      out.a = Q(1);
      out.b = Q(2);
      out.controls = {{Q(0), true}};
      return true;
    case Kind::Ccx: case Kind::C3x: case Kind::C4x: {
      out = Op();
      out.kind = OpKind::X;
      size_t nc = in.qubits.size() - 1;
      for (size_t i = 0; i < nc; ++i) out.controls.push_back({Q(i), true});
      out.target = Q(nc);
      return true;
    }
    case Kind::C3sqrtx:
      out = make_one(Q(3), gates::sx(), {{Q(0), true}, {Q(1), true}, {Q(2), true}});
      return true;
    case Kind::Rxx: out = make_dense({Q(0), Q(1)}, gates::rxx(in.theta)); return true;
    case Kind::Rzz: out = make_dense({Q(0), Q(1)}, gates::rzz(in.theta)); return true;
    case Kind::Rccx: out = make_dense({Q(0), Q(1), Q(2)}, gates::rccx()); return true;
    case Kind::Rc3x: out = make_dense({Q(0), Q(1), Q(2), Q(3)}, gates::rc3x()); return true;
    case Kind::Measure:
      out = Op();
      out.kind = OpKind::Measure;
      out.qubit = Q(0);
      out.cbit = in.cbit;
      return true;
    case Kind::Reset:
      out = Op();
      out.kind = OpKind::Reset;
      out.qubit = Q(0);
      return true;
    case Kind::Conditional: {
      // A conditional no-op stays a (non-unitary) conditional, as in the
      // reference: it still ends the deterministic prefix.
      Op inner;
      if (!compile_one(*in.op, inner)) inner.kind = OpKind::Noop;
      out = Op();
      out.kind = OpKind::Conditional;
      out.cond = in.cond;
      out.inner = std::make_shared<Op>(std::move(inner));
      return true;
    }
    case Kind::Barrier: return false;
    case Kind::Gate: {
      std::string n = lower(in.name);
      std::vector<uint32_t> qs;
      for (size_t i = 0; i < in.qubits.size(); ++i) qs.push_back(Q(i));
      if (n == "remote_link_phi_plus" || n == "remote_epr" || n == "epr") out = make_dense(qs, gates::phi_plus());
      else if (n == "remote_link_psi_plus") out = make_dense(qs, gates::psi_plus());
      else if (n == "remote_link_psi_minus") out = make_dense(qs, gates::psi_minus());
      else if (n == "nonlocal_cz" || n == "remote_cz") out = controlled(gates::Z());
      else if (n == "remote_cx") out = controlled(gates::X());
      else if (n == "remote_rzz") out = make_dense(qs, gates::rzz(in.params[0]));
      else if (n == "remote_cu1") out = controlled(gates::p(in.params[0]));
      else if (n == "remote_barrier") return false;
      else throw ValueErr("Unsupported generic gate \"" + in.name + "\"; decompose it before simulating");
      return true;
    }
    default: throw ValueErr("Unsupported instruction");
  }
}

}  // namespace

std::vector<Op> compile(const std::vector<Instruction>& insts) {
  std::map<uint32_t, Mat2> pending;  // ordered like the reference's BTreeMap
  std::vector<Op> result;
  const Mat2 ident = {cplx(1, 0), cplx(0, 0), cplx(0, 0), cplx(1, 0)};
  auto flush = [&](uint32_t q) {
    auto it = pending.find(q);
    if (it == pending.end()) return;
    Mat2 m = it->second;
    pending.erase(it);
    bool is_ident = true;
    for (int i = 0; i < 4; ++i) is_ident = is_ident && ceq(m[i], ident[i]);
    if (!is_ident) result.push_back(make_one(q, m, {}));
  };
  for (const Instruction& in : insts) {
    uint32_t q;
    Mat2 m;
    if (matrix_1q(in, q, m)) {
      auto it = pending.find(q);
      Mat2 old = it == pending.end() ? ident : it->second;
      Mat2 prod;
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
          cplx acc(0, 0);
          for (int k = 0; k < 2; ++k) acc = cadd(acc, cmul(m[i * 2 + k], old[k * 2 + j]));
          prod[i * 2 + j] = acc;
        }
      pending[q] = prod;
      continue;
    }
    if (in.kind == Kind::Barrier || in.kind == Kind::Measure) {
      std::vector<uint32_t> keys;
      for (auto& kv : pending) keys.push_back(kv.first);
      for (uint32_t k : keys) flush(k);
    } else {
      for (uint64_t t : in.qubits) flush((uint32_t)t);
    }
    Op op;
    if (compile_one(in, op)) result.push_back(std::move(op));
  }
  std::vector<uint32_t> keys;
  for (auto& kv : pending) keys.push_back(kv.first);
  for (uint32_t k : keys) flush(k);
  return result;
}

// Helpers shared with the P-block backend (pblock.cpp).
std::vector<std::pair<uint64_t, uint64_t>> quantum_ranges(const Circuit& c) {
  return register_ranges(c.qregs, "quantum");
}
void validate_circuit_rules(const Circuit& c) {
  Ranges q = register_ranges(c.qregs, "quantum");
  Ranges b = register_ranges(c.cregs, "classical");
  for (size_t i = 0; i < c.instructions.size(); ++i) {
    try {
      validate_instruction(c.instructions[i], q, b);
    } catch (const ValueErr& e) {
      throw ValueErr("Instruction " + std::to_string(i) + ": " + e.msg);
    }
  }
}
bool one_qubit_matrix(const Instruction& in, uint32_t& q, Mat2& m) { return matrix_1q(in, q, m); }
bool compile_instruction(const Instruction& in, Op& op) { return compile_one(in, op); }

}  // namespace dq
