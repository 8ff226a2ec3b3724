/*
 * CPU FP64 oracle for the dense statevector path -- TEST INFRASTRUCTURE ONLY.
 * See sv_oracle.h for what it restates and the canonical choices it shares
 * with the engine. Built with -ffp-contract=off (oracle/Makefile).
 */
#include "sv_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Complex helpers (num-complex operation order)                            */
/* ------------------------------------------------------------------------ */
typedef struct {
  double re, im;
} cx;

static inline cx cmul(cx a, cx b) {
  cx r;
  r.re = a.re * b.re - a.im * b.im;
  r.im = a.re * b.im + a.im * b.re;
  return r;
}
static inline cx cadd(cx a, cx b) {
  cx r;
  r.re = a.re + b.re;
  r.im = a.im + b.im;
  return r;
}
static inline cx C(double re, double im) {
  cx r;
  r.re = re;
  r.im = im;
  return r;
}
static inline int ceq(cx a, cx b) { return a.re == b.re && a.im == b.im; }

void or_set_threads(int t) {
#ifdef _OPENMP
  if (t > 0) omp_set_num_threads(t);
#else
  (void)t;
#endif
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon et al., SC'11)                                     */
/* ------------------------------------------------------------------------ */
void or_philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int round = 0; round < 10; ++round) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c[0] = n0;
    c[1] = n1;
    c[2] = n2;
    c[3] = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

double or_uniform53(uint64_t seed, uint64_t stream, uint64_t draw) {
  uint32_t c[4] = {(uint32_t)draw, (uint32_t)(draw >> 32), (uint32_t)stream, (uint32_t)(stream >> 32)};
  or_philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  uint64_t bits = ((uint64_t)c[1] << 32) | c[0];
  return (double)(bits >> 11) * (1.0 / 9007199254740992.0);
}

/* ------------------------------------------------------------------------ */
/* Gate matrices (src/gates.rs:23-274)                                      */
/* ------------------------------------------------------------------------ */
#define S2 0.70710678118654752440
#define PI_ 3.14159265358979323846

static cx ei(double x) { return C(cos(x), sin(x)); }

static void m2_set(cx m[4], cx a, cx b, cx c, cx d) {
  m[0] = a;
  m[1] = b;
  m[2] = c;
  m[3] = d;
}

static void u3(double th, double ph, double la, cx m[4]) {
  double ct = cos(th / 2.0), st = sin(th / 2.0);
  cx a = cmul(ei(la), C(st, 0.0));
  m2_set(m, C(ct, 0.0), C(-a.re, -a.im), cmul(ei(ph), C(st, 0.0)), cmul(ei(ph + la), C(ct, 0.0)));
}

/* gate_matrix_1q (src/types.rs:303-344) */
static int gate1q(int kind, double th, double ph, double la, cx m[4]) {
  switch (kind) {
    case OR_ID:
    case OR_U0: m2_set(m, C(1, 0), C(0, 0), C(0, 0), C(1, 0)); return 1;
    case OR_X: m2_set(m, C(0, 0), C(1, 0), C(1, 0), C(0, 0)); return 1;
    case OR_Y: m2_set(m, C(0, 0), C(0, -1), C(0, 1), C(0, 0)); return 1;
    case OR_Z: m2_set(m, C(1, 0), C(0, 0), C(0, 0), C(-1, 0)); return 1;
    case OR_H: m2_set(m, C(S2, 0), C(S2, 0), C(S2, 0), C(-S2, 0)); return 1;
    case OR_S: m2_set(m, C(1, 0), C(0, 0), C(0, 0), C(0, 1)); return 1;
    case OR_SDG: m2_set(m, C(1, 0), C(0, 0), C(0, 0), C(0, -1)); return 1;
    case OR_T: m2_set(m, C(1, 0), C(0, 0), C(0, 0), C(S2, S2)); return 1;
    case OR_TDG: m2_set(m, C(1, 0), C(0, 0), C(0, 0), C(S2, -S2)); return 1;
    case OR_SX: m2_set(m, C(0.5, 0.5), C(0.5, -0.5), C(0.5, -0.5), C(0.5, 0.5)); return 1;
    case OR_SXDG: m2_set(m, C(0.5, -0.5), C(0.5, 0.5), C(0.5, 0.5), C(0.5, -0.5)); return 1;
    case OR_U3:
    case OR_U: u3(th, ph, la, m); return 1;
    case OR_U2: u3(PI_ / 2.0, ph, la, m); return 1;
    case OR_U1:
    case OR_P: m2_set(m, C(1, 0), C(0, 0), C(0, 0), ei(la)); return 1;
    case OR_RX: {
      double c = cos(th / 2.0), s = sin(th / 2.0);
      m2_set(m, C(c, 0), C(0, -s), C(0, -s), C(c, 0));
      return 1;
    }
    case OR_RY: {
      double c = cos(th / 2.0), s = sin(th / 2.0);
      m2_set(m, C(c, 0), C(-s, 0), C(s, 0), C(c, 0));
      return 1;
    }
    case OR_RZ: m2_set(m, ei(-ph / 2.0), C(0, 0), C(0, 0), ei(ph / 2.0)); return 1;
    default: return 0;
  }
}

int or_gate_1q(int kind, double th, double ph, double la, double out[8]) {
  cx m[4];
  if (!gate1q(kind, th, ph, la, m)) return 0;
  for (int i = 0; i < 4; ++i) {
    out[2 * i] = m[i].re;
    out[2 * i + 1] = m[i].im;
  }
  return 1;
}

/* ------------------------------------------------------------------------ */
/* Compiled operations (execution.rs:21-56)                                 */
/* ------------------------------------------------------------------------ */
enum { K_ONE, K_X, K_SWAP, K_DIAG, K_DENSE, K_MEAS, K_RESET, K_NOOP };

typedef struct {
  int kind;
  uint32_t target, a, b, qubit;
  uint64_t cbit;
  int nctl;
  uint32_t cq[5];
  uint8_t cv[5];
  int k;
  uint32_t qubits[5];
  cx* mat; /* One: 4, Diagonal: 2^k, Dense: 4^k */
  int ncond;
  uint64_t cbase[OR_MAX_COND], csize[OR_MAX_COND], cvalue[OR_MAX_COND];
} op_t;

typedef struct {
  op_t* v;
  int64_t n, cap;
} oplist;

static void ol_push(oplist* l, op_t o) {
  if (l->n == l->cap) {
    l->cap = l->cap ? 2 * l->cap : 64;
    l->v = (op_t*)realloc(l->v, (size_t)l->cap * sizeof(op_t));
  }
  l->v[l->n++] = o;
}
static void ol_free(oplist* l) {
  for (int64_t i = 0; i < l->n; ++i) free(l->v[i].mat);
  free(l->v);
}

static int is_unitary(const op_t* o) { return o->ncond == 0 && o->kind != K_MEAS && o->kind != K_RESET; }

/* Operation::one (execution.rs:66-87) */
static op_t op_one(uint32_t target, const cx m[4], const uint32_t* cq, const uint8_t* cv, int nctl) {
  op_t o;
  memset(&o, 0, sizeof o);
  cx x[4] = {C(0, 0), C(1, 0), C(1, 0), C(0, 0)};
  int is_x = 1;
  for (int i = 0; i < 4; ++i) is_x &= ceq(m[i], x[i]);
  o.nctl = nctl;
  for (int i = 0; i < nctl; ++i) {
    o.cq[i] = cq[i];
    o.cv[i] = cv[i];
  }
  if (is_x) {
    o.kind = K_X;
    o.target = target;
  } else if (ceq(m[1], C(0, 0)) && ceq(m[2], C(0, 0))) {
    o.kind = K_DIAG;
    o.k = nctl + 1;
    for (int i = 0; i < nctl; ++i) o.qubits[i] = cq[i];
    o.qubits[nctl] = target;
    int dim = 1 << o.k;
    o.mat = (cx*)malloc(sizeof(cx) * dim);
    for (int j = 0; j < dim; ++j) o.mat[j] = C(1, 0);
    int index = 0;
    for (int i = 0; i < nctl; ++i) index = (index << 1) | (cv[i] ? 1 : 0);
    index <<= 1;
    o.mat[index] = m[0];
    o.mat[index + 1] = m[3];
    o.nctl = 0;
  } else {
    o.kind = K_ONE;
    o.target = target;
    o.mat = (cx*)malloc(sizeof(cx) * 4);
    memcpy(o.mat, m, sizeof(cx) * 4);
  }
  return o;
}

/* Operation::dense (execution.rs:89-106) */
static op_t op_dense(const uint32_t* qubits, int k, const cx* m) {
  op_t o;
  memset(&o, 0, sizeof o);
  int dim = 1 << k, diag = 1;
  for (int i = 0; i < dim; ++i)
    for (int j = 0; j < dim; ++j)
      if (i != j && !ceq(m[i * dim + j], C(0, 0))) diag = 0;
  o.k = k;
  for (int i = 0; i < k; ++i) o.qubits[i] = qubits[i];
  if (diag) {
    o.kind = K_DIAG;
    o.mat = (cx*)malloc(sizeof(cx) * dim);
    for (int i = 0; i < dim; ++i) o.mat[i] = m[i * dim + i];
  } else {
    o.kind = K_DENSE;
    o.mat = (cx*)malloc(sizeof(cx) * dim * dim);
    memcpy(o.mat, m, sizeof(cx) * dim * dim);
  }
  return o;
}

static void ident(cx* m, int dim) {
  for (int i = 0; i < dim * dim; ++i) m[i] = C(0, 0);
  for (int i = 0; i < dim; ++i) m[i * dim + i] = C(1, 0);
}

/* Operation::compile for non-1q instructions (execution.rs:108-274).
 * Returns 0 and sets *o, or 1 for a no-op, or -1 for an unsupported gate. */
static int compile_inst(const or_inst* in, op_t* o) {
  cx m[4];
  uint32_t q[OR_MAX_OPERANDS];
  for (int i = 0; i < in->nq && i < OR_MAX_OPERANDS; ++i) q[i] = (uint32_t)in->q[i];
  uint8_t on[5] = {1, 1, 1, 1, 1};
  if (gate1q(in->kind, in->theta, in->phi, in->lam, m)) {
    *o = op_one(q[0], m, NULL, NULL, 0);
    return 0;
  }
  switch (in->kind) {
    case OR_CX: gate1q(OR_X, 0, 0, 0, m); *o = op_one(q[1], m, q, on, 1); return 0;
    case OR_CY: gate1q(OR_Y, 0, 0, 0, m); *o = op_one(q[1], m, q, on, 1); return 0;
    case OR_CZ: gate1q(OR_Z, 0, 0, 0, m); *o = op_one(q[1], m, q, on, 1); return 0;
    case OR_CH: gate1q(OR_H, 0, 0, 0, m); *o = op_one(q[1], m, q, on, 1); return 0;
    case OR_CSX: gate1q(OR_SX, 0, 0, 0, m); *o = op_one(q[1], m, q, on, 1); return 0;
    case OR_CRX: gate1q(OR_RX, in->theta, 0, 0, m); *o = op_one(q[1], m, q, on, 1); return 0;
    case OR_CRY: gate1q(OR_RY, in->theta, 0, 0, m); *o = op_one(q[1], m, q, on, 1); return 0;
    case OR_CRZ: gate1q(OR_RZ, 0, in->lam, 0, m); *o = op_one(q[1], m, q, on, 1); return 0;
    case OR_CU1:
    case OR_CP: gate1q(OR_P, 0, 0, in->lam, m); *o = op_one(q[1], m, q, on, 1); return 0;
    case OR_CU3: u3(in->theta, in->phi, in->lam, m); *o = op_one(q[1], m, q, on, 1); return 0;
    case OR_CU: {
      u3(in->theta, in->phi, in->lam, m);
      cx ph = C(cos(in->gamma), sin(in->gamma));
      for (int i = 0; i < 4; ++i) m[i] = cmul(m[i], ph);
      *o = op_one(q[1], m, q, on, 1);
      return 0;
    }
    case OR_SWAP:
      memset(o, 0, sizeof *o);
      o->kind = K_SWAP;
      o->a = q[0];
      o->b = q[1];
      return 0;
    case OR_CSWAP:
      memset(o, 0, sizeof *o);
      o->kind = K_SWAP;
      o->a = q[1];
      o->b = q[2];
      o->nctl = 1;
      o->cq[0] = q[0];
      o->cv[0] = 1;
      return 0;
    case OR_CCX:
    case OR_C3X:
    case OR_C4X: {
      memset(o, 0, sizeof *o);
      o->kind = K_X;
      o->nctl = in->nq - 1;
      for (int i = 0; i < o->nctl; ++i) {
        o->cq[i] = q[i];
        o->cv[i] = 1;
      }
      o->target = q[in->nq - 1];
      return 0;
    }
    case OR_C3SQRTX: gate1q(OR_SX, 0, 0, 0, m); *o = op_one(q[3], m, q, on, 3); return 0;
    case OR_RXX: {
      double c = cos(in->theta / 2.0), s = sin(in->theta / 2.0);
      cx z = C(0, 0), a = C(c, 0), b = C(0, -s);
      cx mm[16] = {a, z, z, b, z, a, b, z, z, b, a, z, b, z, z, a};
      *o = op_dense(q, 2, mm);
      return 0;
    }
    case OR_RZZ: {
      cx ep = C(cos(in->theta / 2.0), sin(in->theta / 2.0));
      cx em = C(cos(in->theta / 2.0), -sin(in->theta / 2.0));
      cx z = C(0, 0);
      cx mm[16] = {em, z, z, z, z, ep, z, z, z, z, ep, z, z, z, z, em};
      *o = op_dense(q, 2, mm);
      return 0;
    }
    case OR_RCCX: {
      cx mm[64];
      ident(mm, 8);
      mm[5 * 8 + 5] = C(-1, 0);
      mm[6 * 8 + 6] = C(0, 0);
      mm[7 * 8 + 7] = C(0, 0);
      mm[6 * 8 + 7] = C(0, -1);
      mm[7 * 8 + 6] = C(0, 1);
      *o = op_dense(q, 3, mm);
      return 0;
    }
    case OR_RC3X: {
      cx mm[256];
      ident(mm, 16);
      mm[12 * 16 + 12] = C(0, 1);
      mm[13 * 16 + 13] = C(0, -1);
      mm[14 * 16 + 14] = C(0, 0);
      mm[15 * 16 + 15] = C(0, 0);
      mm[14 * 16 + 15] = C(1, 0);
      mm[15 * 16 + 14] = C(-1, 0);
      *o = op_dense(q, 4, mm);
      return 0;
    }
    case OR_MEASURE:
      memset(o, 0, sizeof *o);
      o->kind = K_MEAS;
      o->qubit = q[0];
      o->cbit = in->cbit;
      return 0;
    case OR_RESET:
      memset(o, 0, sizeof *o);
      o->kind = K_RESET;
      o->qubit = q[0];
      return 0;
    case OR_BARRIER: return 1;
    case OR_GATE: {
      const char* nm = in->name;
      const cx z = C(0, 0), s = C(S2, 0), ns = C(-S2, 0);
      if (!strcmp(nm, "remote_link_phi_plus") || !strcmp(nm, "remote_epr") || !strcmp(nm, "epr")) {
        cx mm[16] = {s, z, s, z, z, s, z, s, z, s, z, ns, s, z, ns, z};
        *o = op_dense(q, 2, mm);
        return 0;
      }
      if (!strcmp(nm, "remote_link_psi_minus") || !strcmp(nm, "remote_link_psi_plus")) {
        cx mm[16] = {z, s, z, s, s, z, s, z, ns, z, s, z, z, ns, z, s};
        if (!strcmp(nm, "remote_link_psi_plus"))
          for (int j = 8; j < 16; ++j) mm[j] = C(-mm[j].re, -mm[j].im);
        *o = op_dense(q, 2, mm);
        return 0;
      }
      if (!strcmp(nm, "nonlocal_cz") || !strcmp(nm, "remote_cz")) {
        gate1q(OR_Z, 0, 0, 0, m);
        *o = op_one(q[1], m, q, on, 1);
        return 0;
      }
      if (!strcmp(nm, "remote_cx")) {
        gate1q(OR_X, 0, 0, 0, m);
        *o = op_one(q[1], m, q, on, 1);
        return 0;
      }
      if (!strcmp(nm, "remote_rzz")) {
        or_inst t = *in;
        t.kind = OR_RZZ;
        t.theta = in->params[0];
        return compile_inst(&t, o);
      }
      if (!strcmp(nm, "remote_cu1")) {
        gate1q(OR_P, 0, 0, in->params[0], m);
        *o = op_one(q[1], m, q, on, 1);
        return 0;
      }
      if (!strcmp(nm, "remote_barrier")) return 1;
      return -1;
    }
    default: return -1;
  }
}

/* compile (execution.rs:368-416) */
static int compile_all(const or_inst* insts, int64_t count, oplist* out) {
  cx pend[64][4];
  int has[64];
  memset(has, 0, sizeof has);
  const cx I[4] = {{1, 0}, {0, 0}, {0, 0}, {1, 0}};
#define FLUSH(qq)                                       \
  do {                                                  \
    int q_ = (qq);                                      \
    if (has[q_]) {                                      \
      has[q_] = 0;                                      \
      int id_ = 1;                                      \
      for (int t_ = 0; t_ < 4; ++t_) id_ &= ceq(pend[q_][t_], I[t_]); \
      if (!id_) ol_push(out, op_one((uint32_t)q_, pend[q_], NULL, NULL, 0)); \
    }                                                   \
  } while (0)
  for (int64_t i = 0; i < count; ++i) {
    const or_inst* in = &insts[i];
    cx m[4];
    if (in->ncond == 0 && gate1q(in->kind, in->theta, in->phi, in->lam, m)) {
      int q = (int)in->q[0];
      const cx* old = has[q] ? pend[q] : I;
      cx prod[4];
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
          cx acc = C(0, 0);
          for (int k = 0; k < 2; ++k) acc = cadd(acc, cmul(m[a * 2 + k], old[k * 2 + b]));
          prod[a * 2 + b] = acc;
        }
      memcpy(pend[q], prod, sizeof prod);
      has[q] = 1;
      continue;
    }
    if (in->ncond == 0 && (in->kind == OR_BARRIER || in->kind == OR_MEASURE)) {
      for (int q = 0; q < 64; ++q) FLUSH(q);
    } else {
      for (int t = 0; t < in->nq; ++t) FLUSH((int)in->q[t]);
    }
    op_t o;
    int r = compile_inst(in, &o);
    if (r < 0) return -1;
    if (r == 1) {
      if (in->ncond == 0) continue;
      memset(&o, 0, sizeof o);
      o.kind = K_NOOP;
    }
    o.ncond = in->ncond;
    for (int c = 0; c < in->ncond; ++c) {
      o.cbase[c] = in->cbase[c];
      o.csize[c] = in->csize[c];
      o.cvalue[c] = in->cvalue[c];
    }
    ol_push(out, o);
  }
  for (int q = 0; q < 64; ++q) FLUSH(q);
#undef FLUSH
  return 0;
}

int64_t or_compile_kinds(const or_inst* insts, int64_t count, int32_t* kinds, int64_t cap) {
  oplist l = {0};
  if (compile_all(insts, count, &l) != 0) {
    ol_free(&l);
    return -1;
  }
  for (int64_t i = 0; i < l.n && i < cap; ++i) kinds[i] = l.v[i].ncond ? 7 : l.v[i].kind;
  int64_t n = l.n;
  ol_free(&l);
  return n;
}

/* One instruction compiled alone, for the P-block oracle (pblock.py):
 * execution.rs's per-instruction rules without the statevector's
 * single-qubit fusion. Returns 1 and fills the descriptor (matrix entries
 * re, im interleaved into mat: One 4, Diagonal 2^k, Dense 4^k), 0 for a
 * no-op, -1 on error. Conditions are those of the flattened instruction. */

int or_compile_one(const or_inst* in, or_opdesc* d, double* mat, int64_t mat_cap) {
  op_t o;
  memset(&o, 0, sizeof o);
  const int r = compile_inst(in, &o);
  if (r < 0) return -1;
  if (r == 1) return 0;
  memset(d, 0, sizeof *d);
  d->kind = o.kind;
  d->target = o.target;
  d->a = o.a;
  d->b = o.b;
  d->qubit = o.qubit;
  d->cbit = o.cbit;
  d->nctl = o.nctl;
  for (int i = 0; i < 5; ++i) {
    d->cq[i] = o.cq[i];
    d->cv[i] = o.cv[i];
    d->qubits[i] = o.qubits[i];
  }
  d->k = o.k;
  int nm = 0;
  if (o.kind == K_ONE) nm = 4;
  else if (o.kind == K_DIAG) nm = 1 << o.k;
  else if (o.kind == K_DENSE) nm = 1 << (2 * o.k);
  d->nmat = nm;
  if (nm > mat_cap / 2) {
    free(o.mat);
    return -1;
  }
  for (int i = 0; i < nm; ++i) {
    mat[2 * i] = o.mat[i].re;
    mat[2 * i + 1] = o.mat[i].im;
  }
  free(o.mat);
  return 1;
}

/* ------------------------------------------------------------------------ */
/* Kernels (src/engine.rs:64-453)                                           */
/* ------------------------------------------------------------------------ */
static void masks(const uint32_t* cq, const uint8_t* cv, int nctl, uint64_t* mask, uint64_t* desired) {
  *mask = 0;
  *desired = 0;
  for (int i = 0; i < nctl; ++i) {
    *mask |= 1ull << cq[i];
    if (cv[i]) *desired |= 1ull << cq[i];
  }
}

void or_apply_one(double* s, int n, uint32_t target, const double u[8], const uint32_t* cq, const uint8_t* cv,
                  int nctl) {
  uint64_t mask, desired;
  masks(cq, cv, nctl, &mask, &desired);
  const uint64_t tb = 1ull << target, len = 1ull << n;
  cx* st = (cx*)s;
  const cx u0 = C(u[0], u[1]), u1 = C(u[2], u[3]), u2 = C(u[4], u[5]), u3v = C(u[6], u[7]);
#pragma omp parallel for schedule(static) if (n >= 12)
  for (int64_t ii = 0; ii < (int64_t)len; ++ii) {
    uint64_t i = (uint64_t)ii;
    if ((i & tb) || (i & mask) != desired) continue;
    cx a = st[i], b = st[i | tb];
    st[i] = cadd(cmul(u0, a), cmul(u1, b));
    st[i | tb] = cadd(cmul(u2, a), cmul(u3v, b));
  }
}

void or_apply_x(double* s, int n, uint32_t target, const uint32_t* cq, const uint8_t* cv, int nctl) {
  uint64_t mask, desired;
  masks(cq, cv, nctl, &mask, &desired);
  const uint64_t tb = 1ull << target, len = 1ull << n;
  cx* st = (cx*)s;
#pragma omp parallel for schedule(static) if (n >= 12)
  for (int64_t ii = 0; ii < (int64_t)len; ++ii) {
    uint64_t i = (uint64_t)ii;
    if ((i & tb) || (i & mask) != desired) continue;
    cx t = st[i];
    st[i] = st[i | tb];
    st[i | tb] = t;
  }
}

void or_apply_swap(double* s, int n, uint32_t a, uint32_t b, const uint32_t* cq, const uint8_t* cv, int nctl) {
  uint64_t mask, desired;
  masks(cq, cv, nctl, &mask, &desired);
  const uint64_t ab = 1ull << a, bb = 1ull << b, len = 1ull << n;
  cx* st = (cx*)s;
#pragma omp parallel for schedule(static) if (n >= 12)
  for (int64_t ii = 0; ii < (int64_t)len; ++ii) {
    uint64_t i = (uint64_t)ii;
    if (!(i & ab) || (i & bb) || (i & mask) != desired) continue;
    uint64_t j = (i & ~ab) | bb;
    cx t = st[i];
    st[i] = st[j];
    st[j] = t;
  }
}

void or_apply_diagonal(double* s, int n, const uint32_t* qubits, int k, const double* diag) {
  const uint64_t len = 1ull << n;
  cx* st = (cx*)s;
#pragma omp parallel for schedule(static) if (n >= 12)
  for (int64_t ii = 0; ii < (int64_t)len; ++ii) {
    uint64_t i = (uint64_t)ii, j = 0;
    for (int t = 0; t < k; ++t) j = (j << 1) | ((i >> qubits[t]) & 1);
    st[i] = cmul(st[i], C(diag[2 * j], diag[2 * j + 1]));
  }
}

void or_apply_dense(double* s, int n, const uint32_t* qubits, int k, const double* mat) {
  const int dim = 1 << k;
  uint64_t offsets[32];
  uint32_t sorted[5];
  for (int i = 0; i < k; ++i) sorted[i] = qubits[i];
  for (int i = 0; i < k; ++i)
    for (int j = i + 1; j < k; ++j)
      if (sorted[j] < sorted[i]) {
        uint32_t t = sorted[i];
        sorted[i] = sorted[j];
        sorted[j] = t;
      }
  for (int j = 0; j < dim; ++j) {
    uint64_t off = 0;
    for (int b = 0; b < k; ++b) off |= (uint64_t)((j >> (k - 1 - b)) & 1) << qubits[b];
    offsets[j] = off;
  }
  const uint64_t nb = (1ull << n) >> k;
  cx* st = (cx*)s;
#pragma omp parallel for schedule(static) if (n >= 12)
  for (int64_t cc = 0; cc < (int64_t)nb; ++cc) {
    uint64_t base = (uint64_t)cc;
    for (int t = 0; t < k; ++t) { /* insert_target_zeros (engine.rs:245-251) */
      uint64_t low = (1ull << sorted[t]) - 1;
      base = (base & low) | ((base & ~low) << 1);
    }
    cx v[32], r[32];
    for (int j = 0; j < dim; ++j) v[j] = st[base | offsets[j]];
    for (int row = 0; row < dim; ++row) {
      cx acc = C(0, 0);
      for (int col = 0; col < dim; ++col)
        acc = cadd(acc, cmul(C(mat[2 * (row * dim + col)], mat[2 * (row * dim + col) + 1]), v[col]));
      r[row] = acc;
    }
    for (int j = 0; j < dim; ++j) st[base | offsets[j]] = r[j];
  }
}

/* Pairwise tree over a power-of-two count, bottom-up (same tree as the
 * recursive halves definition). */
static void ptree_inplace(double* x, uint64_t len) {
  for (uint64_t w = 1; w < len; w *= 2)
    for (uint64_t i = 0; i + w < len; i += 2 * w) x[i] = x[i] + x[i + w];
}

double or_ptree(const double* x, uint64_t len) {
  if (len == 0) return 0.0;
  double* t = (double*)malloc(sizeof(double) * len);
  memcpy(t, x, sizeof(double) * len);
  ptree_inplace(t, len);
  double r = t[0];
  free(t);
  return r;
}

static inline double prob(const cx* st, uint64_t i) { return st[i].re * st[i].re + st[i].im * st[i].im; }

/* measure_qubit (engine.rs:367-435) with the pairwise-tree branch sums. */
/* f32 != 0: the state holds complex64 values (exactly upcast) and the kept
 * amplitudes are rounded back to FP32 after scaling, as a complex64 state is
 * collapsed (scale computed in FP64, product rounded to float). */
static int measure_impl(double* s, int n, uint32_t qubit, double r, double* p0o, double* p1o, int f32) {
  const uint64_t len = 1ull << n, bit = 1ull << qubit;
  const uint64_t chunk = len < 2048 ? len : 2048, nch = len / chunk;
  double* c0 = (double*)malloc(sizeof(double) * nch);
  double* c1 = (double*)malloc(sizeof(double) * nch);
  cx* st = (cx*)s;
#pragma omp parallel for schedule(static) if (n >= 12)
  for (int64_t cc = 0; cc < (int64_t)nch; ++cc) {
    double a[2048], b[2048];
    for (uint64_t j = 0; j < chunk; ++j) {
      uint64_t i = (uint64_t)cc * chunk + j;
      double p = prob(st, i);
      a[j] = (i & bit) ? 0.0 : p;
      b[j] = (i & bit) ? p : 0.0;
    }
    ptree_inplace(a, chunk);
    ptree_inplace(b, chunk);
    c0[cc] = a[0];
    c1[cc] = b[0];
  }
  ptree_inplace(c0, nch);
  ptree_inplace(c1, nch);
  double p0 = c0[0], p1 = c1[0];
  free(c0);
  free(c1);
  if (p0o) *p0o = p0;
  if (p1o) *p1o = p1;
  double total = p0 + p1;
  if (!(isfinite(total) && total > 0.0)) return -1;
  int outcome = r < p1 / total ? 1 : 0;
  double scale = 1.0 / sqrt(outcome ? p1 : p0);
#pragma omp parallel for schedule(static) if (n >= 12)
  for (int64_t ii = 0; ii < (int64_t)len; ++ii) {
    uint64_t i = (uint64_t)ii;
    if ((int)((i & bit) != 0) == outcome) {
      st[i].re = st[i].re * scale;
      st[i].im = st[i].im * scale;
      if (f32) {
        st[i].re = (double)(float)st[i].re;
        st[i].im = (double)(float)st[i].im;
      }
    } else {
      st[i] = C(0, 0);
    }
  }
  return outcome;
}

int or_measure(double* s, int n, uint32_t qubit, double r, double* p0o, double* p1o) {
  return measure_impl(s, n, qubit, r, p0o, p1o, 0);
}

int or_measure_f32(double* s, int n, uint32_t qubit, double r, double* p0o, double* p1o) {
  return measure_impl(s, n, qubit, r, p0o, p1o, 1);
}

/* Virtual index map of a marginal: low bits walk the unmeasured qubits in
 * ascending order, then output bit i sits on qubits[k-1-i]. */
static void marg_map(int n, const uint64_t* qubits, int k, uint8_t* map) {
  int used[64] = {0}, j = 0;
  for (int i = 0; i < k; ++i) used[qubits[i]] = 1;
  for (int q = 0; q < n; ++q)
    if (!used[q]) map[j++] = (uint8_t)q;
  for (int i = 0; i < k; ++i) map[j++] = (uint8_t)qubits[k - 1 - i];
}
static inline uint64_t phys(uint64_t v, const uint8_t* map, int n) {
  uint64_t p = 0;
  for (int j = 0; j < n; ++j) p |= ((v >> j) & 1ull) << map[j];
  return p;
}

/* Pairwise sums of consecutive virtual segments of 2^seg elements. */
static void seg_sums(const double* s, int n, const uint8_t* map, int seg, double* out) {
  const cx* st = (const cx*)s;
  const uint64_t nseg = 1ull << (n - seg), w = 1ull << seg;
#pragma omp parallel for schedule(static) if (n >= 12)
  for (int64_t o = 0; o < (int64_t)nseg; ++o) {
    double* t = (double*)malloc(sizeof(double) * w);
    for (uint64_t r = 0; r < w; ++r) t[r] = prob(st, phys(((uint64_t)o << seg) | r, map, n));
    ptree_inplace(t, w);
    out[o] = t[0];
    free(t);
  }
}

void or_marginal(const double* s, int n, const uint64_t* qubits, int k, double* out) {
  uint8_t map[64];
  marg_map(n, qubits, k, map);
  seg_sums(s, n, map, n - k, out);
}

/* Canonical inclusive prefix (the engine's block scan, kernels.hip
 * k_scan_block / scan_into): blocks of 2048 bins; in a block, thread t sums
 * its 8 bins in order (p), the 256 thread totals are scanned in rows of 16
 * lanes by Hillis-Steele steps d = 1, 2, 4, 8 (lane i adds lane i-d of its
 * row), rows are offset by the sequential prefix of the row totals, waves by
 * the sequential prefix of the wave totals, and L = (wave offset + previous
 * lane's row-scanned value) + p. Block totals (last outputs) are scanned the
 * same way and added to every later block. Returns total = C[count-1]. */
#define SCAN_BLOCK 2048
static void block_scan(const double* y, uint64_t ystride, uint32_t m, double* out) {
  double p[256][8], s[256], ex[256];
  for (int t = 0; t < 256; ++t) {
    double acc = 0.0;
    for (int e = 0; e < 8; ++e) {
      uint32_t i = 8 * t + e;
      double v = i < m ? y[(uint64_t)i * ystride] : 0.0;
      acc = e == 0 ? v : acc + v;
      p[t][e] = acc;
    }
    s[t] = acc;
  }
  for (int w = 0; w < 4; ++w) {
    double* r = s + 64 * w;
    for (int d = 1; d <= 8; d *= 2) {
      double prev[64];
      memcpy(prev, r, sizeof prev);
      for (int i = 0; i < 64; ++i)
        if (i % 16 >= d) r[i] = prev[i] + prev[i - d];
    }
    double ro1 = r[15], ro2 = ro1 + r[31], ro3 = ro2 + r[47];
    double ro[4] = {0.0, ro1, ro2, ro3};
    for (int i = 0; i < 64; ++i) r[i] = ro[i / 16] + r[i];
    for (int i = 0; i < 64; ++i) ex[64 * w + i] = i == 0 ? 0.0 : r[i - 1];
  }
  double wo1 = s[63], wo2 = wo1 + s[127], wo3 = wo2 + s[191];
  double wo[4] = {0.0, wo1, wo2, wo3};
  for (int t = 0; t < 256; ++t) {
    double to = wo[t / 64] + ex[t];
    for (int e = 0; e < 8; ++e) {
      uint32_t i = 8 * t + e;
      if (i < m) out[i] = to + p[t][e];
    }
  }
}

static void scan_into(const double* x, uint64_t xstride, uint64_t count, double* P) {
  const uint64_t nb = (count + SCAN_BLOCK - 1) / SCAN_BLOCK;
  double* bt = (double*)malloc(sizeof(double) * nb);
  double* tmp = (double*)malloc(sizeof(double) * SCAN_BLOCK);
  for (uint64_t g = 0; g < nb; ++g) {
    const uint64_t lo = g * SCAN_BLOCK;
    const uint32_t m = (uint32_t)(count - lo < SCAN_BLOCK ? count - lo : SCAN_BLOCK);
    block_scan(x + lo * xstride, xstride, m, tmp);
    memcpy(P + lo, tmp, sizeof(double) * m);
    bt[g] = tmp[m - 1];
  }
  free(tmp);
  if (nb > 1) {
    double* q = (double*)malloc(sizeof(double) * nb);
    scan_into(bt, 1, nb, q);
    for (uint64_t j = SCAN_BLOCK; j < count; ++j) P[j] = q[j / SCAN_BLOCK - 1] + P[j];
    free(q);
  }
  free(bt);
}

static double cdf_build(const double* bins, uint64_t count, double* Cc) {
  scan_into(bins, 1, count, Cc);
  return Cc[count - 1];
}

/* The canonical prefix of x[0..count) (tests pin it). */
double or_cdf(const double* x, uint64_t count, double* out) { return cdf_build(x, count, out); }

static uint64_t partition(const double* Cc, uint64_t count, double total, double r) {
  uint64_t lo = 0, hi = count;
  while (lo < hi) {
    uint64_t mid = lo + (hi - lo) / 2;
    double u = mid + 1 == count ? 1.0 : Cc[mid] / total;
    if (u <= r) lo = mid + 1;
    else hi = mid;
  }
  return lo < count ? lo : count - 1;
}

#define MAT_MAX 24

/* Sampling over a marginal: shot s draws Philox(seed, stream s, draw 0). */
int or_sample_marginal(const double* s, int n, const uint64_t* qubits, int k, uint64_t shots, uint64_t seed,
                       uint64_t* out) {
  uint8_t map[64];
  marg_map(n, qubits, k, map);
  if (k <= MAT_MAX) {
    const uint64_t count = 1ull << k;
    double* bins = (double*)malloc(sizeof(double) * count);
    double* Cc = (double*)malloc(sizeof(double) * count);
    seg_sums(s, n, map, n - k, bins);
    double total = cdf_build(bins, count, Cc);
    if (!(isfinite(total) && total > 0.0)) {
      free(bins);
      free(Cc);
      return -1;
    }
    for (uint64_t sh = 0; sh < shots; ++sh) out[sh] = partition(Cc, count, total, or_uniform53(seed, sh, 0));
    free(bins);
    free(Cc);
    return 0;
  }
  const int binseg = n - k;
  const uint64_t nblocks = 1ull << (k - 11);
  double* blk = (double*)malloc(sizeof(double) * nblocks);
  double* Cc = (double*)malloc(sizeof(double) * nblocks);
  seg_sums(s, n, map, binseg + 11, blk);
  double total = cdf_build(blk, nblocks, Cc);
  if (!(isfinite(total) && total > 0.0)) {
    free(blk);
    free(Cc);
    return -1;
  }
  const cx* st = (const cx*)s;
  double* t = (double*)malloc(sizeof(double) * (1ull << binseg));
  for (uint64_t sh = 0; sh < shots; ++sh) {
    double r = or_uniform53(seed, sh, 0);
    uint64_t kb = partition(Cc, nblocks, total, r);
    double acc = kb == 0 ? 0.0 : Cc[kb - 1];
    uint64_t found = ~0ull, last_pos = kb * 2048;
    for (uint64_t j = 0; j < 2048; ++j) {
      uint64_t bin = kb * 2048 + j;
      for (uint64_t e = 0; e < (1ull << binseg); ++e) t[e] = prob(st, phys((bin << binseg) | e, map, n));
      ptree_inplace(t, 1ull << binseg);
      double p = t[0];
      if (p > 0.0) last_pos = bin;
      acc = acc + p;
      if (acc / total > r) {
        found = bin;
        break;
      }
    }
    out[sh] = found != ~0ull ? found : last_pos;
  }
  free(t);
  free(blk);
  free(Cc);
  return 0;
}

int or_batch_sample(const double* s, int n, uint64_t shots, uint64_t seed, int block, uint64_t* out) {
  const cx* st = (const cx*)s;
  const uint64_t len = 1ull << n;
  const uint64_t nch = len < (uint64_t)block ? len : (uint64_t)block, cl = len / nch;
  double* part = (double*)malloc(sizeof(double) * nch);
  for (uint64_t t = 0; t < nch; ++t) {
    double acc = 0.0;
    for (uint64_t i = t * cl; i < (t + 1) * cl; ++i) acc = acc + prob(st, i);
    part[t] = acc;
  }
  double run = 0.0;
  for (uint64_t t = 0; t < nch; ++t) {
    run = run + part[t];
    part[t] = run;
  }
  const double total = part[nch - 1];
  for (uint64_t sh = 0; sh < shots; ++sh) {
    const double r = or_uniform53(seed, sh, 0);
    uint64_t lo = 0, hi = nch;
    while (lo < hi) {
      uint64_t mid = (lo + hi) / 2;
      double u = mid + 1 == nch ? 1.0 : part[mid] / total;
      if (u <= r) lo = mid + 1;
      else hi = mid;
    }
    uint64_t t = lo < nch ? lo : nch - 1;
    double acc = t ? part[t - 1] : 0.0;
    uint64_t found = ~0ull, lastpos = t * cl;
    for (uint64_t i = t * cl; i < (t + 1) * cl; ++i) {
      double p = prob(st, i);
      if (p > 0.0) lastpos = i;
      acc = acc + p;
      if (acc / total > r) {
        found = i;
        break;
      }
    }
    out[sh] = found != ~0ull ? found : lastpos;
  }
  free(part);
  return 0;
}

double or_fidelity(const double* a, const double* b, int n) {
  const uint64_t len = 1ull << n;
  double* re = (double*)malloc(sizeof(double) * len);
  double* im = (double*)malloc(sizeof(double) * len);
  const cx* x = (const cx*)a;
  const cx* y = (const cx*)b;
  for (uint64_t i = 0; i < len; ++i) {
    cx c = cmul(C(x[i].re, -x[i].im), y[i]);
    re[i] = c.re;
    im[i] = c.im;
  }
  ptree_inplace(re, len);
  ptree_inplace(im, len);
  double r = re[0] * re[0] + im[0] * im[0];
  free(re);
  free(im);
  return r;
}

/* ------------------------------------------------------------------------ */
/* Execution (execution.rs:276-366)                                         */
/* ------------------------------------------------------------------------ */
typedef struct {
  uint64_t seed, stream, draw;
} rng_t;

static int exec_op(const op_t* o, double* s, int n, uint8_t* cbits, rng_t* rng) {
  for (int c = 0; c < o->ncond; ++c) {
    uint64_t actual = 0;
    for (uint64_t bit = 0; bit < o->csize[c]; ++bit) actual |= (uint64_t)cbits[o->cbase[c] + bit] << bit;
    if (actual != o->cvalue[c]) return 0;
  }
  switch (o->kind) {
    case K_ONE: {
      double u[8];
      for (int i = 0; i < 4; ++i) {
        u[2 * i] = o->mat[i].re;
        u[2 * i + 1] = o->mat[i].im;
      }
      or_apply_one(s, n, o->target, u, o->cq, o->cv, o->nctl);
      return 0;
    }
    case K_X: or_apply_x(s, n, o->target, o->cq, o->cv, o->nctl); return 0;
    case K_SWAP: or_apply_swap(s, n, o->a, o->b, o->cq, o->cv, o->nctl); return 0;
    case K_DIAG: or_apply_diagonal(s, n, o->qubits, o->k, (const double*)o->mat); return 0;
    case K_DENSE: or_apply_dense(s, n, o->qubits, o->k, (const double*)o->mat); return 0;
    case K_MEAS:
    case K_RESET: {
      double r = or_uniform53(rng->seed, rng->stream, rng->draw);
      rng->draw++;
      int out = or_measure(s, n, o->qubit, r, NULL, NULL);
      if (out < 0) return 3;
      if (o->kind == K_MEAS) cbits[o->cbit] = (uint8_t)out;
      else if (out == 1) or_apply_x(s, n, o->qubit, NULL, NULL, 0);
      return 0;
    }
    default: return 0;
  }
}

static void zero_state(double* s, int n) {
  memset(s, 0, sizeof(double) * 2 * (1ull << n));
  s[0] = 1.0;
}

int or_simulate(const or_inst* insts, int64_t count, int n, uint64_t nc, uint64_t seed, double* state,
                uint8_t* cbits) {
  oplist l = {0};
  if (compile_all(insts, count, &l) != 0) {
    ol_free(&l);
    return 1;
  }
  zero_state(state, n);
  memset(cbits, 0, nc ? nc : 1);
  rng_t rng = {seed, 0, 0};
  for (int64_t i = 0; i < l.n; ++i) {
    int e = exec_op(&l.v[i], state, n, cbits, &rng);
    if (e) {
      ol_free(&l);
      return e;
    }
  }
  ol_free(&l);
  return 0;
}

/* run_shots (shots.rs:58-260) */
int or_run_shots(const or_inst* insts, int64_t count, int n, uint64_t nc, uint64_t shots, uint64_t seed,
                 uint64_t budget, int has_mps, uint64_t mps, int sample_terminal, uint64_t max_batch, int prec,
                 uint8_t* bits_out, or_shot_info* info) {
  memset(info, 0, sizeof *info);
  oplist l = {0};
  if (compile_all(insts, count, &l) != 0) {
    ol_free(&l);
    return 1;
  }
  const uint64_t ncs = nc ? nc : 1;
  if (shots == 0) {
    ol_free(&l);
    return 0;
  }
  const uint64_t len = 1ull << n, sbytes = len * (uint64_t)prec, worker = sbytes + nc;
  if (worker > budget) {
    ol_free(&l);
    return 2;
  }
  int64_t pe = 0;
  while (pe < l.n && is_unitary(&l.v[pe])) ++pe;
  int terminal = 1;
  for (int64_t i = pe; i < l.n; ++i)
    if (!(l.v[i].kind == K_MEAS && l.v[i].ncond == 0)) terminal = 0;
  double* s = NULL;
  if (sample_terminal && terminal) {
    /* cbit -> qubit, last write wins */
    uint64_t* cq = (uint64_t*)malloc(sizeof(uint64_t) * ncs);
    uint8_t* mapped = (uint8_t*)calloc(ncs, 1);
    int any = 0;
    for (int64_t i = pe; i < l.n; ++i) {
      cq[l.v[i].cbit] = l.v[i].qubit;
      mapped[l.v[i].cbit] = 1;
      any = 1;
    }
    if (!any) {
      info->strategy = 1;
      memset(bits_out, 0, shots * ncs);
      free(cq);
      free(mapped);
      ol_free(&l);
      return 0;
    }
    uint64_t qs[64];
    int m = 0;
    for (int q = n - 1; q >= 0; --q) {
      int used = 0;
      for (uint64_t c = 0; c < nc; ++c)
        if (mapped[c] && cq[c] == (uint64_t)q) used = 1;
      if (used) qs[m++] = (uint64_t)q;
    }
    uint64_t cdfb = m <= MAT_MAX ? (1ull << m) * 8 : (1ull << (m - 11)) * 8;
    if (worker + cdfb <= budget) {
      s = (double*)malloc(sizeof(double) * 2 * len);
      zero_state(s, n);
      uint8_t dummy = 0;
      rng_t rng = {seed, 0, 0};
      for (int64_t i = 0; i < pe; ++i) exec_op(&l.v[i], s, n, &dummy, &rng);
      uint64_t* idx = (uint64_t*)malloc(sizeof(uint64_t) * shots);
      if (or_sample_marginal(s, n, qs, m, shots, seed, idx) != 0) {
        free(idx);
        free(s);
        free(cq);
        free(mapped);
        ol_free(&l);
        return 3;
      }
      for (uint64_t sh = 0; sh < shots; ++sh) {
        uint8_t* row = bits_out + sh * ncs;
        memset(row, 0, ncs);
        for (uint64_t c = 0; c < nc; ++c) {
          if (!mapped[c]) continue;
          int pos = 0;
          while (qs[pos] != cq[c]) ++pos;
          row[c] = (uint8_t)((idx[sh] >> (m - 1 - pos)) & 1);
        }
      }
      info->strategy = 1;
      info->parallel_shots = 1;
      info->working_bytes = worker + cdfb;
      info->prefix_ops = (uint64_t)pe;
      free(idx);
      free(s);
      free(cq);
      free(mapped);
      ol_free(&l);
      return 0;
    }
    free(cq);
    free(mapped);
  }
  const int cache = shots > 1 && pe > 0 && sbytes <= budget - worker;
  const uint64_t pbytes = cache ? sbytes : 0;
  uint64_t workers = has_mps ? mps : max_batch;
  if (workers > shots) workers = shots;
  if (workers > max_batch) workers = max_batch;
  if (workers > (budget - pbytes) / worker) workers = (budget - pbytes) / worker;
  if (workers < 1) workers = 1;
  double* prefix = NULL;
  uint8_t* cb = (uint8_t*)malloc(ncs);
  if (cache) {
    prefix = (double*)malloc(sizeof(double) * 2 * len);
    zero_state(prefix, n);
    rng_t rng = {seed, 0, 0};
    memset(cb, 0, ncs);
    for (int64_t i = 0; i < pe; ++i) exec_op(&l.v[i], prefix, n, cb, &rng);
  }
  s = (double*)malloc(sizeof(double) * 2 * len);
  int rc = 0;
  for (uint64_t sh = 0; sh < shots && rc == 0; ++sh) {
    if (cache) memcpy(s, prefix, sizeof(double) * 2 * len);
    else zero_state(s, n);
    memset(cb, 0, ncs);
    rng_t rng = {seed, sh, 0};
    for (int64_t i = cache ? pe : 0; i < l.n && rc == 0; ++i) rc = exec_op(&l.v[i], s, n, cb, &rng);
    memcpy(bits_out + sh * ncs, cb, ncs);
  }
  info->strategy = 2;
  info->parallel_shots = workers;
  info->working_bytes = pbytes + workers * worker;
  info->prefix_ops = cache ? (uint64_t)pe : 0;
  free(prefix);
  free(s);
  free(cb);
  ol_free(&l);
  return rc;
}
