/* CPU oracle for the ket dense state-vector path (test infrastructure only).
 *
 * A plain-C restatement of the reference simulator, written from its
 * behaviour. Build with -ffp-contract=off (see oracle/Makefile): the
 * reference is compiled as ISO C++20 on x86-64, where products and sums are
 * rounded one operation at a time, and this file must round the same way.
 * Citations are file:line into the reference source tree.
 */
#define _GNU_SOURCE
#include "ket_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
  double re, im;
} cx;

/* GCC's expansion of a complex product (what std::complex<double>::operator*
 * compiles to for finite operands). */
static inline cx cmul(cx a, cx b) {
  cx r;
  r.re = a.re * b.re - a.im * b.im;
  r.im = a.re * b.im + a.im * b.re;
  return r;
}
static inline cx cadd(cx a, cx b) {
  cx r = {a.re + b.re, a.im + b.im};
  return r;
}
static inline cx mk(double re, double im) {
  cx r = {re, im};
  return r;
}

/* The reference is built with g++, which turns each cos(x)/sin(x) pair of
 * one argument into a single glibc sincos() call; sincos can differ from
 * separate sin() and cos() in the last bit, so the oracle calls it
 * explicitly (returns (cos x, sin x)). */
static inline cx cos_sin(double x) {
  double s, c;
  sincos(x, &s, &c);
  return mk(c, s);
}

/* ---- parallel_for: src/parallelism.cpp:25 (threshold), 73-79 (contiguous
 * chunks of ceil(count/T), caller = chunk 0), 142-150 (default 1 thread). */
static int g_threads = 1;
#define PAR_THRESHOLD ((size_t)1 << 14)

void oracle_set_threads(int threads) {
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_num_procs();
#else
  if (threads <= 0) threads = 1;
#endif
  g_threads = threads;
}
int oracle_num_threads(void) { return g_threads; }

typedef void (*body_fn)(void* ctx, size_t lo, size_t hi);

static void parallel_for(size_t count, body_fn body, void* ctx) {
  if (count == 0) return;
  if (g_threads <= 1 || count < PAR_THRESHOLD) {
    body(ctx, 0, count);
    return;
  }
#ifdef _OPENMP
  const size_t T = (size_t)g_threads;
  const size_t chunk = (count + T - 1) / T;
#pragma omp parallel num_threads(g_threads)
  {
    const size_t idx = (size_t)omp_get_thread_num();
    const size_t b = chunk * idx;
    if (b < count) {
      const size_t e = (b + chunk < count) ? b + chunk : count;
      body(ctx, b, e);
    }
  }
#else
  body(ctx, 0, count);
#endif
}

/* ---- kernels: src/simulator.cpp:23-254 ---- */
typedef struct {
  cx* s;
  size_t m0, m1, m2; /* masks: meaning per kernel */
  cx a, b, c, d;     /* 2x2 matrix or phase */
} kctx;

/* apply_single: src/simulator.cpp:23-36 */
static void k_single(void* vp, size_t lo, size_t hi) {
  kctx* k = (kctx*)vp;
  cx* s = k->s;
  for (size_t i = lo; i < hi; ++i) {
    if (i & k->m0) continue;
    const size_t i1 = i | k->m0;
    const cx a = s[i], b = s[i1];
    s[i] = cadd(cmul(k->a, a), cmul(k->b, b));
    s[i1] = cadd(cmul(k->c, a), cmul(k->d, b));
  }
}
static void apply_single(cx* s, size_t dim, int q, cx m00, cx m01, cx m10,
                         cx m11) {
  kctx k = {s, (size_t)1 << q, 0, 0, m00, m01, m10, m11};
  parallel_for(dim, k_single, &k);
}

/* apply_cx: src/simulator.cpp:99-109 (m0 = control, m1 = target) */
static void k_cx(void* vp, size_t lo, size_t hi) {
  kctx* k = (kctx*)vp;
  cx* s = k->s;
  for (size_t i = lo; i < hi; ++i) {
    if (!(i & k->m0)) continue;
    if (i & k->m1) continue;
    const cx t = s[i];
    s[i] = s[i | k->m1];
    s[i | k->m1] = t;
  }
}

/* apply_cy: src/simulator.cpp:112-126 */
static void k_cy(void* vp, size_t lo, size_t hi) {
  kctx* k = (kctx*)vp;
  cx* s = k->s;
  for (size_t i = lo; i < hi; ++i) {
    if (!(i & k->m0) || (i & k->m1)) continue;
    const size_t i1 = i | k->m1;
    const cx a = s[i], b = s[i1];
    s[i] = cmul(mk(0.0, -1.0), b);
    s[i1] = cmul(mk(0.0, 1.0), a);
  }
}

/* apply_ch: src/simulator.cpp:129-144; (a+b) times a real scalar is a
 * component-wise product. */
static void k_ch(void* vp, size_t lo, size_t hi) {
  kctx* k = (kctx*)vp;
  cx* s = k->s;
  const double r = k->a.re;
  for (size_t i = lo; i < hi; ++i) {
    if (!(i & k->m0) || (i & k->m1)) continue;
    const size_t i1 = i | k->m1;
    const cx a = s[i], b = s[i1];
    s[i] = mk((a.re + b.re) * r, (a.im + b.im) * r);
    s[i1] = mk((a.re - b.re) * r, (a.im - b.im) * r);
  }
}

/* apply_controlled_single: src/simulator.cpp:148-164 */
static void k_csingle(void* vp, size_t lo, size_t hi) {
  kctx* k = (kctx*)vp;
  cx* s = k->s;
  for (size_t i = lo; i < hi; ++i) {
    if (!(i & k->m0) || (i & k->m1)) continue;
    const size_t i1 = i | k->m1;
    const cx a = s[i], b = s[i1];
    s[i] = cadd(cmul(k->a, a), cmul(k->b, b));
    s[i1] = cadd(cmul(k->c, a), cmul(k->d, b));
  }
}

/* apply_swap: src/simulator.cpp:204-212 (m0 = a, m1 = b);
 * apply_cswap: 232-242 adds the control m2 (0 = none). */
static void k_swap(void* vp, size_t lo, size_t hi) {
  kctx* k = (kctx*)vp;
  cx* s = k->s;
  for (size_t i = lo; i < hi; ++i) {
    if (k->m2 && !(i & k->m2)) continue;
    if ((i & k->m0) && !(i & k->m1)) {
      const size_t j = (i ^ k->m0) | k->m1;
      const cx t = s[i];
      s[i] = s[j];
      s[j] = t;
    }
  }
}

/* apply_ccx: src/simulator.cpp:215-228 (m0, m1 controls, m2 target) */
static void k_ccx(void* vp, size_t lo, size_t hi) {
  kctx* k = (kctx*)vp;
  cx* s = k->s;
  for (size_t i = lo; i < hi; ++i) {
    if (!(i & k->m0) || !(i & k->m1)) continue;
    if (i & k->m2) continue;
    const cx t = s[i];
    s[i] = s[i | k->m2];
    s[i | k->m2] = t;
  }
}

/* apply_cphase: src/simulator.cpp:246-254 (s[k] *= phase) */
static void k_cphase(void* vp, size_t lo, size_t hi) {
  kctx* k = (kctx*)vp;
  cx* s = k->s;
  for (size_t i = lo; i < hi; ++i) {
    if ((i & k->m0) && (i & k->m1)) s[i] = cmul(s[i], k->a);
  }
}

static void run_kernel(body_fn f, cx* s, size_t dim, size_t m0, size_t m1,
                       size_t m2, cx a, cx b, cx c, cx d) {
  kctx k = {s, m0, m1, m2, a, b, c, d};
  parallel_for(dim, f, &k);
}

#define BIT(q) ((size_t)1 << (q))

/* apply_gate_on: src/simulator.cpp:259-338, with the per-gate matrices of
 * 38-97 and 168-200. */
static int apply_gate_on(cx* s, size_t dim, int type, const int* on,
                         const double* p) {
  const cx Z0 = mk(0.0, 0.0), O1 = mk(1.0, 0.0);
  switch (type) {
    case OR_H: {
      const double r = 1.0 / sqrt(2.0);
      apply_single(s, dim, on[0], mk(r, 0), mk(r, 0), mk(r, 0), mk(-r, 0));
      return 0;
    }
    case OR_X:
      apply_single(s, dim, on[0], Z0, O1, O1, Z0);
      return 0;
    case OR_Y:
      apply_single(s, dim, on[0], Z0, mk(0.0, -1.0), mk(0.0, 1.0), Z0);
      return 0;
    case OR_Z:
      apply_single(s, dim, on[0], O1, Z0, Z0, mk(-1.0, 0.0));
      return 0;
    case OR_S:
      apply_single(s, dim, on[0], O1, Z0, Z0, mk(0.0, 1.0));
      return 0;
    case OR_SDG:
      apply_single(s, dim, on[0], O1, Z0, Z0, mk(0.0, -1.0));
      return 0;
    case OR_T: {
      const double r = 1.0 / sqrt(2.0);
      apply_single(s, dim, on[0], O1, Z0, Z0, mk(r, r));
      return 0;
    }
    case OR_TDG: {
      const double r = 1.0 / sqrt(2.0);
      apply_single(s, dim, on[0], O1, Z0, Z0, mk(r, -r));
      return 0;
    }
    case OR_RX: {
      const cx h = cos_sin(p[0] / 2.0);
      const double c = h.re, sn = h.im;
      apply_single(s, dim, on[0], mk(c, 0), mk(0.0, -sn), mk(0.0, -sn),
                   mk(c, 0));
      return 0;
    }
    case OR_RY: {
      const cx h = cos_sin(p[0] / 2.0);
      const double c = h.re, sn = h.im;
      apply_single(s, dim, on[0], mk(c, 0), mk(-sn, 0), mk(sn, 0), mk(c, 0));
      return 0;
    }
    case OR_RZ: {
      const cx h = cos_sin(p[0] / 2.0);
      const double c = h.re, sn = h.im;
      apply_single(s, dim, on[0], mk(c, -sn), Z0, Z0, mk(c, sn));
      return 0;
    }
    case OR_U: {
      const cx h = cos_sin(p[0] / 2.0);
      const double c = h.re, sn = h.im;
      /* -polar(sn, lambda), polar(sn, phi), polar(c, phi + lambda) */
      const cx l = cos_sin(p[2]), f = cos_sin(p[1]), fl = cos_sin(p[1] + p[2]);
      const cx m01 = mk(-(sn * l.re), -(sn * l.im));
      const cx m10 = mk(sn * f.re, sn * f.im);
      const cx m11 = mk(c * fl.re, c * fl.im);
      apply_single(s, dim, on[0], mk(c, 0.0), m01, m10, m11);
      return 0;
    }
    case OR_CH: {
      const double r = 1.0 / sqrt(2.0);
      run_kernel(k_ch, s, dim, BIT(on[0]), BIT(on[1]), 0, mk(r, 0), Z0, Z0,
                 Z0);
      return 0;
    }
    case OR_CX:
      run_kernel(k_cx, s, dim, BIT(on[0]), BIT(on[1]), 0, Z0, Z0, Z0, Z0);
      return 0;
    case OR_CY:
      run_kernel(k_cy, s, dim, BIT(on[0]), BIT(on[1]), 0, Z0, Z0, Z0, Z0);
      return 0;
    case OR_CZ:
      run_kernel(k_cphase, s, dim, BIT(on[0]), BIT(on[1]), 0, mk(-1.0, 0.0),
                 Z0, Z0, Z0);
      return 0;
    case OR_CRX: {
      const cx h = cos_sin(p[0] / 2.0);
      const double c = h.re, sn = h.im;
      run_kernel(k_csingle, s, dim, BIT(on[0]), BIT(on[1]), 0, mk(c, 0),
                 mk(0.0, -sn), mk(0.0, -sn), mk(c, 0));
      return 0;
    }
    case OR_CRY: {
      const cx h = cos_sin(p[0] / 2.0);
      const double c = h.re, sn = h.im;
      run_kernel(k_csingle, s, dim, BIT(on[0]), BIT(on[1]), 0, mk(c, 0),
                 mk(-sn, 0), mk(sn, 0), mk(c, 0));
      return 0;
    }
    case OR_CRZ: {
      const cx h = cos_sin(p[0] / 2.0);
      const double c = h.re, sn = h.im;
      run_kernel(k_csingle, s, dim, BIT(on[0]), BIT(on[1]), 0, mk(c, -sn),
                 Z0, Z0, mk(c, sn));
      return 0;
    }
    case OR_CU: {
      const cx h = cos_sin(p[0] / 2.0);
      const double c = h.re, sn = h.im;
      const cx l = cos_sin(p[2]), f = cos_sin(p[1]), fl = cos_sin(p[1] + p[2]);
      const cx m01 = mk(-(sn * l.re), -(sn * l.im));
      const cx m10 = mk(sn * f.re, sn * f.im);
      const cx m11 = mk(c * fl.re, c * fl.im);
      run_kernel(k_csingle, s, dim, BIT(on[0]), BIT(on[1]), 0, mk(c, 0.0),
                 m01, m10, m11);
      return 0;
    }
    case OR_CP:
      run_kernel(k_cphase, s, dim, BIT(on[0]), BIT(on[1]), 0, cos_sin(p[0]),
                 Z0, Z0, Z0);
      return 0;
    case OR_SWAP:
      run_kernel(k_swap, s, dim, BIT(on[0]), BIT(on[1]), 0, Z0, Z0, Z0, Z0);
      return 0;
    case OR_CCX:
      run_kernel(k_ccx, s, dim, BIT(on[0]), BIT(on[1]), BIT(on[2]), Z0, Z0,
                 Z0, Z0);
      return 0;
    case OR_CSWAP:
      run_kernel(k_swap, s, dim, BIT(on[1]), BIT(on[2]), BIT(on[0]), Z0, Z0,
                 Z0, Z0);
      return 0;
    default:
      return -1;
  }
}

static int is_unitary(int type) { return type >= OR_H && type <= OR_CSWAP; }

static int gate_arity(int type) {
  if (type <= OR_U) return 1;
  if (type <= OR_SWAP) return 2;
  if (type <= OR_CSWAP) return 3;
  return -1;
}

void oracle_ground_state(int n, double* state) {
  const size_t dim = (size_t)1 << n;
  memset(state, 0, dim * 2 * sizeof(double));
  state[0] = 1.0;
}

int oracle_apply_gate(int n, double* state, const oracle_gate* g) {
  if (!is_unitary(g->type) || g->nq != gate_arity(g->type)) return -1;
  for (int i = 0; i < g->nq; ++i)
    if (g->q[i] < 0 || g->q[i] >= n) return -1;
  return apply_gate_on((cx*)state, (size_t)1 << n, g->type, g->q, g->p);
}

/* ---- apply_matrix: src/simulator.cpp:357-393 ---- */
typedef struct {
  cx* s;
  int m;
  const int* sorted;
  const size_t* offset;
  const cx* mat;
} mctx;

static void k_matrix(void* vp, size_t lo, size_t hi) {
  mctx* c = (mctx*)vp;
  const size_t dim = (size_t)1 << c->m;
  cx* in = (cx*)malloc(dim * sizeof(cx));
  cx* out = (cx*)malloc(dim * sizeof(cx));
  for (size_t o = lo; o < hi; ++o) {
    size_t base = o;
    for (int b = 0; b < c->m; ++b) {
      const int q = c->sorted[b];
      base = ((base >> q) << (q + 1)) | (base & (((size_t)1 << q) - 1));
    }
    for (size_t i = 0; i < dim; ++i) in[i] = c->s[base | c->offset[i]];
    for (size_t r = 0; r < dim; ++r) {
      cx acc = {0.0, 0.0};
      const cx* row = &c->mat[r * dim];
      for (size_t k = 0; k < dim; ++k) acc = cadd(acc, cmul(row[k], in[k]));
      out[r] = acc;
    }
    for (size_t i = 0; i < dim; ++i) c->s[base | c->offset[i]] = out[i];
  }
  free(in);
  free(out);
}

static int cmp_int(const void* a, const void* b) {
  return *(const int*)a - *(const int*)b;
}

void oracle_apply_matrix(int n, double* state, const int* qubits, int m,
                         const double* mat) {
  const size_t dim = (size_t)1 << m;
  int sorted[64];
  size_t* offset = (size_t*)calloc(dim, sizeof(size_t));
  memcpy(sorted, qubits, (size_t)m * sizeof(int));
  qsort(sorted, (size_t)m, sizeof(int), cmp_int);
  for (size_t i = 0; i < dim; ++i)
    for (int b = 0; b < m; ++b)
      if (i & ((size_t)1 << b)) offset[i] |= (size_t)1 << qubits[b];
  mctx c = {(cx*)state, m, sorted, offset, (const cx*)mat};
  parallel_for(((size_t)1 << n) >> m, k_matrix, &c);
  free(offset);
}

/* ---- Fuser: src/simulator.cpp:400-477 ---- */
typedef struct {
  int qubits[8];
  int nq;
  const oracle_gate* gates[4096];
  size_t ng;
} fuser;

static int local_index(const int* qubits, int m, int q) {
  for (int i = 0; i < m; ++i)
    if (qubits[i] == q) return i;
  return 0;
}

/* Fuser::block_matrix: column c = the block applied to |c> in the block's
 * local qubit order. */
static void block_matrix(const oracle_gate* const* gates, size_t ng,
                         const int* qubits, int m, cx* mat) {
  const size_t dim = (size_t)1 << m;
  cx* e = (cx*)malloc(dim * sizeof(cx));
  for (size_t col = 0; col < dim; ++col) {
    for (size_t i = 0; i < dim; ++i) e[i] = mk(0.0, 0.0);
    e[col] = mk(1.0, 0.0);
    for (size_t gi = 0; gi < ng; ++gi) {
      const oracle_gate* g = gates[gi];
      int local[3];
      for (int i = 0; i < g->nq; ++i) local[i] = local_index(qubits, m, g->q[i]);
      apply_gate_on(e, dim, g->type, local, g->p);
    }
    for (size_t row = 0; row < dim; ++row) mat[row * dim + col] = e[row];
  }
  free(e);
}

int oracle_block_matrix(const oracle_gate* g, size_t ng, const int* qubits,
                        int m, double* mat_out) {
  const oracle_gate** ptrs =
      (const oracle_gate**)malloc((ng ? ng : 1) * sizeof(*ptrs));
  for (size_t i = 0; i < ng; ++i) ptrs[i] = &g[i];
  block_matrix(ptrs, ng, qubits, m, (cx*)mat_out);
  free(ptrs);
  return 0;
}

static void fuser_flush(fuser* f, cx* s, size_t dim) {
  if (f->ng == 0) return;
  if (f->ng == 1) {
    apply_gate_on(s, dim, f->gates[0]->type, f->gates[0]->q, f->gates[0]->p);
  } else {
    const size_t bd = (size_t)1 << f->nq;
    cx* mat = (cx*)malloc(bd * bd * sizeof(cx));
    block_matrix(f->gates, f->ng, f->qubits, f->nq, mat);
    if (f->nq == 1) {
      apply_single(s, dim, f->qubits[0], mat[0], mat[1], mat[2], mat[3]);
    } else {
      int n = 0;
      while (((size_t)1 << n) < dim) ++n;
      oracle_apply_matrix(n, (double*)s, f->qubits, f->nq, (const double*)mat);
    }
    free(mat);
  }
  f->ng = 0;
  f->nq = 0;
}

static int is_subset(const int* a, int na, const int* b, int nb) {
  for (int i = 0; i < na; ++i) {
    int found = 0;
    for (int j = 0; j < nb; ++j)
      if (a[i] == b[j]) found = 1;
    if (!found) return 0;
  }
  return 1;
}

/* apply_circuit: src/simulator.cpp:482-517. Structural records flush. */
int oracle_run(int n, double* state, const oracle_gate* g, size_t ng,
               int fuse) {
  const size_t dim = (size_t)1 << n;
  cx* s = (cx*)state;
  fuser* f = (fuser*)calloc(1, sizeof(fuser));
  for (size_t i = 0; i < ng; ++i) {
    const oracle_gate* x = &g[i];
    if (!is_unitary(x->type)) {
      fuser_flush(f, s, dim);
      continue;
    }
    if (x->nq != gate_arity(x->type)) {
      free(f);
      return -1;
    }
    for (int k = 0; k < x->nq; ++k)
      if (x->q[k] < 0 || x->q[k] >= n) {
        free(f);
        return -1;
      }
    if (!fuse) {
      apply_gate_on(s, dim, x->type, x->q, x->p);
      continue;
    }
    if (f->ng > 0 && !is_subset(x->q, x->nq, f->qubits, f->nq))
      fuser_flush(f, s, dim);
    if (f->ng == 0) {
      memcpy(f->qubits, x->q, (size_t)x->nq * sizeof(int));
      f->nq = x->nq;
    }
    if (f->ng == sizeof(f->gates) / sizeof(f->gates[0])) {
      /* A block longer than the pending buffer: the block matrix is the
       * product either way, but to stay exact we never split silently. */
      free(f);
      return -1;
    }
    f->gates[f->ng++] = x;
  }
  fuser_flush(f, s, dim);
  free(f);
  return 0;
}

/* ---- measure: src/measurement.cpp:12-21 ---- */
uint64_t oracle_measure_u(const double* state, uint64_t dim, double u) {
  double cumulative = 0.0;
  for (uint64_t i = 0; i < dim; ++i) {
    const double x = state[2 * i], y = state[2 * i + 1];
    cumulative += x * x + y * y; /* std::norm in libstdc++ (no fast-math) */
    if (u < cumulative) return i;
  }
  return dim - 1;
}

/* mt19937 from [rand.eng.mers]: w=32 n=624 m=397 r=31 a=0x9908b0df u=11
 * d=0xffffffff s=7 b=0x9d2c5680 t=15 c=0xefc60000 l=18 f=1812433253. */
void oracle_mt_seed(oracle_mt19937* r, uint32_t seed) {
  r->mt[0] = seed;
  for (int i = 1; i < 624; ++i)
    r->mt[i] = 1812433253u * (r->mt[i - 1] ^ (r->mt[i - 1] >> 30)) + (uint32_t)i;
  r->idx = 624;
}

uint32_t oracle_mt_next(oracle_mt19937* r) {
  if (r->idx >= 624) {
    for (int i = 0; i < 624; ++i) {
      const uint32_t y =
          (r->mt[i] & 0x80000000u) | (r->mt[(i + 1) % 624] & 0x7fffffffu);
      uint32_t v = r->mt[(i + 397) % 624] ^ (y >> 1);
      if (y & 1u) v ^= 0x9908b0dfu;
      r->mt[i] = v;
    }
    r->idx = 0;
  }
  uint32_t y = r->mt[r->idx++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

/* generate_canonical<double, 53> over a 32-bit engine: two draws,
 * (r0 + r1 * 2^32) / 2^64, clamped below 1 (libstdc++ random.tcc). */
double oracle_canonical(oracle_mt19937* r) {
  double sum = 0.0, tmp = 1.0;
  for (int k = 0; k < 2; ++k) {
    sum += (double)oracle_mt_next(r) * tmp;
    tmp *= 4294967296.0;
  }
  double ret = sum / tmp;
  if (ret >= 1.0) ret = nextafter(1.0, 0.0);
  return ret;
}

void oracle_sample(const double* state, uint64_t dim, uint32_t seed,
                   uint64_t shots, uint64_t* out) {
  oracle_mt19937 r;
  oracle_mt_seed(&r, seed);
  for (uint64_t i = 0; i < shots; ++i) {
    const double u = oracle_canonical(&r);
    out[i] = oracle_measure_u(state, dim, u);
  }
}

/* ---- expval: src/observable.cpp:18-83 ---- */
int oracle_pauli_masks(const char* pauli, int n, uint64_t* flip,
                       uint64_t* sign, int* n_y) {
  if ((int)strlen(pauli) != n) return 1;
  *flip = 0;
  *sign = 0;
  *n_y = 0;
  for (int i = 0; i < n; ++i) {
    char p = pauli[n - 1 - i];
    if (p >= 'a' && p <= 'z') p = (char)(p - ('a' - 'A'));
    const uint64_t bit = (uint64_t)1 << i;
    switch (p) {
      case 'I':
        break;
      case 'X':
        *flip |= bit;
        break;
      case 'Y':
        *flip |= bit;
        *sign |= bit;
        ++*n_y;
        break;
      case 'Z':
        *sign |= bit;
        break;
      default:
        return 2;
    }
  }
  return 0;
}

double oracle_expval(const double* state, int n, const char* pauli, int* err) {
  uint64_t flip, sign;
  int n_y;
  *err = oracle_pauli_masks(pauli, n, &flip, &sign, &n_y);
  if (*err) return 0.0;
  cx phase = mk(1.0, 0.0);
  switch (n_y % 4) {
    case 1:
      phase = mk(0.0, 1.0);
      break;
    case 2:
      phase = mk(-1.0, 0.0);
      break;
    case 3:
      phase = mk(0.0, -1.0);
      break;
    default:
      break;
  }
  const cx* s = (const cx*)state;
  const size_t dim = (size_t)1 << n;
  cx sum = {0.0, 0.0};
  for (size_t x = 0; x < dim; ++x) {
    const cx c = (__builtin_popcountll(x & sign) % 2 == 0)
                     ? phase
                     : mk(-phase.re, -phase.im);
    const cx t = cmul(c, s[x]);
    const cx o = s[x ^ flip];
    sum = cadd(sum, cmul(t, mk(o.re, -o.im)));
  }
  return sum.re;
}
