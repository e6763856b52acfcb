// Host FP64 gate math for the planner.
//
// Gate matrices follow ket's per-gate builders (src/simulator.cpp:38-97,
// 168-200, 322-325) and use the same libm calls, so they are the same
// doubles. `apply_on_small` runs one gate over a tiny state with ket's kernel
// arithmetic (src/simulator.cpp:23-254); the Fuser's block matrix
// (src/simulator.cpp:448-465) is built from it column by column, which makes
// fused blocks in reference mode equal to ket's value for value. This file is
// compiled with -ffp-contract=off like the kernels.
#pragma once

#include <cmath>
#include <math.h>
#include <array>
#include <cstddef>
#include <cstdint>
#include <utility>
#include <vector>

#include "ket_amd.h"

namespace kethost {

struct cx {
  double re, im;
};

// cos and sin of one argument from glibc's sincos(). ket is built with g++,
// which merges each std::cos(x)/std::sin(x) pair of the same argument into
// one sincos() call; sincos can differ from separate sin() and cos() calls in
// the last bit, so the planner calls it explicitly to get ket's doubles
// whatever compiler builds this file.
struct CosSin {
  double c, s;
};
inline CosSin cos_sin(double x) {
  CosSin r;
  ::sincos(x, &r.s, &r.c);
  return r;
}
inline cx mul(cx a, cx b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
inline cx add(cx a, cx b) { return {a.re + b.re, a.im + b.im}; }

inline int arity(int type) {
  if (type >= KET_H && type <= KET_U) return 1;
  if (type >= KET_CH && type <= KET_SWAP) return 2;
  if (type >= KET_CCX && type <= KET_CSWAP) return 3;
  return 0;
}
inline bool is_unitary(int type) { return arity(type) > 0; }

// The 2x2 matrix ket uses for a single-qubit gate (m00, m01, m10, m11).
inline void single_matrix(int type, const double* p, cx m[4]) {
  const cx Z0{0.0, 0.0}, O1{1.0, 0.0};
  const double r = 1.0 / std::sqrt(2.0);
  switch (type) {
    case KET_H: m[0] = {r, 0}; m[1] = {r, 0}; m[2] = {r, 0}; m[3] = {-r, 0}; return;
    case KET_X: m[0] = Z0; m[1] = O1; m[2] = O1; m[3] = Z0; return;
    case KET_Y: m[0] = Z0; m[1] = {0.0, -1.0}; m[2] = {0.0, 1.0}; m[3] = Z0; return;
    case KET_Z: m[0] = O1; m[1] = Z0; m[2] = Z0; m[3] = {-1.0, 0.0}; return;
    case KET_S: m[0] = O1; m[1] = Z0; m[2] = Z0; m[3] = {0.0, 1.0}; return;
    case KET_SDG: m[0] = O1; m[1] = Z0; m[2] = Z0; m[3] = {0.0, -1.0}; return;
    case KET_T: m[0] = O1; m[1] = Z0; m[2] = Z0; m[3] = {r, r}; return;
    case KET_TDG: m[0] = O1; m[1] = Z0; m[2] = Z0; m[3] = {r, -r}; return;
    default: break;
  }
  const CosSin h = cos_sin(p[0] / 2.0);
  const double c = h.c, sn = h.s;
  switch (type) {
    case KET_RX: case KET_CRX:
      m[0] = {c, 0}; m[1] = {0.0, -sn}; m[2] = {0.0, -sn}; m[3] = {c, 0}; return;
    case KET_RY: case KET_CRY:
      m[0] = {c, 0}; m[1] = {-sn, 0}; m[2] = {sn, 0}; m[3] = {c, 0}; return;
    case KET_RZ: case KET_CRZ:
      m[0] = {c, -sn}; m[1] = Z0; m[2] = Z0; m[3] = {c, sn}; return;
    case KET_U: case KET_CU:
    {
      const CosSin l = cos_sin(p[2]), f = cos_sin(p[1]), fl = cos_sin(p[1] + p[2]);
      m[0] = {c, 0.0};
      m[1] = {-(sn * l.c), -(sn * l.s)};
      m[2] = {sn * f.c, sn * f.s};
      m[3] = {c * fl.c, c * fl.s};
      return;
    }
    default:
      m[0] = O1; m[1] = Z0; m[2] = Z0; m[3] = O1; return;
  }
}

// One gate over a small state (size 2^m) with ket's kernel arithmetic; `on`
// are local qubit indices in the gate's operand order.
inline void apply_on_small(std::vector<cx>& s, int type, const int* on,
                           const double* p) {
  const std::size_t dim = s.size();
  auto bit = [](int q) { return std::size_t{1} << q; };
  auto single = [&](std::size_t tm, std::size_t cm, const cx m[4]) {
    for (std::size_t k = 0; k < dim; ++k) {
      if ((k & tm) || (k & cm) != cm) continue;
      const std::size_t k1 = k | tm;
      const cx a = s[k], b = s[k1];
      s[k] = add(mul(m[0], a), mul(m[1], b));
      s[k1] = add(mul(m[2], a), mul(m[3], b));
    }
  };
  cx m[4];
  switch (type) {
    case KET_H: case KET_X: case KET_Y: case KET_Z: case KET_S: case KET_SDG:
    case KET_T: case KET_TDG: case KET_RX: case KET_RY: case KET_RZ: case KET_U:
      single_matrix(type, p, m);
      single(bit(on[0]), 0, m);
      return;
    case KET_CH: {
      const double r = 1.0 / std::sqrt(2.0);
      const std::size_t cm = bit(on[0]), tm = bit(on[1]);
      for (std::size_t k = 0; k < dim; ++k) {
        if (!(k & cm) || (k & tm)) continue;
        const cx a = s[k], b = s[k | tm];
        s[k] = {(a.re + b.re) * r, (a.im + b.im) * r};
        s[k | tm] = {(a.re - b.re) * r, (a.im - b.im) * r};
      }
      return;
    }
    case KET_CX: case KET_CCX: {
      const int nc = type == KET_CX ? 1 : 2;
      std::size_t cm = 0;
      for (int i = 0; i < nc; ++i) cm |= bit(on[i]);
      const std::size_t tm = bit(on[nc]);
      for (std::size_t k = 0; k < dim; ++k)
        if ((k & cm) == cm && !(k & tm)) std::swap(s[k], s[k | tm]);
      return;
    }
    case KET_CY: {
      const std::size_t cm = bit(on[0]), tm = bit(on[1]);
      for (std::size_t k = 0; k < dim; ++k) {
        if (!(k & cm) || (k & tm)) continue;
        const cx a = s[k], b = s[k | tm];
        s[k] = mul(cx{0.0, -1.0}, b);
        s[k | tm] = mul(cx{0.0, 1.0}, a);
      }
      return;
    }
    case KET_CRX: case KET_CRY: case KET_CRZ: case KET_CU:
      single_matrix(type, p, m);
      single(bit(on[1]), bit(on[0]), m);
      return;
    case KET_CZ: case KET_CP: {
      const CosSin t = cos_sin(p[0]);
      const cx ph = type == KET_CZ ? cx{-1.0, 0.0} : cx{t.c, t.s};
      const std::size_t ma = bit(on[0]), mb = bit(on[1]);
      for (std::size_t k = 0; k < dim; ++k)
        if ((k & ma) && (k & mb)) s[k] = mul(s[k], ph);
      return;
    }
    case KET_SWAP: case KET_CSWAP: {
      const int o = type == KET_SWAP ? 0 : 1;
      const std::size_t cm = type == KET_SWAP ? 0 : bit(on[0]);
      const std::size_t ma = bit(on[o]), mb = bit(on[o + 1]);
      for (std::size_t k = 0; k < dim; ++k) {
        if ((k & cm) != cm) continue;
        if ((k & ma) && !(k & mb)) std::swap(s[k], s[(k ^ ma) | mb]);
      }
      return;
    }
    default:
      return;
  }
}

// Row-major 2^m x 2^m matrix of a gate run on block qubits `block`: column c
// is the run applied to |c> in the block's local qubit order.
struct GateRef {
  int type;
  int local[3];
  const double* params;
  const cx* mat2 = nullptr;  // if set: a one-qubit 2x2 (m00, m01, m10, m11) on local[0]
};

// Fused blocks (not ket's Fuser): each run of one-qubit gates on the same
// block qubit, with no other gate on that qubit in between, becomes one 2x2
// (later gates on the left), so block_matrix applies it to the 2^k columns
// once. `store` holds the products and must outlive the returned refs.
inline std::vector<GateRef> merge_single_runs(const std::vector<GateRef>& in,
                                              std::vector<std::array<cx, 4>>& store) {
  std::vector<GateRef> out;
  std::vector<long> slot;          // out[i]'s product in `store`, or -1
  long open[8] = {-1, -1, -1, -1, -1, -1, -1, -1};  // local qubit -> pending out index
  store.clear();
  for (const GateRef& g : in) {
    if (arity(g.type) == 1) {
      cx m[4];
      single_matrix(g.type, g.params, m);
      const int q = g.local[0];
      if (open[q] >= 0) {
        std::array<cx, 4>& a = store[(std::size_t)slot[(std::size_t)open[q]]];
        a = {add(mul(m[0], a[0]), mul(m[1], a[2])), add(mul(m[0], a[1]), mul(m[1], a[3])),
             add(mul(m[2], a[0]), mul(m[3], a[2])), add(mul(m[2], a[1]), mul(m[3], a[3]))};
      } else {
        store.push_back({m[0], m[1], m[2], m[3]});
        open[q] = (long)out.size();
        out.push_back(GateRef{KET_U, {q, 0, 0}, nullptr});
        slot.push_back((long)store.size() - 1);
      }
      continue;
    }
    for (int j = 0; j < arity(g.type); ++j) open[g.local[j]] = -1;
    out.push_back(g);
    slot.push_back(-1);
  }
  for (std::size_t i = 0; i < out.size(); ++i)
    if (slot[i] >= 0) out[i].mat2 = store[(std::size_t)slot[i]].data();
  return out;
}

// All d columns advance together, each with apply_on_small's arithmetic (the
// same products and sums per element), and each gate's 2x2 matrix or phase
// is computed once instead of once per column.
inline std::vector<cx> block_matrix(const std::vector<GateRef>& gates, int m) {
  const std::size_t d = std::size_t{1} << m;
  std::vector<cx> s(d * d, cx{0.0, 0.0});  // s[row * d + col]: column col
  for (std::size_t col = 0; col < d; ++col) s[col * d + col] = {1.0, 0.0};
  auto bit = [](int q) { return std::size_t{1} << q; };
  auto single = [&](std::size_t tm, std::size_t cm, const cx mm[4]) {
    for (std::size_t k = 0; k < d; ++k) {
      if ((k & tm) || (k & cm) != cm) continue;
      cx* a = &s[k * d];
      cx* b = &s[(k | tm) * d];
      for (std::size_t c = 0; c < d; ++c) {
        const cx x = a[c], y = b[c];
        a[c] = add(mul(mm[0], x), mul(mm[1], y));
        b[c] = add(mul(mm[2], x), mul(mm[3], y));
      }
    }
  };
  auto swap_rows = [&](std::size_t k0, std::size_t k1) {
    for (std::size_t c = 0; c < d; ++c) std::swap(s[k0 * d + c], s[k1 * d + c]);
  };
  for (const GateRef& g : gates) {
    const int* on = g.local;
    cx mm[4];
    switch (g.type) {
      case KET_H: case KET_X: case KET_Y: case KET_Z: case KET_S: case KET_SDG:
      case KET_T: case KET_TDG: case KET_RX: case KET_RY: case KET_RZ: case KET_U:
        if (g.mat2) {
          single(bit(on[0]), 0, g.mat2);
          break;
        }
        single_matrix(g.type, g.params, mm);
        single(bit(on[0]), 0, mm);
        break;
      case KET_CRX: case KET_CRY: case KET_CRZ: case KET_CU:
        single_matrix(g.type, g.params, mm);
        single(bit(on[1]), bit(on[0]), mm);
        break;
      case KET_CH: {
        const double r = 1.0 / std::sqrt(2.0);
        const std::size_t cm = bit(on[0]), tm = bit(on[1]);
        for (std::size_t k = 0; k < d; ++k) {
          if (!(k & cm) || (k & tm)) continue;
          cx* a = &s[k * d];
          cx* b = &s[(k | tm) * d];
          for (std::size_t c = 0; c < d; ++c) {
            const cx x = a[c], y = b[c];
            a[c] = {(x.re + y.re) * r, (x.im + y.im) * r};
            b[c] = {(x.re - y.re) * r, (x.im - y.im) * r};
          }
        }
        break;
      }
      case KET_CX: case KET_CCX: {
        const int nc = g.type == KET_CX ? 1 : 2;
        std::size_t cm = 0;
        for (int i = 0; i < nc; ++i) cm |= bit(on[i]);
        const std::size_t tm = bit(on[nc]);
        for (std::size_t k = 0; k < d; ++k)
          if ((k & cm) == cm && !(k & tm)) swap_rows(k, k | tm);
        break;
      }
      case KET_CY: {
        const std::size_t cm = bit(on[0]), tm = bit(on[1]);
        for (std::size_t k = 0; k < d; ++k) {
          if (!(k & cm) || (k & tm)) continue;
          cx* a = &s[k * d];
          cx* b = &s[(k | tm) * d];
          for (std::size_t c = 0; c < d; ++c) {
            const cx x = a[c], y = b[c];
            a[c] = mul(cx{0.0, -1.0}, y);
            b[c] = mul(cx{0.0, 1.0}, x);
          }
        }
        break;
      }
      case KET_CZ: case KET_CP: {
        const CosSin t = cos_sin(g.params[0]);
        const cx ph = g.type == KET_CZ ? cx{-1.0, 0.0} : cx{t.c, t.s};
        const std::size_t ma = bit(on[0]), mb = bit(on[1]);
        for (std::size_t k = 0; k < d; ++k)
          if ((k & ma) && (k & mb))
            for (std::size_t c = 0; c < d; ++c) s[k * d + c] = mul(s[k * d + c], ph);
        break;
      }
      case KET_SWAP: case KET_CSWAP: {
        const int o = g.type == KET_SWAP ? 0 : 1;
        const std::size_t cm = g.type == KET_SWAP ? 0 : bit(on[0]);
        const std::size_t ma = bit(on[o]), mb = bit(on[o + 1]);
        for (std::size_t k = 0; k < d; ++k) {
          if ((k & cm) != cm) continue;
          if ((k & ma) && !(k & mb)) swap_rows(k, (k ^ ma) | mb);
        }
        break;
      }
      default:
        break;
    }
  }
  return s;  // row-major: s[row * d + col]
}

}  // namespace kethost
