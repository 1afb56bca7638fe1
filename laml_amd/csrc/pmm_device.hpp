// Device-side building blocks of the PMM engine (gfx950, FP64).
//
// Values are kept in linear space as (mantissa, power-of-two exponent) pairs so
// that the per-(node, site) work is multiplications and adds; exp/log appear
// only per branch (transition coefficients) and per site (the root's log).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace laml {

// value = m * 2^e
struct SD {
  double m;
  int e;
};

__device__ __forceinline__ SD sd(double m, int e) {
  SD r;
  r.m = m;
  r.e = e;
  return r;
}
__device__ __forceinline__ SD sd_zero() { return sd(0.0, 0); }
__device__ __forceinline__ SD sd_mul(SD a, SD b) { return sd(a.m * b.m, a.e + b.e); }
__device__ __forceinline__ SD sd_scale(SD a, double c) { return sd(a.m * c, a.e); }
__device__ __forceinline__ SD sd_add(SD a, SD b) {
  if (a.m == 0.0) return b;
  if (b.m == 0.0) return a;
  const int e = a.e > b.e ? a.e : b.e;
  return sd(ldexp(a.m, a.e - e) + ldexp(b.m, b.e - e), e);
}
// exponent t with |value| in [2^(t-1), 2^t); a very small sentinel for zero
constexpr int kZeroTop = -0x3fffffff;
__device__ __forceinline__ int sd_top(SD a) {
  int ex;
  (void)frexp(a.m, &ex);
  return a.m != 0.0 ? a.e + ex : kZeroTop;
}
// a / b as a plain double (used for posteriors, which are O(1))
__device__ __forceinline__ double sd_div(SD a, SD b) {
  if (a.m == 0.0) return 0.0;
  return ldexp(a.m / b.m, a.e - b.e);
}
// the value as a plain double (used for O(1) posteriors)
__device__ __forceinline__ double sd_val(SD a) { return ldexp(a.m, a.e); }
__device__ __forceinline__ double sd_log(SD a) {
  return log(a.m) + (double)a.e * 0.69314718055994530942;
}

// transition coefficients of one branch of length d (rate 1, silencing nu):
//   p1n = P(0->0) = e^{-(1+nu)d}
//   mut = P(0->a)/q_a = (1-e^{-d}) e^{-nu d}
//   pn  = P(a->a) = e^{-nu d}
//   sil = P(0->-1) = P(a->-1) = 1-e^{-nu d}
struct Coef {
  double p1n, mut, pn, sil;
};

// label codes: 0 = 'z', -1 = '?', j>0 = edit state j of the site
__device__ __forceinline__ int code_label(int code) { return code > 0 ? code : (code == 0 ? 0 : -1); }

__device__ __forceinline__ int combine_label(int a, int b) {
  if (a == 0 || b == 0) return 0;
  if (a < 0) return b;
  if (b < 0) return a;
  return a == b ? a : 0;
}

// inside vector of a leaf below its edge: beta(0), beta(alpha); beta(-1) is
// 1 for '?'/-1 codes and 0 otherwise (implied by the label)
__device__ __forceinline__ void leaf_beta(int code, double phi, int& lab, double& b0, double& bA) {
  if (code == 0) {
    lab = 0;
    b0 = 1.0 - phi;
    bA = 0.0;
  } else if (code > 0) {
    lab = code;
    b0 = 0.0;
    bA = 1.0 - phi;
  } else if (code == -1) {
    lab = -1;
    b0 = phi;
    bA = phi;
  } else {
    lab = -1;
    b0 = 0.0;
    bA = 0.0;
  }
}

// edge-inside of a child seen from its parent: e0 = P(D_c | parent 0),
// eA = P(D_c | parent in the child's edit state) ('?' child: any edit).
// q is the prior of the child's specific state (1 for '?').
__device__ __forceinline__ void edge_eps(int lab, double b0, double bA, int E, double q, const Coef& c,
                                         SD& e0, SD& eA) {
  if (lab == 0) {
    e0 = sd(c.p1n * b0, E);
    eA = sd_zero();
    return;
  }
  e0 = sd(fma(c.mut * q, bA, c.p1n * b0), E);
  eA = sd(c.pn * bA, E);
  if (lab < 0) {
    e0 = sd_add(e0, sd(c.sil, 0));
    eA = sd_add(eA, sd(c.sil, 0));
  }
}

template <int W>
__device__ __forceinline__ double seg_sum(double x) {
#pragma unroll
  for (int off = W / 2; off > 0; off >>= 1) x += __shfl_xor(x, off, W);
  return x;
}

}  // namespace laml
