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
//   sil = P(0->-1) = P(a->-1) = 1-e^{-nu d} = msil * 2^esil (msil in [0.5, 1) or 0)
struct Coef {
  double p1n, mut, pn, msil;
  int esil, pad;
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

// Edge-inside of a child seen from its parent, both values on one exponent:
// eps(0) = e0 2^e = P(D_c | parent 0); eps(alpha) = eA 2^e = P(D_c | parent in
// the child's edit state) ('?' child: any edit state).  The child's inside
// vector is (b0, bA) 2^E with beta(-1) = 1 for '?' children; q is the prior
// of the child's specific state.  Branch-free over the three label kinds.
struct Eps {
  double e0, eA;
  int e;
};
__device__ __forceinline__ Eps edge(int lab, double b0, double bA, int E, double q, const Coef& c) {
  const bool z = lab == 0;
  const double t0 = fma(z ? 0.0 : c.mut * q, bA, c.p1n * b0);
  const double tA = z ? 0.0 : c.pn * bA;
  const int es = (lab < 0 && c.msil != 0.0) ? c.esil : kZeroTop;
  const int e = max(E, es);
  const double s0 = ldexp(1.0, E - e);
  const double s1 = ldexp(c.msil, es - e);
  Eps r;
  r.e0 = fma(t0, s0, s1);
  r.eA = fma(tA, s0, s1);
  r.e = e;
  return r;
}

// Streaming (non-temporal) row accesses.  A pass writes each inside or
// outside row once and the next level reads it once, so these accesses carry
// the non-temporal hint and leave the caches to the per-site tables (priors,
// site offsets, likelihoods) that every node of a tree reads.
typedef double nt_double2 __attribute__((ext_vector_type(2)));
typedef double nt_double4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ double2 load_stream(const double2* p) {
  const nt_double2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_double2*>(p));
  return make_double2(v.x, v.y);
}
__device__ __forceinline__ void store_stream(double2* p, const double2& a) {
  const nt_double2 v = {a.x, a.y};
  __builtin_nontemporal_store(v, reinterpret_cast<nt_double2*>(p));
}
__device__ __forceinline__ void store_stream(double4* p, const double4& a) {
  const nt_double4 v = {a.x, a.y, a.z, a.w};
  __builtin_nontemporal_store(v, reinterpret_cast<nt_double4*>(p));
}

// binary exponent x of m (m in [2^(x-1), 2^x)); 0 for m == 0
__device__ __forceinline__ int exp_of(double m) {
  int ex;
  (void)frexp(m, &ex);
  return ex;
}

// 64-bit DPP move of a double (rows not in ROW_MASK read 0)
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ double dpp_mov(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, ROW_MASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, ROW_MASK, 0xf, false);
  return __hiloint2double(hi, lo);
}

// 64-bit DPP move for patterns that write every lane (row rotations, quad
// permutations): no old value to preserve
template <int CTRL>
__device__ __forceinline__ double dpp_all(double x) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

// Lane swaps of two doubles: swap32 trades lanes 32..63 of x with lanes
// 0..31 of y; swap16 trades the odd 16-lane rows of x with the even rows of
// y (v_permlane32_swap / v_permlane16_swap).
__device__ __forceinline__ void swap32(double& x, double& y) {
  const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)__double2loint(x), (unsigned)__double2loint(y),
                                                   false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)__double2hiint(x), (unsigned)__double2hiint(y),
                                                   false, false);
  x = __hiloint2double((int)hi[0], (int)lo[0]);
  y = __hiloint2double((int)hi[1], (int)lo[1]);
}
__device__ __forceinline__ void swap16(double& x, double& y) {
  const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)__double2loint(x), (unsigned)__double2loint(y),
                                                   false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)__double2hiint(x), (unsigned)__double2hiint(y),
                                                   false, false);
  x = __hiloint2double((int)hi[0], (int)lo[0]);
  y = __hiloint2double((int)hi[1], (int)lo[1]);
}

// Totals over the 64 lanes of a wavefront of twelve values v[0..11].  Two
// butterfly steps pair lanes 32 and then 16 apart while halving the values
// each lane carries (12 -> 6 -> 3), then rotations within each 16-lane row
// finish the sums: on return every lane of row r (lanes 16r .. 16r + 15)
// holds in v[0..2] the totals of values 3r .. 3r + 2.  About 60
// instructions instead of 12 separate reductions.
__device__ __forceinline__ void wave_sum12(double (&v)[12]) {
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    // lanes 0..31: value j of lanes L and L + 32; lanes 32..63: value j + 6
    swap32(v[j], v[j + 6]);
    v[j] += v[j + 6];
  }
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    // rows 0 / 2 keep v[j], rows 1 / 3 take v[j + 3], each summed with the
    // row 16 lanes away: row r now holds value j + 3r
    swap16(v[j], v[j + 3]);
    v[j] += v[j + 3];
  }
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    v[j] += dpp_all<0x128>(v[j]);  // row_ror:8
    v[j] += dpp_all<0x124>(v[j]);  // row_ror:4
    v[j] += dpp_all<0x4e>(v[j]);   // quad_perm [2,3,0,1]
    v[j] += dpp_all<0xb1>(v[j]);   // quad_perm [1,0,3,2]
  }
}

// the value of lane 63 of the wavefront, in every lane
__device__ __forceinline__ double bcast_last(double x) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(x), 63);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(x), 63);
  return __hiloint2double(hi, lo);
}

// Sum over an aligned segment of W lanes (W = 8..64) with DPP: quad
// permutations, row rotations and row broadcasts.  The sum is valid in the
// LAST lane of each segment.
template <int W>
__device__ __forceinline__ double seg_sum_last(double x) {
  x += dpp_mov<0xb1>(x);  // quad_perm [1,0,3,2]
  x += dpp_mov<0x4e>(x);  // quad_perm [2,3,0,1]
  if (W >= 8) x += dpp_mov<0x124>(x);  // row_ror:4
  if (W >= 16) x += dpp_mov<0x128>(x);  // row_ror:8
  if (W >= 32) x += dpp_mov<0x142, 0xa>(x);  // row_bcast:15 into rows 1 and 3
  if (W >= 64) x += dpp_mov<0x143, 0xc>(x);  // row_bcast:31 into rows 2 and 3
  return x;
}

template <int W>
__device__ __forceinline__ double seg_sum(double x) {
#pragma unroll
  for (int off = W / 2; off > 0; off >>= 1) x += __shfl_xor(x, off, W);
  return x;
}

}  // namespace laml
