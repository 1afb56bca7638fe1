// PMM inside / outside / posterior / M-step kernels for gfx950.
//
// Layout (one forest = many trees, grouped into problems):
//  - nodes are numbered in height-level order; a level is a contiguous range.
//  - every (node, site) value lives in a row of K_t contiguous sites; leaf rows
//    (int16 codes) and internal rows (beta, alpha, labels) are packed
//    separately, row_off[v] giving the row start in the matching array.
//  - internal rows hold beta = (b0, bA) mantissas with a shared exponent bE
//    (inside, below the node), alpha = (o0, oS, Y, Z) with exponent aE and
//    specific state aG (outside at the node), and the az-partition label.
//
// A segment of W lanes handles one node and strides over its K sites, so
// per-node reductions over sites are segment shuffles.
#include "pmm_kernels.hpp"

namespace laml {

__global__ void k_coef(DevForest F, const double* __restrict__ d, const double* __restrict__ nu,
                       const int32_t* __restrict__ done) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= F.num_nodes) return;
  const int p = F.tree_prob[F.tree_of[v]];
  if (done && done[p]) return;
  const double dv = d[v];
  const double nv = nu[p];
  Coef c;
  c.p1n = exp(-(1.0 + nv) * dv);
  c.pn = exp(-nv * dv);
  c.mut = -expm1(-dv) * c.pn;
  c.sil = -expm1(-nv * dv);
  F.coef[v] = c;
}

template <int W>
__global__ void __launch_bounds__(256) k_labels(DevForest F, int beg, int cnt) {
  const int seg = blockIdx.x * (256 / W) + (int)(threadIdx.x / W);
  const int lane = threadIdx.x & (W - 1);
  if (seg >= cnt) return;
  const int v = beg + seg;
  const int t = F.tree_of[v];
  const int K = F.tree_nsites[t];
  const int c1 = F.child0[v], c2 = F.child1[v];
  const bool l1 = c1 < F.num_leaves, l2 = c2 < F.num_leaves;
  const int64_t rv = F.row_off[v], r1 = F.row_off[c1], r2 = F.row_off[c2];
  for (int k = lane; k < K; k += W) {
    const int la = l1 ? code_label(F.codes[r1 + k]) : (int)F.labels[r1 + k];
    const int lb = l2 ? code_label(F.codes[r2 + k]) : (int)F.labels[r2 + k];
    F.labels[rv + k] = (int16_t)combine_label(la, lb);
  }
}

// load the inside state of child c at site k
__device__ __forceinline__ void load_child(const DevForest& F, bool leaf, int64_t r, int k, double phi,
                                           int& lab, double& b0, double& bA, int& E, int& code) {
  if (leaf) {
    code = F.codes[r + k];
    leaf_beta(code, phi, lab, b0, bA);
    E = 0;
  } else {
    code = 0;
    lab = F.labels[r + k];
    const double2 bb = F.beta[r + k];
    b0 = bb.x;
    bA = bb.y;
    E = F.bE[r + k];
  }
}

template <int W>
__global__ void __launch_bounds__(256) k_inside(DevForest F, int beg, int cnt, const int32_t* __restrict__ done) {
  const int seg = blockIdx.x * (256 / W) + (int)(threadIdx.x / W);
  const int lane = threadIdx.x & (W - 1);
  if (seg >= cnt) return;
  const int v = beg + seg;
  const int t = F.tree_of[v];
  const int p = F.tree_prob[t];
  if (done && done[p]) return;
  const int K = F.tree_nsites[t];
  const int sb = F.tree_site_base[t];
  const int c1 = F.child0[v], c2 = F.child1[v];
  const Coef cf1 = F.coef[c1], cf2 = F.coef[c2];
  const double phi = F.phi[p];
  const bool l1 = c1 < F.num_leaves, l2 = c2 < F.num_leaves;
  const int64_t rv = F.row_off[v], r1 = F.row_off[c1], r2 = F.row_off[c2];
  for (int k = lane; k < K; k += W) {
    const int lv = F.labels[rv + k];
    int la, lb, Ea, Eb, ca, cb;
    double a0, aA, b0, bA;
    load_child(F, l1, r1, k, phi, la, a0, aA, Ea, ca);
    load_child(F, l2, r2, k, phi, lb, b0, bA, Eb, cb);
    const int qo = F.site_qoff[sb + k] - 1;
    const double qa = la > 0 ? F.prior[qo + la] : 1.0;
    const double qb = lb > 0 ? F.prior[qo + lb] : 1.0;
    SD e0a, eAa, e0b, eAb;
    edge_eps(la, a0, aA, Ea, qa, cf1, e0a, eAa);
    edge_eps(lb, b0, bA, Eb, qb, cf2, e0b, eAb);
    const SD B0 = sd_mul(e0a, e0b);
    const SD BA = lv != 0 ? sd_mul(eAa, eAb) : sd_zero();
    int E = max(sd_top(B0), sd_top(BA));
    if (E == kZeroTop) E = 0;
    F.beta[rv + k] = make_double2(ldexp(B0.m, B0.e - E), ldexp(BA.m, BA.e - E));
    F.bE[rv + k] = E;
  }
}

// Root of every tree: site likelihood L = P(D | stem starts unedited), the
// outside state at the root, and the root branch's posterior counts.
template <int W>
__global__ void __launch_bounds__(256) k_root(DevForest F, double* __restrict__ site_llh_out,
                                              double* __restrict__ counts, const int32_t* __restrict__ done) {
  const int t = blockIdx.x * (256 / W) + (int)(threadIdx.x / W);
  const int lane = threadIdx.x & (W - 1);
  if (t >= F.num_trees) return;
  const int p = F.tree_prob[t];
  if (done && done[p]) return;
  const int r = F.tree_root[t];
  const int K = F.tree_nsites[t];
  const int sb = F.tree_site_base[t];
  const Coef cf = F.coef[r];
  const double phi = F.phi[p];
  const bool leaf = r < F.num_leaves;
  const int64_t rr = F.row_off[r];
  double acc0 = 0, acc1 = 0, acc2 = 0;
  for (int k = lane; k < K; k += W) {
    int lab, E, code;
    double b0, bA;
    load_child(F, leaf, rr, k, phi, lab, b0, bA, E, code);
    const int qo = F.site_qoff[sb + k] - 1;
    const double q = lab > 0 ? F.prior[qo + lab] : 1.0;
    SD e0, eA;
    edge_eps(lab, b0, bA, E, q, cf, e0, eA);
    int LE = sd_top(e0);
    if (LE == kZeroTop) LE = 0;
    const SD L = sd(ldexp(e0.m, e0.e - LE), LE);
    F.Lm[sb + k] = L.m;
    F.LE[sb + k] = L.e;
    const double sl = sd_log(L);
    F.site_llh[sb + k] = sl;
    if (site_llh_out) site_llh_out[sb + k] = sl;
    if (!leaf) {
      F.alpha[rr + k] = make_double4(cf.p1n, cf.sil, cf.mut, 0.0);
      F.aE[rr + k] = 0;
      F.aG[rr + k] = 0;
    }
    if (lab == 0) {
      acc0 += 1.0;
    } else {
      acc0 += sd_div(sd(cf.p1n * b0, E), L);
      acc1 += sd_div(sd(cf.mut * q * bA, E), L);
      if (lab < 0) acc2 += sd_div(sd(cf.sil, 0), L);
    }
  }
  acc0 = seg_sum<W>(acc0);
  acc1 = seg_sum<W>(acc1);
  acc2 = seg_sum<W>(acc2);
  if (lane == 0 && counts) {
    double* c = counts + (int64_t)r * 5;
    c[0] = acc0;
    c[1] = acc1;
    c[2] = acc2;
    c[3] = 0.0;
    c[4] = 0.0;
  }
}

// Outside pass for the children of the internal nodes of one level, with the
// posterior expected transition counts of both child branches and the
// dropout responsibilities of missing leaf children.
template <int W>
__global__ void __launch_bounds__(256) k_outside(DevForest F, int beg, int cnt, double* __restrict__ counts,
                                                 const int32_t* __restrict__ done) {
  const int seg = blockIdx.x * (256 / W) + (int)(threadIdx.x / W);
  const int lane = threadIdx.x & (W - 1);
  if (seg >= cnt) return;
  const int u = beg + seg;
  const int t = F.tree_of[u];
  const int p = F.tree_prob[t];
  if (done && done[p]) return;
  const int K = F.tree_nsites[t];
  const int sb = F.tree_site_base[t];
  const int ch[2] = {F.child0[u], F.child1[u]};
  const Coef cf[2] = {F.coef[ch[0]], F.coef[ch[1]]};
  const bool lf[2] = {ch[0] < F.num_leaves, ch[1] < F.num_leaves};
  const int64_t rc[2] = {F.row_off[ch[0]], F.row_off[ch[1]]};
  const int64_t ru = F.row_off[u];
  const double phi = F.phi[p];
  double acc[2][5];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 5; ++j) acc[i][j] = 0.0;
  double rt = 0.0;
  for (int k = lane; k < K; k += W) {
    const double4 au = F.alpha[ru + k];
    const int Eo = F.aE[ru + k];
    const int gu = F.aG[ru + k];
    const SD L = sd(F.Lm[sb + k], F.LE[sb + k]);
    const int qo = F.site_qoff[sb + k] - 1;
    int lab[2], E[2], code[2];
    double b0[2], bA[2], q[2];
    SD e0[2], eA[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      load_child(F, lf[i], rc[i], k, phi, lab[i], b0[i], bA[i], E[i], code[i]);
      q[i] = lab[i] > 0 ? F.prior[qo + lab[i]] : 1.0;
      edge_eps(lab[i], b0[i], bA[i], E[i], q[i], cf[i], e0[i], eA[i]);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int w = 1 - i;
      // A = alpha_u(0) * eps_w(0): joint of the data outside v with u = 0
      const SD A = sd(au.x * e0[w].m, Eo + e0[w].e);
      // G(a) = alpha_u(a) eps_w(a) = q_a Gy + [a = gg] Gz ; GS = alpha_u(-1) eps_w(-1)
      SD Gy = sd_zero(), Gz = sd_zero(), GS = sd_zero();
      int gg = 0;
      if (lab[w] < 0) {
        Gy = sd(au.z * eA[w].m, Eo + eA[w].e);
        Gz = sd(au.w * eA[w].m, Eo + eA[w].e);
        gg = gu;
        GS = sd(au.y, Eo);
      } else if (lab[w] > 0) {
        const double tq = fma(q[w], au.z, lab[w] == gu ? au.w : 0.0);
        Gz = sd(tq * eA[w].m, Eo + eA[w].e);
        gg = lab[w];
      }
      const SD sumG = sd_add(Gy, Gz);
      const SD av0 = sd_scale(A, cf[i].p1n);
      const SD base = sd_scale(A, cf[i].mut);
      const SD Yv = sd_add(base, sd_scale(Gy, cf[i].pn));
      const SD Zv = sd_scale(Gz, cf[i].pn);
      if (!lf[i]) {
        const SD avS = sd_add(GS, sd_scale(sd_add(A, sumG), cf[i].sil));
        int Ev = max(max(sd_top(av0), sd_top(avS)), max(sd_top(Yv), sd_top(Zv)));
        if (Ev == kZeroTop) Ev = 0;
        F.alpha[rc[i] + k] = make_double4(ldexp(av0.m, av0.e - Ev), ldexp(avS.m, avS.e - Ev),
                                          ldexp(Yv.m, Yv.e - Ev), ldexp(Zv.m, Zv.e - Ev));
        F.aE[rc[i] + k] = Ev;
        F.aG[rc[i] + k] = (int16_t)gg;
      }
      const int lv = lab[i];
      if (lv == 0) {
        acc[i][0] += 1.0;
      } else {
        const SD bv0 = sd(b0[i], E[i]);
        const SD bvA = sd(bA[i], E[i]);
        acc[i][0] += sd_div(sd_mul(av0, bv0), L);
        const double qv = q[i];
        acc[i][1] += sd_div(sd_mul(sd_scale(base, qv), bvA), L);
        SD t3;
        if (lv > 0) {
          t3 = sd_add(sd_scale(Gy, qv), lv == gg ? Gz : sd_zero());
        } else {
          t3 = sumG;
        }
        acc[i][3] += sd_div(sd_mul(sd_scale(t3, cf[i].pn), bvA), L);
        if (lv < 0) {
          acc[i][2] += sd_div(sd_scale(A, cf[i].sil), L);
          acc[i][4] += sd_div(sd_scale(sumG, cf[i].sil), L);
          if (lf[i] && code[i] == -1) rt += phi * sd_div(sd_add(av0, sd_add(Yv, Zv)), L);
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 5; ++j) acc[i][j] = seg_sum<W>(acc[i][j]);
  rt = seg_sum<W>(rt);
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      if (counts) {
        double* c = counts + (int64_t)ch[i] * 5;
#pragma unroll
        for (int j = 0; j < 5; ++j) c[j] = acc[i][j];
      }
    }
    F.rt_part[u] = rt;
  }
}

// ---------------------------------------------------------------------------
// per-problem reductions (one block per problem, fixed order => deterministic)
// ---------------------------------------------------------------------------
template <int NT>
__device__ __forceinline__ double block_sum(double x, double* sh) {
  x = seg_sum<64>(x);
  const int wid = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sh[wid] = x;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < NT / 64; ++i) s += sh[i];
  __syncthreads();
  return s;
}

// Sum of per-site log-likelihoods of every tree of a problem, plus the EM
// bookkeeping of the device loop when em != nullptr.
__global__ void __launch_bounds__(256) k_llh(DevForest F, double* __restrict__ nll, EmState em, int step,
                                             int max_iter, double tol) {
  __shared__ double sh[4];
  const int p = blockIdx.x;
  if (em.done && em.done[p]) return;
  double s = 0.0;
  for (int t = F.prob_tree_off[p]; t < F.prob_tree_off[p + 1]; ++t) {
    const int K = F.tree_nsites[t];
    const int sb = F.tree_site_base[t];
    for (int k = threadIdx.x; k < K; k += 256) s += F.site_llh[sb + k];
  }
  const double llh = block_sum<256>(s, sh);
  if (threadIdx.x != 0) return;
  if (nll) nll[p] = -llh;
  if (!em.done) return;
  if (em.trace) em.trace[(int64_t)step * F.num_problems + p] = -llh;
  if (step > 0) {
    const double pre = em.pre[p];
    if (tol > 0.0 && fabs((llh - pre) / pre) < tol) {
      em.done[p] = 1;
      em.iters[p] = step;
      return;
    }
  }
  em.pre[p] = llh;
  if (step >= max_iter) {
    em.done[p] = 1;
    em.iters[p] = max_iter + 1;
  }
}

// Sum of the dropout responsibilities of missing leaf entries of a problem.
__global__ void __launch_bounds__(256) k_rtilde(DevForest F, double* __restrict__ rtilde,
                                                const int32_t* __restrict__ done) {
  __shared__ double sh[4];
  const int p = blockIdx.x;
  if (done && done[p]) return;
  double s = 0.0;
  for (int i = F.prob_node_off[p] + threadIdx.x; i < F.prob_node_off[p + 1]; i += 256) {
    const int v = F.prob_nodes[i];
    if (v >= F.num_leaves) s += F.rt_part[v];
  }
  const double rt = block_sum<256>(s, sh);
  if (threadIdx.x == 0) rtilde[p] = rt;
}

// ---------------------------------------------------------------------------
// M-step
// ---------------------------------------------------------------------------
constexpr double kEpsNu = 1e-5;   // EM_solver.py:307 eps_nu
constexpr double kEpsS = 1e-6;    // EM_solver.py:307 eps_s

struct BranchObj {
  double A, S1, B, nu;  // f'(d) = -A + S1/expm1(d) + B nu/expm1(nu d)
};
__device__ __forceinline__ void branch_g(const BranchObj& o, double d, double& g, double& h) {
  const double em = expm1(d);
  g = -o.A + o.S1 / em;
  h = -o.S1 * (em + 1.0) / (em * em);
  if (o.B > 0.0) {
    const double en = expm1(o.nu * d);
    g += o.B * o.nu / en;
    h -= o.B * o.nu * o.nu * (en + 1.0) / (en * en);
  }
}
// argmax over [lo, hi] of a strictly concave objective given its derivative:
// safeguarded Newton with bisection fallback, to full double precision.
__device__ double solve_branch(const BranchObj& o, double dmin, double dmax, double x0) {
  double g, h;
  branch_g(o, dmin, g, h);
  if (!(g > 0.0)) return dmin;
  branch_g(o, dmax, g, h);
  if (!(g < 0.0)) return dmax;
  double lo = dmin, hi = dmax;
  double x = (x0 > lo && x0 < hi) ? x0 : 0.5 * (lo + hi);
  for (int it = 0; it < 200; ++it) {
    branch_g(o, x, g, h);
    if (g > 0.0) lo = x; else if (g < 0.0) hi = x; else return x;
    double xn = x - g / h;
    if (!(xn > lo && xn < hi)) xn = 0.5 * (lo + hi);
    if (fabs(xn - x) <= 4e-16 * x || hi - lo <= 4e-16 * hi) return xn;
    x = xn;
  }
  return x;
}

__global__ void __launch_bounds__(256) k_mstep(DevForest F, const double* __restrict__ counts,
                                               const double* __restrict__ rtilde, double* __restrict__ d,
                                               double* __restrict__ nu, double* __restrict__ phi,
                                               const int32_t* __restrict__ done) {
  __shared__ double sh[4];
  __shared__ double s_nu;
  const int p = blockIdx.x;
  if (done && done[p]) return;
  const int nb = F.prob_node_off[p], ne = F.prob_node_off[p + 1];
  const uint32_t flags = F.prob_flags[p];
  const double K = (double)F.tree_nsites[F.prob_tree_off[p]];
  // phi (EM_solver.py:318-326)
  double phi_star = phi[p];
  if (flags & 1u) {
    const double rt = rtilde[p];
    phi_star = rt / (F.prob_nobs[p] + rt);
    if (fabs(phi_star) < F.prob_phimin[p]) phi_star = 0.0;
  }
  // branch lengths (EM_solver.py:328-366), exact per-branch maximisation
  const double nu0 = nu[p];
  const bool use_nu = nu0 > kEpsNu;
  double nfree = 0.0;
  for (int i = nb + threadIdx.x; i < ne; i += 256) {
    const int v = F.prob_nodes[i];
    if (!F.free_branch[v]) continue;
    const double* c = counts + (int64_t)v * 5;
    double s[5];
    double tot = 0.0;
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      s[j] = fmax(kEpsS, c[j]);
      tot += s[j];
    }
#pragma unroll
    for (int j = 0; j < 5; ++j) s[j] = s[j] / tot * K;
    double* sn = F.snorm + (int64_t)v * 5;
#pragma unroll
    for (int j = 0; j < 5; ++j) sn[j] = s[j];
    BranchObj o;
    o.nu = nu0;
    o.A = (nu0 + 1.0) * s[0] + nu0 * s[1] + nu0 * s[3];
    o.S1 = s[1];
    o.B = use_nu ? (s[2] + s[4]) : 0.0;
    d[v] = solve_branch(o, F.dmin, F.dmax, d[v]);
    nfree += 1.0;
  }
  nfree = block_sum<256>(nfree, sh);
  __syncthreads();
  // nu (EM_solver.py:392-407): maximise over [0, 1] given the new lengths
  if ((flags & 2u) && nfree > 0.0) {
    // h'(x) = sum_v -d (S0+S1+S3) + (S2+S4) d / expm1(x d)
    double lo = 0.0, hi = 1.0;
    double x = 1.0;
    bool fin = false;
    for (int it = 0; it < 400 && !fin; ++it) {
      double g = 0.0, h = 0.0;
      for (int i = nb + threadIdx.x; i < ne; i += 256) {
        const int v = F.prob_nodes[i];
        if (!F.free_branch[v]) continue;
        const double* s = F.snorm + (int64_t)v * 5;
        const double dv = d[v];
        const double en = expm1(x * dv);
        const double B = s[2] + s[4];
        g += -dv * (s[0] + s[1] + s[3]) + B * dv / en;
        h -= B * dv * dv * (en + 1.0) / (en * en);
      }
      g = block_sum<256>(g, sh);
      h = block_sum<256>(h, sh);
      if (it == 0) {
        if (!(g < 0.0)) {  // optimum at the upper bound
          x = 1.0;
          break;
        }
        hi = 1.0;
        x = (nu0 > 0.0 && nu0 < 1.0) ? nu0 : 0.5;
        continue;
      }
      if (g > 0.0) lo = x; else if (g < 0.0) hi = x; else break;
      double xn = x - g / h;
      if (!(xn > lo && xn < hi)) xn = 0.5 * (lo + hi);
      if (fabs(xn - x) <= 4e-16 * x || hi - lo <= 4e-16 * hi) {
        x = xn;
        fin = true;
      } else {
        x = xn;
      }
    }
    if (threadIdx.x == 0) s_nu = x;
    __syncthreads();
    if (threadIdx.x == 0) nu[p] = s_nu;
  }
  if (threadIdx.x == 0) phi[p] = phi_star;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <int W>
static void launch_seg(void (*kern)(DevForest, int, int, const int32_t*), const DevForest& F, int beg, int cnt,
                       const int32_t* done, hipStream_t s) {
  const int segs = 256 / W;
  const int grid = (cnt + segs - 1) / segs;
  if (grid > 0) hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, s, F, beg, cnt, done);
}

int seg_width(int max_k) {
  if (max_k <= 8) return 8;
  if (max_k <= 16) return 16;
  if (max_k <= 32) return 32;
  return 64;
}

#define LAML_DISPATCH_W(W, BODY)         \
  switch (W) {                           \
    case 8: { constexpr int WW = 8; BODY; } break;   \
    case 16: { constexpr int WW = 16; BODY; } break; \
    case 32: { constexpr int WW = 32; BODY; } break; \
    default: { constexpr int WW = 64; BODY; } break; \
  }

void run_labels(const DevForest& F, const std::vector<int32_t>& lvl, int W, hipStream_t s) {
  for (size_t h = 1; h + 1 < lvl.size(); ++h) {
    const int beg = lvl[h], cnt = lvl[h + 1] - lvl[h];
    LAML_DISPATCH_W(W, {
      const int grid = (cnt + 256 / WW - 1) / (256 / WW);
      if (grid > 0) hipLaunchKernelGGL(k_labels<WW>, dim3(grid), dim3(256), 0, s, F, beg, cnt);
    });
  }
}

void run_coef(const DevForest& F, const double* d, const double* nu, const int32_t* done, hipStream_t s) {
  const int grid = (F.num_nodes + 255) / 256;
  hipLaunchKernelGGL(k_coef, dim3(grid), dim3(256), 0, s, F, d, nu, done);
}

void run_inside(const DevForest& F, const std::vector<int32_t>& lvl, int W, const int32_t* done, hipStream_t s) {
  for (size_t h = 1; h + 1 < lvl.size(); ++h) {
    const int beg = lvl[h], cnt = lvl[h + 1] - lvl[h];
    LAML_DISPATCH_W(W, { launch_seg<WW>(k_inside<WW>, F, beg, cnt, done, s); });
  }
}

void run_root(const DevForest& F, int W, double* site_llh_out, double* counts, const int32_t* done, hipStream_t s) {
  LAML_DISPATCH_W(W, {
    const int grid = (F.num_trees + 256 / WW - 1) / (256 / WW);
    hipLaunchKernelGGL(k_root<WW>, dim3(grid), dim3(256), 0, s, F, site_llh_out, counts, done);
  });
}

void run_outside(const DevForest& F, const std::vector<int32_t>& lvl, int W, double* counts,
                 const int32_t* done, hipStream_t s) {
  for (size_t h = lvl.size() - 2; h >= 1; --h) {
    const int beg = lvl[h], cnt = lvl[h + 1] - lvl[h];
    LAML_DISPATCH_W(W, {
      const int grid = (cnt + 256 / WW - 1) / (256 / WW);
      if (grid > 0) hipLaunchKernelGGL(k_outside<WW>, dim3(grid), dim3(256), 0, s, F, beg, cnt, counts, done);
    });
  }
}

void run_llh(const DevForest& F, double* nll, EmState em, int step, int max_iter, double tol, hipStream_t s) {
  hipLaunchKernelGGL(k_llh, dim3(F.num_problems), dim3(256), 0, s, F, nll, em, step, max_iter, tol);
}

void run_rtilde(const DevForest& F, double* rtilde, const int32_t* done, hipStream_t s) {
  hipLaunchKernelGGL(k_rtilde, dim3(F.num_problems), dim3(256), 0, s, F, rtilde, done);
}

void run_mstep(const DevForest& F, const double* counts, const double* rtilde, double* d, double* nu, double* phi,
               const int32_t* done, hipStream_t s) {
  hipLaunchKernelGGL(k_mstep, dim3(F.num_problems), dim3(256), 0, s, F, counts, rtilde, d, nu, phi, done);
}

}  // namespace laml
