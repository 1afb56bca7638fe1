/*
 * CPU restatement of LAML's EM_solver E-step and of the oracle's exact
 * M-step, in C with OpenMP over sites.  TEST INFRASTRUCTURE ONLY: it is the
 * CPU baseline that bench.py times and a second restatement that tests hold
 * against the Python oracle (oracle/em_oracle.py); the product never links
 * or calls it.
 *
 * The arithmetic follows the reference function by function, in the same
 * log space and with the same operations, so that it agrees with the Python
 * restatement to rounding:
 *   az_partition                laml_libs/ML_solver.py:125-154
 *   Estep_in_llh                laml_libs/EM_solver.py:91-133
 *   lineage_llh                 laml_libs/EM_solver.py:135-141
 *   Estep_out_llh               laml_libs/EM_solver.py:143-204
 *   __out_alpha_up__            laml_libs/EM_solver.py:206-245
 *   Estep_posterior             laml_libs/EM_solver.py:247-300
 *   Mstep (phi, branch rows)    laml_libs/EM_solver.py:307-348
 *   EM_optimization             laml_libs/EM_solver.py:461-512
 * The branch and nu subproblems of the M-step (solved with MOSEK in the
 * reference, EM_solver.py:349-407) are solved exactly by bisection on their
 * derivatives, as oracle/em_oracle.py does.
 *
 * Every site's inside, outside and posterior recursions are independent of
 * the other sites, so one thread runs whole sites over node-sized arrays and
 * the per-branch sums are reduced over threads at the end.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <omp.h>

#define LAB_Z 0      /* 'z' */
#define LAB_Q (-1)   /* '?' */
#define CODE_MISSING (-1)
#define CODE_SILENCED (-2)

typedef struct {
  int32_t n;                /* nodes in post-order (children before parents) */
  int32_t K;                /* sites */
  const int32_t* c0;        /* first child or -1 */
  const int32_t* c1;        /* second child or -1 */
  const int32_t* parent;    /* parent or -1 (a root) */
  const int32_t* root_of;   /* root of the node's tree */
  const int32_t* leaf_row;  /* row of `codes` for a leaf, -1 otherwise */
  const int16_t* codes;     /* [leaves x K]: 0, 1..m (local state), -1 '?', -2 an explicit -1 */
  const int32_t* qoff;      /* [K+1]: priors of site k at q[qoff[k] + state - 1] */
  const double* q;          /* normalised priors */
  const uint8_t* free_;     /* branch optimised by the M-step */
  double phimin;            /* phi* below this becomes 0 (EM_solver.py:324-326) */
} pmm_problem;

static const double NEG_INF = -INFINITY;

/* EM_solver.py:8-13: max + log(sum(exp(x - max))); all -inf gives NaN */
static inline double lse2(double a, double b) {
  const double m = a > b ? a : b;
  return m + log(exp(a - m) + exp(b - m));
}
static inline double lse3(double a, double b, double c) {
  double m = a > b ? a : b;
  m = m > c ? m : c;
  return m + log(exp(a - m) + exp(b - m) + exp(c - m));
}
static inline double pseudo_log(double x) { return x > 0 ? log(x) : NEG_INF; }

static inline double qval(const pmm_problem* P, int k, int a) { return P->q[P->qoff[k] + a - 1]; }

/* ---------------- (node, state) -> out_alpha map of one site ---------------- */
typedef struct {
  int64_t* key;
  double* val;
  uint32_t* gen;
  uint32_t cap, used, cur;
} oa_map;

static void oa_init(oa_map* m, uint32_t cap) {
  m->cap = cap;
  m->key = (int64_t*)malloc(sizeof(int64_t) * cap);
  m->val = (double*)malloc(sizeof(double) * cap);
  m->gen = (uint32_t*)calloc(cap, sizeof(uint32_t));
  m->used = 0;
  m->cur = 1;
}
static void oa_free(oa_map* m) {
  free(m->key);
  free(m->val);
  free(m->gen);
}
static void oa_clear(oa_map* m) {
  m->cur++;
  m->used = 0;
  if (m->cur == 0) { /* generation counter wrapped */
    memset(m->gen, 0, sizeof(uint32_t) * m->cap);
    m->cur = 1;
  }
}
static inline uint32_t oa_hash(int64_t k, uint32_t cap) {
  uint64_t x = (uint64_t)k * 0x9E3779B97F4A7C15ull;
  return (uint32_t)(x >> 32) & (cap - 1);
}
static double* oa_find(oa_map* m, int v, int a) {
  const int64_t k = ((int64_t)v << 32) | (uint32_t)a;
  for (uint32_t h = oa_hash(k, m->cap);; h = (h + 1) & (m->cap - 1)) {
    if (m->gen[h] != m->cur) return NULL;
    if (m->key[h] == k) return &m->val[h];
  }
}
static void oa_set(oa_map* m, int v, int a, double x);
static void oa_grow(oa_map* m) {
  oa_map n;
  oa_init(&n, m->cap * 2);
  for (uint32_t i = 0; i < m->cap; ++i)
    if (m->gen[i] == m->cur) oa_set(&n, (int)(m->key[i] >> 32), (int)(uint32_t)m->key[i], m->val[i]);
  oa_free(m);
  *m = n;
}
static void oa_set(oa_map* m, int v, int a, double x) {
  if (2 * (m->used + 1) > m->cap) oa_grow(m);
  const int64_t k = ((int64_t)v << 32) | (uint32_t)a;
  for (uint32_t h = oa_hash(k, m->cap);; h = (h + 1) & (m->cap - 1)) {
    if (m->gen[h] != m->cur) {
      m->gen[h] = m->cur;
      m->key[h] = k;
      m->val[h] = x;
      m->used++;
      return;
    }
    if (m->key[h] == k) {
      m->val[h] = x;
      return;
    }
  }
}

/* ---------------- per-thread work arrays ---------------- */
typedef struct {
  int* lab;
  double *L0, *L1, *A, *X, *out0, *out1, *post0, *post1;
  oa_map oa;
} site_ws;

static void ws_init(site_ws* w, int n) {
  w->lab = (int*)malloc(sizeof(int) * n);
  double* b = (double*)malloc(sizeof(double) * 8 * (size_t)n);
  w->L0 = b;
  w->L1 = b + n;
  w->A = b + 2 * (size_t)n;
  w->X = b + 3 * (size_t)n;
  w->out0 = b + 4 * (size_t)n;
  w->out1 = b + 5 * (size_t)n;
  w->post0 = b + 6 * (size_t)n;
  w->post1 = b + 7 * (size_t)n;
  uint32_t cap = 1024;
  while (cap < 4u * (uint32_t)n) cap <<= 1;
  oa_init(&w->oa, cap);
}
static void ws_free(site_ws* w) {
  free(w->lab);
  free(w->L0);
  oa_free(&w->oa);
}

static inline int sibling(const pmm_problem* P, int v) {
  const int u = P->parent[v];
  return P->c0[u] == v ? P->c1[u] : P->c0[u];
}

/* az_partition + Estep_in_llh for site k */
static void inside_site(const pmm_problem* P, const double* d, double nu, double phi, int k, site_ws* w) {
  const int n = P->n, K = P->K;
  for (int v = 0; v < n; ++v) {
    const double dv = d[v];
    const double p = exp(-dv);
    int a;
    if (P->c0[v] < 0) {
      const int c = P->codes[(size_t)P->leaf_row[v] * K + k];
      a = c == 0 ? LAB_Z : (c < 0 ? LAB_Q : c);
      w->lab[v] = a;
      if (a == LAB_Q) {
        const double t = c == CODE_MISSING ? 1 - (1 - phi) * pow(p, nu) : 1 - pow(p, nu);
        w->L0[v] = w->L1[v] = t > 0 ? log(t) : NEG_INF;
      } else if (a == LAB_Z) {
        w->L0[v] = 1 - phi > 0 ? (nu + 1) * (-dv) + log(1 - phi) : NEG_INF;
        w->L1[v] = NEG_INF;
      } else {
        const double q = qval(P, k, a);
        w->L0[v] = (1 - p) * q * (1 - phi) > 0 ? nu * (-dv) + log(1 - p) + log(q) + log(1 - phi) : NEG_INF;
        w->L1[v] = (1 - phi) > 0 ? nu * (-dv) + log(1 - phi) : NEG_INF;
      }
      continue;
    }
    const int x = P->c0[v], y = P->c1[v];
    const int la = w->lab[x], lb = w->lab[y];
    if (la == LAB_Z || lb == LAB_Z || (la > 0 && lb > 0 && la != lb))
      a = LAB_Z;
    else if (la > 0)
      a = la;
    else if (lb > 0)
      a = lb;
    else
      a = LAB_Q;
    w->lab[v] = a;
    const double q = a > 0 ? qval(P, k, a) : 1.0;
    const double l0 = w->L0[x] + w->L0[y];
    const double l1 = w->L1[x] + w->L1[y];
    const double t0 = l0 + (nu + 1) * (-dv);
    const int has1 = a != LAB_Z && q * (1 - p) > 0;
    const double pn = pow(p, nu);
    const int has2 = a == LAB_Q && (1 - pn) > 0;
    if (has1 && has2)
      w->L0[v] = lse3(t0, l1 + log(1 - p) + log(q) + nu * (-dv), log(1 - pn));
    else if (has1)
      w->L0[v] = lse2(t0, l1 + log(1 - p) + log(q) + nu * (-dv));
    else if (has2)
      w->L0[v] = lse2(t0, log(1 - pn));
    else
      w->L0[v] = t0 + log(exp(t0 - t0)); /* log_sum_exp of one term */
    if (a == LAB_Z)
      w->L1[v] = NEG_INF;
    else if (a != LAB_Q || nu == 0 || p == 1)
      w->L1[v] = l1 + nu * (-dv);
    else
      w->L1[v] = lse2(l1 + nu * (-dv), log(1 - pn));
  }
}

/* __out_alpha_up__ for (node, a0) at site k */
static void out_alpha_up(const pmm_problem* P, const double* d, double nu, int k, site_ws* w, int node,
                         int a0, int* path) {
  int v = node, np = 0;
  const double lq = pseudo_log(qval(P, k, a0));
  int stopped = 0;
  while (P->parent[v] >= 0) {
    const int ws = sibling(P, v);
    const int wa = w->lab[ws];
    if (wa != LAB_Q && wa != a0) {
      oa_set(&w->oa, v, a0, w->A[v] - nu * d[v] + pseudo_log(1 - exp(-d[v])) + lq);
      stopped = 1;
      break;
    }
    if (np > 0 && oa_find(&w->oa, v, a0)) { /* value already known above: unwind from here */
      stopped = 1;
      break;
    }
    path[np++] = v;
    v = P->parent[v];
  }
  if (!stopped && !oa_find(&w->oa, v, a0))
    oa_set(&w->oa, v, a0, -nu * d[v] + pseudo_log(1 - exp(-d[v])) + lq);
  while (np > 0) {
    v = path[--np];
    const int u = P->parent[v];
    const int ws = sibling(P, v);
    const double B = *oa_find(&w->oa, u, a0) + nu * (-d[v]) + w->L1[ws];
    const double C = w->A[v] - nu * d[v] + pseudo_log(1 - exp(-d[v])) + lq;
    oa_set(&w->oa, v, a0, lse2(B, C));
  }
}

/* Estep_out_llh for site k (parents before children: descending index) */
static void outside_site(const pmm_problem* P, const double* d, double nu, int k, site_ws* w, int* path) {
  oa_clear(&w->oa);
  for (int v = P->n - 1; v >= 0; --v) {
    const double dv = d[v];
    if (P->parent[v] < 0) {
      w->A[v] = 0.0;
      if (nu * dv > 0) {
        w->X[v] = -nu * dv + log(1 - exp(-dv));
        w->out1[v] = log(1 - exp(-dv * nu));
      } else {
        w->X[v] = NEG_INF;
        w->out1[v] = NEG_INF;
      }
      w->out0[v] = -(1 + nu) * dv;
      continue;
    }
    const int u = P->parent[v];
    const int ws = sibling(P, v);
    w->out0[v] = w->out0[u] + w->L0[ws] - (1 + nu) * dv;
    const double A = w->out0[u] + w->L0[ws];
    w->A[v] = A;
    double X = dv > 0 ? (-nu * dv + log(1 - exp(-dv)) + A) : NEG_INF;
    const int wa = w->lab[ws];
    if (wa == LAB_Z) {
      w->out1[v] = nu * dv > 0 ? log(1 - exp(-dv * nu)) + A : NEG_INF;
    } else if (wa == LAB_Q) {
      X = lse2(X, w->X[u] + w->L1[ws] - nu * dv);
      const double pp = 1 - exp(-dv * nu);
      const double pl = pp > 0 ? log(pp) : NEG_INF;
      if (nu == 0)
        w->out1[v] = w->out1[u];
      else
        w->out1[v] = lse3(pl + A, pl + w->X[u] + w->L1[ws], w->out1[u]);
    } else {
      const int a0 = wa;
      if (!oa_find(&w->oa, u, a0)) out_alpha_up(P, d, nu, k, w, u, a0, path);
      const double oau = *oa_find(&w->oa, u, a0);
      const double B = oau + nu * (-dv) + w->L1[ws];
      const double q = qval(P, k, a0);
      const double C = (dv > 0 && q > 0) ? A - nu * dv + log(1 - exp(-dv)) + log(q) : NEG_INF;
      oa_set(&w->oa, v, a0, lse2(B, C));
      X = lse2(X, B);
      w->out1[v] = nu * dv > 0 ? log(1 - exp(-dv * nu)) + lse2(A, w->L1[ws] + oau) : NEG_INF;
    }
    w->X[v] = X;
  }
}

/* Estep_posterior for site k; adds S0..S4 to S[v*5 + j] and the phi sums */
static void posterior_site(const pmm_problem* P, const double* d, double nu, double phi, int k, site_ws* w,
                           double* S, double* Rt) {
  const int K = P->K;
  for (int v = P->n - 1; v >= 0; --v) {
    const double dv = d[v];
    const double full = w->L0[P->root_of[v]];
    const int a = w->lab[v];
    double vin0;
    int code = 0;
    if (P->c0[v] < 0) {
      code = P->codes[(size_t)P->leaf_row[v] * K + k];
      if (code == 0)
        vin0 = log(1 - phi);
      else
        vin0 = (code == CODE_MISSING && phi > 0) ? log(phi) : NEG_INF;
    } else {
      vin0 = w->L0[P->c0[v]] + w->L0[P->c1[v]];
    }
    const double p0 = vin0 + w->out0[v] - full;
    const double p1 = a == LAB_Q ? 0.0 + w->out1[v] - full : NEG_INF;
    w->post0[v] = p0;
    w->post1[v] = p1;
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0;
    if (a == LAB_Z) {
      s0 = 1.0;
    } else if (P->parent[v] < 0) {
      s0 = exp(vin0 + (1.0 + nu) * (-dv) - w->L0[v]);
      s2 = a != LAB_Q ? 0.0 : (1.0 - exp(-nu * dv)) / exp(w->L0[v]);
      s1 = 1.0 - s0 - s2;
    } else {
      const int u = P->parent[v];
      const double up0 = w->post0[u], up1 = w->post1[u];
      s0 = exp(up0 + vin0 + (1.0 + nu) * (-dv) - w->L0[v]);
      if (a != LAB_Q) {
        s2 = 0.0;
        s4 = 0.0;
      } else {
        s2 = exp(up0 - w->L0[v]) * (1.0 - exp(-nu * dv));
        const double t = 1.0 - exp(up0) - exp(up1);
        s4 = (t == 0 || (1.0 - exp(-nu * dv)) == 0) ? 0.0 : t * (1.0 - exp(-nu * dv)) / exp(w->L1[v]);
      }
      s1 = exp(up0) - s0 - s2;
      s3 = 1.0 - s0 - s1 - exp(p1);
    }
    double* o = S + (size_t)v * 5;
    o[0] += s0;
    o[1] += s1;
    o[2] += s2;
    o[3] += s3;
    o[4] += s4;
    if (P->c0[v] < 0 && code == CODE_MISSING) *Rt += 1 - exp(p1);
  }
}

static int is_root_count(const pmm_problem* P) {
  int r = 0;
  for (int v = 0; v < P->n; ++v) r += P->parent[v] < 0;
  return r;
}

int pmm_ref_threads(void) { return omp_get_max_threads(); }

/* lineage_llh: -sum over trees and sites of root.L0 */
/* per-thread partial sums are combined in thread order, so results do not
 * depend on which thread finishes first */
double pmm_ref_nll(const pmm_problem* P, const double* d, double nu, double phi) {
  const int T = omp_get_max_threads();
  double* part = (double*)calloc(T, sizeof(double));
#pragma omp parallel
  {
    const int t = omp_get_thread_num();
    site_ws w;
    ws_init(&w, P->n);
    double llh = 0.0;
#pragma omp for schedule(static)
    for (int k = 0; k < P->K; ++k) {
      inside_site(P, d, nu, phi, k, &w);
      for (int v = 0; v < P->n; ++v)
        if (P->parent[v] < 0) llh += w.L0[v];
    }
    part[t] = llh;
    ws_free(&w);
  }
  double llh = 0.0;
  for (int t = 0; t < T; ++t) llh += part[t];
  free(part);
  return -llh;
}

/* Estep: inside, outside and posterior of every site.  counts[v*5 + j] =
 * sum over sites of S_j on v's branch; *R = leaf entries other than '?';
 * *Rt = sum over '?' leaf entries of P(not silenced | D).  Returns the NLL. */
double pmm_ref_estep(const pmm_problem* P, const double* d, double nu, double phi, double* counts, double* R,
                     double* Rt) {
  const int n = P->n;
  const int T = omp_get_max_threads();
  double* acc = (double*)calloc((size_t)T * n * 5, sizeof(double));
  double* rt_acc = (double*)calloc(T, sizeof(double));
  double* llh_acc = (double*)calloc(T, sizeof(double));
#pragma omp parallel
  {
    const int t = omp_get_thread_num();
    site_ws w;
    ws_init(&w, n);
    int* path = (int*)malloc(sizeof(int) * n);
    double llh_t = 0.0;
#pragma omp for schedule(static)
    for (int k = 0; k < P->K; ++k) {
      inside_site(P, d, nu, phi, k, &w);
      for (int v = 0; v < n; ++v)
        if (P->parent[v] < 0) llh_t += w.L0[v];
      outside_site(P, d, nu, k, &w, path);
      posterior_site(P, d, nu, phi, k, &w, acc + (size_t)t * n * 5, &rt_acc[t]);
    }
    llh_acc[t] = llh_t;
    free(path);
    ws_free(&w);
  }
  double llh = 0.0;
  for (int t = 0; t < T; ++t) llh += llh_acc[t];
  free(llh_acc);
  memset(counts, 0, sizeof(double) * 5 * (size_t)n);
  double rt = 0.0;
  for (int t = 0; t < T; ++t) {
    const double* a = acc + (size_t)t * n * 5;
    for (size_t i = 0; i < (size_t)n * 5; ++i) counts[i] += a[i];
    rt += rt_acc[t];
  }
  double r = 0.0;
  for (int v = 0; v < n; ++v) {
    if (P->c0[v] >= 0) continue;
    const int16_t* row = P->codes + (size_t)P->leaf_row[v] * P->K;
    for (int k = 0; k < P->K; ++k) r += row[k] != CODE_MISSING;
  }
  *R = r;
  *Rt = rt;
  free(acc);
  free(rt_acc);
  return -llh;
}

/* ---------------- M-step ---------------- */
static double branch_grad(const double* s, double nu, int use_nu, int use_nu4, double d) {
  double g = -(nu + 1) * s[0] - nu * s[1] - nu * s[3] + s[1] / expm1(d);
  if (use_nu) g += s[2] * nu / expm1(nu * d);
  if (use_nu4) g += s[4] * nu / expm1(nu * d);
  return g;
}

static double solve_branch(const double* s, double nu, int use_nu, int use_nu4, double dmin, double dmax) {
  if (branch_grad(s, nu, use_nu, use_nu4, dmin) <= 0) return dmin;
  if (branch_grad(s, nu, use_nu, use_nu4, dmax) >= 0) return dmax;
  double lo = dmin, hi = dmax;
  for (;;) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) return mid;
    if (branch_grad(s, nu, use_nu, use_nu4, mid) > 0)
      lo = mid;
    else
      hi = mid;
  }
}

static double nu_grad(const double* rows, const double* d, int m, int use2, int use4, double nu) {
  double g = 0.0;
  for (int i = 0; i < m; ++i) {
    const double* s = rows + (size_t)i * 5;
    const double dv = d[i];
    g -= dv * (s[0] + s[1] + s[3]);
    if (use2) g += s[2] * dv / expm1(nu * dv);
    if (use4) g += s[4] * dv / expm1(nu * dv);
  }
  return g;
}

static double solve_nu(const double* rows, const double* d, int m) {
  int use2 = 0, use4 = 0;
  double s2 = 0, s4 = 0;
  for (int i = 0; i < m; ++i) {
    s2 += rows[(size_t)i * 5 + 2];
    s4 += rows[(size_t)i * 5 + 4];
  }
  use2 = s2 > 0;
  use4 = s4 > 0;
  if (nu_grad(rows, d, m, use2, use4, 1.0) >= 0) return 1.0;
  double lo = 0.0, hi = 1.0;
  for (;;) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) return mid;
    if (nu_grad(rows, d, m, use2, use4, mid) > 0)
      lo = mid;
    else
      hi = mid;
  }
}

/* Mstep from E-step sums: phi closed form, per-branch rows (clamp 1e-6,
 * renormalise to K), branch lengths, then nu.  Updates d (free branches),
 * *nu and *phi. */
void pmm_ref_mstep(const pmm_problem* P, const double* counts, double R, double Rt, double dmin, double dmax,
                   int opt_nu, int opt_phi, double* d, double* nu, double* phi) {
  const int n = P->n, K = P->K;
  if (opt_phi) {
    double ps = Rt / (R + Rt);
    if (fabs(ps) < P->phimin) ps = 0;
    *phi = ps;
  }
  int m = 0;
  for (int v = 0; v < n; ++v) m += P->free_[v] != 0;
  double* rows = (double*)malloc(sizeof(double) * 5 * (size_t)(m ? m : 1));
  int* idx = (int*)malloc(sizeof(int) * (m ? m : 1));
  double* dn = (double*)malloc(sizeof(double) * (m ? m : 1));
  int i = 0;
  double s2 = 0, s4 = 0;
  for (int v = 0; v < n; ++v) {
    if (!P->free_[v]) continue;
    double s[5], tot = 0;
    for (int j = 0; j < 5; ++j) {
      s[j] = counts[(size_t)v * 5 + j] > 1e-6 ? counts[(size_t)v * 5 + j] : 1e-6;
      tot += s[j];
    }
    for (int j = 0; j < 5; ++j) rows[(size_t)i * 5 + j] = s[j] / tot * K;
    s2 += rows[(size_t)i * 5 + 2];
    s4 += rows[(size_t)i * 5 + 4];
    idx[i++] = v;
  }
  const double x = *nu;
  const int use_nu = s2 > 0 && x > 1e-5;
  const int use_nu4 = s4 > 0 && x > 1e-5;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < m; ++j) dn[j] = solve_branch(rows + (size_t)j * 5, x, use_nu, use_nu4, dmin, dmax);
  if (opt_nu) *nu = solve_nu(rows, dn, m);
  for (int j = 0; j < m; ++j) d[idx[j]] = dn[j];
  free(rows);
  free(idx);
  free(dn);
}

/* EM_optimization: NLL, then E-step, M-step and NLL per iteration until the
 * relative change is below tol (or exactly fixed_iters iterations when
 * fixed_iters > 0).  Returns the iteration count; *nll_out the final NLL. */
int pmm_ref_em(const pmm_problem* P, double dmin, double dmax, int opt_nu, int opt_phi, int max_iter, double tol,
               int fixed_iters, double* d, double* nu, double* phi, double* nll_out) {
  double* counts = (double*)malloc(sizeof(double) * 5 * (size_t)P->n);
  double pre = -pmm_ref_nll(P, d, *nu, *phi), curr = pre;
  int it = 1;
  const int limit = fixed_iters > 0 ? fixed_iters : max_iter;
  (void)is_root_count;
  while (it <= limit) {
    double R, Rt;
    pmm_ref_estep(P, d, *nu, *phi, counts, &R, &Rt);
    pmm_ref_mstep(P, counts, R, Rt, dmin, dmax, opt_nu, opt_phi, d, nu, phi);
    curr = -pmm_ref_nll(P, d, *nu, *phi);
    if (fixed_iters <= 0 && fabs((curr - pre) / pre) < tol) break;
    pre = curr;
    it++;
  }
  free(counts);
  *nll_out = -curr;
  return it;
}
