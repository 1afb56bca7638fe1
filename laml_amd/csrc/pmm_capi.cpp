// C ABI of the PMM engine: forest construction (host validation + upload) and
// the stream-ordered entry points declared in include/laml_pmm.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <deque>
#include <tuple>
#include <limits>
#include <mutex>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "laml_pmm.h"
#include "pmm_kernels.hpp"

using laml::DevForest;
using laml::EmState;

namespace {

thread_local std::string g_err;
// Serialises the HIP graph capture of laml_em with every synchronous
// operation on the legacy default stream (the uploads and read-backs of the
// set-up calls): such an operation must not run while another thread
// captures, or it would have to depend on the capturing stream.
std::mutex g_legacy_mu;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define LAML_HIP(expr)                                                                            \
  do {                                                                                            \
    hipError_t e_ = (expr);                                                                       \
    if (e_ != hipSuccess) return fail(LAML_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

}  // namespace

struct EmGraphKey {
  const double *d, *nu, *phi, *nll, *trace;
  double tol;
  int32_t max_iter;
  uint64_t plan;
  bool operator==(const EmGraphKey& o) const {
    return d == o.d && nu == o.nu && phi == o.phi && nll == o.nll && trace == o.trace && tol == o.tol &&
           max_iter == o.max_iter && plan == o.plan;
  }
};

struct laml_forest {
  int device = 0;
  // captured EM step (replayed by laml_em) and the launch-plan version it
  // was captured for; every change of launch shapes, checkpointing or the
  // ultrametric model bumps the version
  bool use_graph = true;
  hipStream_t cap_stream = nullptr;
  hipGraphExec_t em_exec = nullptr;
  EmGraphKey em_key{};
  uint64_t plan_version = 0;
  void drop_graph() {
    if (em_exec) (void)hipGraphExecDestroy(em_exec);
    em_exec = nullptr;
  }
  int W = 64;
  int h_top = 1 << 30;  // first level handled by the fused top-level kernels
  int h_top_default = 1 << 30;
  int ckpt_L = 0;       // level-checkpoint segment length (0: all inside vectors stored)
  // per level: 64-site chunks one wavefront of the row kernels walks (W == 64)
  std::vector<int32_t> in_cpw, out_cpw;
  // pass mode: 0 = level-synchronous launches, 1 = tree walk (one launch),
  // 2 = hybrid (level launches below h_top, one tree walk above; default)
  int pass_mode = 2;
  // cherry fusion (the height-1 level formed and consumed by its parents):
  // requested (off by default; the autotuner tries it), possible for this
  // forest (row kernels, every tree with at least three leaves), and
  // whether the last inside pass ran with it
  bool fuse = false;
  bool cherry_ok = false;
  bool inside_fused = false;
  laml::CherrySplit split;  // per-level node lists of the fused passes
  // heavy-path chains for the levels below the split (requested:
  // LAML_CHAINS_OUTSIDE | LAML_CHAINS_INSIDE), the plans built so far by
  // split level, and whether every internal node's outside row holds the
  // last outside pass's state (outside chains keep heavy children's states
  // in registers)
  int chains = 0;
  // top of the chain region of each pass (levels below min(top, split));
  // 0: up to the split
  int chain_top_in = 5, chain_top_out = 10;
  // steps per outside pack (chains of one tree run back to back by one
  // wavefront; 1: a chain per wavefront)
  int chain_pack = 8;
  int64_t pack_waves = laml::kPackWaves;
  bool alpha_complete = true;
  std::deque<laml::ChainPlan> chain_plans;  // a deque: plans stay put when more are added
  std::vector<int32_t> h_heavy;  // heavy child of each node (-1: no internal child)
  std::vector<int32_t> h_child0, h_child1;
  std::vector<int32_t> h_tree_of, h_tree_prob, h_tree_site_base;
  laml::TreeWalk TW{};
  // host copies for re-layout of the inside vectors
  std::vector<int32_t> h_parent;
  std::vector<int32_t> h_K;         // padded row stride of each node
  std::vector<int32_t> h_Kt;        // sites of each node's tree
  std::vector<int64_t> h_row_off;   // device row offsets (padded)
  std::vector<int64_t> h_desc_off;  // row offsets of the unpadded description layout
  const int64_t* d_desc_off = nullptr;  // device copy (laml_posteriors_noalloc)
  int64_t leaf_sites = 0, int_sites = 0;
  // device address of the posterior staging array on this forest's device
  double* post_scratch = nullptr;
  void* beta_alloc = nullptr;
  void* bE_alloc = nullptr;
  int64_t beta_rows = 0;
  DevForest F{};
  std::vector<int32_t> level_off;  // host copy
  std::vector<int32_t> tree_nsites;
  std::vector<void*> allocs;
  int64_t leaf_rows = 0, int_rows = 0, total_sites = 0;
  // device scratch owned by the forest
  double* counts = nullptr;  // [N*5]
  double* rtilde = nullptr;  // [P]
  double* phi_buf = nullptr; // [P] staging when the caller passes phi
  EmState em{};
  int32_t* em_done = nullptr;
  // ultrametric model (owned separately so it can be replaced)
  bool ultra_on = false;
  laml::UltraDev U{};
  std::vector<void*> ultra_allocs;
  std::vector<int32_t> ultra_status_h;

  void free_ultra() {
    plan_version++;
    for (void* p : ultra_allocs) (void)hipFree(p);
    ultra_allocs.clear();
    ultra_on = false;
    U = laml::UltraDev{};
  }
  template <typename T>
  int ualloc(T** p, size_t n) {
    void* q = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc(&q, n * sizeof(T));
    if (e != hipSuccess) return fail(LAML_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    ultra_allocs.push_back(q);
    *p = static_cast<T*>(q);
    return LAML_OK;
  }
  template <typename T>
  int uupload(const T** dst, const T* src, size_t n) {
    T* p = nullptr;
    int rc = ualloc(&p, n);
    if (rc) return rc;
    if (n) LAML_HIP(hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice));
    *dst = p;
    return LAML_OK;
  }

  template <typename T>
  int alloc(T** p, size_t n) {
    void* q = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc(&q, n * sizeof(T));
    if (e != hipSuccess) return fail(LAML_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    allocs.push_back(q);
    *p = static_cast<T*>(q);
    return LAML_OK;
  }
  template <typename T>
  int upload(const T** dst, const T* src, size_t n) {
    T* p = nullptr;
    int rc = alloc(&p, n);
    if (rc) return rc;
    if (n) LAML_HIP(hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice));
    *dst = p;
    return LAML_OK;
  }
  ~laml_forest() {
    drop_graph();
    if (cap_stream) (void)hipStreamDestroy(cap_stream);
    for (void* p : allocs) (void)hipFree(p);
    for (void* p : ultra_allocs) (void)hipFree(p);
    if (beta_alloc) (void)hipFree(beta_alloc);
    if (bE_alloc) (void)hipFree(bE_alloc);
  }
};

namespace {

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

struct Derived {
  std::vector<int64_t> row_off;
  int64_t leaf_rows = 0, int_rows = 0;
  std::vector<int32_t> prob_tree_off, prob_node_off, prob_nodes, prob_int_off, prob_nfree, node_pos, tree_site_base;
  std::vector<int32_t> prob_blk_off;
  std::vector<double> prob_nobs_phi;
  int max_k = 0;
};

int validate(const laml_forest_desc* d, Derived& D) {
  std::vector<int64_t>& row_off = D.row_off;
  int64_t& leaf_rows = D.leaf_rows;
  int64_t& int_rows = D.int_rows;
  std::vector<int32_t>& prob_tree_off = D.prob_tree_off;
  std::vector<int32_t>& prob_node_off = D.prob_node_off;
  std::vector<int32_t>& prob_nodes = D.prob_nodes;
  std::vector<int32_t>& tree_site_base = D.tree_site_base;
  int& max_k = D.max_k;
  if (!d) return fail(LAML_EINVAL, "null description");
  const int N = d->num_nodes, n = d->num_leaves, T = d->num_trees, P = d->num_problems, H = d->num_levels;
  if (N <= 0 || n <= 0 || T <= 0 || P <= 0 || H <= 0 || n > N)
    return fail(LAML_EINVAL, "non-positive sizes");
  if (!d->level_off || !d->child0 || !d->child1 || !d->parent || !d->tree_of || !d->free_branch ||
      !d->tree_root || !d->tree_problem || !d->tree_nsites || !d->site_qoff || !d->prior || !d->leaf_codes ||
      !d->problem_nobs || !d->problem_phimin || !d->problem_flags)
    return fail(LAML_EINVAL, "null array in description");
  if (d->level_off[0] != 0 || d->level_off[H] != N || d->level_off[1] != n)
    return fail(LAML_EINVAL, "level_off must start at 0, end at num_nodes and have level 0 = leaves");
  std::vector<int32_t> level(N, -1);
  for (int h = 0; h < H; ++h) {
    if (d->level_off[h + 1] < d->level_off[h]) return fail(LAML_EINVAL, "level_off not monotone");
    for (int v = d->level_off[h]; v < d->level_off[h + 1]; ++v) level[v] = h;
  }
  for (int t = 0; t < T; ++t) {
    if (d->tree_problem[t] < 0 || d->tree_problem[t] >= P) return fail(LAML_EINVAL, "tree_problem out of range");
    if (t > 0 && d->tree_problem[t] < d->tree_problem[t - 1]) return fail(LAML_EINVAL, "tree_problem not sorted");
    if (d->tree_nsites[t] <= 0 || d->tree_nsites[t] > (1 << 24)) return fail(LAML_EINVAL, "bad tree_nsites");
    const int r = d->tree_root[t];
    if (r < 0 || r >= N || d->parent[r] != -1 || d->tree_of[r] != t) return fail(LAML_EINVAL, "bad tree_root");
  }
  // every problem owns >= 1 tree and all its trees have the same K
  prob_tree_off.assign(P + 1, 0);
  for (int t = 0; t < T; ++t) prob_tree_off[d->tree_problem[t] + 1]++;
  for (int p = 0; p < P; ++p) {
    if (prob_tree_off[p + 1] == 0) return fail(LAML_EINVAL, "problem without trees");
    prob_tree_off[p + 1] += prob_tree_off[p];
  }
  for (int t = 1; t < T; ++t)
    if (d->tree_problem[t] == d->tree_problem[t - 1] && d->tree_nsites[t] != d->tree_nsites[t - 1])
      return fail(LAML_EINVAL, "trees of one problem differ in site count");
  // topology
  for (int v = 0; v < N; ++v) {
    const int t = d->tree_of[v];
    if (t < 0 || t >= T) return fail(LAML_EINVAL, "tree_of out of range");
    const int c0 = d->child0[v], c1 = d->child1[v];
    if (v < n) {
      if (c0 != -1 || c1 != -1) return fail(LAML_EINVAL, "leaf with children");
    } else {
      if (c0 < 0 || c0 >= N || c1 < 0 || c1 >= N || c0 == c1) return fail(LAML_EINVAL, "bad children");
      if (level[c0] >= level[v] || level[c1] >= level[v]) return fail(LAML_EINVAL, "child not below parent level");
      if (d->parent[c0] != v || d->parent[c1] != v) return fail(LAML_EINVAL, "parent/child mismatch");
      if (d->tree_of[c0] != t || d->tree_of[c1] != t) return fail(LAML_EINVAL, "child in another tree");
    }
    const int pa = d->parent[v];
    if (pa == -1) {
      if (d->tree_root[t] != v) return fail(LAML_EINVAL, "second root in a tree");
    } else if (pa < n || pa >= N || (d->child0[pa] != v && d->child1[pa] != v)) {
      return fail(LAML_EINVAL, "bad parent");
    }
  }
  // rows and sites
  tree_site_base.assign(T + 1, 0);
  max_k = 0;
  for (int t = 0; t < T; ++t) {
    tree_site_base[t + 1] = tree_site_base[t] + d->tree_nsites[t];
    max_k = std::max(max_k, (int)d->tree_nsites[t]);
  }
  row_off.assign(N, 0);
  leaf_rows = 0;
  int_rows = 0;
  for (int v = 0; v < N; ++v) {
    const int K = d->tree_nsites[d->tree_of[v]];
    if (v < n) {
      row_off[v] = leaf_rows;
      leaf_rows += K;
    } else {
      row_off[v] = int_rows;
      int_rows += K;
    }
  }
  const int S = tree_site_base[T];
  if (d->site_qoff[0] != 0) return fail(LAML_EINVAL, "site_qoff[0] != 0");
  // kernels index the prior table with 32-bit byte offsets
  if (d->site_qoff[S] >= (1 << 28)) return fail(LAML_EINVAL, "prior table too large (>= 2^28 entries)");
  // alphabet size of every site
  std::vector<int32_t> alpha_size(std::max(S, 1));
  for (int s = 0; s < S; ++s) {
    alpha_size[s] = d->site_qoff[s + 1] - d->site_qoff[s];
    if (alpha_size[s] < 0) return fail(LAML_EINVAL, "site_qoff not monotone");
  }
  // leaf codes in range for their sites (-2 .. alphabet size), and the
  // entries observed as 0 or an edit per problem (phi gradient); one
  // branch-free pass over each row
  D.prob_nobs_phi.assign(P, 0.0);
  for (int v = 0; v < n; ++v) {
    const int t = d->tree_of[v];
    const int K = d->tree_nsites[t];
    const int16_t* c = d->leaf_codes + row_off[v];
    const int32_t* m = alpha_size.data() + tree_site_base[t];
    int bad = 0, obs = 0;
    for (int k = 0; k < K; ++k) {
      const int x = c[k];
      bad |= (x < -2) | (x > m[k]);
      obs += x >= 0;
    }
    if (bad) return fail(LAML_EINVAL, "leaf code out of range for its site");
    D.prob_nobs_phi[d->tree_problem[t]] += (double)obs;
  }
  // nodes grouped by problem
  prob_node_off.assign(P + 1, 0);
  for (int v = 0; v < N; ++v) prob_node_off[d->tree_problem[d->tree_of[v]] + 1]++;
  for (int p = 0; p < P; ++p) prob_node_off[p + 1] += prob_node_off[p];
  prob_nodes.assign(N, 0);
  std::vector<int32_t> fill(prob_node_off.begin(), prob_node_off.end() - 1);
  for (int v = 0; v < N; ++v) prob_nodes[fill[d->tree_problem[d->tree_of[v]]]++] = v;
  D.node_pos.assign(N, 0);
  for (int i = 0; i < N; ++i) D.node_pos[prob_nodes[i]] = i;
  D.prob_int_off.assign(P, 0);
  D.prob_nfree.assign(P, 0);
  for (int p = 0; p < P; ++p) {
    int i = prob_node_off[p];
    while (i < prob_node_off[p + 1] && prob_nodes[i] < n) ++i;
    D.prob_int_off[p] = i;
    for (int j = prob_node_off[p]; j < prob_node_off[p + 1]; ++j) D.prob_nfree[p] += d->free_branch[prob_nodes[j]] ? 1 : 0;
  }
  // nu-solve blocks: about 2048 branches per block, at most 64 per problem
  D.prob_blk_off.assign(P + 1, 0);
  for (int p = 0; p < P; ++p) {
    const int nb = prob_node_off[p + 1] - prob_node_off[p];
    const int blocks = std::max(1, std::min(64, (nb + 2047) / 2048));
    D.prob_blk_off[p + 1] = D.prob_blk_off[p] + blocks;
  }
  if (!(d->dmin > 0.0) || !(d->dmax > d->dmin)) return fail(LAML_EINVAL, "bad branch-length box");
  return LAML_OK;
}

}  // namespace

extern "C" {

const char* laml_last_error(void) { return g_err.c_str(); }

const char* laml_version(void) { return "laml-amd 0.1 (gfx950, fp64)"; }

int laml_forest_validate(const laml_forest_desc* d) {
  Derived D;
  return validate(d, D);
}

static int upload_node_recs(laml_forest* f, const std::vector<int64_t>& brow);

int laml_forest_create(const laml_forest_desc* d, int device, laml_forest** out) {
  if (!out) return fail(LAML_EINVAL, "null output pointer");
  *out = nullptr;
  Derived D;
  int rc = validate(d, D);
  if (rc) return rc;
  std::vector<int32_t>& prob_tree_off = D.prob_tree_off;
  std::vector<int32_t>& prob_node_off = D.prob_node_off;
  std::vector<int32_t>& prob_nodes = D.prob_nodes;
  std::vector<int32_t>& tree_site_base = D.tree_site_base;
  const int max_k = D.max_k;
  // Device rows are padded to a multiple of the segment width, so that the
  // slice of a row one wavefront touches starts on a cache-line boundary
  // (64 sites = 128 B of int16, 1 KiB of double2).
  const int Wd = laml::seg_width(max_k);
  const int Nn = d->num_nodes, nl = d->num_leaves;
  std::vector<int64_t> row_off(Nn, 0);
  std::vector<int32_t> stride(Nn, 0);
  int64_t leaf_rows = 0, int_rows = 0;
  for (int v = 0; v < Nn; ++v) {
    const int K = d->tree_nsites[d->tree_of[v]];
    stride[v] = (K + Wd - 1) / Wd * Wd;
    if (v < nl) {
      row_off[v] = leaf_rows;
      leaf_rows += stride[v];
    } else {
      row_off[v] = int_rows;
      int_rows += stride[v];
    }
  }
  // host-only work above (validation, row layout) runs outside the lock;
  // the device calls below are serialised with graph capture
  std::lock_guard<std::mutex> legacy_lock(g_legacy_mu);
  int ndev = 0;
  LAML_HIP(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(LAML_EINVAL, "device out of range");
  DeviceGuard guard(device);
  laml_forest* f = new laml_forest();
  f->device = device;
  {
    const char* env = getenv("LAML_EM_GRAPH");
    f->use_graph = !(env && env[0] == '0');
    // private capture stream (default priority)
    if (hipStreamCreateWithFlags(&f->cap_stream, hipStreamNonBlocking) != hipSuccess) {
      delete f;
      return fail(LAML_EDEVICE, "hipStreamCreate");
    }
  }
  f->W = laml::seg_width(max_k);
  f->level_off.assign(d->level_off, d->level_off + d->num_levels + 1);
  f->tree_nsites.assign(d->tree_nsites, d->tree_nsites + d->num_trees);
  // fuse the tail of small upper levels into one launch per pass
  {
    const char* env = getenv("LAML_TOP_MAX_NODES");
    const int top_max = env ? atoi(env) : 48;
    int ht = d->num_levels;
    while (ht - 1 >= 1 && d->level_off[ht] - d->level_off[ht - 1] <= top_max) --ht;
    f->h_top = ht;
    f->h_top_default = ht;
    const char* pm = getenv("LAML_PASS_MODE");
    if (pm && (atoi(pm) == 0 || atoi(pm) == 1)) f->pass_mode = atoi(pm);
  }
  f->leaf_rows = leaf_rows;
  f->int_rows = int_rows;
  const int N = d->num_nodes, T = d->num_trees, P = d->num_problems;
  const int S = tree_site_base[T];
  f->total_sites = S;
  DevForest& F = f->F;
  F.num_nodes = N;
  F.num_leaves = d->num_leaves;
  F.num_trees = T;
  F.num_problems = P;
  F.dmin = d->dmin;
  F.dmax = d->dmax;
  F.chunk_w = f->W;
  F.max_chunks = (max_k + f->W - 1) / f->W;
  F.cherry_end = d->num_leaves;
  {
    // stream the rows past the caches when one pass's inside + outside rows
    // (60 B per internal site) exceed the 256 MiB last-level cache
    const char* env = getenv("LAML_STREAM_ROWS");
    const double row_bytes = 60.0 * (double)int_rows;
    F.stream_rows = env ? (atoi(env) != 0) : (row_bytes > 256.0 * 1024 * 1024);
  }
  {
    // the height-1 level holds exactly the nodes whose children are both
    // leaves; fusing it needs every tree root above it and row kernels
    bool ok = d->num_levels >= 3 && f->W == 64;
    for (int t = 0; ok && t < d->num_trees; ++t) ok = d->tree_root[t] >= d->level_off[2];
    f->cherry_ok = ok;
    const char* env = getenv("LAML_FUSE_CHERRIES");
    f->fuse = env && atoi(env) != 0;
    if (ok) {
      // per level from 2 up: nodes without a cherry child, then with one
      const int H = d->num_levels;
      const int c_beg = d->level_off[1], c_end = d->level_off[2];
      auto is_cherry = [&](int c) { return c >= c_beg && c < c_end; };
      std::vector<int32_t> list;
      list.reserve(d->num_nodes - c_end);
      f->split.off.assign(2 * H + 1, 0);
      for (int h = 2; h < H; ++h) {
        f->split.off[2 * h] = (int32_t)list.size();
        for (int v = d->level_off[h]; v < d->level_off[h + 1]; ++v)
          if (!is_cherry(d->child0[v]) && !is_cherry(d->child1[v])) list.push_back(v);
        f->split.off[2 * h + 1] = (int32_t)list.size();
        for (int v = d->level_off[h]; v < d->level_off[h + 1]; ++v)
          if (is_cherry(d->child0[v]) || is_cherry(d->child1[v])) list.push_back(v);
      }
      f->split.off[2 * H] = (int32_t)list.size();
      for (int h = 0; h < 2; ++h) f->split.off[2 * h] = f->split.off[2 * h + 1] = 0;
      int rc2 = f->upload(&f->split.nodes, list.data(), list.size());
      if (rc2) {
        delete f;
        return rc2;
      }
    }
  }
  f->in_cpw = laml::default_cpw(f->level_off, f->W, F.max_chunks);
  f->out_cpw = f->in_cpw;
  // tree-walk plan: each tree's per-level node ranges and its (tree, chunk)
  // blocks, largest trees first
  std::vector<int32_t> tw_blk_tree, tw_blk_chunk, tw_lvl_base(d->num_trees + 1, 0), tw_lvl_beg, tw_lvl_cnt;
  {
    const int T0 = d->num_trees, H = d->num_levels;
    std::vector<int32_t> height(T0, 0), tsize(T0, 0);
    for (int h = 0; h < H; ++h)
      for (int v = d->level_off[h]; v < d->level_off[h + 1]; ++v) {
        height[d->tree_of[v]] = h + 1;
        tsize[d->tree_of[v]]++;
      }
    for (int t = 0; t < T0; ++t) tw_lvl_base[t + 1] = tw_lvl_base[t] + height[t];
    tw_lvl_beg.assign(tw_lvl_base[T0], 0);
    tw_lvl_cnt.assign(tw_lvl_base[T0], 0);
    for (int h = 0; h < H; ++h)
      for (int v = d->level_off[h]; v < d->level_off[h + 1]; ++v) {
        const int t = d->tree_of[v];
        const int i = tw_lvl_base[t] + h;
        if (tw_lvl_cnt[i] == 0) tw_lvl_beg[i] = v;
        tw_lvl_cnt[i]++;
      }
    std::vector<int32_t> order(T0);
    for (int t = 0; t < T0; ++t) order[t] = t;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
      return (int64_t)tsize[a] * d->tree_nsites[a] > (int64_t)tsize[b] * d->tree_nsites[b];
    });
    for (int t : order) {
      const int nch = (d->tree_nsites[t] + f->W - 1) / f->W;
      for (int c = 0; c < nch; ++c) {
        tw_blk_tree.push_back(t);
        tw_blk_chunk.push_back(c);
      }
    }
  }
  std::vector<int32_t> tree_prob(d->tree_problem, d->tree_problem + T);
#define UP(dst, src, n) \
  if ((rc = f->upload(&F.dst, src, (size_t)(n)))) { delete f; return rc; }
  UP(level_off, d->level_off, d->num_levels + 1);
  UP(child0, d->child0, N);
  UP(child1, d->child1, N);
  UP(tree_of, d->tree_of, N);
  UP(row_off, row_off.data(), N);
  UP(free_branch, d->free_branch, N);
  UP(tree_root, d->tree_root, T);
  UP(tree_prob, tree_prob.data(), T);
  UP(tree_nsites, d->tree_nsites, T);
  UP(tree_site_base, tree_site_base.data(), T);
  UP(prob_tree_off, prob_tree_off.data(), P + 1);
  UP(prob_node_off, prob_node_off.data(), P + 1);
  UP(prob_nodes, prob_nodes.data(), N);
  UP(prob_int_off, D.prob_int_off.data(), P);
  UP(prob_nfree, D.prob_nfree.data(), P);
  UP(node_pos, D.node_pos.data(), N);
  UP(prob_blk_off, D.prob_blk_off.data(), P + 1);
  F.num_nu_blocks = D.prob_blk_off[P];
  F.max_prob_nodes = 0;
  for (int p = 0; p < P; ++p)
    F.max_prob_nodes = std::max(F.max_prob_nodes, D.prob_node_off[p + 1] - D.prob_node_off[p]);
  UP(prob_flags, d->problem_flags, P);
  UP(prob_nobs, d->problem_nobs, P);
  UP(prob_phimin, d->problem_phimin, P);
  UP(prob_nobs_phi, D.prob_nobs_phi.data(), P);
  UP(site_qoff, d->site_qoff, S + 1);
  {
    std::vector<double> qs(std::max(S, 1), 1.0);
    if (d->site_qsum) std::copy(d->site_qsum, d->site_qsum + S, qs.begin());
    UP(site_qsum, qs.data(), qs.size());
  }
  {
    // two entries of padding: the outside chain kernel gathers 16 bytes at a
    // prior entry (entry 0 for labels that name no edit state)
    std::vector<double> pr((size_t)d->site_qoff[S] + 2, 0.0);
    std::copy(d->prior, d->prior + d->site_qoff[S], pr.begin());
    UP(prior, pr.data(), pr.size());
  }
#undef UP
  {
    // padded leaf rows; filled from the caller's packed codes below
    int16_t* codes_dev = nullptr;
    if ((rc = f->alloc(&codes_dev, (size_t)leaf_rows))) {
      delete f;
      return rc;
    }
    F.codes = codes_dev;
  }
  {
    const int32_t *bt = nullptr, *bc = nullptr, *lbase = nullptr, *lbeg = nullptr, *lcnt = nullptr;
    if ((rc = f->upload(&bt, tw_blk_tree.data(), tw_blk_tree.size())) ||
        (rc = f->upload(&bc, tw_blk_chunk.data(), tw_blk_chunk.size())) ||
        (rc = f->upload(&lbase, tw_lvl_base.data(), tw_lvl_base.size())) ||
        (rc = f->upload(&lbeg, tw_lvl_beg.data(), tw_lvl_beg.size())) ||
        (rc = f->upload(&lcnt, tw_lvl_cnt.data(), tw_lvl_cnt.size()))) {
      delete f;
      return rc;
    }
    f->TW.num_blocks = (int32_t)tw_blk_tree.size();
    f->TW.blk_tree = bt;
    f->TW.blk_chunk = bc;
    f->TW.lvl_base = lbase;
    f->TW.lvl_beg = lbeg;
    f->TW.lvl_cnt = lcnt;
  }
#define AL(dst, type, n) \
  if ((rc = f->alloc<type>(&dst, (size_t)(n)))) { delete f; return rc; }
  int16_t* labels;
  laml::Coef* coef;
  double2* beta;
  int32_t *bE, *aE, *LE;
  double4* alpha;
  int16_t* aG;
  double *iLm, *site_llh, *rt_part, *cpart, *nu_a, *nu_b, *nu_d, *nu_part;
  laml::NuState* nu_state;
  AL(labels, int16_t, int_rows);
  // prior-weight rows: read by the chain kernels only, which need 64-site
  // segments
  double *qleaf = nullptr, *qint = nullptr;
  if (f->W == 64) {
    AL(qleaf, double, leaf_rows);
    AL(qint, double, int_rows);
  }
  AL(coef, laml::Coef, N);
  {
    void* q = nullptr;
    if (hipMalloc(&q, sizeof(double2) * std::max<int64_t>(int_rows, 1)) != hipSuccess) {
      delete f;
      return fail(LAML_ENOMEM, "hipMalloc beta");
    }
    f->beta_alloc = q;
    beta = static_cast<double2*>(q);
    if (hipMalloc(&q, sizeof(int32_t) * std::max<int64_t>(int_rows, 1)) != hipSuccess) {
      delete f;
      return fail(LAML_ENOMEM, "hipMalloc bE");
    }
    f->bE_alloc = q;
    bE = static_cast<int32_t*>(q);
    f->beta_rows = int_rows;
  }
  AL(alpha, double4, int_rows);
  AL(aE, int32_t, int_rows);
  AL(aG, int16_t, int_rows);
  AL(iLm, double, S);
  AL(LE, int32_t, S);
  AL(site_llh, double, S);
  AL(rt_part, double, (size_t)N * F.max_chunks);
  AL(cpart, double, (size_t)N * F.max_chunks * 5);
  AL(nu_a, double, N);
  AL(nu_b, double, N);
  AL(nu_d, double, N);
  AL(nu_part, double, (size_t)F.num_nu_blocks * 3);
  AL(nu_state, laml::NuState, P);
  AL(f->counts, double, (size_t)N * 5);
  AL(f->rtilde, double, P);
  AL(f->phi_buf, double, P);
  AL(f->em.done, int32_t, P);
  AL(f->em.iters, int32_t, P);
  AL(f->em.step, int32_t, P);
  AL(f->em.pre, double, P);
#undef AL
  F.labels = labels;
  F.qleaf = qleaf;
  F.qint = qint;
  F.brow = F.row_off;
  f->h_parent.assign(d->parent, d->parent + N);
  f->h_tree_of.assign(d->tree_of, d->tree_of + N);
  f->h_tree_prob = tree_prob;
  f->h_tree_site_base.assign(tree_site_base.begin(), tree_site_base.end());
  {
    // heavy child: the internal child with more leaves below it (children
    // precede their parents in level order)
    std::vector<int32_t> nl(N, 1);
    f->h_heavy.assign(N, -1);
    f->h_child0.assign(d->child0, d->child0 + N);
    f->h_child1.assign(d->child1, d->child1 + N);
    for (int v = d->num_leaves; v < N; ++v) {
      const int a = d->child0[v], b = d->child1[v];
      nl[v] = nl[a] + nl[b];
      const bool ia = a >= d->num_leaves, ib = b >= d->num_leaves;
      if (ia && ib) f->h_heavy[v] = nl[b] > nl[a] ? b : a;
      else if (ia) f->h_heavy[v] = a;
      else if (ib) f->h_heavy[v] = b;
    }
    // default: outside chains below level 10 for forests with enough
    // (internal node, 64-site chunk) pairs to keep the chain rounds busy
    // (32k; smaller forests are latency-bound and keep level launches)
    const char* penv = getenv("LAML_CHAIN_PACK");
    if (penv) f->chain_pack = std::max(1, atoi(penv));
    const char* wenv = getenv("LAML_PACK_WAVES");
    if (wenv) f->pack_waves = std::max(1, atoi(wenv));
    const char* env = getenv("LAML_CHAINS");
    const int64_t units = (int64_t)(N - d->num_leaves) * ((max_k + 63) / 64);
    f->chains = env ? (atoi(env) & (LAML_CHAINS_OUTSIDE | LAML_CHAINS_INSIDE))
                    : (f->W == 64 && units >= 32768 ? LAML_CHAINS_OUTSIDE : 0);
  }
  f->h_row_off = row_off;
  if ((rc = upload_node_recs(f, row_off))) {
    delete f;
    return rc;
  }
  f->h_K = stride;
  f->h_desc_off = D.row_off;
  f->leaf_sites = D.leaf_rows;
  f->int_sites = D.int_rows;
  if ((rc = f->upload(&f->d_desc_off, f->h_desc_off.data(), (size_t)N))) {
    delete f;
    return rc;
  }
  // the staging array of laml_posteriors_noalloc: its address on this device
  // (looked up once per device, with the device current)
  if (const int se = laml::post_scratch(device, &f->post_scratch)) {
    delete f;
    return fail(LAML_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") +
                                  hipGetErrorString(static_cast<hipError_t>(se)));
  }
  f->h_Kt.resize(N);
  for (int v = 0; v < N; ++v) f->h_Kt[v] = d->tree_nsites[d->tree_of[v]];
  F.coef = coef;
  F.beta = beta;
  F.bE = bE;
  F.alpha = alpha;
  F.aE = aE;
  F.aG = aG;
  F.iLm = iLm;
  F.LE = LE;
  F.site_llh = site_llh;
  F.rt_part = rt_part;
  F.cpart = cpart;
  F.nu_a = nu_a;
  F.nu_b = nu_b;
  F.nu_d = nu_d;
  F.nu_part = nu_part;
  F.nu_state = nu_state;
  F.phi = nullptr;
  if (hipMemset(rt_part, 0, sizeof(double) * N * F.max_chunks) != hipSuccess ||
      hipMemset(cpart, 0, sizeof(double) * N * F.max_chunks * 5) != hipSuccess ||
      hipMemset(nu_state, 0, sizeof(laml::NuState) * P) != hipSuccess) {
    delete f;
    return fail(LAML_EDEVICE, "hipMemset failed");
  }
  {
    // The caller's codes go up as they are (rows of K sites back to back)
    // with their row offsets, staged in the inside-row buffer (unused until
    // the first pass, and 8x the leaf codes' bytes or more in a binary
    // tree), and a kernel writes them into the padded rows.  The staging
    // bytes are cleared afterwards (every pass writes an inside row before
    // reading it; the clear keeps code bytes out of the rows' pad sites).
    const size_t off_bytes = sizeof(int64_t) * (size_t)nl;
    const size_t need = off_bytes + sizeof(int16_t) * (size_t)D.leaf_rows;
    char* stage = static_cast<char*>(f->beta_alloc);
    if (sizeof(double2) * (size_t)std::max<int64_t>(int_rows, 1) < need) {
      if ((rc = f->alloc(&stage, need))) {
        delete f;
        return rc;
      }
    }
    if (hipMemcpy(stage, D.row_off.data(), off_bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(stage + off_bytes, d->leaf_codes, need - off_bytes, hipMemcpyHostToDevice) != hipSuccess) {
      delete f;
      return fail(LAML_EDEVICE, "leaf code upload failed");
    }
    laml::run_pad_codes(F, reinterpret_cast<const int16_t*>(stage + off_bytes),
                        reinterpret_cast<const int64_t*>(stage), nullptr);
    if (hipMemsetAsync(stage, 0, need, nullptr) != hipSuccess) {
      delete f;
      return fail(LAML_EDEVICE, "hipMemsetAsync failed");
    }
  }
  laml::run_labels(F, f->level_off, f->W, nullptr);
  if (qleaf) laml::run_qrows(F, qleaf, qint, nullptr);
  // every set-up copy and kernel above ran on the legacy default stream:
  // waiting for it alone leaves other threads' streams (a server's other
  // solves) running
  hipError_t e = hipStreamSynchronize(nullptr);
  if (e != hipSuccess) {
    delete f;
    return fail(LAML_EDEVICE, std::string("label kernels: ") + hipGetErrorString(e));
  }
  e = hipGetLastError();
  if (e != hipSuccess) {
    delete f;
    return fail(LAML_EDEVICE, std::string("label kernels: ") + hipGetErrorString(e));
  }
  *out = f;
  return LAML_OK;
}

void laml_forest_destroy(laml_forest* f) {
  std::lock_guard<std::mutex> legacy_lock(g_legacy_mu);
  if (!f) return;
  DeviceGuard guard(f->device);
  (void)hipDeviceSynchronize();
  delete f;
}

int laml_forest_set_flags(laml_forest* f, const uint32_t* flags_host) {
  std::lock_guard<std::mutex> legacy_lock(g_legacy_mu);
  if (!f || !flags_host) return fail(LAML_EINVAL, "null argument");
  DeviceGuard guard(f->device);
  LAML_HIP(hipMemcpy(const_cast<uint32_t*>(f->F.prob_flags), flags_host, sizeof(uint32_t) * f->F.num_problems,
                     hipMemcpyHostToDevice));
  return LAML_OK;
}

int laml_forest_labels(laml_forest* f, int16_t* labels_host, int64_t capacity) {
  std::lock_guard<std::mutex> legacy_lock(g_legacy_mu);
  if (!f || !labels_host) return fail(LAML_EINVAL, "null argument");
  DeviceGuard guard(f->device);
  const int64_t need = f->leaf_sites + f->int_sites;
  if (capacity < need) return fail(LAML_EINVAL, "labels capacity too small");
  std::vector<int16_t> codes(f->leaf_rows), lab(f->int_rows);
  LAML_HIP(hipMemcpy(codes.data(), f->F.codes, sizeof(int16_t) * f->leaf_rows, hipMemcpyDeviceToHost));
  LAML_HIP(hipMemcpy(lab.data(), f->F.labels, sizeof(int16_t) * f->int_rows, hipMemcpyDeviceToHost));
  const int N = f->F.num_nodes, n = f->F.num_leaves;
  for (int v = 0; v < N; ++v) {
    const int K = f->h_Kt[v];
    int16_t* out = labels_host + (v < n ? 0 : f->leaf_sites) + f->h_desc_off[v];
    if (v < n) {
      for (int k = 0; k < K; ++k) {
        const int c = codes[f->h_row_off[v] + k];
        out[k] = (int16_t)(c > 0 ? c : (c == 0 ? 0 : -1));
      }
    } else {
      std::memcpy(out, lab.data() + f->h_row_off[v], sizeof(int16_t) * K);
    }
  }
  return LAML_OK;
}

int laml_forest_set_checkpoint(laml_forest* f, int32_t levels) {
  std::lock_guard<std::mutex> legacy_lock(g_legacy_mu);
  if (!f) return fail(LAML_EINVAL, "null forest");
  if (levels < 0) return fail(LAML_EINVAL, "negative checkpoint length");
  if (levels > 0 && f->pass_mode == LAML_PASS_TREES)
    return fail(LAML_ESTATE, "level checkpointing needs the level pass mode");
  // the checkpointed outside pass walks level blocks: the hybrid mode's walk
  // over the top levels gives way to level launches
  if (levels > 0 && f->pass_mode == LAML_PASS_HYBRID) f->pass_mode = LAML_PASS_LEVELS;
  DeviceGuard guard(f->device);
  LAML_HIP(hipDeviceSynchronize());
  const int N = f->F.num_nodes, n = f->F.num_leaves;
  const int H = (int)f->level_off.size() - 1;
  std::vector<int64_t> brow(N, 0);
  int64_t rows = 0;
  if (levels == 0) {
    brow = f->h_row_off;
    for (int v = n; v < N; ++v) rows += f->h_K[v];
    f->h_top = f->h_top_default;
  } else {
    std::vector<int32_t> lev(N, 0);
    for (int h = 0; h < H; ++h)
      for (int v = f->level_off[h]; v < f->level_off[h + 1]; ++v) lev[v] = h;
    auto seg = [&](int v) { return (lev[v] - 1) / levels; };
    const int nseg = (H - 1 + levels - 1) / levels;
    std::vector<char> exported(N, 0);
    for (int v = n; v < N; ++v) {
      const int pa = f->h_parent[v];
      exported[v] = (pa < 0 || seg(pa) != seg(v)) ? 1 : 0;
    }
    for (int v = n; v < N; ++v)
      if (exported[v]) {
        brow[v] = rows;
        rows += f->h_K[v];
      }
    const int64_t base = rows;
    std::vector<int64_t> seg_rows(std::max(nseg, 1), 0);
    int64_t work = 0;
    for (int v = n; v < N; ++v)
      if (!exported[v]) {
        const int sg = seg(v);
        brow[v] = base + seg_rows[sg];
        seg_rows[sg] += f->h_K[v];
        work = std::max(work, seg_rows[sg]);
      }
    rows = base + work;
    f->h_top = H;  // the fused top-level kernels do not cross segments
  }
  // new inside-vector storage
  void* nb = nullptr;
  void* ne = nullptr;
  if (hipMalloc(&nb, sizeof(double2) * std::max<int64_t>(rows, 1)) != hipSuccess)
    return fail(LAML_ENOMEM, "hipMalloc beta");
  if (hipMalloc(&ne, sizeof(int32_t) * std::max<int64_t>(rows, 1)) != hipSuccess) {
    (void)hipFree(nb);
    return fail(LAML_ENOMEM, "hipMalloc bE");
  }
  int64_t* db = nullptr;
  int rc = f->alloc(&db, N);
  if (rc) {
    (void)hipFree(nb);
    (void)hipFree(ne);
    return rc;
  }
  LAML_HIP(hipMemcpy(db, brow.data(), sizeof(int64_t) * N, hipMemcpyHostToDevice));
  if ((rc = upload_node_recs(f, brow))) {
    (void)hipFree(nb);
    (void)hipFree(ne);
    return rc;
  }
  if (f->beta_alloc) (void)hipFree(f->beta_alloc);
  if (f->bE_alloc) (void)hipFree(f->bE_alloc);
  f->beta_alloc = nb;
  f->bE_alloc = ne;
  f->beta_rows = rows;
  f->F.beta = static_cast<double2*>(nb);
  f->F.bE = static_cast<int32_t*>(ne);
  f->F.brow = db;
  f->ckpt_L = levels;
  f->plan_version++;
  return LAML_OK;
}

int64_t laml_forest_inside_rows(laml_forest* f) { return f ? f->beta_rows : -1; }

int laml_forest_get_shapes(laml_forest* f, int32_t* inside_cpw, int32_t* outside_cpw, int32_t* top_level) {
  if (!f) return fail(LAML_EINVAL, "null forest");
  const int H = (int)f->level_off.size() - 1;
  for (int h = 0; h < H; ++h) {
    if (inside_cpw) inside_cpw[h] = f->in_cpw.empty() ? 0 : f->in_cpw[h];
    if (outside_cpw) outside_cpw[h] = f->out_cpw.empty() ? 0 : f->out_cpw[h];
  }
  if (top_level) *top_level = f->h_top_default;
  return LAML_OK;
}

int laml_forest_set_pass_mode(laml_forest* f, int32_t mode) {
  if (!f) return fail(LAML_EINVAL, "null forest");
  if (mode != LAML_PASS_LEVELS && mode != LAML_PASS_TREES && mode != LAML_PASS_HYBRID)
    return fail(LAML_EINVAL, "unknown pass mode");
  if (mode != LAML_PASS_LEVELS && f->ckpt_L > 0)
    return fail(LAML_ESTATE, "the tree walk needs full inside storage (level checkpointing is on)");
  f->pass_mode = mode;
  f->plan_version++;
  return LAML_OK;
}

int32_t laml_forest_pass_mode(laml_forest* f) { return f ? f->pass_mode : -1; }

int laml_forest_set_fusion(laml_forest* f, int32_t on) {
  if (!f) return fail(LAML_EINVAL, "null forest");
  f->fuse = on != 0;
  f->plan_version++;
  return LAML_OK;
}

static bool fuse_now(const laml_forest* f);
int32_t laml_forest_fusion(laml_forest* f) { return f ? (fuse_now(f) ? 1 : 0) : -1; }

int laml_forest_set_chains(laml_forest* f, int32_t which, int32_t inside_top, int32_t outside_top) {
  if (!f) return fail(LAML_EINVAL, "null forest");
  if (which & ~(LAML_CHAINS_OUTSIDE | LAML_CHAINS_INSIDE)) return fail(LAML_EINVAL, "unknown chain flags");
  if (inside_top < 0 || outside_top < 0) return fail(LAML_EINVAL, "negative chain region top");
  f->chains = which;
  f->chain_top_in = inside_top;
  f->chain_top_out = outside_top;
  f->plan_version++;
  return LAML_OK;
}

int laml_forest_get_chains(laml_forest* f, int32_t* which, int32_t* inside_top, int32_t* outside_top) {
  if (!f || !which || !inside_top || !outside_top) return fail(LAML_EINVAL, "null argument");
  *which = f->chains;
  *inside_top = f->chain_top_in;
  *outside_top = f->chain_top_out;
  return LAML_OK;
}

static int chains_now(const laml_forest* f);
int32_t laml_forest_chains(laml_forest* f) { return f ? chains_now(f) : -1; }

int laml_forest_set_shapes(laml_forest* f, const int32_t* inside_cpw, const int32_t* outside_cpw, int32_t top_level) {
  if (!f) return fail(LAML_EINVAL, "null forest");
  const int H = (int)f->level_off.size() - 1;
  if (top_level < 1 || top_level > H) return fail(LAML_EINVAL, "top_level out of range [1, num_levels]");
  if (f->W == 64) {
    for (int h = 1; h < H; ++h) {
      if ((inside_cpw && (inside_cpw[h] < 1 || inside_cpw[h] > f->F.max_chunks)) ||
          (outside_cpw && (outside_cpw[h] < 1 || outside_cpw[h] > f->F.max_chunks)))
        return fail(LAML_EINVAL, "chunks per wavefront out of range [1, max_chunks]");
    }
    for (int h = 1; h < H; ++h) {
      if (inside_cpw) f->in_cpw[h] = inside_cpw[h];
      if (outside_cpw) f->out_cpw[h] = outside_cpw[h];
    }
  }
  f->h_top_default = top_level;
  f->plan_version++;
  if (f->ckpt_L == 0) f->h_top = top_level;
  return LAML_OK;
}

int laml_forest_set_ultrametric(laml_forest* f, const laml_ultra_desc* u) {
  std::lock_guard<std::mutex> legacy_lock(g_legacy_mu);
  if (!f) return fail(LAML_EINVAL, "null forest");
  DeviceGuard guard(f->device);
  LAML_HIP(hipDeviceSynchronize());
  f->free_ultra();
  if (!u) return LAML_OK;
  const int N = f->F.num_nodes, P = f->F.num_problems, V = u->num_classes, H = u->num_levels;
  if (V < 0 || H < 0 || (V > 0 && H <= 0)) return fail(LAML_EINVAL, "bad class counts");
  if (!u->level_problem_off || !u->up_cls || !u->dn_cls || !u->up_off || !u->dn_off || !u->cls_parent ||
      !u->cls_pedge || !u->inc_off || !u->inc_branch || !u->inc_sign || !u->problem_has_free)
    return fail(LAML_EINVAL, "null array in ultrametric description");
  const int64_t LP = (int64_t)H * P + 1;
  if (u->level_problem_off[0] != 0 || u->level_problem_off[LP - 1] != V)
    return fail(LAML_EINVAL, "level_problem_off must span the classes");
  for (int64_t i = 0; i + 1 < LP; ++i)
    if (u->level_problem_off[i + 1] < u->level_problem_off[i]) return fail(LAML_EINVAL, "level_problem_off not monotone");
  for (int v = 0; v < N; ++v) {
    if (u->up_cls[v] < -1 || u->up_cls[v] >= V || u->dn_cls[v] < -1 || u->dn_cls[v] >= V)
      return fail(LAML_EINVAL, "branch class out of range");
  }
  if (u->inc_off[0] != 0) return fail(LAML_EINVAL, "inc_off[0] != 0");
  for (int c = 0; c < V; ++c) {
    if (u->inc_off[c + 1] <= u->inc_off[c]) return fail(LAML_EINVAL, "class without incident free branch");
    if (u->cls_parent[c] < -1 || u->cls_parent[c] >= V) return fail(LAML_EINVAL, "cls_parent out of range");
    if (u->cls_parent[c] >= 0 && (u->cls_pedge[c] < 0 || u->cls_pedge[c] >= N))
      return fail(LAML_EINVAL, "cls_pedge out of range");
  }
  const int NI = u->inc_off[V];
  for (int j = 0; j < NI; ++j)
    if (u->inc_branch[j] < 0 || u->inc_branch[j] >= N) return fail(LAML_EINVAL, "inc_branch out of range");
  laml::UltraDev& U = f->U;
  U.num_classes = V;
  U.num_levels = H;
  int rc;
#define UU(dst, src, n) \
  if ((rc = f->uupload(&U.dst, src, (size_t)(n)))) { f->free_ultra(); return rc; }
  UU(plo, u->level_problem_off, LP);
  UU(up_cls, u->up_cls, N);
  UU(dn_cls, u->dn_cls, N);
  UU(up_off, u->up_off, N);
  UU(dn_off, u->dn_off, N);
  UU(cls_parent, u->cls_parent, V);
  UU(cls_pedge, u->cls_pedge, V);
  UU(inc_off, u->inc_off, V + 1);
  UU(inc_branch, u->inc_branch, NI);
  UU(inc_sign, u->inc_sign, NI);
  UU(has_free, u->problem_has_free, P);
#undef UU
#define UA(dst, type, n) \
  if ((rc = f->ualloc<type>(&U.dst, (size_t)(n)))) { f->free_ultra(); return rc; }
  UA(t, double, V);
  UA(dt, double, V);
  UA(Dd, double, V);
  UA(Gg, double, V);
  UA(r, double, V);
  UA(g, double, V);
  UA(gp, double, N);
  UA(w, double, N);
  UA(bA, double, N);
  UA(bS1, double, N);
  UA(bB, double, N);
  UA(status, int32_t, P);
#undef UA
  LAML_HIP(hipMemset(U.status, 0, sizeof(int32_t) * P));
  LAML_HIP(hipMemset(U.t, 0, sizeof(double) * (V > 0 ? V : 1)));
  LAML_HIP(hipMemset(U.dt, 0, sizeof(double) * (V > 0 ? V : 1)));
  f->ultra_on = true;
  f->plan_version++;
  return LAML_OK;
}

int laml_forest_set_heights(laml_forest* f, const double* heights_host) {
  std::lock_guard<std::mutex> legacy_lock(g_legacy_mu);
  if (!f || !heights_host) return fail(LAML_EINVAL, "null argument");
  if (!f->ultra_on) return fail(LAML_ESTATE, "no ultrametric model set");
  DeviceGuard guard(f->device);
  if (f->U.num_classes > 0)
    LAML_HIP(hipMemcpy(f->U.t, heights_host, sizeof(double) * f->U.num_classes, hipMemcpyHostToDevice));
  return LAML_OK;
}

int laml_forest_ultra_status(laml_forest* f, int32_t* status_host) {
  std::lock_guard<std::mutex> legacy_lock(g_legacy_mu);
  if (!f || !status_host) return fail(LAML_EINVAL, "null argument");
  if (!f->ultra_on) return fail(LAML_ESTATE, "no ultrametric model set");
  DeviceGuard guard(f->device);
  LAML_HIP(hipMemcpy(status_host, f->U.status, sizeof(int32_t) * f->F.num_problems, hipMemcpyDeviceToHost));
  return LAML_OK;
}

int laml_forest_mstep_status(laml_forest* f, int32_t* status_host) {
  std::lock_guard<std::mutex> legacy_lock(g_legacy_mu);
  if (!f || !status_host) return fail(LAML_EINVAL, "null argument");
  DeviceGuard guard(f->device);
  const int P = f->F.num_problems;
  std::vector<laml::NuState> st(P);
  LAML_HIP(hipMemcpy(st.data(), f->F.nu_state, sizeof(laml::NuState) * P, hipMemcpyDeviceToHost));
  std::vector<int32_t> us;
  if (f->ultra_on) {
    us.resize(P);
    LAML_HIP(hipMemcpy(us.data(), f->U.status, sizeof(int32_t) * P, hipMemcpyDeviceToHost));
  }
  for (int p = 0; p < P; ++p) {
    int32_t s = st[p].status & 3;
    // a singular Newton system of the ultrametric branch solve is a failed
    // branch solve; a stalled line search ends at the best point found
    if (f->ultra_on && us[p] == 2) s |= LAML_MSTEP_FAILED_D;
    status_host[p] = s;
  }
  return LAML_OK;
}

static int check_launch() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(LAML_EDEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
  return LAML_OK;
}

// uploads the per-node records of the inside row kernel for inside rows at
// brow (rows of the stored inside vectors)
static int upload_node_recs(laml_forest* f, const std::vector<int64_t>& brow) {
  const int N = f->F.num_nodes, n = f->F.num_leaves;
  std::vector<laml::NodeRec> rec(std::max(N - n, 1));
  for (int v = n; v < N; ++v) {
    laml::NodeRec& r = rec[v - n];
    const int t = f->h_tree_of[v];
    r.ch0 = f->h_child0[v];
    r.ch1 = f->h_child1[v];
    r.p = f->h_tree_prob[t];
    r.K = f->tree_nsites[t];
    r.sb = f->h_tree_site_base[t];
    r.pad = 0;
    r.ro = f->h_row_off[v];
    r.bro = brow[v];
    r.ro0 = f->h_row_off[r.ch0];
    r.bro0 = brow[r.ch0];
    r.ro1 = f->h_row_off[r.ch1];
    r.bro1 = brow[r.ch1];
  }
  const laml::NodeRec* d = nullptr;
  int rc = f->upload(&d, rec.data(), rec.size());
  if (rc) return rc;
  f->F.nrec = d;
  return LAML_OK;
}

// whether the next inside pass fuses the cherries (level launches with full
// inside storage only)
static bool fuse_now(const laml_forest* f) {
  return f->fuse && f->cherry_ok && f->pass_mode == LAML_PASS_LEVELS && f->ckpt_L == 0 && f->W == 64;
}

static DevForest with_cherries(const laml_forest* f, const DevForest& F, bool fused) {
  DevForest G = F;
  G.cherry_end = fused ? f->level_off[2] : F.num_leaves;
  return G;
}

// hybrid pass mode: level launches below this level, one tree-walk launch
// (a block per tree and site chunk) for the levels from it up
static int hybrid_split(const laml_forest* f) {
  const int H = (int)f->level_off.size() - 1;
  return std::max(1, std::min(f->h_top, H));
}

// which passes run the levels below their split as heavy-path chains
// (LAML_CHAINS_* bits; row kernels, full inside storage, no cherry fusion;
// the tree walk pass has its own levels)
static int chains_now(const laml_forest* f) {
  return f->W == 64 && f->ckpt_L == 0 && f->pass_mode != LAML_PASS_TREES && !fuse_now(f) ? f->chains : 0;
}

// The chain plan of the internal nodes below level hs (built on first use,
// kept for the forest's life).  Heads: tree roots below hs, nodes whose
// parent lies at hs or above, and light children; a head's round is one
// more than its parent's chain's (0 when the state comes from above).
static const laml::ChainPlan* chain_plan(laml_forest* f, int hs) {
  for (const laml::ChainPlan& cp : f->chain_plans)
    if (cp.h_split == hs && cp.pack_cap == f->chain_pack) return &cp;
  // a new plan is uploaded with synchronous copies: not while another
  // thread captures its EM step (the passes only look plans up: each entry
  // point builds them before any capture)
  std::lock_guard<std::mutex> legacy_lock(g_legacy_mu);
  const int N = f->F.num_nodes;
  const int lo = f->level_off[1], B = f->level_off[hs];
  std::vector<int32_t> chain_of(N, -1), round_of(N, 0);
  std::vector<std::tuple<int, int, int>> heads;  // (round, -len, -head)
  for (int v = B - 1; v >= lo; --v) {
    const int pa = f->h_parent[v];
    const bool above = pa < 0 || pa >= B;
    if (!above && f->h_heavy[pa] == v) continue;
    const int r = above ? 0 : round_of[chain_of[pa]] + 1;
    int len = 0;
    for (int w = v; w >= 0; w = f->h_heavy[w]) {
      chain_of[w] = v;
      ++len;
    }
    round_of[v] = r;
    heads.emplace_back(r, -len, -v);
  }
  std::sort(heads.begin(), heads.end());
  // inside rounds: children precede parents, so a light child's chain has
  // its final round when its parent is visited
  std::vector<int32_t> in_round(N, 0);
  for (int v = lo; v < B; ++v) {
    const int cs[2] = {f->h_child0[v], f->h_child1[v]};
    for (int c : cs)
      if (c >= f->F.num_leaves && c != f->h_heavy[v])
        in_round[chain_of[v]] = std::max(in_round[chain_of[v]], in_round[c] + 1);
  }
  laml::ChainPlan cp;
  cp.h_split = hs;
  cp.pack_cap = f->chain_pack;
  std::vector<int2> ch;
  std::vector<int32_t> nodes, off_of(N, -1), round_off;
  std::vector<std::tuple<int, int, int>> in_heads;  // (inside round, -len, -head)
  ch.reserve(heads.size());
  nodes.reserve(B - lo);
  for (const auto& hd : heads) {
    const int r = std::get<0>(hd), v = -std::get<2>(hd);
    while ((int)round_off.size() <= r) round_off.push_back((int32_t)ch.size());
    off_of[v] = (int32_t)nodes.size();
    ch.push_back(make_int2((int)nodes.size(), -std::get<1>(hd)));
    for (int w = v; w >= 0; w = f->h_heavy[w]) nodes.push_back(w);
    in_heads.emplace_back(in_round[v], std::get<1>(hd), std::get<2>(hd));
  }
  round_off.push_back((int32_t)ch.size());
  std::sort(in_heads.begin(), in_heads.end());
  std::vector<int2> ich;
  ich.reserve(in_heads.size());
  for (const auto& hd : in_heads) {
    const int r = std::get<0>(hd), v = -std::get<2>(hd);
    while ((int)cp.in_round_off.size() <= r) cp.in_round_off.push_back((int32_t)ich.size());
    ich.push_back(make_int2(off_of[v], -std::get<1>(hd)));
  }
  cp.in_round_off.push_back((int32_t)ich.size());
  std::vector<laml::ChainStep> steps(nodes.size());
  for (size_t i = 0; i < nodes.size(); ++i) {
    const int v = nodes[i];
    laml::ChainStep& r = steps[i];
    r.u = v;
    r.ch0 = f->h_child0[v];
    r.ch1 = f->h_child1[v];
    const int hv = f->h_heavy[v];
    r.hv = hv < 0 ? -1 : (hv == r.ch0 ? 0 : 1);
    r.ro0 = f->h_row_off[r.ch0];
    r.ro1 = f->h_row_off[r.ch1];
  }
  // Outside packs: per round, the chains of each tree (longest first) go
  // back to back into packs of at most cap steps (a longer chain is a pack
  // of its own), so that a wavefront's staging runs on from one chain into
  // the next; the cap shrinks to keep kPackWaves wavefronts in the round,
  // and packs are launched longest first.
  std::vector<int2> packs;
  std::vector<laml::ChainStep> psteps;
  std::vector<int64_t> phro;
  psteps.reserve(steps.size());
  phro.reserve(steps.size());
  {
    struct Pack {
      int len;
      std::vector<int32_t> members;
    };
    for (size_t r = 0; r + 1 < round_off.size(); ++r) {
      cp.pack_round_off.push_back((int32_t)packs.size());
      std::vector<int32_t> order;
      int64_t round_steps = 0;
      for (int i = round_off[r]; i < round_off[r + 1]; ++i) {
        order.push_back(i);
        round_steps += ch[i].y;
      }
      const int cap = (int)std::max<int64_t>(
          1, std::min<int64_t>(cp.pack_cap, round_steps * f->F.max_chunks / f->pack_waves));
      // by tree, keeping the longest-first order within a tree
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return f->h_tree_of[nodes[ch[a].x]] < f->h_tree_of[nodes[ch[b].x]];
      });
      std::vector<Pack> rp;
      int cur_tree = -1;
      for (const int i : order) {
        const int tr = f->h_tree_of[nodes[ch[i].x]];
        if (tr != cur_tree || rp.back().len + ch[i].y > cap) {
          rp.push_back(Pack{0, {}});
          cur_tree = tr;
        }
        rp.back().len += ch[i].y;
        rp.back().members.push_back(i);
      }
      std::stable_sort(rp.begin(), rp.end(), [](const Pack& a, const Pack& b) { return a.len > b.len; });
      bool packed = false;
      for (const Pack& pk : rp) packed = packed || pk.members.size() > 1;
      cp.round_packed.push_back(packed ? 1 : 0);
      for (const Pack& pk : rp) {
        packs.push_back(make_int2((int)psteps.size(), pk.len));
        for (const int i : pk.members) {
          for (int j = 0; j < ch[i].y; ++j) {
            laml::ChainStep rec = steps[ch[i].x + j];
            if (j == 0) {
              phro.push_back(f->h_row_off[rec.u]);
              rec.u = ~rec.u;
            } else {
              phro.push_back(0);
            }
            psteps.push_back(rec);
          }
        }
      }
    }
    cp.pack_round_off.push_back((int32_t)packs.size());
  }
  if (f->upload(&cp.in_chains, ich.data(), ich.size()) || f->upload(&cp.nodes, nodes.data(), nodes.size()) ||
      f->upload(&cp.packs, packs.data(), packs.size()) || f->upload(&cp.pack_steps, psteps.data(), psteps.size()) ||
      f->upload(&cp.pack_head_ro, phro.data(), phro.size()))
    return nullptr;
  f->chain_plans.push_back(std::move(cp));
  return &f->chain_plans.back();
}

// the chain plan of a pass below split level hs (nullptr: that pass runs
// no chains)
static const laml::ChainPlan* pass_chains(laml_forest* f, bool inside, int hs) {
  if (!(chains_now(f) & (inside ? LAML_CHAINS_INSIDE : LAML_CHAINS_OUTSIDE))) return nullptr;
  const int t = inside ? f->chain_top_in : f->chain_top_out;
  const int top = t <= 0 ? hs : std::min(t, hs);
  return top > 1 ? chain_plan(f, top) : nullptr;
}

// builds the chain plans the next passes use, outside any stream capture
static int prepare_chains(laml_forest* f) {
  const int hs = hybrid_split(f);
  for (const bool inside : {true, false}) {
    if (!(chains_now(f) & (inside ? LAML_CHAINS_INSIDE : LAML_CHAINS_OUTSIDE))) continue;
    const int t = inside ? f->chain_top_in : f->chain_top_out;
    const int top = t <= 0 ? hs : std::min(t, hs);
    if (top > 1 && !chain_plan(f, top)) return fail(LAML_ENOMEM, "chain plan");
  }
  return LAML_OK;
}

// inside levels [1, hs) of the level and hybrid passes: chains below the
// chain region's top, level launches above it
static void inside_below(laml_forest* f, const DevForest& F, int hs, const int32_t* done, hipStream_t s) {
  const laml::ChainPlan* cp = pass_chains(f, true, hs);
  if (cp) laml::run_inside_chains(F, *cp, done, s);
  laml::run_inside_levels(F, f->level_off, f->W, cp ? cp->h_split : 1, hs, f->in_cpw, done, s);
}

// outside levels [1, hs) of the level and hybrid passes
static void outside_below(laml_forest* f, const DevForest& F, int hs, const int32_t* done, hipStream_t s,
                          bool allow_chains) {
  const laml::ChainPlan* cp = allow_chains ? pass_chains(f, false, hs) : nullptr;
  laml::run_outside_levels(F, f->level_off, f->W, cp ? cp->h_split : 1, hs, f->out_cpw, done, s);
  if (cp) laml::run_outside_chains(F, *cp, done, s);
  f->alpha_complete = cp == nullptr;
}

// inside pass and root step (site likelihoods) in the forest's pass mode
static void pass_inside_root(laml_forest* f, const DevForest& F, double* site_llh, const int32_t* done,
                             hipStream_t s) {
  if (f->pass_mode == LAML_PASS_HYBRID) {
    f->inside_fused = false;
    const int hs = hybrid_split(f);
    inside_below(f, F, hs, done, s);
    laml::run_tree_walk(F, f->TW, f->W, 3, site_llh, done, s, hs);
  } else if (f->pass_mode == 1) {
    f->inside_fused = false;
    laml::run_tree_walk(F, f->TW, f->W, 3, site_llh, done, s);
  } else {
    f->inside_fused = fuse_now(f);
    const DevForest G = with_cherries(f, F, f->inside_fused);
    const laml::ChainPlan* cp = pass_chains(f, true, hybrid_split(f));
    laml::run_inside(G, f->level_off, f->W, f->h_top, f->in_cpw, done, s, f->inside_fused ? &f->split : nullptr, cp);
    laml::run_root(G, f->W, site_llh, done, s);
  }
}

// outside pass with posterior partials (and the root step first when asked);
// after an inside pass with fused cherries the outside pass fuses them too
// (the cherries' inside rows were not written)
static void pass_outside(laml_forest* f, const DevForest& F, bool with_root, const int32_t* done, hipStream_t s,
                         bool allow_chains = true) {
  if (f->inside_fused) {
    const DevForest G = with_cherries(f, F, true);
    if (with_root) laml::run_root(G, f->W, nullptr, done, s);
    laml::run_outside(G, f->level_off, f->W, f->h_top, f->out_cpw, done, s, &f->split);
    f->alpha_complete = false;
    return;
  }
  if (f->pass_mode == 1) {
    laml::run_tree_walk(F, f->TW, f->W, (with_root ? 2 : 0) | 4, nullptr, done, s);
    f->alpha_complete = true;
    return;
  }
  if (f->pass_mode == LAML_PASS_HYBRID) {
    const int hs = hybrid_split(f);
    laml::run_tree_walk(F, f->TW, f->W, (with_root ? 2 : 0) | 4, nullptr, done, s, hs);
    outside_below(f, F, hs, done, s, allow_chains);
    return;
  }
  if (with_root) laml::run_root(F, f->W, nullptr, done, s);
  if (f->ckpt_L > 0) {
    laml::run_outside_ckpt(F, f->level_off, f->W, f->ckpt_L, f->in_cpw, f->out_cpw, done, s);
    f->alpha_complete = true;
    return;
  }
  const laml::ChainPlan* cp = allow_chains ? pass_chains(f, false, hybrid_split(f)) : nullptr;
  laml::run_outside(F, f->level_off, f->W, f->h_top, f->out_cpw, done, s, nullptr, cp);
  f->alpha_complete = cp == nullptr;
}

int laml_nll(laml_forest* f, const double* d, const double* nu, const double* phi, double* nll, double* site_llh,
             void* stream) {
  if (!f || !d || !nu || !phi || !nll) return fail(LAML_EINVAL, "null argument");
  DeviceGuard guard(f->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  int rc = prepare_chains(f);
  if (rc) return rc;
  DevForest F = f->F;
  F.phi = phi;
  laml::run_coef(F, d, nu, nullptr, s);
  pass_inside_root(f, F, site_llh, nullptr, s);
  EmState none{};
  laml::run_llh(F, nll, none, 0, 0.0, s);
  return check_launch();
}

int laml_estep(laml_forest* f, const double* d, const double* nu, const double* phi, double* counts, double* rtilde,
               double* nll, void* stream) {
  if (!f || !d || !nu || !phi || !counts || !rtilde || !nll) return fail(LAML_EINVAL, "null argument");
  DeviceGuard guard(f->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  DevForest F = f->F;
  F.phi = phi;
  laml::run_coef(F, d, nu, nullptr, s);
  int rc = prepare_chains(f);
  if (rc) return rc;
  if (f->pass_mode == 1) {
    f->inside_fused = false;
    laml::run_tree_walk(F, f->TW, f->W, 7, nullptr, nullptr, s);
    f->alpha_complete = true;
  } else if (f->pass_mode == LAML_PASS_HYBRID) {
    f->inside_fused = false;
    const int hs = hybrid_split(f);
    inside_below(f, F, hs, nullptr, s);
    laml::run_tree_walk(F, f->TW, f->W, 7, nullptr, nullptr, s, hs);
    outside_below(f, F, hs, nullptr, s, true);
  } else {
    pass_inside_root(f, F, nullptr, nullptr, s);
    pass_outside(f, F, false, nullptr, s);
  }
  EmState none{};
  laml::run_llh(F, nll, none, 0, 0.0, s);
  laml::run_counts(F, counts, s);
  laml::run_rtilde(F, rtilde, nullptr, s);
  return check_launch();
}

int laml_outside(laml_forest* f, const double* phi, double* counts, double* rtilde, void* stream) {
  if (!f || !phi || !counts || !rtilde) return fail(LAML_EINVAL, "null argument");
  DeviceGuard guard(f->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  int rc = prepare_chains(f);
  if (rc) return rc;
  DevForest F = f->F;
  F.phi = phi;
  pass_outside(f, F, true, nullptr, s);
  laml::run_counts(F, counts, s);
  laml::run_rtilde(F, rtilde, nullptr, s);
  return check_launch();
}

int laml_grad(laml_forest* f, const double* d, const double* nu, const double* phi, double* grad_d,
              double* grad_nu, double* grad_phi, void* stream) {
  if (!f || !d || !nu || !phi || !grad_d || !grad_nu || !grad_phi) return fail(LAML_EINVAL, "null argument");
  DeviceGuard guard(f->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  int rc = prepare_chains(f);
  if (rc) return rc;
  DevForest F = f->F;
  F.phi = phi;
  pass_outside(f, F, true, nullptr, s);
  laml::run_grad(F, d, nu, phi, grad_d, grad_nu, grad_phi, s);
  return check_launch();
}

int laml_mstep(laml_forest* f, const double* counts, const double* rtilde, double* d, double* nu, double* phi,
               void* stream) {
  if (!f || !counts || !rtilde || !d || !nu || !phi) return fail(LAML_EINVAL, "null argument");
  DeviceGuard guard(f->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  laml::run_mstep(f->F, counts, rtilde, d, nu, phi, nullptr, s, f->ultra_on ? &f->U : nullptr);
  return check_launch();
}

// one EM step of every problem not yet converged: inside pass, likelihood
// and convergence test, then (unless this is the final step) outside pass
// with posterior counts and the M-step
static void em_step(laml_forest* f, const DevForest& F, const EmState& em, double* d, double* nu, double* phi,
                    double* nll, int32_t max_iter, double tol, bool last, hipStream_t s) {
  laml::run_coef(F, d, nu, em.done, s);
  if (f->pass_mode == 1 && !last) {
    // one launch for the whole E-step; the outside half of a problem that
    // converges at this step is not used (its M-step is skipped)
    f->inside_fused = false;
    laml::run_tree_walk(F, f->TW, f->W, 7, nullptr, em.done, s);
    f->alpha_complete = true;
    laml::run_llh(F, nll, em, max_iter, tol, s);
  } else if (f->pass_mode == LAML_PASS_HYBRID && !last) {
    // bottom inside levels, one walk for the top of both passes, the
    // convergence test, then the bottom outside levels
    f->inside_fused = false;
    const int hs = hybrid_split(f);
    inside_below(f, F, hs, em.done, s);
    laml::run_tree_walk(F, f->TW, f->W, 7, nullptr, em.done, s, hs);
    laml::run_llh(F, nll, em, max_iter, tol, s);
    outside_below(f, F, hs, em.done, s, true);
  } else {
    pass_inside_root(f, F, nullptr, em.done, s);
    laml::run_llh(F, nll, em, max_iter, tol, s);
    if (last) return;
    pass_outside(f, F, false, em.done, s);
  }
  laml::run_mstep(F, nullptr, nullptr, d, nu, phi, em.done, s, f->ultra_on ? &f->U : nullptr);
}

int laml_em(laml_forest* f, double* d, double* nu, double* phi, int32_t max_iter, double tol, double* nll,
            int32_t* iters, double* trace, void* stream) {
  if (!f || !d || !nu || !phi || !nll || !iters) return fail(LAML_EINVAL, "null argument");
  if (max_iter < 0) return fail(LAML_EINVAL, "max_iter < 0");
  DeviceGuard guard(f->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  DevForest F = f->F;
  F.phi = phi;
  const int P = F.num_problems;
  EmState em = f->em;
  em.trace = trace;
  int prc = prepare_chains(f);
  if (prc) return prc;
  laml::run_em_reset(em, P, s);
  // The EM step is captured once into a HIP graph and replayed: its kernel
  // arguments do not change between steps (the step number lives on the
  // device).  The graph holds kernel nodes only.
  const bool graph = f->use_graph && max_iter > 0;
  if (graph) {
    // replayed steps all have step < max_iter, so the captured likelihood
    // kernel gets a bound it never reaches and the graph serves any max_iter
    const EmGraphKey key{d, nu, phi, nll, trace, tol, 0, f->plan_version};
    if (!f->em_exec || !(f->em_key == key)) {
      std::lock_guard<std::mutex> legacy_lock(g_legacy_mu);
      f->drop_graph();
      hipGraph_t g = nullptr;
      LAML_HIP(hipStreamBeginCapture(f->cap_stream, hipStreamCaptureModeThreadLocal));
      em_step(f, F, em, d, nu, phi, nll, std::numeric_limits<int32_t>::max(), tol, false, f->cap_stream);
      const hipError_t le = hipGetLastError();
      const hipError_t ce = hipStreamEndCapture(f->cap_stream, &g);
      if (le != hipSuccess || ce != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(LAML_EDEVICE, std::string("EM step capture: ") + hipGetErrorString(le != hipSuccess ? le : ce));
      }
      const hipError_t ie = hipGraphInstantiate(&f->em_exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ie != hipSuccess) {
        f->em_exec = nullptr;
        return fail(LAML_EDEVICE, std::string("EM graph instantiate: ") + hipGetErrorString(ie));
      }
      f->em_key = key;
    }
  }
  std::vector<int32_t> done_h(P);
  const int chunk = 8;
  for (int step = 0; step <= max_iter; ++step) {
    const bool last = step == max_iter;
    if (graph && !last)
      LAML_HIP(hipGraphLaunch(f->em_exec, s));
    else
      em_step(f, F, em, d, nu, phi, nll, max_iter, tol, last, s);
    int rc = check_launch();
    if (rc) return rc;
    if (tol > 0.0 && step % chunk == chunk - 1 && step < max_iter) {
      LAML_HIP(hipMemcpyAsync(done_h.data(), em.done, sizeof(int32_t) * P, hipMemcpyDeviceToHost, s));
      LAML_HIP(hipStreamSynchronize(s));
      bool all = true;
      for (int p = 0; p < P; ++p) all = all && done_h[p];
      if (all) break;
    }
  }
  // a replayed step runs no host code: the outside rows are those of the
  // captured step's pass
  if (max_iter > 0 && f->pass_mode != LAML_PASS_TREES)
    f->alpha_complete = !pass_chains(f, false, hybrid_split(f)) && !fuse_now(f);
  LAML_HIP(hipMemcpyAsync(iters, em.iters, sizeof(int32_t) * P, hipMemcpyDeviceToDevice, s));
  return check_launch();
}

int laml_posteriors(laml_forest* f, const double* phi, double* post0_host, double* postS_host, int64_t capacity,
                    void* stream) {
  // allocates, frees and synchronises: not while another thread captures
  std::lock_guard<std::mutex> legacy_lock(g_legacy_mu);
  if (!f || !phi || !post0_host || !postS_host) return fail(LAML_EINVAL, "null argument");
  const int64_t need = f->leaf_sites + f->int_sites;
  if (capacity < need) return fail(LAML_EINVAL, "posterior capacity too small");
  if (f->ckpt_L > 0) return fail(LAML_ESTATE, "posteriors need full inside storage (level checkpointing is on)");
  if (f->inside_fused)
    return fail(LAML_ESTATE, "posteriors need every node's rows: rerun the E-step with cherry fusion off "
                             "(laml_forest_set_fusion)");
  DeviceGuard guard(f->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  DevForest F = f->F;
  F.phi = phi;
  // the last outside pass kept heavy children's outside states in
  // registers: form every node's outside row again from the resident inside
  // state (one outside pass in level launches)
  if (!f->alpha_complete) pass_outside(f, F, true, nullptr, s, false);
  const int64_t rows = f->leaf_rows + f->int_rows;
  double *p0 = nullptr, *pS = nullptr;
  LAML_HIP(hipMalloc(&p0, sizeof(double) * std::max<int64_t>(rows, 1)));
  if (hipMalloc(&pS, sizeof(double) * std::max<int64_t>(rows, 1)) != hipSuccess) {
    (void)hipFree(p0);
    return fail(LAML_ENOMEM, "hipMalloc posteriors");
  }
  laml::run_node_post(F, f->W, f->leaf_rows, p0, pS, s);
  std::vector<double> h0(rows), hS(rows);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(h0.data(), p0, sizeof(double) * rows, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(hS.data(), pS, sizeof(double) * rows, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(p0);
  (void)hipFree(pS);
  if (e != hipSuccess) return fail(LAML_EDEVICE, std::string("posteriors: ") + hipGetErrorString(e));
  const int N = f->F.num_nodes, n = f->F.num_leaves;
  for (int v = 0; v < N; ++v) {
    const int K = f->h_Kt[v];
    const int64_t src = (v < n ? 0 : f->leaf_rows) + f->h_row_off[v];
    const int64_t dst = (v < n ? 0 : f->leaf_sites) + f->h_desc_off[v];
    std::memcpy(post0_host + dst, h0.data() + src, sizeof(double) * K);
    std::memcpy(postS_host + dst, hS.data() + src, sizeof(double) * K);
  }
  return LAML_OK;
}

int laml_posteriors_noalloc(laml_forest* f, const double* phi, double* post0_host, double* postS_host,
                            int64_t capacity, void* stream) {
  if (!f || !phi || !post0_host || !postS_host) return fail(LAML_EINVAL, "null argument");
  const int64_t need = f->leaf_sites + f->int_sites;
  if (capacity < need) return fail(LAML_EINVAL, "posterior capacity too small");
  const int64_t rows = f->leaf_rows + f->int_rows;
  // padded rows of both posteriors, then their packed copies
  if (2 * (rows + need) > laml::kPostScratchDoubles)
    return fail(LAML_EINVAL, "laml_posteriors_noalloc: the forest's posteriors (2 x (padded rows + entries) "
                             "doubles) exceed the fixed staging area of 8388608 doubles; use laml_posteriors");
  if (f->ckpt_L > 0) return fail(LAML_ESTATE, "posteriors need full inside storage (level checkpointing is on)");
  if (f->inside_fused)
    return fail(LAML_ESTATE, "posteriors need every node's rows: rerun the E-step with cherry fusion off "
                             "(laml_forest_set_fusion)");
  if (!f->post_scratch) return fail(LAML_ESTATE, "no staging area on this device");
  DeviceGuard guard(f->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  DevForest F = f->F;
  F.phi = phi;
  if (!f->alpha_complete) pass_outside(f, F, true, nullptr, s, false);
  double* p0 = f->post_scratch;
  double* pS = p0 + rows;
  double* o0 = pS + rows;
  double* oS = o0 + need;
  laml::run_node_post(F, f->W, f->leaf_rows, p0, pS, s);
  laml::run_post_pack(F, f->d_desc_off, f->leaf_rows, f->leaf_sites, p0, pS, o0, oS, s);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && need > 0)
    e = hipMemcpyAsync(post0_host, o0, sizeof(double) * need, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess && need > 0)
    e = hipMemcpyAsync(postS_host, oS, sizeof(double) * need, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail(LAML_EDEVICE, std::string("posteriors: ") + hipGetErrorString(e));
  return LAML_OK;
}

int laml_quartet_em(laml_forest* f, int32_t num_candidates, const int32_t* quartet_nodes, const double* lengths_in,
                    const uint8_t* optimise, const double* llh_const, const double* nu, const double* phi,
                    int32_t max_iter, double tol, double* lengths_out, double* nll, int32_t* iters, void* stream) {
  if (!f || num_candidates < 0) return fail(LAML_EINVAL, "bad arguments");
  if (num_candidates == 0) return LAML_OK;
  if (!quartet_nodes || !lengths_in || !optimise || !llh_const || !nu || !phi || !lengths_out || !nll || !iters)
    return fail(LAML_EINVAL, "null argument");
  if (max_iter < 0) return fail(LAML_EINVAL, "max_iter < 0");
  if (f->ckpt_L > 0) return fail(LAML_ESTATE, "quartet EM needs full inside storage (level checkpointing is on)");
  if (f->inside_fused)
    return fail(LAML_ESTATE, "quartet EM needs every node's rows: rerun the E-step with cherry fusion off "
                             "(laml_forest_set_fusion)");
  DeviceGuard guard(f->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!f->alpha_complete) {
    DevForest F = f->F;
    F.phi = phi;
    pass_outside(f, F, true, nullptr, s, false);
  }
  laml::run_quartet_em(f->F, num_candidates, quartet_nodes, lengths_in, optimise, llh_const, nu, phi, max_iter, tol,
                       lengths_out, nll, iters, s);
  return check_launch();
}

}  // extern "C"
