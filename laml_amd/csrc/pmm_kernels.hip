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

#include <algorithm>
#include <map>
#include <mutex>

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
  int es;
  c.msil = frexp(-expm1(-nv * dv), &es);
  c.esil = c.msil != 0.0 ? es : 0;
  c.pad = 0;
  F.coef[v] = c;
}

// Segment mapping shared by the level kernels: a block of 256 threads holds
// 256/W segments of W lanes; segment s of a level handles node beg+s, and
// blockIdx.y selects the chunk of W consecutive sites of that node.
template <int W>
struct Seg {
  int v, lane, k;
  __device__ __forceinline__ bool init(int beg, int cnt) {
    const int seg = blockIdx.x * (256 / W) + (int)(threadIdx.x / W);
    lane = threadIdx.x & (W - 1);
    if (seg >= cnt) return false;
    v = beg + seg;
    k = (int)blockIdx.y * W + lane;
    return true;
  }
};

template <int W>
__global__ void __launch_bounds__(256) k_labels(DevForest F, int beg, int cnt) {
  Seg<W> g;
  if (!g.init(beg, cnt)) return;
  const int v = g.v, k = g.k;
  const int t = F.tree_of[v];
  if (k >= F.tree_nsites[t]) return;
  const int c1 = F.child0[v], c2 = F.child1[v];
  const int la = c1 < F.num_leaves ? code_label(F.codes[F.row_off[c1] + k]) : (int)F.labels[F.row_off[c1] + k];
  const int lb = c2 < F.num_leaves ? code_label(F.codes[F.row_off[c2] + k]) : (int)F.labels[F.row_off[c2] + k];
  F.labels[F.row_off[v] + k] = (int16_t)combine_label(la, lb);
}

// Prior-weight rows of every node (a wavefront per row; sites past the
// tree's up to the row stride get 0).
__global__ void __launch_bounds__(256) k_qrows(DevForest F, double* __restrict__ qleaf, double* __restrict__ qint) {
  const int lane = (int)(threadIdx.x & 63);
  const int W = F.chunk_w;
  for (int v = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6); v < F.num_nodes; v += (int)gridDim.x * 4) {
    const int t = F.tree_of[v];
    const int K = F.tree_nsites[t];
    const int stride = (K + W - 1) / W * W;
    const int sb = F.tree_site_base[t];
    const int64_t ro = F.row_off[v];
    const bool leaf = v < F.num_leaves;
    double* dst = (leaf ? qleaf : qint) + ro;
    for (int k = lane; k < stride; k += 64) {
      double q = 0.0;
      if (k < K) {
        const int lab = leaf ? code_label(F.codes[ro + k]) : (int)F.labels[ro + k];
        q = lab > 0 ? F.prior[F.site_qoff[sb + k] - 1 + lab] : F.site_qsum[sb + k];
      }
      dst[k] = q;
    }
  }
}

// Leaf code rows from the caller's packed layout (leaf v's K codes at
// packed_off[v], rows back to back) into the device rows, padded to a
// multiple of the segment width with code 0; a wavefront per row.
__global__ void __launch_bounds__(256) k_pad_codes(DevForest F, const int16_t* __restrict__ packed,
                                                   const int64_t* __restrict__ packed_off) {
  const int lane = (int)(threadIdx.x & 63);
  const int W = F.chunk_w;
  int16_t* codes = const_cast<int16_t*>(F.codes);
  for (int v = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6); v < F.num_leaves; v += (int)gridDim.x * 4) {
    const int K = F.tree_nsites[F.tree_of[v]];
    const int stride = (K + W - 1) / W * W;
    const int16_t* src = packed + packed_off[v];
    int16_t* dst = codes + F.row_off[v];
    for (int k = lane; k < stride; k += 64) dst[k] = k < K ? src[k] : (int16_t)0;
  }
}

// load the inside state of node c at site k (labels/codes by row_off, the
// inside vector by brow, which differs from row_off under level checkpointing)
__device__ __forceinline__ void load_child(const DevForest& F, int c, bool leaf, int k, double phi, int& lab,
                                           double& b0, double& bA, int& E, int& code) {
  const int64_t r = F.row_off[c];
  if (leaf) {
    code = F.codes[r + k];
    leaf_beta(code, phi, lab, b0, bA);
    E = 0;
  } else {
    code = 0;
    lab = F.labels[r + k];
    const int64_t br = F.brow[c];
    const double2 bb = F.beta[br + k];
    b0 = bb.x;
    bA = bb.y;
    E = F.bE[br + k];
  }
}

// Element k of a row: the byte offset is formed in 32 bits, so that with a
// wave-uniform row pointer the access is a scalar base plus a 32-bit lane
// offset (rows hold fewer than 2^24 sites; the forest checks it).
template <typename T>
__device__ __forceinline__ T& row_at(T* p, unsigned k) {
  return *reinterpret_cast<T*>(reinterpret_cast<char*>(p) + k * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ const T& row_at(const T* p, unsigned k) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + k * (unsigned)sizeof(T));
}

// A load of data that no kernel writes while it runs (launch plans, node
// records, branch coefficients), through the constant address space: at a
// wave-uniform address the compiler may then issue it as a scalar load into
// SGPRs even after the kernel's own stores, where it must otherwise assume
// the store may have changed it and load it into VGPRs of every lane (other
// addresses stay vector loads).
template <typename T>
__device__ __forceinline__ T uniform_load(const T* p) {
  static_assert(sizeof(T) % 4 == 0, "loaded as 32-bit words");
  typedef const __attribute__((address_space(4))) uint32_t* WordPtr;
  const WordPtr q = (WordPtr)(uintptr_t)p;
  uint32_t w[sizeof(T) / 4];
#pragma unroll
  for (unsigned i = 0; i < sizeof(T) / 4; ++i) w[i] = q[i];
  T out;
  __builtin_memcpy(&out, w, sizeof(T));
  return out;
}

// Kinds of child a pass kernel meets.  A *cherry* is an internal node whose
// children are both leaves; when the pass fuses cherries (F.cherry_end >
// F.num_leaves: nodes [num_leaves, cherry_end) are exactly the cherries, the
// whole height-1 level) no cherry row is stored: its parent forms the
// cherry's inside vector from the two leaf codes, and in the outside pass
// runs the cherry's posterior step itself, so the height-1 level costs no
// launch and no row traffic.
// Template kind arguments: a fixed kind, or kNoCherry (leaf or internal,
// read at run time: kernels that never meet a fused cherry carry no cherry
// code).
constexpr int kInternal = 0, kLeaf = 1, kCherry = 2, kNoCherry = -2;

__device__ __forceinline__ int child_kind(const DevForest& F, int c) {
  return c < F.num_leaves ? kLeaf : (c < F.cherry_end ? kCherry : kInternal);
}

template <int KT>
__device__ __forceinline__ int resolve_kind(int k) {
  return KT >= 0 ? KT : (KT == kNoCherry ? (k == kLeaf ? kLeaf : kInternal) : k);
}

// Rows of one child node at its parent's sites: leaf codes, label + inside
// vector, or (cherry) the codes and branch coefficients of its two leaves.
// Held as pointers to the row start so that, where the node is wave-uniform,
// every per-site access is a scalar base plus a 32-bit lane offset.
struct ChildRows {
  const int16_t* code;   // leaf codes (leaf), az-partition labels (internal), first leaf's codes (cherry)
  const double2* beta;   // inside mantissas (internal)
  const int32_t* bE;     // inside exponent (internal)
  const int16_t* code2;  // second leaf's codes (cherry)
  Coef g[2];             // the cherry's leaf branches
  int gc[2];             // the cherry's leaves
};

__device__ __forceinline__ ChildRows child_rows(const DevForest& F, int c, int kind) {
  ChildRows r;
  r.beta = nullptr;
  r.bE = nullptr;
  r.code2 = nullptr;
  r.gc[0] = r.gc[1] = -1;
  if (kind == kLeaf) {
    r.code = F.codes + F.row_off[c];
  } else if (kind == kCherry) {
    const int a = F.child0[c], b = F.child1[c];
    r.code = F.codes + F.row_off[a];
    r.code2 = F.codes + F.row_off[b];
    r.g[0] = F.coef[a];
    r.g[1] = F.coef[b];
    r.gc[0] = a;
    r.gc[1] = b;
  } else {
    const int64_t br = F.brow[c];
    r.code = F.labels + F.row_off[c];
    r.beta = F.beta + br;
    r.bE = F.bE + br;
  }
  return r;
}
__device__ __forceinline__ ChildRows child_rows(const DevForest& F, int c, bool leaf) {
  return child_rows(F, c, leaf ? kLeaf : kInternal);
}

// inside vector of a leaf at site k, seen from its parent over its branch
__device__ __forceinline__ Eps leaf_edge(int code, double phi, const double* __restrict__ prior, int qo, double qs,
                                         const Coef& cf, int& lab) {
  double b0, bA;
  leaf_beta(code, phi, lab, b0, bA);
  const double q = lab > 0 ? row_at(prior, (unsigned)(qo + lab)) : qs;
  return edge(lab, b0, bA, 0, q, cf);
}

// The inside vector a cherry's node would store at site k: the same
// operations as the inside pass for a node with two leaf children.
__device__ __forceinline__ void cherry_beta(const ChildRows& r, int k, double phi, const double* __restrict__ prior,
                                            int qo, double qs, int& lab, double& b0, double& bA, int& E) {
  int la, lb;
  const Eps ea = leaf_edge(row_at(r.code, k), phi, prior, qo, qs, r.g[0], la);
  const Eps eb = leaf_edge(row_at(r.code2, k), phi, prior, qo, qs, r.g[1], lb);
  lab = combine_label(la, lb);
  const double B0 = ea.e0 * eb.e0;
  const double BA = lab != 0 ? ea.eA * eb.eA : 0.0;
  const int sh = min(-exp_of(fmax(B0, BA)), 1000);
  const double sc = ldexp(1.0, sh);
  b0 = B0 * sc;
  bA = BA * sc;
  E = ea.e + eb.e - sh;
}

// inside state of a child at site k (qo, qs: the site's prior offset and
// mass, used for a cherry's leaves).  stream: read the inside row with the
// non-temporal hint (the inside pass reads each row once).
__device__ __forceinline__ void load_child_rows(const ChildRows& r, int kind, int k, double phi,
                                                const double* __restrict__ prior, int qo, double qs, int& lab,
                                                double& b0, double& bA, int& E, int& code, bool stream = false) {
  if (kind == kLeaf) {
    code = row_at(r.code, k);
    leaf_beta(code, phi, lab, b0, bA);
    E = 0;
  } else if (kind == kCherry) {
    code = 0;
    cherry_beta(r, k, phi, prior, qo, qs, lab, b0, bA, E);
  } else {
    code = 0;
    lab = row_at(r.code, k);
    const double2 bb = stream ? load_stream(&row_at(r.beta, k)) : row_at(r.beta, k);
    b0 = bb.x;
    bA = bb.y;
    E = row_at(r.bE, k);
  }
}
__device__ __forceinline__ void load_child_rows(const ChildRows& r, bool leaf, int k, double phi, int& lab,
                                                double& b0, double& bA, int& E, int& code) {
  load_child_rows(r, leaf ? kLeaf : kInternal, k, phi, nullptr, 0, 0.0, lab, b0, bA, E, code);
}

// per-node values of the inside pass (uniform over the sites of the node)
struct InNode {
  int K;
  int kind[2];
  Coef cf[2];
  ChildRows c[2];
  const int16_t* lab;    // the node's labels
  double2* beta;         // the node's inside row
  int32_t* bE;
  const int32_t* qoff;   // prior offsets of the tree's sites
  const double* qsum;    // prior mass of the tree's sites
  double phi;
  bool stream;           // rows with the non-temporal hint
};

// CH: the node may have a fused cherry child (otherwise children are leaves
// or stored internal nodes, and no cherry code is generated)
template <bool CH>
__device__ __forceinline__ void in_node_load(const DevForest& F, int v, int t, int p, int K, InNode& n) {
  n.K = K;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int c = i == 0 ? F.child0[v] : F.child1[v];
    n.kind[i] = CH ? child_kind(F, c) : (c < F.num_leaves ? kLeaf : kInternal);
    n.cf[i] = F.coef[c];
    n.c[i] = child_rows(F, c, n.kind[i]);
  }
  n.lab = F.labels + F.row_off[v];
  const int64_t bv = F.brow[v];
  n.beta = F.beta + bv;
  n.bE = F.bE + bv;
  n.qoff = F.site_qoff + F.tree_site_base[t];
  n.qsum = F.site_qsum + F.tree_site_base[t];
  n.phi = F.phi[p];
  n.stream = F.stream_rows != 0;
}

// The inside step of one site in three phases: row loads, prior gathers
// (they need the loaded labels), then arithmetic and stores.
struct InSite {
  int lv, qo, la, lb, Ea, Eb, ca, cb;
  double qs, a0, aA, b0, bA, qa, qb;
};

template <int K0, int K1>
__device__ __forceinline__ void inside_load(const DevForest& F, const InNode& n, int k, InSite& x) {
  const int kind[2] = {resolve_kind<K0>(n.kind[0]), resolve_kind<K1>(n.kind[1])};
  x.lv = row_at(n.lab, k);
  x.qo = row_at(n.qoff, k) - 1;
  // a missing ('?') child may be in any edit state: its prior weight is the
  // site's total prior mass
  x.qs = row_at(n.qsum, k);
  load_child_rows(n.c[0], kind[0], k, n.phi, F.prior, x.qo, x.qs, x.la, x.a0, x.aA, x.Ea, x.ca, n.stream);
  load_child_rows(n.c[1], kind[1], k, n.phi, F.prior, x.qo, x.qs, x.lb, x.b0, x.bA, x.Eb, x.cb, n.stream);
}

__device__ __forceinline__ void inside_gather(const DevForest& F, InSite& x) {
  x.qa = x.la > 0 ? row_at(F.prior, (unsigned)(x.qo + x.la)) : x.qs;
  x.qb = x.lb > 0 ? row_at(F.prior, (unsigned)(x.qo + x.lb)) : x.qs;
}

__device__ __forceinline__ void inside_finish(const InNode& n, int k, const InSite& x) {
  const Eps ea = edge(x.la, x.a0, x.aA, x.Ea, x.qa, n.cf[0]);
  const Eps eb = edge(x.lb, x.b0, x.bA, x.Eb, x.qb, n.cf[1]);
  const double B0 = ea.e0 * eb.e0;
  const double BA = x.lv != 0 ? ea.eA * eb.eA : 0.0;
  // renormalise the pair by a power of two (bounded so that a denormal pair
  // cannot produce an infinite scale)
  const int sh = min(-exp_of(fmax(B0, BA)), 1000);
  const double sc = ldexp(1.0, sh);
  const double2 b = make_double2(B0 * sc, BA * sc);
  if (n.stream)
    store_stream(&row_at(n.beta, k), b);
  else
    row_at(n.beta, k) = b;
  row_at(n.bE, k) = ea.e + eb.e - sh;
}

// inside vector (below the node) of internal node n at site k.  K0 / K1: the
// kind of each child (kLeaf, kInternal, kCherry), or kNoCherry to read
// n.kind at run time.
template <int K0, int K1>
__device__ __forceinline__ void inside_site(const DevForest& F, const InNode& n, int k) {
  InSite x;
  inside_load<K0, K1>(F, n, k, x);
  inside_gather(F, x);
  inside_finish(n, k, x);
}
// the same with the children's prior weights read from their prior-weight
// rows q0, q1 (no load waits for the labels)
__device__ __forceinline__ void inside_site_q(const DevForest& F, const InNode& n, int k, const double* q0,
                                              const double* q1) {
  InSite x;
  inside_load<kNoCherry, kNoCherry>(F, n, k, x);
  x.qa = row_at(q0, k);
  x.qb = row_at(q1, k);
  inside_finish(n, k, x);
}

// The inside-pass values of internal node v from its record (children that
// are leaves or stored internal nodes).
__device__ __forceinline__ void in_node_rec(const DevForest& F, const NodeRec& r, InNode& n) {
  n.K = r.K;
  const int ch[2] = {r.ch0, r.ch1};
  const int64_t ro[2] = {r.ro0, r.ro1}, bro[2] = {r.bro0, r.bro1};
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    n.kind[i] = ch[i] < F.num_leaves ? kLeaf : kInternal;
    n.cf[i] = uniform_load(F.coef + ch[i]);
    n.c[i].code = (n.kind[i] == kLeaf ? F.codes : F.labels) + ro[i];
    n.c[i].beta = F.beta + bro[i];
    n.c[i].bE = F.bE + bro[i];
    n.c[i].code2 = nullptr;
    n.c[i].gc[0] = n.c[i].gc[1] = -1;
  }
  n.lab = F.labels + r.ro;
  n.beta = F.beta + r.bro;
  n.bE = F.bE + r.bro;
  n.qoff = F.site_qoff + r.sb;
  n.qsum = F.site_qsum + r.sb;
  n.phi = uniform_load(F.phi + r.p);
  n.stream = F.stream_rows != 0;
}

// segment form: one site k of node v
__device__ __forceinline__ void inside_node(const DevForest& F, int v, int k, const int32_t* __restrict__ done) {
  const NodeRec r = F.nrec[v - F.num_leaves];
  if (k >= r.K) return;
  if (done && done[r.p]) return;
  InNode n;
  in_node_rec(F, r, n);
  inside_site<kNoCherry, kNoCherry>(F, n, k);
}

template <int W>
__global__ void __launch_bounds__(256) k_inside(DevForest F, int beg, int cnt, const int32_t* __restrict__ done) {
  Seg<W> g;
  if (!g.init(beg, cnt)) return;
  inside_node(F, g.v, g.k, done);
}

// Row form (K > 32): one wavefront per node walks chunks [c0, c0 + cpw) of 64
// sites.  The node number is wave-uniform, so the node's metadata are scalar
// loads and the child-kind tests scalar branches.
template <int K0, int K1>
__device__ __forceinline__ void inside_row_run(const DevForest& F, const InNode& n, int c0, int c1, int lane) {
  // software-pipelined over the wave's chunks: the rows and prior gathers of
  // chunk c + 1 are issued before chunk c's arithmetic and stores (they touch
  // other sites, so the order of memory accesses is free)
  InSite x;
  int k = c0 * 64 + lane;
  bool on = k < n.K;
  if (on) {
    inside_load<K0, K1>(F, n, k, x);
    inside_gather(F, x);
  }
  for (int c = c0; c < c1; ++c) {
    InSite y;
    const int kn = k + 64;
    const bool onn = c + 1 < c1 && kn < n.K;
    if (onn) inside_load<K0, K1>(F, n, kn, y);
    if (on) inside_finish(n, k, x);
    if (onn) inside_gather(F, y);
    x = y;
    k = kn;
    on = onn;
  }
}

// Both children leaves (the cherry level, whose rows are mostly written):
// the codes, labels and site values of the wavefront's chunks (at most
// kLeafGroup, their number a template argument so that the code is
// straight-line) are loaded first, then their prior gathers, then the
// arithmetic and stores, so that a wavefront waits twice instead of twice
// per chunk (a load issued after a store waits for the store too).  The
// per-chunk state is kept compact (raw codes until the arithmetic); same
// operations as inside_site, so the same bits.
constexpr int kLeafGroup = 8;
template <int NG>
__device__ __forceinline__ void inside_row_run_leaves_n(const DevForest& F, const InNode& n, int c0, int lane) {
  // lanes past the tree's sites load its last site, so that no load depends
  // on the lane mask
  int lv[NG], ca[NG], cb[NG], qo[NG];
  double qs[NG], ga[NG], gb[NG];
#pragma unroll
  for (int j = 0; j < NG; ++j) {
    const unsigned kk = min((c0 + j) * 64 + lane, n.K - 1);
    lv[j] = row_at(n.lab, kk);
    qo[j] = row_at(n.qoff, kk);
    qs[j] = row_at(n.qsum, kk);
    ca[j] = row_at(n.c[0].code, kk);
    cb[j] = row_at(n.c[1].code, kk);
  }
#pragma unroll
  for (int j = 0; j < NG; ++j) {
    const int la = code_label(ca[j]), lb = code_label(cb[j]);
    ga[j] = row_at(F.prior, (unsigned)(la > 0 ? qo[j] - 1 + la : 0));
    gb[j] = row_at(F.prior, (unsigned)(lb > 0 ? qo[j] - 1 + lb : 0));
  }
#pragma unroll
  for (int j = 0; j < NG; ++j) {
    const int k = (c0 + j) * 64 + lane;
    if (k < n.K) {
      InSite x;
      x.lv = lv[j];
      x.qo = qo[j] - 1;
      x.qs = qs[j];
      x.ca = ca[j];
      x.cb = cb[j];
      leaf_beta(ca[j], n.phi, x.la, x.a0, x.aA);
      leaf_beta(cb[j], n.phi, x.lb, x.b0, x.bA);
      x.Ea = 0;
      x.Eb = 0;
      x.qa = x.la > 0 ? ga[j] : x.qs;
      x.qb = x.lb > 0 ? gb[j] : x.qs;
      inside_finish(n, k, x);
    }
  }
}
__device__ __forceinline__ void inside_row_run_leaves(const DevForest& F, const InNode& n, int c0, int c1, int lane) {
  switch (c1 - c0) {
    case 1: inside_row_run_leaves_n<1>(F, n, c0, lane); break;
    case 2: inside_row_run_leaves_n<2>(F, n, c0, lane); break;
    case 3: inside_row_run_leaves_n<3>(F, n, c0, lane); break;
    case 4: inside_row_run_leaves_n<4>(F, n, c0, lane); break;
    case 5: inside_row_run_leaves_n<5>(F, n, c0, lane); break;
    case 6: inside_row_run_leaves_n<6>(F, n, c0, lane); break;
    case 7: inside_row_run_leaves_n<7>(F, n, c0, lane); break;
    default: inside_row_run_leaves_n<8>(F, n, c0, lane); break;
  }
}

// The wave's node: entry beg + wv of a node list, or of the level range.
// Returns -1 past the end.
__device__ __forceinline__ int row_node(int beg, int cnt, const int32_t* __restrict__ nodes) {
  const int wv = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
  if (wv >= cnt) return -1;
  return nodes ? __builtin_amdgcn_readfirstlane(nodes[beg + wv]) : beg + wv;
}

// Nodes whose children are leaves or stored internal nodes.
__global__ void __launch_bounds__(256) k_inside_rows(DevForest F, int beg, int cnt, int cpw,
                                                      const int32_t* __restrict__ done,
                                                      const int32_t* __restrict__ nodes) {
  const int v = row_node(beg, cnt, nodes);
  if (v < 0) return;
  const int lane = threadIdx.x & 63;
  const NodeRec r = F.nrec[v - F.num_leaves];
  const int nch = (r.K + 63) >> 6;
  const int c0 = (int)blockIdx.y * cpw;
  if (c0 >= nch) return;
  if (done && done[r.p]) return;
  const int c1 = min(c0 + cpw, nch);
  InNode n;
  in_node_rec(F, r, n);
  if (n.kind[0] == kLeaf) {
    if (n.kind[1] == kLeaf) inside_row_run<kLeaf, kLeaf>(F, n, c0, c1, lane);
    else inside_row_run<kLeaf, kInternal>(F, n, c0, c1, lane);
  } else {
    if (n.kind[1] == kLeaf) inside_row_run<kInternal, kLeaf>(F, n, c0, c1, lane);
    else inside_row_run<kInternal, kInternal>(F, n, c0, c1, lane);
  }
}

// The same for level 1, whose nodes have two leaf children (a separate
// kernel: its loads-first path needs more registers than the other levels'
// kernel should carry).
__global__ void __launch_bounds__(256) k_inside_rows_leaves(DevForest F, int beg, int cnt, int cpw,
                                                             const int32_t* __restrict__ done) {
  const int v = row_node(beg, cnt, nullptr);
  if (v < 0) return;
  const int lane = threadIdx.x & 63;
  const NodeRec r = F.nrec[v - F.num_leaves];
  const int nch = (r.K + 63) >> 6;
  const int c0 = (int)blockIdx.y * cpw;
  if (c0 >= nch) return;
  if (done && done[r.p]) return;
  const int c1 = min(c0 + cpw, nch);
  InNode n;
  in_node_rec(F, r, n);
  if (n.kind[0] == kLeaf && n.kind[1] == kLeaf) {
    if (c1 - c0 <= kLeafGroup) inside_row_run_leaves(F, n, c0, c1, lane);
    else inside_row_run<kLeaf, kLeaf>(F, n, c0, c1, lane);
  } else if (n.kind[0] == kLeaf) inside_row_run<kLeaf, kInternal>(F, n, c0, c1, lane);
  else if (n.kind[1] == kLeaf) inside_row_run<kInternal, kLeaf>(F, n, c0, c1, lane);
  else inside_row_run<kInternal, kInternal>(F, n, c0, c1, lane);
}

// Nodes with at least one fused cherry child (a separate kernel, so that the
// cherry code does not raise the register count of the common case).
__global__ void __launch_bounds__(256) k_inside_rows_ch(DevForest F, int beg, int cnt, int cpw,
                                                         const int32_t* __restrict__ done,
                                                         const int32_t* __restrict__ nodes) {
  const int v = row_node(beg, cnt, nodes);
  if (v < 0) return;
  const int lane = threadIdx.x & 63;
  const int t = F.tree_of[v];
  const int p = F.tree_prob[t];
  const int K = F.tree_nsites[t];
  const int nch = (K + 63) >> 6;
  const int c0 = (int)blockIdx.y * cpw;
  if (c0 >= nch) return;
  if (done && done[p]) return;
  const int c1 = min(c0 + cpw, nch);
  InNode n;
  in_node_load<true>(F, v, t, p, K, n);
  if (n.kind[0] == kCherry) {
    if (n.kind[1] == kCherry) inside_row_run<kCherry, kCherry>(F, n, c0, c1, lane);
    else if (n.kind[1] == kLeaf) inside_row_run<kCherry, kLeaf>(F, n, c0, c1, lane);
    else inside_row_run<kCherry, kInternal>(F, n, c0, c1, lane);
  } else if (n.kind[0] == kLeaf) {
    inside_row_run<kLeaf, kCherry>(F, n, c0, c1, lane);
  } else {
    inside_row_run<kInternal, kCherry>(F, n, c0, c1, lane);
  }
}

// Inside pass along heavy-path chains (row form, one 64-site chunk per
// wavefront): the wavefront walks its chain from the deepest node (whose
// children are both leaves) up to the head, keeping each node's inside
// state in registers for its parent, so only light children's inside rows
// are read.  Every node's inside row is still written (the outside pass
// reads it).  The same operations as the level kernels, the same bits.
struct InCarry {
  int lab, E;
  double b0, bA;
};

// one site of node n whose child HV (0 or 1) is the carried heavy child and
// whose other child has kind KL; HV < 0: both children are leaves
template <int HV, int KL>
__device__ __forceinline__ void chain_in_site(const DevForest& F, const InNode& n, int k, int qo, double qs,
                                              InCarry& c) {
  InSite x;
  x.lv = row_at(n.lab, k);
  x.qo = qo;
  x.qs = qs;
  if (HV == 0) {
    x.la = c.lab;
    x.a0 = c.b0;
    x.aA = c.bA;
    x.Ea = c.E;
    load_child_rows(n.c[1], KL, k, n.phi, F.prior, qo, qs, x.lb, x.b0, x.bA, x.Eb, x.cb, n.stream);
  } else if (HV == 1) {
    load_child_rows(n.c[0], KL, k, n.phi, F.prior, qo, qs, x.la, x.a0, x.aA, x.Ea, x.ca, n.stream);
    x.lb = c.lab;
    x.b0 = c.b0;
    x.bA = c.bA;
    x.Eb = c.E;
  } else {
    load_child_rows(n.c[0], kLeaf, k, n.phi, F.prior, qo, qs, x.la, x.a0, x.aA, x.Ea, x.ca, n.stream);
    load_child_rows(n.c[1], kLeaf, k, n.phi, F.prior, qo, qs, x.lb, x.b0, x.bA, x.Eb, x.cb, n.stream);
  }
  inside_gather(F, x);
  const Eps ea = edge(x.la, x.a0, x.aA, x.Ea, x.qa, n.cf[0]);
  const Eps eb = edge(x.lb, x.b0, x.bA, x.Eb, x.qb, n.cf[1]);
  const double B0 = ea.e0 * eb.e0;
  const double BA = x.lv != 0 ? ea.eA * eb.eA : 0.0;
  const int sh = min(-exp_of(fmax(B0, BA)), 1000);
  const double sc = ldexp(1.0, sh);
  const double2 b = make_double2(B0 * sc, BA * sc);
  if (n.stream)
    store_stream(&row_at(n.beta, k), b);
  else
    row_at(n.beta, k) = b;
  const int E = ea.e + eb.e - sh;
  row_at(n.bE, k) = E;
  c.lab = x.lv;
  c.b0 = b.x;
  c.bA = b.y;
  c.E = E;
}

// Chains [cbeg, cbeg + cnt) of an inside chain order (chain c is
// cnodes[chains[c].x .. chains[c].x + chains[c].y), head first; the walk
// starts at the last node).  Wavefront w runs chain w / max_chunks at chunk
// w % max_chunks; light children's inside rows must be complete.
__global__ void __launch_bounds__(256) k_inside_chain(DevForest F, const int2* __restrict__ chains, int cbeg,
                                                       int cnt, const int32_t* __restrict__ cnodes,
                                                       const int32_t* __restrict__ done) {
  const int w = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
  const int wv = w / F.max_chunks;
  if (wv >= cnt) return;
  const int2 cd = chains[cbeg + wv];
  const int nb = __builtin_amdgcn_readfirstlane(cd.x);
  const int len = __builtin_amdgcn_readfirstlane(cd.y);
  int prev = -1;
  int v = __builtin_amdgcn_readfirstlane(cnodes[nb + len - 1]);
  const int t = F.tree_of[v];
  const int p = F.tree_prob[t];
  const int K = F.tree_nsites[t];
  const int chunk = w - wv * F.max_chunks;
  if (chunk * 64 >= K) return;
  if (done && done[p]) return;
  const int k = chunk * 64 + (int)(threadIdx.x & 63);
  const bool on = k < K;
  const int sb = F.tree_site_base[t];
  int qo = 0;
  double qs = 0.0;
  if (on) {
    qo = F.site_qoff[sb + k] - 1;
    qs = F.site_qsum[sb + k];
  }
  InCarry c{0, 0, 0.0, 0.0};
  // each step's node record is loaded one step ahead
  NodeRec r = uniform_load(F.nrec + (v - F.num_leaves));
  for (int j = len - 1;; --j) {
    NodeRec rn = r;
    int vn = v;
    if (j > 0) {
      vn = uniform_load(cnodes + nb + j - 1);
      rn = uniform_load(F.nrec + (vn - F.num_leaves));
    }
    InNode n;
    in_node_rec(F, r, n);
    const int hv = prev < 0 ? -1 : (r.ch0 == prev ? 0 : 1);
    if (on) {
      if (hv < 0) {
        chain_in_site<-1, kLeaf>(F, n, k, qo, qs, c);
      } else if (hv == 0) {
        if (n.kind[1] == kLeaf) chain_in_site<0, kLeaf>(F, n, k, qo, qs, c);
        else chain_in_site<0, kInternal>(F, n, k, qo, qs, c);
      } else {
        if (n.kind[0] == kLeaf) chain_in_site<1, kLeaf>(F, n, k, qo, qs, c);
        else chain_in_site<1, kInternal>(F, n, k, qo, qs, c);
      }
    }
    if (j == 0) break;
    prev = v;
    v = vn;
    r = rn;
  }
}

// Inside pass over the top levels [h0, h1): one block per site chunk walks
// the levels in order (its segments share each level's nodes), so the small
// upper levels cost one launch instead of one each.
template <int W>
__global__ void __launch_bounds__(1024) k_inside_top(DevForest F, int h0, int h1, const int32_t* __restrict__ done) {
  const int segs = 1024 / W;
  const int seg = threadIdx.x / W;
  const int k = (int)blockIdx.x * W + (int)(threadIdx.x & (W - 1));
  for (int h = h0; h < h1; ++h) {
    const int e = F.level_off[h + 1];
    for (int v = F.level_off[h] + seg; v < e; v += segs) inside_node(F, v, k, done);
    __syncthreads();
  }
}

// Root of tree t at the sites of one chunk: site likelihood L = P(D | stem
// starts unedited), the outside state at the root, and the root branch's
// posterior counts (partial number `chunk`).  All W lanes of the segment call
// this together.
template <int W>
__device__ __forceinline__ void root_chunk(const DevForest& F, int t, int chunk, int lane,
                                           double* __restrict__ site_llh_out) {
  const int p = F.tree_prob[t];
  const int K = F.tree_nsites[t];
  const int k = chunk * W + lane;
  const int r = F.tree_root[t];
  const bool act = k < K;
  double acc0 = 0, acc1 = 0, acc2 = 0;
  if (act) {
    const int sb = F.tree_site_base[t];
    const Coef cf = F.coef[r];
    const double phi = F.phi[p];
    const bool leaf = r < F.num_leaves;
    const int64_t rr = F.row_off[r];
    int lab, E, code;
    double b0, bA;
    load_child(F, r, leaf, k, phi, lab, b0, bA, E, code);
    const int qo = F.site_qoff[sb + k] - 1;
    const double q = lab > 0 ? F.prior[qo + lab] : F.site_qsum[sb + k];
    const Eps er = edge(lab, b0, bA, E, q, cf);
    const int ex = exp_of(er.e0);
    const double Lm = ldexp(er.e0, -ex);
    const int LE = er.e + ex;
    F.iLm[sb + k] = 1.0 / Lm;
    F.LE[sb + k] = LE;
    const double sl = log(Lm) + (double)LE * 0.69314718055994530942;
    F.site_llh[sb + k] = sl;
    if (site_llh_out) site_llh_out[sb + k] = sl;
    if (!leaf) {
      // outside state at the root: the stem starts unedited
      F.alpha[rr + k] = make_double4(cf.p1n, ldexp(cf.msil, cf.esil), cf.mut, 0.0);
      F.aE[rr + k] = 0;
      F.aG[rr + k] = 0;
    }
    if (lab == 0) {
      acc0 = 1.0;
    } else {
      const double iL = 1.0 / Lm;
      const double f1 = ldexp(iL, min(E - LE, 1000));
      acc0 = cf.p1n * b0 * f1;
      acc1 = cf.mut * q * bA * f1;
      if (lab < 0) acc2 = ldexp(iL * cf.msil, min(cf.esil - LE, 1000));
    }
  }
  acc0 = seg_sum_last<W>(acc0);
  acc1 = seg_sum_last<W>(acc1);
  acc2 = seg_sum_last<W>(acc2);
  if (lane == W - 1) {
    double* c = F.cpart + ((int64_t)r * F.max_chunks + chunk) * 5;
    c[0] = acc0;
    c[1] = acc1;
    c[2] = acc2;
    c[3] = 0.0;
    c[4] = 0.0;
  }
}

// Root step of every tree (segments index trees, blockIdx.y the site chunk).
template <int W>
__global__ void __launch_bounds__(256) k_root(DevForest F, double* __restrict__ site_llh_out,
                                              const int32_t* __restrict__ done) {
  Seg<W> g;
  if (!g.init(0, F.num_trees)) return;
  const int t = g.v;
  if ((int)blockIdx.y * W >= F.tree_nsites[t]) return;
  if (done && done[F.tree_prob[t]]) return;
  root_chunk<W>(F, t, (int)blockIdx.y, g.lane, site_llh_out);
}

// Outside pass for the children of the internal nodes of one level, with the
// posterior expected transition counts of both child branches and the
// dropout responsibilities of missing leaf children.

// per-node values of the outside pass (uniform over the sites of the node)
struct OutNode {
  int K;
  int ch[2];
  int kind[2];
  Coef cf[2];
  ChildRows c[2];
  double4* ca[2];        // children's outside rows (internal children)
  int32_t* caE[2];
  int16_t* caG[2];
  const double4* a;      // the node's outside row
  const int32_t* aE;
  const int16_t* aG;
  const double* iLm;     // the tree's site likelihoods (1 / mantissa, exponent)
  const int32_t* LE;
  const int32_t* qoff;
  const double* qsum;
  double phi;
  bool stream;           // rows with the non-temporal hint
};

template <bool CH>
__device__ __forceinline__ void out_node_load(const DevForest& F, int u, int t, int p, int K, OutNode& n) {
  n.K = K;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int c = i == 0 ? F.child0[u] : F.child1[u];
    n.ch[i] = c;
    n.kind[i] = CH ? child_kind(F, c) : (c < F.num_leaves ? kLeaf : kInternal);
    n.cf[i] = F.coef[c];
    n.c[i] = child_rows(F, c, n.kind[i]);
    const int64_t rc = F.row_off[c];
    n.ca[i] = F.alpha + rc;
    n.caE[i] = F.aE + rc;
    n.caG[i] = F.aG + rc;
  }
  const int64_t ru = F.row_off[u];
  n.a = F.alpha + ru;
  n.aE = F.aE + ru;
  n.aG = F.aG + ru;
  const int sb = F.tree_site_base[t];
  n.iLm = F.iLm + sb;
  n.LE = F.LE + sb;
  n.qoff = F.site_qoff + sb;
  n.qsum = F.site_qsum + sb;
  n.phi = F.phi[p];
  n.stream = F.stream_rows != 0;
}

// The outside-pass values of internal node u from its record (children that
// are leaves or stored internal nodes).
__device__ __forceinline__ void out_node_rec(const DevForest& F, const NodeRec& r, OutNode& n) {
  n.K = r.K;
  const int ch[2] = {r.ch0, r.ch1};
  const int64_t ro[2] = {r.ro0, r.ro1}, bro[2] = {r.bro0, r.bro1};
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    n.ch[i] = ch[i];
    n.kind[i] = ch[i] < F.num_leaves ? kLeaf : kInternal;
    n.cf[i] = uniform_load(F.coef + ch[i]);
    n.c[i].code = (n.kind[i] == kLeaf ? F.codes : F.labels) + ro[i];
    n.c[i].beta = F.beta + bro[i];
    n.c[i].bE = F.bE + bro[i];
    n.c[i].code2 = nullptr;
    n.c[i].gc[0] = n.c[i].gc[1] = -1;
    n.ca[i] = F.alpha + ro[i];
    n.caE[i] = F.aE + ro[i];
    n.caG[i] = F.aG + ro[i];
  }
  n.a = F.alpha + r.ro;
  n.aE = F.aE + r.ro;
  n.aG = F.aG + r.ro;
  n.iLm = F.iLm + r.sb;
  n.LE = F.LE + r.sb;
  n.qoff = F.site_qoff + r.sb;
  n.qsum = F.site_qsum + r.sb;
  n.phi = uniform_load(F.phi + r.p);
  n.stream = F.stream_rows != 0;
}

// The posterior step of one node at one site.  In: the node's outside state
// au 2^Eo with specific state gu, 1/L 2^-LE of the site, and for both
// children their inside state (lab, b0, bA, E), leaf code, prior q, edge
// values ep and branch coefficients cf.  Adds the counts of both child
// branches to acc and the dropout responsibility of missing leaf children to
// rt, and hands every non-leaf child's outside state to sink(i, quad, ec, gg)
// (the quad carries 2^ec; gg is the specific state).
template <typename Sink>
__device__ __forceinline__ void post_site(const double4& au, int Eo, int gu, double iL, int LE, double qs,
                                          double phi, const int (&lab)[2], const double (&b0)[2],
                                          const double (&bA)[2], const int (&E)[2], const int (&code)[2],
                                          const double (&q)[2], const Eps (&ep)[2], const Coef (&cfs)[2],
                                          const bool (&lf)[2], double (&acc)[2][5], double& rt, Sink&& sink) {
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int w = 1 - i;
    const int lw = lab[w];
    const Coef& cf = cfs[i];
    // values below share the exponent ea = Eo + e_w:
    // A = alpha_u(0) eps_w(0) (joint of the data outside v with u = 0);
    // G(a) = alpha_u(a) eps_w(a) = q_a Gy + [a = gg] Gz
    const double A = au.x * ep[w].e0;
    const double Gy = lw < 0 ? au.z * ep[w].eA : 0.0;
    const double Gz = lw < 0 ? au.w * ep[w].eA
                             : (lw > 0 ? fma(q[w], au.z, lw == gu ? au.w : 0.0) * ep[w].eA : 0.0);
    const int gg = lw < 0 ? gu : (lw > 0 ? lw : 0);
    const int ea = Eo + ep[w].e;
    // sum over edit states a of G(a) = q_a Gy + [a = gg] Gz
    const double sumG = fma(qs, Gy, Gz);
    const double av0 = A * cf.p1n;
    const double base = A * cf.mut;
    const double Yv = fma(cf.pn, Gy, base);
    const double Zv = cf.pn * Gz;
    if (!lf[i]) {
      // alpha_v(-1) = GS 2^Eo + sil (A + sumG) 2^ea, GS = alpha_u(-1) eps_w(-1).
      // The stored exponent comes from the largest value of either term
      // group (GS is often zero), so the quad is normalised in one step.
      const double GS = lw < 0 ? au.y : 0.0;
      const double T = ldexp(cf.msil * (A + sumG), cf.esil);
      const double mx = fmax(fmax(av0, T), fmax(Yv, Zv));
      const int ta = exp_of(mx);
      const int tg = GS != 0.0 ? Eo + exp_of(GS) : kZeroTop;
      int ec = max(mx != 0.0 ? ea + ta : kZeroTop, tg);
      if (ec == kZeroTop) ec = Eo;
      const double sa = ldexp(1.0, min(ea - ec, 1000));
      const double sg = ldexp(1.0, min(Eo - ec, 1000));
      sink(i, make_double4(av0 * sa, fma(GS, sg, T * sa), Yv * sa, Zv * sa), ec, gg);
    }
    const int lv = lab[i];
    if (lv == 0) {
      acc[i][0] += 1.0;
    } else {
      // S0, S1, S3 carry 2^(ea + E_v) / L; S2, S4 carry sil 2^ea / L
      const double f1 = ldexp(iL, min(ea + E[i] - LE, 1000));
      acc[i][0] = fma(av0 * b0[i], f1, acc[i][0]);
      acc[i][1] = fma(base * (q[i] * bA[i]), f1, acc[i][1]);
      const double t3 = lv > 0 ? fma(q[i], Gy, lv == gg ? Gz : 0.0) : sumG;
      acc[i][3] = fma(t3 * (cf.pn * bA[i]), f1, acc[i][3]);
      if (lv < 0) {
        const double f2 = ldexp(iL * cf.msil, min(ea + cf.esil - LE, 1000));
        acc[i][2] = fma(A, f2, acc[i][2]);
        acc[i][4] = fma(sumG, f2, acc[i][4]);
        if (lf[i] && code[i] == -1)
          rt = fma(phi * (av0 + fma(qs, Yv, Zv)), ldexp(iL, min(ea - LE, 1000)), rt);
      }
    }
  }
}

// The posterior step a fused cherry's own node would run at site k, from the
// outside state its parent just formed (au 2^Eo, gu): counts of its two leaf
// branches into cacc, dropout responsibility of its missing leaves into crt.
__device__ __forceinline__ void cherry_post(const DevForest& F, const ChildRows& r, int k, const double4& au, int Eo,
                                            int gu, double iL, int LE, int qo, double qs, double phi,
                                            double (&cacc)[2][5], double& crt) {
  int lab[2], E[2] = {0, 0}, code[2];
  double b0[2], bA[2], q[2];
  Eps ep[2];
  code[0] = row_at(r.code, k);
  code[1] = row_at(r.code2, k);
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    leaf_beta(code[j], phi, lab[j], b0[j], bA[j]);
    q[j] = lab[j] > 0 ? row_at(F.prior, (unsigned)(qo + lab[j])) : qs;
    ep[j] = edge(lab[j], b0[j], bA[j], 0, q[j], r.g[j]);
  }
  const bool lf[2] = {true, true};
  post_site(au, Eo, gu, iL, LE, qs, phi, lab, b0, bA, E, code, q, ep, r.g, lf, cacc, crt,
            [](int, const double4&, int, int) {});
}

// The outside step of one site in three phases: row loads, prior gathers,
// then arithmetic and stores.
struct OutSite {
  double4 au;
  int Eo, gu, LE, qo;
  double iL, qs;
  int lab[2], E[2], code[2];
  double b0[2], bA[2], q[2];
};

// the children's rows at site k (x.qo, x.qs already loaded)
template <int K0, int K1>
__device__ __forceinline__ void outside_load_children(const DevForest& F, const OutNode& n, int k, OutSite& x) {
  const int kind[2] = {resolve_kind<K0>(n.kind[0]), resolve_kind<K1>(n.kind[1])};
#pragma unroll
  for (int i = 0; i < 2; ++i)
    load_child_rows(n.c[i], kind[i], k, n.phi, F.prior, x.qo, x.qs, x.lab[i], x.b0[i], x.bA[i], x.E[i], x.code[i]);
}

template <int K0, int K1>
__device__ __forceinline__ void outside_load(const DevForest& F, const OutNode& n, int k, OutSite& x) {
  x.au = row_at(n.a, k);
  x.Eo = row_at(n.aE, k);
  x.gu = row_at(n.aG, k);
  x.iL = row_at(n.iLm, k);
  x.LE = row_at(n.LE, k);
  x.qo = row_at(n.qoff, k) - 1;
  x.qs = row_at(n.qsum, k);
  outside_load_children<K0, K1>(F, n, k, x);
}

__device__ __forceinline__ void outside_gather(const DevForest& F, OutSite& x) {
#pragma unroll
  for (int i = 0; i < 2; ++i) x.q[i] = x.lab[i] > 0 ? row_at(F.prior, (unsigned)(x.qo + x.lab[i])) : x.qs;
}

// One site k of node n: the children's outside states and the posterior
// counts of both child branches, added to acc; the dropout responsibility of
// missing leaf children added to rt.  A fused cherry child's outside state is
// not stored: the cherry's own posterior step runs here, into cacc[i] and
// crt[i].  K0 / K1: the kind of each child, or kNoCherry to read n.kind.
template <int K0, int K1>
__device__ __forceinline__ void outside_finish(const DevForest& F, const OutNode& n, int k, const OutSite& x,
                                               double (&acc)[2][5], double& rt, double (&cacc)[2][2][5],
                                               double (&crt)[2]) {
  const int kind[2] = {resolve_kind<K0>(n.kind[0]), resolve_kind<K1>(n.kind[1])};
  const bool lf[2] = {kind[0] == kLeaf, kind[1] == kLeaf};
  const double iL = x.iL;
  Eps ep[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) ep[i] = edge(x.lab[i], x.b0[i], x.bA[i], x.E[i], x.q[i], n.cf[i]);
  post_site(x.au, x.Eo, x.gu, iL, x.LE, x.qs, n.phi, x.lab, x.b0, x.bA, x.E, x.code, x.q, ep, n.cf, lf, acc, rt,
            [&](int i, const double4& a, int ec, int gg) {
              if (kind[i] == kCherry) {
                cherry_post(F, n.c[i], k, a, ec, gg, iL, x.LE, x.qo, x.qs, n.phi, cacc[i], crt[i]);
              } else {
                if (n.stream)
                  store_stream(&row_at(n.ca[i], k), a);
                else
                  row_at(n.ca[i], k) = a;
                row_at(n.caE[i], k) = ec;
                row_at(n.caG[i], k) = (int16_t)gg;
              }
            });
}

template <int K0, int K1>
__device__ __forceinline__ void outside_site(const DevForest& F, const OutNode& n, int k, double (&acc)[2][5],
                                             double& rt, double (&cacc)[2][2][5], double (&crt)[2]) {
  OutSite x;
  outside_load<K0, K1>(F, n, k, x);
  outside_gather(F, x);
  outside_finish<K0, K1>(F, n, k, x, acc, rt, cacc, crt);
}
// the same with the children's prior weights read from their prior-weight
// rows q0, q1, for nodes without fused cherry children
__device__ __forceinline__ void outside_site_q(const DevForest& F, const OutNode& n, int k, double (&acc)[2][5],
                                               double& rt, const double* q0, const double* q1) {
  double cacc[2][2][5];
  double crt[2];
  OutSite x;
  outside_load<kNoCherry, kNoCherry>(F, n, k, x);
  x.q[0] = row_at(q0, k);
  x.q[1] = row_at(q1, k);
  outside_finish<kNoCherry, kNoCherry>(F, n, k, x, acc, rt, cacc, crt);
}

// outside_site for kernels that never meet a fused cherry
template <int K0, int K1>
__device__ __forceinline__ void outside_site(const DevForest& F, const OutNode& n, int k, double (&acc)[2][5],
                                             double& rt) {
  double cacc[2][2][5];
  double crt[2];
  outside_site<K0, K1>(F, n, k, acc, rt, cacc, crt);
}

// segment sums of a fused cherry's leaf-branch counts and dropout sum for one
// chunk; the last lane of the segment writes them as partial number `slot`
// of the leaf branches and of the cherry.  All lanes of the segment call this.
template <int W>
__device__ __forceinline__ void cherry_write_partials(const DevForest& F, int c, const ChildRows& r, int lane,
                                                      int slot, double (&cacc)[2][5], double crt) {
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int m = 0; m < 5; ++m) cacc[j][m] = seg_sum_last<W>(cacc[j][m]);
  crt = seg_sum_last<W>(crt);
  if (lane == W - 1) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      double* o = F.cpart + ((int64_t)r.gc[j] * F.max_chunks + slot) * 5;
#pragma unroll
      for (int m = 0; m < 5; ++m) o[m] = cacc[j][m];
    }
    F.rt_part[(int64_t)c * F.max_chunks + slot] = crt;
  }
}

// segment sums of the partial counts; they are written as partial number
// `slot` of each child branch (and of the node's dropout sum), and partials
// slot+1 .. slot_end-1 are cleared.  64-lane segments: one butterfly sum of
// all eleven values, each total stored by one lane, the clearing by
// coalesced vector stores; narrower segments: DPP segment sums, stored by
// the segment's last lane.
template <int W>
__device__ __forceinline__ void out_write_partials(const DevForest& F, int u, const int (&ch)[2], int lane, int slot,
                                                   int slot_end, double (&acc)[2][5], double rt) {
  if constexpr (W == 64) {
    double v[12];
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      v[j] = acc[0][j];
      v[5 + j] = acc[1][j];
    }
    v[10] = rt;
    v[11] = 0.0;
    wave_sum12(v);
    // lane 16r + j (j < 3) stores total 3r + j: counts 0..4 of child 0,
    // counts 0..4 of child 1, the dropout sum
    const int r = lane >> 4, j = lane & 15;
    const int vi = 3 * r + j;
    if (j < 3 && vi < 11) {
      const double x = j == 0 ? v[0] : (j == 1 ? v[1] : v[2]);
      if (vi < 10) {
        const int i = vi >= 5 ? 1 : 0;
        F.cpart[((int64_t)ch[i] * F.max_chunks + slot) * 5 + (vi - 5 * i)] = x;
      } else {
        F.rt_part[(int64_t)u * F.max_chunks + slot] = x;
      }
    }
    const int nz = slot_end - slot - 1;
    if (nz > 0) {
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        double* c = F.cpart + ((int64_t)ch[i] * F.max_chunks + slot + 1) * 5;
        for (int b = lane; b < 5 * nz; b += 64) c[b] = 0.0;
      }
      for (int b = lane; b < nz; b += 64) F.rt_part[(int64_t)u * F.max_chunks + slot + 1 + b] = 0.0;
    }
    return;
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 5; ++j) acc[i][j] = seg_sum_last<W>(acc[i][j]);
  rt = seg_sum_last<W>(rt);
  if (lane == W - 1) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      double* c = F.cpart + ((int64_t)ch[i] * F.max_chunks + slot) * 5;
#pragma unroll
      for (int j = 0; j < 5; ++j) c[j] = acc[i][j];
      for (int s2 = slot + 1; s2 < slot_end; ++s2)
#pragma unroll
        for (int j = 0; j < 5; ++j) F.cpart[((int64_t)ch[i] * F.max_chunks + s2) * 5 + j] = 0.0;
    }
    F.rt_part[(int64_t)u * F.max_chunks + slot] = rt;
    for (int s2 = slot + 1; s2 < slot_end; ++s2) F.rt_part[(int64_t)u * F.max_chunks + s2] = 0.0;
  }
}

__device__ __forceinline__ void zero_acc(double (&acc)[2][5]) {
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 5; ++j) acc[i][j] = 0.0;
}

// Segment form: the W lanes of a segment handle sites chunk*W + lane of node
// u (all lanes of the segment call this together).
template <int W>
__device__ __forceinline__ void outside_node(const DevForest& F, int u, int chunk, int lane,
                                             const int32_t* __restrict__ done) {
  const int k = chunk * W + lane;
  const NodeRec r = F.nrec[u - F.num_leaves];
  const int K = r.K;
  if (chunk * W >= K) return;
  if (done && done[r.p]) return;
  OutNode n;
  out_node_rec(F, r, n);
  double acc[2][5];
  zero_acc(acc);
  double rt = 0.0;
  if (k < K) outside_site<kNoCherry, kNoCherry>(F, n, k, acc, rt);
  out_write_partials<W>(F, u, n.ch, lane, chunk, chunk + 1, acc, rt);
}

// Row form (K > 32): one wavefront per node, so the node number is
// wave-uniform (its metadata are scalar loads and the child-kind tests
// scalar branches); the wave walks chunks [c0, c0 + cpw) of 64 sites,
// accumulating per lane, and reduces once.  The leaf-branch counts of a
// fused cherry child are reduced and written per chunk.
template <int K0, int K1>
__device__ __forceinline__ void outside_row_run(const DevForest& F, int u, const OutNode& n, int c0, int c1,
                                                int lane) {
  double acc[2][5];
  zero_acc(acc);
  double rt = 0.0;
  for (int c = c0; c < c1; ++c) {
    const int k = c * 64 + lane;
    double cacc[2][2][5];
    double crt[2] = {0.0, 0.0};
    if (K0 == kCherry) zero_acc(cacc[0]);
    if (K1 == kCherry) zero_acc(cacc[1]);
    if (k < n.K) outside_site<K0, K1>(F, n, k, acc, rt, cacc, crt);
    if (K0 == kCherry) cherry_write_partials<64>(F, n.ch[0], n.c[0], lane, c, cacc[0], crt[0]);
    if (K1 == kCherry) cherry_write_partials<64>(F, n.ch[1], n.c[1], lane, c, cacc[1], crt[1]);
  }
  out_write_partials<64>(F, u, n.ch, lane, c0, c1, acc, rt);
}

// Nodes whose children are leaves or stored internal nodes.
__global__ void __launch_bounds__(256) k_outside_rows(DevForest F, int beg, int cnt, int cpw,
                                                       const int32_t* __restrict__ done,
                                                       const int32_t* __restrict__ nodes) {
  const int u = row_node(beg, cnt, nodes);
  if (u < 0) return;
  const int lane = threadIdx.x & 63;
  const NodeRec r = F.nrec[u - F.num_leaves];
  const int nch = (r.K + 63) >> 6;
  const int c0 = (int)blockIdx.y * cpw;
  if (c0 >= nch) return;
  if (done && done[r.p]) return;
  const int c1 = min(c0 + cpw, nch);
  OutNode n;
  out_node_rec(F, r, n);
  if (n.kind[0] == kLeaf) {
    if (n.kind[1] == kLeaf) outside_row_run<kLeaf, kLeaf>(F, u, n, c0, c1, lane);
    else outside_row_run<kLeaf, kInternal>(F, u, n, c0, c1, lane);
  } else {
    if (n.kind[1] == kLeaf) outside_row_run<kInternal, kLeaf>(F, u, n, c0, c1, lane);
    else outside_row_run<kInternal, kInternal>(F, u, n, c0, c1, lane);
  }
}

// Nodes with at least one fused cherry child.
__global__ void __launch_bounds__(256) k_outside_rows_ch(DevForest F, int beg, int cnt, int cpw,
                                                          const int32_t* __restrict__ done,
                                                          const int32_t* __restrict__ nodes) {
  const int u = row_node(beg, cnt, nodes);
  if (u < 0) return;
  const int lane = threadIdx.x & 63;
  const int t = F.tree_of[u];
  const int p = F.tree_prob[t];
  const int K = F.tree_nsites[t];
  const int nch = (K + 63) >> 6;
  const int c0 = (int)blockIdx.y * cpw;
  if (c0 >= nch) return;
  if (done && done[p]) return;
  const int c1 = min(c0 + cpw, nch);
  OutNode n;
  out_node_load<true>(F, u, t, p, K, n);
  if (n.kind[0] == kCherry) {
    if (n.kind[1] == kCherry) outside_row_run<kCherry, kCherry>(F, u, n, c0, c1, lane);
    else if (n.kind[1] == kLeaf) outside_row_run<kCherry, kLeaf>(F, u, n, c0, c1, lane);
    else outside_row_run<kCherry, kInternal>(F, u, n, c0, c1, lane);
  } else if (n.kind[0] == kLeaf) {
    outside_row_run<kLeaf, kCherry>(F, u, n, c0, c1, lane);
  } else {
    outside_row_run<kInternal, kCherry>(F, u, n, c0, c1, lane);
  }
}

// Outside pass along heavy-path chains (row form, one 64-site chunk per
// wavefront).  A chain is a node followed by its successive heavy children
// (each the child with the larger subtree among the internal ones).  The
// wavefront runs the posterior step of every node of its chain from the top,
// keeping the heavy child's outside state in registers as the next node's
// own: only light children's outside rows are written and only the head's
// is read, which removes most of the outside-row traffic of the level
// passes.  The site's likelihood and prior values are read once per chain.
// Partial counts go to the chunk's slot as in the level kernels with one
// chunk per wavefront, so the results are the same bits.
//
// The steps are software-pipelined so that each waits for memory once, on
// operations issued a whole step earlier.  The inside rows, branch
// coefficients, prior-weight rows and codes of step j + 1 are copied to the
// wavefront's slice of LDS during step j by direct global-to-LDS loads,
// which occupy no registers in flight.  At the top of step j every
// outstanding access completes, step j's operands are read from LDS and the
// step records from the scalar cache, and the next copies are issued; step
// j itself runs on registers.
// No register written by a memory load is carried from one step to the
// next, so the compiler never waits for one mid-step.  The vector memory
// counter counts loads and stores together, so the partial counts of step j
// are stored at the top of step j + 1, after its wait.

// the codes (leaf) or az-partition labels (internal) of child c at row ro
__device__ __forceinline__ const int16_t* chain_code_row(const DevForest& F, int c, int64_t ro) {
  return (c < F.num_leaves ? F.codes : F.labels) + ro;
}
__device__ __forceinline__ const double* chain_q_row(const DevForest& F, int c, int64_t ro) {
  return (c < F.num_leaves ? F.qleaf : F.qint) + ro;
}

typedef __attribute__((address_space(3))) void lds_void;
typedef const __attribute__((address_space(1))) void glb_void;

// One wavefront's LDS slice: the next step's inside rows of both children,
// both children's prior weights, both children's branch coefficients (10
// words each), both children's codes (the aligned 32-bit word holding each
// lane's 16-bit code), and in a pack the next head's outside state.
constexpr int kChainLdsBeta = 0;       // [2][64] double2
constexpr int kChainLdsBE = 2048;      // [2][64] int32
constexpr int kChainLdsQ = 2560;       // [2][64] double
constexpr int kChainLdsCoef = 3584;    // [64] uint32 (20 used)
constexpr int kChainLdsCode = 3840;    // [2][64] uint32
constexpr int kChainLdsHeadA = 4352;   // [2 halves][64] 16 bytes: a head's outside quad
constexpr int kChainLdsHeadE = 6400;   // [64] int32: its exponent
constexpr int kChainLdsHeadG = 6656;   // [64] uint32: the word holding its specific state
constexpr int kChainLdsBytes = 6912;

// copy `size` (4 or 16) bytes at src (per lane) to dst + size * lane; rows
// read once per pass carry the non-temporal hint, so that the per-site
// tables and priors every chain of a tree reads stay in the L2 cache
__device__ __forceinline__ void to_lds(const void* src, char* dst, unsigned size, bool stream = false) {
  if (size == 16) {
    if (stream)
      __builtin_amdgcn_global_load_lds((glb_void*)src, (lds_void*)dst, 16, 0, 2);
    else
      __builtin_amdgcn_global_load_lds((glb_void*)src, (lds_void*)dst, 16, 0, 0);
  } else {
    if (stream)
      __builtin_amdgcn_global_load_lds((glb_void*)src, (lds_void*)dst, 4, 0, 2);
    else
      __builtin_amdgcn_global_load_lds((glb_void*)src, (lds_void*)dst, 4, 0, 0);
  }
}
template <typename T>
__device__ __forceinline__ T lds_at(const char* base, unsigned byte_off) {
  return *(const __attribute__((address_space(3))) T*)(base + byte_off);
}
template <>
__device__ __forceinline__ double2 lds_at<double2>(const char* base, unsigned byte_off) {
  const nt_double2 v = *(const __attribute__((address_space(3))) nt_double2*)(base + byte_off);
  return make_double2(v.x, v.y);
}

// a wave-uniform double in scalar registers
__device__ __forceinline__ double uniform_double(double x) {
  const int lo = __builtin_amdgcn_readfirstlane(__double2loint(x));
  const int hi = __builtin_amdgcn_readfirstlane(__double2hiint(x));
  return __hiloint2double(hi, lo);
}

// every memory access issued so far has completed; nothing is moved across
__device__ __forceinline__ void chain_wait_all() {
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_sched_barrier(0);
}
// every LDS read and scalar load issued so far has completed
__device__ __forceinline__ void chain_wait_lgkm() {
  __builtin_amdgcn_s_waitcnt(0xc07f);
  __builtin_amdgcn_sched_barrier(0);
}

// the 16-bit code at site kk of a row, from the aligned word that holds it
__device__ __forceinline__ int code_of_word(uint32_t wd, unsigned kk) {
  return (int)(int16_t)(kk & 1u ? wd >> 16 : wd & 0xffffu);
}

// Copy step r's codes.
__device__ __forceinline__ void chain_stage_codes(const DevForest& F, const ChainStep& r, unsigned kk, char* lds) {
  const int16_t* row0 = chain_code_row(F, r.ch0, r.ro0) + (kk & ~1u);
  const int16_t* row1 = chain_code_row(F, r.ch1, r.ro1) + (kk & ~1u);
  to_lds(row0, lds + kChainLdsCode, 4, true);
  to_lds(row1, lds + kChainLdsCode + 256, 4, true);
}

// Copy step r's operands: branch coefficients of both children, inside rows
// of internal children, and both children's prior-weight rows over the
// chunk's 64 sites (k0 .. k0 + 63, inside the padded rows; 16 bytes by each
// of the first 32 lanes).
__device__ __forceinline__ void chain_stage(const DevForest& F, const ChainStep& r, unsigned k0, unsigned kk,
                                            int lane, char* lds) {
  {
    const int c = lane < 10 ? r.ch0 : r.ch1;
    const int wd = lane < 10 ? lane : min(lane - 10, 9);
    to_lds(reinterpret_cast<const uint32_t*>(F.coef + c) + wd, lds + kChainLdsCoef, 4);
  }
  const int ch[2] = {r.ch0, r.ch1};
  const int64_t ro[2] = {r.ro0, r.ro1};
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const bool leaf = ch[i] < F.num_leaves;
    if (!leaf) {
      to_lds(&row_at(F.beta + ro[i], kk), lds + kChainLdsBeta + 1024 * i, 16, true);
      to_lds(&row_at(F.bE + ro[i], kk), lds + kChainLdsBE + 256 * i, 4, true);
    }
    if (lane < 32) to_lds(chain_q_row(F, ch[i], ro[i]) + k0 + 2 * lane, lds + kChainLdsQ + 512 * i, 16, true);
  }
}

// Copy the stored outside state (quad, exponent, specific state) of the node
// whose rows start at ru: the head of the next chain of a pack.
__device__ __forceinline__ void chain_stage_head(const DevForest& F, int64_t ru, unsigned kk, char* lds) {
  const char* a = reinterpret_cast<const char*>(&row_at(F.alpha + ru, kk));
  to_lds(a, lds + kChainLdsHeadA, 16, true);
  to_lds(a + 16, lds + kChainLdsHeadA + 1024, 16, true);
  to_lds(&row_at(F.aE + ru, kk), lds + kChainLdsHeadE, 4, true);
  to_lds(F.aG + ru + (kk & ~1u), lds + kChainLdsHeadG, 4, true);
}

// One chain step at one site, on registers: the children's raw row values
// (leaf codes or labels), prior weights, inside rows (ignored for leaves)
// and branch coefficients.  The heavy child's outside state replaces (au,
// Eo, gu); the light internal child's is stored at row ro[light].
template <int K0, int K1, typename Stage>
__device__ __forceinline__ void chain_compute(const DevForest& F, const char* lds, int lane, const Coef (&cf)[2],
                                              double phi, int hv, double iL, int LE, double qs, const int (&raw)[2],
                                              const int64_t (&ro)[2], unsigned kk, bool on, bool stream,
                                              Stage&& stage, double4& au, int& Eo, int& gu, double (&acc)[2][5],
                                              double& rt) {
  const bool lf[2] = {K0 == kLeaf, K1 == kLeaf};
  int lab[2], E[2], code[2];
  double b0[2], bA[2], q[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    q[i] = lds_at<double>(lds, kChainLdsQ + 512 * i + 8 * lane);
    if (lf[i]) {
      code[i] = raw[i];
      leaf_beta(raw[i], phi, lab[i], b0[i], bA[i]);
      E[i] = 0;
    } else {
      code[i] = 0;
      lab[i] = raw[i];
      const double2 bb = lds_at<double2>(lds, kChainLdsBeta + 1024 * i + 16 * lane);
      b0[i] = bb.x;
      bA[i] = bb.y;
      E[i] = lds_at<int>(lds, kChainLdsBE + 256 * i + 4 * lane);
    }
  }
  // the slice is read: the next step's copies may overwrite it
  chain_wait_lgkm();
  stage();
  if (!on) return;
  Eps ep[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) ep[i] = edge(lab[i], b0[i], bA[i], E[i], q[i], cf[i]);
  const double4 au0 = au;
  const int Eo0 = Eo, gu0 = gu;
  post_site(au0, Eo0, gu0, iL, LE, qs, phi, lab, b0, bA, E, code, q, ep, cf, lf, acc, rt,
            [&](int i, const double4& a, int ec, int gg) {
              if (i == hv) {
                au = a;
                Eo = ec;
                gu = gg;
              } else {
                if (stream)
                  store_stream(&row_at(F.alpha + ro[i], kk), a);
                else
                  row_at(F.alpha + ro[i], kk) = a;
                row_at(F.aE + ro[i], kk) = ec;
                row_at(F.aG + ro[i], kk) = (int16_t)gg;
              }
            });
}

// Packs [cbeg, cbeg + cnt) of a chain plan: pack c's steps are
// steps[packs[c].x .. packs[c].x + packs[c].y), one or more chains of one
// tree back to back, each head first; a head's node id is stored
// complemented (~u) and head_ro holds its row offset.  Wavefront w runs pack
// w / max_chunks at chunk w % max_chunks, so that every chunk of the
// (longest, first) packs starts first.  The first head's outside state is
// loaded in the prologue; a later head's is staged into the slice during the
// step before it, like every other operand.  Lanes past the tree's sites
// read its last site (every address stays inside the rows) and neither
// store nor add to the counts.  Grid: ceil(cnt * max_chunks / 4) blocks of
// 256; the heads' outside rows must be complete.
template <bool kPacked>
__global__ void __launch_bounds__(256) k_outside_chain(DevForest F, const int2* __restrict__ packs, int cbeg,
                                                        int cnt, const ChainStep* __restrict__ steps,
                                                        const int64_t* __restrict__ head_ro,
                                                        const int32_t* __restrict__ done) {
  __shared__ __align__(16) char lds_all[4 * kChainLdsBytes];
  const int w = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
  const int wv = w / F.max_chunks;
  if (wv >= cnt) return;
  const int2 cd = uniform_load(packs + cbeg + wv);
  const int nb = cd.x, len = cd.y;
  const ChainStep* cs = steps + nb;
  const int64_t* hro = head_ro + nb;
  ChainStep r0 = uniform_load(cs);
  const int u0 = r0.u < 0 ? ~r0.u : r0.u;
  const int t = F.tree_of[u0];
  const int p = F.tree_prob[t];
  const int K = F.tree_nsites[t];
  const int chunk = w - wv * F.max_chunks;
  if (chunk * 64 >= K) return;
  if (done && done[p]) return;
  char* lds = lds_all + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) * kChainLdsBytes;
  const int lane = threadIdx.x & 63;
  const int k = chunk * 64 + lane;
  const bool on = k < K;
  const unsigned kk = on ? k : K - 1;
  const int sb = F.tree_site_base[t];
  const double phi = F.phi[p];
  const bool stream = F.stream_rows != 0;
  const double iL = row_at(F.iLm + sb, kk);
  const int LE = row_at(F.LE + sb, kk);
  const double qs = row_at(F.site_qsum + sb, kk);
  const unsigned k0 = (unsigned)chunk * 64u;
  const int64_t ru = F.row_off[u0];
  double4 au = row_at(F.alpha + ru, kk);
  int Eo = row_at(F.aE + ru, kk);
  int gu = row_at(F.aG + ru, kk);

  // prologue: step 0's operands and codes and step 1's codes staged
  {
    chain_stage_codes(F, r0, kk, lds);
    chain_stage(F, r0, k0, kk, lane, lds);
  }
  int touch = 0;  // first word of the record four steps ahead (brings its line into the scalar cache)
  if (len > 2) touch = uniform_load(&cs[min(len - 1, 4)].u);

  double* pp = nullptr;  // partial count held back from the previous step
  double pv = 0.0;
  for (int j = 0;;) {
    chain_wait_all();
    asm volatile("" ::"s"(touch));
    if (pp) *pp = pv;
    const bool more = j + 1 < len;
    r0 = uniform_load(cs + j);
    const ChainStep r1 = uniform_load(cs + (more ? j + 1 : j));
    const bool head1 = kPacked && more && r1.u < 0;
    const int64_t ru1 = head1 ? uniform_load(hro + j + 1) : 0;
    if (j + 4 < len) touch = uniform_load(&cs[j + 4].u);
    // a later chain's head: its outside state, staged during the last step
    const bool head0 = r0.u < 0;
    const int u = head0 ? ~r0.u : r0.u;
    if (kPacked && j > 0 && head0) {
      const double2 xy = lds_at<double2>(lds, kChainLdsHeadA + 16 * lane);
      const double2 zw = lds_at<double2>(lds, kChainLdsHeadA + 1024 + 16 * lane);
      au = make_double4(xy.x, xy.y, zw.x, zw.y);
      Eo = lds_at<int>(lds, kChainLdsHeadE + 4 * lane);
      gu = code_of_word(lds_at<uint32_t>(lds, kChainLdsHeadG + 4 * lane), kk);
    }
    // this step's operands, from the slice (wave-uniform: moved to scalar
    // registers)
    Coef cf[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      cf[i].p1n = uniform_double(lds_at<double>(lds, kChainLdsCoef + 40 * i));
      cf[i].mut = uniform_double(lds_at<double>(lds, kChainLdsCoef + 40 * i + 8));
      cf[i].pn = uniform_double(lds_at<double>(lds, kChainLdsCoef + 40 * i + 16));
      cf[i].msil = uniform_double(lds_at<double>(lds, kChainLdsCoef + 40 * i + 24));
      cf[i].esil = __builtin_amdgcn_readfirstlane(lds_at<int>(lds, kChainLdsCoef + 40 * i + 32));
      cf[i].pad = 0;
    }
    const int raw[2] = {code_of_word(lds_at<uint32_t>(lds, kChainLdsCode + 4 * lane), kk),
                        code_of_word(lds_at<uint32_t>(lds, kChainLdsCode + 256 + 4 * lane), kk)};
    // operands and codes of step j + 1, issued once this step's are read
    // from the slice
    auto stage = [&]() {
      if (more) {
        chain_stage(F, r1, k0, kk, lane, lds);
        chain_stage_codes(F, r1, kk, lds);
        if (head1) chain_stage_head(F, ru1, kk, lds);
      }
    };
    // step j
    double acc[2][5];
    zero_acc(acc);
    double rt = 0.0;
    const int hv = r0.hv;
    const int64_t ro[2] = {r0.ro0, r0.ro1};
    if (r0.ch0 < F.num_leaves) {
      if (r0.ch1 < F.num_leaves)
        chain_compute<kLeaf, kLeaf>(F, lds, lane, cf, phi, hv, iL, LE, qs, raw, ro, kk, on, stream, stage, au, Eo,
                                    gu, acc, rt);
      else
        chain_compute<kLeaf, kInternal>(F, lds, lane, cf, phi, hv, iL, LE, qs, raw, ro, kk, on, stream, stage, au,
                                        Eo, gu, acc, rt);
    } else {
      if (r0.ch1 < F.num_leaves)
        chain_compute<kInternal, kLeaf>(F, lds, lane, cf, phi, hv, iL, LE, qs, raw, ro, kk, on, stream, stage, au,
                                        Eo, gu, acc, rt);
      else
        chain_compute<kInternal, kInternal>(F, lds, lane, cf, phi, hv, iL, LE, qs, raw, ro, kk, on, stream, stage,
                                            au, Eo, gu, acc, rt);
    }
    // partial counts: one butterfly sum of the eleven values; lane 16r + i
    // (i < 3) holds total 3r + i (counts 0..4 of child 0, of child 1, the
    // dropout sum), stored at the top of the next step
    double v[12];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
      v[m] = acc[0][m];
      v[5 + m] = acc[1][m];
    }
    v[10] = rt;
    v[11] = 0.0;
    wave_sum12(v);
    {
      const int rr = lane >> 4, i = lane & 15;
      const int vi = 3 * rr + i;
      pv = i == 0 ? v[0] : (i == 1 ? v[1] : v[2]);
      pp = nullptr;
      if (i < 3 && vi < 10) {
        const int c = vi >= 5 ? 1 : 0;
        pp = F.cpart + ((int64_t)(c ? r0.ch1 : r0.ch0) * F.max_chunks + chunk) * 5 + (vi - 5 * c);
      } else if (i < 3 && vi == 10) {
        pp = F.rt_part + (int64_t)u * F.max_chunks + chunk;
      }
    }
    if (!more) break;
    ++j;
  }
  if (pp) *pp = pv;
}

template <int W>
__global__ void __launch_bounds__(256) k_outside(DevForest F, int beg, int cnt, const int32_t* __restrict__ done) {
  Seg<W> g;
  if (!g.init(beg, cnt)) return;
  outside_node<W>(F, g.v, (int)blockIdx.y, g.lane, done);
}

// Outside pass over the top levels, from level h1-1 down to h0, one block per
// site chunk (see k_inside_top).
template <int W>
__global__ void __launch_bounds__(1024) k_outside_top(DevForest F, int h0, int h1, const int32_t* __restrict__ done) {
  const int segs = 1024 / W;
  const int seg = threadIdx.x / W;
  const int lane = threadIdx.x & (W - 1);
  for (int h = h1 - 1; h >= h0; --h) {
    const int e = F.level_off[h + 1];
    for (int u = F.level_off[h] + seg; u < e; u += segs) outside_node<W>(F, u, (int)blockIdx.x, lane, done);
    __syncthreads();
  }
}



// threads per tree-walk block (16 wavefronts share each level of the tree)
constexpr int kWalkThreads = 1024;

// Tree walk: one block per (tree, chunk of W sites) runs the whole pass of
// that tree at those sites -- inside levels bottom-up, the root step, outside
// levels top-down -- with block barriers between levels.  Sites of one chunk
// never depend on other chunks, so blocks are independent: a forest of many
// trees of uneven sizes is one launch per pass, and each block stops at its
// own tree's height.  phases: 1 inside, 2 root, 4 outside (any combination,
// in that order).  Levels below h_lo are left to level launches (the hybrid
// pass: the walk covers the small top levels of every tree at once).
template <int W>
__global__ void __launch_bounds__(kWalkThreads) k_tree_walk(DevForest F, TreeWalk TW, int phases, int h_lo,
                                                            double* __restrict__ site_llh_out,
                                                            const int32_t* __restrict__ done) {
  const int t = TW.blk_tree[blockIdx.x];
  const int chunk = TW.blk_chunk[blockIdx.x];
  const int p = F.tree_prob[t];
  if (done && done[p]) return;
  const int K = F.tree_nsites[t];
  constexpr int kSegs = kWalkThreads / W;
  const int seg = (int)(threadIdx.x / W);
  const int lane = threadIdx.x & (W - 1);
  const int k = chunk * W + lane;
  const int lb = TW.lvl_base[t];
  const int H = TW.lvl_base[t + 1] - lb;
  const int hb = max(1, h_lo);
  if (phases & 1) {
    for (int h = hb; h < H; ++h) {
      const int beg = uniform_load(TW.lvl_beg + lb + h), cnt = uniform_load(TW.lvl_cnt + lb + h);
      for (int i = seg; i < cnt; i += kSegs) {
        const int v = W == 64 ? __builtin_amdgcn_readfirstlane(beg + i) : beg + i;
        if (k < K) {
          InNode n;
          const NodeRec r = uniform_load(F.nrec + (v - F.num_leaves));
          in_node_rec(F, r, n);
          if constexpr (W == 64)
            inside_site_q(F, n, k, chain_q_row(F, r.ch0, r.ro0), chain_q_row(F, r.ch1, r.ro1));
          else
            inside_site<kNoCherry, kNoCherry>(F, n, k);
        }
      }
      __syncthreads();
    }
  }
  if (phases & 2) {
    if (seg == 0) root_chunk<W>(F, t, chunk, lane, site_llh_out);
    __syncthreads();
  }
  if (phases & 4) {
    for (int h = H - 1; h >= hb; --h) {
      const int beg = uniform_load(TW.lvl_beg + lb + h), cnt = uniform_load(TW.lvl_cnt + lb + h);
      for (int i = seg; i < cnt; i += kSegs) {
        const int u = W == 64 ? __builtin_amdgcn_readfirstlane(beg + i) : beg + i;
        OutNode n;
        const NodeRec r = uniform_load(F.nrec + (u - F.num_leaves));
        out_node_rec(F, r, n);
        double acc[2][5];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int j = 0; j < 5; ++j) acc[a][j] = 0.0;
        double rt = 0.0;
        if constexpr (W == 64) {
          if (k < K) outside_site_q(F, n, k, acc, rt, chain_q_row(F, r.ch0, r.ro0), chain_q_row(F, r.ch1, r.ro1));
        } else {
          if (k < K) outside_site<kNoCherry, kNoCherry>(F, n, k, acc, rt);
        }
        out_write_partials<W>(F, u, n.ch, lane, chunk, chunk + 1, acc, rt);
      }
      __syncthreads();
    }
  }
}

// Per-node counts: fixed-order sum of the per-chunk partials, one thread per
// (node, count), so that a wavefront's loads of one chunk slot cover a few
// whole partial records instead of one value per 200-byte record.
__global__ void __launch_bounds__(256) k_counts(DevForest F, double* __restrict__ counts) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)F.num_nodes * 5) return;
  const int v = (int)(idx / 5);
  const int j = (int)(idx - (int64_t)v * 5);
  const int C = (F.tree_nsites[F.tree_of[v]] + F.chunk_w - 1) / F.chunk_w;
  const double* src = F.cpart + (int64_t)v * F.max_chunks * 5 + j;
  double s = 0.0;
  for (int c = 0; c < C; ++c) s += src[c * 5];
  counts[idx] = s;
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
// bookkeeping of the device loop when em.done != nullptr.
// Per-problem log-likelihood; inside the EM loop also the convergence test
// of EM_solver.EM_optimization (EM_solver.py:483-512).  The EM step number of
// each problem lives on the device (em.step), so that a captured EM step can
// be replayed unchanged.
__global__ void __launch_bounds__(256) k_llh(DevForest F, double* __restrict__ nll, EmState em, int max_iter,
                                             double tol) {
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
  const int step = em.step[p];
  em.step[p] = step + 1;
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

constexpr int kRedThreads = 1024;

// adds node v's per-chunk dropout sums, chunks 0 .. C-1 in order, to s; up
// to 8 chunks are loaded together before the additions
__device__ __forceinline__ void add_node_rt(const DevForest& F, int64_t v, int C, double& s) {
  constexpr int kC = 8;
  const double* src = F.rt_part + v * F.max_chunks;
  if (C <= kC) {
    double x[kC];
#pragma unroll
    for (int c = 0; c < kC; ++c) x[c] = c < C ? src[c] : 0.0;
#pragma unroll
    for (int c = 0; c < kC; ++c)
      if (c < C) s += x[c];
  } else {
    for (int c = 0; c < C; ++c) s += src[c];
  }
}

// sum over the problem's internal nodes of the per-node dropout sums
__device__ __forceinline__ double problem_rtilde(const DevForest& F, int p, double* sh) {
  double s = 0.0;
  const int b = F.prob_int_off[p], e = F.prob_node_off[p + 1];
  const int C = (F.tree_nsites[F.prob_tree_off[p]] + F.chunk_w - 1) / F.chunk_w;
  for (int i = b + threadIdx.x; i < e; i += kRedThreads) add_node_rt(F, F.prob_nodes[i], C, s);
  return block_sum<kRedThreads>(s, sh);
}

// Sum of the dropout responsibilities of missing leaf entries of a problem.
__global__ void __launch_bounds__(kRedThreads) k_rtilde(DevForest F, double* __restrict__ rtilde,
                                                        const int32_t* __restrict__ done) {
  __shared__ double sh[kRedThreads / 64];
  const int p = blockIdx.x;
  if (done && done[p]) return;
  const double rt = problem_rtilde(F, p, sh);
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
// argmax over [dmin, dmax] of the strictly concave branch objective.  Its
// derivative g behaves like S/d - A for small d, so Newton runs on the nearly
// linear d*g(d) (same root and sign), safeguarded by a bracket.
__device__ double solve_branch(const BranchObj& o, double dmin, double dmax, double x0) {
  double g, h;
  branch_g(o, dmin, g, h);
  if (!(g > 0.0)) return dmin;
  branch_g(o, dmax, g, h);
  if (!(g < 0.0)) return dmax;
  double lo = dmin, hi = dmax;
  double x = (x0 > lo && x0 < hi) ? x0 : 0.5 * (lo + hi);
  for (int it = 0; it < 100; ++it) {
    branch_g(o, x, g, h);
    if (g > 0.0) lo = x; else if (g < 0.0) hi = x; else return x;
    const double step = x * g / (g + x * h);
    // stop once the Newton step is at the rounding-noise level of g
    if (fabs(step) <= 1e-13 * x) return fmin(fmax(x - step, dmin), dmax);
    double xn = x - step;
    if (!(xn > lo && xn < hi)) xn = 0.5 * (lo + hi);
    if (hi - lo <= 1e-13 * hi) return xn;
    x = xn;
  }
  return x;
}

// Branch lengths of every free branch (EM_solver.py:328-366): counts are
// clamped at eps_s and renormalised to sum K, then the per-branch concave
// problem is solved exactly with the current nu.  Also stages the nu
// problem's per-branch coefficients at the branch's position in its problem.
__global__ void __launch_bounds__(64) k_branch(DevForest F, const double* __restrict__ counts,
                                                double* __restrict__ d, const double* __restrict__ nu,
                                                const int32_t* __restrict__ done, UltraDev U, int ultra) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= F.num_nodes) return;
  const int t = F.tree_of[v];
  const int p = F.tree_prob[t];
  if (done && done[p]) return;
  const int pos = F.node_pos[v];
  if (!F.free_branch[v]) {
    F.nu_a[pos] = 0.0;
    F.nu_b[pos] = 0.0;
    F.nu_d[pos] = d[v];
    return;
  }
  const int Kt = F.tree_nsites[t];
  const double K = (double)Kt;
  double s[5];
  if (counts) {
#pragma unroll
    for (int j = 0; j < 5; ++j) s[j] = counts[(int64_t)v * 5 + j];
  } else {
    const int C = (Kt + F.chunk_w - 1) / F.chunk_w;
#pragma unroll
    for (int j = 0; j < 5; ++j) s[j] = 0.0;
    for (int c = 0; c < C; ++c)
#pragma unroll
      for (int j = 0; j < 5; ++j) s[j] += F.cpart[((int64_t)v * F.max_chunks + c) * 5 + j];
  }
  double tot = 0.0;
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    s[j] = fmax(kEpsS, s[j]);
    tot += s[j];
  }
#pragma unroll
  for (int j = 0; j < 5; ++j) s[j] = s[j] / tot * K;
  const double nu0 = nu[p];
  BranchObj o;
  o.nu = nu0;
  o.A = (nu0 + 1.0) * s[0] + nu0 * s[1] + nu0 * s[3];
  o.S1 = s[1];
  o.B = nu0 > kEpsNu ? (s[2] + s[4]) : 0.0;
  if (ultra) {
    // the ultrametric solve needs the objective of every free branch; the
    // lengths are set by k_ultra and the nu staging refreshed after it
    U.bA[v] = o.A;
    U.bS1[v] = o.S1;
    U.bB[v] = o.B;
    F.nu_a[pos] = s[0] + s[1] + s[3];
    F.nu_b[pos] = s[2] + s[4];
    return;
  }
  const double dn = solve_branch(o, F.dmin, F.dmax, d[v]);
  // non-finite counts give a non-finite length: reported as the reference's
  // failed branch solve (EM_solver.py:415-418,450-459)
  if (!isfinite(dn)) F.nu_state[p].d_bad = 1;
  d[v] = dn;
  F.nu_a[pos] = s[0] + s[1] + s[3];
  F.nu_b[pos] = s[2] + s[4];
  F.nu_d[pos] = dn;
}

// phi closed form (EM_solver.py:318-326) and nu (EM_solver.py:392-407).
// nu maximises a concave function of one variable over [0, 1] given the new
// branch lengths; its derivative h'(x) = sum_v d_v (b_v / expm1(x d_v) - a_v)
// sums over every branch of the problem.  Each problem's branches are split
// over several blocks; one launch = one Newton evaluation: every block writes
// its partial (h', h''), and the last block of the problem to arrive reduces
// the partials in block order (deterministic) and takes a bracketed Newton
// step.  The first launch also sums the dropout responsibilities for phi.
// A problem still unconverged after the last of these launches is finished
// by k_nu_solve in resume mode (one block per problem).
constexpr int kNuThreads = 256;
constexpr int kNuIters = 5;
// evaluations a nu solve may take in all before its M-step reports failure:
// the bracket halves at least every other evaluation, so this is far more
// than [0, 1] needs to close to 1e-12 relative at any optimum the clamped
// counts allow (nu* >= ~1e-12)
constexpr int kNuMaxEvals = 200;

// One bracketed Newton step on x h'(x), which is nearly linear in x, from the
// evaluation (G, Hs) = (h'(x), h''(x)).  [lo, hi] brackets the root of h'
// (hi > 1 until the bound x = 1 has been evaluated).  Returns true when the
// solve has converged; xn is the next iterate (the answer when converged).
__device__ __forceinline__ bool nu_newton_step(double x, double G, double Hs, double& lo, double& hi, double& xn) {
  if (G > 0.0) lo = x; else if (G < 0.0) hi = x;
  if (G > 0.0 && x >= 1.0) {
    xn = 1.0;  // optimum at the upper bound
    return true;
  }
  if (G == 0.0) {
    xn = x;
    return true;
  }
  const double step = x * G / (G + x * Hs);
  const double top = fmin(hi, 1.0);
  // the sum over branches carries ~1e-14 relative rounding noise near the
  // root; a Newton step below 1e-12 x is converged
  if (fabs(step) <= 1e-12 * x) {
    xn = fmin(fmax(x - step, 0.0), 1.0);
    return true;
  }
  xn = x - step;
  if (xn >= 1.0 && hi > 1.0) xn = 1.0;
  else if (!(xn > lo && xn < top)) xn = 0.5 * (lo + top);
  return top - lo <= 1e-12 * top;
}

// One thread's share (stride kNuThreads from branch position i0) of h'(x)
// and h''(x) over the branch positions [i0 - lane offset, pe) of a span.
// Every nu kernel sums through these two functions and in the same span
// order, so all of them give a problem the same nu bit for bit, whichever
// kernel runs and whatever else the forest holds.
__device__ __forceinline__ void nu_span_terms(const DevForest& F, int pb, int pe, int lt, double x, double& g,
                                              double& h) {
  for (int i = pb + lt; i < pe; i += kNuThreads) {
    const double bb = F.nu_b[i];
    if (bb == 0.0) continue;
    const double dv = F.nu_d[i];
    const double en = expm1(x * dv);
    const double ri = 1.0 / en;
    g += dv * fma(bb, ri, -F.nu_a[i]);
    h -= bb * dv * dv * (en + 1.0) * (ri * ri);
  }
}
__device__ __forceinline__ void rt_span(const DevForest& F, int pb, int pe, int lt, int C, double& rt) {
  for (int i = pb + lt; i < pe; i += kNuThreads) add_node_rt(F, F.prob_nodes[i], C, rt);
}

__global__ void __launch_bounds__(kNuThreads) k_nu_eval(DevForest F, const double* __restrict__ rtilde_in,
                                                        double* __restrict__ nu, double* __restrict__ phi,
                                                        const int32_t* __restrict__ done, int iter) {
  __shared__ double sh[kNuThreads / 64];
  __shared__ int s_last;
  const int blk = blockIdx.x;
  // block -> (problem, part) by binary search over prob_blk_off
  int lo_p = 0, hi_p = F.num_problems;
  while (hi_p - lo_p > 1) {
    const int mid = (lo_p + hi_p) >> 1;
    if (F.prob_blk_off[mid] <= blk) lo_p = mid; else hi_p = mid;
  }
  const int p = lo_p;
  NuState* st = F.nu_state + p;
  if (iter > 0) {
    // set by the first evaluation for problems that are done or keep nu
    // fixed, and by the evaluation that converges: the remaining launches of
    // a captured M-step cost one load
    if (st->fin) return;
  } else if (done && done[p]) {
    if (blk == F.prob_blk_off[p] && threadIdx.x == 0) st->fin = 1;
    return;
  }
  const uint32_t flags = F.prob_flags[p];
  const bool opt_nu = (flags & 2u) && F.prob_nfree[p] > 0;
  const int part = blk - F.prob_blk_off[p];
  const int nparts = F.prob_blk_off[p + 1] - F.prob_blk_off[p];
  const int b = F.prob_node_off[p], e = F.prob_node_off[p + 1];
  const int span = (e - b + nparts - 1) / nparts;
  const int pb = b + part * span, pe = min(e, pb + span);
  // first evaluation point: the current nu (or 1/2 when outside (0, 1))
  const double x = (iter == 0) ? ((nu[p] > 0.0 && nu[p] < 1.0) ? nu[p] : 0.5) : st->x;
  double g = 0.0, h = 0.0, rt = 0.0;
  if (opt_nu) nu_span_terms(F, pb, pe, threadIdx.x, x, g, h);
  if (iter == 0 && (flags & 1u) && !rtilde_in) {
    const int C = (F.tree_nsites[F.prob_tree_off[p]] + F.chunk_w - 1) / F.chunk_w;
    rt_span(F, max(pb, F.prob_int_off[p]), pe, threadIdx.x, C, rt);
  }
  g = block_sum<kNuThreads>(g, sh);
  h = block_sum<kNuThreads>(h, sh);
  rt = block_sum<kNuThreads>(rt, sh);
  if (threadIdx.x == 0) {
    double* pp = F.nu_part + (int64_t)blk * 3;
    pp[0] = g;
    pp[1] = h;
    pp[2] = rt;
    __threadfence();
    const int prev = atomicAdd(&st->count, 1);
    s_last = (prev == nparts - 1);
  }
  __syncthreads();
  if (!s_last || threadIdx.x != 0) return;
  __threadfence();
  st->count = 0;
  double G = 0.0, Hs = 0.0, RT = 0.0;
  for (int q = F.prob_blk_off[p]; q < F.prob_blk_off[p + 1]; ++q) {
    const volatile double* pp = F.nu_part + (int64_t)q * 3;
    G += pp[0];
    Hs += pp[1];
    RT += pp[2];
  }
  if (iter == 0) {
    double phi_star = phi[p];
    if (flags & 1u) {
      const double r = rtilde_in ? rtilde_in[p] : RT;
      phi_star = r / (F.prob_nobs[p] + r);
      if (fabs(phi_star) < F.prob_phimin[p]) phi_star = 0.0;
    }
    phi[p] = phi_star;
    st->status = st->d_bad ? 1 : 0;
    st->d_bad = 0;
    st->lo = 0.0;
    st->hi = 2.0;  // > 1: the bound x = 1 has not been evaluated
    st->fin = opt_nu ? 0 : 1;
    if (!opt_nu) return;
  }
  double lo = st->lo, hi = st->hi, xn;
  const bool fin = nu_newton_step(x, G, Hs, lo, hi, xn);
  st->lo = lo;
  st->hi = hi;
  st->x = xn;
  st->fin = fin ? 1 : 0;
  nu[p] = xn;
}

// The same phi update and nu solve with one block per problem and every
// Newton evaluation in one launch (forests of several problems of up to
// kNuSolveMaxNodes branches).  The block's four groups of kNuThreads threads
// sum four of the problem's spans (those of k_nu_eval's blocks) at a time,
// each span exactly as a k_nu_eval block sums it, and the span sums are added
// in span order, so the result equals k_nu_eval's bit for bit.  resume = 1
// continues the solves that the k_nu_eval launches left unconverged (from
// their Newton state; phi is already done) and does nothing for the others.
// The M-step status of each problem is final after this kernel.
constexpr int kNuSolveThreads = 1024;
constexpr int kNuSolveMaxNodes = 65536;
constexpr int kNuGroups = kNuSolveThreads / kNuThreads;

// sums over the spans of problem p in k_nu_eval's order: (h', h'') at x when
// want_g, the dropout sum when want_rt
__device__ void nu_eval_spans(const DevForest& F, int p, double x, bool want_g, bool want_rt, double& G, double& Hs,
                              double& RT, double* sh) {
  const int nparts = F.prob_blk_off[p + 1] - F.prob_blk_off[p];
  const int b = F.prob_node_off[p], e = F.prob_node_off[p + 1];
  const int span = (e - b + nparts - 1) / nparts;
  const int grp = threadIdx.x / kNuThreads, lt = threadIdx.x % kNuThreads;
  const int wid = threadIdx.x >> 6;
  const int C = want_rt ? (F.tree_nsites[F.prob_tree_off[p]] + F.chunk_w - 1) / F.chunk_w : 0;
  constexpr int kW = kNuThreads / 64;  // waves per span
  G = 0.0;
  Hs = 0.0;
  RT = 0.0;
  for (int q0 = 0; q0 < nparts; q0 += kNuGroups) {
    const int part = q0 + grp;
    double g = 0.0, h = 0.0, rt = 0.0;
    if (part < nparts) {
      const int pb = b + part * span, pe = min(e, pb + span);
      if (want_g) nu_span_terms(F, pb, pe, lt, x, g, h);
      if (want_rt) rt_span(F, max(pb, F.prob_int_off[p]), pe, lt, C, rt);
    }
    g = seg_sum<64>(g);
    h = seg_sum<64>(h);
    rt = seg_sum<64>(rt);
    if ((threadIdx.x & 63) == 0) {
      sh[wid] = g;
      sh[kNuSolveThreads / 64 + wid] = h;
      sh[2 * (kNuSolveThreads / 64) + wid] = rt;
    }
    __syncthreads();
    for (int j = 0; j < kNuGroups && q0 + j < nparts; ++j) {
      double sg = 0.0, shh = 0.0, sr = 0.0;
#pragma unroll
      for (int w = 0; w < kW; ++w) {
        sg += sh[j * kW + w];
        shh += sh[kNuSolveThreads / 64 + j * kW + w];
        sr += sh[2 * (kNuSolveThreads / 64) + j * kW + w];
      }
      G += sg;
      Hs += shh;
      RT += sr;
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kNuSolveThreads) k_nu_solve(DevForest F, const double* __restrict__ rtilde_in,
                                                             double* __restrict__ nu, double* __restrict__ phi,
                                                             const int32_t* __restrict__ done, int resume) {
  __shared__ double sh[3 * (kNuSolveThreads / 64)];
  const int p = blockIdx.x;
  NuState* st = F.nu_state + p;
  if (done && done[p]) return;
  // every thread reads the same flag, written by an earlier launch
  if (resume && st->fin) return;
  const uint32_t flags = F.prob_flags[p];
  const bool opt_nu = (flags & 2u) && F.prob_nfree[p] > 0;
  int status = 0;
  double x, lo, hi;
  int evals;
  bool fin = false;
  if (!resume) {
    x = (nu[p] > 0.0 && nu[p] < 1.0) ? nu[p] : 0.5;
    lo = 0.0;
    hi = 2.0;  // hi > 1: the bound x = 1 has not been evaluated
    // the first evaluation also sums the dropout responsibilities, as
    // k_nu_eval's first launch does
    const bool want_rt = (flags & 1u) && !rtilde_in;
    double G, Hs, RT;
    nu_eval_spans(F, p, x, opt_nu, want_rt, G, Hs, RT, sh);
    if (flags & 1u) {
      const double r = rtilde_in ? rtilde_in[p] : RT;
      double phi_star = r / (F.prob_nobs[p] + r);
      if (fabs(phi_star) < F.prob_phimin[p]) phi_star = 0.0;
      if (threadIdx.x == 0) phi[p] = phi_star;
    }
    status = st->d_bad ? 1 : 0;
    __syncthreads();  // every thread has read d_bad before thread 0 clears it
    if (threadIdx.x == 0) {
      st->d_bad = 0;
      st->status = status;
      st->fin = 1;
    }
    if (!opt_nu) return;
    double xn;
    fin = nu_newton_step(x, G, Hs, lo, hi, xn);
    x = xn;
    evals = 1;
  } else {
    x = st->x;
    lo = st->lo;
    hi = st->hi;
    evals = kNuIters + 1;
    status = st->status;
  }
  for (; evals < kNuMaxEvals && !fin; ++evals) {
    double G, Hs, RT;
    nu_eval_spans(F, p, x, true, false, G, Hs, RT, sh);
    double xn;
    fin = nu_newton_step(x, G, Hs, lo, hi, xn);
    x = xn;
  }
  if (threadIdx.x == 0) {
    nu[p] = x;
    st->x = x;
    st->lo = lo;
    st->hi = hi;
    st->fin = 1;
    st->status = status | (fin ? 0 : 2);
  }
}

// ---------------------------------------------------------------------------
// Gradient of the negative log-likelihood (Fisher identity: the gradient of
// log L is the posterior expectation of the complete-data score).  With raw
// per-branch expected counts S0..S4 from the outside pass:
//   dlogL/dd_v  = -(1+nu) S0 + S1 (1/expm1(d) - nu) + (S2+S4) nu/expm1(nu d) - nu S3
//   dlogL/dnu   = sum_v d_v (-S0 - S1 - S3 + (S2+S4)/expm1(nu d_v))
//   dlogL/dphi  = -n_obs/(1-phi) + Rt/phi
// n_obs counts leaf entries observed as 0 or an edit (an explicit -1 entry
// does not depend on phi); Rt sums P(leaf not silenced | D) over '?' entries.
// ---------------------------------------------------------------------------
__global__ void k_grad_branch(DevForest F, const double* __restrict__ d, const double* __restrict__ nu,
                              double* __restrict__ grad_d) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= F.num_nodes) return;
  const int t = F.tree_of[v];
  const int p = F.tree_prob[t];
  const int C = (F.tree_nsites[t] + F.chunk_w - 1) / F.chunk_w;
  double s[5] = {0, 0, 0, 0, 0};
  for (int c = 0; c < C; ++c)
#pragma unroll
    for (int j = 0; j < 5; ++j) s[j] += F.cpart[((int64_t)v * F.max_chunks + c) * 5 + j];
  const double dv = d[v], x = nu[p];
  double g = -(1.0 + x) * s[0] + s[1] * (1.0 / expm1(dv) - x) - x * s[3];
  if (s[2] + s[4] != 0.0) g += (s[2] + s[4]) * x / expm1(x * dv);
  grad_d[v] = -g;
  F.nu_a[F.node_pos[v]] = s[0] + s[1] + s[3];
  F.nu_b[F.node_pos[v]] = s[2] + s[4];
  F.nu_d[F.node_pos[v]] = dv;
}

__global__ void __launch_bounds__(kRedThreads) k_grad_problem(DevForest F, const double* __restrict__ nu,
                                                              const double* __restrict__ phi,
                                                              double* __restrict__ grad_nu,
                                                              double* __restrict__ grad_phi) {
  __shared__ double sh[kRedThreads / 64];
  const int p = blockIdx.x;
  const double x = nu[p];
  double g = 0.0;
  for (int i = F.prob_node_off[p] + threadIdx.x; i < F.prob_node_off[p + 1]; i += kRedThreads) {
    const double dv = F.nu_d[i];
    const double b = F.nu_b[i];
    g -= dv * F.nu_a[i];
    if (b != 0.0) g += b * dv / expm1(x * dv);
  }
  g = block_sum<kRedThreads>(g, sh);
  const double rt = problem_rtilde(F, p, sh);
  if (threadIdx.x == 0) {
    grad_nu[p] = -g;
    const double ph = phi[p];
    grad_phi[p] = F.prob_nobs_phi[p] / (1.0 - ph) - (rt != 0.0 ? rt / ph : 0.0);
  }
}

// ---------------------------------------------------------------------------
// Ultrametric branch-length M-step (EM_solver.py:349-366 with the equality
// system of EM_solver.py:53-89).  Heights of the variable height classes are
// the unknowns, d_v = t(parent side) - t(child side).  A log-barrier keeps
// every free branch in (dmin, dmax); each Newton system is a weighted tree
// Laplacian over the classes and is solved exactly by eliminating classes
// bottom-up and substituting top-down (two level sweeps).  One block per
// problem; every decision depends on block-reduced values, so control flow is
// uniform across the block.
// ---------------------------------------------------------------------------
constexpr int kUltraThreads = 1024;

template <int NT>
__device__ __forceinline__ double block_min(double x, double* sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x = fmin(x, __shfl_xor(x, off, 64));
  const int wid = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sh[wid] = x;
  __syncthreads();
  double s = sh[0];
#pragma unroll
  for (int i = 1; i < NT / 64; ++i) s = fmin(s, sh[i]);
  __syncthreads();
  return s;
}

__device__ __forceinline__ double ult_height(const UltraDev& U, int c, double off) {
  return c >= 0 ? U.t[c] + off : off;
}

// height at t + a dt (only called with a finite direction)
__device__ __forceinline__ double ult_height_at(const UltraDev& U, int c, double off, double a) {
  return c >= 0 ? fma(a, U.dt[c], U.t[c]) + off : off;
}

// branch objective (barrier excluded): value, first and second derivative
__device__ __forceinline__ void ult_obj(double A, double S1, double B, double nu, double d, double& f, double& g,
                                        double& h) {
  const double em = expm1(d);
  f = -A * d + S1 * log(-expm1(-d));
  g = -A + S1 / em;
  h = -S1 * (em + 1.0) / (em * em);
  if (B > 0.0) {
    const double en = expm1(nu * d);
    f += B * log(-expm1(-nu * d));
    g += B * nu / en;
    h -= B * nu * nu * (en + 1.0) / (en * en);
  }
}

// barrier objective of problem p's free branches at heights t + a dt
__device__ double ult_phi(const DevForest& F, const UltraDev& U, int nb, int ne, double nu0, double mu, double a,
                          double* sh) {
  double s = 0.0;
  for (int i = nb + threadIdx.x; i < ne; i += kUltraThreads) {
    const int v = F.prob_nodes[i];
    if (!F.free_branch[v]) continue;
    const double dv =
        ult_height_at(U, U.up_cls[v], U.up_off[v], a) - ult_height_at(U, U.dn_cls[v], U.dn_off[v], a);
    const double lo = dv - F.dmin, hi = F.dmax - dv;
    if (!(lo > 0.0 && hi > 0.0)) {
      s = -INFINITY;
      continue;
    }
    double f, g, h;
    ult_obj(U.bA[v], U.bS1[v], U.bB[v], nu0, dv, f, g, h);
    s += f + mu * (log(lo) + log(hi));
  }
  return block_sum<kUltraThreads>(s, sh);
}

__global__ void __launch_bounds__(kUltraThreads) k_ultra(DevForest F, UltraDev U, double* __restrict__ d,
                                                         const double* __restrict__ nu,
                                                         const int32_t* __restrict__ done) {
  __shared__ double sh[kUltraThreads / 64];
  const int p = blockIdx.x;
  if ((done && done[p]) || !U.has_free[p]) return;
  const int P = F.num_problems;
  const int nb = F.prob_node_off[p], ne = F.prob_node_off[p + 1];
  const double nu0 = nu[p];
  const double dmin = F.dmin, dmax = F.dmax;
  const double K = (double)F.tree_nsites[F.prob_tree_off[p]];
  const int H = U.num_levels;
  int status = 0;
  double mu = 1e-2 * K;
  const double mu_end = 1e-13 * K;
  for (int outer = 0; outer < 40; ++outer) {
    for (int it = 0; it < 60; ++it) {
      // branch derivatives of the barrier objective
      for (int i = nb + threadIdx.x; i < ne; i += kUltraThreads) {
        const int v = F.prob_nodes[i];
        if (!F.free_branch[v]) continue;
        const double dv = ult_height(U, U.up_cls[v], U.up_off[v]) - ult_height(U, U.dn_cls[v], U.dn_off[v]);
        const double lo = dv - dmin, hi = dmax - dv;
        double f, g, h;
        ult_obj(U.bA[v], U.bS1[v], U.bB[v], nu0, dv, f, g, h);
        U.gp[v] = g + mu * (1.0 / lo - 1.0 / hi);
        U.w[v] = -h + mu * (1.0 / (lo * lo) + 1.0 / (hi * hi));
      }
      __syncthreads();
      // eliminate classes bottom-up.  The eliminated diagonal of class c is
      // D_c = G_c + w(edge to the parent class), where G_c, the conductance of
      // c's subtree to the pinned heights, sums positive terms only: pinned
      // neighbours add w, a child class k adds the series conductance
      // w G_k / (w + G_k).  (Forming D_c as sum(w) - sum(w^2 / D_k) cancels to
      // zero when one branch's barrier weight dominates.)
      for (int lv = 0; lv < H; ++lv) {
        const int cb = U.plo[lv * P + p], ce = U.plo[lv * P + p + 1];
        for (int c = cb + threadIdx.x; c < ce; c += kUltraThreads) {
          double gg = 0.0, G = 0.0, wup = 0.0, rr = 0.0;
          for (int j = U.inc_off[c]; j < U.inc_off[c + 1]; ++j) {
            const int v = U.inc_branch[j];
            const double wv = U.w[v];
            if (U.inc_sign[j] > 0) {
              gg += U.gp[v];
              const int kc = U.dn_cls[v];
              if (kc >= 0) {
                const double Dk = U.Dd[kc];
                G += wv * (U.Gg[kc] / Dk);
                rr += wv * (U.r[kc] / Dk);
              } else {
                G += wv;
              }
            } else {
              gg -= U.gp[v];
              if (U.up_cls[v] >= 0) wup += wv; else G += wv;
            }
          }
          U.g[c] = gg;
          U.Gg[c] = G;
          U.Dd[c] = G + wup;
          U.r[c] = gg + rr;
        }
        __syncthreads();
      }
      // substitute top-down
      for (int lv = H - 1; lv >= 0; --lv) {
        const int cb = U.plo[lv * P + p], ce = U.plo[lv * P + p + 1];
        for (int c = cb + threadIdx.x; c < ce; c += kUltraThreads) {
          const int pc = U.cls_parent[c];
          U.dt[c] = pc < 0 ? U.r[c] / U.Dd[c] : (U.r[c] + U.w[U.cls_pedge[c]] * U.dt[pc]) / U.Dd[c];
        }
        __syncthreads();
      }
      // Newton decrement, largest step keeping every branch inside the box
      double lam = 0.0, step = 0.0;
      const int c0 = U.plo[p], c1 = U.plo[(H - 1) * P + p + 1];
      for (int lv = 0; lv < H; ++lv) {
        const int cb = U.plo[lv * P + p], ce = U.plo[lv * P + p + 1];
        for (int c = cb + threadIdx.x; c < ce; c += kUltraThreads) {
          lam += U.g[c] * U.dt[c];
          step = fmax(step, fabs(U.dt[c]) / (1.0 + fabs(U.t[c])));
        }
      }
      (void)c0;
      (void)c1;
      double amax = 1.0;
      for (int i = nb + threadIdx.x; i < ne; i += kUltraThreads) {
        const int v = F.prob_nodes[i];
        if (!F.free_branch[v]) continue;
        const double dv = ult_height(U, U.up_cls[v], U.up_off[v]) - ult_height(U, U.dn_cls[v], U.dn_off[v]);
        const double dd = (U.up_cls[v] >= 0 ? U.dt[U.up_cls[v]] : 0.0) - (U.dn_cls[v] >= 0 ? U.dt[U.dn_cls[v]] : 0.0);
        if (dd < 0.0) amax = fmin(amax, 0.99 * (dv - dmin) / (-dd));
        else if (dd > 0.0) amax = fmin(amax, 0.99 * (dmax - dv) / dd);
      }
      lam = block_sum<kUltraThreads>(lam, sh);
      step = -block_min<kUltraThreads>(-step, sh);
      amax = block_min<kUltraThreads>(amax, sh);
      if (!isfinite(lam) || !isfinite(step)) {
        status = 2;  // singular Newton system: keep the current heights
        break;
      }
      if (!(lam > 0.0) || step * amax <= 1e-15) break;
      double a = amax;
      const double phi0 = ult_phi(F, U, nb, ne, nu0, mu, 0.0, sh);
      if (lam > 1e-12 * (1.0 + fabs(phi0))) {
        // backtracking (Armijo) while the predicted gain is measurable
        int k = 0;
        for (; k < 40; ++k) {
          const double phi1 = ult_phi(F, U, nb, ne, nu0, mu, a, sh);
          if (phi1 >= phi0 + 0.25 * a * lam) break;
          a *= 0.5;
        }
        if (k == 40) {
          status = 1;
          break;
        }
      }
      for (int lv = 0; lv < H; ++lv) {
        const int cb = U.plo[lv * P + p], ce = U.plo[lv * P + p + 1];
        for (int c = cb + threadIdx.x; c < ce; c += kUltraThreads) U.t[c] += a * U.dt[c];
      }
      __syncthreads();
      if (step * a <= 1e-15) break;
    }
    if (mu <= mu_end) break;
    mu = fmax(mu * 0.1, mu_end);
  }
  for (int i = nb + threadIdx.x; i < ne; i += kUltraThreads) {
    const int v = F.prob_nodes[i];
    if (!F.free_branch[v]) continue;
    const double dv = ult_height(U, U.up_cls[v], U.up_off[v]) - ult_height(U, U.dn_cls[v], U.dn_off[v]);
    d[v] = dv;
    F.nu_d[F.node_pos[v]] = dv;
  }
  if (threadIdx.x == 0) U.status[p] = status;
}

// ---------------------------------------------------------------------------
// Quartet-local EM of NNI candidates (Topology_search.py:242-322 with
// local_brlen_opt): nu, phi and every branch except the five around the move
// are fixed, so the subtrees hanging off the quartet (A and W below u, C below
// v) keep their inside vectors and the data above v keeps its outside state.
// A candidate's EM then needs, per site, only those inputs and the five free
// branches: each EM step is O(K) instead of a pass over the tree.  One
// wavefront per candidate; lanes stride over the tree's sites.
//
// Candidate c (after the move): v has children u and C, u has children A and
// W; p is v's parent and s v's sibling (p = -1: v is the root and its branch
// is the stem from the unedited ancestor).  Free branches in the order
// (A, W, u, C, v); opt bit j set = branch j is optimised.
// ---------------------------------------------------------------------------
namespace {

struct QTop {            // data outside a child branch, at its parent's side
  double A, Gy, Gz, GS, sumG;
  int ea, Eo, gg;
};

struct QAlpha {          // outside state at a node: (o0, oS, Y, Z) 2^E, state g
  double x, y, z, w;
  int E, g;
};

__device__ __forceinline__ QTop q_top(const QAlpha& au, const Eps& ew, int lw, double qw, double qs) {
  QTop t;
  t.A = au.x * ew.e0;
  t.Gy = lw < 0 ? au.z * ew.eA : 0.0;
  t.Gz = lw < 0 ? au.w * ew.eA : (lw > 0 ? fma(qw, au.z, lw == au.g ? au.w : 0.0) * ew.eA : 0.0);
  t.gg = lw < 0 ? au.g : (lw > 0 ? lw : 0);
  t.ea = au.E + ew.e;
  t.GS = lw < 0 ? au.y : 0.0;
  t.sumG = fma(qs, t.Gy, t.Gz);
  t.Eo = au.E;
  return t;
}

__device__ __forceinline__ QAlpha q_child_alpha(const QTop& t, const Coef& cf) {
  const double av0 = t.A * cf.p1n;
  const double Yv = fma(cf.pn, t.Gy, t.A * cf.mut);
  const double Zv = cf.pn * t.Gz;
  const double T = ldexp(cf.msil * (t.A + t.sumG), cf.esil);
  const double mx = fmax(fmax(av0, T), fmax(Yv, Zv));
  const int tg = t.GS != 0.0 ? t.Eo + exp_of(t.GS) : kZeroTop;
  int ec = max(mx != 0.0 ? t.ea + exp_of(mx) : kZeroTop, tg);
  if (ec == kZeroTop) ec = t.Eo;
  const double sa = ldexp(1.0, min(t.ea - ec, 1000));
  const double sg = ldexp(1.0, min(t.Eo - ec, 1000));
  QAlpha a;
  a.x = av0 * sa;
  a.y = fma(t.GS, sg, T * sa);
  a.z = Yv * sa;
  a.w = Zv * sa;
  a.E = ec;
  a.g = t.gg;
  return a;
}

// posterior counts of the branch above a node with label lv and inside
// (b0, bA) 2^E, prior q of its state (qs for '?'), given the top of the
// branch and 1/L = iL 2^-LE
__device__ __forceinline__ void q_post(const QTop& t, const Coef& cf, int lv, double b0, double bA, int E, double q,
                                       double iL, int LE, double* S) {
  if (lv == 0) {
    S[0] += 1.0;
    return;
  }
  const double f1 = ldexp(iL, min(t.ea + E - LE, 1000));
  S[0] = fma(t.A * cf.p1n * b0, f1, S[0]);
  S[1] = fma(t.A * cf.mut * (q * bA), f1, S[1]);
  const double t3 = lv > 0 ? fma(q, t.Gy, lv == t.gg ? t.Gz : 0.0) : t.sumG;
  S[3] = fma(t3 * (cf.pn * bA), f1, S[3]);
  if (lv < 0) {
    const double f2 = ldexp(iL * cf.msil, min(t.ea + cf.esil - LE, 1000));
    S[2] = fma(t.A, f2, S[2]);
    S[4] = fma(t.sumG, f2, S[4]);
  }
}

__device__ __forceinline__ Coef q_coef(double d, double nu) {
  Coef c;
  c.p1n = exp(-(1.0 + nu) * d);
  c.pn = exp(-nu * d);
  c.mut = -expm1(-d) * c.pn;
  int es;
  c.msil = frexp(-expm1(-nu * d), &es);
  c.esil = c.msil != 0.0 ? es : 0;
  c.pad = 0;
  return c;
}

// inside of a node from its two children's edge values; returns the label
__device__ __forceinline__ int q_combine(int la, int lb, const Eps& ea, const Eps& eb, double& B0, double& BA,
                                         int& E) {
  const int lv = combine_label(la, lb);
  const double b0 = ea.e0 * eb.e0;
  const double bA = lv != 0 ? ea.eA * eb.eA : 0.0;
  const int sh = min(-exp_of(fmax(b0, bA)), 1000);
  const double sc = ldexp(1.0, sh);
  B0 = b0 * sc;
  BA = bA * sc;
  E = ea.e + eb.e - sh;
  return lv;
}

}  // namespace

__global__ void __launch_bounds__(256) k_quartet_em(DevForest F, int ncand, const int32_t* __restrict__ qnodes,
                                                    const double* __restrict__ len_in,
                                                    const uint8_t* __restrict__ opt_mask,
                                                    const double* __restrict__ llh_const,
                                                    const double* __restrict__ nu_p,
                                                    const double* __restrict__ phi_p, int max_iter, double tol,
                                                    double* __restrict__ len_out, double* __restrict__ nll_out,
                                                    int32_t* __restrict__ iters_out) {
  const int c = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
  if (c >= ncand) return;
  const int lane = threadIdx.x & 63;
  const int p = qnodes[c * 5 + 0], sib = qnodes[c * 5 + 1];
  const int nA = qnodes[c * 5 + 2], nW = qnodes[c * 5 + 3], nC = qnodes[c * 5 + 4];
  const int t = F.tree_of[nA];
  const int prob = F.tree_prob[t];
  const int K = F.tree_nsites[t];
  const int sb = F.tree_site_base[t];
  const double nu = nu_p[prob], phi = phi_p[prob];
  const uint32_t opt = opt_mask[c];
  double d[5];
#pragma unroll
  for (int j = 0; j < 5; ++j) d[j] = len_in[c * 5 + j];
  const ChildRows rA = child_rows(F, nA, nA < F.num_leaves);
  const ChildRows rW = child_rows(F, nW, nW < F.num_leaves);
  const ChildRows rC = child_rows(F, nC, nC < F.num_leaves);
  const bool lfA = nA < F.num_leaves, lfW = nW < F.num_leaves, lfC = nC < F.num_leaves;
  const bool top_root = p < 0;
  Coef cs{};
  ChildRows rS{};
  bool lfS = false;
  const double4* aP = nullptr;
  const int32_t* aEP = nullptr;
  const int16_t* aGP = nullptr;
  if (!top_root) {
    cs = F.coef[sib];
    lfS = sib < F.num_leaves;
    rS = child_rows(F, sib, lfS);
    const int64_t rp = F.row_off[p];
    aP = F.alpha + rp;
    aEP = F.aE + rp;
    aGP = F.aG + rp;
  }
  const int32_t* qoff = F.site_qoff + sb;
  const double* qsum = F.site_qsum + sb;
  const double kLn2 = 0.69314718055994530942;
  double llh_prev = 0.0, llh = 0.0;
  int iters = max_iter + 1;
  for (int it = 0; it <= max_iter; ++it) {
    Coef cf[5];
#pragma unroll
    for (int j = 0; j < 5; ++j) cf[j] = q_coef(d[j], nu);
    double S[5][5];
#pragma unroll
    for (int j = 0; j < 5; ++j)
#pragma unroll
      for (int m = 0; m < 5; ++m) S[j][m] = 0.0;
    double ll = 0.0;
    for (int k = lane; k < K; k += 64) {
      const int qo = row_at(qoff, (unsigned)k) - 1;
      const double qs = row_at(qsum, (unsigned)k);
      int labA, labW, labC, EA, EW, EC, code;
      double b0A, bAA, b0W, bAW, b0C, bAC;
      load_child_rows(rA, lfA, k, phi, labA, b0A, bAA, EA, code);
      load_child_rows(rW, lfW, k, phi, labW, b0W, bAW, EW, code);
      load_child_rows(rC, lfC, k, phi, labC, b0C, bAC, EC, code);
      const double qA = labA > 0 ? row_at(F.prior, (unsigned)(qo + labA)) : qs;
      const double qW = labW > 0 ? row_at(F.prior, (unsigned)(qo + labW)) : qs;
      const double qC = labC > 0 ? row_at(F.prior, (unsigned)(qo + labC)) : qs;
      const Eps eA = edge(labA, b0A, bAA, EA, qA, cf[0]);
      const Eps eW = edge(labW, b0W, bAW, EW, qW, cf[1]);
      double b0u, bAu;
      int Eu;
      const int labU = q_combine(labA, labW, eA, eW, b0u, bAu, Eu);
      const double qU = labU > 0 ? row_at(F.prior, (unsigned)(qo + labU)) : qs;
      const Eps eU = edge(labU, b0u, bAu, Eu, qU, cf[2]);
      const Eps eC = edge(labC, b0C, bAC, EC, qC, cf[3]);
      double b0v, bAv;
      int Ev;
      const int labV = q_combine(labU, labC, eU, eC, b0v, bAv, Ev);
      const double qV = labV > 0 ? row_at(F.prior, (unsigned)(qo + labV)) : qs;
      // top of v's branch: the unedited ancestor, or p's outside state with
      // the sibling's edge values
      QTop tv;
      if (top_root) {
        tv.A = 1.0;
        tv.Gy = tv.Gz = tv.GS = tv.sumG = 0.0;
        tv.ea = tv.Eo = tv.gg = 0;
      } else {
        int labS, ES;
        double b0S, bAS;
        load_child_rows(rS, lfS, k, phi, labS, b0S, bAS, ES, code);
        const double qS = labS > 0 ? row_at(F.prior, (unsigned)(qo + labS)) : qs;
        const Eps eS = edge(labS, b0S, bAS, ES, qS, cs);
        const double4 a4 = row_at(aP, (unsigned)k);
        QAlpha ap;
        ap.x = a4.x;
        ap.y = a4.y;
        ap.z = a4.z;
        ap.w = a4.w;
        ap.E = row_at(aEP, (unsigned)k);
        ap.g = row_at(aGP, (unsigned)k);
        tv = q_top(ap, eS, labS, qS, qs);
      }
      const QAlpha av = q_child_alpha(tv, cf[4]);
      // site likelihood from v's outside and inside states
      double sumA;
      if (labV > 0) sumA = fma(qV, av.z, labV == av.g ? av.w : 0.0);
      else if (labV < 0) sumA = fma(qs, av.z, av.w);
      else sumA = 0.0;
      const double t1 = fma(av.x, b0v, sumA * bAv);
      const double t2 = labV < 0 ? av.y : 0.0;
      const int e1 = t1 != 0.0 ? Ev + exp_of(t1) : kZeroTop;
      const int e2 = t2 != 0.0 ? exp_of(t2) : kZeroTop;
      const int e = max(e1, e2);
      double Lm = ldexp(t1, min(Ev - e, 1000)) + ldexp(t2, min(-e, 1000));
      const int ex = exp_of(Lm);
      Lm = ldexp(Lm, -ex);
      const int LE = av.E + e + ex;
      ll += log(Lm) + (double)LE * kLn2;
      const double iL = 1.0 / Lm;
      // posteriors: v's branch, then u and C below v, then A and W below u
      q_post(tv, cf[4], labV, b0v, bAv, Ev, qV, iL, LE, S[4]);
      const QTop tu = q_top(av, eC, labC, qC, qs);
      q_post(tu, cf[2], labU, b0u, bAu, Eu, qU, iL, LE, S[2]);
      const QTop tc = q_top(av, eU, labU, qU, qs);
      q_post(tc, cf[3], labC, b0C, bAC, EC, qC, iL, LE, S[3]);
      const QAlpha au = q_child_alpha(tu, cf[2]);
      const QTop ta = q_top(au, eW, labW, qW, qs);
      q_post(ta, cf[0], labA, b0A, bAA, EA, qA, iL, LE, S[0]);
      const QTop tw = q_top(au, eA, labA, qA, qs);
      q_post(tw, cf[1], labW, b0W, bAW, EW, qW, iL, LE, S[1]);
    }
    ll = seg_sum_last<64>(ll);
    ll = __shfl(ll, 63, 64);
    llh = ll + llh_const[c];
    // EM_solver.py:483-512: stop when the relative change is below tol, or
    // after max_iter M-steps (iteration count max_iter + 1)
    if (it > 0 && tol > 0.0 && fabs((llh - llh_prev) / llh_prev) < tol) {
      iters = it;
      break;
    }
    if (it == max_iter) break;
    llh_prev = llh;
    // M-step of the free branches (EM_solver.py:339-366 for one branch)
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      double sj[5];
      double tot = 0.0;
#pragma unroll
      for (int m = 0; m < 5; ++m) {
        const double x = seg_sum_last<64>(S[j][m]);
        sj[m] = fmax(kEpsS, __shfl(x, 63, 64));
        tot += sj[m];
      }
      if (!((opt >> j) & 1u)) continue;
#pragma unroll
      for (int m = 0; m < 5; ++m) sj[m] = sj[m] / tot * (double)K;
      BranchObj o;
      o.nu = nu;
      o.A = (nu + 1.0) * sj[0] + nu * sj[1] + nu * sj[3];
      o.S1 = sj[1];
      o.B = nu > kEpsNu ? (sj[2] + sj[4]) : 0.0;
      d[j] = solve_branch(o, F.dmin, F.dmax, d[j]);
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < 5; ++j) len_out[c * 5 + j] = d[j];
    nll_out[c] = -llh;
    iters_out[c] = iters;
  }
}

void run_quartet_em(const DevForest& F, int ncand, const int32_t* qnodes, const double* len_in,
                    const uint8_t* opt_mask, const double* llh_const, const double* nu, const double* phi,
                    int max_iter, double tol, double* len_out, double* nll_out, int32_t* iters_out, hipStream_t s) {
  if (ncand <= 0) return;
  hipLaunchKernelGGL(k_quartet_em, dim3((ncand + 3) / 4), dim3(256), 0, s, F, ncand, qnodes, len_in, opt_mask,
                     llh_const, nu, phi, max_iter, tol, len_out, nll_out, iters_out);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
int seg_width(int max_k) {
  if (max_k <= 8) return 8;
  if (max_k <= 16) return 16;
  if (max_k <= 32) return 32;
  return 64;
}

#define LAML_DISPATCH_W(W, BODY)                     \
  switch (W) {                                       \
    case 8: { constexpr int WW = 8; BODY; } break;   \
    case 16: { constexpr int WW = 16; BODY; } break; \
    case 32: { constexpr int WW = 32; BODY; } break; \
    default: { constexpr int WW = 64; BODY; } break; \
  }

static dim3 seg_grid(int cnt, int W, int chunks) {
  const int segs = 256 / W;
  return dim3((cnt + segs - 1) / segs, chunks);
}

void run_labels(const DevForest& F, const std::vector<int32_t>& lvl, int W, hipStream_t s) {
  for (size_t h = 1; h + 1 < lvl.size(); ++h) {
    const int beg = lvl[h], cnt = lvl[h + 1] - lvl[h];
    if (cnt <= 0) continue;
    LAML_DISPATCH_W(W, {
      hipLaunchKernelGGL(k_labels<WW>, seg_grid(cnt, WW, F.max_chunks), dim3(256), 0, s, F, beg, cnt);
    });
  }
}

void run_qrows(const DevForest& F, double* qleaf, double* qint, hipStream_t s) {
  const int grid = std::max(1, std::min((F.num_nodes + 3) / 4, 8192));
  hipLaunchKernelGGL(k_qrows, dim3(grid), dim3(256), 0, s, F, qleaf, qint);
}

void run_pad_codes(const DevForest& F, const int16_t* packed, const int64_t* packed_off, hipStream_t s) {
  const int grid = std::max(1, std::min((F.num_leaves + 3) / 4, 8192));
  hipLaunchKernelGGL(k_pad_codes, dim3(grid), dim3(256), 0, s, F, packed, packed_off);
}

void run_coef(const DevForest& F, const double* d, const double* nu, const int32_t* done, hipStream_t s) {
  const int grid = (F.num_nodes + 255) / 256;
  hipLaunchKernelGGL(k_coef, dim3(grid), dim3(256), 0, s, F, d, nu, done);
}

static void launch_inside_rows(dim3 grid, hipStream_t s, const DevForest& F, int beg, int cnt, int cpw,
                               const int32_t* done, const int32_t* nodes) {
  hipLaunchKernelGGL(k_inside_rows, grid, dim3(256), 0, s, F, beg, cnt, cpw, done, nodes);
}

static void launch_outside_rows(dim3 grid, hipStream_t s, const DevForest& F, int beg, int cnt, int cpw,
                                const int32_t* done, const int32_t* nodes) {
  hipLaunchKernelGGL(k_outside_rows, grid, dim3(256), 0, s, F, beg, cnt, cpw, done, nodes);
}

// one level of the inside pass: row kernels when a per-level shape is given
// (W == 64), segment kernels otherwise
static void inside_level(const DevForest& F, const std::vector<int32_t>& lvl, int W, const std::vector<int32_t>& cpw,
                         int h, const int32_t* done, hipStream_t s) {
  const int beg = lvl[h], cnt = lvl[h + 1] - lvl[h];
  if (cnt <= 0) return;
  if (W == 64 && !cpw.empty() && cpw[h] > 0) {
    const int c = cpw[h];
    const dim3 grid((cnt + 3) / 4, (F.max_chunks + c - 1) / c);
    if (h == 1)
      hipLaunchKernelGGL(k_inside_rows_leaves, grid, dim3(256), 0, s, F, beg, cnt, c, done);
    else
      launch_inside_rows(grid, s, F, beg, cnt, c, done, nullptr);
    return;
  }
  LAML_DISPATCH_W(W, {
    hipLaunchKernelGGL(k_inside<WW>, seg_grid(cnt, WW, F.max_chunks), dim3(256), 0, s, F, beg, cnt, done);
  });
}

static void outside_level(const DevForest& F, const std::vector<int32_t>& lvl, int W, const std::vector<int32_t>& cpw,
                          int h, const int32_t* done, hipStream_t s) {
  const int beg = lvl[h], cnt = lvl[h + 1] - lvl[h];
  if (cnt <= 0) return;
  if (W == 64 && !cpw.empty() && cpw[h] > 0) {
    const int c = cpw[h];
    launch_outside_rows(dim3((cnt + 3) / 4, (F.max_chunks + c - 1) / c), s, F, beg, cnt, c, done, nullptr);
    return;
  }
  LAML_DISPATCH_W(W, {
    hipLaunchKernelGGL(k_outside<WW>, seg_grid(cnt, WW, F.max_chunks), dim3(256), 0, s, F, beg, cnt, done);
  });
}

std::vector<int32_t> default_cpw(const std::vector<int32_t>& lvl, int W, int max_chunks) {
  std::vector<int32_t> c;
  if (W != 64) return c;
  const int H = (int)lvl.size() - 1;
  c.assign(H, 1);
  for (int h = 1; h < H; ++h) {
    const int64_t cnt = lvl[h + 1] - lvl[h];
    // enough wavefronts to keep every SIMD busy (kRowTargetWaves), then as
    // many chunks per wavefront as that allows
    const int64_t want = cnt * max_chunks / kRowTargetWaves;
    c[h] = (int32_t)std::max<int64_t>(1, std::min<int64_t>(max_chunks, want));
  }
  return c;
}

// One level of a pass with fused cherries (row kernels): the level's nodes
// without a cherry child, then those with one, each from its node list.
static void split_level(bool outside, const DevForest& F, const CherrySplit& cs, const std::vector<int32_t>& cpw,
                        int h, const int32_t* done, hipStream_t s) {
  const int c = std::max(1, cpw[h]);
  const int pb = cs.off[2 * h], pe = cs.off[2 * h + 1], ce = cs.off[2 * h + 2];
  const unsigned gy = (unsigned)((F.max_chunks + c - 1) / c);
  if (pe > pb) {
    const int cnt = pe - pb;
    if (outside)
      launch_outside_rows(dim3((cnt + 3) / 4, gy), s, F, pb, cnt, c, done, cs.nodes);
    else
      launch_inside_rows(dim3((cnt + 3) / 4, gy), s, F, pb, cnt, c, done, cs.nodes);
  }
  if (ce > pe) {
    const int cnt = ce - pe;
    if (outside)
      hipLaunchKernelGGL(k_outside_rows_ch, dim3((cnt + 3) / 4, gy), dim3(256), 0, s, F, pe, cnt, c, done,
                         cs.nodes);
    else
      hipLaunchKernelGGL(k_inside_rows_ch, dim3((cnt + 3) / 4, gy), dim3(256), 0, s, F, pe, cnt, c, done, cs.nodes);
  }
}

void run_inside(const DevForest& F, const std::vector<int32_t>& lvl, int W, int h_top, const std::vector<int32_t>& cpw,
                const int32_t* done, hipStream_t s, const CherrySplit* cs, const ChainPlan* cp) {
  const int H = (int)lvl.size() - 1;
  if (cs) {
    // fused cherries: row kernels on every level from 2 up, no top kernel
    for (int h = 2; h < H; ++h) split_level(false, F, *cs, cpw, h, done, s);
    return;
  }
  const int ht = std::max(1, std::min(h_top, H));
  // chains below the chain plan's top level, level launches from there
  const int h0 = cp ? std::min(cp->h_split, ht) : 1;
  if (cp) run_inside_chains(F, *cp, done, s);
  for (int h = h0; h < ht; ++h) inside_level(F, lvl, W, cpw, h, done, s);
  if (ht < H) {
    LAML_DISPATCH_W(W, {
      hipLaunchKernelGGL(k_inside_top<WW>, dim3(F.max_chunks), dim3(1024), 0, s, F, ht, H, done);
    });
  }
}

void run_root(const DevForest& F, int W, double* site_llh_out, const int32_t* done, hipStream_t s) {
  LAML_DISPATCH_W(W, {
    hipLaunchKernelGGL(k_root<WW>, seg_grid(F.num_trees, WW, F.max_chunks), dim3(256), 0, s, F, site_llh_out,
                       done);
  });
}

void run_outside_ckpt(const DevForest& F, const std::vector<int32_t>& lvl, int W, int L,
                      const std::vector<int32_t>& in_cpw, const std::vector<int32_t>& out_cpw, const int32_t* done,
                      hipStream_t s) {
  const int H = (int)lvl.size() - 1;
  const int nseg = (H - 1 + L - 1) / L;  // internal levels 1..H-1
  for (int sg = nseg - 1; sg >= 0; --sg) {
    const int a = 1 + sg * L, b = std::min(H, 1 + (sg + 1) * L);
    // recompute the segment's inside vectors (working area), then its
    // outside levels from the top of the segment down
    if (sg != nseg - 1)
      for (int h = a; h < b; ++h) inside_level(F, lvl, W, in_cpw, h, done, s);
    for (int h = b - 1; h >= a; --h) outside_level(F, lvl, W, out_cpw, h, done, s);
  }
}

void run_outside(const DevForest& F, const std::vector<int32_t>& lvl, int W, int h_top,
                 const std::vector<int32_t>& cpw, const int32_t* done, hipStream_t s, const CherrySplit* cs,
                 const ChainPlan* cp) {
  const int H = (int)lvl.size() - 1;
  if (cs) {
    for (int h = H - 1; h >= 2; --h) split_level(true, F, *cs, cpw, h, done, s);
    return;
  }
  const int ht = std::max(1, std::min(h_top, H));
  if (ht < H) {
    LAML_DISPATCH_W(W, {
      hipLaunchKernelGGL(k_outside_top<WW>, dim3(F.max_chunks), dim3(1024), 0, s, F, ht, H, done);
    });
  }
  const int h0 = cp ? std::min(cp->h_split, ht) : 1;
  for (int h = ht - 1; h >= h0; --h) outside_level(F, lvl, W, cpw, h, done, s);
  if (cp) run_outside_chains(F, *cp, done, s);
}

void run_inside_chains(const DevForest& F, const ChainPlan& cp, const int32_t* done, hipStream_t s) {
  for (size_t r = 0; r + 1 < cp.in_round_off.size(); ++r) {
    const int beg = cp.in_round_off[r], cnt = cp.in_round_off[r + 1] - beg;
    if (cnt <= 0) continue;
    const int64_t waves = (int64_t)cnt * F.max_chunks;
    hipLaunchKernelGGL(k_inside_chain, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, F, cp.in_chains, beg, cnt,
                       cp.nodes, done);
  }
}

void run_outside_chains(const DevForest& F, const ChainPlan& cp, const int32_t* done, hipStream_t s) {
  for (size_t r = 0; r + 1 < cp.pack_round_off.size(); ++r) {
    const int beg = cp.pack_round_off[r], cnt = cp.pack_round_off[r + 1] - beg;
    if (cnt <= 0) continue;
    const int64_t waves = (int64_t)cnt * F.max_chunks;
    if (cp.round_packed[r])
      hipLaunchKernelGGL(k_outside_chain<true>, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, F, cp.packs, beg,
                         cnt, cp.pack_steps, cp.pack_head_ro, done);
    else
      hipLaunchKernelGGL(k_outside_chain<false>, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, F, cp.packs, beg,
                         cnt, cp.pack_steps, cp.pack_head_ro, done);
  }
}

void run_tree_walk(const DevForest& F, const TreeWalk& TW, int W, int phases, double* site_llh_out,
                   const int32_t* done, hipStream_t s, int h_lo) {
  LAML_DISPATCH_W(W, {
    hipLaunchKernelGGL(k_tree_walk<WW>, dim3(TW.num_blocks), dim3(kWalkThreads), 0, s, F, TW, phases, h_lo,
                       site_llh_out, done);
  });
}

void run_inside_levels(const DevForest& F, const std::vector<int32_t>& lvl, int W, int h0, int h1,
                       const std::vector<int32_t>& cpw, const int32_t* done, hipStream_t s) {
  for (int h = h0; h < h1; ++h) inside_level(F, lvl, W, cpw, h, done, s);
}

void run_outside_levels(const DevForest& F, const std::vector<int32_t>& lvl, int W, int h0, int h1,
                        const std::vector<int32_t>& cpw, const int32_t* done, hipStream_t s) {
  for (int h = h1 - 1; h >= h0; --h) outside_level(F, lvl, W, cpw, h, done, s);
}

// Posterior state probabilities of every (node, site): P(v = 0 | D) and
// P(v = -1 | D), the post0 / post1 of EM_solver.Estep_posterior
// (EM_solver.py:247-300, in linear space), from the resident inside and
// outside state.  Segments index internal nodes; a leaf child's outside
// state is formed from its parent's as in the outside pass.  Outputs use the
// device row layout: internal node v at leaf_rows + row_off[v], leaf v at
// row_off[v].
template <int W>
__global__ void __launch_bounds__(256) k_node_post(DevForest F, int64_t leaf_rows, double* __restrict__ p0,
                                                   double* __restrict__ pS) {
  Seg<W> g;
  if (!g.init(F.num_leaves, F.num_nodes - F.num_leaves)) return;
  const int u = g.v, k = g.k;
  const int t = F.tree_of[u];
  const int prob = F.tree_prob[t];
  const int K = F.tree_nsites[t];
  if (k >= K) return;
  const int sb = F.tree_site_base[t];
  const double phi = F.phi[prob];
  const double iL = F.iLm[sb + k];
  const int LE = F.LE[sb + k];
  const int64_t ru = F.row_off[u];
  const double4 a4 = F.alpha[ru + k];
  QAlpha au;
  au.x = a4.x;
  au.y = a4.y;
  au.z = a4.z;
  au.w = a4.w;
  au.E = F.aE[ru + k];
  au.g = F.aG[ru + k];
  const int labu = F.labels[ru + k];
  const int64_t bu = F.brow[u];
  const double2 bb = F.beta[bu + k];
  const int Eu = F.bE[bu + k];
  p0[leaf_rows + ru + k] = au.x * bb.x * ldexp(iL, min(au.E + Eu - LE, 1000));
  pS[leaf_rows + ru + k] = labu < 0 ? au.y * ldexp(iL, min(au.E - LE, 1000)) : 0.0;
  const int ch[2] = {F.child0[u], F.child1[u]};
  const int qo = F.site_qoff[sb + k] - 1;
  const double qs = F.site_qsum[sb + k];
  int lab[2], E[2], code[2];
  double b0[2], bA[2], q[2];
  Eps ep[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    load_child(F, ch[i], ch[i] < F.num_leaves, k, phi, lab[i], b0[i], bA[i], E[i], code[i]);
    q[i] = lab[i] > 0 ? F.prior[qo + lab[i]] : qs;
    ep[i] = edge(lab[i], b0[i], bA[i], E[i], q[i], F.coef[ch[i]]);
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    if (ch[i] >= F.num_leaves) continue;
    const QTop tp = q_top(au, ep[1 - i], lab[1 - i], q[1 - i], qs);
    const QAlpha av = q_child_alpha(tp, F.coef[ch[i]]);
    const int64_t rv = F.row_off[ch[i]];
    p0[rv + k] = av.x * b0[i] * ldexp(iL, min(av.E - LE, 1000));
    pS[rv + k] = lab[i] < 0 ? av.y * ldexp(iL, min(av.E - LE, 1000)) : 0.0;
  }
}

void run_node_post(const DevForest& F, int W, int64_t leaf_rows, double* p0, double* pS, hipStream_t s) {
  const int cnt = F.num_nodes - F.num_leaves;
  if (cnt <= 0) return;
  LAML_DISPATCH_W(W, {
    hipLaunchKernelGGL(k_node_post<WW>, seg_grid(cnt, WW, F.max_chunks), dim3(256), 0, s, F, leaf_rows, p0, pS);
  });
}

// ---------------------------------------------------------------------------
// Staging area of laml_posteriors_noalloc: a file-scope device array, so that
// the entry point needs no allocation.  In host code the name `scratch` is a
// host-side stand-in, never a device address: the address every kernel
// argument and copy uses is the one hipGetSymbolAddress returns for the
// device that is current, looked up once per device by post_scratch().
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[kPostScratchDoubles];

int post_scratch(int device, double** out) {
  static std::mutex mu;
  static std::map<int, double*> by_device;
  std::lock_guard<std::mutex> lock(mu);
  auto it = by_device.find(device);
  if (it == by_device.end()) {
    // the caller has made `device` current: each device holds its own copy
    // of the array at its own address
    double* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    it = by_device.emplace(device, ptr).first;
  }
  *out = it->second;
  return 0;
}

// Copies each node's K posterior values from the padded row layout of
// run_node_post (leaf rows, then internal rows at leaf_rows) to the packed
// layout of laml_forest_labels (leaf entries, then internal entries at
// leaf_sites; desc_off[v] = the node's first entry in its part).
__global__ void __launch_bounds__(64) k_post_pack(DevForest F, const int64_t* __restrict__ desc_off,
                                                  int64_t leaf_rows, int64_t leaf_sites,
                                                  const double* __restrict__ p0, const double* __restrict__ pS,
                                                  double* __restrict__ out0, double* __restrict__ outS) {
  const int v = blockIdx.x;
  if (v >= F.num_nodes) return;
  const int K = F.tree_nsites[F.tree_of[v]];
  const bool leaf = v < F.num_leaves;
  const int64_t src = (leaf ? 0 : leaf_rows) + F.row_off[v];
  const int64_t dst = (leaf ? 0 : leaf_sites) + desc_off[v];
  for (int k = threadIdx.x; k < K; k += 64) {
    out0[dst + k] = p0[src + k];
    outS[dst + k] = pS[src + k];
  }
}

void run_post_pack(const DevForest& F, const int64_t* desc_off, int64_t leaf_rows, int64_t leaf_sites,
                   const double* p0, const double* pS, double* out0, double* outS, hipStream_t s) {
  if (F.num_nodes <= 0) return;
  hipLaunchKernelGGL(k_post_pack, dim3(F.num_nodes), dim3(64), 0, s, F, desc_off, leaf_rows, leaf_sites, p0, pS,
                     out0, outS);
}

void run_counts(const DevForest& F, double* counts, hipStream_t s) {
  hipLaunchKernelGGL(k_counts, dim3((unsigned)(((int64_t)F.num_nodes * 5 + 255) / 256)), dim3(256), 0, s, F,
                     counts);
}

// clears the EM bookkeeping of every problem (a kernel rather than a memset,
// so that the call stays valid inside a caller's stream capture)
__global__ void k_em_reset(EmState em, int P) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  em.done[p] = 0;
  em.iters[p] = 0;
  em.step[p] = 0;
}

void run_em_reset(const EmState& em, int P, hipStream_t s) {
  hipLaunchKernelGGL(k_em_reset, dim3((P + 255) / 256), dim3(256), 0, s, em, P);
}

void run_llh(const DevForest& F, double* nll, EmState em, int max_iter, double tol, hipStream_t s) {
  hipLaunchKernelGGL(k_llh, dim3(F.num_problems), dim3(256), 0, s, F, nll, em, max_iter, tol);
}

void run_grad(const DevForest& F, const double* d, const double* nu, const double* phi, double* grad_d,
              double* grad_nu, double* grad_phi, hipStream_t s) {
  hipLaunchKernelGGL(k_grad_branch, dim3((F.num_nodes + 255) / 256), dim3(256), 0, s, F, d, nu, grad_d);
  hipLaunchKernelGGL(k_grad_problem, dim3(F.num_problems), dim3(kRedThreads), 0, s, F, nu, phi, grad_nu, grad_phi);
}

void run_rtilde(const DevForest& F, double* rtilde, const int32_t* done, hipStream_t s) {
  hipLaunchKernelGGL(k_rtilde, dim3(F.num_problems), dim3(kRedThreads), 0, s, F, rtilde, done);
}

void run_mstep(const DevForest& F, const double* counts, const double* rtilde, double* d, double* nu, double* phi,
               const int32_t* done, hipStream_t s, const UltraDev* U) {
  UltraDev none{};
  hipLaunchKernelGGL(k_branch, dim3((F.num_nodes + 63) / 64), dim3(64), 0, s, F, counts, d, nu, done,
                     U ? *U : none, U ? 1 : 0);
  if (U)
    hipLaunchKernelGGL(k_ultra, dim3(F.num_problems), dim3(kUltraThreads), 0, s, F, *U, d, nu, done);
  // one launch when the problems keep enough blocks busy; otherwise split
  // each problem's branches over several blocks per evaluation
  if (F.max_prob_nodes <= kNuSolveMaxNodes && F.num_problems >= 8) {
    hipLaunchKernelGGL(k_nu_solve, dim3(F.num_problems), dim3(kNuSolveThreads), 0, s, F, rtilde, nu, phi, done, 0);
    return;
  }
  for (int it = 0; it <= kNuIters; ++it)
    hipLaunchKernelGGL(k_nu_eval, dim3(F.num_nu_blocks), dim3(kNuThreads), 0, s, F, rtilde, nu, phi, done, it);
  // solves the launches above left unconverged (rare): one block each
  hipLaunchKernelGGL(k_nu_solve, dim3(F.num_problems), dim3(kNuSolveThreads), 0, s, F, rtilde, nu, phi, done, 1);
}

}  // namespace laml
