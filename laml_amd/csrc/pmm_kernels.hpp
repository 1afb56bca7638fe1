// Internal interface between the kernels and the C ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "pmm_device.hpp"

namespace laml {

// Newton state of one problem's nu solve, and the status of its last M-step
struct NuState {
  double x, lo, hi;
  int32_t fin;
  int32_t count;   // blocks arrived in the current evaluation
  int32_t status;  // last M-step: bit 0 a branch length came out non-finite,
                   // bit 1 the nu solve did not converge (kept at its last
                   // iterate)
  int32_t d_bad;   // set by k_branch, folded into status by the nu solve
};

// Per internal node, what the inside row kernel reads before its row loads
// (one scalar load instead of chains of dependent ones): children, problem,
// sites and first site of the tree, and the row offsets of the node's labels
// and inside row and of its children's.
struct NodeRec {
  int32_t ch0, ch1, p, K;
  int32_t sb, pad;
  int64_t ro, bro;
  int64_t ro0, bro0, ro1, bro1;
};

// Device view of a forest; passed by value to every kernel.
struct DevForest {
  int32_t num_nodes, num_leaves, num_trees, num_problems;
  int32_t chunk_w;     // sites per segment (W)
  int32_t max_chunks;  // ceil(max K / W)
  // nodes [num_leaves, cherry_end) are fused cherries in the current pass
  // (their parents form and consume their inside and outside states; no
  // cherry row is read or written); cherry_end == num_leaves: no fusion
  int32_t cherry_end;
  // 1: the inside/outside rows of a pass exceed the last-level cache, and
  // the row kernels read and write them with the non-temporal hint
  int32_t stream_rows;
  double dmin, dmax;
  // topology (level order)
  const int32_t* level_off;  // [levels+1]
  const int32_t* child0;
  const int32_t* child1;
  const int32_t* tree_of;
  const int64_t* row_off;
  const int64_t* brow;  // [N] row of the node's inside vector (beta, bE)
  const NodeRec* nrec;  // [N - num_leaves] by internal node
  const uint8_t* free_branch;
  // trees / problems
  const int32_t* tree_root;
  const int32_t* tree_prob;
  const int32_t* tree_nsites;
  const int32_t* tree_site_base;
  const int32_t* prob_tree_off;  // [P+1]
  const int32_t* prob_node_off;  // [P+1]
  const int32_t* prob_nodes;     // nodes grouped by problem (leaves first)
  const int32_t* prob_int_off;   // [P] first internal node of each problem in prob_nodes
  const int32_t* prob_nfree;     // [P] number of optimised branches
  const int32_t* node_pos;       // [N] position of each node in prob_nodes
  const int32_t* prob_blk_off;   // [P+1] blocks of the nu solve per problem
  int32_t num_nu_blocks;
  int32_t max_prob_nodes;        // branches of the largest problem
  const uint32_t* prob_flags;
  const double* prob_nobs;
  const double* prob_phimin;
  const double* prob_nobs_phi;  // [P] leaf entries observed as 0 or an edit
  // sites
  const int32_t* site_qoff;  // 1-based local state j -> prior[site_qoff[s] + j - 1]
  const double* site_qsum;   // prior mass of each site's alphabet (1 for normalised priors)
  const double* prior;
  const int16_t* codes;  // leaf rows
  // prior weight of each node's label at each site (leaf rows by the leaf
  // row offsets, internal rows by the internal ones): prior[site_qoff[s] - 1 +
  // label] for an edit label, the site's prior mass site_qsum[s] otherwise;
  // fixed for the forest's life (labels and priors are)
  const double* qleaf;
  const double* qint;
  // per-call parameters
  const double* phi;  // [P]
  // state
  int16_t* labels;  // internal rows
  Coef* coef;       // [N]
  double2* beta;    // internal rows
  int32_t* bE;
  double4* alpha;  // internal rows
  int32_t* aE;
  int16_t* aG;
  double* iLm;  // [sites] 1 / mantissa of the site likelihood (L = 2^LE / iLm)
  int32_t* LE;
  double* site_llh;
  double* rt_part;  // [N * max_chunks] dropout sums of leaf children per chunk
  double* cpart;    // [N * max_chunks * 5] per-chunk branch count partials
  double* nu_a;     // [N] by position in prob_nodes: S0+S1+S3 of the branch
  double* nu_b;     // [N] S2+S4 (0 for fixed branches)
  double* nu_d;     // [N] new branch length
  double* nu_part;  // [num_nu_blocks * 3] partial sums of the nu solve
  NuState* nu_state;  // [P]
};

// Height-class model of the ultrametric constraint (one per forest)
struct UltraDev {
  int32_t num_classes, num_levels;
  const int32_t* plo;        // [num_levels * P + 1] class ranges by (level, problem)
  const int32_t* up_cls;     // [N] parent-side class of each free branch (-1 pinned)
  const int32_t* dn_cls;     // [N] child-side class (-1 pinned)
  const double* up_off;      // [N] offset (absolute height when pinned)
  const double* dn_off;      // [N]
  const int32_t* cls_parent; // [V]
  const int32_t* cls_pedge;  // [V] branch joining the class to its parent class
  const int32_t* inc_off;    // [V+1]
  const int32_t* inc_branch; // incident free branches
  const int8_t* inc_sign;    // +1: class is the parent side
  const uint8_t* has_free;   // [P]
  double* t;                 // [V] heights (warm start / solution)
  double* dt;                // [V] Newton direction
  double* Dd;                // [V] eliminated diagonal (= Gg + weight of the edge to the parent class)
  double* Gg;                // [V] conductance of the class subtree to the pinned heights
  double* r;                 // [V] eliminated right-hand side
  double* g;                 // [V] gradient
  double* gp;                // [N] branch first derivative
  double* w;                 // [N] branch curvature (positive)
  double* bA;                // [N] branch objective: -A d + S1 log(1-e^-d) + B log(1-e^-nu d)
  double* bS1;
  double* bB;
  int32_t* status;           // [P] 0 ok, 1 line search stalled, 2 singular Newton system
};

// Tree-walk launch plan: block b runs tree blk_tree[b] at site chunk
// blk_chunk[b]; tree t's internal levels are lvl_base[t] + 1 ..
// lvl_base[t+1] - 1, level h holding the contiguous node range
// [lvl_beg[lvl_base[t] + h], + lvl_cnt[...]).
struct TreeWalk {
  int32_t num_blocks;
  const int32_t* blk_tree;
  const int32_t* blk_chunk;
  const int32_t* lvl_base;  // [T+1]
  const int32_t* lvl_beg;
  const int32_t* lvl_cnt;
};

struct EmState {
  int32_t* done;   // [P] or nullptr outside the EM loop
  int32_t* iters;  // [P]
  int32_t* step;   // [P] EM steps taken (read and advanced on the device)
  double* pre;     // [P]
  double* trace;   // [(max_iter+1)*P] or nullptr
};

int seg_width(int max_k);
void run_labels(const DevForest& F, const std::vector<int32_t>& lvl, int W, hipStream_t s);
// the prior-weight rows qleaf / qint from the codes and labels (after run_labels)
void run_qrows(const DevForest& F, double* qleaf, double* qint, hipStream_t s);
// leaf code rows from the packed host layout (uploaded as it is) into the
// padded device rows of F.codes
void run_pad_codes(const DevForest& F, const int16_t* packed, const int64_t* packed_off, hipStream_t s);
void run_coef(const DevForest& F, const double* d, const double* nu, const int32_t* done, hipStream_t s);
// wavefronts a row-kernel launch aims for before it gives each wavefront
// more than one chunk of 64 sites (8 per SIMD of the 256 CUs)
constexpr int64_t kRowTargetWaves = 8192;
// default chunks per wavefront of the row kernels for each level (empty
// unless W == 64)
std::vector<int32_t> default_cpw(const std::vector<int32_t>& lvl, int W, int max_chunks);
// Per-level node lists of a pass with fused cherries: for level h >= 2,
// nodes[off[2h] .. off[2h+1]) have no cherry child, nodes[off[2h+1] ..
// off[2h+2]) have at least one (device list, host offsets).
struct CherrySplit {
  const int32_t* nodes = nullptr;
  std::vector<int32_t> off;
};

// Heavy-path chains over the levels below h_split (k_outside_chain,
// k_inside_chain): a chain is a node and its successive heavy children,
// head first.  Outside rounds: the heads of round 0 get their outside state
// from the levels above h_split or the root step, those of round r > 0 from
// round r - 1; round r's packs are [pack_round_off[r], pack_round_off[r +
// 1]), longest first, each the steps of one or more of the round's chains of
// one tree.  in_chains lists the chains (offset into nodes, length) in
// inside rounds: a chain runs one round after the last chain hanging off it
// (its light internal children's), round 0 holding chains with none (device
// arrays, host offsets).
// One node of a chain, as the outside chain kernel reads it: the node, its
// children, which child continues the chain (-1: none), and the children's
// row offsets.
struct ChainStep {
  int32_t u, ch0, ch1, hv;
  int64_t ro0, ro1;
};

struct ChainPlan {
  int32_t h_split = -1;
  int32_t pack_cap = 0;              // most steps per outside pack the plan was built for
  const int2* in_chains = nullptr;
  const int32_t* nodes = nullptr;
  // outside packs: (offset into pack_steps, steps), each one or more chains
  // of one tree back to back; a chain's first record holds ~u and
  // pack_head_ro its row offset
  const int2* packs = nullptr;
  const ChainStep* pack_steps = nullptr;
  const int64_t* pack_head_ro = nullptr;
  std::vector<int32_t> pack_round_off, in_round_off;
  std::vector<uint8_t> round_packed;  // per round: some pack holds more than one chain
};
// wavefronts an outside chain round keeps: its packs hold at most
// max(1, min(cap, (steps x chunks) / kPackWaves)) steps, so that packing
// only thins rounds with wavefronts to spare (16 per SIMD)
constexpr int64_t kPackWaves = 16384;

// cp: the levels below the top kernel's run as heavy-path chains
void run_inside(const DevForest& F, const std::vector<int32_t>& lvl, int W, int h_top, const std::vector<int32_t>& cpw,
                const int32_t* done, hipStream_t s, const CherrySplit* cs = nullptr, const ChainPlan* cp = nullptr);
void run_inside_chains(const DevForest& F, const ChainPlan& cp, const int32_t* done, hipStream_t s);
// cp: the levels below the top kernel's run as heavy-path chains
void run_outside(const DevForest& F, const std::vector<int32_t>& lvl, int W, int h_top,
                 const std::vector<int32_t>& cpw, const int32_t* done, hipStream_t s, const CherrySplit* cs = nullptr,
                 const ChainPlan* cp = nullptr);
void run_outside_chains(const DevForest& F, const ChainPlan& cp, const int32_t* done, hipStream_t s);
void run_outside_ckpt(const DevForest& F, const std::vector<int32_t>& lvl, int W, int L,
                      const std::vector<int32_t>& in_cpw, const std::vector<int32_t>& out_cpw, const int32_t* done,
                      hipStream_t s);
void run_root(const DevForest& F, int W, double* site_llh_out, const int32_t* done, hipStream_t s);
void run_counts(const DevForest& F, double* counts, hipStream_t s);
void run_tree_walk(const DevForest& F, const TreeWalk& TW, int W, int phases, double* site_llh_out,
                   const int32_t* done, hipStream_t s, int h_lo = 1);
// levels [h0, h1) of the inside pass (bottom up) / outside pass (top down)
void run_inside_levels(const DevForest& F, const std::vector<int32_t>& lvl, int W, int h0, int h1,
                       const std::vector<int32_t>& cpw, const int32_t* done, hipStream_t s);
void run_outside_levels(const DevForest& F, const std::vector<int32_t>& lvl, int W, int h0, int h1,
                        const std::vector<int32_t>& cpw, const int32_t* done, hipStream_t s);
void run_llh(const DevForest& F, double* nll, EmState em, int max_iter, double tol, hipStream_t s);
void run_em_reset(const EmState& em, int P, hipStream_t s);
void run_node_post(const DevForest& F, int W, int64_t leaf_rows, double* p0, double* pS, hipStream_t s);
// doubles in the file-scope staging array of laml_posteriors_noalloc (64 MiB
// per device, held until the process ends)
constexpr int64_t kPostScratchDoubles = (int64_t)8 << 20;
// device address of that array on `device` (which must be current), looked
// up once per device; 0 or a hipError_t
int post_scratch(int device, double** out);
void run_post_pack(const DevForest& F, const int64_t* desc_off, int64_t leaf_rows, int64_t leaf_sites,
                   const double* p0, const double* pS, double* out0, double* outS, hipStream_t s);
void run_quartet_em(const DevForest& F, int ncand, const int32_t* qnodes, const double* len_in,
                    const uint8_t* opt_mask, const double* llh_const, const double* nu, const double* phi,
                    int max_iter, double tol, double* len_out, double* nll_out, int32_t* iters_out, hipStream_t s);
void run_rtilde(const DevForest& F, double* rtilde, const int32_t* done, hipStream_t s);
void run_grad(const DevForest& F, const double* d, const double* nu, const double* phi, double* grad_d,
              double* grad_nu, double* grad_phi, hipStream_t s);
void run_mstep(const DevForest& F, const double* counts, const double* rtilde, double* d, double* nu, double* phi,
               const int32_t* done, hipStream_t s, const UltraDev* U = nullptr);

}  // namespace laml
