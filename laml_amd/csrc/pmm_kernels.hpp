// Internal interface between the kernels and the C ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "pmm_device.hpp"

namespace laml {

// Device view of a forest; passed by value to every kernel.
struct DevForest {
  int32_t num_nodes, num_leaves, num_trees, num_problems;
  double dmin, dmax;
  // topology (level order)
  const int32_t* child0;
  const int32_t* child1;
  const int32_t* tree_of;
  const int64_t* row_off;
  const uint8_t* free_branch;
  // trees / problems
  const int32_t* tree_root;
  const int32_t* tree_prob;
  const int32_t* tree_nsites;
  const int32_t* tree_site_base;
  const int32_t* prob_tree_off;  // [P+1]
  const int32_t* prob_node_off;  // [P+1]
  const int32_t* prob_nodes;     // nodes grouped by problem
  const uint32_t* prob_flags;
  const double* prob_nobs;
  const double* prob_phimin;
  // sites
  const int32_t* site_qoff;  // 1-based local state j -> prior[site_qoff[s] + j - 1]
  const double* prior;
  const int16_t* codes;  // leaf rows
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
  double* Lm;  // [sites]
  int32_t* LE;
  double* site_llh;
  double* rt_part;  // [N]
  double* snorm;    // [N*5]
};

struct EmState {
  int32_t* done;   // [P] or nullptr outside the EM loop
  int32_t* iters;  // [P]
  double* pre;     // [P]
  double* trace;   // [(max_iter+1)*P] or nullptr
};

int seg_width(int max_k);
void run_labels(const DevForest& F, const std::vector<int32_t>& lvl, int W, hipStream_t s);
void run_coef(const DevForest& F, const double* d, const double* nu, const int32_t* done, hipStream_t s);
void run_inside(const DevForest& F, const std::vector<int32_t>& lvl, int W, const int32_t* done, hipStream_t s);
void run_root(const DevForest& F, int W, double* site_llh_out, double* counts, const int32_t* done, hipStream_t s);
void run_outside(const DevForest& F, const std::vector<int32_t>& lvl, int W, double* counts, const int32_t* done,
                 hipStream_t s);
void run_llh(const DevForest& F, double* nll, EmState em, int step, int max_iter, double tol, hipStream_t s);
void run_rtilde(const DevForest& F, double* rtilde, const int32_t* done, hipStream_t s);
void run_mstep(const DevForest& F, const double* counts, const double* rtilde, double* d, double* nu, double* phi,
               const int32_t* done, hipStream_t s);

}  // namespace laml
