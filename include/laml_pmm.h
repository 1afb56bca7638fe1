/*
 * laml_pmm.h -- C ABI of the MI355X (gfx950) PMM likelihood / EM engine.
 *
 * The reference (raphael-group/LAML 1.0.5) is pure Python and has no FFI; the
 * entry points below are what its solver classes would bind for the hot path.
 * Each declaration names the reference function it replaces.
 *
 * Conventions
 *  - Every function returns LAML_OK (0) or a positive error code; the message
 *    of the last error on the calling thread is laml_last_error().
 *  - No C++ exception crosses this boundary.
 *  - "device" pointers are HIP device memory of the forest's device (e.g. the
 *    data_ptr() of a torch tensor); "host" pointers are ordinary memory.
 *  - Work is enqueued on the given hipStream_t (passed as void*; NULL = the
 *    null stream) and is stream-ordered.  A forest is not thread-safe; distinct
 *    forests are independent.
 *
 * Model objects
 *  - A "forest" holds one or more trees.  Trees are grouped into "problems":
 *    all trees of a problem share nu, phi and one character matrix (the
 *    reference's multi-tree solver, ML_solver.py:19-32); distinct problems are
 *    independent (batched NNI candidates, serving requests).
 *  - Nodes are numbered in height-level order over the whole forest: level h
 *    holds the nodes whose longest path down to a leaf has h edges, so leaves
 *    are nodes [0, num_leaves).  Every node owns the branch above it; a root
 *    owns the stem edge from the unedited ancestor (EM_solver.py:154-158).
 *  - Trees must be binary (polytomies resolved on the host, EM_solver.py:26-43).
 */
#ifndef LAML_PMM_H
#define LAML_PMM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LAML_OK 0
#define LAML_EINVAL 1    /* malformed description or argument */
#define LAML_EDEVICE 2   /* HIP runtime error */
#define LAML_ENOMEM 3    /* device allocation failed */
#define LAML_ESTATE 4    /* call out of order */

/* leaf codes (int16) */
#define LAML_CODE_UNEDITED 0  /* state 0                                    */
                              /* 1..m: local edit index of the site         */
#define LAML_CODE_MISSING (-1) /* '?': dropout or silencing                 */
#define LAML_CODE_SILENCED (-2) /* explicit -1: silencing only             */

/* per-problem optimisation flags */
#define LAML_OPT_PHI 1u
#define LAML_OPT_NU 2u

typedef struct laml_forest laml_forest;

typedef struct {
  int32_t num_problems;
  int32_t num_trees;
  int32_t num_nodes;
  int32_t num_leaves;
  int32_t num_levels;          /* levels 0..num_levels-1                     */
  const int32_t* level_off;    /* [num_levels+1]                            */
  const int32_t* child0;       /* [num_nodes], -1 for leaves                */
  const int32_t* child1;       /* [num_nodes], -1 for leaves                */
  const int32_t* parent;       /* [num_nodes], -1 for roots                 */
  const int32_t* tree_of;      /* [num_nodes]                               */
  const uint8_t* free_branch;  /* [num_nodes] 1 = optimised by the M-step   */
  const int32_t* tree_root;    /* [num_trees]                               */
  const int32_t* tree_problem; /* [num_trees], non-decreasing               */
  const int32_t* tree_nsites;  /* [num_trees] K of the tree's problem       */
  const int32_t* site_qoff;    /* [sum_t K_t + 1] CSR offsets into prior    */
  const double* prior;         /* [site_qoff[last]] q of local states 1..m  */
  const int16_t* leaf_codes;   /* [sum over leaves of K] leaf rows in node order */
  const double* problem_nobs;  /* [num_problems] # non-'?' leaf entries (R) */
  const double* problem_phimin;/* [num_problems] phi cut-off 1/(K*ncells)   */
  const uint32_t* problem_flags; /* [num_problems] LAML_OPT_* bits          */
  double dmin, dmax;           /* branch-length box of the M-step           */
  const double* site_qsum;     /* [sum_t K_t] total prior mass of each site's
                                * packed alphabet, used where a missing ('?')
                                * subtree sums over edit states; NULL = 1.0
                                * (normalised priors, the reference EM model) */
} laml_forest_desc;

/* Host-only check of a description (shapes, topology, codes, offsets); no
 * device is touched.  laml_forest_create runs the same checks first. */
int laml_forest_validate(const laml_forest_desc* desc);

/* Build the device-resident forest: uploads the description, computes the
 * az-partition labels on the device.
 * Replaces: ML_solver.__init__ + EM_solver.__init__ data setup
 * (ML_solver.py:19-86, EM_solver.py:19-24) and az_partition
 * (ML_solver.py:125-154). */
int laml_forest_create(const laml_forest_desc* desc, int device, laml_forest** out);

/* Release every device buffer of the forest. */
void laml_forest_destroy(laml_forest* f);

/* Change the LAML_OPT_* flags of every problem (host array [num_problems]).
 * Replaces the fixed_phi / fixed_nu arguments of EM_solver.optimize_one
 * (EM_solver.py:514-523). */
int laml_forest_set_flags(laml_forest* f, const uint32_t* flags_host);

/* Level checkpointing of the inside vectors: with levels = L > 0 only the
 * inside vectors of nodes whose parent lies in a higher block of L levels
 * (and the roots) are kept; the outside pass walks the blocks top-down and
 * recomputes each block's inside vectors in one reusable working area before
 * using them.  levels = 0 restores full storage.  Inside state must be
 * recomputed (laml_nll) after a change.  Needs the level pass mode (a forest
 * in the hybrid mode is switched to it; the tree-walk mode is refused).
 * No counterpart in the reference (its inside values live in Python
 * objects); it bounds the memory of the adjoint for very large trees. */
int laml_forest_set_checkpoint(laml_forest* f, int32_t levels);

/* Rows (node x site) of inside-vector storage currently allocated. */
int64_t laml_forest_inside_rows(laml_forest* f);

/* Launch shapes of the level-synchronous passes, host arrays of num_levels
 * entries indexed by level (entry 0 unused).  When the widest site row holds
 * more than 32 sites, each level runs one wavefront per node over
 * inside_cpw[h] / outside_cpw[h] consecutive chunks of 64 sites; otherwise
 * the entries read 0 and are ignored.  top_level is the first level that the
 * fused top-level kernel runs (num_levels: none); level checkpointing
 * disables the fused kernel while it is on.  Shapes change the order of
 * floating-point sums over sites only.
 * No counterpart in the reference (its passes are Python loops); the values
 * come from the default heuristic or from laml_amd.autotune. */
/* Pass mode: LAML_PASS_LEVELS runs each pass as level-synchronous launches
 * over the whole forest; LAML_PASS_TREES runs the inside pass, the root step
 * and the outside pass of every (tree, chunk of sites) pair in one block of a
 * single launch, which suits forests of many small or uneven trees;
 * LAML_PASS_HYBRID (the default; environment LAML_PASS_MODE=0 or 1 selects
 * another default) combines them.  The tree walk needs full inside storage:
 * LAML_PASS_TREES refuses level checkpointing, and turning checkpointing on in
 * LAML_PASS_HYBRID switches the forest to LAML_PASS_LEVELS.  All modes
 * compute the same values; only the order of floating-point sums over sites
 * may differ.
 * No counterpart in the reference. */
#define LAML_PASS_LEVELS 0
#define LAML_PASS_TREES 1
/* LAML_PASS_HYBRID: level launches for the levels below top_level (see
 * laml_forest_set_shapes) and one tree-walk launch for the levels from
 * top_level up, the root step included: the small upper levels of every tree
 * run in one launch instead of one launch each. */
#define LAML_PASS_HYBRID 2
int laml_forest_set_pass_mode(laml_forest* f, int32_t mode);
int32_t laml_forest_pass_mode(laml_forest* f);

/* Cherry fusion (off by default; environment LAML_FUSE_CHERRIES=1 turns the
 * default on; laml_amd.autotune tries it): in level mode with row kernels
 * (more than 32 sites) and full inside storage, the nodes whose two children
 * are leaves (the whole height-1 level) get no launch and no rows in the
 * passes: their parents form their inside vectors from the leaf codes and run
 * their posterior step in the outside pass.  Values are identical except for
 * the order of floating-point sums over sites.  Needs every tree to have at
 * least three leaves (otherwise it stays off).  laml_posteriors and
 * laml_quartet_em read every node's rows, so they fail after a fused E-step:
 * turn fusion off and rerun it.  laml_forest_fusion reports whether the next
 * pass fuses (1) or not (0).
 * No counterpart in the reference. */
int laml_forest_set_fusion(laml_forest* f, int32_t on);
int32_t laml_forest_fusion(laml_forest* f);

/* Heavy-path chains (environment LAML_CHAINS=<flags> sets the default flags;
 * without it, outside chains are on for forests of at least 32,768
 * (internal node, 64-site chunk) pairs; laml_amd.autotune tries them): with
 * row kernels (more than 32 sites), full inside storage and no cherry
 * fusion, the levels below a pass's chain-region top (and below the split:
 * top_level, or the hybrid walk's first level) run as chains of nodes, each
 * node followed by its child with more leaves below it; the levels from the
 * region top to the split keep level launches.  One wavefront walks a chain
 * at a 64-site chunk and keeps that child's state in registers:
 * LAML_CHAINS_OUTSIDE, the outside pass walks chains top-down and writes and
 * reads only the other children's outside rows; LAML_CHAINS_INSIDE, the
 * inside pass walks them bottom-up and reads only the other children's
 * inside rows (every inside row is still written).  Chains run in rounds
 * (one launch each) ordered by their dependences; in an outside round with
 * wavefronts to spare one wavefront runs several short chains of one tree
 * back to back, up to 8 steps (environment LAML_CHAIN_PACK sets that cap at
 * forest creation, 1: one chain per wavefront).  Results are
 * bit-identical to level launches with one chunk per wavefront.  After an
 * outside pass with chains the heavy children's outside rows are not
 * stored: laml_posteriors and laml_quartet_em then recompute every outside
 * row first (one outside pass in level launches).
 * which: LAML_CHAINS_* flags; inside_top / outside_top: chain-region top
 * level of each pass (0: up to the split; defaults 5 and 10).  Measured on
 * MI355X, outside chains below level 10 beat level launches on forests of
 * that size; inside chains were slower on every forest tried (opt-in).
 * laml_forest_chains reports the flags the next passes use.
 * No counterpart in the reference. */
#define LAML_CHAINS_OUTSIDE 1
#define LAML_CHAINS_INSIDE 2
int laml_forest_set_chains(laml_forest* f, int32_t which, int32_t inside_top, int32_t outside_top);
int laml_forest_get_chains(laml_forest* f, int32_t* which, int32_t* inside_top, int32_t* outside_top);
int32_t laml_forest_chains(laml_forest* f);

int laml_forest_get_shapes(laml_forest* f, int32_t* inside_cpw, int32_t* outside_cpw, int32_t* top_level);
int laml_forest_set_shapes(laml_forest* f, const int32_t* inside_cpw, const int32_t* outside_cpw,
                           int32_t top_level);

/* Per-node az-partition labels, int16 [num_nodes * K] row-major (0 = 'z',
 * -1 = '?', j>0 = local edit j), copied to host.  Replaces reading node.alpha
 * after ML_solver.az_partition (ML_solver.py:125-154). */
int laml_forest_labels(laml_forest* f, int16_t* labels_host, int64_t capacity);

/* Negative log-likelihood.
 * d: device [num_nodes]; nu, phi: device [num_problems];
 * nll: device [num_problems] (out); site_llh: device [sum_t K_t] (out, may be
 * NULL) per-tree per-site log-likelihood.
 * Replaces: ML_solver.negative_llh -> EM_solver.lineage_llh -> Estep_in_llh
 * (ML_solver.py:252-254, EM_solver.py:91-141). */
int laml_nll(laml_forest* f, const double* d, const double* nu, const double* phi,
             double* nll, double* site_llh, void* stream);

/* E-step: inside, outside and posterior expected counts.
 * counts: device [num_nodes * 5] (out) raw sums over sites of S0..S4 of each
 * branch; rtilde: device [num_problems] (out) sum over missing leaf entries of
 * P(not silenced | D); nll: device [num_problems] (out).
 * Replaces: EM_solver.Estep = Estep_in_llh + Estep_out_llh + __out_alpha_up__
 * + Estep_posterior (EM_solver.py:91-305) and the count sums of Mstep
 * (EM_solver.py:318-348). */
int laml_estep(laml_forest* f, const double* d, const double* nu, const double* phi,
               double* counts, double* rtilde, double* nll, void* stream);

/* Outside pass and posterior counts only, on the inside state left resident
 * by the preceding laml_nll / laml_estep call on this forest with the same
 * parameters (stream-ordered).  Outputs as laml_estep.
 * Replaces: EM_solver.Estep_out_llh + __out_alpha_up__ + Estep_posterior
 * (EM_solver.py:143-300). */
int laml_outside(laml_forest* f, const double* phi, double* counts, double* rtilde, void* stream);

/* Gradient of the negative log-likelihood with respect to every branch length
 * (device [num_nodes]), nu and phi (device [num_problems] each): the outside
 * pass run as the adjoint of the inside pass left resident by the preceding
 * laml_nll with the same parameters, then the Fisher identity on the raw
 * expected transition counts.  Defined for nu > 0 and 0 < phi < 1.
 * No counterpart in the reference, which has no gradient; its expected
 * counts are those of EM_solver.Estep_posterior (EM_solver.py:247-300) and the
 * score terms those of the M-step objective (EM_solver.py:349-357) and of
 * fast_em/laml.py:62-78. */
int laml_grad(laml_forest* f, const double* d, const double* nu, const double* phi, double* grad_d,
              double* grad_nu, double* grad_phi, void* stream);

/* M-step, in place on d, nu, phi (device): phi closed form, per-branch exact
 * maximisation of the branch-length objective over [dmin, dmax] for free
 * branches, then the 1-D nu problem on [0, 1].  counts/rtilde as produced by
 * laml_estep.
 * Replaces: EM_solver.Mstep without ultrametric constraint
 * (EM_solver.py:307-459; MOSEK calls at :365, :405). */
int laml_mstep(laml_forest* f, const double* counts, const double* rtilde,
               double* d, double* nu, double* phi, void* stream);

/* Full EM loop on the device, stopping each problem at
 * |dllh/llh| < tol or after max_iter M-steps (tol <= 0: run max_iter).
 * d, nu, phi: device, updated in place.  nll: device [num_problems] final
 * negative log-likelihood; iters: device int32 [num_problems] EM iteration
 * count with the reference's numbering; trace: device [(max_iter+1) *
 * num_problems] or NULL, negative log-likelihood after each M-step (row 0 =
 * start).
 * Replaces: EM_solver.EM_optimization (EM_solver.py:461-512). */
int laml_em(laml_forest* f, double* d, double* nu, double* phi, int32_t max_iter, double tol,
            double* nll, int32_t* iters, double* trace, void* stream);

/* Posterior state probabilities of every (node, site) on the inside and
 * outside state left by a preceding laml_estep with the same phi:
 * post0 = P(node unedited | D), postS = P(node silenced | D), host double
 * arrays in the layout of laml_forest_labels (leaf rows, then internal rows,
 * K_t entries each).  Needs full inside storage.
 * Replaces: the post0 / post1 of EM_solver.Estep_posterior
 * (EM_solver.py:247-300) read by the annotation writer (runner.py:367-435). */
int laml_posteriors(laml_forest* f, const double* phi, double* post0_host, double* postS_host, int64_t capacity,
                    void* stream);

/* laml_posteriors without any allocation, for callers that cannot allocate
 * device memory during the call (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The same results as
 * laml_posteriors, bit for bit; the staging rows live in a fixed file-scope
 * device array of 8388608 doubles (64 MiB) that each device holds from the
 * library's first forest on it until the process ends.  A forest whose
 * 2 x (padded rows + leaf_sites + internal sites) exceeds that is refused
 * with LAML_EINVAL before any work (laml_posteriors has no such limit).
 * Every call on one device shares the array: calls that overlap in time on
 * one device (other threads, other streams) race and must be serialised by
 * the caller.  The call ends with a synchronisation of `stream`.
 * No counterpart in the reference. */
int laml_posteriors_noalloc(laml_forest* f, const double* phi, double* post0_host, double* postS_host,
                            int64_t capacity, void* stream);

/* Quartet-local EM of NNI candidates.  With nu, phi and every branch but
 * the five around an NNI move fixed, the subtrees hanging off the move keep
 * their inside vectors and the data above it its outside state, so each
 * candidate's EM runs on those resident values alone, O(K) per EM step.
 * Candidate c, after the move: v has children u and C, u has children A and
 * W, p is v's parent and s v's sibling, all nodes of this forest's tree
 * as it was before the move (p = s = -1 when v is a root).
 * quartet_nodes: device int32 [5 * n] (p, s, A, W, C); lengths_in /
 * lengths_out: device [5 * n] branch lengths above (A, W, u, C, v);
 * optimise: device uint8 [n], bit j set = branch j is optimised; llh_const:
 * device [n] log-likelihood of the problem's other trees; nu, phi: device
 * [num_problems] (fixed); nll: device [n] final negative log-likelihood;
 * iters: device int32 [n] with the numbering of laml_em.  Needs the inside
 * and outside state of a preceding laml_estep on this forest with the same
 * parameters (stream-ordered) and full inside storage.
 * Replaces: the per-candidate solver build + EM_optimization of
 * Topology_search.apply_nni / Topology_search_parallel.apply_nni with
 * local_brlen_opt (Topology_search.py:242-322, Topology_search_parallel.py:57-63). */
int laml_quartet_em(laml_forest* f, int32_t num_candidates, const int32_t* quartet_nodes, const double* lengths_in,
                    const uint8_t* optimise, const double* llh_const, const double* nu, const double* phi,
                    int32_t max_iter, double tol, double* lengths_out, double* nll, int32_t* iters, void* stream);

/* Height-class model of the ultrametric constraint (see laml_amd/ultra.py).
 * Classes are sorted by (class level, problem); level_problem_off[l*P + p]
 * is the first class of problem p at class level l. */
typedef struct {
  int32_t num_classes;
  int32_t num_levels;
  const int32_t* level_problem_off; /* [num_levels * num_problems + 1]        */
  const int32_t* up_cls;            /* [num_nodes] parent-side class, -1 pinned */
  const int32_t* dn_cls;            /* [num_nodes] child-side class, -1 pinned  */
  const double* up_off;             /* [num_nodes] offset / absolute height    */
  const double* dn_off;             /* [num_nodes]                             */
  const int32_t* cls_parent;        /* [num_classes] -1 at the top             */
  const int32_t* cls_pedge;         /* [num_classes] branch to the parent class */
  const int32_t* inc_off;           /* [num_classes + 1]                       */
  const int32_t* inc_branch;        /* incident free branches                  */
  const int8_t* inc_sign;           /* +1 where the class is the parent side   */
  const uint8_t* problem_has_free;  /* [num_problems]                          */
} laml_ultra_desc;

/* Switch the M-step of every problem to the ultrametric-constrained branch
 * problem (NULL switches back to the unconstrained one).
 * Replaces: EM_solver.ultrametric_constr + the constrained MOSEK solve
 * (EM_solver.py:53-89, 349-366, used when ultra_constr=True). */
int laml_forest_set_ultrametric(laml_forest* f, const laml_ultra_desc* u);

/* Starting heights of the variable classes (host [num_classes]); they must
 * give every free branch a length strictly inside (dmin, dmax).  Later
 * ultrametric M-steps warm-start from the previous solution. */
int laml_forest_set_heights(laml_forest* f, const double* heights_host);

/* Per-problem status of the last ultrametric M-step (0 ok, 1 line search
 * stalled, 2 singular Newton system; heights are kept at the last accepted
 * point in both failure cases), host int32 [num_problems]. */
int laml_forest_ultra_status(laml_forest* f, int32_t* status_host);

/* Per-problem status of the last M-step each problem ran (by laml_mstep or
 * inside laml_em), host int32 [num_problems]: 0 = both subproblems solved;
 * bit LAML_MSTEP_FAILED_D = a branch length came out non-finite, or the
 * ultrametric branch solve hit a singular Newton system; bit
 * LAML_MSTEP_FAILED_NU = the nu solve did not converge within its evaluation
 * cap (nu is left at its last iterate).  Problems that have not run an
 * M-step read 0.
 * Replaces: the (success, status) pair of EM_solver.Mstep, whose status
 * strings "failed_d" / ",failed_nu" (EM_solver.py:441-459) EM_optimization
 * returns from its last M-step (EM_solver.py:490-512). */
#define LAML_MSTEP_FAILED_D 1
#define LAML_MSTEP_FAILED_NU 2
int laml_forest_mstep_status(laml_forest* f, int32_t* status_host);

/* Text of the last error on this thread ("" if none). */
const char* laml_last_error(void);

/* Library version string. */
const char* laml_version(void);

#ifdef __cplusplus
}
#endif
#endif /* LAML_PMM_H */
