/*
 * phylo_hip.h -- C ABI of the gfx950 (MI355X) implementation of phylocaml's two
 * site-pattern hot paths:
 *
 *   - Fitch parsimony, the NonAdditive_c character type
 *       (reference: lib/nonAdditive_c.ml, lib/bitvector/bv.c:46-55,147-161)
 *   - Felsenstein pruning, the Likelihood_c character type over MlModel
 *       (reference: lib/likelihood_c.ml:1-33, lib/mlModel.ml, lib/mlmodel.c)
 *
 * Every entry point is extern "C", takes plain pointers and sizes, and returns an
 * int status (PHY_OK == 0).  On failure the message is available from
 * phy_last_error(ctx); the OCaml stubs turn a non-zero status into
 * caml_failwith(phy_last_error(ctx)), which is how the reference's C stubs report
 * errors (lib/mlmodel.c:110-120,191-193,249-257).
 *
 * Device pointers ("d_" prefix) must point to device memory of the context's
 * device.  Work is enqueued on the context's stream and is asynchronous unless a
 * function says otherwise; results written to device pointers are valid after
 * phy_ctx_sync() or any later stream-ordered operation.
 *
 * Data layouts in HBM
 *   Packed state sets ("nibble blocks"): one 4-bit IUPAC set per site
 *     (A=1, C=2, G=4, T=8; lib/alphabet.ml:301-307), 16 sites per uint64 word,
 *     LSB first: site s lives in word s/16, bits 4*(s%16) .. 4*(s%16)+3.
 *     A block of n_sites sites occupies phy_words(n_sites) words; padding
 *     nibbles of the last word must be zero on input and are written as zero.
 *   Tree node arrays: node-major, (2N-1) rows of `stride` words each; rows
 *     0..N-1 are the leaves (taxa), rows N..2N-2 the internal nodes.
 *   Partial likelihoods: node-major, then rate-major, then state-major, then
 *     pattern (structure of arrays): L[v][r][i][p], p contiguous.
 */
#ifndef PHYLO_HIP_H
#define PHYLO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------- */
/* Status codes                                                              */
/* ------------------------------------------------------------------------- */
enum {
  PHY_OK = 0,
  PHY_EINVAL = 1,   /* bad argument, malformed topology, size mismatch        */
  PHY_EDEVICE = 2,  /* HIP runtime error                                      */
  PHY_ENOMEM = 3,   /* allocation failure                                     */
  PHY_ENUMERIC = 4, /* NaN in a rate matrix, non-reversible model             */
  PHY_ERANGE = 5    /* a size beyond what a kernel supports                   */
};

/* Likelihood evaluation forms (bit flags). */
enum {
  PHY_F64 = 0,      /* FP64 partials, int32 exponent per (pattern, rate)     */
  PHY_F32 = 1,      /* FP32 partials, int32 exponent per (pattern, rate)     */
  PHY_FUSED = 2     /* keep interior partials on chip (lnL only, no stored
                       partials, no derivatives); a tree whose stack does
                       not fit in LDS runs the stored form instead, with
                       the same result                                       */
};

/* Gamma rate-category modes. */
enum {
  PHY_GAMMA_REF = 0,  /* rate_i = quantile(i/n) of Gamma(shape a, scale b):
                         exactly MlModel.gamma_rates (lib/mlModel.ml:93-99)  */
  PHY_GAMMA_MEAN = 1  /* Yang (1994) mean of each equal-probability bin of
                         Gamma(shape a, rate a): mean rate 1
                         (the documented intent, lib/mlModel.mli:12)          */
};

typedef struct phy_ctx phy_ctx;
typedef struct phy_topo phy_topo;
typedef struct phy_model phy_model;

/* Number of uint64 words of a nibble block of n_sites sites. */
int64_t phy_words(int64_t n_sites);

/* ------------------------------------------------------------------------- */
/* Context: device, stream, error string.                                    */
/* Replaces nothing in the reference (the reference is single-threaded OCaml */
/* with no device); the OCaml stubs keep one context per process.            */
/* ------------------------------------------------------------------------- */
int phy_ctx_create(int device, void* hip_stream, phy_ctx** out);
void phy_ctx_destroy(phy_ctx* ctx);
const char* phy_last_error(const phy_ctx* ctx);
int phy_ctx_sync(phy_ctx* ctx);
int phy_ctx_device(const phy_ctx* ctx);
void* phy_ctx_stream(const phy_ctx* ctx);

/* Measurement hooks (no counterpart in the reference).  With the timer on,
 * each entry point records a HIP event pair on the context's stream around
 * its dominant kernel launches, labelled "prune_down", "prune_deriv",
 * "prune_down_batch", "fitch_down", "fitch_up", "aa_prune", "sankoff_down",
 * "mlp_forward", "mlp_backward" or "copy_stream".  phy_ctx_kernel_timer
 * switches it on or off and clears what was recorded;
 * phy_ctx_kernel_time waits for the recorded spans and returns the total
 * milliseconds and the number of spans of one label since the timer was
 * switched on. */
int phy_ctx_kernel_timer(phy_ctx* ctx, int32_t enable);
int phy_ctx_kernel_time(phy_ctx* ctx, const char* label, double* total_ms, int64_t* count);

/* STREAM-style device copy (16-byte non-temporal loads and stores; both
 * pointers 16-byte aligned): the bandwidth reference the HBM-bound kernels
 * are compared with.  No counterpart in the reference. */
int phy_copy_stream(phy_ctx* ctx, void* d_dst, const void* d_src, size_t bytes);

/* Device memory for callers without their own allocator (the OCaml custom
 * blocks; lib/bitvector/bv.c:163-189 mallocs/frees host vectors the same way). */
int phy_dev_alloc(phy_ctx* ctx, size_t bytes, void** d_out);
int phy_dev_free(phy_ctx* ctx, void* d_ptr);
int phy_copy_h2d(phy_ctx* ctx, void* d_dst, const void* h_src, size_t bytes);
int phy_copy_d2h(phy_ctx* ctx, void* h_dst, const void* d_src, size_t bytes); /* synchronous */

/* ------------------------------------------------------------------------- */
/* Rooted binary topology.                                                   */
/* Replaces the traversal order of Tree.post_order_edges / pre_order_nodes   */
/* (lib/tree.ml:123-187) driving Node.Make1D.median_2 (lib/node.ml:183-198): */
/* the order is compiled once into device-side programs.                     */
/*   n_leaves  N >= 2; nodes 0..N-1 are leaves, N..2N-2 internal.            */
/*   children  2*(N-1) ints: children of internal node N+k are               */
/*             children[2k], children[2k+1].                                 */
/*   root      an internal node id that is nobody's child.                   */
/* Returns PHY_EINVAL for a malformed tree (a node with two parents, an      */
/* unreachable node, a cycle, an id out of range).                           */
/* ------------------------------------------------------------------------- */
int phy_topo_create(phy_ctx* ctx, int32_t n_leaves, const int32_t* children,
                    int32_t root, phy_topo** out);
void phy_topo_destroy(phy_topo* topo);
/* Host-side facts about the compiled programs: LDS stack depths of the
 * post-order and pre-order walks, and the parent of every node (-1 = root). */
int phy_topo_info(const phy_topo* topo, int32_t* down_stack_depth,
                  int32_t* up_stack_depth, int32_t* parent_out /* 2N-1 or NULL */);
/* The compiled programs themselves (for host-side checking): post-order
 * entries {v, A, B, flags} (4 x (N-1) ints), its leaf order (N), pre-order
 * entries {v, a, b, flags} (4 x (N-1)), its node load order (2N-1).  Each
 * argument may be NULL. */
int phy_topo_programs(const phy_topo* topo, int32_t* down_prog, int32_t* down_fifo,
                      int32_t* up_prog, int32_t* up_fifo);
/* The split programs of the Fitch passes (for host-side checking).  A tree of
 * >= 128 leaves is cut under n_parts cut nodes (returned in cut_nodes, at
 * most 32; n_parts = 0 means the tree runs whole).  Program k of pass `up`
 * (0 = post-order: parts 0..n_parts-1, then the top at n_parts; 1 =
 * pre-order: the top at 0, then part k at k + 1) has *n_ops entries (4 ints
 * each) and *n_fifo operand ids; prog and fifo may be NULL to query the
 * sizes.  A pre-order part starts with source SRC_EXT = 3 (its parent's final
 * set is read from memory); the top marks cut children with flag bits 64 (a)
 * and 128 (b).  No counterpart in the reference: its traversal is
 * sequential OCaml (lib/tree.ml:123-187). */
int phy_topo_split(const phy_topo* topo, int32_t* n_parts, int32_t* cut_nodes);
int phy_topo_split_program(const phy_topo* topo, int32_t up, int32_t k, int32_t* prog,
                           int32_t* n_ops, int32_t* fifo, int32_t* n_fifo);

/* ------------------------------------------------------------------------- */
/* Fitch parsimony: NodeData operations of NonAdditive_c on nibble blocks.   */
/* Weights: d_weights[site] (uint32), or NULL for unit weights.  Costs are   */
/* exact unsigned integers (the reference accumulates float weights,         */
/* lib/nonAdditive_c.ml:28-31, exact while the sum stays below 2^53).        */
/* Cost outputs are single uint64 device values that the call overwrites.    */
/* ------------------------------------------------------------------------- */

/* median_2 x y (lib/nonAdditive_c.ml:19-35; bv_fitch lib/bitvector/bv.c:147-161):
 * out_s = x_s & y_s if non-empty, else x_s | y_s; cost = sum_s w_s [x_s & y_s = 0]. */
int phy_fitch_median2(phy_ctx* ctx, const uint64_t* d_x, const uint64_t* d_y,
                      uint64_t* d_out, int64_t n_sites, const uint32_t* d_weights,
                      uint64_t* d_cost);

/* filter_codes / filter_codes_comp of NodeData.S (lib/nodeData.ml:9-10, TODO
 * in lib/nonAdditive_c.ml:9-10): the characters h_idx[0 .. n_idx-1] (host
 * array of site indices, each in 0 .. n_sites-1, checked) of a nibble block,
 * in that order, into a new block of words(n_idx) words, and their weights
 * when d_weights is given (d_weights_out n_idx uint32; both null or both
 * given).  Synchronous on the context's stream. */
int phy_fitch_select(phy_ctx* ctx, const uint64_t* d_states, int64_t n_sites,
                     const uint32_t* d_weights, const int64_t* h_idx, int64_t n_idx,
                     uint64_t* d_states_out, uint32_t* d_weights_out);

/* median_3 a b c (lib/nonAdditive_c.ml signature slot, TODO in the reference):
 * the 3-way Fitch median: per site the states contained in the largest number
 * of {a,b,c}; cost = sum_s w_s (3 - that number). */
int phy_fitch_median3(phy_ctx* ctx, const uint64_t* d_a, const uint64_t* d_b,
                      const uint64_t* d_c, uint64_t* d_out, int64_t n_sites,
                      const uint32_t* d_weights, uint64_t* d_cost);

/* distance_1 a b (bv_distance, lib/bitvector/bv.c:46-55, weighted):
 * sum_s w_s [a_s & b_s = 0]. */
int phy_fitch_distance1(phy_ctx* ctx, const uint64_t* d_a, const uint64_t* d_b,
                        int64_t n_sites, const uint32_t* d_weights, uint64_t* d_cost);

/* distance_2 a b c (TODO in the reference, lib/nonAdditive_c.ml:45): cost of
 * inserting c on the edge (a,b) = distance_1 (median_2 a b) c. */
int phy_fitch_distance2(phy_ctx* ctx, const uint64_t* d_a, const uint64_t* d_b,
                        const uint64_t* d_c, int64_t n_sites,
                        const uint32_t* d_weights, uint64_t* d_cost);

/* adjust_3 (TODO in the reference, lib/nonAdditive_c.ml:39; called from
 * Node.Make1D.readjust_3, lib/node.ml:239-256): Fitch (1971) final set of an
 * internal node v from its preliminary set P_v, its children's preliminary
 * sets P_a, P_b and its parent's final set F_p:
 *   F_p subset of P_v          -> F_v = F_p
 *   P_a & P_b non-empty        -> F_v = P_v | (F_p & (P_a | P_b))
 *   otherwise                  -> F_v = P_v | F_p
 * d_changed (nullable) receives the number of sites where F_v != P_v. */
int phy_fitch_adjust3(phy_ctx* ctx, const uint64_t* d_pv, const uint64_t* d_pa,
                      const uint64_t* d_pb, const uint64_t* d_fp, uint64_t* d_out,
                      int64_t n_sites, uint64_t* d_changed);

/* Batched insertion costs (distance_2 over many queries, e.g. the SPR
 * neighbourhood): for query q = (c, a, b) (d_q: 3 ints per query, row ids),
 * d_costs[q] = sum_s w_s [P_c,s & med(F_a,s, F_b,s) = 0], med the Fitch
 * median, P rows from d_pre and F rows from d_fin (both rows x stride). */
int phy_fitch_insertion_costs(phy_ctx* ctx, const uint64_t* d_pre, const uint64_t* d_fin,
                              int64_t stride, int64_t n_sites, const uint32_t* d_weights,
                              int64_t n_queries, const int32_t* d_q, uint64_t* d_costs);

/* Whole-tree down-pass: Node.Make1D.median_2 (lib/node.ml:183-198) applied in
 * post-order to every internal node, in one launch.
 *   d_sets      (2N-1) rows x stride words; rows 0..N-1 (leaves) are read,
 *               rows N..2N-2 receive the preliminary sets.
 *   d_site_cost n_sites uint32: number of union events at each site.
 *   d_node_cost (nullable) N-1 uint64: weighted local cost of each internal
 *               node N+k (NonAdditive_c.t.cost of that node).
 *   d_total     1 uint64: sum_s w_s site_cost_s.
 * A tree of >= 128 leaves runs split (parts side by side, then the top);
 * the per-part site counts live in a device buffer owned by the topology,
 * allocated on the first call and grown with n_sites, so two down-passes of
 * one topology must not run at the same time on different streams. */
int phy_fitch_downpass(phy_ctx* ctx, const phy_topo* topo, uint64_t* d_sets,
                       int64_t stride, int64_t n_sites, const uint32_t* d_weights,
                       uint32_t* d_site_cost, uint64_t* d_node_cost,
                       uint64_t* d_total);

/* Whole-tree up-pass (final states; Node.Make1D.final_states / uppass_*,
 * TODO at lib/node.ml:258-270): pre-order, adjust_3's rule at internal nodes,
 * F_root = P_root, and at a leaf L with parent final F_p:
 * F_leaf = F_p & L where non-empty, else L.
 *   d_sets   as written by phy_fitch_downpass;  d_final (2N-1) rows x stride. */
int phy_fitch_uppass(phy_ctx* ctx, const phy_topo* topo, const uint64_t* d_sets,
                     uint64_t* d_final, int64_t stride, int64_t n_sites);

/* Whole-tree Sankoff (1975) parsimony under an integer transformation cost
 * matrix (the PTcm.X spec of lib/pTcm.ml:5-50, costs as in the median search
 * of lib/costMatrix.ml:52-65: a child in state j under a parent in state k
 * costs tcm[j][k]; no Sankoff NodeData exists in the reference), in one
 * launch: C_leaf[k] = 0 for k in the leaf's set, else infinite;
 * C_v[k] = sum over the children c of min_j (C_c[j] + tcm[j][k]);
 * site cost = min_k C_root[k].
 *   n_states     K, 2..8;  d_tcm K x K int32 (device), entries >= 0 with
 *                (2N - 2) * max(tcm) < 2^31 (the int32 sums cannot overflow)
 *   d_leaves     N rows x stride bytes: the state set of leaf i at site s is
 *                bit k of d_leaves[i * stride + s]; bits >= K are ignored
 *                and an empty set means every state (missing data)
 *   d_site_cost  n_sites uint32;  d_total 1 uint64: sum_s w_s site_cost_s */
int phy_sankoff_downpass(phy_ctx* ctx, const phy_topo* topo, int32_t n_states,
                         const int32_t* d_tcm, const uint8_t* d_leaves, int64_t stride,
                         int64_t n_sites, const uint32_t* d_weights, uint32_t* d_site_cost,
                         uint64_t* d_total);

/* ------------------------------------------------------------------------- */
/* MlModel native stubs (host, FP64, synchronous; no LAPACK).                */
/* ------------------------------------------------------------------------- */

/* Replaces likelihood_CAML_diagonalize_gtr / diagonalize_gtr
 * (lib/mlmodel.c:208-271; called from MlModel.diagonalize, lib/mlModel.ml:554-583).
 * Q (n x n, row-major) must be a reversible rate matrix.  Writes U, D (n x n,
 * eigenvalues on the diagonal, like the reference's diag(D,n)) and Ui with
 * Q = U diag(D) Ui, computed by cyclic Jacobi on the symmetric
 * Pi^1/2 Q Pi^-1/2.  A non-reversible Q gives PHY_ENUMERIC (the reference's
 * "Imaginary eigenvalues" failure, lib/mlmodel.c:248-250); NaN gives PHY_ENUMERIC
 * (the assert of lib/mlModel.ml:557-568). */
int phy_diagonalize_gtr(int32_t n, const double* Q, double* U, double* D, double* Ui);

/* Replaces likelihood_CAML_diagonalize_sym (lib/mlmodel.c:163-205): A (n x n,
 * symmetric, row-major) is overwritten by its eigenvectors as rows (so that
 * A = U^T diag(D) U), D (n x n) receives the eigenvalues on its diagonal. */
int phy_diagonalize_sym(int32_t n, double* A, double* D);

/* Replaces likelihood_CAML_compose_gtr (lib/mlmodel.c:325-365):
 * P = exp(Q t) row-major (P[i][j] = Pr(j | i, t)); t == -1 -> Q; t < 1e-10 -> I. */
int phy_compose_gtr(int32_t n, const double* U, const double* D, const double* Ui,
                    double t, double* P);

/* Replaces likelihood_CAML_compose_sym (lib/mlmodel.c:280-323), including its
 * single-precision branch length (the `const float t` parameter at :280). */
int phy_compose_sym(int32_t n, const double* U, const double* D, float t, double* P);

/* Replaces MlModel.gamma_rates (lib/mlModel.ml:93-99) for mode PHY_GAMMA_REF
 * (shape a, scale b), and gives Yang's mean-1 categories for PHY_GAMMA_MEAN
 * (shape a; b ignored).  Writes n rates. */
int phy_gamma_rates(int32_t mode, double a, double b, int32_t n, double* rates);

/* Replaces MlModel.m_gtr_independent + m_meanrate (lib/mlModel.ml:390-421,
 * 204-213): Q from priors pi (n) and the (n+1)(n-2)/2 free exchangeabilities
 * co (upper triangle, row-major, the last (n-1,n) rate fixed at 1.0);
 * diagonal = -row sum; Q /= sum_i -Q_ii pi_i. */
int phy_gtr_q(int32_t n, const double* pi, const double* co, double* Q);

/* ------------------------------------------------------------------------- */
/* Likelihood model on the device: MlModel.t (lib/mlModel.ml:53-63) reduced  */
/* to what pruning needs: pi, the eigen-decomposition of Q, Gamma rates and  */
/* their probabilities.  n_states K in {4, 20}; n_rates R in 1..8 (the      */
/* pruning passes take R in {1, 2, 4, 8}, the 20-state batch R in {1, 2, 4};  */
/* other counts give PHY_ERANGE there).                                      */
/* ------------------------------------------------------------------------- */
int phy_model_create(phy_ctx* ctx, int32_t n_states, int32_t n_rates,
                     const double* pi, const double* Q, const double* rates,
                     const double* probs, phy_model** out);
void phy_model_destroy(phy_model* model);

/* P matrices of a set of branches on the device, FP64: for branch b and rate
 * r, P[b][r] = compose(rate_r * t_b) (the reference's caller scales t by the
 * rate: MlModel.compose takes one scalar, lib/mlModel.ml:585-587);
 * P1 = dP/dt and P2 = d2P/dt2 (nullable) = rate_r^k U diag(lambda^k
 * e^(lambda rate_r t_b)) Ui.  Layout [b][r][i][j]. */
int phy_model_pmatrices(phy_ctx* ctx, const phy_model* model, const double* d_t,
                        int32_t n_branches, double* d_P, double* d_P1, double* d_P2);

/* ------------------------------------------------------------------------- */
/* Felsenstein pruning: Likelihood_c (lib/likelihood_c.ml:1-33, all TODO in  */
/* the reference) for DNA (K = 4) with tips as nibble blocks.                */
/* ------------------------------------------------------------------------- */

/* Device workspace (bytes) that phy_lk_tree needs for this topology, model,
 * pattern count and form (PHY_F64 | PHY_F32, optionally | PHY_FUSED). */
int phy_lk_tree_work_size(const phy_topo* topo, const phy_model* model,
                          int64_t n_patterns, int32_t form, int32_t want_derivs,
                          size_t* bytes);

/* Whole-tree evaluation (Node.Make1D.median_2 in post-order + root_cost,
 * lib/node.ml:183-198,272), one launch for the partials (PHY_FUSED on a tree
 * of >= 256 leaves that splits into parts: two, the parts side by side and
 * then the top of the tree, with the same result bit for bit):
 *   d_tips     N rows x stride words: tip state sets (nibble blocks; a gap is
 *              all four bits = missing data).
 *   d_weights  n_patterns uint32 pattern weights.
 *   d_blen     2N-1 branch lengths (FP64); d_blen[v] is the branch above v,
 *              d_blen[root] is ignored.
 *   d_lnl      1 double: sum_p w_p (ln L_p), L_p = sum_r prob_r sum_i pi_i
 *              L_root[r][i][p] 2^e[r][p] (each rate class of a pattern keeps
 *              its own power-of-two exponent).
 *   d_site_lnl (nullable) n_patterns doubles: ln L_p.
 *   d_d1,d_d2  (nullable, not with PHY_FUSED) 2N-1 doubles: dlnL/dt_v and
 *              d2lnL/dt_v^2 for the branch above every non-root node v.  The
 *              two root branches form one unrooted edge: both receive the
 *              derivative for that edge.
 *   d_work     phy_lk_tree_work_size bytes.
 * The internal partials of node N+k (unless PHY_FUSED) are left in d_work and
 * can be read with phy_lk_tree_partials. */
int phy_lk_tree(phy_ctx* ctx, const phy_topo* topo, const phy_model* model,
                int32_t form, int64_t n_patterns, const uint64_t* d_tips,
                int64_t stride, const uint32_t* d_weights, const double* d_blen,
                double* d_lnl, double* d_site_lnl, double* d_d1, double* d_d2,
                void* d_work, size_t work_bytes);

/* No counterpart in the reference.  The whole-tree lnL of phy_lk_tree in
 * its fused form (form = PHY_FUSED, optionally | PHY_F32; no derivatives)
 * with no workspace argument and no allocation: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The workspace (P matrices of
 * every branch and rate class, one partial sum per 64 patterns) lives in a
 * fixed 8 MiB device array per GPU, resolved when a context is created and
 * held until the process ends; a tree and pattern count whose workspace
 * (the P matrices and the sums, at most phy_lk_tree_work_size) exceeds it,
 * or whose stack does not fit the fused form, gives PHY_ERANGE.  It runs the
 * whole tree per launch (no part rows); results are those of phy_lk_tree
 * with the same arguments, bit for bit.  Every call on one GPU shares that array: calls that overlap on
 * one GPU (other streams, other contexts) race. */
int phy_lk_tree_noalloc(phy_ctx* ctx, const phy_topo* topo, const phy_model* model,
                        int32_t form, int64_t n_patterns, const uint64_t* d_tips,
                        int64_t stride, const uint32_t* d_weights, const double* d_blen,
                        double* d_lnl, double* d_site_lnl);

/* Pointers into d_work of the stored partials and exponents of internal node
 * v (N <= v <= 2N-2) after phy_lk_tree without PHY_FUSED: the partial is
 * R*K rows of pitch values (FP32 or FP64 by form), row r*K + i holding
 * L[r][i][p] at column p (pitch = *pitch, the padded pattern count); the
 * exponents are R rows of pitch int32, e[r][p], so the value is
 * L[r][i][p] * 2^e[r][p]. */
int phy_lk_tree_partials(const phy_topo* topo, const phy_model* model,
                         int64_t n_patterns, int32_t form, int32_t want_derivs,
                         void* d_work, int32_t v, void** d_partial,
                         int32_t** d_exp, int64_t* pitch);

/* Per-node NodeData operations of Likelihood_c (lib/likelihood_c.ml:11-24,
 * all TODO in the reference) on values in the layout of the whole-tree
 * stored partials: R*K rows of `pitch` values (row r*K + i holds L[r][i][p],
 * FP32 or FP64 by form) and R rows of `pitch` int32 exponents (e[r][p]), the
 * value being L[r][i][p] * 2^e[r][p].  Every (pattern, rate class) is
 * rescaled on its own; the classes of a pattern are combined only in FP64.
 * A value's branch length t is the length of the branch from its node to the
 * point where the operation joins it.  d_weights (nullable: unit weights)
 * are n_patterns uint32 pattern weights.  All device results are enqueued on
 * the context's stream.
 *   leaf    (median_1 of a leaf, lib/likelihood_c.ml:14): L = indicator of
 *           the tip's state set in every class; e = 0.
 *   median2 (median_2, lib/likelihood_c.ml:15; the step Node.Make1D.median_2
 *           makes per block, lib/node.ml:183-198):
 *           L_out[r] = (P(rate_r t_x) L_x[r]) (*) (P(rate_r t_y) L_y[r]),
 *           e_out[r] = e_x[r] + e_y[r] + rescaling exponent.
 *   median3 (median_3, lib/likelihood_c.ml:16; the unrooted root of
 *           Node.Make1D.median_3, lib/node.ml:200-216): the same over three
 *           neighbours x, y, z; d_out/d_eout (nullable together) receive
 *           the joined value and d_lnl (nullable) its log-likelihood
 *           sum_p w_p ln sum_r prob_r sum_i pi_i L_out[r][i][p] 2^e[r][p],
 *           i.e. -cost of the tree seen from that node.
 *   root    (root_cost, lib/likelihood_c.ml:21): d_lnl = that sum at a root
 *           value (no further branch).
 *   edge    lnL, dlnL/dt and d2lnL/dt2 of the tree made of the conditional
 *           values x and y joined by one edge of length t (reversible model,
 *           so pi sits on either side) into d_out[3]: distance_1
 *           (lib/likelihood_c.ml:23) is -lnL at t = t_x + t_y; a Newton
 *           step on the edge uses d_out[1..2].
 *   adjust3 (adjust_3, lib/likelihood_c.ml:18, as Node.Make1D.readjust_3
 *           calls it, lib/node.ml:239-256): the length of the edge between
 *           a node's own value x (the median of its two children) and the
 *           value y of the rest of the tree beyond its parent, optimised by
 *           Newton-Raphson with step halving (at most max_iter steps, stop
 *           when a step moves t by <= tol * t; t kept in [t_lo, t_hi]).
 *           Synchronises the stream after each evaluation; *t_out receives
 *           the length, *lnl_out (nullable) its lnL, *n_evals (nullable) the
 *           number of edge evaluations. */
int phy_lk_leaf(phy_ctx* ctx, const phy_model* model, int32_t form,
                int64_t n_patterns, const uint64_t* d_tip, void* d_out,
                int32_t* d_eout, int64_t pitch);
int phy_lk_median2(phy_ctx* ctx, const phy_model* model, int32_t form,
                   int64_t n_patterns, const void* d_x, const int32_t* d_ex,
                   double t_x, const void* d_y, const int32_t* d_ey, double t_y,
                   void* d_out, int32_t* d_eout, int64_t pitch);
int phy_lk_median3(phy_ctx* ctx, const phy_model* model, int32_t form,
                   int64_t n_patterns, const void* d_x, const int32_t* d_ex,
                   double t_x, const void* d_y, const int32_t* d_ey, double t_y,
                   const void* d_z, const int32_t* d_ez, double t_z, void* d_out,
                   int32_t* d_eout, int64_t pitch, const uint32_t* d_weights,
                   double* d_lnl);
int phy_lk_root(phy_ctx* ctx, const phy_model* model, int32_t form,
                int64_t n_patterns, const void* d_root, const int32_t* d_eroot,
                int64_t pitch, const uint32_t* d_weights, double* d_lnl);
int phy_lk_edge(phy_ctx* ctx, const phy_model* model, int32_t form,
                int64_t n_patterns, const void* d_x, const int32_t* d_ex,
                const void* d_y, const int32_t* d_ey, int64_t pitch, double t,
                const uint32_t* d_weights, double* d_out);
int phy_lk_adjust3(phy_ctx* ctx, const phy_model* model, int32_t form,
                   int64_t n_patterns, const void* d_x, const int32_t* d_ex,
                   const void* d_y, const int32_t* d_ey, int64_t pitch,
                   const uint32_t* d_weights, double t0, double t_lo, double t_hi,
                   int32_t max_iter, double tol, double* t_out, double* lnl_out,
                   int32_t* n_evals);

/* Branch-length Newton-Raphson with its step control on the device, for the
 * branch-length hooks the reference declares (MlModel.num_parameters /
 * update, lib/mlModel.ml:193-196; unimplemented there).  An iteration is
 * phy_newton_propose, phy_lk_tree with derivatives at the proposal, then
 * phy_newton_accept: kernels only (no host copy, no synchronisation), so
 * iterations can be captured in a HIP graph.  d_state: 4 doubles, [lnL at
 * d_blen, 1 = next proposal is a Newton step / 0 = halve the last proposal
 * towards d_blen, accepted steps, halvings]; phy_newton_init sets it from
 * the lnL at the starting lengths.  Propose: every branch takes -d1/d2 where
 * d2 < 0, else moves half its length along the gradient, clamped to
 * [1e-8, 10]; the two root branches take one step on their sum and keep
 * their ratio.  Accept: when lnL at the proposal is not lower, d_blen, d_d1,
 * d_d2 and the state's lnL take the proposal's values. */
int phy_newton_init(phy_ctx* ctx, double* d_state, const double* d_lnl);
int phy_newton_propose(phy_ctx* ctx, const phy_topo* topo, const double* d_state,
                       const double* d_blen, const double* d_d1, const double* d_d2,
                       double* d_blen_new);
int phy_newton_accept(phy_ctx* ctx, const phy_topo* topo, double* d_state,
                      const double* d_lnl_new, double* d_blen, const double* d_blen_new,
                      double* d_d1, const double* d_d1_new, double* d_d2,
                      const double* d_d2_new);


/* ------------------------------------------------------------------------- */
/* SPR neighbourhood (TreeNeighborhoods.generate_spr, lib/treeNeighborhoods. */
/* ml:27-75, which never compiled; the operation of Tree.break/Tree.join,    */
/* lib/tree.ml:299-494, on the rooted layout of phy_topo_create).  Host.     */
/* A move (s, v): prune non-root node s (its parent p is removed, its        */
/* sibling x takes p's place with t_x += t_p, or becomes the root), then     */
/* regraft on the edge above v (p reused: children (v, s), t_p = t_v/2,      */
/* t_v = t_v/2).  v must lie outside the subtree of s and differ from p, x   */
/* and the pruned tree's root.                                               */
/* ------------------------------------------------------------------------- */
/* Count (and, if moves != NULL, write up to capacity (s, v) pairs of) all
 * moves of the tree. */
int phy_spr_moves(int32_t n_leaves, const int32_t* children, int32_t root,
                  int64_t* n_moves, int32_t* moves, int64_t capacity);
/* The tree after move (s, v): children_out 2(N-1), root_out, blen_out 2N-1. */
int phy_spr_apply(int32_t n_leaves, const int32_t* children, int32_t root,
                  const double* blen, int32_t s, int32_t v, int32_t* children_out,
                  int32_t* root_out, double* blen_out);
/* TBR and NNI (TreeNeighborhoods.generate_tbr / generate_nni, lib/
 * treeNeighborhoods.ml:27-75, never compiled: NNI is TBR with join distance
 * 1 and no reroot, SPR is TBR with no reroot).  A TBR move (s, w, v) reroots
 * the subtree of s on the edge above w (a node of that subtree below the
 * children of s; w = s keeps its root), then regrafts it as the SPR move
 * (s, v).  Rerooting puts the subtree's root at the middle of the edge above
 * w, reverses the path from s with its lengths and suppresses the former
 * root (its two edges add).  v = -1 keeps the subtree where it is (only
 * the reroot acts).  Join distance: edges between the edge the prune point
 * lies on and the edge above v in the unrooted pruned tree (the root's two
 * edges count as one; v = -1 is distance 0, other targets at distance 0
 * give the same tree and are left out).  Reroot distance: edges between the
 * subtree root's two edges and the edge above w (w = s: 0).  The moves are
 * every (s, w, v) within max_join / max_reroot (< 0: unlimited) except the
 * identity (s, s, -1).  The unrooted SPR neighbourhood is the union of the
 * moves with reroot distance 0 and those with join distance 0 (the two
 * sides of each cut edge). */
int phy_tbr_moves(int32_t n_leaves, const int32_t* children, int32_t root, int32_t max_join,
                  int32_t max_reroot, int64_t* n_moves, int32_t* moves, int64_t capacity);
int phy_tbr_apply(int32_t n_leaves, const int32_t* children, int32_t root,
                  const double* blen, int32_t s, int32_t w, int32_t v, int32_t* children_out,
                  int32_t* root_out, double* blen_out);
/* The 2(N - 3) NNI neighbours, one SPR move (s, v) each: for each internal
 * unrooted edge, the swaps of either child of its lower end with the
 * subtree across it. */
int phy_nni_moves(int32_t n_leaves, const int32_t* children, int32_t root, int64_t* n_moves,
                  int32_t* moves, int64_t capacity);
/* Exact lnL of every SPR neighbour, many candidate trees per fused launch
 * (chunk candidates at a time; form PHY_F32 or PHY_F64, fused implied):
 * d_lnl[i] = lnL of the tree after moves[i] (branch lengths as above, not
 * re-optimised).  children/root/blen/moves are host arrays of the base tree;
 * tips/weights are device arrays as for phy_lk_tree. */
int phy_lk_spr_work_size(int32_t n_leaves, const phy_model* model, int64_t n_patterns,
                         int32_t form, int32_t chunk, size_t* bytes);
int phy_lk_spr_scores(phy_ctx* ctx, const phy_model* model, int32_t form,
                      int64_t n_patterns, const uint64_t* d_tips, int64_t stride,
                      const uint32_t* d_weights, int32_t n_leaves, const int32_t* children,
                      int32_t root, const double* blen, int64_t n_moves,
                      const int32_t* moves, int32_t chunk, double* d_lnl, void* d_work,
                      size_t work_bytes);
/* The same scores without rebuilding a tree per move (form PHY_F32 or
 * PHY_F64, n_rates 1, 2, 4 or 8): per pruned subtree s, the base tree's
 * stored partials are repaired along the ancestors of s's parent and one
 * pre-order pass over the pruned tree scores the insertion at the middle of
 * every branch (the lnL of every regraft of s is a sum over patterns of
 * log sum_r prob_r [P(t_v/2)^T A_v] . ([P(t_v/2) D_v] (*) P(t_s) D_s)).
 * About 2N node operations per pruned subtree instead of N per move; the
 * values equal phy_lk_spr_scores' up to rounding.  s_batch pruned subtrees
 * run per launch; the workspace depends on the tree (its longest repair
 * path) and on s_batch.  d_base_lnl (nullable, device) receives the base
 * tree's lnL.  Replaces the per-move evaluation of the candidates of
 * TreeNeighborhoods' SPR generator (lib/treeNeighborhoods.ml:27-75), whose
 * trees the reference scores one Tree.join at a time (lib/tree.ml:400-494). */
int phy_lk_spr_incr_work_size(int32_t n_leaves, const int32_t* children, int32_t root,
                              const phy_model* model, int64_t n_patterns, int32_t form,
                              int32_t s_batch, size_t* bytes);
int phy_lk_spr_scores_incr(phy_ctx* ctx, const phy_model* model, int32_t form,
                           int64_t n_patterns, const uint64_t* d_tips, int64_t stride,
                           const uint32_t* d_weights, int32_t n_leaves, const int32_t* children,
                           int32_t root, const double* blen, int64_t n_moves,
                           const int32_t* moves, int32_t s_batch, double* d_lnl,
                           double* d_base_lnl, void* d_work, size_t work_bytes);
/* The same scoring prepared once and run many times: the plan holds the
 * moves grouped by pruned subtree and every pruned tree's programs on the
 * device (the library allocates them); a run takes device branch lengths
 * and a model with the plan's number of rate classes, and issues kernels
 * only (no host work, copy or synchronisation), so it can be captured in a
 * HIP graph and replayed after the lengths or the model change.  d_lnl
 * receives one lnL per move of the plan, in the order given; d_base_lnl
 * (nullable) the base tree's. */
typedef struct phy_spr_plan phy_spr_plan;
int phy_spr_plan_create(phy_ctx* ctx, const phy_model* model, int32_t form, int64_t n_patterns,
                        int32_t n_leaves, const int32_t* children, int32_t root,
                        int64_t n_moves, const int32_t* moves, int32_t s_batch,
                        phy_spr_plan** out);
int phy_spr_plan_work_size(const phy_spr_plan* plan, size_t* bytes);
int phy_spr_plan_run(phy_ctx* ctx, const phy_spr_plan* plan, const phy_model* model,
                     const uint64_t* d_tips, int64_t stride, const uint32_t* d_weights,
                     const double* d_blen, double* d_lnl, double* d_base_lnl, void* d_work,
                     size_t work_bytes);
void phy_spr_plan_destroy(phy_spr_plan* plan);
/* The same for TBR moves (s, w, v), 3 ints each (workspace as for SPR). */
int phy_lk_tbr_scores(phy_ctx* ctx, const phy_model* model, int32_t form,
                      int64_t n_patterns, const uint64_t* d_tips, int64_t stride,
                      const uint32_t* d_weights, int32_t n_leaves, const int32_t* children,
                      int32_t root, const double* blen, int64_t n_moves,
                      const int32_t* moves, int32_t chunk, double* d_lnl, void* d_work,
                      size_t work_bytes);
/* The TBR scores without rebuilding a tree per move (form PHY_F32 or
 * PHY_F64, n_rates 1, 2, 4 or 8), the incremental SPR scoring extended to
 * rerooted subtrees: moves are grouped into slots (a pruned subtree with up
 * to 32 of its rerootings); per slot the ancestors of s's parent are
 * repaired, one pre-order pass over the subtree of s gives the root vector
 * of each needed rerooting, and one pre-order pass over the pruned tree
 * scores every needed (target, rerooting) pair, keeping the attachment
 * (v = -1) included.  Moves are validated as phy_tbr_apply does; the lnL are
 * those of the trees phy_tbr_apply builds.  The workspace depends on the
 * moves (slots, rerootings per slot, pairs per batch of s_batch slots).
 * A plan made by phy_tbr_plan_create is run, sized and destroyed by the
 * phy_spr_plan functions above.  Replaces the reference's per-candidate TBR
 * scoring through lib/treeNeighborhoods.ml:27-75 and lib/tree.ml:400-494. */
int phy_lk_tbr_incr_work_size(int32_t n_leaves, const int32_t* children, int32_t root,
                              const phy_model* model, int64_t n_patterns, int32_t form,
                              int64_t n_moves, const int32_t* moves, int32_t s_batch,
                              size_t* bytes);
int phy_lk_tbr_scores_incr(phy_ctx* ctx, const phy_model* model, int32_t form,
                           int64_t n_patterns, const uint64_t* d_tips, int64_t stride,
                           const uint32_t* d_weights, int32_t n_leaves, const int32_t* children,
                           int32_t root, const double* blen, int64_t n_moves,
                           const int32_t* moves, int32_t s_batch, double* d_lnl,
                           double* d_base_lnl, void* d_work, size_t work_bytes);
int phy_tbr_plan_create(phy_ctx* ctx, const phy_model* model, int32_t form, int64_t n_patterns,
                        int32_t n_leaves, const int32_t* children, int32_t root,
                        int64_t n_moves, const int32_t* moves, int32_t s_batch,
                        phy_spr_plan** out);

/* ------------------------------------------------------------------------- */
/* Learned SPR proposal: a small MLP (n_in -> H -> H -> 1, GELU) trained and */
/* run by hand-written kernels (BF16-input / FP32-accumulate MFMA GEMMs,     */
/* FP32 master weights, Adam).  No counterpart in the reference; it ranks    */
/* TreeNeighborhoods SPR moves (lib/treeNeighborhoods.ml) before exact       */
/* re-scoring.  Parameters (flat FP32): W1[n_in][H] b1[H] W2[H][H] b2[H]     */
/* w3[H] b3.  Loss 0.5 mean((y - t)^2).  Device arrays: x [B][n_in], t, y [B]. */
/* ------------------------------------------------------------------------- */
typedef struct phy_mlp phy_mlp;
int phy_mlp_create(phy_ctx* ctx, int32_t n_in, int32_t n_hidden, uint64_t seed, phy_mlp** out);
void phy_mlp_destroy(phy_mlp* mlp);
int64_t phy_mlp_n_params(const phy_mlp* mlp);
int phy_mlp_get_params(phy_ctx* ctx, const phy_mlp* mlp, float* h_out);      /* synchronous */
int phy_mlp_set_params(phy_ctx* ctx, phy_mlp* mlp, const float* h_in);       /* synchronous */
int phy_mlp_forward(phy_ctx* ctx, phy_mlp* mlp, const float* d_x, int64_t batch, float* d_y);
/* gradients of the loss into d_grads (n_params floats, nullable) and the
 * loss into d_loss (1 double, nullable); parameters unchanged */
int phy_mlp_backward(phy_ctx* ctx, phy_mlp* mlp, const float* d_x, const float* d_t,
                     int64_t batch, float* d_grads, double* d_loss);
/* one Adam step (beta 0.9 / 0.999, eps 1e-8) on one batch */
int phy_mlp_step(phy_ctx* ctx, phy_mlp* mlp, const float* d_x, const float* d_t, int64_t batch,
                 float lr, double* d_loss);

/* ------------------------------------------------------------------------- */
/* Batched 20-state (amino-acid) pruning: many independent trees, each with  */
/* its own GTR-20 model and Gamma rates, in one persistent launch on BF16    */
/* MFMA (Likelihood_c.median_2 + root_cost over whole trees, lib/node.ml:    */
/* 183-198,272, for n_states = 20; P(t) per MlModel.compose,                  */
/* lib/mlModel.ml:585-587).  All per-tree inputs are concatenated in tree    */
/* order:                                                                    */
/*   children   sum_t 2(N_t - 1) ints (phy_topo_create layout per tree)      */
/*   blen       sum_t (2N_t - 1) branch lengths                              */
/*   pi, Q      T x 20, T x 400;  rates, probs  T x n_rates                  */
/*   tips       sum_t N_t x P_t bytes, row-major per tree: state index       */
/*              0..19, >= 20 = missing (all states)                          */
/*   weights    sum_t P_t pattern weights                                    */
/* waves: waves per workgroup (1, 2, 4 or 8); a workgroup owns 16*waves     */
/* patterns of one tree.  The host inputs are copied to the device here.     */
/* ------------------------------------------------------------------------- */
typedef struct phy_aa_batch phy_aa_batch;
int phy_aa_batch_create(phy_ctx* ctx, int32_t n_trees, const int32_t* n_leaves,
                        const int64_t* n_patterns, const int32_t* children,
                        const int32_t* roots, const double* blen, const double* pi,
                        const double* Q, int32_t n_rates, const double* rates,
                        const double* probs, const uint8_t* tips, const uint32_t* weights,
                        int32_t waves, phy_aa_batch** out);
void phy_aa_batch_destroy(phy_aa_batch* batch);
/* pattern_nodes = sum_t P_t (N_t - 1); n_items = work items of the queue. */
int phy_aa_batch_info(const phy_aa_batch* batch, int64_t* pattern_nodes, int64_t* n_items,
                      int32_t* stack_depth);
/* One evaluation: P operands from the branch lengths, the persistent pruning
 * launch, per-tree compensated sums into d_lnl[n_trees] (device). */
int phy_aa_batch_run(phy_ctx* ctx, phy_aa_batch* batch, double* d_lnl);

#ifdef __cplusplus
}
#endif
#endif /* PHYLO_HIP_H */
