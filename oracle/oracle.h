/*
 * oracle.h -- CPU restatement of the reference's algorithms for the two hot
 * paths.  TEST INFRASTRUCTURE ONLY: used by tests/, by __graft_entry__.smoke()
 * and by bench.py's cpu_baseline leg as the checker; never by the product.
 *
 * Layout: plain per-site arrays (one byte / uint32 bitmask per site), not the
 * packed device layout, so the oracle shares no data-layout code with the
 * library it checks.
 *
 * Pinning (what checks each function; see tests/test_oracle_*.py):
 *   orc_fitch_median2   reference KAT over all Alphabet.dna code pairs
 *                       (test/costMatrixTest.ml:83-108) + cost = distance
 *                       invariant (test/bitvectorTest.ml:59-69)
 *   orc_fitch_downpass  brute-force unit-cost Sankoff minimum over all
 *                       internal assignments (N <= 7)
 *   orc_fitch_uppass    brute-force most-parsimonious-reconstruction sets
 *                       (N <= 7, unambiguous and ambiguous leaves)
 *   orc_fitch_median3   brute-force 3-way minimum per site
 *   orc_gtr_q           closed forms: JC69 / F81 special cases of GTR
 *   orc_compose         closed forms (JC, F81), Chapman-Kolmogorov, rows sum
 *                       to 1, detailed balance, scipy expm goldens
 *   orc_lnl             brute-force sum over all K^(N-1) internal states
 *                       (N <= 5), 2-taxon closed form
 *   orc_lnl derivatives central finite differences of orc_lnl
 */
#ifndef PHY_ORACLE_H
#define PHY_ORACLE_H
#include <stdint.h>

/* ---------------- Fitch (NonAdditive_c) ---------------- */
/* lib/nonAdditive_c.ml:19-35: out_i = x_i & y_i, or x_i | y_i when empty;
 * returns sum_i w_i [x_i & y_i = 0] (w NULL = unit weights). */
double orc_fitch_median2(const uint8_t* x, const uint8_t* y, uint8_t* out, int64_t n,
                         const double* w);
/* lib/bitvector/bv.c:46-55: number of sites with empty intersection. */
int64_t orc_fitch_distance(const uint8_t* a, const uint8_t* b, int64_t n);
/* 3-way median: states in the most of {a,b,c}; returns sum w (3 - count). */
double orc_fitch_median3(const uint8_t* a, const uint8_t* b, const uint8_t* c, uint8_t* out,
                         int64_t n, const double* w);
/* Fitch (1971) final set of an internal node (see phylo_hip.h adjust_3). */
void orc_fitch_adjust3(const uint8_t* pv, const uint8_t* pa, const uint8_t* pb,
                       const uint8_t* fp, uint8_t* out, int64_t n);
/* Whole-tree down-pass over sets[(2N-1) x S] (rows 0..N-1 input leaves, rows
 * N.. written), node by node in post-order (lib/tree.ml:171-187).
 * site_cost[S], node_cost[N-1] (nullable); returns sum_s w_s site_cost_s. */
uint64_t orc_fitch_downpass(int32_t N, const int32_t* children, int32_t root, uint8_t* sets,
                            int64_t S, const uint32_t* w, uint32_t* site_cost,
                            uint64_t* node_cost, int32_t n_threads);
/* Whole-tree up-pass (final sets of every node). */
void orc_fitch_uppass(int32_t N, const int32_t* children, int32_t root, const uint8_t* sets,
                      uint8_t* fin, int64_t S, int32_t n_threads);
/* Brute force for one site on a tiny tree: minimum unit cost over all
 * internal assignments (leaf edges cost 0 when the parent's state is in the
 * leaf set) and the union of each internal node's states over all optimal
 * assignments (mpr[2N-1], leaves untouched).  K <= 8 states, N <= 8. */
int64_t orc_fitch_brute(int32_t N, const int32_t* children, int32_t root, const uint8_t* leaves,
                        int32_t K, uint8_t* mpr);

/* ---------------- model (MlModel) ---------------- */
/* lib/mlModel.ml:390-421 (m_gtr_independent) + :204-213 (m_meanrate). */
void orc_gtr_q(int32_t n, const double* pi, const double* co, double* Q);
/* exp(Q t) by scaling and squaring of a Taylor series (Moler & Van Loan
 * 2003, method 3), with the compose rules of lib/mlmodel.c:325-342:
 * t == -1 -> Q, t < 1e-10 -> I. */
void orc_compose(int32_t n, const double* Q, double t, double* P);

/* ---------------- Felsenstein pruning (Likelihood_c) ---------------- */
/* Textbook Felsenstein (1981) pruning (cited at lib/mlModel.mli:30-31),
 * FP64 with exact power-of-two rescaling and a Neumaier-compensated
 * weighted sum.  tips[leaf * P + p] = state bitmask (bit i = state i
 * possible).  site_lnl (nullable) [P].  d1/d2 (nullable) [2N-1]: derivatives
 * of lnL in the branch above each non-root node (root branches: the shared
 * unrooted edge's value in both), d1 = sum_p w_p g_p, d2 = sum_p w_p h_p
 * with g = L'/L, h = L''/L - g^2.  d1abs/d2abs (nullable, with d1/d2) [2N-1]:
 * sum_p w_p |g_p| and sum_p w_p (|L''/L| + g^2), the magnitudes of the
 * per-pattern terms (for h, of the two parts of the difference): the scale a
 * rounding error of those terms is measured against.  Returns lnL. */
double orc_lnl(int32_t N, const int32_t* children, int32_t root, const double* blen, int32_t K,
               const double* pi, const double* Q, int32_t R, const double* rates,
               const double* probs, const uint32_t* tips, int64_t P, const uint32_t* weights,
               double* site_lnl, double* d1, double* d2, double* d1abs, double* d2abs,
               int32_t n_threads);
/* Brute force: sum over all K^(N-1) internal assignments per pattern. */
double orc_lnl_brute(int32_t N, const int32_t* children, int32_t root, const double* blen,
                     int32_t K, const double* pi, const double* Q, int32_t R,
                     const double* rates, const double* probs, const uint32_t* tips, int64_t P,
                     const uint32_t* weights);
#endif
