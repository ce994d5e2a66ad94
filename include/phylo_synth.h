/*
 * phylo_synth.h -- seeded synthetic workloads for the two hot paths (host only;
 * none of these functions touches the GPU).  The reference has no workload
 * generator (its bench/ holds only a Makefile, bench/Makefile:1-13); these build
 * the inputs the benchmarks and parity tests run on:
 *
 *   - rooted Yule trees with exponential branch lengths,
 *   - IUPAC DNA state sets (lib/alphabet.ml:301-326 codes) for parsimony,
 *   - GTR+Gamma parameters and sequences simulated down a tree,
 *   - site-pattern compression with integer weights.
 *
 * All randomness comes from a counter-based generator keyed by `seed`, so a
 * result does not depend on the number of host threads used.
 */
#ifndef PHYLO_SYNTH_H
#define PHYLO_SYNTH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Yule tree: start from two leaves and split a uniformly chosen leaf until
 * there are N; leaves get ids 0..N-1, internal nodes N..2N-2 (layout of
 * phy_topo_create).  blen[v] ~ Exponential(mean_bl) for every non-root v;
 * blen[root] = 0.  Returns 0, or 1 for N < 2. */
int phy_synth_yule(int32_t n_leaves, uint64_t seed, double mean_bl,
                   int32_t* children, int32_t* root, double* blen);

/* Parsimony leaves: per (leaf, site) 95% one base uniform over {1,2,4,8},
 * 4% one two-base ambiguity uniform over {3,5,6,9,10,12}, 1% gap as missing
 * (0xF).  Written as nibble blocks: row `leaf` of `out` (stride words). */
int phy_synth_fitch_leaves(int32_t n_leaves, int64_t n_sites, uint64_t seed,
                           uint64_t* out, int64_t stride);

/* Dirichlet(conc^n) draw into x[0..n). */
int phy_synth_dirichlet(int32_t n, double conc, uint64_t seed, double* x);

/* GTR parameters as the reference packs them (lib/mlModel.ml:396-410):
 * exchangeabilities ~ Dirichlet(1^(n(n-1)/2)) divided by the last one, whose
 * first (n(n-1)/2 - 1) entries go to co; pi ~ Dirichlet(pi_conc^n). */
int phy_synth_gtr_params(int32_t n_states, double pi_conc, uint64_t seed,
                         double* pi, double* co);

/* Simulate n_sites sites down the tree: root state ~ pi, per-site rate
 * category uniform over the n_rates rates, child state ~ row of
 * P(rate * blen) = exp(Q rate blen).  If unique != 0, sites are drawn until
 * n_sites distinct leaf patterns have been collected (duplicates dropped;
 * fails with 1 after 64*n_sites draws).  states: N rows x n_sites bytes
 * (state index 0..n_states-1). */
int phy_synth_simulate(int32_t n_leaves, const int32_t* children, int32_t root,
                       const double* blen, int32_t n_states, const double* pi,
                       const double* Q, int32_t n_rates, const double* rates,
                       int64_t n_sites, uint64_t seed, int32_t unique,
                       uint8_t* states);

/* Compress N rows x n_sites states into unique columns (first-occurrence
 * order) with integer weights = multiplicities.  patterns: N rows x n_sites
 * capacity (row pitch n_sites).  Writes *n_patterns. */
int phy_synth_compress(int32_t n_leaves, int64_t n_sites, const uint8_t* states,
                       uint8_t* patterns, uint32_t* weights, int64_t* n_patterns);

/* weights[i] = Poisson(lambda) + 1. */
int phy_synth_poisson_weights(int64_t n, double lambda, uint64_t seed,
                              uint32_t* weights);

/* DNA state indices (0..3, 255 = missing) of N rows x n_sites (row pitch
 * `pitch` bytes) into nibble blocks (1 << state, missing -> 0xF). */
int phy_pack_dna(int32_t n_rows, int64_t n_sites, const uint8_t* states,
                 int64_t pitch, uint64_t* out, int64_t stride);

#ifdef __cplusplus
}
#endif
#endif /* PHYLO_SYNTH_H */
