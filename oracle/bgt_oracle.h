/*
 * bgt_oracle.h -- CPU restatement of the BayesianGroupTesting lattice hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the parity checker for the HIP
 * provider and the timed CPU baseline. Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product never does.
 *
 * Every function follows the reference CPU provider's serial code path,
 * loop order for loop order, as the reference builds by default
 * (-O3 -march=native -ffast-math, OpenMP off). Where -ffast-math changes
 * an operation the restatement writes the compiled form explicitly
 * (the in-place normalisation x / D is compiled to x * (1/D)).
 *
 * Coordinates are the reference's *internal* ones: atom a = v*n + i
 * (variant-major), bit a of a state set means atom a is "negative" (bit 1),
 * prior pi[a] = P(bit a == 0).
 */
#ifndef BGT_ORACLE_H
#define BGT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_tree_stats {
	int total_leaves;
	double correct_probability;
	double incorrect_probability;
	double false_positive_probability;
	double false_negative_probability;
	double unclassified_probability;
	double expected_stages;
	double expected_tests;
} orc_tree_stats;

/* selectors: 0 auto, 1 op_bha, 2 op_bbpa, 3 brute_force (same numbering as the provider ABI) */

void orc_dilution_table(int n, double alpha, double h, double *out);
double orc_prior_prob(uint32_t state, int atoms, const double *pi);
void orc_prior(int atoms, const double *pi, double *post);
double orc_response_prob(int dilution, int curr_subjs, int variants, uint32_t e, uint32_t r,
						 uint32_t s, const double *dil, int dil_stride);
double orc_update_in_place(int dilution, int curr_subjs, int variants, double *post, uint32_t e,
						   uint32_t r, const double *dil, int dil_stride);
double orc_update_out(int dilution, int curr_subjs, int variants, const double *post, double *out,
					  uint32_t e, uint32_t r, const double *dil, int dil_stride);
void orc_zeta(double *a, int n);
uint32_t orc_op_bha(const double *post, int n, double *upset_out);
/* op_bha with the reference's OpenMP loops (all-cores CPU baseline only) */
uint32_t orc_op_bha_omp(const double *post, int n, double *upset_out);
int orc_threads(void);
uint32_t orc_bbpa(const double *post, int n, int k, double target, double *best_score);
/* orc_bbpa with every experiment's objective written to scores[0 .. 2^n) */
uint32_t orc_bbpa_scores(const double *post, int n, int k, double target, double *scores);
uint32_t orc_brute_force(const double *post, int n, int k, double *best_score);
uint32_t orc_select(const double *post, int n, int k, int selector);
void orc_atom_masses(const double *post, int atoms, double *out);
void orc_classify(const double *masses, int atoms, double up, double lo, uint32_t *pos, uint32_t *neg);
double orc_upset_mass(const double *post, int atoms, uint32_t e);
int orc_select_experiments(const double *post, int n, int steps, uint32_t *out);
/* every candidate's lookahead score after m selected public experiments (lattice_api.cpp:139-158) */
int orc_lookahead_scores(const double *post, int n, const uint32_t *selected, int m, double *scores);
uint32_t orc_reverse_bits(uint32_t value, int width);

/* shrink plan (shrink_plan.cpp:124-177) and projection (shrink.hpp:70-86) */
typedef struct orc_shrink_plan {
	int original_subjects, current_subjects, variants, source_atoms;
	int base_atom_count, reduced_atom_count, new_current_subjects;
	uint32_t current_classified_atoms, removable_current_atoms, new_classified_subjects;
	int original_to_current_atom_position[32];
	uint32_t base_atom_masks[32];
	uint32_t reduced_atom_masks[32];
} orc_shrink_plan;

void orc_make_shrink_plan(int orig, int curr, int k, uint32_t classified_atoms,
						  uint32_t already_classified_subjects, orc_shrink_plan *plan);
void orc_project_posterior(const double *src, double *dst, const orc_shrink_plan *plan);
uint32_t orc_true_ex(uint32_t halving, uint32_t clas_subjs, int orig_subjs);
/* one update_metadata_with_shrinking step (lattice_base.cpp:266-316): classification from the
 * atom masses of post (curr subjects), then the projection when subjects became fully
 * classified; out (2^(curr k) entries) receives the posterior after it; returns curr' */
int orc_shrink_step(int orig, int k, double up, double lo, const double *post, int curr, uint32_t *pos,
					uint32_t *neg, uint32_t *clas_subjs, double *out);

/* run_simulation -> TreeStats (simulation_runner.cpp:92-132) */
int orc_run_simulation(int dilution, int n, int k, const double *prior, int depth, double up,
					   double lo, double branch, int selector, const double *dil, int dil_stride,
					   orc_tree_stats *stats);

/* The no-shrink restatement of the tree, the reference's own independent
 * check (tests/cpp/regression/no_shrink_tree_reference_test.cpp:37-311). */
int orc_run_simulation_noshrink(int dilution, int n, int k, const double *prior, int depth,
								double up, double lo, double branch, double alpha, double h,
								orc_tree_stats *stats);

/* one binary round (Op-BHA select, update, atom marginals) with timings, used for the CPU baseline */
double orc_now(void);

#ifdef __cplusplus
}
#endif

#endif
