/*
 * bgt_hip.h -- C ABI of the MI355X (gfx950) HIP provider for BayesianGroupTesting.
 *
 * Plain C types only: no exceptions, no C++ or torch types cross this boundary.
 * Every entry point returns an int status:
 *   BGT_HIP_OK            0  success
 *   BGT_HIP_BAD_INPUT     1  invalid argument (also: missing dilution table)
 *   BGT_HIP_UNSUPPORTED   2  shape/selector not supported, or no gfx950 device
 *   BGT_HIP_NO_BACKEND    3  multi-rank entry points (this provider is single GPU)
 *   BGT_HIP_INTERNAL      4  HIP runtime failure; see bgt_hip_last_error()
 * These are the reference CUDA provider's codes (src/kernels/cuda/cuda_provider.cu:2384-2446).
 *
 * Coordinates at this ABI are the reference's *internal* ones: atom a = v*n + i
 * (variant-major), state bit a == 1 means atom a is negative, prior[a] = P(bit a == 0).
 * The public bit reversal for binary lattices (src/core/lattice_api.cpp:17-26) is the
 * caller's job, exactly as bgt::Lattice does it above model::Lattice.
 *
 * Dilution tables are passed flat: subjects x (subjects + 1) doubles, row s-1 holding
 * pool size s -- the storage of bgt::DilutionTable (include/bgt/lattice.hpp:17-32,
 * src/core/model/dilution.cpp:44-67) without the row-pointer array.
 *
 * Threading: one host thread per lattice handle at a time; distinct handles may be
 * used from distinct threads. Process-wide state (mutex-guarded, never torn down): the
 * per-thread last-error string, a cache of released lattices for re-use by later
 * creations of the same shape and storage (1/8 of device memory by default; the
 * environment variable BGT_HIP_LATTICE_CACHE_BYTES sets the budget, 0 disables it),
 * pools of stream-ordered device memory, pinned host blocks and streams, and the host
 * threads of the exact rescoring.
 */
#ifndef BGT_HIP_H
#define BGT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BGT_HIP_OK 0
#define BGT_HIP_BAD_INPUT 1
#define BGT_HIP_UNSUPPORTED 2
#define BGT_HIP_NO_BACKEND 3
#define BGT_HIP_INTERNAL 4

/* selectors: numbering of bgt::SelectorType (include/bgt/types.hpp:89-94) plus brute_force */
#define BGT_HIP_SELECT_AUTO 0
#define BGT_HIP_SELECT_OP_BHA 1
#define BGT_HIP_SELECT_OP_BBPA 2
#define BGT_HIP_SELECT_BRUTE_FORCE 3

/* storage forms */
#define BGT_HIP_FP64 0  /* plain: FP64 posterior and FP64 transform intermediate */
#define BGT_HIP_FP32C 1 /* FP32 posterior and intermediate; FP64 accumulation on chip */

/* same layout as bgt::TreeStats (include/bgt/types.hpp:109-119) */
typedef struct bgt_hip_tree_stats {
	int total_leaves;
	double correct_probability;
	double incorrect_probability;
	double false_positive_probability;
	double false_negative_probability;
	double unclassified_probability;
	double expected_stages;
	double expected_tests;
} bgt_hip_tree_stats;

/* Result of one selection. objective is |M(E) - 1/2| for Op-BHA and the L1 bucket
 * distance for Op-BBPA / brute force. gap = runner_up_objective - objective. When the
 * device's gap is within its error bound, the pick is rounding-determined (near_tie = 1)
 * and, for binary non-dilution lattices, is re-decided by exact FP64 upset masses
 * recomputed from the prior and the update history over every candidate whose objective
 * lies within the error band (redecided = 1; 2 when the band held more candidates than
 * were collected and only the top two were rescored). Candidates that a one-subject
 * extension or reduction in the band beats for certain are not rescored; when a single
 * candidate remains, it is the exact best and objective keeps the device's value.
 * Otherwise, after a re-decision, objective, runner_up, runner_up_objective and gap_bound
 * are the exact-rescoring values and near_tie reports whether even those are within FP64
 * rounding. */
typedef struct bgt_hip_selection {
	uint32_t experiment;
	uint32_t runner_up;
	double objective;
	double runner_up_objective;
	double gap_bound;
	int near_tie;
	int redecided;
	int candidates; /* experiments rescored exactly (0 without a re-decision) */
	int band;       /* candidates the device collected within the error band */
	int collected;  /* entries the device set aside before the final band filter */
} bgt_hip_selection;

/* ---- provider entry points: counterparts of include/bgt/detail/cuda_provider.hpp:5-24 ---- */

/* replaces bgt_cuda_provider_available (cuda_provider.hpp:5) */
int bgt_hip_provider_available(void);

/* replaces bgt_cuda_provider_run (cuda_provider.hpp:6-9); dilution is flat (see above) */
int bgt_hip_provider_run(int use_dilution, int subjects, int variants, const double *prior,
						 int search_depth, double threshold_up, double threshold_lo,
						 double branch_threshold, int selector, const double *dilution,
						 bgt_hip_tree_stats *stats);

/* run_simulation with the storage form chosen (FP64 plain or FP32C) */
int bgt_hip_provider_run_precision(int use_dilution, int subjects, int variants,
								   const double *prior, int search_depth, double threshold_up,
								   double threshold_lo, double branch_threshold, int selector,
								   const double *dilution, int precision,
								   bgt_hip_tree_stats *stats);

/* replaces bgt_cuda_provider_run_distributed (cuda_provider.hpp:10-13): this provider is
 * single-GPU and returns BGT_HIP_NO_BACKEND */
int bgt_hip_provider_run_distributed(int use_dilution, int subjects, int variants,
									 const double *prior, int search_depth, double threshold_up,
									 double threshold_lo, double branch_threshold, int selector,
									 int enable_collectives, const double *dilution,
									 bgt_hip_tree_stats *stats);

/* replaces bgt_cuda_provider_benchmark_select (cuda_provider.hpp:14-24) */
int bgt_hip_provider_benchmark_select(int subjects, int variants, const double *prior,
									  int selector, int iterations, int warmup,
									  uint32_t *candidate, double *init_seconds,
									  double *mean_seconds, double *min_seconds,
									  double *max_seconds, double *stddev_seconds);

/* ---- lattice session: the device-resident counterpart of bgt::Lattice / model::Lattice ---- */

typedef struct bgt_hip_lattice bgt_hip_lattice;

/* replaces model::create_lattice + Lattice::Lattice (lattice_api.cpp:66-77,
 * lattice_base.cpp:71-79): allocates the posterior on the GPU and writes the prior. */
int bgt_hip_lattice_create(int use_dilution, int subjects, int variants, const double *prior,
						   int precision, bgt_hip_lattice **out);
int bgt_hip_lattice_destroy(bgt_hip_lattice *lattice);

/* replaces Lattice::update / update_probs_in_place (lattice_api.cpp:170-174,
 * lattice_base.cpp:230-244). Asynchronous: returns before the GPU finishes. */
int bgt_hip_lattice_update(bgt_hip_lattice *lattice, uint32_t experiment, uint32_t response,
						   const double *dilution);

/* replaces Lattice::select_experiment (lattice_api.cpp:108-114, bbpa.cpp:861-871);
 * selector BGT_HIP_SELECT_BRUTE_FORCE is the explicit sweep of
 * benchmarks/experiment_selection_benchmark.cpp:214-247 */
int bgt_hip_lattice_select(bgt_hip_lattice *lattice, int selector, bgt_hip_selection *out);

/* replaces Lattice::posterior_probability (lattice_api.cpp:89-93, lattice_base.cpp:139-142) */
int bgt_hip_lattice_posterior(bgt_hip_lattice *lattice, uint32_t state, double *out);

/* bulk read of the normalised posterior (count must equal 2^(subjects*variants)) */
int bgt_hip_lattice_read_posterior(bgt_hip_lattice *lattice, double *out, uint64_t count);

/* replaces Lattice::upset_probability_mass / get_prob_mass (lattice_api.cpp:95-99,
 * lattice_base.cpp:318-351) */
int bgt_hip_lattice_upset_mass(bgt_hip_lattice *lattice, uint32_t experiment, double *out);

/* Exact FP64 upset masses M(E) of `count` atom sets E (internal coordinates: subjects*variants
 * bits), recomputed from the prior and the updates applied so far rather than read from the
 * stored posterior: the values near ties are re-decided with. Dilution lattices and lattices
 * of several variants enumerate the states S >= E on the device (at most 64 pools over the
 * history: updates times variants); binary non-dilution lattices run the host's pool
 * program. rel_bound (may be NULL) receives the relative rounding bound. Returns
 * BGT_HIP_UNSUPPORTED for lattices without an exact program (projected lattices, longer
 * histories) or past the program's budget. The reference has no counterpart: its masses are
 * read from its FP64 posterior (get_prob_mass, lattice_base.cpp:318-351), which these equal
 * within their bounds. */
int bgt_hip_lattice_exact_masses(bgt_hip_lattice *lattice, int count, const uint32_t *experiments, double *out,
								 double *rel_bound);

/* atom marginals m_a = sum_{S : bit a of S} x_S (compute_atom_masses, lattice_base.cpp:25-41);
 * out holds subjects*variants doubles */
int bgt_hip_lattice_atom_masses(bgt_hip_lattice *lattice, double *out);

/* replaces Lattice::update_classification (lattice_api.cpp:176-179, lattice_base.cpp:246-264);
 * the masks are sticky and kept by the handle */
int bgt_hip_lattice_classify(bgt_hip_lattice *lattice, double threshold_up, double threshold_lo,
							 uint32_t *positive_mask, uint32_t *negative_mask);

/* replaces Lattice::positive_classification_mask / negative_classification_mask
 * (lattice_api.cpp:181-191) in internal coordinates; reads the handle's sticky masks */
int bgt_hip_lattice_masks(bgt_hip_lattice *lattice, uint32_t *positive_mask, uint32_t *negative_mask);

/* Restrict the experiments the selectors consider to subsets of `subjects_mask` (internal
 * coordinates; all subjects by default). The tree builder sets it to the unclassified
 * subjects, the effect the reference's shrinking tree gets by removing classified subjects
 * from the lattice (update_metadata_with_shrinking, lattice_base.cpp:266-305). */
int bgt_hip_lattice_set_candidates(bgt_hip_lattice *lattice, uint32_t subjects_mask);

/* ---- shrinking: removing classified subjects from a lattice ---- */

/* Plan of a shrink (ShrinkPlan, shrink.hpp:27-46). Atoms are variant-major: atom v*n + i
 * is variant v of subject i, in the original numbering (original_subjects) or the current
 * one (current_subjects, the subjects still in the lattice). The reference's base and
 * reduced atom mask lists are the atoms of the current layout outside and inside
 * removable_current_atoms, in increasing order. */
typedef struct bgt_hip_shrink_plan {
	int original_subjects;
	int current_subjects;
	int variants;
	int source_atoms;          /* current_subjects * variants */
	int base_atom_count;       /* atoms kept */
	int reduced_atom_count;    /* atoms removed */
	int new_current_subjects;  /* subjects left after the shrink */
	uint32_t current_classified_atoms;  /* classified atoms, current numbering */
	uint32_t removable_current_atoms;   /* every atom of each fully classified subject */
	uint32_t new_classified_subjects;   /* fully classified subjects, original numbering */
	int original_to_current_atom_position[32]; /* -1 for atoms of subjects already removed */
} bgt_hip_shrink_plan;

/* replaces make_shrink_plan (shrink_plan.cpp:124-177): classified_atoms in the original
 * numbering, already_classified_subjects the subjects an earlier shrink removed. Host only. */
int bgt_hip_shrink_plan_make(int original_subjects, int current_subjects, int variants,
							 uint32_t classified_atoms, uint32_t already_classified_subjects,
							 bgt_hip_shrink_plan *out);

/* replaces project_posterior (shrink.hpp:70-86) as apply_shrink_plan uses it
 * (lattice_base.cpp:307-316): a new lattice of plan->new_current_subjects subjects whose
 * posterior is the source's summed over the removed atoms, each sum in FP64 in increasing
 * order of the removed atoms' index, as the reference's. The source lattice is unchanged.
 * Classification masks carry over on the kept atoms; the new lattice has no exact
 * re-decision. Fails (status 2) when the plan would remove every subject. */
int bgt_hip_lattice_project(bgt_hip_lattice *source, const bgt_hip_shrink_plan *plan,
							bgt_hip_lattice **out);

/* k-step Op-BHA lookahead (lattice_api.cpp:116-168) in the reference's PUBLIC coordinates
 * (binary lattices only; steps <= 16, status 3 beyond: 2^16 groups at the last step). The
 * first pick is the Op-BHA selection; each later one, after m earlier picks, scores every
 * candidate by the L1 distance of its 2^(m+1) group masses from uniform. Up to m = 4 the
 * masses of a candidate come from down-set sums built once (two 2^n FP64 work arrays for
 * the call; below 13 subjects a pass over all states per candidate); from m = 5 on, from
 * the down-set sums of the posterior masked to each of the 2^m classes of the earlier picks
 * (2^m transforms per step, up to 64 classes of 2^n FP64 entries, at most 256 MiB, held at
 * once). Departure from the reference's tie rule: candidates whose scores lie within twice
 * the score's FP64 rounding bound, 2^(m+1) (2^(m+1) (n+1) + 3) 2^-53, of the lowest go to
 * the lowest public index (the reference decides such ties, frequent and often exact, by
 * its summation rounding). */
int bgt_hip_lattice_select_experiments(bgt_hip_lattice *lattice, int steps, uint32_t *out);

/* One fused round of the adaptive loop: update(experiment, response), atom marginals of the
 * updated posterior and classification, and the next selection -- Op-BHA for binary lattices
 * (two HBM passes), Op-BBPA for multinomial ones (one pass where the feature tables fit on
 * chip: 6..11 subjects, (2^k - 1) 2^n <= 4095; otherwise update, classify and select in
 * turn). Masses of the updated posterior go to masses_out (may be NULL). */
int bgt_hip_lattice_round(bgt_hip_lattice *lattice, uint32_t experiment, uint32_t response,
						  const double *dilution, double threshold_up, double threshold_lo,
						  bgt_hip_selection *next, uint32_t *positive_mask,
						  uint32_t *negative_mask, double *masses_out);

/* Stated error bounds of the storage form (relative, per posterior entry and per mass),
 * accumulated over the updates applied so far. Zero-width for nothing; FP64 reports its
 * own rounding bound. */
int bgt_hip_lattice_error_bound(bgt_hip_lattice *lattice, double *posterior_rel,
								double *mass_rel);

/* Average device time (ms) per launch of each kernel class since the last reset, measured
 * with HIP events on the lattice's stream. names is a ';'-separated list. enable >= 0 resets
 * and switches the clock: 0 off, 1 every kernel class and host phase, 2 light (events only
 * around one transform pass per fused round, pass 1 and pass 2 in turn, the rest unbracketed:
 * each event record costs the stream a few microseconds); enable < 0 reads. */
int bgt_hip_lattice_kernel_times(bgt_hip_lattice *lattice, int enable, double *ms_out,
								 int *launches_out, int capacity, char *names, int names_len);

int bgt_hip_lattice_info(bgt_hip_lattice *lattice, int *subjects, int *variants, int *precision,
						 uint64_t *device_bytes);

/* Device bytes a lattice of this shape and precision holds once it has selected, without
 * creating it (admission control by HBM footprint). No reference counterpart. */
int bgt_hip_lattice_footprint(int use_dilution, int subjects, int variants, int precision, uint64_t *bytes);

/* ---- batches and serving ---- */

/* response value that means "no update for this lattice in this step" (select only) */
#define BGT_HIP_NO_UPDATE 0xFFFFFFFFu

/* Apply one (experiment, response) update to each lattice (none where responses[i] is
 * BGT_HIP_NO_UPDATE), then select with `selector` on each, one descriptor upload and one
 * result download for the whole batch. Lattices of uneven (n, k) share launches: one per
 * kernel for all updates and all marginals of a storage form, and one per kernel per
 * selection regime (transform in LDS up to 14 atoms, one tile at 15, tiles at 16..20) and
 * objective; only the one-pass BBPA tables, brute force and lattices above 20 atoms launch
 * per shape. Each lattice's work is split into the same workgroups as in
 * bgt_hip_lattice_update + bgt_hip_lattice_select, so its results are identical to those
 * per-lattice calls. A lattice may appear at most once per batch. No reference
 * counterpart (the reference steps one lattice at a time). */
int bgt_hip_batch_step(bgt_hip_lattice **lattices, const uint32_t *experiments,
					   const uint32_t *responses, const double *const *dilutions, int count,
					   int selector, bgt_hip_selection *out);

/* As bgt_hip_batch_step, with atom marginals and sticky classification between the update
 * and the selection (the adaptive loop's round for every lattice: the same results as
 * update + classify + select per lattice). positive/negative receive each lattice's masks
 * (may be NULL). */
int bgt_hip_batch_round(bgt_hip_lattice **lattices, const uint32_t *experiments,
						const uint32_t *responses, const double *const *dilutions, int count,
						int selector, double threshold_up, double threshold_lo,
						bgt_hip_selection *out, uint32_t *positive, uint32_t *negative);

/* Graph replay of batched steps (bgt_hip_batch_step, bgt_hip_batch_round), off by default.
 * With enable != 0 a step's copies and launches are captured into a hipGraph once per
 * launch structure (the launch groups and their sizes, the selector and the staging
 * buffers) and replayed by later steps of the same structure; the step's own descriptors,
 * experiments and likelihood tables reach the device through the graph's first copy, so the
 * results equal the launched step's. Up to 16 graphs per calling thread are kept (least
 * recently used out). previous (may be NULL) receives the former setting. No reference
 * counterpart (the reference launches every kernel per call). */
int bgt_hip_batch_graphs(int enable, int *previous);

/* Batched steps that were captured into a new graph, and steps that replayed one, since the
 * process started (either may be NULL). */
int bgt_hip_batch_graph_counts(unsigned long long *captures, unsigned long long *replays);

/* Shrinking in bgt_hip_provider_run_simulation's tree, off by default: on, a child whose
 * classification covers every variant of some subject is replaced by its projection onto the
 * other subjects and selects in that numbering (global_tree_intra.cpp:8-40 with
 * update_metadata_with_shrinking, lattice_base.cpp:266-316); off keeps full-size children
 * with experiments restricted to the subjects still in play (equal in exact arithmetic, as
 * the reference's no_shrink_tree_reference_test states, and faster where the tree's levels
 * are bound by launches). previous (may be NULL) receives the former setting. No reference
 * counterpart (the reference always shrinks). */
int bgt_hip_tree_shrink(int enable, int *previous);

/* ---- autotuning ---- */

/* Measure once per KernelSpec key (include/bgt/types.hpp:132-181) which launch forms of the
 * transform are fastest on this device for lattices of this shape, and use them from then on.
 * For FP32C lattices of 26..30 atoms: the form of the high bits (0: two passes, from 28 atoms
 * the split pass 2 on the blocked intermediate; 1: three passes; 2: two passes with the plain
 * pass 2 on the row-form intermediate, 28..30 atoms) and the grid of pass 1 (0: persistent
 * with a prefetch of the next tile; 1: one workgroup per tile); every other shape has one
 * form. Each pair is timed over fused rounds (update, classification and the next selection,
 * the work of one test); form 0 is kept unless another is faster by more than 3 %, so that
 * forms within the measurement's spread are not chosen by noise. Results are cached in memory
 * and in a text file per key named as the reference's JIT artefacts (kernel_spec_id,
 * src/runtime/jit_compiler.cpp:13-38) in BGT_HIP_TUNE_DIR (default: the temporary
 * directory's bayesian_group_testing_hip_tune); force = 1 measures again. form receives the
 * high-bit form; ms_two_pass / ms_three_pass the measured round time of forms 0 and 1, each
 * with its faster pass-1 grid (0 when taken from a cache or when the shape has one form
 * only); spec_id (may be NULL) the key. Run it before sessions of that
 * shape start: the choice applies process-wide. The reference's KernelRegistry /
 * JitCompiler have no device measurement. */
int bgt_hip_tune(int use_dilution, int subjects, int variants, int precision, int selector, int force, int *form,
				 double *ms_two_pass, double *ms_three_pass, char *spec_id, int spec_len);

/* As bgt_hip_tune, with both choices and every measurement: ms (may be NULL) receives 6
 * round times, ms[pass2_form * 2 + pass1_form]. No reference counterpart. */
int bgt_hip_tune_forms(int use_dilution, int subjects, int variants, int precision, int selector, int force,
					   int *pass2_form, int *pass1_form, double *ms, char *spec_id, int spec_len);

/* Pin the high-bit form (0, 1, 2 as above) or the pass-1 grid (0, 1) process-wide; -1
 * returns to the tuned choice, or the built-in default (form 0, persistent pass 1). */
int bgt_hip_tune_force(int form);
int bgt_hip_tune_force_pass1(int form);

/* ---- selection without device allocation ---- */

/* Largest lattice and static scratch size of bgt_hip_select_noalloc. */
#define BGT_HIP_NOALLOC_MAX_SUBJECTS 20
#define BGT_HIP_NOALLOC_SCRATCH_BYTES 8454144 /* 8.06 MiB per GPU, held until the process ends */

/* No counterpart in the reference. For callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops and harnesses that forbid hipMalloc): the selection
 * of bgt_hip_lattice_select (same kernels, same launch shapes) on a caller-owned posterior of
 * a binary lattice, with every other buffer in a static device array of
 * BGT_HIP_NOALLOC_SCRATCH_BYTES that each GPU holds until the process ends. It calls no
 * hipMalloc, hipFree or allocating library routine.
 *   posterior    device memory, 2^subjects entries in internal coordinates, double
 *                (precision BGT_HIP_FP64) or float (BGT_HIP_FP32C); entry S is read as
 *                posterior[S] * scale (scale > 0: a pending normalisation, 1.0 for none)
 *   subjects     1..BGT_HIP_NOALLOC_MAX_SUBJECTS (status 2 beyond)
 *   selector     BGT_HIP_SELECT_AUTO / _OP_BHA (Op-BHA) or BGT_HIP_SELECT_OP_BBPA (binary)
 *   result       device memory, one bgt_hip_selection: experiment, runner-up and their
 *                objectives as bgt_hip_lattice_select reports them before any exact
 *                re-decision; gap_bound is the transform's own FP64 rounding on the top two
 *                (errors the posterior already carries are the caller's to add), near_tie
 *                whether the gap lies within it; redecided, candidates, band, collected = 0
 *   masses       device memory for the atom marginals (subjects doubles) or NULL
 *   stream       the hipStream_t to run on (NULL: the default stream); the call returns at
 *                once, and its results are ready when the stream reaches them
 * Every call on one GPU shares that GPU's scratch array: calls that overlap on one GPU (on
 * different streams, or from different threads) race and must be serialised by the caller. */
int bgt_hip_select_noalloc(const void *posterior, int subjects, int precision, int selector, double scale,
						   bgt_hip_selection *result, double *masses, void *stream);

/* ---- exact rescoring (host only, no device needed) ---- */

/* Exact FP64 upset masses M(E) of a binary non-dilution lattice after the given updates,
 * for `count` query experiments, recomputed from the prior and the update history (internal
 * coordinates): the values the selections re-decide near ties with. Returns
 * BGT_HIP_UNSUPPORTED when the history is too entangled for the exact program's budget. */
int bgt_hip_exact_upset_mass(int subjects, const double *prior, int updates, const uint32_t *experiments,
							 const uint32_t *responses, int count, const uint32_t *queries, double *out);

/* The same for a binary dilution lattice: dilution is the flat subjects x (subjects + 1)
 * table the updates were applied with (replicated_lattice.cpp:17-28, dilution.cpp:44-67). */
int bgt_hip_exact_upset_mass_dilution(int subjects, const double *prior, const double *dilution, int updates,
									  const uint32_t *experiments, const uint32_t *responses, int count,
									  const uint32_t *queries, double *out);

/* ---- misc ---- */
const char *bgt_hip_last_error(void);
int bgt_hip_device_synchronize(void);
int bgt_hip_device_name(char *out, int len);

#ifdef __cplusplus
}
#endif

#endif
