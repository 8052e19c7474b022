// Internal declarations shared by the host orchestration and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime_api.h>

#include "bgt_hip_exact.hpp"

#include <cstddef>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>

namespace bgt_hip {

// Largest lattice: 2^30 states (the reference's int state_count limit, state_encoding.hpp:26-33).
constexpr int kMaxAtoms = 30;
constexpr int kMaxVariants = 5;
constexpr int kMaxSubjects = 30;

// Register tile of the transform: 1024 threads x 32 values = 2^15 states per workgroup.
constexpr int kTileBits = 15;
// The transform intermediate z holds 2^(A-15) rows of 2^15 entries; rows are spaced by
// 2^15 entries plus 128 bytes so that the strided 16-32 byte reads of pass 2 spread over
// the memory channels instead of landing on the same one. (A layout holding each pass-2
// tile contiguously turns those reads into runs of 64 entries but pass 1's stores into
// 16-byte pieces; at 2^28 states it made pass 1 slower by more than it saved in pass 2.)
//
// Blocked form (FP32, 28..30 atoms, the split pass 2): z is cut into column blocks of
// 2^kZBlockBits columns (128 bytes) by all 2^(A-15) rows, each block contiguous and blocks
// 2^(A-15+kZBlockBits) + kZBlockSkew entries apart: entry (row r, column c) sits at
// (c >> 5) * pitch_b + (r << 5) + (c & 31). Pass 1 stores whole 128-byte lines (a wave's
// 64 columns are two of them); the 8-column units of pass 2 read 32 bytes of each line of
// one block, the block's four units side by side on one XCD. The skew keeps the 1024
// blocks of a row off one memory channel (without it, pass 1's stores at a power-of-two
// stride ran at a third of their speed). Measured at 2^28 states: pass 2 648 -> 574 us,
// pass 1 +8 us.
constexpr int kZSkewBytes = 128;
constexpr int kZBlockBits = 5;
constexpr uint32_t kZBlockSkew = 32; // entries (128 bytes of FP32)
template <typename T>
constexpr uint32_t z_pitch() { return (1u << kTileBits) + kZSkewBytes / sizeof(T); }
inline uint64_t z_bytes(int atoms, size_t elem)
{
	return atoms <= kTileBits ? 0 : (uint64_t)(1ull << (atoms - kTileBits)) * ((1ull << kTileBits) + kZSkewBytes / elem) * elem;
}
constexpr int kTileThreads = 1024;
constexpr int kTileRegs = 32;
// Lattices with at most this many atoms run the transform in one LDS-resident workgroup.
constexpr int kSmallBits = 14;

enum class Storage : int { fp64 = 0, fp32c = 1 };

enum Objective : int {
	kObjHalving = 0,   // Op-BHA: |M(E) - 1/2|
	kObjBinaryBbpa = 1, // BBPA k=1: |M - t| + |(T - M) - t|
	kObjGather = 2,    // BBPA k>=2: scatter F_B(E) = M(spread_B(E)) into feature tables
	kObjNone = 3,      // transform only (upset masses written out)
};

// Per-lattice scalar slots in device memory (double scal[kScalars]).
enum Scalar : int {
	kScalInv0 = 0, // 1/D, double-buffered with kScalInv1 (pending normalisation of y)
	kScalInv1 = 1,
	kScalDen = 2,   // last denominator D
	kScalTotal = 3, // total mass of the last transform input
	kScalUpset = 4, // last upset-mass query
	kScalars = 8,
};

struct SelResult {
	double d1, d2;
	uint32_t e1, e2;
};

// One selection candidate (internal coordinates). `d` is the objective with a sign: its
// magnitude is the objective and its sign says on which side of the target the candidate's
// mass lies (negative: M(E) below 1/2 for Op-BHA, P(0) below P(1) for binary BBPA).
// Candidates are ordered by |d|, then by experiment. flags: kCandSided when the sign is
// meaningful.
struct Cand {
	double d;
	uint32_t e;
	uint32_t flags;
};
constexpr uint32_t kCandSided = 2u;
// host side only: band index + 1 of a candidate from here up in flags
constexpr int kCandIndexShift = 8;
// Candidates within the error band of the best objective, collected for exact rescoring:
// up to kBandSmall of them for lattices of at most 20 atoms, kBandLarge above.
constexpr int kBandSmall = 4096;
constexpr int kBandLarge = 65536;
inline int band_capacity(int atoms) { return atoms > 20 ? kBandLarge : kBandSmall; }
// per-workgroup flags of the final transform launch (at most this many workgroups)
constexpr int kMaxFinalGroups = 1024;
// The first kBandHead of them travel with the 2 KB result head in one device-to-host copy.
constexpr int kBandHead = 64;
// Bytes at the start of every lattice's arena that one result copy brings back: scalars,
// selection result, counters, atom masses and the first kBandHead band candidates.
constexpr size_t kHeadBytes = 1024 + kBandHead * sizeof(Cand);

// One update of a binary non-dilution lattice for the exact program: the pool and the
// likelihood factors when the pool is / is not contained in the state.
struct ExactRec {
	uint32_t e;
	uint32_t pad;
	double fin, fout;
};

// Updates a lattice of several variants keeps for the exact Op-BBPA re-decision on the device.
constexpr int kGatherMaxHist = 32;
// P(r_v = 1) of a non-dilution test whose pool holds no positive subject (replicated_lattice.cpp:6)
constexpr double kNegativeResponse = 0.99;
// counter slot: the selection was re-decided exactly on the device (k_gather_pick)
constexpr int kGatherFlag = 3;

// Everything a kernel needs to know about one lattice. Arrays of these live in device
// memory so that one launch can serve a batch of lattices (blockIdx.y selects).
struct LatDesc {
	void *post;        // unnormalised posterior y (double or float); x = y * inv
	void *z;           // spilled transform intermediate (double or float)
	double *scal;      // kScalars slots
	double *part;      // per-block partial sums
	SelResult *tops;   // per-block top-2 candidates
	SelResult *result; // final selection
	unsigned *counter; // last-block arrival counter
	double *lik;       // likelihood factors f[v * (n+1) + idx]
	double *masses;    // 32 atom masses (singletons of the transform)
	double *ftab;      // (2^k) x 2^n feature tables F_B(E) for BBPA k >= 2
	int n, k, atoms;
	int len;           // |E| of the experiment being applied
	uint32_t e, r;     // experiment / response being applied
	uint32_t active;   // selectable subjects (experiments must be subsets)
	int dilution;
	int inv_in, inv_out; // scalar slot read / written for 1/D
	Cand *cand;        // per-thread top-2 of the last selection (null: not collected)
	Cand *band;        // up to band_cap candidates within `band_width` of the best
	double band_width; // error band of the objective (absolute)
	unsigned *wgflag;  // workgroups of the final transform whose runner-ups reached the band
	unsigned band_cap;
	// out-of-place update (k_update only): read x = src_post * src_scal[src_inv] instead of
	// this lattice's own posterior (null: in place)
	const void *src_post;
	const double *src_scal;
	int src_inv;
	// exact rescoring on the device (k_exact_band): the update history, the prior's
	// (pi_a, 1 - pi_a) pairs and the output array
	ExactRec *hist;
	int hist_len;
	// the newest update's record travels here when it is the only one not yet on the device
	// (hist_new_at >= 0); k_exact_band uses it for that index and stores it for later rounds
	int hist_new_at;
	ExactRec hist_new;
	const double *pairs;
	double *exact_u;
	// Op-BBPA histogram tables (k >= 2): fixed-point bins, three 64-bit limbs each, zero
	// between uses
	unsigned long long *fixtab;
	// Op-BBPA error band of lattices of several variants (k_bbpa_band; null: not collected)
	Cand *gband;
	// their exact re-decision on the device (k_gather_exact, k_gather_pick): the update
	// history (pools and responses), the dilution table it was applied with (null:
	// non-dilution), and the exact objectives of the band entries; gh_len < 0: no exact
	// program (a longer history, or updates with different tables)
	uint32_t gh_e[kGatherMaxHist];
	uint8_t gh_r[kGatherMaxHist];
	int gh_len;
	// the program's frame: gx_n subjects, of which those in gx_keep, in order, are this
	// lattice's (a projected lattice keeps the program of the lattice it came from; pools,
	// pairs and the table are in the frame, band entries are placed in it)
	int gx_n;
	uint32_t gx_keep;
	const double *gh_pairs; // the prior's (pi_a, 1 - pi_a), indexed 2a + bit (frame atoms)
	const double *gh_dil;   // rows of gx_n + 1 entries
	double *gexact;
};

// Device exact rescoring covers the first kExactBand band candidates, histories of at
// most kExactMaxHist updates and candidates whose remaining pools number at most
// kExactMaxPools; the host program takes the rest.
constexpr int kExactBand = 1024;
// A single lattice with exact rescoring brings back a wider head in the same one copy: the
// 1 KB block, the device's exact masses of the first kExactBand band candidates, and the
// first kBandHeadWide band candidates (its arena places the exact masses at 1 KB and the band
// right after them), so that a band of ordinary size needs no second copy.
constexpr int kBandHeadWide = 1024;
constexpr size_t kHeadExactWide = 1024;
constexpr size_t kHeadBandWide = kHeadExactWide + sizeof(double) * kExactBand;
constexpr size_t kHeadBytesWide = kHeadBandWide + kBandHeadWide * sizeof(Cand);
constexpr int kExactMaxHist = kDenseMaxHist;
constexpr int kExactMaxPools = kDenseMaxPools;
// the exact band kernel stages one update and one prior value per lane of its wavefront
static_assert(kExactMaxHist <= 64 && 2 * kMaxAtoms <= 64, "exact band staging is one wavefront wide");
static_assert(kExactMaxPools <= 10, "the device's dense program holds at most 2^10 pool states in registers");

constexpr int kLikCapacity = kMaxVariants * (kMaxSubjects + 1);

// Kernel argument: one lattice by value (single calls need no host-to-device copy) or a
// device array of `count` lattices (batched launches, blockIdx.y selects the lattice).
// ragged (with `many`): the lattices may differ in shape. The grid is sized for the largest;
// each lattice's kernels take its own workgroup count (the functions below, the same counts
// its per-lattice launches use, so its work is split identically) and their sizes from its
// descriptor, and the workgroups past that count return at once.
struct Descs {
	LatDesc one;
	double one_lik[kLikCapacity];
	const LatDesc *many;
	int ragged;
};

// Workgroup counts per lattice of the launches a ragged batch merges.
// k_update: 256 threads x 16 states per workgroup, at most 2048 workgroups
constexpr unsigned update_groups(int atoms)
{
	const uint64_t g = ((1ull << atoms) + 4095ull) / 4096ull;
	return g < 1ull ? 1u : (g > 2048ull ? 2048u : (unsigned)g);
}
// k_atom_masses: chunks of 2^mass_chunk_bits states
constexpr int mass_chunk_bits(int atoms) { return atoms < 14 ? atoms : 14; }
constexpr unsigned mass_groups(int atoms) { return 1u << (atoms - mass_chunk_bits(atoms)); }
// one workgroup per 2^15-state tile (k_zeta_pass1 and k_zeta_final up to 20 atoms)
constexpr unsigned tile_groups(int atoms) { return atoms <= kTileBits ? 1u : 1u << (atoms - kTileBits); }
// k_bbpa_score: one thread per experiment of n subjects, at most 2^20 workgroups of 256
constexpr unsigned score_groups(int n)
{
	const uint64_t g = ((1ull << n) + 255ull) / 256ull;
	return g > (1ull << 20) ? (1u << 20) : (unsigned)g;
}
// Largest lattice of a ragged transform group (pass 1, the final pass, the band rescan):
// above it the pass-1 grid depends on the compute units and pass 2 on the atom count.
constexpr int kRaggedMaxAtoms = 20;
// What a finalize kernel's `count` (the producing launch's workgroups) is for each lattice
// of a ragged launch.
enum CountRule : int {
	kCountGiven = 0, // the argument (shape-uniform launches)
	kCountUpdate = 1, // update_groups(atoms)
	kCountTiles = 2,  // tile_groups(atoms)
	kCountScore = 3,  // score_groups(n)
	kCountMasses = 4, // mass_groups(atoms)
};

// Host-side error type carrying a provider status code.
struct Error : std::runtime_error {
	int code;
	Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void check_hip(hipError_t e, const char *what);
#define BGT_HIP_CHECK(expr) ::bgt_hip::check_hip((expr), #expr)

// ---- kernel launchers (bgt_hip_kernels.hip). `descs` is a device array of `count` LatDesc,
// all of the same (n, k), or with descs.ragged, of uneven shapes whose largest sets the sizes
// passed here. Every launcher is asynchronous on `stream`. ----
void launch_prior(const Descs &descs, int count, int atoms, Storage st, const double *pi_pairs,
				  uint64_t states, hipStream_t stream);
void launch_update(const Descs &descs, int count, uint64_t states, int k, int dilution,
				   Storage st, hipStream_t stream);
// exact path: transform of x = y * inv (n >= kTileBits + 1): pass 1 over the low 15 bits.
// defer_sum: leave the sum of its per-tile partials to a persistent pass 2 (fold below)
// `objective` is the one of the final pass that reads this z (it decides z's layout)
void launch_zeta_pass1(const Descs &descs, int count, int n, Storage st, bool fused_update,
					   int k, int dilution, int objective, hipStream_t stream, bool defer_sum = false);
// z of an n-atom lattice read by a final pass with this objective is in the blocked form
bool blocked_z(int n, Storage st, int objective);
// pass 2 over bits 15..n-1 (or the whole transform when n == 15), with an objective.
// fold > 0 (persistent pass 2 after a fused pass 1 with defer_sum): every workgroup sums
// the `fold` per-tile partials itself, as the separate finalize launch would, and the first
// stores the denominator, its inverse and the total
// after_transform (optional) runs on the host between the transform kernel(s) and the
// merge of their top-2 lists (a timing hook).
unsigned launch_zeta_final(const Descs &descs, int count, int n, int k, Storage st, int objective,
						   bool from_posterior, bool scaled, double target, hipStream_t stream,
						   unsigned fold = 0, const std::function<void()> &after_transform = {});
// n <= kSmallBits: whole transform in one workgroup per lattice
unsigned launch_zeta_small(const Descs &descs, int count, int n, int k, Storage st, int objective,
					   double target, hipStream_t stream);
// workgroups of launch_zeta_final / launch_zeta_small for a lattice of `atoms` (1024 threads
// each; the candidate buffer needs 2 entries per thread)
unsigned zeta_final_groups(int atoms);
// the final transform of `atoms` atoms is the persistent pass 2 (which collects its error
// band in the pass, each state at most once)
bool persistent_pass2(int atoms, int objective, bool from_posterior);
// the FP32 three-pass form of the high bits (pass M over bits 15..24, pass T over the rest)
// rather than one persistent pass 2, for lattices of 26..30 atoms
bool three_pass_form(int atoms, Storage st);
// -1: the tuned or default choice per atom count; 0: always two passes; 1: three where it
// applies (a process-wide override, for tests and the tuner's own measurements)
void set_pass2_form(int form);
// Candidates of the last zeta_final within d.band_width of the best: d.band[0..counter[0]),
// complete unless counter[0] exceeds d.band_cap (entries may repeat). Pass 2 of 21..30
// atoms appends candidates while it scans and this keeps those inside the final band. For
// 15..20 atoms the per-thread top-2 lists are gathered, and every workgroup with a thread
// whose runner-up reached the band is transformed again to add all of its entries inside
// the band. Lattices of at most kSmallBits atoms collect their band inside
// launch_zeta_small and need no call here. The arguments after `entries` are those of the
// launch_zeta_final call that produced the candidates, `entries` its return value.
void launch_band_collect(const Descs &descs, int count, unsigned entries, int n, int k, Storage st, int objective,
						 bool from_posterior, bool scaled, double target, hipStream_t stream);
// copy the kHeadBytes head of each of `count` lattices (arena base = LatDesc::scal) into
// out[i * kHeadBytes]
void launch_gather_heads(const LatDesc *many, int count, unsigned char *out, hipStream_t stream);
void launch_bbpa_score(const Descs &descs, int count, int n, int k, double target, hipStream_t stream,
					   bool scaled = false);
// after launch_bbpa_score: the candidates within d.band_width of the best into d.gband
// (lattices of several variants that collect a band; the others return at once), then, where
// the top two are within the band (a near tie) and the band is complete, their exact FP64
// objectives from the prior and the update history and the exact best two into d.result
// (counter[kGatherFlag] = 1). The posterior is a product over the variant blocks (the prior
// is a product over atoms and each likelihood factor depends on one block,
// replicated_lattice.cpp:8-28), so a feature mass F_B(E) is the product over v in B of the
// binary mass M_v(E) = U_v(E) / U_v({}), U_v summed over the 2^(n - |E|) states of block v.
// scaled (score and band): tables from a final pass scaled by the pending 1/D (fused update)
void launch_bbpa_band(const Descs &descs, int count, int n, int k, double target, hipStream_t stream,
					  bool scaled = false);
// Op-BBPA feature tables of k >= 2 lattices in one pass over the posterior (bbpa.cpp:237-313)
// where each lane can hold its entries of every table (n >= 6, (2^k - 1) 2^(n-6) <= 64, more
// than 14 atoms): d.ftab, the total and the atom masses, ready for launch_bbpa_score. fused:
// apply the pending update (d.e, the likelihood factors) in the same pass and store the new
// denominator (slot inv_out).
bool bbpa_hist_path(int n, int k);
size_t bbpa_fix_bytes(int n, int k);
void launch_bbpa_hist(const Descs &descs, int count, int n, int k, Storage st, bool fused, hipStream_t stream,
					  bool tables = true);
// After the one-pass table kernel alone (launch_bbpa_hist with tables = false): the tables,
// the scores, the top two, the band and the exact re-decision in one workgroup per lattice
// (the one-pass path has at most 10 subjects, 1,024 experiments); the results of
// launch_bbpa_tables + launch_bbpa_score + launch_bbpa_band, the exact objectives summed per
// band entry in state order (bit-identical between per-lattice and batched calls, which both
// take this path).
void launch_bbpa_finish(const Descs &descs, int count, int k, bool fused, double target, hipStream_t stream);
// Unnormalised exact upset masses U(E) of the first kExactBand band candidates of each
// lattice whose top two are a near tie (d.exact_u[i], -1 where the device program does not
// apply, -2 for dominated entries), from d.pairs and d.hist: the quantity ExactHistory
// computes on the host, as a dense dynamic program. per_lattice workgroups (one wavefront
// each) share a lattice's band entries; ragged descriptors carry their own atom counts.
void launch_exact_band(const Descs &descs, int count, int atoms, hipStream_t stream,
					   unsigned per_lattice = kExactBand);

// Exact unnormalised upset masses by enumeration, for binary lattices of either likelihood
// (replicated_lattice.cpp:8-28): U(E) = sum over the states S >= E of
// P(S) prod_t f_t(m_t(S)), with P(S) the prior product and m_t(S) = |E_t \ S| the number of
// positive atoms in pool E_t. Each term is formed in FP64 from the prior pairs and the
// factor tables, and the terms are summed in a fixed order: per thread, the 2^kEnumLowBits
// states that differ in the query's lowest free atoms in a row, rows in turn; then a
// workgroup tree and a tree over the query's workgroups (enum_groups of its free atoms).
// A query's value therefore does not depend on the other queries of the launch.
constexpr int kEnumLowBits = 6;
constexpr int kEnumMaxGroups = 256;
constexpr int kEnumMaxHist = 64;
constexpr unsigned enum_groups(int free_bits)
{
	const int g = free_bits - kEnumLowBits - 10;
	return 1u << (g < 0 ? 0 : (g > 8 ? 8 : g));
}
struct EnumJob {
	const double *pairs;     // 2 atoms: (pi_a, 1 - pi_a), indexed by the state bit
	const uint32_t *pools;   // updates: the pools
	const double *fac;       // updates x stride: factor by the pool's positive count
	const uint32_t *queries; // count experiments
	double *partial;         // count x kEnumMaxGroups
	double *out;             // count
	int atoms, updates, stride, count;
};
void launch_exact_enum(const EnumJob &job, hipStream_t stream);
// The jobs of several lattices in one launch: jobs_dev is a device copy of `jobs`. Every
// query's value is the one launch_exact_enum gives it.
void launch_exact_enum_many(const EnumJob *jobs_dev, const std::vector<EnumJob> &jobs, hipStream_t stream);
// Relative rounding bound of one value of launch_exact_enum: the products (prior and
// likelihood factors), the in-thread sums and the two trees, all terms non-negative.
inline double enum_rel_bound(int atoms, int updates)
{
	const double rows = (double)(1u << kEnumLowBits); // terms per row
	const double per_thread = 256.0;                  // rows per thread at most (30 free atoms)
	return ((double)atoms + (double)updates + 3.0 + rows + per_thread + 16.0) * 0x1p-53;
}
void launch_brute_force(const Descs &descs, int count, int n, int k, Storage st,
						hipStream_t stream);
void launch_atom_masses(const Descs &descs, int count, int atoms, Storage st, hipStream_t stream);
void launch_upset_mass(const Descs &descs, int count, int atoms, uint32_t experiment,
					   Storage st, hipStream_t stream);
// destination entry b = sum over the subsets `red` of `removed` (increasing order) of
// source[base(b) | red] * (1/D), base(b) = b's bits placed at the atoms outside `removed`
// project_posterior: removals of at most 6 atoms sum each kept entry's 2^r source entries in
// the reference's serial order (bit-identical); larger ones in chunks (k_project_wide), whose
// sums go through `partial` (at least partial_cap doubles, up to 2048 used)
void launch_project(const void *src, const double *src_scal, int inv_slot, void *dst, int dst_atoms, uint32_t removed,
					Storage st, hipStream_t stream, double *partial, size_t partial_cap);
void launch_read_posterior(const Descs &descs, int atoms, Storage st, double *out,
						   hipStream_t stream);
void launch_lookahead(const Descs &descs, int n, Storage st, const uint32_t *selected,
					  int count_selected, double *scores, hipStream_t stream);

// tree statistics: one thread per true state walks the flattened tree
struct FlatNode {
	uint32_t ex, res;
	uint32_t pos, neg;
	int stage;
	int nchild;
	int first_child;
	int subtree_end; // one past the last node of this node's subtree (preorder)
};
void launch_tree_stats(const FlatNode *nodes, int node_count, int n, int k, int dilution,
					   const double *dil, const double *pi, int depth, double branch,
					   double *partials, unsigned *counter, double *out, hipStream_t stream);

// make_shrink_plan (shrink_plan.cpp:124-177) into the C ABI's plan struct
void make_shrink_plan(int original_subjects, int current_subjects, int variants, uint32_t classified_atoms,
					  uint32_t already_classified_subjects, struct bgt_hip_shrink_plan &plan);

// One step of the k-step lookahead on down-set sums: the m experiments selected so far
// (internal coordinates) as the intersections over each subset of them (index k: bit i set
// = experiment i in the subset; the empty subset is every atom).
struct LookStep {
	uint32_t inter[16];
	int m;
	int n;
};
// lattices from this many subjects take the down-set lookahead (below, the O(4^n) scan)
constexpr int kLookDownMin = 13;
// the lookahead selects at most this many experiments (2^16 groups at the last step)
constexpr int kLookMaxSteps = 16;
// from this many earlier picks on, a step takes the class-decomposed scores (below, the
// 2^(m+1) group masses of a candidate are kept in registers)
constexpr int kLookClassMin = 5;
// up to this many subjects a class's down-set sums are formed in LDS (2^n doubles)
constexpr int kLookLdsMax = 12;
// The earlier picks of a lookahead step in internal coordinates, for the class-decomposed
// scores: a state's class has bit i set when the state lies inside pick i.
struct LookClasses {
	uint32_t sel[kLookMaxSteps - 1];
	int m;
	int n;
};
// classes whose down-set sums one chunk holds (rows of 2^n doubles)
size_t look_class_rows(int n);
// every candidate's score (internal index order) after m >= 1 earlier picks through the
// class-decomposed sums (P: look_class_rows(n) rows of 2^n doubles), and each of `blocks`
// blocks' lowest score, as launch_lookahead_down
void launch_look_classes(const Descs &descs, Storage st, const LookClasses &lc, double *P, double *scores,
						 double *best_score, unsigned blocks, hipStream_t stream);
// D(X) = sum over s <= X of x(s), internal coordinates, FP64 (2^atoms entries)
void launch_down_sums(const Descs &descs, int atoms, Storage st, double *D, hipStream_t stream);
// every candidate's score (internal index order) and each of `blocks` blocks' lowest score
void launch_lookahead_down(const double *D, const LookStep &step, double *scores, double *best_score,
						   unsigned blocks, hipStream_t stream);
// each block's lowest public index scoring at most `limit` (same grid as the scoring launch,
// whose per-block lowest scores are block_best)
void launch_lookahead_pick(const double *scores, const double *block_best, int n, double limit,
						   uint32_t *best_exp, unsigned blocks, hipStream_t stream);
// Bound on the FP64 rounding of one lookahead score over 2^(m+1) groups of n-subject states,
// with masses formed from down-set sums (n levels) and a Moebius inversion (m+1 levels):
// each group mass carries at most 2^(m+1)(n+1) u, the score sums 2^(m+1) such terms.
inline double lookahead_score_error(int n, int m)
{
	const double g = (double)(1 << (m + 1));
	return g * (g * (double)(n + 1) + 3.0) * 0x1p-53;
}

} // namespace bgt_hip
