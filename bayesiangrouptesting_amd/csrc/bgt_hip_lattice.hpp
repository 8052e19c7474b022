// Device-resident lattice: the HIP counterpart of model::Lattice for one replicated lattice
// (src/core/model/lattice_base.cpp, replicated_lattice.cpp, src/kernels/cpu/bbpa.cpp).
#pragma once

#include "bgt_hip.h"
#include "bgt_hip_exact.hpp"
#include "bgt_hip_internal.hpp"

#include <memory>
#include <string>
#include <vector>

namespace bgt_hip {

// Per-kernel device time from HIP events recorded on the lattice's stream around each
// launch. Recording never blocks the host; resolve() waits once and folds the pairs in.
struct KernelClock {
	bool enabled = false;
	// light: only the launches marked essential (pass 1) are bracketed by events; each event
	// record costs the stream a few microseconds, which a benchmark's timed rounds then carry
	bool light = false;
	struct Entry {
		std::string name;
		double total_ms = 0.0;
		int launches = 0;
	};
	std::vector<Entry> entries;
	void begin(hipStream_t s, bool essential = false);
	void end(hipStream_t s, const char *name);
	// host-side entries: `ms` of host time, or a count carried in total_ms
	void host(const char *name, double ms);
	void resolve();
	void reset();
	~KernelClock();

private:
	struct Pending {
		std::string name;
		hipEvent_t a, b;
	};
	std::vector<Pending> pending_;
	std::vector<hipEvent_t> pool_;
	hipEvent_t open_ = nullptr;
	hipEvent_t take();
};

class DeviceLattice {
public:
	DeviceLattice(bool dilution, int subjects, int variants, const double *prior, Storage storage);
	~DeviceLattice();
	DeviceLattice(const DeviceLattice &) = delete;
	DeviceLattice &operator=(const DeviceLattice &) = delete;

	// Copy of `parent`'s posterior (same shape); the child's classification masks start
	// from the parent's. Used by the tree builder for out-of-place updates.
	DeviceLattice(const DeviceLattice &parent, bool copy_posterior);

	// Lattices through the process-wide cache of released lattices (see LatticeCache):
	// make() gives a lattice holding `prior` (a cached one of this shape re-initialised, or
	// a new one); make_like() a lattice of `like`'s shape and storage whose posterior and
	// state are to be set by the caller (adopt + an out-of-place update); recycle() hands a
	// lattice back to the cache, which keeps it while within its byte budget.
	static std::unique_ptr<DeviceLattice> make(bool dilution, int subjects, int variants, const double *prior,
											   Storage storage);
	static std::unique_ptr<DeviceLattice> make_like(const DeviceLattice &like);
	static void recycle(std::unique_ptr<DeviceLattice> lattice);
	// A lattice without the atoms in `removed` (every variant of some subjects): its
	// posterior is the projection of `source`'s (project_posterior, shrink.hpp:70-86), its
	// classification masks are `source`'s on the kept atoms, every kept subject is a
	// candidate again. It has no exact re-decision (the update history no longer applies)
	// and its error bound counts the projection as one more rounding.
	static std::unique_ptr<DeviceLattice> project(const DeviceLattice &source, uint32_t removed);

	int subjects() const { return n_; }
	int variants() const { return k_; }
	int atoms() const { return atoms_; }
	bool dilution() const { return dilution_; }
	Storage storage() const { return storage_; }
	uint64_t states() const { return states_; }
	uint64_t device_bytes() const { return bytes_; }
	hipStream_t stream() const { return stream_; }

	void update(uint32_t experiment, uint32_t response, const double *dilution_flat);
	bgt_hip_selection select(int selector);
	double posterior(uint32_t state);
	void read_posterior(double *out);
	double upset_mass(uint32_t experiment);
	void atom_masses(double *out);
	void classify(double threshold_up, double threshold_lo);
	std::vector<uint32_t> select_experiments(int steps);
	bgt_hip_selection round(uint32_t experiment, uint32_t response, const double *dilution_flat,
							double threshold_up, double threshold_lo, double *masses_out);

	uint32_t positive_mask() const { return pos_; }
	uint32_t negative_mask() const { return neg_; }
	void set_masks(uint32_t pos, uint32_t neg) { pos_ = pos; neg_ = neg; }
	void set_active(uint32_t active) { active_ = active; }
	uint32_t active() const { return active_; }
	int updates_applied() const { return updates_; }
	void error_bound(double *posterior_rel, double *mass_rel) const;

	KernelClock &clock() { return clock_; }

	// descriptor of the current state (for batched launches)
	LatDesc desc() const;
	const double *lik_host() const { return lik_; }
	void prepare_update(uint32_t experiment, uint32_t response, const double *dilution_flat);
	void commit_update();
	void sync();
	SelResult fetch_result();
	void fetch_head();
	bgt_hip_selection finish_selection(int selector, const SelResult &r, bool scaled);
	void fetch_masses(double *out);

	// parameters for the selection launchers
	static int resolve_selector(int selector, int variants);
	// device bytes a lattice of this shape holds once it has selected (posterior,
	// transform intermediate and arena)
	static uint64_t footprint(bool dilution, int subjects, int variants, Storage storage);

	// ---- hooks for batched steps (bgt_hip_batch.cpp) ----
	void check_selector(int selector) const;
	void ensure_z();
	// wait for work this lattice's own stream may still have in flight
	void settle();
	// error band of the selection objective (device objectives closer than this may be
	// misordered)
	double band_width(int selector) const;
	// take a result head gathered by a batched step (kHeadBytes of the arena's start);
	// exact_staged: the step ran k_exact_band for this lattice (stage_exact)
	void load_head(const unsigned char *head, bool exact_staged = false);
	// Binary non-dilution lattices with a dense exact program: the descriptor fields of
	// k_exact_band (history, prior pairs, output), the history's new records uploaded (a
	// single new record travels in the descriptor; more are copied on this lattice's stream,
	// then synchronised when sync_upload, for kernels on another stream). False when the
	// device program does not apply.
	bool stage_exact(LatDesc &d, bool sync_upload);
	// sticky classification from the masses of the last head (lattice_base.cpp:246-264)
	void classify_from_head(double threshold_up, double threshold_lo, double *masses_out);
	// selection result of the last head
	SelResult head_result() const;
	// host-side preparation of exact rescoring, to overlap with device work
	void prepare_exact() { exact_.prepare(); }
	// After load_head: whether finish_selection will re-decide this selection exactly (a near
	// tie on a lattice with an exact program); if so, its normaliser is computed now, so that
	// the normalisers of many lattices can run side by side on the host threads.
	bool prepare_rescore(int selector);
	// After load_head, binary dilution lattices: the sets (the exact program's frame) whose
	// masses finish_selection's re-decision will first ask the device enumeration for; empty
	// when it will ask for none. The band it worked out is kept for that re-decision.
	std::vector<uint32_t> enum_ahead_sets(int selector);
	// The enumerations of several lattices' sets (enum_ahead_sets) in one launch on `stream`
	// and one synchronisation; each lattice's next re-decision takes its masses from there.
	static void enum_ahead_run(const std::vector<DeviceLattice *> &lats,
							   const std::vector<std::vector<uint32_t>> &sets, hipStream_t stream);
	// Exact FP64 M(E) of `count` experiments recomputed from the prior and the update
	// history (internal coordinates): by enumeration on the device for binary dilution
	// lattices (at most kEnumMaxHist updates), by the host program for binary non-dilution
	// ones. ok[i] is false where the host program's budget was exceeded; returns false when
	// the lattice has no exact program.
	bool exact_masses(const uint32_t *experiments, size_t count, double *out, bool *ok);
	// relative rounding bound of exact_masses
	double exact_rel_bound() const;
	bool uses_exact_enum() const { return uses_enum(); }
	// lattices of several variants: the device's exact program inputs uploaded (batched steps
	// call it while preparing, before their launches)
	void prepare_gather();
	// Take over `parent`'s host-side state (masks, candidate restriction, update history,
	// pending normalisation slot) ahead of an out-of-place update from it; the posterior
	// itself is produced by that update. Both must have the same shape and storage.
	void adopt(const DeviceLattice &parent);
	// source fields of an out-of-place update reading this lattice's posterior
	void as_source(LatDesc &d) const;

private:
	struct Blank {};
	// a lattice of this shape whose posterior and state the caller sets
	DeviceLattice(bool dilution, int subjects, int variants, Storage storage, Blank);
	Descs single() const;
	void allocate();
	// write the prior and reset every piece of state (construction, and re-use from the cache)
	void init_prior(const double *prior);
	void fill_lik(uint32_t experiment, uint32_t response, const double *dilution_flat);
	void rescore(bgt_hip_selection &out, int selector, const SelResult &r);
	void rescore_band(bgt_hip_selection &out, int selector, const SelResult &r);
	// the band of the head now loaded, one entry per experiment, less the dominated entries
	std::vector<Cand> live_band(int selector, const SelResult &r, bool &complete, unsigned &count);
	// masses of the head's first re-decision chunk, enumerated ahead (enum_ahead_run)
	struct Ahead {
		bool valid = false;         // formed for the head now loaded
		std::vector<Cand> live;     // live_band's result
		bool complete = false;      // and its other outputs
		unsigned count = 0;
		std::vector<uint32_t> sets; // frame sets, increasing
		std::vector<double> masses; // their M(E)
	};
	Ahead ahead_;
	bool ahead_masses(const uint32_t *frame_sets, size_t count, double *out) const;
	// lattices of several variants: the Op-BBPA band after k_bbpa_score (k_bbpa_band), and its
	// exact re-decision from exact feature masses by the device enumeration
	void collect_gather_band(double target, const Descs *fused = nullptr);
	double gather_objective_bound() const;
	void copy_gather_history(const DeviceLattice &parent);
	// exact masses of the band on the device (binary non-dilution), after the band kernels
	void launch_exact(const Descs &ds);
	// exact masses come from the device enumeration (binary dilution, short enough history)
	bool uses_enum() const
	{
		return (dilution_ || k_ >= 2) && exact_.usable() && exact_.enum_pools() <= kEnumMaxHist;
	}
	// M(E) = U(E) / U({}) of atom sets by launch_exact_enum, per variant block (the posterior
	// is a product over the blocks), U({}) in the same launches; synchronises
	void enum_masses(const uint32_t *experiments, size_t count, double *out);
	// masses[v][i] = M_v(queries[v][i]), the binary mass of a set of subjects in block v
	void enum_run(const std::vector<std::vector<uint32_t>> &queries, std::vector<std::vector<double>> &masses);
	// the fields of the head now in the host mailbox, with up to band_entries band candidates
	void unpack_head(int band_entries, size_t band_offset);
	// complete band of candidates for exact rescoring (binary non-dilution lattices); the
	// arguments after `entries` are those of the final transform launch
	void collect_band(const Descs &ds, unsigned entries, int objective, bool from_posterior, bool scaled,
					  double target);

	bool dilution_;
	int n_, k_, atoms_;
	uint64_t states_;
	Storage storage_;
	uint64_t rounds_ = 0; // fused rounds run (the light clock alternates its bracket)
	hipStream_t stream_ = nullptr;
	uint64_t bytes_ = 0;

	void *post_ = nullptr;
	void *z_ = nullptr;
	unsigned long long *fixtab_ = nullptr; // Op-BBPA histogram bins (k >= 2, tables on chip)
	const double *pending_dil_ = nullptr;  // dilution table of the update being applied
	size_t fix_bytes_ = 0;
	char *arena_ = nullptr; // scalars, results, partials, tables
	double *scal_ = nullptr;
	double *part_ = nullptr;
	SelResult *tops_ = nullptr;
	SelResult *result_ = nullptr;
	unsigned *counter_ = nullptr;
	double *masses_ = nullptr;
	double *ftab_ = nullptr;
	double *pairs_ = nullptr;
	double *scores_ = nullptr;
	uint32_t *selected_ = nullptr;
	Cand *cand_ = nullptr; // per-thread top-2 of the last transform (exact rescoring only)
	Cand *band_ = nullptr; // candidates within the error band
	unsigned *wgflag_ = nullptr; // workgroups of the final transform to rescan
	unsigned band_cap_ = 0;
	Cand *band_host_ = nullptr; // pinned host copy of the band
	// several variants: the device re-decision's history (pools, responses, the dilution
	// table), exact objectives of the band and the arena copy of the program's inputs
	std::vector<uint32_t> gh_e_;
	std::vector<uint8_t> gh_r_;
	std::vector<double> gh_dil_;
	bool gh_broken_ = false, gh_uploaded_ = false;
	double *gexact_ = nullptr;
	double *ghin_ = nullptr;
	// The exact program's frame: a projected lattice (project) keeps the program of the
	// lattice it came from, in that lattice's numbering of xn_ subjects, of which the kept
	// ones, in order, are this lattice's subjects (xkeep_). The marginal of the original
	// posterior over the kept atoms is the projected posterior, and it stays so under updates
	// whose pools hold only kept subjects, so M(E) of this lattice is the program's M of E
	// placed in the frame. Unprojected: xn_ = n_ and every subject kept.
	int xn_ = 0;
	uint32_t xkeep_ = 0;
	double proj_rel_ = 0.0; // FP64: the projections' summation roundings, in units of 2^-53
	uint32_t to_frame(uint32_t subjects) const; // a set of this lattice's subjects in the frame
	uint32_t atoms_to_frame(uint32_t atoms) const; // per variant block
	// the dilution table of the current update (this lattice's rows) at the frame's row
	// stride, rows beyond n_ zero (pools never hold more than n_ subjects)
	void frame_table(const double *dil, std::vector<double> &out) const;
	unsigned head_band_ = 0;    // band candidates that came back with the last head
	ExactRec *hist_dev_ = nullptr; // update history for the device exact program
	double *exact_dev_ = nullptr;  // its exact masses of the first kExactBand band candidates
	int hist_uploaded_ = 0;        // history records already on the device
	bool exact_pending_ = false;   // exact masses were launched for the band being fetched
	bool exact_valid_ = false;     // the host mailbox holds them for the current band
	bool pairs_own_ = false;       // the prior pairs on the device are this lattice's own
	ExactHistory exact_;

	// pinned host mailbox for results
	void *host_box_ = nullptr;

	double lik_[kLikCapacity] = {};
	uint32_t e_ = 0, r_ = 0;
	int len_ = 0;
	int inv_cur_ = kScalInv0;
	uint32_t pos_ = 0, neg_ = 0;
	uint32_t active_ = 0;
	int updates_ = 0;
	bool busy_ = false; // work may be pending on stream_ (asynchronous update or copy)
	KernelClock clock_;
};

// Host-side selector values (bgt_hip.h)
constexpr int kSelAuto = BGT_HIP_SELECT_AUTO;
constexpr int kSelBha = BGT_HIP_SELECT_OP_BHA;
constexpr int kSelBbpa = BGT_HIP_SELECT_OP_BBPA;
constexpr int kSelBrute = BGT_HIP_SELECT_BRUTE_FORCE;

void require_device();
int device_count();

} // namespace bgt_hip
