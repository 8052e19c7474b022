// Batched steps over many device-resident lattices: the continuous-batching core of the
// serving loop and of the level-batched tree. One step applies an (experiment, response)
// update to each lattice that has one, takes atom marginals and classifies, and selects the
// next experiment, with one upload of descriptors and one download of all results.
// Lattices of uneven (n, k) share launches wherever the kernels take their sizes from the
// descriptors (updates, marginals, and the selections of at most 20 atoms); each lattice's
// work is split into the same workgroups as in the per-lattice calls (update, classify,
// select), so its results are identical whether it steps alone or in a batch.
#pragma once

#include "bgt_hip_lattice.hpp"

#include <vector>

namespace bgt_hip {

struct BatchItem {
	DeviceLattice *lat = nullptr;
	bool update = false; // apply (experiment, response) first
	uint32_t experiment = 0, response = 0;
	// with update: read the posterior to update from this lattice (same shape and storage,
	// adopted by `lat` beforehand) instead of lat's own -- the out-of-place update of the
	// tree's children (lattice_base.cpp:210-228). It must not be updated in the same batch.
	const DeviceLattice *source = nullptr;
	const double *dilution = nullptr;
	bool classify = true; // atom marginals + sticky classification
	bool select = true;   // select the next experiment
};

struct BatchResult {
	bgt_hip_selection selection;
	uint32_t positive, negative;
};

// `selector` applies to every item (resolved per lattice shape as in select()).
void batch_step(std::vector<BatchItem> &items, int selector, double threshold_up, double threshold_lo,
				std::vector<BatchResult> &out);

// Graph replay of batched steps: with it on, a step's copies and launches are captured
// into a hipGraph once per launch structure (the launch groups, their sizes and the staging
// buffers) and replayed by steps of the same structure, with the step's own descriptors
// uploaded by the graph's first copy. Returns the previous setting.
bool set_batch_graphs(bool on);
// steps captured into a new graph, and steps that replayed one, since the process started
void batch_graph_counts(unsigned long long *captures, unsigned long long *replays);

} // namespace bgt_hip
