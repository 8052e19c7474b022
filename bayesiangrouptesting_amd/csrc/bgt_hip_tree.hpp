#pragma once

#include "bgt_hip_lattice.hpp"

namespace bgt_hip {

struct TreeBuildStats {
	int selections = 0;
	int near_ties = 0;
	int projections = 0; // children replaced by their projection (shrinking trees)
};

// Shrink children whose classification covers a subject, or keep full-size lattices with a
// candidate restriction (default); returns the former setting.
bool set_tree_shrink(bool enable);

bgt_hip_tree_stats run_tree(bool dilution, int n, int k, const double *prior, int depth, double up, double lo,
							double branch, int selector, const double *dil, Storage storage,
							TreeBuildStats *bs = nullptr);

} // namespace bgt_hip
