// Autotuner of the transform's launch forms, keyed by the reference's KernelSpec
// (bgt_hip_tune.cpp).
#pragma once

#include "bgt_hip_internal.hpp"

#include <string>

namespace bgt_hip {

struct TuneKey {
	bool dilution = false;
	int subjects = 0;
	int variants = 1;
	int selector = 1;
	Storage storage = Storage::fp64;
	std::string arch = "gfx950";
};

struct TuneResult {
	std::string id; // the key in the reference's kernel_spec_id format
	int form = 0;   // 0: two passes, 1: three passes
	double ms[2] = {0.0, 0.0}; // measured selection time per form (0 when not measured)
	bool from_cache = false;
};

std::string spec_id(const TuneKey &key);
// Measure (or load from the memory and disk caches unless `force`) and remember the form.
TuneResult tune(const TuneKey &key, bool force);
// The remembered form for lattices of this size and storage: 0 or 1, or -1 when untuned.
int tuned_pass2_form(int atoms, Storage st);

} // namespace bgt_hip
