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

// Launch forms of the transform of FP32C lattices of 26..30 atoms.
//   pass-2 form: 0 two passes (from 28 atoms the split pass 2 on the blocked z), 1 three
//                passes, 2 two passes with the plain pass 2 on the row-form z (28..30 atoms)
//   pass-1 form: 0 persistent workgroups (one per CU) that load their next tile while they
//                transform the current one, 1 one workgroup per tile
constexpr int kPass2Forms = 3;
constexpr int kPass1Forms = 2;

struct TuneResult {
	std::string id;    // the key in the reference's kernel_spec_id format
	int form = 0;      // pass-2 form
	int pass1 = 0;     // pass-1 form
	double ms[kPass2Forms][kPass1Forms] = {}; // measured selection time per pair (0: not measured)
	bool from_cache = false;
};

std::string spec_id(const TuneKey &key);
// Measure (or load from the memory and disk caches unless `force`) and remember the forms.
TuneResult tune(const TuneKey &key, bool force);
// The remembered forms for lattices of this size and storage, or -1 when untuned.
int tuned_pass2_form(int atoms, Storage st);
int tuned_pass1_form(int atoms, Storage st);
// -1: the tuned or default choice; otherwise the form, process-wide (tests and the tuner)
void set_pass1_form(int form);

} // namespace bgt_hip
