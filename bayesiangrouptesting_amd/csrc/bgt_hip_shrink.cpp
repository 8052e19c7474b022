// Shrink plans (the host side of removing classified subjects from a lattice).
#include "bgt_hip.h"
#include "bgt_hip_internal.hpp"

#include <cstring>

namespace bgt_hip {

namespace {

uint32_t bit(int i) { return 1u << i; }

// every variant of subject i is in `atoms` (subjects of `subjects` per variant block)
bool subject_covered(uint32_t atoms, int i, int subjects, int variants)
{
	for (int v = 0; v < variants; ++v)
		if (!(atoms & bit(v * subjects + i)))
			return false;
	return true;
}

} // namespace

// make_shrink_plan (shrink_plan.cpp:124-177), with curr_shrinkable_atoms (shrink_plan.cpp:
// 34-56) and update_clas_subj (shrink_plan.cpp:59-79) folded in.
void make_shrink_plan(int orig, int curr, int k, uint32_t classified_atoms, uint32_t already, bgt_hip_shrink_plan &p)
{
	if (orig <= 0 || curr <= 0 || k <= 0 || curr > orig || orig * k > 32)
		throw Error(BGT_HIP_BAD_INPUT, "shrink plan: bad subject or variant counts");
	std::memset(&p, 0, sizeof(p));
	p.original_subjects = orig;
	p.current_subjects = curr;
	p.variants = k;
	p.source_atoms = curr * k;
	for (int &pos : p.original_to_current_atom_position)
		pos = -1;
	// original subject -> current position, skipping the subjects removed before
	int next = 0;
	for (int i = 0; i < orig; ++i) {
		if (already & bit(i))
			continue;
		if (next >= curr)
			throw Error(BGT_HIP_BAD_INPUT, "shrink plan: more subjects left than the current lattice holds");
		for (int v = 0; v < k; ++v) {
			const int a = v * orig + i, c = v * curr + next;
			p.original_to_current_atom_position[a] = c;
			if (classified_atoms & bit(a))
				p.current_classified_atoms |= bit(c);
		}
		++next;
	}
	for (int i = 0; i < curr; ++i)
		if (subject_covered(p.current_classified_atoms, i, curr, k))
			for (int v = 0; v < k; ++v)
				p.removable_current_atoms |= bit(v * curr + i);
	p.reduced_atom_count = __builtin_popcount(p.removable_current_atoms);
	p.base_atom_count = p.source_atoms - p.reduced_atom_count;
	for (int i = 0; i < orig; ++i)
		if (subject_covered(classified_atoms, i, orig, k))
			p.new_classified_subjects |= bit(i);
	p.new_current_subjects = orig - __builtin_popcount(p.new_classified_subjects);
}

} // namespace bgt_hip
