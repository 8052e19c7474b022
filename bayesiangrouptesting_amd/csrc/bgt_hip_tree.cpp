// run_simulation -> TreeStats on the device (the counterpart of LocalTree,
// src/core/tree/global_tree_intra.cpp:8-40, and the true-state loop of
// src/runtime/simulation_runner.cpp:92-132).
//
// The reference shrinks each child lattice by marginalising subjects whose atoms are all
// classified (lattice_base.cpp:266-316). The device tree does the same when shrinking is on
// (set_tree_shrink): a child whose classification covers every variant of some subject is
// replaced by its projection (DeviceLattice::project), selects in its own numbering, and
// the experiment is mapped back by Tree::true_ex (tree.cpp:55-68). By default it keeps
// every lattice at full size and restricts experiments to the still-active subjects
// instead: the two are equal in exact arithmetic (the reference's own
// no_shrink_tree_reference_test pins this at 1e-12), index order is preserved by true_ex,
// so ties resolve the same way, and at tree sizes whose levels are bound by launches (2^16
// states, depth 6) the projections cost more than the smaller children save.
#include "bgt_hip_tree.hpp"

#include "bgt_hip_batch.hpp"
#include "bgt_hip_resources.hpp"

#include <hip/hip_runtime_api.h>

#include <atomic>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

namespace bgt_hip {

namespace {

struct HostNode {
	uint32_t ex = 0, res = 0, pos = 0, neg = 0;
	int stage = 0;
	std::vector<std::unique_ptr<HostNode>> children;
};

struct TreeCtx {
	int n, k, depth, selector;
	double up, lo;
	const double *dil;
	bool shrink;
};

std::atomic<bool> g_tree_shrink{false};

// A node's lattice numbering: the original subjects still in it, in order (the removed ones
// are `removed`); atom v * cur + j of the lattice is original atom v * n + kept[j].
struct Frame {
	std::vector<int> kept;
	uint32_t removed = 0;

	uint32_t to_original(uint32_t x, int n, int k) const
	{
		const int cur = (int)kept.size();
		uint32_t out = 0;
		for (int v = 0; v < k; ++v)
			for (int j = 0; j < cur; ++j)
				if ((x >> (v * cur + j)) & 1u)
					out |= 1u << (v * n + kept[j]);
		return out;
	}
};

uint32_t active_subjects(uint32_t classified, int n, int k)
{
	// subjects with at least one unclassified atom (shrink_plan.cpp:34-79)
	uint32_t active = 0;
	for (int i = 0; i < n; ++i) {
		bool all = true;
		for (int v = 0; v < k; ++v)
			if (!((classified >> (v * n + i)) & 1u))
				all = false;
		if (!all)
			active |= 1u << i;
	}
	return active;
}

bool is_classified(uint32_t pos, uint32_t neg, int atoms) { return __builtin_popcount(pos | neg) == atoms; }

// LocalTree construction (global_tree_intra.cpp:8-40), one level at a time: every
// unfinished node of a level selects in one batched step, and all of their children are
// produced in one more -- each child's posterior written by the out-of-place update of its
// parent's (read the parent once, write the children; update_probs, lattice_base.cpp:210-228)
// and classified. A node becomes a leaf when it is classified or at the depth limit. Device
// lattices are recycled: a leaf's lattice, and a parent's once its children exist, return to
// a pool that the next level draws from, so a tree allocates about two levels' worth.
std::unique_ptr<HostNode> build_levels(const TreeCtx &c, std::unique_ptr<DeviceLattice> root_lat, TreeBuildStats &bs)
{
	struct Open {
		HostNode *node;
		std::unique_ptr<DeviceLattice> lat;
		Frame frame;
	};
	const int atoms = c.n * c.k;
	const int rc = 1 << c.k;
	// lattices to reuse, by subject count (projected children come in every size)
	std::map<int, std::vector<std::unique_ptr<DeviceLattice>>> pool;
	auto take = [&](const DeviceLattice &like) {
		auto &free = pool[like.subjects()];
		if (free.empty())
			return DeviceLattice::make_like(like);
		auto l = std::move(free.back());
		free.pop_back();
		return l;
	};
	auto give = [&](std::unique_ptr<DeviceLattice> l) {
		const int n = l->subjects();
		pool[n].push_back(std::move(l));
	};
	// the dilution table at each subject count: the original's leading rows (pool sizes
	// at most the count), at the row stride of that count
	std::map<int, std::vector<double>> tables;
	auto table_for = [&](int cur) -> const double * {
		if (c.dil == nullptr || cur == c.n)
			return c.dil;
		auto &t = tables[cur];
		if (t.empty()) {
			t.resize((size_t)cur * (cur + 1));
			for (int row = 0; row < cur; ++row)
				std::memcpy(t.data() + (size_t)row * (cur + 1), c.dil + (size_t)row * (c.n + 1),
							sizeof(double) * (cur + 1));
		}
		return t.data();
	};
	auto root = std::make_unique<HostNode>();
	root->pos = root_lat->positive_mask();
	root->neg = root_lat->negative_mask();
	Frame root_frame;
	for (int i = 0; i < c.n; ++i)
		root_frame.kept.push_back(i);
	std::vector<Open> frontier;
	if (!is_classified(root->pos, root->neg, atoms) && c.depth > 0)
		frontier.push_back({root.get(), std::move(root_lat), std::move(root_frame)});
	else
		DeviceLattice::recycle(std::move(root_lat));
	std::vector<BatchItem> items;
	std::vector<BatchResult> res;
	while (!frontier.empty()) {
		// select on every open node of this level
		items.assign(frontier.size(), BatchItem{});
		for (size_t i = 0; i < frontier.size(); ++i) {
			const HostNode &nd = *frontier[i].node;
			// a shrunk lattice holds only subjects with an unclassified atom
			if (!c.shrink)
				frontier[i].lat->set_active(active_subjects(nd.pos | nd.neg, c.n, c.k));
			items[i].lat = frontier[i].lat.get();
			items[i].classify = false;
		}
		batch_step(items, c.selector, c.up, c.lo, res);
		// children: one per response, each a lattice from the pool that adopts its parent's
		// state and receives the parent's posterior times the likelihood
		std::vector<Open> kids;
		kids.reserve(frontier.size() * rc);
		items.clear();
		for (size_t i = 0; i < frontier.size(); ++i) {
			const uint32_t e = res[i].selection.experiment;
			bs.selections += 1;
			bs.near_ties += res[i].selection.near_tie;
			HostNode &nd = *frontier[i].node;
			const DeviceLattice &parent = *frontier[i].lat;
			const Frame &fr = frontier[i].frame;
			// the experiment in the original numbering (true_ex, tree.cpp:55-68)
			const uint32_t full = (e == 0xFFFFFFFFu || !c.shrink) ? e : fr.to_original(e, c.n, 1);
			for (int r = 0; r < rc; ++r) {
				auto child = std::make_unique<HostNode>();
				child->ex = full;
				child->res = (uint32_t)r;
				child->stage = nd.stage + 1;
				HostNode *cp = child.get();
				nd.children.push_back(std::move(child));
				auto lat = take(parent);
				lat->adopt(parent);
				BatchItem it;
				it.lat = lat.get();
				it.update = true;
				it.experiment = e;
				it.response = (uint32_t)r;
				it.source = &parent;
				it.dilution = table_for(parent.subjects());
				it.select = false;
				items.push_back(it);
				kids.push_back({cp, std::move(lat), fr});
			}
		}
		batch_step(items, c.selector, c.up, c.lo, res);
		std::vector<HostNode *> parents(frontier.size());
		for (size_t i = 0; i < frontier.size(); ++i)
			parents[i] = frontier[i].node;
		for (auto &o : frontier)
			give(std::move(o.lat));
		frontier.clear();
		for (size_t i = 0; i < kids.size(); ++i) {
			Open &kid = kids[i];
			HostNode &nd = *kid.node;
			if (c.shrink) {
				// classification in the original numbering: the parent's (removed subjects'
				// atoms included) and this child's sticky masks over its own atoms
				const HostNode &pn = *parents[i / rc];
				nd.pos = pn.pos | kid.frame.to_original(res[i].positive, c.n, c.k);
				nd.neg = pn.neg | kid.frame.to_original(res[i].negative, c.n, c.k);
			} else {
				nd.pos = res[i].positive;
				nd.neg = res[i].negative;
			}
			if (is_classified(nd.pos, nd.neg, atoms) || nd.stage >= c.depth) {
				give(std::move(kid.lat)); // a leaf
				continue;
			}
			if (c.shrink) {
				// update_metadata_with_shrinking's plan (lattice_base.cpp:266-305) and its
				// projection (apply_shrink_plan, 307-316)
				bgt_hip_shrink_plan plan;
				make_shrink_plan(c.n, kid.lat->subjects(), c.k, nd.pos | nd.neg, kid.frame.removed, plan);
				if (plan.removable_current_atoms != 0u && plan.new_current_subjects > 0) {
					auto projected = DeviceLattice::project(*kid.lat, plan.removable_current_atoms);
					give(std::move(kid.lat));
					kid.lat = std::move(projected);
					std::vector<int> kept;
					for (size_t j = 0; j < kid.frame.kept.size(); ++j)
						if (!((plan.removable_current_atoms >> j) & 1u))
							kept.push_back(kid.frame.kept[j]);
					kid.frame.kept = std::move(kept);
					kid.frame.removed = plan.new_classified_subjects;
					bs.projections += 1;
				}
			}
			frontier.push_back(std::move(kid));
		}
	}
	for (auto &entry : pool)
		for (auto &l : entry.second)
			DeviceLattice::recycle(std::move(l));
	return root;
}

int count_leaves(const HostNode &n)
{
	if (n.children.empty())
		return 1;
	int t = 0;
	for (auto &ch : n.children)
		t += count_leaves(*ch);
	return t;
}

void flatten(const HostNode &n, std::vector<FlatNode> &out)
{
	const int me = (int)out.size();
	FlatNode f;
	f.ex = n.ex;
	f.res = n.res;
	f.pos = n.pos;
	f.neg = n.neg;
	f.stage = n.stage;
	f.nchild = (int)n.children.size();
	f.first_child = me + 1;
	f.subtree_end = 0;
	out.push_back(f);
	for (auto &ch : n.children)
		flatten(*ch, out);
	out[me].subtree_end = (int)out.size();
}

} // namespace

bool set_tree_shrink(bool enable) { return g_tree_shrink.exchange(enable); }

bgt_hip_tree_stats run_tree(bool dilution, int n, int k, const double *prior, int depth, double up, double lo,
							double branch, int selector, const double *dil, Storage storage, TreeBuildStats *bs_out)
{
	// validation as CpuProvider::run (providers.cpp:292-303)
	if (n <= 0 || k <= 0 || depth < 0)
		throw Error(BGT_HIP_BAD_INPUT, "invalid simulation dimensions");
	if (depth > 29)
		throw Error(BGT_HIP_UNSUPPORTED, "search depth above 29 exceeds the statistics arrays");
	if (selector < 0 || selector > 3)
		throw Error(BGT_HIP_BAD_INPUT, "unsupported selector");
	if (DeviceLattice::resolve_selector(selector, k) == kSelBha && k != 1)
		throw Error(BGT_HIP_BAD_INPUT, "Op-BHA tree selection is only supported for binary lattices");
	if (dilution && dil == nullptr)
		throw Error(BGT_HIP_BAD_INPUT, "dilution table is required for dilution models");
	TreeCtx c{n, k, depth, DeviceLattice::resolve_selector(selector, k), up, lo, dil, g_tree_shrink.load()};
	auto root_lat = DeviceLattice::make(dilution, n, k, prior, storage);
	hipStream_t s = nullptr;
	TreeBuildStats bs;
	auto root = build_levels(c, std::move(root_lat), bs);
	std::vector<FlatNode> nodes;
	flatten(*root, nodes);
	bgt_hip_tree_stats stats;
	std::memset(&stats, 0, sizeof(stats));
	stats.total_leaves = count_leaves(*root);

	const int atoms = n * k;
	FlatNode *d_nodes = nullptr;
	double *d_dil = nullptr, *d_pi = nullptr, *d_part = nullptr, *d_out = nullptr;
	unsigned *d_counter = nullptr;
	const size_t dil_count = dilution ? (size_t)n * (n + 1) : 1;
	s = stream_acquire();
	d_nodes = static_cast<FlatNode *>(device_alloc(sizeof(FlatNode) * nodes.size(), s));
	d_dil = static_cast<double *>(device_alloc(sizeof(double) * dil_count, s));
	d_pi = static_cast<double *>(device_alloc(sizeof(double) * atoms, s));
	d_part = static_cast<double *>(device_alloc(sizeof(double) * 7 * 4096, s));
	d_out = static_cast<double *>(device_alloc(sizeof(double) * 8, s));
	d_counter = static_cast<unsigned *>(device_alloc(sizeof(unsigned) * 4, s));
	BGT_HIP_CHECK(hipMemsetAsync(d_counter, 0, sizeof(unsigned) * 4, s));
	BGT_HIP_CHECK(hipMemcpyAsync(d_nodes, nodes.data(), sizeof(FlatNode) * nodes.size(), hipMemcpyHostToDevice, s));
	if (dilution)
		BGT_HIP_CHECK(hipMemcpyAsync(d_dil, dil, sizeof(double) * dil_count, hipMemcpyHostToDevice, s));
	BGT_HIP_CHECK(hipMemcpyAsync(d_pi, prior, sizeof(double) * atoms, hipMemcpyHostToDevice, s));
	launch_tree_stats(d_nodes, (int)nodes.size(), n, k, dilution ? 1 : 0, d_dil, d_pi, depth, branch, d_part,
					  d_counter, d_out, s);
	double out[8] = {};
	BGT_HIP_CHECK(hipMemcpyAsync(out, d_out, sizeof(double) * 7, hipMemcpyDeviceToHost, s));
	BGT_HIP_CHECK(hipStreamSynchronize(s));
	device_free(d_nodes, s);
	device_free(d_dil, s);
	device_free(d_pi, s);
	device_free(d_part, s);
	device_free(d_out, s);
	device_free(d_counter, s);
	stream_release(s);
	stats.correct_probability = out[0];
	stats.incorrect_probability = out[1];
	stats.false_positive_probability = out[2];
	stats.false_negative_probability = out[3];
	stats.unclassified_probability = out[4];
	stats.expected_stages = out[5];
	stats.expected_tests = out[6];
	if (bs_out)
		*bs_out = bs;
	return stats;
}

} // namespace bgt_hip
