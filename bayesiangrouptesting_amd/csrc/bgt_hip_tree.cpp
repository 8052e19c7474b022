// run_simulation -> TreeStats on the device (the counterpart of LocalTree,
// src/core/tree/global_tree_intra.cpp:8-40, and the true-state loop of
// src/runtime/simulation_runner.cpp:92-132).
//
// The reference shrinks each child lattice by marginalising subjects whose atoms are all
// classified (lattice_base.cpp:266-316). The device tree keeps every lattice at full size
// and restricts experiments to the still-active subjects instead; the two are equal in
// exact arithmetic (the reference's own no_shrink_tree_reference_test pins this at 1e-12),
// and index order is preserved by Tree::true_ex, so ties resolve the same way.
#include "bgt_hip_tree.hpp"

#include "bgt_hip_batch.hpp"
#include "bgt_hip_resources.hpp"

#include <hip/hip_runtime_api.h>

#include <cstring>
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
	};
	const int atoms = c.n * c.k;
	const int rc = 1 << c.k;
	std::vector<std::unique_ptr<DeviceLattice>> pool;
	auto take = [&](const DeviceLattice &like) {
		if (pool.empty())
			return DeviceLattice::make_like(like);
		auto l = std::move(pool.back());
		pool.pop_back();
		return l;
	};
	auto root = std::make_unique<HostNode>();
	root->pos = root_lat->positive_mask();
	root->neg = root_lat->negative_mask();
	std::vector<Open> frontier;
	if (!is_classified(root->pos, root->neg, atoms) && c.depth > 0)
		frontier.push_back({root.get(), std::move(root_lat)});
	else
		DeviceLattice::recycle(std::move(root_lat));
	std::vector<BatchItem> items;
	std::vector<BatchResult> res;
	while (!frontier.empty()) {
		// select on every open node of this level
		items.assign(frontier.size(), BatchItem{});
		for (size_t i = 0; i < frontier.size(); ++i) {
			const HostNode &nd = *frontier[i].node;
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
			for (int r = 0; r < rc; ++r) {
				auto child = std::make_unique<HostNode>();
				child->ex = e;
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
				it.dilution = c.dil;
				it.select = false;
				items.push_back(it);
				kids.push_back({cp, std::move(lat)});
			}
		}
		batch_step(items, c.selector, c.up, c.lo, res);
		for (auto &o : frontier)
			pool.push_back(std::move(o.lat));
		frontier.clear();
		for (size_t i = 0; i < kids.size(); ++i) {
			HostNode &nd = *kids[i].node;
			nd.pos = res[i].positive;
			nd.neg = res[i].negative;
			if (is_classified(nd.pos, nd.neg, atoms) || nd.stage >= c.depth)
				pool.push_back(std::move(kids[i].lat)); // a leaf
			else
				frontier.push_back(std::move(kids[i]));
		}
	}
	for (auto &l : pool)
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
	TreeCtx c{n, k, depth, DeviceLattice::resolve_selector(selector, k), up, lo, dil};
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
