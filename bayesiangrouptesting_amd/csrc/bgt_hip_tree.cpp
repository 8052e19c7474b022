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

std::unique_ptr<HostNode> build(const TreeCtx &c, std::unique_ptr<DeviceLattice> lat, uint32_t ex, uint32_t res,
								int stage, TreeBuildStats &bs)
{
	auto node = std::make_unique<HostNode>();
	node->ex = ex;
	node->res = res;
	node->stage = stage;
	node->pos = lat->positive_mask();
	node->neg = lat->negative_mask();
	const int atoms = c.n * c.k;
	const bool classified = __builtin_popcount(node->pos | node->neg) == atoms;
	if (classified || stage >= c.depth)
		return node;
	lat->set_active(active_subjects(node->pos | node->neg, c.n, c.k));
	const bgt_hip_selection sel = lat->select(c.selector);
	bs.selections += 1;
	bs.near_ties += sel.near_tie;
	const uint32_t e = sel.experiment;
	const int rc = 1 << c.k;
	for (int r = 0; r < rc; ++r) {
		std::unique_ptr<DeviceLattice> child;
		if (r != rc - 1)
			child = std::make_unique<DeviceLattice>(*lat, true);
		else
			child = std::move(lat);
		child->update(e, (uint32_t)r, c.dil);
		child->classify(c.up, c.lo);
		node->children.push_back(build(c, std::move(child), e, (uint32_t)r, stage + 1, bs));
	}
	return node;
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
	auto root_lat = std::make_unique<DeviceLattice>(dilution, n, k, prior, storage);
	hipStream_t s = nullptr;
	TreeBuildStats bs;
	auto root = build(c, std::move(root_lat), 0, 0, 0, bs);
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
	BGT_HIP_CHECK(hipStreamCreate(&s));
	BGT_HIP_CHECK(hipMalloc(&d_nodes, sizeof(FlatNode) * nodes.size()));
	BGT_HIP_CHECK(hipMalloc(&d_dil, sizeof(double) * dil_count));
	BGT_HIP_CHECK(hipMalloc(&d_pi, sizeof(double) * atoms));
	BGT_HIP_CHECK(hipMalloc(&d_part, sizeof(double) * 7 * 4096));
	BGT_HIP_CHECK(hipMalloc(&d_out, sizeof(double) * 8));
	BGT_HIP_CHECK(hipMalloc(&d_counter, sizeof(unsigned) * 4));
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
	(void)hipFree(d_nodes);
	(void)hipFree(d_dil);
	(void)hipFree(d_pi);
	(void)hipFree(d_part);
	(void)hipFree(d_out);
	(void)hipFree(d_counter);
	(void)hipStreamDestroy(s);
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
