#include "bgt_hip_batch.hpp"
#include "bgt_hip_resources.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <list>
#include <map>
#include <tuple>
#include <utility>
#include <vector>

namespace bgt_hip {

namespace {

std::atomic<int> g_graphs{0};               // replay captured steps (set_batch_graphs)
std::atomic<unsigned long long> g_captures{0}, g_replays{0};

// A step's launch structure: everything its kernel arguments and copies depend on (the
// staging buffers, the upload and download sizes, and per launch group its kind, shape,
// size, descriptor offset and transform form). Steps of equal signature launch the same
// kernels with the same arguments; what differs between them travels in the uploaded block.
using Signature = std::vector<long long>;

// Instantiated graphs of recent step signatures, least recently used first out.
struct GraphCache {
	static constexpr size_t kMax = 16;
	std::list<std::pair<Signature, hipGraphExec_t>> entries; // front: most recent

	hipGraphExec_t find(const Signature &sig)
	{
		for (auto it = entries.begin(); it != entries.end(); ++it)
			if (it->first == sig) {
				entries.splice(entries.begin(), entries, it);
				return entries.front().second;
			}
		return nullptr;
	}
	void insert(Signature sig, hipGraphExec_t ex)
	{
		entries.emplace_front(std::move(sig), ex);
		while (entries.size() > kMax) {
			(void)hipGraphExecDestroy(entries.back().second);
			entries.pop_back();
		}
	}
	void clear()
	{
		for (auto &e : entries)
			(void)hipGraphExecDestroy(e.second);
		entries.clear();
	}
};

// Per-thread staging: one stream, device buffers for descriptors and likelihood tables,
// and the gathered result heads (device and pinned host). Grown on demand, reused.
struct Staging {
	hipStream_t stream = nullptr;
	char *dev = nullptr;
	size_t dev_bytes = 0;
	char *host = nullptr; // pinned: descriptors + tables on the way in, heads on the way out
	size_t host_bytes = 0;
	GraphCache graphs; // graphs name the buffers above: cleared whenever they move

	~Staging()
	{
		graphs.clear();
		if (dev)
			(void)hipFree(dev);
		if (host)
			(void)hipHostFree(host);
		if (stream)
			(void)hipStreamDestroy(stream);
	}
	void reserve(size_t dbytes, size_t hbytes)
	{
		if (stream && dbytes <= dev_bytes && hbytes <= host_bytes)
			return;
		std::lock_guard<std::mutex> gate(capture_gate());
		if (!stream)
			BGT_HIP_CHECK(hipStreamCreate(&stream));
		if (dbytes > dev_bytes || hbytes > host_bytes)
			graphs.clear();
		if (dbytes > dev_bytes) {
			if (dev)
				BGT_HIP_CHECK(hipFree(dev));
			dev = nullptr;
			dev_bytes = std::max(dbytes, 2 * dev_bytes);
			BGT_HIP_CHECK(hipMalloc(&dev, dev_bytes));
		}
		if (hbytes > host_bytes) {
			if (host)
				BGT_HIP_CHECK(hipHostFree(host));
			host = nullptr;
			host_bytes = std::max(hbytes, 2 * host_bytes);
			BGT_HIP_CHECK(hipHostMalloc(&host, host_bytes, hipHostMallocDefault));
		}
	}
};

Staging &staging()
{
	static thread_local Staging s;
	return s;
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Launch groups. A shape group holds lattices of one shape, which share grids and kernel
// templates. A ragged group (Descs::ragged) holds lattices of different shapes whose kernels
// take their sizes and workgroup counts from their descriptors; its grid is sized for its
// largest lattice. Updates and marginals form one ragged group per storage; selections one
// per transform regime (whole transform in LDS, one tile, one to 32 tiles), storage,
// objective and variant count, and a shape group beyond that (one-pass BBPA tables, brute
// force, the persistent pass 2), where templates depend on the shape.
enum Regime : int { kShape = 0, kSmall = 1, kOneTile = 2, kTiles = 3, kAll = 4 };
// regime, storage, dilution | objective, subjects | size class, variants, fused (0: no; 1 + dilution:
// the update fused into the selection's first pass, whose likelihood form must be uniform)
using GroupKey = std::tuple<int, int, int, int, int, int>;

GroupKey shape_key(const DeviceLattice &l)
{
	return {kShape, static_cast<int>(l.storage()), l.dilution() ? 1 : 0, l.subjects(), l.variants(), 0};
}

// Updates and marginals: one ragged group per storage and size class of three atoms. A
// ragged grid is sized for its largest lattice and the workgroups beyond a lattice's own
// count exit at once; a single group for all sizes launched hundreds of thousands of those
// per step (a 2^26-state session beside 2^12-state ones).
GroupKey all_key(const DeviceLattice &l, bool by_size = true)
{
	const int a = l.atoms();
	return {kAll, static_cast<int>(l.storage()), 0, by_size ? (a <= 14 ? 0 : (a - 12) / 3) : 0, 0, 0};
}

int objective_of(int sel, int k) { return sel == kSelBha ? kObjHalving : (k == 1 ? kObjBinaryBbpa : kObjGather); }

GroupKey select_key(const DeviceLattice &l, int selector, bool fused)
{
	const int sel = DeviceLattice::resolve_selector(selector, l.variants());
	const int k = l.variants(), atoms = l.atoms();
	const int objective = objective_of(sel, k);
	const int f = fused ? 1 + (l.dilution() ? 1 : 0) : 0;
	if (sel == kSelBrute || (objective == kObjGather && bbpa_hist_path(l.subjects(), k)) || atoms > kRaggedMaxAtoms) {
		GroupKey key = shape_key(l);
		std::get<5>(key) = f;
		return key;
	}
	const int regime = atoms <= kSmallBits ? kSmall : (atoms == kTileBits ? kOneTile : kTiles);
	// the tiled regime (16..20 atoms) in two size classes, for the same reason as all_key
	return {regime, static_cast<int>(l.storage()), objective, regime == kTiles && atoms >= 18 ? 1 : 0, k, f};
}

// A step that updates, classifies and selects a lattice by auto_select runs as its fused
// round does (DeviceLattice::round) wherever that round fuses, so that the two give the same
// bits: more than one tile (the update in the first pass of the transform, or in the one-pass
// Op-BBPA histogram; the atom masses from the selection's own outputs).
bool fuses(const BatchItem &it, int selector)
{
	return it.update && it.select && it.classify && it.source == nullptr && selector == kSelAuto &&
		   it.lat->atoms() > kTileBits;
}

} // namespace

bool set_batch_graphs(bool on) { return g_graphs.exchange(on ? 1 : 0) != 0; }

void batch_graph_counts(unsigned long long *captures, unsigned long long *replays)
{
	if (captures)
		*captures = g_captures.load();
	if (replays)
		*replays = g_replays.load();
}

void batch_step(std::vector<BatchItem> &items, int selector, double up, double lo, std::vector<BatchResult> &out)
{
	const size_t count = items.size();
	out.assign(count, BatchResult{});
	if (count == 0)
		return;
	// validate everything before any lattice changes state
	std::vector<const DeviceLattice *> seen;
	seen.reserve(count);
	for (auto &it : items) {
		if (it.lat == nullptr)
			throw Error(BGT_HIP_BAD_INPUT, "null lattice in batch");
		seen.push_back(it.lat);
		if (it.select)
			it.lat->check_selector(selector);
		if (it.update)
			it.lat->prepare_update(it.experiment, it.response, it.dilution);
	}
	std::sort(seen.begin(), seen.end());
	if (std::adjacent_find(seen.begin(), seen.end()) != seen.end())
		throw Error(BGT_HIP_BAD_INPUT, "a lattice appears twice in one batch");
	for (auto &it : items) {
		if (it.source == nullptr)
			continue;
		const DeviceLattice &src = *it.source;
		if (!it.update || src.dilution() != it.lat->dilution() || src.subjects() != it.lat->subjects() ||
			src.variants() != it.lat->variants() || src.storage() != it.lat->storage())
			throw Error(BGT_HIP_BAD_INPUT, "an update source must match its lattice's shape and storage");
		for (auto &other : items)
			if (other.update && other.lat == it.source)
				throw Error(BGT_HIP_BAD_INPUT, "an update source is itself updated in the same batch");
	}
	for (auto &it : items) {
		it.lat->settle();
		if (it.source)
			const_cast<DeviceLattice *>(it.source)->settle();
	}
	// host-side preparation in item order: update descriptors carry the pre-update state,
	// selection descriptors the state after it
	std::vector<LatDesc> du(count), ds(count);
	std::vector<const double *> liks(count, nullptr);
	std::vector<char> staged(count, 0); // k_exact_band runs for these lattices' bands
	std::vector<char> fused(count, 0);
	for (size_t i = 0; i < count; ++i) {
		auto &it = items[i];
		if (fuses(it, selector)) {
			// the round's descriptors: the pre-update state with the new experiment's factors
			// (the selection's first pass applies the update), the band width and the exact
			// re-decision's history after it
			fused[i] = 1;
			du[i] = it.lat->desc();
			liks[i] = it.lat->lik_host();
			it.lat->commit_update();
			it.lat->ensure_z();
			it.lat->prepare_gather();
			const LatDesc post = it.lat->desc();
			ds[i] = du[i];
			ds[i].band_width = it.lat->band_width(DeviceLattice::resolve_selector(selector, it.lat->variants()));
			ds[i].gh_len = post.gh_len;
			ds[i].gx_n = post.gx_n;
			ds[i].gx_keep = post.gx_keep;
			std::memcpy(ds[i].gh_e, post.gh_e, sizeof(post.gh_e));
			std::memcpy(ds[i].gh_r, post.gh_r, sizeof(post.gh_r));
			ds[i].gh_pairs = post.gh_pairs;
			ds[i].gh_dil = post.gh_dil;
			ds[i].gexact = post.gexact;
			if (it.lat->storage() == Storage::fp32c)
				staged[i] = it.lat->stage_exact(ds[i], true) ? 1 : 0;
			continue;
		}
		if (it.update) {
			du[i] = it.lat->desc();
			if (it.source)
				it.source->as_source(du[i]);
			liks[i] = it.lat->lik_host();
			it.lat->commit_update();
		}
		if (it.select) {
			it.lat->ensure_z();
			it.lat->prepare_gather();
		}
		ds[i] = it.lat->desc();
		ds[i].band_width =
			it.select ? it.lat->band_width(DeviceLattice::resolve_selector(selector, it.lat->variants())) : 0.0;
		// FP32C binary non-dilution selections: their bands' exact masses on the device, as
		// the per-lattice round computes them (near ties are common in this form)
		if (it.select && it.lat->storage() == Storage::fp32c)
			staged[i] = it.lat->stage_exact(ds[i], true) ? 1 : 0;
	}
	// launch groups (see GroupKey); each group gets its own contiguous block of device
	// descriptors
	struct Group {
		std::vector<size_t> idx;
		size_t offset = 0;  // byte offset of its descriptor block in the staging buffer
		int max_atoms = 0;  // its largest lattice (ragged groups: sizes the grid)
		int max_subjects = 0;
	};
	std::map<GroupKey, Group> gu, gc, gs, gx;
	auto join = [&](std::map<GroupKey, Group> &gm, const GroupKey &key, size_t i) {
		Group &g = gm[key];
		g.idx.push_back(i);
		g.max_atoms = std::max(g.max_atoms, items[i].lat->atoms());
		g.max_subjects = std::max(g.max_subjects, items[i].lat->subjects());
	};
	for (size_t i = 0; i < count; ++i) {
		const DeviceLattice &lat = *items[i].lat;
		if (items[i].update && !fused[i])
			join(gu, all_key(lat), i);
		if (items[i].classify && !fused[i])
			join(gc, all_key(lat), i);
		if (items[i].select)
			join(gs, select_key(lat, selector, fused[i] != 0), i);
		if (staged[i])
			join(gx, all_key(lat, false), i);
	}
	size_t off = 0;
	const size_t o_lik = off;
	off = align256(off + sizeof(double) * kLikCapacity * count);
	for (auto *gm : {&gu, &gc, &gs, &gx})
		for (auto &g : *gm) {
			g.second.offset = off;
			off = align256(off + sizeof(LatDesc) * g.second.idx.size());
		}
	const size_t o_all = off; // selection descriptors of every item, for the head gather
	off = align256(off + sizeof(LatDesc) * count);
	const size_t o_in = off; // bytes uploaded
	const size_t o_heads = off;
	off = align256(off + kHeadBytes * count);
	Staging &st = staging();
	st.reserve(off, off);
	char *h = st.host;
	char *d = st.dev;
	for (size_t i = 0; i < count; ++i)
		if (items[i].update) {
			std::memcpy(h + o_lik + sizeof(double) * kLikCapacity * i, liks[i], sizeof(double) * kLikCapacity);
			du[i].lik = reinterpret_cast<double *>(d + o_lik) + i * kLikCapacity;
			if (fused[i])
				ds[i].lik = du[i].lik;
		}
	auto fill = [&](std::map<GroupKey, Group> &gm, const std::vector<LatDesc> &src) {
		for (auto &g : gm) {
			LatDesc *dst = reinterpret_cast<LatDesc *>(h + g.second.offset);
			for (size_t q = 0; q < g.second.idx.size(); ++q)
				dst[q] = src[g.second.idx[q]];
		}
	};
	fill(gu, du);
	fill(gc, ds);
	fill(gs, ds);
	fill(gx, ds);
	std::memcpy(h + o_all, ds.data(), sizeof(LatDesc) * count);
	hipStream_t s = st.stream;
	// the step's copies and launches, on the stream or into a graph being captured from it
	auto enqueue = [&] {
		BGT_HIP_CHECK(hipMemcpyAsync(d, h, o_in, hipMemcpyHostToDevice, s));
		auto descs_at = [&](size_t offset, bool ragged) {
			Descs x;
			std::memset(&x, 0, sizeof(x));
			x.many = reinterpret_cast<const LatDesc *>(d + offset);
			x.ragged = ragged ? 1 : 0;
			return x;
		};
		// 1. updates (x L and the denominator)
		for (auto &g : gu) {
			const DeviceLattice &lat = *items[g.second.idx[0]].lat;
			launch_update(descs_at(g.second.offset, true), (int)g.second.idx.size(), 1ull << g.second.max_atoms,
						  lat.variants(), lat.dilution(), lat.storage(), s);
		}
		// 2. atom marginals of the updated posteriors
		for (auto &g : gc) {
			const DeviceLattice &lat = *items[g.second.idx[0]].lat;
			launch_atom_masses(descs_at(g.second.offset, true), (int)g.second.idx.size(), g.second.max_atoms,
							   lat.storage(), s);
		}
		// 3. selection, as DeviceLattice::select does for one lattice (a ragged group's sizes
		// are those of its largest lattice; each lattice takes its own from its descriptor)
		for (auto &g : gs) {
			const DeviceLattice &lat = *items[g.second.idx[0]].lat;
			const int cnt = (int)g.second.idx.size();
			const bool ragged = std::get<0>(g.first) != kShape;
			const Descs dg = descs_at(g.second.offset, ragged);
			const int sel = DeviceLattice::resolve_selector(selector, lat.variants());
			if (sel == kSelBrute) {
				launch_brute_force(dg, cnt, lat.subjects(), lat.variants(), lat.storage(), s);
				continue;
			}
			const int k = lat.variants();
			const int objective = objective_of(sel, k);
			const double target = sel == kSelBha ? 0.5 : 1.0 / double(1 << k);
			const int atoms = g.second.max_atoms;
			const int subjects = g.second.max_subjects;
			// fused (as DeviceLattice::round): the group's update runs in the first pass, the
			// outputs are scaled by the new 1/D; its likelihood form is the group's (dilution
			// is part of the key)
			const int fkey = std::get<5>(g.first);
			const bool fused = fkey != 0;
			const int dil = fused ? fkey - 1 : (lat.dilution() ? 1 : 0);
			if (objective == kObjGather && bbpa_hist_path(lat.subjects(), k)) {
				// as DeviceLattice::select: the feature tables from one histogram pass
				launch_bbpa_hist(dg, cnt, lat.subjects(), k, lat.storage(), fused, s, false);
				launch_bbpa_finish(dg, cnt, k, fused, target, s);
				continue;
			}
			unsigned entries = 0;
			if (atoms <= kSmallBits) {
				entries = launch_zeta_small(dg, cnt, atoms, k, lat.storage(), objective, target, s);
			} else if (atoms == kTileBits) {
				entries = launch_zeta_final(dg, cnt, atoms, k, lat.storage(), objective, true, false, target, s);
			} else {
				// a persistent pass 2 sums the fused update's pass-1 partials itself (shape groups)
				const bool fold = fused && !ragged && persistent_pass2(atoms, objective, false);
				launch_zeta_pass1(dg, cnt, atoms, lat.storage(), fused, k, dil, objective, s, fold);
				entries = launch_zeta_final(dg, cnt, atoms, k, lat.storage(), objective, false, fused, target, s,
											fold ? (1u << (atoms - kTileBits)) : 0u);
			}
			if (objective == kObjGather) {
				launch_bbpa_score(dg, cnt, subjects, k, target, s, fused);
				launch_bbpa_band(dg, cnt, subjects, k, target, s, fused);
			} else if (k == 1 && atoms > kSmallBits)
				launch_band_collect(dg, cnt, entries, atoms, k, lat.storage(), objective, atoms == kTileBits, fused,
									target, s);
		}
		// 4. the exact masses of the staged bands (lattices whose top two are no near tie return
		// at once; a few workgroups per lattice share its band entries)
		for (auto &g : gx)
			launch_exact_band(descs_at(g.second.offset, true), (int)g.second.idx.size(), g.second.max_atoms, s, 16);
		// 5. one copy brings every head back
		launch_gather_heads(reinterpret_cast<const LatDesc *>(d + o_all), (int)count,
							reinterpret_cast<unsigned char *>(d + o_heads), s);
		BGT_HIP_CHECK(hipMemcpyAsync(h + o_heads, d + o_heads, kHeadBytes * count, hipMemcpyDeviceToHost, s));
	};
	if (g_graphs.load() != 0) {
		Signature sig{(long long)count,
					  selector,
					  (long long)o_in,
					  (long long)o_heads,
					  (long long)reinterpret_cast<intptr_t>(d),
					  (long long)reinterpret_cast<intptr_t>(h)};
		int kind = 0;
		for (auto *gm : {&gu, &gc, &gs, &gx}) {
			for (auto &g : *gm) {
				const DeviceLattice &lat = *items[g.second.idx[0]].lat;
				sig.insert(sig.end(), {kind, std::get<0>(g.first), std::get<1>(g.first), std::get<2>(g.first),
									   std::get<3>(g.first), std::get<4>(g.first), std::get<5>(g.first),
									   (long long)g.second.idx.size(),
									   (long long)g.second.offset, g.second.max_atoms, g.second.max_subjects,
									   three_pass_form(lat.atoms(), lat.storage()) ? 1 : 0});
			}
			++kind;
		}
		hipGraphExec_t ex = st.graphs.find(sig);
		if (ex == nullptr) {
			// no allocation on any thread while the capture is open (see capture_gate)
			std::lock_guard<std::mutex> gate(capture_gate());
			// relaxed capture: other threads (lattice construction, say) keep allocating; the
			// captured region itself makes stream operations on `s` only
			BGT_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
			hipGraph_t graph = nullptr;
			try {
				enqueue();
			} catch (...) {
				if (hipStreamEndCapture(s, &graph) == hipSuccess && graph != nullptr)
					(void)hipGraphDestroy(graph);
				throw;
			}
			BGT_HIP_CHECK(hipStreamEndCapture(s, &graph));
			const hipError_t e = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
			(void)hipGraphDestroy(graph);
			BGT_HIP_CHECK(e);
			st.graphs.insert(std::move(sig), ex);
			g_captures.fetch_add(1);
		} else {
			g_replays.fetch_add(1);
		}
		BGT_HIP_CHECK(hipGraphLaunch(ex, s));
	} else {
		enqueue();
	}
	for (auto &it : items)
		if (it.select)
			it.lat->prepare_exact(); // on the host while the device works
	BGT_HIP_CHECK(hipStreamSynchronize(s));
	for (size_t i = 0; i < count; ++i)
		items[i].lat->load_head(reinterpret_cast<const unsigned char *>(h + o_heads + kHeadBytes * i), staged[i] != 0);
	// the exact normalisers of the lattices whose selections will be re-decided, side by side
	// on the host threads (each lattice's program is its own; the results do not depend on
	// which thread computes them)
	std::vector<size_t> tied;
	for (size_t i = 0; i < count; ++i)
		if (items[i].select)
			tied.push_back(i);
	if (tied.size() > 1)
		parallel_for(tied.size(), [&](size_t j, int) { items[tied[j]].lat->prepare_rescore(selector); });
	// the binary dilution lattices' first exact masses by one enumeration launch and one
	// synchronisation for the step, where each lattice's re-decision would make its own
	{
		std::vector<DeviceLattice *> enum_lats;
		std::vector<std::vector<uint32_t>> enum_sets;
		for (size_t i : tied) {
			if (!items[i].lat->uses_exact_enum())
				continue;
			std::vector<uint32_t> sets = items[i].lat->enum_ahead_sets(selector);
			if (sets.empty())
				continue;
			enum_lats.push_back(items[i].lat);
			enum_sets.push_back(std::move(sets));
		}
		if (!enum_lats.empty())
			DeviceLattice::enum_ahead_run(enum_lats, enum_sets, s);
	}
	for (size_t i = 0; i < count; ++i) {
		auto &it = items[i];
		if (it.classify)
			it.lat->classify_from_head(up, lo, nullptr);
		if (it.select)
			out[i].selection = it.lat->finish_selection(selector, it.lat->head_result(), false);
		out[i].positive = it.lat->positive_mask();
		out[i].negative = it.lat->negative_mask();
	}
}

} // namespace bgt_hip
