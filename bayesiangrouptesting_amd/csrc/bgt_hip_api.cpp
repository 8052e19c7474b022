// extern "C" boundary of the provider (include/bgt_hip.h). No exception crosses it.
#include "bgt_hip.h"
#include "bgt_hip_batch.hpp"
#include "bgt_hip_exact.hpp"
#include "bgt_hip_lattice.hpp"
#include "bgt_hip_noalloc.hpp"
#include "bgt_hip_tune.hpp"
#include "bgt_hip_tree.hpp"

#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

struct bgt_hip_lattice {
	std::unique_ptr<bgt_hip::DeviceLattice> impl;
};

namespace {

thread_local std::string g_last_error;

// Every entry point runs with the calling thread in relaxed stream-capture interaction: a
// batched step that another thread captures into a graph (bgt_hip_batch_graphs) must not
// make this thread's allocations and copies on its own streams fail. Nothing the library
// does on one thread touches a stream another thread captures.
struct RelaxedCaptureInteraction {
	hipStreamCaptureMode prev = hipStreamCaptureModeRelaxed;
	bool swapped = false;
	RelaxedCaptureInteraction() { swapped = hipThreadExchangeStreamCaptureMode(&prev) == hipSuccess; }
	~RelaxedCaptureInteraction()
	{
		if (swapped)
			(void)hipThreadExchangeStreamCaptureMode(&prev);
	}
};

template <typename F>
int guarded(F &&f)
{
	RelaxedCaptureInteraction relaxed;
	try {
		f();
		g_last_error.clear();
		return BGT_HIP_OK;
	} catch (const bgt_hip::Error &e) {
		g_last_error = e.what();
		return e.code;
	} catch (const std::bad_alloc &) {
		g_last_error = "host allocation failed";
		return BGT_HIP_INTERNAL;
	} catch (const std::exception &e) {
		g_last_error = e.what();
		return BGT_HIP_INTERNAL;
	}
}

bgt_hip::DeviceLattice &L(bgt_hip_lattice *h)
{
	if (h == nullptr || !h->impl)
		throw bgt_hip::Error(BGT_HIP_BAD_INPUT, "null lattice handle");
	return *h->impl;
}

bgt_hip::Storage storage_of(int precision)
{
	if (precision == BGT_HIP_FP64)
		return bgt_hip::Storage::fp64;
	if (precision == BGT_HIP_FP32C)
		return bgt_hip::Storage::fp32c;
	throw bgt_hip::Error(BGT_HIP_BAD_INPUT, "unknown precision");
}

} // namespace

extern "C" {

const char *bgt_hip_last_error(void) { return g_last_error.c_str(); }

int bgt_hip_provider_available(void)
{
	try {
		bgt_hip::require_device();
		return 1;
	} catch (...) {
		return 0;
	}
}

int bgt_hip_device_synchronize(void)
{
	return guarded([&] { BGT_HIP_CHECK(hipDeviceSynchronize()); });
}

int bgt_hip_device_name(char *out, int len)
{
	return guarded([&] {
		if (out == nullptr || len <= 0)
			throw bgt_hip::Error(BGT_HIP_BAD_INPUT, "null buffer");
		if (bgt_hip::device_count() <= 0)
			throw bgt_hip::Error(BGT_HIP_UNSUPPORTED, "no HIP device");
		hipDeviceProp_t p;
		BGT_HIP_CHECK(hipGetDeviceProperties(&p, 0));
		std::snprintf(out, (size_t)len, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
	});
}

int bgt_hip_provider_run_precision(int use_dilution, int subjects, int variants, const double *prior, int search_depth,
								   double threshold_up, double threshold_lo, double branch_threshold, int selector,
								   const double *dilution, int precision, bgt_hip_tree_stats *stats)
{
	return guarded([&] {
		if (stats == nullptr || prior == nullptr)
			throw bgt_hip::Error(BGT_HIP_BAD_INPUT, "null argument");
		*stats = bgt_hip::run_tree(use_dilution != 0, subjects, variants, prior, search_depth, threshold_up,
								   threshold_lo, branch_threshold, selector, dilution, storage_of(precision));
	});
}

int bgt_hip_provider_run(int use_dilution, int subjects, int variants, const double *prior, int search_depth,
						 double threshold_up, double threshold_lo, double branch_threshold, int selector,
						 const double *dilution, bgt_hip_tree_stats *stats)
{
	return bgt_hip_provider_run_precision(use_dilution, subjects, variants, prior, search_depth, threshold_up,
										  threshold_lo, branch_threshold, selector, dilution, BGT_HIP_FP64, stats);
}

int bgt_hip_provider_run_distributed(int, int, int, const double *, int, double, double, double, int, int,
									 const double *, bgt_hip_tree_stats *)
{
	g_last_error = "the HIP provider runs on one GPU; multi-rank evaluation is not provided";
	return BGT_HIP_NO_BACKEND;
}

int bgt_hip_provider_benchmark_select(int subjects, int variants, const double *prior, int selector, int iterations,
									  int warmup, uint32_t *candidate, double *init_seconds, double *mean_seconds,
									  double *min_seconds, double *max_seconds, double *stddev_seconds)
{
	// timing shape of bgt_cuda_provider_benchmark_select (cuda_provider.cu:2449-2551)
	// argument checks and codes as the reference's: 1 for a missing output, 2 for a shape or
	// selector it cannot run (Op-BHA needs one variant)
	if (candidate == nullptr || init_seconds == nullptr || mean_seconds == nullptr || min_seconds == nullptr ||
		max_seconds == nullptr || stddev_seconds == nullptr) {
		g_last_error = "benchmark_select: every output pointer is required";
		return BGT_HIP_BAD_INPUT;
	}
	*candidate = 0;
	*init_seconds = *mean_seconds = *min_seconds = *max_seconds = *stddev_seconds = 0.0;
	if (subjects <= 0 || variants <= 0 || iterations <= 0 || warmup < 0 || prior == nullptr ||
		subjects * variants > 30 || variants > 10 ||
		(variants != 1 && selector == BGT_HIP_SELECT_OP_BHA)) {
		g_last_error = "benchmark_select: unsupported shape or selector";
		return BGT_HIP_UNSUPPORTED;
	}
	return guarded([&] {
		using clk = std::chrono::steady_clock;
		const auto t0 = clk::now();
		bgt_hip::DeviceLattice lat(false, subjects, variants, prior, bgt_hip::Storage::fp64);
		const double init = std::chrono::duration<double>(clk::now() - t0).count();
		bgt_hip_selection sel{};
		for (int i = 0; i < warmup; ++i)
			sel = lat.select(selector);
		std::vector<double> t(iterations);
		for (int i = 0; i < iterations; ++i) {
			const auto a = clk::now();
			sel = lat.select(selector);
			t[i] = std::chrono::duration<double>(clk::now() - a).count();
		}
		double mean = 0, mn = t[0], mx = t[0];
		for (double x : t) {
			mean += x;
			mn = std::min(mn, x);
			mx = std::max(mx, x);
		}
		mean /= iterations;
		double var = 0;
		for (double x : t)
			var += (x - mean) * (x - mean);
		var /= iterations;
		*candidate = sel.experiment;
		*init_seconds = init;
		*mean_seconds = mean;
		*min_seconds = mn;
		*max_seconds = mx;
		*stddev_seconds = std::sqrt(var);
	});
}

int bgt_hip_lattice_create(int use_dilution, int subjects, int variants, const double *prior, int precision,
						   bgt_hip_lattice **out)
{
	return guarded([&] {
		if (out == nullptr)
			throw bgt_hip::Error(BGT_HIP_BAD_INPUT, "null output handle");
		*out = nullptr;
		auto h = std::make_unique<bgt_hip_lattice>();
		h->impl = bgt_hip::DeviceLattice::make(use_dilution != 0, subjects, variants, prior, storage_of(precision));
		*out = h.release();
	});
}

int bgt_hip_lattice_destroy(bgt_hip_lattice *lattice)
{
	return guarded([&] {
		if (lattice != nullptr)
			bgt_hip::DeviceLattice::recycle(std::move(lattice->impl));
		delete lattice;
	});
}

int bgt_hip_lattice_update(bgt_hip_lattice *h, uint32_t experiment, uint32_t response, const double *dilution)
{
	return guarded([&] { L(h).update(experiment, response, dilution); });
}

int bgt_hip_lattice_select(bgt_hip_lattice *h, int selector, bgt_hip_selection *out)
{
	return guarded([&] {
		if (out == nullptr)
			throw bgt_hip::Error(BGT_HIP_BAD_INPUT, "null output");
		*out = L(h).select(selector);
	});
}

int bgt_hip_lattice_posterior(bgt_hip_lattice *h, uint32_t state, double *out)
{
	return guarded([&] {
		if (out == nullptr)
			throw bgt_hip::Error(BGT_HIP_BAD_INPUT, "null output");
		*out = L(h).posterior(state);
	});
}

int bgt_hip_lattice_read_posterior(bgt_hip_lattice *h, double *out, uint64_t count)
{
	return guarded([&] {
		auto &l = L(h);
		if (out == nullptr || count != l.states())
			throw bgt_hip::Error(BGT_HIP_BAD_INPUT, "output must hold 2^(subjects*variants) doubles");
		l.read_posterior(out);
	});
}

int bgt_hip_lattice_upset_mass(bgt_hip_lattice *h, uint32_t experiment, double *out)
{
	return guarded([&] {
		if (out == nullptr)
			throw bgt_hip::Error(BGT_HIP_BAD_INPUT, "null output");
		*out = L(h).upset_mass(experiment);
	});
}

int bgt_hip_lattice_exact_masses(bgt_hip_lattice *h, int count, const uint32_t *experiments, double *out,
								 double *rel_bound)
{
	return guarded([&] {
		if (count < 0 || (count > 0 && (experiments == nullptr || out == nullptr)))
			throw bgt_hip::Error(BGT_HIP_BAD_INPUT, "null experiments or output");
		auto &lat = L(h);
		const int atoms = lat.subjects() * lat.variants();
		const uint32_t full = atoms >= 32 ? 0xFFFFFFFFu : ((1u << atoms) - 1u);
		for (int i = 0; i < count; ++i)
			if (experiments[i] & ~full)
				throw bgt_hip::Error(BGT_HIP_BAD_INPUT, "experiment outside the lattice");
		if ((lat.dilution() || lat.variants() >= 2) && !lat.uses_exact_enum())
			throw bgt_hip::Error(BGT_HIP_UNSUPPORTED, "no exact program for this lattice's history");
		std::unique_ptr<bool[]> ok(new bool[count > 0 ? count : 1]);
		if (!lat.exact_masses(experiments, (size_t)count, out, ok.get()))
			throw bgt_hip::Error(BGT_HIP_UNSUPPORTED, "no exact program for this lattice");
		for (int i = 0; i < count; ++i)
			if (!ok[i])
				throw bgt_hip::Error(BGT_HIP_UNSUPPORTED, "beyond the exact program's budget");
		if (rel_bound)
			*rel_bound = lat.exact_rel_bound();
	});
}

int bgt_hip_lattice_atom_masses(bgt_hip_lattice *h, double *out)
{
	return guarded([&] {
		if (out == nullptr)
			throw bgt_hip::Error(BGT_HIP_BAD_INPUT, "null output");
		L(h).atom_masses(out);
	});
}

int bgt_hip_lattice_classify(bgt_hip_lattice *h, double up, double lo, uint32_t *pos, uint32_t *neg)
{
	return guarded([&] {
		auto &l = L(h);
		l.classify(up, lo);
		if (pos)
			*pos = l.positive_mask();
		if (neg)
			*neg = l.negative_mask();
	});
}

int bgt_hip_lattice_masks(bgt_hip_lattice *h, uint32_t *pos, uint32_t *neg)
{
	return guarded([&] {
		auto &l = L(h);
		if (pos)
			*pos = l.positive_mask();
		if (neg)
			*neg = l.negative_mask();
	});
}

int bgt_hip_exact_upset_mass(int subjects, const double *prior, int updates, const uint32_t *experiments,
							 const uint32_t *responses, int count, const uint32_t *queries, double *out)
{
	return guarded([&] {
		if (subjects <= 0 || subjects > 32 || prior == nullptr || updates < 0 || count < 0 ||
			(updates > 0 && (experiments == nullptr || responses == nullptr)) ||
			(count > 0 && (queries == nullptr || out == nullptr)))
			throw bgt_hip::Error(BGT_HIP_BAD_INPUT, "bad arguments");
		bgt_hip::ExactHistory h;
		h.reset(true, subjects, prior);
		for (int t = 0; t < updates; ++t)
			h.push(experiments[t], responses[t] & 1u);
		std::unique_ptr<bool[]> ok(new bool[count > 0 ? count : 1]);
		if (!h.masses(queries, (size_t)count, out, ok.get()))
			throw bgt_hip::Error(BGT_HIP_UNSUPPORTED, "history exceeds the exact program's state budget");
		for (int i = 0; i < count; ++i)
			if (!ok[i])
				throw bgt_hip::Error(BGT_HIP_UNSUPPORTED, "query exceeds the exact program's state budget");
	});
}

int bgt_hip_exact_upset_mass_dilution(int subjects, const double *prior, const double *dilution, int updates,
									  const uint32_t *experiments, const uint32_t *responses, int count,
									  const uint32_t *queries, double *out)
{
	return guarded([&] {
		if (subjects <= 0 || subjects > 32 || prior == nullptr || dilution == nullptr || updates < 0 || count < 0 ||
			(updates > 0 && (experiments == nullptr || responses == nullptr)) ||
			(count > 0 && (queries == nullptr || out == nullptr)))
			throw bgt_hip::Error(BGT_HIP_BAD_INPUT, "bad arguments");
		bgt_hip::ExactHistory h;
		h.reset(true, subjects, prior, true);
		for (int t = 0; t < updates; ++t)
			h.push(experiments[t], responses[t] & 1u, dilution);
		std::unique_ptr<bool[]> ok(new bool[count > 0 ? count : 1]);
		if (!h.masses(queries, (size_t)count, out, ok.get()))
			throw bgt_hip::Error(BGT_HIP_UNSUPPORTED, "history exceeds the exact program's state budget");
		for (int i = 0; i < count; ++i)
			if (!ok[i])
				throw bgt_hip::Error(BGT_HIP_UNSUPPORTED, "query exceeds the exact program's state budget");
	});
}

int bgt_hip_shrink_plan_make(int original_subjects, int current_subjects, int variants, uint32_t classified_atoms,
							 uint32_t already_classified_subjects, bgt_hip_shrink_plan *out)
{
	return guarded([&] {
		if (out == nullptr)
			throw bgt_hip::Error(BGT_HIP_BAD_INPUT, "null plan");
		bgt_hip::make_shrink_plan(original_subjects, current_subjects, variants, classified_atoms,
								  already_classified_subjects, *out);
	});
}

int bgt_hip_lattice_project(bgt_hip_lattice *h, const bgt_hip_shrink_plan *plan, bgt_hip_lattice **out)
{
	return guarded([&] {
		if (plan == nullptr || out == nullptr)
			throw bgt_hip::Error(BGT_HIP_BAD_INPUT, "null plan or output handle");
		*out = nullptr;
		auto &l = L(h);
		if (plan->current_subjects != l.subjects() || plan->variants != l.variants() ||
			plan->source_atoms != l.atoms())
			throw bgt_hip::Error(BGT_HIP_BAD_INPUT, "the plan is for another lattice shape");
		auto r = std::make_unique<bgt_hip_lattice>();
		r->impl = bgt_hip::DeviceLattice::project(l, plan->removable_current_atoms);
		if (r->impl->subjects() != plan->new_current_subjects)
			throw bgt_hip::Error(BGT_HIP_BAD_INPUT, "the plan's subject count does not match its removable atoms");
		*out = r.release();
	});
}

int bgt_hip_lattice_set_candidates(bgt_hip_lattice *h, uint32_t mask)
{
	return guarded([&] {
		auto &l = L(h);
		const uint32_t full = l.subjects() >= 32 ? 0xFFFFFFFFu : ((1u << l.subjects()) - 1u);
		if (mask & ~full)
			throw bgt_hip::Error(BGT_HIP_BAD_INPUT, "candidate mask has bits beyond the subject count");
		l.set_active(mask);
	});
}

int bgt_hip_lattice_select_experiments(bgt_hip_lattice *h, int steps, uint32_t *out)
{
	return guarded([&] {
		if (out == nullptr && steps > 0)
			throw bgt_hip::Error(BGT_HIP_BAD_INPUT, "null output");
		auto v = L(h).select_experiments(steps);
		for (size_t i = 0; i < v.size(); ++i)
			out[i] = v[i];
	});
}

int bgt_hip_lattice_round(bgt_hip_lattice *h, uint32_t experiment, uint32_t response, const double *dilution,
						  double up, double lo, bgt_hip_selection *next, uint32_t *pos, uint32_t *neg,
						  double *masses_out)
{
	return guarded([&] {
		auto &l = L(h);
		const bgt_hip_selection s = l.round(experiment, response, dilution, up, lo, masses_out);
		if (next)
			*next = s;
		if (pos)
			*pos = l.positive_mask();
		if (neg)
			*neg = l.negative_mask();
	});
}

int bgt_hip_lattice_error_bound(bgt_hip_lattice *h, double *posterior_rel, double *mass_rel)
{
	return guarded([&] {
		double a = 0, b = 0;
		L(h).error_bound(&a, &b);
		if (posterior_rel)
			*posterior_rel = a;
		if (mass_rel)
			*mass_rel = b;
	});
}

int bgt_hip_lattice_kernel_times(bgt_hip_lattice *h, int enable, double *ms_out, int *launches_out, int capacity,
								 char *names, int names_len)
{
	return guarded([&] {
		auto &clk = L(h).clock();
		if (enable >= 0) {
			clk.reset();
			clk.enabled = enable != 0;
			clk.light = enable == 2;
			return;
		}
		clk.resolve();
		std::string joined;
		for (int i = 0; i < (int)clk.entries.size() && i < capacity; ++i) {
			const auto &e = clk.entries[i];
			if (ms_out)
				ms_out[i] = e.launches ? e.total_ms / e.launches : 0.0;
			if (launches_out)
				launches_out[i] = e.launches;
			if (i)
				joined += ";";
			joined += e.name;
		}
		if (names && names_len > 0) {
			std::strncpy(names, joined.c_str(), (size_t)names_len - 1);
			names[names_len - 1] = 0;
		}
	});
}

int bgt_hip_lattice_info(bgt_hip_lattice *h, int *subjects, int *variants, int *precision, uint64_t *device_bytes)
{
	return guarded([&] {
		auto &l = L(h);
		if (subjects)
			*subjects = l.subjects();
		if (variants)
			*variants = l.variants();
		if (precision)
			*precision = l.storage() == bgt_hip::Storage::fp64 ? BGT_HIP_FP64 : BGT_HIP_FP32C;
		if (device_bytes)
			*device_bytes = l.device_bytes();
	});
}

int bgt_hip_lattice_footprint(int use_dilution, int subjects, int variants, int precision, uint64_t *bytes)
{
	return guarded([&] {
		if (bytes == nullptr || subjects <= 0 || variants <= 0)
			throw bgt_hip::Error(BGT_HIP_BAD_INPUT, "bad arguments");
		if (variants > bgt_hip::kMaxVariants || subjects * variants > bgt_hip::kMaxAtoms)
			throw bgt_hip::Error(BGT_HIP_UNSUPPORTED, "lattice larger than 2^30 states or more than 5 variants");
		*bytes = bgt_hip::DeviceLattice::footprint(use_dilution != 0, subjects, variants, storage_of(precision));
	});
}

static void run_batch(bgt_hip_lattice **lattices, const uint32_t *experiments, const uint32_t *responses,
					  const double *const *dilutions, int count, int selector, bool classify, double up, double lo,
					  bgt_hip_selection *out, uint32_t *pos, uint32_t *neg)
{
	if (count < 0 || (count > 0 && (lattices == nullptr || responses == nullptr || out == nullptr)))
		throw bgt_hip::Error(BGT_HIP_BAD_INPUT, "null argument");
	std::vector<bgt_hip::BatchItem> items(count);
	for (int i = 0; i < count; ++i) {
		auto &it = items[i];
		it.lat = &L(lattices[i]);
		it.update = responses[i] != BGT_HIP_NO_UPDATE;
		if (it.update && experiments == nullptr)
			throw bgt_hip::Error(BGT_HIP_BAD_INPUT, "experiments are required for updates");
		it.experiment = it.update ? experiments[i] : 0u;
		it.response = it.update ? responses[i] : 0u;
		it.dilution = dilutions ? dilutions[i] : nullptr;
		it.classify = classify;
		it.select = true;
	}
	std::vector<bgt_hip::BatchResult> res;
	bgt_hip::batch_step(items, selector, up, lo, res);
	for (int i = 0; i < count; ++i) {
		out[i] = res[i].selection;
		if (pos)
			pos[i] = res[i].positive;
		if (neg)
			neg[i] = res[i].negative;
	}
}

int bgt_hip_batch_step(bgt_hip_lattice **lattices, const uint32_t *experiments, const uint32_t *responses,
					   const double *const *dilutions, int count, int selector, bgt_hip_selection *out)
{
	return guarded([&] {
		run_batch(lattices, experiments, responses, dilutions, count, selector, false, 0.0, 0.0, out, nullptr,
				  nullptr);
	});
}

int bgt_hip_batch_round(bgt_hip_lattice **lattices, const uint32_t *experiments, const uint32_t *responses,
						const double *const *dilutions, int count, int selector, double threshold_up,
						double threshold_lo, bgt_hip_selection *out, uint32_t *positive, uint32_t *negative)
{
	return guarded([&] {
		run_batch(lattices, experiments, responses, dilutions, count, selector, true, threshold_up, threshold_lo,
				  out, positive, negative);
	});
}

int bgt_hip_batch_graphs(int enable, int *previous)
{
	return guarded([&] {
		const bool was = bgt_hip::set_batch_graphs(enable != 0);
		if (previous)
			*previous = was ? 1 : 0;
	});
}

int bgt_hip_batch_graph_counts(unsigned long long *captures, unsigned long long *replays)
{
	return guarded([&] { bgt_hip::batch_graph_counts(captures, replays); });
}

int bgt_hip_tree_shrink(int enable, int *previous)
{
	return guarded([&] {
		const bool was = bgt_hip::set_tree_shrink(enable != 0);
		if (previous)
			*previous = was ? 1 : 0;
	});
}

namespace {

bgt_hip::TuneResult run_tune(int use_dilution, int subjects, int variants, int precision, int selector, int force,
							 char *spec_id, int spec_len)
{
	if (subjects <= 0 || variants <= 0 || subjects * variants > bgt_hip::kMaxAtoms)
		throw bgt_hip::Error(BGT_HIP_BAD_INPUT, "tune: shape out of range");
	bgt_hip::require_device();
	bgt_hip::TuneKey key;
	key.dilution = use_dilution != 0;
	key.subjects = subjects;
	key.variants = variants;
	key.selector = selector == BGT_HIP_SELECT_AUTO ? (variants == 1 ? BGT_HIP_SELECT_OP_BHA : BGT_HIP_SELECT_OP_BBPA)
												   : selector;
	key.storage = storage_of(precision);
	const bgt_hip::TuneResult r = bgt_hip::tune(key, force != 0);
	if (spec_id && spec_len > 0) {
		std::strncpy(spec_id, r.id.c_str(), (size_t)spec_len - 1);
		spec_id[spec_len - 1] = '\0';
	}
	return r;
}

} // namespace

int bgt_hip_tune(int use_dilution, int subjects, int variants, int precision, int selector, int force, int *form,
				 double *ms_two_pass, double *ms_three_pass, char *spec_id, int spec_len)
{
	return guarded([&] {
		const bgt_hip::TuneResult r = run_tune(use_dilution, subjects, variants, precision, selector, force, spec_id,
											   spec_len);
		if (form)
			*form = r.form;
		// each form with its faster pass-1 grid
		auto best = [&](int f2) {
			const double a = r.ms[f2][0], b = r.ms[f2][1];
			return (a > 0.0 && (b <= 0.0 || a <= b)) ? a : b;
		};
		if (ms_two_pass)
			*ms_two_pass = best(0);
		if (ms_three_pass)
			*ms_three_pass = best(1);
	});
}

int bgt_hip_tune_forms(int use_dilution, int subjects, int variants, int precision, int selector, int force,
					   int *pass2_form, int *pass1_form, double *ms, char *spec_id, int spec_len)
{
	return guarded([&] {
		const bgt_hip::TuneResult r = run_tune(use_dilution, subjects, variants, precision, selector, force, spec_id,
											   spec_len);
		if (pass2_form)
			*pass2_form = r.form;
		if (pass1_form)
			*pass1_form = r.pass1;
		if (ms)
			for (int f2 = 0; f2 < bgt_hip::kPass2Forms; ++f2)
				for (int f1 = 0; f1 < bgt_hip::kPass1Forms; ++f1)
					ms[f2 * bgt_hip::kPass1Forms + f1] = r.ms[f2][f1];
	});
}

int bgt_hip_tune_force(int form)
{
	return guarded([&] {
		if (form < -1 || form >= bgt_hip::kPass2Forms)
			throw bgt_hip::Error(BGT_HIP_BAD_INPUT, "tune_force: form is -1, 0, 1 or 2");
		bgt_hip::set_pass2_form(form);
	});
}

int bgt_hip_tune_force_pass1(int form)
{
	return guarded([&] {
		if (form < -1 || form >= bgt_hip::kPass1Forms)
			throw bgt_hip::Error(BGT_HIP_BAD_INPUT, "tune_force_pass1: form is -1, 0 or 1");
		bgt_hip::set_pass1_form(form);
	});
}

int bgt_hip_select_noalloc(const void *posterior, int subjects, int precision, int selector, double scale,
						   bgt_hip_selection *result, double *masses, void *stream)
{
	return guarded([&] {
		bgt_hip::select_noalloc(posterior, subjects, precision, selector, scale, result, masses,
								static_cast<hipStream_t>(stream));
	});
}

} // extern "C"
