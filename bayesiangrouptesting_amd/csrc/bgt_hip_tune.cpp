// Autotuner of the transform's launch forms, keyed by the reference's KernelSpec.
//
// The reference keys compiled kernels by KernelSpec (include/bgt/types.hpp:132-181: provider,
// subjects, variants, state bits, dilution, selector, precisions, architecture) and keeps
// one text artefact per key in a cache directory (src/runtime/jit_compiler.cpp:13-110,
// "kernel-<hash>-p..-n..-v..." names). This provider compiles every kernel ahead of time;
// what depends on the shape is the choice among launch forms of the same result, which is
// measured once per key on the device and kept in memory and in a cache file of the same
// naming scheme, so a later process starts from the measured choice.
//
// Tuned, for FP32 lattices of 26..30 atoms: the form of the high-bit transform (two passes
// with the split pass 2 on the blocked z, reading 32-byte pieces; three passes reading whole
// 128-byte lines; two passes with the plain pass 2 on the row-form z, reading 2^(30-A)-entry
// pieces) and the grid of pass 1 (persistent with a prefetch of the next tile, or one
// workgroup per tile); see the kernels. Everything else is fixed by the shape.
#include "bgt_hip_tune.hpp"

#include "bgt_hip_lattice.hpp"

#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <functional>
#include <map>
#include <mutex>
#include <sstream>
#include <vector>

namespace bgt_hip {

namespace {

std::mutex g_mu;
int g_forced = -1;  // set_pass2_form: -1 none
int g_forced1 = -1; // set_pass1_form: -1 none
struct Choice {
	int pass2 = 0, pass1 = 0;
};
std::map<std::string, Choice> g_choices; // spec id -> forms

// Provider value of this provider in the reference's enum once a `hip` value is added after
// auto_select, cpu, cuda (types.hpp:96-101); precision values of Precision (float32 = 0,
// float64 = 1) plus float32_compensated = 2.
constexpr int kProviderHip = 3;
int precision_value(Storage st) { return st == Storage::fp64 ? 1 : 2; }

std::string device_arch()
{
	int dev = 0;
	hipDeviceProp_t p;
	if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess)
		return "gfx950";
	std::string a = p.gcnArchName;
	const size_t colon = a.find(':');
	return colon == std::string::npos ? a : a.substr(0, colon);
}

std::filesystem::path cache_dir()
{
	if (const char *d = std::getenv("BGT_HIP_TUNE_DIR"); d != nullptr && *d)
		return d;
	return std::filesystem::temp_directory_path() / "bayesian_group_testing_hip_tune";
}

// KernelSpecHash's combination (types.hpp:160-178) over the same fields; the compile
// options of the reference's JIT do not apply (kernels are compiled ahead of time)
size_t spec_hash(const TuneKey &k)
{
	size_t seed = 0;
	auto combine = [&seed](size_t h) { seed ^= h + 0x9e3779b9 + (seed << 6) + (seed >> 2); };
	combine(std::hash<int>{}(kProviderHip));
	combine(std::hash<int>{}(k.subjects));
	combine(std::hash<int>{}(k.variants));
	combine(std::hash<int>{}(32));
	combine(std::hash<int>{}(k.dilution ? 1 : 0));
	combine(std::hash<int>{}(k.selector));
	combine(std::hash<int>{}(precision_value(k.storage)));
	combine(std::hash<int>{}(1));
	combine(std::hash<std::string>{}(k.arch));
	return seed;
}

bool applies(int atoms, Storage st) { return st == Storage::fp32c && atoms >= 26 && atoms <= 30; }

bool read_choice(const std::filesystem::path &f, Choice &c)
{
	std::ifstream in(f);
	std::string line;
	int p2 = -1, p1 = -1;
	while (std::getline(in, line)) {
		if (line.rfind("pass2_form=", 0) == 0)
			p2 = std::atoi(line.c_str() + 11);
		else if (line.rfind("pass1_form=", 0) == 0)
			p1 = std::atoi(line.c_str() + 11);
	}
	if (p2 < 0 || p2 >= kPass2Forms || p1 < 0 || p1 >= kPass1Forms)
		return false;
	c.pass2 = p2;
	c.pass1 = p1;
	return true;
}

// the remembered choice of any tuned key of this atom count and storage (the transform does
// not depend on dilution or selector)
const Choice *choice_for(int atoms, Storage st)
{
	for (const auto &kv : g_choices)
		if (kv.first.find("-n" + std::to_string(atoms) + "-v1-") != std::string::npos &&
			kv.first.find("-r" + std::to_string(precision_value(st)) + "-") != std::string::npos)
			return &kv.second;
	return nullptr;
}

} // namespace

std::string spec_id(const TuneKey &k)
{
	// kernel_spec_id (jit_compiler.cpp:13-38): hash, then the fields
	std::ostringstream out;
	out << "kernel-" << spec_hash(k) << "-p" << kProviderHip << "-n" << k.subjects << "-v" << k.variants << "-s32"
		<< "-d" << (k.dilution ? 1 : 0) << "-q" << k.selector << "-r" << precision_value(k.storage) << "-a1-"
		<< k.arch;
	return out.str();
}

int tuned_pass2_form(int atoms, Storage st)
{
	if (!applies(atoms, st))
		return 0;
	std::lock_guard<std::mutex> lock(g_mu);
	int f = g_forced;
	if (f < 0) {
		const Choice *c = choice_for(atoms, st);
		f = c ? c->pass2 : -1;
	}
	return (f == 2 && atoms < 28) ? 0 : f; // below 28 atoms the plain pass 2 is form 0
}

int tuned_pass1_form(int atoms, Storage st)
{
	if (!applies(atoms, st))
		return 0;
	std::lock_guard<std::mutex> lock(g_mu);
	if (g_forced1 >= 0)
		return g_forced1;
	const Choice *c = choice_for(atoms, st);
	return c ? c->pass1 : -1;
}

void set_pass2_form(int form)
{
	std::lock_guard<std::mutex> lock(g_mu);
	g_forced = form;
}

void set_pass1_form(int form)
{
	std::lock_guard<std::mutex> lock(g_mu);
	g_forced1 = form;
}

TuneResult tune(const TuneKey &key_in, bool force)
{
	TuneKey key = key_in;
	key.arch = device_arch();
	TuneResult r;
	r.id = spec_id(key);
	if (key.variants != 1 || !applies(key.subjects, key.storage)) {
		r.form = 0;
		r.pass1 = 0;
		r.from_cache = false;
		return r;
	}
	const std::filesystem::path file = cache_dir() / (r.id + ".bgthip");
	if (!force) {
		std::lock_guard<std::mutex> lock(g_mu);
		auto it = g_choices.find(r.id);
		if (it != g_choices.end()) {
			r.form = it->second.pass2;
			r.pass1 = it->second.pass1;
			r.from_cache = true;
			return r;
		}
		std::error_code ec;
		Choice c;
		if (std::filesystem::exists(file, ec) && read_choice(file, c)) {
			g_choices[r.id] = c;
			r.form = c.pass2;
			r.pass1 = c.pass1;
			r.from_cache = true;
			return r;
		}
	}
	// measure: fused rounds (update, classification, next selection: the work of a session's
	// test) on a lattice of this shape under each pair of forms; the band work is common to
	// all, pass 1's grid and the high-bit transform differ. The thresholds classify nothing,
	// the responses alternate, each round tests the previous round's pick.
	const int n = key.subjects;
	std::vector<double> prior(n);
	for (int a = 0; a < n; ++a)
		prior[a] = 0.005 + 0.045 * (a + 0.5) / n;
	std::vector<double> dil;
	if (key.dilution) {
		// DilutionTable(n, 0.05, 0.9) (dilution.cpp:44-67): any valid table times the same
		dil.assign((size_t)n * (n + 1), 0.0);
		for (int s = 1; s <= n; ++s) {
			dil[(size_t)(s - 1) * (n + 1)] = 0.05;
			for (int m = 1; m <= s; ++m)
				dil[(size_t)(s - 1) * (n + 1) + m] = 1.0 - 0.05 * m / ((s - m) * 0.9 + m);
		}
	}
	auto lat = DeviceLattice::make(key.dilution, n, 1, prior.data(), key.storage);
	uint32_t e = lat->select(key.selector).experiment;
	int saved2, saved1;
	{
		std::lock_guard<std::mutex> lock(g_mu);
		saved2 = g_forced;
		saved1 = g_forced1;
	}
	double best = 1e300, best0 = 1e300;
	int pass1_0 = 0;
	for (int f2 = 0; f2 < kPass2Forms; ++f2) {
		if (f2 == 2 && n < 28)
			continue; // the plain pass 2 is form 0 there
		for (int f1 = 0; f1 < kPass1Forms; ++f1) {
			set_pass2_form(f2);
			set_pass1_form(f1);
			double mn = 1e30;
			for (int i = 0; i < 9; ++i) {
				const auto t0 = std::chrono::steady_clock::now();
				const bgt_hip_selection s = lat->round(e, (uint32_t)(i & 1), dil.empty() ? nullptr : dil.data(), 0.0,
													   0.0, nullptr);
				const double ms =
					std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
				e = (s.experiment == 0u || s.experiment == 0xFFFFFFFFu) ? 1u : s.experiment;
				if (i >= 2)
					mn = std::min(mn, ms);
			}
			r.ms[f2][f1] = mn;
			if (mn < best) {
				best = mn;
				r.form = f2;
				r.pass1 = f1;
			}
			if (f2 == 0 && mn < best0) {
				best0 = mn;
				pass1_0 = f1;
			}
		}
	}
	// form 0 (the split pass 2 where it applies) stays unless another is faster by more than
	// the measurement's spread (3 %): near-equal forms would otherwise be picked by noise
	if (r.form != 0 && best0 <= 1.03 * best) {
		r.form = 0;
		r.pass1 = pass1_0;
	}
	set_pass2_form(saved2);
	set_pass1_form(saved1);
	DeviceLattice::recycle(std::move(lat));
	r.from_cache = false;
	{
		std::lock_guard<std::mutex> lock(g_mu);
		g_choices[r.id] = Choice{r.form, r.pass1};
	}
	std::error_code ec;
	std::filesystem::create_directories(cache_dir(), ec);
	std::ofstream out(file, std::ios::trunc);
	if (out) {
		out << "provider=" << kProviderHip << '\n'
			<< "subjects=" << key.subjects << '\n'
			<< "variants=" << key.variants << '\n'
			<< "state_bits=32\n"
			<< "dilution=" << (key.dilution ? 1 : 0) << '\n'
			<< "selector=" << key.selector << '\n'
			<< "precision=" << precision_value(key.storage) << '\n'
			<< "gpu_arch=" << key.arch << '\n'
			<< "pass2_form=" << r.form << '\n'
			<< "pass1_form=" << r.pass1 << '\n';
		for (int f2 = 0; f2 < kPass2Forms; ++f2)
			for (int f1 = 0; f1 < kPass1Forms; ++f1)
				out << "select_ms_" << f2 << "_" << f1 << "=" << r.ms[f2][f1] << '\n';
	}
	return r;
}

} // namespace bgt_hip
