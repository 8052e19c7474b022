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
// Tuned: the form of the high-bit transform for FP32 lattices of 26..30 atoms (two passes,
// the second reading 2^(30-A)- or 2^(31-A)-entry pieces, or three passes reading whole
// 128-byte lines; see the kernels). Everything else is fixed by the shape.
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
int g_forced = -1;                    // set_pass2_form: -1 none
std::map<std::string, int> g_choices; // spec id -> pass-2 form

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

int read_choice(const std::filesystem::path &f)
{
	std::ifstream in(f);
	std::string line;
	while (std::getline(in, line))
		if (line.rfind("pass2_form=", 0) == 0)
			return std::atoi(line.c_str() + 11);
	return -1;
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
	if (g_forced >= 0)
		return g_forced;
	// the transform does not depend on dilution or selector: any tuned key of this atom
	// count and storage decides
	for (const auto &kv : g_choices)
		if (kv.first.find("-n" + std::to_string(atoms) + "-v1-") != std::string::npos &&
			kv.first.find("-r" + std::to_string(precision_value(st)) + "-") != std::string::npos)
			return kv.second;
	return -1;
}

void set_pass2_form(int form)
{
	std::lock_guard<std::mutex> lock(g_mu);
	g_forced = form;
}

TuneResult tune(const TuneKey &key_in, bool force)
{
	TuneKey key = key_in;
	key.arch = device_arch();
	TuneResult r;
	r.id = spec_id(key);
	if (key.variants != 1 || !applies(key.subjects, key.storage)) {
		r.form = 0;
		r.from_cache = false;
		return r;
	}
	const std::filesystem::path file = cache_dir() / (r.id + ".bgthip");
	if (!force) {
		std::lock_guard<std::mutex> lock(g_mu);
		auto it = g_choices.find(r.id);
		if (it != g_choices.end()) {
			r.form = it->second;
			r.from_cache = true;
			return r;
		}
		std::error_code ec;
		if (std::filesystem::exists(file, ec)) {
			const int c = read_choice(file);
			if (c == 0 || c == 1) {
				g_choices[r.id] = c;
				r.form = c;
				r.from_cache = true;
				return r;
			}
		}
	}
	// measure: the selection of a prior lattice of this shape under each form (pass 1 and
	// the band work are common to both; only the high-bit transform differs)
	const int n = key.subjects;
	std::vector<double> prior(n);
	for (int a = 0; a < n; ++a)
		prior[a] = 0.005 + 0.045 * (a + 0.5) / n;
	auto lat = DeviceLattice::make(key.dilution, n, 1, prior.data(), key.storage);
	int saved;
	{
		std::lock_guard<std::mutex> lock(g_mu);
		saved = g_forced;
	}
	double best[2] = {0.0, 0.0};
	for (int form = 0; form < 2; ++form) {
		set_pass2_form(form);
		double mn = 1e30;
		for (int i = 0; i < 7; ++i) {
			const auto t0 = std::chrono::steady_clock::now();
			lat->select(key.selector);
			const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
			if (i >= 2)
				mn = std::min(mn, ms);
		}
		best[form] = mn;
	}
	set_pass2_form(saved);
	DeviceLattice::recycle(std::move(lat));
	r.form = best[1] < best[0] ? 1 : 0;
	r.ms[0] = best[0];
	r.ms[1] = best[1];
	r.from_cache = false;
	{
		std::lock_guard<std::mutex> lock(g_mu);
		g_choices[r.id] = r.form;
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
			<< "select_ms_two_pass=" << best[0] << '\n'
			<< "select_ms_three_pass=" << best[1] << '\n';
	}
	return r;
}

} // namespace bgt_hip
