// Selection on a caller-owned posterior without any device allocation.
//
// Everything the selection kernels need beyond the caller's input and output (scalars, the
// selection result, per-tile partial sums and top-2 lists, the atom singletons and, above
// 2^15 states, the spilled transform intermediate) lives in `scratch`, a file-scope device
// array of kNoAllocScratchBytes that every GPU the library runs on holds until the process
// ends. Host code never passes the array's name to a kernel: its device address is looked
// up once per device with hipGetSymbolAddress and kept per device. Every call on a GPU
// shares that GPU's array, so calls that overlap on one GPU race; the array keeps the last
// call's contents, and each call first rewrites what must start from a known value.
#include <hip/hip_runtime.h>

#include "bgt_hip.h"
#include "bgt_hip_internal.hpp"
#include "bgt_hip_noalloc.hpp"

#include <cmath>
#include <mutex>

namespace bgt_hip {

namespace {

// layout of the scratch array (byte offsets), sized for kNoAllocMaxSubjects
constexpr size_t kNaScal = 0;                                     // kScalars doubles
constexpr size_t kNaResult = 256;                                 // SelResult
constexpr size_t kNaCounter = 512;                                // 32 unsigned
constexpr size_t kNaMasses = 768;                                 // 32 doubles
constexpr size_t kNaPart = 1024;                                  // one double per pass-1 tile
constexpr size_t kNaTiles = size_t(1) << (kNoAllocMaxSubjects - kTileBits);
constexpr size_t kNaTops = kNaPart + 256 * ((kNaTiles * sizeof(double) + 255) / 256); // one SelResult per workgroup
constexpr size_t kNaZ = 4096;                                     // transform intermediate
static_assert(kNaTops + kNaTiles * sizeof(SelResult) <= kNaZ, "scratch header overlaps the intermediate");
constexpr size_t kNaZBytes = (size_t(1) << (kNoAllocMaxSubjects - kTileBits)) *
							 ((size_t(1) << kTileBits) + kZSkewBytes / sizeof(double)) * sizeof(double);
static_assert(kNaZ + kNaZBytes <= kNoAllocScratchBytes, "scratch array too small for its largest lattice");

__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};

// The device address of this device's copy of `scratch`, looked up on the first call with
// the device current.
unsigned char *scratch_ptr()
{
	int dev = 0;
	check_hip(hipGetDevice(&dev), "hipGetDevice");
	if (dev < 0 || dev >= kMaxDevices)
		throw Error(BGT_HIP_UNSUPPORTED, "device index beyond the scratch table");
	std::lock_guard<std::mutex> lock(g_mu);
	if (g_scratch[dev] == nullptr) {
		void *p = nullptr;
		check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
		if (p == nullptr)
			throw Error(BGT_HIP_INTERNAL, "hipGetSymbolAddress returned no address for the scratch array");
		g_scratch[dev] = p;
	}
	return static_cast<unsigned char *>(g_scratch[dev]);
}

// The scale every posterior entry is read with (the kernels' pending 1/D slot) and the
// counters the transform kernels expect cleared; a kernel rather than a memset, so that
// the call stays capturable.
__global__ void k_noalloc_init(double *scal, unsigned *counter, double scale)
{
	const unsigned t = threadIdx.x;
	if (t < kScalars)
		scal[t] = t == kScalInv0 ? scale : 0.0;
	if (t < 32)
		counter[t] = 0u;
}

// The selection record and the atom singletons into the caller's device memory. Op-BHA
// returns 0 and BBPA invalid_state when nothing beats the initial best (bbpa.cpp:539,
// lattice.hpp:250), as bgt_hip_lattice_select does.
__global__ void k_noalloc_result(const SelResult *res, const double *masses, int atoms, int bbpa, double gap_bound,
								 bgt_hip_selection *out, double *masses_out)
{
	const int t = threadIdx.x;
	if (t == 0) {
		const SelResult r = *res;
		const bool found = r.e1 != 0xFFFFFFFFu && r.d1 < 2.0;
		bgt_hip_selection s;
		s.experiment = found ? r.e1 : (bbpa ? 0xFFFFFFFFu : 0u);
		s.runner_up = r.e2;
		s.objective = r.d1;
		s.runner_up_objective = r.d2;
		s.gap_bound = gap_bound;
		s.near_tie = (r.d2 - r.d1) <= gap_bound ? 1 : 0;
		s.redecided = 0;
		s.candidates = 0;
		s.band = 0;
		s.collected = 0;
		*out = s;
	}
	if (masses_out != nullptr && t < atoms)
		masses_out[t] = masses[t];
}

} // namespace

void select_noalloc(const void *posterior, int subjects, int precision, int selector, double scale,
					bgt_hip_selection *result, double *masses, hipStream_t stream)
{
	if (posterior == nullptr || result == nullptr)
		throw Error(BGT_HIP_BAD_INPUT, "select_noalloc: posterior and result are required");
	if (subjects < 1)
		throw Error(BGT_HIP_BAD_INPUT, "select_noalloc: at least one subject");
	if (subjects > kNoAllocMaxSubjects)
		throw Error(BGT_HIP_UNSUPPORTED, "select_noalloc: at most 20 subjects (2^20 states) fit the static scratch array");
	if (precision != BGT_HIP_FP64 && precision != BGT_HIP_FP32C)
		throw Error(BGT_HIP_BAD_INPUT, "select_noalloc: unknown precision");
	if (selector != BGT_HIP_SELECT_AUTO && selector != BGT_HIP_SELECT_OP_BHA && selector != BGT_HIP_SELECT_OP_BBPA)
		throw Error(BGT_HIP_BAD_INPUT, "select_noalloc: Op-BHA or Op-BBPA only");
	if (!(scale > 0.0) || !std::isfinite(scale))
		throw Error(BGT_HIP_BAD_INPUT, "select_noalloc: scale must be positive and finite");
	unsigned char *base = scratch_ptr();
	const Storage st = precision == BGT_HIP_FP64 ? Storage::fp64 : Storage::fp32c;
	const bool bbpa = selector == BGT_HIP_SELECT_OP_BBPA;
	const int objective = bbpa ? kObjBinaryBbpa : kObjHalving;
	const double target = 0.5;
	const int atoms = subjects;

	Descs ds{};
	LatDesc &d = ds.one;
	d.post = const_cast<void *>(posterior);
	d.z = base + kNaZ;
	d.scal = reinterpret_cast<double *>(base + kNaScal);
	d.part = reinterpret_cast<double *>(base + kNaPart);
	d.tops = reinterpret_cast<SelResult *>(base + kNaTops);
	d.result = reinterpret_cast<SelResult *>(base + kNaResult);
	d.counter = reinterpret_cast<unsigned *>(base + kNaCounter);
	d.masses = reinterpret_cast<double *>(base + kNaMasses);
	d.n = subjects;
	d.k = 1;
	d.atoms = atoms;
	d.active = subjects >= 32 ? 0xFFFFFFFFu : ((1u << subjects) - 1u);
	d.inv_in = kScalInv0;
	d.inv_out = kScalInv1;
	ds.many = nullptr;

	hipLaunchKernelGGL(k_noalloc_init, dim3(1), dim3(64), 0, stream, d.scal, d.counter, scale);
	check_hip(hipGetLastError(), "k_noalloc_init");
	if (atoms <= kSmallBits) {
		launch_zeta_small(ds, 1, atoms, 1, st, objective, target, stream);
	} else if (atoms == kTileBits) {
		launch_zeta_final(ds, 1, atoms, 1, st, objective, true, false, target, stream);
	} else {
		launch_zeta_pass1(ds, 1, atoms, st, false, 1, 0, objective, stream);
		launch_zeta_final(ds, 1, atoms, 1, st, objective, false, false, target, stream);
	}
	// the transform's own rounding (FP64 sums along chains of at most `atoms` additions) on
	// one objective, twice for BBPA's two masses, for both candidates of the top two
	const double gap_bound = 2.0 * (bbpa ? 2.0 : 1.0) * (atoms + 1) * std::ldexp(1.0, -53);
	hipLaunchKernelGGL(k_noalloc_result, dim3(1), dim3(64), 0, stream, d.result, d.masses, atoms, bbpa ? 1 : 0,
					   gap_bound, result, masses);
	check_hip(hipGetLastError(), "k_noalloc_result");
}

} // namespace bgt_hip
