// Host orchestration of one device-resident lattice.
#include "bgt_hip_lattice.hpp"

#include "bgt_hip_resources.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>

namespace bgt_hip {

void check_hip(hipError_t e, const char *what)
{
	if (e != hipSuccess)
		throw Error(BGT_HIP_INTERNAL, std::string(what) + ": " + hipGetErrorString(e));
}

int device_count()
{
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess)
		return 0;
	return n;
}

void require_device()
{
	static thread_local int checked = -1;
	if (checked == 1)
		return;
	if (device_count() <= 0)
		throw Error(BGT_HIP_UNSUPPORTED, "no HIP device is visible");
	hipDeviceProp_t prop;
	BGT_HIP_CHECK(hipGetDeviceProperties(&prop, 0));
	if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
		throw Error(BGT_HIP_UNSUPPORTED, std::string("this provider is built for gfx950, found ") + prop.gcnArchName);
	checked = 1;
}

// ------------------------------------------------------------------------ kernel clock

KernelClock::~KernelClock()
{
	for (auto &p : pending_) {
		(void)hipEventDestroy(p.a);
		(void)hipEventDestroy(p.b);
	}
	for (auto e : pool_)
		(void)hipEventDestroy(e);
	if (open_)
		(void)hipEventDestroy(open_);
}

hipEvent_t KernelClock::take()
{
	if (!pool_.empty()) {
		hipEvent_t e = pool_.back();
		pool_.pop_back();
		return e;
	}
	// timing-only events: no system-scope fence (which would write back and invalidate the
	// L2 of every XCD at each record)
	hipEvent_t e;
	BGT_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
	return e;
}

void KernelClock::begin(hipStream_t s, bool essential)
{
	if (!enabled || (light && !essential)) {
		open_ = nullptr;
		return;
	}
	if (pending_.size() >= 4096)
		resolve();
	open_ = take();
	BGT_HIP_CHECK(hipEventRecord(open_, s));
}

void KernelClock::end(hipStream_t s, const char *name)
{
	if (!enabled || !open_)
		return;
	hipEvent_t b = take();
	BGT_HIP_CHECK(hipEventRecord(b, s));
	pending_.push_back({name, open_, b});
	open_ = nullptr;
}

void KernelClock::host(const char *name, double ms)
{
	if (!enabled)
		return;
	for (auto &e : entries)
		if (e.name == name) {
			e.total_ms += ms;
			e.launches += 1;
			return;
		}
	entries.push_back({name, ms, 1});
}

void KernelClock::resolve()
{
	if (pending_.empty())
		return;
	BGT_HIP_CHECK(hipEventSynchronize(pending_.back().b));
	for (auto &p : pending_) {
		float ms = 0.0f;
		BGT_HIP_CHECK(hipEventElapsedTime(&ms, p.a, p.b));
		bool found = false;
		for (auto &e : entries)
			if (e.name == p.name) {
				e.total_ms += ms;
				e.launches += 1;
				found = true;
				break;
			}
		if (!found)
			entries.push_back({p.name, ms, 1});
		pool_.push_back(p.a);
		pool_.push_back(p.b);
	}
	pending_.clear();
}

void KernelClock::reset()
{
	resolve();
	entries.clear();
}

namespace {


size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Arena head: [0,256) scalars, [256,512) selection result, [512,768) counters,
// [768,1024) atom masses. Lattices with exact rescoring continue with the device's exact
// masses of the band (kExactBand doubles) and then the band candidates; the others have
// neither. The head copy of a lattice brings back this 1 KB block with the first band
// candidates: kHeadBytes (band at 1 KB: the batched heads gather it there) or, for a single
// lattice with exact rescoring, kHeadBytesWide (exact masses at kHeadExactWide, band at
// kHeadBandWide).
constexpr size_t kHeadResult = 256;
constexpr size_t kHeadCounter = 512;
constexpr size_t kHeadMasses = 768;
constexpr size_t kHeadBand = 1024;

struct HostBox {
	SelResult result;
	unsigned counter[4];
	double masses[32];
	double scal[kScalars];
	double value;
	double raw;
	ExactRec hist[kExactMaxHist];   // staging of the update history for the device
	double exact_u[kExactBand];     // exact masses of band candidates from the device
	alignas(256) unsigned char head[kHeadBytesWide];
};

} // namespace

// ---------------------------------------------------------------------- construction

DeviceLattice::DeviceLattice(bool dilution, int subjects, int variants, const double *prior, Storage storage)
	: dilution_(dilution), n_(subjects), k_(variants), atoms_(subjects * variants), storage_(storage)
{
	if (subjects <= 0 || variants <= 0)
		throw Error(BGT_HIP_BAD_INPUT, "subjects and variants must be positive");
	if (variants > kMaxVariants || subjects > kMaxSubjects || atoms_ > kMaxAtoms)
		throw Error(BGT_HIP_UNSUPPORTED, "lattice larger than 2^30 states or more than 5 variants");
	if (prior == nullptr)
		throw Error(BGT_HIP_BAD_INPUT, "prior is required");
	require_device();
	states_ = 1ull << atoms_;
	stream_ = stream_acquire();
	allocate();
	init_prior(prior);
}

DeviceLattice::DeviceLattice(bool dilution, int subjects, int variants, Storage storage, Blank)
	: dilution_(dilution), n_(subjects), k_(variants), atoms_(subjects * variants), storage_(storage)
{
	if (subjects <= 0 || variants <= 0)
		throw Error(BGT_HIP_BAD_INPUT, "subjects and variants must be positive");
	if (variants > kMaxVariants || subjects > kMaxSubjects || atoms_ > kMaxAtoms)
		throw Error(BGT_HIP_UNSUPPORTED, "lattice larger than 2^30 states or more than 5 variants");
	require_device();
	states_ = 1ull << atoms_;
	stream_ = stream_acquire();
	allocate();
}

void DeviceLattice::init_prior(const double *prior)
{
	pos_ = neg_ = 0;
	active_ = (n_ >= 32) ? 0xFFFFFFFFu : ((1u << n_) - 1u);
	updates_ = 0;
	hist_uploaded_ = 0;
	exact_pending_ = exact_valid_ = false;
	inv_cur_ = kScalInv0;
	e_ = r_ = 0;
	len_ = 0;
	clock_.reset();
	clock_.enabled = false;
	exact_.reset(true, atoms_, prior, dilution_, k_);
	proj_rel_ = 0.0;
	xn_ = n_;
	xkeep_ = (n_ >= 32) ? 0xFFFFFFFFu : ((1u << n_) - 1u);
	gh_e_.clear();
	gh_r_.clear();
	gh_dil_.clear();
	gh_broken_ = false;
	gh_uploaded_ = false;
	pairs_own_ = true;
	std::vector<double> pairs(2 * kMaxAtoms, 0.0);
	for (int a = 0; a < atoms_; ++a) {
		pairs[2 * a] = prior[a];
		pairs[2 * a + 1] = 1.0 - prior[a];
	}
	const double init[kScalars] = {1.0, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, 0.0};
	BGT_HIP_CHECK(hipMemcpyAsync(scal_, init, sizeof(init), hipMemcpyHostToDevice, stream_));
	BGT_HIP_CHECK(hipMemcpyAsync(pairs_, pairs.data(), sizeof(double) * pairs.size(), hipMemcpyHostToDevice, stream_));
	// the histogram bins start at zero (each use clears them after reading; a lattice taken
	// back from the cache after a failed call may not have)
	if (fixtab_)
		BGT_HIP_CHECK(hipMemsetAsync(fixtab_, 0, fix_bytes_, stream_));
	launch_prior(single(), 1, atoms_, storage_, pairs_, states_, stream_);
	sync();
}

// ------------------------------------------------------------------- lattice cache

namespace {

// Released lattices by shape and storage, kept for re-use while their device bytes stay
// within a budget (1/8 of the device's memory, or BGT_HIP_LATTICE_CACHE_BYTES bytes when
// that environment variable is set; 0 disables the cache): creating a lattice allocates device and
// pinned host memory and a stream, and releasing them synchronises the device, which
// dominates small lattices' lifetimes (tree children, serving sessions). The cache lives
// until the process ends and is never destroyed, so no HIP call runs during static
// destruction.
class LatticeCache {
public:
	static LatticeCache &get()
	{
		static LatticeCache *cache = new LatticeCache();
		return *cache;
	}
	std::unique_ptr<DeviceLattice> take(bool dilution, int n, int k, Storage st)
	{
		std::lock_guard<std::mutex> lk(m_);
		auto it = free_.find(Key{dilution, n, k, (int)st});
		if (it == free_.end() || it->second.empty())
			return nullptr;
		auto l = std::move(it->second.back());
		it->second.pop_back();
		--count_;
		bytes_ -= l->device_bytes();
		return l;
	}
	void give(std::unique_ptr<DeviceLattice> l)
	{
		{
			std::lock_guard<std::mutex> lk(m_);
			if (!budget_set_) {
				size_t free_b = 0, total_b = 0;
				budget_ = hipMemGetInfo(&free_b, &total_b) == hipSuccess ? total_b / 8 : (1ull << 30);
				if (const char *env = std::getenv("BGT_HIP_LATTICE_CACHE_BYTES"))
					budget_ = std::strtoull(env, nullptr, 10);
				budget_set_ = true;
			}
			if (bytes_ + l->device_bytes() <= budget_ && count_ < kMaxCached) {
				bytes_ += l->device_bytes();
				++count_;
				free_[Key{l->dilution(), l->subjects(), l->variants(), (int)l->storage()}].push_back(std::move(l));
				return;
			}
		}
		l.reset(); // over budget: released outside the lock
	}

private:
	using Key = std::tuple<bool, int, int, int>;
	// small lattices cost little device memory but each holds pinned host buffers (about
	// 64 KiB): the count is bounded too
	static constexpr size_t kMaxCached = 4096;
	std::mutex m_;
	std::map<Key, std::vector<std::unique_ptr<DeviceLattice>>> free_;
	size_t count_ = 0;
	uint64_t bytes_ = 0, budget_ = 0;
	bool budget_set_ = false;
};

} // namespace

std::unique_ptr<DeviceLattice> DeviceLattice::make(bool dilution, int subjects, int variants, const double *prior,
												   Storage storage)
{
	if (subjects > 0 && variants > 0 && prior != nullptr && subjects * variants <= kMaxAtoms) {
		if (auto l = LatticeCache::get().take(dilution, subjects, variants, storage)) {
			l->settle();
			l->init_prior(prior);
			return l;
		}
	}
	return std::make_unique<DeviceLattice>(dilution, subjects, variants, prior, storage);
}

std::unique_ptr<DeviceLattice> DeviceLattice::make_like(const DeviceLattice &like)
{
	if (auto l = LatticeCache::get().take(like.dilution_, like.n_, like.k_, like.storage_)) {
		l->settle();
		return l;
	}
	return std::make_unique<DeviceLattice>(like, false);
}

void DeviceLattice::recycle(std::unique_ptr<DeviceLattice> lattice)
{
	if (lattice)
		LatticeCache::get().give(std::move(lattice));
}

std::unique_ptr<DeviceLattice> DeviceLattice::project(const DeviceLattice &src, uint32_t removed)
{
	const uint32_t all = src.atoms_ >= 32 ? 0xFFFFFFFFu : ((1u << src.atoms_) - 1u);
	if (removed & ~all)
		throw Error(BGT_HIP_BAD_INPUT, "removed atoms outside the lattice");
	int removed_subjects = 0;
	for (int i = 0; i < src.n_; ++i) {
		uint32_t subject = 0;
		for (int v = 0; v < src.k_; ++v)
			subject |= 1u << (v * src.n_ + i);
		if ((removed & subject) == subject)
			++removed_subjects;
		else if (removed & subject)
			throw Error(BGT_HIP_BAD_INPUT, "removed atoms must cover every variant of each removed subject");
	}
	const int n2 = src.n_ - removed_subjects;
	if (n2 < 1)
		throw Error(BGT_HIP_UNSUPPORTED, "projection would remove every subject");
	std::unique_ptr<DeviceLattice> l = LatticeCache::get().take(src.dilution_, n2, src.k_, src.storage_);
	if (l)
		l->settle();
	else
		l.reset(new DeviceLattice(src.dilution_, n2, src.k_, src.storage_, Blank{}));
	const uint32_t keep = all & ~removed;
	// the source's pending work on its own stream (an update) comes first; work of batched
	// steps has finished when they return. (The stream is shared with other lattices: a wait
	// for the whole stream would also wait for theirs.)
	const_cast<DeviceLattice &>(src).settle();
	launch_project(src.post_, src.scal_, src.inv_cur_, l->post_, l->atoms_, removed, src.storage_, l->stream_, l->part_,
				   2048);
	HostBox *box = static_cast<HostBox *>(l->host_box_);
	for (int s = 0; s < kScalars; ++s)
		box->scal[s] = s < 4 ? 1.0 : 0.0; // pending normalisation 1: the sums are the values
	BGT_HIP_CHECK(hipMemcpyAsync(l->scal_, box->scal, sizeof(double) * kScalars, hipMemcpyHostToDevice, l->stream_));
	auto compact = [&](uint32_t mask) { // the kept atoms' bits, in order
		uint32_t out = 0;
		int at = 0;
		for (uint32_t m = keep; m; m &= m - 1u, ++at)
			if (mask & (m & (0u - m)))
				out |= 1u << at;
		return out;
	};
	l->pos_ = compact(src.pos_);
	l->neg_ = compact(src.neg_);
	l->active_ = (1u << n2) - 1u;
	l->updates_ = src.updates_ + 1;
	l->inv_cur_ = kScalInv0;
	l->e_ = l->r_ = 0;
	l->len_ = 0;
	l->hist_uploaded_ = 0;
	l->exact_pending_ = l->exact_valid_ = false;
	l->pairs_own_ = false;
	// the exact program stays the source's, in its frame: the kept subjects of this
	// projection, placed among the source's own kept subjects
	uint32_t kept_subjects = 0;
	for (int i = 0; i < src.n_; ++i)
		if (!((removed >> i) & 1u))
			kept_subjects |= 1u << i;
	l->exact_ = src.exact_;
	l->xn_ = src.xn_;
	l->xkeep_ = src.to_frame(kept_subjects);
	// each kept entry sums 2^r source entries in increasing order, each times 1/D: at most
	// 2^r roundings relative (the FP32C form's one rounding to FP32 counts as the extra update)
	l->proj_rel_ = src.proj_rel_ + std::ldexp(1.0, __builtin_popcount(removed));
	l->copy_gather_history(src);
	l->clock_.reset();
	l->clock_.enabled = false;
	l->busy_ = true;
	return l;
}

DeviceLattice::DeviceLattice(const DeviceLattice &parent, bool copy_posterior)
	: dilution_(parent.dilution_), n_(parent.n_), k_(parent.k_), atoms_(parent.atoms_), states_(parent.states_),
	  storage_(parent.storage_)
{
	stream_ = stream_acquire();
	allocate();
	pos_ = parent.pos_;
	neg_ = parent.neg_;
	active_ = parent.active_;
	updates_ = parent.updates_;
	inv_cur_ = parent.inv_cur_;
	exact_ = parent.exact_;
	xn_ = parent.xn_;
	xkeep_ = parent.xkeep_;
	proj_rel_ = parent.proj_rel_;
	copy_gather_history(parent);
	if (copy_posterior) {
		// order the copy after the parent's pending work
		BGT_HIP_CHECK(hipStreamSynchronize(parent.stream_));
		const size_t eb = storage_ == Storage::fp64 ? 8 : 4;
		BGT_HIP_CHECK(hipMemcpyAsync(post_, parent.post_, eb * states_, hipMemcpyDeviceToDevice, stream_));
		BGT_HIP_CHECK(hipMemcpyAsync(scal_, parent.scal_, sizeof(double) * kScalars, hipMemcpyDeviceToDevice, stream_));
		busy_ = true;
	}
}

namespace {

struct ArenaLayout {
	size_t scal, res, cnt, mass, band, cand, flags, hist, exact_u, pairs, part, tops, ftab, fix, fix_bytes, scores, sel,
		gexact, ghin, total;
	bool exact;
};

ArenaLayout arena_layout(bool dilution, int n, int k)
{
	const int atoms = n * k;
	ArenaLayout a{};
	const uint64_t tiles = atoms > kTileBits ? (1ull << (atoms - kTileBits)) : 1ull;
	const uint64_t chunks = atoms > 14 ? (1ull << (atoms - 14)) : 1ull;
	const uint64_t ec = 1ull << n;
	const size_t part = std::max<uint64_t>({2048, tiles, (chunks + 1) * kMaxAtoms});
	size_t tops = std::max<uint64_t>({tiles, ec / 256 + 1, 1});
	if (n <= 16)
		tops = std::max<uint64_t>(tops, ec);
	const size_t ftab = k >= 2 ? (size_t(1) << k) * ec : 1;
	const size_t scores = (k == 1 && n <= 20) ? ec : 1;
	size_t off = 0;
	auto take = [&](size_t bytes) {
		const size_t o = off;
		off += align256(bytes);
		return o;
	};
	// binary lattices collect their error band for exact re-decision; the device's dense
	// program (k_exact_band) covers the non-dilution ones. Lattices of several variants
	// collect their Op-BBPA band (k_bbpa_band) in the same place, right after the 1 KB block,
	// with room for every experiment (2^n, n <= 15), so it never overflows.
	a.exact = k == 1;
	const size_t gband = k >= 2 ? sizeof(Cand) * ec : 0;
	a.scal = take(sizeof(double) * kScalars); // 0
	a.res = take(sizeof(SelResult));         // 256
	a.cnt = take(sizeof(unsigned) * 4);      // 512
	a.mass = take(sizeof(double) * 32);      // 768
	a.exact_u = take(a.exact ? sizeof(double) * kExactBand : 0);        // kHeadExactWide
	a.band = take(a.exact ? sizeof(Cand) * band_capacity(atoms) : gband); // kHeadBandWide (exact), else kHeadBand
	a.cand = take(a.exact ? sizeof(Cand) * 2 * kTileThreads * zeta_final_groups(atoms) : 0);
	a.flags = take(a.exact ? sizeof(unsigned) * kMaxFinalGroups : 0);
	a.hist = take(a.exact ? sizeof(ExactRec) * kExactMaxHist : 0);
	a.pairs = take(sizeof(double) * 2 * kMaxAtoms);
	a.part = take(sizeof(double) * part);
	a.tops = take(sizeof(SelResult) * tops);
	a.ftab = take(sizeof(double) * ftab);
	a.fix_bytes = bbpa_fix_bytes(n, k);
	a.fix = take(a.fix_bytes);
	a.scores = take(sizeof(double) * scores);
	a.sel = take(sizeof(uint32_t) * 16);
	// several variants: exact objectives of the band, and the exact program's inputs (the
	// prior pairs and the dilution table)
	a.gexact = take(k >= 2 ? sizeof(double) * ec : 0);
	// (sized for the largest frame a projection can keep: kMaxAtoms atoms, kMaxAtoms / k subjects)
	a.ghin = take(k >= 2 ? sizeof(double) * (2 * kMaxAtoms + (kMaxAtoms / k) * (kMaxAtoms / k + 1)) : 0);
	a.total = std::max<size_t>(off, a.exact ? kHeadBytesWide : kHeadBytes);
	return a;
}

} // namespace

uint64_t DeviceLattice::footprint(bool dilution, int subjects, int variants, Storage storage)
{
	const int atoms = subjects * variants;
	const size_t eb = storage == Storage::fp64 ? 8 : 4;
	const uint64_t zb = bbpa_hist_path(subjects, variants) ? 0 : z_bytes(atoms, eb);
	return (uint64_t)eb * (1ull << atoms) + zb + arena_layout(dilution, subjects, variants).total;
}

void DeviceLattice::allocate()
{
	const size_t eb = storage_ == Storage::fp64 ? 8 : 4;
	post_ = device_alloc(eb * states_, stream_);
	bytes_ += eb * states_;
	const ArenaLayout a = arena_layout(dilution_, n_, k_);
	if (a.exact ? (a.exact_u != kHeadExactWide || a.band != kHeadBandWide) : a.band != kHeadBand)
		throw Error(BGT_HIP_INTERNAL, "arena layout does not match the head copy");
	arena_ = static_cast<char *>(device_alloc(a.total, stream_));
	bytes_ += a.total;
	BGT_HIP_CHECK(hipMemsetAsync(arena_, 0, a.total, stream_));
	scal_ = reinterpret_cast<double *>(arena_ + a.scal);
	result_ = reinterpret_cast<SelResult *>(arena_ + a.res);
	counter_ = reinterpret_cast<unsigned *>(arena_ + a.cnt);
	masses_ = reinterpret_cast<double *>(arena_ + a.mass);
	pairs_ = reinterpret_cast<double *>(arena_ + a.pairs);
	part_ = reinterpret_cast<double *>(arena_ + a.part);
	tops_ = reinterpret_cast<SelResult *>(arena_ + a.tops);
	ftab_ = reinterpret_cast<double *>(arena_ + a.ftab);
	if (a.fix_bytes) {
		fixtab_ = reinterpret_cast<unsigned long long *>(arena_ + a.fix);
		fix_bytes_ = a.fix_bytes;
	}
	scores_ = reinterpret_cast<double *>(arena_ + a.scores);
	selected_ = reinterpret_cast<uint32_t *>(arena_ + a.sel);
	if (a.exact) {
		band_ = reinterpret_cast<Cand *>(arena_ + a.band);
		cand_ = reinterpret_cast<Cand *>(arena_ + a.cand);
		wgflag_ = reinterpret_cast<unsigned *>(arena_ + a.flags);
		hist_dev_ = reinterpret_cast<ExactRec *>(arena_ + a.hist);
		exact_dev_ = reinterpret_cast<double *>(arena_ + a.exact_u);
		band_cap_ = (unsigned)band_capacity(atoms_);
		band_host_ = static_cast<Cand *>(pinned_alloc(sizeof(Cand) * band_cap_));
	} else if (k_ >= 2) {
		band_ = reinterpret_cast<Cand *>(arena_ + a.band);
		band_cap_ = 1u << n_; // every experiment
		band_host_ = static_cast<Cand *>(pinned_alloc(sizeof(Cand) * band_cap_));
		gexact_ = reinterpret_cast<double *>(arena_ + a.gexact);
		ghin_ = reinterpret_cast<double *>(arena_ + a.ghin);
	}
	double init[kScalars] = {1.0, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, 0.0};
	BGT_HIP_CHECK(hipMemcpyAsync(scal_, init, sizeof(init), hipMemcpyHostToDevice, stream_));
	host_box_ = pinned_alloc(sizeof(HostBox));
	sync(); // the allocations are complete before any other stream uses them
}

void DeviceLattice::ensure_z()
{
	// lattices whose Op-BBPA tables come from the histogram pass never transform
	if (z_ || atoms_ <= kTileBits || bbpa_hist_path(n_, k_))
		return;
	const size_t eb = storage_ == Storage::fp64 ? 8 : 4;
	const uint64_t zb = z_bytes(atoms_, eb);
	z_ = device_alloc(zb, stream_);
	bytes_ += zb;
	sync(); // batched steps use it from another stream
}

DeviceLattice::~DeviceLattice()
{
	// every use of this lattice's memory is on its own stream or finished (batched steps
	// synchronise before they return), so the frees need only order after its own stream
	if (stream_)
		(void)hipStreamSynchronize(stream_);
	device_free(post_, stream_);
	device_free(z_, stream_);
	device_free(arena_, stream_);
	pinned_release(host_box_, sizeof(HostBox));
	pinned_release(band_host_, sizeof(Cand) * band_cap_);
	stream_release(stream_);
}

void DeviceLattice::adopt(const DeviceLattice &parent)
{
	if (parent.dilution_ != dilution_ || parent.n_ != n_ || parent.k_ != k_ || parent.storage_ != storage_)
		throw Error(BGT_HIP_BAD_INPUT, "adopt: lattices differ in shape or storage");
	pos_ = parent.pos_;
	neg_ = parent.neg_;
	active_ = parent.active_;
	updates_ = parent.updates_;
	inv_cur_ = parent.inv_cur_;
	exact_ = parent.exact_;
	xn_ = parent.xn_;
	xkeep_ = parent.xkeep_;
	proj_rel_ = parent.proj_rel_;
	copy_gather_history(parent);
	hist_uploaded_ = 0;
	exact_pending_ = exact_valid_ = false;
	pairs_own_ = false; // the device prior pairs are those this lattice was created with
}

void DeviceLattice::as_source(LatDesc &d) const
{
	d.src_post = post_;
	d.src_scal = scal_;
	d.src_inv = inv_cur_;
}

LatDesc DeviceLattice::desc() const
{
	LatDesc d;
	std::memset(&d, 0, sizeof(d));
	d.post = post_;
	d.z = z_;
	d.scal = scal_;
	d.part = part_;
	d.tops = tops_;
	d.result = result_;
	d.counter = counter_;
	d.lik = nullptr;
	d.masses = masses_;
	d.ftab = ftab_;
	d.fixtab = fixtab_;
	d.n = n_;
	d.k = k_;
	d.atoms = atoms_;
	d.len = len_;
	d.e = e_;
	d.r = r_;
	d.active = active_;
	d.dilution = dilution_ ? 1 : 0;
	d.inv_in = inv_cur_;
	d.inv_out = inv_cur_ == kScalInv0 ? kScalInv1 : kScalInv0;
	d.cand = cand_;
	d.band = k_ == 1 ? band_ : nullptr;
	d.gband = k_ >= 2 ? band_ : nullptr;
	d.gh_len = -1;
	d.gx_n = xn_;
	d.gx_keep = xkeep_;
	if (k_ >= 2 && ghin_ != nullptr && gh_uploaded_ && !gh_broken_ && exact_.usable()) {
		d.gh_len = (int)gh_e_.size();
		for (size_t t = 0; t < gh_e_.size(); ++t) {
			d.gh_e[t] = gh_e_[t];
			d.gh_r[t] = gh_r_[t];
		}
		d.gh_pairs = ghin_;
		d.gh_dil = dilution_ ? ghin_ + 2 * (size_t)xn_ * k_ : nullptr;
		d.gexact = gexact_;
	}
	d.band_width = 0.0;
	d.wgflag = wgflag_;
	d.band_cap = band_cap_;
	d.hist = hist_dev_;
	d.hist_len = exact_.updates();
	d.hist_new_at = -1;
	d.pairs = pairs_;
	d.exact_u = nullptr; // set by the launches that want exact masses
	return d;
}

Descs DeviceLattice::single() const
{
	Descs ds;
	std::memset(&ds, 0, sizeof(ds));
	ds.one = desc();
	std::memcpy(ds.one_lik, lik_, sizeof(lik_));
	ds.many = nullptr;
	return ds;
}

void DeviceLattice::sync()
{
	BGT_HIP_CHECK(hipStreamSynchronize(stream_));
	busy_ = false;
}

void DeviceLattice::settle()
{
	if (busy_)
		sync();
}

// ----------------------------------------------------------------------- likelihoods

void DeviceLattice::fill_lik(uint32_t experiment, uint32_t response, const double *dil)
{
	// factor tables for likelihood() in the kernels (replicated_lattice.cpp:8-28)
	std::fill(std::begin(lik_), std::end(lik_), 0.0);
	const int len = __builtin_popcount(experiment);
	for (int v = 0; v < k_; ++v) {
		const bool rb = (response >> v) & 1u;
		double *f = lik_ + v * (n_ + 1);
		if (!dilution_) {
			f[0] = rb ? kNegativeResponse : (1.0 - kNegativeResponse);
			f[1] = rb ? (1.0 - kNegativeResponse) : kNegativeResponse;
		} else {
			for (int m = 0; m <= len; ++m) {
				// a pool of size 0 has no row (dilution.cpp indexes row -1); P(r=1 | E=0) := 0
				const double p1 = len > 0 ? dil[(size_t)(len - 1) * (n_ + 1) + m] : 0.0;
				f[m] = rb ? p1 : 1.0 - p1;
			}
		}
	}
	e_ = experiment;
	r_ = response;
	len_ = len;
}

void DeviceLattice::prepare_update(uint32_t experiment, uint32_t response, const double *dil)
{
	const uint32_t emask = (n_ >= 32) ? 0xFFFFFFFFu : ((1u << n_) - 1u);
	if (experiment & ~emask)
		throw Error(BGT_HIP_BAD_INPUT, "experiment has bits beyond the subject count");
	if (response >> k_)
		throw Error(BGT_HIP_BAD_INPUT, "response has bits beyond the variant count");
	if (dilution_ && dil == nullptr)
		throw Error(BGT_HIP_BAD_INPUT, "dilution table is required for dilution lattices");
	fill_lik(experiment, response, dil);
	pending_dil_ = dil;
}

void DeviceLattice::commit_update()
{
	inv_cur_ = inv_cur_ == kScalInv0 ? kScalInv1 : kScalInv0;
	++updates_;
	// the exact program works in its frame (a projected lattice's pools placed among the
	// subjects of the lattice it came from, its table at the frame's row stride)
	const uint32_t xe = to_frame(e_);
	std::vector<double> ftab;
	const double *xdil = pending_dil_;
	if (dilution_ && pending_dil_ != nullptr && xn_ != n_) {
		frame_table(pending_dil_, ftab);
		xdil = ftab.data();
	}
	exact_.push(xe, r_, xdil);
	if (k_ >= 2 && !gh_broken_) {
		// the history of the exact Op-BBPA re-decision on the device (one dilution table only:
		// the rows this lattice can use must equal those recorded)
		if ((int)gh_e_.size() >= kGatherMaxHist) {
			gh_broken_ = true;
		} else {
			gh_e_.push_back(xe);
			gh_r_.push_back((uint8_t)r_);
			if (dilution_) {
				const size_t cells = (size_t)xn_ * (xn_ + 1);
				if (gh_dil_.empty()) {
					gh_dil_.assign(xdil, xdil + cells);
					gh_uploaded_ = false;
				} else {
					for (int row = 0; row < n_ && !gh_broken_; ++row)
						gh_broken_ = std::memcmp(gh_dil_.data() + (size_t)row * (xn_ + 1), xdil + (size_t)row * (xn_ + 1),
												 sizeof(double) * (n_ + 1)) != 0;
				}
			}
		}
	}
	pending_dil_ = nullptr;
}

void DeviceLattice::copy_gather_history(const DeviceLattice &parent)
{
	gh_e_ = parent.gh_e_;
	gh_r_ = parent.gh_r_;
	gh_dil_ = parent.gh_dil_;
	gh_broken_ = parent.gh_broken_;
	gh_uploaded_ = false;
}

uint32_t DeviceLattice::to_frame(uint32_t subjects) const
{
	if (xn_ == n_)
		return subjects;
	uint32_t out = 0;
	for (uint32_t r = xkeep_; subjects != 0u && r != 0u; r &= r - 1u, subjects >>= 1)
		if (subjects & 1u)
			out |= r & (0u - r);
	return out;
}

uint32_t DeviceLattice::atoms_to_frame(uint32_t atoms) const
{
	if (xn_ == n_)
		return atoms;
	const uint32_t smask = (1u << n_) - 1u;
	uint32_t out = 0;
	for (int v = 0; v < k_; ++v)
		out |= to_frame((atoms >> (v * n_)) & smask) << (v * xn_);
	return out;
}

void DeviceLattice::frame_table(const double *dil, std::vector<double> &out) const
{
	out.assign((size_t)xn_ * (xn_ + 1), 0.0);
	for (int row = 0; row < n_; ++row)
		std::memcpy(out.data() + (size_t)row * (xn_ + 1), dil + (size_t)row * (n_ + 1), sizeof(double) * (n_ + 1));
}

void DeviceLattice::prepare_gather()
{
	// the exact program's inputs into the arena (the prior pairs, the dilution table), once
	// per lattice and table; synchronous, so that any stream's kernels may read them
	if (k_ < 2 || ghin_ == nullptr || gh_broken_ || gh_uploaded_ || !exact_.usable())
		return;
	const size_t xatoms = (size_t)xn_ * k_;
	const size_t cells = 2 * xatoms + (size_t)xn_ * (xn_ + 1);
	double *h = static_cast<double *>(pinned_alloc(sizeof(double) * cells));
	exact_.prior_pairs(h);
	for (size_t i = 0; i < (size_t)xn_ * (xn_ + 1); ++i)
		h[2 * xatoms + i] = i < gh_dil_.size() ? gh_dil_[i] : 0.0;
	const hipError_t e1 = hipMemcpyAsync(ghin_, h, sizeof(double) * cells, hipMemcpyHostToDevice, stream_);
	const hipError_t e2 = e1 == hipSuccess ? hipStreamSynchronize(stream_) : e1;
	pinned_release(h, sizeof(double) * cells);
	BGT_HIP_CHECK(e2);
	gh_uploaded_ = true;
}

void DeviceLattice::update(uint32_t experiment, uint32_t response, const double *dil)
{
	prepare_update(experiment, response, dil);
	clock_.begin(stream_);
	launch_update(single(), 1, states_, k_, dilution_, storage_, stream_);
	clock_.end(stream_, "update");
	commit_update();
	busy_ = true;
}

// ------------------------------------------------------------------------- selection

int DeviceLattice::resolve_selector(int selector, int variants)
{
	// simulation_helpers.cpp:40-45
	if (selector == kSelAuto)
		return variants == 1 ? kSelBha : kSelBbpa;
	return selector;
}

void DeviceLattice::check_selector(int selector) const
{
	if (selector < kSelAuto || selector > kSelBrute)
		throw Error(BGT_HIP_BAD_INPUT, "unsupported selector");
	if (resolve_selector(selector, k_) == kSelBha && k_ != 1)
		throw Error(BGT_HIP_BAD_INPUT, "Op-BHA experiment selection is only supported for binary lattices");
	if (selector == kSelBrute && (n_ > 16 || k_ > 4 || n_ + atoms_ > 40))
		throw Error(BGT_HIP_UNSUPPORTED, "brute-force selection is limited to 16 subjects, 4 variants and 2^40 work");
}

void DeviceLattice::fetch_head()
{
	HostBox *box = static_cast<HostBox *>(host_box_);
	// lattices with an error band (their arena holds at least kHeadBytesWide) bring back
	// the wide head, and the band's exact masses when they were computed
	// in one copy (a second one, issued on another engine, waited about 30 us for the first)
	const bool wide = exact_dev_ != nullptr;
	BGT_HIP_CHECK(hipMemcpyAsync(box->head, arena_, wide ? kHeadBytesWide : kHeadBytes, hipMemcpyDeviceToHost,
								 stream_));
	if (clock_.enabled) {
		const auto t0 = std::chrono::steady_clock::now();
		sync();
		clock_.host("host_fetch_wait",
					std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
	} else {
		sync();
	}
	exact_valid_ = exact_pending_;
	exact_pending_ = false;
	if (exact_valid_)
		std::memcpy(box->exact_u, box->head + kHeadExactWide, sizeof(double) * kExactBand);
	unpack_head(wide ? kBandHeadWide : kBandHead, wide ? kHeadBandWide : kHeadBand);
}

void DeviceLattice::load_head(const unsigned char *head, bool exact_staged)
{
	HostBox *box = static_cast<HostBox *>(host_box_);
	std::memcpy(box->head, head, kHeadBytes);
	exact_valid_ = false;
	exact_pending_ = exact_staged; // the band's exact masses stay on the device until a near tie wants them
	unpack_head(kBandHead, kHeadBand);
}

bool DeviceLattice::stage_exact(LatDesc &d, bool sync_upload)
{
	exact_pending_ = false;
	if (exact_dev_ == nullptr || dilution_ || k_ != 1 || !pairs_own_ || xn_ != n_ || !exact_.usable() ||
		exact_.updates() > kExactMaxHist)
		return false;
	HostBox *box = static_cast<HostBox *>(host_box_);
	const int n = exact_.updates();
	int new_at = -1;
	if (hist_uploaded_ < n) {
		for (int t = hist_uploaded_; t < n; ++t) {
			ExactRec &rec = box->hist[t];
			exact_.record(t, rec.e, rec.fin, rec.fout);
			rec.pad = 0;
		}
		if (n - hist_uploaded_ == 1) {
			new_at = n - 1; // one new record: a kernel argument, no copy ahead of the kernel
		} else {
			BGT_HIP_CHECK(hipMemcpyAsync(hist_dev_ + hist_uploaded_, box->hist + hist_uploaded_,
										 sizeof(ExactRec) * (n - hist_uploaded_), hipMemcpyHostToDevice, stream_));
			if (sync_upload)
				sync(); // the kernel runs on another stream
		}
		hist_uploaded_ = n;
	}
	d.hist = hist_dev_;
	d.hist_len = n;
	d.hist_new_at = new_at;
	if (new_at >= 0)
		d.hist_new = box->hist[new_at];
	d.pairs = pairs_;
	d.exact_u = exact_dev_;
	exact_pending_ = true;
	return true;
}

void DeviceLattice::launch_exact(const Descs &ds_in)
{
	Descs ds = ds_in;
	if (!stage_exact(ds.one, false))
		return;
	clock_.begin(stream_);
	launch_exact_band(ds, 1, atoms_, stream_);
	clock_.end(stream_, "exact_band");
}

void DeviceLattice::unpack_head(int band_entries, size_t band_offset)
{
	ahead_ = Ahead{}; // formed for the previous head
	HostBox *box = static_cast<HostBox *>(host_box_);
	std::memcpy(box->scal, box->head, sizeof(double) * kScalars);
	std::memcpy(&box->result, box->head + kHeadResult, sizeof(SelResult));
	std::memcpy(box->counter, box->head + kHeadCounter, sizeof(box->counter));
	std::memcpy(box->masses, box->head + kHeadMasses, sizeof(double) * 32);
	head_band_ = 0;
	if (band_host_) {
		head_band_ = (unsigned)std::min<unsigned>({(unsigned)band_entries, box->counter[0], band_cap_});
		std::memcpy(band_host_, box->head + band_offset, sizeof(Cand) * head_band_);
	}
}

SelResult DeviceLattice::fetch_result()
{
	HostBox *box = static_cast<HostBox *>(host_box_);
	BGT_HIP_CHECK(hipMemcpyAsync(&box->result, result_, sizeof(SelResult), hipMemcpyDeviceToHost, stream_));
	sync();
	return box->result;
}

void DeviceLattice::fetch_masses(double *out)
{
	HostBox *box = static_cast<HostBox *>(host_box_);
	BGT_HIP_CHECK(hipMemcpyAsync(box->masses, masses_, sizeof(double) * 32, hipMemcpyDeviceToHost, stream_));
	sync();
	std::memcpy(out, box->masses, sizeof(double) * atoms_);
}

bgt_hip_selection DeviceLattice::finish_selection(int selector, const SelResult &r, bool)
{
	const int sel = resolve_selector(selector, k_);
	bgt_hip_selection out;
	std::memset(&out, 0, sizeof(out));
	const bool found = r.e1 != 0xFFFFFFFFu && r.d1 < (sel == kSelBrute ? INFINITY : 2.0);
	// reference defaults when nothing beats the initial best: Op-BHA returns 0
	// (bbpa.cpp:539), BBPA returns invalid_state (lattice.hpp:250), brute force 0.
	out.experiment = found ? r.e1 : (sel == kSelBbpa ? 0xFFFFFFFFu : 0u);
	out.objective = r.d1;
	out.runner_up = r.e2;
	out.runner_up_objective = r.d2;
	out.gap_bound = band_width(sel);
	out.near_tie = (r.d2 - r.d1) <= out.gap_bound ? 1 : 0;
	if (cand_ != nullptr)
		out.collected = (int)static_cast<const HostBox *>(host_box_)->counter[2];
	if (k_ >= 2 && sel == kSelBbpa && found) {
		const HostBox *box = static_cast<const HostBox *>(host_box_);
		if (box->counter[kGatherFlag] == 1u) {
			// re-decided on the device (k_gather_pick): the result holds the exact best two
			out.gap_bound = 2.0 * gather_objective_bound();
			out.near_tie = (r.d2 - r.d1) <= out.gap_bound ? 1 : 0;
			out.redecided = 1;
			out.candidates = (int)box->counter[0];
			out.band = (int)box->counter[0];
		}
		return out;
	}
	if (out.near_tie && found && exact_.usable() && k_ == 1 && (sel == kSelBha || sel == kSelBbpa))
		rescore(out, sel, r);
	return out;
}

bool DeviceLattice::prepare_rescore(int selector)
{
	const int sel = resolve_selector(selector, k_);
	if (!(sel == kSelBha || sel == kSelBbpa) || !exact_.usable())
		return false;
	const SelResult r = static_cast<const HostBox *>(host_box_)->result;
	const bool found = r.e1 != 0xFFFFFFFFu && r.d1 < 2.0;
	if (!found || !((r.d2 - r.d1) <= band_width(sel)))
		return false;
	if (!uses_enum())
		exact_.prepare(true); // the device enumeration forms its own normaliser
	return true;
}

bool DeviceLattice::exact_masses(const uint32_t *es, size_t count, double *out, bool *ok)
{
	if (!exact_.usable())
		return false;
	std::vector<uint32_t> xs;
	if (xn_ != n_) { // a projected lattice: the sets in the program's frame
		xs.resize(count);
		for (size_t i = 0; i < count; ++i)
			xs[i] = atoms_to_frame(es[i]);
		es = xs.data();
	}
	if (!uses_enum())
		return exact_.masses(es, count, out, ok);
	if (count > 0 && !(ahead_.valid && ahead_masses(es, count, out)))
		enum_masses(es, count, out);
	for (size_t i = 0; i < count; ++i)
		ok[i] = true;
	return true;
}

double DeviceLattice::exact_rel_bound() const
{
	// M = U(E) / U({}): both values' bounds and the division
	// per block M_v = U_v(E) / U_v({}) (two enumerations and a division); the product of
	// the k blocks' masses adds k - 1 roundings
	if (uses_enum())
		return (double)k_ * (2.0 * enum_rel_bound(xn_, exact_.enum_pools()) + 0x1p-53) + (double)(k_ - 1) * 0x1p-53;
	return exact_.rel_bound();
}

void DeviceLattice::enum_run(const std::vector<std::vector<uint32_t>> &queries, std::vector<std::vector<double>> &masses)
{
	// one launch_exact_enum per variant block, all on the stream, one copy of the values back
	// and one synchronisation; query 0 of each block is the empty set, its normaliser U({})
	const int T = exact_.enum_pools();
	const int xn = xn_; // the program's frame
	const int stride = xn + 1;
	auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
	struct Part {
		size_t q, o_pairs, o_fac, o_pool, o_q, o_part, o_out, h_out;
	};
	std::vector<Part> parts(queries.size());
	size_t in_bytes = 0;
	for (size_t v = 0; v < queries.size(); ++v) {
		Part &p = parts[v];
		p.q = queries[v].size() + 1;
		if (p.q > 65535)
			throw Error(BGT_HIP_INTERNAL, "too many exact masses for one launch");
		p.o_pairs = in_bytes;
		p.o_fac = p.o_pairs + up(sizeof(double) * 2 * xn);
		p.o_pool = p.o_fac + up(sizeof(double) * (size_t)T * stride + 8);
		p.o_q = p.o_pool + up(sizeof(uint32_t) * (size_t)T + 4);
		in_bytes = p.o_q + up(sizeof(uint32_t) * p.q);
	}
	size_t dev_bytes = in_bytes, out_bytes = 0;
	for (Part &p : parts) {
		p.o_part = dev_bytes;
		dev_bytes += up(sizeof(double) * p.q * kEnumMaxGroups);
		p.o_out = dev_bytes;
		dev_bytes += up(sizeof(double) * p.q);
		p.h_out = in_bytes + out_bytes;
		out_bytes += up(sizeof(double) * p.q);
	}
	const size_t host_bytes = in_bytes + out_bytes;
	char *h = static_cast<char *>(pinned_alloc(host_bytes));
	char *d = static_cast<char *>(device_alloc(dev_bytes, stream_));
	for (size_t v = 0; v < queries.size(); ++v) {
		const Part &p = parts[v];
		exact_.enum_inputs((int)v, reinterpret_cast<double *>(h + p.o_pairs), reinterpret_cast<uint32_t *>(h + p.o_pool),
						   reinterpret_cast<double *>(h + p.o_fac), stride);
		uint32_t *qh = reinterpret_cast<uint32_t *>(h + p.o_q);
		qh[0] = 0u;
		if (!queries[v].empty())
			std::memcpy(qh + 1, queries[v].data(), sizeof(uint32_t) * queries[v].size());
	}
	try {
		BGT_HIP_CHECK(hipMemcpyAsync(d, h, in_bytes, hipMemcpyHostToDevice, stream_));
		clock_.begin(stream_);
		for (size_t v = 0; v < queries.size(); ++v) {
			const Part &p = parts[v];
			EnumJob job;
			job.pairs = reinterpret_cast<const double *>(d + p.o_pairs);
			job.pools = reinterpret_cast<const uint32_t *>(d + p.o_pool);
			job.fac = reinterpret_cast<const double *>(d + p.o_fac);
			job.queries = reinterpret_cast<const uint32_t *>(d + p.o_q);
			job.partial = reinterpret_cast<double *>(d + p.o_part);
			job.out = reinterpret_cast<double *>(d + p.o_out);
			job.atoms = xn;
			job.updates = T;
			job.stride = stride;
			job.count = (int)p.q;
			launch_exact_enum(job, stream_);
		}
		clock_.end(stream_, "exact_enum");
		for (const Part &p : parts)
			BGT_HIP_CHECK(hipMemcpyAsync(h + p.h_out, d + p.o_out, sizeof(double) * p.q, hipMemcpyDeviceToHost, stream_));
		sync();
	} catch (...) {
		device_free(d, stream_);
		pinned_release(h, host_bytes);
		throw;
	}
	device_free(d, stream_);
	masses.resize(queries.size());
	for (size_t v = 0; v < queries.size(); ++v) {
		const double *u = reinterpret_cast<const double *>(h + parts[v].h_out);
		masses[v].resize(queries[v].size());
		for (size_t i = 0; i < queries[v].size(); ++i)
			masses[v][i] = u[i + 1] / u[0];
	}
	pinned_release(h, host_bytes);
}

void DeviceLattice::enum_masses(const uint32_t *es, size_t count, double *out)
{
	// the mass of an atom set (in the program's frame) is the product of its variant blocks'
	// masses (the posterior is a product over the blocks), in variant order
	const uint32_t smask = (xn_ >= 32) ? 0xFFFFFFFFu : ((1u << xn_) - 1u);
	std::vector<std::vector<uint32_t>> q(k_, std::vector<uint32_t>(count));
	for (int v = 0; v < k_; ++v)
		for (size_t i = 0; i < count; ++i)
			q[v][i] = (es[i] >> (v * xn_)) & smask;
	std::vector<std::vector<double>> m;
	enum_run(q, m);
	for (size_t i = 0; i < count; ++i) {
		double x = m[0][i];
		for (int v = 1; v < k_; ++v)
			x *= m[v][i];
		out[i] = x;
	}
}

double DeviceLattice::band_width(int sel) const
{
	// Two candidates' device objectives can be misordered only when they are closer than
	// twice the objective's error eps.
	// FP32C: each stored entry is the exact unnormalised posterior times (1 + e_S),
	// |e_S| <= (t+1) u32 (one FP32 rounding for the prior and per update; the FP64 factors
	// and 1/D are common to all entries). With x = y / sum(y), M(E) / M*(E) =
	// (1 + avg_{S>=E} e) / (1 + avg_all e), so M - M* = M* (1 - M*) (avg_E e - avg_notE e)
	// and |M - M*| <= M*(1 - M*) 2 (t+1) u32 <= (t+1) u32 / 2, plus one FP32 rounding of the
	// transform intermediate (M* u32) and FP64 sums far below; BBPA's |M - t| + |(T - M) - t|
	// doubles it (T = sum x is 1 up to FP64 rounding). With k >= 2 variants each of the 2^k
	// response masses is an inclusion-exclusion over 2^|R| feature masses, 3^k in all.
	// FP64: the per-mass relative bound of error_bound().
	double eps;
	if (storage_ == Storage::fp32c) {
		const double u32 = std::ldexp(1.0, -24);
		// the three-pass transform rounds its intermediate to FP32 once more (M* u32 more)
		const double z_roundings = three_pass_form(atoms_, storage_) ? 2.0 : 1.0;
		eps = (0.5 * (updates_ + 1) + 0.5 + z_roundings) * u32 * 1.001 + 1e-15;
	} else {
		double post_rel = 0.0, mass_rel = 0.0;
		error_bound(&post_rel, &mass_rel);
		eps = mass_rel;
	}
	double scale = 1.0;
	if (sel != kSelBha)
		scale = k_ == 1 ? 2.0 : std::pow(3.0, k_);
	return 2.0 * scale * eps;
}

void DeviceLattice::collect_band(const Descs &ds, unsigned entries, int objective, bool from_posterior, bool scaled,
								 double target)
{
	// lattices of at most kSmallBits atoms collect their band inside the transform kernel, and
	// after a persistent pass 2 the final-kernel launch already compacted it
	if (cand_ == nullptr || atoms_ <= kSmallBits || persistent_pass2(atoms_, objective, from_posterior))
		return;
	clock_.begin(stream_);
	launch_band_collect(ds, 1, entries, atoms_, k_, storage_, objective, from_posterior, scaled, target, stream_);
	clock_.end(stream_, "band_collect");
}

void DeviceLattice::collect_gather_band(double target, const Descs *fused)
{
	// lattices of several variants: the band, and its exact re-decision where it applies
	// (fused: the descriptors of a fused round, whose tables are scaled by the pending 1/D)
	if (k_ < 2 || band_ == nullptr)
		return;
	prepare_gather();
	Descs ds = fused ? *fused : single();
	ds.one.band_width = band_width(kSelBbpa);
	clock_.begin(stream_);
	launch_bbpa_band(ds, 1, n_, k_, target, stream_, fused != nullptr);
	clock_.end(stream_, "bbpa_band");
}

double DeviceLattice::gather_objective_bound() const
{
	// each exact feature mass (at most 1) within exact_rel_bound; an objective sums 3^k of
	// them with signs, in sums of at most 2^k terms each, and adds 2^k distances
	const double terms = std::pow(3.0, k_), resp = double(1u << k_);
	return terms * exact_rel_bound() + (terms * resp + resp) * 0x1p-53;
}

namespace {

// Open-addressing set of band experiments with their side of the target (+1 above, -1
// below, 0 not known for sure).
class SideTable {
public:
	explicit SideTable(size_t n)
	{
		size_t cap = 16;
		while (cap < 2 * n)
			cap <<= 1;
		key_.assign(cap, kEmpty);
		side_.assign(cap, 0);
	}
	void put(uint32_t e, int side)
	{
		size_t h = slot(e);
		while (key_[h] != kEmpty && key_[h] != e)
			h = (h + 1) & (key_.size() - 1);
		key_[h] = e;
		side_[h] = (int8_t)side;
	}
	int side(uint32_t e) const
	{
		size_t h = slot(e);
		while (key_[h] != kEmpty) {
			if (key_[h] == e)
				return side_[h];
			h = (h + 1) & (key_.size() - 1);
		}
		return 0;
	}

private:
	static constexpr uint32_t kEmpty = 0xFFFFFFFFu;
	size_t slot(uint32_t e) const { return (size_t)((e * 0x9E3779B1u) >> 7) & (key_.size() - 1); }
	std::vector<uint32_t> key_;
	std::vector<int8_t> side_;
};

} // namespace

void DeviceLattice::rescore(bgt_hip_selection &out, int sel, const SelResult &r)
{
	if (!clock_.enabled) {
		rescore_band(out, sel, r);
		return;
	}
	const auto t0 = std::chrono::steady_clock::now();
	rescore_band(out, sel, r);
	clock_.host("host_rescore", std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
}

std::vector<Cand> DeviceLattice::live_band(int sel, const SelResult &r, bool &complete_out, unsigned &count_out)
{
	// The band's candidates, one entry per experiment, less those the dominance test
	// removes (see rescore_band); complete_out: the band was collected whole.
	HostBox *box = static_cast<HostBox *>(host_box_);
	const double eps = 0.5 * band_width(sel);
	const unsigned count = box->counter[0];
	auto tp = std::chrono::steady_clock::now();
	auto phase = [&](const char *name) {
		if (!clock_.enabled)
			return;
		const auto now = std::chrono::steady_clock::now();
		clock_.host(name, std::chrono::duration<double, std::milli>(now - tp).count());
		tp = now;
	};
	bool complete = cand_ != nullptr && count <= band_cap_;
	std::vector<Cand> cands;
	if (complete) {
		const unsigned inhead = std::min<unsigned>(count, head_band_);
		bool wait = false;
		if (count > inhead) {
			BGT_HIP_CHECK(hipMemcpyAsync(band_host_ + inhead, band_ + inhead, sizeof(Cand) * (count - inhead),
										 hipMemcpyDeviceToHost, stream_));
			wait = true;
		}
		if (exact_pending_ && !exact_valid_ && count > 0) {
			// a batched step left the band's exact masses on the device (k_exact_band)
			BGT_HIP_CHECK(hipMemcpyAsync(box->exact_u, exact_dev_,
										 sizeof(double) * std::min<unsigned>(count, kExactBand), hipMemcpyDeviceToHost,
										 stream_));
			exact_valid_ = true;
			wait = true;
		}
		exact_pending_ = false;
		if (wait)
			sync();
		cands.assign(band_host_, band_host_ + count);
		if (exact_valid_) {
			// band entry i carries i + 1 in the upper flag bits: its exact mass came back
			// from the device when box->exact_u[i] is not negative
			const unsigned tagged = std::min<unsigned>(count, kExactBand);
			for (unsigned i = 0; i < tagged; ++i)
				cands[i].flags |= (i + 1u) << kCandIndexShift;
		}
	}
	// one entry per experiment: the band of the persistent pass 2 holds each state once, so
	// only the device's top two need a look; a rescan repeats what the lists held, so there
	// the entries are sorted and the signed copy, then the one with a device mass, is kept
	const int objective = sel == kSelBha ? kObjHalving : (k_ == 1 ? kObjBinaryBbpa : kObjGather);
	if (complete && persistent_pass2(atoms_, objective, false)) {
		const size_t nb = cands.size();
		for (const Cand c : {Cand{r.d1, r.e1, 0}, Cand{r.d2, r.e2, 0}}) {
			if (c.e == 0xFFFFFFFFu)
				continue;
			bool seen = false;
			for (size_t q = 0; q < nb && !seen; ++q)
				seen = cands[q].e == c.e;
			if (!seen && !(cands.size() > nb && cands.back().e == c.e))
				cands.push_back(c);
		}
	} else {
		for (const Cand c : {Cand{r.d1, r.e1, 0}, Cand{r.d2, r.e2, 0}})
			if (c.e != 0xFFFFFFFFu)
				cands.push_back(c);
		auto rank = [](const Cand &c) {
			return (c.flags & kCandSided) * 2u + ((c.flags >> kCandIndexShift) != 0u);
		};
		std::sort(cands.begin(), cands.end(),
				  [&](const Cand &a, const Cand &b) { return a.e < b.e || (a.e == b.e && rank(a) > rank(b)); });
		cands.erase(
			std::unique(cands.begin(), cands.end(), [](const Cand &a, const Cand &b) { return a.e == b.e; }),
			cands.end());
	}
	phase("t_dedupe");
	auto side_of = [&](const Cand &c) {
		if (!(c.flags & kCandSided) || !(std::fabs(c.d) > eps))
			return 0;
		return c.d > 0.0 ? 1 : -1;
	};
	// the device marked the dominated band entries (exact mass -2) when it computed the
	// band's masses over the whole band; otherwise the same test runs here
	const bool device_dom = complete && exact_valid_ && count <= (unsigned)kExactBand;
	size_t sided = 0;
	if (!device_dom)
		for (const Cand &c : cands)
			sided += side_of(c) != 0;
	std::vector<Cand> live;
	if (device_dom) {
		live.reserve(cands.size());
		for (const Cand &c : cands) {
			const uint32_t tag = c.flags >> kCandIndexShift;
			if (tag == 0u || box->exact_u[tag - 1] != -2.0)
				live.push_back(c);
		}
	} else if (sided < 2) {
		live = std::move(cands); // no candidate can dominate another
	} else {
		SideTable sides(cands.size());
		for (const Cand &c : cands)
			sides.put(c.e, side_of(c));
		std::vector<char> dominated(cands.size(), 0);
		auto check = [&](size_t i) {
			const Cand &c = cands[i];
			const int sd = side_of(c);
			bool dom = false;
			if (sd > 0) {
				for (uint32_t rest = active_ & ~c.e; rest && !dom; rest &= rest - 1u)
					dom = sides.side(c.e | (rest & (0u - rest))) > 0;
			} else if (sd < 0) {
				for (uint32_t rest = c.e; rest && !dom; rest &= rest - 1u)
					dom = sides.side(c.e & ~(rest & (0u - rest))) < 0;
			}
			dominated[i] = dom;
		};
		if (cands.size() >= 256)
			parallel_for(cands.size(), [&](size_t i, int) { check(i); });
		else
			for (size_t i = 0; i < cands.size(); ++i)
				check(i);
		live.reserve(cands.size());
		for (size_t i = 0; i < cands.size(); ++i)
			if (!dominated[i])
				live.push_back(cands[i]);
	}
	phase("t_dominance");
	clock_.host("n_band_unique", (double)cands.size());
	clock_.host("n_band_live", (double)live.size());
	complete_out = complete;
	count_out = count;
	return live;
}

// most candidates one rescoring computes exact masses for
constexpr int kMaxRescored = 4096;

// the live band sorted by device objective; returns how many lead it that may be needed
// before the exact runner-up is known (the first chunk of rescore_band's scan)
static size_t first_chunk(std::vector<Cand> &live, double eps, size_t cap)
{
	std::sort(live.begin(), live.end(), [](const Cand &a, const Cand &b) {
		const double x = std::fabs(a.d), y = std::fabs(b.d);
		return x < y || (x == y && a.e < b.e);
	});
	if (live.size() < 2)
		return 1;
	const double reach = std::fabs(live[1].d) + 2.0 * eps;
	size_t need = 0;
	while (need < live.size() && need < cap && !(std::fabs(live[need].d) > reach))
		++need;
	return std::max<size_t>(need, 1);
}

void DeviceLattice::rescore_band(bgt_hip_selection &out, int sel, const SelResult &r)
{
	// Exact FP64 objectives for the candidates within the error band. The device best's
	// true objective is at most d1 + eps, so the true best has a device objective of at most
	// d1 + 2 eps: the band, which the device collects completely (the per-thread top-2 lists
	// plus a rescan of every workgroup whose lists may have hidden more) unless it overflows
	// its capacity.
	// Dominance: M(E) is strictly decreasing under inclusion (every state has positive
	// mass), and both objectives are increasing in |M - 1/2| (binary BBPA's is 2|M - 1/2|).
	// A candidate whose mass is surely above 1/2 therefore loses to any one-subject extension
	// whose mass is surely above 1/2, and one surely below 1/2 to any one-subject reduction
	// surely below 1/2. Such candidates are not rescored; the rest are, in order of device
	// objective, until the next one's lower bound exceeds the exact runner-up.
	HostBox *box = static_cast<HostBox *>(host_box_);
	const double eps = 0.5 * band_width(sel);
	// masses enumerated ahead by a batched step serve this rescoring only
	struct Spent {
		Ahead &a;
		~Spent() { a = Ahead{}; }
	} spent{ahead_};
	bool complete = false;
	unsigned count = 0;
	std::vector<Cand> live;
	if (ahead_.valid) {
		live = std::move(ahead_.live);
		complete = ahead_.complete;
		count = ahead_.count;
	} else {
		live = live_band(sel, r, complete, count);
	}
	auto tp = std::chrono::steady_clock::now();
	auto phase = [&](const char *name) {
		if (!clock_.enabled)
			return;
		const auto now = std::chrono::steady_clock::now();
		clock_.host(name, std::chrono::duration<double, std::milli>(now - tp).count());
		tp = now;
	};
	if (complete && live.size() == 1) {
		// every other candidate loses to one that beats it strictly: the survivor is the
		// exact best, decided without its exact mass (objective: the device's value)
		out.experiment = live[0].e;
		out.objective = std::fabs(live[0].d);
		if (out.runner_up == out.experiment) {
			out.runner_up = r.e1 == out.experiment ? r.e2 : r.e1;
			out.runner_up_objective = r.e1 == out.experiment ? r.d2 : r.d1;
		}
		out.near_tie = 0;
		out.redecided = 1;
		out.candidates = 0;
		out.band = (int)count;
		return;
	}
	const double t = 1.0 / double(1 << k_);
	double b1 = INFINITY, b2 = INFINITY;
	uint32_t e1 = 0xFFFFFFFFu, e2 = 0xFFFFFFFFu;
	int rescored = 0;
	auto objective_of = [&](double m) {
		return sel == kSelBha ? std::fabs(m - 0.5) : std::fabs(m - t) + std::fabs((1.0 - m) - t);
	};
	// the best two by (exact objective, experiment): a strict total order over distinct
	// experiments, so the result does not depend on the order candidates arrive in
	auto consider = [&](double d, uint32_t e) {
		if (d < b1 || (d == b1 && e < e1)) {
			b2 = b1;
			e2 = e1;
			b1 = d;
			e1 = e;
		} else if (d < b2 || (d == b2 && e < e2)) {
			b2 = d;
			e2 = e;
		}
	};
	// When the device returned the exact mass of every live candidate, one pass over them
	// gives the same best two as the ordered scan below (which only spares exact masses that
	// would have to be computed on the host), without sorting the band.
	bool linear = exact_valid_ && !live.empty() && live.size() <= (size_t)kMaxRescored;
	for (size_t i = 0; linear && i < live.size(); ++i) {
		const uint32_t tag = live[i].flags >> kCandIndexShift;
		linear = tag != 0u && box->exact_u[tag - 1] >= 0.0;
	}
	double m0;
	if (linear)
		linear = exact_.normalise(box->exact_u[(live[0].flags >> kCandIndexShift) - 1], m0);
	if (linear) {
		for (const Cand &c : live) {
			double m;
			exact_.normalise(box->exact_u[(c.flags >> kCandIndexShift) - 1], m);
			consider(objective_of(m), c.e);
		}
		rescored = (int)live.size();
		clock_.host("n_exact_device", (double)live.size());
		clock_.host("n_exact_host", 0.0);
	} else {
		// The scan below is sequential; the exact masses it needs are computed ahead, in
		// parallel, so the result is that of the sequential scan. It never goes past a device
		// objective of |d_2| + 2 eps (d_2: the second candidate's): the exact objectives of the
		// first two candidates are at most |d_2| + eps, so the exact runner-up is too, and the
		// scan stops at the first candidate whose lower bound |d| - eps exceeds it. Those masses
		// are computed in one parallel run (or were, by a batched step's enumeration); later
		// chunks are only a guard.
		std::vector<uint32_t> ex;
		std::vector<double> mass;
		std::unique_ptr<bool[]> okv;
		size_t ready = 0, chunk = first_chunk(live, eps, (size_t)kMaxRescored);
		for (size_t i = 0; i < live.size(); ++i) {
			const Cand &c = live[i];
			if (std::fabs(c.d) - eps > b2)
				break; // its exact objective exceeds the current runner-up, as do all later ones
			if (rescored == kMaxRescored) {
				complete = false;
				break;
			}
			if (i == ready) {
				// next chunk: candidates that may still be needed under the current runner-up
				size_t j = i;
				while (j < live.size() && j - i < chunk && !(std::fabs(live[j].d) - eps > b2))
					++j;
				ex.resize(j - i);
				mass.resize(j - i);
				okv.reset(new bool[j - i]);
				// masses the device computed are normalised here; the rest run on the host
				std::vector<size_t> host;
				for (size_t q = i; q < j; ++q) {
					const uint32_t tag = live[q].flags >> kCandIndexShift;
					const double u = tag ? box->exact_u[tag - 1] : -1.0;
					if (u >= 0.0 && exact_.normalise(u, mass[q - i])) {
						okv[q - i] = true;
					} else {
						ex[host.size()] = live[q].e;
						host.push_back(q - i);
					}
				}
				clock_.host("n_exact_device", (double)(j - i - host.size()));
				clock_.host("n_exact_host", (double)host.size());
				if (!host.empty()) {
					std::vector<double> hm(host.size());
					std::unique_ptr<bool[]> hok(new bool[host.size()]);
					if (!exact_masses(ex.data(), host.size(), hm.data(), hok.get()))
						return; // the normaliser is beyond the exact program's budget: keep the device result
					for (size_t h = 0; h < host.size(); ++h) {
						mass[host[h]] = hm[h];
						okv[host[h]] = hok[h];
					}
				}
				ready = j;
				chunk *= 2;
			}
			const size_t slot = i - (ready - mass.size());
			if (!okv[slot])
				return; // beyond the exact program's budget: keep the device result
			const double m = mass[slot];
			++rescored;
			consider(objective_of(m), c.e);
		}
	}
	phase("t_scan");
	if (e1 == 0xFFFFFFFFu)
		return;
	out.experiment = e1;
	out.objective = b1;
	out.runner_up = e2;
	out.runner_up_objective = b2;
	if (complete) {
		out.gap_bound = 2.0 * (sel == kSelBha ? 1.0 : 2.0) * exact_rel_bound();
		out.near_tie = (b2 - b1) <= out.gap_bound ? 1 : 0;
	} else {
		// the band held more candidates than were collected: the exact order of the top
		// two is known, not that no other candidate beats them
		out.near_tie = 1;
	}
	out.redecided = complete ? 1 : 2;
	out.candidates = rescored;
	out.band = (int)count;
}

std::vector<uint32_t> DeviceLattice::enum_ahead_sets(int selector)
{
	// the conditions under which finish_selection calls rescore_band, and the sets of that
	// scan's first request for exact masses (no candidate has a device mass here)
	ahead_ = Ahead{};
	std::vector<uint32_t> sets;
	const int sel = resolve_selector(selector, k_);
	if (k_ != 1 || !uses_enum() || !(sel == kSelBha || sel == kSelBbpa))
		return sets;
	const SelResult r = static_cast<const HostBox *>(host_box_)->result;
	const bool found = r.e1 != 0xFFFFFFFFu && r.d1 < 2.0;
	if (!found || !((r.d2 - r.d1) <= band_width(sel)))
		return sets;
	bool complete = false;
	unsigned count = 0;
	std::vector<Cand> live = live_band(sel, r, complete, count);
	if (!(complete && live.size() == 1) && !live.empty()) {
		const size_t need = std::min(first_chunk(live, 0.5 * band_width(sel), (size_t)kMaxRescored), live.size());
		sets.reserve(need);
		for (size_t i = 0; i < need; ++i)
			sets.push_back(xn_ != n_ ? atoms_to_frame(live[i].e) : live[i].e);
	}
	ahead_.live = std::move(live);
	ahead_.complete = complete;
	ahead_.count = count;
	ahead_.valid = true;
	return sets;
}

bool DeviceLattice::ahead_masses(const uint32_t *frame_sets, size_t count, double *out) const
{
	for (size_t i = 0; i < count; ++i) {
		const auto it = std::lower_bound(ahead_.sets.begin(), ahead_.sets.end(), frame_sets[i]);
		if (it == ahead_.sets.end() || *it != frame_sets[i])
			return false;
		out[i] = ahead_.masses[(size_t)(it - ahead_.sets.begin())];
	}
	return true;
}

void DeviceLattice::enum_ahead_run(const std::vector<DeviceLattice *> &lats,
								   const std::vector<std::vector<uint32_t>> &sets, hipStream_t stream)
{
	// enum_run's inputs and outputs for every lattice with sets, one job each (one variant
	// block), the job table itself uploaded with them; query 0 of a job is U({})
	auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
	struct Part {
		DeviceLattice *lat;
		const std::vector<uint32_t> *q;
		size_t o_pairs, o_fac, o_pool, o_q, o_part, o_out;
	};
	std::vector<Part> parts;
	for (size_t i = 0; i < lats.size(); ++i)
		if (!sets[i].empty() && lats[i]->ahead_.valid)
			parts.push_back(Part{lats[i], &sets[i], 0, 0, 0, 0, 0, 0});
	if (parts.empty())
		return;
	const size_t o_jobs = 0;
	size_t in_bytes = up(sizeof(EnumJob) * parts.size());
	for (Part &p : parts) {
		const int xn = p.lat->xn_, T = p.lat->exact_.enum_pools();
		if (p.q->size() + 1 > 65535)
			throw Error(BGT_HIP_INTERNAL, "too many exact masses for one launch");
		p.o_pairs = in_bytes;
		p.o_fac = p.o_pairs + up(sizeof(double) * 2 * xn);
		p.o_pool = p.o_fac + up(sizeof(double) * (size_t)T * (xn + 1) + 8);
		p.o_q = p.o_pool + up(sizeof(uint32_t) * (size_t)T + 4);
		in_bytes = p.o_q + up(sizeof(uint32_t) * (p.q->size() + 1));
	}
	size_t dev_bytes = in_bytes;
	for (Part &p : parts) {
		p.o_part = dev_bytes;
		dev_bytes += up(sizeof(double) * (p.q->size() + 1) * kEnumMaxGroups);
	}
	const size_t o_outs = dev_bytes; // every job's values, contiguous: one copy back
	for (Part &p : parts) {
		p.o_out = dev_bytes;
		dev_bytes += sizeof(double) * (p.q->size() + 1);
	}
	const size_t out_bytes = dev_bytes - o_outs;
	const size_t host_bytes = in_bytes + out_bytes;
	char *h = static_cast<char *>(pinned_alloc(host_bytes));
	char *d = nullptr;
	try {
		d = static_cast<char *>(device_alloc(dev_bytes, stream));
		std::vector<EnumJob> jobs(parts.size());
		for (size_t j = 0; j < parts.size(); ++j) {
			const Part &p = parts[j];
			const int xn = p.lat->xn_;
			p.lat->exact_.enum_inputs(0, reinterpret_cast<double *>(h + p.o_pairs),
									  reinterpret_cast<uint32_t *>(h + p.o_pool), reinterpret_cast<double *>(h + p.o_fac),
									  xn + 1);
			uint32_t *qh = reinterpret_cast<uint32_t *>(h + p.o_q);
			qh[0] = 0u;
			std::memcpy(qh + 1, p.q->data(), sizeof(uint32_t) * p.q->size());
			EnumJob &job = jobs[j];
			job.pairs = reinterpret_cast<const double *>(d + p.o_pairs);
			job.pools = reinterpret_cast<const uint32_t *>(d + p.o_pool);
			job.fac = reinterpret_cast<const double *>(d + p.o_fac);
			job.queries = reinterpret_cast<const uint32_t *>(d + p.o_q);
			job.partial = reinterpret_cast<double *>(d + p.o_part);
			job.out = reinterpret_cast<double *>(d + p.o_out);
			job.atoms = xn;
			job.updates = p.lat->exact_.enum_pools();
			job.stride = xn + 1;
			job.count = (int)(p.q->size() + 1);
		}
		std::memcpy(h + o_jobs, jobs.data(), sizeof(EnumJob) * jobs.size());
		BGT_HIP_CHECK(hipMemcpyAsync(d, h, in_bytes, hipMemcpyHostToDevice, stream));
		launch_exact_enum_many(reinterpret_cast<const EnumJob *>(d + o_jobs), jobs, stream);
		BGT_HIP_CHECK(hipMemcpyAsync(h + in_bytes, d + o_outs, out_bytes, hipMemcpyDeviceToHost, stream));
		BGT_HIP_CHECK(hipStreamSynchronize(stream));
	} catch (...) {
		device_free(d, stream);
		pinned_release(h, host_bytes);
		throw;
	}
	device_free(d, stream);
	for (const Part &p : parts) {
		const double *u = reinterpret_cast<const double *>(h + in_bytes + (p.o_out - o_outs));
		std::vector<size_t> order(p.q->size());
		for (size_t i = 0; i < order.size(); ++i)
			order[i] = i;
		std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return (*p.q)[a] < (*p.q)[b]; });
		Ahead &a = p.lat->ahead_;
		a.sets.resize(order.size());
		a.masses.resize(order.size());
		for (size_t i = 0; i < order.size(); ++i) {
			a.sets[i] = (*p.q)[order[i]];
			a.masses[i] = u[order[i] + 1] / u[0]; // as enum_run forms M
		}
	}
	pinned_release(h, host_bytes);
}

bgt_hip_selection DeviceLattice::select(int selector)
{
	check_selector(selector);
	const int sel = resolve_selector(selector, k_);
	const Descs ds = single();
	if (sel == kSelBrute) {
		clock_.begin(stream_);
		launch_brute_force(ds, 1, n_, k_, storage_, stream_);
		clock_.end(stream_, "brute_force");
		return finish_selection(selector, fetch_result(), false);
	}
	const int objective = sel == kSelBha ? kObjHalving : (k_ == 1 ? kObjBinaryBbpa : kObjGather);
	const double target = sel == kSelBha ? 0.5 : 1.0 / double(1 << k_);
	if (objective == kObjGather && bbpa_hist_path(n_, k_)) {
		// the feature tables from one histogram pass (bbpa.cpp:237-313), no transform
		prepare_gather();
		Descs dh = single();
		dh.one.band_width = band_width(sel);
		clock_.begin(stream_, true);
		launch_bbpa_hist(dh, 1, n_, k_, storage_, false, stream_, false);
		clock_.end(stream_, "bbpa_hist");
		clock_.begin(stream_);
		launch_bbpa_finish(dh, 1, k_, false, target, stream_);
		clock_.end(stream_, "bbpa_finish");
		fetch_head();
		return finish_selection(selector, static_cast<HostBox *>(host_box_)->result, false);
	}
	ensure_z();
	Descs dz = single();
	dz.one.band_width = band_width(sel);
	unsigned entries = 0;
	if (atoms_ <= kSmallBits) {
		clock_.begin(stream_);
		entries = launch_zeta_small(dz, 1, atoms_, k_, storage_, objective, target, stream_);
		clock_.end(stream_, "zeta_small");
	} else if (atoms_ == kTileBits) {
		clock_.begin(stream_);
		entries = launch_zeta_final(dz, 1, atoms_, k_, storage_, objective, true, false, target, stream_);
		clock_.end(stream_, "zeta_final");
		if (objective != kObjGather)
			collect_band(dz, entries, objective, true, false, target);
	} else {
		clock_.begin(stream_, true);
		launch_zeta_pass1(dz, 1, atoms_, storage_, false, k_, dilution_, objective, stream_);
		clock_.end(stream_, "zeta_pass1");
		clock_.begin(stream_);
		entries = launch_zeta_final(dz, 1, atoms_, k_, storage_, objective, false, false, target, stream_);
		clock_.end(stream_, "zeta_pass2");
		if (objective != kObjGather)
			collect_band(dz, entries, objective, false, false, target);
	}
	if (objective == kObjGather) {
		clock_.begin(stream_);
		launch_bbpa_score(dz, 1, n_, k_, target, stream_);
		clock_.end(stream_, "bbpa_score");
		collect_gather_band(target);
	} else {
		launch_exact(dz);
	}
	exact_.prepare(); // on the host while the device selects
	fetch_head();
	return finish_selection(selector, static_cast<HostBox *>(host_box_)->result, false);
}

// ------------------------------------------------------------------- mass queries

double DeviceLattice::posterior(uint32_t state)
{
	if (state >= states_)
		throw Error(BGT_HIP_BAD_INPUT, "state out of range");
	HostBox *box = static_cast<HostBox *>(host_box_);
	const size_t eb = storage_ == Storage::fp64 ? 8 : 4;
	BGT_HIP_CHECK(hipMemcpyAsync(&box->raw, static_cast<char *>(post_) + eb * state, eb, hipMemcpyDeviceToHost, stream_));
	BGT_HIP_CHECK(hipMemcpyAsync(box->scal, scal_, sizeof(double) * kScalars, hipMemcpyDeviceToHost, stream_));
	sync();
	double y;
	if (storage_ == Storage::fp64) {
		y = box->raw;
	} else {
		float f;
		std::memcpy(&f, &box->raw, sizeof(f));
		y = f;
	}
	return y * box->scal[inv_cur_];
}

void DeviceLattice::read_posterior(double *out)
{
	HostBox *box = static_cast<HostBox *>(host_box_);
	BGT_HIP_CHECK(hipMemcpyAsync(box->scal, scal_, sizeof(double) * kScalars, hipMemcpyDeviceToHost, stream_));
	if (storage_ == Storage::fp64) {
		BGT_HIP_CHECK(hipMemcpyAsync(out, post_, 8 * states_, hipMemcpyDeviceToHost, stream_));
		sync();
		const double inv = box->scal[inv_cur_];
		for (uint64_t s = 0; s < states_; ++s)
			out[s] = out[s] * inv;
	} else {
		std::vector<float> tmp(states_);
		BGT_HIP_CHECK(hipMemcpyAsync(tmp.data(), post_, 4 * states_, hipMemcpyDeviceToHost, stream_));
		sync();
		const double inv = box->scal[inv_cur_];
		for (uint64_t s = 0; s < states_; ++s)
			out[s] = static_cast<double>(tmp[s]) * inv;
	}
}

double DeviceLattice::upset_mass(uint32_t experiment)
{
	const uint32_t full = (atoms_ >= 32) ? 0xFFFFFFFFu : ((1u << atoms_) - 1u);
	if (experiment & ~full)
		throw Error(BGT_HIP_BAD_INPUT, "experiment has bits beyond the atom count");
	clock_.begin(stream_);
	launch_upset_mass(single(), 1, atoms_, experiment, storage_, stream_);
	clock_.end(stream_, "upset_mass");
	HostBox *box = static_cast<HostBox *>(host_box_);
	BGT_HIP_CHECK(hipMemcpyAsync(box->scal, scal_, sizeof(double) * kScalars, hipMemcpyDeviceToHost, stream_));
	sync();
	return box->scal[kScalUpset];
}

void DeviceLattice::atom_masses(double *out)
{
	clock_.begin(stream_);
	launch_atom_masses(single(), 1, atoms_, storage_, stream_);
	clock_.end(stream_, "atom_masses");
	fetch_masses(out);
}

void DeviceLattice::classify(double up, double lo)
{
	// lattice_base.cpp:246-264
	double m[kMaxAtoms];
	atom_masses(m);
	for (int i = 0; i < atoms_; ++i) {
		const uint32_t placement = 1u << i;
		if ((pos_ | neg_) & placement)
			continue;
		if (m[i] < lo)
			pos_ |= placement;
		else if (m[i] > (1 - up))
			neg_ |= placement;
	}
}

std::vector<uint32_t> DeviceLattice::select_experiments(int steps)
{
	// lattice_api.cpp:116-168 (PUBLIC coordinates; binary lattices)
	if (steps < 0)
		throw Error(BGT_HIP_BAD_INPUT, "k must be non-negative");
	std::vector<uint32_t> out;
	if (steps == 0)
		return out;
	if (k_ != 1)
		throw Error(BGT_HIP_BAD_INPUT, "Op-BHA multi-experiment selection is only supported for binary lattices");
	if (steps > kLookMaxSteps)
		throw Error(BGT_HIP_UNSUPPORTED, "lookahead is limited to 16 steps (2^16 groups at the last step)");
	auto rev = [&](uint32_t v) {
		uint32_t r = 0;
		for (int i = 0; i < n_; ++i) {
			r <<= 1;
			r |= (v >> i) & 1u;
		}
		return r;
	};
	out.push_back(rev(select(kSelBha).experiment));
	if (steps == 1)
		return out;
	// Tie rule: the lowest public index among the candidates whose score lies within twice
	// the score's rounding bound of the lowest score. Candidates that tie in exact arithmetic
	// (common: whenever both halves of every earlier group fall on one side of the target,
	// the score does not depend on the candidate) are then decided by index, not by rounding.
	//
	// Scores per step: up to 4 earlier picks, every candidate's 2^(m+1) group masses in
	// registers (from 13 subjects through down-set sums built once, below by a pass over all
	// states per candidate); from 5 on, the class-decomposed sums (2^m masked down-set
	// transforms per step).
	const uint64_t states = 1ull << n_;
	const unsigned blocks = (unsigned)std::min<uint64_t>(2048, std::max<uint64_t>(1, states / 256));
	struct Buffers {
		hipStream_t s;
		std::vector<void *> held;
		void *get(size_t bytes)
		{
			held.push_back(device_alloc(bytes, s));
			return held.back();
		}
		~Buffers()
		{
			for (auto it = held.rbegin(); it != held.rend(); ++it)
				device_free(*it, s);
		}
	} buf{stream_, {}};
	double *D = nullptr, *P = nullptr, *sc = nullptr, *bsc = nullptr;
	uint32_t *bex = nullptr;
	std::vector<double> hs(blocks), host_scores;
	std::vector<uint32_t> he(blocks);
	// the lowest public index within the tie limit, from per-block lowest scores on the device
	auto pick_on_device = [&](int m) {
		BGT_HIP_CHECK(hipMemcpyAsync(hs.data(), bsc, sizeof(double) * blocks, hipMemcpyDeviceToHost, stream_));
		sync();
		double best = INFINITY;
		for (double v : hs)
			best = std::min(best, v);
		const double limit = best + 2.0 * lookahead_score_error(n_, m);
		launch_lookahead_pick(sc, bsc, n_, limit, bex, blocks, stream_);
		BGT_HIP_CHECK(hipMemcpyAsync(he.data(), bex, sizeof(uint32_t) * blocks, hipMemcpyDeviceToHost, stream_));
		sync();
		uint32_t best_c = 0xFFFFFFFFu;
		for (uint32_t v : he)
			best_c = std::min(best_c, v);
		if (best_c == 0xFFFFFFFFu || best_c == 0u)
			throw Error(BGT_HIP_INTERNAL, "lookahead found no candidate");
		return best_c;
	};
	const bool down = n_ >= kLookDownMin;
	if (down || steps > kLookClassMin) {
		sc = static_cast<double *>(buf.get(sizeof(double) * states));
		bsc = static_cast<double *>(buf.get(sizeof(double) * blocks));
		bex = static_cast<uint32_t *>(buf.get(sizeof(uint32_t) * blocks));
	}
	if (down) {
		D = static_cast<double *>(buf.get(sizeof(double) * states));
		launch_down_sums(single(), atoms_, storage_, D, stream_);
	}
	const uint32_t full = (uint32_t)(states - 1);
	for (int step = 1; step < steps; ++step) {
		const int m = (int)out.size();
		if (m >= kLookClassMin) {
			if (P == nullptr)
				P = static_cast<double *>(buf.get(sizeof(double) * look_class_rows(n_) * states));
			LookClasses lc{};
			lc.m = m;
			lc.n = n_;
			for (int i = 0; i < m; ++i)
				lc.sel[i] = rev(out[i]);
			launch_look_classes(single(), storage_, lc, P, sc, bsc, blocks, stream_);
			out.push_back(pick_on_device(m));
		} else if (down) {
			LookStep st{};
			st.n = n_;
			st.m = m;
			for (int k = 0; k < (1 << m); ++k) {
				uint32_t x = full;
				for (int i = 0; i < m; ++i)
					if ((k >> i) & 1)
						x &= rev(out[i]);
				st.inter[k] = x;
			}
			launch_lookahead_down(D, st, sc, bsc, blocks, stream_);
			out.push_back(pick_on_device(m));
		} else {
			host_scores.resize(states);
			BGT_HIP_CHECK(
				hipMemcpyAsync(selected_, out.data(), sizeof(uint32_t) * out.size(), hipMemcpyHostToDevice, stream_));
			launch_lookahead(single(), n_, storage_, selected_, m, scores_, stream_);
			BGT_HIP_CHECK(
				hipMemcpyAsync(host_scores.data(), scores_, sizeof(double) * states, hipMemcpyDeviceToHost, stream_));
			sync();
			double best = INFINITY;
			for (uint32_t c = 1; c < states; ++c)
				best = std::min(best, host_scores[c]);
			const double limit = best + 2.0 * lookahead_score_error(n_, m);
			uint32_t best_c = 1;
			while (best_c < states && !(host_scores[best_c] <= limit))
				++best_c;
			if (best_c == states)
				throw Error(BGT_HIP_INTERNAL, "lookahead found no candidate");
			out.push_back(best_c);
		}
	}
	return out;
}

// ------------------------------------------------------------------------ fused round

bgt_hip_selection DeviceLattice::round(uint32_t experiment, uint32_t response, const double *dil, double up,
									   double lo, double *masses_out)
{
	if (k_ != 1) {
		// multinomial lattices: Op-BBPA (auto_select's choice, simulation_helpers.cpp:40-45);
		// with tables on chip, the update and the histogram of the updated posterior are one
		// pass, and the atom masses come with the single-variant tables
		if (!bbpa_hist_path(n_, k_) && atoms_ <= kTileBits) {
			update(experiment, response, dil);
			classify(up, lo);
			if (masses_out)
				fetch_masses(masses_out);
			return select(kSelBbpa);
		}
		if (!bbpa_hist_path(n_, k_)) {
			// the transform path: the update fused into pass 1 (the general likelihood), then
			// the feature tables and the atom masses of the updated posterior from the final
			// pass, scaled by the new 1/D (as the binary round below)
			const double target = 1.0 / double(1 << k_);
			prepare_update(experiment, response, dil);
			ensure_z();
			Descs ds = single(); // pre-update descriptors (the final pass applies inv_out)
			clock_.begin(stream_, true);
			launch_zeta_pass1(ds, 1, atoms_, storage_, true, k_, dilution_, kObjGather, stream_, false);
			clock_.end(stream_, "round_pass1");
			commit_update();
			prepare_gather();
			{
				const LatDesc post = desc();
				ds.one.band_width = band_width(kSelBbpa);
				ds.one.gh_len = post.gh_len;
				ds.one.gx_n = post.gx_n;
				ds.one.gx_keep = post.gx_keep;
				std::memcpy(ds.one.gh_e, post.gh_e, sizeof(post.gh_e));
				std::memcpy(ds.one.gh_r, post.gh_r, sizeof(post.gh_r));
				ds.one.gh_pairs = post.gh_pairs;
				ds.one.gh_dil = post.gh_dil;
				ds.one.gexact = post.gexact;
			}
			clock_.begin(stream_);
			launch_zeta_final(ds, 1, atoms_, k_, storage_, kObjGather, false, true, target, stream_);
			clock_.end(stream_, "round_final");
			clock_.begin(stream_);
			launch_bbpa_score(ds, 1, n_, k_, target, stream_, true);
			clock_.end(stream_, "bbpa_score");
			collect_gather_band(target, &ds);
			fetch_head();
			classify_from_head(up, lo, masses_out);
			return finish_selection(kSelBbpa, static_cast<HostBox *>(host_box_)->result, true);
		}
		prepare_update(experiment, response, dil);
		const Descs ds = single(); // the new experiment's factors, the pre-update 1/D
		clock_.begin(stream_, true);
		launch_bbpa_hist(ds, 1, n_, k_, storage_, true, stream_, false);
		clock_.end(stream_, "round_hist");
		commit_update();
		// the tables of the pre-update descriptors (their inv_out slot receives 1/D), the band
		// and the exact program of the updated lattice
		prepare_gather();
		Descs df = ds;
		{
			const LatDesc post = desc();
			df.one.band_width = band_width(kSelBbpa);
			df.one.gh_len = post.gh_len;
			df.one.gx_n = post.gx_n;
			df.one.gx_keep = post.gx_keep;
			std::memcpy(df.one.gh_e, post.gh_e, sizeof(post.gh_e));
			std::memcpy(df.one.gh_r, post.gh_r, sizeof(post.gh_r));
			df.one.gh_pairs = post.gh_pairs;
			df.one.gh_dil = post.gh_dil;
			df.one.gexact = post.gexact;
		}
		clock_.begin(stream_);
		launch_bbpa_finish(df, 1, k_, true, 1.0 / double(1 << k_), stream_);
		clock_.end(stream_, "bbpa_finish");
		fetch_head();
		classify_from_head(up, lo, masses_out);
		return finish_selection(kSelBbpa, static_cast<HostBox *>(host_box_)->result, true);
	}
	if (atoms_ <= kTileBits) {
		update(experiment, response, dil);
		classify(up, lo);
		if (masses_out)
			fetch_masses(masses_out);
		return select(kSelBha);
	}
	prepare_update(experiment, response, dil);
	ensure_z();
	Descs ds = single(); // descriptors of the pre-update state (pass 2 applies inv_out)
	// a persistent pass 2 sums pass 1's partials itself (one launch and its gap fewer)
	const bool fold = persistent_pass2(atoms_, kObjHalving, false);
	// the light clock brackets one transform pass per round, pass 1 and pass 2 in turn
	const int light_pass = 1 + (int)(rounds_++ & 1u);
	clock_.begin(stream_, light_pass == 1);
	launch_zeta_pass1(ds, 1, atoms_, storage_, true, k_, dilution_, kObjHalving, stream_, fold);
	clock_.end(stream_, "round_pass1");
	commit_update();
	ds.one.band_width = band_width(kSelBha); // error band of the updated posterior
	clock_.begin(stream_, light_pass == 2);
	const unsigned entries = launch_zeta_final(ds, 1, atoms_, k_, storage_, kObjHalving, false, true, 0.5, stream_,
											   fold ? (1u << (atoms_ - kTileBits)) : 0u, [&] {
												   clock_.end(stream_, "round_pass2");
												   clock_.begin(stream_);
											   });
	clock_.end(stream_, "top2_compact");
	collect_band(ds, entries, kObjHalving, false, true, 0.5);
	launch_exact(ds);
	exact_.prepare(); // on the host while the device runs the round
	fetch_head();
	classify_from_head(up, lo, masses_out);
	return finish_selection(kSelBha, static_cast<HostBox *>(host_box_)->result, true);
}

SelResult DeviceLattice::head_result() const { return static_cast<const HostBox *>(host_box_)->result; }

void DeviceLattice::classify_from_head(double up, double lo, double *masses_out)
{
	// lattice_base.cpp:246-264
	const HostBox *box = static_cast<const HostBox *>(host_box_);
	for (int i = 0; i < atoms_; ++i) {
		const uint32_t placement = 1u << i;
		if ((pos_ | neg_) & placement)
			continue;
		if (box->masses[i] < lo)
			pos_ |= placement;
		else if (box->masses[i] > (1 - up))
			neg_ |= placement;
	}
	if (masses_out)
		std::memcpy(masses_out, box->masses, sizeof(double) * atoms_);
}

void DeviceLattice::error_bound(double *posterior_rel, double *mass_rel) const
{
	// Relative error of a stored posterior entry after `updates_` updates, against exact
	// arithmetic: one rounding for the prior, two per update (x * L, then * 1/D), plus the
	// common-scale error of the summed denominator. Masses add the transform's roundings.
	const double u64 = std::ldexp(1.0, -53);
	if (storage_ == Storage::fp64) {
		*posterior_rel = (2.0 * updates_ + 1.0 + atoms_ + proj_rel_) * u64;
		*mass_rel = *posterior_rel + 2.0 * atoms_ * u64;
	} else {
		const double u32 = std::ldexp(1.0, -24);
		*posterior_rel = (2.0 * updates_ + 2.0) * u32;
		*mass_rel = *posterior_rel + (three_pass_form(atoms_, storage_) ? 3.0 : 2.0) * u32;
	}
}

} // namespace bgt_hip
