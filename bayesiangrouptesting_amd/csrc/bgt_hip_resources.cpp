#include "bgt_hip_resources.hpp"

#include "bgt_hip_internal.hpp"

#include <cstdint>
#include <mutex>
#include <vector>

namespace bgt_hip {

namespace {

// Shared state, created on first use and never destroyed: no HIP call may run during
// static destruction, after the runtime has gone.
struct Resources {
	std::mutex m;
	int async_alloc = -1; // -1 unknown, 0 unsupported, 1 stream-ordered allocation in use
	std::vector<void *> pinned_free[24]; // by log2 of the block size
	std::vector<hipStream_t> streams; // the shared lattice streams, created on demand
	size_t next_stream = 0;
};

// Lattices share a bounded set of streams. Every stream holds runtime queue state and
// completion signals, of which the kernel driver grants a process a fixed number (4096 signal
// events), while a process can hold thousands of lattices at once (a tree level's children,
// lattices kept for re-use). Work that lattices put on one stream runs in the order it was
// enqueued, so sharing removes overlap between them, never ordering within one.
constexpr size_t kLatticeStreams = 32;

Resources &res()
{
	static Resources *r = new Resources();
	return *r;
}

int pinned_class(size_t bytes)
{
	int c = 8; // 256 bytes
	while ((size_t(1) << c) < bytes)
		++c;
	return c;
}

bool async_supported()
{
	Resources &r = res();
	std::lock_guard<std::mutex> lk(r.m);
	if (r.async_alloc < 0) {
		int dev = 0, ok = 0;
		r.async_alloc = 0;
		if (hipGetDevice(&dev) == hipSuccess &&
			hipDeviceGetAttribute(&ok, hipDeviceAttributeMemoryPoolsSupported, dev) == hipSuccess && ok) {
			hipMemPool_t pool;
			if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
				// keep freed memory in the pool instead of returning it at each synchronisation
				uint64_t keep = UINT64_MAX;
				if (hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep) == hipSuccess)
					r.async_alloc = 1;
			}
		}
	}
	return r.async_alloc == 1;
}

} // namespace

std::mutex &capture_gate()
{
	static std::mutex *m = new std::mutex();
	return *m;
}

void *device_alloc(size_t bytes, hipStream_t stream)
{
	std::lock_guard<std::mutex> gate(capture_gate());
	void *p = nullptr;
	if (async_supported())
		BGT_HIP_CHECK(hipMallocAsync(&p, bytes, stream));
	else
		BGT_HIP_CHECK(hipMalloc(&p, bytes));
	return p;
}

void device_free(void *p, hipStream_t stream)
{
	if (p == nullptr)
		return;
	std::lock_guard<std::mutex> gate(capture_gate());
	if (async_supported())
		(void)hipFreeAsync(p, stream);
	else
		(void)hipFree(p);
}

void *pinned_alloc(size_t bytes)
{
	const int c = pinned_class(bytes);
	Resources &r = res();
	{
		std::lock_guard<std::mutex> lk(r.m);
		auto &fl = r.pinned_free[c];
		if (!fl.empty()) {
			void *p = fl.back();
			fl.pop_back();
			return p;
		}
	}
	// a fresh chunk of at least 64 KiB, carved into blocks of this class
	const size_t block = size_t(1) << c;
	const size_t chunk = block < (size_t(1) << 16) ? (size_t(1) << 16) : block;
	void *base = nullptr;
	{
		std::lock_guard<std::mutex> gate(capture_gate());
		BGT_HIP_CHECK(hipHostMalloc(&base, chunk, hipHostMallocDefault));
	}
	std::lock_guard<std::mutex> lk(r.m);
	for (size_t off = block; off < chunk; off += block)
		r.pinned_free[c].push_back(static_cast<char *>(base) + off);
	return base;
}

void pinned_release(void *p, size_t bytes)
{
	if (p == nullptr)
		return;
	Resources &r = res();
	std::lock_guard<std::mutex> lk(r.m);
	r.pinned_free[pinned_class(bytes)].push_back(p);
}

hipStream_t stream_acquire()
{
	Resources &r = res();
	{
		std::lock_guard<std::mutex> lk(r.m);
		if (r.streams.size() >= kLatticeStreams)
			return r.streams[r.next_stream++ % r.streams.size()];
	}
	hipStream_t s = nullptr;
	{
		std::lock_guard<std::mutex> gate(capture_gate());
		BGT_HIP_CHECK(hipStreamCreate(&s));
	}
	std::lock_guard<std::mutex> lk(r.m);
	if (r.streams.size() >= kLatticeStreams) {
		// another thread filled the set meanwhile: use one of it, destroy the spare
		(void)hipStreamDestroy(s);
		return r.streams[r.next_stream++ % r.streams.size()];
	}
	r.streams.push_back(s);
	return s;
}

void stream_release(hipStream_t)
{
	// the streams are shared and live until the process ends
}

} // namespace bgt_hip
