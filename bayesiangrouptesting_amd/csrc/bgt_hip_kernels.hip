// gfx950 kernels of the BayesianGroupTesting HIP provider.
//
// Layout in HBM: one lattice = 2^A posterior entries (A = subjects * variants), entry S at
// offset S (the reference's state index, variant-major atoms). The posterior is kept
// unnormalised as y with a pending 1/D in a device scalar slot; every reader forms
// x = y * (1/D), which is exactly the value the reference stores after
// update_probs_in_place's `x * (1/D)` (lattice_base.cpp:230-244 under -ffast-math).
//
// The subset zeta transform (bbpa.cpp:497-512) runs on register tiles of 2^15 entries
// per 1024-thread workgroup: tile bits 0..5 across the lanes of a wave, bits 6..9 and
// 14 across registers, bits 10..13 across the 16 waves (one LDS transpose). Bits are
// always swept in increasing order, so every output is the same chain of two-operand
// adds as the reference's serial loop and the transform of identical inputs is
// bit-identical. Lattices with more than 15 atoms take two HBM passes: pass 1 sweeps
// bits 0..14 of contiguous tiles, pass 2 sweeps bits 15..A-1 on tiles of
// 2^(30-A) contiguous columns x 2^(A-15) strided rows and fuses the selection objective,
// the arg-min and the atom singletons, so the transform's output is never written.
#include <hip/hip_runtime.h>

#include "bgt_hip_internal.hpp"
#include "bgt_hip_tune.hpp"

#include <cmath>
#include <functional>
#include <cstdlib>
#include <string>

namespace bgt_hip {

namespace {

constexpr unsigned kNone = 0xFFFFFFFFu;

// A pointer every lane holds alike (a descriptor field), moved to scalar registers so
// that accesses take the scalar-base form with a 32-bit lane offset.
template <typename T>
__device__ __forceinline__ T *uniform_ptr(T *p)
{
	const unsigned long long b = (unsigned long long)p;
	const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
	const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
	return reinterpret_cast<T *>(((unsigned long long)hi << 32) | lo);
}

// A value every lane holds alike, moved to scalar registers.
__device__ __forceinline__ double uniform(double x)
{
	const unsigned long long b = (unsigned long long)__double_as_longlong(x);
	const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
	const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
	return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Lane l's value, for a uniform l (no memory access).
__device__ __forceinline__ uint32_t lane_u(uint32_t v, int l)
{
	return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}
__device__ __forceinline__ double lane_d(double v, int l)
{
	const unsigned long long b = (unsigned long long)__double_as_longlong(v);
	const uint32_t lo = lane_u((uint32_t)b, l), hi = lane_u((uint32_t)(b >> 32), l);
	return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Lattice buffers arrive through a descriptor struct, so the compiler cannot infer their
// address space; cast to global so loads and stores are global_* rather than flat_*.
template <typename T>
using gptr = __attribute__((address_space(1))) T *;
template <typename T>
__device__ __forceinline__ gptr<T> G(T *p)
{
	return (gptr<T>)p;
}

// Load/store at a wave-uniform base plus a 32-bit per-lane byte offset: the compiler keeps
// the base in SGPRs (global_* with saddr) instead of one 64-bit address per access.
template <typename T>
__device__ __forceinline__ T ld_u(const T *ubase, uint32_t byte_off)
{
	return *(gptr<const T>)(reinterpret_cast<const char *>(ubase) + byte_off);
}
template <typename T>
__device__ __forceinline__ void st_u(T *ubase, uint32_t byte_off, T v)
{
	*(gptr<T>)(reinterpret_cast<char *>(ubase) + byte_off) = v;
}
// The same as a streaming (non-temporal) store: for data not read again soon, such as the
// updated posterior, which the next round reads after a pass over the whole lattice.
template <typename T>
__device__ __forceinline__ void st_u_nt(T *ubase, uint32_t byte_off, T v)
{
	__builtin_nontemporal_store(v, (gptr<T>)(reinterpret_cast<char *>(ubase) + byte_off));
}
// A streaming (non-temporal) load at a wave-uniform base plus a lane offset: for data read
// once per pass, such as the posterior in pass 1.
template <typename T>
__device__ __forceinline__ T ld_u_nt(const T *ubase, uint32_t byte_off)
{
	return __builtin_nontemporal_load((gptr<const T>)(reinterpret_cast<const char *>(ubase) + byte_off));
}

struct Top2 {
	double d1, d2;
	uint32_t e1, e2;
};

// Objectives travel with a sign (the side of the target their mass lies on, see Cand) and
// are ordered by magnitude, then by experiment: the lowest experiment wins ties, as in the
// reference's ascending scans with strict comparisons.
__device__ __forceinline__ bool better(double a, uint32_t ea, double b, uint32_t eb)
{
	const double x = fabs(a), y = fabs(b);
	return x < y || (x == y && ea < eb);
}

__device__ __forceinline__ void top2_init(Top2 &t)
{
	t.d1 = INFINITY;
	t.d2 = INFINITY;
	t.e1 = kNone;
	t.e2 = kNone;
}

__device__ __forceinline__ void top2_insert(Top2 &t, double d, uint32_t e)
{
	if (better(d, e, t.d1, t.e1)) {
		t.d2 = t.d1;
		t.e2 = t.e1;
		t.d1 = d;
		t.e1 = e;
	} else if (better(d, e, t.d2, t.e2)) {
		t.d2 = d;
		t.e2 = e;
	}
}

__device__ __forceinline__ void top2_wave(Top2 &t)
{
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) {
		const double d1 = __shfl_xor(t.d1, o);
		const double d2 = __shfl_xor(t.d2, o);
		const uint32_t e1 = __shfl_xor(t.e1, o);
		const uint32_t e2 = __shfl_xor(t.e2, o);
		top2_insert(t, d1, e1);
		top2_insert(t, d2, e2);
	}
}

// Block-wide top-2; the result is valid in thread 0.
__device__ Top2 top2_block(Top2 t, Top2 *red)
{
	top2_wave(t);
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	if (lane == 0)
		red[w] = t;
	__syncthreads();
	Top2 r;
	top2_init(r);
	if (threadIdx.x == 0) {
		const int nw = (blockDim.x + 63) >> 6;
		for (int i = 0; i < nw; ++i) {
			top2_insert(r, red[i].d1, red[i].e1);
			top2_insert(r, red[i].d2, red[i].e2);
		}
	}
	__syncthreads();
	return r;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
	for (int o = 32; o > 0; o >>= 1)
		v += __shfl_xor(v, o);
	return v;
}

// Block-wide sum in a fixed order (lanes by xor tree, waves in index order); valid in thread 0.
__device__ double block_sum(double v, double *red)
{
	v = wave_sum(v);
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	if (lane == 0)
		red[w] = v;
	__syncthreads();
	double s = 0.0;
	if (threadIdx.x == 0) {
		const int nw = (blockDim.x + 63) >> 6;
		for (int i = 0; i < nw; ++i)
			s += red[i];
	}
	__syncthreads();
	return s;
}

// Tile order: the 8 XCDs receive workgroups round-robin, so give XCD x the contiguous
// tile range [x*G/8, (x+1)*G/8): neighbouring tiles (and the 64 B runs of pass 2) are
// then read by CUs that share one L2.
__device__ __forceinline__ unsigned xcd_tile(unsigned b, unsigned g)
{
	if (g < 8 || (g & 7))
		return b;
	return (b & 7) * (g >> 3) + (b >> 3);
}

// Likelihood L(E, r, S) (replicated_lattice.cpp:8-28): product over variants, in variant
// order, of a per-variant factor chosen by the number of pooled subjects missing from
// that variant's block. tab[v*(n+1)+idx] holds the factor; idx is the missing count for
// dilution and [missing > 0] otherwise.
__device__ __forceinline__ double likelihood(uint32_t s, uint32_t e, int n, int k, int len,
											 int dilution, const double *tab)
{
	double ret = 1.0;
	for (int v = 0; v < k; ++v) {
		const uint32_t sv = s >> (v * n);
		const int m = len - __popc(e & sv);
		const int idx = dilution ? m : (m != 0);
		ret *= tab[v * (n + 1) + idx];
	}
	return ret;
}

__device__ __forceinline__ void load_lik_table(const Descs &descs, const LatDesc &d, double *s_lik)
{
	const double *src = descs.many ? d.lik : descs.one_lik;
	const int cnt = d.k * (d.n + 1);
	for (int i = threadIdx.x; i < cnt; i += blockDim.x)
		s_lik[i] = src[i];
	__syncthreads();
}

// ------------------------------------------------------------------ register tile zeta

template <int B>
__device__ __forceinline__ void bfly_reg(double (&v)[kTileRegs])
{
#pragma unroll
	for (int j = 0; j < kTileRegs; ++j)
		if (!(j & (1 << B)))
			v[j] += v[j | (1 << B)];
}

// Load layout: tile index i = lane | jlo<<6 | wave<<10 | jhi<<14 (j = jlo + 16*jhi).
__device__ __forceinline__ uint32_t tile_index_a(int lane, int wave, int j)
{
	return (uint32_t)lane | ((uint32_t)(j & 15) << 6) | ((uint32_t)wave << 10) | ((uint32_t)(j >> 4) << 14);
}

// Output layout after zeta_tile: lane bits 0..5 hold tile bits 4,5,2,3,0,1; wave bits hold
// tile bits 6..9; jlo holds 10..13 and jhi bit 14. For a fixed register a wave still
// covers one contiguous run of 64 states, so stores stay fully coalesced.
__device__ __forceinline__ uint32_t lane_bits_f(int lane)
{
	const uint32_t l = (uint32_t)lane;
	return ((l & 1u) << 4) | (((l >> 1) & 1u) << 5) | (l & 0xCu) | ((l >> 4) & 3u);
}
__device__ __forceinline__ uint32_t tile_index_f(int lane, int wave, int j)
{
	return lane_bits_f(lane) | ((uint32_t)wave << 6) | ((uint32_t)(j & 15) << 10) | ((uint32_t)(j >> 4) << 14);
}

constexpr int kWaveStride = 65; // padded rows of the wave-local transpose (bank-conflict free)
constexpr int kTileLds = 16 * 16 * kWaveStride; // doubles; also holds the 16x16x64 block transpose

__device__ __forceinline__ void wave_sync()
{
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Swap lane bits 0..3 with register bits 0..3 (jlo) inside one wave, through its own LDS
// rows: thread lane L, register jlo' receives the value lane (L & 48 | jlo') held in
// register (L & 15).
__device__ __forceinline__ void wave_transpose(double (&v)[kTileRegs], double *lw, int lane)
{
#pragma unroll
	for (int h = 0; h < 2; ++h) {
#pragma unroll
		for (int j = 0; j < 16; ++j)
			lw[j * kWaveStride + lane] = v[j + 16 * h];
		wave_sync();
		const int rd = (lane & 15) * kWaveStride + (lane & 48);
#pragma unroll
		for (int j = 0; j < 16; ++j)
			v[j + 16 * h] = lw[rd + j];
		wave_sync();
	}
}

// Exchange lane bit 4 (W32 = false, v_permlane16_swap) or lane bit 5 (W32 = true,
// v_permlane32_swap) with register bit RB: afterwards the pair (j, j | 1<<RB) of every
// lane holds two states that differ in the former lane bit.
template <bool W32, int RB>
__device__ __forceinline__ void perm_swap(double (&v)[kTileRegs])
{
#pragma unroll
	for (int j = 0; j < kTileRegs; ++j) {
		if (j & (1 << RB))
			continue;
		const unsigned long long a = (unsigned long long)__double_as_longlong(v[j]);
		const unsigned long long b = (unsigned long long)__double_as_longlong(v[j | (1 << RB)]);
		const unsigned alo = (unsigned)a, ahi = (unsigned)(a >> 32);
		const unsigned blo = (unsigned)b, bhi = (unsigned)(b >> 32);
		unsigned nlo0, nlo1, nhi0, nhi1;
		if (W32) {
			const auto lo = __builtin_amdgcn_permlane32_swap(alo, blo, false, false);
			const auto hi = __builtin_amdgcn_permlane32_swap(ahi, bhi, false, false);
			nlo0 = lo[0]; nlo1 = lo[1]; nhi0 = hi[0]; nhi1 = hi[1];
		} else {
			const auto lo = __builtin_amdgcn_permlane16_swap(alo, blo, false, false);
			const auto hi = __builtin_amdgcn_permlane16_swap(ahi, bhi, false, false);
			nlo0 = lo[0]; nlo1 = lo[1]; nhi0 = hi[0]; nhi1 = hi[1];
		}
		v[j] = __longlong_as_double((long long)(((unsigned long long)nhi0 << 32) | nlo0));
		v[j | (1 << RB)] = __longlong_as_double((long long)(((unsigned long long)nhi1 << 32) | nlo1));
	}
}

// Swap wave bits with register bits 0..3 across the workgroup (two halves through LDS).
__device__ __forceinline__ void block_transpose(double (&v)[kTileRegs], double *lds, int lane, int wave)
{
#pragma unroll
	for (int h = 0; h < 2; ++h) {
#pragma unroll
		for (int j = 0; j < 16; ++j)
			lds[((wave * 16 + j) << 6) + lane] = v[j + 16 * h];
		__syncthreads();
#pragma unroll
		for (int j = 0; j < 16; ++j)
			v[j + 16 * h] = lds[((j * 16 + wave) << 6) + lane];
		__syncthreads();
	}
}

// Sweep tile bits q0..14 in increasing order; every butterfly is register-local.
// Input in the load layout, output in the layout of tile_index_f.
__device__ __forceinline__ void zeta_tile(double (&v)[kTileRegs], int q0, int lane, int wave, double *lds)
{
	double *lw = lds + wave * (16 * kWaveStride);
	wave_transpose(v, lw, lane); // registers: tile bits 0..3
	if (q0 <= 0) bfly_reg<0>(v);
	if (q0 <= 1) bfly_reg<1>(v);
	if (q0 <= 2) bfly_reg<2>(v);
	if (q0 <= 3) bfly_reg<3>(v);
	perm_swap<false, 0>(v); // register bit 0: tile bit 4
	if (q0 <= 4) bfly_reg<0>(v);
	perm_swap<true, 1>(v); // register bit 1: tile bit 5
	if (q0 <= 5) bfly_reg<1>(v);
	wave_transpose(v, lw, lane); // registers: tile bits 6..9
	if (q0 <= 6) bfly_reg<0>(v);
	if (q0 <= 7) bfly_reg<1>(v);
	if (q0 <= 8) bfly_reg<2>(v);
	if (q0 <= 9) bfly_reg<3>(v);
	__syncthreads(); // the block transpose reuses the wave rows
	block_transpose(v, lds, lane, wave); // registers: tile bits 10..13
	if (q0 <= 10) bfly_reg<0>(v);
	if (q0 <= 11) bfly_reg<1>(v);
	if (q0 <= 12) bfly_reg<2>(v);
	if (q0 <= 13) bfly_reg<3>(v);
	bfly_reg<4>(v);
}

// --------------------------------------------------------------- selection objective

// Op-BHA: M - 1/2 (objective |M - 1/2|, bbpa.cpp:537-551). Binary BBPA:
// |P0 - t| + |P1 - t| with P0 = M, P1 = T - M (bbpa.cpp:201-220), signed by P0 - P1.
__device__ __forceinline__ double signed_objective(int objective, double m, double total, double target)
{
	if (objective == kObjHalving)
		return m - 0.5;
	const double negative = total - m;
	return copysign(fabs(m - target) + fabs(negative - target), m - negative);
}

struct ObjCtx {
	int objective;
	uint32_t active;
	double target, total;
	int ns, k; // subjects, variants (gather)
	double *ftab;
	double *masses;
};

__device__ __forceinline__ void consume(const ObjCtx &c, uint32_t g, double m, Top2 &t)
{
	if (__popc(g) == 1)
		c.masses[__ffs(g) - 1] = m;
	if (c.objective == kObjHalving || c.objective == kObjBinaryBbpa) {
		if ((g & ~c.active) == 0)
			top2_insert(t, signed_objective(c.objective, m, c.total, c.target), g);
	} else if (c.objective == kObjGather) {
		// F_B(E) = M(spread_B(E)): g's non-zero variant blocks all equal E, B = those blocks
		const uint32_t mask = (1u << c.ns) - 1u;
		const uint32_t ec = 1u << c.ns;
		if (g == 0) {
			for (int b = 1; b < (1 << c.k); ++b)
				c.ftab[(size_t)b * ec] = m;
			return;
		}
		uint32_t e = 0, bset = 0;
		bool ok = true;
		for (int v = 0; v < c.k; ++v) {
			const uint32_t blk = (g >> (v * c.ns)) & mask;
			if (blk) {
				if (e == 0)
					e = blk;
				else if (blk != e)
					ok = false;
				bset |= 1u << v;
			}
		}
		if (ok)
			c.ftab[(size_t)bset * ec + e] = m;
	}
}

// Each workgroup publishes its top-2; k_finalize_top2 merges them after the kernel ends
// (a kernel boundary orders the writes without a device-scope fence per workgroup, which
// on a multi-XCD part would write back and invalidate the L2 of every XCD).
__device__ __forceinline__ void publish_top2(const LatDesc &d, unsigned slot, Top2 t, Top2 *red)
{
	Top2 b = top2_block(t, red);
	if (threadIdx.x == 0)
		reinterpret_cast<Top2 *>(d.tops)[slot] = b;
}

// Per-thread top-2 for the band collection (exact rescoring of near ties): thread `slot`
// of the launch writes entries 2*slot and 2*slot+1.
__device__ __forceinline__ void write_cand(const LatDesc &d, unsigned slot, const Top2 &t)
{
	if (d.cand == nullptr)
		return;
	Cand a, b;
	a.d = t.d1;
	a.e = t.e1;
	a.flags = 0;
	b.d = t.d2;
	b.e = t.e2;
	b.flags = 0;
	d.cand[2 * (size_t)slot] = a;
	d.cand[2 * (size_t)slot + 1] = b;
}

// Append the lanes' candidates with `in` set to buf[0..cap), counting in *counter: one
// atomic per wave. Once the buffer has overflowed only the count matters, so the atomic is
// skipped (a posterior so concentrated that every objective sits at 1/2 would otherwise
// serialise on one address). Every lane of the wave must call this.
__device__ __forceinline__ void cand_append(Cand *buf, unsigned *counter, unsigned cap, bool in, double sd,
											uint32_t e)
{
	const unsigned long long mask = __ballot(in);
	if (mask == 0ull)
		return;
	const int lane = threadIdx.x & 63;
	const int leader = __ffsll((long long)mask) - 1;
	unsigned base = 0;
	if (lane == leader) {
		const unsigned seen = __hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		base = seen > cap ? seen : atomicAdd(counter, (unsigned)__popcll(mask));
	}
	base = __shfl(base, leader);
	if (in && base <= cap) {
		const unsigned idx = base + (unsigned)__popcll(mask & ((1ull << lane) - 1ull));
		if (idx < cap) {
			Cand c;
			c.d = sd;
			c.e = e;
			c.flags = kCandSided;
			buf[idx] = c;
		}
	}
}

__device__ __forceinline__ void band_append(const LatDesc &d, bool in, double sd, uint32_t e)
{
	cand_append(d.band, &d.counter[0], d.band_cap, in, sd, e);
}

// Pass 2 of 21..30 atoms collects its band while it scans: counter[kRawCount] counts the
// entries it appended to the `cand` buffer, and eight slots from counter[kBestSlots] hold
// running bests of |objective|, one per XCD. A slot holds the high word of a non-negative
// double plus one: the double that key decodes to (key << 32) is at least the value
// recorded, and keys order like the values. Pass 1 resets both.
constexpr int kRawCount = 2;
constexpr int kBestSlots = 16;
constexpr unsigned kKeyInf = 0x7FF00000u; // +inf
__device__ __forceinline__ unsigned *best_slots(const LatDesc &d) { return d.counter + kBestSlots; }
__device__ __forceinline__ unsigned best_key(double v)
{
	return (unsigned)((unsigned long long)__double_as_longlong(v) >> 32) + 1u;
}
__device__ __forceinline__ double key_value(unsigned key)
{
	return key >= kKeyInf ? INFINITY : __longlong_as_double((long long)((unsigned long long)key << 32));
}

// Rescan launches (collect mode) transform only the workgroups flagged by k_band_collect,
// and nothing at all once the band has overflowed.
__device__ __forceinline__ bool rescan_skip(const LatDesc &d, unsigned group)
{
	return d.band == nullptr || d.wgflag[group] == 0u ||
		   __hip_atomic_load(&d.counter[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > d.band_cap;
}

// ----------------------------------------------------------------------------- kernels

template <typename TP>
__global__ __launch_bounds__(256) void k_prior(Descs descs, int atoms, const double *pairs, uint64_t states)
{
	// lattice_base.cpp:184-208: P(S) = prod over atoms a = 0..A-1, in order, from 1.0
	const LatDesc &d = descs.many ? descs.many[blockIdx.y] : descs.one;
	__shared__ double s_pairs[2 * kMaxAtoms];
	if (threadIdx.x < 2 * kMaxAtoms)
		s_pairs[threadIdx.x] = threadIdx.x < 2u * atoms ? pairs[threadIdx.x] : 1.0;
	__syncthreads();
	gptr<TP> post = G(static_cast<TP *>(d.post));
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < states; s += stride) {
		double prob = 1.0;
		for (int a = 0; a < atoms; ++a)
			prob *= s_pairs[2 * a + ((s >> a) & 1u)];
		post[s] = static_cast<TP>(prob);
	}
}

template <typename TP>
__global__ __launch_bounds__(256) void k_update(Descs descs, uint64_t states)
{
	// lattice_base.cpp:230-244: y' = x * L(E, r, S); D = sum y'; the pending 1/D is applied
	// by the next reader (x' = y' * (1/D)). With d.src_post set the input x is another
	// lattice's posterior (update_probs, the out-of-place form the tree uses,
	// lattice_base.cpp:210-228).
	const LatDesc &d = descs.many ? descs.many[blockIdx.y] : descs.one;
	const unsigned groups = descs.ragged ? update_groups(d.atoms) : gridDim.x;
	if (blockIdx.x >= groups)
		return;
	if (descs.ragged)
		states = 1ull << d.atoms;
	__shared__ double s_lik[kMaxVariants * (kMaxSubjects + 1)];
	__shared__ double red[16];
	load_lik_table(descs, d, s_lik);
	gptr<TP> post = G(static_cast<TP *>(d.post));
	gptr<const TP> src = G(static_cast<const TP *>(d.src_post ? d.src_post : d.post));
	const double inv = d.src_post ? d.src_scal[d.src_inv] : d.scal[d.inv_in];
	double acc = 0.0;
	const uint64_t stride = (uint64_t)groups * blockDim.x;
	for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < states; s += stride) {
		const double x = static_cast<double>(src[s]) * inv;
		const TP y = static_cast<TP>(x * likelihood((uint32_t)s, d.e, d.n, d.k, d.len, d.dilution, s_lik));
		post[s] = y;
		acc += static_cast<double>(y);
	}
	const double bs = block_sum(acc, red);
	if (threadIdx.x == 0)
		d.part[blockIdx.x] = bs;
}

// Pass 1: sweep bits 0..14 of contiguous 2^15 tiles. Input is x = y * inv; with FUSED the
// update y' = x * L is applied first (stored back in place) and the transform runs on y'.
// LIK: 0 = general table lookup, 1 = binary non-dilution (two factors chosen by E subset S).
// The intermediate is written in state order.
// The 32 posterior entries of one thread in pass-1 tile `tile` (load layout).
template <typename TP>
__device__ __forceinline__ void pass1_load(const LatDesc &d, unsigned tile, int lane, int wave, TP (&raw)[kTileRegs])
{
	// wave-uniform tile base plus one 32-bit lane offset: no per-register address VGPRs
	const TP *tb = uniform_ptr(static_cast<const TP *>(d.post)) + ((uint64_t)tile << kTileBits);
	const uint32_t lob = ((uint32_t)lane | ((uint32_t)wave << 10)) * (uint32_t)sizeof(TP);
#pragma unroll
	for (int j = 0; j < kTileRegs; ++j)
		raw[j] = ld_u_nt(tb + (((j & 15) << 6) | ((j >> 4) << 14)), lob); // read once per round
}

// Pass-1 variants that run persistent with a prefetch of the next tile (FP32 storage, 32
// registers per thread): all but the fused general likelihood on the blocked intermediate,
// which would spill (its table lookups beside the blocked store's offsets). The launcher
// picks the grid by the same rule.
template <typename TP, bool FUSED, int LIK, bool ZB>
constexpr bool pass1_persistent()
{
	return sizeof(TP) == 4 && (!FUSED || LIK == 1 || !ZB);
}

// One pass-1 tile from its loaded entries: the fused update (stored in place), the sweep
// of bits 0..14, the intermediate z and the tile total.
template <typename TP, typename TZ, bool FUSED, int LIK, bool ZB>
__device__ __forceinline__ void pass1_tile(const LatDesc &d, unsigned tile, const TP (&raw)[kTileRegs], int lane,
										   int wave, double inv, double f_in, double f_out, const double *s_lik,
										   double *lds)
{
	const uint64_t base = (uint64_t)tile << kTileBits;
	TP *tb = uniform_ptr(static_cast<TP *>(d.post)) + base;
	const uint32_t lob = ((uint32_t)lane | ((uint32_t)wave << 10)) * (uint32_t)sizeof(TP);
	double v[kTileRegs];
#pragma unroll
	for (int j = 0; j < kTileRegs; ++j) {
		const uint32_t off = ((uint32_t)(j & 15) << 6) | ((uint32_t)(j >> 4) << 14);
		double x = static_cast<double>(raw[j]) * inv;
		if (FUSED) {
			const uint32_t g = (uint32_t)(base + tile_index_a(lane, wave, j));
			double lk;
			if (LIK == 1)
				lk = ((d.e & g) == d.e) ? f_in : f_out;
			else
				lk = likelihood(g, d.e, d.n, d.k, d.len, d.dilution, s_lik);
			const TP y = static_cast<TP>(x * lk);
			st_u_nt(tb + off, lob, y); // read again only by the next round
			x = static_cast<double>(y);
		}
		v[j] = x;
	}
	zeta_tile(v, 0, lane, wave, lds);
	// FP64 z streams past the caches (measured faster in both passes); FP32 z stays cached,
	// as pass 2 reads it next (streaming it moved time from pass 2 to pass 1)
	if constexpr (ZB) {
		// blocked form (FP32): row `tile`, column c = lane_bits_f + 64 wave + 1024 jlo + 16384 jhi
		// at (c >> 5) * pitch_b + (tile << 5) + (c & 31); the lane offsets stay below 32
		// pitches, 2^27 bytes at 30 atoms
		static_assert(sizeof(TZ) == 4, "the blocked form is FP32");
		const uint32_t lbf = lane_bits_f(lane);
		const int atoms = __builtin_amdgcn_readfirstlane(d.atoms);
		const uint64_t pitch_b = ((uint64_t)1 << (atoms - kTileBits + kZBlockBits)) + kZBlockSkew;
		TZ *zt = uniform_ptr(static_cast<TZ *>(d.z)) + ((uint64_t)tile << kZBlockBits);
		const uint32_t pitch_bytes = (uint32_t)pitch_b * (uint32_t)sizeof(TZ);
		const uint32_t loz = (lbf & 31u) * (uint32_t)sizeof(TZ) + ((lbf >> 5) + ((uint32_t)wave << 1)) * pitch_bytes;
#pragma unroll
		for (int j = 0; j < kTileRegs; ++j)
			st_u(uniform_ptr(zt + (uint64_t)(((j & 15) << 5) | ((j >> 4) << 9)) * pitch_b), loz, static_cast<TZ>(v[j]));
	} else {
		TZ *zt = uniform_ptr(static_cast<TZ *>(d.z)) + (uint64_t)tile * z_pitch<TZ>();
		const uint32_t loz = (lane_bits_f(lane) + ((uint32_t)wave << 6)) * (uint32_t)sizeof(TZ);
#pragma unroll
		for (int j = 0; j < kTileRegs; ++j) {
			if constexpr (sizeof(TZ) == 8)
				st_u_nt(zt + (((j & 15) << 10) | ((j >> 4) << 14)), loz, static_cast<TZ>(v[j]));
			else
				st_u(zt + (((j & 15) << 10) | ((j >> 4) << 14)), loz, static_cast<TZ>(v[j]));
		}
	}
	// after the full 15-bit sweep, tile index 0 (thread 0, register 0) holds the tile's total
	if (threadIdx.x == 0)
		d.part[tile] = v[0];
}

// Pass 1. With one workgroup per tile (gridDim.x = tile count) each workgroup loads, then
// transforms, its tile. With fewer workgroups (persistent, a multiple of 8: FP32 storage,
// where the next tile's entries fit in the registers the transform leaves free) workgroup
// b serves the contiguous tile range of XCD b % 8 in turn and loads its next tile while it
// transforms the current one; the workgroup barriers of the transform wait for LDS only,
// so those loads stay in flight across them.
// ZB: z in the blocked form (FP32 only).
// ragged (lattices of at most kRaggedMaxAtoms atoms): one workgroup per tile of each lattice.
template <typename TP, typename TZ, bool FUSED, int LIK, bool ZB = false>
__global__ __launch_bounds__(kTileThreads) void k_zeta_pass1(Descs descs)
{
	const LatDesc &d = descs.many ? descs.many[blockIdx.y] : descs.one;
	if (descs.ragged && blockIdx.x >= tile_groups(d.atoms))
		return;
	__shared__ double lds[kTileLds];
	__shared__ double s_lik[kMaxVariants * (kMaxSubjects + 1)];
	if (FUSED)
		load_lik_table(descs, d, s_lik);
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	// uniform by construction; said so, the tile loop stays scalar
	const unsigned tiles = 1u << (__builtin_amdgcn_readfirstlane(d.atoms) - kTileBits);
	if (blockIdx.x == 0 && d.band != nullptr) {
		// the band collection of the pass 2 that follows (raw appends, running bests) and of
		// the compaction after it (band count and its companion counter)
		if (threadIdx.x < 8)
			best_slots(d)[threadIdx.x] = kKeyInf;
		else if (threadIdx.x == 8)
			d.counter[kRawCount] = 0u;
		else if (threadIdx.x == 9)
			d.counter[0] = 0u;
		else if (threadIdx.x == 10)
			d.counter[1] = 0u;
	}
	const double inv = uniform(d.scal[d.inv_in]);
	const double f_in = FUSED ? uniform(s_lik[0]) : 0.0, f_out = FUSED ? uniform(s_lik[1]) : 0.0;
	if (!pass1_persistent<TP, FUSED, LIK, ZB>() || gridDim.x >= tiles) {
		const unsigned tile = xcd_tile(blockIdx.x, tiles);
		TP raw[kTileRegs];
		pass1_load(d, tile, lane, wave, raw);
		pass1_tile<TP, TZ, FUSED, LIK, ZB>(d, tile, raw, lane, wave, inv, f_in, f_out, s_lik, lds);
		return;
	}
	if constexpr (pass1_persistent<TP, FUSED, LIK, ZB>()) {
		const unsigned wg_per_xcd = gridDim.x >> 3;
		const unsigned per_xcd = tiles >> 3;
		const unsigned xcd = blockIdx.x & 7u;
		const unsigned first = xcd * per_xcd + (blockIdx.x >> 3), last = (xcd + 1) * per_xcd;
		TP next[kTileRegs];
		if (first < last)
			pass1_load(d, first, lane, wave, next);
		for (unsigned T = first; T < last; T += wg_per_xcd) {
			TP cur[kTileRegs];
#pragma unroll
			for (int j = 0; j < kTileRegs; ++j)
				cur[j] = next[j];
			if (T + wg_per_xcd < last)
				pass1_load(d, T + wg_per_xcd, lane, wave, next);
			pass1_tile<TP, TZ, FUSED, LIK, ZB>(d, T, cur, lane, wave, inv, f_in, f_out, s_lik, lds);
		}
	}
}

// Final pass: sweep tile bits q0..14 and evaluate the objective on every output entry.
// Tile index i maps to state g = tile*c + (i & (c-1)) + ((i >> logc) << 15), c = 2^logc.
// FROM_POST reads x = y * inv from the posterior (single-pass lattices, A == 15);
// otherwise it reads the pass-1 intermediate, scaled by inv_out when `scaled`.
// With `collect` (a rescan after k_band_collect) only flagged tiles run, and every entry
// within the band of the final best is appended to the band; nothing else is written.
template <typename TIN, bool FROM_POST>
__global__ __launch_bounds__(kTileThreads) void k_zeta_final(Descs descs, int logc, int q0,
															 int objective, int scaled, double target, int collect)
{
	const LatDesc &d = descs.many ? descs.many[blockIdx.y] : descs.one;
	// ragged: each lattice's own tiles and column split (15..20 atoms)
	const unsigned tiles = descs.ragged ? (FROM_POST ? 1u : tile_groups(d.atoms)) : gridDim.x;
	if (blockIdx.x >= tiles)
		return;
	if (descs.ragged && !FROM_POST) {
		logc = 2 * kTileBits - d.atoms;
		q0 = logc;
	}
	__shared__ double lds[kTileLds];
	__shared__ Top2 red2[16];
	__shared__ double red[16];
	__shared__ double s_total;
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const unsigned tile = xcd_tile(blockIdx.x, tiles);
	if (collect && rescan_skip(d, tile))
		return;
	const uint32_t cmask = (1u << logc) - 1u;
	const uint64_t cbase = (uint64_t)tile << logc;
	gptr<const TIN> in = G(FROM_POST ? static_cast<const TIN *>(d.post) : static_cast<const TIN *>(d.z));
	const double inv = FROM_POST ? d.scal[d.inv_in] : 1.0;
	double v[kTileRegs];
#pragma unroll
	for (int j = 0; j < kTileRegs; ++j) {
		const uint32_t i = tile_index_a(lane, wave, j);
		// state cbase + (i & cmask) + (row << 15); z rows are z_pitch apart
		const uint64_t g = FROM_POST ? cbase + (i & cmask) + ((uint64_t)(i >> logc) << kTileBits)
									 : cbase + (i & cmask) + (uint64_t)(i >> logc) * z_pitch<TIN>();
		v[j] = static_cast<double>(in[g]) * inv;
	}
	ObjCtx c;
	c.objective = objective;
	c.active = d.active;
	c.target = target;
	c.ns = d.n;
	c.k = d.k;
	c.ftab = d.ftab;
	c.masses = d.masses;
	zeta_tile(v, q0, lane, wave, lds);
	if (collect) {
		const double sc = scaled ? d.scal[d.inv_out] : 1.0;
		const double total = d.scal[kScalTotal] * sc;
		const double lim = fabs(d.result->d1) + d.band_width;
#pragma unroll
		for (int j = 0; j < kTileRegs; ++j) {
			const uint32_t i = tile_index_f(lane, wave, j);
			const uint32_t g = (uint32_t)(cbase + (i & cmask) + ((uint64_t)(i >> logc) << kTileBits));
			const double sd = signed_objective(objective, scaled ? v[j] * sc : v[j], total, target);
			band_append(d, (g & ~d.active) == 0u && fabs(sd) <= lim, sd, g);
		}
		return;
	}
	if (FROM_POST) {
		// the whole lattice is this one tile: its total is the transform at index 0
		if (threadIdx.x == 0) {
			s_total = v[0];
			d.scal[kScalTotal] = v[0];
		}
		__syncthreads();
		c.total = s_total;
	} else {
		c.total = d.scal[kScalTotal];
	}
	const double scale = scaled ? d.scal[d.inv_out] : 1.0;
	if (scaled)
		c.total = c.total * scale;
	Top2 t;
	top2_init(t);
#pragma unroll
	for (int j = 0; j < kTileRegs; ++j) {
		const uint32_t i = tile_index_f(lane, wave, j);
		const uint32_t g = (uint32_t)(cbase + (i & cmask) + ((uint64_t)(i >> logc) << kTileBits));
		const double m = scaled ? v[j] * scale : v[j];
		consume(c, g, m, t);
	}
	write_cand(d, tile * blockDim.x + threadIdx.x, t);
	publish_top2(d, tile, t, red2);
}

// ------------------------------------------- pass 2 for lattices of 21..30 atoms (n >= 21)
//
// A pass-2 tile is 2^LC contiguous columns (state bits 0..LC-1) by 2^(15-LC) rows (state
// bits 15..A-1) of one column block (state bits LC..14 = tile id), LC = 30 - A; only the
// row bits are swept. Tile bit t (t < LC: column bit t; t >= LC: state bit t - LC + 15) is
// placed at load time as:
//   lanes:     column bits 0..min(LC,6)-1, then (LC < 6) tile bits 9+LC..14
//   registers: tile bits LC..LC+4 (the five lowest row bits)
//   waves:     tile bits LC+5..LC+8 (LC <= 6), else column bits 6..LC-1 and rows 11+w
// so a wave load reads runs of 2^min(LC,6) contiguous entries and the sweep needs one block
// transpose (waves <-> registers) plus, for LC <= 3, one wave transpose, and permlane swaps
// for lane bits 4 and 5. Bits are still swept in increasing order.
namespace p2 {

struct Map {
	int lane[6];
	int wave[4];
	int reg[5];
};

constexpr Map initial_map(int lc)
{
	Map m{};
	for (int l = 0; l < 6; ++l)
		m.lane[l] = (lc >= 6 || l < lc) ? l : 9 + l;
	for (int w = 0; w < 4; ++w)
		m.wave[w] = lc <= 6 ? lc + 5 + w : (w < lc - 6 ? 6 + w : 11 + w);
	for (int r = 0; r < 5; ++r)
		m.reg[r] = lc + r;
	return m;
}

// after the block transpose (register bits 0..3 <-> wave bits 0..3)
constexpr Map block_map(int lc)
{
	Map m = initial_map(lc);
	for (int b = 0; b < 4; ++b) {
		const int t = m.reg[b];
		m.reg[b] = m.wave[b];
		m.wave[b] = t;
	}
	return m;
}

// after the wave transpose (lane bits 0..3 <-> register bits 0..3), LC <= 3
constexpr Map wave_map(int lc)
{
	Map m = block_map(lc);
	if (lc <= 3)
		for (int b = 0; b < 4; ++b) {
			const int t = m.reg[b];
			m.reg[b] = m.lane[b];
			m.lane[b] = t;
		}
	return m;
}

// after the permlane swaps (lane bit 4 <-> register bit 0 for LC <= 4, lane bit 5 <->
// register bit 1 for LC <= 5): the layout the objective is evaluated in
constexpr Map final_map(int lc)
{
	Map m = wave_map(lc);
	if (lc <= 4) {
		const int t = m.reg[0];
		m.reg[0] = m.lane[4];
		m.lane[4] = t;
	}
	if (lc <= 5) {
		const int t = m.reg[1];
		m.reg[1] = m.lane[5];
		m.lane[5] = t;
	}
	return m;
}

// offset of tile bit t in the state index, relative to the tile's first column
constexpr uint32_t weight(int lc, int t) { return t < lc ? (1u << t) : (1u << (t - lc + kTileBits)); }
// offset of tile bit t in the intermediate z (rows are `pitch` entries apart)
constexpr uint64_t zweight(int lc, int t, uint32_t pitch) { return t < lc ? (1ull << t) : ((uint64_t)pitch << (t - lc)); }

constexpr uint32_t reg_offset(const Map &m, int lc, int j)
{
	uint32_t o = 0;
	for (int r = 0; r < 5; ++r)
		if ((j >> r) & 1)
			o += weight(lc, m.reg[r]);
	return o;
}

// registers in increasing state order (the scan keeps the first of equal objectives,
// which is then the lowest experiment, as the reference's serial scan does)
struct Order {
	int j[kTileRegs];
};
constexpr Order scan_order(int lc)
{
	Order o{};
	const Map m = final_map(lc);
	uint32_t key[kTileRegs] = {};
	for (int j = 0; j < kTileRegs; ++j) {
		o.j[j] = j;
		key[j] = reg_offset(m, lc, j);
	}
	for (int a = 1; a < kTileRegs; ++a)
		for (int b = a; b > 0 && key[b - 1] > key[b]; --b) {
			const uint32_t tk = key[b];
			key[b] = key[b - 1];
			key[b - 1] = tk;
			const int tj = o.j[b];
			o.j[b] = o.j[b - 1];
			o.j[b - 1] = tj;
		}
	return o;
}

// state-index offset of (lane, wave) in the layout the objective is evaluated in
template <int LC>
__device__ __forceinline__ uint32_t thread_state_offset(int lane, int wave)
{
	constexpr Map m = final_map(LC);
	uint32_t o = 0;
#pragma unroll
	for (int l = 0; l < 6; ++l)
		o += ((lane >> l) & 1) ? weight(LC, m.lane[l]) : 0u;
#pragma unroll
	for (int w = 0; w < 4; ++w)
		o += ((wave >> w) & 1) ? weight(LC, m.wave[w]) : 0u;
	return o;
}

// z offset of (lane, wave) in the load layout
template <int LC, uint32_t PITCH>
__device__ __forceinline__ uint64_t thread_z_offset(int lane, int wave)
{
	constexpr Map m = initial_map(LC);
	uint64_t o = 0;
#pragma unroll
	for (int l = 0; l < 6; ++l)
		o += ((lane >> l) & 1) ? zweight(LC, m.lane[l], PITCH) : 0ull;
#pragma unroll
	for (int w = 0; w < 4; ++w)
		o += ((wave >> w) & 1) ? zweight(LC, m.wave[w], PITCH) : 0ull;
	return o;
}

// largest z offset of any (lane, wave) in the load layout
constexpr uint64_t max_thread_z_offset(int lc, uint32_t pitch)
{
	const Map m = initial_map(lc);
	uint64_t o = 0;
	for (int l = 0; l < 6; ++l)
		o += zweight(lc, m.lane[l], pitch);
	for (int w = 0; w < 4; ++w)
		o += zweight(lc, m.wave[w], pitch);
	return o;
}

template <int LC>
__device__ __forceinline__ uint32_t reg_offset_rt(int j)
{
	constexpr Map m = final_map(LC);
	uint32_t o = 0;
#pragma unroll
	for (int r = 0; r < 5; ++r)
		o += ((j >> r) & 1) ? weight(LC, m.reg[r]) : 0u;
	return o;
}

// Per-thread top-2 of one tile in scan order: strict comparisons of magnitudes keep the
// earlier (lower) state on equal objectives. Signed objectives as signed_objective():
// BBPA |p - t| + |(T - p) - t| signed by p - (T - p); otherwise p - 1/2.
template <int LC, bool BBPA, bool MASKED>
__device__ __forceinline__ void scan_tile(const double (&v)[kTileRegs], double scale, double total, double target,
										  uint32_t gt, uint32_t inactive, double &d1, double &d2, int &j1, int &j2)
{
	constexpr Order ord = scan_order(LC);
	constexpr Map mf = final_map(LC);
#pragma unroll
	for (int q = 0; q < kTileRegs; ++q) {
		const int j = ord.j[q];
		const double p = v[j] * scale;
		double dd;
		if (BBPA) {
			const double negative = total - p;
			dd = copysign(fabs(p - target) + fabs(negative - target), p - negative);
		} else {
			dd = p - 0.5;
		}
		if (MASKED) {
			const uint32_t g = gt + reg_offset(mf, LC, j);
			dd = (g & inactive) ? INFINITY : dd;
		}
		// one select per update keeps this branch-free
		const bool c1 = fabs(dd) < fabs(d1), c2 = fabs(dd) < fabs(d2);
		j2 = c2 ? j : j2;
		d2 = c2 ? dd : d2;
		j2 = c1 ? j1 : j2;
		d2 = c1 ? d1 : d2;
		j1 = c1 ? j : j1;
		d1 = c1 ? dd : d1;
		// keep the scheduler from hoisting every product ahead of the selects
		__builtin_amdgcn_sched_barrier(0);
	}
}

} // namespace p2

// Pass 2 for A = 30 - LC atoms, persistent: workgroup b of G serves the tiles of XCD b % 8
// (contiguous range, neighbours run on CUs sharing one L2). Objective: kObjHalving or
// kObjBinaryBbpa; atom singletons are written by the tiles that hold them (tile 0 and
// power-of-two tiles).
// Band collection (lattices with exact rescoring, d.band set): every entry within
// band_width of the best |objective| seen so far is appended to the `cand` buffer. The
// best seen is a running minimum over this wave's tiles and its XCD's slot, never below the
// final best, so every entry within band_width of the final best is appended;
// k_top2_compact keeps those after the merge.
//
// SPLIT (FP32 intermediate, A >= 28): the lattice is run as one of A - 1 atoms (this LC is
// 31 - A) whose top row bit, state bit A - 1, is split into two work units per tile. The
// unit of the lower half (bit A-1 clear) sums each entry with its partner in the upper half
// before the sweep, which is the sweep of bit A-1 done first; the unit of the upper half
// sweeps its own half. Tiles are then twice as wide (runs of 32 bytes instead of 16), so
// the pass needs 3/4 of the memory requests for 3/2 of the bytes, and the two units of a
// tile run side by side on one XCD, where the half both read is one L2 fetch. The sum order
// differs from the increasing bit order only in FP64 additions of FP32 values, inside the
// FP32C error band; FP64 storage keeps the exact order.
template <typename TIN, int LC, bool SPLIT = false>
__global__ __launch_bounds__(kTileThreads) void k_zeta_pass2(Descs descs, int objective, int scaled, double target,
																unsigned fold)
{
	constexpr p2::Map mb = p2::block_map(LC);
	constexpr p2::Map mw = p2::wave_map(LC);
	constexpr unsigned tiles = 1u << (kTileBits - LC);
	constexpr unsigned units = SPLIT ? 2 * tiles : tiles;
	const LatDesc &d = descs.many ? descs.many[blockIdx.y] : descs.one;
	__shared__ double lds[kTileLds];
	__shared__ Top2 red2[16];
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const unsigned wg_per_xcd = gridDim.x >> 3;
	const unsigned per_xcd = units >> 3;
	const unsigned xcd = blockIdx.x & 7u;
	const unsigned first = xcd * per_xcd + (blockIdx.x >> 3), last = (xcd + 1) * per_xcd;
	// SPLIT reads z in the blocked form: rows of a column block are 2^kZBlockBits entries
	// apart, blocks pitch_b apart; tile T is columns (T & (2^(5-LC) - 1)) << LC.. of block
	// T >> (5 - LC)
	static_assert(!SPLIT || (sizeof(TIN) == 4 && LC >= 1 && LC <= 3), "the split form is FP32, 28..30 atoms");
	constexpr uint32_t pitch = SPLIT ? (1u << kZBlockBits) : z_pitch<TIN>();
	// SPLIT: the upper half's rows start this many entries after the lower half's
	constexpr size_t half_rows = (size_t)pitch << (kTileBits - LC);
	constexpr uint64_t pitch_b = 2 * half_rows + kZBlockSkew;
	const uint64_t lo = p2::thread_z_offset<LC, pitch>(lane, wave);
	const uint32_t lf = p2::thread_state_offset<LC>(lane, wave);
	const TIN *in = static_cast<const TIN *>(d.z);
	// per-lane byte offset from a wave-uniform base; at 30 atoms (and FP64 at 29) it no
	// longer fits 32 bits and the lane's own 64-bit pointer is used instead
	constexpr bool wide = p2::max_thread_z_offset(LC, pitch) * sizeof(TIN) >= (1ull << 32);
	const uint32_t lob = wide ? 0u : (uint32_t)(lo * sizeof(TIN));
	const TIN *inw = wide ? in + lo : in;
	double scale, total;
	if (fold != 0u) {
		// the fused update's denominator D from pass 1's per-tile partials, formed exactly as
		// k_finalize_sum forms it (1024 threads, the same strides and reduction), so the
		// values are identical to the separate launch's; the first workgroup stores them
		double t = 0.0;
		for (unsigned i = threadIdx.x; i < fold; i += blockDim.x)
			t += d.part[i];
		t = block_sum(t, lds);
		if (threadIdx.x == 0)
			lds[16] = t;
		__syncthreads();
		t = uniform(lds[16]); // workgroup-uniform values live in scalar registers
		__syncthreads();
		const double inv = uniform(1.0 / t);
		if (blockIdx.x == 0 && threadIdx.x == 0) {
			d.scal[kScalDen] = t;
			d.scal[d.inv_out] = inv;
			d.scal[kScalTotal] = t;
		}
		scale = uniform(scaled ? inv : 1.0);
		total = uniform(scaled ? t * scale : t);
	} else {
		scale = uniform(scaled ? d.scal[d.inv_out] : 1.0);
		total = uniform(scaled ? d.scal[kScalTotal] * scale : d.scal[kScalTotal]);
	}
	const uint32_t full = (d.n >= 32) ? 0xFFFFFFFFu : ((1u << d.n) - 1u);
	const uint32_t inactive = full & ~d.active;
	const bool bbpa = objective == kObjBinaryBbpa;
	const bool banding = d.band != nullptr;
	unsigned *slot = best_slots(d) + xcd;
	const unsigned raw_cap = 2u * gridDim.x * blockDim.x;
	double seen = INFINITY; // wave-uniform: the best |objective| this wave knows of
	bool raw_full = false;  // wave-uniform: the candidate buffer has overflowed
	Top2 run;
	top2_init(run);
	for (unsigned U = first; U < last; U += wg_per_xcd) {
		// (prefetching the next tile into registers costs more in spills than it hides:
		// the tile alone takes 64 of the 128 VGPRs a 1024-thread workgroup allows)
		// units 2T and 2T+1 are tile T's halves; which is which alternates from one turn of a
		// workgroup to the next, so every workgroup does as many two-half units as one-half
		const unsigned T = SPLIT ? U >> 1 : U;
		const bool upper = SPLIT && (((U & 1u) ^ ((T / (wg_per_xcd >> 1)) & 1u)) != 0u);
		double v[kTileRegs];
		{
			const uint64_t tile_at = SPLIT ? ((uint64_t)(T >> (kZBlockBits - LC)) * pitch_b +
											  ((uint64_t)(T & ((1u << (kZBlockBits - LC)) - 1u)) << LC))
										   : ((uint64_t)T << LC);
			const TIN *tb = inw + tile_at + (upper ? half_rows : 0);
			if (SPLIT && !upper) {
#pragma unroll
				for (int j = 0; j < kTileRegs; ++j)
					v[j] = static_cast<double>(ld_u(tb + (size_t)j * pitch, lob)) +
						   static_cast<double>(ld_u(tb + half_rows + (size_t)j * pitch, lob));
			} else {
#pragma unroll
				for (int j = 0; j < kTileRegs; ++j)
					v[j] = static_cast<double>(ld_u(tb + (size_t)j * pitch, lob));
			}
		}
		// registers: tile bits LC..LC+4
		bfly_reg<0>(v);
		bfly_reg<1>(v);
		bfly_reg<2>(v);
		bfly_reg<3>(v);
		bfly_reg<4>(v);
		__syncthreads(); // the previous tile's wave transposes used the same LDS
		block_transpose(v, lds, lane, wave);
		if constexpr (mb.reg[0] >= LC)
			bfly_reg<0>(v);
		if constexpr (mb.reg[1] >= LC)
			bfly_reg<1>(v);
		if constexpr (mb.reg[2] >= LC)
			bfly_reg<2>(v);
		if constexpr (mb.reg[3] >= LC)
			bfly_reg<3>(v);
		if constexpr (LC <= 3) {
			wave_transpose(v, lds + wave * (16 * kWaveStride), lane);
			if constexpr (mw.reg[0] >= LC)
				bfly_reg<0>(v);
			if constexpr (mw.reg[1] >= LC)
				bfly_reg<1>(v);
			if constexpr (mw.reg[2] >= LC)
				bfly_reg<2>(v);
			if constexpr (mw.reg[3] >= LC)
				bfly_reg<3>(v);
		}
		if constexpr (LC <= 4) {
			perm_swap<false, 0>(v);
			bfly_reg<0>(v);
		}
		if constexpr (LC <= 5) {
			perm_swap<true, 1>(v);
			bfly_reg<1>(v);
		}
		const uint32_t gt = ((uint32_t)T << LC) + lf + (upper ? (1u << (2 * kTileBits - LC)) : 0u);
		if ((T & (T - 1u)) == 0u) {
			// tile 0 and power-of-two tiles hold the singletons {a}
			constexpr p2::Map mf = p2::final_map(LC);
#pragma unroll
			for (int j = 0; j < kTileRegs; ++j) {
				const uint32_t g = gt + p2::reg_offset(mf, LC, j);
				if (__popc(g) == 1)
					d.masses[__ffs(g) - 1] = v[j] * scale;
			}
		}
		double d1 = INFINITY, d2 = INFINITY;
		int j1 = 0, j2 = 0;
		if (inactive == 0u) {
			if (bbpa)
				p2::scan_tile<LC, true, false>(v, scale, total, target, gt, inactive, d1, d2, j1, j2);
			else
				p2::scan_tile<LC, false, false>(v, scale, total, target, gt, inactive, d1, d2, j1, j2);
		} else {
			if (bbpa)
				p2::scan_tile<LC, true, true>(v, scale, total, target, gt, inactive, d1, d2, j1, j2);
			else
				p2::scan_tile<LC, false, true>(v, scale, total, target, gt, inactive, d1, d2, j1, j2);
		}
		if (fabs(d1) < INFINITY)
			top2_insert(run, d1, gt + p2::reg_offset_rt<LC>(j1));
		if (fabs(d2) < INFINITY)
			top2_insert(run, d2, gt + p2::reg_offset_rt<LC>(j2));
		if (banding) {
			double wm = fabs(d1);
#pragma unroll
			for (int o = 32; o > 0; o >>= 1)
				wm = fmin(wm, __shfl_xor(wm, o));
			seen = fmin(seen, key_value(__builtin_amdgcn_readfirstlane(
								  __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))));
			if (wm < seen) {
				seen = wm;
				if (lane == 0)
					atomicMin(slot, best_key(wm));
			}
			const double thr = seen + d.band_width;
			if (!raw_full && __ballot(fabs(d1) <= thr) != 0ull) {
				constexpr p2::Map mf = p2::final_map(LC);
				// count the wave's entries over all registers first (bit j of `inbits`: register
				// j is in), so the wave takes its slots with one atomic: on plateaus many waves
				// append, and per-register atomics on one counter serialise
				uint32_t inbits = 0u;
				unsigned count_in = 0u;
#pragma unroll
				for (int j = 0; j < kTileRegs; ++j) {
					const uint32_t g = gt + p2::reg_offset(mf, LC, j);
					const double sd = signed_objective(objective, v[j] * scale, total, target);
					const bool in = (g & inactive) == 0u && fabs(sd) <= thr;
					inbits |= (uint32_t)in << j;
					count_in += (unsigned)__popcll(__ballot(in));
				}
				if (count_in != 0u) {
					unsigned base = 0u;
					if (lane == 0)
						base = atomicAdd(&d.counter[kRawCount], count_in);
					base = __builtin_amdgcn_readfirstlane(base);
					raw_full = base + count_in > raw_cap;
					if (!raw_full) {
						const unsigned long long below = (1ull << lane) - 1ull;
						// objectives and states are recomputed (same operations, same bits)
						// from opaque copies of the scale and the tile base, so the compiler
						// cannot keep all 32 of the first loop's alive across the atomic
						double scale2 = scale;
						uint32_t gt2 = gt;
						asm volatile("" : "+v"(scale2), "+v"(gt2));
#pragma unroll
						for (int j = 0; j < kTileRegs; ++j) {
							const bool in = (inbits >> j) & 1u;
							const unsigned long long mask = __ballot(in);
							if (in) {
								Cand c;
								c.d = signed_objective(objective, v[j] * scale2, total, target);
								c.e = gt2 + p2::reg_offset(mf, LC, j);
								c.flags = kCandSided;
								d.cand[base + (unsigned)__popcll(mask & below)] = c;
							}
							base += (unsigned)__popcll(mask);
						}
					}
				}
			}
		}
	}
	publish_top2(d, blockIdx.x, run, red2);
}

// ----------------------------- three-pass transform of FP32 lattices of 26..30 atoms
//
// With 2^15-entry tiles the two passes of the transform share at most 30 - A state bits, so
// pass 2 of a large FP32 lattice reads pieces of 16 or 32 bytes (see k_zeta_pass2). The
// three-pass form gives the high bits two passes that both read whole 128-byte lines:
//   pass M (k_zeta_mid): state bits 15..24 on tiles of 32 contiguous columns (128 bytes of
//     z per row) by 1024 rows, in place in z, persistent with the next tile's entries loaded
//     during the current sweep (as pass 1);
//   pass T (k_zeta_top): bits 25..A-1, 2^(A-25) rows per column, consecutive lanes on
//     consecutive columns, fused with the objective, the top-2, the error band and the
//     atom singletons (as pass 2).
// It moves 2 S b_z bytes more than the two-pass form. Sums stay FP64 on chip; z is rounded
// to FP32 once more, between pass M and pass T. Bits are still swept in increasing order.
constexpr int kMidBits = 10;                  // pass M sweeps bits 15..24
constexpr int kMidCols = 5;                   // 32 columns per pass-M tile
constexpr int kTopLo = kTileBits + kMidBits;  // pass T sweeps bits 25..A-1
constexpr unsigned kTopGroups = 512;          // pass T: 512 x 512 threads (2^18, a power of two)
// Untuned default: the two-pass form everywhere. Measured at 2^28 FP32 states (rocprofv3,
// fused rounds): pass M 0.48 ms and pass T 0.50-0.66 ms against 0.64 ms for the split pass 2,
// so the three-pass form is a candidate of the tuner only.
constexpr int kThreePassFrom = 31;

// z entries of a pass-M tile: column block T & 1023 (32 columns), row group T >> 10 (1024 rows)
template <typename TZ>
__device__ __forceinline__ TZ *mid_tile(TZ *z, unsigned T)
{
	return z + ((uint64_t)(T >> kMidBits) << kMidBits) * z_pitch<TZ>() + ((uint64_t)(T & 1023u) << kMidCols);
}

// Load layout (tile_index_a): column = lane & 31; row = lane bit 5 | jlo << 1 | wave << 5 | jhi << 9.
template <typename TZ>
__device__ __forceinline__ void mid_load(const TZ *tb, uint32_t lob, TZ (&raw)[kTileRegs])
{
	constexpr uint32_t pitch = z_pitch<TZ>();
#pragma unroll
	for (int j = 0; j < kTileRegs; ++j)
		raw[j] = ld_u(tb + (size_t)((((j & 15) << 1) | ((j >> 4) << 9)) * pitch), lob);
}

template <typename TZ>
__global__ __launch_bounds__(kTileThreads) void k_zeta_mid(Descs descs)
{
	const LatDesc &d = descs.many ? descs.many[blockIdx.y] : descs.one;
	__shared__ double lds[kTileLds];
	constexpr uint32_t pitch = z_pitch<TZ>();
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const unsigned tiles = 1u << (__builtin_amdgcn_readfirstlane(d.atoms) - kTileBits);
	TZ *z = uniform_ptr(static_cast<TZ *>(d.z));
	const uint32_t lob = ((uint32_t)(lane & 31) + ((uint32_t)(lane >> 5) + ((uint32_t)wave << 5)) * pitch) *
						 (uint32_t)sizeof(TZ);
	// output layout (tile_index_f): column = lane_bits_f & 31; row = its bit 5 | wave << 1 | jlo << 5 | jhi << 9
	const uint32_t lbf = lane_bits_f(lane);
	const uint32_t los = ((lbf & 31u) + ((lbf >> 5) + ((uint32_t)wave << 1)) * pitch) * (uint32_t)sizeof(TZ);
	// workgroup b serves the contiguous tile range of XCD b % 8: neighbouring column blocks
	// (adjacent 128-byte pieces of the same rows) are read behind one L2
	const unsigned wg_per_xcd = gridDim.x >> 3;
	const unsigned per_xcd = tiles >> 3;
	const unsigned xcd = blockIdx.x & 7u;
	const unsigned first = xcd * per_xcd + (blockIdx.x >> 3), last = (xcd + 1) * per_xcd;
	TZ next[kTileRegs];
	if (first < last)
		mid_load(mid_tile(z, first), lob, next);
	for (unsigned T = first; T < last; T += wg_per_xcd) {
		double v[kTileRegs];
#pragma unroll
		for (int j = 0; j < kTileRegs; ++j)
			v[j] = static_cast<double>(next[j]);
		if (T + wg_per_xcd < last)
			mid_load(mid_tile(z, T + wg_per_xcd), lob, next);
		zeta_tile(v, kMidCols, lane, wave, lds); // tile bits 5..14: state bits 15..24
		TZ *tb = mid_tile(z, T);
#pragma unroll
		for (int j = 0; j < kTileRegs; ++j)
			st_u(tb + (size_t)((((j & 15) << 5) | ((j >> 4) << 9)) * pitch), los, static_cast<TZ>(v[j]));
	}
}

// Pass T: thread t of the launch's kTopThreads threads takes positions p = (it*P + q)*NT + t
// (column p & 32767, row p >> 15 below bit 25) and their 2^HB rows of bits 25..A-1,
// P = 32 / 2^HB positions per turn, so each turn holds 32 entries (state p + (h << 25) in
// register q * 2^HB + h). NT is a multiple of 2^15, so every load is a wave-uniform base
// plus the thread's own 32-bit offset. Objective, band and singletons as in k_zeta_pass2;
// ties resolve by the order (|objective|, experiment), as everywhere.
constexpr unsigned kTopThreads = kTopGroups * 512u;
constexpr int kTopRegs = 16;
template <typename TZ, int HB>
__global__ __launch_bounds__(512) void k_zeta_top(Descs descs, int objective, int scaled, double target,
												  unsigned fold, unsigned raw_cap)
{
	constexpr int H = 1 << HB;
	constexpr int R = H > kTopRegs ? H : kTopRegs; // entries per thread per turn
	constexpr int P = R / H;
	constexpr uint32_t pitch = z_pitch<TZ>();
	constexpr uint64_t hstride = (uint64_t)pitch << kMidBits;
	constexpr uint32_t positions = 1u << kTopLo;
	constexpr uint32_t NT = kTopThreads;
	static_assert((NT & ((1u << kTileBits) - 1u)) == 0u, "pass T thread count must be a multiple of 2^15");
	const LatDesc &d = descs.many ? descs.many[blockIdx.y] : descs.one;
	__shared__ double red[16];
	__shared__ Top2 red2[16];
	const int lane = threadIdx.x & 63;
	double scale, total;
	if (fold != 0u) {
		// the fused update's denominator from pass 1's per-tile partials, formed exactly as
		// k_finalize_sum (1024 threads) forms it: this launch has 512 threads per workgroup,
		// so each thread folds the two strides of k_finalize_sum's threads t and t + 512 in
		// their order, and the sum over waves takes them in that kernel's order
		double t0 = 0.0, t1 = 0.0;
		for (unsigned i = threadIdx.x; i < fold; i += 1024u)
			t0 += d.part[i];
		for (unsigned i = threadIdx.x + 512u; i < fold; i += 1024u)
			t1 += d.part[i];
		const double w0 = wave_sum(t0), w1 = wave_sum(t1);
		const int w = threadIdx.x >> 6;
		if (lane == 0) {
			red[w] = w0;
			red[w + 8] = w1;
		}
		__syncthreads();
		if (threadIdx.x == 0) {
			double s = 0.0;
			for (int i = 0; i < 16; ++i)
				s += red[i];
			red[0] = s;
		}
		__syncthreads();
		const double tsum = uniform(red[0]);
		const double inv = uniform(1.0 / tsum);
		if (blockIdx.x == 0 && threadIdx.x == 0) {
			d.scal[kScalDen] = tsum;
			d.scal[d.inv_out] = inv;
			d.scal[kScalTotal] = tsum;
		}
		scale = uniform(scaled ? inv : 1.0);
		total = uniform(scaled ? tsum * scale : tsum);
	} else {
		scale = uniform(scaled ? d.scal[d.inv_out] : 1.0);
		total = uniform(scaled ? d.scal[kScalTotal] * scale : d.scal[kScalTotal]);
	}
	const uint32_t full = (d.n >= 32) ? 0xFFFFFFFFu : ((1u << d.n) - 1u);
	const uint32_t inactive = full & ~d.active;
	const bool banding = d.band != nullptr;
	unsigned *slot = best_slots(d) + (blockIdx.x & 7u);
	const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
	const TZ *z = uniform_ptr(static_cast<const TZ *>(d.z));
	const uint32_t lob = ((t >> kTileBits) * pitch + (t & ((1u << kTileBits) - 1u))) * (uint32_t)sizeof(TZ);
	double seen = INFINITY; // wave-uniform: the best |objective| this wave knows of
	bool raw_full = false;  // wave-uniform: the candidate buffer has overflowed
	Top2 run;
	top2_init(run);
	unsigned turn = 0;
	for (uint32_t base = 0; base < positions; base += P * NT, ++turn) {
		// the XCD's running best, read every eighth turn (its latency would otherwise stall
		// each turn): a stale value only widens this wave's threshold
		unsigned key = kKeyInf;
		if (banding && (turn & 7u) == 0u)
			key = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		double v[R];
		const TZ *zb = z + (uint64_t)(base >> kTileBits) * pitch;
#pragma unroll
		for (int q = 0; q < P; ++q)
#pragma unroll
			for (int h = 0; h < H; ++h)
				v[q * H + h] = static_cast<double>(
					ld_u(zb + (uint64_t)q * (NT >> kTileBits) * pitch + (uint64_t)h * hstride, lob));
		// bits 25..A-1 in increasing order
#pragma unroll
		for (int b = 0; b < HB; ++b)
#pragma unroll
			for (int r = 0; r < R; ++r)
				if (!(r & (1 << b)))
					v[r] += v[r | (1 << b)];
		double best = INFINITY;
#pragma unroll
		for (int r = 0; r < R; ++r) {
			const uint32_t g = base + (uint32_t)(r / H) * NT + t + ((uint32_t)(r % H) << kTopLo);
			const double m = v[r] * scale;
			if (__popc(g) == 1)
				d.masses[__ffs(g) - 1] = m;
			if ((g & inactive) == 0u) {
				const double dd = signed_objective(objective, m, total, target);
				top2_insert(run, dd, g);
				best = fmin(best, fabs(dd));
			}
		}
		if (banding) {
			double wm = best;
#pragma unroll
			for (int o = 32; o > 0; o >>= 1)
				wm = fmin(wm, __shfl_xor(wm, o));
			if ((turn & 7u) == 0u)
				seen = fmin(seen, key_value(__builtin_amdgcn_readfirstlane(key)));
			if (wm < seen) {
				seen = wm;
				if (lane == 0)
					atomicMin(slot, best_key(wm));
			}
			const double thr = seen + d.band_width;
			if (!raw_full && __ballot(best <= thr) != 0ull) {
				uint32_t inbits = 0u;
				unsigned count_in = 0u;
#pragma unroll
				for (int r = 0; r < R; ++r) {
					const uint32_t g = base + (uint32_t)(r / H) * NT + t + ((uint32_t)(r % H) << kTopLo);
					const double sd = signed_objective(objective, v[r] * scale, total, target);
					const bool in = (g & inactive) == 0u && fabs(sd) <= thr;
					inbits |= (uint32_t)in << r;
					count_in += (unsigned)__popcll(__ballot(in));
				}
				if (count_in != 0u) {
					unsigned at = 0u;
					if (lane == 0)
						at = atomicAdd(&d.counter[kRawCount], count_in);
					at = __builtin_amdgcn_readfirstlane(at);
					raw_full = at + count_in > raw_cap;
					if (!raw_full) {
						const unsigned long long below = (1ull << lane) - 1ull;
						double scale2 = scale;
						uint32_t base2 = base;
						asm volatile("" : "+v"(scale2), "+v"(base2));
#pragma unroll
						for (int r = 0; r < R; ++r) {
							const bool in = (inbits >> r) & 1u;
							const unsigned long long mask = __ballot(in);
							if (in) {
								Cand c;
								c.d = signed_objective(objective, v[r] * scale2, total, target);
								c.e = base2 + (uint32_t)(r / H) * NT + t + ((uint32_t)(r % H) << kTopLo);
								c.flags = kCandSided;
								d.cand[at + (unsigned)__popcll(mask & below)] = c;
							}
							at += (unsigned)__popcll(mask);
						}
					}
				}
			}
		}
	}
	publish_top2(d, blockIdx.x, run, red2);
}

// Lattices of at most 2^14 states: the whole transform in the LDS of one workgroup.
template <typename TP>
__global__ __launch_bounds__(1024) void k_zeta_small(Descs descs, int atoms, int objective, double target)
{
	const LatDesc &d = descs.many ? descs.many[blockIdx.y] : descs.one;
	if (descs.ragged)
		atoms = d.atoms;
	__shared__ double a[1 << kSmallBits];
	__shared__ Top2 red2[16];
	__shared__ double red[16];
	__shared__ double s_total;
	__shared__ double s_lim;
	__shared__ unsigned s_count;
	const uint32_t states = 1u << atoms;
	gptr<const TP> post = G(static_cast<const TP *>(d.post));
	const double inv = d.scal[d.inv_in];
	double acc = 0.0;
	for (uint32_t s = threadIdx.x; s < states; s += blockDim.x) {
		const double x = static_cast<double>(post[s]) * inv;
		a[s] = x;
		acc += x;
	}
	const double t = block_sum(acc, red);
	if (threadIdx.x == 0) {
		s_total = t;
		d.scal[kScalTotal] = t;
	}
	__syncthreads();
	for (int b = 0; b < atoms; ++b) {
		const uint32_t bit = 1u << b;
		for (uint32_t p = threadIdx.x; p < (states >> 1); p += blockDim.x) {
			const uint32_t lo = p & (bit - 1u);
			const uint32_t e = ((p >> b) << (b + 1)) | lo;
			a[e] += a[e | bit];
		}
		__syncthreads();
	}
	ObjCtx c;
	c.objective = objective;
	c.active = d.active;
	c.target = target;
	c.total = s_total;
	c.ns = d.n;
	c.k = d.k;
	c.ftab = d.ftab;
	c.masses = d.masses;
	Top2 tt;
	top2_init(tt);
	for (uint32_t s = threadIdx.x; s < states; s += blockDim.x)
		consume(c, s, a[s], tt);
	Top2 r = top2_block(tt, red2);
	if (threadIdx.x == 0) {
		d.result->d1 = fabs(r.d1);
		d.result->d2 = fabs(r.d2);
		d.result->e1 = r.e1;
		d.result->e2 = r.e2;
		d.counter[0] = 0;
		d.counter[1] = 0;
		s_lim = fabs(r.d1) + d.band_width;
		s_count = 0u;
	}
	__syncthreads();
	// the whole transform is in this workgroup's LDS: collect the complete band directly
	if (d.band != nullptr && (objective == kObjHalving || objective == kObjBinaryBbpa)) {
		const double lim = s_lim;
		for (uint32_t s = threadIdx.x; s < states; s += blockDim.x) {
			const double sd = signed_objective(objective, a[s], s_total, target);
			if ((s & ~d.active) == 0u && fabs(sd) <= lim) {
				const unsigned idx = atomicAdd(&s_count, 1u);
				if (idx < d.band_cap) {
					Cand cd;
					cd.d = sd;
					cd.e = s;
					cd.flags = kCandSided;
					d.band[idx] = cd;
				}
			}
		}
		__syncthreads();
		if (threadIdx.x == 0)
			d.counter[0] = s_count;
	}
}

// ----------------------------------------- Op-BBPA feature tables in one pass (k >= 2)
//
// The reference builds, for every non-empty variant subset B, the table
// H_B[I] = sum of x_S over the states whose variant blocks in B intersect in I
// (bbpa.cpp:237-285), and zetas each table over the n subject bits (bbpa.cpp:292-313), giving
// F_B(E) = sum of x_S over the states with S_v >= E for every v in B.
//
// Here the zeta comes first, per row: a row is the 2^n states of fixed blocks S_1..S_{k-1},
// and its own zeta over the subject bits, Z(E) = sum over S_0 >= E of the row, gives every
// table's contribution at once: F_B(E) += Z(E) when B holds variant 0 and E lies inside the
// AND c of B's other blocks (all subjects for B = {0}), and F_B(E) += Z(0) when B does not
// hold variant 0 and E lies inside the AND of its blocks. Lane l, register r of a wave hold
// E = l + 64 r (n >= 6), so each lane accumulates its own entries of every table over the
// rows it visits, with no exchange; the row's zeta runs over the lane bits by butterflies and
// over the register bits in registers, bits in increasing order.
//
// The workgroup's waves then add their tables in wave order, and workgroups add theirs to
// the device table in fixed point, 96 fractional bits as three 32-bit digits in 64-bit
// integer limbs: integer additions commute, so the tables do not depend on the order in
// which workgroups finish (the same bits alone or batched, run after run).
// Workgroup size: the largest whose waves fit 4 per SIMD beside their tables (one workgroup
// of 16 waves per CU where each lane holds at most 32 table entries, 4 waves otherwise):
// more waves per workgroup hide the row loads' latency and share one merge into the device
// table (the merge's atomics are per workgroup, not per wave). Four variants over two row
// bits (30 entries) would spill at the 128-register limit of 16 waves, so they keep 4.
constexpr int hist_threads(int k, int nh)
{
	return ((((1 << k) - 1) << nh) <= 32 && !(k == 4 && nh == 1)) ? 1024 : 256;
}

__device__ __forceinline__ void fix_add(unsigned long long *bin3, double x)
{
	// x in [0, 2): the digits of floor(x 2^96) (each step exact: t - floor(t) is the
	// fraction of a double, times 2^32 is a power-of-two scaling)
	const double t0 = x * 4294967296.0;
	const double f0 = floor(t0);
	const double t1 = (t0 - f0) * 4294967296.0;
	const double f1 = floor(t1);
	const double f2 = floor((t1 - f1) * 4294967296.0);
	if (f0 != 0.0)
		atomicAdd(&bin3[0], (unsigned long long)f0);
	if (f1 != 0.0)
		atomicAdd(&bin3[1], (unsigned long long)f1);
	if (f2 != 0.0)
		atomicAdd(&bin3[2], (unsigned long long)f2);
}

// value of a bin's three limbs: L0 2^-32 + L1 2^-64 + L2 2^-96 after the carries
__device__ __forceinline__ double fix_value(const unsigned long long *bin3)
{
	unsigned long long l0 = bin3[0], l1 = bin3[1], l2 = bin3[2];
	l1 += l2 >> 32;
	l2 &= 0xFFFFFFFFull;
	l0 += l1 >> 32;
	l1 &= 0xFFFFFFFFull;
	return ((double)l0 * 0x1p-32 + (double)l1 * 0x1p-64) + (double)l2 * 0x1p-96;
}

// FUSED: apply the update first (y' = x L, stored in place) and tabulate y'; otherwise
// x = y * inv. K variants, NH = n - 6 (2^NH registers per lane). The device table d.fixtab
// holds (2^K - 1) tables of 2^n bins (table B at (B - 1) 2^n) and the total, three limbs each.
template <typename TP, bool FUSED, int K, int NH>
__global__ __launch_bounds__(hist_threads(K, NH)) void k_bbpa_hist(Descs descs)
{
	constexpr int R = 1 << NH;
	constexpr int T = (1 << K) - 1;
	constexpr int kHistWaves = hist_threads(K, NH) / 64;
	const LatDesc &d = descs.many ? descs.many[blockIdx.y] : descs.one;
	extern __shared__ double s_acc[]; // T * 2^n + 1 doubles: the workgroup's tables
	__shared__ double s_lik[kLikCapacity];
	const int n = NH + 6;
	const uint32_t ec = 1u << n, mask = ec - 1u;
	if (FUSED)
		load_lik_table(descs, d, s_lik);
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const unsigned gwave = blockIdx.x * kHistWaves + wave;
	const unsigned nwaves = gridDim.x * kHistWaves;
	const uint64_t rows = 1ull << (n * (K - 1));
	TP *post = static_cast<TP *>(d.post);
	const double inv = d.scal[d.inv_in];
	double f[T][R];
#pragma unroll
	for (int b = 0; b < T; ++b)
#pragma unroll
		for (int r = 0; r < R; ++r)
			f[b][r] = 0.0;
	double total = 0.0; // lane 0's row totals
	for (uint64_t row = gwave; row < rows; row += nwaves) {
		const uint64_t base = row << n;
		double v[R];
#pragma unroll
		for (int r = 0; r < R; ++r) {
			const uint64_t st = base + (uint64_t)lane + 64u * r;
			double x = static_cast<double>(post[st]) * inv;
			if (FUSED) {
				const TP y = static_cast<TP>(x * likelihood((uint32_t)st, d.e, n, K, d.len, d.dilution, s_lik));
				post[st] = y;
				x = static_cast<double>(y);
			}
			v[r] = x;
		}
		// the row's zeta over S_0: lane bits 0..5, then register bits 6..n-1
#pragma unroll
		for (int lb = 0; lb < 6; ++lb) {
			const bool low = !((lane >> lb) & 1);
#pragma unroll
			for (int r = 0; r < R; ++r) {
				const double o = __shfl_xor(v[r], 1 << lb);
				v[r] = low ? v[r] + o : v[r];
			}
		}
#pragma unroll
		for (int hb = 0; hb < NH; ++hb)
#pragma unroll
			for (int r = 0; r < R; ++r)
				if (!(r & (1 << hb)))
					v[r] += v[r | (1 << hb)];
		const double z0 = lane_d(v[0], 0); // the row's total, Z(0)
		total += z0;
		uint32_t blk[K];
		blk[0] = mask;
#pragma unroll
		for (int q = 1; q < K; ++q)
			blk[q] = (uint32_t)(row >> ((q - 1) * n)) & mask;
#pragma unroll
		for (int b = 1; b <= T; ++b) {
			uint32_t c = mask;
#pragma unroll
			for (int q = 1; q < K; ++q)
				if ((b >> q) & 1)
					c &= blk[q];
#pragma unroll
			for (int r = 0; r < R; ++r) {
				const uint32_t e = (uint32_t)lane + 64u * r;
				const double add = (b & 1) ? v[r] : z0;
				f[b - 1][r] += ((e & ~c) == 0u) ? add : 0.0;
			}
		}
	}
	// the workgroup's waves in wave order, then the device table in fixed point
	const unsigned cells = (unsigned)T * ec;
	for (int w = 0; w < kHistWaves; ++w) {
		if (wave == w) {
#pragma unroll
			for (int b = 0; b < T; ++b)
#pragma unroll
				for (int r = 0; r < R; ++r) {
					const unsigned i = (unsigned)b * ec + (unsigned)lane + 64u * r;
					s_acc[i] = (w == 0) ? f[b][r] : s_acc[i] + f[b][r];
				}
			if (lane == 0)
				s_acc[cells] = (w == 0) ? total : s_acc[cells] + total;
		}
		__syncthreads();
	}
	unsigned long long *g = d.fixtab;
	for (unsigned i = threadIdx.x; i <= cells; i += blockDim.x)
		fix_add(&g[3u * i], s_acc[i]);
}

// The tables in FP64 into d.ftab (table B at B * 2^n), the total into the scalars, the atom
// masses (F_{v}({i})) into d.masses; the fixed-point table is cleared for the next use.
// FUSED: the tables were of the updated, unnormalised posterior: their total is the
// denominator D, stored with 1/D in the inv_out slot, and the tables are scaled by 1/D.
__global__ __launch_bounds__(1024) void k_bbpa_tables(Descs descs, int fused)
{
	const LatDesc &d = descs.many ? descs.many[blockIdx.y] : descs.one;
	const int n = d.n, k = d.k;
	const uint32_t ec = 1u << n;
	const unsigned cells = ((1u << k) - 1u) * ec;
	unsigned long long *g = d.fixtab;
	const double t = fix_value(&g[3u * cells]);
	const double scale = fused ? 1.0 / t : 1.0;
	for (unsigned i = threadIdx.x; i < cells; i += blockDim.x)
		d.ftab[(size_t)ec + i] = fix_value(&g[3u * i]) * scale;
	for (unsigned i = threadIdx.x; i < (unsigned)(k * n); i += blockDim.x) {
		const unsigned v = i / n, a = i % n;
		d.masses[i] = fix_value(&g[3u * (((1u << v) - 1u) * ec + (1u << a))]) * scale;
	}
	if (threadIdx.x == 0) {
		if (fused) {
			d.scal[kScalDen] = t;
			d.scal[d.inv_out] = scale;
			d.scal[kScalTotal] = t * scale;
		} else {
			d.scal[kScalTotal] = t;
		}
	}
	__syncthreads();
	for (unsigned i = threadIdx.x; i < 3u * (cells + 1u); i += blockDim.x)
		g[i] = 0ull;
}

// BBPA k >= 2 scoring from the feature tables (bbpa.cpp:332-365 for k = 2 and 416-460
// otherwise); same operation order per experiment.
// Op-BBPA objective of experiment e from the feature tables (bbpa.cpp:416-460): each response
// mask's mass by inclusion-exclusion over the tables F_B (F_{} = the total), and the sum of
// their distances to the target 1/2^k.
__device__ __forceinline__ double bbpa_objective(const double *fm, uint32_t ec, uint32_t e, int k, double total,
												 double target)
{
	if (k == 2) {
		const double p0 = fm[ec + e], p1 = fm[2 * ec + e], pb = fm[3 * ec + e];
		const double r0 = pb;
		const double r1 = p1 - pb;
		const double r2 = p0 - pb;
		const double r3 = total - p0 - p1 + pb;
		return fabs(r0 - target) + fabs(r1 - target) + fabs(r2 - target) + fabs(r3 - target);
	}
	const uint32_t rc = 1u << k, all = rc - 1u;
	double score = 0.0;
	for (uint32_t resp = 0; resp < rc; ++resp) {
		const uint32_t positive = (~resp) & all;
		const uint32_t missing = resp;
		double exact = 0.0;
		for (uint32_t extra = missing;; extra = (extra - 1u) & missing) {
			const uint32_t fs = positive | extra;
			const double f = fs == 0 ? total : fm[(size_t)fs * ec + e];
			if ((__popc(extra) & 1) == 0)
				exact += f;
			else
				exact -= f;
			if (extra == 0)
				break;
		}
		score += fabs(exact - target);
	}
	return score;
}

// scaled: the tables come from a final pass that scaled the transform by the pending 1/D (a
// fused update), so the total is scaled the same way (the final pass's own total)
__global__ __launch_bounds__(256) void k_bbpa_score(Descs descs, int ns, int k, double target, int scaled)
{
	const LatDesc &d = descs.many ? descs.many[blockIdx.y] : descs.one;
	// ragged: lattices of one variant count (one target), each with its own subjects
	if (descs.ragged) {
		ns = d.n;
		if (blockIdx.x >= score_groups(ns))
			return;
	}
	__shared__ Top2 red2[16];
	const uint32_t ec = 1u << ns;
	const double total = scaled ? d.scal[kScalTotal] * d.scal[d.inv_out] : d.scal[kScalTotal];
	Top2 t;
	top2_init(t);
	const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e < ec && (e & ~d.active) == 0)
		top2_insert(t, bbpa_objective(d.ftab, ec, e, k, total, target), e);
	publish_top2(d, blockIdx.x, t, red2);
}

// The Op-BBPA error band of lattices of several variants: after k_bbpa_score and the merge of
// its top-2 lists, every selectable experiment whose objective (recomputed from the same
// tables, so the same value) lies within band_width of the best is appended to d.gband,
// one atomic per wave on counter[0] (cleared by the merge). Appends past band_cap are
// counted, not stored: counter[0] > band_cap reports an incomplete band.
__global__ __launch_bounds__(256) void k_bbpa_band(Descs descs, int ns, int k, double target, int scaled)
{
	const LatDesc &d = descs.many ? descs.many[blockIdx.y] : descs.one;
	if (descs.ragged) {
		ns = d.n;
		if (blockIdx.x >= score_groups(ns))
			return;
	}
	if (d.gband == nullptr)
		return;
	const uint32_t ec = 1u << ns;
	const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
	const double lim = d.result->d1 + d.band_width;
	double s = INFINITY;
	const double total = scaled ? d.scal[kScalTotal] * d.scal[d.inv_out] : d.scal[kScalTotal];
	if (e < ec && (e & ~d.active) == 0)
		s = bbpa_objective(d.ftab, ec, e, k, total, target);
	const bool in = s <= lim;
	const unsigned long long mask = __ballot(in);
	if (mask == 0ull)
		return;
	const int lane = threadIdx.x & 63;
	unsigned base = 0u;
	if (lane == __ffsll((long long)mask) - 1)
		base = atomicAdd(&d.counter[0], (unsigned)__popcll(mask));
	base = __shfl(base, __ffsll((long long)mask) - 1);
	if (in) {
		const unsigned at = base + (unsigned)__popcll(mask & ((1ull << lane) - 1ull));
		if (at < d.band_cap) {
			Cand c;
			c.d = s;
			c.e = e;
			c.flags = 0u;
			d.gband[at] = c;
		}
	}
}

// Exact re-decision of an Op-BBPA near tie on a lattice of several variants (see
// launch_bbpa_band): each workgroup takes band entries blockIdx.x, + gridDim.x, ... and writes
// their exact objectives to d.gexact. Factors as the update kernels' likelihood tables
// (fill_lik: the same FP64 expressions).
constexpr unsigned kGatherGroups = 256;

__device__ __forceinline__ bool gather_redecides(const LatDesc &d)
{
	if (d.gband == nullptr || d.gh_len < 0)
		return false;
	const double d1 = d.result->d1, d2 = d.result->d2;
	return (d2 - d1) <= d.band_width && d.counter[0] <= d.band_cap;
}

// A set of a projected lattice's subjects placed in the frame of its exact program: bit i
// of e goes to the i-th set bit of keep (keep = every subject: e itself).
__device__ __forceinline__ uint32_t frame_deposit(uint32_t e, uint32_t keep)
{
	uint32_t out = 0u;
	for (uint32_t r = keep; e != 0u && r != 0u; r &= r - 1u, e >>= 1)
		out |= (e & 1u) ? (r & (0u - r)) : 0u;
	return out;
}

// The Op-BBPA objective from the per-variant binary masses M_v(E): bbpa.cpp:416-460 over
// F_B = prod_{v in B} M_v (F_{} = 1), in variant order
__device__ double gather_objective(const double *mv, int k, double target)
{
	const uint32_t all = (1u << k) - 1u;
	double score = 0.0;
	for (uint32_t resp = 0; resp <= all; ++resp) {
		const uint32_t positive = (~resp) & all;
		double exact = 0.0;
		for (uint32_t extra = resp;; extra = (extra - 1u) & resp) {
			const uint32_t fs = positive | extra;
			double fv = 1.0;
			bool first = true;
			for (int v = 0; v < k; ++v)
				if ((fs >> v) & 1u) {
					fv = first ? mv[v] : fv * mv[v];
					first = false;
				}
			if ((__popc(extra) & 1) == 0)
				exact += fv;
			else
				exact -= fv;
			if (extra == 0)
				break;
		}
		score += fabs(exact - target);
	}
	return score;
}

// The exact program's update history for the block-weight sums, staged in LDS by
// stage_gather (several variants: frames of at most kMaxAtoms / 2 subjects)
constexpr int kGatherDil = (kMaxAtoms / 2) * (kMaxAtoms / 2 + 1);
struct GatherHist {
	uint32_t e[kGatherMaxHist]; // pools (frame subjects)
	uint32_t r[kGatherMaxHist]; // responses, bit v for variant v
	double dil[kGatherDil];     // dilution table rows of n + 1 entries
	int len, n, dilution;
};

// every thread of the workgroup; ends with a barrier
__device__ void stage_gather(const LatDesc &d, GatherHist &h)
{
	const int n = d.gx_n, len = d.gh_len < 0 ? 0 : d.gh_len;
	for (int i = threadIdx.x; i < len; i += blockDim.x) {
		h.e[i] = d.gh_e[i];
		h.r[i] = d.gh_r[i];
	}
	const int cells = min(n * (n + 1), kGatherDil);
	if (d.gh_dil != nullptr)
		for (int i = threadIdx.x; i < cells; i += blockDim.x)
			h.dil[i] = d.gh_dil[i];
	if (threadIdx.x == 0) {
		h.len = len;
		h.n = n;
		h.dilution = d.gh_dil != nullptr ? 1 : 0;
	}
	__syncthreads();
}

// The weight of state S of block v (in the program's frame of h.n subjects): the prior
// factors of its atoms times the likelihood factors of the history's updates for variant v
__device__ __forceinline__ double gather_weight(const GatherHist &h, const double *s_pairs, int v, uint32_t s)
{
	const int n = h.n;
	double p = 1.0;
	for (int i = 0; i < n; ++i)
		p *= s_pairs[2 * (v * n + i) + ((s >> i) & 1u)];
	for (int t = 0; t < h.len; ++t) {
		const uint32_t pool = h.e[t];
		const int len = __popc(pool);
		const int m = __popc(pool & ~s); // positive pooled subjects of block v
		const bool rb = (h.r[t] >> v) & 1u;
		double fv;
		if (h.dilution) {
			const double p1 = len > 0 ? h.dil[(len - 1) * (n + 1) + m] : 0.0;
			fv = rb ? p1 : 1.0 - p1;
		} else if (m == 0) {
			fv = rb ? kNegativeResponse : (1.0 - kNegativeResponse);
		} else {
			fv = rb ? (1.0 - kNegativeResponse) : kNegativeResponse;
		}
		p *= fv;
	}
	return p;
}

// the superset j of E: j's bits deposited at the subjects outside E, in increasing order
__device__ __forceinline__ uint32_t gather_superset(uint32_t e, uint32_t freeb, int f, uint32_t j)
{
	uint32_t s = e, r = freeb;
	for (int i = 0; i < f; ++i) {
		const uint32_t b = r & (0u - r);
		r &= r - 1u;
		s |= ((j >> i) & 1u) ? b : 0u;
	}
	return s;
}

// U_v(E) of block v (E in the program's frame): the sum of gather_weight over the states
// S >= E, every thread a strided share of the states in increasing order, then the
// workgroup tree
__device__ double gather_block_mass(const GatherHist &h, const double *s_pairs, int v, uint32_t e, double *red)
{
	const uint32_t freeb = ((1u << h.n) - 1u) & ~e;
	const int f = __popc(freeb);
	double acc = 0.0;
	for (uint32_t j = threadIdx.x; j < (1u << f); j += blockDim.x)
		acc += gather_weight(h, s_pairs, v, gather_superset(e, freeb, f, j));
	const double sum = block_sum(acc, red);
	__shared__ double s_sum;
	if (threadIdx.x == 0)
		s_sum = sum;
	__syncthreads();
	const double out = s_sum;
	__syncthreads();
	return out;
}

// The same sum by one wavefront: lane l the states l, l + 64, ... in increasing order, then
// the wavefront's butterfly (each step's two partners add the same two values, so every lane
// ends with the same sum). Frames of at most kWaveMassBits subjects: a lane then adds at
// most 256 terms in sequence, the share the exact masses' rounding bound allows a thread.
constexpr int kWaveMassBits = 14;
__device__ double gather_block_mass_wave(const GatherHist &h, const double *s_pairs, int v, uint32_t e, int lane)
{
	const uint32_t freeb = ((1u << h.n) - 1u) & ~e;
	const int f = __popc(freeb);
	double acc = 0.0;
	for (uint32_t j = (uint32_t)lane; j < (1u << f); j += 64u)
		acc += gather_weight(h, s_pairs, v, gather_superset(e, freeb, f, j));
	for (int o = 32; o > 0; o >>= 1)
		acc += __shfl_xor(acc, o);
	return acc;
}

__global__ __launch_bounds__(256) void k_gather_exact(Descs descs, int k, double target)
{
	const LatDesc &d = descs.many ? descs.many[blockIdx.y] : descs.one;
	if (!gather_redecides(d))
		return;
	const unsigned count = d.counter[0];
	if (blockIdx.x >= count)
		return;
	__shared__ double s_pairs[2 * kMaxAtoms];
	__shared__ double red[16];
	__shared__ GatherHist s_hist;
	for (int i = threadIdx.x; i < 2 * k * d.gx_n; i += blockDim.x)
		s_pairs[i] = d.gh_pairs[i];
	stage_gather(d, s_hist);
	double u0[kMaxVariants];
	for (int v = 0; v < k; ++v)
		u0[v] = gather_block_mass(s_hist, s_pairs, v, 0u, red);
	for (unsigned i = blockIdx.x; i < count; i += gridDim.x) {
		const uint32_t e = frame_deposit(d.gband[i].e, d.gx_keep);
		double mv[kMaxVariants];
		for (int v = 0; v < k; ++v)
			mv[v] = gather_block_mass(s_hist, s_pairs, v, e, red) / u0[v];
		const double score = gather_objective(mv, k, target);
		if (threadIdx.x == 0)
			d.gexact[i] = score;
	}
}

// the exact best two of the band by (objective, experiment) into d.result
__global__ __launch_bounds__(1024) void k_gather_pick(Descs descs)
{
	const LatDesc &d = descs.many ? descs.many[blockIdx.y] : descs.one;
	if (!gather_redecides(d))
		return;
	__shared__ Top2 red2[16];
	const unsigned count = d.counter[0];
	Top2 t;
	top2_init(t);
	for (unsigned i = threadIdx.x; i < count; i += blockDim.x)
		top2_insert(t, d.gexact[i], d.gband[i].e);
	t = top2_block(t, red2);
	if (threadIdx.x == 0) {
		d.result->d1 = t.d1;
		d.result->d2 = t.d2;
		d.result->e1 = t.e1;
		d.result->e2 = t.e2;
		d.counter[kGatherFlag] = 1u;
	}
}

// The rest of an Op-BBPA selection on the one-pass tables, in one workgroup per lattice
// (every experiment of at most 10 subjects one thread): the tables as k_bbpa_tables, each
// experiment's objective as k_bbpa_score (the same operations), the top two as
// k_finalize_top2 (the order (|objective|, experiment) is total, so the same two), the band as
// k_bbpa_band, and at a near tie the exact re-decision: every state S of every variant block v
// gets its weight w_v(S) (the prior factors of the block's atoms times the history's factors
// for variant v) in LDS, the superset sums U_v(E) of every block are formed there in place
// (one subject bit at a time, so each is a sum tree of depth n), and thread i takes band
// entry i's per-variant binary masses U_v(E) / U_v({}) and forms its exact objective; the
// exact best two go into d.result (counter[kGatherFlag] = 1). Six launches become one. The
// one-pass path holds at most 2^10 states per block and 2048 weights in all
// ((2^k - 1) 2^(n-6) <= 64).
constexpr int kFinishWeights = 2048;
__global__ __launch_bounds__(1024) void k_bbpa_finish(Descs descs, int fused, int k, double target)
{
	const LatDesc &d = descs.many ? descs.many[blockIdx.y] : descs.one;
	__shared__ Top2 red2[16];
	__shared__ double red[16];
	__shared__ double s_pairs[2 * kMaxAtoms];
	__shared__ double s_w[kFinishWeights];
	__shared__ GatherHist s_hist;
	__shared__ unsigned s_count;
	__shared__ double s_lim;
	__shared__ int s_redecide;
	const int n = d.n;
	const uint32_t ec = 1u << n;
	const unsigned cells = ((1u << k) - 1u) * ec;
	unsigned long long *g = d.fixtab;
	// 1. the tables (k_bbpa_tables)
	const double t = fix_value(&g[3u * cells]);
	const double scale = fused ? 1.0 / t : 1.0;
	for (unsigned i = threadIdx.x; i < cells; i += blockDim.x)
		d.ftab[(size_t)ec + i] = fix_value(&g[3u * i]) * scale;
	for (unsigned i = threadIdx.x; i < (unsigned)(k * n); i += blockDim.x) {
		const unsigned v = i / n, a = i % n;
		d.masses[i] = fix_value(&g[3u * (((1u << v) - 1u) * ec + (1u << a))]) * scale;
	}
	const double total = fused ? t * scale : t;
	if (threadIdx.x == 0) {
		if (fused) {
			d.scal[kScalDen] = t;
			d.scal[d.inv_out] = scale;
			d.scal[kScalTotal] = total;
		} else {
			d.scal[kScalTotal] = t;
		}
		s_count = 0u;
	}
	__syncthreads(); // every thread has read the bins; the table writes are visible
	for (unsigned i = threadIdx.x; i < 3u * (cells + 1u); i += blockDim.x)
		g[i] = 0ull;
	// 2. objectives and the top two (k_bbpa_score, k_finalize_top2)
	const uint32_t e = threadIdx.x;
	double score = INFINITY;
	const bool mine = e < ec && (e & ~d.active) == 0u;
	if (mine)
		score = bbpa_objective(d.ftab, ec, e, k, total, target);
	Top2 tt;
	top2_init(tt);
	if (mine)
		top2_insert(tt, score, e);
	tt = top2_block(tt, red2);
	if (threadIdx.x == 0) {
		d.result->d1 = fabs(tt.d1);
		d.result->d2 = fabs(tt.d2);
		d.result->e1 = tt.e1;
		d.result->e2 = tt.e2;
		d.counter[0] = 0;
		d.counter[1] = 0;
		d.counter[kGatherFlag] = 0;
		s_lim = fabs(tt.d1) + d.band_width;
		s_redecide = d.gband != nullptr && d.gh_len >= 0 && (fabs(tt.d2) - fabs(tt.d1)) <= d.band_width;
	}
	__syncthreads();
	if (d.gband == nullptr)
		return;
	// 3. the band (k_bbpa_band): every selectable experiment within band_width of the best
	const bool in = mine && score <= s_lim;
	unsigned at = 0u;
	if (in)
		at = atomicAdd(&s_count, 1u);
	if (in)
		d.gband[at] = Cand{score, e, 0u};
	__syncthreads();
	const unsigned count = s_count;
	if (threadIdx.x == 0)
		d.counter[0] = count;
	if (!s_redecide)
		return;
	// 4. the exact re-decision in the program's frame (xn subjects; a projected lattice's band
	// entries are placed in it): the block weights, their normalisers, then one band entry per
	// thread; a frame whose blocks exceed the weights' room (a projected lattice keeps its
	// original frame) sums the normalisers with the whole workgroup and each entry's block
	// masses with one wavefront instead (with the whole workgroup beyond kWaveMassBits
	// subjects)
	const int xn = d.gx_n;
	for (int i = threadIdx.x; i < 2 * k * xn; i += blockDim.x)
		s_pairs[i] = d.gh_pairs[i];
	stage_gather(d, s_hist);
	Top2 ex;
	top2_init(ex);
	if ((k << xn) <= kFinishWeights) {
		const uint32_t xc = 1u << xn;
		for (uint32_t idx = threadIdx.x; idx < (uint32_t)k * xc; idx += blockDim.x) {
			const int v = (int)(idx >> xn);
			const uint32_t st = idx & (xc - 1u);
			s_w[idx] = gather_weight(s_hist, s_pairs, v, st);
		}
		__syncthreads();
		// every block's superset sums in place, one subject bit at a time (each sum a tree of
		// depth xn): afterwards s_w[v][S] = U_v(S) and s_w[v][0] = U_v({}), the normaliser
		const uint32_t half = xc >> 1;
		for (int b = 0; b < xn; ++b) {
			const uint32_t low = (1u << b) - 1u;
			for (uint32_t idx = threadIdx.x; idx < (uint32_t)k * half; idx += blockDim.x) {
				const uint32_t v = idx / half, r = idx % half;
				const uint32_t st = ((r & ~low) << 1) | (r & low); // bit b clear
				double *w = s_w + (v << xn);
				w[st] += w[st | (1u << b)];
			}
			__syncthreads();
		}
		if (threadIdx.x < count) {
			const uint32_t ce = d.gband[threadIdx.x].e;
			const uint32_t xe = frame_deposit(ce, d.gx_keep);
			double mv[kMaxVariants];
			for (int v = 0; v < k; ++v)
				mv[v] = s_w[((uint32_t)v << xn) + xe] / s_w[(uint32_t)v << xn];
			top2_insert(ex, gather_objective(mv, k, target), ce);
		}
	} else if (xn <= kWaveMassBits) {
		// the normalisers by the whole workgroup, then one wavefront per band entry (no
		// workgroup barrier per entry)
		double u0[kMaxVariants];
		for (int v = 0; v < k; ++v)
			u0[v] = gather_block_mass(s_hist, s_pairs, v, 0u, red);
		const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
		for (unsigned i = (unsigned)wave; i < count; i += (unsigned)waves) {
			const uint32_t ce = d.gband[i].e;
			const uint32_t xe = frame_deposit(ce, d.gx_keep);
			double mv[kMaxVariants];
			for (int v = 0; v < k; ++v)
				mv[v] = gather_block_mass_wave(s_hist, s_pairs, v, xe, lane) / u0[v];
			if (lane == 0)
				top2_insert(ex, gather_objective(mv, k, target), ce);
		}
	} else {
		double u0[kMaxVariants];
		for (int v = 0; v < k; ++v)
			u0[v] = gather_block_mass(s_hist, s_pairs, v, 0u, red);
		for (unsigned i = 0; i < count; ++i) {
			const uint32_t ce = d.gband[i].e;
			const uint32_t xe = frame_deposit(ce, d.gx_keep);
			double mv[kMaxVariants];
			for (int v = 0; v < k; ++v)
				mv[v] = gather_block_mass(s_hist, s_pairs, v, xe, red) / u0[v];
			if (threadIdx.x == 0)
				top2_insert(ex, gather_objective(mv, k, target), ce);
		}
	}
	ex = top2_block(ex, red2);
	if (threadIdx.x == 0) {
		d.result->d1 = ex.d1;
		d.result->d2 = ex.d2;
		d.result->e1 = ex.e1;
		d.result->e2 = ex.e2;
		d.counter[kGatherFlag] = 1u;
	}
}

// Explicit (experiment, state) sweep (benchmarks/experiment_selection_benchmark.cpp:214-247):
// one workgroup per experiment, buckets per thread, fixed-order reductions.
template <typename TP, int K>
__global__ __launch_bounds__(256) void k_brute_force(Descs descs, int ns)
{
	constexpr int RC = 1 << K;
	const LatDesc &d = descs.many ? descs.many[blockIdx.y] : descs.one;
	__shared__ double red[16];
	__shared__ Top2 red2[16];
	const uint32_t e = blockIdx.x;
	const uint32_t states = 1u << (ns * K);
	const uint32_t mask = (1u << ns) - 1u;
	gptr<const TP> post = G(static_cast<const TP *>(d.post));
	const double inv = d.scal[d.inv_in];
	double mass[RC];
#pragma unroll
	for (int b = 0; b < RC; ++b)
		mass[b] = 0.0;
	for (uint32_t s = threadIdx.x; s < states; s += blockDim.x) {
		const double x = static_cast<double>(post[s]) * inv;
		int id = 0;
#pragma unroll
		for (int v = 0; v < K; ++v)
			id |= ((e & ~((s >> (v * ns)) & mask)) != 0) << v;
#pragma unroll
		for (int b = 0; b < RC; ++b)
			mass[b] += (id == b) ? x : 0.0;
	}
	const double target = 1.0 / (double)RC;
	double score = 0.0;
	for (int b = 0; b < RC; ++b) {
		const double mb = block_sum(mass[b], red);
		score += fabs(mb - target); // meaningful in thread 0
	}
	Top2 t;
	top2_init(t);
	if (threadIdx.x == 0 && (e & ~d.active) == 0)
		top2_insert(t, score, e);
	publish_top2(d, blockIdx.x, t, red2);
}

// Atom marginals m_a = sum_{S : bit a} x_S (lattice_base.cpp:25-41). A workgroup covers
// 2^14 contiguous states: thread bits are state bits 0..7, register index bits are 8..13,
// and higher bits are uniform over the chunk.
template <typename TP>
__global__ __launch_bounds__(256) void k_atom_masses(Descs descs, int atoms)
{
	const LatDesc &d = descs.many ? descs.many[blockIdx.y] : descs.one;
	if (descs.ragged) {
		atoms = d.atoms;
		if (blockIdx.x >= mass_groups(atoms))
			return;
	}
	__shared__ double red[16];
	const int cb = mass_chunk_bits(atoms); // chunk bits
	const int jb = cb > 8 ? cb - 8 : 0;     // register bits
	const uint32_t chunk = blockIdx.x;
	const uint64_t base = (uint64_t)chunk << cb;
	gptr<const TP> post = G(static_cast<const TP *>(d.post));
	const double inv = d.scal[d.inv_in];
	const uint32_t tid = threadIdx.x;
	double total = 0.0;
	double sq[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
	const bool live = (cb >= 8) || tid < (1u << cb);
	if (live) {
		const uint32_t J = 1u << jb;
		for (uint32_t j0 = 0; j0 < J; j0 += 16) {
			double x[16];
#pragma unroll
			for (int q = 0; q < 16; ++q)
				x[q] = (j0 + q < J) ? static_cast<double>(post[base + tid + ((j0 + q) << 8)]) : 0.0;
#pragma unroll
			for (int q = 0; q < 16; ++q) {
				const uint32_t j = j0 + q;
				const double xv = x[q] * inv;
				if (j < J) {
					total += xv;
#pragma unroll
					for (int b = 0; b < 6; ++b)
						if (b < jb && ((j >> b) & 1u))
							sq[b] += xv;
				}
			}
		}
	}
	for (int a = 0; a < atoms; ++a) {
		double contrib;
		if (a < 8) {
			contrib = ((tid >> a) & 1u) ? total : 0.0;
		} else if (a < 8 + jb) {
			contrib = 0.0;
#pragma unroll
			for (int q = 0; q < 6; ++q)
				if (a - 8 == q)
					contrib = sq[q];
		} else {
			contrib = ((chunk >> (a - cb)) & 1u) ? total : 0.0;
		}
		const double s = block_sum(contrib, red);
		if (threadIdx.x == 0)
			d.part[(size_t)chunk * kMaxAtoms + a] = s;
	}
}

// Upset mass M(E) = sum over supersets of E (lattice_base.cpp:318-351), enumerating the
// free bits of E.
template <typename TP>
__global__ __launch_bounds__(256) void k_upset_mass(Descs descs, int atoms, uint32_t e)
{
	const LatDesc &d = descs.many ? descs.many[blockIdx.y] : descs.one;
	__shared__ double red[16];
	__shared__ int pos[kMaxAtoms];
	const uint32_t full = atoms == 32 ? 0xFFFFFFFFu : ((1u << atoms) - 1u);
	const uint32_t freeb = full & ~e;
	const int m = __popc(freeb);
	if (threadIdx.x == 0) {
		int c = 0;
		for (int b = 0; b < atoms; ++b)
			if ((freeb >> b) & 1u)
				pos[c++] = b;
	}
	__syncthreads();
	gptr<const TP> post = G(static_cast<const TP *>(d.post));
	const double inv = d.scal[d.inv_in];
	const uint64_t count = 1ull << m;
	double acc = 0.0;
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
		uint32_t s = e;
		for (int b = 0; b < m; ++b)
			if ((i >> b) & 1u)
				s |= 1u << pos[b];
		acc += static_cast<double>(post[s]) * inv;
	}
	const double bs = block_sum(acc, red);
	if (threadIdx.x == 0)
		d.part[blockIdx.x] = bs;
}

// Projection of a posterior onto the atoms it keeps (project_posterior, shrink.hpp:70-86):
// entry b of the destination is the sum, in increasing order of the removed atoms' index,
// of the source entries base(b) | red, where base(b) places b's bits at the kept atoms and
// red runs over the subsets of the removed atoms. Sources are read as x = y * (1/D) and
// summed in FP64 (the reference's accumulator_t); the destination holds the sums.
template <typename TS, typename TD>
__global__ __launch_bounds__(256) void k_project(const TS *src, const double *src_scal, int inv_slot, TD *dst,
												 uint64_t dst_states, uint32_t removed)
{
	const double inv = src_scal[inv_slot];
	for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; b < dst_states;
		 b += (uint64_t)gridDim.x * blockDim.x) {
		// b's bits at the kept atoms: a zero inserted at each removed atom, lowest first
		uint32_t base = (uint32_t)b;
		for (uint32_t m = removed; m; m &= m - 1u) {
			const uint32_t low = (m & (0u - m)) - 1u;
			base = ((base & ~low) << 1) | (base & low);
		}
		double sum = 0.0;
		// subsets of the removed atoms in increasing order (0, then (red - removed) & removed)
		uint32_t red = 0;
		do {
			sum += static_cast<double>(src[base | red]) * inv;
			red = (red - removed) & removed;
		} while (red != 0u);
		dst[b] = static_cast<TD>(sum);
	}
}

// Projections that sum many source entries into each kept entry (2^r terms, r > kProjectSerialBits:
// typically most subjects classified at once, leaving few kept entries whose serial sums
// above would each run 2^r dependent additions in a handful of threads): a workgroup sums a
// contiguous chunk of one kept entry's terms (the subsets of the removed atoms by index),
// each thread its strided share in index order, then the workgroup tree; k_project_join adds
// an entry's chunk sums in chunk order. Every source entry is read once, in a fixed order of
// additions, so the result does not depend on timing.
constexpr int kProjectSerialBits = 6;
template <typename TS, typename TD>
__global__ __launch_bounds__(256) void k_project_wide(const TS *src, const double *src_scal, int inv_slot, TD *dst,
													  double *partial, uint64_t dst_states, uint32_t removed, int r,
													  uint32_t chunks)
{
	__shared__ double red[4];
	const double inv = src_scal[inv_slot];
	const uint64_t per = (1ull << r) / chunks;
	for (uint64_t w = blockIdx.x; w < dst_states * chunks; w += gridDim.x) {
		const uint64_t b = w / chunks;
		const uint64_t g = w % chunks;
		uint32_t base = (uint32_t)b;
		for (uint32_t m = removed; m; m &= m - 1u) {
			const uint32_t low = (m & (0u - m)) - 1u;
			base = ((base & ~low) << 1) | (base & low);
		}
		double sum = 0.0;
		for (uint64_t j = g * per + threadIdx.x; j < (g + 1) * per; j += blockDim.x) {
			// j's bits deposited at the removed atoms, lowest first
			uint32_t sub = 0u;
			uint64_t q = j;
			for (uint32_t m = removed; q != 0u; m &= m - 1u, q >>= 1)
				sub |= (q & 1u) ? (m & (0u - m)) : 0u;
			sum += static_cast<double>(src[base | sub]) * inv;
		}
		const double s = block_sum(sum, red);
		if (threadIdx.x == 0) {
			if (chunks == 1)
				dst[b] = static_cast<TD>(s);
			else
				partial[w] = s;
		}
	}
}

template <typename TD>
__global__ __launch_bounds__(256) void k_project_join(const double *partial, TD *dst, uint64_t dst_states,
													  uint32_t chunks)
{
	for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; b < dst_states;
		 b += (uint64_t)gridDim.x * blockDim.x) {
		double sum = 0.0;
		for (uint32_t g = 0; g < chunks; ++g)
			sum += partial[b * chunks + g];
		dst[b] = static_cast<TD>(sum);
	}
}

template <typename TP>
__global__ __launch_bounds__(256) void k_read_posterior(Descs descs, uint64_t states, double *out)
{
	const LatDesc &d = descs.many ? descs.many[0] : descs.one;
	gptr<const TP> post = G(static_cast<const TP *>(d.post));
	const double inv = d.scal[d.inv_in];
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < states; s += stride)
		out[s] = static_cast<double>(post[s]) * inv;
}

// k-step lookahead (lattice_api.cpp:116-168), PUBLIC coordinates: for candidate c, the
// mass of the 2^(m+1) groups formed by the m selected experiments and c.
template <typename TP>
__global__ __launch_bounds__(256) void k_lookahead(Descs descs, int n, const uint32_t *selected,
												   int m, double *scores)
{
	const LatDesc &d = descs.many ? descs.many[0] : descs.one;
	__shared__ double red[16];
	__shared__ uint32_t sel[16];
	if (threadIdx.x < (unsigned)m)
		sel[threadIdx.x] = selected[threadIdx.x];
	__syncthreads();
	const uint32_t c = blockIdx.x;
	const uint32_t states = 1u << n;
	gptr<const TP> post = G(static_cast<const TP *>(d.post));
	const double inv = d.scal[d.inv_in];
	const int groups = 1 << (m + 1);
	double mass[32];
#pragma unroll
	for (int g = 0; g < 32; ++g)
		mass[g] = 0.0;
	for (uint32_t s = threadIdx.x; s < states; s += blockDim.x) {
		// public state s reads internal state reverse_bits(s, n)
		const uint32_t internal = __brev(s) >> (32 - n);
		const double x = static_cast<double>(post[internal]) * inv;
		int gi = 0;
		for (int i = 0; i < m; ++i)
			if ((sel[i] & s) == s)
				gi |= 1 << i;
		if ((c & s) == s)
			gi |= 1 << m;
#pragma unroll
		for (int g = 0; g < 32; ++g)
			if (g == gi)
				mass[g] += x;
	}
	const double target = 1.0 / (double)groups;
	double score = 0.0;
#pragma unroll
	for (int g = 0; g < 32; ++g) {
		if (g < groups) {
			const double mg = block_sum(mass[g], red);
			score += fabs(mg - target);
		}
	}
	if (threadIdx.x == 0)
		scores[c] = score;
}

// ------------------------------------------- k-step lookahead through down-set sums
//
// The reference scores each candidate c by the masses of the 2^(m+1) groups the states fall
// into by which of the m selected experiments and c contain them (s <= e_i), summing the
// posterior once per candidate (lattice_api.cpp:116-168, O(4^n)). With the down-set sums
// D(X) = sum_{s <= X} x(s), the states contained in every experiment of a set K have mass
// F(K) = D(the intersection of K's experiments), and a group's mass is the Moebius
// inversion sum_{K >= J} (-1)^{|K \ J|} F(K): O(2^n) for D once, then 2^(m+1) lookups per
// candidate. Subset relations are kept by the public bit reversal, so D is built in internal
// coordinates; candidates are ranked by public index, as the reference's loop.

// Pass A: D over bits 0..11. A wave owns 4096 contiguous entries: lane = bits 0..5 (every
// load and store a 512-byte run in FP64), register r = bits 6..11. x = y * (1/D) is read from
// the posterior and summed in FP64; no LDS and no barrier.
constexpr int kDownFirstBits = 12;
template <typename TP>
__global__ __launch_bounds__(256) void k_down_first(Descs descs, double *D)
{
	const LatDesc &d = descs.many ? descs.many[0] : descs.one;
	const uint32_t lane = threadIdx.x & 63u;
	const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
	const uint64_t base = (wave << kDownFirstBits) | lane;
	gptr<const TP> post = G(static_cast<const TP *>(d.post));
	const double inv = d.scal[d.inv_in];
	double v[64];
#pragma unroll
	for (int r = 0; r < 64; ++r)
		v[r] = static_cast<double>(post[base + ((uint64_t)r << 6)]) * inv;
#pragma unroll
	for (int b = 0; b < 6; ++b)
#pragma unroll
		for (int r = 0; r < 64; ++r)
			if (r & (1 << b))
				v[r] += v[r ^ (1 << b)];
#pragma unroll
	for (int b = 0; b < 6; ++b) {
		const bool up = (lane >> b) & 1u; // bit b set: add the partner's value (bit b clear)
#pragma unroll
		for (int r = 0; r < 64; ++r) {
			const double o = __shfl_xor(v[r], 1 << b);
			v[r] += up ? o : 0.0;
		}
	}
#pragma unroll
	for (int r = 0; r < 64; ++r)
		D[base + ((uint64_t)r << 6)] = v[r];
}

// Pass B: D over bits b0..b0+R-1 (b0 >= 12, R <= 6) in registers: a thread owns one column
// below b0 (consecutive columns across the lanes, so 512-byte runs) and its 2^R entries; grid
// row y works on the transform at D + y * row_stride.
template <int R>
__global__ __launch_bounds__(256) void k_down_more(double *D, int b0, uint64_t row_stride)
{
	D += blockIdx.y * row_stride; // several masked transforms side by side (grid y)
	const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	const uint64_t col = t & ((1ull << b0) - 1ull), hi = t >> b0;
	const uint64_t base = col | (hi << (b0 + R));
	double v[1 << R];
#pragma unroll
	for (int r = 0; r < (1 << R); ++r)
		v[r] = D[base + ((uint64_t)r << b0)];
#pragma unroll
	for (int b = 0; b < R; ++b)
#pragma unroll
		for (int r = 0; r < (1 << R); ++r)
			if (r & (1 << b))
				v[r] += v[r ^ (1 << b)];
#pragma unroll
	for (int r = 0; r < (1 << R); ++r)
		D[base + ((uint64_t)r << b0)] = v[r];
}

// Scores of every candidate into scores[internal index] (candidate 0 scores +inf), and each
// block's lowest score into best_score[blockIdx.x]. M experiments are selected (1..4), so
// every register index below is a compile-time constant.
template <int M>
__global__ __launch_bounds__(256) void k_lookahead_down(const double *D, LookStep st, double *scores,
														double *best_score)
{
	constexpr int K = 1 << M;       // subsets of the selected experiments
	constexpr int GROUPS = 2 * K;   // with the candidate
	__shared__ double s_sc[256];
	const uint32_t states = 1u << st.n;
	const double target = 1.0 / (double)GROUPS;
	double bs = INFINITY;
	// the candidate-independent masses F(K), K without the candidate
	double fk[K];
	uint32_t inter[K];
#pragma unroll
	for (int k = 0; k < K; ++k) {
		inter[k] = st.inter[k];
		fk[k] = D[inter[k]];
	}
	for (uint32_t ci = blockIdx.x * blockDim.x + threadIdx.x; ci < states; ci += gridDim.x * blockDim.x) {
		// g[J]: J's bit M is the candidate; F(K) first, then Moebius inversion over the
		// superset lattice: mass(J) = sum_{K >= J} (-1)^|K\J| F(K)
		double g[GROUPS];
#pragma unroll
		for (int k = 0; k < K; ++k) {
			g[k] = fk[k];
			g[k | K] = D[inter[k] & ci];
		}
#pragma unroll
		for (int b = 0; b <= M; ++b) {
#pragma unroll
			for (int j = 0; j < GROUPS; ++j)
				if (!(j & (1 << b)))
					g[j] -= g[j | (1 << b)];
		}
		double score = 0.0;
#pragma unroll
		for (int j = 0; j < GROUPS; ++j)
			score += fabs(g[j] - target);
		score = ci == 0u ? INFINITY : score; // the empty experiment is not a candidate
		scores[ci] = score;
		bs = fmin(bs, score);
	}
	s_sc[threadIdx.x] = bs;
	__syncthreads();
	for (int o = blockDim.x >> 1; o > 0; o >>= 1) {
		if ((int)threadIdx.x < o)
			s_sc[threadIdx.x] = fmin(s_sc[threadIdx.x], s_sc[threadIdx.x + o]);
		__syncthreads();
	}
	if (threadIdx.x == 0)
		best_score[blockIdx.x] = s_sc[0];
}

// Each block's lowest public index whose score is at most `limit` (0xFFFFFFFF if none). The
// grid is the scoring launch's, so a block whose lowest score exceeds the limit has nothing.
__global__ __launch_bounds__(256) void k_lookahead_pick(const double *scores, const double *block_best, int n,
														double limit, uint32_t *best_exp)
{
	__shared__ uint32_t s_ex[256];
	if (block_best[blockIdx.x] > limit) {
		if (threadIdx.x == 0)
			best_exp[blockIdx.x] = 0xFFFFFFFFu;
		return;
	}
	const uint32_t states = 1u << n;
	uint32_t be = 0xFFFFFFFFu;
	for (uint32_t ci = blockIdx.x * blockDim.x + threadIdx.x; ci < states; ci += gridDim.x * blockDim.x)
		if (scores[ci] <= limit)
			be = min(be, __brev(ci) >> (32 - n));
	s_ex[threadIdx.x] = be;
	__syncthreads();
	for (int o = blockDim.x >> 1; o > 0; o >>= 1) {
		if ((int)threadIdx.x < o)
			s_ex[threadIdx.x] = min(s_ex[threadIdx.x], s_ex[threadIdx.x + o]);
		__syncthreads();
	}
	if (threadIdx.x == 0)
		best_exp[blockIdx.x] = s_ex[0];
}

// ------------------------------- k-step lookahead after five or more earlier picks
//
// With m >= 5 earlier picks the 2^(m+1) group masses of a candidate no longer fit in
// registers. The score is regrouped by the class J of a state under the m earlier picks (bit
// i: the state lies inside pick i). A candidate c splits class J into the states inside c,
// of mass H_J(c) = sum_{s <= c, class(s) = J} x(s) (the down-set sums of the posterior masked
// to class J), and the rest, of mass H_J(full) - H_J(c); the score is
// sum_J |H_J(full) - H_J(c) - t| + |H_J(c) - t| with t = 2^-(m+1). A step costs 2^m masked
// down-set transforms of 2^n entries, every sum over non-negative terms.

__device__ __forceinline__ uint32_t look_class(uint32_t s, const LookClasses &lc)
{
	uint32_t j = 0;
	for (int i = 0; i < lc.m; ++i)
		j |= ((lc.sel[i] & s) == s ? 1u : 0u) << i;
	return j;
}

// Up to kLookLdsMax subjects: one workgroup per class (j0 + blockIdx.x), the masked posterior
// and its down-set sums in LDS (2^n doubles), the class's row of sums into P (row stride 2^n).
template <typename TP>
__global__ __launch_bounds__(256) void k_look_class_lds(Descs descs, LookClasses lc, uint32_t j0, double *P)
{
	extern __shared__ double L[];
	const LatDesc &d = descs.many ? descs.many[0] : descs.one;
	const uint32_t states = 1u << lc.n;
	const uint32_t j = j0 + blockIdx.x;
	gptr<const TP> post = G(static_cast<const TP *>(d.post));
	const double inv = d.scal[d.inv_in];
	for (uint32_t s = threadIdx.x; s < states; s += blockDim.x)
		L[s] = look_class(s, lc) == j ? static_cast<double>(post[s]) * inv : 0.0;
	__syncthreads();
	for (int b = 0; b < lc.n; ++b) {
		const uint32_t bit = 1u << b;
		for (uint32_t h = threadIdx.x; h < (states >> 1); h += blockDim.x) {
			// the h-th state with bit b set: h's bits below b, bit b, h's higher bits above it
			const uint32_t s = ((h >> b) << (b + 1)) | bit | (h & (bit - 1u));
			L[s] += L[s ^ bit];
		}
		__syncthreads();
	}
	double *row = P + (size_t)blockIdx.x * states;
	for (uint32_t s = threadIdx.x; s < states; s += blockDim.x)
		row[s] = L[s];
}

// From 13 subjects: the first 12 bits of the down-set sums of the posterior masked to class
// j0 + blockIdx.y, into row blockIdx.y of P (row stride 2^atoms); the layout of k_down_first.
template <typename TP>
__global__ __launch_bounds__(256) void k_look_class_first(Descs descs, LookClasses lc, uint32_t j0, double *P)
{
	const LatDesc &d = descs.many ? descs.many[0] : descs.one;
	const uint32_t lane = threadIdx.x & 63u;
	const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
	const uint64_t base = (wave << kDownFirstBits) | lane;
	const uint32_t j = j0 + blockIdx.y;
	double *D = P + ((uint64_t)blockIdx.y << lc.n);
	gptr<const TP> post = G(static_cast<const TP *>(d.post));
	const double inv = d.scal[d.inv_in];
	double v[64];
#pragma unroll
	for (int r = 0; r < 64; ++r) {
		const uint64_t s = base + ((uint64_t)r << 6);
		v[r] = look_class((uint32_t)s, lc) == j ? static_cast<double>(post[s]) * inv : 0.0;
	}
#pragma unroll
	for (int b = 0; b < 6; ++b)
#pragma unroll
		for (int r = 0; r < 64; ++r)
			if (r & (1 << b))
				v[r] += v[r ^ (1 << b)];
#pragma unroll
	for (int b = 0; b < 6; ++b) {
		const bool up = (lane >> b) & 1u;
#pragma unroll
		for (int r = 0; r < 64; ++r) {
			const double o = __shfl_xor(v[r], 1 << b);
			v[r] += up ? o : 0.0;
		}
	}
#pragma unroll
	for (int r = 0; r < 64; ++r)
		D[base + ((uint64_t)r << 6)] = v[r];
}

// Adds classes 0..rows-1 of P (row stride 2^n, row r holding H_J for J = j0 + r) to every
// candidate's running score, scores[internal index], class by class in ascending order.
// `first` starts the scores at zero; `last` closes them: the empty experiment scores +inf and
// each block's lowest score goes to best_score[blockIdx.x] (the grid of k_lookahead_pick).
__global__ __launch_bounds__(256) void k_look_class_accum(const double *P, int rows, int n, double target, int first,
														  int last, double *scores, double *best_score)
{
	__shared__ double s_sc[256];
	const uint32_t states = 1u << n;
	double bs = INFINITY;
	for (uint32_t ci = blockIdx.x * blockDim.x + threadIdx.x; ci < states; ci += gridDim.x * blockDim.x) {
		double sc = first ? 0.0 : scores[ci];
		for (int r = 0; r < rows; ++r) {
			const double *row = P + ((size_t)r << n);
			const double h = row[ci];
			sc += fabs(row[states - 1u] - h - target) + fabs(h - target);
		}
		if (last)
			sc = ci == 0u ? INFINITY : sc;
		scores[ci] = sc;
		bs = fmin(bs, sc);
	}
	if (!last)
		return;
	s_sc[threadIdx.x] = bs;
	__syncthreads();
	for (int o = blockDim.x >> 1; o > 0; o >>= 1) {
		if ((int)threadIdx.x < o)
			s_sc[threadIdx.x] = fmin(s_sc[threadIdx.x], s_sc[threadIdx.x + o]);
		__syncthreads();
	}
	if (threadIdx.x == 0)
		best_score[blockIdx.x] = s_sc[0];
}

// ------------------------------------------------------------------ tree statistics

__device__ __forceinline__ double orig_likelihood(uint32_t e, uint32_t r, uint32_t s, int n, int k,
												  int dilution, const double *dil)
{
	// replicated_lattice.cpp:8-28 on the original (unshrunk) lattice
	double ret = 1.0;
	const int len = __popc(e);
	for (int v = 0; v < k; ++v) {
		const uint32_t sv = s >> (v * n);
		const bool rb = (r >> v) & 1u;
		if (!dilution) {
			const bool contained = (e & sv) == e;
			ret *= contained ? (rb ? 0.99 : (1.0 - 0.99)) : (rb ? (1.0 - 0.99) : 0.99);
		} else {
			double p1 = 0.0;
			if (len > 0)
				p1 = dil[(len - 1) * (n + 1) + (len - __popc(e & sv))];
			ret *= rb ? p1 : 1.0 - p1;
		}
	}
	return ret;
}

// One thread per true state: apply_true_state + parse (tree.cpp:70-199) on a tree
// flattened in preorder. Statistics are accumulated per state and reduced in a fixed order.
__global__ __launch_bounds__(256) void k_tree_stats(const FlatNode *nodes, int node_count, int n, int k,
													int dilution, const double *dil, const double *pi,
													int depth, double branch, double *partials,
													unsigned *counter, double *out)
{
	__shared__ double red[16];
	const int atoms = n * k;
	const uint32_t states = 1u << atoms;
	double st[7] = {0, 0, 0, 0, 0, 0, 0};
	const uint32_t stride = gridDim.x * blockDim.x;
	for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < states; t += stride) {
		double coef = 1.0;
		for (int a = 0; a < atoms; ++a)
			coef *= ((t >> a) & 1u) ? (1.0 - pi[a]) : pi[a];
		coef = coef * 1.0;
		double pathb[32];
		double correct = 0.0, incorrect = 0.0, fp = 0.0, fn = 0.0, unclas = 0.0, es = 0.0, et = 0.0;
		int i = 0;
		pathb[0] = 1.0;
		while (i < node_count) {
			const FlatNode nd = nodes[i];
			double b;
			if (i == 0) {
				b = 1.0;
			} else {
				const double cp = pathb[nd.stage - 1] * orig_likelihood(nd.ex, nd.res, t, n, k, dilution, dil);
				if (!(cp > branch)) {
					i = nd.subtree_end;
					continue;
				}
				b = cp;
			}
			pathb[nd.stage] = b;
			if (b < branch) {
				i = nd.subtree_end;
				continue;
			}
			if (nd.nchild == 0) {
				const int index = nd.stage;
				const bool classified = __popc(nd.pos | nd.neg) == atoms;
				if (classified && nd.neg == t) {
					correct += b * coef;
				} else if (classified) {
					incorrect += b * coef;
					const double tp = (double)__popc(nd.pos), tn = (double)__popc(nd.neg);
					const double fpv = tp == 0.0 ? 0.0 : __popc((nd.neg ^ t) & t) / tp;
					const double fnv = tn == 0.0 ? 0.0 : __popc((nd.neg ^ t) & ~t) / tn;
					fp += fpv * coef * b;
					fn += fnv * coef * b;
				} else {
					unclas = unclas + b;
				}
				es = es + ceil((double)index / 1.0) * b;
				et = et + index * b;
			}
			++i;
		}
		st[0] += correct;
		st[1] += incorrect;
		st[2] += fp;
		st[3] += fn;
		st[4] += unclas * coef;
		st[5] += es * coef;
		st[6] += et * coef;
	}
	double bs[7];
	for (int q = 0; q < 7; ++q)
		bs[q] = block_sum(st[q], red);
	if (threadIdx.x == 0)
		for (int q = 0; q < 7; ++q)
			partials[(size_t)blockIdx.x * 7 + q] = bs[q];
}

// ------------------------------------------------------------------- finalize kernels
// One workgroup per lattice folds the per-workgroup partials in index order.

enum SumMode : int { kSumTotal = 0, kSumDenominator = 1, kSumUpset = 2 };

// `count` of a ragged launch: each lattice's workgroups of the producing launch
__device__ __forceinline__ unsigned ragged_count(const Descs &descs, const LatDesc &d, unsigned count, int rule)
{
	if (!descs.ragged)
		return count;
	switch (rule) {
	case kCountUpdate: return update_groups(d.atoms);
	case kCountTiles: return tile_groups(d.atoms);
	case kCountScore: return score_groups(d.n);
	case kCountMasses: return mass_groups(d.atoms);
	default: return count;
	}
}

__global__ __launch_bounds__(1024) void k_finalize_sum(Descs descs, unsigned count, int mode, int rule)
{
	const LatDesc &d = descs.many ? descs.many[blockIdx.y] : descs.one;
	count = ragged_count(descs, d, count, rule);
	__shared__ double red[16];
	double t = 0.0;
	for (unsigned i = threadIdx.x; i < count; i += blockDim.x)
		t += d.part[i];
	t = block_sum(t, red);
	if (threadIdx.x == 0) {
		if (mode == kSumDenominator) {
			d.scal[kScalDen] = t;
			d.scal[d.inv_out] = 1.0 / t;
			d.scal[kScalTotal] = t;
		} else if (mode == kSumUpset) {
			d.scal[kScalUpset] = t;
		} else {
			d.scal[kScalTotal] = t;
		}
	}
}

__global__ __launch_bounds__(1024) void k_finalize_top2(Descs descs, unsigned count, int rule)
{
	const LatDesc &d = descs.many ? descs.many[blockIdx.y] : descs.one;
	count = ragged_count(descs, d, count, rule);
	__shared__ Top2 red2[16];
	Top2 r;
	top2_init(r);
	for (unsigned i = threadIdx.x; i < count; i += blockDim.x) {
		const Top2 p = reinterpret_cast<const Top2 *>(d.tops)[i];
		top2_insert(r, p.d1, p.e1);
		top2_insert(r, p.d2, p.e2);
	}
	r = top2_block(r, red2);
	if (threadIdx.x == 0) {
		d.result->d1 = fabs(r.d1);
		d.result->d2 = fabs(r.d2);
		d.result->e1 = r.e1;
		d.result->e2 = r.e2;
		d.counter[0] = 0; // band collection (runs next, in stream order)
		d.counter[1] = 0;
		d.counter[kGatherFlag] = 0;
	}
	// workgroup flags of the band rescan, one per workgroup of the launch just finished
	if (d.wgflag != nullptr)
		for (unsigned i = threadIdx.x; i < count && i < (unsigned)kMaxFinalGroups; i += blockDim.x)
			d.wgflag[i] = 0u;
}

// Candidates within band_width of the best objective from the per-thread top-2 lists
// (entry 2i + 1 is thread i's runner-up). A thread whose runner-up is inside the band may
// hold further candidates: its workgroup is flagged for the rescan that follows, and
// counter[1] counts such runner-ups.
__global__ __launch_bounds__(256) void k_band_collect(Descs descs, unsigned entries)
{
	const LatDesc &d = descs.many ? descs.many[blockIdx.y] : descs.one;
	if (descs.ragged)
		entries = 2u * tile_groups(d.atoms) * kTileThreads; // two per thread of the final pass
	const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
	Cand c;
	c.d = INFINITY;
	c.e = kNone;
	if (i < entries && d.cand != nullptr)
		c = d.cand[i];
	const double lim = fabs(d.result->d1) + d.band_width;
	const bool in = c.e != kNone && fabs(c.d) <= lim;
	if (__ballot(in) == 0ull)
		return;
	band_append(d, in, c.d, c.e);
	if (in && (i & 1u))
		d.wgflag[i / (2u * kTileThreads)] = 1u;
	const unsigned long long ru = __ballot(in && (i & 1u));
	if (ru != 0ull && (threadIdx.x & 63) == __ffsll((long long)ru) - 1)
		atomicAdd(&d.counter[1], (unsigned)__popcll(ru));
}

// The band of a pass 2 that collected while scanning: its appended entries (relative to
// running bests) that lie within band_width of the final best. counter[0] was cleared by
// k_finalize_top2; an overflow of the appended entries is reported as an overflowed band.
// After a persistent pass 2: the merge of its workgroups' top-2 lists (k_finalize_top2's
// result; the top two by (|objective|, experiment) do not depend on how the lists are
// split, so every block finds the same) and, for lattices with a band, the compaction of
// the candidates it appended (those within band_width of the final best), in one launch.
// Pass 1 reset the counters.
__global__ __launch_bounds__(256) void k_top2_compact(Descs descs, unsigned count, unsigned raw_cap)
{
	const LatDesc &d = descs.many ? descs.many[blockIdx.y] : descs.one;
	__shared__ Top2 red2[4];
	__shared__ double s_lim;
	Top2 r;
	top2_init(r);
	for (unsigned i = threadIdx.x; i < count; i += blockDim.x) {
		const Top2 p = reinterpret_cast<const Top2 *>(d.tops)[i];
		top2_insert(r, p.d1, p.e1);
		top2_insert(r, p.d2, p.e2);
	}
	r = top2_block(r, red2);
	if (threadIdx.x == 0) {
		if (blockIdx.x == 0) {
			d.result->d1 = fabs(r.d1);
			d.result->d2 = fabs(r.d2);
			d.result->e1 = r.e1;
			d.result->e2 = r.e2;
			if (d.band == nullptr) {
				d.counter[0] = 0;
				d.counter[1] = 0;
			}
		}
		s_lim = fabs(r.d1) + d.band_width;
	}
	__syncthreads();
	if (d.band == nullptr)
		return;
	const unsigned raw = d.counter[kRawCount];
	if (raw > raw_cap) {
		if (blockIdx.x == 0 && threadIdx.x == 0)
			d.counter[0] = d.band_cap + 1u;
		return;
	}
	const double lim = s_lim;
	for (unsigned base = blockIdx.x * blockDim.x; base < raw; base += gridDim.x * blockDim.x) {
		const unsigned i = base + threadIdx.x;
		Cand c;
		c.d = INFINITY;
		c.e = kNone;
		if (i < raw)
			c = d.cand[i];
		band_append(d, c.e != kNone && fabs(c.d) <= lim, c.d, c.e);
	}
}

// Exact unnormalised upset masses of band candidates, one wave per candidate:
//   U(E) = Q(E) * prod_{t : E_t <= E} f_t(in) * sum over the free atoms of the factors of
//          the pools A_t = E_t \ E that remain
// (the program of the host's ExactHistory, replicated_lattice.cpp:8-28 likelihoods), as a
// dense dynamic program over the 2^c subsets of pools that can still be contained: atom a
// negative keeps the state (weight 1 - pi_a), positive removes its pools (weight pi_a); a
// pool's factor is applied, and its bit cleared, after its last atom. Both steps run as
// sweeps over one pool bit at a time, 2^(c-1) independent pairs per sweep spread over the
// wavefront. Every term is non-negative and every sum runs in a fixed order. Candidates
// with more than kExactMaxPools distinct pools, or histories longer than kExactMaxHist,
// get -1; candidates dominated by another band entry get -2 (see below).
// The dense pool-state program of k_exact_band with the 2^c states in registers: state
// k = lane + 64 r, RB = max(0, c - 6) register bits (R = 2^RB doubles per lane). A sweep
// over a lane bit exchanges partners with a lane shuffle, one over a register bit pairs
// registers; each element sees exactly the operations of the host's dense program, in the
// same order, so the result is bit-identical to it. Lanes beyond 2^c (c < 6) hold zeros
// and pair only with each other. Lane j < c holds pool j (pool, fin, fout); lane a < atoms
// holds atom a's prior pair (pv, qv).
template <int RB>
__device__ __forceinline__ double exact_dense_regs(int c, uint32_t relevant, uint32_t pool, double pfin,
												   double pfout, double pv, double qv)
{
	constexpr int R = 1 << RB;
	const int lane = threadIdx.x;
	const uint32_t states = 1u << c;
	double a[R], b[R];
#pragma unroll
	for (int r = 0; r < R; ++r)
		a[r] = ((uint32_t)lane + ((uint32_t)r << 6)) == states - 1u ? 1.0 : 0.0;
	// one sweep over pool bit j: OP(low, high) gives the new value of the bit-clear
	// element; the bit-set element becomes 0
	auto sweep = [&](double *v, int j, auto op) {
		if (j < 6) {
			const bool hi = (lane >> j) & 1;
#pragma unroll
			for (int r = 0; r < R; ++r) {
				const double o = __shfl_xor(v[r], 1 << j);
				v[r] = hi ? 0.0 : op(v[r], o);
			}
		} else {
			const int rj = j - 6;
#pragma unroll
			for (int bit = 0; bit < RB; ++bit) {
				if (bit == rj) {
#pragma unroll
					for (int r = 0; r < R; ++r)
						if (!((r >> bit) & 1)) {
							v[r] = op(v[r], v[r | (1 << bit)]);
							v[r | (1 << bit)] = 0.0;
						}
				}
			}
		}
	};
	const bool mine = lane < c;
	uint32_t done = 0;
	for (uint32_t left = relevant; left; left &= left - 1u) {
		const int at = __ffs(left) - 1;
		done |= 1u << at;
		const bool in = mine && ((pool >> at) & 1u);
		const uint32_t kill = (uint32_t)__ballot(in);
		const uint32_t close = (uint32_t)__ballot(in && (pool & ~done) == 0u);
		const double p = lane_d(pv, at), q = lane_d(qv, at);
#pragma unroll
		for (int r = 0; r < R; ++r)
			b[r] = a[r];
		for (uint32_t bits = kill; bits; bits &= bits - 1u)
			sweep(b, __ffs(bits) - 1, [](double lo, double hi) { return lo + hi; });
#pragma unroll
		for (int r = 0; r < R; ++r)
			a[r] = a[r] * q + b[r] * p;
		for (uint32_t bits = close; bits; bits &= bits - 1u) {
			const int j = __ffs(bits) - 1;
			const double fi = lane_d(pfin, j), fo = lane_d(pfout, j);
			sweep(a, j, [fi, fo](double lo, double hi) { return lo * fo + hi * fi; });
		}
	}
	return lane_d(a[0], 0); // every pool has closed: the weight sits in state 0 (lane 0)
}

__global__ __launch_bounds__(64) void k_exact_band(Descs descs, int atoms)
{
	const LatDesc &d = descs.many ? descs.many[blockIdx.y] : descs.one;
	if (d.exact_u == nullptr || d.band == nullptr)
		return;
	if (descs.ragged)
		atoms = d.atoms;
	// the newest record arrived as an argument: keep it on the device for later rounds (no
	// block reads that slot in this launch; every lane takes the argument instead)
	if (d.hist_new_at >= 0 && blockIdx.x == 0 && threadIdx.x == 0)
		d.hist[d.hist_new_at] = d.hist_new;
	// only a near tie is re-decided: the host's own test on the same two values
	if (!(d.result->d2 - d.result->d1 <= d.band_width))
		return;
	const unsigned count = min(d.counter[0], d.band_cap);
	__shared__ uint32_t s_pool[kExactMaxPools];
	__shared__ double s_fin[kExactMaxPools], s_fout[kExactMaxPools];
	const int lane = threadIdx.x; // one wavefront
	const int h = d.hist_len;
	// band entries blockIdx.x, + gridDim.x, ...: every decision below is uniform over the
	// wavefront, so the early continues keep the barriers matched
	for (unsigned i = blockIdx.x; i < count && i < (unsigned)kExactBand; i += gridDim.x) {
		const Cand me = d.band[i];
		const uint32_t e = me.e;
		// Dominance (every state has positive mass, so M strictly decreases under inclusion):
		// a candidate surely above 1/2 loses to a one-atom extension in the band that is surely
		// above 1/2, one surely below to a one-atom reduction surely below. Such candidates
		// need no mass: -2. Sure side: signed and |d| beyond half the band width. (The host
		// uses these marks only for bands of at most kExactBand entries.)
		const double half_band = 0.5 * d.band_width;
		const int side = ((me.flags & kCandSided) && fabs(me.d) > half_band) ? (me.d > 0.0 ? 1 : -1) : 0;
		if (side != 0 && count <= (unsigned)kExactBand) {
			bool dom = false;
			for (unsigned base = 0; base < count; base += 64 * 8) {
				Cand o[8]; // eight loads in flight per lane
	#pragma unroll
				for (int u = 0; u < 8; ++u) {
					const unsigned j = base + (unsigned)lane + 64u * u;
					o[u] = j < count ? d.band[j] : me;
				}
	#pragma unroll
				for (int u = 0; u < 8; ++u) {
					const uint32_t x = o[u].e ^ e;
					if (x == 0u || (x & (x - 1u)) != 0u)
						continue;
					const int os =
						((o[u].flags & kCandSided) && fabs(o[u].d) > half_band) ? (o[u].d > 0.0 ? 1 : -1) : 0;
					dom = dom || (side > 0 && os > 0 && (o[u].e & x) != 0u) ||
						  (side < 0 && os < 0 && (o[u].e & x) == 0u);
				}
			}
			if (__ballot(dom) != 0ull) {
				if (lane == 0)
					d.exact_u[i] = -2.0;
				continue;
			}
		}
		if (h > kExactMaxHist) {
			if (lane == 0)
				d.exact_u[i] = -1.0;
			continue;
		}
		// lane t stages update t (its remainder outside E and factors), lane a atom a's pair
		uint32_t rest = 0;
		double fin0 = 1.0, fout0 = 1.0;
		if (lane < h) {
			const ExactRec r = lane == d.hist_new_at ? d.hist_new : d.hist[lane];
			rest = r.e & ~e;
			fin0 = r.fin;
			fout0 = r.fout;
		}
		const double pv = lane < atoms ? d.pairs[2 * lane] : 1.0;
		const double qv = lane < atoms ? d.pairs[2 * lane + 1] : 1.0;
		// Q(E) times the always-contained factors (updates whose pool lies inside E see it
		// contained), in the host's order
		double lead = 1.0;
		for (int at = 0; at < atoms; ++at) {
			const double x = lane_d(qv, at);
			if ((e >> at) & 1u)
				lead *= x;
		}
		// The other updates group by their remaining atoms, one pool per distinct remainder in
		// order of first occurrence, with the factors of the group multiplied in update order.
		bool first = lane < h && rest != 0u;
		double fin = fin0, fout = fout0;
		for (int t = 0; t < h; ++t) {
			const uint32_t rt = lane_u(rest, t);
			const double fi = lane_d(fin0, t), fo = lane_d(fout0, t);
			if (rt == 0u)
				lead *= fi;
			if (rt == rest) {
				if (t < lane)
					first = false;
				else if (t > lane) {
					fin *= fi;
					fout *= fo;
				}
			}
		}
		const unsigned long long firsts = __ballot(first);
		const int c = __popcll(firsts);
		if (c > kExactMaxPools) {
			if (lane == 0)
				d.exact_u[i] = -1.0;
			continue;
		}
		if (first) {
			const int j = __popcll(firsts & ((1ull << lane) - 1ull));
			s_pool[j] = rest;
			s_fin[j] = fin;
			s_fout[j] = fout;
		}
		__syncthreads();
		const uint32_t pool = lane < c ? s_pool[lane] : 0u;
		const double pfin = lane < c ? s_fin[lane] : 1.0, pfout = lane < c ? s_fout[lane] : 1.0;
		uint32_t relevant = 0;
		for (int j = 0; j < c; ++j)
			relevant |= lane_u(pool, j);
		double w;
		switch (c <= 6 ? 0 : c - 6) {
		case 0: w = exact_dense_regs<0>(c, relevant, pool, pfin, pfout, pv, qv); break;
		case 1: w = exact_dense_regs<1>(c, relevant, pool, pfin, pfout, pv, qv); break;
		case 2: w = exact_dense_regs<2>(c, relevant, pool, pfin, pfout, pv, qv); break;
		case 3: w = exact_dense_regs<3>(c, relevant, pool, pfin, pfout, pv, qv); break;
		default: w = exact_dense_regs<4>(c, relevant, pool, pfin, pfout, pv, qv); break;
		}
		if (lane == 0)
			d.exact_u[i] = lead * w;
		__syncthreads(); // the next entry restages the pools
	}
}

// Exact unnormalised upset masses by enumeration (see EnumJob): workgroup (x, y) serves query
// y's share x of the rows. A row is the candidate's atoms, one assignment of the free atoms
// above the lowest kEnumLowBits, and every assignment of those lowest ones; its prior
// product and the factors of the pools without a low free atom are formed once per row.
// MANY: job blockIdx.z of a device array (several lattices' jobs in one launch, the grid
// sized for the largest; a query's workgroups do the same work as in a launch of its own).
template <bool MANY>
__global__ __launch_bounds__(256) void k_exact_enum(EnumJob one, const EnumJob *many)
{
	const EnumJob job = MANY ? many[blockIdx.z] : one;
	if (MANY && blockIdx.y >= (unsigned)job.count)
		return;
	const uint32_t e = job.queries[blockIdx.y];
	const int atoms = job.atoms;
	const uint32_t full = (1u << atoms) - 1u; // atoms <= 30
	const uint32_t freeb = full & ~e;
	const int f = __popc(freeb);
	const unsigned groups = enum_groups(f);
	if (blockIdx.x >= groups)
		return;
	__shared__ double s_pairs[2 * kMaxAtoms];
	__shared__ double s_fac[kEnumMaxHist * (kMaxSubjects + 1)];
	__shared__ uint32_t s_pool[kEnumMaxHist];
	__shared__ double s_plow[1 << kEnumLowBits];
	__shared__ uint32_t s_slow[1 << kEnumLowBits];
	__shared__ double red[16];
	const int T = job.updates, stride = job.stride;
	for (int i = threadIdx.x; i < 2 * atoms; i += blockDim.x)
		s_pairs[i] = job.pairs[i];
	for (int i = threadIdx.x; i < T * stride; i += blockDim.x)
		s_fac[i] = job.fac[i];
	for (int i = threadIdx.x; i < T; i += blockDim.x)
		s_pool[i] = job.pools[i];
	const int L = f < kEnumLowBits ? f : kEnumLowBits;
	uint32_t lowmask = 0u;
	{
		uint32_t r = freeb;
		for (int i = 0; i < L; ++i) {
			lowmask |= r & (0u - r);
			r &= r - 1u;
		}
	}
	const uint32_t highmask = freeb & ~lowmask;
	__syncthreads();
	if (threadIdx.x < (1u << L)) {
		// bit i of j assigns the i-th lowest free atom; factors in increasing atom order
		const unsigned j = threadIdx.x;
		uint32_t sl = 0u, r = lowmask;
		double p = 1.0;
		for (int i = 0; i < L; ++i) {
			const uint32_t b = r & (0u - r);
			r &= r - 1u;
			const unsigned bit = (j >> i) & 1u;
			sl |= bit ? b : 0u;
			p *= s_pairs[2 * (__ffs(b) - 1) + bit];
		}
		s_slow[j] = sl;
		s_plow[j] = p;
	}
	// pools holding a low free atom change within a row; the others once per row
	uint64_t inner = 0ull;
	for (int t = 0; t < T; ++t)
		if (s_pool[t] & lowmask)
			inner |= 1ull << t;
	// the candidate's atoms are negative in every state counted: their factors lead
	double lead = 1.0;
	for (uint32_t r = e & full; r != 0u; r &= r - 1u)
		lead *= s_pairs[2 * (__ffs(r) - 1) + 1];
	__syncthreads();
	const int fh = f - L;
	const uint64_t rows = 1ull << fh;
	const uint64_t step = (uint64_t)groups * blockDim.x;
	const unsigned lows = 1u << L;
	double acc = 0.0;
	for (uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; h < rows; h += step) {
		uint32_t sh = 0u, r = highmask;
		double p = lead;
		for (int i = 0; i < fh; ++i) {
			const uint32_t b = r & (0u - r);
			r &= r - 1u;
			const unsigned bit = (unsigned)(h >> i) & 1u;
			sh |= bit ? b : 0u;
			p *= s_pairs[2 * (__ffs(b) - 1) + bit];
		}
		const uint32_t srow = e | sh;
		for (int t = 0; t < T; ++t)
			if (!((inner >> t) & 1ull))
				p *= s_fac[t * stride + __popc(s_pool[t] & ~srow)];
		double row = 0.0;
		for (unsigned j = 0; j < lows; ++j) {
			const uint32_t st = srow | s_slow[j];
			double term = p * s_plow[j];
			for (uint64_t w = inner; w != 0ull; w &= w - 1ull) {
				const int t = __ffsll((long long)w) - 1;
				term *= s_fac[t * stride + __popc(s_pool[t] & ~st)];
			}
			row += term;
		}
		acc += row;
	}
	const double bs = block_sum(acc, red);
	if (threadIdx.x == 0)
		job.partial[(size_t)blockIdx.y * kEnumMaxGroups + blockIdx.x] = bs;
}

// the query's workgroup sums, in one fixed tree (MANY: query x of job y)
template <bool MANY>
__global__ __launch_bounds__(kEnumMaxGroups) void k_exact_enum_final(EnumJob one, const EnumJob *many)
{
	const EnumJob job = MANY ? many[blockIdx.y] : one;
	if (MANY && blockIdx.x >= (unsigned)job.count)
		return;
	const uint32_t full = (1u << job.atoms) - 1u;
	const unsigned groups = enum_groups(__popc(full & ~job.queries[blockIdx.x]));
	__shared__ double red[16];
	const double v = threadIdx.x < groups ? job.partial[(size_t)blockIdx.x * kEnumMaxGroups + threadIdx.x] : 0.0;
	const double sum = block_sum(v, red);
	if (threadIdx.x == 0)
		job.out[blockIdx.x] = sum;
}

// Result heads of a batch of lattices into one staging buffer (one copy to the host).
__global__ __launch_bounds__(256) void k_gather_heads(const LatDesc *many, unsigned char *out)
{
	// the 1 KB block from the arena's start, then the first kBandHead band candidates from
	// wherever the lattice's band lives (lattices without a band: the next 1 KB, unused)
	static_assert(kHeadBytes == 2048 && kBandHead * sizeof(Cand) == 1024, "head: 1 KB block + 1 KB of band");
	const LatDesc &d = many[blockIdx.x];
	const uint4 *src = reinterpret_cast<const uint4 *>(d.scal);
	const uint4 *band = d.band ? reinterpret_cast<const uint4 *>(d.band) : src + 1024 / sizeof(uint4);
	uint4 *dst = reinterpret_cast<uint4 *>(out + (size_t)blockIdx.x * kHeadBytes);
	for (unsigned i = threadIdx.x; i < kHeadBytes / sizeof(uint4); i += blockDim.x)
		dst[i] = i < 1024 / sizeof(uint4) ? src[i] : band[i - 1024 / sizeof(uint4)];
}

__global__ __launch_bounds__(1024) void k_finalize_masses(Descs descs, unsigned count, int atoms)
{
	const LatDesc &d = descs.many ? descs.many[blockIdx.y] : descs.one;
	if (descs.ragged) {
		atoms = d.atoms;
		count = mass_groups(atoms);
	}
	__shared__ double red[16];
	for (int a = 0; a < atoms; ++a) {
		double s = 0.0;
		for (unsigned i = threadIdx.x; i < count; i += blockDim.x)
			s += d.part[(size_t)i * kMaxAtoms + a];
		s = block_sum(s, red);
		if (threadIdx.x == 0)
			d.masses[a] = s;
	}
}

__global__ __launch_bounds__(1024) void k_finalize_tree(const double *partials, unsigned count, double *out)
{
	__shared__ double red[16];
	for (int q = 0; q < 7; ++q) {
		double s = 0.0;
		for (unsigned i = threadIdx.x; i < count; i += blockDim.x)
			s += partials[(size_t)i * 7 + q];
		s = block_sum(s, red);
		if (threadIdx.x == 0)
			out[q] = s;
	}
}

template <typename F>
void launch_checked(const char *name, F f)
{
	f();
	check_hip(hipGetLastError(), name);
}

int compute_units()
{
	static const int cus = [] {
		int dev = 0, v = 0;
		if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0)
			v = 256;
		return v;
	}();
	return cus;
}

unsigned grid_for(uint64_t work, unsigned per_block, unsigned cap)
{
	uint64_t g = (work + per_block - 1) / per_block;
	if (g < 1)
		g = 1;
	if (g > cap)
		g = cap;
	return (unsigned)g;
}

} // namespace

// ------------------------------------------------------------------------- launchers

void launch_prior(const Descs &descs, int count, int atoms, Storage st, const double *pairs,
				  uint64_t states, hipStream_t stream)
{
	const dim3 grid(grid_for(states, 256 * 8, 4096), count);
	launch_checked("k_prior", [&] {
		if (st == Storage::fp64)
			hipLaunchKernelGGL(k_prior<double>, grid, dim3(256), 0, stream, descs, atoms, pairs, states);
		else
			hipLaunchKernelGGL(k_prior<float>, grid, dim3(256), 0, stream, descs, atoms, pairs, states);
	});
}

void launch_update(const Descs &descs, int count, uint64_t states, int, int, Storage st,
				   hipStream_t stream)
{
	const dim3 grid(grid_for(states, 256 * 16, 2048), count);
	launch_checked("k_update", [&] {
		if (st == Storage::fp64)
			hipLaunchKernelGGL(k_update<double>, grid, dim3(256), 0, stream, descs, states);
		else
			hipLaunchKernelGGL(k_update<float>, grid, dim3(256), 0, stream, descs, states);
		hipLaunchKernelGGL(k_finalize_sum, dim3(1, count), dim3(1024), 0, stream, descs, grid.x, (int)kSumDenominator,
						   (int)kCountUpdate);
	});
}

void launch_zeta_pass1(const Descs &descs, int count, int n, Storage st, bool fused, int k, int dilution,
					   int objective, hipStream_t stream, bool defer_sum)
{
	const bool zblk = blocked_z(n, st, objective);
	// FP32 storage: persistent workgroups (one per CU) that prefetch their next tile, once
	// there are more tiles than CUs; otherwise one workgroup per tile
	const unsigned tiles = 1u << (n - kTileBits);
	const bool persistent =
		st == Storage::fp32c && (!fused || (k == 1 && !dilution) || !zblk) && tuned_pass1_form(n, st) != 1;
	const unsigned groups = persistent ? zeta_final_groups(n) : tiles;
	const dim3 grid(groups, count);
	const bool binary_nd = (k == 1 && !dilution);
	launch_checked("k_zeta_pass1", [&] {
		if (st == Storage::fp64) {
			if (fused && binary_nd)
				hipLaunchKernelGGL((k_zeta_pass1<double, double, true, 1>), grid, dim3(kTileThreads), 0, stream, descs);
			else if (fused)
				hipLaunchKernelGGL((k_zeta_pass1<double, double, true, 0>), grid, dim3(kTileThreads), 0, stream, descs);
			else
				hipLaunchKernelGGL((k_zeta_pass1<double, double, false, 0>), grid, dim3(kTileThreads), 0, stream, descs);
		} else if (zblk) {
			if (fused && binary_nd)
				hipLaunchKernelGGL((k_zeta_pass1<float, float, true, 1, true>), grid, dim3(kTileThreads), 0, stream, descs);
			else if (fused)
				hipLaunchKernelGGL((k_zeta_pass1<float, float, true, 0, true>), grid, dim3(kTileThreads), 0, stream, descs);
			else
				hipLaunchKernelGGL((k_zeta_pass1<float, float, false, 0, true>), grid, dim3(kTileThreads), 0, stream, descs);
		} else {
			if (fused && binary_nd)
				hipLaunchKernelGGL((k_zeta_pass1<float, float, true, 1>), grid, dim3(kTileThreads), 0, stream, descs);
			else if (fused)
				hipLaunchKernelGGL((k_zeta_pass1<float, float, true, 0>), grid, dim3(kTileThreads), 0, stream, descs);
			else
				hipLaunchKernelGGL((k_zeta_pass1<float, float, false, 0>), grid, dim3(kTileThreads), 0, stream, descs);
		}
		if (!(defer_sum && fused))
			hipLaunchKernelGGL(k_finalize_sum, dim3(1, count), dim3(1024), 0, stream, descs, tiles,
							   (int)(fused ? kSumDenominator : kSumTotal), (int)kCountTiles);
	});
}

unsigned zeta_final_groups(int atoms)
{
	if (atoms <= kTileBits)
		return 1;
	const unsigned tiles = 1u << (atoms - kTileBits);
	if (atoms >= 21) {
		unsigned g = std::min<unsigned>(tiles, (unsigned)compute_units());
		return std::max(8u, g & ~7u);
	}
	return tiles;
}

bool three_pass_form(int n, Storage st)
{
	if (st != Storage::fp32c || n < kTopLo + 1 || n > 30)
		return false;
	const int tuned = tuned_pass2_form(n, st);
	if (tuned >= 0)
		return tuned == 1;
	return n >= kThreePassFrom;
}

bool persistent_pass2(int n, int objective, bool from_posterior)
{
	return !from_posterior && n >= 21 && n <= 30 && (objective == kObjHalving || objective == kObjBinaryBbpa);
}

// the split pass 2 (FP32, 28..30 atoms, two-pass form 0) reads z in the blocked form; the
// plain pass 2 (two-pass form 2, chosen by the tuner or pinned) reads the row form
bool blocked_z(int n, Storage st, int objective)
{
	return st == Storage::fp32c && n >= 28 && persistent_pass2(n, objective, false) && !three_pass_form(n, st) &&
		   tuned_pass2_form(n, st) != 2;
}

// The final transform kernel of a selection (pass 2, or the single pass of 15..20 atoms),
// without the merge of the workgroups' top-2 lists; returns its workgroup count. `collect`
// (single pass of 15..20 atoms only) rescans the workgroups flagged by k_band_collect.
unsigned launch_final_kernel(const Descs &descs, int count, int n, Storage st, int objective, bool from_posterior,
							 bool scaled, double target, int collect, hipStream_t stream, unsigned fold = 0)
{
	if (persistent_pass2(n, objective, from_posterior) && three_pass_form(n, st)) {
		// FP32, 26..30 atoms: pass M over bits 15..24, then pass T over the rest with the objective
		const dim3 gm(zeta_final_groups(n), count);
		const dim3 gt(kTopGroups, count);
		const unsigned raw_cap = 2u * zeta_final_groups(n) * kTileThreads;
		const int sc = scaled ? 1 : 0;
		launch_checked("k_zeta_mid", [&] {
			hipLaunchKernelGGL(k_zeta_mid<float>, gm, dim3(kTileThreads), 0, stream, descs);
		});
		launch_checked("k_zeta_top", [&] {
			switch (n - kTopLo) {
			case 1: hipLaunchKernelGGL((k_zeta_top<float, 1>), gt, dim3(512), 0, stream, descs, objective, sc, target, fold, raw_cap); break;
			case 2: hipLaunchKernelGGL((k_zeta_top<float, 2>), gt, dim3(512), 0, stream, descs, objective, sc, target, fold, raw_cap); break;
			case 3: hipLaunchKernelGGL((k_zeta_top<float, 3>), gt, dim3(512), 0, stream, descs, objective, sc, target, fold, raw_cap); break;
			case 4: hipLaunchKernelGGL((k_zeta_top<float, 4>), gt, dim3(512), 0, stream, descs, objective, sc, target, fold, raw_cap); break;
			case 5: hipLaunchKernelGGL((k_zeta_top<float, 5>), gt, dim3(512), 0, stream, descs, objective, sc, target, fold, raw_cap); break;
			default: throw Error(4, "three-pass transform needs 26..30 atoms");
			}
		});
		return kTopGroups;
	}
	if (persistent_pass2(n, objective, from_posterior)) {
		// persistent pass 2: one workgroup per CU (a multiple of 8, one share per XCD)
		const dim3 grid(zeta_final_groups(n), count);
		if (blocked_z(n, st, objective)) {
			// FP32 runs of 16 bytes or less: the split form, one atom smaller and twice as wide,
			// on z in the blocked form
			launch_checked("k_zeta_pass2", [&] {
				const int sc = scaled ? 1 : 0;
				if (n == 28)
					hipLaunchKernelGGL((k_zeta_pass2<float, 3, true>), grid, dim3(kTileThreads), 0, stream, descs,
									   objective, sc, target, fold);
				else if (n == 29)
					hipLaunchKernelGGL((k_zeta_pass2<float, 2, true>), grid, dim3(kTileThreads), 0, stream, descs,
									   objective, sc, target, fold);
				else
					hipLaunchKernelGGL((k_zeta_pass2<float, 1, true>), grid, dim3(kTileThreads), 0, stream, descs,
									   objective, sc, target, fold);
			});
			return grid.x;
		}
		launch_checked("k_zeta_pass2", [&] {
#define BGT_P2(LCV)                                                                                            \
	case LCV:                                                                                                  \
		if (st == Storage::fp64)                                                                               \
			hipLaunchKernelGGL((k_zeta_pass2<double, LCV>), grid, dim3(kTileThreads), 0, stream, descs, objective, \
							   scaled ? 1 : 0, target, fold);                                                  \
		else                                                                                                   \
			hipLaunchKernelGGL((k_zeta_pass2<float, LCV>), grid, dim3(kTileThreads), 0, stream, descs, objective,  \
							   scaled ? 1 : 0, target, fold);                                                  \
		break;
			switch (2 * kTileBits - n) {
				BGT_P2(0)
				BGT_P2(1)
				BGT_P2(2)
				BGT_P2(3)
				BGT_P2(4)
				BGT_P2(5)
				BGT_P2(6)
				BGT_P2(7)
				BGT_P2(8)
				BGT_P2(9)
			}
#undef BGT_P2
		});
		return grid.x;
	}
	// n == 15: one tile, contiguous, read from the posterior.
	// n > 15: 2^(n-15) tiles of 2^(30-n) columns by 2^(n-15) rows, read from z.
	const int logc = from_posterior ? kTileBits : (2 * kTileBits - n);
	const int q0 = from_posterior ? 0 : logc;
	const dim3 grid(from_posterior ? 1u : (1u << (n - kTileBits)), count);
	launch_checked("k_zeta_final", [&] {
		if (from_posterior) {
			if (st == Storage::fp64)
				hipLaunchKernelGGL((k_zeta_final<double, true>), grid, dim3(kTileThreads), 0, stream, descs, logc, q0,
								   objective, 0, target, collect);
			else
				hipLaunchKernelGGL((k_zeta_final<float, true>), grid, dim3(kTileThreads), 0, stream, descs, logc, q0,
								   objective, 0, target, collect);
		} else {
			if (st == Storage::fp64)
				hipLaunchKernelGGL((k_zeta_final<double, false>), grid, dim3(kTileThreads), 0, stream, descs, logc,
								   q0, objective, scaled ? 1 : 0, target, collect);
			else
				hipLaunchKernelGGL((k_zeta_final<float, false>), grid, dim3(kTileThreads), 0, stream, descs, logc,
								   q0, objective, scaled ? 1 : 0, target, collect);
		}
	});
	return grid.x;
}

unsigned launch_zeta_final(const Descs &descs, int count, int n, int, Storage st, int objective,
						   bool from_posterior, bool scaled, double target, hipStream_t stream, unsigned fold,
						   const std::function<void()> &after_transform)
{
	if (fold != 0u && !persistent_pass2(n, objective, from_posterior))
		throw Error(4, "folding the pass-1 partials needs the persistent pass 2");
	const unsigned groups = launch_final_kernel(descs, count, n, st, objective, from_posterior, scaled, target, 0,
												stream, fold);
	if (after_transform)
		after_transform();
	const unsigned entries = 2u * groups * kTileThreads;
	if (persistent_pass2(n, objective, from_posterior)) {
		// the merge of the top-2 lists and the band compaction in one launch; the raw candidate
		// buffer holds two entries per thread of a persistent pass 2
		const unsigned raw_cap = 2u * zeta_final_groups(n) * kTileThreads;
		launch_checked("k_top2_compact", [&] {
			hipLaunchKernelGGL(k_top2_compact, dim3(64, count), dim3(256), 0, stream, descs, groups, raw_cap);
		});
		return raw_cap;
	}
	launch_checked("k_finalize_top2", [&] {
		hipLaunchKernelGGL(k_finalize_top2, dim3(1, count), dim3(1024), 0, stream, descs, groups, (int)kCountTiles);
	});
	return entries;
}

unsigned launch_zeta_small(const Descs &descs, int count, int n, int, Storage st, int objective, double target,
					   hipStream_t stream)
{
	const dim3 grid(1, count);
	launch_checked("k_zeta_small", [&] {
		if (st == Storage::fp64)
			hipLaunchKernelGGL(k_zeta_small<double>, grid, dim3(1024), 0, stream, descs, n, objective, target);
		else
			hipLaunchKernelGGL(k_zeta_small<float>, grid, dim3(1024), 0, stream, descs, n, objective, target);
	});
	return 2u * 1024u;
}

void launch_gather_heads(const LatDesc *many, int count, unsigned char *out, hipStream_t stream)
{
	launch_checked("k_gather_heads", [&] {
		hipLaunchKernelGGL(k_gather_heads, dim3(count), dim3(256), 0, stream, many, out);
	});
}

void launch_band_collect(const Descs &descs, int count, unsigned entries, int n, int, Storage st, int objective,
						 bool from_posterior, bool scaled, double target, hipStream_t stream)
{
	if (persistent_pass2(n, objective, from_posterior)) {
		// pass 2 appended its candidates while scanning; launch_zeta_final's k_top2_compact
		// already kept those inside the final band
		(void)entries;
		return;
	}
	const dim3 grid((entries + 255) / 256, count);
	launch_checked("k_band_collect", [&] {
		hipLaunchKernelGGL(k_band_collect, grid, dim3(256), 0, stream, descs, entries);
	});
	// rescan of the flagged workgroups (the others return at once)
	launch_final_kernel(descs, count, n, st, objective, from_posterior, scaled, target, 1, stream);
}



void launch_exact_band(const Descs &descs, int count, int atoms, hipStream_t stream, unsigned per_lattice)
{
	per_lattice = std::max(1u, std::min(per_lattice, (unsigned)kExactBand));
	launch_checked("k_exact_band", [&] {
		hipLaunchKernelGGL(k_exact_band, dim3(per_lattice, count), dim3(64), 0, stream, descs, atoms);
	});
}

static void check_enum_job(const EnumJob &job)
{
	if (job.atoms < 1 || job.atoms > kMaxAtoms || job.updates < 0 || job.updates > kEnumMaxHist || job.count < 1 ||
		job.count > 65535 || job.stride < 1 || job.stride > kMaxSubjects + 1)
		throw Error(4, "exact enumeration: shape out of range");
}

void launch_exact_enum(const EnumJob &job, hipStream_t stream)
{
	check_enum_job(job);
	// sized for the most free atoms a query can have (the empty one); each query uses its own
	const unsigned groups = enum_groups(job.atoms);
	launch_checked("k_exact_enum", [&] {
		hipLaunchKernelGGL(k_exact_enum<false>, dim3(groups, job.count), dim3(256), 0, stream, job, nullptr);
		hipLaunchKernelGGL(k_exact_enum_final<false>, dim3(job.count), dim3(kEnumMaxGroups), 0, stream, job,
						   nullptr);
	});
}

void launch_exact_enum_many(const EnumJob *jobs_dev, const std::vector<EnumJob> &jobs, hipStream_t stream)
{
	if (jobs.empty())
		return;
	if (jobs.size() > 65535)
		throw Error(4, "exact enumeration: too many jobs for one launch");
	unsigned groups = 1u, count = 1u;
	for (const EnumJob &j : jobs) {
		check_enum_job(j);
		groups = std::max(groups, enum_groups(j.atoms));
		count = std::max(count, (unsigned)j.count);
	}
	const EnumJob none{};
	launch_checked("k_exact_enum_many", [&] {
		hipLaunchKernelGGL(k_exact_enum<true>, dim3(groups, count, (unsigned)jobs.size()), dim3(256), 0, stream,
						   none, jobs_dev);
		hipLaunchKernelGGL(k_exact_enum_final<true>, dim3(count, (unsigned)jobs.size()), dim3(kEnumMaxGroups), 0,
						   stream, none, jobs_dev);
	});
}

void launch_bbpa_score(const Descs &descs, int count, int n, int k, double target, hipStream_t stream, bool scaled)
{
	const dim3 grid(grid_for(1ull << n, 256, 1u << 20), count);
	launch_checked("k_bbpa_score", [&] {
		hipLaunchKernelGGL(k_bbpa_score, grid, dim3(256), 0, stream, descs, n, k, target, scaled ? 1 : 0);
		hipLaunchKernelGGL(k_finalize_top2, dim3(1, count), dim3(1024), 0, stream, descs, grid.x, (int)kCountScore);
	});
}

void launch_bbpa_finish(const Descs &descs, int count, int k, bool fused, double target, hipStream_t stream)
{
	const int n = descs.many ? 0 : descs.one.n; // batched groups hold one shape of the one-pass path
	if (k < 2 || k > kMaxVariants || (n > 0 && ((unsigned)k << n) > (unsigned)kFinishWeights))
		throw Error(4, "launch_bbpa_finish: the one-pass path's shapes only");
	launch_checked("k_bbpa_finish", [&] {
		hipLaunchKernelGGL(k_bbpa_finish, dim3(1, count), dim3(1024), 0, stream, descs, fused ? 1 : 0, k, target);
	});
}

void launch_bbpa_band(const Descs &descs, int count, int n, int k, double target, hipStream_t stream, bool scaled)
{
	const dim3 grid(score_groups(n), count);
	launch_checked("k_bbpa_band", [&] {
		hipLaunchKernelGGL(k_bbpa_band, grid, dim3(256), 0, stream, descs, n, k, target, scaled ? 1 : 0);
		hipLaunchKernelGGL(k_gather_exact, dim3(kGatherGroups, count), dim3(256), 0, stream, descs, k, target);
		hipLaunchKernelGGL(k_gather_pick, dim3(1, count), dim3(1024), 0, stream, descs);
	});
}

bool bbpa_hist_path(int n, int k)
{
	// per lane (2^k - 1) 2^(n-6) FP64 accumulators: at most 64 (128 registers)
	if (k < 2 || k > kMaxVariants || n < 6 || n * k <= kSmallBits)
		return false;
	return ((1u << k) - 1u) << (n - 6) <= 64u;
}

size_t bbpa_fix_bytes(int n, int k)
{
	return bbpa_hist_path(n, k) ? sizeof(unsigned long long) * 3u * (((1u << k) - 1u) * (1u << n) + 1u) : 0;
}

void launch_bbpa_hist(const Descs &descs, int count, int n, int k, Storage st, bool fused, hipStream_t stream,
					  bool tables)
{
	if (!bbpa_hist_path(n, k))
		throw Error(4, "one-pass feature tables need k >= 2, n >= 6 and (2^k - 1) 2^(n-6) <= 64");
	const unsigned cells = ((1u << k) - 1u) * (1u << n);
	const size_t lds = sizeof(double) * (cells + 1u);
	const uint64_t rows = 1ull << (n * k - n);
	const int threads = hist_threads(k, n - 6);
	const unsigned waves = (unsigned)threads / 64u;
	// one 1024-thread workgroup per CU, or two of 256 threads
	const uint64_t cap = (threads == 1024 ? 1ull : 2ull) * compute_units();
	const unsigned groups = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((rows + waves - 1) / waves, cap));
	const dim3 grid(groups, count);
	launch_checked("k_bbpa_hist", [&] {
#define BGT_HIST(KV, NHV)                                                                                       \
	if (k == KV && n - 6 == NHV) {                                                                              \
		if (st == Storage::fp64) {                                                                              \
			if (fused)                                                                                          \
				hipLaunchKernelGGL((k_bbpa_hist<double, true, KV, NHV>), grid, dim3(threads), lds, stream, descs); \
			else                                                                                                \
				hipLaunchKernelGGL((k_bbpa_hist<double, false, KV, NHV>), grid, dim3(threads), lds, stream, descs); \
		} else {                                                                                                \
			if (fused)                                                                                          \
				hipLaunchKernelGGL((k_bbpa_hist<float, true, KV, NHV>), grid, dim3(threads), lds, stream, descs); \
			else                                                                                                \
				hipLaunchKernelGGL((k_bbpa_hist<float, false, KV, NHV>), grid, dim3(threads), lds, stream, descs); \
		}                                                                                                       \
		return;                                                                                                 \
	}
		BGT_HIST(2, 0) BGT_HIST(2, 1) BGT_HIST(2, 2) BGT_HIST(2, 3) BGT_HIST(2, 4)
		BGT_HIST(3, 0) BGT_HIST(3, 1) BGT_HIST(3, 2) BGT_HIST(3, 3)
		BGT_HIST(4, 0) BGT_HIST(4, 1) BGT_HIST(4, 2)
		BGT_HIST(5, 0) BGT_HIST(5, 1)
#undef BGT_HIST
		throw Error(4, "no one-pass table kernel for this shape");
	});
	if (!tables)
		return; // launch_bbpa_finish converts them
	launch_checked("k_bbpa_tables", [&] {
		hipLaunchKernelGGL(k_bbpa_tables, dim3(1, count), dim3(1024), 0, stream, descs, fused ? 1 : 0);
	});
}

void launch_brute_force(const Descs &descs, int count, int n, int k, Storage st, hipStream_t stream)
{
	const dim3 grid(1u << n, count);
	launch_checked("k_brute_force", [&] {
#define BGT_BF(KK)                                                                                  \
	if (k == KK) {                                                                                  \
		if (st == Storage::fp64)                                                                    \
			hipLaunchKernelGGL((k_brute_force<double, KK>), grid, dim3(256), 0, stream, descs, n); \
		else                                                                                        \
			hipLaunchKernelGGL((k_brute_force<float, KK>), grid, dim3(256), 0, stream, descs, n);  \
		hipLaunchKernelGGL(k_finalize_top2, dim3(1, count), dim3(1024), 0, stream, descs, grid.x, 0); \
		return;                                                                                     \
	}
		BGT_BF(1)
		BGT_BF(2)
		BGT_BF(3)
		BGT_BF(4)
#undef BGT_BF
		throw Error(2, "brute force supports at most 4 variants");
	});
}

void launch_atom_masses(const Descs &descs, int count, int atoms, Storage st, hipStream_t stream)
{
	const int cb = atoms < 14 ? atoms : 14;
	const dim3 grid(1u << (atoms - cb), count);
	launch_checked("k_atom_masses", [&] {
		if (st == Storage::fp64)
			hipLaunchKernelGGL(k_atom_masses<double>, grid, dim3(256), 0, stream, descs, atoms);
		else
			hipLaunchKernelGGL(k_atom_masses<float>, grid, dim3(256), 0, stream, descs, atoms);
		hipLaunchKernelGGL(k_finalize_masses, dim3(1, count), dim3(1024), 0, stream, descs, grid.x, atoms);
	});
}

void launch_upset_mass(const Descs &descs, int count, int atoms, uint32_t e, Storage st, hipStream_t stream)
{
	const int freebits = atoms - __builtin_popcount(e & ((atoms >= 32) ? 0xFFFFFFFFu : ((1u << atoms) - 1u)));
	const dim3 grid(grid_for(1ull << freebits, 256 * 8, 2048), count);
	launch_checked("k_upset_mass", [&] {
		if (st == Storage::fp64)
			hipLaunchKernelGGL(k_upset_mass<double>, grid, dim3(256), 0, stream, descs, atoms, e);
		else
			hipLaunchKernelGGL(k_upset_mass<float>, grid, dim3(256), 0, stream, descs, atoms, e);
		hipLaunchKernelGGL(k_finalize_sum, dim3(1, count), dim3(1024), 0, stream, descs, grid.x, (int)kSumUpset, 0);
	});
}

void launch_project(const void *src, const double *src_scal, int inv_slot, void *dst, int dst_atoms, uint32_t removed,
					Storage st, hipStream_t stream, double *partial, size_t partial_cap)
{
	const uint64_t states = 1ull << dst_atoms;
	const int r = __builtin_popcount(removed);
	if (r > kProjectSerialBits) {
		// about 2048 workgroups: kept entries times chunks of at least 256 terms each, the
		// chunk sums within the partial buffer
		uint64_t chunks = 1;
		while (states * chunks * 2 <= 2048 && (1ull << r) / (chunks * 2) >= 256 && states * chunks * 2 <= partial_cap)
			chunks *= 2;
		if (chunks > 1 && partial == nullptr)
			throw Error(4, "wide projection without a partial buffer");
		const dim3 grid((unsigned)std::min<uint64_t>(states * chunks, 8192));
		const dim3 join(grid_for(states, 256, 1024));
		launch_checked("k_project_wide", [&] {
			if (st == Storage::fp64) {
				hipLaunchKernelGGL((k_project_wide<double, double>), grid, dim3(256), 0, stream,
								   static_cast<const double *>(src), src_scal, inv_slot, static_cast<double *>(dst),
								   partial, states, removed, r, (uint32_t)chunks);
				if (chunks > 1)
					hipLaunchKernelGGL(k_project_join<double>, join, dim3(256), 0, stream, partial,
									   static_cast<double *>(dst), states, (uint32_t)chunks);
			} else {
				hipLaunchKernelGGL((k_project_wide<float, float>), grid, dim3(256), 0, stream,
								   static_cast<const float *>(src), src_scal, inv_slot, static_cast<float *>(dst),
								   partial, states, removed, r, (uint32_t)chunks);
				if (chunks > 1)
					hipLaunchKernelGGL(k_project_join<float>, join, dim3(256), 0, stream, partial,
									   static_cast<float *>(dst), states, (uint32_t)chunks);
			}
		});
		return;
	}
	const dim3 grid(grid_for(states, 256, 8192));
	launch_checked("k_project", [&] {
		if (st == Storage::fp64)
			hipLaunchKernelGGL((k_project<double, double>), grid, dim3(256), 0, stream, static_cast<const double *>(src),
							   src_scal, inv_slot, static_cast<double *>(dst), states, removed);
		else
			hipLaunchKernelGGL((k_project<float, float>), grid, dim3(256), 0, stream, static_cast<const float *>(src),
							   src_scal, inv_slot, static_cast<float *>(dst), states, removed);
	});
}

void launch_read_posterior(const Descs &descs, int atoms, Storage st, double *out, hipStream_t stream)
{
	const uint64_t states = 1ull << atoms;
	const dim3 grid(grid_for(states, 256 * 4, 8192));
	launch_checked("k_read_posterior", [&] {
		if (st == Storage::fp64)
			hipLaunchKernelGGL(k_read_posterior<double>, grid, dim3(256), 0, stream, descs, states, out);
		else
			hipLaunchKernelGGL(k_read_posterior<float>, grid, dim3(256), 0, stream, descs, states, out);
	});
}

void launch_lookahead(const Descs &descs, int n, Storage st, const uint32_t *selected, int m, double *scores,
					  hipStream_t stream)
{
	const dim3 grid(1u << n);
	launch_checked("k_lookahead", [&] {
		if (st == Storage::fp64)
			hipLaunchKernelGGL(k_lookahead<double>, grid, dim3(256), 0, stream, descs, n, selected, m, scores);
		else
			hipLaunchKernelGGL(k_lookahead<float>, grid, dim3(256), 0, stream, descs, n, selected, m, scores);
	});
}

namespace {

// the bits of the down-set sums above the first 12, for `rows` transforms of 2^atoms entries
// side by side: ceil(rest / 6) register passes of near-equal width; each pass runs
// 2^(atoms-R) threads per transform, 256 per workgroup (atoms - R >= 12)
void launch_down_rest(double *D, int atoms, unsigned rows, hipStream_t stream)
{
	const int rest = atoms - kDownFirstBits;
	const int passes = (rest + 5) / 6;
	const uint64_t stride = 1ull << atoms;
	for (int p = 0, b0 = kDownFirstBits; p < passes; ++p) {
		const int w = (rest - (b0 - kDownFirstBits) + (passes - p) - 1) / (passes - p);
		const dim3 grid((unsigned)(1ull << (atoms - w - 8)), rows);
		launch_checked("k_down_more", [&] {
			switch (w) {
			case 1: hipLaunchKernelGGL(k_down_more<1>, grid, dim3(256), 0, stream, D, b0, stride); break;
			case 2: hipLaunchKernelGGL(k_down_more<2>, grid, dim3(256), 0, stream, D, b0, stride); break;
			case 3: hipLaunchKernelGGL(k_down_more<3>, grid, dim3(256), 0, stream, D, b0, stride); break;
			case 4: hipLaunchKernelGGL(k_down_more<4>, grid, dim3(256), 0, stream, D, b0, stride); break;
			case 5: hipLaunchKernelGGL(k_down_more<5>, grid, dim3(256), 0, stream, D, b0, stride); break;
			case 6: hipLaunchKernelGGL(k_down_more<6>, grid, dim3(256), 0, stream, D, b0, stride); break;
			default: throw Error(4, "down-set pass width out of range");
			}
		});
		b0 += w;
	}
}

} // namespace

void launch_down_sums(const Descs &descs, int atoms, Storage st, double *D, hipStream_t stream)
{
	// pass A: 2^(atoms-12) waves of 4096 entries, up to four waves per workgroup
	if (atoms < kDownFirstBits || atoms > 32)
		throw Error(4, "down-set sums need 12..32 atoms");
	launch_checked("k_down_first", [&] {
		const uint64_t threads = 1ull << (atoms - 6);
		const unsigned block = (unsigned)std::min<uint64_t>(256, threads);
		const dim3 grid((unsigned)(threads / block));
		if (st == Storage::fp64)
			hipLaunchKernelGGL(k_down_first<double>, grid, dim3(block), 0, stream, descs, D);
		else
			hipLaunchKernelGGL(k_down_first<float>, grid, dim3(block), 0, stream, descs, D);
	});
	launch_down_rest(D, atoms, 1, stream);
}

size_t look_class_rows(int n)
{
	// classes per chunk: at most 64, and at most 256 MiB of rows
	const uint64_t row = sizeof(double) << n;
	return (size_t)std::max<uint64_t>(1, std::min<uint64_t>(64, (256ull << 20) / row));
}

void launch_look_classes(const Descs &descs, Storage st, const LookClasses &lc, double *P, double *scores,
						 double *best_score, unsigned blocks, hipStream_t stream)
{
	const int n = lc.n;
	if (lc.m < 1 || lc.m >= kLookMaxSteps || n < 1 || n > 31)
		throw Error(4, "lookahead classes out of range");
	if (n > kLookLdsMax && n < kDownFirstBits)
		throw Error(4, "lookahead classes: no transform for this subject count");
	const uint32_t classes = 1u << lc.m;
	const uint32_t chunk = (uint32_t)look_class_rows(n);
	const double target = 1.0 / (double)(2u * classes);
	for (uint32_t j0 = 0; j0 < classes; j0 += chunk) {
		const uint32_t rows = std::min(chunk, classes - j0);
		if (n <= kLookLdsMax) {
			const size_t lds = sizeof(double) << n;
			launch_checked("k_look_class_lds", [&] {
				if (st == Storage::fp64)
					hipLaunchKernelGGL(k_look_class_lds<double>, dim3(rows), dim3(256), lds, stream, descs, lc, j0, P);
				else
					hipLaunchKernelGGL(k_look_class_lds<float>, dim3(rows), dim3(256), lds, stream, descs, lc, j0, P);
			});
		} else {
			launch_checked("k_look_class_first", [&] {
				const uint64_t threads = 1ull << (n - 6);
				const unsigned block = (unsigned)std::min<uint64_t>(256, threads);
				const dim3 grid((unsigned)(threads / block), rows);
				if (st == Storage::fp64)
					hipLaunchKernelGGL(k_look_class_first<double>, grid, dim3(block), 0, stream, descs, lc, j0, P);
				else
					hipLaunchKernelGGL(k_look_class_first<float>, grid, dim3(block), 0, stream, descs, lc, j0, P);
			});
			launch_down_rest(P, n, rows, stream);
		}
		const int first = j0 == 0 ? 1 : 0, last = j0 + rows == classes ? 1 : 0;
		launch_checked("k_look_class_accum", [&] {
			hipLaunchKernelGGL(k_look_class_accum, dim3(blocks), dim3(256), 0, stream, P, (int)rows, n, target, first,
							   last, scores, best_score);
		});
	}
}

void launch_lookahead_down(const double *D, const LookStep &step, double *scores, double *best_score,
						   unsigned blocks, hipStream_t stream)
{
	if (step.m < 1 || step.m > 4)
		throw Error(4, "lookahead step out of range");
	launch_checked("k_lookahead_down", [&] {
		switch (step.m) {
		case 1: hipLaunchKernelGGL(k_lookahead_down<1>, dim3(blocks), dim3(256), 0, stream, D, step, scores, best_score); break;
		case 2: hipLaunchKernelGGL(k_lookahead_down<2>, dim3(blocks), dim3(256), 0, stream, D, step, scores, best_score); break;
		case 3: hipLaunchKernelGGL(k_lookahead_down<3>, dim3(blocks), dim3(256), 0, stream, D, step, scores, best_score); break;
		default: hipLaunchKernelGGL(k_lookahead_down<4>, dim3(blocks), dim3(256), 0, stream, D, step, scores, best_score); break;
		}
	});
}

void launch_lookahead_pick(const double *scores, const double *block_best, int n, double limit,
						   uint32_t *best_exp, unsigned blocks, hipStream_t stream)
{
	launch_checked("k_lookahead_pick", [&] {
		hipLaunchKernelGGL(k_lookahead_pick, dim3(blocks), dim3(256), 0, stream, scores, block_best, n, limit,
						   best_exp);
	});
}

void launch_tree_stats(const FlatNode *nodes, int node_count, int n, int k, int dilution, const double *dil,
					   const double *pi, int depth, double branch, double *partials, unsigned *counter, double *out,
					   hipStream_t stream)
{
	const uint64_t states = 1ull << (n * k);
	const dim3 grid(grid_for(states, 256, 4096));
	launch_checked("k_tree_stats", [&] {
		hipLaunchKernelGGL(k_tree_stats, grid, dim3(256), 0, stream, nodes, node_count, n, k, dilution, dil, pi,
						   depth, branch, partials, counter, out);
		hipLaunchKernelGGL(k_finalize_tree, dim3(1), dim3(1024), 0, stream, partials, grid.x, out);
	});
}

} // namespace bgt_hip
