#include "bgt_hip_exact.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>

namespace bgt_hip {

namespace {
constexpr double kNegative = 0.99;     // replicated_lattice.cpp:6
constexpr size_t kMaxStates = 1u << 18; // program budget (pool-set states per step)

inline uint32_t mix(uint64_t k)
{
	k ^= k >> 33;
	k *= 0xff51afd7ed558ccdull;
	k ^= k >> 33;
	return (uint32_t)k;
}
// A fixed pool of host threads for exact rescoring: run(n, f) calls f(i, worker) for
// i in [0, n) on the pool and the calling thread (worker 0) and returns when all are done.
// Calls from different host threads are serialised.
class Pool {
public:
	static Pool &get()
	{
		static Pool pool;
		return pool;
	}
	int size() const { return (int)threads_.size() + 1; }
	void run(size_t n, const std::function<void(size_t, int)> &f)
	{
		std::lock_guard<std::mutex> call(call_);
		{
			std::lock_guard<std::mutex> lk(m_);
			job_ = &f;
			n_ = n;
			next_.store(0);
			busy_ = (int)threads_.size();
			++gen_;
		}
		cv_.notify_all();
		work(0);
		std::unique_lock<std::mutex> lk(m_);
		done_.wait(lk, [&] { return busy_ == 0; });
		job_ = nullptr;
	}
	~Pool()
	{
		{
			std::lock_guard<std::mutex> lk(m_);
			stop_ = true;
			++gen_;
		}
		cv_.notify_all();
		for (auto &t : threads_)
			t.join();
	}

private:
	Pool()
	{
		const unsigned hw = std::thread::hardware_concurrency();
		const int extra = (int)std::min(15u, hw > 1 ? hw - 1 : 0u);
		for (int w = 1; w <= extra; ++w)
			threads_.emplace_back([this, w] { loop(w); });
	}
	void work(int w)
	{
		// blocks of 8 items: fewer atomics, and neighbouring outputs stay on one thread
		constexpr size_t kBlock = 8;
		for (size_t b = next_.fetch_add(kBlock); b < n_; b = next_.fetch_add(kBlock))
			for (size_t i = b; i < std::min(n_, b + kBlock); ++i)
				(*job_)(i, w);
	}
	void loop(int w)
	{
		uint64_t seen = 0;
		while (true) {
			{
				std::unique_lock<std::mutex> lk(m_);
				cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
				if (stop_)
					return;
				seen = gen_;
			}
			work(w);
			std::lock_guard<std::mutex> lk(m_);
			if (--busy_ == 0)
				done_.notify_one();
		}
	}
	std::vector<std::thread> threads_;
	std::mutex call_, m_;
	std::condition_variable cv_, done_;
	const std::function<void(size_t, int)> *job_ = nullptr;
	size_t n_ = 0;
	std::atomic<size_t> next_{0};
	int busy_ = 0;
	uint64_t gen_ = 0;
	bool stop_ = false;
};

} // namespace

void parallel_for(size_t n, const std::function<void(size_t, int)> &f) { Pool::get().run(n, f); }
int parallel_width() { return Pool::get().size(); }

void ExactHistory::Table::clear(size_t expect, int pools)
{
	if (pools <= kDenseBits) {
		const size_t cap = size_t(1) << pools;
		if (!dense || slot.size() < cap) {
			slot.assign(std::max(cap, slot.size()), -1);
		} else {
			for (uint64_t k : key)
				slot[k] = -1;
		}
		dense = true;
		key.clear();
		w.clear();
		return;
	}
	key.clear();
	w.clear();
	size_t cap = 16;
	while (cap < 2 * expect)
		cap <<= 1;
	if (slot.size() != cap || dense)
		slot.assign(cap, -1);
	else
		std::fill(slot.begin(), slot.end(), -1);
	dense = false;
}

void ExactHistory::Table::add(uint64_t k, double v)
{
	if (dense) {
		int32_t &s = slot[k];
		if (s < 0) {
			s = (int32_t)key.size();
			key.push_back(k);
			w.push_back(v);
		} else {
			w[s] += v;
		}
		return;
	}
	const size_t m = slot.size() - 1;
	size_t h = mix(k) & m;
	while (true) {
		const int32_t s = slot[h];
		if (s < 0) {
			slot[h] = (int32_t)key.size();
			key.push_back(k);
			w.push_back(v);
			return;
		}
		if (key[s] == k) {
			w[s] += v;
			return;
		}
		h = (h + 1) & m;
	}
}

void ExactHistory::reset(bool usable, int atoms, const double *prior, bool dilution, int variants)
{
	usable_ = usable && atoms <= 32 && variants >= 1 && atoms % variants == 0;
	broken_ = false;
	dilution_ = dilution;
	atoms_ = atoms;
	variants_ = variants;
	subjects_ = variants >= 1 ? atoms / variants : atoms;
	pi_.assign(prior, prior + atoms);
	q_.resize(atoms);
	for (int a = 0; a < atoms; ++a)
		q_[a] = 1.0 - pi_[a];
	hist_.clear();
	z_ = 1.0;
	z_ok_ = true;
}

void ExactHistory::push(uint32_t e, uint32_t r, const double *dil)
{
	if (!usable_)
		return;
	if (variants_ > 1) {
		// per variant the factors of the device likelihood tables (replicated_lattice.cpp:8-28,
		// same FP64 expressions): dilution P(r_v = 1) = D[s-1][m] (the empty pool 0),
		// non-dilution P(r_v = 1) = 0.99 when no pooled atom is positive, 0.01 otherwise
		if (dilution_ && dil == nullptr) {
			broken_ = true;
			return;
		}
		const int len = __builtin_popcount(e);
		const int stride = subjects_ + 1;
		Upd u{e, 0.0, 0.0, {}, {}};
		u.fv.assign((size_t)variants_ * stride, 0.0);
		for (int v = 0; v < variants_; ++v) {
			const bool rv = (r >> v) & 1u;
			for (int m = 0; m <= len; ++m) {
				double f;
				if (dilution_) {
					const double p1 = len > 0 ? dil[(size_t)(len - 1) * stride + m] : 0.0;
					f = rv ? p1 : 1.0 - p1;
				} else if (m == 0) {
					f = rv ? kNegative : (1.0 - kNegative);
				} else {
					f = rv ? (1.0 - kNegative) : kNegative;
				}
				u.fv[(size_t)v * stride + m] = f;
			}
		}
		hist_.push_back(std::move(u));
		z_ok_ = false;
		return;
	}
	if (dilution_) {
		if (dil == nullptr) {
			broken_ = true;
			return;
		}
		// the factors of the device likelihood tables (same FP64 expressions): pool size s,
		// m positive atoms, P(r = 1) = D[s-1][m]; the empty pool has P(r = 1) = 0
		const int len = __builtin_popcount(e);
		Upd u{e, 0.0, 0.0, {}, {}};
		u.f.resize(len + 1);
		for (int m = 0; m <= len; ++m) {
			const double p1 = len > 0 ? dil[(size_t)(len - 1) * (atoms_ + 1) + m] : 0.0;
			u.f[m] = r ? p1 : 1.0 - p1;
		}
		u.f_in = u.f[0];
		hist_.push_back(std::move(u));
		z_ok_ = false;
		return;
	}
	// the factors of the device likelihood tables (same FP64 expressions)
	const double contained = r ? kNegative : (1.0 - kNegative);
	const double missing = r ? (1.0 - kNegative) : kNegative;
	hist_.push_back({e, contained, missing, {}, {}});
	z_ok_ = false;
}

void ExactHistory::enum_inputs(int variant, double *pairs, uint32_t *pools, double *fac, int stride) const
{
	for (int i = 0; i < subjects_; ++i) {
		pairs[2 * i] = pi_[variant * subjects_ + i];
		pairs[2 * i + 1] = q_[variant * subjects_ + i];
	}
	if (variants_ > 1) {
		for (size_t t = 0; t < hist_.size(); ++t) {
			pools[t] = hist_[t].e;
			const int len = __builtin_popcount(hist_[t].e);
			for (int m = 0; m < stride; ++m)
				fac[t * stride + m] = m <= len ? hist_[t].fv[(size_t)variant * (subjects_ + 1) + m] : 0.0;
		}
		return;
	}
	for (size_t t = 0; t < hist_.size(); ++t) {
		const Upd &u = hist_[t];
		pools[t] = u.e;
		const int len = __builtin_popcount(u.e);
		for (int m = 0; m < stride; ++m) {
			double v = 0.0;
			if (m <= len)
				v = dilution_ ? u.f[m] : (m == 0 ? u.f_in : u.f_out);
			fac[t * (size_t)stride + m] = v;
		}
	}
}

double ExactHistory::rel_bound() const
{
	// products of up to `atoms` prior factors and `updates` likelihood factors and sums of
	// the program's non-negative terms, for U(E) and U({}): a generous multiple of the
	// unit roundoff
	return 8.0 * (atoms_ + 2.0 * hist_.size() + 8.0) * std::ldexp(1.0, -53);
}

bool ExactHistory::counts(uint32_t e, double &out, Scratch &sc) const
{
	double lead = 1.0;
	for (int a = 0; a < atoms_; ++a)
		if ((e >> a) & 1u)
			lead *= q_[a];
	sc.pool.clear();
	sc.g.resize(0);
	for (const Upd &u : hist_) {
		const uint32_t rest = u.e & ~e;
		if (rest == 0) {
			lead *= u.f[0]; // every pooled atom is negative: no positive atom in the pool
			continue;
		}
		// the candidate's atoms are negative, so the pool's count is over its remaining atoms
		size_t i = 0;
		while (i < sc.pool.size() && sc.pool[i] != rest)
			++i;
		const int len = __builtin_popcount(rest);
		if (i == sc.pool.size()) {
			sc.pool.push_back(rest);
			sc.g.emplace_back(u.f.begin(), u.f.begin() + len + 1);
		} else {
			for (int m = 0; m <= len; ++m)
				sc.g[i][m] *= u.f[m];
		}
	}
	const size_t c = sc.pool.size();
	uint32_t relevant = 0;
	for (uint32_t p : sc.pool)
		relevant |= p;
	// Fields of the 64-bit state key: each partly processed pool holds its count of positive
	// atoms so far in a field just wide enough for its size, taken (lowest free bits first)
	// at the pool's first atom and given back after its last.
	std::vector<int> first(c, -1), last(c, -1), shift(c, -1), width(c, 0);
	std::vector<int> pos;
	for (int a : order_)
		if ((relevant >> a) & 1u)
			pos.push_back(a);
	for (size_t t = 0; t < pos.size(); ++t)
		for (size_t i = 0; i < c; ++i)
			if ((sc.pool[i] >> pos[t]) & 1u) {
				if (first[i] < 0)
					first[i] = (int)t;
				last[i] = (int)t;
			}
	for (size_t i = 0; i < c; ++i)
		width[i] = 32 - __builtin_clz((uint32_t)__builtin_popcount(sc.pool[i]));
	Table *cur = &sc.a, *nxt = &sc.b;
	cur->clear(1, 64);
	cur->add(0ull, 1.0);
	uint64_t used = 0; // key bits in use
	for (size_t t = 0; t < pos.size(); ++t) {
		const int a = pos[t];
		uint64_t inc = 0; // key increment of a positive atom
		std::vector<int> closing;
		for (size_t i = 0; i < c; ++i) {
			if (!((sc.pool[i] >> a) & 1u))
				continue;
			if (first[i] == (int)t) {
				const uint64_t field = (1ull << width[i]) - 1ull;
				int at = 0;
				while (at + width[i] <= 64 && ((used >> at) & field) != 0ull)
					++at;
				if (at + width[i] > 64)
					return false; // more partly processed pools than the key holds
				shift[i] = at;
				used |= field << at;
			}
			inc += 1ull << shift[i];
			if (last[i] == (int)t)
				closing.push_back((int)i);
		}
		nxt->clear(2 * cur->key.size(), 64);
		for (size_t s = 0; s < cur->key.size(); ++s) {
			const uint64_t k = cur->key[s];
			const double wv = cur->w[s];
			uint64_t k0 = k, k1 = k + inc;
			double w0 = wv * q_[a], w1 = wv * pi_[a];
			for (int i : closing) {
				const uint64_t fm = (1ull << width[i]) - 1ull;
				w0 *= sc.g[i][(k0 >> shift[i]) & fm];
				w1 *= sc.g[i][(k1 >> shift[i]) & fm];
				k0 &= ~(fm << shift[i]);
				k1 &= ~(fm << shift[i]);
			}
			nxt->add(k0, w0);
			nxt->add(k1, w1);
		}
		for (int i : closing)
			used &= ~(((1ull << width[i]) - 1ull) << shift[i]);
		if (nxt->key.size() > kMaxStates)
			return false;
		std::swap(cur, nxt);
	}
	double sum = 0.0;
	for (size_t s = 0; s < cur->key.size(); ++s)
		sum += cur->w[s];
	out = lead * sum;
	return true;
}

bool ExactHistory::unnormalised(uint32_t e, double &out, Scratch &sc) const
{
	if (dilution_)
		return counts(e, out, sc);
	double qe = 1.0;
	for (int a = 0; a < atoms_; ++a)
		if ((e >> a) & 1u)
			qe *= q_[a];
	double always = 1.0;
	sc.pool.clear();
	sc.fin.clear();
	sc.fout.clear();
	for (const Upd &u : hist_) {
		const uint32_t rest = u.e & ~e;
		if (rest == 0) {
			always *= u.f_in;
			continue;
		}
		// pools with the same remaining atoms are contained together: merge their factors
		size_t i = 0;
		while (i < sc.pool.size() && sc.pool[i] != rest)
			++i;
		if (i == sc.pool.size()) {
			sc.pool.push_back(rest);
			sc.fin.push_back(u.f_in);
			sc.fout.push_back(u.f_out);
		} else {
			sc.fin[i] *= u.f_in;
			sc.fout[i] *= u.f_out;
		}
	}
	const size_t c = sc.pool.size();
	if (c <= (size_t)kDenseMaxPools && hist_.size() <= (size_t)kDenseMaxHist) {
		out = dense(e, sc);
		return true;
	}
	if (c > 64)
		return false;
	uint32_t relevant = 0;
	for (uint32_t p : sc.pool)
		relevant |= p;
	// State: bit i set while pool i can still be contained. Atoms are processed in order_;
	// once a pool's last atom has been processed its factor is applied and its bit cleared,
	// so only the pools that are partly processed tell states apart.
	Table *cur = &sc.a, *nxt = &sc.b;
	cur->clear(1, (int)c);
	cur->add(c == 64 ? ~0ull : ((1ull << c) - 1ull), 1.0);
	uint32_t done = 0;
	for (int a : order_) {
		if (!((relevant >> a) & 1u))
			continue;
		done |= 1u << a;
		uint64_t kill = 0, close = 0;
		for (size_t i = 0; i < c; ++i) {
			if ((sc.pool[i] >> a) & 1u) {
				kill |= 1ull << i;
				if ((sc.pool[i] & ~done) == 0)
					close |= 1ull << i;
			}
		}
		nxt->clear(2 * cur->key.size(), (int)c);
		for (size_t s = 0; s < cur->key.size(); ++s) {
			const uint64_t k = cur->key[s];
			const double wv = cur->w[s];
			// atom negative: pools unchanged; atom positive: its pools are out
			uint64_t k0 = k, k1 = k & ~kill;
			double w0 = wv * q_[a], w1 = wv * pi_[a];
			for (uint64_t rest = close; rest; rest &= rest - 1) {
				const int i = __builtin_ctzll(rest);
				w0 *= ((k0 >> i) & 1u) ? sc.fin[i] : sc.fout[i];
				w1 *= ((k1 >> i) & 1u) ? sc.fin[i] : sc.fout[i];
			}
			nxt->add(k0 & ~close, w0);
			nxt->add(k1 & ~close, w1);
		}
		if (nxt->key.size() > kMaxStates)
			return false;
		std::swap(cur, nxt);
	}
	// every pool has closed: one state remains
	double sum = 0.0;
	for (size_t s = 0; s < cur->key.size(); ++s)
		sum += cur->w[s];
	out = qe * always * sum;
	return true;
}

double ExactHistory::dense(uint32_t e, Scratch &sc) const
{
	// Q(E) and the always-contained factors, multiplied in atom and then update order
	double lead = 1.0;
	for (int a = 0; a < atoms_; ++a)
		if ((e >> a) & 1u)
			lead *= q_[a];
	for (const Upd &u : hist_)
		if ((u.e & ~e) == 0u)
			lead *= u.f_in;
	const int c = (int)sc.pool.size();
	const uint32_t states = 1u << c, half = states >> 1;
	std::vector<double> &A = sc.da, &B = sc.db;
	A.assign(states, 0.0);
	A[states - 1] = 1.0;
	B.resize(states);
	uint32_t relevant = 0;
	for (uint32_t p : sc.pool)
		relevant |= p;
	// atoms in increasing order; state bit j set while pool j can still be contained
	uint32_t done = 0;
	for (uint32_t left = relevant; left; left &= left - 1u) {
		const int at = __builtin_ctz(left);
		done |= 1u << at;
		uint32_t kill = 0, close = 0;
		for (int j = 0; j < c; ++j)
			if ((sc.pool[j] >> at) & 1u) {
				kill |= 1u << j;
				if ((sc.pool[j] & ~done) == 0u)
					close |= 1u << j;
			}
		const double p = pi_[at], q = q_[at];
		// atom positive: its pools are out (the kill bits projected away, one per sweep)
		std::copy(A.begin(), A.end(), B.begin());
		for (uint32_t bits = kill; bits; bits &= bits - 1u) {
			const int j = __builtin_ctz(bits);
			const uint32_t low = (1u << j) - 1u;
			for (uint32_t x = 0; x < half; ++x) {
				const uint32_t k = ((x & ~low) << 1) | (x & low);
				B[k] += B[k | (1u << j)];
				B[k | (1u << j)] = 0.0;
			}
		}
		for (uint32_t k = 0; k < states; ++k)
			A[k] = A[k] * q + B[k] * p;
		// pools whose last atom this was take their factor and leave the state
		for (uint32_t bits = close; bits; bits &= bits - 1u) {
			const int j = __builtin_ctz(bits);
			const uint32_t low = (1u << j) - 1u;
			const double fi = sc.fin[j], fo = sc.fout[j];
			for (uint32_t x = 0; x < half; ++x) {
				const uint32_t k = ((x & ~low) << 1) | (x & low);
				A[k] = A[k] * fo + A[k | (1u << j)] * fi;
				A[k | (1u << j)] = 0.0;
			}
		}
	}
	return lead * A[0];
}

void ExactHistory::plan_order()
{
	// Atom order of the program: greedily the atom after which the fewest pools are partly
	// processed (lowest index on ties), for the pools of the whole history; irrelevant
	// atoms last. Fewer partly processed pools mean fewer distinct states.
	std::vector<uint32_t> pools;
	uint32_t relevant = 0;
	for (const Upd &u : hist_) {
		if (u.e == 0 || std::find(pools.begin(), pools.end(), u.e) != pools.end())
			continue;
		pools.push_back(u.e);
		relevant |= u.e;
	}
	order_.clear();
	uint32_t done = 0;
	while (relevant & ~done) {
		int best = -1, best_open = 1 << 30;
		for (int a = 0; a < atoms_; ++a) {
			if (!((relevant & ~done) >> a & 1u))
				continue;
			const uint32_t after = done | (1u << a);
			int open = 0;
			for (uint32_t p : pools)
				open += (p & after) != 0 && (p & ~after) != 0;
			if (open < best_open) {
				best_open = open;
				best = a;
			}
		}
		order_.push_back(best);
		done |= 1u << best;
	}
	for (int a = 0; a < atoms_; ++a)
		if (!((done >> a) & 1u))
			order_.push_back(a);
}

bool ExactHistory::normaliser()
{
	if (variants_ != 1)
		return false; // the pool program is binary
	if (z_ok_)
		return true;
	plan_order();
	if (!unnormalised(0u, z_, scratch_)) {
		broken_ = true;
		return false;
	}
	z_ok_ = true;
	return true;
}

bool ExactHistory::mass(uint32_t e, double &out)
{
	if (!usable() || !normaliser())
		return false;
	double u = 0.0;
	if (!unnormalised(e, u, scratch_))
		return false;
	out = u / z_;
	return true;
}

bool ExactHistory::masses(const uint32_t *e, size_t count, double *out, bool *ok)
{
	if (!usable() || !normaliser())
		return false;
	auto one = [&](size_t i, Scratch &sc) {
		double u = 0.0;
		ok[i] = unnormalised(e[i], u, sc);
		out[i] = ok[i] ? u / z_ : NAN;
	};
	constexpr size_t kParallelFrom = 16;
	if (count < kParallelFrom) {
		for (size_t i = 0; i < count; ++i)
			one(i, scratch_);
		return true;
	}
	Pool &pool = Pool::get();
	if (workers_.size() < (size_t)pool.size())
		workers_.resize(pool.size());
	pool.run(count, [&](size_t i, int w) { one(i, workers_[w]); });
	return true;
}

} // namespace bgt_hip
