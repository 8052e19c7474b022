#include "bgt_hip_exact.hpp"

#include <algorithm>
#include <cmath>

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
} // namespace

// Open-addressing map from pool-set state to weight; entries keep insertion order, so the
// final summation order is deterministic.
void ExactHistory::Table::clear(size_t expect)
{
	key.clear();
	w.clear();
	size_t cap = 16;
	while (cap < 2 * expect)
		cap <<= 1;
	if (slot.size() != cap)
		slot.assign(cap, -1);
	else
		std::fill(slot.begin(), slot.end(), -1);
}

void ExactHistory::Table::add(uint64_t k, double v)
{
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

void ExactHistory::reset(bool usable, int atoms, const double *prior)
{
	usable_ = usable && atoms <= 32;
	broken_ = false;
	atoms_ = atoms;
	pi_.assign(prior, prior + atoms);
	q_.resize(atoms);
	for (int a = 0; a < atoms; ++a)
		q_[a] = 1.0 - pi_[a];
	hist_.clear();
	z_ = 1.0;
	z_ok_ = true;
}

void ExactHistory::push(uint32_t e, uint32_t r)
{
	if (!usable_)
		return;
	// the factors of the device likelihood tables (same FP64 expressions)
	const double contained = r ? kNegative : (1.0 - kNegative);
	const double missing = r ? (1.0 - kNegative) : kNegative;
	hist_.push_back({e, contained, missing});
	z_ok_ = false;
}

double ExactHistory::rel_bound() const
{
	// products of up to `atoms` prior factors and `updates` likelihood factors and sums of
	// the program's non-negative terms, for U(E) and U({}): a generous multiple of the
	// unit roundoff
	return 8.0 * (atoms_ + 2.0 * hist_.size() + 8.0) * std::ldexp(1.0, -53);
}

bool ExactHistory::unnormalised(uint32_t e, double &out, Scratch &sc) const
{
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
	if (c == 0) {
		out = qe * always;
		return true;
	}
	if (c > 64)
		return false;
	uint32_t relevant = 0;
	for (uint32_t p : sc.pool)
		relevant |= p;
	Table *cur = &sc.a, *nxt = &sc.b;
	cur->clear(1);
	cur->add(c == 64 ? ~0ull : ((1ull << c) - 1ull), 1.0);
	for (int a = 0; a < atoms_; ++a) {
		if (!((relevant >> a) & 1u))
			continue;
		uint64_t kill = 0;
		for (size_t i = 0; i < c; ++i)
			if ((sc.pool[i] >> a) & 1u)
				kill |= 1ull << i;
		nxt->clear(2 * cur->key.size());
		for (size_t s = 0; s < cur->key.size(); ++s) {
			const uint64_t k = cur->key[s];
			const double wv = cur->w[s];
			nxt->add(k, wv * q_[a]);          // atom negative: pools unchanged
			nxt->add(k & ~kill, wv * pi_[a]); // atom positive: its pools are out
		}
		if (nxt->key.size() > kMaxStates)
			return false;
		std::swap(cur, nxt);
	}
	double sum = 0.0;
	for (size_t s = 0; s < cur->key.size(); ++s) {
		const uint64_t k = cur->key[s];
		double f = 1.0;
		for (size_t i = 0; i < c; ++i)
			f *= ((k >> i) & 1u) ? sc.fin[i] : sc.fout[i];
		sum += cur->w[s] * f;
	}
	out = qe * always * sum;
	return true;
}

bool ExactHistory::normaliser()
{
	if (z_ok_)
		return true;
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
	for (size_t i = 0; i < count; ++i) {
		double u = 0.0;
		ok[i] = unnormalised(e[i], u, scratch_);
		out[i] = ok[i] ? u / z_ : NAN;
	}
	return true;
}

} // namespace bgt_hip
