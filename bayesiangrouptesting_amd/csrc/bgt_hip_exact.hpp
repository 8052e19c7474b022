// Exact upset masses of a binary lattice, recomputed in FP64 from the prior and the applied
// updates rather than from the stored posterior. Used to re-decide selections
// whose best candidates are closer than the stored form's error (FP32C storage carries
// ~1e-7 relative error per entry while top-two gaps at 2^28 states are ~1e-9).
//
// With prior P(S) = prod_a (bit a of S ? q_a : pi_a), q_a = 1 - pi_a, and updates
// (E_t, r_t) whose likelihood depends only on c_t(S) = [S contains E_t]
// (replicated_lattice.cpp:8-28, non-dilution), the unnormalised upset mass is
//   U(E) = sum_{S >= E} P(S) prod_t f_t(c_t(S))
//        = Q(E) * prod_{t : E_t <= E} f_t(1) * sum over the free atoms of the factors of
//          the remaining pools A_t = E_t \ E,
// which a dynamic program over the set of pools that can still be contained evaluates as a
// sum of non-negative terms. The normaliser is the same program for E = {} (all atoms free),
// so M(E) = U(E) / U({}) carries only a few ulp of rounding. (A recurrence on the per-update
// denominators would subtract nearly equal masses whenever a pool is likely negative.) The
// programs stay small in practice: adaptive experiments reuse the same atoms, so the pool
// sets that can still be contained number in the tens, not 2^updates.
//
// Dilution lattices (replicated_lattice.cpp:17-28): the factor of update t depends on the
// number m_t of positive atoms in the pool, f_t(m) = D[|E_t| - 1][m] or 1 - D[|E_t| - 1][m]
// (dilution.cpp:44-67). The program then carries, for each pool that is partly processed,
// its count of positive atoms so far (pools with the same remaining atoms merge, their factor
// vectors multiplied entry by entry), and applies f(count) when the pool's last atom has been
// processed.
#pragma once

#include <cstdint>
#include <functional>
#include <vector>

namespace bgt_hip {

// Candidates whose remaining pools number at most kDenseMaxPools, under histories of at
// most kDenseMaxHist updates, run the dense program: the same operations in the same
// order as the device's band kernel, so host and device masses are bit-identical.
constexpr int kDenseMaxPools = 10;
constexpr int kDenseMaxHist = 64;

// f(i, worker) for i in [0, n) on the host threads of the exact rescoring (up to 16, the
// calling thread included); returns when all are done. worker < parallel_width().
void parallel_for(size_t n, const std::function<void(size_t, int)> &f);
int parallel_width();

class ExactHistory {
public:
	ExactHistory() = default;
	// A copy carries the program (prior, history, atom order, normaliser), not the working
	// memory of its evaluations, which grows to megabytes per worker after rescoring and
	// would otherwise be copied with every child lattice and projection.
	ExactHistory(const ExactHistory &o) { copy_program(o); }
	ExactHistory &operator=(const ExactHistory &o)
	{
		if (this != &o)
			copy_program(o);
		return *this;
	}
	// dilution: the pushes carry dilution tables (subjects x (subjects + 1), row s-1 for pool
	// size s). The host's pool program covers binary lattices (variants == 1); lattices of
	// several variants keep their history for the device enumeration only (enum_inputs).
	void reset(bool usable, int atoms, const double *prior, bool dilution = false, int variants = 1);
	bool usable() const { return usable_ && !broken_; }
	bool dilution() const { return dilution_; }
	int variants() const { return variants_; }
	// pools of the device enumeration: one per update (each variant block separately)
	int enum_pools() const { return (int)hist_.size(); }
	void push(uint32_t experiment, uint32_t response, const double *dilution_table = nullptr);
	int updates() const { return (int)hist_.size(); }
	// M(E) after all pushed updates; false if the state budget of the program is exceeded
	bool mass(uint32_t experiment, double &out);
	// M for `count` experiments (ok[i] false where the budget was exceeded); false if the
	// normaliser itself is out of budget. Large requests are spread over a pool of host
	// worker threads; each result is the same as mass() would give.
	bool masses(const uint32_t *experiments, size_t count, double *out, bool *ok);
	// relative rounding bound of mass() (a few ulp per atom and per update)
	double rel_bound() const;
	// update t of the history: its pool and its factors when contained / not contained
	void record(int t, uint32_t &experiment, double &f_in, double &f_out) const
	{
		experiment = hist_[t].e;
		f_in = hist_[t].f_in;
		f_out = hist_[t].f_out;
	}
	// inputs of the device enumeration (launch_exact_enum) over variant v's block of subjects
	// atoms: pairs[2i + bit] of its atoms (bit 1: atom v * subjects + i is negative), each
	// update's pool and variant v's factor by the pool's count m of positive atoms in the
	// block, fac[t * stride + m] for m <= |pool| < stride (zero beyond). The prior is a
	// product over atoms and every likelihood factor depends on one block
	// (replicated_lattice.cpp:8-28), so the posterior is a product over the variant blocks
	// and the mass of an atom set is the product of its blocks' masses.
	void enum_inputs(int variant, double *pairs, uint32_t *pools, double *fac, int stride) const;
	int subjects() const { return subjects_; }
	// the prior's (pi_a, 1 - pi_a) for every atom, pairs[2a + bit]
	void prior_pairs(double *pairs) const
	{
		for (int a = 0; a < atoms_; ++a) {
			pairs[2 * a] = pi_[a];
			pairs[2 * a + 1] = q_[a];
		}
	}
	// U(E) / U({}) for an unnormalised mass computed elsewhere; false without a normaliser
	bool normalise(double u, double &out)
	{
		if (!usable() || !normaliser())
			return false;
		out = u / z_;
		return true;
	}
	// compute the normaliser of the current history now (host work that can overlap with
	// device work; mass() would otherwise do it on first use). Dilution programs, whose
	// normaliser carries counts of every pool and costs far more, wait for a re-decision
	// that needs them.
	void prepare(bool dilution_too = false)
	{
		if (usable() && (!dilution_ || dilution_too))
			normaliser();
	}

private:
	struct Upd {
		uint32_t e;
		double f_in, f_out;   // factor when the pool is / is not contained in S
		std::vector<double> f; // dilution: factor by the pool's count of positive atoms
		// several variants: variant v's factor by its count m of positive atoms in the pool,
		// fv[v * (subjects + 1) + m]
		std::vector<double> fv;
	};
	// Pool-set states of the program in insertion order with their weights. Up to
	// kDenseBits pools, states index a direct map (reset entry by entry); beyond that an
	// open-addressing hash. Either way entries keep insertion order, so the summation order
	// is the same.
	static constexpr int kDenseBits = 16;
	struct Table {
		std::vector<uint64_t> key;
		std::vector<double> w;
		std::vector<int32_t> slot;
		bool dense = false;
		void clear(size_t expect, int pools);
		void add(uint64_t k, double v);
	};
	struct alignas(128) Scratch { // aligned: the pool's threads each grow their own
		std::vector<uint32_t> pool;
		std::vector<double> fin, fout;
		std::vector<std::vector<double>> g; // dilution: merged factor vectors
		Table a, b;
		std::vector<double> da, db; // dense program
	};
	void copy_program(const ExactHistory &o)
	{
		usable_ = o.usable_;
		broken_ = o.broken_;
		dilution_ = o.dilution_;
		atoms_ = o.atoms_;
		variants_ = o.variants_;
		subjects_ = o.subjects_;
		pi_ = o.pi_;
		q_ = o.q_;
		hist_ = o.hist_;
		order_ = o.order_;
		z_ = o.z_;
		z_ok_ = o.z_ok_;
	}
	bool unnormalised(uint32_t e, double &out, Scratch &sc) const;
	bool counts(uint32_t e, double &out, Scratch &sc) const; // the dilution program
	double dense(uint32_t e, Scratch &sc) const; // sc.pool/fin/fout built by unnormalised
	void plan_order(); // atom order of the program for the current history (normaliser())
	std::vector<Scratch> workers_; // one per pool thread (masses())
	bool normaliser();

	bool usable_ = false, broken_ = false, dilution_ = false;
	int atoms_ = 0;
	int variants_ = 1, subjects_ = 0;
	std::vector<double> pi_, q_;
	std::vector<Upd> hist_;
	std::vector<int> order_; // atom order of the program
	double z_ = 1.0;   // U({}) for the current history
	bool z_ok_ = true; // z_ is up to date
	Scratch scratch_;
};

} // namespace bgt_hip
