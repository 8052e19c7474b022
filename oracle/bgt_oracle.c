/*
 * bgt_oracle.c -- CPU restatement of the BayesianGroupTesting hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker and the timed CPU baseline);
 * the HIP provider never links or calls it.
 *
 * Each function cites the reference file:line whose serial loop it restates.
 * Built twice by oracle/Makefile:
 *   liboracle.so       strict IEEE (no contraction), the parity checker;
 *   liboracle_fast.so  the reference's default flags (-O3 -ffast-math), timed
 *                      as the CPU baseline; with OpenMP, used only by
 *                      orc_op_bha_omp (the reference's own OpenMP loops).
 */
#define _POSIX_C_SOURCE 199309L
#ifdef _OPENMP
#include <omp.h>
#endif
#include "bgt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define ORC_NEGATIVE_RESPONSE 0.99 /* replicated_lattice.cpp:6 */

static inline int popc(uint32_t v) { return __builtin_popcount(v); }
static inline uint32_t bit(int i) { return (uint32_t)1u << i; }

double orc_now(void)
{
	struct timespec ts;
	clock_gettime(CLOCK_MONOTONIC, &ts);
	return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

uint32_t orc_reverse_bits(uint32_t value, int width)
{
	/* lattice_api.cpp:17-26 */
	uint32_t ret = 0;
	for (int i = 0; i < width; i++) {
		ret <<= 1;
		ret |= (value >> i) & 1u;
	}
	return ret;
}

void orc_dilution_table(int n, double alpha, double h, double *out)
{
	/* dilution.cpp:44-67: row rk-1 holds pool size rk; stride n+1 */
	for (int i = 0; i < n * (n + 1); i++)
		out[i] = 0.0;
	for (int rk = 1; rk <= n; rk++) {
		double *row = out + (size_t)(rk - 1) * (size_t)(n + 1);
		row[0] = alpha;
		for (int r = 1; r <= rk; r++) {
			const int k = rk - r;
			row[r] = 1 - alpha * r / (k * h + r);
		}
	}
}

double orc_prior_prob(uint32_t state, int atoms, const double *pi)
{
	/* lattice_base.cpp:197-208 */
	double prob = 1.0;
	for (int i = 0; i < atoms; i++) {
		if ((state & bit(i)) == 0)
			prob *= pi[i];
		else
			prob *= (1.0 - pi[i]);
	}
	return prob;
}

void orc_prior(int atoms, const double *pi, double *post)
{
	/* lattice_base.cpp:184-195 */
	const uint32_t count = bit(atoms);
	for (uint32_t s = 0; s < count; s++)
		post[s] = orc_prior_prob(s, atoms, pi);
}

double orc_response_prob(int dilution, int curr_subjs, int variants, uint32_t e, uint32_t r,
						 uint32_t s, const double *dil, int dil_stride)
{
	double ret = 1.0;
	if (!dilution) {
		/* replicated_lattice.cpp:8-14 */
		for (int v = 0; v < variants; v++) {
			const int contained = (e & (s >> (v * curr_subjs))) == e;
			const int rbit = (r & (1u << v)) != 0;
			ret *= contained ? (rbit ? ORC_NEGATIVE_RESPONSE : (1.0 - ORC_NEGATIVE_RESPONSE))
							 : (rbit ? (1.0 - ORC_NEGATIVE_RESPONSE) : ORC_NEGATIVE_RESPONSE);
		}
		return ret;
	}
	/* replicated_lattice.cpp:17-28. A pool of size 0 indexes row -1 there
	 * (undefined); this restatement defines P(r_v = 1 | E = 0) = 0. */
	const int len = popc(e);
	for (int v = 0; v < variants; v++) {
		const int rbit = (r & (1u << v)) != 0;
		double p1 = 0.0;
		if (len > 0) {
			const int missing = len - popc(e & (s >> (v * curr_subjs)));
			p1 = dil[(size_t)(len - 1) * (size_t)dil_stride + (size_t)missing];
		}
		ret *= rbit ? p1 : 1.0 - p1;
	}
	return ret;
}

double orc_update_in_place(int dilution, int curr_subjs, int variants, double *post, uint32_t e,
						   uint32_t r, const double *dil, int dil_stride)
{
	/* lattice_base.cpp:230-244; the `/ denominator` loop compiles to a
	 * multiply by one reciprocal under the reference's -ffast-math. */
	const uint32_t count = bit(curr_subjs * variants);
	double denominator = 0.0;
	for (uint32_t i = 0; i < count; i++) {
		post[i] = post[i] * orc_response_prob(dilution, curr_subjs, variants, e, r, i, dil, dil_stride);
		denominator += post[i];
	}
	const double inv = 1.0 / denominator;
	for (uint32_t i = 0; i < count; i++)
		post[i] = post[i] * inv;
	return denominator;
}

double orc_update_out(int dilution, int curr_subjs, int variants, const double *post, double *out,
					  uint32_t e, uint32_t r, const double *dil, int dil_stride)
{
	/* lattice_base.cpp:210-228 */
	const uint32_t count = bit(curr_subjs * variants);
	double denominator = 0.0;
	for (uint32_t i = 0; i < count; i++) {
		out[i] = post[i] * orc_response_prob(dilution, curr_subjs, variants, e, r, i, dil, dil_stride);
		denominator += out[i];
	}
	const double inv = 1.0 / denominator;
	for (uint32_t i = 0; i < count; i++)
		out[i] = out[i] * inv;
	return denominator;
}

void orc_zeta(double *a, int n)
{
	/* bbpa.cpp:497-512: bit sweeps 0..n-1, a[e] += a[e | bit] for e with bit clear */
	const uint32_t count = bit(n);
	for (int subject = 0; subject < n; subject++) {
		const uint32_t b = bit(subject);
		const uint32_t step = b << 1;
		for (uint32_t base = 0; base < count; base += step)
			for (uint32_t off = 0; off < b; off++)
				a[base + off] += a[base + off + b];
	}
}

uint32_t orc_op_bha(const double *post, int n, double *upset_out)
{
	/* bbpa.cpp:902-916 (copy, zeta) and 537-551 (argmin |M - 0.5|, lowest
	 * index wins ties, initial best distance 2.0) */
	const uint32_t count = bit(n);
	double *a = upset_out ? upset_out : (double *)malloc(sizeof(double) * count);
	for (uint32_t s = 0; s < count; s++)
		a[s] = post[s];
	orc_zeta(a, n);
	uint32_t candidate = 0;
	double best = 2.0;
	for (uint32_t e = 0; e < count; e++) {
		const double d = fabs(a[e] - 0.5);
		if (d < best) {
			best = d;
			candidate = e;
		}
	}
	if (!upset_out)
		free(a);
	return candidate;
}

uint32_t orc_op_bha_omp(const double *post, int n, double *upset_out)
{
	/* bbpa.cpp:920-940 (op_bha_omp): the copy, subset_zeta_transform_omp (bbpa.cpp:517-533:
	 * sweeps in order, the blocks of one sweep in parallel) and best_op_bha_candidate_omp
	 * (bbpa.cpp:555-571: the lowest distance, ties to the lowest index), from 2^16 states.
	 * Built without OpenMP it is orc_op_bha. Used only as the all-cores CPU baseline. */
	const uint32_t count = bit(n);
	if (count < (1u << 16))
		return orc_op_bha(post, n, upset_out);
	double *a = upset_out ? upset_out : (double *)malloc(sizeof(double) * count);
	const long long cnt = (long long)count;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
	for (long long s = 0; s < cnt; s++)
		a[s] = post[s];
	for (int subject = 0; subject < n; subject++) {
		const long long b = (long long)bit(subject);
		const long long step = b << 1;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
		for (long long base = 0; base < cnt; base += step)
			for (long long off = 0; off < b; off++)
				a[base + off] += a[base + off + b];
	}
	uint32_t candidate = 0;
	double best = 2.0;
#ifdef _OPENMP
#pragma omp parallel
#endif
	{
		uint32_t lc = 0;
		double lb = 2.0;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
		for (long long e = 0; e < cnt; e++) {
			const double d = fabs(a[e] - 0.5);
			if (d < lb || (d == lb && (uint32_t)e < lc)) {
				lb = d;
				lc = (uint32_t)e;
			}
		}
#ifdef _OPENMP
#pragma omp critical
#endif
		{
			if (lb < best || (lb == best && lc < candidate)) {
				best = lb;
				candidate = lc;
			}
		}
	}
	if (!upset_out)
		free(a);
	return candidate;
}

int orc_threads(void)
{
#ifdef _OPENMP
	int t = 1;
#pragma omp parallel
	{
#pragma omp single
		t = omp_get_num_threads();
	}
	return t;
#else
	return 1;
#endif
}

static uint32_t bbpa_core(const double *post, int n, int k, double target, double *best_score, double *scores)
{
	/* bbpa.cpp:950-970 (BBPA_serial); scores[e], when given, receives every experiment's
	 * objective as the scorers form it */
	const uint32_t ec = bit(n);
	const uint32_t rc = bit(k);
	const uint32_t mask = ec - 1;
	const uint32_t states = bit(n * k);
	double *fm = (double *)calloc((size_t)ec * rc, sizeof(double));
	double best = 2.0;
	uint32_t cand = 0xFFFFFFFFu;
	double total = 0.0;
	if (k == 1) {
		/* bbpa.cpp:79-96, 201-220 */
		for (uint32_t s = 0; s < states; s++) {
			fm[s & mask] += post[s];
			total += post[s];
		}
		orc_zeta(fm, n);
		for (uint32_t e = 0; e < ec; e++) {
			const double positive = fm[e];
			const double negative = total - positive;
			const double score = fabs(positive - target) + fabs(negative - target);
			if (scores)
				scores[e] = score;
			if (score < best || (score == best && e < cand)) {
				best = score;
				cand = e;
			}
		}
		free(fm);
		if (best_score)
			*best_score = best;
		return cand;
	}
	if (k == 2) {
		/* bbpa.cpp:113-141 */
		double *v0 = fm + ec, *v1 = fm + 2 * ec, *both = fm + 3 * ec;
		for (uint32_t s = 0; s < states; s++) {
			const uint32_t m0 = s & mask;
			const uint32_t m1 = (s >> n) & mask;
			const double mass = post[s];
			v0[m0] += mass;
			v1[m1] += mass;
			both[m0 & m1] += mass;
			total += mass;
		}
		fm[0] = total;
	} else {
		/* bbpa.cpp:237-285 */
		uint32_t vm[32], inter[1 << 10];
		for (uint32_t s = 0; s < states; s++) {
			const double mass = post[s];
			for (int v = 0; v < k; v++)
				vm[v] = (s >> (v * n)) & mask;
			inter[0] = mask;
			for (uint32_t sub = 1; sub < rc; sub++) {
				const uint32_t low = sub & (0u - sub);
				const int v = __builtin_ctz(low);
				const uint32_t prev = sub ^ low;
				const uint32_t x = inter[prev] & vm[v];
				inter[sub] = x;
				fm[(size_t)sub * ec + x] += mass;
			}
			total += mass;
		}
		fm[0] = total;
	}
	/* bbpa.cpp:292-300 */
	for (uint32_t sub = 1; sub < rc; sub++)
		orc_zeta(fm + (size_t)sub * ec, n);
	if (k == 2) {
		/* bbpa.cpp:332-365 */
		const double *v0 = fm + ec, *v1 = fm + 2 * ec, *both = fm + 3 * ec;
		for (uint32_t e = 0; e < ec; e++) {
			const double p0 = v0[e], p1 = v1[e], pb = both[e];
			const double r0 = pb;
			const double r1 = p1 - pb;
			const double r2 = p0 - pb;
			const double r3 = total - p0 - p1 + pb;
			const double score = fabs(r0 - target) + fabs(r1 - target) + fabs(r2 - target) + fabs(r3 - target);
			if (scores)
				scores[e] = score;
			if (score < best || (score == best && e < cand)) {
				best = score;
				cand = e;
			}
		}
	} else {
		/* bbpa.cpp:416-460 */
		const uint32_t all = rc - 1;
		for (uint32_t e = 0; e < ec; e++) {
			double score = 0.0;
			for (uint32_t resp = 0; resp < rc; resp++) {
				const uint32_t positive = (~resp) & all;
				const uint32_t missing = resp;
				double exact = 0.0;
				for (uint32_t extra = missing;; extra = (extra - 1) & missing) {
					const uint32_t fs = positive | extra;
					const double f = fs == 0 ? total : fm[(size_t)fs * ec + e];
					if ((popc(extra) & 1) == 0)
						exact += f;
					else
						exact -= f;
					if (extra == 0)
						break;
				}
				score += fabs(exact - target);
			}
			if (scores)
				scores[e] = score;
			if (score < best || (score == best && e < cand)) {
				best = score;
				cand = e;
			}
		}
	}
	free(fm);
	if (best_score)
		*best_score = best;
	return cand;
}

uint32_t orc_bbpa(const double *post, int n, int k, double target, double *best_score)
{
	return bbpa_core(post, n, k, target, best_score, NULL);
}

uint32_t orc_bbpa_scores(const double *post, int n, int k, double target, double *scores)
{
	return bbpa_core(post, n, k, target, NULL, scores);
}

static inline uint32_t response_bucket(uint32_t e, uint32_t s, int n, int k)
{
	/* benchmarks/experiment_selection_benchmark.cpp:200-212 */
	const uint32_t mask = bit(n) - 1;
	uint32_t r = 0;
	for (int v = 0; v < k; v++) {
		const uint32_t vs = (s >> (v * n)) & mask;
		r |= (uint32_t)((e & ~vs) != 0) << v;
	}
	return r;
}

uint32_t orc_brute_force(const double *post, int n, int k, double *best_score)
{
	/* benchmarks/experiment_selection_benchmark.cpp:214-247: explicit
	 * (experiment, state) bucket sweep, strict < from +inf */
	const uint32_t ec = bit(n), rc = bit(k), states = bit(n * k);
	const double target = 1.0 / (double)rc;
	double mass[1 << 10];
	uint32_t cand = 0;
	double best = INFINITY;
	for (uint32_t e = 0; e < ec; e++) {
		for (uint32_t i = 0; i < rc; i++)
			mass[i] = 0.0;
		for (uint32_t s = 0; s < states; s++)
			mass[response_bucket(e, s, n, k)] += post[s];
		double score = 0.0;
		for (uint32_t i = 0; i < rc; i++)
			score += fabs(mass[i] - target);
		if (score < best) {
			best = score;
			cand = e;
		}
	}
	if (best_score)
		*best_score = best;
	return cand;
}

uint32_t orc_select(const double *post, int n, int k, int selector)
{
	/* bbpa.cpp:861-871 with resolve_selector (simulation_helpers.cpp:40-45) */
	if (selector == 0)
		selector = k == 1 ? 1 : 2;
	if (selector == 1)
		return k == 1 ? orc_op_bha(post, n, NULL) : 0xFFFFFFFFu;
	if (selector == 2)
		return orc_bbpa(post, n, k, 1.0 / (double)bit(k), NULL);
	return orc_brute_force(post, n, k, NULL);
}

void orc_atom_masses(const double *post, int atoms, double *out)
{
	/* lattice_base.cpp:25-41: state-major, per-atom sequential sums */
	const uint32_t count = bit(atoms);
	for (int a = 0; a < atoms; a++)
		out[a] = 0.0;
	for (uint32_t s = 0; s < count; s++) {
		const double p = post[s];
		for (int a = 0; a < atoms; a++)
			if (s & bit(a))
				out[a] += p;
	}
}

void orc_classify(const double *masses, int atoms, double up, double lo, uint32_t *pos, uint32_t *neg)
{
	/* lattice_base.cpp:246-264 (masks are sticky) */
	for (int i = 0; i < atoms; i++) {
		const uint32_t placement = bit(i);
		if ((*pos | *neg) & placement)
			continue;
		const double m = masses[i];
		if (m < lo)
			*pos |= placement;
		else if (m > (1 - up))
			*neg |= placement;
	}
}

double orc_upset_mass(const double *post, int atoms, uint32_t e)
{
	/* lattice_base.cpp:318-351 (active generation over free bits) */
	int idx[32], m = 0;
	for (int i = 0; i < atoms; i++)
		if (!(e & bit(i)))
			idx[m++] = i;
	double ret = 0.0;
	const uint32_t pw = bit(m);
	for (uint32_t i = 0; i < pw; i++) {
		uint32_t t = e;
		for (int j = 0; j < m; j++)
			if (i & bit(j))
				t |= bit(idx[j]);
		ret += post[t];
	}
	return ret;
}

int orc_select_experiments(const double *post, int n, int steps, uint32_t *out)
{
	/* lattice_api.cpp:116-168, binary Op-BHA lookahead in PUBLIC
	 * coordinates: the first pick is Op-BHA, later picks are a greedy
	 * brute-force sweep over the partition into 2^|selected| buckets. */
	if (steps <= 0)
		return 0;
	const uint32_t total = bit(n);
	out[0] = orc_reverse_bits(orc_op_bha(post, n, NULL), n);
	double *pub = (double *)malloc(sizeof(double) * total);
	for (uint32_t s = 0; s < total; s++)
		pub[s] = post[orc_reverse_bits(s, n)];
	double *mass = (double *)malloc(sizeof(double) * ((size_t)1 << steps));
	for (int step = 1; step < steps; step++) {
		double min_score = 1.7976931348623157e308;
		uint32_t min_e = 0;
		const int groups = step + 1;
		for (uint32_t c = 1; c < total; c++) {
			for (int g = 0; g < (1 << groups); g++)
				mass[g] = 0.0;
			for (uint32_t s = 0; s < total; s++) {
				int gi = 0;
				for (int i = 0; i < step; i++)
					if ((out[i] & s) == s)
						gi |= 1 << i;
				if ((c & s) == s)
					gi |= 1 << step;
				mass[gi] += pub[s];
			}
			const double target = 1.0 / (double)(1 << groups);
			double score = 0.0;
			for (int g = 0; g < (1 << groups); g++)
				score += fabs(mass[g] - target);
			if (score < min_score) {
				min_score = score;
				min_e = c;
			}
		}
		out[step] = min_e;
	}
	free(mass);
	free(pub);
	return steps;
}

int orc_lookahead_scores(const double *post, int n, const uint32_t *selected, int m, double *scores)
{
	/* lattice_api.cpp:139-158: the score of every public candidate c (scores[c], c >= 1)
	 * after the public experiments selected[0..m-1], in the reference's summation order:
	 * states ascending, each added to its group's mass, then |mass - 1/2^(m+1)| summed in
	 * group order. scores[0] is left untouched. */
	if (m < 1 || m > 15)
		return -1;
	const uint32_t total = bit(n);
	const int groups = 1 << (m + 1);
	double *mass = (double *)malloc(sizeof(double) * (size_t)groups);
	int *cls = (int *)malloc(sizeof(int) * total);
	double *pub = (double *)malloc(sizeof(double) * total);
	for (uint32_t s = 0; s < total; s++) {
		pub[s] = post[orc_reverse_bits(s, n)];
		int gi = 0;
		for (int i = 0; i < m; i++)
			if ((selected[i] & s) == s)
				gi |= 1 << i;
		cls[s] = gi;
	}
	const double target = 1.0 / (double)groups;
	for (uint32_t c = 1; c < total; c++) {
		for (int g = 0; g < groups; g++)
			mass[g] = 0.0;
		for (uint32_t s = 0; s < total; s++)
			mass[cls[s] | ((c & s) == s ? 1 << m : 0)] += pub[s];
		double score = 0.0;
		for (int g = 0; g < groups; g++)
			score += fabs(mass[g] - target);
		scores[c] = score;
	}
	free(pub);
	free(cls);
	free(mass);
	return 0;
}

/* ------------------------------------------------------------------ shrink */

static uint32_t curr_shrinkable_atoms(uint32_t curr_clas_atoms, int curr, int k)
{
	/* shrink_plan.cpp:34-56 */
	uint32_t ret = 0;
	for (int i = 0; i < curr; i++) {
		int all = 1;
		for (int j = 0; j < k; j++)
			if (!(curr_clas_atoms & bit(j * curr + i)))
				all = 0;
		if (all)
			for (int j = 0; j < k; j++)
				ret |= bit(j * curr + i);
	}
	return ret;
}

static uint32_t update_clas_subj(uint32_t clas_atoms, int orig, int k)
{
	/* shrink_plan.cpp:59-79 */
	uint32_t ret = 0;
	for (int i = 0; i < orig; i++) {
		int all = 1;
		for (int j = 0; j < k; j++)
			if (!(clas_atoms & bit(j * orig + i)))
				all = 0;
		if (all)
			ret |= bit(i);
	}
	return ret;
}

void orc_make_shrink_plan(int orig, int curr, int k, uint32_t classified_atoms,
						  uint32_t already_classified_subjects, orc_shrink_plan *p)
{
	/* shrink_plan.cpp:124-177 */
	memset(p, 0, sizeof(*p));
	p->original_subjects = orig;
	p->current_subjects = curr;
	p->variants = k;
	p->source_atoms = curr * k;
	for (int i = 0; i < 32; i++)
		p->original_to_current_atom_position[i] = -1;
	int o2c[32];
	int cs = 0;
	for (int s = 0; s < orig; s++) {
		o2c[s] = -1;
		if (already_classified_subjects & bit(s))
			continue;
		o2c[s] = cs++;
	}
	for (int s = 0; s < orig; s++) {
		const int csp = o2c[s];
		if (csp < 0)
			continue;
		for (int v = 0; v < k; v++) {
			const int oa = v * orig + s;
			const int ca = v * curr + csp;
			p->original_to_current_atom_position[oa] = ca;
			if (classified_atoms & bit(oa))
				p->current_classified_atoms |= bit(ca);
		}
	}
	p->removable_current_atoms = curr_shrinkable_atoms(p->current_classified_atoms, curr, k);
	p->reduced_atom_count = popc(p->removable_current_atoms);
	p->base_atom_count = p->source_atoms - p->reduced_atom_count;
	int nb = 0, nr = 0;
	for (int a = 0; a < p->source_atoms; a++) {
		if (p->removable_current_atoms & bit(a))
			p->reduced_atom_masks[nr++] = bit(a);
		else
			p->base_atom_masks[nb++] = bit(a);
	}
	p->new_classified_subjects = update_clas_subj(classified_atoms, orig, k);
	p->new_current_subjects = orig - popc(p->new_classified_subjects);
}

static inline uint32_t expand_compact(uint32_t compact, const uint32_t *masks, int count)
{
	/* shrink.hpp:56-65 */
	uint32_t x = 0;
	for (int i = 0; i < count; i++)
		if (compact & bit(i))
			x |= masks[i];
	return x;
}

void orc_project_posterior(const double *src, double *dst, const orc_shrink_plan *p)
{
	/* shrink.hpp:70-86 */
	const uint32_t base_states = bit(p->base_atom_count);
	const uint32_t red_states = bit(p->reduced_atom_count);
	for (uint32_t b = 0; b < base_states; b++) {
		const uint32_t bs = expand_compact(b, p->base_atom_masks, p->base_atom_count);
		double sum = 0.0;
		for (uint32_t r = 0; r < red_states; r++)
			sum += src[bs | expand_compact(r, p->reduced_atom_masks, p->reduced_atom_count)];
		dst[b] = sum;
	}
}

uint32_t orc_true_ex(uint32_t halving, uint32_t clas_subjs, int orig_subjs)
{
	/* tree.cpp:55-68 */
	uint32_t ret = 0, pos = 0;
	for (int i = 0; i < orig_subjs; i++) {
		if (!(clas_subjs & bit(i))) {
			if (halving & bit((int)pos))
				ret |= bit(i);
			++pos;
		}
	}
	return ret;
}

/* -------------------------------------------------------------------- tree */

typedef struct orc_node {
	uint32_t ex, res;
	int stage;
	double b;
	uint32_t pos, neg;
	int nchild;
	struct orc_node *child[1 << 5];
} orc_node;

typedef struct orc_tree_cfg {
	int dilution, orig, k, depth, selector;
	double up, lo, branch;
	const double *dil;
	int dil_stride;
	const double *pi;
} orc_tree_cfg;

static void update_metadata_with_shrinking(const orc_tree_cfg *c, double **post, int *curr,
										   uint32_t *pos, uint32_t *neg, uint32_t *clas_subjs)
{
	/* lattice_base.cpp:266-305 and apply_shrink_plan 307-316 */
	const uint32_t clas_atoms = *pos | *neg;
	orc_shrink_plan old_plan;
	orc_make_shrink_plan(c->orig, *curr, c->k, clas_atoms, *clas_subjs, &old_plan);
	const int ca = *curr * c->k;
	double masses[32];
	orc_atom_masses(*post, ca, masses);
	uint32_t np = 0, nn = 0;
	for (int i = 0; i < c->orig * c->k; i++) {
		if (clas_atoms & bit(i))
			continue;
		const int cp = old_plan.original_to_current_atom_position[i];
		if (cp < 0)
			continue;
		const double m = masses[cp];
		if (m < c->lo)
			np |= bit(i);
		else if (m > (1 - c->up))
			nn |= bit(i);
	}
	*pos |= np;
	*neg |= nn;
	orc_shrink_plan plan;
	orc_make_shrink_plan(c->orig, *curr, c->k, *pos | *neg, *clas_subjs, &plan);
	if (plan.removable_current_atoms != 0) {
		double *proj = (double *)malloc(sizeof(double) * bit(plan.base_atom_count));
		orc_project_posterior(*post, proj, &plan);
		free(*post);
		*post = proj;
		*clas_subjs = plan.new_classified_subjects;
		*curr = plan.new_current_subjects;
	}
}

int orc_shrink_step(int orig, int k, double up, double lo, const double *post, int curr, uint32_t *pos,
					uint32_t *neg, uint32_t *clas_subjs, double *out)
{
	/* one update_metadata_with_shrinking (lattice_base.cpp:266-316) on a copy of post; out
	 * receives the posterior after it, 2^(curr' k) entries for the returned curr' */
	orc_tree_cfg c;
	memset(&c, 0, sizeof c);
	c.orig = orig;
	c.k = k;
	c.up = up;
	c.lo = lo;
	const size_t n = (size_t)1 << (curr * k);
	double *p = (double *)malloc(n * sizeof(double));
	memcpy(p, post, n * sizeof(double));
	int cc = curr;
	update_metadata_with_shrinking(&c, &p, &cc, pos, neg, clas_subjs);
	memcpy(out, p, ((size_t)1 << (cc * k)) * sizeof(double));
	free(p);
	return cc;
}

static orc_node *build_local_tree(const orc_tree_cfg *c, double *post, int curr, uint32_t pos,
								  uint32_t neg, uint32_t clas_subjs, uint32_t ex, uint32_t res,
								  int stage)
{
	/* global_tree_intra.cpp:8-40 */
	orc_node *node = (orc_node *)calloc(1, sizeof(orc_node));
	node->ex = ex;
	node->res = res;
	node->stage = stage;
	node->b = 0.0;
	node->pos = pos;
	node->neg = neg;
	const int classified = popc(pos | neg) == c->orig * c->k;
	if (!classified && stage < c->depth) {
		const int rc = 1 << c->k;
		node->nchild = rc;
		const uint32_t halving = orc_select(post, curr, c->k, c->selector);
		const uint32_t full = orc_true_ex(halving, clas_subjs, c->orig);
		const uint32_t states = bit(curr * c->k);
		for (int r = 0; r < rc; r++) {
			double *cp;
			if (r != rc - 1) {
				cp = (double *)malloc(sizeof(double) * states);
				orc_update_out(c->dilution, curr, c->k, post, cp, halving, (uint32_t)r, c->dil, c->dil_stride);
			} else {
				cp = post;
				post = NULL;
				orc_update_in_place(c->dilution, curr, c->k, cp, halving, (uint32_t)r, c->dil, c->dil_stride);
			}
			int ccurr = curr;
			uint32_t cpos = pos, cneg = neg, cclas = clas_subjs;
			update_metadata_with_shrinking(c, &cp, &ccurr, &cpos, &cneg, &cclas);
			node->child[r] = build_local_tree(c, cp, ccurr, cpos, cneg, cclas, full, (uint32_t)r, stage + 1);
		}
	} else {
		free(post);
	}
	return node;
}

static void free_tree(orc_node *n)
{
	if (!n)
		return;
	for (int i = 0; i < n->nchild; i++)
		free_tree(n->child[i]);
	free(n);
}

static int count_leaves(const orc_node *n)
{
	/* tree.cpp:134-152 */
	if (n->nchild == 0)
		return 1;
	int t = 0;
	for (int i = 0; i < n->nchild; i++)
		t += count_leaves(n->child[i]);
	return t;
}

static void apply_true_state(const orc_tree_cfg *c, orc_node *node, uint32_t t, double prob)
{
	/* tree.cpp:174-199 (the original lattice keeps all subjects) */
	node->b = prob;
	for (int i = 0; i < node->nchild; i++) {
		orc_node *ch = node->child[i];
		const double cp = prob * orc_response_prob(c->dilution, c->orig, c->k, ch->ex, ch->res, t, c->dil, c->dil_stride);
		if (cp > c->branch)
			apply_true_state(c, ch, t, cp);
		else
			ch->b = 0.0;
	}
}

typedef struct orc_acc {
	double correct[30], incorrect[30], fp[30], fn[30];
	double unclas, exp_stage, exp_test;
} orc_acc;

static void find_all_stat(const orc_tree_cfg *c, const orc_node *n, const orc_node **leaves, int *count)
{
	/* tree.cpp:154-172 */
	if (n == NULL || n->b < c->branch)
		return;
	if (n->nchild == 0) {
		leaves[(*count)++] = n;
		return;
	}
	for (int i = 0; i < n->nchild; i++)
		find_all_stat(c, n->child[i], leaves, count);
}

static void parse(const orc_tree_cfg *c, const orc_node *root, uint32_t t, const orc_node **leaves, orc_acc *st)
{
	/* tree.cpp:70-132 (standard deviations are computed there but never
	 * reported, so they are not restated) */
	memset(st, 0, sizeof(*st));
	const double coef = orc_prior_prob(t, c->orig * c->k, c->pi) * 1.0;
	int count = 0;
	find_all_stat(c, root, leaves, &count);
	const int atoms = c->orig * c->k;
	for (int i = 0; i < count; i++) {
		const orc_node *leaf = leaves[i];
		const int index = leaf->stage;
		const int classified = popc(leaf->pos | leaf->neg) == atoms;
		if (classified && leaf->neg == t) {
			st->correct[index] += leaf->b * coef;
		} else if (classified && leaf->neg != t) {
			st->incorrect[index] += leaf->b * coef;
			const double tp = (double)popc(leaf->pos);
			const double tn = (double)popc(leaf->neg);
			const double fp = tp == 0.0 ? 0.0 : popc((leaf->neg ^ t) & t) / tp;
			const double fn = tn == 0.0 ? 0.0 : popc((leaf->neg ^ t) & (~t)) / tn;
			st->fp[index] += fp * coef * leaf->b;
			st->fn[index] += fn * coef * leaf->b;
		} else if (!classified) {
			st->unclas = st->unclas + leaf->b;
		}
		st->exp_stage = st->exp_stage + ceil((double)index / 1.0) * leaf->b;
		st->exp_test = st->exp_test + index * leaf->b;
	}
	st->unclas = st->unclas * coef;
	st->exp_stage = st->exp_stage * coef;
	st->exp_test = st->exp_test * coef;
}

int orc_run_simulation(int dilution, int n, int k, const double *prior, int depth, double up,
					   double lo, double branch, int selector, const double *dil, int dil_stride,
					   orc_tree_stats *stats)
{
	/* simulation_runner.cpp:92-132 and providers.cpp:292-332 */
	if (n <= 0 || k <= 0 || depth < 0 || depth > 29 || n * k > 30 || k > 5)
		return 1;
	if (selector == 0)
		selector = k == 1 ? 1 : 2;
	if (selector == 1 && k != 1)
		return 1;
	if (dilution && dil == NULL)
		return 1;
	orc_tree_cfg c = {dilution, n, k, depth, selector, up, lo, branch, dil, dil_stride, prior};
	const int atoms = n * k;
	double *post = (double *)malloc(sizeof(double) * bit(atoms));
	orc_prior(atoms, prior, post);
	orc_node *root = build_local_tree(&c, post, n, 0, 0, 0, 0, 0, 0);
	memset(stats, 0, sizeof(*stats));
	stats->total_leaves = count_leaves(root);
	const orc_node **leaves = (const orc_node **)malloc(sizeof(orc_node *) * (size_t)(stats->total_leaves + 1));
	const uint32_t total = bit(atoms);
	orc_acc st;
	for (uint32_t t = 0; t < total; t++) {
		apply_true_state(&c, root, t, 1.0);
		parse(&c, root, t, leaves, &st);
		/* simulation_helpers.cpp:11-23 */
		stats->unclassified_probability += st.unclas;
		stats->expected_stages += st.exp_stage;
		stats->expected_tests += st.exp_test;
		for (int i = 0; i <= depth * 1; i++) {
			stats->correct_probability += st.correct[i];
			stats->incorrect_probability += st.incorrect[i];
			stats->false_positive_probability += st.fp[i];
			stats->false_negative_probability += st.fn[i];
		}
	}
	free(leaves);
	free_tree(root);
	return 0;
}

/* ------------------------------------------------- no-shrink restatement */

typedef struct ns_node {
	uint32_t e, r, pos, neg;
	int stage, nchild;
	double *post;
	struct ns_node *child[1 << 5];
} ns_node;

typedef struct ns_cfg {
	int n, k, depth, dilution;
	double up, lo, branch, alpha, h;
} ns_cfg;

static double ns_dilution_probability(int pool, int missing, double alpha, double h)
{
	/* no_shrink_tree_reference_test.cpp:45-52 */
	if (missing == 0)
		return alpha;
	const int positives = pool - missing;
	return 1.0 - alpha * (double)missing / ((double)positives * h + (double)missing);
}

static double ns_response(const ns_cfg *c, uint32_t e, uint32_t r, uint32_t s)
{
	/* no_shrink_tree_reference_test.cpp:54-78 */
	double p = 1.0;
	const int pool = popc(e);
	for (int v = 0; v < c->k; v++) {
		const uint32_t vs = s >> (v * c->n);
		const int rb = (r & bit(v)) != 0;
		if (c->dilution) {
			const int missing = pool - popc(e & vs);
			const double pr = ns_dilution_probability(pool, missing, c->alpha, c->h);
			p *= rb ? pr : 1.0 - pr;
		} else {
			const int all = (e & vs) == e;
			p *= all ? (rb ? ORC_NEGATIVE_RESPONSE : 1.0 - ORC_NEGATIVE_RESPONSE)
					 : (rb ? 1.0 - ORC_NEGATIVE_RESPONSE : ORC_NEGATIVE_RESPONSE);
		}
	}
	return p;
}

static void ns_build(const ns_cfg *c, ns_node *node)
{
	/* no_shrink_tree_reference_test.cpp:219-242 */
	const int atoms = c->n * c->k;
	if (popc(node->pos | node->neg) == atoms || node->stage >= c->depth)
		return;
	const uint32_t cls = node->pos | node->neg;
	int active[32], na = 0;
	for (int s = 0; s < c->n; s++) {
		int all = 1;
		for (int v = 0; v < c->k; v++)
			if (!(cls & bit(v * c->n + s)))
				all = 0;
		if (!all)
			active[na++] = s;
	}
	const uint32_t states = bit(atoms);
	double *proj = (double *)calloc(bit(na * c->k), sizeof(double));
	for (uint32_t s = 0; s < states; s++) {
		uint32_t cpt = 0;
		for (int v = 0; v < c->k; v++)
			for (int i = 0; i < na; i++)
				if (s & bit(v * c->n + active[i]))
					cpt |= bit(v * na + i);
		proj[cpt] += node->post[s];
	}
	/* select_compact_experiment (lines 169-191) */
	const uint32_t ec = bit(na), rc = bit(c->k);
	const double target = 1.0 / (double)rc;
	uint32_t best_e = 0;
	double best = 2.0;
	double mass[32];
	for (uint32_t e = 0; e < ec; e++) {
		for (uint32_t i = 0; i < rc; i++)
			mass[i] = 0.0;
		for (uint32_t s = 0; s < bit(na * c->k); s++) {
			uint32_t id = 0;
			for (int v = 0; v < c->k; v++) {
				const uint32_t vs = s >> (v * na);
				if ((e & vs) != e)
					id |= bit(v);
			}
			mass[id] += proj[s];
		}
		double score = 0.0;
		for (uint32_t i = 0; i < rc; i++)
			score += fabs(mass[i] - target);
		if (score < best) {
			best = score;
			best_e = e;
		}
	}
	free(proj);
	uint32_t full = 0;
	for (int i = 0; i < na; i++)
		if (best_e & bit(i))
			full |= bit(active[i]);
	node->nchild = (int)rc;
	for (uint32_t r = 0; r < rc; r++) {
		ns_node *ch = (ns_node *)calloc(1, sizeof(ns_node));
		ch->e = full;
		ch->r = r;
		ch->stage = node->stage + 1;
		ch->post = (double *)malloc(sizeof(double) * states);
		double den = 0.0;
		for (uint32_t s = 0; s < states; s++) {
			ch->post[s] = node->post[s] * ns_response(c, full, r, s);
			den += ch->post[s];
		}
		for (uint32_t s = 0; s < states; s++)
			ch->post[s] /= den;
		ch->pos = node->pos;
		ch->neg = node->neg;
		/* update_classification (lines 118-132) */
		const uint32_t cl = ch->pos | ch->neg;
		for (int a = 0; a < atoms; a++) {
			if (cl & bit(a))
				continue;
			double m = 0.0;
			for (uint32_t s = 0; s < states; s++)
				if (s & bit(a))
					m += ch->post[s];
			if (m < c->lo)
				ch->pos |= bit(a);
			else if (m > 1.0 - c->up)
				ch->neg |= bit(a);
		}
		ns_build(c, ch);
		node->child[r] = ch;
	}
}

static int ns_leaves(const ns_node *n)
{
	if (n->nchild == 0)
		return 1;
	int t = 0;
	for (int i = 0; i < n->nchild; i++)
		t += ns_leaves(n->child[i]);
	return t;
}

static void ns_acc(const ns_cfg *c, const ns_node *node, uint32_t t, double bp, double prior, orc_tree_stats *st)
{
	/* no_shrink_tree_reference_test.cpp:254-293 */
	if (bp < c->branch)
		return;
	const int atoms = c->n * c->k;
	if (node->nchild == 0) {
		const double w = bp * prior;
		const int classified = popc(node->pos | node->neg) == atoms;
		if (classified && node->neg == t) {
			st->correct_probability += w;
		} else if (classified) {
			st->incorrect_probability += w;
			const double tp = (double)popc(node->pos), tn = (double)popc(node->neg);
			if (tp != 0.0)
				st->false_positive_probability += (double)popc((node->neg ^ t) & t) / tp * w;
			if (tn != 0.0)
				st->false_negative_probability += (double)popc((node->neg ^ t) & ~t) / tn * w;
		} else {
			st->unclassified_probability += w;
		}
		st->expected_stages += (double)node->stage * w;
		st->expected_tests += (double)node->stage * w;
		return;
	}
	for (int i = 0; i < node->nchild; i++) {
		const ns_node *ch = node->child[i];
		ns_acc(c, ch, t, bp * ns_response(c, ch->e, ch->r, t), prior, st);
	}
}

static void ns_free(ns_node *n)
{
	for (int i = 0; i < n->nchild; i++)
		ns_free(n->child[i]);
	free(n->post);
	free(n);
}

int orc_run_simulation_noshrink(int dilution, int n, int k, const double *prior, int depth,
								double up, double lo, double branch, double alpha, double h,
								orc_tree_stats *stats)
{
	/* no_shrink_tree_reference_test.cpp:295-311 */
	if (n <= 0 || k <= 0 || n * k > 24 || k > 5)
		return 1;
	ns_cfg c = {n, k, depth, dilution, up, lo, branch, alpha, h};
	const int atoms = n * k;
	ns_node *root = (ns_node *)calloc(1, sizeof(ns_node));
	root->post = (double *)malloc(sizeof(double) * bit(atoms));
	for (uint32_t s = 0; s < bit(atoms); s++)
		root->post[s] = orc_prior_prob(s, atoms, prior);
	ns_build(&c, root);
	memset(stats, 0, sizeof(*stats));
	stats->total_leaves = ns_leaves(root);
	for (uint32_t s = 0; s < bit(atoms); s++)
		ns_acc(&c, root, s, 1.0, orc_prior_prob(s, atoms, prior), stats);
	ns_free(root);
	return 0;
}
