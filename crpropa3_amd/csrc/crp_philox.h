/*
 * Philox4x32-10 counter-based random numbers (J. K. Salmon, M. A. Moraes,
 * R. O. Dror, D. E. Shaw, "Parallel random numbers: as easy as 1, 2, 3",
 * SC'11): ten rounds of two 32x32->64 multiplications and a key schedule
 * bumped by the golden-ratio Weyl constants.  Stateless: the output depends
 * only on (counter, key), so particle i at step k can draw its numbers without
 * any sequence shared with other particles.
 *
 * The same source compiles as C11 (host; the oracle) and as HIP device code.
 */
#ifndef CRP_PHILOX_H
#define CRP_PHILOX_H

#include <stdint.h>

#if defined(__HIPCC__)
#define CRP_PHILOX_FN static inline __host__ __device__
#else
#define CRP_PHILOX_FN static inline
#endif

typedef struct {
	uint32_t v[4];
} crp_philox_ctr;

CRP_PHILOX_FN uint32_t crp_mulhilo32(uint32_t a, uint32_t b, uint32_t *hi) {
	uint64_t p = (uint64_t)a * (uint64_t)b;
	*hi = (uint32_t)(p >> 32);
	return (uint32_t)p;
}

CRP_PHILOX_FN crp_philox_ctr crp_philox4x32_10(crp_philox_ctr c, uint32_t k0, uint32_t k1) {
	for (int round = 0; round < 10; round++) {
		uint32_t hi0, hi1;
		uint32_t lo0 = crp_mulhilo32(0xD2511F53u, c.v[0], &hi0);
		uint32_t lo1 = crp_mulhilo32(0xCD9E8D57u, c.v[2], &hi1);
		crp_philox_ctr n;
		n.v[0] = hi1 ^ c.v[1] ^ k0;
		n.v[1] = lo1;
		n.v[2] = hi0 ^ c.v[3] ^ k1;
		n.v[3] = lo0;
		c = n;
		k0 += 0x9E3779B9u;
		k1 += 0xBB67AE85u;
	}
	return c;
}

/*
 * The uniform stream of one particle in one step: draw d uses block d/2 of the
 * counter (stream low word, step, block, stream high word) with key
 * (seed, 0x43525053 'CRPS'); each block yields two doubles with 53 random bits,
 * in [0, 1).  `stream` names the particle (its index, or a serial number that
 * stays with it from call to call).
 */
typedef struct {
	uint32_t seed, particle, step, next, streamHi;
	double cache;
	int cached;
} crp_sde_rng;

CRP_PHILOX_FN void crp_sde_rng_init(crp_sde_rng *r, uint32_t seed, uint64_t stream, uint32_t step) {
	r->seed = seed;
	r->particle = (uint32_t)stream;
	r->streamHi = (uint32_t)(stream >> 32);
	r->step = step;
	r->next = 0;
	r->cached = 0;
	r->cache = 0;
}

CRP_PHILOX_FN double crp_sde_uniform(crp_sde_rng *r) {
	if (r->cached) {
		r->cached = 0;
		return r->cache;
	}
	crp_philox_ctr c;
	c.v[0] = r->particle;
	c.v[1] = r->step;
	c.v[2] = r->next++;
	c.v[3] = r->streamHi;
	crp_philox_ctr o = crp_philox4x32_10(c, r->seed, 0x43525053u);
	const double s = 1.0 / 9007199254740992.0; /* 2^-53 */
	double u0 = (double)(((uint64_t)o.v[0] << 21) | (o.v[1] >> 11)) * s;
	double u1 = (double)(((uint64_t)o.v[2] << 21) | (o.v[3] >> 11)) * s;
	r->cache = u1;
	r->cached = 1;
	return u0;
}

#endif
