/* Philox4x32-10 counter-based generator -- oracle restatement.
 *
 * Algorithm: Salmon et al., "Parallel Random Numbers: As Easy as 1, 2, 3"
 * (SC'11), section 3: ten rounds of two 32x32->64 multiplies with the round
 * constants M0 = 0xD2511F53, M1 = 0xCD9E8D57 and Weyl key increments
 * W0 = 0x9E3779B9, W1 = 0xBB67AE85. The reference project has no RNG of this
 * kind (it uses std::mt19937, src/Renderer.cpp:1139); this is the generator the
 * GPU kernels use, restated independently of them. */
#include "oracle.h"

static void round_once(uint32_t c[4], const uint32_t k[2]) {
	uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
	uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
	uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
	uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
	uint32_t n0 = hi1 ^ c[1] ^ k[0];
	uint32_t n2 = hi0 ^ c[3] ^ k[1];
	c[0] = n0;
	c[1] = lo1;
	c[2] = n2;
	c[3] = lo0;
}

void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
	uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
	uint32_t k[2] = {key[0], key[1]};
	for (int r = 0; r < 10; r++) {
		if (r > 0) {
			k[0] += 0x9E3779B9u;
			k[1] += 0xBB67AE85u;
		}
		round_once(c, k);
	}
	out[0] = c[0];
	out[1] = c[1];
	out[2] = c[2];
	out[3] = c[3];
}

void or_philox_batch(int n, const uint32_t *ctr, const uint32_t *key, uint32_t *out) {
	for (int i = 0; i < n; i++)
		or_philox4x32_10(ctr + 4 * i, key + 2 * i, out + 4 * i);
}
