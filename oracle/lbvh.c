/* Oracle: LBVH stages (Morton codes, binary radix tree). The reference has no
 * GPU BVH; its Morton bit order (x at bit 3i, y at 3i+1, z at 3i+2) is taken
 * from src/Morton.cpp:37-44,68-70. The radix tree is built here top-down --
 * a different construction from the GPU's bottom-up Karras (2012) kernel --
 * and numbered with Karras's convention, so both must agree node for node.
 *
 * TEST INFRASTRUCTURE ONLY -- see oracle.h. */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>

static uint32_t spread10(uint32_t x) { /* bit i -> bit 3i for the low 10 bits */
	uint32_t r = 0;
	for (int i = 0; i < 10; i++)
		r |= ((x >> i) & 1u) << (3 * i);
	return r;
}

uint32_t or_morton3(uint32_t x, uint32_t y, uint32_t z) {
	return spread10(x) | (spread10(y) << 1) | (spread10(z) << 2);
}

/* Centroid c = ((v0 + v1) + v2) / 3; per axis q = (uint)min(max((c - cmin) *
 * (1024 / ext), 0), 1023) with ext = cmax - cmin over all centroids
 * (scale 0 when ext == 0). */
void or_morton_codes(const float *verts, int n, uint32_t *codes) {
	float *c = (float *)malloc(sizeof(float) * 3 * (size_t)(n > 0 ? n : 1));
	float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
	for (int i = 0; i < n; i++) {
		for (int a = 0; a < 3; a++) {
			float v = (verts[9 * i + a] + verts[9 * i + 3 + a] + verts[9 * i + 6 + a]) / 3.0f;
			c[3 * i + a] = v;
			if (v < lo[a])
				lo[a] = v;
			if (v > hi[a])
				hi[a] = v;
		}
	}
	float scale[3];
	for (int a = 0; a < 3; a++) {
		float ext = hi[a] - lo[a];
		scale[a] = ext > 0.0f ? 1024.0f / ext : 0.0f;
	}
	for (int i = 0; i < n; i++) {
		uint32_t q[3];
		for (int a = 0; a < 3; a++) {
			float f = (c[3 * i + a] - lo[a]) * scale[a];
			f = fminf(fmaxf(f, 0.0f), 1023.0f);
			q[a] = (uint32_t)f;
		}
		codes[i] = or_morton3(q[0], q[1], q[2]);
	}
	free(c);
}

static int clz64(uint64_t x) {
	return x ? __builtin_clzll(x) : 64;
}

static uint64_t akey(const uint32_t *codes, int i) { /* index-augmented key */
	return ((uint64_t)codes[i] << 32) | (uint32_t)i;
}

static void split_rec(const uint32_t *codes, int first, int last, int id, int *left, int *right) {
	uint64_t kf = akey(codes, first);
	int common = clz64(kf ^ akey(codes, last));
	/* Largest g in [first, last-1] whose key shares more than `common` leading
	 * bits with the first key (linear scan: the oracle favours obviousness). */
	int g = first;
	for (int i = first; i < last; i++)
		if (clz64(kf ^ akey(codes, i)) > common)
			g = i;
	int lid, rid;
	if (g == first)
		lid = ~g;
	else
		lid = g;
	if (g + 1 == last)
		rid = ~(g + 1);
	else
		rid = g + 1;
	left[id] = lid;
	right[id] = rid;
	if (lid >= 0)
		split_rec(codes, first, g, lid, left, right);
	if (rid >= 0)
		split_rec(codes, g + 1, last, rid, left, right);
}

void or_radix_tree(const uint32_t *sorted_codes, int n, int *left, int *right) {
	if (n < 2)
		return;
	split_rec(sorted_codes, 0, n - 1, 0, left, right);
}
