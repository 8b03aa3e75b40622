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

/* PLOC (Meister & Bittner 2018, "Parallel Locally-Ordered Clustering for
 * Bounding Volume Hierarchy Construction") over the sorted order, written as
 * a plain sequential loop over rounds. The reference has no such builder; the
 * rules restated here are the GPU builder's own:
 *  - a cluster's box is the min/max of its triangles' vertices;
 *  - its nearest neighbour is the cluster within OR_PLOC_R positions whose
 *    merged box has the least half-area dx*dy + dy*dz + dz*dx (evaluated as
 *    fmaf(dz, dx, fmaf(dx, dy, dy * dz))), the first in increasing position
 *    on equal area;
 *  - a mutual pair merges; its smaller position owns the new node, whose left
 *    child is the owner's reference, and the new cluster takes the owner's
 *    place in the next round's order, other clusters keeping theirs;
 *  - new internal ids count down from n - 2: a round's owners, in increasing
 *    position, take id_top, id_top - 1, ...
 * Returns the number of rounds, or -1 when a round merges nothing. */
#define OR_PLOC_R 8

static float ploc_area(const float *a, const float *b) {
	float dx = fmaxf(a[3], b[3]) - fminf(a[0], b[0]);
	float dy = fmaxf(a[4], b[4]) - fminf(a[1], b[1]);
	float dz = fmaxf(a[5], b[5]) - fminf(a[2], b[2]);
	return fmaf(dz, dx, fmaf(dx, dy, dy * dz));
}

int or_ploc_tree(const float *verts, const int *order, int n, int *left, int *right) {
	if (n < 2)
		return 0;
	int *ref = (int *)malloc(sizeof(int) * 2 * (size_t)n);
	float *box = (float *)malloc(sizeof(float) * 12 * (size_t)n); /* lo xyz, hi xyz */
	int *nn = (int *)malloc(sizeof(int) * (size_t)n);
	int *ref2 = ref + n;
	float *box2 = box + 6 * (size_t)n;
	for (int p = 0; p < n; p++) {
		const float *t = verts + 9 * (size_t)order[p];
		for (int a = 0; a < 3; a++) {
			box[6 * p + a] = fminf(fminf(t[a], t[3 + a]), t[6 + a]);
			box[6 * p + 3 + a] = fmaxf(fmaxf(t[a], t[3 + a]), t[6 + a]);
		}
		ref[p] = ~p;
	}
	int m = n, id_top = n - 2, rounds = 0;
	while (m > 1) {
		for (int i = 0; i < m; i++) {
			float best = INFINITY;
			int bj = -1;
			for (int j = i - OR_PLOC_R; j <= i + OR_PLOC_R; j++) {
				if (j < 0 || j >= m || j == i)
					continue;
				float a = ploc_area(box + 6 * i, box + 6 * j);
				/* (area, parity of the pair's smaller position, smaller
				 * position, larger position): candidates come in increasing
				 * j, so the last two keys increase along the loop */
				int par = (j < i ? j : i) & 1;
				int bpar = bj < 0 ? 2 : (bj < i ? bj : i) & 1;
				if (a < best || (a == best && par < bpar)) {
					best = a;
					bj = j;
				}
			}
			nn[i] = bj;
		}
		int merges = 0, kept = 0;
		for (int i = 0; i < m; i++) {
			int j = nn[i];
			int mutual = j >= 0 && nn[j] == i;
			if (mutual && j < i)
				continue; /* merged into its owner j */
			float *nb = box2 + 6 * kept;
			if (mutual) {
				int id = id_top - merges;
				left[id] = ref[i];
				right[id] = ref[j];
				for (int a = 0; a < 3; a++) {
					nb[a] = fminf(box[6 * i + a], box[6 * j + a]);
					nb[3 + a] = fmaxf(box[6 * i + 3 + a], box[6 * j + 3 + a]);
				}
				ref2[kept] = id;
				merges++;
			} else {
				for (int a = 0; a < 6; a++)
					nb[a] = box[6 * i + a];
				ref2[kept] = ref[i];
			}
			kept++;
		}
		rounds++;
		if (merges == 0) {
			rounds = -1;
			break;
		}
		id_top -= merges;
		m = kept;
		int *tr = ref; ref = ref2; ref2 = tr;
		float *tb = box; box = box2; box2 = tb;
	}
	free(ref < ref2 ? ref : ref2);
	free(box < box2 ? box : box2);
	free(nn);
	return rounds;
}
