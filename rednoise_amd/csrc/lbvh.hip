// GPU LBVH build: centroid bounds -> 30-bit Morton codes -> stable LSD radix
// sort (wave64 ballot ranking) -> Karras (HPG 2012) binary radix tree ->
// bottom-up AABB refit (LDS inside 1024-leaf chunks, agent-scope atomics above) -> top-down collapse into the
// compressed 8-wide BVH of rn_types.h (greedy largest-area opening until nodes are
// full, leaves of <= 3 triangles, octant-ordered child slots, 8-bit
// conservative boxes).
#include "rn_device.h"
#include "rn_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>

namespace rn {

namespace {

#ifndef RN_LEAF_MAX
#define RN_LEAF_MAX 3
#endif
constexpr int kLeafMax = RN_LEAF_MAX; // leaf triangles at most (a node holds three tmask bits per slot)
static_assert(kLeafMax >= 1 && kLeafMax <= 3, "a slot's tmask bits hold at most 3 triangles");
constexpr int kBlock = 256;

// Total order on floats as unsigned ints (for atomic min/max of bounds).
__device__ __forceinline__ uint32_t ford(float f) {
	uint32_t u = __float_as_uint(f);
	return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float fordinv(uint32_t u) {
	return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

__device__ __forceinline__ float3 centroid(const float *v, int i) {
#pragma clang fp contract(off)
	const float *t = v + 9 * (size_t)i;
	return f3((t[0] + t[3] + t[6]) / 3.0f, (t[1] + t[4] + t[7]) / 3.0f, (t[2] + t[5] + t[8]) / 3.0f);
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, size_t n) {
	size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	for (; i < n; i += (size_t)gridDim.x * blockDim.x)
		p[i] = v;
}

// Bounds words are kBoundStride apart (one 128-byte line each): atomics on
// one line serialise, so six bounds in one line would queue every block's
// updates behind each other.
constexpr int kBoundStride = 32;

__global__ void k_centroid_bounds(const float *v, int n, uint32_t *b) {
	__shared__ float red[6][kBlock / 64];
	float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
		float3 c = centroid(v, i);
		lo[0] = fminf(lo[0], c.x);
		lo[1] = fminf(lo[1], c.y);
		lo[2] = fminf(lo[2], c.z);
		hi[0] = fmaxf(hi[0], c.x);
		hi[1] = fmaxf(hi[1], c.y);
		hi[2] = fmaxf(hi[2], c.z);
	}
#pragma unroll
	for (int k = 0; k < 3; k++) {
		for (int off = 32; off > 0; off >>= 1) {
			lo[k] = fminf(lo[k], __shfl_xor(lo[k], off));
			hi[k] = fmaxf(hi[k], __shfl_xor(hi[k], off));
		}
	}
	const int wave = threadIdx.x >> 6;
	if ((threadIdx.x & 63) == 0) {
		for (int k = 0; k < 3; k++) {
			red[k][wave] = lo[k];
			red[3 + k][wave] = hi[k];
		}
	}
	__syncthreads();
	if (threadIdx.x < 6) { // one atomic per block and bound
		const int k = threadIdx.x;
		float r = red[k][0];
		for (int w = 1; w < kBlock / 64; w++)
			r = k < 3 ? fminf(r, red[k][w]) : fmaxf(r, red[k][w]);
		if (k < 3)
			atomicMin(&b[k * kBoundStride], ford(r));
		else
			atomicMax(&b[k * kBoundStride], ford(r));
	}
}

__device__ __forceinline__ uint32_t spread10(uint32_t x) { // bit i -> bit 3i
	x &= 0x3FFu;
	x = (x | (x << 16)) & 0x030000FFu;
	x = (x | (x << 8)) & 0x0300F00Fu;
	x = (x | (x << 4)) & 0x030C30C3u;
	x = (x | (x << 2)) & 0x09249249u;
	return x;
}

__global__ void k_morton(const float *v, int n, const uint32_t *b, uint32_t *keys, uint32_t *vals) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	float3 c = centroid(v, i);
	uint32_t q[3];
#pragma unroll
	for (int k = 0; k < 3; k++) {
		float lo = fordinv(b[k * kBoundStride]), hi = fordinv(b[(3 + k) * kBoundStride]);
		float ext = hi - lo;
		float scale = ext > 0.0f ? 1024.0f / ext : 0.0f;
		float f = (get3(c, k) - lo) * scale;
		f = fminf(fmaxf(f, 0.0f), 1023.0f);
		q[k] = (uint32_t)f;
	}
	keys[i] = spread10(q[0]) | (spread10(q[1]) << 1) | (spread10(q[2]) << 2);
	vals[i] = (uint32_t)i;
}

// ------------------------------------------------------------ radix sort --

constexpr int kRsItems = 8;
constexpr int kRsTile = kBlock * kRsItems;

__global__ void k_rs_hist(const uint32_t *keys, int n, int shift, uint32_t *hist, int nblocks) {
	__shared__ uint32_t h[256];
	h[threadIdx.x] = 0;
	__syncthreads();
	int base = blockIdx.x * kRsTile;
#pragma unroll
	for (int r = 0; r < kRsItems; r++) {
		int p = base + r * kBlock + threadIdx.x;
		if (p < n)
			atomicAdd(&h[(keys[p] >> shift) & 0xFFu], 1u);
	}
	__syncthreads();
	hist[(size_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter. Tile position p = base + r*256 + tid; within a round the
// lanes of a wave are ranked by an 8-ballot digit match, waves are ordered by
// per-digit prefix sums in LDS, rounds by a running per-digit count.
__global__ void k_rs_scatter(const uint32_t *keys, const uint32_t *vals, int n, int shift, const uint32_t *offs,
                             int nblocks, uint32_t *keys_out, uint32_t *vals_out) {
	__shared__ uint32_t run[256];
	__shared__ uint32_t wc[4][256];
	__shared__ uint32_t wb[4][256];
	const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
	run[tid] = 0;
	for (int w = 0; w < 4; w++)
		wc[w][tid] = 0;
	__syncthreads();
	const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
	const int base = blockIdx.x * kRsTile;
	const uint32_t gbase = offs[(size_t)tid * nblocks + blockIdx.x];
	__shared__ uint32_t gb[256];
	gb[tid] = gbase;
	for (int r = 0; r < kRsItems; r++) {
		int p = base + r * kBlock + tid;
		bool valid = p < n;
		uint32_t key = valid ? keys[p] : 0u, val = valid ? vals[p] : 0u;
		uint32_t digit = (key >> shift) & 0xFFu;
		uint64_t m = __ballot(valid);
#pragma unroll
		for (int b = 0; b < 8; b++) {
			uint64_t bb = __ballot((digit >> b) & 1u);
			m &= ((digit >> b) & 1u) ? bb : ~bb;
		}
		uint32_t rank = (uint32_t)__popcll(m & lt);
		uint32_t cnt = (uint32_t)__popcll(m);
		if (valid && rank == cnt - 1)
			wc[wave][digit] = cnt;
		__syncthreads();
		uint32_t acc = run[tid];
#pragma unroll
		for (int w = 0; w < 4; w++) {
			uint32_t c = wc[w][tid];
			wb[w][tid] = acc;
			acc += c;
			wc[w][tid] = 0;
		}
		run[tid] = acc;
		__syncthreads();
		if (valid) {
			uint32_t dst = gb[digit] + wb[wave][digit] + rank;
			keys_out[dst] = key;
			vals_out[dst] = val;
		}
	}
}

// Exclusive scan of uint32 (1024 elements per block, recursive over blocks).
constexpr int kScanItems = 4;
constexpr int kScanTile = kBlock * kScanItems;

__global__ void k_scan_tile(const uint32_t *in, uint32_t *out, size_t n, uint32_t *sums) {
	__shared__ uint32_t s[kBlock];
	size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
	uint32_t v[kScanItems], tot = 0;
#pragma unroll
	for (int k = 0; k < kScanItems; k++) {
		v[k] = (base + k < n) ? in[base + k] : 0u;
		tot += v[k];
	}
	s[threadIdx.x] = tot;
	__syncthreads();
	for (int off = 1; off < kBlock; off <<= 1) {
		uint32_t x = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0u;
		__syncthreads();
		s[threadIdx.x] += x;
		__syncthreads();
	}
	uint32_t excl = s[threadIdx.x] - tot;
#pragma unroll
	for (int k = 0; k < kScanItems; k++) {
		if (base + k < n)
			out[base + k] = excl;
		excl += v[k];
	}
	if (threadIdx.x == kBlock - 1 && sums)
		sums[blockIdx.x] = s[kBlock - 1];
}

__global__ void k_scan_add(uint32_t *out, size_t n, const uint32_t *sums) {
	size_t i = (size_t)blockIdx.x * kScanTile + threadIdx.x;
	uint32_t add = sums[blockIdx.x];
	for (int k = 0; k < kScanItems; k++, i += kBlock)
		if (i < n)
			out[i] += add;
}

void exclusive_scan(const uint32_t *in, uint32_t *out, size_t n, uint32_t *tmp, hipStream_t s) {
	size_t blocks = (n + kScanTile - 1) / kScanTile;
	if (blocks <= 1) {
		k_scan_tile<<<1, kBlock, 0, s>>>(in, out, n, nullptr);
		return;
	}
	uint32_t *sums = tmp, *sums_scan = tmp + blocks, *rest = tmp + 2 * blocks;
	k_scan_tile<<<(unsigned)blocks, kBlock, 0, s>>>(in, out, n, sums);
	exclusive_scan(sums, sums_scan, blocks, rest, s);
	k_scan_add<<<(unsigned)blocks, kBlock, 0, s>>>(out, n, sums_scan);
}

size_t scan_tmp_words(size_t n) {
	size_t total = 0;
	while (true) {
		size_t blocks = (n + kScanTile - 1) / kScanTile;
		if (blocks <= 1)
			break;
		total += 2 * blocks;
		n = blocks;
	}
	return total + 16;
}

// ---------------------------------------------------------------- Karras --

__device__ __forceinline__ int delta(const uint32_t *k, int n, int i, int j) {
	if (j < 0 || j >= n)
		return -1;
	uint32_t a = k[i], b = k[j];
	if (a == b)
		return 32 + __clz((uint32_t)(i ^ j));
	return __clz(a ^ b);
}

// Internal node i of n-1; parent[] holds internal nodes at [0, n-1) and leaf
// p at n-1+p; children >= 0 are internal, < 0 are leaves ~p.
__global__ void k_karras(const uint32_t *k, int n, int *left, int *right, int *parent, int2 *range, uint32_t *count) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n - 1)
		return;
	int d = (delta(k, n, i, i + 1) - delta(k, n, i, i - 1)) >= 0 ? 1 : -1;
	int dmin = delta(k, n, i, i - d);
	int lmax = 2;
	while (delta(k, n, i, i + lmax * d) > dmin)
		lmax <<= 1;
	int l = 0;
	for (int t = lmax >> 1; t >= 1; t >>= 1)
		if (delta(k, n, i, i + (l + t) * d) > dmin)
			l += t;
	int j = i + l * d;
	int dnode = delta(k, n, i, j);
	int s = 0, t = l;
	do {
		t = (t + 1) >> 1;
		if (delta(k, n, i, i + (s + t) * d) > dnode)
			s += t;
	} while (t > 1);
	int g = i + s * d + (d < 0 ? -1 : 0);
	int lo = min(i, j), hi = max(i, j);
	int lc = (lo == g) ? ~g : g;
	int rc = (hi == g + 1) ? ~(g + 1) : g + 1;
	left[i] = lc;
	right[i] = rc;
	parent[lc >= 0 ? lc : n - 1 + ~lc] = i;
	parent[rc >= 0 ? rc : n - 1 + ~rc] = i;
	range[i] = make_int2(lo, hi);
	count[i] = (uint32_t)(hi - lo + 1);
	if (i == 0)
		parent[0] = -1;
}

// Bottom-up refit. Block b owns leaves [b*kRefitChunk, (b+1)*kRefitChunk).
// An internal node whose leaf range lies inside one chunk is "local": Karras
// node i covers a range that contains i, so it sits at LDS slot i - base, and
// its children are local too. Local nodes are refitted in LDS with
// workgroup-scope counters (the second child to arrive computes the parent).
// Only nodes whose range crosses a chunk boundary (about two per chunk) use
// agent-scope acquire-release counters in global memory; those are costly on
// MI355X because agent scope spans eight separately cached XCDs.
constexpr int kRefitChunk = 1024;

__global__ __launch_bounds__(kRefitChunk) void k_refit(const float *v, const uint32_t *order, int n,
                                                       const int *parent, const int *left, const int *right,
                                                       const int2 *range, float4 *box, uint32_t *visit) {
	__shared__ uint32_t cnt[kRefitChunk];
	__shared__ float4 lbox[2 * kRefitChunk]; // local internal node boxes (lo, hi)
	__shared__ float4 leafb[2 * kRefitChunk];
	const int base = blockIdx.x * kRefitChunk;
	const int end = min(base + kRefitChunk, n);
	const int p = base + (int)threadIdx.x;
	cnt[threadIdx.x] = 0u;
	float4 lo = make_float4(0.f, 0.f, 0.f, 0.f), hi = lo;
	if (p < n) {
		const float *t = v + 9 * (size_t)order[p];
		lo = make_float4(fminf(fminf(t[0], t[3]), t[6]), fminf(fminf(t[1], t[4]), t[7]),
		                 fminf(fminf(t[2], t[5]), t[8]), 0.0f);
		hi = make_float4(fmaxf(fmaxf(t[0], t[3]), t[6]), fmaxf(fmaxf(t[1], t[4]), t[7]),
		                 fmaxf(fmaxf(t[2], t[5]), t[8]), 0.0f);
		size_t li = (size_t)(n - 1 + p);
		box[2 * li] = lo;
		box[2 * li + 1] = hi;
		leafb[2 * threadIdx.x] = lo;
		leafb[2 * threadIdx.x + 1] = hi;
	}
	__syncthreads();
	if (p >= n || n < 2)
		return;
	int node = parent[n - 1 + p];
	// local part
	while (node >= 0) {
		const int2 rg = range[node];
		if (rg.x < base || rg.y >= end)
			break;
		const int slot = node - base;
		uint32_t old = __hip_atomic_fetch_add(&cnt[slot], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
		if (old == 0)
			return;
		const int l = left[node], r = right[node];
		const float4 *ca = l >= 0 ? &lbox[2 * (l - base)] : &leafb[2 * (~l - base)];
		const float4 *cb = r >= 0 ? &lbox[2 * (r - base)] : &leafb[2 * (~r - base)];
		const float4 al = ca[0], ah = ca[1], bl = cb[0], bh = cb[1];
		lo = make_float4(fminf(al.x, bl.x), fminf(al.y, bl.y), fminf(al.z, bl.z), 0.0f);
		hi = make_float4(fmaxf(ah.x, bh.x), fmaxf(ah.y, bh.y), fmaxf(ah.z, bh.z), 0.0f);
		lbox[2 * slot] = lo;
		lbox[2 * slot + 1] = hi;
		box[2 * (size_t)node] = lo;
		box[2 * (size_t)node + 1] = hi;
		node = parent[node];
	}
	// nodes spanning chunks: the box just written is published by the
	// release half of the agent-scope counter update
	while (node >= 0) {
		uint32_t old = __hip_atomic_fetch_add(&visit[node], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
		if (old == 0)
			return;
		int l = left[node], r = right[node];
		size_t a = l >= 0 ? (size_t)l : (size_t)(n - 1 + ~l);
		size_t b = r >= 0 ? (size_t)r : (size_t)(n - 1 + ~r);
		float4 al = box[2 * a], ah = box[2 * a + 1], bl = box[2 * b], bh = box[2 * b + 1];
		lo = make_float4(fminf(al.x, bl.x), fminf(al.y, bl.y), fminf(al.z, bl.z), 0.0f);
		hi = make_float4(fmaxf(ah.x, bh.x), fmaxf(ah.y, bh.y), fmaxf(ah.z, bh.z), 0.0f);
		box[2 * (size_t)node] = lo;
		box[2 * (size_t)node + 1] = hi;
		node = parent[node];
	}
}

// ------------------------------------------------------------------ PLOC --
// Parallel locally-ordered clustering (Meister and Bittner, "Parallel
// Locally-Ordered Clustering for Bounding Volume Hierarchy Construction",
// IEEE TVCG 2018): clusters start as the Morton-sorted triangles; each round
// every cluster finds the neighbour within kPlocR positions whose merged box
// has the least surface area, mutual nearest neighbours merge, and the
// survivors are compacted in order. The result has the node layout of the
// Karras tree (internal ids 0..n-2 with root 0, leaves ~position, boxes at
// 2*id and 2*(n-1+position), triangle counts), so the collapse takes either.

#ifndef RN_PLOC_R
#define RN_PLOC_R 8
#endif
constexpr int kPlocR = RN_PLOC_R; // PLOC search radius (clusters on each side)

__global__ void k_leaf_boxes(const float *v, const uint32_t *order, int n, float4 *box, int *cref, float4 *cbox,
                             uint32_t *ccnt) {
	int p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= n)
		return;
	const float *t = v + 9 * (size_t)order[p];
	float4 lo = make_float4(fminf(fminf(t[0], t[3]), t[6]), fminf(fminf(t[1], t[4]), t[7]),
	                        fminf(fminf(t[2], t[5]), t[8]), 0.0f);
	float4 hi = make_float4(fmaxf(fmaxf(t[0], t[3]), t[6]), fmaxf(fmaxf(t[1], t[4]), t[7]),
	                        fmaxf(fmaxf(t[2], t[5]), t[8]), 0.0f);
	size_t li = (size_t)(n - 1 + p);
	box[2 * li] = lo;
	box[2 * li + 1] = hi;
	cref[p] = ~p;
	cbox[2 * (size_t)p] = lo;
	cbox[2 * (size_t)p + 1] = hi;
	ccnt[p] = 1u;
}

__device__ __forceinline__ float merged_area(float4 alo, float4 ahi, float4 blo, float4 bhi) {
	float dx = fmaxf(ahi.x, bhi.x) - fminf(alo.x, blo.x);
	float dy = fmaxf(ahi.y, bhi.y) - fminf(alo.y, blo.y);
	float dz = fmaxf(ahi.z, bhi.z) - fminf(alo.z, blo.z);
	return fmaf(dz, dx, fmaf(dx, dy, dy * dz)); // fixed evaluation order: the tree is bit-reproducible
}

// Nearest neighbour of every cluster. Pairs are ordered by (merged area,
// parity of the smaller index, smaller index, larger index), a strict total
// order, so the least pair of the round is always mutual and every round
// merges at least once. The parity key matters on runs of equal areas (a
// strip of identical quads): with it every other pair of such a run is
// mutual, where ordering by index alone leaves only the run's first pair
// mutual and the run takes a round per merge. Candidates come in increasing
// j, so along the loop the last two keys increase and a candidate replaces
// the best on a smaller area, or on an equal one with a smaller parity.
__device__ __forceinline__ bool ploc_better(float a, int j, float best, int bj, int i) {
	const int par = (j < i ? j : i) & 1;
	const int bpar = bj < 0 ? 2 : ((bj < i ? bj : i) & 1);
	return a < best || (a == best && par < bpar);
}
__global__ __launch_bounds__(kBlock) void k_ploc_nn(const uint32_t *pm, const float4 *cbox, int *nn) {
	__shared__ float4 slo[kBlock + 2 * kPlocR], shi[kBlock + 2 * kPlocR];
	const int m = (int)*pm;
	if (m <= 1 || (int)(blockIdx.x * kBlock) >= m)
		return;
	const int base = blockIdx.x * kBlock - kPlocR;
	for (int k = threadIdx.x; k < kBlock + 2 * kPlocR; k += kBlock) {
		int g = base + k;
		if (g >= 0 && g < m) {
			slo[k] = cbox[2 * (size_t)g];
			shi[k] = cbox[2 * (size_t)g + 1];
		}
	}
	__syncthreads();
	const int i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= m)
		return;
	const int li = i - base;
	const float4 alo = slo[li], ahi = shi[li];
	float best = INFINITY;
	int bj = -1;
	const int j0 = max(0, i - kPlocR), j1 = min(m - 1, i + kPlocR);
	for (int j = j0; j <= j1; j++) {
		if (j == i)
			continue;
		float a = merged_area(alo, ahi, slo[j - base], shi[j - base]);
		if (ploc_better(a, j, best, bj, i)) {
			best = a;
			bj = j;
		}
	}
	nn[i] = bj;
}

// A cluster's flags: mf when it merges with its nearest neighbour and has
// the smaller index of the pair (it then owns the new node), kf when it
// keeps a place in the next round (every cluster except the larger index
// of a merging pair).
__device__ __forceinline__ void ploc_flags(const int *nn, int i, uint32_t &mf, uint32_t &kf) {
	const int j = nn[i];
	const bool mutual = j >= 0 && nn[j] == i;
	mf = (mutual && i < j) ? 1u : 0u;
	kf = (mutual && i > j) ? 0u : 1u;
}

// Round state on the device, so that rounds run without the host:
// tot[0] merges and tot[1] clusters kept by the last round, tot[2] set when a
// round with m > 1 merged nothing, tot[3] m, tot[4] id_top (the next
// internal id, counting down), tot[5] and tot[6] the m and id_top the
// current round's merge works from.
constexpr int kPlocMerges = 0, kPlocKept = 1, kPlocErr = 2, kPlocM = 3, kPlocTop = 4, kPlocM0 = 5, kPlocTop0 = 6;
constexpr int kPlocStateWords = 7;

// A round after the neighbour search is three kernels: the merge and keep
// counts of each tile of kPlocTile clusters, one work-group's scan of the
// tile counts (which also advances the round state), and the merge, which
// recomputes its tile's flags and their offsets. Each thread of the tile
// kernels takes kPlocTile / kBlock clusters of its tile, kBlock apart.
constexpr int kPlocTile = 4 * kBlock;
constexpr int kPlocPer = kPlocTile / kBlock;
constexpr int kPlocScanBlock = 1024;

// Exclusive scan of a pair of counts over a block of BS threads (a multiple
// of 64); *total gets the block's sums. `wsum` holds BS / 64 entries; a
// second call on the same array needs a barrier between the two.
template <int BS>
__device__ __forceinline__ uint2 block_scan2(uint2 v, uint2 *total, uint2 *wsum) {
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	uint2 x = v;
#pragma unroll
	for (int off = 1; off < 64; off <<= 1) {
		const uint32_t a = __shfl_up(x.x, off), b = __shfl_up(x.y, off);
		if (lane >= off) {
			x.x += a;
			x.y += b;
		}
	}
	if (lane == 63)
		wsum[w] = x;
	__syncthreads();
	uint2 pre = make_uint2(0u, 0u), tot = make_uint2(0u, 0u);
	for (int k = 0; k < BS / 64; k++) {
		const uint2 s = wsum[k];
		if (k < w) {
			pre.x += s.x;
			pre.y += s.y;
		}
		tot.x += s.x;
		tot.y += s.y;
	}
	*total = tot;
	return make_uint2(pre.x + x.x - v.x, pre.y + x.y - v.y);
}

// Per tile: (merges, kept) of its clusters below m.
__global__ __launch_bounds__(kBlock) void k_ploc_count(const uint32_t *tot, const int *nn, uint2 *tsum) {
	const int m = (int)tot[kPlocM];
	const int t0 = blockIdx.x * kPlocTile;
	if (m <= 1 || t0 >= m)
		return;
	uint32_t sm = 0, sk = 0;
#pragma unroll
	for (int r = 0; r < kPlocPer; r++) {
		const int i = t0 + r * kBlock + threadIdx.x;
		if (i < m) {
			uint32_t mf, kf;
			ploc_flags(nn, i, mf, kf);
			sm += mf;
			sk += kf;
		}
	}
	__shared__ uint2 wsum[kBlock / 64];
	uint2 total;
	block_scan2<kBlock>(make_uint2(sm, sk), &total, wsum);
	if (threadIdx.x == 0)
		tsum[blockIdx.x] = total;
}

// Exclusive offsets of the tiles' counts, then the round's totals: the
// current (m, id_top) are kept for the merge and the next round's are set.
__global__ __launch_bounds__(kPlocScanBlock) void k_ploc_scan(uint32_t *tot, const uint2 *tsum, uint2 *toff) {
	const int m = (int)tot[kPlocM], id_top = (int)tot[kPlocTop];
	if (m <= 1) {
		if (threadIdx.x == 0) {
			tot[kPlocM0] = (uint32_t)m;
			tot[kPlocTop0] = (uint32_t)id_top;
		}
		return;
	}
	const int nt = (m + kPlocTile - 1) / kPlocTile;
	const int per = (nt + kPlocScanBlock - 1) / kPlocScanBlock;
	const int b0 = min(nt, (int)threadIdx.x * per), b1 = min(nt, b0 + per);
	uint2 s = make_uint2(0u, 0u);
	for (int t = b0; t < b1; t++) {
		s.x += tsum[t].x;
		s.y += tsum[t].y;
	}
	__shared__ uint2 wsum[kPlocScanBlock / 64];
	uint2 total;
	uint2 o = block_scan2<kPlocScanBlock>(s, &total, wsum);
	for (int t = b0; t < b1; t++) {
		toff[t] = o;
		o.x += tsum[t].x;
		o.y += tsum[t].y;
	}
	if (threadIdx.x == 0) {
		tot[kPlocM0] = (uint32_t)m;
		tot[kPlocTop0] = (uint32_t)id_top;
		tot[kPlocMerges] = total.x;
		tot[kPlocKept] = total.y;
		if (total.x == 0u || total.y != (uint32_t)m - total.x)
			tot[kPlocErr] = 1u;
		tot[kPlocTop] = (uint32_t)(id_top - (int)total.x);
		tot[kPlocM] = total.y;
	}
}

// The merge: a kept cluster moves to its keep offset; a merging pair's
// owner writes internal node id_top - (its merge offset), with the pair's
// union box, and the new cluster refers to it. Pass r of a tile takes its
// clusters t0 + r * kBlock + threadIdx.x (coalesced), so each pass scans
// its flags and adds the counts of the passes before it.
__global__ __launch_bounds__(kBlock) void k_ploc_merge(const uint32_t *tot, const int *nn, const uint2 *toff,
                                                       const int *cref, const float4 *cbox, const uint32_t *ccnt,
                                                       int *cref2, float4 *cbox2, uint32_t *ccnt2, int *left,
                                                       int *right, uint32_t *count, float4 *box) {
	__shared__ uint2 wsum[2][kBlock / 64];
	const int m = (int)tot[kPlocM0], id_top = (int)tot[kPlocTop0];
	const int t0 = blockIdx.x * kPlocTile;
	if (m <= 1 || t0 >= m)
		return;
	uint2 base = toff[blockIdx.x];
#pragma unroll
	for (int r = 0; r < kPlocPer; r++) {
		const int i = t0 + r * kBlock + threadIdx.x;
		uint32_t mf = 0, kf = 0;
		if (i < m)
			ploc_flags(nn, i, mf, kf);
		uint2 total;
		// alternate the scan's LDS: a pass's writes must not meet the
		// previous pass's reads
		const uint2 ex = block_scan2<kBlock>(make_uint2(mf, kf), &total, wsum[r & 1]);
		if (kf) {
			const uint32_t pos = base.y + ex.y;
			float4 lo = cbox[2 * (size_t)i], hi = cbox[2 * (size_t)i + 1];
			if (mf) {
				const int j = nn[i];
				const int id = id_top - (int)(base.x + ex.x);
				const float4 blo = cbox[2 * (size_t)j], bhi = cbox[2 * (size_t)j + 1];
				lo = make_float4(fminf(lo.x, blo.x), fminf(lo.y, blo.y), fminf(lo.z, blo.z), 0.0f);
				hi = make_float4(fmaxf(hi.x, bhi.x), fmaxf(hi.y, bhi.y), fmaxf(hi.z, bhi.z), 0.0f);
				left[id] = cref[i];
				right[id] = cref[j];
				const uint32_t c = ccnt[i] + ccnt[j];
				count[id] = c;
				box[2 * (size_t)id] = lo;
				box[2 * (size_t)id + 1] = hi;
				cref2[pos] = id;
				ccnt2[pos] = c;
			} else {
				cref2[pos] = cref[i];
				ccnt2[pos] = ccnt[i];
			}
			cbox2[2 * (size_t)pos] = lo;
			cbox2[2 * (size_t)pos + 1] = hi;
		}
		base.x += total.x;
		base.y += total.y;
	}
}

// The last rounds of PLOC, from kPlocTail clusters down to the root, in one
// work-group: the same rounds as the kernels above (same windows, pairs,
// ids and order), with block barriers in place of kernel boundaries, so the
// tree is the same and the four launches (and a share of a host round
// trip) of each small round are saved. The clusters' boxes and nearest neighbours stay in
// LDS for all of these rounds (the neighbour search reads 2 * kPlocR boxes
// per cluster); their references and triangle counts stay in the two global
// buffers. On entry: m clusters in buffer `cur`; st[0] = m, st[1] = id_top
// (the next internal id, counting down). On exit st[0] = 1 and st[1] = -1
// when every cluster has merged. The cluster buffers hold no boxes of these
// rounds afterwards (nothing reads them: the tree is left, right, count and
// box).
constexpr int kPlocTailBlock = 1024;
#ifndef RN_PLOC_TAIL
#define RN_PLOC_TAIL 4096
#endif
constexpr int kPlocTail = RN_PLOC_TAIL;
constexpr int kPlocTailPer = kPlocTail / kPlocTailBlock; // clusters per thread at most
static_assert(kPlocTail % kPlocTailBlock == 0 && kPlocTailPer <= 8, "a thread merges at most 8 clusters");
static_assert(kPlocTail * 36 + kPlocTailBlock / 8 <= 160 * 1024, "the tail's LDS fits a work-group");

__global__ __launch_bounds__(kPlocTailBlock) void k_ploc_tail(int n, int cur, int *cref, const float4 *cbox,
                                                                uint32_t *ccnt, int *left, int *right,
                                                                uint32_t *count, float4 *box, int *st) {
	__shared__ float4 sbox[2 * kPlocTail]; // lo, hi of cluster i at 2i, 2i + 1
	__shared__ int snn[kPlocTail];
	__shared__ uint2 wsum[kPlocTailBlock / 64];
	const int tid = threadIdx.x;
	int m = st[0], id_top = st[1];
	if (m > kPlocTail)
		return; // the host calls the tail at m <= kPlocTail; st[0] > 1 reports it
	for (int i = tid; i < 2 * m; i += kPlocTailBlock)
		sbox[i] = cbox[(size_t)cur * 2 * n + i];
	__syncthreads();
	while (m > 1) {
		const int *cr = cref + (size_t)cur * n;
		const uint32_t *cc = ccnt + (size_t)cur * n;
		int *cr2 = cref + (size_t)(cur ^ 1) * n;
		uint32_t *cc2 = ccnt + (size_t)(cur ^ 1) * n;
		// nearest neighbours (k_ploc_nn)
		for (int i = tid; i < m; i += kPlocTailBlock) {
			const float4 alo = sbox[2 * i], ahi = sbox[2 * i + 1];
			float best = INFINITY;
			int bj = -1;
			const int j0 = max(0, i - kPlocR), j1 = min(m - 1, i + kPlocR);
			for (int j = j0; j <= j1; j++) {
				if (j == i)
					continue;
				float a = merged_area(alo, ahi, sbox[2 * j], sbox[2 * j + 1]);
				if (ploc_better(a, j, best, bj, i)) {
					best = a;
					bj = j;
				}
			}
			snn[i] = bj;
		}
		__syncthreads();
		// flags (ploc_flags) over this thread's contiguous chunk (bit r for
		// cluster b0 + r), and its sums
		const int per = (m + kPlocTailBlock - 1) / kPlocTailBlock;
		const int b0 = min(m, tid * per), b1 = min(m, b0 + per);
		uint32_t fm = 0, fk = 0;
		for (int i = b0; i < b1; i++) {
			const int j = snn[i];
			const bool mutual = j >= 0 && snn[j] == i;
			fm |= (mutual && i < j) ? 1u << (i - b0) : 0u;
			fk |= (mutual && i > j) ? 0u : 1u << (i - b0);
		}
		uint2 total;
		const uint2 ex = block_scan2<kPlocTailBlock>(
		    make_uint2((uint32_t)__builtin_popcount(fm), (uint32_t)__builtin_popcount(fk)), &total, wsum);
		const uint32_t merges = total.x, kept = total.y;
		// merge (k_ploc_merge) with this thread's exclusive offsets; the new
		// boxes wait in registers until every thread has read the old ones
		uint32_t mo = ex.x, ko = ex.y;
		float4 nlo[kPlocTailPer], nhi[kPlocTailPer];
#pragma unroll
		for (int r = 0; r < kPlocTailPer; r++) {
			const int i = b0 + r;
			const bool kf = (fk >> r) & 1u, mf = (fm >> r) & 1u;
			nlo[r] = nhi[r] = make_float4(0.f, 0.f, 0.f, 0.f);
			if (i < b1 && kf) {
				const uint32_t pos = ko;
				float4 lo = sbox[2 * i], hi = sbox[2 * i + 1];
				if (mf) {
					const int j = snn[i];
					const int id = id_top - (int)mo;
					const float4 blo = sbox[2 * j], bhi = sbox[2 * j + 1];
					lo = make_float4(fminf(lo.x, blo.x), fminf(lo.y, blo.y), fminf(lo.z, blo.z), 0.0f);
					hi = make_float4(fmaxf(hi.x, bhi.x), fmaxf(hi.y, bhi.y), fmaxf(hi.z, bhi.z), 0.0f);
					left[id] = cr[i];
					right[id] = cr[j];
					const uint32_t c = cc[i] + cc[j];
					count[id] = c;
					box[2 * (size_t)id] = lo;
					box[2 * (size_t)id + 1] = hi;
					cr2[pos] = id;
					cc2[pos] = c;
				} else {
					cr2[pos] = cr[i];
					cc2[pos] = cc[i];
				}
				nlo[r] = lo;
				nhi[r] = hi;
			}
			if (i < b1) {
				mo += mf;
				ko += kf;
			}
		}
		__syncthreads();
		ko = ex.y;
#pragma unroll
		for (int r = 0; r < kPlocTailPer; r++) {
			if (b0 + r < b1 && ((fk >> r) & 1u)) {
				sbox[2 * ko] = nlo[r];
				sbox[2 * ko + 1] = nhi[r];
				ko++;
			}
		}
		__syncthreads();
		if (merges == 0u || (int)kept != m - (int)merges)
			break; // cannot happen (the least pair of a round is mutual); the host check reports it
		id_top -= (int)merges;
		m = (int)kept;
		cur ^= 1;
	}
	if (tid == 0) {
		st[0] = m;
		st[1] = id_top;
		st[2] = cur;
	}
}

// -------------------------------------------------------------- collapse --

struct Entry {
	int id; // >= 0 binary internal, < 0 leaf ~pos
	float3 lo, hi;
	int cnt; // triangles below
};

__device__ __forceinline__ Entry make_entry(int c, int n, const float4 *box, const uint32_t *count) {
	Entry e;
	e.id = c;
	size_t bi = c >= 0 ? (size_t)c : (size_t)(n - 1 + ~c);
	float4 lo = box[2 * bi], hi = box[2 * bi + 1];
	e.lo = f3(lo.x, lo.y, lo.z);
	e.hi = f3(hi.x, hi.y, hi.z);
	e.cnt = c >= 0 ? (int)count[c] : 1;
	return e;
}

__device__ __forceinline__ float half_area(const Entry &e) {
	float3 d = e.hi - e.lo;
	return d.x * d.y + d.y * d.z + d.z * d.x;
}

// Smallest power of two s with 255 * s >= ext, as an exponent byte (>= 1).
__device__ __forceinline__ uint32_t scale_exp(float ext) {
	float v = ext / 255.0f;
	uint32_t bits = __float_as_uint(v);
	uint32_t e = (bits >> 23) & 0xFFu;
	if (bits & 0x7FFFFFu)
		e += 1;
	return e < 1 ? 1u : e;
}

// Grid origin of one axis of a node: child planes are stored as p + q s,
// q = 0..255, and any p <= lo with p + 255 s >= hi holds all children. A
// child thinner than s along the axis (a wall, an emitter, an axis-aligned
// triangle) gains up to two steps of thickness unless its plane lies on a
// grid line; rays that start or end on such a surface then enter its box.
// The origin chosen puts the most flat child area (the child box's face
// across the axis) on grid lines; p = lo when there is no flat child.
// Entries are read at compile-time indices only, so they stay in registers.
#ifndef RN_GRID_ALIGN
#define RN_GRID_ALIGN 1
#endif
__device__ __forceinline__ float grid_origin(const Entry (&e)[8], int m, int a, float lo, float hi, float s) {
	float plo[8], area[8];
	uint32_t flat = 0;
#pragma unroll
	for (int j = 0; j < 8; j++) {
		const float3 d = e[j].hi - e[j].lo;
		const float ext = a == 0 ? d.x : a == 1 ? d.y : d.z;
		plo[j] = a == 0 ? e[j].lo.x : a == 1 ? e[j].lo.y : e[j].lo.z;
		area[j] = a == 0 ? d.y * d.z : a == 1 ? d.x * d.z : d.x * d.y;
		if (j < m && ext < s)
			flat |= 1u << j;
	}
	if (!flat)
		return lo;
	const float is = 1.0f / s; // a power of two: exact
	auto on_grid = [&](float p) {
		float w = 0.0f;
#pragma unroll
		for (int k = 0; k < 8; k++) {
			const float q = (plo[k] - p) * is;
			w += ((flat >> k) & 1u) && q == floorf(q) ? area[k] : 0.0f;
		}
		return w;
	};
	float best = lo, bw = on_grid(lo);
#pragma unroll
	for (int j = 0; j < 8; j++) {
		if (!((flat >> j) & 1u))
			continue;
		float p = plo[j] - s * ceilf((plo[j] - lo) * is);
		if (p > lo || (double)p + 255.0 * (double)s < (double)hi) // exact: every child must fit
			continue;
		const float w = on_grid(p);
		if (w > bw) {
			bw = w;
			best = p;
		}
	}
	return best;
}

// One BVH8 node per thread, one tree level per launch. The entry list, the
// slot scores and the node record are only indexed at compile-time indices
// (loops unrolled over the 8 slots with predicates), so they live in
// registers rather than in per-thread scratch memory.
constexpr int kCollapseBlock = 64;

__global__ __launch_bounds__(kCollapseBlock) void k_collapse(int k0, int k1, int width, int *task, const float4 *box,
                                                             const int *left, const int *right,
                           const uint32_t *count, const uint32_t *order, const float *verts, int n, Bvh8Node *nodes,
                           float4 *tris, uint32_t *counters, int node_cap) {
	int k = k0 + blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= k1)
		return;
	int b = task[k];
	Entry e[8];
	int m;
	if (b < 0) {
		e[0] = make_entry(b, n, box, count);
		m = 1;
	} else {
		e[0] = make_entry(left[b], n, box, count);
		e[1] = make_entry(right[b], n, box, count);
		m = 2;
		// Open the largest-area internal entry until the node is full: large
		// entries first (fewer levels), then small ones (tighter leaf boxes
		// at no extra node visit).
		while (m < width) {
			int best = -1, id = 0;
			float ba = -1.0f;
#pragma unroll
			for (int j = 0; j < 8; j++) {
				if (j < m && e[j].id >= 0) {
					float a = half_area(e[j]);
					if (a > ba) {
						ba = a;
						best = j;
						id = e[j].id;
					}
				}
			}
			if (best < 0)
				break;
			const Entry l = make_entry(left[id], n, box, count), r = make_entry(right[id], n, box, count);
#pragma unroll
			for (int j = 0; j < 8; j++) {
				if (j == best)
					e[j] = l;
				else if (j == m)
					e[j] = r;
			}
			m++;
		}
	}
	float3 lo = e[0].lo, hi = e[0].hi;
#pragma unroll
	for (int j = 1; j < 8; j++) {
		if (j < m) {
			lo = f3(fminf(lo.x, e[j].lo.x), fminf(lo.y, e[j].lo.y), fminf(lo.z, e[j].lo.z));
			hi = f3(fmaxf(hi.x, e[j].hi.x), fmaxf(hi.y, e[j].hi.y), fmaxf(hi.z, e[j].hi.z));
		}
	}
	// Octant slot assignment: greedily give each (entry, slot) pair with the
	// largest projection of the entry centre onto the slot diagonal (the
	// first strict maximum in entry-major order each round).
	float3 c = 0.5f * (lo + hi);
	float score[8][8];
#pragma unroll
	for (int j = 0; j < 8; j++) {
		float3 dc = 0.5f * (e[j].lo + e[j].hi) - c;
#pragma unroll
		for (int s = 0; s < 8; s++)
			score[j][s] = ((s & 1) ? dc.x : -dc.x) + ((s & 2) ? dc.y : -dc.y) + ((s & 4) ? dc.z : -dc.z);
	}
	uint32_t slot_of = 0; // 4 bits per entry: its slot
	uint32_t used = 0, taken = 0;
	for (int it = 0; it < m; it++) {
		float bs = -INFINITY;
		int bj = 0, bsl = 0;
#pragma unroll
		for (int j = 0; j < 8; j++) {
#pragma unroll
			for (int s = 0; s < 8; s++) {
				if (j < m && !((used >> j) & 1u) && !((taken >> s) & 1u) && score[j][s] > bs) {
					bs = score[j][s];
					bj = j;
					bsl = s;
				}
			}
		}
		if (bs == -INFINITY) { // no ordered score (non-finite boxes): first free entry and slot
			bj = __builtin_ctz(~used);
			bsl = __builtin_ctz(~taken);
		}
		used |= 1u << bj;
		taken |= 1u << bsl;
		slot_of |= (uint32_t)bsl << (4 * bj);
	}
	int n_int = 0, n_tri = 0;
#pragma unroll
	for (int j = 0; j < 8; j++) {
		if (j < m) {
			if (e[j].cnt > kLeafMax)
				n_int++;
			else
				n_tri += e[j].cnt;
		}
	}
	uint32_t child_base = n_int ? atomicAdd(&counters[0], (uint32_t)n_int) : 0u;
	uint32_t tri_base = n_tri ? atomicAdd(&counters[1], (uint32_t)n_tri) : 0u;
	if (child_base + (uint32_t)n_int > (uint32_t)node_cap) {
		atomicOr(&counters[2], 1u); // capacity overflow: the host rebuilds with more room
		return;
	}
	if (tri_base + (uint32_t)n_tri > (uint32_t)n) {
		atomicOr(&counters[2], 2u); // more leaf triangles than input: malformed hierarchy
		return;
	}
	Bvh8Node nd;
	uint32_t ex = scale_exp(hi.x - lo.x), ey = scale_exp(hi.y - lo.y), ez = scale_exp(hi.z - lo.z);
	double sx = ldexp(1.0, (int)ex - 127), sy = ldexp(1.0, (int)ey - 127), sz = ldexp(1.0, (int)ez - 127);
	// the scales are powers of two: multiplying by 1 / s is the exact division
	const double rx = ldexp(1.0, 127 - (int)ex), ry = ldexp(1.0, 127 - (int)ey), rz = ldexp(1.0, 127 - (int)ez);
	if (RN_GRID_ALIGN) {
		lo = f3(grid_origin(e, m, 0, lo.x, hi.x, (float)sx), grid_origin(e, m, 1, lo.y, hi.y, (float)sy),
		        grid_origin(e, m, 2, lo.z, hi.z, (float)sz));
	}
	nd.p[0] = lo.x;
	nd.p[1] = lo.y;
	nd.p[2] = lo.z;
	nd.e[0] = (uint8_t)ex;
	nd.e[1] = (uint8_t)ey;
	nd.e[2] = (uint8_t)ez;
	nd.child_base = child_base;
	nd.tri_base = tri_base;
	nd.tmask = 0;
	nd.pad = 0;
#pragma unroll
	for (int s = 0; s < 8; s++) {
		for (int a = 0; a < 3; a++) {
			nd.qlo[a][s] = 255;
			nd.qhi[a][s] = 0;
		}
	}
	// entries in slot order: internal children take consecutive nodes from
	// child_base and leaf triangles consecutive records from tri_base
	uint32_t imask = 0, ic = 0, off = 0;
#pragma unroll
	for (int s = 0; s < 8; s++) {
		int j = -1;
#pragma unroll
		for (int jj = 0; jj < 8; jj++)
			if (jj < m && ((slot_of >> (4 * jj)) & 15u) == (uint32_t)s)
				j = jj;
		if (j >= 0) {
			Entry q = e[0];
#pragma unroll
			for (int jj = 1; jj < 8; jj++)
				if (jj == j)
					q = e[jj];
			double qlo[3] = {floor(((double)q.lo.x - (double)lo.x) * rx), floor(((double)q.lo.y - (double)lo.y) * ry),
			                 floor(((double)q.lo.z - (double)lo.z) * rz)};
			double qhi[3] = {ceil(((double)q.hi.x - (double)lo.x) * rx), ceil(((double)q.hi.y - (double)lo.y) * ry),
			                 ceil(((double)q.hi.z - (double)lo.z) * rz)};
			for (int a = 0; a < 3; a++) {
				nd.qlo[a][s] = (uint8_t)fmin(fmax(qlo[a], 0.0), 255.0);
				nd.qhi[a][s] = (uint8_t)fmin(fmax(qhi[a], 0.0), 255.0);
			}
			const int cnt = q.cnt;
			if (cnt > kLeafMax) {
				imask |= 1u << s;
				task[child_base + ic] = q.id;
				ic++;
			} else {
				nd.tmask |= ((1u << cnt) - 1u) << (3 * s);
				// the subtree's (<= kLeafMax = 3) sorted positions, left to
				// right: an internal child of a 3-triangle subtree has two
				// leaf children
				int p0 = 0, p1 = 0, p2 = 0, np = 0;
				auto put = [&](int pos) {
					if (np == 0)
						p0 = pos;
					else if (np == 1)
						p1 = pos;
					else
						p2 = pos;
					np++;
				};
				if (q.id < 0) {
					put(~q.id);
				} else {
					const int l = left[q.id], r = right[q.id];
					if (l < 0) {
						put(~l);
					} else {
						put(~left[l]);
						put(~right[l]);
					}
					if (r < 0) {
						put(~r);
					} else {
						put(~left[r]);
						put(~right[r]);
					}
				}
				for (int r = 0; r < kLeafMax; r++) {
					if (r >= cnt || r >= np)
						break;
					uint32_t prim = order[r == 0 ? p0 : r == 1 ? p1 : p2];
					const float *t = verts + 9 * (size_t)prim;
					size_t dst = 3 * (size_t)(tri_base + off + r);
					tris[dst] = make_float4(t[0], t[1], t[2], __int_as_float((int)prim));
					tris[dst + 1] = make_float4(t[3], t[4], t[5], 0.0f);
					tris[dst + 2] = make_float4(t[6], t[7], t[8], 0.0f);
				}
				off += cnt;
			}
		}
	}
	nd.imask = (uint8_t)imask;
	const uint4 *src = reinterpret_cast<const uint4 *>(&nd);
	uint4 *dst = reinterpret_cast<uint4 *>(nodes + k);
#pragma unroll
	for (int q = 0; q < 5; q++)
		dst[q] = src[q];
}

__global__ void k_empty_root(Bvh8Node *nodes) {
	Bvh8Node nd;
	for (int a = 0; a < 3; a++) {
		nd.p[a] = 0.0f;
		nd.e[a] = 127;
		for (int s = 0; s < 8; s++) {
			nd.qlo[a][s] = 255;
			nd.qhi[a][s] = 0;
		}
	}
	nd.tmask = 0;
	nd.pad = 0;
	nd.imask = 0;
	nd.child_base = 0;
	nd.tri_base = 0;
	*nodes = nd;
}

__global__ void k_set_root_task(int *task, int n, uint32_t *counters) {
	task[0] = n >= 2 ? 0 : ~0;
	counters[0] = 1;
	counters[1] = 0;
	counters[2] = 0;
}

unsigned grid_for(size_t n, int block = kBlock) { return (unsigned)std::max<size_t>(1, (n + block - 1) / block); }

void fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t s) {
	k_fill_u32<<<(unsigned)std::min<size_t>(4096, grid_for(n)), kBlock, 0, s>>>(p, v, n);
}

// Binary tree for the collapse with PLOC over the Morton order (n >= 2).
void ploc_tree(Bvh &b, const float *d_verts, int n, hipStream_t s) {
	const uint32_t *order = b.vals.as<uint32_t>();
	int *left = static_cast<int *>(b.left.reserve((size_t)(n - 1) * 4));
	int *right = static_cast<int *>(b.right.reserve((size_t)(n - 1) * 4));
	uint32_t *count = static_cast<uint32_t *>(b.count.reserve((size_t)(n - 1) * 4));
	float4 *box = static_cast<float4 *>(b.boxes.reserve((size_t)(2 * n - 1) * 32));
	int *cref = static_cast<int *>(b.ploc_ref.reserve((size_t)n * 2 * 4));
	float4 *cbox = static_cast<float4 *>(b.ploc_box.reserve((size_t)n * 2 * 32));
	uint32_t *ccnt = static_cast<uint32_t *>(b.ploc_cnt.reserve((size_t)n * 2 * 4));
	int *nn = static_cast<int *>(b.ploc_nn.reserve((size_t)n * 4));
	const size_t nt_max = ((size_t)n + kPlocTile - 1) / kPlocTile;
	uint32_t *fl = static_cast<uint32_t *>(b.ploc_flags.reserve((4 * nt_max + 8) * 4));
	uint2 *tsum = reinterpret_cast<uint2 *>(fl), *toff = reinterpret_cast<uint2 *>(fl + 2 * nt_max);
	uint32_t *tot = fl + 4 * nt_max;
	k_leaf_boxes<<<grid_for(n), kBlock, 0, s>>>(d_verts, order, n, box, cref, cbox, ccnt);
	RN_HIP(hipGetLastError());
	// Rounds run in batches of kPlocBatch with m and id_top on the device
	// (tot); the host reads them back once per batch. Kernels are sized for
	// m_ub, the count at the last read-back, and stop at the device's m.
	// Rounds after m <= kPlocTail within a batch are ordinary rounds: the
	// tree does not depend on where the work-group tail takes over.
	constexpr int kPlocBatch = 8; // 4 / 6 / 8 / 12 rounds per read-back: 1.89 / 1.81 / 1.77 / 1.77 ms at 1M triangles
	int m_ub = n, cur = 0;
	uint32_t hs[8] = {0u, 0u, 0u, (uint32_t)n, (uint32_t)(n - 2), 0u, 0u, 0u};
	RN_HIP(hipMemcpyAsync(tot, hs, kPlocStateWords * sizeof(uint32_t), hipMemcpyHostToDevice, s));
	while (m_ub > 1) {
		if (m_ub <= kPlocTail) { // the rest in one work-group
			int *st = reinterpret_cast<int *>(tot + kPlocM);
			k_ploc_tail<<<1, kPlocTailBlock, 0, s>>>(n, cur, cref, cbox, ccnt, left, right, count, box, st);
			RN_HIP(hipGetLastError());
			break;
		}
		for (int r = 0; r < kPlocBatch; r++) {
			int *cr = cref + (size_t)cur * n, *cr2 = cref + (size_t)(cur ^ 1) * n;
			float4 *cb = cbox + (size_t)cur * 2 * n, *cb2 = cbox + (size_t)(cur ^ 1) * 2 * n;
			uint32_t *cc = ccnt + (size_t)cur * n, *cc2 = ccnt + (size_t)(cur ^ 1) * n;
			const unsigned nt = grid_for(m_ub, kPlocTile);
			k_ploc_nn<<<grid_for(m_ub), kBlock, 0, s>>>(tot + kPlocM, cb, nn);
			k_ploc_count<<<nt, kBlock, 0, s>>>(tot, nn, tsum);
			k_ploc_scan<<<1, kPlocScanBlock, 0, s>>>(tot, tsum, toff);
			k_ploc_merge<<<nt, kBlock, 0, s>>>(tot, nn, toff, cr, cb, cc, cr2, cb2, cc2, left, right, count, box);
			cur ^= 1;
		}
		RN_HIP(hipMemcpyAsync(hs, tot, 5 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
		RN_HIP(hipStreamSynchronize(s));
		if (hs[kPlocErr])
			throw Error("PLOC: a clustering round merged nothing (internal error)");
		m_ub = (int)hs[kPlocM];
	}
	RN_HIP(hipMemcpyAsync(hs, tot, 5 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
	RN_HIP(hipStreamSynchronize(s));
	if (hs[kPlocErr] || hs[kPlocM] != 1u)
		throw Error("PLOC: a clustering round merged nothing (internal error)");
	const int id_top = (int)hs[kPlocTop];
	if (id_top != -1)
		throw Error("PLOC: node count mismatch (internal error)");
}

// Codes + sorted order + (n >= 2) binary tree and refitted boxes.
void lbvh_front(Bvh &b, const float *d_verts, int n, hipStream_t s, bool karras = true) {
	uint32_t *bounds = static_cast<uint32_t *>(b.bounds.reserve(6 * kBoundStride * sizeof(uint32_t)));
	uint32_t *keys = static_cast<uint32_t *>(b.keys.reserve((size_t)n * 4));
	uint32_t *vals = static_cast<uint32_t *>(b.vals.reserve((size_t)n * 4));
	uint32_t *keys2 = static_cast<uint32_t *>(b.keys2.reserve((size_t)n * 4));
	uint32_t *vals2 = static_cast<uint32_t *>(b.vals2.reserve((size_t)n * 4));
	fill_u32(bounds, 0xFFFFFFFFu, 3 * kBoundStride, s);
	fill_u32(bounds + 3 * kBoundStride, 0u, 3 * kBoundStride, s);
	k_centroid_bounds<<<std::min(1024u, grid_for(n)), kBlock, 0, s>>>(d_verts, n, bounds);
	k_morton<<<grid_for(n), kBlock, 0, s>>>(d_verts, n, bounds, keys, vals);
	radix_sort_pairs(keys, vals, keys2, vals2, n, 30, b.hist, b.scan_tmp, s);
	size_t nn = (size_t)(n > 1 ? 2 * n - 1 : 1);
	int *left = static_cast<int *>(b.left.reserve((size_t)std::max(1, n - 1) * 4));
	int *right = static_cast<int *>(b.right.reserve((size_t)std::max(1, n - 1) * 4));
	int *parent = static_cast<int *>(b.parent.reserve(nn * 4));
	int2 *range = static_cast<int2 *>(b.range.reserve((size_t)std::max(1, n - 1) * 8));
	float4 *box = static_cast<float4 *>(b.boxes.reserve(nn * 32));
	uint32_t *visit = static_cast<uint32_t *>(b.visit.reserve((size_t)std::max(1, n - 1) * 4));
	if (!karras) {
		if (n >= 2) {
			ploc_tree(b, d_verts, n, s);
		} else {
			int *cref = static_cast<int *>(b.ploc_ref.reserve(8));
			float4 *cbox = static_cast<float4 *>(b.ploc_box.reserve(64));
			uint32_t *ccnt = static_cast<uint32_t *>(b.ploc_cnt.reserve(8));
			k_leaf_boxes<<<1, kBlock, 0, s>>>(d_verts, vals, n, box, cref, cbox, ccnt);
		}
		return;
	}
	uint32_t *count = static_cast<uint32_t *>(b.count.reserve((size_t)std::max(1, n - 1) * 4));
	if (n >= 2) {
		k_karras<<<grid_for(n - 1), kBlock, 0, s>>>(keys, n, left, right, parent, range, count);
		fill_u32(visit, 0u, (size_t)(n - 1), s);
	}
	k_refit<<<grid_for(n, kRefitChunk), kRefitChunk, 0, s>>>(d_verts, vals, n, parent, left, right, range, box,
	                                                         visit);
}

} // namespace

void device_exclusive_scan(const uint32_t *in, uint32_t *out, size_t n, uint32_t *tmp, hipStream_t s) {
	exclusive_scan(in, out, n, tmp, s);
}

size_t device_scan_tmp_words(size_t n) { return scan_tmp_words(n); }

void radix_sort_pairs(uint32_t *d_keys, uint32_t *d_vals, uint32_t *d_keys_tmp, uint32_t *d_vals_tmp, int n,
                      int bits, DevBuf &hist, DevBuf &scan_tmp, hipStream_t s) {
	if (n <= 1)
		return;
	int nblocks = (n + kRsTile - 1) / kRsTile;
	size_t hn = (size_t)256 * nblocks;
	uint32_t *h = static_cast<uint32_t *>(hist.reserve(2 * hn * 4));
	uint32_t *offs = h + hn;
	uint32_t *tmp = static_cast<uint32_t *>(scan_tmp.reserve(scan_tmp_words(hn) * 4));
	uint32_t *ki = d_keys, *vi = d_vals, *ko = d_keys_tmp, *vo = d_vals_tmp;
	int passes = (bits + 7) / 8;
	for (int p = 0; p < passes; p++) {
		k_rs_hist<<<nblocks, kBlock, 0, s>>>(ki, n, 8 * p, h, nblocks);
		exclusive_scan(h, offs, hn, tmp, s);
		k_rs_scatter<<<nblocks, kBlock, 0, s>>>(ki, vi, n, 8 * p, offs, nblocks, ko, vo);
		std::swap(ki, ko);
		std::swap(vi, vo);
	}
	if (ki != d_keys) {
		RN_HIP(hipMemcpyAsync(d_keys, ki, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
		RN_HIP(hipMemcpyAsync(d_vals, vi, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
	}
}

void build_bvh(Bvh &b, const float *d_verts, int n, int width, hipStream_t s, int builder) {
	auto t0 = std::chrono::steady_clock::now();
	if (width != 4 && width != 8)
		throw Error("bvh width must be 4 or 8");
	if (builder != 0 && builder != 1)
		throw Error("unknown BVH builder");
	b.width = width;
	b.n_tris = n;
	b.builder = builder;
	uint32_t *counters = static_cast<uint32_t *>(b.counters.reserve(4 * sizeof(uint32_t)));
	if (n == 0) {
		Bvh8Node *nodes = static_cast<Bvh8Node *>(b.nodes.reserve(sizeof(Bvh8Node)));
		b.tris.reserve(48);
		k_empty_root<<<1, 1, 0, s>>>(nodes);
		RN_HIP(hipStreamSynchronize(s));
		b.n_nodes = 1;
		b.depth = 1;
		b.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
		return;
	}
	lbvh_front(b, d_verts, n, s, builder == 1);
	float4 *tris = static_cast<float4 *>(b.tris.reserve((size_t)n * 48));
	int cap = n / 2 + 64;
	for (int attempt = 0; attempt < 3; attempt++) {
		Bvh8Node *nodes = static_cast<Bvh8Node *>(b.nodes.reserve((size_t)cap * sizeof(Bvh8Node)));
		int *task = static_cast<int *>(b.task.reserve((size_t)cap * 4));
		k_set_root_task<<<1, 1, 0, s>>>(task, n, counters);
		uint32_t lo = 0, hi = 1, h[4];
		int depth = 0;
		bool overflow = false;
		while (lo < hi) {
			k_collapse<<<grid_for(hi - lo, kCollapseBlock), kCollapseBlock, 0, s>>>(
			    (int)lo, (int)hi, width, task, b.boxes.as<float4>(), b.left.as<int>(), b.right.as<int>(),
			    b.count.as<uint32_t>(), b.vals.as<uint32_t>(), d_verts, n, nodes, tris, counters, cap);
			RN_HIP(hipMemcpyAsync(h, counters, 3 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
			RN_HIP(hipStreamSynchronize(s));
			depth++;
			if (h[2] & 2u)
				throw Error("BVH8 collapse produced more leaf triangles than the input has");
			if (h[2]) {
				overflow = true;
				break;
			}
			lo = hi;
			hi = h[0];
		}
		if (!overflow) {
			if ((int)h[1] != n)
				throw Error("BVH8 collapse lost triangles");
			b.n_nodes = (int)hi;
			b.depth = depth;
			b.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
			return;
		}
		cap = 2 * cap + n;
	}
	throw Error("BVH8 collapse exceeded its node capacity");
}

void lbvh_stages(const float *d_verts, int n, uint32_t *h_codes, int *h_order, int *h_left, int *h_right,
                 hipStream_t s) {
	Bvh b;
	lbvh_front(b, d_verts, n, s);
	RN_HIP(hipMemcpyAsync(h_codes, b.keys.ptr, (size_t)n * 4, hipMemcpyDeviceToHost, s));
	RN_HIP(hipMemcpyAsync(h_order, b.vals.ptr, (size_t)n * 4, hipMemcpyDeviceToHost, s));
	if (n >= 2) {
		RN_HIP(hipMemcpyAsync(h_left, b.left.ptr, (size_t)(n - 1) * 4, hipMemcpyDeviceToHost, s));
		RN_HIP(hipMemcpyAsync(h_right, b.right.ptr, (size_t)(n - 1) * 4, hipMemcpyDeviceToHost, s));
	}
	RN_HIP(hipStreamSynchronize(s));
}

void ploc_stages(const float *d_verts, int n, int *h_order, int *h_left, int *h_right, hipStream_t s) {
	Bvh b;
	lbvh_front(b, d_verts, n, s, /*karras=*/false);
	RN_HIP(hipMemcpyAsync(h_order, b.vals.ptr, (size_t)n * 4, hipMemcpyDeviceToHost, s));
	if (n >= 2) {
		RN_HIP(hipMemcpyAsync(h_left, b.left.ptr, (size_t)(n - 1) * 4, hipMemcpyDeviceToHost, s));
		RN_HIP(hipMemcpyAsync(h_right, b.right.ptr, (size_t)(n - 1) * 4, hipMemcpyDeviceToHost, s));
	}
	RN_HIP(hipStreamSynchronize(s));
}

} // namespace rn
