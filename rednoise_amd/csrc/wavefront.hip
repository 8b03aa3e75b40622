// Wavefront form of the physical estimator (Laine, Karras, Aila,
// "Megakernels Considered Harmful", HPG 2013), for one scene or a batch of
// scenes of uneven size that share one path queue.
//
// Per iteration four kernels run:
//   generate - appends freshly generated camera paths to the queue;
//   trace    - closest hit of every queued path (SoA state), then the any-hit
//              rays the previous iteration's shade stage pushed (an
//              unoccluded ray stores its path's new radiance in the path's
//              slot; without slots it adds its contribution to the frame
//              buffer with float atomics);
//   shade    - emission, NEE light sample, bounce and Russian roulette;
//              shadow rays and continuing paths are pushed with one wave64
//              ballot per wave and one atomic per block (compaction keeps
//              the next queue dense);
//   advance  - one lane: rotates the queues and plans the next generation so
//              the queue stays full until the work runs out.
// The trace kernel is persistent: a wave takes chunks of queue entries from
// an atomic counter and gives a lane whose ray has finished the next ray of
// the chunk as soon as enough lanes are idle, so a wave no longer runs as
// long as its slowest ray once per 64 rays.
// Queue sizes live in device memory, so the iteration can be captured once in
// a hipGraph and replayed; the host only polls a done flag between replays.
#include "rn_pathp.h"
#include "wavefront.h"

#include <algorithm>
#include <cstddef>
#include <cstring>
#include <vector>

namespace rn {

namespace {

constexpr int kPixTile = 8; // 8x8 pixel tiles in generation order

// Unsigned division by a divisor fixed for a render, as a multiply-high,
// two shifts and an add: Granlund and Montgomery, "Division by invariant
// integers using multiplication" (PLDI 1994), figure 4.1, exact for every
// 32-bit dividend. m = floor(2^32 (2^l - d) / d) + 1 with l = ceil(log2 d),
// s1 = min(l, 1), s2 = max(l - 1, 0); set on the host (make_fastdiv).
struct FastDiv {
	uint32_t m, s1, s2;
};
__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv &f) {
	const uint32_t t = __umulhi(n, f.m);
	return (t + ((n - t) >> f.s1)) >> f.s2;
}

struct WScene {
	DevScene sc;
	DevCamera cam;
	DevParams p;
	float *out;
	unsigned long long first, count; // global work range
	int tiles_x, npix_pad;
	FastDiv fd_width, fd_tiles, fd_npix; // division by p.width, tiles_x, npix_pad
};

// Chunk counters of the trace kernels are striped over kStripes addresses,
// one 128-byte line each, so waves contend on eight counters instead of one.
constexpr int kStripes = 8;
constexpr int kRing = 72; // > the largest max_depth a render accepts

// The any-hit rays an iteration's shade stage pushes are traced in the next
// iteration's trace launch, after its closest-hit rays (they are
// independent: the shadow queue is rewritten only by the next shade stage).
// One persistent launch per iteration instead of two has one drain tail
// instead of two: 38.0 against 39.0 ms of tracing per 1M-triangle 1080p frame.
constexpr int kStripePad = 32; // unsigneds per counter line

struct WCtl {
	unsigned cur_count, gen_count;
	unsigned next_count, sh_count; // one 64-bit word: k_shade bumps both with one atomic
	unsigned long long gen_base, work_next, work_total;
	unsigned done, iter, done_iter; // done_iter: value of iter when done was first set
	unsigned sh_pending;            // any-hit rays to trace in the coming iteration
	unsigned pad[18];
	unsigned trace_ctr[kStripes * kStripePad];
	unsigned shadow_ctr[kStripes * kStripePad];
	// Scene-index range [ring_lo, ring_hi] of the paths generated in each of
	// the last kRing iterations (ring_hi < 0: none), and the one scene every
	// path alive in the current iteration belongs to (-1: several or none).
	int single;
	int ring_lo[kRing], ring_hi[kRing];
};
static_assert(offsetof(WCtl, next_count) % 8 == 0 && offsetof(WCtl, sh_count) == offsetof(WCtl, next_count) + 4,
              "next_count and sh_count must form one aligned 64-bit word");
static_assert(offsetof(WCtl, trace_ctr) % 128 == 0, "stripe counters start on a 128-byte line");

RN_DEV void reset_stripes(WCtl *c) {
	for (int k = 0; k < kStripes; k++) {
		c->trace_ctr[k * kStripePad] = 0;
		c->shadow_ctr[k * kStripePad] = 0;
	}
}

struct WQ {
	float4 *o[2], *d[2], *b[2];
	int *hitp;     // per queue entry: hit triangle id, -1 = miss
	float4 *hitn;  // per queue entry: cross(v1 - v0, v2 - v0) of the hit triangle, t
	float4 *so, *sd, *sc;
	WCtl *ctl;
	// Radiance slots, 3 floats per path of the render's work range (global
	// work index = slot), or null: contributions go to the frame buffers
	// with float atomics. A path's depth-0 shade stage writes its slot (the
	// camera ray's emission, or zero); a shade stage that pushes an any-hit
	// ray reads the slot and puts its value plus the ray's contribution in
	// the ray's record, and the next trace launch stores that if the ray is
	// unoccluded (one any-hit ray per path per launch, and only shade stages
	// and trace launches between them touch the slot, so no write races);
	// k_reduce_slots sums a pixel's slots in sample order.
	float *acc;
	const WScene *scenes;
	int nscenes;
	int max_depth; // the largest of the scenes' max_depth (paths live at most this many iterations)
	unsigned capacity;
	DevCounters *cnt;
};

__global__ void k_fill(uint32_t *p, size_t n) {
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
		p[i] = 0u;
}

__global__ void k_init(WCtl *ctl, unsigned long long total) {
	for (int k = 0; k < kRing; k++)
		ctl->ring_hi[k] = -1;
	ctl->single = -1;
	ctl->cur_count = ctl->gen_count = ctl->next_count = ctl->sh_count = 0;
	reset_stripes(ctl);
	ctl->gen_base = 0;
	ctl->work_next = 0;
	ctl->work_total = total;
	ctl->done = 0;
	ctl->iter = 0;
	ctl->done_iter = 0;
	ctl->sh_pending = 0;
}

RN_DEV int find_scene(const WScene *sc, int n, unsigned long long g) {
	int lo = 0, hi = n - 1;
	while (lo < hi) {
		int mid = (lo + hi + 1) >> 1;
		if (sc[mid].first <= g)
			lo = mid;
		else
			hi = mid - 1;
	}
	return lo;
}

// Rotate queues and plan the next generation.
__global__ void k_advance(WQ q, int first) {
	WCtl *c = q.ctl;
	if (!first)
		c->cur_count = c->next_count;
	c->sh_pending = first ? 0u : c->sh_count;
	c->next_count = 0;
	c->sh_count = 0;
	reset_stripes(c);
	unsigned long long remaining = c->work_total - c->work_next;
	unsigned long long room = (unsigned long long)(q.capacity - c->cur_count);
	unsigned gen = (unsigned)(remaining < room ? remaining : room);
	c->gen_base = c->work_next;
	c->gen_count = gen;
	c->work_next += gen;
	// Once done, every later launch of the iteration finds empty queues and
	// does nothing, so the host may enqueue more iterations than are needed.
	const unsigned was_done = c->done;
	c->done = (c->cur_count == 0 && gen == 0 && c->sh_pending == 0) ? 1u : 0u;
	c->iter++;
	if (!was_done && c->done)
		c->done_iter = c->iter;
	// A path generated in iteration j is traced in iterations j .. j +
	// max_depth - 1 at most (its last any-hit ray one iteration later when
	// that ray shares the next trace launch), so the scenes of the rays of
	// the coming iteration are among those generated in the last `window`
	// iterations, this one included.
	const int window = q.max_depth + 1;
	const int slot = (int)(c->iter % (unsigned)window);
	c->ring_lo[slot] = gen > 0 ? find_scene(q.scenes, q.nscenes, c->gen_base) : 0;
	c->ring_hi[slot] = gen > 0 ? find_scene(q.scenes, q.nscenes, c->gen_base + gen - 1) : -1;
	int lo = 0x7FFFFFFF, hi = -1;
	for (int k = 0; k < window; k++)
		if (c->ring_hi[k] >= 0) {
			lo = min(lo, c->ring_lo[k]);
			hi = max(hi, c->ring_hi[k]);
		}
	c->single = (hi >= 0 && lo == hi) ? lo : -1;
}


// Queue records are written once and read back by the next stage, after
// gigabytes of other traffic: non-temporal stores keep them from displacing
// the BVH nodes and triangles the trace kernels hold in the caches (1M
// triangles at 1080p: trace stages 0.2-0.3 ms faster per frame).
#ifndef RN_NT_STORES
#define RN_NT_STORES 1
#endif
RN_DEV void put4(float4 *p, float4 v) {
	if (RN_NT_STORES) {
		__builtin_nontemporal_store(v.x, &p->x);
		__builtin_nontemporal_store(v.y, &p->y);
		__builtin_nontemporal_store(v.z, &p->z);
		__builtin_nontemporal_store(v.w, &p->w);
	} else {
		*p = v;
	}
}

// The hit records and the ray records read once by the trace refill and by
// the shade stage move non-temporally too (another 0.2 ms per frame).
#ifndef RN_NT_MORE
#define RN_NT_MORE 1
#endif
RN_DEV float4 get4(const float4 *p) {
	if (RN_NT_MORE)
		return make_float4(__builtin_nontemporal_load(&p->x), __builtin_nontemporal_load(&p->y),
		                   __builtin_nontemporal_load(&p->z), __builtin_nontemporal_load(&p->w));
	return *p;
}
RN_DEV void put4m(float4 *p, float4 v) {
	if (RN_NT_MORE)
		put4(p, v);
	else
		*p = v;
}
RN_DEV int geti(const int *p) { return RN_NT_MORE ? __builtin_nontemporal_load(p) : *p; }
RN_DEV void puti(int *p, int v) {
	if (RN_NT_MORE)
		__builtin_nontemporal_store(v, p);
	else
		*p = v;
}

// Set bits of a wave-wide mask in the lanes below this one.
RN_DEV unsigned lanes_below(uint64_t m) {
	return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// meta word: scene index (low 16 bits) | depth << 16; depth 255 = dead slot.
constexpr uint32_t kDead = 255u;

// The scene table seen through the constant address space: the kernels
// never write it, and a load through it at a uniform index is a scalar load
// wherever it stands (through a generic pointer the compiler makes a load
// that follows a store in the same loop a vector load).
typedef __attribute__((address_space(4))) const WScene cWScene;

// Field-wise copies out of the constant-address-space table (a struct copy
// would need a copy constructor taking the qualified type).
RN_DEV DevScene scene_of(const cWScene &w) {
	DevScene d;
	d.nodes = w.sc.nodes;
	d.tris = w.sc.tris;
	d.albedo = w.sc.albedo;
	d.emission = w.sc.emission;
	d.light_tri = w.sc.light_tri;
	d.light_cdf = w.sc.light_cdf;
	d.light_geom = w.sc.light_geom;
	d.n = w.sc.n;
	d.nl = w.sc.nl;
	d.total_light_area = w.sc.total_light_area;
	d.depth = w.sc.depth;
	return d;
}
RN_DEV DevParams params_of(const cWScene &w) {
	DevParams d;
	d.width = w.p.width;
	d.height = w.p.height;
	d.spp = w.p.spp;
	d.max_depth = w.p.max_depth;
	d.rr_depth = w.p.rr_depth;
	d.rr_prob = w.p.rr_prob;
	d.seed = w.p.seed;
	d.scene_id = w.p.scene_id;
	d.ray_eps = w.p.ray_eps;
	d.sample_base = w.p.sample_base;
	return d;
}

// Compact camera paths. In a one-scene render a camera path's queue entry
// holds only its direction record (d.xyz, sample): its origin is the camera,
// its throughput 1, its meta word (scene 0, depth 0), and its pixel follows
// from the work index gen_base + (i - cur_count) of entry i >= cur_count (the
// camera paths k_generate appended). A sample of kDeadSample marks an entry
// outside the image (tile padding). Batches keep full records.
constexpr uint32_t kDeadSample = 0xFFFFFFFFu;
RN_DEV bool compact_camera_paths(const WQ &q) { return q.nscenes == 1; }

// Radiance slot of the path of pixel `pix`, sample `smp`: its global work
// index, i.e. the inverse of k_generate's order (sample-major, 8x8 tiles,
// rows within a tile). A camera path's slot is the work index k_generate
// gave its queue entry; later vertices find it from the pixel and sample
// they carry. The host keeps the work range below 2^32 in slot mode.
struct SlotMap {
	uint32_t first, width, tiles_x, npix_pad, sbase;
	FastDiv fd_width, fd_tiles;
};
RN_DEV SlotMap slot_map(const WScene &ws) {
	return SlotMap{(uint32_t)ws.first, (uint32_t)ws.p.width, (uint32_t)ws.tiles_x, (uint32_t)ws.npix_pad,
	               ws.p.sample_base, ws.fd_width, ws.fd_tiles};
}
// The pixel of the camera path of work index g and sample smp (the inverse
// of k_generate's tiling: smp is the sample plane, the remainder the tiled
// pixel index).
RN_DEV uint32_t camera_pixel(const SlotMap &m, unsigned long long g, uint32_t smp) {
	const uint32_t pp = (uint32_t)(g - m.first - (unsigned long long)(smp - m.sbase) * m.npix_pad);
	const uint32_t tile = pp >> 6, t = pp & 63u;
	const uint32_t ty = fdiv(tile, m.fd_tiles), tx = tile - ty * m.tiles_x;
	const uint32_t x = tx * kPixTile + (t & 7u), y = ty * kPixTile + (t >> 3);
	return y * m.width + x;
}
RN_DEV uint32_t path_slot(const SlotMap &m, uint32_t pix, uint32_t smp) {
	const uint32_t y = fdiv(pix, m.fd_width), x = pix - y * m.width;
	const uint32_t pp = ((((y >> 3) * m.tiles_x) + (x >> 3)) << 6) | ((y & 7u) << 3) | (x & 7u);
	return m.first + (smp - m.sbase) * m.npix_pad + pp;
}

// a + b rounded on its own, never fused with the product that made b (the
// megakernel adds its contributions the same way).
RN_DEV float3 add_uncontracted(float3 a, float3 b) {
#pragma clang fp contract(off)
	return f3(a.x + b.x, a.y + b.y, a.z + b.z);
}

// Plain loads and stores of a radiance slot (global memory).
RN_DEV float3 slot_load(const float *p) {
	typedef __attribute__((address_space(1))) const float gf;
	return f3(((gf *)p)[0], ((gf *)p)[1], ((gf *)p)[2]);
}
RN_DEV void slot_store(float *p, float3 v) {
	typedef __attribute__((address_space(1))) float gf;
	((gf *)p)[0] = v.x;
	((gf *)p)[1] = v.y;
	((gf *)p)[2] = v.z;
}

// New camera paths for queue entries [cur_count, cur_count + gen_count).
// The scene table is a kernel argument of its own, marked restrict: nothing
// the kernel stores aliases it, so loads at a uniform index are scalar loads.
__global__ __launch_bounds__(256) void k_generate(WQ q, int cur, const WScene *__restrict__ scenes) {
	WCtl *c = q.ctl;
	const unsigned ncur = c->cur_count, n = ncur + c->gen_count;
	const unsigned long long gbase = c->gen_base;
	float4 *O = q.o[cur], *D = q.d[cur], *B = q.b[cur];
	unsigned long long npaths = 0;
	for (unsigned base = ncur + blockIdx.x * 256u; base < n; base += gridDim.x * 256u) {
		const unsigned i = base + threadIdx.x;
		if (i >= n)
			continue;
		unsigned long long g = gbase + (i - ncur);
		const int si = q.nscenes == 1 ? 0 : find_scene(scenes, q.nscenes, g);
		const WScene &ws = scenes[si];
		unsigned long long local = g - ws.first;
		uint32_t smp, pp;
		if ((local >> 32) == 0) { // 32-bit division for work ranges below 2^32
			const uint32_t s = fdiv((uint32_t)local, ws.fd_npix);
			smp = ws.p.sample_base + s;
			pp = (uint32_t)local - s * (uint32_t)ws.npix_pad;
		} else {
			smp = ws.p.sample_base + (uint32_t)(local / (unsigned long long)ws.npix_pad);
			pp = (uint32_t)(local % (unsigned long long)ws.npix_pad);
		}
		uint32_t tile = pp >> 6, t = pp & 63u;
		const uint32_t ty = fdiv(tile, ws.fd_tiles);
		int x = (int)(tile - ty * (uint32_t)ws.tiles_x) * kPixTile + (int)(t & 7u);
		int y = (int)ty * kPixTile + (int)(t >> 3);
		float4 o4 = make_float4(0.f, 0.f, 0.f, 0.f), d4 = o4;
		uint32_t meta = (uint32_t)si | (kDead << 16);
		if (x < ws.p.width && y < ws.p.height) {
			uint32_t pix = (uint32_t)(y * ws.p.width + x);
			u32x4 r = philox4x32_10(pix, smp, 0u, 0u, ws.p.seed, ws.p.scene_id);
			float3 d = camera_dir(ws.cam, x, y, ws.p.width, ws.p.height, u01(r.x), u01(r.y));
			o4 = make_float4(ws.cam.eye[0], ws.cam.eye[1], ws.cam.eye[2], __uint_as_float(pix));
			d4 = make_float4(d.x, d.y, d.z, __uint_as_float(smp));
			meta = (uint32_t)si;
			npaths++;
		}
		if (compact_camera_paths(q)) {
			if ((meta >> 16) == kDead)
				d4.w = __uint_as_float(kDeadSample);
			put4(D + i, d4);
		} else {
			put4(O + i, o4);
			put4(D + i, d4);
			put4(B + i, make_float4(1.f, 1.f, 1.f, __uint_as_float(meta)));
		}
	}
	wave_add_u64(&q.cnt->shard[blockIdx.x % kCounterShards].paths, npaths);
}

// Queue entries a wave takes from the chunk counter at a time: about
// 1/kChunksPerWave of a wave's share of the queue, within [kChunkMin,
// kChunkMax] (fewer counter atomics on long queues, balance on short ones);
// and the number of idle lanes that triggers a refill.
#ifndef RN_CHUNKS_PER_WAVE
#define RN_CHUNKS_PER_WAVE 16
#endif
#ifndef RN_CHUNK_MIN
#define RN_CHUNK_MIN 256 // 128 / 256 / 384 / 512: 51.0-52.0 / 50.4-51.6 / 51.7 / 52.0 ms per 1M-triangle 1080p frame
#endif
#ifndef RN_CHUNK_MAX
#define RN_CHUNK_MAX 1024
#endif
#ifndef RN_REFILL_MIN
#define RN_REFILL_MIN 16
#endif
constexpr unsigned kChunksPerWave = RN_CHUNKS_PER_WAVE, kChunkMin = RN_CHUNK_MIN, kChunkMax = RN_CHUNK_MAX;
constexpr int kRefillMin = RN_REFILL_MIN;
static_assert(kRefillMin >= 1 && kRefillMin <= 64, "refill threshold is a lane count");

// What the single-scene trace instance keeps of its scene (unused by the
// batch instance, whose lanes read their own scene's at each refill).
struct TraceScene {
	const Bvh8Node *nodes;
	const float4 *tris;
	float *out;
	float4 eye;
};

// Persistent trace over a queue. QUERY 0: closest hit of path-queue entries
// [0, cur_count + gen_count), written to q.hit; dead entries are skipped.
// QUERY 1: any hit of shadow-queue entries [0, sh_count); unoccluded
// contributions go to the frame buffer.
//
// Every branch that decides on refills or on leaving the loop depends only on
// ballots and on values read from lane 0, so it is uniform over the wave and
// all 64 lanes reach each __ballot/__shfl together. The loop ends once the
// chunk counter has passed the queue length and every lane is idle.
template <int QUERY, int LDS, int SPILL, int BS, bool ONE>
RN_DEV void trace_queue(const WQ &q, int cur, uint64_t *stk, const uint8_t *octab, const TraceScene &ts) {
	WCtl *c = q.ctl;
	const unsigned n = QUERY == 0 ? c->cur_count + c->gen_count : c->sh_pending;
	unsigned *ctrs = QUERY == 0 ? c->trace_ctr : c->shadow_ctr;
	// chunk j = stripe + kStripes * k is the k-th chunk of stripe j % kStripes;
	// a wave starts on its own stripe and moves on when it is exhausted
	const unsigned waves = gridDim.x * (unsigned)(BS / 64);
	const unsigned chunk = min(kChunkMax, max(kChunkMin, n / (waves * kChunksPerWave)));
	const unsigned nch = (n + chunk - 1) / chunk;
	// wave-uniform values are taken with readfirstlane (every lane is active
	// where they are read), so they live in scalar registers
	unsigned stripe =
	    __builtin_amdgcn_readfirstlane((blockIdx.x * (unsigned)(BS / 64) + (threadIdx.x >> 6)) % (unsigned)kStripes);
	const float4 *O = QUERY == 0 ? q.o[cur] : q.so;
	const float4 *D = QUERY == 0 ? q.d[cur] : q.sd;
	const float4 *M = QUERY == 0 ? q.b[cur] : q.sc; // .w: meta (closest) or scene index (shadow)
	const int lane = threadIdx.x & 63;
	unsigned next = 0, end = 0;
	bool more = true, valid = false;
	unsigned idx = 0;
	// Any hit, read at the refill so that an unoccluded ray's write waits
	// for no load: the ray's radiance slot and the slot's value with the
	// ray's contribution added (or its pixel and contribution without
	// slots), and (batches) its scene.
	unsigned aux = 0, fin_scene = 0;
	float3 fin_c = f3(0.f, 0.f, 0.f);
	// ONE: every queued ray belongs to scene c->single, whose BVH, triangles
	// and frame buffer are held in scalar registers
	const Bvh8Node *nodes = ts.nodes;
	const float4 *tris = ts.tris;
	float *const out_one = ts.out;
	// compact camera paths (entries from cur_count on) start at the eye
	const bool compact = compact_camera_paths(q);
	const unsigned ncur = c->cur_count;
	const float4 eye = ts.eye;
	Trav<SPILL> t;
	unsigned nrays = 0; // this lane's rays (at most the queue length, below 2^31)
	while (true) {
		const uint64_t idle = __ballot(!valid);
		if (more && __popcll(idle) >= kRefillMin) {
			if (next >= end) {
				unsigned cb = 0xFFFFFFFFu;
				if (lane == 0) {
					for (int tries = 0; tries < kStripes; tries++) {
						unsigned *ctr = ctrs + stripe * kStripePad;
						// skip a stripe seen exhausted without an atomic
						unsigned k0 = __hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
						if (stripe + (unsigned)kStripes * k0 < nch) {
							unsigned k = atomicAdd(ctr, 1u);
							unsigned j = stripe + (unsigned)kStripes * k;
							if (j < nch) {
								cb = j * chunk;
								break;
							}
						}
						stripe = (stripe + 1u) % (unsigned)kStripes;
					}
				}
				cb = __builtin_amdgcn_readfirstlane(cb);
				stripe = __builtin_amdgcn_readfirstlane(stripe);
				if (cb == 0xFFFFFFFFu) {
					more = false;
				} else {
					next = cb;
					end = min(cb + chunk, n);
				}
			}
			if (next < end) {
				// idle lanes below this one (v_mbcnt: no lane mask held in registers)
				const unsigned my = next + lanes_below(idle);
				if (!valid && my < end) {
					// the entry's loads are issued together
					// path records stream past non-temporally; shadow records
					// load plainly (0.26 ms per frame faster than streamed)
					// a compact camera path: its direction record only
					const bool cam = QUERY == 0 && ONE && compact && my >= ncur;
					const float4 d4 = QUERY == 0 ? get4(D + my) : D[my];
					const float4 o4 = cam ? eye : (QUERY == 0 ? get4(O + my) : O[my]);
					// closest hit: the path's meta word (a one-scene render's
					// paths below cur_count, pushed by the shade stage, are
					// live: their meta is not read); any hit: the
					// contribution and scene index
					const float4 m4 = QUERY == 1 ? M[my]
					                 : make_float4(0.f, 0.f, 0.f,
					                               cam ? __uint_as_float(__float_as_uint(d4.w) == kDeadSample
					                                                         ? kDead << 16 : 0u)
					                               : (ONE && compact) ? 0.f
					                                                  : M[my].w);
					const uint32_t meta = __float_as_uint(m4.w);
					if (QUERY == 1 || (meta >> 16) != kDead) {
						if (!ONE) {
							const DevScene &sc = q.scenes[meta & 0xFFFFu].sc;
							nodes = sc.nodes;
							tris = sc.tris;
						}
						trav_init<0>(t, xyz(o4), xyz(d4), QUERY == 0 ? INFINITY : d4.w);
						valid = true;
						idx = my;
						if (QUERY == 1) {
							aux = __float_as_uint(o4.w);
							fin_c = xyz(m4);
							if (!ONE)
								fin_scene = meta & 0xFFFFu;
						}
						nrays++;
					}
				}
				next = min(next + (unsigned)__popcll(idle), end);
			}
		}
		// one back-edge: with a second one (a `continue` when no lane holds a
		// ray) the compiler copied the traversal state in and out of spare
		// registers on every step (70 against 31 moves in the loop)
		if (__ballot(valid) == 0 && !more)
			break;
		if (valid && trav_step<QUERY, 0, LDS, SPILL, false, QUERY == 0>(t, nodes, tris, -1, stk, BS, threadIdx.x, nullptr,
		                                                                  octab)) {
			valid = false;
			if (QUERY == 0) {
				puti(q.hitp + idx, t.h.prim);
				if (t.h.prim >= 0)
					put4m(q.hitn + idx, make_float4(t.h.ng.x, t.h.ng.y, t.h.ng.z, t.h.t));
			} else if (t.h.prim == 0) {
				if (q.acc) {
					slot_store(q.acc + 3 * (size_t)aux, fin_c);
				} else {
					float *out = ONE ? out_one : q.scenes[fin_scene].out;
					float *fb = out + 3 * (size_t)aux;
					atomic_add_global(fb, fin_c.x);
					atomic_add_global(fb + 1, fin_c.y);
					atomic_add_global(fb + 2, fin_c.z);
				}
			}
		}
	}
	CounterSet *cs = &q.cnt->shard[blockIdx.x % kCounterShards];
	wave_add_u64(QUERY == 0 ? &cs->closest : &cs->shadow, (unsigned long long)nrays);
}

// Occupancy target of the trace kernels without a spill stack, in waves per
// SIMD: 6 fits both instances in 80 VGPRs with no scratch (the single-scene
// one 2% faster at 1M triangles than at the compiler's 87 VGPRs and 5 waves,
// the batch one 1.4% on 64 scenes against its 95 VGPRs); 0 = the compiler's
// choice.
#ifndef RN_TRACE_WPE
#define RN_TRACE_WPE 6
#endif
#define RN_TRACE_ATTR \
	__attribute__((amdgpu_waves_per_eu(SPILL == 0 ? RN_TRACE_WPE : 0, SPILL == 0 ? RN_TRACE_WPE : 0)))

// The octant permutation of a node's hit mask as a table in LDS (one byte
// load per node visit instead of three conditional bit swaps).
#ifndef RN_OCT_TABLE
#define RN_OCT_TABLE 1
#endif

// Which instance runs an iteration: the one-scene instance when every queued
// path belongs to one scene (always, for a one-scene render), the batch
// instance otherwise; the other returns at once.
template <bool ONE> RN_DEV bool instance_skips(const WQ &q) {
	const int single = q.ctl->single;
	return ONE ? single < 0 : single >= 0;
}

// Both queries in one persistent launch: a wave drains the closest-hit
// queue, then the any-hit queue of the previous iteration.
template <int LDS, int SPILL, int BS, bool ONE>
__global__ __launch_bounds__(BS) RN_TRACE_ATTR void k_trace(WQ q, int cur) {
	if (instance_skips<ONE>(q))
		return;
	__shared__ uint64_t stk[LDS * BS];
	__shared__ uint8_t octab[RN_OCT_TABLE ? 2048 : 1];
	if (RN_OCT_TABLE) {
		fill_oct_table(octab, threadIdx.x, BS);
		__syncthreads();
	}
	// the single scene's pointers and camera, read once before any store (so
	// into scalar registers) for both queries
	TraceScene ts{nullptr, nullptr, nullptr, make_float4(0.f, 0.f, 0.f, 0.f)};
	if (ONE) {
		const cWScene &w = ((const cWScene *)q.scenes)[q.ctl->single];
		ts.nodes = w.sc.nodes;
		ts.tris = w.sc.tris;
		ts.out = (float *)w.out;
		ts.eye = make_float4(w.cam.eye[0], w.cam.eye[1], w.cam.eye[2], 0.f);
	}
	trace_queue<0, LDS, SPILL, BS, ONE>(q, cur, stk, RN_OCT_TABLE ? octab : nullptr, ts);
	trace_queue<1, LDS, SPILL, BS, ONE>(q, 0, stk, RN_OCT_TABLE ? octab : nullptr, ts);
}

// Shade every queued path at its closest hit: emission (camera rays only),
// the NEE light sample, the bounce and the roulette; push the shadow ray and
// the continuing path. Pushes are counted per block: one atomic per queue per
// 1024 entries (same-address atomics serialise at about 11 ns each on
// MI355X, so one per wave would bound this kernel).
#ifndef RN_SHADE_BLOCK
#define RN_SHADE_BLOCK 1024
#endif
constexpr int kShadeBlock = RN_SHADE_BLOCK;
constexpr int kShadeWaves = kShadeBlock / 64;

// Two blocks per CU (8 waves per SIMD) need at most 64 VGPRs and 80 SGPRs:
// one SGPR more halves the single-scene instance's occupancy (shade 17-25%
// slower per iteration at 1M triangles, 1080p), so the occupancy is asked
// for. The batch instance, at 74 VGPRs, runs one block per CU either way.
template <bool ONE>
__global__ __launch_bounds__(kShadeBlock) __attribute__((amdgpu_waves_per_eu(ONE ? 8 : 0, ONE ? 8 : 0))) void k_shade(
    WQ q, int cur, const WScene *__restrict__ scenes) {
	if (instance_skips<ONE>(q))
		return;
	const unsigned one = ONE ? (unsigned)q.ctl->single : 0u;
	const bool compact = compact_camera_paths(q);
	// the camera and the scene's slot map (a compact camera path's pixel, a
	// path's radiance slot), kept in LDS: scalar registers are at their limit
	// for two blocks per CU
	__shared__ float cam_eye[3];
	__shared__ SlotMap cam_map;
	if (ONE) {
		if (threadIdx.x < 3)
			cam_eye[threadIdx.x] = scenes[one].cam.eye[threadIdx.x];
		if (threadIdx.x == 0)
			cam_map = slot_map(scenes[one]);
		__syncthreads();
	}
	static_assert(kShadeWaves <= 64 && kShadeBlock < 65536, "block counts are packed in 16-bit halves");
	__shared__ unsigned wcnt[kShadeWaves]; // per wave: shadow pushes | continuing pushes << 16
	__shared__ unsigned bbase[2];
	WCtl *c = q.ctl;
	// entries [cur_count, cur_count + gen_count) are the camera paths of
	// work indices gen_base + (i - cur_count)
	const unsigned ncur = c->cur_count, n = ncur + c->gen_count;
	const unsigned long long gbase = c->gen_base;
	const float4 *O = q.o[cur], *D = q.d[cur], *B = q.b[cur];
	float4 *NO = q.o[cur ^ 1], *ND = q.d[cur ^ 1], *NB = q.b[cur ^ 1];
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
	for (unsigned base = blockIdx.x * (unsigned)kShadeBlock; base < n; base += gridDim.x * (unsigned)kShadeBlock) {
		const unsigned i = base + threadIdx.x;
		uint32_t meta = kDead << 16;
		float4 b4 = make_float4(0.f, 0.f, 0.f, 0.f);
		// a compact camera path: its direction record only (scene 0, depth 0,
		// throughput 1)
		const bool cam = ONE && compact && i >= ncur && i < n;
		if (cam) {
			meta = __float_as_uint(D[i].w) == kDeadSample ? kDead << 16 : 0u;
			b4 = make_float4(1.f, 1.f, 1.f, __uint_as_float(meta));
		} else if (i < n) {
			b4 = get4(B + i);
			meta = __float_as_uint(b4.w);
		}
		const uint32_t depth = meta >> 16;
		{
			bool push_sh = false, push_next = false;
			float3 so = f3(0.f, 0.f, 0.f), wo = so, beta = so, contrib = so, shw = so;
			float shmax = 0.f;
			float4 o4 = make_float4(0.f, 0.f, 0.f, 0.f), d4 = o4;
			uint32_t aux = 0; // slot mode: the path's radiance slot
			auto shade_vertex = [&](const cWScene &ws) {
				const int prim = geti(q.hitp + i);
				if (prim >= 0) {
					const float4 hn = get4(q.hitn + i);
					d4 = get4(D + i);
					o4 = cam ? make_float4(cam_eye[0], cam_eye[1], cam_eye[2],
					                       __uint_as_float(camera_pixel(cam_map, gbase + (i - ncur), __float_as_uint(d4.w))))
					         : get4(O + i);
					beta = xyz(b4);
					// copies through the constant address space: scalar loads
					// at the round's (uniform) scene index
					const DevScene sc = scene_of(ws);
					const DevParams p = params_of(ws);
					float3 o = xyz(o4), d = xyz(d4);
					uint32_t pix = __float_as_uint(o4.w), smp = __float_as_uint(d4.w);
					float3 P = o + d * hn.w;
					float3 ng = facing_from_cross(xyz(hn), d);
					// slot mode: the path's slot and its value once this vertex's
					// emission is in (the any-hit ray pushed below carries that
					// value plus its contribution, stored if it is unoccluded)
					float3 sv = f3(0.f, 0.f, 0.f);
					if (depth == 0) {
						float3 le = load3(sc.emission, prim);
						if (q.acc) {
							// the slot's first value: 0 + beta * le (the
							// megakernel's first L + beta * le)
							aux = (uint32_t)(gbase + (i - ncur));
							sv = f3(beta.x * le.x, beta.y * le.y, beta.z * le.z);
							slot_store(q.acc + 3 * (size_t)aux, sv);
						} else if (le.x != 0.f || le.y != 0.f || le.z != 0.f) {
							float *fb = (float *)ws.out + 3 * (size_t)pix;
							atomic_add_global(fb, beta.x * le.x);
							atomic_add_global(fb + 1, beta.y * le.y);
							atomic_add_global(fb + 2, beta.z * le.z);
						}
					} else if (q.acc) {
						aux = path_slot(ONE ? cam_map
						                        : SlotMap{(uint32_t)ws.first, (uint32_t)ws.p.width, (uint32_t)ws.tiles_x,
						                                  (uint32_t)ws.npix_pad, ws.p.sample_base,
						                                  FastDiv{ws.fd_width.m, ws.fd_width.s1, ws.fd_width.s2},
						                                  FastDiv{ws.fd_tiles.m, ws.fd_tiles.s1, ws.fd_tiles.s2}},
						                pix, smp);
					}
					float3 a = load3(sc.albedo, prim);
					so = P + ng * p.ray_eps;
					u32x4 rn = philox4x32_10(pix, smp, depth, 1u, p.seed, p.scene_id);
					if (sc.nl > 0) {
						NeeSample ns;
						if (sample_light(sc, so, ng, u01(rn.x), u01(rn.y), u01(rn.z), ns)) {
							float3 le = load3(sc.emission, ns.light);
							contrib = beta * a * (le * ns.Y);
							if (q.acc) // the slot's value: this vertex's emission at depth 0, else read here
								contrib = add_uncontracted(depth == 0 ? sv : slot_load(q.acc + 3 * (size_t)aux), contrib);
							shw = ns.wi;
							shmax = ns.tmax;
							push_sh = true;
						}
					}
					if ((int)depth + 1 < p.max_depth) {
						u32x4 rb = philox4x32_10(pix, smp, depth, 2u, p.seed, p.scene_id);
						wo = cosine_dir(ng, u01(rb.x), u01(rb.y));
						beta = beta * a;
						push_next = true;
						if ((int)depth >= p.rr_depth) {
							float qs = rr_survival(p, beta);
							if (!(u01(rn.w) < qs))
								push_next = false;
							else
								beta = beta * (1.0f / qs);
						}
					}
				} else if (depth == 0 && q.acc) {
					// a camera ray that left the scene: its slot holds zero
					slot_store(q.acc + 3 * (size_t)(uint32_t)(gbase + (i - ncur)), f3(0.f, 0.f, 0.f));
				}
					};
			// the single-scene instance reads its scene's fields with scalar
			// loads; the batch instance reads each lane's scene (a loop over
			// the scenes of a wave, to shade each with scalar loads, was 8%
			// slower on the 64-scene batch: one more barrier per block)
			if (depth != kDead)
				shade_vertex(((const cWScene *)scenes)[ONE ? one : (meta & 0xFFFFu)]);
			// Block-wide compaction: each wave's two push counts packed in one
			// word (both at most 64, block sums below 2^16); thread 0 sums them
			// and takes the block's entries with one 64-bit atomic; each wave adds
			// the words of the waves before it.
			const uint64_t msh = __ballot(push_sh), mnx = __ballot(push_next);
			if (lane == 0)
				wcnt[wave] = (unsigned)__popcll(msh) | ((unsigned)__popcll(mnx) << 16);
			__syncthreads();
			if (threadIdx.x == 0) {
				unsigned tot = 0;
				for (int v = 0; v < kShadeWaves; v++)
					tot += wcnt[v];
				unsigned long long old = 0;
				if (tot)
					old = atomicAdd(reinterpret_cast<unsigned long long *>(&c->next_count),
					                (unsigned long long)(tot >> 16) | ((unsigned long long)(tot & 0xFFFFu) << 32));
				bbase[0] = (unsigned)(old >> 32);
				bbase[1] = (unsigned)old;
			}
			__syncthreads();
			unsigned before = 0;
			for (int v = 0; v < wave; v++)
				before += wcnt[v];
			const unsigned ss = bbase[0] + (before & 0xFFFFu) + (unsigned)__popcll(msh & lt);
			const unsigned ns = bbase[1] + (before >> 16) + (unsigned)__popcll(mnx & lt);
			__syncthreads(); // wcnt and bbase are rewritten by the next iteration
			// the shadow record's .w: the path's radiance slot, or its pixel
			// without slots
			if (push_sh) {
				put4(q.so + ss, make_float4(so.x, so.y, so.z, q.acc ? __uint_as_float(aux) : o4.w));
				put4(q.sd + ss, make_float4(shw.x, shw.y, shw.z, shmax));
				put4(q.sc + ss, make_float4(contrib.x, contrib.y, contrib.z, __uint_as_float(meta & 0xFFFFu)));
			}
			if (push_next) {
				put4(NO + ns, make_float4(so.x, so.y, so.z, o4.w));
				put4(ND + ns, make_float4(wo.x, wo.y, wo.z, d4.w));
				put4(NB + ns, make_float4(beta.x, beta.y, beta.z, __uint_as_float((meta & 0xFFFFu) | ((depth + 1) << 16))));
			}
		}
	}
}

__global__ void k_resolve(float *out, size_t n, float scale) {
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
		out[i] *= scale;
}

// Each pixel of one scene: the sum of its paths' radiance slots in sample
// order (from zero, as the megakernel sums its samples), times `scale`.
// `acc` is the scene's first slot; a thread takes one tiled pixel index, so
// a wave reads 64 consecutive slots per sample.
__global__ void k_reduce_slots(float *out, const float *acc, int width, int height, int tiles_x, unsigned npix_pad,
                               int spp, float scale) {
	for (unsigned pp = blockIdx.x * blockDim.x + threadIdx.x; pp < npix_pad; pp += gridDim.x * blockDim.x) {
		const unsigned tile = pp >> 6, t = pp & 63u;
		const int x = (int)(tile % (unsigned)tiles_x) * kPixTile + (int)(t & 7u);
		const int y = (int)(tile / (unsigned)tiles_x) * kPixTile + (int)(t >> 3);
		if (x >= width || y >= height)
			continue;
		float3 sum = f3(0.f, 0.f, 0.f);
		const float *a = acc + 3 * (size_t)pp;
#pragma unroll 8
		for (int s = 0; s < spp; s++) {
			sum = sum + slot_load(a);
			a += 3 * (size_t)npix_pad;
		}
		float *o = out + 3 * ((size_t)y * width + x);
		o[0] = sum.x * scale;
		o[1] = sum.y * scale;
		o[2] = sum.z * scale;
	}
}

struct WaveWork {
	DevBuf state, hit, shadow, ctl, scenes, acc;
	unsigned *pinned_flag = nullptr; // host-pinned done, iter, done_iter; kept between renders
	// Instantiated graphs of the last graph render, with the launch arguments
	// and iteration counts they have baked in; reused while those stay the
	// same. `group` holds a group of iterations, `first` the iterations a
	// render enqueues before it first reads the done flag.
	struct Cached {
		hipGraphExec_t exec = nullptr;
		std::vector<uint64_t> key;
		void drop() {
			if (exec)
				(void)hipGraphExecDestroy(exec);
			exec = nullptr;
			key.clear();
		}
	} group, first;
	// Iterations the last render needed, with its launch arguments: a render
	// with the same ones enqueues that many before it first reads the flag.
	std::vector<uint64_t> pred_key;
	uint64_t pred_iters = 0;
	std::vector<hipEvent_t> events; // pool for per-stage timing
	hipStream_t capture_stream = nullptr;
	hipEvent_t join[2] = {nullptr, nullptr};
};

hipEvent_t pool_event(WaveWork &ww, size_t i) {
	while (ww.events.size() <= i) {
		hipEvent_t e;
		RN_HIP(hipEventCreate(&e));
		ww.events.push_back(e);
	}
	return ww.events[i];
}

WaveWork &work() {
	static WaveWork w;
	return w;
}

} // namespace

// The constants of FastDiv for a divisor d >= 1.
FastDiv make_fastdiv(uint32_t d) {
	if (d == 0)
		throw Error("wavefront: division by zero in the work-index map");
	uint32_t l = 0;
	while (l < 32 && (1ull << l) < d)
		l++;
	const uint64_t m = (((1ull << l) - d) << 32) / d + 1; // (2^l - d) < 2^32: no overflow
	return FastDiv{(uint32_t)m, l < 1 ? l : 1u, l > 1 ? l - 1 : 0u};
}

void wavefront_release() {
	WaveWork &ww = work();
	// the graphs refer to the queues freed below
	ww.group.drop();
	ww.first.drop();
	ww.pred_key.clear();
	ww.state.release();
	ww.hit.release();
	ww.shadow.release();
	ww.acc.release();
}

void fill_zero(void *p, size_t bytes, hipStream_t s) {
	size_t n = bytes / 4;
	if (n == 0)
		return;
	k_fill<<<(unsigned)std::min<size_t>(2048, (n + 255) / 256), 256, 0, s>>>(static_cast<uint32_t *>(p), n);
	RN_HIP(hipGetLastError());
}

void wavefront_render_on(const WaveItem *items, int count, const WaveConfig &cfg, DevCounters *cnt, hipStream_t s,
                         WaveStats &st, const WaveScratch *ext);

// Per path in flight: path state double buffered (6 float4), closest-hit
// record (float4 + int) and shadow ray (3 float4).
constexpr size_t kEntryBytes = 10 * sizeof(float4) + sizeof(int);
constexpr size_t kScratchAlign = 256;

static size_t align_up(size_t x) { return (x + kScratchAlign - 1) / kScratchAlign * kScratchAlign; }

static size_t scratch_fixed_bytes(int count) {
	return align_up(sizeof(WCtl)) + align_up((size_t)count * sizeof(WScene));
}

size_t wave_scratch_bytes(unsigned capacity, int count) {
	const size_t c = capacity;
	return align_up(c * 6 * sizeof(float4)) + align_up(c * (sizeof(float4) + sizeof(int))) +
	       align_up(c * 3 * sizeof(float4)) + scratch_fixed_bytes(count);
}

unsigned wave_scratch_capacity(size_t bytes, int count) {
	const size_t fixed = scratch_fixed_bytes(count) + 3 * kScratchAlign; // alignment slack of the three arrays
	if (bytes <= fixed)
		return 0;
	size_t cap = (bytes - fixed) / kEntryBytes;
	cap = std::min<size_t>(cap, 1u << 30);
	while (cap > 0 && wave_scratch_bytes((unsigned)cap, count) > bytes)
		cap--;
	return (unsigned)cap;
}

// The legacy default stream cannot be captured: a graph requested there runs
// on a library stream ordered after, and before, the caller's work.
void wavefront_render(const WaveItem *items, int count, const WaveConfig &cfg, DevCounters *cnt, hipStream_t s,
                      WaveStats &st, const WaveScratch *scratch) {
	if (scratch) {
		if (cfg.use_graph || cfg.time_stages)
			throw Error("wavefront: a render in caller memory takes no graph and no stage timing");
		wavefront_render_on(items, count, cfg, cnt, s, st, scratch);
		return;
	}
	if (!(cfg.use_graph && s == nullptr)) {
		wavefront_render_on(items, count, cfg, cnt, s, st, nullptr);
		return;
	}
	WaveWork &ww = work();
	if (!ww.capture_stream) {
		RN_HIP(hipStreamCreateWithFlags(&ww.capture_stream, hipStreamNonBlocking));
		RN_HIP(hipEventCreateWithFlags(&ww.join[0], hipEventDisableTiming));
		RN_HIP(hipEventCreateWithFlags(&ww.join[1], hipEventDisableTiming));
	}
	RN_HIP(hipEventRecord(ww.join[0], s));
	RN_HIP(hipStreamWaitEvent(ww.capture_stream, ww.join[0], 0));
	wavefront_render_on(items, count, cfg, cnt, ww.capture_stream, st, nullptr);
	RN_HIP(hipEventRecord(ww.join[1], ww.capture_stream));
	RN_HIP(hipStreamWaitEvent(s, ww.join[1], 0));
}

void wavefront_render_on(const WaveItem *items, int count, const WaveConfig &cfg, DevCounters *cnt, hipStream_t s,
                         WaveStats &st, const WaveScratch *ext) {
	if (count <= 0 || count > 65535)
		throw Error("wavefront: batch size must be in [1, 65535]");
	// a one-scene render marks padding entries with the sample kDeadSample
	if (count == 1 && (unsigned long long)items[0].params.sample_base + (unsigned long long)items[0].params.spp >=
	                      (unsigned long long)kDeadSample)
		throw Error("wavefront: sample_base + spp must stay below 2^32 - 1");
	std::vector<WScene> hs((size_t)count);
	unsigned long long total = 0;
	int max_depth = 0;
	for (int i = 0; i < count; i++) {
		const WaveItem &it = items[i];
		WScene &w = hs[(size_t)i];
		w.sc = it.scene;
		w.cam = it.cam;
		w.p = it.params;
		w.out = it.out;
		w.tiles_x = (it.params.width + kPixTile - 1) / kPixTile;
		int tiles_y = (it.params.height + kPixTile - 1) / kPixTile;
		w.npix_pad = w.tiles_x * tiles_y * kPixTile * kPixTile;
		w.fd_width = make_fastdiv((uint32_t)it.params.width);
		w.fd_tiles = make_fastdiv((uint32_t)w.tiles_x);
		w.fd_npix = make_fastdiv((uint32_t)w.npix_pad);
		w.first = total;
		w.count = (unsigned long long)w.npix_pad * (unsigned long long)it.params.spp;
		total += w.count;
		max_depth = std::max(max_depth, it.scene.depth);
	}
	WaveWork &ww = work();
	// Radiance slots (12 bytes per path of the work range) when the range
	// has 32-bit indices and its slots take at most a third of the free
	// device memory: no float atomics, and each pixel is the same sum as the
	// megakernel's. Otherwise (or in caller memory, or when asked) the
	// contributions go to the zeroed frame buffers with float atomics.
	bool slots = !ext && !cfg.fb_atomics && total <= 0xFFFFFFFFull;
	if (slots) {
		size_t freeb = 0, totb = 0;
		RN_HIP(hipMemGetInfo(&freeb, &totb));
		slots = total * 12ull <= (freeb + ww.acc.bytes) / 3;
	}
	if (!slots)
		for (int i = 0; i < count; i++)
			fill_zero(items[i].out, (size_t)items[i].params.width * items[i].params.height * 3 * 4, s);
	// Paths in flight. Each launch of the persistent trace kernels ends with
	// a tail in which its waves drain, so fewer, larger launches win: at
	// 1,048,588 triangles / 1080p / 64 spp a frame took 142 ms with 2M paths
	// in flight, 81 ms with 16M and 71 ms with the whole frame (133M) in
	// flight. By default up to 2^27 paths (164 bytes each, 22 GB) are kept,
	// but no more than a quarter of the free device memory.
	unsigned long long want;
	const unsigned ext_cap = ext ? wave_scratch_capacity(ext->bytes, count) : 0u;
	if (ext) {
		if (ext_cap < 64)
			throw Error("wavefront: the caller's work memory holds fewer than 64 paths");
		if (cfg.queue_paths > 0 && (unsigned long long)cfg.queue_paths > ext_cap)
			throw Error("wavefront: queue_paths " + std::to_string(cfg.queue_paths) +
			            " exceeds the caller's work memory (" + std::to_string(ext_cap) + " paths)");
		want = cfg.queue_paths > 0 ? (unsigned long long)cfg.queue_paths : ext_cap;
	} else if (cfg.queue_paths > 0) {
		want = std::min<unsigned long long>((unsigned long long)cfg.queue_paths, 1ull << 30);
	} else {
		size_t freeb = 0, totb = 0;
		RN_HIP(hipMemGetInfo(&freeb, &totb));
		freeb += ww.state.bytes + ww.hit.bytes + ww.shadow.bytes; // reusable
		want = std::min<unsigned long long>(1ull << 27, (unsigned long long)(freeb / 4 / kEntryBytes));
		want = std::max<unsigned long long>(want, 1ull << 16);
	}
	const unsigned cap = (unsigned)std::max<unsigned long long>(std::min(total, want), 64ull);
	WQ q;
	std::memset(&q, 0, sizeof q); // padding too: the bytes of q key the cached graph
	// carve the queues from the caller's memory, or reserve the library's own
	char *carve = ext ? static_cast<char *>(ext->base) : nullptr;
	auto take = [&](DevBuf &own, size_t bytes) -> void * {
		if (!ext)
			return own.reserve(bytes);
		void *p = carve;
		carve += align_up(bytes);
		return p;
	};
	float4 *st4 = static_cast<float4 *>(take(ww.state, (size_t)cap * 6 * sizeof(float4)));
	q.o[0] = st4;
	q.d[0] = st4 + cap;
	q.b[0] = st4 + 2 * (size_t)cap;
	q.o[1] = st4 + 3 * (size_t)cap;
	q.d[1] = st4 + 4 * (size_t)cap;
	q.b[1] = st4 + 5 * (size_t)cap;
	float4 *hit4 = static_cast<float4 *>(take(ww.hit, (size_t)cap * (sizeof(float4) + sizeof(int))));
	q.hitn = hit4;
	q.hitp = reinterpret_cast<int *>(hit4 + cap);
	float4 *sh4 = static_cast<float4 *>(take(ww.shadow, (size_t)cap * 3 * sizeof(float4)));
	q.so = sh4;
	q.sd = sh4 + cap;
	q.sc = sh4 + 2 * (size_t)cap;
	q.ctl = static_cast<WCtl *>(take(ww.ctl, sizeof(WCtl)));
	q.acc = slots ? static_cast<float *>(ww.acc.reserve((size_t)total * 3 * sizeof(float))) : nullptr;
	WScene *dsc = static_cast<WScene *>(take(ww.scenes, hs.size() * sizeof(WScene)));
	if (ext && (size_t)(carve - static_cast<char *>(ext->base)) > ext->bytes)
		throw Error("wavefront: internal error, caller work memory overrun"); // excluded by wave_scratch_capacity
	RN_HIP(hipMemcpyAsync(dsc, hs.data(), hs.size() * sizeof(WScene), hipMemcpyHostToDevice, s));
	q.scenes = dsc;
	q.nscenes = count;
	{
		int md = 1;
		for (int i = 0; i < count; i++)
			md = std::max(md, items[i].params.max_depth);
		if (md > kRing)
			throw Error("wavefront: max_depth above " + std::to_string(kRing));
		q.max_depth = md;
	}
	q.capacity = cap;
	q.cnt = cnt;

	int dev = 0, n_cu = 0;
	RN_HIP(hipGetDevice(&dev));
	RN_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
	const bool deep = max_depth - 1 > kLdsDepth;

	// Stage timing: one event at each boundary between stages (3 per
	// iteration; every event costs GPU time); interval k, from event k to
	// event k + 1, belongs to stage stage_of[k] and the intervals are summed
	// at the end. The advance kernel counts with the shade stage.
	const bool timed = cfg.time_stages && !cfg.use_graph;
	std::vector<int> stage_of;
	size_t nev = 0;
	auto mark = [&](int stage_before) {
		if (!timed)
			return;
		RN_HIP(hipEventRecord(pool_event(ww, nev++), s));
		stage_of.push_back(stage_before);
	};
	const int bs = cfg.block == 128 ? 128 : 256;
	// Trace kernels: enough resident waves to fill every SIMD; each wave
	// pulls chunks until the queue is exhausted. Shade and generate: plain
	// grid-stride launches.
	unsigned grid = (unsigned)std::max(1, n_cu) * (bs == 128 ? 16u : 8u);
	const unsigned chunks = (cap + kChunkMin - 1) / kChunkMin, waves_per_block = (unsigned)bs / 64u;
	grid = std::min<unsigned>(grid, (chunks + waves_per_block - 1) / waves_per_block);
	const unsigned sgrid = std::min<unsigned>((unsigned)std::max(1, n_cu) * 8u, (cap + 255) / 256);
	const unsigned shgrid = std::min<unsigned>((unsigned)std::max(1, n_cu) * (2048u / kShadeBlock),
	                                           (cap + kShadeBlock - 1) / kShadeBlock);
	// gen: the iteration's k_generate (left out when the queue took the whole
	// work range in the first iteration: later ones would generate nothing)
	auto iteration = [&](int cur, bool gen = true) {
		if (timed && nev == 0)
			RN_HIP(hipEventRecord(pool_event(ww, nev++), s));
		if (gen)
			k_generate<<<sgrid, 256, 0, s>>>(q, cur, q.scenes);
		mark(1);
#define RN_LAUNCH1(K, LDSD, SPD, DEEP, ONE, ...)                                                                      \
	do {                                                                                                               \
		if (DEEP && bs == 128)                                                                                         \
			K<LDSD, SPD, 128, ONE><<<grid, 128, 0, s>>>(__VA_ARGS__);                                                  \
		else if (DEEP)                                                                                                 \
			K<LDSD, SPD, 256, ONE><<<grid, 256, 0, s>>>(__VA_ARGS__);                                                  \
		else if (bs == 128)                                                                                            \
			K<LDSD, 0, 128, ONE><<<grid, 128, 0, s>>>(__VA_ARGS__);                                                    \
		else                                                                                                           \
			K<LDSD, 0, 256, ONE><<<grid, 256, 0, s>>>(__VA_ARGS__);                                                    \
	} while (0)
#define RN_LAUNCH(K, LDSD, SPD, DEEP, ...)                                                                            \
	do {                                                                                                               \
		RN_LAUNCH1(K, LDSD, SPD, DEEP, true, __VA_ARGS__);                                                             \
		if (count > 1)                                                                                                 \
			RN_LAUNCH1(K, LDSD, SPD, DEEP, false, __VA_ARGS__);                                                        \
	} while (0)
		RN_LAUNCH(k_trace, kLdsDepth, kSpillDepth, deep, q, cur);
		mark(0);
		k_shade<true><<<shgrid, kShadeBlock, 0, s>>>(q, cur, q.scenes);
		if (count > 1)
			k_shade<false><<<shgrid, kShadeBlock, 0, s>>>(q, cur, q.scenes);
#undef RN_LAUNCH
#undef RN_LAUNCH1
		k_advance<<<1, 1, 0, s>>>(q, 0);
		mark(1);
	};


	k_init<<<1, 1, 0, s>>>(q.ctl, total);
	k_advance<<<1, 1, 0, s>>>(q, 1);
	RN_HIP(hipGetLastError());
	// done, iter and done_iter are read into pinned memory (allocated once),
	// or into the stack when the render must allocate nothing
	unsigned stack_flag[4] = {0u, 0u, 0u, 0u};
	unsigned *flag = stack_flag;
	if (!ext) {
		if (!ww.pinned_flag) {
			if (fault_injected("alloc"))
				throw Error("injected allocation failure (RN_FAULT_INJECT=alloc): pinned flag");
			RN_HIP(hipHostMalloc(reinterpret_cast<void **>(&ww.pinned_flag), sizeof(unsigned) * 4,
			                     hipHostMallocDefault));
		}
		flag = ww.pinned_flag;
	}
	int per = cfg.iters_per_graph > 0 ? cfg.iters_per_graph : 8;
	per += per & 1;
	// Everything an iteration's launches bake in: the queue descriptor (all
	// pointers and sizes), the launch shapes and the kernel instances.
	std::vector<uint64_t> key;
	{
		const unsigned char *qb = reinterpret_cast<const unsigned char *>(&q);
		for (size_t off = 0; off < sizeof(WQ); off += 8) {
			uint64_t w = 0;
			std::memcpy(&w, qb + off, std::min<size_t>(8, sizeof(WQ) - off));
			key.push_back(w);
		}
		key.push_back(((uint64_t)grid << 32) | sgrid);
		key.push_back(((uint64_t)shgrid << 32) | (uint32_t)bs);
		key.push_back(((uint64_t)deep << 2) | ((uint64_t)(count == 1) << 1) | (uint64_t)cfg.use_graph);
		key.push_back((uint64_t)per);
	}
	// A graph of n iterations (n even: every iteration swaps the path
	// queues, so a graph starts and ends on the same one), cached in c.
	auto graph_of = [&](WaveWork::Cached &c, uint64_t n) {
		std::vector<uint64_t> k = key;
		k.push_back(n);
		if (c.exec && c.key == k)
			return c.exec;
		c.drop();
		hipGraph_t graph = nullptr;
		RN_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
		for (uint64_t it = 0; it < n; it++)
			iteration((int)(it & 1));
		RN_HIP(hipStreamEndCapture(s, &graph));
		hipGraphExec_t e = nullptr;
		const hipError_t rc = hipGraphInstantiate(&e, graph, nullptr, nullptr, 0);
		(void)hipGraphDestroy(graph);
		RN_HIP(rc);
		c.exec = e;
		c.key = std::move(k);
		return e;
	};
	uint64_t iters = 0;
	// Every iteration retires or advances each queued path, so the loop
	// ends within (batches of queue refills) x (max depth + 2) iterations.
	int maxd = 1;
	for (int i = 0; i < count; i++)
		maxd = std::max(maxd, items[i].params.max_depth);
	uint64_t limit = ((total + cap - 1) / cap + 1) * (uint64_t)(maxd + 2) + 64;
	if (fault_injected("wavefront_iterations"))
		limit = 0; // the check below fires after the first launch
	// Iterations enqueued before the done flag is first read: as many as the
	// last render with the same launch arguments needed (with a graph rounded
	// up to even, since a graph starts and ends on the same queue), else one
	// group; after that, one group per read of the flag.
	uint64_t n = (uint64_t)per;
	if (ww.pred_key == key && ww.pred_iters > 0)
		n = ww.pred_iters + (cfg.use_graph ? (ww.pred_iters & 1) : 0);
	while (true) {
		if (iters > limit)
			throw Error("wavefront: iteration limit exceeded (queue accounting error)");
		if (cfg.use_graph) {
			RN_HIP(hipGraphLaunch(iters == 0 ? graph_of(ww.first, n) : graph_of(ww.group, n), s));
		} else {
			for (uint64_t it = 0; it < n; it++)
				iteration((int)((iters + it) & 1), iters + it == 0 || total > cap);
		}
		iters += n;
		n = (uint64_t)per;
		// The images from the slots, enqueued before the done flag is read:
		// the reduce only reads the slots and overwrites the images, so when
		// more iterations turn out to be needed it simply runs again after
		// them (no host round trip between the last iteration and the reduce).
		if (slots)
			for (int i = 0; i < count; i++) {
				const WScene &w = hs[(size_t)i];
				const unsigned np = (unsigned)w.npix_pad;
				k_reduce_slots<<<(unsigned)std::min<size_t>(8192, (np + 255) / 256), 256, 0, s>>>(
				    items[i].out, q.acc + 3 * (size_t)w.first, items[i].params.width, items[i].params.height,
				    w.tiles_x, np, items[i].params.spp, 1.0f / (float)items[i].params.spp);
			}
		RN_HIP(hipGetLastError());
		RN_HIP(hipMemcpyAsync(flag, &q.ctl->done, 3 * sizeof(unsigned), hipMemcpyDeviceToHost, s));
		RN_HIP(hipStreamSynchronize(s));
		if (flag[0])
			break;
	}
	// iterations the render needed (the first advance only plans)
	const uint64_t needed = flag[2] > 0 ? (uint64_t)flag[2] - 1 : iters;
	ww.pred_key = key;
	ww.pred_iters = needed;
	st.iterations = needed;
	st.extend_launches = iters;
	if (timed) {
		double acc[3] = {0, 0, 0};
		for (size_t k = 0; k < stage_of.size(); k++) {
			float ms = 0;
			RN_HIP(hipEventElapsedTime(&ms, ww.events[k], ww.events[k + 1]));
			acc[stage_of[k]] += ms;
		}
		st.extend_ms = acc[0];
		st.shade_ms = acc[1];
		st.shadow_ms = acc[2];
	}
	if (!slots)
		for (int i = 0; i < count; i++) {
			size_t n = (size_t)items[i].params.width * items[i].params.height * 3;
			k_resolve<<<(unsigned)std::min<size_t>(2048, (n + 255) / 256), 256, 0, s>>>(
			    items[i].out, n, 1.0f / (float)items[i].params.spp);
		}
	RN_HIP(hipGetLastError());
}

} // namespace rn
