// Wavefront form of the physical estimator (Laine, Karras, Aila,
// "Megakernels Considered Harmful", HPG 2013), for one scene or a batch of
// scenes of uneven size that share one path queue.
//
// Per iteration three persistent, grid-stride kernels run:
//   extend  - appends freshly generated camera paths to the queue, traces the
//             closest hit of every queued path (SoA state), then shades it:
//             emission, NEE light sample, bounce and Russian roulette; shadow
//             rays and continuing paths are pushed with one wave64 ballot and
//             one atomic per wave (compaction keeps the next queue dense);
//   shadow  - any-hit queries; unoccluded contributions are added to the
//             frame buffer with float atomics;
//   advance - one lane: rotates the queues and plans the next generation so
//             the queue stays full until the work runs out.
// Queue sizes live in device memory, so the iteration can be captured once in
// a hipGraph and replayed; the host only polls a done flag between replays.
#include "rn_pathp.h"
#include "wavefront.h"

#include <algorithm>
#include <vector>

namespace rn {

namespace {

constexpr int kPixTile = 8; // 8x8 pixel tiles in generation order

struct WScene {
	DevScene sc;
	DevCamera cam;
	DevParams p;
	float *out;
	unsigned long long first, count; // global work range
	int tiles_x, npix_pad;
};

struct WCtl {
	unsigned cur_count, gen_count, next_count, sh_count;
	unsigned long long gen_base, work_next, work_total;
	unsigned done, iter;
};

struct WQ {
	float4 *o[2], *d[2], *b[2];
	float4 *so, *sd, *sc;
	WCtl *ctl;
	const WScene *scenes;
	int nscenes;
	unsigned capacity;
	DevCounters *cnt;
};

__global__ void k_fill(uint32_t *p, size_t n) {
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
		p[i] = 0u;
}

__global__ void k_init(WCtl *ctl, unsigned long long total) {
	ctl->cur_count = ctl->gen_count = ctl->next_count = ctl->sh_count = 0;
	ctl->gen_base = 0;
	ctl->work_next = 0;
	ctl->work_total = total;
	ctl->done = 0;
	ctl->iter = 0;
}

// Rotate queues and plan the next generation.
__global__ void k_advance(WQ q, int first) {
	WCtl *c = q.ctl;
	if (!first)
		c->cur_count = c->next_count;
	c->next_count = 0;
	c->sh_count = 0;
	unsigned long long remaining = c->work_total - c->work_next;
	unsigned long long room = (unsigned long long)(q.capacity - c->cur_count);
	unsigned gen = (unsigned)(remaining < room ? remaining : room);
	c->gen_base = c->work_next;
	c->gen_count = gen;
	c->work_next += gen;
	c->done = (c->cur_count == 0 && gen == 0) ? 1u : 0u;
	c->iter++;
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

// meta word: scene index (low 16 bits) | depth << 16; depth 255 = dead slot.
constexpr uint32_t kDead = 255u;

// Push with one atomic per wave: returns this lane's slot (valid when push).
RN_DEV unsigned wave_push(unsigned *counter, bool push) {
	uint64_t m = __ballot(push);
	if (m == 0)
		return 0;
	int lane = threadIdx.x & 63;
	int leader = __ffsll((unsigned long long)m) - 1;
	unsigned base = 0;
	if (lane == leader)
		base = atomicAdd(counter, (unsigned)__popcll(m));
	base = __shfl(base, leader);
	uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
	return base + (unsigned)__popcll(m & lt);
}

// Extend + shade: generate the new camera paths appended to the queue, trace
// every queued path's closest hit, then (all lanes of the wave together)
// add emission, draw the NEE light sample, the bounce and the roulette, and
// push the shadow ray and the continuing path.
template <int LDS, int SPILL, int BS>
__global__ __launch_bounds__(BS) void k_extend(WQ q, int cur) {
	__shared__ uint64_t stk[LDS * BS];
	WCtl *c = q.ctl;
	const unsigned ncur = c->cur_count, n = ncur + c->gen_count;
	const unsigned long long gbase = c->gen_base;
	float4 *O = q.o[cur], *D = q.d[cur], *B = q.b[cur];
	float4 *NO = q.o[cur ^ 1], *ND = q.d[cur ^ 1], *NB = q.b[cur ^ 1];
	unsigned long long ncl = 0, npaths = 0;
	for (unsigned base = blockIdx.x * BS; base < n; base += gridDim.x * BS) {
		const unsigned i = base + threadIdx.x;
		float4 o4 = make_float4(0.f, 0.f, 0.f, 0.f), d4 = o4, b4 = make_float4(0.f, 0.f, 0.f, 0.f);
		uint32_t meta = kDead << 16;
		if (i < n) {
			if (i >= ncur) { // new camera path
				unsigned long long g = gbase + (i - ncur);
				int si = find_scene(q.scenes, q.nscenes, g);
				const WScene &ws = q.scenes[si];
				unsigned long long local = g - ws.first;
				uint32_t smp = (uint32_t)(local / (unsigned long long)ws.npix_pad);
				uint32_t pp = (uint32_t)(local % (unsigned long long)ws.npix_pad);
				uint32_t tile = pp >> 6, t = pp & 63u;
				int x = (int)(tile % (uint32_t)ws.tiles_x) * kPixTile + (int)(t & 7u);
				int y = (int)(tile / (uint32_t)ws.tiles_x) * kPixTile + (int)(t >> 3);
				uint32_t pix = (uint32_t)(y * ws.p.width + x);
				meta = (uint32_t)si | (kDead << 16);
				if (x < ws.p.width && y < ws.p.height) {
					u32x4 r = philox4x32_10(pix, smp, 0u, 0u, ws.p.seed, ws.p.scene_id);
					float3 d = camera_dir(ws.cam, x, y, ws.p.width, ws.p.height, u01(r.x), u01(r.y));
					o4 = make_float4(ws.cam.eye[0], ws.cam.eye[1], ws.cam.eye[2], __uint_as_float(pix));
					d4 = make_float4(d.x, d.y, d.z, __uint_as_float(smp));
					b4 = make_float4(1.f, 1.f, 1.f, 0.f);
					meta = (uint32_t)si;
					npaths++;
				}
			} else {
				o4 = O[i];
				d4 = D[i];
				b4 = B[i];
				meta = __float_as_uint(b4.w);
			}
		}
		if (i < n && i >= ncur) { // publish the new path so it can be reloaded after traversal
			O[i] = o4;
			D[i] = d4;
			B[i] = make_float4(b4.x, b4.y, b4.z, __uint_as_float(meta));
		}
		const uint32_t depth = meta >> 16;
		bool push_sh = false, push_next = false;
		float3 so = f3(0.f, 0.f, 0.f), wo = so, beta = so, contrib = so, shw = so;
		float shmax = 0.f;
		if (depth != kDead) {
			const DevScene &sc0 = q.scenes[meta & 0xFFFFu].sc;
			Hit h = traverse<0, 0, LDS, SPILL>(sc0.nodes, sc0.tris, xyz(o4), xyz(d4), INFINITY, -1, stk, BS,
			                                   threadIdx.x);
			ncl++;
			if (h.prim >= 0) {
				// reload the state rather than keeping it live across traversal
				o4 = O[i];
				d4 = D[i];
				b4 = B[i];
				beta = xyz(b4);
				const WScene &ws = q.scenes[meta & 0xFFFFu];
				const DevScene &sc = ws.sc;
				const DevParams &p = ws.p;
				float3 o = xyz(o4), d = xyz(d4);
				uint32_t pix = __float_as_uint(o4.w), smp = __float_as_uint(d4.w);
				float3 P = o + d * h.t;
				float3 ng = facing_normal(sc.tris, h.leaf, d);
				if (depth == 0) {
					float3 le = load3(sc.emission, h.prim);
					if (le.x != 0.f || le.y != 0.f || le.z != 0.f) {
						float *fb = ws.out + 3 * (size_t)pix;
						atomicAdd(fb, beta.x * le.x);
						atomicAdd(fb + 1, beta.y * le.y);
						atomicAdd(fb + 2, beta.z * le.z);
					}
				}
				float3 a = load3(sc.albedo, h.prim);
				so = P + ng * p.ray_eps;
				u32x4 rn = philox4x32_10(pix, smp, depth, 1u, p.seed, p.scene_id);
				if (sc.nl > 0) {
					NeeSample ns;
					if (sample_light(sc, so, ng, u01(rn.x), u01(rn.y), u01(rn.z), ns)) {
						float3 le = load3(sc.emission, ns.light);
						contrib = beta * a * (le * ns.Y);
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
			}
		}
		unsigned ss = wave_push(&c->sh_count, push_sh);
		if (push_sh) {
			q.so[ss] = make_float4(so.x, so.y, so.z, o4.w);
			q.sd[ss] = make_float4(shw.x, shw.y, shw.z, shmax);
			q.sc[ss] = make_float4(contrib.x, contrib.y, contrib.z, __uint_as_float(meta & 0xFFFFu));
		}
		unsigned ns = wave_push(&c->next_count, push_next);
		if (push_next) {
			NO[ns] = make_float4(so.x, so.y, so.z, o4.w);
			ND[ns] = make_float4(wo.x, wo.y, wo.z, d4.w);
			NB[ns] = make_float4(beta.x, beta.y, beta.z, __uint_as_float((meta & 0xFFFFu) | ((depth + 1) << 16)));
		}
	}
	wave_add_u64(&q.cnt->closest, ncl);
	wave_add_u64(&q.cnt->paths, npaths);
}

template <int LDS, int SPILL, int BS>
__global__ __launch_bounds__(BS) void k_shadow(WQ q) {
	__shared__ uint64_t stk[LDS * BS];
	const unsigned n = q.ctl->sh_count;
	unsigned long long nsh = 0;
	for (unsigned base = blockIdx.x * BS; base < n; base += gridDim.x * BS) {
		unsigned i = base + threadIdx.x;
		if (i >= n)
			continue;
		float4 o4 = q.so[i], d4 = q.sd[i], c4 = q.sc[i];
		const WScene &ws = q.scenes[__float_as_uint(c4.w)];
		Hit h = traverse<1, 0, LDS, SPILL>(ws.sc.nodes, ws.sc.tris, xyz(o4), xyz(d4), d4.w, -1, stk, BS,
		                                   threadIdx.x);
		nsh++;
		if (h.prim == 0) {
			float *fb = ws.out + 3 * (size_t)__float_as_uint(o4.w);
			atomicAdd(fb, c4.x);
			atomicAdd(fb + 1, c4.y);
			atomicAdd(fb + 2, c4.z);
		}
	}
	wave_add_u64(&q.cnt->shadow, nsh);
}

__global__ void k_resolve(float *out, size_t n, float scale) {
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
		out[i] *= scale;
}

struct WaveWork {
	DevBuf state, hit, shadow, ctl, scenes, host_flag;
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

void fill_zero(void *p, size_t bytes, hipStream_t s) {
	size_t n = bytes / 4;
	if (n == 0)
		return;
	k_fill<<<(unsigned)std::min<size_t>(2048, (n + 255) / 256), 256, 0, s>>>(static_cast<uint32_t *>(p), n);
	RN_HIP(hipGetLastError());
}

void wavefront_render_on(const WaveItem *items, int count, const WaveConfig &cfg, DevCounters *cnt, hipStream_t s,
                         WaveStats &st);

// The legacy default stream cannot be captured: a graph requested there runs
// on a library stream ordered after, and before, the caller's work.
void wavefront_render(const WaveItem *items, int count, const WaveConfig &cfg, DevCounters *cnt, hipStream_t s,
                      WaveStats &st) {
	if (!(cfg.use_graph && s == nullptr)) {
		wavefront_render_on(items, count, cfg, cnt, s, st);
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
	wavefront_render_on(items, count, cfg, cnt, ww.capture_stream, st);
	RN_HIP(hipEventRecord(ww.join[1], ww.capture_stream));
	RN_HIP(hipStreamWaitEvent(s, ww.join[1], 0));
}

void wavefront_render_on(const WaveItem *items, int count, const WaveConfig &cfg, DevCounters *cnt, hipStream_t s,
                         WaveStats &st) {
	if (count <= 0 || count > 65535)
		throw Error("wavefront: batch size must be in [1, 65535]");
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
		w.first = total;
		w.count = (unsigned long long)w.npix_pad * (unsigned long long)it.params.spp;
		total += w.count;
		max_depth = std::max(max_depth, it.scene.depth);
		fill_zero(it.out, (size_t)it.params.width * it.params.height * 3 * 4, s);
	}
	WaveWork &ww = work();
	unsigned cap = (unsigned)std::min<unsigned long long>(total, 1ull << 21);
	if (cfg.queue_paths > 0)
		cap = (unsigned)std::min<unsigned long long>(total, (unsigned long long)cfg.queue_paths);
	cap = std::max(cap, 64u);
	WQ q;
	float4 *st4 = static_cast<float4 *>(ww.state.reserve((size_t)cap * 6 * sizeof(float4)));
	q.o[0] = st4;
	q.d[0] = st4 + cap;
	q.b[0] = st4 + 2 * (size_t)cap;
	q.o[1] = st4 + 3 * (size_t)cap;
	q.d[1] = st4 + 4 * (size_t)cap;
	q.b[1] = st4 + 5 * (size_t)cap;
	float4 *sh4 = static_cast<float4 *>(ww.shadow.reserve((size_t)cap * 3 * sizeof(float4)));
	q.so = sh4;
	q.sd = sh4 + cap;
	q.sc = sh4 + 2 * (size_t)cap;
	q.ctl = static_cast<WCtl *>(ww.ctl.reserve(sizeof(WCtl)));
	WScene *dsc = static_cast<WScene *>(ww.scenes.reserve(hs.size() * sizeof(WScene)));
	RN_HIP(hipMemcpyAsync(dsc, hs.data(), hs.size() * sizeof(WScene), hipMemcpyHostToDevice, s));
	q.scenes = dsc;
	q.nscenes = count;
	q.capacity = cap;
	q.cnt = cnt;

	hipDeviceProp_t prop;
	int dev = 0;
	RN_HIP(hipGetDevice(&dev));
	RN_HIP(hipGetDeviceProperties(&prop, dev));
	bool deep = max_depth - 1 > kLdsDepth;

	// Stage timing: events bracket each launch; pairs are summed at the end.
	const bool timed = cfg.time_stages && !cfg.use_graph;
	std::vector<int> stage_of;
	size_t nev = 0;
	auto mark = [&](int stage, bool begin) {
		if (!timed)
			return;
		RN_HIP(hipEventRecord(pool_event(ww, nev++), s));
		if (begin)
			stage_of.push_back(stage);
	};
	const int bs = cfg.block == 128 ? 128 : 256;
	unsigned grid = (unsigned)std::max(1, prop.multiProcessorCount) * (bs == 128 ? 16u : 8u);
	grid = std::min<unsigned>(grid, (cap + bs - 1) / bs);
	auto iteration = [&](int cur) {
		mark(0, true);
#define RN_LAUNCH(K, ...)                                                                                             \
	do {                                                                                                               \
		if (deep && bs == 128)                                                                                         \
			K<kLdsDepth, kSpillDepth, 128><<<grid, 128, 0, s>>>(__VA_ARGS__);                                          \
		else if (deep)                                                                                                 \
			K<kLdsDepth, kSpillDepth, 256><<<grid, 256, 0, s>>>(__VA_ARGS__);                                          \
		else if (bs == 128)                                                                                            \
			K<kLdsDepth, 0, 128><<<grid, 128, 0, s>>>(__VA_ARGS__);                                                    \
		else                                                                                                           \
			K<kLdsDepth, 0, 256><<<grid, 256, 0, s>>>(__VA_ARGS__);                                                    \
	} while (0)
		RN_LAUNCH(k_extend, q, cur);
		mark(0, false);
		mark(2, true);
		RN_LAUNCH(k_shadow, q);
		mark(2, false);
#undef RN_LAUNCH
		k_advance<<<1, 1, 0, s>>>(q, 0);
	};


	k_init<<<1, 1, 0, s>>>(q.ctl, total);
	k_advance<<<1, 1, 0, s>>>(q, 1);
	RN_HIP(hipGetLastError());
	unsigned *flag = nullptr;
	RN_HIP(hipHostMalloc(reinterpret_cast<void **>(&flag), sizeof(unsigned) * 2, hipHostMallocDefault));
	int per = cfg.iters_per_graph > 0 ? cfg.iters_per_graph : 8;
	per += per & 1;
	hipGraphExec_t exec = nullptr;
	hipGraph_t graph = nullptr;
	try {
		if (cfg.use_graph) {
			RN_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
			for (int it = 0; it < per; it++)
				iteration(it & 1);
			RN_HIP(hipStreamEndCapture(s, &graph));
			RN_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
		}
		uint64_t iters = 0;
		// Every iteration retires or advances each queued path, so the loop
		// ends within (batches of queue refills) x (max depth + 2) iterations.
		int maxd = 1;
		for (int i = 0; i < count; i++)
			maxd = std::max(maxd, items[i].params.max_depth);
		uint64_t limit = ((total + cap - 1) / cap + 1) * (uint64_t)(maxd + 2) + 64;
		while (true) {
			if (iters > limit)
				throw Error("wavefront: iteration limit exceeded (queue accounting error)");
			if (cfg.use_graph) {
				RN_HIP(hipGraphLaunch(exec, s));
			} else {
				for (int it = 0; it < per; it++)
					iteration(it & 1);
			}
			iters += (uint64_t)per;
			RN_HIP(hipGetLastError());
			RN_HIP(hipMemcpyAsync(flag, &q.ctl->done, sizeof(unsigned), hipMemcpyDeviceToHost, s));
			RN_HIP(hipStreamSynchronize(s));
			if (*flag)
				break;
		}
		st.iterations = iters;
		st.extend_launches = iters;
		if (timed) {
			double acc[3] = {0, 0, 0};
			for (size_t k = 0; k < stage_of.size(); k++) {
				float ms = 0;
				RN_HIP(hipEventElapsedTime(&ms, ww.events[2 * k], ww.events[2 * k + 1]));
				acc[stage_of[k]] += ms;
			}
			st.extend_ms = acc[0];
			st.shade_ms = acc[1];
			st.shadow_ms = acc[2];
		}
	} catch (...) {
		if (exec)
			(void)hipGraphExecDestroy(exec);
		if (graph)
			(void)hipGraphDestroy(graph);
		(void)hipHostFree(flag);
		throw;
	}
	if (exec)
		RN_HIP(hipGraphExecDestroy(exec));
	if (graph)
		RN_HIP(hipGraphDestroy(graph));
	RN_HIP(hipHostFree(flag));
	for (int i = 0; i < count; i++) {
		size_t n = (size_t)items[i].params.width * items[i].params.height * 3;
		k_resolve<<<(unsigned)std::min<size_t>(2048, (n + 255) / 256), 256, 0, s>>>(items[i].out, n,
		                                                                            1.0f / (float)items[i].params.spp);
	}
	RN_HIP(hipGetLastError());
}

} // namespace rn
