// Physical estimator, plain form: one lane per pixel looping over its samples
// (the "per-pixel megakernel"), and the path-replay adjoint built on the same
// per-sample code.
#include "rn_internal.h"
#include "rn_kernels.h"
#include "rn_pathp.h"

namespace rn {

namespace {

constexpr int kMBlock = 128; // 16 x 8 pixel tile, two waves
constexpr int kTileW = 16, kTileH = 8;
constexpr int kMaxAdjDepth = 16;
#ifndef RN_ADJ_WAVES_PER_CU
#define RN_ADJ_WAVES_PER_CU 256
#endif
constexpr int kAdjWavesPerCU = RN_ADJ_WAVES_PER_CU;

struct AdjRec {
	int tri;
	float3 a, X, beta;
	float invq;
};

struct GradSink {
	float *g_albedo, *g_emission; // global, 3 per triangle
	float *l_albedo, *l_emission; // LDS, 3 per hot slot
	const int *hot_slot;
};

RN_DEV void grad_add(float *glob, float *lds, const int *hot_slot, int prim, float3 g) {
	int hs = gload_i(hot_slot + prim);
	if (hs >= 0) {
		atomic_add_lds(&lds[3 * hs], g.x);
		atomic_add_lds(&lds[3 * hs + 1], g.y);
		atomic_add_lds(&lds[3 * hs + 2], g.z);
	} else {
		atomic_add_global(&glob[3 * prim], g.x);
		atomic_add_global(&glob[3 * prim + 1], g.y);
		atomic_add_global(&glob[3 * prim + 2], g.z);
	}
}

// L + c with each sum rounded on its own (no fused multiply-add with the
// product that made c): the wavefront integrator adds a stored contribution
// to a path's radiance slot, so both integrators round alike.
RN_DEV float3 add_uncontracted(float3 L, float3 c) {
#pragma clang fp contract(off)
	return f3(L.x + c.x, L.y + c.y, L.z + c.z);
}

// One camera path of the physical estimator. With ADJ the path is a replay:
// emission derivatives are pushed as they appear and the albedo derivatives
// by the reverse sweep over the vertex records at the end.
template <int LDS, int SPILL, bool ADJ, bool COUNT = false>
RN_DEV float3 trace_sample(const DevScene &sc, const DevCamera &cam, const DevParams &p, int x, int y, uint32_t smp,
                           uint64_t *lds, int lane, unsigned long long &ncl, unsigned long long &nsh, float3 wgt,
                           const GradSink *gs, uint32_t *cc = nullptr, uint32_t *cs = nullptr) {
	const uint32_t pix = (uint32_t)(y * p.width + x);
	u32x4 r = philox4x32_10(pix, smp, 0u, 0u, p.seed, p.scene_id);
	float3 o = f3(cam.eye[0], cam.eye[1], cam.eye[2]);
	float3 d = camera_dir(cam, x, y, p.width, p.height, u01(r.x), u01(r.y));
	float3 L = f3(0.f, 0.f, 0.f), beta = f3(1.f, 1.f, 1.f);
	AdjRec rec[ADJ ? kMaxAdjDepth : 1];
	int nrec = 0;
	for (int depth = 0; depth < p.max_depth; depth++) {
		ncl++;
		Hit h = traverse<0, 0, LDS, SPILL, COUNT>(sc.nodes, sc.tris, o, d, INFINITY, -1, lds, kMBlock, lane, cc);
		if (h.prim < 0)
			break;
		float3 P = o + d * h.t;
		float3 ng = facing_normal(sc.tris, h.leaf, d);
		if (depth == 0) {
			float3 le = load3(sc.emission, h.prim);
			L = L + beta * le;
			if (ADJ)
				grad_add(gs->g_emission, gs->l_emission, gs->hot_slot, h.prim, wgt);
		}
		float3 a = load3(sc.albedo, h.prim);
		float3 so = P + ng * p.ray_eps;
		u32x4 rn = philox4x32_10(pix, smp, (uint32_t)depth, 1u, p.seed, p.scene_id);
		float3 X = f3(0.f, 0.f, 0.f);
		if (sc.nl > 0) {
			NeeSample ns;
			if (sample_light(sc, so, ng, u01(rn.x), u01(rn.y), u01(rn.z), ns)) {
				nsh++;
				Hit sh = traverse<1, 0, LDS, SPILL, COUNT>(sc.nodes, sc.tris, so, ns.wi, ns.tmax, -1, lds, kMBlock,
				                                           lane, cs);
				if (sh.prim == 0) {
					float3 le = load3(sc.emission, ns.light);
					X = le * ns.Y;
					L = add_uncontracted(L, beta * a * X);
					if (ADJ)
						grad_add(gs->g_emission, gs->l_emission, gs->hot_slot, ns.light, wgt * beta * a * ns.Y);
				}
			}
		}
		if (ADJ) {
			rec[nrec].tri = h.prim;
			rec[nrec].a = a;
			rec[nrec].X = X;
			rec[nrec].beta = beta;
			rec[nrec].invq = 1.0f;
			nrec++;
		}
		if (depth + 1 >= p.max_depth)
			break;
		u32x4 rb = philox4x32_10(pix, smp, (uint32_t)depth, 2u, p.seed, p.scene_id);
		float3 wo = cosine_dir(ng, u01(rb.x), u01(rb.y));
		beta = beta * a;
		if (depth >= p.rr_depth) {
			float q = rr_survival(p, beta);
			if (!(u01(rn.w) < q))
				break;
			beta = beta * (1.0f / q);
			if (ADJ)
				rec[nrec - 1].invq = 1.0f / q;
		}
		o = so;
		d = wo;
	}
	if (ADJ) {
		float3 H = f3(0.f, 0.f, 0.f);
		for (int j = nrec - 1; j >= 0; j--) {
			float3 D = rec[j].X + H * rec[j].invq;
			grad_add(gs->g_albedo, gs->l_albedo, gs->hot_slot, rec[j].tri, wgt * rec[j].beta * D);
			H = rec[j].a * D;
		}
	}
	return L;
}

template <int LDS, int SPILL, bool COUNT>
__global__ __launch_bounds__(kMBlock) void k_mega(DevScene sc, DevCamera cam, DevParams p, float *out,
                                                  DevCounters *cnt) {
	__shared__ uint64_t stk[LDS * kMBlock];
	uint32_t b = xcd_block(blockIdx.x, gridDim.x);
	int tiles_x = (p.width + kTileW - 1) / kTileW;
	int x = (int)(b % tiles_x) * kTileW + (threadIdx.x % kTileW);
	int y = (int)(b / tiles_x) * kTileH + (threadIdx.x / kTileW);
	bool valid = x < p.width && y < p.height;
	unsigned long long ncl = 0, nsh = 0;
	uint32_t cc[2] = {0u, 0u}, cs[2] = {0u, 0u};
	if (valid) {
		float3 sum = f3(0.f, 0.f, 0.f);
		for (int s = 0; s < p.spp; s++)
			sum = sum + trace_sample<LDS, SPILL, false, COUNT>(sc, cam, p, x, y, p.sample_base + (uint32_t)s, stk, threadIdx.x, ncl,
			                                                   nsh, sum, nullptr, cc, cs);
		float inv = 1.0f / (float)p.spp;
		size_t o = 3 * ((size_t)y * p.width + x);
		out[o] = sum.x * inv;
		out[o + 1] = sum.y * inv;
		out[o + 2] = sum.z * inv;
	}
	CounterSet *cs_ = &cnt->shard[blockIdx.x % kCounterShards];
	wave_add_u64(&cs_->closest, ncl);
	wave_add_u64(&cs_->shadow, nsh);
	wave_add_u64(&cs_->paths, valid ? (unsigned long long)p.spp : 0ull);
	if (COUNT) {
		wave_add_u64(&cs_->nodes_closest, cc[0]);
		wave_add_u64(&cs_->tris_closest, cc[1]);
		wave_add_u64(&cs_->nodes_shadow, cs[0]);
		wave_add_u64(&cs_->tris_shadow, cs[1]);
	}
}

// Block (tile, g) replays samples [g * spl, (g + 1) * spl) of its tile's
// pixels: the gradients are sums of atomics, so a pixel's samples may be
// spread over blocks, and more, shorter blocks balance the load.
template <int LDS, int SPILL>
__global__ __launch_bounds__(kMBlock) void k_adjoint(DevScene sc, DevCamera cam, DevParams p, const float *dl_dimage,
                                                     const int *hot_slot, const int *hot_prim, int n_hot,
                                                     float *g_albedo, float *g_emission, int spl) {
	__shared__ uint64_t stk[LDS * kMBlock];
	__shared__ float la[3 * kHotSlots], le[3 * kHotSlots];
	for (int i = threadIdx.x; i < 3 * kHotSlots; i += kMBlock) {
		la[i] = 0.f;
		le[i] = 0.f;
	}
	__syncthreads();
	uint32_t b = xcd_block(blockIdx.x, gridDim.x);
	int tiles_x = (p.width + kTileW - 1) / kTileW;
	int x = (int)(b % tiles_x) * kTileW + (threadIdx.x % kTileW);
	int y = (int)(b / tiles_x) * kTileH + (threadIdx.x / kTileW);
	if (x < p.width && y < p.height) {
		size_t pi = 3 * ((size_t)y * p.width + x);
		float inv = 1.0f / (float)p.spp;
		float3 w = f3(dl_dimage[pi] * inv, dl_dimage[pi + 1] * inv, dl_dimage[pi + 2] * inv);
		GradSink gs{g_albedo, g_emission, la, le, hot_slot};
		unsigned long long ncl = 0, nsh = 0;
		const int s0 = (int)blockIdx.y * spl, s1 = min(p.spp, s0 + spl);
		if (w.x != 0.f || w.y != 0.f || w.z != 0.f)
			for (int s = s0; s < s1; s++)
				trace_sample<LDS, SPILL, true>(sc, cam, p, x, y, p.sample_base + (uint32_t)s, stk, threadIdx.x, ncl, nsh, w, &gs);
	}
	__syncthreads();
	for (int i = threadIdx.x; i < 3 * n_hot; i += kMBlock) {
		int prim = hot_prim[i / 3], c = i % 3;
		if (la[i] != 0.f)
			atomic_add_global(&g_albedo[3 * prim + c], la[i]);
		if (le[i] != 0.f)
			atomic_add_global(&g_emission[3 * prim + c], le[i]);
	}
}

unsigned mega_grid(const DevParams &p) {
	return (unsigned)(((p.width + kTileW - 1) / kTileW) * ((p.height + kTileH - 1) / kTileH));
}

} // namespace

void launch_mega(const DevScene &sc, const DevCamera &cam, const DevParams &p, float *out, DevCounters *cnt,
                 bool count_traversal, hipStream_t s) {
	bool deep = sc.depth - 1 > kLdsDepth;
	if (count_traversal) {
		if (deep)
			k_mega<kLdsDepth, kSpillDepth, true><<<mega_grid(p), kMBlock, 0, s>>>(sc, cam, p, out, cnt);
		else
			k_mega<kLdsDepth, 0, true><<<mega_grid(p), kMBlock, 0, s>>>(sc, cam, p, out, cnt);
	} else {
		if (deep)
			k_mega<kLdsDepth, kSpillDepth, false><<<mega_grid(p), kMBlock, 0, s>>>(sc, cam, p, out, cnt);
		else
			k_mega<kLdsDepth, 0, false><<<mega_grid(p), kMBlock, 0, s>>>(sc, cam, p, out, cnt);
	}
	RN_HIP(hipGetLastError());
}

void launch_adjoint(const DevScene &sc, const DevCamera &cam, const DevParams &p, const float *dl_dimage,
                    const int *hot_slot, const int *hot_prim, int n_hot, float *g_albedo, float *g_emission,
                    hipStream_t s) {
	if (p.max_depth > kMaxAdjDepth)
		throw Error("rn_adjoint supports max_depth <= 16");
	if (n_hot > kHotSlots)
		throw Error("too many hot triangles");
	// Sample groups: enough blocks for about kAdjWavesPerCU waves per CU
	// (16 are resident at a time), each group at least one sample.
	int dev = 0, n_cu = 0;
	RN_HIP(hipGetDevice(&dev));
	RN_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
	const long long waves = (long long)mega_grid(p) * (kMBlock / 64);
	const long long want = (long long)std::max(1, n_cu) * kAdjWavesPerCU;
	const int groups = (int)std::min<long long>(std::max(1, p.spp), std::max<long long>(1, (want + waves - 1) / waves));
	const int spl = (p.spp + groups - 1) / groups;
	const dim3 grid(mega_grid(p), (unsigned)((p.spp + spl - 1) / spl));
	if (sc.depth - 1 > kLdsDepth)
		k_adjoint<kLdsDepth, kSpillDepth>
		    <<<grid, kMBlock, 0, s>>>(sc, cam, p, dl_dimage, hot_slot, hot_prim, n_hot, g_albedo, g_emission, spl);
	else
		k_adjoint<kLdsDepth, 0>
		    <<<grid, kMBlock, 0, s>>>(sc, cam, p, dl_dimage, hot_slot, hot_prim, n_hot, g_albedo, g_emission, spl);
	RN_HIP(hipGetLastError());
}

} // namespace rn
