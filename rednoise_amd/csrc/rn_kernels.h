// Host launchers of the kernels (defined in the .hip files).
#pragma once

#include <hip/hip_runtime.h>

#include "rn_compat.h"
#include "rn_types.h"

namespace rn {

// Traversal stack: kLdsDepth entries per lane in LDS, kSpillDepth more in a
// lane-private array for BVHs deeper than kLdsDepth + 1 levels.
#ifndef RN_LDS_DEPTH
#define RN_LDS_DEPTH 12
#endif
constexpr int kLdsDepth = RN_LDS_DEPTH;
constexpr int kSpillDepth = 52;
constexpr int kMaxBvhDepth = kLdsDepth + kSpillDepth + 1;

// Counters written by the integrators (device memory, 64-bit). Block b adds
// into shard b % kCounterShards and the host sums the shards: atomics on one
// address serialise (about 11 ns each on MI355X), so a single counter bumped
// once per wave would cost tens of microseconds per launch.
constexpr int kCounterShards = 64;
struct CounterSet {
	unsigned long long closest, shadow, paths, pad;
	unsigned long long nodes_closest, tris_closest, nodes_shadow, tris_shadow;
};
struct DevCounters {
	CounterSet shard[kCounterShards];
	CounterSet total() const {
		CounterSet t{};
		for (const CounterSet &c : shard) {
			t.closest += c.closest;
			t.shadow += c.shadow;
			t.paths += c.paths;
			t.nodes_closest += c.nodes_closest;
			t.tris_closest += c.tris_closest;
			t.nodes_shadow += c.nodes_shadow;
			t.tris_shadow += c.tris_shadow;
		}
		return t;
	}
};

void launch_trace_rays(const DevScene &sc, int n, const float *o, const float *d, const float *tmax,
                       const int *ignore, int query, int tri, int *prim, float *t, float *u, float *v, hipStream_t s);
void launch_primary_compat(const DevScene &sc, const CompatDev &cd, int W, int H, float cam_z, int *prim, float *t,
                           float *u, float *v, int *texel, hipStream_t s);
void launch_philox(int n, const uint32_t *ctr, const uint32_t *key, uint32_t *out, hipStream_t s);

// Plain per-pixel megakernel of the physical estimator: out = W*H*3 mean radiance.
// With count_traversal the instrumented instance also sums BVH nodes visited
// and triangles tested per ray type into the counters.
void launch_mega(const DevScene &sc, const DevCamera &cam, const DevParams &p, float *out, DevCounters *cnt,
                 bool count_traversal, hipStream_t s);

// Path-replay adjoint. hot_slot[prim] in [0, kHotSlots) selects an LDS
// accumulator for heavily hit triangles, -1 otherwise.
constexpr int kHotSlots = 256;
void launch_adjoint(const DevScene &sc, const DevCamera &cam, const DevParams &p, const float *dl_dimage,
                    const int *hot_slot, const int *hot_prim, int n_hot, float *g_albedo, float *g_emission,
                    hipStream_t s);

// rn_render(RN_RAYTRACED): the reference's Whitted renderer, RGBA8 output.
void launch_whitted_render(const DevScene &sc, const CompatDev &cd, int W, int H, float cam_z, unsigned char *rgba,
                           hipStream_t s);

// Reference estimator (rn_render RN_PATHTRACED): RGBA8 output.
// rn_render(RN_PATHTRACED) with the reference's per-pixel std::mt19937 streams.
void launch_compat_render_mt(const DevScene &sc, const CompatDev &cd, int W, int H, float cam_z, int spp,
                             unsigned char *rgba, hipStream_t s);
void launch_compat_render(const DevScene &sc, const CompatDev &cd, int W, int H, float cam_z, int spp, uint32_t key0,
                          uint32_t key1, unsigned char *rgba, hipStream_t s);

// Progressive accumulation (float64 sum += weight * mean) and the fused
// tone-map + RGBA8 pack of sum * inv_count; the mean as FP32.
void launch_accum_add(double *sum, const float *mean, size_t n, double weight, hipStream_t s);
void launch_tonemap_rgba8(const double *sum, size_t npx, double inv_count, float exposure, float gamma,
                          unsigned char *rgba, hipStream_t s);
void launch_mean(const double *sum, size_t n, double inv_count, float *out, hipStream_t s);

} // namespace rn
