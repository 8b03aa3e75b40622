// Device pieces of the physical estimator shared by the megakernel, the
// wavefront stages and the adjoint. Sampling layout (identical on every path):
//   key = (seed, scene_id)
//   camera jitter      ctr = (pixel, sample, 0, 0)  words 0,1
//   NEE + roulette     ctr = (pixel, sample, b, 1)  words 0 (light), 1,2
//                                                   (point), 3 (roulette)
//   bounce direction   ctr = (pixel, sample, b, 2)  words 0,1
// where b is the vertex depth (0 = first hit).
#pragma once

#include "rn_device.h"

#define RN_INV_PI 0.31830988618379067f

RN_DEV float3 camera_dir(const DevCamera &c, int x, int y, int W, int H, float jx, float jy) {
	float fx = ((float)x + jx) / (float)W * 2.0f - 1.0f;
	float fy = 1.0f - ((float)y + jy) / (float)H * 2.0f;
	float3 du = f3(c.du[0], c.du[1], c.du[2]), dv = f3(c.dv[0], c.dv[1], c.dv[2]), w = f3(c.w[0], c.w[1], c.w[2]);
	return normalize3(du * fx + dv * fy + w);
}

// Cosine-weighted direction about n; branchless orthonormal basis of Duff et
// al., "Building an Orthonormal Basis, Revisited" (JCGT 6(1), 2017).
RN_DEV float3 cosine_dir(float3 n, float u1, float u2) {
	float r = sqrtf(u1);
	float phi = 6.28318530717958647692f * u2;
	float sp, cp;
	__sincosf(phi, &sp, &cp);
	float x = r * cp, y = r * sp, z = sqrtf(fmaxf(0.0f, 1.0f - u1));
	float sign = n.z >= 0.0f ? 1.0f : -1.0f;
	float a = -1.0f / (sign + n.z);
	float b = n.x * n.y * a;
	float3 t = f3(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
	float3 bb = f3(b, sign + n.y * n.y * a, -n.y);
	return t * x + bb * y + n * z;
}

RN_DEV int pick_light(const float *cdf, int nl, float u) {
	int lo = 0, hi = nl - 1;
	while (lo < hi) {
		int mid = (lo + hi) >> 1;
		if (u < gload_f(cdf + mid))
			hi = mid;
		else
			lo = mid + 1;
	}
	return lo;
}

// Scene buffers are global memory (see gload_f)
RN_DEV float3 load3(const float *p, int i) {
	return f3(gload_f(p + 3 * i), gload_f(p + 3 * i + 1), gload_f(p + 3 * i + 2));
}

// Geometric normal of leaf-ordered triangle `leaf`, facing against d.
// Unit geometric normal facing against d, from cross(v1 - v0, v2 - v0).
RN_DEV float3 facing_from_cross(float3 c, float3 d) {
	float3 ng = normalize3(c);
	return dot3(ng, d) > 0.0f ? ng * -1.0f : ng;
}

RN_DEV float3 facing_normal(const float4 *tris, uint32_t leaf, float3 d) {
	float4 a = tris[3 * leaf], b = tris[3 * leaf + 1], c = tris[3 * leaf + 2];
	return facing_from_cross(cross3(xyz(b) - xyz(a), xyz(c) - xyz(a)), d);
}

// Light sample for NEE from shading origin `so` with normal ng. Returns the
// shadow ray (wi, tmax) and the radiance factor X = Le * cos_s * cos_l / d^2 *
// A_total / pi, or false when the light faces away.
struct NeeSample {
	float3 wi;
	float tmax;
	float Y;   // geometric factor without Le
	int light; // light triangle id
};

RN_DEV bool sample_light(const DevScene &sc, float3 so, float3 ng, float ul, float u1, float u2, NeeSample &ns) {
	int li = pick_light(sc.light_cdf, sc.nl, ul);
	float4 g0 = gload_f4(sc.light_geom, 3 * li), g1 = gload_f4(sc.light_geom, 3 * li + 1),
	       g2 = gload_f4(sc.light_geom, 3 * li + 2);
	float su = sqrtf(u1);
	float b1 = su * (1.0f - u2), b2 = su * u2;
	float3 e1 = xyz(g1), e2 = xyz(g2);
	float3 q = xyz(g0) + e1 * b1 + e2 * b2;
	float3 nl = normalize3(cross3(e1, e2));
	float3 wi = q - so;
	float d2 = dot3(wi, wi);
	float dist = sqrtf(d2);
	wi = wi * (1.0f / dist);
	float cs = dot3(ng, wi);
	float cl = fabsf(dot3(nl, wi));
	if (!(cs > 0.0f && cl > 0.0f))
		return false;
	ns.wi = wi;
	ns.tmax = dist * 0.9999f;
	ns.Y = cs * cl / d2 * sc.total_light_area * RN_INV_PI;
	ns.light = gload_i(sc.light_tri + li);
	return true;
}

RN_DEV float rr_survival(const DevParams &p, float3 beta) {
	return p.rr_prob > 0.0f ? p.rr_prob : fminf(0.95f, maxc3(beta));
}

// Remap the hardware block id so that consecutive logical blocks land on the
// same XCD (blocks are dispatched round-robin over the 8 XCDs of MI355X):
// neighbouring image tiles then share one XCD's L2.
RN_DEV uint32_t xcd_block(uint32_t b, uint32_t nblocks) {
	const uint32_t nx = 8;
	uint32_t per = nblocks / nx;
	if (per == 0 || b >= per * nx)
		return b;
	return (b % nx) * per + b / nx;
}

// Wave-aggregated 64-bit counter increment.
RN_DEV void wave_add_u64(unsigned long long *dst, unsigned long long v) {
	for (int off = 32; off > 0; off >>= 1)
		v += __shfl_xor(v, off);
	if ((threadIdx.x & 63) == 0 && v)
		atomicAdd(dst, v);
}
