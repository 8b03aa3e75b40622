// Device helpers shared by the reference-mode renderers (pt_compat.hip:
// rn_render's path tracer and Whitted tracer; photon.hip: photon mapping and
// bidirectional path tracing): noise, surface colour, the default light and
// the sky, with the reference's FP32 operation order where a result must be
// bit-identical (the functions with contraction switched off).
#pragma once

#include "rn_compat.h"
#include "rn_internal.h"
#include "rn_kernels.h"

namespace rn {

namespace {

constexpr int kCBlock = 128;
constexpr int kTileW = 16, kTileH = 8;

// Ken Perlin's improved-noise permutation (the published table; the
// reference uses the same one, src/Noise.cpp:7-30).
__constant__ uint8_t c_perm[256] = {
    151, 160, 137, 91,  90,  15,  131, 13,  201, 95,  96,  53,  194, 233, 7,   225, 140, 36,  103, 30,  69,  142,
    8,   99,  37,  240, 21,  10,  23,  190, 6,   148, 247, 120, 234, 75,  0,   26,  197, 62,  94,  252, 219, 203,
    117, 35,  11,  32,  57,  177, 33,  88,  237, 149, 56,  87,  174, 20,  125, 136, 171, 168, 68,  175, 74,  165,
    71,  134, 139, 48,  27,  166, 77,  146, 158, 231, 83,  111, 229, 122, 60,  211, 133, 230, 220, 105, 92,  41,
    55,  46,  245, 40,  244, 102, 143, 54,  65,  25,  63,  161, 1,   216, 80,  73,  209, 76,  132, 187, 208, 89,
    18,  169, 200, 196, 135, 130, 116, 188, 159, 86,  164, 100, 109, 198, 173, 186, 3,   64,  52,  217, 226, 250,
    124, 123, 5,   202, 38,  147, 118, 126, 255, 82,  85,  212, 207, 206, 59,  227, 47,  16,  58,  17,  182, 189,
    28,  42,  223, 183, 170, 213, 119, 248, 152, 2,   44,  154, 163, 70,  221, 153, 101, 155, 167, 43,  172, 9,
    129, 22,  39,  253, 19,  98,  108, 110, 79,  113, 224, 232, 178, 185, 112, 104, 218, 246, 97,  228, 251, 34,
    242, 193, 238, 210, 144, 12,  191, 179, 162, 241, 81,  51,  145, 235, 249, 14,  239, 107, 49,  192, 214, 31,
    181, 199, 106, 157, 184, 84,  204, 176, 115, 121, 50,  45,  127, 4,   150, 254, 138, 236, 205, 93,  222, 114,
    67,  29,  24,  72,  243, 141, 128, 195, 78,  66,  215, 61,  156, 180};

RN_DEV int perm(int i) { return c_perm[i & 255]; }
RN_DEV float fade(float t) {
#pragma clang fp contract(off)
	return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f);
}
RN_DEV float lerpf(float t, float a, float b) {
#pragma clang fp contract(off)
	return a + t * (b - a);
}
RN_DEV float grad(int hash, float x, float y, float z) {
#pragma clang fp contract(off)
	int h = hash & 15;
	float u = h < 8 ? x : y;
	float v = h < 4 ? y : (h == 12 || h == 14 ? x : z);
	return ((h & 1) == 0 ? u : -u) + ((h & 2) == 0 ? v : -v);
}

// src/Noise.cpp:44-61
RN_DEV float perlin(float x, float y, float z) {
#pragma clang fp contract(off)
	int X = (int)floorf(x) & 255, Y = (int)floorf(y) & 255, Z = (int)floorf(z) & 255;
	x -= floorf(x);
	y -= floorf(y);
	z -= floorf(z);
	float u = fade(x), v = fade(y), w = fade(z);
	int A = perm(X) + Y, AA = perm(A) + Z, AB = perm(A + 1) + Z;
	int B = perm(X + 1) + Y, BA = perm(B) + Z, BB = perm(B + 1) + Z;
	return lerpf(w,
	             lerpf(v, lerpf(u, grad(perm(AA), x, y, z), grad(perm(BA), x - 1, y, z)),
	                   lerpf(u, grad(perm(AB), x, y - 1, z), grad(perm(BB), x - 1, y - 1, z))),
	             lerpf(v, lerpf(u, grad(perm(AA + 1), x, y, z - 1), grad(perm(BA + 1), x - 1, y, z - 1)),
	                   lerpf(u, grad(perm(AB + 1), x, y - 1, z - 1), grad(perm(BB + 1), x - 1, y - 1, z - 1))));
}

// src/Noise.cpp:246-255 and the wood colour of src/Renderer.cpp:566-571.
RN_DEV float3 procedural(float3 p) {
#pragma clang fp contract(off)
	float3 q = gscale(p, 3.0f);
	float sum = 0.0f, amp = 1.0f, freq = 1.0f, norm = 0.0f;
	for (int i = 0; i < 4; i++) {
		sum += amp * perlin(q.x * freq, q.y * freq, q.z * freq);
		norm += amp;
		amp *= 0.5f;
		freq *= 2.0f;
	}
	float n = 0.5f + 0.5f * (sum / norm);
	float rings = 0.5f + 0.5f * sinf((q.x + q.y) * 3.0f + n * 8.0f);
	return gadd(gscale(f3(60.0f, 30.0f, 12.0f), 1.0f - rings), gscale(f3(205.0f, 150.0f, 90.0f), rings));
}

RN_DEV float3 texel_rgb(const CompatDev &cd, int tex, float tu, float tv) {
	int4 ti = cd.tex_info[tex];
	uint32_t p = cd.texels[(size_t)ti.x + (size_t)compat_texel_of(ti.y, ti.z, tu, tv)];
	return f3((float)((p >> 16) & 0xFFu), (float)((p >> 8) & 0xFFu), (float)(p & 0xFFu));
}

// src/Renderer.cpp:595-612
RN_DEV void parallax(const CompatDev &cd, const float4 *tris, uint32_t leaf, int prim, int tex, float &tu, float &tv,
                     float3 view) {
#pragma clang fp contract(off)
	float3 v0 = xyz(tris[3 * leaf]), v1 = xyz(tris[3 * leaf + 1]), v2 = xyz(tris[3 * leaf + 2]);
	float3 e1 = gsub(v1, v0), e2 = gsub(v2, v0);
	const float *tp = cd.ctc + 6 * (size_t)prim;
	float d1x = tp[2] - tp[0], d1y = tp[3] - tp[1], d2x = tp[4] - tp[0], d2y = tp[5] - tp[1];
	float det = d1x * d2y - d2x * d1y;
	if (fabsf(det) < 1e-8f)
		return;
	float f = 1.0f / det;
	float3 T = gnormalize(gscale(gsub(gscale(e1, d2y), gscale(e2, d1y)), f));
	float3 B = gnormalize(gscale(gadd(gscale(e1, -d2x), gscale(e2, d1x)), f));
	float3 n = xyz(cd.cnorm[prim]);
	float vx = gdot(view, T), vy = gdot(view, B), vz = gdot(view, n);
	if (fabsf(vz) < 1e-4f)
		return;
	float h = texel_rgb(cd, tex, tu, tv).x / 255.0f;
	tu = tu - vx / vz * (h * 0.05f);
	tv = tv - vy / vz * (h * 0.05f);
}

// src/Renderer.cpp:642-660
RN_DEV float3 base_colour(const CompatDev &cd, const float4 *tris, uint32_t leaf, int prim, float u, float v,
                          float3 point, float3 view) {
	float4 col = cd.ccol[prim];
	int tex = __float_as_int(col.w);
	int mat = __float_as_int(cd.cnorm[prim].w);
	if (tex >= 0) {
		float tu, tv;
		compat_uv(cd, prim, u, v, tu, tv);
		if (mat == CM_PARALLAX)
			parallax(cd, tris, leaf, prim, tex, tu, tv, view);
		return texel_rgb(cd, tex, tu, tv);
	}
	if (mat == CM_PROCEDURAL)
		return procedural(point);
	return xyz(col);
}

RN_DEV float3 face_viewer(float3 n, float3 d) { return gdot(n, d) > 0.0f ? gscale(n, -1.0f) : n; }

// src/Renderer.cpp:778-788
RN_DEV float3 cos_hemi(float3 n, float u1, float u2) {
#pragma clang fp contract(off)
	float r = sqrtf(u1);
	float theta = 2.0f * 3.14159265f * u2;
	float x = r * cosf(theta), y = r * sinf(theta), z = sqrtf(fmaxf(0.0f, 1.0f - u1));
	float3 up = fabsf(n.z) < 0.99f ? f3(0, 0, 1) : f3(1, 0, 0);
	float3 t = gnormalize(gcross(up, n));
	float3 b = gcross(n, t);
	return gnormalize(gadd(gadd(gscale(t, x), gscale(b, y)), gscale(n, z)));
}

// glm::reflect (func_geometric.inl:104-110): I - N * dot(N, I) * 2.
RN_DEV float3 greflect(float3 i, float3 n) {
#pragma clang fp contract(off)
	float k = gdot(n, i);
	return gsub(i, gscale(gscale(n, k), 2.0f));
}
RN_DEV float3 reflectv(float3 i, float3 n) { return greflect(i, n); }

// Default light of renderPathTraced (src/Renderer.cpp:1129-1134, src/Light.h:13-29).
#define RN_LPOS f3(0.0f, 0.95f, 0.0f)
constexpr float kLRadius = 0.15f, kLIntensity = 40.0f, kLKq = 12.566370f;

// The 9 any-hit queries of visibility as one packet: a node is loaded once
// for every ray whose box test reached it, and a leaf triangle once for the
// rays that reached its leaf. Each ray still meets exactly the boxes and
// triangles of its own traversal (a child is entered with the rays that hit
// its box, and a ray leaves the packet at its first hit), so each answer is
// the one traverse<1, 1> gives and the count is exact. Returns the rays that
// reach their light point unoccluded. Entries of the depth-first stack hold
// a node and its ray mask; a level pops one entry and pushes at most 8.
#ifndef RN_VIS_PACKET
#define RN_VIS_PACKET 1
#endif
constexpr int kVisRays = 9;
constexpr int kPacketStack = 7 * kMaxBvhDepth + 1;

RN_DEV int packet_unoccluded(const Bvh8Node *__restrict__ nodes, const float4 *__restrict__ tris, float3 o,
                             const float3 (&dir)[kVisRays], const float (&tmax)[kVisRays], int ignore) {
	TRay r[kVisRays];
#pragma unroll
	for (int i = 0; i < kVisRays; i++)
		r[i] = make_tray(o, dir[i]);
	uint32_t active = (1u << kVisRays) - 1u;
	uint64_t st[kPacketStack];
	int sp = 0;
	st[sp++] = (uint64_t)active << 32; // the root, node 0
	while (sp > 0 && active) {
		const uint64_t e = st[--sp];
		const uint32_t node = (uint32_t)e;
		const uint32_t m = (uint32_t)(e >> 32) & active;
		if (!m)
			continue;
		uint32_t w[20];
		load_node_g(nodes, node, w);
		uint32_t th[kVisRays], tri_any = 0, smask[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
		for (int i = 0; i < kVisRays; i++) {
			th[i] = 0;
			if ((m >> i) & 1u) {
				uint32_t ih, t;
				intersect_children(w, r[i], tmax[i], ih, t, nullptr);
				const uint32_t slots = permute_oct(ih, r[i].oct); // back to slot order
				th[i] = t;
				tri_any |= t;
#pragma unroll
				for (int s = 0; s < 8; s++)
					smask[s] |= ((slots >> s) & 1u) << i;
			}
		}
		const uint32_t tri_base = w[5], tmask = w[6];
		while (tri_any && active) {
			const uint32_t k = __builtin_ctz(tri_any);
			tri_any &= tri_any - 1u;
			const uint32_t ti = leaf_record(tri_base, tmask, k);
			const float4 a = gload_f4(tris, 3 * ti), b4 = gload_f4(tris, 3 * ti + 1), c4 = gload_f4(tris, 3 * ti + 2);
			if (__float_as_int(a.w) == ignore)
				continue;
#pragma unroll
			for (int i = 0; i < kVisRays; i++) {
				if (((active & m) >> i) & 1u && ((th[i] >> k) & 1u)) {
					float tt, uu, vv;
					if (tri_compat(o, dir[i], xyz(a), xyz(b4), xyz(c4), tt, uu, vv) && tt < tmax[i])
						active &= ~(1u << i);
				}
			}
		}
		const uint32_t imask = w[3] >> 24, child_base = w[4];
#pragma unroll
		for (int s = 0; s < 8; s++) {
			const uint32_t cm = smask[s] & active;
			if (cm && sp < kPacketStack)
				st[sp++] = ((uint64_t)cm << 32) | (uint64_t)(child_base + __builtin_popcount(imask & ((1u << s) - 1u)));
		}
	}
	return __builtin_popcount(active);
}

// src/Renderer.cpp:483-497 (the default light's 9-point disk) in glm's
// operation order: normalize, cross, length and the division to / dist.
template <int LDS, int SPILL>
RN_DEV float visibility(const DevScene &sc, float3 point, int ignore, uint64_t *lds, int lane) {
#pragma clang fp contract(off)
	const float ks[9][2] = {{0, 0},       {1, 0},        {-1, 0},       {0, 1},        {0, -1},
	                        {0.7f, 0.7f}, {-0.7f, 0.7f}, {0.7f, -0.7f}, {-0.7f, -0.7f}};
	float3 lp = RN_LPOS;
	float3 axis = gnormalize(gsub(point, lp));
	float3 up = fabsf(axis.y) < 0.99f ? f3(0, 1, 0) : f3(1, 0, 0);
	float3 tangent = gnormalize(gcross(up, axis));
	float3 bitangent = gcross(axis, tangent);
	int un = 0;
	if (RN_VIS_PACKET) {
		float3 dir[kVisRays];
		float dist[kVisRays];
#pragma unroll
		for (int i = 0; i < kVisRays; i++) {
			float3 spos = gadd(lp, gscale(gadd(gscale(tangent, ks[i][0]), gscale(bitangent, ks[i][1])), kLRadius));
			float3 to = gsub(spos, point);
			dist[i] = sqrtf(gdot(to, to));
			dir[i] = f3(to.x / dist[i], to.y / dist[i], to.z / dist[i]);
		}
		un = packet_unoccluded(sc.nodes, sc.tris, point, dir, dist, ignore);
	} else {
#pragma unroll 1
		for (int i = 0; i < 9; i++) {
			float3 spos = gadd(lp, gscale(gadd(gscale(tangent, ks[i][0]), gscale(bitangent, ks[i][1])), kLRadius));
			float3 to = gsub(spos, point);
			float dist = sqrtf(gdot(to, to));
			float3 dir = f3(to.x / dist, to.y / dist, to.z / dist);
			Hit h = traverse<1, 1, LDS, SPILL>(sc.nodes, sc.tris, point, dir, dist, ignore, lds, kCBlock, lane);
			if (h.prim == 0)
				un++;
		}
	}
	return (float)un / 9.0f;
}

// directLight (src/Renderer.cpp:791-814) for the single default light.
template <int LDS, int SPILL>
RN_DEV float3 direct_light(const DevScene &sc, float3 point, float3 n, float3 albedo, int ignore, uint64_t *lds,
                           int lane) {
#pragma clang fp contract(off)
	float3 to = gsub(RN_LPOS, point);
	float dist = sqrtf(gdot(to, to));
	float3 ldir = gdiv(to, dist);
	float att = kLIntensity / (0.0f + 0.0f * dist + kLKq * dist * dist);
	float vis = visibility<LDS, SPILL>(sc, point, ignore, lds, lane);
	float inc = fmaxf(0.0f, gdot(n, ldir));
	float s = vis * 1.0f * att * inc;
	return gmul(albedo, f3(fminf(s, 1.0f), fminf(s, 1.0f), fminf(s, 1.0f)));
}

RN_DEV float3 gmix(float3 a, float3 b, float t) {
#pragma clang fp contract(off)
	return gadd(gscale(a, 1.0f - t), gscale(b, t));
}

RN_DEV float3 wh_environment(float3 d) { // src/Renderer.cpp:555-563
#pragma clang fp contract(off)
	float t = 0.5f * (d.y + 1.0f);
	return gmix(f3(200.0f, 210.0f, 235.0f), f3(70.0f, 120.0f, 205.0f), t);
}

} // namespace

} // namespace rn
