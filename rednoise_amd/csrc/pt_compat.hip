// rn_render(RN_PATHTRACED): the reference estimator (src/Renderer.cpp:
// 778-907,1119-1170) on the GPU. One lane per pixel loops over the samples;
// the recursion of pathTrace becomes a loop with a throughput. The per-pixel
// std::mt19937 becomes Philox4x32-10 with the layout
//   camera jitter: ctr = (pixel, sample, 0xFFFFFFFF, 0) words 0,1
//   level L:       ctr = (pixel, sample, L, 0) words 0,1 (hemisphere), 2 (Fresnel)
// Geometry (camera, triangle test) keeps the reference's FP32 operation order.
#include "rn_internal.h"
#include "rn_kernels.h"
#include "rn_pathp.h"

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
RN_DEV float fade(float t) { return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f); }
RN_DEV float lerpf(float t, float a, float b) { return a + t * (b - a); }
RN_DEV float grad(int hash, float x, float y, float z) {
	int h = hash & 15;
	float u = h < 8 ? x : y;
	float v = h < 4 ? y : (h == 12 || h == 14 ? x : z);
	return ((h & 1) == 0 ? u : -u) + ((h & 2) == 0 ? v : -v);
}

// src/Noise.cpp:44-61
RN_DEV float perlin(float x, float y, float z) {
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
	float3 q = p * 3.0f;
	float sum = 0.0f, amp = 1.0f, freq = 1.0f, norm = 0.0f;
	for (int i = 0; i < 4; i++) {
		sum += amp * perlin(q.x * freq, q.y * freq, q.z * freq);
		norm += amp;
		amp *= 0.5f;
		freq *= 2.0f;
	}
	float n = 0.5f + 0.5f * (sum / norm);
	float rings = 0.5f + 0.5f * sinf((q.x + q.y) * 3.0f + n * 8.0f);
	return f3(60.0f, 30.0f, 12.0f) * (1.0f - rings) + f3(205.0f, 150.0f, 90.0f) * rings;
}

RN_DEV float3 texel_rgb(const CompatDev &cd, int tex, float tu, float tv) {
	int4 ti = cd.tex_info[tex];
	uint32_t p = cd.texels[(size_t)ti.x + (size_t)compat_texel_of(ti.y, ti.z, tu, tv)];
	return f3((float)((p >> 16) & 0xFFu), (float)((p >> 8) & 0xFFu), (float)(p & 0xFFu));
}

// src/Renderer.cpp:595-612
RN_DEV void parallax(const CompatDev &cd, const float4 *tris, uint32_t leaf, int prim, int tex, float &tu, float &tv,
                     float3 view) {
	float3 v0 = xyz(tris[3 * leaf]), v1 = xyz(tris[3 * leaf + 1]), v2 = xyz(tris[3 * leaf + 2]);
	float3 e1 = v1 - v0, e2 = v2 - v0;
	const float *tp = cd.ctc + 6 * (size_t)prim;
	float d1x = tp[2] - tp[0], d1y = tp[3] - tp[1], d2x = tp[4] - tp[0], d2y = tp[5] - tp[1];
	float det = d1x * d2y - d2x * d1y;
	if (fabsf(det) < 1e-8f)
		return;
	float f = 1.0f / det;
	float3 T = normalize3((e1 * d2y - e2 * d1y) * f);
	float3 B = normalize3((e1 * -d2x + e2 * d1x) * f);
	float3 n = xyz(cd.cnorm[prim]);
	float vx = dot3(view, T), vy = dot3(view, B), vz = dot3(view, n);
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

RN_DEV float3 face_viewer(float3 n, float3 d) { return dot3(n, d) > 0.0f ? n * -1.0f : n; }

// src/Renderer.cpp:778-788
RN_DEV float3 cos_hemi(float3 n, float u1, float u2) {
	float r = sqrtf(u1);
	float theta = 2.0f * 3.14159265f * u2;
	float x = r * cosf(theta), y = r * sinf(theta), z = sqrtf(fmaxf(0.0f, 1.0f - u1));
	float3 up = fabsf(n.z) < 0.99f ? f3(0, 0, 1) : f3(1, 0, 0);
	float3 t = normalize3(cross3(up, n));
	float3 b = cross3(n, t);
	return normalize3(t * x + b * y + n * z);
}

RN_DEV float3 reflectv(float3 i, float3 n) { return i - n * (dot3(n, i) * 2.0f); }

// Default light of renderPathTraced (src/Renderer.cpp:1129-1134, src/Light.h:13-29).
#define RN_LPOS f3(0.0f, 0.95f, 0.0f)
constexpr float kLRadius = 0.15f, kLIntensity = 40.0f, kLKq = 12.566370f;

// src/Renderer.cpp:483-497 (the default light's 9-point disk) in glm's
// operation order: normalize, cross, length and the division to / dist.
template <int LDS, int SPILL>
RN_DEV float visibility(const DevScene &sc, float3 point, int ignore, uint64_t *lds, int lane) {
#pragma clang fp contract(off)
	const float ks[9][2] = {{0, 0},       {1, 0},        {-1, 0},       {0, 1},        {0, -1},
	                        {0.7f, 0.7f}, {-0.7f, 0.7f}, {0.7f, -0.7f}, {-0.7f, -0.7f}};
	float3 lp = RN_LPOS;
	float3 axis = gnormalize(point - lp);
	float3 up = fabsf(axis.y) < 0.99f ? f3(0, 1, 0) : f3(1, 0, 0);
	float3 tangent = gnormalize(gcross(up, axis));
	float3 bitangent = gcross(axis, tangent);
	int un = 0;
#pragma unroll 1
	for (int i = 0; i < 9; i++) {
		float3 spos = lp + (tangent * ks[i][0] + bitangent * ks[i][1]) * kLRadius;
		float3 to = spos - point;
		float dist = sqrtf(gdot(to, to));
		float3 dir = f3(to.x / dist, to.y / dist, to.z / dist);
		Hit h = traverse<1, 1, LDS, SPILL>(sc.nodes, sc.tris, point, dir, dist, ignore, lds, kCBlock, lane);
		if (h.prim == 0)
			un++;
	}
	return (float)un / 9.0f;
}

template <int LDS, int SPILL>
RN_DEV float3 direct_light(const DevScene &sc, float3 point, float3 n, float3 albedo, int ignore, uint64_t *lds,
                           int lane) {
	float3 to = RN_LPOS - point;
	float dist = sqrtf(dot3(to, to));
	float3 ldir = to * (1.0f / dist);
	float att = kLIntensity / (kLKq * dist * dist);
	float vis = visibility<LDS, SPILL>(sc, point, ignore, lds, lane);
	float inc = fmaxf(0.0f, dot3(n, ldir));
	float s = fminf(vis * att * inc, 1.0f);
	return albedo * s;
}

template <int LDS, int SPILL>
RN_DEV float3 compat_path(const DevScene &sc, const CompatDev &cd, float3 o, float3 d, uint32_t pix, uint32_t smp,
                          uint32_t k0, uint32_t k1, uint64_t *lds, int lane) {
	float3 L = f3(0, 0, 0), thr = f3(1, 1, 1);
	int depth = 4; // maxDepth, src/Renderer.cpp:1126
	for (uint32_t level = 0;; level++) {
		Hit h = traverse<0, 1, LDS, SPILL>(sc.nodes, sc.tris, o, d, INFINITY, -1, lds, kCBlock, lane);
		if (h.prim < 0) { // environment, src/Renderer.cpp:555-563
			float t = 0.5f * (d.y + 1.0f);
			float3 env = f3(200.0f, 210.0f, 235.0f) * (1.0f - t) + f3(70.0f, 120.0f, 205.0f) * t;
			L = L + thr * (env * (1.0f / 255.0f));
			break;
		}
		int ti = h.prim;
		u32x4 r = philox4x32_10(pix, smp, level, 0u, k0, k1);
		float3 point = o + d * h.t;
		float4 nm = cd.cnorm[ti];
		float3 N = xyz(nm);
		int mat = __float_as_int(nm.w);
		if (mat == CM_EMISSIVE) {
			L = L + thr * (xyz(cd.ccol[ti]) * (4.0f / 255.0f));
			break;
		}
		if (mat == CM_MIRROR && depth > 0) {
			float3 n = face_viewer(N, d);
			float3 rr = reflectv(d, n);
			o = point + rr * 1e-4f;
			d = rr;
			depth--;
			continue;
		}
		if ((mat == CM_GLASS || mat == CM_DISPERSIVE) && depth > 0) {
			float3 n = N;
			float cosi = dot3(d, n);
			float etai = 1.0f, etat = 1.5f;
			if (cosi < 0.0f) {
				cosi = -cosi;
			} else {
				etai = 1.5f;
				etat = 1.0f;
				n = n * -1.0f;
			}
			float r0 = (etai - etat) / (etai + etat);
			r0 *= r0;
			float x = 1.0f - cosi;
			float fres = r0 + (1.0f - r0) * (x * x * x * x * x);
			float eta = etai / etat;
			float dv = dot3(n, d);
			float k = 1.0f - eta * eta * (1.0f - dv * dv);
			float3 refr = k >= 0.0f ? d * eta - n * (eta * dv + sqrtf(k)) : f3(0, 0, 0);
			float3 nd = (sqrtf(dot3(refr, refr)) < 1e-6f || u01(r.z) < fres) ? reflectv(d, n) : refr;
			o = point + nd * 1e-4f;
			d = nd;
			depth--;
			continue;
		}
		if (mat == CM_METAL && depth > 0) {
			float3 n = face_viewer(N, d);
			float3 refl = reflectv(d, n);
			float3 diff = cos_hemi(n, u01(r.x), u01(r.y));
			const float r2 = 0.2f * 0.2f; // ModelTriangle::roughness default (framework/ModelTriangle.h:31)
			float3 nd = normalize3(refl * (1.0f - r2) + diff * r2);
			if (dot3(nd, n) <= 0.0f)
				nd = refl;
			float3 albedo = base_colour(cd, sc.tris, h.leaf, ti, h.u, h.v, point, d * -1.0f) * (1.0f / 255.0f);
			thr = thr * albedo;
			o = point + nd * 1e-4f;
			d = nd;
			depth--;
			continue;
		}
		if (mat == CM_SUBSURFACE) { // src/Renderer.cpp:872-897
			float3 n = face_viewer(N, d);
			float3 albedo = base_colour(cd, sc.tris, h.leaf, ti, h.u, h.v, point, d * -1.0f) * (1.0f / 255.0f);
			float3 to = RN_LPOS - point;
			float dist = sqrtf(dot3(to, to));
			float3 ldir = to * (1.0f / dist);
			float att = kLIntensity / (kLKq * dist * dist);
			float wd = fmaxf(0.0f, (dot3(n, ldir) + 0.5f) / 1.5f);
			float3 shade = albedo * (visibility<LDS, SPILL>(sc, point, ti, lds, lane) * att * wd);
			Hit ex = traverse<0, 1, LDS, SPILL>(sc.nodes, sc.tris, point - n * 1e-3f, ldir, INFINITY, ti, lds, kCBlock,
			                                    lane);
			float thick = ex.prim >= 0 ? ex.t : 0.0f;
			shade = shade + albedo * (att * expf(-3.0f * thick) * 0.6f);
			float3 direct = f3(fminf(shade.x, 1.5f), fminf(shade.y, 1.5f), fminf(shade.z, 1.5f));
			L = L + thr * direct;
			if (depth <= 0)
				break;
			float3 sb = cos_hemi(n, u01(r.x), u01(r.y));
			thr = thr * albedo;
			o = point + sb * 1e-4f;
			d = sb;
			depth--;
			continue;
		}
		// Diffuse and every material that falls through to it
		// (src/Renderer.cpp:899-906).
		float3 n = face_viewer(N, d);
		float3 albedo = base_colour(cd, sc.tris, h.leaf, ti, h.u, h.v, point, d * -1.0f) * (1.0f / 255.0f);
		L = L + thr * direct_light<LDS, SPILL>(sc, point, n, albedo, ti, lds, lane);
		if (depth <= 0)
			break;
		float3 bounce = cos_hemi(n, u01(r.x), u01(r.y));
		thr = thr * albedo;
		o = point + bounce * 1e-4f;
		d = bounce;
		depth--;
	}
	return L;
}

template <int LDS, int SPILL>
__global__ __launch_bounds__(kCBlock) void k_compat(DevScene sc, CompatDev cd, int W, int H, float cam_z, int spp,
                                                    uint32_t k0, uint32_t k1, unsigned char *rgba) {
	__shared__ uint64_t stk[LDS * kCBlock];
	uint32_t b = xcd_block(blockIdx.x, gridDim.x);
	int tiles_x = (W + kTileW - 1) / kTileW;
	int x = (int)(b % tiles_x) * kTileW + (threadIdx.x % kTileW);
	int y = (int)(b / tiles_x) * kTileH + (threadIdx.x / kTileW);
	if (x >= W || y >= H)
		return;
	uint32_t pix = (uint32_t)(y * W + x);
	float3 sum = f3(0, 0, 0);
	for (int s = 0; s < spp; s++) {
		u32x4 r = philox4x32_10(pix, (uint32_t)s, 0xFFFFFFFFu, 0u, k0, k1);
		float3 o, d;
		compat_primary_ray(x, y, W, H, cam_z, u01(r.x) - 0.5f, u01(r.y) - 0.5f, o, d);
		sum = sum + compat_path<LDS, SPILL>(sc, cd, o, d, pix, (uint32_t)s, k0, k1, stk, threadIdx.x);
	}
	float3 c = sum * (1.0f / (float)spp) * 255.0f;
	int rr = min(255, (int)c.x), gg = min(255, (int)c.y), bb = min(255, (int)c.z);
	uchar4 px = make_uchar4((unsigned char)(rr & 0xFF), (unsigned char)(gg & 0xFF), (unsigned char)(bb & 0xFF), 255);
	reinterpret_cast<uchar4 *>(rgba)[pix] = px;
}

// ------------------------------------------------------------ Whitted --
// rn_render(RN_RAYTRACED): renderRaytraced (src/Renderer.cpp:909-948) with
// the arguments rn_render passes (src/capi.cpp:59-61): Phong shading, the
// default soft point light, maxDepth 4, no fog. traceRay's recursion
// (src/Renderer.cpp:735-772) is a depth-first walk over an explicit stack of
// weighted rays: every branch (mirror, the two glass rays weighted by
// Fresnel, the three dispersive channels, metal tinted by its colour) is
// linear in the colours returned below it, so the pixel is the sum of
// weight x leaf colour. The leaf colours keep glm's FP32 operation order.

struct WhRay {
	float3 o, d, w;
	int depth;
};
constexpr int kWhStack = 24; // depth 4 x (6 dispersive children - 1) + 1

RN_DEV float3 gmix(float3 a, float3 b, float t) {
#pragma clang fp contract(off)
	return a * (1.0f - t) + b * t;
}

RN_DEV float3 wh_environment(float3 d) { // src/Renderer.cpp:555-563
#pragma clang fp contract(off)
	float t = 0.5f * (d.y + 1.0f);
	return gmix(f3(200.0f, 210.0f, 235.0f), f3(70.0f, 120.0f, 205.0f), t);
}

RN_DEV float3 greflect(float3 i, float3 n) {
#pragma clang fp contract(off)
	float k = gdot(n, i);
	return i - n * k * 2.0f;
}

// src/Renderer.cpp:574-581
RN_DEV float3 bump_normal(float3 n, float3 p) {
#pragma clang fp contract(off)
	float3 q = p * 6.0f;
	const float e = 0.15f;
	float3 g = f3(perlin(q.x + e, q.y, q.z) - perlin(q.x - e, q.y, q.z),
	              perlin(q.x, q.y + e, q.z) - perlin(q.x, q.y - e, q.z),
	              perlin(q.x, q.y, q.z + e) - perlin(q.x, q.y, q.z - e));
	return gnormalize(n + g * 2.0f);
}

// src/Renderer.cpp:618-639
RN_DEV float3 tangent_normal(const CompatDev &cd, const float4 *tris, uint32_t leaf, int prim, int tex, float u,
                             float v, float3 gn) {
#pragma clang fp contract(off)
	float3 v0 = xyz(tris[3 * leaf]), v1 = xyz(tris[3 * leaf + 1]), v2 = xyz(tris[3 * leaf + 2]);
	float3 e1 = v1 - v0, e2 = v2 - v0;
	const float *tp = cd.ctc + 6 * (size_t)prim;
	float d1x = tp[2] - tp[0], d1y = tp[3] - tp[1], d2x = tp[4] - tp[0], d2y = tp[5] - tp[1];
	float det = d1x * d2y - d2x * d1y;
	if (fabsf(det) < 1e-8f)
		return gn;
	float f = 1.0f / det;
	float3 T = gnormalize((e1 * d2y - e2 * d1y) * f);
	float3 B = gnormalize((e1 * -d2x + e2 * d1x) * f);
	int4 ti = cd.tex_info[tex];
	float e = 1.0f / (float)ti.y;
	auto lum = [&](float uu, float vv) {
		float3 c = texel_rgb(cd, tex, uu, vv);
		return (c.x + c.y + c.z) / 765.0f;
	};
	float hu = lum(u + e, v) - lum(u - e, v);
	float hv = lum(u, v + e) - lum(u, v - e);
	float3 tn = gnormalize(f3(-hu * 4.0f, -hv * 4.0f, 1.0f));
	return gnormalize(T * tn.x + B * tn.y + gn * tn.z);
}

// shadeSurface (src/Renderer.cpp:503-534) for the default light.
template <int LDS, int SPILL>
RN_DEV float3 shade_surface(const DevScene &sc, float3 point, float3 n, float3 view, int ignore, float3 base,
                            uint64_t *lds, int lane) {
#pragma clang fp contract(off)
	const float ambient = 0.2f;
	float3 to = RN_LPOS - point;
	float dist = sqrtf(gdot(to, to));
	float3 ldir = f3(to.x / dist, to.y / dist, to.z / dist);
	float att = kLIntensity / (0.0f + 0.0f * dist + kLKq * dist * dist);
	float vis = visibility<LDS, SPILL>(sc, point, ignore, lds, lane);
	float inc = fmaxf(0.0f, gdot(n, ldir));
	float bright = vis * 1.0f * att * inc;
	float3 diff = f3(0.0f, 0.0f, 0.0f) + f3(1.0f, 1.0f, 1.0f) * bright;
	float spec = 0.0f;
	if (bright > 0.0f) {
		float3 rd = n * (2.0f * gdot(n, ldir)) - ldir;
		spec += vis * 1.0f * powf(fmaxf(0.0f, gdot(rd, view)), 64.0f);
	}
	float3 lit = diff + f3(ambient, ambient, ambient);
	lit = f3(fminf(1.0f, lit.x), fminf(1.0f, lit.y), fminf(1.0f, lit.z));
	float s255 = 255.0f * fminf(1.0f, spec);
	return f3(base.x * lit.x + s255, base.y * lit.y + s255, base.z * lit.z + s255);
}

// shadeDiffuse with Phong normals (src/Renderer.cpp:662-699).
template <int LDS, int SPILL>
RN_DEV float3 shade_diffuse(const DevScene &sc, const CompatDev &cd, const Hit &h, float3 point, float3 d,
                            uint64_t *lds, int lane) {
#pragma clang fp contract(off)
	const int prim = h.prim;
	float w0 = 1.0f - h.u - h.v, w1 = h.u, w2 = h.v;
	float3 view = d * -1.0f;
	float3 base = base_colour(cd, sc.tris, h.leaf, prim, h.u, h.v, point, view);
	const float *vn = cd.vnorm + 9 * (size_t)prim;
	float3 n0 = f3(vn[0], vn[1], vn[2]), n1 = f3(vn[3], vn[4], vn[5]), n2 = f3(vn[6], vn[7], vn[8]);
	float3 n = gnormalize(n0 * w0 + n1 * w1 + n2 * w2);
	n = face_viewer(n, d);
	int mat = __float_as_int(cd.cnorm[prim].w);
	if (mat == CM_BUMP)
		n = bump_normal(n, point);
	int tex = __float_as_int(cd.ccol[prim].w);
	if (mat == CM_NORMALMAP && tex >= 0) {
		float tu, tv;
		compat_uv(cd, prim, h.u, h.v, tu, tv);
		n = tangent_normal(cd, sc.tris, h.leaf, prim, tex, tu, tv, n);
	}
	return shade_surface<LDS, SPILL>(sc, point, n, view, prim, base, lds, lane);
}

// glassTrace (src/Renderer.cpp:710-731): push the reflected and refracted
// rays weighted by Fresnel (the reflected ray alone on total internal
// reflection).
RN_DEV void glass_push(WhRay *st, int &sp, float3 point, float3 d, float3 N, float3 w, int depth, float ior) {
#pragma clang fp contract(off)
	float3 n = N;
	float cosi = gdot(d, n);
	float etai = 1.0f, etat = ior;
	if (cosi < 0.0f) {
		cosi = -cosi;
	} else {
		etai = ior;
		etat = 1.0f;
		n = n * -1.0f;
	}
	float3 refl = greflect(d, n);
	float eta = etai / etat;
	float dv = gdot(n, d);
	float k = 1.0f - eta * eta * (1.0f - dv * dv);
	float3 refr = k >= 0.0f ? d * eta - n * (eta * dv + sqrtf(k)) : f3(0.0f, 0.0f, 0.0f);
	if (sqrtf(gdot(refr, refr)) < 1e-6f) {
		if (sp < kWhStack)
			st[sp++] = WhRay{point + refl * 1e-4f, refl, w, depth - 1};
		return;
	}
	float r0 = (etai - etat) / (etai + etat);
	r0 *= r0;
	float fres = r0 + (1.0f - r0) * powf(1.0f - cosi, 5.0f);
	if (sp < kWhStack)
		st[sp++] = WhRay{point + refr * 1e-4f, refr, w * (1.0f - fres), depth - 1};
	if (sp < kWhStack)
		st[sp++] = WhRay{point + refl * 1e-4f, refl, w * fres, depth - 1};
}

template <int LDS, int SPILL>
RN_DEV float3 whitted_pixel(const DevScene &sc, const CompatDev &cd, float3 o, float3 d, uint64_t *lds, int lane) {
#pragma clang fp contract(off)
	WhRay st[kWhStack];
	int sp = 0;
	st[sp++] = WhRay{o, d, f3(1.0f, 1.0f, 1.0f), 4};
	float3 col = f3(0.0f, 0.0f, 0.0f);
	while (sp > 0) {
		const WhRay r = st[--sp];
		Hit h = traverse<0, 1, LDS, SPILL>(sc.nodes, sc.tris, r.o, r.d, INFINITY, -1, lds, kCBlock, lane);
		if (h.prim < 0) {
			col = col + r.w * wh_environment(r.d);
			continue;
		}
		const int ti = h.prim;
		const float3 point = r.o + r.d * h.t;
		const float4 nm = cd.cnorm[ti];
		const float3 N = xyz(nm);
		const int mat = __float_as_int(nm.w);
		const float3 kd = xyz(cd.ccol[ti]);
		if (mat == CM_EMISSIVE) {
			col = col + r.w * (kd * 1.3f);
			continue;
		}
		if (r.depth > 0 && mat == CM_MIRROR) {
			float3 n = face_viewer(N, r.d);
			float3 rr = greflect(r.d, n);
			if (sp < kWhStack)
				st[sp++] = WhRay{point + rr * 1e-4f, rr, r.w, r.depth - 1};
			continue;
		}
		if (r.depth > 0 && mat == CM_GLASS) {
			glass_push(st, sp, point, r.d, N, r.w, r.depth, 1.5f);
			continue;
		}
		if (r.depth > 0 && mat == CM_DISPERSIVE) {
			glass_push(st, sp, point, r.d, N, r.w * f3(1.0f, 0.0f, 0.0f), r.depth, 1.50f);
			glass_push(st, sp, point, r.d, N, r.w * f3(0.0f, 1.0f, 0.0f), r.depth, 1.53f);
			glass_push(st, sp, point, r.d, N, r.w * f3(0.0f, 0.0f, 1.0f), r.depth, 1.56f);
			continue;
		}
		if (r.depth > 0 && mat == CM_METAL) {
			float3 n = face_viewer(N, r.d);
			float3 rr = greflect(r.d, n);
			if (sp < kWhStack)
				st[sp++] = WhRay{point + rr * 1e-4f, rr, r.w * (kd * (1.0f / 255.0f)), r.depth - 1};
			continue;
		}
		col = col + r.w * shade_diffuse<LDS, SPILL>(sc, cd, h, point, r.d, lds, lane);
	}
	return col;
}

template <int LDS, int SPILL>
__global__ __launch_bounds__(kCBlock) void k_whitted(DevScene sc, CompatDev cd, int W, int H, float cam_z,
                                                     unsigned char *rgba) {
	__shared__ uint64_t stk[LDS * kCBlock];
	uint32_t b = xcd_block(blockIdx.x, gridDim.x);
	int tiles_x = (W + kTileW - 1) / kTileW;
	int x = (int)(b % tiles_x) * kTileW + (threadIdx.x % kTileW);
	int y = (int)(b / tiles_x) * kTileH + (threadIdx.x / kTileW);
	if (x >= W || y >= H)
		return;
	float3 o, d;
	// renderRaytraced's pixel position x - W/2 (no half-pixel offset)
	compat_primary_ray(x, y, W, H, cam_z, -0.5f, -0.5f, o, d);
	float3 c = whitted_pixel<LDS, SPILL>(sc, cd, o, d, stk, threadIdx.x);
	int rr = min(255, (int)c.x), gg = min(255, (int)c.y), bb = min(255, (int)c.z);
	reinterpret_cast<uchar4 *>(rgba)[y * W + x] =
	    make_uchar4((unsigned char)(rr & 0xFF), (unsigned char)(gg & 0xFF), (unsigned char)(bb & 0xFF), 255);
}

} // namespace

void launch_whitted_render(const DevScene &sc, const CompatDev &cd, int W, int H, float cam_z, unsigned char *rgba,
                           hipStream_t s) {
	unsigned g = (unsigned)(((W + kTileW - 1) / kTileW) * ((H + kTileH - 1) / kTileH));
	if (sc.depth - 1 > kLdsDepth)
		k_whitted<kLdsDepth, kSpillDepth><<<g, kCBlock, 0, s>>>(sc, cd, W, H, cam_z, rgba);
	else
		k_whitted<kLdsDepth, 0><<<g, kCBlock, 0, s>>>(sc, cd, W, H, cam_z, rgba);
	RN_HIP(hipGetLastError());
}

void launch_compat_render(const DevScene &sc, const CompatDev &cd, int W, int H, float cam_z, int spp, uint32_t key0,
                          uint32_t key1, unsigned char *rgba, hipStream_t s) {
	unsigned g = (unsigned)(((W + kTileW - 1) / kTileW) * ((H + kTileH - 1) / kTileH));
	if (sc.depth - 1 > kLdsDepth)
		k_compat<kLdsDepth, kSpillDepth><<<g, kCBlock, 0, s>>>(sc, cd, W, H, cam_z, spp, key0, key1, rgba);
	else
		k_compat<kLdsDepth, 0><<<g, kCBlock, 0, s>>>(sc, cd, W, H, cam_z, spp, key0, key1, rgba);
	RN_HIP(hipGetLastError());
}

} // namespace rn
