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
#include "compat_shade.h"

namespace rn {

namespace {

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

// ------------------------------------------- the reference's own streams --
// rn_render(RN_PATHTRACED) as the reference computes it: one std::mt19937
// per pixel seeded (y * W + x) * 9781 + 1, read through libstdc++'s
// uniform_real_distribution<float> (src/Renderer.cpp:1138-1160), and
// pathTrace's colours combined as the recursion returns them
// (direct + albedo * bounce, innermost first). The generator's 624-word state
// lives in the lane's private (scratch) memory.

// MT19937 (Matsumoto and Nishimura 1998) with std::mt19937's seeding.
struct Mt19937 {
	uint32_t x[624];
	int i;
	RN_DEV void seed(uint32_t s) {
		x[0] = s;
#pragma unroll 1
		for (int k = 1; k < 624; k++)
			x[k] = 1812433253u * (x[k - 1] ^ (x[k - 1] >> 30)) + (uint32_t)k;
		i = 624;
	}
	RN_DEV static uint32_t mix(uint32_t a, uint32_t b, uint32_t c) {
		uint32_t y = (a & 0x80000000u) | (b & 0x7FFFFFFFu);
		return c ^ (y >> 1) ^ ((y & 1u) ? 0x9908B0DFu : 0u);
	}
	RN_DEV void twist() {
#pragma unroll 1
		for (int k = 0; k < 227; k++)
			x[k] = mix(x[k], x[k + 1], x[k + 397]);
#pragma unroll 1
		for (int k = 227; k < 623; k++)
			x[k] = mix(x[k], x[k + 1], x[k - 227]);
		x[623] = mix(x[623], x[0], x[396]);
		i = 0;
	}
	RN_DEV uint32_t next() {
		if (i >= 624)
			twist();
		uint32_t y = x[i++];
		y ^= y >> 11;
		y ^= (y << 7) & 0x9D2C5680u;
		y ^= (y << 15) & 0xEFC60000u;
		y ^= y >> 18;
		return y;
	}
	// generate_canonical<float, 24>: float(x) / 2^32, below 1
	RN_DEV float unif() {
		float r = (float)next() / 4294967296.0f;
		return r >= 1.0f ? 0x1.fffffep-1f : r;
	}
};

constexpr int kMtLevels = 5; // pathTrace depth 4 .. 0

// pathTrace (src/Renderer.cpp:820-907) for one camera ray: walks down to the
// end of the path recording each level's (direct, albedo) and how it
// combines with the level below, then folds the colours back up.
template <int LDS, int SPILL>
RN_DEV float3 compat_path_mt(const DevScene &sc, const CompatDev &cd, float3 o, float3 d, Mt19937 &g, uint64_t *lds,
                             int lane) {
#pragma clang fp contract(off)
	float3 dl[kMtLevels], al[kMtLevels];
	int kind[kMtLevels]; // 0: pass the bounce through, 1: albedo * bounce, 2: direct + albedo * bounce
	int nl = 0;
	float3 res;
	int depth = 4;
	for (;;) {
		Hit h = traverse<0, 1, LDS, SPILL>(sc.nodes, sc.tris, o, d, INFINITY, -1, lds, kCBlock, lane);
		if (h.prim < 0) {
			res = gdiv(wh_environment(d), 255.0f);
			break;
		}
		const int ti = h.prim;
		const float3 point = gadd(o, gscale(d, h.t));
		const float4 nm = cd.cnorm[ti];
		const float3 N = xyz(nm);
		const int mat = __float_as_int(nm.w);
		if (mat == CM_EMISSIVE) {
			res = gscale(gdiv(xyz(cd.ccol[ti]), 255.0f), 4.0f);
			break;
		}
		float3 nd;
		if (mat == CM_MIRROR && depth > 0) {
			nd = greflect(d, face_viewer(N, d));
			kind[nl] = 0;
		} else if ((mat == CM_GLASS || mat == CM_DISPERSIVE) && depth > 0) {
			float3 n = N;
			float cosi = gdot(d, n);
			float etai = 1.0f, etat = 1.5f;
			if (cosi < 0.0f) {
				cosi = -cosi;
			} else {
				etai = 1.5f;
				etat = 1.0f;
				n = gscale(n, -1.0f);
			}
			float r0 = (etai - etat) / (etai + etat);
			r0 *= r0;
			float fres = r0 + (1.0f - r0) * powf(1.0f - cosi, 5.0f);
			float eta = etai / etat;
			float dv = gdot(n, d);
			float k = 1.0f - eta * eta * (1.0f - dv * dv);
			float3 refr = k >= 0.0f ? gsub(gscale(d, eta), gscale(n, eta * dv + sqrtf(k))) : f3(0.0f, 0.0f, 0.0f);
			nd = (sqrtf(gdot(refr, refr)) < 1e-6f || g.unif() < fres) ? greflect(d, n) : refr;
			kind[nl] = 0;
		} else if (mat == CM_METAL && depth > 0) {
			const float3 n = face_viewer(N, d);
			const float3 refl = greflect(d, n);
			const float u1 = g.unif();
			const float u2 = g.unif();
			const float3 diff = cos_hemi(n, u1, u2);
			const float r2 = 0.2f * 0.2f; // ModelTriangle::roughness default (framework/ModelTriangle.h:31)
			nd = gnormalize(gmix(refl, diff, r2));
			if (gdot(nd, n) <= 0.0f)
				nd = refl;
			al[nl] = gdiv(base_colour(cd, sc.tris, h.leaf, ti, h.u, h.v, point, gscale(d, -1.0f)), 255.0f);
			kind[nl] = 1;
		} else {
			const float3 n = face_viewer(N, d);
			const float3 albedo = gdiv(base_colour(cd, sc.tris, h.leaf, ti, h.u, h.v, point, gscale(d, -1.0f)), 255.0f);
			float3 direct;
			if (mat == CM_SUBSURFACE) { // src/Renderer.cpp:872-897
				const float3 to = gsub(RN_LPOS, point);
				const float dist = sqrtf(gdot(to, to));
				const float3 ldir = gdiv(to, dist);
				const float att = kLIntensity / (0.0f + 0.0f * dist + kLKq * dist * dist);
				const float wd = fmaxf(0.0f, (gdot(n, ldir) + 0.5f) / (1.0f + 0.5f));
				float3 shade = gscale(albedo, visibility<LDS, SPILL>(sc, point, ti, lds, lane) * att * wd);
				Hit ex = traverse<0, 1, LDS, SPILL>(sc.nodes, sc.tris, gsub(point, gscale(n, 1e-3f)), ldir, INFINITY, ti,
				                                    lds, kCBlock, lane);
				const float thick = ex.prim >= 0 ? ex.t : 0.0f;
				shade = gadd(shade, gscale(albedo, att * expf(-3.0f * thick) * 0.6f));
				direct = f3(fminf(shade.x, 1.5f), fminf(shade.y, 1.5f), fminf(shade.z, 1.5f));
			} else { // diffuse and every material that falls through to it (src/Renderer.cpp:899-906)
				direct = direct_light<LDS, SPILL>(sc, point, n, albedo, ti, lds, lane);
			}
			if (depth <= 0) {
				res = direct;
				break;
			}
			const float u1 = g.unif();
			const float u2 = g.unif();
			nd = cos_hemi(n, u1, u2);
			dl[nl] = direct;
			al[nl] = albedo;
			kind[nl] = 2;
		}
		nl++;
		o = gadd(point, gscale(nd, 1e-4f));
		d = nd;
		depth--;
	}
	for (int k = nl - 1; k >= 0; k--) {
		if (kind[k] == 1)
			res = gmul(al[k], res);
		else if (kind[k] == 2)
			res = gadd(dl[k], gmul(al[k], res));
	}
	return res;
}

template <int LDS, int SPILL>
__global__ __launch_bounds__(kCBlock) void k_compat_mt(DevScene sc, CompatDev cd, int W, int H, float cam_z, int spp,
                                                       unsigned char *rgba) {
#pragma clang fp contract(off)
	__shared__ uint64_t stk[LDS * kCBlock];
	uint32_t b = xcd_block(blockIdx.x, gridDim.x);
	int tiles_x = (W + kTileW - 1) / kTileW;
	int x = (int)(b % tiles_x) * kTileW + (threadIdx.x % kTileW);
	int y = (int)(b / tiles_x) * kTileH + (threadIdx.x / kTileW);
	if (x >= W || y >= H)
		return;
	const uint32_t pix = (uint32_t)(y * W + x);
	Mt19937 g;
	g.seed(pix * 9781u + 1u);
	float3 sum = f3(0.0f, 0.0f, 0.0f);
	for (int s = 0; s < spp; s++) {
		const float jx = g.unif() * 1.0f + -0.5f; // uniform_real_distribution<float>(-0.5, 0.5)
		const float jy = g.unif() * 1.0f + -0.5f;
		float3 o, d;
		compat_primary_ray(x, y, W, H, cam_z, jx, jy, o, d);
		(void)g.unif(); // the camera-motion draw (the motion is zero)
		sum = gadd(sum, compat_path_mt<LDS, SPILL>(sc, cd, o, d, g, stk, threadIdx.x));
	}
	const float3 c = gscale(gdiv(sum, (float)spp), 255.0f);
	int rr = min(255, (int)c.x), gg = min(255, (int)c.y), bb = min(255, (int)c.z);
	reinterpret_cast<uchar4 *>(rgba)[pix] =
	    make_uchar4((unsigned char)(rr & 0xFF), (unsigned char)(gg & 0xFF), (unsigned char)(bb & 0xFF), 255);
}

// ------------------------------------------------------------ Whitted --
// rn_render(RN_RAYTRACED): renderRaytraced (src/Renderer.cpp:909-948) with
// the arguments rn_render passes (src/capi.cpp:59-61): Phong shading, the
// default soft point light, maxDepth 4, no fog. traceRay's recursion
// (src/Renderer.cpp:704-772) runs on an explicit stack of open calls (mirror,
// metal, glass, dispersive); each combines its children's colours in the
// reference's order and FP32 operations once they have returned (the
// reflected ray of glass before the refracted one, the three dispersive
// channels in turn), so the pixel is the reference's sum, not a reordering.

// src/Renderer.cpp:574-581
RN_DEV float3 bump_normal(float3 n, float3 p) {
#pragma clang fp contract(off)
	float3 q = gscale(p, 6.0f);
	const float e = 0.15f;
	float3 g = f3(perlin(q.x + e, q.y, q.z) - perlin(q.x - e, q.y, q.z),
	              perlin(q.x, q.y + e, q.z) - perlin(q.x, q.y - e, q.z),
	              perlin(q.x, q.y, q.z + e) - perlin(q.x, q.y, q.z - e));
	return gnormalize(gadd(n, gscale(g, 2.0f)));
}

// src/Renderer.cpp:618-639
RN_DEV float3 tangent_normal(const CompatDev &cd, const float4 *tris, uint32_t leaf, int prim, int tex, float u,
                             float v, float3 gn) {
#pragma clang fp contract(off)
	float3 v0 = xyz(tris[3 * leaf]), v1 = xyz(tris[3 * leaf + 1]), v2 = xyz(tris[3 * leaf + 2]);
	float3 e1 = gsub(v1, v0), e2 = gsub(v2, v0);
	const float *tp = cd.ctc + 6 * (size_t)prim;
	float d1x = tp[2] - tp[0], d1y = tp[3] - tp[1], d2x = tp[4] - tp[0], d2y = tp[5] - tp[1];
	float det = d1x * d2y - d2x * d1y;
	if (fabsf(det) < 1e-8f)
		return gn;
	float f = 1.0f / det;
	float3 T = gnormalize(gscale(gsub(gscale(e1, d2y), gscale(e2, d1y)), f));
	float3 B = gnormalize(gscale(gadd(gscale(e1, -d2x), gscale(e2, d1x)), f));
	int4 ti = cd.tex_info[tex];
	float e = 1.0f / (float)ti.y;
	auto lum = [&](float uu, float vv) {
		float3 c = texel_rgb(cd, tex, uu, vv);
		return (c.x + c.y + c.z) / 765.0f;
	};
	float hu = lum(u + e, v) - lum(u - e, v);
	float hv = lum(u, v + e) - lum(u, v - e);
	float3 tn = gnormalize(f3(-hu * 4.0f, -hv * 4.0f, 1.0f));
	return gnormalize(gadd(gadd(gscale(T, tn.x), gscale(B, tn.y)), gscale(gn, tn.z)));
}

// shadeSurface (src/Renderer.cpp:503-534) for the default light.
template <int LDS, int SPILL>
RN_DEV float3 shade_surface(const DevScene &sc, float3 point, float3 n, float3 view, int ignore, float3 base,
                            uint64_t *lds, int lane) {
#pragma clang fp contract(off)
	const float ambient = 0.2f;
	float3 to = gsub(RN_LPOS, point);
	float dist = sqrtf(gdot(to, to));
	float3 ldir = gdiv(to, dist);
	float att = kLIntensity / (0.0f + 0.0f * dist + kLKq * dist * dist);
	float vis = visibility<LDS, SPILL>(sc, point, ignore, lds, lane);
	float inc = fmaxf(0.0f, gdot(n, ldir));
	float bright = vis * 1.0f * att * inc;
	float3 diff = gadd(f3(0.0f, 0.0f, 0.0f), gscale(f3(1.0f, 1.0f, 1.0f), bright));
	float spec = 0.0f;
	if (bright > 0.0f) {
		float3 rd = gsub(gscale(n, 2.0f * gdot(n, ldir)), ldir);
		spec += vis * 1.0f * powf(fmaxf(0.0f, gdot(rd, view)), 64.0f);
	}
	float3 lit = gadd(diff, f3(ambient, ambient, ambient));
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
	float3 view = gscale(d, -1.0f);
	float3 base = base_colour(cd, sc.tris, h.leaf, prim, h.u, h.v, point, view);
	const float *vn = cd.vnorm + 9 * (size_t)prim;
	float3 n0 = f3(vn[0], vn[1], vn[2]), n1 = f3(vn[3], vn[4], vn[5]), n2 = f3(vn[6], vn[7], vn[8]);
	float3 n = gnormalize(gadd(gadd(gscale(n0, w0), gscale(n1, w1)), gscale(n2, w2)));
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

// One open call of traceRay / glassTrace waiting for its children.
struct WhFrame {
	float3 point, refr, acc; // glass: refracted ray, reflected colour; metal: tint; dispersive: channels so far
	float3 d, N;             // dispersive: the incoming ray and the face normal, for its next glassTrace
	float fresnel;
	int kind; // 0 mirror, 1 metal, 2 glass, 3 dispersive
	int stage;
	int depth;
};
constexpr int kWhFrames = 8; // per level at most a dispersive call and one of its glassTrace calls

// glassTrace (src/Renderer.cpp:710-731) up to its first child: the frame
// holds the refracted ray and the Fresnel weight; returns the reflected ray.
RN_DEV float3 glass_open(WhFrame &f, float3 point, float3 d, float3 N, int depth, float ior) {
#pragma clang fp contract(off)
	float3 n = N;
	float cosi = gdot(d, n);
	float etai = 1.0f, etat = ior;
	if (cosi < 0.0f) {
		cosi = -cosi;
	} else {
		etai = ior;
		etat = 1.0f;
		n = gscale(n, -1.0f);
	}
	const float3 refl = greflect(d, n);
	const float eta = etai / etat;
	const float dv = gdot(n, d);
	const float k = 1.0f - eta * eta * (1.0f - dv * dv);
	f.refr = k >= 0.0f ? gsub(gscale(d, eta), gscale(n, eta * dv + sqrtf(k))) : f3(0.0f, 0.0f, 0.0f);
	float r0 = (etai - etat) / (etai + etat);
	r0 *= r0;
	f.fresnel = r0 + (1.0f - r0) * powf(1.0f - cosi, 5.0f);
	f.kind = 2;
	f.stage = 0;
	f.point = point;
	f.depth = depth;
	return refl;
}

template <int LDS, int SPILL>
RN_DEV float3 whitted_pixel(const DevScene &sc, const CompatDev &cd, float3 o, float3 d, uint64_t *lds, int lane) {
#pragma clang fp contract(off)
	WhFrame st[kWhFrames];
	int sp = 0, depth = 4;
	float3 res;
	for (;;) {
		// traceRay(o, d, depth) down to its first child or a leaf colour
		Hit h = traverse<0, 1, LDS, SPILL>(sc.nodes, sc.tris, o, d, INFINITY, -1, lds, kCBlock, lane);
		bool leaf = true;
		if (h.prim < 0) {
			res = wh_environment(d);
		} else {
			const int ti = h.prim;
			const float3 point = gadd(o, gscale(d, h.t));
			const float4 nm = cd.cnorm[ti];
			const float3 N = xyz(nm);
			const int mat = __float_as_int(nm.w);
			const float3 kd = xyz(cd.ccol[ti]);
			if (mat == CM_EMISSIVE) {
				res = gscale(kd, 1.3f);
			} else if (depth > 0 && sp + 2 <= kWhFrames &&
			           (mat == CM_MIRROR || mat == CM_GLASS || mat == CM_DISPERSIVE || mat == CM_METAL)) {
				float3 nd;
				if (mat == CM_MIRROR || mat == CM_METAL) {
					WhFrame &f = st[sp++];
					f.kind = mat == CM_MIRROR ? 0 : 1;
					f.acc = gdiv(kd, 255.0f);
					nd = greflect(d, face_viewer(N, d));
				} else if (mat == CM_GLASS) {
					nd = glass_open(st[sp++], point, d, N, depth, 1.5f);
				} else {
					WhFrame &f = st[sp++];
					f.kind = 3;
					f.stage = 0;
					f.point = point;
					f.d = d;
					f.N = N;
					f.depth = depth;
					f.acc = f3(0.0f, 0.0f, 0.0f);
					nd = glass_open(st[sp++], point, d, N, depth, 1.50f);
				}
				o = gadd(point, gscale(nd, 1e-4f));
				d = nd;
				depth--;
				leaf = false;
			} else {
				res = shade_diffuse<LDS, SPILL>(sc, cd, h, point, d, lds, lane);
			}
		}
		if (!leaf)
			continue;
		// return res to the open calls
		bool descend = false;
		while (sp > 0 && !descend) {
			WhFrame &f = st[sp - 1];
			if (f.kind == 0) {
				sp--;
			} else if (f.kind == 1) {
				res = gmul(f.acc, res);
				sp--;
			} else if (f.kind == 2) {
				if (f.stage == 0 && sqrtf(gdot(f.refr, f.refr)) >= 1e-6f) {
					f.acc = res;
					f.stage = 1;
					o = gadd(f.point, gscale(f.refr, 1e-4f));
					d = f.refr;
					depth = f.depth - 1;
					descend = true;
				} else {
					if (f.stage == 1)
						res = gadd(gscale(f.acc, f.fresnel), gscale(res, 1.0f - f.fresnel));
					sp--;
				}
			} else { // dispersive: res is glassTrace's colour for channel f.stage
				if (f.stage == 0)
					f.acc.x = res.x;
				else if (f.stage == 1)
					f.acc.y = res.y;
				else
					f.acc.z = res.z;
				if (f.stage < 2) {
					f.stage++;
					const float3 nd = glass_open(st[sp++], f.point, f.d, f.N, f.depth, f.stage == 1 ? 1.53f : 1.56f);
					o = gadd(f.point, gscale(nd, 1e-4f));
					d = nd;
					depth = f.depth - 1;
					descend = true;
				} else {
					res = f.acc;
					sp--;
				}
			}
		}
		if (!descend)
			return res;
	}
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

void launch_compat_render_mt(const DevScene &sc, const CompatDev &cd, int W, int H, float cam_z, int spp,
                             unsigned char *rgba, hipStream_t s) {
	unsigned g = (unsigned)(((W + kTileW - 1) / kTileW) * ((H + kTileH - 1) / kTileH));
	if (sc.depth - 1 > kLdsDepth)
		k_compat_mt<kLdsDepth, kSpillDepth><<<g, kCBlock, 0, s>>>(sc, cd, W, H, cam_z, spp, rgba);
	else
		k_compat_mt<kLdsDepth, 0><<<g, kCBlock, 0, s>>>(sc, cd, W, H, cam_z, spp, rgba);
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
