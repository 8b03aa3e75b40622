// Photon mapping (renderPhotonMapped, src/Renderer.cpp:1172-1346, and the
// PhotonMap of src/Photon.cpp:5-46) and bidirectional path tracing
// (renderBidirectional, src/BDPT.cpp:15-144) on the GPU.
//
// Random numbers: the reference draws from std::mt19937 streams that run
// across photons (one stream) or across the pixels of a row; here each photon,
// pixel or pixel sample owns a counter-based stream, so every lane is
// independent: draw k of stream (a, b, tag) is word k % 4 of Philox4x32-10 at
// counter (a, b, tag, k / 4), drawn in the reference's order of use.
//   photon i:             (i, 0, 'phot')
//   final gather, pixel:  (pixel, 0, 'gath')
//   bidirectional sample: (pixel, sample, 'bdpt')
// Everything else keeps the reference's FP32 operation order (no fused
// multiply-adds), so that with the same photons the photon-mapped image is
// the reference's.
#include "compat_shade.h"
#include "photon.h"
#include "rn_pathp.h"

#include <algorithm>

namespace rn {

namespace {

constexpr uint32_t kTagPhoton = 0x70686F74u, kTagGather = 0x67617468u, kTagBdpt = 0x62647074u;
constexpr int kMaxDeposits = 6; // tracePhoton's depth runs 5..0, one deposit at most per level
constexpr int kPmBlock = 256;
#ifndef RN_PM_CULL
#define RN_PM_CULL 1
#endif

struct Stream {
	uint32_t a, b, tag, k, k0, k1;
	u32x4 w;
	RN_DEV void start(uint32_t sa, uint32_t sb, uint32_t st) {
		a = sa;
		b = sb;
		tag = st;
		k = 0;
	}
	RN_DEV float next() {
		const uint32_t j = k & 3u;
		if (j == 0u)
			w = philox4x32_10(a, b, tag, k >> 2, k0, k1);
		k++;
		return u01(j == 0u ? w.x : (j == 1u ? w.y : (j == 2u ? w.z : w.w)));
	}
};

// cosineHemisphere (src/Renderer.cpp:778-788, src/BDPT.cpp:21-31): u1 then u2.
RN_DEV float3 cosine_hemisphere(float3 n, Stream &r) {
	float u1 = r.next();
	float u2 = r.next();
	return cos_hemi(n, u1, u2);
}

// The Fresnel term and the refracted ray of the glass branches
// (src/Renderer.cpp:1200-1214,1254-1268; glm::refract, func_geometric.inl:113-123).
RN_DEV float glass_setup(float3 d, float3 N, float3 &nn, float3 &refr) {
#pragma clang fp contract(off)
	const float ior = 1.5f;
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
	float r0 = (etai - etat) / (etai + etat);
	r0 *= r0;
	float fresnel = r0 + (1.0f - r0) * powf(1.0f - cosi, 5.0f);
	float eta = etai / etat;
	float dv = gdot(n, d);
	float k = 1.0f - eta * eta * (1.0f - dv * dv);
	refr = k >= 0.0f ? gsub(gscale(d, eta), gscale(n, eta * dv + sqrtf(k))) : f3(0.0f, 0.0f, 0.0f);
	nn = n;
	return fresnel;
}

RN_DEV float glength(float3 v) { return sqrtf(gdot(v, v)); }

RN_DEV float3 albedo_at(const DevScene &sc, const CompatDev &cd, const Hit &h, float3 point, float3 d) {
	return gdiv(base_colour(cd, sc.tris, h.leaf, h.prim, h.u, h.v, point, gscale(d, -1.0f)), 255.0f);
}

// ---------------------------------------------------------- pass 1 ----

// tracePhoton (src/Renderer.cpp:1177-1226) for photon i: every branch of
// its recursion is a tail call, so it is a loop. Deposits go to the photon's
// six slots; count[i] says how many it filled.
template <int LDS, int SPILL>
__global__ __launch_bounds__(kCBlock) void k_photon_trace(DevScene sc, CompatDev cd, int n_photons, uint32_t k0,
                                                          uint32_t k1, float4 *dep_a, float4 *dep_b, float *dep_c,
                                                          uint32_t *count) {
#pragma clang fp contract(off)
	__shared__ uint64_t stk[LDS * kCBlock];
	const int i = blockIdx.x * kCBlock + threadIdx.x;
	if (i >= n_photons)
		return;
	Stream r;
	r.k0 = k0;
	r.k1 = k1;
	r.start((uint32_t)i, 0u, kTagPhoton);
	// src/Renderer.cpp:1314-1322 with the default light
	float z = 1.0f - 2.0f * r.next();
	float a = 2.0f * 3.14159265f * r.next();
	float rr = sqrtf(fmaxf(0.0f, 1.0f - z * z));
	float3 d = f3(rr * cosf(a), z, rr * sinf(a));
	float e = kLIntensity * 1.8f / (float)n_photons;
	float3 power = f3(1.0f * e, 1.0f * e, 1.0f * e);
	float3 o = RN_LPOS;
	int depth = 5, nd = 0;
	const size_t slot0 = (size_t)i * kMaxDeposits;
	while (depth >= 0) {
		Hit h = traverse<0, 1, LDS, SPILL>(sc.nodes, sc.tris, o, d, INFINITY, -1, stk, kCBlock, threadIdx.x);
		if (h.prim < 0)
			break;
		const float3 point = gadd(o, gscale(d, h.t));
		const float4 nm = cd.cnorm[h.prim];
		const float3 N = xyz(nm);
		const float3 n = face_viewer(N, d);
		const int mat = __float_as_int(nm.w);
		float3 nd3;
		if (mat == CM_MIRROR) {
			nd3 = greflect(d, n);
		} else if (mat == CM_METAL) {
			power = gmul(power, albedo_at(sc, cd, h, point, d));
			nd3 = greflect(d, n);
		} else if (mat == CM_GLASS) {
			float3 nn, refr;
			float fresnel = glass_setup(d, N, nn, refr);
			nd3 = (glength(refr) < 1e-6f || r.next() < fresnel) ? greflect(d, nn) : refr;
		} else {
			// deposit on the diffuse surface, then maybe bounce
			if (nd < kMaxDeposits) {
				dep_a[slot0 + nd] = make_float4(point.x, point.y, point.z, power.x);
				dep_b[slot0 + nd] = make_float4(d.x, d.y, d.z, power.y);
				dep_c[slot0 + nd] = power.z;
				nd++;
			}
			const float3 albedo = albedo_at(sc, cd, h, point, d);
			const float p = fmaxf(fmaxf(albedo.x, albedo.y), albedo.z);
			if (!(depth > 0 && r.next() < p))
				break;
			nd3 = cosine_hemisphere(n, r);
			power = gdiv(gmul(power, albedo), p);
		}
		o = gadd(point, gscale(nd3, 1e-4f));
		d = nd3;
		depth--;
	}
	count[i] = (uint32_t)nd;
}

// Photons in the reference's order (by photon, then by deposit).
__global__ void k_photon_compact(int n_photons, const uint32_t *count, const uint32_t *offs, const float4 *dep_a,
                                 const float4 *dep_b, const float *dep_c, float4 *a, float4 *b, float *c) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n_photons)
		return;
	const uint32_t m = count[i], o = offs[i];
	const size_t slot0 = (size_t)i * kMaxDeposits;
	for (uint32_t j = 0; j < m; j++) {
		a[o + j] = dep_a[slot0 + j];
		b[o + j] = dep_b[slot0 + j];
		c[o + j] = dep_c[slot0 + j];
	}
}

__global__ void k_photon_total(int n, const uint32_t *count, const uint32_t *offs, uint32_t *total) {
	total[0] = n > 0 ? offs[n - 1] + count[n - 1] : 0u;
}

// ---------------------------------------------------------- the map ---

// PhotonMap::cellKey (src/Photon.cpp:5-8) of the cell holding p (cell side
// `cell`, src/Photon.cpp:14-16 and :22-25).
RN_DEV unsigned long long cell_key(int x, int y, int z) {
	auto f = [](int v) -> unsigned long long { return (unsigned long long)((v + (1 << 20)) & 0x1FFFFF); };
	return (f(x) << 42) | (f(y) << 21) | f(z);
}

RN_DEV int3 cell_of(float3 p, float cell) {
	float3 c = gdiv(p, cell);
	return make_int3((int)floorf(c.x), (int)floorf(c.y), (int)floorf(c.z));
}

RN_DEV unsigned long long key_of(float4 a, float cell) {
	int3 c = cell_of(xyz(a), cell);
	return cell_key(c.x, c.y, c.z);
}

__global__ void k_photon_keys_lo(int n, const float4 *a, float cell, uint32_t *lo, uint32_t *perm) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	lo[i] = (uint32_t)key_of(a[i], cell);
	perm[i] = (uint32_t)i;
}

__global__ void k_photon_keys_hi(int n, const float4 *a, float cell, const uint32_t *perm, uint32_t *hi) {
	const int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= n)
		return;
	hi[s] = (uint32_t)(key_of(a[perm[s]], cell) >> 32);
}

// Sorted photons, their keys, and the first slot of every run.
__global__ void k_photon_sorted(int n, const uint32_t *perm, const float4 *a, const float4 *b, const float *c,
                                float cell, float4 *sa, float4 *sb, float *sc, unsigned long long *skey,
                                uint32_t *flag) {
	const int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= n)
		return;
	const uint32_t p = perm[s];
	const float4 pa = a[p];
	sa[s] = pa;
	sb[s] = b[p];
	sc[s] = c[p];
	const unsigned long long k = key_of(pa, cell);
	skey[s] = k;
	flag[s] = (s == 0 || key_of(a[perm[s - 1]], cell) != k) ? 1u : 0u;
}

__global__ void k_photon_cells(int n, const unsigned long long *skey, const uint32_t *flag, const uint32_t *offs,
                               unsigned long long *cell_key_out, uint32_t *cell_start, uint32_t *n_cells) {
	const int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= n)
		return;
	if (flag[s]) {
		cell_key_out[offs[s]] = skey[s];
		cell_start[offs[s]] = (uint32_t)s;
	}
	if (s == n - 1) {
		const uint32_t total = offs[s] + flag[s];
		cell_start[total] = (uint32_t)n;
		n_cells[0] = total;
	}
}

// PhotonMap::gather (src/Photon.cpp:21-46): the 27 cells around the point in
// (dx, dy, dz) order, each run in photon-index order.
RN_DEV float3 pm_gather(const PhotonMapDev &m, float3 point, float3 normal, float radius) {
#pragma clang fp contract(off)
	const int3 c = cell_of(point, m.cell);
	float sx = 0.0f, sy = 0.0f, sz = 0.0f;
	const float r2 = radius * radius;
	// A neighbour cell whose box (widened by a margin far above the rounding
	// of a photon's cell index) lies beyond the radius holds no photon that
	// passes the distance test, so skipping it leaves the sum unchanged.
	const bool cull = RN_PM_CULL && radius <= 4.0f * m.cell;
	const float margin = 1e-3f * m.cell + 1e-6f * (fabsf(point.x) + fabsf(point.y) + fabsf(point.z));
	auto gap = [&](int k, float p) {
		const float lo = (float)k * m.cell - margin, hi = (float)(k + 1) * m.cell + margin;
		return p < lo ? lo - p : (p > hi ? p - hi : 0.0f);
	};
	for (int dx = -1; dx <= 1; dx++)
		for (int dy = -1; dy <= 1; dy++)
			for (int dz = -1; dz <= 1; dz++) {
				if (cull) {
					const float gx = gap(c.x + dx, point.x), gy = gap(c.y + dy, point.y), gz = gap(c.z + dz, point.z);
					if (gx * gx + gy * gy + gz * gz > r2 * 1.0001f)
						continue;
				}
				const unsigned long long key = cell_key(c.x + dx, c.y + dy, c.z + dz);
				int lo = 0, hi = m.n_cells;
				while (lo < hi) {
					const int mid = (lo + hi) >> 1;
					if (m.key[mid] < key)
						lo = mid + 1;
					else
						hi = mid;
				}
				if (lo >= m.n_cells || m.key[lo] != key)
					continue;
				const uint32_t e = m.start[lo + 1];
				for (uint32_t s = m.start[lo]; s < e; s++) {
					const float4 pa = m.a[s];
					const float3 d = gsub(xyz(pa), point);
					if (gdot(d, d) > r2)
						continue;
					const float4 pb = m.b[s];
					if (gdot(normal, gscale(xyz(pb), -1.0f)) <= 0.0f)
						continue;
					sx = sx + pa.w;
					sy = sy + pb.w;
					sz = sz + m.c[s];
				}
			}
	return gdiv(f3(sx, sy, sz), 3.14159265f * r2);
}

__global__ void k_photon_gather(PhotonMapDev m, int count, const float *points, const float *normals, float radius,
                                float *out) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= count)
		return;
	const float *p = points + 3 * (size_t)i, *q = normals + 3 * (size_t)i;
	float3 g = pm_gather(m, f3(p[0], p[1], p[2]), f3(q[0], q[1], q[2]), radius);
	out[3 * (size_t)i] = g.x;
	out[3 * (size_t)i + 1] = g.y;
	out[3 * (size_t)i + 2] = g.z;
}

// ---------------------------------------------------------- pass 2 ----

// One open call of photonShade waiting for its children.
struct PsFrame {
	float3 point, refr, acc; // glass: refracted ray, reflected colour; metal: albedo in acc
	float fresnel;
	int kind; // 0 mirror, 1 metal, 2 glass
	int stage;
	int depth;
};
constexpr int kPsFrames = 4; // the recursion depth of photonShade (depth 4 .. 0)

// The diffuse end of photonShade (src/Renderer.cpp:1278-1294).
template <int LDS, int SPILL>
RN_DEV float3 ps_diffuse(const DevScene &sc, const CompatDev &cd, const PhotonMapDev &m, const Hit &h, float3 point,
                         float3 d, float3 n, float radius, int gather_rays, Stream &r, uint64_t *lds, int lane) {
#pragma clang fp contract(off)
	const float3 albedo = albedo_at(sc, cd, h, point, d);
	const float3 direct = direct_light<LDS, SPILL>(sc, point, n, albedo, h.prim, lds, lane);
	float3 indirect;
	if (gather_rays > 0) {
		float3 gathered = f3(0.0f, 0.0f, 0.0f);
		for (int i = 0; i < gather_rays; i++) {
			const float3 gdir = cosine_hemisphere(n, r);
			const float3 go = gadd(point, gscale(gdir, 1e-4f));
			Hit gh = traverse<0, 1, LDS, SPILL>(sc.nodes, sc.tris, go, gdir, INFINITY, -1, lds, kCBlock, lane);
			if (gh.prim >= 0) {
				const float3 gp = gadd(go, gscale(gdir, gh.t));
				gathered = gadd(gathered, pm_gather(m, gp, face_viewer(xyz(cd.cnorm[gh.prim]), gdir), radius));
			} else {
				gathered = gadd(gathered, gdiv(wh_environment(gdir), 255.0f));
			}
		}
		indirect = gdiv(gmul(albedo, gathered), (float)gather_rays);
	} else {
		indirect = gmul(albedo, pm_gather(m, point, n, radius));
	}
	return gadd(direct, indirect);
}

// photonShade (src/Renderer.cpp:1233-1294) from the camera ray (o, d) at
// depth 4: the recursion runs on an explicit stack of open calls and
// combines the children's colours in the reference's order (the reflected
// ray of glass before the refracted one).
template <int LDS, int SPILL>
RN_DEV float3 photon_shade(const DevScene &sc, const CompatDev &cd, const PhotonMapDev &m, float3 o, float3 d,
                           float radius, int gather_rays, Stream &r, uint64_t *lds, int lane) {
#pragma clang fp contract(off)
	PsFrame st[kPsFrames];
	int sp = 0, depth = 4;
	float3 res;
	for (;;) {
		// evaluate photonShade(o, d, depth) down to a leaf
		Hit h = traverse<0, 1, LDS, SPILL>(sc.nodes, sc.tris, o, d, INFINITY, -1, lds, kCBlock, lane);
		if (h.prim < 0) {
			res = gdiv(wh_environment(d), 255.0f);
		} else {
			const float3 point = gadd(o, gscale(d, h.t));
			const float4 nm = cd.cnorm[h.prim];
			const float3 N = xyz(nm);
			const float3 n = face_viewer(N, d);
			const int mat = __float_as_int(nm.w);
			if (depth > 0 && sp < kPsFrames && (mat == CM_MIRROR || mat == CM_METAL || mat == CM_GLASS)) {
				PsFrame &f = st[sp++];
				f.point = point;
				f.depth = depth;
				f.stage = 0;
				float3 nd;
				if (mat == CM_MIRROR) {
					f.kind = 0;
					nd = greflect(d, n);
				} else if (mat == CM_METAL) {
					f.kind = 1;
					f.acc = albedo_at(sc, cd, h, point, d);
					nd = greflect(d, n);
				} else {
					f.kind = 2;
					float3 nn;
					f.fresnel = glass_setup(d, N, nn, f.refr);
					nd = greflect(d, nn);
				}
				o = gadd(point, gscale(nd, 1e-4f));
				d = nd;
				depth--;
				continue;
			}
			res = ps_diffuse<LDS, SPILL>(sc, cd, m, h, point, d, n, radius, gather_rays, r, lds, lane);
		}
		// return res to the open calls
		bool descend = false;
		while (sp > 0) {
			PsFrame &f = st[sp - 1];
			if (f.kind == 1) {
				res = gmul(f.acc, res);
			} else if (f.kind == 2) {
				if (f.stage == 0 && glength(f.refr) >= 1e-6f) {
					f.acc = res;
					f.stage = 1;
					o = gadd(f.point, gscale(f.refr, 1e-4f));
					d = f.refr;
					depth = f.depth - 1;
					descend = true;
					break;
				}
				if (f.stage == 1)
					res = gadd(gscale(f.acc, f.fresnel), gscale(res, 1.0f - f.fresnel));
			}
			sp--;
		}
		if (!descend)
			return res;
	}
}

RN_DEV void store_rgba(unsigned char *rgba, int idx, float3 c) {
	int rr = min(255, (int)c.x), gg = min(255, (int)c.y), bb = min(255, (int)c.z);
	reinterpret_cast<uchar4 *>(rgba)[idx] =
	    make_uchar4((unsigned char)(rr & 0xFF), (unsigned char)(gg & 0xFF), (unsigned char)(bb & 0xFF), 255);
}

template <int LDS, int SPILL>
__global__ __launch_bounds__(kCBlock) void k_photon_render(DevScene sc, CompatDev cd, PhotonMapDev m, int W, int H,
                                                           float cam_z, int gather_rays, uint32_t k0, uint32_t k1,
                                                           unsigned char *rgba) {
	__shared__ uint64_t stk[LDS * kCBlock];
	uint32_t b = xcd_block(blockIdx.x, gridDim.x);
	int tiles_x = (W + kTileW - 1) / kTileW;
	int x = (int)(b % tiles_x) * kTileW + (threadIdx.x % kTileW);
	int y = (int)(b / tiles_x) * kTileH + (threadIdx.x / kTileW);
	if (x >= W || y >= H)
		return;
	Stream r;
	r.k0 = k0;
	r.k1 = k1;
	r.start((uint32_t)(y * W + x), 0u, kTagGather);
	float3 o, d;
	compat_primary_ray(x, y, W, H, cam_z, -0.5f, -0.5f, o, d); // x - W/2, src/Renderer.cpp:1339
	float3 c = photon_shade<LDS, SPILL>(sc, cd, m, o, d, m.cell, gather_rays, r, stk, threadIdx.x);
	store_rgba(rgba, y * W + x, gscale(c, 255.0f));
}

// ---------------------------------------------------- bidirectional ---

struct BdVertex {
	float3 position, normal, albedo, throughput;
};

// faceViewer of src/BDPT.cpp:34-36 (a zero dot product flips the normal,
// unlike the renderer's).
RN_DEV float3 face_viewer_bd(float3 n, float3 d) { return gdot(n, d) < 0.0f ? n : gscale(n, -1.0f); }

// tracePath (src/BDPT.cpp:42-62): appends up to max_depth - (start) vertices.
template <int LDS, int SPILL, int MAXV>
RN_DEV int trace_subpath(const DevScene &sc, const CompatDev &cd, float3 o, float3 d, float3 beta, BdVertex *out,
                         int nv, Stream &r, uint64_t *lds, int lane) {
#pragma clang fp contract(off)
	const float eps = 1e-3f;
	for (int depth = 0; depth < 4; depth++) {
		Hit h = traverse<0, 1, LDS, SPILL>(sc.nodes, sc.tris, o, d, INFINITY, -1, lds, kCBlock, lane);
		if (h.prim < 0)
			break;
		BdVertex v;
		v.position = gadd(o, gscale(d, h.t));
		v.normal = face_viewer_bd(xyz(cd.cnorm[h.prim]), d);
		v.albedo = gdiv(xyz(cd.ccol[h.prim]), 255.0f);
		v.throughput = beta;
		if (nv < MAXV)
			out[nv++] = v;
		const float3 nd = cosine_hemisphere(v.normal, r);
		beta = gmul(beta, v.albedo);
		o = gadd(v.position, gscale(nd, eps));
		d = nd;
	}
	return nv;
}

template <int LDS, int SPILL>
__global__ __launch_bounds__(kCBlock) void k_bdpt(DevScene sc, CompatDev cd, int W, int H, float cam_z, int samples,
                                                  float exposure, uint32_t k0, uint32_t k1, unsigned char *rgba) {
#pragma clang fp contract(off)
	__shared__ uint64_t stk[LDS * kCBlock];
	uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
	int tiles_x = (W + kTileW - 1) / kTileW;
	int x = (int)(blk % tiles_x) * kTileW + (threadIdx.x % kTileW);
	int y = (int)(blk / tiles_x) * kTileH + (threadIdx.x / kTileW);
	if (x >= W || y >= H)
		return;
	const float kPi = 3.14159265f, eps = 1e-3f;
	const float3 light_centre = f3(0.0f, 0.33f, 0.0f), light_normal = f3(0.0f, -1.0f, 0.0f);
	const float light_half = 0.18f;
	const float light_area = (2.0f * light_half) * (2.0f * light_half);
	const float3 le_area = gscale(f3(30.0f, 30.0f, 30.0f), light_area);
	const uint32_t pix = (uint32_t)(y * W + x);
	Stream r;
	r.k0 = k0;
	r.k1 = k1;
	float3 accum = f3(0.0f, 0.0f, 0.0f);
	for (int si = 0; si < samples; si++) {
		r.start(pix, (uint32_t)si, kTagBdpt);
		// src/BDPT.cpp:90-91: the y jitter is drawn first (right-to-left
		// argument evaluation of the reference's build)
		const float jy = r.next();
		const float jx = r.next();
		float3 o, d;
		compat_camera_ray((float)x + jx - (float)W / 2.0f, -((float)y + jy - (float)H / 2.0f), cam_z, o, d);
		BdVertex cam[4], lig[5];
		const int nc = trace_subpath<LDS, SPILL, 4>(sc, cd, o, d, f3(1.0f, 1.0f, 1.0f), cam, 0, r, stk, threadIdx.x);
		const float lz = r.next();
		const float lx = r.next();
		const float3 lp = gadd(light_centre, f3((lx * 2.0f - 1.0f) * light_half, 0.0f, (lz * 2.0f - 1.0f) * light_half));
		lig[0] = BdVertex{lp, light_normal, f3(0.0f, 0.0f, 0.0f), le_area};
		const float3 ldir = cosine_hemisphere(light_normal, r);
		const int nl = trace_subpath<LDS, SPILL, 5>(sc, cd, gadd(lp, gscale(ldir, eps)), ldir, le_area, lig, 1, r, stk,
		                                            threadIdx.x);
		for (int i = 0; i < nc; i++) {
			const BdVertex cv = cam[i];
			for (int j = 0; j < nl; j++) {
				const BdVertex lv = lig[j];
				float3 dd = gsub(lv.position, cv.position);
				const float dist2 = gdot(dd, dd);
				const float dist = sqrtf(dist2);
				dd = gdiv(dd, dist);
				const float cos_c = gdot(cv.normal, dd);
				const float cos_l = gdot(lv.normal, gscale(dd, -1.0f));
				if (cos_c <= 0.0f || cos_l <= 0.0f)
					continue;
				Hit sh = traverse<1, 1, LDS, SPILL>(sc.nodes, sc.tris, gadd(cv.position, gscale(cv.normal, eps)), dd,
				                                    dist - 2.0f * eps, -1, stk, kCBlock, threadIdx.x);
				if (sh.prim != 0)
					continue;
				const float G = cos_c * cos_l / dist2;
				const float3 fc = gdiv(cv.albedo, kPi);
				float3 contrib = gmul(gscale(gmul(cv.throughput, fc), G), lv.throughput);
				if (j > 0)
					contrib = gmul(contrib, gdiv(lv.albedo, kPi));
				const float w = 1.0f / (float)(i + j + 2);
				accum = gadd(accum, gscale(contrib, w));
			}
		}
	}
	float3 col = gscale(accum, exposure / (float)samples);
	col = gscale(f3(col.x / (col.x + 1.0f), col.y / (col.y + 1.0f), col.z / (col.z + 1.0f)), 255.0f);
	store_rgba(rgba, (int)pix, col);
}

unsigned grid_of(size_t n, int block) { return (unsigned)std::max<size_t>(1, (n + block - 1) / block); }

} // namespace

void photon_map_build(PhotonMap &m, float cell, hipStream_t s) {
	m.cell = cell;
	const int n = m.n;
	const size_t nn = (size_t)std::max(n, 1);
	uint32_t *lo = static_cast<uint32_t *>(m.keys_lo.reserve(nn * 4));
	uint32_t *hi = static_cast<uint32_t *>(m.keys_hi.reserve(nn * 4));
	uint32_t *perm = static_cast<uint32_t *>(m.perm.reserve(nn * 4));
	uint32_t *tk = static_cast<uint32_t *>(m.tmp_k.reserve(nn * 4));
	uint32_t *tv = static_cast<uint32_t *>(m.tmp_v.reserve(nn * 4));
	float4 *sa = static_cast<float4 *>(m.sa.reserve(nn * 16));
	float4 *sb = static_cast<float4 *>(m.sb.reserve(nn * 16));
	float *sc = static_cast<float *>(m.sc.reserve(nn * 4));
	unsigned long long *skey = static_cast<unsigned long long *>(m.dep_a.reserve(nn * 8));
	uint32_t *flag = static_cast<uint32_t *>(m.flags.reserve(nn * 4));
	uint32_t *offs = static_cast<uint32_t *>(m.offs.reserve(nn * 4));
	unsigned long long *ckey = static_cast<unsigned long long *>(m.cell_key.reserve(nn * 8));
	uint32_t *cstart = static_cast<uint32_t *>(m.cell_start.reserve((nn + 1) * 4));
	uint32_t *total = static_cast<uint32_t *>(m.total.reserve(16));
	uint32_t *tmp = static_cast<uint32_t *>(m.scan_tmp.reserve(device_scan_tmp_words(nn) * 4 + 16));
	if (n == 0) {
		m.n_cells = 0;
		RN_HIP(hipMemsetAsync(cstart, 0, 4, s));
		return;
	}
	const float4 *a = m.a.as<float4>();
	// stable sort by (high word, low word) of the 63-bit key: low word first
	k_photon_keys_lo<<<grid_of(n, kPmBlock), kPmBlock, 0, s>>>(n, a, cell, lo, perm);
	radix_sort_pairs(lo, perm, tk, tv, n, 32, m.hist, m.scan_tmp, s);
	k_photon_keys_hi<<<grid_of(n, kPmBlock), kPmBlock, 0, s>>>(n, a, cell, perm, hi);
	radix_sort_pairs(hi, perm, tk, tv, n, 31, m.hist, m.scan_tmp, s);
	tmp = static_cast<uint32_t *>(m.scan_tmp.reserve(device_scan_tmp_words(nn) * 4 + 16));
	k_photon_sorted<<<grid_of(n, kPmBlock), kPmBlock, 0, s>>>(n, perm, a, m.b.as<float4>(), m.c.as<float>(), cell, sa,
	                                                           sb, sc, skey, flag);
	device_exclusive_scan(flag, offs, (size_t)n, tmp, s);
	k_photon_cells<<<grid_of(n, kPmBlock), kPmBlock, 0, s>>>(n, skey, flag, offs, ckey, cstart, total);
	RN_HIP(hipGetLastError());
	uint32_t h = 0;
	RN_HIP(hipMemcpyAsync(&h, total, 4, hipMemcpyDeviceToHost, s));
	RN_HIP(hipStreamSynchronize(s));
	m.n_cells = (int)h;
}

void photon_trace(PhotonMap &m, const DevScene &sc, const CompatDev &cd, int num_photons, uint32_t key0,
                  uint32_t key1, hipStream_t s) {
	if (num_photons < 0 || num_photons > (1 << 26))
		throw Error("photon count must be in [0, 2^26]");
	const size_t np = (size_t)std::max(num_photons, 1);
	const size_t slots = np * kMaxDeposits;
	float4 *da = static_cast<float4 *>(m.dep_a.reserve(slots * 16));
	float4 *db = static_cast<float4 *>(m.dep_b.reserve(slots * 16));
	float *dc = static_cast<float *>(m.dep_c.reserve(slots * 4));
	uint32_t *cnt = static_cast<uint32_t *>(m.dep_n.reserve(np * 4));
	uint32_t *offs = static_cast<uint32_t *>(m.offs.reserve(np * 4));
	uint32_t *total = static_cast<uint32_t *>(m.total.reserve(16));
	uint32_t *tmp = static_cast<uint32_t *>(m.scan_tmp.reserve(device_scan_tmp_words(np) * 4 + 16));
	if (num_photons > 0) {
		unsigned g = grid_of((size_t)num_photons, kCBlock);
		if (sc.depth - 1 > kLdsDepth)
			k_photon_trace<kLdsDepth, kSpillDepth><<<g, kCBlock, 0, s>>>(sc, cd, num_photons, key0, key1, da, db, dc,
			                                                             cnt);
		else
			k_photon_trace<kLdsDepth, 0><<<g, kCBlock, 0, s>>>(sc, cd, num_photons, key0, key1, da, db, dc, cnt);
		RN_HIP(hipGetLastError());
		device_exclusive_scan(cnt, offs, (size_t)num_photons, tmp, s);
		k_photon_total<<<1, 1, 0, s>>>(num_photons, cnt, offs, total);
	}
	uint32_t h = 0;
	if (num_photons > 0) {
		RN_HIP(hipMemcpyAsync(&h, total, 4, hipMemcpyDeviceToHost, s));
		RN_HIP(hipStreamSynchronize(s));
	}
	m.n = (int)h;
	const size_t nn = (size_t)std::max(m.n, 1);
	float4 *a = static_cast<float4 *>(m.a.reserve(nn * 16));
	float4 *b = static_cast<float4 *>(m.b.reserve(nn * 16));
	float *c = static_cast<float *>(m.c.reserve(nn * 4));
	if (m.n > 0) {
		k_photon_compact<<<grid_of((size_t)num_photons, kPmBlock), kPmBlock, 0, s>>>(num_photons, cnt, offs, da, db,
		                                                                             dc, a, b, c);
		RN_HIP(hipGetLastError());
	}
	photon_map_build(m, 0.25f, s);
}

void launch_photon_gather(const PhotonMap &m, int count, const float *points, const float *normals, float radius,
                          float *out, hipStream_t s) {
	if (count <= 0)
		return;
	k_photon_gather<<<grid_of((size_t)count, kPmBlock), kPmBlock, 0, s>>>(photon_map_dev(m), count, points, normals,
	                                                                      radius, out);
	RN_HIP(hipGetLastError());
}

void launch_photon_render(const DevScene &sc, const CompatDev &cd, const PhotonMap &m, int W, int H, float cam_z,
                          int gather_rays, uint32_t key0, uint32_t key1, unsigned char *rgba, hipStream_t s) {
	unsigned g = (unsigned)(((W + kTileW - 1) / kTileW) * ((H + kTileH - 1) / kTileH));
	PhotonMapDev md = photon_map_dev(m);
	if (sc.depth - 1 > kLdsDepth)
		k_photon_render<kLdsDepth, kSpillDepth><<<g, kCBlock, 0, s>>>(sc, cd, md, W, H, cam_z, gather_rays, key0, key1,
		                                                              rgba);
	else
		k_photon_render<kLdsDepth, 0><<<g, kCBlock, 0, s>>>(sc, cd, md, W, H, cam_z, gather_rays, key0, key1, rgba);
	RN_HIP(hipGetLastError());
}

void launch_bdpt_render(const DevScene &sc, const CompatDev &cd, int W, int H, float cam_z, int samples,
                        float exposure, uint32_t key0, uint32_t key1, unsigned char *rgba, hipStream_t s) {
	unsigned g = (unsigned)(((W + kTileW - 1) / kTileW) * ((H + kTileH - 1) / kTileH));
	if (sc.depth - 1 > kLdsDepth)
		k_bdpt<kLdsDepth, kSpillDepth><<<g, kCBlock, 0, s>>>(sc, cd, W, H, cam_z, samples, exposure, key0, key1, rgba);
	else
		k_bdpt<kLdsDepth, 0><<<g, kCBlock, 0, s>>>(sc, cd, W, H, cam_z, samples, exposure, key0, key1, rgba);
	RN_HIP(hipGetLastError());
}

} // namespace rn
