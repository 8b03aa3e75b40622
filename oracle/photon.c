/* Oracle: restatement of the reference's photon mapper (src/Renderer.cpp:
 * 1172-1327, src/Photon.cpp:5-46) and bidirectional path tracer
 * (src/BDPT.cpp:15-144) in plain C, FP32, glm's operation order.
 *
 * Random numbers come from one of two generators, drawn in the reference's
 * order of use:
 *   OR_RNG_MT      the reference's own: std::mt19937 (seed 777 for the
 *                  photons; 1000 + y per row for the final gather; 9871 + y
 *                  per row for the bidirectional tracer) read through
 *                  libstdc++'s uniform_real_distribution<float>(0, 1);
 *   OR_RNG_PHILOX  counter-based streams, one per photon / pixel / pixel
 *                  sample, so that every one can be computed in parallel:
 *                  draw k of stream (a, b, tag) is word k % 4 of
 *                  Philox4x32-10 at counter (a, b, tag, k / 4).
 * With OR_RNG_MT the images equal the compiled reference's; with
 * OR_RNG_PHILOX they are what the GPU computes.
 *
 * TEST INFRASTRUCTURE ONLY -- see oracle.h. */
#include "compat_impl.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ----------------------------------------------------------- generators -- */

enum { TAG_PHOTON = 0x70686F74u, TAG_GATHER = 0x67617468u, TAG_BDPT = 0x62647074u };

typedef struct {
	int kind;
	mt19937 *mt;
	uint32_t key[2], a, b, tag, k;
} rng_t;

static float philox_unif(rng_t *r) {
	uint32_t ctr[4] = {r->a, r->b, r->tag, r->k >> 2}, out[4];
	or_philox4x32_10(ctr, r->key, out);
	float u = unif(out[r->k & 3u]);
	r->k++;
	return u;
}

static float U(rng_t *r) {
	return r->kind == OR_RNG_MT ? mt_unif(r->mt) : philox_unif(r);
}

static void stream(rng_t *r, uint32_t a, uint32_t b, uint32_t tag) {
	r->a = a;
	r->b = b;
	r->tag = tag;
	r->k = 0;
}

/* src/Renderer.cpp:778-788 drawing (u1, u2) in order. */
static v3 cosine_hemisphere(v3 n, rng_t *r) {
	float u1 = U(r);
	float u2 = U(r);
	return cos_hemi(n, u1, u2);
}

static int closest(const cctx *c, v3 o, v3 d, float *t, float *u, float *v) {
	float oa[3] = {o.x, o.y, o.z}, da[3] = {d.x, d.y, d.z};
	return or_compat_bvh_intersect(c->bvh, oa, da, -1, t, u, v);
}

/* The Fresnel term and the refracted ray of the glass branches
 * (src/Renderer.cpp:1200-1214, 1255-1268). */
static float glass_setup(v3 d, v3 N, v3 *nn, v3 *refr) {
	const float ior = 1.5f;
	v3 n = N;
	float cosi = dot(d, n);
	float etai = 1.0f, etat = ior;
	if (cosi < 0.0f) {
		cosi = -cosi;
	} else {
		float tmp = etai;
		etai = etat;
		etat = tmp;
		n = muls(n, -1.0f);
	}
	float r0 = (etai - etat) / (etai + etat);
	r0 *= r0;
	float fresnel = r0 + (1.0f - r0) * powf(1.0f - cosi, 5.0f);
	*nn = n;
	*refr = refract(d, n, etai / etat);
	return fresnel;
}

/* ------------------------------------------------------------- photons -- */

typedef struct {
	float *pos, *power, *dir;
	int n, cap;
} plist;

static void deposit(plist *p, v3 x, v3 power, v3 d) {
	if (p->n < p->cap) {
		float *a = p->pos + 3 * (size_t)p->n, *b = p->power + 3 * (size_t)p->n, *c = p->dir + 3 * (size_t)p->n;
		a[0] = x.x, a[1] = x.y, a[2] = x.z;
		b[0] = power.x, b[1] = power.y, b[2] = power.z;
		c[0] = d.x, c[1] = d.y, c[2] = d.z;
	}
	p->n++;
}

/* tracePhoton, src/Renderer.cpp:1177-1226 (its recursion is a tail call on
 * every branch, so it is a loop here). */
static void trace_photon(const cctx *c, v3 o, v3 d, v3 power, int depth, plist *out, rng_t *r) {
	const or_compat_scene *s = c->s;
	for (;;) {
		if (depth < 0)
			return;
		float t, hu, hv;
		int ti = closest(c, o, d, &t, &hu, &hv);
		if (ti < 0)
			return;
		v3 point = add(o, muls(d, t));
		v3 N = ld(s->normals + 3 * ti);
		v3 n = face_viewer(N, d);
		int mat = s->material[ti];
		if (mat == M_MIRROR) {
			v3 rr = reflect(d, n);
			o = add(point, muls(rr, 1e-4f));
			d = rr;
			depth--;
			continue;
		}
		if (mat == M_METAL) {
			v3 albedo = divs(base_colour(s, ti, hu, hv, point, muls(d, -1.0f)), 255.0f);
			v3 rr = reflect(d, n);
			power = mul(power, albedo);
			o = add(point, muls(rr, 1e-4f));
			d = rr;
			depth--;
			continue;
		}
		if (mat == M_GLASS) {
			v3 nn, refr;
			float fresnel = glass_setup(d, N, &nn, &refr);
			v3 nd = (length(refr) < 1e-6f || U(r) < fresnel) ? reflect(d, nn) : refr;
			o = add(point, muls(nd, 1e-4f));
			d = nd;
			depth--;
			continue;
		}
		deposit(out, point, power, d);
		v3 albedo = divs(base_colour(s, ti, hu, hv, point, muls(d, -1.0f)), 255.0f);
		float p = fmaxf(fmaxf(albedo.x, albedo.y), albedo.z); /* std::max({r, g, b}) */
		if (!(depth > 0 && U(r) < p))
			return;
		v3 nd = cosine_hemisphere(n, r);
		power = divs(mul(power, albedo), p);
		o = add(point, muls(nd, 1e-4f));
		d = nd;
		depth--;
	}
}

/* Pass 1 of renderPhotonMapped (src/Renderer.cpp:1310-1323) with the
 * default light (src/Renderer.cpp:1303-1308, src/Light.h:13-29). */
int or_photon_trace(const or_compat_scene *s, int num_photons, int rng, uint32_t key0, uint32_t key1, float *pos,
                    float *power, float *dir, int cap) {
	cctx c;
	c.s = s;
	c.bvh = or_compat_bvh_build(s->verts, s->n);
	plist out = {pos, power, dir, 0, cap};
	mt19937 *g = (mt19937 *)malloc(sizeof(mt19937));
	mt_seed(g, 777u);
	rng_t r = {rng, g, {key0, key1}, 0, 0, 0, 0};
	const float emit_scale = 1.8f;
	v3 lp = ld(kLightPos);
	for (int i = 0; i < num_photons; i++) {
		if (rng == OR_RNG_PHILOX)
			stream(&r, (uint32_t)i, 0u, TAG_PHOTON);
		float z = 1.0f - 2.0f * U(&r);
		float a = 2.0f * 3.14159265f * U(&r);
		float rr = sqrtf(fmaxf(0.0f, 1.0f - z * z));
		v3 d = mk(rr * cosf(a), z, rr * sinf(a));
		float e = kLightIntensity * emit_scale / (float)num_photons;
		v3 pw = mk(1.0f * e, 1.0f * e, 1.0f * e);
		trace_photon(&c, lp, d, pw, 5, &out, &r);
	}
	free(g);
	or_compat_bvh_free((void *)c.bvh);
	return out.n;
}

/* ---------------------------------------------------------- photon map -- */

/* PhotonMap::cellKey (src/Photon.cpp:5-8). */
static long long cell_key(int x, int y, int z) {
#define RN_F(v) ((long long)(((v) + (1 << 20)) & 0x1FFFFF))
	return (RN_F(x) << 42) | (RN_F(y) << 21) | RN_F(z);
#undef RN_F
}

typedef struct {
	const float *pos, *power, *dir;
	float cell;
	int n;
	long long *key; /* per sorted slot */
	int *idx;       /* photon index per sorted slot */
} pmap;

static const long long *g_sort_keys;
static int by_key(const void *a, const void *b) {
	int i = *(const int *)a, j = *(const int *)b;
	long long ka = g_sort_keys[i], kb = g_sort_keys[j];
	if (ka != kb)
		return ka < kb ? -1 : 1;
	return i < j ? -1 : (i > j);
}

static v3 cell_of(v3 p, float cell) {
	return mk(p.x / cell, p.y / cell, p.z / cell);
}

/* PhotonMap::build (src/Photon.cpp:10-19): the hash map's buckets, each
 * holding its photons in index order, become runs of (key, index)-sorted
 * slots. */
static void pmap_build(pmap *m, int n, const float *pos, const float *power, const float *dir, float cell) {
	m->pos = pos, m->power = power, m->dir = dir, m->cell = cell, m->n = n;
	long long *k = (long long *)malloc(sizeof(long long) * (size_t)(n > 0 ? n : 1));
	m->idx = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
	m->key = (long long *)malloc(sizeof(long long) * (size_t)(n > 0 ? n : 1));
	for (int i = 0; i < n; i++) {
		v3 c = cell_of(ld(pos + 3 * i), cell);
		k[i] = cell_key((int)floorf(c.x), (int)floorf(c.y), (int)floorf(c.z));
		m->idx[i] = i;
	}
	g_sort_keys = k;
	qsort(m->idx, (size_t)n, sizeof(int), by_key);
	for (int i = 0; i < n; i++)
		m->key[i] = k[m->idx[i]];
	free(k);
}

static void pmap_free(pmap *m) {
	free(m->idx);
	free(m->key);
}

/* PhotonMap::gather (src/Photon.cpp:21-46). */
static v3 pmap_gather(const pmap *m, v3 point, v3 normal, float radius) {
	v3 c = cell_of(point, m->cell);
	int cx = (int)floorf(c.x), cy = (int)floorf(c.y), cz = (int)floorf(c.z);
	v3 sum = mk(0, 0, 0);
	float r2 = radius * radius;
	for (int dx = -1; dx <= 1; dx++)
		for (int dy = -1; dy <= 1; dy++)
			for (int dz = -1; dz <= 1; dz++) {
				long long key = cell_key(cx + dx, cy + dy, cz + dz);
				int lo = 0, hi = m->n;
				while (lo < hi) {
					int mid = lo + (hi - lo) / 2;
					if (m->key[mid] < key)
						lo = mid + 1;
					else
						hi = mid;
				}
				for (int s = lo; s < m->n && m->key[s] == key; s++) {
					int i = m->idx[s];
					v3 d = sub(ld(m->pos + 3 * i), point);
					if (dot(d, d) > r2)
						continue;
					if (dot(normal, muls(ld(m->dir + 3 * i), -1.0f)) <= 0.0f)
						continue;
					sum = add(sum, ld(m->power + 3 * i));
				}
			}
	return divs(sum, 3.14159265f * r2);
}

void or_photon_gather(int n, const float *pos, const float *power, const float *dir, float cell, int m,
                      const float *points, const float *normals, float radius, float *out) {
	pmap pm;
	pmap_build(&pm, n, pos, power, dir, cell);
	for (int i = 0; i < m; i++) {
		v3 g = pmap_gather(&pm, ld(points + 3 * i), ld(normals + 3 * i), radius);
		out[3 * i] = g.x, out[3 * i + 1] = g.y, out[3 * i + 2] = g.z;
	}
	pmap_free(&pm);
}

/* ---------------------------------------------------------- pass 2 ---- */

/* photonShade, src/Renderer.cpp:1233-1294. */
static v3 photon_shade(const cctx *c, const pmap *m, v3 o, v3 d, float radius, int depth, int gather_rays,
                       rng_t *r) {
	const or_compat_scene *s = c->s;
	float t, hu, hv;
	int ti = closest(c, o, d, &t, &hu, &hv);
	if (ti < 0)
		return divs(environment(d), 255.0f);
	v3 point = add(o, muls(d, t));
	v3 N = ld(s->normals + 3 * ti);
	v3 n = face_viewer(N, d);
	int mat = s->material[ti];
	if (depth > 0 && mat == M_MIRROR) {
		v3 rr = reflect(d, n);
		return photon_shade(c, m, add(point, muls(rr, 1e-4f)), rr, radius, depth - 1, gather_rays, r);
	}
	if (depth > 0 && mat == M_METAL) {
		v3 albedo = divs(base_colour(s, ti, hu, hv, point, muls(d, -1.0f)), 255.0f);
		v3 rr = reflect(d, n);
		return mul(albedo, photon_shade(c, m, add(point, muls(rr, 1e-4f)), rr, radius, depth - 1, gather_rays, r));
	}
	if (depth > 0 && mat == M_GLASS) {
		v3 nn, refr;
		float fresnel = glass_setup(d, N, &nn, &refr);
		v3 reflected = reflect(d, nn);
		v3 refl_col =
		    photon_shade(c, m, add(point, muls(reflected, 1e-4f)), reflected, radius, depth - 1, gather_rays, r);
		if (length(refr) < 1e-6f)
			return refl_col;
		v3 refr_col = photon_shade(c, m, add(point, muls(refr, 1e-4f)), refr, radius, depth - 1, gather_rays, r);
		return add(muls(refl_col, fresnel), muls(refr_col, 1.0f - fresnel));
	}
	v3 albedo = divs(base_colour(s, ti, hu, hv, point, muls(d, -1.0f)), 255.0f);
	v3 direct = direct_light(c, point, n, albedo, ti);
	v3 indirect;
	if (gather_rays > 0) {
		v3 gathered = mk(0, 0, 0);
		for (int i = 0; i < gather_rays; i++) {
			v3 gdir = cosine_hemisphere(n, r);
			float gt, gu, gv;
			v3 go = add(point, muls(gdir, 1e-4f));
			int gi = closest(c, go, gdir, &gt, &gu, &gv);
			if (gi >= 0) {
				v3 gp = add(go, muls(gdir, gt));
				gathered = add(gathered, pmap_gather(m, gp, face_viewer(ld(s->normals + 3 * gi), gdir), radius));
			} else {
				gathered = add(gathered, divs(environment(gdir), 255.0f));
			}
		}
		indirect = divs(mul(albedo, gathered), (float)gather_rays);
	} else {
		indirect = mul(albedo, pmap_gather(m, point, n, radius));
	}
	return add(direct, indirect);
}

static void put_pixel(unsigned char *rgba, int W, int x, int y, v3 col) {
	int r = (int)col.x < 255 ? (int)col.x : 255;
	int g = (int)col.y < 255 ? (int)col.y : 255;
	int b = (int)col.z < 255 ? (int)col.z : 255;
	unsigned char *px = rgba + 4 * ((size_t)y * (size_t)W + (size_t)x);
	px[0] = (unsigned char)(r & 0xFF);
	px[1] = (unsigned char)(g & 0xFF);
	px[2] = (unsigned char)(b & 0xFF);
	px[3] = 255;
}

/* Pass 2 of renderPhotonMapped (src/Renderer.cpp:1328-1345) over rows
 * [y0, y1) with rn_render's camera and the gather radius 0.25 (also the
 * cell size of the map). */
void or_photon_render(const or_compat_scene *s, int n, const float *pos, const float *power, const float *dir,
                      int W, int H, float cam_z, int gather_rays, int rng, uint32_t key0, uint32_t key1, int y0,
                      int y1, unsigned char *rgba) {
	const float radius = 0.25f;
	cctx c;
	c.s = s;
	c.bvh = or_compat_bvh_build(s->verts, s->n);
	pmap pm;
	pmap_build(&pm, n, pos, power, dir, radius);
#pragma omp parallel for schedule(dynamic, 4)
	for (int y = y0; y < y1; y++) {
		mt19937 *g = (mt19937 *)malloc(sizeof(mt19937));
		mt_seed(g, 1000u + (uint32_t)y);
		rng_t r = {rng, g, {key0, key1}, 0, 0, 0, 0};
		for (int x = 0; x < W; x++) {
			if (rng == OR_RNG_PHILOX)
				stream(&r, (uint32_t)(y * W + x), 0u, TAG_GATHER);
			float o[3], d[3];
			or_compat_primary_ray(x, y, W, H, cam_z, -0.5f, -0.5f, o, d);
			v3 col = muls(photon_shade(&c, &pm, ld(o), ld(d), radius, 4, gather_rays, &r), 255.0f);
			put_pixel(rgba, W, x, y, col);
		}
		free(g);
	}
	pmap_free(&pm);
	or_compat_bvh_free((void *)c.bvh);
}

/* ------------------------------------------------------- bidirectional -- */

typedef struct {
	v3 position, normal, albedo, throughput;
	int is_light;
} vertex;

/* faceViewer of src/BDPT.cpp:34-36 (the flip on a zero dot product differs
 * from the renderer's). */
static v3 face_viewer_bd(v3 n, v3 d) {
	return dot(n, d) < 0.0f ? n : muls(n, -1.0f);
}

/* tracePath, src/BDPT.cpp:42-62. */
static int trace_subpath(const cctx *c, v3 o, v3 d, v3 beta, int max_depth, vertex *out, int nv, rng_t *r) {
	const or_compat_scene *s = c->s;
	const float eps = 1e-3f;
	for (int depth = 0; depth < max_depth; depth++) {
		float t, hu, hv;
		int ti = closest(c, o, d, &t, &hu, &hv);
		if (ti < 0)
			return nv;
		vertex v;
		v.position = add(o, muls(d, t));
		v.normal = face_viewer_bd(ld(s->normals + 3 * ti), d);
		v.albedo = divs(mk((float)s->colour[3 * ti], (float)s->colour[3 * ti + 1], (float)s->colour[3 * ti + 2]),
		                255.0f);
		v.throughput = beta;
		v.is_light = 0;
		out[nv++] = v;
		v3 nd = cosine_hemisphere(v.normal, r);
		beta = mul(beta, v.albedo);
		o = add(v.position, muls(nd, eps));
		d = nd;
	}
	return nv;
}

/* renderBidirectional, src/BDPT.cpp:66-144, over rows [y0, y1) with
 * rn_render's camera. The reference draws the two jitters of the camera ray
 * and the two coordinates of the light point inside one constructor call
 * each; GCC evaluates those arguments right to left, so the second
 * coordinate takes the first draw. */
void or_bdpt_render(const or_compat_scene *s, int W, int H, float cam_z, int samples, float exposure, int rng,
                    uint32_t key0, uint32_t key1, int y0, int y1, unsigned char *rgba) {
	const float kPi = 3.14159265f, eps = 1e-3f;
	const v3 light_centre = mk(0.0f, 0.33f, 0.0f), light_normal = mk(0.0f, -1.0f, 0.0f);
	const float light_half = 0.18f;
	const float light_area = (2.0f * light_half) * (2.0f * light_half);
	const v3 Le = mk(30.0f, 30.0f, 30.0f);
	const int max_depth = 4;
	cctx c;
	c.s = s;
	c.bvh = or_compat_bvh_build(s->verts, s->n);
#pragma omp parallel for schedule(dynamic, 4)
	for (int y = y0; y < y1; y++) {
		mt19937 *g = (mt19937 *)malloc(sizeof(mt19937));
		mt_seed(g, 9871u + (uint32_t)y);
		rng_t r = {rng, g, {key0, key1}, 0, 0, 0, 0};
		for (int x = 0; x < W; x++) {
			v3 accum = mk(0, 0, 0);
			for (int si = 0; si < samples; si++) {
				if (rng == OR_RNG_PHILOX)
					stream(&r, (uint32_t)(y * W + x), (uint32_t)si, TAG_BDPT);
				float jy = U(&r);
				float jx = U(&r);
				float o[3], d[3]; /* src/BDPT.cpp:90-91 */
				or_compat_camera_ray((float)x + jx - (float)W / 2.0f, -((float)y + jy - (float)H / 2.0f), cam_z, o, d);
				vertex cam[4], lig[5];
				int nc = trace_subpath(&c, ld(o), ld(d), mk(1, 1, 1), max_depth, cam, 0, &r);
				float lz = U(&r);
				float lx = U(&r);
				v3 lp = add(light_centre, mk((lx * 2.0f - 1.0f) * light_half, 0.0f, (lz * 2.0f - 1.0f) * light_half));
				vertex l0;
				l0.position = lp;
				l0.normal = light_normal;
				l0.albedo = mk(0, 0, 0);
				l0.throughput = muls(Le, light_area);
				l0.is_light = 1;
				lig[0] = l0;
				v3 ldir = cosine_hemisphere(light_normal, &r);
				int nl = trace_subpath(&c, add(lp, muls(ldir, eps)), ldir, muls(Le, light_area), max_depth, lig, 1, &r);
				for (int i = 0; i < nc; i++) {
					const vertex *cv = &cam[i];
					for (int j = 0; j < nl; j++) {
						const vertex *lv = &lig[j];
						v3 dd = sub(lv->position, cv->position);
						float dist2 = dot(dd, dd);
						float dist = sqrtf(dist2);
						dd = divs(dd, dist);
						float cos_c = dot(cv->normal, dd);
						float cos_l = dot(lv->normal, muls(dd, -1.0f));
						if (cos_c <= 0.0f || cos_l <= 0.0f)
							continue;
						v3 so = add(cv->position, muls(cv->normal, eps));
						float sa[3] = {so.x, so.y, so.z}, da[3] = {dd.x, dd.y, dd.z};
						if (or_compat_bvh_occluded(c.bvh, sa, da, dist - 2.0f * eps, -1))
							continue;
						float G = cos_c * cos_l / dist2;
						v3 fc = divs(cv->albedo, kPi);
						v3 contrib = mul(muls(mul(cv->throughput, fc), G), lv->throughput);
						if (!lv->is_light)
							contrib = mul(contrib, divs(lv->albedo, kPi));
						float w = 1.0f / (float)(i + j + 2);
						accum = add(accum, muls(contrib, w));
					}
				}
			}
			v3 col = muls(accum, exposure / (float)samples);
			col = muls(mk(col.x / (col.x + 1.0f), col.y / (col.y + 1.0f), col.z / (col.z + 1.0f)), 255.0f);
			put_pixel(rgba, W, x, y, col);
		}
		free(g);
	}
	or_compat_bvh_free((void *)c.bvh);
}
