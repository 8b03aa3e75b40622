/* Helpers shared by the oracle's restatements of the reference renderers
 * (compat.c: the path-trace call; photon.c: photon mapping and
 * bidirectional path tracing): glm vector operations in glm's FP32
 * operation order, surface colour, the default light and the sky.
 *
 * TEST INFRASTRUCTURE ONLY -- see oracle.h. */
#pragma once

#include "oracle.h"

#include <math.h>
#include <stddef.h>
#include <stdint.h>

typedef struct {
	float x, y, z;
} v3;

static v3 mk(float x, float y, float z) {
	v3 r = {x, y, z};
	return r;
}
static v3 ld(const float *p) {
	return mk(p[0], p[1], p[2]);
}
static v3 add(v3 a, v3 b) {
	return mk(a.x + b.x, a.y + b.y, a.z + b.z);
}
static v3 sub(v3 a, v3 b) {
	return mk(a.x - b.x, a.y - b.y, a.z - b.z);
}
static v3 mul(v3 a, v3 b) {
	return mk(a.x * b.x, a.y * b.y, a.z * b.z);
}
static v3 muls(v3 a, float s) {
	return mk(a.x * s, a.y * s, a.z * s);
}
static v3 divs(v3 a, float s) {
	return mk(a.x / s, a.y / s, a.z / s);
}
/* glm dot: tmp = a*b; tmp.x + tmp.y + tmp.z (func_geometric.inl:48-55). */
static float dot(v3 a, v3 b) {
	v3 t = mul(a, b);
	return t.x + t.y + t.z;
}
/* glm cross (func_geometric.inl:68-79). */
static v3 cross(v3 x, v3 y) {
	return mk(x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y);
}
/* glm normalize: v * (1 / sqrt(dot(v,v))) (func_geometric.inl:82-90,
 * func_exponential.inl:44-49). */
static v3 normalize(v3 v) {
	float s = 1.0f / sqrtf(dot(v, v));
	return muls(v, s);
}
static float length(v3 v) {
	return sqrtf(dot(v, v));
}
/* glm reflect: I - N * dot(N, I) * 2 (func_geometric.inl:104-110). */
static v3 reflect(v3 i, v3 n) {
	return sub(i, muls(muls(n, dot(n, i)), 2.0f));
}
/* glm refract, vector form (func_geometric.inl:113-123). */
static v3 refract(v3 i, v3 n, float eta) {
	float dv = dot(n, i);
	float k = 1.0f - eta * eta * (1.0f - dv * dv);
	if (k >= 0.0f)
		return sub(muls(i, eta), muls(n, eta * dv + sqrtf(k)));
	return mk(0, 0, 0);
}
/* glm mix: x*(1-a) + y*a (func_common.inl:104-112). */
static v3 mixv(v3 x, v3 y, float a) {
	return add(muls(x, 1.0f - a), muls(y, a));
}
static float gmin(float x, float y) { /* glm::min: (y < x) ? y : x */
	return (y < x) ? y : x;
}
static float gmax(float x, float y) { /* glm::max: (x < y) ? y : x */
	return (x < y) ? y : x;
}
static v3 vmin(v3 a, v3 b) {
	return mk(gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z));
}
static v3 vmax(v3 a, v3 b) {
	return mk(gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z));
}
static float comp(v3 v, int a) {
	return a == 0 ? v.x : (a == 1 ? v.y : v.z);
}


/* ------------------------------------------------- reference estimator -- */

enum { M_DIFFUSE, M_MIRROR, M_GLASS, M_PROCEDURAL, M_BUMP, M_PARALLAX, M_METAL, M_DISPERSIVE, M_SUBSURFACE,
       M_NORMALMAP, M_EMISSIVE };

static float unif(uint32_t x) {
	return (float)(x >> 8) * (1.0f / 16777216.0f);
}

typedef struct {
	const or_compat_scene *s;
	const void *bvh;
} cctx;

static v3 tex_lookup(const or_compat_scene *s, int id, float u, float v) { /* src/Renderer.cpp:584-591 */
	int tw = s->tex_w[id], th = s->tex_h[id];
	int tx = (int)(u * (float)(tw - 1));
	int ty = (int)((1.0f - v) * (float)(th - 1));
	tx = tx < 0 ? 0 : (tx > tw - 1 ? tw - 1 : tx);
	ty = ty < 0 ? 0 : (ty > th - 1 ? th - 1 : ty);
	uint32_t p = s->tex[id][(size_t)ty * (size_t)tw + (size_t)tx];
	return mk((float)((p >> 16) & 0xFF), (float)((p >> 8) & 0xFF), (float)(p & 0xFF));
}

/* src/Renderer.cpp:595-612 */
static void parallax(const or_compat_scene *s, int ti, float *tu, float *tv, v3 view) {
	const float *V = s->verts + 9 * ti, *tp = s->texcoords + 6 * ti;
	v3 e1 = sub(ld(V + 3), ld(V)), e2 = sub(ld(V + 6), ld(V));
	float d1x = tp[2] - tp[0], d1y = tp[3] - tp[1], d2x = tp[4] - tp[0], d2y = tp[5] - tp[1];
	float det = d1x * d2y - d2x * d1y;
	if (fabsf(det) < 1e-8f)
		return;
	float f = 1.0f / det;
	v3 T = normalize(muls(sub(muls(e1, d2y), muls(e2, d1y)), f));
	v3 B = normalize(muls(add(muls(e1, -d2x), muls(e2, d1x)), f));
	v3 n = ld(s->normals + 3 * ti);
	float vx = dot(view, T), vy = dot(view, B), vz = dot(view, n);
	if (fabsf(vz) < 1e-4f)
		return;
	float h = tex_lookup(s, s->texture_id[ti], *tu, *tv).x / 255.0f;
	*tu = *tu - vx / vz * (h * 0.05f);
	*tv = *tv - vy / vz * (h * 0.05f);
}

/* src/Renderer.cpp:566-571 */
static v3 procedural(v3 p) {
	v3 q = muls(p, 3.0f);
	float qa[3] = {q.x, q.y, q.z};
	float n = or_fractal_noise(qa, 4);
	float rings = 0.5f + 0.5f * sinf((q.x + q.y) * 3.0f + n * 8.0f);
	return mixv(mk(60.0f, 30.0f, 12.0f), mk(205.0f, 150.0f, 90.0f), rings);
}

/* src/Renderer.cpp:642-660 */
static v3 base_colour(const or_compat_scene *s, int ti, float u, float v, v3 point, v3 view) {
	if (s->texture_id[ti] >= 0) {
		const float *tp = s->texcoords + 6 * ti;
		float w0 = 1.0f - u - v, w1 = u, w2 = v;
		float tu = w0 * tp[0] + w1 * tp[2] + w2 * tp[4];
		float tv = w0 * tp[1] + w1 * tp[3] + w2 * tp[5];
		if (s->material[ti] == M_PARALLAX)
			parallax(s, ti, &tu, &tv, view);
		return tex_lookup(s, s->texture_id[ti], tu, tv);
	}
	if (s->material[ti] == M_PROCEDURAL)
		return procedural(point);
	return mk((float)s->colour[3 * ti], (float)s->colour[3 * ti + 1], (float)s->colour[3 * ti + 2]);
}

/* src/Renderer.cpp:538-540 */
static v3 face_viewer(v3 n, v3 d) {
	return (dot(n, d) > 0.0f) ? muls(n, -1.0f) : n;
}

/* src/Renderer.cpp:778-788 with (u1,u2) supplied. */
static v3 cos_hemi(v3 n, float u1, float u2) {
	float r = sqrtf(u1);
	float theta = 2.0f * 3.14159265f * u2;
	float x = r * cosf(theta), y = r * sinf(theta), z = sqrtf(fmaxf(0.0f, 1.0f - u1));
	v3 up = fabsf(n.z) < 0.99f ? mk(0, 0, 1) : mk(1, 0, 0);
	v3 t = normalize(cross(up, n));
	v3 b = cross(n, t);
	return normalize(add(add(muls(t, x), muls(b, y)), muls(n, z)));
}

/* Default light of renderPathTraced (src/Renderer.cpp:1129-1134, src/Light.h:13-29). */
static const float kLightPos[3] = {0.0f, 0.95f, 0.0f};
static const float kLightRadius = 0.15f, kLightIntensity = 40.0f, kLightKq = 12.566370f;

/* Disk branch of visibility (src/Renderer.cpp:483-497). */
static float visibility(const cctx *c, v3 point, int ignore) {
	static const float ks[9][2] = {{0, 0},       {1, 0},        {-1, 0},      {0, 1},       {0, -1},
	                               {0.7f, 0.7f}, {-0.7f, 0.7f}, {0.7f, -0.7f}, {-0.7f, -0.7f}};
	v3 lp = ld(kLightPos);
	v3 axis = normalize(sub(point, lp));
	v3 up = fabsf(axis.y) < 0.99f ? mk(0, 1, 0) : mk(1, 0, 0);
	v3 tangent = normalize(cross(up, axis));
	v3 bitangent = cross(axis, tangent);
	int un = 0;
	for (int i = 0; i < 9; i++) {
		v3 sp = add(lp, muls(add(muls(tangent, ks[i][0]), muls(bitangent, ks[i][1])), kLightRadius));
		v3 to = sub(sp, point);
		float dist = length(to);
		v3 dir = divs(to, dist);
		float pa[3] = {point.x, point.y, point.z}, da[3] = {dir.x, dir.y, dir.z};
		if (!or_compat_bvh_occluded(c->bvh, pa, da, dist, ignore))
			un++;
	}
	return (float)un / 9.0f;
}

/* src/Renderer.cpp:791-814 for the single default light. */
static v3 direct_light(const cctx *c, v3 point, v3 n, v3 albedo, int ignore) {
	v3 to = sub(ld(kLightPos), point);
	float dist = length(to);
	v3 ldir = divs(to, dist);
	float att = kLightIntensity / (0.0f + 0.0f * dist + kLightKq * dist * dist);
	float vis = visibility(c, point, ignore);
	float inc = fmaxf(0.0f, dot(n, ldir));
	float s = vis * 1.0f * att * inc;
	v3 sum = mk(s, s, s);
	return mul(albedo, vmin(sum, mk(1, 1, 1)));
}

/* src/Renderer.cpp:555-563 (sky model off). */
static v3 environment(v3 d) {
	float t = 0.5f * (d.y + 1.0f);
	return mixv(mk(200.0f, 210.0f, 235.0f), mk(70.0f, 120.0f, 205.0f), t);
}

/* ---------------------------------------------------- std::mt19937 -- */

/* MT19937 (Matsumoto and Nishimura, ACM TOMACS 8(1), 1998) with the seeding
 * of std::mersenne_twister_engine ([rand.eng.mers]). */
typedef struct {
	uint32_t x[624];
	int i;
} mt19937;

static void mt_seed(mt19937 *g, uint32_t s) {
	g->x[0] = s;
	for (int i = 1; i < 624; i++)
		g->x[i] = 1812433253u * (g->x[i - 1] ^ (g->x[i - 1] >> 30)) + (uint32_t)i;
	g->i = 624;
}

static uint32_t mt_next(mt19937 *g) {
	if (g->i >= 624) {
		for (int k = 0; k < 624; k++) {
			uint32_t y = (g->x[k] & 0x80000000u) | (g->x[(k + 1) % 624] & 0x7FFFFFFFu);
			uint32_t v = g->x[(k + 397) % 624] ^ (y >> 1);
			if (y & 1u)
				v ^= 0x9908B0DFu;
			g->x[k] = v;
		}
		g->i = 0;
	}
	uint32_t y = g->x[g->i++];
	y ^= y >> 11;
	y ^= (y << 7) & 0x9D2C5680u;
	y ^= (y << 15) & 0xEFC60000u;
	y ^= y >> 18;
	return y;
}

/* libstdc++ generate_canonical<float, 24> over a 32-bit engine
 * (bits/random.tcc): one draw, float(x) / 2^32, and the largest float
 * below 1 when the rounding reaches 1. */
static float mt_unif(mt19937 *g) {
	float r = (float)mt_next(g) / 4294967296.0f;
	if (r >= 1.0f)
		r = nextafterf(1.0f, 0.0f);
	return r;
}
