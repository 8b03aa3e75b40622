/* Oracle: restatement of the reference path-trace call in plain C, FP32, with
 * glm's operation order kept wherever the GPU result must be bit-identical
 * (ray generation, triangle test, texel selection). Compiled with
 * -ffp-contract=off so that no multiply-add is fused, like the reference's
 * x86-64 build without -march (CMakeLists.txt:157-209).
 *
 * TEST INFRASTRUCTURE ONLY -- see oracle.h. */
#include "compat_impl.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------ triangle test -- */

/* src/Geometry.cpp:11-25. M = mat3(-d, e0, e1) in column order m[c][r];
 * det by func_matrix.inl:253-262, inverse by :322-344, M^-1 * sp by
 * type_mat3x3.inl:468-474. */
int or_compat_intersect_tri(const float o[3], const float d[3], const float v0p[3], const float v1p[3],
                            const float v2p[3], float *t, float *u, float *v) {
	v3 v0 = ld(v0p), v1 = ld(v1p), v2 = ld(v2p);
	v3 e0 = sub(v1, v0), e1 = sub(v2, v0), sp = sub(ld(o), v0);
	float m00 = -d[0], m01 = -d[1], m02 = -d[2];
	float m10 = e0.x, m11 = e0.y, m12 = e0.z;
	float m20 = e1.x, m21 = e1.y, m22 = e1.z;
	float det = +m00 * (m11 * m22 - m21 * m12) - m10 * (m01 * m22 - m21 * m02) + m20 * (m01 * m12 - m11 * m02);
	if (fabsf(det) < 1e-8f)
		return 0;
	float ood = 1.0f / (+m00 * (m11 * m22 - m21 * m12) - m10 * (m01 * m22 - m21 * m02) +
	                    m20 * (m01 * m12 - m11 * m02));
	float i00 = +(m11 * m22 - m21 * m12) * ood;
	float i10 = -(m10 * m22 - m20 * m12) * ood;
	float i20 = +(m10 * m21 - m20 * m11) * ood;
	float i01 = -(m01 * m22 - m21 * m02) * ood;
	float i11 = +(m00 * m22 - m20 * m02) * ood;
	float i21 = -(m00 * m21 - m20 * m01) * ood;
	float i02 = +(m01 * m12 - m11 * m02) * ood;
	float i12 = -(m00 * m12 - m10 * m02) * ood;
	float i22 = +(m00 * m11 - m10 * m01) * ood;
	float rx = i00 * sp.x + i10 * sp.y + i20 * sp.z;
	float ry = i01 * sp.x + i11 * sp.y + i21 * sp.z;
	float rz = i02 * sp.x + i12 * sp.y + i22 * sp.z;
	*t = rx;
	*u = ry;
	*v = rz;
	return (ry >= 0.0f) && (rz >= 0.0f) && (ry + rz <= 1.0f) && (rx > 1e-5f);
}

/* ------------------------------------------------------------- camera -- */

/* Camera(w,h,2,(0,0,cam_z)).lookAt(0) (src/capi.cpp:49-50, src/Camera.cpp:27-33)
 * then primaryRay (src/Camera.cpp:69-79) with f = focalLength*scale = 300
 * (src/Camera.h:17-18) and the image-plane offsets of src/Renderer.cpp:1148. */
void or_compat_camera_ray(float sx, float sy, float cam_z, float o[3], float d[3]) {
	v3 pos = mk(0.0f, 0.0f, cam_z);
	v3 fwd = normalize(sub(pos, mk(0, 0, 0)));
	v3 right = normalize(cross(mk(0.0f, 1.0f, 0.0f), fwd));
	v3 up = cross(fwd, right);
	float f = 2.0f * 150.0f;
	float a = sx / f, b = sy / f, c = -1.0f;
	/* mat3(right, up, fwd) * (a,b,c): row sums m[0][r]*a + m[1][r]*b + m[2][r]*c */
	v3 dir = mk(right.x * a + up.x * b + fwd.x * c, right.y * a + up.y * b + fwd.y * c,
	            right.z * a + up.z * b + fwd.z * c);
	dir = normalize(dir);
	o[0] = pos.x;
	o[1] = pos.y;
	o[2] = pos.z;
	d[0] = dir.x;
	d[1] = dir.y;
	d[2] = dir.z;
}

void or_compat_primary_ray(int x, int y, int W, int H, float cam_z, float jx, float jy, float o[3], float d[3]) {
	float sx = (float)x + 0.5f + jx - (float)W / 2.0f;
	float sy = -((float)y + 0.5f + jy - (float)H / 2.0f);
	or_compat_camera_ray(sx, sy, cam_z, o, d);
}

/* ---------------------------------------------------------- brute force -- */

int or_compat_closest_brute(const float *verts, int n, const float o[3], const float d[3], int ignore, float *t,
                            float *u, float *v) {
	int best = -1;
	float bt = INFINITY, bu = 0, bv = 0;
	for (int i = 0; i < n; i++) {
		if (i == ignore)
			continue;
		float tt, uu, vv;
		if (or_compat_intersect_tri(o, d, verts + 9 * i, verts + 9 * i + 3, verts + 9 * i + 6, &tt, &uu, &vv) &&
		    tt < bt) {
			best = i;
			bt = tt;
			bu = uu;
			bv = vv;
		}
	}
	*t = bt;
	*u = bu;
	*v = bv;
	return best;
}

/* ------------------------------------------------- median-split BVH -- */

typedef struct {
	v3 bmin, bmax;
	int left, right, start, count;
} cnode;

typedef struct {
	const float *verts;
	int n;
	v3 *centroids;
	int *order;
	cnode *nodes;
	int nnodes, cap;
	int root;
} cbvh;

static int push_node(cbvh *b, cnode nd) {
	if (b->nnodes == b->cap) {
		b->cap = b->cap ? 2 * b->cap : 64;
		b->nodes = (cnode *)realloc(b->nodes, sizeof(cnode) * (size_t)b->cap);
	}
	b->nodes[b->nnodes] = nd;
	return b->nnodes++;
}

/* Quickselect on order[lo,hi) so that order[k] holds the element of rank k
 * by centroid[axis] (strict <), smaller ones before and larger after --
 * the partition std::nth_element guarantees (src/BVH.cpp:43-44). */
static void select_k(cbvh *b, int lo, int hi, int k, int axis) {
	int *a = b->order;
	while (hi - lo > 1) {
		int mid = lo + (hi - lo) / 2;
		float pa = comp(b->centroids[a[lo]], axis), pb = comp(b->centroids[a[mid]], axis),
		      pc = comp(b->centroids[a[hi - 1]], axis);
		float pivot = (pa < pb) ? ((pb < pc) ? pb : ((pa < pc) ? pc : pa)) : ((pa < pc) ? pa : ((pb < pc) ? pc : pb));
		int i = lo, lt = lo, gt = hi; /* three-way partition */
		while (i < gt) {
			float c = comp(b->centroids[a[i]], axis);
			if (c < pivot) {
				int tmp = a[lt];
				a[lt] = a[i];
				a[i] = tmp;
				lt++;
				i++;
			} else if (pivot < c) {
				gt--;
				int tmp = a[gt];
				a[gt] = a[i];
				a[i] = tmp;
			} else {
				i++;
			}
		}
		if (k < lt)
			hi = lt;
		else if (k >= gt)
			lo = gt;
		else
			return;
	}
}

/* src/BVH.cpp:18-63 */
static int cbuild(cbvh *b, int start, int count) {
	cnode nd;
	memset(&nd, 0, sizeof nd);
	nd.left = nd.right = -1;
	v3 bmin = mk(1e30f, 1e30f, 1e30f), bmax = mk(-1e30f, -1e30f, -1e30f);
	for (int i = start; i < start + count; i++) {
		const float *t = b->verts + 9 * b->order[i];
		for (int k = 0; k < 3; k++) {
			bmin = vmin(bmin, ld(t + 3 * k));
			bmax = vmax(bmax, ld(t + 3 * k));
		}
	}
	nd.bmin = bmin;
	nd.bmax = bmax;
	if (count <= 4) {
		nd.start = start;
		nd.count = count;
		return push_node(b, nd);
	}
	v3 cmin = mk(1e30f, 1e30f, 1e30f), cmax = mk(-1e30f, -1e30f, -1e30f);
	for (int i = start; i < start + count; i++) {
		cmin = vmin(cmin, b->centroids[b->order[i]]);
		cmax = vmax(cmax, b->centroids[b->order[i]]);
	}
	v3 ext = sub(cmax, cmin);
	int axis = (ext.x > ext.y && ext.x > ext.z) ? 0 : (ext.y > ext.z ? 1 : 2);
	int mid = start + count / 2;
	select_k(b, start, start + count, mid, axis);
	nd.count = 0;
	int me = push_node(b, nd);
	int l = cbuild(b, start, mid - start);
	int r = cbuild(b, mid, start + count - mid);
	b->nodes[me].left = l;
	b->nodes[me].right = r;
	return me;
}

int or_compat_all_hits(const float *verts, int n, const float o[3], const float d[3], int *prims, float *ts, int cap) {
	int k = 0;
	for (int i = 0; i < n; i++) {
		float tt, uu, vv;
		if (or_compat_intersect_tri(o, d, verts + 9 * i, verts + 9 * i + 3, verts + 9 * i + 6, &tt, &uu, &vv)) {
			if (k < cap) {
				prims[k] = i;
				ts[k] = tt;
			}
			k++;
		}
	}
	return k;
}

void *or_compat_bvh_build(const float *verts, int n) {
	cbvh *b = (cbvh *)calloc(1, sizeof(cbvh));
	b->verts = verts;
	b->n = n;
	b->root = -1;
	b->centroids = (v3 *)malloc(sizeof(v3) * (size_t)(n > 0 ? n : 1));
	b->order = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
	for (int i = 0; i < n; i++) {
		b->order[i] = i;
		const float *t = verts + 9 * i;
		b->centroids[i] = divs(add(add(ld(t), ld(t + 3)), ld(t + 6)), 3.0f);
	}
	if (n > 0)
		b->root = cbuild(b, 0, n);
	return b;
}

void or_compat_bvh_free(void *p) {
	cbvh *b = (cbvh *)p;
	if (!b)
		return;
	free(b->centroids);
	free(b->order);
	free(b->nodes);
	free(b);
}

/* src/BVH.cpp:66-75 with std::min/max({..}) and glm::min/max comparison forms. */
static int slab(v3 o, v3 inv, v3 bmin, v3 bmax, float tmax) {
	v3 t0 = mul(sub(bmin, o), inv);
	v3 t1 = mul(sub(bmax, o), inv);
	v3 ts = vmin(t0, t1), tb = vmax(t0, t1);
	float te = ts.x;
	if (te < ts.y)
		te = ts.y;
	if (te < ts.z)
		te = ts.z;
	float tx = tb.x;
	if (tb.y < tx)
		tx = tb.y;
	if (tb.z < tx)
		tx = tb.z;
	float lo = (te < 0.0f) ? 0.0f : te;
	return tx >= lo && te < tmax;
}

int or_compat_bvh_intersect(const void *p, const float o[3], const float d[3], int ignore, float *t, float *u,
                            float *v) {
	const cbvh *b = (const cbvh *)p;
	int best = -1;
	float bt = INFINITY, bu = 0, bv = 0;
	*t = bt;
	*u = 0;
	*v = 0;
	if (b->root < 0)
		return -1;
	v3 ro = ld(o), rd = ld(d);
	v3 inv = mk(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
	int stack[64], sp = 0;
	stack[sp++] = b->root;
	while (sp > 0) {
		const cnode *nd = &b->nodes[stack[--sp]];
		if (!slab(ro, inv, nd->bmin, nd->bmax, bt))
			continue;
		if (nd->count > 0) {
			for (int i = nd->start; i < nd->start + nd->count; i++) {
				int ti = b->order[i];
				if (ti == ignore)
					continue;
				float tt, uu, vv;
				const float *tv = b->verts + 9 * ti;
				if (or_compat_intersect_tri(o, d, tv, tv + 3, tv + 6, &tt, &uu, &vv) && tt < bt) {
					best = ti;
					bt = tt;
					bu = uu;
					bv = vv;
				}
			}
		} else {
			stack[sp++] = nd->left;
			stack[sp++] = nd->right;
		}
	}
	*t = bt;
	*u = bu;
	*v = bv;
	return best;
}

int or_compat_bvh_occluded(const void *p, const float o[3], const float d[3], float maxd, int ignore) {
	const cbvh *b = (const cbvh *)p;
	if (b->root < 0)
		return 0;
	v3 ro = ld(o), rd = ld(d);
	v3 inv = mk(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
	int stack[64], sp = 0;
	stack[sp++] = b->root;
	while (sp > 0) {
		const cnode *nd = &b->nodes[stack[--sp]];
		if (!slab(ro, inv, nd->bmin, nd->bmax, maxd))
			continue;
		if (nd->count > 0) {
			for (int i = nd->start; i < nd->start + nd->count; i++) {
				int ti = b->order[i];
				if (ti == ignore)
					continue;
				float tt, uu, vv;
				const float *tv = b->verts + 9 * ti;
				if (or_compat_intersect_tri(o, d, tv, tv + 3, tv + 6, &tt, &uu, &vv) && tt < maxd)
					return 1;
			}
		} else {
			stack[sp++] = nd->left;
			stack[sp++] = nd->right;
		}
	}
	return 0;
}

void or_compat_bvh_intersect_batch(const void *bvh, int nrays, const float *o, const float *d, const int *ignore,
                                   int *prim, float *t, float *u, float *v) {
	for (int i = 0; i < nrays; i++)
		prim[i] = or_compat_bvh_intersect(bvh, o + 3 * i, d + 3 * i, ignore ? ignore[i] : -1, t + i, u + i, v + i);
}

/* src/Renderer.cpp:642-656 (texture branch, no parallax) + :584-591. */
int or_compat_texel(float u, float v, const float tp[6], int tw, int th) {
	float w0 = 1.0f - u - v, w1 = u, w2 = v;
	float tu = w0 * tp[0] + w1 * tp[2] + w2 * tp[4];
	float tv = w0 * tp[1] + w1 * tp[3] + w2 * tp[5];
	int tx = (int)(tu * (float)(tw - 1));
	int ty = (int)((1.0f - tv) * (float)(th - 1));
	tx = tx < 0 ? 0 : (tx > tw - 1 ? tw - 1 : tx);
	ty = ty < 0 ? 0 : (ty > th - 1 ? th - 1 : ty);
	return ty * tw + tx;
}

/* --------------------------------------------------------------- noise -- */

/* Ken Perlin's improved-noise permutation (the published 2002 table that
 * src/Noise.cpp:7-30 also uses). */
static const unsigned char kPerm[256] = {
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

static int P(int i) {
	return kPerm[i & 255];
}
static float fade(float t) {
	return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f);
}
static float lerpf(float t, float a, float b) {
	return a + t * (b - a);
}
static float grad(int hash, float x, float y, float z) {
	int h = hash & 15;
	float u = h < 8 ? x : y;
	float v = h < 4 ? y : (h == 12 || h == 14 ? x : z);
	return ((h & 1) == 0 ? u : -u) + ((h & 2) == 0 ? v : -v);
}

/* src/Noise.cpp:44-61 */
float or_perlin(float x, float y, float z) {
	int X = (int)floorf(x) & 255, Y = (int)floorf(y) & 255, Z = (int)floorf(z) & 255;
	x -= floorf(x);
	y -= floorf(y);
	z -= floorf(z);
	float u = fade(x), v = fade(y), w = fade(z);
	int A = P(X) + Y, AA = P(A) + Z, AB = P(A + 1) + Z;
	int B = P(X + 1) + Y, BA = P(B) + Z, BB = P(B + 1) + Z;
	return lerpf(w,
	             lerpf(v, lerpf(u, grad(P(AA), x, y, z), grad(P(BA), x - 1, y, z)),
	                   lerpf(u, grad(P(AB), x, y - 1, z), grad(P(BB), x - 1, y - 1, z))),
	             lerpf(v, lerpf(u, grad(P(AA + 1), x, y, z - 1), grad(P(BA + 1), x - 1, y, z - 1)),
	                   lerpf(u, grad(P(AB + 1), x, y - 1, z - 1), grad(P(BB + 1), x - 1, y - 1, z - 1))));
}

/* src/Noise.cpp:246-255 */
float or_fractal_noise(const float p[3], int octaves) {
	float sum = 0.0f, amp = 1.0f, freq = 1.0f, norm = 0.0f;
	for (int i = 0; i < octaves; i++) {
		sum += amp * or_perlin(p[0] * freq, p[1] * freq, p[2] * freq);
		norm += amp;
		amp *= 0.5f;
		freq *= 2.0f;
	}
	return 0.5f + 0.5f * (sum / norm);
}

/* Philox layout of the reference estimator:
 *   key = (key0, key1); camera jitter: ctr = (pixel, sample, 0xFFFFFFFF, 0),
 *   words 0,1 -> jx, jy in [-0.5, 0.5); recursion level L (0 = camera ray):
 *   ctr = (pixel, sample, L, 0), words 0,1 -> cosine hemisphere, word 2 ->
 *   the Fresnel choice of glass. */
static v3 trace_path(const cctx *c, v3 o, v3 d, uint32_t pix, uint32_t smp, const uint32_t key[2]) {
	const or_compat_scene *s = c->s;
	v3 L = mk(0, 0, 0), thr = mk(1, 1, 1);
	int depth = 4; /* maxDepth, src/Renderer.cpp:1126 */
	for (uint32_t level = 0;; level++) {
		float oa[3] = {o.x, o.y, o.z}, da[3] = {d.x, d.y, d.z};
		float t, hu, hv;
		int ti = or_compat_bvh_intersect(c->bvh, oa, da, -1, &t, &hu, &hv);
		if (ti < 0) {
			L = add(L, mul(thr, divs(environment(d), 255.0f)));
			break;
		}
		uint32_t ctr[4] = {pix, smp, level, 0}, r[4];
		or_philox4x32_10(ctr, key, r);
		v3 point = add(o, muls(d, t));
		int mat = s->material[ti];
		v3 N = ld(s->normals + 3 * ti);
		v3 col = mk((float)s->colour[3 * ti], (float)s->colour[3 * ti + 1], (float)s->colour[3 * ti + 2]);
		if (mat == M_EMISSIVE) {
			L = add(L, mul(thr, muls(divs(col, 255.0f), 4.0f)));
			break;
		}
		if (mat == M_MIRROR && depth > 0) {
			v3 n = face_viewer(N, d);
			v3 rr = reflect(d, n);
			o = add(point, muls(rr, 1e-4f));
			d = rr;
			depth--;
			continue;
		}
		if ((mat == M_GLASS || mat == M_DISPERSIVE) && depth > 0) {
			float ior = 1.5f;
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
			float fres = r0 + (1.0f - r0) * powf(1.0f - cosi, 5.0f);
			v3 refr = refract(d, n, etai / etat);
			v3 nd = (length(refr) < 1e-6f || unif(r[2]) < fres) ? reflect(d, n) : refr;
			o = add(point, muls(nd, 1e-4f));
			d = nd;
			depth--;
			continue;
		}
		if (mat == M_METAL && depth > 0) {
			v3 n = face_viewer(N, d);
			v3 refl = reflect(d, n);
			v3 diff = cos_hemi(n, unif(r[0]), unif(r[1]));
			float rough = 0.2f; /* ModelTriangle::roughness default, framework/ModelTriangle.h:31 */
			float r2 = rough * rough;
			v3 nd = normalize(mixv(refl, diff, r2));
			if (dot(nd, n) <= 0.0f)
				nd = refl;
			v3 albedo = divs(base_colour(s, ti, hu, hv, point, muls(d, -1.0f)), 255.0f);
			thr = mul(thr, albedo);
			o = add(point, muls(nd, 1e-4f));
			d = nd;
			depth--;
			continue;
		}
		if (mat == M_SUBSURFACE) { /* src/Renderer.cpp:872-897 */
			v3 n = face_viewer(N, d);
			v3 albedo = divs(base_colour(s, ti, hu, hv, point, muls(d, -1.0f)), 255.0f);
			v3 to = sub(ld(kLightPos), point);
			float dist = length(to);
			v3 ldir = divs(to, dist);
			float att = kLightIntensity / (0.0f + 0.0f * dist + kLightKq * dist * dist);
			float wrap = 0.5f;
			float wd = fmaxf(0.0f, (dot(n, ldir) + wrap) / (1.0f + wrap));
			v3 shade = muls(albedo, visibility(c, point, ti) * att * wd);
			v3 eo = sub(point, muls(n, 1e-3f));
			float eoa[3] = {eo.x, eo.y, eo.z}, lda[3] = {ldir.x, ldir.y, ldir.z};
			float et, eu, ev;
			int eh = or_compat_bvh_intersect(c->bvh, eoa, lda, ti, &et, &eu, &ev);
			float thick = eh >= 0 ? et : 0.0f;
			shade = add(shade, muls(albedo, att * expf(-3.0f * thick) * 0.6f));
			v3 direct = vmin(shade, mk(1.5f, 1.5f, 1.5f));
			L = add(L, mul(thr, direct));
			if (depth <= 0)
				break;
			v3 sb = cos_hemi(n, unif(r[0]), unif(r[1]));
			thr = mul(thr, albedo);
			o = add(point, muls(sb, 1e-4f));
			d = sb;
			depth--;
			continue;
		}
		/* Diffuse (also Procedural/Bump/Parallax/NormalMap and depth-0
		 * Mirror/Glass/Metal), src/Renderer.cpp:899-906. */
		v3 n = face_viewer(N, d);
		v3 albedo = divs(base_colour(s, ti, hu, hv, point, muls(d, -1.0f)), 255.0f);
		v3 direct = direct_light(c, point, n, albedo, ti);
		L = add(L, mul(thr, direct));
		if (depth <= 0)
			break;
		v3 bounce = cos_hemi(n, unif(r[0]), unif(r[1]));
		thr = mul(thr, albedo);
		o = add(point, muls(bounce, 1e-4f));
		d = bounce;
		depth--;
	}
	return L;
}

/* src/Renderer.cpp:1119-1170 over rows [y0, y1). */
void or_compat_render(const or_compat_scene *s, int W, int H, float cam_z, int spp, uint32_t key0, uint32_t key1,
                      int y0, int y1, unsigned char *rgba) {
	cctx c;
	c.s = s;
	c.bvh = or_compat_bvh_build(s->verts, s->n);
	uint32_t key[2] = {key0, key1};
#pragma omp parallel for schedule(dynamic, 4)
	for (int y = y0; y < y1; y++) {
		for (int x = 0; x < W; x++) {
			uint32_t pix = (uint32_t)(y * W + x);
			v3 sum = mk(0, 0, 0);
			for (int si = 0; si < spp; si++) {
				uint32_t ctr[4] = {pix, (uint32_t)si, 0xFFFFFFFFu, 0}, r[4];
				or_philox4x32_10(ctr, key, r);
				float jx = unif(r[0]) - 0.5f, jy = unif(r[1]) - 0.5f;
				float o[3], d[3];
				or_compat_primary_ray(x, y, W, H, cam_z, jx, jy, o, d);
				sum = add(sum, trace_path(&c, ld(o), ld(d), pix, (uint32_t)si, key));
			}
			v3 col = muls(divs(sum, (float)spp), 255.0f);
			int r = (int)col.x < 255 ? (int)col.x : 255;
			int g = (int)col.y < 255 ? (int)col.y : 255;
			int b = (int)col.z < 255 ? (int)col.z : 255;
			unsigned char *px = rgba + 4 * ((size_t)y * (size_t)W + (size_t)x);
			px[0] = (unsigned char)(r & 0xFF);
			px[1] = (unsigned char)(g & 0xFF);
			px[2] = (unsigned char)(b & 0xFF);
			px[3] = 255;
		}
	}
	or_compat_bvh_free((void *)c.bvh);
}

/* ------------------------------------- reference estimator, mt19937 -- */

static float mt_jitter(mt19937 *g) { /* uniform_real_distribution<float>(-0.5, 0.5) */
	return mt_unif(g) * 1.0f + -0.5f;
}

static v3 cosine_hemisphere_mt(v3 n, mt19937 *g) {
	float u1 = mt_unif(g);
	float u2 = mt_unif(g);
	return cos_hemi(n, u1, u2);
}

/* pathTrace, src/Renderer.cpp:820-907, as written: recursion, colours
 * combined as direct + albedo * (what the bounce returns). */
static v3 path_trace_mt(const cctx *c, v3 o, v3 d, int depth, mt19937 *g) {
	const or_compat_scene *s = c->s;
	float oa[3] = {o.x, o.y, o.z}, da[3] = {d.x, d.y, d.z};
	float t, hu, hv;
	int ti = or_compat_bvh_intersect(c->bvh, oa, da, -1, &t, &hu, &hv);
	if (ti < 0)
		return divs(environment(d), 255.0f);
	v3 point = add(o, muls(d, t));
	int mat = s->material[ti];
	v3 N = ld(s->normals + 3 * ti);
	v3 col = mk((float)s->colour[3 * ti], (float)s->colour[3 * ti + 1], (float)s->colour[3 * ti + 2]);
	if (mat == M_EMISSIVE)
		return muls(divs(col, 255.0f), 4.0f);
	if (mat == M_MIRROR && depth > 0) {
		v3 rr = reflect(d, face_viewer(N, d));
		return path_trace_mt(c, add(point, muls(rr, 1e-4f)), rr, depth - 1, g);
	}
	if ((mat == M_GLASS || mat == M_DISPERSIVE) && depth > 0) {
		float ior = 1.5f;
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
		float fres = r0 + (1.0f - r0) * powf(1.0f - cosi, 5.0f);
		v3 refr = refract(d, n, etai / etat);
		v3 nd = (length(refr) < 1e-6f || mt_unif(g) < fres) ? reflect(d, n) : refr;
		return path_trace_mt(c, add(point, muls(nd, 1e-4f)), nd, depth - 1, g);
	}
	if (mat == M_METAL && depth > 0) {
		v3 n = face_viewer(N, d);
		v3 refl = reflect(d, n);
		v3 diff = cosine_hemisphere_mt(n, g);
		float rough = 0.2f; /* ModelTriangle::roughness default, framework/ModelTriangle.h:31 */
		v3 nd = normalize(mixv(refl, diff, rough * rough));
		if (dot(nd, n) <= 0.0f)
			nd = refl;
		v3 albedo = divs(base_colour(s, ti, hu, hv, point, muls(d, -1.0f)), 255.0f);
		return mul(albedo, path_trace_mt(c, add(point, muls(nd, 1e-4f)), nd, depth - 1, g));
	}
	if (mat == M_SUBSURFACE) {
		v3 n = face_viewer(N, d);
		v3 albedo = divs(base_colour(s, ti, hu, hv, point, muls(d, -1.0f)), 255.0f);
		v3 to = sub(ld(kLightPos), point);
		float dist = length(to);
		v3 ldir = divs(to, dist);
		float att = kLightIntensity / (0.0f + 0.0f * dist + kLightKq * dist * dist);
		float wd = fmaxf(0.0f, (dot(n, ldir) + 0.5f) / (1.0f + 0.5f));
		v3 shade = add(mk(0, 0, 0), muls(mul(albedo, mk(1, 1, 1)), visibility(c, point, ti) * att * wd));
		v3 eo = sub(point, muls(n, 1e-3f));
		float eoa[3] = {eo.x, eo.y, eo.z}, lda[3] = {ldir.x, ldir.y, ldir.z};
		float et, eu, ev;
		int eh = or_compat_bvh_intersect(c->bvh, eoa, lda, ti, &et, &eu, &ev);
		float thick = eh >= 0 ? et : 0.0f;
		shade = add(shade, muls(mul(albedo, mk(1, 1, 1)), att * expf(-3.0f * thick) * 0.6f));
		v3 direct = vmin(shade, mk(1.5f, 1.5f, 1.5f));
		if (depth <= 0)
			return direct;
		v3 sb = cosine_hemisphere_mt(n, g);
		return add(direct, mul(albedo, path_trace_mt(c, add(point, muls(sb, 1e-4f)), sb, depth - 1, g)));
	}
	v3 n = face_viewer(N, d);
	v3 albedo = divs(base_colour(s, ti, hu, hv, point, muls(d, -1.0f)), 255.0f);
	v3 direct = direct_light(c, point, n, albedo, ti);
	if (depth <= 0)
		return direct;
	v3 bounce = cosine_hemisphere_mt(n, g);
	v3 indirect = mul(albedo, path_trace_mt(c, add(point, muls(bounce, 1e-4f)), bounce, depth - 1, g));
	return add(direct, indirect);
}

/* renderPathTraced (src/Renderer.cpp:1119-1170) with its own random
 * streams: std::mt19937 seeded (y * W + x) * 9781 + 1 per pixel; per
 * sample two jitters, one draw for the (zero) camera motion, then the
 * path's draws. */
void or_compat_render_mt(const or_compat_scene *s, int W, int H, float cam_z, int spp, int y0, int y1,
                         unsigned char *rgba) {
	cctx c;
	c.s = s;
	c.bvh = or_compat_bvh_build(s->verts, s->n);
#pragma omp parallel for schedule(dynamic, 4)
	for (int y = y0; y < y1; y++) {
		mt19937 *g = (mt19937 *)malloc(sizeof(mt19937));
		for (int x = 0; x < W; x++) {
			mt_seed(g, (uint32_t)(y * W + x) * 9781u + 1u);
			v3 sum = mk(0, 0, 0);
			for (int si = 0; si < spp; si++) {
				float jx = mt_jitter(g);
				float jy = mt_jitter(g);
				float o[3], d[3];
				or_compat_primary_ray(x, y, W, H, cam_z, jx, jy, o, d);
				(void)mt_unif(g); /* camera motion (zero here) */
				sum = add(sum, path_trace_mt(&c, ld(o), ld(d), 4, g));
			}
			v3 col = muls(divs(sum, (float)spp), 255.0f);
			int r = (int)col.x < 255 ? (int)col.x : 255;
			int gg = (int)col.y < 255 ? (int)col.y : 255;
			int b = (int)col.z < 255 ? (int)col.z : 255;
			unsigned char *px = rgba + 4 * ((size_t)y * (size_t)W + (size_t)x);
			px[0] = (unsigned char)(r & 0xFF);
			px[1] = (unsigned char)(gg & 0xFF);
			px[2] = (unsigned char)(b & 0xFF);
			px[3] = 255;
		}
		free(g);
	}
	or_compat_bvh_free((void *)c.bvh);
}
