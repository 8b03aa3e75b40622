/* Oracle: the physical path-tracing estimator ("Mode P") and its adjoint,
 * written once for a real type R and included twice (float, double) by
 * phys.c. The reference's renderPathTraced (src/Renderer.cpp:1119-1170)
 * cannot express physical quad lights, albedo/pi or Russian roulette; this
 * estimator follows the structure of the reference's own physically based NEE
 * tracer (src/Metropolis.cpp:53-81: NEE with G = cosS cosL / d^2, albedo/pi,
 * cosine bounce, beta *= albedo) generalised to area-sampled emissive
 * triangles. Sampling layout (shared with the GPU kernels):
 *
 *   key = (seed, scene_id)
 *   camera jitter:  ctr = (pixel, sample, 0, 0)      words 0,1
 *   NEE + RR at vertex b: ctr = (pixel, sample, b, 1) words 0 (light), 1,2
 *                                                     (point), 3 (roulette)
 *   bounce at vertex b:   ctr = (pixel, sample, b, 2) words 0,1
 *   u(word) = (word >> 8) * 2^-24
 *
 * TEST INFRASTRUCTURE ONLY -- see oracle.h. */

#ifndef R
#error "define R and SFX before including phys_impl.h"
#endif

typedef struct {
	R x, y, z;
} SFX(V);

#define VT SFX(V)

static inline VT SFX(vmk)(R x, R y, R z) {
	VT r = {x, y, z};
	return r;
}
static inline VT SFX(vadd)(VT a, VT b) {
	return SFX(vmk)(a.x + b.x, a.y + b.y, a.z + b.z);
}
static inline VT SFX(vsub)(VT a, VT b) {
	return SFX(vmk)(a.x - b.x, a.y - b.y, a.z - b.z);
}
static inline VT SFX(vmul)(VT a, VT b) {
	return SFX(vmk)(a.x * b.x, a.y * b.y, a.z * b.z);
}
static inline VT SFX(vs)(VT a, R s) {
	return SFX(vmk)(a.x * s, a.y * s, a.z * s);
}
static inline R SFX(vdot)(VT a, VT b) {
	return a.x * b.x + a.y * b.y + a.z * b.z;
}
static inline VT SFX(vcross)(VT a, VT b) {
	return SFX(vmk)(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline VT SFX(vnorm)(VT a) {
	R l = (R)1 / (R)sqrt((double)SFX(vdot)(a, a));
	return SFX(vs)(a, l);
}
static inline R SFX(vget)(VT a, int k) {
	return k == 0 ? a.x : (k == 1 ? a.y : a.z);
}
static inline R SFX(maxc)(VT a) {
	R m = a.x > a.y ? a.x : a.y;
	return m > a.z ? m : a.z;
}

typedef struct {
	R lo[3], hi[3];
	int left, right, start, count;
} SFX(Node);

typedef struct {
	int n;
	VT *v0, *v1, *v2;
	VT *albedo, *emission;
	int nl;
	int *light_tri;
	R *light_cdf;
	R total_area;
	SFX(Node) * nodes;
	int nnodes, cap, root;
	int *order;
	VT *cent;
} SFX(Scene);

#define ST SFX(Scene)

static int SFX(push)(ST *s, SFX(Node) nd) {
	if (s->nnodes == s->cap) {
		s->cap = s->cap ? 2 * s->cap : 1024;
		s->nodes = (SFX(Node) *)realloc(s->nodes, sizeof(SFX(Node)) * (size_t)s->cap);
	}
	s->nodes[s->nnodes] = nd;
	return s->nnodes++;
}

static void SFX(select)(ST *s, int lo, int hi, int k, int axis) {
	int *a = s->order;
	while (hi - lo > 1) {
		int mid = lo + (hi - lo) / 2;
		R pa = SFX(vget)(s->cent[a[lo]], axis), pb = SFX(vget)(s->cent[a[mid]], axis),
		  pc = SFX(vget)(s->cent[a[hi - 1]], axis);
		R pivot = (pa < pb) ? ((pb < pc) ? pb : ((pa < pc) ? pc : pa)) : ((pa < pc) ? pa : ((pb < pc) ? pc : pb));
		int i = lo, lt = lo, gt = hi;
		while (i < gt) {
			R c = SFX(vget)(s->cent[a[i]], axis);
			if (c < pivot) {
				int t = a[lt];
				a[lt] = a[i];
				a[i] = t;
				lt++;
				i++;
			} else if (pivot < c) {
				gt--;
				int t = a[gt];
				a[gt] = a[i];
				a[i] = t;
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

/* Median-split binary BVH, leaves <= 4 triangles (the reference's build rule,
 * src/BVH.cpp:18-63), in precision R. */
static int SFX(build)(ST *s, int start, int count) {
	SFX(Node) nd;
	for (int k = 0; k < 3; k++) {
		nd.lo[k] = (R)INFINITY;
		nd.hi[k] = (R)-INFINITY;
	}
	R clo[3] = {(R)INFINITY, (R)INFINITY, (R)INFINITY}, chi[3] = {(R)-INFINITY, (R)-INFINITY, (R)-INFINITY};
	for (int i = start; i < start + count; i++) {
		int t = s->order[i];
		VT vs[3] = {s->v0[t], s->v1[t], s->v2[t]};
		for (int j = 0; j < 3; j++)
			for (int k = 0; k < 3; k++) {
				R x = SFX(vget)(vs[j], k);
				if (x < nd.lo[k])
					nd.lo[k] = x;
				if (x > nd.hi[k])
					nd.hi[k] = x;
			}
		for (int k = 0; k < 3; k++) {
			R x = SFX(vget)(s->cent[t], k);
			if (x < clo[k])
				clo[k] = x;
			if (x > chi[k])
				chi[k] = x;
		}
	}
	nd.left = nd.right = -1;
	nd.start = start;
	nd.count = count;
	if (count <= 4)
		return SFX(push)(s, nd);
	R ex = chi[0] - clo[0], ey = chi[1] - clo[1], ez = chi[2] - clo[2];
	int axis = (ex > ey && ex > ez) ? 0 : (ey > ez ? 1 : 2);
	int mid = start + count / 2;
	SFX(select)(s, start, start + count, mid, axis);
	nd.count = 0;
	int me = SFX(push)(s, nd);
	int l = SFX(build)(s, start, mid - start);
	int r = SFX(build)(s, mid, start + count - mid);
	s->nodes[me].left = l;
	s->nodes[me].right = r;
	return me;
}

void *SFX(or_phys_prepare)(const float *verts, int n, const float *albedo, const float *emission) {
	ST *s = (ST *)calloc(1, sizeof(ST));
	s->n = n;
	size_t N = (size_t)(n > 0 ? n : 1);
	s->v0 = (VT *)malloc(sizeof(VT) * N);
	s->v1 = (VT *)malloc(sizeof(VT) * N);
	s->v2 = (VT *)malloc(sizeof(VT) * N);
	s->albedo = (VT *)malloc(sizeof(VT) * N);
	s->emission = (VT *)malloc(sizeof(VT) * N);
	s->order = (int *)malloc(sizeof(int) * N);
	s->cent = (VT *)malloc(sizeof(VT) * N);
	s->light_tri = (int *)malloc(sizeof(int) * N);
	s->light_cdf = (R *)malloc(sizeof(R) * N);
	double *acc = (double *)malloc(sizeof(double) * N);
	double total = 0.0;
	for (int i = 0; i < n; i++) {
		const float *t = verts + 9 * i;
		s->v0[i] = SFX(vmk)(t[0], t[1], t[2]);
		s->v1[i] = SFX(vmk)(t[3], t[4], t[5]);
		s->v2[i] = SFX(vmk)(t[6], t[7], t[8]);
		s->albedo[i] = SFX(vmk)(albedo[3 * i], albedo[3 * i + 1], albedo[3 * i + 2]);
		if (emission)
			s->emission[i] = SFX(vmk)(emission[3 * i], emission[3 * i + 1], emission[3 * i + 2]);
		else
			s->emission[i] = SFX(vmk)(0, 0, 0);
		s->order[i] = i;
		s->cent[i] = SFX(vs)(SFX(vadd)(SFX(vadd)(s->v0[i], s->v1[i]), s->v2[i]), (R)(1.0 / 3.0));
		if (s->emission[i].x > 0 || s->emission[i].y > 0 || s->emission[i].z > 0) {
			VT c = SFX(vcross)(SFX(vsub)(s->v1[i], s->v0[i]), SFX(vsub)(s->v2[i], s->v0[i]));
			double area = 0.5 * sqrt((double)SFX(vdot)(c, c));
			total += area;
			acc[s->nl] = total;
			s->light_tri[s->nl++] = i;
		}
	}
	for (int i = 0; i < s->nl; i++)
		s->light_cdf[i] = (R)(acc[i] / total);
	if (s->nl > 0)
		s->light_cdf[s->nl - 1] = (R)1;
	s->total_area = (R)total;
	free(acc);
	s->root = n > 0 ? SFX(build)(s, 0, n) : -1;
	return s;
}

void SFX(or_phys_free)(void *p) {
	ST *s = (ST *)p;
	if (!s)
		return;
	free(s->v0);
	free(s->v1);
	free(s->v2);
	free(s->albedo);
	free(s->emission);
	free(s->order);
	free(s->cent);
	free(s->light_tri);
	free(s->light_cdf);
	free(s->nodes);
	free(s);
}

/* Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2(1) 2013): shear
 * into ray space with the dominant axis as z, 2-D edge functions, double
 * fallback when an edge function is exactly zero. */
typedef struct {
	int kx, ky, kz;
	R sx, sy, sz;
	VT o;
} SFX(WRay);

static SFX(WRay) SFX(wray)(VT o, VT d) {
	SFX(WRay) w;
	R ax = d.x < 0 ? -d.x : d.x, ay = d.y < 0 ? -d.y : d.y, az = d.z < 0 ? -d.z : d.z;
	int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
	int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
	if (SFX(vget)(d, kz) < 0) {
		int t = kx;
		kx = ky;
		ky = t;
	}
	w.kx = kx;
	w.ky = ky;
	w.kz = kz;
	w.sx = SFX(vget)(d, kx) / SFX(vget)(d, kz);
	w.sy = SFX(vget)(d, ky) / SFX(vget)(d, kz);
	w.sz = (R)1 / SFX(vget)(d, kz);
	w.o = o;
	return w;
}

static int SFX(wtest)(const SFX(WRay) * w, VT v0, VT v1, VT v2, R tmax, R *tt, R *uu, R *vv) {
	VT A = SFX(vsub)(v0, w->o), B = SFX(vsub)(v1, w->o), C = SFX(vsub)(v2, w->o);
	R Az0 = SFX(vget)(A, w->kz), Bz0 = SFX(vget)(B, w->kz), Cz0 = SFX(vget)(C, w->kz);
	R Ax = SFX(vget)(A, w->kx) - w->sx * Az0, Ay = SFX(vget)(A, w->ky) - w->sy * Az0;
	R Bx = SFX(vget)(B, w->kx) - w->sx * Bz0, By = SFX(vget)(B, w->ky) - w->sy * Bz0;
	R Cx = SFX(vget)(C, w->kx) - w->sx * Cz0, Cy = SFX(vget)(C, w->ky) - w->sy * Cz0;
	R U = Cx * By - Cy * Bx, Vv = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
	if (U == (R)0 || Vv == (R)0 || W == (R)0) {
		double cxby = (double)Cx * (double)By, cybx = (double)Cy * (double)Bx;
		double axcy = (double)Ax * (double)Cy, aycx = (double)Ay * (double)Cx;
		double bxay = (double)Bx * (double)Ay, byax = (double)By * (double)Ax;
		U = (R)(cxby - cybx);
		Vv = (R)(axcy - aycx);
		W = (R)(bxay - byax);
		/* An edge function still exactly zero takes the sign it would have
		 * with the ray's point moved by (e1, e2), 0 < e2 << e1: the sign of
		 * its x derivative, else of its y derivative, kept as the sign of the
		 * zero. The two triangles of a shared edge see opposite derivatives,
		 * so exactly one keeps the point; around a vertex exactly one
		 * triangle of the fan keeps it. */
		if (U == (R)0)
			U = (R)copysign(0.0, (double)(Cy != By ? Cy - By : Bx - Cx));
		if (Vv == (R)0)
			Vv = (R)copysign(0.0, (double)(Ay != Cy ? Ay - Cy : Cx - Ax));
		if (W == (R)0)
			W = (R)copysign(0.0, (double)(By != Ay ? By - Ay : Ax - Bx));
	}
	/* a hit needs U, V and W of one sign */
	if (!signbit(U) != !signbit(Vv) || !signbit(U) != !signbit(W))
		return 0;
	R det = U + Vv + W;
	if (det == (R)0)
		return 0;
	R Az = w->sz * Az0, Bz = w->sz * Bz0, Cz = w->sz * Cz0;
	R T = U * Az + Vv * Bz + W * Cz;
	if (det < 0) {
		T = -T;
		det = -det;
		U = -U;
		Vv = -Vv;
		W = -W;
	}
	if (!(T > (R)0) || !(T < tmax * det))
		return 0;
	R rcp = (R)1 / det;
	*tt = T * rcp;
	*uu = Vv * rcp;
	*vv = W * rcp;
	return 1;
}

static int SFX(slab)(const SFX(Node) * nd, VT o, VT inv, R tmax) {
	R te = 0, tx = tmax;
	for (int k = 0; k < 3; k++) {
		R oi = SFX(vget)(o, k), ii = SFX(vget)(inv, k);
		if (ii == (R)INFINITY || ii == (R)-INFINITY) {
			/* the ray is parallel to this slab: inside it (boundary planes
			 * included) at every t, or never */
			if (oi < nd->lo[k] || oi > nd->hi[k])
				return 0;
			continue;
		}
		R a = (nd->lo[k] - oi) * ii, b = (nd->hi[k] - oi) * ii;
		R lo = a < b ? a : b, hi = a < b ? b : a;
		te = lo > te ? lo : te;
		tx = hi < tx ? hi : tx;
	}
	return te <= tx * (R)(1 + 4e-7);
}

/* Closest hit; returns prim or -1. */
static int SFX(closest)(const ST *s, VT o, VT d, R tmax, R *t, R *u, R *v) {
	int best = -1;
	R bt = tmax;
	if (s->root < 0)
		return -1;
	SFX(WRay) w = SFX(wray)(o, d);
	VT inv = SFX(vmk)((R)1 / d.x, (R)1 / d.y, (R)1 / d.z);
	int stack[128], sp = 0;
	stack[sp++] = s->root;
	while (sp > 0) {
		const SFX(Node) *nd = &s->nodes[stack[--sp]];
		if (!SFX(slab)(nd, o, inv, bt))
			continue;
		if (nd->count > 0) {
			for (int i = nd->start; i < nd->start + nd->count; i++) {
				int ti = s->order[i];
				R tt, uu, vv;
				if (SFX(wtest)(&w, s->v0[ti], s->v1[ti], s->v2[ti], bt, &tt, &uu, &vv)) {
					if (tt < bt || (tt == bt && ti < best)) {
						bt = tt;
						best = ti;
						*u = uu;
						*v = vv;
					}
				}
			}
		} else {
			stack[sp++] = nd->left;
			stack[sp++] = nd->right;
		}
	}
	*t = bt;
	return best;
}

static int SFX(any)(const ST *s, VT o, VT d, R tmax) {
	if (s->root < 0)
		return 0;
	SFX(WRay) w = SFX(wray)(o, d);
	VT inv = SFX(vmk)((R)1 / d.x, (R)1 / d.y, (R)1 / d.z);
	int stack[128], sp = 0;
	stack[sp++] = s->root;
	while (sp > 0) {
		const SFX(Node) *nd = &s->nodes[stack[--sp]];
		if (!SFX(slab)(nd, o, inv, tmax))
			continue;
		if (nd->count > 0) {
			for (int i = nd->start; i < nd->start + nd->count; i++) {
				int ti = s->order[i];
				R tt, uu, vv;
				if (SFX(wtest)(&w, s->v0[ti], s->v1[ti], s->v2[ti], tmax, &tt, &uu, &vv))
					return 1;
			}
		} else {
			stack[sp++] = nd->left;
			stack[sp++] = nd->right;
		}
	}
	return 0;
}

/* Every triangle the watertight test accepts on (o, d) with 0 < t < tmax, by
 * brute force in triangle order: ids into prims, distances into ts (at most
 * cap of each); returns the number of hits. */
int SFX(or_phys_all_hits)(const void *p, const R o[3], const R d[3], R tmax, int *prims, R *ts, int cap) {
	const ST *s = (const ST *)p;
	SFX(WRay) w = SFX(wray)(SFX(vmk)(o[0], o[1], o[2]), SFX(vmk)(d[0], d[1], d[2]));
	int n = 0;
	for (int i = 0; i < s->n; i++) {
		R tt, uu, vv;
		if (SFX(wtest)(&w, s->v0[i], s->v1[i], s->v2[i], tmax, &tt, &uu, &vv)) {
			if (n < cap) {
				prims[n] = i;
				ts[n] = tt;
			}
			n++;
		}
	}
	return n;
}

int SFX(or_phys_closest)(const void *p, const R o[3], const R d[3], R tmax, int brute, R *t, R *u, R *v) {
	const ST *s = (const ST *)p;
	VT O = SFX(vmk)(o[0], o[1], o[2]), D = SFX(vmk)(d[0], d[1], d[2]);
	if (!brute)
		return SFX(closest)(s, O, D, tmax, t, u, v);
	SFX(WRay) w = SFX(wray)(O, D);
	int best = -1;
	R bt = tmax;
	for (int i = 0; i < s->n; i++) {
		R tt, uu, vv;
		if (SFX(wtest)(&w, s->v0[i], s->v1[i], s->v2[i], bt, &tt, &uu, &vv) && tt < bt) {
			bt = tt;
			best = i;
			*u = uu;
			*v = vv;
		}
	}
	*t = bt;
	return best;
}

static inline R SFX(unif)(uint32_t x) {
	return (R)((float)(x >> 8) * (1.0f / 16777216.0f));
}

/* Uniform point on light `li` and its unit normal. */
static VT SFX(light_point)(const ST *s, int tri, R r1, R r2, VT *nl) {
	VT a = s->v0[tri], e1 = SFX(vsub)(s->v1[tri], a), e2 = SFX(vsub)(s->v2[tri], a);
	R su = (R)sqrt((double)r1);
	R b1 = su * ((R)1 - r2), b2 = su * r2;
	*nl = SFX(vnorm)(SFX(vcross)(e1, e2));
	return SFX(vadd)(SFX(vadd)(a, SFX(vs)(e1, b1)), SFX(vs)(e2, b2));
}

static int SFX(pick_light)(const ST *s, R u) {
	int lo = 0, hi = s->nl - 1;
	while (lo < hi) {
		int mid = (lo + hi) / 2;
		if (u < s->light_cdf[mid])
			hi = mid;
		else
			lo = mid + 1;
	}
	return lo;
}

/* Cosine-weighted direction about n with the branchless orthonormal basis of
 * Duff et al., "Building an Orthonormal Basis, Revisited", JCGT 6(1) 2017. */
static VT SFX(cosine_dir)(VT n, R u1, R u2) {
	R r = (R)sqrt((double)u1);
	R phi = (R)(2.0 * 3.14159265358979323846) * u2;
	R x = r * (R)cos((double)phi), y = r * (R)sin((double)phi);
	R z = (R)sqrt((double)((R)1 - u1 > 0 ? (R)1 - u1 : 0));
	R sign = n.z >= 0 ? (R)1 : (R)-1;
	R a = (R)-1 / (sign + n.z);
	R b = n.x * n.y * a;
	VT t = SFX(vmk)((R)1 + sign * n.x * n.x * a, sign * b, -sign * n.x);
	VT bb = SFX(vmk)(b, sign + n.y * n.y * a, -n.y);
	return SFX(vadd)(SFX(vadd)(SFX(vs)(t, x), SFX(vs)(bb, y)), SFX(vs)(n, z));
}

static VT SFX(camera_dir)(const or_cam *c, int x, int y, int W, int H, R jx, R jy) {
	R fx = ((R)x + jx) / (R)W * (R)2 - (R)1;
	R fy = (R)1 - ((R)y + jy) / (R)H * (R)2;
	VT du = SFX(vmk)(c->du[0], c->du[1], c->du[2]), dv = SFX(vmk)(c->dv[0], c->dv[1], c->dv[2]);
	VT w = SFX(vmk)(c->w[0], c->w[1], c->w[2]);
	return SFX(vnorm)(SFX(vadd)(SFX(vadd)(SFX(vs)(du, fx), SFX(vs)(dv, fy)), w));
}

/* Per-vertex record kept for the reverse sweep of the adjoint. */
typedef struct {
	int tri;
	VT a, X, beta;
	R invq;
} SFX(Vtx);

/* One camera path. When `rec` is non-NULL the vertex records are stored and
 * the NEE emission derivatives are added to g_emission (scaled by w). */
static VT SFX(path)(const ST *s, const or_cam *cam, const or_params *p, int x, int y, uint32_t smp, or_counts *cnt,
                    SFX(Vtx) * rec, int *nrec, VT wgt, double *g_emission) {
	uint32_t pix = (uint32_t)(y * p->width + x);
	uint32_t key[2] = {p->seed, p->scene_id};
	uint32_t ctr[4] = {pix, smp, 0, 0}, r[4];
	or_philox4x32_10(ctr, key, r);
	VT o = SFX(vmk)(cam->eye[0], cam->eye[1], cam->eye[2]);
	VT d = SFX(camera_dir)(cam, x, y, p->width, p->height, SFX(unif)(r[0]), SFX(unif)(r[1]));
	VT L = SFX(vmk)(0, 0, 0), beta = SFX(vmk)(1, 1, 1);
	R eps = (R)p->ray_eps;
	if (nrec)
		*nrec = 0;
	for (int depth = 0; depth < p->max_depth; depth++) {
		R t, u, v;
		cnt->closest++;
		int ti = SFX(closest)(s, o, d, (R)INFINITY, &t, &u, &v);
		if (ti < 0)
			break;
		VT P = SFX(vadd)(o, SFX(vs)(d, t));
		VT ng = SFX(vnorm)(SFX(vcross)(SFX(vsub)(s->v1[ti], s->v0[ti]), SFX(vsub)(s->v2[ti], s->v0[ti])));
		if (SFX(vdot)(ng, d) > 0)
			ng = SFX(vs)(ng, (R)-1);
		if (depth == 0) {
			L = SFX(vadd)(L, SFX(vmul)(beta, s->emission[ti]));
			if (g_emission) {
				g_emission[3 * ti] += (double)wgt.x;
				g_emission[3 * ti + 1] += (double)wgt.y;
				g_emission[3 * ti + 2] += (double)wgt.z;
			}
		}
		VT a = s->albedo[ti];
		VT so = SFX(vadd)(P, SFX(vs)(ng, eps));
		ctr[2] = (uint32_t)depth;
		ctr[3] = 1;
		or_philox4x32_10(ctr, key, r);
		VT X = SFX(vmk)(0, 0, 0);
		if (s->nl > 0) {
			int li = SFX(pick_light)(s, SFX(unif)(r[0]));
			int lt = s->light_tri[li];
			VT nl;
			VT q = SFX(light_point)(s, lt, SFX(unif)(r[1]), SFX(unif)(r[2]), &nl);
			VT wi = SFX(vsub)(q, so);
			R d2 = SFX(vdot)(wi, wi);
			R dist = (R)sqrt((double)d2);
			wi = SFX(vs)(wi, (R)1 / dist);
			R cs = SFX(vdot)(ng, wi);
			R cl = SFX(vdot)(nl, wi);
			cl = cl < 0 ? -cl : cl;
			if (cs > 0 && cl > 0) {
				cnt->shadow++;
				if (!SFX(any)(s, so, wi, dist * (R)0.9999)) {
					R Y = cs * cl / d2 * s->total_area * (R)(1.0 / 3.14159265358979323846);
					X = SFX(vs)(s->emission[lt], Y);
					L = SFX(vadd)(L, SFX(vmul)(SFX(vmul)(beta, a), X));
					if (g_emission) {
						VT g = SFX(vs)(SFX(vmul)(SFX(vmul)(wgt, beta), a), Y);
						g_emission[3 * lt] += (double)g.x;
						g_emission[3 * lt + 1] += (double)g.y;
						g_emission[3 * lt + 2] += (double)g.z;
					}
				}
			}
		}
		if (rec) {
			rec[*nrec].tri = ti;
			rec[*nrec].a = a;
			rec[*nrec].X = X;
			rec[*nrec].beta = beta;
			rec[*nrec].invq = (R)1;
			(*nrec)++;
		}
		if (depth + 1 >= p->max_depth)
			break;
		uint32_t c2[4] = {pix, smp, (uint32_t)depth, 2}, b[4];
		or_philox4x32_10(c2, key, b);
		VT wo = SFX(cosine_dir)(ng, SFX(unif)(b[0]), SFX(unif)(b[1]));
		beta = SFX(vmul)(beta, a);
		if (depth >= p->rr_depth) {
			R q = p->rr_prob > 0 ? (R)p->rr_prob : (SFX(maxc)(beta) < (R)0.95 ? SFX(maxc)(beta) : (R)0.95);
			if (!(SFX(unif)(r[3]) < q))
				break;
			beta = SFX(vs)(beta, (R)1 / q);
			if (rec)
				rec[*nrec - 1].invq = (R)1 / q;
		}
		o = so;
		d = wo;
	}
	return L;
}

static void SFX(render_pixel)(const ST *s, const or_cam *cam, const or_params *p, int pixel, R *out, or_counts *cnt) {
	int x = pixel % p->width, y = pixel / p->width;
	VT sum = SFX(vmk)(0, 0, 0);
	for (int k = 0; k < p->spp; k++) {
		cnt->paths++;
		sum = SFX(vadd)(sum, SFX(path)(s, cam, p, x, y, (uint32_t)k, cnt, NULL, NULL, sum, NULL));
	}
	R inv = (R)1 / (R)p->spp;
	out[0] = sum.x * inv;
	out[1] = sum.y * inv;
	out[2] = sum.z * inv;
}

void SFX(or_phys_render)(const void *sp, const or_cam *cam, const or_params *p, int p0, int p1, int nthreads, R *out,
                         or_counts *counts) {
	const ST *s = (const ST *)sp;
	uint64_t c0 = 0, c1 = 0, c2 = 0;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads > 0 ? nthreads : 1) reduction(+ : c0, c1, c2)
	for (int i = p0; i < p1; i++) {
		or_counts c = {0, 0, 0};
		SFX(render_pixel)(s, cam, p, i, out + 3 * (size_t)i, &c);
		c0 += c.closest;
		c1 += c.shadow;
		c2 += c.paths;
	}
	if (counts) {
		counts->closest = c0;
		counts->shadow = c1;
		counts->paths = c2;
	}
}

void SFX(or_phys_render_list)(const void *sp, const or_cam *cam, const or_params *p, int npix, const int *pix,
                              int nthreads, R *out, or_counts *counts) {
	const ST *s = (const ST *)sp;
	uint64_t c0 = 0, c1 = 0, c2 = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1) reduction(+ : c0, c1, c2)
	for (int i = 0; i < npix; i++) {
		or_counts c = {0, 0, 0};
		SFX(render_pixel)(s, cam, p, pix[i], out + 3 * (size_t)i, &c);
		c0 += c.closest;
		c1 += c.shadow;
		c2 += c.paths;
	}
	if (counts) {
		counts->closest = c0;
		counts->shadow = c1;
		counts->paths = c2;
	}
}
