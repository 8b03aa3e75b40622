// Device-side building blocks shared by the kernels: vector helpers, Philox,
// the BVH8 node layout, the two ray/triangle tests and wide-BVH traversal.
// Everything here is gfx950 (CDNA4, wave64) code; there is no other target.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rn_types.h"

#define RN_DEV __device__ __forceinline__

// ------------------------------------------------------------------ math --

RN_DEV float3 f3(float x, float y, float z) { return make_float3(x, y, z); }
RN_DEV float3 operator+(float3 a, float3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
RN_DEV float3 operator-(float3 a, float3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
RN_DEV float3 operator*(float3 a, float3 b) { return f3(a.x * b.x, a.y * b.y, a.z * b.z); }
RN_DEV float3 operator*(float3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
RN_DEV float3 operator*(float s, float3 a) { return f3(a.x * s, a.y * s, a.z * s); }
RN_DEV float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RN_DEV float3 cross3(float3 a, float3 b) {
	return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
RN_DEV float3 normalize3(float3 a) { return a * (1.0f / sqrtf(dot3(a, a))); }
RN_DEV float maxc3(float3 a) { return fmaxf(fmaxf(a.x, a.y), a.z); }
RN_DEV float get3(float3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
RN_DEV float3 xyz(float4 a) { return f3(a.x, a.y, a.z); }

// ---------------------------------------------------------------- Philox --
// Philox4x32-10 (Salmon et al., SC'11): ten rounds of two 32x32->64-bit
// multiplies. Each product is one v_mad_u64_u32 (high and low halves
// together; v_mul_hi_u32 + v_mul_lo_u32 would be two quarter-rate
// instructions) and each three-way xor one v_bitop3_b32 (truth table 0x96).

struct u32x4 {
	uint32_t x, y, z, w;
};

RN_DEV u32x4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
#pragma unroll
	for (int r = 0; r < 10; r++) {
		if (r > 0) {
			k0 += 0x9E3779B9u;
			k1 += 0xBB67AE85u;
		}
		const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
		const uint32_t n0 = __builtin_amdgcn_bitop3_b32((uint32_t)(p1 >> 32), c1, k0, 0x96);
		const uint32_t n2 = __builtin_amdgcn_bitop3_b32((uint32_t)(p0 >> 32), c3, k1, 0x96);
		c0 = n0;
		c1 = (uint32_t)p1;
		c2 = n2;
		c3 = (uint32_t)p0;
	}
	return u32x4{c0, c1, c2, c3};
}

// 24-bit uniform in [0,1): exactly representable, identical on host and device.
RN_DEV float u01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

// ------------------------------------------------------------------ BVH8 --

RN_DEV void load_node(const Bvh8Node *__restrict__ nodes, uint32_t i, uint32_t w[20]) {
	const uint4 *p = reinterpret_cast<const uint4 *>(nodes + i);
#pragma unroll
	for (int k = 0; k < 5; k++) {
		uint4 q = p[k];
		w[4 * k] = q.x;
		w[4 * k + 1] = q.y;
		w[4 * k + 2] = q.z;
		w[4 * k + 3] = q.w;
	}
}

// Ray state shared by the traversal loop.
struct TRay {
	float3 o, d;
	float3 inv;  // 1/d with zero components replaced by +-1e-30
	uint32_t oct; // bit k set when d[k] < 0
};

RN_DEV TRay make_tray(float3 o, float3 d) {
	TRay r;
	r.o = o;
	r.d = d;
	// |1/d| stays below 2^40, so node slab terms scaled by 2^24 (below) stay
	// finite for any node scale up to 2^60. A component of magnitude at most
	// 2^-40 (zero included, of either sign) becomes +2^-40: for the slabs the
	// ray then drifts to the + side of every axis it does not move along, so
	// a box whose upper face lies in the ray's plane is culled and one whose
	// lower face does is kept. The watertight test's tie rule resolves a
	// crossing on such a plane in favour of the triangle on the + side of
	// both sheared axes, so the box of that triangle is always one the slab
	// test keeps: exact crossings on box faces are found. (An infinite
	// reciprocal would keep every box of that axis: NaN slab terms.)
	const float tiny = 0x1p-40f;
	float dx = fabsf(d.x) > tiny ? d.x : tiny;
	float dy = fabsf(d.y) > tiny ? d.y : tiny;
	float dz = fabsf(d.z) > tiny ? d.z : tiny;
	r.inv = f3(1.0f / dx, 1.0f / dy, 1.0f / dz);
	r.oct = (dx < 0.0f ? 1u : 0u) | (dy < 0.0f ? 2u : 0u) | (dz < 0.0f ? 4u : 0u);
	return r;
}

typedef _Float16 h16x2 __attribute__((ext_vector_type(2)));

// Two bytes of w (chosen by the v_perm selector) as two f16 denormals
// byte * 2^-24, exactly.
RN_DEV h16x2 h16_pair(uint32_t w, uint32_t sel) {
	return __builtin_bit_cast(h16x2, __builtin_amdgcn_perm(0u, w, sel));
}

// Bit b of the result is bit (b ^ oct) of x (x has 8 significant bits).
// fill_oct_table writes the same map as a 2 KB table, entry (oct << 8) | x.
RN_DEV uint32_t permute_oct(uint32_t x, uint32_t oct) {
	x = (oct & 1u) ? (((x & 0x55u) << 1) | ((x >> 1) & 0x55u)) : x;
	x = (oct & 2u) ? (((x & 0x33u) << 2) | ((x >> 2) & 0x33u)) : x;
	x = (oct & 4u) ? (((x & 0x0Fu) << 4) | (x >> 4)) : x;
	return x;
}

RN_DEV void fill_oct_table(uint8_t *tab, int tid, int nthreads) {
	for (int i = tid; i < 2048; i += nthreads)
		tab[i] = (uint8_t)permute_oct((uint32_t)i & 0xFFu, (uint32_t)i >> 8);
}

// Intersect the (up to) 8 child boxes of a node. Returns the internal-child hit
// bits in traversal order (bit b <-> slot b ^ oct, so bit 0 is the slot built
// for the ray's own octant, i.e. the nearest) and the leaf triangle bits in
// the node's tmask layout (bits 3s .. 3s + 2 for slot s; leaf_record maps a
// bit to its triangle). A hit slot sets all three of its bits, a constant
// per slot, and the node's tmask keeps those of its triangles.
RN_DEV void intersect_children(const uint32_t w[20], const TRay &r, float tmax, uint32_t &ihits, uint32_t &thits,
                               const uint8_t *octab = nullptr) {
	float px = __uint_as_float(w[0]), py = __uint_as_float(w[1]), pz = __uint_as_float(w[2]);
	uint32_t e = w[3];
	// Plane bytes are widened to f16 denormals q * 2^-24 with one v_perm per
	// two bytes and fed to v_fma_mix_f32 against the slab scale times 2^24:
	// q * 2^-24 * (b * 2^24) is exactly q * b, so each t is the same single
	// rounding of q * b + a as with a separate byte-to-float conversion.
	float sx = __uint_as_float(((e & 0xFFu) + 24u) << 23);
	float sy = __uint_as_float((((e >> 8) & 0xFFu) + 24u) << 23);
	float sz = __uint_as_float((((e >> 16) & 0xFFu) + 24u) << 23);
	uint32_t imask = e >> 24;
	float ax = (px - r.o.x) * r.inv.x, ay = (py - r.o.y) * r.inv.y, az = (pz - r.o.z) * r.inv.z;
	float bx = sx * r.inv.x, by = sy * r.inv.y, bz = sz * r.inv.z;
	// words: 4 child_base, 5 tri_base, 6 tmask, 7 pad, 8-9 qlo.x, 10-11 qlo.y,
	// 12-13 qlo.z, 14-15 qhi.x, 16-17 qhi.y, 18-19 qhi.z
	uint32_t nx0 = r.inv.x >= 0 ? w[8] : w[14], nx1 = r.inv.x >= 0 ? w[9] : w[15];
	uint32_t fx0 = r.inv.x >= 0 ? w[14] : w[8], fx1 = r.inv.x >= 0 ? w[15] : w[9];
	uint32_t ny0 = r.inv.y >= 0 ? w[10] : w[16], ny1 = r.inv.y >= 0 ? w[11] : w[17];
	uint32_t fy0 = r.inv.y >= 0 ? w[16] : w[10], fy1 = r.inv.y >= 0 ? w[17] : w[11];
	uint32_t nz0 = r.inv.z >= 0 ? w[12] : w[18], nz1 = r.inv.z >= 0 ? w[13] : w[19];
	uint32_t fz0 = r.inv.z >= 0 ? w[18] : w[12], fz1 = r.inv.z >= 0 ? w[19] : w[13];
	const float tfar_lim = tmax * 1.0000005f;
	// acc bits 0-23: leaf triangles hit; bits 24-31: slots hit (slot order)
	uint32_t acc = 0;
#pragma unroll
	for (int h = 0; h < 4; h++) { // two slots per pass: 2h and 2h + 1
		const int wi = h >> 1;
		const uint32_t sel = (h & 1) ? 0x0C030C02u : 0x0C010C00u;
		const h16x2 qnx = h16_pair(wi ? nx1 : nx0, sel), qfx = h16_pair(wi ? fx1 : fx0, sel);
		const h16x2 qny = h16_pair(wi ? ny1 : ny0, sel), qfy = h16_pair(wi ? fy1 : fy0, sel);
		const h16x2 qnz = h16_pair(wi ? nz1 : nz0, sel), qfz = h16_pair(wi ? fz1 : fz0, sel);
#pragma unroll
		for (int k = 0; k < 2; k++) {
			const int s = 2 * h + k;
			float tnx = fmaf((float)qnx[k], bx, ax);
			float tny = fmaf((float)qny[k], by, ay);
			float tnz = fmaf((float)qnz[k], bz, az);
			float tfx = fmaf((float)qfx[k], bx, ax);
			float tfy = fmaf((float)qfy[k], by, ay);
			float tfz = fmaf((float)qfz[k], bz, az);
			float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, 0.0f));
			float tf = fminf(fminf(tfx, tfy), fminf(tfz, tfar_lim));
			bool hit = tn <= tf * 1.0000005f;
			acc |= hit ? (7u << (3 * s)) | (1u << (24 + s)) : 0u;
		}
	}
	ihits = octab ? (uint32_t)octab[(r.oct << 8) | ((acc >> 24) & imask)] : permute_oct((acc >> 24) & imask, r.oct);
	thits = acc & w[6]; // tmask has no bits above 23
}

// The leaf-ordered triangle record of bit k of a node's tmask layout.
RN_DEV uint32_t leaf_record(uint32_t tri_base, uint32_t tmask, uint32_t k) {
	return tri_base + __builtin_popcount(tmask & ((1u << k) - 1u));
}

// -------------------------------------------------------- triangle tests --

// Watertight test (Woop, Benthin, Wald, JCGT 2(1) 2013). Ray-constant shear
// terms are computed once per ray.
struct WRay {
	uint32_t k; // the axes kx | ky << 2 | kz << 4, packed: one register per ray instead of three
	float sx, sy, sz;
	RN_DEV int kx() const { return (int)(k & 3u); }
	RN_DEV int ky() const { return (int)((k >> 2) & 3u); }
	RN_DEV int kz() const { return (int)(k >> 4); }
};

RN_DEV WRay make_wray(float3 d) {
	WRay w;
	float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
	int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
	int kx = kz == 2 ? 0 : kz + 1;
	int ky = kx == 2 ? 0 : kx + 1;
	if (get3(d, kz) < 0.0f) {
		int t = kx;
		kx = ky;
		ky = t;
	}
	w.k = (uint32_t)kx | ((uint32_t)ky << 2) | ((uint32_t)kz << 4);
	float dz = get3(d, kz);
	w.sx = get3(d, kx) / dz;
	w.sy = get3(d, ky) / dz;
	w.sz = 1.0f / dz;
	return w;
}

// Returns true for a hit with 0 < t < tmax; u = weight of v1, v = weight of v2.
RN_DEV bool tri_watertight(const WRay &w, float3 o, float3 v0, float3 v1, float3 v2, float tmax, float &t, float &u,
                           float &v) {
#pragma clang fp contract(off)
	float3 A = v0 - o, B = v1 - o, C = v2 - o;
	const int kx = w.kx(), ky = w.ky(), kz = w.kz();
	float Az0 = get3(A, kz), Bz0 = get3(B, kz), Cz0 = get3(C, kz);
	float Ax = get3(A, kx) - w.sx * Az0, Ay = get3(A, ky) - w.sy * Az0;
	float Bx = get3(B, kx) - w.sx * Bz0, By = get3(B, ky) - w.sy * Bz0;
	float Cx = get3(C, kx) - w.sx * Cz0, Cy = get3(C, ky) - w.sy * Cz0;
	float U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
	if (U == 0.0f || V == 0.0f || W == 0.0f) {
		// Edge functions recomputed exactly (float products are exact in
		// double). One that is still zero takes the sign it would have if
		// the ray's point moved by (e1, e2), 0 < e2 << e1: the sign of its x
		// derivative, or of its y derivative when that is zero, kept as the
		// sign of the zero. The two triangles of a shared edge see opposite
		// derivatives, so exactly one of them keeps a point on the edge, and
		// around a vertex exactly one triangle of a fan keeps it: a ray
		// crossing a closed mesh at an edge or a vertex hits one triangle
		// there, not none and not two. The derivatives' signs are exact (a
		// difference of floats is zero only when they are equal).
		U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
		V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
		W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
		if (U == 0.0f)
			U = copysignf(0.0f, Cy != By ? Cy - By : Bx - Cx);
		if (V == 0.0f)
			V = copysignf(0.0f, Ay != Cy ? Ay - Cy : Cx - Ax);
		if (W == 0.0f)
			W = copysignf(0.0f, By != Ay ? By - Ay : Ax - Bx);
	}
	// a hit needs U, V and W of one sign
	if (((__float_as_uint(U) ^ __float_as_uint(V)) | (__float_as_uint(U) ^ __float_as_uint(W))) & 0x80000000u)
		return false;
	float det = U + V + W;
	if (det == 0.0f)
		return false;
	float T = U * (w.sz * Az0) + V * (w.sz * Bz0) + W * (w.sz * Cz0);
	if (det < 0.0f) {
		T = -T;
		det = -det;
		V = -V;
		W = -W;
	}
	if (!(T > 0.0f) || !(T < tmax * det))
		return false;
	float rcp = 1.0f / det;
	t = T * rcp;
	u = V * rcp;
	v = W * rcp;
	return true;
}

// The reference test (src/Geometry.cpp:11-25) in glm's FP32 operation order
// (glm/detail/func_matrix.inl:253-262,322-344, type_mat3x3.inl:468-474), with
// contraction off so (t,u,v) are bit-identical to the reference's x86 build.
RN_DEV bool tri_compat(float3 o, float3 d, float3 v0, float3 v1, float3 v2, float &t, float &u, float &v) {
#pragma clang fp contract(off)
	float3 e0 = v1 - v0, e1 = v2 - v0, sp = o - v0;
	float m00 = -d.x, m01 = -d.y, m02 = -d.z;
	float m10 = e0.x, m11 = e0.y, m12 = e0.z;
	float m20 = e1.x, m21 = e1.y, m22 = e1.z;
	float c0 = m11 * m22 - m21 * m12, c1 = m01 * m22 - m21 * m02, c2 = m01 * m12 - m11 * m02;
	float det = m00 * c0 - m10 * c1 + m20 * c2;
	if (fabsf(det) < 1e-8f)
		return false;
	float ood = 1.0f / det;
	float i00 = c0 * ood;
	float i10 = -(m10 * m22 - m20 * m12) * ood;
	float i20 = (m10 * m21 - m20 * m11) * ood;
	float i01 = -c1 * ood;
	float i11 = (m00 * m22 - m20 * m02) * ood;
	float i21 = -(m00 * m21 - m20 * m01) * ood;
	float i02 = c2 * ood;
	float i12 = -(m00 * m12 - m10 * m02) * ood;
	float i22 = (m00 * m11 - m10 * m01) * ood;
	float rx = i00 * sp.x + i10 * sp.y + i20 * sp.z;
	float ry = i01 * sp.x + i11 * sp.y + i21 * sp.z;
	float rz = i02 * sp.x + i12 * sp.y + i22 * sp.z;
	t = rx;
	u = ry;
	v = rz;
	return (ry >= 0.0f) && (rz >= 0.0f) && (ry + rz <= 1.0f) && (rx > 1e-5f);
}

// ------------------------------------------------------------- traversal --

// A node group on the stack: child_base | imask << 32 | pending hit bits << 40.
RN_DEV uint64_t mk_group(uint32_t base, uint32_t imask, uint32_t hits) {
	return (uint64_t)base | ((uint64_t)imask << 32) | ((uint64_t)hits << 40);
}

struct Hit {
	int prim;      // original triangle id, -1 = miss (any-hit: 1 = occluded)
	uint32_t leaf; // leaf-ordered triangle index of the hit
	float t, u, v;
	float3 ng;     // with NG: cross(v1 - v0, v2 - v0) of the hit triangle
};

// Scene buffers are always global memory; loading through address space 1
// gives global_load (vmcnt only) instead of flat_load, which would also hold
// lgkmcnt and serialise against the LDS stack.
typedef unsigned int nu4 __attribute__((ext_vector_type(4)));
typedef float nf4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const nu4 g_nu4;
typedef __attribute__((address_space(1))) const nf4 g_nf4;

RN_DEV float4 gload_f4(const float4 *base, uint32_t i) {
	nf4 v = ((g_nf4 *)base)[i];
	return make_float4(v.x, v.y, v.z, v.w);
}

// The x, y, z words of float4 record i (one three-word load).
RN_DEV float3 gload_f3(const float4 *base, uint32_t i) {
	typedef float nf3 __attribute__((ext_vector_type(3)));
	const nf3 v = *(__attribute__((address_space(1))) const nf3 *)((const float *)base + 4 * (size_t)i);
	return f3(v.x, v.y, v.z);
}

RN_DEV float gload_f(const float *p) { return *(__attribute__((address_space(1))) const float *)p; }
RN_DEV int gload_i(const int *p) { return *(__attribute__((address_space(1))) const int *)p; }

// Float adds into global memory and into LDS through pointers the compiler
// sees as generic (read from a scene table, or held in a struct). Through a
// generic pointer an atomic is a flat instruction, which the hardware also
// counts on the LDS counter: the next wait for an LDS access (the traversal
// stack) then waits for the atomic's round trip to memory too. Naming the
// address space makes them global and LDS atomics. Relaxed, like atomicAdd.
RN_DEV void atomic_add_global(float *p, float v) {
	__hip_atomic_fetch_add((__attribute__((address_space(1))) float *)p, v, __ATOMIC_RELAXED,
	                       __HIP_MEMORY_SCOPE_AGENT);
}
RN_DEV void atomic_add_lds(float *p, float v) {
	__hip_atomic_fetch_add((__attribute__((address_space(3))) float *)p, v, __ATOMIC_RELAXED,
	                       __HIP_MEMORY_SCOPE_WORKGROUP);
}

RN_DEV void load_node_g(const Bvh8Node *nodes, uint32_t i, uint32_t w[20]) {
	g_nu4 *p = (g_nu4 *)(nodes + i);
#pragma unroll
	for (int k = 0; k < 5; k++) {
		nu4 q = p[k];
		w[4 * k] = q.x;
		w[4 * k + 1] = q.y;
		w[4 * k + 2] = q.z;
		w[4 * k + 3] = q.w;
	}
}

// Resumable wide-BVH traversal of one ray, one ray per lane. trav_step runs
// one iteration (visit at most one node, test the triangles of the leaves it
// hit, pop the stack when the current group is exhausted) and returns true
// when the query is answered, so a lane can be refilled with a new ray while
// the other lanes of its wave are still traversing.
//
// The short stack lives in LDS: entry k of lane `lane` is lds[k * ldsStride +
// lane]; entries beyond LDS_DEPTH spill to a lane-private array of SPILL
// entries. The stack never holds more than one entry per tree level below the
// root, so the host picks an instance with LDS_DEPTH + SPILL >= depth - 1; a
// deeper push is dropped rather than written out of bounds. QUERY 0 = closest
// hit, 1 = any hit (early exit). TRI 0 = watertight, 1 = reference (compat)
// test. Triangles equal to `ignore` (original id) are skipped. With COUNT,
// cnt[0] += nodes visited and cnt[1] += triangles tested.
template <int SPILL>
struct Trav {
	TRay r;
	WRay wr;
	uint64_t grp;
	uint32_t tri_base, thits, tmask;
	uint32_t tri_base2, thits2, tmask2; // a second pending leaf group (kTriSlots 2)
	int sp;
	Hit h;
	uint64_t spill[SPILL > 0 ? SPILL : 1];
};

template <int TRI, int SPILL>
RN_DEV void trav_init(Trav<SPILL> &s, float3 o, float3 d, float tmax) {
	s.r = make_tray(o, d);
	if (TRI == 0)
		s.wr = make_wray(d);
	s.h.prim = -1;
	s.h.leaf = 0;
	s.h.t = tmax;
	s.h.u = 0.0f;
	s.h.v = 0.0f;
	s.sp = 0;
	// A ray with a NaN or infinite origin or direction component starts with
	// nothing to visit: it misses (any hit: unoccluded), as it does against
	// every triangle. (Left to the slab test, NaN slab distances vanish in
	// the min/max and such a ray would enter every box of the tree.)
	const bool finite = fabsf(o.x) < INFINITY && fabsf(o.y) < INFINITY && fabsf(o.z) < INFINITY &&
	                    fabsf(d.x) < INFINITY && fabsf(d.y) < INFINITY && fabsf(d.z) < INFINITY;
	s.grp = finite ? mk_group(0u, 1u, 1u << (0u ^ s.r.oct)) : 0ull;
	s.tri_base = 0;
	s.thits = 0;
	s.tmask = 0;
	s.tri_base2 = 0;
	s.thits2 = 0;
	s.tmask2 = 0;
}

// Triangles a lane tests per step (0 = all of the node's hit leaves at
// once). With a budget, a lane with triangles left tests them in the next
// steps before it visits another node, so one lane with many leaf hits does
// not stall its wave for all of them in one step. Measured on MI355X at 1M
// triangles (1080p, 64 spp, depth 8): budget 1 gives the fastest wavefront
// (165 ms per frame against 202 ms with no budget; 2, 3, 4 and 6 fall in
// between in that order) and the fastest megakernel.
#ifndef RN_TRI_BUDGET
#define RN_TRI_BUDGET 1
#endif
constexpr int kTriBudget = RN_TRI_BUDGET;

// Pending leaf groups a lane holds (1 or 2). With 2, a lane with triangles
// still to test may visit its next node in the same step as long as the
// second slot is free, so node visits and triangle tests overlap.
#ifndef RN_TRI_SLOTS
#define RN_TRI_SLOTS 2
#endif
constexpr int kTriSlots = RN_TRI_SLOTS;
static_assert(kTriSlots == 1 || (kTriSlots == 2 && kTriBudget > 0), "two slots need a triangle budget");

template <int QUERY, int TRI, int LDS_DEPTH, int SPILL, bool COUNT = false, bool NG = false>
RN_DEV bool trav_step(Trav<SPILL> &s, const Bvh8Node *__restrict__ nodes, const float4 *__restrict__ tris, int ignore,
                      uint64_t *lds, int ldsStride, int lane, uint32_t *cnt = nullptr, const uint8_t *octab = nullptr) {
	uint32_t hits = (uint32_t)(s.grp >> 40);
	if (hits && (kTriBudget == 0 || (kTriSlots == 1 ? s.thits == 0 : s.thits2 == 0))) {
		uint32_t b = __builtin_ctz(hits);
		hits &= hits - 1u;
		uint32_t slot = b ^ s.r.oct;
		uint32_t imask = (uint32_t)(s.grp >> 32) & 0xFFu;
		uint32_t node = (uint32_t)s.grp + __builtin_popcount(imask & ((1u << slot) - 1u));
		if (hits) {
			uint64_t rest = (s.grp & 0xFFFFFFFFFFull) | ((uint64_t)hits << 40);
			if (s.sp < LDS_DEPTH) {
				lds[s.sp * ldsStride + lane] = rest;
				s.sp++;
			} else if (SPILL > 0 && s.sp < LDS_DEPTH + SPILL) {
				s.spill[s.sp - LDS_DEPTH] = rest;
				s.sp++;
			}
		}
		uint32_t w[20];
		load_node_g(nodes, node, w);
		if (COUNT)
			cnt[0]++;
		uint32_t ih, th;
		intersect_children(w, s.r, s.h.t, ih, th, octab);
		s.grp = mk_group(w[4], w[3] >> 24, ih);
		if (kTriSlots == 1 || s.thits == 0) {
			s.thits = th;
			s.tri_base = w[5];
			s.tmask = w[6];
		} else {
			s.thits2 = th;
			s.tri_base2 = w[5];
			s.tmask2 = w[6];
		}
	}
	if (s.thits) {
		uint32_t m = s.thits;
		int budget = kTriBudget;
		while (m && (kTriBudget == 0 || budget-- > 0)) {
			uint32_t k = __builtin_ctz(m);
			m &= m - 1u;
			uint32_t ti = leaf_record(s.tri_base, s.tmask, k);
			// The id test does not branch before the triangle test: a branch
			// there lets the compiler sink the second and third loads behind
			// the first, which costs a second memory round trip per test.
			// v1 and v2 load three words each: with a fourth, unused word the
			// register allocator may reuse its register while the load is in
			// flight and wait for it before issuing the next load
			const float4 a = gload_f4(tris, 3 * ti);
			const float3 b4 = gload_f3(tris, 3 * ti + 1), c4 = gload_f3(tris, 3 * ti + 2);
			const int prim = __float_as_int(a.w);
			const bool skip = prim == ignore;
			if (COUNT && !skip)
				cnt[1]++;
			float tt, uu, vv;
			bool ok;
			if (TRI == 0)
				ok = tri_watertight(s.wr, s.r.o, xyz(a), b4, c4, s.h.t, tt, uu, vv);
			else
				ok = tri_compat(s.r.o, s.r.d, xyz(a), b4, c4, tt, uu, vv) && tt < s.h.t;
			if (ok && !skip) {
				if (QUERY == 1) {
					s.h.prim = 1;
					return true;
				}
				s.h.prim = prim;
				s.h.leaf = ti;
				s.h.t = tt;
				s.h.u = uu;
				s.h.v = vv;
				if (NG) // the vertices are in registers now; the shader needs no reload
					s.h.ng = cross3(b4 - xyz(a), c4 - xyz(a));
			}
		}
		s.thits = m;
		if (kTriSlots == 2 && m == 0) {
			s.thits = s.thits2;
			s.tri_base = s.tri_base2;
			s.tmask = s.tmask2;
			s.thits2 = 0;
		}
	}
	if ((s.grp >> 40) == 0 && (kTriBudget == 0 || kTriSlots == 2 || s.thits == 0)) {
		if (s.sp == 0) {
			if (kTriSlots == 2 && s.thits != 0)
				return false; // triangles left: finish them first
			if (QUERY == 1)
				s.h.prim = 0;
			return true;
		}
		s.sp--;
		// the LDS entry is read through an LDS pointer: a select between the
		// two addresses would otherwise become one flat load
		if (SPILL == 0 || s.sp < LDS_DEPTH)
			s.grp = *(__attribute__((address_space(3))) uint64_t *)(lds + s.sp * ldsStride + lane);
		else
			s.grp = s.spill[s.sp - LDS_DEPTH];
	}
	return false;
}

// Whole-query form of the traversal above.
template <int QUERY, int TRI, int LDS_DEPTH, int SPILL, bool COUNT = false>
RN_DEV Hit traverse(const Bvh8Node *__restrict__ nodes, const float4 *__restrict__ tris, float3 o, float3 d,
                    float tmax, int ignore, uint64_t *lds, int ldsStride, int lane, uint32_t *cnt = nullptr) {
	Trav<SPILL> s;
	trav_init<TRI>(s, o, d, tmax);
	while (!trav_step<QUERY, TRI, LDS_DEPTH, SPILL, COUNT>(s, nodes, tris, ignore, lds, ldsStride, lane, cnt)) {
	}
	return s.h;
}
