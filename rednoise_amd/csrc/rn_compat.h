// Device code of the reference estimator ("compat" path of rn_render):
// camera, texel selection and surface colour in the reference's FP32
// operation order.
#pragma once

#include "rn_device.h"

// Per-triangle attributes of the reference path, original triangle order.
struct CompatDev {
	const float4 *cnorm;     // face normal xyz, w = material id (int bits)
	const float *vnorm;      // smoothed vertex normals, 9 per triangle
	const float4 *ccol;      // Kd colour ints as floats 0..255, w = texture id (int bits, -1 none)
	const float *ctc;        // texture coordinates, 6 per triangle
	const uint32_t *texels;  // every texture's 0xAARRGGBB pixels, concatenated
	const int4 *tex_info;    // per texture: (first texel, width, height, 0)
};

enum CompatMaterial {
	CM_DIFFUSE = 0,
	CM_MIRROR,
	CM_GLASS,
	CM_PROCEDURAL,
	CM_BUMP,
	CM_PARALLAX,
	CM_METAL,
	CM_DISPERSIVE,
	CM_SUBSURFACE,
	CM_NORMALMAP,
	CM_EMISSIVE
};

// glm's dot/normalize/cross (func_geometric.inl:48-55,68-79,82-90).
RN_DEV float gdot(float3 a, float3 b) {
#pragma clang fp contract(off)
	float x = a.x * b.x, y = a.y * b.y, z = a.z * b.z;
	return x + y + z;
}
RN_DEV float3 gnormalize(float3 v) {
#pragma clang fp contract(off)
	float s = 1.0f / sqrtf(gdot(v, v));
	return f3(v.x * s, v.y * s, v.z * s);
}
RN_DEV float3 gcross(float3 x, float3 y) {
#pragma clang fp contract(off)
	return f3(x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y);
}
// Component-wise vector arithmetic that is never fused into multiply-adds
// (the float3 operators may be, once inlined into a caller).
RN_DEV float3 gadd(float3 a, float3 b) {
#pragma clang fp contract(off)
	return f3(a.x + b.x, a.y + b.y, a.z + b.z);
}
RN_DEV float3 gsub(float3 a, float3 b) {
#pragma clang fp contract(off)
	return f3(a.x - b.x, a.y - b.y, a.z - b.z);
}
RN_DEV float3 gmul(float3 a, float3 b) {
#pragma clang fp contract(off)
	return f3(a.x * b.x, a.y * b.y, a.z * b.z);
}
RN_DEV float3 gscale(float3 a, float s) {
#pragma clang fp contract(off)
	return f3(a.x * s, a.y * s, a.z * s);
}
RN_DEV float3 gdiv(float3 a, float s) {
#pragma clang fp contract(off)
	return f3(a.x / s, a.y / s, a.z / s);
}

// rn_render's camera: Camera(w,h,2,(0,0,cam_z)).lookAt(0) then primaryRay
// (src/capi.cpp:49-50, src/Camera.cpp:27-33,69-79) through the screen offset
// (sx, sy) pixels from the image centre.
RN_DEV void compat_camera_ray(float sx, float sy, float cam_z, float3 &o, float3 &d) {
#pragma clang fp contract(off)
	float3 pos = f3(0.0f, 0.0f, cam_z);
	float3 fwd = gnormalize(f3(pos.x - 0.0f, pos.y - 0.0f, pos.z - 0.0f));
	float3 right = gnormalize(gcross(f3(0.0f, 1.0f, 0.0f), fwd));
	float3 up = gcross(fwd, right);
	float f = 2.0f * 150.0f;
	float a = sx / f, b = sy / f, c = -1.0f;
	float3 dir = f3(right.x * a + up.x * b + fwd.x * c, right.y * a + up.y * b + fwd.y * c,
	                right.z * a + up.z * b + fwd.z * c);
	o = pos;
	d = gnormalize(dir);
}

// The pixel (x, y) with jitter (jx, jy) (src/Renderer.cpp:1148).
RN_DEV void compat_primary_ray(int x, int y, int W, int H, float cam_z, float jx, float jy, float3 &o, float3 &d) {
#pragma clang fp contract(off)
	float sx = (float)x + 0.5f + jx - (float)W / 2.0f;
	float sy = -((float)y + 0.5f + jy - (float)H / 2.0f);
	compat_camera_ray(sx, sy, cam_z, o, d);
}

// Barycentric texture coordinate (src/Renderer.cpp:642-649).
RN_DEV void compat_uv(const CompatDev &cd, int prim, float u, float v, float &tu, float &tv) {
#pragma clang fp contract(off)
	const float *tp = cd.ctc + 6 * (size_t)prim;
	float w0 = 1.0f - u - v, w1 = u, w2 = v;
	tu = w0 * tp[0] + w1 * tp[2] + w2 * tp[4];
	tv = w0 * tp[1] + w1 * tp[3] + w2 * tp[5];
}

// Nearest texel (src/Renderer.cpp:584-591) as a linear index.
RN_DEV int compat_texel_of(int tw, int th, float tu, float tv) {
#pragma clang fp contract(off)
	int tx = (int)(tu * (float)(tw - 1));
	int ty = (int)((1.0f - tv) * (float)(th - 1));
	tx = tx < 0 ? 0 : (tx > tw - 1 ? tw - 1 : tx);
	ty = ty < 0 ? 0 : (ty > th - 1 ? th - 1 : ty);
	return ty * tw + tx;
}

RN_DEV int compat_texel_index(const CompatDev &cd, int prim, float u, float v) {
	int tex = __float_as_int(cd.ccol[prim].w);
	if (tex < 0)
		return -1;
	int4 ti = cd.tex_info[tex];
	float tu, tv;
	compat_uv(cd, prim, u, v, tu, tv);
	return compat_texel_of(ti.y, ti.z, tu, tv);
}
