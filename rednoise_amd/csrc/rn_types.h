// Plain data shared by host code and kernels. Layouts are fixed: the BVH8
// node is exported through rn_scene_export_bvh and checked on the CPU.
#pragma once

#include <stdint.h>

// Compressed 8-wide BVH node, 80 bytes (five 16-byte loads). Child boxes are
// stored relative to the origin p with one power-of-two scale per axis
// (2^(e-127)) and 8-bit conservative bounds, the quantisation of Ylitie,
// Karras and Laine, "Efficient Incoherent Ray Traversal on GPUs Through
// Compressed Wide BVHs" (HPG 2017).
//   imask bit s   : slot s is an internal child; internal children are stored
//                   contiguously from child_base in slot order.
//   tmask         : leaf triangles, three bits per slot: bit 3s + j is set
//                   when slot s is a leaf of more than j triangles (none for
//                   an internal or empty slot). Leaves follow each other in
//                   slot order from tri_base, so the triangle of bit k is
//                   record tri_base + (number of tmask bits below bit k).
//   pad           : zero.
//   empty slot    : qlo = 255, qhi = 0 on every axis (never hit).
// Slot s is assigned to the child lying furthest along the diagonal whose
// component k is +1 when bit k of s is set and -1 otherwise. A ray with sign
// octant o (bit k set when d[k] < 0) enters a node from the side of diagonal
// o, so it visits slot o first and slot o^7 last.
struct alignas(16) Bvh8Node {
	float p[3];
	uint8_t e[3];
	uint8_t imask;
	uint32_t child_base;
	uint32_t tri_base;
	uint32_t tmask;
	uint32_t pad;
	uint8_t qlo[3][8];
	uint8_t qhi[3][8];
};
static_assert(sizeof(Bvh8Node) == 80, "BVH8 node must be 80 bytes");

// Pinhole camera in the form the kernels use (see rn_camera_basis).
struct DevCamera {
	float eye[3];
	float du[3];
	float dv[3];
	float w[3];
};

// Estimator parameters as the kernels see them.
struct DevParams {
	int width, height, spp, max_depth, rr_depth;
	float rr_prob;
	uint32_t seed, scene_id;
	float ray_eps;
	uint32_t sample_base; // first sample index: a progressive render continues where it stopped
};

// Device view of one scene (pointers into HBM).
struct DevScene {
	const Bvh8Node *nodes;
	const float4 *tris;       // 3 float4 per leaf-ordered triangle, v0.w = prim id
	const float *albedo;      // 3n, original order
	const float *emission;    // 3n, original order
	const int *light_tri;     // nl light triangle ids
	const float *light_cdf;   // nl, last = 1
	const float4 *light_geom; // 3 float4 per light: v0, e1, e2 (xyz)
	int n, nl;
	float total_light_area;
	int depth;                // BVH8 levels
};
