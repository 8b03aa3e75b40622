// Ray-query kernels: batches of arbitrary rays, the reference camera's
// primary rays (bit-exact parity path), and Philox for host cross-checks.
#include "rn_device.h"
#include "rn_internal.h"
#include "rn_kernels.h"

namespace rn {

namespace {

constexpr int kQBlock = 128;

template <int QUERY, int TRI, int LDS, int SPILL>
__global__ __launch_bounds__(kQBlock) void k_trace_rays(const Bvh8Node *nodes, const float4 *tris, int n,
                                                        const float *o, const float *d, const float *tmax,
                                                        const int *ignore, int *prim, float *t, float *u, float *v) {
	__shared__ uint64_t stk[LDS * kQBlock];
	int i = blockIdx.x * kQBlock + threadIdx.x;
	if (i >= n)
		return;
	float3 ro = f3(o[3 * i], o[3 * i + 1], o[3 * i + 2]);
	float3 rd = f3(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
	float tm = tmax ? tmax[i] : INFINITY;
	int ig = ignore ? ignore[i] : -1;
	Hit h = traverse<QUERY, TRI, LDS, SPILL>(nodes, tris, ro, rd, tm, ig, stk, kQBlock, threadIdx.x);
	prim[i] = h.prim;
	if (QUERY == 0) {
		t[i] = h.prim >= 0 ? h.t : INFINITY;
		u[i] = h.u;
		v[i] = h.v;
	}
}

template <int LDS, int SPILL>
__global__ __launch_bounds__(kQBlock) void k_primary_compat(const Bvh8Node *nodes, const float4 *tris, int W, int H,
                                                            float cam_z, const CompatDev cd, int *prim, float *t,
                                                            float *u, float *v, int *texel) {
	__shared__ uint64_t stk[LDS * kQBlock];
	int i = blockIdx.x * kQBlock + threadIdx.x;
	if (i >= W * H)
		return;
	float3 o, d;
	compat_primary_ray(i % W, i / W, W, H, cam_z, 0.0f, 0.0f, o, d);
	Hit h = traverse<0, 1, LDS, SPILL>(nodes, tris, o, d, INFINITY, -1, stk, kQBlock, threadIdx.x);
	prim[i] = h.prim;
	t[i] = h.prim >= 0 ? h.t : INFINITY;
	u[i] = h.u;
	v[i] = h.v;
	int tx = -1;
	if (h.prim >= 0)
		tx = compat_texel_index(cd, h.prim, h.u, h.v);
	texel[i] = tx;
}

__global__ void k_philox(int n, const uint32_t *ctr, const uint32_t *key, uint32_t *out) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	u32x4 r = philox4x32_10(ctr[4 * i], ctr[4 * i + 1], ctr[4 * i + 2], ctr[4 * i + 3], key[2 * i], key[2 * i + 1]);
	out[4 * i] = r.x;
	out[4 * i + 1] = r.y;
	out[4 * i + 2] = r.z;
	out[4 * i + 3] = r.w;
}

} // namespace

void launch_trace_rays(const DevScene &sc, int n, const float *o, const float *d, const float *tmax,
                       const int *ignore, int query, int tri, int *prim, float *t, float *u, float *v,
                       hipStream_t s) {
	unsigned g = (unsigned)((n + kQBlock - 1) / kQBlock);
	if (n <= 0)
		return;
	bool deep = sc.depth - 1 > kLdsDepth;
#define RN_Q(Q, T)                                                                                                    \
	if (deep)                                                                                                          \
		k_trace_rays<Q, T, kLdsDepth, kSpillDepth>                                                                     \
		    <<<g, kQBlock, 0, s>>>(sc.nodes, sc.tris, n, o, d, tmax, ignore, prim, t, u, v);                           \
	else                                                                                                               \
		k_trace_rays<Q, T, kLdsDepth, 0><<<g, kQBlock, 0, s>>>(sc.nodes, sc.tris, n, o, d, tmax, ignore, prim, t, u, v);
	if (query == 0 && tri == 0) {
		RN_Q(0, 0)
	} else if (query == 0 && tri == 1) {
		RN_Q(0, 1)
	} else if (query == 1 && tri == 0) {
		RN_Q(1, 0)
	} else {
		RN_Q(1, 1)
	}
#undef RN_Q
	RN_HIP(hipGetLastError());
}

void launch_primary_compat(const DevScene &sc, const CompatDev &cd, int W, int H, float cam_z, int *prim, float *t,
                           float *u, float *v, int *texel, hipStream_t s) {
	unsigned g = (unsigned)((W * H + kQBlock - 1) / kQBlock);
	if (sc.depth - 1 > kLdsDepth)
		k_primary_compat<kLdsDepth, kSpillDepth>
		    <<<g, kQBlock, 0, s>>>(sc.nodes, sc.tris, W, H, cam_z, cd, prim, t, u, v, texel);
	else
		k_primary_compat<kLdsDepth, 0><<<g, kQBlock, 0, s>>>(sc.nodes, sc.tris, W, H, cam_z, cd, prim, t, u, v, texel);
	RN_HIP(hipGetLastError());
}

void launch_philox(int n, const uint32_t *ctr, const uint32_t *key, uint32_t *out, hipStream_t s) {
	if (n <= 0)
		return;
	k_philox<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, ctr, key, out);
	RN_HIP(hipGetLastError());
}

} // namespace rn
