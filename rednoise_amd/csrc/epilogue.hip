// Progressive accumulation and the fused tone-map + RGBA8 pack.
//
// A progressive render adds frames of s samples per pixel to a float64 sum;
// resolving divides by the sample count, applies exposure, the Reinhard
// operator c / (c + 1) and gamma, and packs RGBA8 with alpha 255 in the same
// pass (the steps of toneMap, src/Renderer.cpp:1357-1368, applied to the HDR
// mean instead of an 8-bit canvas; with exposure 1 and gamma 1 it is the
// OpenCL path's 255 * c / (c + 1), gpu/pathtracer.cl:122-132, and the byte
// order is rn_render's RGBA8, src/capi.cpp:69-77).
#include "rn_device.h"
#include "rn_internal.h"
#include "rn_kernels.h"

#include <algorithm>

namespace rn {

namespace {

__global__ void k_accum_add(double *sum, const float *mean, size_t n, double weight) {
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
		sum[i] += (double)mean[i] * weight;
}

__device__ __forceinline__ unsigned char tone_byte(float c, float exposure, float inv_gamma, bool gamma_one) {
	c = fmaxf(c, 0.0f) * exposure;
	c = c / (c + 1.0f);
	c = (gamma_one ? c : powf(c, inv_gamma)) * 255.0f;
	return (unsigned char)min(255, (int)c);
}

__global__ void k_tonemap_rgba8(const double *sum, size_t npx, double inv_count, float exposure, float inv_gamma,
                                int gamma_one, uchar4 *rgba) {
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < npx; i += (size_t)gridDim.x * blockDim.x) {
		const float r = (float)(sum[3 * i] * inv_count), g = (float)(sum[3 * i + 1] * inv_count),
		            b = (float)(sum[3 * i + 2] * inv_count);
		rgba[i] = make_uchar4(tone_byte(r, exposure, inv_gamma, gamma_one), tone_byte(g, exposure, inv_gamma, gamma_one),
		                      tone_byte(b, exposure, inv_gamma, gamma_one), 255);
	}
}

__global__ void k_mean(const double *sum, size_t n, double inv_count, float *out) {
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
		out[i] = (float)(sum[i] * inv_count);
}

unsigned grid_of(size_t n) { return (unsigned)std::min<size_t>(4096, (n + 255) / 256 + 1); }

} // namespace

void launch_accum_add(double *sum, const float *mean, size_t n, double weight, hipStream_t s) {
	k_accum_add<<<grid_of(n), 256, 0, s>>>(sum, mean, n, weight);
	RN_HIP(hipGetLastError());
}

void launch_tonemap_rgba8(const double *sum, size_t npx, double inv_count, float exposure, float gamma,
                          unsigned char *rgba, hipStream_t s) {
	k_tonemap_rgba8<<<grid_of(npx), 256, 0, s>>>(sum, npx, inv_count, exposure, 1.0f / gamma, gamma == 1.0f ? 1 : 0,
	                                             reinterpret_cast<uchar4 *>(rgba));
	RN_HIP(hipGetLastError());
}

void launch_mean(const double *sum, size_t n, double inv_count, float *out, hipStream_t s) {
	k_mean<<<grid_of(n), 256, 0, s>>>(sum, n, inv_count, out);
	RN_HIP(hipGetLastError());
}

} // namespace rn
