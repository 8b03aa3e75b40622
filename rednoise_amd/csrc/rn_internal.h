// Host-side internals shared by the translation units of librednoise.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "rn_types.h"

namespace rn {

// Error raised anywhere below the C ABI; every extern "C" entry catches it and
// turns it into a 0/NULL return plus rn_last_error().
struct Error : std::runtime_error {
	explicit Error(const std::string &m) : std::runtime_error(m) {}
};

void hip_check(hipError_t e, const char *what, const char *file, int line);

// Failure injection for tests of the error paths. RN_FAULT_INJECT holds a
// comma-separated list, read at every check:
//   alloc                 every device and pinned allocation fails
//   bvh_depth             every built BVH counts as deeper than the
//                         traversal stack (the guard before any traversal)
//   wavefront_iterations  the wavefront's iteration limit is one launch group
// Each failure is raised on the host, before a kernel launch or between
// synchronised launches, and reported like a real one.
bool fault_injected(const char *what);
#define RN_HIP(x) ::rn::hip_check((x), #x, __FILE__, __LINE__)

// Grow-only device allocation owned by whoever holds it.
struct DevBuf {
	void *ptr = nullptr;
	size_t bytes = 0;
	DevBuf() = default;
	DevBuf(const DevBuf &) = delete;
	DevBuf &operator=(const DevBuf &) = delete;
	~DevBuf() { release(); }
	void release();
	// Ensure capacity >= n bytes (contents are not preserved on growth).
	void *reserve(size_t n);
	template <class T> T *as() const { return static_cast<T *>(ptr); }
};

// BVH8 build products and scratch, owned by a scene and reused by rebuilds.
struct Bvh {
	DevBuf nodes, tris;
	int n_nodes = 0, depth = 0, width = 8, n_tris = 0;
	double build_ms = 0.0;
	// scratch
	DevBuf bounds, keys, vals, keys2, vals2, hist, scan_tmp, left, right, parent, range, boxes, visit, task, counters;
	DevBuf count, ploc_ref, ploc_box, ploc_cnt, ploc_nn, ploc_flags;
	int builder = 0; // 0 = PLOC, 1 = Karras LBVH (the builder of the last build)
};

// Build the BVH of n triangles whose vertices (9n floats) are already on the
// device. Stream-ordered; the call synchronises once per BVH level to learn
// the level sizes.
#ifndef RN_DEFAULT_BUILDER
#define RN_DEFAULT_BUILDER 0
#endif
void build_bvh(Bvh &bvh, const float *d_verts, int n, int width, hipStream_t s, int builder = RN_DEFAULT_BUILDER);

// Stages exposed for testing (device buffers in, host copies out).
void lbvh_stages(const float *d_verts, int n, uint32_t *h_codes, int *h_order, int *h_left, int *h_right,
                 hipStream_t s);
// The PLOC builder's sorted order and binary tree (same layout as lbvh_stages).
void ploc_stages(const float *d_verts, int n, int *h_order, int *h_left, int *h_right, hipStream_t s);
void radix_sort_pairs(uint32_t *d_keys, uint32_t *d_vals, uint32_t *d_keys_tmp, uint32_t *d_vals_tmp, int n,
                      int bits, DevBuf &hist, DevBuf &scan_tmp, hipStream_t s);
// Exclusive prefix sum of n words; tmp holds device_scan_tmp_words(n) words.
void device_exclusive_scan(const uint32_t *in, uint32_t *out, size_t n, uint32_t *tmp, hipStream_t s);
size_t device_scan_tmp_words(size_t n);

} // namespace rn
