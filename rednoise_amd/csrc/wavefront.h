// Wavefront integrator of the physical estimator, for one scene or a batch of
// scenes sharing one path queue.
#pragma once

#include "rn_internal.h"
#include "rn_kernels.h"

namespace rn {

struct WaveItem {
	DevScene scene;
	DevCamera cam;
	DevParams params;
	float *out; // width*height*3, device
};

struct WaveConfig {
	int queue_paths = 0;     // <= 0: min(total paths, 2^27, free memory / 4 / 160 B)
	int block = 0;           // <= 0: 256
	bool use_graph = false;  // capture the iteration loop in a hipGraph
	int iters_per_graph = 0; // <= 0: 8 (rounded up to even)
	bool time_stages = false;
	bool fb_atomics = false; // contributions by frame-buffer float atomics instead of per-path radiance slots
};

struct WaveStats {
	double extend_ms = 0.0, shade_ms = 0.0, shadow_ms = 0.0;
	uint64_t extend_launches = 0, iterations = 0;
};

// Device memory handed in by the caller in place of the library's own
// queues: paths in flight, the control block and the scene table are carved
// from [base, base + bytes). With it the render allocates nothing (no graph,
// no stage timing; the done flag is read into host stack memory).
struct WaveScratch {
	void *base = nullptr;
	size_t bytes = 0;
};
// Bytes of caller memory a render of `count` scenes with `capacity` paths in
// flight needs, and the largest capacity that fits in `bytes`.
size_t wave_scratch_bytes(unsigned capacity, int count);
unsigned wave_scratch_capacity(size_t bytes, int count);

// Renders every item (outputs zeroed, accumulated and divided by spp).
// Counters receive closest-hit rays, shadow rays and paths.
void wavefront_render(const WaveItem *items, int count, const WaveConfig &cfg, DevCounters *cnt, hipStream_t s,
                      WaveStats &st, const WaveScratch *scratch = nullptr);

void fill_zero(void *p, size_t bytes, hipStream_t s);

// Frees the wavefront queues (they are reallocated by the next render).
void wavefront_release();

} // namespace rn
