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
};

struct WaveStats {
	double extend_ms = 0.0, shade_ms = 0.0, shadow_ms = 0.0;
	uint64_t extend_launches = 0, iterations = 0;
};

// Renders every item (outputs zeroed, accumulated and divided by spp).
// Counters receive closest-hit rays, shadow rays and paths.
void wavefront_render(const WaveItem *items, int count, const WaveConfig &cfg, DevCounters *cnt, hipStream_t s,
                      WaveStats &st);

void fill_zero(void *p, size_t bytes, hipStream_t s);

// Frees the wavefront queues (they are reallocated by the next render).
void wavefront_release();

} // namespace rn
