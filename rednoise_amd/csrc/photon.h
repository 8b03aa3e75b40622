// Photon mapping (renderPhotonMapped, src/Renderer.cpp:1172-1346, with the
// PhotonMap of src/Photon.cpp) and bidirectional path tracing
// (renderBidirectional, src/BDPT.cpp:66-144) on the GPU: photon.hip.
#pragma once

#include "rn_internal.h"
#include "rn_kernels.h"

namespace rn {

// A photon map in device memory. Photons are kept twice: in emission order
// (what PhotonMap::photons holds) and sorted by (cell key, index), which
// turns the reference's hash buckets into runs; the runs are indexed by a
// sorted table of the occupied cells' keys.
struct PhotonMap {
	int n = 0;           // photons
	int n_cells = 0;     // occupied cells
	float cell = 0.25f;  // cell side (PhotonMap::build's radius)
	DevBuf a, b, c;      // emission order: (position, power.x), (direction, power.y), power.z
	DevBuf sa, sb, sc;   // the same, sorted by (cell key, index)
	DevBuf cell_key;     // uint64 per occupied cell, ascending
	DevBuf cell_start;   // n_cells + 1 offsets into the sorted photons
	// scratch
	DevBuf dep_a, dep_b, dep_c, dep_n, offs, keys_lo, keys_hi, perm, tmp_k, tmp_v, hist, scan_tmp, flags, total;
};

// Device view of a PhotonMap.
struct PhotonMapDev {
	const float4 *a, *b;
	const float *c;
	const unsigned long long *key;
	const uint32_t *start;
	int n_cells;
	float cell;
};

inline PhotonMapDev photon_map_dev(const PhotonMap &m) {
	return PhotonMapDev{m.sa.as<float4>(), m.sb.as<float4>(), m.sc.as<float>(), m.cell_key.as<unsigned long long>(),
	                    m.cell_start.as<uint32_t>(), m.n_cells, m.cell};
}

// Pass 1 of renderPhotonMapped: emit num_photons from the default light,
// trace them and build the map with cell side 0.25. Synchronises once (to
// learn the photon count).
void photon_trace(PhotonMap &m, const DevScene &sc, const CompatDev &cd, int num_photons, uint32_t key0,
                  uint32_t key1, hipStream_t s);
// A map from photons already in m.a / m.b / m.c (m.n of them): sort and index.
void photon_map_build(PhotonMap &m, float cell, hipStream_t s);
// PhotonMap::gather at `count` device points and normals (3 floats each).
void launch_photon_gather(const PhotonMap &m, int count, const float *points, const float *normals, float radius,
                          float *out, hipStream_t s);
// Pass 2 of renderPhotonMapped with rn_render's camera.
void launch_photon_render(const DevScene &sc, const CompatDev &cd, const PhotonMap &m, int W, int H, float cam_z,
                          int gather_rays, uint32_t key0, uint32_t key1, unsigned char *rgba, hipStream_t s);
// renderBidirectional with rn_render's camera.
void launch_bdpt_render(const DevScene &sc, const CompatDev &cd, int W, int H, float cam_z, int samples,
                        float exposure, uint32_t key0, uint32_t key1, unsigned char *rgba, hipStream_t s);

} // namespace rn
