/* rednoise (MI355X) -- extended C ABI.
 *
 * The reference ABI (rednoise.h) can only express its own fixed estimator and
 * camera (src/capi.cpp:44-79: f = 300 px, depth 4, no Russian roulette). The
 * entry points below carry what the physical workloads need: arrays-in scenes,
 * an explicit camera, an FP32 HDR output, the path-replay adjoint, batched
 * multi-scene rendering and raw ray queries. None of them has a counterpart in
 * the reference ABI; each says which reference function it generalises.
 *
 * All pointers are host pointers unless the function name ends in _device, in
 * which case they are device (HBM) pointers and `stream` is a hipStream_t
 * (NULL = the legacy default stream). Functions return 1 on success and 0 on
 * failure, with the reason in rn_last_error().
 */
#ifndef REDNOISE_EX_H
#define REDNOISE_EX_H

#include "rednoise.h"
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Thread-local description of the last failure on the calling thread ("" if
 * none). New: the reference ABI has return codes only (src/capi.cpp:46-47). */
const char *rn_last_error(void);

/* ---------------------------------------------------------------- scenes -- */

/* Scene from plain arrays: verts = 9*n floats (v0,v1,v2 per triangle),
 * albedo = 3*n (linear, 0..1), emission = 3*n radiance or NULL (no emitters).
 * Triangles with any emission component > 0 form the light set; the light set
 * is fixed at creation. Generalises rn_scene_load_obj (src/capi.cpp:22-34).
 * Every vertex coordinate must be finite: a scene holding a NaN or infinite
 * one (from any constructor) fails its first build or render with an error
 * naming the triangle. */
rn_scene *rn_scene_from_arrays(const float *verts, int n, const float *albedo, const float *emission);

/* Replace per-triangle albedo / emission (3*n each; either may be NULL). The
 * light set does not change. */
int rn_scene_set_materials(rn_scene *scene, const float *albedo, const float *emission);

/* Binary scene cache. rn_scene_save writes everything a scene holds
 * (geometry, both material sets, the reference estimator's attributes and
 * textures) as raw arrays with a checksum; rn_scene_load reads it back, so a
 * scene parsed once from OBJ (rn_scene_load_obj, src/capi.cpp:22-34 and
 * src/ObjLoader.cpp:102-214 in the reference) reloads at file-read speed.
 * A loaded scene renders exactly as the saved one. No reference counterpart:
 * it replaces re-parsing the OBJ on every start. Save returns 1/0; load
 * returns NULL on a missing, truncated, corrupt or foreign file, with the
 * reason in rn_last_error(). */
int rn_scene_save(const rn_scene *scene, const char *path);
rn_scene *rn_scene_load(const char *path);

/* Copy the scene's triangle data out: verts 9*n, albedo 3*n, emission 3*n
 * (any may be NULL). For an OBJ scene albedo = Kd colour int / 255. */
int rn_scene_get_triangles(const rn_scene *scene, float *verts, float *albedo, float *emission);

/* Per-triangle attributes of the reference estimator, as loadOBJ builds them
 * (src/ObjLoader.cpp:169-211): face normal (3*n), colour ints (3*n), material
 * id (n, enum order of framework/ModelTriangle.h:12-24), texture id (n, -1 =
 * none), texture coordinates (6*n). Any pointer may be NULL. */
int rn_scene_get_compat(const rn_scene *scene, float *normals, int *colour, int *material, int *texture_id,
                        float *texcoords);

/* Texture `id` of an OBJ scene: writes width/height, and if argb != NULL the
 * width*height 0xAARRGGBB pixels (framework/TextureMap.cpp:4-31). */
/* Smoothed vertex normals, 9 floats per triangle (loadOBJ's vertexNormals,
 * src/ObjLoader.cpp:178-205; a scene from arrays gets its face normals). */
int rn_scene_get_vertex_normals(const rn_scene *scene, float *vnormals);
int rn_scene_get_texture(const rn_scene *scene, int id, int *width, int *height, uint32_t *argb);

/* ------------------------------------------------------------ rendering -- */

/* rn_render(scene, RN_PATHTRACED, ...) with counter-based Philox4x32-10
 * streams under the 64-bit key (key0, key1) in place of the reference's
 * per-pixel std::mt19937 (which rn_render keeps): the same estimator, and
 * independent keys give independent renders, from which per-tile standard
 * errors are estimated. Generalises src/capi.cpp:44-79. */
int rn_render_seeded(const rn_scene *scene, int width, int height, float cam_z, int samples, uint32_t key0,
                     uint32_t key1, unsigned char *rgba);

/* Pinhole camera: eye, look-at target, up vector, vertical field of view. */
typedef struct {
	float eye[3];
	float target[3];
	float up[3];
	float vfov_deg;
} rn_camera;

enum { RN_INTEGRATOR_MEGAKERNEL = 0, RN_INTEGRATOR_WAVEFRONT = 1 };

typedef struct {
	int width, height;
	int spp;          /* samples per pixel, >= 1 */
	int max_depth;    /* path segments traced per sample, >= 1 */
	int rr_depth;     /* Russian roulette at vertices with depth >= rr_depth (>= max_depth: off) */
	float rr_prob;    /* > 0: fixed survival probability; <= 0: min(0.95, max throughput) */
	uint32_t seed;    /* Philox4x32-10 key word 0 */
	float ray_eps;    /* origin offset along the normal; <= 0: 1e-4 * scene bounding radius */
	int integrator;   /* RN_INTEGRATOR_* */
	int queue_paths;  /* wavefront paths in flight; <= 0: chosen by the library */
	int block_size;   /* threads per block for the trace kernels; <= 0: default */
	int bvh_width;    /* 4 or 8 children per node; <= 0: 8 */
	int flags;        /* RN_FLAG_* */
} rn_render_params;

/* rn_render_params.flags */
enum {
	RN_FLAG_COUNT_TRAVERSAL = 1, /* megakernel: also count BVH nodes / triangles per ray type (slower) */
	RN_FLAG_TIME_STAGES = 2,     /* wavefront: time every stage launch with HIP events */
	RN_FLAG_USE_GRAPH = 4,       /* wavefront: replay the iteration from a captured hipGraph */
	RN_FLAG_FB_ATOMICS = 8       /* wavefront: add contributions to the frame buffer with float atomics
	                              * (arbitrary order) instead of per-path radiance slots summed in sample
	                              * order; slots are also left out when the render's paths number 2^32 or
	                              * more or their 12 bytes each exceed a third of the free device memory,
	                              * and by rn_render_ex_device_noalloc */
};

/* Physical ("Mode P") estimator: Lambertian albedo/pi, next-event estimation to
 * the emissive triangles (area-proportional choice), Russian roulette, black
 * background. Writes the per-pixel mean radiance, width*height*3 FP32 (row 0 =
 * top). Generalises renderPathTraced (src/Renderer.cpp:1119-1170), whose NEE
 * structure follows src/Metropolis.cpp:53-81. */
int rn_render_ex(const rn_scene *scene, const rn_camera *camera, const rn_render_params *params, float *hdr_rgb);

/* As rn_render_ex with device pointers for albedo/emission (3*n, NULL = the
 * scene's own) and the output image; enqueued on `stream`, not synchronised. */
int rn_render_ex_device(const rn_scene *scene, const rn_camera *camera, const rn_render_params *params,
                        const float *d_albedo, const float *d_emission, float *d_hdr_rgb, void *stream);

/* Allocation-free form of rn_render_ex_device. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (serving
 * loops, harnesses that forbid hipMalloc, and -- with the megakernel, which
 * only enqueues work -- code under HIP graph capture). It calls no hipMalloc,
 * hipFree, stream or event creation: all of its work memory (counters and the
 * wavefront's queues, control block and scene table) lives in one device
 * array of about 86 MB per GPU that the library reserves when it is loaded
 * and holds until the process ends.
 *   - The scene must already hold its device copy and a BVH of
 *     params->bvh_width (call rn_scene_build first); otherwise it returns 0.
 *   - RN_INTEGRATOR_MEGAKERNEL enqueues on `stream` and returns.
 *     RN_INTEGRATOR_WAVEFRONT keeps at most RN_NOALLOC_MAX_QUEUE_PATHS paths
 *     in flight (any image size; larger frames refill the queue) and
 *     synchronises `stream` before it returns; queue_paths above that
 *     maximum, RN_FLAG_USE_GRAPH and RN_FLAG_TIME_STAGES are rejected.
 *   - Every call on one GPU shares that GPU's array: calls whose GPU work
 *     overlaps (on different streams, from different threads) race on it.
 * The image equals rn_render_ex_device's (same Philox streams). */
#define RN_NOALLOC_MAX_QUEUE_PATHS 524288
int rn_render_ex_device_noalloc(const rn_scene *scene, const rn_camera *camera, const rn_render_params *params,
                                const float *d_albedo, const float *d_emission, float *d_hdr_rgb, void *stream);

/* Path-replay adjoint of rn_render_ex: given dL/dI (width*height*3, the loss
 * gradient with respect to each pixel's mean radiance) writes dL/d albedo and
 * dL/d emission (3*n each; either may be NULL). Re-traces every path from its
 * Philox counter; no per-path state is kept between the forward pass and this
 * call. Uses the megakernel sampling order regardless of params->integrator. */
int rn_adjoint(const rn_scene *scene, const rn_camera *camera, const rn_render_params *params,
               const float *dl_dimage, float *dl_dalbedo, float *dl_demission);

/* Device-pointer form of rn_adjoint; gradients are accumulated into
 * d_dl_dalbedo / d_dl_demission (the caller zeroes them). */
int rn_adjoint_device(const rn_scene *scene, const rn_camera *camera, const rn_render_params *params,
                      const float *d_albedo, const float *d_emission, const float *d_dl_dimage,
                      float *d_dl_dalbedo, float *d_dl_demission, void *stream);

/* One scene of a batch. */
typedef struct {
	const rn_scene *scene;
	rn_camera camera;
	rn_render_params params; /* integrator/queue/block fields are taken from the batch options */
	float *hdr_rgb;          /* host output, width*height*3 */
} rn_batch_item;

typedef struct {
	int queue_paths;  /* paths in flight shared by all scenes; <= 0: autotuned/default */
	int block_size;   /* <= 0: autotuned/default */
	int bvh_width;    /* 4 or 8; <= 0: 8 */
	int use_graph;    /* capture the wavefront iteration in a hipGraph */
	int iters_per_graph; /* wavefront iterations per graph launch; <= 0: 8 */
} rn_batch_options;

/* Render many scenes of uneven size in one set of persistent wavefront
 * launches that share one path queue. Synchronous. */
int rn_render_batch(const rn_batch_item *items, int count, const rn_batch_options *opts);

/* Per-call measurements of the last rn_render_ex / rn_render_batch /
 * rn_adjoint on the calling thread. */
typedef struct {
	double build_ms;        /* LBVH + BVH8 build (0 if cached) */
	double render_ms;       /* integrator kernels, excluding the build */
	double total_ms;        /* wall time of the call */
	uint64_t rays_closest;  /* closest-hit queries issued */
	uint64_t rays_shadow;   /* any-hit queries issued */
	uint64_t paths;         /* camera paths traced */
	double trace_kernel_ms; /* summed duration of the dominant trace kernel */
	uint64_t trace_kernel_launches; /* wavefront: iterations enqueued, including any that found the queues empty */
	int bvh_nodes;
	int bvh_depth;
	uint64_t nodes_closest, tris_closest; /* with RN_FLAG_COUNT_TRAVERSAL */
	uint64_t nodes_shadow, tris_shadow;
	double extend_ms, shade_ms, shadow_ms; /* wavefront stage times (RN_FLAG_TIME_STAGES): extend_ms is the
	                                        * trace launches (closest hits, then the previous iteration's any-hit
	                                        * rays, in one launch), shade_ms generate + shade + the one-lane
	                                        * queue advance; shadow_ms is 0 */
	uint64_t iterations;                  /* wavefront iterations */
} rn_stats;

int rn_get_last_stats(rn_stats *out);

/* ---------------------------------------------------------- ray queries -- */

enum { RN_QUERY_CLOSEST = 0, RN_QUERY_ANY = 1 };
enum { RN_TRI_WATERTIGHT = 0, RN_TRI_COMPAT = 1 };

/* Trace n rays (origins 3n, directions 3n, tmax n or NULL = +inf) against the
 * scene's BVH8 on the GPU. `ignore` (n or NULL) skips one triangle per ray
 * (src/BVH.cpp:77-111 ignoreIndex). Closest: prim (-1 = miss), t, u (weight of
 * v1), v (weight of v2). Any: prim = 1 if occluded else 0. RN_TRI_COMPAT uses
 * the reference triangle test (src/Geometry.cpp:11-25) in its exact FP32
 * operation order; RN_TRI_WATERTIGHT the watertight test. */
int rn_trace_rays(const rn_scene *scene, int n, const float *orig, const float *dir, const float *tmax,
                  const int *ignore, int query, int tri_test, int *prim, float *t, float *u, float *v);

/* Primary rays of rn_render's camera (pixel centres, no jitter,
 * src/Camera.cpp:69-79 with f = 300) traced with the reference triangle test:
 * per pixel prim, t, u, v and the nearest texel index ty*tw+tx of a textured
 * hit (-1 if untextured), as selected by src/Renderer.cpp:584-591,642-656. */
int rn_trace_primary_compat(const rn_scene *scene, int width, int height, float cam_z, int *prim, float *t,
                            float *u, float *v, int *texel);

/* -------------------------------------------------------- build / debug -- */

typedef struct {
	int nodes;        /* BVH8 nodes */
	int depth;        /* levels, root = 1 */
	int leaf_tris;    /* triangles referenced by leaves (== n) */
	int width;        /* max children per node */
	double build_ms;
	int builder;      /* RN_BVH_PLOC or RN_BVH_LBVH */
} rn_bvh_info;

/* Binary-hierarchy builders ahead of the BVH8 collapse. The reference builds
 * its BVH on the CPU by median splits (src/BVH.cpp:18-59); both builders here give
 * the same closest hits, they differ in tree quality and build time. */
enum { RN_BVH_PLOC = 0, RN_BVH_LBVH = 1 };

/* Select the scene's builder (default RN_BVH_PLOC: parallel locally-ordered
 * clustering over the Morton order; RN_BVH_LBVH: Karras radix tree, faster to
 * build, about 1.8x more node visits per ray). A change drops the built BVH;
 * the next render or rn_scene_build rebuilds it. Returns 1, or 0 with
 * rn_last_error() set. */
int rn_scene_set_builder(rn_scene *scene, int builder);

/* Build (or rebuild) the scene's device BVH with the given width (4 or 8). */
int rn_scene_build(rn_scene *scene, int bvh_width, rn_bvh_info *info);

/* Export the device BVH: nodes as 80-byte records (nodes*80 bytes) and the
 * leaf-ordered primitive ids (n ints). */
int rn_scene_export_bvh(const rn_scene *scene, void *nodes, int *leaf_prims);

/* LBVH stages on the GPU for n triangles (verts 9n): 30-bit Morton codes in
 * sorted order, the sorted primitive order, and the binary hierarchy as
 * left/right child arrays of the n-1 internal nodes (value >= 0: internal
 * node, value < 0: leaf ~value, i.e. sorted position -value-1). */
int rn_lbvh_stages(const float *verts, int n, uint32_t *morton_sorted, int *order, int *left, int *right);

/* The default builder's stages on the GPU for n triangles: the sorted
 * primitive order and the PLOC binary hierarchy (parallel locally-ordered
 * clustering, search radius 8) in the layout of rn_lbvh_stages, internal ids
 * 0 .. n-2 with the root at 0. Exposes the tree the BVH8 collapse reads; the
 * reference builds its BVH on the CPU (src/BVH.cpp:18-59) and has no
 * counterpart of this stage. */
int rn_ploc_stages(const float *verts, int n, int *order, int *left, int *right);

/* Stable LSD radix sort of (key, value) pairs on the GPU over the low `bits`
 * key bits. */
int rn_radix_sort_pairs(uint32_t *keys, uint32_t *values, int n, int bits);

/* Philox4x32-10 on the GPU: out[4i..4i+3] = philox(ctr[4i..4i+3], key[2i..2i+1]). */
int rn_philox4x32_10(int n, const uint32_t *ctr, const uint32_t *key, uint32_t *out);

/* Progressive rendering: an accumulator holds the float64 running sum of
 * the physical estimator over frames. rn_accum_render adds params->spp
 * samples per pixel, numbered from the samples already held, so k frames of
 * s samples render the same paths as one frame of k*s (params->width and
 * height must equal the accumulator's). rn_accum_resolve_rgba8 writes the
 * running mean tone-mapped and packed in one GPU pass: c = mean * exposure,
 * c = c / (c + 1), c = c^(1/gamma) * 255, byte = min(255, (int)c), RGBA
 * order, alpha 255 -- the steps of the reference's toneMap
 * (src/Renderer.cpp:1357-1368) on the HDR mean, its OpenCL epilogue
 * (gpu/pathtracer.cl:122-132) for exposure = gamma = 1, and rn_render's
 * byte order (src/capi.cpp:69-77). rn_accum_mean writes the mean as FP32
 * (3 floats per pixel). Functions return 1 / non-NULL on success, else 0 /
 * NULL with rn_last_error() set; rn_accum_samples returns the count. */
typedef struct rn_accum rn_accum;
rn_accum *rn_accum_create(int width, int height);
void rn_accum_free(rn_accum *acc);
int rn_accum_reset(rn_accum *acc);
int rn_accum_render(rn_accum *acc, const rn_scene *scene, const rn_camera *camera, const rn_render_params *params);
int rn_accum_samples(const rn_accum *acc);
int rn_accum_resolve_rgba8(const rn_accum *acc, float exposure, float gamma, unsigned char *rgba);
int rn_accum_mean(const rn_accum *acc, float *hdr_rgb);

/* Name of the GPU the library runs on (e.g. "gfx950"), or "" if none. */
const char *rn_device_name(void);

/* Frees the device work buffers the library keeps between calls (wavefront
 * path and shadow queues, staging for host outputs). The wavefront keeps up
 * to 2^27 paths in flight by default (about 21 GB on a large GPU; see
 * rn_render_params.queue_paths); the next render allocates again. Scenes and
 * their BVHs are untouched. Returns 1, or 0 with rn_last_error() set. */
int rn_release_workspace(void);

/* ---------------------------------------------- other estimators -- */

/* Photon mapping, renderPhotonMapped (src/Renderer.cpp:1297-1346) split in
 * its two passes, with the PhotonMap of src/Photon.cpp:5-46 on the GPU.
 * The reference draws from one std::mt19937 stream across photons (seed 777)
 * and one per image row for the final gather (seed 1000 + y); here photon i
 * and pixel p own the Philox4x32-10 streams (i, 0, 'phot') and
 * (p, 0, 'gath') under the key (key0, key1). Functions return 1 / non-NULL on
 * success, else 0 / NULL with rn_last_error() set. */
typedef struct rn_photon_map rn_photon_map;

/* Pass 1 (src/Renderer.cpp:1310-1326): emit num_photons (0..2^26) from the
 * default light (src/Renderer.cpp:1303-1308), trace them through mirrors,
 * metal and glass (tracePhoton, :1177-1226) and build the map with cell side
 * 0.25 (:1301). */
rn_photon_map *rn_photon_map_trace(const rn_scene *scene, int num_photons, uint32_t key0, uint32_t key1);

/* A map over given photons (3n floats each: position, power, direction of
 * arrival), PhotonMap::photons + PhotonMap::build(cell) (src/Photon.cpp:10-19). */
rn_photon_map *rn_photon_map_create(int n, const float *position, const float *power, const float *direction,
                                    float cell);
void rn_photon_map_free(rn_photon_map *map);
/* Photons held (PhotonMap::photons.size()). */
int rn_photon_map_count(const rn_photon_map *map);
/* The photons in emission order (the reference's push order); NULL arrays
 * are skipped. */
int rn_photon_map_get(const rn_photon_map *map, float *position, float *power, float *direction);
/* PhotonMap::gather (src/Photon.cpp:21-46) at m points with their normals
 * (3m floats each): irradiance, 3m floats. */
int rn_photon_map_gather(const rn_photon_map *map, int m, const float *points, const float *normals, float radius,
                         float *irradiance);

/* Pass 2 (src/Renderer.cpp:1328-1345, photonShade :1233-1294) with rn_render's
 * camera (src/capi.cpp:49-50): direct light plus the photon irradiance at the
 * first diffuse hit through mirrors, metal and glass (depth 4), gathered
 * within the map's cell side; gather_rays > 0 (at most 4096) gathers one
 * bounce away instead (final gather). RGBA8 as rn_render writes it. */
int rn_render_photon(const rn_scene *scene, const rn_photon_map *map, int width, int height, float cam_z,
                     int gather_rays, uint32_t key0, uint32_t key1, unsigned char *rgba);

/* Bidirectional path tracing, renderBidirectional (src/BDPT.cpp:66-144), with
 * rn_render's camera: per sample a camera and a light subpath of up to 4
 * diffuse bounces each, the fixed ceiling light of src/BDPT.cpp:77-83, every
 * vertex pair connected with a shadow ray and weighted 1 / (i + j + 2), then
 * exposure and Reinhard. The reference's per-row std::mt19937 (seed 9871 + y)
 * becomes the Philox stream (pixel, sample, 'bdpt'). samples in [1, 2^20]. */
int rn_render_bidirectional(const rn_scene *scene, int width, int height, float cam_z, int samples, float exposure,
                            uint32_t key0, uint32_t key1, unsigned char *rgba);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* REDNOISE_EX_H */
