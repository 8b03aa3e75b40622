/* CPU oracle for the rednoise MI355X path tracer -- TEST INFRASTRUCTURE ONLY.
 *
 * This library restates, in plain C, the reference algorithms of the path-trace
 * render call (ReverseZoom2151/rednoise: src/capi.cpp, src/Renderer.cpp,
 * src/BVH.cpp, src/Geometry.cpp, src/Camera.cpp) and specifies the algorithms
 * the reference does not have (LBVH stages, Philox4x32-10, the physical
 * estimator and its adjoint). It is the checker: only the test-suite, the
 * smoke check and the CPU-baseline leg of the benchmark load it. The product
 * (librednoise.so) never links or calls it.
 *
 * Every function cites the reference file:line it follows, or names the
 * closed form / published algorithm that pins it. Pinning status per function
 * is recorded next to its declaration:
 *   [pinned: ...]  what checks it independently of this file.
 */
#ifndef RN_ORACLE_H
#define RN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- Philox -- */

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
 * [pinned: Random123 known-answer vectors; rocRAND's independent header.] */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void or_philox_batch(int n, const uint32_t *ctr, const uint32_t *key, uint32_t *out);

/* ------------------------------------------------- reference compat (FP32) -- */

/* src/Geometry.cpp:11-25 with glm's exact FP32 operation order
 * (third_party/glm/detail/func_matrix.inl:253-262,322-344,
 *  type_mat3x3.inl:468-474). Returns 1 on a hit.
 * [pinned: tests/tests.cpp:150-165 closed-form hits; :324-329 Cramer agreement;
 *  bit-exact against BVH::intersect goldens from the compiled reference.] */
int or_compat_intersect_tri(const float o[3], const float d[3], const float v0[3], const float v1[3],
                            const float v2[3], float *t, float *u, float *v);

/* Camera of rn_render (src/capi.cpp:49-50, src/Camera.cpp:27-33,69-79):
 * pixel-centre (jitter jx,jy) primary ray for a camera at (0,0,cam_z) looking
 * at the origin, f = 300 px.
 * [pinned: tests/tests.cpp:113-148 projection/orthonormality; goldens.] */
void or_compat_primary_ray(int x, int y, int W, int H, float cam_z, float jx, float jy, float o[3], float d[3]);
/* The same camera for a screen offset (sx, sy) from the image centre, pixels. */
void or_compat_camera_ray(float sx, float sy, float cam_z, float o[3], float d[3]);

/* Brute-force closest hit (src/Geometry.cpp:45-63): first minimum in index
 * order, strict <. verts = 9n floats. Returns prim or -1. */
int or_compat_closest_brute(const float *verts, int n, const float o[3], const float d[3], int ignore, float *t,
                            float *u, float *v);

/* Median-split binary BVH (src/BVH.cpp:6-63) over 9n verts; opaque handle. */
void *or_compat_bvh_build(const float *verts, int n);
void or_compat_bvh_free(void *bvh);
/* Closest hit (src/BVH.cpp:77-111) and any hit (:113-139).
 * [pinned: brute force agreement, the reference's own testBVH method.] */
int or_compat_bvh_intersect(const void *bvh, const float o[3], const float d[3], int ignore, float *t, float *u,
                            float *v);
int or_compat_bvh_occluded(const void *bvh, const float o[3], const float d[3], float maxd, int ignore);
/* Batch form of or_compat_bvh_intersect. */
void or_compat_bvh_intersect_batch(const void *bvh, int nrays, const float *o, const float *d, const int *ignore,
                                   int *prim, float *t, float *u, float *v);

/* Nearest texel (src/Renderer.cpp:584-591 + 642-656): returns ty*tw+tx. */
int or_compat_texel(float u, float v, const float tp[6], int tw, int th);

/* Perlin improved noise and 4-octave fBm (src/Noise.cpp:7-61,246-255). */
float or_perlin(float x, float y, float z);
float or_fractal_noise(const float p[3], int octaves);

/* Reference estimator ("Mode R", src/Renderer.cpp:778-907,1119-1170) with the
 * per-pixel std::mt19937 replaced by Philox4x32-10 draws (layout documented in
 * compat.c). Scene arrays as exported by rn_scene_get_compat. Writes RGBA8.
 * [pinned statistically: 3 standard errors per 16x16 tile against images of
 *  the compiled reference renderPathTraced.] */
typedef struct {
	int n;
	const float *verts;     /* 9n */
	const float *normals;   /* 3n face normals */
	const int *colour;      /* 3n ints 0..255 */
	const int *material;    /* n */
	const int *texture_id;  /* n, -1 none */
	const float *texcoords; /* 6n */
	int n_textures;
	const int *tex_w, *tex_h;    /* n_textures */
	const uint32_t *const *tex;  /* n_textures pointers to ARGB pixels */
} or_compat_scene;

void or_compat_render(const or_compat_scene *scene, int W, int H, float cam_z, int spp, uint32_t key0,
                      uint32_t key1, int y0, int y1, unsigned char *rgba);
/* The same estimator with the reference's own per-pixel std::mt19937 streams
 * and pathTrace's recursive colour sums: the reference's image.
 * [pinned: pixel for pixel against renderPathTraced of the compiled
 *  reference (tests/golden/modeR_cornell.npz).] */
void or_compat_render_mt(const or_compat_scene *scene, int W, int H, float cam_z, int spp, int y0, int y1,
                         unsigned char *rgba);

/* Photon mapping and bidirectional path tracing (photon.c). `rng` selects
 * the reference's std::mt19937 streams (OR_RNG_MT; the images then equal the
 * compiled reference's) or the GPU's Philox layout (OR_RNG_PHILOX).
 * [pinned: OR_RNG_MT images against renderPhotonMapped / renderBidirectional
 *  of the compiled reference, pixel for pixel; or_photon_gather against
 *  PhotonMap::gather.] */
enum { OR_RNG_MT = 0, OR_RNG_PHILOX = 1 };
/* Pass 1 of renderPhotonMapped: returns the photon count (arrays 3 floats
 * per photon, at most `cap` written). */
int or_photon_trace(const or_compat_scene *scene, int num_photons, int rng, uint32_t key0, uint32_t key1,
                    float *pos, float *power, float *dir, int cap);
/* PhotonMap::build(cell) + gather(point, normal, radius) at m points. */
void or_photon_gather(int n, const float *pos, const float *power, const float *dir, float cell, int m,
                      const float *points, const float *normals, float radius, float *out);
/* Pass 2 of renderPhotonMapped with rn_render's camera over rows [y0, y1). */
void or_photon_render(const or_compat_scene *scene, int n, const float *pos, const float *power, const float *dir,
                      int W, int H, float cam_z, int gather_rays, int rng, uint32_t key0, uint32_t key1, int y0,
                      int y1, unsigned char *rgba);
/* renderBidirectional with rn_render's camera over rows [y0, y1). */
void or_bdpt_render(const or_compat_scene *scene, int W, int H, float cam_z, int samples, float exposure, int rng,
                    uint32_t key0, uint32_t key1, int y0, int y1, unsigned char *rgba);

/* ------------------------------------------------------------------ LBVH -- */

/* 30-bit Morton code of the triangle centroids normalised to the centroid
 * bounds, 10 bits per axis, x at bit 3i, y at 3i+1, z at 3i+2
 * (src/Morton.cpp:37-44,68-70 bit order).
 * [pinned: tests/tests.cpp:422-432 round trip; known codes.] */
uint32_t or_morton3(uint32_t x, uint32_t y, uint32_t z);
void or_morton_codes(const float *verts, int n, uint32_t *codes);

/* Binary radix tree of sorted codes built top-down by splitting at the highest
 * differing bit (index-augmented keys for duplicates), numbered with the
 * Karras 2012 convention so it must equal the GPU's bottom-up construction.
 * left/right: >= 0 internal node id, < 0 leaf ~sorted position.
 * [pinned: structural invariants + brute-force range checks in tests.] */
void or_radix_tree(const uint32_t *sorted_codes, int n, int *left, int *right);

/* PLOC binary tree over the sorted order (order[p] = triangle at sorted
 * position p), search radius 8, same numbering as or_radix_tree's output
 * arrays (internal ids 0 .. n-2 with the root at 0, leaves ~position).
 * Returns the number of clustering rounds (-1: a round merged nothing).
 * [pinned: hand-worked trees (two pairs, a caterpillar and its mirror,
 *  equal gaps, duplicates); structural invariants; brute-force mutual-pair
 *  check of the first round.] */
int or_ploc_tree(const float *verts, const int *order, int n, int *left, int *right);

/* -------------------------------------------------- physical estimator -- */

/* Physical estimator ("Mode P"): Lambertian albedo/pi, NEE to emissive
 * triangles chosen in proportion to area, cosine bounce, Russian roulette,
 * black background. Sampling layout identical to the GPU kernels (phys_impl.h).
 * [pinned: white-furnace closed form L = Le (1-a^(D+1))/(1-a); Lambert
 *  polygon irradiance for one bounce; FP32 against FP64.] */
typedef struct {
	float eye[3];
	float du[3]; /* right * tan(vfov/2) * aspect */
	float dv[3]; /* up * tan(vfov/2) */
	float w[3];  /* forward */
} or_cam;

typedef struct {
	int width, height, spp, max_depth, rr_depth;
	float rr_prob;
	uint32_t seed, scene_id;
	float ray_eps;
} or_params;

typedef struct {
	uint64_t closest, shadow, paths;
} or_counts;

void or_camera_basis(const float eye[3], const float target[3], const float up[3], float vfov_deg, int W, int H,
                     or_cam *cam);

/* Opaque prepared scene (BVH in the given precision). */
void *or_phys_prepare_f32(const float *verts, int n, const float *albedo, const float *emission);
void *or_phys_prepare_f64(const float *verts, int n, const float *albedo, const float *emission);
void or_phys_free_f32(void *s);
void or_phys_free_f64(void *s);
/* Render pixels [p0,p1) (row-major index) into out (3 per pixel, indexed by
 * absolute pixel). OpenMP over pixels when nthreads > 1. */
void or_phys_render_f32(const void *s, const or_cam *cam, const or_params *p, int p0, int p1, int nthreads,
                        float *out, or_counts *counts);
void or_phys_render_f64(const void *s, const or_cam *cam, const or_params *p, int p0, int p1, int nthreads,
                        double *out, or_counts *counts);
/* Render an explicit pixel list. */
void or_phys_render_list_f64(const void *s, const or_cam *cam, const or_params *p, int npix, const int *pix,
                             int nthreads, double *out, or_counts *counts);
void or_phys_render_list_f32(const void *s, const or_cam *cam, const or_params *p, int npix, const int *pix,
                             int nthreads, float *out, or_counts *counts);
/* FP64 path-replay adjoint: dL/dI (3 per pixel) -> dL/d albedo, dL/d emission
 * (3n each, accumulated). [pinned: central finite differences of
 * or_phys_render_f64 with common random numbers.] */
void or_phys_adjoint_f64(const void *s, const or_cam *cam, const or_params *p, const float *dl_dimage, int nthreads,
                         double *g_albedo, double *g_emission);
/* Closest hit in the given precision (brute force when brute != 0). */
int or_phys_all_hits_f32(const void *s, const float o[3], const float d[3], float tmax, int *prims, float *ts,
                         int cap);
int or_phys_all_hits_f64(const void *s, const double o[3], const double d[3], double tmax, int *prims, double *ts,
                         int cap);
/* Every triangle the reference test (src/Geometry.cpp:11-25) accepts on
 * (o, d), by brute force: ids and t (at most cap each); returns the count. */
int or_compat_all_hits(const float *verts, int n, const float o[3], const float d[3], int *prims, float *ts, int cap);
int or_phys_closest_f64(const void *s, const double o[3], const double d[3], double tmax, int brute, double *t,
                        double *u, double *v);
int or_phys_closest_f32(const void *s, const float o[3], const float d[3], float tmax, int brute, float *t, float *u,
                        float *v);

#ifdef __cplusplus
}
#endif
#endif
