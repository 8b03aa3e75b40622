/* rednoise (MI355X) -- the stable C ABI of the path-trace render call.
 *
 * Prototypes, enum values and version macros are identical to the reference
 * header include/rednoise/rednoise.h:16-48 so that existing callers (the ctypes
 * binding bindings/python/rednoise/__init__.py:197-234, the Rust FFI
 * bindings/rust/src/lib.rs:34-64, examples/c_consumer.c) link unchanged.
 * Behaviour differences are listed per function.
 *
 * Every entry point is exception-free and returns 0 / NULL on failure; the
 * reason is available from rn_last_error() (declared in rednoise_ex.h).
 */
#ifndef REDNOISE_H
#define REDNOISE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Reference: include/rednoise/rednoise.h:16-18. */
#define RN_VERSION_MAJOR 0
#define RN_VERSION_MINOR 1
#define RN_VERSION_PATCH 2

/* Opaque scene: host triangle/material arrays plus lazily built device copies
 * (LBVH/BVH8, SoA triangle buffers). Reference: include/rednoise/rednoise.h:21. */
typedef struct rn_scene rn_scene;

/* Reference: include/rednoise/rednoise.h:24. RN_PATHTRACED and RN_RAYTRACED
 * are implemented by this library; the other modes make rn_render return 0
 * with an "unsupported mode" error. */
typedef enum { RN_WIREFRAME = 0, RN_RASTERISED = 1, RN_RAYTRACED = 2, RN_PATHTRACED = 3 } rn_render_mode;

/* Load a Wavefront OBJ (+ .mtl, + P6 PPM map_Kd textures), scaling vertices by
 * `scale`. NULL on a NULL path, an unreadable file or an empty model.
 * Replaces src/capi.cpp:22-34 (loader semantics of src/ObjLoader.cpp:102-214). */
rn_scene *rn_scene_load_obj(const char *path, float scale);

/* Release a scene and every device buffer it owns; NULL is ignored.
 * Replaces src/capi.cpp:36-38. */
void rn_scene_free(rn_scene *scene);

/* Triangle count, 0 for NULL. Replaces src/capi.cpp:40-42. */
int rn_scene_triangle_count(const rn_scene *scene);

/* Render into the caller's width*height*4 RGBA8 buffer with the fixed camera at
 * (0,0,cam_z) looking at the origin (focal length 300 px). RN_PATHTRACED runs the
 * reference estimator (src/Renderer.cpp:820-907,1119-1170) on the GPU with the
 * reference's own random streams (one std::mt19937 per pixel, seeded
 * (y * W + x) * 9781 + 1), so it computes the reference's image; RN_RAYTRACED
 * runs the Whitted renderer (src/Renderer.cpp:909-948). Synchronous. Returns 1
 * on success, 0 on a NULL scene/buffer, width/height <= 0, samples <= 0 for
 * RN_PATHTRACED, an unsupported mode or a device error.
 * Replaces src/capi.cpp:44-79. */
int rn_render(const rn_scene *scene, rn_render_mode mode, int width, int height, float cam_z, int samples,
              unsigned char *rgba);

/* Write an RGBA8 buffer as a PNG (stored/uncompressed deflate). 1 on success.
 * Replaces src/capi.cpp:81-85. */
int rn_save_png(const char *path, int width, int height, const unsigned char *rgba);

/* Library version string, "0.1.2". Replaces src/capi.cpp:87-89. */
const char *rn_version(void);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* REDNOISE_H */
