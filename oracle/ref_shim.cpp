// C entry points over the reference's own C++ functions, compiled together
// with the reference sources into oracle/_ref/librednoise_ref.so (see the
// Makefile). Used only by tests/golden/make_golden.py to write golden vectors
// and by the test-suite when the library is present. The reference code is
// called, never modified or copied.
#include "BDPT.h"
#include "BVH.h"
#include "Camera.h"
#include "Geometry.h"
#include "Noise.h"
#include "ObjLoader.h"
#include "Photon.h"
#include "Renderer.h"
#include <Canvas.h>
#include <ModelTriangle.h>
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

static std::vector<ModelTriangle> from_verts(const float *v, int n) {
	std::vector<ModelTriangle> tris;
	tris.reserve(static_cast<size_t>(n));
	for (int i = 0; i < n; i++) {
		const float *t = v + 9 * i;
		tris.emplace_back(glm::vec3(t[0], t[1], t[2]), glm::vec3(t[3], t[4], t[5]), glm::vec3(t[6], t[7], t[8]),
		                  Colour(255, 255, 255));
	}
	return tris;
}

extern "C" {

// The smoothed vertex normals of loadOBJ (src/ObjLoader.cpp:178-205), 9 per
// triangle; returns the triangle count.
int ref_load_obj_vnormals(const char *path, float scale, float *vn) {
	std::vector<ModelTriangle> m = loadOBJ(path, scale);
	int n = static_cast<int>(m.size());
	if (vn)
		for (int i = 0; i < n; i++)
			for (int k = 0; k < 3; k++)
				for (int a = 0; a < 3; a++)
					vn[9 * i + 3 * k + a] = m[static_cast<size_t>(i)].vertexNormals[k][a];
	return n;
}

// loadOBJ (src/ObjLoader.cpp:102-214) flattened: returns the triangle count;
// fills the arrays when non-NULL (verts 9n, normals 3n, colour 3n, material n,
// has_texture n, texcoords 6n).
int ref_load_obj(const char *path, float scale, float *verts, float *normals, int *colour, int *material,
                 int *has_texture, float *texcoords) {
	std::vector<ModelTriangle> m = loadOBJ(path, scale);
	int n = static_cast<int>(m.size());
	for (int i = 0; i < n; i++) {
		const ModelTriangle &t = m[static_cast<size_t>(i)];
		for (int k = 0; k < 3; k++)
			for (int a = 0; a < 3; a++) {
				if (verts)
					verts[9 * i + 3 * k + a] = t.vertices[k][a];
			}
		for (int a = 0; a < 3; a++)
			if (normals)
				normals[3 * i + a] = t.normal[a];
		if (colour) {
			colour[3 * i] = t.colour.red;
			colour[3 * i + 1] = t.colour.green;
			colour[3 * i + 2] = t.colour.blue;
		}
		if (material)
			material[i] = static_cast<int>(t.material);
		if (has_texture)
			has_texture[i] = t.texture ? 1 : 0;
		if (texcoords)
			for (int k = 0; k < 3; k++) {
				texcoords[6 * i + 2 * k] = t.texturePoints[k].x;
				texcoords[6 * i + 2 * k + 1] = t.texturePoints[k].y;
			}
	}
	return n;
}

// BVH::intersect (src/BVH.cpp:77-111) over n triangles for nr rays.
void ref_bvh_intersect(const float *verts, int n, int nr, const float *o, const float *d, const int *ignore,
                       int *prim, float *t, float *u, float *v) {
	std::vector<ModelTriangle> tris = from_verts(verts, n);
	BVH bvh(tris);
	for (int i = 0; i < nr; i++) {
		RayTriangleIntersection h = bvh.intersect(glm::vec3(o[3 * i], o[3 * i + 1], o[3 * i + 2]),
		                                          glm::vec3(d[3 * i], d[3 * i + 1], d[3 * i + 2]),
		                                          ignore ? ignore[i] : -1);
		prim[i] = h.hit ? static_cast<int>(h.triangleIndex) : -1;
		t[i] = h.distanceFromCamera;
		u[i] = h.u;
		v[i] = h.v;
	}
}

// BVH::occluded (src/BVH.cpp:113-139).
void ref_bvh_occluded(const float *verts, int n, int nr, const float *o, const float *d, const float *maxd,
                      const int *ignore, int *occ) {
	std::vector<ModelTriangle> tris = from_verts(verts, n);
	BVH bvh(tris);
	for (int i = 0; i < nr; i++)
		occ[i] = bvh.occluded(glm::vec3(o[3 * i], o[3 * i + 1], o[3 * i + 2]),
		                      glm::vec3(d[3 * i], d[3 * i + 1], d[3 * i + 2]), maxd[i], ignore ? ignore[i] : -1)
		             ? 1
		             : 0;
}

// getClosestIntersection (src/Geometry.cpp:45-63).
void ref_brute_closest(const float *verts, int n, int nr, const float *o, const float *d, int *prim, float *t,
                       float *u, float *v) {
	std::vector<ModelTriangle> tris = from_verts(verts, n);
	for (int i = 0; i < nr; i++) {
		RayTriangleIntersection h = getClosestIntersection(glm::vec3(o[3 * i], o[3 * i + 1], o[3 * i + 2]),
		                                                   glm::vec3(d[3 * i], d[3 * i + 1], d[3 * i + 2]), tris);
		prim[i] = h.hit ? static_cast<int>(h.triangleIndex) : -1;
		t[i] = h.distanceFromCamera;
		u[i] = h.u;
		v[i] = h.v;
	}
}

// intersectTriangle (src/Geometry.cpp:11-25) on one triangle per ray.
void ref_intersect_triangle(int nr, const float *o, const float *d, const float *tri9, int *hit, float *t, float *u,
                            float *v) {
	for (int i = 0; i < nr; i++) {
		const float *q = tri9 + 9 * i;
		ModelTriangle tri(glm::vec3(q[0], q[1], q[2]), glm::vec3(q[3], q[4], q[5]), glm::vec3(q[6], q[7], q[8]),
		                  Colour());
		hit[i] = intersectTriangle(glm::vec3(o[3 * i], o[3 * i + 1], o[3 * i + 2]),
		                           glm::vec3(d[3 * i], d[3 * i + 1], d[3 * i + 2]), tri, t[i], u[i], v[i])
		             ? 1
		             : 0;
	}
}

// rn_render's camera (src/capi.cpp:49-50) primary ray through image-plane
// offsets (src/Renderer.cpp:1148) with jitter (jx, jy).
void ref_primary_rays(int w, int h, float cam_z, const float *jitter, float *o, float *d) {
	Camera cam(w, h, 2.0f, glm::vec3(0.0f, 0.0f, cam_z));
	cam.lookAt(glm::vec3(0.0f));
	float f = cam.focalLength * cam.scale;
	for (int y = 0; y < h; y++)
		for (int x = 0; x < w; x++) {
			size_t i = static_cast<size_t>(y) * static_cast<size_t>(w) + static_cast<size_t>(x);
			float jx = jitter ? jitter[2 * i] : 0.0f, jy = jitter ? jitter[2 * i + 1] : 0.0f;
			glm::vec3 ro, rd;
			cam.primaryRay(x + 0.5f + jx - w / 2.0f, -(y + 0.5f + jy - h / 2.0f), f, ro, rd);
			for (int a = 0; a < 3; a++) {
				o[3 * i + a] = ro[a];
				d[3 * i + a] = rd[a];
			}
		}
}

// fractalNoise (src/Noise.cpp:246-255).
void ref_fractal_noise(int np, const float *p, int octaves, float *out) {
	for (int i = 0; i < np; i++)
		out[i] = fractalNoise(glm::vec3(p[3 * i], p[3 * i + 1], p[3 * i + 2]), octaves);
}

// Canvas 0xAARRGGBB -> RGBA8 as rn_render unpacks it (src/capi.cpp:68-76).
static void unpack(const Canvas &canvas, int count, unsigned char *rgba) {
	for (int i = 0; i < count; i++) {
		uint32_t p = canvas.pixels[static_cast<size_t>(i)];
		rgba[i * 4 + 0] = static_cast<unsigned char>((p >> 16) & 0xFF);
		rgba[i * 4 + 1] = static_cast<unsigned char>((p >> 8) & 0xFF);
		rgba[i * 4 + 2] = static_cast<unsigned char>(p & 0xFF);
		rgba[i * 4 + 3] = static_cast<unsigned char>((p >> 24) & 0xFF);
	}
}

// renderPhotonMapped (src/Renderer.cpp:1297-1346) of an OBJ with rn_render's
// camera (src/capi.cpp:49-50) and the default light.
int ref_render_photon(const char *path, float scale, int w, int h, float cam_z, int num_photons, int gather_rays,
                      unsigned char *rgba) {
	std::vector<ModelTriangle> model = loadOBJ(path, scale);
	if (model.empty())
		return 0;
	Camera camera(w, h, 2.0f, glm::vec3(0.0f, 0.0f, cam_z));
	camera.lookAt(glm::vec3(0.0f));
	Canvas canvas(w, h);
	renderPhotonMapped(model, camera, canvas, num_photons, {}, {}, gather_rays);
	unpack(canvas, w * h, rgba);
	return 1;
}

// renderBidirectional (src/BDPT.cpp:66-144) of an OBJ with rn_render's camera.
int ref_render_bidirectional(const char *path, float scale, int w, int h, float cam_z, int samples, float exposure,
                             unsigned char *rgba) {
	std::vector<ModelTriangle> model = loadOBJ(path, scale);
	if (model.empty())
		return 0;
	Camera camera(w, h, 2.0f, glm::vec3(0.0f, 0.0f, cam_z));
	camera.lookAt(glm::vec3(0.0f));
	Canvas canvas(w, h);
	renderBidirectional(model, camera, canvas, samples, exposure);
	unpack(canvas, w * h, rgba);
	return 1;
}

// PhotonMap::build(cell) then gather(point, normal, radius) at m points
// (src/Photon.cpp:10-46).
void ref_photon_gather(int n, const float *pos, const float *power, const float *dir, float cell, int m,
                       const float *points, const float *normals, float radius, float *out) {
	PhotonMap map;
	map.photons.resize(static_cast<size_t>(n));
	for (int i = 0; i < n; i++)
		for (int a = 0; a < 3; a++) {
			map.photons[static_cast<size_t>(i)].position[a] = pos[3 * i + a];
			map.photons[static_cast<size_t>(i)].power[a] = power[3 * i + a];
			map.photons[static_cast<size_t>(i)].direction[a] = dir[3 * i + a];
		}
	map.build(cell);
	for (int i = 0; i < m; i++) {
		glm::vec3 g = map.gather(glm::vec3(points[3 * i], points[3 * i + 1], points[3 * i + 2]),
		                         glm::vec3(normals[3 * i], normals[3 * i + 1], normals[3 * i + 2]), radius);
		for (int a = 0; a < 3; a++)
			out[3 * i + a] = g[a];
	}
}

} // extern "C"
