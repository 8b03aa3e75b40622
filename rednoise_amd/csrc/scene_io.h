// Host scene representation and file formats.
#pragma once

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

namespace rn {

struct Texture {
	int width = 0, height = 0;
	std::vector<uint32_t> argb;
};

// Everything the two estimators need, in original triangle order.
struct HostScene {
	int n = 0;
	std::vector<float> verts;     // 9n
	std::vector<float> albedo;    // 3n, physical estimator
	std::vector<float> emission;  // 3n, physical estimator
	std::vector<float> normals;   // 3n face normals (reference estimator)
	std::vector<float> vnormals;  // 9n smoothed vertex normals (Whitted shading)
	std::vector<int> colour;      // 3n Kd ints
	std::vector<int> material;    // n
	std::vector<int> texture_id;  // n
	std::vector<float> texcoords; // 6n
	std::vector<Texture> textures;
};

bool load_obj(const std::string &path, float scale, HostScene &sc, std::string &err);
bool load_ppm(const std::string &path, Texture &tex, std::string &err);
void face_normal(const float *tri9, float *out3);
// Vertex normals of a scene without shared-vertex information: each vertex
// gets its face's normal.
void flat_vertex_normals(HostScene &sc);
bool write_png(const std::string &path, int w, int h, const unsigned char *rgba);

// Binary scene cache (raw arrays + checksum); see scene_io.cpp for the layout.
bool save_scene(const std::string &path, const HostScene &sc, std::string &err);
bool load_scene(const std::string &path, HostScene &sc, std::string &err);

} // namespace rn
