// Scene input/output on the host: the Wavefront OBJ/MTL loader of
// rn_scene_load_obj (semantics of src/ObjLoader.cpp:23-214 and
// framework/TextureMap.cpp:4-31) and the PNG writer of rn_save_png.
#include "scene_io.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>

namespace rn {

namespace {

struct MtlDef {
	int r = 0, g = 0, b = 0;
	bool has_colour = false;
	std::string map_kd;
};

// src/ObjLoader.cpp:23-48: Kd scaled by 255 and truncated to int.
std::map<std::string, MtlDef> load_mtl(const std::string &path) {
	std::map<std::string, MtlDef> out;
	std::ifstream f(path);
	if (!f.is_open())
		return out;
	std::string line, current;
	while (std::getline(f, line)) {
		std::stringstream ss(line);
		std::string type;
		ss >> type;
		if (type == "newmtl") {
			ss >> current;
		} else if (type == "Kd" && !current.empty()) {
			float r = 0, g = 0, b = 0;
			ss >> r >> g >> b;
			MtlDef &m = out[current];
			m.r = static_cast<int>(r * 255);
			m.g = static_cast<int>(g * 255);
			m.b = static_cast<int>(b * 255);
			m.has_colour = true;
		} else if (type == "map_Kd" && !current.empty()) {
			ss >> out[current].map_kd;
		}
	}
	return out;
}

// src/ObjLoader.cpp:51-73
int material_for(const std::string &n) {
	if (n == "Mirror")
		return 1;
	if (n == "Glass")
		return 2;
	if (n == "Marble")
		return 3;
	if (n == "Bumpy")
		return 4;
	if (n == "Parallax")
		return 5;
	if (n == "Metal")
		return 6;
	if (n == "Dispersive" || n == "Prism")
		return 7;
	if (n == "Subsurface" || n == "SSS" || n == "Wax" || n == "Jade")
		return 8;
	if (n == "NormalMap" || n == "Normal")
		return 9;
	if (n == "Emissive" || n == "Emit" || n == "Glow")
		return 10;
	return 0;
}

// "v", "v/vt", "v/vt/vn", "v//vn" -> 0-based indices (src/ObjLoader.cpp:76-90).
bool parse_face_token(const std::string &tok, int &v, int &vt) {
	v = -1;
	vt = -1;
	try {
		size_t s1 = tok.find('/');
		if (s1 == std::string::npos) {
			v = std::stoi(tok) - 1;
			return true;
		}
		v = std::stoi(tok.substr(0, s1)) - 1;
		size_t s2 = tok.find('/', s1 + 1);
		std::string vts = (s2 == std::string::npos) ? tok.substr(s1 + 1) : tok.substr(s1 + 1, s2 - s1 - 1);
		if (!vts.empty())
			vt = std::stoi(vts) - 1;
	} catch (const std::exception &) {
		return false;
	}
	return true;
}

struct Face {
	int a, b, c, ta, tb, tc;
	int r, g, bl;
	int material;
	int texture;
};

} // namespace

// P6 PPM as 0xAARRGGBB (framework/TextureMap.cpp:4-31).
bool load_ppm(const std::string &path, Texture &tex, std::string &err) {
	std::ifstream in(path, std::ifstream::binary);
	if (!in.is_open()) {
		err = "cannot open texture " + path;
		return false;
	}
	std::string line;
	std::getline(in, line); // magic
	std::getline(in, line);
	while (!line.empty() && line[0] == '#')
		std::getline(in, line);
	std::vector<std::string> tok;
	{
		std::string cur;
		for (char ch : line) {
			if (ch == ' ') {
				tok.push_back(cur);
				cur.clear();
			} else {
				cur += ch;
			}
		}
		tok.push_back(cur);
	}
	if (tok.size() != 2) {
		err = "bad PPM size line in " + path;
		return false;
	}
	try {
		tex.width = std::stoi(tok[0]);
		tex.height = std::stoi(tok[1]);
	} catch (const std::exception &) {
		err = "bad PPM size in " + path;
		return false;
	}
	if (tex.width <= 0 || tex.height <= 0) {
		err = "bad PPM size in " + path;
		return false;
	}
	std::getline(in, line); // max value, assumed 255
	size_t n = (size_t)tex.width * (size_t)tex.height;
	tex.argb.resize(n);
	for (size_t i = 0; i < n; i++) {
		int r = in.get(), g = in.get(), b = in.get();
		tex.argb[i] = (255u << 24) + ((uint32_t)(r & 0xFF) << 16) + ((uint32_t)(g & 0xFF) << 8) + (uint32_t)(b & 0xFF);
	}
	return true;
}

// src/ObjLoader.cpp:102-214. Face normals use glm's cross/length order; the
// smoothed vertex normals are not built (the path tracer uses face normals).
bool load_obj(const std::string &path, float scale, HostScene &sc, std::string &err) {
	std::ifstream file(path);
	if (!file.is_open()) {
		err = "cannot open " + path;
		return false;
	}
	std::string dir;
	size_t slash = path.find_last_of("/\\");
	if (slash != std::string::npos)
		dir = path.substr(0, slash + 1);
	std::map<std::string, MtlDef> mtls;
	std::map<std::string, int> tex_cache;
	std::vector<float> vx;
	std::vector<float> vt;
	std::vector<Face> faces;
	int cr = 0, cg = 0, cb = 0, cmat = 0, ctex = -1;
	std::string line;
	while (std::getline(file, line)) {
		std::stringstream ss(line);
		std::string type;
		ss >> type;
		if (type == "mtllib") {
			std::string name;
			ss >> name;
			mtls = load_mtl(dir + name);
		} else if (type == "usemtl") {
			std::string name;
			ss >> name;
			cmat = material_for(name);
			ctex = -1;
			auto it = mtls.find(name);
			if (it != mtls.end()) {
				if (it->second.has_colour) {
					cr = it->second.r;
					cg = it->second.g;
					cb = it->second.b;
				} else {
					cr = cg = cb = 0;
				}
				if (!it->second.map_kd.empty()) {
					std::string tp = dir + it->second.map_kd;
					auto c = tex_cache.find(tp);
					if (c == tex_cache.end()) {
						Texture t;
						if (!load_ppm(tp, t, err))
							return false;
						sc.textures.push_back(std::move(t));
						ctex = (int)sc.textures.size() - 1;
						tex_cache[tp] = ctex;
					} else {
						ctex = c->second;
					}
				}
			}
		} else if (type == "v") {
			float x = 0, y = 0, z = 0;
			ss >> x >> y >> z;
			vx.push_back(x * scale);
			vx.push_back(y * scale);
			vx.push_back(z * scale);
		} else if (type == "vt") {
			float u = 0, v = 0;
			ss >> u >> v;
			vt.push_back(u);
			vt.push_back(v);
		} else if (type == "f") {
			std::string t1, t2, t3;
			ss >> t1 >> t2 >> t3;
			Face f;
			if (!parse_face_token(t1, f.a, f.ta) || !parse_face_token(t2, f.b, f.tb) ||
			    !parse_face_token(t3, f.c, f.tc)) {
				err = "malformed face in " + path;
				return false;
			}
			f.r = cr;
			f.g = cg;
			f.bl = cb;
			f.material = cmat;
			f.texture = ctex;
			faces.push_back(f);
		}
	}
	int nv = (int)(vx.size() / 3), nt = (int)(vt.size() / 2);
	size_t n = faces.size();
	sc.n = (int)n;
	sc.verts.resize(9 * n);
	sc.normals.resize(3 * n);
	sc.colour.resize(3 * n);
	sc.material.resize(n);
	sc.texture_id.resize(n);
	sc.texcoords.assign(6 * n, 0.0f);
	for (size_t i = 0; i < n; i++) {
		const Face &f = faces[i];
		int idx[3] = {f.a, f.b, f.c};
		for (int k = 0; k < 3; k++) {
			if (idx[k] < 0 || idx[k] >= nv) {
				err = "face index out of range in " + path;
				return false;
			}
			for (int a = 0; a < 3; a++)
				sc.verts[9 * i + 3 * k + a] = vx[3 * (size_t)idx[k] + a];
		}
		face_normal(&sc.verts[9 * i], &sc.normals[3 * i]);
		sc.colour[3 * i] = f.r;
		sc.colour[3 * i + 1] = f.g;
		sc.colour[3 * i + 2] = f.bl;
		sc.material[i] = f.material;
		sc.texture_id[i] = -1;
		if (f.ta >= 0 && f.tb >= 0 && f.tc >= 0) {
			int tix[3] = {f.ta, f.tb, f.tc};
			for (int k = 0; k < 3; k++) {
				if (tix[k] >= nt) {
					err = "texture coordinate index out of range in " + path;
					return false;
				}
				sc.texcoords[6 * i + 2 * k] = vt[2 * (size_t)tix[k]];
				sc.texcoords[6 * i + 2 * k + 1] = vt[2 * (size_t)tix[k] + 1];
			}
			sc.texture_id[i] = f.texture;
		}
	}
	// Physical-estimator view of an OBJ scene: albedo = Kd int / 255, the
	// Emissive material radiates its colour * 4 / 255 (src/Renderer.cpp:829-833).
	sc.albedo.resize(3 * n);
	sc.emission.assign(3 * n, 0.0f);
	for (size_t i = 0; i < n; i++)
		for (int a = 0; a < 3; a++) {
			sc.albedo[3 * i + a] = (float)sc.colour[3 * i + a] / 255.0f;
			if (sc.material[i] == 10)
				sc.emission[3 * i + a] = (float)sc.colour[3 * i + a] / 255.0f * 4.0f;
		}
	return true;
}

// Face normal with glm's operation order: cross(v1-v0, v2-v0) / length, or
// (0,1,0) for a degenerate face (src/ObjLoader.cpp:169-176).
void face_normal(const float *t, float *out) {
#pragma STDC FP_CONTRACT OFF
	float ax = t[3] - t[0], ay = t[4] - t[1], az = t[5] - t[2];
	float bx = t[6] - t[0], by = t[7] - t[1], bz = t[8] - t[2];
	float cx = ay * bz - by * az, cy = az * bx - bz * ax, cz = ax * by - bx * ay;
	float px = cx * cx, py = cy * cy, pz = cz * cz;
	float len = std::sqrt(px + py + pz);
	if (len > 1e-8f) {
		out[0] = cx / len;
		out[1] = cy / len;
		out[2] = cz / len;
	} else {
		out[0] = 0.0f;
		out[1] = 1.0f;
		out[2] = 0.0f;
	}
}

// ------------------------------------------------------------------ PNG --

namespace {

uint32_t crc_table[256];
bool crc_ready = false;

uint32_t crc32(const unsigned char *p, size_t n, uint32_t c = 0xFFFFFFFFu) {
	if (!crc_ready) {
		for (uint32_t i = 0; i < 256; i++) {
			uint32_t v = i;
			for (int k = 0; k < 8; k++)
				v = (v & 1) ? 0xEDB88320u ^ (v >> 1) : v >> 1;
			crc_table[i] = v;
		}
		crc_ready = true;
	}
	for (size_t i = 0; i < n; i++)
		c = crc_table[(c ^ p[i]) & 0xFF] ^ (c >> 8);
	return c;
}

void put32(std::vector<unsigned char> &o, uint32_t v) {
	o.push_back((unsigned char)(v >> 24));
	o.push_back((unsigned char)(v >> 16));
	o.push_back((unsigned char)(v >> 8));
	o.push_back((unsigned char)v);
}

void chunk(std::vector<unsigned char> &out, const char *type, const std::vector<unsigned char> &data) {
	put32(out, (uint32_t)data.size());
	size_t start = out.size();
	out.insert(out.end(), type, type + 4);
	out.insert(out.end(), data.begin(), data.end());
	uint32_t c = crc32(&out[start], out.size() - start) ^ 0xFFFFFFFFu;
	put32(out, c);
}

} // namespace

// RGBA8 PNG with an uncompressed (stored-block) zlib stream.
bool write_png(const std::string &path, int w, int h, const unsigned char *rgba) {
	static std::mutex mu;
	std::lock_guard<std::mutex> lock(mu);
	std::vector<unsigned char> raw;
	raw.reserve((size_t)h * ((size_t)w * 4 + 1));
	for (int y = 0; y < h; y++) {
		raw.push_back(0);
		raw.insert(raw.end(), rgba + (size_t)y * w * 4, rgba + (size_t)(y + 1) * w * 4);
	}
	std::vector<unsigned char> z = {0x78, 0x01};
	size_t pos = 0;
	do {
		size_t len = std::min<size_t>(65535, raw.size() - pos);
		bool last = pos + len == raw.size();
		z.push_back(last ? 1 : 0);
		z.push_back((unsigned char)(len & 0xFF));
		z.push_back((unsigned char)(len >> 8));
		z.push_back((unsigned char)(~len & 0xFF));
		z.push_back((unsigned char)((~len >> 8) & 0xFF));
		z.insert(z.end(), raw.begin() + (long)pos, raw.begin() + (long)(pos + len));
		pos += len;
	} while (pos < raw.size());
	uint32_t a = 1, b = 0;
	for (unsigned char ch : raw) {
		a = (a + ch) % 65521u;
		b = (b + a) % 65521u;
	}
	put32(z, (b << 16) | a);
	std::vector<unsigned char> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
	std::vector<unsigned char> ihdr;
	put32(ihdr, (uint32_t)w);
	put32(ihdr, (uint32_t)h);
	ihdr.push_back(8); // bit depth
	ihdr.push_back(6); // RGBA
	ihdr.push_back(0);
	ihdr.push_back(0);
	ihdr.push_back(0);
	chunk(out, "IHDR", ihdr);
	chunk(out, "IDAT", z);
	chunk(out, "IEND", {});
	FILE *f = std::fopen(path.c_str(), "wb");
	if (!f)
		return false;
	bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
	ok = (std::fclose(f) == 0) && ok;
	return ok;
}

} // namespace rn
