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
// smoothed vertex normals (used by the Whitted renderer) follow
// src/ObjLoader.cpp:178-205 with glm's dot/length/normalize order.
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
	// Smoothed vertex normals: per vertex, the normals of the adjacent faces
	// (in face order) within ~60 degrees of this face's (dot > 0.5), summed
	// and normalised; the face normal when the sum has zero length.
	{
#pragma STDC FP_CONTRACT OFF
		std::vector<std::vector<int>> vfaces((size_t)nv);
		for (size_t i = 0; i < n; i++) {
			vfaces[(size_t)faces[i].a].push_back((int)i);
			vfaces[(size_t)faces[i].b].push_back((int)i);
			vfaces[(size_t)faces[i].c].push_back((int)i);
		}
		sc.vnormals.resize(9 * n);
		auto dot = [](const float *a, const float *b) {
			float x = a[0] * b[0], y = a[1] * b[1], z = a[2] * b[2];
			return x + y + z;
		};
		for (size_t i = 0; i < n; i++) {
			const float *fi = &sc.normals[3 * i];
			int idx[3] = {faces[i].a, faces[i].b, faces[i].c};
			for (int k = 0; k < 3; k++) {
				float sum[3] = {0.0f, 0.0f, 0.0f};
				for (int g : vfaces[(size_t)idx[k]]) {
					const float *fg = &sc.normals[3 * (size_t)g];
					if (dot(fg, fi) > 0.5f)
						for (int a = 0; a < 3; a++)
							sum[a] += fg[a];
				}
				float *out = &sc.vnormals[9 * i + 3 * k];
				float len = std::sqrt(dot(sum, sum));
				if (len > 0.0f) {
					float inv = 1.0f / std::sqrt(dot(sum, sum));
					for (int a = 0; a < 3; a++)
						out[a] = sum[a] * inv;
				} else {
					for (int a = 0; a < 3; a++)
						out[a] = fi[a];
				}
			}
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

void flat_vertex_normals(HostScene &sc) {
	const size_t n = (size_t)sc.n;
	sc.vnormals.resize(9 * n);
	for (size_t i = 0; i < n; i++)
		for (int k = 0; k < 3; k++)
			for (int a = 0; a < 3; a++)
				sc.vnormals[9 * i + 3 * k + a] = sc.normals[3 * i + a];
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

// ------------------------------------------------------- binary scene cache --
//
// A scene cache file holds a HostScene as raw little-endian arrays so a
// large scene loads at file-read speed instead of OBJ-parsing speed:
//   bytes 0-7    "RNSCENE1"
//   u32          format version (1)
//   u32          n (triangles)
//   u32          number of textures
//   u32          reserved (0)
//   u64          checksum of everything after this header
//   f32 verts[9n], f32 albedo[3n], f32 emission[3n], f32 normals[3n],
//   f32 vertex normals[9n], i32 colour[3n], i32 material[n],
//   i32 texture_id[n], f32 texcoords[6n],
//   then per texture: i32 width, i32 height, u32 argb[width * height].
// The checksum folds the payload as 64-bit words (the tail zero-padded):
// h = rotl(h ^ w, 29) * 0x9E3779B97F4A7C15, starting from the byte count.

namespace {

constexpr char kCacheMagic[8] = {'R', 'N', 'S', 'C', 'E', 'N', 'E', '1'};
constexpr uint32_t kCacheVersion = 1;

struct CacheHash {
	uint64_t h = 0, bytes = 0;
	unsigned char tail[8];
	int ntail = 0;
	void word(uint64_t w) {
		h ^= w;
		h = (h << 29) | (h >> 35);
		h *= 0x9E3779B97F4A7C15ull;
	}
	void add(const void *p, size_t n) {
		const unsigned char *b = static_cast<const unsigned char *>(p);
		bytes += n;
		while (n > 0 && ntail > 0 && ntail < 8) {
			tail[ntail++] = *b++;
			n--;
			if (ntail == 8) {
				uint64_t w;
				std::memcpy(&w, tail, 8);
				word(w);
				ntail = 0;
			}
		}
		while (n >= 8) {
			uint64_t w;
			std::memcpy(&w, b, 8);
			word(w);
			b += 8;
			n -= 8;
		}
		while (n > 0) {
			tail[ntail++] = *b++;
			n--;
		}
	}
	uint64_t finish() {
		if (ntail > 0) {
			std::memset(tail + ntail, 0, (size_t)(8 - ntail));
			uint64_t w;
			std::memcpy(&w, tail, 8);
			word(w);
			ntail = 0;
		}
		word(bytes);
		return h;
	}
};

template <class T> bool put(std::FILE *f, CacheHash &hs, const std::vector<T> &v) {
	if (v.empty())
		return true;
	hs.add(v.data(), v.size() * sizeof(T));
	return std::fwrite(v.data(), sizeof(T), v.size(), f) == v.size();
}

template <class T> bool get(std::FILE *f, CacheHash &hs, std::vector<T> &v, size_t count) {
	v.resize(count);
	if (count == 0)
		return true;
	if (std::fread(v.data(), sizeof(T), count, f) != count)
		return false;
	hs.add(v.data(), count * sizeof(T));
	return true;
}

} // namespace

bool save_scene(const std::string &path, const HostScene &sc, std::string &err) {
	const size_t n = (size_t)sc.n;
	if (sc.verts.size() != 9 * n || sc.albedo.size() != 3 * n || sc.emission.size() != 3 * n ||
	    sc.normals.size() != 3 * n || sc.vnormals.size() != 9 * n || sc.colour.size() != 3 * n ||
	    sc.material.size() != n ||
	    sc.texture_id.size() != n || sc.texcoords.size() != 6 * n) {
		err = "scene arrays are inconsistent";
		return false;
	}
	const std::string tmp = path + ".tmp";
	std::FILE *f = std::fopen(tmp.c_str(), "wb");
	if (!f) {
		err = "cannot open " + tmp + " for writing";
		return false;
	}
	// header with a placeholder checksum, patched after the payload
	uint32_t hdr[4] = {kCacheVersion, (uint32_t)n, (uint32_t)sc.textures.size(), 0u};
	uint64_t sum = 0;
	bool ok = std::fwrite(kCacheMagic, 1, 8, f) == 8 && std::fwrite(hdr, 4, 4, f) == 4 && std::fwrite(&sum, 8, 1, f) == 1;
	CacheHash hs;
	ok = ok && put(f, hs, sc.verts) && put(f, hs, sc.albedo) && put(f, hs, sc.emission) && put(f, hs, sc.normals) && put(f, hs, sc.vnormals) &&
	     put(f, hs, sc.colour) && put(f, hs, sc.material) && put(f, hs, sc.texture_id) && put(f, hs, sc.texcoords);
	for (const Texture &t : sc.textures) {
		if (!ok)
			break;
		int32_t wh[2] = {t.width, t.height};
		hs.add(wh, sizeof wh);
		ok = std::fwrite(wh, 4, 2, f) == 2 && put(f, hs, t.argb);
	}
	sum = hs.finish();
	ok = ok && std::fseek(f, 24, SEEK_SET) == 0 && std::fwrite(&sum, 8, 1, f) == 1;
	ok = (std::fclose(f) == 0) && ok;
	if (!ok || std::rename(tmp.c_str(), path.c_str()) != 0) {
		std::remove(tmp.c_str());
		err = "write failed: " + path;
		return false;
	}
	return true;
}

bool load_scene(const std::string &path, HostScene &sc, std::string &err) {
	std::FILE *f = std::fopen(path.c_str(), "rb");
	if (!f) {
		err = "cannot open " + path;
		return false;
	}
	struct Closer {
		std::FILE *f;
		~Closer() { std::fclose(f); }
	} closer{f};
	char magic[8];
	uint32_t hdr[4];
	uint64_t sum = 0;
	if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, kCacheMagic, 8) != 0) {
		err = path + " is not a scene cache file";
		return false;
	}
	if (std::fread(hdr, 4, 4, f) != 4 || std::fread(&sum, 8, 1, f) != 1) {
		err = path + ": truncated header";
		return false;
	}
	// Bytes after the 32-byte header. Every count read below is checked
	// against them before any array is sized, so a truncated file or a
	// corrupt header cannot make the loader allocate what the file does not
	// hold.
	long long left = -1;
	if (std::fseek(f, 0, SEEK_END) == 0) {
		const long long end = std::ftell(f);
		if (end >= 32 && std::fseek(f, 32, SEEK_SET) == 0)
			left = end - 32;
	}
	if (left < 0) {
		err = path + ": cannot determine the file size";
		return false;
	}
	if (hdr[0] != kCacheVersion) {
		err = path + ": unsupported scene cache version " + std::to_string(hdr[0]);
		return false;
	}
	const size_t n = hdr[1], nt = hdr[2];
	if (n == 0 || n > (1u << 30) || nt > (1u << 20)) {
		err = path + ": implausible triangle or texture count";
		return false;
	}
	// per triangle: verts 9, albedo 3, emission 3, normals 3, vertex normals 9,
	// texcoords 6 floats; colour 3, material 1, texture id 1 ints
	constexpr long long kTriBytes = (9 + 3 + 3 + 3 + 9 + 6) * 4 + (3 + 1 + 1) * 4;
	if ((long long)n * kTriBytes > left) {
		err = path + ": truncated (the header counts more triangles than the file holds)";
		return false;
	}
	left -= (long long)n * kTriBytes;
	HostScene h;
	h.n = (int)n;
	CacheHash hs;
	bool ok = get(f, hs, h.verts, 9 * n) && get(f, hs, h.albedo, 3 * n) && get(f, hs, h.emission, 3 * n) &&
	          get(f, hs, h.normals, 3 * n) && get(f, hs, h.vnormals, 9 * n) && get(f, hs, h.colour, 3 * n) && get(f, hs, h.material, n) &&
	          get(f, hs, h.texture_id, n) && get(f, hs, h.texcoords, 6 * n);
	h.textures.resize(nt);
	for (size_t k = 0; ok && k < nt; k++) {
		int32_t wh[2];
		ok = std::fread(wh, 4, 2, f) == 2;
		if (!ok)
			break;
		hs.add(wh, sizeof wh);
		left -= 8;
		// the OBJ loader's textures are at least 1x1 (load_ppm rejects empty
		// images), and a texel lookup into an empty texture would read
		// outside the texel buffer
		if (wh[0] <= 0 || wh[1] <= 0 || (int64_t)wh[0] * wh[1] > (int64_t)1 << 28) {
			err = path + ": implausible texture size";
			return false;
		}
		if ((long long)wh[0] * wh[1] * 4 > left) {
			err = path + ": truncated";
			return false;
		}
		left -= (long long)wh[0] * wh[1] * 4;
		h.textures[k].width = wh[0];
		h.textures[k].height = wh[1];
		ok = get(f, hs, h.textures[k].argb, (size_t)wh[0] * (size_t)wh[1]);
	}
	if (!ok) {
		err = path + ": truncated";
		return false;
	}
	if (std::fgetc(f) != EOF) {
		err = path + ": trailing bytes after the scene";
		return false;
	}
	if (hs.finish() != sum) {
		err = path + ": checksum mismatch (corrupt scene cache)";
		return false;
	}
	for (int id : h.texture_id)
		if (id < -1 || id >= (int)nt) {
			err = path + ": texture id out of range";
			return false;
		}
	sc = std::move(h);
	return true;
}

} // namespace rn
