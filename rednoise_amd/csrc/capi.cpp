// The C ABI of librednoise.so (include/rednoise/rednoise.h, rednoise_ex.h):
// scene ownership, device residency, argument checking and the error
// contract. Every entry point catches everything and reports through
// rn_last_error(); no exception crosses the ABI.
#include "rednoise/rednoise.h"
#include "rednoise/rednoise_ex.h"

#include "rn_internal.h"
#include "rn_kernels.h"
#include "scene_io.h"
#include "photon.h"
#include "wavefront.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <numeric>

namespace rn {

void hip_check(hipError_t e, const char *what, const char *file, int line) {
	if (e != hipSuccess)
		throw Error(std::string(hipGetErrorString(e)) + " in " + what + " (" + file + ":" + std::to_string(line) + ")");
}

void DevBuf::release() {
	if (ptr)
		(void)hipFree(ptr);
	ptr = nullptr;
	bytes = 0;
}

bool fault_injected(const char *what) {
	const char *env = std::getenv("RN_FAULT_INJECT");
	if (!env || !*env)
		return false;
	const size_t len = std::strlen(what);
	for (const char *p = env; *p;) {
		const char *q = std::strchr(p, ',');
		const size_t n = q ? (size_t)(q - p) : std::strlen(p);
		if (n == len && std::strncmp(p, what, len) == 0)
			return true;
		if (!q)
			break;
		p = q + 1;
	}
	return false;
}

void *DevBuf::reserve(size_t n) {
	if (n == 0)
		n = 16;
	if (n > bytes) {
		if (fault_injected("alloc"))
			throw Error("injected allocation failure (RN_FAULT_INJECT=alloc): " + std::to_string(n) + " bytes");
		release();
		RN_HIP(hipMalloc(&ptr, n));
		bytes = n;
	}
	return ptr;
}

namespace {

thread_local std::string t_error;
thread_local rn_stats t_stats;

void set_error(const std::string &m) { t_error = m; }

// Process-wide device context: the library's own stream, and one lock that
// serialises every call using the shared workspace.
struct Context {
	std::mutex mu;
	hipStream_t stream = nullptr;
	std::string arch;
	int device = -1;
	DevBuf out, in0, in1, in2, in3, o0, o1, o2, o3, o4, counters;
	hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

Context &ctx() {
	static Context c;
	if (c.device < 0) {
		int count = 0;
		if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
			throw Error("no HIP device available (librednoise.so needs an MI355X / gfx950 GPU)");
		RN_HIP(hipGetDevice(&c.device));
		hipDeviceProp_t prop;
		RN_HIP(hipGetDeviceProperties(&prop, c.device));
		c.arch = prop.gcnArchName;
		RN_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
		RN_HIP(hipEventCreate(&c.ev0));
		RN_HIP(hipEventCreate(&c.ev1));
	}
	return c;
}

// The BVH builders order and cluster triangles by their boxes and centroids:
// a NaN or infinite coordinate has no place in that order (the clustering
// would never merge its triangle), so geometry is checked before it reaches
// the device.
void check_finite(const float *verts, int n, const char *who) {
	for (size_t i = 0; i < 9 * (size_t)n; i++)
		if (!std::isfinite(verts[i]))
			throw Error(std::string(who) + ": triangle " + std::to_string(i / 9) +
			            " has a non-finite vertex coordinate");
}

template <class T> T *upload(DevBuf &b, const T *h, size_t count, hipStream_t s) {
	T *d = static_cast<T *>(b.reserve(count * sizeof(T)));
	if (count)
		RN_HIP(hipMemcpyAsync(d, h, count * sizeof(T), hipMemcpyHostToDevice, s));
	return d;
}

template <class T> void download(T *h, const void *d, size_t count, hipStream_t s) {
	if (count)
		RN_HIP(hipMemcpyAsync(h, d, count * sizeof(T), hipMemcpyDeviceToHost, s));
}

double ms_since(std::chrono::steady_clock::time_point t0) {
	return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

} // namespace
} // namespace rn

using namespace rn;

struct rn_scene {
	HostScene h;
	std::mutex mu;
	bool uploaded = false;
	bool bvh_valid = false;
	int builder = RN_DEFAULT_BUILDER; // RN_BVH_PLOC or RN_BVH_LBVH
	int nl = 0;
	float total_area = 0.0f;
	float radius = 1.0f;
	int n_hot = 0;
	DevBuf d_verts, d_albedo, d_emission, d_light_tri, d_light_cdf, d_light_geom, d_hot_slot, d_hot_prim;
	DevBuf d_cnorm, d_vnorm, d_ccol, d_ctc, d_texels, d_texinfo;
	Bvh bvh;

	// Device copies of geometry and materials, built on first use.
	void ensure_device(hipStream_t s) {
		if (uploaded)
			return;
		const int n = h.n;
		check_finite(h.verts.data(), n, "scene");
		upload(d_verts, h.verts.data(), h.verts.size(), s);
		upload(d_albedo, h.albedo.data(), h.albedo.size(), s);
		upload(d_emission, h.emission.data(), h.emission.size(), s);
		// Light set: triangles emitting in any channel, chosen by area.
		std::vector<int> lt;
		std::vector<double> acc;
		std::vector<float> geom;
		double total = 0.0;
		std::vector<double> area((size_t)n);
		float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
		for (int i = 0; i < n; i++) {
			const float *t = &h.verts[9 * (size_t)i];
			double e1[3], e2[3];
			for (int a = 0; a < 3; a++) {
				e1[a] = (double)t[3 + a] - t[a];
				e2[a] = (double)t[6 + a] - t[a];
				for (int k = 0; k < 3; k++) {
					lo[a] = std::min(lo[a], t[3 * k + a]);
					hi[a] = std::max(hi[a], t[3 * k + a]);
				}
			}
			double c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
			area[(size_t)i] = 0.5 * std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
			const float *em = &h.emission[3 * (size_t)i];
			if (em[0] > 0.0f || em[1] > 0.0f || em[2] > 0.0f) {
				total += area[(size_t)i];
				acc.push_back(total);
				lt.push_back(i);
				for (int k = 0; k < 3; k++) {
					geom.push_back(k == 0 ? t[0] : (float)(k == 1 ? e1[0] : e2[0]));
					geom.push_back(k == 0 ? t[1] : (float)(k == 1 ? e1[1] : e2[1]));
					geom.push_back(k == 0 ? t[2] : (float)(k == 1 ? e1[2] : e2[2]));
					geom.push_back(0.0f);
				}
			}
		}
		// Light geometry uses float edges computed as in the kernels' own
		// vertex differences so that a light sample lies on the triangle.
		for (size_t l = 0; l < lt.size(); l++) {
			const float *t = &h.verts[9 * (size_t)lt[l]];
			for (int a = 0; a < 3; a++) {
				geom[12 * l + 4 + a] = t[3 + a] - t[a];
				geom[12 * l + 8 + a] = t[6 + a] - t[a];
			}
		}
		nl = (int)lt.size();
		total_area = (float)total;
		std::vector<float> cdf(lt.size());
		for (size_t i = 0; i < lt.size(); i++)
			cdf[i] = (float)(acc[i] / total);
		if (!cdf.empty())
			cdf.back() = 1.0f;
		upload(d_light_tri, lt.data(), lt.size(), s);
		upload(d_light_cdf, cdf.data(), cdf.size(), s);
		upload(d_light_geom, geom.data(), geom.size(), s);
		float ext = 0.0f;
		for (int a = 0; a < 3; a++)
			if (n > 0)
				ext = std::max(ext, hi[a] - lo[a]);
		radius = n > 0 ? 0.5f * ext * std::sqrt(3.0f) : 1.0f;
		// Hot set for the adjoint's LDS gradient accumulators: lights first,
		// then the largest triangles.
		std::vector<int> order((size_t)n);
		std::iota(order.begin(), order.end(), 0);
		std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return area[(size_t)a] > area[(size_t)b]; });
		std::vector<int> hot;
		std::vector<int> slot((size_t)std::max(n, 1), -1);
		for (int l : lt)
			if ((int)hot.size() < kHotSlots) {
				slot[(size_t)l] = (int)hot.size();
				hot.push_back(l);
			}
		for (int i : order) {
			if ((int)hot.size() >= kHotSlots)
				break;
			if (slot[(size_t)i] < 0) {
				slot[(size_t)i] = (int)hot.size();
				hot.push_back(i);
			}
		}
		n_hot = (int)hot.size();
		upload(d_hot_slot, slot.data(), slot.size(), s);
		upload(d_hot_prim, hot.data(), hot.size(), s);
		// Reference-estimator attributes.
		std::vector<float4> cn((size_t)n), cc((size_t)n);
		for (int i = 0; i < n; i++) {
			int mat = h.material.empty() ? 0 : h.material[(size_t)i];
			int tex = h.texture_id.empty() ? -1 : h.texture_id[(size_t)i];
			float nx = h.normals[3 * (size_t)i], ny = h.normals[3 * (size_t)i + 1], nz = h.normals[3 * (size_t)i + 2];
			cn[(size_t)i] = make_float4(nx, ny, nz, 0.0f);
			std::memcpy(&cn[(size_t)i].w, &mat, 4);
			cc[(size_t)i] = make_float4((float)h.colour[3 * (size_t)i], (float)h.colour[3 * (size_t)i + 1],
			                            (float)h.colour[3 * (size_t)i + 2], 0.0f);
			std::memcpy(&cc[(size_t)i].w, &tex, 4);
		}
		upload(d_cnorm, cn.data(), cn.size(), s);
		upload(d_ccol, cc.data(), cc.size(), s);
		upload(d_ctc, h.texcoords.data(), h.texcoords.size(), s);
		upload(d_vnorm, h.vnormals.data(), h.vnormals.size(), s);
		std::vector<uint32_t> texels;
		std::vector<int4> info;
		for (const Texture &t : h.textures) {
			info.push_back(make_int4((int)texels.size(), t.width, t.height, 0));
			texels.insert(texels.end(), t.argb.begin(), t.argb.end());
		}
		upload(d_texels, texels.data(), texels.size(), s);
		upload(d_texinfo, info.data(), info.size(), s);
		RN_HIP(hipStreamSynchronize(s));
		uploaded = true;
	}

	void ensure_bvh(int width, hipStream_t s) {
		ensure_device(s);
		if (width <= 0)
			width = 8;
		if (bvh_valid && bvh.width == width && bvh.builder == builder)
			return;
		build_bvh(bvh, d_verts.as<float>(), h.n, width, s, builder);
		// checked before any traversal kernel runs on this BVH: a deeper tree
		// would overflow the traversal stack (LDS entries + the spill stack)
		if (bvh.depth > kMaxBvhDepth || fault_injected("bvh_depth"))
			throw Error("BVH deeper than the traversal stack supports (depth " + std::to_string(bvh.depth) +
			            ", limit " + std::to_string(kMaxBvhDepth) + ")");
		bvh_valid = true;
	}

	DevScene dev(const float *albedo = nullptr, const float *emission = nullptr) const {
		DevScene d;
		d.nodes = bvh.nodes.as<Bvh8Node>();
		d.tris = bvh.tris.as<float4>();
		d.albedo = albedo ? albedo : d_albedo.as<float>();
		d.emission = emission ? emission : d_emission.as<float>();
		d.light_tri = d_light_tri.as<int>();
		d.light_cdf = d_light_cdf.as<float>();
		d.light_geom = d_light_geom.as<float4>();
		d.n = h.n;
		d.nl = nl;
		d.total_light_area = total_area;
		d.depth = bvh.depth;
		return d;
	}

	CompatDev compat() const {
		CompatDev c;
		c.cnorm = d_cnorm.as<float4>();
		c.vnorm = d_vnorm.as<float>();
		c.ccol = d_ccol.as<float4>();
		c.ctc = d_ctc.as<float>();
		c.texels = d_texels.as<uint32_t>();
		c.tex_info = d_texinfo.as<int4>();
		return c;
	}
};

namespace rn {

DevCamera make_camera(const rn_camera &c, int W, int H) {
	double w[3] = {(double)c.target[0] - c.eye[0], (double)c.target[1] - c.eye[1], (double)c.target[2] - c.eye[2]};
	double lw = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
	if (!(lw > 0.0))
		throw Error("camera eye and target coincide");
	for (double &x : w)
		x /= lw;
	double u[3] = {w[1] * c.up[2] - w[2] * c.up[1], w[2] * c.up[0] - w[0] * c.up[2], w[0] * c.up[1] - w[1] * c.up[0]};
	double lu = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
	if (!(lu > 0.0))
		throw Error("camera up vector is parallel to the view direction");
	for (double &x : u)
		x /= lu;
	double v[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
	if (!(c.vfov_deg > 0.0f && c.vfov_deg < 180.0f))
		throw Error("camera vfov_deg must be in (0, 180)");
	double th = std::tan(0.5 * (double)c.vfov_deg * 3.14159265358979323846 / 180.0);
	double aspect = (double)W / (double)H;
	DevCamera d;
	for (int k = 0; k < 3; k++) {
		d.eye[k] = c.eye[k];
		d.w[k] = (float)w[k];
		d.du[k] = (float)(u[k] * th * aspect);
		d.dv[k] = (float)(v[k] * th);
	}
	return d;
}

DevParams make_params(const rn_render_params &p, const rn_scene *sc, uint32_t scene_id) {
	if (p.width <= 0 || p.height <= 0)
		throw Error("width and height must be positive");
	if ((long long)p.width * p.height > (1ll << 31) / 4)
		throw Error("image too large");
	if (p.spp <= 0)
		throw Error("spp must be >= 1");
	if (p.max_depth <= 0 || p.max_depth > 64)
		throw Error("max_depth must be in [1, 64]");
	DevParams d;
	d.width = p.width;
	d.height = p.height;
	d.spp = p.spp;
	d.max_depth = p.max_depth;
	d.rr_depth = p.rr_depth;
	d.rr_prob = p.rr_prob;
	d.seed = p.seed;
	d.scene_id = scene_id;
	d.ray_eps = p.ray_eps > 0.0f ? p.ray_eps : 1e-4f * sc->radius;
	d.sample_base = 0;
	return d;
}

} // namespace rn

struct rn_photon_map {
	PhotonMap m;
};

#define RN_TRY try {
#define RN_CATCH(ret)                                                                                                  \
	}                                                                                                                  \
	catch (const std::exception &e) {                                                                                  \
		set_error(e.what());                                                                                           \
		return ret;                                                                                                    \
	}                                                                                                                  \
	catch (...) {                                                                                                      \
		set_error("unknown error");                                                                                    \
		return ret;                                                                                                    \
	}

extern "C" {

const char *rn_last_error(void) { return t_error.c_str(); }

const char *rn_version(void) { return "0.1.2"; }

int rn_release_workspace(void) {
	RN_TRY
	t_error.clear();
	Context &c = ctx();
	std::lock_guard<std::mutex> lk(c.mu);
	if (c.stream)
		RN_HIP(hipStreamSynchronize(c.stream));
	RN_HIP(hipDeviceSynchronize());
	wavefront_release();
	for (DevBuf *b : {&c.out, &c.in0, &c.in1, &c.in2, &c.in3, &c.o0, &c.o1, &c.o2, &c.o3, &c.o4})
		b->release();
	return 1;
	RN_CATCH(0)
}

const char *rn_device_name(void) {
	RN_TRY
	static std::string name;
	name = ctx().arch;
	return name.c_str();
	RN_CATCH("")
}

rn_scene *rn_scene_load_obj(const char *path, float scale) {
	RN_TRY
	t_error.clear();
	if (!path) {
		set_error("path is NULL");
		return nullptr;
	}
	std::unique_ptr<rn_scene> s(new (std::nothrow) rn_scene());
	if (!s) {
		set_error("out of memory");
		return nullptr;
	}
	std::string err;
	if (!load_obj(path, scale, s->h, err)) {
		set_error(err);
		return nullptr;
	}
	if (s->h.n == 0) {
		set_error("OBJ has no triangles");
		return nullptr;
	}
	return s.release();
	RN_CATCH(nullptr)
}

rn_scene *rn_scene_from_arrays(const float *verts, int n, const float *albedo, const float *emission) {
	RN_TRY
	t_error.clear();
	if (!verts || !albedo || n <= 0) {
		set_error("verts/albedo must be non-NULL and n > 0");
		return nullptr;
	}
	std::unique_ptr<rn_scene> s(new rn_scene());
	HostScene &h = s->h;
	h.n = n;
	h.verts.assign(verts, verts + 9 * (size_t)n);
	h.albedo.assign(albedo, albedo + 3 * (size_t)n);
	if (emission)
		h.emission.assign(emission, emission + 3 * (size_t)n);
	else
		h.emission.assign(3 * (size_t)n, 0.0f);
	h.normals.resize(3 * (size_t)n);
	h.colour.resize(3 * (size_t)n);
	h.material.assign((size_t)n, 0);
	h.texture_id.assign((size_t)n, -1);
	h.texcoords.assign(6 * (size_t)n, 0.0f);
	for (size_t i = 0; i < (size_t)n; i++) {
		face_normal(&h.verts[9 * i], &h.normals[3 * i]);
		for (int a = 0; a < 3; a++)
			h.colour[3 * i + a] = (int)std::lround(std::min(1.0f, std::max(0.0f, albedo[3 * i + a])) * 255.0f);
	}
	flat_vertex_normals(h);
	return s.release();
	RN_CATCH(nullptr)
}

int rn_scene_save(const rn_scene *scene, const char *path) {
	RN_TRY
	t_error.clear();
	if (!scene || !path) {
		set_error("scene and path must be non-NULL");
		return 0;
	}
	std::lock_guard<std::mutex> lk(const_cast<rn_scene *>(scene)->mu);
	std::string err;
	if (!save_scene(path, scene->h, err)) {
		set_error(err);
		return 0;
	}
	return 1;
	RN_CATCH(0)
}

rn_scene *rn_scene_load(const char *path) {
	RN_TRY
	t_error.clear();
	if (!path) {
		set_error("path is NULL");
		return nullptr;
	}
	std::unique_ptr<rn_scene> s(new (std::nothrow) rn_scene());
	if (!s) {
		set_error("out of memory");
		return nullptr;
	}
	std::string err;
	if (!load_scene(path, s->h, err)) {
		set_error(err);
		return nullptr;
	}
	return s.release();
	RN_CATCH(nullptr)
}

void rn_scene_free(rn_scene *scene) {
	if (!scene)
		return;
	if (scene->uploaded || scene->bvh_valid) {
		// device work queued by an asynchronous entry point may still read
		// the scene's buffers: drain the library stream first
		try {
			Context &c = ctx();
			std::lock_guard<std::mutex> lk(c.mu);
			(void)hipStreamSynchronize(c.stream);
			delete scene;
			return;
		} catch (...) {
		}
	}
	delete scene;
}

int rn_scene_triangle_count(const rn_scene *scene) { return scene ? scene->h.n : 0; }

int rn_scene_set_materials(rn_scene *scene, const float *albedo, const float *emission) {
	RN_TRY
	if (!scene) {
		set_error("scene is NULL");
		return 0;
	}
	std::lock_guard<std::mutex> lk(scene->mu);
	size_t n3 = 3 * (size_t)scene->h.n;
	if (albedo)
		scene->h.albedo.assign(albedo, albedo + n3);
	if (emission)
		scene->h.emission.assign(emission, emission + n3);
	if (scene->uploaded) {
		Context &c = ctx();
		if (albedo)
			RN_HIP(hipMemcpyAsync(scene->d_albedo.ptr, albedo, n3 * 4, hipMemcpyHostToDevice, c.stream));
		if (emission)
			RN_HIP(hipMemcpyAsync(scene->d_emission.ptr, emission, n3 * 4, hipMemcpyHostToDevice, c.stream));
		RN_HIP(hipStreamSynchronize(c.stream));
	}
	return 1;
	RN_CATCH(0)
}

int rn_scene_get_triangles(const rn_scene *scene, float *verts, float *albedo, float *emission) {
	if (!scene) {
		set_error("scene is NULL");
		return 0;
	}
	const HostScene &h = scene->h;
	if (verts)
		std::memcpy(verts, h.verts.data(), h.verts.size() * 4);
	if (albedo)
		std::memcpy(albedo, h.albedo.data(), h.albedo.size() * 4);
	if (emission)
		std::memcpy(emission, h.emission.data(), h.emission.size() * 4);
	return 1;
}

int rn_scene_get_compat(const rn_scene *scene, float *normals, int *colour, int *material, int *texture_id,
                        float *texcoords) {
	if (!scene) {
		set_error("scene is NULL");
		return 0;
	}
	const HostScene &h = scene->h;
	if (normals)
		std::memcpy(normals, h.normals.data(), h.normals.size() * 4);
	if (colour)
		std::memcpy(colour, h.colour.data(), h.colour.size() * 4);
	if (material)
		std::memcpy(material, h.material.data(), h.material.size() * 4);
	if (texture_id)
		std::memcpy(texture_id, h.texture_id.data(), h.texture_id.size() * 4);
	if (texcoords)
		std::memcpy(texcoords, h.texcoords.data(), h.texcoords.size() * 4);
	return 1;
}

int rn_scene_get_vertex_normals(const rn_scene *scene, float *vnormals) {
	if (!scene || !vnormals) {
		set_error("scene and vnormals must be non-NULL");
		return 0;
	}
	std::memcpy(vnormals, scene->h.vnormals.data(), scene->h.vnormals.size() * 4);
	return 1;
}

int rn_scene_get_texture(const rn_scene *scene, int id, int *width, int *height, uint32_t *argb) {
	if (!scene || id < 0 || id >= (int)scene->h.textures.size()) {
		set_error("no such texture");
		return 0;
	}
	const Texture &t = scene->h.textures[(size_t)id];
	if (width)
		*width = t.width;
	if (height)
		*height = t.height;
	if (argb)
		std::memcpy(argb, t.argb.data(), t.argb.size() * 4);
	return 1;
}

// RN_PATHTRACED: with mt = true the reference's own per-pixel std::mt19937
// streams (the reference's image); otherwise Philox under the key (k0, k1).
static int render_compat(const rn_scene *scene, int width, int height, float cam_z, int samples, uint32_t k0,
                         uint32_t k1, bool mt, unsigned char *rgba) {
	if (samples <= 0)
		throw Error("rn_render: samples must be >= 1 for RN_PATHTRACED");
	if ((long long)width * height > (1ll << 29))
		throw Error("rn_render: image too large");
	auto t0 = std::chrono::steady_clock::now();
	Context &c = ctx();
	std::lock_guard<std::mutex> lk(c.mu);
	rn_scene *sc = const_cast<rn_scene *>(scene);
	std::lock_guard<std::mutex> sl(sc->mu);
	bool had = sc->bvh_valid;
	sc->ensure_bvh(8, c.stream);
	unsigned char *d_rgba = static_cast<unsigned char *>(c.out.reserve((size_t)width * height * 4));
	RN_HIP(hipEventRecord(c.ev0, c.stream));
	if (mt)
		launch_compat_render_mt(sc->dev(), sc->compat(), width, height, cam_z, samples, d_rgba, c.stream);
	else
		launch_compat_render(sc->dev(), sc->compat(), width, height, cam_z, samples, k0, k1, d_rgba, c.stream);
	RN_HIP(hipEventRecord(c.ev1, c.stream));
	download<unsigned char>(rgba, d_rgba, (size_t)width * height * 4, c.stream);
	RN_HIP(hipStreamSynchronize(c.stream));
	float ms = 0;
	RN_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
	std::memset(&t_stats, 0, sizeof t_stats);
	t_stats.build_ms = had ? 0.0 : sc->bvh.build_ms;
	t_stats.render_ms = ms;
	t_stats.total_ms = ms_since(t0);
	t_stats.paths = (uint64_t)width * height * samples;
	t_stats.trace_kernel_ms = ms;
	t_stats.trace_kernel_launches = 1;
	t_stats.bvh_nodes = sc->bvh.n_nodes;
	t_stats.bvh_depth = sc->bvh.depth;
	return 1;
}

// RN_RAYTRACED: the reference's Whitted renderer (deterministic; samples
// is ignored, as in src/capi.cpp:59-61).
static int render_whitted(const rn_scene *scene, int width, int height, float cam_z, unsigned char *rgba) {
	if ((long long)width * height > (1ll << 29))
		throw Error("rn_render: image too large");
	auto t0 = std::chrono::steady_clock::now();
	Context &c = ctx();
	std::lock_guard<std::mutex> lk(c.mu);
	rn_scene *sc = const_cast<rn_scene *>(scene);
	std::lock_guard<std::mutex> sl(sc->mu);
	bool had = sc->bvh_valid;
	sc->ensure_bvh(8, c.stream);
	unsigned char *d_rgba = static_cast<unsigned char *>(c.out.reserve((size_t)width * height * 4));
	RN_HIP(hipEventRecord(c.ev0, c.stream));
	launch_whitted_render(sc->dev(), sc->compat(), width, height, cam_z, d_rgba, c.stream);
	RN_HIP(hipEventRecord(c.ev1, c.stream));
	download<unsigned char>(rgba, d_rgba, (size_t)width * height * 4, c.stream);
	RN_HIP(hipStreamSynchronize(c.stream));
	float ms = 0;
	RN_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
	std::memset(&t_stats, 0, sizeof t_stats);
	t_stats.build_ms = had ? 0.0 : sc->bvh.build_ms;
	t_stats.render_ms = ms;
	t_stats.total_ms = ms_since(t0);
	t_stats.trace_kernel_ms = ms;
	t_stats.trace_kernel_launches = 1;
	t_stats.bvh_nodes = sc->bvh.n_nodes;
	t_stats.bvh_depth = sc->bvh.depth;
	return 1;
}

int rn_render(const rn_scene *scene, rn_render_mode mode, int width, int height, float cam_z, int samples,
              unsigned char *rgba) {
	RN_TRY
	t_error.clear();
	if (!scene || !rgba || width <= 0 || height <= 0) {
		set_error("rn_render: NULL scene/buffer or non-positive size");
		return 0;
	}
	if (mode == RN_RAYTRACED)
		return render_whitted(scene, width, height, cam_z, rgba);
	if (mode != RN_PATHTRACED) {
		set_error("rn_render: only RN_PATHTRACED and RN_RAYTRACED are implemented by this library");
		return 0;
	}
	return render_compat(scene, width, height, cam_z, samples, 0u, 0u, true, rgba);
	RN_CATCH(0)
}

int rn_render_seeded(const rn_scene *scene, int width, int height, float cam_z, int samples, uint32_t key0,
                     uint32_t key1, unsigned char *rgba) {
	RN_TRY
	t_error.clear();
	if (!scene || !rgba || width <= 0 || height <= 0)
		throw Error("rn_render_seeded: NULL scene/buffer or non-positive size");
	return render_compat(scene, width, height, cam_z, samples, key0, key1, false, rgba);
	RN_CATCH(0)
}

int rn_save_png(const char *path, int width, int height, const unsigned char *rgba) {
	RN_TRY
	if (!path || !rgba || width <= 0 || height <= 0) {
		set_error("rn_save_png: bad arguments");
		return 0;
	}
	if (!write_png(path, width, height, rgba)) {
		set_error("rn_save_png: cannot write file");
		return 0;
	}
	return 1;
	RN_CATCH(0)
}

// Device-pointer entry points run on the caller's stream (NULL = the legacy
// default stream, which orders them with torch's default stream); host-pointer
// entry points run on the library's own stream.
struct rn_accum {
	int width = 0, height = 0;
	long long samples = 0;
	DevBuf sum; // 3 doubles per pixel
	std::mutex mu;
};

// With acc, the frame (spp samples numbered from acc->samples) is rendered
// into the library's staging buffer and added to the accumulator's sum.
static int render_common(const rn_scene *scene, const rn_camera *camera, const rn_render_params *params,
                         const float *d_albedo, const float *d_emission, float *d_out, hipStream_t stream,
                         bool caller_stream, float *h_out, rn_accum *acc = nullptr) {
	if (!scene || !camera || !params)
		throw Error("NULL scene, camera or params");
	auto t0 = std::chrono::steady_clock::now();
	Context &c = ctx();
	std::lock_guard<std::mutex> lk(c.mu);
	rn_scene *sc = const_cast<rn_scene *>(scene);
	std::lock_guard<std::mutex> sl(sc->mu);
	DevParams p = make_params(*params, sc, 0);
	if (acc) {
		if (p.width != acc->width || p.height != acc->height)
			throw Error("params width/height differ from the accumulator's");
		if (acc->samples + p.spp > 0x7FFFFFFFll)
			throw Error("accumulator sample count would exceed 2^31 - 1");
		p.sample_base = (uint32_t)acc->samples;
	}
	DevCamera cam = make_camera(*camera, p.width, p.height);
	hipStream_t s = caller_stream ? stream : c.stream;
	bool had = sc->bvh_valid && sc->bvh.width == (params->bvh_width > 0 ? params->bvh_width : 8);
	sc->ensure_bvh(params->bvh_width, s);
	size_t npx = (size_t)p.width * p.height;
	float *out = d_out ? d_out : static_cast<float *>(c.out.reserve(npx * 3 * 4));
	DevCounters *cnt = static_cast<DevCounters *>(c.counters.reserve(sizeof(DevCounters)));
	fill_zero(cnt, sizeof(DevCounters), s);
	DevScene ds = sc->dev(d_albedo, d_emission);
	RN_HIP(hipEventRecord(c.ev0, s));
	WaveStats ws{};
	const bool wave = params->integrator == RN_INTEGRATOR_WAVEFRONT;
	if (wave) {
		WaveItem item{ds, cam, p, out};
		WaveConfig cfg;
		cfg.queue_paths = params->queue_paths;
		cfg.block = params->block_size;
		cfg.time_stages = (params->flags & RN_FLAG_TIME_STAGES) != 0;
		cfg.use_graph = (params->flags & RN_FLAG_USE_GRAPH) != 0;
		cfg.fb_atomics = (params->flags & RN_FLAG_FB_ATOMICS) != 0;
		wavefront_render(&item, 1, cfg, cnt, s, ws);
	} else {
		launch_mega(ds, cam, p, out, cnt, (params->flags & RN_FLAG_COUNT_TRAVERSAL) != 0, s);
	}
	RN_HIP(hipEventRecord(c.ev1, s));
	if (acc)
		launch_accum_add(acc->sum.as<double>(), out, npx * 3, (double)p.spp, s);
	DevCounters hcs;
	download<DevCounters>(&hcs, cnt, 1, s);
	const CounterSet hc = hcs.total();
	if (h_out)
		download<float>(h_out, out, npx * 3, s);
	RN_HIP(hipStreamSynchronize(s));
	if (acc)
		acc->samples += p.spp;
	float ms = 0;
	RN_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
	std::memset(&t_stats, 0, sizeof t_stats);
	t_stats.build_ms = had ? 0.0 : sc->bvh.build_ms;
	t_stats.render_ms = ms;
	t_stats.total_ms = ms_since(t0);
	t_stats.rays_closest = hc.closest;
	t_stats.rays_shadow = hc.shadow;
	t_stats.paths = hc.paths;
	t_stats.trace_kernel_ms = wave ? std::max(ws.extend_ms, ws.shadow_ms) : ms;
	t_stats.trace_kernel_launches = wave ? ws.extend_launches : 1;
	t_stats.bvh_nodes = sc->bvh.n_nodes;
	t_stats.bvh_depth = sc->bvh.depth;
	t_stats.nodes_closest = hc.nodes_closest;
	t_stats.tris_closest = hc.tris_closest;
	t_stats.nodes_shadow = hc.nodes_shadow;
	t_stats.tris_shadow = hc.tris_shadow;
	t_stats.extend_ms = ws.extend_ms;
	t_stats.shade_ms = ws.shade_ms;
	t_stats.shadow_ms = ws.shadow_ms;
	t_stats.iterations = ws.iterations;
	return 1;
}

int rn_render_ex(const rn_scene *scene, const rn_camera *camera, const rn_render_params *params, float *hdr_rgb) {
	RN_TRY
	t_error.clear();
	if (!hdr_rgb)
		throw Error("hdr_rgb is NULL");
	return render_common(scene, camera, params, nullptr, nullptr, nullptr, nullptr, false, hdr_rgb);
	RN_CATCH(0)
}

int rn_render_ex_device(const rn_scene *scene, const rn_camera *camera, const rn_render_params *params,
                        const float *d_albedo, const float *d_emission, float *d_hdr_rgb, void *stream) {
	RN_TRY
	t_error.clear();
	if (!d_hdr_rgb)
		throw Error("d_hdr_rgb is NULL");
	return render_common(scene, camera, params, d_albedo, d_emission, d_hdr_rgb, static_cast<hipStream_t>(stream),
	                     true, nullptr);
	RN_CATCH(0)
}

} // extern "C"

// ------------------------------------------------ allocation-free render --
//
// rn_render_ex_device_noalloc keeps all of its work memory in `scratch`, one
// file-scope device array per GPU, reserved when the library's code object
// is loaded and held until the process ends: the integrators' counters, then
// the wavefront's queues for up to RN_NOALLOC_MAX_QUEUE_PATHS paths in flight,
// its control block and its one-entry scene table. In host code `scratch`
// names a host-side stand-in, never the device memory: the device address is
// looked up once per GPU with hipGetSymbolAddress (with that GPU current)
// and only that address is handed to kernels and fills.
constexpr size_t kCountersBytes = (sizeof(DevCounters) + 255) / 256 * 256;
constexpr size_t kNoAllocPaths = RN_NOALLOC_MAX_QUEUE_PATHS;
// counters + 164 bytes per path + control block, scene table and alignment
// (86 MB in all)
constexpr size_t kNoAllocBytes = kCountersBytes + kNoAllocPaths * (10 * 16 + 4) + 16384;
static_assert(sizeof(DevCounters) <= kCountersBytes, "counter block");

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;     // guards g_scratch_ptr
void *g_scratch_ptr[kMaxDevices] = {};
std::mutex g_noalloc_mu;     // one allocation-free render at a time (host side)

// Device address of `scratch` on the current GPU.
void *scratch_on_current_device() {
	int dev = -1;
	RN_HIP(hipGetDevice(&dev));
	if (dev < 0 || dev >= kMaxDevices)
		throw Error("rn_render_ex_device_noalloc: device index " + std::to_string(dev) + " out of range");
	std::lock_guard<std::mutex> lk(g_scratch_mu);
	if (!g_scratch_ptr[dev]) {
		void *p = nullptr;
		RN_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
		if (!p)
			throw Error("rn_render_ex_device_noalloc: no device address for the work memory");
		g_scratch_ptr[dev] = p;
	}
	return g_scratch_ptr[dev];
}

} // namespace

extern "C" {

int rn_render_ex_device_noalloc(const rn_scene *scene, const rn_camera *camera, const rn_render_params *params,
                                const float *d_albedo, const float *d_emission, float *d_hdr_rgb, void *stream) {
	RN_TRY
	t_error.clear();
	if (!scene || !camera || !params || !d_hdr_rgb)
		throw Error("rn_render_ex_device_noalloc: NULL scene, camera, params or output");
	if (params->flags & (RN_FLAG_USE_GRAPH | RN_FLAG_TIME_STAGES))
		throw Error("rn_render_ex_device_noalloc: RN_FLAG_USE_GRAPH and RN_FLAG_TIME_STAGES are not available "
		            "(both need memory of their own)");
	const bool wave = params->integrator == RN_INTEGRATOR_WAVEFRONT;
	if (params->integrator != RN_INTEGRATOR_WAVEFRONT && params->integrator != RN_INTEGRATOR_MEGAKERNEL)
		throw Error("rn_render_ex_device_noalloc: unknown integrator");
	if (wave && params->queue_paths > (int)kNoAllocPaths)
		throw Error("rn_render_ex_device_noalloc: queue_paths " + std::to_string(params->queue_paths) +
		            " exceeds the maximum of " + std::to_string(kNoAllocPaths) + " paths in flight");
	auto t0 = std::chrono::steady_clock::now();
	std::lock_guard<std::mutex> nl(g_noalloc_mu);
	rn_scene *sc = const_cast<rn_scene *>(scene);
	std::lock_guard<std::mutex> sl(sc->mu);
	const int width = params->bvh_width > 0 ? params->bvh_width : 8;
	if (!sc->uploaded || !sc->bvh_valid || sc->bvh.width != width)
		throw Error("rn_render_ex_device_noalloc: the scene has no device copy or no BVH" + std::to_string(width) +
		            " yet; call rn_scene_build(scene, " + std::to_string(width) +
		            ") first (this entry point neither uploads nor builds)");
	DevParams p = make_params(*params, sc, 0);
	DevCamera cam = make_camera(*camera, p.width, p.height);
	hipStream_t s = static_cast<hipStream_t>(stream);
	char *base = static_cast<char *>(scratch_on_current_device());
	DevCounters *cnt = reinterpret_cast<DevCounters *>(base);
	fill_zero(cnt, sizeof(DevCounters), s);
	DevScene ds = sc->dev(d_albedo, d_emission);
	WaveStats ws{};
	if (wave) {
		WaveItem item{ds, cam, p, d_hdr_rgb};
		WaveConfig cfg;
		cfg.queue_paths = params->queue_paths;
		cfg.block = params->block_size;
		WaveScratch mem;
		mem.base = base + kCountersBytes;
		mem.bytes = kNoAllocBytes - kCountersBytes;
		if (wave_scratch_capacity(mem.bytes, 1) < kNoAllocPaths)
			throw Error("rn_render_ex_device_noalloc: internal error, work memory smaller than its queue");
		wavefront_render(&item, 1, cfg, cnt, s, ws, &mem);
	} else {
		launch_mega(ds, cam, p, d_hdr_rgb, cnt, (params->flags & RN_FLAG_COUNT_TRAVERSAL) != 0, s);
	}
	std::memset(&t_stats, 0, sizeof t_stats);
	t_stats.total_ms = ms_since(t0);
	t_stats.paths = (uint64_t)p.width * p.height * p.spp;
	t_stats.iterations = ws.iterations;
	t_stats.trace_kernel_launches = wave ? ws.extend_launches : 1;
	t_stats.bvh_nodes = sc->bvh.n_nodes;
	t_stats.bvh_depth = sc->bvh.depth;
	return 1;
	RN_CATCH(0)
}

rn_accum *rn_accum_create(int width, int height) {
	RN_TRY
	t_error.clear();
	if (width <= 0 || height <= 0 || (long long)width * height > (1ll << 31) / 4) {
		set_error("width and height must be positive and the image at most 2^29 pixels");
		return nullptr;
	}
	Context &c = ctx();
	std::unique_ptr<rn_accum> a(new rn_accum());
	a->width = width;
	a->height = height;
	std::lock_guard<std::mutex> lk(c.mu);
	a->sum.reserve((size_t)width * height * 3 * sizeof(double));
	fill_zero(a->sum.ptr, (size_t)width * height * 3 * sizeof(double), c.stream);
	RN_HIP(hipStreamSynchronize(c.stream));
	return a.release();
	RN_CATCH(nullptr)
}

void rn_accum_free(rn_accum *acc) { delete acc; }

int rn_accum_reset(rn_accum *acc) {
	RN_TRY
	t_error.clear();
	if (!acc)
		throw Error("accumulator is NULL");
	// lock order everywhere: accumulator, then library context
	std::lock_guard<std::mutex> al(acc->mu);
	Context &c = ctx();
	std::lock_guard<std::mutex> lk(c.mu);
	fill_zero(acc->sum.ptr, (size_t)acc->width * acc->height * 3 * sizeof(double), c.stream);
	RN_HIP(hipStreamSynchronize(c.stream));
	acc->samples = 0;
	return 1;
	RN_CATCH(0)
}

int rn_accum_render(rn_accum *acc, const rn_scene *scene, const rn_camera *camera, const rn_render_params *params) {
	RN_TRY
	t_error.clear();
	if (!acc)
		throw Error("accumulator is NULL");
	std::lock_guard<std::mutex> al(acc->mu);
	return render_common(scene, camera, params, nullptr, nullptr, nullptr, nullptr, false, nullptr, acc);
	RN_CATCH(0)
}

int rn_accum_samples(const rn_accum *acc) { return acc ? (int)acc->samples : 0; }

int rn_accum_resolve_rgba8(const rn_accum *acc, float exposure, float gamma, unsigned char *rgba) {
	RN_TRY
	t_error.clear();
	if (!acc || !rgba)
		throw Error("accumulator and rgba must be non-NULL");
	if (!(exposure > 0.0f) || !(gamma > 0.0f))
		throw Error("exposure and gamma must be positive");
	std::lock_guard<std::mutex> al(const_cast<rn_accum *>(acc)->mu);
	if (acc->samples <= 0)
		throw Error("the accumulator holds no samples");
	Context &c = ctx();
	std::lock_guard<std::mutex> lk(c.mu);
	const size_t npx = (size_t)acc->width * acc->height;
	unsigned char *d = static_cast<unsigned char *>(c.o4.reserve(npx * 4));
	launch_tonemap_rgba8(acc->sum.as<double>(), npx, 1.0 / (double)acc->samples, exposure, gamma, d, c.stream);
	download<unsigned char>(rgba, d, npx * 4, c.stream);
	RN_HIP(hipStreamSynchronize(c.stream));
	return 1;
	RN_CATCH(0)
}

int rn_accum_mean(const rn_accum *acc, float *hdr_rgb) {
	RN_TRY
	t_error.clear();
	if (!acc || !hdr_rgb)
		throw Error("accumulator and hdr_rgb must be non-NULL");
	std::lock_guard<std::mutex> al(const_cast<rn_accum *>(acc)->mu);
	if (acc->samples <= 0)
		throw Error("the accumulator holds no samples");
	Context &c = ctx();
	std::lock_guard<std::mutex> lk(c.mu);
	const size_t n = (size_t)acc->width * acc->height * 3;
	float *d = static_cast<float *>(c.o3.reserve(n * 4));
	launch_mean(acc->sum.as<double>(), n, 1.0 / (double)acc->samples, d, c.stream);
	download<float>(hdr_rgb, d, n, c.stream);
	RN_HIP(hipStreamSynchronize(c.stream));
	return 1;
	RN_CATCH(0)
}

static int adjoint_common(const rn_scene *scene, const rn_camera *camera, const rn_render_params *params,
                          const float *d_albedo, const float *d_emission, const float *d_dl, const float *h_dl,
                          float *d_ga, float *d_ge, float *h_ga, float *h_ge, hipStream_t stream,
                          bool caller_stream) {
	if (!scene || !camera || !params)
		throw Error("NULL scene, camera or params");
	auto t0 = std::chrono::steady_clock::now();
	Context &c = ctx();
	std::lock_guard<std::mutex> lk(c.mu);
	rn_scene *sc = const_cast<rn_scene *>(scene);
	std::lock_guard<std::mutex> sl(sc->mu);
	DevParams p = make_params(*params, sc, 0);
	DevCamera cam = make_camera(*camera, p.width, p.height);
	hipStream_t s = caller_stream ? stream : c.stream;
	sc->ensure_bvh(params->bvh_width, s);
	size_t npx = (size_t)p.width * p.height, n3 = 3 * (size_t)sc->h.n;
	const float *dl = d_dl ? d_dl : upload(c.in0, h_dl, npx * 3, s);
	float *ga = d_ga ? d_ga : static_cast<float *>(c.o0.reserve(n3 * 4));
	float *ge = d_ge ? d_ge : static_cast<float *>(c.o1.reserve(n3 * 4));
	if (!d_ga)
		fill_zero(ga, n3 * 4, s);
	if (!d_ge)
		fill_zero(ge, n3 * 4, s);
	RN_HIP(hipEventRecord(c.ev0, s));
	launch_adjoint(sc->dev(d_albedo, d_emission), cam, p, dl, sc->d_hot_slot.as<int>(), sc->d_hot_prim.as<int>(),
	               sc->n_hot, ga, ge, s);
	RN_HIP(hipEventRecord(c.ev1, s));
	if (h_ga)
		download<float>(h_ga, ga, n3, s);
	if (h_ge)
		download<float>(h_ge, ge, n3, s);
	RN_HIP(hipStreamSynchronize(s));
	float ms = 0;
	RN_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
	std::memset(&t_stats, 0, sizeof t_stats);
	t_stats.render_ms = ms;
	t_stats.total_ms = ms_since(t0);
	t_stats.trace_kernel_ms = ms;
	t_stats.trace_kernel_launches = 1;
	t_stats.bvh_nodes = sc->bvh.n_nodes;
	t_stats.bvh_depth = sc->bvh.depth;
	return 1;
}

int rn_adjoint(const rn_scene *scene, const rn_camera *camera, const rn_render_params *params, const float *dl_dimage,
               float *dl_dalbedo, float *dl_demission) {
	RN_TRY
	t_error.clear();
	if (!dl_dimage)
		throw Error("dl_dimage is NULL");
	return adjoint_common(scene, camera, params, nullptr, nullptr, nullptr, dl_dimage, nullptr, nullptr, dl_dalbedo,
	                      dl_demission, nullptr, false);
	RN_CATCH(0)
}

int rn_adjoint_device(const rn_scene *scene, const rn_camera *camera, const rn_render_params *params,
                      const float *d_albedo, const float *d_emission, const float *d_dl_dimage, float *d_dl_dalbedo,
                      float *d_dl_demission, void *stream) {
	RN_TRY
	t_error.clear();
	if (!d_dl_dimage || !d_dl_dalbedo || !d_dl_demission)
		throw Error("NULL device pointer");
	return adjoint_common(scene, camera, params, d_albedo, d_emission, d_dl_dimage, nullptr, d_dl_dalbedo,
	                      d_dl_demission, nullptr, nullptr, static_cast<hipStream_t>(stream), true);
	RN_CATCH(0)
}

int rn_render_batch(const rn_batch_item *items, int count, const rn_batch_options *opts) {
	RN_TRY
	t_error.clear();
	if (!items || count <= 0)
		throw Error("rn_render_batch: no items");
	auto t0 = std::chrono::steady_clock::now();
	Context &c = ctx();
	std::lock_guard<std::mutex> lk(c.mu);
	rn_batch_options o{};
	if (opts)
		o = *opts;
	for (int i = 0; i < count; i++)
		if (!items[i].scene || !items[i].hdr_rgb)
			throw Error("rn_render_batch: NULL scene or output");
	// Every distinct scene stays locked for the whole batch, as one scene is
	// for a single render: set_materials / set_builder cannot change a scene
	// between its upload and the end of the render. Locks are taken in
	// address order so two batches over the same scenes cannot deadlock.
	std::vector<rn_scene *> distinct((size_t)count);
	for (int i = 0; i < count; i++)
		distinct[(size_t)i] = const_cast<rn_scene *>(items[i].scene);
	std::sort(distinct.begin(), distinct.end(), std::less<rn_scene *>());
	distinct.erase(std::unique(distinct.begin(), distinct.end()), distinct.end());
	std::vector<std::unique_lock<std::mutex>> scene_locks;
	scene_locks.reserve(distinct.size());
	for (rn_scene *sc : distinct)
		scene_locks.emplace_back(sc->mu);
	std::vector<WaveItem> w((size_t)count);
	std::vector<size_t> offs((size_t)count + 1, 0);
	double build_ms = 0.0;
	int max_depth_bvh = 0, nodes = 0;
	for (int i = 0; i < count; i++) {
		const rn_batch_item &it = items[i];
		rn_scene *sc = const_cast<rn_scene *>(it.scene);
		bool had = sc->bvh_valid && sc->bvh.width == (o.bvh_width > 0 ? o.bvh_width : 8);
		sc->ensure_bvh(o.bvh_width, c.stream);
		if (!had)
			build_ms += sc->bvh.build_ms;
		max_depth_bvh = std::max(max_depth_bvh, sc->bvh.depth);
		nodes += sc->bvh.n_nodes;
		w[(size_t)i].scene = sc->dev();
		w[(size_t)i].params = make_params(it.params, sc, 0);
		w[(size_t)i].cam = make_camera(it.camera, it.params.width, it.params.height);
		offs[(size_t)i + 1] = offs[(size_t)i] + (size_t)it.params.width * it.params.height * 3;
	}
	float *out = static_cast<float *>(c.out.reserve(offs.back() * 4));
	for (int i = 0; i < count; i++)
		w[(size_t)i].out = out + offs[(size_t)i];
	DevCounters *cnt = static_cast<DevCounters *>(c.counters.reserve(sizeof(DevCounters)));
	fill_zero(cnt, sizeof(DevCounters), c.stream);
	WaveConfig cfg;
	cfg.queue_paths = o.queue_paths;
	cfg.block = o.block_size;
	cfg.use_graph = o.use_graph != 0;
	cfg.iters_per_graph = o.iters_per_graph;
	cfg.time_stages = !cfg.use_graph;
	WaveStats ws{};
	RN_HIP(hipEventRecord(c.ev0, c.stream));
	wavefront_render(w.data(), count, cfg, cnt, c.stream, ws);
	RN_HIP(hipEventRecord(c.ev1, c.stream));
	DevCounters hcs;
	download<DevCounters>(&hcs, cnt, 1, c.stream);
	const CounterSet hc = hcs.total();
	for (int i = 0; i < count; i++)
		download<float>(items[i].hdr_rgb, out + offs[(size_t)i], offs[(size_t)i + 1] - offs[(size_t)i], c.stream);
	RN_HIP(hipStreamSynchronize(c.stream));
	float ms = 0;
	RN_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
	std::memset(&t_stats, 0, sizeof t_stats);
	t_stats.build_ms = build_ms;
	t_stats.render_ms = ms;
	t_stats.total_ms = ms_since(t0);
	t_stats.rays_closest = hc.closest;
	t_stats.rays_shadow = hc.shadow;
	t_stats.paths = hc.paths;
	t_stats.trace_kernel_ms = std::max(ws.extend_ms, ws.shadow_ms);
	t_stats.trace_kernel_launches = ws.extend_launches;
	t_stats.bvh_nodes = nodes;
	t_stats.bvh_depth = max_depth_bvh;
	t_stats.extend_ms = ws.extend_ms;
	t_stats.shade_ms = ws.shade_ms;
	t_stats.shadow_ms = ws.shadow_ms;
	t_stats.iterations = ws.iterations;
	return 1;
	RN_CATCH(0)
}

int rn_get_last_stats(rn_stats *out) {
	if (!out)
		return 0;
	*out = t_stats;
	return 1;
}

int rn_trace_rays(const rn_scene *scene, int n, const float *orig, const float *dir, const float *tmax,
                  const int *ignore, int query, int tri_test, int *prim, float *t, float *u, float *v) {
	RN_TRY
	t_error.clear();
	if (!scene || n < 0 || !orig || !dir || !prim)
		throw Error("rn_trace_rays: bad arguments");
	if (query == RN_QUERY_CLOSEST && (!t || !u || !v))
		throw Error("rn_trace_rays: closest-hit queries need t, u, v outputs");
	if (n == 0)
		return 1;
	Context &c = ctx();
	std::lock_guard<std::mutex> lk(c.mu);
	rn_scene *sc = const_cast<rn_scene *>(scene);
	std::lock_guard<std::mutex> sl(sc->mu);
	sc->ensure_bvh(8, c.stream);
	size_t N = (size_t)n;
	const float *d_o = upload(c.in0, orig, 3 * N, c.stream);
	const float *d_d = upload(c.in1, dir, 3 * N, c.stream);
	const float *d_t = tmax ? upload(c.in2, tmax, N, c.stream) : nullptr;
	const int *d_i = ignore ? upload(c.in3, ignore, N, c.stream) : nullptr;
	int *d_p = static_cast<int *>(c.o0.reserve(N * 4));
	float *d_tt = static_cast<float *>(c.o1.reserve(N * 4));
	float *d_u = static_cast<float *>(c.o2.reserve(N * 4));
	float *d_v = static_cast<float *>(c.o3.reserve(N * 4));
	launch_trace_rays(sc->dev(), n, d_o, d_d, d_t, d_i, query, tri_test, d_p, d_tt, d_u, d_v, c.stream);
	download<int>(prim, d_p, N, c.stream);
	if (query == RN_QUERY_CLOSEST) {
		download<float>(t, d_tt, N, c.stream);
		download<float>(u, d_u, N, c.stream);
		download<float>(v, d_v, N, c.stream);
	}
	RN_HIP(hipStreamSynchronize(c.stream));
	return 1;
	RN_CATCH(0)
}

int rn_trace_primary_compat(const rn_scene *scene, int width, int height, float cam_z, int *prim, float *t, float *u,
                            float *v, int *texel) {
	RN_TRY
	t_error.clear();
	if (!scene || width <= 0 || height <= 0 || !prim || !t || !u || !v || !texel)
		throw Error("rn_trace_primary_compat: bad arguments");
	Context &c = ctx();
	std::lock_guard<std::mutex> lk(c.mu);
	rn_scene *sc = const_cast<rn_scene *>(scene);
	std::lock_guard<std::mutex> sl(sc->mu);
	sc->ensure_bvh(8, c.stream);
	size_t N = (size_t)width * height;
	int *d_p = static_cast<int *>(c.o0.reserve(N * 4));
	float *d_t = static_cast<float *>(c.o1.reserve(N * 4));
	float *d_u = static_cast<float *>(c.o2.reserve(N * 4));
	float *d_v = static_cast<float *>(c.o3.reserve(N * 4));
	int *d_x = static_cast<int *>(c.o4.reserve(N * 4));
	launch_primary_compat(sc->dev(), sc->compat(), width, height, cam_z, d_p, d_t, d_u, d_v, d_x, c.stream);
	download<int>(prim, d_p, N, c.stream);
	download<float>(t, d_t, N, c.stream);
	download<float>(u, d_u, N, c.stream);
	download<float>(v, d_v, N, c.stream);
	download<int>(texel, d_x, N, c.stream);
	RN_HIP(hipStreamSynchronize(c.stream));
	return 1;
	RN_CATCH(0)
}

int rn_scene_build(rn_scene *scene, int bvh_width, rn_bvh_info *info) {
	RN_TRY
	t_error.clear();
	if (!scene)
		throw Error("scene is NULL");
	if (bvh_width <= 0)
		bvh_width = 8;
	Context &c = ctx();
	std::lock_guard<std::mutex> lk(c.mu);
	std::lock_guard<std::mutex> sl(scene->mu);
	scene->bvh_valid = false;
	scene->ensure_bvh(bvh_width, c.stream);
	if (info) {
		info->nodes = scene->bvh.n_nodes;
		info->depth = scene->bvh.depth;
		info->leaf_tris = scene->bvh.n_tris;
		info->width = scene->bvh.width;
		info->build_ms = scene->bvh.build_ms;
		info->builder = scene->bvh.builder;
	}
	return 1;
	RN_CATCH(0)
}

int rn_scene_set_builder(rn_scene *scene, int builder) {
	RN_TRY
	t_error.clear();
	if (!scene)
		throw Error("scene is NULL");
	if (builder != RN_BVH_PLOC && builder != RN_BVH_LBVH)
		throw Error("unknown BVH builder " + std::to_string(builder));
	std::lock_guard<std::mutex> sl(scene->mu);
	if (scene->builder != builder)
		scene->bvh_valid = false;
	scene->builder = builder;
	return 1;
	RN_CATCH(0)
}

int rn_scene_export_bvh(const rn_scene *scene, void *nodes, int *leaf_prims) {
	RN_TRY
	if (!scene || !scene->bvh_valid)
		throw Error("scene has no built BVH");
	Context &c = ctx();
	std::lock_guard<std::mutex> lk(c.mu);
	if (nodes)
		download<Bvh8Node>(static_cast<Bvh8Node *>(nodes), scene->bvh.nodes.ptr, (size_t)scene->bvh.n_nodes,
		                   c.stream);
	std::vector<float4> tris(3 * (size_t)scene->h.n);
	download<float4>(tris.data(), scene->bvh.tris.ptr, tris.size(), c.stream);
	RN_HIP(hipStreamSynchronize(c.stream));
	if (leaf_prims)
		for (size_t i = 0; i < (size_t)scene->h.n; i++)
			std::memcpy(&leaf_prims[i], &tris[3 * i].w, 4);
	return 1;
	RN_CATCH(0)
}

int rn_lbvh_stages(const float *verts, int n, uint32_t *morton_sorted, int *order, int *left, int *right) {
	RN_TRY
	t_error.clear();
	if (!verts || n <= 0 || !morton_sorted || !order || (n >= 2 && (!left || !right)))
		throw Error("rn_lbvh_stages: bad arguments");
	check_finite(verts, n, "rn_lbvh_stages");
	Context &c = ctx();
	std::lock_guard<std::mutex> lk(c.mu);
	const float *d_v = upload(c.in0, verts, 9 * (size_t)n, c.stream);
	lbvh_stages(d_v, n, morton_sorted, order, left, right, c.stream);
	return 1;
	RN_CATCH(0)
}

int rn_ploc_stages(const float *verts, int n, int *order, int *left, int *right) {
	RN_TRY
	t_error.clear();
	if (!verts || n <= 0 || !order || (n >= 2 && (!left || !right)))
		throw Error("rn_ploc_stages: bad arguments");
	check_finite(verts, n, "rn_ploc_stages");
	Context &c = ctx();
	std::lock_guard<std::mutex> lk(c.mu);
	const float *d_v = upload(c.in0, verts, 9 * (size_t)n, c.stream);
	ploc_stages(d_v, n, order, left, right, c.stream);
	return 1;
	RN_CATCH(0)
}

int rn_radix_sort_pairs(uint32_t *keys, uint32_t *values, int n, int bits) {
	RN_TRY
	t_error.clear();
	if (!keys || !values || n < 0 || bits <= 0 || bits > 32)
		throw Error("rn_radix_sort_pairs: bad arguments");
	Context &c = ctx();
	std::lock_guard<std::mutex> lk(c.mu);
	size_t N = (size_t)n;
	uint32_t *dk = upload(c.in0, keys, N, c.stream);
	uint32_t *dv = upload(c.in1, values, N, c.stream);
	uint32_t *tk = static_cast<uint32_t *>(c.o0.reserve(N * 4 + 16));
	uint32_t *tv = static_cast<uint32_t *>(c.o1.reserve(N * 4 + 16));
	radix_sort_pairs(dk, dv, tk, tv, n, bits, c.o2, c.o3, c.stream);
	download<uint32_t>(keys, dk, N, c.stream);
	download<uint32_t>(values, dv, N, c.stream);
	RN_HIP(hipStreamSynchronize(c.stream));
	return 1;
	RN_CATCH(0)
}

int rn_philox4x32_10(int n, const uint32_t *ctr, const uint32_t *key, uint32_t *out) {
	RN_TRY
	t_error.clear();
	if (n < 0 || !ctr || !key || !out)
		throw Error("rn_philox4x32_10: bad arguments");
	Context &c = ctx();
	std::lock_guard<std::mutex> lk(c.mu);
	size_t N = (size_t)n;
	const uint32_t *dc = upload(c.in0, ctr, 4 * N, c.stream);
	const uint32_t *dk = upload(c.in1, key, 2 * N, c.stream);
	uint32_t *dout = static_cast<uint32_t *>(c.o0.reserve(4 * N * 4));
	launch_philox(n, dc, dk, dout, c.stream);
	download<uint32_t>(out, dout, 4 * N, c.stream);
	RN_HIP(hipStreamSynchronize(c.stream));
	return 1;
	RN_CATCH(0)
}

// ---------------------------------------------------- photon mapping --

rn_photon_map *rn_photon_map_trace(const rn_scene *scene, int num_photons, uint32_t key0, uint32_t key1) {
	RN_TRY
	t_error.clear();
	if (!scene)
		throw Error("rn_photon_map_trace: scene is NULL");
	if (num_photons < 0 || num_photons > (1 << 26))
		throw Error("rn_photon_map_trace: photon count must be in [0, 2^26]");
	auto t0 = std::chrono::steady_clock::now();
	Context &c = ctx();
	std::lock_guard<std::mutex> lk(c.mu);
	rn_scene *sc = const_cast<rn_scene *>(scene);
	std::lock_guard<std::mutex> sl(sc->mu);
	sc->ensure_bvh(8, c.stream);
	std::unique_ptr<rn_photon_map> map(new rn_photon_map());
	RN_HIP(hipEventRecord(c.ev0, c.stream));
	photon_trace(map->m, sc->dev(), sc->compat(), num_photons, key0, key1, c.stream);
	RN_HIP(hipEventRecord(c.ev1, c.stream));
	RN_HIP(hipStreamSynchronize(c.stream));
	float ms = 0;
	RN_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
	std::memset(&t_stats, 0, sizeof t_stats);
	t_stats.render_ms = ms;
	t_stats.total_ms = ms_since(t0);
	return map.release();
	RN_CATCH(nullptr)
}

rn_photon_map *rn_photon_map_create(int n, const float *position, const float *power, const float *direction,
                                    float cell) {
	RN_TRY
	t_error.clear();
	if (n < 0 || (n > 0 && (!position || !power || !direction)))
		throw Error("rn_photon_map_create: bad photon arrays");
	if (!(cell > 0.0f) || !std::isfinite(cell))
		throw Error("rn_photon_map_create: cell side must be positive and finite");
	Context &c = ctx();
	std::lock_guard<std::mutex> lk(c.mu);
	std::unique_ptr<rn_photon_map> map(new rn_photon_map());
	PhotonMap &m = map->m;
	m.n = n;
	size_t N = (size_t)std::max(n, 1);
	std::vector<float4> a(N), b(N);
	std::vector<float> cc(N);
	for (size_t i = 0; i < (size_t)n; i++) {
		a[i] = make_float4(position[3 * i], position[3 * i + 1], position[3 * i + 2], power[3 * i]);
		b[i] = make_float4(direction[3 * i], direction[3 * i + 1], direction[3 * i + 2], power[3 * i + 1]);
		cc[i] = power[3 * i + 2];
	}
	upload(m.a, a.data(), (size_t)n, c.stream);
	upload(m.b, b.data(), (size_t)n, c.stream);
	upload(m.c, cc.data(), (size_t)n, c.stream);
	if (n == 0) {
		m.a.reserve(16);
		m.b.reserve(16);
		m.c.reserve(4);
	}
	photon_map_build(m, cell, c.stream);
	RN_HIP(hipStreamSynchronize(c.stream));
	return map.release();
	RN_CATCH(nullptr)
}

void rn_photon_map_free(rn_photon_map *map) {
	if (!map)
		return;
	try {
		Context &c = ctx();
		std::lock_guard<std::mutex> lk(c.mu);
		RN_HIP(hipStreamSynchronize(c.stream));
		delete map;
	} catch (...) {
	}
}

int rn_photon_map_count(const rn_photon_map *map) { return map ? map->m.n : 0; }

int rn_photon_map_get(const rn_photon_map *map, float *position, float *power, float *direction) {
	RN_TRY
	t_error.clear();
	if (!map)
		throw Error("rn_photon_map_get: map is NULL");
	Context &c = ctx();
	std::lock_guard<std::mutex> lk(c.mu);
	const PhotonMap &m = map->m;
	size_t n = (size_t)m.n;
	std::vector<float4> a(n), b(n);
	std::vector<float> cc(n);
	download<float4>(a.data(), m.a.ptr, n, c.stream);
	download<float4>(b.data(), m.b.ptr, n, c.stream);
	download<float>(cc.data(), m.c.ptr, n, c.stream);
	RN_HIP(hipStreamSynchronize(c.stream));
	for (size_t i = 0; i < n; i++) {
		if (position) {
			position[3 * i] = a[i].x;
			position[3 * i + 1] = a[i].y;
			position[3 * i + 2] = a[i].z;
		}
		if (power) {
			power[3 * i] = a[i].w;
			power[3 * i + 1] = b[i].w;
			power[3 * i + 2] = cc[i];
		}
		if (direction) {
			direction[3 * i] = b[i].x;
			direction[3 * i + 1] = b[i].y;
			direction[3 * i + 2] = b[i].z;
		}
	}
	return 1;
	RN_CATCH(0)
}

int rn_photon_map_gather(const rn_photon_map *map, int m, const float *points, const float *normals, float radius,
                         float *irradiance) {
	RN_TRY
	t_error.clear();
	if (!map || m < 0 || (m > 0 && (!points || !normals || !irradiance)))
		throw Error("rn_photon_map_gather: bad arguments");
	if (!(radius >= 0.0f) || !std::isfinite(radius))
		throw Error("rn_photon_map_gather: radius must be finite and non-negative");
	if (m == 0)
		return 1;
	Context &c = ctx();
	std::lock_guard<std::mutex> lk(c.mu);
	const float *dp = upload(c.in0, points, 3 * (size_t)m, c.stream);
	const float *dn = upload(c.in1, normals, 3 * (size_t)m, c.stream);
	float *dout = static_cast<float *>(c.o0.reserve(3 * (size_t)m * 4));
	launch_photon_gather(map->m, m, dp, dn, radius, dout, c.stream);
	download<float>(irradiance, dout, 3 * (size_t)m, c.stream);
	RN_HIP(hipStreamSynchronize(c.stream));
	return 1;
	RN_CATCH(0)
}

static void check_image(const rn_scene *scene, int width, int height, const unsigned char *rgba, const char *who) {
	if (!scene || !rgba || width <= 0 || height <= 0)
		throw Error(std::string(who) + ": NULL scene/buffer or non-positive size");
	if ((long long)width * height > (1ll << 29))
		throw Error(std::string(who) + ": image too large");
}

int rn_render_photon(const rn_scene *scene, const rn_photon_map *map, int width, int height, float cam_z,
                     int gather_rays, uint32_t key0, uint32_t key1, unsigned char *rgba) {
	RN_TRY
	t_error.clear();
	check_image(scene, width, height, rgba, "rn_render_photon");
	if (!map)
		throw Error("rn_render_photon: map is NULL");
	if (gather_rays < 0 || gather_rays > 4096)
		throw Error("rn_render_photon: gather_rays must be in [0, 4096]");
	auto t0 = std::chrono::steady_clock::now();
	Context &c = ctx();
	std::lock_guard<std::mutex> lk(c.mu);
	rn_scene *sc = const_cast<rn_scene *>(scene);
	std::lock_guard<std::mutex> sl(sc->mu);
	sc->ensure_bvh(8, c.stream);
	unsigned char *d_rgba = static_cast<unsigned char *>(c.out.reserve((size_t)width * height * 4));
	RN_HIP(hipEventRecord(c.ev0, c.stream));
	launch_photon_render(sc->dev(), sc->compat(), map->m, width, height, cam_z, gather_rays, key0, key1, d_rgba,
	                     c.stream);
	RN_HIP(hipEventRecord(c.ev1, c.stream));
	download<unsigned char>(rgba, d_rgba, (size_t)width * height * 4, c.stream);
	RN_HIP(hipStreamSynchronize(c.stream));
	float ms = 0;
	RN_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
	std::memset(&t_stats, 0, sizeof t_stats);
	t_stats.render_ms = ms;
	t_stats.total_ms = ms_since(t0);
	t_stats.trace_kernel_ms = ms;
	t_stats.trace_kernel_launches = 1;
	return 1;
	RN_CATCH(0)
}

int rn_render_bidirectional(const rn_scene *scene, int width, int height, float cam_z, int samples, float exposure,
                            uint32_t key0, uint32_t key1, unsigned char *rgba) {
	RN_TRY
	t_error.clear();
	check_image(scene, width, height, rgba, "rn_render_bidirectional");
	if (samples <= 0 || samples > (1 << 20))
		throw Error("rn_render_bidirectional: samples must be in [1, 2^20]");
	auto t0 = std::chrono::steady_clock::now();
	Context &c = ctx();
	std::lock_guard<std::mutex> lk(c.mu);
	rn_scene *sc = const_cast<rn_scene *>(scene);
	std::lock_guard<std::mutex> sl(sc->mu);
	sc->ensure_bvh(8, c.stream);
	unsigned char *d_rgba = static_cast<unsigned char *>(c.out.reserve((size_t)width * height * 4));
	RN_HIP(hipEventRecord(c.ev0, c.stream));
	launch_bdpt_render(sc->dev(), sc->compat(), width, height, cam_z, samples, exposure, key0, key1, d_rgba, c.stream);
	RN_HIP(hipEventRecord(c.ev1, c.stream));
	download<unsigned char>(rgba, d_rgba, (size_t)width * height * 4, c.stream);
	RN_HIP(hipStreamSynchronize(c.stream));
	float ms = 0;
	RN_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
	std::memset(&t_stats, 0, sizeof t_stats);
	t_stats.render_ms = ms;
	t_stats.total_ms = ms_since(t0);
	t_stats.trace_kernel_ms = ms;
	t_stats.trace_kernel_launches = 1;
	return 1;
	RN_CATCH(0)
}

} // extern "C"
