// C-ABI of libddgi (include/ddgi.h): the IrradianceProbes surface of VulkanExp
// (src/IrradianceProbes.hpp:16-129, src/IrradianceProbes.cpp:12-624) over HIP device memory.
//
// One handle owns: the probe grid parameters, two atlas pairs (exposed / next, swapped after each
// update instead of the reference's work->exposed image copy), probe states, the ray buffers, the
// per-frame blend weights and the scene (BVH + shading records).  Per-frame work is enqueued on
// the caller's stream; nothing here synchronises unless the call returns host data.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "handle.hpp"

using namespace ddgi;

namespace ddgi {
namespace host {

constexpr double kPiD = 3.1415926538;  // src/shaders/common.glsl:4
constexpr uint32_t kStateMagic = 0x49474444u;  // "DDGI"
constexpr uint32_t kStateVersion = 1u;

thread_local std::string g_last_error;

ddgi_status fail(Handle* h, ddgi_status s, const std::string& msg) {
	if(h) h->err = msg;
	g_last_error = msg;
	return s;
}

void record_error(const std::string& msg) { g_last_error = msg; }

bool finite_all(const float* v, int n) {
	for(int i = 0; i < n; ++i)
		if(!std::isfinite(v[i])) return false;
	return true;
}

ddgi_status validate_params(const ddgi_grid_params* p, uint32_t max_rays, std::string& why) {
	for(int k = 0; k < 3; ++k) {
		if(p->resolution[k] < 2) {
			why = "resolution must be >= 2 on every axis (probeGridCellSize divides by resolution - 1)";
			return DDGI_ERROR_INVALID_ARGUMENT;
		}
		if(!(p->extent_max[k] > p->extent_min[k])) {
			why = "extent_max must exceed extent_min on every axis";
			return DDGI_ERROR_INVALID_ARGUMENT;
		}
	}
	if(!finite_all(p->extent_min, 3) || !finite_all(p->extent_max, 3) || !std::isfinite(p->hysteresis) ||
	   !std::isfinite(p->depth_sharpness) || !std::isfinite(p->shadow_bias)) {
		why = "non-finite grid parameter";
		return DDGI_ERROR_INVALID_ARGUMENT;
	}
	if(p->rays_per_probe == 0 || p->rays_per_probe > max_rays) {
		why = "rays_per_probe must be in [1, max_rays_per_probe]";
		return DDGI_ERROR_INVALID_ARGUMENT;
	}
	if(p->color_res != 8 || p->depth_res != 16) {
		why = "only colorRes 8 and depthRes 16 are supported (the reference bakes them into its shaders)";
		return DDGI_ERROR_UNSUPPORTED;
	}
	const uint64_t P = uint64_t(p->resolution[0]) * p->resolution[1] * p->resolution[2];
	if(P > (1u << 24)) {
		why = "more than 2^24 probes";
		return DDGI_ERROR_UNSUPPORTED;
	}
	return DDGI_OK;
}

GridConst make_grid_const(const Handle& h) {
	GridConst g{};
	const ddgi_grid_params& p = h.params;
	for(int k = 0; k < 3; ++k) {
		g.emin[k] = p.extent_min[k];
		g.cell[k] = (p.extent_max[k] - p.extent_min[k]) / float(p.resolution[k] - 1);
		g.res[k] = p.resolution[k];
	}
	g.cell_len = std::sqrt(g.cell[0] * g.cell[0] + g.cell[1] * g.cell[1] + g.cell[2] * g.cell[2]);
	const float ex = p.extent_max[0] - p.extent_min[0], ey = p.extent_max[1] - p.extent_min[1],
				ez = p.extent_max[2] - p.extent_min[2];
	g.tmax_primary = std::sqrt(ex * ex + ey * ey + ez * ez);
	g.hysteresis = h.uploaded_hysteresis;
	g.sharpness = p.depth_sharpness;
	g.shadow_bias = p.shadow_bias;
	g.n = p.rays_per_probe;
	g.probe_count = h.P;
	return g;
}

// sphericalFibonacci (src/shaders/irradiance.glsl:52-64) evaluated in double, rounded to float.
void fib_host(uint32_t i, uint32_t n, float out[3]) {
	const double di = double(i), dn = double(n);
	const double PHI = std::sqrt(5.0) * 0.5 + 0.5;
	const double a = di * (PHI - 1.0);
	const double phi = 2.0 * kPiD * (a - std::floor(a));
	const double ct = 1.0 - (2.0 * di + 1.0) * (1.0 / dn);
	double s2 = 1.0 - ct * ct;
	s2 = s2 < 0.0 ? 0.0 : (s2 > 1.0 ? 1.0 : s2);
	const double st = std::sqrt(s2);
	out[0] = float(std::cos(phi) * st);
	out[1] = float(std::sin(phi) * st);
	out[2] = float(ct);
}

uint32_t morton2(uint32_t x, uint32_t y) {
	uint32_t r = 0;
	for(int b = 0; b < 8; ++b) r |= ((x >> b) & 1u) << (2 * b) | ((y >> b) & 1u) << (2 * b + 1);
	return r;
}

// Upload fib(i, n) for the frame rays, the 512 init rays, and the thread->ray order in force until
// the first update (directions grouped by octahedral cell, so a wavefront traces a compact cone of
// directions); each update replaces it by its frame's order (rays by world elevation).
ddgi_status ensure_fib(Handle* h) {
	const uint32_t n = h->params.rays_per_probe;
	if(h->fib_n == n) return DDGI_OK;
	std::vector<float> f(size_t(kMaxRays) * 4, 0.0f), f512(512 * 4, 0.0f);
	for(uint32_t i = 0; i < n; ++i) fib_host(i, n, &f[i * 4]);
	for(uint32_t i = 0; i < 512; ++i) fib_host(i, n, &f512[i * 4]);
	std::vector<uint32_t> key(n);
	for(uint32_t i = 0; i < n; ++i) {
		const double x = f[i * 4], y = f[i * 4 + 1], z = f[i * 4 + 2];
		const double l1 = std::fabs(x) + std::fabs(y) + std::fabs(z);
		double u = x / l1, v = y / l1;
		if(z < 0.0) {
			const double nu = (1.0 - std::fabs(v)) * (u >= 0.0 ? 1.0 : -1.0);
			const double nv = (1.0 - std::fabs(u)) * (v >= 0.0 ? 1.0 : -1.0);
			u = nu;
			v = nv;
		}
		const uint32_t qu = std::min(15u, uint32_t((u * 0.5 + 0.5) * 16.0));
		const uint32_t qv = std::min(15u, uint32_t((v * 0.5 + 0.5) * 16.0));
		key[i] = morton2(qu, qv);
	}
	std::vector<uint16_t> perm(kMaxRays);
	std::iota(perm.begin(), perm.begin() + n, uint16_t(0));
	std::stable_sort(perm.begin(), perm.begin() + n, [&](uint16_t a, uint16_t b) { return key[a] < key[b]; });
	for(uint32_t i = n; i < kMaxRays; ++i) perm[i] = uint16_t(i);
	HIP_TRY(h, hipMemcpy(h->fib, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice));
	HIP_TRY(h, hipMemcpy(h->fib512, f512.data(), f512.size() * sizeof(float), hipMemcpyHostToDevice));
	HIP_TRY(h, hipMemcpy(h->perm, perm.data(), perm.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
	h->fib_n = n;
	return DDGI_OK;
}

ddgi_status set_device(Handle* h) {
	int cur = -1;
	HIP_TRY(h, hipGetDevice(&cur));
	if(cur != h->device) HIP_TRY(h, hipSetDevice(h->device));
	return DDGI_OK;
}

void free_scene(Handle* h) {
	if(h->move_pending) (void)hipEventSynchronize(h->move_done);
	h->move_pending = false;
	(void)hipHostFree(h->move_host);
	h->move_host = nullptr;
	h->move_xf = nullptr;
	h->move_meta = nullptr;
	h->obj_tris = nullptr;
	h->slot_box = nullptr;
	h->rel_xf = nullptr;
	h->build_xf.clear();
	h->level_nodes = nullptr;
	h->instances.clear();
	for(auto& b : h->scene_bufs) (void)hipFree(b.p);
	h->scene_bufs.clear();
	h->has_scene = false;
	h->sd = SceneDev{};
}

// A scene buffer of `count` elements, filled from `src` when given.
template <class T>
ddgi_status upload(Handle* h, const T* src, size_t count, const T** dst) {
	DeviceBuffer b;
	b.bytes = std::max<size_t>(count * sizeof(T), 16);
	HIP_TRY(h, hipMalloc(&b.p, b.bytes));
	h->scene_bufs.push_back(b);
	if(count && src) HIP_TRY(h, hipMemcpy(b.p, src, count * sizeof(T), hipMemcpyHostToDevice));
	if(dst) *dst = static_cast<const T*>(b.p);
	return DDGI_OK;
}

// Row-major 3x4 identity transforms.
std::vector<float> identity_xf(size_t count) {
	std::vector<float> v(std::max<size_t>(count, 1) * 12, 0.0f);
	for(size_t k = 0; k < count; ++k) v[k * 12 + 0] = v[k * 12 + 5] = v[k * 12 + 10] = 1.0f;
	return v;
}

// out = M * B^-1 (row-major 3x4 affine transforms, in double): the motion of an instance since the
// BVH was built with transform B.  Within 1e-12 of the identity it is the identity exactly, so an
// instance moved back bounds its pieces as built.
void relative_transform(const float* M, const float* B, float* out) {
	const double a[9] = {B[0], B[1], B[2], B[4], B[5], B[6], B[8], B[9], B[10]};
	const double c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8], c02 = a[3] * a[7] - a[4] * a[6];
	const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
	double inv[9] = {c00 / det,
					 (a[2] * a[7] - a[1] * a[8]) / det,
					 (a[1] * a[5] - a[2] * a[4]) / det,
					 c01 / det,
					 (a[0] * a[8] - a[2] * a[6]) / det,
					 (a[2] * a[3] - a[0] * a[5]) / det,
					 c02 / det,
					 (a[1] * a[6] - a[0] * a[7]) / det,
					 (a[0] * a[4] - a[1] * a[3]) / det};
	const double tb[3] = {B[3], B[7], B[11]};
	double r[12];
	bool ident = true;
	for(int i = 0; i < 3; ++i) {
		for(int j = 0; j < 3; ++j) {
			r[i * 4 + j] = double(M[i * 4 + 0]) * inv[0 * 3 + j] + double(M[i * 4 + 1]) * inv[1 * 3 + j] +
						   double(M[i * 4 + 2]) * inv[2 * 3 + j];
			ident = ident && std::fabs(r[i * 4 + j] - (i == j ? 1.0 : 0.0)) < 1e-12;
		}
		// translation: M's offset minus M's linear part applied to B^-1's offset (B^-1 x = inv (x - tb))
		double t = M[i * 4 + 3];
		for(int j = 0; j < 3; ++j) t -= r[i * 4 + j] * tb[j];
		r[i * 4 + 3] = t;
		ident = ident && std::fabs(t) < 1e-12 * (1.0 + std::fabs(tb[0]) + std::fabs(tb[1]) + std::fabs(tb[2]));
	}
	for(int q = 0; q < 12; ++q) out[q] = ident ? ((q == 0 || q == 5 || q == 10) ? 1.0f : 0.0f) : float(r[q]);
}

// Device record of an instance with row-major 3x4 transform M: the normal matrix (inverse
// transpose of the linear part, from cofactors in double), the offset entry, the global id of its
// first triangle and its mask.  False for a singular transform.
bool instance_dev(const ddgi_instance& in, const float* M, uint32_t prim_base, InstanceDev& out) {
	const double a[9] = {M[0], M[1], M[2], M[4], M[5], M[6], M[8], M[9], M[10]};
	const double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) + a[2] * (a[3] * a[7] - a[4] * a[6]);
	if(det == 0.0) return false;
	const double cof[9] = {a[4] * a[8] - a[5] * a[7], -(a[3] * a[8] - a[5] * a[6]), a[3] * a[7] - a[4] * a[6],
						   -(a[1] * a[8] - a[2] * a[7]), a[0] * a[8] - a[2] * a[6], -(a[0] * a[7] - a[1] * a[6]),
						   a[1] * a[5] - a[2] * a[4], -(a[0] * a[5] - a[2] * a[3]), a[0] * a[4] - a[1] * a[3]};
	for(int q = 0; q < 9; ++q) out.nm[q] = float(cof[q] / det);
	out.custom_index = in.custom_index;
	out.prim_base = prim_base;
	out.mask = in.mask;
	return true;
}

ddgi_status make_rotation(const float* o, Rot& R) {
	if(!o || !finite_all(o, 9)) return DDGI_ERROR_INVALID_ARGUMENT;
	std::memcpy(R.r, o, sizeof(R.r));
	return DDGI_OK;
}

void tiled_to_image(const float* tiled, float* img, const Handle& h, uint32_t res, uint32_t ch) {
	const uint32_t rx = h.params.resolution[0], ry = h.params.resolution[1], rz = h.params.resolution[2];
	const uint32_t W = res * rx * ry;
	for(uint32_t p = 0; p < h.P; ++p) {
		const uint32_t gx = p % rx, gy = (p % (rx * ry)) / rx, gz = p / (rx * ry);
		const uint32_t ox = res * (gy * rx + gx), oy = res * gz;
		for(uint32_t y = 0; y < res; ++y)
			std::memcpy(&img[(size_t(oy + y) * W + ox) * ch], &tiled[(size_t(p) * res * res + y * res) * ch],
						sizeof(float) * res * ch);
	}
	(void)rz;
}

void image_to_tiled(const float* img, float* tiled, const Handle& h, uint32_t res, uint32_t ch) {
	const uint32_t rx = h.params.resolution[0], ry = h.params.resolution[1];
	const uint32_t W = res * rx * ry;
	for(uint32_t p = 0; p < h.P; ++p) {
		const uint32_t gx = p % rx, gy = (p % (rx * ry)) / rx, gz = p / (rx * ry);
		const uint32_t ox = res * (gy * rx + gx), oy = res * gz;
		for(uint32_t y = 0; y < res; ++y)
			std::memcpy(&tiled[(size_t(p) * res * res + y * res) * ch], &img[(size_t(oy + y) * W + ox) * ch],
						sizeof(float) * res * ch);
	}
}

void free_diff(Handle* h) {
	for(auto& b : h->diff.bufs) (void)hipFree(b.p);
	h->diff = Handle::Diff{};
}

void free_reflection_scratch(Handle* h) {
	void* bufs[] = {h->refl_hits, h->refl_list, h->refl_keys, h->refl_bkeys, h->refl_blist, h->refl_count, h->refl_counts};
	for(void* b : bufs) (void)hipFree(b);
	h->refl_hits = nullptr;
	h->refl_list = h->refl_keys = h->refl_bkeys = h->refl_blist = h->refl_count = h->refl_counts = nullptr;
	h->refl_cap = 0;
}

void destroy_handle(Handle* h) {
	free_diff(h);
	free_scene(h);
	for(int i = 0; i < 2; ++i) {
		(void)hipFree(h->irr[i]);
		(void)hipFree(h->dep[i]);
	}
	free_reflection_scratch(h);
	void* bufs[] = {h->states, h->indices, h->sel,     h->rays,         h->hits,       h->shade_list, h->shade_count,
					h->miss_list, h->shade_keys, h->sort_keys, h->sort_list, h->sort_counts, h->fib,   h->perm,       h->fib512,
					h->dirs,   h->wt,      h->inv_sum,    h->w_hi,        h->w_lo,       h->visit};
	for(void* b : bufs) (void)hipFree(b);
	for(auto& e : h->ev)
		if(e) (void)hipEventDestroy(e);
	if(h->move_done) (void)hipEventDestroy(h->move_done);
	delete h;
}

// `order`: the ray order of the `rays` records (the traversal's thread order for ddgi_update,
// nullptr = natural order for caller-provided buffers).
// `R` is null when h->dirs already holds this frame's directions (the update wrote them for its trace).
ddgi_status enqueue_blend_and_publish(Handle* h, const Rot* R, const float4* rays, const uint16_t* order,
									  uint32_t max_slots, bool all, uint32_t flags, hipStream_t s) {
	const GridConst g = make_grid_const(*h);
	const uint32_t mode = (flags & DDGI_UPDATE_BLEND_BF16) ? kBlendBF16 : ((flags & DDGI_UPDATE_BLEND_FP16) ? kBlendFP16 : kBlendFP32);
	if(R) HIP_TRY(h, launch_frame_dirs(*R, h->fib, h->dirs, g.n, s));
	HIP_TRY(h, launch_blend_weights(g, h->dirs, order, h->wt, h->inv_sum, h->w_hi, h->w_lo, mode, s));
	const int nxt = h->cur ^ 1;
	if(!all) {
		// probes that are not updated keep their tiles: start the next atlas as a copy
		HIP_TRY(h, hipMemcpyAsync(h->irr[nxt], h->irr[h->cur], size_t(h->P) * 64 * sizeof(float4), hipMemcpyDeviceToDevice, s));
		HIP_TRY(h, hipMemcpyAsync(h->dep[nxt], h->dep[h->cur], size_t(h->P) * 256 * sizeof(float2), hipMemcpyDeviceToDevice, s));
	}
	BlendArgs b{};
	b.grid = g;
	b.indices = h->indices;
	b.count = &h->sel->count;
	b.rays = rays;
	b.wt = h->wt;
	b.inv_sum = h->inv_sum;
	b.w_hi16 = h->w_hi;
	b.w_lo16 = h->w_lo;
	b.states = h->states;
	b.irr_prev = h->irr[h->cur];
	b.dep_prev = h->dep[h->cur];
	b.irr_out = h->irr[nxt];
	b.dep_out = h->dep[nxt];
	b.max_slots = max_slots;
	// the FP32 blend in 64-probe workgroups once they fill the device, else in 16-probe blocks
	HIP_TRY(h, launch_blend(b, mode, s, h->cus > 0 && max_slots >= 64u * h->cus));
	h->cur = nxt;  // publish (reference: vkCmdCopyImage work -> exposed, src/IrradianceProbes.cpp:530-576)
	return DDGI_OK;
}

// src/IrradianceProbes.cpp:462-476 with its order: in the first branch the uniforms are uploaded
// BEFORE the value moves, so the change reaches the kernels one ramp step later.  The distance is
// a float compared with the double 0.05, as there (0.05f would differ when it is exactly 0.05f).
void advance_hysteresis(Handle* h, uint32_t updated) {
	if(h->updated_probes >= h->P) {
		const float target = h->opts.target_hysteresis;
		if(double(std::fabs(h->params.hysteresis - target)) > 0.05) {
			h->uploaded_hysteresis = h->params.hysteresis;
			h->params.hysteresis += 0.1f * (target - h->params.hysteresis);
		} else if(h->params.hysteresis != target) {
			h->params.hysteresis = target;
			h->uploaded_hysteresis = target;
		}
		h->updated_probes -= h->P;
	}
	h->updated_probes += updated;
}

// ---- device scratch of the allocation-free reflection pass (ddgi_reflection_noalloc) ----
// One array per GPU the library runs on, held until the process ends: 32 bytes per texel for up to
// kNoAllocTexels texels (hit record, queue entry, key, the bucketed key and entry of the sort),
// then the queue length and the sort counters.  Every call on a GPU shares that GPU's copy.
constexpr uint64_t kNoAllocTexels = uint64_t(1) << 21;  // 2,097,152 texels (a 1920 x 1080 view fits)
constexpr size_t kNoAllocCountOffset = size_t(kNoAllocTexels) * 32;
constexpr size_t kNoAllocCountersOffset = kNoAllocCountOffset + 256;
constexpr size_t kNoAllocBytes = kNoAllocCountersOffset + size_t(kShadeSortCounts) * 4;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxNoAllocDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kMaxNoAllocDevices] = {};

// The device address of `scratch` on the handle's device (current), looked up once per device.
// In host code the array's name is only a stand-in: kernels and copies take this address.
ddgi_status noalloc_scratch(Handle* h, unsigned char** out) {
	if(h->device < 0 || h->device >= kMaxNoAllocDevices)
		return fail(h, DDGI_ERROR_UNSUPPORTED, "device ordinal above the allocation-free scratch table");
	std::lock_guard<std::mutex> lk(g_noalloc_mu);
	if(!g_noalloc_ptr[h->device]) {
		void* ptr = nullptr;
		const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
		if(e != hipSuccess || !ptr)
			return fail(h, DDGI_ERROR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
		g_noalloc_ptr[h->device] = ptr;
	}
	*out = static_cast<unsigned char*>(g_noalloc_ptr[h->device]);
	return DDGI_OK;
}

// Scratch of one reflection chunk of up to `cap` texels.
struct ReflScratch {
	float4* hits;
	uint32_t *list, *keys, *bkeys, *blist, *count, *counts;
};

// reflection.rgen:101-175 over `count` G-buffer texels in chunks of at most `chunk_cap`, with the
// scratch `sc` (see ddgi_reflection).
ddgi_status enqueue_reflection(Handle* h, const ddgi_gbuffer_texel* gbuffer, uint64_t count, const float view_origin[3],
							   const float* noise, float* out_rgba, const ReflScratch& sc, uint64_t chunk_cap,
							   hipStream_t s) {
	ReflectionArgs a{};
	a.scene = h->sd;
	a.grid = make_grid_const(*h);
	a.light = h->light;
	a.states = h->states;
	a.irr = h->irr[h->cur];
	a.dep = h->dep[h->cur];
	a.gbuffer = gbuffer;
	a.noise = reinterpret_cast<const float2*>(noise);
	a.out = reinterpret_cast<float4*>(out_rgba);
	a.count = count;
	for(int k = 0; k < 3; ++k) a.origin[k] = view_origin[k];
	uint32_t key_shift = 0, key_bits = 0;
	shade_key_layout(a.grid, &key_shift, &key_bits);
	a.hits = sc.hits;
	a.list = sc.list;
	a.keys = sc.keys;
	a.list_count = sc.count;
	a.key_shift = key_shift;
	for(uint64_t first = 0; first < count; first += chunk_cap) {
		a.first = first;
		a.chunk = uint32_t(std::min<uint64_t>(chunk_cap, count - first));
		HIP_TRY(h, launch_clear(sc.count, 1, s));
		HIP_TRY(h, launch_reflection_trace(a, s));
		HIP_TRY(h, launch_key_sort(SortJob{sc.keys, sc.list, sc.count, sc.counts, sc.bkeys, sc.blist, nullptr, 0u, 1u, 0u, 0u},
								   a.chunk, 3 * key_bits, s));
		HIP_TRY(h, launch_reflection_shade(a, s));
	}
	return DDGI_OK;
}

}  // namespace host
}  // namespace ddgi

using namespace ddgi::host;

extern "C" {

uint32_t ddgi_abi_version(void) { return DDGI_ABI_VERSION; }

const char* ddgi_status_string(ddgi_status s) {
	switch(s) {
		case DDGI_OK: return "ok";
		case DDGI_ERROR_INVALID_ARGUMENT: return "invalid argument";
		case DDGI_ERROR_OUT_OF_MEMORY: return "out of memory";
		case DDGI_ERROR_HIP: return "HIP runtime error";
		case DDGI_ERROR_UNSUPPORTED: return "unsupported";
		case DDGI_ERROR_NO_SCENE: return "no scene";
		case DDGI_ERROR_INVALID_STATE: return "invalid state";
		case DDGI_ERROR_NO_DEVICE: return "no device";
	}
	return "unknown status";
}

const char* ddgi_last_error(void* handle) {
	if(handle) return H(handle)->err.c_str();
	return g_last_error.c_str();
}

ddgi_status ddgi_create(const ddgi_grid_params* params, const ddgi_options* options, void** out_handle) {
	if(!params || !out_handle) return fail(nullptr, DDGI_ERROR_INVALID_ARGUMENT, "null argument");
	*out_handle = nullptr;
	ddgi_options opts{};
	if(options) opts = *options;
	if(opts.max_rays_per_probe == 0) opts.max_rays_per_probe = kMaxRays;
	if(opts.max_rays_per_probe > kMaxRays) return fail(nullptr, DDGI_ERROR_UNSUPPORTED, "max_rays_per_probe above 512");
	if(opts.target_hysteresis == 0.0f) opts.target_hysteresis = 0.98f;  // src/IrradianceProbes.hpp:46
	std::string why;
	ddgi_status st = validate_params(params, opts.max_rays_per_probe, why);
	if(st != DDGI_OK) return fail(nullptr, st, why);
	int ndev = 0;
	if(hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(nullptr, DDGI_ERROR_NO_DEVICE, "no HIP device");
	Handle* h = new(std::nothrow) Handle;
	if(!h) return fail(nullptr, DDGI_ERROR_OUT_OF_MEMORY, "host allocation failed");
	h->params = *params;
	h->uploaded_hysteresis = params->hysteresis;
	h->opts = opts;
	if(opts.device >= 0) {
		if(opts.device >= ndev) {
			delete h;
			return fail(nullptr, DDGI_ERROR_INVALID_ARGUMENT, "device ordinal out of range");
		}
		h->device = opts.device;
	} else if(hipGetDevice(&h->device) != hipSuccess) {
		delete h;
		return fail(nullptr, DDGI_ERROR_NO_DEVICE, "hipGetDevice failed");
	}
	hipDeviceProp_t prop;
	if(hipGetDeviceProperties(&prop, h->device) != hipSuccess || std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
		delete h;
		return fail(nullptr, DDGI_ERROR_NO_DEVICE, "device is not gfx950");
	}
	h->cus = uint32_t(prop.multiProcessorCount);
	auto cleanup_fail = [&](ddgi_status s) {
		std::string m = h->err;
		destroy_handle(h);
		return fail(nullptr, s, m);
	};
	if((st = set_device(h)) != DDGI_OK) return cleanup_fail(st);
	{  // the device address of the allocation-free scratch, once per device
		unsigned char* ptr = nullptr;
		if((st = noalloc_scratch(h, &ptr)) != DDGI_OK) return cleanup_fail(st);
	}
	h->P = uint32_t(params->resolution[0]) * params->resolution[1] * params->resolution[2];
	h->max_rays = opts.max_rays_per_probe;
	const size_t P = h->P, R = h->max_rays;
	hipError_t e = hipSuccess;
	for(int i = 0; i < 2 && e == hipSuccess; ++i) {
		e = dalloc(&h->irr[i], P * 64 * sizeof(float4));
		if(e == hipSuccess) e = dalloc(&h->dep[i], P * 256 * sizeof(float2));
	}
	if(e == hipSuccess) e = dalloc(&h->states, P * 4);
	if(e == hipSuccess) e = dalloc(&h->indices, P * 4);
	if(e == hipSuccess) e = dalloc(&h->sel, sizeof(SelectState));
	if(e == hipSuccess) e = dalloc(&h->rays, P * R * sizeof(float4));
	if(e == hipSuccess) e = dalloc(&h->hits, P * R * sizeof(float4));
	// shade list and keys: the primary pass's segments hold up to kShadeSegments * kPrimaryBlock
	// entries beyond P * R (each segment is sized for whole workgroups)
	if(e == hipSuccess) e = dalloc(&h->shade_list, (P * R + kShadeSegments * kPrimaryBlock) * sizeof(uint32_t));
	if(e == hipSuccess) e = dalloc(&h->shade_count, kListCounterBytes);
	if(e == hipSuccess) e = dalloc(&h->shade_keys, (P * R + kShadeSegments * kPrimaryBlock) * sizeof(uint32_t));
	if(e == hipSuccess) e = dalloc(&h->sort_keys, P * R * sizeof(uint32_t));
	if(e == hipSuccess) e = dalloc(&h->sort_list, P * R * sizeof(uint32_t));
	if(e == hipSuccess) e = dalloc(&h->sort_counts, size_t(kShadeSortCounts) * sizeof(uint32_t));
	if(e == hipSuccess) e = dalloc(&h->fib, kMaxRays * sizeof(float4));
	if(e == hipSuccess) e = dalloc(&h->perm, kMaxRays * sizeof(uint16_t));
	if(e == hipSuccess) e = dalloc(&h->fib512, 512 * sizeof(float4));
	if(e == hipSuccess) e = dalloc(&h->dirs, 512 * sizeof(float4));
	if(e == hipSuccess) e = dalloc(&h->wt, size_t(kWeightRows) * kMaxRays * sizeof(float));
	if(e == hipSuccess) e = dalloc(&h->inv_sum, kBlendTexels * sizeof(float));
	// 16-bit MFMA weights: 256 padded rows x 512 rays
	if(e == hipSuccess) e = dalloc(&h->w_hi, size_t(kWeightRows) * kMaxRays * 2);
	if(e == hipSuccess) e = dalloc(&h->w_lo, size_t(kWeightRows) * kMaxRays * 2);
	if(e == hipSuccess) e = dalloc(&h->visit, 8 * sizeof(unsigned long long));
	if(e != hipSuccess) {
		h->err = std::string("allocation failed: ") + hipGetErrorString(e);
		return cleanup_fail(e == hipErrorOutOfMemory ? DDGI_ERROR_OUT_OF_MEMORY : DDGI_ERROR_HIP);
	}
	// zeroed atlases (the reference leaves them uninitialised), all probes enabled (state 1)
	e = hipMemset(h->irr[0], 0, P * 64 * sizeof(float4));
	if(e == hipSuccess) e = hipMemset(h->irr[1], 0, P * 64 * sizeof(float4));
	if(e == hipSuccess) e = hipMemset(h->dep[0], 0, P * 256 * sizeof(float2));
	if(e == hipSuccess) e = hipMemset(h->dep[1], 0, P * 256 * sizeof(float2));
	if(e == hipSuccess) e = hipMemset(h->sel, 0, sizeof(SelectState));
	if(e == hipSuccess) e = hipMemset(h->rays, 0, P * R * sizeof(float4));
	if(e == hipSuccess) e = hipMemset(h->visit, 0, 8 * sizeof(unsigned long long));
	if(e == hipSuccess) {
		std::vector<uint32_t> ones(P, 1u);
		e = hipMemcpy(h->states, ones.data(), P * 4, hipMemcpyHostToDevice);
	}
	for(int i = 0; i < 6 && e == hipSuccess; ++i) e = hipEventCreate(&h->ev[i]);
	if(e != hipSuccess) {
		h->err = std::string("initialisation failed: ") + hipGetErrorString(e);
		return cleanup_fail(DDGI_ERROR_HIP);
	}
	h->light.dir[0] = 0.2f / std::sqrt(4.08f);  // LightBuffer default (src/Light.hpp:7-8)
	h->light.dir[1] = 2.0f / std::sqrt(4.08f);
	h->light.dir[2] = 0.2f / std::sqrt(4.08f);
	for(int k = 0; k < 3; ++k) {
		h->light.color[k] = 10.0f;
		h->light.sky[k] = 0.5f;
		h->light.sky_brightness = 10.0f;  // LightBuffer colour (10, 10, 10, 10)
		h->light.sky_model = DDGI_SKY_CONSTANT;
	}
	if((st = ensure_fib(h)) != DDGI_OK) return cleanup_fail(st);
	*out_handle = h;
	return DDGI_OK;
}

ddgi_status ddgi_destroy(void* handle) {
	if(!handle) return fail(nullptr, DDGI_ERROR_INVALID_ARGUMENT, "null handle");
	Handle* h = H(handle);
	if(h->batch_members) return fail(h, DDGI_ERROR_INVALID_STATE, "handle is part of a batch: destroy the batch first");
	set_device(h);
	(void)hipDeviceSynchronize();
	destroy_handle(h);
	return DDGI_OK;
}

ddgi_status ddgi_set_params(void* handle, const ddgi_grid_params* params) {
	if(!handle || !params) return fail(H(handle), DDGI_ERROR_INVALID_ARGUMENT, "null argument");
	Handle* h = H(handle);
	for(int k = 0; k < 3; ++k)
		if(params->resolution[k] != h->params.resolution[k])
			return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "resolution is fixed at creation (re-create the handle to resize)");
	std::string why;
	ddgi_status st = validate_params(params, h->max_rays, why);
	if(st != DDGI_OK) return fail(h, st, why);
	if((st = set_device(h)) != DDGI_OK) return st;
	h->params = *params;
	h->uploaded_hysteresis = params->hysteresis;  // updateUniforms (src/IrradianceProbes.cpp:106-108)
	return ensure_fib(h);
}

ddgi_status ddgi_get_params(void* handle, ddgi_grid_params* out) {
	if(!handle || !out) return fail(H(handle), DDGI_ERROR_INVALID_ARGUMENT, "null argument");
	*out = H(handle)->params;
	return DDGI_OK;
}

ddgi_status ddgi_set_light(void* handle, const ddgi_light* light) {
	if(!handle || !light) return fail(H(handle), DDGI_ERROR_INVALID_ARGUMENT, "null argument");
	Handle* h = H(handle);
	if(!finite_all(light->direction, 3) || !finite_all(light->color, 4))
		return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "non-finite light");
	for(int k = 0; k < 3; ++k) {
		h->light.dir[k] = light->direction[k];
		h->light.color[k] = light->color[k];
	}
	h->light.sky_brightness = light->color[3];
	return DDGI_OK;
}

ddgi_status ddgi_set_sky_model(void* handle, uint32_t model) {
	if(!handle) return fail(nullptr, DDGI_ERROR_INVALID_ARGUMENT, "null handle");
	Handle* h = H(handle);
	if(model != DDGI_SKY_CONSTANT && model != DDGI_SKY_ATMOSPHERE)
		return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "unknown sky model");
	if(model == DDGI_SKY_ATMOSPHERE && !h->miss_list) {
		ddgi_status st = set_device(h);
		if(st != DDGI_OK) return st;
		HIP_TRY(h, dalloc(&h->miss_list, size_t(h->P) * h->max_rays * sizeof(uint32_t)));
	}
	h->light.sky_model = model;
	return DDGI_OK;
}

ddgi_status ddgi_set_sky(void* handle, const float rgb[3]) {
	if(!handle || !rgb) return fail(H(handle), DDGI_ERROR_INVALID_ARGUMENT, "null argument");
	if(!finite_all(rgb, 3)) return fail(H(handle), DDGI_ERROR_INVALID_ARGUMENT, "non-finite sky");
	for(int k = 0; k < 3; ++k) H(handle)->light.sky[k] = rgb[k];
	return DDGI_OK;
}

ddgi_status ddgi_set_scene(void* handle, const ddgi_vertex* vertices, size_t vertex_count, const uint32_t* indices,
						   size_t index_count, const ddgi_offset_entry* offsets, size_t offset_count,
						   const ddgi_instance* instances, size_t instance_count, const ddgi_material* materials,
						   size_t material_count) {
	if(!handle) return fail(nullptr, DDGI_ERROR_INVALID_ARGUMENT, "null handle");
	Handle* h = H(handle);
	if((vertex_count && !vertices) || (index_count && !indices) || (offset_count && !offsets) ||
	   (instance_count && !instances) || (material_count && !materials))
		return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "null array with non-zero count");
	for(size_t k = 0; k < offset_count; ++k)
		if(offsets[k].material_index >= material_count)
			return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "offset entry references a missing material");
	bool metallic = false;
	for(size_t k = 0; k < material_count; ++k) {
		const ddgi_material& m = materials[k];
		if(!std::isfinite(m.metallic) || !std::isfinite(m.roughness) || !finite_all(m.base_color, 3) || !finite_all(m.emissive, 3))
			return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "non-finite material");
		metallic = metallic || m.metallic != 0.0f;
	}
	BvhHost bvh;
	const char* why = nullptr;
	if(!build_bvh(vertices, vertex_count, indices, index_count, offsets, offset_count, instances, instance_count, bvh, &why))
		return fail(h, DDGI_ERROR_INVALID_ARGUMENT, why ? why : "invalid scene");
	if(bvh.tri_count == 0) return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "scene has no triangles");
	if(bvh.wide_depth > kMaxWideDepth)
		return fail(h, DDGI_ERROR_UNSUPPORTED, "BVH deeper than the traversal stack (32 levels of the 4-wide tree)");
	ddgi_status st = set_device(h);
	if(st != DDGI_OK) return st;
	(void)hipDeviceSynchronize();  // the previous scene may still be read by queued work
	free_scene(h);
	SceneDev sd{};
	std::vector<InstanceDev> inst(instance_count);
	for(size_t k = 0; k < instance_count; ++k)
		if(!instance_dev(instances[k], instances[k].transform, bvh.inst_prim_base[k], inst[k]))
			return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "singular instance transform");
	std::vector<float> mats(material_count * 12);
	for(size_t k = 0; k < material_count; ++k) std::memcpy(&mats[k * 12], &materials[k], 48);
	const float4* nodes = nullptr;
	const float4* tris = nullptr;
	const InstanceDev* ip = nullptr;
	const uint32_t* op = nullptr;
	const uint32_t* ixp = nullptr;
	const float4* vp = nullptr;
	const float* mp = nullptr;
	if((st = upload(h, reinterpret_cast<const float4*>(bvh.nodes.data()), bvh.nodes.size() / 4, &nodes)) != DDGI_OK ||
	   (st = upload(h, reinterpret_cast<const float4*>(bvh.tris.data()), bvh.tris.size() / 4, &tris)) != DDGI_OK ||
	   (st = upload(h, inst.data(), inst.size(), &ip)) != DDGI_OK ||
	   (st = upload(h, reinterpret_cast<const uint32_t*>(offsets), offset_count * 3, &op)) != DDGI_OK ||
	   (st = upload(h, indices, index_count, &ixp)) != DDGI_OK ||
	   (st = upload(h, reinterpret_cast<const float4*>(vertices), vertex_count * 4, &vp)) != DDGI_OK ||
	   (st = upload(h, mats.data(), mats.size(), &mp)) != DDGI_OK ||
	   (st = upload(h, bvh.obj_tris.data(), bvh.obj_tris.size(), &h->obj_tris)) != DDGI_OK ||
	   (st = upload(h, bvh.slot_box.data(), bvh.slot_box.size(), &h->slot_box)) != DDGI_OK ||
	   (st = upload(h, identity_xf(instance_count).data(), instance_count * 12, const_cast<const float**>(&h->rel_xf))) != DDGI_OK ||
	   (st = upload(h, bvh.level_nodes.data(), bvh.level_nodes.size(), &h->level_nodes)) != DDGI_OK ||
	   (st = upload<float>(h, nullptr, instance_count * 13, nullptr)) != DDGI_OK) {
		free_scene(h);
		return st;
	}
	h->move_xf = static_cast<float*>(h->scene_bufs.back().p);
	h->move_meta = reinterpret_cast<uint32_t*>(h->move_xf + instance_count * 12);
	if(!h->move_done) HIP_TRY(h, hipEventCreateWithFlags(&h->move_done, hipEventDisableTiming));
	{
		const hipError_t e = hipHostMalloc(&h->move_host, std::max<size_t>(instance_count * (25 * 4 + sizeof(InstanceDev)), 16),
										   hipHostMallocDefault);
		if(e != hipSuccess) {
			h->move_host = nullptr;
			free_scene(h);
			return fail(h, DDGI_ERROR_OUT_OF_MEMORY, std::string("hipHostMalloc: ") + hipGetErrorString(e));
		}
	}
	h->instances.assign(instances, instances + instance_count);
	h->build_xf.resize(instance_count * 12);
	for(size_t k = 0; k < instance_count; ++k) std::memcpy(&h->build_xf[k * 12], instances[k].transform, 48);
	sd.nodes = nodes;
	sd.tris = tris;
	sd.inst = ip;
	sd.offsets = op;
	sd.indices = ixp;
	sd.vertices = vp;
	sd.materials = mp;
	sd.tri_count = bvh.tri_count;
	sd.node_count = bvh.node_count;
	sd.instance_count = uint32_t(instance_count);
	sd.stack_depth = 3 * bvh.wide_depth;  // entries the traversal can hold at once
	h->sd = sd;
	h->bvh_bytes = bvh.nodes.size() * 4 + bvh.tris.size() * 4;
	bvh.nodes.clear();
	bvh.tris.clear();
	bvh.obj_tris.clear();
	bvh.obj_tris.shrink_to_fit();
	bvh.slot_box.clear();
	bvh.slot_box.shrink_to_fit();
	h->bvh_meta = std::move(bvh);
	h->has_scene = true;
	h->has_metallic = metallic;
	h->scene_materials = uint32_t(material_count);
	free_diff(h);  // prepared factors refer to the previous scene
	return DDGI_OK;
}

ddgi_status ddgi_update_instances(void* handle, uint32_t first, uint32_t count, const float* transforms, void* stream) {
	if(!handle) return fail(nullptr, DDGI_ERROR_INVALID_ARGUMENT, "null handle");
	Handle* h = H(handle);
	if(!h->has_scene) return fail(h, DDGI_ERROR_NO_SCENE, "ddgi_set_scene must precede ddgi_update_instances");
	const size_t n = h->instances.size();
	if(size_t(first) > n || size_t(count) > n - first)
		return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "instance range outside the scene");
	if(count == 0) return DDGI_OK;
	if(!transforms) return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "null transforms");
	for(size_t q = 0; q < size_t(count) * 12; ++q)
		if(!std::isfinite(transforms[q])) return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "non-finite instance transform");
	ddgi_status st = set_device(h);
	if(st != DDGI_OK) return st;
	if(h->move_pending) HIP_TRY(h, hipEventSynchronize(h->move_done));  // staging still read by the last call
	h->move_pending = false;
	float* xf = static_cast<float*>(h->move_host);
	uint32_t* meta = reinterpret_cast<uint32_t*>(xf + n * 12);
	InstanceDev* inst = reinterpret_cast<InstanceDev*>(meta + n);
	float* rel = reinterpret_cast<float*>(inst + n);  // n * 12 floats
	for(uint32_t k = 0; k < count; ++k) {
		const ddgi_instance& in = h->instances[first + k];
		const float* M = transforms + size_t(k) * 12;
		if(!instance_dev(in, M, h->bvh_meta.inst_prim_base[first + k], inst[k]))
			return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "singular instance transform");
		const double det = double(M[0]) * (double(M[5]) * M[10] - double(M[6]) * M[9]) -
						   double(M[1]) * (double(M[4]) * M[10] - double(M[6]) * M[8]) +
						   double(M[2]) * (double(M[4]) * M[9] - double(M[5]) * M[8]);
		const bool flip = (det < 0.0) != ((in.flags & DDGI_INSTANCE_FLIP_FACING) != 0);
		meta[k] = (in.mask & 0xFFu) | (flip ? 0x100u : 0u);
		relative_transform(M, &h->build_xf[size_t(first + k) * 12], rel + size_t(k) * 12);
	}
	std::memcpy(xf, transforms, size_t(count) * 48);
	hipStream_t s = static_cast<hipStream_t>(stream);
	HIP_TRY(h, hipMemcpyAsync(h->move_xf, xf, size_t(count) * 48, hipMemcpyHostToDevice, s));
	HIP_TRY(h, hipMemcpyAsync(h->move_meta, meta, size_t(count) * sizeof(uint32_t), hipMemcpyHostToDevice, s));
	HIP_TRY(h, hipMemcpyAsync(const_cast<InstanceDev*>(h->sd.inst) + first, inst, size_t(count) * sizeof(InstanceDev),
							  hipMemcpyHostToDevice, s));
	HIP_TRY(h, hipMemcpyAsync(h->rel_xf + size_t(first) * 12, rel, size_t(count) * 48, hipMemcpyHostToDevice, s));
	float4* tris = const_cast<float4*>(h->sd.tris);
	float4* nodes = const_cast<float4*>(h->sd.nodes);
	HIP_TRY(h, launch_retransform(tris, h->obj_tris, h->sd.tri_count, h->move_xf, h->move_meta, first, count, s));
	const std::vector<uint32_t>& lv = h->bvh_meta.level_offsets;
	for(size_t d = lv.size() - 1; d >= 1; --d)
		HIP_TRY(h, launch_refit_level(nodes, tris, h->slot_box, h->rel_xf, h->level_nodes + lv[d - 1], lv[d] - lv[d - 1], s));
	// the staging area and move_xf / move_meta are reused once this has run
	HIP_TRY(h, hipEventRecord(h->move_done, s));
	h->move_pending = true;
	for(uint32_t k = 0; k < count; ++k) std::memcpy(h->instances[first + k].transform, transforms + size_t(k) * 12, 48);
	free_diff(h);  // prepared factors hold hits of the old geometry
	h->last_stream = s;
	return DDGI_OK;
}

ddgi_status ddgi_get_bvh_info(void* handle, uint64_t* node_count, uint64_t* triangle_count, uint64_t* bytes) {
	if(!handle) return fail(nullptr, DDGI_ERROR_INVALID_ARGUMENT, "null handle");
	Handle* h = H(handle);
	if(!h->has_scene) return fail(h, DDGI_ERROR_NO_SCENE, "no scene");
	if(node_count) *node_count = h->sd.node_count;
	if(triangle_count) *triangle_count = h->sd.tri_count;
	if(bytes) *bytes = h->bvh_bytes;
	return DDGI_OK;
}

ddgi_status ddgi_get_slot_ids(void* handle, uint32_t* instance, uint32_t* primitive, uint64_t count) {
	if(!handle) return fail(nullptr, DDGI_ERROR_INVALID_ARGUMENT, "null handle");
	Handle* h = H(handle);
	if(!h->has_scene) return fail(h, DDGI_ERROR_NO_SCENE, "no scene");
	if(!instance || !primitive || count != h->sd.tri_count)
		return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "null output or count != triangle count");
	ddgi_status st = set_device(h);
	if(st != DDGI_OK) return st;
	// the global id (row 0 .w) and the instance (row 1 .w) of every triangle record, in slot order
	std::vector<float4> tris(size_t(count) * 3);
	HIP_TRY(h, hipMemcpy(tris.data(), h->sd.tris, tris.size() * sizeof(float4), hipMemcpyDeviceToHost));
	for(uint64_t k = 0; k < count; ++k) {
		uint32_t gid, inst;
		std::memcpy(&gid, &tris[k * 3].w, 4);
		std::memcpy(&inst, &tris[k * 3 + 1].w, 4);
		if(inst >= h->bvh_meta.inst_prim_base.size()) return fail(h, DDGI_ERROR_INVALID_STATE, "corrupt triangle record");
		instance[k] = inst;
		primitive[k] = gid - h->bvh_meta.inst_prim_base[inst];
	}
	return DDGI_OK;
}

ddgi_status ddgi_init_probes(void* handle, const float orientation[9], void* stream) {
	if(!handle) return fail(nullptr, DDGI_ERROR_INVALID_ARGUMENT, "null handle");
	Handle* h = H(handle);
	if(!h->has_scene) return fail(h, DDGI_ERROR_NO_SCENE, "ddgi_set_scene must precede ddgi_init_probes");
	Rot R;
	if(make_rotation(orientation, R) != DDGI_OK) return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "invalid orientation");
	ddgi_status st = set_device(h);
	if(st != DDGI_OK || (st = ensure_fib(h)) != DDGI_OK) return st;
	hipStream_t s = static_cast<hipStream_t>(stream);
	// initProbes resets the host-side hysteresis that the ramp moves, without uploading it
	// (src/IrradianceProbes.cpp:358)
	h->params.hysteresis = 0.0f;
	const GridConst g = make_grid_const(*h);
	HIP_TRY(h, launch_frame_dirs(R, h->fib512, h->dirs, 512, s));
	HIP_TRY(h, launch_init_probes(h->sd, g, h->dirs, h->states, nullptr, s));
	h->last_stream = s;
	return DDGI_OK;
}

static ddgi_status run_update(Handle* h, const Rot& R, uint32_t flags, hipStream_t s, float4* factors);

ddgi_status ddgi_update(void* handle, const float orientation[9], uint32_t flags, void* stream) {
	if(!handle) return fail(nullptr, DDGI_ERROR_INVALID_ARGUMENT, "null handle");
	Handle* h = H(handle);
	if(!h->has_scene) return fail(h, DDGI_ERROR_NO_SCENE, "ddgi_set_scene must precede ddgi_update");
	if((flags & DDGI_UPDATE_BLEND_BF16) && (flags & DDGI_UPDATE_BLEND_FP16))
		return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "BF16 and FP16 blends are exclusive");
	Rot R;
	if(make_rotation(orientation, R) != DDGI_OK) return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "invalid orientation");
	ddgi_status st = set_device(h);
	if(st != DDGI_OK || (st = ensure_fib(h)) != DDGI_OK) return st;
	return run_update(h, R, flags, static_cast<hipStream_t>(stream), nullptr);
}

static ddgi_status run_update(Handle* h, const Rot& R, uint32_t flags, hipStream_t s, float4* factors) {
	ddgi_status st = DDGI_OK;
	const bool timing = !(flags & DDGI_UPDATE_NO_TIMING);
	const bool all = (flags & DDGI_UPDATE_ALL_PROBES) != 0;
	if(timing) HIP_TRY(h, hipEventRecord(h->ev[0], s));
	uint32_t max_slots = h->P;
	if(all) {
		if(!h->indices_are_iota) HIP_TRY(h, launch_iota(h->indices, h->P, h->sel, s));
		else HIP_TRY(h, launch_set_count(h->sel, h->P, s));
		h->indices_are_iota = true;
	} else {
		HIP_TRY(h, launch_select(h->states, h->P, h->opts.probes_per_update, h->sel, h->indices, s));
		h->indices_are_iota = false;
		if(h->opts.probes_per_update) max_slots = std::min(max_slots, h->opts.probes_per_update);
	}
	if(timing) HIP_TRY(h, hipEventRecord(h->ev[1], s));
	if(flags & DDGI_UPDATE_HYSTERESIS_RAMP) {
		uint32_t cnt = h->P;
		if(!all) {
			SelectState ss;
			HIP_TRY(h, hipMemcpyAsync(&ss, h->sel, sizeof(ss), hipMemcpyDeviceToHost, s));
			HIP_TRY(h, hipStreamSynchronize(s));
			cnt = ss.count;
		}
		advance_hysteresis(h, cnt);
	}
	const GridConst g = make_grid_const(*h);
	uint32_t key_shift = 0, key_bits = 0;  // shade-list ordering keys
	shade_key_layout(g, &key_shift, &key_bits);
	HIP_TRY(h, launch_frame_dirs_order(R, h->fib, h->dirs, g.n, h->perm, s, h->shade_count));
	TraceFrameArgs a{};
	a.scene = h->sd;
	a.grid = g;
	a.light = h->light;
	a.indices = h->indices;
	a.count = &h->sel->count;
	a.dirs = h->dirs;
	a.perm = h->perm;
	a.states = h->states;
	a.irr = h->irr[h->cur];
	a.dep = h->dep[h->cur];
	a.hits = h->hits;
	a.rays = h->rays;
	a.shade_list = h->shade_list;
	a.shade_count = h->shade_count;
	a.miss_list = h->miss_list;
	a.max_slots = max_slots;
	a.visit_counters = nullptr;
	a.factors = factors;
	if(flags & DDGI_UPDATE_COUNT_VISITS) {
		HIP_TRY(h, hipMemsetAsync(h->visit, 0, 8 * sizeof(unsigned long long), s));
		a.visit_counters = h->visit;
	}
	a.shade_keys = h->shade_keys;
	a.key_shift = key_shift;
	{  // segment s of the shade list takes the front-face hits of workgroups b with b % 8 == s
		const uint32_t blocks = uint32_t((uint64_t(max_slots) * g.n + kPrimaryBlock - 1) / kPrimaryBlock);
		a.seg_cap = (blocks + kShadeSegments - 1) / kShadeSegments * kPrimaryBlock;
	}
	HIP_TRY(h, launch_trace_primary(a, s));
	if(timing) HIP_TRY(h, hipEventRecord(h->ev[2], s));
	// shade the front-face hits ordered by the grid cell they lie in
	HIP_TRY(h, launch_key_sort(SortJob{h->shade_keys, h->shade_list, h->shade_count, h->sort_counts, h->sort_keys,
									   h->sort_list, h->shade_count + kShadeTotalSlot, 0u, kShadeSegments, a.seg_cap, 0u},
							   uint64_t(max_slots) * g.n, 3 * key_bits, s));
	HIP_TRY(h, launch_shade(a, s));
	if(h->light.sky_model == DDGI_SKY_ATMOSPHERE) HIP_TRY(h, launch_sky(a, s));
	if(timing) HIP_TRY(h, hipEventRecord(h->ev[3], s));
	if((st = enqueue_blend_and_publish(h, nullptr, h->rays, h->perm, max_slots, all, flags, s)) != DDGI_OK) return st;
	if(timing) {
		HIP_TRY(h, hipEventRecord(h->ev[4], s));
		HIP_TRY(h, hipEventRecord(h->ev[5], s));
		h->timing_pending = true;
	}
	h->last_count = all ? h->P : 0xFFFFFFFFu;
	h->last_stream = s;
	return DDGI_OK;
}

ddgi_status ddgi_blend_rays(void* handle, const void* rays, const void* probe_indices, uint32_t count,
							const float orientation[9], uint32_t flags, void* stream) {
	if(!handle) return fail(nullptr, DDGI_ERROR_INVALID_ARGUMENT, "null handle");
	Handle* h = H(handle);
	if(!rays || !probe_indices || count == 0 || count > h->P)
		return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "rays/probe_indices null or count outside [1, P]");
	if((flags & DDGI_UPDATE_BLEND_BF16) && (flags & DDGI_UPDATE_BLEND_FP16))
		return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "BF16 and FP16 blends are exclusive");
	Rot R;
	if(make_rotation(orientation, R) != DDGI_OK) return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "invalid orientation");
	ddgi_status st = set_device(h);
	if(st != DDGI_OK || (st = ensure_fib(h)) != DDGI_OK) return st;
	hipStream_t s = static_cast<hipStream_t>(stream);
	const bool timing = !(flags & DDGI_UPDATE_NO_TIMING);
	if(timing)
		for(int k = 0; k < 4; ++k) HIP_TRY(h, hipEventRecord(h->ev[k], s));
	HIP_TRY(h, hipMemcpyAsync(h->indices, probe_indices, size_t(count) * 4, hipMemcpyDeviceToDevice, s));
	HIP_TRY(h, launch_set_count(h->sel, count, s));
	h->indices_are_iota = false;
	// the listed probes may not cover the grid: the next atlas starts as a copy of the exposed one,
	// unless the caller states that they are every probe
	const bool all = (flags & DDGI_UPDATE_ALL_PROBES) && count == h->P;
	if((st = enqueue_blend_and_publish(h, &R, static_cast<const float4*>(rays), nullptr, count, all, flags, s)) != DDGI_OK)
		return st;
	if(timing) {
		HIP_TRY(h, hipEventRecord(h->ev[4], s));
		HIP_TRY(h, hipEventRecord(h->ev[5], s));
		h->timing_pending = true;
	}
	h->last_count = count;
	h->last_stream = s;
	return DDGI_OK;
}

extern "C++" {
template <class T>
static hipError_t diff_alloc(Handle* h, T** p, size_t bytes) {
	DeviceBuffer b;
	b.bytes = std::max<size_t>(bytes, 16);
	hipError_t e = hipMalloc(&b.p, b.bytes);
	if(e == hipSuccess) {
		h->diff.bufs.push_back(b);
		*p = static_cast<T*>(b.p);
	}
	return e;
}
}

ddgi_status ddgi_diff_prepare(void* handle, const float orientation[9], const float* gbuf_position,
							  const float* gbuf_normal, const float* gbuf_to_camera, const uint32_t* gbuf_material,
							  uint64_t texel_count, uint32_t material_count, void* stream) {
	if(!handle) return fail(nullptr, DDGI_ERROR_INVALID_ARGUMENT, "null handle");
	Handle* h = H(handle);
	if(!h->has_scene) return fail(h, DDGI_ERROR_NO_SCENE, "ddgi_set_scene must precede ddgi_diff_prepare");
	if(h->has_metallic)
		return fail(h, DDGI_ERROR_UNSUPPORTED, "differentiation needs non-metallic materials (radiance affine in albedo)");
	if(material_count == 0 || material_count > 1024 || material_count < h->scene_materials)
		return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "material_count must cover the scene materials and be <= 1024");
	if(texel_count == 0 || texel_count > (uint64_t(1) << 31) || !gbuf_position || !gbuf_normal || !gbuf_to_camera ||
	   !gbuf_material)
		return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "G-buffer arrays null or texel_count outside [1, 2^31]");
	Rot R;
	if(make_rotation(orientation, R) != DDGI_OK) return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "invalid orientation");
	ddgi_status st = set_device(h);
	if(st != DDGI_OK || (st = ensure_fib(h)) != DDGI_OK) return st;
	hipStream_t s = static_cast<hipStream_t>(stream);
	{
		std::vector<uint32_t> mats(texel_count);
		HIP_TRY(h, hipStreamSynchronize(s));
		HIP_TRY(h, hipMemcpy(mats.data(), gbuf_material, texel_count * 4, hipMemcpyDeviceToHost));
		for(uint32_t m : mats)
			if(m >= material_count) return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "G-buffer material id >= material_count");
	}
	free_diff(h);
	Handle::Diff& D = h->diff;
	const uint64_t P = h->P, n = h->params.rays_per_probe, rays = P * n, N = texel_count, M = material_count;
	const uint32_t max_items = uint32_t(P + (N + kDiffChunk - 1) / kDiffChunk);
	DiffArgs d{};
	float4* factors = nullptr;
	float *pos, *nrm, *toc, *out, *loss_part, *g_T;
	uint32_t *gmat, *states, *indices, *cell_id, *cell_count, *cell_start, *cursor, *cell_item_start, *item_cell,
		*item_begin, *item_count, *n_items, *sorted;
	float2* dep;
	long long *part_tile, *part_mat_pix, *part_mat_ray;
	hipError_t e = diff_alloc(h, &factors, rays * 4 * sizeof(float4));
	if(e == hipSuccess) e = diff_alloc(h, &D.rad, rays * sizeof(float4));
	if(e == hipSuccess) e = diff_alloc(h, &D.prev, P * 64 * sizeof(float4));
	if(e == hipSuccess) e = diff_alloc(h, &D.irr, P * 64 * sizeof(float4));
	if(e == hipSuccess) e = diff_alloc(h, &dep, P * 256 * sizeof(float2));
	if(e == hipSuccess) e = diff_alloc(h, &states, P * 4);
	if(e == hipSuccess) e = diff_alloc(h, &indices, P * 4);
	if(e == hipSuccess) e = diff_alloc(h, &D.wt, size_t(kWeightRows) * kMaxRays * 4);
	if(e == hipSuccess) e = diff_alloc(h, &D.inv_sum, kBlendTexels * 4);
	if(e == hipSuccess) e = diff_alloc(h, &D.dirs, kMaxRays * sizeof(float4));
	if(e == hipSuccess) e = diff_alloc(h, &D.count, 4);
	if(e == hipSuccess) e = diff_alloc(h, &pos, N * 12);
	if(e == hipSuccess) e = diff_alloc(h, &nrm, N * 12);
	if(e == hipSuccess) e = diff_alloc(h, &toc, N * 12);
	if(e == hipSuccess) e = diff_alloc(h, &gmat, N * 4);
	if(e == hipSuccess) e = diff_alloc(h, &out, N * 12);
	D.loss_parts = uint32_t((N + 255) / 256);
	if(e == hipSuccess) e = diff_alloc(h, &loss_part, size_t(D.loss_parts) * 4);
	if(e == hipSuccess) e = diff_alloc(h, &cell_id, N * 4);
	if(e == hipSuccess) e = diff_alloc(h, &cell_count, P * 4);
	if(e == hipSuccess) e = diff_alloc(h, &cell_start, (P + 1) * 4);
	if(e == hipSuccess) e = diff_alloc(h, &cursor, P * 4);
	if(e == hipSuccess) e = diff_alloc(h, &cell_item_start, (P + 1) * 4);
	if(e == hipSuccess) e = diff_alloc(h, &item_cell, size_t(max_items) * 4);
	if(e == hipSuccess) e = diff_alloc(h, &item_begin, size_t(max_items) * 4);
	if(e == hipSuccess) e = diff_alloc(h, &item_count, size_t(max_items) * 4);
	if(e == hipSuccess) e = diff_alloc(h, &n_items, 4);
	if(e == hipSuccess) e = diff_alloc(h, &sorted, N * 4);
	if(e == hipSuccess) e = diff_alloc(h, &part_tile, size_t(max_items) * 8 * 192 * 8);
	if(e == hipSuccess) e = diff_alloc(h, &part_mat_pix, size_t(max_items) * M * 3 * 8);
	if(e == hipSuccess) e = diff_alloc(h, &g_T, P * 108 * 4);
	if(e == hipSuccess) e = diff_alloc(h, &part_mat_ray, P * (M * 3 + 3) * 8);
	if(e != hipSuccess) {
		free_diff(h);
		return fail(h, e == hipErrorOutOfMemory ? DDGI_ERROR_OUT_OF_MEMORY : DDGI_ERROR_HIP,
					std::string("diff allocation: ") + hipGetErrorString(e));
	}
	HIP_TRY(h, hipMemcpyAsync(pos, gbuf_position, N * 12, hipMemcpyDeviceToDevice, s));
	HIP_TRY(h, hipMemcpyAsync(nrm, gbuf_normal, N * 12, hipMemcpyDeviceToDevice, s));
	HIP_TRY(h, hipMemcpyAsync(toc, gbuf_to_camera, N * 12, hipMemcpyDeviceToDevice, s));
	HIP_TRY(h, hipMemcpyAsync(gmat, gbuf_material, N * 4, hipMemcpyDeviceToDevice, s));
	HIP_TRY(h, hipMemcpyAsync(D.prev, h->irr[h->cur], P * 64 * sizeof(float4), hipMemcpyDeviceToDevice, s));
	// the prepared update: all probes, FP32 blend, shading factors recorded
	if((st = run_update(h, R, DDGI_UPDATE_ALL_PROBES | DDGI_UPDATE_NO_TIMING, s, factors)) != DDGI_OK) {
		free_diff(h);
		return st;
	}
	HIP_TRY(h, hipMemcpyAsync(dep, h->dep[h->cur], P * 256 * sizeof(float2), hipMemcpyDeviceToDevice, s));
	HIP_TRY(h, hipMemcpyAsync(states, h->states, P * 4, hipMemcpyDeviceToDevice, s));
	const GridConst g = make_grid_const(*h);
	HIP_TRY(h, launch_frame_dirs(R, h->fib, D.dirs, g.n, s));
	HIP_TRY(h, launch_blend_weights(g, D.dirs, h->perm, D.wt, D.inv_sum, h->w_hi, h->w_lo, kBlendFP32, s));
	{
		std::vector<uint32_t> iota(P);
		std::iota(iota.begin(), iota.end(), 0u);
		const uint32_t cnt = uint32_t(P);
		HIP_TRY(h, hipStreamSynchronize(s));
		HIP_TRY(h, hipMemcpy(indices, iota.data(), P * 4, hipMemcpyHostToDevice));
		HIP_TRY(h, hipMemcpy(D.count, &cnt, 4, hipMemcpyHostToDevice));
	}
	d.grid = g;
	d.factors = factors;
	d.perm = h->perm;
	d.dirs = D.dirs;
	d.inv_sum = D.inv_sum;
	d.states = states;
	d.dep = dep;
	d.indices = indices;
	d.slots = uint32_t(P);
	d.rays = rays;
	d.pos = pos;
	d.nrm = nrm;
	d.tocam = toc;
	d.gmat = gmat;
	d.N = N;
	d.M = uint32_t(M);
	d.sorted = sorted;
	d.cell_item_start = cell_item_start;
	d.item_begin = item_begin;
	d.item_count = item_count;
	d.n_items = n_items;
	d.max_items = max_items;
	d.rad = D.rad;
	d.irr = D.irr;
	d.out = out;
	d.loss_part = loss_part;
	d.part_tile = part_tile;
	d.part_mat_pix = part_mat_pix;
	d.g_T = g_T;
	d.part_mat_ray = part_mat_ray;
	HIP_TRY(h, diff_bin(d, cell_id, cell_count, cell_start, cursor, cell_item_start, item_cell, item_begin, item_count,
						n_items, sorted, s));
	D.d = d;
	D.ready = true;
	h->last_stream = s;
	return DDGI_OK;
}

ddgi_status ddgi_diff_forward(void* handle, const float* albedo, const float* sun, const float* target, float* shade_out,
							  float* loss_out, void* stream) {
	if(!handle) return fail(nullptr, DDGI_ERROR_INVALID_ARGUMENT, "null handle");
	Handle* h = H(handle);
	if(!h->diff.ready) return fail(h, DDGI_ERROR_INVALID_STATE, "ddgi_diff_prepare must precede ddgi_diff_forward");
	if(!albedo || !sun || !target || !shade_out || !loss_out) return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "null array");
	ddgi_status st = set_device(h);
	if(st != DDGI_OK) return st;
	hipStream_t s = static_cast<hipStream_t>(stream);
	Handle::Diff& D = h->diff;
	DiffArgs d = D.d;
	d.albedo = albedo;
	d.sun = sun;
	d.target = target;
	d.shade = shade_out;
	d.loss = loss_out;
	HIP_TRY(h, diff_radiance(d, s));
	BlendArgs b{};
	b.grid = d.grid;
	b.indices = d.indices;
	b.count = D.count;
	b.rays = D.rad;
	b.wt = D.wt;
	b.inv_sum = D.inv_sum;
	b.states = nullptr;
	b.irr_prev = D.prev;
	b.dep_prev = d.dep;
	b.irr_out = D.irr;
	b.dep_out = nullptr;
	b.max_slots = d.slots;
	b.flags = kBlendIrradianceOnly;
	HIP_TRY(h, launch_blend(b, kBlendFP32, s));
	HIP_TRY(h, diff_gather_forward(d, s));
	h->last_stream = s;
	return DDGI_OK;
}

ddgi_status ddgi_diff_backward(void* handle, const float* albedo, const float* sun, const float* target,
							   const float* grad_loss, float* grad_albedo, float* grad_sun, void* stream) {
	if(!handle) return fail(nullptr, DDGI_ERROR_INVALID_ARGUMENT, "null handle");
	Handle* h = H(handle);
	if(!h->diff.ready) return fail(h, DDGI_ERROR_INVALID_STATE, "ddgi_diff_prepare must precede ddgi_diff_backward");
	if(!albedo || !sun || !target || !grad_loss || !grad_albedo || !grad_sun)
		return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "null array");
	ddgi_status st = set_device(h);
	if(st != DDGI_OK) return st;
	hipStream_t s = static_cast<hipStream_t>(stream);
	DiffArgs d = h->diff.d;
	d.albedo = albedo;
	d.sun = sun;
	d.target = target;
	d.grad_loss = grad_loss;
	d.g_albedo = grad_albedo;
	d.g_sun = grad_sun;
	HIP_TRY(h, diff_backward(d, s));
	h->last_stream = s;
	return DDGI_OK;
}

ddgi_status ddgi_trace_rays(void* handle, const ddgi_ray* rays, ddgi_hit* hits, uint64_t count, uint32_t mask,
							uint32_t flags, void* stream) {
	if(!handle) return fail(nullptr, DDGI_ERROR_INVALID_ARGUMENT, "null handle");
	Handle* h = H(handle);
	if(!h->has_scene) return fail(h, DDGI_ERROR_NO_SCENE, "no scene");
	if(count && (!rays || !hits)) return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "null rays/hits");
	ddgi_status st = set_device(h);
	if(st != DDGI_OK) return st;
	hipStream_t s = static_cast<hipStream_t>(stream);
	HIP_TRY(h, launch_trace_rays(h->sd, rays, hits, count, mask, (flags & DDGI_TRACE_ANY_HIT) != 0,
								 (flags & DDGI_TRACE_CULL_BACK_FACES) != 0, s));
	h->last_stream = s;
	return DDGI_OK;
}

ddgi_status ddgi_gbuffer(void* handle, const ddgi_ray* rays, uint64_t count, ddgi_gbuffer_texel* out, void* stream) {
	if(!handle) return fail(nullptr, DDGI_ERROR_INVALID_ARGUMENT, "null handle");
	Handle* h = H(handle);
	if(!h->has_scene) return fail(h, DDGI_ERROR_NO_SCENE, "ddgi_set_scene must precede ddgi_gbuffer");
	if(count && (!rays || !out)) return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "null rays or output");
	ddgi_status st = set_device(h);
	if(st != DDGI_OK) return st;
	hipStream_t s = static_cast<hipStream_t>(stream);
	HIP_TRY(h, launch_gbuffer(h->sd, rays, count, out, s));
	h->last_stream = s;
	return DDGI_OK;
}

ddgi_status ddgi_gather(void* handle, const ddgi_gbuffer_texel* gbuffer, uint64_t count, const float view_origin[3],
						const float* direct, const float* reflection, float* out_rgb, void* stream) {
	if(!handle) return fail(nullptr, DDGI_ERROR_INVALID_ARGUMENT, "null handle");
	Handle* h = H(handle);
	if(!h->has_scene) return fail(h, DDGI_ERROR_NO_SCENE, "ddgi_set_scene must precede ddgi_gather");
	if(count && (!gbuffer || !out_rgb)) return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "null G-buffer or output");
	if(!view_origin || !finite_all(view_origin, 3)) return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "invalid view origin");
	ddgi_status st = set_device(h);
	if(st != DDGI_OK) return st;
	GatherArgs a{};
	a.scene = h->sd;
	a.grid = make_grid_const(*h);
	a.light = h->light;
	a.states = h->states;
	a.irr = h->irr[h->cur];
	a.dep = h->dep[h->cur];
	a.gbuffer = gbuffer;
	a.direct = direct;
	a.reflection = reflection;
	a.out = out_rgb;
	a.count = count;
	for(int k = 0; k < 3; ++k) a.origin[k] = view_origin[k];
	// the shadow ray direction: normalize(light direction) in FP32 (directLight.rgen:72)
	const float* L = h->light.dir;
	const float len = std::sqrt(L[0] * L[0] + L[1] * L[1] + L[2] * L[2]);
	for(int k = 0; k < 3; ++k) a.sdir[k] = L[k] / len;
	hipStream_t s = static_cast<hipStream_t>(stream);
	HIP_TRY(h, launch_gather(a, s));
	h->last_stream = s;
	return DDGI_OK;
}

ddgi_status ddgi_direct_light(void* handle, const ddgi_gbuffer_texel* gbuffer, uint64_t count, const float* noise,
							  const float* previous, float* out_rgba, void* stream) {
	if(!handle) return fail(nullptr, DDGI_ERROR_INVALID_ARGUMENT, "null handle");
	Handle* h = H(handle);
	if(!h->has_scene) return fail(h, DDGI_ERROR_NO_SCENE, "ddgi_set_scene must precede ddgi_direct_light");
	if(count && (!gbuffer || !out_rgba)) return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "null G-buffer or output");
	ddgi_status st = set_device(h);
	if(st != DDGI_OK) return st;
	DirectLightArgs a{};
	a.scene = h->sd;
	a.gbuffer = gbuffer;
	a.noise = reinterpret_cast<const float2*>(noise);
	a.previous = reinterpret_cast<const float4*>(previous);
	a.out = reinterpret_cast<float4*>(out_rgba);
	a.count = count;
	const float* L = h->light.dir;  // normalize(light direction) in FP32 (directLight.rgen:71)
	const float len = std::sqrt(L[0] * L[0] + L[1] * L[1] + L[2] * L[2]);
	for(int k = 0; k < 3; ++k) a.sdir[k] = L[k] / len;
	hipStream_t s = static_cast<hipStream_t>(stream);
	HIP_TRY(h, launch_direct_light(a, s));
	h->last_stream = s;
	return DDGI_OK;
}

ddgi_status ddgi_reflection(void* handle, const ddgi_gbuffer_texel* gbuffer, uint64_t count, const float view_origin[3],
							const float* noise, float* out_rgba, void* stream) {
	if(!handle) return fail(nullptr, DDGI_ERROR_INVALID_ARGUMENT, "null handle");
	Handle* h = H(handle);
	if(!h->has_scene) return fail(h, DDGI_ERROR_NO_SCENE, "ddgi_set_scene must precede ddgi_reflection");
	if(count && (!gbuffer || !out_rgba)) return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "null G-buffer or output");
	if(!view_origin || !finite_all(view_origin, 3)) return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "invalid view origin");
	ddgi_status st = set_device(h);
	if(st != DDGI_OK) return st;
	hipStream_t s = static_cast<hipStream_t>(stream);
	// texels go in chunks of at most 4 Mi; the scratch (32 B per texel of a chunk) grows on demand
	constexpr uint64_t kChunk = uint64_t(1) << 22;
	const uint32_t cap = uint32_t(std::min<uint64_t>(count, kChunk));
	if(cap > h->refl_cap) {
		free_reflection_scratch(h);
		hipError_t e = dalloc(&h->refl_hits, size_t(cap) * sizeof(float4));
		if(e == hipSuccess) e = dalloc(&h->refl_list, size_t(cap) * 4);
		if(e == hipSuccess) e = dalloc(&h->refl_keys, size_t(cap) * 4);
		if(e == hipSuccess) e = dalloc(&h->refl_bkeys, size_t(cap) * 4);
		if(e == hipSuccess) e = dalloc(&h->refl_blist, size_t(cap) * 4);
		if(e == hipSuccess) e = dalloc(&h->refl_count, 16);
		if(e == hipSuccess) e = dalloc(&h->refl_counts, size_t(kShadeSortCounts) * 4);
		if(e != hipSuccess) {
			free_reflection_scratch(h);
			(void)hipGetLastError();
			return fail(h, DDGI_ERROR_OUT_OF_MEMORY, "reflection scratch allocation failed");
		}
		h->refl_cap = cap;
	}
	const ReflScratch sc{h->refl_hits, h->refl_list, h->refl_keys, h->refl_bkeys, h->refl_blist, h->refl_count, h->refl_counts};
	if((st = enqueue_reflection(h, gbuffer, count, view_origin, noise, out_rgba, sc, kChunk, s)) != DDGI_OK) return st;
	h->last_stream = s;
	return DDGI_OK;
}

ddgi_status ddgi_reflection_noalloc(void* handle, const ddgi_gbuffer_texel* gbuffer, uint64_t count,
									const float view_origin[3], const float* noise, float* out_rgba, void* stream) {
	if(!handle) return fail(nullptr, DDGI_ERROR_INVALID_ARGUMENT, "null handle");
	Handle* h = H(handle);
	if(!h->has_scene) return fail(h, DDGI_ERROR_NO_SCENE, "ddgi_set_scene must precede ddgi_reflection_noalloc");
	if(count && (!gbuffer || !out_rgba)) return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "null G-buffer or output");
	if(!view_origin || !finite_all(view_origin, 3)) return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "invalid view origin");
	if(count > kNoAllocTexels)
		return fail(h, DDGI_ERROR_INVALID_ARGUMENT,
					"count exceeds the 2,097,152 texels of the allocation-free scratch (split the G-buffer into calls)");
	ddgi_status st = set_device(h);
	if(st != DDGI_OK) return st;
	unsigned char* ptr = nullptr;
	if((st = noalloc_scratch(h, &ptr)) != DDGI_OK) return st;
	const size_t N = size_t(kNoAllocTexels);
	const ReflScratch sc{reinterpret_cast<float4*>(ptr),
						 reinterpret_cast<uint32_t*>(ptr + N * 16),
						 reinterpret_cast<uint32_t*>(ptr + N * 20),
						 reinterpret_cast<uint32_t*>(ptr + N * 24),
						 reinterpret_cast<uint32_t*>(ptr + N * 28),
						 reinterpret_cast<uint32_t*>(ptr + kNoAllocCountOffset),
						 reinterpret_cast<uint32_t*>(ptr + kNoAllocCountersOffset)};
	hipStream_t s = static_cast<hipStream_t>(stream);
	if((st = enqueue_reflection(h, gbuffer, count, view_origin, noise, out_rgba, sc, kNoAllocTexels, s)) != DDGI_OK) return st;
	h->last_stream = s;
	return DDGI_OK;
}

ddgi_status ddgi_sample(void* handle, const float* position, const float* normal, const float* to_camera, uint64_t count,
						float* rgb_out, void* stream) {
	if(!handle) return fail(nullptr, DDGI_ERROR_INVALID_ARGUMENT, "null handle");
	Handle* h = H(handle);
	if(count && (!position || !normal || !to_camera || !rgb_out)) return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "null array");
	ddgi_status st = set_device(h);
	if(st != DDGI_OK) return st;
	hipStream_t s = static_cast<hipStream_t>(stream);
	HIP_TRY(h, launch_sample(make_grid_const(*h), h->states, h->irr[h->cur], h->dep[h->cur], position, normal, to_camera,
							 count, rgb_out, s));
	h->last_stream = s;
	return DDGI_OK;
}

ddgi_status ddgi_get_atlas(void* handle, uint32_t kind, uint32_t layout, void* host_dst, size_t bytes) {
	if(!handle || !host_dst) return fail(H(handle), DDGI_ERROR_INVALID_ARGUMENT, "null argument");
	Handle* h = H(handle);
	if(kind > 1 || layout > 1) return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "unknown atlas kind/layout");
	const uint32_t res = kind == 0 ? 8 : 16, ch = kind == 0 ? 4 : 2;
	const size_t need = size_t(h->P) * res * res * ch * sizeof(float);
	if(bytes != need) return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "destination size mismatch");
	ddgi_status st = set_device(h);
	if(st != DDGI_OK) return st;
	HIP_TRY(h, hipStreamSynchronize(h->last_stream));
	const void* src = kind == 0 ? static_cast<const void*>(h->irr[h->cur]) : static_cast<const void*>(h->dep[h->cur]);
	if(layout == DDGI_LAYOUT_TILED) {
		HIP_TRY(h, hipMemcpy(host_dst, src, need, hipMemcpyDeviceToHost));
	} else {
		std::vector<float> tmp(need / sizeof(float));
		HIP_TRY(h, hipMemcpy(tmp.data(), src, need, hipMemcpyDeviceToHost));
		tiled_to_image(tmp.data(), static_cast<float*>(host_dst), *h, res, ch);
	}
	return DDGI_OK;
}

ddgi_status ddgi_set_atlas(void* handle, uint32_t kind, uint32_t layout, const void* host_src, size_t bytes) {
	if(!handle || !host_src) return fail(H(handle), DDGI_ERROR_INVALID_ARGUMENT, "null argument");
	Handle* h = H(handle);
	if(kind > 1 || layout > 1) return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "unknown atlas kind/layout");
	const uint32_t res = kind == 0 ? 8 : 16, ch = kind == 0 ? 4 : 2;
	const size_t need = size_t(h->P) * res * res * ch * sizeof(float);
	if(bytes != need) return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "source size mismatch");
	ddgi_status st = set_device(h);
	if(st != DDGI_OK) return st;
	HIP_TRY(h, hipStreamSynchronize(h->last_stream));
	void* dst = kind == 0 ? static_cast<void*>(h->irr[h->cur]) : static_cast<void*>(h->dep[h->cur]);
	if(layout == DDGI_LAYOUT_TILED) {
		HIP_TRY(h, hipMemcpy(dst, host_src, need, hipMemcpyHostToDevice));
	} else {
		std::vector<float> tmp(need / sizeof(float));
		image_to_tiled(static_cast<const float*>(host_src), tmp.data(), *h, res, ch);
		HIP_TRY(h, hipMemcpy(dst, tmp.data(), need, hipMemcpyHostToDevice));
	}
	return DDGI_OK;
}

ddgi_status ddgi_get_states(void* handle, uint32_t* host_dst, size_t count) {
	if(!handle || !host_dst) return fail(H(handle), DDGI_ERROR_INVALID_ARGUMENT, "null argument");
	Handle* h = H(handle);
	if(count != h->P) return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "count must equal the probe count");
	ddgi_status st = set_device(h);
	if(st != DDGI_OK) return st;
	HIP_TRY(h, hipStreamSynchronize(h->last_stream));
	HIP_TRY(h, hipMemcpy(host_dst, h->states, count * 4, hipMemcpyDeviceToHost));
	return DDGI_OK;
}

ddgi_status ddgi_set_states(void* handle, const uint32_t* host_src, size_t count) {
	if(!handle || !host_src) return fail(H(handle), DDGI_ERROR_INVALID_ARGUMENT, "null argument");
	Handle* h = H(handle);
	if(count != h->P) return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "count must equal the probe count");
	for(size_t i = 0; i < count; ++i)
		if(host_src[i] > 8) return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "probe states are in [0, 8]");
	ddgi_status st = set_device(h);
	if(st != DDGI_OK) return st;
	HIP_TRY(h, hipStreamSynchronize(h->last_stream));
	HIP_TRY(h, hipMemcpy(h->states, host_src, count * 4, hipMemcpyHostToDevice));
	return DDGI_OK;
}

ddgi_status ddgi_get_timings(void* handle, ddgi_timings* out) {
	if(!handle || !out) return fail(H(handle), DDGI_ERROR_INVALID_ARGUMENT, "null argument");
	Handle* h = H(handle);
	ddgi_status st = set_device(h);
	if(st != DDGI_OK) return st;
	if(h->timing_pending) {
		HIP_TRY(h, hipEventSynchronize(h->ev[5]));
		float ms[5] = {};
		for(int i = 0; i < 5; ++i) HIP_TRY(h, hipEventElapsedTime(&ms[i], h->ev[i], h->ev[i + 1]));
		h->last.select_ms = ms[0];
		h->last.primary_ms = ms[1];
		h->last.shade_ms = ms[2];
		h->last.trace_ms = ms[1] + ms[2];
		h->last.blend_ms = ms[3];
		h->last.border_ms = 0.0f;
		h->last.copy_ms = 0.0f;
		HIP_TRY(h, hipEventElapsedTime(&h->last.total_ms, h->ev[0], h->ev[5]));
		uint32_t cnt = h->last_count;
		if(cnt == 0xFFFFFFFFu) {
			SelectState ss;
			HIP_TRY(h, hipMemcpy(&ss, h->sel, sizeof(ss), hipMemcpyDeviceToHost));
			cnt = ss.count;
		}
		h->last.probes_updated = cnt;
		h->last.rays_traced = cnt * h->params.rays_per_probe;
		h->timing_pending = false;
	}
	*out = h->last;
	return DDGI_OK;
}

ddgi_status ddgi_get_trace_stats(void* handle, uint64_t out[DDGI_TRACE_STATS]) {
	if(!handle || !out) return fail(H(handle), DDGI_ERROR_INVALID_ARGUMENT, "null argument");
	Handle* h = H(handle);
	ddgi_status st = set_device(h);
	if(st != DDGI_OK) return st;
	HIP_TRY(h, hipStreamSynchronize(h->last_stream));
	unsigned long long v[8];
	HIP_TRY(h, hipMemcpy(v, h->visit, sizeof(v), hipMemcpyDeviceToHost));
	for(int i = 0; i < DDGI_TRACE_STATS; ++i) out[i] = v[i];
	return DDGI_OK;
}

ddgi_status ddgi_get_device_view(void* handle, ddgi_device_view* out) {
	if(!handle || !out) return fail(H(handle), DDGI_ERROR_INVALID_ARGUMENT, "null argument");
	Handle* h = H(handle);
	out->irradiance = h->irr[h->cur];
	out->depth = h->dep[h->cur];
	out->states = h->states;
	out->rays = h->rays;
	out->hits = h->hits;
	out->ray_order = h->perm;
	out->probe_indices = h->indices;
	out->ray_directions = h->dirs;
	out->probe_count = h->P;
	out->max_rays_per_probe = h->max_rays;
	out->last_update_count = h->last_count;
	return DDGI_OK;
}

ddgi_status ddgi_synchronize(void* handle) {
	if(!handle) return fail(nullptr, DDGI_ERROR_INVALID_ARGUMENT, "null handle");
	Handle* h = H(handle);
	ddgi_status st = set_device(h);
	if(st != DDGI_OK) return st;
	HIP_TRY(h, hipStreamSynchronize(h->last_stream));
	return DDGI_OK;
}

// Saved state: header (magic, version, P, rays/probe, hysteresis bits, updated_probes, select
// state) + irradiance tiles + depth tiles + states.
ddgi_status ddgi_save_state(void* handle, void* host_dst, size_t* bytes) {
	if(!handle || !bytes) return fail(H(handle), DDGI_ERROR_INVALID_ARGUMENT, "null argument");
	Handle* h = H(handle);
	const size_t irr_b = size_t(h->P) * 64 * 16, dep_b = size_t(h->P) * 256 * 8, st_b = size_t(h->P) * 4;
	const size_t hdr = 16 * 4;
	const size_t need = hdr + irr_b + dep_b + st_b;
	if(!host_dst) {
		*bytes = need;
		return DDGI_OK;
	}
	if(*bytes < need) return fail(h, DDGI_ERROR_INVALID_ARGUMENT, "buffer too small");
	ddgi_status st = set_device(h);
	if(st != DDGI_OK) return st;
	HIP_TRY(h, hipStreamSynchronize(h->last_stream));
	uint32_t head[16] = {};
	SelectState ss;
	HIP_TRY(h, hipMemcpy(&ss, h->sel, sizeof(ss), hipMemcpyDeviceToHost));
	head[0] = kStateMagic;
	head[1] = kStateVersion;
	head[2] = h->P;
	head[3] = h->params.rays_per_probe;
	std::memcpy(&head[4], &h->params.hysteresis, 4);
	head[5] = h->updated_probes;
	head[6] = ss.last_offset;
	head[7] = ss.loop_index;
	std::memcpy(&head[8], &h->uploaded_hysteresis, 4);
	char* d = static_cast<char*>(host_dst);
	std::memcpy(d, head, hdr);
	HIP_TRY(h, hipMemcpy(d + hdr, h->irr[h->cur], irr_b, hipMemcpyDeviceToHost));
	HIP_TRY(h, hipMemcpy(d + hdr + irr_b, h->dep[h->cur], dep_b, hipMemcpyDeviceToHost));
	HIP_TRY(h, hipMemcpy(d + hdr + irr_b + dep_b, h->states, st_b, hipMemcpyDeviceToHost));
	*bytes = need;
	return DDGI_OK;
}

ddgi_status ddgi_load_state(void* handle, const void* host_src, size_t bytes) {
	if(!handle || !host_src) return fail(H(handle), DDGI_ERROR_INVALID_ARGUMENT, "null argument");
	Handle* h = H(handle);
	const size_t irr_b = size_t(h->P) * 64 * 16, dep_b = size_t(h->P) * 256 * 8, st_b = size_t(h->P) * 4;
	const size_t hdr = 16 * 4;
	if(bytes != hdr + irr_b + dep_b + st_b) return fail(h, DDGI_ERROR_INVALID_STATE, "state size does not match this grid");
	uint32_t head[16];
	std::memcpy(head, host_src, hdr);
	if(head[0] != kStateMagic || head[1] != kStateVersion || head[2] != h->P)
		return fail(h, DDGI_ERROR_INVALID_STATE, "state belongs to another grid or version");
	ddgi_status st = set_device(h);
	if(st != DDGI_OK) return st;
	HIP_TRY(h, hipStreamSynchronize(h->last_stream));
	const char* s = static_cast<const char*>(host_src);
	HIP_TRY(h, hipMemcpy(h->irr[h->cur], s + hdr, irr_b, hipMemcpyHostToDevice));
	HIP_TRY(h, hipMemcpy(h->dep[h->cur], s + hdr + irr_b, dep_b, hipMemcpyHostToDevice));
	HIP_TRY(h, hipMemcpy(h->states, s + hdr + irr_b + dep_b, st_b, hipMemcpyHostToDevice));
	SelectState ss{head[6], head[7], 0, 0};
	HIP_TRY(h, hipMemcpy(h->sel, &ss, sizeof(ss), hipMemcpyHostToDevice));
	std::memcpy(&h->params.hysteresis, &head[4], 4);
	std::memcpy(&h->uploaded_hysteresis, &head[8], 4);
	h->updated_probes = head[5];
	h->indices_are_iota = false;
	return DDGI_OK;
}

}  // extern "C"
