// Host-side state of one probe volume (the object behind a `void*` handle of include/ddgi.h) and
// the helpers shared by the single-volume entry points (api.hip) and the batched update
// (batch.hip).  Internal to libddgi.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "bvh.hpp"
#include "diff.hpp"
#include "kernels.hpp"

namespace ddgi {
namespace host {

struct DeviceBuffer {
	void* p = nullptr;
	size_t bytes = 0;
};

struct Handle {
	ddgi_grid_params params{};
	ddgi_options opts{};
	int device = 0;
	uint32_t cus = 0;  // compute units of the device
	uint32_t P = 0;
	uint32_t max_rays = 0;
	// atlases: [cur] is exposed, [cur ^ 1] receives the next update
	float4* irr[2] = {nullptr, nullptr};
	float2* dep[2] = {nullptr, nullptr};
	int cur = 0;
	uint32_t* states = nullptr;
	uint32_t* indices = nullptr;
	SelectState* sel = nullptr;
	bool indices_are_iota = false;
	float4* rays = nullptr;
	float4* hits = nullptr;
	uint32_t* shade_list = nullptr;   // P * max_rays entries
	uint32_t* shade_count = nullptr;  // [0] shade list, [kMissCountSlot] miss list
	uint32_t* miss_list = nullptr;    // P * max_rays entries, allocated with the atmospheric sky
	uint32_t* shade_keys = nullptr;   // P * max_rays: cell key of each shade-list entry
	uint32_t* sort_keys = nullptr;    // P * max_rays: keys / entries bucketed by coarse digit
	uint32_t* sort_list = nullptr;
	uint32_t* sort_counts = nullptr;  // kShadeSortCounts
	// reflection pass scratch (ddgi_reflection), grown to the largest chunk of texels seen
	uint32_t refl_cap = 0;
	float4* refl_hits = nullptr;
	uint32_t* refl_list = nullptr;
	uint32_t* refl_keys = nullptr;
	uint32_t* refl_bkeys = nullptr;
	uint32_t* refl_blist = nullptr;
	uint32_t* refl_count = nullptr;
	uint32_t* refl_counts = nullptr;
	float4* fib = nullptr;      // kMaxRays entries of fib(i, n) for the current n
	uint16_t* perm = nullptr;   // kMaxRays
	float4* fib512 = nullptr;   // probe-init directions fib(i, n), i < 512
	float4* dirs = nullptr;     // world directions, kMaxRays
	uint32_t fib_n = 0;
	float* wt = nullptr;
	float* inv_sum = nullptr;
	void* w_hi = nullptr;
	void* w_lo = nullptr;
	unsigned long long* visit = nullptr;  // traversal counters (statistics runs)
	// scene
	std::vector<DeviceBuffer> scene_bufs;
	SceneDev sd{};
	bool has_scene = false;
	BvhHost bvh_meta;
	size_t bvh_bytes = 0;
	// instance moves (ddgi_update_instances): the scene's instance records, the triangles'
	// object-space vertices in BVH slot order and the wide nodes by level; the moved instances'
	// transforms (12 floats), triangle meta words and InstanceDev records pass through a pinned
	// staging area (`move_host`, reused once `move_done` has fired) to `move_xf` / `move_meta`
	std::vector<ddgi_instance> instances;
	const float* obj_tris = nullptr;
	const float* slot_box = nullptr;  // the builder's piece bounds per slot (BvhHost::slot_box)
	float* rel_xf = nullptr;          // per instance: current transform after the inverse of the build's (3x4)
	std::vector<float> build_xf;      // per instance: the transform the BVH was built with
	const uint32_t* level_nodes = nullptr;
	float* move_xf = nullptr;
	uint32_t* move_meta = nullptr;
	void* move_host = nullptr;
	hipEvent_t move_done = nullptr;
	bool move_pending = false;
	LightConst light{};
	// timing
	hipEvent_t ev[6] = {};
	bool timing_pending = false;
	ddgi_timings last{};
	uint32_t last_count = 0xFFFFFFFFu;
	uint32_t updated_probes = 0;  // hysteresis ramp bookkeeping (src/IrradianceProbes.cpp:463)
	float uploaded_hysteresis = 0.0f;  // the value the kernels use (the reference's UBO copy)
	hipStream_t last_stream = nullptr;
	std::string err;
	bool has_metallic = false;  // any scene material with metallic != 0
	uint32_t scene_materials = 0;
	uint32_t batch_members = 0;  // batches holding this handle (destroy is refused while > 0)
	// differentiable update (ddgi_diff_*): buffers owned while prepared
	struct Diff {
		bool ready = false;
		std::vector<DeviceBuffer> bufs;
		DiffArgs d{};
		float4* prev = nullptr;     // irradiance before the prepared update
		float4* irr = nullptr;      // differentiated irradiance (forward output)
		float4* rad = nullptr;
		float* wt = nullptr;        // blend weights of the prepared frame (ray_order order)
		float* inv_sum = nullptr;
		float4* dirs = nullptr;
		uint32_t* count = nullptr;  // = P (all probes)
		float2* dep_scratch = nullptr;
		uint32_t loss_parts = 0;
	} diff;
};

// Records `msg` as the handle's (and the thread's) last error and returns `s`.
ddgi_status fail(Handle* h, ddgi_status s, const std::string& msg);

#define HIP_TRY(h, expr)                                                                               \
	do {                                                                                               \
		hipError_t e_ = (expr);                                                                        \
		if(e_ != hipSuccess) {                                                                         \
			return ::ddgi::host::fail((h), e_ == hipErrorOutOfMemory ? DDGI_ERROR_OUT_OF_MEMORY : DDGI_ERROR_HIP, \
									  std::string(#expr) + ": " + hipGetErrorString(e_));              \
		}                                                                                              \
	} while(0)

template <class T>
hipError_t dalloc(T** p, size_t bytes) {
	void* q = nullptr;
	hipError_t e = hipMalloc(&q, bytes < 16 ? 16 : bytes);
	*p = static_cast<T*>(q);
	return e;
}

inline Handle* H(void* p) { return static_cast<Handle*>(p); }

// Grid constants of the kernels (cell size, ray tmax, the hysteresis the kernels use).
GridConst make_grid_const(const Handle& h);
// Makes the handle's device current.
ddgi_status set_device(Handle* h);
// Uploads the spherical-Fibonacci tables and the ray order for the current rays_per_probe.
ddgi_status ensure_fib(Handle* h);
// Validates a row-major 3x3 orientation.
ddgi_status make_rotation(const float* o, Rot& R);
// Hysteresis ramp of src/IrradianceProbes.cpp:462-476 after an update of `updated` probes.
void advance_hysteresis(Handle* h, uint32_t updated);

}  // namespace host
}  // namespace ddgi
