// Batched update of many probe volumes (ddgi_batch_* of include/ddgi.h).
//
// The reference updates one probe grid per IrradianceProbes::update call
// (src/IrradianceProbes.cpp:426-594).  Small volumes cannot fill 256 CUs and a sequence of
// per-volume launches pays a launch gap and a tail per kernel.  Here one update of N volumes is a
// single hipGraph launch of nine kernels:
//   k_batch_begin    frame directions of every volume, reset of the work-queue heads
//   k_batch_weights  blend weights of every volume (grid: 256 rows x volumes)
//   k_batch_trace    persistent primary traversal over the work items of all volumes
//   k_sort_*         (four kernels, grid y = volume) every volume's shade list ordered by hit cell
//   k_batch_shade    shading + shadow rays, one thread per entry every volume's list can hold
//   k_batch_blend    blend + border copy, one workgroup per probe tile of every volume
// The graph's kernel arguments are fixed; what changes per update (orientation, which atlas is
// exposed, hysteresis, light, scene pointers) lives in a device array of BatchVolume records that
// is copied from a pinned ring of host records before each launch.  Each per-ray and per-probe
// computation is the device function of the single-volume path, so results are bit-identical.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <numeric>
#include <vector>

#include "handle.hpp"

using namespace ddgi;
using namespace ddgi::host;

namespace {

constexpr int kRing = 4;  // pinned host copies of the volume records in flight

struct Batch {
	std::vector<Handle*> vols;
	int device = 0;
	int cus = 0;
	int wall_khz = 0;
	ddgi_batch_config cfg{};
	BatchVolume* d_vols = nullptr;
	BatchVolume* h_ring = nullptr;  // kRing x count records, pinned
	hipEvent_t ring_ev[kRing] = {};
	int ring_pos = 0;
	BatchItem* d_trace_items = nullptr;
	size_t trace_cap = 0;
	BatchItem* d_shade_items = nullptr;
	size_t shade_cap = 0;
	BatchItem* d_blend_items = nullptr;
	size_t blend_cap = 0;
	BatchControl* d_ctl = nullptr;
	// graph of the current configuration
	hipGraphExec_t exec = nullptr;
	bool exec_valid = false;
	uint32_t exec_mode = 0;
	std::vector<uint64_t> exec_shape;  // (P, n) per volume the work items were built for
	hipStream_t cap = nullptr;
	hipEvent_t ev0 = nullptr, ev1 = nullptr;
	bool timing_pending = false;
	bool has_run = false;
};

Batch* B(void* p) { return static_cast<Batch*>(p); }

ddgi_status bfail(ddgi_status s, const std::string& msg) { return fail(nullptr, s, msg); }

#define BHIP_TRY(expr)                                                                              \
	do {                                                                                            \
		hipError_t e_ = (expr);                                                                     \
		if(e_ != hipSuccess)                                                                        \
			return bfail(e_ == hipErrorOutOfMemory ? DDGI_ERROR_OUT_OF_MEMORY : DDGI_ERROR_HIP,     \
						 std::string(#expr) + ": " + hipGetErrorString(e_));                        \
	} while(0)

ddgi_status check_config(const ddgi_batch_config& c) {
	if(c.trace_block != 64 && c.trace_block != 128 && c.trace_block != 256)
		return bfail(DDGI_ERROR_INVALID_ARGUMENT, "trace_block must be 64, 128 or 256");
	if(c.trace_item_blocks < 1 || c.trace_item_blocks > 64)
		return bfail(DDGI_ERROR_INVALID_ARGUMENT, "trace_item_blocks must be in [1, 64]");
	if(c.blend_probes != 4 && c.blend_probes != 8 && c.blend_probes != 16 && c.blend_probes != 32 && c.blend_probes != 64)
		return bfail(DDGI_ERROR_INVALID_ARGUMENT, "blend_probes must be 4, 8, 16, 32 or 64");
	return DDGI_OK;
}

uint32_t mode_of(uint32_t flags) {
	return (flags & DDGI_UPDATE_BLEND_BF16) ? kBlendBF16 : ((flags & DDGI_UPDATE_BLEND_FP16) ? kBlendFP16 : kBlendFP32);
}

std::vector<uint64_t> batch_shape(const Batch& b) {
	std::vector<uint64_t> v;
	for(const Handle* h : b.vols) v.push_back(uint64_t(h->P) << 32 | h->params.rays_per_probe);
	return v;
}

void destroy_exec(Batch* b) {
	if(b->exec) (void)hipGraphExecDestroy(b->exec);
	b->exec = nullptr;
	b->exec_valid = false;
}

template <class T>
ddgi_status ensure_items(T** buf, size_t* cap, size_t need) {
	if(need <= *cap) return DDGI_OK;
	if(*buf) BHIP_TRY(hipFree(*buf));
	*buf = nullptr;
	*cap = 0;
	BHIP_TRY(dalloc(buf, need * sizeof(T)));
	*cap = need;
	return DDGI_OK;
}

// Work items of the current configuration.  Volumes are queued largest BVH first (their rays
// cost the most), so the small volumes fill the tail of the persistent traversal.
ddgi_status build_items(Batch* b, uint32_t mode, BatchLaunch& L) {
	const uint32_t nv = uint32_t(b->vols.size());
	std::vector<uint32_t> order(nv);
	std::iota(order.begin(), order.end(), 0u);
	std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
		return b->vols[x]->sd.tri_count > b->vols[y]->sd.tri_count;
	});
	const uint32_t per_item = b->cfg.trace_block * b->cfg.trace_item_blocks;
	const uint32_t probes = mode == kBlendFP32 ? b->cfg.blend_probes : kMfmaProbes;
	std::vector<BatchItem> ti, si, bi;
	for(uint32_t v : order) {
		const Handle* h = b->vols[v];
		const uint64_t threads = uint64_t(h->P) * h->params.rays_per_probe;  // also the shade list's capacity
		for(uint64_t f = 0; f < threads; f += per_item) ti.push_back(BatchItem{v, uint32_t(f)});
		for(uint64_t f = 0; f < threads; f += kShadeBlock) si.push_back(BatchItem{v, uint32_t(f)});
		for(uint32_t s = 0; s < h->P; s += probes) bi.push_back(BatchItem{v, s});
	}
	ddgi_status st = ensure_items(&b->d_trace_items, &b->trace_cap, ti.size());
	if(st != DDGI_OK || (st = ensure_items(&b->d_shade_items, &b->shade_cap, si.size())) != DDGI_OK ||
	   (st = ensure_items(&b->d_blend_items, &b->blend_cap, bi.size())) != DDGI_OK)
		return st;
	BHIP_TRY(hipMemcpy(b->d_trace_items, ti.data(), ti.size() * sizeof(BatchItem), hipMemcpyHostToDevice));
	BHIP_TRY(hipMemcpy(b->d_shade_items, si.data(), si.size() * sizeof(BatchItem), hipMemcpyHostToDevice));
	BHIP_TRY(hipMemcpy(b->d_blend_items, bi.data(), bi.size() * sizeof(BatchItem), hipMemcpyHostToDevice));
	L.vols = b->d_vols;
	L.trace_items = b->d_trace_items;
	L.blend_items = b->d_blend_items;
	L.ctl = b->d_ctl;
	L.volume_count = nv;
	L.trace_item_count = uint32_t(ti.size());
	L.trace_item_threads = per_item;
	L.blend_item_count = uint32_t(bi.size());
	L.blend_probes = probes;
	L.blend_mode = mode;
	L.shade_items = b->d_shade_items;
	L.shade_item_count = uint32_t(si.size());
	return DDGI_OK;
}

// (Re)captures the graph when the configuration, blend mode or a volume's shape changed since the
// last capture.
ddgi_status ensure_graph(Batch* b, uint32_t mode) {
	const std::vector<uint64_t> shape = batch_shape(*b);
	if(b->exec_valid && b->exec_mode == mode && b->exec_shape == shape) return DDGI_OK;
	BHIP_TRY(hipDeviceSynchronize());  // the item buffers may be in use by an earlier launch
	destroy_exec(b);
	BatchLaunch L{};
	ddgi_status st = build_items(b, mode, L);
	if(st != DDGI_OK) return st;
	int r_primary = 0;
	BHIP_TRY(batch_trace_resident(b->cfg.trace_block, &r_primary));
	const uint32_t g_primary = uint32_t(std::max(1, r_primary) * b->cus);
	BHIP_TRY(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
	hipError_t e = launch_batch_begin(L, b->cap);
	if(e == hipSuccess) e = launch_batch_weights(L, b->cap);
	uint64_t max_list = 0;
	for(const Handle* h : b->vols) max_list = std::max<uint64_t>(max_list, uint64_t(h->P) * h->params.rays_per_probe);
	if(e == hipSuccess) e = launch_batch_trace(L, b->cfg.trace_block, std::min(g_primary, L.trace_item_count), b->cap);
	if(e == hipSuccess) e = launch_batch_sort(L, max_list, b->cap);
	if(e == hipSuccess) e = launch_batch_shade(L, b->cap);
	if(e == hipSuccess) e = launch_batch_blend(L, b->cap);
	hipGraph_t graph = nullptr;
	const hipError_t e_end = hipStreamEndCapture(b->cap, &graph);
	if(e != hipSuccess || e_end != hipSuccess) {
		if(graph) (void)hipGraphDestroy(graph);
		return bfail(DDGI_ERROR_HIP, std::string("graph capture failed: ") + hipGetErrorString(e != hipSuccess ? e : e_end));
	}
	e = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
	(void)hipGraphDestroy(graph);
	if(e != hipSuccess) {
		b->exec = nullptr;
		return bfail(DDGI_ERROR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
	}
	b->exec_valid = true;
	b->exec_mode = mode;
	b->exec_shape = shape;
	return DDGI_OK;
}

void fill_volume(const Handle& h, const Rot& R, BatchVolume& v) {
	const GridConst g = make_grid_const(h);
	const int nxt = h.cur ^ 1;
	v = BatchVolume{};
	TraceFrameArgs& a = v.trace;
	a.scene = h.sd;
	a.grid = g;
	a.light = h.light;
	a.light.sky_model = DDGI_SKY_CONSTANT;  // batches run no sky pass (check_update_args refuses the atmosphere)
	a.miss_list = nullptr;
	a.indices = h.indices;
	a.count = &h.sel->count;
	a.dirs = h.dirs;
	a.perm = h.perm;
	a.states = h.states;
	a.irr = h.irr[h.cur];
	a.dep = h.dep[h.cur];
	a.hits = h.hits;
	a.rays = h.rays;
	a.shade_list = h.shade_list;
	a.shade_count = h.shade_count;
	a.max_slots = h.P;
	uint32_t key_shift = 0, key_bits = 0;  // the shade list is ordered by hit cell, as in ddgi_update
	shade_key_layout(g, &key_shift, &key_bits);
	a.shade_keys = h.shade_keys;
	a.key_shift = key_shift;
	a.seg_cap = 0;  // one shade list per volume
	v.sort = SortJob{h.shade_keys, h.shade_list, h.shade_count, h.sort_counts, h.sort_keys, h.sort_list,
					 h.shade_count + kShadeTotalSlot, sort_fine_bits(3 * key_bits), 1u, 0u, 0u};
	BlendArgs& bl = v.blend;
	bl.grid = g;
	bl.indices = h.indices;
	bl.count = &h.sel->count;
	bl.rays = h.rays;
	bl.wt = h.wt;
	bl.inv_sum = h.inv_sum;
	bl.w_hi16 = h.w_hi;
	bl.w_lo16 = h.w_lo;
	bl.states = h.states;
	bl.irr_prev = h.irr[h.cur];
	bl.dep_prev = h.dep[h.cur];
	bl.irr_out = h.irr[nxt];
	bl.dep_out = h.dep[nxt];
	bl.max_slots = h.P;
	bl.flags = 0;
	v.R = R;
	v.fib = h.fib;
	v.dirs = h.dirs;
	v.perm = h.perm;
	v.wt = h.wt;
	v.inv_sum = h.inv_sum;
	v.w_hi = h.w_hi;
	v.w_lo = h.w_lo;
	v.threads = h.P * g.n;
	v.n_pad = (g.n + 63u) & ~63u;
}

void destroy_batch(Batch* b) {
	destroy_exec(b);
	for(Handle* h : b->vols) --h->batch_members;
	(void)hipFree(b->d_vols);
	(void)hipFree(b->d_trace_items);
	(void)hipFree(b->d_shade_items);
	(void)hipFree(b->d_blend_items);
	(void)hipFree(b->d_ctl);
	if(b->h_ring) (void)hipHostFree(b->h_ring);
	for(auto& e : b->ring_ev)
		if(e) (void)hipEventDestroy(e);
	if(b->ev0) (void)hipEventDestroy(b->ev0);
	if(b->ev1) (void)hipEventDestroy(b->ev1);
	if(b->cap) (void)hipStreamDestroy(b->cap);
	delete b;
}

ddgi_status run_batch(Batch* b, const float* orientations, uint32_t flags, hipStream_t s) {
	const uint32_t mode = mode_of(flags);
	const uint32_t nv = uint32_t(b->vols.size());
	for(uint32_t v = 0; v < nv; ++v) {
		Handle* h = b->vols[v];
		if(!h->has_scene) return bfail(DDGI_ERROR_NO_SCENE, "a volume of the batch has no scene");
		if(h->device != b->device) return bfail(DDGI_ERROR_INVALID_STATE, "volume moved to another device");
		ddgi_status st = ensure_fib(h);
		if(st != DDGI_OK) return bfail(st, h->err);
	}
	ddgi_status st = ensure_graph(b, mode);
	if(st != DDGI_OK) return st;
	for(Handle* h : b->vols) {
		if(!h->indices_are_iota) {
			BHIP_TRY(launch_iota(h->indices, h->P, h->sel, s));
			h->indices_are_iota = true;
		}
		if(flags & DDGI_UPDATE_HYSTERESIS_RAMP) advance_hysteresis(h, h->P);
	}
	BatchVolume* ring = b->h_ring + size_t(b->ring_pos) * nv;
	BHIP_TRY(hipEventSynchronize(b->ring_ev[b->ring_pos]));  // its previous copy has been consumed
	for(uint32_t v = 0; v < nv; ++v) {
		Rot R;
		std::memcpy(R.r, orientations + size_t(v) * 9, sizeof(R.r));
		fill_volume(*b->vols[v], R, ring[v]);
	}
	BHIP_TRY(hipMemcpyAsync(b->d_vols, ring, sizeof(BatchVolume) * nv, hipMemcpyHostToDevice, s));
	BHIP_TRY(hipEventRecord(b->ring_ev[b->ring_pos], s));
	b->ring_pos = (b->ring_pos + 1) % kRing;
	BHIP_TRY(hipEventRecord(b->ev0, s));
	BHIP_TRY(hipGraphLaunch(b->exec, s));
	BHIP_TRY(hipEventRecord(b->ev1, s));
	for(Handle* h : b->vols) {
		h->cur ^= 1;  // publish
		h->last_count = h->P;
		h->last_stream = s;
		h->timing_pending = false;
		h->last = ddgi_timings{};  // phase times of a batched update are the batch's (ddgi_batch_get_timings)
		h->last.probes_updated = h->P;
		h->last.rays_traced = h->P * h->params.rays_per_probe;
	}
	b->timing_pending = true;
	b->has_run = true;
	return DDGI_OK;
}

ddgi_status check_update_args(Batch* b, const float* orientations, uint32_t flags) {
	if(!(flags & DDGI_UPDATE_ALL_PROBES))
		return bfail(DDGI_ERROR_UNSUPPORTED, "batched updates trace every probe: pass DDGI_UPDATE_ALL_PROBES");
	if(flags & DDGI_UPDATE_COUNT_VISITS) return bfail(DDGI_ERROR_UNSUPPORTED, "visit counting is not batched");
	if((flags & DDGI_UPDATE_BLEND_BF16) && (flags & DDGI_UPDATE_BLEND_FP16))
		return bfail(DDGI_ERROR_INVALID_ARGUMENT, "BF16 and FP16 blends are exclusive");
	if(!orientations) return bfail(DDGI_ERROR_INVALID_ARGUMENT, "null orientations");
	for(const Handle* h : b->vols)
		if(h->light.sky_model != DDGI_SKY_CONSTANT)
			return bfail(DDGI_ERROR_UNSUPPORTED, "batched updates support the constant sky only");
	for(size_t i = 0; i < b->vols.size() * 9; ++i)
		if(!std::isfinite(orientations[i])) return bfail(DDGI_ERROR_INVALID_ARGUMENT, "non-finite orientation");
	return DDGI_OK;
}

ddgi_status set_batch_device(Batch* b) {
	int cur = -1;
	BHIP_TRY(hipGetDevice(&cur));
	if(cur != b->device) BHIP_TRY(hipSetDevice(b->device));
	return DDGI_OK;
}

// Device copies of what an update changes, for the autotuner to put back.
struct Snapshot {
	std::vector<void*> bufs;
	struct Host {
		float hysteresis, uploaded;
		uint32_t updated_probes;
	};
	std::vector<Host> host;
	~Snapshot() {
		for(void* p : bufs) (void)hipFree(p);
	}
};

ddgi_status take_snapshot(Batch* b, Snapshot& snap, hipStream_t s) {
	for(Handle* h : b->vols) {
		const size_t sz[3] = {size_t(h->P) * 64 * sizeof(float4), size_t(h->P) * 256 * sizeof(float2), size_t(h->P) * 4};
		const void* src[3] = {h->irr[h->cur], h->dep[h->cur], h->states};
		for(int k = 0; k < 3; ++k) {
			void* p = nullptr;
			BHIP_TRY(hipMalloc(&p, sz[k]));
			snap.bufs.push_back(p);
			BHIP_TRY(hipMemcpyAsync(p, src[k], sz[k], hipMemcpyDeviceToDevice, s));
		}
		snap.host.push_back(Snapshot::Host{h->params.hysteresis, h->uploaded_hysteresis, h->updated_probes});
	}
	return DDGI_OK;
}

ddgi_status restore_snapshot(Batch* b, const Snapshot& snap, hipStream_t s) {
	size_t k = 0;
	for(size_t v = 0; v < b->vols.size(); ++v) {
		Handle* h = b->vols[v];
		const size_t sz[3] = {size_t(h->P) * 64 * sizeof(float4), size_t(h->P) * 256 * sizeof(float2), size_t(h->P) * 4};
		void* dst[3] = {h->irr[h->cur], h->dep[h->cur], h->states};
		for(int i = 0; i < 3; ++i, ++k) BHIP_TRY(hipMemcpyAsync(dst[i], snap.bufs[k], sz[i], hipMemcpyDeviceToDevice, s));
		h->params.hysteresis = snap.host[v].hysteresis;
		h->uploaded_hysteresis = snap.host[v].uploaded;
		h->updated_probes = snap.host[v].updated_probes;
	}
	BHIP_TRY(hipStreamSynchronize(s));
	return DDGI_OK;
}

ddgi_status time_config(Batch* b, const ddgi_batch_config& c, const float* orientations, uint32_t flags,
						uint32_t iterations, hipStream_t s, float* ms) {
	b->cfg = c;
	b->exec_valid = false;
	ddgi_status st = run_batch(b, orientations, flags, s);  // capture + warm-up
	if(st != DDGI_OK) return st;
	hipEvent_t a = nullptr, z = nullptr;
	BHIP_TRY(hipEventCreate(&a));
	hipError_t e = hipEventCreate(&z);
	if(e == hipSuccess) e = hipEventRecord(a, s);
	for(uint32_t i = 0; i < iterations && e == hipSuccess && st == DDGI_OK; ++i) st = run_batch(b, orientations, flags, s);
	if(e == hipSuccess) e = hipEventRecord(z, s);
	if(e == hipSuccess) e = hipEventSynchronize(z);
	float total = 0.0f;
	if(e == hipSuccess) e = hipEventElapsedTime(&total, a, z);
	(void)hipEventDestroy(a);
	if(z) (void)hipEventDestroy(z);
	if(st != DDGI_OK) return st;
	if(e != hipSuccess) return bfail(DDGI_ERROR_HIP, std::string("autotune timing: ") + hipGetErrorString(e));
	*ms = total / float(iterations);
	return DDGI_OK;
}

}  // namespace

extern "C" {

ddgi_status ddgi_batch_create(void* const* handles, uint32_t count, const ddgi_batch_config* config, void** out_batch) {
	if(!handles || !out_batch || count == 0) return bfail(DDGI_ERROR_INVALID_ARGUMENT, "null argument or empty batch");
	*out_batch = nullptr;
	if(count > kMaxBatchVolumes) return bfail(DDGI_ERROR_UNSUPPORTED, "more than 256 volumes in one batch");
	ddgi_batch_config cfg{128, 4, 16, {}};
	if(config) cfg = *config;
	ddgi_status st = check_config(cfg);
	if(st != DDGI_OK) return st;
	for(uint32_t i = 0; i < count; ++i) {
		if(!handles[i]) return bfail(DDGI_ERROR_INVALID_ARGUMENT, "null handle in batch");
		for(uint32_t j = 0; j < i; ++j)
			if(handles[j] == handles[i]) return bfail(DDGI_ERROR_INVALID_ARGUMENT, "a handle appears twice in the batch");
		if(H(handles[i])->device != H(handles[0])->device)
			return bfail(DDGI_ERROR_INVALID_ARGUMENT, "all volumes of a batch must be on one device");
	}
	Batch* b = new(std::nothrow) Batch;
	if(!b) return bfail(DDGI_ERROR_OUT_OF_MEMORY, "host allocation failed");
	b->cfg = cfg;
	b->device = H(handles[0])->device;
	for(uint32_t i = 0; i < count; ++i) {
		b->vols.push_back(H(handles[i]));
		++H(handles[i])->batch_members;
	}
	auto cleanup = [&](ddgi_status s2) {  // the message is already recorded
		destroy_batch(b);
		return s2;
	};
	if((st = set_batch_device(b)) != DDGI_OK) return cleanup(st);
	hipDeviceProp_t prop;
	hipError_t e = hipGetDeviceProperties(&prop, b->device);
	if(e == hipSuccess) b->cus = prop.multiProcessorCount;
	if(e == hipSuccess) e = hipDeviceGetAttribute(&b->wall_khz, hipDeviceAttributeWallClockRate, b->device);
	if(e == hipSuccess) e = dalloc(&b->d_vols, sizeof(BatchVolume) * count);
	if(e == hipSuccess) e = dalloc(&b->d_ctl, sizeof(BatchControl));
	if(e == hipSuccess) e = hipMemset(b->d_ctl, 0, sizeof(BatchControl));
	if(e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&b->h_ring), sizeof(BatchVolume) * count * kRing, hipHostMallocDefault);
	for(int i = 0; i < kRing && e == hipSuccess; ++i) e = hipEventCreate(&b->ring_ev[i]);
	if(e == hipSuccess) e = hipEventCreate(&b->ev0);
	if(e == hipSuccess) e = hipEventCreate(&b->ev1);
	if(e == hipSuccess) e = hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking);
	if(e != hipSuccess) {
		fail(nullptr, e == hipErrorOutOfMemory ? DDGI_ERROR_OUT_OF_MEMORY : DDGI_ERROR_HIP,
			 std::string("batch set-up: ") + hipGetErrorString(e));
		return cleanup(e == hipErrorOutOfMemory ? DDGI_ERROR_OUT_OF_MEMORY : DDGI_ERROR_HIP);
	}
	if(b->wall_khz <= 0) b->wall_khz = 100000;  // the 100 MHz constant clock of CDNA parts
	*out_batch = b;
	return DDGI_OK;
}

ddgi_status ddgi_batch_destroy(void* batch) {
	if(!batch) return bfail(DDGI_ERROR_INVALID_ARGUMENT, "null batch");
	Batch* b = B(batch);
	(void)set_batch_device(b);
	(void)hipDeviceSynchronize();
	destroy_batch(b);
	return DDGI_OK;
}

ddgi_status ddgi_batch_set_config(void* batch, const ddgi_batch_config* config) {
	if(!batch || !config) return bfail(DDGI_ERROR_INVALID_ARGUMENT, "null argument");
	ddgi_status st = check_config(*config);
	if(st != DDGI_OK) return st;
	Batch* b = B(batch);
	b->cfg = *config;
	b->exec_valid = false;
	return DDGI_OK;
}

ddgi_status ddgi_batch_get_config(void* batch, ddgi_batch_config* out) {
	if(!batch || !out) return bfail(DDGI_ERROR_INVALID_ARGUMENT, "null argument");
	*out = B(batch)->cfg;
	return DDGI_OK;
}

ddgi_status ddgi_batch_update(void* batch, const float* orientations, uint32_t flags, void* stream) {
	if(!batch) return bfail(DDGI_ERROR_INVALID_ARGUMENT, "null batch");
	Batch* b = B(batch);
	ddgi_status st = check_update_args(b, orientations, flags);
	if(st != DDGI_OK || (st = set_batch_device(b)) != DDGI_OK) return st;
	return run_batch(b, orientations, flags, static_cast<hipStream_t>(stream));
}

ddgi_status ddgi_batch_get_timings(void* batch, float* total_ms, float* volume_ms, uint32_t count) {
	if(!batch || !total_ms) return bfail(DDGI_ERROR_INVALID_ARGUMENT, "null argument");
	Batch* b = B(batch);
	if(!b->has_run) return bfail(DDGI_ERROR_INVALID_STATE, "no batched update has run");
	if(volume_ms && count != b->vols.size()) return bfail(DDGI_ERROR_INVALID_ARGUMENT, "count must equal the volume count");
	ddgi_status st = set_batch_device(b);
	if(st != DDGI_OK) return st;
	BHIP_TRY(hipEventSynchronize(b->ev1));
	BHIP_TRY(hipEventElapsedTime(total_ms, b->ev0, b->ev1));
	if(volume_ms) {
		BatchControl c;
		BHIP_TRY(hipMemcpy(&c, b->d_ctl, sizeof(c), hipMemcpyDeviceToHost));
		for(uint32_t v = 0; v < count; ++v)
			volume_ms[v] = c.done[v] > c.t0 ? float(double(c.done[v] - c.t0) / double(b->wall_khz)) : 0.0f;
	}
	return DDGI_OK;
}

ddgi_status ddgi_batch_autotune(void* batch, const float* orientations, uint32_t flags, uint32_t iterations, void* stream,
								ddgi_batch_config* best, ddgi_batch_tune_record* records, uint32_t max_records,
								uint32_t* record_count) {
	if(!batch) return bfail(DDGI_ERROR_INVALID_ARGUMENT, "null batch");
	Batch* b = B(batch);
	ddgi_status st = check_update_args(b, orientations, flags);
	if(st != DDGI_OK || (st = set_batch_device(b)) != DDGI_OK) return st;
	if(iterations == 0) iterations = 3;
	hipStream_t s = static_cast<hipStream_t>(stream);
	if(record_count) *record_count = 0;
	BHIP_TRY(hipStreamSynchronize(s));
	Snapshot snap;
	if((st = take_snapshot(b, snap, s)) != DDGI_OK) return st;
	const ddgi_batch_config start = b->cfg;
	uint32_t nrec = 0;
	auto record = [&](const ddgi_batch_config& c, float ms, uint32_t phase) {
		if(records && nrec < max_records) records[nrec] = ddgi_batch_tune_record{c, ms, phase};
		++nrec;
	};
	ddgi_batch_config win = start;
	float win_ms = 3.4e38f;
	// phase 0: traversal workgroup size x work-item size
	const uint32_t blocks[] = {64, 128, 256};
	const uint32_t items[] = {8, 16, 32};
	for(uint32_t bs : blocks)
		for(uint32_t it : items) {
			ddgi_batch_config c = start;
			c.trace_block = bs;
			c.trace_item_blocks = it;
			float ms = 0.0f;
			if((st = time_config(b, c, orientations, flags, iterations, s, &ms)) != DDGI_OK) break;
			record(c, ms, 0);
			if(ms < win_ms) {
				win_ms = ms;
				win = c;
			}
		}
	// phase 1: FP32 blend tile (the 16-bit blends have one tile shape); 64 = the 16-bit blends'
	// 64-probe workgroups, which give the same sums as the smaller tiles
	if(st == DDGI_OK && mode_of(flags) == kBlendFP32) {
		const ddgi_batch_config base = win;
		const uint32_t tiles[] = {4, 8, 16, 32, 64};
		for(uint32_t p : tiles) {
			ddgi_batch_config c = base;
			c.blend_probes = p;
			float ms = 0.0f;
			if((st = time_config(b, c, orientations, flags, iterations, s, &ms)) != DDGI_OK) break;
			record(c, ms, 1);
			if(ms < win_ms) {
				win_ms = ms;
				win = c;
			}
		}
	}
	const ddgi_status rs = restore_snapshot(b, snap, s);
	b->cfg = st == DDGI_OK ? win : start;
	b->exec_valid = false;
	b->has_run = false;
	if(st != DDGI_OK) return st;
	if(rs != DDGI_OK) return rs;
	if(best) *best = win;
	if(record_count) *record_count = std::min(nrec, records ? max_records : 0u);
	return DDGI_OK;
}

}  // extern "C"
