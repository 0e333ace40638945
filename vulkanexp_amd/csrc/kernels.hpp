// Kernel-argument structs and launcher declarations shared by the HIP translation units and the
// C-ABI implementation.  Everything a kernel reads is passed by value or as raw device pointers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ddgi.h"
#include "bvh.hpp"

namespace ddgi {

constexpr float kPiF = 3.1415926538f;  // src/shaders/common.glsl:4
constexpr uint32_t kMaxRays = 512;
constexpr uint32_t kIrrTexels = 36;    // 6x6 interior of the 8x8 tile
constexpr uint32_t kDepTexels = 196;   // 14x14 interior of the 16x16 tile
constexpr uint32_t kBlendTexels = kIrrTexels + kDepTexels;
constexpr uint32_t kWeightRows = 256;  // blend weight rows: 36 irradiance + 12 zero + 196 depth + 12 zero
constexpr uint32_t kTraceBlock = 128;  // threads per traversal block
constexpr uint32_t kPrimaryBlock = 256;  // threads per block of the probe update's primary traversal
constexpr uint32_t kShadeBlock = 64;     // threads per block of the shading kernels (k_shade, k_reflection_shade, k_batch_shade)

// Grid constants derived on the host (FP32, same operations as the oracle's ray generation).
struct GridConst {
	float emin[3];
	float cell[3];        // (max - min) / (res - 1)
	float cell_len;       // length(cell)
	float hysteresis;
	float sharpness;
	float shadow_bias;
	float tmax_primary;   // length(max - min)
	int32_t res[3];
	uint32_t n;           // rays per probe
	uint32_t probe_count;
};

struct Rot {
	float r[9];  // row-major, rows X, Y, Z
};

struct LightConst {
	float dir[3];    // toward the light
	float color[3];  // radiance
	float sky[3];    // constant miss radiance
	float sky_brightness;  // the light colour's 4th component: sun scale of the atmospheric sky
	uint32_t sky_model;    // DDGI_SKY_CONSTANT or DDGI_SKY_ATMOSPHERE
};

// Per-instance shading record: normal matrix transpose(inverse(M)) row-major, custom index
// (offset-table entry) and the global id of the instance's first triangle.
struct InstanceDev {
	float nm[9];
	uint32_t custom_index;
	uint32_t prim_base;
	uint32_t mask;
};

struct SceneDev {
	const float4* nodes;
	const float4* tris;
	const InstanceDev* inst;
	const uint32_t* offsets;   // 3 per entry (material, vertex offset, index offset)
	const uint32_t* indices;
	const float4* vertices;    // 4 float4 per vertex (64-byte record)
	const float* materials;    // 12 floats per material
	uint32_t tri_count;
	uint32_t node_count;
	uint32_t instance_count;
	uint32_t stack_depth;      // most stack entries a traversal needs (3 per level of the 4-wide tree)
};

// Internal hit record: (t, slot | back << 31, u, v); slot = triangle position in the BVH array,
// 0xFFFFFFFF = miss (t = -1).
// List counters, each on its own 128-byte line (device-scope atomics on one line serialise): the
// primary pass appends its front-face hits to kShadeSegments segments of the shade list, segment s
// counted at [kCounterStride * s] (workgroup b appends to segment b % kShadeSegments, so each XCD
// has its own counter); the miss list is counted at [kMissCountSlot]; the key sort stores the
// ordered list's length at [kShadeTotalSlot], which the shading reads.
constexpr uint32_t kShadeSegments = 8;
constexpr uint32_t kCounterStride = 32;
constexpr uint32_t kMissCountSlot = kCounterStride * kShadeSegments;
constexpr uint32_t kShadeTotalSlot = kMissCountSlot + kCounterStride;
constexpr uint32_t kListCounterBytes = (kShadeTotalSlot + kCounterStride) * 4;

struct TraceFrameArgs {
	SceneDev scene;
	GridConst grid;
	LightConst light;
	const uint32_t* indices;   // probe linear index per slot
	const uint32_t* count;     // device count of slots
	const float4* dirs;        // world ray directions of the frame
	const uint16_t* perm;      // thread order -> ray index (directionally coherent)
	const uint32_t* states;    // probe states (read by sampleProbes)
	const float4* irr;         // exposed irradiance atlas (tiled)
	const float2* dep;         // exposed depth atlas (tiled)
	float4* hits;
	float4* rays;              // (rgb, depth)
	uint32_t* shade_list;      // ray positions (slot * n + k) of the front-face hits, filled by the primary pass
	uint32_t* shade_count;     // list counters (kShadeSegments segments, miss list, ordered total)
	uint32_t* miss_list;       // atmospheric sky only: ray positions of the misses
	uint32_t max_slots;        // grid sized for this many slots
	// optional (statistics runs only): [0]/[1] primary node / triangle fetches, [2]/[3] shadow
	// node / triangle fetches, [4] shadow rays
	unsigned long long* visit_counters;
	// optional (differentiable prepare): 4 float4 per ray, rad = c0 + A a + (B a + C) L:
	// (c0, material bits), (A, lit bits), (B, 0), (C, 0); material 0xFFFFFFFF for misses/back faces
	float4* factors;
	// optional (shade-list ordering): the primary pass stores the cell key of each shade-list entry
	// (cell coordinates shifted right by key_shift, so that they stay below 64 and keys below 2^18)
	uint32_t* shade_keys;
	uint32_t key_shift;
	// > 0: the shade list is kShadeSegments segments of seg_cap entries (k_trace_primary); 0: one list
	uint32_t seg_cap;
};
constexpr uint32_t kShadeSortCounts = 1024 * 128 + 1024;  // (list tile, coarse digit) counters + digit totals

struct SelectState {
	uint32_t last_offset;
	uint32_t loop_index;
	uint32_t count;
	uint32_t pad;
};

struct BlendArgs {
	GridConst grid;
	const uint32_t* indices;
	const uint32_t* count;
	const float4* rays;         // slot-major (rgb, depth)
	const float* wt;            // FP32 weights [kWeightRows][n_pad]: irradiance rows 0..35, depth rows 48..243
	const float* inv_sum;       // [232]
	const void* w_lo16;         // MFMA paths: weights split hi/lo in 16-bit types
	const void* w_hi16;
	uint32_t* states;
	const float4* irr_prev;
	const float2* dep_prev;
	float4* irr_out;
	float2* dep_out;
	uint32_t max_slots;
	uint32_t flags;             // kBlendIrradianceOnly: no depth stores, no state update
};
constexpr uint32_t kBlendIrradianceOnly = 0x1;

// Final gather over a G-buffer (FinalGather.frag:38-77).
struct GatherArgs {
	SceneDev scene;
	GridConst grid;
	LightConst light;
	const uint32_t* states;
	const float4* irr;
	const float2* dep;
	const ddgi_gbuffer_texel* gbuffer;
	const float* direct;      // optional, one value per texel
	const float* reflection;  // optional, rgb per texel
	float* out;               // rgb per texel
	uint64_t count;
	float origin[3];          // camera position (the view vectors' origin)
	float sdir[3];            // normalize(light direction), computed on the host
};

// Ray-traced shadow pass (directLight.rgen) over G-buffer texels.
struct DirectLightArgs {
	SceneDev scene;
	const ddgi_gbuffer_texel* gbuffer;
	const float2* noise;     // optional: the blue-noise sample per texel
	const float4* previous;  // optional: the previous frame's output (its g, b pass through)
	float4* out;
	uint64_t count;
	float sdir[3];           // normalize(light direction), computed on the host
};

// Reflection pass (reflection.rgen) over G-buffer texels.
struct ReflectionArgs {
	SceneDev scene;
	GridConst grid;
	LightConst light;
	const uint32_t* states;
	const float4* irr;
	const float2* dep;
	const ddgi_gbuffer_texel* gbuffer;
	const float2* noise;  // optional: the blue-noise sample per texel
	float4* out;
	uint64_t count;
	float origin[3];      // camera position
	// the chunk of texels [first, first + chunk) being traced; front-face hits of the chunk are
	// queued (chunk-local texel index, cell key) and shaded in key order
	uint64_t first;
	uint32_t chunk;
	float4* hits;         // [chunk] hit records of the queued texels
	uint32_t* list;       // [chunk]
	uint32_t* keys;       // [chunk]
	uint32_t* list_count;
	uint32_t key_shift;
};

// A list to order by key (launch_key_sort): keys, entries and the device length; scratch: counts
// (kShadeSortCounts) and bucket_keys / bucket_list (the list's capacity each).
struct SortJob {
	const uint32_t* keys;
	uint32_t* list;
	const uint32_t* count;  // segment s's length at count[kCounterStride * s]
	uint32_t* counts;
	uint32_t* bucket_keys;
	uint32_t* bucket_list;
	uint32_t* total;        // optional: receives the list's length
	uint32_t fine_bits;     // key bits below the 10-bit coarse digit
	uint32_t segments;      // the input list: `segments` segments of seg_cap entries (1: contiguous)
	uint32_t seg_cap;
	uint32_t pad;
};
// The sort kernels' argument: one job, or (many != nullptr) job y of workgroup (x, y) at
// many + y * stride.
struct SortJobs {
	SortJob one;
	const char* many;
	uint32_t stride;
};
uint32_t sort_fine_bits(uint32_t key_bits);

// ---- launchers (return hipGetLastError of the launch) ----
// The frame's directions (as launch_frame_dirs) and thread -> ray order (kMaxRays entries; the first
// n sorted by the world elevation of the directions, ties by ray index, the rest unchanged).
hipError_t launch_frame_dirs_order(const Rot& R, const float4* fib, float4* dirs, uint32_t n, uint16_t* perm,
								   hipStream_t s, uint32_t* zero = nullptr);
// zero (optional): a counter cleared by the same launch
hipError_t launch_frame_dirs(const Rot& R, const float4* fib, float4* dirs, uint32_t count, hipStream_t s,
							 uint32_t* zero = nullptr);
// Orders job.list (keys job.keys of key_bits <= 18 bits; input in job.segments segments, output
// contiguous, at most `capacity` entries) in place: by the keys' high 10 bits over the whole list,
// and by the whole key inside each fixed chunk of 4096 entries (a bucket of equal high bits longer
// than a chunk is ordered chunk by chunk); entries with equal keys come in any order.  Scratch:
// job.counts (kShadeSortCounts), job.bucket_keys and job.bucket_list (capacity each).  The
// caller's job.fine_bits is ignored (set from key_bits).
hipError_t launch_key_sort(SortJob job, uint64_t capacity, uint32_t key_bits, hipStream_t s);
// Cell-key layout of the shade ordering for a grid: coordinate shift and bits per axis.
void shade_key_layout(const GridConst& g, uint32_t* shift, uint32_t* bits_per_axis);
hipError_t launch_clear(uint32_t* words, uint32_t n, hipStream_t s);
hipError_t launch_trace_primary(const TraceFrameArgs& a, hipStream_t s);
hipError_t launch_shade(const TraceFrameArgs& a, hipStream_t s);
// Atmospheric-sky radiance of the frame's misses (the miss list).
hipError_t launch_sky(const TraceFrameArgs& a, hipStream_t s);
hipError_t launch_trace_rays(const SceneDev& sc, const ddgi_ray* rays, ddgi_hit* hits, uint64_t count, uint32_t mask,
							 bool any_hit, bool cull, hipStream_t s);
// Instance moves: re-transform the triangles of instances [first, first + count) (xf: 12 floats,
// meta: mask | flip << 8 per moved instance), then refit the wide tree level by level, deepest
// first (slot_box: the builder's piece bounds per slot; rel_xf: per instance its motion since the
// build, 3x4).
hipError_t launch_retransform(float4* tris, const float* obj, uint32_t tri_count, const float* xf, const uint32_t* meta,
							  uint32_t first, uint32_t count, hipStream_t s);
hipError_t launch_refit_level(float4* nodes, const float4* tris, const float* slot_box, const float* rel_xf,
							  const uint32_t* level_nodes, uint32_t count, hipStream_t s);
hipError_t launch_gbuffer(const SceneDev& sc, const ddgi_ray* rays, uint64_t count, ddgi_gbuffer_texel* out, hipStream_t s);
hipError_t launch_gather(const GatherArgs& a, hipStream_t s);
hipError_t launch_direct_light(const DirectLightArgs& a, hipStream_t s);
hipError_t launch_reflection_trace(const ReflectionArgs& a, hipStream_t s);
hipError_t launch_reflection_shade(const ReflectionArgs& a, hipStream_t s);
hipError_t launch_init_probes(const SceneDev& sc, const GridConst& g, const float4* dirs512, uint32_t* states,
							  uint32_t* backface_counts, hipStream_t s);
hipError_t launch_select(const uint32_t* states, uint32_t P, uint32_t cap, SelectState* st, uint32_t* indices,
						 hipStream_t s);
hipError_t launch_iota(uint32_t* indices, uint32_t P, SelectState* st, hipStream_t s);
hipError_t launch_set_count(SelectState* st, uint32_t count, hipStream_t s);
hipError_t launch_sample(const GridConst& g, const uint32_t* states, const float4* irr, const float2* dep,
						 const float* pos, const float* nrm, const float* tocam, uint64_t count, float* out,
						 hipStream_t s);
// Weight row k uses direction dirs[order[k]] (order = nullptr: identity), so that the blend can
// consume ray records stored in the traversal's thread order.
hipError_t launch_blend_weights(const GridConst& g, const float4* dirs, const uint16_t* order, float* wt, float* inv_sum,
								void* w_hi16, void* w_lo16, uint32_t mode, hipStream_t s);
// FP32 (mode kBlendFP32): fp32_wide selects the 64-probe workgroups of the 16-bit modes
// (k_blend_wide) over the 16-probe blocks (k_blend_fp32); both give the same sums (the same MFMAs
// per output element, in the same order).  The irradiance-only blend always takes the blocks.
hipError_t launch_blend(const BlendArgs& a, uint32_t mode, hipStream_t s, bool fp32_wide = false);

// blend modes
constexpr uint32_t kBlendFP32 = 0;
constexpr uint32_t kBlendBF16 = 1;
constexpr uint32_t kBlendFP16 = 2;


// ---- batched update of many probe volumes (one graph launch per update) ----
constexpr uint32_t kMaxBatchVolumes = 256;
constexpr uint32_t kMfmaProbes = 64;   // probes per MFMA blend workgroup (BF16 / FP16)

// Everything the kernels of one volume read during a batched update.
struct BatchVolume {
	TraceFrameArgs trace;  // trace.dirs == dirs
	BlendArgs blend;       // blend.wt / inv_sum / w_hi16 / w_lo16 == the buffers below
	Rot R;                 // orientation of this update
	const float4* fib;     // fib(i, n)
	float4* dirs;          // world directions, written by k_batch_begin
	uint16_t* perm;        // ray order of the traversal threads, written by k_batch_begin
	float* wt;
	float* inv_sum;
	void* w_hi;
	void* w_lo;
	uint32_t threads;      // traversal threads: slots * rays
	uint32_t n_pad;        // rays rounded up to the MFMA chunk
	SortJob sort;          // ordering of the shade list (trace.shade_keys) before shading
};

struct BatchItem {
	uint32_t volume;
	uint32_t first;  // first traversal thread (trace items) or first slot (blend items)
};

struct BatchControl {
	uint32_t next[4];        // work-queue heads: [0] primary traversal
	unsigned long long t0;   // wall clock when the update started on the device
	unsigned long long pad;
	unsigned long long done[kMaxBatchVolumes];  // wall clock when a volume's last blend tile ended
};

struct BatchLaunch {
	const BatchVolume* vols;
	const BatchItem* trace_items;
	const BatchItem* blend_items;
	BatchControl* ctl;
	uint32_t volume_count;
	uint32_t trace_item_count;
	uint32_t trace_item_threads;  // traversal threads per item
	uint32_t blend_item_count;
	uint32_t blend_probes;        // slots per blend item
	uint32_t blend_mode;
	const BatchItem* shade_items;  // one workgroup of kShadeBlock shade-list entries each
	uint32_t shade_item_count;
};

hipError_t launch_batch_begin(const BatchLaunch& b, hipStream_t s);
hipError_t launch_batch_weights(const BatchLaunch& b, hipStream_t s);
hipError_t launch_batch_trace(const BatchLaunch& b, uint32_t block, uint32_t grid, hipStream_t s);
// Orders the shade list of every volume (BatchVolume::sort) in one set of launches; max_capacity
// = the largest list capacity.
hipError_t launch_batch_sort(const BatchLaunch& b, uint64_t max_capacity, hipStream_t s);
// Shades every volume's ordered shade list: one workgroup per shade item, one thread per entry.
hipError_t launch_batch_shade(const BatchLaunch& b, hipStream_t s);
hipError_t launch_batch_blend(const BatchLaunch& b, hipStream_t s);
// Resident workgroups per CU of the persistent kernels (for sizing their grids).
hipError_t batch_trace_resident(uint32_t block, int* per_cu);

}  // namespace ddgi
