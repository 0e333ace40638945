/*
 * ddgi.h — C-ABI of the MI355X (gfx950) DDGI irradiance-probe library (libddgi.so).
 *
 * The surface mirrors VulkanExp's `class IrradianceProbes`
 * (src/IrradianceProbes.hpp:16-129, src/IrradianceProbes.cpp:12-624): a handle owns the probe
 * grid, its atlases, the probe-state buffer and the scene acceleration structure; `ddgi_update`
 * is one call of `IrradianceProbes::update` (trace -> irradiance/depth blend -> border copy ->
 * publish).  Every entry point returns a `ddgi_status`; nothing throws across the ABI.
 *
 * Conventions
 *  - Pointers documented as "device" are HIP device pointers (hipMalloc / torch CUDA tensors);
 *    "host" pointers are ordinary CPU memory.  Streams are `hipStream_t` passed as `void*`
 *    (NULL = the legacy default stream).
 *  - A handle is externally synchronised: one thread and one stream at a time.  Independent
 *    handles may be used concurrently from different threads on different streams.
 *  - Atlases are kept on the device in a probe-major tiled layout (DDGI_LAYOUT_TILED); the
 *    reference's 2D image layout (DDGI_LAYOUT_IMAGE: width colorRes*Rx*Ry, height colorRes*Rz)
 *    is available on read-back.
 */
#ifndef DDGI_H
#define DDGI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DDGI_ABI_VERSION 1u

typedef enum ddgi_status {
	DDGI_OK = 0,
	DDGI_ERROR_INVALID_ARGUMENT = 1, /* null handle/pointer, bad size, bad grid, NaN/Inf input */
	DDGI_ERROR_OUT_OF_MEMORY = 2,    /* device or host allocation failed */
	DDGI_ERROR_HIP = 3,              /* a HIP runtime call failed; see ddgi_last_error */
	DDGI_ERROR_UNSUPPORTED = 4,      /* valid request the library does not implement */
	DDGI_ERROR_NO_SCENE = 5,         /* update/init/trace before ddgi_set_scene */
	DDGI_ERROR_INVALID_STATE = 6,    /* call order violated (e.g. saved state of another grid) */
	DDGI_ERROR_NO_DEVICE = 7         /* no usable gfx950 device */
} ddgi_status;

/* Grid parameters: 64 bytes, same field order and meaning as the std140 UBO `GridInfo`
 * (src/IrradianceProbes.hpp:49-60, mirrored by src/shaders/ProbeGrid.glsl:4-15). */
typedef struct ddgi_grid_params {
	float extent_min[3];    /* world position of probe (0,0,0) */
	float depth_sharpness;  /* exponent of the depth-blend weight (reference default 12) */
	float extent_max[3];    /* world position of probe (Rx-1,Ry-1,Rz-1) */
	float hysteresis;       /* weight of the previous texel value in the blend */
	int32_t resolution[3];  /* probes per axis, each >= 2 */
	uint32_t rays_per_probe;/* 1 .. ddgi_options.max_rays_per_probe */
	uint32_t color_res;     /* irradiance tile edge incl. 1-texel border; only 8 is supported */
	uint32_t depth_res;     /* depth tile edge incl. 1-texel border; only 16 is supported */
	float shadow_bias;      /* sampleProbes position bias (reference default 0.3) */
	uint32_t padding;
} ddgi_grid_params;

/* 64-byte vertex record (src/vulkan/Vertex.hpp:8-15, unpacked by src/shaders/Vertex.glsl:16-33). */
typedef struct ddgi_vertex {
	float pos[3];
	float color[3];
	float normal[3];
	float tangent[4];
	float tex_coord[2];
	uint32_t padding;
} ddgi_vertex;

/* 48-byte material record (src/vulkan/Material.hpp:16-25, src/shaders/unpackMaterial.glsl:7-19).
 * Textures are not sampled by this library: texture ids are carried but ignored. */
typedef struct ddgi_material {
	float metallic;
	float roughness;
	float base_color[3];
	float emissive[3];
	uint32_t albedo_texture;
	uint32_t normal_texture;
	uint32_t metallic_roughness_texture;
	uint32_t emissive_texture;
} ddgi_material;

/* Mesh offset-table entry (src/Renderer.hpp:21-25; built at src/Renderer.cpp:97-126). */
typedef struct ddgi_offset_entry {
	uint32_t material_index;
	uint32_t vertex_offset; /* in vertices */
	uint32_t index_offset;  /* in indices */
} ddgi_offset_entry;

/* One TLAS instance (src/Renderer.cpp:532-551): 3x4 row-major object-to-world transform, the
 * custom index (= offset-table entry of its mesh), the 8-bit visibility mask, flags and the
 * mesh's triangle count (the BLAS extent, src/Renderer.cpp:323-362). */
typedef struct ddgi_instance {
	float transform[12];
	uint32_t custom_index;
	uint32_t mask;
	uint32_t flags;          /* DDGI_INSTANCE_FLIP_FACING */
	uint32_t triangle_count;
} ddgi_instance;

#define DDGI_INSTANCE_FLIP_FACING 0x1u
#define DDGI_MASK_STATIC 0x01u  /* src/shaders/InstanceMasks.glsl:1-3 */
#define DDGI_MASK_DYNAMIC 0x02u

/* Directional light UBO (src/Light.hpp:6-9): direction points TOWARD the light, color.rgb is
 * the radiance, color.a the sky brightness of the reference's sky model. */
typedef struct ddgi_light {
	float direction[4];
	float color[4];
} ddgi_light;

typedef struct ddgi_options {
	uint32_t max_rays_per_probe; /* capacity of the ray buffers (reference: 256); <= 512; 0 = 512 */
	uint32_t probes_per_update;  /* reference ProbesPerUpdate (src/IrradianceProbes.hpp:44); 0 = no cap */
	float target_hysteresis;     /* reference TargetHysteresis (0.98); used with DDGI_UPDATE_HYSTERESIS_RAMP */
	int32_t device;              /* HIP device ordinal; -1 = current device */
	uint32_t reserved[4];
} ddgi_options;

/* ddgi_update flags */
#define DDGI_UPDATE_ALL_PROBES 0x1u       /* trace+blend every probe of the grid (indices 0..P-1) */
#define DDGI_UPDATE_BLEND_BF16 0x2u       /* blend on BF16 MFMA (hi/lo compensated), FP32 accumulation */
#define DDGI_UPDATE_BLEND_FP16 0x4u       /* blend on FP16 MFMA, FP32 accumulation */
#define DDGI_UPDATE_HYSTERESIS_RAMP 0x8u  /* reference hysteresis ramp (src/IrradianceProbes.cpp:462-476) */
#define DDGI_UPDATE_NO_TIMING 0x10u       /* skip per-phase HIP events */
#define DDGI_UPDATE_COUNT_VISITS 0x20u    /* count BVH node/triangle fetches (slower; statistics) */

/* Atlas kinds / layouts */
#define DDGI_ATLAS_IRRADIANCE 0u /* float4 texels (rgb, 1) */
#define DDGI_ATLAS_DEPTH 1u      /* float2 texels (mean, mean of squares) */
#define DDGI_LAYOUT_TILED 0u     /* [probe][row][col] texels, probe tiles of res x res */
#define DDGI_LAYOUT_IMAGE 1u     /* reference 2D image, row-major, width res*Rx*Ry */

/* Ray query record for ddgi_trace_rays (a device array). */
typedef struct ddgi_ray {
	float origin[3];
	float tmin;
	float direction[3];
	float tmax;
} ddgi_ray;

/* Hit record written by ddgi_trace_rays: instance = TLAS instance index, primitive = triangle
 * index within the instance's mesh (gl_PrimitiveID), u/v = barycentric weights of vertices 1
 * and 2 (the closest-hit `attribs`), flags bit0 = back-facing.  Misses: instance = 0xFFFFFFFF,
 * t = -1.  For any-hit queries only `instance != 0xFFFFFFFF` is meaningful. */
typedef struct ddgi_hit {
	float t;
	float u;
	float v;
	uint32_t instance;
	uint32_t primitive;
	uint32_t flags;
} ddgi_hit;

#define DDGI_TRACE_ANY_HIT 0x1u        /* terminate on first hit (shadow-ray semantics) */
#define DDGI_TRACE_CULL_BACK_FACES 0x2u /* ignore back-facing triangles (rasterizer back-face culling) */

/* G-buffer texel (the reference's four attachments, GBuffer.frag:63-66): 64 bytes. */
typedef struct ddgi_gbuffer_texel {
	float position[3];  /* world position of the nearest front-facing hit */
	float depth;        /* its distance from the ray origin; 0 on a miss */
	float normal[3];    /* normalize(normal matrix * interpolated vertex normal); on a miss the ray
	                       direction (the gather's sky lookup) */
	float metalness;
	float albedo[3];    /* interpolated vertex colour * base colour */
	float roughness;
	float emissive[3];
	uint32_t material;  /* material index, 0xFFFFFFFF on a miss */
} ddgi_gbuffer_texel;

/* Per-phase GPU times of the last completed update, in milliseconds (the reference's timestamp
 * split, src/IrradianceProbes.cpp:441-452). */
typedef struct ddgi_timings {
	float select_ms;
	float trace_ms;   /* primary traversal + shading + shadow rays */
	float blend_ms;   /* irradiance + depth blend with fused border copy */
	float border_ms;  /* always 0: the border copy is fused into the blend epilogue */
	float copy_ms;    /* always 0: the work->exposed copy is a pointer swap */
	float total_ms;
	uint32_t probes_updated;
	uint32_t rays_traced;
	float primary_ms; /* closest-hit traversal of the probe rays (part of trace_ms) */
	float shade_ms;   /* closest-hit shading + shadow rays (part of trace_ms) */
} ddgi_timings;

/* Traversal statistics of the last update run with DDGI_UPDATE_COUNT_VISITS:
 * [0] primary node fetches, [1] primary triangle fetches, [2] shadow node fetches,
 * [3] shadow triangle fetches, [4] shadow rays.  Node = 128 bytes (4-wide), triangle = 48 bytes. */
#define DDGI_TRACE_STATS 5

/* Raw device view of a handle's buffers for zero-copy interop (torch). Valid until the next
 * ddgi_update / ddgi_set_scene / ddgi_destroy on the handle. */
typedef struct ddgi_device_view {
	void* irradiance;        /* exposed irradiance atlas, tiled float4 */
	void* depth;             /* exposed depth atlas, tiled float2 */
	void* states;            /* uint32[P] probe states */
	void* rays;              /* float4[P*max_rays] (rgb, depth) of the last update, slot-major, entry
	                            k of a slot holds ray ray_order[k] */
	void* hits;              /* internal hit records of the last update, same order as rays */
	void* ray_order;         /* uint16[max_rays]: ray index stored at each position of a slot (set by
	                            each update: its rays sorted by the world elevation of their direction) */
	void* probe_indices;     /* uint32[P] probes traced by the last update */
	void* ray_directions;    /* float4[max_rays] world directions of the last update */
	uint64_t probe_count;
	uint32_t max_rays_per_probe;
	uint32_t last_update_count; /* probes updated by the last update (host copy; -1 if unknown) */
} ddgi_device_view;

/* ---- lifetime (IrradianceProbes::init / destroy, src/IrradianceProbes.cpp:12-104,596-624) ---- */
ddgi_status ddgi_create(const ddgi_grid_params* params, const ddgi_options* options, void** out_handle);
ddgi_status ddgi_destroy(void* handle);
uint32_t ddgi_abi_version(void);
const char* ddgi_status_string(ddgi_status status);
const char* ddgi_last_error(void* handle);

/* ---- configuration (updateUniforms / setLightBuffer / writeDescriptorSet) ----
 * ddgi_set_params: IrradianceProbes::updateUniforms (src/IrradianceProbes.cpp:106-108); the
 *   resolution must equal the one given at creation (the reference re-inits to resize).
 * ddgi_set_light: the light UBO of binding 10 (src/IrradianceProbes.cpp:327-336).
 *   color[3] scales the sun colour in the atmospheric sky (src/shaders/miss.rmiss:22).
 * ddgi_set_sky: constant radiance returned by misses under DDGI_SKY_CONSTANT.
 * ddgi_set_sky_model: what misses return (src/shaders/miss.rmiss:17-23): DDGI_SKY_CONSTANT (the
 *   default) or DDGI_SKY_ATMOSPHERE, the reference's sky() (src/shaders/sky.glsl:60-127) seen from
 *   the ray origin with the light's direction and colour * color[3], sun disc included; the
 *   gather's camera misses use colour * 1 (src/shaders/FinalGather.frag:52-54).  Batched updates
 *   (ddgi_batch_update) support the constant sky only. */
#define DDGI_SKY_CONSTANT 0u
#define DDGI_SKY_ATMOSPHERE 1u
ddgi_status ddgi_set_params(void* handle, const ddgi_grid_params* params);
ddgi_status ddgi_get_params(void* handle, ddgi_grid_params* out_params);
ddgi_status ddgi_set_light(void* handle, const ddgi_light* light);
ddgi_status ddgi_set_sky(void* handle, const float rgb[3]);
ddgi_status ddgi_set_sky_model(void* handle, uint32_t model);

/* ---- scene (Renderer::allocateMeshes + createAccelerationStructures + createTLAS,
 * src/Renderer.cpp:57-95,272-449,525-642).  All arrays are host memory and are copied.
 * Builds the binned-SAH BVH on the host and uploads it. */
ddgi_status ddgi_set_scene(void* handle, const ddgi_vertex* vertices, size_t vertex_count, const uint32_t* indices,
						   size_t index_count, const ddgi_offset_entry* offsets, size_t offset_count,
						   const ddgi_instance* instances, size_t instance_count, const ddgi_material* materials,
						   size_t material_count);
/* Moves instances [first, first + count) of the current scene to new row-major 3x4 transforms
 * (12 floats each, host memory, copied), as Renderer::updateAccelerationStructureInstances +
 * updateTLAS (src/Renderer.cpp:671-742: new instance transforms, then an acceleration-structure
 * update in place).  The moved instances' triangles are re-transformed on the device exactly as
 * ddgi_set_scene would flatten them, and the BVH boxes are refitted bottom-up (the tree's topology
 * is kept, as a TLAS update keeps it), so every query afterwards returns the hits a scene built
 * from the new transforms returns.  Masks, offset entries and mesh data are kept.  Asynchronous on
 * `stream`; the transform array may be reused when the call returns.  The probe states keep their
 * values: ddgi_init_probes re-classifies them for the moved geometry (the "Update State" action,
 * src/UI.cpp:991-993). */
ddgi_status ddgi_update_instances(void* handle, uint32_t first, uint32_t count, const float* transforms, void* stream);
/* Node count, triangle slots and bytes of the uploaded BVH (the builder's spatial splits give a
 * triangle cut by a split plane one slot per piece, so the slots can outnumber the triangles). */
ddgi_status ddgi_get_bvh_info(void* handle, uint64_t* node_count, uint64_t* triangle_count, uint64_t* bytes);
/* For every BVH triangle slot (count = triangle_count of ddgi_get_bvh_info; host arrays): the TLAS
 * instance and the primitive index within its mesh (gl_InstanceID / gl_PrimitiveID of a hit on it,
 * src/shaders/closesthit.glsl:147-150).  The internal hit records of an update (the `hits` of
 * ddgi_device_view: t, slot | back-face << 31, u, v) name the slot; this maps it to the ids a
 * ddgi_trace_rays record carries.  Synchronous. */
ddgi_status ddgi_get_slot_ids(void* handle, uint32_t* instance, uint32_t* primitive, uint64_t count);

/* ---- scene ingestion (no counterpart in IrradianceProbes: the reference's Scene::load / save,
 * src/Scene.cpp:30-49 (.scene, .gltf, .glb, .obj), :115-533 (glTF 2.0), :534-646 (OBJ), :647-684 (.mat), :710-934
 * (.scene), flattened as Renderer::updateMeshOffsetTable + createTLAS build their tables,
 * src/Renderer.cpp:97-126, :512-551).  Host-only; a scene object owns the tables it returns. ----
 * The flattening: one offset entry per mesh (= glTF primitive) with indices, meshes packed back to
 *   back; one instance per node that renders a mesh, ordered by (material, mesh) and then node
 *   order, with the node's global transform, mask DDGI_MASK_STATIC and no facing flip; a default
 *   material (metallic 1, roughness 1, white) is appended when a mesh refers to a missing one.
 *   Skinned primitives are loaded but not instanced.  Textures are kept as records only.
 * ddgi_scene_load: .scene, .gltf (external or base64 buffers), .glb or .obj (Scene::loadOBJ,
 *   src/Scene.cpp:534-646: triangles, groups as sub-meshes, vertex normals recomputed as the
 *   reference does).
 * ddgi_scene_load_material: adds the materials (and texture records) of a .mat file.
 * ddgi_scene_create: a scene from flattened tables (static instances only), e.g. to save one.
 * ddgi_scene_save: writes the .scene container (header "SCEN", a JSON chunk, then per mesh a
 *   vertex chunk and an index chunk, README.md:75-91). */
typedef struct ddgi_scene_tables {
	const ddgi_vertex* vertices;
	uint64_t vertex_count;
	const uint32_t* indices;
	uint64_t index_count;
	const ddgi_offset_entry* offsets;
	uint64_t offset_count;
	const ddgi_instance* instances;
	uint64_t instance_count;
	const ddgi_material* materials;
	uint64_t material_count;
	float bounds_min[3]; /* Scene::computeBounds (src/Scene.cpp:1077-1094) */
	float bounds_max[3];
	uint64_t mesh_count;
	uint64_t node_count;
	uint64_t texture_count;
} ddgi_scene_tables;
ddgi_status ddgi_scene_load(const char* path, void** out_scene);
ddgi_status ddgi_scene_load_material(void* scene, const char* path);
ddgi_status ddgi_scene_create(const ddgi_vertex* vertices, size_t vertex_count, const uint32_t* indices,
							  size_t index_count, const ddgi_offset_entry* offsets, size_t offset_count,
							  const ddgi_instance* instances, size_t instance_count, const ddgi_material* materials,
							  size_t material_count, void** out_scene);
ddgi_status ddgi_scene_get_tables(void* scene, ddgi_scene_tables* out);
ddgi_status ddgi_scene_save(void* scene, const char* path);
ddgi_status ddgi_scene_destroy(void* scene);

/* ---- per-frame work ----
 * ddgi_init_probes: IrradianceProbes::initProbes (src/IrradianceProbes.cpp:357-394) with the
 *   shader src/shaders/probesInit.rgen:31-64; orientation = row-major 3x3 R (rows X,Y,Z of
 *   genBasis, src/IrradianceProbes.cpp:347-367), ray i = R * sphericalFibonacci(i, n).
 * ddgi_update: IrradianceProbes::update (src/IrradianceProbes.cpp:426-594).  Asynchronous on
 *   `stream`; selection of the probes happens on the device. */
ddgi_status ddgi_init_probes(void* handle, const float orientation[9], void* stream);
ddgi_status ddgi_update(void* handle, const float orientation[9], uint32_t flags, void* stream);

/* Blend only: feed an externally produced ray buffer `rays` (device float4[count*n], slot-major
 * (rgb, depth) in ray-index order, n = rays_per_probe) for probes `probe_indices` (device uint32[count]) and the
 * frame orientation, then run blend + border + state update + publish like ddgi_update does.
 * Probes not listed keep their tiles: the next atlas starts as a copy of the exposed one.  With
 * DDGI_UPDATE_ALL_PROBES in `flags` and count = the probe count, the caller states that
 * probe_indices lists every probe once, and that copy is skipped. */
ddgi_status ddgi_blend_rays(void* handle, const void* rays, const void* probe_indices, uint32_t count,
							const float orientation[9], uint32_t flags, void* stream);

/* Trace arbitrary rays against the scene BVH (device arrays).  `mask` is the ray cull mask. */
ddgi_status ddgi_trace_rays(void* handle, const ddgi_ray* rays, ddgi_hit* hits, uint64_t count, uint32_t mask,
							uint32_t flags, void* stream);

/* sampleProbes (src/shaders/irradiance.glsl:145-237) at `count` points: device float3 arrays
 * (packed xyz) position, normal, to_camera; writes device float3 rgb. */
/* ---- full-frame consumer: G-buffer + final gather (no counterpart in IrradianceProbes; the
 * reference's GBufferPipeline / GatherPipeline, src/RenderPasses/GBufferPipeline.cpp,
 * src/RenderPasses/GatherPipeline.cpp:3-383, src/shaders/FinalGather.frag:38-77) ---- */
/* G-buffer of `count` camera rays (device ddgi_ray array) into device texels: the nearest
 * front-facing hit (back faces culled like GBufferPipeline.cpp:212), mask 0xFF. */
ddgi_status ddgi_gbuffer(void* handle, const ddgi_ray* rays, uint64_t count, ddgi_gbuffer_texel* out, void* stream);
/* Final gather of device G-buffer texels into device rgb (count * 3 floats), with the handle's
 * exposed atlases, light and sky.  view_origin (host, 3 floats): the camera position.  direct
 * (device, count floats, may be NULL): sun visibility per texel; NULL = a hard shadow ray toward
 * the sun.  reflection (device, count * 3 floats, may be NULL = black): the reflection pass. */
ddgi_status ddgi_gather(void* handle, const ddgi_gbuffer_texel* gbuffer, uint64_t count, const float view_origin[3],
						const float* direct, const float* reflection, float* out_rgb, void* stream);
ddgi_status ddgi_sample(void* handle, const float* position, const float* normal, const float* to_camera,
						uint64_t count, float* rgb_out, void* stream);
/* Ray-traced shadow pass (src/shaders/directLight.rgen:42-99, DirectLightPipeline.cpp:5-247) over
 * device G-buffer texels into device float4 records (the reference's rgba32f directLightOut):
 * (-1,-1,-1,-1) on a miss texel; otherwise the sun direction normalize(light.direction), jittered
 * inside a 0.02 rad cone by the texel's noise sample (device float2 per texel in [0,1)^2, the
 * blue-noise texel the shader reads; NULL = no jitter), gives (0,0,0,0) when it faces away from the
 * normal, else an any-hit ray (mask 0xFF, tmin 0.01, tmax 10000) gives (visibility, previous.g,
 * previous.b, 1) with visibility 1 on escape (faded below the horizon) or 0.  previous: device
 * float4 per texel (previousDirectLight), NULL = zeros.  The record's first channel is what
 * ddgi_gather's `direct` takes (after the reference's filter passes, which are not part of this
 * library). */
ddgi_status ddgi_direct_light(void* handle, const ddgi_gbuffer_texel* gbuffer, uint64_t count, const float* noise,
							  const float* previous, float* out_rgba, void* stream);
/* Reflection pass (src/shaders/reflection.rgen:101-175) over device G-buffer texels into device
 * float4 records (rgb, roughness): texels with depth > 0 and (roughness < 0.4 or metalness > 0.01)
 * trace one ray (closest hit, mask 0xFF, tmin 0.1, tmax 10000) from the texel along
 * reflect(-normalize(view_origin - position), normal), rotated within the roughness lobe by the
 * texel's noise sample (device float2 per texel, NULL = the mirror direction); the hit is shaded as
 * closesthit.glsl at recursion depth 1 (specular from sampleProbes along the reflected ray), a back
 * face is black, a miss sees the sky model; rgb = colorCompression(radiance).  Other texels: 0.
 * The rgb part is what ddgi_gather's `reflection` takes.  The handle keeps device scratch of 32
 * bytes per texel (of at most 4 Mi texels at a time), allocated by the first call that needs more
 * than it holds. */
ddgi_status ddgi_reflection(void* handle, const ddgi_gbuffer_texel* gbuffer, uint64_t count, const float view_origin[3],
							const float* noise, float* out_rgba, void* stream);
/* ddgi_reflection without any device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and results as ddgi_reflection, for at most
 * 2,097,152 texels per call (a 1920 x 1080 view fits; larger counts are refused with
 * DDGI_ERROR_INVALID_ARGUMENT: split the G-buffer into calls).  Its scratch is one 64.5 MiB array
 * per GPU, part of the library's code object and held until the process ends; every call on a GPU
 * shares that GPU's array, so calls that overlap on one GPU (other streams, other handles, other
 * threads) race and must be serialised by the caller. */
ddgi_status ddgi_reflection_noalloc(void* handle, const ddgi_gbuffer_texel* gbuffer, uint64_t count,
									const float view_origin[3], const float* noise, float* out_rgba, void* stream);

/* ---- differentiable probes (no counterpart in the reference, which is not differentiable) ----
 * Gradients of a shaded-G-buffer loss with respect to per-material albedo and the sun radiance,
 * through one probe update (ray radiance -> irradiance blend + border copy) and probe sampling
 * (sampleProbes) at G-buffer texels, shaded as irradiance * albedo * (1 - 0.004) like the
 * reference's gather (src/shaders/FinalGather.frag:62-72).  loss = mean over texels and channels
 * of (shade - target)^2.
 * ddgi_diff_prepare: runs one update of every probe with `orientation` (like ddgi_update with
 *   DDGI_UPDATE_ALL_PROBES) and keeps what the parameters do not change: ray hits and shading
 *   factors, the irradiance atlas before the update, the depth atlas and the probe states after
 *   it, and the G-buffer (device arrays: position/normal/to_camera float[3*count], material
 *   uint32[count], copied).  Requires non-metallic materials (radiance is then affine in albedo
 *   and sun); material_count (<= 1024) albedo slots, every material id < material_count.
 * ddgi_diff_forward: device albedo float[3*material_count], sun float[3], target float[3*count]
 *   -> shade_out float[3*count], loss_out float[1].
 * ddgi_diff_backward: differentiates the most recent ddgi_diff_forward of the handle (same albedo,
 *   sun, target); grad_loss float[1] is d(total)/d(loss); writes grad_albedo float[3*material_count]
 *   and grad_sun float[3].  Deterministic (fixed-point reductions). */
ddgi_status ddgi_diff_prepare(void* handle, const float orientation[9], const float* gbuf_position,
							  const float* gbuf_normal, const float* gbuf_to_camera, const uint32_t* gbuf_material,
							  uint64_t texel_count, uint32_t material_count, void* stream);
ddgi_status ddgi_diff_forward(void* handle, const float* albedo, const float* sun, const float* target, float* shade_out,
							  float* loss_out, void* stream);
ddgi_status ddgi_diff_backward(void* handle, const float* albedo, const float* sun, const float* target,
							   const float* grad_loss, float* grad_albedo, float* grad_sun, void* stream);

/* ---- state access ---- */
ddgi_status ddgi_get_atlas(void* handle, uint32_t kind, uint32_t layout, void* host_dst, size_t bytes);
ddgi_status ddgi_set_atlas(void* handle, uint32_t kind, uint32_t layout, const void* host_src, size_t bytes);
ddgi_status ddgi_get_states(void* handle, uint32_t* host_dst, size_t count);
ddgi_status ddgi_set_states(void* handle, const uint32_t* host_src, size_t count);
ddgi_status ddgi_get_timings(void* handle, ddgi_timings* out);
ddgi_status ddgi_get_trace_stats(void* handle, uint64_t out[DDGI_TRACE_STATS]);
ddgi_status ddgi_get_device_view(void* handle, ddgi_device_view* out);
/* Save/restore atlases + states + schedule counters + hysteresis.  Call with dst NULL to query
 * the size. */
ddgi_status ddgi_save_state(void* handle, void* host_dst, size_t* bytes);
ddgi_status ddgi_load_state(void* handle, const void* host_src, size_t bytes);
/* Block the host until all work queued on the handle's last stream is done. */
ddgi_status ddgi_synchronize(void* handle);

/* ---- many probe volumes per update ----
 * No counterpart in the reference, which updates one IrradianceProbes per scene
 * (src/IrradianceProbes.cpp:426-594).  A batch groups handles (same device, each with a scene, no
 * handle twice) and updates all of them with ONE hipGraph launch: frame directions and blend
 * weights of every volume, a persistent work-queue kernel for the primary traversal over the work
 * items of all volumes, the shade-list ordering, then the shading + shadow rays (one thread per
 * list entry) and the blend (one workgroup per probe tile) over all volumes.  The atlases,
 * states and ray records are bit-identical to ddgi_update(h, R_h, DDGI_UPDATE_ALL_PROBES | mode)
 * on each handle.  The handles stay usable on their own; ddgi_destroy is refused while a batch
 * holds the handle.  Errors of the batch calls are reported by ddgi_last_error(NULL). */
typedef struct ddgi_batch_config {
	uint32_t trace_block;       /* primary traversal workgroup size: 64, 128 or 256 threads */
	uint32_t trace_item_blocks; /* work item = trace_block * trace_item_blocks consecutive rays of one volume (1..64) */
	uint32_t blend_probes;      /* FP32 blend tile: probes per workgroup, 4, 8, 16, 32 or 64 (64: the
	                               512-thread workgroups of the 16-bit blends, which always use 64) */
	uint32_t reserved[5];
} ddgi_batch_config;

typedef struct ddgi_batch_tune_record {
	ddgi_batch_config config;
	float ms;                   /* mean time of one batched update with this configuration */
	uint32_t phase;             /* 0: traversal candidates, 1: blend candidates */
} ddgi_batch_tune_record;

/* config NULL = defaults (128, 4, 16). */
ddgi_status ddgi_batch_create(void* const* handles, uint32_t count, const ddgi_batch_config* config, void** out_batch);
ddgi_status ddgi_batch_destroy(void* batch);
ddgi_status ddgi_batch_set_config(void* batch, const ddgi_batch_config* config);
ddgi_status ddgi_batch_get_config(void* batch, ddgi_batch_config* out);
/* orientations: host, count * 9 floats (row-major 3x3 per volume, in handle order).  flags must
 * contain DDGI_UPDATE_ALL_PROBES; DDGI_UPDATE_BLEND_BF16 / _FP16 / _HYSTERESIS_RAMP apply as in
 * ddgi_update; DDGI_UPDATE_COUNT_VISITS is not supported. */
ddgi_status ddgi_batch_update(void* batch, const float* orientations, uint32_t flags, void* stream);
/* Times of the last batched update: total (HIP events around the graph launch) and, per volume,
 * the time from the start of the update on the device until the volume's last blend tile ended
 * (device wall clock).  Waits for the update. volume_ms may be NULL. */
ddgi_status ddgi_batch_get_timings(void* batch, float* total_ms, float* volume_ms, uint32_t count);
/* Project autotuner: times traversal configurations (workgroup size x work-item size), then FP32
 * blend tiles, each over `iterations` real batched updates, keeps the fastest, and restores the
 * volumes' atlases, states and hysteresis afterwards.  records (may be NULL) receives up to
 * max_records measurements; *record_count the number written. */
ddgi_status ddgi_batch_autotune(void* batch, const float* orientations, uint32_t flags, uint32_t iterations, void* stream,
								ddgi_batch_config* best, ddgi_batch_tune_record* records, uint32_t max_records,
								uint32_t* record_count);

#ifdef __cplusplus
}
#endif

#endif /* DDGI_H */
