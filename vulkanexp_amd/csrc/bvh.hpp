// Host-built binned-SAH BVH that replaces the driver's BLAS/TLAS build
// (src/Renderer.cpp:272-449, :525-642).  Instances are flattened into world space at build time,
// so traversal is single-level.
//
// Device layout (all float4 / 16-byte aligned, read with 128-bit loads):
//   node k  : kNodeFloats floats, a kWide-wide node collapsed from the binary SAH tree, holding the
//             boxes of its (up to) kWide children side by side so that one node fetch tests all of
//             them, in rows of kWide floats (one float4 per 4 children):
//               row 0 = lo.x of children 0..kWide-1   row 1 = hi.x   row 2 = lo.y
//               row 3 = hi.y                          row 4 = lo.z   row 5 = hi.z
//               row 6 = child codes as int bits: inner node index, or
//                       leaf = 0x80000000 | first_triangle << 3 | (count - 1), or kEmptyChild
//             then padding (4-wide: 128 bytes, 8-wide: 256 bytes)
//   triangle: 3 float4 = 48 bytes, leaf order (a slot; a triangle cut by spatial splits fills one
//             slot per piece, each a full copy of the triangle whose leaf box bounds only its piece):
//               [0] = (v0.xyz, global triangle id)   global id = instance-major flattened index
//               [1] = (v1.xyz, instance index)
//               [2] = (v2.xyz, mask | flip_facing << 8)
// Node boxes are inflated by a few ulps so that FP32 slab tests stay conservative; an empty child
// slot has the code kEmptyChild and a point box far outside any scene.
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/ddgi.h"

namespace ddgi {

constexpr uint32_t kLeafFlag = 0x80000000u;
constexpr uint32_t kEmptyChild = 0xFFFFFFFFu;
constexpr uint32_t kMaxLeafSize = 8;
// children per node: 4 (8-wide nodes visit 26 % fewer nodes per probe ray but trace slower: the
// wider fetch, slab tests and sorting network cost registers and occupancy)
constexpr uint32_t kWide = 4;
constexpr uint32_t kNodeFloats = kWide == 4 ? 32 : 64;  // 6 box rows + 1 code row of kWide, padded
constexpr uint32_t kCodeRow = 6 * kWide;               // float offset of the child codes
constexpr uint32_t kMaxWideDepth = 32;   // traversal stack: kWide - 1 entries per level
static_assert(kWide == 4 || kWide == 8, "nodes are 4- or 8-wide");

struct BvhHost {
	std::vector<float> nodes;     // kNodeFloats per node
	std::vector<float> tris;      // 12 floats per triangle (leaf order)
	std::vector<uint32_t> inst_prim_base;  // per instance: global id of its triangle 0
	uint32_t node_count = 0;
	uint32_t tri_count = 0;
	uint32_t max_depth = 0;       // binary tree
	uint32_t wide_depth = 0;      // levels of the wide tree
	double sah_cost = 0.0;
	// for refits after instance moves: object-space vertices of each triangle slot (9 floats, leaf
	// order) and the wide nodes grouped by depth (level d = nodes level_nodes[level_offsets[d-1] ..
	// level_offsets[d]), root at level 1)
	std::vector<float> obj_tris;
	// per slot the bounds of the part of its triangle that the slot holds (the whole triangle, or
	// the piece a spatial split left there; 6 floats: lo xyz, hi xyz), which a refit keeps for
	// instances that have not moved
	std::vector<float> slot_box;
	std::vector<uint32_t> level_nodes;
	std::vector<uint32_t> level_offsets;
};

// Box padding of the device layout: the float just below v - (|v| 1e-6 + 1e-6) and just above
// v + (|v| 1e-6 + 1e-6), computed in double (the traversal's slab tests stay conservative).
float pad_down(float v);
float pad_up(float v);

// World-space FP32 vertex of instance row-major 3x4 transform M applied to object position p:
// each row evaluated in double left to right and rounded once to float.
void transform_point(const float* M, const float* p, float out[3]);

// Flatten the scene and build the BVH.  Returns false (with `err`) on invalid input.
bool build_bvh(const ddgi_vertex* vertices, size_t vertex_count, const uint32_t* indices, size_t index_count,
			   const ddgi_offset_entry* offsets, size_t offset_count, const ddgi_instance* instances,
			   size_t instance_count, BvhHost& out, const char** err);

}  // namespace ddgi
