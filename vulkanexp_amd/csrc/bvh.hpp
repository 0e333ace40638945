// Host-built binned-SAH BVH that replaces the driver's BLAS/TLAS build
// (src/Renderer.cpp:272-449, :525-642).  Instances are flattened into world space at build time,
// so traversal is single-level.
//
// Device layout (all float4 / 16-byte aligned, read with 128-bit loads):
//   node k  : 4 float4 = 64 bytes, holding the boxes of BOTH children, so one node fetch decides
//             which children a ray enters (layout after Aila & Laine, HPG 2009):
//               [0] = (c0.lo.x, c0.hi.x, c0.lo.y, c0.hi.y)
//               [1] = (c1.lo.x, c1.hi.x, c1.lo.y, c1.hi.y)
//               [2] = (c0.lo.z, c0.hi.z, c1.lo.z, c1.hi.z)
//               [3] = (child0, child1, -, -) as int bits; child >= 0: inner node index,
//                     child < 0: leaf = 0x80000000 | first_triangle << 3 | (count - 1)
//   triangle: 3 float4 = 48 bytes, leaf order:
//               [0] = (v0.xyz, global triangle id)   global id = instance-major flattened index
//               [1] = (v1.xyz, instance index)
//               [2] = (v2.xyz, mask | flip_facing << 8)
// Node boxes are inflated by a few ulps so that FP32 slab tests stay conservative.
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/ddgi.h"

namespace ddgi {

constexpr uint32_t kLeafFlag = 0x80000000u;
constexpr uint32_t kMaxLeafSize = 8;

struct BvhHost {
	std::vector<float> nodes;     // 16 floats per node
	std::vector<float> tris;      // 12 floats per triangle (leaf order)
	std::vector<uint32_t> inst_prim_base;  // per instance: global id of its triangle 0
	uint32_t node_count = 0;
	uint32_t tri_count = 0;
	uint32_t max_depth = 0;
	double sah_cost = 0.0;
};

// World-space FP32 vertex of instance row-major 3x4 transform M applied to object position p:
// each row evaluated in double left to right and rounded once to float.
void transform_point(const float* M, const float* p, float out[3]);

// Flatten the scene and build the BVH.  Returns false (with `err`) on invalid input.
bool build_bvh(const ddgi_vertex* vertices, size_t vertex_count, const uint32_t* indices, size_t index_count,
			   const ddgi_offset_entry* offsets, size_t offset_count, const ddgi_instance* instances,
			   size_t instance_count, BvhHost& out, const char** err);

}  // namespace ddgi
