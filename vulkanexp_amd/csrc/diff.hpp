// Arguments and launchers of the differentiable probe update (diff.hip).
#pragma once

#include "kernels.hpp"

namespace ddgi {

constexpr uint32_t kDiffChunk = 2048;  // G-buffer texels per backward work item (one cell)

struct DiffArgs {
	GridConst grid;
	// fixed by prepare
	const float4* factors;          // [rays][4], thread (ray_order) order within a slot
	const uint16_t* perm;           // ray order
	const float4* dirs;             // world directions of the prepared frame (ray index order)
	const float* inv_sum;           // blend normalisation per texel (first 36: irradiance)
	const uint32_t* states;         // probe states after the prepared update
	const float2* dep;              // depth atlas after the prepared update
	const uint32_t* indices;        // probe per slot
	uint32_t slots;
	uint64_t rays;                  // slots * n
	const float* pos;               // G-buffer [N][3]
	const float* nrm;
	const float* tocam;
	const uint32_t* gmat;
	uint64_t N;
	uint32_t M;                     // material slots
	const uint32_t* sorted;         // texels sorted by cell
	const uint32_t* cell_item_start;  // [cells + 1]
	const uint32_t* item_begin;
	const uint32_t* item_count;
	const uint32_t* n_items;
	uint32_t max_items;
	// per call
	const float* albedo;            // [M][3]
	const float* sun;               // [3]
	const float* target;            // [N][3]
	float4* rad;                    // [rays]
	const float4* irr;              // differentiated irradiance atlas (tiled)
	float* out;                     // [N][3] sampled irradiance
	float* shade;                   // [N][3]
	float* loss_part;
	float* loss;
	long long* part_tile;           // [max_items][8][64][3]
	long long* part_mat_pix;        // [max_items][M][3]
	float* g_T;                     // [P][36][3]
	long long* part_mat_ray;        // [slots][M * 3 + 3]
	const float* grad_loss;         // [1]
	float* g_albedo;                // [M][3]
	float* g_sun;                   // [3]
};

hipError_t diff_bin(const DiffArgs& d, uint32_t* cell_id, uint32_t* cell_count, uint32_t* cell_start, uint32_t* cursor,
					uint32_t* cell_item_start, uint32_t* item_cell, uint32_t* item_begin, uint32_t* item_count,
					uint32_t* n_items, uint32_t* sorted, hipStream_t s);
hipError_t diff_radiance(const DiffArgs& d, hipStream_t s);
hipError_t diff_gather_forward(const DiffArgs& d, hipStream_t s);
hipError_t diff_backward(const DiffArgs& d, hipStream_t s);

}  // namespace ddgi
