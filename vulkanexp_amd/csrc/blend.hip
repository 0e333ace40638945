// Irradiance + depth blend with hysteresis (src/shaders/probesUpdate.glsl:49-119), border copy
// (src/shaders/probesCopyBorders.comp:21-246) and the probe state machine, on gfx950.
//
// The reference recomputes max(0, dot(texelDir, rayDir)) (and its pow) per probe, although the
// weights depend only on the frame's rotation.  Here one small kernel computes, once per frame, the
// weight matrix in 256 padded rows of n_pad rays (irradiance texels at rows 0..35, depth texels at
// rows 48..243, the other rows and the rays beyond n zero)
//   W[t][i] = max(0, texelDir_t . rayDir_i)               irradiance texel t
//   W[48 + t][i] = max(0, texelDir_t . rayDir_i)^sharpness  depth texel t
//   inv_sum[t] = 1 / sum_i W (1 when the sum <= 1e-3, i.e. no normalisation)
// and the blend of a probe is a [232 x n] x [n x 3] product, on the matrix cores.  The border copy is fused into the
// epilogue: each interior texel also writes the border texels that mirror it, so the separate
// border pass and its re-read of the atlas disappear.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include "device_common.hpp"

namespace ddgi {

namespace {

constexpr int kProbesPerBlock = 16;

// Column of ray i in the 16-bit weight rows, which only the wide MFMA blend reads: inside each
// block of 32 rays, ray 4 j + g is stored at 8 g + j, so that the 8 consecutive weights a lane
// of K block g takes are the rays 4 j + g that its A fragment holds (loaded as 4 consecutive rays
// of each of 16 probes per instruction).
__device__ __forceinline__ uint32_t wide_column(uint32_t i) { return (i & ~31u) | ((i & 3u) << 3) | ((i & 31u) >> 2); }

// One block per padded weight row: the weight of every ray and their sum, stored in [row][n_pad]
// layout as FP32, or as 16-bit copies (BF16: hi/lo split; FP16: one copy) in wide_column order.
__device__ __forceinline__ void blend_weight_row(const GridConst& g, const float4* __restrict__ dirs,
												 const uint16_t* __restrict__ order, float* __restrict__ wt,
												 float* __restrict__ inv_sum, void* __restrict__ w_hi_v,
												 void* __restrict__ w_lo_v, uint32_t mode, uint32_t n_pad, uint32_t row) {
	__shared__ float red[256];
	const bool irr_row = row < kIrrTexels;
	const bool dep_row = row >= 48u && row < 48u + kDepTexels;
	const bool valid = irr_row || dep_row;
	const uint32_t t = irr_row ? row : row - 48u + kIrrTexels;  // texel index 0..231 when valid
	float tx = 0.0f, ty = 0.0f, tz = 0.0f;
	if(valid) texel_direction(t, tx, ty, tz);
	__bf16* w_hi = static_cast<__bf16*>(w_hi_v);
	__bf16* w_lo = static_cast<__bf16*>(w_lo_v);
	_Float16* w_h = static_cast<_Float16*>(w_hi_v);
	float sum = 0.0f;
	for(uint32_t i = threadIdx.x; i < n_pad; i += 256) {
		float w = 0.0f;
		if(valid && i < g.n) {
			const float4 d = dirs[order ? order[i] : i];
			w = fmaxf(0.0f, tx * d.x + ty * d.y + tz * d.z);
			if(dep_row) w = powf(w, g.sharpness);
		}
		if(mode == kBlendFP32) {
			wt[size_t(row) * n_pad + i] = w;
		} else if(mode == kBlendBF16) {
			// compensated split: w = hi + lo, both bf16; normalise by the sum of what the MFMA
			// multiplies, so the rounding of the weights cancels in the normalisation
			const __bf16 hi = __bf16(w);
			const __bf16 lo = __bf16(w - float(hi));
			w_hi[size_t(row) * n_pad + wide_column(i)] = hi;
			w_lo[size_t(row) * n_pad + wide_column(i)] = lo;
			w = float(hi) + float(lo);
		} else if(mode == kBlendFP16) {
			const _Float16 h = _Float16(w);
			w_h[size_t(row) * n_pad + wide_column(i)] = h;
			w = float(h);
		}
		sum += w;
	}
	red[threadIdx.x] = sum;
	__syncthreads();
	for(int o = 128; o > 0; o >>= 1) {
		if(int(threadIdx.x) < o) red[threadIdx.x] += red[threadIdx.x + o];
		__syncthreads();
	}
	if(threadIdx.x == 0 && valid) inv_sum[t] = red[0] > 1e-3f ? 1.0f / red[0] : 1.0f;
}

__global__ __launch_bounds__(256) void k_blend_weights(GridConst g, const float4* __restrict__ dirs,
													   const uint16_t* __restrict__ order, float* __restrict__ wt,
													   float* __restrict__ inv_sum, void* __restrict__ w_hi_v,
													   void* __restrict__ w_lo_v, uint32_t mode, uint32_t n_pad) {
	blend_weight_row(g, dirs, order, wt, inv_sum, w_hi_v, w_lo_v, mode, n_pad, blockIdx.x);
}

// The interior texel whose blended value texel (x, y) of an R x R tile holds after the border copy:
// itself inside; on the border, the octahedral mirror that probesCopyBorders.comp's tables encode
// (an edge texel copies the interior texel next to it, reflected along the edge; a corner copies
// the opposite interior corner).
template <int R>
__device__ __forceinline__ void border_source(int x, int y, int& sx, int& sy) {
	const bool bx = x == 0 || x == R - 1, by = y == 0 || y == R - 1;
	if(bx && by) {
		sx = x == 0 ? R - 2 : 1;
		sy = y == 0 ? R - 2 : 1;
	} else if(bx) {
		sx = x == 0 ? 1 : R - 2;
		sy = R - 1 - y;
	} else if(by) {
		sx = R - 1 - x;
		sy = y == 0 ? 1 : R - 2;
	} else {
		sx = x;
		sy = y;
	}
}

// Accumulator layout in LDS between the K loop and the epilogue, per probe slot of the block:
// irradiance sums [texel][4] (rgb + pad) and depth sums [texel][2], each probe's block padded by
// 4 floats so that consecutive probes start on different banks.
constexpr int kIrrStride = 4 * int(kIrrTexels) + 4;
constexpr int kDepStride = 2 * int(kDepTexels) + 4;

// Accumulator element (texel row, data column) of a blend GEMM into the LDS layout above: irradiance
// columns are 3 probe + channel, depth columns 2 probe + channel; padding rows and columns are
// dropped.
__device__ __forceinline__ void store_irr_sum(float* s_irr, int tex, int col, int np_max, float v) {
	const int p = col / 3, ch = col - 3 * p;
	if(tex < int(kIrrTexels) && p < np_max) s_irr[p * kIrrStride + tex * 4 + ch] = v;
}
__device__ __forceinline__ void store_dep_sum(float* s_dep, int tex, int col, int np_max, float v) {
	const int p = col >> 1, ch = col & 1;
	if(tex < int(kDepTexels) && p < np_max) s_dep[p * kDepStride + tex * 2 + ch] = v;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
	for(int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
	return v;
}

// Epilogue shared by the blends, for the block's np probe slots (atlas indices in s_idx): whole
// tiles, borders included.  Each texel takes the interior texel it holds after the border copy
// (border_source), normalises that texel's sum, applies hysteresis against that texel's exposed
// value and stores into the next atlas; a wavefront writes one tile row-block at a time, so the
// stores are contiguous.  Irradiance: one wavefront per 8x8 tile (lane = texel), maxChange per
// probe over the interior texels by a wave reduction.  Depth: a wavefront per 16x16 tile, four
// texels per lane.  A pass takes kTilesPerPass tiles per wavefront (tiles p0 + waves u); its
// exposed-atlas loads, irradiance and depth together, depend on the probe indices only, so they
// are issued before the stores (epilogue_load), and a kernel may issue them before its sums are
// ready.

// The lane's texels: the interior texel its irradiance texel and its four depth texels hold after
// the border copy, and those texels' weight normalisers.
struct EpiLane {
	uint32_t irr_src, irr_stex, dep_src[4], dep_stex[4];
	float irr_inv, dep_inv[4];
	bool interior;
};

__device__ __forceinline__ EpiLane epilogue_lane(const BlendArgs& a) {
	const uint32_t lane = threadIdx.x & 63u;
	EpiLane L;
	int sx, sy;
	border_source<8>(int(lane & 7u), int(lane >> 3), sx, sy);
	L.interior = sx == int(lane & 7u) && sy == int(lane >> 3);
	L.irr_stex = uint32_t(sy - 1) * 6u + uint32_t(sx - 1);
	L.irr_src = uint32_t(sy) * 8u + uint32_t(sx);
	L.irr_inv = a.inv_sum[L.irr_stex];
#pragma unroll
	for(uint32_t i = 0; i < 4; ++i) {
		border_source<16>(int(lane & 15u), int(lane >> 4) + 4 * int(i), sx, sy);
		L.dep_stex[i] = uint32_t(sy - 1) * 14u + uint32_t(sx - 1);
		L.dep_src[i] = uint32_t(sy) * 16u + uint32_t(sx);
		L.dep_inv[i] = a.inv_sum[kIrrTexels + L.dep_stex[i]];
	}
	return L;
}

// exposed-atlas values of one pass
template <uint32_t kTilesPerPass>
struct EpiPrev {
	float4 irr[kTilesPerPass];
	float2 dep[kTilesPerPass][4];
};

template <uint32_t kTilesPerPass>
__device__ __forceinline__ void epilogue_load(const BlendArgs& a, const EpiLane& L, const uint32_t* s_idx, uint32_t np,
											  uint32_t p0, bool irr_only, EpiPrev<kTilesPerPass>& e) {
	const uint32_t waves = blockDim.x >> 6;
#pragma unroll
	for(uint32_t u = 0; u < kTilesPerPass; ++u) {
		const uint32_t p = p0 + waves * u;
		if(p < np) {
			const size_t probe = s_idx[p];
			e.irr[u] = a.irr_prev[probe * 64 + L.irr_src];
			if(!irr_only) {
#pragma unroll
				for(uint32_t i = 0; i < 4; ++i) e.dep[u][i] = a.dep_prev[probe * 256 + L.dep_src[i]];
			}
		}
	}
}

template <uint32_t kTilesPerPass>
__device__ __forceinline__ void epilogue_store(const BlendArgs& a, const EpiLane& L, const uint32_t* s_idx, uint32_t np,
											   uint32_t p0, const float* s_irr, const float* s_dep, uint32_t* s_maxc,
											   bool irr_only, const EpiPrev<kTilesPerPass>& e) {
	const float h = a.grid.hysteresis;
	float4* __restrict__ irr_out = a.irr_out;
	float2* __restrict__ dep_out = a.dep_out;
	const uint32_t lane = threadIdx.x & 63u, waves = blockDim.x >> 6;
#pragma unroll
	for(uint32_t u = 0; u < kTilesPerPass; ++u) {
		const uint32_t p = p0 + waves * u;
		if(p >= np) break;
		const size_t probe = s_idx[p];
		const float4 sum = *reinterpret_cast<const float4*>(s_irr + p * kIrrStride + L.irr_stex * 4);
		const float r0 = sum.x * L.irr_inv, r1 = sum.y * L.irr_inv, r2 = sum.z * L.irr_inv;
		const float4 pv = e.irr[u];
		const float mc = L.interior ? fmaxf(fmaxf(fabsf(r0 - pv.x), fabsf(r1 - pv.y)), fabsf(r2 - pv.z)) : 0.0f;
		irr_out[probe * 64 + lane] = make_float4(r0 * (1.0f - h) + pv.x * h, r1 * (1.0f - h) + pv.y * h,
												 r2 * (1.0f - h) + pv.z * h, 1.0f);
		const float m = wave_max(mc);
		if(lane == 0) s_maxc[p] = __float_as_uint(m);
		if(irr_only) continue;
#pragma unroll
		for(uint32_t i = 0; i < 4; ++i) {
			const float2 dsum = *reinterpret_cast<const float2*>(s_dep + p * kDepStride + L.dep_stex[i] * 2);
			const float d0 = dsum.x * L.dep_inv[i], d1 = dsum.y * L.dep_inv[i];
			const float2 dv = e.dep[u][i];
			dep_out[probe * 256 + lane + 64 * i] = make_float2(d0 * (1.0f - h) + dv.x * h, d1 * (1.0f - h) + dv.y * h);
		}
	}
}

template <uint32_t kTilesPerPass = 4>
__device__ __forceinline__ void blend_epilogue(const BlendArgs& a, const uint32_t* s_idx, uint32_t np,
											   const float* s_irr, const float* s_dep, uint32_t* s_maxc, bool irr_only) {
	const EpiLane L = epilogue_lane(a);
	const uint32_t wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
	for(uint32_t p0 = wave; p0 < np; p0 += waves * kTilesPerPass) {
		EpiPrev<kTilesPerPass> e;
		epilogue_load(a, L, s_idx, np, p0, irr_only, e);
		epilogue_store(a, L, s_idx, np, p0, s_irr, s_dep, s_maxc, irr_only, e);
	}
}

// State machine (probesUpdate.glsl:110-118) with the max over all irradiance texels.
__device__ __forceinline__ void update_state(uint32_t* states, uint32_t probe, float max_change, uint32_t oor, uint32_t n) {
	uint32_t st = states[probe];
	if(oor >= n) {
		st = 8;
	} else {
		const float s = float(st);
		if(max_change < 0.02f / s) st = min(st + 1u, 8u);
		else if(max_change > 0.04f / s) st = max(st - 1u, 1u);
		else if(max_change > 0.25f) st = 1u;
	}
	states[probe] = st;
}

// FP32 blend on the matrix cores (v_mfma_f32_16x16x4_f32, FP32 operands and accumulation).  A
// block owns NP probe slots from slot0.  Output tiles of 16 x 16: M = the 16 padded texel row tiles
// of the weight matrix (3 irradiance, 13 depth; wave w takes tiles w, w+4, w+8, w+12), N = the
// block's data columns (irradiance: 3 per probe, depth: (d, d^2) per probe, each padded to a
// multiple of 16), K = rays, staged in LDS in chunks of 64 as [column][ray].  A group of 16 rays
// k0 .. k0+15 takes 4 MFMAs: lane l reads the float4 of rays k0 + 4 (l / 16) .. +3 of its weight row
// and of its data column, and step s multiplies component s, so instruction K index l / 16 is ray
// k0 + 4 (l / 16) + s for both operands.  Every output element sees the same sequence of MFMAs for
// any NP, so the result does not depend on the tile size.
typedef float floatx4_t __attribute__((ext_vector_type(4)));

constexpr int kMfmaChunk = 64;              // rays staged per LDS chunk
constexpr int kWRows = 256;                 // padded weight rows: irr 0..47, depth 48..255
constexpr int kDepRow0 = 48;

template <int NP>
__device__ __forceinline__ void blend_fp32_tile(const BlendArgs& a, uint32_t slot0, uint32_t n_pad) {
	constexpr int kNI = (3 * NP + 15) / 16, kND = (2 * NP + 15) / 16;  // N tiles
	constexpr int kCI = 16 * kNI, kCD = 16 * kND;                       // padded columns
	constexpr int kStride = kMfmaChunk + 4;  // floats per staged column: 16-byte skew between columns
	constexpr int kStageFloats = (kCI + kCD) * kStride;
	constexpr int kOutFloats = NP * (kIrrStride + kDepStride);
	__shared__ __attribute__((aligned(16))) float smem[kStageFloats > kOutFloats ? kStageFloats : kOutFloats];
	__shared__ uint32_t s_oor[NP];
	__shared__ uint32_t s_maxc[NP];
	__shared__ uint32_t s_idx[NP];
	float* sb = smem;                            // [column][kStride]: irradiance 3p + c, then depth kCI + 2p + c
	float* s_irr = smem;                         // after the K loop: [probe][kIrrStride]
	float* s_dep = smem + NP * kIrrStride;       //                   [probe][kDepStride]

	// count and n steer loops that contain barriers: make them provably uniform (count is read
	// through a pointer the compiler cannot prove to be global, which it would treat as divergent)
	const uint32_t count = __builtin_amdgcn_readfirstlane(*a.count);
	if(slot0 >= count) return;
	const uint32_t np = min(uint32_t(NP), count - slot0);
	const uint32_t n = __builtin_amdgcn_readfirstlane(a.grid.n);
	n_pad = __builtin_amdgcn_readfirstlane(n_pad);
	const float cell_len = a.grid.cell_len;
	const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
	const int lrow = lane & 15, lgrp = lane >> 4;
	if(t < NP) {
		s_oor[t] = 0;
		s_idx[t] = uint32_t(t) < np ? a.indices[slot0 + t] : 0u;
	}
	floatx4_t acc[4][kNI];
#pragma unroll
	for(int q = 0; q < 4; ++q)
#pragma unroll
		for(int c = 0; c < kNI; ++c) acc[q][c] = floatx4_t{0.0f, 0.0f, 0.0f, 0.0f};

	for(uint32_t k0 = 0; k0 < n_pad; k0 += kMfmaChunk) {
		__syncthreads();
		// staging: 16 consecutive threads per probe, 4 consecutive rays each
		for(int e = t; e < NP * 16; e += 256) {
			const int sp = e >> 4, sk = (e & 15) * 4;
			float v[5][4];
			uint32_t oor = 0;
#pragma unroll
			for(int q = 0; q < 4; ++q) {
				const uint32_t k = k0 + sk + q;
				float4 r = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
				float dc = 0.0f;
				if(uint32_t(sp) < np && k < n) {
					r = a.rays[size_t(slot0 + sp) * n + k];
					if(r.w < 0.0f || r.w > cell_len) ++oor;
					dc = fminf(cell_len, r.w);
					if(dc < 0.0f) dc = cell_len;
				}
				v[0][q] = r.x;
				v[1][q] = r.y;
				v[2][q] = r.z;
				v[3][q] = dc;
				v[4][q] = dc * dc;
			}
			const int col[5] = {sp * 3, sp * 3 + 1, sp * 3 + 2, kCI + sp * 2, kCI + sp * 2 + 1};
#pragma unroll
			for(int c = 0; c < 5; ++c)
				*reinterpret_cast<float4*>(sb + col[c] * kStride + sk) = make_float4(v[c][0], v[c][1], v[c][2], v[c][3]);
			for(int off = 8; off > 0; off >>= 1) oor += __shfl_xor(oor, off, 16);
			if((e & 15) == 0 && oor != 0u) atomicAdd(&s_oor[sp], oor);
		}
		__syncthreads();
#pragma unroll
		for(int kb = 0; kb < kMfmaChunk; kb += 16) {
			const int kk = kb + 4 * lgrp;
			// the data fragments serve all of the wave's row tiles
			float4 bi[kNI], bd[kND];
			if(wave < 3) {
#pragma unroll
				for(int c = 0; c < kNI; ++c) bi[c] = *reinterpret_cast<const float4*>(sb + (c * 16 + lrow) * kStride + kk);
			}
#pragma unroll
			for(int c = 0; c < kND; ++c) bd[c] = *reinterpret_cast<const float4*>(sb + (kCI + c * 16 + lrow) * kStride + kk);
#pragma unroll
			for(int q = 0; q < 4; ++q) {
				const int mt = wave + 4 * q;  // 0..2 irradiance, 3..15 depth
				if(mt < 3) {
					const float4 w = *reinterpret_cast<const float4*>(a.wt + size_t(mt * 16 + lrow) * n_pad + k0 + kk);
#pragma unroll
					for(int c = 0; c < kNI; ++c) {
						acc[q][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(w.x, bi[c].x, acc[q][c], 0, 0, 0);
						acc[q][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(w.y, bi[c].y, acc[q][c], 0, 0, 0);
						acc[q][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(w.z, bi[c].z, acc[q][c], 0, 0, 0);
						acc[q][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(w.w, bi[c].w, acc[q][c], 0, 0, 0);
					}
				} else {
					const float4 w = *reinterpret_cast<const float4*>(
						a.wt + size_t(kDepRow0 + (mt - 3) * 16 + lrow) * n_pad + k0 + kk);
#pragma unroll
					for(int c = 0; c < kND; ++c) {
						acc[q][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(w.x, bd[c].x, acc[q][c], 0, 0, 0);
						acc[q][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(w.y, bd[c].y, acc[q][c], 0, 0, 0);
						acc[q][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(w.z, bd[c].z, acc[q][c], 0, 0, 0);
						acc[q][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(w.w, bd[c].w, acc[q][c], 0, 0, 0);
					}
				}
			}
		}
	}
	__syncthreads();  // staging done: reuse LDS for the accumulators (lane l, register r = D[4 (l/16) + r][l % 16])
#pragma unroll
	for(int q = 0; q < 4; ++q) {
		const int mt = wave + 4 * q;
#pragma unroll
		for(int c = 0; c < kNI; ++c) {
			const int col = c * 16 + lrow;
#pragma unroll
			for(int r = 0; r < 4; ++r) {
				const int row = 4 * lgrp + r;
				if(mt < 3) store_irr_sum(s_irr, mt * 16 + row, col, NP, acc[q][c][r]);
				else if(c < kND) store_dep_sum(s_dep, (mt - 3) * 16 + row, col, NP, acc[q][c][r]);
			}
		}
	}
	__syncthreads();
	const bool irr_only = (a.flags & kBlendIrradianceOnly) != 0u;
	blend_epilogue<>(a, s_idx, np, s_irr, s_dep, s_maxc, irr_only);
	__syncthreads();
	if(uint32_t(t) < np && !irr_only) update_state(a.states, s_idx[t], __uint_as_float(s_maxc[t]), s_oor[t], n);
}

// 4 waves/SIMD: the 4 blocks of a CU that a 16-probe tiling of a large grid needs stay resident
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_blend_fp32(BlendArgs a, uint32_t n_pad) {
	blend_fp32_tile<kProbesPerBlock>(a, blockIdx.x * kProbesPerBlock, n_pad);
}

// ---------------------------------------------------------------------------------------------
// 16-bit MFMA operand types (v_mfma_f32_16x16x32_{bf16,f16}: 8 K values per lane and operand).
//   BF16: w = w_hi + w_lo, x = x_hi + x_lo, acc += w_hi x_hi + w_hi x_lo + w_lo x_hi (3 MFMAs;
//         relative error ~2^-16), normalised by the sum of the split weights.
//   FP16: one product of fp16 weights and data (relative error ~2^-11 per term).
// Accumulator layout (verified for gfx950): lane l, register r holds D[4*(l/16)+r][l%16].
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));

template <uint32_t MODE>
struct MfmaTypes;
template <>
struct MfmaTypes<kBlendBF16> {
	typedef __bf16 T;
	typedef bf16x8_t V;
	static constexpr int kParts = 2;
	__device__ static floatx4_t mma(V a, V b, floatx4_t c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
};
template <>
struct MfmaTypes<kBlendFP16> {
	typedef _Float16 T;
	typedef f16x8_t V;
	static constexpr int kParts = 1;
	__device__ static floatx4_t mma(V a, V b, floatx4_t c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
};
// FP32 on the wide structure: v_mfma_f32_16x16x4_f32, 8 per accumulator and 32-ray K step; the
// step's weights are staged as two 16-ray halves (the LDS geometry of BF16's two parts)
template <>
struct MfmaTypes<kBlendFP32> {
	typedef float T;
	typedef float V;
	static constexpr int kParts = 2;
	__device__ static floatx4_t mma(float a, float b, floatx4_t c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
};

// ---------------------------------------------------------------------------------------------
// MFMA blend, transposed and pipelined (BF16 / FP16 modes, and FP32 on grids whose 64-probe
// workgroups fill the device: there v_mfma_f32_16x16x4_f32 runs 8 per accumulator and K step on
// the unconverted ray data, pairing rays as blend_fp32_tile does).  Per channel the blend of a group of 16
// probes is C^T [16 probes x texels] = X^T [16 probes x n rays] W^T [n x texels]: the ray data is
// the A operand, read by each lane straight from HBM (row = one of the group's probes, 8 rays per
// lane and K step), and the weights are the B operand, shared by the workgroup's eight waves
// through LDS.  A workgroup owns 64 probe slots (four groups of 16); each group's texel tiles are
// split between two waves (irradiance: 3 channels x 3 texel tiles plus depth tiles 3..6; depth
// tiles 7..15), whose accumulators take up to 72 registers per lane, so it runs two waves per SIMD
// and hides latency by prefetching: the ray data of K step s + 2 is in flight in registers and
// the weights of step s + 1 are being copied from L2 into LDS (global_load_lds, no registers)
// while the matrix cores work on step s.
// Products and their order are those of the blend above (hi x hi, lo x hi, hi x lo into FP32).
// LDS weight layout per K step: [part][texel tile][K block (4)][row (16)] x 8 elements, so that a
// wave's B fragment (lane l: texel row l % 16, K block l / 16) is one contiguous 1 KB read and one
// 64-lane global_load_lds instruction fills one (part, tile).
// One 16-byte global -> LDS copy per lane (global_load_lds_dwordx4): the lanes' data lands at
// lds + 16 * lane (lds must be wave-uniform), without passing through registers.
// x = hi + lo for two floats at once: hi = bf16(x) (round to nearest even, one packed conversion),
// lo = bf16(x - hi), where hi's float value is its bits shifted (or masked) into the high half.
__device__ __forceinline__ void split_bf16x2(float a, float b, uint32_t& hi, uint32_t& lo) {
	typedef float f32x2 __attribute__((ext_vector_type(2)));
	typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
	hi = __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{a, b}, bf16x2));
	const float ha = __uint_as_float(hi << 16), hb = __uint_as_float(hi & 0xFFFF0000u);
	lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{a - ha, b - hb}, bf16x2));
}

__device__ __forceinline__ void lds_copy16(const void* src, void* lds) {
	__builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)lds, 16, 0, 0);
}

// Workgroup barrier for LDS traffic: waits for this wave's LDS reads and writes (lgkmcnt) and
// then meets the other waves.  Unlike __syncthreads() it has no release fence, which would also
// wait for every global load in flight (the prefetches); LDS filled by global_load_lds must be
// waited for by the caller (vmcnt) before it.  The empty asm statements keep the compiler from
// moving memory accesses across the barrier.
__device__ __forceinline__ void lds_barrier() {
	__asm__ volatile("" ::: "memory");
	__builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0), vmcnt and expcnt untouched
	__builtin_amdgcn_s_barrier();
	__asm__ volatile("" ::: "memory");
}

constexpr int kWideProbes = 64;   // probe slots per workgroup (16 per probe group)
constexpr int kWideThreads = 512; // eight waves: four probe groups x two halves of the texel tiles
constexpr int kKStep = 32;       // rays per MFMA K step

constexpr int kWideWBytes = 2 * 16 * 64 * 16;  // weights of one K step: 2 parts x 16 tiles x 64 lanes x 16 B
constexpr int kWideHalf = kWideProbes / 2;      // probes whose sums the epilogue takes at a time

// LDS of a wide-blend workgroup, declared by the kernels (one copy however many K-step variants
// a kernel holds).  The two weight buffers are separate objects, so that the compiler sees that
// reading one does not have to wait for the copy into the other; the sums of the epilogue (half of
// the workgroup's probes at a time) have their own.
struct WideLds {
	unsigned char* w[2];  // kWideWBytes each
	float* sums;          // kWideHalf x (kIrrStride + kDepStride)
	uint32_t* s_oor;      // [kWideProbes]
	uint32_t* s_maxc;
	uint32_t* s_idx;
};

// K loop of one wave of the wide blend: its 16 probes (group q) times the texel tiles of half H
// (0: the 3 irradiance tiles and depth tiles 3..6; 1: depth tiles 7..15), summed in `acc`; then
// the sums go to LDS when the epilogue takes this probe group.
template <uint32_t MODE, int STEPS, int H>
__device__ __forceinline__ void wide_wave(const BlendArgs& a, uint32_t slot0, uint32_t np, const WideLds& lds) {
	typedef MfmaTypes<MODE> M;
	typedef typename M::T T;
	typedef typename M::V V;
	constexpr int P = M::kParts;
	constexpr int kTile0 = H == 0 ? 0 : 7, kTile1 = H == 0 ? 7 : 16;  // texel tiles [kTile0, kTile1)
	constexpr int kCh0 = H == 0 ? 0 : 3;                                 // first data channel used
	constexpr int kAcc = H == 0 ? 3 * 3 + 2 * 4 : 2 * 9;                 // accumulator tiles
	const uint32_t n = __builtin_amdgcn_readfirstlane(a.grid.n);
	constexpr uint32_t steps = STEPS;
	const uint32_t n_pad = STEPS * kKStep;
	const float cell_len = a.grid.cell_len;
	const int t = threadIdx.x, lane = t & 63, wave = t >> 6, q = wave & 3;
	const int lrow = lane & 15, lgrp = lane >> 4;
	const uint32_t my_slot = uint32_t(16 * q + lrow);  // this lane's A row: a probe slot of the workgroup
	// the workgroup's rays as a buffer resource (np probes x n rays): loads past its end return
	// zero, so no address needs a clamp, and the lane's offset stays in one register for the whole
	// K loop (per-load clamped addresses were reused as MFMA results right after the loads, which
	// made the compiler wait for every load in flight)
	const __amdgpu_buffer_rsrc_t rays_rsrc = __builtin_amdgcn_make_buffer_rsrc(
		const_cast<float4*>(a.rays + size_t(slot0) * n), short(0), int(np * n * 16u), 0x00020000);
	constexpr bool kF32 = MODE == kBlendFP32;
	// ray of K step s that load j of lane group g = l / 16 takes: 16-bit, 4 j + g (load j reads 4
	// consecutive rays, 64 bytes, of each of the wave's 16 probes; the weights are stored in the
	// matching wide_column order); FP32, 16 (j / 4) + 4 g + j % 4, so that MFMA j pairs the rays the
	// 16-ray groups of blend_fp32_tile pair, in the same order (the same sums)
	auto ray_of = [&](int j) { return kF32 ? uint32_t(16 * (j >> 2) + 4 * lgrp + (j & 3)) : uint32_t(4 * j + lgrp); };
	const uint32_t my_off = (my_slot * n + (kF32 ? 4u * uint32_t(lgrp) : uint32_t(lgrp))) * 16u;
	const T* __restrict__ wsrc[2] = {static_cast<const T*>(kF32 ? static_cast<const void*>(a.wt) : a.w_hi16),
									  static_cast<const T*>(a.w_lo16)};

	// weights of K step s into LDS buffer b: (part, tile) j by wave j % 8, lane l copies texel row
	// 16 tile + l % 16, K block l / 16: 16-bit, 8 weights in wide_column order; FP32, the 4 weights
	// of rays 16 part + 4 (l / 16) + 0..3 (the natural order of the FP32 weight rows)
	auto stage_weights = [&](uint32_t s, int b) {
#pragma unroll
		for(int i = 0; i < P * 2; ++i) {
			const int j = wave + 8 * i, part = j >> 4, tile = j & 15;
			const T* src = kF32 ? wsrc[0] + size_t(16 * tile + lrow) * n_pad + s * kKStep + part * 16 + lgrp * 4
								: wsrc[part] + size_t(16 * tile + lrow) * n_pad + s * kKStep + lgrp * 8;
			unsigned char* dst = lds.w[b] + (part * 16 + tile) * 1024;
			lds_copy16(src, dst);
		}
	};
	// ray data of K step s: lane (probe l % 16, K block g = l / 16) takes rays 32 s + ray_of(j),
	// j = 0..7.  Rays that do not exist
	// load as zero, which contributes nothing to any column: absent probes lie past the buffer's
	// records, and in the last, partial K step of a ray count that is not a multiple of 32 the
	// rays beyond n get an offset past the records as well.  The whole offset is in the range-
	// checked vector operand (the scalar offset is not checked against the records).
	auto load_rays = [&](uint32_t s, float4* r) {
		typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
		const uint32_t base = my_off + s * (kKStep * 16u);
		// rays of step s that exist (uniform): all 32 before the last two steps (the steps cover n
		// rounded up to 64), so that the selects below fold away there
		const uint32_t lim = s + 2 < steps ? uint32_t(kKStep) : (n > s * kKStep ? min(n - s * kKStep, uint32_t(kKStep)) : 0u);
#pragma unroll
		for(int j = 0; j < 8; ++j) {
			const uint32_t rel = kF32 ? uint32_t(16 * (16 * (j >> 2) + (j & 3))) : uint32_t(64 * j);  // ray_of(j) - ray_of(0), bytes
			const uint32_t off = ray_of(j) < lim ? base + rel : 0x80000000u;
			const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(rays_rsrc, off, 0, 0);
			r[j] = make_float4(__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z), __uint_as_float(v.w));
		}
	};

	floatx4_t acc[kAcc];
#pragma unroll
	for(int k = 0; k < kAcc; ++k) acc[k] = floatx4_t{0.0f, 0.0f, 0.0f, 0.0f};
	uint32_t my_oor = 0;

	// one K step on the data in r (rays of step s) and LDS buffer b; then r receives step s + 2
	auto step = [&](uint32_t s, int b, float4* r) {
		// the weights of step s (and everything but the newest ray prefetch) have landed; the other
		// waves are done with buffer b ^ 1
		if(s + 1 < steps) __builtin_amdgcn_s_waitcnt(0x0F78);  // vmcnt(8): the step s + 1 rays may stay in flight
		else __builtin_amdgcn_s_waitcnt(0x0F70);               // vmcnt(0)
		lds_barrier();
		// the depth half's ray data of step s (distances only) counts as produced here, after the
		// wait: otherwise the compiler hoists its conversion into an earlier step, where the
		// in-order load counter (with the weight copies in flight as well) can only be waited on
		// down to zero
		if constexpr(H == 1) {
#pragma unroll
			for(int j = 0; j < 8; ++j) {
				float w = r[j].w;
				__asm__ volatile("" : "+v"(w));
				r[j].w = w;
			}
		} else if constexpr(kF32) {
			// the same for the FP32 irradiance half, whose MFMAs take the ray data unconverted
#pragma unroll
			for(int j = 0; j < 8; ++j) {
				floatx4_t x = floatx4_t{r[j].x, r[j].y, r[j].z, r[j].w};
				__asm__ volatile("" : "+v"(x));
				r[j] = make_float4(x[0], x[1], x[2], x[3]);
			}
		}
		if(s + 1 < steps) stage_weights(s + 1, b ^ 1);
		// A fragments (hi / lo) of the data channels this half uses (16-bit)
		V A[5][2];
		float v[8][5];
#pragma unroll
		for(int j = 0; j < 8; ++j) {
			const float4 x = r[j];  // zero where the ray does not exist: in range, nothing in any column
			if(H == 0 && (x.w < 0.0f || x.w > cell_len)) ++my_oor;
			float dc = fminf(cell_len, x.w);
			if(dc < 0.0f) dc = cell_len;
			v[j][0] = x.x;
			v[j][1] = x.y;
			v[j][2] = x.z;
			v[j][3] = dc;
			v[j][4] = dc * dc;
		}
		// the count is taken in its step: otherwise the compiler sinks the comparisons to the end
		// of the K loop and keeps every step's distances live until then
		if constexpr(H == 0) __asm__ volatile("" : "+v"(my_oor));
#pragma unroll
		for(int c = kCh0; c < 5; ++c) {
			if constexpr(kF32) {
				// FP32: the operands are the channel values themselves
			} else if constexpr(P == 2) {
				// two rays at a time: one packed rounding for the high parts, their exact float
				// values by a shift and a mask, one packed rounding for the residuals
				uint32_t hi[4], lo[4];
#pragma unroll
				for(int i = 0; i < 4; ++i) split_bf16x2(v[2 * i][c], v[2 * i + 1][c], hi[i], lo[i]);
				A[c][0] = __builtin_bit_cast(V, hi);
				A[c][1] = __builtin_bit_cast(V, lo);
			} else {
#pragma unroll
				for(int j = 0; j < 8; ++j) A[c][0][j] = T(v[j][c]);
			}
		}
		if constexpr(!kF32) {
			if(s + 2 < steps) load_rays(s + 2, r);
		}
		const unsigned char* wb = lds.w[b] + lane * 16;
		// per texel tile, the products in the order hi x hi, lo x hi, hi x lo for every channel, so
		// that consecutive MFMAs update different accumulators
		// FP32: the next tile's weights are read while this tile's MFMAs issue
		float4 nb0, nb1;
		if constexpr(kF32) {
			nb0 = *reinterpret_cast<const float4*>(wb + kTile0 * 1024);
			nb1 = *reinterpret_cast<const float4*>(wb + (16 + kTile0) * 1024);
		}
#pragma unroll
		for(int tt = kTile0; tt < kTile1; ++tt) {
			const int c0 = tt < 3 ? 0 : 3, nc = tt < 3 ? 3 : 2;
			// accumulator of (channel c, tile tt): irradiance 3 tt + c; depth base + 2 (tt - first) + c
			const int k0 = tt < 3 ? 3 * tt : (H == 0 ? 9 + 2 * (tt - 3) : 2 * (tt - 7));
			if constexpr(kF32) {
				// 8 MFMAs per accumulator: MFMA j multiplies ray ray_of(j) (K index l / 16) of each
				// probe row with its weights, the channels interleaved
				const float4 B0 = nb0, B1 = nb1;
				if(tt + 1 < kTile1) {
					nb0 = *reinterpret_cast<const float4*>(wb + (tt + 1) * 1024);
					nb1 = *reinterpret_cast<const float4*>(wb + (16 + tt + 1) * 1024);
				}
				const float bw[8] = {B0.x, B0.y, B0.z, B0.w, B1.x, B1.y, B1.z, B1.w};
#pragma unroll
				for(int j = 0; j < 8; ++j)
#pragma unroll
					for(int c = 0; c < nc; ++c) acc[k0 + c] = M::mma(v[j][c0 + c], bw[j], acc[k0 + c]);
				continue;
			}
			V B[2];
#pragma unroll
			for(int p = 0; p < P; ++p) B[p] = *reinterpret_cast<const V*>(wb + (p * 16 + tt) * 1024);
#pragma unroll
			for(int c = 0; c < nc; ++c) acc[k0 + c] = M::mma(A[c0 + c][0], B[0], acc[k0 + c]);
			if(P == 2) {
#pragma unroll
				for(int c = 0; c < nc; ++c) acc[k0 + c] = M::mma(A[c0 + c][1], B[0], acc[k0 + c]);
#pragma unroll
				for(int c = 0; c < nc; ++c) acc[k0 + c] = M::mma(A[c0 + c][0], B[1], acc[k0 + c]);
			}
		}
		// FP32: the MFMAs read the ray registers themselves, so the next loads into them follow
		if constexpr(kF32) {
			if(s + 2 < steps) load_rays(s + 2, r);
		}
	};

	float4 ra[8], rb[8];
	stage_weights(0, 0);
	load_rays(0, ra);
	if(steps > 1) load_rays(1, rb);
	// unrolled up to 8 steps (256 rays); beyond that the compiler's scheduling across the unrolled
	// steps spills (216 bytes per lane at 16 steps), and a loop over step pairs runs faster
	if constexpr(STEPS <= 8) {
#pragma unroll
		for(uint32_t s = 0; s < steps; s += 2) {
			step(s, 0, ra);
			if(s + 1 < steps) step(s + 1, 1, rb);
		}
	} else {
#pragma unroll 1
		for(uint32_t s = 0; s < steps; s += 2) {
			step(s, 0, ra);
			if(s + 1 < steps) step(s + 1, 1, rb);
		}
	}

	// the epilogue's exposed-atlas loads for both halves of the probes, issued now: they need only
	// the probe indices, and their latency then passes while the sums go through LDS (one pass
	// per half: 32 probes over eight waves, four tiles each)
	static_assert(kWideHalf == (kWideThreads / 64) * 4, "one epilogue pass per half");
	const EpiLane el = epilogue_lane(a);
	EpiPrev<4> pre[2];
#pragma unroll
	for(int half = 0; half < 2; ++half) {
		const uint32_t first = uint32_t(half * kWideHalf);
		if(first < np) epilogue_load(a, el, lds.s_idx + first, min(uint32_t(kWideHalf), np - first), uint32_t(wave), false, pre[half]);
	}
	// out-of-range counts: the four lanes of a probe row (l % 16 equal)
	if(H == 0) {
		my_oor += __shfl_xor(my_oor, 16);
		my_oor += __shfl_xor(my_oor, 32);
		if(lgrp == 0) lds.s_oor[my_slot] = my_oor;
	}
	// epilogue, half of the probes at a time: groups 2h, 2h + 1 store their sums (D row 4 (l / 16) +
	// r = probe, column l % 16 = texel), then all eight waves blend, store and reduce those tiles
	float* s_irr = lds.sums;
	float* s_dep = s_irr + kWideHalf * kIrrStride;
#pragma unroll
	for(int half = 0; half < 2; ++half) {
		const uint32_t first = uint32_t(half * kWideHalf);
		if(first >= np) break;
		__syncthreads();
		if((q >> 1) == half) {
#pragma unroll
			for(int r = 0; r < 4; ++r) {
				const int slot = 16 * (q & 1) + 4 * lgrp + r;
#pragma unroll
				for(int tt = kTile0; tt < kTile1; ++tt) {
					const int tex = 16 * (tt < 3 ? tt : tt - 3) + lrow;
					const int k0 = tt < 3 ? 3 * tt : (H == 0 ? 9 + 2 * (tt - 3) : 2 * (tt - 7));
					if(tt < 3) {
						if(tex < int(kIrrTexels)) {
#pragma unroll
							for(int c = 0; c < 3; ++c) s_irr[slot * kIrrStride + tex * 4 + c] = acc[k0 + c][r];
						}
					} else if(tex < int(kDepTexels)) {
						s_dep[slot * kDepStride + tex * 2 + 0] = acc[k0][r];
						s_dep[slot * kDepStride + tex * 2 + 1] = acc[k0 + 1][r];
					}
				}
			}
		}
		__syncthreads();
		epilogue_store(a, el, lds.s_idx + first, min(uint32_t(kWideHalf), np - first), uint32_t(wave), s_irr, s_dep,
					   lds.s_maxc + first, false, pre[half]);
	}
}

template <uint32_t MODE, int STEPS>
__device__ __forceinline__ void blend_mfma_wide(const BlendArgs& a, uint32_t slot0, const WideLds& lds) {
	const uint32_t count = __builtin_amdgcn_readfirstlane(*a.count);  // uniform: see blend_fp32_tile
	if(slot0 >= count) return;
	const uint32_t np = min(uint32_t(kWideProbes), count - slot0);
	const int t = threadIdx.x;
	if(t < kWideProbes) lds.s_idx[t] = uint32_t(t) < np ? a.indices[slot0 + t] : 0u;
	// waves 0..3 take the irradiance tiles and the first depth tiles of probe groups 0..3, waves
	// 4..7 the other depth tiles of the same groups (wave-uniform branch: every wave meets the same
	// barriers)
	if(__builtin_amdgcn_readfirstlane(t >> 6) < 4) wide_wave<MODE, STEPS, 0>(a, slot0, np, lds);
	else wide_wave<MODE, STEPS, 1>(a, slot0, np, lds);
	__syncthreads();
	if(uint32_t(t) < np)
		update_state(a.states, lds.s_idx[t], __uint_as_float(lds.s_maxc[t]), lds.s_oor[t], __builtin_amdgcn_readfirstlane(a.grid.n));
}

// eight waves, two per SIMD (a workgroup per CU): up to 256 registers per lane for a wave's
// accumulators and two K steps of prefetched ray data
#define DDGI_WIDE_LDS                                                                             \
	__shared__ __attribute__((aligned(16))) unsigned char wide_w0[kWideWBytes];                   \
	__shared__ __attribute__((aligned(16))) unsigned char wide_w1[kWideWBytes];                   \
	__shared__ __attribute__((aligned(16))) float wide_sums[kWideHalf * (kIrrStride + kDepStride)]; \
	__shared__ uint32_t wide_small[3 * kWideProbes];                                              \
	const WideLds lds{{wide_w0, wide_w1}, wide_sums, wide_small, wide_small + kWideProbes, wide_small + 2 * kWideProbes}

template <uint32_t MODE, int STEPS>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_blend_wide(BlendArgs a) {
	DDGI_WIDE_LDS;
	blend_mfma_wide<MODE, STEPS>(a, blockIdx.x * kWideProbes, lds);
}

// ---- batched update ----
// Weight rows of every volume: grid (256 rows, volumes).
__global__ __launch_bounds__(256) void k_batch_weights(BatchLaunch b) {
	const BatchVolume& v = b.vols[blockIdx.y];
	blend_weight_row(v.blend.grid, v.dirs, v.perm, v.wt, v.inv_sum, v.w_hi, v.w_lo, b.blend_mode, v.n_pad, blockIdx.x);
}

// Blend over the tiles of all volumes: one workgroup per blend item (a tile of one volume's probe
// slots); thread 0 stamps the volume's completion time once the tile is done.  (A persistent form
// taking tiles from a work queue was 0.1 ms slower per batched update.)
template <uint32_t MODE, int NP>
__global__ __launch_bounds__(256) void k_batch_blend(BatchLaunch b) {
	const uint32_t vol = __builtin_amdgcn_readfirstlane(b.blend_items[blockIdx.x].volume);
	const uint32_t first = __builtin_amdgcn_readfirstlane(b.blend_items[blockIdx.x].first);
	const BatchVolume& v = b.vols[vol];
	blend_fp32_tile<NP>(v.blend, first, v.n_pad);
	__syncthreads();
	if(threadIdx.x == 0) atomicMax(&b.ctl->done[vol], (unsigned long long)wall_clock64());
}

// the MFMA blends of a batch: one workgroup per 64-slot item, as k_blend_wide
template <uint32_t MODE>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_batch_blend_wide(BatchLaunch b) {
	const uint32_t vol = __builtin_amdgcn_readfirstlane(b.blend_items[blockIdx.x].volume);
	const uint32_t first = __builtin_amdgcn_readfirstlane(b.blend_items[blockIdx.x].first);
	const BatchVolume& v = b.vols[vol];
	DDGI_WIDE_LDS;
	switch(__builtin_amdgcn_readfirstlane(v.n_pad) / kKStep) {
		case 2: blend_mfma_wide<MODE, 2>(v.blend, first, lds); break;
		case 4: blend_mfma_wide<MODE, 4>(v.blend, first, lds); break;
		case 6: blend_mfma_wide<MODE, 6>(v.blend, first, lds); break;
		case 8: blend_mfma_wide<MODE, 8>(v.blend, first, lds); break;
		case 10: blend_mfma_wide<MODE, 10>(v.blend, first, lds); break;
		case 12: blend_mfma_wide<MODE, 12>(v.blend, first, lds); break;
		case 14: blend_mfma_wide<MODE, 14>(v.blend, first, lds); break;
		case 16: blend_mfma_wide<MODE, 16>(v.blend, first, lds); break;
	}
	__syncthreads();
	if(threadIdx.x == 0) atomicMax(&b.ctl->done[vol], (unsigned long long)wall_clock64());
}

template <uint32_t MODE>
const void* wide_blend_kernel_of(uint32_t steps) {
	switch(steps) {
		case 2: return reinterpret_cast<const void*>(&k_blend_wide<MODE, 2>);
		case 4: return reinterpret_cast<const void*>(&k_blend_wide<MODE, 4>);
		case 6: return reinterpret_cast<const void*>(&k_blend_wide<MODE, 6>);
		case 8: return reinterpret_cast<const void*>(&k_blend_wide<MODE, 8>);
		case 10: return reinterpret_cast<const void*>(&k_blend_wide<MODE, 10>);
		case 12: return reinterpret_cast<const void*>(&k_blend_wide<MODE, 12>);
		case 14: return reinterpret_cast<const void*>(&k_blend_wide<MODE, 14>);
		case 16: return reinterpret_cast<const void*>(&k_blend_wide<MODE, 16>);
	}
	return nullptr;
}

// The wide MFMA blend for `steps` K steps of 32 rays (n_pad = 32 steps, a multiple of 64 up to 512).
const void* wide_blend_kernel(uint32_t mode, uint32_t steps) {
	if(mode == kBlendBF16) return wide_blend_kernel_of<kBlendBF16>(steps);
	if(mode == kBlendFP16) return wide_blend_kernel_of<kBlendFP16>(steps);
	if(mode == kBlendFP32) return wide_blend_kernel_of<kBlendFP32>(steps);
	return nullptr;
}

const void* batch_blend_kernel(uint32_t mode, uint32_t probes) {
	if(mode == kBlendBF16) return probes == kMfmaProbes ? reinterpret_cast<const void*>(&k_batch_blend_wide<kBlendBF16>) : nullptr;
	if(mode == kBlendFP16) return probes == kMfmaProbes ? reinterpret_cast<const void*>(&k_batch_blend_wide<kBlendFP16>) : nullptr;
	switch(probes) {
		case kMfmaProbes: return reinterpret_cast<const void*>(&k_batch_blend_wide<kBlendFP32>);
		case 4: return reinterpret_cast<const void*>(&k_batch_blend<kBlendFP32, 4>);
		case 8: return reinterpret_cast<const void*>(&k_batch_blend<kBlendFP32, 8>);
		case 16: return reinterpret_cast<const void*>(&k_batch_blend<kBlendFP32, 16>);
		case 32: return reinterpret_cast<const void*>(&k_batch_blend<kBlendFP32, 32>);
	}
	return nullptr;
}

}  // namespace

hipError_t launch_blend_weights(const GridConst& g, const float4* dirs, const uint16_t* order, float* wt, float* inv_sum,
								void* w_hi16, void* w_lo16, uint32_t mode, hipStream_t s) {
	const uint32_t n_pad = (g.n + kMfmaChunk - 1) & ~uint32_t(kMfmaChunk - 1);
	hipLaunchKernelGGL(k_blend_weights, dim3(kWRows), dim3(256), 0, s, g, dirs, order, wt, inv_sum, w_hi16, w_lo16, mode,
					   n_pad);
	return hipGetLastError();
}

hipError_t launch_blend(const BlendArgs& a, uint32_t mode, hipStream_t s, bool fp32_wide) {
	if(a.max_slots == 0) return hipSuccess;
	const uint32_t n_pad = (a.grid.n + kMfmaChunk - 1) & ~uint32_t(kMfmaChunk - 1);
	if(mode == kBlendFP32 && (!fp32_wide || (a.flags & kBlendIrradianceOnly))) {
		hipLaunchKernelGGL(k_blend_fp32, dim3((a.max_slots + kProbesPerBlock - 1) / kProbesPerBlock), dim3(256), 0, s, a, n_pad);
		return hipGetLastError();
	}
	const void* f = wide_blend_kernel(mode, n_pad / kKStep);
	if(!f) return hipErrorInvalidValue;
	BlendArgs arg = a;
	void* args[] = {&arg};
	return hipLaunchKernel(f, dim3((a.max_slots + kWideProbes - 1) / kWideProbes), dim3(kWideThreads), args, 0, s);
}

hipError_t launch_batch_weights(const BatchLaunch& b, hipStream_t s) {
	hipLaunchKernelGGL(k_batch_weights, dim3(kWRows, b.volume_count), dim3(256), 0, s, b);
	return hipGetLastError();
}

hipError_t launch_batch_blend(const BatchLaunch& b, hipStream_t s) {
	const void* f = batch_blend_kernel(b.blend_mode, b.blend_probes);
	if(!f || b.blend_item_count == 0) return hipErrorInvalidValue;
	BatchLaunch arg = b;
	void* args[] = {&arg};
	return hipLaunchKernel(f, dim3(b.blend_item_count), dim3(b.blend_probes == kMfmaProbes ? kWideThreads : 256), args, 0, s);
}

}  // namespace ddgi
