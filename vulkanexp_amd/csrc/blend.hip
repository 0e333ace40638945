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

// One block per padded weight row: the weight of every ray and their sum, stored in [row][n_pad]
// layout as FP32, or as 16-bit copies (BF16: hi/lo split; FP16: one copy).
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
			w_hi[size_t(row) * n_pad + i] = hi;
			w_lo[size_t(row) * n_pad + i] = lo;
			w = float(hi) + float(lo);
		} else if(mode == kBlendFP16) {
			const _Float16 h = _Float16(w);
			w_h[size_t(row) * n_pad + i] = h;
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

// Border destinations mirroring interior texel (x, y) of an R x R tile (probesCopyBorders.comp
// tables == octahedral mirror).  Returns the count, writes up to 3 destinations.
template <int R>
__device__ __forceinline__ int border_targets(int x, int y, int dx[3], int dy[3]) {
	int k = 0;
	if(x == 1) {
		dx[k] = 0;
		dy[k] = R - 1 - y;
		++k;
	}
	if(x == R - 2) {
		dx[k] = R - 1;
		dy[k] = R - 1 - y;
		++k;
	}
	if(y == 1) {
		dx[k] = R - 1 - x;
		dy[k] = 0;
		++k;
	}
	if(y == R - 2) {
		dx[k] = R - 1 - x;
		dy[k] = R - 1;
		++k;
	}
	// corners
	if(x == R - 2 && y == R - 2 && k < 3) {
		dx[k] = 0;
		dy[k] = 0;
		++k;
	}
	if(x == R - 2 && y == 1 && k < 3) {
		dx[k] = 0;
		dy[k] = R - 1;
		++k;
	}
	if(x == 1 && y == R - 2 && k < 3) {
		dx[k] = R - 1;
		dy[k] = 0;
		++k;
	}
	if(x == 1 && y == 1 && k < 3) {
		dx[k] = R - 1;
		dy[k] = R - 1;
		++k;
	}
	return k;
}

// Epilogue shared by the blends, for the block's np probes (atlas indices in s_idx): normalise the
// sums the K loop left in LDS ([48][ci] irradiance rows, [208][cd] depth rows; probe p's channels in
// columns 3p + c and 2p + c), apply hysteresis against the exposed atlas, store interior + mirrored
// border texels into the next atlas and collect maxChange per probe in s_maxc.  A thread takes four
// items at a time and issues their four exposed-atlas loads before any of their stores, so the
// load latency is paid once per four items.
__device__ __forceinline__ void blend_epilogue(const BlendArgs& a, const uint32_t* s_idx, uint32_t np,
											   const float* s_out_irr, int ci, const float* s_out_dep, int cd,
											   uint32_t* s_maxc, bool irr_only) {
	constexpr uint32_t kU = 4;
	const float h = a.grid.hysteresis;
	const float4* __restrict__ irr_prev = a.irr_prev;
	const float2* __restrict__ dep_prev = a.dep_prev;
	float4* __restrict__ irr_out = a.irr_out;
	float2* __restrict__ dep_out = a.dep_out;
	const uint32_t ni = np * kIrrTexels;
	for(uint32_t e0 = threadIdx.x; e0 < ni; e0 += 256u * kU) {
		float4 prev[kU];
#pragma unroll
		for(uint32_t u = 0; u < kU; ++u) {
			const uint32_t e = e0 + 256u * u, p = e / kIrrTexels, tex = e % kIrrTexels;
			if(e < ni) prev[u] = irr_prev[size_t(s_idx[p]) * 64 + (tex / 6u + 1u) * 8 + tex % 6u + 1u];
		}
#pragma unroll
		for(uint32_t u = 0; u < kU; ++u) {
			const uint32_t e = e0 + 256u * u, p = e / kIrrTexels, tex = e % kIrrTexels;
			if(e >= ni) break;
			const int x = int(tex % 6u) + 1, y = int(tex / 6u) + 1;
			const float inv = a.inv_sum[tex];
			const float* sum = s_out_irr + tex * ci + p * 3;
			const float r0 = sum[0] * inv, r1 = sum[1] * inv, r2 = sum[2] * inv;
			const float4 pv = prev[u];
			const float mc = fmaxf(fmaxf(fabsf(r0 - pv.x), fabsf(r1 - pv.y)), fabsf(r2 - pv.z));
			const float4 v = make_float4(r0 * (1.0f - h) + pv.x * h, r1 * (1.0f - h) + pv.y * h,
										 r2 * (1.0f - h) + pv.z * h, 1.0f);
			float4* tile = irr_out + size_t(s_idx[p]) * 64;
			tile[y * 8 + x] = v;
			int dx[3], dy[3];
			const int nb = border_targets<8>(x, y, dx, dy);
			for(int k = 0; k < nb; ++k) tile[dy[k] * 8 + dx[k]] = v;
			atomicMax(&s_maxc[p], __float_as_uint(mc));
		}
	}
	if(irr_only) return;
	const uint32_t nd = np * kDepTexels;
	for(uint32_t e0 = threadIdx.x; e0 < nd; e0 += 256u * kU) {
		float2 prev[kU];
#pragma unroll
		for(uint32_t u = 0; u < kU; ++u) {
			const uint32_t e = e0 + 256u * u, p = e / kDepTexels, tex = e % kDepTexels;
			if(e < nd) prev[u] = dep_prev[size_t(s_idx[p]) * 256 + (tex / 14u + 1u) * 16 + tex % 14u + 1u];
		}
#pragma unroll
		for(uint32_t u = 0; u < kU; ++u) {
			const uint32_t e = e0 + 256u * u, p = e / kDepTexels, tex = e % kDepTexels;
			if(e >= nd) break;
			const int x = int(tex % 14u) + 1, y = int(tex / 14u) + 1;
			const float inv = a.inv_sum[kIrrTexels + tex];
			const float* sum = s_out_dep + tex * cd + p * 2;
			const float r0 = sum[0] * inv, r1 = sum[1] * inv;
			const float2 pv = prev[u];
			const float2 v = make_float2(r0 * (1.0f - h) + pv.x * h, r1 * (1.0f - h) + pv.y * h);
			float2* tile = dep_out + size_t(s_idx[p]) * 256;
			tile[y * 16 + x] = v;
			int dx[3], dy[3];
			const int nb = border_targets<16>(x, y, dx, dy);
			for(int k = 0; k < nb; ++k) tile[dy[k] * 16 + dx[k]] = v;
		}
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
	constexpr int kOutFloats = 48 * kCI + 208 * kCD;
	__shared__ __attribute__((aligned(16))) float smem[kStageFloats > kOutFloats ? kStageFloats : kOutFloats];
	__shared__ uint32_t s_oor[NP];
	__shared__ uint32_t s_maxc[NP];
	__shared__ uint32_t s_idx[NP];
	float* sb = smem;                    // [column][kStride]: irradiance 3p + c, then depth kCI + 2p + c
	float* s_out_irr = smem;             // after the K loop: [48][kCI]
	float* s_out_dep = smem + 48 * kCI;  //                   [208][kCD]

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
		s_maxc[t] = 0;
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
#pragma unroll
			for(int r = 0; r < 4; ++r) {
				const int row = 4 * lgrp + r;
				if(mt < 3) s_out_irr[(mt * 16 + row) * kCI + c * 16 + lrow] = acc[q][c][r];
				else if(c < kND) s_out_dep[((mt - 3) * 16 + row) * kCD + c * 16 + lrow] = acc[q][c][r];
			}
		}
	}
	__syncthreads();
	const bool irr_only = (a.flags & kBlendIrradianceOnly) != 0u;
	blend_epilogue(a, s_idx, np, s_out_irr, kCI, s_out_dep, kCD, s_maxc, irr_only);
	__syncthreads();
	if(uint32_t(t) < np && !irr_only) update_state(a.states, s_idx[t], __uint_as_float(s_maxc[t]), s_oor[t], n);
}

// 4 waves/SIMD: the 4 blocks of a CU that a 16-probe tiling of a large grid needs stay resident
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_blend_fp32(BlendArgs a, uint32_t n_pad) {
	blend_fp32_tile<kProbesPerBlock>(a, blockIdx.x * kProbesPerBlock, n_pad);
}

// ---------------------------------------------------------------------------------------------
// MFMA blend.  Per block: 16 probe slots.  GEMM shapes per block, as 16x16 output tiles of
// v_mfma_f32_16x16x32_{bf16,f16}:
//   irradiance: M = 48 (36 texels + 12 zero rows) x N = 48 (16 probes x rgb)   -> 3 x 3 tiles
//   depth:      M = 208 (196 + 12)                x N = 32 (16 probes x d, d^2) -> 13 x 2 tiles
// K = rays, in steps of 32.  The 16 M-tiles (3 irradiance, 13 depth) are dealt to the 4 waves
// (wave w: tiles w, w+4, w+8, w+12), so each wave loads a weight fragment once per K step and
// reuses it for all N-tiles of its row.  Weights (A) come straight from L2 in the padded
// [256][n_pad] 16-bit layout; ray data (B) is staged in LDS per 64-ray chunk, column-major so a
// lane's 8 consecutive K values are one 16-byte read.
//   BF16: w = w_hi + w_lo, x = x_hi + x_lo, acc += w_hi x_hi + w_hi x_lo + w_lo x_hi (3 MFMAs;
//         relative error ~2^-16), normalised by the sum of the split weights.
//   FP16: one product of fp16 weights and data (relative error ~2^-11 per term).
// Accumulator layout (verified for gfx950): lane l, register r holds D[4*(l/16)+r][l%16].
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));

constexpr int kBStride = kMfmaChunk + 8;    // 16-bit elements per LDS column (padded)
constexpr int kIrrCols = kProbesPerBlock * 3, kDepCols = kProbesPerBlock * 2;

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

template <uint32_t MODE>
__device__ __forceinline__ void blend_mfma_tile(const BlendArgs& a, uint32_t slot0, uint32_t n_pad) {
	typedef MfmaTypes<MODE> M;
	typedef typename M::T T;
	typedef typename M::V V;
	constexpr int P = M::kParts;
	// LDS: staging (B operand) and, after the K loop, the accumulator tiles (aliased)
	constexpr int kStageBytes = P * (kIrrCols + kDepCols) * kBStride * 2;
	constexpr int kOutBytes = (48 * kIrrCols + 208 * kDepCols) * 4;
	__shared__ __attribute__((aligned(16))) unsigned char smem[kStageBytes > kOutBytes ? kStageBytes : kOutBytes];
	__shared__ uint32_t s_oor[kProbesPerBlock];
	__shared__ uint32_t s_maxc[kProbesPerBlock];
	__shared__ uint32_t s_idx[kProbesPerBlock];
	T* sb = reinterpret_cast<T*>(smem);  // [part][col][kBStride], irr cols first then depth cols
	float* s_out_irr = reinterpret_cast<float*>(smem);          // [48][48]
	float* s_out_dep = s_out_irr + 48 * kIrrCols;               // [208][32]

	const uint32_t count = __builtin_amdgcn_readfirstlane(*a.count);  // uniform: see blend_fp32_tile
	if(slot0 >= count) return;
	const uint32_t np = min(uint32_t(kProbesPerBlock), count - slot0);
	const uint32_t n = __builtin_amdgcn_readfirstlane(a.grid.n);
	n_pad = __builtin_amdgcn_readfirstlane(n_pad);
	const float cell_len = a.grid.cell_len;
	const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
	const int lrow = lane & 15, lgrp = lane >> 4;
	const T* __restrict__ wparts[2] = {static_cast<const T*>(a.w_hi16), static_cast<const T*>(a.w_lo16)};

	floatx4_t acc[4][3];
#pragma unroll
	for(int q = 0; q < 4; ++q)
#pragma unroll
		for(int c = 0; c < 3; ++c) acc[q][c] = floatx4_t{0.0f, 0.0f, 0.0f, 0.0f};

	// staging assignment: thread t loads rays 4*(t%16) .. +3 of probe t/16 of the chunk
	const int sp = t >> 4, sk = (t & 15) * 4;
	uint32_t my_oor = 0;
	for(uint32_t k0 = 0; k0 < n_pad; k0 += kMfmaChunk) {
		__syncthreads();
		for(int q = 0; q < 4; ++q) {
			const uint32_t k = k0 + sk + q;
			float4 r = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
			float dc = 0.0f;
			if(uint32_t(sp) < np && k < n) {
				r = a.rays[size_t(slot0 + sp) * n + k];
				if(r.w < 0.0f || r.w > cell_len) ++my_oor;
				dc = fminf(cell_len, r.w);
				if(dc < 0.0f) dc = cell_len;
			}
			const float v[5] = {r.x, r.y, r.z, dc, dc * dc};
			const int col[5] = {sp * 3, sp * 3 + 1, sp * 3 + 2, kIrrCols + sp * 2, kIrrCols + sp * 2 + 1};
#pragma unroll
			for(int c = 0; c < 5; ++c) {
				const T hi = T(v[c]);
				sb[(0 * (kIrrCols + kDepCols) + col[c]) * kBStride + sk + q] = hi;
				if(P == 2) sb[(1 * (kIrrCols + kDepCols) + col[c]) * kBStride + sk + q] = T(v[c] - float(hi));
			}
		}
		__syncthreads();
#pragma unroll
		for(int ks = 0; ks < kMfmaChunk; ks += 32) {
			const uint32_t kg = k0 + ks + lgrp * 8;  // this lane's 8 K values (same mapping for A and B)
#pragma unroll
			for(int q = 0; q < 4; ++q) {
				const int mt = wave + 4 * q;            // 0..2 irradiance, 3..15 depth
				const bool irr = mt < 3;
				const int wrow = irr ? mt * 16 + lrow : kDepRow0 + (mt - 3) * 16 + lrow;
				V A[2];
#pragma unroll
				for(int p = 0; p < P; ++p) A[p] = *reinterpret_cast<const V*>(wparts[p] + size_t(wrow) * n_pad + kg);
				const int nt_count = irr ? 3 : 2;
#pragma unroll
				for(int nt = 0; nt < 3; ++nt) {
					if(nt >= nt_count) break;
					const int col = irr ? nt * 16 + lrow : kIrrCols + nt * 16 + lrow;
					V B[2];
#pragma unroll
					for(int p = 0; p < P; ++p)
						B[p] = *reinterpret_cast<const V*>(sb + (p * (kIrrCols + kDepCols) + col) * kBStride + ks + lgrp * 8);
					acc[q][nt] = M::mma(A[0], B[0], acc[q][nt]);
					if(P == 2) {
						acc[q][nt] = M::mma(A[0], B[1], acc[q][nt]);
						acc[q][nt] = M::mma(A[1], B[0], acc[q][nt]);
					}
				}
			}
		}
	}
	__syncthreads();  // staging done: reuse LDS for the accumulators
#pragma unroll
	for(int q = 0; q < 4; ++q) {
		const int mt = wave + 4 * q;
		const bool irr = mt < 3;
#pragma unroll
		for(int nt = 0; nt < 3; ++nt) {
			if(!irr && nt == 2) break;
#pragma unroll
			for(int r = 0; r < 4; ++r) {
				const int row = 4 * lgrp + r;
				if(irr) s_out_irr[(mt * 16 + row) * kIrrCols + nt * 16 + lrow] = acc[q][nt][r];
				else s_out_dep[((mt - 3) * 16 + row) * kDepCols + nt * 16 + lrow] = acc[q][nt][r];
			}
		}
	}
	// out-of-range counts: the 16 stagers of a probe are 16 consecutive lanes of one wave
	for(int off = 8; off > 0; off >>= 1) my_oor += __shfl_xor(my_oor, off, 16);
	if((t & 15) == 0) s_oor[sp] = my_oor;
	if(t < kProbesPerBlock) {
		s_maxc[t] = 0;
		s_idx[t] = uint32_t(t) < np ? a.indices[slot0 + t] : 0u;
	}
	__syncthreads();
	blend_epilogue(a, s_idx, np, s_out_irr, kIrrCols, s_out_dep, kDepCols, s_maxc, false);
	__syncthreads();
	if(uint32_t(t) < np) update_state(a.states, s_idx[t], __uint_as_float(s_maxc[t]), s_oor[t], n);
}

template <uint32_t MODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_blend_mfma(BlendArgs a, uint32_t n_pad) {
	blend_mfma_tile<MODE>(a, blockIdx.x * kProbesPerBlock, n_pad);
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
	if constexpr(MODE == kBlendFP32) blend_fp32_tile<NP>(v.blend, first, v.n_pad);
	else blend_mfma_tile<MODE>(v.blend, first, v.n_pad);
	__syncthreads();
	if(threadIdx.x == 0) atomicMax(&b.ctl->done[vol], (unsigned long long)wall_clock64());
}

const void* batch_blend_kernel(uint32_t mode, uint32_t probes) {
	if(mode == kBlendBF16) return probes == kMfmaProbes ? reinterpret_cast<const void*>(&k_batch_blend<kBlendBF16, 16>) : nullptr;
	if(mode == kBlendFP16) return probes == kMfmaProbes ? reinterpret_cast<const void*>(&k_batch_blend<kBlendFP16, 16>) : nullptr;
	switch(probes) {
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

hipError_t launch_blend(const BlendArgs& a, uint32_t mode, hipStream_t s) {
	if(a.max_slots == 0) return hipSuccess;
	const uint32_t blocks = (a.max_slots + kProbesPerBlock - 1) / kProbesPerBlock;
	const uint32_t n_pad = (a.grid.n + kMfmaChunk - 1) & ~uint32_t(kMfmaChunk - 1);
	if(mode == kBlendBF16) hipLaunchKernelGGL(k_blend_mfma<kBlendBF16>, dim3(blocks), dim3(256), 0, s, a, n_pad);
	else if(mode == kBlendFP16) hipLaunchKernelGGL(k_blend_mfma<kBlendFP16>, dim3(blocks), dim3(256), 0, s, a, n_pad);
	else hipLaunchKernelGGL(k_blend_fp32, dim3(blocks), dim3(256), 0, s, a, n_pad);
	return hipGetLastError();
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
	return hipLaunchKernel(f, dim3(b.blend_item_count), dim3(256), args, 0, s);
}

}  // namespace ddgi
