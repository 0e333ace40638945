// Irradiance + depth blend with hysteresis (src/shaders/probesUpdate.glsl:49-119), border copy
// (src/shaders/probesCopyBorders.comp:21-246) and the probe state machine, on gfx950.
//
// The reference recomputes max(0, dot(texelDir, rayDir)) (and its pow) per probe, although the
// weights depend only on the frame's rotation.  Here one small kernel computes, once per frame,
//   Wt[i][t] = max(0, texelDir_t . rayDir_i)               t < 36   (irradiance texels)
//   Wt[i][36 + t] = max(0, texelDir_t . rayDir_i)^sharpness  t < 196  (depth texels)
//   inv_sum[t] = 1 / sum_i W (1 when the sum <= 1e-3, i.e. no normalisation)
// and the blend of a probe is a [232 x n] x [n x 3] product.  The border copy is fused into the
// epilogue: each interior texel also writes the border texels that mirror it, so the separate
// border pass and its re-read of the atlas disappear.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include "kernels.hpp"

namespace ddgi {

namespace {

constexpr int kBlendThreads = 256;  // one thread per texel (232 used)
constexpr int kProbesPerBlock = 16;
constexpr int kRayChunk = 32;

// octDecode(specializedNormalizeLocalTexelCoord(c)) in FP32 (irradiance.glsl:107-112, probesUpdate.glsl:33-42)
__device__ __forceinline__ void texel_direction(uint32_t t, float& x, float& y, float& z) {
	float u, v;
	if(t < kIrrTexels) {
		const int cx = int(t % 6u), cy = int(t / 6u);
		u = 0.33333f * (float(cx) - 2.5f);
		v = 0.33333f * (float(cy) - 2.5f);
	} else {
		const uint32_t d = t - kIrrTexels;
		const int cx = int(d % 14u), cy = int(d / 14u);
		u = 0.142857f * (float(cx) - 6.5f);
		v = 0.142857f * (float(cy) - 6.5f);
	}
	x = u;
	y = v;
	z = 1.0f - fabsf(u) - fabsf(v);
	if(z < 0.0f) {
		const float nx = (1.0f - fabsf(y)) * (x >= 0.0f ? 1.0f : -1.0f);
		const float ny = (1.0f - fabsf(x)) * (y >= 0.0f ? 1.0f : -1.0f);
		x = nx;
		y = ny;
	}
	const float l = sqrtf(x * x + y * y + z * z);
	x /= l;
	y /= l;
	z /= l;
}

// One block per texel: weights for every ray, their sum, and the 16-bit hi/lo split.
__global__ __launch_bounds__(256) void k_blend_weights(GridConst g, const float4* __restrict__ dirs, float* __restrict__ wt,
													   float* __restrict__ inv_sum, __hip_bfloat16* __restrict__ w_hi,
													   __hip_bfloat16* __restrict__ w_lo, __half* __restrict__ w_h,
													   uint32_t mode, uint32_t n_pad) {
	__shared__ float red[256];
	const uint32_t t = blockIdx.x;
	float tx, ty, tz;
	texel_direction(t, tx, ty, tz);
	const bool depth = t >= kIrrTexels;
	float sum = 0.0f, sum16 = 0.0f;
	for(uint32_t i = threadIdx.x; i < n_pad; i += 256) {
		float w = 0.0f;
		if(i < g.n) {
			const float4 d = dirs[i];
			w = fmaxf(0.0f, tx * d.x + ty * d.y + tz * d.z);
			if(depth) w = powf(w, g.sharpness);
			sum += w;
			wt[size_t(i) * kBlendTexels + t] = w;
		}
		if(mode == kBlendBF16) {
			// compensated split: w = hi + lo with both parts in bf16; normalise by the sum of what
			// the MFMA actually multiplies
			const __hip_bfloat16 hi = __float2bfloat16(w);
			const float r = w - __bfloat162float(hi);
			const __hip_bfloat16 lo = __float2bfloat16(r);
			w_hi[size_t(t) * n_pad + i] = hi;
			w_lo[size_t(t) * n_pad + i] = lo;
			sum16 += __bfloat162float(hi) + __bfloat162float(lo);
		} else if(mode == kBlendFP16) {
			const __half h = __float2half(w);
			w_h[size_t(t) * n_pad + i] = h;
			sum16 += __half2float(h);
		}
	}
	const float s = (mode == kBlendFP32) ? sum : sum16;
	red[threadIdx.x] = s;
	__syncthreads();
	for(int o = 128; o > 0; o >>= 1) {
		if(int(threadIdx.x) < o) red[threadIdx.x] += red[threadIdx.x + o];
		__syncthreads();
	}
	if(threadIdx.x == 0) inv_sum[t] = red[0] > 1e-3f ? 1.0f / red[0] : 1.0f;
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

// Epilogue shared by the FP32 and MFMA blends: normalise, hysteresis against the exposed atlas,
// store interior + mirrored border texels into the next atlas, collect maxChange.
__device__ __forceinline__ void store_texel(const BlendArgs& a, uint32_t t, uint32_t probe, const float acc[3],
											float inv, float* max_change) {
	const float h = a.grid.hysteresis;
	if(t < kIrrTexels) {
		const int x = int(t % 6u) + 1, y = int(t / 6u) + 1;
		const float4 prev = a.irr_prev[size_t(probe) * 64 + y * 8 + x];
		const float r0 = acc[0] * inv, r1 = acc[1] * inv, r2 = acc[2] * inv;
		*max_change = fmaxf(fmaxf(fabsf(r0 - prev.x), fabsf(r1 - prev.y)), fabsf(r2 - prev.z));
		const float4 v = make_float4(r0 * (1.0f - h) + prev.x * h, r1 * (1.0f - h) + prev.y * h,
									 r2 * (1.0f - h) + prev.z * h, 1.0f);
		float4* tile = a.irr_out + size_t(probe) * 64;
		tile[y * 8 + x] = v;
		int dx[3], dy[3];
		const int nb = border_targets<8>(x, y, dx, dy);
		for(int k = 0; k < nb; ++k) tile[dy[k] * 8 + dx[k]] = v;
	} else {
		const uint32_t d = t - kIrrTexels;
		const int x = int(d % 14u) + 1, y = int(d / 14u) + 1;
		const float2 prev = a.dep_prev[size_t(probe) * 256 + y * 16 + x];
		const float r0 = acc[0] * inv, r1 = acc[1] * inv;
		const float2 v = make_float2(r0 * (1.0f - h) + prev.x * h, r1 * (1.0f - h) + prev.y * h);
		float2* tile = a.dep_out + size_t(probe) * 256;
		tile[y * 16 + x] = v;
		int dx[3], dy[3];
		const int nb = border_targets<16>(x, y, dx, dy);
		for(int k = 0; k < nb; ++k) tile[dy[k] * 16 + dx[k]] = v;
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

// FP32 VALU blend: a block owns 16 probe slots; thread t owns texel t for all 16 probes and
// accumulates 3 channels (rgb for irradiance, (d, d^2, 0) for depth) over ray chunks staged in LDS.
__global__ __launch_bounds__(kBlendThreads) void k_blend_fp32(BlendArgs a) {
	__shared__ float4 s_irr[kProbesPerBlock][kRayChunk];
	__shared__ float4 s_dep[kProbesPerBlock][kRayChunk];
	__shared__ uint32_t s_oor[kProbesPerBlock];
	__shared__ uint32_t s_maxc[kProbesPerBlock];
	const uint32_t count = *a.count;
	const uint32_t slot0 = blockIdx.x * kProbesPerBlock;
	if(slot0 >= count) return;
	const uint32_t np = min(uint32_t(kProbesPerBlock), count - slot0);
	const uint32_t n = a.grid.n;
	const uint32_t t = threadIdx.x;
	const float cell_len = a.grid.cell_len;
	if(t < kProbesPerBlock) {
		s_oor[t] = 0;
		s_maxc[t] = 0;
	}
	float acc[kProbesPerBlock][3];
#pragma unroll
	for(int p = 0; p < kProbesPerBlock; ++p) acc[p][0] = acc[p][1] = acc[p][2] = 0.0f;
	const bool is_depth = t >= kIrrTexels;
	const bool active = t < kBlendTexels;
	for(uint32_t i0 = 0; i0 < n; i0 += kRayChunk) {
		__syncthreads();
		// stage 16 probes x 32 rays
		for(uint32_t e = t; e < kProbesPerBlock * kRayChunk; e += kBlendThreads) {
			const uint32_t p = e / kRayChunk, i = e % kRayChunk;
			float4 r = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
			float4 dd = r;
			if(p < np && i0 + i < n) {
				r = a.rays[size_t(slot0 + p) * n + i0 + i];
				float dc = fminf(cell_len, r.w);
				if(dc < 0.0f) dc = cell_len;
				dd = make_float4(dc, dc * dc, 0.0f, 0.0f);
				if(r.w < 0.0f || r.w > cell_len) atomicAdd(&s_oor[p], 1u);
				r.w = 0.0f;
			}
			s_irr[p][i] = r;
			s_dep[p][i] = dd;
		}
		__syncthreads();
		if(active) {
			const uint32_t iend = min(uint32_t(kRayChunk), n - i0);
			const float4(*src)[kRayChunk] = is_depth ? s_dep : s_irr;
			for(uint32_t i = 0; i < iend; ++i) {
				const float w = a.wt[size_t(i0 + i) * kBlendTexels + t];
#pragma unroll
				for(int p = 0; p < kProbesPerBlock; ++p) {
					const float4 x = src[p][i];
					acc[p][0] += w * x.x;
					acc[p][1] += w * x.y;
					acc[p][2] += w * x.z;
				}
			}
		}
	}
	if(active) {
		const float inv = a.inv_sum[t];
#pragma unroll
		for(int p = 0; p < kProbesPerBlock; ++p) {
			if(uint32_t(p) < np) {
				float mc = 0.0f;
				store_texel(a, t, a.indices[slot0 + p], acc[p], inv, &mc);
				if(!is_depth) atomicMax(&s_maxc[p], __float_as_uint(mc));
			}
		}
	}
	__syncthreads();
	if(t < np) update_state(a.states, a.indices[slot0 + t], __uint_as_float(s_maxc[t]), s_oor[t], n);
}

}  // namespace

hipError_t launch_blend_weights(const GridConst& g, const float4* dirs, float* wt, float* inv_sum, void* w_hi16,
								void* w_lo16, uint32_t mode, hipStream_t s) {
	const uint32_t n_pad = (g.n + 31u) & ~31u;
	hipLaunchKernelGGL(k_blend_weights, dim3(kBlendTexels), dim3(256), 0, s, g, dirs, wt, inv_sum,
					   static_cast<__hip_bfloat16*>(w_hi16), static_cast<__hip_bfloat16*>(w_lo16),
					   static_cast<__half*>(w_hi16), mode, n_pad);
	return hipGetLastError();
}

hipError_t launch_blend(const BlendArgs& a, uint32_t mode, hipStream_t s) {
	if(a.max_slots == 0) return hipSuccess;
	const uint32_t blocks = (a.max_slots + kProbesPerBlock - 1) / kProbesPerBlock;
	(void)mode;
	hipLaunchKernelGGL(k_blend_fp32, dim3(blocks), dim3(kBlendThreads), 0, s, a);
	return hipGetLastError();
}

}  // namespace ddgi
