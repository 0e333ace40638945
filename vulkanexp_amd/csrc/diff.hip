// Differentiable probe update: gradients of a shaded-G-buffer loss with respect to per-material
// albedo and the sun radiance, through probe sampling (src/shaders/irradiance.glsl:145-237), the
// border copy and the ray-to-texel irradiance blend (src/shaders/probesUpdate.glsl:70-103).
//
// Per-ray radiance is affine in the parameters (factors from the shading kernel):
//   rad = c0 + A (.) a[m] + (B (.) a[m] + C) (.) sun
// Forward:  rad -> irradiance blend (FP32 kernel, irradiance only) -> sampleProbes at every
//           G-buffer texel -> shade = out (.) a[g] * 0.996 (FinalGather.frag:62-72) -> MSE.
// Backward: hand-written adjoint.  All reductions are deterministic: partial sums are 64-bit
//           fixed-point integers (order-independent) combined in a fixed order.  The sampling
//           weights (visibility, trilinear) and the blend weights are recomputed, not stored.
// G-buffer texels are binned by probe cell at prepare time, so a work item (a chunk of one cell's
// texels) scatters into only the 8 probe tiles of that cell's cage, accumulated in LDS.
#include "device_common.hpp"
#include "diff.hpp"

namespace ddgi {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ bool cell_of(const GridConst& g, const float* pos, uint64_t x, uint32_t& cell) {
	const float ax = fabsf(g.cell[0]), ay = fabsf(g.cell[1]), az = fabsf(g.cell[2]);
	const float gx = (pos[x * 3] - g.emin[0]) / ax, gy = (pos[x * 3 + 1] - g.emin[1]) / ay, gz = (pos[x * 3 + 2] - g.emin[2]) / az;
	if(gx < 0.0f || gy < 0.0f || gz < 0.0f) return false;
	const int f0 = int(gx), f1 = int(gy), f2 = int(gz);
	if(f0 > g.res[0] - 1 || f1 > g.res[1] - 1 || f2 > g.res[2] - 1) return false;
	cell = uint32_t(f0) + uint32_t(g.res[0]) * (uint32_t(f1) + uint32_t(g.res[1]) * uint32_t(f2));
	return true;
}

__device__ __forceinline__ Vec3 load3(const float* p, uint64_t i) { return v3(p[i * 3], p[i * 3 + 1], p[i * 3 + 2]); }

// ---- prepare: bin texels by cell ----
__global__ void k_cells(DiffArgs d, uint32_t* __restrict__ cell_id, uint32_t* __restrict__ cell_count) {
	const uint64_t x = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if(x >= d.N) return;
	uint32_t c;
	if(cell_of(d.grid, d.pos, x, c)) {
		cell_id[x] = c;
		atomicAdd(&cell_count[c], 1u);
	} else {
		cell_id[x] = 0xFFFFFFFFu;
	}
}

// single block: exclusive scans over cells -> pixel offsets and work items of <= kDiffChunk texels
__global__ __launch_bounds__(1024) void k_items(DiffArgs d, const uint32_t* __restrict__ cell_count,
												uint32_t* __restrict__ cell_start, uint32_t* __restrict__ cell_item_start,
												uint32_t* __restrict__ item_cell, uint32_t* __restrict__ item_begin,
												uint32_t* __restrict__ item_count, uint32_t* __restrict__ n_items) {
	__shared__ uint32_t s_pix, s_items;
	if(threadIdx.x == 0) {
		s_pix = 0;
		s_items = 0;
	}
	__syncthreads();
	const uint32_t cells = d.grid.probe_count;
	for(uint32_t base = 0; base < cells; base += 1024) {
		// serial scan by thread 0 over this chunk (cells <= 2^24, a few thousand here)
		if(threadIdx.x == 0) {
			for(uint32_t c = base; c < min(cells, base + 1024u); ++c) {
				const uint32_t cnt = cell_count[c];
				cell_start[c] = s_pix;
				cell_item_start[c] = s_items;
				const uint32_t ni = (cnt + kDiffChunk - 1) / kDiffChunk;
				for(uint32_t k = 0; k < ni; ++k) {
					item_cell[s_items + k] = c;
					item_begin[s_items + k] = s_pix + k * kDiffChunk;
					item_count[s_items + k] = min(kDiffChunk, cnt - k * kDiffChunk);
				}
				s_pix += cnt;
				s_items += ni;
			}
		}
		__syncthreads();
	}
	if(threadIdx.x == 0) {
		cell_start[cells] = s_pix;
		cell_item_start[cells] = s_items;
		*n_items = s_items;
	}
}

__global__ void k_scatter(DiffArgs d, const uint32_t* __restrict__ cell_id, const uint32_t* __restrict__ cell_start,
						  uint32_t* __restrict__ cursor, uint32_t* __restrict__ sorted) {
	const uint64_t x = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if(x >= d.N) return;
	const uint32_t c = cell_id[x];
	if(c == 0xFFFFFFFFu) return;
	sorted[cell_start[c] + atomicAdd(&cursor[c], 1u)] = uint32_t(x);
}

// ---- forward ----
__global__ void k_radiance(DiffArgs d) {
	const uint64_t r = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if(r >= d.rays) return;
	const float4* f = d.factors + r * 4;
	const float4 c0 = f[0];
	const uint32_t m = __float_as_uint(c0.w);
	float4 out = make_float4(c0.x, c0.y, c0.z, 0.0f);
	if(m != 0xFFFFFFFFu) {
		const float4 A = f[1], B = f[2], C = f[3];
		const float* a = d.albedo + size_t(m) * 3;
		out.x += A.x * a[0] + (B.x * a[0] + C.x) * d.sun[0];
		out.y += A.y * a[1] + (B.y * a[1] + C.y) * d.sun[1];
		out.z += A.z * a[2] + (B.z * a[2] + C.z) * d.sun[2];
	}
	d.rad[r] = out;
}

__global__ __launch_bounds__(kBlock) void k_gather_fwd(DiffArgs d) {
	__shared__ float red[kBlock];
	const uint64_t x = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
	float l = 0.0f;
	if(x < d.N) {
		SampleCollect c;
		sample_collect(d.grid, d.states, d.dep, load3(d.pos, x), load3(d.nrm, x), load3(d.tocam, x), c);
		const Vec3 o = sample_eval(c, d.irr);
		const float* a = d.albedo + size_t(d.gmat[x]) * 3;
		const float s[3] = {o.x * a[0] * 0.996f, o.y * a[1] * 0.996f, o.z * a[2] * 0.996f};
		d.out[x * 3] = o.x;
		d.out[x * 3 + 1] = o.y;
		d.out[x * 3 + 2] = o.z;
		for(int ch = 0; ch < 3; ++ch) {
			d.shade[x * 3 + ch] = s[ch];
			const float e = s[ch] - d.target[x * 3 + ch];
			l += e * e;
		}
	}
	red[threadIdx.x] = l;
	__syncthreads();
	for(int o = kBlock / 2; o > 0; o >>= 1) {
		if(int(threadIdx.x) < o) red[threadIdx.x] += red[threadIdx.x + o];
		__syncthreads();
	}
	if(threadIdx.x == 0) d.loss_part[blockIdx.x] = red[0];
}

__global__ void k_loss(DiffArgs d, uint32_t parts) {
	if(threadIdx.x != 0 || blockIdx.x != 0) return;
	double s = 0.0;
	for(uint32_t i = 0; i < parts; ++i) s += double(d.loss_part[i]);
	d.loss[0] = float(s / double(3ull * d.N));
}

// ---- backward ----
// One block per work item: the item's texels share one cage of 8 probes.
__global__ __launch_bounds__(kBlock) void k_gather_bwd(DiffArgs d) {
	extern __shared__ unsigned long long lds[];
	unsigned long long* tile = lds;                  // [8][64][3]
	unsigned long long* mat = lds + 8 * 192;         // [M][3]
	const uint32_t item = blockIdx.x;
	if(item >= *d.n_items) return;
	for(uint32_t k = threadIdx.x; k < 8 * 192 + d.M * 3; k += kBlock) lds[k] = 0ull;
	__syncthreads();
	const uint32_t begin = d.item_begin[item], cnt = d.item_count[item];
	const float inv3n = 2.0f / float(3ull * d.N);
	for(uint32_t k = threadIdx.x; k < cnt; k += kBlock) {
		const uint64_t x = d.sorted[begin + k];
		SampleCollect c;
		sample_collect(d.grid, d.states, d.dep, load3(d.pos, x), load3(d.nrm, x), load3(d.tocam, x), c);
		const uint32_t m = d.gmat[x];
		const float* a = d.albedo + size_t(m) * 3;
		float go[3];
		for(int ch = 0; ch < 3; ++ch) {
			const float o = d.out[x * 3 + ch];
			const float gs = (o * a[ch] * 0.996f - d.target[x * 3 + ch]) * inv3n;
			atomicAdd(&mat[m * 3 + ch], (unsigned long long)to_fixed(gs * o * 0.996f));
			go[ch] = gs * a[ch] * 0.996f;
		}
		sample_adjoint(c, d.irr, v3(go[0], go[1], go[2]), tile);
	}
	__syncthreads();
	long long* pt = d.part_tile + size_t(item) * 8 * 192;
	for(uint32_t k = threadIdx.x; k < 8 * 192; k += kBlock) pt[k] = (long long)tile[k];
	long long* pm = d.part_mat_pix + size_t(item) * d.M * 3;
	for(uint32_t k = threadIdx.x; k < d.M * 3; k += kBlock) pm[k] = (long long)mat[k];
}

// One block per probe: sum the partial tiles of the (up to 8) cells whose cage contains the
// probe, fold the border gradients onto the interior texels they copy, scale to float.
__global__ __launch_bounds__(192) void k_probe_grad(DiffArgs d) {
	__shared__ long long g[64 * 3];
	const uint32_t p = blockIdx.x;
	const uint32_t rx = uint32_t(d.grid.res[0]), ry = uint32_t(d.grid.res[1]);
	const int px = int(p % rx), py = int((p % (rx * ry)) / rx), pz = int(p / (rx * ry));
	const uint32_t e = threadIdx.x;  // texel * 3 + channel
	long long s = 0;
	for(int o = 0; o < 8; ++o) {
		const int cx = px - (o & 1), cy = py - ((o >> 1) & 1), cz = pz - ((o >> 2) & 1);
		if(cx < 0 || cy < 0 || cz < 0) continue;
		const uint32_t cell = uint32_t(cx) + rx * (uint32_t(cy) + ry * uint32_t(cz));
		for(uint32_t it = d.cell_item_start[cell]; it < d.cell_item_start[cell + 1]; ++it)
			s += d.part_tile[(size_t(it) * 8 + o) * 192 + e];
	}
	g[e] = s;
	__syncthreads();
	// border fold (adjoint of probesCopyBorders: dst <- src)
	if(e < 3) {
		const int R = 8;
		auto fold = [&](int dx, int dy, int sx, int sy) { g[(sy * R + sx) * 3 + e] += g[(dy * R + dx) * 3 + e]; };
		fold(0, 0, R - 2, R - 2);
		fold(0, R - 1, R - 2, 1);
		fold(R - 1, 0, 1, R - 2);
		fold(R - 1, R - 1, 1, 1);
		for(int j = 1; j + 1 < R; ++j) {
			fold(0, j, 1, R - 1 - j);
			fold(R - 1, j, R - 2, R - 1 - j);
			fold(j, 0, R - 1 - j, 1);
			fold(j, R - 1, R - 1 - j, R - 2);
		}
	}
	__syncthreads();
	if(e < 36 * 3) {
		const uint32_t t = e / 3, ch = e % 3;
		const int x = int(t % 6) + 1, y = int(t / 6) + 1;
		d.g_T[(size_t(p) * 36 + t) * 3 + ch] = float(double(g[(y * 8 + x) * 3 + ch]) / kFixScale);
	}
}

// One block per probe slot: d loss / d rad_i = (1 - h) sum_t w_ti inv_t dT_t with the weights
// recomputed from the ray directions, then the radiance adjoint into per-material / sun partials.
__global__ __launch_bounds__(kBlock) void k_blend_bwd(DiffArgs d) {
	extern __shared__ unsigned long long lds[];
	unsigned long long* mat = lds;                             // [M][3] then sun[3]
	__shared__ float s_g[36 * 3], s_dir[36 * 3];
	const uint32_t slot = blockIdx.x;
	const uint32_t probe = d.indices[slot];
	for(uint32_t k = threadIdx.x; k < d.M * 3 + 3; k += kBlock) lds[k] = 0ull;
	const float h = d.grid.hysteresis;
	if(threadIdx.x < 36) {
		float tx, ty, tz;
		texel_direction(threadIdx.x, tx, ty, tz);
		s_dir[threadIdx.x * 3] = tx;
		s_dir[threadIdx.x * 3 + 1] = ty;
		s_dir[threadIdx.x * 3 + 2] = tz;
	}
	if(threadIdx.x < 108) {
		const uint32_t t = threadIdx.x / 3;
		s_g[threadIdx.x] = d.g_T[size_t(probe) * 108 + threadIdx.x] * (1.0f - h) * d.inv_sum[t];
	}
	__syncthreads();
	const uint32_t n = d.grid.n;
	for(uint32_t j = threadIdx.x; j < n; j += kBlock) {
		const float4 dr = d.dirs[d.perm[j]];
		float gr[3] = {0.0f, 0.0f, 0.0f};
		for(int t = 0; t < 36; ++t) {
			const float w = fmaxf(0.0f, s_dir[t * 3] * dr.x + s_dir[t * 3 + 1] * dr.y + s_dir[t * 3 + 2] * dr.z);
			gr[0] += w * s_g[t * 3];
			gr[1] += w * s_g[t * 3 + 1];
			gr[2] += w * s_g[t * 3 + 2];
		}
		const float4* f = d.factors + (size_t(slot) * n + j) * 4;
		const uint32_t m = __float_as_uint(f[0].w);
		if(m == 0xFFFFFFFFu) continue;
		const float4 A = f[1], B = f[2], C = f[3];
		const float* a = d.albedo + size_t(m) * 3;
		const float Av[3] = {A.x, A.y, A.z}, Bv[3] = {B.x, B.y, B.z}, Cv[3] = {C.x, C.y, C.z};
		for(int ch = 0; ch < 3; ++ch) {
			atomicAdd(&mat[m * 3 + ch], (unsigned long long)to_fixed(gr[ch] * (Av[ch] + Bv[ch] * d.sun[ch])));
			atomicAdd(&mat[d.M * 3 + ch], (unsigned long long)to_fixed(gr[ch] * (Bv[ch] * a[ch] + Cv[ch])));
		}
	}
	__syncthreads();
	long long* pr = d.part_mat_ray + size_t(slot) * (d.M * 3 + 3);
	for(uint32_t k = threadIdx.x; k < d.M * 3 + 3; k += kBlock) pr[k] = (long long)mat[k];
}

__global__ void k_grad_final(DiffArgs d) {
	const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
	if(k >= d.M * 3 + 3) return;
	long long s = 0;
	if(k < d.M * 3) {
		const uint32_t items = *d.n_items;
		for(uint32_t it = 0; it < items; ++it) s += d.part_mat_pix[size_t(it) * d.M * 3 + k];
	}
	for(uint32_t sl = 0; sl < d.slots; ++sl) s += d.part_mat_ray[size_t(sl) * (d.M * 3 + 3) + k];
	const float v = float(double(s) / kFixScale) * d.grad_loss[0];
	if(k < d.M * 3) d.g_albedo[k] = v;
	else d.g_sun[k - d.M * 3] = v;
}

inline uint32_t nblocks(uint64_t n, uint32_t b) { return uint32_t((n + b - 1) / b); }

}  // namespace

hipError_t diff_bin(const DiffArgs& d, uint32_t* cell_id, uint32_t* cell_count, uint32_t* cell_start, uint32_t* cursor,
					uint32_t* cell_item_start, uint32_t* item_cell, uint32_t* item_begin, uint32_t* item_count,
					uint32_t* n_items, uint32_t* sorted, hipStream_t s) {
	const uint32_t cells = d.grid.probe_count;
	hipError_t e;
	if((e = hipMemsetAsync(cell_count, 0, cells * 4, s)) != hipSuccess) return e;
	if((e = hipMemsetAsync(cursor, 0, cells * 4, s)) != hipSuccess) return e;
	if(d.N) hipLaunchKernelGGL(k_cells, dim3(nblocks(d.N, 256)), dim3(256), 0, s, d, cell_id, cell_count);
	hipLaunchKernelGGL(k_items, dim3(1), dim3(1024), 0, s, d, cell_count, cell_start, cell_item_start, item_cell, item_begin,
					   item_count, n_items);
	if(d.N) hipLaunchKernelGGL(k_scatter, dim3(nblocks(d.N, 256)), dim3(256), 0, s, d, cell_id, cell_start, cursor, sorted);
	return hipGetLastError();
}

hipError_t diff_radiance(const DiffArgs& d, hipStream_t s) {
	hipLaunchKernelGGL(k_radiance, dim3(nblocks(d.rays, 256)), dim3(256), 0, s, d);
	return hipGetLastError();
}

hipError_t diff_gather_forward(const DiffArgs& d, hipStream_t s) {
	const uint32_t parts = nblocks(d.N, kBlock);
	if(d.N) hipLaunchKernelGGL(k_gather_fwd, dim3(parts), dim3(kBlock), 0, s, d);
	hipLaunchKernelGGL(k_loss, dim3(1), dim3(64), 0, s, d, parts);
	return hipGetLastError();
}

hipError_t diff_backward(const DiffArgs& d, hipStream_t s) {
	const size_t lds1 = (8 * 192 + size_t(d.M) * 3) * sizeof(unsigned long long);
	hipLaunchKernelGGL(k_gather_bwd, dim3(d.max_items), dim3(kBlock), lds1, s, d);
	hipLaunchKernelGGL(k_probe_grad, dim3(d.grid.probe_count), dim3(192), 0, s, d);
	const size_t lds2 = (size_t(d.M) * 3 + 3) * sizeof(unsigned long long);
	hipLaunchKernelGGL(k_blend_bwd, dim3(d.slots), dim3(kBlock), lds2, s, d);
	hipLaunchKernelGGL(k_grad_final, dim3(nblocks(d.M * 3 + 3, 256)), dim3(256), 0, s, d);
	return hipGetLastError();
}

}  // namespace ddgi
