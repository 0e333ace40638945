// Probe ray tracing on gfx950: ray generation (src/shaders/traceProbes.rgen:29-49), BVH traversal
// replacing the RT cores, closest-hit shading (src/shaders/closesthit.glsl:133-289, NO_REFLECTION),
// shadow rays (closesthit.glsl:252-281, shadow.rmiss), the constant-sky miss, probe-state init
// (src/shaders/probesInit.rgen:31-64), device-side probe selection
// (src/IrradianceProbes.cpp:396-424) and standalone sampleProbes queries.
//
// Built with -ffp-contract=off (see device_common.hpp).
#include <algorithm>
#include <type_traits>
#include <vector>

#include "device_common.hpp"

namespace ddgi {

namespace {

__device__ __forceinline__ void probe_origin(const GridConst& g, uint32_t linear, float o[3]) {
	const uint32_t rx = uint32_t(g.res[0]), ry = uint32_t(g.res[1]);
	const uint32_t gx = linear % rx, gy = (linear % (rx * ry)) / rx, gz = linear / (rx * ry);
	o[0] = float(gx) * g.cell[0] + g.emin[0];
	o[1] = float(gy) * g.cell[1] + g.emin[1];
	o[2] = float(gz) * g.cell[2] + g.emin[2];
}

__device__ __forceinline__ float4 pack_hit(const HitResult& h) {
	if(h.slot == 0xFFFFFFFFu) return make_float4(-1.0f, __uint_as_float(0xFFFFFFFFu), 0.0f, 0.0f);
	const bool back = (h.det < 0.0f) != ((h.meta & 0x100u) != 0u);
	return make_float4(h.t, __uint_as_float(h.slot | (back ? 0x80000000u : 0u)), h.V / h.det, h.W / h.det);
}

// ---- per-frame world directions: dirs[i] = R * fib[i] (FP32, no FMA) ----
__device__ __forceinline__ float4 frame_dir(const Rot& R, float4 f) {
	float d[3];
	for(int k = 0; k < 3; ++k) d[k] = R.r[k * 3 + 0] * f.x + R.r[k * 3 + 1] * f.y + R.r[k * 3 + 2] * f.z;
	return make_float4(d[0], d[1], d[2], 0.0f);
}

// `zero` (optional): two counters the next kernels start from (the shade- and miss-list lengths),
// cleared here so that no separate memset is needed.
__global__ void k_clear(uint32_t* __restrict__ words, uint32_t n) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i < n) words[i] = 0u;
}

__global__ void k_frame_dirs(Rot R, const float4* __restrict__ fib, float4* __restrict__ dirs, uint32_t count,
							 uint32_t* zero) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if(zero && i < kShadeSegments) zero[kCounterStride * i] = 0u;
	if(zero && i == 0) zero[kMissCountSlot] = 0u;
	if(i >= count) return;
	dirs[i] = frame_dir(R, fib[i]);
}

// Thread -> ray order of a frame, in one workgroup of kMaxRays threads (thread i holds ray i's world
// direction y): rays sorted by the elevation of their world direction, ties by ray index, so that
// the wavefronts of a probe's rays take one band of elevations each.  A band's rays have more
// alike lengths than a lune from pole to pole has where the scene is wider than it is tall (the
// rays into the floor and the ceiling end early); the sums of the blend follow the order through
// its weight rows, and the ray records are stored in it.
// s_key: kMaxRays floats, 16-byte aligned; entries past n hold +infinity, which no finite key
// counts.  Reads go four keys at a time to the same address in every lane (broadcast).
__device__ __forceinline__ void frame_order_block(float y, uint32_t n, float* s_key, uint16_t* perm) {
	const uint32_t i = threadIdx.x;
	s_key[i] = i < n ? y : __int_as_float(0x7f800000);
	__syncthreads();
	if(i < n) {
		uint32_t rank = 0;
		const float4* k4 = reinterpret_cast<const float4*>(s_key);
#pragma unroll 4
		for(uint32_t j = 0; j < (n + 3u) / 4u; ++j) {
			const float4 k = k4[j];
			const uint32_t j0 = 4u * j;
			rank += (k.x < y || (k.x == y && j0 < i)) ? 1u : 0u;
			rank += (k.y < y || (k.y == y && j0 + 1u < i)) ? 1u : 0u;
			rank += (k.z < y || (k.z == y && j0 + 2u < i)) ? 1u : 0u;
			rank += (k.w < y || (k.w == y && j0 + 3u < i)) ? 1u : 0u;
		}
		perm[rank] = uint16_t(i);
	}
}

// The update's frame directions and thread -> ray order in one workgroup of kMaxRays threads
// (thread i: ray i); `zero` as in k_frame_dirs.
__global__ __launch_bounds__(kMaxRays) void k_frame_dirs_order(Rot R, const float4* __restrict__ fib,
															   float4* __restrict__ dirs, uint32_t n, uint16_t* perm,
															   uint32_t* zero) {
	__shared__ __attribute__((aligned(16))) float s_key[kMaxRays];
	const uint32_t i = threadIdx.x;
	if(zero && i < kShadeSegments) zero[kCounterStride * i] = 0u;
	if(zero && i == 0) zero[kMissCountSlot] = 0u;
	float y = 0.0f;
	if(i < n) {
		const float4 d = frame_dir(R, fib[i]);
		dirs[i] = d;
		y = d.y;
	}
	frame_order_block(y, n, s_key, perm);
}

// Ray records of rays that need no shading: misses.rmiss (sky radiance, depth -1) and back faces
// (closesthit.glsl:137-141: black, depth 0.8 t); with `factors` also their (constant) factor rows.
__device__ __forceinline__ void write_unshaded(const TraceFrameArgs& a, size_t ri, bool miss, float t) {
	const float4 rec = miss ? make_float4(a.light.sky[0], a.light.sky[1], a.light.sky[2], -1.0f)
							: make_float4(0.0f, 0.0f, 0.0f, t * 0.8f);
	a.rays[ri] = rec;
	if(a.factors) {
		float4* f = a.factors + ri * 4;
		f[0] = make_float4(rec.x, rec.y, rec.z, __uint_as_float(0xFFFFFFFFu));
		f[1] = f[2] = f[3] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
	}
}

// Shade-list ordering key: the Morton code of the half-cell of the probe grid that holds p
// (half-cell coordinates clamped to the grid and shifted right by `shift`; a NaN coordinate counts
// as 0).  Half cells order the hits more finely than the cage cells they share.
__device__ __forceinline__ uint32_t cell_key(const GridConst& g, float px, float py, float pz, uint32_t shift) {
	const float p[3] = {px, py, pz};
	uint32_t key = 0;
#pragma unroll
	for(int ax = 0; ax < 3; ++ax) {
		float f = floorf((p[ax] - g.emin[ax]) / g.cell[ax] * 2.0f);
		f = fminf(fmaxf(f, 0.0f), float(2 * g.res[ax] - 1));
		const uint32_t c = uint32_t(f) >> shift;
#pragma unroll
		for(int b = 0; b < 6; ++b) key |= ((c >> b) & 1u) << (3 * b + ax);
	}
	return key;
}

// ---- primary probe rays: closest hit, mask STATIC|DYNAMIC, tmin 0.01, tmax = |max - min| ----
// Result of the ray at position ri (= slot * n + k, thread order): its hit record, and either its
// final ray record (miss, back face) or a shade-list entry (front-face hit).
__device__ __forceinline__ void finish_primary_ray(const TraceFrameArgs& a, uint32_t ri, const HitResult& h,
												   uint32_t key) {
	const float4 packed = pack_hit(h);
	a.hits[ri] = packed;
	const uint32_t code = __float_as_uint(packed.y);
	const bool front = code != 0xFFFFFFFFu && !(code & 0x80000000u);
	// front-face hits are queued for shading, in this workgroup's segment of the list
	const uint32_t seg = a.seg_cap ? blockIdx.x % kShadeSegments : 0u;
	const size_t base = size_t(seg) * a.seg_cap;
	const uint32_t pos = wave_append(a.shade_count + kCounterStride * seg, front);
	if(front) {
		a.shade_list[base + pos] = ri;
		if(a.shade_keys) a.shade_keys[base + pos] = key;
	}
	const bool miss = code == 0xFFFFFFFFu;
	if(a.light.sky_model == DDGI_SKY_ATMOSPHERE) {  // misses are queued for the sky kernel
		const uint32_t mpos = wave_append(a.shade_count + kMissCountSlot, miss);
		if(miss) a.miss_list[mpos] = ri;
		else if(!front) write_unshaded(a, ri, false, h.t);
	} else if(!front) {
		write_unshaded(a, ri, miss, h.t);
	}
}

// Traversal thread `gtid` = slot * n + j traces ray perm[j] of the probe in `slot`; `stack` is the
// thread's column of an LDS stack with `stride` columns.
template <int STACK, bool COUNT>
__device__ __forceinline__ void trace_primary_ray(const TraceFrameArgs& a, uint32_t gtid, uint32_t* stack, int stride) {
	const uint32_t n = a.grid.n;
	const uint32_t slot = gtid / n, j = gtid - slot * n;
	if(slot >= *a.count) return;
	const uint32_t i = a.perm[j];
	float o[3];
	probe_origin(a.grid, a.indices[slot], o);
	const float4 d = a.dirs[i];
	RayPrep r;
	ray_prepare(o[0], o[1], o[2], d.x, d.y, d.z, r);
	uint32_t nn = 0, nt = 0;
	const HitResult h = traverse<false, STACK, COUNT>(a.scene.nodes, a.scene.tris, r, 0.01f, a.grid.tmax_primary,
													  DDGI_MASK_STATIC | DDGI_MASK_DYNAMIC, stack, stride, nn, nt);
	const uint32_t key = a.shade_keys ? cell_key(a.grid, o[0] + h.t * d.x, o[1] + h.t * d.y, o[2] + h.t * d.z, a.key_shift)
									  : 0u;
	finish_primary_ray(a, slot * n + j, h, key);  // thread order: coalesced stores
	if(COUNT) {
		atomicAdd(&a.visit_counters[0], (unsigned long long)nn);
		atomicAdd(&a.visit_counters[1], (unsigned long long)nt);
	}
}

template <int STACK, bool COUNT>
// seven waves per SIMD (72 VGPRs): the occupancy at which the traversal ran fastest, 5 to 8 measured
__global__ __launch_bounds__(kPrimaryBlock) __attribute__((amdgpu_waves_per_eu(7, 7))) void k_trace_primary(TraceFrameArgs a) {
	__shared__ uint32_t stack[STACK * kPrimaryBlock];
	trace_primary_ray<STACK, COUNT>(a, xcd_remap(blockIdx.x, gridDim.x) * kPrimaryBlock + threadIdx.x,
									stack + threadIdx.x, kPrimaryBlock);
}

// ---- closest-hit shading + shadow ray (closesthit.glsl:143-289) ----
// Shading of the front-face hit (t, BVH slot, barycentrics b1 b2) of the ray o + t d at recursion
// depth >= 1, where the specular term reads the probes along the reflected direction: emissive +
// specular * sampleProbes(reflect) + diffuse * sampleProbes(normal) + the sun through an any-hit
// shadow ray (mask 0xFF, tmin 0.1, tmax 10000), faded when the sun is below the horizon.  The
// parts are returned too (the differentiable update decomposes them).
struct SurfaceShade {
	Vec3 color;
	Vec3 em_spec;   // emissive + specular * reflection
	Vec3 indirect;  // sampleProbes along the normal
	Vec3 normal, to_cam;
	float metal, rough, fade;
	uint32_t mat;
	bool lit;
};

template <int STACK, bool COUNT>
__device__ __forceinline__ SurfaceShade shade_surface(const SceneDev& sc, const GridConst& grid, const LightConst& light,
													  const uint32_t* states, const float4* irr, const float2* dep,
													  const float o[3], Vec3 d, float t, uint32_t tslot, float b1,
													  float b2, uint32_t* stack, int stride,
													  unsigned long long* visit_counters) {
	SurfaceShade r;
	const float b0 = 1.0f - b1 - b2;
	const Vec3 position = v3(d.x * t + o[0], d.y * t + o[1], d.z * t + o[2]);
	// shadow ray toward the light: any hit, mask 0xFF, tmin 0.1, tmax 10000.  It goes first: it needs
	// only the hit point, so few values stay live across the traversal (the shading terms, live
	// across it when it came last, spilled 48 bytes per lane at 6 waves per SIMD).
	RayPrep sr;
	ray_prepare(position.x, position.y, position.z, light.dir[0], light.dir[1], light.dir[2], sr);
	uint32_t nn = 0, nt = 0;
	const HitResult sh = traverse<true, STACK, COUNT>(sc.nodes, sc.tris, sr, 0.1f, 10000.0f, 0xFFu, stack, stride, nn, nt);
	if(COUNT) {
		atomicAdd(&visit_counters[2], (unsigned long long)nn);
		atomicAdd(&visit_counters[3], (unsigned long long)nt);
		atomicAdd(&visit_counters[4], 1ull);
	}
	const bool lit = sh.slot == 0xFFFFFFFFu;
	// fetch geometry and material records (closesthit.glsl:147-159)
	const float4* tr = sc.tris + size_t(tslot) * 3;
	const uint32_t gid = __float_as_uint(tr[0].w), inst = __float_as_uint(tr[1].w);
	const InstanceDev I = sc.inst[inst];
	const uint32_t prim = gid - I.prim_base;
	const uint32_t* oe = sc.offsets + size_t(I.custom_index) * 3;
	const uint32_t mat = oe[0], voff = oe[1], ib = oe[2] + 3u * prim;
	Vec3 nrm[3];
	for(int k = 0; k < 3; ++k) {
		const float4* vr = sc.vertices + size_t(voff + sc.indices[ib + k]) * 4;
		const float4 q1 = vr[1], q2 = vr[2];
		nrm[k] = v3(q1.z, q1.w, q2.x);
	}
	const Vec3 tsn = normalize3(nrm[0] * b0 + nrm[1] * b1 + nrm[2] * b2);
	const Vec3 normal = normalize3(v3(I.nm[0] * tsn.x + I.nm[1] * tsn.y + I.nm[2] * tsn.z,
									  I.nm[3] * tsn.x + I.nm[4] * tsn.y + I.nm[5] * tsn.z,
									  I.nm[6] * tsn.x + I.nm[7] * tsn.y + I.nm[8] * tsn.z));
	const float* m = sc.materials + size_t(mat) * 12;
	const float metal = m[0], rough = m[1];
	const Vec3 albedo = v3(m[2], m[3], m[4]);
	Vec3 color = v3(m[5], m[6], m[7]);
	const float f0 = 0.04f;
	const Vec3 diffuse = albedo * ((1.0f - f0) * (1.0f - metal));
	const Vec3 spec = v3(f0 * (1.0f - metal) + albedo.x * metal, f0 * (1.0f - metal) + albedo.y * metal,
						 f0 * (1.0f - metal) + albedo.z * metal);
	const float ndd = dot3(normal, d);
	const Vec3 refl = d - normal * (2.0f * ndd);
	const Vec3 to_cam = v3(-d.x, -d.y, -d.z);
	const Vec3 reflection = sample_probes(grid, states, irr, dep, position, refl, to_cam);
	color = color + spec * reflection;
	r.em_spec = color;
	const Vec3 indirect = sample_probes(grid, states, irr, dep, position, normal, to_cam);
	color = color + indirect * diffuse;
	float fade = 1.0f;
	if(lit) {
		const Vec3 L = v3(light.dir[0], light.dir[1], light.dir[2]);
		color = color + pbr_metallic_roughness(normal, normalize3(to_cam), v3(light.color[0], light.color[1], light.color[2]),
											   L, albedo, metal, rough);
		if(L.y < 0.0f) {
			fade = 1.0f - fminf(fmaxf(-L.y, 0.0f), 0.1f) / 0.1f;
			color = color * fade;
		}
	}
	r.color = color;
	r.indirect = indirect;
	r.normal = normal;
	r.to_cam = to_cam;
	r.metal = metal;
	r.rough = rough;
	r.fade = fade;
	r.mat = mat;
	r.lit = lit;
	return r;
}

// ---- shade-list ordering ----
// The primary pass appends front-face hits in ray order, so the hits one shading wave takes lie
// wherever one probe's rays ended.  Sorting the list by the Morton code of the probe-grid half-cell
// holding the hit point (cell_key) gives consecutive waves hits that share cage probes (the same
// atlas tiles) and cast shadow rays from nearby origins along the same direction (the same BVH
// nodes), which the caches then serve.  Each entry's shading depends on its own record only, so
// the order changes no output, and entries with equal keys may come in any order.
//
// Keys of up to 18 bits = a coarse digit (the high 10 bits) and a fine digit (the rest):
//  1. k_sort_count: kSortTiles workgroups, one per contiguous tile of the list (of one of its
//     segments), count the coarse digits of their tile in LDS and store them as counts[tile][digit];
//  2. k_sort_columns: per digit, the exclusive prefix sum over the tiles in place, and the digit's
//     total;
//  3. k_sort_bucket: every tile moves its entries into their coarse buckets: a run starts at the
//     sum of the totals of the smaller digits plus the column prefix (positions inside a run from
//     LDS counters);
//  4. k_sort_fine: fixed chunks of the bucketed list are ordered by their local keys in LDS, back
//     into the list.
// Every position comes from a prefix sum or an LDS counter: no global atomics, whose contention
// on a few thousand hot counters costs more than the whole sort.
constexpr uint32_t kSortTiles = 128, kSortDigits = 1024, kSortCoarseBits = 10, kSortUnroll = 8;
static_assert(kSortTiles * kSortDigits + kSortDigits == kShadeSortCounts && kSortDigits == 1u << kSortCoarseBits &&
				  kSortTiles % 4u == 0u,
			  "sort counters");

// Exclusive prefix sum over a workgroup of 1024 threads (16 waves): wave scans by lane shuffles,
// then the wave totals through LDS.  Returns this thread's prefix; `total` gets the sum.
__device__ __forceinline__ uint32_t block_exclusive_scan_1024(uint32_t v, uint32_t* s_wave, uint32_t& total) {
	const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
	uint32_t inc = v;
#pragma unroll
	for(int off = 1; off < 64; off <<= 1) {
		const uint32_t o = uint32_t(__shfl_up(int(inc), off));
		if(lane >= uint32_t(off)) inc += o;
	}
	if(lane == 63u) s_wave[wave] = inc;
	__syncthreads();
	uint32_t before = 0;
	total = 0;
#pragma unroll
	for(uint32_t w = 0; w < 16u; ++w) {
		const uint32_t x = s_wave[w];
		before += w < wave ? x : 0u;
		total += x;
	}
	__syncthreads();  // s_wave may be reused after this
	return before + inc - v;
}

// Length of a job's input list (the sum of its segments).
__device__ __forceinline__ uint32_t job_length(const SortJob& j) {
	uint32_t n = 0;
	for(uint32_t q = 0; q < j.segments; ++q) n += __builtin_amdgcn_readfirstlane(j.count[kCounterStride * q]);
	return n;
}
// Tile b of a job's input: entries [lo, hi) of segment `seg` (kSortTiles / segments tiles per
// segment; the tiles only have to partition the entries, in any order).
__device__ __forceinline__ void sort_tile(const SortJob& j, uint32_t b, size_t& base, uint32_t& lo, uint32_t& hi) {
	const uint32_t per_seg = kSortTiles / j.segments, seg = b / per_seg, sub = b - seg * per_seg;
	const uint32_t count = seg < j.segments ? __builtin_amdgcn_readfirstlane(j.count[kCounterStride * seg]) : 0u;
	const uint32_t per = (count + per_seg - 1u) / per_seg;
	base = size_t(seg) * j.seg_cap;
	lo = min(count, sub * per);
	hi = min(count, lo + per);
}

// The list workgroup (x, y) works on: the single job, or job y of a batch's array.
__device__ __forceinline__ SortJob sort_job(const SortJobs& J) {
	if(!J.many) return J.one;
	return *reinterpret_cast<const SortJob*>(J.many + size_t(blockIdx.y) * J.stride);
}

__global__ __launch_bounds__(1024) void k_sort_count(SortJobs J) {
	const SortJob j = sort_job(J);
	const uint32_t* __restrict__ keys = j.keys;
	const uint32_t fine_bits = j.fine_bits;
	uint32_t* __restrict__ counts = j.counts;
	__shared__ uint32_t s_cnt[kSortDigits];
	const uint32_t t = threadIdx.x;
	s_cnt[t] = 0u;
	__syncthreads();
	size_t base;
	uint32_t lo, hi;
	sort_tile(j, blockIdx.x, base, lo, hi);
	for(uint32_t i0 = lo; i0 < hi; i0 += 1024u * kSortUnroll) {
		uint32_t k[kSortUnroll];
#pragma unroll
		for(uint32_t u = 0; u < kSortUnroll; ++u) {
			const uint32_t i = i0 + 1024u * u + t;
			k[u] = i < hi ? keys[base + i] : 0xFFFFFFFFu;
		}
#pragma unroll
		for(uint32_t u = 0; u < kSortUnroll; ++u)
			if(k[u] != 0xFFFFFFFFu) atomicAdd(&s_cnt[k[u] >> fine_bits], 1u);
	}
	__syncthreads();
	counts[blockIdx.x * kSortDigits + t] = s_cnt[t];
}

// Column prefix sums: workgroup g takes digits 64 g .. 64 g + 63 (one per lane); wave q of the
// workgroup walks tiles 32 q .. 32 q + 31 of them (loads coalesced over the lanes, all in flight).
__global__ __launch_bounds__(256) void k_sort_columns(SortJobs J) {
	const SortJob j = sort_job(J);
	uint32_t* __restrict__ counts = j.counts;
	uint32_t* __restrict__ totals = j.counts + kSortTiles * kSortDigits;
	if(j.total && blockIdx.x == 0 && threadIdx.x == 0) {  // the ordered list's length, for the shading
		*j.total = job_length(j);
	}
	constexpr uint32_t kQ = kSortTiles / 4u;
	__shared__ uint32_t s_part[4][64];
	const uint32_t lane = __lane_id(), q = threadIdx.x >> 6;
	const uint32_t d = blockIdx.x * 64u + lane;
	uint32_t x[kQ];
	uint32_t sum = 0;
#pragma unroll
	for(uint32_t b = 0; b < kQ; ++b) x[b] = counts[(q * kQ + b) * kSortDigits + d];
#pragma unroll
	for(uint32_t b = 0; b < kQ; ++b) sum += x[b];
	s_part[q][lane] = sum;
	__syncthreads();
	uint32_t run = 0;
	for(uint32_t w = 0; w < q; ++w) run += s_part[w][lane];
	if(q == 3u) totals[d] = run + sum;
#pragma unroll
	for(uint32_t b = 0; b < kQ; ++b) {
		counts[(q * kQ + b) * kSortDigits + d] = run;
		run += x[b];
	}
}

__global__ __launch_bounds__(1024) void k_sort_bucket(SortJobs J) {
	const SortJob j = sort_job(J);
	const uint32_t* __restrict__ keys = j.keys;
	const uint32_t* __restrict__ list = j.list;
	const uint32_t fine_bits = j.fine_bits;
	const uint32_t* __restrict__ counts = j.counts;
	const uint32_t* __restrict__ totals = j.counts + kSortTiles * kSortDigits;
	uint32_t* __restrict__ bucket_keys = j.bucket_keys;
	uint32_t* __restrict__ bucket_list = j.bucket_list;
	__shared__ uint32_t s_pos[kSortDigits];
	__shared__ uint32_t s_wave[16];
	const uint32_t t = threadIdx.x;
	uint32_t all;
	const uint32_t digit_start = block_exclusive_scan_1024(totals[t], s_wave, all);
	s_pos[t] = digit_start + counts[blockIdx.x * kSortDigits + t];
	__syncthreads();
	size_t base;
	uint32_t lo, hi;
	sort_tile(j, blockIdx.x, base, lo, hi);
	for(uint32_t i0 = lo; i0 < hi; i0 += 1024u * kSortUnroll) {
		uint32_t k[kSortUnroll], e[kSortUnroll];
#pragma unroll
		for(uint32_t u = 0; u < kSortUnroll; ++u) {
			const uint32_t i = i0 + 1024u * u + t;
			k[u] = i < hi ? keys[base + i] : 0xFFFFFFFFu;
			e[u] = i < hi ? list[base + i] : 0u;
		}
#pragma unroll
		for(uint32_t u = 0; u < kSortUnroll; ++u) {
			if(k[u] == 0xFFFFFFFFu) continue;
			const uint32_t p = atomicAdd(&s_pos[k[u] >> fine_bits], 1u);
			bucket_keys[p] = k[u];
			bucket_list[p] = e[u];
		}
	}
}

// Fine pass over fixed chunks of kFineChunk entries of the bucketed list (equal work per workgroup
// however the keys are spread): a chunk is ordered in LDS by its local key = (coarse digit - the
// chunk's first coarse digit, fine digit), local keys from 1023 up sharing the last counter.  A
// chunk inside one bucket is ordered by the fine digit; one spanning buckets keeps their order.
constexpr uint32_t kFineChunk = 4096;
__global__ __launch_bounds__(1024) void k_sort_fine(SortJobs J) {
	const SortJob j = sort_job(J);
	const uint32_t* __restrict__ bucket_keys = j.bucket_keys;
	const uint32_t* __restrict__ bucket_list = j.bucket_list;
	const uint32_t fine_bits = j.fine_bits;
	uint32_t* __restrict__ out = j.list;
	constexpr uint32_t kU = kFineChunk / 1024u;
	__shared__ uint32_t s_cnt[1024];
	__shared__ uint32_t s_wave[16];
	const uint32_t t = threadIdx.x;
	const uint32_t count = job_length(j);
	const uint32_t c0 = blockIdx.x * kFineChunk;
	if(c0 >= count) return;  // uniform over the workgroup
	const uint32_t c1 = min(count, c0 + kFineChunk);
	const uint32_t first = bucket_keys[c0] >> fine_bits, mask = (1u << fine_bits) - 1u;
	s_cnt[t] = 0u;
	__syncthreads();
	uint32_t lk[kU], e[kU];
#pragma unroll
	for(uint32_t u = 0; u < kU; ++u) {
		const uint32_t i = c0 + 1024u * u + t;
		lk[u] = 0xFFFFFFFFu;
		e[u] = 0u;
		if(i < c1) {
			const uint32_t k = bucket_keys[i];
			const uint32_t d = (k >> fine_bits) - first;  // buckets come in increasing digit order
			lk[u] = d >= (1024u >> fine_bits) ? 1023u : min(1023u, (d << fine_bits) | (k & mask));
			e[u] = bucket_list[i];
		}
	}
#pragma unroll
	for(uint32_t u = 0; u < kU; ++u)
		if(lk[u] != 0xFFFFFFFFu) atomicAdd(&s_cnt[lk[u]], 1u);
	__syncthreads();
	uint32_t total;
	const uint32_t before = block_exclusive_scan_1024(s_cnt[t], s_wave, total);
	s_cnt[t] = before;
	__syncthreads();
#pragma unroll
	for(uint32_t u = 0; u < kU; ++u)
		if(lk[u] != 0xFFFFFFFFu) out[c0 + atomicAdd(&s_cnt[lk[u]], 1u)] = e[u];
}

// Entry k of the frame's shade list (the front-face hits queued by the primary traversal), so that
// every lane of a wave shades and casts a shadow ray.  Each ray's result depends only on its own
// record: the order of the list does not change any output.
template <int STACK, bool COUNT>
__device__ __forceinline__ void shade_ray(const TraceFrameArgs& a, uint32_t k, uint32_t* stack, int stride) {
	if(k >= a.shade_count[kShadeTotalSlot]) return;
	const uint32_t n = a.grid.n;
	const uint32_t ri = a.shade_list[k];  // hits and rays are stored in thread (ray_order) order
	const uint32_t slot = ri / n, j = ri - slot * n;
	const uint32_t i = a.perm[j];
	const float4 hit = a.hits[ri];
	const uint32_t code = __float_as_uint(hit.y);
	const float t = hit.x;
	float o[3];
	probe_origin(a.grid, a.indices[slot], o);
	const float4 d4 = a.dirs[i];
	const SurfaceShade r = shade_surface<STACK, COUNT>(a.scene, a.grid, a.light, a.states, a.irr, a.dep, o,
														v3(d4.x, d4.y, d4.z), t, code & 0x7FFFFFFFu, hit.z, hit.w, stack,
														stride, a.visit_counters);
	a.rays[ri] = make_float4(r.color.x, r.color.y, r.color.z, t);
	if(a.factors) {
		// affine decomposition in (albedo, light radiance) for non-metallic materials
		const float f0 = 0.04f;
		const Vec3 c0 = r.em_spec * r.fade;
		const Vec3 A = r.indirect * ((1.0f - f0) * (1.0f - r.metal) * r.fade);
		Vec3 B = v3(0.0f, 0.0f, 0.0f), C = v3(0.0f, 0.0f, 0.0f);
		if(r.lit) {
			const Vec3 L = v3(a.light.dir[0], a.light.dir[1], a.light.dir[2]);
			const Vec3 one = v3(1.0f, 1.0f, 1.0f), zero = v3(0.0f, 0.0f, 0.0f);
			const Vec3 with = pbr_metallic_roughness(r.normal, normalize3(r.to_cam), one, L, one, r.metal, r.rough);
			const Vec3 without = pbr_metallic_roughness(r.normal, normalize3(r.to_cam), one, L, zero, r.metal, r.rough);
			B = (with - without) * r.fade;
			C = without * r.fade;
		}
		float4* f = a.factors + ri * 4;
		f[0] = make_float4(c0.x, c0.y, c0.z, __uint_as_float(r.mat));
		f[1] = make_float4(A.x, A.y, A.z, __uint_as_float(r.lit ? 1u : 0u));
		f[2] = make_float4(B.x, B.y, B.z, 0.0f);
		f[3] = make_float4(C.x, C.y, C.z, 0.0f);
	}
}

// One thread per entry of the ordered shade list, over a grid sized for the list's capacity (its
// length is known only on the device; workgroups past it end at once).  The hardware refills a CU
// as soon as one of its workgroups ends, so the CUs stay busy to the end of the list, which a
// resident grid striding over the list does not do.  At least 6 waves per SIMD (80 registers): the
// shadow traversal is latency-bound.
template <int STACK, bool COUNT>
__global__ __launch_bounds__(kShadeBlock, 6) void k_shade(TraceFrameArgs a) {
	__shared__ uint32_t stack[STACK * kShadeBlock];
	const uint32_t k = blockIdx.x * kShadeBlock + threadIdx.x;
	if(k < a.shade_count[kShadeTotalSlot]) shade_ray<STACK, COUNT>(a, k, stack + threadIdx.x, kShadeBlock);
}

// ---- misses under the atmospheric sky (miss.rmiss:17-23 with sky.glsl:60-127) ----
// Entry k of the miss list: the sky seen from the probe along the ray, with the light's colour
// scaled by its 4th component and the sun's Mie lobe; depth -1.
__global__ __launch_bounds__(256) void k_sky(TraceFrameArgs a) {
	const uint32_t count = a.shade_count[kMissCountSlot];
	const uint32_t n = a.grid.n;
	const Vec3 L = v3(a.light.dir[0], a.light.dir[1], a.light.dir[2]);
	const Vec3 C = v3(a.light.color[0], a.light.color[1], a.light.color[2]);
	for(uint32_t k = blockIdx.x * 256 + threadIdx.x; k < count; k += gridDim.x * 256) {
		const uint32_t ri = a.miss_list[k];
		const uint32_t slot = ri / n, j = ri - slot * n;
		float o[3];
		probe_origin(a.grid, a.indices[slot], o);
		const float4 d = a.dirs[a.perm[j]];
		const Vec3 c = sky_atmosphere(v3(o[0], o[1], o[2]), v3(d.x, d.y, d.z), L, C, a.light.sky_brightness, true);
		const float4 rec = make_float4(c.x, c.y, c.z, -1.0f);
		a.rays[ri] = rec;
		if(a.factors) {
			float4* f = a.factors + size_t(ri) * 4;
			f[0] = make_float4(rec.x, rec.y, rec.z, __uint_as_float(0xFFFFFFFFu));
			f[1] = f[2] = f[3] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
		}
	}
}

// ---- batched update of many volumes ----
// Start of a batched update: block v writes volume v's frame directions and thread -> ray order;
// block 0 also resets the work-queue heads and completion stamps and records the start time.
__global__ __launch_bounds__(kMaxRays) void k_batch_begin(BatchLaunch b) {
	static_assert(kMaxRays == 512, "one thread per ray of a volume");
	__shared__ __attribute__((aligned(16))) float s_key[kMaxRays];
	const BatchVolume& v = b.vols[blockIdx.x];
	const uint32_t n = v.trace.grid.n;
	float y = 0.0f;
	if(threadIdx.x < n) {
		const float4 d = frame_dir(v.R, v.fib[threadIdx.x]);
		v.dirs[threadIdx.x] = d;
		y = d.y;
	}
	frame_order_block(y, n, s_key, v.perm);
	if(threadIdx.x == 0) v.trace.shade_count[0] = v.trace.shade_count[kMissCountSlot] = 0u;
	if(blockIdx.x == 0) {
		if(threadIdx.x < 4) b.ctl->next[threadIdx.x] = 0u;
		for(uint32_t k = threadIdx.x; k < kMaxBatchVolumes; k += 512) b.ctl->done[k] = 0ull;
		if(threadIdx.x == 0) b.ctl->t0 = wall_clock64();
	}
}

// Persistent primary traversal over the work items of all volumes: each workgroup takes the next
// item (a run of consecutive traversal threads of one volume) from a queue until the queue is
// empty.  Every workgroup leaves once the head passes the item count, so the grid always drains.
// five waves per SIMD (96 VGPRs, no spills beyond the traversal's private stack)
template <int STACK, int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(5, 5))) void k_batch_trace(BatchLaunch b) {
	__shared__ uint32_t stack[STACK * BLOCK];
	__shared__ uint32_t s_item;
	uint32_t* head = &b.ctl->next[0];
	// bounded: a workgroup can take at most every item once, plus the fetch that ends the loop
	for(uint32_t round = 0; round <= b.trace_item_count; ++round) {
		if(threadIdx.x == 0) s_item = atomicAdd(head, 1u);
		__syncthreads();
		const uint32_t it = __builtin_amdgcn_readfirstlane(s_item);
		__syncthreads();
		if(it >= b.trace_item_count) break;
		const uint32_t vol = __builtin_amdgcn_readfirstlane(b.trace_items[it].volume);
		const uint32_t first = __builtin_amdgcn_readfirstlane(b.trace_items[it].first);
		const BatchVolume& v = b.vols[vol];
		const uint32_t end = min(first + b.trace_item_threads, v.threads);
		for(uint32_t g = first + threadIdx.x; g < end; g += BLOCK)
			trace_primary_ray<STACK, false>(v.trace, g, stack + threadIdx.x, BLOCK);
	}
}

// Shading of every volume's ordered shade list, one thread per entry that the list can hold (as
// k_shade): workgroup i takes the kShadeBlock entries of shade item i; entries past a list's length
// return at once.
template <int STACK>
__global__ __launch_bounds__(kShadeBlock, 6) void k_batch_shade(BatchLaunch b) {
	__shared__ uint32_t stack[STACK * kShadeBlock];
	const uint32_t vol = __builtin_amdgcn_readfirstlane(b.shade_items[blockIdx.x].volume);
	const uint32_t first = __builtin_amdgcn_readfirstlane(b.shade_items[blockIdx.x].first);
	shade_ray<STACK, false>(b.vols[vol].trace, first + threadIdx.x, stack + threadIdx.x, kShadeBlock);
}

// ---- arbitrary ray queries (public ddgi_hit records) ----
template <int STACK, bool ANY, bool CULL>
__global__ __launch_bounds__(kTraceBlock) void k_trace_rays(SceneDev sc, const ddgi_ray* __restrict__ rays,
															ddgi_hit* __restrict__ hits, uint64_t count, uint32_t mask) {
	__shared__ uint32_t stack[STACK * kTraceBlock];
	const uint64_t id = uint64_t(blockIdx.x) * kTraceBlock + threadIdx.x;
	if(id >= count) return;
	const ddgi_ray q = rays[id];
	RayPrep r;
	ray_prepare(q.origin[0], q.origin[1], q.origin[2], q.direction[0], q.direction[1], q.direction[2], r);
	uint32_t nn, nt;
	const HitResult h = traverse<ANY, STACK, false, CULL>(sc.nodes, sc.tris, r, q.tmin, q.tmax, mask, stack + threadIdx.x,
														   kTraceBlock, nn, nt);
	ddgi_hit out;
	if(h.slot == 0xFFFFFFFFu) {
		out.t = -1.0f;
		out.u = out.v = 0.0f;
		out.instance = out.primitive = 0xFFFFFFFFu;
		out.flags = 0;
	} else {
		const float4 p = pack_hit(h);
		const uint32_t inst = __float_as_uint(sc.tris[size_t(h.slot) * 3 + 1].w);
		out.t = h.t;
		out.u = p.z;
		out.v = p.w;
		out.instance = inst;
		out.primitive = h.gid - sc.inst[inst].prim_base;
		out.flags = (__float_as_uint(p.y) & 0x80000000u) ? 1u : 0u;
	}
	hits[id] = out;
}

// ---- G-buffer of camera rays (GBuffer.vert.glsl:35-48, GBuffer.frag:32-66) ----
// The nearest front-facing hit (the G-buffer pass culls back faces,
// src/RenderPasses/GBufferPipeline.cpp:212): world position, its distance to the ray origin,
// normal = normalize(normal matrix * interpolated vertex normal), albedo = interpolated vertex
// colour * base colour, metalness, roughness, emissive, material.  A miss leaves depth 0.
// seven waves per SIMD (72 VGPRs, no spills beyond the traversal's private stack)
template <int STACK>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(7, 7))) void k_gbuffer(SceneDev sc, const ddgi_ray* __restrict__ rays, uint64_t count,
														 ddgi_gbuffer_texel* __restrict__ out) {
	__shared__ uint32_t stack[STACK * kTraceBlock];
	const uint64_t id = uint64_t(blockIdx.x) * kTraceBlock + threadIdx.x;
	if(id >= count) return;
	const ddgi_ray q = rays[id];
	RayPrep r;
	ray_prepare(q.origin[0], q.origin[1], q.origin[2], q.direction[0], q.direction[1], q.direction[2], r);
	uint32_t nn, nt;
	const HitResult h = traverse<false, STACK, false, true>(sc.nodes, sc.tris, r, q.tmin, q.tmax, 0xFFu,
															 stack + threadIdx.x, kTraceBlock, nn, nt);
	ddgi_gbuffer_texel g;
	if(h.slot == 0xFFFFFFFFu) {  // depth 0; the normal slot carries the ray direction (for the sky)
		for(int k = 0; k < 3; ++k) {
			g.position[k] = g.albedo[k] = g.emissive[k] = 0.0f;
			g.normal[k] = q.direction[k];
		}
		g.depth = g.metalness = g.roughness = 0.0f;
		g.material = 0xFFFFFFFFu;
		out[id] = g;
		return;
	}
	const float t = h.t;
	const Vec3 position = v3(q.direction[0] * t + q.origin[0], q.direction[1] * t + q.origin[1],
							 q.direction[2] * t + q.origin[2]);
	const Vec3 rel = position - v3(q.origin[0], q.origin[1], q.origin[2]);
	const float b1 = h.V / h.det, b2 = h.W / h.det, b0 = 1.0f - b1 - b2;
	const float4* tr = sc.tris + size_t(h.slot) * 3;
	const uint32_t gid = __float_as_uint(tr[0].w), inst = __float_as_uint(tr[1].w);
	const InstanceDev I = sc.inst[inst];
	const uint32_t* oe = sc.offsets + size_t(I.custom_index) * 3;
	const uint32_t mat = oe[0], voff = oe[1], ib = oe[2] + 3u * (gid - I.prim_base);
	Vec3 nsum = v3(0.0f, 0.0f, 0.0f), csum = v3(0.0f, 0.0f, 0.0f);
	const float b[3] = {b0, b1, b2};
	for(int k = 0; k < 3; ++k) {
		const float4* vr = sc.vertices + size_t(voff + sc.indices[ib + k]) * 4;  // 64-byte vertex record
		const float4 q0 = vr[0], q1 = vr[1], q2 = vr[2];
		nsum = nsum + v3(q1.z, q1.w, q2.x) * b[k];
		csum = csum + v3(q0.w, q1.x, q1.y) * b[k];
	}
	const Vec3 normal = normalize3(v3(I.nm[0] * nsum.x + I.nm[1] * nsum.y + I.nm[2] * nsum.z,
									  I.nm[3] * nsum.x + I.nm[4] * nsum.y + I.nm[5] * nsum.z,
									  I.nm[6] * nsum.x + I.nm[7] * nsum.y + I.nm[8] * nsum.z));
	const float* m = sc.materials + size_t(mat) * 12;
	g.position[0] = position.x;
	g.position[1] = position.y;
	g.position[2] = position.z;
	g.depth = sqrtf(dot3(rel, rel));
	g.normal[0] = normal.x;
	g.normal[1] = normal.y;
	g.normal[2] = normal.z;
	g.metalness = m[0];
	g.albedo[0] = csum.x * m[2];
	g.albedo[1] = csum.y * m[3];
	g.albedo[2] = csum.z * m[4];
	g.roughness = m[1];
	g.emissive[0] = m[5];
	g.emissive[1] = m[6];
	g.emissive[2] = m[7];
	g.material = mat;
	out[id] = g;
}

// ---- final gather (FinalGather.frag:38-77) over a G-buffer ----
// Miss: the sky (the constant, or the atmospheric model along the pixel ray that the G-buffer texel
// carries in its normal).  Hit: sun visibility * pbrMetallicRoughness + mix(f0, albedo, metal) *
// reflection + sampleProbes(position, normal, view) * albedo (1 - f0)(1 - metal) + emissive, with
// f0 = 0.004.  Sun visibility: `direct[i]` when given, else a hard shadow toward the sun as
// directLight.rgen:66-90 without its jitter: when dot(sdir, n) > 0, an any-hit ray from the
// position (tmin 0.01, tmax 10000) that escapes gives 1, faded below the horizon; otherwise 0.

template <int STACK>
__global__ __launch_bounds__(kTraceBlock) void k_gather(GatherArgs a) {
	__shared__ uint32_t stack[STACK * kTraceBlock];
	const uint64_t id = uint64_t(blockIdx.x) * kTraceBlock + threadIdx.x;
	if(id >= a.count) return;
	const ddgi_gbuffer_texel g = a.gbuffer[id];
	float* o = a.out + id * 3;
	if(g.depth <= 0.0f) {  // FinalGather.frag:52-54: the sky (sun scale 1) along the pixel's ray
		Vec3 c = v3(a.light.sky[0], a.light.sky[1], a.light.sky[2]);
		if(a.light.sky_model == DDGI_SKY_ATMOSPHERE)
			c = sky_atmosphere(v3(a.origin[0], a.origin[1], a.origin[2]), v3(g.normal[0], g.normal[1], g.normal[2]),
							   v3(a.light.dir[0], a.light.dir[1], a.light.dir[2]),
							   v3(a.light.color[0], a.light.color[1], a.light.color[2]), 1.0f, true);
		o[0] = c.x;
		o[1] = c.y;
		o[2] = c.z;
		return;
	}
	const Vec3 position = v3(g.position[0], g.position[1], g.position[2]);
	const Vec3 normal = normalize3(v3(g.normal[0], g.normal[1], g.normal[2]));
	const Vec3 albedo = v3(g.albedo[0], g.albedo[1], g.albedo[2]);
	const float metal = g.metalness;
	const Vec3 view = normalize3(v3(a.origin[0], a.origin[1], a.origin[2]) - position);
	float visible = 0.0f;
	if(a.direct) {
		visible = a.direct[id];
	} else {
		const Vec3 sd = v3(a.sdir[0], a.sdir[1], a.sdir[2]);
		if(dot3(sd, normal) > 0.0f) {
			RayPrep sr;
			ray_prepare(position.x, position.y, position.z, sd.x, sd.y, sd.z, sr);
			uint32_t nn, nt;
			const HitResult sh = traverse<true, STACK>(a.scene.nodes, a.scene.tris, sr, 0.01f, 10000.0f, 0xFFu,
														stack + threadIdx.x, kTraceBlock, nn, nt);
			if(sh.slot == 0xFFFFFFFFu) {
				visible = 1.0f;
				if(sd.y < 0.0f) visible *= 1.0f - fminf(fmaxf(-sd.y, 0.0f), 0.1f) / 0.1f;
			}
		}
	}
	Vec3 color = pbr_metallic_roughness(normal, view, v3(a.light.color[0], a.light.color[1], a.light.color[2]),
										v3(a.light.dir[0], a.light.dir[1], a.light.dir[2]), albedo, metal, g.roughness) *
				 visible;
	const float f0 = 0.004f;
	const Vec3 diffuse = albedo * (1.0f - f0) * (1.0f - metal);
	if(a.reflection) {
		const Vec3 spec = v3(f0 * (1.0f - metal) + albedo.x * metal, f0 * (1.0f - metal) + albedo.y * metal,
							 f0 * (1.0f - metal) + albedo.z * metal);
		const float* rf = a.reflection + id * 3;
		color = color + spec * v3(rf[0], rf[1], rf[2]);
	}
	const Vec3 indirect = sample_probes(a.grid, a.states, a.irr, a.dep, position, normal, view);
	color = color + indirect * diffuse;
	color = color + v3(g.emissive[0], g.emissive[1], g.emissive[2]);
	o[0] = color.x;
	o[1] = color.y;
	o[2] = color.z;
}

// ---- ray-traced shadow pass (directLight.rgen:42-99) ----
// Miss texel: -1.  Otherwise the sun direction, jittered inside a 0.02 rad cone by the texel's
// noise sample (none: the exact direction); facing away from the normal: 0; else an any-hit ray
// (mask 0xFF, tmin 0.01, tmax 10000) gives (visibility, previous g, previous b, 1), visibility 1 on
// escape, faded while the sun is below the horizon.
template <int STACK>
__global__ __launch_bounds__(kTraceBlock) void k_direct_light(DirectLightArgs a) {
	__shared__ uint32_t stack[STACK * kTraceBlock];
	const uint64_t id = uint64_t(blockIdx.x) * kTraceBlock + threadIdx.x;
	if(id >= a.count) return;
	const ddgi_gbuffer_texel g = a.gbuffer[id];
	float4 out = make_float4(-1.0f, -1.0f, -1.0f, -1.0f);
	if(g.depth > 0.0f) {
		const Vec3 normal = v3(g.normal[0], g.normal[1], g.normal[2]);
		Vec3 dir = v3(a.sdir[0], a.sdir[1], a.sdir[2]);
		if(a.noise) {
			const float2 nz = a.noise[id];
			const Vec3 temp = rotate_axis(dir, normalize3(cross3(normal, dir)), 2.0f * (nz.x - 0.5f) * 0.02f);
			dir = rotate_axis(temp, dir, (2.0f * sky::kPi) * nz.y);
		}
		out = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
		if(dot3(dir, normal) > 0.0f) {
			RayPrep sr;
			ray_prepare(g.position[0], g.position[1], g.position[2], dir.x, dir.y, dir.z, sr);
			uint32_t nn, nt;
			const HitResult sh = traverse<true, STACK>(a.scene.nodes, a.scene.tris, sr, 0.01f, 10000.0f, 0xFFu,
														stack + threadIdx.x, kTraceBlock, nn, nt);
			float vis = 0.0f;
			if(sh.slot == 0xFFFFFFFFu) {
				vis = 1.0f;
				if(dir.y < 0.0f) vis *= 1.0f - fminf(fmaxf(-dir.y, 0.0f), 0.1f) / 0.1f;
			}
			const float4 prev = a.previous ? a.previous[id] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
			out = make_float4(vis, prev.y, prev.z, 1.0f);
		}
	}
	a.out[id] = out;
}

// ---- reflection pass (reflection.rgen:101-175) ----
// Texels with depth > 0 and (roughness < 0.4 or metalness > 0.01) trace one ray along the mirror
// direction of the view vector, rotated by the texel's noise sample within the roughness lobe (no
// sample: the mirror direction), closest hit, mask 0xFF, tmin 0.1, tmax 10000.  The hit is shaded
// by closesthit.glsl at recursion depth 1 (its specular term reads the probes), a back face is
// black, a miss sees the sky; out = (colorCompression(radiance), roughness), other texels 0.
// k_reflection_trace finishes misses, back faces and untraced texels and queues front-face hits
// with the cell key of the hit point; k_reflection_shade shades the queue after it is ordered by
// key (as the probe update's shade list), recomputing each ray's direction with the same operations.

// The reflection ray of texel id, or false when the texel traces none.
__device__ __forceinline__ bool reflection_ray(const ReflectionArgs& a, uint64_t id, const ddgi_gbuffer_texel& g,
											   Vec3& dir) {
	if(!(g.depth > 0.0f && (g.roughness < 0.4f || g.metalness > 0.01f))) return false;
	const Vec3 position = v3(g.position[0], g.position[1], g.position[2]);
	const Vec3 normal = v3(g.normal[0], g.normal[1], g.normal[2]);
	const Vec3 to_origin = normalize3(v3(a.origin[0], a.origin[1], a.origin[2]) - position);
	const Vec3 incident = v3(-to_origin.x, -to_origin.y, -to_origin.z);
	const Vec3 rdir = normalize3(incident - normal * (2.0f * dot3(normal, incident)));
	dir = rdir;
	if(a.noise) {
		const float2 nz = a.noise[id];
		const float theta = g.roughness * (nz.x - 0.5f) * 2.0f * sky::kPi;
		const float phi = (nz.y - 0.5f) * 2.0f * sky::kPi;
		const Vec3 tangent = dot3(rdir, normal) < 0.9f ? normalize3(cross3(rdir, normal))
													   : normalize3(cross3(rdir, v3(1.0f, 0.0f, 0.0f)));
		dir = rotate_axis(rdir, tangent, theta);
		dir = rotate_axis(dir, rdir, phi);
	}
	return true;
}

// colorCompression = reinhard_whitepoint(v, 1): v (1 + v / 1) / (1 + v), with the roughness in w
__device__ __forceinline__ float4 compress_reflection(const Vec3& c, float roughness) {
	return make_float4(c.x * (1.0f + c.x / 1.0f) / (1.0f + c.x), c.y * (1.0f + c.y / 1.0f) / (1.0f + c.y),
					   c.z * (1.0f + c.z / 1.0f) / (1.0f + c.z), roughness);
}

template <int STACK>
__global__ __launch_bounds__(kTraceBlock) void k_reflection_trace(ReflectionArgs a) {
	__shared__ uint32_t stack[STACK * kTraceBlock];
	const uint32_t local = blockIdx.x * kTraceBlock + threadIdx.x;
	const uint64_t id = a.first + local;
	if(local >= a.chunk || id >= a.count) return;
	const ddgi_gbuffer_texel g = a.gbuffer[id];
	float4 out = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
	bool queued = false;
	uint32_t key = 0;
	Vec3 dir;
	if(reflection_ray(a, id, g, dir)) {
		const float o[3] = {g.position[0], g.position[1], g.position[2]};
		RayPrep r;
		ray_prepare(o[0], o[1], o[2], dir.x, dir.y, dir.z, r);
		uint32_t nn, nt;
		const HitResult h = traverse<false, STACK>(a.scene.nodes, a.scene.tris, r, 0.1f, 10000.0f, 0xFFu,
													stack + threadIdx.x, kTraceBlock, nn, nt);
		const float4 hit = pack_hit(h);
		const uint32_t code = __float_as_uint(hit.y);
		Vec3 c = v3(0.0f, 0.0f, 0.0f);
		if(code == 0xFFFFFFFFu) {  // miss.rmiss
			c = v3(a.light.sky[0], a.light.sky[1], a.light.sky[2]);
			if(a.light.sky_model == DDGI_SKY_ATMOSPHERE)
				c = sky_atmosphere(v3(o[0], o[1], o[2]), dir, v3(a.light.dir[0], a.light.dir[1], a.light.dir[2]),
								   v3(a.light.color[0], a.light.color[1], a.light.color[2]), a.light.sky_brightness, true);
		} else if(!(code & 0x80000000u)) {
			queued = true;
			a.hits[local] = hit;
			key = cell_key(a.grid, o[0] + hit.x * dir.x, o[1] + hit.x * dir.y, o[2] + hit.x * dir.z, a.key_shift);
		}
		out = compress_reflection(c, g.roughness);
	}
	const uint32_t pos = wave_append(a.list_count, queued);
	if(queued) {
		a.list[pos] = local;
		a.keys[pos] = key;
	} else {
		a.out[id] = out;
	}
}

// One thread per entry of the key-ordered queue of front-face reflection hits, over a grid sized
// for the chunk (as k_shade).
template <int STACK>
__global__ __launch_bounds__(kShadeBlock, 5) void k_reflection_shade(ReflectionArgs a) {
	__shared__ uint32_t stack[STACK * kShadeBlock];
	const uint32_t count = *a.list_count;
	for(uint32_t k = blockIdx.x * kShadeBlock + threadIdx.x; k < count; k += gridDim.x * kShadeBlock) {
		const uint32_t local = a.list[k];
		const uint64_t id = a.first + local;
		const ddgi_gbuffer_texel g = a.gbuffer[id];
		Vec3 dir;
		reflection_ray(a, id, g, dir);  // queued texels trace a ray
		const float4 hit = a.hits[local];
		const float o[3] = {g.position[0], g.position[1], g.position[2]};
		const Vec3 c = shade_surface<STACK, false>(a.scene, a.grid, a.light, a.states, a.irr, a.dep, o, dir, hit.x,
												   __float_as_uint(hit.y), hit.z, hit.w, stack + threadIdx.x, kShadeBlock,
												   nullptr)
						   .color;
		a.out[id] = compress_reflection(c, g.roughness);
	}
}

// ---- instance moves (the TLAS instance update + refit of src/Renderer.cpp:671-742) ----
// World vertices of the triangles of instances [first, first + count): each row of the row-major
// 3x4 transform evaluated in double left to right and rounded once to float, exactly as the host
// flattening (bvh.cpp transform_point), so a moved scene traces like a freshly built one.  `meta`
// holds mask | flip << 8 per moved instance.
__global__ void k_retransform(float4* __restrict__ tris, const float* __restrict__ obj, uint32_t tri_count,
							  const float* __restrict__ xf, const uint32_t* __restrict__ meta, uint32_t first,
							  uint32_t count) {
	const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
	if(slot >= tri_count) return;
	float4* t = tris + size_t(slot) * 3;
	const uint32_t k = __float_as_uint(t[1].w) - first;  // wraps for instances below `first`
	if(k >= count) return;
	const float* M = xf + size_t(k) * 12;
	for(int v = 0; v < 3; ++v) {
		const float* p = obj + size_t(slot) * 9 + v * 3;
		float w[3];
		for(int r = 0; r < 3; ++r) {
			double acc = double(M[r * 4 + 0]) * double(p[0]);
			acc = acc + double(M[r * 4 + 1]) * double(p[1]);
			acc = acc + double(M[r * 4 + 2]) * double(p[2]);
			acc = acc + double(M[r * 4 + 3]);
			w[r] = float(acc);
		}
		const float4 old = t[v];
		t[v] = make_float4(w[0], w[1], w[2], v == 2 ? __uint_as_float(meta[k]) : old.w);
	}
}

__device__ __forceinline__ float pad_down_dev(float v) {
	const double d = double(v) - (fabs(double(v)) * 1e-6 + 1e-6);
	float f = float(d);
	if(double(f) > d) f = nextafterf(f, -__int_as_float(0x7f800000));
	return f;
}
__device__ __forceinline__ float pad_up_dev(float v) {
	const double d = double(v) + (fabs(double(v)) * 1e-6 + 1e-6);
	float f = float(d);
	if(double(f) < d) f = nextafterf(f, __int_as_float(0x7f800000));
	return f;
}

// Refit of one level of the wide tree (deeper levels already refitted): every child box of the
// level's nodes becomes the padded bounds of its triangles (leaf) or of its child node's boxes.
__global__ void k_refit_level(float4* __restrict__ nodes, const float4* __restrict__ tris,
							  const float* __restrict__ slot_box, const float* __restrict__ rel_xf,
							  const uint32_t* __restrict__ level_nodes, uint32_t count) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= count) return;
	float* rec = reinterpret_cast<float*>(nodes + size_t(level_nodes[i]) * (kNodeFloats / 4));
	for(int k = 0; k < int(kWide); ++k) {
		const uint32_t code = __float_as_uint(rec[kCodeRow + k]);
		if(code == kEmptyChild) continue;
		float lo[3] = {3.402823466e+38f, 3.402823466e+38f, 3.402823466e+38f};
		float hi[3] = {-3.402823466e+38f, -3.402823466e+38f, -3.402823466e+38f};
		if(code & kLeafFlag) {
			const uint32_t firstt = (code & ~kLeafFlag) >> 3, cnt = (code & 7u) + 1u;
			for(uint32_t q = 0; q < cnt; ++q) {
				const size_t slot = size_t(firstt + q);
				float tlo[3] = {3.402823466e+38f, 3.402823466e+38f, 3.402823466e+38f};
				float thi[3] = {-3.402823466e+38f, -3.402823466e+38f, -3.402823466e+38f};
				for(int v = 0; v < 3; ++v) {
					const float4 p = tris[slot * 3 + v];
					const float c[3] = {p.x, p.y, p.z};
					for(int a = 0; a < 3; ++a) {
						tlo[a] = fminf(tlo[a], c[a]);
						thi[a] = fmaxf(thi[a], c[a]);
					}
				}
				// the slot bounds only its piece of the triangle (a spatial split may have cut it): the
				// piece's box at the build, carried by the instance's motion since (the box of its 8
				// moved corners, in double, rounded outward), clips the triangle's box
				{
					const float* M = rel_xf + size_t(__float_as_uint(tris[slot * 3 + 1].w)) * 12;
					const float* bx = slot_box + slot * 6;
					double mlo[3] = {1e300, 1e300, 1e300}, mhi[3] = {-1e300, -1e300, -1e300};
					for(int c = 0; c < 8; ++c) {
						const double p[3] = {bx[(c & 1) ? 3 : 0], bx[(c & 2) ? 4 : 1], bx[(c & 4) ? 5 : 2]};
						for(int r = 0; r < 3; ++r) {
							const double q = double(M[r * 4 + 0]) * p[0] + double(M[r * 4 + 1]) * p[1] +
											 double(M[r * 4 + 2]) * p[2] + double(M[r * 4 + 3]);
							mlo[r] = fmin(mlo[r], q);
							mhi[r] = fmax(mhi[r], q);
						}
					}
					for(int a = 0; a < 3; ++a) {
						float flo = float(mlo[a]), fhi = float(mhi[a]);
						if(double(flo) > mlo[a]) flo = nextafterf(flo, -__int_as_float(0x7f800000));
						if(double(fhi) < mhi[a]) fhi = nextafterf(fhi, __int_as_float(0x7f800000));
						tlo[a] = fmaxf(tlo[a], flo);
						thi[a] = fminf(thi[a], fhi);
					}
				}
				for(int a = 0; a < 3; ++a) {
					lo[a] = fminf(lo[a], tlo[a]);
					hi[a] = fmaxf(hi[a], thi[a]);
				}
			}
			for(int a = 0; a < 3; ++a) {
				lo[a] = pad_down_dev(lo[a]);
				hi[a] = pad_up_dev(hi[a]);
			}
		} else {  // the child's padded boxes already bound it
			const float* ch = reinterpret_cast<const float*>(nodes + size_t(code) * (kNodeFloats / 4));
			for(int c = 0; c < int(kWide); ++c) {
				if(__float_as_uint(ch[kCodeRow + c]) == kEmptyChild) continue;
				for(int a = 0; a < 3; ++a) {
					lo[a] = fminf(lo[a], ch[(2 * a) * kWide + c]);
					hi[a] = fmaxf(hi[a], ch[(2 * a + 1) * kWide + c]);
				}
			}
		}
		for(int a = 0; a < 3; ++a) {
			rec[(2 * a) * kWide + k] = lo[a];
			rec[(2 * a + 1) * kWide + k] = hi[a];
		}
	}
}

// ---- probe-state init: one block per probe, 512 rays (probesInit.rgen:31-64) ----
template <int STACK>
__global__ __launch_bounds__(kTraceBlock) void k_init_probes(SceneDev sc, GridConst g, const float4* __restrict__ dirs512,
															 uint32_t* __restrict__ states, uint32_t* __restrict__ backface_out) {
	__shared__ uint32_t stack[STACK * kTraceBlock];
	__shared__ uint32_t s_back, s_affect;
	const uint32_t probe = blockIdx.x;
	if(threadIdx.x == 0) {
		s_back = 0;
		s_affect = 0;
	}
	__syncthreads();
	float o[3];
	probe_origin(g, probe, o);
	const float max_distance = g.cell_len;
	const float tmax = 1.5f * max_distance;
	uint32_t back = 0, affect = 0;
	for(uint32_t i = threadIdx.x; i < 512u; i += kTraceBlock) {
		const float4 d = dirs512[i];
		RayPrep r;
		ray_prepare(o[0], o[1], o[2], d.x, d.y, d.z, r);
		uint32_t nn, nt;
		const HitResult h = traverse<false, STACK>(sc.nodes, sc.tris, r, 0.01f, tmax, DDGI_MASK_STATIC, stack + threadIdx.x,
												   kTraceBlock, nn, nt);
		const float depth = (h.slot == 0xFFFFFFFFu) ? 3.402823466e+38f : h.t;
		if(depth < max_distance) {
			const bool is_back = (h.det < 0.0f) != ((h.meta & 0x100u) != 0u);
			if(is_back) ++back;
			const float px = o[0] + depth * d.x, py = o[1] + depth * d.y, pz = o[2] + depth * d.z;
			if(fabsf(px - o[0]) < g.cell[0] && fabsf(py - o[1]) < g.cell[1] && fabsf(pz - o[2]) < g.cell[2]) affect = 1;
		}
	}
	atomicAdd(&s_back, back);
	if(affect) atomicOr(&s_affect, 1u);
	__syncthreads();
	if(threadIdx.x == 0) {
		uint32_t st;
		if(float(s_back) > 0.5f * float(g.n)) st = 0;
		else st = s_affect ? 1u : 8u;
		states[probe] = st;
		if(backface_out) backface_out[probe] = s_back;
	}
}

// ---- device-side selectProbesToUpdate: one block of 1024 threads scans the round-robin order ----
constexpr int kSelectBlock = 1024;

__global__ __launch_bounds__(kSelectBlock) void k_select(const uint32_t* __restrict__ states, uint32_t P, uint32_t cap,
														 SelectState* st, uint32_t* __restrict__ out) {
	__shared__ uint32_t wave_sums[kSelectBlock / 64];
	__shared__ uint32_t s_base, s_stop_q, s_done;
	const uint32_t off = st->last_offset, loop = st->loop_index;
	if(threadIdx.x == 0) {
		s_base = 0;
		s_stop_q = 0xFFFFFFFFu;
		s_done = 0;
	}
	__syncthreads();
	const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
	for(uint32_t chunk = 0; chunk < P; chunk += kSelectBlock) {
		const uint32_t q = chunk + threadIdx.x;
		bool pred = false;
		uint32_t idx = 0;
		if(q < P) {
			idx = off + q;
			uint32_t L = loop;
			if(idx >= P) {
				idx -= P;
				L += 1u;
			}
			const uint32_t s = states[idx];
			pred = (s != 0u) && ((idx + L) % s) == 0u;
		}
		// block-wide exclusive scan of pred
		const uint64_t ballot = __ballot(pred);
		const uint32_t in_wave = __popcll(ballot & ((uint64_t(1) << lane) - 1u));
		if(lane == 0) wave_sums[wave] = __popcll(ballot);
		__syncthreads();
		uint32_t wave_off = 0;
		for(uint32_t w = 0; w < wave; ++w) wave_off += wave_sums[w];
		const uint32_t pos = s_base + wave_off + in_wave;
		if(pred && (cap == 0u || pos < cap)) {
			out[pos] = idx;
			if(cap != 0u && pos == cap - 1u) s_stop_q = q;  // exactly one thread
		}
		__syncthreads();
		if(threadIdx.x == 0) {
			uint32_t tot = 0;
			for(uint32_t w = 0; w < kSelectBlock / 64; ++w) tot += wave_sums[w];
			s_base += tot;
			if(cap != 0u && s_base >= cap) s_done = 1;
		}
		__syncthreads();
		if(s_done) break;
	}
	if(threadIdx.x == 0) {
		uint32_t cnt = s_base;
		if(cap != 0u && cnt > cap) cnt = cap;
		st->count = cnt;
		if(s_stop_q != 0xFFFFFFFFu) {
			const uint32_t next = off + s_stop_q + 1u;
			st->last_offset = next >= P ? next - P : next;
			st->loop_index = loop + (next >= P ? 1u : 0u);
		} else {
			st->last_offset = off;
			st->loop_index = loop + 1u;
		}
	}
}

__global__ void k_set_count(SelectState* st, uint32_t c) {
	if(blockIdx.x == 0 && threadIdx.x == 0) st->count = c;
}

__global__ void k_iota(uint32_t* __restrict__ out, uint32_t P, SelectState* st) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i < P) out[i] = i;
	if(i == 0) st->count = P;
}

// ---- sampleProbes at arbitrary points ----
__global__ void k_sample(GridConst g, const uint32_t* __restrict__ states, const float4* __restrict__ irr,
						 const float2* __restrict__ dep, const float* __restrict__ pos, const float* __restrict__ nrm,
						 const float* __restrict__ tocam, uint64_t count, float* __restrict__ out) {
	const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if(i >= count) return;
	const Vec3 r = sample_probes(g, states, irr, dep, v3(pos[i * 3], pos[i * 3 + 1], pos[i * 3 + 2]),
								 v3(nrm[i * 3], nrm[i * 3 + 1], nrm[i * 3 + 2]), v3(tocam[i * 3], tocam[i * 3 + 1], tocam[i * 3 + 2]));
	out[i * 3] = r.x;
	out[i * 3 + 1] = r.y;
	out[i * 3 + 2] = r.z;
}

inline uint32_t blocks_for(uint64_t n, uint32_t b) { return uint32_t((n + b - 1) / b); }

}  // namespace

hipError_t launch_frame_dirs_order(const Rot& R, const float4* fib, float4* dirs, uint32_t n, uint16_t* perm,
								   hipStream_t s, uint32_t* zero) {
	static_assert(kShadeSegments <= kMaxRays, "one thread clears each shade-list counter");
	if(n == 0 || n > kMaxRays) return hipErrorInvalidValue;
	hipLaunchKernelGGL(k_frame_dirs_order, dim3(1), dim3(kMaxRays), 0, s, R, fib, dirs, n, perm, zero);
	return hipGetLastError();
}

hipError_t launch_frame_dirs(const Rot& R, const float4* fib, float4* dirs, uint32_t count, hipStream_t s,
							 uint32_t* zero) {
	hipLaunchKernelGGL(k_frame_dirs, dim3(blocks_for(count, 256)), dim3(256), 0, s, R, fib, dirs, count, zero);
	return hipGetLastError();
}

static hipError_t launch_sort_jobs(const SortJobs& J, uint32_t jobs, uint64_t capacity, hipStream_t s) {
	if(capacity == 0 || jobs == 0) return hipSuccess;
	const uint32_t chunks = uint32_t((capacity + kFineChunk - 1) / kFineChunk);
	hipLaunchKernelGGL(k_sort_count, dim3(kSortTiles, jobs), dim3(1024), 0, s, J);
	hipLaunchKernelGGL(k_sort_columns, dim3(kSortDigits / 64u, jobs), dim3(256), 0, s, J);
	hipLaunchKernelGGL(k_sort_bucket, dim3(kSortTiles, jobs), dim3(1024), 0, s, J);
	hipLaunchKernelGGL(k_sort_fine, dim3(chunks, jobs), dim3(1024), 0, s, J);
	return hipGetLastError();
}

uint32_t sort_fine_bits(uint32_t key_bits) { return key_bits > kSortCoarseBits ? key_bits - kSortCoarseBits : 0u; }

hipError_t launch_key_sort(SortJob job, uint64_t capacity, uint32_t key_bits, hipStream_t s) {
	if(key_bits > kSortCoarseBits + 8u || capacity > 0xFFFFFFFFull || job.segments == 0 ||
	   job.segments > kShadeSegments || kSortTiles % job.segments != 0)
		return hipErrorInvalidValue;
	job.fine_bits = sort_fine_bits(key_bits);
	SortJobs J{};
	J.one = job;
	return launch_sort_jobs(J, 1, capacity, s);
}

static_assert(std::is_standard_layout<BatchVolume>::value, "the batched sort finds each volume's job by offsetof");

hipError_t launch_batch_sort(const BatchLaunch& b, uint64_t max_capacity, hipStream_t s) {
	if(max_capacity > 0xFFFFFFFFull) return hipErrorInvalidValue;
	SortJobs J{};
	J.many = reinterpret_cast<const char*>(b.vols) + offsetof(BatchVolume, sort);
	J.stride = uint32_t(sizeof(BatchVolume));
	return launch_sort_jobs(J, b.volume_count, max_capacity, s);
}

void shade_key_layout(const GridConst& g, uint32_t* shift, uint32_t* bits_per_axis) {
	// half-cell coordinates below 64 per axis
	const uint32_t max_cell = uint32_t(2 * std::max({g.res[0], g.res[1], g.res[2]}) - 1);
	uint32_t sh = 0, b = 0;
	while((max_cell >> sh) >= 64u) ++sh;
	while(((max_cell >> sh) >> b) != 0u) ++b;
	*shift = sh;
	*bits_per_axis = b;
}

hipError_t launch_clear(uint32_t* words, uint32_t n, hipStream_t s) {
	if(n == 0) return hipSuccess;
	hipLaunchKernelGGL(k_clear, dim3(blocks_for(n, 256)), dim3(256), 0, s, words, n);
	return hipGetLastError();
}

hipError_t launch_trace_primary(const TraceFrameArgs& a, hipStream_t s) {
	const uint64_t threads = uint64_t(a.max_slots) * a.grid.n;
	if(threads == 0) return hipSuccess;
	const dim3 grid(blocks_for(threads, kPrimaryBlock));
	if(a.visit_counters) hipLaunchKernelGGL((k_trace_primary<kLdsStack, true>), grid, dim3(kPrimaryBlock), 0, s, a);
	else hipLaunchKernelGGL((k_trace_primary<kLdsStack, false>), grid, dim3(kPrimaryBlock), 0, s, a);
	return hipGetLastError();
}

hipError_t launch_sky(const TraceFrameArgs& a, hipStream_t s) {
	// one thread per entry the miss list can hold (the miss count is known only on the device)
	const uint32_t cap = a.max_slots * a.grid.n;
	hipLaunchKernelGGL(k_sky, dim3(std::max(1u, blocks_for(cap, 256))), dim3(256), 0, s, a);
	return hipGetLastError();
}

hipError_t launch_shade(const TraceFrameArgs& a, hipStream_t s) {
	const uint64_t threads = uint64_t(a.max_slots) * a.grid.n;  // upper bound of the shade list
	if(threads == 0) return hipSuccess;
	const bool cnt = a.visit_counters != nullptr;
	const void* f = cnt ? reinterpret_cast<const void*>(&k_shade<kLdsStack, true>)
						: reinterpret_cast<const void*>(&k_shade<kLdsStack, false>);
	const dim3 grid(blocks_for(threads, kShadeBlock));  // one thread per entry the list can hold
	TraceFrameArgs arg = a;
	void* args[] = {&arg};
	return hipLaunchKernel(f, grid, dim3(kShadeBlock), args, 0, s);
}

hipError_t launch_trace_rays(const SceneDev& sc, const ddgi_ray* rays, ddgi_hit* hits, uint64_t count, uint32_t mask,
							 bool any_hit, bool cull, hipStream_t s) {
	if(count == 0) return hipSuccess;
	const dim3 grid(blocks_for(count, kTraceBlock));
	if(any_hit) {
		if(cull) hipLaunchKernelGGL((k_trace_rays<kLdsStack, true, true>), grid, dim3(kTraceBlock), 0, s, sc, rays, hits, count, mask);
		else hipLaunchKernelGGL((k_trace_rays<kLdsStack, true, false>), grid, dim3(kTraceBlock), 0, s, sc, rays, hits, count, mask);
	} else {
		if(cull) hipLaunchKernelGGL((k_trace_rays<kLdsStack, false, true>), grid, dim3(kTraceBlock), 0, s, sc, rays, hits, count, mask);
		else hipLaunchKernelGGL((k_trace_rays<kLdsStack, false, false>), grid, dim3(kTraceBlock), 0, s, sc, rays, hits, count, mask);
	}
	return hipGetLastError();
}

hipError_t launch_retransform(float4* tris, const float* obj, uint32_t tri_count, const float* xf, const uint32_t* meta,
							  uint32_t first, uint32_t count, hipStream_t s) {
	if(tri_count == 0 || count == 0) return hipSuccess;
	hipLaunchKernelGGL(k_retransform, dim3(blocks_for(tri_count, 256)), dim3(256), 0, s, tris, obj, tri_count, xf, meta,
					   first, count);
	return hipGetLastError();
}

hipError_t launch_refit_level(float4* nodes, const float4* tris, const float* slot_box, const float* rel_xf,
							  const uint32_t* level_nodes, uint32_t count, hipStream_t s) {
	if(count == 0) return hipSuccess;
	hipLaunchKernelGGL(k_refit_level, dim3(blocks_for(count, 128)), dim3(128), 0, s, nodes, tris, slot_box, rel_xf,
					   level_nodes, count);
	return hipGetLastError();
}

hipError_t launch_gbuffer(const SceneDev& sc, const ddgi_ray* rays, uint64_t count, ddgi_gbuffer_texel* out, hipStream_t s) {
	if(count == 0) return hipSuccess;
	hipLaunchKernelGGL(k_gbuffer<kLdsStack>, dim3(blocks_for(count, kTraceBlock)), dim3(kTraceBlock), 0, s, sc, rays, count,
					   out);
	return hipGetLastError();
}

hipError_t launch_gather(const GatherArgs& a, hipStream_t s) {
	if(a.count == 0) return hipSuccess;
	hipLaunchKernelGGL(k_gather<kLdsStack>, dim3(blocks_for(a.count, kTraceBlock)), dim3(kTraceBlock), 0, s, a);
	return hipGetLastError();
}

hipError_t launch_direct_light(const DirectLightArgs& a, hipStream_t s) {
	if(a.count == 0) return hipSuccess;
	hipLaunchKernelGGL(k_direct_light<kLdsStack>, dim3(blocks_for(a.count, kTraceBlock)), dim3(kTraceBlock), 0, s, a);
	return hipGetLastError();
}

hipError_t launch_reflection_trace(const ReflectionArgs& a, hipStream_t s) {
	if(a.chunk == 0) return hipSuccess;
	hipLaunchKernelGGL(k_reflection_trace<kLdsStack>, dim3(blocks_for(a.chunk, kTraceBlock)), dim3(kTraceBlock), 0, s, a);
	return hipGetLastError();
}

hipError_t launch_reflection_shade(const ReflectionArgs& a, hipStream_t s) {
	if(a.chunk == 0) return hipSuccess;
	hipLaunchKernelGGL(k_reflection_shade<kLdsStack>, dim3(blocks_for(a.chunk, kShadeBlock)), dim3(kShadeBlock), 0, s, a);
	return hipGetLastError();
}

hipError_t launch_init_probes(const SceneDev& sc, const GridConst& g, const float4* dirs512, uint32_t* states,
							  uint32_t* backface_counts, hipStream_t s) {
	if(g.probe_count == 0) return hipSuccess;
	hipLaunchKernelGGL(k_init_probes<kLdsStack>, dim3(g.probe_count), dim3(kTraceBlock), 0, s, sc, g, dirs512, states, backface_counts);
	return hipGetLastError();
}

hipError_t launch_select(const uint32_t* states, uint32_t P, uint32_t cap, SelectState* st, uint32_t* indices,
						 hipStream_t s) {
	hipLaunchKernelGGL(k_select, dim3(1), dim3(kSelectBlock), 0, s, states, P, cap, st, indices);
	return hipGetLastError();
}

hipError_t launch_set_count(SelectState* st, uint32_t count, hipStream_t s) {
	hipLaunchKernelGGL(k_set_count, dim3(1), dim3(64), 0, s, st, count);
	return hipGetLastError();
}

hipError_t launch_iota(uint32_t* indices, uint32_t P, SelectState* st, hipStream_t s) {
	hipLaunchKernelGGL(k_iota, dim3(blocks_for(P, 256)), dim3(256), 0, s, indices, P, st);
	return hipGetLastError();
}

hipError_t launch_sample(const GridConst& g, const uint32_t* states, const float4* irr, const float2* dep, const float* pos,
						 const float* nrm, const float* tocam, uint64_t count, float* out, hipStream_t s) {
	if(count == 0) return hipSuccess;
	hipLaunchKernelGGL(k_sample, dim3(blocks_for(count, 256)), dim3(256), 0, s, g, states, irr, dep, pos, nrm, tocam, count, out);
	return hipGetLastError();
}

// ---- batched update ----
namespace {

const void* batch_trace_kernel(uint32_t block) {
	switch(block) {
		case 64: return reinterpret_cast<const void*>(&k_batch_trace<kLdsStack, 64>);
		case 128: return reinterpret_cast<const void*>(&k_batch_trace<kLdsStack, 128>);
		case 256: return reinterpret_cast<const void*>(&k_batch_trace<kLdsStack, 256>);
	}
	return nullptr;
}

}  // namespace

hipError_t launch_batch_begin(const BatchLaunch& b, hipStream_t s) {
	hipLaunchKernelGGL(k_batch_begin, dim3(b.volume_count), dim3(512), 0, s, b);
	return hipGetLastError();
}

hipError_t batch_trace_resident(uint32_t block, int* per_cu) {
	const void* f = batch_trace_kernel(block);
	if(!f) return hipErrorInvalidValue;
	return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, f, int(block), 0);
}

hipError_t launch_batch_shade(const BatchLaunch& b, hipStream_t s) {
	if(b.shade_item_count == 0) return hipSuccess;
	hipLaunchKernelGGL(k_batch_shade<kLdsStack>, dim3(b.shade_item_count), dim3(kShadeBlock), 0, s, b);
	return hipGetLastError();
}

hipError_t launch_batch_trace(const BatchLaunch& b, uint32_t block, uint32_t grid, hipStream_t s) {
	const void* f = batch_trace_kernel(block);
	if(!f || grid == 0) return hipErrorInvalidValue;
	BatchLaunch arg = b;
	void* args[] = {&arg};
	return hipLaunchKernel(f, dim3(grid), dim3(block), args, 0, s);
}

}  // namespace ddgi
