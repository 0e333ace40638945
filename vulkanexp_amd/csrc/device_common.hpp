// Device-side building blocks shared by the probe kernels (gfx950).
//
// Compiled with -ffp-contract=off in the trace translation unit: the intersection arithmetic
// must be exactly the FP32 operations listed here, in this order, with no fused multiply-add,
// so that hit ids, t and barycentrics are reproducible bit for bit on the host.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.hpp"

namespace ddgi {

// ---------------------------------------------------------------------------------------------
// Ray precomputation for the watertight test (Woop, Benthin, Wald, JCGT 2(1) 2013): kz is the
// dominant axis of the direction (ties resolved x < y < z as written below), (kx, ky) the other
// two in cyclic order, swapped when d[kz] < 0 to keep the winding; shear constants Sx, Sy, Sz.
struct RayPrep {
	float o[3];
	int kx, ky, kz;
	float sx, sy, sz;
	float inv[3];  // 1/d for the slab tests
	float okx, oky, okz;
};

__device__ __forceinline__ float sel3(float a, float b, float c, int k) { return k == 0 ? a : (k == 1 ? b : c); }

// Workgroups are dealt round-robin to the 8 XCDs of an MI355X (workgroup i -> XCD i % 8), and each
// XCD has its own L2.  Remap so that XCD x processes one contiguous range of logical blocks (a
// slab of neighbouring probes whose rays touch overlapping BVH nodes and atlas tiles).  A
// bijection on [0, nblocks) for any nblocks.
constexpr uint32_t kNumXcd = 8;
__device__ __forceinline__ uint32_t xcd_remap(uint32_t block, uint32_t nblocks) {
	const uint32_t q = nblocks / kNumXcd, r = nblocks % kNumXcd;
	const uint32_t x = block % kNumXcd, k = block / kNumXcd;
	return x * q + min(x, r) + k;
}

// List append from a wavefront: every active lane with `pred` gets a distinct position in the list
// whose length is *counter, with one atomic per wave (issued by the lowest such lane).  Explicit, so
// that it does not depend on the compiler proving `counter` uniform.
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool pred) {
	const uint64_t mask = __ballot(pred);
	const uint32_t lane = __lane_id();
	const uint32_t before = uint32_t(__popcll(mask & ((uint64_t(1) << lane) - 1u)));
	uint32_t base = 0;
	if(mask != 0u) {
		const int leader = __ffsll((unsigned long long)mask) - 1;
		if(int(lane) == leader) base = atomicAdd(counter, uint32_t(__popcll(mask)));
		base = __shfl(base, leader);
	}
	return base + before;
}

__device__ __forceinline__ void ray_prepare(float ox, float oy, float oz, float dx, float dy, float dz, RayPrep& r) {
	const float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
	const int kz = (ax > ay) ? ((ax > az) ? 0 : 2) : ((ay > az) ? 1 : 2);
	int kx = kz + 1;
	if(kx == 3) kx = 0;
	int ky = kx + 1;
	if(ky == 3) ky = 0;
	const float dkz = sel3(dx, dy, dz, kz);
	if(dkz < 0.0f) {
		const int t = kx;
		kx = ky;
		ky = t;
	}
	r.kx = kx;
	r.ky = ky;
	r.kz = kz;
	r.sx = sel3(dx, dy, dz, kx) / dkz;
	r.sy = sel3(dx, dy, dz, ky) / dkz;
	r.sz = 1.0f / dkz;
	r.o[0] = ox;
	r.o[1] = oy;
	r.o[2] = oz;
	r.inv[0] = 1.0f / dx;
	r.inv[1] = 1.0f / dy;
	r.inv[2] = 1.0f / dz;
	r.okx = sel3(ox, oy, oz, kx);
	r.oky = sel3(ox, oy, oz, ky);
	r.okz = sel3(ox, oy, oz, kz);
}

// Returns true when the triangle is hit with tmin <= t <= tmax; U,V,W,det are the edge functions.
__device__ __forceinline__ bool tri_intersect(const RayPrep& r, float4 a, float4 b, float4 c, float tmin, float tmax,
											  float& t, float& V, float& W, float& det) {
	const float Akx = sel3(a.x, a.y, a.z, r.kx) - r.okx, Aky = sel3(a.x, a.y, a.z, r.ky) - r.oky,
				Akz = sel3(a.x, a.y, a.z, r.kz) - r.okz;
	const float Bkx = sel3(b.x, b.y, b.z, r.kx) - r.okx, Bky = sel3(b.x, b.y, b.z, r.ky) - r.oky,
				Bkz = sel3(b.x, b.y, b.z, r.kz) - r.okz;
	const float Ckx = sel3(c.x, c.y, c.z, r.kx) - r.okx, Cky = sel3(c.x, c.y, c.z, r.ky) - r.oky,
				Ckz = sel3(c.x, c.y, c.z, r.kz) - r.okz;
	const float Ax = Akx - r.sx * Akz, Ay = Aky - r.sy * Akz;
	const float Bx = Bkx - r.sx * Bkz, By = Bky - r.sy * Bkz;
	const float Cx = Ckx - r.sx * Ckz, Cy = Cky - r.sy * Ckz;
	float U = Cx * By - Cy * Bx;
	V = Ax * Cy - Ay * Cx;
	W = Bx * Ay - By * Ax;
	if(U == 0.0f || V == 0.0f || W == 0.0f) {
		U = float(double(Cx) * double(By) - double(Cy) * double(Bx));
		V = float(double(Ax) * double(Cy) - double(Ay) * double(Cx));
		W = float(double(Bx) * double(Ay) - double(By) * double(Ax));
	}
	if((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
	det = U + V + W;
	if(det == 0.0f) return false;
	const float Az = r.sz * Akz, Bz = r.sz * Bkz, Cz = r.sz * Ckz;
	const float T = U * Az + V * Bz + W * Cz;
	t = T / det;
	return t >= tmin && t <= tmax;
}

// Conservative slab test of one child box: [tn, tf] clipped to [tmin, tmax], far end widened by
// a few ulps (robust traversal, Ize 2013); boxes are also inflated at build time.  The planes come
// as the near and far plane of each axis for this ray (lo where 1/d is positive, hi where it is
// negative): (p - o) / d rounds monotonically in p, so the near plane's distance is the smaller of
// the two and no min / max pair per axis is needed.  With d = +-0 the distances are infinite and
// the test keeps the box whenever the origin lies within its slab.
__device__ __forceinline__ bool slab(const RayPrep& r, float nx, float fx, float ny, float fy, float nz, float fz,
									 float tmin, float tmax, float& tn) {
	const float xn = (nx - r.o[0]) * r.inv[0], xf = (fx - r.o[0]) * r.inv[0];
	const float yn = (ny - r.o[1]) * r.inv[1], yf = (fy - r.o[1]) * r.inv[1];
	const float zn = (nz - r.o[2]) * r.inv[2], zf = (fz - r.o[2]) * r.inv[2];
	tn = fmaxf(fmaxf(xn, yn), fmaxf(zn, tmin));
	const float tf = fminf(fminf(xf, yf), fminf(zf, tmax));
	return tn <= tf * 1.000001f;
}

struct HitResult {
	float t;
	uint32_t slot;  // triangle position in the BVH triangle array, 0xFFFFFFFF = miss
	uint32_t gid;   // global triangle id (ties go to the lower one)
	uint32_t meta;  // mask | flip_facing << 8 of the hit triangle
	float V, W, det;
};

// Closest-hit (ANY = false; ties -> lowest global id) or any-hit traversal of the BVH.
//
// Nodes are 4-wide (bvh.hpp): one 128-byte fetch tests four child boxes.  The children the ray
// enters are sorted by entry distance (a five-comparator network on registers); the nearest is
// visited next and the others are pushed farthest first, so the next pop is the nearest remaining.
// The result does not depend on the visiting order: closest hit is the minimum of (t, global id)
// over every triangle the conservative box tests let through.
//
// Stack: the newest STACK entries live in an LDS ring laid out [entry][thread] (conflict-free;
// `stack` points at this thread's column, `stride` is the block size); when the ring is full its
// oldest entry moves to a private spill array, and pops drain the ring first, then the spill, so
// the order stays LIFO.  At most kWide - 1 entries per level of the wide tree are pending, and
// STACK + kSpill covers the deepest tree accepted (kMaxWideDepth levels), while the LDS footprint
// stays small enough for full occupancy.
//
// Leaves are tested two triangles at a time, each triangle's 48 bytes loaded in one go, so that
// two loads are in flight and the mask word costs no extra round trip.
//
// With COUNT, n_nodes / n_tris receive the number of node (kNodeFloats * 4 bytes) and 48-byte
// triangle fetches
// (the algorithmic bytes of the traversal).
constexpr int kLdsStack = 16;  // LDS ring entries per thread
// private spill entries: kLdsStack + kSpill = (kWide - 1) * kMaxWideDepth
constexpr int kSpill = (int(kWide) - 1) * int(kMaxWideDepth) - kLdsStack;
static_assert(kLdsStack + kSpill >= (int(kWide) - 1) * int(kMaxWideDepth), "traversal stack too small for the deepest tree");

// CULL: back-facing triangles are ignored (rasterizer back-face culling).
template <bool ANY, bool CULL>
__device__ __forceinline__ bool test_triangle(const RayPrep& r, float tmin, uint32_t mask, uint32_t slot, float4 ta,
											  float4 tb, float4 tc, HitResult& best) {
	if((__float_as_uint(tc.w) & mask) == 0u) return false;
	float t, V, W, det;
	if(!tri_intersect(r, ta, tb, tc, tmin, best.t, t, V, W, det)) return false;
	if(CULL && ((det < 0.0f) != ((__float_as_uint(tc.w) & 0x100u) != 0u))) return false;
	const uint32_t gid = __float_as_uint(ta.w);
	if(best.slot != 0xFFFFFFFFu && !(t < best.t || gid < best.gid)) return false;
	best.t = t;
	best.slot = slot;
	best.gid = gid;
	best.meta = __float_as_uint(tc.w);
	best.V = V;
	best.W = W;
	best.det = det;
	return true;
}

template <bool ANY, int STACK, bool COUNT = false, bool CULL = false>
__device__ __forceinline__ HitResult traverse(const float4* __restrict__ nodes, const float4* __restrict__ tris,
											  const RayPrep& r, float tmin, float tmax, uint32_t mask,
											  uint32_t* stack, int stride, uint32_t& n_nodes_out, uint32_t& n_tris_out) {
	static_assert((STACK & (STACK - 1)) == 0, "the LDS ring size must be a power of two");
	HitResult best;
	best.t = tmax;
	best.slot = 0xFFFFFFFFu;
	best.gid = 0xFFFFFFFFu;
	best.meta = 0u;
	best.V = best.W = best.det = 0.0f;
	uint32_t spill[kSpill];
	int head = 0, count = 0, nspill = 0;  // ring write position, ring entries, spilled entries
// pop into `v`, or run `on_empty` when ring and spill are both empty
#define DDGI_STACK_POP(v, on_empty)                    \
	if(count > 0) {                                    \
		head = (head - 1) & (STACK - 1);               \
		--count;                                       \
		v = stack[head * stride];                      \
	} else if(nspill > 0) {                            \
		v = spill[--nspill];                           \
	} else {                                           \
		on_empty;                                      \
	}
// push `v`; when the ring is full its oldest entry (at `head`) moves to the spill array first
#define DDGI_STACK_PUSH(v)                             \
	{                                                  \
		if(count == STACK) {                           \
			spill[nspill++] = stack[head * stride];    \
			--count;                                   \
		}                                              \
		stack[head * stride] = (v);                    \
		head = (head + 1) & (STACK - 1);               \
		++count;                                       \
	}
// order (key, code) pairs i, j by key
#define DDGI_CSWAP(i, j)                               \
	{                                                  \
		const bool sw_ = key[j] < key[i];              \
		const float ki_ = key[i], kj_ = key[j];        \
		const uint32_t ci_ = code[i], cj_ = code[j];   \
		key[i] = sw_ ? kj_ : ki_;                      \
		key[j] = sw_ ? ki_ : kj_;                      \
		code[i] = sw_ ? cj_ : ci_;                     \
		code[j] = sw_ ? ci_ : cj_;                     \
	}
	// Leaves are postponed (Aila and Laine's speculative traversal): a lane that reaches a leaf keeps
	// it in `leaf` and goes on descending from its stack until at least half of the active lanes of
	// the wavefront hold a leaf (or it has run out of nodes); the wavefront then tests leaves
	// together.  Waiting for every lane idles the lanes that hold one; testing at once idles the
	// others; half measured fastest (the probe update's primary pass 0.92 ms when waiting for all,
	// 0.86 ms at a half; an eighth 0.90, three quarters 0.89).  Closest hit is the minimum of
	// (t, global id) over all triangles, so the order changes no result.
	constexpr uint32_t kNone = 0xFFFFFFFFu;  // no node (empty stack) / no postponed leaf
	uint32_t node = 0;  // root (always an inner node)
	uint32_t leaf = kNone;
	uint32_t n_nodes = 0, n_tris = 0;
	const float kInf = __int_as_float(0x7f800000);
	// byte offset in a node of this ray's near plane row per axis (rows lo.x, hi.x, lo.y, hi.y, lo.z,
	// hi.z): the near plane is lo where 1/d is positive (sign bit clear), hi where not; the far row
	// is the other one of the pair, its offset the near one's with bit kRowBytes flipped (a node
	// starts at a multiple of 4 rows)
	constexpr uint32_t kRowBytes = 16u * (kWide / 4u);
	uint32_t near_off[3];
#pragma unroll
	for(int a = 0; a < 3; ++a) near_off[a] = kRowBytes * (2u * uint32_t(a) + (__float_as_uint(r.inv[a]) >> 31));
	const char* __restrict__ node_bytes = reinterpret_cast<const char*>(nodes);
	for(;;) {
		// descend through inner nodes
		while(!(node & kLeafFlag)) {
			// 32-bit offsets from the array's base (the node array stays far below 4 GiB)
			const uint32_t node_off = node * uint32_t(kNodeFloats * 4);
			constexpr int G = int(kWide) / 4;  // float4s per row
			float bx[6][kWide];  // near x, far x, near y, far y, near z, far z
			uint32_t code[kWide];
#pragma unroll
			for(int row = 0; row < 6; ++row) {
				const uint32_t near = node_off + near_off[row >> 1];
				const uint32_t off = (row & 1) ? near ^ kRowBytes : near;
#pragma unroll
				for(int g = 0; g < G; ++g) {
					const float4 v = *reinterpret_cast<const float4*>(node_bytes + (off + 16u * uint32_t(g)));
					bx[row][4 * g] = v.x;
					bx[row][4 * g + 1] = v.y;
					bx[row][4 * g + 2] = v.z;
					bx[row][4 * g + 3] = v.w;
				}
			}
#pragma unroll
			for(int g = 0; g < G; ++g) {
				const float4 v = *reinterpret_cast<const float4*>(node_bytes + (node_off + kRowBytes * 6u + 16u * uint32_t(g)));
				code[4 * g] = __float_as_uint(v.x);
				code[4 * g + 1] = __float_as_uint(v.y);
				code[4 * g + 2] = __float_as_uint(v.z);
				code[4 * g + 3] = __float_as_uint(v.w);
			}
			++n_nodes;
			// entry distance of each child the ray enters, infinity for the others (slab distances are
			// finite or clipped to [tmin, tmax], so a finite key is an entered child)
			float key[kWide];
#pragma unroll
			for(int k = 0; k < int(kWide); ++k) {
				float tn;
				key[k] = (slab(r, bx[0][k], bx[1][k], bx[2][k], bx[3][k], bx[4][k], bx[5][k], tmin, best.t, tn) &&
						  code[k] != kEmptyChild)
							 ? tn
							 : kInf;
			}
			// sorted by key, the entered children come first
			{
				if constexpr(kWide == 4) {
					DDGI_CSWAP(0, 1)
					DDGI_CSWAP(2, 3)
					DDGI_CSWAP(0, 2)
					DDGI_CSWAP(1, 3)
					DDGI_CSWAP(1, 2)
				} else {  // Batcher's odd-even merge network for 8 keys (19 comparators)
					DDGI_CSWAP(0, 1)
					DDGI_CSWAP(2, 3)
					DDGI_CSWAP(4, 5)
					DDGI_CSWAP(6, 7)
					DDGI_CSWAP(0, 2)
					DDGI_CSWAP(1, 3)
					DDGI_CSWAP(4, 6)
					DDGI_CSWAP(5, 7)
					DDGI_CSWAP(1, 2)
					DDGI_CSWAP(5, 6)
					DDGI_CSWAP(0, 4)
					DDGI_CSWAP(1, 5)
					DDGI_CSWAP(2, 6)
					DDGI_CSWAP(3, 7)
					DDGI_CSWAP(2, 4)
					DDGI_CSWAP(3, 5)
					DDGI_CSWAP(1, 2)
					DDGI_CSWAP(3, 4)
					DDGI_CSWAP(5, 6)
				}
			}
			if(key[0] == kInf) {
				DDGI_STACK_POP(node, node = kNone)
			} else {
				node = code[0];
#pragma unroll
				for(int k = int(kWide) - 1; k >= 1; --k)
					if(key[k] != kInf) DDGI_STACK_PUSH(code[k])
			}
			if(leaf == kNone && node != kNone && (node & kLeafFlag)) {  // postpone the leaf
				leaf = node;
				DDGI_STACK_POP(node, node = kNone)
			}
			// at least half of the active lanes hold a leaf
			if(uint32_t(__popcll(__ballot(leaf != kNone))) * 2u >= uint32_t(__popcll(__ballot(true)))) break;
		}
		// leaves: the postponed one, then any leaves next on the stack
		while(leaf != kNone) {
			const uint32_t first = (leaf & ~kLeafFlag) >> 3, cnt = (leaf & 7u) + 1u;
			for(uint32_t k = 0; k < cnt; k += 2) {
				const float4* tr = tris + size_t(first + k) * 3;
				const float4 ta = tr[0], tb = tr[1], tc = tr[2];
				const bool two = k + 1 < cnt;
				float4 ua = ta, ub = tb, uc = tc;
				if(two) {
					ua = tr[3];
					ub = tr[4];
					uc = tr[5];
				}
				n_tris += two ? 2u : 1u;
				// in index order, so that equal-t hits resolve exactly as a sequential scan would
				if(test_triangle<ANY, CULL>(r, tmin, mask, first + k, ta, tb, tc, best) && ANY) goto done;
				if(two && test_triangle<ANY, CULL>(r, tmin, mask, first + k + 1, ua, ub, uc, best) && ANY) goto done;
			}
			leaf = kNone;
			if(node != kNone && (node & kLeafFlag)) {
				leaf = node;
				DDGI_STACK_POP(node, node = kNone)
			}
		}
		if(node == kNone) break;
	}
done:
#undef DDGI_STACK_POP
#undef DDGI_STACK_PUSH
#undef DDGI_CSWAP
	if(COUNT) {
		n_nodes_out = n_nodes;
		n_tris_out = n_tris;
	}
	return best;
}

// ---------------------------------------------------------------------------------------------
// Octahedral maps
__device__ __forceinline__ float gsign(float v) { return v > 0.0f ? 1.0f : (v < 0.0f ? -1.0f : 0.0f); }

// spherePointToOctohedralUV (src/shaders/irradiance.glsl:119-138), GLSL sign() semantics.
__device__ __forceinline__ void oct_uv(float dx, float dy, float dz, float& u, float& v) {
	const float sx = gsign(dx), sy = gsign(dy), sz = gsign(dz);
	const float sum = dx * sx + dy * sy + dz * sz;
	float ox = dx / sum, oy = dy / sum;
	const float oz = dz / sum;
	if(oz < 0.0f) {
		const float ax = fabsf(ox), ay = fabsf(oy);
		ox = sx * (1.0f - ay);
		oy = sy * (1.0f - ax);
	}
	u = ox * 0.5f + 0.5f;
	v = oy * 0.5f + 0.5f;
}

// Direction of blend texel t: octDecode(specializedNormalizeLocalTexelCoord(c)) in FP32
// (irradiance.glsl:107-112, probesUpdate.glsl:33-42, with the reference's constants 0.33333 and
// 0.142857); t < 36: irradiance texel (x, y) = (t % 6, t / 6), else depth texel t - 36 of 14x14.
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

// ---------------------------------------------------------------------------------------------
// Bilinear fetch inside one probe tile (texel coords 1 + (res-2)*uv, centres at k + 0.5).
template <int RES>
__device__ __forceinline__ void tile_coords(float u, float v, int& x0, int& y0, int& x1, int& y1, float& wx, float& wy) {
	const float sx = 1.0f + float(RES - 2) * u - 0.5f;
	const float sy = 1.0f + float(RES - 2) * v - 0.5f;
	const float fx = floorf(sx), fy = floorf(sy);
	wx = sx - fx;
	wy = sy - fy;
	x0 = min(max(int(fx), 0), RES - 1);
	y0 = min(max(int(fy), 0), RES - 1);
	x1 = min(x0 + 1, RES - 1);
	y1 = min(y0 + 1, RES - 1);
}

struct Vec3 {
	float x, y, z;
};
__device__ __forceinline__ Vec3 v3(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ Vec3 operator+(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ Vec3 operator-(Vec3 a, Vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ Vec3 operator*(Vec3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ Vec3 operator*(Vec3 a, Vec3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ float dot3(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ Vec3 normalize3(Vec3 a) {
	const float l = sqrtf(dot3(a, a));
	return {a.x / l, a.y / l, a.z / l};
}

__device__ __forceinline__ Vec3 cross3(Vec3 a, Vec3 b) {
	return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// sin and cos of an FP32 angle by a fixed sequence of FP32 operations, so that a host replica gives
// the same bits: Cody-Waite reduction by the nearest multiple of pi/2 (pi/2 split into a float
// and its remainder), then the Taylor series of sin (to r^11) and cos (to r^12) on |r| <= pi/4,
// whose truncation error is below 1e-10, and the quadrant's sign/swap.
__device__ __forceinline__ void sincos_series(float x, float& s, float& c) {
	const float q = rintf(x * 0.636619772f);
	const float r = (x - q * 1.57079637f) - q * -4.37113883e-08f;
	const float r2 = r * r;
	const float sr = r + r * r2 * (-1.0f / 6.0f + r2 * (1.0f / 120.0f + r2 * (-1.0f / 5040.0f + r2 * (1.0f / 362880.0f +
																									r2 * (-1.0f / 39916800.0f)))));
	const float cr = 1.0f + r2 * (-0.5f + r2 * (1.0f / 24.0f + r2 * (-1.0f / 720.0f + r2 * (1.0f / 40320.0f +
															r2 * (-1.0f / 3628800.0f + r2 * (1.0f / 479001600.0f))))));
	switch(int(q) & 3) {
		case 0: s = sr; c = cr; break;
		case 1: s = cr; c = -sr; break;
		case 2: s = -sr; c = -cr; break;
		default: s = -cr; c = sr; break;
	}
}

// rotateAxis (common.glsl:6-8): mix(dot(axis, p) axis, p, cos a) + cross(axis, p) sin a.
__device__ __forceinline__ Vec3 rotate_axis(Vec3 p, Vec3 axis, float angle) {
	float s, c;
	sincos_series(angle, s, c);
	const Vec3 along = axis * dot3(axis, p);
	return (along * (1.0f - c) + p * c) + cross3(axis, p) * s;
}

// ---- atmospheric sky (src/shaders/sky.glsl:60-127, after O'Neil, GPU Gems 2 ch. 16) ----
// Planet of radius 100 km whose surface lies at y = -100, atmosphere 2.5 km thick, scale height a
// quarter of it; Rayleigh (Kr 0.0025, 1/lambda^4 for lambda = 650/570/475 nm) and Mie (Km 0.001,
// g = -0.99) single scattering integrated with 64 midpoint samples.  The reference's own choices
// are kept: the exit distance of a ray leaving the atmosphere adds the closest-approach offset
// where it should subtract it (and the reverse), a sample whose attenuation overflows is skipped
// without advancing the sample point, and a colour with an infinite channel gives black.
// Sample points are formed as position + (k + 1/2) * step instead of by repeated addition (the
// same point, without the FP32 drift of 64 additions at 1e5 m).  Exponentials use the hardware
// exp2 (__expf: relative error about 6e-8 |x| for the arguments here, |x| < 90).
namespace sky {
constexpr float kInner = 100000.0f;
constexpr float kOuter = 102500.0f;
constexpr float kScale = 1.0f / (kOuter - kInner);
constexpr float kScaleDepth = 0.25f;
constexpr float kPi = 3.1415926538f;  // common.glsl:4
constexpr float kKr = 0.0025f, kKm = 0.0010f, kG = -0.990f;
constexpr int kSamples = 64;

__device__ __forceinline__ float scale_fn(float c) {
	const float x = 1.0f - c;
	return kScaleDepth * __expf(fmaf(x, fmaf(x, fmaf(x, fmaf(x, 5.25f, -6.80f), 3.83f), 0.459f), -0.00287f));
}
// First crossing ahead of the sphere (radius r, centred at the origin) by a ray from outside, -1 if none.
__device__ __forceinline__ float enter_sphere(Vec3 o, Vec3 d, float r) {
	const float a = dot3(d, d), b = dot3(d, o), c = dot3(o, o) - r * r;
	const float g = b * b - a * c;
	if(g > 0.0f) {
		const float dis = (-sqrtf(g) - b) / a;
		if(dis > 0.0f) return dis;
	}
	return -1.0f;
}
// Exit distance from the sphere of a ray inside it, as sky.glsl:33-42 computes it.
__device__ __forceinline__ float leave_sphere(Vec3 o, Vec3 d, float r) {
	const float along = -dot3(o, d);
	const Vec3 pc = o + d * along;
	const float lpc = sqrtf(dot3(pc, pc));
	const float dist = sqrtf(r * r - lpc * lpc);
	const Vec3 off = pc - o;
	const float loff = sqrtf(dot3(off, off));
	return along > 0.0f ? dist - loff : dist + loff;
}
}  // namespace sky

__device__ __forceinline__ Vec3 sky_atmosphere(Vec3 origin, Vec3 dir, Vec3 sun_dir, Vec3 sun_color, float brightness,
											   bool show_sun) {
	using namespace sky;
	sun_color = sun_color * brightness;
	Vec3 pos = v3(origin.x, origin.y + (kInner + 100.0f), origin.z);
	float height = sqrtf(dot3(pos, pos));
	const Vec3 L = normalize3(sun_dir);
	if(fabsf(height - kInner) < 1e-3f) {
		pos = pos + normalize3(pos) * 1e-2f;
		height = sqrtf(dot3(pos, pos));
	}
	const Vec3 zero = v3(0.0f, 0.0f, 0.0f);
	if(height < kOuter) {
		if(height > kInner) {
			const float pd = enter_sphere(pos, dir, kInner);
			if(pd >= 0.0f) {
				const float c = fmaxf(0.1f, dot3(L, normalize3(pos + dir * pd)));
				return v3(c * 0.05f, c * 0.05f, c * 0.05f);
			}
		} else {
			return zero;
		}
		const float ray_depth = leave_sphere(pos, dir, kOuter);
		if(isinf(ray_depth) || isnan(ray_depth)) return zero;
		const float start_depth = __expf(kScale / kScaleDepth * (kInner - height));
		const float start_offset = start_depth * scale_fn(dot3(dir, pos) / height);
		const float seg = ray_depth / float(kSamples);
		const float scaled = seg * kScale;
		const Vec3 step = dir * seg;
		// 1 / lambda^4 per channel, times 4 pi Kr, plus 4 pi Km
		const Vec3 inv_wl = v3(1.0f / (0.650f * 0.650f * 0.650f * 0.650f), 1.0f / (0.570f * 0.570f * 0.570f * 0.570f),
							   1.0f / (0.475f * 0.475f * 0.475f * 0.475f));
		const Vec3 ext = v3(inv_wl.x * (kKr * 4.0f * kPi) + kKm * 4.0f * kPi, inv_wl.y * (kKr * 4.0f * kPi) + kKm * 4.0f * kPi,
							inv_wl.z * (kKr * 4.0f * kPi) + kKm * 4.0f * kPi);
		Vec3 color = zero;
		int k = 0;  // samples advanced so far
		for(int i = 0; i < kSamples; ++i) {
			const Vec3 sp = pos + step * (float(k) + 0.5f);
			// one reciprocal square root instead of a square root and two divisions (the loop is
			// VALU-bound; the sky's accuracy is set by the hardware exponentials, not by these)
			const float h2 = dot3(sp, sp), inv_h = rsqrtf(h2), h = h2 * inv_h;
			const float depth = __expf(kScale / kScaleDepth * (kInner - h));
			const float scatter = start_offset + depth * (scale_fn(dot3(L, sp) * inv_h) - scale_fn(dot3(dir, sp) * inv_h));
			const Vec3 att = v3(__expf(-scatter * ext.x), __expf(-scatter * ext.y), __expf(-scatter * ext.z));
			if(!isfinite(att.x) || !isfinite(att.y) || !isfinite(att.z)) continue;
			color = color + att * (depth * scaled);
			++k;
		}
		const Vec3 secondary = color * sun_color * kKm;
		color = color * inv_wl * sun_color * kKr;
		if(show_sun) {
			const float mc = -dot3(L, dir);
			const float m = fmaxf(1e-3f, 1.0f + kG * kG - 2.0f * kG * mc);
			const float phase = 1.5f * ((1.0f - kG * kG) / (2.0f + kG * kG)) * (1.0f + mc * mc) / (m * sqrtf(m));
			color = color + secondary * phase;
		}
		if(!isinf(color.x) && !isinf(color.y) && !isinf(color.z)) return color;
	} else {
		const float d = enter_sphere(pos, dir, kOuter);
		if(d > 0.0f) {
			const float c = dot3(L, normalize3(pos + dir * d)) * 0.5f;
			return v3(c * 0.5294117647f, c * 0.80784313725f, c * 0.92156862745f);
		}
	}
	return zero;
}

// Weight (visibility x backface x trilinear) and fallback weight of cage probe i of a
// sampleProbes call (irradiance.glsl:164-206); false when the probe is off-grid or disabled.
__device__ __forceinline__ bool cage_probe(const GridConst& g, const uint32_t* __restrict__ states,
										   const float2* __restrict__ dep, Vec3 position, Vec3 normal, Vec3 biased,
										   Vec3 emin, Vec3 alpha, int f0, int f1, int f2, int i, uint32_t& lin,
										   float& weight, float& fbw) {
	const int o0 = i & 1, o1 = (i >> 1) & 1, o2 = (i >> 2) & 1;
	const int p0i = f0 + o0, p1i = f1 + o1, p2i = f2 + o2;
	if(p0i > g.res[0] - 1 || p1i > g.res[1] - 1 || p2i > g.res[2] - 1) return false;
	lin = uint32_t(p0i) + uint32_t(g.res[0]) * (uint32_t(p1i) + uint32_t(g.res[1]) * uint32_t(p2i));
	if(states[lin] == 0u) return false;
	const Vec3 pp = v3(float(p0i) * g.cell[0] + emin.x, float(p1i) * g.cell[1] + emin.y, float(p2i) * g.cell[2] + emin.z);
	const Vec3 dtp = normalize3(pp - position);
	const Vec3 bdir = pp - biased;
	const Vec3 nb = normalize3(bdir);
	float du, dv;
	oct_uv(-nb.x, -nb.y, -nb.z, du, dv);
	const float trx = o0 ? alpha.x : 1.0f - alpha.x;
	const float try_ = o1 ? alpha.y : 1.0f - alpha.y;
	const float trz = o2 ? alpha.z : 1.0f - alpha.z;
	weight = 1.0f;
	const float bfw = fmaxf(0.0001f, (dot3(dtp, normal) + 1.0f) * 0.5f);
	weight *= bfw * bfw + 0.2f;
	fbw = weight;
	// Chebyshev visibility from the bilinear depth moments
	int dx0, dy0, dx1, dy1;
	float dwx, dwy;
	tile_coords<16>(du, dv, dx0, dy0, dx1, dy1, dwx, dwy);
	const float2* dt = dep + size_t(lin) * 256;
	const float2 m00 = dt[dy0 * 16 + dx0], m10 = dt[dy0 * 16 + dx1], m01 = dt[dy1 * 16 + dx0], m11 = dt[dy1 * 16 + dx1];
	const float mean = (m00.x * (1.0f - dwx) + m10.x * dwx) * (1.0f - dwy) + (m01.x * (1.0f - dwx) + m11.x * dwx) * dwy;
	const float m2 = (m00.y * (1.0f - dwx) + m10.y * dwx) * (1.0f - dwy) + (m01.y * (1.0f - dwx) + m11.y * dwx) * dwy;
	const float variance = fabsf(mean * mean - m2);
	const float bdist = sqrtf(dot3(bdir, bdir));
	const float dm = fmaxf(bdist - mean, 0.0001f);
	float cheb = variance / (variance + dm * dm);
	cheb = fmaxf(cheb * cheb * cheb, 0.0f);
	weight *= (bdist <= mean) ? 1.0f : cheb;
	weight = fmaxf(0.000001f, weight);
	if(weight < 0.2f) weight *= weight * weight * (1.0f / (0.2f * 0.2f));
	const float trw = trx * try_ * trz + 0.001f;
	weight *= trw;
	fbw *= trw;
	return true;
}

// sampleProbes (src/shaders/irradiance.glsl:145-237) on the tiled FP32 atlases.
__device__ __forceinline__ Vec3 sample_probes(const GridConst& g, const uint32_t* __restrict__ states,
											  const float4* __restrict__ irr, const float2* __restrict__ dep,
											  Vec3 position, Vec3 normal, Vec3 to_camera) {
	const Vec3 acell = v3(fabsf(g.cell[0]), fabsf(g.cell[1]), fabsf(g.cell[2]));
	const Vec3 emin = v3(g.emin[0], g.emin[1], g.emin[2]);
	const Vec3 gc = v3((position.x - emin.x) / acell.x, (position.y - emin.y) / acell.y, (position.z - emin.z) / acell.z);
	if(gc.x < 0.0f || gc.y < 0.0f || gc.z < 0.0f) return v3(0.0f, 0.0f, 0.0f);
	const Vec3 biased = position + (normal + to_camera) * g.shadow_bias;
	const int f0 = int(gc.x), f1 = int(gc.y), f2 = int(gc.z);
	const Vec3 p0 = v3(float(f0) * g.cell[0] + emin.x, float(f1) * g.cell[1] + emin.y, float(f2) * g.cell[2] + emin.z);
	Vec3 alpha = v3((position.x - p0.x) / acell.x, (position.y - p0.y) / acell.y, (position.z - p0.z) / acell.z);
	alpha = v3(fminf(fmaxf(alpha.x, 0.0f), 1.0f), fminf(fmaxf(alpha.y, 0.0f), 1.0f), fminf(fmaxf(alpha.z, 0.0f), 1.0f));
	float cu, cv;
	oct_uv(normal.x, normal.y, normal.z, cu, cv);
	int cx0, cy0, cx1, cy1;
	float cwx, cwy;
	tile_coords<8>(cu, cv, cx0, cy0, cx1, cy1, cwx, cwy);
	Vec3 fin = v3(0, 0, 0), fb = v3(0, 0, 0);
	float tw = 0.0f, tfb = 0.0f;
	for(int i = 0; i < 8; ++i) {
		uint32_t lin;
		float weight, fbw;
		if(!cage_probe(g, states, dep, position, normal, biased, emin, alpha, f0, f1, f2, i, lin, weight, fbw)) continue;
		const float4* ct = irr + size_t(lin) * 64;
		const float4 c00 = ct[cy0 * 8 + cx0], c10 = ct[cy0 * 8 + cx1], c01 = ct[cy1 * 8 + cx0], c11 = ct[cy1 * 8 + cx1];
		const float cr = (c00.x * (1.0f - cwx) + c10.x * cwx) * (1.0f - cwy) + (c01.x * (1.0f - cwx) + c11.x * cwx) * cwy;
		const float cg = (c00.y * (1.0f - cwx) + c10.y * cwx) * (1.0f - cwy) + (c01.y * (1.0f - cwx) + c11.y * cwx) * cwy;
		const float cb = (c00.z * (1.0f - cwx) + c10.z * cwx) * (1.0f - cwy) + (c01.z * (1.0f - cwx) + c11.z * cwx) * cwy;
		const Vec3 col = v3(sqrtf(cr), sqrtf(cg), sqrtf(cb));
		fin = fin + col * weight;
		tw += weight;
		fb = fb + col * fbw;
		tfb += fbw;
	}
	if(tw > 1e-3f) fin = fin * (1.0f / tw);
	if(tfb > 1e-3f) fb = fb * (1.0f / tfb);
	fin = fin * fin;
	fb = fb * fb;
	const float k = 8.0f * fminf(fmaxf(tw, 0.0f), 1.0f / 8.0f);
	return v3(fb.x * (1.0f - k) + fin.x * k, fb.y * (1.0f - k) + fin.y * k, fb.z * (1.0f - k) + fin.z * k);
}

// The irradiance-independent part of a sampleProbes call, split out for the adjoint: per cage
// slot i (0..7) the probe, its weight and fallback weight; the colour tap footprint (the same
// tile coordinates for all 8 probes: they depend on the normal only); the weight totals.
struct SampleCollect {
	uint32_t mask;  // bit i: cage probe i contributes
	uint32_t lin[8];
	float w[8], fw[8];
	int cx0, cy0, cx1, cy1;
	float cwx, cwy, tw, tfb;
	bool outside;
};

__device__ __forceinline__ void sample_collect(const GridConst& g, const uint32_t* __restrict__ states,
											   const float2* __restrict__ dep, Vec3 position, Vec3 normal,
											   Vec3 to_camera, SampleCollect& c) {
	c.mask = 0;
	c.tw = c.tfb = 0.0f;
	c.outside = true;
	const Vec3 acell = v3(fabsf(g.cell[0]), fabsf(g.cell[1]), fabsf(g.cell[2]));
	const Vec3 emin = v3(g.emin[0], g.emin[1], g.emin[2]);
	const Vec3 gc = v3((position.x - emin.x) / acell.x, (position.y - emin.y) / acell.y, (position.z - emin.z) / acell.z);
	if(gc.x < 0.0f || gc.y < 0.0f || gc.z < 0.0f) return;
	c.outside = false;
	const Vec3 biased = position + (normal + to_camera) * g.shadow_bias;
	const int f0 = int(gc.x), f1 = int(gc.y), f2 = int(gc.z);
	const Vec3 p0 = v3(float(f0) * g.cell[0] + emin.x, float(f1) * g.cell[1] + emin.y, float(f2) * g.cell[2] + emin.z);
	Vec3 alpha = v3((position.x - p0.x) / acell.x, (position.y - p0.y) / acell.y, (position.z - p0.z) / acell.z);
	alpha = v3(fminf(fmaxf(alpha.x, 0.0f), 1.0f), fminf(fmaxf(alpha.y, 0.0f), 1.0f), fminf(fmaxf(alpha.z, 0.0f), 1.0f));
	float cu, cv;
	oct_uv(normal.x, normal.y, normal.z, cu, cv);
	tile_coords<8>(cu, cv, c.cx0, c.cy0, c.cx1, c.cy1, c.cwx, c.cwy);
#pragma unroll
	for(int i = 0; i < 8; ++i) {
		c.lin[i] = 0;
		c.w[i] = c.fw[i] = 0.0f;
		uint32_t lin;
		float w, fw;
		if(!cage_probe(g, states, dep, position, normal, biased, emin, alpha, f0, f1, f2, i, lin, w, fw)) continue;
		c.mask |= 1u << i;
		c.lin[i] = lin;
		c.w[i] = w;
		c.fw[i] = fw;
		c.tw += w;
		c.tfb += fw;
	}
}

__device__ __forceinline__ float tap3(const float4* __restrict__ irr, const SampleCollect& c, uint32_t lin, int ch) {
	const float4* ct = irr + size_t(lin) * 64;
	const float4 c00 = ct[c.cy0 * 8 + c.cx0], c10 = ct[c.cy0 * 8 + c.cx1], c01 = ct[c.cy1 * 8 + c.cx0],
				 c11 = ct[c.cy1 * 8 + c.cx1];
	const float a00 = ch == 0 ? c00.x : (ch == 1 ? c00.y : c00.z), a10 = ch == 0 ? c10.x : (ch == 1 ? c10.y : c10.z);
	const float a01 = ch == 0 ? c01.x : (ch == 1 ? c01.y : c01.z), a11 = ch == 0 ? c11.x : (ch == 1 ? c11.y : c11.z);
	return (a00 * (1.0f - c.cwx) + a10 * c.cwx) * (1.0f - c.cwy) + (a01 * (1.0f - c.cwx) + a11 * c.cwx) * c.cwy;
}

__device__ __forceinline__ Vec3 sample_eval(const SampleCollect& c, const float4* __restrict__ irr) {
	if(c.outside) return v3(0.0f, 0.0f, 0.0f);
	float out[3];
#pragma unroll
	for(int ch = 0; ch < 3; ++ch) {
		float fin = 0.0f, fb = 0.0f;
#pragma unroll
		for(int i = 0; i < 8; ++i) {
			if(!(c.mask & (1u << i))) continue;
			const float s = sqrtf(tap3(irr, c, c.lin[i], ch));
			fin += s * c.w[i];
			fb += s * c.fw[i];
		}
		if(c.tw > 1e-3f) fin *= 1.0f / c.tw;
		if(c.tfb > 1e-3f) fb *= 1.0f / c.tfb;
		const float k = 8.0f * fminf(fmaxf(c.tw, 0.0f), 1.0f / 8.0f);
		out[ch] = fb * fb * (1.0f - k) + fin * fin * k;
	}
	return v3(out[0], out[1], out[2]);
}

// Order-independent (hence deterministic) accumulation: contributions are added as 64-bit
// fixed-point integers with 2^-50 resolution.
constexpr double kFixScale = 1125899906842624.0;  // 2^50
__device__ __forceinline__ long long to_fixed(float x) {
	const float s = x * 1125899906842624.0f;
	const float lim = 9.0e18f;
	return __float2ll_rn(fminf(fmaxf(s, -lim), lim));
}

// Adjoint of sample_eval with respect to the irradiance atlas.  acc[i*192 + texel*3 + ch] is the
// fixed-point gradient accumulator of cage slot i's tile (64 texels x rgb).  sqrt'(x) is taken at
// max(x, 1e-12).
__device__ __forceinline__ void sample_adjoint(const SampleCollect& c, const float4* __restrict__ irr, Vec3 g_out,
											   unsigned long long* acc) {
	if(c.outside || c.mask == 0u) return;
	const float k = 8.0f * fminf(fmaxf(c.tw, 0.0f), 1.0f / 8.0f);
	const float inv_w = c.tw > 1e-3f ? 1.0f / c.tw : 1.0f;
	const float inv_fb = c.tfb > 1e-3f ? 1.0f / c.tfb : 1.0f;
	const float b00 = (1.0f - c.cwx) * (1.0f - c.cwy), b10 = c.cwx * (1.0f - c.cwy), b01 = (1.0f - c.cwx) * c.cwy,
				b11 = c.cwx * c.cwy;
	const int t00 = c.cy0 * 8 + c.cx0, t10 = c.cy0 * 8 + c.cx1, t01 = c.cy1 * 8 + c.cx0, t11 = c.cy1 * 8 + c.cx1;
#pragma unroll
	for(int ch = 0; ch < 3; ++ch) {
		const float go = ch == 0 ? g_out.x : (ch == 1 ? g_out.y : g_out.z);
		float s[8];
		float fin = 0.0f, fb = 0.0f;
#pragma unroll
		for(int i = 0; i < 8; ++i) {
			s[i] = 0.0f;
			if(!(c.mask & (1u << i))) continue;
			s[i] = sqrtf(tap3(irr, c, c.lin[i], ch));
			fin += s[i] * c.w[i];
			fb += s[i] * c.fw[i];
		}
		fin *= inv_w;
		fb *= inv_fb;
#pragma unroll
		for(int i = 0; i < 8; ++i) {
			if(!(c.mask & (1u << i))) continue;
			const float dds = (1.0f - k) * 2.0f * fb * c.fw[i] * inv_fb + k * 2.0f * fin * c.w[i] * inv_w;
			const float gcol = go * dds * 0.5f / sqrtf(fmaxf(s[i] * s[i], 1e-12f));
			unsigned long long* a = acc + i * 192 + ch;
			atomicAdd(a + t00 * 3, (unsigned long long)to_fixed(gcol * b00));
			atomicAdd(a + t10 * 3, (unsigned long long)to_fixed(gcol * b10));
			atomicAdd(a + t01 * 3, (unsigned long long)to_fixed(gcol * b01));
			atomicAdd(a + t11 * 3, (unsigned long long)to_fixed(gcol * b11));
		}
	}
}

// pbrMetallicRoughness (src/shaders/pbrMetallicRoughness.glsl:43-84)
__device__ __forceinline__ Vec3 pbr_metallic_roughness(Vec3 n, Vec3 v, Vec3 light_color, Vec3 light_dir, Vec3 albedo,
													   float metal, float rough) {
	const float f0 = 0.04f;
	const Vec3 diffuse = albedo * ((1.0f - f0) * (1.0f - metal));
	const float alpha = rough * rough;
	const Vec3 spec = v3(f0 * (1.0f - metal) + albedo.x * metal, f0 * (1.0f - metal) + albedo.y * metal,
						 f0 * (1.0f - metal) + albedo.z * metal);
	const float reflectance = fmaxf(fmaxf(spec.x, spec.y), spec.z);
	const float r90 = fminf(fmaxf(reflectance * 25.0f, 0.0f), 1.0f);
	const Vec3 l = normalize3(light_dir);
	const Vec3 h = normalize3(l + v);
	const float NdotL = fminf(fmaxf(dot3(n, l), 0.001f), 1.0f);
	const float NdotV = fminf(fmaxf(fabsf(dot3(n, v)), 0.001f), 1.0f);
	const float NdotH = fminf(fmaxf(dot3(n, h), 0.0f), 1.0f);
	const float VdotH = fminf(fmaxf(dot3(v, h), 0.0f), 1.0f);
	const float om = fminf(fmaxf(1.0f - VdotH, 0.0f), 1.0f);
	const float fw = om * om * om * om * om;
	const Vec3 F = v3(spec.x + (r90 - spec.x) * fw, spec.y + (r90 - spec.y) * fw, spec.z + (r90 - spec.z) * fw);
	const float a2 = alpha * alpha;
	const float attL = 2.0f * NdotL / (NdotL + sqrtf(a2 + (1.0f - a2) * (NdotL * NdotL)));
	const float attV = 2.0f * NdotV / (NdotV + sqrtf(a2 + (1.0f - a2) * (NdotV * NdotV)));
	const float G = attL * attV;
	const float ah = NdotH * alpha;
	const float k = alpha / ((1.0f - NdotH * NdotH) + ah * ah);
	const float D = fminf(fmaxf(k * k * (1.0f / kPiF), 0.0f), 4.0f);
	const float sp = G * D / (4.0f * NdotL * NdotV);
	const float inv_pi = 1.0f / kPiF;
	return v3(NdotL * light_color.x * ((1.0f - F.x) * diffuse.x * inv_pi + F.x * sp),
			  NdotL * light_color.y * ((1.0f - F.y) * diffuse.y * inv_pi + F.y * sp),
			  NdotL * light_color.z * ((1.0f - F.z) * diffuse.z * inv_pi + F.z * sp));
}

}  // namespace ddgi
