// TEST INFRASTRUCTURE — the CPU oracle of the DDGI probe update.
//
// This is a restatement, in C++ on the host, of VulkanExp's probe-update shaders; it exists to
// check the HIP library and to serve as the timed CPU baseline.  Only tests/, the smoke test and
// bench.py's cpu_baseline leg load it.  The product library never links or calls it.
//
// Precision model
//   * Ray/triangle decisions (which triangle, t, barycentrics, facing, shadow visibility) use an
//     FP32 bit-replica of the watertight test (Woop, Benthin, Wald, JCGT 2(1) 2013) with a fixed
//     operation order and no FMA contraction (built with -ffp-contract=off); the same decisions are
//     what the GPU makes, so hit ids compare bit-exactly.
//   * Everything continuous (hit position, normals, shading, probe sampling, blend, hysteresis) is
//     templated on the real type T and instantiated in double (the oracle) and float (context).
//
// Parity status of each part is recorded in tests/ (see the pinning tests named in each header).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/ddgi.h"

namespace orc {

// src/shaders/common.glsl:4 — the reference's value of pi (H9: not M_PI).
constexpr double kPi = 3.1415926538;

template <class T>
struct V3 {
	T x, y, z;
	V3() : x(0), y(0), z(0) {}
	V3(T a, T b, T c) : x(a), y(b), z(c) {}
	template <class U>
	explicit V3(const V3<U>& o) : x(T(o.x)), y(T(o.y)), z(T(o.z)) {}
	T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
	T& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline V3<T> operator*(V3<T> a, V3<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class T> inline V3<T> operator/(V3<T> a, V3<T> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator*(T s, V3<T> a) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline T length(V3<T> a) { return std::sqrt(dot(a, a)); }
template <class T> inline V3<T> normalize(V3<T> a) { T l = length(a); return {a.x / l, a.y / l, a.z / l}; }
template <class T> inline V3<T> vabs(V3<T> a) { return {std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)}; }
template <class T> inline T clampT(T v, T lo, T hi) { return v < lo ? lo : (v > hi ? hi : v); }
template <class T> inline T mixT(T a, T b, T w) { return a * (T(1) - w) + b * w; }
template <class T> inline V3<T> mix3(V3<T> a, V3<T> b, T w) { return {mixT(a.x, b.x, w), mixT(a.y, b.y, w), mixT(a.z, b.z, w)}; }
// GLSL sign(): 0 at 0 (src/shaders/irradiance.glsl:123 uses it; H3)
template <class T> inline T gsign(T v) { return v > T(0) ? T(1) : (v < T(0) ? T(-1) : T(0)); }
// signNotZero (src/shaders/irradiance.glsl:88-90)
template <class T> inline T sign_not_zero(T v) { return v >= T(0) ? T(1) : T(-1); }

// ---------------------------------------------------------------------------------------------
// Probe-grid index maps (src/shaders/irradiance.glsl:7-38)
struct Grid {
	ddgi_grid_params p;
	uint32_t rx() const { return uint32_t(p.resolution[0]); }
	uint32_t ry() const { return uint32_t(p.resolution[1]); }
	uint32_t rz() const { return uint32_t(p.resolution[2]); }
	uint64_t count() const { return uint64_t(rx()) * ry() * rz(); }
	void grid_index(uint32_t linear, int g[3]) const {  // probeLinearIndexToGridIndex :7-13
		g[0] = int(linear % rx());
		g[1] = int((linear % (rx() * ry())) / rx());
		g[2] = int(linear / (rx() * ry()));
	}
	uint32_t linear(const int g[3]) const { return uint32_t(g[0]) + rx() * uint32_t(g[1]) + rx() * ry() * uint32_t(g[2]); }  // :15-17
	template <class T> V3<T> cell() const {  // probeGridCellSize :19-21
		return {(T(p.extent_max[0]) - T(p.extent_min[0])) / T(p.resolution[0] - 1),
				(T(p.extent_max[1]) - T(p.extent_min[1])) / T(p.resolution[1] - 1),
				(T(p.extent_max[2]) - T(p.extent_min[2])) / T(p.resolution[2] - 1)};
	}
	template <class T> V3<T> position(const int g[3]) const {  // probeIndexToWorldPosition :23-26
		V3<T> c = cell<T>();
		return {T(g[0]) * c.x + T(p.extent_min[0]), T(g[1]) * c.y + T(p.extent_min[1]), T(g[2]) * c.z + T(p.extent_min[2])};
	}
	// probeIndexTo{Color,Depth}UVOffset :32-38 — texel origin of the probe tile in the 2D atlas image
	void tile_origin(const int g[3], uint32_t res, uint32_t& ox, uint32_t& oy) const {
		ox = res * (uint32_t(g[1]) * rx() + uint32_t(g[0]));
		oy = res * uint32_t(g[2]);
	}
	uint32_t image_width(uint32_t res) const { return res * rx() * ry(); }
	uint32_t image_height(uint32_t res) const { return res * rz(); }
};

// sphericalFibonacci (src/shaders/irradiance.glsl:52-64)
template <class T>
inline V3<T> spherical_fibonacci(T i, T n) {
	const T PHI = std::sqrt(T(5)) * T(0.5) + T(0.5);
	T a = i * (PHI - T(1));
	T phi = T(2) * T(kPi) * (a - std::floor(a));
	T cos_t = T(1) - (T(2) * i + T(1)) * (T(1) / n);
	T sin_t = std::sqrt(clampT(T(1) - cos_t * cos_t, T(0), T(1)));
	return {std::cos(phi) * sin_t, std::sin(phi) * sin_t, cos_t};
}

// octDecode (src/shaders/irradiance.glsl:107-112)
template <class T>
inline V3<T> oct_decode(T ox, T oy) {
	V3<T> v(ox, oy, T(1) - std::fabs(ox) - std::fabs(oy));
	if(v.z < T(0)) {
		T nx = (T(1) - std::fabs(v.y)) * sign_not_zero(v.x);
		T ny = (T(1) - std::fabs(v.x)) * sign_not_zero(v.y);
		v.x = nx;
		v.y = ny;
	}
	return normalize(v);
}

// octEncode (src/shaders/irradiance.glsl:97-103) — not on the probe path; kept for closed-form tests.
template <class T>
inline void oct_encode(V3<T> v, T& ox, T& oy) {
	T l1 = std::fabs(v.x) + std::fabs(v.y) + std::fabs(v.z);
	ox = v.x * (T(1) / l1);
	oy = v.y * (T(1) / l1);
	if(v.z < T(0)) {
		T nx = (T(1) - std::fabs(oy)) * sign_not_zero(ox);
		T ny = (T(1) - std::fabs(ox)) * sign_not_zero(oy);
		ox = nx;
		oy = ny;
	}
}

// spherePointToOctohedralUV (src/shaders/irradiance.glsl:119-138), with GLSL sign() (H3).
template <class T>
inline void sphere_to_oct_uv(V3<T> d, T& u, T& v) {
	V3<T> octant(gsign(d.x), gsign(d.y), gsign(d.z));
	T sum = dot(d, octant);
	V3<T> o(d.x / sum, d.y / sum, d.z / sum);
	if(o.z < T(0)) {
		V3<T> a = vabs(o);
		T nx = octant.x * (T(1) - a.y);
		T ny = octant.y * (T(1) - a.x);
		o.x = nx;
		o.y = ny;
	}
	u = o.x * T(0.5) + T(0.5);
	v = o.y * T(0.5) + T(0.5);
}

// specializedNormalizeLocalTexelCoord (src/shaders/probesUpdate.glsl:33-42): the reference's
// constants 0.33333 and 0.142857, not 1/3 and 1/7.
template <class T>
inline T texel_coord_irradiance(int c) { return T(0.33333) * (T(c) - T(2.5)); }
template <class T>
inline T texel_coord_depth(int c) { return T(0.142857) * (T(c) - T(6.5)); }

// ---------------------------------------------------------------------------------------------
// Scene: flattened world-space FP32 triangles with (instance, primitive) ids, plus the reference
// records the closest-hit shader reads (src/shaders/closesthit.glsl:147-159).
struct Scene {
	// flattened, instance-major (TLAS order), global id = index into these arrays
	std::vector<float> tri;          // ntri * 9 world positions (FP32)
	std::vector<uint32_t> tri_inst;  // instance index
	std::vector<uint32_t> tri_prim;  // primitive index within the instance's mesh
	std::vector<uint8_t> tri_flip;   // facing inverted (negative-determinant transform XOR flag)
	std::vector<uint32_t> inst_mask;
	std::vector<ddgi_vertex> vertices;
	std::vector<uint32_t> indices;
	std::vector<ddgi_offset_entry> offsets;
	std::vector<ddgi_instance> instances;
	std::vector<ddgi_material> materials;
	std::vector<double> inst_normal;  // 9 per instance: transpose(inverse(M_linear)) in FP64

	// oracle BVH (binned SAH, independent of the library's builder)
	struct Node {
		float lo[3], hi[3];
		uint32_t first;  // leaf: first index into order; inner: left child index (right = left + 1)
		uint32_t count;  // > 0: leaf
	};
	std::vector<Node> nodes;
	std::vector<uint32_t> order;  // triangle ids in leaf order
	void build_bvh();
	size_t tri_count() const { return tri_inst.size(); }
};

// FP32 ray with the watertight-test precomputation.
struct Ray32 {
	float o[3], d[3];
	float tmin, tmax;
	int kx, ky, kz;
	float sx, sy, sz;
};
void ray32_prepare(Ray32& r);

struct Hit32 {
	float t = -1.0f;
	float u = 0.0f, v = 0.0f;  // barycentric weights of vertex 1 and vertex 2
	uint32_t tri = 0xFFFFFFFFu;
	bool back = false;
	bool hit() const { return tri != 0xFFFFFFFFu; }
};

// Intersect one triangle; on acceptance returns true with t and the edge functions.
bool tri_intersect32(const Ray32& r, const float* v0, const float* v1, const float* v2, float& t, float& U, float& V,
					 float& W, float& det);

// Closest hit (ties: lowest global triangle id) or any hit, through the oracle BVH or by brute
// force over every triangle. `stats` (optional) counts node and triangle visits.
// cull: ignore back-facing triangles (the G-buffer pass culls them)
Hit32 trace32(const Scene& s, const Ray32& r, uint32_t mask, bool any_hit, bool brute, uint64_t* stats = nullptr,
			  bool cull = false);

// ---------------------------------------------------------------------------------------------
// Atlases in the reference's image layout (2D, row-major), in double.
struct Atlas {
	uint32_t width = 0, height = 0, channels = 0;
	const double* data = nullptr;
	double at(uint32_t x, uint32_t y, uint32_t c) const { return data[(uint64_t(y) * width + x) * channels + c]; }
};

struct Light {
	V3<double> dir;    // toward the light (src/Light.hpp:7)
	V3<double> color;  // radiance (rgb)
	double sky_brightness = 10.0;  // color.a: the atmospheric sky's sun scale (src/shaders/miss.rmiss:22)
};

// What misses see: the constant radiance, or (model 1) the atmospheric sky of
// src/shaders/sky.glsl:60-127 with the light's direction and colour * sky_brightness.
struct SkyCfg {
	V3<double> constant;
	int model = 0;
};

// sky() of src/shaders/sky.glsl:60-127 in double precision (see ddgi_oracle.cpp).
V3<double> sky_atmosphere(V3<double> origin, V3<double> dir, V3<double> sun_dir, V3<double> sun_color, double brightness,
						  bool show_sun);

}  // namespace orc
