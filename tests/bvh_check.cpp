// Host-side check of the BVH builder (vulkanexp_amd/csrc/bvh.cpp), compiled and run by
// tests/test_bvh_host.py:   bvh_check SCENE RAYS SEED
// Builds the device layout from a .scene file and checks
//  * structure: every child code names an existing node or a leaf range inside the triangle
//    slots, every node is reached once, the wide depth is within the traversal's limit;
//  * pieces: every slot's piece box lies inside its triangle's box, and the slot's triangle record
//    is its triangle (global id, instance, vertices);
//  * closest hits: RAYS random rays (origins inside the scene bounds, directions uniform on the
//    sphere) traced through the tree give the same nearest t as a loop over every triangle (both
//    with the same double-precision test), and any-hit rays the same decision.
// Prints one summary line; exit status 0 when every check holds.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../vulkanexp_amd/csrc/bvh.hpp"

namespace ddgi {
namespace host {
void record_error(const std::string&) {}  // the library's error slot (api.hip), not linked here
}  // namespace host
}  // namespace ddgi

using namespace ddgi;

namespace {

uint32_t f2u(float f) {
	uint32_t u;
	std::memcpy(&u, &f, 4);
	return u;
}

// Moller-Trumbore in double; t in (1e-6, tmax)
bool hit(const float* v0, const float* v1, const float* v2, const double o[3], const double d[3], double tmax, double& t) {
	double e1[3], e2[3], s[3], p[3], q[3];
	for(int k = 0; k < 3; ++k) {
		e1[k] = double(v1[k]) - v0[k];
		e2[k] = double(v2[k]) - v0[k];
		s[k] = o[k] - v0[k];
	}
	p[0] = d[1] * e2[2] - d[2] * e2[1];
	p[1] = d[2] * e2[0] - d[0] * e2[2];
	p[2] = d[0] * e2[1] - d[1] * e2[0];
	const double det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
	if(std::fabs(det) < 1e-30) return false;
	const double u = (s[0] * p[0] + s[1] * p[1] + s[2] * p[2]) / det;
	if(u < 0.0 || u > 1.0) return false;
	q[0] = s[1] * e1[2] - s[2] * e1[1];
	q[1] = s[2] * e1[0] - s[0] * e1[2];
	q[2] = s[0] * e1[1] - s[1] * e1[0];
	const double v = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) / det;
	if(v < 0.0 || u + v > 1.0) return false;
	t = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) / det;
	return t > 1e-6 && t < tmax;
}

double trace_tree(const BvhHost& b, const double o[3], const double d[3], double tmax, bool any) {
	std::vector<uint32_t> stack{0};
	double best = tmax;
	while(!stack.empty()) {
		const uint32_t code = stack.back();
		stack.pop_back();
		if(code & kLeafFlag) {
			const uint32_t first = (code & ~kLeafFlag) >> 3, cnt = (code & 7u) + 1u;
			for(uint32_t i = 0; i < cnt; ++i) {
				const float* r = &b.tris[size_t(first + i) * 12];
				double t;
				if(hit(r, r + 4, r + 8, o, d, best, t)) {
					best = t;
					if(any) return best;
				}
			}
			continue;
		}
		const float* n = &b.nodes[size_t(code) * kNodeFloats];
		for(uint32_t k = 0; k < kWide; ++k) {
			const uint32_t c = f2u(n[kCodeRow + k]);
			if(c == kEmptyChild) continue;
			double t0 = 0.0, t1 = best;
			for(int a = 0; a < 3; ++a) {
				const double lo = double(n[(2 * a) * kWide + k]), hi = double(n[(2 * a + 1) * kWide + k]);
				if(d[a] == 0.0) {
					if(o[a] < lo || o[a] > hi) t0 = 1.0, t1 = 0.0;
					continue;
				}
				double ta = (lo - o[a]) / d[a], tb = (hi - o[a]) / d[a];
				if(ta > tb) std::swap(ta, tb);
				t0 = std::max(t0, ta);
				t1 = std::min(t1, tb);
			}
			if(t0 <= t1) stack.push_back(c);
		}
	}
	return best;
}

}  // namespace

int main(int argc, char** argv) {
	if(argc != 4) {
		std::fprintf(stderr, "usage: %s SCENE RAYS SEED\n", argv[0]);
		return 2;
	}
	void* sc = nullptr;
	if(ddgi_scene_load(argv[1], &sc) != DDGI_OK) {
		std::fprintf(stderr, "cannot load %s\n", argv[1]);
		return 2;
	}
	ddgi_scene_tables T;
	ddgi_scene_get_tables(sc, &T);
	BvhHost b;
	const char* err = nullptr;
	if(!build_bvh(T.vertices, T.vertex_count, T.indices, T.index_count, T.offsets, T.offset_count, T.instances,
				  T.instance_count, b, &err)) {
		std::fprintf(stderr, "build failed: %s\n", err ? err : "?");
		return 1;
	}
	size_t tris = 0;
	for(size_t k = 0; k < T.instance_count; ++k) tris += T.instances[k].triangle_count;
	// the flattened world triangles, by global id (instance-major), as the builder flattens them
	std::vector<float> world(tris * 9);
	std::vector<uint32_t> inst_of(tris);
	for(size_t k = 0, g = 0; k < T.instance_count; ++k) {
		const ddgi_instance& in = T.instances[k];
		const ddgi_offset_entry& oe = T.offsets[in.custom_index];
		for(uint32_t p = 0; p < in.triangle_count; ++p, ++g) {
			for(int v = 0; v < 3; ++v)
				transform_point(in.transform, T.vertices[oe.vertex_offset + T.indices[oe.index_offset + 3 * p + v]].pos,
								&world[g * 9 + v * 3]);
			inst_of[g] = uint32_t(k);
		}
	}
	int bad = 0;
	// structure
	std::vector<uint32_t> seen(b.node_count, 0);
	std::vector<uint32_t> slot_seen(b.tri_count, 0);
	seen[0] = 1;
	for(uint32_t w = 0; w < b.node_count; ++w)
		for(uint32_t k = 0; k < kWide; ++k) {
			const uint32_t c = f2u(b.nodes[size_t(w) * kNodeFloats + kCodeRow + k]);
			if(c == kEmptyChild) continue;
			if(c & kLeafFlag) {
				const uint32_t first = (c & ~kLeafFlag) >> 3, cnt = (c & 7u) + 1u;
				if(size_t(first) + cnt > b.tri_count) ++bad;
				else
					for(uint32_t i = 0; i < cnt; ++i) ++slot_seen[first + i];
			} else if(c >= b.node_count || c == 0) {
				++bad;
			} else {
				++seen[c];
			}
		}
	for(uint32_t w = 0; w < b.node_count; ++w) bad += seen[w] != 1;
	for(uint32_t s = 0; s < b.tri_count; ++s) bad += slot_seen[s] != 1;
	bad += b.wide_depth > kMaxWideDepth;
	const int structure_bad = bad;
	// pieces
	std::vector<uint32_t> pieces(tris, 0);
	for(uint32_t s = 0; s < b.tri_count; ++s) {
		const float* r = &b.tris[size_t(s) * 12];
		const uint32_t g = f2u(r[3]);
		if(g >= tris || f2u(r[7]) != inst_of[g]) {
			++bad;
			continue;
		}
		++pieces[g];
		for(int v = 0; v < 3; ++v)
			for(int a = 0; a < 3; ++a) bad += r[v * 4 + a] != world[size_t(g) * 9 + v * 3 + a];
		for(int a = 0; a < 3; ++a) {
			float lo = world[size_t(g) * 9 + a], hi = lo;
			for(int v = 1; v < 3; ++v) {
				lo = std::min(lo, world[size_t(g) * 9 + v * 3 + a]);
				hi = std::max(hi, world[size_t(g) * 9 + v * 3 + a]);
			}
			bad += b.slot_box[size_t(s) * 6 + a] < lo || b.slot_box[size_t(s) * 6 + 3 + a] > hi ||
				   b.slot_box[size_t(s) * 6 + a] > b.slot_box[size_t(s) * 6 + 3 + a];
		}
	}
	size_t split = 0;
	for(size_t g = 0; g < tris; ++g) {
		bad += pieces[g] == 0;
		split += pieces[g] > 1;
	}
	const int piece_bad = bad - structure_bad;
	// closest and any hits against every triangle
	std::mt19937_64 rng(std::strtoull(argv[3], nullptr, 10));
	std::uniform_real_distribution<double> U(0.0, 1.0);
	const long rays = std::atol(argv[2]);
	long hit_bad = 0, hits = 0;
	for(long r = 0; r < rays; ++r) {
		double o[3], d[3];
		for(int a = 0; a < 3; ++a) o[a] = T.bounds_min[a] + (double(T.bounds_max[a]) - T.bounds_min[a]) * U(rng);
		const double z = 2.0 * U(rng) - 1.0, ph = 2.0 * M_PI * U(rng), s = std::sqrt(1.0 - z * z);
		d[0] = std::cos(ph) * s;
		d[1] = std::sin(ph) * s;
		d[2] = z;
		double brute = 1e30, t;
		for(size_t g = 0; g < tris; ++g)
			if(hit(&world[g * 9], &world[g * 9 + 3], &world[g * 9 + 6], o, d, brute, t)) brute = t;
		const double tree = trace_tree(b, o, d, 1e30, false);
		hit_bad += tree != brute;
		hits += brute < 1e30;
		const bool any = trace_tree(b, o, d, 1e30, true) < 1e30;
		hit_bad += any != (brute < 1e30);
	}
	std::printf("triangles %zu slots %u (split %zu) nodes %u wide_depth %u | structure errors %d, piece errors %d, "
				"hit mismatches %ld of %ld rays (%ld hits)\n",
				tris, b.tri_count, split, b.node_count, b.wide_depth, structure_bad, piece_bad, hit_bad, rays, hits);
	return (bad == 0 && hit_bad == 0) ? 0 : 1;
}
