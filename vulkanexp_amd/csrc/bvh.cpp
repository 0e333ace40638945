// Binned-SAH BVH builder (host).  See bvh.hpp for the device layout.
#include "bvh.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

namespace ddgi {

namespace {

constexpr int kBins = 32;
constexpr double kCostTraverse = 1.0;  // one 64-byte node fetch + 2 slab tests
constexpr double kCostTriangle = 1.0;  // one 48-byte triangle fetch + watertight test

struct Aabb {
	float lo[3], hi[3];
	Aabb() {
		for(int k = 0; k < 3; ++k) {
			lo[k] = std::numeric_limits<float>::max();
			hi[k] = -std::numeric_limits<float>::max();
		}
	}
	void grow(const float* p) {
		for(int k = 0; k < 3; ++k) {
			lo[k] = std::min(lo[k], p[k]);
			hi[k] = std::max(hi[k], p[k]);
		}
	}
	void grow(const Aabb& b) {
		for(int k = 0; k < 3; ++k) {
			lo[k] = std::min(lo[k], b.lo[k]);
			hi[k] = std::max(hi[k], b.hi[k]);
		}
	}
	bool empty() const { return lo[0] > hi[0]; }
	double half_area() const {
		if(empty()) return 0.0;
		double e0 = double(hi[0]) - lo[0], e1 = double(hi[1]) - lo[1], e2 = double(hi[2]) - lo[2];
		return e0 * e1 + e1 * e2 + e2 * e0;
	}
};

struct BuildNode {
	Aabb box;
	int32_t left = -1, right = -1;  // children (inner)
	uint32_t first = 0, count = 0;  // leaf range in `order`
	bool leaf() const { return left < 0; }
};

struct Builder {
	const std::vector<Aabb>& tri_box;
	const std::vector<float>& centroid;  // 3 per triangle
	std::vector<uint32_t> order;
	std::vector<BuildNode> nodes;
	uint32_t max_depth = 0;

	Builder(const std::vector<Aabb>& tb, const std::vector<float>& c) : tri_box(tb), centroid(c) {}

	int bin_of(uint32_t tri, int axis, float cmin, double scale) const {
		int b = int((double(centroid[tri * 3 + axis]) - cmin) * scale);
		return std::min(std::max(b, 0), kBins - 1);
	}

	void build(uint32_t n) {
		order.resize(n);
		for(uint32_t i = 0; i < n; ++i) order[i] = i;
		nodes.reserve(2 * size_t(n) / 2 + 8);
		nodes.emplace_back();
		struct Job {
			uint32_t node, begin, end, depth;
		};
		std::vector<Job> stack{{0, 0, n, 0}};
		while(!stack.empty()) {
			Job j = stack.back();
			stack.pop_back();
			max_depth = std::max(max_depth, j.depth);
			Aabb box, cbox;
			for(uint32_t i = j.begin; i < j.end; ++i) {
				box.grow(tri_box[order[i]]);
				cbox.grow(&centroid[order[i] * 3]);
			}
			nodes[j.node].box = box;
			const uint32_t cnt = j.end - j.begin;
			const double leaf_cost = kCostTriangle * cnt;
			double best = std::numeric_limits<double>::max();
			int best_axis = -1, best_bin = -1;
			const double parent_area = std::max(box.half_area(), 1e-30);
			for(int ax = 0; ax < 3 && cnt > 1; ++ax) {
				const double ext = double(cbox.hi[ax]) - cbox.lo[ax];
				if(!(ext > 0.0)) continue;
				const double scale = kBins / ext;
				Aabb bb[kBins];
				uint32_t bc[kBins] = {};
				for(uint32_t i = j.begin; i < j.end; ++i) {
					const uint32_t t = order[i];
					const int b = bin_of(t, ax, cbox.lo[ax], scale);
					bb[b].grow(tri_box[t]);
					++bc[b];
				}
				double ra[kBins];
				uint32_t rc[kBins];
				Aabb acc;
				uint32_t c = 0;
				for(int b = kBins - 1; b > 0; --b) {
					acc.grow(bb[b]);
					c += bc[b];
					ra[b] = acc.half_area();
					rc[b] = c;
				}
				Aabb lacc;
				uint32_t lc = 0;
				for(int b = 0; b < kBins - 1; ++b) {
					lacc.grow(bb[b]);
					lc += bc[b];
					if(lc == 0 || rc[b + 1] == 0) continue;
					const double cost = kCostTraverse + kCostTriangle * (lacc.half_area() * lc + ra[b + 1] * rc[b + 1]) / parent_area;
					if(cost < best) {
						best = cost;
						best_axis = ax;
						best_bin = b;
					}
				}
			}
			const bool can_leaf = cnt <= kMaxLeafSize;
			if(cnt == 1 || (can_leaf && leaf_cost <= best) || (best_axis < 0 && can_leaf)) {
				nodes[j.node].first = j.begin;
				nodes[j.node].count = cnt;
				continue;
			}
			uint32_t mid;
			if(best_axis < 0) {
				mid = j.begin + cnt / 2;  // coincident centroids: split by order
			} else {
				const double scale = kBins / (double(cbox.hi[best_axis]) - cbox.lo[best_axis]);
				const float cmin = cbox.lo[best_axis];
				auto it = std::partition(order.begin() + j.begin, order.begin() + j.end,
										 [&](uint32_t t) { return bin_of(t, best_axis, cmin, scale) <= best_bin; });
				mid = uint32_t(it - order.begin());
				if(mid == j.begin || mid == j.end) mid = j.begin + cnt / 2;
			}
			const int32_t l = int32_t(nodes.size());
			nodes.emplace_back();
			nodes.emplace_back();
			nodes[j.node].left = l;
			nodes[j.node].right = l + 1;
			stack.push_back({uint32_t(l + 1), mid, j.end, j.depth + 1});
			stack.push_back({uint32_t(l), j.begin, mid, j.depth + 1});
		}
	}
};

inline uint32_t f2u(float f) {
	uint32_t u;
	std::memcpy(&u, &f, 4);
	return u;
}
inline float u2f(uint32_t u) {
	float f;
	std::memcpy(&f, &u, 4);
	return f;
}

}  // namespace

float pad_down(float v) {
	double d = double(v) - (std::fabs(double(v)) * 1e-6 + 1e-6);
	float f = float(d);
	if(double(f) > d) f = std::nextafter(f, -std::numeric_limits<float>::infinity());
	return f;
}

float pad_up(float v) {
	double d = double(v) + (std::fabs(double(v)) * 1e-6 + 1e-6);
	float f = float(d);
	if(double(f) < d) f = std::nextafter(f, std::numeric_limits<float>::infinity());
	return f;
}

void transform_point(const float* M, const float* p, float out[3]) {
	for(int r = 0; r < 3; ++r) {
		double acc = double(M[r * 4 + 0]) * double(p[0]);
		acc = acc + double(M[r * 4 + 1]) * double(p[1]);
		acc = acc + double(M[r * 4 + 2]) * double(p[2]);
		acc = acc + double(M[r * 4 + 3]);
		out[r] = float(acc);
	}
}

bool build_bvh(const ddgi_vertex* vertices, size_t vertex_count, const uint32_t* indices, size_t index_count,
			   const ddgi_offset_entry* offsets, size_t offset_count, const ddgi_instance* instances,
			   size_t instance_count, BvhHost& out, const char** err) {
	out = BvhHost{};
	// ---- flatten instances (TLAS order) into world-space triangles ----
	size_t total = 0;
	for(size_t k = 0; k < instance_count; ++k) total += instances[k].triangle_count;
	if(total >= (size_t(1) << 28)) {
		*err = "scene exceeds 2^28 triangles";
		return false;
	}
	std::vector<float> world(total * 9), object(total * 9);
	std::vector<uint32_t> inst_of(total), meta(total);
	out.inst_prim_base.resize(instance_count);
	size_t g = 0;
	for(size_t k = 0; k < instance_count; ++k) {
		const ddgi_instance& in = instances[k];
		if(in.custom_index >= offset_count) {
			*err = "instance custom_index outside the offset table";
			return false;
		}
		const ddgi_offset_entry& oe = offsets[in.custom_index];
		if(size_t(oe.index_offset) + size_t(in.triangle_count) * 3 > index_count) {
			*err = "instance triangles exceed the index buffer";
			return false;
		}
		for(int q = 0; q < 12; ++q) {
			if(!std::isfinite(in.transform[q])) {
				*err = "non-finite instance transform";
				return false;
			}
		}
		const float* M = in.transform;
		const double det = double(M[0]) * (double(M[5]) * M[10] - double(M[6]) * M[9]) -
						   double(M[1]) * (double(M[4]) * M[10] - double(M[6]) * M[8]) +
						   double(M[2]) * (double(M[4]) * M[9] - double(M[5]) * M[8]);
		const bool flip = (det < 0.0) != ((in.flags & DDGI_INSTANCE_FLIP_FACING) != 0);
		out.inst_prim_base[k] = uint32_t(g);
		for(uint32_t p = 0; p < in.triangle_count; ++p, ++g) {
			for(int v = 0; v < 3; ++v) {
				const uint32_t idx = indices[oe.index_offset + 3 * p + v];
				if(size_t(oe.vertex_offset) + idx >= vertex_count) {
					*err = "index outside the vertex buffer";
					return false;
				}
				const float* pos = vertices[oe.vertex_offset + idx].pos;
				if(!std::isfinite(pos[0]) || !std::isfinite(pos[1]) || !std::isfinite(pos[2])) {
					*err = "non-finite vertex position";
					return false;
				}
				transform_point(M, pos, &world[g * 9 + v * 3]);
				for(int q = 0; q < 3; ++q) object[g * 9 + v * 3 + q] = pos[q];
			}
			inst_of[g] = uint32_t(k);
			meta[g] = (in.mask & 0xFFu) | (flip ? 0x100u : 0u);
		}
	}
	out.tri_count = uint32_t(total);
	if(total == 0) {
		out.node_count = 0;
		return true;
	}
	// ---- build ----
	std::vector<Aabb> tb(total);
	std::vector<float> cen(total * 3);
	for(size_t i = 0; i < total; ++i) {
		for(int v = 0; v < 3; ++v) tb[i].grow(&world[i * 9 + v * 3]);
		for(int k = 0; k < 3; ++k) cen[i * 3 + k] = 0.5f * (tb[i].lo[k] + tb[i].hi[k]);
	}
	Builder b(tb, cen);
	b.build(uint32_t(total));
	out.max_depth = b.max_depth;

	// triangles in leaf order
	out.tris.resize(total * 12);
	out.obj_tris.resize(total * 9);
	for(size_t i = 0; i < total; ++i) {
		const uint32_t t = b.order[i];
		std::memcpy(&out.obj_tris[i * 9], &object[t * 9], 9 * sizeof(float));
		float* d = &out.tris[i * 12];
		for(int v = 0; v < 3; ++v) {
			d[v * 4 + 0] = world[t * 9 + v * 3 + 0];
			d[v * 4 + 1] = world[t * 9 + v * 3 + 1];
			d[v * 4 + 2] = world[t * 9 + v * 3 + 2];
		}
		d[3] = u2f(t);
		d[7] = u2f(inst_of[t]);
		d[11] = u2f(meta[t]);
	}
	// ---- collapse to a kWide-wide tree: each wide node takes the children of a binary node and
	// keeps opening its largest-area inner child until it has kWide ----
	auto leaf_code = [&](const BuildNode& n) { return kLeafFlag | (n.first << 3) | (n.count - 1); };
	auto wide_children = [&](int32_t n, int32_t ch[kWide]) -> int {
		int cnt = 0;
		if(b.nodes[n].leaf()) {  // the whole scene is one leaf: a root with that leaf as its only child
			ch[cnt++] = n;
			return cnt;
		}
		ch[cnt++] = b.nodes[n].left;
		ch[cnt++] = b.nodes[n].right;
		while(cnt < int(kWide)) {
			int pick = -1;
			double area = -1.0;
			for(int k = 0; k < cnt; ++k) {
				const BuildNode& c = b.nodes[ch[k]];
				if(!c.leaf() && c.box.half_area() > area) {
					area = c.box.half_area();
					pick = k;
				}
			}
			if(pick < 0) break;
			const int32_t open = ch[pick];
			ch[pick] = b.nodes[open].left;
			ch[cnt++] = b.nodes[open].right;
		}
		return cnt;
	};
	struct Wide {
		int32_t bin;       // binary node it was made from
		uint32_t depth;
		int32_t kids[kWide];  // binary child nodes, -1 = empty slot
	};
	// depth first: the children of a node get consecutive indices, then each child's subtree in turn
	Wide root{0, 1, {}};
	for(uint32_t k = 0; k < kWide; ++k) root.kids[k] = -1;
	std::vector<Wide> wide{root};
	std::vector<uint32_t> todo{0};
	uint32_t wide_depth = 0;
	while(!todo.empty()) {
		const uint32_t w = todo.back();
		todo.pop_back();
		int32_t ch[kWide];
		const int cnt = wide_children(wide[w].bin, ch);
		wide_depth = std::max(wide_depth, wide[w].depth);
		uint32_t created[kWide];
		int ncreated = 0;
		for(int k = 0; k < int(kWide); ++k) {
			wide[w].kids[k] = k < cnt ? ch[k] : -1;
			if(k < cnt && !b.nodes[ch[k]].leaf()) {
				created[ncreated++] = uint32_t(wide.size());
				Wide child{ch[k], wide[w].depth + 1, {}};
				for(uint32_t q = 0; q < kWide; ++q) child.kids[q] = -1;
				wide.push_back(child);
			}
		}
		for(int k = ncreated - 1; k >= 0; --k) todo.push_back(created[k]);
	}
	out.wide_depth = wide_depth;
	// node indices in the order wide nodes were created
	std::vector<uint32_t> wide_index(b.nodes.size(), 0xFFFFFFFFu);
	for(size_t w = 0; w < wide.size(); ++w) wide_index[wide[w].bin] = uint32_t(w);
	out.nodes.assign(wide.size() * kNodeFloats, 0.0f);
	for(size_t w = 0; w < wide.size(); ++w) {
		float* rec = &out.nodes[w * kNodeFloats];
		for(int k = 0; k < int(kWide); ++k) {
			const int32_t c = wide[w].kids[k];
			float lo[3], hi[3];
			uint32_t code = kEmptyChild;
			if(c >= 0) {
				const BuildNode& n = b.nodes[c];
				for(int a = 0; a < 3; ++a) {
					lo[a] = pad_down(n.box.lo[a]);
					hi[a] = pad_up(n.box.hi[a]);
				}
				code = n.leaf() ? leaf_code(n) : wide_index[c];
			} else {
				for(int a = 0; a < 3; ++a) lo[a] = hi[a] = 1e30f;  // unreachable point box
			}
			for(int a = 0; a < 3; ++a) {
				rec[(2 * a) * kWide + k] = lo[a];
				rec[(2 * a + 1) * kWide + k] = hi[a];
			}
			rec[kCodeRow + k] = u2f(code);
		}
	}
	out.node_count = uint32_t(wide.size());
	// nodes by level, for bottom-up refits
	out.level_offsets.assign(wide_depth + 1, 0);
	for(const Wide& w : wide) ++out.level_offsets[w.depth];
	for(uint32_t d = 1; d <= wide_depth; ++d) out.level_offsets[d] += out.level_offsets[d - 1];
	out.level_nodes.resize(wide.size());
	{
		std::vector<uint32_t> fill(out.level_offsets.begin(), out.level_offsets.end() - 1);
		for(size_t w = 0; w < wide.size(); ++w) out.level_nodes[fill[wide[w].depth - 1]++] = uint32_t(w);
	}
	// SAH cost of the emitted tree (for reporting)
	double cost = 0.0;
	const double root_area = std::max(b.nodes[0].box.half_area(), 1e-30);
	for(const BuildNode& n : b.nodes) {
		const double a = n.box.half_area() / root_area;
		cost += n.leaf() ? a * kCostTriangle * n.count : a * kCostTraverse;
	}
	out.sah_cost = cost;
	return true;
}

}  // namespace ddgi
