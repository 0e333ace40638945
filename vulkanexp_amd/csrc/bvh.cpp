// Binned-SAH BVH builder (host) with spatial splits.  See bvh.hpp for the device layout.
#include "bvh.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <new>
#include <system_error>
#include <thread>

namespace ddgi {

namespace {

constexpr int kBins = 32;
constexpr double kCostTraverse = 1.0;  // one 64-byte node fetch + 2 slab tests
constexpr double kCostTriangle = 1.0;  // one 48-byte triangle fetch + watertight test
// Spatial splits (Stich, Friedrich and Dietrich, "Spatial Splits in Bounding Volume Hierarchies",
// HPG 2009) are tried where the best object split's children overlap by more than kSplitAlpha of
// the root's surface, while the references stay within kSplitBudget times the triangles; below
// kSplitMaxDepth binary levels the builder uses object splits only.
constexpr double kSplitAlpha = 1e-3;
constexpr double kSplitBudget = 1.5;
constexpr uint32_t kSplitMaxDepth = 48;

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

// A reference to triangle `tri`: the bounds of the part of it this subtree holds (the whole
// triangle, or the piece a spatial split left on one side).
struct Ref {
	uint32_t tri;
	Aabb box;
	float centroid(int a) const { return 0.5f * (box.lo[a] + box.hi[a]); }
};

// Convex polygon (a triangle cut by axis-aligned planes: at most 3 + 2 vertices), in double.
struct Poly {
	double v[8][3];
	int n = 0;
};

// The part of `in` with x[axis] <= plane (keep_low) or >= plane.
void clip_poly(const Poly& in, int axis, double plane, bool keep_low, Poly& out) {
	out.n = 0;
	for(int i = 0; i < in.n; ++i) {
		const double* a = in.v[i];
		const double* b = in.v[(i + 1) % in.n];
		const double da = keep_low ? plane - a[axis] : a[axis] - plane;
		const double db = keep_low ? plane - b[axis] : b[axis] - plane;
		if(da >= 0.0) {
			for(int k = 0; k < 3; ++k) out.v[out.n][k] = a[k];
			++out.n;
		}
		if((da >= 0.0) != (db >= 0.0)) {
			const double u = da / (da - db);
			for(int k = 0; k < 3; ++k) out.v[out.n][k] = a[k] + (b[k] - a[k]) * u;
			out.v[out.n][axis] = plane;
			++out.n;
		}
	}
}

// Float bounds containing the polygon (rounded outward), intersected with `within`; empty if the
// polygon is empty or misses `within`.
Aabb poly_bounds(const Poly& p, const Aabb& within) {
	Aabb out;
	if(p.n == 0) return out;
	for(int i = 0; i < p.n; ++i)
		for(int k = 0; k < 3; ++k) {
			float lo = float(p.v[i][k]), hi = lo;
			if(double(lo) > p.v[i][k]) lo = std::nextafter(lo, -std::numeric_limits<float>::infinity());
			if(double(hi) < p.v[i][k]) hi = std::nextafter(hi, std::numeric_limits<float>::infinity());
			out.lo[k] = std::min(out.lo[k], lo);
			out.hi[k] = std::max(out.hi[k], hi);
		}
	for(int k = 0; k < 3; ++k) {
		out.lo[k] = std::max(out.lo[k], within.lo[k]);
		out.hi[k] = std::min(out.hi[k], within.hi[k]);
		if(out.lo[k] > out.hi[k]) return Aabb();
	}
	return out;
}

struct Builder {
	const std::vector<float>& world;  // 9 floats per triangle
	std::vector<uint32_t> order;      // triangle of each slot, leaf order (a split triangle repeats)
	std::vector<Aabb> order_box;      // the bounds of the piece each slot holds
	std::vector<BuildNode> nodes;
	uint32_t max_depth = 0;
	bool spatial = true;
	uint32_t depth0 = 0;     // depth of this builder's root in the whole tree
	double root_area = 0.0;  // the whole tree's root surface (0: set from this builder's root)
	size_t budget = 0;       // references this builder may reach (0: kSplitBudget x its input)
	size_t grain = 0;        // subtrees of at most this many references go to worker threads (0: none)

	explicit Builder(const std::vector<float>& w) : world(w) {}

	// Subtree `sub` (built for node `at` of this tree) into this tree: its root becomes node `at`,
	// its other nodes and its slots are appended.
	void splice(uint32_t at, const Builder& sub) {
		const int32_t base = int32_t(nodes.size()) - 1;  // sub node k >= 1 lands at base + k
		const uint32_t obase = uint32_t(order.size());
		auto mapped = [&](BuildNode n) {
			if(n.leaf()) n.first += obase;
			else n.left += base, n.right += base;
			return n;
		};
		nodes[at] = mapped(sub.nodes[0]);
		for(size_t k = 1; k < sub.nodes.size(); ++k) nodes.push_back(mapped(sub.nodes[k]));
		order.insert(order.end(), sub.order.begin(), sub.order.end());
		order_box.insert(order_box.end(), sub.order_box.begin(), sub.order_box.end());
		max_depth = std::max(max_depth, sub.max_depth);
	}

	Poly triangle(uint32_t t) const {
		Poly p;
		p.n = 3;
		for(int v = 0; v < 3; ++v)
			for(int k = 0; k < 3; ++k) p.v[v][k] = world[size_t(t) * 9 + v * 3 + k];
		return p;
	}

	// Bounds of the pieces of reference r below and above `plane` on `axis`.
	void split_ref(const Ref& r, int axis, double plane, Aabb& lo_box, Aabb& hi_box) const {
		const Poly tri = triangle(r.tri);
		Poly part;
		clip_poly(tri, axis, plane, true, part);
		lo_box = poly_bounds(part, r.box);
		clip_poly(tri, axis, plane, false, part);
		hi_box = poly_bounds(part, r.box);
	}

	void build(std::vector<Ref> all) {
		const size_t n = all.size();
		if(budget == 0) budget = std::min(size_t(double(n) * kSplitBudget), (size_t(1) << 28) - 1);
		size_t refs = n;
		nodes.reserve(2 * n + 8);
		nodes.emplace_back();
		struct Job {
			uint32_t node, depth;
			std::vector<Ref> refs;
		};
		std::vector<Job> stack, deferred;
		stack.push_back(Job{0, depth0, std::move(all)});
		while(!stack.empty()) {
			Job j = std::move(stack.back());
			stack.pop_back();
			if(grain && j.node != 0 && j.refs.size() <= grain) {
				deferred.push_back(std::move(j));
				continue;
			}
			max_depth = std::max(max_depth, j.depth);
			std::vector<Ref>& R = j.refs;
			Aabb box, cbox;
			for(const Ref& r : R) {
				box.grow(r.box);
				const float c[3] = {r.centroid(0), r.centroid(1), r.centroid(2)};
				cbox.grow(c);
			}
			nodes[j.node].box = box;
			if(j.node == 0 && root_area == 0.0) root_area = std::max(box.half_area(), 1e-30);
			const uint32_t cnt = uint32_t(R.size());
			const double leaf_cost = kCostTriangle * cnt;
			const double parent_area = std::max(box.half_area(), 1e-30);
			// object split: references binned by the centroids of their bounds
			double best = std::numeric_limits<double>::max();
			int best_axis = -1, best_bin = -1;
			Aabb best_l, best_r;
			auto bin_of = [&](const Ref& r, int ax, double scale) {
				const int b = int((double(r.centroid(ax)) - cbox.lo[ax]) * scale);
				return std::min(std::max(b, 0), kBins - 1);
			};
			for(int ax = 0; ax < 3 && cnt > 1; ++ax) {
				const double ext = double(cbox.hi[ax]) - cbox.lo[ax];
				if(!(ext > 0.0)) continue;
				const double scale = kBins / ext;
				Aabb bb[kBins];
				uint32_t bc[kBins] = {};
				for(const Ref& r : R) {
					const int b = bin_of(r, ax, scale);
					bb[b].grow(r.box);
					++bc[b];
				}
				Aabb racc[kBins];
				uint32_t rc[kBins];
				Aabb acc;
				uint32_t c = 0;
				for(int b = kBins - 1; b > 0; --b) {
					acc.grow(bb[b]);
					c += bc[b];
					racc[b] = acc;
					rc[b] = c;
				}
				Aabb lacc;
				uint32_t lc = 0;
				for(int b = 0; b < kBins - 1; ++b) {
					lacc.grow(bb[b]);
					lc += bc[b];
					if(lc == 0 || rc[b + 1] == 0) continue;
					const double cost =
						kCostTraverse + kCostTriangle * (lacc.half_area() * lc + racc[b + 1].half_area() * rc[b + 1]) / parent_area;
					if(cost < best) {
						best = cost;
						best_axis = ax;
						best_bin = b;
						best_l = lacc;
						best_r = racc[b + 1];
					}
				}
			}
			// spatial split: bins of equal width over the node's bounds; a reference enters the bin
			// of its low end and leaves from the bin of its high end, and every bin it spans grows
			// by the bounds of the piece of the triangle inside it (cut bin by bin)
			double best_s = std::numeric_limits<double>::max();
			int s_axis = -1;
			double s_plane = 0.0;
			bool try_spatial = spatial && best_axis >= 0 && refs < budget && j.depth < kSplitMaxDepth;
			if(try_spatial) {
				Aabb ov;
				for(int k = 0; k < 3; ++k) {
					ov.lo[k] = std::max(best_l.lo[k], best_r.lo[k]);
					ov.hi[k] = std::min(best_l.hi[k], best_r.hi[k]);
				}
				const bool overlap = ov.lo[0] <= ov.hi[0] && ov.lo[1] <= ov.hi[1] && ov.lo[2] <= ov.hi[2];
				try_spatial = overlap && ov.half_area() > kSplitAlpha * root_area;
			}
			// per axis its cheapest plane; the three axes of a large node on threads of their own (the
			// results are combined in axis order, as one thread would)
			double axis_cost[3] = {best_s, best_s, best_s}, axis_plane[3] = {0.0, 0.0, 0.0};
			auto spatial_axis = [&](int ax) {
				const double lo = box.lo[ax], ext = double(box.hi[ax]) - lo;
				if(!(ext > 0.0)) return;
				const double w = ext / kBins;
				auto plane_at = [&](int b) { return lo + w * b; };  // low plane of bin b
				Aabb bb[kBins];
				uint32_t entry[kBins] = {}, leave[kBins] = {};
				for(const Ref& r : R) {
					int b0 = int((double(r.box.lo[ax]) - lo) / w), b1 = int((double(r.box.hi[ax]) - lo) / w);
					b0 = std::min(std::max(b0, 0), kBins - 1);
					b1 = std::min(std::max(b1, b0), kBins - 1);
					++entry[b0];
					++leave[b1];
					if(b0 == b1) {
						bb[b0].grow(r.box);
						continue;
					}
					Poly rest = triangle(r.tri), piece, next;
					for(int b = b0; b < b1; ++b) {
						clip_poly(rest, ax, plane_at(b + 1), true, piece);
						const Aabb pb = poly_bounds(piece, r.box);
						if(!pb.empty()) bb[b].grow(pb);
						clip_poly(rest, ax, plane_at(b + 1), false, next);
						rest = next;
					}
					const Aabb pb = poly_bounds(rest, r.box);
					if(!pb.empty()) bb[b1].grow(pb);
				}
				Aabb racc[kBins];
				uint32_t rc[kBins];
				Aabb acc;
				uint32_t c = 0;
				for(int b = kBins - 1; b > 0; --b) {
					acc.grow(bb[b]);
					c += leave[b];
					racc[b] = acc;
					rc[b] = c;
				}
				Aabb lacc;
				uint32_t lc = 0;
				for(int b = 0; b < kBins - 1; ++b) {
					lacc.grow(bb[b]);
					lc += entry[b];
					if(lc == 0 || rc[b + 1] == 0) continue;
					const double cost =
						kCostTraverse + kCostTriangle * (lacc.half_area() * lc + racc[b + 1].half_area() * rc[b + 1]) / parent_area;
					if(cost < axis_cost[ax]) {
						axis_cost[ax] = cost;
						axis_plane[ax] = plane_at(b + 1);
					}
				}
			};
			if(try_spatial) {
				if(cnt >= 16384u) {
					// an axis whose thread cannot start is binned here
					std::thread t[2];
					for(int k = 0; k < 2; ++k) {
						try {
							t[k] = std::thread(spatial_axis, k + 1);
						} catch(const std::system_error&) {
						}
					}
					spatial_axis(0);
					for(int k = 0; k < 2; ++k) {
						if(t[k].joinable()) t[k].join();
						else spatial_axis(k + 1);
					}
				} else {
					for(int ax = 0; ax < 3; ++ax) spatial_axis(ax);
				}
				for(int ax = 0; ax < 3; ++ax)
					if(axis_cost[ax] < best_s) {
						best_s = axis_cost[ax];
						s_axis = ax;
						s_plane = axis_plane[ax];
					}
			}
			const bool can_leaf = cnt <= kMaxLeafSize;
			if(cnt == 1 || (can_leaf && leaf_cost <= std::min(best, best_s)) || (best_axis < 0 && s_axis < 0 && can_leaf)) {
				nodes[j.node].first = uint32_t(order.size());
				nodes[j.node].count = cnt;
				for(const Ref& r : R) {
					order.push_back(r.tri);
					order_box.push_back(r.box);
				}
				continue;
			}
			std::vector<Ref> L, H;
			if(s_axis >= 0 && best_s < best) {
				// a reference on one side goes there whole; one across the plane is cut in two
				size_t cut = 0;
				for(const Ref& r : R) {
					if(double(r.box.hi[s_axis]) <= s_plane) {
						L.push_back(r);
					} else if(double(r.box.lo[s_axis]) >= s_plane) {
						H.push_back(r);
					} else {
						Aabb lb, hb;
						split_ref(r, s_axis, s_plane, lb, hb);
						if(!lb.empty()) L.push_back(Ref{r.tri, lb});
						if(!hb.empty()) H.push_back(Ref{r.tri, hb});
						cut += !lb.empty() && !hb.empty();
					}
				}
				refs += cut;
			}
			if(L.empty() || H.empty()) {  // object split (or halves when the centroids coincide)
				L.clear();
				H.clear();
				if(best_axis >= 0) {
					const double scale = kBins / (double(cbox.hi[best_axis]) - cbox.lo[best_axis]);
					for(const Ref& r : R) (bin_of(r, best_axis, scale) <= best_bin ? L : H).push_back(r);
				}
				if(L.empty() || H.empty()) {
					L.clear();
					H.clear();
					for(uint32_t i = 0; i < cnt; ++i) (i < cnt / 2 ? L : H).push_back(R[i]);
				}
			}
			R.clear();
			R.shrink_to_fit();
			const int32_t l = int32_t(nodes.size());
			nodes.emplace_back();
			nodes.emplace_back();
			nodes[j.node].left = l;
			nodes[j.node].right = l + 1;
			stack.push_back(Job{uint32_t(l + 1), j.depth + 1, std::move(H)});
			stack.push_back(Job{uint32_t(l), j.depth + 1, std::move(L)});
		}
		if(deferred.empty()) return;
		// the deferred subtrees, each by its own builder on a pool of threads (the work each does is
		// fixed by its input, and the subtrees are spliced in a fixed order: the tree does not
		// depend on the scheduling)
		std::vector<Builder> subs;
		subs.reserve(deferred.size());
		for(Job& j : deferred) {
			subs.emplace_back(world);
			Builder& b = subs.back();
			b.spatial = spatial;
			b.depth0 = j.depth;
			b.root_area = root_area;
			b.budget = std::min(size_t(double(j.refs.size()) * kSplitBudget), (size_t(1) << 28) - 1);
		}
		std::atomic<size_t> next{0};
		std::atomic<bool> out_of_memory{false};
		auto work = [&]() {
			for(size_t i = next++; i < subs.size(); i = next++) {
				try {
					subs[i].build(std::move(deferred[i].refs));
				} catch(const std::bad_alloc&) {
					out_of_memory = true;
				}
			}
		};
		const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
		std::vector<std::thread> pool;
		try {
			for(unsigned t = 1; t < std::min<size_t>(hw, subs.size()); ++t) pool.emplace_back(work);
		} catch(const std::system_error&) {  // fewer threads share the subtrees
		}
		work();
		for(std::thread& t : pool) t.join();
		if(out_of_memory) throw std::bad_alloc();
		for(size_t i = 0; i < subs.size(); ++i) splice(deferred[i].node, subs[i]);
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

// Build the tree over the flattened triangles and emit the device layout into `out` (its node,
// triangle and level fields).
static void emit_tree(const std::vector<float>& world, const std::vector<float>& object, const std::vector<uint32_t>& inst_of,
			   const std::vector<uint32_t>& meta, bool spatial, BvhHost& out) {
	const size_t total = inst_of.size();
	std::vector<Ref> refs(total);
	for(size_t i = 0; i < total; ++i) {
		refs[i].tri = uint32_t(i);
		for(int v = 0; v < 3; ++v) refs[i].box.grow(&world[i * 9 + v * 3]);
	}
	Builder b(world);
	b.spatial = spatial;
	b.grain = std::max<size_t>(total / 16, 4096);  // subtrees below this size are built in parallel
	b.build(std::move(refs));
	out.max_depth = b.max_depth;
	const size_t slots = b.order.size();
	out.tri_count = uint32_t(slots);

	// triangles in leaf order
	out.tris.resize(slots * 12);
	out.obj_tris.resize(slots * 9);
	out.slot_box.resize(slots * 6);
	for(size_t i = 0; i < slots; ++i) {
		const uint32_t t = b.order[i];
		for(int a = 0; a < 3; ++a) {
			out.slot_box[i * 6 + a] = b.order_box[i].lo[a];
			out.slot_box[i * 6 + 3 + a] = b.order_box[i].hi[a];
		}
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
}

static bool build_bvh_unguarded(const ddgi_vertex* vertices, size_t vertex_count, const uint32_t* indices, size_t index_count,
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
	// ---- build: with spatial splits, unless that tree is deeper than the traversal stack allows or
	// its slots overflow the leaf codes' 28 bits ----
	emit_tree(world, object, inst_of, meta, true, out);
	if(out.wide_depth > kMaxWideDepth || out.tri_count >= (1u << 28)) emit_tree(world, object, inst_of, meta, false, out);
	return true;
}

bool build_bvh(const ddgi_vertex* vertices, size_t vertex_count, const uint32_t* indices, size_t index_count,
			   const ddgi_offset_entry* offsets, size_t offset_count, const ddgi_instance* instances,
			   size_t instance_count, BvhHost& out, const char** err) {
	try {
		return build_bvh_unguarded(vertices, vertex_count, indices, index_count, offsets, offset_count, instances,
								   instance_count, out, err);
	} catch(const std::bad_alloc&) {
		out = BvhHost{};
		*err = "out of host memory building the BVH";
		return false;
	}
}

}  // namespace ddgi
