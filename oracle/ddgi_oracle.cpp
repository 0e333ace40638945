// TEST INFRASTRUCTURE — the CPU oracle of the DDGI probe update (see oracle.hpp for the
// precision model).  Loaded by tests/, the smoke test and bench.py's cpu_baseline leg only.
//
// Reference files restated here, function by function:
//   scene flattening ............ src/Renderer.cpp:97-126 (offset table), :532-551 (TLAS instances),
//                                 src/shaders/closesthit.glsl:147-159 (record fetch, normal transform)
//   traceProbes ................. src/shaders/traceProbes.rgen:27-49
//   closest hit (NO_REFLECTION) . src/shaders/closesthit.glsl:133-289
//   pbrMetallicRoughness ........ src/shaders/pbrMetallicRoughness.glsl:8-84
//   miss / shadow miss .......... src/shaders/miss.rmiss:17-23 (constant sky, or sky.glsl:60-127),
//                                 src/shaders/shadow.rmiss:6-9
//   sampleProbes ................ src/shaders/irradiance.glsl:145-237
//   blend + state machine ....... src/shaders/probesUpdate.glsl:49-119
//   border copy ................. src/shaders/probesCopyBorders.comp:21-246 (as the mirror formula;
//                                 tests/golden/border_tables.json holds the reference tables)
//   probe init .................. src/shaders/probesInit.rgen:29-64, backfaceTest.rchit:11-15,
//                                 probeInitMiss.rmiss:11-15
//   probe selection ............. src/IrradianceProbes.cpp:396-424
#include <algorithm>
#include <array>
#include <chrono>
#include <cstdio>
#include <limits>

#include "oracle.hpp"
#include "threadpool.hpp"

namespace orc {

// =============================================================================================
// FP32 watertight ray/triangle test (Woop, Benthin, Wald 2013), fixed operation order.
// Front facing <=> det > 0 <=> dot(direction, cross(v1-v0, v2-v0)) < 0.
void ray32_prepare(Ray32& r) {
	float ax = std::fabs(r.d[0]), ay = std::fabs(r.d[1]), az = std::fabs(r.d[2]);
	int kz = (ax > ay) ? ((ax > az) ? 0 : 2) : ((ay > az) ? 1 : 2);
	int kx = kz + 1;
	if(kx == 3) kx = 0;
	int ky = kx + 1;
	if(ky == 3) ky = 0;
	if(r.d[kz] < 0.0f) std::swap(kx, ky);
	r.kx = kx;
	r.ky = ky;
	r.kz = kz;
	r.sx = r.d[kx] / r.d[kz];
	r.sy = r.d[ky] / r.d[kz];
	r.sz = 1.0f / r.d[kz];
}

bool tri_intersect32(const Ray32& r, const float* v0, const float* v1, const float* v2, float& t, float& U, float& V,
					 float& W, float& det) {
	const int kx = r.kx, ky = r.ky, kz = r.kz;
	const float Akx = v0[kx] - r.o[kx], Aky = v0[ky] - r.o[ky], Akz = v0[kz] - r.o[kz];
	const float Bkx = v1[kx] - r.o[kx], Bky = v1[ky] - r.o[ky], Bkz = v1[kz] - r.o[kz];
	const float Ckx = v2[kx] - r.o[kx], Cky = v2[ky] - r.o[ky], Ckz = v2[kz] - r.o[kz];
	const float Ax = Akx - r.sx * Akz, Ay = Aky - r.sy * Akz;
	const float Bx = Bkx - r.sx * Bkz, By = Bky - r.sy * Bkz;
	const float Cx = Ckx - r.sx * Ckz, Cy = Cky - r.sy * Ckz;
	U = Cx * By - Cy * Bx;
	V = Ax * Cy - Ay * Cx;
	W = Bx * Ay - By * Ax;
	if(U == 0.0f || V == 0.0f || W == 0.0f) {
		// Edge case: redo the 2D edge functions in double (exact products), as in the paper.
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
	return t >= r.tmin && t <= r.tmax;
}

// =============================================================================================
// Oracle BVH: binned SAH over triangle centroids, 16 bins, leaves of <= 4 triangles.
namespace {
struct Box {
	float lo[3] = {std::numeric_limits<float>::max(), std::numeric_limits<float>::max(), std::numeric_limits<float>::max()};
	float hi[3] = {-std::numeric_limits<float>::max(), -std::numeric_limits<float>::max(), -std::numeric_limits<float>::max()};
	void grow(const float* p) {
		for(int k = 0; k < 3; ++k) {
			lo[k] = std::min(lo[k], p[k]);
			hi[k] = std::max(hi[k], p[k]);
		}
	}
	void grow(const Box& b) {
		for(int k = 0; k < 3; ++k) {
			lo[k] = std::min(lo[k], b.lo[k]);
			hi[k] = std::max(hi[k], b.hi[k]);
		}
	}
	double area() const {
		double e[3];
		for(int k = 0; k < 3; ++k) e[k] = std::max(0.0, double(hi[k]) - double(lo[k]));
		return 2.0 * (e[0] * e[1] + e[1] * e[2] + e[2] * e[0]);
	}
};
}  // namespace

void Scene::build_bvh() {
	const size_t n = tri_count();
	nodes.clear();
	order.resize(n);
	for(size_t i = 0; i < n; ++i) order[i] = uint32_t(i);
	if(n == 0) return;
	std::vector<Box> tb(n);
	std::vector<std::array<float, 3>> cen(n);
	for(size_t i = 0; i < n; ++i) {
		for(int v = 0; v < 3; ++v) tb[i].grow(&tri[i * 9 + v * 3]);
		for(int k = 0; k < 3; ++k) cen[i][k] = 0.5f * (tb[i].lo[k] + tb[i].hi[k]);
	}
	struct Job {
		uint32_t node, begin, end;
	};
	nodes.push_back({});
	std::vector<Job> stack{{0, 0, uint32_t(n)}};
	constexpr int kBins = 16;
	while(!stack.empty()) {
		Job j = stack.back();
		stack.pop_back();
		Box bb, cb;
		for(uint32_t i = j.begin; i < j.end; ++i) {
			bb.grow(tb[order[i]]);
			cb.grow(cen[order[i]].data());
		}
		Node& nd = nodes[j.node];
		for(int k = 0; k < 3; ++k) {
			nd.lo[k] = bb.lo[k];
			nd.hi[k] = bb.hi[k];
		}
		const uint32_t cnt = j.end - j.begin;
		int best_axis = -1;
		int best_split = -1;
		double best_cost = double(cnt);  // leaf cost (one unit per triangle)
		if(cnt > 4) best_cost = std::numeric_limits<double>::max();
		for(int ax = 0; ax < 3 && cnt > 1; ++ax) {
			double ext = double(cb.hi[ax]) - double(cb.lo[ax]);
			if(!(ext > 0.0)) continue;
			Box bins[kBins];
			uint32_t bc[kBins] = {};
			for(uint32_t i = j.begin; i < j.end; ++i) {
				int b = int(kBins * ((double(cen[order[i]][ax]) - cb.lo[ax]) / ext));
				b = std::min(std::max(b, 0), kBins - 1);
				bins[b].grow(tb[order[i]]);
				++bc[b];
			}
			double right_area[kBins];
			uint32_t right_cnt[kBins];
			Box acc;
			uint32_t c = 0;
			for(int b = kBins - 1; b > 0; --b) {
				acc.grow(bins[b]);
				c += bc[b];
				right_area[b] = acc.area();
				right_cnt[b] = c;
			}
			Box lacc;
			uint32_t lc = 0;
			const double inv_parent = 1.0 / std::max(bb.area(), 1e-30);
			for(int b = 0; b < kBins - 1; ++b) {
				lacc.grow(bins[b]);
				lc += bc[b];
				if(lc == 0 || right_cnt[b + 1] == 0) continue;
				double cost = 1.0 + (lacc.area() * lc + right_area[b + 1] * right_cnt[b + 1]) * inv_parent;
				if(cost < best_cost) {
					best_cost = cost;
					best_axis = ax;
					best_split = b;
				}
			}
		}
		if(best_axis < 0) {
			if(cnt <= 4 || cnt == 1) {
				nd.first = j.begin;
				nd.count = cnt;
				continue;
			}
			// no useful split (coincident centroids): split the range in half
			uint32_t mid = j.begin + cnt / 2;
			uint32_t left = uint32_t(nodes.size());
			nodes.push_back({});
			nodes.push_back({});
			nodes[j.node].first = left;
			nodes[j.node].count = 0;
			stack.push_back({left, j.begin, mid});
			stack.push_back({left + 1, mid, j.end});
			continue;
		}
		const double ext = double(cb.hi[best_axis]) - double(cb.lo[best_axis]);
		auto mid_it = std::partition(order.begin() + j.begin, order.begin() + j.end, [&](uint32_t id) {
			int b = int(kBins * ((double(cen[id][best_axis]) - cb.lo[best_axis]) / ext));
			b = std::min(std::max(b, 0), kBins - 1);
			return b <= best_split;
		});
		uint32_t mid = uint32_t(mid_it - order.begin());
		if(mid == j.begin || mid == j.end) mid = j.begin + cnt / 2;
		uint32_t left = uint32_t(nodes.size());
		nodes.push_back({});
		nodes.push_back({});
		nodes[j.node].first = left;
		nodes[j.node].count = 0;
		stack.push_back({left, j.begin, mid});
		stack.push_back({left + 1, mid, j.end});
	}
}

// Conservative slab test in double: the box is inflated and the interval padded, so that a node
// is never culled when one of its triangles could be accepted by the FP32 test.
static inline bool box_interval(const Scene::Node& nd, const double o[3], const double d[3], const double inv[3],
								double t0, double t1, double& tnear) {
	for(int k = 0; k < 3; ++k) {
		double pad = 1e-5 * (std::fabs(double(nd.lo[k])) + std::fabs(double(nd.hi[k]))) + 1e-5;
		double lo = double(nd.lo[k]) - pad, hi = double(nd.hi[k]) + pad;
		if(d[k] == 0.0) {
			if(o[k] < lo || o[k] > hi) return false;
			continue;
		}
		double a = (lo - o[k]) * inv[k], b = (hi - o[k]) * inv[k];
		if(a > b) std::swap(a, b);
		t0 = std::max(t0, a);
		t1 = std::min(t1, b);
		if(t0 > t1) return false;
	}
	tnear = t0;
	return true;
}

static inline bool better(float t, uint32_t id, const Hit32& best) {
	return !best.hit() || t < best.t || (t == best.t && id < best.tri);
}

// `cull`: back-facing triangles are ignored (the G-buffer rasterizer's VK_CULL_MODE_BACK_BIT,
// src/RenderPasses/GBufferPipeline.cpp:212).
static inline bool try_tri(const Scene& s, const Ray32& r, uint32_t id, uint32_t mask, Hit32& best, bool cull) {
	if((s.inst_mask[s.tri_inst[id]] & mask) == 0) return false;
	const float* v = &s.tri[uint64_t(id) * 9];
	float t, U, V, W, det;
	if(!tri_intersect32(r, v, v + 3, v + 6, t, U, V, W, det)) return false;
	if(cull && ((det < 0.0f) != (s.tri_flip[id] != 0))) return false;
	if(!better(t, id, best)) return false;
	best.t = t;
	best.tri = id;
	best.u = V / det;
	best.v = W / det;
	best.back = (det < 0.0f) != (s.tri_flip[id] != 0);
	return true;
}

Hit32 trace32(const Scene& s, const Ray32& r, uint32_t mask, bool any_hit, bool brute, uint64_t* stats, bool cull) {
	Hit32 best;
	const size_t n = s.tri_count();
	if(brute || s.nodes.empty()) {
		for(size_t i = 0; i < n; ++i) {
			if(try_tri(s, r, uint32_t(i), mask, best, cull) && any_hit) return best;
		}
		if(stats) stats[1] += n;
		return best;
	}
	double o[3], d[3], inv[3];
	for(int k = 0; k < 3; ++k) {
		o[k] = r.o[k];
		d[k] = r.d[k];
		inv[k] = d[k] != 0.0 ? 1.0 / d[k] : 0.0;
	}
	const double tlo = double(r.tmin) - 1e-4 * std::fabs(double(r.tmin)) - 1e-4;
	uint32_t stack[128];
	int sp = 0;
	stack[sp++] = 0;
	while(sp > 0) {
		const uint32_t ni = stack[--sp];
		const Scene::Node& nd = s.nodes[ni];
		double thi = best.hit() ? double(best.t) : double(r.tmax);
		thi += 1e-4 * std::fabs(thi) + 1e-4;
		double tn;
		if(stats) stats[0] += 1;
		if(!box_interval(nd, o, d, inv, tlo, thi, tn)) continue;
		if(nd.count > 0) {
			for(uint32_t k = 0; k < nd.count; ++k) {
				if(stats) stats[1] += 1;
				if(try_tri(s, r, s.order[nd.first + k], mask, best, cull) && any_hit) return best;
			}
			continue;
		}
		// push the farther child first
		const Scene::Node& a = s.nodes[nd.first];
		const Scene::Node& b = s.nodes[nd.first + 1];
		double ta = 0, tb = 0;
		bool ha = box_interval(a, o, d, inv, tlo, thi, ta);
		bool hb = box_interval(b, o, d, inv, tlo, thi, tb);
		if(ha && hb) {
			if(ta <= tb) {
				stack[sp++] = nd.first + 1;
				stack[sp++] = nd.first;
			} else {
				stack[sp++] = nd.first;
				stack[sp++] = nd.first + 1;
			}
		} else if(ha) {
			stack[sp++] = nd.first;
		} else if(hb) {
			stack[sp++] = nd.first + 1;
		}
		if(sp > 126) {  // cannot happen for depth <= 64 trees; degrade to brute force if it does
			return trace32(s, r, mask, any_hit, true, stats, cull);
		}
	}
	return best;
}

// =============================================================================================
// sampleProbes (src/shaders/irradiance.glsl:145-237) on image-layout double atlases.
template <class T, int C>
static void bilinear_tile(const Atlas& a, uint32_t ox, uint32_t oy, uint32_t res, T u, T v, T out[C]) {
	// texel coordinates inside the probe tile: 1 + (res-2)*uv, centres at k+0.5 (H8: exact
	// FP bilinear weights instead of the texture unit's fixed-point ones)
	T sx = T(1) + T(res - 2) * u - T(0.5);
	T sy = T(1) + T(res - 2) * v - T(0.5);
	T fx = std::floor(sx), fy = std::floor(sy);
	T wx = sx - fx, wy = sy - fy;
	int x0 = std::min(std::max(int(fx), 0), int(res) - 1);
	int y0 = std::min(std::max(int(fy), 0), int(res) - 1);
	int x1 = std::min(x0 + 1, int(res) - 1), y1 = std::min(y0 + 1, int(res) - 1);
	for(int c = 0; c < C; ++c) {
		T t00 = T(a.at(ox + x0, oy + y0, c)), t10 = T(a.at(ox + x1, oy + y0, c));
		T t01 = T(a.at(ox + x0, oy + y1, c)), t11 = T(a.at(ox + x1, oy + y1, c));
		out[c] = (t00 * (T(1) - wx) + t10 * wx) * (T(1) - wy) + (t01 * (T(1) - wx) + t11 * wx) * wy;
	}
}

// The irradiance-independent part of one sampleProbes call: for each of the 8 cage probes the
// visibility/trilinear weight, the fallback weight and the bilinear footprint in the irradiance
// tile.  (Everything the adjoint needs to recompute; nothing here depends on the irradiance.)
template <class T>
struct SampleCtx {
	bool outside = true;
	int used = 0;
	uint32_t ox[8], oy[8];
	int x0[8], y0[8], x1[8], y1[8];
	T wx[8], wy[8], w[8], fw[8];
	T total_w = 0, total_fb = 0;
};

template <class T>
void sample_collect(const Grid& g, const uint32_t* states, const Atlas& dep, V3<T> position, V3<T> normal,
					V3<T> to_camera, SampleCtx<T>& ctx) {
	ctx = SampleCtx<T>();
	const V3<T> cell = g.cell<T>();
	const V3<T> emin(T(g.p.extent_min[0]), T(g.p.extent_min[1]), T(g.p.extent_min[2]));
	const V3<T> acell = vabs(cell);
	const V3<T> gc = (position - emin) / acell;
	if(gc.x < T(0) || gc.y < T(0) || gc.z < T(0)) return;
	ctx.outside = false;
	const V3<T> bias = (normal + to_camera) * T(g.p.shadow_bias);
	const V3<T> biased = position + bias;
	int first[3] = {int(gc.x), int(gc.y), int(gc.z)};
	const V3<T> p0 = g.position<T>(first);
	V3<T> alpha = (position - p0) / acell;
	alpha = {clampT(alpha.x, T(0), T(1)), clampT(alpha.y, T(0), T(1)), clampT(alpha.z, T(0), T(1))};
	const uint32_t cres = g.p.color_res, dres = g.p.depth_res;
	T cu, cv;
	sphere_to_oct_uv(normal, cu, cv);
	for(int i = 0; i < 8; ++i) {
		int off[3] = {i & 1, (i >> 1) & 1, (i >> 2) & 1};
		int pc[3] = {first[0] + off[0], first[1] + off[1], first[2] + off[2]};
		if(pc[0] > g.p.resolution[0] - 1 || pc[1] > g.p.resolution[1] - 1 || pc[2] > g.p.resolution[2] - 1) continue;
		if(states[g.linear(pc)] == 0) continue;
		const V3<T> pp = g.position<T>(pc);
		const V3<T> dir_to_probe = normalize(pp - position);
		const V3<T> bdir = pp - biased;
		T du, dv;
		sphere_to_oct_uv(-normalize(bdir), du, dv);
		const V3<T> tri(off[0] ? alpha.x : T(1) - alpha.x, off[1] ? alpha.y : T(1) - alpha.y,
						off[2] ? alpha.z : T(1) - alpha.z);
		T weight = 1;
		T bfw = std::max(T(0.0001), (dot(dir_to_probe, normal) + T(1)) * T(0.5));
		weight *= bfw * bfw + T(0.2);
		T fb_weight = weight;
		uint32_t ox, oy;
		g.tile_origin(pc, dres, ox, oy);
		T mom[2];
		bilinear_tile<T, 2>(dep, ox, oy, dres, du, dv, mom);
		const T mean = mom[0];
		const T variance = std::fabs(mom[0] * mom[0] - mom[1]);
		const T bdist = length(pp - biased);
		const T dm = std::max(bdist - mean, T(0.0001));
		T cheb = variance / (variance + dm * dm);
		cheb = std::max(cheb * cheb * cheb, T(0));
		weight *= (bdist <= mean) ? T(1) : cheb;
		weight = std::max(T(0.000001), weight);
		const T crush = T(0.2);
		if(weight < crush) weight *= weight * weight * (T(1) / (crush * crush));
		const T tw = tri.x * tri.y * tri.z + T(0.001);
		weight *= tw;
		fb_weight *= tw;
		const int k = ctx.used++;
		g.tile_origin(pc, cres, ctx.ox[k], ctx.oy[k]);
		// bilinear footprint of the colour tap (same arithmetic as bilinear_tile)
		T sx = T(1) + T(cres - 2) * cu - T(0.5), sy = T(1) + T(cres - 2) * cv - T(0.5);
		T fx = std::floor(sx), fy = std::floor(sy);
		ctx.wx[k] = sx - fx;
		ctx.wy[k] = sy - fy;
		ctx.x0[k] = std::min(std::max(int(fx), 0), int(cres) - 1);
		ctx.y0[k] = std::min(std::max(int(fy), 0), int(cres) - 1);
		ctx.x1[k] = std::min(ctx.x0[k] + 1, int(cres) - 1);
		ctx.y1[k] = std::min(ctx.y0[k] + 1, int(cres) - 1);
		ctx.w[k] = weight;
		ctx.fw[k] = fb_weight;
		ctx.total_w += weight;
		ctx.total_fb += fb_weight;
	}
}

template <class T>
static T tap_color(const SampleCtx<T>& ctx, int k, const Atlas& irr, int c) {
	const T wx = ctx.wx[k], wy = ctx.wy[k];
	const uint32_t ox = ctx.ox[k], oy = ctx.oy[k];
	T t00 = T(irr.at(ox + ctx.x0[k], oy + ctx.y0[k], c)), t10 = T(irr.at(ox + ctx.x1[k], oy + ctx.y0[k], c));
	T t01 = T(irr.at(ox + ctx.x0[k], oy + ctx.y1[k], c)), t11 = T(irr.at(ox + ctx.x1[k], oy + ctx.y1[k], c));
	return (t00 * (T(1) - wx) + t10 * wx) * (T(1) - wy) + (t01 * (T(1) - wx) + t11 * wx) * wy;
}

template <class T>
V3<T> sample_eval(const SampleCtx<T>& ctx, const Atlas& irr) {
	if(ctx.outside) return {};
	V3<T> final_color, fallback_color;
	for(int k = 0; k < ctx.used; ++k) {
		V3<T> color(std::sqrt(tap_color(ctx, k, irr, 0)), std::sqrt(tap_color(ctx, k, irr, 1)),
					std::sqrt(tap_color(ctx, k, irr, 2)));
		final_color = final_color + color * ctx.w[k];
		fallback_color = fallback_color + color * ctx.fw[k];
	}
	if(ctx.total_w > T(1e-3)) final_color = final_color * (T(1) / ctx.total_w);
	if(ctx.total_fb > T(1e-3)) fallback_color = fallback_color * (T(1) / ctx.total_fb);
	final_color = final_color * final_color;
	fallback_color = fallback_color * fallback_color;
	return mix3(fallback_color, final_color, T(8) * clampT(ctx.total_w, T(0), T(1) / T(8)));
}

// Adjoint of sample_eval with respect to the irradiance atlas: accumulates d out . g_out into
// g_irr (image layout, 3 channels).  sqrt'(x) = 1 / (2 sqrt(x)) is evaluated at max(x, 1e-12).
template <class T>
void sample_adjoint(const SampleCtx<T>& ctx, const Atlas& irr, V3<T> g_out, double* g_irr) {
	if(ctx.outside || ctx.used == 0) return;
	const T k8 = T(8) * clampT(ctx.total_w, T(0), T(1) / T(8));
	const T inv_w = ctx.total_w > T(1e-3) ? T(1) / ctx.total_w : T(1);
	const T inv_fb = ctx.total_fb > T(1e-3) ? T(1) / ctx.total_fb : T(1);
	for(int c = 0; c < 3; ++c) {
		T s[8];
		T fin = 0, fb = 0;
		for(int k = 0; k < ctx.used; ++k) {
			s[k] = std::sqrt(tap_color(ctx, k, irr, c));
			fin += s[k] * ctx.w[k];
			fb += s[k] * ctx.fw[k];
		}
		fin *= inv_w;  // normalised sqrt-space sums
		fb *= inv_fb;
		// out = (1 - k8) fb^2 + k8 fin^2
		for(int k = 0; k < ctx.used; ++k) {
			const T dout_ds = (T(1) - k8) * T(2) * fb * ctx.fw[k] * inv_fb + k8 * T(2) * fin * ctx.w[k] * inv_w;
			const T col = s[k] * s[k];
			const T dout_dcol = dout_ds * T(0.5) / std::sqrt(std::max(col, T(1e-12)));
			const T gc = dout_dcol * g_out[c];
			const T wx = ctx.wx[k], wy = ctx.wy[k];
			auto add = [&](int x, int y, T wgt) {
				g_irr[(uint64_t(ctx.oy[k] + y) * irr.width + ctx.ox[k] + x) * 3 + c] += double(gc * wgt);
			};
			add(ctx.x0[k], ctx.y0[k], (T(1) - wx) * (T(1) - wy));
			add(ctx.x1[k], ctx.y0[k], wx * (T(1) - wy));
			add(ctx.x0[k], ctx.y1[k], (T(1) - wx) * wy);
			add(ctx.x1[k], ctx.y1[k], wx * wy);
		}
	}
}

template <class T>
V3<T> sample_probes(const Grid& g, const uint32_t* states, const Atlas& irr, const Atlas& dep, V3<T> position,
					V3<T> normal, V3<T> to_camera) {
	SampleCtx<T> ctx;
	sample_collect(g, states, dep, position, normal, to_camera, ctx);
	return sample_eval(ctx, irr);
}

// =============================================================================================
// pbrMetallicRoughness (src/shaders/pbrMetallicRoughness.glsl:8-84)
template <class T>
V3<T> pbr_metallic_roughness(V3<T> n, V3<T> v, V3<T> light_color, V3<T> light_dir, V3<T> albedo, T metal, T rough) {
	const V3<T> f0(T(0.04), T(0.04), T(0.04));
	V3<T> diffuse = albedo * (V3<T>(1, 1, 1) - f0);
	diffuse = diffuse * (T(1) - metal);
	const T alpha = rough * rough;
	const V3<T> spec = mix3(f0, albedo, metal);
	const T reflectance = std::max(std::max(spec.x, spec.y), spec.z);
	const T r90 = clampT(reflectance * T(25), T(0), T(1));
	const V3<T> R0 = spec, R90(r90, r90, r90);
	const V3<T> l = normalize(light_dir);
	const V3<T> h = normalize(l + v);
	const T NdotL = clampT(dot(n, l), T(0.001), T(1));
	const T NdotV = clampT(std::fabs(dot(n, v)), T(0.001), T(1));
	const T NdotH = clampT(dot(n, h), T(0), T(1));
	const T VdotH = clampT(dot(v, h), T(0), T(1));
	const T fw = std::pow(clampT(T(1) - VdotH, T(0), T(1)), T(5));
	const V3<T> F = R0 + (R90 - R0) * fw;
	const T a2 = alpha * alpha;
	const T attL = T(2) * NdotL / (NdotL + std::sqrt(a2 + (T(1) - a2) * (NdotL * NdotL)));
	const T attV = T(2) * NdotV / (NdotV + std::sqrt(a2 + (T(1) - a2) * (NdotV * NdotV)));
	const T G = attL * attV;
	const T ah = NdotH * alpha;
	const T k = alpha / ((T(1) - NdotH * NdotH) + ah * ah);
	const T D = clampT(k * k * (T(1) / T(kPi)), T(0), T(4));
	const V3<T> diffuse_contrib = (V3<T>(1, 1, 1) - F) * diffuse * (T(1) / T(kPi));
	const V3<T> spec_contrib = F * (G * D / (T(4) * NdotL * NdotV));
	return (light_color * NdotL) * (diffuse_contrib + spec_contrib);
}

// =============================================================================================
// Atmospheric sky, src/shaders/sky.glsl:60-127 (O'Neil, "Accurate Atmospheric Scattering", GPU
// Gems 2 ch. 16), evaluated in double.  The reference's choices are kept as they are:
//   - traceSphereInside (:33-42) adds the closest-approach offset when the closest approach lies
//     ahead of the origin and subtracts it otherwise (the reverse of the geometric exit distance);
//   - a sample whose attenuation is infinite or NaN is skipped and the sample point is not
//     advanced (:106, `continue` before `samplePoint += sampleRay`);
//   - a colour with an infinite channel gives black (:118-126).
// "Infinite" means what it means in the shader's FP32: exp overflows above ln(FLT_MAX), and a
// colour channel above FLT_MAX counts as infinite.  Sample k lies at position + (k + 1/2) step,
// the point the shader reaches by repeated addition.
namespace {
constexpr double kSkyInner = 100000.0, kSkyOuter = 102500.0;
constexpr double kSkyScale = 1.0 / (kSkyOuter - kSkyInner), kSkyScaleDepth = 0.25;
constexpr double kSkyPi = double(3.1415926538f);  // common.glsl:4 (a float constant)
constexpr double kSkyKr = double(0.0025f), kSkyKm = double(0.0010f), kSkyG = double(-0.990f);

double sky_scale(double c) {  // sky.glsl:44-48
	const double x = 1.0 - c;
	return kSkyScaleDepth * std::exp(-0.00287 + x * (0.459 + x * (3.83 + x * (-6.80 + x * 5.25))));
}
double sky_enter(V3<double> o, V3<double> d, double r) {  // traceSphereOutside, centre 0 (:14-30)
	const double a = dot(d, d), b = dot(d, o), c = dot(o, o) - r * r;
	const double g = b * b - a * c;
	if(g > 0.0) {
		const double dis = (-std::sqrt(g) - b) / a;
		if(dis > 0.0) return dis;
	}
	return -1.0;
}
double sky_leave(V3<double> o, V3<double> d, double r) {  // traceSphereInside, centre 0 (:33-42)
	const double docdir = dot(-o, d);
	const V3<double> pc = o + docdir * d;
	const double lpc = length(pc);
	const double dist = std::sqrt(r * r - lpc * lpc);
	return docdir > 0.0 ? dist - length(pc - o) : dist + length(pc - o);
}
bool f32_inf(double v) { return std::isinf(v) || std::fabs(v) > double(std::numeric_limits<float>::max()); }
}  // namespace

V3<double> sky_atmosphere(V3<double> origin, V3<double> dir, V3<double> sun_dir, V3<double> sun_color, double brightness,
						  bool show_sun) {
	sun_color = sun_color * brightness;
	V3<double> pos(origin.x, origin.y + (kSkyInner + 100.0), origin.z);
	double height = length(pos);
	const V3<double> L = normalize(sun_dir);
	if(std::fabs(height - kSkyInner) < 1e-3) {
		pos = pos + 1e-2 * normalize(pos);
		height = length(pos);
	}
	if(height < kSkyOuter) {
		if(height > kSkyInner) {
			const double pd = sky_enter(pos, dir, kSkyInner);
			if(pd >= 0.0) {
				const double c = std::max(0.1, dot(L, normalize(pos + pd * dir))) * 0.05;
				return {c, c, c};
			}
		} else {
			return {};
		}
		const double ray_depth = sky_leave(pos, dir, kSkyOuter);
		if(std::isinf(ray_depth) || std::isnan(ray_depth)) return {};
		const double start_offset = std::exp(kSkyScale / kSkyScaleDepth * (kSkyInner - height)) *
									sky_scale(dot(dir, pos) / height);
		const double seg = ray_depth / 64.0;
		const V3<double> step = dir * seg;
		const V3<double> inv_wl(1.0 / std::pow(double(0.650f), 4.0), 1.0 / std::pow(double(0.570f), 4.0),
								1.0 / std::pow(double(0.475f), 4.0));
		const double k_r4 = kSkyKr * 4.0 * kSkyPi, k_m4 = kSkyKm * 4.0 * kSkyPi;
		const double ln_max = std::log(double(std::numeric_limits<float>::max()));
		V3<double> color;
		int advanced = 0;
		for(int i = 0; i < 64; ++i) {
			const V3<double> sp = pos + (double(advanced) + 0.5) * step;
			const double h = length(sp);
			const double depth = std::exp(kSkyScale / kSkyScaleDepth * (kSkyInner - h));
			const double scatter = start_offset + depth * (sky_scale(dot(L, sp) / h) - sky_scale(dot(dir, sp) / h));
			bool bad = false;
			V3<double> att;
			for(int c = 0; c < 3; ++c) {
				const double e = -scatter * (inv_wl[c] * k_r4 + k_m4);
				if(std::isnan(e) || e > ln_max) bad = true;
				att[c] = std::exp(e);
			}
			if(bad) continue;
			color = color + att * (depth * seg * kSkyScale);
			++advanced;
		}
		const V3<double> secondary = color * kSkyKm * sun_color;
		color = color * inv_wl * kSkyKr * sun_color;
		if(show_sun) {
			const double mc = dot(L, -dir);
			const double g2 = kSkyG * kSkyG;
			const double phase = 1.5 * ((1.0 - g2) / (2.0 + g2)) * (1.0 + mc * mc) /
								 std::pow(std::max(1e-3, 1.0 + g2 - 2.0 * kSkyG * mc), 1.5);
			color = color + phase * secondary;
		}
		if(!f32_inf(color.x) && !f32_inf(color.y) && !f32_inf(color.z)) return color;
	} else {
		const double d = sky_enter(pos, dir, kSkyOuter);
		if(d > 0.0) {
			const double c = dot(L, normalize(pos + d * dir)) * 0.5;
			return {c * 0.5294117647, c * 0.80784313725, c * 0.92156862745};
		}
	}
	return {};
}

// =============================================================================================
// Closest-hit shading of one probe ray (src/shaders/closesthit.glsl:133-289, NO_REFLECTION) plus
// the miss shader with a constant sky.  Writes (rgb, depth).
// Per-ray radiance factors: for a non-metallic hit material the closest-hit radiance is affine in
// the material albedo a and the light radiance L (the sun colour):
//   rad = c0 + A (.) a + (B (.) a + C) (.) L
// with the shadow visibility and the sun-below-horizon fade folded in.  Misses: c0 = sky,
// back faces: 0.  (Metallic materials make F depend on a; the factors are only affine for
// metallic == 0.)
struct ShadeAux {
	double c0[3];        // emissive + specular * reflection probes
	double A[3];         // (1 - f0)(1 - metal) * sampleProbes(position, normal)
	double B[3];         // lit: albedo-linear part of pbrMetallicRoughness per unit light radiance
	double C[3];         // lit: albedo-independent part per unit light radiance
	uint32_t lit;        // shadow ray missed
	uint32_t material;   // material index of the hit (0xFFFFFFFF: miss / back face)
};

template <class T>
static void shade_ray(const Scene& s, const Grid& g, const Light& L, const SkyCfg& skycfg, const uint32_t* states,
					  const Atlas& irr, const Atlas& dep, const Ray32& r, const Hit32& h, T out[4], ShadeAux* aux) {
	V3<double> sky = skycfg.constant;
	if(!h.hit() && skycfg.model == 1)
		sky = sky_atmosphere(V3<double>(r.o[0], r.o[1], r.o[2]), V3<double>(r.d[0], r.d[1], r.d[2]), L.dir, L.color,
							 L.sky_brightness, true);
	if(aux) {
		std::memset(aux, 0, sizeof(*aux));
		aux->material = 0xFFFFFFFFu;
		if(!h.hit()) {
			aux->c0[0] = sky.x;
			aux->c0[1] = sky.y;
			aux->c0[2] = sky.z;
		}
	}
	if(!h.hit()) {  // miss.rmiss: depth -1, sky radiance
		out[0] = T(sky.x);
		out[1] = T(sky.y);
		out[2] = T(sky.z);
		out[3] = T(-1);
		return;
	}
	const float* tv = &s.tri[uint64_t(h.tri) * 9];
	const V3<T> o(T(r.o[0]), T(r.o[1]), T(r.o[2]));
	const V3<T> d(T(r.d[0]), T(r.d[1]), T(r.d[2]));
	T t, b1, b2;
	if(sizeof(T) == sizeof(double)) {
		// FP64 re-intersection of the chosen triangle (Moller-Trumbore) for t and barycentrics
		V3<double> v0(tv[0], tv[1], tv[2]), v1(tv[3], tv[4], tv[5]), v2(tv[6], tv[7], tv[8]);
		V3<double> od(r.o[0], r.o[1], r.o[2]), dd(r.d[0], r.d[1], r.d[2]);
		V3<double> e1 = v1 - v0, e2 = v2 - v0;
		V3<double> p(dd.y * e2.z - dd.z * e2.y, dd.z * e2.x - dd.x * e2.z, dd.x * e2.y - dd.y * e2.x);
		double det = dot(e1, p);
		V3<double> tvec = od - v0;
		V3<double> q(tvec.y * e1.z - tvec.z * e1.y, tvec.z * e1.x - tvec.x * e1.z, tvec.x * e1.y - tvec.y * e1.x);
		if(det != 0.0) {
			t = T(dot(e2, q) / det);
			b1 = T(dot(tvec, p) / det);
			b2 = T(dot(dd, q) / det);
		} else {
			t = T(h.t);
			b1 = T(h.u);
			b2 = T(h.v);
		}
	} else {
		t = T(h.t);
		b1 = T(h.u);
		b2 = T(h.v);
	}
	if(h.back) {  // closesthit.glsl:137-141
		out[0] = out[1] = out[2] = T(0);
		out[3] = t * T(0.8);
		return;
	}
	const V3<T> b(T(1) - b1 - b2, b1, b2);
	const V3<T> position = d * t + o;
	const uint32_t inst = s.tri_inst[h.tri], prim = s.tri_prim[h.tri];
	const ddgi_offset_entry& oe = s.offsets[s.instances[inst].custom_index];
	const uint32_t ib = oe.index_offset + 3 * prim;
	const ddgi_vertex& w0 = s.vertices[oe.vertex_offset + s.indices[ib + 0]];
	const ddgi_vertex& w1 = s.vertices[oe.vertex_offset + s.indices[ib + 1]];
	const ddgi_vertex& w2 = s.vertices[oe.vertex_offset + s.indices[ib + 2]];
	const ddgi_material& m = s.materials[oe.material_index];
	const V3<T> n0(T(w0.normal[0]), T(w0.normal[1]), T(w0.normal[2]));
	const V3<T> n1(T(w1.normal[0]), T(w1.normal[1]), T(w1.normal[2]));
	const V3<T> n2(T(w2.normal[0]), T(w2.normal[1]), T(w2.normal[2]));
	const V3<T> tsn = normalize(n0 * b.x + n1 * b.y + n2 * b.z);
	const double* N = &s.inst_normal[inst * 9];
	const V3<T> normal = normalize(V3<T>(T(N[0]) * tsn.x + T(N[1]) * tsn.y + T(N[2]) * tsn.z,
										 T(N[3]) * tsn.x + T(N[4]) * tsn.y + T(N[5]) * tsn.z,
										 T(N[6]) * tsn.x + T(N[7]) * tsn.y + T(N[8]) * tsn.z));
	const V3<T> albedo(T(m.base_color[0]), T(m.base_color[1]), T(m.base_color[2]));
	const T metal = T(m.metallic), rough = T(m.roughness);
	V3<T> color(T(m.emissive[0]), T(m.emissive[1]), T(m.emissive[2]));
	const V3<T> f0(T(0.04), T(0.04), T(0.04));
	const V3<T> diffuse = albedo * (V3<T>(1, 1, 1) - f0) * (T(1) - metal);
	const V3<T> spec = mix3(f0, albedo, metal);
	const V3<T> refl_dir = d - normal * (T(2) * dot(normal, d));
	const V3<T> reflection = sample_probes<T>(g, states, irr, dep, position, refl_dir, -d);
	color = color + spec * reflection;
	const V3<T> indirect = sample_probes<T>(g, states, irr, dep, position, normal, -d);
	color = color + indirect * diffuse;
	// shadow ray from the FP32 hit position, exactly as traced on the GPU (closesthit.glsl:270-281)
	Ray32 sr;
	for(int k = 0; k < 3; ++k) sr.o[k] = r.d[k] * h.t + r.o[k];
	sr.d[0] = float(L.dir.x);
	sr.d[1] = float(L.dir.y);
	sr.d[2] = float(L.dir.z);
	sr.tmin = 0.1f;
	sr.tmax = 10000.0f;
	ray32_prepare(sr);
	const bool shadowed = trace32(s, sr, 0xFFu, true, false).hit();
	T fade = T(1);
	if(!shadowed) {
		const V3<T> ldir(T(L.dir.x), T(L.dir.y), T(L.dir.z));
		color = color + pbr_metallic_roughness<T>(normal, normalize(-d), V3<T>(L.color), ldir, albedo, metal, rough);
		if(ldir.y < T(0)) fade = T(1) - clampT(-ldir.y, T(0), T(0.1)) / T(0.1);
		color = color * fade;
	}
	if(aux) {
		aux->material = oe.material_index;
		aux->lit = shadowed ? 0u : 1u;
		const T f = shadowed ? T(1) : fade;
		const V3<T> em(T(m.emissive[0]), T(m.emissive[1]), T(m.emissive[2]));
		const V3<T> c0 = (em + spec * reflection) * f;
		const V3<T> A = indirect * ((T(1) - T(0.04)) * (T(1) - metal) * f);
		V3<T> B, Cc;
		if(!shadowed) {
			const V3<T> ldir(T(L.dir.x), T(L.dir.y), T(L.dir.z));
			const V3<T> one(1, 1, 1);
			const V3<T> with = pbr_metallic_roughness<T>(normal, normalize(-d), one, ldir, one, metal, rough);
			const V3<T> zero = pbr_metallic_roughness<T>(normal, normalize(-d), one, ldir, V3<T>(), metal, rough);
			B = (with - zero) * f;
			Cc = zero * f;
		}
		for(int c = 0; c < 3; ++c) {
			aux->c0[c] = double(c0[c]);
			aux->A[c] = double(A[c]);
			aux->B[c] = double(B[c]);
			aux->C[c] = double(Cc[c]);
		}
	}
	out[0] = color.x;
	out[1] = color.y;
	out[2] = color.z;
	out[3] = t;
}

// =============================================================================================
// Ray generation for probe rays (src/shaders/traceProbes.rgen:29-43) in FP32.
static void fib_table32(uint32_t count, uint32_t n, std::vector<float>& f) {
	f.resize(size_t(count) * 3);
	for(uint32_t i = 0; i < count; ++i) {
		V3<double> v = spherical_fibonacci<double>(double(i), double(n));
		f[i * 3 + 0] = float(v.x);
		f[i * 3 + 1] = float(v.y);
		f[i * 3 + 2] = float(v.z);
	}
}

static void rotate32(const float R[9], const float* f, float out[3]) {
	for(int k = 0; k < 3; ++k) out[k] = R[k * 3 + 0] * f[0] + R[k * 3 + 1] * f[1] + R[k * 3 + 2] * f[2];
}

static void probe_origin32(const Grid& g, uint32_t linear, float o[3]) {
	int gi[3];
	g.grid_index(linear, gi);
	for(int k = 0; k < 3; ++k) {
		float cell = (g.p.extent_max[k] - g.p.extent_min[k]) / float(g.p.resolution[k] - 1);
		o[k] = float(gi[k]) * cell + g.p.extent_min[k];
	}
}


// =============================================================================================
// Blend (src/shaders/probesUpdate.glsl:49-119) + border copy (probesCopyBorders.comp) in T,
// updating double image-layout atlases in place.  Also returns each probe's maxChange and the
// out-of-range ray count that drive the state machine.
template <class T>
static void blend_probe(const Grid& g, const float* dirs, const double* rays, uint32_t slot, uint32_t linear,
						uint32_t* states, double* irr, double* dep, double* max_change_out, uint32_t* oor_out) {
	const uint32_t n = g.p.rays_per_probe;
	const V3<T> cell = g.cell<T>();
	const T cell_len = length(cell);
	const T h = T(g.p.hysteresis);
	int gi[3];
	g.grid_index(linear, gi);
	const double* rr = rays + uint64_t(slot) * n * 4;
	uint32_t out_of_range = 0;
	for(uint32_t i = 0; i < n; ++i) {
		T depth = T(rr[i * 4 + 3]);
		if(depth < T(0) || depth > cell_len) ++out_of_range;
	}
	// irradiance: interior 6x6 of the 8x8 tile
	{
		const uint32_t res = g.p.color_res, W = g.image_width(res);
		uint32_t ox, oy;
		g.tile_origin(gi, res, ox, oy);
		T max_change = 0;
		for(int cy = 0; cy < int(res) - 2; ++cy)
			for(int cx = 0; cx < int(res) - 2; ++cx) {
				const V3<T> td = oct_decode<T>(texel_coord_irradiance<T>(cx), texel_coord_irradiance<T>(cy));
				T acc[4] = {0, 0, 0, 0};
				for(uint32_t i = 0; i < n; ++i) {
					const V3<T> rd(T(dirs[i * 3 + 0]), T(dirs[i * 3 + 1]), T(dirs[i * 3 + 2]));
					const T w = std::max(T(0), dot(td, rd));
					acc[0] += w * T(rr[i * 4 + 0]);
					acc[1] += w * T(rr[i * 4 + 1]);
					acc[2] += w * T(rr[i * 4 + 2]);
					acc[3] += w;
				}
				if(acc[3] > T(1e-3))
					for(int c = 0; c < 3; ++c) acc[c] /= acc[3];
				double* px = &irr[(uint64_t(oy + 1 + cy) * W + ox + 1 + cx) * 3];
				T mc = 0;
				for(int c = 0; c < 3; ++c) {
					const T prev = T(px[c]);
					mc = std::max(mc, T(std::fabs(acc[c] - prev)));
					px[c] = double(mixT(acc[c], prev, h));
				}
				max_change = std::max(max_change, mc);
			}
		if(max_change_out) *max_change_out = double(max_change);
		// state machine (probesUpdate.glsl:110-118), with the intended max over all texels (H5)
		uint32_t& st = states[linear];
		if(out_of_range >= n) {
			st = 8;
		} else {
			const T mc = max_change;
			if(mc < T(0.02) / T(st)) st = std::min(st + 1u, 8u);
			else if(mc > T(0.04) / T(st)) st = std::max(st - 1u, 1u);
			else if(mc > T(0.25)) st = 1;
		}
	}
	if(oor_out) *oor_out = out_of_range;
	// depth: interior 14x14 of the 16x16 tile
	{
		const uint32_t res = g.p.depth_res, W = g.image_width(res);
		uint32_t ox, oy;
		g.tile_origin(gi, res, ox, oy);
		const T sharp = T(g.p.depth_sharpness);
		for(int cy = 0; cy < int(res) - 2; ++cy)
			for(int cx = 0; cx < int(res) - 2; ++cx) {
				const V3<T> td = oct_decode<T>(texel_coord_depth<T>(cx), texel_coord_depth<T>(cy));
				T acc[3] = {0, 0, 0};
				for(uint32_t i = 0; i < n; ++i) {
					const V3<T> rd(T(dirs[i * 3 + 0]), T(dirs[i * 3 + 1]), T(dirs[i * 3 + 2]));
					T depth = std::min(cell_len, T(rr[i * 4 + 3]));
					if(depth < T(0)) depth = cell_len;
					const T w = std::pow(std::max(T(0), dot(td, rd)), sharp);
					acc[0] += w * depth;
					acc[1] += w * depth * depth;
					acc[2] += w;
				}
				if(acc[2] > T(1e-3)) {
					acc[0] /= acc[2];
					acc[1] /= acc[2];
				}
				double* px = &dep[(uint64_t(oy + 1 + cy) * W + ox + 1 + cx) * 2];
				for(int c = 0; c < 2; ++c) px[c] = double(mixT(acc[c], T(px[c]), h));
			}
	}
}

// Border copy of one probe tile: the reference tables (probesCopyBorders.comp:21-220) are the
// octahedral mirror  (0,j)<-(1,R-1-j), (R-1,j)<-(R-2,R-1-j), (j,0)<-(R-1-j,1), (j,R-1)<-(R-1-j,R-2)
// and corners (0,0)<-(R-2,R-2), (0,R-1)<-(R-2,1), (R-1,0)<-(1,R-2), (R-1,R-1)<-(1,1), coordinates
// (x, y).  tests/golden/border_tables.json pins this against the tables.
static void border_copy(double* img, uint32_t W, uint32_t ch, uint32_t ox, uint32_t oy, uint32_t R) {
	auto px = [&](uint32_t x, uint32_t y) { return &img[(uint64_t(oy + y) * W + ox + x) * ch]; };
	auto cp = [&](uint32_t dx, uint32_t dy, uint32_t sx, uint32_t sy) {
		std::memcpy(px(dx, dy), px(sx, sy), sizeof(double) * ch);
	};
	cp(0, 0, R - 2, R - 2);
	cp(0, R - 1, R - 2, 1);
	cp(R - 1, 0, 1, R - 2);
	cp(R - 1, R - 1, 1, 1);
	for(uint32_t j = 1; j + 1 < R; ++j) {
		cp(0, j, 1, R - 1 - j);
		cp(R - 1, j, R - 2, R - 1 - j);
		cp(j, 0, R - 1 - j, 1);
		cp(j, R - 1, R - 1 - j, R - 2);
	}
}

static double ms_since(std::chrono::steady_clock::time_point t0) {
	return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace orc

using namespace orc;

namespace {
struct OracleScene {
	Scene s;
};

Grid make_grid(const ddgi_grid_params* p) {
	Grid g;
	g.p = *p;
	return g;
}

Light make_light(const ddgi_light* l) {
	Light L;
	L.dir = V3<double>(l->direction[0], l->direction[1], l->direction[2]);
	L.color = V3<double>(l->color[0], l->color[1], l->color[2]);
	L.sky_brightness = l->color[3];
	return L;
}

void fill_hit_record(const Scene& s, const Hit32& h, uint32_t* rec) {
	uint32_t tb;
	float t = h.hit() ? h.t : -1.0f;
	std::memcpy(&tb, &t, 4);
	rec[0] = tb;
	rec[1] = h.hit() ? s.tri_inst[h.tri] : 0xFFFFFFFFu;
	rec[2] = h.hit() ? s.tri_prim[h.tri] : 0xFFFFFFFFu;
	rec[3] = (h.hit() && h.back) ? 1u : 0u;
	std::memcpy(&rec[4], &h.u, 4);
	std::memcpy(&rec[5], &h.v, 4);
	if(!h.hit()) rec[4] = rec[5] = 0;
}
}  // namespace

extern "C" {

int orc_version() { return 1; }

void orc_set_threads(uint32_t n) { ThreadPool::GetInstance().resize(n); }
uint32_t orc_get_threads() { return ThreadPool::GetInstance().size(); }

// ---- math (closed-form tests) ----
void orc_fib64(uint32_t i, uint32_t n, double* out) {
	V3<double> v = spherical_fibonacci<double>(double(i), double(n));
	out[0] = v.x;
	out[1] = v.y;
	out[2] = v.z;
}
// float table: the FP64 formula rounded to FP32 (the ray directions both sides trace)
void orc_fib_table32(uint32_t count, uint32_t n, float* out) {
	std::vector<float> f;
	fib_table32(count, n, f);
	std::memcpy(out, f.data(), f.size() * sizeof(float));
}
void orc_oct_decode(double ox, double oy, double* out) {
	V3<double> v = oct_decode<double>(ox, oy);
	out[0] = v.x;
	out[1] = v.y;
	out[2] = v.z;
}
void orc_oct_encode(const double* v, double* out) { oct_encode<double>(V3<double>(v[0], v[1], v[2]), out[0], out[1]); }
void orc_oct_uv(const double* d, double* out) { sphere_to_oct_uv<double>(V3<double>(d[0], d[1], d[2]), out[0], out[1]); }
void orc_texel_dirs(uint32_t kind, double* out) {  // kind 0: 6x6 irradiance, 1: 14x14 depth; row-major (y, x)
	int R = kind == 0 ? 6 : 14;
	for(int y = 0; y < R; ++y)
		for(int x = 0; x < R; ++x) {
			double u = kind == 0 ? texel_coord_irradiance<double>(x) : texel_coord_depth<double>(x);
			double v = kind == 0 ? texel_coord_irradiance<double>(y) : texel_coord_depth<double>(y);
			V3<double> d = oct_decode<double>(u, v);
			out[(y * R + x) * 3 + 0] = d.x;
			out[(y * R + x) * 3 + 1] = d.y;
			out[(y * R + x) * 3 + 2] = d.z;
		}
}
int orc_tri_intersect32(const float* ray /*o3,tmin,d3,tmax*/, const float* v /*9*/, float* out /*t,u,v,det*/) {
	Ray32 r;
	for(int k = 0; k < 3; ++k) {
		r.o[k] = ray[k];
		r.d[k] = ray[4 + k];
	}
	r.tmin = ray[3];
	r.tmax = ray[7];
	ray32_prepare(r);
	float t, U, V, W, det;
	bool ok = tri_intersect32(r, v, v + 3, v + 6, t, U, V, W, det);
	if(!ok) return 0;
	out[0] = t;
	out[1] = V / det;
	out[2] = W / det;
	out[3] = det;
	return 1;
}

// ---- scene ----
void* orc_scene_create(const ddgi_vertex* vertices, uint64_t nv, const uint32_t* indices, uint64_t ni,
					   const ddgi_offset_entry* offsets, uint64_t no, const ddgi_instance* instances, uint64_t ninst,
					   const ddgi_material* materials, uint64_t nm) {
	auto* os = new OracleScene;
	Scene& s = os->s;
	s.vertices.assign(vertices, vertices + nv);
	s.indices.assign(indices, indices + ni);
	s.offsets.assign(offsets, offsets + no);
	s.instances.assign(instances, instances + ninst);
	s.materials.assign(materials, materials + nm);
	s.inst_mask.resize(ninst);
	s.inst_normal.resize(ninst * 9);
	for(uint64_t k = 0; k < ninst; ++k) {
		const ddgi_instance& in = instances[k];
		const float* M = in.transform;  // 3x4 row-major
		double a[9] = {M[0], M[1], M[2], M[4], M[5], M[6], M[8], M[9], M[10]};
		double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) + a[2] * (a[3] * a[7] - a[4] * a[6]);
		// normal matrix = transpose(inverse(A)) = cofactor(A) / det
		double cof[9] = {a[4] * a[8] - a[5] * a[7], -(a[3] * a[8] - a[5] * a[6]), a[3] * a[7] - a[4] * a[6],
						 -(a[1] * a[8] - a[2] * a[7]), a[0] * a[8] - a[2] * a[6], -(a[0] * a[7] - a[1] * a[6]),
						 a[1] * a[5] - a[2] * a[4], -(a[0] * a[5] - a[2] * a[3]), a[0] * a[4] - a[1] * a[3]};
		for(int q = 0; q < 9; ++q) s.inst_normal[k * 9 + q] = cof[q] / det;
		s.inst_mask[k] = in.mask & 0xFFu;
		const bool flip = (det < 0.0) != ((in.flags & DDGI_INSTANCE_FLIP_FACING) != 0);
		const ddgi_offset_entry& oe = offsets[in.custom_index];
		for(uint32_t p = 0; p < in.triangle_count; ++p) {
			float w[9];
			for(int vtx = 0; vtx < 3; ++vtx) {
				const ddgi_vertex& vv = vertices[oe.vertex_offset + indices[oe.index_offset + 3 * p + vtx]];
				for(int r = 0; r < 3; ++r) {
					// (double) row . (x, y, z, 1), left to right, rounded once to FP32
					double acc = double(M[r * 4 + 0]) * double(vv.pos[0]);
					acc = acc + double(M[r * 4 + 1]) * double(vv.pos[1]);
					acc = acc + double(M[r * 4 + 2]) * double(vv.pos[2]);
					acc = acc + double(M[r * 4 + 3]);
					w[vtx * 3 + r] = float(acc);
				}
			}
			s.tri.insert(s.tri.end(), w, w + 9);
			s.tri_inst.push_back(uint32_t(k));
			s.tri_prim.push_back(p);
			s.tri_flip.push_back(flip ? 1 : 0);
		}
	}
	s.build_bvh();
	return os;
}

void orc_scene_destroy(void* h) { delete static_cast<OracleScene*>(h); }
uint64_t orc_scene_tri_count(void* h) { return static_cast<OracleScene*>(h)->s.tri_count(); }
void orc_scene_world_tris(void* h, float* out) {
	const Scene& s = static_cast<OracleScene*>(h)->s;
	std::memcpy(out, s.tri.data(), s.tri.size() * sizeof(float));
}

// ---- ray queries: rays = n * 8 floats (o.xyz, tmin, d.xyz, tmax); out = n * 6 uint32 records
//      {t bits, instance, primitive, flags(back), u bits, v bits}; stats = {node visits, tri tests};
//      flags: 1 = any hit, 2 = cull back faces
void orc_trace(void* h, const float* rays, uint64_t n, uint32_t mask, int flags, int brute, uint32_t* out,
			   uint64_t* stats) {
	const bool any_hit = (flags & 1) != 0, cull = (flags & 2) != 0;
	const Scene& s = static_cast<OracleScene*>(h)->s;
	std::vector<uint64_t> st(2 * ((n + 255) / 256), 0);
	ThreadPool::parallel_for(n, 256, [&](uint64_t b, uint64_t e) {
		uint64_t* ls = &st[2 * (b / 256)];
		for(uint64_t i = b; i < e; ++i) {
			Ray32 r;
			for(int k = 0; k < 3; ++k) {
				r.o[k] = rays[i * 8 + k];
				r.d[k] = rays[i * 8 + 4 + k];
			}
			r.tmin = rays[i * 8 + 3];
			r.tmax = rays[i * 8 + 7];
			ray32_prepare(r);
			Hit32 hh = trace32(s, r, mask, any_hit, brute != 0, ls, cull);
			fill_hit_record(s, hh, &out[i * 6]);
		}
	});
	if(stats) {
		stats[0] = stats[1] = 0;
		for(size_t i = 0; i < st.size(); i += 2) {
			stats[0] += st[i];
			stats[1] += st[i + 1];
		}
	}
}

// ---- probe rays of one frame (traceProbes.rgen): origins/directions exactly as traced ----
void orc_probe_rays(const ddgi_grid_params* p, const float* R, const uint32_t* indices, uint32_t count, float* rays_out) {
	Grid g = make_grid(p);
	const uint32_t n = p->rays_per_probe;
	std::vector<float> fib;
	fib_table32(n, n, fib);
	float dx = p->extent_max[0] - p->extent_min[0], dy = p->extent_max[1] - p->extent_min[1],
		  dz = p->extent_max[2] - p->extent_min[2];
	const float tmax = std::sqrt(dx * dx + dy * dy + dz * dz);
	for(uint32_t s = 0; s < count; ++s) {
		float o[3];
		probe_origin32(g, indices[s], o);
		for(uint32_t i = 0; i < n; ++i) {
			float* r = &rays_out[(uint64_t(s) * n + i) * 8];
			float d[3];
			rotate32(R, &fib[i * 3], d);
			r[0] = o[0];
			r[1] = o[1];
			r[2] = o[2];
			r[3] = 0.01f;
			r[4] = d[0];
			r[5] = d[1];
			r[6] = d[2];
			r[7] = tmax;
		}
	}
}

// ---- one probe-update frame: trace + shade (T) ; blend + border + state machine (T) ----
// atlases are image layout double: irr W*H*3, dep W*H*2; updated in place.  rays_out (optional)
// count*n*4 doubles (rgb, depth); hits_out (optional) count*n*6 uint32; aux_out (optional)
// count*n ShadeAux; max_change/oor (optional) per slot.  phase_ms[3] = trace, blend, border.
int orc_update(void* h, const ddgi_grid_params* p, const ddgi_light* light, const double* sky, const float* R,
			   const uint32_t* indices, uint32_t count, uint32_t* states, double* irr, double* dep, double* rays_out,
			   uint32_t* hits_out, void* aux_out, double* max_change, uint32_t* oor, int fp32, double* phase_ms,
			   int sky_model) {
	const Scene& s = static_cast<OracleScene*>(h)->s;
	Grid g = make_grid(p);
	const Light L = make_light(light);
	SkyCfg skyc;
	skyc.constant = V3<double>(sky[0], sky[1], sky[2]);
	skyc.model = sky_model;
	const uint32_t n = p->rays_per_probe;
	const uint64_t nr = uint64_t(count) * n;
	std::vector<double> rays_local;
	double* rays = rays_out;
	if(!rays) {
		rays_local.resize(nr * 4);
		rays = rays_local.data();
	}
	Atlas A_irr{g.image_width(p->color_res), g.image_height(p->color_res), 3, irr};
	Atlas A_dep{g.image_width(p->depth_res), g.image_height(p->depth_res), 2, dep};
	std::vector<uint32_t> states_before(states, states + g.count());
	std::vector<float> fib;
	fib_table32(n, n, fib);
	std::vector<float> dirs(size_t(n) * 3);
	for(uint32_t i = 0; i < n; ++i) rotate32(R, &fib[i * 3], &dirs[i * 3]);
	float ex = p->extent_max[0] - p->extent_min[0], ey = p->extent_max[1] - p->extent_min[1],
		  ez = p->extent_max[2] - p->extent_min[2];
	const float tmax = std::sqrt(ex * ex + ey * ey + ez * ez);
	ShadeAux* aux = static_cast<ShadeAux*>(aux_out);

	auto t0 = std::chrono::steady_clock::now();
	// trace: parallel over probes, 4 probes per task
	ThreadPool::parallel_for(count, 4, [&](uint64_t b, uint64_t e) {
		for(uint64_t sl = b; sl < e; ++sl) {
			float o[3];
			probe_origin32(g, indices[sl], o);
			for(uint32_t i = 0; i < n; ++i) {
				Ray32 r;
				for(int k = 0; k < 3; ++k) {
					r.o[k] = o[k];
					r.d[k] = dirs[i * 3 + k];
				}
				r.tmin = 0.01f;
				r.tmax = tmax;
				ray32_prepare(r);
				const Hit32 hh = trace32(s, r, DDGI_MASK_STATIC | DDGI_MASK_DYNAMIC, false, false);
				const uint64_t ri = sl * n + i;
				if(hits_out) fill_hit_record(s, hh, &hits_out[ri * 6]);
				ShadeAux* ax = aux ? &aux[ri] : nullptr;
				if(fp32) {
					float c[4];
					shade_ray<float>(s, g, L, skyc, states_before.data(), A_irr, A_dep, r, hh, c, ax);
					for(int q = 0; q < 4; ++q) rays[ri * 4 + q] = c[q];
				} else {
					double c[4];
					shade_ray<double>(s, g, L, skyc, states_before.data(), A_irr, A_dep, r, hh, c, ax);
					for(int q = 0; q < 4; ++q) rays[ri * 4 + q] = c[q];
				}
			}
		}
	});
	const double trace_ms = ms_since(t0);
	t0 = std::chrono::steady_clock::now();
	// blend: parallel over probes, 16 probes per task (each probe touches only its own tiles/state)
	ThreadPool::parallel_for(count, 16, [&](uint64_t b, uint64_t e) {
		for(uint64_t sl = b; sl < e; ++sl) {
			double mc = 0;
			uint32_t o = 0;
			if(fp32)
				blend_probe<float>(g, dirs.data(), rays, uint32_t(sl), indices[sl], states, irr, dep, &mc, &o);
			else
				blend_probe<double>(g, dirs.data(), rays, uint32_t(sl), indices[sl], states, irr, dep, &mc, &o);
			if(max_change) max_change[sl] = mc;
			if(oor) oor[sl] = o;
		}
	});
	const double blend_ms = ms_since(t0);
	t0 = std::chrono::steady_clock::now();
	ThreadPool::parallel_for(count, 64, [&](uint64_t b, uint64_t e) {
		for(uint64_t sl = b; sl < e; ++sl) {
			int gi[3];
			g.grid_index(indices[sl], gi);
			uint32_t ox, oy;
			g.tile_origin(gi, p->color_res, ox, oy);
			border_copy(irr, g.image_width(p->color_res), 3, ox, oy, p->color_res);
			g.tile_origin(gi, p->depth_res, ox, oy);
			border_copy(dep, g.image_width(p->depth_res), 2, ox, oy, p->depth_res);
		}
	});
	const double border_ms = ms_since(t0);
	if(phase_ms) {
		phase_ms[0] = trace_ms;
		phase_ms[1] = blend_ms;
		phase_ms[2] = border_ms;
	}
	return 0;
}

uint64_t orc_shade_aux_size() { return sizeof(ShadeAux); }

// ---- blend only, from a given ray buffer (count*n*4 doubles) and FP32 world directions n*3 ----
void orc_blend(const ddgi_grid_params* p, const float* dirs, const double* rays, const uint32_t* indices, uint32_t count,
			   uint32_t* states, double* irr, double* dep, double* max_change, uint32_t* oor, int fp32) {
	Grid g = make_grid(p);
	ThreadPool::parallel_for(count, 16, [&](uint64_t b, uint64_t e) {
		for(uint64_t sl = b; sl < e; ++sl) {
			double mc = 0;
			uint32_t o = 0;
			if(fp32)
				blend_probe<float>(g, dirs, rays, uint32_t(sl), indices[sl], states, irr, dep, &mc, &o);
			else
				blend_probe<double>(g, dirs, rays, uint32_t(sl), indices[sl], states, irr, dep, &mc, &o);
			if(max_change) max_change[sl] = mc;
			if(oor) oor[sl] = o;
			int gi[3];
			g.grid_index(indices[sl], gi);
			uint32_t ox, oy;
			g.tile_origin(gi, p->color_res, ox, oy);
			border_copy(irr, g.image_width(p->color_res), 3, ox, oy, p->color_res);
			g.tile_origin(gi, p->depth_res, ox, oy);
			border_copy(dep, g.image_width(p->depth_res), 2, ox, oy, p->depth_res);
		}
	});
}

// ---- border copy only (probesCopyBorders.comp) of the listed probes, in place ----
void orc_border_copy(const ddgi_grid_params* p, const uint32_t* indices, uint32_t count, double* irr, double* dep) {
	Grid g = make_grid(p);
	for(uint32_t sl = 0; sl < count; ++sl) {
		int gi[3];
		g.grid_index(indices[sl], gi);
		uint32_t ox, oy;
		g.tile_origin(gi, p->color_res, ox, oy);
		border_copy(irr, g.image_width(p->color_res), 3, ox, oy, p->color_res);
		g.tile_origin(gi, p->depth_res, ox, oy);
		border_copy(dep, g.image_width(p->depth_res), 2, ox, oy, p->depth_res);
	}
}

// ---- sampleProbes at n points (double xyz arrays) ----
void orc_sample(const ddgi_grid_params* p, const uint32_t* states, const double* irr, const double* dep,
				const double* pos, const double* nrm, const double* tocam, uint64_t n, double* out, int fp32) {
	Grid g = make_grid(p);
	Atlas A_irr{g.image_width(p->color_res), g.image_height(p->color_res), 3, irr};
	Atlas A_dep{g.image_width(p->depth_res), g.image_height(p->depth_res), 2, dep};
	ThreadPool::parallel_for(n, 1024, [&](uint64_t b, uint64_t e) {
		for(uint64_t i = b; i < e; ++i) {
			if(fp32) {
				V3<float> r = sample_probes<float>(g, states, A_irr, A_dep, V3<float>(pos[i * 3], pos[i * 3 + 1], pos[i * 3 + 2]),
												   V3<float>(nrm[i * 3], nrm[i * 3 + 1], nrm[i * 3 + 2]),
												   V3<float>(tocam[i * 3], tocam[i * 3 + 1], tocam[i * 3 + 2]));
				out[i * 3] = r.x;
				out[i * 3 + 1] = r.y;
				out[i * 3 + 2] = r.z;
			} else {
				V3<double> r = sample_probes<double>(g, states, A_irr, A_dep, V3<double>(pos[i * 3], pos[i * 3 + 1], pos[i * 3 + 2]),
													 V3<double>(nrm[i * 3], nrm[i * 3 + 1], nrm[i * 3 + 2]),
													 V3<double>(tocam[i * 3], tocam[i * 3 + 1], tocam[i * 3 + 2]));
				out[i * 3] = r.x;
				out[i * 3 + 1] = r.y;
				out[i * 3 + 2] = r.z;
			}
		}
	});
}

// ---- pbrMetallicRoughness at one point (closed-form tests) ----
// ---- camera-effect passes: ray directions in FP32, operation by operation as the kernels form
// them (so that the traced rays are the same bits), shading in double ----
// sin / cos of an FP32 angle: Cody-Waite reduction by the nearest multiple of pi/2, Taylor series
// of sin (to r^11) and cos (to r^12) on |r| <= pi/4, quadrant sign/swap.
static void sincos_series32(float x, float& s, float& c) {
	const float q = std::nearbyint(x * 0.636619772f);
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
static V3<float> cross32(V3<float> a, V3<float> b) {
	return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
static float dot32(V3<float> a, V3<float> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static V3<float> normalize32(V3<float> a) {
	const float l = std::sqrt(dot32(a, a));
	return {a.x / l, a.y / l, a.z / l};
}
// rotateAxis (src/shaders/common.glsl:6-8)
static V3<float> rotate_axis32(V3<float> p, V3<float> axis, float angle) {
	float sn, cs;
	sincos_series32(angle, sn, cs);
	const V3<float> along = axis * dot32(axis, p);
	const V3<float> m = along * (1.0f - cs) + p * cs;
	return m + cross32(axis, p) * sn;
}

// scale() of src/shaders/sky.glsl:44-48 (O'Neil's fit of the optical depth).
double orc_sky_scale(double c) { return sky_scale(c); }

// sky() (src/shaders/sky.glsl:60-127) for `count` rays: origins and dirs float3 (FP32 values, as
// a shader sees them), sun direction / colour float3; writes count * 3 doubles.
void orc_sky(uint64_t count, const float* origins, const float* dirs, const float* sun_dir, const float* sun_color,
			 double brightness, int show_sun, double* out) {
	const V3<double> L(sun_dir[0], sun_dir[1], sun_dir[2]), C(sun_color[0], sun_color[1], sun_color[2]);
	ThreadPool::parallel_for(count, 1024, [&](uint64_t b, uint64_t e) {
		for(uint64_t i = b; i < e; ++i) {
			const V3<double> c =
				sky_atmosphere(V3<double>(origins[i * 3], origins[i * 3 + 1], origins[i * 3 + 2]),
							   V3<double>(dirs[i * 3], dirs[i * 3 + 1], dirs[i * 3 + 2]), L, C, brightness, show_sun != 0);
			out[i * 3] = c.x;
			out[i * 3 + 1] = c.y;
			out[i * 3 + 2] = c.z;
		}
	});
}

double orc_sin_series(float x) {
	float sn, cs;
	sincos_series32(x, sn, cs);
	return sn;
}
double orc_cos_series(float x) {
	float sn, cs;
	sincos_series32(x, sn, cs);
	return cs;
}

// directLight.rgen:42-99 for `count` G-buffer texels (16 floats each, ddgi_gbuffer_texel); sdir =
// normalize(light direction) in FP32; noise count * 2 or null; previous count * 4 or null; writes
// count * 4 floats.  All arithmetic FP32 as the shader forms it: the result is exact.
void orc_direct_light(void* h, const float* sdir, const float* gbuffer, uint64_t count, const float* noise,
					  const float* previous, float* out) {
	const Scene& s = static_cast<OracleScene*>(h)->s;
	const float kPi = 3.1415926538f;
	ThreadPool::parallel_for(count, 256, [&](uint64_t b, uint64_t e) {
		for(uint64_t i = b; i < e; ++i) {
			const float* g = gbuffer + i * 16;
			float* o = out + i * 4;
			if(!(g[3] > 0.0f)) {
				o[0] = o[1] = o[2] = o[3] = -1.0f;
				continue;
			}
			const V3<float> normal(g[4], g[5], g[6]);
			V3<float> dir(sdir[0], sdir[1], sdir[2]);
			if(noise) {
				const float nx = noise[i * 2], ny = noise[i * 2 + 1];
				const V3<float> temp = rotate_axis32(dir, normalize32(cross32(normal, dir)), 2.0f * (nx - 0.5f) * 0.02f);
				dir = rotate_axis32(temp, dir, (2.0f * kPi) * ny);
			}
			o[0] = o[1] = o[2] = o[3] = 0.0f;
			if(dot32(dir, normal) > 0.0f) {
				Ray32 r;
				for(int k = 0; k < 3; ++k) {
					r.o[k] = g[k];
					r.d[k] = dir[k];
				}
				r.tmin = 0.01f;
				r.tmax = 10000.0f;
				ray32_prepare(r);
				const Hit32 hh = trace32(s, r, 0xFFu, true, false);
				float vis = 0.0f;
				if(!hh.hit()) {
					vis = 1.0f;
					if(dir.y < 0.0f) vis *= 1.0f - std::min(std::max(-dir.y, 0.0f), 0.1f) / 0.1f;
				}
				o[0] = vis;
				o[1] = previous ? previous[i * 4 + 1] : 0.0f;
				o[2] = previous ? previous[i * 4 + 2] : 0.0f;
				o[3] = 1.0f;
			}
		}
	});
}

// reflection.rgen:101-175 for `count` G-buffer texels: the reflected ray in FP32 as the shader
// forms it, traced with the FP32 replica, shaded by closesthit.glsl (recursion depth 1) in double;
// writes count * 4 doubles (colorCompression(rgb), roughness) and, when hits_out is given, the
// ray's hit record (6 uint32, as orc_trace) for traced texels (0xFFFFFFFF instance otherwise).
void orc_reflection(void* h, const ddgi_grid_params* p, const ddgi_light* light, const double* sky, int sky_model,
					const uint32_t* states, const double* irr, const double* dep, const float* gbuffer, uint64_t count,
					const float* origin, const float* noise, double* out, uint32_t* hits_out) {
	const Scene& s = static_cast<OracleScene*>(h)->s;
	const Grid g = make_grid(p);
	const Light L = make_light(light);
	SkyCfg skyc;
	skyc.constant = V3<double>(sky[0], sky[1], sky[2]);
	skyc.model = sky_model;
	const Atlas A_irr{g.image_width(p->color_res), g.image_height(p->color_res), 3, irr};
	const Atlas A_dep{g.image_width(p->depth_res), g.image_height(p->depth_res), 2, dep};
	const float kPi = 3.1415926538f;
	ThreadPool::parallel_for(count, 64, [&](uint64_t b, uint64_t e) {
		for(uint64_t i = b; i < e; ++i) {
			const float* t = gbuffer + i * 16;
			double* o = out + i * 4;
			o[0] = o[1] = o[2] = o[3] = 0.0;
			if(hits_out) {
				uint32_t* hr = hits_out + i * 6;
				const float neg = -1.0f;
				std::memcpy(&hr[0], &neg, 4);
				hr[1] = hr[2] = 0xFFFFFFFFu;
				hr[3] = hr[4] = hr[5] = 0u;
			}
			const float rough = t[11], metal = t[7];
			if(!(t[3] > 0.0f && (rough < 0.4f || metal > 0.01f))) continue;
			const V3<float> position(t[0], t[1], t[2]), normal(t[4], t[5], t[6]);
			const V3<float> to_origin = normalize32(V3<float>(origin[0], origin[1], origin[2]) - position);
			const V3<float> incident(-to_origin.x, -to_origin.y, -to_origin.z);
			const V3<float> rdir = normalize32(incident - normal * (2.0f * dot32(normal, incident)));
			V3<float> dir = rdir;
			if(noise) {
				const float nx = noise[i * 2], ny = noise[i * 2 + 1];
				const float theta = rough * (nx - 0.5f) * 2.0f * kPi;
				const float phi = (ny - 0.5f) * 2.0f * kPi;
				const V3<float> tangent = dot32(rdir, normal) < 0.9f ? normalize32(cross32(rdir, normal))
																	 : normalize32(cross32(rdir, V3<float>(1.0f, 0.0f, 0.0f)));
				dir = rotate_axis32(rdir, tangent, theta);
				dir = rotate_axis32(dir, rdir, phi);
			}
			Ray32 r;
			for(int k = 0; k < 3; ++k) {
				r.o[k] = position[k];
				r.d[k] = dir[k];
			}
			r.tmin = 0.1f;
			r.tmax = 10000.0f;
			ray32_prepare(r);
			const Hit32 hh = trace32(s, r, 0xFFu, false, false);
			if(hits_out) fill_hit_record(s, hh, &hits_out[i * 6]);
			double c[4];
			shade_ray<double>(s, g, L, skyc, states, A_irr, A_dep, r, hh, c, nullptr);
			for(int k = 0; k < 3; ++k) o[k] = c[k] * (1.0 + c[k] / 1.0) / (1.0 + c[k]);  // reinhard_whitepoint(v, 1)
			o[3] = rough;
		}
	});
}

void orc_pbr(const double* n, const double* v, const double* lc, const double* ld, const double* albedo, double metal,
			 double rough, double* out) {
	V3<double> r = pbr_metallic_roughness<double>(V3<double>(n[0], n[1], n[2]), V3<double>(v[0], v[1], v[2]),
												  V3<double>(lc[0], lc[1], lc[2]), V3<double>(ld[0], ld[1], ld[2]),
												  V3<double>(albedo[0], albedo[1], albedo[2]), metal, rough);
	out[0] = r.x;
	out[1] = r.y;
	out[2] = r.z;
}

// ---- probe-state init (probesInit.rgen:31-64), FP32 replica (all decisions are discrete) ----
void orc_init_probes(void* h, const ddgi_grid_params* p, const float* R, uint32_t* states_out, uint32_t* backface_out) {
	const Scene& s = static_cast<OracleScene*>(h)->s;
	Grid g = make_grid(p);
	const uint32_t n = p->rays_per_probe;
	std::vector<float> fib;
	fib_table32(512, n, fib);
	float cell[3];
	for(int k = 0; k < 3; ++k) cell[k] = (p->extent_max[k] - p->extent_min[k]) / float(p->resolution[k] - 1);
	const float max_distance = std::sqrt(cell[0] * cell[0] + cell[1] * cell[1] + cell[2] * cell[2]);
	const float tmax = 1.5f * max_distance;
	ThreadPool::parallel_for(g.count(), 8, [&](uint64_t b, uint64_t e) {
		for(uint64_t lin = b; lin < e; ++lin) {
			float o[3];
			probe_origin32(g, uint32_t(lin), o);
			uint32_t back = 0;
			bool affect = false;
			for(uint32_t i = 0; i < 512; ++i) {
				Ray32 r;
				float d[3];
				rotate32(R, &fib[i * 3], d);
				for(int k = 0; k < 3; ++k) {
					r.o[k] = o[k];
					r.d[k] = d[k];
				}
				r.tmin = 0.01f;
				r.tmax = tmax;
				ray32_prepare(r);
				const Hit32 hh = trace32(s, r, DDGI_MASK_STATIC, false, false);
				const float depth = hh.hit() ? hh.t : 3.402823466e+38f;
				if(depth < max_distance) {
					if(hh.back) ++back;
					bool inside = true;
					for(int k = 0; k < 3; ++k) {
						float pos = o[k] + depth * d[k];
						float dist = std::fabs(pos - o[k]);
						inside = inside && (dist < cell[k]);
					}
					if(inside) affect = true;
				}
			}
			uint32_t st;
			if(float(back) > 0.5f * float(n)) st = 0;
			else st = affect ? 1u : 8u;
			states_out[lin] = st;
			if(backface_out) backface_out[lin] = back;
		}
	});
}

// ---- selectProbesToUpdate (src/IrradianceProbes.cpp:396-424); loop_index and last_offset are the
//      reference's function-static / member counters, passed in and out explicitly ----
uint32_t orc_select(const uint32_t* states, uint32_t P, uint32_t cap, uint32_t* last_offset, uint32_t* loop_index,
					uint32_t* out) {
	uint32_t idx = *last_offset, checked = 0, cnt = 0;
	while(checked < P && (cap == 0 || cnt < cap)) {
		if(states[idx] != 0 && ((idx + *loop_index) % states[idx]) == 0) out[cnt++] = idx;
		++idx;
		if(idx >= P) {
			idx = 0;
			++*loop_index;
		}
		++checked;
	}
	*last_offset = idx;
	return cnt;
}

// ---- differentiable probes: one update step as a function of (albedo[M][3], sun[3]) ----
// rad_i = c0 + A (.) a[m_i] + (B (.) a[m_i] + C) (.) sun                  (ShadeAux factors)
// irradiance texel = (1 - h) * sum_i w_ti rad_i / sum_i w_ti + h * prev   (probesUpdate.glsl)
//                    then the border copy
// out_x = sampleProbes(pos_x, n_x, v_x) on that atlas (depth atlas and states fixed)
// shade_x = out_x (.) a[g_x] * (1 - 0.004)                               (FinalGather.frag:62-72)
// loss = sum (shade - target)^2 / (3 N)
// Returns the loss; with g_albedo / g_sun non-null also the exact FP64 gradient (hand-written
// adjoint of every step).  irr_out (optional) receives the differentiated irradiance atlas.
double orc_diff(const ddgi_grid_params* p, const uint32_t* states, const double* irr_prev, const double* dep,
				const float* dirs, const uint32_t* indices, uint32_t count, const void* aux_v, const double* albedo,
				uint32_t M, const double* sun, const double* pos, const double* nrm, const double* tocam,
				const uint32_t* gmat, const double* target, uint64_t N, double* irr_out, double* shade_out,
				double* g_albedo, double* g_sun) {
	Grid g = make_grid(p);
	const ShadeAux* aux = static_cast<const ShadeAux*>(aux_v);
	const uint32_t n = p->rays_per_probe;
	const double h = double(p->hysteresis);
	const uint32_t W8 = g.image_width(8), H8 = g.image_height(8);
	const size_t irr_size = size_t(W8) * H8 * 3;
	// radiance
	std::vector<double> rad(size_t(count) * n * 3);
	for(size_t r = 0; r < size_t(count) * n; ++r) {
		const ShadeAux& x = aux[r];
		for(int c = 0; c < 3; ++c) {
			double v = x.c0[c];
			if(x.material != 0xFFFFFFFFu) {
				const double a = albedo[size_t(x.material) * 3 + c];
				v += x.A[c] * a + (x.B[c] * a + x.C[c]) * sun[c];
			}
			rad[r * 3 + c] = v;
		}
	}
	// texel directions, weights and normalisation (per frame)
	std::vector<double> wts(size_t(36) * n), inv(36);
	for(int t = 0; t < 36; ++t) {
		const V3<double> td = oct_decode<double>(texel_coord_irradiance<double>(t % 6), texel_coord_irradiance<double>(t / 6));
		double s = 0;
		for(uint32_t i = 0; i < n; ++i) {
			const V3<double> rd(dirs[i * 3], dirs[i * 3 + 1], dirs[i * 3 + 2]);
			wts[size_t(t) * n + i] = std::max(0.0, dot(td, rd));
			s += wts[size_t(t) * n + i];
		}
		inv[t] = s > 1e-3 ? 1.0 / s : 1.0;
	}
	std::vector<double> irr(irr_prev, irr_prev + irr_size);
	for(uint32_t sl = 0; sl < count; ++sl) {
		int gi[3];
		g.grid_index(indices[sl], gi);
		uint32_t ox, oy;
		g.tile_origin(gi, 8, ox, oy);
		for(int t = 0; t < 36; ++t) {
			double S[3] = {0, 0, 0};
			for(uint32_t i = 0; i < n; ++i)
				for(int c = 0; c < 3; ++c) S[c] += wts[size_t(t) * n + i] * rad[(size_t(sl) * n + i) * 3 + c];
			double* px = &irr[(size_t(oy + 1 + t / 6) * W8 + ox + 1 + t % 6) * 3];
			for(int c = 0; c < 3; ++c) px[c] = (1.0 - h) * (S[c] * inv[t]) + h * px[c];
		}
		border_copy(irr.data(), W8, 3, ox, oy, 8);
	}
	if(irr_out) std::memcpy(irr_out, irr.data(), irr_size * sizeof(double));
	Atlas A_irr{W8, H8, 3, irr.data()};
	Atlas A_dep{g.image_width(16), g.image_height(16), 2, dep};
	// G-buffer: forward (and per-chunk adjoint partials, reduced in chunk order)
	const uint64_t chunk = 4096, nchunks = (N + chunk - 1) / chunk;
	std::vector<double> loss_part(nchunks, 0.0);
	const bool grad = g_albedo != nullptr;
	std::vector<std::vector<double>> girr_part(grad ? nchunks : 0), galb_part(grad ? nchunks : 0);
	ThreadPool::parallel_for(N, chunk, [&](uint64_t b, uint64_t e) {
		const uint64_t ci = b / chunk;
		double lsum = 0;
		std::vector<double>* gi = nullptr;
		std::vector<double>* ga = nullptr;
		if(grad) {
			girr_part[ci].assign(irr_size, 0.0);
			galb_part[ci].assign(size_t(M) * 3, 0.0);
			gi = &girr_part[ci];
			ga = &galb_part[ci];
		}
		for(uint64_t x = b; x < e; ++x) {
			SampleCtx<double> ctx;
			sample_collect<double>(g, states, A_dep, V3<double>(pos[x * 3], pos[x * 3 + 1], pos[x * 3 + 2]),
								   V3<double>(nrm[x * 3], nrm[x * 3 + 1], nrm[x * 3 + 2]),
								   V3<double>(tocam[x * 3], tocam[x * 3 + 1], tocam[x * 3 + 2]), ctx);
			const V3<double> out = sample_eval(ctx, A_irr);
			const uint32_t m = gmat[x];
			V3<double> gout;
			for(int c = 0; c < 3; ++c) {
				const double a = albedo[size_t(m) * 3 + c];
				const double sh = out[c] * a * (1.0 - 0.004);
				if(shade_out) shade_out[x * 3 + c] = sh;
				const double diff = sh - target[x * 3 + c];
				lsum += diff * diff;
				if(grad) {
					const double gs = 2.0 * diff / double(3 * N);
					(*ga)[size_t(m) * 3 + c] += gs * out[c] * (1.0 - 0.004);
					gout[c] = gs * a * (1.0 - 0.004);
				}
			}
			if(grad) sample_adjoint(ctx, A_irr, gout, gi->data());
		}
		loss_part[ci] = lsum;
	});
	double loss = 0;
	for(double v : loss_part) loss += v;
	loss /= double(3 * N);
	if(!grad) return loss;
	std::vector<double> girr(irr_size, 0.0);
	std::fill(g_albedo, g_albedo + size_t(M) * 3, 0.0);
	for(uint64_t ci = 0; ci < nchunks; ++ci) {
		for(size_t k = 0; k < irr_size; ++k) girr[k] += girr_part[ci][k];
		for(size_t k = 0; k < size_t(M) * 3; ++k) g_albedo[k] += galb_part[ci][k];
	}
	for(int c = 0; c < 3; ++c) g_sun[c] = 0.0;
	// adjoint of the border copy, hysteresis, normalisation and ray sum, then of the radiance
	for(uint32_t sl = 0; sl < count; ++sl) {
		int gi[3];
		g.grid_index(indices[sl], gi);
		uint32_t ox, oy;
		g.tile_origin(gi, 8, ox, oy);
		auto G = [&](uint32_t x, uint32_t y) { return &girr[(size_t(oy + y) * W8 + ox + x) * 3]; };
		auto fold = [&](uint32_t dx, uint32_t dy, uint32_t sx, uint32_t sy) {
			for(int c = 0; c < 3; ++c) G(sx, sy)[c] += G(dx, dy)[c];
		};
		const uint32_t R = 8;
		fold(0, 0, R - 2, R - 2);
		fold(0, R - 1, R - 2, 1);
		fold(R - 1, 0, 1, R - 2);
		fold(R - 1, R - 1, 1, 1);
		for(uint32_t j = 1; j + 1 < R; ++j) {
			fold(0, j, 1, R - 1 - j);
			fold(R - 1, j, R - 2, R - 1 - j);
			fold(j, 0, R - 1 - j, 1);
			fold(j, R - 1, R - 1 - j, R - 2);
		}
		for(uint32_t i = 0; i < n; ++i) {
			double grad_rad[3] = {0, 0, 0};
			for(int t = 0; t < 36; ++t) {
				const double* gt = G(1 + t % 6, 1 + t / 6);
				const double k = (1.0 - h) * inv[t] * wts[size_t(t) * n + i];
				for(int c = 0; c < 3; ++c) grad_rad[c] += k * gt[c];
			}
			const ShadeAux& x = aux[size_t(sl) * n + i];
			if(x.material == 0xFFFFFFFFu) continue;
			for(int c = 0; c < 3; ++c) {
				const double a = albedo[size_t(x.material) * 3 + c];
				g_albedo[size_t(x.material) * 3 + c] += grad_rad[c] * (x.A[c] + x.B[c] * sun[c]);
				g_sun[c] += grad_rad[c] * (x.B[c] * a + x.C[c]);
			}
		}
	}
	return loss;
}

}  // extern "C"
