// Scene ingestion: the reference's `.scene` container and glTF 2.0 (.gltf / .glb) loaders, its
// `.mat` material files and `.scene` writer (src/Scene.cpp:30-49, :115-533, :647-684, :710-934),
// flattened into the tables ddgi_set_scene takes the way the reference's renderer builds its offset
// table and TLAS instances (src/Renderer.cpp:97-126, :512-551).  Host-only C++17 (no GPU code).
//
// The scene model is the reference's: meshes (one per glTF primitive, each with a default material),
// a node tree under a root node (local column-major 4x4 transforms; a node may render one mesh) and
// materials / texture records.  Flattening:
//   offset table  one entry per mesh with indices, in mesh order: {default material, first vertex,
//                 first index}, meshes packed back to back;
//   instances     one per node that renders a mesh, ordered by (material, mesh) with ties in node
//                 order (the reference sorts its renderer pool with the same key, ties unspecified);
//                 transform = the node's global transform (parent chain below the root, the root's
//                 own transform not applied), mask STATIC, no facing flip, custom index = the mesh's
//                 offset entry;
//   bounds        union over rendering nodes of the mesh bounds' two corners mapped by the node's
//                 transform from the root (src/Scene.cpp:1077-1103, src/Bounds.hpp:38-45).
#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <limits>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/ddgi.h"

namespace ddgi {
namespace host {
void record_error(const std::string& msg);  // api.hip: the thread's last error (ddgi_last_error(NULL))
}
}  // namespace ddgi

namespace {

namespace fs = std::filesystem;

struct Failure : std::runtime_error {
	ddgi_status status;
	Failure(ddgi_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};
[[noreturn]] void bad(const std::string& m) { throw Failure(DDGI_ERROR_INVALID_ARGUMENT, m); }

// ============================================================================================
// JSON document (objects keep their key order, so a written file reads back the same way)
struct Json {
	enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
	bool b = false;
	double num = 0.0;
	std::string str;
	std::vector<Json> arr;
	std::vector<std::pair<std::string, Json>> obj;

	static Json number(double v) {
		Json j;
		j.kind = Number;
		j.num = v;
		return j;
	}
	static Json string(std::string s) {
		Json j;
		j.kind = String;
		j.str = std::move(s);
		return j;
	}
	static Json array() {
		Json j;
		j.kind = Array;
		return j;
	}
	static Json object() {
		Json j;
		j.kind = Object;
		return j;
	}
	const Json* find(const std::string& k) const {
		if(kind != Object) return nullptr;
		for(const auto& kv : obj)
			if(kv.first == k) return &kv.second;
		return nullptr;
	}
	bool has(const std::string& k) const { return find(k) != nullptr; }
	const Json& at(const std::string& k) const {
		const Json* j = find(k);
		if(!j) bad("JSON: missing key '" + k + "'");
		return *j;
	}
	const Json& at(size_t i) const {
		if(kind != Array || i >= arr.size()) bad("JSON: index " + std::to_string(i) + " out of range");
		return arr[i];
	}
	Json& set(const std::string& k, Json v) {
		for(auto& kv : obj)
			if(kv.first == k) return kv.second = std::move(v);
		obj.emplace_back(k, std::move(v));
		return obj.back().second;
	}
	double as_number() const {
		if(kind != Number) bad("JSON: number expected");
		return num;
	}
	long long as_int() const {
		const double v = as_number();
		if(v != std::floor(v) || std::fabs(v) > 9.0e15) bad("JSON: integer expected");
		return (long long)v;
	}
	const std::string& as_string() const {
		if(kind != String) bad("JSON: string expected");
		return str;
	}
	const std::vector<Json>& items() const {
		static const std::vector<Json> none;
		if(kind == Null) return none;
		if(kind != Array) bad("JSON: array expected");
		return arr;
	}
	long long get_int(const std::string& k, long long dflt) const {
		const Json* j = find(k);
		return j ? j->as_int() : dflt;
	}
	std::string get_string(const std::string& k, const std::string& dflt) const {
		const Json* j = find(k);
		return j ? j->as_string() : dflt;
	}
};

class JsonParser {
  public:
	JsonParser(const char* p, size_t n) : p_(p), end_(p + n) {}
	Json document() {
		Json j = value(0);
		ws();
		if(p_ != end_) fail("trailing characters");
		return j;
	}

  private:
	const char* p_;
	const char* end_;
	[[noreturn]] void fail(const char* what) { bad(std::string("JSON parse error: ") + what); }
	void ws() {
		while(p_ != end_ && (*p_ == ' ' || *p_ == '\t' || *p_ == '\n' || *p_ == '\r')) ++p_;
	}
	bool lit(const char* w) {
		const size_t n = std::strlen(w);
		if(size_t(end_ - p_) < n || std::memcmp(p_, w, n) != 0) return false;
		p_ += n;
		return true;
	}
	Json value(int depth) {
		if(depth > 256) fail("nesting too deep");
		ws();
		if(p_ == end_) fail("unexpected end");
		Json j;
		switch(*p_) {
			case '{': {
				++p_;
				j.kind = Json::Object;
				ws();
				if(p_ != end_ && *p_ == '}') {
					++p_;
					return j;
				}
				for(;;) {
					ws();
					if(p_ == end_ || *p_ != '"') fail("object key expected");
					std::string k = string_body();
					ws();
					if(p_ == end_ || *p_ != ':') fail("':' expected");
					++p_;
					j.obj.emplace_back(std::move(k), value(depth + 1));
					ws();
					if(p_ != end_ && *p_ == ',') {
						++p_;
						continue;
					}
					if(p_ != end_ && *p_ == '}') {
						++p_;
						return j;
					}
					fail("',' or '}' expected");
				}
			}
			case '[': {
				++p_;
				j.kind = Json::Array;
				ws();
				if(p_ != end_ && *p_ == ']') {
					++p_;
					return j;
				}
				for(;;) {
					j.arr.push_back(value(depth + 1));
					ws();
					if(p_ != end_ && *p_ == ',') {
						++p_;
						continue;
					}
					if(p_ != end_ && *p_ == ']') {
						++p_;
						return j;
					}
					fail("',' or ']' expected");
				}
			}
			case '"': j.kind = Json::String; j.str = string_body(); return j;
			case 't': if(!lit("true")) fail("bad literal"); j.kind = Json::Bool; j.b = true; return j;
			case 'f': if(!lit("false")) fail("bad literal"); j.kind = Json::Bool; j.b = false; return j;
			case 'n': if(!lit("null")) fail("bad literal"); return j;
			default: {
				const char* s = p_;
				if(p_ != end_ && (*p_ == '-' || *p_ == '+')) ++p_;
				while(p_ != end_ && (std::isdigit((unsigned char)*p_) || *p_ == '.' || *p_ == 'e' || *p_ == 'E' ||
									 *p_ == '-' || *p_ == '+'))
					++p_;
				if(p_ == s) fail("unexpected character");
				std::string t(s, p_);
				char* e = nullptr;
				const double v = std::strtod(t.c_str(), &e);
				if(!e || *e != '\0') fail("bad number");
				j.kind = Json::Number;
				j.num = v;
				return j;
			}
		}
	}
	static void utf8(std::string& out, uint32_t c) {
		if(c < 0x80) out += char(c);
		else if(c < 0x800) {
			out += char(0xC0 | (c >> 6));
			out += char(0x80 | (c & 0x3F));
		} else if(c < 0x10000) {
			out += char(0xE0 | (c >> 12));
			out += char(0x80 | ((c >> 6) & 0x3F));
			out += char(0x80 | (c & 0x3F));
		} else {
			out += char(0xF0 | (c >> 18));
			out += char(0x80 | ((c >> 12) & 0x3F));
			out += char(0x80 | ((c >> 6) & 0x3F));
			out += char(0x80 | (c & 0x3F));
		}
	}
	uint32_t hex4() {
		if(end_ - p_ < 4) fail("bad \\u escape");
		uint32_t v = 0;
		for(int i = 0; i < 4; ++i) {
			const char c = *p_++;
			v <<= 4;
			if(c >= '0' && c <= '9') v |= uint32_t(c - '0');
			else if(c >= 'a' && c <= 'f') v |= uint32_t(c - 'a' + 10);
			else if(c >= 'A' && c <= 'F') v |= uint32_t(c - 'A' + 10);
			else fail("bad \\u escape");
		}
		return v;
	}
	std::string string_body() {
		++p_;  // opening quote
		std::string out;
		while(p_ != end_ && *p_ != '"') {
			char c = *p_++;
			if(c != '\\') {
				out += c;
				continue;
			}
			if(p_ == end_) fail("bad escape");
			c = *p_++;
			switch(c) {
				case '"': out += '"'; break;
				case '\\': out += '\\'; break;
				case '/': out += '/'; break;
				case 'b': out += '\b'; break;
				case 'f': out += '\f'; break;
				case 'n': out += '\n'; break;
				case 'r': out += '\r'; break;
				case 't': out += '\t'; break;
				case 'u': {
					uint32_t cp = hex4();
					if(cp >= 0xD800 && cp < 0xDC00 && end_ - p_ >= 6 && p_[0] == '\\' && p_[1] == 'u') {
						p_ += 2;
						const uint32_t lo = hex4();
						cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
					}
					utf8(out, cp);
					break;
				}
				default: fail("bad escape");
			}
		}
		if(p_ == end_) fail("unterminated string");
		++p_;
		return out;
	}
};

void dump(const Json& j, std::string& out) {
	switch(j.kind) {
		case Json::Null: out += "null"; break;
		case Json::Bool: out += j.b ? "true" : "false"; break;
		case Json::Number: {
			char buf[40];
			if(j.num == std::floor(j.num) && std::fabs(j.num) < 1e15) std::snprintf(buf, sizeof buf, "%.0f", j.num);
			else std::snprintf(buf, sizeof buf, "%.17g", j.num);
			out += buf;
			break;
		}
		case Json::String: {
			out += '"';
			for(unsigned char c : j.str) {
				if(c == '"') out += "\\\"";
				else if(c == '\\') out += "\\\\";
				else if(c < 0x20) {
					char buf[8];
					std::snprintf(buf, sizeof buf, "\\u%04x", c);
					out += buf;
				} else out += char(c);
			}
			out += '"';
			break;
		}
		case Json::Array:
			out += '[';
			for(size_t i = 0; i < j.arr.size(); ++i) {
				if(i) out += ',';
				dump(j.arr[i], out);
			}
			out += ']';
			break;
		case Json::Object:
			out += '{';
			for(size_t i = 0; i < j.obj.size(); ++i) {
				if(i) out += ',';
				dump(Json::string(j.obj[i].first), out);
				out += ':';
				dump(j.obj[i].second, out);
			}
			out += '}';
			break;
	}
}

std::vector<char> decode_base64(const std::string& s) {
	auto val = [](char c) -> int {
		if(c >= 'A' && c <= 'Z') return c - 'A';
		if(c >= 'a' && c <= 'z') return c - 'a' + 26;
		if(c >= '0' && c <= '9') return c - '0' + 52;
		if(c == '+' || c == '-') return 62;
		if(c == '/' || c == '_') return 63;
		return -1;
	};
	std::vector<char> out;
	uint32_t acc = 0;
	int bits = 0;
	for(char c : s) {
		if(c == '=') break;
		const int v = val(c);
		if(v < 0) {
			if(c == '\n' || c == '\r' || c == ' ') continue;
			bad("invalid base64 data");
		}
		acc = (acc << 6) | uint32_t(v);
		bits += 6;
		if(bits >= 8) {
			bits -= 8;
			out.push_back(char((acc >> bits) & 0xFF));
		}
	}
	return out;
}

std::vector<char> read_file(const fs::path& p) {
	std::ifstream f(p, std::ios::binary | std::ios::ate);
	if(!f) throw Failure(DDGI_ERROR_INVALID_ARGUMENT, "cannot open '" + p.string() + "'");
	const std::streamsize n = f.tellg();
	f.seekg(0);
	std::vector<char> buf(size_t(std::max<std::streamsize>(n, 0)));
	if(n > 0 && !f.read(buf.data(), n)) throw Failure(DDGI_ERROR_INVALID_ARGUMENT, "cannot read '" + p.string() + "'");
	return buf;
}

// ============================================================================================
// Scene model
using Mat4 = std::array<float, 16>;  // column-major (glm::mat4): element (row r, column c) at [c * 4 + r]

Mat4 identity() {
	Mat4 m{};
	m[0] = m[5] = m[10] = m[15] = 1.0f;
	return m;
}
// glm's mat4 product: column c of A * B = sum_k column k of A * B[c][k], summed in k order.
Mat4 mul(const Mat4& A, const Mat4& B) {
	Mat4 R{};
	for(int c = 0; c < 4; ++c)
		for(int r = 0; r < 4; ++r) {
			float s = A[0 * 4 + r] * B[c * 4 + 0];
			s = s + A[1 * 4 + r] * B[c * 4 + 1];
			s = s + A[2 * 4 + r] * B[c * 4 + 2];
			s = s + A[3 * 4 + r] * B[c * 4 + 3];
			R[c * 4 + r] = s;
		}
	return R;
}
// M * (p, 1), first three components (glm's mat4 * vec4 in column order).
void apply(const Mat4& M, const float p[3], float out[3]) {
	for(int r = 0; r < 3; ++r) {
		float s = M[0 * 4 + r] * p[0];
		s = s + M[1 * 4 + r] * p[1];
		s = s + M[2 * 4 + r] * p[2];
		s = s + M[3 * 4 + r] * 1.0f;
		out[r] = s;
	}
}

constexpr uint32_t kNoTexture = 0xFFFFFFFFu;
constexpr int kFormatSrgb = 43;   // VK_FORMAT_R8G8B8A8_SRGB
constexpr int kFormatUnorm = 37;  // VK_FORMAT_R8G8B8A8_UNORM

struct SMesh {
	std::string name;
	std::vector<ddgi_vertex> vertices;
	std::vector<uint32_t> indices;
	uint32_t material = 0;  // default material (src/vulkan/Mesh.hpp:33)
	float bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};
	bool skinned = false;
	void compute_bounds() {
		for(int a = 0; a < 3; ++a) {
			bmin[a] = std::numeric_limits<float>::max();
			bmax[a] = -std::numeric_limits<float>::max();
		}
		for(const auto& v : vertices)
			for(int a = 0; a < 3; ++a) {
				bmin[a] = std::min(bmin[a], v.pos[a]);
				bmax[a] = std::max(bmax[a], v.pos[a]);
			}
	}
	bool bounds_valid() const { return bmin[0] <= bmax[0] && bmin[1] <= bmax[1] && bmin[2] <= bmax[2]; }
};

struct SNode {
	std::string name = "Unamed Node";
	Mat4 local = identity();
	int parent = -1;
	std::vector<int> children;
	int mesh = -1;  // rendered mesh (MeshRendererComponent), -1 none
	uint32_t material = 0xFFFFFFFFu;
};

struct SMaterial {
	std::string name = "NoName";
	ddgi_material p{1.0f, 1.0f, {1, 1, 1}, {0, 0, 0}, kNoTexture, kNoTexture, kNoTexture, kNoTexture};
};

struct STexture {
	std::string source;
	int format = kFormatSrgb;
	Json sampler = Json::object();
};

struct Scene {
	std::vector<SMesh> meshes;
	std::vector<SNode> nodes;
	std::vector<SMaterial> materials;
	std::vector<STexture> textures;
	int root = -1;
	// flattened tables (valid after flatten())
	std::vector<ddgi_vertex> t_vertices;
	std::vector<uint32_t> t_indices;
	std::vector<ddgi_offset_entry> t_offsets;
	std::vector<ddgi_instance> t_instances;
	std::vector<ddgi_material> t_materials;
	float bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};

	int add_node(const std::string& name) {
		nodes.emplace_back();
		nodes.back().name = name;
		return int(nodes.size()) - 1;
	}
	void add_child(int parent, int child) {  // Scene::addChild (src/Scene.cpp:1029-1048): appended last
		if(parent == child) bad("a node cannot be its own child");
		if(nodes[child].parent != -1) bad("node '" + nodes[child].name + "' has two parents");
		nodes[parent].children.push_back(child);
		nodes[child].parent = parent;
	}
	void flatten();
};

void Scene::flatten() {
	// offset table: meshes with indices, packed (Renderer::updateMeshOffsetTable)
	t_vertices.clear();
	t_indices.clear();
	t_offsets.clear();
	std::vector<int32_t> entry(meshes.size(), -1);
	for(size_t m = 0; m < meshes.size(); ++m) {
		const SMesh& me = meshes[m];
		if(me.indices.empty()) continue;
		for(uint32_t i : me.indices)
			if(i >= me.vertices.size()) bad("mesh '" + me.name + "' has an index outside its vertices");
		if(me.indices.size() % 3) bad("mesh '" + me.name + "' index count is not a multiple of 3");
		entry[m] = int32_t(t_offsets.size());
		t_offsets.push_back({me.material, uint32_t(t_vertices.size()), uint32_t(t_indices.size())});
		t_vertices.insert(t_vertices.end(), me.vertices.begin(), me.vertices.end());
		t_indices.insert(t_indices.end(), me.indices.begin(), me.indices.end());
	}
	// materials; a mesh whose default material does not exist gets a default material appended
	t_materials.clear();
	for(const auto& m : materials) t_materials.push_back(m.p);
	bool need_default = false;
	for(const auto& oe : t_offsets) need_default = need_default || oe.material_index >= t_materials.size();
	if(need_default) {
		const uint32_t d = uint32_t(t_materials.size());
		t_materials.push_back(SMaterial{}.p);
		for(auto& oe : t_offsets)
			if(oe.material_index >= d) oe.material_index = d;
	}
	// global transforms below the root (Scene::update, src/Scene.cpp:936-957); nodes outside the
	// root's tree keep the identity
	std::vector<Mat4> global(nodes.size(), identity());
	std::vector<Mat4> visit(nodes.size(), identity());  // forEachNode: the root's transform included
	if(root >= 0) {
		std::vector<int> stack{root};
		visit[root] = nodes[root].local;
		while(!stack.empty()) {
			const int n = stack.back();
			stack.pop_back();
			for(int c : nodes[n].children) {
				global[c] = mul(global[n], nodes[c].local);
				visit[c] = mul(visit[n], nodes[c].local);
				stack.push_back(c);
			}
		}
	}
	// instances: rendering nodes ordered by (material, mesh), ties in node order
	std::vector<int> rend;
	for(size_t n = 0; n < nodes.size(); ++n) {
		const int m = nodes[n].mesh;
		if(m < 0) continue;
		if(size_t(m) >= meshes.size()) bad("node '" + nodes[n].name + "' renders a missing mesh");
		if(entry[m] < 0 || meshes[m].skinned) continue;
		rend.push_back(int(n));
	}
	std::stable_sort(rend.begin(), rend.end(), [&](int a, int b) {
		if(nodes[a].material != nodes[b].material) return nodes[a].material < nodes[b].material;
		return nodes[a].mesh < nodes[b].mesh;
	});
	t_instances.clear();
	for(int n : rend) {
		ddgi_instance in{};
		const Mat4& G = global[n];
		for(int r = 0; r < 3; ++r)
			for(int c = 0; c < 4; ++c) in.transform[r * 4 + c] = G[c * 4 + r];
		const SMesh& me = meshes[nodes[n].mesh];
		in.custom_index = uint32_t(entry[nodes[n].mesh]);
		in.mask = DDGI_MASK_STATIC;
		in.flags = 0;
		in.triangle_count = uint32_t(me.indices.size() / 3);
		t_instances.push_back(in);
	}
	// bounds (Scene::computeBounds)
	bool init = false;
	for(size_t n = 0; n < nodes.size(); ++n) {
		if(nodes[n].mesh < 0) continue;
		SMesh& me = meshes[nodes[n].mesh];
		if(!me.bounds_valid()) continue;
		float a[3], b[3];
		apply(visit[n], me.bmin, a);
		apply(visit[n], me.bmax, b);
		for(int k = 0; k < 3; ++k) {
			const float lo = std::min(a[k], b[k]), hi = std::max(a[k], b[k]);
			bmin[k] = init ? std::min(bmin[k], lo) : lo;
			bmax[k] = init ? std::max(bmax[k], hi) : hi;
		}
		init = true;
	}
	if(!init)
		for(int k = 0; k < 3; ++k) bmin[k] = bmax[k] = 0.0f;
}

// ---- materials (parseMaterial, src/vulkan/Material.cpp:7-28; Scene::loadMaterial :666-682) ----
float jf(const Json& j) { return float(j.as_number()); }

void add_material(Scene& sc, const Json& mat, uint32_t tex_offset) {
	SMaterial m;
	m.name = mat.get_string("name", "NoName");
	if(const Json* pbr = mat.find("pbrMetallicRoughness")) {
		if(const Json* c = pbr->find("baseColorFactor"))  // a vec4 read into the vec3 field
			for(int k = 0; k < 3; ++k) m.p.base_color[k] = jf(c->at(k));
		if(const Json* v = pbr->find("metallicFactor")) m.p.metallic = jf(*v);
		if(const Json* v = pbr->find("roughnessFactor")) m.p.roughness = jf(*v);
		if(const Json* t = pbr->find("baseColorTexture")) m.p.albedo_texture = uint32_t(tex_offset + t->at("index").as_int());
		if(const Json* t = pbr->find("metallicRoughnessTexture"))
			m.p.metallic_roughness_texture = uint32_t(tex_offset + t->at("index").as_int());
	}
	if(const Json* e = mat.find("emissiveFactor"))
		for(int k = 0; k < 3; ++k) m.p.emissive[k] = jf(e->at(k));
	if(const Json* t = mat.find("emissiveTexture")) m.p.emissive_texture = uint32_t(tex_offset + t->at("index").as_int());
	if(const Json* t = mat.find("normalTexture")) m.p.normal_texture = uint32_t(tex_offset + t->at("index").as_int());
	// the normal and metallic-roughness maps are linear data
	for(uint32_t t : {m.p.normal_texture, m.p.metallic_roughness_texture})
		if(t != kNoTexture && t < sc.textures.size()) sc.textures[t].format = kFormatUnorm;
	sc.materials.push_back(std::move(m));
}

// ---- textures (Scene::loadTextures, src/Scene.cpp:684-708): records only, never decoded ----
void add_textures(Scene& sc, const fs::path& file, const Json& doc) {
	for(const Json& t : doc.find("textures") ? doc.at("textures").items() : std::vector<Json>{}) {
		STexture tx;
		const Json& image = doc.at("images").at(size_t(t.at("source").as_int()));
		if(const Json* s = doc.find("samplers")) {
			const long long si = t.get_int("sampler", 0);
			if(si >= 0 && size_t(si) < s->items().size()) tx.sampler = s->at(size_t(si));
		}
		if(image.has("uri")) tx.source = (file.parent_path() / image.at("uri").as_string()).lexically_normal().string();
		else tx.source = "data/blank.png";  // embedded images are replaced by a blank image
		sc.textures.push_back(std::move(tx));
	}
}

// ============================================================================================
// glTF 2.0 (src/Scene.cpp:115-533)
constexpr int kFloat = 5126, kUShort = 5123, kUInt = 5125;

struct Accessor {
	const char* data = nullptr;  // first element
	size_t stride = 0, count = 0, avail = 0;  // avail: bytes from `data` to the end of the buffer
};

Accessor accessor(const Json& doc, const std::vector<std::vector<char>>& buffers, long long index, int want_component,
				  const char* want_type, size_t default_stride) {
	const Json& a = doc.at("accessors").at(size_t(index));
	if(want_component && a.at("componentType").as_int() != want_component)
		bad("glTF accessor " + std::to_string(index) + ": unsupported component type");
	if(want_type && a.at("type").as_string() != want_type)
		bad("glTF accessor " + std::to_string(index) + ": expected type " + want_type);
	if(!a.has("bufferView")) bad("glTF accessor without a buffer view (sparse accessors are not supported)");
	const Json& bv = doc.at("bufferViews").at(size_t(a.at("bufferView").as_int()));
	const long long b = bv.at("buffer").as_int();
	if(b < 0 || size_t(b) >= buffers.size()) bad("glTF buffer view refers to a missing buffer");
	const std::vector<char>& buf = buffers[size_t(b)];
	Accessor r;
	const size_t off = size_t(a.get_int("byteOffset", 0) + bv.get_int("byteOffset", 0));
	r.stride = size_t(bv.get_int("byteStride", (long long)default_stride));
	r.count = size_t(a.at("count").as_int());
	if(off > buf.size()) bad("glTF accessor starts past its buffer");
	r.data = buf.data() + off;
	r.avail = buf.size() - off;
	if(r.count && (r.count - 1) * r.stride + default_stride > r.avail) bad("glTF accessor runs past its buffer");
	return r;
}

template <class T>
T load(const char* p) {
	T v;
	std::memcpy(&v, p, sizeof(T));
	return v;
}

Mat4 node_transform(const Json& n) {
	if(const Json* m = n.find("matrix")) {
		Mat4 M;
		for(int i = 0; i < 16; ++i) M[i] = jf(m->at(size_t(i)));
		return M;
	}
	float s[3] = {1, 1, 1}, t[3] = {0, 0, 0}, q[4] = {0, 0, 0, 1};  // q = (x, y, z, w)
	if(const Json* v = n.find("scale"))
		for(int k = 0; k < 3; ++k) s[k] = jf(v->at(size_t(k)));
	if(const Json* v = n.find("translation"))
		for(int k = 0; k < 3; ++k) t[k] = jf(v->at(size_t(k)));
	if(const Json* v = n.find("rotation"))
		for(int k = 0; k < 4; ++k) q[k] = jf(v->at(size_t(k)));
	// rotation matrix of the unit quaternion (glm::toMat4), then translation * rotation * scale
	const float x = q[0], y = q[1], z = q[2], w = q[3];
	const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, xz = x * z, yz = y * z, wx = w * x, wy = w * y, wz = w * z;
	Mat4 R = identity();
	R[0] = 1.0f - 2.0f * (yy + zz);
	R[1] = 2.0f * (xy + wz);
	R[2] = 2.0f * (xz - wy);
	R[4] = 2.0f * (xy - wz);
	R[5] = 1.0f - 2.0f * (xx + zz);
	R[6] = 2.0f * (yz + wx);
	R[8] = 2.0f * (xz + wy);
	R[9] = 2.0f * (yz - wx);
	R[10] = 1.0f - 2.0f * (xx + yy);
	Mat4 T = identity(), S = identity();
	T[12] = t[0];
	T[13] = t[1];
	T[14] = t[2];
	S[0] = s[0];
	S[5] = s[1];
	S[10] = s[2];
	return mul(mul(T, R), S);
}

void load_gltf_document(Scene& sc, const fs::path& file, const Json& doc, std::vector<std::vector<char>>& buffers) {
	const uint32_t tex_offset = uint32_t(sc.textures.size());
	add_textures(sc, file, doc);
	const uint32_t mat_offset = uint32_t(sc.materials.size());
	for(const Json& m : doc.find("materials") ? doc.at("materials").items() : std::vector<Json>{}) add_material(sc, m, tex_offset);
	// meshes: one per primitive
	std::vector<std::vector<int>> prim_meshes;
	for(const Json& m : doc.at("meshes").items()) {
		const std::string base = m.get_string("name", "UnamedMesh");
		prim_meshes.emplace_back();
		for(const Json& p : m.at("primitives").items()) {
			if(p.get_int("mode", 4) != 4) bad("glTF primitive mode other than triangles");
			prim_meshes.back().push_back(int(sc.meshes.size()));
			sc.meshes.emplace_back();
			SMesh& me = sc.meshes.back();
			me.name = base + "_" + p.get_string("name", "Unamed");
			if(p.has("material")) me.material = uint32_t(mat_offset + p.at("material").as_int());
			const Json& attr = p.at("attributes");
			const Accessor pos = accessor(doc, buffers, attr.at("POSITION").as_int(), kFloat, "VEC3", 12);
			Accessor nrm, tan, uv;
			if(attr.has("NORMAL")) nrm = accessor(doc, buffers, attr.at("NORMAL").as_int(), kFloat, "VEC3", 12);
			if(attr.has("TANGENT")) tan = accessor(doc, buffers, attr.at("TANGENT").as_int(), kFloat, "VEC4", 16);
			if(attr.has("TEXCOORD_0")) uv = accessor(doc, buffers, attr.at("TEXCOORD_0").as_int(), kFloat, "VEC2", 8);
			me.skinned = attr.has("WEIGHTS_0");
			for(const Accessor* a : {&nrm, &tan, &uv})
				if(a->data && a->count < pos.count) bad("glTF attribute shorter than POSITION");
			me.vertices.resize(pos.count);
			for(size_t i = 0; i < pos.count; ++i) {
				ddgi_vertex v{};
				std::memcpy(v.pos, pos.data + i * pos.stride, 12);
				v.color[0] = v.color[1] = v.color[2] = 1.0f;
				if(nrm.data) std::memcpy(v.normal, nrm.data + i * nrm.stride, 12);
				if(tan.data) std::memcpy(v.tangent, tan.data + i * tan.stride, 16);
				else v.tangent[0] = v.tangent[1] = v.tangent[2] = v.tangent[3] = 1.0f;
				if(uv.data) std::memcpy(v.tex_coord, uv.data + i * uv.stride, 8);
				me.vertices[i] = v;
			}
			const Json& pa = doc.at("accessors").at(size_t(attr.at("POSITION").as_int()));
			if(pa.has("min") && pa.has("max")) {
				for(int k = 0; k < 3; ++k) {
					me.bmin[k] = jf(pa.at("min").at(size_t(k)));
					me.bmax[k] = jf(pa.at("max").at(size_t(k)));
				}
			} else {
				me.compute_bounds();
			}
			if(!me.bounds_valid()) me.compute_bounds();
			if(p.has("indices")) {
				const long long ia = p.at("indices").as_int();
				const long long ct = doc.at("accessors").at(size_t(ia)).at("componentType").as_int();
				if(ct != kUShort && ct != kUInt) bad("glTF indices must be unsigned short or unsigned int");
				const Accessor ix = accessor(doc, buffers, ia, 0, "SCALAR", ct == kUShort ? 2 : 4);
				me.indices.resize(ix.count);
				for(size_t i = 0; i < ix.count; ++i)
					me.indices[i] = ct == kUShort ? load<uint16_t>(ix.data + i * ix.stride) : load<uint32_t>(ix.data + i * ix.stride);
			} else {  // non-indexed: consecutive triplets
				for(size_t i = 0; i < me.vertices.size(); i += 3) {
					me.indices.push_back(uint32_t(i));
					me.indices.push_back(uint32_t(i + 1));
					me.indices.push_back(uint32_t(i + 2));
				}
			}
		}
	}
	// nodes
	const std::vector<Json>& gnodes = doc.find("nodes") ? doc.at("nodes").items() : std::vector<Json>{};
	std::vector<int> ent(gnodes.size());
	std::vector<std::vector<int>> kids;
	for(size_t i = 0; i < gnodes.size(); ++i) {
		const Json& n = gnodes[i];
		ent[i] = sc.add_node(n.get_string("name", "Unamed Node"));
		sc.nodes[ent[i]].local = node_transform(n);
		kids.emplace_back();
		for(const Json& c : n.find("children") ? n.at("children").items() : std::vector<Json>{}) kids.back().push_back(int(c.as_int()));
		const long long mi = n.get_int("mesh", -1);
		if(mi < 0) continue;
		if(size_t(mi) >= prim_meshes.size()) bad("glTF node refers to a missing mesh");
		const std::vector<int>& pm = prim_meshes[size_t(mi)];
		if(pm.size() == 1) {
			sc.nodes[ent[i]].mesh = pm[0];
			sc.nodes[ent[i]].material = sc.meshes[pm[0]].material;
		} else {
			for(int m : pm) {  // one child node per primitive
				const int sub = sc.add_node("Unamed Node");
				sc.add_child(ent[i], sub);
				sc.nodes[sub].mesh = m;
				sc.nodes[sub].material = sc.meshes[m].material;
			}
		}
	}
	for(const Json& s : doc.find("scenes") ? doc.at("scenes").items() : std::vector<Json>{}) {
		const int e = sc.add_node(s.get_string("name", "Unamed Scene"));
		sc.add_child(sc.root, e);
		for(const Json& n : s.find("nodes") ? s.at("nodes").items() : std::vector<Json>{}) {
			const long long k = n.as_int();
			if(k < 0 || size_t(k) >= ent.size()) bad("glTF scene refers to a missing node");
			sc.add_child(e, ent[size_t(k)]);
		}
	}
	for(size_t i = 0; i < kids.size(); ++i)
		for(int c : kids[i]) {
			if(c < 0 || size_t(c) >= ent.size()) bad("glTF node refers to a missing child");
			sc.add_child(ent[i], ent[size_t(c)]);
		}
}

struct ChunkHeader {
	uint32_t length, type;
};
constexpr uint32_t kMagicGltf = 0x46546C67u, kMagicScene = 0x4E454353u;
constexpr uint32_t kChunkJson = 0x4E4F534Au, kChunkBin = 0x004E4942u;

// Header + chunks of a GLB-style container: the JSON chunk's text, then the binary chunks.
std::pair<std::string, std::vector<std::vector<char>>> read_container(const std::vector<char>& f, uint32_t magic) {
	if(f.size() < 12) bad("file too short for its header");
	const uint32_t m = load<uint32_t>(f.data()), len = load<uint32_t>(f.data() + 8);
	if(m != magic) bad("wrong file magic");
	if(len > f.size()) bad("header length exceeds the file");
	size_t off = 12;
	if(off + 8 > len) bad("missing JSON chunk");
	const ChunkHeader jc = load<ChunkHeader>(f.data() + off);
	if(jc.type != kChunkJson) bad("first chunk is not JSON");
	if(off + 8 + size_t(jc.length) > len) bad("JSON chunk runs past the file");
	std::string json(f.data() + off + 8, jc.length);
	off += 8 + jc.length;
	std::vector<std::vector<char>> bins;
	while(off < len) {
		if(off + 8 > len) bad("truncated chunk header");
		const ChunkHeader c = load<ChunkHeader>(f.data() + off);
		if(c.type != kChunkBin) bad("chunk after the JSON chunk is not BIN");
		if(off + 8 + size_t(c.length) > len) bad("BIN chunk runs past the file");
		bins.emplace_back(f.data() + off + 8, f.data() + off + 8 + c.length);
		off += 8 + c.length;
	}
	return {std::move(json), std::move(bins)};
}

void load_gltf(Scene& sc, const fs::path& file) {
	Json doc;
	std::vector<std::vector<char>> buffers;
	const std::string ext = file.extension().string();
	if(ext == ".gltf") {
		const std::vector<char> text = read_file(file);
		doc = JsonParser(text.data(), text.size()).document();
		for(const Json& b : doc.find("buffers") ? doc.at("buffers").items() : std::vector<Json>{}) {
			const size_t len = size_t(b.at("byteLength").as_int());
			const std::string uri = b.at("uri").as_string();
			const std::string prefix = "data:application/octet-stream;base64,";
			if(uri.rfind("data:", 0) == 0) {
				if(uri.rfind(prefix, 0) != 0) bad("unsupported data URI in a glTF buffer");
				buffers.push_back(decode_base64(uri.substr(prefix.size())));
			} else {
				std::vector<char> data = read_file(file.parent_path() / uri);
				if(data.size() < len) bad("glTF buffer file shorter than its byteLength");
				data.resize(len);
				buffers.push_back(std::move(data));
			}
		}
	} else {
		auto c = read_container(read_file(file), kMagicGltf);
		doc = JsonParser(c.first.data(), c.first.size()).document();
		buffers = std::move(c.second);
	}
	load_gltf_document(sc, file, doc, buffers);
}

// ============================================================================================
// Wavefront OBJ (src/Scene.cpp:534-646), with the reference's behaviour kept: `o` starts a new mesh
// node but keeps filling the current mesh until a `g` starts a new sub-mesh; face indices are
// absolute in the file and made relative to the current sub-mesh; only triangles; `vt`/`vn`
// indices in a face set that vertex's attributes; every finished sub-mesh gets its vertex normals
// recomputed (Mesh::computeVertexNormals, src/vulkan/Mesh.cpp:69-94: the average of the adjacent
// triangles' normalize(cross(p0 - p1, p2 - p0)), which for counter-clockwise faces points inward,
// and tangents normalize(p0 - p1)), overwriting the file's normals.  Unlike the reference, an index
// outside the data read so far is an error instead of an out-of-bounds write.
void compute_vertex_normals(SMesh& m) {
	for(auto& v : m.vertices) {
		for(int k = 0; k < 3; ++k) v.normal[k] = 0.0f;
		for(int k = 0; k < 4; ++k) v.tangent[k] = 0.0f;
	}
	for(size_t i = 0; i + 2 < m.indices.size(); i += 3) {
		const float* p0 = m.vertices[m.indices[i]].pos;
		const float* p1 = m.vertices[m.indices[i + 1]].pos;
		const float* p2 = m.vertices[m.indices[i + 2]].pos;
		const float a[3] = {p0[0] - p1[0], p0[1] - p1[1], p0[2] - p1[2]};
		const float b[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
		float n[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
		const float nl = 1.0f / std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
		const float al = 1.0f / std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
		for(int j = 0; j < 3; ++j) {
			ddgi_vertex& v = m.vertices[m.indices[i + j]];
			for(int k = 0; k < 3; ++k) {
				v.normal[k] += n[k] * nl;
				v.tangent[k] += a[k] * al;
			}
			v.tangent[3] += 1.0f;
		}
	}
	for(auto& v : m.vertices) {  // a vertex no triangle uses divides by zero, as in the reference
		for(int k = 0; k < 3; ++k) {
			v.normal[k] = v.normal[k] / v.tangent[3];
			v.tangent[k] = v.tangent[k] / v.tangent[3];
		}
		v.tangent[3] = 1.0f;
	}
}

void load_obj(Scene& sc, const fs::path& file) {
	std::ifstream f(file);
	if(!f) throw Failure(DDGI_ERROR_INVALID_ARGUMENT, "cannot open '" + file.string() + "'");
	const int root = sc.add_node("Unamed Node");
	sc.add_child(sc.root, root);
	int mesh_node = -1, cur = -1;  // current mesh node, current mesh
	uint32_t vertex_offset = 0;
	std::vector<std::array<float, 2>> uvs;
	std::vector<std::array<float, 3>> normals;
	auto next_mesh = [&] {
		mesh_node = sc.add_node("Unamed Node");
		sc.add_child(root, mesh_node);
	};
	auto next_submesh = [&] {
		if(cur >= 0) vertex_offset += uint32_t(sc.meshes[cur].vertices.size());
		sc.meshes.emplace_back();
		sc.meshes.back().name = "AnonMesh";
		cur = int(sc.meshes.size()) - 1;
		const int sub = sc.add_node("Unamed Node");
		sc.add_child(mesh_node, sub);
		sc.nodes[sub].mesh = cur;
		sc.nodes[sub].material = 0xFFFFFFFFu;  // a renderer without a material of its own
	};
	next_mesh();
	next_submesh();
	ddgi_vertex v{};
	v.color[0] = v.color[1] = v.color[2] = 1.0f;
	std::string line;
	size_t lineno = 0;
	auto where = [&] { return file.filename().string() + ":" + std::to_string(lineno); };
	while(std::getline(f, line)) {
		++lineno;
		if(!line.empty() && line.back() == '\r') line.pop_back();
		if(line.empty()) continue;
		const size_t first = line.find_first_not_of(" \t");
		if(first == std::string::npos || line[first] == '#') continue;
		SMesh& m = sc.meshes[cur];
		if(line[0] == 'v') {
			if(line.size() > 1 && line[1] == 't') {
				char* p = line.data() + std::min<size_t>(3, line.size());
				std::array<float, 2> uv{};
				for(int k = 0; k < 2; ++k) uv[k] = std::strtof(p, &p);
				uvs.push_back(uv);
			} else if(line.size() > 1 && line[1] == 'n') {
				char* p = line.data() + std::min<size_t>(3, line.size());
				std::array<float, 3> n{};
				for(int k = 0; k < 3; ++k) n[k] = std::strtof(p, &p);
				const float l = 1.0f / std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
				for(int k = 0; k < 3; ++k) n[k] *= l;
				normals.push_back(n);
			} else {
				char* p = line.data() + std::min<size_t>(2, line.size());
				for(int k = 0; k < 3; ++k) v.pos[k] = std::strtof(p, &p);
				m.vertices.push_back(v);
			}
		} else if(line[0] == 'f') {
			char* p = line.data() + std::min<size_t>(2, line.size());
			for(int k = 0; k < 3; ++k) {  // triangles only
				const long vi = std::strtol(p, &p, 10);
				if(vi <= 0) bad(where() + ": OBJ face index must be a positive absolute index");
				const long rel = vi - 1 - long(vertex_offset);
				if(rel < 0 || size_t(rel) >= m.vertices.size()) bad(where() + ": OBJ face refers to a vertex outside its group");
				m.indices.push_back(uint32_t(rel));
				if(*p == '/') {
					++p;
					const long ti = std::strtol(p, &p, 10);
					if(ti > 0) {
						if(size_t(ti) > uvs.size()) bad(where() + ": OBJ face refers to a missing texture coordinate");
						m.vertices[rel].tex_coord[0] = uvs[ti - 1][0];
						m.vertices[rel].tex_coord[1] = uvs[ti - 1][1];
					}
					if(*p == '/') {
						++p;
						const long ni = std::strtol(p, &p, 10);
						if(ni > 0) {
							if(size_t(ni) > normals.size()) bad(where() + ": OBJ face refers to a missing normal");
							for(int q = 0; q < 3; ++q) m.vertices[rel].normal[q] = normals[ni - 1][q];
						}
					}
				}
			}
		} else if(line[0] == 'o') {
			if(!m.vertices.empty()) next_mesh();
		} else if(line[0] == 'g') {
			if(!m.vertices.empty()) {
				compute_vertex_normals(m);
				m.compute_bounds();
				next_submesh();
			}
		}  // other statements (usemtl, mtllib, s, ...) are ignored, as the reference does with a warning
	}
	compute_vertex_normals(sc.meshes[cur]);
	sc.meshes[cur].compute_bounds();
}

// ============================================================================================
// .scene (src/Scene.cpp:710-934)
void load_scene_file(Scene& sc, const fs::path& file) {
	auto c = read_container(read_file(file), kMagicScene);
	const Json doc = JsonParser(c.first.data(), c.first.size()).document();
	const auto& bins = c.second;
	sc = Scene{};
	const std::vector<Json>& ents = doc.at("entities").items();
	for(const Json& e : ents) {
		const int n = sc.add_node(e.at("name").as_string());
		const Json& t = e.at("transform");
		for(int i = 0; i < 16; ++i) sc.nodes[n].local[i] = jf(t.at(size_t(i)));
		if(const Json* r = e.find("meshRenderer")) {
			sc.nodes[n].mesh = int(r->at("meshIndex").as_int());
			sc.nodes[n].material = uint32_t(r->at("materialIndex").as_int());
		}
	}
	for(size_t i = 0; i < ents.size(); ++i)
		for(const Json& ch : ents[i].at("children").items()) {
			const long long k = ch.as_int();
			if(k < 0 || size_t(k) >= ents.size()) bad(".scene entity refers to a missing child");
			sc.add_child(int(i), int(k));
		}
	for(const Json& t : doc.find("textures") ? doc.at("textures").items() : std::vector<Json>{}) {
		STexture tx;
		tx.source = (file.parent_path() / t.at("source").as_string()).lexically_normal().string();
		tx.format = int(t.at("format").as_int());
		tx.sampler = t.at("sampler");
		sc.textures.push_back(std::move(tx));
	}
	for(const Json& m : doc.at("materials").items()) add_material(sc, m, 0);
	for(const Json& m : doc.at("meshes").items()) {
		sc.meshes.emplace_back();
		SMesh& me = sc.meshes.back();
		me.name = m.at("name").as_string();
		me.material = uint32_t(m.get_int("material", 0));
		const long long va = m.at("vertexArray").as_int() - 1, ia = m.at("indexArray").as_int() - 1;  // chunk 0 is the JSON
		if(va < 0 || size_t(va) >= bins.size() || ia < 0 || size_t(ia) >= bins.size()) bad(".scene mesh refers to a missing chunk");
		const auto& vb = bins[size_t(va)];
		const auto& ib = bins[size_t(ia)];
		if(vb.size() % sizeof(ddgi_vertex) || ib.size() % 4) bad(".scene mesh chunk size is not a whole number of records");
		me.vertices.resize(vb.size() / sizeof(ddgi_vertex));
		if(!vb.empty()) std::memcpy(me.vertices.data(), vb.data(), vb.size());
		me.indices.resize(ib.size() / 4);
		if(!ib.empty()) std::memcpy(me.indices.data(), ib.data(), ib.size());
		me.compute_bounds();
	}
	for(size_t n = 0; n < sc.nodes.size(); ++n)
		if(sc.nodes[n].parent == -1) {  // the first entity without a parent is the root
			sc.root = int(n);
			break;
		}
}

Json vec_json(const float* v, int n) {
	Json a = Json::array();
	for(int i = 0; i < n; ++i) a.arr.push_back(Json::number(v[i]));
	return a;
}

Json material_json(const SMaterial& m) {  // toJSON(Material), src/vulkan/Material.cpp:30-53
	auto tex = [](uint32_t id) {
		Json t = Json::object();
		t.set("index", Json::number(double(id)));
		return t;
	};
	Json o = Json::object();
	o.set("name", Json::string(m.name));
	Json pbr = Json::object();
	const float bc[4] = {m.p.base_color[0], m.p.base_color[1], m.p.base_color[2], 1.0f};
	pbr.set("baseColorFactor", vec_json(bc, 4));
	pbr.set("metallicFactor", Json::number(m.p.metallic));
	pbr.set("roughnessFactor", Json::number(m.p.roughness));
	pbr.set("baseColorTexture", tex(m.p.albedo_texture));
	pbr.set("metallicRoughnessTexture", tex(m.p.metallic_roughness_texture));
	o.set("pbrMetallicRoughness", std::move(pbr));
	o.set("normalTexture", tex(m.p.normal_texture));
	o.set("emissiveFactor", vec_json(m.p.emissive, 3));
	o.set("emissiveTexture", tex(m.p.emissive_texture));
	return o;
}

void save_scene_file(const Scene& sc, const fs::path& file) {
	Json root = Json::object();
	Json mats = Json::array();
	for(const auto& m : sc.materials) mats.arr.push_back(material_json(m));
	root.set("materials", std::move(mats));
	Json ents = Json::array();
	for(const SNode& n : sc.nodes) {
		Json e = Json::object();
		e.set("name", Json::string(n.name));
		e.set("transform", vec_json(n.local.data(), 16));
		e.set("parent", Json::number(-1));
		Json ch = Json::array();
		for(int c : n.children) ch.arr.push_back(Json::number(c));
		e.set("children", std::move(ch));
		if(n.mesh >= 0) {
			Json r = Json::object();
			r.set("meshIndex", Json::number(n.mesh));
			r.set("materialIndex", Json::number(double(int32_t(n.material))));
			e.set("meshRenderer", std::move(r));
		}
		ents.arr.push_back(std::move(e));
	}
	root.set("entities", std::move(ents));
	Json meshes = Json::array();
	for(size_t m = 0; m < sc.meshes.size(); ++m) {
		Json o = Json::object();
		o.set("name", Json::string(sc.meshes[m].name));
		o.set("material", Json::number(sc.meshes[m].material));
		o.set("vertexArray", Json::number(double(1 + 2 * m)));
		o.set("indexArray", Json::number(double(2 + 2 * m)));
		meshes.arr.push_back(std::move(o));
	}
	root.set("meshes", std::move(meshes));
	Json texs = Json::array();
	for(const auto& t : sc.textures) {
		Json o = Json::object();
		o.set("source", Json::string(fs::path(t.source).lexically_relative(file.parent_path()).string()));
		o.set("format", Json::number(t.format));
		o.set("sampler", t.sampler);
		texs.arr.push_back(std::move(o));
	}
	root.set("textures", std::move(texs));
	std::string text;
	dump(root, text);
	std::vector<std::pair<uint32_t, std::pair<const char*, size_t>>> chunks;
	chunks.push_back({kChunkJson, {text.data(), text.size()}});
	for(const auto& m : sc.meshes) {
		chunks.push_back({kChunkBin, {reinterpret_cast<const char*>(m.vertices.data()), m.vertices.size() * sizeof(ddgi_vertex)}});
		chunks.push_back({kChunkBin, {reinterpret_cast<const char*>(m.indices.data()), m.indices.size() * 4}});
	}
	uint64_t total = 12;
	for(const auto& c : chunks) total += 8 + c.second.second;
	if(total > 0xFFFFFFFFull) throw Failure(DDGI_ERROR_UNSUPPORTED, ".scene files are limited to 4 GiB");
	std::ofstream f(file, std::ios::binary);
	if(!f) throw Failure(DDGI_ERROR_INVALID_ARGUMENT, "cannot open '" + file.string() + "' for writing");
	const uint32_t hdr[3] = {kMagicScene, 0u, uint32_t(total)};
	f.write(reinterpret_cast<const char*>(hdr), sizeof hdr);
	for(const auto& c : chunks) {
		const uint32_t ch[2] = {uint32_t(c.second.second), c.first};
		f.write(reinterpret_cast<const char*>(ch), sizeof ch);
		if(c.second.second) f.write(c.second.first, std::streamsize(c.second.second));
	}
	if(!f) throw Failure(DDGI_ERROR_INVALID_ARGUMENT, "write to '" + file.string() + "' failed");
}

Scene* S(void* p) { return static_cast<Scene*>(p); }

template <class F>
ddgi_status guarded(F&& f) {
	try {
		f();
		return DDGI_OK;
	} catch(const Failure& e) {
		ddgi::host::record_error(e.what());
		return e.status;
	} catch(const std::bad_alloc&) {
		ddgi::host::record_error("out of host memory");
		return DDGI_ERROR_OUT_OF_MEMORY;
	} catch(const std::exception& e) {
		ddgi::host::record_error(e.what());
		return DDGI_ERROR_INVALID_ARGUMENT;
	}
}

}  // namespace

extern "C" {

ddgi_status ddgi_scene_load(const char* path, void** out_scene) {
	if(!path || !out_scene) {
		ddgi::host::record_error("null argument");
		return DDGI_ERROR_INVALID_ARGUMENT;
	}
	*out_scene = nullptr;
	Scene* sc = new(std::nothrow) Scene;
	if(!sc) return DDGI_ERROR_OUT_OF_MEMORY;
	const ddgi_status st = guarded([&] {
		const fs::path file(path);
		const std::string ext = file.extension().string();
		if(ext == ".scene") {
			load_scene_file(*sc, file);
		} else if(ext == ".gltf" || ext == ".glb") {
			sc->root = sc->add_node("Root");
			load_gltf(*sc, file);
		} else if(ext == ".obj") {
			sc->root = sc->add_node("Root");
			load_obj(*sc, file);
		} else {
			throw Failure(DDGI_ERROR_UNSUPPORTED, "unsupported scene file extension '" + ext + "'");
		}
		sc->flatten();
	});
	if(st != DDGI_OK) {
		delete sc;
		return st;
	}
	*out_scene = sc;
	return DDGI_OK;
}

ddgi_status ddgi_scene_load_material(void* scene, const char* path) {
	if(!scene || !path) {
		ddgi::host::record_error("null argument");
		return DDGI_ERROR_INVALID_ARGUMENT;
	}
	Scene* sc = S(scene);
	Scene backup = *sc;
	const ddgi_status st = guarded([&] {
		const fs::path file(path);
		if(file.extension() != ".mat") throw Failure(DDGI_ERROR_UNSUPPORTED, "material files end in .mat");
		const std::vector<char> text = read_file(file);
		const Json doc = JsonParser(text.data(), text.size()).document();
		const uint32_t tex_offset = uint32_t(sc->textures.size());
		add_textures(*sc, file, doc);
		for(const Json& m : doc.find("materials") ? doc.at("materials").items() : std::vector<Json>{}) add_material(*sc, m, tex_offset);
		sc->flatten();
	});
	if(st != DDGI_OK) *sc = std::move(backup);
	return st;
}

ddgi_status ddgi_scene_create(const ddgi_vertex* vertices, size_t vertex_count, const uint32_t* indices, size_t index_count,
							  const ddgi_offset_entry* offsets, size_t offset_count, const ddgi_instance* instances,
							  size_t instance_count, const ddgi_material* materials, size_t material_count, void** out_scene) {
	if(!out_scene || (vertex_count && !vertices) || (index_count && !indices) || (offset_count && !offsets) ||
	   (instance_count && !instances) || (material_count && !materials)) {
		ddgi::host::record_error("null argument");
		return DDGI_ERROR_INVALID_ARGUMENT;
	}
	*out_scene = nullptr;
	Scene* sc = new(std::nothrow) Scene;
	if(!sc) return DDGI_ERROR_OUT_OF_MEMORY;
	const ddgi_status st = guarded([&] {
		for(size_t m = 0; m < material_count; ++m) {
			SMaterial mt;
			mt.name = "Material" + std::to_string(m);
			mt.p = materials[m];
			sc->materials.push_back(mt);
		}
		// a mesh per offset entry: its vertices and indices run to the next larger offset
		std::vector<uint32_t> vstarts, istarts;
		for(size_t k = 0; k < offset_count; ++k) {
			vstarts.push_back(offsets[k].vertex_offset);
			istarts.push_back(offsets[k].index_offset);
		}
		std::sort(vstarts.begin(), vstarts.end());
		std::sort(istarts.begin(), istarts.end());
		auto next_after = [](const std::vector<uint32_t>& s, uint32_t v, size_t end) {
			auto it = std::upper_bound(s.begin(), s.end(), v);
			return it == s.end() ? end : size_t(*it);
		};
		for(size_t k = 0; k < offset_count; ++k) {
			const ddgi_offset_entry& oe = offsets[k];
			const size_t v0 = oe.vertex_offset, v1 = next_after(vstarts, oe.vertex_offset, vertex_count);
			const size_t i0 = oe.index_offset, i1 = next_after(istarts, oe.index_offset, index_count);
			if(v0 > vertex_count || v1 > vertex_count || i0 > index_count || i1 > index_count)
				bad("offset entry outside the vertex or index table");
			SMesh me;
			me.name = "Mesh" + std::to_string(k);
			me.material = oe.material_index;
			me.vertices.assign(vertices + v0, vertices + v1);
			me.indices.assign(indices + i0, indices + i1);
			me.compute_bounds();
			sc->meshes.push_back(std::move(me));
		}
		sc->root = sc->add_node("Root");
		for(size_t i = 0; i < instance_count; ++i) {
			const ddgi_instance& in = instances[i];
			if(in.custom_index >= offset_count) bad("instance custom_index outside the offset table");
			if(in.mask != DDGI_MASK_STATIC || in.flags != 0)
				throw Failure(DDGI_ERROR_UNSUPPORTED, "the .scene format stores static instances without facing flips only");
			if(size_t(in.triangle_count) * 3 != sc->meshes[in.custom_index].indices.size())
				bad("instance triangle_count differs from its mesh");
			const int n = sc->add_node("Instance" + std::to_string(i));
			Mat4 M = identity();
			for(int r = 0; r < 3; ++r)
				for(int c = 0; c < 4; ++c) M[c * 4 + r] = in.transform[r * 4 + c];
			sc->nodes[n].local = M;
			sc->nodes[n].mesh = int(in.custom_index);
			sc->nodes[n].material = offsets[in.custom_index].material_index;
			sc->add_child(sc->root, n);
		}
		sc->flatten();
	});
	if(st != DDGI_OK) {
		delete sc;
		return st;
	}
	*out_scene = sc;
	return DDGI_OK;
}

ddgi_status ddgi_scene_get_tables(void* scene, ddgi_scene_tables* out) {
	if(!scene || !out) {
		ddgi::host::record_error("null argument");
		return DDGI_ERROR_INVALID_ARGUMENT;
	}
	const Scene* sc = S(scene);
	out->vertices = sc->t_vertices.data();
	out->vertex_count = sc->t_vertices.size();
	out->indices = sc->t_indices.data();
	out->index_count = sc->t_indices.size();
	out->offsets = sc->t_offsets.data();
	out->offset_count = sc->t_offsets.size();
	out->instances = sc->t_instances.data();
	out->instance_count = sc->t_instances.size();
	out->materials = sc->t_materials.data();
	out->material_count = sc->t_materials.size();
	for(int k = 0; k < 3; ++k) {
		out->bounds_min[k] = sc->bmin[k];
		out->bounds_max[k] = sc->bmax[k];
	}
	out->mesh_count = sc->meshes.size();
	out->node_count = sc->nodes.size();
	out->texture_count = sc->textures.size();
	return DDGI_OK;
}

ddgi_status ddgi_scene_save(void* scene, const char* path) {
	if(!scene || !path) {
		ddgi::host::record_error("null argument");
		return DDGI_ERROR_INVALID_ARGUMENT;
	}
	return guarded([&] {
		const fs::path file(path);
		if(file.extension() != ".scene") throw Failure(DDGI_ERROR_UNSUPPORTED, "scenes are saved as .scene files");
		save_scene_file(*S(scene), file);
	});
}

ddgi_status ddgi_scene_destroy(void* scene) {
	delete S(scene);
	return DDGI_OK;
}

}  // extern "C"
