/*
 * On-disk scene ingest (include/lrh_ingest.h): a PLY reader and the geometry
 * part of LuxCore's scene-file parser, producing the mesh and object arrays
 * the intersection device builds from.
 *
 * Compiled with -ffp-contract=off and no fast-math: shape transformations
 * are applied to vertices and motion keys are inverted exactly as the
 * reference's CPU code rounds them (transform.h:131-142,
 * matrix4x4.cpp:118-175), because the device's results depend on them.
 */
#include "../../include/lrh_ingest.h"
#include "../csrc/lrh_math.h"

#include <algorithm>
#include <cerrno>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include <sys/stat.h>

namespace {

thread_local std::string gLastError;

struct IngestError : std::runtime_error {
	int code;
	IngestError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void fail(int code, const std::string &msg) { throw IngestError(code, msg); }

// ---------------------------------------------------------------------------
// PLY (exttrianglemeshfile.cpp:259-452 over rply, src/luxrays/utils/ply/rply.cpp)
// ---------------------------------------------------------------------------
enum PType { P_NONE, P_I8, P_U8, P_I16, P_U16, P_I32, P_U32, P_F32, P_F64 };

PType plyType(const std::string &s) {
	static const std::unordered_map<std::string, PType> names = {
		{"char", P_I8}, {"int8", P_I8}, {"uchar", P_U8}, {"uint8", P_U8},
		{"short", P_I16}, {"int16", P_I16}, {"ushort", P_U16}, {"uint16", P_U16},
		{"int", P_I32}, {"int32", P_I32}, {"uint", P_U32}, {"uint32", P_U32},
		{"float", P_F32}, {"float32", P_F32}, {"double", P_F64}, {"float64", P_F64}};
	auto it = names.find(s);
	return it == names.end() ? P_NONE : it->second;
}

struct PlyProp {
	std::string name;
	bool list = false;
	PType countType = P_NONE, type = P_NONE;
};

struct PlyElem {
	std::string name;
	uint64_t count = 0;
	std::vector<PlyProp> props;
};

struct MeshData {
	std::string name;
	std::vector<float> verts;      // 3 per vertex
	std::vector<uint32_t> tris;    // 3 per triangle
};

// Body reader.  Values come out as double, as rply hands them to the
// callbacks; the quirks of the reference's rply copy are kept: binary
// float32 NaN/Inf read as 0, ASCII integers and floats range-checked.
class PlyBody {
public:
	PlyBody(const char *p, const char *end, int format) : p_(p), end_(end), format_(format) {}

	bool value(PType t, double *v) { return format_ == 0 ? ascii(t, v) : binary(t, v); }

private:
	bool word(std::string *w) {
		while (p_ < end_ && (*p_ == ' ' || *p_ == '\t' || *p_ == '\n' || *p_ == '\r')) ++p_;
		const char *s = p_;
		while (p_ < end_ && !(*p_ == ' ' || *p_ == '\t' || *p_ == '\n' || *p_ == '\r')) ++p_;
		if (p_ == s) return false;
		w->assign(s, p_ - s);
		return true;
	}

	bool ascii(PType t, double *v) {
		if (!word(&w_)) return false;
		char *e = nullptr;
		const char *s = w_.c_str();
		switch (t) {
		case P_I8: { *v = (double)strtol(s, &e, 10); if (*e || *v > CHAR_MAX || *v < CHAR_MIN) return false; return true; }
		case P_U8: { *v = (double)strtol(s, &e, 10); if (*e || *v > UCHAR_MAX || *v < 0) return false; return true; }
		case P_I16: { *v = (double)strtol(s, &e, 10); if (*e || *v > SHRT_MAX || *v < SHRT_MIN) return false; return true; }
		case P_U16: { *v = (double)strtol(s, &e, 10); if (*e || *v > USHRT_MAX || *v < 0) return false; return true; }
		case P_I32: { *v = (double)strtol(s, &e, 10); if (*e) return false; return true; }
		case P_U32: { *v = (double)strtol(s, &e, 10); if (*e || *v < 0) return false; return true; }
		case P_F32: { *v = strtod(s, &e); if (*e || *v < -FLT_MAX || *v > FLT_MAX) return false; return true; }
		case P_F64: { *v = strtod(s, &e); if (*e || *v < -DBL_MAX || *v > DBL_MAX) return false; return true; }
		default: return false;
		}
	}

	bool chunk(void *dst, size_t n) {
		if ((size_t)(end_ - p_) < n) return false;
		unsigned char *d = (unsigned char *)dst;
		if (format_ == 1) {
			std::memcpy(d, p_, n);
		} else {
			for (size_t i = 0; i < n; ++i) d[i] = (unsigned char)p_[n - 1 - i];
		}
		p_ += n;
		return true;
	}

	bool binary(PType t, double *v) {
		switch (t) {
		case P_I8: { int8_t x; if (!chunk(&x, 1)) return false; *v = x; return true; }
		case P_U8: { uint8_t x; if (!chunk(&x, 1)) return false; *v = x; return true; }
		case P_I16: { int16_t x; if (!chunk(&x, 2)) return false; *v = x; return true; }
		case P_U16: { uint16_t x; if (!chunk(&x, 2)) return false; *v = x; return true; }
		case P_I32: { int32_t x; if (!chunk(&x, 4)) return false; *v = x; return true; }
		case P_U32: { uint32_t x; if (!chunk(&x, 4)) return false; *v = x; return true; }
		case P_F32: {
			float x;
			if (!chunk(&x, 4)) return false;
			*v = x;
			if (std::isnan(*v) || std::isinf(*v)) *v = 0.0;
			return true;
		}
		case P_F64: { double x; if (!chunk(&x, 8)) return false; *v = x; return true; }
		default: return false;
		}
	}

	const char *p_, *end_;
	int format_;   // 0 ascii, 1 binary little endian, 2 binary big endian
	std::string w_;
};

bool hostLittleEndian() {
	const uint16_t one = 1;
	unsigned char b;
	std::memcpy(&b, &one, 1);
	return b == 1;
}

std::vector<char> readFile(const std::string &path, const char *what) {
	FILE *f = std::fopen(path.c_str(), "rb");
	if (!f) fail(LRH_EIO, std::string("Unable to read ") + what + " '" + path + "'");
	std::vector<char> data;
	char buf[1 << 16];
	size_t n;
	while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) data.insert(data.end(), buf, buf + n);
	const bool bad = std::ferror(f) != 0;
	std::fclose(f);
	if (bad) fail(LRH_EIO, std::string("Error while reading ") + what + " '" + path + "'");
	return data;
}

std::vector<std::string> splitWords(const std::string &line) {
	std::vector<std::string> w;
	size_t i = 0;
	while (i < line.size()) {
		while (i < line.size() && (line[i] == ' ' || line[i] == '\t' || line[i] == '\r')) ++i;
		size_t j = i;
		while (j < line.size() && !(line[j] == ' ' || line[j] == '\t' || line[j] == '\r')) ++j;
		if (j > i) w.push_back(line.substr(i, j - i));
		i = j;
	}
	return w;
}

MeshData loadPly(const std::string &path) {
	const std::vector<char> data = readFile(path, "PLY mesh file");
	const char *p = data.data(), *end = data.data() + data.size();
	auto headerFail = [&]() { fail(LRH_EIO, "Unable to read PLY header from '" + path + "'"); };

	// header: one keyword per line up to end_header
	auto nextLine = [&](std::string *line) {
		if (p >= end) return false;
		const char *s = p;
		while (p < end && *p != '\n') ++p;
		line->assign(s, p - s);
		if (p < end) ++p;
		return true;
	};
	std::string line;
	if (!nextLine(&line) || splitWords(line) != std::vector<std::string>{"ply"}) headerFail();
	int format = -1;
	std::vector<PlyElem> elems;
	bool done = false;
	while (!done && nextLine(&line)) {
		const std::vector<std::string> w = splitWords(line);
		if (w.empty()) headerFail();
		if (w[0] == "format") {
			if (w.size() != 3 || w[2] != "1.0") headerFail();
			if (w[1] == "ascii") format = 0;
			else if (w[1] == "binary_little_endian") format = hostLittleEndian() ? 1 : 2;
			else if (w[1] == "binary_big_endian") format = hostLittleEndian() ? 2 : 1;
			else headerFail();
		} else if (w[0] == "comment" || w[0] == "obj_info") {
		} else if (w[0] == "element") {
			if (w.size() != 3) headerFail();
			char *e = nullptr;
			const long long c = strtoll(w[2].c_str(), &e, 10);
			if (*e || c < 0) headerFail();
			PlyElem el;
			el.name = w[1];
			el.count = (uint64_t)c;
			elems.push_back(el);
		} else if (w[0] == "property") {
			if (elems.empty()) headerFail();
			PlyProp pr;
			if (w.size() == 5 && w[1] == "list") {
				pr.list = true;
				pr.countType = plyType(w[2]);
				pr.type = plyType(w[3]);
				pr.name = w[4];
				if (pr.countType == P_NONE || pr.type == P_NONE) headerFail();
			} else if (w.size() == 3) {
				pr.type = plyType(w[1]);
				pr.name = w[2];
				if (pr.type == P_NONE) headerFail();
			} else
				headerFail();
			elems.back().props.push_back(pr);
		} else if (w[0] == "end_header") {
			done = true;
		} else
			headerFail();
	}
	if (!done || format < 0) headerFail();

	// which element/property feeds what (ply_set_read_cb lookups)
	auto findProp = [&](const char *elem, const char *prop, size_t *ei, size_t *pi) {
		for (size_t i = 0; i < elems.size(); ++i) {
			if (elems[i].name != elem) continue;
			for (size_t j = 0; j < elems[i].props.size(); ++j)
				if (elems[i].props[j].name == prop) { *ei = i; *pi = j; return true; }
			return false;
		}
		return false;
	};
	size_t vx = 0, px = 0, vy = 0, py = 0, vz = 0, pz = 0, fe = 0, pf = 0;
	const bool hasX = findProp("vertex", "x", &vx, &px);
	if (!hasX || elems[vx].count == 0) fail(LRH_EIO, "No vertices found in '" + path + "'");
	if (!findProp("vertex", "y", &vy, &py) || !findProp("vertex", "z", &vz, &pz))
		fail(LRH_EIO, "Vertex y or z missing in '" + path + "'");
	const bool hasFaces = findProp("face", "vertex_indices", &fe, &pf);
	if (!hasFaces || elems[fe].count == 0) fail(LRH_EIO, "No faces found in '" + path + "'");
	if (elems[vx].count > 0xffffffffull) fail(LRH_ERANGE, "Too many vertices in '" + path + "'");

	MeshData mesh;
	mesh.name = path;
	mesh.verts.assign(3 * elems[vx].count, 0.f);
	std::vector<double> raw;   // face indices as read, validated after the body
	raw.reserve(3 * elems[fe].count);

	PlyBody body(p, end, format);
	auto parseFail = [&]() { fail(LRH_EIO, "Unable to parse PLY file '" + path + "'"); };
	for (size_t ei = 0; ei < elems.size(); ++ei) {
		const PlyElem &el = elems[ei];
		for (uint64_t inst = 0; inst < el.count; ++inst) {
			for (size_t pi = 0; pi < el.props.size(); ++pi) {
				const PlyProp &pr = el.props[pi];
				double v;
				if (!pr.list) {
					if (!body.value(pr.type, &v)) parseFail();
					if (ei == vx) {
						if (pi == px) mesh.verts[3 * inst + 0] = (float)v;
						else if (pi == py) mesh.verts[3 * inst + 1] = (float)v;
						else if (pi == pz) mesh.verts[3 * inst + 2] = (float)v;
					}
					continue;
				}
				double len;
				if (!body.value(pr.countType, &len)) parseFail();
				if (!(len >= 0) || len != std::floor(len)) parseFail();
				const long n = (long)len;
				const bool face = (ei == fe && pi == pf);
				double q[4] = {0, 0, 0, 0};
				for (long k = 0; k < n; ++k) {
					if (!body.value(pr.type, &v)) parseFail();
					if (face && k < 4) q[k] = v;
				}
				if (!face) continue;
				// FaceCB: triangles as they are, quads split on the v0-v2
				// diagonal, every other length skipped
				if (n == 3) {
					raw.push_back(q[0]); raw.push_back(q[1]); raw.push_back(q[2]);
				} else if (n == 4) {
					raw.push_back(q[0]); raw.push_back(q[1]); raw.push_back(q[2]);
					raw.push_back(q[0]); raw.push_back(q[2]); raw.push_back(q[3]);
				}
			}
		}
	}

	const double nv = (double)elems[vx].count;
	mesh.tris.resize(raw.size());
	for (size_t i = 0; i < raw.size(); ++i) {
		const double v = raw[i];
		if (!(v >= 0) || !(v < nv) || v != std::floor(v))
			fail(LRH_EIO, "Vertex index out of range in '" + path + "'");
		mesh.tris[i] = (uint32_t)v;
	}
	return mesh;
}

// ---------------------------------------------------------------------------
// Properties (properties.cpp:652-749, 843-875, 951-966, 1099-1113)
// ---------------------------------------------------------------------------
std::string trim(const std::string &s) {
	size_t a = 0, b = s.size();
	while (a < b && std::isspace((unsigned char)s[a])) ++a;
	while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
	return s.substr(a, b - a);
}

class Properties {
public:
	void set(const std::string &name, std::vector<std::string> values) {
		if (!map_.count(name)) names_.push_back(name);
		map_[name] = std::move(values);
	}
	bool defined(const std::string &name) const { return map_.count(name) != 0; }
	const std::vector<std::string> &get(const std::string &name) const {
		auto it = map_.find(name);
		if (it == map_.end()) fail(LRH_EINVAL, "Missing property: " + name);
		return it->second;
	}
	// first-appearance order of "prefix.<field>" (GetAllUniqueSubNames)
	std::vector<std::string> uniqueSubNames(const std::string &prefix) const {
		std::vector<std::string> out;
		std::unordered_set<std::string> seen;
		const size_t want = (size_t)std::count(prefix.begin(), prefix.end(), '.') + 2;
		for (const std::string &n : names_) {
			if (n.compare(0, prefix.size(), prefix) != 0) continue;
			size_t idx = 0;
			bool ok = true;
			for (size_t i = 0; i < want; ++i) {
				if (idx >= n.size()) { ok = false; break; }
				idx = n.find('.', idx);
				if (idx == std::string::npos) { ok = false; break; }
				++idx;
			}
			if (!ok) continue;
			const std::string s = n.substr(0, idx - 1);
			if (!s.empty() && seen.insert(s).second) out.push_back(s);
		}
		return out;
	}

	void parse(const std::string &text, const std::string &where) {
		size_t pos = 0;
		int lineNumber = 0;
		while (pos <= text.size()) {
			size_t nl = text.find('\n', pos);
			if (nl == std::string::npos) nl = text.size();
			std::string line = trim(text.substr(pos, nl - pos));
			pos = nl + 1;
			++lineNumber;
			if (line.empty() || line[0] == '#') continue;
			const size_t eq = line.find('=');
			if (eq == std::string::npos)
				fail(LRH_EINVAL, "Syntax error in a Properties at line " + std::to_string(lineNumber) + " of " + where);
			const std::string name = trim(line.substr(0, eq));
			set(name, fields(name, trim(line.substr(eq + 1))));
		}
	}

private:
	static std::vector<std::string> fields(const std::string &name, const std::string &value) {
		std::vector<std::string> out;
		const size_t len = value.size();
		size_t first = 0;
		auto skipBlanks = [&](size_t i) {
			while (i < len && (value[i] == ' ' || value[i] == '\t')) ++i;
			return i;
		};
		while (first < len) {
			size_t last;
			if (first + 5 < len && value[first] == '{' && value[first + 1] == '[') {
				last = first;
				bool found = false;
				while (last < len - 1) {
					if (value[last] == ']' || value[last + 1] == '}') {
						out.push_back(value.substr(first, last - first + 2));
						last = skipBlanks(last + 2);
						found = true;
						break;
					}
					++last;
				}
				if (!found) fail(LRH_EINVAL, "Unterminated blob in property: " + name);
			} else if (value[first] == '"' || value[first] == '\'') {
				++first;
				last = first;
				while (last < len && value[last] != '"' && value[last] != '\'') ++last;
				if (last >= len) fail(LRH_EINVAL, "Unterminated quote in property: " + name);
				out.push_back(value.substr(first, last - first));
				last = skipBlanks(last + 1);
			} else {
				last = first;
				while (last < len && value[last] != ' ' && value[last] != '\t') ++last;
				out.push_back(value.substr(first, last - first));
				last = skipBlanks(last);
			}
			first = last;
		}
		return out;
	}

	std::vector<std::string> names_;
	std::unordered_map<std::string, std::vector<std::string>> map_;
};

float toFloat(const std::string &prop, const std::string &s) {
	char *e = nullptr;
	errno = 0;
	const float v = strtof(s.c_str(), &e);
	if (s.empty() || *e) fail(LRH_EINVAL, "Bad number '" + s + "' in property: " + prop);
	return v;
}

unsigned toUint(const std::string &prop, const std::string &s) {
	char *e = nullptr;
	const unsigned long long v = strtoull(s.c_str(), &e, 10);
	if (s.empty() || *e || v > 0xffffffffull) fail(LRH_EINVAL, "Bad index '" + s + "' in property: " + prop);
	return (unsigned)v;
}

typedef float Mat[4][4];

// Property::Get<Matrix4x4>: the 16 values are the matrix column by column
void getMatrix(const Properties &props, const std::string &name, Mat out) {
	const std::vector<std::string> &v = props.get(name);
	if (v.size() != 16) fail(LRH_EINVAL, "Wrong number of values in property: " + name);
	for (int i = 0; i < 4; ++i)
		for (int j = 0; j < 4; ++j) out[i][j] = toFloat(name, v[4 * j + i]);
}

void getPoint(const Properties &props, const std::string &name, const float def[3], float out[3]) {
	if (!props.defined(name)) { for (int k = 0; k < 3; ++k) out[k] = def[k]; return; }
	const std::vector<std::string> &v = props.get(name);
	if (v.size() != 3) fail(LRH_EINVAL, "Wrong number of values in property: " + name);
	for (int k = 0; k < 3; ++k) out[k] = toFloat(name, v[k]);
}

std::string getString(const Properties &props, const std::string &name, const std::string &def) {
	if (!props.defined(name)) return def;
	const std::vector<std::string> &v = props.get(name);
	return v.empty() ? std::string() : v[0];
}

bool fileExists(const std::string &p) {
	struct stat st;
	return ::stat(p.c_str(), &st) == 0;
}

std::string dirName(const std::string &p) {
	const size_t s = p.find_last_of('/');
	if (s == std::string::npos) return std::string();
	return s == 0 ? std::string("/") : p.substr(0, s);
}

std::string baseName(const std::string &p) {
	const size_t s = p.find_last_of('/');
	return s == std::string::npos ? p : p.substr(s + 1);
}

std::string joinPath(const std::string &d, const std::string &f) {
	if (d.empty()) return f;
	return d.back() == '/' ? d + f : d + "/" + f;
}

}  // namespace

// ---------------------------------------------------------------------------
// the scene
// ---------------------------------------------------------------------------
struct lrh_scenefile {
	std::vector<MeshData> meshes;
	struct Object {
		std::string name;
		unsigned mesh = 0;
		unsigned kind = 0;
		float m[16];
		std::vector<float> keyTimes;
		std::vector<float> keyMats;   // 16 per key, global->local
		float kd[3] = {0.75f, 0.75f, 0.75f};
		float le[3] = {0.f, 0.f, 0.f};
	};
	std::vector<Object> objects;
	lrh_scenefile_camera camera;
	std::vector<std::string> searchDirs;

	lrh_scenefile() {
		const float orig[3] = {0.f, 10.f, 0.f}, target[3] = {0.f, 0.f, 0.f}, up[3] = {0.f, 0.f, 1.f};
		for (int k = 0; k < 3; ++k) { camera.orig[k] = orig[k]; camera.target[k] = target[k]; camera.up[k] = up[k]; }
		camera.fieldOfView = 45.f;
	}

	std::string resolve(const std::string &name) const {   // FileNameResolver::ResolveFile
		if (name.empty() || name[0] == '/' || fileExists(name)) return name;
		for (const std::string &d : searchDirs) {
			const std::string p = joinPath(d, name);
			if (fileExists(p)) return p;
		}
		for (const std::string &d : searchDirs) {
			const std::string p = joinPath(d, baseName(name));
			if (fileExists(p)) return p;
		}
		return name;
	}

	void build(const Properties &props);
};

namespace {

// shape types whose Refine keeps the source mesh's vertices and triangles
bool geometryPreserving(const std::string &t) {
	return t == "pointiness" || t == "harlequin" || t == "islandaov" || t == "randomtriangleaov" ||
	       t == "edgedetectoraov" || t == "cameraprojuv";
}

// a constant color: 1 or 3 numbers, or a constfloat1 / constfloat3 texture
// (parsetextures.cpp); false for anything else
bool constColor(const Properties &props, const std::string &name, float out[3]) {
	if (!props.defined(name)) return false;
	const std::vector<std::string> &v = props.get(name);
	auto num = [](const std::string &t, float *f) {
		char *e = nullptr;
		*f = strtof(t.c_str(), &e);
		return !t.empty() && !*e;
	};
	float f[3];
	if (v.size() == 3 && num(v[0], &f[0]) && num(v[1], &f[1]) && num(v[2], &f[2])) {
		for (int k = 0; k < 3; ++k) out[k] = f[k];
		return true;
	}
	if (v.size() == 1 && num(v[0], &f[0])) {
		for (int k = 0; k < 3; ++k) out[k] = f[0];
		return true;
	}
	if (v.size() == 1) {
		const std::string tex = "scene.textures." + v[0];
		const std::string type = getString(props, tex + ".type", "");
		if (type == "constfloat3" || type == "constfloat1") return constColor(props, tex + ".value", out);
	}
	return false;
}

void materialColors(const Properties &props, const std::string &mat, float kd[3], float le[3]) {
	for (int k = 0; k < 3; ++k) { kd[k] = 0.75f; le[k] = 0.f; }
	constColor(props, mat + ".kd", kd);
	float e[3], g[3] = {1.f, 1.f, 1.f};
	if (constColor(props, mat + ".emission", e)) {
		constColor(props, mat + ".emission.gain", g);
		for (int k = 0; k < 3; ++k) le[k] = e[k] * g[k];
	}
}

void applyTransform(MeshData *mesh, const Mat m) {   // TriangleMesh::ApplyTransform (trianglemesh.cpp:85-93)
	for (size_t i = 0; i + 2 < mesh->verts.size(); i += 3) {
		const float p[3] = {mesh->verts[i], mesh->verts[i + 1], mesh->verts[i + 2]};
		float r[3];
		lrh::xfPoint4(m, p, r);
		mesh->verts[i] = r[0]; mesh->verts[i + 1] = r[1]; mesh->verts[i + 2] = r[2];
	}
}

MeshData inlinedMesh(const Properties &props, const std::string &propName, const std::string &shapeName) {
	// Scene::CreateInlinedMesh (parseshapes.cpp:78-139)
	MeshData m;
	m.name = shapeName;
	if (!props.defined(propName + ".vertices")) fail(LRH_EINVAL, "Missing shape vertex list: " + shapeName);
	const std::vector<std::string> &v = props.get(propName + ".vertices");
	if (v.empty() || v.size() % 3 != 0) fail(LRH_EINVAL, "Wrong shape vertex list length: " + shapeName);
	m.verts.resize(v.size());
	for (size_t i = 0; i < v.size(); ++i) m.verts[i] = toFloat(propName + ".vertices", v[i]);
	if (!props.defined(propName + ".faces")) fail(LRH_EINVAL, "Missing shape face list: " + shapeName);
	const std::vector<std::string> &f = props.get(propName + ".faces");
	if (f.empty() || f.size() % 3 != 0) fail(LRH_EINVAL, "Wrong shape face list length: " + shapeName);
	m.tris.resize(f.size());
	const size_t nv = v.size() / 3;
	for (size_t i = 0; i < f.size(); ++i) {
		m.tris[i] = toUint(propName + ".faces", f[i]);
		if (m.tris[i] >= nv) fail(LRH_EINVAL, "Vertex index out of range in shape: " + shapeName);
	}
	return m;
}

}  // namespace

void lrh_scenefile::build(const Properties &props) {
	// camera (parsecamera.cpp:50-71): the deprecated 6-value lookat forces up
	// to +z; without lookat.orig every vector stays zero (identity camera)
	for (int k = 0; k < 3; ++k) camera.orig[k] = camera.target[k] = camera.up[k] = 0.f;
	if (props.defined("scene.camera.lookat")) {
		const std::vector<std::string> &v = props.get("scene.camera.lookat");
		if (v.size() < 6) fail(LRH_EINVAL, "Out of bound index in property: scene.camera.lookat");
		for (int k = 0; k < 3; ++k) {
			camera.orig[k] = toFloat("scene.camera.lookat", v[k]);
			camera.target[k] = toFloat("scene.camera.lookat", v[3 + k]);
		}
		camera.up[2] = 1.f;
	} else if (props.defined("scene.camera.lookat.orig")) {
		const float o[3] = {0.f, 10.f, 0.f}, t[3] = {0.f, 0.f, 0.f}, u[3] = {0.f, 0.f, 1.f};
		getPoint(props, "scene.camera.lookat.orig", o, camera.orig);
		getPoint(props, "scene.camera.lookat.target", t, camera.target);
		getPoint(props, "scene.camera.up", u, camera.up);
	}
	if (props.defined("scene.camera.fieldofview"))
		camera.fieldOfView = toFloat("scene.camera.fieldofview", props.get("scene.camera.fieldofview").at(0));

	// materials: only their names matter here (CreateObject checks them)
	std::unordered_set<std::string> materials;
	for (const std::string &k : props.uniqueSubNames("scene.materials"))
		materials.insert(k.substr(std::string("scene.materials.").size()));

	// shapes, in definition order (Scene::ParseShapes / CreateShape)
	std::unordered_map<std::string, unsigned> meshIndex;   // extMeshCache
	auto defineMesh = [&](MeshData m) {
		auto it = meshIndex.find(m.name);
		if (it != meshIndex.end()) {
			meshes[it->second] = std::move(m);
			return it->second;
		}
		const unsigned idx = (unsigned)meshes.size();
		meshIndex[m.name] = idx;
		meshes.push_back(std::move(m));
		return idx;
	};
	for (const std::string &key : props.uniqueSubNames("scene.shapes")) {
		const std::string shapeName = key.substr(std::string("scene.shapes.").size());
		const std::string type = getString(props, key + ".type", "mesh");
		MeshData m;
		if (type == "mesh" || type == "inlinedmesh") {
			if (type == "mesh")
				m = loadPly(resolve(getString(props, key + ".ply", "")));
			else
				m = inlinedMesh(props, key, shapeName);
			if (props.defined(key + ".transformation")) {
				Mat t;
				getMatrix(props, key + ".transformation", t);
				applyTransform(&m, t);
			}
		} else if (geometryPreserving(type)) {
			const std::string src = getString(props, key + ".source", "");
			auto it = meshIndex.find(src);
			if (it == meshIndex.end())
				fail(LRH_EINVAL, "Unknown shape name in a " + type + " shape: " + shapeName);
			m = meshes[it->second];
		} else if (type == "strands" || type == "group" || type == "subdiv" || type == "displacement" ||
		           type == "simplify" || type == "bevel") {
			fail(LRH_EINVAL, "Shape type " + type + " changes geometry and is not supported by the scene reader: " +
			                     shapeName);
		} else
			fail(LRH_EINVAL, "Unknown shape type: " + type);
		m.name = shapeName;
		defineMesh(std::move(m));
	}

	// objects, in definition order (Scene::ParseObjects / CreateObject)
	std::unordered_map<std::string, unsigned> objIndex;
	for (const std::string &key : props.uniqueSubNames("scene.objects")) {
		const std::string objName = key.substr(std::string("scene.objects.").size());
		const std::string matName = getString(props, key + ".material", "");
		if (matName.empty()) fail(LRH_EINVAL, "Syntax error in object material reference: " + objName);
		if (!materials.count(matName)) fail(LRH_EINVAL, "Unknown material: " + matName);

		std::string shapeName;
		if (props.defined(key + ".ply")) {
			shapeName = getString(props, key + ".ply", "");
			if (!meshIndex.count(shapeName)) {
				MeshData m = loadPly(resolve(shapeName));
				m.name = shapeName;
				defineMesh(std::move(m));
			}
		} else if (props.defined(key + ".vertices")) {
			shapeName = "InlinedMesh-" + objName;
			if (!meshIndex.count(shapeName)) defineMesh(inlinedMesh(props, key, shapeName));
		} else if (props.defined(key + ".shape")) {
			shapeName = getString(props, key + ".shape", "");
			if (!meshIndex.count(shapeName)) fail(LRH_EINVAL, "Unknown shape: " + shapeName);
		} else
			fail(LRH_EINVAL, "Missing shape in object definition: " + objName);

		Object o;
		o.name = objName;
		o.mesh = meshIndex.at(shapeName);
		materialColors(props, "scene.materials." + matName, o.kd, o.le);
		for (int i = 0; i < 16; ++i) o.m[i] = (i % 5 == 0) ? 1.f : 0.f;
		if (props.defined(key + ".motion.0.time")) {
			o.kind = 2;
			for (unsigned i = 0;; ++i) {
				const std::string prefix = key + ".motion." + std::to_string(i);
				if (!props.defined(prefix + ".time")) break;
				const float t = toFloat(prefix + ".time", props.get(prefix + ".time").at(0));
				if (i > 0 && t <= o.keyTimes.back()) fail(LRH_EINVAL, objName + " motion time must be monotonic");
				o.keyTimes.push_back(t);
				Mat l2w = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
				if (props.defined(prefix + ".transformation")) getMatrix(props, prefix + ".transformation", l2w);
				// keys are global->local: Inverse(Transform(mat)).m = mat's inverse
				Mat w2l;
				if (!lrh::invert4(l2w, w2l)) fail(LRH_EINVAL, "Singular matrix in " + prefix + ".transformation");
				o.keyMats.insert(o.keyMats.end(), &w2l[0][0], &w2l[0][0] + 16);
			}
		} else if (props.defined(key + ".transformation")) {
			o.kind = 1;
			Mat l2w;
			getMatrix(props, key + ".transformation", l2w);
			Mat inv;
			if (!lrh::invert4(l2w, inv)) fail(LRH_EINVAL, "Singular matrix in " + key + ".transformation");
			std::memcpy(o.m, l2w, sizeof(o.m));
		}
		auto it = objIndex.find(objName);
		if (it != objIndex.end())
			objects[it->second] = std::move(o);   // a redefinition replaces the object in place
		else {
			objIndex[objName] = (unsigned)objects.size();
			objects.push_back(std::move(o));
		}
	}
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
namespace {

template <class F>
int guard(F &&f) {
	try {
		f();
		return LRH_OK;
	} catch (const IngestError &e) {
		gLastError = e.what();
		return e.code;
	} catch (const std::bad_alloc &) {
		gLastError = "out of host memory";
		return LRH_ENOMEM;
	} catch (const std::exception &e) {
		gLastError = e.what();
		return LRH_EINVAL;
	}
}

void loadText(lrh_scenefile *sf, Properties *props, const std::string &text, const std::string &where) {
	props->parse(text, where);
	// a render configuration names its scene with scene.file
	if (props->defined("scene.file")) {
		const std::string scn = sf->resolve(getString(*props, "scene.file", ""));
		const std::vector<char> data = readFile(scn, "scene file");
		const std::string dir = dirName(scn);
		if (!dir.empty()) sf->searchDirs.push_back(dir);
		Properties inner;
		inner.parse(std::string(data.begin(), data.end()), scn);
		*props = inner;
	}
}

}  // namespace

extern "C" {

int lrh_ply_load(const char *path, lrh_scenefile **out) {
	return guard([&] {
		if (!path || !out) fail(LRH_EINVAL, "null argument");
		*out = nullptr;
		std::unique_ptr<lrh_scenefile> sf(new lrh_scenefile());
		MeshData m = loadPly(path);
		sf->meshes.push_back(std::move(m));
		lrh_scenefile::Object o;
		o.name = path;
		for (int i = 0; i < 16; ++i) o.m[i] = (i % 5 == 0) ? 1.f : 0.f;
		sf->objects.push_back(o);
		*out = sf.release();
	});
}

int lrh_scenefile_load(const char *path, const char *const *searchDirs, unsigned int nDirs, lrh_scenefile **out) {
	return guard([&] {
		if (!path || !out || (nDirs && !searchDirs)) fail(LRH_EINVAL, "null argument");
		*out = nullptr;
		std::unique_ptr<lrh_scenefile> sf(new lrh_scenefile());
		sf->searchDirs.push_back(dirName(path));
		for (unsigned i = 0; i < nDirs; ++i)
			if (searchDirs[i]) sf->searchDirs.push_back(searchDirs[i]);
		const std::vector<char> data = readFile(path, "scene file");
		Properties props;
		loadText(sf.get(), &props, std::string(data.begin(), data.end()), path);
		sf->build(props);
		*out = sf.release();
	});
}

int lrh_scenefile_parse(const char *text, size_t len, const char *baseDir, lrh_scenefile **out) {
	return guard([&] {
		if ((!text && len) || !out) fail(LRH_EINVAL, "null argument");
		*out = nullptr;
		std::unique_ptr<lrh_scenefile> sf(new lrh_scenefile());
		sf->searchDirs.push_back(baseDir ? baseDir : "");
		Properties props;
		loadText(sf.get(), &props, std::string(text ? text : "", len), "<text>");
		sf->build(props);
		*out = sf.release();
	});
}

int lrh_scenefile_get_info(const lrh_scenefile *sf, lrh_scenefile_info *info) {
	return guard([&] {
		if (!sf || !info) fail(LRH_EINVAL, "null argument");
		std::memset(info, 0, sizeof(*info));
		info->nMeshes = (unsigned)sf->meshes.size();
		info->nObjects = (unsigned)sf->objects.size();
		for (const MeshData &m : sf->meshes) {
			info->nVerts += m.verts.size() / 3;
			info->nTris += m.tris.size() / 3;
		}
		for (const lrh_scenefile::Object &o : sf->objects) {
			info->nKeys += (unsigned)o.keyTimes.size();
			if (o.kind != 0) info->twoLevel = 1;
			info->nWorldTris += sf->meshes[o.mesh].tris.size() / 3;
		}
	});
}

int lrh_scenefile_meshes(const lrh_scenefile *sf, lrh_mesh *out, unsigned int n) {
	return guard([&] {
		if (!sf || (n && !out)) fail(LRH_EINVAL, "null argument");
		if (n != sf->meshes.size()) fail(LRH_EINVAL, "mesh count mismatch");
		for (unsigned i = 0; i < n; ++i) {
			const MeshData &m = sf->meshes[i];
			out[i].verts = m.verts.data();
			out[i].nVerts = (unsigned)(m.verts.size() / 3);
			out[i].tris = m.tris.data();
			out[i].nTris = (unsigned)(m.tris.size() / 3);
		}
	});
}

int lrh_scenefile_objects(const lrh_scenefile *sf, lrh_object *out, unsigned int n) {
	return guard([&] {
		if (!sf || (n && !out)) fail(LRH_EINVAL, "null argument");
		if (n != sf->objects.size()) fail(LRH_EINVAL, "object count mismatch");
		for (unsigned i = 0; i < n; ++i) {
			const lrh_scenefile::Object &o = sf->objects[i];
			std::memset(&out[i], 0, sizeof(out[i]));
			out[i].mesh = o.mesh;
			out[i].kind = o.kind;
			if (o.kind == 1) out[i].m = o.m;
			if (o.kind == 2) {
				out[i].nKeys = (unsigned)o.keyTimes.size();
				out[i].keyTimes = o.keyTimes.data();
				out[i].keyMats = o.keyMats.data();
			}
		}
	});
}

int lrh_scenefile_object_name(const lrh_scenefile *sf, unsigned int obj, const char **name, const char **meshName) {
	return guard([&] {
		if (!sf) fail(LRH_EINVAL, "null argument");
		if (obj >= sf->objects.size()) fail(LRH_EINVAL, "object index out of range");
		if (name) *name = sf->objects[obj].name.c_str();
		if (meshName) *meshName = sf->meshes[sf->objects[obj].mesh].name.c_str();
	});
}

int lrh_scenefile_object_material(const lrh_scenefile *sf, unsigned int obj, float kd[3], float le[3]) {
	return guard([&] {
		if (!sf || !kd || !le) fail(LRH_EINVAL, "null argument");
		if (obj >= sf->objects.size()) fail(LRH_EINVAL, "object index out of range");
		for (int k = 0; k < 3; ++k) {
			kd[k] = sf->objects[obj].kd[k];
			le[k] = sf->objects[obj].le[k];
		}
	});
}

int lrh_scenefile_get_camera(const lrh_scenefile *sf, lrh_scenefile_camera *cam) {
	return guard([&] {
		if (!sf || !cam) fail(LRH_EINVAL, "null argument");
		*cam = sf->camera;
	});
}

int lrh_scenefile_destroy(lrh_scenefile *sf) {
	delete sf;
	return LRH_OK;
}

const char *lrh_ingest_last_error(void) { return gLastError.c_str(); }

}  // extern "C"
