/*
 * C-ABI layer (include/lrhip.h) of the HIP intersection device: devices,
 * buffers and queues (HardwareDevice), scene build + upload (BVHAccel /
 * MBVHAccel Init plus the BVHKernel / MBVHKernel upload), and ray-buffer
 * traces (HardwareIntersectionKernel::EnqueueTraceRayBuffer).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <atomic>
#include <exception>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <map>
#include <set>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/lrhip.h"
#include "lrh_build.h"
#include "lrh_gpubuild.h"
#include "lrh_kernels.h"
#include "lrh_paths.h"
#include "lrh_refit.h"
#include "lrh_layout.h"
#include "lrh_internal.h"

using lrh::Box;
using lrh::Error;

static_assert(sizeof(lrh_node4q) == 64, "node4q must be 64 bytes");
static_assert(sizeof(lrh_node4f) == 128, "node4f must be 128 bytes");
static_assert(sizeof(lrh_node8q) == 128, "node8q must be 128 bytes");
static_assert(sizeof(lrh_tri) == 48, "tri record must be 48 bytes");
static_assert(sizeof(lrh_motion) == 32, "motion record must be 32 bytes");
static_assert(sizeof(lrh_motion_seg) == 304 && offsetof(lrh_motion_seg, start) == LRH_SEG_HOT_BYTES,
		"motion segment: 112 hot bytes, then the matrices");
static_assert(sizeof(lrh_ray_rec) == LRH_RAY_BYTES, "Ray must be 48 bytes");
static_assert(sizeof(lrh_hit_rec) == LRH_HIT_BYTES, "RayHit must be 20 bytes");

namespace {

thread_local std::string g_lastError;

int fail(int code, const std::string &msg) {
	g_lastError = msg;
	return code;
}

#define HIPCHK(expr)                                                                            \
	do {                                                                                        \
		hipError_t _e = (expr);                                                                 \
		if (_e != hipSuccess)                                                                   \
			throw Error(LRH_EHIP, std::string(#expr) + ": " + hipGetErrorString(_e));           \
	} while (0)

template <class F>
int guarded(F fn) {
	try {
		fn();
		return LRH_OK;
	} catch (const Error &e) {
		return fail(e.code, e.what());
	} catch (const std::bad_alloc &) {
		return fail(LRH_ENOMEM, "host allocation failed");
	} catch (const std::exception &e) {
		return fail(LRH_EINVAL, e.what());
	}
}

double now() {
	return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

// Scratch memory of lrh_trace_noalloc: the ray-ticket counter (first 256
// bytes) and the spill stack, in one statically sized array that every GPU
// the library runs on holds from the loading of this code object until the
// process ends.  In host code `scratch` is only the host-side stand-in of the
// array; its device address on each GPU comes from hipGetSymbolAddress,
// taken once per device in lrh_device_open with that device current.
static __device__ __align__(256) unsigned char scratch[LRH_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr size_t kScratchCounterBytes = 256;
constexpr int kMaxOrdinals = 64;
std::mutex g_scratchMu;
void *g_scratchPtr[kMaxOrdinals] = {};

// device address of `scratch` on the current device (ordinal), resolved once
int resolveScratch(int ordinal, void **out) {
	if (ordinal < 0 || ordinal >= kMaxOrdinals) return fail(LRH_ERANGE, "device ordinal beyond the scratch table");
	std::lock_guard<std::mutex> g(g_scratchMu);
	if (!g_scratchPtr[ordinal]) {
		void *p = nullptr;
		const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (e != hipSuccess || !p)
			return fail(LRH_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
		g_scratchPtr[ordinal] = p;
	}
	*out = g_scratchPtr[ordinal];
	return LRH_OK;
}
}  // namespace

// the message of lrh_last_error for the library's other host files
int lrhSetError(int code, const std::string &msg) { return fail(code, msg); }

namespace {

}  // namespace

struct lrh_device {
	int ordinal = 0;
	hipStream_t stream = nullptr;
	int cuCount = 0;
	size_t totalMem = 0;
	std::string name;
	uint32_t *dError = nullptr;
	// per-stream trace workspace: ray-ticket counter + spill stack, so that
	// traces issued on different streams never share scratch memory
	struct Workspace {
		uint32_t *counter = nullptr;      // [0] tickets, [1] finished blocks
		unsigned long long *stats = nullptr;
		uint2 *spill = nullptr;
		size_t spillBytes = 0;
		uint8_t *table = nullptr;         // batched-launch descriptor table
		size_t tableBytes = 0;
	};
	std::map<hipStream_t, Workspace> ws;
	std::atomic<size_t> used{0};
	std::mutex mu;
	std::map<size_t, size_t> allocs;  // lrh_alloc'd pointers -> bytes
	std::map<uint64_t, int> occ;      // occupancy cache
	std::set<lrh_scene *> scenes;     // live scenes (destroyed with the device)
	std::atomic<unsigned long long> rays{0};
	double t0 = 0;
	void *scratch = nullptr;  // device address of the no-alloc scratch array on this GPU
	std::string scratchError;  // why it could not be resolved

	void activate() const { HIPCHK(hipSetDevice(ordinal)); }
	void *devAlloc(size_t bytes) {
		void *p = nullptr;
		if (bytes == 0) return nullptr;
		const hipError_t e = hipMalloc(&p, bytes);
		if (e != hipSuccess) throw Error(LRH_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
		used += bytes;
		return p;
	}
	void devFree(void *p, size_t bytes) {
		if (!p) return;
		(void)hipFree(p);
		used -= bytes;
	}
	// blocks a device build released, kept (up to kBuildPoolMax, counted in
	// `used`) for the next build asking for the same sizes: a build's scratch
	// runs to gigabytes, and releasing it and mapping it again on every build
	// came to several times the build once the runtime's own cache was full
	std::mutex poolMu;
	std::multimap<size_t, void *> buildPool;
	size_t pooledBytes = 0;
	static constexpr size_t kBuildPoolMax = size_t(32) << 30;
	void *buildAlloc(size_t bytes) {
		{
			std::lock_guard<std::mutex> g(poolMu);
			auto it = buildPool.find(bytes);
			if (it != buildPool.end()) {
				void *p = it->second;
				buildPool.erase(it);
				pooledBytes -= bytes;
				return p;
			}
		}
		return devAlloc(bytes);
	}
	void buildRelease(void *p, size_t bytes) {
		if (!p) return;
		std::lock_guard<std::mutex> g(poolMu);
		if (bytes > kBuildPoolMax) {
			devFree(p, bytes);
			return;
		}
		while (pooledBytes + bytes > kBuildPoolMax && !buildPool.empty()) {  // largest blocks go first
			auto it = std::prev(buildPool.end());
			devFree(it->second, it->first);
			pooledBytes -= it->first;
			buildPool.erase(it);
		}
		buildPool.emplace(bytes, p);
		pooledBytes += bytes;
	}
	void drainPool() {
		std::lock_guard<std::mutex> g(poolMu);
		for (auto &b : buildPool) devFree(b.second, b.first);
		buildPool.clear();
		pooledBytes = 0;
	}
};

struct DevBuf {
	lrh_device *dev = nullptr;
	void *p = nullptr;
	size_t bytes = 0;
	void alloc(lrh_device *d, size_t b) {
		release();
		dev = d;
		p = d->devAlloc(b);
		bytes = b;
	}
	void adopt(lrh_device *d, void *ptr, size_t b) {  // take ownership of a devAlloc'ed block
		release();
		dev = d;
		p = ptr;
		bytes = b;
	}
	void upload(const void *src, size_t b, hipStream_t s) {
		if (b) HIPCHK(hipMemcpyAsync(p, src, b, hipMemcpyHostToDevice, s));
	}
	void release() {
		if (p && dev) dev->devFree(p, bytes);
		p = nullptr;
		bytes = 0;
	}
	~DevBuf() { release(); }
};

namespace {
uint64_t nextSceneSerial() {
	static std::atomic<uint64_t> n{0};
	return ++n;
}
}  // namespace

struct lrh_event {
	hipEvent_t ev = nullptr;
	lrh_device *dev = nullptr;
};

struct lrh_scene {
	lrh_device *dev = nullptr;
	bool twoLevel = false;
	unsigned layout = 0;
	uint32_t root = LRH_CHILD_EMPTY;
	uint32_t maxDepth = 0;
	DevBuf nodes, tris, insts, xforms, motions, segs, times;
	DevBuf invPerm, meshTriBase;  // single level: (mesh, triangle) -> record, for hit shading
	unsigned nMeshes = 0;
	lrh_scene_info info{};
	// two-level state kept for lrh_scene_update
	lrh::BuildParams params;
	std::vector<unsigned> objMesh, objKind, objLeafRoot;
	std::vector<Box> meshBox;
	size_t rootReserve = 0;  // node slots reserved for the root tree
	uint32_t leafDepth = 0;
	// root topology kept for lrh_scene_refit: leaf slot -> object, node
	// indices by level (deepest first), object boxes of the last host build
	std::vector<uint32_t> rootOrder, levelNodes, levelStart;
	std::vector<Box> objBoxes;  // object boxes of the last accepted build, update or refit
	// bumped whenever an edit moves one of the scene's device arrays or deepens
	// its tree: a resident batch built before then would read freed memory or
	// overflow its spill stack, so it refuses to launch
	uint64_t generation = 0;
	// process-wide creation number: a new scene at a destroyed one's address
	// is told apart from it
	uint64_t serial = nextSceneSerial();
	DevBuf rSlotBoxes, rNodeBoxes, rLevelNodes, rMeshBoxes, rMotionObjs, rSampleBase, rKeys, rError, rMotions, rSegs;
	// host-only scenes (created with dev == NULL) keep their arrays here so
	// the builder can be inspected without a GPU; they cannot be traced
	std::vector<uint8_t> hostNodes, hostPrims;
};

// ---------------------------------------------------------------------------
// scene construction helpers
// ---------------------------------------------------------------------------
namespace {

void checkMesh(const lrh_mesh &m, unsigned i) {
	if (m.nTris && (!m.verts || !m.tris)) throw Error(LRH_EINVAL, "mesh " + std::to_string(i) + ": null arrays");
	for (size_t k = 0; k < 3 * (size_t)m.nTris; ++k)
		if (m.tris[k] >= m.nVerts)
			throw Error(LRH_EINVAL, "mesh " + std::to_string(i) + ": vertex index out of range");
}

lrh::BuildParams toParams(const lrh_build_opts *o) {
	lrh::BuildParams p;
	if (o) {
		if (o->layout > 2) throw Error(LRH_EINVAL, "layout must be 0, 1 or 2");
		p.layout = o->layout;
		if (o->maxLeafSize) p.maxLeaf = o->maxLeafSize;
		if (o->nThreads) p.nThreads = o->nThreads;
		if (o->travCost > 0.f) p.travCost = o->travCost;
	}
	// early split clipping: 0 selects the default budget, a negative value turns it off
	p.splitBudget = (o && o->splitBudget != 0.f) ? o->splitBudget : 4.0f;
	if (!(p.splitBudget <= 64.f)) throw Error(LRH_EINVAL, "splitBudget must be at most 64");
	if (p.maxLeaf < 1 || p.maxLeaf > LRH_LEAF_MAX) throw Error(LRH_EINVAL, "maxLeafSize must be 1..8");
	return p;
}

template <class F>
void parallelRange(size_t n, unsigned nt, F fn) {
	if (nt <= 1 || n < 100000) { fn(0, n); return; }
	std::vector<std::thread> th;
	const size_t chunk = (n + nt - 1) / nt;
	for (unsigned t = 0; t < nt; ++t) {
		const size_t b = t * chunk, e = std::min(n, b + chunk);
		if (b >= e) break;
		th.emplace_back([=, &fn]() { fn(b, e); });
	}
	for (auto &t : th) t.join();
}

struct HostTree {
	std::vector<lrh_node4q> q;
	std::vector<lrh_node4f> f;
	std::vector<lrh_node8q> q8;
	size_t count(unsigned layout) const { return layout == 2 ? q8.size() : layout ? f.size() : q.size(); }
	void append(const lrh::WideBVH &w, unsigned layout) {
		if (layout == 2) q8.insert(q8.end(), w.q8nodes.begin(), w.q8nodes.end());
		else if (layout) f.insert(f.end(), w.fnodes.begin(), w.fnodes.end());
		else q.insert(q.end(), w.qnodes.begin(), w.qnodes.end());
	}
	const void *data(unsigned layout) const {
		return layout == 2 ? (const void *)q8.data() : layout ? (const void *)f.data() : (const void *)q.data();
	}
	size_t bytes(unsigned layout) const {
		return layout == 2 ? q8.size() * sizeof(lrh_node8q)
				: layout ? f.size() * sizeof(lrh_node4f) : q.size() * sizeof(lrh_node4q);
	}
	void reserve(size_t n, unsigned layout) {
		if (layout == 2) q8.resize(q8.size() + n);
		else if (layout) f.resize(f.size() + n);
		else q.resize(q.size() + n);
	}
};

// build a mesh's triangle tree; appends nodes and triangle records
uint32_t buildMeshTree(const lrh_mesh *meshes, const std::vector<unsigned> &meshList, bool localMeshIndex,
		const lrh::BuildParams &p, HostTree &ht, std::vector<lrh_tri> &tris, uint32_t *maxDepth,
		std::vector<uint32_t> *invPerm = nullptr) {
	// global primitive -> (mesh, tri)
	std::vector<size_t> off(meshList.size() + 1, 0);
	for (size_t i = 0; i < meshList.size(); ++i) off[i + 1] = off[i] + meshes[meshList[i]].nTris;
	const size_t n = off.back();
	if (n == 0) return LRH_CHILD_EMPTY;
	if (tris.size() + n > LRH_MAX_LEAF_START) throw Error(LRH_ERANGE, "scene exceeds 268,435,455 triangle records");
	std::vector<Box> boxes(n);
	const unsigned nt = p.nThreads ? p.nThreads : lrh::hardwareThreads();
	for (size_t i = 0; i < meshList.size(); ++i) {
		const lrh_mesh &m = meshes[meshList[i]];
		parallelRange(m.nTris, nt, [&](size_t b, size_t e) {
			for (size_t k = b; k < e; ++k) {
				const unsigned *t = m.tris + 3 * k;
				boxes[off[i] + k] = lrh::triangleBox(m.verts + 3 * (size_t)t[0], m.verts + 3 * (size_t)t[1],
						m.verts + 3 * (size_t)t[2]);
			}
		});
	}
	// large triangles cut into several references (early split clipping)
	struct Source : lrh::TriangleSource {
		const lrh_mesh *meshes;
		const std::vector<unsigned> *list;
		const std::vector<size_t> *off;
		void get(size_t prim, float q[3][3]) const override {
			const size_t mi = std::upper_bound(off->begin(), off->end(), prim) - off->begin() - 1;
			const lrh_mesh &m = meshes[(*list)[mi]];
			const unsigned *t = m.tris + 3 * (prim - (*off)[mi]);
			for (int c = 0; c < 3; ++c)
				for (int k = 0; k < 3; ++k) q[c][k] = m.verts[3 * (size_t)t[c] + k];
		}
	} src;
	src.meshes = meshes;
	src.list = &meshList;
	src.off = &off;
	std::vector<Box> refBoxes;
	std::vector<uint32_t> refPrim;
	// the extra references must fit the 28-bit record offsets of the leaves
	const double room = (double)LRH_MAX_LEAF_START - (double)tris.size() - (double)n;
	const float budget = (float)std::min((double)p.splitBudget, room / (double)n);
	lrh::splitLargeTriangles(boxes, src, budget, nt, &refBoxes, &refPrim);
	if (!refPrim.empty() && (double)tris.size() + (double)refPrim.size() > (double)LRH_MAX_LEAF_START) {
		// the pieces came out above the estimate and past the record limit: keep the triangles whole
		refBoxes.clear();
		refPrim.clear();
	}
	const bool cut = !refPrim.empty();
	const size_t nRefs = cut ? refPrim.size() : n;
	if (tris.size() + nRefs > LRH_MAX_LEAF_START)
		throw Error(LRH_ERANGE, "scene exceeds 268,435,455 triangle records");
	lrh::WideBVH w;
	const uint32_t nodeBase = (uint32_t)ht.count(p.layout);
	const uint32_t primBase = (uint32_t)tris.size();
	lrh::buildWideBVH(cut ? refBoxes : boxes, p, nodeBase, primBase, &w);
	boxes.clear();
	boxes.shrink_to_fit();
	refBoxes.clear();
	refBoxes.shrink_to_fit();
	ht.append(w, p.layout);
	tris.resize(primBase + nRefs);
	parallelRange(nRefs, nt, [&](size_t b, size_t e) {
		for (size_t i = b; i < e; ++i) {
			const size_t prim = cut ? refPrim[w.order[i]] : w.order[i];
			const size_t mi = std::upper_bound(off.begin(), off.end(), prim) - off.begin() - 1;
			const lrh_mesh &m = meshes[meshList[mi]];
			const size_t k = prim - off[mi];
			const unsigned *t = m.tris + 3 * k;
			const float *p0 = m.verts + 3 * (size_t)t[0], *p1 = m.verts + 3 * (size_t)t[1], *p2 = m.verts + 3 * (size_t)t[2];
			lrh_tri &r = tris[primBase + i];
			for (int c = 0; c < 3; ++c) {
				r.p0[c] = p0[c];
				r.e1[c] = p1[c] - p0[c];
				r.e2[c] = p2[c] - p0[c];
			}
			r.triIndex = (uint32_t)k;
			r.meshIndex = localMeshIndex ? 0u : meshList[mi];
			r.pad = 0;
		}
	});
	if (invPerm) {
		// triangle -> its record (the first one of a cut triangle)
		invPerm->assign(n, UINT32_MAX);
		for (size_t i = 0; i < nRefs; ++i) {
			const size_t prim = cut ? refPrim[w.order[i]] : w.order[i];
			if ((*invPerm)[prim] == UINT32_MAX) (*invPerm)[prim] = (uint32_t)(primBase + i);
		}
	}
	if (w.maxDepth > *maxDepth) *maxDepth = w.maxDepth;
	return w.root;
}

Box meshBBox(const lrh_mesh &m) {  // TriangleMesh::GetBBox over all vertices
	Box b = Box::Empty();
	for (size_t i = 0; i < m.nVerts; ++i) b.growPoint(m.verts + 3 * i);
	return b;
}

struct ObjectData {
	std::vector<float> xf;                   // world->local, 16 per static instance
	std::vector<lrh_motion> mot;
	std::vector<lrh_motion_seg> segs;
	std::vector<float> times;
	std::vector<lrh_inst> inst;              // in object order (before root permutation)
	std::vector<Box> boxes;                  // epsilon-expanded world boxes
};

void prepareObjects(lrh_scene *s, const lrh_object *objs, unsigned n, ObjectData &od, bool motionBounds = true) {
	// motion bounds (1,025 samples per segment, each with a matrix inverse)
	// are computed after the loop, in parallel over the objects
	struct Deferred { size_t slot; lrh::MotionSystemHost ms; Box local; };
	std::vector<Deferred> deferred;
	for (unsigned o = 0; o < n; ++o) {
		const lrh_object &ob = objs[o];
		if (ob.kind != s->objKind[o] || ob.mesh != s->objMesh[o])
			throw Error(LRH_EINVAL, "object " + std::to_string(o) + ": kind/mesh differ from the scene");
		lrh_inst in;
		in.root = s->objLeafRoot[o];
		in.kind = ob.kind;
		in.index = 0;
		in.object = o;
		Box wb = s->meshBox[ob.mesh];
		if (ob.kind == 1) {
			if (!ob.m) throw Error(LRH_EINVAL, "instance without matrix");
			lrh::M4 m, inv;
			std::memcpy(m.m, ob.m, 64);
			if (ob.mInv) std::memcpy(inv.m, ob.mInv, 64);
			else if (!lrh::invert(m, &inv)) throw Error(LRH_EINVAL, "singular instance matrix");
			in.index = (uint32_t)(od.xf.size() / 16);
			od.xf.insert(od.xf.end(), &inv.m[0][0], &inv.m[0][0] + 16);
			wb = lrh::transformBox(m, wb);  // InstanceTriangleMesh::GetBBox
		} else if (ob.kind == 2) {
			if (!ob.nKeys || !ob.keyTimes || !ob.keyMats) throw Error(LRH_EINVAL, "motion object without keys");
			lrh::MotionSystemHost ms;
			ms.init(ob.nKeys, ob.keyTimes, ob.keyMats);
			lrh_motion md;
			md.firstTime = (uint32_t)od.times.size();
			md.nTimes = (uint32_t)ms.times.size();
			md.firstSeg = (uint32_t)od.segs.size();
			md.pad = 0;
			for (unsigned k = 0; k < LRH_MOTION_INLINE_TIMES; ++k)
				md.t[k] = k < ms.times.size() ? ms.times[k] : 0.f;
			od.times.insert(od.times.end(), ms.times.begin(), ms.times.end());
			od.segs.insert(od.segs.end(), ms.segs.begin(), ms.segs.end());
			in.index = (uint32_t)od.mot.size();
			od.mot.push_back(md);
			if (motionBounds) deferred.push_back(Deferred{od.boxes.size(), std::move(ms), wb});
			// filled below (or, without motionBounds, by the device refit)
			od.boxes.push_back(Box::Empty());
			od.inst.push_back(in);
			continue;
		} else if (ob.kind != 0)
			throw Error(LRH_EINVAL, "object kind must be 0, 1 or 2");
		const float e = lrh::machineEpsilon(wb);
		for (int k = 0; k < 3; ++k) { wb.lo[k] -= e; wb.hi[k] += e; }
		od.boxes.push_back(wb);
		od.inst.push_back(in);
	}
	if (deferred.empty()) return;
	const unsigned nt = std::max(1u, std::min<unsigned>(s->params.nThreads ? s->params.nThreads : lrh::hardwareThreads(),
			(unsigned)deferred.size()));
	std::atomic<size_t> next{0};
	std::exception_ptr failed;
	std::mutex failMu;
	std::vector<std::thread> th;
	for (unsigned t = 0; t < nt; ++t)
		th.emplace_back([&]() {
			try {
				for (size_t i; (i = next.fetch_add(1)) < deferred.size();) {
					Box wb = deferred[i].ms.bound(deferred[i].local);  // MotionTriangleMesh::GetBBox
					const float e = lrh::machineEpsilon(wb);
					for (int k = 0; k < 3; ++k) { wb.lo[k] -= e; wb.hi[k] += e; }
					od.boxes[deferred[i].slot] = wb;
				}
			} catch (...) {
				// a singular motion sample: stop the others and report it after the join
				next.store(deferred.size());
				std::lock_guard<std::mutex> g(failMu);
				if (!failed) failed = std::current_exception();
			}
		});
	for (auto &x : th) x.join();
	if (failed) std::rethrow_exception(failed);
}

// root tree over the objects, written into the reserved node slots [0, rootReserve)
void buildRoot(lrh_scene *s, ObjectData &od, HostTree &ht, std::vector<lrh_inst> &instOut, uint32_t *depth) {
	lrh::BuildParams rp = s->params;
	rp.maxLeaf = 1;
	lrh::WideBVH w;
	lrh::buildWideBVH(od.boxes, rp, 0, 0, &w);
	if (w.nodeCount() > s->rootReserve) throw Error(LRH_EINVAL, "root tree larger than its reservation");
	if (s->layout == 2) std::copy(w.q8nodes.begin(), w.q8nodes.end(), ht.q8.begin());
	else if (s->layout) std::copy(w.fnodes.begin(), w.fnodes.end(), ht.f.begin());
	else std::copy(w.qnodes.begin(), w.qnodes.end(), ht.q.begin());
	instOut.resize(od.inst.size());
	for (size_t i = 0; i < w.order.size(); ++i) instOut[i] = od.inst[w.order[i]];
	*depth = w.maxDepth;
	s->root = w.root;
	// refit topology: slot order and the root nodes grouped by depth
	s->rootOrder = w.order;
	s->objBoxes = od.boxes;
	const size_t nn = w.nodeCount();
	std::vector<uint32_t> depthOf(nn, 0), bfs(1, w.root);
	uint32_t maxD = 0;
	for (size_t q = 0; q < bfs.size(); ++q) {
		const uint32_t i = bfs[q];
		for (int c = 0; c < (s->layout == 2 ? 8 : 4); ++c) {
			const uint32_t r = s->layout == 2 ? w.q8nodes[i].child[c] : s->layout ? w.fnodes[i].child[c] : w.qnodes[i].child[c];
			if (r == LRH_CHILD_EMPTY || (r & LRH_CHILD_LEAF) || r >= nn) continue;
			depthOf[r] = depthOf[i] + 1;
			maxD = std::max(maxD, depthOf[r]);
			bfs.push_back(r);
		}
	}
	s->levelNodes.clear();
	s->levelStart.assign(1, 0);
	for (int d = (int)maxD; d >= 0; --d) {
		for (uint32_t i : bfs)
			if (depthOf[i] == (uint32_t)d) s->levelNodes.push_back(i);
		s->levelStart.push_back((uint32_t)s->levelNodes.size());
	}
}

void keepHost(lrh_scene *s, const HostTree &ht, const void *tris, size_t triBytes, const void *insts,
		size_t instBytes) {
	const uint8_t *n = (const uint8_t *)ht.data(s->layout);
	s->hostNodes.assign(n, n + ht.bytes(s->layout));
	s->hostPrims.resize(triBytes + instBytes);
	if (triBytes) std::memcpy(s->hostPrims.data(), tris, triBytes);
	if (instBytes) std::memcpy(s->hostPrims.data() + triBytes, insts, instBytes);
}

// uploads an edit's instance matrices and motion systems; true if one of the
// arrays had to grow (a new device address)
bool uploadEdit(lrh_scene *s, const ObjectData &od, hipStream_t st) {
	lrh_device *dev = s->dev;
	bool moved = false;
	auto put = [&](DevBuf &b, const void *src, size_t bytes) {
		if (bytes > b.bytes) {
			b.alloc(dev, bytes);
			moved = true;
		}
		b.upload(src, bytes, st);
	};
	put(s->xforms, od.xf.data(), od.xf.size() * 4);
	put(s->motions, od.mot.data(), od.mot.size() * sizeof(lrh_motion));
	put(s->segs, od.segs.data(), od.segs.size() * sizeof(lrh_motion_seg));
	put(s->times, od.times.data(), od.times.size() * 4);
	return moved;
}

void finishInfo(lrh_scene *s, size_t nNodes, size_t nPrims, size_t primBytes, double tb, double tu) {
	s->info.nodeCount = nNodes;
	s->info.nodeBytes = s->dev ? s->nodes.bytes : s->hostNodes.size();
	s->info.primCount = nPrims;
	if (!s->dev) primBytes = s->hostPrims.size();
	s->info.primBytes = primBytes;
	s->info.deviceBytes = s->nodes.bytes + s->tris.bytes + s->insts.bytes + s->xforms.bytes + s->motions.bytes +
			s->segs.bytes + s->times.bytes + s->invPerm.bytes + s->meshTriBase.bytes;
	s->info.maxDepth = s->maxDepth;
	s->info.layout = s->layout;
	s->info.buildSeconds = tb;
	s->info.uploadSeconds = tu;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char *lrh_last_error(void) { return g_lastError.c_str(); }
const char *lrh_version(void) { return "lrhip 0.1 (gfx950)"; }

int lrh_device_count(int *count) {
	if (!count) return fail(LRH_EINVAL, "null count");
	int n = 0;
	const hipError_t e = hipGetDeviceCount(&n);
	if (e != hipSuccess) {
		*count = 0;
		return fail(LRH_ENODEV, std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
	}
	*count = n;
	return LRH_OK;
}

int lrh_device_open(int ordinal, lrh_device **out) {
	if (!out) return fail(LRH_EINVAL, "null out");
	*out = nullptr;
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(LRH_ENODEV, "no HIP device");
	if (ordinal < 0 || ordinal >= n) return fail(LRH_EINVAL, "device ordinal out of range");
	std::unique_ptr<lrh_device> d(new lrh_device());
	const int rc = guarded([&]() {
		d->ordinal = ordinal;
		d->activate();
		hipDeviceProp_t prop;
		HIPCHK(hipGetDeviceProperties(&prop, ordinal));
		d->cuCount = prop.multiProcessorCount;
		d->totalMem = prop.totalGlobalMem;
		d->name = prop.gcnArchName;
		HIPCHK(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
		d->dError = (uint32_t *)d->devAlloc(sizeof(uint32_t));
		HIPCHK(hipMemsetAsync(d->dError, 0, sizeof(uint32_t), d->stream));
		HIPCHK(hipStreamSynchronize(d->stream));
		// the no-alloc scratch of this GPU (device current: its own copy); a
		// failure here only disables lrh_trace_noalloc, which reports it
		if (resolveScratch(ordinal, &d->scratch) != LRH_OK) d->scratchError = g_lastError;
		d->t0 = now();
	});
	if (rc != LRH_OK) return rc;
	*out = d.release();
	return LRH_OK;
}

int lrh_device_close(lrh_device *dev) {
	if (!dev) return fail(LRH_EINVAL, "null device");
	(void)hipSetDevice(dev->ordinal);
	if (dev->stream) (void)hipStreamSynchronize(dev->stream);
	(void)hipDeviceSynchronize();
	// scenes still alive are released with their device; their handles become invalid
	std::set<lrh_scene *> live;
	{
		std::lock_guard<std::mutex> g(dev->mu);
		live.swap(dev->scenes);
	}
	for (lrh_scene *s : live) delete s;
	for (auto &w : dev->ws) {
		dev->devFree(w.second.spill, w.second.spillBytes);
		dev->devFree(w.second.counter, 2 * sizeof(uint32_t));
		dev->devFree(w.second.stats, LRH_TRACE_STATS * sizeof(unsigned long long));
		dev->devFree(w.second.table, w.second.tableBytes);
	}
	dev->devFree(dev->dError, sizeof(uint32_t));
	dev->drainPool();
	for (auto &a : dev->allocs) (void)hipFree((void *)a.first);
	if (dev->stream) (void)hipStreamDestroy(dev->stream);
	delete dev;
	return LRH_OK;
}

int lrh_device_info(lrh_device *dev, char *name, size_t nameLen, int *cuCount, size_t *totalMem) {
	if (!dev) return fail(LRH_EINVAL, "null device");
	if (name && nameLen) {
		std::strncpy(name, dev->name.c_str(), nameLen - 1);
		name[nameLen - 1] = 0;
	}
	if (cuCount) *cuCount = dev->cuCount;
	if (totalMem) *totalMem = dev->totalMem;
	return LRH_OK;
}

int lrh_alloc(lrh_device *dev, size_t bytes, void **dptr) {
	if (!dev || !dptr) return fail(LRH_EINVAL, "null argument");
	return guarded([&]() {
		dev->activate();
		*dptr = dev->devAlloc(bytes ? bytes : 1);
		std::lock_guard<std::mutex> g(dev->mu);
		dev->allocs[(size_t)*dptr] = bytes ? bytes : 1;
	});
}

int lrh_free(lrh_device *dev, void *dptr) {
	if (!dev) return fail(LRH_EINVAL, "null device");
	if (!dptr) return LRH_OK;
	return guarded([&]() {
		size_t bytes = 0;
		{
			std::lock_guard<std::mutex> g(dev->mu);
			auto it = dev->allocs.find((size_t)dptr);
			if (it == dev->allocs.end()) throw Error(LRH_EINVAL, "pointer not allocated by lrh_alloc");
			bytes = it->second;
			dev->allocs.erase(it);
		}
		dev->activate();
		dev->devFree(dptr, bytes);
	});
}

static hipMemcpyKind toKind(int kind) {
	switch (kind) {
	case LRH_H2D: return hipMemcpyHostToDevice;
	case LRH_D2H: return hipMemcpyDeviceToHost;
	case LRH_D2D: return hipMemcpyDeviceToDevice;
	default: throw Error(LRH_EINVAL, "bad copy kind");
	}
}

int lrh_memcpy(lrh_device *dev, void *dst, const void *src, size_t bytes, int kind, void *stream) {
	if (!dev || (bytes && (!dst || !src))) return fail(LRH_EINVAL, "null argument");
	return guarded([&]() {
		dev->activate();
		hipStream_t s = stream ? (hipStream_t)stream : dev->stream;
		if (bytes) HIPCHK(hipMemcpyAsync(dst, src, bytes, toKind(kind), s));
		HIPCHK(hipStreamSynchronize(s));
	});
}

int lrh_memcpy_async(lrh_device *dev, void *dst, const void *src, size_t bytes, int kind, void *stream) {
	if (!dev || (bytes && (!dst || !src))) return fail(LRH_EINVAL, "null argument");
	return guarded([&]() {
		dev->activate();
		if (bytes) HIPCHK(hipMemcpyAsync(dst, src, bytes, toKind(kind), stream ? (hipStream_t)stream : dev->stream));
	});
}

int lrh_host_alloc(lrh_device *dev, size_t bytes, void **hptr) {
	if (!dev || !hptr) return fail(LRH_EINVAL, "null argument");
	return guarded([&]() {
		dev->activate();
		HIPCHK(hipHostMalloc(hptr, bytes ? bytes : 1, hipHostMallocDefault));
	});
}

int lrh_host_free(lrh_device *dev, void *hptr) {
	if (!dev) return fail(LRH_EINVAL, "null device");
	return guarded([&]() {
		if (hptr) HIPCHK(hipHostFree(hptr));
	});
}

int lrh_stream_create(lrh_device *dev, void **stream) {
	if (!dev || !stream) return fail(LRH_EINVAL, "null argument");
	return guarded([&]() {
		dev->activate();
		hipStream_t s;
		HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
		*stream = (void *)s;
	});
}

int lrh_stream_destroy(lrh_device *dev, void *stream) {
	if (!dev) return fail(LRH_EINVAL, "null device");
	return guarded([&]() {
		if (stream) HIPCHK(hipStreamDestroy((hipStream_t)stream));
	});
}

int lrh_sync(lrh_device *dev, void *stream) {
	if (!dev) return fail(LRH_EINVAL, "null device");
	return guarded([&]() {
		dev->activate();
		HIPCHK(hipStreamSynchronize(stream ? (hipStream_t)stream : dev->stream));
		uint32_t err = 0;
		HIPCHK(hipMemcpy(&err, dev->dError, sizeof(err), hipMemcpyDeviceToHost));
		if (err) {
			const uint32_t zero = 0;
			HIPCHK(hipMemcpy(dev->dError, &zero, sizeof(zero), hipMemcpyHostToDevice));
			throw Error(LRH_EDEVICE, (err & 1u) ? "traversal stack overflow" : "traversal step guard hit");
		}
	});
}

size_t lrh_used_memory(lrh_device *dev) { return dev ? dev->used.load() : 0; }

int lrh_event_create(lrh_device *dev, lrh_event **ev) {
	if (!dev || !ev) return fail(LRH_EINVAL, "null argument");
	return guarded([&]() {
		dev->activate();
		std::unique_ptr<lrh_event> e(new lrh_event());
		e->dev = dev;
		HIPCHK(hipEventCreate(&e->ev));
		*ev = e.release();
	});
}

int lrh_event_destroy(lrh_event *ev) {
	if (!ev) return LRH_OK;
	(void)hipEventDestroy(ev->ev);
	delete ev;
	return LRH_OK;
}

int lrh_event_record(lrh_event *ev, void *stream) {
	if (!ev) return fail(LRH_EINVAL, "null event");
	return guarded([&]() { HIPCHK(hipEventRecord(ev->ev, stream ? (hipStream_t)stream : ev->dev->stream)); });
}

int lrh_event_sync(lrh_event *ev) {
	if (!ev) return fail(LRH_EINVAL, "null event");
	return guarded([&]() { HIPCHK(hipEventSynchronize(ev->ev)); });
}

int lrh_event_elapsed(lrh_event *a, lrh_event *b, float *ms) {
	if (!a || !b || !ms) return fail(LRH_EINVAL, "null argument");
	return guarded([&]() { HIPCHK(hipEventElapsedTime(ms, a->ev, b->ev)); });
}

namespace {

void *gbAlloc(void *ctx, size_t bytes) { return static_cast<lrh_device *>(ctx)->buildAlloc(bytes); }
void gbRelease(void *ctx, void *p, size_t bytes) { static_cast<lrh_device *>(ctx)->buildRelease(p, bytes); }

// single-level scene built on the device (LRH_BUILD_DEVICE_LBVH / _PLOC):
// vertices and triangles are uploaded mesh by mesh, the tree and the records
// never leave HBM
void deviceBuild(lrh_scene *s, const lrh_mesh *meshes, unsigned nMeshes, unsigned builder) {
	lrh_device *dev = s->dev;
	dev->activate();
	const double t0 = now();
	std::vector<uint32_t> tbase(nMeshes + 1, 0), vbase(nMeshes + 1, 0);
	for (unsigned i = 0; i < nMeshes; ++i) {
		if ((uint64_t)tbase[i] + meshes[i].nTris > LRH_MAX_LEAF_START)
			throw Error(LRH_ERANGE, "scene exceeds 268,435,455 triangle records");
		if ((uint64_t)vbase[i] + meshes[i].nVerts > 0xffffffffull) throw Error(LRH_ERANGE, "too many vertices");
		tbase[i + 1] = tbase[i] + meshes[i].nTris;
		vbase[i + 1] = vbase[i] + meshes[i].nVerts;
	}
	s->nMeshes = nMeshes;
	s->meshTriBase.alloc(dev, tbase.size() * 4);
	s->meshTriBase.upload(tbase.data(), tbase.size() * 4, dev->stream);
	const uint32_t n = tbase.back();
	if (n == 0) {
		s->root = LRH_CHILD_EMPTY;
		s->nodes.alloc(dev, 64);
		s->tris.alloc(dev, 48);
		s->invPerm.alloc(dev, 4);
		HIPCHK(hipStreamSynchronize(dev->stream));
		finishInfo(s, 0, 0, 0, now() - t0, 0.0);
		return;
	}
	DevBuf verts, tris;
	verts.alloc(dev, (size_t)vbase.back() * 12 + 16);
	tris.alloc(dev, (size_t)n * 12 + 16);
	for (unsigned i = 0; i < nMeshes; ++i) {
		const lrh_mesh &m = meshes[i];
		if (m.nVerts)
			HIPCHK(hipMemcpyAsync((char *)verts.p + (size_t)vbase[i] * 12, m.verts, (size_t)m.nVerts * 12,
					hipMemcpyHostToDevice, dev->stream));
		if (m.nTris)
			HIPCHK(hipMemcpyAsync((char *)tris.p + (size_t)tbase[i] * 12, m.tris, (size_t)m.nTris * 12,
					hipMemcpyHostToDevice, dev->stream));
	}
	HIPCHK(hipStreamSynchronize(dev->stream));
	const double t1 = now();
	LrhGpuBuildIn in;
	in.verts = (const float *)verts.p;
	in.tris = (const uint32_t *)tris.p;
	in.meshTriBase = tbase.data();
	in.meshVertBase = vbase.data();
	in.nMeshes = nMeshes;
	in.nTris = n;
	in.maxLeaf = s->params.maxLeaf;
	in.layout = s->layout;
	in.method = builder == LRH_BUILD_DEVICE_PLOC ? 1u : 0u;
	in.travCost = s->params.travCost;
	LrhGpuAlloc a{gbAlloc, gbRelease, dev};
	LrhGpuBuildOut out;
	try {
		lrhGpuBuildBVH(in, a, dev->stream, &out);
	} catch (const std::runtime_error &e) {
		throw Error(LRH_EHIP, e.what());
	}
	s->nodes.adopt(dev, out.nodes, out.nodeBytes);
	s->tris.adopt(dev, out.tris, out.triBytes);
	s->invPerm.adopt(dev, out.invPerm, out.invPermBytes);
	s->root = out.root;
	s->maxDepth = out.maxDepth;
	finishInfo(s, out.nodeCount, n, out.triBytes, now() - t1, t1 - t0);
}

}  // namespace

int lrh_scene_create_bvh(lrh_device *dev, const lrh_mesh *meshes, unsigned int nMeshes,
		const lrh_build_opts *opts, lrh_scene **out) {
	if (!out || (nMeshes && !meshes)) return fail(LRH_EINVAL, "null argument");
	*out = nullptr;
	std::unique_ptr<lrh_scene> s(new lrh_scene());
	const int rc = guarded([&]() {
		for (unsigned i = 0; i < nMeshes; ++i) checkMesh(meshes[i], i);
		s->dev = dev;
		s->params = toParams(opts);
		s->layout = s->params.layout;
		s->twoLevel = false;
		const unsigned builder = opts ? opts->builder : LRH_BUILD_HOST_SAH;
		if (builder > LRH_BUILD_DEVICE_PLOC)
			throw Error(LRH_EINVAL,
					"builder must be LRH_BUILD_HOST_SAH, LRH_BUILD_DEVICE_LBVH or LRH_BUILD_DEVICE_PLOC");
		if (builder != LRH_BUILD_HOST_SAH) {
			if (!dev) throw Error(LRH_EINVAL, "the device builder needs a device");
			deviceBuild(s.get(), meshes, nMeshes, builder);
			return;
		}
		const double t0 = now();
		HostTree ht;
		std::vector<lrh_tri> tris;
		std::vector<unsigned> all(nMeshes);
		for (unsigned i = 0; i < nMeshes; ++i) all[i] = i;
		std::vector<uint32_t> inv;
		s->root = buildMeshTree(meshes, all, false, s->params, ht, tris, &s->maxDepth, &inv);
		std::vector<uint32_t> base(nMeshes + 1, 0);
		for (unsigned i = 0; i < nMeshes; ++i) base[i + 1] = base[i] + meshes[i].nTris;
		s->nMeshes = nMeshes;
		const double t1 = now();
		if (!dev) {
			keepHost(s.get(), ht, tris.data(), tris.size() * sizeof(lrh_tri), nullptr, 0);
			finishInfo(s.get(), ht.count(s->layout), tris.size(), tris.size() * sizeof(lrh_tri), t1 - t0, 0.0);
			return;
		}
		dev->activate();
		s->nodes.alloc(dev, std::max<size_t>(ht.bytes(s->layout), 64));
		s->nodes.upload(ht.data(s->layout), ht.bytes(s->layout), dev->stream);
		s->tris.alloc(dev, std::max<size_t>(tris.size() * sizeof(lrh_tri), 48));
		s->tris.upload(tris.data(), tris.size() * sizeof(lrh_tri), dev->stream);
		s->invPerm.alloc(dev, std::max<size_t>(inv.size() * 4, 4));
		s->invPerm.upload(inv.data(), inv.size() * 4, dev->stream);
		s->meshTriBase.alloc(dev, base.size() * 4);
		s->meshTriBase.upload(base.data(), base.size() * 4, dev->stream);
		HIPCHK(hipStreamSynchronize(dev->stream));
		finishInfo(s.get(), ht.count(s->layout), tris.size(), tris.size() * sizeof(lrh_tri), t1 - t0, now() - t1);
	});
	if (rc != LRH_OK) return rc;
	if (dev) {
		std::lock_guard<std::mutex> g(dev->mu);
		dev->scenes.insert(s.get());
	}
	*out = s.release();
	return LRH_OK;
}

int lrh_scene_create_mbvh(lrh_device *dev, const lrh_mesh *meshes, unsigned int nMeshes,
		const lrh_object *objects, unsigned int nObjects, const lrh_build_opts *opts, lrh_scene **out) {
	if (!out || (nMeshes && !meshes) || (nObjects && !objects)) return fail(LRH_EINVAL, "null argument");
	*out = nullptr;
	std::unique_ptr<lrh_scene> s(new lrh_scene());
	const int rc = guarded([&]() {
		for (unsigned i = 0; i < nMeshes; ++i) checkMesh(meshes[i], i);
		s->dev = dev;
		s->params = toParams(opts);
		s->layout = s->params.layout;
		s->twoLevel = true;
		const double t0 = now();
		HostTree ht;
		std::vector<lrh_tri> tris;
		s->rootReserve = std::max<size_t>(nObjects, 1);
		ht.reserve(s->rootReserve, s->layout);
		// one leaf tree per referenced base mesh (mbvhaccel.cpp:90-180)
		std::map<unsigned, uint32_t> leafRoot;
		s->objMesh.resize(nObjects);
		s->objKind.resize(nObjects);
		s->objLeafRoot.resize(nObjects);
		s->meshBox.resize(nMeshes);
		for (unsigned i = 0; i < nMeshes; ++i) s->meshBox[i] = meshBBox(meshes[i]);
		uint32_t depth = 0;
		for (unsigned o = 0; o < nObjects; ++o) {
			const unsigned m = objects[o].mesh;
			if (m >= nMeshes) throw Error(LRH_EINVAL, "object mesh index out of range");
			s->objMesh[o] = m;
			s->objKind[o] = objects[o].kind;
			auto it = leafRoot.find(m);
			if (it == leafRoot.end() || objects[o].kind == 0) {
				const std::vector<unsigned> one(1, m);
				const uint32_t r = buildMeshTree(meshes, one, true, s->params, ht, tris, &depth);
				leafRoot[m] = r;
				s->objLeafRoot[o] = r;
			} else
				s->objLeafRoot[o] = it->second;
		}
		ObjectData od;
		std::vector<lrh_inst> instOut;
		uint32_t rootDepth = 0;
		if (nObjects) {
			prepareObjects(s.get(), objects, nObjects, od);
			buildRoot(s.get(), od, ht, instOut, &rootDepth);
		} else
			s->root = LRH_CHILD_EMPTY;
		s->leafDepth = depth;
		s->maxDepth = rootDepth + depth + 1;
		const double t1 = now();
		if (!dev) {
			keepHost(s.get(), ht, tris.data(), tris.size() * sizeof(lrh_tri), instOut.data(),
					instOut.size() * sizeof(lrh_inst));
			finishInfo(s.get(), ht.count(s->layout), tris.size() + instOut.size(),
					tris.size() * sizeof(lrh_tri) + instOut.size() * sizeof(lrh_inst), t1 - t0, 0.0);
			return;
		}
		dev->activate();
		s->nodes.alloc(dev, std::max<size_t>(ht.bytes(s->layout), 64));
		s->nodes.upload(ht.data(s->layout), ht.bytes(s->layout), dev->stream);
		s->tris.alloc(dev, std::max<size_t>(tris.size() * sizeof(lrh_tri), 48));
		s->tris.upload(tris.data(), tris.size() * sizeof(lrh_tri), dev->stream);
		s->insts.alloc(dev, std::max<size_t>(instOut.size() * sizeof(lrh_inst), 16));
		s->insts.upload(instOut.data(), instOut.size() * sizeof(lrh_inst), dev->stream);
		s->xforms.alloc(dev, std::max<size_t>(od.xf.size() * 4, 64));
		s->xforms.upload(od.xf.data(), od.xf.size() * 4, dev->stream);
		s->motions.alloc(dev, std::max<size_t>(od.mot.size() * sizeof(lrh_motion), sizeof(lrh_motion)));
		s->motions.upload(od.mot.data(), od.mot.size() * sizeof(lrh_motion), dev->stream);
		s->segs.alloc(dev, std::max<size_t>(od.segs.size() * sizeof(lrh_motion_seg), sizeof(lrh_motion_seg)));
		s->segs.upload(od.segs.data(), od.segs.size() * sizeof(lrh_motion_seg), dev->stream);
		s->times.alloc(dev, std::max<size_t>(od.times.size() * 4, 16));
		s->times.upload(od.times.data(), od.times.size() * 4, dev->stream);
		HIPCHK(hipStreamSynchronize(dev->stream));
		finishInfo(s.get(), ht.count(s->layout), tris.size() + instOut.size(),
				tris.size() * sizeof(lrh_tri) + instOut.size() * sizeof(lrh_inst), t1 - t0, now() - t1);
	});
	if (rc != LRH_OK) return rc;
	if (dev) {
		std::lock_guard<std::mutex> g(dev->mu);
		dev->scenes.insert(s.get());
	}
	*out = s.release();
	return LRH_OK;
}

int lrh_scene_update(lrh_scene *s, const lrh_object *objects, unsigned int nObjects) {
	if (!s || (nObjects && !objects)) return fail(LRH_EINVAL, "null argument");
	if (!s->twoLevel) return fail(LRH_EINVAL, "only two-level scenes support update (Accelerator::DoesSupportUpdate)");
	if (nObjects != s->objMesh.size()) return fail(LRH_EINVAL, "object count differs from the scene");
	if (!s->dev) return fail(LRH_EINVAL, "host-only scene");
	return guarded([&]() {
		if (!nObjects) return;
		ObjectData od;
		prepareObjects(s, objects, nObjects, od);
		HostTree ht;
		ht.reserve(s->rootReserve, s->layout);
		std::vector<lrh_inst> instOut;
		uint32_t rootDepth = 0;
		buildRoot(s, od, ht, instOut, &rootDepth);
		lrh_device *dev = s->dev;
		dev->activate();
		HIPCHK(hipStreamSynchronize(dev->stream));
		// root nodes live in the reserved prefix of the node array
		const size_t nb = s->rootReserve * (s->layout == 2 ? sizeof(lrh_node8q) : s->layout ? sizeof(lrh_node4f) : sizeof(lrh_node4q));
		HIPCHK(hipMemcpyAsync(s->nodes.p, ht.data(s->layout), nb, hipMemcpyHostToDevice, dev->stream));
		HIPCHK(hipMemcpyAsync(s->insts.p, instOut.data(), instOut.size() * sizeof(lrh_inst), hipMemcpyHostToDevice,
				dev->stream));
		const bool moved = uploadEdit(s, od, dev->stream);
		HIPCHK(hipStreamSynchronize(dev->stream));
		const uint32_t depth = rootDepth + s->leafDepth + 1;
		if (moved || depth > s->maxDepth) ++s->generation;
		s->maxDepth = depth;
		s->info.maxDepth = s->maxDepth;
	});
}

int lrh_scene_refit(lrh_scene *s, const lrh_object *objects, unsigned int nObjects) {
	if (!s || (nObjects && !objects)) return fail(LRH_EINVAL, "null argument");
	if (!s->twoLevel) return fail(LRH_EINVAL, "only two-level scenes can be refitted");
	if (nObjects != s->objMesh.size()) return fail(LRH_EINVAL, "object count differs from the scene");
	if (!s->dev) return fail(LRH_EINVAL, "host-only scene");
	return guarded([&]() {
		if (!nObjects) return;
		ObjectData od;
		prepareObjects(s, objects, nObjects, od, false);
		lrh_device *dev = s->dev;
		dev->activate();
		const hipStream_t st = dev->stream;
		HIPCHK(hipStreamSynchronize(st));
		// leaf-slot boxes of plain and instance objects, the list of motion objects
		std::vector<float> slot(6 * (size_t)nObjects, 0.f);
		std::vector<LrhRefitMotion> mo;
		std::vector<uint32_t> sampleBase(1, 0);
		for (size_t j = 0; j < s->rootOrder.size(); ++j) {
			const uint32_t o = s->rootOrder[j];
			if (s->objKind[o] == 2) {
				const uint32_t mi = od.inst[o].index;
				mo.push_back({(uint32_t)j, mi, s->objMesh[o], 0});
				const uint64_t next = (uint64_t)sampleBase.back() + (uint64_t)(od.mot[mi].nTimes + 1) * 1025u;
				if (next > 0xffffffffull) throw Error(LRH_ERANGE, "too many motion samples for one refit");
				sampleBase.push_back((uint32_t)next);
			} else {
				for (int k = 0; k < 3; ++k) {
					slot[6 * j + k] = od.boxes[o].lo[k];
					slot[6 * j + 3 + k] = od.boxes[o].hi[k];
				}
			}
		}
		auto ensure = [&](DevBuf &b, size_t bytes) {
			if (b.bytes < bytes) b.alloc(dev, bytes);
		};
		// refit inputs; the edit's motion systems go to staging buffers first,
		// so a rejected edit (a singular motion sample) leaves the scene as it was
		ensure(s->rSlotBoxes, slot.size() * 4);
		s->rSlotBoxes.upload(slot.data(), slot.size() * 4, st);
		ensure(s->rLevelNodes, std::max<size_t>(s->levelNodes.size(), 1) * 4);
		s->rLevelNodes.upload(s->levelNodes.data(), s->levelNodes.size() * 4, st);
		const size_t nRoot = s->levelNodes.empty() ? 1 : *std::max_element(s->levelNodes.begin(), s->levelNodes.end()) + 1;
		ensure(s->rNodeBoxes, 6 * 4 * nRoot);
		ensure(s->rError, 4);
		LrhRefitArgs a;
		if (!mo.empty()) {
			std::vector<float> mb(6 * s->meshBox.size());
			for (size_t m = 0; m < s->meshBox.size(); ++m)
				for (int k = 0; k < 3; ++k) {
					mb[6 * m + k] = s->meshBox[m].lo[k];
					mb[6 * m + 3 + k] = s->meshBox[m].hi[k];
				}
			ensure(s->rMeshBoxes, mb.size() * 4);
			s->rMeshBoxes.upload(mb.data(), mb.size() * 4, st);
			ensure(s->rMotionObjs, mo.size() * sizeof(LrhRefitMotion));
			s->rMotionObjs.upload(mo.data(), mo.size() * sizeof(LrhRefitMotion), st);
			ensure(s->rSampleBase, sampleBase.size() * 4);
			s->rSampleBase.upload(sampleBase.data(), sampleBase.size() * 4, st);
			ensure(s->rKeys, 6 * 4 * mo.size());
			ensure(s->rMotions, od.mot.size() * sizeof(lrh_motion));
			s->rMotions.upload(od.mot.data(), od.mot.size() * sizeof(lrh_motion), st);
			ensure(s->rSegs, od.segs.size() * sizeof(lrh_motion_seg));
			s->rSegs.upload(od.segs.data(), od.segs.size() * sizeof(lrh_motion_seg), st);
			a.motionObjs = (const LrhRefitMotion *)s->rMotionObjs.p;
			a.nMotion = (uint32_t)mo.size();
			a.sampleBase = (const uint32_t *)s->rSampleBase.p;
			a.nSamples = sampleBase.back();
			a.motions = (const lrh_motion *)s->rMotions.p;
			a.segs = (const lrh_motion_seg *)s->rSegs.p;
			a.meshBoxes = (const float *)s->rMeshBoxes.p;
			a.keys = (uint32_t *)s->rKeys.p;
			a.error = (uint32_t *)s->rError.p;
		}
		a.slotBoxes = (float *)s->rSlotBoxes.p;
		a.nodeBoxes = (float *)s->rNodeBoxes.p;
		a.levelNodes = (const uint32_t *)s->rLevelNodes.p;
		a.levelStart = s->levelStart.data();
		a.nLevels = (uint32_t)s->levelStart.size() - 1;
		a.layout = s->layout;
		a.nodes = s->nodes.p;
		HIPCHK(lrhLaunchRefitMotion(a, st));
		if (!mo.empty()) {
			uint32_t err = 0;
			HIPCHK(hipMemcpyAsync(&err, s->rError.p, 4, hipMemcpyDeviceToHost, st));
			HIPCHK(hipStreamSynchronize(st));
			if (err) throw Error(LRH_EINVAL, "singular motion sample");
		}
		// accepted: the edit's matrices and motion systems, as lrh_scene_update uploads them
		if (uploadEdit(s, od, st)) ++s->generation;
		HIPCHK(lrhLaunchRefitLevels(a, st));
		// the accepted object boxes (a rejected edit never gets here, so the
		// scene's boxes stay those of the last accepted edit)
		std::vector<float> out(6 * (size_t)nObjects);
		HIPCHK(hipMemcpyAsync(out.data(), s->rSlotBoxes.p, out.size() * 4, hipMemcpyDeviceToHost, st));
		HIPCHK(hipStreamSynchronize(st));
		for (size_t j = 0; j < s->rootOrder.size(); ++j) {
			Box &b = s->objBoxes[s->rootOrder[j]];
			for (int k = 0; k < 3; ++k) {
				b.lo[k] = out[6 * j + k];
				b.hi[k] = out[6 * j + 3 + k];
			}
		}
	});
}

int lrh_scene_object_boxes(lrh_scene *s, float *boxes, unsigned int nObjects) {
	if (!s || (nObjects && !boxes)) return fail(LRH_EINVAL, "null argument");
	if (!s->twoLevel) return fail(LRH_EINVAL, "object boxes exist for two-level scenes only");
	if (nObjects != s->objMesh.size()) return fail(LRH_EINVAL, "object count differs from the scene");
	return guarded([&]() {
		for (unsigned o = 0; o < nObjects; ++o)
			for (int k = 0; k < 3; ++k) {
				boxes[6 * o + k] = s->objBoxes[o].lo[k];
				boxes[6 * o + 3 + k] = s->objBoxes[o].hi[k];
			}
	});
}

int lrh_scene_destroy(lrh_scene *s) {
	if (!s) return LRH_OK;
	if (s->dev) {
		(void)hipSetDevice(s->dev->ordinal);
		(void)hipStreamSynchronize(s->dev->stream);
		std::lock_guard<std::mutex> g(s->dev->mu);
		s->dev->scenes.erase(s);
	}
	delete s;
	return LRH_OK;
}

int lrh_scene_get_info(lrh_scene *s, lrh_scene_info *info) {
	if (!s || !info) return fail(LRH_EINVAL, "null argument");
	*info = s->info;
	return LRH_OK;
}

int lrh_scene_download(lrh_scene *s, void *nodes, size_t nodeBytes, void *prims, size_t primBytes) {
	if (!s) return fail(LRH_EINVAL, "null scene");
	return guarded([&]() {
		if (!s->dev) {
			if (nodes) std::memcpy(nodes, s->hostNodes.data(), std::min(nodeBytes, s->hostNodes.size()));
			if (prims) std::memcpy(prims, s->hostPrims.data(), std::min(primBytes, s->hostPrims.size()));
			return;
		}
		s->dev->activate();
		HIPCHK(hipStreamSynchronize(s->dev->stream));
		if (nodes) HIPCHK(hipMemcpy(nodes, s->nodes.p, std::min(nodeBytes, s->nodes.bytes), hipMemcpyDeviceToHost));
		if (prims) HIPCHK(hipMemcpy(prims, s->tris.p, std::min(primBytes, s->tris.bytes), hipMemcpyDeviceToHost));
	});
}

}  // extern "C"

// ---------------------------------------------------------------------------
// tracing
// ---------------------------------------------------------------------------
namespace {

struct LaunchShape {
	unsigned block, lds, blocks, refill, spillDepth, leafTris;
	size_t ldsBytes;
};

unsigned pow2floor(unsigned v) {
	unsigned p = 1;
	while (p * 2 <= v) p *= 2;
	return p;
}

// launch shape of one persistent trace: block size, LDS stack depth, grid
// from the occupancy of the kernel variant, and the spill depth the tree
// depth needs (3 pushes per level plus the object sentinel)
LaunchShape chooseShape(lrh_device *dev, unsigned layout, bool two, bool any, bool batch, unsigned count,
		unsigned maxDepth, size_t extraLds, const lrh_trace_opts *o) {
	LaunchShape L;
	L.block = (o && o->blockSize) ? o->blockSize : 256;
	if (L.block != 64 && L.block != 128 && L.block != 256) throw Error(LRH_EINVAL, "blockSize must be 64, 128 or 256");
	L.lds = (o && o->ldsStack) ? o->ldsStack : 8;
	if (L.lds < 4 || L.lds > 64 || pow2floor(L.lds) != L.lds)
		throw Error(LRH_EINVAL, "ldsStack must be a power of two in 4..64");
	L.refill = (o && o->refillThreshold) ? std::min(64u, o->refillThreshold) : 8;
	L.leafTris = (o && o->leafTris) ? o->leafTris : 8u;
	if (L.leafTris > 8) throw Error(LRH_EINVAL, "leafTris must be 0..8");
	L.ldsBytes = (size_t)L.lds * L.block * sizeof(uint2) + extraLds;
	if (L.ldsBytes > 65536) throw Error(LRH_ERANGE, "LDS stack, node cache and item table exceed 64 KiB per block");
	const unsigned pct = (o && o->gridPercent) ? o->gridPercent : 100u;
	if (pct > 100) throw Error(LRH_EINVAL, "gridPercent must be 0..100");
	int perCU = (o && o->blocksPerCU) ? (int)o->blocksPerCU : 0;
	if (!perCU) {
		const uint64_t key = ((uint64_t)layout << 40) | ((uint64_t)two << 39) | ((uint64_t)any << 38) |
				((uint64_t)batch << 37) | ((uint64_t)L.ldsBytes << 8) | (L.block >> 6);
		std::lock_guard<std::mutex> g(dev->mu);
		auto it = dev->occ.find(key);
		if (it == dev->occ.end()) {
			int nb = 0;
			HIPCHK(lrhOccupancy(layout, two, any, batch, L.block, L.ldsBytes, &nb));
			dev->occ[key] = nb > 0 ? nb : 1;
			perCU = nb > 0 ? nb : 1;
		} else
			perCU = it->second;
		perCU = std::max(1, (int)(((unsigned)perCU * pct + 50u) / 100u));
	}
	const unsigned persistent = (unsigned)perCU * (unsigned)dev->cuCount;
	const unsigned need = (unsigned)(((uint64_t)count + L.block - 1) / L.block);
	L.blocks = std::max(1u, std::min(persistent, need));
	// a visit pushes at most width - 1 entries
	const unsigned worst = (layout == 2 ? 7u : 3u) * (maxDepth + 2u) + 4u;
	L.spillDepth = worst > L.lds ? worst - L.lds : 1u;
	return L;
}

lrh_device::Workspace &workspace(lrh_device *dev, hipStream_t st, size_t spillBytes) {
	std::lock_guard<std::mutex> g(dev->mu);
	lrh_device::Workspace &w = dev->ws[st];
	if (!w.counter) {
		w.counter = (uint32_t *)dev->devAlloc(2 * sizeof(uint32_t));
		w.stats = (unsigned long long *)dev->devAlloc(LRH_TRACE_STATS * sizeof(unsigned long long));
		HIPCHK(hipMemsetAsync(w.counter, 0, 2 * sizeof(uint32_t), st));
	}
	if (w.spillBytes < spillBytes) {
		// the previous spill array may still be read by work queued on st
		HIPCHK(hipStreamSynchronize(st));
		dev->devFree(w.spill, w.spillBytes);
		w.spill = nullptr;
		w.spillBytes = 0;
		w.spill = (uint2 *)dev->devAlloc(spillBytes);
		w.spillBytes = spillBytes;
	}
	return w;
}

void fillSceneParams(LrhTraceParams &P, const lrh_scene *s) {
	P.nodes = s->nodes.p;
	P.tris = (const lrh_tri *)s->tris.p;
	P.insts = (const lrh_inst *)s->insts.p;
	P.xforms = (const float *)s->xforms.p;
	P.motions = (const lrh_motion *)s->motions.p;
	P.segs = (const lrh_motion_seg *)s->segs.p;
	P.times = (const float *)s->times.p;
	P.root = s->root;
}

// argument checks shared by every single-scene trace entry point
int checkTrace(const lrh_scene *s, const void *rays, const void *hits, unsigned int count, unsigned int query) {
	if (!s) return fail(LRH_EINVAL, "null scene");
	if (count && (!rays || !hits)) return fail(LRH_EINVAL, "null ray or hit buffer");
	if (query > 1) return fail(LRH_EINVAL, "query must be LRH_QUERY_CLOSEST or LRH_QUERY_ANY");
	if (count > 0x7fffffffu) return fail(LRH_ERANGE, "at most 2^31-1 rays per trace");
	if (!s->dev) return fail(LRH_EINVAL, "host-only scene cannot be traced");
	if (((uintptr_t)rays & 15u) || ((uintptr_t)hits & 3u))
		return fail(LRH_EINVAL, "ray buffer must be 16-byte aligned, hit buffer 4-byte aligned");
	return LRH_OK;
}

int traceImpl(lrh_scene *s, const void *rays, void *hits, unsigned int count, unsigned int query,
		const lrh_trace_opts *opts, void *stream, unsigned long long *statsOut) {
	const int rc = checkTrace(s, rays, hits, count, query);
	if (rc != LRH_OK) return rc;
	if (count == 0) return LRH_OK;
	lrh_device *dev = s->dev;
	return guarded([&]() {
		dev->activate();
		hipStream_t st = stream ? (hipStream_t)stream : dev->stream;
		const bool any = query == LRH_QUERY_ANY;
		const LaunchShape L = chooseShape(dev, s->layout, s->twoLevel, any, false, count, s->maxDepth, 0, opts);
		const unsigned nLanes = L.blocks * L.block;
		lrh_device::Workspace &w = workspace(dev, st, (size_t)nLanes * L.spillDepth * sizeof(uint2));
		LrhTraceParams P;
		std::memset(&P, 0, sizeof(P));
		fillSceneParams(P, s);
		P.rays = (const lrh_ray_rec *)rays;
		P.hits = (lrh_hit_rec *)hits;
		P.nRays = count;
		P.counter = w.counter;
		P.spill = w.spill;
		P.spillDepth = L.spillDepth;
		P.nLanes = nLanes;
		P.ldsDepth = L.lds;
		P.refill = L.refill;
		P.leafTris = L.leafTris;
		P.error = dev->dError;
		P.stats = w.stats;
		if (statsOut) HIPCHK(hipMemsetAsync(w.stats, 0, LRH_TRACE_STATS * sizeof(unsigned long long), st));
		HIPCHK(lrhLaunchTrace(s->layout, s->twoLevel, any, statsOut != nullptr, false, dim3(L.blocks), dim3(L.block),
				L.ldsBytes, st, P));
		dev->rays += count;
		if (statsOut) {
			HIPCHK(hipMemcpyAsync(statsOut, w.stats, LRH_TRACE_STATS * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
			HIPCHK(hipStreamSynchronize(st));
		}
	});
}

// descriptor table of a batched launch, validated against its device
struct BatchTable {
	std::vector<LrhBatchDesc> descs;
	std::vector<uint32_t> starts;
	unsigned layout = 0;
	bool two = false;
	unsigned maxDepth = 0;
	uint64_t total = 0;
	// the distinct scenes of the batch with their generation at creation
	struct SceneStamp {
		const lrh_scene *scene;
		uint64_t serial, generation;
	};
	std::vector<SceneStamp> sceneGen;
};

BatchTable makeTable(lrh_device *dev, const lrh_batch_item *items, unsigned n) {
	BatchTable T;
	if (n == 0) throw Error(LRH_EINVAL, "empty batch");
	if (n > 4096) throw Error(LRH_ERANGE, "at most 4096 items per batch");
	T.descs.resize(n);
	T.starts.resize(n + 1);
	for (unsigned i = 0; i < n; ++i)
		if (!items[i].scene || items[i].scene->dev != dev)
			throw Error(LRH_EINVAL, "batch item " + std::to_string(i) + ": scene of another device");
	// ticket order: closest-hit items before any-hit ones, larger scenes first
	// and the items of one scene together, so the costliest rays start first
	// (the launch ends on short rays) and the grid works on one tree at a time
	std::vector<unsigned> order(n);
	for (unsigned i = 0; i < n; ++i) order[i] = i;
	std::stable_sort(order.begin(), order.end(), [items](unsigned a, unsigned b) {
		const lrh_batch_item &x = items[a], &y = items[b];
		if ((x.query == LRH_QUERY_ANY) != (y.query == LRH_QUERY_ANY)) return y.query == LRH_QUERY_ANY;
		if (x.scene->info.primCount != y.scene->info.primCount) return x.scene->info.primCount > y.scene->info.primCount;
		return x.scene->serial < y.scene->serial;
	});
	for (unsigned j = 0; j < n; ++j) {
		const unsigned i = order[j];
		const lrh_batch_item &it = items[i];
		const lrh_scene *s = it.scene;
		if (it.count && (!it.rays || !it.hits)) throw Error(LRH_EINVAL, "batch item " + std::to_string(i) + ": null buffer");
		if (((uintptr_t)it.rays & 15u) || ((uintptr_t)it.hits & 3u))
			throw Error(LRH_EINVAL, "batch item " + std::to_string(i) + ": misaligned buffer");
		if (it.query > 1) throw Error(LRH_EINVAL, "batch item " + std::to_string(i) + ": bad query");
		if (j == 0) T.layout = s->layout;
		else if (s->layout != T.layout) throw Error(LRH_EINVAL, "all scenes of a batch must share the node layout");
		T.two = T.two || s->twoLevel;
		T.maxDepth = std::max(T.maxDepth, s->maxDepth);
		bool seen = false;
		for (const auto &sg : T.sceneGen) seen = seen || sg.scene == s;
		if (!seen) T.sceneGen.push_back({s, s->serial, s->generation});
		LrhBatchDesc &d = T.descs[j];
		std::memset(&d, 0, sizeof(d));
		d.nodes = s->nodes.p;
		d.tris = (const lrh_tri *)s->tris.p;
		d.insts = (const lrh_inst *)s->insts.p;
		d.xforms = (const float *)s->xforms.p;
		d.motions = (const lrh_motion *)s->motions.p;
		d.segs = (const lrh_motion_seg *)s->segs.p;
		d.times = (const float *)s->times.p;
		d.rays = (const lrh_ray_rec *)it.rays;
		d.hits = (lrh_hit_rec *)it.hits;
		d.root = s->root;
		d.count = it.count;
		d.flags = (it.query == LRH_QUERY_ANY ? LRH_ITEM_ANY : 0u) | (s->twoLevel ? LRH_ITEM_TWO : 0u);
		T.starts[j] = (uint32_t)T.total;
		T.total += it.count;
		if (T.total > 0x7fffffffull) throw Error(LRH_ERANGE, "at most 2^31-1 rays per batch");
	}
	T.starts[n] = (uint32_t)T.total;
	return T;
}

}  // namespace

struct lrh_batch {
	lrh_device *dev = nullptr;
	BatchTable T;
	LaunchShape L{};
	DevBuf descs, starts, counter, spill;
	LrhTraceParams P{};
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
};

namespace {
// a resident batch holds its scenes' device addresses and spill depth: it may
// launch only while every scene is alive and unchanged in those respects
void checkBatchScenes(const lrh_batch *b) {
	std::lock_guard<std::mutex> g(b->dev->mu);
	for (const auto &sg : b->T.sceneGen) {
		if (!b->dev->scenes.count(const_cast<lrh_scene *>(sg.scene)) || sg.scene->serial != sg.serial)
			throw Error(LRH_EINVAL, "a scene of the batch was destroyed: recreate the batch");
		if (sg.scene->generation != sg.generation)
			throw Error(LRH_EINVAL,
					"a scene of the batch changed since the batch was created (an edit moved its device arrays "
					"or deepened its tree): recreate the batch");
	}
}
}  // namespace

extern "C" {

int lrh_trace_ex(lrh_scene *s, const void *rays, void *hits, unsigned int count, unsigned int query,
		const lrh_trace_opts *opts, void *stream) {
	return traceImpl(s, rays, hits, count, query, opts, stream, nullptr);
}

int lrh_trace(lrh_scene *s, const void *rays, void *hits, unsigned int count, unsigned int query, void *stream) {
	return traceImpl(s, rays, hits, count, query, nullptr, stream, nullptr);
}

int lrh_trace_stats(lrh_scene *s, const void *rays, void *hits, unsigned int count, unsigned int query,
		const lrh_trace_opts *opts, void *stream, unsigned long long out[LRH_TRACE_STATS]) {
	if (!out) return fail(LRH_EINVAL, "null stats output");
	for (int i = 0; i < LRH_TRACE_STATS; ++i) out[i] = 0;
	return traceImpl(s, rays, hits, count, query, opts, stream, out);
}

int lrh_trace_noalloc(lrh_scene *s, const void *rays, void *hits, unsigned int count, unsigned int query,
		void *stream) {
	const int rc = checkTrace(s, rays, hits, count, query);
	if (rc != LRH_OK) return rc;
	if (count == 0) return LRH_OK;
	lrh_device *dev = s->dev;
	if (!dev->scratch) return fail(LRH_EHIP, "the no-alloc scratch is unavailable on this device: " + dev->scratchError);
	return guarded([&]() {
		dev->activate();
		hipStream_t st = stream ? (hipStream_t)stream : dev->stream;
		const bool any = query == LRH_QUERY_ANY;
		LaunchShape L = chooseShape(dev, s->layout, s->twoLevel, any, false, count, s->maxDepth, 0, nullptr);
		// the persistent grid is capped at the lanes whose spill stacks fit
		const size_t blockBytes = (size_t)L.spillDepth * sizeof(uint2) * L.block;
		const size_t maxBlocks = (LRH_NOALLOC_SCRATCH_BYTES - kScratchCounterBytes) / blockBytes;
		if (maxBlocks == 0)
			throw Error(LRH_ERANGE, "tree depth " + std::to_string(s->maxDepth) + " needs " +
					std::to_string(blockBytes) + " bytes of spill stack per block, beyond the " +
					std::to_string((unsigned long long)LRH_NOALLOC_SCRATCH_BYTES) + "-byte no-alloc scratch");
		L.blocks = (unsigned)std::min<size_t>(L.blocks, maxBlocks);
		uint32_t *counter = static_cast<uint32_t *>(dev->scratch);
		uint2 *spill = reinterpret_cast<uint2 *>(static_cast<unsigned char *>(dev->scratch) + kScratchCounterBytes);
		// the ticket counters start at zero whatever the last call left (a fill
		// kernel, not a memset node, so the call may be graph-captured)
		HIPCHK(lrhLaunchReset(counter, 2, st));
		LrhTraceParams P;
		std::memset(&P, 0, sizeof(P));
		fillSceneParams(P, s);
		P.rays = (const lrh_ray_rec *)rays;
		P.hits = (lrh_hit_rec *)hits;
		P.nRays = count;
		P.counter = counter;
		P.spill = spill;
		P.spillDepth = L.spillDepth;
		P.nLanes = L.blocks * L.block;
		P.ldsDepth = L.lds;
		P.refill = L.refill;
		P.leafTris = L.leafTris;
		P.error = dev->dError;
		P.stats = nullptr;
		HIPCHK(lrhLaunchTrace(s->layout, s->twoLevel, any, false, false, dim3(L.blocks), dim3(L.block), L.ldsBytes, st,
				P));
		dev->rays += count;
	});
}

int lrh_batch_create(lrh_device *dev, const lrh_batch_item *items, unsigned int nItems, const lrh_trace_opts *opts,
		lrh_batch **out) {
	if (!dev || !out || (nItems && !items)) return fail(LRH_EINVAL, "null argument");
	*out = nullptr;
	std::unique_ptr<lrh_batch> b(new lrh_batch());
	const int rc = guarded([&]() {
		dev->activate();
		b->dev = dev;
		b->T = makeTable(dev, items, nItems);
		const size_t extra = ((nItems + 1) * sizeof(uint32_t) + 15) & ~(size_t)15;
		b->L = chooseShape(dev, b->T.layout, b->T.two, false, true, (unsigned)std::max<uint64_t>(b->T.total, 1),
				b->T.maxDepth, extra, opts);
		const unsigned nLanes = b->L.blocks * b->L.block;
		b->descs.alloc(dev, b->T.descs.size() * sizeof(LrhBatchDesc));
		b->starts.alloc(dev, b->T.starts.size() * sizeof(uint32_t));
		b->counter.alloc(dev, 2 * sizeof(uint32_t));
		b->spill.alloc(dev, (size_t)nLanes * b->L.spillDepth * sizeof(uint2));
		b->descs.upload(b->T.descs.data(), b->descs.bytes, dev->stream);
		b->starts.upload(b->T.starts.data(), b->starts.bytes, dev->stream);
		HIPCHK(hipMemsetAsync(b->counter.p, 0, 2 * sizeof(uint32_t), dev->stream));
		HIPCHK(hipStreamSynchronize(dev->stream));
		LrhTraceParams &P = b->P;
		std::memset(&P, 0, sizeof(P));
		P.nRays = (uint32_t)b->T.total;
		P.root = LRH_CHILD_EMPTY;
		P.counter = (uint32_t *)b->counter.p;
		P.spill = (uint2 *)b->spill.p;
		P.spillDepth = b->L.spillDepth;
		P.nLanes = nLanes;
		P.ldsDepth = b->L.lds;
		P.refill = b->L.refill;
		P.leafTris = b->L.leafTris;
		P.error = dev->dError;
		P.items = (const LrhBatchDesc *)b->descs.p;
		P.itemStart = (const uint32_t *)b->starts.p;
		P.nItems = nItems;
	});
	if (rc != LRH_OK) return rc;
	*out = b.release();
	return LRH_OK;
}

int lrh_batch_launch(lrh_batch *b, void *stream) {
	if (!b) return fail(LRH_EINVAL, "null batch");
	return guarded([&]() {
		b->dev->activate();
		if (b->T.total == 0) return;
		checkBatchScenes(b);
		hipStream_t st = stream ? (hipStream_t)stream : b->dev->stream;
		HIPCHK(lrhLaunchTrace(b->T.layout, b->T.two, false, false, true, dim3(b->L.blocks), dim3(b->L.block),
				b->L.ldsBytes, st, b->P));
		b->dev->rays += b->T.total;
	});
}

int lrh_batch_capture(lrh_batch *b, void *stream) {
	if (!b) return fail(LRH_EINVAL, "null batch");
	return guarded([&]() {
		b->dev->activate();
		checkBatchScenes(b);
		hipStream_t st = stream ? (hipStream_t)stream : b->dev->stream;
		if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
		if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
		HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
		hipError_t le = hipSuccess;
		if (b->T.total)
			le = lrhLaunchTrace(b->T.layout, b->T.two, false, false, true, dim3(b->L.blocks), dim3(b->L.block),
					b->L.ldsBytes, st, b->P);
		const hipError_t ce = hipStreamEndCapture(st, &b->graph);
		HIPCHK(le);
		HIPCHK(ce);
		HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
	});
}

int lrh_batch_replay(lrh_batch *b, void *stream) {
	if (!b) return fail(LRH_EINVAL, "null batch");
	if (!b->exec) return fail(LRH_EINVAL, "batch has no captured graph (lrh_batch_capture)");
	return guarded([&]() {
		b->dev->activate();
		checkBatchScenes(b);
		HIPCHK(hipGraphLaunch(b->exec, stream ? (hipStream_t)stream : b->dev->stream));
		b->dev->rays += b->T.total;
	});
}

int lrh_batch_info(lrh_batch *b, unsigned long long *rays, unsigned int *blocks, unsigned int *blockSize) {
	if (!b) return fail(LRH_EINVAL, "null batch");
	if (rays) *rays = b->T.total;
	if (blocks) *blocks = b->L.blocks;
	if (blockSize) *blockSize = b->L.block;
	return LRH_OK;
}

int lrh_batch_destroy(lrh_batch *b) {
	if (!b) return LRH_OK;
	(void)hipSetDevice(b->dev->ordinal);
	(void)hipDeviceSynchronize();
	if (b->exec) (void)hipGraphExecDestroy(b->exec);
	if (b->graph) (void)hipGraphDestroy(b->graph);
	delete b;
	return LRH_OK;
}

int lrh_trace_batch(lrh_device *dev, const lrh_batch_item *items, unsigned int nItems, const lrh_trace_opts *opts,
		void *stream) {
	if (!dev || (nItems && !items)) return fail(LRH_EINVAL, "null argument");
	if (nItems == 0) return LRH_OK;
	return guarded([&]() {
		dev->activate();
		hipStream_t st = stream ? (hipStream_t)stream : dev->stream;
		BatchTable T = makeTable(dev, items, nItems);
		if (T.total == 0) return;
		const size_t extra = ((nItems + 1) * sizeof(uint32_t) + 15) & ~(size_t)15;
		const LaunchShape L = chooseShape(dev, T.layout, T.two, false, true, (unsigned)T.total, T.maxDepth, extra, opts);
		const unsigned nLanes = L.blocks * L.block;
		lrh_device::Workspace &w = workspace(dev, st, (size_t)nLanes * L.spillDepth * sizeof(uint2));
		{
			// descriptor table of this call; stream order keeps the previous
			// call's kernel ahead of the copy that overwrites it
			std::lock_guard<std::mutex> g(dev->mu);
			const size_t need = T.descs.size() * sizeof(LrhBatchDesc) + T.starts.size() * sizeof(uint32_t);
			if (w.tableBytes < need) {
				HIPCHK(hipStreamSynchronize(st));
				dev->devFree(w.table, w.tableBytes);
				w.table = nullptr;
				w.tableBytes = 0;
				w.table = (uint8_t *)dev->devAlloc(need);
				w.tableBytes = need;
			}
		}
		LrhBatchDesc *dDesc = (LrhBatchDesc *)w.table;
		uint32_t *dStart = (uint32_t *)(w.table + T.descs.size() * sizeof(LrhBatchDesc));
		HIPCHK(hipMemcpyAsync(dDesc, T.descs.data(), T.descs.size() * sizeof(LrhBatchDesc), hipMemcpyHostToDevice, st));
		HIPCHK(hipMemcpyAsync(dStart, T.starts.data(), T.starts.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st));
		LrhTraceParams P;
		std::memset(&P, 0, sizeof(P));
		P.nRays = (uint32_t)T.total;
		P.root = LRH_CHILD_EMPTY;
		P.counter = w.counter;
		P.spill = w.spill;
		P.spillDepth = L.spillDepth;
		P.nLanes = nLanes;
		P.ldsDepth = L.lds;
		P.refill = L.refill;
		P.leafTris = L.leafTris;
		P.error = dev->dError;
		P.stats = w.stats;
		P.items = dDesc;
		P.itemStart = dStart;
		P.nItems = nItems;
		HIPCHK(lrhLaunchTrace(T.layout, T.two, false, false, true, dim3(L.blocks), dim3(L.block), L.ldsBytes, st, P));
		dev->rays += T.total;
	});
}

int lrh_rays_camera(lrh_device *dev, const lrh_camera *cam, void *rays, void *stream) {
	if (!dev || !cam || !rays) return fail(LRH_EINVAL, "null argument");
	if (!cam->width || !cam->height || (uint64_t)cam->width * cam->height > 0x7fffffffull)
		return fail(LRH_ERANGE, "bad camera size");
	if ((uintptr_t)rays & 15u) return fail(LRH_EINVAL, "ray buffer must be 16-byte aligned");
	return guarded([&]() {
		dev->activate();
		LrhCameraParams C;
		double f[3], r[3], u[3];
		for (int k = 0; k < 3; ++k) f[k] = (double)cam->look[k] - cam->eye[k];
		double fl = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
		if (!(fl > 0)) throw Error(LRH_EINVAL, "camera eye equals look");
		for (int k = 0; k < 3; ++k) f[k] /= fl;
		r[0] = f[1] * cam->up[2] - f[2] * cam->up[1];
		r[1] = f[2] * cam->up[0] - f[0] * cam->up[2];
		r[2] = f[0] * cam->up[1] - f[1] * cam->up[0];
		double rl = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
		if (!(rl > 0)) throw Error(LRH_EINVAL, "camera up parallel to the view direction");
		for (int k = 0; k < 3; ++k) r[k] /= rl;
		u[0] = r[1] * f[2] - r[2] * f[1];
		u[1] = r[2] * f[0] - r[0] * f[2];
		u[2] = r[0] * f[1] - r[1] * f[0];
		for (int k = 0; k < 3; ++k) {
			C.eye[k] = cam->eye[k];
			C.fwd[k] = (float)f[k];
			C.right[k] = (float)r[k];
			C.up[k] = (float)u[k];
		}
		C.th = (float)tan(cam->vfovDeg * 3.14159265358979 / 360.0);
		C.tw = C.th * (float)cam->width / (float)cam->height;
		C.mint = cam->mint;
		C.width = cam->width;
		C.height = cam->height;
		HIPCHK(lrhLaunchCamera(C, (lrh_ray_rec *)rays, stream ? (hipStream_t)stream : dev->stream));
	});
}

int lrh_rays_bounce(lrh_scene *s, const void *rays, const void *hits, unsigned int count, unsigned int bounce,
		unsigned long long seed, unsigned int pixelBase, unsigned int rrStart, float rrSurvive,
		const float lightCenter[3], void *nextRays, void *shadowRays, unsigned int *liveCount, void *stream) {
	if (!s || !s->dev || (count && (!rays || !hits || !nextRays)) || !lightCenter)
		return fail(LRH_EINVAL, "null argument");
	if (s->twoLevel) return fail(LRH_EINVAL, "bounce generation needs a single-level scene");
	if (((uintptr_t)rays & 15u) || ((uintptr_t)nextRays & 15u) || ((uintptr_t)shadowRays & 15u) ||
			((uintptr_t)hits & 3u))
		return fail(LRH_EINVAL, "misaligned buffer");
	if (count == 0) return LRH_OK;
	return guarded([&]() {
		s->dev->activate();
		LrhBounceParams P;
		P.rays = (const lrh_ray_rec *)rays;
		P.hits = (const lrh_hit_rec *)hits;
		P.next = (lrh_ray_rec *)nextRays;
		P.shadow = (lrh_ray_rec *)shadowRays;
		P.tris = (const lrh_tri *)s->tris.p;
		P.invPerm = (const uint32_t *)s->invPerm.p;
		P.meshTriBase = (const uint32_t *)s->meshTriBase.p;
		P.nMeshes = s->nMeshes;
		P.count = count;
		P.bounce = bounce;
		P.rrStart = rrStart;
		P.rrSurvive = rrSurvive;
		for (int k = 0; k < 3; ++k) P.light[k] = lightCenter[k];
		P.seed = seed;
		P.pixelBase = pixelBase;
		P.liveCount = liveCount;
		HIPCHK(lrhLaunchBounce(P, stream ? (hipStream_t)stream : s->dev->stream));
	});
}

static_assert(sizeof(lrh_material) == sizeof(LrhMaterial), "lrh_material layout");
static_assert(sizeof(lrh_light_tri) == sizeof(LrhLightTri), "lrh_light_tri layout");
static_assert(sizeof(lrh_path) == sizeof(LrhPath) && sizeof(lrh_path) == 48, "lrh_path layout");

}  // extern "C"

namespace {
// checks of one lrh_shade call and its kernel parameters
int shadeParams(lrh_scene *s, const lrh_material *materials, unsigned int nMaterials, const lrh_light_tri *lights,
		unsigned int nLights, const void *rays, const void *hits, const void *shadowHits, void *paths,
		void *nextRays, void *shadowRays, const lrh_shade_params &prm, unsigned int *liveCount, LrhShadeParams &P) {
	if (s->twoLevel) return fail(LRH_EINVAL, "shading needs a single-level scene");
	const unsigned count = prm.count;
	if (count && (!rays || !hits || !paths || !nextRays || !shadowRays))
		return fail(LRH_EINVAL, "null ray, hit or path buffer");
	if ((nMaterials && !materials) || (nLights && !lights)) return fail(LRH_EINVAL, "null material or light array");
	if (((uintptr_t)rays & 15u) || ((uintptr_t)nextRays & 15u) || ((uintptr_t)shadowRays & 15u) ||
			((uintptr_t)hits & 3u) || ((uintptr_t)shadowHits & 3u) || ((uintptr_t)paths & 3u) ||
			((uintptr_t)materials & 3u) || ((uintptr_t)lights & 3u))
		return fail(LRH_EINVAL, "misaligned buffer");
	if (prm.maxDepth == 0) return fail(LRH_EINVAL, "maxDepth must be at least 1");
	if (!(prm.rrSurvive > 0.f && prm.rrSurvive <= 1.f)) return fail(LRH_EINVAL, "rrSurvive must be in (0, 1]");
	P.rays = (const lrh_ray_rec *)rays;
	P.hits = (const lrh_hit_rec *)hits;
	P.shadowHits = (const lrh_hit_rec *)shadowHits;
	P.paths = (LrhPath *)paths;
	P.next = (lrh_ray_rec *)nextRays;
	P.shadow = (lrh_ray_rec *)shadowRays;
	P.materials = (const LrhMaterial *)materials;
	P.nMaterials = nMaterials;
	P.lights = (const LrhLightTri *)lights;
	P.nLights = nLights;
	P.tris = (const lrh_tri *)s->tris.p;
	P.invPerm = (const uint32_t *)s->invPerm.p;
	P.meshTriBase = (const uint32_t *)s->meshTriBase.p;
	P.nMeshes = s->nMeshes;
	P.count = count;
	P.maxDepth = prm.maxDepth;
	P.rrDepth = prm.rrDepth;
	P.rrSurvive = prm.rrSurvive;
	P.seed = prm.seed;
	P.pixelBase = prm.pixelBase;
	P.liveCount = liveCount;
	return LRH_OK;
}
}  // namespace

extern "C" {

int lrh_paths_init(lrh_device *dev, void *paths, unsigned int count, void *stream) {
	if (!dev || (count && !paths)) return fail(LRH_EINVAL, "null argument");
	if ((uintptr_t)paths & 3u) return fail(LRH_EINVAL, "misaligned path buffer");
	if (count == 0) return LRH_OK;
	return guarded([&]() {
		dev->activate();
		HIPCHK(lrhLaunchPathsInit((LrhPath *)paths, count, stream ? (hipStream_t)stream : dev->stream));
	});
}

int lrh_shade(lrh_scene *s, const lrh_material *materials, unsigned int nMaterials, const lrh_light_tri *lights,
		unsigned int nLights, const void *rays, const void *hits, const void *shadowHits, void *paths,
		void *nextRays, void *shadowRays, const lrh_shade_params *prm, unsigned int *liveCount, void *stream) {
	if (!s || !s->dev || !prm) return fail(LRH_EINVAL, "null argument");
	LrhShadeParams P;
	const int rc = shadeParams(s, materials, nMaterials, lights, nLights, rays, hits, shadowHits, paths, nextRays,
			shadowRays, *prm, liveCount, P);
	if (rc != LRH_OK || prm->count == 0) return rc;
	return guarded([&]() {
		s->dev->activate();
		HIPCHK(lrhLaunchShade(P, stream ? (hipStream_t)stream : s->dev->stream));
	});
}

int lrh_shade_batch(const lrh_shade_item *items, unsigned int nItems, const lrh_material *materials,
		unsigned int nMaterials, const lrh_light_tri *lights, unsigned int nLights, unsigned int *liveCount,
		void *stream) {
	if (!items || nItems == 0) return fail(LRH_EINVAL, "null or empty item list");
	if (nItems > LRH_SHADE_BATCH_MAX) return fail(LRH_ERANGE, "more than LRH_SHADE_BATCH_MAX items");
	LrhShadeBatchArgs A{};
	lrh_device *dev = nullptr;
	uint64_t blocks = 0;
	for (unsigned k = 0; k < nItems; ++k) {
		const lrh_shade_item &it = items[k];
		if (!it.scene || !it.scene->dev) return fail(LRH_EINVAL, "null scene in item " + std::to_string(k));
		if (dev && it.scene->dev != dev) return fail(LRH_EINVAL, "items on different devices");
		dev = it.scene->dev;
		LrhShadeParams P;
		const int rc = shadeParams(it.scene, materials, nMaterials, lights, nLights, it.rays, it.hits,
				it.shadowHits, it.paths, it.nextRays, it.shadowRays, it.params, liveCount, P);
		if (rc != LRH_OK) return rc;
		if (P.count == 0) continue;  // nothing to shade: no blocks
		A.blockStart[A.n] = (uint32_t)blocks;
		A.p[A.n++] = P;
		blocks += (P.count + 255u) / 256u;
	}
	if (A.n == 0) return LRH_OK;
	if (blocks > 0xFFFFFFFFull) return fail(LRH_ERANGE, "batch too large");
	A.blockStart[A.n] = (uint32_t)blocks;
	return guarded([&]() {
		dev->activate();
		HIPCHK(lrhLaunchShadeBatch(A, stream ? (hipStream_t)stream : dev->stream));
	});
}

int lrh_event_query(lrh_event *ev) {
	if (!ev) return fail(LRH_EINVAL, "null event");
	const hipError_t e = hipEventQuery(ev->ev);
	if (e == hipSuccess) return 1;
	if (e == hipErrorNotReady) return 0;
	return fail(LRH_EHIP, std::string("hipEventQuery: ") + hipGetErrorString(e));
}

int lrh_device_stats(lrh_device *dev, unsigned long long *rays, double *seconds) {
	if (!dev) return fail(LRH_EINVAL, "null device");
	if (rays) *rays = dev->rays.load();
	if (seconds) *seconds = now() - dev->t0;
	return LRH_OK;
}

}  // extern "C"
