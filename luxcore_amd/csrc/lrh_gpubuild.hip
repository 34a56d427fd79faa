/*
 * Linear BVH build on the device, written into the 4-wide node records of
 * lrh_layout.h.  Steps (one stream, no host round trip except one count per
 * level of the collapse):
 *   1. per triangle: the reference's leaf box (Triangle::WorldBound expanded
 *      by MachineEpsilon::E, src/luxrays/accelerators/bvhaccel.cpp:116-122),
 *      rounded exactly as the host builder rounds it, and the bounds of the
 *      box centres;
 *   2. 63-bit Morton codes of the centres, radix-sorted with the triangle id
 *      as value (stable, so equal codes keep id order);
 *   3. binary radix tree over the sorted codes (Karras, "Maximizing
 *      Parallelism in the Construction of BVHs, Octrees, and k-d Trees",
 *      HPG 2012), equal codes ordered by position;
 *   4. boxes bottom-up: the second child to finish computes its parent;
 *   5. breadth-first collapse into 4-wide nodes: each node takes a binary
 *      node's two children and keeps splitting its largest-area child until
 *      it has four; a subtree of at most maxLeaf triangles becomes a leaf
 *      (its triangles are consecutive in Morton order); child slots come
 *      from a prefix sum, so the node order is deterministic and the top of
 *      the tree comes first;
 *   6. triangle records in leaf order, bit for bit the host records.
 */
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "lrh_devnode.h"
#include "lrh_gpubuild.h"
#include "lrh_layout.h"

namespace {

using namespace lrhdev;

#define GB_CHECK(x)                                                                                   \
	do {                                                                                              \
		const hipError_t e_ = (x);                                                                    \
		if (e_ != hipSuccess) throw std::runtime_error(std::string("device build: ") + hipGetErrorString(e_)); \
	} while (0)

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kNone = 0xffffffffu;

__device__ __forceinline__ uint32_t meshOf(const uint32_t *base, uint32_t nMeshes, uint32_t i) {
	uint32_t lo = 0, hi = nMeshes;  // last mesh whose base <= i
	while (hi - lo > 1) {
		const uint32_t mid = (lo + hi) >> 1;
		if (base[mid] <= i) lo = mid; else hi = mid;
	}
	return lo;
}

// triangle i (global id: the triangle arrays are concatenated) of mesh m;
// its vertex indices are local to the mesh
__device__ __forceinline__ void loadTri(const float *verts, const uint32_t *tris, const uint32_t *vbase, uint32_t m,
		uint32_t i, float p[3][3]) {
	for (int c = 0; c < 3; ++c) {
		const size_t v = (size_t)vbase[m] + tris[3 * (size_t)i + c];
		for (int k = 0; k < 3; ++k) p[c][k] = verts[3 * v + k];
	}
}

__global__ void k_initBounds(uint32_t *cb) {
	if (threadIdx.x < 3) {
		cb[threadIdx.x] = 0xffffffffu;  // min keys
		cb[3 + threadIdx.x] = 0u;       // max keys
	}
}

// 1. leaf boxes (triangleBox of the host builder) and centre bounds
__global__ void k_prims(const float *verts, const uint32_t *tris, const uint32_t *tbase, const uint32_t *vbase,
		uint32_t nMeshes, uint32_t n, GBox *boxes, uint32_t *cb) {
	__shared__ uint32_t smin[3], smax[3];
	if (threadIdx.x < 3) {
		smin[threadIdx.x] = 0xffffffffu;
		smax[threadIdx.x] = 0u;
	}
	__syncthreads();
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) {
		const uint32_t m = meshOf(tbase, nMeshes, i);
		float p[3][3];
		loadTri(verts, tris, vbase, m, i, p);
		GBox b;
		for (int k = 0; k < 3; ++k) {
			b.lo[k] = minT(p[0][k], p[1][k]);
			b.hi[k] = maxT(p[0][k], p[1][k]);
		}
		for (int k = 0; k < 3; ++k) {
			b.lo[k] = minT(b.lo[k], p[2][k]);
			b.hi[k] = maxT(b.hi[k], p[2][k]);
		}
		const float ea = maxT(machineEps(b.lo[0]), maxT(machineEps(b.lo[1]), machineEps(b.lo[2])));
		const float ec = maxT(machineEps(b.hi[0]), maxT(machineEps(b.hi[1]), machineEps(b.hi[2])));
		const float e = maxT(ea, ec);
		for (int k = 0; k < 3; ++k) {
			b.lo[k] -= e;
			b.hi[k] += e;
		}
		boxes[i] = b;
		for (int k = 0; k < 3; ++k) {
			const uint32_t c = fkey(b.lo[k] + b.hi[k]);
			atomicMin(&smin[k], c);
			atomicMax(&smax[k], c);
		}
	}
	__syncthreads();
	if (threadIdx.x < 3) {
		atomicMin(&cb[threadIdx.x], smin[threadIdx.x]);
		atomicMax(&cb[3 + threadIdx.x], smax[threadIdx.x]);
	}
}

__device__ __forceinline__ uint64_t spread21(uint32_t v) {
	uint64_t x = v & 0x1fffffu;
	x = (x | (x << 32)) & 0x1f00000000ffffull;
	x = (x | (x << 16)) & 0x1f0000ff0000ffull;
	x = (x | (x << 8)) & 0x100f00f00f00f00full;
	x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
	x = (x | (x << 2)) & 0x1249249249249249ull;
	return x;
}

// 2. Morton codes of the box centres (21 bits per axis)
__global__ void k_morton(const GBox *boxes, uint32_t n, const uint32_t *cb, uint64_t *keys, uint32_t *vals) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const GBox b = boxes[i];
	uint32_t q[3];
	for (int k = 0; k < 3; ++k) {
		const float lo = unkey(cb[k]), hi = unkey(cb[3 + k]);
		const float ext = hi - lo;
		const float u = ext > 0.f ? ((b.lo[k] + b.hi[k]) - lo) / ext : 0.f;
		float s = u * 2097152.f;
		s = s < 0.f ? 0.f : (s > 2097151.f ? 2097151.f : s);
		q[k] = (uint32_t)s;
	}
	keys[i] = (spread21(q[0]) << 2) | (spread21(q[1]) << 1) | spread21(q[2]);
	vals[i] = i;
}

__device__ __forceinline__ int delta(const uint64_t *k, int n, int i, int j) {
	if (j < 0 || j >= n) return -1;
	const uint64_t a = k[i], b = k[j];
	if (a != b) return __clzll((long long)(a ^ b));
	return 64 + __clz((int)((uint32_t)i ^ (uint32_t)j));
}

// 3. binary radix tree: internal node i, children (kLeafBit | leaf) or internal index
__global__ void k_radixTree(const uint64_t *keys, int n, uint32_t *left, uint32_t *right, uint32_t *first,
		uint32_t *last, uint32_t *iparent, uint32_t *lparent) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n - 1) return;
	const int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
	const int dmin = delta(keys, n, i, i - d);
	int lmax = 2;
	while (delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
	int l = 0;
	for (int t = lmax / 2; t >= 1; t /= 2)
		if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
	const int j = i + l * d;
	const int dnode = delta(keys, n, i, j);
	int s = 0;
	for (int t = (l + 1) / 2;; t = (t + 1) / 2) {
		if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
		if (t == 1) break;
	}
	const int g = i + s * d + (d < 0 ? -1 : 0);
	const int lo = i < j ? i : j, hi = i < j ? j : i;
	const uint32_t lc = (lo == g) ? (kLeafBit | (uint32_t)g) : (uint32_t)g;
	const uint32_t rc = (hi == g + 1) ? (kLeafBit | (uint32_t)(g + 1)) : (uint32_t)(g + 1);
	left[i] = lc;
	right[i] = rc;
	first[i] = (uint32_t)lo;
	last[i] = (uint32_t)hi;
	if (lc & kLeafBit) lparent[lc & ~kLeafBit] = (uint32_t)i; else iparent[lc] = (uint32_t)i;
	if (rc & kLeafBit) lparent[rc & ~kLeafBit] = (uint32_t)i; else iparent[rc] = (uint32_t)i;
	if (i == 0) iparent[0] = kNone;
}

__device__ __forceinline__ float loadF(const float *p) {
	// read another thread's result past this CU's L1
	return __uint_as_float(__hip_atomic_load(reinterpret_cast<const uint32_t *>(p), __ATOMIC_RELAXED,
			__HIP_MEMORY_SCOPE_AGENT));
}

__device__ GBox childBox(uint32_t c, const GBox *leafBox, const GBox *nodeBox) {
	const float *p = (c & kLeafBit) ? leafBox[c & ~kLeafBit].lo : nodeBox[c].lo;
	GBox b;
	for (int k = 0; k < 3; ++k) {
		b.lo[k] = loadF(p + k);
		b.hi[k] = loadF(p + 3 + k);
	}
	return b;
}

// 4. boxes bottom-up; also the leaf boxes in sorted order
__global__ void k_boxesUp(const GBox *boxes, const uint32_t *vals, uint32_t n, const uint32_t *left,
		const uint32_t *right, const uint32_t *iparent, const uint32_t *lparent, uint32_t *flags, GBox *leafBox,
		GBox *nodeBox) {
	const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= n) return;
	leafBox[j] = boxes[vals[j]];
	uint32_t p = lparent[j];
	// a binary radix tree over n keys is at most 96 levels deep (64 code bits
	// + 32 position bits); the bound also stops the walk on a malformed tree
	for (int steps = 0; p < n - 1 && steps < 128; ++steps) {
		__threadfence();
		if (atomicAdd(&flags[p], 1u) == 0u) return;  // the sibling finishes this node
		__threadfence();
		const GBox a = childBox(left[p], leafBox, nodeBox), b = childBox(right[p], leafBox, nodeBox);
		GBox u;
		for (int k = 0; k < 3; ++k) {
			u.lo[k] = minT(a.lo[k], b.lo[k]);
			u.hi[k] = maxT(a.hi[k], b.hi[k]);
		}
		nodeBox[p] = u;
		p = iparent[p];
	}
}

__device__ __forceinline__ uint32_t leafRef(uint32_t start, uint32_t count) {
	return LRH_CHILD_LEAF | ((count - 1u) << LRH_LEAF_START_BITS) | start;
}

__device__ __forceinline__ uint32_t countOf(uint32_t c, const uint32_t *first, const uint32_t *last) {
	return (c & kLeafBit) ? 1u : last[c] - first[c] + 1u;
}

__device__ __forceinline__ float areaOf(const GBox &b) {
	const float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
	if (!(dx >= 0.f)) return 0.f;
	return 2.f * (dx * dy + dy * dz + dz * dx);
}

// 5a. children of each frontier node (binary refs, Morton order) and how many become inner nodes
__global__ void k_expand(const uint32_t *frontier, uint32_t nf, uint32_t nInner, const uint32_t *left,
		const uint32_t *right, const uint32_t *first, const uint32_t *last, const GBox *leafBox, const GBox *nodeBox,
		uint32_t maxLeaf, uint32_t *kids, uint32_t *innerCount) {
	const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= nf) return;
	const uint32_t b = frontier[e];
	if (b >= nInner) {  // not an internal node: nothing to expand (malformed input)
		for (int k = 0; k < 4; ++k) kids[4 * e + k] = kNone;
		innerCount[e] = 0;
		return;
	}
	uint32_t ch[4] = {left[b], right[b], kNone, kNone};
	int nch = 2;
	while (nch < 4) {
		int pick = -1;
		float best = -1.f;
		for (int k = 0; k < nch; ++k) {
			if ((ch[k] & kLeafBit) || countOf(ch[k], first, last) <= maxLeaf) continue;
			const float a = areaOf(nodeBox[ch[k]]);
			if (a > best) {
				best = a;
				pick = k;
			}
		}
		if (pick < 0) break;
		const uint32_t c = ch[pick];
		for (int k = nch; k > pick + 1; --k) ch[k] = ch[k - 1];
		ch[pick] = left[c];
		ch[pick + 1] = right[c];
		++nch;
	}
	uint32_t inner = 0;
	for (int k = 0; k < 4; ++k) {
		kids[4 * e + k] = ch[k];
		if (ch[k] != kNone && countOf(ch[k], first, last) > maxLeaf) ++inner;
	}
	innerCount[e] = inner;
}

// 5b. write the frontier's wide nodes; inner children get slots after the level
__global__ void k_emit(const uint32_t *frontier, uint32_t nf, const uint32_t *kids, const uint32_t *innerOff,
		const uint32_t *first, const uint32_t *last, const GBox *leafBox, const GBox *nodeBox, uint32_t maxLeaf,
		uint32_t levelBase, uint32_t layout, void *nodes, uint32_t *next) {
	const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= nf) return;
	GBox cb[4];
	uint32_t ref[4] = {LRH_CHILD_EMPTY, LRH_CHILD_EMPTY, LRH_CHILD_EMPTY, LRH_CHILD_EMPTY};
	int n = 0;
	uint32_t slot = innerOff[e];
	for (int k = 0; k < 4; ++k) {
		const uint32_t c = kids[4 * e + k];
		if (c == kNone) continue;
		cb[n] = (c & kLeafBit) ? leafBox[c & ~kLeafBit] : nodeBox[c];
		const uint32_t cnt = countOf(c, first, last);
		if (cnt <= maxLeaf) {
			const uint32_t start = (c & kLeafBit) ? (c & ~kLeafBit) : first[c];
			ref[n] = leafRef(start, cnt);
		} else {
			ref[n] = levelBase + nf + slot;
			next[slot] = c;
			++slot;
		}
		++n;
	}
	if (layout == 0)
		quantizeNode(cb, n, ref, reinterpret_cast<lrh_node4q *>(nodes) + levelBase + e);
	else
		floatNode(cb, n, ref, reinterpret_cast<lrh_node4f *>(nodes) + levelBase + e);
}

// a scene of at most maxLeaf triangles: one node with one leaf
__global__ void k_smallRoot(const GBox *leafBox, uint32_t n, uint32_t layout, void *nodes) {
	if (threadIdx.x != 0 || blockIdx.x != 0) return;
	GBox b;
	for (int k = 0; k < 3; ++k) {
		b.lo[k] = INFINITY;
		b.hi[k] = -INFINITY;
	}
	for (uint32_t i = 0; i < n; ++i)
		for (int k = 0; k < 3; ++k) {
			b.lo[k] = minT(b.lo[k], leafBox[i].lo[k]);
			b.hi[k] = maxT(b.hi[k], leafBox[i].hi[k]);
		}
	const uint32_t ref[4] = {leafRef(0, n), LRH_CHILD_EMPTY, LRH_CHILD_EMPTY, LRH_CHILD_EMPTY};
	if (layout == 0)
		quantizeNode(&b, 1, ref, reinterpret_cast<lrh_node4q *>(nodes));
	else
		floatNode(&b, 1, ref, reinterpret_cast<lrh_node4f *>(nodes));
}

__global__ void k_identity(const GBox *boxes, uint32_t n, uint32_t *vals, GBox *leafBox) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	vals[i] = i;
	leafBox[i] = boxes[i];
}

// 6. triangle records in leaf order, as the host builder writes them
__global__ void k_records(const float *verts, const uint32_t *tris, const uint32_t *tbase, const uint32_t *vbase,
		uint32_t nMeshes, const uint32_t *vals, uint32_t n, lrh_tri *recs, uint32_t *invPerm) {
	const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= n) return;
	const uint32_t i = vals[j];
	const uint32_t m = meshOf(tbase, nMeshes, i);
	float p[3][3];
	loadTri(verts, tris, vbase, m, i, p);
	lrh_tri r;
	for (int k = 0; k < 3; ++k) {
		r.p0[k] = p[0][k];
		r.e1[k] = p[1][k] - p[0][k];
		r.e2[k] = p[2][k] - p[0][k];
	}
	r.triIndex = i - tbase[m];
	r.meshIndex = m;
	r.pad = 0;
	recs[j] = r;
	invPerm[i] = j;
}

struct Scratch {
	const LrhGpuAlloc &a;
	std::vector<std::pair<void *, size_t>> held;
	explicit Scratch(const LrhGpuAlloc &al) : a(al) {}
	template <class T>
	T *get(size_t count) {
		const size_t bytes = std::max<size_t>(count * sizeof(T), 16);
		void *p = a.alloc(a.ctx, bytes);
		held.push_back({p, bytes});
		return static_cast<T *>(p);
	}
	~Scratch() {
		for (auto &h : held) a.release(a.ctx, h.first, h.second);
	}
};

inline unsigned grid(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

void lrhGpuBuildBVH(const LrhGpuBuildIn &in, const LrhGpuAlloc &a, hipStream_t st, LrhGpuBuildOut *out) {
	const uint32_t n = in.nTris;
	if (n == 0) throw std::runtime_error("device build: empty scene");
	if (in.maxLeaf < 1 || in.maxLeaf > LRH_LEAF_MAX) throw std::runtime_error("device build: maxLeaf must be 1..8");
	const size_t nodeSz = in.layout == 0 ? sizeof(lrh_node4q) : sizeof(lrh_node4f);
	Scratch S(a);
	hipEvent_t e0, e1;
	GB_CHECK(hipEventCreate(&e0));
	GB_CHECK(hipEventCreate(&e1));
	GB_CHECK(hipEventRecord(e0, st));

	uint32_t *tbase = S.get<uint32_t>(in.nMeshes + 1), *vbase = S.get<uint32_t>(in.nMeshes + 1);
	GB_CHECK(hipMemcpyAsync(tbase, in.meshTriBase, (in.nMeshes + 1) * 4, hipMemcpyHostToDevice, st));
	GB_CHECK(hipMemcpyAsync(vbase, in.meshVertBase, (in.nMeshes + 1) * 4, hipMemcpyHostToDevice, st));
	GBox *boxes = S.get<GBox>(n);
	uint32_t *cb = S.get<uint32_t>(6);
	hipLaunchKernelGGL(k_initBounds, dim3(1), dim3(64), 0, st, cb);
	hipLaunchKernelGGL(k_prims, dim3(grid(n, 256)), dim3(256), 0, st, in.verts, in.tris, tbase, vbase, in.nMeshes, n,
			boxes, cb);
	GB_CHECK(hipGetLastError());

	uint32_t *vals = S.get<uint32_t>(n);
	GBox *leafBox = S.get<GBox>(n);
	// output buffers (owned by the caller from here on)
	const size_t capNodes = std::max<uint32_t>(n > 1 ? n - 1 : 1, 1);
	void *nodes = a.alloc(a.ctx, capNodes * nodeSz);
	uint32_t nodeCount = 0, depth = 0;
	try {
		if (n <= in.maxLeaf) {
			hipLaunchKernelGGL(k_identity, dim3(grid(n, 256)), dim3(256), 0, st, boxes, n, vals, leafBox);
			hipLaunchKernelGGL(k_smallRoot, dim3(1), dim3(64), 0, st, leafBox, n, in.layout, nodes);
			GB_CHECK(hipGetLastError());
			nodeCount = 1;
		} else {
			uint64_t *keys = S.get<uint64_t>(n), *keysS = S.get<uint64_t>(n);
			uint32_t *valsIn = S.get<uint32_t>(n);
			hipLaunchKernelGGL(k_morton, dim3(grid(n, 256)), dim3(256), 0, st, boxes, n, cb, keys, valsIn);
			GB_CHECK(hipGetLastError());
			size_t tmpBytes = 0;
			GB_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmpBytes, keys, keysS, valsIn, vals, (int)n, 0, 63, st));
			void *tmp = S.get<uint8_t>(tmpBytes);
			GB_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, tmpBytes, keys, keysS, valsIn, vals, (int)n, 0, 63, st));

			const uint32_t ni = n - 1;
			uint32_t *left = S.get<uint32_t>(ni), *right = S.get<uint32_t>(ni), *first = S.get<uint32_t>(ni),
					 *last = S.get<uint32_t>(ni), *iparent = S.get<uint32_t>(ni), *lparent = S.get<uint32_t>(n),
					 *flags = S.get<uint32_t>(ni);
			GBox *nodeBox = S.get<GBox>(ni);
			GB_CHECK(hipMemsetAsync(flags, 0, (size_t)ni * 4, st));
			GB_CHECK(hipMemsetAsync(iparent, 0xff, (size_t)ni * 4, st));  // unset parents end the climb
			GB_CHECK(hipMemsetAsync(lparent, 0xff, (size_t)n * 4, st));
			hipLaunchKernelGGL(k_radixTree, dim3(grid(ni, 256)), dim3(256), 0, st, keysS, (int)n, left, right, first,
					last, iparent, lparent);
			hipLaunchKernelGGL(k_boxesUp, dim3(grid(n, 256)), dim3(256), 0, st, boxes, vals, n, left, right, iparent,
					lparent, flags, leafBox, nodeBox);
			GB_CHECK(hipGetLastError());

			// breadth-first collapse: level frontiers of binary node ids
			uint32_t *fA = S.get<uint32_t>(ni), *fB = S.get<uint32_t>(ni);
			uint32_t *kids = S.get<uint32_t>(4 * (size_t)ni), *cnt = S.get<uint32_t>(ni + 1),
					 *off = S.get<uint32_t>(ni + 1);
			size_t scanBytes = 0;
			GB_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, scanBytes, cnt, off, (int)(ni + 1), st));
			void *scanTmp = S.get<uint8_t>(scanBytes);
			const uint32_t root = 0;
			GB_CHECK(hipMemcpyAsync(fA, &root, 4, hipMemcpyHostToDevice, st));
			uint32_t nf = 1, levelBase = 0;
			while (nf) {
				if ((size_t)levelBase + nf > capNodes) throw std::runtime_error("device build: node capacity");
				hipLaunchKernelGGL(k_expand, dim3(grid(nf, 256)), dim3(256), 0, st, fA, nf, ni, left, right, first,
						last, leafBox, nodeBox, in.maxLeaf, kids, cnt);
				GB_CHECK(hipMemsetAsync(cnt + nf, 0, 4, st));
				GB_CHECK(hipcub::DeviceScan::ExclusiveSum(scanTmp, scanBytes, cnt, off, (int)(nf + 1), st));
				hipLaunchKernelGGL(k_emit, dim3(grid(nf, 256)), dim3(256), 0, st, fA, nf, kids, off, first, last,
						leafBox, nodeBox, in.maxLeaf, levelBase, in.layout, nodes, fB);
				GB_CHECK(hipGetLastError());
				uint32_t nextN = 0;
				GB_CHECK(hipMemcpyAsync(&nextN, off + nf, 4, hipMemcpyDeviceToHost, st));
				GB_CHECK(hipStreamSynchronize(st));
				levelBase += nf;
				nf = nextN;
				std::swap(fA, fB);
				++depth;
			}
			nodeCount = levelBase;
		}
		// triangle records in leaf order
		out->triBytes = (size_t)n * sizeof(lrh_tri);
		out->tris = a.alloc(a.ctx, out->triBytes);
		out->invPermBytes = (size_t)n * 4;
		out->invPerm = static_cast<uint32_t *>(a.alloc(a.ctx, out->invPermBytes));
		hipLaunchKernelGGL(k_records, dim3(grid(n, 256)), dim3(256), 0, st, in.verts, in.tris, tbase, vbase,
				in.nMeshes, vals, n, static_cast<lrh_tri *>(out->tris), out->invPerm);
		GB_CHECK(hipGetLastError());
		// right-size the node array
		out->nodeBytes = (size_t)nodeCount * nodeSz;
		out->nodes = a.alloc(a.ctx, out->nodeBytes);
		GB_CHECK(hipMemcpyAsync(out->nodes, nodes, out->nodeBytes, hipMemcpyDeviceToDevice, st));
		GB_CHECK(hipEventRecord(e1, st));
		GB_CHECK(hipStreamSynchronize(st));
	} catch (...) {
		(void)hipStreamSynchronize(st);  // nothing in flight may still use what is released below
		a.release(a.ctx, nodes, capNodes * nodeSz);
		if (out->tris) a.release(a.ctx, out->tris, out->triBytes);
		if (out->invPerm) a.release(a.ctx, out->invPerm, out->invPermBytes);
		if (out->nodes) a.release(a.ctx, out->nodes, out->nodeBytes);
		out->tris = out->nodes = nullptr;
		out->invPerm = nullptr;
		(void)hipEventDestroy(e0);
		(void)hipEventDestroy(e1);
		throw;
	}
	a.release(a.ctx, nodes, capNodes * nodeSz);
	float ms = 0.f;
	(void)hipEventElapsedTime(&ms, e0, e1);
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	out->nodeCount = nodeCount;
	out->root = 0;
	out->maxDepth = depth ? depth - 1 : 0;
	out->kernelMs = ms;
}
