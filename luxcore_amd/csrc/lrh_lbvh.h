/*
 * The per-element steps of the device BVH builds, shared by the device
 * kernels (lrh_gpubuild.hip) and their sequential CPU emulation (the test
 * suite's check of the build's algorithm and of its error paths):
 *   1. a triangle's leaf box: Triangle::WorldBound expanded by
 *      MachineEpsilon::E (src/luxrays/accelerators/bvhaccel.cpp:116-122),
 *      rounded as the host builder rounds it;
 *   2. the 63-bit Morton code of its centre (21 bits per axis);
 *   3. a node of the binary radix tree over the sorted codes (Karras,
 *      "Maximizing Parallelism in the Construction of BVHs, Octrees, and k-d
 *      Trees", HPG 2012), equal codes ordered by position; or the nearest
 *      neighbour and partner of a cluster in a PLOC round;
 *   4. a node's surface-area cost and leaf decision, the expansion of a
 *      binary node into up to four or eight children (largest area first)
 *      and the wide record it becomes;
 *   5. the triangle record, bit for bit the host builder's.
 * Every search loop is bounded: keys that are not in order make the range
 * search fail (radixNode returns false) instead of looping.
 */
#ifndef LRH_LBVH_H
#define LRH_LBVH_H

#include <stdint.h>

#include "lrh_devnode.h"
#include "lrh_layout.h"

namespace lrhlbvh {

using lrhdev::GBox;
using lrhdev::maxT;
using lrhdev::minT;

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kNone = 0xffffffffu;

// bits of the build's error word
enum : uint32_t {
	kErrUnsorted = 1u,   // sorted codes out of order
	kErrRange = 2u,      // a radix-tree node found no range (keys not in order around it)
	kErrClimb = 4u,      // a radix-tree node deeper than any valid tree (or not below the root)
	kErrExpand = 8u,     // a frontier entry that is not an internal node
};

// radix sort: 8-bit digits, tiles of kSortTile keys per block
constexpr uint32_t kSortTile = 2048u;
constexpr uint32_t kSortBlock = 256u;
// exclusive scan: tiles of kScanTile values per block
constexpr uint32_t kScanTile = 2048u;

LRH_DN inline uint32_t meshOf(const uint32_t *base, uint32_t nMeshes, uint32_t i) {
	uint32_t lo = 0, hi = nMeshes;  // last mesh whose base <= i
	while (hi - lo > 1) {
		const uint32_t mid = (lo + hi) >> 1;
		if (base[mid] <= i) lo = mid; else hi = mid;
	}
	return lo;
}

// triangle i (global id over the concatenated meshes) of mesh m; its vertex
// indices are local to the mesh
LRH_DN inline void loadTri(const float *verts, const uint32_t *tris, const uint32_t *vbase, uint32_t m, uint32_t i,
		float p[3][3]) {
	for (int c = 0; c < 3; ++c) {
		const size_t v = (size_t)vbase[m] + tris[3 * (size_t)i + c];
		for (int k = 0; k < 3; ++k) p[c][k] = verts[3 * v + k];
	}
}

// 1. leaf box of triangle i (the host builder's triangleBox)
LRH_DN inline GBox primBox(const float *verts, const uint32_t *tris, const uint32_t *tbase, const uint32_t *vbase,
		uint32_t nMeshes, uint32_t i) {
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
	const float ea = maxT(lrhdev::machineEps(b.lo[0]), maxT(lrhdev::machineEps(b.lo[1]), lrhdev::machineEps(b.lo[2])));
	const float ec = maxT(lrhdev::machineEps(b.hi[0]), maxT(lrhdev::machineEps(b.hi[1]), lrhdev::machineEps(b.hi[2])));
	const float e = maxT(ea, ec);
	for (int k = 0; k < 3; ++k) {
		b.lo[k] -= e;
		b.hi[k] += e;
	}
	return b;
}

LRH_DN inline uint64_t spread21(uint32_t v) {
	uint64_t x = v & 0x1fffffu;
	x = (x | (x << 32)) & 0x1f00000000ffffull;
	x = (x | (x << 16)) & 0x1f0000ff0000ffull;
	x = (x | (x << 8)) & 0x100f00f00f00f00full;
	x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
	x = (x | (x << 2)) & 0x1249249249249249ull;
	return x;
}

// 2. Morton code of the box centre within the centre bounds (ordered keys cb:
// min x y z, max x y z), every axis scaled by the largest extent: a cube of
// cells, so that a flat scene (a terrain) is not ordered by height as
// finely as by position
LRH_DN inline uint64_t mortonKey(const GBox &b, const uint32_t cb[6]) {
	uint32_t q[3];
	float ext = 0.f;
	for (int k = 0; k < 3; ++k) ext = maxT(ext, lrhdev::unkey(cb[3 + k]) - lrhdev::unkey(cb[k]));
	for (int k = 0; k < 3; ++k) {
		const float lo = lrhdev::unkey(cb[k]);
		const float u = ext > 0.f ? ((b.lo[k] + b.hi[k]) - lo) / ext : 0.f;
		float s = u * 2097152.f;
		s = s < 0.f ? 0.f : (s > 2097151.f ? 2097151.f : s);
		q[k] = (uint32_t)s;
	}
	return (spread21(q[0]) << 2) | (spread21(q[1]) << 1) | spread21(q[2]);
}

LRH_DN inline int clz64(uint64_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
	return __clzll((long long)x);
#else
	return x ? __builtin_clzll(x) : 64;
#endif
}
LRH_DN inline int clz32(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
	return __clz((int)x);
#else
	return x ? __builtin_clz(x) : 32;
#endif
}

// common prefix length of keys i and j (positions break ties); -1 outside [0, n)
LRH_DN inline int delta(const uint64_t *k, int64_t n, int64_t i, int64_t j) {
	if (j < 0 || j >= n) return -1;
	const uint64_t a = k[i], b = k[j];
	if (a != b) return clz64(a ^ b);
	return 64 + clz32((uint32_t)i ^ (uint32_t)j);
}

// 3. internal node i of the radix tree: children lc / rc (kLeafBit | leaf, or
// an internal index) over the key range [lo, hi].  With keys in order the
// range search finds at least one more key and each loop ends within 33
// steps; false when the keys around i are not in order.
LRH_DN inline bool radixNode(const uint64_t *keys, int64_t n, int64_t i, uint32_t &lc, uint32_t &rc, uint32_t &lo,
		uint32_t &hi) {
	const int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
	const int dmin = delta(keys, n, i, i - d);
	int64_t lmax = 2;
	for (int steps = 0; delta(keys, n, i, i + lmax * d) > dmin; ++steps) {
		if (steps >= 33) return false;
		lmax *= 2;
	}
	int64_t l = 0;
	for (int64_t t = lmax / 2; t >= 1; t /= 2)
		if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
	if (l < 1) return false;  // no key beyond i shares more than dmin bits: out of order
	const int64_t j = i + l * d;
	const int dnode = delta(keys, n, i, j);
	int64_t s = 0;
	// split: t = ceil(l/2), ceil(t/2), ..., 1
	for (int64_t t = (l + 1) / 2, steps = 0;; t = (t + 1) / 2, ++steps) {
		if (steps >= 40) return false;
		if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
		if (t <= 1) break;
	}
	const int64_t g = i + s * d + (d < 0 ? -1 : 0);
	const int64_t a = i < j ? i : j, b = i < j ? j : i;
	if (g < a || g + 1 > b) return false;
	lc = (a == g) ? (kLeafBit | (uint32_t)g) : (uint32_t)g;
	rc = (b == g + 1) ? (kLeafBit | (uint32_t)(g + 1)) : (uint32_t)(g + 1);
	lo = (uint32_t)a;
	hi = (uint32_t)b;
	return true;
}

LRH_DN inline uint32_t countOf(uint32_t c, const uint32_t *first, const uint32_t *last) {
	return (c & kLeafBit) ? 1u : last[c] - first[c] + 1u;
}

LRH_DN inline float areaOf(const GBox &b) {
	const float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
	if (!(dx >= 0.f)) return 0.f;
	return 2.f * (dx * dy + dy * dz + dz * dx);
}

// surface-area cost of internal node x of the binary tree (N triangles, box
// area A, children costs cl, cr; a one-triangle leaf costs its box area):
// one leaf of its N triangles when N <= maxLeaf and that costs no more than
// a node of cost ci per area plus its children's costs
LRH_DN inline float nodeCost(float A, float cl, float cr, uint32_t N, uint32_t maxLeaf, float ci, uint32_t &leaf) {
	const float split = ci * A + (cl + cr);
	const float asLeaf = (float)N * A;
	leaf = (N <= maxLeaf && asLeaf <= split) ? 1u : 0u;
	return leaf ? asLeaf : split;
}

// a child that becomes a leaf of the wide tree: a triangle, or a node the
// cost chose as a leaf
LRH_DN inline bool leafChild(uint32_t c, const uint32_t *leafOK) { return (c & kLeafBit) || leafOK[c]; }

// 4a. the children of internal node b (Morton order): its two, then the
// largest-area child that is not a leaf split until there are `width` (4 or
// 8); returns how many of them become inner nodes
constexpr int kMaxWide = 8;
LRH_DN inline uint32_t expandNode(uint32_t b, const uint32_t *left, const uint32_t *right, const uint32_t *leafOK,
		const GBox *nodeBox, int width, uint32_t kids[kMaxWide]) {
	uint32_t ch[kMaxWide] = {left[b], right[b], kNone, kNone, kNone, kNone, kNone, kNone};
	int nch = 2;
	while (nch < width) {
		int pick = -1;
		float best = -1.f;
		for (int k = 0; k < nch; ++k) {
			if (leafChild(ch[k], leafOK)) continue;
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
	for (int k = 0; k < kMaxWide; ++k) {
		kids[k] = ch[k];
		if (ch[k] != kNone && !leafChild(ch[k], leafOK)) ++inner;
	}
	return inner;
}

LRH_DN inline uint32_t leafRef(uint32_t start, uint32_t count) {
	return LRH_CHILD_LEAF | ((count - 1u) << LRH_LEAF_START_BITS) | start;
}

// 4b. the wide node of a frontier entry: leaf children reference their key
// range, inner children the slots innerOff, innerOff+1, ... after the level
// (their binary ids go to next[slot])
LRH_DN inline void emitNode(const uint32_t kids[kMaxWide], uint32_t innerOff, const uint32_t *first,
		const uint32_t *last, const uint32_t *leafOK, const GBox *leafBox, const GBox *nodeBox, uint32_t childBase,
		uint32_t layout, void *nodes, uint32_t nodeIndex, uint32_t *next) {
	GBox cb[kMaxWide];
	uint32_t ref[kMaxWide];
	for (int k = 0; k < kMaxWide; ++k) ref[k] = LRH_CHILD_EMPTY;
	int n = 0;
	uint32_t slot = innerOff;
	for (int k = 0; k < kMaxWide; ++k) {
		const uint32_t c = kids[k];
		if (c == kNone) continue;
		cb[n] = (c & kLeafBit) ? leafBox[c & ~kLeafBit] : nodeBox[c];
		if (leafChild(c, leafOK)) {
			const uint32_t start = (c & kLeafBit) ? (c & ~kLeafBit) : first[c];
			ref[n] = leafRef(start, countOf(c, first, last));
		} else {
			ref[n] = childBase + slot;
			next[slot] = c;
			++slot;
		}
		++n;
	}
	if (layout == 0)
		lrhdev::quantizeNode(cb, n, ref, reinterpret_cast<lrh_node4q *>(nodes) + nodeIndex);
	else if (layout == 2)
		lrhdev::quantizeNode8(cb, n, ref, reinterpret_cast<lrh_node8q *>(nodes) + nodeIndex);
	else
		lrhdev::floatNode(cb, n, ref, reinterpret_cast<lrh_node4f *>(nodes) + nodeIndex);
}

// 5. triangle record of global triangle i, as the host builder writes it
LRH_DN inline lrh_tri triRecord(const float *verts, const uint32_t *tris, const uint32_t *tbase, const uint32_t *vbase,
		uint32_t nMeshes, uint32_t i) {
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
	return r;
}

LRH_DN inline uint32_t digitOf(uint64_t key, uint32_t shift) { return (uint32_t)(key >> shift) & 0xffu; }

// ---- agglomerative clustering of the Morton-ordered boxes (PLOC: Meister &
// Bittner, "Parallel Locally-Ordered Clustering for Bounding Volume
// Hierarchy Construction", IEEE TVCG 2018).  Each round every cluster finds
// its nearest neighbour within kPlocRadius places (the smallest merged box
// area); mutual pairs merge into a new binary node at the lower place and the
// array is compacted in order.  The binary tree then gets contiguous leaf
// ranges (subtree sizes bottom-up, first places top-down) and is collapsed
// like the radix tree. ----
constexpr uint32_t kPlocRadius = 16;

LRH_DN inline GBox unionBox(const GBox &a, const GBox &b) {
	GBox u;
	for (int k = 0; k < 3; ++k) {
		u.lo[k] = minT(a.lo[k], b.lo[k]);
		u.hi[k] = maxT(a.hi[k], b.hi[k]);
	}
	return u;
}

// area of the merged box, a NaN (boxes of NaN vertices) as +inf
LRH_DN inline float mergedArea(const GBox &a, const GBox &b) {
	const float x = areaOf(unionBox(a, b));
	return x == x ? x : INFINITY;
}

// nearest neighbour of cluster i among places [i - r, i + r]: the smallest
// merged area, equal areas by the pair's (lower, upper) place, so the pair
// with the smallest such key of all is mutual and every round merges
// (box(j) returns cluster j's box: the array itself, or a tile of it staged
// in LDS; the arithmetic is the same either way)
template <class BoxAt>
LRH_DN inline uint32_t plocNearestAt(const BoxAt &box, uint32_t nc, uint32_t i, uint32_t r) {
	const uint32_t lo = i > r ? i - r : 0u;
	const uint32_t hi = (nc - 1u - i > r) ? i + r : nc - 1u;
	float best = INFINITY;
	uint32_t arg = i;
	const GBox ci = box(i);
	for (uint32_t j = lo; j <= hi; ++j) {
		if (j == i) continue;
		const float a = mergedArea(ci, box(j));
		bool better = arg == i || a < best;
		if (arg != i && a == best) {
			const uint32_t p0 = i < j ? i : j, q0 = i < j ? j : i;
			const uint32_t p1 = i < arg ? i : arg, q1 = i < arg ? arg : i;
			better = p0 < p1 || (p0 == p1 && q0 < q1);
		}
		if (better) {
			best = a;
			arg = j;
		}
	}
	return arg;
}

LRH_DN inline uint32_t plocNearest(const GBox *c, uint32_t nc, uint32_t i, uint32_t r) {
	return plocNearestAt([c](uint32_t j) -> const GBox & { return c[j]; }, nc, i, r);
}

// cluster i's partner this round (kNone: none): its nearest neighbour when
// that is mutual; in a forced round (kept for inputs on which the
// neighbour rounds stall) the places 2k and 2k+1
LRH_DN inline uint32_t plocPartner(const uint32_t *nn, uint32_t nc, uint32_t i, bool forced) {
	if (forced) {
		const uint32_t j = i ^ 1u;
		return j < nc ? j : kNone;
	}
	const uint32_t j = nn[i];
	return (j != i && nn[j] == i) ? j : kNone;
}

LRH_DN inline uint32_t plocSize(uint32_t c, const uint32_t *count) { return (c & kLeafBit) ? 1u : count[c]; }

}  // namespace lrhlbvh

#endif
