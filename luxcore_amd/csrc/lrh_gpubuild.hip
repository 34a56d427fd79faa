/*
 * BVH builds on the device (linear BVH and PLOC), written into the 4-wide or
 * 8-wide node records of lrh_layout.h.  The per-element steps are
 * lrh_lbvh.h's (shared with the CPU emulation); here they run as kernels:
 *   1. leaf boxes and the bounds of their centres;
 *   2. 63-bit Morton codes over a cube of cells, sorted with the triangle id
 *      as value by a stable LSD radix sort (8-bit digits: per-tile digit
 *      counts, one exclusive scan of the [digit][tile] table, a scatter that
 *      ranks equal digits in index order with wave ballots, stages the tile
 *      in LDS in sorted order and writes each digit's run coalesced), then
 *      checked to be in order;
 *   3. a binary tree over the sorted leaves: the radix tree (LBVH) or
 *      agglomerative clustering in rounds of mutual nearest neighbours
 *      within 16 places (PLOC), brought into the radix tree's form;
 *   4. boxes bottom-up -- LBVH: one launch per depth (depths by pointer
 *      jumping, nodes ordered by depth); PLOC: one launch per round -- and
 *      surface-area costs, which decide whether a node of at most maxLeaf
 *      triangles becomes one leaf;
 *   5. breadth-first collapse into 4- or 8-wide nodes, child slots from an
 *      exclusive scan, so the node order is deterministic;
 *   6. triangle records in leaf order, bit for bit the host records.
 * Every kernel runs a bounded amount of work: no block waits on another
 * (the sort and the scans are sequences of kernels, not decoupled
 * look-back), every search loop has a step bound, and a violated
 * precondition sets a bit of an error word (or fails a host check of a
 * count read back) that the host reports as an error instead of going on.
 */
#include <hip/hip_runtime.h>

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "lrh_devnode.h"
#include "lrh_gpubuild.h"
#include "lrh_layout.h"
#include "lrh_lbvh.h"

namespace {

using namespace lrhlbvh;

#define GB_CHECK(x)                                                                                   \
	do {                                                                                              \
		const hipError_t e_ = (x);                                                                    \
		if (e_ != hipSuccess) throw std::runtime_error(std::string("device build: ") + hipGetErrorString(e_)); \
	} while (0)

__global__ void k_fill(uint32_t *p, uint32_t n, uint32_t v) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) p[i] = v;
}

__global__ void k_initBounds(uint32_t *cb) {
	if (threadIdx.x < 3) {
		cb[threadIdx.x] = 0xffffffffu;  // min keys
		cb[3 + threadIdx.x] = 0u;       // max keys
	}
}

// 1. leaf boxes and centre bounds: a grid-stride loop, the bounds reduced
// in registers, across the wave by lane exchanges and across the block in
// LDS, then six global atomics per block
constexpr uint32_t kPrimBlocks = 4096;
__device__ __forceinline__ uint32_t minU(uint32_t a, uint32_t b) { return a < b ? a : b; }
__device__ __forceinline__ uint32_t maxU(uint32_t a, uint32_t b) { return a > b ? a : b; }

__global__ __launch_bounds__(256) void k_prims(const float *verts, const uint32_t *tris, const uint32_t *tbase,
		const uint32_t *vbase, uint32_t nMeshes, uint32_t n, GBox *boxes, uint32_t *cb) {
	uint32_t mn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, mx[3] = {0u, 0u, 0u};
	for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
		const GBox b = primBox(verts, tris, tbase, vbase, nMeshes, i);
		boxes[i] = b;
		for (int k = 0; k < 3; ++k) {
			const uint32_t c = lrhdev::fkey(b.lo[k] + b.hi[k]);
			mn[k] = minU(mn[k], c);
			mx[k] = maxU(mx[k], c);
		}
	}
	for (int k = 0; k < 3; ++k)
		for (int o = 32; o > 0; o >>= 1) {
			mn[k] = minU(mn[k], (uint32_t)__shfl_xor((int)mn[k], o));
			mx[k] = maxU(mx[k], (uint32_t)__shfl_xor((int)mx[k], o));
		}
	__shared__ uint32_t smin[4][3], smax[4][3];
	const uint32_t wave = threadIdx.x >> 6;
	if ((threadIdx.x & 63u) == 0)
		for (int k = 0; k < 3; ++k) {
			smin[wave][k] = mn[k];
			smax[wave][k] = mx[k];
		}
	__syncthreads();
	if (threadIdx.x < 3) {
		const uint32_t k = threadIdx.x;
		uint32_t a = smin[0][k], b = smax[0][k];
		for (uint32_t w = 1; w < blockDim.x / 64; ++w) {
			a = minU(a, smin[w][k]);
			b = maxU(b, smax[w][k]);
		}
		atomicMin(&cb[k], a);
		atomicMax(&cb[3 + k], b);
	}
}

// 2. Morton codes of the box centres
__global__ void k_morton(const GBox *boxes, uint32_t n, const uint32_t *cbIn, uint64_t *keys, uint32_t *vals) {
	__shared__ uint32_t cb[6];
	if (threadIdx.x < 6) cb[threadIdx.x] = cbIn[threadIdx.x];
	__syncthreads();
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	keys[i] = mortonKey(boxes[i], cb);
	vals[i] = i;
}

// ---- radix sort pass: per-tile digit counts, table [digit][tile] ----
__global__ __launch_bounds__(kSortBlock) void k_sortCount(const uint64_t *keys, uint32_t n, uint32_t shift,
		uint32_t nTiles, uint32_t *table) {
	__shared__ uint32_t cnt[256];
	cnt[threadIdx.x] = 0;
	__syncthreads();
	const uint64_t base = (uint64_t)blockIdx.x * kSortTile;
	for (uint32_t j = threadIdx.x; j < kSortTile; j += kSortBlock) {
		const uint64_t idx = base + j;
		if (idx < n) atomicAdd(&cnt[digitOf(keys[idx], shift)], 1u);
	}
	__syncthreads();
	table[(size_t)threadIdx.x * nTiles + blockIdx.x] = cnt[threadIdx.x];
}

// ---- radix sort pass: stable scatter.  A tile is taken in rounds of 256
// consecutive keys, one per thread in index order; a key's place within the
// tile is its digit's start in the tile (a scan of the tile's digit counts)
// + the tile's earlier keys of that digit (earlier rounds) + the earlier
// waves' keys of that digit in this round + its rank among the lanes of its
// wave with the same digit (ballots over the eight digit bits).  Keys and
// values are first placed in LDS in that order, then written out in place
// order, so consecutive threads write consecutive addresses of one digit's
// run (its global start from the scanned [digit][tile] table). ----
__global__ __launch_bounds__(kSortBlock) void k_sortScatter(const uint64_t *keysIn, const uint32_t *valsIn, uint32_t n,
		uint32_t shift, uint32_t nTiles, const uint32_t *offsets, uint64_t *keysOut, uint32_t *valsOut, uint32_t *err) {
	constexpr uint32_t kWaves = kSortBlock / 64;
	__shared__ uint32_t run[256];
	__shared__ uint32_t wcnt[kWaves][256];
	__shared__ uint32_t off[256];    // global start of each digit's run for this tile
	__shared__ uint32_t start[256];  // each digit's start within the tile
	__shared__ uint64_t sKey[kSortTile];
	__shared__ uint32_t sVal[kSortTile];
	const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63u;
	const uint64_t base = (uint64_t)blockIdx.x * kSortTile;
	const uint32_t tileN = (uint32_t)((n - base) < kSortTile ? (n - base) : kSortTile);
	// the tile's digit counts, scanned to starts
	run[tid] = 0;
	for (uint32_t w = 0; w < kWaves; ++w) wcnt[w][tid] = 0;
	off[tid] = offsets[(size_t)tid * nTiles + blockIdx.x];
	__syncthreads();
	for (uint32_t j = tid; j < tileN; j += kSortBlock) atomicAdd(&run[digitOf(keysIn[base + j], shift)], 1u);
	__syncthreads();
	start[tid] = run[tid];
	__syncthreads();
	for (uint32_t o = 1; o < 256; o <<= 1) {  // inclusive scan (Hillis-Steele)
		const uint32_t add = tid >= o ? start[tid - o] : 0u;
		__syncthreads();
		start[tid] += add;
		__syncthreads();
	}
	start[tid] -= run[tid];  // exclusive
	run[tid] = 0;
	__syncthreads();
	const uint64_t lanesBelow = (1ull << lane) - 1ull;
	for (uint32_t r = 0; r < kSortTile / kSortBlock; ++r) {
		const uint32_t j = r * kSortBlock + tid;
		const bool valid = j < tileN;
		const uint64_t key = valid ? keysIn[base + j] : 0ull;
		const uint32_t val = valid ? valsIn[base + j] : 0u;
		const uint32_t dg = valid ? digitOf(key, shift) : 0u;
		uint64_t same = __ballot(valid);
		for (uint32_t b = 0; b < 8; ++b) {
			const bool bit = (dg >> b) & 1u;
			const uint64_t m = __ballot(bit);
			same &= bit ? m : ~m;
		}
		const uint32_t rank = (uint32_t)__popcll(same & lanesBelow);
		if (valid && rank == 0) wcnt[wave][dg] = (uint32_t)__popcll(same);
		__syncthreads();
		if (valid) {
			uint32_t before = run[dg];
			for (uint32_t w = 0; w < wave; ++w) before += wcnt[w][dg];
			const uint32_t p = start[dg] + before + rank;
			if (p < kSortTile) {
				sKey[p] = key;
				sVal[p] = val;
			} else {
				atomicOr(err, kErrUnsorted);  // a place outside the tile: never written
			}
		}
		__syncthreads();
		uint32_t add = 0;
		for (uint32_t w = 0; w < kWaves; ++w) {
			add += wcnt[w][tid];
			wcnt[w][tid] = 0;
		}
		run[tid] += add;
		__syncthreads();
	}
	// out in place order: one digit's run is contiguous in the tile and in
	// the output
	for (uint32_t p = tid; p < tileN; p += kSortBlock) {
		const uint64_t key = sKey[p];
		const uint32_t dg = digitOf(key, shift);
		const uint32_t pos = off[dg] + (p - start[dg]);
		if (pos < n) {
			keysOut[pos] = key;
			valsOut[pos] = sVal[p];
		} else {
			atomicOr(err, kErrUnsorted);
		}
	}
}

// ---- exclusive scan of uint32 values: per-tile scans, then the tile
// totals scanned the same way, then added ----
__global__ __launch_bounds__(256) void k_scanTile(const uint32_t *in, uint32_t *out, uint32_t n, uint32_t *tileSums) {
	constexpr uint32_t kPer = kScanTile / 256;
	__shared__ uint32_t part[256];
	const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kPer;
	uint32_t v[kPer];
	uint32_t sum = 0;
	for (uint32_t k = 0; k < kPer; ++k) {
		v[k] = (base + k < n) ? in[base + k] : 0u;
		sum += v[k];
	}
	part[threadIdx.x] = sum;
	__syncthreads();
	// inclusive scan of the 256 partial sums (Hillis-Steele)
	for (uint32_t o = 1; o < 256; o <<= 1) {
		const uint32_t add = threadIdx.x >= o ? part[threadIdx.x - o] : 0u;
		__syncthreads();
		part[threadIdx.x] += add;
		__syncthreads();
	}
	uint32_t run = part[threadIdx.x] - sum;
	for (uint32_t k = 0; k < kPer; ++k) {
		if (base + k < n) out[base + k] = run;
		run += v[k];
	}
	if (threadIdx.x == 255) tileSums[blockIdx.x] = part[255];
}

__global__ void k_scanAdd(uint32_t *out, uint32_t n, const uint32_t *tileOffsets) {
	const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) out[i] += tileOffsets[i / kScanTile];
}

__global__ void k_checkSorted(const uint64_t *keys, uint32_t n, uint32_t *err) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i + 1 < n && keys[i] > keys[i + 1]) atomicOr(err, kErrUnsorted);
}

// 3. binary radix tree
__global__ void k_radixTree(const uint64_t *keys, uint32_t n, uint32_t *left, uint32_t *right, uint32_t *first,
		uint32_t *last, uint32_t *iparent, uint32_t *err) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i + 1 >= n) return;
	uint32_t lc, rc, lo, hi;
	if (!radixNode(keys, n, i, lc, rc, lo, hi)) {
		atomicOr(err, kErrRange);
		return;
	}
	left[i] = lc;
	right[i] = rc;
	first[i] = lo;
	last[i] = hi;
	if (!(lc & kLeafBit)) iparent[lc] = i;
	if (!(rc & kLeafBit)) iparent[rc] = i;
}

// 4. node depths by pointer jumping: after round r every node knows its
// distance to the root if that is below 2^(r+1), and jump[x] is its ancestor
// 2^(r+1) levels up, or the root
__global__ void k_depthInit(const uint32_t *iparent, uint32_t ni, uint32_t *depth, uint32_t *jump) {
	const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
	if (x >= ni) return;
	const uint32_t p = iparent[x];
	depth[x] = (x == 0 || p >= ni) ? 0u : 1u;
	jump[x] = (x == 0 || p >= ni) ? 0u : p;
}

__global__ void k_depthJump(uint32_t ni, const uint32_t *depth, const uint32_t *jump, uint32_t *depth2,
		uint32_t *jump2) {
	const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
	if (x >= ni) return;
	const uint32_t j = jump[x];
	depth2[x] = depth[x] + depth[j];
	jump2[x] = jump[j];
}

// a node not yet at the root after the doublings, or deeper than any valid
// tree: an error
constexpr uint32_t kMaxDepth = 128;
__global__ void k_depthCheck(uint32_t ni, const uint32_t *depth, const uint32_t *jump, uint32_t *err) {
	const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
	if (x < ni && (jump[x] != 0 || depth[x] >= kMaxDepth)) atomicOr(err, kErrClimb);
}

// the first place of each depth (digit) in the sorted order: the scanned
// digit table's entry of the first tile; out[kMaxDepth] counts the nodes
// of depth below kMaxDepth
__global__ void k_digitStarts(const uint32_t *table, uint32_t tiles, uint32_t *out) {
	const uint32_t d = threadIdx.x;
	if (d <= kMaxDepth) out[d] = table[(size_t)d * tiles];
}

// sort keys (depths) and values (node ids) for the counting sort by depth
__global__ void k_depthKeys(uint32_t ni, const uint32_t *depth, uint64_t *keys, uint32_t *ids) {
	const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
	if (x >= ni) return;
	keys[x] = depth[x];
	ids[x] = x;
}

__global__ void k_leafBoxes(const GBox *boxes, const uint32_t *vals, uint32_t n, GBox *leafBox) {
	const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j < n) leafBox[j] = boxes[vals[j]];
}

// boxes and surface-area costs of the nodes of one depth, deepest level
// first: the children's are the previous launch's
__global__ void k_boxLevel(const uint32_t *order, uint32_t cnt, const uint32_t *left, const uint32_t *right,
		const uint32_t *first, const uint32_t *last, const GBox *leafBox, uint32_t maxLeaf, float ci, GBox *nodeBox,
		float *cost, uint32_t *leafOK) {
	const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= cnt) return;
	const uint32_t x = order[e], l = left[x], r = right[x];
	const GBox a = (l & kLeafBit) ? leafBox[l & ~kLeafBit] : nodeBox[l];
	const GBox b = (r & kLeafBit) ? leafBox[r & ~kLeafBit] : nodeBox[r];
	const GBox u = unionBox(a, b);
	const float cl = (l & kLeafBit) ? areaOf(a) : cost[l];
	const float cr = (r & kLeafBit) ? areaOf(b) : cost[r];
	uint32_t lf;
	cost[x] = nodeCost(areaOf(u), cl, cr, last[x] - first[x] + 1u, maxLeaf, ci, lf);
	leafOK[x] = lf;
	nodeBox[x] = u;
}

// 3'. clustering (PLOC): clusters start as the sorted leaves
__global__ void k_plocInit(const GBox *boxes, const uint32_t *vals, uint32_t n, uint32_t *cid, GBox *cbox,
		GBox *leafBox) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const GBox b = boxes[vals[i]];
	leafBox[i] = b;
	cid[i] = kLeafBit | i;
	cbox[i] = b;
}

// nearest neighbours of a tile of 256 clusters, their boxes and the
// kPlocRadius on either side staged in LDS (each box is read by 2r+1 threads)
constexpr uint32_t kPlocTile = 256;
__global__ __launch_bounds__(kPlocTile) void k_plocNearest(const GBox *cbox, uint32_t nc, uint32_t *nn) {
	constexpr uint32_t kSpan = kPlocTile + 2 * kPlocRadius;
	__shared__ float tile[kSpan * 6];
	const uint32_t t0 = blockIdx.x * kPlocTile;
	const int64_t base = (int64_t)t0 - kPlocRadius;  // cluster of tile entry 0
	const float *src = reinterpret_cast<const float *>(cbox);
	for (uint32_t k = threadIdx.x; k < kSpan * 6; k += kPlocTile) {
		const int64_t j = base + k / 6;
		tile[k] = (j >= 0 && j < (int64_t)nc) ? src[j * 6 + k % 6] : 0.f;
	}
	__syncthreads();
	const uint32_t i = t0 + threadIdx.x;
	if (i >= nc) return;
	const GBox *lds = reinterpret_cast<const GBox *>(tile);
	nn[i] = plocNearestAt([&](uint32_t j) -> const GBox & { return lds[(int64_t)j - base]; }, nc, i, kPlocRadius);
}

// lead: the lower place of a merging pair (a new node); keep: the place
// survives the round
__global__ void k_plocFlags(const uint32_t *nn, uint32_t nc, uint32_t forced, uint32_t *lead, uint32_t *keep) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i > nc) return;
	if (i == nc) {  // the scan's total slot
		lead[i] = 0;
		keep[i] = 0;
		return;
	}
	const uint32_t j = plocPartner(nn, nc, i, forced != 0);
	lead[i] = (j != kNone && i < j) ? 1u : 0u;
	keep[i] = (j != kNone && i > j) ? 0u : 1u;
}

__global__ void k_plocApply(const uint32_t *nn, uint32_t nc, uint32_t forced, const uint32_t *cid, const GBox *cbox,
		const uint32_t *leadOff, const uint32_t *keepOff, uint32_t nodeBase, uint32_t *outCid, GBox *outBox,
		uint32_t *left, uint32_t *right, GBox *nodeBox) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nc) return;
	const uint32_t j = plocPartner(nn, nc, i, forced != 0);
	if (j != kNone && i > j) return;  // merged into its lower partner
	const uint32_t o = keepOff[i];
	if (j != kNone) {
		const uint32_t node = nodeBase + leadOff[i];
		const GBox u = unionBox(cbox[i], cbox[j]);
		left[node] = cid[i];
		right[node] = cid[j];
		nodeBox[node] = u;
		outCid[o] = node;
		outBox[o] = u;
	} else {
		outCid[o] = cid[i];
		outBox[o] = cbox[i];
	}
}

// leaves per subtree and surface-area costs, the nodes of one round (their
// children are older)
__global__ void k_plocCount(uint32_t b, uint32_t e, const uint32_t *left, const uint32_t *right,
		const GBox *leafBox, const GBox *nodeBox, uint32_t maxLeaf, float ci, uint32_t *count, float *cost,
		uint32_t *leafOK) {
	const uint32_t x = b + blockIdx.x * blockDim.x + threadIdx.x;
	if (x >= e) return;
	const uint32_t l = left[x], r = right[x];
	const uint32_t c = plocSize(l, count) + plocSize(r, count);
	count[x] = c;
	const float cl = (l & kLeafBit) ? areaOf(leafBox[l & ~kLeafBit]) : cost[l];
	const float cr = (r & kLeafBit) ? areaOf(leafBox[r & ~kLeafBit]) : cost[r];
	uint32_t lf;
	cost[x] = nodeCost(areaOf(nodeBox[x]), cl, cr, c, maxLeaf, ci, lf);
	leafOK[x] = lf;
}

// first leaf place of each subtree, top down (the nodes of one round; their
// parents are younger); leafFirst: the place of each sorted leaf
__global__ void k_plocFirst(uint32_t b, uint32_t e, const uint32_t *left, const uint32_t *right, const uint32_t *count,
		uint32_t *first, uint32_t *leafFirst) {
	const uint32_t x = b + blockIdx.x * blockDim.x + threadIdx.x;
	if (x >= e) return;
	const uint32_t f = first[x], l = left[x], r = right[x];
	const uint32_t fr = f + plocSize(l, count);
	if (l & kLeafBit) leafFirst[l & ~kLeafBit] = f; else first[l] = f;
	if (r & kLeafBit) leafFirst[r & ~kLeafBit] = fr; else first[r] = fr;
}

// the clustering's tree in the radix tree's form: internal node x becomes
// n-2-x (root 0), leaves their places; leaf values and boxes in place order
__global__ void k_plocFinish(uint32_t n, const uint32_t *left, const uint32_t *right, const uint32_t *count,
		const uint32_t *first, const GBox *nodeBox, const uint32_t *leafOK, const uint32_t *leafFirst,
		const uint32_t *vals, const GBox *leafBoxIn, uint32_t *left2, uint32_t *right2, uint32_t *first2,
		uint32_t *last2, GBox *nodeBox2, uint32_t *leafOK2, uint32_t *vals2, GBox *leafBox2) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const uint32_t p = leafFirst[i];
	vals2[p] = vals[i];
	leafBox2[p] = leafBoxIn[i];
	if (i + 1 >= n) return;  // n - 1 internal nodes
	const uint32_t y = n - 2 - i;
	auto remap = [&](uint32_t c) { return (c & kLeafBit) ? (kLeafBit | leafFirst[c & ~kLeafBit]) : n - 2 - c; };
	left2[y] = remap(left[i]);
	right2[y] = remap(right[i]);
	first2[y] = first[i];
	last2[y] = first[i] + count[i] - 1;
	nodeBox2[y] = nodeBox[i];
	leafOK2[y] = leafOK[i];
}

// 5a. children of each frontier node and how many become inner nodes
__global__ void k_expand(const uint32_t *frontier, uint32_t nf, uint32_t nInner, const uint32_t *left,
		const uint32_t *right, const uint32_t *leafOK, const GBox *nodeBox, int width, uint32_t *kids,
		uint32_t *innerCount, uint32_t *err) {
	const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= nf) return;
	const uint32_t b = frontier[e];
	if (b >= nInner) {  // not an internal node
		atomicOr(err, kErrExpand);
		for (int k = 0; k < kMaxWide; ++k) kids[(size_t)kMaxWide * e + k] = kNone;
		innerCount[e] = 0;
		return;
	}
	uint32_t ch[kMaxWide];
	innerCount[e] = expandNode(b, left, right, leafOK, nodeBox, width, ch);
	for (int k = 0; k < kMaxWide; ++k) kids[(size_t)kMaxWide * e + k] = ch[k];
}

// 5b. the frontier's wide nodes; inner children get slots after the level
__global__ void k_emit(uint32_t nf, const uint32_t *kids, const uint32_t *innerOff, const uint32_t *first,
		const uint32_t *last, const uint32_t *leafOK, const GBox *leafBox, const GBox *nodeBox, uint32_t levelBase,
		uint32_t layout, void *nodes, uint32_t *next) {
	const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= nf) return;
	uint32_t ch[kMaxWide];
	for (int k = 0; k < kMaxWide; ++k) ch[k] = kids[(size_t)kMaxWide * e + k];
	emitNode(ch, innerOff[e], first, last, leafOK, leafBox, nodeBox, levelBase + nf, layout, nodes, levelBase + e,
			next);
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
	uint32_t ref[kMaxWide];
	for (int k = 0; k < kMaxWide; ++k) ref[k] = k ? LRH_CHILD_EMPTY : leafRef(0, n);
	if (layout == 0)
		lrhdev::quantizeNode(&b, 1, ref, reinterpret_cast<lrh_node4q *>(nodes));
	else if (layout == 2)
		lrhdev::quantizeNode8(&b, 1, ref, reinterpret_cast<lrh_node8q *>(nodes));
	else
		lrhdev::floatNode(&b, 1, ref, reinterpret_cast<lrh_node4f *>(nodes));
}

__global__ void k_identity(const GBox *boxes, uint32_t n, uint32_t *vals, GBox *leafBox) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	vals[i] = i;
	leafBox[i] = boxes[i];
}

// 6. triangle records in leaf order
__global__ void k_records(const float *verts, const uint32_t *tris, const uint32_t *tbase, const uint32_t *vbase,
		uint32_t nMeshes, const uint32_t *vals, uint32_t n, lrh_tri *recs, uint32_t *invPerm) {
	const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= n) return;
	const uint32_t i = vals[j];
	recs[j] = triRecord(verts, tris, tbase, vbase, nMeshes, i);
	invPerm[i] = j;
}

// the build's scratch arrays, carved from chunks of at least 256 MiB (an
// array above that is a chunk of its own) and released together at the end:
// the clustering's rounds ask for hundreds of small scan arrays, and a
// device allocation and release each would cost more than the kernels
struct Scratch {
	static constexpr size_t kChunk = size_t(256) << 20;
	const LrhGpuAlloc &a;
	std::vector<std::pair<void *, size_t>> held;
	char *cur = nullptr;
	size_t left = 0;
	uint32_t *pin = nullptr;  // 2 KiB of pinned host memory for the readbacks (word 0: error, 1-3: counts, 128+: depths)
	explicit Scratch(const LrhGpuAlloc &al) : a(al) {
		GB_CHECK(hipHostMalloc(reinterpret_cast<void **>(&pin), 2048, hipHostMallocDefault));
	}
	template <class T>
	T *get(size_t count) {
		const size_t bytes = (std::max<size_t>(count * sizeof(T), 16) + 255) & ~size_t(255);
		if (bytes > left) {
			const size_t chunk = std::max(bytes, kChunk);
			void *p = a.alloc(a.ctx, chunk);
			held.push_back({p, chunk});
			if (bytes == chunk) return static_cast<T *>(p);  // a chunk of its own; the current one stays
			cur = static_cast<char *>(p);
			left = chunk;
		}
		T *r = reinterpret_cast<T *>(cur);
		cur += bytes;
		left -= bytes;
		return r;
	}
	~Scratch() {
		for (auto &h : held) a.release(a.ctx, h.first, h.second);
		(void)hipHostFree(pin);
	}
};

inline unsigned grid(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

// exclusive scan of n values in place of out (in may equal out); the tile
// sums are scanned by the same routine, so the depth is log_2048(n)
void exclusiveScan(Scratch &S, const uint32_t *in, uint32_t *out, uint32_t n, hipStream_t st) {
	if (n == 0) return;
	const uint32_t tiles = grid(n, kScanTile);
	uint32_t *sums = S.get<uint32_t>(tiles);
	hipLaunchKernelGGL(k_scanTile, dim3(tiles), dim3(256), 0, st, in, out, n, sums);
	GB_CHECK(hipGetLastError());
	if (tiles > 1) {
		exclusiveScan(S, sums, sums, tiles, st);
		hipLaunchKernelGGL(k_scanAdd, dim3(grid(n, 256)), dim3(256), 0, st, out, n, sums);
		GB_CHECK(hipGetLastError());
	}
}

// waits for the stream (and the readbacks queued before) and returns the
// error word
uint32_t readError(Scratch &S, const uint32_t *err, hipStream_t st) {
	GB_CHECK(hipMemcpyAsync(S.pin, err, 4, hipMemcpyDeviceToHost, st));
	GB_CHECK(hipStreamSynchronize(st));
	return S.pin[0];
}

std::string errorText(uint32_t e) {
	std::string s = "device build: ";
	if (e & kErrUnsorted) s += "sorted Morton codes out of order; ";
	if (e & kErrRange) s += "radix tree node without a key range (codes not in order); ";
	if (e & kErrClimb) s += "radix tree deeper than any valid tree (or not rooted); ";
	if (e & kErrExpand) s += "collapse reached a non-internal node; ";
	return s;
}

// boxes, costs and leaf decisions of the radix tree's nodes bottom-up with a
// kernel boundary between levels instead of a fenced walk per leaf (a walk
// publishes each box to the other XCDs' caches by an agent-scope fence,
// which writes back and invalidates L2 on every step): depths by pointer
// jumping, nodes sorted by depth, then one launch per level from the
// deepest.
void boxesByDepth(Scratch &S, const GBox *boxes, const uint32_t *vals, uint32_t n, const uint32_t *left,
		const uint32_t *right, const uint32_t *first, const uint32_t *last, const uint32_t *iparent, uint32_t maxLeaf,
		float ci, hipStream_t st, uint32_t *err, GBox *leafBox, GBox *nodeBox, uint32_t *leafOK) {
	const uint32_t ni = n - 1;
	uint32_t *dA = S.get<uint32_t>(ni), *dB = S.get<uint32_t>(ni), *jA = S.get<uint32_t>(ni),
			 *jB = S.get<uint32_t>(ni), *order = S.get<uint32_t>(ni), *starts = S.get<uint32_t>(kMaxDepth + 1);
	uint64_t *dKeys = S.get<uint64_t>(ni), *dKeysOut = S.get<uint64_t>(ni);
	float *cost = S.get<float>(ni);
	hipLaunchKernelGGL(k_depthInit, dim3(grid(ni, 256)), dim3(256), 0, st, iparent, ni, dA, jA);
	// depth < kMaxDepth = 2^7: seven doublings reach the root
	for (int r = 0; r < 7; ++r) {
		hipLaunchKernelGGL(k_depthJump, dim3(grid(ni, 256)), dim3(256), 0, st, ni, dA, jA, dB, jB);
		std::swap(dA, dB);
		std::swap(jA, jB);
	}
	hipLaunchKernelGGL(k_depthCheck, dim3(grid(ni, 256)), dim3(256), 0, st, ni, dA, jA, err);
	GB_CHECK(hipGetLastError());
	if (const uint32_t e = readError(S, err, st)) throw std::runtime_error(errorText(e));
	// nodes in depth order: one stable pass of the radix sort (depth < 2^8
	// is one digit) with the node ids as values, so no counter is contended;
	// the scanned digit table gives where each depth starts
	const uint32_t tiles = grid(ni, kSortTile);
	uint32_t *table = S.get<uint32_t>((size_t)256 * tiles), *ids = S.get<uint32_t>(ni);
	hipLaunchKernelGGL(k_depthKeys, dim3(grid(ni, 256)), dim3(256), 0, st, ni, dA, dKeys, ids);
	hipLaunchKernelGGL(k_sortCount, dim3(tiles), dim3(kSortBlock), 0, st, dKeys, ni, 0u, tiles, table);
	GB_CHECK(hipGetLastError());
	exclusiveScan(S, table, table, 256u * tiles, st);
	hipLaunchKernelGGL(k_sortScatter, dim3(tiles), dim3(kSortBlock), 0, st, dKeys, ids, ni, 0u, tiles, table, dKeysOut,
			order, err);
	hipLaunchKernelGGL(k_digitStarts, dim3(1), dim3(256), 0, st, table, tiles, starts);
	GB_CHECK(hipGetLastError());
	GB_CHECK(hipMemcpyAsync(S.pin + 128, starts, (kMaxDepth + 1) * 4, hipMemcpyDeviceToHost, st));
	if (const uint32_t e = readError(S, err, st)) throw std::runtime_error(errorText(e));
	uint32_t off[kMaxDepth + 1], h[kMaxDepth];
	std::copy(S.pin + 128, S.pin + 128 + kMaxDepth + 1, off);
	for (uint32_t d = 0; d < kMaxDepth; ++d) {
		if (off[d + 1] < off[d]) throw std::runtime_error("device build: depth table out of order");
		h[d] = off[d + 1] - off[d];
	}
	if (off[0] != 0 || off[kMaxDepth] != ni)
		throw std::runtime_error("device build: radix tree nodes not all below the root");
	hipLaunchKernelGGL(k_leafBoxes, dim3(grid(n, 256)), dim3(256), 0, st, boxes, vals, n, leafBox);
	GB_CHECK(hipGetLastError());
	for (uint32_t d = kMaxDepth; d-- > 0;) {
		const uint32_t cnt = h[d];
		if (!cnt) continue;
		hipLaunchKernelGGL(k_boxLevel, dim3(grid(cnt, 256)), dim3(256), 0, st, order + off[d], cnt, left, right, first,
				last, leafBox, maxLeaf, ci, nodeBox, cost, leafOK);
	}
	GB_CHECK(hipGetLastError());
	if (const uint32_t e = readError(S, err, st)) throw std::runtime_error(errorText(e));
}

// the clustering's binary tree over the n >= 2 sorted leaves, written in the
// radix tree's form (left / right / first / last / nodeBox, root 0, leaf
// places contiguous per subtree); vals and leafBox are repointed to the
// leaves in place order.  Each round: nearest neighbours, the merge and keep
// flags scanned, then the pairs merged and the clusters compacted.  The host
// reads the round's merge count; a round that merges fewer than 1/64 of the
// clusters is followed by a forced round that pairs places 2k, 2k+1, so the
// number of rounds is bounded by the size alone.  (Natural rounds merge a
// third of the clusters at first and a tenth or less near the top, where
// boxes overlap; forcing there would cost tree quality.)
void plocTree(Scratch &S, const GBox *boxes, uint32_t n, uint32_t maxLeaf, float ci, hipStream_t st, uint32_t *&vals,
		GBox *&leafBox, uint32_t *left, uint32_t *right, uint32_t *first, uint32_t *last, GBox *nodeBox,
		uint32_t *leafOK, uint32_t *roundsOut) {
	const uint32_t ni = n - 1;
	uint32_t *cidA = S.get<uint32_t>(n), *cidB = S.get<uint32_t>(n), *nn = S.get<uint32_t>(n);
	GBox *cboxA = S.get<GBox>(n), *cboxB = S.get<GBox>(n);
	uint32_t *lead = S.get<uint32_t>((size_t)n + 1), *keep = S.get<uint32_t>((size_t)n + 1),
			 *leadOff = S.get<uint32_t>((size_t)n + 1), *keepOff = S.get<uint32_t>((size_t)n + 1);
	hipLaunchKernelGGL(k_plocInit, dim3(grid(n, 256)), dim3(256), 0, st, boxes, vals, n, cidA, cboxA, leafBox);
	GB_CHECK(hipGetLastError());
	std::vector<uint32_t> roundStart{0};
	uint32_t nc = n, nodeBase = 0;
	bool forced = false;
	// fewer than 1/64 merged forces the next round, which halves: at most
	// 2 * 31 slow rounds, 63 single merges below 64 clusters, and
	// log(2^31) / log(64/63) others
	const uint32_t kMaxRounds = 4096;
	while (nc > 1) {
		if (roundStart.size() > kMaxRounds) throw std::runtime_error("device build: clustering did not converge");
		if (!forced)
			hipLaunchKernelGGL(k_plocNearest, dim3(grid(nc, kPlocTile)), dim3(kPlocTile), 0, st, cboxA, nc, nn);
		hipLaunchKernelGGL(k_plocFlags, dim3(grid((size_t)nc + 1, 256)), dim3(256), 0, st, nn, nc, forced ? 1u : 0u,
				lead, keep);
		GB_CHECK(hipGetLastError());
		exclusiveScan(S, lead, leadOff, nc + 1, st);
		exclusiveScan(S, keep, keepOff, nc + 1, st);
		hipLaunchKernelGGL(k_plocApply, dim3(grid(nc, 256)), dim3(256), 0, st, nn, nc, forced ? 1u : 0u, cidA, cboxA,
				leadOff, keepOff, nodeBase, cidB, cboxB, left, right, nodeBox);
		GB_CHECK(hipGetLastError());
		GB_CHECK(hipMemcpyAsync(S.pin + 1, leadOff + nc, 4, hipMemcpyDeviceToHost, st));
		GB_CHECK(hipMemcpyAsync(S.pin + 2, keepOff + nc, 4, hipMemcpyDeviceToHost, st));
		GB_CHECK(hipStreamSynchronize(st));
		const uint32_t merged = S.pin[1], kept = S.pin[2];
		if (merged == 0 || merged > nc / 2 || kept != nc - merged || nodeBase + merged > ni)
			throw std::runtime_error("device build: clustering round without a consistent merge count");
		nodeBase += merged;
		roundStart.push_back(nodeBase);
		forced = !forced && (uint64_t)merged * 64u < nc;
		nc = kept;
		std::swap(cidA, cidB);
		std::swap(cboxA, cboxB);
	}
	if (nodeBase != ni) throw std::runtime_error("device build: clustering left nodes unmade");
	*roundsOut = (uint32_t)roundStart.size() - 1;
	// leaves per subtree, oldest round first; first places, newest first
	uint32_t *count = S.get<uint32_t>(ni), *firstP = S.get<uint32_t>(ni), *leafFirst = S.get<uint32_t>(n),
			 *leafP = S.get<uint32_t>(ni);
	float *cost = S.get<float>(ni);
	for (size_t r = 0; r + 1 < roundStart.size(); ++r) {
		const uint32_t b = roundStart[r], e = roundStart[r + 1];
		hipLaunchKernelGGL(k_plocCount, dim3(grid(e - b, 256)), dim3(256), 0, st, b, e, left, right, leafBox, nodeBox,
				maxLeaf, ci, count, cost, leafP);
	}
	hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, st, firstP + (ni - 1), 1u, 0u);  // the root, node n-2
	for (size_t r = roundStart.size() - 1; r-- > 0;) {
		const uint32_t b = roundStart[r], e = roundStart[r + 1];
		hipLaunchKernelGGL(k_plocFirst, dim3(grid(e - b, 256)), dim3(256), 0, st, b, e, left, right, count, firstP,
				leafFirst);
	}
	GB_CHECK(hipGetLastError());
	// the radix tree's form, written into the caller's arrays; the binary
	// tree's own are copied aside first
	uint32_t *leftP = S.get<uint32_t>(ni), *rightP = S.get<uint32_t>(ni);
	GBox *boxP = S.get<GBox>(ni);
	GB_CHECK(hipMemcpyAsync(leftP, left, (size_t)ni * 4, hipMemcpyDeviceToDevice, st));
	GB_CHECK(hipMemcpyAsync(rightP, right, (size_t)ni * 4, hipMemcpyDeviceToDevice, st));
	GB_CHECK(hipMemcpyAsync(boxP, nodeBox, (size_t)ni * sizeof(GBox), hipMemcpyDeviceToDevice, st));
	uint32_t *vals2 = S.get<uint32_t>(n);
	GBox *leafBox2 = S.get<GBox>(n);
	hipLaunchKernelGGL(k_plocFinish, dim3(grid(n, 256)), dim3(256), 0, st, n, leftP, rightP, count, firstP, boxP,
			leafP, leafFirst, vals, leafBox, left, right, first, last, nodeBox, leafOK, vals2, leafBox2);
	GB_CHECK(hipGetLastError());
	vals = vals2;
	leafBox = leafBox2;
}

}  // namespace

void lrhGpuBuildBVH(const LrhGpuBuildIn &in, const LrhGpuAlloc &a, hipStream_t st, LrhGpuBuildOut *out) {
	const uint32_t n = in.nTris;
	if (n == 0) throw std::runtime_error("device build: empty scene");
	if (n > 0x7fffffffu) throw std::runtime_error("device build: at most 2^31-1 triangles");
	if (in.maxLeaf < 1 || in.maxLeaf > LRH_LEAF_MAX) throw std::runtime_error("device build: maxLeaf must be 1..8");
	if (!(in.travCost >= 0.f) || !std::isfinite(in.travCost))
		throw std::runtime_error("device build: travCost must be finite and >= 0");
	if (in.layout > 2) throw std::runtime_error("device build: layout must be 0, 1 or 2");
	const size_t nodeSz = in.layout == 0 ? sizeof(lrh_node4q) : in.layout == 2 ? sizeof(lrh_node8q) : sizeof(lrh_node4f);
	const int width = in.layout == 2 ? 8 : 4;
	Scratch S(a);
	hipEvent_t e0, e1;
	GB_CHECK(hipEventCreate(&e0));
	GB_CHECK(hipEventCreate(&e1));
	GB_CHECK(hipEventRecord(e0, st));

	uint32_t *err = S.get<uint32_t>(1);
	hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, st, err, 1u, 0u);
	uint32_t *tbase = S.get<uint32_t>(in.nMeshes + 1), *vbase = S.get<uint32_t>(in.nMeshes + 1);
	GB_CHECK(hipMemcpyAsync(tbase, in.meshTriBase, (in.nMeshes + 1) * 4, hipMemcpyHostToDevice, st));
	GB_CHECK(hipMemcpyAsync(vbase, in.meshVertBase, (in.nMeshes + 1) * 4, hipMemcpyHostToDevice, st));
	GBox *boxes = S.get<GBox>(n);
	uint32_t *cb = S.get<uint32_t>(6);
	hipLaunchKernelGGL(k_initBounds, dim3(1), dim3(64), 0, st, cb);
	hipLaunchKernelGGL(k_prims, dim3(std::min<unsigned>(grid(n, 256), kPrimBlocks)), dim3(256), 0, st, in.verts, in.tris, tbase, vbase, in.nMeshes, n,
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
			uint64_t *keys = S.get<uint64_t>(n), *keysB = S.get<uint64_t>(n);
			uint32_t *valsB = S.get<uint32_t>(n);
			hipLaunchKernelGGL(k_morton, dim3(grid(n, 256)), dim3(256), 0, st, boxes, n, cb, keys, vals);
			GB_CHECK(hipGetLastError());
			// stable LSD radix sort, 8 passes of 8 bits; an even number of
			// passes leaves the result in keys / vals
			const uint32_t tiles = grid(n, kSortTile);
			uint32_t *table = S.get<uint32_t>((size_t)256 * tiles);
			uint64_t *kIn = keys, *kOut = keysB;
			uint32_t *vIn = vals, *vOut = valsB;
			for (uint32_t shift = 0; shift < 64; shift += 8) {
				hipLaunchKernelGGL(k_sortCount, dim3(tiles), dim3(kSortBlock), 0, st, kIn, n, shift, tiles, table);
				GB_CHECK(hipGetLastError());
				exclusiveScan(S, table, table, 256u * tiles, st);
				hipLaunchKernelGGL(k_sortScatter, dim3(tiles), dim3(kSortBlock), 0, st, kIn, vIn, n, shift, tiles, table,
						kOut, vOut, err);
				GB_CHECK(hipGetLastError());
				std::swap(kIn, kOut);
				std::swap(vIn, vOut);
			}
			hipLaunchKernelGGL(k_checkSorted, dim3(grid(n, 256)), dim3(256), 0, st, keys, n, err);
			GB_CHECK(hipGetLastError());
			if (const uint32_t e = readError(S, err, st)) throw std::runtime_error(errorText(e));

			const uint32_t ni = n - 1;
			uint32_t *left = S.get<uint32_t>(ni), *right = S.get<uint32_t>(ni), *first = S.get<uint32_t>(ni),
					 *last = S.get<uint32_t>(ni);
			GBox *nodeBox = S.get<GBox>(ni);
			uint32_t *leafOK = S.get<uint32_t>(ni);
			if (in.method == 1) {
				plocTree(S, boxes, n, in.maxLeaf, in.travCost, st, vals, leafBox, left, right, first, last, nodeBox,
						leafOK, &out->plocRounds);
			} else {
				uint32_t *iparent = S.get<uint32_t>(ni);
				hipLaunchKernelGGL(k_fill, dim3(grid(ni, 256)), dim3(256), 0, st, iparent, ni, kNone);  // root: none
				hipLaunchKernelGGL(k_radixTree, dim3(grid(ni, 256)), dim3(256), 0, st, keys, n, left, right, first,
						last, iparent, err);
				GB_CHECK(hipGetLastError());
				if (const uint32_t e = readError(S, err, st)) throw std::runtime_error(errorText(e));
				boxesByDepth(S, boxes, vals, n, left, right, first, last, iparent, in.maxLeaf, in.travCost, st, err,
						leafBox, nodeBox, leafOK);
			}

			// breadth-first collapse: level frontiers of binary node ids
			uint32_t *fA = S.get<uint32_t>(ni), *fB = S.get<uint32_t>(ni);
			uint32_t *kids = S.get<uint32_t>((size_t)kMaxWide * ni), *cnt = S.get<uint32_t>(ni + 1),
					 *off = S.get<uint32_t>(ni + 1);
			hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, st, fA, 1u, 0u);  // the root
			uint32_t nf = 1, levelBase = 0;
			while (nf) {
				if ((size_t)levelBase + nf > capNodes) throw std::runtime_error("device build: node capacity");
				if (depth > 128) throw std::runtime_error("device build: collapse deeper than any valid tree");
				hipLaunchKernelGGL(k_expand, dim3(grid(nf, 256)), dim3(256), 0, st, fA, nf, ni, left, right, leafOK,
						nodeBox, width, kids, cnt, err);
				hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, st, cnt + nf, 1u, 0u);
				GB_CHECK(hipGetLastError());
				exclusiveScan(S, cnt, off, nf + 1, st);
				hipLaunchKernelGGL(k_emit, dim3(grid(nf, 256)), dim3(256), 0, st, nf, kids, off, first, last, leafOK,
						leafBox, nodeBox, levelBase, in.layout, nodes, fB);
				GB_CHECK(hipGetLastError());
				GB_CHECK(hipMemcpyAsync(S.pin + 3, off + nf, 4, hipMemcpyDeviceToHost, st));
				if (const uint32_t e = readError(S, err, st)) throw std::runtime_error(errorText(e));
				const uint32_t nextN = S.pin[3];
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
