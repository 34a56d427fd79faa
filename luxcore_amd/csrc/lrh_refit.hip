/*
 * Two-level root refit on the device: after an edit of instance matrices or
 * motion keys (MBVHAccel::Update, src/luxrays/accelerators/mbvhaccel.cpp:237-250,
 * which rebuilds the root tree on the host, and MBVHKernel::Update,
 * src/luxrays/accelerators/mbvhaccelhw.cpp:442-466, which re-uploads it), the
 * root tree keeps its topology and only its boxes are recomputed here:
 *
 *   1. motion objects' world boxes: MotionSystem::Bound(box, true)
 *      (src/luxrays/core/geometry/motionsystem.cpp:75-87, 356-365), i.e. the
 *      union over every segment of 1,025 samples of Sample(t).Inverse() * box,
 *      one thread per (object, segment, sample), combined with exact atomic
 *      min/max on ordered float keys, then expanded by MachineEpsilon::E
 *      (mbvhaccel.cpp:192-193); rounded as the host computes them (this file
 *      builds with -ffp-contract=off and correctly rounded division);
 *   2. the root nodes bottom-up, one launch per level, deepest first: each
 *      node takes the union of its children (object boxes for leaf slots,
 *      nodes of the level below) and is rewritten with the host builder's
 *      outward-rounded quantization.
 *
 * Instance and plain objects' boxes are cheap and come from the host.
 */
#include <hip/hip_runtime.h>

#include "lrh_devnode.h"
#include "lrh_layout.h"
#include "lrh_motion_dev.h"
#include "lrh_refit.h"

namespace {

using namespace lrhdev;

constexpr uint32_t kSamples = 1025;  // 0..1024 over each segment

// Gauss-Jordan with full pivoting, the operation order of Matrix4x4::Inverse
// (src/luxrays/core/geometry/matrix4x4.cpp:118-175); false if singular
__device__ bool invert4d(const float in[16], float out[16]) {
	int indxc[4], indxr[4], ipiv[4] = {0, 0, 0, 0};
	float a[4][4];
	for (int i = 0; i < 4; ++i)
		for (int j = 0; j < 4; ++j) a[i][j] = in[4 * i + j];
	for (int i = 0; i < 4; ++i) {
		int irow = -1, icol = -1;
		float big = 0.f;
		for (int j = 0; j < 4; ++j) {
			if (ipiv[j] == 1) continue;
			for (int k = 0; k < 4; ++k) {
				if (ipiv[k] == 0) {
					if (fabsf(a[j][k]) >= big) { big = fabsf(a[j][k]); irow = j; icol = k; }
				} else if (ipiv[k] > 1)
					return false;
			}
		}
		if (icol < 0) return false;
		++ipiv[icol];
		if (irow != icol)
			for (int k = 0; k < 4; ++k) { const float t = a[irow][k]; a[irow][k] = a[icol][k]; a[icol][k] = t; }
		indxr[i] = irow;
		indxc[i] = icol;
		if (a[icol][icol] == 0.f) return false;
		const float pivinv = 1.f / a[icol][icol];
		a[icol][icol] = 1.f;
		for (int j = 0; j < 4; ++j) a[icol][j] *= pivinv;
		for (int j = 0; j < 4; ++j) {
			if (j == icol) continue;
			const float save = a[j][icol];
			a[j][icol] = 0;
			for (int k = 0; k < 4; ++k) a[j][k] -= a[icol][k] * save;
		}
	}
	for (int j = 3; j >= 0; --j)
		if (indxr[j] != indxc[j])
			for (int k = 0; k < 4; ++k) {
				const float t = a[k][indxr[j]];
				a[k][indxr[j]] = a[k][indxc[j]];
				a[k][indxc[j]] = t;
			}
	for (int i = 0; i < 4; ++i)
		for (int j = 0; j < 4; ++j) out[4 * i + j] = a[i][j];
	return true;
}

// Matrix * Point with the projective divide (transform.h:131-142)
__device__ void xfPointD(const float M[16], const float p[3], float r[3]) {
	const float x = p[0], y = p[1], z = p[2];
	r[0] = M[0] * x + M[1] * y + M[2] * z + M[3];
	r[1] = M[4] * x + M[5] * y + M[6] * z + M[7];
	r[2] = M[8] * x + M[9] * y + M[10] * z + M[11];
	const float w = M[12] * x + M[13] * y + M[14] * z + M[15];
	if (w != 1.f) {
		const float inv = 1.f / w;
		r[0] = inv * r[0]; r[1] = inv * r[1]; r[2] = inv * r[2];
	}
}

// Transform * BBox over the 8 corners in the order of transform.h:271-280
__device__ GBox transformBoxD(const float M[16], const GBox &b) {
	const float c[8][3] = {
		{b.lo[0], b.lo[1], b.lo[2]}, {b.hi[0], b.hi[1], b.hi[2]},
		{b.hi[0], b.lo[1], b.lo[2]}, {b.lo[0], b.hi[1], b.lo[2]}, {b.lo[0], b.lo[1], b.hi[2]},
		{b.hi[0], b.hi[1], b.lo[2]}, {b.hi[0], b.lo[1], b.hi[2]}, {b.lo[0], b.hi[1], b.hi[2]}};
	float p0[3], p1[3];
	xfPointD(M, c[0], p0);
	xfPointD(M, c[1], p1);
	GBox r;
	for (int k = 0; k < 3; ++k) { r.lo[k] = minT(p0[k], p1[k]); r.hi[k] = maxT(p0[k], p1[k]); }
	for (int i = 2; i < 8; ++i) {
		float p[3];
		xfPointD(M, c[i], p);
		for (int k = 0; k < 3; ++k) { r.lo[k] = minT(r.lo[k], p[k]); r.hi[k] = maxT(r.hi[k], p[k]); }
	}
	return r;
}

__global__ void k_fill(uint32_t *p, uint32_t n, uint32_t lowValue, uint32_t highValue) {
	// keys: 6 per box, three minima then three maxima
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) p[i] = (i % 6u) < 3u ? lowValue : highValue;
}

__global__ void k_clear(uint32_t *p) {
	if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0u;
}

// one thread per (motion object, segment, sample)
__global__ void k_motionBounds(const LrhRefitMotion *mo, uint32_t nMo, const uint32_t *sampleBase,
		const lrh_motion *motions, const lrh_motion_seg *segs, const float *meshBoxes, uint32_t *keys,
		uint32_t *error) {
	const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
	if (tid >= sampleBase[nMo]) return;
	uint32_t lo = 0, hi = nMo;  // motion object m with sampleBase[m] <= tid < sampleBase[m + 1]
	while (hi - lo > 1) {
		const uint32_t mid = (lo + hi) >> 1;
		if (sampleBase[mid] <= tid) lo = mid; else hi = mid;
	}
	const uint32_t m = lo;
	const uint32_t local = tid - sampleBase[m];
	const uint32_t seg = local / kSamples, i = local % kSamples;
	const lrh_motion &md = motions[mo[m].motion];
	const lrh_motion_seg &g = segs[md.firstSeg + seg];
	const float t = lrhLerp((float)i / 1024.f, g.startTime, g.endTime);
	float M[16], inv[16];
	lrhSegSample(g, t, M);
	if (!invert4d(M, inv)) {
		atomicOr(error, 1u);
		return;
	}
	const float *mb = meshBoxes + 6 * (size_t)mo[m].mesh;
	GBox ib;
	for (int k = 0; k < 3; ++k) { ib.lo[k] = mb[k]; ib.hi[k] = mb[3 + k]; }
	const GBox b = transformBoxD(inv, ib);
	uint32_t *kk = keys + 6 * (size_t)m;
	for (int k = 0; k < 3; ++k) {
		atomicMin(&kk[k], fkey(b.lo[k]));
		atomicMax(&kk[3 + k], fkey(b.hi[k]));
	}
}

// world box of each motion object, expanded by MachineEpsilon::E, into its leaf slot
__global__ void k_motionFinish(const LrhRefitMotion *mo, uint32_t nMo, const uint32_t *keys, float *slotBoxes) {
	const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
	if (m >= nMo) return;
	GBox b;
	for (int k = 0; k < 3; ++k) {
		b.lo[k] = unkey(keys[6 * (size_t)m + k]);
		b.hi[k] = unkey(keys[6 * (size_t)m + 3 + k]);
	}
	const float ea = maxT(machineEps(b.lo[0]), maxT(machineEps(b.lo[1]), machineEps(b.lo[2])));
	const float ec = maxT(machineEps(b.hi[0]), maxT(machineEps(b.hi[1]), machineEps(b.hi[2])));
	const float e = maxT(ea, ec);
	float *out = slotBoxes + 6 * (size_t)mo[m].slot;
	for (int k = 0; k < 3; ++k) {
		out[k] = b.lo[k] - e;
		out[3 + k] = b.hi[k] + e;
	}
}

__device__ GBox loadBox(const float *p) {
	GBox b;
	for (int k = 0; k < 3; ++k) { b.lo[k] = p[k]; b.hi[k] = p[3 + k]; }
	return b;
}

// one root node per thread; its children are leaf slots or nodes of a deeper level
__global__ void k_refitLevel(const uint32_t *levelNodes, uint32_t count, uint32_t layout, void *nodes,
		const float *slotBoxes, float *nodeBoxes) {
	const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= count) return;
	const uint32_t idx = levelNodes[e];
	uint32_t child[8];
	for (int c = 0; c < 8; ++c) child[c] = LRH_CHILD_EMPTY;
	int n = 0, width = 4;
	if (layout == 0) {
		const lrh_node4q *q = reinterpret_cast<const lrh_node4q *>(nodes) + idx;
		n = q->nchild;
		for (int c = 0; c < 4; ++c) child[c] = q->child[c];
	} else if (layout == 2) {
		const lrh_node8q *q = reinterpret_cast<const lrh_node8q *>(nodes) + idx;
		n = q->nchild;
		width = 8;
		for (int c = 0; c < 8; ++c) child[c] = q->child[c];
	} else {
		const lrh_node4f *f = reinterpret_cast<const lrh_node4f *>(nodes) + idx;
		for (int c = 0; c < 4; ++c) {
			child[c] = f->child[c];
			if (child[c] != LRH_CHILD_EMPTY) n = c + 1;
		}
	}
	if (n > width) n = width;
	GBox cb[8];
	GBox u;
	for (int k = 0; k < 3; ++k) { u.lo[k] = INFINITY; u.hi[k] = -INFINITY; }
	for (int c = 0; c < n; ++c) {
		const uint32_t r = child[c];
		if (r & LRH_CHILD_LEAF) {
			const uint32_t start = r & LRH_LEAF_START_MASK;
			const uint32_t cnt = ((r >> LRH_LEAF_START_BITS) & 7u) + 1u;
			GBox b = loadBox(slotBoxes + 6 * (size_t)start);
			for (uint32_t j = 1; j < cnt; ++j) {
				const GBox o = loadBox(slotBoxes + 6 * (size_t)(start + j));
				for (int k = 0; k < 3; ++k) { b.lo[k] = minT(b.lo[k], o.lo[k]); b.hi[k] = maxT(b.hi[k], o.hi[k]); }
			}
			cb[c] = b;
		} else {
			cb[c] = loadBox(nodeBoxes + 6 * (size_t)r);
		}
		for (int k = 0; k < 3; ++k) { u.lo[k] = minT(u.lo[k], cb[c].lo[k]); u.hi[k] = maxT(u.hi[k], cb[c].hi[k]); }
	}
	for (int k = 0; k < 3; ++k) {
		nodeBoxes[6 * (size_t)idx + k] = u.lo[k];
		nodeBoxes[6 * (size_t)idx + 3 + k] = u.hi[k];
	}
	if (layout == 0)
		quantizeNode(cb, n, child, reinterpret_cast<lrh_node4q *>(nodes) + idx);
	else if (layout == 2)
		quantizeNode8(cb, n, child, reinterpret_cast<lrh_node8q *>(nodes) + idx);
	else
		floatNode(cb, n, child, reinterpret_cast<lrh_node4f *>(nodes) + idx);
}

inline unsigned grid(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

hipError_t lrhLaunchRefitMotion(const LrhRefitArgs &a, hipStream_t st) {
	if (!a.nMotion) return hipSuccess;
	hipLaunchKernelGGL(k_fill, dim3(grid(6 * (size_t)a.nMotion, 256)), dim3(256), 0, st, a.keys, 6u * a.nMotion,
			0xffffffffu, 0u);
	hipLaunchKernelGGL(k_clear, dim3(1), dim3(64), 0, st, a.error);
	hipLaunchKernelGGL(k_motionBounds, dim3(grid(a.nSamples, 256)), dim3(256), 0, st, a.motionObjs, a.nMotion,
			a.sampleBase, a.motions, a.segs, a.meshBoxes, a.keys, a.error);
	hipLaunchKernelGGL(k_motionFinish, dim3(grid(a.nMotion, 256)), dim3(256), 0, st, a.motionObjs, a.nMotion, a.keys,
			a.slotBoxes);
	return hipGetLastError();
}

hipError_t lrhLaunchRefitLevels(const LrhRefitArgs &a, hipStream_t st) {
	for (uint32_t l = 0; l < a.nLevels; ++l) {
		const uint32_t b = a.levelStart[l], c = a.levelStart[l + 1] - b;
		if (c)
			hipLaunchKernelGGL(k_refitLevel, dim3(grid(c, 256)), dim3(256), 0, st, a.levelNodes + b, c, a.layout,
					a.nodes, a.slotBoxes, a.nodeBoxes);
	}
	return hipGetLastError();
}
