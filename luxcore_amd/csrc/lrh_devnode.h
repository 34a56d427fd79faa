/*
 * Device helpers shared by the device-side tree code (the linear BVH build
 * and the two-level root refit): boxes, MachineEpsilon, ordered float keys
 * for atomic min/max, and the 4-wide node writers with the host builder's
 * outward-rounded quantization (lrh_build.cpp quantize / toFloatNode).
 */
#ifndef LRH_DEVNODE_H
#define LRH_DEVNODE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lrh_layout.h"

namespace lrhdev {

struct GBox {
	float lo[3], hi[3];
};

__device__ __forceinline__ float minT(float a, float b) { return a < b ? a : b; }  // utils.h Min
__device__ __forceinline__ float maxT(float a, float b) { return a > b ? a : b; }  // utils.h Max

// MachineEpsilon::E (include/luxrays/core/epsilon.h:50-94)
__device__ inline float machineEps(float v) {
	const float n = __uint_as_float(__float_as_uint(v) + 0x80u);
	const float e = fabsf(n - v);
	return e > 1e-5f ? (e < 1e-1f ? e : 1e-1f) : 1e-5f;
}

// ordered keys so unsigned min/max order floats
__device__ __forceinline__ uint32_t fkey(float f) {
	const uint32_t u = __float_as_uint(f);
	return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unkey(uint32_t k) {
	return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

__device__ __forceinline__ float expScale(uint32_t e) { return __uint_as_float(e << 23); }
__device__ __forceinline__ float decodeQ(float org, uint32_t q, float s) { return org + (float)q * s; }

// outward-rounded 8-bit child bounds, the rule of the host builder's quantize
__device__ inline void quantizeNode(const GBox *cb, int n, const uint32_t child[4], lrh_node4q *out) {
	lrh_node4q q;
	GBox pb;
	for (int k = 0; k < 3; ++k) {
		pb.lo[k] = INFINITY;
		pb.hi[k] = -INFINITY;
	}
	for (int c = 0; c < n; ++c)
		for (int k = 0; k < 3; ++k) {
			pb.lo[k] = minT(pb.lo[k], cb[c].lo[k]);
			pb.hi[k] = maxT(pb.hi[k], cb[c].hi[k]);
		}
	uint8_t *qlo[3] = {q.qlo_x, q.qlo_y, q.qlo_z};
	uint8_t *qhi[3] = {q.qhi_x, q.qhi_y, q.qhi_z};
	uint8_t *ex[3] = {&q.ex, &q.ey, &q.ez};
	for (int k = 0; k < 3; ++k) {
		const float o = pb.lo[k];
		q.org[k] = o;
		const float ext = pb.hi[k] - o;
		int e2 = -126;
		if (ext > 0.f) frexpf(ext / 255.f, &e2);
		int e = e2 + 127;
		if (e < 1) e = 1;
		while (e < 254 && decodeQ(o, 255, expScale((uint32_t)e)) < pb.hi[k]) ++e;
		const float s = expScale((uint32_t)e);
		*ex[k] = (uint8_t)e;
		for (int c = 0; c < 4; ++c) {
			if (c >= n) {
				qlo[k][c] = 255;  // empty slot: inverted box
				qhi[k][c] = 0;
				continue;
			}
			const float lo = cb[c].lo[k], hi = cb[c].hi[k];
			long ql = (long)floorf((lo - o) / s);
			if (ql < 0) ql = 0;
			if (ql > 255) ql = 255;
			while (ql > 0 && decodeQ(o, (uint32_t)ql, s) > lo) --ql;
			long qh = (long)ceilf((hi - o) / s);
			if (qh < 0) qh = 0;
			if (qh > 255) qh = 255;
			while (qh < 255 && decodeQ(o, (uint32_t)qh, s) < hi) ++qh;
			qlo[k][c] = (uint8_t)ql;
			qhi[k][c] = (uint8_t)qh;
		}
	}
	q.nchild = (uint8_t)n;
	for (int c = 0; c < 4; ++c) q.child[c] = child[c];
	q.pad[0] = q.pad[1] = 0;
	*out = q;
}

__device__ inline void floatNode(const GBox *cb, int n, const uint32_t child[4], lrh_node4f *out) {
	lrh_node4f f;
	for (int c = 0; c < 4; ++c) {
		const bool v = c < n;
		f.lo_x[c] = v ? cb[c].lo[0] : INFINITY;
		f.lo_y[c] = v ? cb[c].lo[1] : INFINITY;
		f.lo_z[c] = v ? cb[c].lo[2] : INFINITY;
		f.hi_x[c] = v ? cb[c].hi[0] : -INFINITY;
		f.hi_y[c] = v ? cb[c].hi[1] : -INFINITY;
		f.hi_z[c] = v ? cb[c].hi[2] : -INFINITY;
		f.child[c] = child[c];
		f.pad[c] = 0;
	}
	*out = f;
}

}  // namespace lrhdev

#endif
