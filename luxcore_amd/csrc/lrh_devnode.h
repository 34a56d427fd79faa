/*
 * Device helpers shared by the device-side tree code (the linear BVH build
 * and the two-level root refit): boxes, MachineEpsilon, ordered float keys
 * for atomic min/max, and the 4-wide node writers with the host builder's
 * outward-rounded quantization (lrh_build.cpp quantize / toFloatNode).
 */
#ifndef LRH_DEVNODE_H
#define LRH_DEVNODE_H

#include <stdint.h>

#include <cmath>
#include <cstring>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define LRH_DN __host__ __device__
#else
#define LRH_DN
#endif

#include "lrh_layout.h"

namespace lrhdev {

// float <-> bits on either side (the host build is the CPU emulation of the
// device tree code)
LRH_DN inline uint32_t bitsOf(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
	return __float_as_uint(f);
#else
	uint32_t u;
	std::memcpy(&u, &f, 4);
	return u;
#endif
}
LRH_DN inline float fromBits(uint32_t u) {
#if defined(__HIP_DEVICE_COMPILE__)
	return __uint_as_float(u);
#else
	float f;
	std::memcpy(&f, &u, 4);
	return f;
#endif
}

struct GBox {
	float lo[3], hi[3];
};

LRH_DN inline float minT(float a, float b) { return a < b ? a : b; }  // utils.h Min
LRH_DN inline float maxT(float a, float b) { return a > b ? a : b; }  // utils.h Max

// MachineEpsilon::E (include/luxrays/core/epsilon.h:50-94)
LRH_DN inline float machineEps(float v) {
	const float n = fromBits(bitsOf(v) + 0x80u);
	const float e = fabsf(n - v);
	return e > 1e-5f ? (e < 1e-1f ? e : 1e-1f) : 1e-5f;
}

// ordered keys so unsigned min/max order floats
LRH_DN inline uint32_t fkey(float f) {
	const uint32_t u = bitsOf(f);
	return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
LRH_DN inline float unkey(uint32_t k) {
	return fromBits((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

LRH_DN inline float expScale(uint32_t e) { return fromBits(e << 23); }
LRH_DN inline float decodeQ(float org, uint32_t q, float s) { return org + (float)q * s; }

// outward-rounded 8-bit child bounds, the rule of the host builder's quantize
// W-wide quantized node (lrh_node4q, lrh_node8q): the host builder's
// outward rounding
template <int W, class Node>
LRH_DN inline void quantizeNodeW(const GBox *cb, int n, const uint32_t *child, Node *out) {
	Node q;
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
		for (int c = 0; c < W; ++c) {
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
	for (int c = 0; c < W; ++c) q.child[c] = child[c];
	for (auto &x : q.pad) x = 0;
	*out = q;
}

LRH_DN inline void quantizeNode(const GBox *cb, int n, const uint32_t child[4], lrh_node4q *out) {
	quantizeNodeW<4>(cb, n, child, out);
}

LRH_DN inline void quantizeNode8(const GBox *cb, int n, const uint32_t child[8], lrh_node8q *out) {
	quantizeNodeW<8>(cb, n, child, out);
}

LRH_DN inline void floatNode(const GBox *cb, int n, const uint32_t child[4], lrh_node4f *out) {
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
