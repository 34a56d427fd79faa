/*
 * InterpolatedTransform::Sample on the device
 * (src/luxrays/core/geometry/motionsystem.cpp:89-159), rounded as written:
 * the files that include it build with -ffp-contract=off and correctly
 * rounded division; slerp's two per-sample sines are the C library's sinf
 * restated (lrh_sinf.h), and its time-independent constants (sign,
 * phi = acosf(cos_phi), sinf(phi)) come with the segment from the host, which
 * calls the C library as the reference does, so host and device produce the
 * reference's bits.
 */
#ifndef LRH_MOTION_DEV_H
#define LRH_MOTION_DEV_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lrh_layout.h"
#include "lrh_sinf.h"

__device__ __forceinline__ float lrhLerp(float t, float a, float b) { return a + t * (b - a); }

// The first LRH_SEG_HOT_BYTES of a segment in registers: the seven loads
// are issued together and waited for once (read field by field behind the
// branches, they would wait one after another).
struct LrhSegHot {
	float startTime, endTime;
	uint32_t flags, slerpArc;
	float qsign, phi, sinPhi, pad0;
	float startQ[4], endQ[4];
	float sS[3], eS[3], sT[3], eT[3];
};
static_assert(sizeof(LrhSegHot) == LRH_SEG_HOT_BYTES, "hot part of lrh_motion_seg");

__device__ __forceinline__ LrhSegHot lrhLoadSegHot(const lrh_motion_seg &g) {
	const uint4 *p = reinterpret_cast<const uint4 *>(&g);
	uint32_t w[28];
#pragma unroll
	for (int k = 0; k < 7; ++k) {
		const uint4 q = p[k];
		w[4 * k] = q.x;
		w[4 * k + 1] = q.y;
		w[4 * k + 2] = q.z;
		w[4 * k + 3] = q.w;
	}
	// an empty asm that takes the first 32 bytes (times, flags, slerp
	// constants): those loads stay here, ahead of the branches below
#pragma unroll
	for (int k = 0; k < 8; ++k) asm volatile("" : "+v"(w[k]));
	LrhSegHot h;
	__builtin_memcpy(&h, w, sizeof(h));
	return h;
}

// one segment at `time` -> row-major M (global->local for scene objects)
__device__ __forceinline__ void lrhSegSample(const lrh_motion_seg &seg, float time, float M[16]) {
	const LrhSegHot g = lrhLoadSegHot(seg);
	const uint32_t f = g.flags;
	if (!(f & LRH_MF_ACTIVE) || time <= g.startTime) {
		for (int i = 0; i < 16; ++i) M[i] = seg.start[i];
		return;
	}
	if (time >= g.endTime) {
		for (int i = 0; i < 16; ++i) M[i] = seg.end[i];
		return;
	}
	const float w = g.endTime - g.startTime;
	const float d = time - g.startTime;
	const float le = d / w;
	if ((f & LRH_MF_TRANS) && !(f & (LRH_MF_SCALE | LRH_MF_ROT))) {
		for (int i = 0; i < 16; ++i) M[i] = seg.start[i];
		if (f & LRH_MF_TX) M[3] = lrhLerp(le, g.sT[0], g.eT[0]);
		if (f & LRH_MF_TY) M[7] = lrhLerp(le, g.sT[1], g.eT[1]);
		if (f & LRH_MF_TZ) M[11] = lrhLerp(le, g.sT[2], g.eT[2]);
		return;
	}
	if (f & LRH_MF_ROT) {
		// Slerp (src/luxrays/core/geometry/quaternion.cpp:143-164) + ToMatrix (:115-140);
		// the segment holds the shorter-arc sign, phi and sin(phi)
		float f1, f2;
		if (g.slerpArc) {
			// sinf as the reference's C library computes it (lrh_sinf.h): the
			// same float for every argument (the slerp's lie in [0, pi/2])
			f1 = lrhSinf((1.f - le) * g.phi) / g.sinPhi;
			f2 = lrhSinf(le * g.phi) / g.sinPhi;
		} else {
			f1 = 1.f - le;
			f2 = le;
		}
		const float gg = g.qsign * f2;
		float q[4];
		for (int i = 0; i < 4; ++i) q[i] = g.startQ[i] * f1 + g.endQ[i] * gg;
		const float x = q[1], y = q[2], z = q[3], qw = q[0];
		const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, xz = x * z, yz = y * z;
		const float xw = x * qw, yw = y * qw, zw = z * qw;
		M[0] = 1.f - 2.f * (yy + zz);
		M[4] = 2.f * (xy - zw);
		M[8] = 2.f * (xz + yw);
		M[1] = 2.f * (xy + zw);
		M[5] = 1.f - 2.f * (xx + zz);
		M[9] = 2.f * (yz - xw);
		M[2] = 2.f * (xz - yw);
		M[6] = 2.f * (yz + xw);
		M[10] = 1.f - 2.f * (xx + yy);
		M[3] = M[7] = M[11] = 0.f;
		M[12] = M[13] = M[14] = 0.f;
		M[15] = 1.f;
	} else {
		for (int i = 0; i < 16; ++i) M[i] = seg.startR[i];
	}
	float S[3];
	for (int k = 0; k < 3; ++k) S[k] = (f & LRH_MF_SCALE) ? lrhLerp(le, g.sS[k], g.eS[k]) : g.sS[k];
	for (int j = 0; j < 3; ++j) {
		M[0 + j] = S[0] * M[0 + j];
		M[4 + j] = S[1] * M[4 + j];
		M[8 + j] = S[2] * M[8 + j];
	}
	M[3] = (f & LRH_MF_TX) ? lrhLerp(le, g.sT[0], g.eT[0]) : g.sT[0];
	M[7] = (f & LRH_MF_TY) ? lrhLerp(le, g.sT[1], g.eT[1]) : g.sT[1];
	M[11] = (f & LRH_MF_TZ) ? lrhLerp(le, g.sT[2], g.eT[2]) : g.sT[2];
}


#endif
