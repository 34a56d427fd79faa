/*
 * Host-side scene preparation (see lrh_build.h).
 *
 * Compiled with -ffp-contract=off and no fast-math: the epsilon expansion,
 * the instance boxes and the motion-system bounds must round exactly as the
 * reference's CPU build does, because they decide which objects the device
 * may cull (src/luxrays/accelerators/mbvhaccel.cpp:187-202).
 */
#include "lrh_build.h"
#include "lrh_math.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
#include <thread>

namespace lrh {

// ---------------------------------------------------------------------------
// boxes and epsilon
// ---------------------------------------------------------------------------
static inline float fmin_t(float a, float b) { return a < b ? a : b; }  // utils.h Min
static inline float fmax_t(float a, float b) { return a > b ? a : b; }  // utils.h Max

Box Box::Empty() {
	Box b;
	for (int k = 0; k < 3; ++k) { b.lo[k] = INFINITY; b.hi[k] = -INFINITY; }
	return b;
}
void Box::grow(const Box &b) {
	for (int k = 0; k < 3; ++k) { lo[k] = fmin_t(lo[k], b.lo[k]); hi[k] = fmax_t(hi[k], b.hi[k]); }
}
void Box::growPoint(const float p[3]) {
	for (int k = 0; k < 3; ++k) { lo[k] = fmin_t(lo[k], p[k]); hi[k] = fmax_t(hi[k], p[k]); }
}
float Box::area() const {
	const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
	if (!(dx >= 0.f)) return 0.f;
	return 2.f * (dx * dy + dy * dz + dz * dx);
}

float machineEpsilon(float v) {
	uint32_t i;
	std::memcpy(&i, &v, 4);
	i += 0x80u;  // DEFAULT_EPSILON_DISTANCE_FROM_VALUE, epsilon_types.cl:29
	float n;
	std::memcpy(&n, &i, 4);
	const float e = fabsf(n - v);
	return e > 1e-5f ? (e < 1e-1f ? e : 1e-1f) : 1e-5f;  // Clamp(e, min, max)
}

float machineEpsilon(const Box &b) {
	const float a = fmax_t(machineEpsilon(b.lo[0]), fmax_t(machineEpsilon(b.lo[1]), machineEpsilon(b.lo[2])));
	const float c = fmax_t(machineEpsilon(b.hi[0]), fmax_t(machineEpsilon(b.hi[1]), machineEpsilon(b.hi[2])));
	return fmax_t(a, c);
}

Box triangleBox(const float *p0, const float *p1, const float *p2) {
	Box b;
	for (int k = 0; k < 3; ++k) {
		b.lo[k] = fmin_t(p0[k], p1[k]);
		b.hi[k] = fmax_t(p0[k], p1[k]);
	}
	for (int k = 0; k < 3; ++k) {
		b.lo[k] = fmin_t(b.lo[k], p2[k]);
		b.hi[k] = fmax_t(b.hi[k], p2[k]);
	}
	const float e = machineEpsilon(b);
	for (int k = 0; k < 3; ++k) { b.lo[k] -= e; b.hi[k] += e; }
	return b;
}

// ---------------------------------------------------------------------------
// matrices
// ---------------------------------------------------------------------------
M4 identity() {
	M4 r;
	for (int i = 0; i < 4; ++i)
		for (int j = 0; j < 4; ++j) r.m[i][j] = (i == j) ? 1.f : 0.f;
	return r;
}

bool invert(const M4 &in, M4 *out) { return invert4(in.m, out->m); }

static void xfPoint(const M4 &M, const float p[3], float r[3]) { xfPoint4(M.m, p, r); }

Box transformBox(const M4 &m, const Box &b) {
	// corner order of transform.h:271-280: (pMin,pMax) first, then six more
	const float c[8][3] = {
		{b.lo[0], b.lo[1], b.lo[2]}, {b.hi[0], b.hi[1], b.hi[2]},
		{b.hi[0], b.lo[1], b.lo[2]}, {b.lo[0], b.hi[1], b.lo[2]}, {b.lo[0], b.lo[1], b.hi[2]},
		{b.hi[0], b.hi[1], b.lo[2]}, {b.hi[0], b.lo[1], b.hi[2]}, {b.lo[0], b.hi[1], b.hi[2]}};
	float p0[3], p1[3];
	xfPoint(m, c[0], p0);
	xfPoint(m, c[1], p1);
	Box r;
	for (int k = 0; k < 3; ++k) { r.lo[k] = fmin_t(p0[k], p1[k]); r.hi[k] = fmax_t(p0[k], p1[k]); }
	for (int i = 2; i < 8; ++i) {
		float p[3];
		xfPoint(m, c[i], p);
		r.growPoint(p);
	}
	return r;
}

// ---------------------------------------------------------------------------
// motion systems
// ---------------------------------------------------------------------------
namespace {

struct V3 { float x, y, z; };
inline float dot3(const V3 &a, const V3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross3(const V3 &a, const V3 &b) {
	return V3{(a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x)};
}
inline V3 mul3(const V3 &a, float f) { return V3{a.x * f, a.y * f, a.z * f}; }
inline V3 sub3(const V3 &a, const V3 &b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline float len3(const V3 &v) { return sqrtf(v.x * v.x + v.y * v.y + v.z * v.z); }
inline float lerpf(float t, float a, float b) { return a + t * (b - a); }

float det3(float A[3][3]) {
	auto d2 = [](float a00, float a01, float a10, float a11) { return a00 * a11 - a01 * a10; };
	return A[0][0] * d2(A[1][1], A[1][2], A[2][1], A[2][2]) - A[0][1] * d2(A[1][0], A[1][2], A[2][0], A[2][2]) +
			A[0][2] * d2(A[1][0], A[1][1], A[2][0], A[2][1]);
}

float det4(const M4 &m) {  // expansion along the last row (matrix4x4.cpp Determinant)
	float result = 0, s = -1, A[3][3];
	for (int i = 0; i < 3; i++)
		for (int j = 0; j < 3; j++) A[i][j] = m.m[i][j + 1];
	for (int k = 0;; ++k) {
		if (m.m[3][k] != 0.f) result += s * m.m[3][k] * det3(A);
		if (k >= 3) break;
		s *= -1;
		for (int i = 0; i < 3; i++) A[i][k] = m.m[i][k];
	}
	return result;
}

struct Decomp {
	float S[3] = {0, 0, 0}, T[3] = {0, 0, 0};
	M4 R;
	bool valid = false;
};

// unmatrix-style decomposition (motionsystem.cpp:168-276); perspective
// terms are removed from R as there, shear is dropped by Sample.
Decomp decompose(const M4 &m) {
	Decomp d;
	d.R = m;
	M4 &R = d.R;
	if (R.m[3][3] == 0) return d;
	const float w = R.m[3][3];
	for (int i = 0; i < 4; ++i)
		for (int j = 0; j < 4; ++j) R.m[i][j] /= w;
	M4 pmat = R;
	for (int i = 0; i < 3; ++i) pmat.m[i][3] = 0.f;
	pmat.m[3][3] = 1.f;
	if (det4(pmat) == 0.f) return d;
	if (R.m[3][0] != 0.f || R.m[3][1] != 0.f || R.m[3][2] != 0.f) {
		R.m[3][0] = R.m[3][1] = R.m[3][2] = 0.f;
		R.m[3][3] = 1.f;
	}
	d.T[0] = R.m[0][3]; d.T[1] = R.m[1][3]; d.T[2] = R.m[2][3];
	for (int i = 0; i < 3; ++i) R.m[i][3] = 0.f;
	V3 row[3];
	for (int i = 0; i < 3; ++i) row[i] = V3{R.m[i][0], R.m[i][1], R.m[i][2]};
	float Sx = len3(row[0]);
	row[0] = mul3(row[0], 1.f / Sx);
	float Sxy = dot3(row[0], row[1]);
	row[1] = sub3(row[1], mul3(row[0], Sxy));
	float Sy = len3(row[1]);
	row[1] = mul3(row[1], 1.f / Sy);
	float Sxz = dot3(row[0], row[2]);
	row[2] = sub3(row[2], mul3(row[0], Sxz));
	float Syz = dot3(row[1], row[2]);
	row[2] = sub3(row[2], mul3(row[1], Syz));
	float Sz = len3(row[2]);
	row[2] = mul3(row[2], 1.f / Sz);
	if (dot3(row[0], cross3(row[1], row[2])) < 0.f) {
		Sx *= -1.f; Sy *= -1.f; Sz *= -1.f;
		for (int i = 0; i < 3; ++i) row[i] = mul3(row[i], -1.f);
	}
	for (int i = 0; i < 3; ++i) { R.m[i][0] = row[i].x; R.m[i][1] = row[i].y; R.m[i][2] = row[i].z; }
	d.S[0] = Sx; d.S[1] = Sy; d.S[2] = Sz;
	d.valid = true;
	return d;
}

void orthoNormalize(float m[4][4]) {  // quaternion.cpp:29-66
	float len, t[3][3];
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j) t[i][j] = m[i][j];
	len = sqrtf(t[0][0] * t[0][0] + t[0][1] * t[0][1] + t[0][2] * t[0][2]);
	len = (len == 0.f) ? 1.f : 1.f / len;
	t[0][0] *= len; t[0][1] *= len; t[0][2] *= len;
	t[2][0] = (t[0][1] * t[1][2] - t[0][2] * t[1][1]);
	t[2][1] = (t[0][2] * t[1][0] - t[0][0] * t[1][2]);
	t[2][2] = (t[0][0] * t[1][1] - t[0][1] * t[1][0]);
	len = sqrtf(t[2][0] * t[2][0] + t[2][1] * t[2][1] + t[2][2] * t[2][2]);
	len = (len == 0.f) ? 1.f : 1.f / len;
	t[2][0] *= len; t[2][1] *= len; t[2][2] *= len;
	t[1][0] = (t[2][1] * t[0][2] - t[2][2] * t[0][1]);
	t[1][1] = (t[2][2] * t[0][0] - t[2][0] * t[0][2]);
	t[1][2] = (t[2][0] * t[0][1] - t[2][1] * t[0][0]);
	len = sqrtf(t[1][0] * t[1][0] + t[1][1] * t[1][1] + t[1][2] * t[1][2]);
	len = (len == 0.f) ? 1.f : 1.f / len;
	t[1][0] *= len; t[1][1] *= len; t[1][2] *= len;
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j) m[i][j] = t[i][j];
}

// Quaternion(Matrix4x4) (quaternion.cpp:69-106); q = {w, x, y, z}
void quatFromMatrix(const M4 &mat, float q[4]) {
	float o[4][4];
	std::memcpy(o, mat.m, sizeof(o));
	orthoNormalize(o);
	const float trace = o[0][0] + o[1][1] + o[2][2] + 1.f;
	if (trace > 1e-6f) {
		const float s = sqrtf(trace) * 2.f;
		q[1] = (o[1][2] - o[2][1]) / s;
		q[2] = (o[2][0] - o[0][2]) / s;
		q[3] = (o[0][1] - o[1][0]) / s;
		q[0] = 0.25f * s;
	} else if (o[0][0] > o[1][1] && o[0][0] > o[2][2]) {
		const float s = sqrtf(1.f + o[0][0] - o[1][1] - o[2][2]) * 2.f;
		q[1] = 0.25f * s;
		q[2] = (o[0][1] + o[1][0]) / s;
		q[3] = (o[2][0] + o[0][2]) / s;
		q[0] = (o[1][2] - o[2][1]) / s;
	} else if (o[1][1] > o[2][2]) {
		const float s = sqrtf(1.f + o[1][1] - o[0][0] - o[2][2]) * 2.f;
		q[1] = (o[0][1] + o[1][0]) / s;
		q[2] = 0.25f * s;
		q[3] = (o[1][2] + o[2][1]) / s;
		q[0] = (o[2][0] - o[0][2]) / s;
	} else {
		const float s = sqrtf(1.f + o[2][2] - o[0][0] - o[1][1]) * 2.f;
		q[1] = (o[2][0] + o[0][2]) / s;
		q[2] = (o[1][2] + o[2][1]) / s;
		q[3] = 0.25f * s;
		q[0] = (o[0][1] - o[1][0]) / s;
	}
}

inline float qdot(const float a[4], const float b[4]) {  // w*w + Dot(v, v)
	return a[0] * b[0] + (a[1] * b[1] + a[2] * b[2] + a[3] * b[3]);
}

void qnormalize(float q[4]) {
	const float f = 1.f / sqrtf(qdot(q, q));
	for (int i = 0; i < 4; ++i) q[i] = q[i] * f;
}

void quatToMatrix(const float q[4], float m[4][4]) {  // quaternion.cpp:115-140
	const float x = q[1], y = q[2], z = q[3], w = q[0];
	const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, xz = x * z, yz = y * z;
	const float xw = x * w, yw = y * w, zw = z * w;
	m[0][0] = 1.f - 2.f * (yy + zz);
	m[1][0] = 2.f * (xy - zw);
	m[2][0] = 2.f * (xz + yw);
	m[0][1] = 2.f * (xy + zw);
	m[1][1] = 1.f - 2.f * (xx + zz);
	m[2][1] = 2.f * (yz - xw);
	m[0][2] = 2.f * (xz - yw);
	m[1][2] = 2.f * (yz + xw);
	m[2][2] = 1.f - 2.f * (xx + yy);
	m[0][3] = m[1][3] = m[2][3] = 0.f;
	m[3][0] = m[3][1] = m[3][2] = 0.f;
	m[3][3] = 1.f;
}

void slerpSeg(float t, const lrh_motion_seg &g, float out[4]) {  // quaternion.cpp:143-164
	float f1, f2;
	if (g.slerpArc) {
		// the C library's sinf, as the reference calls it; the device's
		// restatement (lrh_sinf.h) gives the same floats; phi and sin(phi)
		// come with the segment
		f1 = sinf((1.f - t) * g.phi) / g.sinPhi;
		f2 = sinf(t * g.phi) / g.sinPhi;
	} else {
		f1 = 1.f - t;
		f2 = t;
	}
	const float gg = g.qsign * f2;
	for (int i = 0; i < 4; ++i) out[i] = g.startQ[i] * f1 + g.endQ[i] * gg;
}

// the time-independent part of Slerp: the shorter-arc sign, phi, sin(phi)
void slerpConstants(lrh_motion_seg &g) {
	float cos_phi = qdot(g.startQ, g.endQ);
	g.qsign = (cos_phi > 0.f) ? 1.f : -1.f;
	cos_phi *= g.qsign;
	g.slerpArc = (1.f - cos_phi > 1e-6f) ? 1u : 0u;
	if (g.slerpArc) {
		// acosf and sinf of the C library, as quaternion.cpp:151-153 calls them
		g.phi = acosf(cos_phi);
		g.sinPhi = sinf(g.phi);
	}
}

lrh_motion_seg makeSeg(float st, float et, const M4 &s, const M4 &e) {
	lrh_motion_seg g;
	std::memset(&g, 0, sizeof(g));
	g.startTime = st;
	g.endTime = et;
	std::memcpy(g.start, s.m, 64);
	std::memcpy(g.end, e.m, 64);
	g.startQ[0] = g.endQ[0] = 1.f;
	const M4 I = identity();
	std::memcpy(g.startR, I.m, 64);
	if (st == et) return g;  // InterpolatedTransform ctor: static, inactive
	const Decomp ds = decompose(s), de = decompose(e);
	if (!ds.valid || !de.valid)
		throw Error(-1, "singular matrix in a motion key");
	quatFromMatrix(ds.R, g.startQ);
	qnormalize(g.startQ);
	quatFromMatrix(de.R, g.endQ);
	qnormalize(g.endQ);
	std::memcpy(g.startR, ds.R.m, 64);
	for (int k = 0; k < 3; ++k) {
		g.sS[k] = ds.S[k]; g.eS[k] = de.S[k];
		g.sT[k] = ds.T[k]; g.eT[k] = de.T[k];
	}
	uint32_t f = 0;
	if (ds.T[0] != de.T[0]) f |= LRH_MF_TX;
	if (ds.T[1] != de.T[1]) f |= LRH_MF_TY;
	if (ds.T[2] != de.T[2]) f |= LRH_MF_TZ;
	if (f & (LRH_MF_TX | LRH_MF_TY | LRH_MF_TZ)) f |= LRH_MF_TRANS;
	if (ds.S[0] != de.S[0]) f |= LRH_MF_SX;
	if (ds.S[1] != de.S[1]) f |= LRH_MF_SY;
	if (ds.S[2] != de.S[2]) f |= LRH_MF_SZ;
	if (f & (LRH_MF_SX | LRH_MF_SY | LRH_MF_SZ)) f |= LRH_MF_SCALE;
	if (fabsf(qdot(g.startQ, g.endQ) - 1.f) >= 1e-6f) f |= LRH_MF_ROT;
	if (f & (LRH_MF_TRANS | LRH_MF_SCALE | LRH_MF_ROT)) f |= LRH_MF_ACTIVE;
	g.flags = f;
	slerpConstants(g);
	return g;
}

M4 segSample(const lrh_motion_seg &g, float time) {  // InterpolatedTransform::Sample
	M4 r;
	if (!(g.flags & LRH_MF_ACTIVE) || time <= g.startTime) { std::memcpy(r.m, g.start, 64); return r; }
	if (time >= g.endTime) { std::memcpy(r.m, g.end, 64); return r; }
	const float w = g.endTime - g.startTime;
	const float d = time - g.startTime;
	const float le = d / w;
	if ((g.flags & LRH_MF_TRANS) && !(g.flags & (LRH_MF_SCALE | LRH_MF_ROT))) {
		std::memcpy(r.m, g.start, 64);
		if (g.flags & LRH_MF_TX) r.m[0][3] = lerpf(le, g.sT[0], g.eT[0]);
		if (g.flags & LRH_MF_TY) r.m[1][3] = lerpf(le, g.sT[1], g.eT[1]);
		if (g.flags & LRH_MF_TZ) r.m[2][3] = lerpf(le, g.sT[2], g.eT[2]);
		return r;
	}
	if (g.flags & LRH_MF_ROT) {
		float q[4];
		slerpSeg(le, g, q);
		quatToMatrix(q, r.m);
	} else
		std::memcpy(r.m, g.startR, 64);
	float S[3];
	for (int k = 0; k < 3; ++k) S[k] = (g.flags & LRH_MF_SCALE) ? lerpf(le, g.sS[k], g.eS[k]) : g.sS[k];
	for (int j = 0; j < 3; ++j) {
		r.m[0][j] = S[0] * r.m[0][j];
		r.m[1][j] = S[1] * r.m[1][j];
		r.m[2][j] = S[2] * r.m[2][j];
	}
	r.m[0][3] = (g.flags & LRH_MF_TX) ? lerpf(le, g.sT[0], g.eT[0]) : g.sT[0];
	r.m[1][3] = (g.flags & LRH_MF_TY) ? lerpf(le, g.sT[1], g.eT[1]) : g.sT[1];
	r.m[2][3] = (g.flags & LRH_MF_TZ) ? lerpf(le, g.sT[2], g.eT[2]) : g.sT[2];
	return r;
}

}  // namespace

void MotionSystemHost::init(unsigned nKeys, const float *keyTimes, const float *keyMats) {
	if (nKeys == 0) throw Error(-1, "motion system without keys");
	times.assign(keyTimes, keyTimes + nKeys);
	for (unsigned k = 1; k < nKeys; ++k)
		if (!(times[k] > times[k - 1])) throw Error(-1, "motion key times must increase strictly");
	std::vector<M4> ms(nKeys);
	for (unsigned k = 0; k < nKeys; ++k) std::memcpy(ms[k].m, keyMats + 16 * (size_t)k, 64);
	segs.clear();
	unsigned prev = 0;
	for (unsigned k = 0; k < nKeys; ++k) {
		segs.push_back(makeSeg(times[prev], times[k], ms[prev], ms[k]));
		prev = k;
	}
	segs.push_back(makeSeg(times[prev], times[prev], ms[prev], ms[prev]));
}

M4 MotionSystemHost::sample(float time) const {  // MotionSystem::Sample
	size_t index = std::upper_bound(times.begin(), times.end(), time) - times.begin();
	if (index > times.size() - 1) index = times.size() - 1;
	return segSample(segs[index], time);
}

Box MotionSystemHost::bound(const Box &ibox) const {  // MotionSystem::Bound(box, true)
	Box result = Box::Empty();
	for (const lrh_motion_seg &g : segs) {
		Box tbox = Box::Empty();
		const float N = 1024.f;
		for (float i = 0; i <= N; ++i) {
			const float t = lerpf(i / N, g.startTime, g.endTime);
			M4 inv;
			if (!invert(segSample(g, t), &inv)) throw Error(-1, "singular motion sample");
			tbox.grow(transformBox(inv, ibox));
		}
		result.grow(tbox);
	}
	return result;
}

unsigned hardwareThreads() {
	const unsigned n = std::thread::hardware_concurrency();
	return n ? n : 1;
}

// ---------------------------------------------------------------------------
// early split clipping
// ---------------------------------------------------------------------------
namespace {

struct Poly {
	// a triangle clipped to an axis-aligned box has at most 9 vertices;
	// rounding may repeat some, hence the room
	float v[16][3];
	int n = 0;
};

// the part of `in` on one side of the plane x[axis] = c (keepLow: x <= c);
// false if the result would not fit (the caller then keeps the piece whole)
bool clipPoly(const Poly &in, int axis, float c, bool keepLow, Poly *out) {
	out->n = 0;
	for (int i = 0; i < in.n; ++i) {
		const float *a = in.v[i], *b = in.v[(i + 1) % in.n];
		const bool ina = keepLow ? a[axis] <= c : a[axis] >= c;
		const bool inb = keepLow ? b[axis] <= c : b[axis] >= c;
		if (ina) {
			if (out->n == 16) return false;
			for (int k = 0; k < 3; ++k) out->v[out->n][k] = a[k];
			++out->n;
		}
		if (ina != inb) {
			if (out->n == 16) return false;
			const float t = (c - a[axis]) / (b[axis] - a[axis]);
			for (int k = 0; k < 3; ++k) out->v[out->n][k] = a[k] + t * (b[k] - a[k]);
			out->v[out->n][axis] = c;
			++out->n;
		}
	}
	return true;
}

// box of the polygon, clamped to `cell`, expanded by MachineEpsilon::E
Box polyBox(const Poly &p, const Box &cell) {
	Box b = Box::Empty();
	for (int i = 0; i < p.n; ++i) b.growPoint(p.v[i]);
	for (int k = 0; k < 3; ++k) {
		b.lo[k] = fmax_t(b.lo[k], cell.lo[k]);
		b.hi[k] = fmin_t(b.hi[k], cell.hi[k]);
	}
	const float e = machineEpsilon(b);
	for (int k = 0; k < 3; ++k) {
		b.lo[k] -= e;
		b.hi[k] += e;
	}
	return b;
}

// cut one triangle's box until each piece's area is below `target` or
// `pieces` references are made; returns the references' boxes
void cutTriangle(const float p[3][3], const Box &box, float target, unsigned maxPieces, std::vector<Box> &out) {
	struct Item { Poly poly; Box cell; };
	std::vector<Item> work;
	Item first;
	first.poly.n = 3;
	for (int i = 0; i < 3; ++i)
		for (int k = 0; k < 3; ++k) first.poly.v[i][k] = p[i][k];
	first.cell = box;
	work.push_back(first);
	std::vector<Box> done;
	while (!work.empty()) {
		Item it = work.back();
		work.pop_back();
		const Box pb = polyBox(it.poly, it.cell);
		if (pb.area() <= target || done.size() + work.size() + 2 > maxPieces) {
			done.push_back(pb);
			continue;
		}
		int axis = 0;
		for (int k = 1; k < 3; ++k)
			if (pb.hi[k] - pb.lo[k] > pb.hi[axis] - pb.lo[axis]) axis = k;
		const float c = 0.5f * (pb.lo[axis] + pb.hi[axis]);
		Item lo, hi;
		if (!clipPoly(it.poly, axis, c, true, &lo.poly) || !clipPoly(it.poly, axis, c, false, &hi.poly)) {
			done.push_back(pb);
			continue;
		}
		lo.cell = pb;
		lo.cell.hi[axis] = c;
		hi.cell = pb;
		hi.cell.lo[axis] = c;
		if (lo.poly.n >= 3) work.push_back(lo);
		if (hi.poly.n >= 3) work.push_back(hi);
		if (lo.poly.n < 3 && hi.poly.n < 3) done.push_back(pb);  // degenerate: keep the piece whole
	}
	out.insert(out.end(), done.begin(), done.end());
}

}  // namespace

void splitLargeTriangles(const std::vector<Box> &boxes, const TriangleSource &tris, float budget, unsigned nThreads,
		std::vector<Box> *refBoxes, std::vector<uint32_t> *refPrim) {
	refBoxes->clear();
	refPrim->clear();
	const size_t n = boxes.size();
	if (!(budget > 0.f) || n < 2) return;
	// typical box: the median area; a box is cut when it is above `factor`
	// times that (2 to start with), the factor raised until the cut pieces
	// fit the budget
	std::vector<float> area(n);
	for (size_t i = 0; i < n; ++i) area[i] = boxes[i].area();
	std::vector<float> tmp(area);
	std::nth_element(tmp.begin(), tmp.begin() + n / 2, tmp.end());
	const float median = tmp[n / 2];
	if (!(median > 0.f)) return;
	const double allowed = (double)budget * (double)n;
	float factor = 2.f;
	// only boxes that fit their triangle poorly are worth cutting: box area
	// above 2.5x the triangle's two-sided area (a triangle lying in an axis
	// plane has exactly 2x: its box is a flat slab already, and cutting such
	// triangles was measured to cost up to 16%; triangles of random
	// orientation have 4.5x at the median)
	// threads only for large meshes: scenes may hold thousands of small ones
	const unsigned ntFit = n < 65536 ? 1u : std::max(1u, std::min<unsigned>(nThreads ? nThreads : hardwareThreads(), 64));
	std::vector<uint8_t> loose(n, 0);
	{
		auto fit = [&](unsigned t) {
				for (size_t i = t; i < n; i += ntFit) {
					if (!(area[i] > factor * median)) continue;
					float p[3][3];
					tris.get(i, p);
					double e1[3], e2[3];
					for (int k = 0; k < 3; ++k) {
						e1[k] = (double)p[1][k] - p[0][k];
						e2[k] = (double)p[2][k] - p[0][k];
					}
					const double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
							cz = e1[0] * e2[1] - e1[1] * e2[0];
					const double twoSided = std::sqrt(cx * cx + cy * cy + cz * cz);  // 2 x triangle area
					loose[i] = (double)area[i] > 2.5 * twoSided ? 1 : 0;
				}
			};
		if (ntFit == 1) {
			fit(0);
		} else {
			std::vector<std::thread> th;
			for (unsigned t = 0; t < ntFit; ++t) th.emplace_back(fit, t);
			for (auto &x : th) x.join();
		}
	}
	std::vector<uint32_t> cand;
	for (int round = 0; round < 24; ++round, factor *= 2.f) {
		// pieces needed: about area / target per candidate (halving a box
		// roughly halves its area), at most 64 per triangle
		const float target = factor * median;
		double extra = 0.0;
		cand.clear();
		for (size_t i = 0; i < n; ++i)
			if (loose[i] && area[i] > target) {
				cand.push_back((uint32_t)i);
				extra += std::min(63.0, (double)area[i] / target);
			}
		if (extra <= allowed) break;
	}
	if (cand.empty()) return;
	const float target = factor * median;
	// cut the candidates in parallel, in chunks, then place the references:
	// every triangle keeps its slot, its extra pieces follow all triangles
	const unsigned nt = cand.size() < 1024 ? 1u
			: std::max(1u, std::min<unsigned>(nThreads ? nThreads : hardwareThreads(), 64));
	std::vector<std::vector<Box>> pieces(cand.size());
	std::atomic<size_t> next{0};
	auto cut = [&]() {
		for (;;) {
			const size_t c0 = next.fetch_add(256);
			if (c0 >= cand.size()) break;
			const size_t c1 = std::min(cand.size(), c0 + 256);
			for (size_t c = c0; c < c1; ++c) {
				float p[3][3];
				tris.get(cand[c], p);
				cutTriangle(p, boxes[cand[c]], target, 64, pieces[c]);
			}
		}
	};
	if (nt == 1) {
		cut();
	} else {
		std::vector<std::thread> th;
		for (unsigned t = 0; t < nt; ++t) th.emplace_back(cut);
		for (auto &x : th) x.join();
	}
	*refBoxes = boxes;
	refPrim->resize(n);
	for (size_t i = 0; i < n; ++i) (*refPrim)[i] = (uint32_t)i;
	for (size_t c = 0; c < cand.size(); ++c) {
		const std::vector<Box> &pc = pieces[c];
		if (pc.empty()) continue;
		(*refBoxes)[cand[c]] = pc[0];
		for (size_t j = 1; j < pc.size(); ++j) {
			refBoxes->push_back(pc[j]);
			refPrim->push_back(cand[c]);
		}
	}
}

// ---------------------------------------------------------------------------
// 4-wide binned-SAH builder
// ---------------------------------------------------------------------------
namespace {

constexpr int kMaxWidth = 8;
struct WNode {
	uint32_t n;
	uint32_t child[kMaxWidth];
	Box cbox[kMaxWidth];
};

class WideBuilder {
public:
	WideBuilder(const std::vector<Box> &boxes, const BuildParams &p) : B(boxes), P(p) {
		const size_t n = B.size();
		ids.resize(n);
		cen.resize(3 * n);
		for (size_t i = 0; i < n; ++i) {
			ids[i] = (uint32_t)i;
			for (int k = 0; k < 3; ++k) cen[3 * i + k] = B[i].lo[k] + B[i].hi[k];
		}
		cap = n + 1;
		nodes.reset(new WNode[cap]);
		count = 1;
		maxDepth = 0;
		const unsigned t = p.nThreads ? p.nThreads : hardwareThreads();
		budget = (int)t - 1;
	}

	uint32_t run() {
		const size_t n = B.size();
		Box all = Box::Empty();
		for (size_t i = 0; i < n; ++i) all.grow(B[i]);
		if (n <= P.maxLeaf) {
			// a root node with one leaf child keeps the root an inner node
			WNode &r = nodes[0];
			r.n = 1;
			r.child[0] = lrh_leaf_ref(0, (uint32_t)n);
			r.cbox[0] = all;
			count = 1;
			return 0;
		}
		buildNode(0, 0, (uint32_t)n, all, 1);
		return 0;
	}

	const std::vector<Box> &B;
	const BuildParams &P;
	std::vector<uint32_t> ids;
	std::vector<float> cen;
	std::unique_ptr<WNode[]> nodes;
	size_t cap;
	std::atomic<size_t> count;
	std::atomic<uint32_t> maxDepth;
	std::atomic<int> budget;

private:
	struct Range {
		uint32_t b, e;
		Box box;
		bool final;
	};

	// best binned SAH split of [b,e); returns false if no split exists
	bool findSplit(uint32_t b, uint32_t e, const Box &box, bool median, uint32_t *mid, Box *lb, Box *rb,
			float *cost) {
		const uint32_t n = e - b;
		Box cb = Box::Empty();
		for (uint32_t i = b; i < e; ++i) cb.growPoint(&cen[3 * (size_t)ids[i]]);
		int axis = 0;
		float ext = cb.hi[0] - cb.lo[0];
		for (int k = 1; k < 3; ++k)
			if (cb.hi[k] - cb.lo[k] > ext) { ext = cb.hi[k] - cb.lo[k]; axis = k; }
		if (!(ext > 0.f) || median) {
			// object median along the widest centroid axis
			const uint32_t m = b + n / 2;
			const int ax = axis;
			std::nth_element(ids.begin() + b, ids.begin() + m, ids.begin() + e, [&](uint32_t x, uint32_t y) {
				return cen[3 * (size_t)x + ax] < cen[3 * (size_t)y + ax];
			});
			*mid = m;
		} else {
			const int NB = n > 4096 ? 32 : 16;
			float bestCost = INFINITY;
			int bestAxis = -1, bestBin = -1;
			for (int k = 0; k < 3; ++k) {
				const float ek = cb.hi[k] - cb.lo[k];
				if (!(ek > 0.f)) continue;
				const float s = (float)NB * 0.9999f / ek;
				Box bb[32];
				uint32_t cnt[32];
				for (int i = 0; i < NB; ++i) { bb[i] = Box::Empty(); cnt[i] = 0; }
				for (uint32_t i = b; i < e; ++i) {
					const uint32_t id = ids[i];
					int bin = (int)((cen[3 * (size_t)id + k] - cb.lo[k]) * s);
					bin = bin < 0 ? 0 : (bin >= NB ? NB - 1 : bin);
					cnt[bin]++;
					bb[bin].grow(B[id]);
				}
				float rA[32];
				uint32_t rN[32];
				Box acc = Box::Empty();
				uint32_t c = 0;
				for (int i = NB - 1; i > 0; --i) {
					acc.grow(bb[i]);
					c += cnt[i];
					rA[i] = acc.area();
					rN[i] = c;
				}
				acc = Box::Empty();
				c = 0;
				for (int i = 0; i < NB - 1; ++i) {
					acc.grow(bb[i]);
					c += cnt[i];
					if (c == 0 || rN[i + 1] == 0) continue;
					const float cst = acc.area() * (float)c + rA[i + 1] * (float)rN[i + 1];
					if (cst < bestCost) { bestCost = cst; bestAxis = k; bestBin = i; }
				}
			}
			if (bestAxis < 0) return findSplit(b, e, box, true, mid, lb, rb, cost);
			const int k = bestAxis;
			const float s = (float)NB * 0.9999f / (cb.hi[k] - cb.lo[k]);
			const float lo = cb.lo[k];
			auto it = std::partition(ids.begin() + b, ids.begin() + e, [&](uint32_t id) {
				int bin = (int)((cen[3 * (size_t)id + k] - lo) * s);
				bin = bin < 0 ? 0 : (bin >= NB ? NB - 1 : bin);
				return bin <= bestBin;
			});
			*mid = (uint32_t)(it - ids.begin());
			if (*mid == b || *mid == e) return findSplit(b, e, box, true, mid, lb, rb, cost);
		}
		Box l = Box::Empty(), r = Box::Empty();
		for (uint32_t i = b; i < *mid; ++i) l.grow(B[ids[i]]);
		for (uint32_t i = *mid; i < e; ++i) r.grow(B[ids[i]]);
		*lb = l;
		*rb = r;
		const float A = box.area();
		*cost = P.travCost + (A > 0.f ? (l.area() * (float)(*mid - b) + r.area() * (float)(e - *mid)) / A : (float)n);
		return true;
	}

	void buildNode(uint32_t me, uint32_t b, uint32_t e, const Box &box, uint32_t depth) {
		uint32_t md = maxDepth.load();
		while (depth > md && !maxDepth.compare_exchange_weak(md, depth)) {}
		const bool median = depth > 40;
		const int width = P.layout == 2 ? 8 : 4;
		Range ch[kMaxWidth];
		int nch = 1;
		ch[0] = Range{b, e, box, false};
		while (nch < width) {
			int pick = -1;
			float best = -1.f;
			for (int i = 0; i < nch; ++i) {
				if (ch[i].final || ch[i].e - ch[i].b < 2) continue;
				const float a = ch[i].box.area();
				if (a > best) { best = a; pick = i; }
			}
			if (pick < 0) break;
			const Range r = ch[pick];
			uint32_t mid;
			Box lb, rb;
			float cost;
			findSplit(r.b, r.e, r.box, median, &mid, &lb, &rb, &cost);
			const uint32_t n = r.e - r.b;
			if (n <= P.maxLeaf && (float)n <= cost && nch > 1) {
				ch[pick].final = true;  // SAH prefers this leaf
				continue;
			}
			for (int i = nch; i > pick + 1; --i) ch[i] = ch[i - 1];
			ch[pick] = Range{r.b, mid, lb, false};
			ch[pick + 1] = Range{mid, r.e, rb, false};
			++nch;
		}
		int nInner = 0;
		for (int i = 0; i < nch; ++i)
			if (ch[i].e - ch[i].b > P.maxLeaf) ++nInner;
		const size_t base = nInner ? count.fetch_add(nInner) : 0;
		if (base + nInner > cap) throw Error(-3, "node pool exhausted");
		WNode &node = nodes[me];
		node.n = nch;
		std::vector<std::thread> th;
		int k = 0;
		for (int i = 0; i < nch; ++i) {
			node.cbox[i] = ch[i].box;
			const uint32_t n = ch[i].e - ch[i].b;
			if (n <= P.maxLeaf) {
				node.child[i] = lrh_leaf_ref(ch[i].b, n);
				continue;
			}
			const uint32_t ci = (uint32_t)(base + k++);
			node.child[i] = ci;
			const Range r = ch[i];
			const bool big = n > 65536;
			if (big && budget.fetch_sub(1) > 0) {
				th.emplace_back([this, ci, r, depth]() {
					buildNode(ci, r.b, r.e, r.box, depth + 1);
					budget.fetch_add(1);
				});
			} else {
				if (big) budget.fetch_add(1);
				buildNode(ci, r.b, r.e, r.box, depth + 1);
			}
		}
		for (auto &t : th) t.join();
	}
};

// outward-rounded 8-bit child bounds (see lrh_node4q)
inline float decodeQ(float org, uint32_t q, float s) { return org + (float)q * s; }
inline float expScale(uint32_t e) { uint32_t b = e << 23; float f; std::memcpy(&f, &b, 4); return f; }

void quantize(const WNode &w, lrh_node4q *q) {
	std::memset(q, 0, sizeof(*q));
	Box pb = Box::Empty();
	for (uint32_t c = 0; c < w.n; ++c) pb.grow(w.cbox[c]);
	uint8_t *qlo[3] = {q->qlo_x, q->qlo_y, q->qlo_z};
	uint8_t *qhi[3] = {q->qhi_x, q->qhi_y, q->qhi_z};
	uint8_t *ex[3] = {&q->ex, &q->ey, &q->ez};
	for (int k = 0; k < 3; ++k) {
		const float o = pb.lo[k];
		q->org[k] = o;
		const float ext = pb.hi[k] - o;
		int e2 = 0;
		if (ext > 0.f) {
			frexpf(ext / 255.f, &e2);  // ext/255 = f * 2^e2 with f in [0.5, 1)
		} else
			e2 = -126;
		int e = e2 + 127;
		if (e < 1) e = 1;
		while (e < 254 && decodeQ(o, 255, expScale((uint32_t)e)) < pb.hi[k]) ++e;
		const float s = expScale((uint32_t)e);
		*ex[k] = (uint8_t)e;
		for (uint32_t c = 0; c < 4; ++c) {
			// an empty slot gets an inverted box (lo at the top, hi at the
			// bottom of the node): the slab test rejects it without a check of
			// the child reference
			if (c >= w.n) { qlo[k][c] = 255; qhi[k][c] = 0; continue; }
			const float lo = w.cbox[c].lo[k], hi = w.cbox[c].hi[k];
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
	q->nchild = (uint8_t)w.n;
	for (uint32_t c = 0; c < 4; ++c) q->child[c] = c < w.n ? w.child[c] : LRH_CHILD_EMPTY;
}

// the 8-wide node: the same outward rounding as quantize, eight slots
void quantize8(const WNode &w, lrh_node8q *q) {
	std::memset(q, 0, sizeof(*q));
	Box pb = Box::Empty();
	for (uint32_t c = 0; c < w.n; ++c) pb.grow(w.cbox[c]);
	uint8_t *qlo[3] = {q->qlo_x, q->qlo_y, q->qlo_z};
	uint8_t *qhi[3] = {q->qhi_x, q->qhi_y, q->qhi_z};
	uint8_t *ex[3] = {&q->ex, &q->ey, &q->ez};
	for (int k = 0; k < 3; ++k) {
		const float o = pb.lo[k];
		q->org[k] = o;
		const float ext = pb.hi[k] - o;
		int e2 = 0;
		if (ext > 0.f)
			frexpf(ext / 255.f, &e2);
		else
			e2 = -126;
		int e = e2 + 127;
		if (e < 1) e = 1;
		while (e < 254 && decodeQ(o, 255, expScale((uint32_t)e)) < pb.hi[k]) ++e;
		const float s = expScale((uint32_t)e);
		*ex[k] = (uint8_t)e;
		for (uint32_t c = 0; c < 8; ++c) {
			if (c >= w.n) { qlo[k][c] = 255; qhi[k][c] = 0; continue; }
			const float lo = w.cbox[c].lo[k], hi = w.cbox[c].hi[k];
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
	q->nchild = (uint8_t)w.n;
	for (uint32_t c = 0; c < 8; ++c) q->child[c] = c < w.n ? w.child[c] : LRH_CHILD_EMPTY;
}

void toFloatNode(const WNode &w, lrh_node4f *f) {
	std::memset(f, 0, sizeof(*f));
	for (uint32_t c = 0; c < 4; ++c) {
		const bool v = c < w.n;
		f->lo_x[c] = v ? w.cbox[c].lo[0] : INFINITY;
		f->lo_y[c] = v ? w.cbox[c].lo[1] : INFINITY;
		f->lo_z[c] = v ? w.cbox[c].lo[2] : INFINITY;
		f->hi_x[c] = v ? w.cbox[c].hi[0] : -INFINITY;
		f->hi_y[c] = v ? w.cbox[c].hi[1] : -INFINITY;
		f->hi_z[c] = v ? w.cbox[c].hi[2] : -INFINITY;
		f->child[c] = v ? w.child[c] : LRH_CHILD_EMPTY;
	}
}

template <class F>
void parallelFor(size_t n, unsigned nThreads, F fn) {
	if (nThreads <= 1 || n < 65536) { for (size_t i = 0; i < n; ++i) fn(i); return; }
	std::vector<std::thread> th;
	for (unsigned t = 0; t < nThreads; ++t)
		th.emplace_back([=, &fn]() {
			for (size_t i = t; i < n; i += nThreads) fn(i);
		});
	for (auto &t : th) t.join();
}

}  // namespace

void buildWideBVH(const std::vector<Box> &boxes, const BuildParams &p, uint32_t nodeBase,
		uint32_t primBase, WideBVH *out) {
	if (boxes.empty()) throw Error(-1, "empty primitive set");
	if (p.maxLeaf < 1 || p.maxLeaf > LRH_LEAF_MAX) throw Error(-1, "maxLeafSize must be 1..8");
	if ((uint64_t)primBase + boxes.size() > LRH_MAX_LEAF_START)
		throw Error(-4, "too many primitives for 28-bit leaf offsets");
	WideBuilder wb(boxes, p);
	wb.run();
	const size_t nNodes = wb.count.load();
	if ((uint64_t)nodeBase + nNodes >= 0x80000000ull) throw Error(-4, "too many nodes");
	// top of the tree first, breadth first: node i < K is at depth <= that of
	// node K, so a prefix of the array holds the most visited nodes (the
	// trace kernel keeps such a prefix in LDS); the rest keep build order
	std::vector<uint32_t> newIndex(nNodes, UINT32_MAX), oldOf;
	oldOf.reserve(nNodes);
	{
		const size_t topMax = std::min<size_t>(nNodes, LRH_TOP_NODES);
		oldOf.push_back(0);
		newIndex[0] = 0;
		for (size_t q = 0; q < oldOf.size() && oldOf.size() < topMax; ++q) {
			const WNode &w = wb.nodes[oldOf[q]];
			for (uint32_t c = 0; c < w.n && oldOf.size() < topMax; ++c) {
				const uint32_t ch = w.child[c];
				if (ch & LRH_CHILD_LEAF) continue;
				newIndex[ch] = (uint32_t)oldOf.size();
				oldOf.push_back(ch);
			}
		}
		for (size_t i = 0; i < nNodes; ++i)
			if (newIndex[i] == UINT32_MAX) {
				newIndex[i] = (uint32_t)oldOf.size();
				oldOf.push_back((uint32_t)i);
			}
	}
	// relocate child references
	auto reloc = [&](uint32_t c) -> uint32_t {
		if (c == LRH_CHILD_EMPTY) return c;
		if (c & LRH_CHILD_LEAF) {
			const uint32_t cntm1 = (c >> LRH_LEAF_START_BITS) & 7u;
			const uint32_t start = (c & LRH_LEAF_START_MASK) + primBase;
			return LRH_CHILD_LEAF | (cntm1 << LRH_LEAF_START_BITS) | start;
		}
		return newIndex[c] + nodeBase;
	};
	const unsigned nt = p.nThreads ? p.nThreads : hardwareThreads();
	for (size_t i = 0; i < nNodes; ++i)
		for (uint32_t c = 0; c < wb.nodes[i].n; ++c) wb.nodes[i].child[c] = reloc(wb.nodes[i].child[c]);
	out->qnodes.clear();
	out->fnodes.clear();
	out->q8nodes.clear();
	if (p.layout == 0) {
		out->qnodes.resize(nNodes);
		parallelFor(nNodes, nt, [&](size_t i) { quantize(wb.nodes[oldOf[i]], &out->qnodes[i]); });
	} else if (p.layout == 2) {
		out->q8nodes.resize(nNodes);
		parallelFor(nNodes, nt, [&](size_t i) { quantize8(wb.nodes[oldOf[i]], &out->q8nodes[i]); });
	} else {
		out->fnodes.resize(nNodes);
		parallelFor(nNodes, nt, [&](size_t i) { toFloatNode(wb.nodes[oldOf[i]], &out->fnodes[i]); });
	}
	out->order = std::move(wb.ids);
	out->root = nodeBase;
	out->maxDepth = wb.maxDepth.load();
}

}  // namespace lrh
