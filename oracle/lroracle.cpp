/*
 * lroracle.cpp — CPU restatement of LuxRays' BVHAccel / MBVHAccel path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): see lroracle.h for who may load it
 * and for what pins each function.  Paths cited are relative to the LuxCore
 * v2.6 source tree.
 *
 * Build flags (oracle/Makefile): -O3 -msse -msse2 -msse3 -mssse3
 * -ffp-contract=off, as cmake/PlatformSpecific.cmake:265-273 plus no FMA
 * contraction, so every float expression below rounds like the reference.
 * Expression shapes (operand order, where a reciprocal is taken, ternary
 * Min/Max) are kept identical to the cited lines on purpose: they decide the
 * last bit of t/b1/b2 and the CLASSIC tree topology.
 */
#include "lroracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <thread>
#include <vector>

namespace {

const unsigned NULL_INDEX = 0xffffffffu;  // luxrays_types.cl:21

// ---------------------------------------------------------------------------
// Scalar helpers: include/luxrays/utils/utils.h:122-161
// ---------------------------------------------------------------------------
template <class T> inline T Min(T a, T b) { return a < b ? a : b; }
template <class T> inline T Max(T a, T b) { return a > b ? a : b; }
template <class T> inline T Clamp(T v, T lo, T hi) { return v > lo ? (v < hi ? v : hi) : lo; }
inline float Lerp(float t, float v1, float v2) { return v1 + t * (v2 - v1); }

// The reference's slerp calls the platform's acosf/sinf
// (src/luxrays/core/geometry/quaternion.cpp:151-155), whose last bit is
// libm-specific (glibc's are not correctly rounded).  The oracle calls the
// host libm, as the reference does; orc_set_libm(1) switches to correctly
// rounded results (computed in double, then rounded once) instead, for
// checks that must not depend on the platform's library.
int g_crMath = 0;
inline float AcosF(float x) { return g_crMath ? (float)acos((double)x) : acosf(x); }
inline float SinF(float x) { return g_crMath ? (float)sin((double)x) : sinf(x); }

// ---------------------------------------------------------------------------
// Point / Vector: include/luxrays/core/geometry/{point,vector}.h
// ---------------------------------------------------------------------------
struct V3 {
	float x, y, z;
	float operator[](int i) const { return (&x)[i]; }
	float &operator[](int i) { return (&x)[i]; }
};
inline V3 mk(float x, float y, float z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
inline V3 sub(const V3 &a, const V3 &b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 add(const V3 &a, const V3 &b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 scale(const V3 &a, float f) { return mk(a.x * f, a.y * f, a.z * f); }
// vector.h:151-153
inline float Dot(const V3 &a, const V3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// vector.h:159-163
inline V3 Cross(const V3 &a, const V3 &b) {
	return mk((a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x));
}
inline float Length(const V3 &v) { return sqrtf(v.x * v.x + v.y * v.y + v.z * v.z); }
// point.h: operator/ multiplies by the reciprocal
inline V3 divPoint(const V3 &p, float f) { const float inv = 1.f / f; return mk(inv * p.x, inv * p.y, inv * p.z); }

// ---------------------------------------------------------------------------
// BBox: include/luxrays/core/geometry/bbox.h:42-100, src/.../bbox.cpp:29-49
// ---------------------------------------------------------------------------
struct BBox {
	V3 pMin, pMax;
	BBox() : pMin(mk(INFINITY, INFINITY, INFINITY)), pMax(mk(-INFINITY, -INFINITY, -INFINITY)) {}
	BBox(const V3 &p1, const V3 &p2) {
		pMin = mk(Min(p1.x, p2.x), Min(p1.y, p2.y), Min(p1.z, p2.z));
		pMax = mk(Max(p1.x, p2.x), Max(p1.y, p2.y), Max(p1.z, p2.z));
	}
	void Expand(float d) {
		pMin.x -= d; pMin.y -= d; pMin.z -= d;
		pMax.x += d; pMax.y += d; pMax.z += d;
	}
	float SurfaceArea() const {
		const V3 d = sub(pMax, pMin);
		return 2.f * (d.x * d.y + d.y * d.z + d.z * d.x);
	}
};
inline BBox Union(const BBox &b, const V3 &p) {
	BBox r = b;
	r.pMin.x = Min(b.pMin.x, p.x); r.pMin.y = Min(b.pMin.y, p.y); r.pMin.z = Min(b.pMin.z, p.z);
	r.pMax.x = Max(b.pMax.x, p.x); r.pMax.y = Max(b.pMax.y, p.y); r.pMax.z = Max(b.pMax.z, p.z);
	return r;
}
inline BBox Union(const BBox &b, const BBox &b2) {
	BBox r;
	r.pMin.x = Min(b.pMin.x, b2.pMin.x); r.pMin.y = Min(b.pMin.y, b2.pMin.y); r.pMin.z = Min(b.pMin.z, b2.pMin.z);
	r.pMax.x = Max(b.pMax.x, b2.pMax.x); r.pMax.y = Max(b.pMax.y, b2.pMax.y); r.pMax.z = Max(b.pMax.z, b2.pMax.z);
	return r;
}

// ---------------------------------------------------------------------------
// MachineEpsilon: include/luxrays/core/epsilon.h:50-94, epsilon_types.cl:21-30
// ---------------------------------------------------------------------------
const float EPS_MIN = 1e-5f, EPS_MAX = 1e-1f;
const unsigned EPS_ULPS = 0x80u;
inline float NextFloat(float v) {
	uint32_t i; std::memcpy(&i, &v, 4); i += EPS_ULPS; float r; std::memcpy(&r, &i, 4); return r;
}
inline float E(float v) { const float e = fabsf(NextFloat(v) - v); return Clamp(e, EPS_MIN, EPS_MAX); }
inline float E(const V3 &p) { return Max(E(p.x), Max(E(p.y), E(p.z))); }
inline float E(const BBox &b) { return Max(E(b.pMin), E(b.pMax)); }

// ---------------------------------------------------------------------------
// Ray
// ---------------------------------------------------------------------------
struct Ray {
	V3 o, d;
	float mint, maxt, time;
};
inline Ray toRay(const orc_ray &r) {
	Ray x; x.o = mk(r.ox, r.oy, r.oz); x.d = mk(r.dx, r.dy, r.dz);
	x.mint = r.mint; x.maxt = r.maxt; x.time = r.time; return x;
}

// Triangle::Intersect: include/luxrays/core/geometry/triangle.h:55-89
inline bool TriangleIntersect(const Ray &ray, const V3 &p0, const V3 &p1, const V3 &p2,
		float *t, float *b1, float *b2) {
	const V3 e1 = sub(p1, p0);
	const V3 e2 = sub(p2, p0);
	const V3 s1 = Cross(ray.d, e2);
	const float divisor = Dot(s1, e1);
	if (divisor == 0.f)
		return false;
	const float invDivisor = 1.f / divisor;
	const V3 d = sub(ray.o, p0);
	*b1 = Dot(d, s1) * invDivisor;
	if (*b1 < 0.f)
		return false;
	const V3 s2 = Cross(d, e1);
	*b2 = Dot(ray.d, s2) * invDivisor;
	if (*b2 < 0.f)
		return false;
	const float b0 = 1.f - *b1 - *b2;
	if (b0 < 0.f)
		return false;
	*t = Dot(e2, s2) * invDivisor;
	if (*t < ray.mint || *t > ray.maxt)
		return false;
	return true;
}

// BBox::IntersectP: src/luxrays/core/geometry/bbox.cpp:147-165
inline bool BBoxIntersectP(const Ray &ray, const V3 &pMin, const V3 &pMax,
		float *hitt0 = nullptr, float *hitt1 = nullptr) {
	float t0 = ray.mint, t1 = ray.maxt;
	for (int i = 0; i < 3; ++i) {
		const float invRayDir = 1.f / ray.d[i];
		float tNear = (pMin[i] - ray.o[i]) * invRayDir;
		float tFar = (pMax[i] - ray.o[i]) * invRayDir;
		if (tNear > tFar) std::swap(tNear, tFar);
		t0 = tNear > t0 ? tNear : t0;
		t1 = tFar < t1 ? tFar : t1;
		if (t0 > t1) return false;
	}
	if (hitt0) *hitt0 = t0;
	if (hitt1) *hitt1 = t1;
	return true;
}

// ---------------------------------------------------------------------------
// Matrix4x4 / Transform: matrix4x4.h, matrix4x4.cpp:118-175, matrix4x4op.h:31-104
// ---------------------------------------------------------------------------
struct M4 {
	float m[4][4];
	static M4 Identity() { M4 r; for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) r.m[i][j] = (i == j) ? 1.f : 0.f; return r; }
};
inline M4 mul(const M4 &a, const M4 &b) {  // matrix4x4.h operator*
	M4 r;
	for (int i = 0; i < 4; ++i)
		for (int j = 0; j < 4; ++j)
			r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j] + a.m[i][3] * b.m[3][j];
	return r;
}
inline M4 Transpose(const M4 &a) { M4 r; for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) r.m[i][j] = a.m[j][i]; return r; }

// matrix4x4.cpp:118-175 (full-pivot Gauss-Jordan); false on a singular matrix
bool Inverse(const M4 &in, M4 *out) {
	int indxc[4], indxr[4];
	int ipiv[4] = {0, 0, 0, 0};
	float minv[4][4];
	std::memcpy(minv, in.m, sizeof(minv));
	for (int i = 0; i < 4; ++i) {
		int irow = -1, icol = -1;
		float big = 0.f;
		for (int j = 0; j < 4; ++j) {
			if (ipiv[j] != 1) {
				for (int k = 0; k < 4; ++k) {
					if (ipiv[k] == 0) {
						if (fabsf(minv[j][k]) >= big) {
							big = fabsf(minv[j][k]);
							irow = j;
							icol = k;
						}
					} else if (ipiv[k] > 1)
						return false;
				}
			}
		}
		if (icol < 0) return false;
		++ipiv[icol];
		if (irow != icol) {
			for (int k = 0; k < 4; ++k)
				std::swap(minv[irow][k], minv[icol][k]);
		}
		indxr[i] = irow;
		indxc[i] = icol;
		if (minv[icol][icol] == 0.f)
			return false;
		const float pivinv = 1.f / minv[icol][icol];
		minv[icol][icol] = 1.f;
		for (int j = 0; j < 4; ++j)
			minv[icol][j] *= pivinv;
		for (int j = 0; j < 4; ++j) {
			if (j != icol) {
				const float save = minv[j][icol];
				minv[j][icol] = 0;
				for (int k = 0; k < 4; ++k)
					minv[j][k] -= minv[icol][k] * save;
			}
		}
	}
	for (int j = 3; j >= 0; --j) {
		if (indxr[j] != indxc[j]) {
			for (int k = 0; k < 4; ++k)
				std::swap(minv[k][indxr[j]], minv[k][indxc[j]]);
		}
	}
	std::memcpy(out->m, minv, sizeof(minv));
	return true;
}

// matrix4x4.cpp Det2x2/Det3x3/Determinant (expansion along the last row)
inline float Det2x2(float a00, float a01, float a10, float a11) { return a00 * a11 - a01 * a10; }
inline float Det3x3(float A[3][3]) {
	return A[0][0] * Det2x2(A[1][1], A[1][2], A[2][1], A[2][2]) -
			A[0][1] * Det2x2(A[1][0], A[1][2], A[2][0], A[2][2]) +
			A[0][2] * Det2x2(A[1][0], A[1][1], A[2][0], A[2][1]);
}
float Determinant(const M4 &mm) {
	float result = 0, s = -1;
	float A[3][3];
	for (int i = 0; i < 3; i++)
		for (int j = 0; j < 3; j++)
			A[i][j] = mm.m[i][j + 1];
	int k = 0;
	while (true) {
		if (mm.m[3][k] != 0.f)
			result += s * mm.m[3][k] * Det3x3(A);
		if (k >= 3) break;
		s *= -1;
		for (int i = 0; i < 3; i++)
			A[i][k] = mm.m[i][k];
		k++;
	}
	return result;
}

// matrix4x4op.h:31-41 (point with the /w branch) and :54-59 (vector)
inline V3 xfPoint(const M4 &M, const V3 &p) {
	const float x = p.x, y = p.y, z = p.z;
	const V3 pr = mk(M.m[0][0] * x + M.m[0][1] * y + M.m[0][2] * z + M.m[0][3],
			M.m[1][0] * x + M.m[1][1] * y + M.m[1][2] * z + M.m[1][3],
			M.m[2][0] * x + M.m[2][1] * y + M.m[2][2] * z + M.m[2][3]);
	const float w = M.m[3][0] * x + M.m[3][1] * y + M.m[3][2] * z + M.m[3][3];
	if (w != 1.f)
		return divPoint(pr, w);
	return pr;
}
inline V3 xfVector(const M4 &M, const V3 &v) {
	const float x = v.x, y = v.y, z = v.z;
	return mk(M.m[0][0] * x + M.m[0][1] * y + M.m[0][2] * z,
			M.m[1][0] * x + M.m[1][1] * y + M.m[1][2] * z,
			M.m[2][0] * x + M.m[2][1] * y + M.m[2][2] * z);
}
// matrix4x4op.h:94-103 / transform.h:271-280 (8 corners, this union order)
BBox xfBBox(const M4 &M, const BBox &b) {
	BBox r(xfPoint(M, b.pMin), xfPoint(M, b.pMax));
	r = Union(r, xfPoint(M, mk(b.pMax.x, b.pMin.y, b.pMin.z)));
	r = Union(r, xfPoint(M, mk(b.pMin.x, b.pMax.y, b.pMin.z)));
	r = Union(r, xfPoint(M, mk(b.pMin.x, b.pMin.y, b.pMax.z)));
	r = Union(r, xfPoint(M, mk(b.pMax.x, b.pMax.y, b.pMin.z)));
	r = Union(r, xfPoint(M, mk(b.pMax.x, b.pMin.y, b.pMax.z)));
	r = Union(r, xfPoint(M, mk(b.pMin.x, b.pMax.y, b.pMax.z)));
	return r;
}
inline Ray xfRay(const M4 &M, const Ray &r) {
	Ray o; o.o = xfPoint(M, r.o); o.d = xfVector(M, r.d);
	o.mint = r.mint; o.maxt = r.maxt; o.time = r.time; return o;
}

// ---------------------------------------------------------------------------
// Quaternion: include/luxrays/core/geometry/quaternion.h, quaternion.cpp:29-164
// ---------------------------------------------------------------------------
struct Quat { float w; V3 v; };
inline float QDot(const Quat &a, const Quat &b) { return a.w * b.w + Dot(a.v, b.v); }
inline Quat QScale(float f, const Quat &q) { Quat r; r.w = q.w * f; r.v = scale(q.v, f); return r; }
inline Quat QAdd(const Quat &a, const Quat &b) { Quat r; r.w = a.w + b.w; r.v = add(a.v, b.v); return r; }
inline Quat QNormalize(const Quat &q) { return QScale(1.f / sqrtf(QDot(q, q)), q); }

void orthoNormalize(float m[4][4]) {  // quaternion.cpp:29-66
	float len, temp[3][3];
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
			temp[i][j] = m[i][j];
	len = sqrtf(temp[0][0] * temp[0][0] + temp[0][1] * temp[0][1] + temp[0][2] * temp[0][2]);
	len = (len == 0.f) ? 1.f : 1.f / len;
	temp[0][0] *= len; temp[0][1] *= len; temp[0][2] *= len;
	temp[2][0] = (temp[0][1] * temp[1][2] - temp[0][2] * temp[1][1]);
	temp[2][1] = (temp[0][2] * temp[1][0] - temp[0][0] * temp[1][2]);
	temp[2][2] = (temp[0][0] * temp[1][1] - temp[0][1] * temp[1][0]);
	len = sqrtf(temp[2][0] * temp[2][0] + temp[2][1] * temp[2][1] + temp[2][2] * temp[2][2]);
	len = (len == 0.f) ? 1.f : 1.f / len;
	temp[2][0] *= len; temp[2][1] *= len; temp[2][2] *= len;
	temp[1][0] = (temp[2][1] * temp[0][2] - temp[2][2] * temp[0][1]);
	temp[1][1] = (temp[2][2] * temp[0][0] - temp[2][0] * temp[0][2]);
	temp[1][2] = (temp[2][0] * temp[0][1] - temp[2][1] * temp[0][0]);
	len = sqrtf(temp[1][0] * temp[1][0] + temp[1][1] * temp[1][1] + temp[1][2] * temp[1][2]);
	len = (len == 0.f) ? 1.f : 1.f / len;
	temp[1][0] *= len; temp[1][1] *= len; temp[1][2] *= len;
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
			m[i][j] = temp[i][j];
}

Quat QuatFromMatrix(const M4 &mat) {  // quaternion.cpp:69-106
	float o[4][4];
	std::memcpy(o, mat.m, sizeof(o));
	orthoNormalize(o);
	Quat q;
	const float trace = o[0][0] + o[1][1] + o[2][2] + 1.f;
	if (trace > 1e-6f) {
		const float s = sqrtf(trace) * 2.f;
		q.v.x = (o[1][2] - o[2][1]) / s;
		q.v.y = (o[2][0] - o[0][2]) / s;
		q.v.z = (o[0][1] - o[1][0]) / s;
		q.w = 0.25f * s;
	} else {
		if (o[0][0] > o[1][1] && o[0][0] > o[2][2]) {
			const float s = sqrtf(1.f + o[0][0] - o[1][1] - o[2][2]) * 2.f;
			q.v.x = 0.25f * s;
			q.v.y = (o[0][1] + o[1][0]) / s;
			q.v.z = (o[2][0] + o[0][2]) / s;
			q.w = (o[1][2] - o[2][1]) / s;
		} else if (o[1][1] > o[2][2]) {
			const float s = sqrtf(1.f + o[1][1] - o[0][0] - o[2][2]) * 2.f;
			q.v.x = (o[0][1] + o[1][0]) / s;
			q.v.y = 0.25f * s;
			q.v.z = (o[1][2] + o[2][1]) / s;
			q.w = (o[2][0] - o[0][2]) / s;
		} else {
			const float s = sqrtf(1.f + o[2][2] - o[0][0] - o[1][1]) * 2.f;
			q.v.x = (o[2][0] + o[0][2]) / s;
			q.v.y = (o[1][2] + o[2][1]) / s;
			q.v.z = 0.25f * s;
			q.w = (o[0][1] - o[1][0]) / s;
		}
	}
	return q;
}

void QuatToMatrix(const Quat &q, float m[4][4]) {  // quaternion.cpp:115-140
	const float xx = q.v.x * q.v.x, yy = q.v.y * q.v.y, zz = q.v.z * q.v.z;
	const float xy = q.v.x * q.v.y, xz = q.v.x * q.v.z, yz = q.v.y * q.v.z;
	const float xw = q.v.x * q.w, yw = q.v.y * q.w, zw = q.v.z * q.w;
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

Quat Slerp(float t, const Quat &q1, const Quat &q2) {  // quaternion.cpp:143-164
	float cos_phi = QDot(q1, q2);
	const float sign = (cos_phi > 0.f) ? 1.f : -1.f;
	cos_phi *= sign;
	float f1, f2;
	if (1.f - cos_phi > 1e-6f) {
		const float phi = AcosF(cos_phi);
		const float sin_phi = SinF(phi);
		f1 = SinF((1.f - t) * phi) / sin_phi;
		f2 = SinF(t * phi) / sin_phi;
	} else {
		f1 = 1.f - t;
		f2 = t;
	}
	return QAdd(QScale(f1, q1), QScale(sign * f2, q2));
}

// ---------------------------------------------------------------------------
// InterpolatedTransform / MotionSystem: src/luxrays/core/geometry/motionsystem.cpp
// ---------------------------------------------------------------------------
struct Decomposed {  // motionsystem.cpp:168-276
	float Sx = 0, Sy = 0, Sz = 0, Sxy = 0, Sxz = 0, Syz = 0;
	M4 R;
	float Tx = 0, Ty = 0, Tz = 0, Px = 0, Py = 0, Pz = 0, Pw = 0;
	bool Valid = false;
	Decomposed() { R = M4::Identity(); }
	explicit Decomposed(const M4 &m) {
		R = m;
		if (R.m[3][3] == 0) return;
		for (int i = 0; i < 4; ++i)
			for (int j = 0; j < 4; ++j)
				R.m[i][j] /= R.m[3][3];
		M4 pmat = R;
		for (int i = 0; i < 3; ++i) pmat.m[i][3] = 0.f;
		pmat.m[3][3] = 1.f;
		if (Determinant(pmat) == 0.f) return;
		if (R.m[3][0] != 0.f || R.m[3][1] != 0.f || R.m[3][2] != 0.f) {
			const float prhs[4] = {R.m[3][0], R.m[3][1], R.m[3][2], R.m[3][3]};
			M4 inv;
			if (!Inverse(pmat, &inv)) return;
			const M4 A = Transpose(inv);
			float psol[4];
			for (int r = 0; r < 4; ++r)
				psol[r] = A.m[r][0] * prhs[0] + A.m[r][1] * prhs[1] + A.m[r][2] * prhs[2] + A.m[r][3] * prhs[3];
			Px = psol[0]; Py = psol[1]; Pz = psol[2]; Pw = psol[3];
			R.m[3][0] = R.m[3][1] = R.m[3][2] = 0.f;
			R.m[3][3] = 1.f;
		}
		Tx = R.m[0][3]; Ty = R.m[1][3]; Tz = R.m[2][3];
		for (int i = 0; i < 3; ++i) R.m[i][3] = 0.f;
		V3 row[3];
		for (int i = 0; i < 3; ++i) row[i] = mk(R.m[i][0], R.m[i][1], R.m[i][2]);
		Sx = Length(row[0]);
		row[0] = scale(row[0], 1.f / Sx);
		Sxy = Dot(row[0], row[1]);
		row[1] = sub(row[1], scale(row[0], Sxy));
		Sy = Length(row[1]);
		row[1] = scale(row[1], 1.f / Sy);
		Sxy /= Sy;
		Sxz = Dot(row[0], row[2]);
		row[2] = sub(row[2], scale(row[0], Sxz));
		Syz = Dot(row[1], row[2]);
		row[2] = sub(row[2], scale(row[1], Syz));
		Sz = Length(row[2]);
		row[2] = scale(row[2], 1.f / Sz);
		Sxz /= Sz;
		Syz /= Sz;
		if (Dot(row[0], Cross(row[1], row[2])) < 0.f) {
			Sx *= -1.f; Sy *= -1.f; Sz *= -1.f;
			for (int i = 0; i < 3; ++i) row[i] = scale(row[i], -1.f);
		}
		for (int i = 0; i < 3; ++i) {
			R.m[i][0] = row[i].x; R.m[i][1] = row[i].y; R.m[i][2] = row[i].z;
		}
		Valid = true;
	}
};

struct Interp {  // motionsystem.cpp:37-73 (ctor), 89-159 (Sample), 75-87 (Bound)
	float startTime = 0, endTime = 0;
	M4 start, end;
	Decomposed startT, endT;
	Quat startQ, endQ;
	bool hasRotation = false, hasTranslation = false, hasScale = false;
	bool hasTranslationX = false, hasTranslationY = false, hasTranslationZ = false;
	bool hasScaleX = false, hasScaleY = false, hasScaleZ = false;
	bool isActive = false;

	Interp(float st, float et, const M4 &s, const M4 &e) {
		startTime = st; endTime = et; start = s; end = e;
		startQ.w = endQ.w = 1.f; startQ.v = endQ.v = mk(0, 0, 0);
		if (startTime == endTime) return;
		startT = Decomposed(start);
		endT = Decomposed(end);
		if (!startT.Valid || !endT.Valid)
			throw std::runtime_error("Singular matrix in InterpolatedTransform");
		startQ = QNormalize(QuatFromMatrix(startT.R));
		endQ = QNormalize(QuatFromMatrix(endT.R));
		hasTranslationX = startT.Tx != endT.Tx;
		hasTranslationY = startT.Ty != endT.Ty;
		hasTranslationZ = startT.Tz != endT.Tz;
		hasTranslation = hasTranslationX || hasTranslationY || hasTranslationZ;
		hasScaleX = startT.Sx != endT.Sx;
		hasScaleY = startT.Sy != endT.Sy;
		hasScaleZ = startT.Sz != endT.Sz;
		hasScale = hasScaleX || hasScaleY || hasScaleZ;
		hasRotation = fabsf(QDot(startQ, endQ) - 1.f) >= 1e-6f;
		isActive = hasTranslation || hasScale || hasRotation;
	}

	M4 Sample(float time) const {
		if (!isActive) return start;
		if (time <= startTime) return start;
		if (time >= endTime) return end;
		const float w = endTime - startTime;
		const float d = time - startTime;
		const float le = d / w;
		M4 r;
		if (hasTranslation && !(hasScale || hasRotation)) {
			r = start;
			if (hasTranslationX) r.m[0][3] = Lerp(le, startT.Tx, endT.Tx);
			if (hasTranslationY) r.m[1][3] = Lerp(le, startT.Ty, endT.Ty);
			if (hasTranslationZ) r.m[2][3] = Lerp(le, startT.Tz, endT.Tz);
			return r;
		}
		if (hasRotation) {
			const Quat q = Slerp(le, startQ, endQ);
			QuatToMatrix(q, r.m);
		} else
			r = startT.R;
		if (hasScale) {
			const float Sx = Lerp(le, startT.Sx, endT.Sx);
			const float Sy = Lerp(le, startT.Sy, endT.Sy);
			const float Sz = Lerp(le, startT.Sz, endT.Sz);
			for (int j = 0; j < 3; ++j) {
				r.m[0][j] = Sx * r.m[0][j];
				r.m[1][j] = Sy * r.m[1][j];
				r.m[2][j] = Sz * r.m[2][j];
			}
		} else {
			for (int j = 0; j < 3; ++j) {
				r.m[0][j] = startT.Sx * r.m[0][j];
				r.m[1][j] = startT.Sy * r.m[1][j];
				r.m[2][j] = startT.Sz * r.m[2][j];
			}
		}
		r.m[0][3] = hasTranslationX ? Lerp(le, startT.Tx, endT.Tx) : startT.Tx;
		r.m[1][3] = hasTranslationY ? Lerp(le, startT.Ty, endT.Ty) : startT.Ty;
		r.m[2][3] = hasTranslationZ ? Lerp(le, startT.Tz, endT.Tz) : startT.Tz;
		return r;
	}

	BBox Bound(const BBox &ibox, bool storingGlobal2Local) const {
		BBox tbox;
		const float N = 1024.f;
		for (float i = 0; i <= N; ++i) {
			const float t = Lerp(i / N, startTime, endTime);
			M4 m = Sample(t);
			if (storingGlobal2Local) {
				M4 inv;
				if (!Inverse(m, &inv)) throw std::runtime_error("Singular matrix in Bound");
				m = inv;
			}
			tbox = Union(tbox, xfBBox(m, ibox));
		}
		return tbox;
	}
};

struct Motion {  // motionsystem.cpp:296-365
	std::vector<float> times;
	std::vector<Interp> its;
	Motion(const std::vector<float> &t, const std::vector<M4> &tr) {
		times = t;
		size_t prev = 0;
		for (size_t k = 0; k < times.size(); ++k) {
			its.push_back(Interp(times[prev], times[k], tr[prev], tr[k]));
			prev = k;
		}
		its.push_back(Interp(times[prev], times[prev], tr[prev], tr[prev]));
	}
	M4 Sample(float time) const {
		size_t index = std::upper_bound(times.begin(), times.end(), time) - times.begin();
		index = Min(index, times.size() - 1);
		return its[index].Sample(time);
	}
	BBox Bound(const BBox &ibox, bool storingGlobal2Local) const {
		BBox r;
		for (const Interp &it : its)
			r = Union(r, it.Bound(ibox, storingGlobal2Local));
		return r;
	}
};

// ---------------------------------------------------------------------------
// Meshes (flattened arrays; mesh index = add order, dataset.cpp:66-81)
// ---------------------------------------------------------------------------
struct MeshSet {
	std::vector<unsigned> vertOff, triOff;
	std::vector<V3> verts;
	std::vector<unsigned> tris;  // local indices, 3 per triangle
	unsigned nMeshes() const { return (unsigned)vertOff.size() - 1; }
	const V3 &vertex(unsigned mesh, unsigned v) const { return verts[vertOff[mesh] + v]; }
	const unsigned *tri(unsigned mesh, unsigned t) const { return &tris[3 * ((size_t)triOff[mesh] + t)]; }
	void load(unsigned n, const unsigned *vo, const unsigned *to, const float *v, const unsigned *t) {
		vertOff.assign(vo, vo + n + 1);
		triOff.assign(to, to + n + 1);
		verts.resize(vertOff[n]);
		std::memcpy(verts.data(), v, sizeof(float) * 3 * (size_t)vertOff[n]);
		tris.assign(t, t + 3 * (size_t)triOff[n]);
	}
	// TriangleMesh::GetBBox: trianglemesh.cpp:72-83
	BBox meshBBox(unsigned mesh) const {
		BBox b;
		for (unsigned i = vertOff[mesh]; i < vertOff[mesh + 1]; ++i)
			b = Union(b, verts[i]);
		return b;
	}
};

// ---------------------------------------------------------------------------
// BVH build types: include/luxrays/core/bvh/bvhbuild.h:40-65
// ---------------------------------------------------------------------------
struct BVHParams { unsigned treeType; int costSamples, isectCost, traversalCost; float emptyBonus; };

struct TreeNode {
	BBox bbox;
	unsigned meshIndex = 0, triangleIndex = 0;                             // triangleLeaf
	unsigned leafIndex = 0, transformIndex = 0, motionIndex = 0, meshOffsetIndex = 0;  // bvhLeaf
	TreeNode *leftChild = nullptr, *rightSibling = nullptr;
};

// ---------------------------------------------------------------------------
// CLASSIC builder: src/luxrays/core/bvh/bvhclassicbuild.cpp:35-233
// ---------------------------------------------------------------------------
inline bool ltf(const TreeNode *n, unsigned axis, float v) {  // :35-47
	return n->bbox.pMax[axis] + n->bbox.pMin[axis] < v;
}

void FindBestSplit(const BVHParams &params, std::vector<TreeNode *> &list,
		unsigned begin, unsigned end, float *splitValue, unsigned *bestAxis) {  // :51-118
	if (end - begin == 2) {
		*splitValue = (list[begin]->bbox.pMax[0] + list[begin]->bbox.pMin[0] +
				list[end - 1]->bbox.pMax[0] + list[end - 1]->bbox.pMin[0]) / 2;
		*bestAxis = 0;
		return;
	}
	V3 mean2 = mk(0, 0, 0), var = mk(0, 0, 0);
	for (unsigned i = begin; i < end; i++)
		mean2 = add(mean2, add(list[i]->bbox.pMax, list[i]->bbox.pMin));
	{
		const float inv = 1.f / static_cast<float>(end - begin);
		mean2.x *= inv; mean2.y *= inv; mean2.z *= inv;
	}
	for (unsigned i = begin; i < end; i++) {
		V3 v = sub(add(list[i]->bbox.pMax, list[i]->bbox.pMin), mean2);
		v.x *= v.x; v.y *= v.y; v.z *= v.z;
		var = add(var, v);
	}
	if (var.x > var.y && var.x > var.z) *bestAxis = 0;
	else if (var.y > var.z) *bestAxis = 1;
	else *bestAxis = 2;

	if (params.costSamples > 1) {
		BBox nodeBounds;
		for (unsigned i = begin; i < end; i++)
			nodeBounds = Union(nodeBounds, list[i]->bbox);
		const V3 d = sub(nodeBounds.pMax, nodeBounds.pMin);
		const float invTotalSA = 1.f / nodeBounds.SurfaceArea();
		const float increment = 2 * d[*bestAxis] / (params.costSamples + 1);
		float bestCost = INFINITY;
		*splitValue = mean2[*bestAxis];
		for (float splitVal = 2 * nodeBounds.pMin[*bestAxis] + increment;
				splitVal < 2 * nodeBounds.pMax[*bestAxis]; splitVal += increment) {
			int nBelow = 0, nAbove = 0;
			BBox bbBelow, bbAbove;
			for (unsigned j = begin; j < end; j++) {
				if ((list[j]->bbox.pMax[*bestAxis] + list[j]->bbox.pMin[*bestAxis]) < splitVal) {
					nBelow++;
					bbBelow = Union(bbBelow, list[j]->bbox);
				} else {
					nAbove++;
					bbAbove = Union(bbAbove, list[j]->bbox);
				}
			}
			const float pBelow = bbBelow.SurfaceArea() * invTotalSA;
			const float pAbove = bbAbove.SurfaceArea() * invTotalSA;
			const float eb = (nAbove == 0 || nBelow == 0) ? params.emptyBonus : 0.f;
			const float cost = params.traversalCost + params.isectCost * (1.f - eb) * (pBelow * nBelow + pAbove * nAbove);
			if (cost < bestCost) {
				bestCost = cost;
				*splitValue = splitVal;
			}
		}
	} else
		*splitValue = mean2[*bestAxis];
}

TreeNode *BuildClassic(unsigned *nNodes, const BVHParams &params,
		std::vector<TreeNode *> &leafList, unsigned begin, unsigned end, unsigned axis,
		std::vector<std::unique_ptr<TreeNode>> &pool) {  // :121-178
	unsigned splitAxis = axis;
	float splitValue = 0.f;
	*nNodes += 1;
	if (end - begin == 1) {
		pool.emplace_back(new TreeNode(*leafList[begin]));
		return pool.back().get();
	}
	pool.emplace_back(new TreeNode());
	TreeNode *parent = pool.back().get();
	std::vector<unsigned> splits;
	splits.reserve(params.treeType + 1);
	splits.push_back(begin);
	splits.push_back(end);
	for (unsigned i = 2; i <= params.treeType; i *= 2) {
		for (unsigned j = 0, offset = 0; j + offset < i && splits.size() > j + 1; j += 2) {
			if (splits[j + 1] - splits[j] < 2) {
				j--;  // unsigned wrap then += 2, exactly as the reference loop
				offset++;
				continue;
			}
			FindBestSplit(params, leafList, splits[j], splits[j + 1], &splitValue, &splitAxis);
			const unsigned ax = splitAxis;
			const float sv = splitValue;
			auto it = std::partition(leafList.begin() + splits[j], leafList.begin() + splits[j + 1],
					[ax, sv](TreeNode *n) { return ltf(n, ax, sv); });
			unsigned middle = (unsigned)std::distance(leafList.begin(), it);
			middle = Max(splits[j] + 1, Min(splits[j + 1] - 1, middle));
			splits.insert(splits.begin() + j + 1, middle);
		}
	}
	TreeNode *child = BuildClassic(nNodes, params, leafList, splits[0], splits[1], splitAxis, pool);
	parent->leftChild = child;
	parent->bbox = child->bbox;
	TreeNode *lastChild = child;
	for (size_t i = 1; i < splits.size() - 1; i++) {
		child = BuildClassic(nNodes, params, leafList, splits[i], splits[i + 1], splitAxis, pool);
		lastChild->rightSibling = child;
		parent->bbox = Union(parent->bbox, child->bbox);
		lastChild = child;
	}
	return parent;
}

// BuildBVHArray: bvhclassicbuild.cpp:181-220
unsigned BuildClassicArray(const MeshSet *meshes, TreeNode *node, unsigned offset, orc_bvh_node *arr) {
	while (node) {
		orc_bvh_node *a = &arr[offset];
		if (node->leftChild) {
			std::memcpy(&a->bvhNode.bboxMin[0], &node->bbox.pMin.x, sizeof(float) * 3);
			std::memcpy(&a->bvhNode.bboxMax[0], &node->bbox.pMax.x, sizeof(float) * 3);
			offset = BuildClassicArray(meshes, node->leftChild, offset + 1, arr);
			a->nodeData = offset;
		} else {
			if (meshes) {
				const unsigned *t = meshes->tri(node->meshIndex, node->triangleIndex);
				a->triangleLeaf.v[0] = t[0];
				a->triangleLeaf.v[1] = t[1];
				a->triangleLeaf.v[2] = t[2];
				a->triangleLeaf.meshIndex = node->meshIndex;
				a->triangleLeaf.triangleIndex = node->triangleIndex;
			} else {
				a->bvhLeaf.leafIndex = node->leafIndex;
				a->bvhLeaf.transformIndex = node->transformIndex;
				a->bvhLeaf.motionIndex = node->motionIndex;
				a->bvhLeaf.meshOffsetIndex = node->meshOffsetIndex;
			}
			++offset;
			a->nodeData = offset | 0x80000000u;
		}
		a->pad0 = 0;
		node = node->rightSibling;
	}
	return offset;
}

std::vector<orc_bvh_node> BuildClassicBVH(const BVHParams &params, const MeshSet *meshes,
		std::vector<TreeNode *> &leafList) {
	unsigned nNodes = 0;
	std::vector<std::unique_ptr<TreeNode>> pool;
	TreeNode *root = BuildClassic(&nNodes, params, leafList, 0, (unsigned)leafList.size(), 2, pool);
	std::vector<orc_bvh_node> arr(nNodes);
	BuildClassicArray(meshes, root, 0, arr.data());
	return arr;
}

// ---------------------------------------------------------------------------
// Binned-SAH N-ary builder, the stand-in for EMBREE_BINNED_SAH
// (rtcBuildBVH with maxLeafSize=1, maxBranchingFactor=treeType,
// bvhembreebuild.cpp:218-280).  Embree 3.12.2 is not vendored; this is the
// oracle's own binned SAH.  Its output is flattened with the reference's
// Embree flattening rules (BuildEmbreeBVHArray, bvhembreebuild.cpp:105-163).
// ---------------------------------------------------------------------------
struct SahInner {
	unsigned n;
	unsigned child[8];  // high bit set = leaf (low bits = leafList index)
	BBox cbox[8];
};

class SahBuilder {
public:
	SahBuilder(const std::vector<TreeNode *> &leafs, unsigned width, int nThreads)
		: L(leafs), W(width), budget(nThreads > 1 ? nThreads - 1 : 0) {
		const size_t n = L.size();
		ids.resize(n);
		cen.resize(n);
		for (size_t i = 0; i < n; ++i) {
			ids[i] = (unsigned)i;
			cen[i] = add(L[i]->bbox.pMin, L[i]->bbox.pMax);
		}
		inner.reset(new SahInner[n > 1 ? n - 1 : 1]);
		innerCount = 0;
	}

	unsigned build() {
		if (L.size() == 1) return 0x80000000u;
		BBox all;
		for (size_t i = 0; i < L.size(); ++i) all = Union(all, L[i]->bbox);
		return buildRange(0, (unsigned)L.size(), all);
	}

	const std::vector<TreeNode *> &L;
	const unsigned W;
	std::vector<unsigned> ids;
	std::vector<V3> cen;
	std::unique_ptr<SahInner[]> inner;
	std::atomic<unsigned> innerCount;
	std::atomic<int> budget;

private:
	struct Range { unsigned b, e; BBox box; };

	// binned SAH split of [b,e); returns the split point and the two boxes
	unsigned split(unsigned b, unsigned e, BBox *lb, BBox *rb) {
		const int NB = 32;
		BBox cb;
		for (unsigned i = b; i < e; ++i) cb = Union(cb, cen[ids[i]]);
		float bestCost = INFINITY;
		int bestAxis = -1, bestBin = -1;
		for (int axis = 0; axis < 3; ++axis) {
			const float ext = cb.pMax[axis] - cb.pMin[axis];
			if (!(ext > 0.f)) continue;
			const float k = (float)NB * 0.99999f / ext;
			BBox bbox[NB];
			unsigned cnt[NB] = {0};
			for (unsigned i = b; i < e; ++i) {
				int bin = (int)((cen[ids[i]][axis] - cb.pMin[axis]) * k);
				bin = bin < 0 ? 0 : (bin >= NB ? NB - 1 : bin);
				cnt[bin]++;
				bbox[bin] = Union(bbox[bin], L[ids[i]]->bbox);
			}
			float rArea[NB];
			unsigned rCnt[NB];
			BBox acc;
			unsigned c = 0;
			for (int i = NB - 1; i > 0; --i) {
				acc = Union(acc, bbox[i]);
				c += cnt[i];
				rArea[i] = acc.SurfaceArea();
				rCnt[i] = c;
			}
			acc = BBox();
			c = 0;
			for (int i = 0; i < NB - 1; ++i) {
				acc = Union(acc, bbox[i]);
				c += cnt[i];
				if (c == 0 || rCnt[i + 1] == 0) continue;
				const float cost = acc.SurfaceArea() * (float)c + rArea[i + 1] * (float)rCnt[i + 1];
				if (cost < bestCost) { bestCost = cost; bestAxis = axis; bestBin = i; }
			}
		}
		unsigned mid;
		if (bestAxis < 0) {
			mid = b + (e - b) / 2;
		} else {
			const float ext = cb.pMax[bestAxis] - cb.pMin[bestAxis];
			const float k = (float)NB * 0.99999f / ext;
			const float mn = cb.pMin[bestAxis];
			const int ax = bestAxis, bb = bestBin;
			auto it = std::partition(ids.begin() + b, ids.begin() + e, [&](unsigned id) {
				int bin = (int)((cen[id][ax] - mn) * k);
				bin = bin < 0 ? 0 : (bin >= NB ? NB - 1 : bin);
				return bin <= bb;
			});
			mid = (unsigned)(it - ids.begin());
			if (mid == b || mid == e) mid = b + (e - b) / 2;
		}
		BBox l, r;
		for (unsigned i = b; i < mid; ++i) l = Union(l, L[ids[i]]->bbox);
		for (unsigned i = mid; i < e; ++i) r = Union(r, L[ids[i]]->bbox);
		*lb = l; *rb = r;
		return mid;
	}

	unsigned buildRange(unsigned b, unsigned e, const BBox &box) {
		std::vector<Range> ch;
		ch.push_back(Range{b, e, box});
		while (ch.size() < W) {
			int pick = -1;
			float best = -1.f;
			for (size_t i = 0; i < ch.size(); ++i) {
				if (ch[i].e - ch[i].b < 2) continue;
				const float a = ch[i].box.SurfaceArea();
				if (a > best) { best = a; pick = (int)i; }
			}
			if (pick < 0) break;
			const Range r = ch[pick];
			BBox lb, rb;
			const unsigned mid = split(r.b, r.e, &lb, &rb);
			ch[pick] = Range{r.b, mid, lb};
			ch.insert(ch.begin() + pick + 1, Range{mid, r.e, rb});
		}
		const unsigned me = innerCount++;
		SahInner &node = inner[me];
		node.n = (unsigned)ch.size();
		std::vector<std::thread> threads;
		for (size_t i = 0; i < ch.size(); ++i) {
			node.cbox[i] = ch[i].box;
			if (ch[i].e - ch[i].b == 1) {
				node.child[i] = 0x80000000u | ids[ch[i].b];
				continue;
			}
			const bool big = ch[i].e - ch[i].b > 50000;
			if (big && budget.fetch_sub(1) > 0) {
				const Range r = ch[i];
				unsigned *slot = &node.child[i];
				threads.emplace_back([this, r, slot]() { *slot = buildRange(r.b, r.e, r.box); budget.fetch_add(1); });
			} else {
				if (big) budget.fetch_add(1);
				node.child[i] = buildRange(ch[i].b, ch[i].e, ch[i].box);
			}
		}
		for (auto &t : threads) t.join();
		return me;
	}
};

// BuildEmbreeBVHArray: bvhembreebuild.cpp:105-163
unsigned FlattenSah(const MeshSet *meshes, const SahBuilder &sb, unsigned ref,
		unsigned offset, orc_bvh_node *arr) {
	orc_bvh_node *a = &arr[offset];
	a->pad0 = 0;
	if (!(ref & 0x80000000u)) {
		const SahInner &node = sb.inner[ref];
		++offset;
		BBox bbox;
		for (unsigned i = 0; i < node.n; ++i) {
			const unsigned childIndex = offset;
			offset = FlattenSah(meshes, sb, node.child[i], childIndex, arr);
			if (!(node.child[i] & 0x80000000u))
				arr[childIndex].nodeData = offset;
			bbox = Union(bbox, node.cbox[i]);
		}
		std::memcpy(&a->bvhNode.bboxMin[0], &bbox.pMin.x, sizeof(float) * 3);
		std::memcpy(&a->bvhNode.bboxMax[0], &bbox.pMax.x, sizeof(float) * 3);
	} else {
		const TreeNode *leaf = sb.L[ref & 0x7fffffffu];
		if (meshes) {
			const unsigned *t = meshes->tri(leaf->meshIndex, leaf->triangleIndex);
			a->triangleLeaf.v[0] = t[0];
			a->triangleLeaf.v[1] = t[1];
			a->triangleLeaf.v[2] = t[2];
			a->triangleLeaf.meshIndex = leaf->meshIndex;
			a->triangleLeaf.triangleIndex = leaf->triangleIndex;
		} else {
			a->bvhLeaf.leafIndex = leaf->leafIndex;
			a->bvhLeaf.transformIndex = leaf->transformIndex;
			a->bvhLeaf.motionIndex = leaf->motionIndex;
			a->bvhLeaf.meshOffsetIndex = leaf->meshOffsetIndex;
		}
		++offset;
		a->nodeData = offset | 0x80000000u;
	}
	return offset;
}

std::vector<orc_bvh_node> BuildSahBVH(const BVHParams &params, const MeshSet *meshes,
		std::vector<TreeNode *> &leafList, int nThreads) {
	SahBuilder sb(leafList, params.treeType, nThreads);
	const unsigned root = sb.build();
	const size_t nNodes = (size_t)sb.innerCount.load() + leafList.size();
	std::vector<orc_bvh_node> arr(nNodes);
	arr[0].nodeData = FlattenSah(meshes, sb, root, 0, arr.data());
	if (root & 0x80000000u)
		arr[0].nodeData |= 0x80000000u;
	return arr;
}

std::vector<orc_bvh_node> BuildTree(int builder, const BVHParams &params, const MeshSet *meshes,
		std::vector<TreeNode *> &leafList, int nThreads) {
	if (builder == ORC_BUILDER_CLASSIC)
		return BuildClassicBVH(params, meshes, leafList);
	if (builder == ORC_BUILDER_SAH)
		return BuildSahBVH(params, meshes, leafList, nThreads);
	throw std::runtime_error("unknown builder");
}

BVHParams MakeParams(int treeType, int costSamples, int isectCost, int travCost, float emptyBonus) {
	// bvhaccel.cpp:49-70
	BVHParams p;
	if (treeType <= 2) p.treeType = 2;
	else if (treeType <= 4) p.treeType = 4;
	else p.treeType = 8;
	p.costSamples = costSamples;
	p.isectCost = isectCost;
	p.traversalCost = travCost;
	p.emptyBonus = emptyBonus;
	return p;
}

// ---------------------------------------------------------------------------
// Parallel batch driver: static chunks of 4096 rays, round-robin over threads
// ---------------------------------------------------------------------------
template <class F>
void ParallelChunks(size_t n, int nThreads, F fn) {
	const size_t CH = 4096;
	const size_t nChunks = (n + CH - 1) / CH;
	if (nThreads < 1) nThreads = 1;
	if ((size_t)nThreads > nChunks) nThreads = (int)(nChunks ? nChunks : 1);
	auto work = [&](int tid) {
		for (size_t c = tid; c < nChunks; c += nThreads) {
			const size_t b = c * CH, e = Min(n, b + CH);
			fn(b, e);
		}
	};
	if (nThreads == 1) { work(0); return; }
	std::vector<std::thread> th;
	for (int t = 0; t < nThreads; ++t) th.emplace_back(work, t);
	for (auto &t : th) t.join();
}

void writeMiss(const orc_ray &r, orc_rayhit *h) {
	h->t = r.maxt; h->b1 = 0.f; h->b2 = 0.f;
	h->meshIndex = NULL_INDEX; h->triangleIndex = NULL_INDEX;
}

}  // namespace

// ===========================================================================
// Single-level BVH
// ===========================================================================
struct orc_bvh {
	MeshSet meshes;
	std::vector<orc_bvh_node> tree;
	unsigned nNodes = 0;

	// BVHAccel::Intersect: src/luxrays/accelerators/bvhaccel.cpp:170-260
	bool Intersect(const Ray &initialRay, orc_rayhit *hit, unsigned long long *cnt) const {
		hit->t = initialRay.maxt;
		hit->meshIndex = NULL_INDEX;
		if (!nNodes) return false;
		Ray ray = initialRay;
		unsigned currentNode = 0;
		const unsigned stopNode = tree[0].nodeData & 0x7fffffffu;
		float t, b1, b2;
		unsigned long long nIn = 0, nLeaf = 0;
		while (currentNode < stopNode) {
			const orc_bvh_node &node = tree[currentNode];
			const unsigned nodeData = node.nodeData;
			if (nodeData & 0x80000000u) {
				++nLeaf;
				const unsigned mi = node.triangleLeaf.meshIndex;
				const V3 &p0 = meshes.vertex(mi, node.triangleLeaf.v[0]);
				const V3 &p1 = meshes.vertex(mi, node.triangleLeaf.v[1]);
				const V3 &p2 = meshes.vertex(mi, node.triangleLeaf.v[2]);
				if (TriangleIntersect(ray, p0, p1, p2, &t, &b1, &b2)) {
					if (t < hit->t) {
						ray.maxt = t;
						hit->t = t;
						hit->b1 = b1;
						hit->b2 = b2;
						hit->meshIndex = mi;
						hit->triangleIndex = node.triangleLeaf.triangleIndex;
					}
				}
				++currentNode;
			} else {
				++nIn;
				const V3 pmin = mk(node.bvhNode.bboxMin[0], node.bvhNode.bboxMin[1], node.bvhNode.bboxMin[2]);
				const V3 pmax = mk(node.bvhNode.bboxMax[0], node.bvhNode.bboxMax[1], node.bvhNode.bboxMax[2]);
				if (BBoxIntersectP(ray, pmin, pmax))
					++currentNode;
				else
					currentNode = nodeData;
			}
		}
		if (cnt) { cnt[0] += nIn; cnt[1] += nLeaf; }
		return hit->meshIndex != NULL_INDEX;
	}
};

extern "C" {

int orc_triangle_intersect(const orc_ray *r, const float p0[3], const float p1[3],
		const float p2[3], float *t, float *b1, float *b2) {
	return TriangleIntersect(toRay(*r), mk(p0[0], p0[1], p0[2]), mk(p1[0], p1[1], p1[2]),
			mk(p2[0], p2[1], p2[2]), t, b1, b2) ? 1 : 0;
}

int orc_bbox_intersectp(const orc_ray *r, const float pmin[3], const float pmax[3],
		float *t0, float *t1) {
	return BBoxIntersectP(toRay(*r), mk(pmin[0], pmin[1], pmin[2]), mk(pmax[0], pmax[1], pmax[2]), t0, t1) ? 1 : 0;
}

float orc_epsilon(float v) { return E(v); }

void orc_set_libm(int correctlyRounded) { g_crMath = correctlyRounded ? 1 : 0; }

float orc_epsilon_bbox(const float pmin[3], const float pmax[3]) {
	BBox b;
	b.pMin = mk(pmin[0], pmin[1], pmin[2]);
	b.pMax = mk(pmax[0], pmax[1], pmax[2]);
	return E(b);
}

orc_bvh *orc_bvh_create(unsigned nMeshes, const unsigned *vertOff, const unsigned *triOff,
		const float *verts, const unsigned *tris, int builder, int treeType,
		int costSamples, int isectCost, int travCost, float emptyBonus, int nThreads) {
	try {
		std::unique_ptr<orc_bvh> b(new orc_bvh());
		b->meshes.load(nMeshes, vertOff, triOff, verts, tris);
		const size_t nTris = b->meshes.triOff[nMeshes];
		if (nTris == 0) { b->nNodes = 0; return b.release(); }  // bvhaccel.cpp:81-89
		// bvhaccel.cpp:96-134: leaf boxes, expanded by MachineEpsilon::E
		std::vector<TreeNode> nodes(nTris);
		std::vector<TreeNode *> list(nTris);
		for (unsigned m = 0; m < nMeshes; ++m) {
			const unsigned base = b->meshes.triOff[m];
			const unsigned cnt = b->meshes.triOff[m + 1] - base;
			for (unsigned i = 0; i < cnt; ++i) {
				TreeNode *n = &nodes[base + i];
				const unsigned *t = b->meshes.tri(m, i);
				n->bbox = Union(BBox(b->meshes.vertex(m, t[0]), b->meshes.vertex(m, t[1])), b->meshes.vertex(m, t[2]));
				n->bbox.Expand(E(n->bbox));
				n->meshIndex = m;
				n->triangleIndex = i;
				list[base + i] = n;
			}
		}
		const BVHParams p = MakeParams(treeType, costSamples, isectCost, travCost, emptyBonus);
		b->tree = BuildTree(builder, p, &b->meshes, list, nThreads);
		b->nNodes = (unsigned)b->tree.size();
		return b.release();
	} catch (...) {
		return nullptr;
	}
}

void orc_bvh_destroy(orc_bvh *b) { delete b; }
unsigned orc_bvh_node_count(const orc_bvh *b) { return b->nNodes; }
const orc_bvh_node *orc_bvh_nodes(const orc_bvh *b) { return b->tree.data(); }

void orc_bvh_intersect(const orc_bvh *b, const orc_ray *rays, orc_rayhit *hits,
		size_t n, int nThreads, unsigned long long *counters) {
	std::atomic<unsigned long long> cIn(0), cLeaf(0);
	ParallelChunks(n, nThreads, [&](size_t s, size_t e) {
		unsigned long long c[2] = {0, 0};
		for (size_t i = s; i < e; ++i) {
			if (rays[i].flags & 1u) continue;  // RAY_FLAGS_MASKED, bvh.cl:243-244
			orc_rayhit h;
			if (b->Intersect(toRay(rays[i]), &h, counters ? c : nullptr))
				hits[i] = h;
			else
				writeMiss(rays[i], &hits[i]);
		}
		cIn += c[0];
		cLeaf += c[1];
	});
	if (counters) { counters[0] = cIn.load(); counters[1] = cLeaf.load(); }
}

void orc_bruteforce(unsigned nMeshes, const unsigned *vertOff, const unsigned *triOff,
		const float *verts, const unsigned *tris, const orc_ray *rays, orc_rayhit *hits,
		size_t n, int nThreads) {
	MeshSet ms;
	ms.load(nMeshes, vertOff, triOff, verts, tris);
	ParallelChunks(n, nThreads, [&](size_t s, size_t e) {
		for (size_t i = s; i < e; ++i) {
			if (rays[i].flags & 1u) continue;
			Ray ray = toRay(rays[i]);
			orc_rayhit h;
			h.t = ray.maxt; h.b1 = 0.f; h.b2 = 0.f;
			h.meshIndex = NULL_INDEX; h.triangleIndex = NULL_INDEX;
			float t, b1, b2;
			for (unsigned m = 0; m < nMeshes; ++m) {
				const unsigned cnt = ms.triOff[m + 1] - ms.triOff[m];
				for (unsigned k = 0; k < cnt; ++k) {
					const unsigned *tv = ms.tri(m, k);
					if (TriangleIntersect(ray, ms.vertex(m, tv[0]), ms.vertex(m, tv[1]), ms.vertex(m, tv[2]), &t, &b1, &b2) &&
							t < h.t) {
						ray.maxt = t;
						h.t = t; h.b1 = b1; h.b2 = b2; h.meshIndex = m; h.triangleIndex = k;
					}
				}
			}
			hits[i] = h;
		}
	});
}

int orc_matrix_inverse(const float m[16], float out[16]) {
	M4 a, r;
	std::memcpy(a.m, m, sizeof(a.m));
	if (!Inverse(a, &r)) return 0;
	std::memcpy(out, r.m, sizeof(r.m));
	return 1;
}

static bool makeMotion(unsigned nKeys, const float *times, const float *mats, std::unique_ptr<Motion> *out) {
	std::vector<float> t(times, times + nKeys);
	std::vector<M4> ms(nKeys);
	for (unsigned k = 0; k < nKeys; ++k) std::memcpy(ms[k].m, mats + 16 * k, sizeof(ms[k].m));
	try {
		out->reset(new Motion(t, ms));
	} catch (...) {
		return false;
	}
	return true;
}

int orc_motion_sample(unsigned nKeys, const float *times, const float *mats, float time, float out[16]) {
	std::unique_ptr<Motion> m;
	if (!makeMotion(nKeys, times, mats, &m)) return 0;
	const M4 r = m->Sample(time);
	std::memcpy(out, r.m, sizeof(r.m));
	return 1;
}

int orc_motion_bound(unsigned nKeys, const float *times, const float *mats,
		const float pmin[3], const float pmax[3], float omin[3], float omax[3]) {
	std::unique_ptr<Motion> m;
	if (!makeMotion(nKeys, times, mats, &m)) return 0;
	BBox b;
	b.pMin = mk(pmin[0], pmin[1], pmin[2]);
	b.pMax = mk(pmax[0], pmax[1], pmax[2]);
	try {
		const BBox r = m->Bound(b, true);
		for (int i = 0; i < 3; ++i) { omin[i] = r.pMin[i]; omax[i] = r.pMax[i]; }
	} catch (...) {
		return 0;
	}
	return 1;
}

}  // extern "C"

// ===========================================================================
// Two-level BVH: src/luxrays/accelerators/mbvhaccel.cpp:58-357
// ===========================================================================
struct orc_mbvh {
	MeshSet base;                           // unique base meshes
	std::vector<unsigned> objMesh, objKind; // per DataSet mesh (object)
	std::vector<unsigned> leafIndex, transformIndex, motionIndex;
	std::vector<std::unique_ptr<orc_bvh>> uniqueLeafs;
	std::vector<M4> xfM, xfInv;             // Transform m / mInv per instance
	std::vector<std::unique_ptr<Motion>> motions;
	std::vector<BBox> leafBox;              // epsilon-expanded world boxes
	std::vector<orc_bvh_node> root;
	unsigned nRootNodes = 0;

	Ray leafRay(unsigned o, const Ray &ray) const {
		if (transformIndex[o] != NULL_INDEX)
			return xfRay(xfInv[transformIndex[o]], ray);
		if (motionIndex[o] != NULL_INDEX)
			return xfRay(motions[motionIndex[o]]->Sample(ray.time), ray);
		return ray;
	}

	bool Intersect(const Ray &ray, orc_rayhit *rayHit, unsigned long long *cnt) const {
		rayHit->t = ray.maxt;
		rayHit->meshIndex = NULL_INDEX;
		if (!nRootNodes) return false;
		bool insideLeafTree = false;
		unsigned currentRootNode = 0;
		const unsigned rootStopNode = root[0].nodeData & 0x7fffffffu;
		unsigned currentNode = currentRootNode;
		unsigned currentStopNode = rootStopNode;
		unsigned currentMeshOffset = 0;
		const orc_bvh_node *currentTree = root.data();
		const orc_bvh *currentLeaf = nullptr;
		Ray currentRay = ray;
		unsigned long long nIn = 0, nLeaf = 0;
		for (;;) {
			if (currentNode >= currentStopNode) {
				if (insideLeafTree) {
					currentTree = root.data();
					currentNode = currentRootNode;
					currentStopNode = rootStopNode;
					currentRay = ray;
					currentRay.maxt = rayHit->t;
					insideLeafTree = false;
					if (currentNode >= currentStopNode) break;
				} else
					break;
			}
			const orc_bvh_node &node = currentTree[currentNode];
			const unsigned nodeData = node.nodeData;
			if (nodeData & 0x80000000u) {
				++nLeaf;
				if (insideLeafTree) {
					const unsigned absoluteMeshIndex = node.triangleLeaf.meshIndex + currentMeshOffset;
					const unsigned bm = objMesh[absoluteMeshIndex];
					const V3 &p0 = base.vertex(bm, node.triangleLeaf.v[0]);
					const V3 &p1 = base.vertex(bm, node.triangleLeaf.v[1]);
					const V3 &p2 = base.vertex(bm, node.triangleLeaf.v[2]);
					float t, b1, b2;
					if (TriangleIntersect(currentRay, p0, p1, p2, &t, &b1, &b2)) {
						if (t < rayHit->t) {
							currentRay.maxt = t;
							rayHit->t = t;
							rayHit->b1 = b1;
							rayHit->b2 = b2;
							rayHit->meshIndex = absoluteMeshIndex;
							rayHit->triangleIndex = node.triangleLeaf.triangleIndex;
						}
					}
					++currentNode;
				} else {
					const unsigned o = node.bvhLeaf.meshOffsetIndex;
					currentLeaf = uniqueLeafs[node.bvhLeaf.leafIndex].get();
					currentTree = currentLeaf->tree.data();
					currentRay = leafRay(o, ray);
					currentRay.maxt = rayHit->t;
					currentMeshOffset = node.bvhLeaf.meshOffsetIndex;
					currentRootNode = currentNode + 1;
					currentNode = 0;
					currentStopNode = currentTree[0].nodeData & 0x7fffffffu;
					insideLeafTree = true;
				}
			} else {
				++nIn;
				const V3 pmin = mk(node.bvhNode.bboxMin[0], node.bvhNode.bboxMin[1], node.bvhNode.bboxMin[2]);
				const V3 pmax = mk(node.bvhNode.bboxMax[0], node.bvhNode.bboxMax[1], node.bvhNode.bboxMax[2]);
				if (BBoxIntersectP(currentRay, pmin, pmax))
					++currentNode;
				else
					currentNode = nodeData;
			}
		}
		if (cnt) { cnt[0] += nIn; cnt[1] += nLeaf; }
		return rayHit->meshIndex != NULL_INDEX;
	}
};

extern "C" {

orc_mbvh *orc_mbvh_create(unsigned nMeshes, const unsigned *vertOff, const unsigned *triOff,
		const float *verts, const unsigned *tris,
		unsigned nObjs, const unsigned *objMesh, const unsigned *objKind, const float *objMat,
		const unsigned *objKeyOff, const float *keyTimes, const float *keyMats,
		int builder, int treeType, int nThreads) {
	try {
		std::unique_ptr<orc_mbvh> b(new orc_mbvh());
		b->base.load(nMeshes, vertOff, triOff, verts, tris);
		b->objMesh.assign(objMesh, objMesh + nObjs);
		b->objKind.assign(objKind, objKind + nObjs);
		if (nObjs == 0) { b->nRootNodes = 0; return b.release(); }
		const BVHParams p = MakeParams(treeType, 0, 80, 10, 0.5f);
		std::map<unsigned, unsigned> uniqueByMesh;
		// one leaf BVH over a single base mesh (mesh index 0 inside it)
		auto newLeaf = [&](unsigned bm) {
			const unsigned nv = b->base.vertOff[bm + 1] - b->base.vertOff[bm];
			const unsigned nt = b->base.triOff[bm + 1] - b->base.triOff[bm];
			const unsigned vo[2] = {0, nv}, to[2] = {0, nt};
			orc_bvh *leaf = orc_bvh_create(1, vo, to, &b->base.verts[b->base.vertOff[bm]].x,
					b->base.tri(bm, 0), builder, p.treeType, 0, 80, 10, 0.5f, nThreads);
			if (!leaf) throw std::runtime_error("leaf build failed");
			b->uniqueLeafs.emplace_back(leaf);
			return (unsigned)b->uniqueLeafs.size() - 1;
		};
		for (unsigned o = 0; o < nObjs; ++o) {  // mbvhaccel.cpp:90-180
			const unsigned bm = objMesh[o];
			if (objKind[o] == 0) {
				const unsigned li = newLeaf(bm);
				uniqueByMesh[bm] = li;
				b->leafIndex.push_back(li);
				b->transformIndex.push_back(NULL_INDEX);
				b->motionIndex.push_back(NULL_INDEX);
			} else {
				auto it = uniqueByMesh.find(bm);
				if (it == uniqueByMesh.end()) {
					const unsigned li = newLeaf(bm);
					uniqueByMesh[bm] = li;
					b->leafIndex.push_back(li);
				} else
					b->leafIndex.push_back(it->second);
				if (objKind[o] == 1) {
					M4 m, inv;
					std::memcpy(m.m, objMat + 16 * (size_t)o, sizeof(m.m));
					if (!Inverse(m, &inv)) throw std::runtime_error("singular instance matrix");
					b->transformIndex.push_back((unsigned)b->xfM.size());
					b->xfM.push_back(m);
					b->xfInv.push_back(inv);
					b->motionIndex.push_back(NULL_INDEX);
				} else {
					const unsigned k0 = objKeyOff[o], k1 = objKeyOff[o + 1];
					std::unique_ptr<Motion> ms;
					if (!makeMotion(k1 - k0, keyTimes + k0, keyMats + 16 * (size_t)k0, &ms))
						throw std::runtime_error("bad motion keys");
					b->motionIndex.push_back((unsigned)b->motions.size());
					b->motions.push_back(std::move(ms));
					b->transformIndex.push_back(NULL_INDEX);
				}
			}
		}
		// root leaves: mbvhaccel.cpp:187-202, boxes from Mesh::GetBBox
		std::vector<TreeNode> leafs(nObjs);
		std::vector<TreeNode *> list(nObjs);
		b->leafBox.resize(nObjs);
		for (unsigned o = 0; o < nObjs; ++o) {
			const BBox mb = b->base.meshBBox(objMesh[o]);
			BBox wb;
			if (b->transformIndex[o] != NULL_INDEX)
				wb = xfBBox(b->xfM[b->transformIndex[o]], mb);         // trianglemesh.cpp:241-248
			else if (b->motionIndex[o] != NULL_INDEX)
				wb = b->motions[b->motionIndex[o]]->Bound(mb, true);  // trianglemesh.cpp:268-275
			else
				wb = mb;
			wb.Expand(E(wb));
			b->leafBox[o] = wb;
			TreeNode *n = &leafs[o];
			n->bbox = wb;
			n->leafIndex = b->leafIndex[o];
			n->transformIndex = b->transformIndex[o];
			n->motionIndex = b->motionIndex[o];
			n->meshOffsetIndex = o;
			list[o] = n;
		}
		b->root = BuildTree(builder, p, nullptr, list, nThreads);
		b->nRootNodes = (unsigned)b->root.size();
		return b.release();
	} catch (...) {
		return nullptr;
	}
}

void orc_mbvh_destroy(orc_mbvh *b) { delete b; }

void orc_mbvh_leaf_bbox(const orc_mbvh *b, unsigned o, float pmin[3], float pmax[3]) {
	for (int i = 0; i < 3; ++i) { pmin[i] = b->leafBox[o].pMin[i]; pmax[i] = b->leafBox[o].pMax[i]; }
}

void orc_mbvh_intersect(const orc_mbvh *b, const orc_ray *rays, orc_rayhit *hits,
		size_t n, int nThreads, unsigned long long *counters) {
	std::atomic<unsigned long long> cIn(0), cLeaf(0);
	ParallelChunks(n, nThreads, [&](size_t s, size_t e) {
		unsigned long long c[2] = {0, 0};
		for (size_t i = s; i < e; ++i) {
			if (rays[i].flags & 1u) continue;
			orc_rayhit h;
			if (b->Intersect(toRay(rays[i]), &h, counters ? c : nullptr))
				hits[i] = h;
			else
				writeMiss(rays[i], &hits[i]);
		}
		cIn += c[0];
		cLeaf += c[1];
	});
	if (counters) { counters[0] = cIn.load(); counters[1] = cLeaf.load(); }
}

void orc_mbvh_bruteforce(const orc_mbvh *b, const orc_ray *rays, orc_rayhit *hits,
		size_t n, int nThreads) {
	const unsigned nObjs = (unsigned)b->objMesh.size();
	ParallelChunks(n, nThreads, [&](size_t s, size_t e) {
		for (size_t i = s; i < e; ++i) {
			if (rays[i].flags & 1u) continue;
			const Ray ray = toRay(rays[i]);
			orc_rayhit h;
			h.t = ray.maxt; h.b1 = 0.f; h.b2 = 0.f;
			h.meshIndex = NULL_INDEX; h.triangleIndex = NULL_INDEX;
			for (unsigned o = 0; o < nObjs; ++o) {
				Ray lr = b->leafRay(o, ray);
				lr.maxt = h.t;
				const unsigned bm = b->objMesh[o];
				const unsigned cnt = b->base.triOff[bm + 1] - b->base.triOff[bm];
				float t, b1, b2;
				for (unsigned k = 0; k < cnt; ++k) {
					const unsigned *tv = b->base.tri(bm, k);
					if (TriangleIntersect(lr, b->base.vertex(bm, tv[0]), b->base.vertex(bm, tv[1]),
							b->base.vertex(bm, tv[2]), &t, &b1, &b2) && t < h.t) {
						lr.maxt = t;
						h.t = t; h.b1 = b1; h.b2 = b2; h.meshIndex = o; h.triangleIndex = k;
					}
				}
			}
			hits[i] = h;
		}
	});
}

}  // extern "C"
