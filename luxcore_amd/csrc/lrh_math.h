/*
 * 4x4 matrix helpers shared by the host builder and the scene-file reader.
 * Both must round exactly as the reference's CPU code does (compiled with
 * -ffp-contract=off, no fast-math), so they live in one place.
 */
#ifndef LRH_MATH_H
#define LRH_MATH_H

#include <cmath>
#include <cstring>
#include <utility>

namespace lrh {

// Gauss-Jordan elimination with full pivoting, in the operation order of
// Matrix4x4::Inverse (src/luxrays/core/geometry/matrix4x4.cpp:118-175).
// Returns false for a singular matrix.
inline bool invert4(const float in[4][4], float out[4][4]) {
	int indxc[4], indxr[4], ipiv[4] = {0, 0, 0, 0};
	float a[4][4];
	std::memcpy(a, in, sizeof(a));
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
			for (int k = 0; k < 4; ++k) std::swap(a[irow][k], a[icol][k]);
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
			for (int k = 0; k < 4; ++k) std::swap(a[k][indxr[j]], a[k][indxc[j]]);
	std::memcpy(out, a, sizeof(a));
	return true;
}

// Matrix * Point with the projective divide by w when w != 1
// (include/luxrays/core/geometry/transform.h:131-142, point.h:100-103).
inline void xfPoint4(const float M[4][4], const float p[3], float r[3]) {
	const float x = p[0], y = p[1], z = p[2];
	r[0] = M[0][0] * x + M[0][1] * y + M[0][2] * z + M[0][3];
	r[1] = M[1][0] * x + M[1][1] * y + M[1][2] * z + M[1][3];
	r[2] = M[2][0] * x + M[2][1] * y + M[2][2] * z + M[2][3];
	const float w = M[3][0] * x + M[3][1] * y + M[3][2] * z + M[3][3];
	if (w != 1.f) {
		const float inv = 1.f / w;
		r[0] = inv * r[0]; r[1] = inv * r[1]; r[2] = inv * r[2];
	}
}

}  // namespace lrh

#endif
