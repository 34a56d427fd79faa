/*
 * sinf as the reference's platform computes it, for the per-ray sines of a
 * motion segment's slerp (src/luxrays/core/geometry/quaternion.cpp:151-155
 * calls sinf; its argument there lies in [0, pi/2]).
 *
 * The C library of the supported platform (glibc 2.28 and later, x86-64)
 * implements sinf with the single-precision sine of Arm's Optimized Routines
 * (MIT licensed, github.com/ARM-software/optimized-routines, math/sinf.c and
 * math/sincosf.h): the float argument is widened to double; below 0.75 in
 * magnitude (the top 12 bits of the argument below those of pi/4) a degree-7
 * odd polynomial is evaluated, otherwise the argument is reduced by the
 * nearest multiple n of pi/2 (n from the argument times 2^24 * 2/pi, with
 * a float-to-int conversion and rounding by 0x800000 >> 24) and the sine or
 * cosine polynomial of the remainder is evaluated, signs by quadrant; the
 * double result is rounded once to float.  On x86-64 processors with FMA the
 * library selects its FMA build, in which the compiler contracts each
 * "a + b * c" of the polynomials and the reduction into one fused
 * multiply-add: written out here with fma().  The coefficients are the
 * published minimax coefficients of that routine.
 *
 * Valid for |x| < 120 (the routine's fast-reduction range); the slerp needs
 * [0, pi/2].  Host and device evaluate the same IEEE double operations, so
 * both give the library's float for every argument (exhaustive check against
 * the host's sinf in the test suite).
 */
#ifndef LRH_SINF_H
#define LRH_SINF_H

#include <stdint.h>

#include <cmath>
#include <cstring>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define LRH_SINF_HD __host__ __device__
#else
#define LRH_SINF_HD
#endif

namespace lrhsinf {

// 2^24 * 2/pi and pi/2 (doubles)
constexpr double kHpiInv = 0x1.45F306DC9C883p+23;
constexpr double kHpi = 0x1.921FB54442D18p0;
// sin(x) ~ x + x^3 (s1 + x^2 s2 + x^4 s3); cos(x) ~ c0 + x^2 c1 + x^4 c2 + x^6 (c3 + x^2 c4)
constexpr double kS1 = -0x1.555545995a603p-3;
constexpr double kS2 = 0x1.1107605230bc4p-7;
constexpr double kS3 = -0x1.994eb3774cf24p-13;
constexpr double kC0 = 0x1p0;
constexpr double kC1 = -0x1.ffffffd0c621cp-2;
constexpr double kC2 = 0x1.55553e1068f19p-5;
constexpr double kC3 = -0x1.6c087e89a359dp-10;
constexpr double kC4 = 0x1.99343027bf8c3p-16;

LRH_SINF_HD inline uint32_t bitsOf(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
	return __float_as_uint(f);
#else
	uint32_t u;
	std::memcpy(&u, &f, 4);
	return u;
#endif
}

LRH_SINF_HD inline double fmad(double a, double b, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
	return __fma_rn(a, b, c);
#else
	return std::fma(a, b, c);
#endif
}

// the sine (n even) or cosine (n odd) polynomial; cosSign = -1 negates the
// cosine (quadrants with n & 2)
LRH_SINF_HD inline float poly(double x, double x2, int n, double cosSign) {
	if ((n & 1) == 0) {
		const double x3 = x * x2;
		const double s1 = fmad(x2, kS3, kS2);
		const double x7 = x3 * x2;
		const double s = fmad(x3, kS1, x);
		return (float)fmad(x7, s1, s);
	}
	const double x4 = x2 * x2;
	const double c2 = fmad(x2, cosSign * kC4, cosSign * kC3);
	const double c1 = fmad(x2, cosSign * kC1, cosSign * kC0);
	const double x6 = x4 * x2;
	const double c = fmad(x4, cosSign * kC2, c1);
	return (float)fmad(x6, c2, c);
}

}  // namespace lrhsinf

LRH_SINF_HD inline float lrhSinf(float y) {
	using namespace lrhsinf;
	const uint32_t top = (bitsOf(y) >> 20) & 0x7ffu;
	double x = (double)y;
	if (top < 0x3f4u) {  // below the top bits of pi/4 as a float (|y| < 0.75)
		if (top < 0x398u) return y;  // |y| < 2^-12
		return poly(x, x * x, 0, 1.0);
	}
	// |y| < 120: one multiply-subtract reduction by pi/2
	const double r = x * kHpiInv;
	const int n = ((int32_t)r + 0x800000) >> 24;
	x = fmad(-(double)n, kHpi, x);
	const double sgn = ((n & 3) == 1 || (n & 3) == 2) ? -1.0 : 1.0;
	const double cosSign = (n & 2) ? -1.0 : 1.0;
	return poly(x * sgn, x * x, n, cosSign);
}

#endif
