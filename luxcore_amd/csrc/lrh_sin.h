/*
 * sin(x) in double for the slerp of a motion segment, x in [0, pi] (a float
 * argument; pi rounded up to float, 3.14159274, is inside).  Plain IEEE double
 * multiplies and adds (the files that include it build with
 * -ffp-contract=off), so the GPU and a host compiler produce the same bits.
 *
 * Method: reduction to [-pi/4, pi/4] by Sterbenz-exact subtraction of the
 * leading double of pi/2 or pi (x/2 <= c <= 2x holds on each branch), the
 * trailing double carried as a tail; Taylor polynomials of sin (to x^17) and
 * cos (to x^18), whose truncation errors on [0, pi/4] are below 1e-19
 * relative, evaluated around the exact leading term with the tail corrections
 * of the classic kernel-sin / kernel-cos split (K. C. Ng, "Argument reduction
 * for huge arguments: good to the last bit", 1992, and the freely
 * distributed fdlibm kernels it describes).  The result rounded to float
 * equals the float rounding of the C library's sin for every float argument
 * in the domain (checked exhaustively against glibc by the test suite).
 */
#ifndef LRH_SIN_H
#define LRH_SIN_H

#if defined(__HIPCC__)
#define LRH_HD __host__ __device__
#else
#define LRH_HD
#endif

namespace lrhsin {

// 1/(2k+1)! and 1/(2k)! with alternating signs, rounded to double
constexpr double S3 = -1.0 / 6.0;
constexpr double S5 = 1.0 / 120.0;
constexpr double S7 = -1.0 / 5040.0;
constexpr double S9 = 1.0 / 362880.0;
constexpr double S11 = -1.0 / 39916800.0;
constexpr double S13 = 1.0 / 6227020800.0;
constexpr double S15 = -1.0 / 1307674368000.0;
constexpr double S17 = 1.0 / 355687428096000.0;
constexpr double C4 = 1.0 / 24.0;
constexpr double C6 = -1.0 / 720.0;
constexpr double C8 = 1.0 / 40320.0;
constexpr double C10 = -1.0 / 3628800.0;
constexpr double C12 = 1.0 / 479001600.0;
constexpr double C14 = -1.0 / 87178291200.0;
constexpr double C16 = 1.0 / 20922789888000.0;
constexpr double C18 = -1.0 / 6402373705728000.0;
// pi/2 and pi as a leading double and the remainder
constexpr double PIO2_HI = 1.5707963267948966;      // 0x3FF921FB54442D18
constexpr double PIO2_LO = 6.123233995736766e-17;   // pi/2 - PIO2_HI
constexpr double PI_HI = 3.141592653589793;         // 0x400921FB54442D18
constexpr double PI_LO = 1.2246467991473532e-16;    // pi - PI_HI
constexpr double PIO4 = 0.7853981633974483;
constexpr double PI3O4 = 2.356194490192345;

// sin(a + t) for |a| <= pi/4 and |t| tiny against ulp(a)
LRH_HD inline double kernelSin(double a, double t) {
	const double z = a * a;
	const double v = z * a;
	const double r = S5 + z * (S7 + z * (S9 + z * (S11 + z * (S13 + z * (S15 + z * S17)))));
	// a + t + v*S3 + v*z*r, the small terms summed before the leading one
	return a - (((z * (0.5 * t - v * r) - t) - v * S3));
}

// cos(a + t) for |a| <= pi/4 and |t| tiny against ulp(a)
LRH_HD inline double kernelCos(double a, double t) {
	const double z = a * a;
	const double r = z * z * (C4 + z * (C6 + z * (C8 + z * (C10 + z * (C12 + z * (C14 + z * (C16 + z * C18)))))));
	const double hz = 0.5 * z;
	const double w = 1.0 - hz;
	// 1 - z/2 + r - a*t, with the rounding error of 1 - z/2 recovered
	return w + (((1.0 - w) - hz) + (r - a * t));
}

}  // namespace lrhsin

// sin(x) for x in [0, 3.1415927] (see above)
LRH_HD inline double lrhSin0Pi(double x) {
	using namespace lrhsin;
	if (x <= PIO4) return kernelSin(x, 0.0);
	if (x <= PI3O4) {
		// sin(x) = cos(x - pi/2); x - PIO2_HI is exact (Sterbenz)
		const double r = x - PIO2_HI;
		const double y = r - PIO2_LO;
		const double t = (r - y) - PIO2_LO;
		return kernelCos(y, t);
	}
	// sin(x) = sin(pi - x); PI_HI - x is exact (Sterbenz)
	const double r = PI_HI - x;
	const double y = r + PI_LO;
	const double t = PI_LO - (y - r);
	return kernelSin(y, t);
}

#endif
