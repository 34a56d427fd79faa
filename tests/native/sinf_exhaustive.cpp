/*
 * lrhSinf (luxcore_amd/csrc/lrh_sinf.h), the sine of the device's motion
 * slerp, against the host C library's sinf, which the reference calls
 * (src/luxrays/core/geometry/quaternion.cpp:151-155): every float x with
 * 0 <= x < 120 (the slerp needs [0, pi/2]), and every float in [-pi/2, 0),
 * must give the same float bit for bit.  Prints the number of arguments and
 * of differing ones (and the first few) and exits 1 when any differs.
 * Build: g++ -O2 -ffp-contract=off -pthread.
 */
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "../../luxcore_amd/csrc/lrh_sinf.h"

int main(int argc, char **argv) {
	const uint32_t posEnd = 0x42F00000u;    // 120.0f: [0, 120)
	const uint32_t negBegin = 0x80000001u;  // -denorm_min ..
	const uint32_t negEnd = 0xBFC90FDCu;    // .. -1.57079637f (pi/2 rounded up), inclusive
	const uint64_t nPos = posEnd, nNeg = (uint64_t)negEnd - negBegin + 1;
	unsigned nThreads = argc > 1 ? (unsigned)atoi(argv[1]) : std::thread::hardware_concurrency();
	if (nThreads == 0) nThreads = 1;
	if (nThreads > 64) nThreads = 64;
	std::atomic<unsigned long long> bad{0}, badHalfPi{0};
	std::vector<uint32_t> first[64];
	std::vector<std::thread> th;
	for (unsigned t = 0; t < nThreads; ++t)
		th.emplace_back([&, t]() {
			const uint64_t n = nPos + nNeg;
			const uint64_t b = n * t / nThreads, e = n * (t + 1) / nThreads;
			unsigned long long local = 0, localHalfPi = 0;
			for (uint64_t k = b; k < e; ++k) {
				const uint32_t bits = k < nPos ? (uint32_t)k : (uint32_t)(negBegin + (k - nPos));
				float x;
				std::memcpy(&x, &bits, 4);
				const float a = lrhSinf(x);
				const float c = sinf(x);
				if (std::memcmp(&a, &c, 4) != 0) {
					++local;
					if (std::fabs(x) <= 1.57079637f) ++localHalfPi;
					if (first[t].size() < 4) first[t].push_back(bits);
				}
			}
			bad += local;
			badHalfPi += localHalfPi;
		});
	for (auto &x : th) x.join();
	printf("arguments %llu differing %llu (within pi/2: %llu)\n", (unsigned long long)(nPos + nNeg), bad.load(),
			badHalfPi.load());
	for (unsigned t = 0; t < nThreads; ++t)
		for (uint32_t bits : first[t]) {
			float x;
			std::memcpy(&x, &bits, 4);
			printf("  x=%.9g (0x%08x): lrhSinf %.9g libc %.9g\n", x, bits, lrhSinf(x), sinf(x));
		}
	return bad.load() ? 1 : 0;
}
