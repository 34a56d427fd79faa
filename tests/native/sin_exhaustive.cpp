/*
 * Every float x in [0, 3.14159274] (pi rounded up to float, the largest angle
 * a motion segment's slerp evaluates): (float)lrhSin0Pi(x), the device's sine,
 * must equal (float)sin((double)x) of the C library, which the oracle uses.
 * Prints the number of differing arguments (and the first few) and exits 1
 * when there is any.  Build: g++ -O2 -ffp-contract=off -pthread.
 */
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include "../../luxcore_amd/csrc/lrh_sin.h"

int main(int argc, char **argv) {
	const uint32_t last = 0x40490FDBu;  // 3.14159274f
	unsigned nThreads = argc > 1 ? (unsigned)atoi(argv[1]) : std::thread::hardware_concurrency();
	if (nThreads == 0) nThreads = 1;
	std::atomic<unsigned long long> bad{0};
	std::vector<uint32_t> first[64];
	std::vector<std::thread> th;
	for (unsigned t = 0; t < nThreads && t < 64; ++t)
		th.emplace_back([&, t]() {
			const uint64_t n = (uint64_t)last + 1;
			const uint64_t b = n * t / nThreads, e = n * (t + 1) / nThreads;
			unsigned long long local = 0;
			for (uint64_t u = b; u < e; ++u) {
				float x;
				const uint32_t bits = (uint32_t)u;
				std::memcpy(&x, &bits, 4);
				const float a = (float)lrhSin0Pi((double)x);
				const float c = (float)std::sin((double)x);
				if (std::memcmp(&a, &c, 4) != 0) {
					++local;
					if (first[t].size() < 4) first[t].push_back(bits);
				}
			}
			bad += local;
		});
	for (auto &x : th) x.join();
	printf("arguments %llu differing %llu\n", (unsigned long long)last + 1, bad.load());
	for (unsigned t = 0; t < nThreads && t < 64; ++t)
		for (uint32_t bits : first[t]) {
			float x;
			std::memcpy(&x, &bits, 4);
			printf("  x=%.9g (0x%08x): lrhSin0Pi %.17g libc %.17g\n", x, bits, lrhSin0Pi((double)x), std::sin((double)x));
		}
	return bad.load() ? 1 : 0;
}
