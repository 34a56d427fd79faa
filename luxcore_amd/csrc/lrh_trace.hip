/*
 * gfx950 traversal kernels of the HIP intersection device.
 *
 * One lane = one ray, wave64, persistent grid.  Per wave, a ballot over idle
 * lanes takes a block of ray tickets with one atomic and hands them out by
 * lane rank, so lanes whose ray terminated are refilled while the others keep
 * traversing (the reference launches one thread per ray with no refill,
 * include/luxrays/accelerators/bvh.cl:228-260).
 *
 * Traversal is ordered (children sorted by entry distance) over 4-wide or
 * 8-wide nodes, with a per-lane stack whose top `ldsDepth` entries live in LDS and
 * deeper entries spill to a global scratch array; popped entries whose entry
 * distance is not below the current hit are dropped without a fetch.  The
 * reference walks a skip-pointer array in memory order instead
 * (bvh.cl:136-217, src/luxrays/accelerators/bvhaccel.cpp:184-257); both visit
 * a superset of the boxes containing the closest hit, so the closest hit is
 * the same up to ties at equal t.
 *
 * The triangle test is Triangle::Intersect
 * (include/luxrays/core/geometry/triangle.h:55-89) with every operation
 * rounded as written (the file builds with -ffp-contract=off and correctly
 * rounded division) and the acceptance rule of BVHAccel::Intersect
 * (bvhaccel.cpp:173, 232-233): a hit is kept iff mint <= t < best, best
 * starting at maxt.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lrh_layout.h"
#include "lrh_kernels.h"
#include "lrh_motion_dev.h"

#define LRH_SENTINEL 0xfffffffeu
#define DEVI __device__ __forceinline__

namespace {

DEVI float asf(uint32_t u) { return __uint_as_float(u); }
DEVI uint32_t asu(float f) { return __float_as_uint(f); }

DEVI float safeRcp(float d) {
	// a zero component becomes +-1e-20 so slab products stay finite
	const float dd = fabsf(d) < 1e-20f ? copysignf(1e-20f, d) : d;
	return 1.0f / dd;
}

DEVI float byteF(uint32_t w, int c) {
	// compiles to v_cvt_f32_ubyteN
	return (float)((w >> (8 * c)) & 0xffu);
}

struct LocalRay {
	float ox, oy, oz, dx, dy, dz;
	float ix, iy, iz;     // 1/d
	float nx, ny, nz;     // -o/d
};

DEVI void setupRay(LocalRay &r) {
	r.ix = safeRcp(r.dx);
	r.iy = safeRcp(r.dy);
	r.iz = safeRcp(r.dz);
	r.nx = -r.ox * r.ix;
	r.ny = -r.oy * r.iy;
	r.nz = -r.oz * r.iz;
}

// Matrix * Point with the projective divide and Matrix * Vector
// (include/luxrays/core/geometry/matrix4x4op.h:31-59), rounded as written.
DEVI void xfRay(const float *M, const float wo[3], const float wd[3], LocalRay &r) {
	const float x = wo[0], y = wo[1], z = wo[2];
	float px = M[0] * x + M[1] * y + M[2] * z + M[3];
	float py = M[4] * x + M[5] * y + M[6] * z + M[7];
	float pz = M[8] * x + M[9] * y + M[10] * z + M[11];
	const float w = M[12] * x + M[13] * y + M[14] * z + M[15];
	if (w != 1.f) {
		const float inv = 1.f / w;
		px = inv * px; py = inv * py; pz = inv * pz;
	}
	r.ox = px; r.oy = py; r.oz = pz;
	const float a = wd[0], b = wd[1], c = wd[2];
	r.dx = M[0] * a + M[1] * b + M[2] * c;
	r.dy = M[4] * a + M[5] * b + M[6] * c;
	r.dz = M[8] * a + M[9] * b + M[10] * c;
	setupRay(r);
}

// ---- motion sampling: MotionSystem::Sample picks the segment
// (src/luxrays/core/geometry/motionsystem.cpp:340-346), lrhSegSample
// evaluates it (InterpolatedTransform::Sample, :89-159) ----
__device__ __forceinline__ void motionSample(const lrh_motion *mot, const lrh_motion_seg *segs, const float *times,
		float time, float M[16]) {
	// upper_bound(times, time), clamped to the last key; the record's
	// inline key times serve systems of up to LRH_MOTION_INLINE_TIMES keys
	const uint4 a = reinterpret_cast<const uint4 *>(mot)[0];
	const uint4 b = reinterpret_cast<const uint4 *>(mot)[1];
	// both halves of the record in one wait
	uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
	for (int k = 0; k < 8; ++k) asm volatile("" : "+v"(w[k]));
	const uint32_t n = w[1];
	uint32_t idx = 0;
	if (n <= LRH_MOTION_INLINE_TIMES) {
		const float tk[4] = {__uint_as_float(w[4]), __uint_as_float(w[5]), __uint_as_float(w[6]),
				__uint_as_float(w[7])};
#pragma unroll
		for (uint32_t k = 0; k < 4; ++k)
			if (idx == k && k < n && !(time < tk[k])) idx = k + 1;
	} else {
		const float *tk = times + w[0];
		while (idx < n && !(time < tk[idx])) ++idx;
	}
	if (idx > n - 1) idx = n - 1;
	lrhSegSample(segs[w[2] + idx], time, M);
}

// ---- node tests: up to 4 children -> (entry distance, child ref) ----
// conservative widening of the slab interval by 2^-20 of |t| on each side
// (entry t - |t| eps, exit t + |t| eps: wider for either sign of t, which a
// plain scale by 1 -/+ eps is not for negative t, i.e. negative mint)
constexpr float kWiden = 1.0f / 1048576.0f;
DEVI float widenLo(float t) { return fmaf(-fabsf(t), kWiden, t); }
DEVI float widenHi(float t) { return fmaf(fabsf(t), kWiden, t); }

DEVI void testNodeQ(const lrh_node4q *__restrict__ nodes, uint32_t idx, const LocalRay &r, float mint, float best,
		float t[4], uint32_t c[4]) {
	const uint4 *p = reinterpret_cast<const uint4 *>(nodes + idx);
	const uint4 q0 = p[0], q1 = p[1], q2 = p[2], q3 = p[3];
	const float sx = asf((q0.w & 0xffu) << 23), sy = asf(((q0.w >> 8) & 0xffu) << 23),
			sz = asf(((q0.w >> 16) & 0xffu) << 23);
	const float Ax = sx * r.ix, Ay = sy * r.iy, Az = sz * r.iz;
	const float Bx = fmaf(asf(q0.x), r.ix, r.nx), By = fmaf(asf(q0.y), r.iy, r.ny), Bz = fmaf(asf(q0.z), r.iz, r.nz);
	const uint32_t nX = r.ix >= 0.f ? q1.x : q1.y, fX = r.ix >= 0.f ? q1.y : q1.x;
	const uint32_t nY = r.iy >= 0.f ? q1.z : q1.w, fY = r.iy >= 0.f ? q1.w : q1.z;
	const uint32_t nZ = r.iz >= 0.f ? q2.x : q2.y, fZ = r.iz >= 0.f ? q2.y : q2.x;
	c[0] = q2.z; c[1] = q2.w; c[2] = q3.x; c[3] = q3.y;
#pragma unroll
	for (int k = 0; k < 4; ++k) {
		const float tnx = fmaf(byteF(nX, k), Ax, Bx), tfx = fmaf(byteF(fX, k), Ax, Bx);
		const float tny = fmaf(byteF(nY, k), Ay, By), tfy = fmaf(byteF(fY, k), Ay, By);
		const float tnz = fmaf(byteF(nZ, k), Az, Bz), tfz = fmaf(byteF(fZ, k), Az, Bz);
		const float tn = widenLo(fmaxf(fmaxf(tnx, tny), tnz));
		const float tf = widenHi(fminf(fminf(tfx, tfy), tfz));
		const float a = fmaxf(tn, mint), b = fminf(tf, best);
		// empty slots hold inverted boxes; one that still passes (a far node
		// whose inverted extent is below the widening) is dropped when its
		// LRH_CHILD_EMPTY reference is taken or popped
		t[k] = (a <= b) ? a : INFINITY;
	}
}

// 8-wide quantized node: six 16-byte loads (origin and scales, the three
// axes' bytes, eight child references)
DEVI void testNode8Q(const lrh_node8q *__restrict__ nodes, uint32_t idx, const LocalRay &r, float mint, float best,
		float t[8], uint32_t c[8]) {
	const uint4 *p = reinterpret_cast<const uint4 *>(nodes + idx);
	const uint4 q0 = p[0], q1 = p[1], q2 = p[2], q3 = p[3], q4 = p[4], q5 = p[5];
	const float sx = asf((q0.w & 0xffu) << 23), sy = asf(((q0.w >> 8) & 0xffu) << 23),
			sz = asf(((q0.w >> 16) & 0xffu) << 23);
	const float Ax = sx * r.ix, Ay = sy * r.iy, Az = sz * r.iz;
	const float Bx = fmaf(asf(q0.x), r.ix, r.nx), By = fmaf(asf(q0.y), r.iy, r.ny), Bz = fmaf(asf(q0.z), r.iz, r.nz);
	const bool px = r.ix >= 0.f, py = r.iy >= 0.f, pz = r.iz >= 0.f;
	const uint32_t nX[2] = {px ? q1.x : q1.z, px ? q1.y : q1.w}, fX[2] = {px ? q1.z : q1.x, px ? q1.w : q1.y};
	const uint32_t nY[2] = {py ? q2.x : q2.z, py ? q2.y : q2.w}, fY[2] = {py ? q2.z : q2.x, py ? q2.w : q2.y};
	const uint32_t nZ[2] = {pz ? q3.x : q3.z, pz ? q3.y : q3.w}, fZ[2] = {pz ? q3.z : q3.x, pz ? q3.w : q3.y};
	c[0] = q4.x; c[1] = q4.y; c[2] = q4.z; c[3] = q4.w;
	c[4] = q5.x; c[5] = q5.y; c[6] = q5.z; c[7] = q5.w;
#pragma unroll
	for (int k = 0; k < 8; ++k) {
		const int w = k >> 2, b = k & 3;
		const float tnx = fmaf(byteF(nX[w], b), Ax, Bx), tfx = fmaf(byteF(fX[w], b), Ax, Bx);
		const float tny = fmaf(byteF(nY[w], b), Ay, By), tfy = fmaf(byteF(fY[w], b), Ay, By);
		const float tnz = fmaf(byteF(nZ[w], b), Az, Bz), tfz = fmaf(byteF(fZ[w], b), Az, Bz);
		const float tn = widenLo(fmaxf(fmaxf(tnx, tny), tnz));
		const float tf = widenHi(fminf(fminf(tfx, tfy), tfz));
		const float a = fmaxf(tn, mint), bb = fminf(tf, best);
		t[k] = (a <= bb) ? a : INFINITY;
	}
}

DEVI void testNodeF(const lrh_node4f *__restrict__ nodes, uint32_t idx, const LocalRay &r, float mint, float best,
		float t[4], uint32_t c[4]) {
	const float4 *p = reinterpret_cast<const float4 *>(nodes + idx);
	const float4 lx = p[0], hx = p[1], ly = p[2], hy = p[3], lz = p[4], hz = p[5];
	const uint4 cc = reinterpret_cast<const uint4 *>(p)[6];
	c[0] = cc.x; c[1] = cc.y; c[2] = cc.z; c[3] = cc.w;
	const float LX[4] = {lx.x, lx.y, lx.z, lx.w}, HX[4] = {hx.x, hx.y, hx.z, hx.w};
	const float LY[4] = {ly.x, ly.y, ly.z, ly.w}, HY[4] = {hy.x, hy.y, hy.z, hy.w};
	const float LZ[4] = {lz.x, lz.y, lz.z, lz.w}, HZ[4] = {hz.x, hz.y, hz.z, hz.w};
	const bool sx = r.ix >= 0.f, sy = r.iy >= 0.f, sz = r.iz >= 0.f;
#pragma unroll
	for (int k = 0; k < 4; ++k) {
		const float tnx = fmaf(sx ? LX[k] : HX[k], r.ix, r.nx), tfx = fmaf(sx ? HX[k] : LX[k], r.ix, r.nx);
		const float tny = fmaf(sy ? LY[k] : HY[k], r.iy, r.ny), tfy = fmaf(sy ? HY[k] : LY[k], r.iy, r.ny);
		const float tnz = fmaf(sz ? LZ[k] : HZ[k], r.iz, r.nz), tfz = fmaf(sz ? HZ[k] : LZ[k], r.iz, r.nz);
		const float tn = widenLo(fmaxf(fmaxf(tnx, tny), tnz));
		const float tf = widenHi(fminf(fminf(tfx, tfy), tfz));
		const float a = fmaxf(tn, mint), b = fminf(tf, best);
		t[k] = (a <= b) ? a : INFINITY;  // empty slots: lo = +inf, hi = -inf
	}
}

DEVI void cswap(float &ta, uint32_t &ca, float &tb, uint32_t &cb) {
	const bool s = tb < ta;
	const float t0 = s ? tb : ta, t1 = s ? ta : tb;
	const uint32_t c0 = s ? cb : ca, c1 = s ? ca : cb;
	ta = t0; tb = t1; ca = c0; cb = c1;
}

// ---- stack: top ldsDepth entries in LDS (ring), the rest spilled ----
struct Stack {
	uint2 *lds;
	uint2 *spill;
	uint32_t stride;    // blockDim (LDS) / nLanes (spill)
	uint32_t lane;      // threadIdx / global lane
	uint32_t gstride;
	uint32_t glane;
	uint32_t mask;      // ldsDepth - 1
	uint32_t depth;     // ldsDepth
	uint32_t spillDepth;
	uint32_t sp;
};

DEVI bool push(Stack &s, uint32_t ref, float t, uint32_t *err) {
	uint2 *e = &s.lds[(s.sp & s.mask) * s.stride + s.lane];
	if (s.sp >= s.depth) {
		const uint32_t k = s.sp - s.depth;
		if (k >= s.spillDepth) {
			atomicOr(err, 1u);
			return false;
		}
		s.spill[(size_t)k * s.gstride + s.glane] = *e;
	}
	*e = make_uint2(ref, asu(t));
	++s.sp;
	return true;
}

DEVI uint2 pop(Stack &s) {
	--s.sp;
	uint2 *e = &s.lds[(s.sp & s.mask) * s.stride + s.lane];
	const uint2 v = *e;
	if (s.sp >= s.depth) *e = s.spill[(size_t)(s.sp - s.depth) * s.gstride + s.glane];
	return v;
}

constexpr uint32_t kNoRay = 0xffffffffu;
constexpr uint32_t kMaxSteps = 1u << 24;  // per-ray guard against corrupt trees

}  // namespace

// Per-lane view of the scene and buffers a ray belongs to.  In a plain trace
// every field comes from the launch parameters (uniform, kept in SGPRs); in a
// batched trace each lane loads its item's descriptor when it takes a ray.
struct SceneView {
	const void *nodes;
	const lrh_tri *tris;
	const lrh_inst *insts;
	const float *xforms;
	const lrh_motion *motions;
	const lrh_motion_seg *segs;
	const float *times;
	lrh_hit_rec *hits;
	uint32_t root;
	uint32_t flags;  // LRH_ITEM_ANY | LRH_ITEM_TWO
};

template <int LAYOUT, bool TWO, bool ANY, bool COUNT, bool BATCH>
__device__ __forceinline__ void traceBody(const LrhTraceParams &P) {
	extern __shared__ uint2 ldsStack[];  // [ldsDepth][blockDim] stack; batched: then item starts
	const uint32_t lane = __lane_id();
	Stack st;
	st.lds = ldsStack;
	st.spill = P.spill;
	st.stride = blockDim.x;
	st.lane = threadIdx.x;
	st.gstride = P.nLanes;
	st.glane = blockIdx.x * blockDim.x + threadIdx.x;
	st.mask = P.ldsDepth - 1;
	st.depth = P.ldsDepth;
	st.spillDepth = P.spillDepth;
	st.sp = 0;

	uint32_t *itemStart = nullptr;
	if (BATCH) {
		itemStart = reinterpret_cast<uint32_t *>(ldsStack + (size_t)P.ldsDepth * blockDim.x);
		for (uint32_t i = threadIdx.x; i <= P.nItems; i += blockDim.x) itemStart[i] = P.itemStart[i];
		__syncthreads();
	}

	SceneView v;
	v.nodes = P.nodes;
	v.tris = P.tris;
	v.insts = P.insts;
	v.xforms = P.xforms;
	v.motions = P.motions;
	v.segs = P.segs;
	v.times = P.times;
	v.hits = P.hits;
	v.root = P.root;
	v.flags = (ANY ? LRH_ITEM_ANY : 0u) | (TWO ? LRH_ITEM_TWO : 0u);

	uint32_t rid = kNoRay;  // ray index within its buffer
	LocalRay r;
	float wo[3], wd[3];
	float mint = 0.f, best = 0.f, time = 0.f, hb1 = 0.f, hb2 = 0.f;
	uint32_t hTri = LRH_NULL_INDEX, hMesh = LRH_NULL_INDEX, cur = LRH_CHILD_EMPTY, obj = 0, steps = 0;
	float curT = 0.f;
	bool inInst = false;
	bool exhausted = false;
	// COUNT variant only: per lane node tests / primitive tests / rays; per
	// wave passes, live lanes summed over passes, passes with a node test,
	// passes with a leaf test, refills
	uint32_t cNode = 0, cPrim = 0, cRay = 0, cPass = 0, cLive = 0, cNodePass = 0, cLeafPass = 0, cRefill = 0;
	uint32_t cObj = 0, cXf = 0, cMot = 0;  // per lane: object entries, through a matrix, through a motion system


	for (;;) {
		// the ray record a lane takes this pass
		uint4 ldA = make_uint4(0u, 0u, 0u, 0u), ldB = make_uint4(0u, 0u, 0u, 0u);
		uint32_t ldTime = 0u, ldFlags = 0u;
		bool loading = false;
		// ---- refill: idle lanes take tickets when fewer than `refill` lanes are live ----
		if (!exhausted) {
			const uint64_t idle = __ballot(rid == kNoRay);
			const uint32_t nIdle = (uint32_t)__popcll(idle);
			if (64u - nIdle < P.refill) {
				if (COUNT) ++cRefill;
				// one atomic per refill for exactly the idle lanes: tickets are
				// taken on demand so the grid drains evenly at the end
				const int leader = __ffsll((long long)idle) - 1;
				uint32_t base = 0;
				if ((int)lane == leader) base = atomicAdd(P.counter, nIdle);
				base = __shfl(base, leader);
				if (base + nIdle >= P.nRays) exhausted = true;
				const uint32_t ticket = base + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
				if (rid == kNoRay && ticket < P.nRays) {
					uint32_t local = ticket;
					const lrh_ray_rec *rays = P.rays;
					if (BATCH) {
						// item of this ticket: the last item whose start is <= ticket
						uint32_t lo = 0, hi = P.nItems;
						while (hi - lo > 1) {
							const uint32_t mid = (lo + hi) >> 1;
							if (itemStart[mid] <= ticket) lo = mid; else hi = mid;
						}
						const LrhBatchDesc &d = P.items[lo];
						local = ticket - itemStart[lo];
						rays = d.rays;
						v.nodes = d.nodes;
						v.tris = d.tris;
						v.insts = d.insts;
						v.xforms = d.xforms;
						v.motions = d.motions;
						v.segs = d.segs;
						v.times = d.times;
						v.hits = d.hits;
						v.root = d.root;
						v.flags = d.flags & (TWO ? (LRH_ITEM_ANY | LRH_ITEM_TWO) : LRH_ITEM_ANY);
					}
					const uint4 *rp = reinterpret_cast<const uint4 *>(rays + local);
					rid = local;
					ldA = rp[0];
					ldB = rp[1];
					const uint2 c2 = reinterpret_cast<const uint2 *>(rp + 2)[0];
					ldTime = c2.x;
					ldFlags = c2.y;
					loading = true;
					cur = LRH_CHILD_EMPTY;
					st.sp = 0;
				}
			}
		}
		// ---- set up the rays taken this pass ----
		if (loading) {
			loading = false;
			if (ldFlags & 1u) {  // RAY_FLAGS_MASKED: skipped, hit record untouched
				rid = kNoRay;
			} else {
				r.ox = asf(ldA.x); r.oy = asf(ldA.y); r.oz = asf(ldA.z);
				r.dx = asf(ldA.w); r.dy = asf(ldB.x); r.dz = asf(ldB.y);
				if (TWO) {
					wo[0] = r.ox; wo[1] = r.oy; wo[2] = r.oz;
					wd[0] = r.dx; wd[1] = r.dy; wd[2] = r.dz;
				}
				mint = asf(ldB.z);
				best = asf(ldB.w);
				time = asf(ldTime);
				setupRay(r);
				hb1 = 0.f; hb2 = 0.f;
				hTri = LRH_NULL_INDEX; hMesh = LRH_NULL_INDEX;
				// a NaN origin or direction component makes every triangle test
				// fail (the divisor or the distance is NaN), while every box
				// would be entered: the ray misses without a traversal
				const bool nanRay = isnan(r.ox) || isnan(r.oy) || isnan(r.oz) || isnan(r.dx) || isnan(r.dy) ||
						isnan(r.dz);
				cur = nanRay ? LRH_CHILD_EMPTY : v.root;
				curT = -INFINITY;
				inInst = false;
				steps = 0;
				if (COUNT) ++cRay;
			}
		}
		if (!__ballot(rid != kNoRay)) {
			if (exhausted) break;
			continue;
		}
		// passes until fewer than `refill` lanes are live (after the last
		// ticket: until the wave is empty)
		const uint32_t minLive = exhausted ? 1u : P.refill;
		do {
		if (COUNT) {
			++cPass;
			cLive += (uint32_t)__popcll(__ballot(rid != kNoRay));
		}

		// ---- node phase: one node test per lane and pass ----
		bool done = false, tested = false;
		if (rid != kNoRay) {
			for (;;) {
				if (cur == LRH_CHILD_EMPTY) {
					if (st.sp == 0) break;
					const uint2 e = pop(st);
					cur = e.x;
					curT = asf(e.y);
					if (TWO && cur == LRH_SENTINEL) {
						// leaving an object's tree: back to the world-space ray
						r.ox = wo[0]; r.oy = wo[1]; r.oz = wo[2];
						r.dx = wd[0]; r.dy = wd[1]; r.dz = wd[2];
						setupRay(r);
						inInst = false;
						cur = LRH_CHILD_EMPTY;
					} else if (!(curT < best)) {
						cur = LRH_CHILD_EMPTY;  // entry not before the current hit
					}
					if (cur == LRH_CHILD_EMPTY) continue;
				}
				if (cur & LRH_CHILD_LEAF) break;
				if (++steps > kMaxSteps) {
					atomicOr(P.error, 2u);
					done = true;
					break;
				}
				if (COUNT) {
					++cNode;
					tested = true;
				}
				if (LAYOUT == 2) {
					float t8[8];
					uint32_t c8[8];
					testNode8Q(reinterpret_cast<const lrh_node8q *>(v.nodes), cur, r, mint, best, t8, c8);
					// 19-comparator sorting network for eight entries
					cswap(t8[0], c8[0], t8[2], c8[2]); cswap(t8[1], c8[1], t8[3], c8[3]);
					cswap(t8[4], c8[4], t8[6], c8[6]); cswap(t8[5], c8[5], t8[7], c8[7]);
					cswap(t8[0], c8[0], t8[4], c8[4]); cswap(t8[1], c8[1], t8[5], c8[5]);
					cswap(t8[2], c8[2], t8[6], c8[6]); cswap(t8[3], c8[3], t8[7], c8[7]);
					cswap(t8[0], c8[0], t8[1], c8[1]); cswap(t8[2], c8[2], t8[3], c8[3]);
					cswap(t8[4], c8[4], t8[5], c8[5]); cswap(t8[6], c8[6], t8[7], c8[7]);
					cswap(t8[2], c8[2], t8[4], c8[4]); cswap(t8[3], c8[3], t8[5], c8[5]);
					cswap(t8[1], c8[1], t8[4], c8[4]); cswap(t8[3], c8[3], t8[6], c8[6]);
					cswap(t8[1], c8[1], t8[2], c8[2]); cswap(t8[3], c8[3], t8[4], c8[4]);
					cswap(t8[5], c8[5], t8[6], c8[6]);
					bool ok = true;
					// sorted: the lanes that push entry k also push k-1..1; a
					// position no lane of the wave reaches is skipped as a whole
#pragma unroll
					for (int k = 7; k >= 1; --k)
						if (__ballot(t8[k] != INFINITY))
							if (t8[k] != INFINITY) ok = ok && push(st, c8[k], t8[k], P.error);
					if (t8[0] != INFINITY) {
						cur = c8[0];
						curT = t8[0];
					} else
						cur = LRH_CHILD_EMPTY;
					if (!ok) done = true;
					break;
				}
				float t4[4];
				uint32_t c4[4];
				if (LAYOUT == 0)
					testNodeQ(reinterpret_cast<const lrh_node4q *>(v.nodes), cur, r, mint, best, t4, c4);
				else
					testNodeF(reinterpret_cast<const lrh_node4f *>(v.nodes), cur, r, mint, best, t4, c4);
				cswap(t4[0], c4[0], t4[1], c4[1]);
				cswap(t4[2], c4[2], t4[3], c4[3]);
				cswap(t4[0], c4[0], t4[2], c4[2]);
				cswap(t4[1], c4[1], t4[3], c4[3]);
				cswap(t4[1], c4[1], t4[2], c4[2]);
				bool ok = true;
				if (t4[3] != INFINITY) ok = ok && push(st, c4[3], t4[3], P.error);
				if (t4[2] != INFINITY) ok = ok && push(st, c4[2], t4[2], P.error);
				if (t4[1] != INFINITY) ok = ok && push(st, c4[1], t4[1], P.error);
				if (t4[0] != INFINITY) {
					cur = c4[0];
					curT = t4[0];
				} else
					cur = LRH_CHILD_EMPTY;
				if (!ok) done = true;
				break;
			}
		}

		// ---- leaf phase ----
		const bool atLeaf = rid != kNoRay && !done && cur != LRH_CHILD_EMPTY &&
				(cur & LRH_CHILD_LEAF);
		if (COUNT) {
			cNodePass += __ballot(tested) ? 1u : 0u;
			cLeafPass += __ballot(atLeaf) ? 1u : 0u;
		}
		if (atLeaf) {
			const bool two = TWO && (v.flags & LRH_ITEM_TWO);
			const bool anyq = (BATCH || ANY) && (v.flags & LRH_ITEM_ANY);
			const uint32_t start = cur & LRH_LEAF_START_MASK;
			uint32_t cnt = ((cur >> LRH_LEAF_START_BITS) & 7u) + 1u;
			cur = LRH_CHILD_EMPTY;
			if (!two && cnt > P.leafTris) {
				// test leafTris triangles now and keep the rest of the leaf for
				// the next pass: a wave then never loops over the longest leaf
				// of its lanes while the others wait
				cur = LRH_CHILD_LEAF | ((cnt - P.leafTris - 1u) << LRH_LEAF_START_BITS) | (start + P.leafTris);
				cnt = P.leafTris;
			}
			if (COUNT) cPrim += cnt;
			if (two && !inInst) {
				// root leaf: enter the object's tree in its local space
				const lrh_inst in = v.insts[start];
				if (COUNT) {
					++cObj;
					cXf += in.kind == 1u ? 1u : 0u;
					cMot += in.kind == 2u ? 1u : 0u;
				}
				if (in.kind == 1u) {
					xfRay(v.xforms + 16 * (size_t)in.index, wo, wd, r);
				} else if (in.kind == 2u) {
					float M[16];
					motionSample(v.motions + in.index, v.segs, v.times, time, M);
					xfRay(M, wo, wd, r);
				}
				obj = in.object;
				inInst = true;
				if (push(st, LRH_SENTINEL, -INFINITY, P.error)) {
					cur = in.root;
					curT = -INFINITY;
				} else
					done = true;
			} else {
				for (uint32_t k = 0; k < cnt; ++k) {
					const uint4 *tp = reinterpret_cast<const uint4 *>(v.tris + start + k);
					const uint4 a = tp[0], b = tp[1], c = tp[2];
					// Triangle::Intersect, rounded as written
					const float p0x = asf(a.x), p0y = asf(a.y), p0z = asf(a.z);
					const float e1x = asf(a.w), e1y = asf(b.x), e1z = asf(b.y);
					const float e2x = asf(b.z), e2y = asf(b.w), e2z = asf(c.x);
					const float s1x = (r.dy * e2z) - (r.dz * e2y);
					const float s1y = (r.dz * e2x) - (r.dx * e2z);
					const float s1z = (r.dx * e2y) - (r.dy * e2x);
					const float div = s1x * e1x + s1y * e1y + s1z * e1z;
					const float inv = 1.f / div;
					const float ddx = r.ox - p0x, ddy = r.oy - p0y, ddz = r.oz - p0z;
					const float b1 = (ddx * s1x + ddy * s1y + ddz * s1z) * inv;
					const float s2x = (ddy * e1z) - (ddz * e1y);
					const float s2y = (ddz * e1x) - (ddx * e1z);
					const float s2z = (ddx * e1y) - (ddy * e1x);
					const float b2 = (r.dx * s2x + r.dy * s2y + r.dz * s2z) * inv;
					const float b0 = 1.f - b1 - b2;
					const float t = (e2x * s2x + e2y * s2y + e2z * s2z) * inv;
					const bool ok = (div != 0.f) && !(b1 < 0.f) && !(b2 < 0.f) && !(b0 < 0.f) &&
							!(t < mint) && (t < best);
					if (ok) {
						best = t;
						hb1 = b1;
						hb2 = b2;
						hTri = c.y;
						hMesh = two ? obj : c.z;
						if (anyq) {
							done = true;
							break;
						}
					}
				}
			}
		}
		if (rid != kNoRay && (done || (cur == LRH_CHILD_EMPTY && st.sp == 0))) {
			lrh_hit_rec *h = v.hits + rid;
			h->t = best;
			h->b1 = hb1;
			h->b2 = hb2;
			h->meshIndex = hMesh;
			h->triangleIndex = hTri;
			rid = kNoRay;
			cur = LRH_CHILD_EMPTY;
		}
		} while ((uint32_t)__popcll(__ballot(rid != kNoRay)) >= minLive);
	}
	if (COUNT) {
		unsigned long long a = cNode, b = cPrim, c = cRay, e = cObj, f = cXf, g = cMot;
		for (int o = 32; o > 0; o >>= 1) {
			a += __shfl_xor(a, o);
			b += __shfl_xor(b, o);
			c += __shfl_xor(c, o);
			e += __shfl_xor(e, o);
			f += __shfl_xor(f, o);
			g += __shfl_xor(g, o);
		}
		if (lane == 0) {
			// wave counters are uniform: lane 0 adds them once per wave
			atomicAdd(P.stats + 0, a);
			atomicAdd(P.stats + 1, b);
			atomicAdd(P.stats + 2, c);
			atomicAdd(P.stats + 3, (unsigned long long)cPass);
			atomicAdd(P.stats + 4, (unsigned long long)cLive);
			atomicAdd(P.stats + 5, (unsigned long long)cNodePass);
			atomicAdd(P.stats + 6, (unsigned long long)cLeafPass);
			atomicAdd(P.stats + 7, (unsigned long long)cRefill);
			if (TWO) {
				atomicAdd(P.stats + 8, e);
				atomicAdd(P.stats + 9, f);
				atomicAdd(P.stats + 10, g);
			}
		}
	}
	// the last block to finish rewinds the ticket counter for the next launch
	__syncthreads();
	if (threadIdx.x == 0) {
		__threadfence();
		const uint32_t d = atomicAdd(P.counter + 1, 1u);
		if (d == gridDim.x - 1) {
			atomicExch(P.counter + 0, 0u);
			atomicExch(P.counter + 1, 0u);
		}
	}
}

// Single-level traces are built for 7 waves per SIMD (at most 72 VGPRs, no
// spills), measured 1-3% faster than the 6 the compiler picks on its own.
// Two-level traces over 64-byte nodes are built for 6 waves (80 VGPRs, 8
// spilled): the motion sample is inlined and reads its motion record and
// the hot part of its segment with all loads in flight at once; measured on
// moving instances, 6 waves are 3% faster than 5 (90 VGPRs, no spills);
// the other two-level variants (fp32 nodes, batched, counting) keep the
// compiler's choice, which does not spill.
template <int LAYOUT, bool ANY, bool COUNT, bool BATCH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) void lrh_trace_kernel(LrhTraceParams P) {
	traceBody<LAYOUT, false, ANY, COUNT, BATCH>(P);
}

template <int LAYOUT, bool ANY, bool COUNT, bool BATCH>
__global__ __launch_bounds__(256) void lrh_trace2_kernel(LrhTraceParams P) {
	traceBody<LAYOUT, true, ANY, COUNT, BATCH>(P);
}

// batched traces over 8-wide nodes: at 7 waves per SIMD the per-lane item
// view and the eight-entry sort spill; built for 6
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void lrh_trace8b_kernel(LrhTraceParams P) {
	traceBody<2, false, false, false, true>(P);
}

// two-level traces over 8-wide nodes: the compiler's choice is 105 VGPRs
// (4 waves per SIMD); built for 5 (94 VGPRs, no spills; at 6 it spills 10),
// the batched form (per-lane item view) for 4 (at 5 it spills 37)
template <bool ANY, bool BATCH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(BATCH ? 4 : 5))) void lrh_trace28_kernel(
		LrhTraceParams P) {
	traceBody<2, true, ANY, false, BATCH>(P);
}

template <bool ANY, bool COUNT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void lrh_trace2q_kernel(LrhTraceParams P) {
	traceBody<0, true, ANY, COUNT, false>(P);
}  // (the counting variants keep lrh_trace2_kernel: at 80 VGPRs they would spill)

__global__ void lrh_reset_kernel(uint32_t *p, uint32_t n) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) p[i] = 0u;
}

// ---------------------------------------------------------------------------
// host-side launch table
// ---------------------------------------------------------------------------
typedef void (*KernelFn)(LrhTraceParams);

#define LRH_K(L, T, A, C, B)                                                                  \
	((T) ? ((L) == 0 && !(B) && !(C) ? (const void *)&lrh_trace2q_kernel<A, false>            \
	                         : (const void *)&lrh_trace2_kernel<L, A, C, B>)                  \
	     : (const void *)&lrh_trace_kernel<L, A, C, B>)
static const void *kernelPtr(int layout, bool two, bool any, bool count, bool batch) {
	// 8-wide nodes
	static const void *wide8[2][2] = {{LRH_K(2, false, false, false, false), LRH_K(2, false, false, true, false)},
			{LRH_K(2, false, true, false, false), LRH_K(2, false, true, true, false)}};
	static const void *wide8Two[2][2] = {
			{(const void *)&lrh_trace28_kernel<false, false>, (const void *)&lrh_trace2_kernel<2, false, true, false>},
			{(const void *)&lrh_trace28_kernel<true, false>, (const void *)&lrh_trace2_kernel<2, true, true, false>}};
	if (layout == 2) {
		if (batch) return two ? (const void *)&lrh_trace28_kernel<false, true> : (const void *)&lrh_trace8b_kernel;
		return two ? wide8Two[any ? 1 : 0][count ? 1 : 0] : wide8[any ? 1 : 0][count ? 1 : 0];
	}
	static const void *plain[2][2][2][2] = {
		{{{LRH_K(0, false, false, false, false), LRH_K(0, false, false, true, false)},
		  {LRH_K(0, false, true, false, false), LRH_K(0, false, true, true, false)}},
		 {{LRH_K(0, true, false, false, false), LRH_K(0, true, false, true, false)},
		  {LRH_K(0, true, true, false, false), LRH_K(0, true, true, true, false)}}},
		{{{LRH_K(1, false, false, false, false), LRH_K(1, false, false, true, false)},
		  {LRH_K(1, false, true, false, false), LRH_K(1, false, true, true, false)}},
		 {{LRH_K(1, true, false, false, false), LRH_K(1, true, false, true, false)},
		  {LRH_K(1, true, true, false, false), LRH_K(1, true, true, true, false)}}}};
	// batched: the query is per item, so ANY is unused; COUNT is not offered
	static const void *batched[2][2] = {{LRH_K(0, false, false, false, true), LRH_K(0, true, false, false, true)},
			{LRH_K(1, false, false, false, true), LRH_K(1, true, false, false, true)}};
	if (batch) return batched[layout ? 1 : 0][two ? 1 : 0];
	return plain[layout ? 1 : 0][two ? 1 : 0][any ? 1 : 0][count ? 1 : 0];
}

hipError_t lrhLaunchReset(uint32_t *p, uint32_t n, hipStream_t s) {
	hipLaunchKernelGGL(lrh_reset_kernel, dim3((n + 63) / 64), dim3(64), 0, s, p, n);
	return hipGetLastError();
}

hipError_t lrhOccupancy(int layout, bool two, bool any, bool batch, int block, size_t lds,
		int *blocksPerCU) {
	const void *k = kernelPtr(layout, two, any, false, batch);
	if (!k) return hipErrorInvalidDeviceFunction;
	return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocksPerCU, k, block, lds);
}

hipError_t lrhLaunchTrace(int layout, bool two, bool any, bool count, bool batch, dim3 grid, dim3 block,
		size_t lds, hipStream_t s, const LrhTraceParams &P) {
	void *args[] = {(void *)&P};
	const void *k = kernelPtr(layout, two, any, count, batch);
	if (!k) return hipErrorInvalidDeviceFunction;
	return hipLaunchKernel(k, grid, block, args, lds, s);
}
