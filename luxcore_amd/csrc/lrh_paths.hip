/*
 * Ray generation on either side of a trace, for the served path-tracing
 * sessions: pinhole camera rays, and from a wavefront's hits the next
 * cosine-weighted bounce (with Russian roulette) plus one shadow ray toward a
 * 1x1 area light.  They play the part of SLG's AdvancePaths micro-kernels that
 * generate the camera and next-vertex rays
 * (include/slg/engines/pathoclbase/kernels/pathoclbase_kernels_micro.cl:
 * 723-844, 1045-1075), reduced to what a synthetic session needs.
 *
 * The shading step (lrh_shade, and lrh_shade_batch for several wavefronts in
 * one launch) consumes hits as SLG's hit-consumption and direct-light
 * micro-kernels do, reduced to Lambertian surfaces and triangle lights.
 *
 * Terminated paths become masked rays (RAY_FLAGS_MASKED), so a wavefront
 * keeps its size and the tracer skips them; a path that ended before the
 * previous step is not touched again (its slots are masked already).
 * Random numbers are a hash of
 * (session seed, pixel, bounce, dimension): a session's rays do not depend on
 * what else is in flight.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lrh_layout.h"
#include "lrh_paths.h"

namespace {

__device__ __forceinline__ uint32_t pcg(uint32_t v) {
	const uint32_t s = v * 747796405u + 2891336453u;
	const uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
	return (w >> 22u) ^ w;
}

__device__ __forceinline__ float rnd(uint64_t seed, uint32_t pixel, uint32_t bounce, uint32_t dim) {
	const uint32_t h = pcg((uint32_t)seed ^ pcg((uint32_t)(seed >> 32) ^ pcg(pixel ^ pcg(bounce * 16u + dim))));
	return (float)(h >> 8) * (1.0f / 16777216.0f);
}

// MachineEpsilon::E of a point (include/luxrays/core/epsilon.h:50-94)
__device__ __forceinline__ float epsE(float v) {
	const float n = __uint_as_float(__float_as_uint(v) + 0x80u);
	const float e = fabsf(n - v);
	return e > 1e-5f ? (e < 1e-1f ? e : 1e-1f) : 1e-5f;
}

__device__ __forceinline__ void writeMasked(lrh_ray_rec *r) {
	r->ox = r->oy = r->oz = 0.f;
	r->dx = 0.f; r->dy = 0.f; r->dz = 1.f;
	r->mint = 0.f; r->maxt = 0.f; r->time = 0.f;
	r->flags = 1u;
	r->pad[0] = r->pad[1] = 0.f;
}

__global__ void lrh_camera_kernel(LrhCameraParams C, lrh_ray_rec *rays) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	const uint32_t n = C.width * C.height;
	if (i >= n) return;
	const uint32_t x = i % C.width, y = i / C.width;
	const float px = ((float)x + 0.5f) / (float)C.width * 2.f - 1.f;
	const float py = 1.f - ((float)y + 0.5f) / (float)C.height * 2.f;
	float d[3];
	for (int k = 0; k < 3; ++k) d[k] = C.fwd[k] + px * C.tw * C.right[k] + py * C.th * C.up[k];
	const float inv = 1.f / sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
	lrh_ray_rec r;
	r.ox = C.eye[0]; r.oy = C.eye[1]; r.oz = C.eye[2];
	r.dx = d[0] * inv; r.dy = d[1] * inv; r.dz = d[2] * inv;
	r.mint = C.mint;
	r.maxt = INFINITY;
	r.time = 0.f;
	r.flags = 0u;
	r.pad[0] = r.pad[1] = 0.f;
	rays[i] = r;
}

// one bounce of ray i; returns the number of live rays it wrote (0..2)
__device__ uint32_t bounceOne(const LrhBounceParams &P, uint32_t i) {
	const lrh_ray_rec ray = P.rays[i];
	const lrh_hit_rec hit = P.hits[i];
	uint32_t live = 0;
	lrh_ray_rec *next = P.next + i;
	lrh_ray_rec *shadow = P.shadow ? P.shadow + i : nullptr;
	if ((ray.flags & 1u) || hit.meshIndex >= P.nMeshes ||
			hit.triangleIndex >= P.meshTriBase[hit.meshIndex + 1] - P.meshTriBase[hit.meshIndex]) {
		writeMasked(next);
		if (shadow) writeMasked(shadow);
		return 0;
	}
	const uint32_t rec = P.invPerm[P.meshTriBase[hit.meshIndex] + hit.triangleIndex];
	const lrh_tri tr = P.tris[rec];
	// geometric normal, facing the incoming ray
	float n[3] = {tr.e1[1] * tr.e2[2] - tr.e1[2] * tr.e2[1], tr.e1[2] * tr.e2[0] - tr.e1[0] * tr.e2[2],
			tr.e1[0] * tr.e2[1] - tr.e1[1] * tr.e2[0]};
	const float nl = sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
	const float s = (n[0] * ray.dx + n[1] * ray.dy + n[2] * ray.dz > 0.f ? -1.f : 1.f) / fmaxf(nl, 1e-30f);
	for (int k = 0; k < 3; ++k) n[k] *= s;
	const float p[3] = {ray.ox + ray.dx * hit.t, ray.oy + ray.dy * hit.t, ray.oz + ray.dz * hit.t};
	const float eps = fmaxf(epsE(p[0]), fmaxf(epsE(p[1]), epsE(p[2])));
	const uint32_t pix = P.pixelBase + i;
	// Russian roulette from bounce rrStart on
	const bool alive = P.bounce < P.rrStart || rnd(P.seed, pix, P.bounce, 2) < P.rrSurvive;
	if (alive) {
		// cosine-weighted direction around n (Malley), branchless basis (Duff et al. 2017)
		const float u1 = rnd(P.seed, pix, P.bounce, 0), u2 = rnd(P.seed, pix, P.bounce, 1);
		const float r = sqrtf(u1), phi = 6.2831853f * u2;
		const float lx = r * cosf(phi), ly = r * sinf(phi), lz = sqrtf(fmaxf(0.f, 1.f - u1));
		const float sg = n[2] >= 0.f ? 1.f : -1.f;
		const float a = -1.f / (sg + n[2]);
		const float b = n[0] * n[1] * a;
		const float t0[3] = {1.f + sg * n[0] * n[0] * a, sg * b, -sg * n[0]};
		const float t1[3] = {b, sg + n[1] * n[1] * a, -n[1]};
		float d[3];
		for (int k = 0; k < 3; ++k) d[k] = lx * t0[k] + ly * t1[k] + lz * n[k];
		const float dl = 1.f / sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
		lrh_ray_rec o;
		o.ox = p[0]; o.oy = p[1]; o.oz = p[2];
		o.dx = d[0] * dl; o.dy = d[1] * dl; o.dz = d[2] * dl;
		o.mint = eps;
		o.maxt = INFINITY;
		o.time = ray.time;
		o.flags = 0u;
		o.pad[0] = o.pad[1] = 0.f;
		*next = o;
		++live;
	} else
		writeMasked(next);
	if (shadow) {
		const float lx = P.light[0] + rnd(P.seed, pix, P.bounce, 3) - 0.5f;
		const float ly = P.light[1] + rnd(P.seed, pix, P.bounce, 4) - 0.5f;
		const float v[3] = {lx - p[0], ly - p[1], P.light[2] - p[2]};
		const float dist = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
		lrh_ray_rec o;
		o.ox = p[0]; o.oy = p[1]; o.oz = p[2];
		o.dx = v[0] / dist; o.dy = v[1] / dist; o.dz = v[2] / dist;
		o.mint = eps;
		o.maxt = dist - 2.f * eps;
		o.time = ray.time;
		o.flags = dist > 4.f * eps ? 0u : 1u;
		o.pad[0] = o.pad[1] = 0.f;
		*shadow = o;
		live += o.flags ? 0u : 1u;
	}
	return live;
}

// ---- one shading step (see lrh_shade in lrhip.h) ----
__device__ __forceinline__ void cosineDir(const float n[3], float u1, float u2, float d[3]) {
	// cosine-weighted direction around n (Malley), branchless basis (Duff et al. 2017)
	const float r = sqrtf(u1), phi = 6.2831853f * u2;
	const float lx = r * cosf(phi), ly = r * sinf(phi), lz = sqrtf(fmaxf(0.f, 1.f - u1));
	const float sg = n[2] >= 0.f ? 1.f : -1.f;
	const float a = -1.f / (sg + n[2]);
	const float b = n[0] * n[1] * a;
	const float t0[3] = {1.f + sg * n[0] * n[0] * a, sg * b, -sg * n[0]};
	const float t1[3] = {b, sg + n[1] * n[1] * a, -n[1]};
	float v[3];
	for (int k = 0; k < 3; ++k) v[k] = lx * t0[k] + ly * t1[k] + lz * n[k];
	const float dl = 1.f / sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
	for (int k = 0; k < 3; ++k) d[k] = v[k] * dl;
}

__device__ uint32_t shadeOne(const LrhShadeParams &P, uint32_t i) {
	LrhPath path = P.paths[i];
	const bool pending = path.dl[0] != 0.f || path.dl[1] != 0.f || path.dl[2] != 0.f;
	// a path already ended when the previous step ran (pad = 1) and with no
	// light pending: that step and the one that ended it masked its slots of
	// both ray buffers and of the shadow buffer, so there is nothing to do
	if (!path.alive && path.pad && !pending) return 0;
	// the previous step's shadow ray: its light counts if nothing blocked it
	if (P.shadowHits && pending && P.shadowHits[i].meshIndex == 0xffffffffu)
		for (int k = 0; k < 3; ++k) path.rad[k] = path.rad[k] + path.dl[k];
	for (int k = 0; k < 3; ++k) path.dl[k] = 0.f;
	lrh_ray_rec *next = P.next + i;
	lrh_ray_rec *shadow = P.shadow + i;
	if (!path.alive) {
		path.pad = 1u;
		writeMasked(next);
		writeMasked(shadow);
		P.paths[i] = path;
		return 0;
	}
	const lrh_ray_rec ray = P.rays[i];
	const lrh_hit_rec hit = P.hits[i];
	if ((ray.flags & 1u) || hit.meshIndex >= P.nMeshes ||
			hit.triangleIndex >= P.meshTriBase[hit.meshIndex + 1] - P.meshTriBase[hit.meshIndex]) {
		path.alive = 0u;
		writeMasked(next);
		writeMasked(shadow);
		P.paths[i] = path;
		return 0;
	}
	LrhMaterial m;
	if (hit.meshIndex < P.nMaterials)
		m = P.materials[hit.meshIndex];
	else
		for (int k = 0; k < 3; ++k) m.kd[k] = m.le[k] = 0.f;
	const uint32_t rec = P.invPerm[P.meshTriBase[hit.meshIndex] + hit.triangleIndex];
	const lrh_tri tr = P.tris[rec];
	float n[3] = {tr.e1[1] * tr.e2[2] - tr.e1[2] * tr.e2[1], tr.e1[2] * tr.e2[0] - tr.e1[0] * tr.e2[2],
			tr.e1[0] * tr.e2[1] - tr.e1[1] * tr.e2[0]};
	const float nl = sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
	const float s = (n[0] * ray.dx + n[1] * ray.dy + n[2] * ray.dz > 0.f ? -1.f : 1.f) / fmaxf(nl, 1e-30f);
	for (int k = 0; k < 3; ++k) n[k] *= s;
	const float p[3] = {ray.ox + ray.dx * hit.t, ray.oy + ray.dy * hit.t, ray.oz + ray.dz * hit.t};
	const float eps = fmaxf(epsE(p[0]), fmaxf(epsE(p[1]), epsE(p[2])));
	const uint32_t pix = P.pixelBase + i;
	const uint32_t depth = path.depth;
	if (depth == 0)  // a light seen directly
		for (int k = 0; k < 3; ++k) path.rad[k] = path.rad[k] + path.thr[k] * m.le[k];
	const bool diffuse = m.kd[0] > 0.f || m.kd[1] > 0.f || m.kd[2] > 0.f;
	uint32_t live = 0;
	bool shadowWritten = false;
	if (P.nLights && diffuse) {
		const float u0 = rnd(P.seed, pix, depth, 3), u1 = rnd(P.seed, pix, depth, 4), u2 = rnd(P.seed, pix, depth, 5);
		uint32_t j = (uint32_t)(u0 * (float)P.nLights);
		if (j > P.nLights - 1) j = P.nLights - 1;
		const LrhLightTri L = P.lights[j];
		const float su = sqrtf(u1), a = su * (1.f - u2), b = su * u2;
		float w[3];
		for (int k = 0; k < 3; ++k) w[k] = (L.p0[k] + a * L.e1[k] + b * L.e2[k]) - p[k];
		const float d2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
		const float dist = sqrtf(d2);
		float wn[3];
		for (int k = 0; k < 3; ++k) wn[k] = w[k] / dist;
		const float ln[3] = {L.e1[1] * L.e2[2] - L.e1[2] * L.e2[1], L.e1[2] * L.e2[0] - L.e1[0] * L.e2[2],
				L.e1[0] * L.e2[1] - L.e1[1] * L.e2[0]};
		const float lnl = sqrtf(ln[0] * ln[0] + ln[1] * ln[1] + ln[2] * ln[2]);
		const float cosS = n[0] * wn[0] + n[1] * wn[1] + n[2] * wn[2];
		const float cosL = fabsf(ln[0] * wn[0] + ln[1] * wn[1] + ln[2] * wn[2]) / lnl;
		if (cosS > 0.f && cosL > 0.f && dist > 4.f * eps) {
			// uniform light pick (1/n) and point (1/area): 1/pdf = n * area
			const float g = ((cosS * cosL) * (0.5f * lnl * (float)P.nLights)) / d2 * 0.31830988f;
			for (int k = 0; k < 3; ++k) path.dl[k] = path.thr[k] * m.kd[k] * L.le[k] * g;
			lrh_ray_rec o;
			o.ox = p[0]; o.oy = p[1]; o.oz = p[2];
			o.dx = wn[0]; o.dy = wn[1]; o.dz = wn[2];
			o.mint = eps;
			o.maxt = dist - 2.f * eps;
			o.time = ray.time;
			o.flags = 0u;
			o.pad[0] = o.pad[1] = 0.f;
			*shadow = o;
			shadowWritten = true;
			++live;
		}
	}
	if (!shadowWritten) writeMasked(shadow);
	const uint32_t depth1 = depth + 1;
	bool alive = depth1 < P.maxDepth && diffuse;
	if (alive && depth1 >= P.rrDepth) {
		if (rnd(P.seed, pix, depth, 2) < P.rrSurvive)
			for (int k = 0; k < 3; ++k) path.thr[k] = path.thr[k] / P.rrSurvive;
		else
			alive = false;
	}
	if (alive) {
		float d[3];
		cosineDir(n, rnd(P.seed, pix, depth, 0), rnd(P.seed, pix, depth, 1), d);
		// Lambertian: f cos / pdf = kd for cosine-weighted directions
		for (int k = 0; k < 3; ++k) path.thr[k] = path.thr[k] * m.kd[k];
		lrh_ray_rec o;
		o.ox = p[0]; o.oy = p[1]; o.oz = p[2];
		o.dx = d[0]; o.dy = d[1]; o.dz = d[2];
		o.mint = eps;
		o.maxt = INFINITY;
		o.time = ray.time;
		o.flags = 0u;
		o.pad[0] = o.pad[1] = 0.f;
		*next = o;
		++live;
	} else
		writeMasked(next);
	path.depth = depth1;
	path.alive = alive ? 1u : 0u;
	P.paths[i] = path;
	return live;
}

__device__ __forceinline__ void countLive(uint32_t *counter, uint32_t live) {
	// one atomic per wave
	unsigned long long v = live;
	for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
	if (counter && __lane_id() == 0 && v) atomicAdd(counter, (uint32_t)v);
}

__global__ void lrh_bounce_kernel(LrhBounceParams P) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	uint32_t live = 0;
	if (i < P.count) live = bounceOne(P, i);
	countLive(P.liveCount, live);
}

__global__ void lrh_shade_kernel(LrhShadeParams P) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	uint32_t live = 0;
	if (i < P.count) live = shadeOne(P, i);
	countLive(P.liveCount, live);
}

__global__ void lrh_shade_batch_kernel(LrhShadeBatchArgs A) {
	// the item of this block (uniform over the block: scalar loads of the
	// kernel arguments)
	uint32_t k = 0;
	while (k + 1 < A.n && blockIdx.x >= A.blockStart[k + 1]) ++k;
	const LrhShadeParams &P = A.p[k];
	const uint32_t i = (blockIdx.x - A.blockStart[k]) * blockDim.x + threadIdx.x;
	uint32_t live = 0;
	if (i < P.count) live = shadeOne(P, i);
	countLive(P.liveCount, live);
}

__global__ void lrh_paths_init_kernel(LrhPath *paths, uint32_t count) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= count) return;
	LrhPath p;
	for (int k = 0; k < 3; ++k) {
		p.thr[k] = 1.f;
		p.rad[k] = 0.f;
		p.dl[k] = 0.f;
	}
	p.depth = 0u;
	p.alive = 1u;
	p.pad = 0u;
	paths[i] = p;
}

}  // namespace

hipError_t lrhLaunchCamera(const LrhCameraParams &C, lrh_ray_rec *rays, hipStream_t s) {
	const uint32_t n = C.width * C.height;
	hipLaunchKernelGGL(lrh_camera_kernel, dim3((n + 255) / 256), dim3(256), 0, s, C, rays);
	return hipGetLastError();
}

hipError_t lrhLaunchPathsInit(LrhPath *paths, uint32_t count, hipStream_t s) {
	hipLaunchKernelGGL(lrh_paths_init_kernel, dim3((count + 255) / 256), dim3(256), 0, s, paths, count);
	return hipGetLastError();
}

hipError_t lrhLaunchShade(const LrhShadeParams &P, hipStream_t s) {
	hipLaunchKernelGGL(lrh_shade_kernel, dim3((P.count + 255) / 256), dim3(256), 0, s, P);
	return hipGetLastError();
}

hipError_t lrhLaunchShadeBatch(const LrhShadeBatchArgs &A, hipStream_t s) {
	if (A.n == 0 || A.n > (uint32_t)kShadeBatchMax) return hipErrorInvalidValue;
	hipLaunchKernelGGL(lrh_shade_batch_kernel, dim3(A.blockStart[A.n]), dim3(256), 0, s, A);
	return hipGetLastError();
}

hipError_t lrhLaunchBounce(const LrhBounceParams &P, hipStream_t s) {
	hipLaunchKernelGGL(lrh_bounce_kernel, dim3((P.count + 255) / 256), dim3(256), 0, s, P);
	return hipGetLastError();
}
