/* Launch interface of the ray-generation kernels (camera, bounce + shadow). */
#ifndef LRH_PATHS_H
#define LRH_PATHS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lrh_layout.h"

struct LrhCameraParams {
	float eye[3], fwd[3], right[3], up[3];
	float tw, th;        /* tan of the half field of view, horizontal / vertical */
	float mint;
	uint32_t width, height;
};

struct LrhBounceParams {
	const lrh_ray_rec *rays;
	const lrh_hit_rec *hits;
	lrh_ray_rec *next;
	lrh_ray_rec *shadow;          /* may be null */
	const lrh_tri *tris;
	const uint32_t *invPerm;      /* triangle (global, DataSet order) -> record */
	const uint32_t *meshTriBase;  /* first global triangle of each mesh */
	uint32_t nMeshes;
	uint32_t count;
	uint32_t bounce;
	uint32_t rrStart;
	float rrSurvive;
	float light[3];
	uint64_t seed;
	uint32_t pixelBase;
	uint32_t *liveCount;          /* may be null: += live next + shadow rays */
};

struct LrhMaterial { float kd[3], le[3]; };
struct LrhLightTri { float p0[3], e1[3], e2[3], le[3]; };
struct LrhPath { float thr[3], rad[3], dl[3]; uint32_t depth, alive, pad; };

struct LrhShadeParams {
	const lrh_ray_rec *rays;
	const lrh_hit_rec *hits;
	const lrh_hit_rec *shadowHits;  /* may be null */
	LrhPath *paths;
	lrh_ray_rec *next;
	lrh_ray_rec *shadow;
	const LrhMaterial *materials;
	uint32_t nMaterials;
	const LrhLightTri *lights;
	uint32_t nLights;
	const lrh_tri *tris;
	const uint32_t *invPerm;
	const uint32_t *meshTriBase;
	uint32_t nMeshes;
	uint32_t count;
	uint32_t maxDepth;
	uint32_t rrDepth;
	float rrSurvive;
	uint64_t seed;
	uint32_t pixelBase;
	uint32_t *liveCount;
};

hipError_t lrhLaunchCamera(const LrhCameraParams &C, lrh_ray_rec *rays, hipStream_t s);
hipError_t lrhLaunchPathsInit(LrhPath *paths, uint32_t count, hipStream_t s);
hipError_t lrhLaunchShade(const LrhShadeParams &P, hipStream_t s);
// several shading steps in one launch: item k owns blocks [blockStart[k], blockStart[k+1])
constexpr int kShadeBatchMax = 16;
struct LrhShadeBatchArgs {
	uint32_t n;
	uint32_t blockStart[kShadeBatchMax + 1];
	LrhShadeParams p[kShadeBatchMax];
};
// passed by value as the kernel's argument: within the 4 KiB argument limit
static_assert(sizeof(LrhShadeBatchArgs) <= 4096, "shade batch arguments exceed 4 KiB");
hipError_t lrhLaunchShadeBatch(const LrhShadeBatchArgs &A, hipStream_t s);
hipError_t lrhLaunchBounce(const LrhBounceParams &P, hipStream_t s);

#endif
