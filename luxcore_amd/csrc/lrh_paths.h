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

hipError_t lrhLaunchCamera(const LrhCameraParams &C, lrh_ray_rec *rays, hipStream_t s);
hipError_t lrhLaunchBounce(const LrhBounceParams &P, hipStream_t s);

#endif
