/*
 * Device-side BVH build: Morton order of the triangle boxes, a binary tree
 * over them (the radix tree of a linear BVH, or agglomerative clustering of
 * neighbours in that order), then a breadth-first collapse into the same
 * 4-wide node records the host builder writes.  It stands in for the
 * host build of BVHAccel::Init (src/luxrays/accelerators/bvhaccel.cpp:72-168)
 * when the scene is built for a device.
 */
#ifndef LRH_GPUBUILD_H
#define LRH_GPUBUILD_H

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

struct LrhGpuAlloc {
	void *(*alloc)(void *ctx, size_t bytes);  // throws on failure
	void (*release)(void *ctx, void *p, size_t bytes);
	void *ctx;
};

struct LrhGpuBuildIn {
	const float *verts;          // device, all meshes' vertices concatenated (3 floats each)
	const uint32_t *tris;        // device, all meshes' triangles (3 mesh-local vertex indices each)
	const uint32_t *meshTriBase; // host, nMeshes + 1 prefix sums of triangle counts
	const uint32_t *meshVertBase;// host, nMeshes + 1 prefix sums of vertex counts
	uint32_t nMeshes;
	uint32_t nTris;
	uint32_t maxLeaf;            // 1..8
	uint32_t layout;             // 0 quantized 64 B nodes, 1 fp32 128 B nodes
	uint32_t method = 0;         // binary tree over the sorted codes: 0 radix tree (LBVH), 1 clustering (PLOC)
	float travCost = 1.f;        // cost of a binary node per box area, relative to one triangle test:
	                             // a node of at most maxLeaf triangles is one leaf when that costs no more
};

struct LrhGpuBuildOut {
	void *nodes = nullptr;       // device node records, breadth first from the root
	size_t nodeBytes = 0;
	uint32_t nodeCount = 0;
	void *tris = nullptr;        // device lrh_tri records in leaf order
	size_t triBytes = 0;
	uint32_t *invPerm = nullptr; // device, global triangle id -> record index
	size_t invPermBytes = 0;
	uint32_t root = 0xffffffffu;
	uint32_t maxDepth = 0;
	uint32_t plocRounds = 0;     // clustering rounds (method 1)
	double kernelMs = 0.0;
};

// Builds on `stream` and waits for it.  Throws std::runtime_error on HIP errors.
void lrhGpuBuildBVH(const LrhGpuBuildIn &in, const LrhGpuAlloc &a, hipStream_t stream, LrhGpuBuildOut *out);

#endif
