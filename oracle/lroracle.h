/*
 * lroracle — CPU restatement of LuxRays' CPU intersection path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * intersection device: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  Product code never links,
 * imports or calls it.
 *
 * Every function restates the reference algorithm it names (paths are
 * relative to the LuxCore v2.6 source tree) and is compiled with the
 * reference's CPU flags (-O3 -msse -msse2 -msse3 -mssse3, no fast-math)
 * plus -ffp-contract=off, so each fp32 operation rounds exactly as the
 * reference's SSE build does.
 *
 * Pinning (what checks each function against something other than itself):
 *   orc_triangle_intersect  closed-form dyadic vectors (tests/golden)
 *   orc_bbox_intersectp     closed-form slab cases incl. zero directions
 *   orc_epsilon*            closed forms: E(1)=2^-16, clamps at 1e-5 / 0.1
 *   orc_bvh_* (CLASSIC/SAH) tree invariants + brute force (orc_bruteforce)
 *   orc_bruteforce          dyadic known answers and a planted-mistake check
 *   orc_mbvh_*              identity instance == plain BVH; static instance
 *                           == pre-transformed copy; brute force per instance
 *   orc_motion_sample       closed forms (endpoints, translation lerp,
 *                           axis rotation angle lerp, TRS round trip); its
 *                           slerp sines are the C library's sin in double,
 *                           which the device's own sine matches on every
 *                           float argument in [0, pi] (exhaustive test)
 *   orc_mbvh_* with projective instance matrices (last row not 0,0,0,1):
 *                           parity unpinned — MBVHAccel's result then depends
 *                           on its root-tree topology (root leaves are not
 *                           box-tested, only their parents)
 */
#ifndef LRORACLE_H
#define LRORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Ray (48 B) and RayHit (20 B): include/luxrays/core/geometry/ray.h:35-103 */
typedef struct {
	float ox, oy, oz;
	float dx, dy, dz;
	float mint, maxt, time;
	unsigned int flags;
	float pad[2];
} orc_ray;

typedef struct {
	float t, b1, b2;
	unsigned int meshIndex, triangleIndex;
} orc_rayhit;

/* BVHArrayNode (32 B): include/luxrays/core/bvh/bvhbuild_types.cl:21-42 */
typedef struct {
	union {
		struct { float bboxMin[3]; float bboxMax[3]; } bvhNode;
		struct { unsigned int v[3]; unsigned int meshIndex, triangleIndex; } triangleLeaf;
		struct { unsigned int leafIndex, transformIndex, motionIndex, meshOffsetIndex; } bvhLeaf;
	};
	unsigned int nodeData;
	int pad0;
} orc_bvh_node;

enum { ORC_BUILDER_CLASSIC = 0, ORC_BUILDER_SAH = 1 };

typedef struct orc_bvh orc_bvh;
typedef struct orc_mbvh orc_mbvh;

/* ---- geometry kernels ---- */
int orc_triangle_intersect(const orc_ray *ray, const float p0[3], const float p1[3],
		const float p2[3], float *t, float *b1, float *b2);
int orc_bbox_intersectp(const orc_ray *ray, const float pmin[3], const float pmax[3],
		float *t0, float *t1);
float orc_epsilon(float v);
float orc_epsilon_bbox(const float pmin[3], const float pmax[3]);

/* ---- single-level BVH (BVHAccel) ----
 * Meshes are given as concatenated arrays: mesh m owns vertices
 * [vertOff[m], vertOff[m+1]) and triangles [triOff[m], triOff[m+1]); the
 * triangle vertex indices are local to the mesh.  Mesh index = add order. */
orc_bvh *orc_bvh_create(unsigned nMeshes, const unsigned *vertOff, const unsigned *triOff,
		const float *verts, const unsigned *tris, int builder, int treeType,
		int costSamples, int isectCost, int travCost, float emptyBonus, int nThreads);
void orc_bvh_destroy(orc_bvh *b);
unsigned orc_bvh_node_count(const orc_bvh *b);
const orc_bvh_node *orc_bvh_nodes(const orc_bvh *b);
/* Batch wrapper with the ray-buffer semantics of the device kernel
 * (masked rays untouched, miss -> t=maxt, both indices 0xffffffff, b1=b2=0).
 * counters (may be NULL) receives {inner-node visits, leaf visits}. */
void orc_bvh_intersect(const orc_bvh *b, const orc_ray *rays, orc_rayhit *hits,
		size_t n, int nThreads, unsigned long long *counters);

/* O(N*R) closest hit over all triangles, same acceptance rule (t < best). */
void orc_bruteforce(unsigned nMeshes, const unsigned *vertOff, const unsigned *triOff,
		const float *verts, const unsigned *tris, const orc_ray *rays, orc_rayhit *hits,
		size_t n, int nThreads);

/* ---- transforms and motion ---- */
/* Slerp's acosf/sinf: 0 (default) = the host libm's acosf/sinf, as the
 * reference calls them; 1 = correctly rounded fp32 results.  Process-wide;
 * set before building a two-level oracle (motion bounds) and before tracing
 * it. */
void orc_set_libm(int correctlyRounded);
/* Gauss-Jordan inverse; returns 0 on a singular matrix. */
int orc_matrix_inverse(const float m[16], float out[16]);
/* MotionSystem built from nKeys (time, global->local matrix) keys; samples
 * the interpolated matrix at `time`. Returns 0 on a singular key. */
int orc_motion_sample(unsigned nKeys, const float *times, const float *mats,
		float time, float out[16]);
/* MotionSystem::Bound(box, true) */
int orc_motion_bound(unsigned nKeys, const float *times, const float *mats,
		const float pmin[3], const float pmax[3], float omin[3], float omax[3]);

/* ---- two-level BVH (MBVHAccel) ----
 * Base meshes as above.  Each object o refers to base mesh objMesh[o] and
 * has kind objKind[o]: 0 = plain mesh (only allowed once per base mesh and
 * then that mesh is not shared), 1 = instance with local->world matrix
 * objMat[16*o], 2 = motion with keys objKeyOff[o]..objKeyOff[o+1] into
 * keyTimes[] / keyMats[16*k] (global->local matrices). */
orc_mbvh *orc_mbvh_create(unsigned nMeshes, const unsigned *vertOff, const unsigned *triOff,
		const float *verts, const unsigned *tris,
		unsigned nObjs, const unsigned *objMesh, const unsigned *objKind, const float *objMat,
		const unsigned *objKeyOff, const float *keyTimes, const float *keyMats,
		int builder, int treeType, int nThreads);
void orc_mbvh_destroy(orc_mbvh *b);
/* world-space root-leaf box of object o after epsilon expansion */
void orc_mbvh_leaf_bbox(const orc_mbvh *b, unsigned o, float pmin[3], float pmax[3]);
void orc_mbvh_intersect(const orc_mbvh *b, const orc_ray *rays, orc_rayhit *hits,
		size_t n, int nThreads, unsigned long long *counters);
/* brute force over all objects: ray transformed per object exactly as at
 * the leaf-tree entry, every triangle tested, t < best kept */
void orc_mbvh_bruteforce(const orc_mbvh *b, const orc_ray *rays, orc_rayhit *hits,
		size_t n, int nThreads);

#ifdef __cplusplus
}
#endif

#endif
