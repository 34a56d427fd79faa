/*
 * Device-resident record layouts of the HIP intersection device (host and
 * device share this header).
 *
 * The reference keeps one flattened, depth-first skip-pointer array of
 * 32-byte BVHArrayNode records and gathers three vertices per leaf
 * (include/luxrays/core/bvh/bvhbuild_types.cl:21-42,
 * include/luxrays/accelerators/bvh.cl:136-217).  Here the tree is 4-wide and
 * every node is one 64-byte record = one half of an L2 line, so one visit is
 * one aligned 64-byte fetch; triangle leaves are runs of 48-byte records that
 * carry their vertices inline (no index gather).
 */
#ifndef LRH_LAYOUT_H
#define LRH_LAYOUT_H

#include <stdint.h>

#define LRH_NULL_INDEX 0xffffffffu

/* Child references (uint32) inside a node:
 *   0xffffffff                  empty slot
 *   bit31 = 0                   inner child: node index
 *   bit31 = 1                   leaf: bits 28..30 = count-1 (1..8 records),
 *                               bits 0..27 = first record index
 * In a two-level root tree a leaf record is an instance record. */
#define LRH_CHILD_EMPTY 0xffffffffu
#define LRH_CHILD_LEAF 0x80000000u
#define LRH_LEAF_MAX 8u
#define LRH_TOP_NODES 4096u   /* the builder stores this many top nodes breadth first */
#define LRH_LEAF_START_BITS 28u
#define LRH_LEAF_START_MASK 0x0fffffffu
#define LRH_MAX_LEAF_START LRH_LEAF_START_MASK

static inline uint32_t lrh_leaf_ref(uint32_t start, uint32_t count) {
	return LRH_CHILD_LEAF | ((count - 1u) << LRH_LEAF_START_BITS) | start;
}

/* 64-byte 4-wide node with 8-bit child bounds rounded outward.
 * Child c, axis k: lo = org[k] + qlo[k][c] * 2^(e[k]-127),
 *                  hi = org[k] + qhi[k][c] * 2^(e[k]-127);
 * q*2^e is exact, so the decode rounds once and the builder checks the
 * decoded value against the true bound with the same fp32 operations. */
typedef struct __attribute__((aligned(64))) {
	float org[3];        /*  0 */
	uint8_t ex, ey, ez;  /* 12 */
	uint8_t nchild;      /* 15 */
	uint8_t qlo_x[4];    /* 16 */
	uint8_t qhi_x[4];    /* 20 */
	uint8_t qlo_y[4];    /* 24 */
	uint8_t qhi_y[4];    /* 28 */
	uint8_t qlo_z[4];    /* 32 */
	uint8_t qhi_z[4];    /* 36 */
	uint32_t child[4];   /* 40 */
	uint32_t pad[2];     /* 56 */
} lrh_node4q;

/* 128-byte 8-wide node with 8-bit child bounds, one L2 line: the same
 * quantization as lrh_node4q for eight children (single-level scenes). */
typedef struct __attribute__((aligned(128))) {
	float org[3];        /*  0 */
	uint8_t ex, ey, ez;  /* 12 */
	uint8_t nchild;      /* 15 */
	uint8_t qlo_x[8];    /* 16 */
	uint8_t qhi_x[8];    /* 24 */
	uint8_t qlo_y[8];    /* 32 */
	uint8_t qhi_y[8];    /* 40 */
	uint8_t qlo_z[8];    /* 48 */
	uint8_t qhi_z[8];    /* 56 */
	uint32_t child[8];   /* 64 */
	uint32_t pad[8];     /* 96 */
} lrh_node8q;

/* 128-byte 4-wide node with fp32 child bounds (autotuner alternative). */
typedef struct __attribute__((aligned(64))) {
	float lo_x[4], hi_x[4], lo_y[4], hi_y[4], lo_z[4], hi_z[4]; /* 0..95 */
	uint32_t child[4];                                           /* 96 */
	uint32_t pad[4];                                             /* 112 */
} lrh_node4f;

/* 48-byte triangle record: p0 and the two edges exactly as
 * Triangle::Intersect computes them (e1 = p1 - p0, e2 = p2 - p0 in fp32,
 * include/luxrays/core/geometry/triangle.h:57-58), so the device test rounds
 * bit for bit like the reference. */
typedef struct __attribute__((aligned(16))) {
	float p0[3];
	float e1[3];
	float e2[3];
	uint32_t triIndex;
	uint32_t meshIndex;
	uint32_t pad;
} lrh_tri;

/* Two-level instance record (root-tree leaf).  kind: 0 = no transform,
 * 1 = static world->local matrix xf[index], 2 = motion system motion[index].
 * root = child reference of the object's leaf tree. */
typedef struct __attribute__((aligned(16))) {
	uint32_t root;
	uint32_t kind;
	uint32_t index;
	uint32_t object;
} lrh_inst;

/* Restated InterpolatedTransform for the device
 * (src/luxrays/core/geometry/motionsystem.cpp:37-159).  flags bits:
 * 0 active, 1 rotation, 2 translation, 3 scale, 4/5/6 translation x/y/z,
 * 7/8/9 scale x/y/z.  The first 112 bytes hold everything a sample between
 * the key times reads (seven 16-byte loads issued together); the endpoint
 * matrices and startT's rotation follow. */
typedef struct __attribute__((aligned(16))) {
	float startTime, endTime;
	uint32_t flags;
	/* Slerp's per-segment constants (quaternion.cpp:143-164): the sign that
	 * takes the shorter arc, and, when 1 - cos_phi > 1e-6 (slerpArc = 1),
	 * phi = acos(cos_phi) and sin(phi), each evaluated in double and rounded
	 * once to float; only sin((1-t) phi) and sin(t phi) depend on the time */
	uint32_t slerpArc;
	float qsign, phi, sinPhi;
	uint32_t pad0;
	float startQ[4];    /* w, x, y, z */
	float endQ[4];
	float sS[3], eS[3]; /* Sx Sy Sz of startT / endT */
	float sT[3], eT[3]; /* Tx Ty Tz of startT / endT */
	float start[16];    /* start.m */
	float end[16];      /* end.m */
	float startR[16];   /* startT.R */
} lrh_motion_seg;

#define LRH_SEG_HOT_BYTES 112

/* A motion system: keys times[0..nTimes) and nTimes+1 segments.  The first
 * LRH_MOTION_INLINE_TIMES key times are repeated in the record, so choosing
 * the segment of a ray's time needs no load beyond the record. */
#define LRH_MOTION_INLINE_TIMES 4
typedef struct __attribute__((aligned(16))) {
	uint32_t firstTime;
	uint32_t nTimes;
	uint32_t firstSeg;
	uint32_t pad;
	float t[LRH_MOTION_INLINE_TIMES];
} lrh_motion;

enum {
	LRH_MF_ACTIVE = 1u << 0, LRH_MF_ROT = 1u << 1, LRH_MF_TRANS = 1u << 2, LRH_MF_SCALE = 1u << 3,
	LRH_MF_TX = 1u << 4, LRH_MF_TY = 1u << 5, LRH_MF_TZ = 1u << 6,
	LRH_MF_SX = 1u << 7, LRH_MF_SY = 1u << 8, LRH_MF_SZ = 1u << 9
};

/* Ray / RayHit device views (48 / 20 bytes, include/luxrays/core/geometry/ray.h:35-103) */
typedef struct {
	float ox, oy, oz, dx, dy, dz, mint, maxt, time;
	uint32_t flags;
	float pad[2];
} lrh_ray_rec;

typedef struct {
	float t, b1, b2;
	uint32_t meshIndex, triangleIndex;
} lrh_hit_rec;

#endif
