/* Launch interface between the C-ABI layer and the traversal kernels. */
#ifndef LRH_KERNELS_H
#define LRH_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lrh_layout.h"

#define LRH_ITEM_ANY 1u
#define LRH_ITEM_TWO 2u

/* One ray buffer of a batched launch, as the kernel reads it (96 bytes). */
struct LrhBatchDesc {
	const void *nodes;
	const lrh_tri *tris;
	const lrh_inst *insts;
	const float *xforms;
	const lrh_motion *motions;
	const lrh_motion_seg *segs;
	const float *times;
	const lrh_ray_rec *rays;
	lrh_hit_rec *hits;
	uint32_t root;
	uint32_t count;
	uint32_t flags;                 /* LRH_ITEM_ANY | LRH_ITEM_TWO */
	uint32_t pad[3];
};

struct LrhTraceParams {
	const void *nodes;              /* lrh_node4q[] or lrh_node4f[] */
	const lrh_tri *tris;
	const lrh_inst *insts;
	const float *xforms;            /* 16 floats (world->local) per static instance */
	const lrh_motion *motions;
	const lrh_motion_seg *segs;
	const float *times;
	const lrh_ray_rec *rays;
	lrh_hit_rec *hits;
	uint32_t nRays;
	uint32_t root;                  /* child reference of the (root) tree */
	uint32_t *counter;              /* [0] ray tickets, [1] finished blocks; the last block rewinds both */
	uint2 *spill;                   /* spill stack: spillDepth entries x nLanes */
	uint32_t spillDepth;
	uint32_t nLanes;
	uint32_t ldsDepth;              /* power of two */
	uint32_t refill;                /* refill a wave when fewer lanes than this are live */
	uint32_t leafTris;              /* triangles a lane tests per pass (1..8); the rest of a leaf waits a pass */
	uint32_t *error;                /* bit 0 stack overflow, bit 1 step guard */
	unsigned long long *stats;      /* counting variant: LRH_TRACE_STATS counters */
	const LrhBatchDesc *items;      /* batched: item descriptors */
	const uint32_t *itemStart;      /* batched: nItems+1 ticket offsets (prefix sums of counts) */
	uint32_t nItems;
};

hipError_t lrhLaunchReset(uint32_t *p, uint32_t n, hipStream_t s);
hipError_t lrhOccupancy(int layout, bool two, bool any, bool batch, int block, size_t lds,
		int *blocksPerCU);
hipError_t lrhLaunchTrace(int layout, bool two, bool any, bool count, bool batch, dim3 grid, dim3 block,
		size_t lds, hipStream_t s, const LrhTraceParams &P);

#endif
