/*
 * Device refit of a two-level scene's root tree (see lrh_refit.hip).
 */
#ifndef LRH_REFIT_H
#define LRH_REFIT_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lrh_layout.h"

// a motion object: its leaf slot in the root tree, its lrh_motion index and
// the mesh whose box it moves
struct LrhRefitMotion {
	uint32_t slot, motion, mesh, pad;
};

struct LrhRefitArgs {
	// motion bounds
	const LrhRefitMotion *motionObjs = nullptr;  // device, nMotion
	uint32_t nMotion = 0;
	const uint32_t *sampleBase = nullptr;        // device, nMotion + 1 prefix of segments x 1025
	uint32_t nSamples = 0;
	const lrh_motion *motions = nullptr;         // device
	const lrh_motion_seg *segs = nullptr;        // device
	const float *meshBoxes = nullptr;            // device, 6 per mesh (lo xyz, hi xyz)
	uint32_t *keys = nullptr;                    // device scratch, 6 per motion object
	uint32_t *error = nullptr;                   // device word: nonzero if a sample is singular
	// root nodes
	float *slotBoxes = nullptr;                  // device, 6 per leaf slot (instance and plain boxes filled in)
	float *nodeBoxes = nullptr;                  // device scratch, 6 per root node
	const uint32_t *levelNodes = nullptr;        // device, root node indices, deepest level first
	const uint32_t *levelStart = nullptr;        // host, nLevels + 1 offsets into levelNodes
	uint32_t nLevels = 0;
	uint32_t layout = 0;
	void *nodes = nullptr;                       // device node array (root tree at its start)
};

// phase 1: motion objects' boxes into slotBoxes (sets *error on a singular sample)
hipError_t lrhLaunchRefitMotion(const LrhRefitArgs &a, hipStream_t stream);
// phase 2: root nodes bottom-up from slotBoxes
hipError_t lrhLaunchRefitLevels(const LrhRefitArgs &a, hipStream_t stream);

#endif
