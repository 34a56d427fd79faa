/*
 * Host-side scene preparation for the HIP intersection device: build-time
 * geometry (MachineEpsilon, transforms, motion systems) and the 4-wide
 * binned-SAH tree builder that emits the device node/record layouts of
 * lrh_layout.h.
 */
#ifndef LRH_BUILD_H
#define LRH_BUILD_H

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "lrh_layout.h"

namespace lrh {

struct Error : std::runtime_error {
	int code;
	Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

struct Box {
	float lo[3], hi[3];
	static Box Empty();
	void grow(const Box &b);
	void growPoint(const float p[3]);
	float area() const;
};

// MachineEpsilon::E (include/luxrays/core/epsilon.h:50-94)
float machineEpsilon(float v);
float machineEpsilon(const Box &b);
// Triangle::WorldBound + BBox::Expand(MachineEpsilon::E(bbox))
// (src/luxrays/accelerators/bvhaccel.cpp:116-122)
Box triangleBox(const float *p0, const float *p1, const float *p2);

// 4x4 row-major matrices
struct M4 { float m[4][4]; };
M4 identity();
// Gauss-Jordan full-pivot inverse (src/luxrays/core/geometry/matrix4x4.cpp:118-175)
bool invert(const M4 &in, M4 *out);
// Transform * BBox over 8 corners (include/luxrays/core/geometry/transform.h:271-280)
Box transformBox(const M4 &m, const Box &b);

// Restated InterpolatedTransform / MotionSystem
// (src/luxrays/core/geometry/motionsystem.cpp:37-365)
struct MotionSystemHost {
	std::vector<float> times;
	std::vector<lrh_motion_seg> segs;  // times.size() + 1 segments
	void init(unsigned nKeys, const float *keyTimes, const float *keyMats);
	M4 sample(float time) const;
	Box bound(const Box &ibox) const;  // Bound(ibox, storingGlobal2Local = true)
};

struct BuildParams {
	unsigned maxLeaf = 4;
	unsigned nThreads = 0;
	float travCost = 1.0f;
	unsigned layout = 0;  // 0 quantized 64 B, 1 fp32 128 B, 2 quantized 8-wide 128 B
	float splitBudget = 0.f;  // extra triangle references of early split clipping, per triangle; 0 = none
};

// Early split clipping (Ernst & Greiner, "Early Split Clipping for Bounding
// Volume Hierarchies", 2007): a triangle whose box is larger than the typical
// one is cut, by halving its box along the longest axis and clipping
// the triangle to each half, into several references with tighter boxes.
// Each reference box holds the triangle's part inside it (expanded by
// MachineEpsilon::E), so the union still covers the triangle.  boxes: one
// per triangle on input; refBoxes / refPrim: the references and their
// triangle, or both empty when nothing is cut.
struct TriangleSource {
	virtual ~TriangleSource() {}
	virtual void get(size_t prim, float p[3][3]) const = 0;
};
void splitLargeTriangles(const std::vector<Box> &boxes, const TriangleSource &tris, float budget, unsigned nThreads,
		std::vector<Box> *refBoxes, std::vector<uint32_t> *refPrim);

struct WideBVH {
	std::vector<lrh_node4q> qnodes;
	std::vector<lrh_node4f> fnodes;
	std::vector<lrh_node8q> q8nodes;  // layout 2: 8-wide
	std::vector<uint32_t> order;  // primitive order of the leaf ranges
	uint32_t root = LRH_CHILD_EMPTY;
	uint32_t maxDepth = 0;
	size_t nodeCount() const { return qnodes.size() + fnodes.size() + q8nodes.size(); }
};

// Builds a 4-wide tree over primitive boxes. Node indices start at nodeBase,
// leaf ranges at primBase (so several trees can share one array).
void buildWideBVH(const std::vector<Box> &boxes, const BuildParams &p, uint32_t nodeBase,
		uint32_t primBase, WideBVH *out);

unsigned hardwareThreads();

}  // namespace lrh

#endif
