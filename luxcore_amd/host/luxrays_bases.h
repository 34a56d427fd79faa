/*
 * luxrays_bases — LuxRays' interface for this path restated: the geometry
 * records and meshes a DataSet holds, the configuration properties, and the
 * abstract device / accelerator / kernel classes with the reference's exact
 * virtual signatures, so that a device deriving from them (luxrays_hip.h,
 * and the binding a maintainer adds to LuxRays) is checked by `override`
 * against the same interface LuxRays declares:
 *
 *   Ray / RayHit                  include/luxrays/core/geometry/ray.h:35-103
 *   Mesh, TriangleMesh, InstanceTriangleMesh, MotionTriangleMesh
 *                                 include/luxrays/core/trianglemesh.h:42-391
 *   Transform, MotionSystem       include/luxrays/core/geometry/{transform,motionsystem}.h
 *   Properties (keys and typed values)
 *                                 include/luxrays/utils/properties.h
 *   DeviceType, DeviceDescription include/luxrays/core/device.h:39-82
 *   Device                        include/luxrays/core/device.h:107-138
 *   HardwareDeviceRange / Kernel / Program / Buffer, BufferType, HardwareDevice
 *                                 include/luxrays/core/hardwaredevice.h:31-183
 *   IntersectionDevice            include/luxrays/core/intersectiondevice.h:30-85
 *   HardwareIntersectionDevice, HardwareIntersectionKernel
 *                                 include/luxrays/core/hardwareintersectiondevice.h:30-72
 *   AcceleratorType, Accelerator  include/luxrays/core/accelerator.h:31-59
 *
 * Namespace luxrays_hip, so the restatement links next to a real LuxRays
 * build.  DeviceType carries the bit the HIP device registers under.
 */
#ifndef LUXRAYS_BASES_H
#define LUXRAYS_BASES_H

#include <chrono>
#include <cstdlib>
#include <deque>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace luxrays_hip {

typedef unsigned int u_int;
typedef unsigned long long u_longlong;

#define RAY_FLAGS_NONE 0x00000000u
#define RAY_FLAGS_MASKED 0x00000001u

struct Point {
	float x, y, z;
	Point() : x(0.f), y(0.f), z(0.f) {}
	Point(float a, float b, float c) : x(a), y(b), z(c) {}
};
typedef Point Vector;

// 48-byte Ray (ray.h:35-88)
class Ray {
public:
	Ray() : mint(1e-5f), maxt(INFINITY_F()), time(0.f), flags(RAY_FLAGS_NONE) { pad[0] = pad[1] = 0.f; }
	Ray(const Point &origin, const Vector &direction, float start, float end = INFINITY_F(), float t = 0.f)
		: o(origin), d(direction), mint(start), maxt(end), time(t), flags(RAY_FLAGS_NONE) { pad[0] = pad[1] = 0.f; }
	static float INFINITY_F() { return __builtin_inff(); }
	Point o;
	Vector d;
	float mint, maxt, time;
	u_int flags;
	float pad[2];
};

// 20-byte RayHit (ray.h:95-103)
class RayHit {
public:
	float t, b1, b2;
	u_int meshIndex, triangleIndex;
	void SetMiss() { meshIndex = 0xffffffffu; }
	bool Miss() const { return meshIndex == 0xffffffffu; }
};

static_assert(sizeof(Ray) == 48, "Ray is 48 bytes");
static_assert(sizeof(RayHit) == 20, "RayHit is 20 bytes");

struct Triangle {
	u_int v[3];
};

// row-major 4x4 (matrix4x4.h)
struct Matrix4x4 {
	float m[4][4];
	Matrix4x4() {
		for (int i = 0; i < 4; ++i)
			for (int j = 0; j < 4; ++j) m[i][j] = (i == j) ? 1.f : 0.f;
	}
	explicit Matrix4x4(const float mat[4][4]) {
		for (int i = 0; i < 4; ++i)
			for (int j = 0; j < 4; ++j) m[i][j] = mat[i][j];
	}
};

// transform.h: m and its inverse.  Built from m alone, the inverse is the
// reference's Gauss-Jordan inverse (matrix4x4.cpp:118-175), computed by the
// device library when the scene is built.
class Transform {
public:
	Transform() : hasInverse(true) {}
	explicit Transform(const Matrix4x4 &mat) : m(mat), hasInverse(false) {}
	Transform(const Matrix4x4 &mat, const Matrix4x4 &inv) : m(mat), mInv(inv), hasInverse(true) {}
	Matrix4x4 m, mInv;
	bool hasInverse;
};

// keyframed transforms; keys are global->local as MotionTriangleMesh stores them
class MotionSystem {
public:
	MotionSystem() {}
	MotionSystem(const std::vector<float> &t, const std::vector<Transform> &transforms) : times(t), keys(transforms) {
		if (times.empty() || times.size() != keys.size())
			throw std::runtime_error("MotionSystem: times and transforms differ in size");
	}
	std::vector<float> times;
	std::vector<Transform> keys;
};

typedef enum {
	TYPE_TRIANGLE, TYPE_TRIANGLE_INSTANCE, TYPE_TRIANGLE_MOTION,
	TYPE_EXT_TRIANGLE, TYPE_EXT_TRIANGLE_INSTANCE, TYPE_EXT_TRIANGLE_MOTION
} MeshType;

class Mesh {
public:
	virtual ~Mesh() {}
	virtual MeshType GetType() const = 0;
	virtual Point *GetVertices() const = 0;
	virtual Triangle *GetTriangles() const = 0;
	virtual u_int GetTotalVertexCount() const = 0;
	virtual u_int GetTotalTriangleCount() const = 0;
};

class TriangleMesh : public Mesh {
public:
	TriangleMesh(u_int nVerts, u_int nTris, Point *verts, Triangle *tris)
		: vertCount(nVerts), triCount(nTris), vertices(verts), tris(tris) {}
	MeshType GetType() const override { return TYPE_TRIANGLE; }
	Point *GetVertices() const override { return vertices; }
	Triangle *GetTriangles() const override { return tris; }
	u_int GetTotalVertexCount() const override { return vertCount; }
	u_int GetTotalTriangleCount() const override { return triCount; }

private:
	u_int vertCount, triCount;
	Point *vertices;
	Triangle *tris;
};

class InstanceTriangleMesh : public Mesh {
public:
	InstanceTriangleMesh(TriangleMesh *m, const Transform &t) : mesh(m), trans(t) {}
	MeshType GetType() const override { return TYPE_TRIANGLE_INSTANCE; }
	Point *GetVertices() const override { return mesh->GetVertices(); }
	Triangle *GetTriangles() const override { return mesh->GetTriangles(); }
	u_int GetTotalVertexCount() const override { return mesh->GetTotalVertexCount(); }
	u_int GetTotalTriangleCount() const override { return mesh->GetTotalTriangleCount(); }
	const Transform &GetTransformation() const { return trans; }
	void SetTransformation(const Transform &t) { trans = t; }
	TriangleMesh *GetTriangleMesh() const { return mesh; }

private:
	TriangleMesh *mesh;
	Transform trans;
};

class MotionTriangleMesh : public Mesh {
public:
	MotionTriangleMesh(TriangleMesh *m, const MotionSystem &ms) : mesh(m), motionSystem(ms) {}
	MeshType GetType() const override { return TYPE_TRIANGLE_MOTION; }
	Point *GetVertices() const override { return mesh->GetVertices(); }
	Triangle *GetTriangles() const override { return mesh->GetTriangles(); }
	u_int GetTotalVertexCount() const override { return mesh->GetTotalVertexCount(); }
	u_int GetTotalTriangleCount() const override { return mesh->GetTotalTriangleCount(); }
	const MotionSystem &GetMotionSystem() const { return motionSystem; }
	void SetMotionSystem(const MotionSystem &ms) { motionSystem = ms; }
	TriangleMesh *GetTriangleMesh() const { return mesh; }

private:
	TriangleMesh *mesh;
	MotionSystem motionSystem;
};

// Configuration properties: names (dotted keys, as in the reference's
// .cfg/.scn files) with text values and typed reads; a value read as a
// number or flag that does not parse as one throws.
class Properties {
public:
	Properties() {}
	Properties &Set(const std::string &name, const std::string &value);
	// "name = value" lines ('#' starts a comment); a line without '=' throws
	Properties &SetFromString(const std::string &text);
	bool IsDefined(const std::string &name) const { return values.count(name) != 0; }
	const std::vector<std::string> &GetAllNames() const { return names; }
	std::string Get(const std::string &name, const std::string &def) const;
	int GetInt(const std::string &name, int def) const;
	float GetFloat(const std::string &name, float def) const;
	bool GetBool(const std::string &name, bool def) const;

private:
	std::map<std::string, std::string> values;
	std::vector<std::string> names;  // first definition order
};

// device.h:39-53, plus the bit this device registers under.  The HIP GPU is
// both a hardware and an intersection device, so it is part of ALL,
// ALL_HARDWARE and ALL_INTERSECTION.
typedef enum {
	DEVICE_TYPE_NATIVE = 1 << 0,
	DEVICE_TYPE_OPENCL_DEFAULT = 1 << 1,
	DEVICE_TYPE_OPENCL_CPU = 1 << 2,
	DEVICE_TYPE_OPENCL_GPU = 1 << 3,
	DEVICE_TYPE_OPENCL_UNKNOWN = 1 << 4,
	DEVICE_TYPE_CUDA_GPU = 1 << 5,
	DEVICE_TYPE_HIP_GPU = 1 << 6,
	DEVICE_TYPE_OPENCL_ALL = DEVICE_TYPE_OPENCL_DEFAULT | DEVICE_TYPE_OPENCL_CPU | DEVICE_TYPE_OPENCL_GPU |
			DEVICE_TYPE_OPENCL_UNKNOWN,
	DEVICE_TYPE_CUDA_ALL = DEVICE_TYPE_CUDA_GPU,
	DEVICE_TYPE_HIP_ALL = DEVICE_TYPE_HIP_GPU,
	DEVICE_TYPE_ALL = DEVICE_TYPE_NATIVE | DEVICE_TYPE_OPENCL_ALL | DEVICE_TYPE_HIP_ALL,
	DEVICE_TYPE_ALL_HARDWARE = DEVICE_TYPE_OPENCL_ALL | DEVICE_TYPE_CUDA_ALL | DEVICE_TYPE_HIP_ALL,
	DEVICE_TYPE_ALL_INTERSECTION = DEVICE_TYPE_NATIVE | DEVICE_TYPE_OPENCL_ALL | DEVICE_TYPE_HIP_ALL,
} DeviceType;

// device.h:55-82
class DeviceDescription {
public:
	DeviceDescription(const std::string &deviceName, const DeviceType deviceType)
		: name(deviceName), type(deviceType), forceWorkGroupSize(0) {}
	virtual ~DeviceDescription() {}

	const std::string &GetName() const { return name; }
	const DeviceType GetType() const { return type; }
	virtual int GetComputeUnits() const { return 1; }
	virtual u_int GetNativeVectorWidthFloat() const { return 4; }
	virtual size_t GetMaxMemory() const { return ~(size_t)0; }
	virtual size_t GetMaxMemoryAllocSize() const { return ~(size_t)0; }
	virtual bool HasOutOfCoreMemorySupport() const { return false; }
	virtual u_int GetForceWorkGroupSize() const { return forceWorkGroupSize; }
	virtual void SetForceWorkGroupSize(const u_int size) { forceWorkGroupSize = size; }

	// device.cpp:28-52: keep the first GPU description, or none.  (The
	// reference's CPU fallback branch repeats the GPU test and is never
	// taken, so without a GPU the list ends empty; kept so.)
	static void FilterOne(std::vector<DeviceDescription *> &deviceDescriptions);
	// device.cpp:54-66: drop the descriptions with no bit of `type` (ALL keeps all)
	static void Filter(const DeviceType type, std::vector<DeviceDescription *> &deviceDescriptions);
	// device.cpp:68-89, plus "HIP_GPU"; other values throw
	static std::string GetDeviceType(const DeviceType type);

protected:
	std::string name;
	DeviceType type;
	u_int forceWorkGroupSize;
};

class Context;
class DataSet;
class Accelerator;

inline double WallClockTime() {
	return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// device.h:107-138
class Device {
public:
	const std::string &GetName() const { return deviceName; }
	const Context *GetContext() const { return deviceContext; }
	virtual const DeviceDescription *GetDeviceDesc() const = 0;
	const size_t GetDeviceIndex() const { return deviceIndex; }

	virtual bool IsRunning() const { return started; }

	// Set the thread current device
	virtual void PushThreadCurrentDevice() {}
	virtual void PopThreadCurrentDevice() {}

	friend class Context;

protected:
	Device() : deviceContext(nullptr), deviceIndex(0), started(false) {}
	Device(const Context *context, const size_t index) : deviceContext(context), deviceIndex(index), started(false) {}
	virtual ~Device() {}

	virtual void Start() { started = true; }
	virtual void Interrupt() {}
	virtual void Stop() { started = false; }

	const Context *deviceContext;
	size_t deviceIndex;

	std::string deviceName;

	bool started;
};

// hardwaredevice.h:31-56
class HardwareDeviceRange {
public:
	HardwareDeviceRange(const size_t s0) : dimensions(1) {
		sizes[0] = s0;
		sizes[1] = sizes[2] = 0;
	}
	HardwareDeviceRange(const size_t s0, const size_t s1) : dimensions(2) {
		sizes[0] = s0;
		sizes[1] = s1;
		sizes[2] = 0;
	}
	HardwareDeviceRange(const size_t s0, const size_t s1, const size_t s2) : dimensions(3) {
		sizes[0] = s0;
		sizes[1] = s1;
		sizes[2] = s2;
	}
	virtual ~HardwareDeviceRange() {}

	size_t sizes[3];
	u_int dimensions;
};

// hardwaredevice.h:61-70
class HardwareDeviceKernel {
public:
	virtual ~HardwareDeviceKernel() {}
	virtual bool IsNull() const = 0;

protected:
	HardwareDeviceKernel() {}
};

// hardwaredevice.h:76-85
class HardwareDeviceProgram {
public:
	virtual ~HardwareDeviceProgram() {}
	virtual bool IsNull() const = 0;

protected:
	HardwareDeviceProgram() {}
};

// hardwaredevice.h:90-95.  The HIP device accepts OUT_OF_CORE and allocates
// it in HBM like the others (288 GB per device; no managed memory here).
typedef enum {
	BUFFER_TYPE_NONE = 0,
	BUFFER_TYPE_READ_ONLY = 1 << 0,
	BUFFER_TYPE_READ_WRITE = 1 << 1,
	BUFFER_TYPE_OUT_OF_CORE = 1 << 2
} BufferType;

// hardwaredevice.h:97-105
class HardwareDeviceBuffer {
public:
	virtual ~HardwareDeviceBuffer() {}
	virtual bool IsNull() const = 0;
	virtual size_t GetSize() const = 0;

protected:
	HardwareDeviceBuffer() {}
};

// hardwaredevice.h:113-183
class HardwareDevice : virtual public Device {
public:
	void SetAdditionalCompileOpts(const std::vector<std::string> &opts) { additionalCompileOpts = opts; }
	const std::vector<std::string> &GetAdditionalCompileOpts() { return additionalCompileOpts; }

	virtual void CompileProgram(HardwareDeviceProgram **program, const std::vector<std::string> &programParameters,
			const std::string &programSource, const std::string &programName) = 0;
	virtual void GetKernel(HardwareDeviceProgram *program, HardwareDeviceKernel **kernel,
			const std::string &kernelName) = 0;
	virtual u_int GetKernelWorkGroupSize(HardwareDeviceKernel *kernel) = 0;

	virtual void SetKernelArg(HardwareDeviceKernel *kernel, const u_int index, const size_t size, const void *arg) = 0;

protected:
	virtual void SetKernelArgBuffer(HardwareDeviceKernel *kernel, const u_int index,
			const HardwareDeviceBuffer *buff) = 0;

public:
	template <typename T>
	void SetKernelArg(HardwareDeviceKernel *kernel, const u_int index, const T &arg) {
		SetKernelArg(kernel, index, KernelArgumentHandler<T>::Size(arg), KernelArgumentHandler<T>::Ptr(arg));
	}

	virtual void EnqueueKernel(HardwareDeviceKernel *kernel, const HardwareDeviceRange &globalSize,
			const HardwareDeviceRange &workGroupSize) = 0;
	virtual void EnqueueReadBuffer(const HardwareDeviceBuffer *buff, const bool blocking, const size_t size,
			void *ptr) = 0;
	virtual void EnqueueWriteBuffer(const HardwareDeviceBuffer *buff, const bool blocking, const size_t size,
			const void *ptr) = 0;
	virtual void FlushQueue() = 0;
	virtual void FinishQueue() = 0;

	size_t GetUsedMemory() const { return usedMemory; }

	virtual void AllocBuffer(HardwareDeviceBuffer **buff, const BufferType type, void *src, const size_t size,
			const std::string &desc = "") = 0;
	virtual void AllocBufferRO(HardwareDeviceBuffer **buff, void *src, const size_t size,
			const std::string &desc = "") {
		AllocBuffer(buff, BUFFER_TYPE_READ_ONLY, src, size, desc);
	}
	virtual void AllocBufferRW(HardwareDeviceBuffer **buff, void *src, const size_t size,
			const std::string &desc = "") {
		AllocBuffer(buff, BUFFER_TYPE_READ_WRITE, src, size, desc);
	}
	virtual void FreeBuffer(HardwareDeviceBuffer **buff) = 0;

protected:
	template <typename T>
	struct KernelArgumentHandler {
		static size_t Size(const T &) { return sizeof(T); }
		static const T *Ptr(const T &value) { return &value; }
	};

	HardwareDevice() : usedMemory(0) {}
	virtual ~HardwareDevice() {}

	void AllocMemory(const size_t s) { usedMemory += s; }
	void FreeMemory(const size_t s) { usedMemory -= s; }

	std::vector<std::string> additionalCompileOpts;
	size_t usedMemory;
};

typedef HardwareDeviceBuffer *HardwareDeviceBufferPtr;
// hardwaredevice.h:185-186: a buffer argument goes through SetKernelArgBuffer
template <>
inline void HardwareDevice::SetKernelArg<HardwareDeviceBufferPtr>(HardwareDeviceKernel *kernel, const u_int index,
		const HardwareDeviceBufferPtr &buff) {
	SetKernelArgBuffer(kernel, index, buff);
}

// accelerator.h:31-33
typedef enum { ACCEL_AUTO, ACCEL_BVH, ACCEL_MBVH, ACCEL_EMBREE, ACCEL_OPTIX } AcceleratorType;

class IntersectionDevice;
class HardwareIntersectionDevice;
class HardwareIntersectionKernel;

// accelerator.h:39-59
class Accelerator {
public:
	Accelerator() {}
	virtual ~Accelerator() {}

	virtual AcceleratorType GetType() const = 0;

	virtual bool HasNativeSupport(const IntersectionDevice &device) const = 0;
	virtual bool HasHWSupport(const IntersectionDevice &device) const = 0;

	virtual HardwareIntersectionKernel *NewHardwareIntersectionKernel(HardwareIntersectionDevice &device) const = 0;

	virtual void Init(const std::deque<const Mesh *> &meshes, const u_longlong totalVertexCount,
			const u_longlong totalTriangleCount) = 0;
	virtual bool DoesSupportUpdate() const { return false; }
	virtual void Update() { throw std::runtime_error("Internal error in Accelerator::Update()"); }

	virtual bool Intersect(const Ray *ray, RayHit *hit) const = 0;

	// src/luxrays/core/accelerator.cpp:25-55: "AUTO", "BVH", "MBVH", "EMBREE", "OPTIX"; others throw
	static std::string AcceleratorType2String(const AcceleratorType type);
	static AcceleratorType String2AcceleratorType(const std::string &type);
};

// intersectiondevice.h:31-80
class IntersectionDevice : virtual public Device {
public:
	// Returns true if it support HardwareDevice ray tracing
	virtual bool HasHWSupport() const { return false; }

	const Accelerator *GetAccelerator() const { return accel; }

	virtual double GetTotalRaysCount() const { return statsTotalSerialRayCount + statsTotalDataParallelRayCount; }
	virtual double GetTotalPerformance() const {
		const double dt = WallClockTime() - statsStartTime;
		return (dt == 0.0) ? 1.0 : ((statsTotalSerialRayCount + statsTotalDataParallelRayCount) / dt);
	}
	virtual u_longlong GetSerialPerformance() const {
		const double dt = WallClockTime() - statsStartTime;
		return (dt == 0.0) ? 1 : (u_longlong)(statsTotalSerialRayCount / dt);
	}
	virtual u_longlong GetDataParallelPerformance() const {
		const double dt = WallClockTime() - statsStartTime;
		return (dt == 0.0) ? 1 : (u_longlong)(statsTotalDataParallelRayCount / dt);
	}
	virtual void ResetPerformaceStats() {
		statsStartTime = WallClockTime();
		statsTotalSerialRayCount = 0;
		statsTotalDataParallelRayCount = 0;
	}

	// Serial interface: to trace a single ray
	virtual bool TraceRay(const Ray *ray, RayHit *rayHit) {
		statsTotalSerialRayCount += 1;
		return accel->Intersect(ray, rayHit);
	}

	friend class Context;

protected:
	IntersectionDevice()
		: dataSet(nullptr), accel(nullptr), statsStartTime(0.0), statsTotalSerialRayCount(0),
		  statsTotalDataParallelRayCount(0) {}
	virtual ~IntersectionDevice() {}

	virtual void SetDataSet(DataSet *newDataSet) {
		if (started) throw std::runtime_error("SetDataSet on a running device");
		dataSet = newDataSet;
	}
	void Start() override {
		if (!dataSet) throw std::runtime_error("Start of an intersection device without a DataSet");
		Device::Start();
		statsStartTime = WallClockTime();
		statsTotalSerialRayCount = 0;
		statsTotalDataParallelRayCount = 0;
	}

	DataSet *dataSet;
	const Accelerator *accel;
	double statsStartTime;
	u_longlong statsTotalSerialRayCount, statsTotalDataParallelRayCount;
};

// hardwareintersectiondevice.h:32-53
class HardwareIntersectionDevice : public IntersectionDevice, virtual public HardwareDevice {
public:
	// Returns true if it support HardwareDevice ray tracing
	virtual bool HasHWSupport() const override { return true; }

	// Data parallel interface: to trace a multiple rays
	virtual void EnqueueTraceRayBuffer(HardwareDeviceBuffer *rayBuff, HardwareDeviceBuffer *rayHitBuff,
			const unsigned int rayCount) {
		(void)rayBuff;
		(void)rayHitBuff;
		(void)rayCount;
		throw std::runtime_error("Called EnqueueTraceRayBuffer() on a device without parallel support");
	}

	friend class Context;

protected:
	virtual void Update() = 0;

	HardwareIntersectionDevice() {}
	virtual ~HardwareIntersectionDevice() {}
};

// hardwareintersectiondevice.h:59-72
class HardwareIntersectionKernel {
public:
	HardwareIntersectionKernel(HardwareIntersectionDevice &dev) : device(dev) {}
	virtual ~HardwareIntersectionKernel() {}

	virtual void Update(const DataSet *newDataSet) = 0;
	virtual void EnqueueTraceRayBuffer(HardwareDeviceBuffer *rayBuff, HardwareDeviceBuffer *rayHitBuff,
			const unsigned int rayCount) = 0;

protected:
	HardwareIntersectionDevice &device;
};

}  // namespace luxrays_hip

#endif
