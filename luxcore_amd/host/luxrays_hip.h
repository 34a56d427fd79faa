/*
 * luxrays_hip — the host side of the HIP intersection device, written as the
 * part of LuxRays' C++ interface this path uses, over the C ABI of
 * include/lrhip.h.  Class and method names, argument meaning and error
 * behaviour (std::runtime_error) follow the reference:
 *
 *   Ray / RayHit                 include/luxrays/core/geometry/ray.h:35-103
 *   Mesh, TriangleMesh, InstanceTriangleMesh, MotionTriangleMesh
 *                                include/luxrays/core/trianglemesh.h:42-391
 *   Transform, MotionSystem      include/luxrays/core/geometry/{transform,motionsystem}.h
 *   DataSet (Add, Preprocess, GetAccelerator, HasInstances, HasMotionBlur)
 *                                include/luxrays/core/dataset.h:44-70, src/luxrays/core/dataset.cpp
 *   Accelerator / BVHAccel / MBVHAccel
 *                                include/luxrays/core/accelerator.h:39-59
 *   HardwareDeviceBuffer, HardwareDevice buffers and queue
 *                                include/luxrays/core/hardwaredevice.h:90-183
 *   HardwareIntersectionDevice::EnqueueTraceRayBuffer, HardwareIntersectionKernel
 *                                include/luxrays/core/hardwareintersectiondevice.h:35-72
 *   IntersectionDevice::TraceRay + statistics
 *                                include/luxrays/core/intersectiondevice.h:35-85
 *   DeviceType, DeviceDescription (+ DEVICE_TYPE_HIP_GPU, HIPDeviceDescription)
 *                                include/luxrays/core/device.h:39-82, src/luxrays/core/device.cpp
 *   Context (descriptions, AddIntersectionDevices, SetDataSet, UpdateDataSet, Start/Stop)
 *                                include/luxrays/core/context.h, src/luxrays/core/context.cpp:44-330
 *
 * The namespace is luxrays_hip so the mirror can be linked next to a real
 * LuxRays build.  Meshes are owned by the application, as in the reference.
 */
#ifndef LUXRAYS_HIP_H
#define LUXRAYS_HIP_H

#include <deque>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/lrhip.h"

namespace luxrays_hip {

typedef unsigned int u_int;

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
	MeshType GetType() const { return TYPE_TRIANGLE; }
	Point *GetVertices() const { return vertices; }
	Triangle *GetTriangles() const { return tris; }
	u_int GetTotalVertexCount() const { return vertCount; }
	u_int GetTotalTriangleCount() const { return triCount; }

private:
	u_int vertCount, triCount;
	Point *vertices;
	Triangle *tris;
};

class InstanceTriangleMesh : public Mesh {
public:
	InstanceTriangleMesh(TriangleMesh *m, const Transform &t) : mesh(m), trans(t) {}
	MeshType GetType() const { return TYPE_TRIANGLE_INSTANCE; }
	Point *GetVertices() const { return mesh->GetVertices(); }
	Triangle *GetTriangles() const { return mesh->GetTriangles(); }
	u_int GetTotalVertexCount() const { return mesh->GetTotalVertexCount(); }
	u_int GetTotalTriangleCount() const { return mesh->GetTotalTriangleCount(); }
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
	MeshType GetType() const { return TYPE_TRIANGLE_MOTION; }
	Point *GetVertices() const { return mesh->GetVertices(); }
	Triangle *GetTriangles() const { return mesh->GetTriangles(); }
	u_int GetTotalVertexCount() const { return mesh->GetTotalVertexCount(); }
	u_int GetTotalTriangleCount() const { return mesh->GetTotalTriangleCount(); }
	const MotionSystem &GetMotionSystem() const { return motionSystem; }
	void SetMotionSystem(const MotionSystem &ms) { motionSystem = ms; }
	TriangleMesh *GetTriangleMesh() const { return mesh; }

private:
	TriangleMesh *mesh;
	MotionSystem motionSystem;
};

typedef enum { ACCEL_AUTO, ACCEL_BVH, ACCEL_MBVH } AcceleratorType;

// device.h:39-53 with the bit this device registers under.  The HIP GPU is
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
		: name(deviceName), type(deviceType) {}
	virtual ~DeviceDescription() {}

	const std::string &GetName() const { return name; }
	DeviceType GetType() const { return type; }
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
	u_int forceWorkGroupSize = 0;
};

// The HIP counterpart of CUDADeviceDescription (cudadevice.h:36-66,
// cudadevice.cpp:124-136): one description per visible GPU, in ordinal order.
class HIPDeviceDescription : public DeviceDescription {
public:
	HIPDeviceDescription(int ordinal, const std::string &deviceName, int computeUnits, size_t totalMemory)
		: DeviceDescription(deviceName, DEVICE_TYPE_HIP_GPU), ordinal(ordinal), cus(computeUnits),
		  memory(totalMemory) {}
	int GetOrdinal() const { return ordinal; }
	int GetComputeUnits() const override { return cus; }
	size_t GetMaxMemory() const override { return memory; }
	// one hipMalloc may take the whole of HBM
	size_t GetMaxMemoryAllocSize() const override { return memory; }
	// adds nothing when no GPU is visible (the C ABI reports no device)
	static void AddDeviceDescs(std::vector<DeviceDescription *> &descriptions);

private:
	int ordinal;
	int cus;
	size_t memory;
};

class HIPIntersectionDevice;
class HardwareIntersectionKernel;
class DataSet;

// accelerator.h:39-59: the device-side accelerator.  Init builds and uploads
// the scene; Intersect is the serial query (one ray traced on the device).
class Accelerator {
public:
	Accelerator(HIPIntersectionDevice &dev, AcceleratorType t) : device(dev), type(t) {}
	virtual ~Accelerator();
	AcceleratorType GetType() const { return type; }
	void Init(const std::deque<const Mesh *> &meshes, unsigned long long totalVertexCount,
			unsigned long long totalTriangleCount);
	bool DoesSupportUpdate() const { return type == ACCEL_MBVH; }
	// Update() rebuilds the root tree on the host as MBVHAccel::Update does
	// (UPDATE_REBUILD, the default) or refits it on the GPU (UPDATE_REFIT:
	// the same object boxes, the old grouping; lrh_scene_refit)
	enum UpdateMode { UPDATE_REBUILD, UPDATE_REFIT };
	void SetUpdateMode(UpdateMode m) { updateMode = m; }
	void Update();
	bool Intersect(const Ray *ray, RayHit *hit) const;
	HardwareIntersectionKernel *NewHardwareIntersectionKernel(HIPIntersectionDevice &device) const;
	lrh_scene *GetScene() const { return scene; }

private:
	void fillObjects(std::vector<lrh_object> &objs, std::vector<float> &store) const;
	HIPIntersectionDevice &device;
	UpdateMode updateMode = UPDATE_REBUILD;
	AcceleratorType type;
	std::deque<const Mesh *> meshes;
	std::vector<const TriangleMesh *> baseMeshes;
	std::vector<u_int> objMesh;
	lrh_scene *scene = nullptr;
};

// dataset.h: meshes in add order; mesh index = add order
class DataSet {
public:
	DataSet() {}
	void Add(const Mesh *mesh);
	void Preprocess() { preprocessed = true; }
	bool IsPreprocessed() const { return preprocessed; }
	bool HasInstances() const { return hasInstances; }
	bool HasMotionBlur() const { return hasMotion; }
	unsigned long long GetTotalVertexCount() const { return totalVertexCount; }
	unsigned long long GetTotalTriangleCount() const { return totalTriangleCount; }
	const std::deque<const Mesh *> &GetMeshes() const { return meshes; }
	// AUTO as the reference's hardware devices choose it (oclintersectiondevice.cpp:42-56)
	AcceleratorType ResolveType(AcceleratorType t) const {
		if (t != ACCEL_AUTO) return t;
		return (hasInstances || hasMotion) ? ACCEL_MBVH : ACCEL_BVH;
	}

private:
	std::deque<const Mesh *> meshes;
	bool preprocessed = false, hasInstances = false, hasMotion = false;
	unsigned long long totalVertexCount = 0, totalTriangleCount = 0;
};

// hardwaredevice.h:90-95.  OUT_OF_CORE is accepted and allocated in HBM like
// the others (288 GB per device; no managed memory on this path).
typedef enum {
	BUFFER_TYPE_NONE = 0,
	BUFFER_TYPE_READ_ONLY = 1 << 0,
	BUFFER_TYPE_READ_WRITE = 1 << 1,
	BUFFER_TYPE_OUT_OF_CORE = 1 << 2
} BufferType;

class HardwareDeviceBuffer {
public:
	virtual ~HardwareDeviceBuffer() {}
	bool IsNull() const { return ptr == nullptr; }
	size_t GetSize() const { return size; }
	void *GetPtr() const { return ptr; }

private:
	friend class HIPIntersectionDevice;
	void *ptr = nullptr;
	size_t size = 0;
};

// hardwareintersectiondevice.h:59-72
class HardwareIntersectionKernel {
public:
	HardwareIntersectionKernel(HIPIntersectionDevice &dev, const Accelerator &acc) : device(dev), accel(acc) {}
	virtual ~HardwareIntersectionKernel() {}
	virtual void Update(const DataSet *newDataSet);
	virtual void EnqueueTraceRayBuffer(HardwareDeviceBuffer *rayBuff, HardwareDeviceBuffer *rayHitBuff,
			const unsigned int rayCount);
	// extension: any-hit query (first hit with mint <= t < maxt)
	virtual void EnqueueTraceRayBufferAnyHit(HardwareDeviceBuffer *rayBuff, HardwareDeviceBuffer *rayHitBuff,
			const unsigned int rayCount);

protected:
	HIPIntersectionDevice &device;
	const Accelerator &accel;
};

// the HIP counterpart of OpenCLIntersectionDevice / CUDAIntersectionDevice
class HIPIntersectionDevice {
public:
	explicit HIPIntersectionDevice(int ordinal = 0);
	~HIPIntersectionDevice();

	const std::string &GetName() const { return name; }
	static int GetDeviceCount();

	// HardwareDevice buffers (hardwaredevice.h:145-167)
	void AllocBuffer(HardwareDeviceBuffer **buff, int type, void *src, size_t size, const std::string &desc = "");
	void AllocBufferRO(HardwareDeviceBuffer **buff, void *src, size_t size, const std::string &desc = "") {
		AllocBuffer(buff, BUFFER_TYPE_READ_ONLY, src, size, desc);
	}
	void AllocBufferRW(HardwareDeviceBuffer **buff, void *src, size_t size, const std::string &desc = "") {
		AllocBuffer(buff, BUFFER_TYPE_READ_WRITE, src, size, desc);
	}
	void FreeBuffer(HardwareDeviceBuffer **buff);
	void EnqueueReadBuffer(const HardwareDeviceBuffer *buff, bool blocking, size_t size, void *ptr);
	void EnqueueWriteBuffer(const HardwareDeviceBuffer *buff, bool blocking, size_t size, const void *ptr);
	void FlushQueue() {}
	void FinishQueue();
	size_t GetUsedMemory() const;

	// IntersectionDevice / HardwareIntersectionDevice
	void SetDataSet(DataSet *ds, AcceleratorType type = ACCEL_AUTO);
	void Start();
	void Stop();
	bool IsRunning() const { return running; }
	void Update();  // DataSet edit (instance matrices / motion keys)
	void EnqueueTraceRayBuffer(HardwareDeviceBuffer *rayBuff, HardwareDeviceBuffer *rayHitBuff, const unsigned int rayCount);
	void EnqueueTraceRayBufferAnyHit(HardwareDeviceBuffer *rayBuff, HardwareDeviceBuffer *rayHitBuff,
			const unsigned int rayCount);
	bool TraceRay(const Ray *ray, RayHit *rayHit);
	const Accelerator *GetAccelerator() const { return accel.get(); }

	double GetTotalRaysCount() const { return (double)(statsSerial + statsDataParallel); }
	unsigned long long GetSerialRayCount() const { return statsSerial; }
	unsigned long long GetDataParallelRayCount() const { return statsDataParallel; }
	void ResetPerformaceStats() { statsSerial = statsDataParallel = 0; }

	lrh_device *GetHandle() const { return dev; }
	// the description the Context created this device from (null when the
	// device was constructed directly from an ordinal)
	const DeviceDescription *GetDeviceDesc() const { return desc; }
	size_t GetDeviceIndex() const { return deviceIndex; }

private:
	friend class Context;
	lrh_device *dev = nullptr;
	const DeviceDescription *desc = nullptr;
	size_t deviceIndex = 0;
	std::string name;
	DataSet *dataSet = nullptr;
	AcceleratorType accelType = ACCEL_AUTO;
	std::unique_ptr<Accelerator> accel;
	std::unique_ptr<HardwareIntersectionKernel> kernel;
	bool running = false;
	unsigned long long statsSerial = 0, statsDataParallel = 0;
	friend class Accelerator;
	friend class HardwareIntersectionKernel;
};

typedef void (*LuxRaysDebugHandler)(const char *msg);

// context.h / context.cpp:44-330, reduced to the intersection-device part:
// the constructor enumerates the device descriptions, AddIntersectionDevices
// creates devices from (filtered) descriptions, SetDataSet / UpdateDataSet /
// Start / Stop drive every device.  Descriptions and devices are owned by the
// Context.  The reference's native-thread devices are not part of this
// library (the CPU path is LuxRays' own), so only HIP GPUs are enumerated.
class Context {
public:
	explicit Context(LuxRaysDebugHandler handler = nullptr, bool verbose = true);
	~Context();

	const std::vector<DeviceDescription *> &GetAvailableDeviceDescriptions() const { return deviceDescriptions; }
	const std::vector<HIPIntersectionDevice *> &GetIntersectionDevices() const { return idevices; }
	// context.cpp:250-291: a description of another type throws
	// "Unknown device type in Context::CreateIntersectionDevices(): <type>"
	std::vector<HIPIntersectionDevice *> AddIntersectionDevices(std::vector<DeviceDescription *> &deviceDesc);

	void SetDataSet(DataSet *dataSet, AcceleratorType type = ACCEL_AUTO);  // before Start()
	DataSet *GetCurrentDataSet() const { return currentDataSet; }
	void UpdateDataSet();  // after a DataSet edit, while started
	void Start();
	void Interrupt() {}
	void Stop();
	bool IsRunning() const { return started; }

	bool HasDebugHandler() const { return debugHandler != nullptr; }
	bool IsVerbose() const { return verbose; }
	void PrintDebugMsg(const std::string &msg) const {
		if (debugHandler && verbose) debugHandler(msg.c_str());
	}

private:
	LuxRaysDebugHandler debugHandler;
	bool verbose;
	bool started = false;
	DataSet *currentDataSet = nullptr;
	AcceleratorType accelType = ACCEL_AUTO;
	std::vector<DeviceDescription *> deviceDescriptions;
	std::vector<HIPIntersectionDevice *> idevices;
};

void CheckLrh(int rc);

}  // namespace luxrays_hip

#endif
