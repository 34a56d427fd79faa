/*
 * luxrays_hip — the host side of the HIP intersection device: LuxRays'
 * device, accelerator and kernel classes for this path, derived from the
 * reference's abstract interfaces (restated in luxrays_bases.h, every
 * overriding member marked `override`) over the C ABI of include/lrhip.h.
 * Class and method names, argument meaning and error behaviour
 * (std::runtime_error) follow the reference:
 *
 *   HIPIntersectionDevice         the counterpart of OpenCLIntersectionDevice / CUDAIntersectionDevice
 *                                 (src/luxrays/devices/oclintersectiondevice.cpp:30-95,
 *                                  cudadevice.cpp:243-524 for the buffers and queue)
 *   HIPDeviceDescription          the counterpart of CUDADeviceDescription (cudadevice.h:36-66)
 *   HIPAccelerator                BVHAccel / MBVHAccel as a device accelerator
 *                                 (src/luxrays/accelerators/{bvhaccel,mbvhaccel}.cpp)
 *   HIPIntersectionKernel         BVHKernel / MBVHKernel (bvhaccelhw.cpp:38-268, mbvhaccelhw.cpp:41-499)
 *   DataSet                       include/luxrays/core/dataset.h:40-110, src/luxrays/core/dataset.cpp:36-176
 *   Context                       include/luxrays/core/context.h, src/luxrays/core/context.cpp:44-330
 *
 * Configuration keys read (Context config, as the reference reads them):
 *   accelerator.type                  AUTO | BVH | MBVH (EMBREE, OPTIX: not on this device)   dataset.cpp:56
 *   accelerator.instances.enable      true                                                   dataset.cpp:57
 *   accelerator.motionblur.enable     true                                                   dataset.cpp:58
 *   accelerator.bvh.builder.type      EMBREE_BINNED_SAH (default) | CLASSIC | EMBREE_MORTON  bvhaccel.cpp:144-156
 *   accelerator.bvh.treetype, .costsamples, .isectcost, .travcost, .emptybonus          bvhaccel.cpp:49-70
 *   context.verbose                   true                                                   context.cpp:49
 * and the device's own:
 *   hip.bvh.layout                    QUANTIZED (64-byte nodes, default) | FP32 (128-byte nodes) |
 *                                     WIDE8 (128-byte 8-wide nodes)
 *   hip.bvh.maxleafsize               1..8 (default 4)
 *   hip.bvh.travcost                  SAH traversal cost per triangle test (default 1.0)
 *   hip.bvh.splitbudget               early split clipping, extra references per triangle (default 4; < 0 off)
 *   hip.kernel.blocksize / hip.kernel.ldsstack / hip.kernel.refill / hip.kernel.leaftris /
 *   hip.kernel.gridpercent            launch shape of the trace (defaults: the library's)
 *   hip.accelerator.update            REBUILD (default: MBVHAccel::Update's host rebuild) | REFIT (GPU refit)
 *
 * Meshes are owned by the application, accelerators by the DataSet, devices
 * and descriptions by the Context, as in the reference.
 */
#ifndef LUXRAYS_HIP_H
#define LUXRAYS_HIP_H

#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/lrhip.h"
#include "luxrays_bases.h"

namespace luxrays_hip {

void CheckLrh(int rc);

// One description per visible GPU, in ordinal order.
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

// A device allocation (lrh_alloc).
class HIPDeviceBuffer : public HardwareDeviceBuffer {
public:
	bool IsNull() const override { return ptr == nullptr; }
	size_t GetSize() const override { return size; }
	void *GetPtr() const { return ptr; }

private:
	friend class HIPIntersectionDevice;
	void *ptr = nullptr;
	size_t size = 0;
};

class HIPIntersectionDevice;
class HIPIntersectionKernel;

// BVHAccel / MBVHAccel as a device accelerator.  Init keeps the meshes (and,
// for MBVH, the distinct base meshes and each object's one); the tree is
// built and uploaded per device by NewHardwareIntersectionKernel, as the
// reference's BVHKernel / MBVHKernel constructors upload theirs.  The serial
// Intersect traces the ray on the device of the first live kernel (the
// library has no CPU traversal; the reference's CPU path is LuxRays' own).
class HIPAccelerator : public Accelerator {
public:
	HIPAccelerator(const Context *ctx, AcceleratorType t);
	~HIPAccelerator() override;

	AcceleratorType GetType() const override { return type; }
	bool HasNativeSupport(const IntersectionDevice &) const override { return false; }
	bool HasHWSupport(const IntersectionDevice &device) const override;
	HardwareIntersectionKernel *NewHardwareIntersectionKernel(HardwareIntersectionDevice &device) const override;
	void Init(const std::deque<const Mesh *> &meshes, const u_longlong totalVertexCount,
			const u_longlong totalTriangleCount) override;
	bool DoesSupportUpdate() const override { return type == ACCEL_MBVH; }
	// MBVHAccel::Update: checks the edit (same objects); each device's kernel
	// applies it in its Update (rebuild or GPU refit, hip.accelerator.update)
	void Update() override;
	bool Intersect(const Ray *ray, RayHit *hit) const override;

	// the device-side edit of an Update: UPDATE_REBUILD rebuilds the root tree
	// on the host as MBVHAccel::Update does (lrh_scene_update), UPDATE_REFIT
	// keeps its topology and refits it on the GPU (lrh_scene_refit)
	enum UpdateMode { UPDATE_REBUILD, UPDATE_REFIT };
	void SetUpdateMode(UpdateMode m) { updateMode = m; }
	UpdateMode GetUpdateMode() const { return updateMode; }
	const lrh_build_opts &GetBuildOptions() const { return buildOpts; }
	const lrh_trace_opts &GetTraceOptions() const { return traceOpts; }

private:
	friend class HIPIntersectionKernel;
	lrh_scene *build(lrh_device *dev) const;
	void fillObjects(std::vector<lrh_object> &objs, std::vector<float> &store) const;
	const Context *ctx;
	AcceleratorType type;
	UpdateMode updateMode = UPDATE_REBUILD;
	lrh_build_opts buildOpts{};
	lrh_trace_opts traceOpts{};
	std::deque<const Mesh *> meshes;
	std::vector<const TriangleMesh *> baseMeshes;
	std::vector<u_int> objMesh;
	bool initialized = false;
	mutable std::mutex kernelsMutex;
	mutable std::vector<HIPIntersectionKernel *> kernels;  // live kernels, oldest first
};

// BVHKernel / MBVHKernel: the accelerator's tree on one device.
class HIPIntersectionKernel : public HardwareIntersectionKernel {
public:
	HIPIntersectionKernel(HIPIntersectionDevice &dev, const HIPAccelerator &acc);
	~HIPIntersectionKernel() override;

	void Update(const DataSet *newDataSet) override;
	void EnqueueTraceRayBuffer(HardwareDeviceBuffer *rayBuff, HardwareDeviceBuffer *rayHitBuff,
			const unsigned int rayCount) override;
	// extension: any-hit query (the first hit with mint <= t < maxt)
	void EnqueueTraceRayBufferAnyHit(HardwareDeviceBuffer *rayBuff, HardwareDeviceBuffer *rayHitBuff,
			const unsigned int rayCount);
	lrh_scene *GetScene() const { return scene; }
	lrh_device *GetDeviceHandle() const { return devHandle; }

private:
	void trace(HardwareDeviceBuffer *rayBuff, HardwareDeviceBuffer *rayHitBuff, const unsigned int rayCount,
			unsigned int query);
	const HIPAccelerator &accel;
	lrh_device *devHandle = nullptr;
	lrh_scene *scene = nullptr;
};

// The HIP GPU as a LuxRays hardware intersection device: buffers and queue
// of HardwareDevice, the data-parallel trace of HardwareIntersectionDevice.
// Run-time compiled programs (CompileProgram, GetKernel, SetKernelArg,
// EnqueueKernel) are not offered: this device's kernels are built into
// liblrhip, and those calls throw.
class HIPIntersectionDevice : public HardwareIntersectionDevice {
public:
	HIPIntersectionDevice(const Context *context, HIPDeviceDescription *desc, const size_t devIndex);
	~HIPIntersectionDevice() override;

	const DeviceDescription *GetDeviceDesc() const override { return deviceDesc; }

	// HardwareDevice: programs and kernels (not supported)
	void CompileProgram(HardwareDeviceProgram **program, const std::vector<std::string> &programParameters,
			const std::string &programSource, const std::string &programName) override;
	void GetKernel(HardwareDeviceProgram *program, HardwareDeviceKernel **kernel,
			const std::string &kernelName) override;
	u_int GetKernelWorkGroupSize(HardwareDeviceKernel *kernel) override;
	void SetKernelArg(HardwareDeviceKernel *kernel, const u_int index, const size_t size, const void *arg) override;
	using HardwareDevice::SetKernelArg;
	void EnqueueKernel(HardwareDeviceKernel *kernel, const HardwareDeviceRange &globalSize,
			const HardwareDeviceRange &workGroupSize) override;

	// HardwareDevice: buffers and queue (hardwaredevice.h:145-167; semantics of cudadevice.cpp:448-524)
	void EnqueueReadBuffer(const HardwareDeviceBuffer *buff, const bool blocking, const size_t size,
			void *ptr) override;
	void EnqueueWriteBuffer(const HardwareDeviceBuffer *buff, const bool blocking, const size_t size,
			const void *ptr) override;
	void FlushQueue() override {}
	void FinishQueue() override;
	void AllocBuffer(HardwareDeviceBuffer **buff, const BufferType type, void *src, const size_t size,
			const std::string &desc = "") override;
	void FreeBuffer(HardwareDeviceBuffer **buff) override;

	// HardwareIntersectionDevice
	void EnqueueTraceRayBuffer(HardwareDeviceBuffer *rayBuff, HardwareDeviceBuffer *rayHitBuff,
			const unsigned int rayCount) override;
	// extension: any-hit query
	void EnqueueTraceRayBufferAnyHit(HardwareDeviceBuffer *rayBuff, HardwareDeviceBuffer *rayHitBuff,
			const unsigned int rayCount);

	lrh_device *GetHandle() const { return dev; }
	static int GetDeviceCount();

	friend class Context;

protected:
	void SetKernelArgBuffer(HardwareDeviceKernel *kernel, const u_int index,
			const HardwareDeviceBuffer *buff) override;
	// IntersectionDevice / HardwareIntersectionDevice life cycle (oclintersectiondevice.cpp:42-82)
	void SetDataSet(DataSet *newDataSet) override;
	void Start() override;
	void Stop() override;
	void Update() override;

private:
	friend class HIPIntersectionKernel;
	friend class HIPAccelerator;
	lrh_device *dev = nullptr;
	const HIPDeviceDescription *deviceDesc = nullptr;
	HIPIntersectionKernel *kernel = nullptr;
};

// dataset.h / dataset.cpp: meshes in add order (mesh index = add order) and
// the accelerators built over them, one per type, owned by the DataSet.
class DataSet {
public:
	explicit DataSet(const Context *luxRaysContext);
	~DataSet();

	AcceleratorType GetAcceleratorType() const { return accelType; }
	void SetAcceleratorType(AcceleratorType type) { accelType = type; }

	bool GetInstanceSupport() const { return enableInstanceSupport; }
	bool RequiresInstanceSupport() const { return enableInstanceSupport && hasInstances; }
	bool HasInstances() const { return hasInstances; }
	bool GetMotionBlurSupport() const { return hasMotionBlur; }
	bool RequiresMotionBlurSupport() const { return enableMotionBlurSupport && hasMotionBlur; }
	bool HasMotionBlur() const { return hasMotionBlur; }

	u_int Add(const Mesh *mesh);
	void Preprocess() { preprocessed = true; }
	bool IsPreprocessed() const { return preprocessed; }

	bool HasAccelerator(const AcceleratorType type) const;
	const Accelerator *GetAccelerator(const AcceleratorType type);
	bool DoesAllAcceleratorsSupportUpdate() const;
	void UpdateAccelerators();

	u_longlong GetTotalVertexCount() const { return totalVertexCount; }
	u_longlong GetTotalTriangleCount() const { return totalTriangleCount; }
	const std::deque<const Mesh *> &GetMeshes() const { return meshes; }
	const Context *GetContext() const { return context; }

private:
	const Context *context;
	std::deque<const Mesh *> meshes;
	u_longlong totalVertexCount = 0, totalTriangleCount = 0;
	bool preprocessed = false, hasInstances = false, hasMotionBlur = false;
	AcceleratorType accelType = ACCEL_AUTO;
	bool enableInstanceSupport = true, enableMotionBlurSupport = true;
	mutable std::mutex accelsMutex;
	std::map<AcceleratorType, std::unique_ptr<HIPAccelerator>> accels;
};

typedef void (*LuxRaysDebugHandler)(const char *msg);

// context.h / context.cpp:44-330, the intersection-device part: the
// constructor enumerates the device descriptions (HIP GPUs; the native-thread
// CPU devices are LuxRays' own), AddIntersectionDevices creates devices from
// (filtered) descriptions, SetDataSet / UpdateDataSet / Start / Stop drive
// every device.
class Context {
public:
	explicit Context(LuxRaysDebugHandler handler = nullptr, const Properties &config = Properties());
	~Context();

	const Properties &GetConfig() const { return cfg; }
	const std::vector<DeviceDescription *> &GetAvailableDeviceDescriptions() const { return deviceDescriptions; }
	const std::vector<IntersectionDevice *> &GetIntersectionDevices() const { return idevices; }
	// context.cpp:250-291: a description of another type throws
	// "Unknown device type in Context::CreateIntersectionDevices(): <type>"
	std::vector<IntersectionDevice *> AddIntersectionDevices(std::vector<DeviceDescription *> &deviceDesc);

	void SetDataSet(DataSet *dataSet);  // before Start()
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
	Properties cfg;
	bool verbose;
	bool started = false;
	DataSet *currentDataSet = nullptr;
	std::vector<DeviceDescription *> deviceDescriptions;
	std::vector<IntersectionDevice *> idevices;
};

}  // namespace luxrays_hip

#endif
