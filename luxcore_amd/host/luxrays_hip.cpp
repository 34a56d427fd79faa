/*
 * luxrays_hip — the reference's device / accelerator / kernel classes for
 * this path, over the C ABI of include/lrhip.h.  Every failing C call becomes
 * a std::runtime_error carrying lrh_last_error(), as the reference's
 * HardwareDevice calls throw on OpenCL/CUDA errors
 * (src/luxrays/devices/ocldevice.cpp, cudadevice.cpp CHECK_* macros).
 */
#include "luxrays_hip.h"

#include <cstring>
#include <sstream>

namespace luxrays_hip {

void CheckLrh(int rc) {
	if (rc != LRH_OK) throw std::runtime_error(std::string("lrhip: ") + lrh_last_error());
}

//------------------------------------------------------------------------------
// DataSet (src/luxrays/core/dataset.cpp: Add)
//------------------------------------------------------------------------------

void DataSet::Add(const Mesh *mesh) {
	if (!mesh) throw std::runtime_error("DataSet::Add: null mesh");
	if (preprocessed) throw std::runtime_error("DataSet::Add: the data set has already been preprocessed");
	meshes.push_back(mesh);
	totalVertexCount += mesh->GetTotalVertexCount();
	totalTriangleCount += mesh->GetTotalTriangleCount();
	switch (mesh->GetType()) {
		case TYPE_TRIANGLE_INSTANCE:
		case TYPE_EXT_TRIANGLE_INSTANCE:
			hasInstances = true;
			break;
		case TYPE_TRIANGLE_MOTION:
		case TYPE_EXT_TRIANGLE_MOTION:
			hasMotion = true;
			break;
		default:
			break;
	}
}

//------------------------------------------------------------------------------
// Accelerator: BVHAccel / MBVHAccel Init + the kernel upload
//------------------------------------------------------------------------------

static const TriangleMesh *baseOf(const Mesh *m) {
	switch (m->GetType()) {
		case TYPE_TRIANGLE:
		case TYPE_EXT_TRIANGLE:
			return static_cast<const TriangleMesh *>(m);
		case TYPE_TRIANGLE_INSTANCE:
		case TYPE_EXT_TRIANGLE_INSTANCE:
			return static_cast<const InstanceTriangleMesh *>(m)->GetTriangleMesh();
		case TYPE_TRIANGLE_MOTION:
		case TYPE_EXT_TRIANGLE_MOTION:
			return static_cast<const MotionTriangleMesh *>(m)->GetTriangleMesh();
	}
	throw std::runtime_error("Accelerator: unknown mesh type");
}

static lrh_mesh toLrh(const TriangleMesh *m) {
	lrh_mesh r;
	r.verts = reinterpret_cast<const float *>(m->GetVertices());
	r.nVerts = m->GetTotalVertexCount();
	r.tris = reinterpret_cast<const unsigned int *>(m->GetTriangles());
	r.nTris = m->GetTotalTriangleCount();
	return r;
}

Accelerator::~Accelerator() {
	if (scene) lrh_scene_destroy(scene);
}

void Accelerator::fillObjects(std::vector<lrh_object> &objs, std::vector<float> &store) const {
	// one pass to size the float store so pointers into it stay valid
	size_t need = 0;
	for (const Mesh *m : meshes) {
		if (m->GetType() == TYPE_TRIANGLE_INSTANCE || m->GetType() == TYPE_EXT_TRIANGLE_INSTANCE)
			need += 32;
		else if (m->GetType() == TYPE_TRIANGLE_MOTION || m->GetType() == TYPE_EXT_TRIANGLE_MOTION) {
			const MotionSystem &ms = static_cast<const MotionTriangleMesh *>(m)->GetMotionSystem();
			need += ms.times.size() * 17;
		}
	}
	store.assign(need, 0.f);
	size_t at = 0;
	objs.clear();
	for (size_t i = 0; i < meshes.size(); ++i) {
		const Mesh *m = meshes[i];
		lrh_object o;
		std::memset(&o, 0, sizeof(o));
		o.mesh = objMesh[i];
		const MeshType t = m->GetType();
		if (t == TYPE_TRIANGLE_INSTANCE || t == TYPE_EXT_TRIANGLE_INSTANCE) {
			const Transform &tr = static_cast<const InstanceTriangleMesh *>(m)->GetTransformation();
			o.kind = 1;
			std::memcpy(&store[at], &tr.m.m[0][0], 64);
			o.m = &store[at];
			if (tr.hasInverse) {
				std::memcpy(&store[at + 16], &tr.mInv.m[0][0], 64);
				o.mInv = &store[at + 16];
			}
			at += 32;
		} else if (t == TYPE_TRIANGLE_MOTION || t == TYPE_EXT_TRIANGLE_MOTION) {
			const MotionSystem &ms = static_cast<const MotionTriangleMesh *>(m)->GetMotionSystem();
			const size_t k = ms.times.size();
			o.kind = 2;
			o.nKeys = (unsigned int)k;
			std::memcpy(&store[at], ms.times.data(), k * sizeof(float));
			o.keyTimes = &store[at];
			for (size_t j = 0; j < k; ++j) std::memcpy(&store[at + k + 16 * j], &ms.keys[j].m.m[0][0], 64);
			o.keyMats = &store[at + k];
			at += 17 * k;
		} else
			o.kind = 0;
		objs.push_back(o);
	}
}

void Accelerator::Init(const std::deque<const Mesh *> &ms, unsigned long long, unsigned long long) {
	if (scene) throw std::runtime_error("Accelerator::Init called twice");
	meshes = ms;
	baseMeshes.clear();
	objMesh.clear();
	if (type == ACCEL_AUTO) throw std::runtime_error("Accelerator::Init: resolve ACCEL_AUTO first");
	std::vector<lrh_mesh> lm;
	if (type == ACCEL_BVH) {
		// BVHAccel flattens the triangles of every mesh (bvhaccel.cpp:72-168):
		// instances and motion meshes are not accepted
		for (const Mesh *m : meshes) {
			const MeshType t = m->GetType();
			if (t != TYPE_TRIANGLE && t != TYPE_EXT_TRIANGLE)
				throw std::runtime_error("BVHAccel supports only plain triangle meshes; use MBVH");
			lm.push_back(toLrh(static_cast<const TriangleMesh *>(m)));
		}
		CheckLrh(lrh_scene_create_bvh(device.dev, lm.data(), (unsigned int)lm.size(), nullptr, &scene));
	} else {
		// MBVHAccel: one leaf tree per distinct base mesh (mbvhaccel.cpp:90-150)
		for (const Mesh *m : meshes) {
			const TriangleMesh *b = baseOf(m);
			size_t k = 0;
			while (k < baseMeshes.size() && baseMeshes[k] != b) ++k;
			if (k == baseMeshes.size()) baseMeshes.push_back(b);
			objMesh.push_back((u_int)k);
		}
		for (const TriangleMesh *b : baseMeshes) lm.push_back(toLrh(b));
		std::vector<lrh_object> objs;
		std::vector<float> store;
		fillObjects(objs, store);
		CheckLrh(lrh_scene_create_mbvh(device.dev, lm.data(), (unsigned int)lm.size(), objs.data(),
				(unsigned int)objs.size(), nullptr, &scene));
	}
}

void Accelerator::Update() {
	if (!DoesSupportUpdate()) throw std::runtime_error("BVHAccel does not support Update()");
	if (!scene) throw std::runtime_error("Accelerator::Update before Init");
	std::vector<lrh_object> objs;
	std::vector<float> store;
	fillObjects(objs, store);
	if (updateMode == UPDATE_REFIT)
		CheckLrh(lrh_scene_refit(scene, objs.data(), (unsigned int)objs.size()));
	else
		CheckLrh(lrh_scene_update(scene, objs.data(), (unsigned int)objs.size()));
}

bool Accelerator::Intersect(const Ray *ray, RayHit *hit) const {
	// one ray through the same kernel: device staging, synchronous
	void *d = nullptr;
	CheckLrh(lrh_alloc(device.dev, sizeof(Ray) + sizeof(RayHit), &d));
	int rc = lrh_memcpy(device.dev, d, ray, sizeof(Ray), LRH_H2D, nullptr);
	if (rc == LRH_OK)
		rc = lrh_trace(scene, d, static_cast<char *>(d) + sizeof(Ray), 1, LRH_QUERY_CLOSEST, nullptr);
	if (rc == LRH_OK)
		rc = lrh_memcpy(device.dev, hit, static_cast<char *>(d) + sizeof(Ray), sizeof(RayHit), LRH_D2H, nullptr);
	lrh_free(device.dev, d);
	CheckLrh(rc);
	return !hit->Miss();
}

HardwareIntersectionKernel *Accelerator::NewHardwareIntersectionKernel(HIPIntersectionDevice &dev) const {
	return new HardwareIntersectionKernel(dev, *this);
}

//------------------------------------------------------------------------------
// HardwareIntersectionKernel (bvhaccelhw.cpp:239-268, mbvhaccelhw.cpp:442-496)
//------------------------------------------------------------------------------

void HardwareIntersectionKernel::Update(const DataSet *) {
	const_cast<Accelerator &>(accel).Update();
}

void HardwareIntersectionKernel::EnqueueTraceRayBuffer(HardwareDeviceBuffer *rayBuff,
		HardwareDeviceBuffer *rayHitBuff, const unsigned int rayCount) {
	if (rayBuff->GetSize() < (size_t)rayCount * sizeof(Ray) || rayHitBuff->GetSize() < (size_t)rayCount * sizeof(RayHit))
		throw std::runtime_error("EnqueueTraceRayBuffer: buffer smaller than rayCount records");
	CheckLrh(lrh_trace(accel.GetScene(), rayBuff->GetPtr(), rayHitBuff->GetPtr(), rayCount, LRH_QUERY_CLOSEST, nullptr));
}

void HardwareIntersectionKernel::EnqueueTraceRayBufferAnyHit(HardwareDeviceBuffer *rayBuff,
		HardwareDeviceBuffer *rayHitBuff, const unsigned int rayCount) {
	if (rayBuff->GetSize() < (size_t)rayCount * sizeof(Ray) || rayHitBuff->GetSize() < (size_t)rayCount * sizeof(RayHit))
		throw std::runtime_error("EnqueueTraceRayBufferAnyHit: buffer smaller than rayCount records");
	CheckLrh(lrh_trace(accel.GetScene(), rayBuff->GetPtr(), rayHitBuff->GetPtr(), rayCount, LRH_QUERY_ANY, nullptr));
}

//------------------------------------------------------------------------------
// HIPIntersectionDevice
//------------------------------------------------------------------------------

int HIPIntersectionDevice::GetDeviceCount() {
	int n = 0;
	if (lrh_device_count(&n) != LRH_OK) return 0;
	return n;
}

HIPIntersectionDevice::HIPIntersectionDevice(int ordinal) {
	CheckLrh(lrh_device_open(ordinal, &dev));
	char buf[256];
	int cus = 0;
	size_t mem = 0;
	CheckLrh(lrh_device_info(dev, buf, sizeof(buf), &cus, &mem));
	name = std::string("HIP-") + buf;
}

HIPIntersectionDevice::~HIPIntersectionDevice() {
	kernel.reset();
	accel.reset();
	if (dev) lrh_device_close(dev);
}

void HIPIntersectionDevice::AllocBuffer(HardwareDeviceBuffer **buff, int type, void *src, size_t size,
		const std::string &) {
	// cudadevice.cpp:448-520: a buffer of the same size is reused (content
	// updated from src); another size frees and reallocates; size 0 leaves a
	// null buffer
	if (!*buff) *buff = new HardwareDeviceBuffer();
	HardwareDeviceBuffer *b = *buff;
	if (b->ptr) {
		if (size == b->size) {
			if (src) EnqueueWriteBuffer(b, false, size, src);
			return;
		}
		lrh_free(dev, b->ptr);
		b->ptr = nullptr;
		b->size = 0;
	}
	if (!size) return;
	if (!(type & (BUFFER_TYPE_READ_ONLY | BUFFER_TYPE_READ_WRITE)))
		throw std::runtime_error("AllocBuffer: buffer type must be read-only or read-write");
	CheckLrh(lrh_alloc(dev, size, &b->ptr));
	b->size = size;
	if (src) EnqueueWriteBuffer(b, false, size, src);
}

void HIPIntersectionDevice::FreeBuffer(HardwareDeviceBuffer **buff) {
	if (!*buff) return;
	if ((*buff)->ptr) lrh_free(dev, (*buff)->ptr);
	delete *buff;
	*buff = nullptr;
}

void HIPIntersectionDevice::EnqueueReadBuffer(const HardwareDeviceBuffer *buff, bool blocking, size_t size, void *ptr) {
	if (size > buff->size) throw std::runtime_error("EnqueueReadBuffer: size beyond the buffer");
	CheckLrh(blocking ? lrh_memcpy(dev, ptr, buff->ptr, size, LRH_D2H, nullptr)
					  : lrh_memcpy_async(dev, ptr, buff->ptr, size, LRH_D2H, nullptr));
}

void HIPIntersectionDevice::EnqueueWriteBuffer(const HardwareDeviceBuffer *buff, bool blocking, size_t size,
		const void *ptr) {
	if (size > buff->size) throw std::runtime_error("EnqueueWriteBuffer: size beyond the buffer");
	CheckLrh(blocking ? lrh_memcpy(dev, buff->ptr, ptr, size, LRH_H2D, nullptr)
					  : lrh_memcpy_async(dev, buff->ptr, ptr, size, LRH_H2D, nullptr));
}

void HIPIntersectionDevice::FinishQueue() { CheckLrh(lrh_sync(dev, nullptr)); }

size_t HIPIntersectionDevice::GetUsedMemory() const { return lrh_used_memory(dev); }

void HIPIntersectionDevice::SetDataSet(DataSet *ds, AcceleratorType type) {
	if (running) throw std::runtime_error("SetDataSet on a running device");
	dataSet = ds;
	accelType = type;
}

void HIPIntersectionDevice::Start() {
	if (running) throw std::runtime_error("Start on a running device");
	if (!dataSet) throw std::runtime_error("Start: no DataSet set");
	if (!dataSet->IsPreprocessed()) dataSet->Preprocess();
	const AcceleratorType t = dataSet->ResolveType(accelType);
	accel.reset(new Accelerator(*this, t));
	accel->Init(dataSet->GetMeshes(), dataSet->GetTotalVertexCount(), dataSet->GetTotalTriangleCount());
	kernel.reset(accel->NewHardwareIntersectionKernel(*this));
	running = true;
}

void HIPIntersectionDevice::Stop() {
	if (!running) return;
	FinishQueue();
	kernel.reset();
	accel.reset();
	running = false;
}

void HIPIntersectionDevice::Update() {
	if (!running) throw std::runtime_error("Update on a stopped device");
	kernel->Update(dataSet);
}

void HIPIntersectionDevice::EnqueueTraceRayBuffer(HardwareDeviceBuffer *rayBuff, HardwareDeviceBuffer *rayHitBuff,
		const unsigned int rayCount) {
	if (!running) throw std::runtime_error("EnqueueTraceRayBuffer on a stopped device");
	kernel->EnqueueTraceRayBuffer(rayBuff, rayHitBuff, rayCount);
	statsDataParallel += rayCount;
}

void HIPIntersectionDevice::EnqueueTraceRayBufferAnyHit(HardwareDeviceBuffer *rayBuff,
		HardwareDeviceBuffer *rayHitBuff, const unsigned int rayCount) {
	if (!running) throw std::runtime_error("EnqueueTraceRayBufferAnyHit on a stopped device");
	kernel->EnqueueTraceRayBufferAnyHit(rayBuff, rayHitBuff, rayCount);
	statsDataParallel += rayCount;
}

bool HIPIntersectionDevice::TraceRay(const Ray *ray, RayHit *rayHit) {
	if (!running) throw std::runtime_error("TraceRay on a stopped device");
	statsSerial += 1;
	return accel->Intersect(ray, rayHit);
}

//------------------------------------------------------------------------------
// DeviceDescription (src/luxrays/core/device.cpp:28-89)
//------------------------------------------------------------------------------

void DeviceDescription::FilterOne(std::vector<DeviceDescription *> &deviceDescriptions) {
	for (DeviceDescription *d : deviceDescriptions) {
		if (d->GetType() & (DEVICE_TYPE_OPENCL_GPU | DEVICE_TYPE_HIP_GPU)) {
			deviceDescriptions.assign(1, d);
			return;
		}
	}
	deviceDescriptions.clear();
}

void DeviceDescription::Filter(const DeviceType type, std::vector<DeviceDescription *> &deviceDescriptions) {
	if (type == DEVICE_TYPE_ALL) return;
	std::vector<DeviceDescription *> kept;
	for (DeviceDescription *d : deviceDescriptions)
		if (d->GetType() & type) kept.push_back(d);
	deviceDescriptions.swap(kept);
}

std::string DeviceDescription::GetDeviceType(const DeviceType type) {
	switch (type) {
		case DEVICE_TYPE_ALL: return "ALL";
		case DEVICE_TYPE_NATIVE: return "NATIVE_THREAD";
		case DEVICE_TYPE_OPENCL_ALL: return "OPENCL_ALL";
		case DEVICE_TYPE_OPENCL_DEFAULT: return "OPENCL_DEFAULT";
		case DEVICE_TYPE_OPENCL_CPU: return "OPENCL_CPU";
		case DEVICE_TYPE_OPENCL_GPU: return "OPENCL_GPU";
		case DEVICE_TYPE_OPENCL_UNKNOWN: return "OPENCL_UNKNOWN";
		case DEVICE_TYPE_CUDA_GPU: return "CUDA_GPU";
		case DEVICE_TYPE_HIP_GPU: return "HIP_GPU";
		default:
			throw std::runtime_error("Unknown device type in DeviceDescription::GetDeviceType(): " +
					std::to_string((int)type));
	}
}

void HIPDeviceDescription::AddDeviceDescs(std::vector<DeviceDescription *> &descriptions) {
	const int n = HIPIntersectionDevice::GetDeviceCount();
	for (int i = 0; i < n; ++i) {
		lrh_device *d = nullptr;
		CheckLrh(lrh_device_open(i, &d));
		char buf[256];
		int cus = 0;
		size_t mem = 0;
		const int rc = lrh_device_info(d, buf, sizeof(buf), &cus, &mem);
		lrh_device_close(d);
		CheckLrh(rc);
		descriptions.push_back(new HIPDeviceDescription(i, std::string("HIP-") + buf, cus, mem));
	}
}

//------------------------------------------------------------------------------
// Context (src/luxrays/core/context.cpp:44-330)
//------------------------------------------------------------------------------

namespace {
template <class F>
void ctxLog(const Context *c, F &&f) {
	if (!c->HasDebugHandler() || !c->IsVerbose()) return;
	std::ostringstream ss;
	f(ss);
	c->PrintDebugMsg(ss.str());
}
}  // namespace

Context::Context(LuxRaysDebugHandler handler, bool verb) : debugHandler(handler), verbose(verb) {
	ctxLog(this, [](std::ostream &o) { o << "HIP support: enabled"; });
	HIPDeviceDescription::AddDeviceDescs(deviceDescriptions);
	ctxLog(this, [&](std::ostream &o) { o << "HIP device count: " << deviceDescriptions.size(); });
	for (size_t i = 0; i < deviceDescriptions.size(); ++i) {
		const DeviceDescription *d = deviceDescriptions[i];
		ctxLog(this, [&](std::ostream &o) { o << "Device " << i << " name: " << d->GetName(); });
		ctxLog(this, [&](std::ostream &o) {
			o << "Device " << i << " type: " << DeviceDescription::GetDeviceType(d->GetType());
		});
		ctxLog(this, [&](std::ostream &o) { o << "Device " << i << " compute units: " << d->GetComputeUnits(); });
		ctxLog(this, [&](std::ostream &o) {
			o << "Device " << i << " preferred float vector width: " << d->GetNativeVectorWidthFloat();
		});
		ctxLog(this, [&](std::ostream &o) {
			o << "Device " << i << " max allocable memory: " << d->GetMaxMemory() / (1024 * 1024) << "MBytes";
		});
		ctxLog(this, [&](std::ostream &o) {
			o << "Device " << i << " max allocable memory block size: " << d->GetMaxMemoryAllocSize() / (1024 * 1024)
			  << "MBytes";
		});
		ctxLog(this, [&](std::ostream &o) {
			o << "Device " << i << " has out of core memory support: " << d->HasOutOfCoreMemorySupport();
		});
	}
}

Context::~Context() {
	try {
		if (started) Stop();
	} catch (...) {
	}
	for (HIPIntersectionDevice *d : idevices) delete d;
	for (DeviceDescription *d : deviceDescriptions) delete d;
}

std::vector<HIPIntersectionDevice *> Context::AddIntersectionDevices(std::vector<DeviceDescription *> &deviceDesc) {
	if (started) throw std::runtime_error("Context::AddIntersectionDevices on a started context");
	ctxLog(this, [&](std::ostream &o) { o << "Creating " << deviceDesc.size() << " intersection device(s)"; });
	// check every description first, so a bad list creates nothing
	for (DeviceDescription *d : deviceDesc)
		if (!(d->GetType() & DEVICE_TYPE_HIP_ALL))
			throw std::runtime_error("Unknown device type in Context::CreateIntersectionDevices(): " +
					std::to_string((int)d->GetType()));
	std::vector<HIPIntersectionDevice *> created;
	try {
		for (size_t i = 0; i < deviceDesc.size(); ++i) {
			const HIPDeviceDescription *hd = static_cast<const HIPDeviceDescription *>(deviceDesc[i]);
			ctxLog(this, [&](std::ostream &o) {
				o << "Allocating intersection device " << i << ": " << hd->GetName()
				  << " (Type = " << DeviceDescription::GetDeviceType(hd->GetType()) << ")";
			});
			std::unique_ptr<HIPIntersectionDevice> dev(new HIPIntersectionDevice(hd->GetOrdinal()));
			dev->desc = hd;
			dev->deviceIndex = idevices.size() + created.size();
			if (currentDataSet) dev->SetDataSet(currentDataSet, accelType);
			created.push_back(dev.release());
		}
	} catch (...) {
		for (HIPIntersectionDevice *d : created) delete d;
		throw;
	}
	idevices.insert(idevices.end(), created.begin(), created.end());
	return created;
}

void Context::SetDataSet(DataSet *dataSet, AcceleratorType type) {
	if (started) throw std::runtime_error("Context::SetDataSet on a started context");
	currentDataSet = dataSet;
	accelType = type;
	for (HIPIntersectionDevice *d : idevices) d->SetDataSet(dataSet, type);
}

void Context::UpdateDataSet() {
	if (!started) throw std::runtime_error("Context::UpdateDataSet on a stopped context");
	for (HIPIntersectionDevice *d : idevices) d->Update();
}

void Context::Start() {
	if (started) throw std::runtime_error("Context::Start on a started context");
	for (HIPIntersectionDevice *d : idevices) d->Start();
	started = true;
}

void Context::Stop() {
	if (!started) throw std::runtime_error("Context::Stop on a stopped context");
	Interrupt();
	for (HIPIntersectionDevice *d : idevices) d->Stop();
	started = false;
}

}  // namespace luxrays_hip
