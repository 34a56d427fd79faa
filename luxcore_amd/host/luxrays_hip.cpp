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
// Properties
//------------------------------------------------------------------------------

namespace {
std::string trim(const std::string &s) {
	const size_t a = s.find_first_not_of(" \t\r\n");
	if (a == std::string::npos) return "";
	const size_t b = s.find_last_not_of(" \t\r\n");
	return s.substr(a, b - a + 1);
}
}  // namespace

Properties &Properties::Set(const std::string &name, const std::string &value) {
	if (!values.count(name)) names.push_back(name);
	values[name] = value;
	return *this;
}

Properties &Properties::SetFromString(const std::string &text) {
	std::istringstream in(text);
	std::string line;
	while (std::getline(in, line)) {
		const size_t hash = line.find('#');
		if (hash != std::string::npos) line = line.substr(0, hash);
		line = trim(line);
		if (line.empty()) continue;
		const size_t eq = line.find('=');
		if (eq == std::string::npos) throw std::runtime_error("Syntax error in a Properties line: " + line);
		std::string v = trim(line.substr(eq + 1));
		if (v.size() >= 2 && v.front() == '"' && v.back() == '"') v = v.substr(1, v.size() - 2);
		Set(trim(line.substr(0, eq)), v);
	}
	return *this;
}

std::string Properties::Get(const std::string &name, const std::string &def) const {
	auto it = values.find(name);
	return it == values.end() ? def : it->second;
}

int Properties::GetInt(const std::string &name, int def) const {
	auto it = values.find(name);
	if (it == values.end()) return def;
	char *end = nullptr;
	const long v = std::strtol(it->second.c_str(), &end, 10);
	if (end == it->second.c_str() || *end) throw std::runtime_error("Property " + name + " is not an integer: " + it->second);
	return (int)v;
}

float Properties::GetFloat(const std::string &name, float def) const {
	auto it = values.find(name);
	if (it == values.end()) return def;
	char *end = nullptr;
	const float v = std::strtof(it->second.c_str(), &end);
	if (end == it->second.c_str() || *end) throw std::runtime_error("Property " + name + " is not a number: " + it->second);
	return v;
}

bool Properties::GetBool(const std::string &name, bool def) const {
	auto it = values.find(name);
	if (it == values.end()) return def;
	const std::string &v = it->second;
	if (v == "1" || v == "true") return true;
	if (v == "0" || v == "false") return false;
	throw std::runtime_error("Property " + name + " is not a flag: " + v);
}

//------------------------------------------------------------------------------
// Accelerator type names (src/luxrays/core/accelerator.cpp:25-55)
//------------------------------------------------------------------------------

std::string Accelerator::AcceleratorType2String(const AcceleratorType type) {
	switch (type) {
		case ACCEL_AUTO: return "AUTO";
		case ACCEL_BVH: return "BVH";
		case ACCEL_MBVH: return "MBVH";
		case ACCEL_EMBREE: return "EMBREE";
		case ACCEL_OPTIX: return "OPTIX";
	}
	throw std::runtime_error("Unknown accelerator type in AcceleratorType2String(): " + std::to_string((int)type));
}

AcceleratorType Accelerator::String2AcceleratorType(const std::string &type) {
	if (type == "AUTO") return ACCEL_AUTO;
	if (type == "BVH") return ACCEL_BVH;
	if (type == "MBVH") return ACCEL_MBVH;
	if (type == "EMBREE") return ACCEL_EMBREE;
	if (type == "OPTIX") return ACCEL_OPTIX;
	throw std::runtime_error("Unknown accelerator type in String2AcceleratorType(): " + type);
}

//------------------------------------------------------------------------------
// DataSet (src/luxrays/core/dataset.cpp:39-176)
//------------------------------------------------------------------------------

DataSet::DataSet(const Context *luxRaysContext) : context(luxRaysContext) {
	if (context) {
		const Properties &cfg = context->GetConfig();
		accelType = Accelerator::String2AcceleratorType(cfg.Get("accelerator.type", "AUTO"));
		enableInstanceSupport = cfg.GetBool("accelerator.instances.enable", true);
		enableMotionBlurSupport = cfg.GetBool("accelerator.motionblur.enable", true);
	}
}

DataSet::~DataSet() {}

u_int DataSet::Add(const Mesh *mesh) {
	if (!mesh) throw std::runtime_error("DataSet::Add: null mesh");
	if (preprocessed) throw std::runtime_error("DataSet::Add: the data set has already been preprocessed");
	const u_int id = (u_int)meshes.size();
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
			hasMotionBlur = true;
			break;
		default:
			break;
	}
	return id;
}

bool DataSet::HasAccelerator(const AcceleratorType type) const {
	std::lock_guard<std::mutex> g(accelsMutex);
	return accels.count(type) != 0;
}

const Accelerator *DataSet::GetAccelerator(const AcceleratorType type) {
	std::lock_guard<std::mutex> g(accelsMutex);
	auto it = accels.find(type);
	if (it != accels.end()) return it->second.get();
	if (type != ACCEL_BVH && type != ACCEL_MBVH)
		throw std::runtime_error("The HIP intersection device supports the BVH and MBVH accelerators, not " +
				Accelerator::AcceleratorType2String(type));
	std::unique_ptr<HIPAccelerator> a(new HIPAccelerator(context, type));
	a->Init(meshes, totalVertexCount, totalTriangleCount);
	const Accelerator *r = a.get();
	accels[type] = std::move(a);
	return r;
}

bool DataSet::DoesAllAcceleratorsSupportUpdate() const {
	std::lock_guard<std::mutex> g(accelsMutex);
	for (const auto &a : accels)
		if (!a.second->DoesSupportUpdate()) return false;
	return true;
}

void DataSet::UpdateAccelerators() {
	std::lock_guard<std::mutex> g(accelsMutex);
	for (auto &a : accels) a.second->Update();
}

//------------------------------------------------------------------------------
// HIPAccelerator: BVHAccel / MBVHAccel Init + the per-device upload
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

HIPAccelerator::HIPAccelerator(const Context *c, AcceleratorType t) : ctx(c), type(t) {
	if (type != ACCEL_BVH && type != ACCEL_MBVH) throw std::runtime_error("HIPAccelerator: BVH or MBVH only");
	const Properties none;
	const Properties &cfg = ctx ? ctx->GetConfig() : none;
	// accelerator.bvh.* as BVHAccel::ToBVHParams reads them (bvhaccel.cpp:49-70):
	// the tree type is validated (the device's nodes are 4-wide whatever it
	// says: hits do not depend on it); the SAH costs apply to the CLASSIC
	// builder, as in the reference, whose Embree builders ignore them
	(void)cfg.GetInt("accelerator.bvh.treetype", 4);
	(void)cfg.GetInt("accelerator.bvh.costsamples", 0);
	const int isectCost = cfg.GetInt("accelerator.bvh.isectcost", 80);
	const int travCost = cfg.GetInt("accelerator.bvh.travcost", 10);
	(void)cfg.GetFloat("accelerator.bvh.emptybonus", 0.5f);
	const std::string builder = cfg.Get("accelerator.bvh.builder.type", "EMBREE_BINNED_SAH");
	buildOpts.builder = LRH_BUILD_HOST_SAH;
	if (builder == "CLASSIC") {
		if (isectCost <= 0 || travCost <= 0)
			throw std::runtime_error("accelerator.bvh.isectcost and .travcost must be positive");
		buildOpts.travCost = (float)travCost / (float)isectCost;
	} else if (builder == "EMBREE_MORTON") {
		// Morton-order build: the device LBVH (single-level scenes)
		if (type == ACCEL_BVH) buildOpts.builder = LRH_BUILD_DEVICE_LBVH;
	} else if (builder != "EMBREE_BINNED_SAH") {
		throw std::runtime_error("Unknown BVH builder type in BVHAccel::Init(): " + builder);
	}
	const std::string layout = cfg.Get("hip.bvh.layout", "QUANTIZED");
	if (layout == "QUANTIZED") buildOpts.layout = 0;
	else if (layout == "FP32") buildOpts.layout = 1;
	else if (layout == "WIDE8") buildOpts.layout = 2;  // both levels of an MBVH scene too
	else throw std::runtime_error("hip.bvh.layout must be QUANTIZED, FP32 or WIDE8: " + layout);
	const int leaf = cfg.GetInt("hip.bvh.maxleafsize", 4);
	if (leaf < 1 || leaf > 8) throw std::runtime_error("hip.bvh.maxleafsize must be 1..8");
	buildOpts.maxLeafSize = (unsigned)leaf;
	if (cfg.IsDefined("hip.bvh.travcost")) {
		const float tc = cfg.GetFloat("hip.bvh.travcost", 1.f);
		if (!(tc > 0.f)) throw std::runtime_error("hip.bvh.travcost must be positive");
		buildOpts.travCost = tc;
	}
	if (cfg.IsDefined("hip.bvh.splitbudget")) {
		const float sb = cfg.GetFloat("hip.bvh.splitbudget", 0.f);
		buildOpts.splitBudget = sb == 0.f ? -1.f : sb;  // 0 means off here; the C ABI's 0 is its default
	}
	traceOpts.blockSize = (unsigned)cfg.GetInt("hip.kernel.blocksize", 0);
	traceOpts.ldsStack = (unsigned)cfg.GetInt("hip.kernel.ldsstack", 0);
	traceOpts.refillThreshold = (unsigned)cfg.GetInt("hip.kernel.refill", 0);
	traceOpts.leafTris = (unsigned)cfg.GetInt("hip.kernel.leaftris", 0);
	traceOpts.gridPercent = (unsigned)cfg.GetInt("hip.kernel.gridpercent", 0);
	const std::string upd = cfg.Get("hip.accelerator.update", "REBUILD");
	if (upd == "REBUILD") updateMode = UPDATE_REBUILD;
	else if (upd == "REFIT") updateMode = UPDATE_REFIT;
	else throw std::runtime_error("hip.accelerator.update must be REBUILD or REFIT: " + upd);
}

HIPAccelerator::~HIPAccelerator() {}

bool HIPAccelerator::HasHWSupport(const IntersectionDevice &device) const {
	return dynamic_cast<const HIPIntersectionDevice *>(&device) != nullptr;
}

void HIPAccelerator::fillObjects(std::vector<lrh_object> &objs, std::vector<float> &store) const {
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

void HIPAccelerator::Init(const std::deque<const Mesh *> &ms, const u_longlong, const u_longlong) {
	if (initialized) throw std::runtime_error("Accelerator::Init called twice");
	meshes = ms;
	baseMeshes.clear();
	objMesh.clear();
	if (type == ACCEL_BVH) {
		// BVHAccel flattens the triangles of every mesh (bvhaccel.cpp:72-168):
		// instances and motion meshes are not accepted
		for (const Mesh *m : meshes) {
			const MeshType t = m->GetType();
			if (t != TYPE_TRIANGLE && t != TYPE_EXT_TRIANGLE)
				throw std::runtime_error("BVHAccel supports only plain triangle meshes; use MBVH");
		}
	} else {
		// MBVHAccel: one leaf tree per distinct base mesh (mbvhaccel.cpp:90-150)
		for (const Mesh *m : meshes) {
			const TriangleMesh *b = baseOf(m);
			size_t k = 0;
			while (k < baseMeshes.size() && baseMeshes[k] != b) ++k;
			if (k == baseMeshes.size()) baseMeshes.push_back(b);
			objMesh.push_back((u_int)k);
		}
	}
	initialized = true;
}

lrh_scene *HIPAccelerator::build(lrh_device *dev) const {
	if (!initialized) throw std::runtime_error("Accelerator used before Init");
	lrh_scene *scene = nullptr;
	std::vector<lrh_mesh> lm;
	if (type == ACCEL_BVH) {
		for (const Mesh *m : meshes) lm.push_back(toLrh(static_cast<const TriangleMesh *>(m)));
		CheckLrh(lrh_scene_create_bvh(dev, lm.data(), (unsigned int)lm.size(), &buildOpts, &scene));
	} else {
		for (const TriangleMesh *b : baseMeshes) lm.push_back(toLrh(b));
		std::vector<lrh_object> objs;
		std::vector<float> store;
		fillObjects(objs, store);
		CheckLrh(lrh_scene_create_mbvh(dev, lm.data(), (unsigned int)lm.size(), objs.data(), (unsigned int)objs.size(),
				&buildOpts, &scene));
	}
	return scene;
}

void HIPAccelerator::Update() {
	if (!DoesSupportUpdate()) throw std::runtime_error("BVHAccel does not support Update()");
	if (!initialized) throw std::runtime_error("Accelerator::Update before Init");
	// the objects must stay the same (mbvhaccel.cpp:237-250 rebuilds the root
	// over the same leaves); their matrices and keys are read by each kernel
	for (size_t i = 0; i < meshes.size(); ++i)
		if (meshes[i]->GetType() != TYPE_TRIANGLE && meshes[i]->GetType() != TYPE_EXT_TRIANGLE &&
				baseOf(meshes[i]) != baseMeshes[objMesh[i]])
			throw std::runtime_error("MBVHAccel::Update: an object changed its base mesh");
}

bool HIPAccelerator::Intersect(const Ray *ray, RayHit *hit) const {
	lrh_scene *scene = nullptr;
	lrh_device *dev = nullptr;
	{
		std::lock_guard<std::mutex> g(kernelsMutex);
		if (kernels.empty())
			throw std::runtime_error("HIPAccelerator::Intersect needs a started HIP intersection device");
		scene = kernels.front()->GetScene();
		dev = kernels.front()->GetDeviceHandle();
	}
	// one ray through the same kernel: device staging, synchronous
	void *d = nullptr;
	CheckLrh(lrh_alloc(dev, sizeof(Ray) + sizeof(RayHit), &d));
	int rc = lrh_memcpy(dev, d, ray, sizeof(Ray), LRH_H2D, nullptr);
	if (rc == LRH_OK) rc = lrh_trace(scene, d, static_cast<char *>(d) + sizeof(Ray), 1, LRH_QUERY_CLOSEST, nullptr);
	if (rc == LRH_OK) rc = lrh_memcpy(dev, hit, static_cast<char *>(d) + sizeof(Ray), sizeof(RayHit), LRH_D2H, nullptr);
	lrh_free(dev, d);
	CheckLrh(rc);
	return !hit->Miss();
}

HardwareIntersectionKernel *HIPAccelerator::NewHardwareIntersectionKernel(HardwareIntersectionDevice &device) const {
	HIPIntersectionDevice *hip = dynamic_cast<HIPIntersectionDevice *>(&device);
	if (!hip) throw std::runtime_error("HIPAccelerator::NewHardwareIntersectionKernel: not a HIP device");
	return new HIPIntersectionKernel(*hip, *this);
}

//------------------------------------------------------------------------------
// HIPIntersectionKernel (bvhaccelhw.cpp:239-268, mbvhaccelhw.cpp:442-496)
//------------------------------------------------------------------------------

HIPIntersectionKernel::HIPIntersectionKernel(HIPIntersectionDevice &dev, const HIPAccelerator &acc)
	: HardwareIntersectionKernel(dev), accel(acc), devHandle(dev.dev) {
	scene = accel.build(dev.dev);
	std::lock_guard<std::mutex> g(accel.kernelsMutex);
	accel.kernels.push_back(this);
}

HIPIntersectionKernel::~HIPIntersectionKernel() {
	{
		std::lock_guard<std::mutex> g(accel.kernelsMutex);
		for (size_t i = 0; i < accel.kernels.size(); ++i)
			if (accel.kernels[i] == this) {
				accel.kernels.erase(accel.kernels.begin() + (long)i);
				break;
			}
	}
	if (scene) lrh_scene_destroy(scene);
}

void HIPIntersectionKernel::Update(const DataSet *) {
	if (!accel.DoesSupportUpdate()) throw std::runtime_error("BVHKernel does not support Update()");
	std::vector<lrh_object> objs;
	std::vector<float> store;
	accel.fillObjects(objs, store);
	if (accel.GetUpdateMode() == HIPAccelerator::UPDATE_REFIT)
		CheckLrh(lrh_scene_refit(scene, objs.data(), (unsigned int)objs.size()));
	else
		CheckLrh(lrh_scene_update(scene, objs.data(), (unsigned int)objs.size()));
}

void HIPIntersectionKernel::trace(HardwareDeviceBuffer *rayBuff, HardwareDeviceBuffer *rayHitBuff,
		const unsigned int rayCount, unsigned int query) {
	const HIPDeviceBuffer *rb = dynamic_cast<const HIPDeviceBuffer *>(rayBuff);
	const HIPDeviceBuffer *hb = dynamic_cast<const HIPDeviceBuffer *>(rayHitBuff);
	if (!rb || !hb) throw std::runtime_error("EnqueueTraceRayBuffer: buffers not allocated by a HIP device");
	if (rb->GetSize() < (size_t)rayCount * sizeof(Ray) || hb->GetSize() < (size_t)rayCount * sizeof(RayHit))
		throw std::runtime_error("EnqueueTraceRayBuffer: buffer smaller than rayCount records");
	const lrh_trace_opts &o = accel.GetTraceOptions();
	CheckLrh(lrh_trace_ex(scene, rb->GetPtr(), hb->GetPtr(), rayCount, query, &o, nullptr));
}

void HIPIntersectionKernel::EnqueueTraceRayBuffer(HardwareDeviceBuffer *rayBuff, HardwareDeviceBuffer *rayHitBuff,
		const unsigned int rayCount) {
	trace(rayBuff, rayHitBuff, rayCount, LRH_QUERY_CLOSEST);
}

void HIPIntersectionKernel::EnqueueTraceRayBufferAnyHit(HardwareDeviceBuffer *rayBuff, HardwareDeviceBuffer *rayHitBuff,
		const unsigned int rayCount) {
	trace(rayBuff, rayHitBuff, rayCount, LRH_QUERY_ANY);
}

//------------------------------------------------------------------------------
// HIPIntersectionDevice
//------------------------------------------------------------------------------

int HIPIntersectionDevice::GetDeviceCount() {
	int n = 0;
	if (lrh_device_count(&n) != LRH_OK) return 0;
	return n;
}

HIPIntersectionDevice::HIPIntersectionDevice(const Context *context, HIPDeviceDescription *desc, const size_t devIndex)
	: Device(context, devIndex), deviceDesc(desc) {
	if (!desc) throw std::runtime_error("HIPIntersectionDevice: null description");
	CheckLrh(lrh_device_open(desc->GetOrdinal(), &dev));
	deviceName = desc->GetName() + " Intersect";
}

HIPIntersectionDevice::~HIPIntersectionDevice() {
	try {
		if (started) Stop();
	} catch (...) {
	}
	if (dev) lrh_device_close(dev);
}

static void noPrograms(const char *what) {
	throw std::runtime_error(std::string("HIPIntersectionDevice::") + what +
			": run-time compiled programs are not supported (the HIP device's kernels are built into liblrhip)");
}

void HIPIntersectionDevice::CompileProgram(HardwareDeviceProgram **, const std::vector<std::string> &,
		const std::string &, const std::string &) {
	noPrograms("CompileProgram");
}
void HIPIntersectionDevice::GetKernel(HardwareDeviceProgram *, HardwareDeviceKernel **, const std::string &) {
	noPrograms("GetKernel");
}
u_int HIPIntersectionDevice::GetKernelWorkGroupSize(HardwareDeviceKernel *) {
	noPrograms("GetKernelWorkGroupSize");
	return 0;
}
void HIPIntersectionDevice::SetKernelArg(HardwareDeviceKernel *, const u_int, const size_t, const void *) {
	noPrograms("SetKernelArg");
}
void HIPIntersectionDevice::SetKernelArgBuffer(HardwareDeviceKernel *, const u_int, const HardwareDeviceBuffer *) {
	noPrograms("SetKernelArgBuffer");
}
void HIPIntersectionDevice::EnqueueKernel(HardwareDeviceKernel *, const HardwareDeviceRange &,
		const HardwareDeviceRange &) {
	noPrograms("EnqueueKernel");
}

void HIPIntersectionDevice::AllocBuffer(HardwareDeviceBuffer **buff, const BufferType type, void *src,
		const size_t size, const std::string &) {
	// cudadevice.cpp:448-520: a buffer of the same size is reused (content
	// updated from src); another size frees and reallocates; size 0 leaves a
	// null buffer
	if (*buff && !dynamic_cast<HIPDeviceBuffer *>(*buff))
		throw std::runtime_error("AllocBuffer: buffer not allocated by a HIP device");
	if (!(type & (BUFFER_TYPE_READ_ONLY | BUFFER_TYPE_READ_WRITE)))
		throw std::runtime_error("AllocBuffer: buffer type must be read-only or read-write");
	if (!*buff) *buff = new HIPDeviceBuffer();
	HIPDeviceBuffer *b = static_cast<HIPDeviceBuffer *>(*buff);
	if (b->ptr) {
		if (size == b->size) {
			if (src) EnqueueWriteBuffer(b, false, size, src);
			return;
		}
		lrh_free(dev, b->ptr);
		FreeMemory(b->size);
		b->ptr = nullptr;
		b->size = 0;
	}
	if (!size) return;
	CheckLrh(lrh_alloc(dev, size, &b->ptr));
	b->size = size;
	AllocMemory(size);
	if (src) EnqueueWriteBuffer(b, false, size, src);
}

void HIPIntersectionDevice::FreeBuffer(HardwareDeviceBuffer **buff) {
	if (!*buff) return;
	HIPDeviceBuffer *b = dynamic_cast<HIPDeviceBuffer *>(*buff);
	if (!b) throw std::runtime_error("FreeBuffer: buffer not allocated by a HIP device");
	if (b->ptr) {
		lrh_free(dev, b->ptr);
		FreeMemory(b->size);
	}
	delete b;
	*buff = nullptr;
}

void HIPIntersectionDevice::EnqueueReadBuffer(const HardwareDeviceBuffer *buff, const bool blocking, const size_t size,
		void *ptr) {
	const HIPDeviceBuffer *b = dynamic_cast<const HIPDeviceBuffer *>(buff);
	if (!b) throw std::runtime_error("EnqueueReadBuffer: buffer not allocated by a HIP device");
	if (size > b->size) throw std::runtime_error("EnqueueReadBuffer: size beyond the buffer");
	CheckLrh(blocking ? lrh_memcpy(dev, ptr, b->ptr, size, LRH_D2H, nullptr)
					  : lrh_memcpy_async(dev, ptr, b->ptr, size, LRH_D2H, nullptr));
}

void HIPIntersectionDevice::EnqueueWriteBuffer(const HardwareDeviceBuffer *buff, const bool blocking, const size_t size,
		const void *ptr) {
	const HIPDeviceBuffer *b = dynamic_cast<const HIPDeviceBuffer *>(buff);
	if (!b) throw std::runtime_error("EnqueueWriteBuffer: buffer not allocated by a HIP device");
	if (size > b->size) throw std::runtime_error("EnqueueWriteBuffer: size beyond the buffer");
	CheckLrh(blocking ? lrh_memcpy(dev, b->ptr, ptr, size, LRH_H2D, nullptr)
					  : lrh_memcpy_async(dev, b->ptr, ptr, size, LRH_H2D, nullptr));
}

void HIPIntersectionDevice::FinishQueue() { CheckLrh(lrh_sync(dev, nullptr)); }

void HIPIntersectionDevice::SetDataSet(DataSet *newDataSet) {
	IntersectionDevice::SetDataSet(newDataSet);
	accel = nullptr;
	if (dataSet) {
		// oclintersectiondevice.cpp:42-56: AUTO is MBVH with instances or motion, else BVH
		const AcceleratorType t = dataSet->GetAcceleratorType();
		if (t != ACCEL_AUTO)
			accel = dataSet->GetAccelerator(t);
		else if (dataSet->RequiresInstanceSupport() || dataSet->RequiresMotionBlurSupport())
			accel = dataSet->GetAccelerator(ACCEL_MBVH);
		else
			accel = dataSet->GetAccelerator(ACCEL_BVH);
	}
}

void HIPIntersectionDevice::Start() {
	if (started) throw std::runtime_error("Start on a running device");
	if (!accel) throw std::runtime_error("Start of an intersection device without a DataSet");
	IntersectionDevice::Start();
	try {
		kernel = static_cast<HIPIntersectionKernel *>(accel->NewHardwareIntersectionKernel(*this));
	} catch (...) {
		Device::Stop();
		throw;
	}
}

void HIPIntersectionDevice::Stop() {
	if (!started) return;
	FinishQueue();
	delete kernel;
	kernel = nullptr;
	Device::Stop();
}

void HIPIntersectionDevice::Update() {
	if (!started) throw std::runtime_error("Update on a stopped device");
	kernel->Update(dataSet);
}

void HIPIntersectionDevice::EnqueueTraceRayBuffer(HardwareDeviceBuffer *rayBuff, HardwareDeviceBuffer *rayHitBuff,
		const unsigned int rayCount) {
	if (!started) throw std::runtime_error("EnqueueTraceRayBuffer on a stopped device");
	kernel->EnqueueTraceRayBuffer(rayBuff, rayHitBuff, rayCount);
	statsTotalDataParallelRayCount += rayCount;
}

void HIPIntersectionDevice::EnqueueTraceRayBufferAnyHit(HardwareDeviceBuffer *rayBuff,
		HardwareDeviceBuffer *rayHitBuff, const unsigned int rayCount) {
	if (!started) throw std::runtime_error("EnqueueTraceRayBufferAnyHit on a stopped device");
	kernel->EnqueueTraceRayBufferAnyHit(rayBuff, rayHitBuff, rayCount);
	statsTotalDataParallelRayCount += rayCount;
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

Context::Context(LuxRaysDebugHandler handler, const Properties &config)
	: debugHandler(handler), cfg(config), verbose(config.GetBool("context.verbose", true)) {
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
	for (IntersectionDevice *d : idevices) delete d;
	for (DeviceDescription *d : deviceDescriptions) delete d;
}

std::vector<IntersectionDevice *> Context::AddIntersectionDevices(std::vector<DeviceDescription *> &deviceDesc) {
	if (started) throw std::runtime_error("Context::AddIntersectionDevices on a started context");
	ctxLog(this, [&](std::ostream &o) { o << "Creating " << deviceDesc.size() << " intersection device(s)"; });
	// check every description first, so a bad list creates nothing
	for (DeviceDescription *d : deviceDesc)
		if (!(d->GetType() & DEVICE_TYPE_HIP_ALL))
			throw std::runtime_error("Unknown device type in Context::CreateIntersectionDevices(): " +
					std::to_string((int)d->GetType()));
	for (DeviceDescription *d : deviceDesc)
		if (!dynamic_cast<HIPDeviceDescription *>(d))
			throw std::runtime_error("Context::CreateIntersectionDevices(): " + d->GetName() +
					" is not a description this Context enumerated");
	std::vector<IntersectionDevice *> created;
	try {
		for (size_t i = 0; i < deviceDesc.size(); ++i) {
			HIPDeviceDescription *hd = static_cast<HIPDeviceDescription *>(deviceDesc[i]);
			ctxLog(this, [&](std::ostream &o) {
				o << "Allocating intersection device " << i << ": " << hd->GetName()
				  << " (Type = " << DeviceDescription::GetDeviceType(hd->GetType()) << ")";
			});
			std::unique_ptr<HIPIntersectionDevice> dev(new HIPIntersectionDevice(this, hd, idevices.size() + created.size()));
			if (currentDataSet) dev->SetDataSet(currentDataSet);
			created.push_back(dev.release());
		}
	} catch (...) {
		for (IntersectionDevice *d : created) delete d;
		throw;
	}
	idevices.insert(idevices.end(), created.begin(), created.end());
	return created;
}

void Context::SetDataSet(DataSet *dataSet) {
	if (started) throw std::runtime_error("Context::SetDataSet on a started context");
	currentDataSet = dataSet;
	for (IntersectionDevice *d : idevices) d->SetDataSet(dataSet);
}

void Context::UpdateDataSet() {
	// context.cpp:182-196: the accelerators first, then every hardware device
	if (!started) throw std::runtime_error("Context::UpdateDataSet on a stopped context");
	currentDataSet->UpdateAccelerators();
	for (IntersectionDevice *d : idevices) {
		HardwareIntersectionDevice *hw = dynamic_cast<HardwareIntersectionDevice *>(d);
		if (hw) hw->Update();
	}
}

void Context::Start() {
	if (started) throw std::runtime_error("Context::Start on a started context");
	for (IntersectionDevice *d : idevices) d->Start();
	started = true;
}

void Context::Stop() {
	if (!started) throw std::runtime_error("Context::Stop on a stopped context");
	Interrupt();
	for (IntersectionDevice *d : idevices) d->Stop();
	started = false;
}

}  // namespace luxrays_hip
