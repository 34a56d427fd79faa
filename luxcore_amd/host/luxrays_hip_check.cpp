/*
 * luxrays_hip_check — drives the C++ device classes the way an application of
 * the reference drives an intersection device (DataSet::Add for every mesh,
 * SetDataSet/Start, AllocBufferRO/RW, EnqueueTraceRayBuffer, EnqueueReadBuffer,
 * FinishQueue, TraceRay, Update after a transform edit) on a scene read from
 * files, and writes the hit records.  The device comes from a Context, as in
 * an application: the HIP GPU descriptions, filtered to one, then
 * AddIntersectionDevices, SetDataSet, Start, UpdateDataSet, Stop.
 *
 *   luxrays_hip_check <dir> [--accel auto|bvh|mbvh] [--anyhit] [--serial N] [--scene FILE] [--refit]
 *   luxrays_hip_check --devices     (print the Context's device descriptions)
 *   luxrays_hip_check --registry    (DeviceDescription::Filter / FilterOne /
 *                                    GetDeviceType and the Context's type check
 *                                    on synthetic descriptions; no GPU needed)
 *
 * --refit            Update() refits the root tree on the GPU instead of
 *                    rebuilding it on the host
 * --scene FILE       meshes and objects from a LuxCore scene file (.scn/.cfg,
 *                    PLY meshes) through libluxrays_ingest instead of the two
 *                    files below
 * <dir>/meshes.bin   u32 nMeshes; per mesh u32 nVerts, u32 nTris,
 *                    f32 verts[3 nVerts], u32 tris[3 nTris]
 * <dir>/objects.bin  optional; u32 nObjects; per object u32 mesh, u32 kind
 *                    (0 mesh, 1 instance, 2 motion), f32 m[16], u32 nKeys,
 *                    f32 times[nKeys], f32 mats[16 nKeys]
 *                    Without it every mesh is one object of kind 0.
 * <dir>/objects2.bin optional; same objects with new matrices / keys: traced
 *                    again after the edit and Update()
 * <dir>/rays.bin     Ray[n] (48-byte records)
 * writes <dir>/hits.bin (RayHit[n]), hits2.bin after an update, and
 * serial.bin (RayHit[N] of TraceRay on the first N rays).
 */
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <iterator>

#include "luxrays_hip.h"
#include "../../include/lrh_ingest.h"

using namespace luxrays_hip;

namespace {

std::vector<char> readAll(const std::string &path, bool required) {
	std::ifstream f(path, std::ios::binary);
	if (!f) {
		if (required) throw std::runtime_error("cannot open " + path);
		return {};
	}
	return std::vector<char>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

void writeAll(const std::string &path, const void *p, size_t n) {
	std::ofstream f(path, std::ios::binary);
	f.write(static_cast<const char *>(p), (std::streamsize)n);
	if (!f) throw std::runtime_error("cannot write " + path);
}

struct Reader {
	const std::vector<char> &b;
	size_t at = 0;
	explicit Reader(const std::vector<char> &buf) : b(buf) {}
	template <typename T>
	void get(T *out, size_t n) {
		if (at + n * sizeof(T) > b.size()) throw std::runtime_error("truncated input file");
		std::memcpy(out, b.data() + at, n * sizeof(T));
		at += n * sizeof(T);
	}
	u_int u() {
		u_int v;
		get(&v, 1);
		return v;
	}
};

struct ObjDesc {
	u_int mesh, kind;
	float m[16];
	std::vector<float> times, mats;
};

std::vector<ObjDesc> readObjects(const std::vector<char> &buf) {
	Reader r(buf);
	std::vector<ObjDesc> out(r.u());
	for (ObjDesc &o : out) {
		o.mesh = r.u();
		o.kind = r.u();
		r.get(o.m, 16);
		const u_int k = r.u();
		o.times.resize(k);
		o.mats.resize(16 * (size_t)k);
		r.get(o.times.data(), k);
		r.get(o.mats.data(), 16 * (size_t)k);
	}
	return out;
}

Matrix4x4 mat(const float *p) {
	float a[4][4];
	std::memcpy(a, p, 64);
	return Matrix4x4(a);
}

MotionSystem motion(const ObjDesc &o) {
	std::vector<Transform> keys;
	for (size_t j = 0; j < o.times.size(); ++j) keys.push_back(Transform(mat(&o.mats[16 * j])));
	return MotionSystem(o.times, keys);
}

void logMsg(const char *msg) { std::cout << "[LuxRays] " << msg << std::endl; }

int listDevices() {
	try {
		Context ctx(logMsg, Properties().Set("context.verbose", "true"));
		const std::vector<DeviceDescription *> &all = ctx.GetAvailableDeviceDescriptions();
		std::cout << all.size() << " device description(s)" << std::endl;
		std::vector<DeviceDescription *> gpus = all;
		DeviceDescription::Filter(DEVICE_TYPE_HIP_GPU, gpus);
		DeviceDescription::FilterOne(gpus);
		for (const DeviceDescription *d : gpus)
			std::cout << "selected " << d->GetName() << " " << DeviceDescription::GetDeviceType(d->GetType())
					  << std::endl;
	} catch (const std::exception &e) {
		std::cerr << "error: " << e.what() << std::endl;
		return 1;
	}
	return 0;
}

int registry() {
	const DeviceType types[] = {DEVICE_TYPE_NATIVE, DEVICE_TYPE_OPENCL_CPU, DEVICE_TYPE_OPENCL_GPU, DEVICE_TYPE_CUDA_GPU,
			DEVICE_TYPE_HIP_GPU, DEVICE_TYPE_HIP_GPU};
	std::vector<std::unique_ptr<DeviceDescription>> owned;
	std::vector<DeviceDescription *> all;
	for (size_t i = 0; i < sizeof(types) / sizeof(types[0]); ++i) {
		owned.emplace_back(new DeviceDescription("dev" + std::to_string(i), types[i]));
		all.push_back(owned.back().get());
	}
	auto show = [](const char *what, const std::vector<DeviceDescription *> &v) {
		std::cout << what;
		for (const DeviceDescription *d : v) std::cout << " " << d->GetName();
		std::cout << std::endl;
	};
	for (DeviceType t : {DEVICE_TYPE_ALL, DEVICE_TYPE_NATIVE, DEVICE_TYPE_OPENCL_ALL, DEVICE_TYPE_OPENCL_DEFAULT,
				 DEVICE_TYPE_OPENCL_CPU, DEVICE_TYPE_OPENCL_GPU, DEVICE_TYPE_OPENCL_UNKNOWN, DEVICE_TYPE_CUDA_GPU,
				 DEVICE_TYPE_HIP_GPU})
		std::cout << "type " << (int)t << " " << DeviceDescription::GetDeviceType(t) << std::endl;
	try {
		DeviceDescription::GetDeviceType(DEVICE_TYPE_ALL_HARDWARE);
		std::cout << "type ALL_HARDWARE named" << std::endl;
	} catch (const std::exception &e) {
		std::cout << "type error: " << e.what() << std::endl;
	}
	for (DeviceType t : {DEVICE_TYPE_ALL, DEVICE_TYPE_HIP_GPU, DEVICE_TYPE_ALL_HARDWARE, DEVICE_TYPE_ALL_INTERSECTION,
				 DEVICE_TYPE_OPENCL_ALL}) {
		std::vector<DeviceDescription *> v = all;
		DeviceDescription::Filter(t, v);
		show(("filter " + std::to_string((int)t) + ":").c_str(), v);
	}
	std::vector<DeviceDescription *> one = all;
	DeviceDescription::FilterOne(one);
	show("filterone all:", one);
	std::vector<DeviceDescription *> hip = all;
	DeviceDescription::Filter(DEVICE_TYPE_HIP_GPU, hip);
	DeviceDescription::FilterOne(hip);
	show("filterone hip:", hip);
	std::vector<DeviceDescription *> cpu = {all[0], all[1]};
	DeviceDescription::FilterOne(cpu);
	show("filterone cpu-only:", cpu);
	try {
		Context ctx(nullptr, Properties().Set("context.verbose", "false"));
		std::vector<DeviceDescription *> bad = {all[4], all[0]};
		ctx.AddIntersectionDevices(bad);
		std::cout << "context accepted" << std::endl;
	} catch (const std::exception &e) {
		std::cout << "context error: " << e.what() << std::endl;
	}
	return 0;
}

}  // namespace

int main(int argc, char **argv) {
	if (argc == 2 && std::string(argv[1]) == "--devices") return listDevices();
	if (argc == 2 && std::string(argv[1]) == "--registry") return registry();
	if (argc < 2) {
		std::cerr << "usage: luxrays_hip_check <dir> [--accel auto|bvh|mbvh] [--anyhit] [--serial N] [--scene FILE] "
		             "[--refit] [--prop key=value]...\n";
		return 2;
	}
	const std::string dir = argv[1];
	// the Context configuration, as LuxRays reads it (accelerator.*, hip.*)
	Properties cfg;
	cfg.Set("context.verbose", "false");
	bool anyHit = false;
	u_int serial = 0;
	std::string sceneFile;
	bool refit = false;
	for (int i = 2; i < argc; ++i) {
		const std::string a = argv[i];
		if (a == "--accel" && i + 1 < argc) {
			const std::string v = argv[++i];
			cfg.Set("accelerator.type", v == "bvh" ? "BVH" : v == "mbvh" ? "MBVH" : "AUTO");
		} else if (a == "--anyhit")
			anyHit = true;
		else if (a == "--serial" && i + 1 < argc)
			serial = (u_int)std::stoul(argv[++i]);
		else if (a == "--scene" && i + 1 < argc)
			sceneFile = argv[++i];
		else if (a == "--refit")
			refit = true;
		else if (a == "--prop" && i + 1 < argc)
			cfg.SetFromString(argv[++i]);
		else {
			std::cerr << "unknown argument " << a << "\n";
			return 2;
		}
	}
	try {
		// meshes and object descriptions, from the scene file or the binary files
		std::vector<std::vector<Point>> verts;
		std::vector<std::vector<Triangle>> tris;
		std::vector<ObjDesc> descs;
		bool haveObjects = false;
		if (!sceneFile.empty()) {
			lrh_scenefile *sf = nullptr;
			if (lrh_scenefile_load(sceneFile.c_str(), nullptr, 0, &sf) != LRH_OK)
				throw std::runtime_error(lrh_ingest_last_error());
			std::unique_ptr<lrh_scenefile, int (*)(lrh_scenefile *)> hold(sf, lrh_scenefile_destroy);
			lrh_scenefile_info info;
			if (lrh_scenefile_get_info(sf, &info) != LRH_OK) throw std::runtime_error(lrh_ingest_last_error());
			std::vector<lrh_mesh> ms(info.nMeshes);
			std::vector<lrh_object> os(info.nObjects);
			if (lrh_scenefile_meshes(sf, ms.data(), info.nMeshes) != LRH_OK ||
			    lrh_scenefile_objects(sf, os.data(), info.nObjects) != LRH_OK)
				throw std::runtime_error(lrh_ingest_last_error());
			for (const lrh_mesh &m : ms) {
				verts.emplace_back(m.nVerts);
				tris.emplace_back(m.nTris);
				for (u_int v = 0; v < m.nVerts; ++v)
					verts.back()[v] = Point(m.verts[3 * v], m.verts[3 * v + 1], m.verts[3 * v + 2]);
				std::memcpy(tris.back().data(), m.tris, (size_t)m.nTris * sizeof(Triangle));
			}
			for (const lrh_object &o : os) {
				ObjDesc d;
				d.mesh = o.mesh;
				d.kind = o.kind;
				for (int k = 0; k < 16; ++k) d.m[k] = (k % 5 == 0) ? 1.f : 0.f;
				if (o.kind == 1) std::memcpy(d.m, o.m, 64);
				if (o.kind == 2) {
					d.times.assign(o.keyTimes, o.keyTimes + o.nKeys);
					d.mats.assign(o.keyMats, o.keyMats + 16 * (size_t)o.nKeys);
				}
				descs.push_back(d);
			}
			haveObjects = true;
			std::cout << "scene " << sceneFile << " meshes " << info.nMeshes << " objects " << info.nObjects
			          << " world triangles " << info.nWorldTris << std::endl;
		} else {
			const std::vector<char> mb = readAll(dir + "/meshes.bin", true);
			Reader r(mb);
			const u_int nMeshes = r.u();
			for (u_int i = 0; i < nMeshes; ++i) {
				const u_int nv = r.u(), nt = r.u();
				verts.emplace_back(nv);
				tris.emplace_back(nt);
				r.get(verts.back().data(), nv);
				r.get(tris.back().data(), nt);
			}
			const std::vector<char> ob = readAll(dir + "/objects.bin", false);
			if (!ob.empty()) {
				descs = readObjects(ob);
				haveObjects = true;
			}
		}
		const u_int nMeshes = (u_int)verts.size();
		std::vector<std::unique_ptr<TriangleMesh>> base;
		for (u_int i = 0; i < nMeshes; ++i)
			base.emplace_back(new TriangleMesh((u_int)verts[i].size(), (u_int)tris[i].size(), verts[i].data(),
			                                   tris[i].data()));
		// objects
		std::vector<std::unique_ptr<Mesh>> owned;
		std::vector<Mesh *> objects;
		if (!haveObjects) {
			for (auto &m : base) objects.push_back(m.get());
		} else {
			for (const ObjDesc &o : descs) {
				if (o.mesh >= nMeshes) throw std::runtime_error("object mesh index out of range");
				if (o.kind == 0)
					objects.push_back(base[o.mesh].get());
				else if (o.kind == 1) {
					owned.emplace_back(new InstanceTriangleMesh(base[o.mesh].get(), Transform(mat(o.m))));
					objects.push_back(owned.back().get());
				} else {
					owned.emplace_back(new MotionTriangleMesh(base[o.mesh].get(), motion(o)));
					objects.push_back(owned.back().get());
				}
			}
		}
		if (refit) cfg.Set("hip.accelerator.update", "REFIT");
		Context ctx(nullptr, cfg);
		DataSet ds(&ctx);
		for (Mesh *m : objects) ds.Add(m);
		ds.Preprocess();

		// rays
		std::vector<char> rb = readAll(dir + "/rays.bin", true);
		if (rb.size() % sizeof(Ray)) throw std::runtime_error("rays.bin is not a whole number of Ray records");
		const u_int n = (u_int)(rb.size() / sizeof(Ray));
		std::vector<RayHit> hits(n);

		std::vector<DeviceDescription *> devDescs = ctx.GetAvailableDeviceDescriptions();
		DeviceDescription::Filter(DEVICE_TYPE_HIP_GPU, devDescs);
		DeviceDescription::FilterOne(devDescs);
		if (devDescs.empty()) throw std::runtime_error("lrhip: no HIP GPU device available");
		IntersectionDevice *idev = ctx.AddIntersectionDevices(devDescs)[0];
		HIPIntersectionDevice &device = dynamic_cast<HIPIntersectionDevice &>(*idev);
		ctx.SetDataSet(&ds);
		ctx.Start();
		std::cout << "device " << device.GetName() << " accel "
				  << (device.GetAccelerator()->GetType() == ACCEL_BVH ? "BVH" : "MBVH") << " rays " << n << std::endl;

		HardwareDeviceBuffer *rayBuff = nullptr, *hitBuff = nullptr;
		device.AllocBufferRO(&rayBuff, rb.data(), rb.size(), "Ray");
		device.AllocBufferRW(&hitBuff, nullptr, (size_t)n * sizeof(RayHit), "RayHit");
		auto traceAll = [&](const char *name) {
			if (anyHit)
				device.EnqueueTraceRayBufferAnyHit(rayBuff, hitBuff, n);
			else
				device.EnqueueTraceRayBuffer(rayBuff, hitBuff, n);
			device.EnqueueReadBuffer(hitBuff, false, (size_t)n * sizeof(RayHit), hits.data());
			device.FinishQueue();
			writeAll(dir + "/" + name, hits.data(), (size_t)n * sizeof(RayHit));
		};
		traceAll("hits.bin");

		if (serial) {
			const u_int k = serial < n ? serial : n;
			std::vector<RayHit> sh(k);
			const Ray *rays = reinterpret_cast<const Ray *>(rb.data());
			u_int nHit = 0;
			for (u_int i = 0; i < k; ++i) nHit += device.TraceRay(&rays[i], &sh[i]) ? 1 : 0;
			writeAll(dir + "/serial.bin", sh.data(), (size_t)k * sizeof(RayHit));
			std::cout << "serial " << k << " hits " << nHit << std::endl;
		}

		// DataSet edit: new transforms / keys, then Update()
		const std::vector<char> ob2 = readAll(dir + "/objects2.bin", false);
		if (!ob2.empty()) {
			const std::vector<ObjDesc> o2 = readObjects(ob2);
			if (o2.size() != objects.size()) throw std::runtime_error("objects2.bin has another object count");
			for (size_t i = 0; i < o2.size(); ++i) {
				if (o2[i].kind == 1)
					static_cast<InstanceTriangleMesh *>(objects[i])->SetTransformation(Transform(mat(o2[i].m)));
				else if (o2[i].kind == 2)
					static_cast<MotionTriangleMesh *>(objects[i])->SetMotionSystem(motion(o2[i]));
			}
			ctx.UpdateDataSet();
			traceAll("hits2.bin");
		}

		std::cout << "rays traced " << (unsigned long long)device.GetTotalRaysCount() << " used memory "
				  << device.GetUsedMemory() << std::endl;
		device.FreeBuffer(&rayBuff);
		device.FreeBuffer(&hitBuff);
		ctx.Stop();
	} catch (const std::exception &e) {
		std::cerr << "error: " << e.what() << std::endl;
		return 1;
	}
	return 0;
}
