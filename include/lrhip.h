/*
 * lrhip.h — C ABI of the MI355X (gfx950) ray-intersection device for LuxRays.
 *
 * Plain pointers and sizes only.  Every call returns LRH_OK (0) or a negative
 * LRH_E* code; lrh_last_error() gives a thread-local message for the last
 * failing call on the calling thread.  Device pointers are ordinary HIP
 * device addresses; `stream` is a hipStream_t (NULL = the device's own
 * stream, created at default priority).
 *
 * Reference interfaces replaced (paths relative to the LuxCore v2.6 tree):
 *   device/buffer/queue  HardwareDevice            include/luxrays/core/hardwaredevice.h:113-183
 *   scene build          Accelerator::Init         include/luxrays/core/accelerator.h:39-59
 *                        BVHAccel::Init            src/luxrays/accelerators/bvhaccel.cpp:72-168
 *                        MBVHAccel::Init           src/luxrays/accelerators/mbvhaccel.cpp:58-216
 *   upload               BVHKernel ctor            src/luxrays/accelerators/bvhaccelhw.cpp:40-237
 *                        MBVHKernel ctor           src/luxrays/accelerators/mbvhaccelhw.cpp:41-306
 *   edit                 MBVHAccel::Update + MBVHKernel::Update
 *                                                  mbvhaccel.cpp:237-250, mbvhaccelhw.cpp:442-466
 *   trace                HardwareIntersectionKernel::EnqueueTraceRayBuffer
 *                                                  include/luxrays/core/hardwareintersectiondevice.h:59-72
 *                        (kernel Accelerator_Intersect_RayBuffer, include/luxrays/accelerators/bvh.cl:228-260,
 *                         mbvh.cl:351-383)
 *   stats                IntersectionDevice ray counters include/luxrays/core/intersectiondevice.h:42-63
 */
#ifndef LRHIP_H
#define LRHIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LRH_OK 0
#define LRH_EINVAL (-1)     /* bad argument */
#define LRH_EHIP (-2)       /* HIP runtime error */
#define LRH_ENOMEM (-3)     /* allocation failed / over budget */
#define LRH_ERANGE (-4)     /* size beyond a documented limit */
#define LRH_ENODEV (-5)     /* no GPU / device not available */
#define LRH_EDEVICE (-6)    /* a kernel reported an internal error (e.g. stack overflow) */

/* Ray / RayHit records: include/luxrays/core/geometry/ray.h:35-103 */
#define LRH_RAY_BYTES 48
#define LRH_HIT_BYTES 20
#define LRH_RAY_FLAGS_MASKED 1u
#define LRH_NULL_INDEX 0xffffffffu

/* Query types.  Closest hit is the reference's only query; any hit (the first
 * accepted hit with mint <= t < maxt ends the ray) is the extension that SLG's
 * shadow rays would use instead of closest hit + meshIndex test
 * (include/slg/engines/pathoclbase/kernels/pathoclbase_kernels_micro.cl:490-507). */
#define LRH_QUERY_CLOSEST 0u
#define LRH_QUERY_ANY 1u

/* Memory copy kinds */
#define LRH_H2D 1
#define LRH_D2H 2
#define LRH_D2D 3

typedef struct lrh_device lrh_device;
typedef struct lrh_scene lrh_scene;
typedef struct lrh_event lrh_event;
typedef struct lrh_batch lrh_batch;

/* One triangle mesh (TriangleMesh, include/luxrays/core/trianglemesh.h:79-134):
 * verts = 3*nVerts floats, tris = 3*nTris vertex indices local to the mesh. */
typedef struct {
	const float *verts;
	unsigned int nVerts;
	const unsigned int *tris;
	unsigned int nTris;
} lrh_mesh;

/* One DataSet object of a two-level scene (MBVHAccel root leaf,
 * mbvhaccel.cpp:90-202).  kind 0: the base mesh itself (TriangleMesh);
 * kind 1: InstanceTriangleMesh with local->world matrix `m` (row major) and,
 * optionally, its inverse `mInv` (NULL: computed with the reference's
 * Gauss-Jordan inverse, src/luxrays/core/geometry/matrix4x4.cpp:118-175);
 * kind 2: MotionTriangleMesh with nKeys keys (times ascending, matrices
 * global->local, row major) as MotionSystem takes them
 * (src/luxrays/core/geometry/motionsystem.cpp:296-331). */
typedef struct {
	unsigned int mesh;
	unsigned int kind;
	const float *m;
	const float *mInv;
	unsigned int nKeys;
	const float *keyTimes;
	const float *keyMats;
} lrh_object;

/* Build options (defaults when the pointer is NULL or a field is 0).
 * layout: 0 = 64-byte quantized 4-wide nodes (default), 1 = 128-byte fp32
 * 4-wide, 2 = 128-byte quantized 8-wide (one L2 line; every scene kind,
 * builder and trace entry point).
 * builder: LRH_BUILD_HOST_SAH (default) builds on the host (binned SAH, the
 * stand-in for the reference's EMBREE_BINNED_SAH, bvhaccel.cpp:144-156);
 * LRH_BUILD_DEVICE_LBVH builds on the GPU (Morton order + radix tree,
 * collapsed to the same 4-wide records, the triangle records bit for bit the
 * host's; single-level scenes with a device; no early split clipping, so
 * splitBudget is ignored; the replacement of the reference's EMBREE_MORTON
 * builder).  LRH_BUILD_DEVICE_PLOC builds on the GPU too, with the radix
 * tree replaced by agglomerative clustering of the Morton-ordered boxes
 * (rounds of mutual nearest neighbours within 16 places): a tree of lower
 * surface-area cost for a somewhat longer build, same records and limits.
 * A violated internal precondition of a device build (codes out of order, a
 * malformed tree) returns LRH_EHIP with the reason. */
#define LRH_BUILD_HOST_SAH 0u
#define LRH_BUILD_DEVICE_LBVH 1u
#define LRH_BUILD_DEVICE_PLOC 2u
typedef struct {
	unsigned int layout;
	unsigned int maxLeafSize;  /* 1..8, default 4 */
	unsigned int nThreads;     /* host build threads, default = hardware threads */
	float travCost;            /* SAH traversal cost relative to one triangle test, default 1.0 */
	unsigned int builder;      /* LRH_BUILD_HOST_SAH, LRH_BUILD_DEVICE_LBVH or LRH_BUILD_DEVICE_PLOC */
	float splitBudget;         /* early split clipping: triangles whose box area is above twice the
	                              median and above 2.5x the triangle's two-sided area (a poor fit) are
	                              cut into up to 64 references with tighter boxes, at most
	                              this many extra references per triangle over the scene (the size
	                              threshold rises until they fit); 0 = default (4.0), negative = off.
	                              Results are unchanged (a cut triangle is tested through any of its
	                              references); records grow by the extra references */
} lrh_build_opts;

typedef struct {
	unsigned long long nodeCount;
	unsigned long long nodeBytes;
	unsigned long long primCount;      /* triangle records (+ instance records) */
	unsigned long long primBytes;
	unsigned long long deviceBytes;    /* everything the scene holds on the device */
	unsigned int maxDepth;
	unsigned int layout;
	double buildSeconds;
	double uploadSeconds;
} lrh_scene_info;

/* Tuning of the trace kernel; all zero = built-in defaults.
 * blockSize in {64,128,256}; ldsStack a power of two in 4..64 entries per
 * lane; blocksPerCU 0 = occupancy-derived persistent grid, of which a launch
 * takes gridPercent (1..100, 0 = all of it): a caller whose launches run
 * beside its own work on other streams (the served cohorts) leaves that work
 * room to start while a launch drains. */
typedef struct {
	unsigned int blockSize;
	unsigned int ldsStack;
	unsigned int blocksPerCU;
	unsigned int refillThreshold;  /* refill a wave when fewer than this many lanes are live; 0 = default */
	unsigned int leafTris;         /* triangles a lane tests per traversal pass; the rest of its leaf waits
	                                  for the next pass; 0 = default (the whole leaf) */
	unsigned int gridPercent;      /* share of the occupancy-derived grid (blocksPerCU 0); 0 = 100 */
} lrh_trace_opts;

/* One ray buffer of a batched launch (lrh_trace_batch). */
typedef struct {
	lrh_scene *scene;
	const void *rays;     /* device Ray[count] */
	void *hits;           /* device RayHit[count] */
	unsigned int count;
	unsigned int query;
} lrh_batch_item;

const char *lrh_last_error(void);
const char *lrh_version(void);

/* ---- device, buffers, queue (HardwareDevice) ---- */
int lrh_device_count(int *count);
int lrh_device_open(int ordinal, lrh_device **out);
int lrh_device_close(lrh_device *dev);
int lrh_device_info(lrh_device *dev, char *name, size_t nameLen, int *cuCount, size_t *totalMem);
int lrh_alloc(lrh_device *dev, size_t bytes, void **dptr);
int lrh_free(lrh_device *dev, void *dptr);
int lrh_memcpy(lrh_device *dev, void *dst, const void *src, size_t bytes, int kind, void *stream);
int lrh_memcpy_async(lrh_device *dev, void *dst, const void *src, size_t bytes, int kind, void *stream);
int lrh_host_alloc(lrh_device *dev, size_t bytes, void **hptr);   /* pinned */
int lrh_host_free(lrh_device *dev, void *hptr);
int lrh_stream_create(lrh_device *dev, void **stream);
int lrh_stream_destroy(lrh_device *dev, void *stream);
int lrh_sync(lrh_device *dev, void *stream);
/* Device memory the library holds on this device: scenes, buffers, trace
 * workspaces, and the scratch of device BVH builds, which is kept (at most
 * 32 GiB) for the next build of the same sizes and released with the device. */
size_t lrh_used_memory(lrh_device *dev);
int lrh_event_create(lrh_device *dev, lrh_event **ev);
int lrh_event_destroy(lrh_event *ev);
int lrh_event_record(lrh_event *ev, void *stream);
int lrh_event_sync(lrh_event *ev);
int lrh_event_elapsed(lrh_event *start, lrh_event *end, float *ms);

/* ---- scenes (Accelerator::Init + HardwareIntersectionKernel upload) ---- */
int lrh_scene_create_bvh(lrh_device *dev, const lrh_mesh *meshes, unsigned int nMeshes,
		const lrh_build_opts *opts, lrh_scene **out);
int lrh_scene_create_mbvh(lrh_device *dev, const lrh_mesh *meshes, unsigned int nMeshes,
		const lrh_object *objects, unsigned int nObjects, const lrh_build_opts *opts, lrh_scene **out);
/* MBVHAccel::Update: new instance matrices / motion keys for the same objects
 * (same count, kinds and meshes; key counts may change); root tree rebuilt,
 * leaf trees kept.  Synchronous on the device's stream; no trace of this
 * scene may run on another stream meanwhile (the root nodes, instance
 * records and motion arrays are rewritten in place, and arrays that must grow
 * are reallocated).  An edit that reallocates an array or deepens the tree
 * invalidates resident batches over the scene (lrh_batch_launch/replay then
 * return LRH_EINVAL: recreate the batch). */
int lrh_scene_update(lrh_scene *scene, const lrh_object *objects, unsigned int nObjects);
/* The same edit without a rebuild (MBVHAccel::Update, mbvhaccel.cpp:237-250,
 * rebuilds the root tree on the host; MBVHKernel::Update,
 * mbvhaccelhw.cpp:442-466, re-uploads it): the root tree keeps its topology
 * and the GPU recomputes its boxes -- motion objects' MotionSystem::Bound
 * (1,025 samples per segment, motionsystem.cpp:75-87) and every root node,
 * bottom-up.  The object boxes are the ones lrh_scene_update would compute;
 * only the tree's grouping is the old one.  Synchronous on the device's
 * stream; no trace of this scene may run on another stream meanwhile (the
 * root nodes are rewritten in place).  LRH_EINVAL if a motion sample is
 * singular; the scene is then left as it was. */
int lrh_scene_refit(lrh_scene *scene, const lrh_object *objects, unsigned int nObjects);
/* The culling box of every object of a two-level scene (6 floats each: lo
 * xyz, hi xyz; MachineEpsilon-expanded world boxes, mbvhaccel.cpp:187-202),
 * in object order, as the last build, update or refit computed them. */
int lrh_scene_object_boxes(lrh_scene *scene, float *boxes, unsigned int nObjects);
int lrh_scene_destroy(lrh_scene *scene);
int lrh_scene_get_info(lrh_scene *scene, lrh_scene_info *info);
/* Copy of the device node / primitive arrays to host memory (for tests). */
int lrh_scene_download(lrh_scene *scene, void *nodes, size_t nodeBytes, void *prims, size_t primBytes);

/* ---- tracing (EnqueueTraceRayBuffer) ----
 * Asynchronous on `stream`.  Masked rays (flags & 1) leave their hit record
 * untouched; a miss writes t = maxt, b1 = b2 = 0 and both indices 0xffffffff.
 * meshIndex is the DataSet add order (two-level: the object index);
 * triangleIndex is local to that mesh. */
int lrh_trace(lrh_scene *scene, const void *rays, void *hits, unsigned int count,
		unsigned int query, void *stream);
int lrh_trace_ex(lrh_scene *scene, const void *rays, void *hits, unsigned int count,
		unsigned int query, const lrh_trace_opts *opts, void *stream);
/* Same trace through a counting variant of the kernel, synchronous:
 * out[0] = inner-node visits, out[1] = primitive records tested (triangles,
 * plus instance records in two-level scenes), out[2] = rays traced — the
 * bytes-per-ray roofline accounting; and the SIMD-efficiency counters, summed
 * over waves: out[3] = traversal passes, out[4] = live lanes summed over those
 * passes, out[5] = passes with a node test, out[6] = passes with a leaf test,
 * out[7] = refills; two-level scenes: out[8] = object entries (root leaves
 * entered, each an instance record read; out[1] counts them too), out[9] =
 * of those, entries through a static matrix, out[10] = through a motion
 * system. */
#define LRH_TRACE_STATS 11
int lrh_trace_stats(lrh_scene *scene, const void *rays, void *hits, unsigned int count,
		unsigned int query, const lrh_trace_opts *opts, void *stream, unsigned long long out[LRH_TRACE_STATS]);
/* lrh_trace that allocates nothing.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory while they trace (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, results and asynchrony as lrh_trace with the default launch
 * shape; the ticket counter and the spill stack live in one static scratch
 * array of LRH_NOALLOC_SCRATCH_BYTES per GPU (held from the library's first
 * use on that GPU until the process ends), whose counter each call clears
 * with a fill kernel.  The persistent grid is capped at the lanes whose spill
 * stacks fit the array (3·(treeDepth+2)+4-8 entries of 8 bytes per lane), so
 * deep trees run fewer lanes than lrh_trace; a tree too deep for even one
 * block of 256 lanes returns LRH_ERANGE.  Every call on one GPU shares the
 * array: calls that overlap on a GPU (other streams, other threads) race and
 * must be serialised by the caller. */
#define LRH_NOALLOC_SCRATCH_BYTES (128u * 1024u * 1024u)
int lrh_trace_noalloc(lrh_scene *scene, const void *rays, void *hits, unsigned int count,
		unsigned int query, void *stream);
/* Many ray buffers over one or more resident scenes of the same device in one
 * persistent launch: the rays of all items form one ticket space, every lane
 * takes the next ticket, finds its item and traces in that item's scene with
 * that item's query.  All scenes of a batch share one node layout; at most
 * 4,096 items and 2^31-1 rays.  (No reference counterpart: the reference
 * launches one kernel per buffer, bvhaccelhw.cpp:259-268.) */
int lrh_trace_batch(lrh_device *dev, const lrh_batch_item *items, unsigned int nItems,
		const lrh_trace_opts *opts, void *stream);
/* A resident batch: descriptor table and scratch uploaded once, then launched
 * many times, directly or as a captured hipGraph (one kernel node).  Replays
 * of one batch must not overlap; the kernel rewinds its own ticket counter, so
 * the graph holds no memset node.  The batch holds its scenes' device
 * addresses and spill depth: launch, capture and replay return LRH_EINVAL once
 * a scene was destroyed or edited in a way that moved its arrays or deepened
 * its tree (lrh_scene_update / lrh_scene_refit with more motion keys, or a
 * deeper rebuilt root); a refit that keeps the key counts leaves it valid. */
int lrh_batch_create(lrh_device *dev, const lrh_batch_item *items, unsigned int nItems,
		const lrh_trace_opts *opts, lrh_batch **out);
int lrh_batch_launch(lrh_batch *batch, void *stream);
int lrh_batch_capture(lrh_batch *batch, void *stream);
int lrh_batch_replay(lrh_batch *batch, void *stream);
int lrh_batch_info(lrh_batch *batch, unsigned long long *rays, unsigned int *blocks, unsigned int *blockSize);
int lrh_batch_destroy(lrh_batch *batch);
/* ---- ray generation around a trace (served path-tracing sessions) ----
 * Camera rays of a pinhole camera (pixel centres, row-major, width*height
 * rays) and, from a wavefront's rays + hits on a single-level scene, the next
 * cosine-weighted bounce (Russian roulette from bounce rrStart on, survival
 * rrSurvive; terminated paths become masked rays) and one shadow ray per hit
 * toward a uniform point of a 1x1 light centred at lightCenter (masked where
 * the path missed).  Random numbers hash (seed, pixelBase + i, bounce).  They
 * stand for SLG's camera / next-vertex ray generation
 * (pathoclbase_kernels_micro.cl:723-844, 1045-1075). */
typedef struct {
	float eye[3], look[3], up[3];
	float vfovDeg;
	float mint;
	unsigned int width, height;
} lrh_camera;
int lrh_rays_camera(lrh_device *dev, const lrh_camera *cam, void *rays, void *stream);
int lrh_rays_bounce(lrh_scene *scene, const void *rays, const void *hits, unsigned int count,
		unsigned int bounce, unsigned long long seed, unsigned int pixelBase, unsigned int rrStart,
		float rrSurvive, const float lightCenter[3], void *nextRays, void *shadowRays,
		unsigned int *liveCount /* device counter or NULL: += live rays written */, void *stream);
/* One shading step of a path-tracing wavefront: the part of SLG's
 * AdvancePaths micro-kernels that consumes a hit (MK_HIT_OBJECT,
 * pathoclbase_kernels_micro.cl:204-330), samples direct light (MK_DL_ILLUMINATE /
 * MK_DL_SAMPLE_BSDF, :450-560) and generates the next vertex ray
 * (MK_GENERATE_NEXT_VERTEX_RAY, :1045-1075), reduced to Lambertian surfaces
 * (MatteMaterial kd) and triangle area lights with next-event estimation:
 *   - a shadow ray of the previous step that missed adds its pending light;
 *   - emission is added where the camera sees a light directly (depth 0);
 *   - one light triangle is picked uniformly, a point on it uniformly by
 *     area; its contribution kd/pi * Le * cosS * cosL / d^2 / pdf waits for
 *     the shadow ray this step writes (masked when there is none);
 *   - the next ray is cosine-weighted (throughput *= kd), with Russian
 *     roulette from rrDepth on; paths end at maxDepth vertices.
 * Random numbers hash (seed, pixelBase + i, depth, dimension).  Single-level
 * scenes; materials are per DataSet mesh; all arrays are device memory.
 * Pass the hits of the previous step's shadow rays whenever that step wrote
 * shadow rays: light still pending without them is dropped.  One step after
 * the last vertex (all paths ended, next rays masked) gathers the light of
 * the last shadow rays.  A path that had already ended when the previous
 * step ran (lrh_path.pad = 1) and has no light pending is skipped: its slots
 * still hold the masked rays the two steps before wrote, so nextRays must be
 * the buffer of the step before the previous one (the ray buffers alternate,
 * as in the wavefront loop) and shadowRays the same buffer every step. */
typedef struct {
	float kd[3];      /* diffuse reflectance */
	float le[3];      /* emitted radiance (0: not a light) */
} lrh_material;
typedef struct {
	float p0[3], e1[3], e2[3];   /* world-space triangle: p0, p1 - p0, p2 - p0 */
	float le[3];                 /* emitted radiance */
} lrh_light_tri;
typedef struct {
	float thr[3];     /* path throughput */
	float rad[3];     /* radiance gathered so far (the pixel sample) */
	float dl[3];      /* direct light waiting for its shadow ray */
	unsigned int depth, alive;
	unsigned int pad;  /* 1 once a shading step has seen the path ended (lrh_paths_init: 0) */
} lrh_path;
typedef struct {
	unsigned int count;          /* paths in the wavefront */
	unsigned int maxDepth;       /* path vertices */
	unsigned int rrDepth;        /* Russian roulette from this vertex on */
	float rrSurvive;             /* survival probability */
	unsigned long long seed;
	unsigned int pixelBase;
	unsigned int pad;
} lrh_shade_params;
/* Fresh paths (throughput 1, nothing gathered). */
int lrh_paths_init(lrh_device *dev, void *paths, unsigned int count, void *stream);
int lrh_shade(lrh_scene *scene, const lrh_material *materials, unsigned int nMaterials,
		const lrh_light_tri *lights, unsigned int nLights, const void *rays, const void *hits,
		const void *shadowHits /* hits of the previous step's shadow rays, or NULL */, void *paths,
		void *nextRays, void *shadowRays, const lrh_shade_params *params,
		unsigned int *liveCount /* device counter or NULL: += live rays written */, void *stream);
/* One shading step of up to LRH_SHADE_BATCH_MAX ray buffers in one launch:
 * item k gets exactly what lrh_shade gives on its arguments (same materials,
 * lights and live counter for every item; scenes on one device).  No
 * reference counterpart: it is the hit-consumption step of several
 * wavefronts at once, as lrh_trace_batch is their trace. */
#define LRH_SHADE_BATCH_MAX 16
typedef struct {
	lrh_scene *scene;
	const void *rays, *hits;
	const void *shadowHits;           /* or NULL */
	void *paths, *nextRays, *shadowRays;
	lrh_shade_params params;
} lrh_shade_item;
int lrh_shade_batch(const lrh_shade_item *items, unsigned int nItems, const lrh_material *materials,
		unsigned int nMaterials, const lrh_light_tri *lights, unsigned int nLights, unsigned int *liveCount,
		void *stream);
/* ---- served path-tracing sessions (continuous batching) ----
 * No reference counterpart as a component: it batches what SLG's path tracer
 * does one device at a time, one ray buffer per wavefront traced and its hits
 * consumed (src/slg/engines/pathocl/pathoclopenclthread.cpp:183-189).
 * A session = width x height camera rays, then up to maxBounces bounce
 * wavefronts, one shadow ray per path vertex (any-hit), shaded on the GPU by
 * lrh_shade (mode LRH_SERVE_SHADE: grey Lambertian surfaces, a square area
 * light at `light`, Russian roulette after bounce rrStart) or lrh_rays_bounce
 * (LRH_SERVE_BOUNCE).  Sessions are admitted while their buffers fit
 * hbmBudget and their cohort's tick stays within maxRaysPerTick; each tick
 * traces one cohort's ready wavefronts in one lrh_trace_batch launch.
 * A session's rays and results do not depend on what else is in flight.
 * Host-side calls of one server must not overlap (one driving thread). */
#define LRH_SERVE_SHADE 0u
#define LRH_SERVE_BOUNCE 1u
typedef struct {
	unsigned int width, height;       /* session image, default 512 x 512 */
	unsigned int maxBounces;          /* bounce wavefronts after the camera one, default 8 */
	unsigned int rrStart;             /* Russian roulette after this bounce, default 2 */
	float rrSurvive;                  /* survival probability, default 0.7 */
	unsigned long long hbmBudget;     /* bytes of session buffers resident at once, default 4 GiB */
	unsigned int maxRaysPerTick;      /* rays one cohort's tick may carry, default 3 Mi */
	unsigned int cohorts;             /* streams the sessions are split over, 1..8, default 2 */
	unsigned int mode;                /* LRH_SERVE_SHADE (default) or LRH_SERVE_BOUNCE */
	float extent;                     /* scene box [0, extent]^3 for camera placement, default 100 */
	float light[3];                   /* light centre, default (extent/2, extent/2, 1.5 extent) */
	lrh_trace_opts opts;              /* launch shape of the batched trace (zeros: defaults; with more than
	                                     one cohort a zero gridPercent and blocksPerCU become gridPercent 80) */
} lrh_server_config;
typedef struct {
	unsigned int sid;                 /* caller's session id (unique per server) */
	unsigned int scene;               /* index into the server's scenes */
	unsigned long long seed;          /* camera placement and path random numbers */
	double arrival;                   /* seconds after the lrh_server_submit call */
} lrh_session_spec;
typedef struct {
	unsigned long long sessionsDone, sessionsQueued, ticks, wavefronts;
	unsigned long long liveRays;      /* camera rays + live rays the shading steps wrote */
	unsigned int maxActive;           /* most sessions live at once */
	unsigned int sessionBuffersAllocated;
	double busySeconds;               /* union of the ticks' GPU intervals (trace + shading) */
	double traceSeconds;              /* union of the batched traces' GPU intervals */
	unsigned long long latencySamples;
	double p50Ms, p99Ms, maxMs;       /* wavefront latency: ready -> completion of its tick */
	double queueP99Ms;                /* arrival -> admission */
} lrh_server_stats;
typedef struct {
	unsigned int buffers;             /* captured ray buffers */
	unsigned long long rays, liveRays;
	double batchMs;                   /* one lrh_trace_batch over every captured buffer */
	double perBufferMs;               /* one lrh_trace per captured buffer */
} lrh_plain_batch_result;
typedef struct lrh_server lrh_server;
int lrh_server_create(lrh_device *dev, lrh_scene *const *scenes, unsigned int nScenes, const lrh_server_config *cfg,
		lrh_server **out);
int lrh_server_submit(lrh_server *srv, const lrh_session_spec *specs, unsigned int n);
/* Keep every wavefront's rays and hits (and, in shade mode, the final
 * per-pixel radiance) of session `sid`, for checking; call before it is admitted. */
int lrh_server_record(lrh_server *srv, unsigned int sid);
/* One tick of the next cohort; *traced = 1 if it traced anything. */
int lrh_server_tick(lrh_server *srv, int *traced);
/* Ticks until every submitted session has finished (or timeLimit seconds). */
int lrh_server_run(lrh_server *srv, double timeLimit, double *wallSeconds);
int lrh_server_get_stats(lrh_server *srv, lrh_server_stats *stats);
/* Every wavefront latency in ms, in completion order; *n = how many there are. */
int lrh_server_latencies(lrh_server *srv, double *ms, size_t cap, size_t *n);
/* Recorded wavefront of step `step` (0 = camera): kind 0 the closest-hit
 * rays and hits, kind 1 the shadow rays (traced from step 1 on); host buffers
 * of width*height records (either may be NULL). */
int lrh_server_recorded(lrh_server *srv, unsigned int sid, unsigned int step, unsigned int kind, void *rays, void *hits,
		unsigned int *count);
/* Final per-pixel radiance (3 floats per pixel) of a recorded session. */
int lrh_server_radiance(lrh_server *srv, unsigned int sid, float *rgb, unsigned int count);
/* Copy every ray buffer the server traces into a device arena of maxBytes
 * (call before the first session is admitted), then time the same buffers as
 * one plain lrh_trace_batch and as one lrh_trace each: the served form's
 * throughput next to a single plain batch of the same rays (opts NULL: the
 * server's launch shape over the whole grid). */
int lrh_server_capture(lrh_server *srv, unsigned long long maxBytes);
int lrh_server_plain_batch(lrh_server *srv, const lrh_trace_opts *opts, unsigned int reps,
		lrh_plain_batch_result *result);
int lrh_server_destroy(lrh_server *srv);
/* Launch-shape autotuner of a batched trace: times each candidate shape
 * (NULL: the LRH_AUTOTUNE_SHAPES built-in ones, block size {128, 256} x LDS
 * stack {4, 8, 16} x refill threshold {4, 8, 12, 16, 24}) on the actual items
 * (resident batches, median of 3 launches after 2 warm-ups) and keeps the
 * fastest; candidateMs (optional, one per candidate) gets each time, NaN for
 * a shape the device refuses.  With cachePath and key (no spaces), the result
 * is read from / stored in a text cache ("key block lds refill ms blocksPerCU
 * leafTris gridPercent" per line; the last three may be absent),
 * so a serving process tunes once per workload signature; a cached answer
 * times nothing (candidateMs all NaN).  (No reference counterpart: the
 * reference's work-group size is a fixed setting,
 * src/slg/engines/oclrenderengine.cpp:193.) */
#define LRH_AUTOTUNE_SHAPES 30
int lrh_autotune_batch(lrh_device *dev, const lrh_batch_item *items, unsigned int nItems,
		const lrh_trace_opts *candidates, unsigned int nCandidates, const char *cachePath, const char *key,
		lrh_trace_opts *best, float *bestMs, float *candidateMs);
/* Build autotuner of a batched trace: for each candidate build option set
 * (node layout, leaf size, ...) builds the nScenes single-level scenes from
 * their meshes (scene j: sceneMeshes[j][0 .. sceneMeshCount[j])), runs
 * lrh_autotune_batch over the items with each item's scene replaced by the
 * candidate's build of scene itemScene[i] (the items' own scene fields are
 * ignored; shapes / nShapes as lrh_autotune_batch's candidates, no cache),
 * releases the scenes, and keeps the fastest candidate: bestCandidate gets
 * its index, bestShape its launch shape, candidateMs (optional, one per
 * candidate) each candidate's best time.  (No reference counterpart: the
 * reference's tree layout is fixed, bvhaccel.cpp:144-156.) */
int lrh_autotune_build(lrh_device *dev, const lrh_mesh *const *sceneMeshes, const unsigned int *sceneMeshCount,
		unsigned int nScenes, const lrh_batch_item *items, const unsigned int *itemScene, unsigned int nItems,
		const lrh_build_opts *candidates, unsigned int nCandidates, const lrh_trace_opts *shapes,
		unsigned int nShapes, unsigned int *bestCandidate, lrh_trace_opts *bestShape, float *candidateMs);

/* 1 if all work before the event has completed, 0 if not yet, < 0 on error. */
int lrh_event_query(lrh_event *ev);

/* Data-parallel ray count traced so far (IntersectionDevice stats). */
int lrh_device_stats(lrh_device *dev, unsigned long long *rays, double *seconds);

#ifdef __cplusplus
}
#endif

#endif
