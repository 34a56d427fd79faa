/*
 * lrh_ingest.h — on-disk scene ingest for the HIP intersection device:
 * PLY meshes and the geometry part of LuxCore scene files (.scn / .cfg),
 * turned into the lrh_mesh / lrh_object arrays that lrh_scene_create_bvh and
 * lrh_scene_create_mbvh take.  Host code only (libluxrays_ingest.so, no HIP).
 *
 * Every call returns LRH_OK (0) or a negative LRH_E* code (lrhip.h);
 * lrh_ingest_last_error() gives a thread-local message for the last failing
 * call on the calling thread.
 *
 * Reference interfaces replaced (paths relative to the LuxCore v2.6 tree):
 *   PLY mesh            ExtTriangleMesh::LoadPly     src/luxrays/core/exttrianglemeshfile.cpp:259-452
 *                       (face split rule, FaceCB     exttrianglemeshfile.cpp:207-239)
 *   properties file     Properties::SetFromStream    src/luxrays/utils/properties.cpp:843-875
 *                       Property::FromString         properties.cpp:652-749
 *                       Property::Get<Matrix4x4>     properties.cpp:606-614
 *   shapes              Scene::CreateShape           src/slg/scene/parseshapes.cpp:141-400
 *                       (geometry-preserving types only; see below)
 *   objects             Scene::ParseObjects / CreateObject
 *                                                    src/slg/scene/parseobjects.cpp:30-200
 *   DataSet order       Scene::Preprocess            src/slg/scene/scenepreprocess.cpp:52-61
 *   file names          FileNameResolver::ResolveFile src/slg/utils/filenameresolver.cpp:48-95
 */
#ifndef LRH_INGEST_H
#define LRH_INGEST_H

#include "lrhip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define LRH_EIO (-7)        /* file missing, unreadable or malformed */

typedef struct lrh_scenefile lrh_scenefile;

/* Counts of a loaded scene.  twoLevel = 1 when any object is an instance or
 * has motion: the DataSet's AUTO policy then picks MBVH
 * (src/luxrays/core/dataset.cpp:66-81, src/luxrays/devices/oclintersectiondevice.cpp:42-56),
 * otherwise BVH over one mesh per object in object order. */
typedef struct {
	unsigned int nMeshes;          /* unique meshes (shapes and PLY files) */
	unsigned int nObjects;         /* scene objects = DataSet entries */
	unsigned int nKeys;            /* motion keys over all objects */
	unsigned int twoLevel;
	unsigned long long nVerts;     /* over unique meshes */
	unsigned long long nTris;
	unsigned long long nWorldTris; /* over objects (instances counted once each) */
} lrh_scenefile_info;

/* scene.camera.lookat (the first 6 values; up is then +z) or .lookat.orig /
 * .lookat.target / .up, and .fieldofview (default 45), as
 * src/slg/scene/parsecamera.cpp:50-117 reads them; with neither lookat form
 * every vector is zero (the reference's identity-camera case). */
typedef struct {
	float orig[3];
	float target[3];
	float up[3];
	float fieldOfView;             /* degrees */
} lrh_scenefile_camera;

/* One PLY file as a scene of one mesh and one plain object.  ASCII and
 * binary (both byte orders) bodies, every PLY scalar type; faces with 3
 * vertices become one triangle, with 4 vertices two triangles (v0 v1 v2),
 * (v0 v2 v3), other lengths are skipped, as in the reference.  Unlike the
 * reference, a vertex index outside the vertex list is an error. */
int lrh_ply_load(const char *path, lrh_scenefile **out);

/* A .scn file (or a .cfg naming one with scene.file).  Relative file names
 * are resolved as the reference resolves them: as given (relative to the
 * working directory), then under each search directory — the scene file's
 * own directory first, then searchDirs — with the full relative name, then
 * with its last component alone.  Shape types that only add AOVs or UVs
 * (pointiness, harlequin, islandaov, randomtriangleaov, edgedetectoraov,
 * cameraprojuv) keep their source's geometry; types that change geometry
 * (subdiv, displacement, bevel, simplify, strands, group) return LRH_EINVAL. */
int lrh_scenefile_load(const char *path, const char *const *searchDirs, unsigned int nDirs,
                       lrh_scenefile **out);

/* The same from text in memory; relative names resolve against baseDir. */
int lrh_scenefile_parse(const char *text, size_t len, const char *baseDir, lrh_scenefile **out);

int lrh_scenefile_get_info(const lrh_scenefile *sf, lrh_scenefile_info *info);

/* Unique meshes (n = nMeshes).  The pointers stay valid until
 * lrh_scenefile_destroy. */
int lrh_scenefile_meshes(const lrh_scenefile *sf, lrh_mesh *out, unsigned int n);

/* Objects in DataSet order (n = nObjects): kind 0 plain, 1 instance (m =
 * local->world, row major, as Transform(mat) holds it; mInv NULL), 2 motion
 * (keys global->local = Inverse(Transform(mat)) per key, as
 * parseobjects.cpp:138-159 builds them).  Pointers valid until destroy. */
int lrh_scenefile_objects(const lrh_scenefile *sf, lrh_object *out, unsigned int n);

/* Object name and the name of its mesh (shape name or PLY path as written). */
int lrh_scenefile_object_name(const lrh_scenefile *sf, unsigned int obj, const char **name,
                              const char **meshName);

int lrh_scenefile_get_camera(const lrh_scenefile *sf, lrh_scenefile_camera *cam);

/* The object's material reduced to what the Lambertian shading step of
 * lrh_shade uses (src/slg/scene/parsematerials.cpp:143-186, 555-563):
 * kd = .kd (default 0.75 grey; one or three numbers, or a constfloat1 /
 * constfloat3 texture), le = .emission x .emission.gain (0 when .emission is
 * absent or black).  Material types other than matte keep their .kd (or the
 * default) as a Lambertian stand-in; textures other than constants read as
 * the default.  The emission power/efficiency normalisation is not applied. */
int lrh_scenefile_object_material(const lrh_scenefile *sf, unsigned int obj, float kd[3], float le[3]);

int lrh_scenefile_destroy(lrh_scenefile *sf);

const char *lrh_ingest_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
