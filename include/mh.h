/*
 * mh.h — C ABI of the MI355X ray-traversal library (libmh.so) behind Manta's
 * DynBVH plugin interface.
 *
 * Every entry point returns MH_OK (0) or a negative MH_E* code; on failure
 * mh_last_error() returns a thread-local message. No C++ exception crosses this
 * boundary. All array arguments are plain pointers with explicit sizes; device
 * pointers are HIP device addresses on the device chosen by mh_init().
 *
 * Reference interfaces replaced (paths are in the rfManta tree):
 *   AccelerationStructure  Interface/AccelerationStructure.h:11-66
 *   Object::intersect      Interface/Object.h:34
 *   DynBVH                 Model/Groups/DynBVH.h:22-391, DynBVH.cc:101-2501
 *   RayPacket hit/mask     Interface/RayPacket.h:407-442,488-502,531-657
 */
#ifndef MH_H_
#define MH_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MH_ABI_VERSION 2

/* Status codes. */
#define MH_OK 0
#define MH_E_INVALID (-1)     /* bad argument (null pointer, size, flag) */
#define MH_E_HIP (-2)         /* a HIP runtime call failed */
#define MH_E_NOMEM (-3)       /* device or host allocation failed */
#define MH_E_STATE (-4)       /* call out of order (e.g. trace before build) */
#define MH_E_UNSUPPORTED (-5) /* input outside a documented limit */
#define MH_E_IO (-6)          /* file could not be read or written */
#define MH_E_CANCELED (-7)    /* request cancelled (mh_server_cancel) */

/* Trace flags. */
#define MH_CLOSEST_HIT 0u
#define MH_ANY_HIT 1u /* RayPacket::AnyHit (Interface/RayPacket.h:148); opaque occluders */

/* Triangle tests (Model/Primitives). */
#define MH_PRIM_KENSLER_SHIRLEY 0 /* KenslerShirleyTriangle.cc:134-377 */
#define MH_PRIM_WALD 1            /* WaldTriangle.cc:64-555 */

/* Sentinel values of mh_hits.prim. */
#define MH_NO_HIT (-1)
#define MH_MASKED (-2)

/* Reference constants (include/CMakeLists.txt:185,188). */
#define MH_T_EPSILON 1.e-4f
#define MH_MAXT 1.e19f

typedef struct mh_scene mh_scene;

/* Rays in structure-of-arrays form, one float per ray per array, exactly the
 * RayPacketData origin/direction/minT fields (Interface/RayPacket.h:92-97).
 * tmax is the ray's initial minT (MH_MAXT for camera rays, 1 - MH_T_EPSILON for
 * shadow rays, Model/Lights/PointLight.cc:42). masked may be NULL; a nonzero byte
 * marks a ray as masked (RayPacket::maskRay, RayPacket.h:488-495): it never hits
 * and reports prim == MH_MASKED and, for MH_ANY_HIT, occluded == 1. */
typedef struct mh_rays {
  const float* ox;
  const float* oy;
  const float* oz;
  const float* dx;
  const float* dy;
  const float* dz;
  const float* tmax;
  const uint8_t* masked;
} mh_rays;

/* Hit records in structure-of-arrays form. For MH_CLOSEST_HIT: prim (triangle
 * index in the mesh, or MH_NO_HIT / MH_MASKED), t (the final minT), u and v
 * (barycentric weights of vertex 1 and vertex 2, the SCRATCH_U/SCRATCH_V
 * scratchpad slots). For MH_ANY_HIT: occluded (RayPacket::wasHit) and, if
 * non-NULL, prim (the occluder found first). Any pointer except the one the
 * mode needs may be NULL. */
typedef struct mh_hits {
  int32_t* prim;
  float* t;
  float* u;
  float* v;
  uint8_t* occluded;
} mh_hits;

/* Node layouts of the single-ray traversal. All give DynBVH::intersect's one-ray results
 * and visit order. */
#define MH_LAYOUT_PAIR 0 /* 64-byte nodes holding both child boxes */
#define MH_LAYOUT_REF 1  /* the reference's 32-byte BVHNode records (frames only) */
#define MH_LAYOUT_SOA 2  /* the 64-byte nodes as four arrays of their 16-byte fields (frames
                            only; the scene keeps the copy, refreshed after builds, refits
                            and updates) */

/* Trace options; NULL means the defaults given in brackets. */
typedef struct mh_trace_opts {
  /* [1] 1 = each ray traverses alone (DynBVH::intersect on a one-ray packet);
   * 2..64 = consecutive groups of packet_size rays traverse as one RayPacket on
   * one 64-lane wave (DynBVH::intersectNode packet semantics). */
  int packet_size;
  /* [0] per-lane traversal stack entries held in LDS (8,16,32; 0 = autotuned). */
  int stack_lds;
  /* [0] persistent blocks per CU (0 = autotuned). */
  int blocks_per_cu;
  /* [NULL] optional device int64[2]: box tests, triangle tests (accumulated). */
  int64_t* counters;
} mh_trace_opts;

/* Build options. */
typedef struct mh_build_opts {
  int prim_kind;   /* MH_PRIM_* [MH_PRIM_KENSLER_SHIRLEY] */
  int rtsah;       /* 1 = compute isLeftCheaper (DynBVH.cc:1583-1678) [1] */
  int threads;     /* host threads for the full SAH build, 0 = all cores */
  /* 0 = the reference's default build: full-sweep SAH (partitionSAH), on the host.
   * 1 = the reference's approximate build (USE_APPROXIMATE_BUILD, the
   *     MANTA_USE_DYNBVH_APPROXIMATE option, include/CMakeLists.txt:134-147,
   *     DynBVH.cc:1815-1822): 16-bin SAH at every node (partitionApproxSAH,
   *     DynBVH.cc:2125-2275), built on the device; the same tree record for record. */
  int approximate;
} mh_build_opts;

/* Library. */
int mh_init(int device);                 /* select the HIP device; idempotent */
const char* mh_last_error(void);          /* thread-local message of the last failure */
int mh_abi_version(void);
int mh_device_info(char* name, int name_len, int* cu_count, size_t* hbm_bytes);

/* Scene life cycle — AccelerationStructure::setGroup (AccelerationStructure.h:20)
 * on a Mesh (Model/Groups/Mesh.h:17-118): copies the host vertex (3 floats each)
 * and index (3 uint32 each) arrays. */
int mh_scene_create(const float* verts, uint32_t n_verts, const uint32_t* idx, uint32_t n_tris,
                    mh_scene** out);
int mh_scene_destroy(mh_scene* scene);

/* DynBVH::rebuild(proc, numProcs) with group_changed (DynBVH.cc:1699-1793): full SAH
 * build on the host (same split choices as partitionSAH/buildEvents, DynBVH.cc:2310-2440),
 * optional RTSAH, then upload. */
int mh_scene_build(mh_scene* scene, const mh_build_opts* opts);

/* The render-thread count (numProcs) the reference's DynBVH would see:
 * largeSubtreeSize = records / (20 * numProcs) (DynBVH.cc:945, 1776). Sets the
 * isLargeSubtree bits of the current tree as computeSubTreeSizes does (DynBVH.cc:941-950,
 * 1971-1996), and of later builds (rebuild(proc, numProcs)) and updates. With several
 * workers the reference updates seeded subtrees first and the large nodes above them
 * after (parallelUpdateBounds, DynBVH.cc:2012-2066); every node is still refitted and
 * rotated once after its children, so mh_scene_update's tree is the same for any count
 * and only these bits depend on it. [1] */
int mh_scene_set_workers(mh_scene* scene, int workers);

/* The same host build without a scene or a device: DynBVH::rebuild (DynBVH.cc:1699-1793)
 * writing object_ids (n_tris ints) and 2*n_tris 32-byte BVHNode records (DynBVH.h:37-77,
 * the array size DynBVH::allocate uses, DynBVH.cc:1680-1697). *num_nodes receives the
 * number of records in use. Touches no GPU. */
int mh_dynbvh_build_host(const float* verts, uint32_t n_verts, const uint32_t* idx,
                         uint32_t n_tris, const mh_build_opts* opts, int32_t* ids, void* nodes,
                         uint32_t* num_nodes);

/* PLY mesh ingest, as readPlyFile builds a Mesh (Model/Readers/PlyReader.cc:136-269):
 * vertex x, y, z (any scalar type, stored as float) transformed by the 3x4 row-major
 * `transform` (AffineTransform::multiply_point; NULL = identity), faces from the
 * "vertex_indices" list as triangle fans, then Mesh::removeDegenerateTriangles
 * (Model/Groups/Mesh.cc:315-359). ascii and binary PLY. With verts and idx NULL, only
 * *n_verts and *n_tris are set; otherwise the arrays (verts_cap vertices, tris_cap
 * triangles) receive the mesh. Host only. */
int mh_ply_read(const char* path, const float* transform, float* verts, uint32_t verts_cap,
                uint32_t* idx, uint32_t tris_cap, uint32_t* n_verts, uint32_t* n_tris);

/* OBJ mesh ingest, as ObjGroup builds a Mesh (Model/Groups/ObjGroup.cc:184-316) from the
 * file glmReadOBJ reads (Model/Readers/glm/glm.cc:575-960, 1599-1670): vertices, normals
 * and texture coordinates (u, v, 1) in file order; faces as triangle fans, emitted group by
 * group in glm's list order; per corner a normal index (a face normal is appended for
 * corners without one when the file has normals) and a texture index
 * (MH_NO_TEXTURE_INDEX where a corner has none and the file has texture coordinates); per
 * triangle a material slot, Mesh::face_material: 0 = the caller's default material, k >= 1
 * = glm material k-1 (material_names[k]; glm material 0 is NO_ASSIGNED_MATERIAL, then the
 * newmtl entries of the material library); then Mesh::removeDegenerateTriangles
 * (Model/Groups/Mesh.cc:315-359). Material parameters and textures are not read. The
 * normal and texture arrays are NULL when the file has none; normals, normal_idx,
 * texcoords and tex_idx are the arguments mh_scene_set_attributes takes. Host only. */
typedef struct mh_obj_mesh {
  uint32_t n_verts, n_tris, n_normals, n_texcoords, n_materials;
  const float* verts;                 /* [3 n_verts] */
  const uint32_t* idx;                /* [3 n_tris], 0-based */
  const float* normals;               /* [3 n_normals] or NULL */
  const uint32_t* normal_idx;         /* [3 n_tris] or NULL */
  const float* texcoords;             /* [3 n_texcoords] or NULL */
  const uint32_t* tex_idx;            /* [3 n_tris] or NULL */
  const uint32_t* face_material;      /* [n_tris] */
  const char* const* material_names;  /* [n_materials]; [0] is "" (the default slot) */
} mh_obj_mesh;

/* Reads an OBJ file (and the material library it names, beside it) into a mesh the
 * library owns until mh_obj_free. */
int mh_obj_read(const char* path, mh_obj_mesh** mesh);
void mh_obj_free(mh_obj_mesh* mesh);

/* DynBVH::buildFromFile (DynBVH.cc:2442-2472) from memory: n_ids object ids and n_nodes
 * 32-byte BVHNode records in the reference layout (DynBVH.h:37-77). */
int mh_scene_load_dynbvh(mh_scene* scene, const int32_t* ids, uint32_t n_ids, const void* nodes,
                         uint32_t n_nodes, int prim_kind);
/* DynBVH::buildFromFile / saveToFile (DynBVH.cc:2442-2501) on a file path. */
int mh_scene_load_dynbvh_file(mh_scene* scene, const char* path, int prim_kind);
int mh_scene_save_dynbvh_file(const mh_scene* scene, const char* path);
/* Copies the current tree out in the reference layout (sizes first when arrays are NULL). */
int mh_scene_export_dynbvh(const mh_scene* scene, int32_t* ids, uint32_t* n_ids, void* nodes,
                           uint32_t* n_nodes);

/* Scene queries: bounds (Object::computeBounds, Object.h:33) and tree statistics. */
int mh_scene_bounds(const mh_scene* scene, float lo[3], float hi[3]);
int mh_scene_stats(const mh_scene* scene, uint32_t* n_nodes, uint32_t* n_inner, uint32_t* depth,
                   size_t* device_bytes);

/* Per-frame update — AccelerationStructure::update/rebuild without group change
 * (DynBVH.cc:1574-1581, 2012-2123): replace the vertex positions (same topology), then
 * update the tree on the GPU. set_vertices copies from host memory (on_device = 0) or
 * from device memory (on_device = 1) on `stream`.
 * mh_scene_update is DynBVH::update as the reference builds it (TREE_ROT set,
 * DynBVH.h:377): refit with tree rotations and approximate-SAH leaf splits
 * (updateBounds<false> + rotateNode, DynBVH.cc:2078-2123, 3051-3372); the tree's
 * topology changes from frame to frame exactly as the reference's (record numbering of
 * split leaves may differ). It synchronises `stream` once (level bookkeeping).
 * mh_scene_refit is the bounds-only refit (updateBounds<false> without rotations): the
 * topology never changes; asynchronous on `stream`. */
int mh_scene_set_vertices(mh_scene* scene, const float* verts, int on_device, void* stream);
int mh_scene_update(mh_scene* scene, void* stream);
int mh_scene_refit(mh_scene* scene, void* stream);
/* Device pointer to the scene's vertex array (3 floats per vertex), for callers that
 * animate in place; call mh_scene_refit afterwards. */
int mh_scene_vertices_device(mh_scene* scene, float** verts);
/* Sine displacement used by the dynamic-scene workload: y = y0 + amp*sin(freq*x + phase)
 * from the base positions given at create time, then no refit. */
int mh_scene_displace_sine(mh_scene* scene, float amp, float freq, float phase, void* stream);

/* Per-vertex shading attributes of the mesh, Mesh::vertexNormals + normal_indices and
 * Mesh::texCoords + texture_indices (Model/Groups/Mesh.h:28-31): n_normals normals
 * (3 floats each) indexed by normal_idx (3 per triangle), n_texcoords texture coordinates
 * (3 floats each) indexed by tex_idx (3 per triangle; a triangle whose first entry is
 * MH_NO_TEXTURE_INDEX, Mesh::kNoTextureIndex, has none). A null array pair clears that
 * attribute. Host arrays, copied to the device. */
#define MH_NO_TEXTURE_INDEX 0xFFFFFFFFu
int mh_scene_set_attributes(mh_scene* scene, const float* normals, uint32_t n_normals,
                            const uint32_t* normal_idx, const float* texcoords,
                            uint32_t n_texcoords, const uint32_t* tex_idx);

/* Outputs of mh_shade, device arrays of n floats; any may be NULL. */
typedef struct mh_shading {
  float *px, *py, *pz; /* hit position o + d * t */
  float *nx, *ny, *nz; /* shading normal */
  float *gx, *gy, *gz; /* geometric normal */
  float *tu, *tv, *tw; /* texture coordinates */
} mh_shading;

/* What Manta's materials read from a hit, for every hit ray of a traced set, on the
 * device: RayPacket::computeHitPositions (Interface/RayPacket.cc:114-146), computeNormal
 * and computeGeometricNormal (KenslerShirleyTriangle.cc:381-420, WaldTriangle.cc:558-609,
 * by the scene's triangle kind) and computeTexCoords2 (KenslerShirleyTriangle.cc:85-123,
 * WaldTriangle.cc:131-169). hits must hold prim, t, u and v of a closest-hit trace of the
 * same rays; rays that missed get zeros. Asynchronous on `stream`. */
int mh_shade(const mh_scene* scene, const mh_rays* rays, const mh_hits* hits, int64_t n,
             const mh_shading* out, void* stream);

/* DynBVH::intersect (DynBVH.cc:101-404) on n rays resident in device memory.
 * Asynchronous on `stream` (NULL = the null stream). */
int mh_trace(const mh_scene* scene, const mh_rays* rays, const mh_hits* hits, int64_t n,
             uint32_t flags, const mh_trace_opts* opts, void* stream);

/* Same, from and to host memory; stages through the scene's device buffers and
 * returns when the results are on the host. */
int mh_trace_host(const mh_scene* scene, const mh_rays* rays, const mh_hits* hits, int64_t n,
                  uint32_t flags, const mh_trace_opts* opts);

/* No counterpart in the reference: mh_trace for callers that cannot allocate device
 * memory (HIP graph capture, serving loops or harnesses that forbid hipMalloc). Same
 * arguments and results as mh_trace; the single-ray kernel's traversal stack spill lives
 * in a fixed device array of MH_NOALLOC_SCRATCH_BYTES per device instead of a buffer the
 * scene allocates, and the call allocates nothing. The workgroup count is lowered until
 * every lane's spill fits (results are unchanged); a tree too deep for one workgroup per
 * compute unit to fit returns MH_E_UNSUPPORTED. All calls on one device share its array:
 * calls that overlap on one device (other streams, other threads) race. */
#define MH_NOALLOC_SCRATCH_BYTES (32u << 20)
int mh_trace_noalloc(const mh_scene* scene, const mh_rays* rays, const mh_hits* hits, int64_t n,
                     uint32_t flags, const mh_trace_opts* opts, void* stream);

/* ---- Persistent multi-batch frames ------------------------------------------------
 * No reference counterpart: Manta traces one 64-ray packet per call. A frame is a fixed
 * table of ray batches (uneven sizes, each closest-hit or any-hit, device-resident SoA
 * arrays) traced by one persistent launch in which every wave refills finished lanes
 * from a shared ray counter. Per-ray results equal mh_trace with packet_size 1
 * (DynBVH::intersect on one-ray packets). A frame can be captured once into a HIP graph
 * and replayed; the batch table may be rewritten between replays (same batch count). */
#define MH_MAX_BATCHES 256

typedef struct mh_batch {
  mh_rays rays;
  mh_hits hits;
  int64_t count;
  uint32_t flags; /* MH_CLOSEST_HIT or MH_ANY_HIT */
} mh_batch;

typedef struct mh_launch_config {
  int layout;        /* MH_LAYOUT_* */
  int stack_lds;     /* per-lane LDS stack ring: 8, 16 or 32 entries */
  int block;         /* lanes per workgroup: 64, 128 or 256 */
  int blocks_per_cu; /* persistent workgroups per compute unit */
  int refill;        /* idle lanes of a wave that trigger a refill, 1..63; 64 = whole-wave
                        chunks (a wave traces `chunk` consecutive rays 64 at a time) */
  int chunk;         /* rays a wave takes per grab from the shared counter, a multiple of 64
                        (its first grab is assigned statically); 0 = static stride over the
                        rays with no counter (whole-wave mode only) */
  int xcd_remap;     /* 1 = give each XCD (every 8th workgroup) a contiguous share of the
                        rays so that neighbouring rays share an L2 */
  int reverse;       /* 1 = hand rays out from the end of the table (last batch first, in
                        chunks whose rays ascend), so that a table whose costly batches come
                        last still ends on short rays; results are the same */
} mh_launch_config;

typedef struct mh_frame mh_frame;

int mh_frame_create(const mh_scene* scene, const mh_batch* batches, int n_batches,
                    mh_frame** out);
/* Rewrites the batch table (n_batches must equal the frame's), on `stream`. A change
 * between all closest-hit, all any-hit and mixed batches drops the captured graph. */
int mh_frame_set_batches(mh_frame* frame, const mh_batch* batches, int n_batches, void* stream);
int mh_frame_set_config(mh_frame* frame, const mh_launch_config* cfg); /* NULL = default */
int mh_frame_get_config(const mh_frame* frame, mh_launch_config* cfg);
int mh_frame_launch(mh_frame* frame, void* stream);
/* Captures one launch of the current configuration into a HIP graph (stream capture). */
int mh_frame_capture(mh_frame* frame, void* stream);
int mh_frame_replay(mh_frame* frame, void* stream);
/* Average milliseconds per launch over `iters` launches (graph replays when use_graph). */
int mh_frame_time(mh_frame* frame, int iters, int use_graph, void* stream, float* ms);
/* Sweeps layout x stack x block x blocks_per_cu x chunk x xcd_remap x refill, keeps the
 * fastest and stores it
 * in `cache_path` (JSON, may be NULL) keyed by the scene and batch mix; a cached entry
 * is reused without sweeping unless `force`. */
int mh_frame_autotune(mh_frame* frame, int iters, const char* cache_path, int force,
                      mh_launch_config* best, float* best_ms);
int mh_frame_destroy(mh_frame* frame);

/* ---- Ray generation on the device -----------------------------------------------
 * Writable SoA ray arrays (masked may be NULL for camera rays). */
typedef struct mh_ray_buffers {
  float* ox;
  float* oy;
  float* oz;
  float* dx;
  float* dy;
  float* dz;
  float* tmax;
  uint8_t* masked;
} mh_ray_buffers;

/* Pinhole camera, PinholeCamera::setup/makeRays (Model/Cameras/PinholeCamera.cc:167-230).
 * hfov_deg 0 derives the horizontal field of view from the aspect ratio. */
typedef struct mh_camera {
  float eye[3];
  float lookat[3];
  float up[3];
  float vfov_deg;
  float hfov_deg;
  int xres, yres; /* multiples of 8 */
  int spp;
  uint64_t seed;
} mh_camera;

/* xres*yres*spp normalised camera rays: sample-major, pixels in 8x8 tiles, row-major
 * inside a tile (one 64-ray packet per tile and sample). */
int mh_gen_camera_rays(const mh_camera* cam, const mh_ray_buffers* out, void* stream);
/* One shadow ray per primary ray toward `light` (PointLight::computeLight +
 * HardShadows, Model/Lights/PointLight.cc:31-75, Engine/Shadows/HardShadows.cc:87-124);
 * misses and back-facing rays are masked. Reads the scene's current vertices. */
int mh_gen_shadow_rays(const mh_scene* scene, const mh_rays* primary, const mh_hits* hits,
                       int64_t n, const float light[3], const mh_ray_buffers* out, void* stream);
/* One cosine-weighted diffuse bounce per primary ray; misses are masked. */
int mh_gen_bounce_rays(const mh_scene* scene, const mh_rays* primary, const mh_hits* hits,
                       int64_t n, uint64_t seed, const mh_ray_buffers* out, void* stream);

/* ---- Serving loop -----------------------------------------------------------------
 * No reference counterpart. A server keeps scenes resident and renders requests (a
 * camera over one of its scenes: closest-hit primary rays and, optionally, one shadow ray
 * and one diffuse bounce ray per primary ray) with continuous batching: each round of its
 * scheduler thread admits waiting requests in arrival order while their device buffers
 * fit the HBM budget, then queues one persistent batch trace per scene over a chunk of
 * every admitted request's remaining rays (scenes on separate streams, overlapping; the
 * next table of a scene is built while its launch runs, without waiting for it);
 * requests whose rays are all queued have their results copied to pinned host staging on
 * a copy stream that overlaps later launches, and their device buffers freed. Per-ray results equal mh_gen_camera_rays +
 * mh_trace (packet_size 1) + mh_gen_shadow_rays / mh_gen_bounce_rays + mh_trace. */
typedef struct mh_server mh_server;

typedef struct mh_server_opts {
  size_t hbm_budget;   /* device bytes of request buffers in flight [1 GiB] */
  size_t host_budget;  /* pinned host bytes of finished, uncollected results [hbm_budget/4] */
  int64_t chunk_rays;  /* rays of one ray set of a request traced per round [1<<20] */
  int max_active;      /* requests in flight [64, at most 128] */
  int shadows;         /* 1 = one shadow ray per primary ray toward `light` (any-hit) [0] */
  int bounces;         /* 1 = one diffuse bounce ray per primary ray (closest-hit), as
                          mh_gen_bounce_rays with seed = camera.seed + 1 [0] */
  float light[3];
  mh_launch_config launch; /* zero = default configuration: PAIR layout, LDS ring 8,
                              256-lane workgroups, 16 per CU, static stride */
} mh_server_opts;

typedef struct mh_request {
  int scene; /* index in the server's scene list */
  mh_camera camera;
} mh_request;

/* Result arrays of one request, xres*yres*spp entries each in camera ray order; any
 * pointer may be NULL. */
typedef struct mh_request_result {
  int32_t* prim;        /* primary closest hit: triangle id, MH_NO_HIT */
  float* t;
  uint8_t* occluded;    /* shadow ray of each primary ray (shadows) */
  int32_t* bounce_prim; /* bounce ray of each primary ray (bounces) */
  float* bounce_t;
} mh_request_result;

typedef struct mh_request_info {
  uint64_t id;
  int status;        /* MH_OK or the error that ended the request */
  int64_t rays;      /* rays traced (primary + shadow + bounce) */
  double submit_ms;  /* server clock */
  double admit_ms;
  double done_ms;    /* results in host memory */
} mh_request_info;

typedef struct mh_server_stats {
  uint64_t submitted, completed, rejected;
  int64_t rays;        /* rays traced since creation */
  double busy_ms;      /* wall time with requests in flight */
  int active, pending;
  size_t hbm_in_use;
  size_t host_in_use;
} mh_server_stats;

int mh_server_create(const mh_scene* const* scenes, int n_scenes, const mh_server_opts* opts,
                     mh_server** out);
/* Queues a request; returns at once. Requests larger than the budgets are rejected. The
 * scenes must not be changed while a request over them is in flight. */
int mh_server_submit(mh_server* server, const mh_request* req, uint64_t* id);
/* Waits for a request and copies its results out (out may be NULL); each id once. */
int mh_server_wait(mh_server* server, uint64_t id, const mh_request_result* out,
                   mh_request_info* info);
/* Cancels a request: a queued one at once, one in flight once the device work already
 * queued for it has passed (no further chunks are handed out). mh_server_wait then reports
 * MH_E_CANCELED. MH_E_STATE if it has already finished. */
int mh_server_cancel(mh_server* server, uint64_t id);
int mh_server_get_stats(mh_server* server, mh_server_stats* stats);
double mh_server_clock_ms(const mh_server* server);
/* Finishes the requests already submitted, then stops; results not yet collected are
 * discarded. */
int mh_server_destroy(mh_server* server);

/* Device memory helpers for callers without their own allocator. */
int mh_malloc(void** ptr, size_t bytes);
int mh_free(void* ptr);
int mh_memcpy_h2d(void* dst, const void* src, size_t bytes);
int mh_memcpy_d2h(void* dst, const void* src, size_t bytes);
/* Device-to-device copy by a streaming kernel (16-byte vector loads/stores) on `stream`. */
int mh_device_copy(void* dst, const void* src, size_t bytes, void* stream);
int mh_stream_create(void** stream);
int mh_stream_destroy(void* stream);
int mh_stream_sync(void* stream);
/* Elapsed milliseconds of `iters` back-to-back mh_trace calls measured with HIP events
 * recorded on `stream` (the stream the kernels run on). */
int mh_time_trace(const mh_scene* scene, const mh_rays* rays, const mh_hits* hits, int64_t n,
                  uint32_t flags, const mh_trace_opts* opts, void* stream, int iters, float* ms);

#ifdef __cplusplus
}
#endif

#endif /* MH_H_ */
