/*
 * xyz_hip.h -- C ABI of the MI355X (gfx950) kd-tree closest-hit and
 * path-tracing library.
 *
 * Conventions
 *   - Every function returns XYZ_OK (0) or a negative XYZ_E_* code; the
 *     message of the last failure on the calling thread is xyz_last_error().
 *   - No exception crosses this boundary; no torch or HIP type appears in it.
 *     `stream` arguments are hipStream_t values passed as void* (NULL = the
 *     null stream).
 *   - Primitive ids are the caller's triangle order (the index in the vectors
 *     given to xyz_scene_create), i.e. the index a caller would have in the
 *     vector handed to KDTree::Build (reference xyz/geom/kdtree.cpp:207).
 *   - Scene descriptors and tree exports are HOST memory.  Ray, hit, film and
 *     gradient buffers of the compute entry points are DEVICE memory (any
 *     pointer the GPU can dereference), so the inputs of a hot loop stay in HBM.
 *   - A scene handle is immutable after creation; concurrent calls with one
 *     handle on different streams are allowed.
 */
#ifndef XYZ_HIP_H_
#define XYZ_HIP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  XYZ_OK = 0,
  XYZ_E_INVALID = -1,     /* bad argument (NULL, size, non-finite vertex, ...) */
  XYZ_E_NOMEM = -2,       /* host or device allocation failed */
  XYZ_E_DEVICE = -3,      /* a HIP call failed */
  XYZ_E_UNSUPPORTED = -4  /* a size beyond a documented limit */
};

/* BSDF kinds, same order as Bsdf::type (reference xyz/bsdf/bsdf.hpp:21-28). */
enum { XYZ_LAMBERT = 0, XYZ_MIRROR = 1, XYZ_GLASS = 2, XYZ_METAL = 3, XYZ_LIGHT = 4 };

typedef struct xyz_kdtree xyz_kdtree;
typedef struct xyz_scene xyz_scene;

/* One material.  Replaces the Bsdf subclasses' constructor arguments
 * (reference xyz/bsdf/lambert.hpp:12, metal.hpp:15, glass.cpp:119, light.hpp).
 *   rgb: albedo (LAMBERT, METAL) or emitted radiance L_e (LIGHT); the light
 *        power the integrator uses is Phi = L_e * pi * A_lights, so equal L_e on
 *        every light reproduces the reference's single global power
 *        (xyz/core/config.hpp:91-94).
 *   shininess: METAL lobe exponent.   ior: GLASS index (reference: 1.5).
 *   texture: -1, or the index of an RGB texture that replaces a LAMBERT albedo
 *        (nearest texel at the hit's uv; the reference stores uv but never
 *        reads it, xyz/geom/triangle.hpp:111-113). */
typedef struct {
  uint32_t type;
  float rgb[3];
  float shininess;
  float ior;
  int32_t texture;
} xyz_bsdf_desc;

/* Scene description (host memory).  Replaces Scene::Load + Scene::Build
 * (reference xyz/core/scene_mqo.cpp:547-777, xyz/core/scene.cpp:35).
 *   v0,v1,v2: 3*n_tris floats each.   n0,n1,n2: optional per-vertex shading
 *   normals (NULL = flat, reference triangle.cpp:10-33).  uv: optional
 *   6*n_tris.  bsdf: per-triangle material index (NULL = all 0).
 *   textures: tex_wh[2*i] = width, tex_wh[2*i+1] = height, texels are the
 *   textures' RGB floats concatenated, row 0 first.
 *   kd_max_bucket / kd_max_depth: KDTree(max_bucket=16, max_depth=16)
 *   (reference xyz/geom/kdtree.hpp:19-20); 0 selects the default 16.
 *   kd_builder: XYZ_KD_HOST (0) builds the tree on the host, XYZ_KD_DEVICE (1)
 *   on the current GPU (xyz_kdtree_build_device); both give the same tree. */
typedef struct {
  uint32_t n_tris;
  const float *v0, *v1, *v2;
  const float *n0, *n1, *n2;
  const float *uv;
  const uint32_t *bsdf;
  uint32_t n_bsdfs;
  const xyz_bsdf_desc *bsdfs;
  uint32_t n_textures;
  const uint32_t *tex_wh;
  const float *texels;
  uint32_t kd_max_bucket;
  uint32_t kd_max_depth;
  uint32_t kd_builder;
} xyz_scene_desc;

enum { XYZ_KD_HOST = 0, XYZ_KD_DEVICE = 1 };

typedef struct {
  uint64_t n_tris;
  uint64_t n_nodes;       /* interior + leaf nodes, 8 bytes each on the device */
  uint64_t n_leaves;      /* including empty leaves (KDNull) */
  uint64_t n_leaf_refs;   /* triangle references summed over leaves */
  uint32_t depth;         /* deepest interior level + 1 */
  uint32_t max_leaf;      /* largest leaf */
  float bmin[3], bmax[3]; /* root box (bmax is the builder's upper bound) */
  double build_seconds;   /* SAH build, host or device */
  uint64_t device_bytes;  /* HBM held by the scene (0 for a host tree) */
} xyz_scene_info;

/* ---- host-only kd-tree (no GPU needed) ------------------------------------
 * SAH build with the reference's exact split rules (xyz/geom/kdtree.cpp:207-357)
 * flattened to the device layout:
 *   nodes[2*i], nodes[2*i+1], pre-order, left child = i + 1;
 *     interior: nodes[2*i] = (right_child << 2) | axis, nodes[2*i+1] = split bits
 *     leaf:     nodes[2*i] = (count << 2) | 3,         nodes[2*i+1] = first ref
 *   leaf_prims[ref] = primitive id, each leaf's ids ascending (the reference
 *   keeps caller order in leaves). */
int xyz_kdtree_build(uint32_t n_tris, const float *v0, const float *v1, const float *v2,
                     uint32_t max_bucket, uint32_t max_depth, xyz_kdtree **out);
/* The same build on the current GPU (level by level: sorted events, segmented
 * scans, an atomic first-minimum per node and axis), returning the same tree
 * as xyz_kdtree_build for the same input; a split of +-0 may differ in the
 * sign bit only.  Host arrays in, host tree out; synchronous.  Replaces the
 * same KDTree::Build call (xyz/geom/kdtree.cpp:207-357). */
int xyz_kdtree_build_device(uint32_t n_tris, const float *v0, const float *v1, const float *v2,
                            uint32_t max_bucket, uint32_t max_depth, xyz_kdtree **out);
int xyz_kdtree_info(const xyz_kdtree *tree, xyz_scene_info *info);
int xyz_kdtree_export(const xyz_kdtree *tree, uint32_t *nodes, uint32_t *leaf_prims);
void xyz_kdtree_destroy(xyz_kdtree *tree);

/* ---- scenes ---------------------------------------------------------------- */
int xyz_scene_create(const xyz_scene_desc *desc, xyz_scene **out);
int xyz_scene_get_info(const xyz_scene *scene, xyz_scene_info *info);
void xyz_scene_destroy(xyz_scene *scene);

/* ---- Metasequoia scenes -----------------------------------------------------
 * Replaces Scene::Load (reference xyz/core/scene_mqo.cpp:547-777): parses a
 * Metasequoia "Format Text Ver 1.0" document with the reference's rules.
 *   materials -> BSDFs (scene_mqo.cpp:126-235): emi(1) -> light; col alpha
 *     != 1 -> glass; spc(1) with power(100) -> mirror, spc(1) with another
 *     positive power -> metal (shininess = power); otherwise Lambert.  Lambert
 *     and metal colours are a Macbeth patch named in the material name
 *     ("mcc-white", ...), else pow(col * spc, 2.2) through the CCIR 601-1
 *     RGB -> XYZ matrix, clamped to [0, 1]: film channels are CIE XYZ, as in
 *     the reference;
 *   vertices are scaled from centimetres to metres; objects with "shading 1"
 *     get angle-weighted smooth vertex normals, flat ones n0 = n1 = n2 = ng;
 *   every light emits the reference's global power Phi (mercury arc lamp,
 *     xyz/core/config.hpp:81-94) as L_e = Phi / (pi * total light area);
 *   the Scene chunk's pos / lookat / head / pich give the camera eye and at
 *     (default eye (0,1,4), at (0,1,0)).
 * Triangles keep file order (primitive id = order of faces in the file; the
 * reference's final Morton sort is not reproduced).  xyz_mqo_describe fills a
 * scene description whose arrays belong to the xyz_mqo handle (valid until
 * xyz_mqo_destroy), ready for xyz_scene_create. */
typedef struct xyz_mqo xyz_mqo;
typedef struct {
  uint32_t n_tris, n_bsdfs;
  uint32_t n_objects, n_smooth_objects, n_lights;
  uint32_t has_camera;    /* 1 if the document had a Scene chunk */
  float eye[3], at[3];    /* camera position and gazing point, metres */
  float light_power[3];   /* Phi, CIE XYZ */
  float surface_area;     /* of all visible triangles, m^2 */
  float light_area;
} xyz_mqo_info;
int xyz_mqo_load(const char *text, uint64_t len, xyz_mqo **out);
int xyz_mqo_load_file(const char *path, xyz_mqo **out);
int xyz_mqo_describe(const xyz_mqo *mqo, xyz_scene_desc *desc, xyz_mqo_info *info);
void xyz_mqo_destroy(xyz_mqo *mqo);

/* ---- closest hit -------------------------------------------------------------
 * Replaces Scene::FindIntersection / KDTree::FindIntersection
 * (reference xyz/core/scene.hpp:236-238,277-295; xyz/geom/kdtree.cpp:366-372)
 * for a batch of n rays.  org/dir: 3*n floats (xyz per ray).  Outputs (n each):
 *   prim: primitive id or -1 (the reference's nullptr),
 *   t:    Intersection::t_max() at the hit, +inf on a miss,
 *   f, g, back: barycentrics and back-face flag (valid on a hit; may be NULL),
 *   counters: optional 2*n uint32: kd nodes visited and triangles tested per
 *   ray (replaces Scene::n_FindIntersection, xyz/core/scene.hpp:232-234).
 * Interval [1e-5, +inf) as Intersection::Init (xyz/geom/intersection.hpp:11-14).
 * Results are bit-identical to the reference traversal in IEEE binary32. */
int xyz_closest_hit(xyz_scene *scene, uint32_t n, const float *org, const float *dir,
                    int32_t *prim, float *t, float *f, float *g, uint8_t *back,
                    uint32_t *counters, void *stream);

/* Instrumented xyz_closest_hit (no counterpart in the reference; it extends
 * the query counter of xyz/core/scene.hpp:232-234): the same results in prim
 * and t (t may be NULL), and work (device, 4 x uint64, ADDED to) receives the
 * launch's totals: kd nodes visited, triangle tests the reference performs,
 * triangle tests executed, cluster box tests executed.  The last two differ
 * from the second because leaf clusters that provably cannot hold an accepted
 * hit are skipped. */
/* Allocation-free xyz_closest_hit (no counterpart in the reference), for
 * callers that cannot allocate device memory: code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc.  Same arguments and results
 * as xyz_closest_hit; the kernel's work counters (its only scratch memory) live
 * in a fixed 2 KiB device array instead of pool memory, so the call makes no
 * allocation.  At most 2^31 rays per call (XYZ_E_UNSUPPORTED beyond).  Every
 * call on one GPU shares that GPU's array: calls that overlap on one GPU (on
 * different streams) race and must not be issued. */
int xyz_closest_hit_noalloc(xyz_scene *scene, uint32_t n, const float *org, const float *dir,
                            int32_t *prim, float *t, float *f, float *g, uint8_t *back,
                            uint32_t *counters, void *stream);

int xyz_closest_hit_work(xyz_scene *scene, uint32_t n, const float *org, const float *dir,
                         int32_t *prim, float *t, unsigned long long *work, void *stream);

/* Self-check of the device arithmetic under xyz_closest_hit (no counterpart in
 * the reference): the triangle test takes 1/det (xyz/geom/triangle.cpp:139) from the
 * hardware reciprocal and one Newton step when 1e-6 <= |det| < 2^126 (below
 * 1e-6 it rejects the triangle whatever the value).  This counts, on the
 * current device, the binary32 values of that range for which the result
 * differs from the correctly rounded IEEE quotient, into *mismatches (host);
 * 0 means the closest hit is exact on this device.  Synchronous. */
int xyz_check_reciprocal(unsigned long long *mismatches, void *stream);

/* Self-check of the Metal lobe's pow (no counterpart in the reference): the
 * reference evaluates std::pow on floats (xyz/stdafx.hpp:160 sampling cos
 * theta_h = xi^(1/s), xyz/bsdf/metal.cpp:101,146 the lobe density's nh^s);
 * the device evaluates it in double by a fixed series so that every build
 * gives the same bits.  out[i] = pow(x[i], y[i]) for n device floats. */
int xyz_eval_pow(uint32_t n, const float *x, const float *y, float *out, void *stream);

/* ---- path tracing -------------------------------------------------------------
 * Camera (reference ThinLensCamera, xyz/core/thinlenscamera.cpp:71-205):
 *   focal_length == 0: pinhole with an explicit vertical field of view, the
 *     lens-film distance is V = 0.0125 / tan(fov_y / 2);
 *   focal_length > 0: the reference's default thin lens (its defaults are
 *     focal_length 0.035, f_number 1.4): focus distance L = |eye - at|,
 *     V = L f / (L - f), lens radius f / (2 f_number); fov_y_deg is unused.
 *     The lens position takes two numbers of a separate random stream. */
typedef struct {
  float eye[3], at[3], up[3];
  float fov_y_deg;
  uint32_t width, height;
  float focal_length, f_number;
} xyz_camera;

/* Integrator settings (reference constants xyz/core/config.hpp:30,60-61 made
 * runtime):  max_depth D caps paths at D segments (reference
 * XYZ_CONFIG_kMaxRandomWalkDepth = D + 1; the reference's 16 is D = 15);
 * Russian roulette (threshold rr_threshold, reference 0.2) runs at vertices
 * j > rr_start (0 = every vertex, as the reference).  Samples
 * [sample_offset, sample_offset + spp) of every pixel are rendered; a sample's
 * random numbers depend only on (seed, pixel, sample index), so progressive
 * passes and resumes are exact. */
typedef struct {
  uint32_t spp;
  uint32_t sample_offset;
  uint32_t max_depth;
  uint32_t rr_start;
  float rr_threshold;
  uint32_t engine;        /* xyz_render_bdpt only: 0 = by the scene's size, 1 = a thread
                             per sample traces its queries, 2 = the lanes of a wave trace
                             their subpaths' rays together and the connections' queries
                             are queued and traced together; the other renderers ignore it */
  uint64_t seed;
} xyz_render_params;

typedef struct {
  uint64_t paths;          /* camera paths traced */
  uint64_t queries;        /* closest-hit queries (extension + shadow) */
  double seconds;          /* device time of the call (HIP events) */
} xyz_render_stats;

/* Replaces PathTracer::run's pixel loop around Scene::EvalutePathTracing
 * (reference xyz/integrator/integrator.cpp:119-128,
 * xyz/core/scene_pathtracing.cpp:11-144).  film: device float[width*height*3],
 * row 0 = bottom image row (as the reference film), receives += the SUM of the
 * samples' radiance per pixel (divide by the total spp for the mean, as
 * ImageFilm::SaveAsPfm does, xyz/core/imagefilm.cpp:153-206).
 * Wavefront: camera, extend, shade/NEE, shadow and scatter kernels with
 * per-bounce compaction.  stats may be NULL. */
int xyz_render(xyz_scene *scene, const xyz_camera *camera, const xyz_render_params *params,
               float *film, xyz_render_stats *stats, void *stream);

/* Bidirectional path tracing: replaces BidirectionalPathTracer::run's pixel loop
 * around Scene::EvaluateBidirectionalPathTracing (reference
 * xyz/integrator/integrator.cpp:1025-1081,
 * xyz/core/scene_bidirectionalpathtracing.cpp:8-522), on the same hit kernel.
 * Each sample of a pixel adds its own colour to film; its light-tracing
 * (t = 1) contributions go to the pixels they land on, in splat_film, or in
 * film when splat_film is NULL (both device float[width*height*3], row 0 =
 * bottom, SUMS: divide by the total spp).  The estimator is the reference's,
 * with its multiple-importance weights as it computes them (a connection's
 * density written into a path vertex stays there for later strategies), its
 * Lambertian connection factors and its pinhole eye path (a thin-lens camera
 * contributes its lens-film distance and focal length only).  Russian
 * roulette is the BSDFs' (albedo), so rr_start and rr_threshold are unused;
 * max_depth <= 16.  Random numbers: Philox streams 2 (light path) and 3 (eye
 * path), keyed like xyz_render's.  Splats are float atomics, so splat sums
 * may differ in the last bits between runs; own colours do not. */
int xyz_render_bdpt(xyz_scene *scene, const xyz_camera *camera, const xyz_render_params *params,
                    float *film, float *splat_film, xyz_render_stats *stats, void *stream);

/* ------------------------------------------- importance-driven path tracing */
typedef struct xyz_importon_map xyz_importon_map;

typedef struct {
  uint64_t n_importons;
  uint32_t n_emit;       /* light paths traced (0 for a map of caller importons) */
  uint32_t grid[3];      /* cells per axis of the query grid */
  float cell;            /* cell width (>= radius) */
  float radius;          /* search radius: 0.1 (reference CONFIG_SEARCH_RADIUS) */
  double seconds;        /* wall time of the build */
} xyz_importon_info;

/* Replaces Scene::CreateLightImportanceMap (reference
 * xyz/core/scene_photonmap.cpp:272-334) with Bsdf::BoundsImportance
 * (xyz/bsdf/lambert.cpp:16-40, mirror.cpp:18-48, glass.cpp:16-91,
 * metal.cpp:22-66, light.cpp:11-16): n_emit light paths (reference
 * CONFIG_EMITED_SIZE = 100000) of at most max_depth + 1 segments (the
 * reference's XYZ_CONFIG_kMaxRandomWalkDepth) leave an importon -- hit
 * position and incoming direction -- at every Lambert hit reached through a
 * specular bounce after the first segment.  Light paths are traced on the
 * device; random numbers come from Philox stream 4 keyed by (path index,
 * seed).  The importons keep path order, then vertex order, and are indexed
 * by a uniform grid for the fixed-radius query.  Synchronous.  n_emit <=
 * 2^28; max_depth in [1, 255].  A map refers to its scene: destroy it before
 * the scene. */
int xyz_importon_map_create(xyz_scene *scene, uint32_t n_emit, uint32_t max_depth, uint64_t seed,
                            xyz_importon_map **out);
/* A map of caller importons: host float[6*n], position then incoming
 * direction per importon (no counterpart in the reference, which only builds
 * its map from light paths). */
int xyz_importon_map_from_host(xyz_scene *scene, uint64_t n, const float *importons,
                               xyz_importon_map **out);
int xyz_importon_map_info(const xyz_importon_map *map, xyz_importon_info *info);
/* host float[6 * n_importons], in map order */
int xyz_importon_map_export(const xyz_importon_map *map, float *importons);
/* ImportanceQuery (reference xyz/core/photonmap.cpp:321-363) at n vertices:
 * basis = device float[12*n] (position, shading normal, tangent, binormal);
 * cdf = device float[17*n], the cumulative 2 x 8 histogram of cos^2 theta and
 * azimuth octant over the importons within the radius, 16 per bin to start
 * with (cdf[0] = 0). */
int xyz_importance_query(const xyz_importon_map *map, uint32_t n, const float *basis, float *cdf,
                         void *stream);
void xyz_importon_map_destroy(xyz_importon_map *map);

/* Importance-driven path tracing: replaces the pixel loop around
 * Scene::EvaluteImportanceDrivenPathTracing (reference
 * xyz/core/scene_importancedrivenpathtracing.cpp:11-179).  xyz_render's
 * estimator, except at a Lambert vertex reached through a specular bounce
 * (not the first hit): there the importance histogram of the map's importons
 * within 0.1 (Lambert::CalculateWeightedDirectLighting and
 * NextScatteringDirection with importons, xyz/bsdf/lambert.cpp:189-259)
 * scales the NEE weight's implicit density and chooses the next direction (a
 * bin from the cdf, then uniform in the bin's cos^2 theta and azimuth), and
 * the path power is divided by the bin's density.  Same film contract as
 * xyz_render; map must be built for this scene on the current device. */
int xyz_render_idpt(xyz_scene *scene, const xyz_importon_map *map, const xyz_camera *camera,
                    const xyz_render_params *params, float *film, xyz_render_stats *stats,
                    void *stream);

/* Path-replay adjoint of xyz_render (no counterpart in the reference).  Given
 * d_film = dLoss/d(film SUM) (device float[w*h*3]), adds dLoss/d(texel) to
 * d_texels (device, laid out like the scene's texels) and dLoss/dL_e to
 * d_emission (device float[3*n_bsdfs]; lights only).  Paths are re-traced from
 * their counter-based seeds, so nothing is stored between the passes. */
int xyz_render_backward(xyz_scene *scene, const xyz_camera *camera,
                        const xyz_render_params *params, const float *d_film, float *d_texels,
                        float *d_emission, void *stream);

/* ---- batched rendering (no counterpart in the reference) ---------------------
 * Renders n independent (scene, camera, params, film) jobs in one call, with
 * one of two engines:
 *   engine 0: one persistent launch over a global work queue of pixel tiles
 *             (tile_pixels, blocks_per_cu, block_threads; use_graph captures
 *             the launch in a hipGraph that later calls replay);
 *   engine 1: the wavefront kernels of xyz_render over all jobs at once: the
 *             jobs' samples share waves of up to 2^25 paths, so that small jobs
 *             do not pay per-bounce costs on their own (the other fields are
 *             unused).
 * Films are bit-identical to xyz_render's with either engine.  tuning may be
 * NULL (engine 1) or come from xyz_batch_autotune. */
typedef struct {
  uint32_t tile_pixels;    /* engine 0: pixels per work item */
  uint32_t blocks_per_cu;  /* engine 0: persistent workgroups per compute unit */
  uint32_t block_threads;  /* engine 0: threads per workgroup */
  uint32_t use_graph;      /* engine 0: capture the launch in a hipGraph and replay it */
  uint32_t engine;         /* 0: persistent megakernel, 1: shared wavefront */
} xyz_batch_tuning;

int xyz_render_batch(uint32_t n, xyz_scene *const *scenes, const xyz_camera *cameras,
                     const xyz_render_params *params, float *const *films,
                     const xyz_batch_tuning *tuning, xyz_render_stats *stats, void *stream);

/* Times candidate tunings on the given jobs (films are overwritten) and writes
 * the fastest to *best. */
int xyz_batch_autotune(uint32_t n, xyz_scene *const *scenes, const xyz_camera *cameras,
                       const xyz_render_params *params, float *const *films,
                       xyz_batch_tuning *best, void *stream);

/* ---- misc ---------------------------------------------------------------------- */
const char *xyz_last_error(void);
const char *xyz_version(void);

#ifdef __cplusplus
}
#endif
#endif /* XYZ_HIP_H_ */
