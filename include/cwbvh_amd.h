/*
 * cwbvh_amd.h -- C ABI of the MI355X (gfx950) ray-traversal engine.
 *
 * Two traversal paths over triangle meshes, each with closest-hit and any-hit:
 *   - CWBVH: 8-wide compressed wide BVH (quantized child boxes, octant-ordered
 *     child visits), persistent wave64 kernel with work fetching;
 *   - BVH2:  Aila-Laine while-while traversal with Woop triangle tests.
 *
 * Data contract (identical bytes to the reference engine):
 *   ray  = float4(origin.xyz, tmin), float4(dir.xyz, tmax)     32 B
 *          (reference: src/helper_math.h:1853-1857)
 *   hit  = float4(t, triAddr as int32 bits, u, v)               16 B
 *          (reference: src/helper_math.h:1859-1865)
 *          miss: t = ray tmax, triAddr = -1, u = v = 0.
 *          triAddr is the float4 index of the hit triangle's first Woop row;
 *          prim id = prim_idx[triAddr] (or written directly with
 *          CWBVH_TRACE_PRIM_ID).
 *   BVH2 node  = 4 float4 (64 B)     (reference: src/GPUBVHConverter.cpp:10-121)
 *   CWBVH node = 5 float4 (80 B), optionally padded to 8 float4 (128 B, one
 *                HBM/L2 line per node)   (reference: src/GPUBVHConverter.cpp:123-378)
 *   Woop triangle = 3 float4 (48 B)  (reference: src/GPUBVHConverter.cpp:380-403)
 *
 * Every function returns a cwbvh_status (0 = OK) unless stated otherwise.
 * Device pointers are plain HIP device addresses; "stream" arguments are
 * hipStream_t passed as void* (NULL = the null stream).  Work counters and
 * sort workspaces are kept per stream handle, so the special handle
 * hipStreamPerThread (a different stream on every host thread) is refused
 * with CWBVH_ERR_INVALID_ARGUMENT wherever a stream is taken.
 */
#ifndef CWBVH_AMD_H
#define CWBVH_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CWBVH_ABI_VERSION 2

/* ------------------------------------------------------------------ types */

typedef struct cwbvh_float4 { float x, y, z, w; } cwbvh_float4;

typedef struct cwbvh_ray {
    cwbvh_float4 origin_tmin;
    cwbvh_float4 dir_tmax;
} cwbvh_ray;

typedef struct cwbvh_hit {
    cwbvh_float4 t_triId_u_v;
} cwbvh_hit;

typedef enum cwbvh_status {
    CWBVH_OK = 0,
    CWBVH_ERR_INVALID_ARGUMENT = 1,
    CWBVH_ERR_HIP = 2,
    CWBVH_ERR_OUT_OF_MEMORY = 3,
    CWBVH_ERR_STACK_DEPTH = 4,     /* tree deeper than the traversal stack */
    CWBVH_ERR_NOT_BOUND = 5,       /* compat trace before bind */
    CWBVH_ERR_NO_DEVICE = 6,
    CWBVH_ERR_BUDGET = 7,          /* request larger than the memory budget */
    CWBVH_ERR_CANCELLED = 8,
    CWBVH_ERR_INTERNAL = 9,
    CWBVH_ERR_IO = 10              /* file cannot be opened, read or written, or its size is wrong */
} cwbvh_status;

enum {
    CWBVH_KIND_BVH2 = 0,
    CWBVH_KIND_CWBVH = 1
};

/* trace flags */
enum {
    CWBVH_TRACE_CLOSEST = 0,      /* closest hit */
    CWBVH_TRACE_ANY = 1,          /* any hit: stop at the first accepted triangle */
    CWBVH_TRACE_PRIM_ID = 2,      /* write prim_idx[triAddr] instead of triAddr */
    CWBVH_TRACE_STATS = 4         /* count visited nodes / tested triangles */
};

/* Largest traversal stack (entries) the kernels carry: LDS part + scratch spill. */
#define CWBVH_STACK_CAPACITY 64
#define CWBVH_BVH2_STACK_CAPACITY 64

/* ------------------------------------------------------------ diagnostics */

int32_t cwbvh_abi_version(void);
const char* cwbvh_status_string(int32_t status);
/* Detail message of the last failing call on this host thread. */
const char* cwbvh_last_error(void);

/* ------------------------------------------------------- meshes (host) --- */

typedef struct cwbvh_mesh cwbvh_mesh;

/* Copy caller arrays: vertices = float[3*nv], indices = int32[3*nt]. */
int32_t cwbvh_mesh_from_arrays(const float* vertices, int64_t num_vertices,
                               const int32_t* indices, int64_t num_triangles,
                               cwbvh_mesh** out);
int32_t cwbvh_mesh_view(const cwbvh_mesh* mesh, const float** vertices, int64_t* num_vertices,
                        const int32_t** indices, int64_t* num_triangles);
/* Axis-aligned bounds: lo[3], hi[3]. */
int32_t cwbvh_mesh_bounds(const cwbvh_mesh* mesh, float* lo3, float* hi3);
void cwbvh_mesh_free(cwbvh_mesh* mesh);

/* Seeded heightfield over [0,1]^2: nx*ny cells, 2 triangles each, height =
 * sum_{k<octaves} amp0*2^-k * sin(2*pi*2^k*(x cos th_k + y sin th_k) + ph_k). */
int32_t cwbvh_mesh_heightfield(int32_t nx, int32_t ny, int32_t octaves, float amp0,
                               uint64_t seed, cwbvh_mesh** out);
/* gx*gy*gz randomly rotated geodesic spheres (20*freq^2 triangles each) with
 * radius U[scale_min, scale_max] on a lattice of pitch `spacing`, plus a
 * 2-triangle ground square below the lattice reaching ground_margin lattice
 * extents (+ one pitch) beyond it on every side (ground_margin < 0: none). */
int32_t cwbvh_mesh_icosphere_grid(int32_t gx, int32_t gy, int32_t gz, int32_t frequency,
                                  float spacing, float scale_min, float scale_max,
                                  float ground_margin, uint64_t seed, cwbvh_mesh** out);
/* Triangle soup in Gaussian clusters inside the unit cube: cluster sigma
 * log-uniform in [sigma_min, sigma_max], edge lengths lognormal(mu, s). */
int32_t cwbvh_mesh_cluster_soup(int64_t num_triangles, int32_t num_clusters,
                                float sigma_min, float sigma_max, float edge_mu,
                                float edge_sigma, uint64_t seed, cwbvh_mesh** out);
/* The reference's scene file (src/main.cpp:31-45): int32 vertex count,
 * int32 triangle count, float3 positions, int3 indices, int32 material per
 * triangle (little-endian, packed).  The reader refuses files whose size does
 * not match the header (CWBVH_ERR_IO) and out-of-range indices; materials
 * are kept with the mesh.  The writer stores `materials` (nullable: the
 * mesh's own, else zeros). */
int32_t cwbvh_mesh_read_dat(const char* path, cwbvh_mesh** out);
int32_t cwbvh_mesh_write_dat(const cwbvh_mesh* mesh, const int32_t* materials, const char* path);
/* Per-triangle materials (count 0 unless the mesh came from a mesh file). */
int32_t cwbvh_mesh_materials(const cwbvh_mesh* mesh, const int32_t** materials, int64_t* count);

/* The reference's surfel file (src/main.cpp:52-64): int32 size x, int32 size
 * y, then size x * size y float4 world positions, float4 world normals and
 * float texel radii. */
typedef struct cwbvh_samples cwbvh_samples;
int32_t cwbvh_samples_read(const char* path, cwbvh_samples** out);
int32_t cwbvh_samples_write(const char* path, int32_t size_x, int32_t size_y, const cwbvh_float4* positions,
                            const cwbvh_float4* normals, const float* radius);
int32_t cwbvh_samples_view(const cwbvh_samples* samples, int32_t* size_x, int32_t* size_y,
                           const cwbvh_float4** positions, const cwbvh_float4** normals, const float** radius);
void cwbvh_samples_free(cwbvh_samples* samples);

/* Known-answer scene: triangle i lies in plane z = i with vertices
 * (-1.1,-1.1,i), (1.1,-1.1,i), (1.1,1.1,i). */
int32_t cwbvh_mesh_validation_planes(int32_t count, cwbvh_mesh** out);

/* --------------------------------------------------------- rays (host) --- */

/* AO rays from area-uniform surface points: num_points * rays_per_point rays,
 * origin offset `offset` along the geometric normal (front side = the side the
 * normal (v1-v0)x(v2-v0) points to), cosine-weighted directions. */
int32_t cwbvh_rays_ao_surface(const cwbvh_mesh* mesh, int64_t num_points,
                              int32_t rays_per_point, float offset, float tmin,
                              float tmax, uint64_t seed, cwbvh_ray* out);
/* Pinhole camera rays, row-major width*height, vertical fov in degrees. */
int32_t cwbvh_rays_pinhole(int32_t width, int32_t height, const float* eye3,
                           const float* target3, const float* up3, float fov_deg,
                           float tmin, float tmax, cwbvh_ray* out);
/* AO rays from primary hits: for every hit (prim id >= 0), rays_per_hit
 * cosine-weighted rays around the geometric normal facing the incoming ray.
 * prim_ids[i] is the mesh triangle of hit i (-1 = miss).  out must hold
 * n*rays_per_hit rays; *out_count receives the number written. */
int32_t cwbvh_rays_ao_from_hits(const cwbvh_mesh* mesh, const cwbvh_ray* primary,
                                const cwbvh_hit* hits, const int32_t* prim_ids, int64_t n,
                                int32_t rays_per_hit, float offset, float tmax,
                                uint64_t seed, cwbvh_ray* out, int64_t* out_count);
/* AO rays from surfels, as the reference generates them from its sample file
 * (src/ValidationKernels.cu:63-101): for texel i and k < rays_per_texel, a
 * cosine-weighted direction (concentric disk lifted to the hemisphere) in the
 * frame t1 = normalize(n x z) (n x y when that is shorter than 0.1),
 * t2 = normalize(t1 x n), and origin p + (n * radius) * 0.5.  The reference
 * uses tmin 0.01, tmax 1e20 and curand; the random stream here is the
 * library's own (seeded), so the same distribution, not the same rays.
 * out[i * rays_per_texel + k].  The host generator, the device generator
 * (cwbvh_rays_ao_from_samples_device) and the fused trace (cwbvh_trace_surfels)
 * produce the same bits. */
int32_t cwbvh_rays_ao_from_samples(const cwbvh_float4* positions, const cwbvh_float4* normals,
                                   const float* radius, int64_t num_texels, int32_t rays_per_texel,
                                   float tmin, float tmax, uint64_t seed, cwbvh_ray* out);
/* Diffuse secondary rays: area-uniform surface points, uniform sphere
 * directions, origin offset along the normal on the side of the direction. */
int32_t cwbvh_rays_sphere_from_surface(const cwbvh_mesh* mesh, int64_t num_rays, float offset,
                                       float tmin, float tmax, uint64_t seed, cwbvh_ray* out);
/* Known-answer rays for cwbvh_mesh_validation_planes(num_triangles): ray r
 * starts at (0.5, -0.5, k - 0.5), k = (r * 12345678) mod num_triangles computed
 * in 64-bit, direction normalize(-0.01, 0.5, 1), tmin 0, tmax 1e20. */
int32_t cwbvh_rays_validation(int64_t count, int32_t num_triangles, cwbvh_ray* out);
/* Fisher-Yates shuffle of a ray array (seeded). */
int32_t cwbvh_rays_shuffle(cwbvh_ray* rays, int64_t n, uint64_t seed);

/* ----------------------------------------------- BVH build + layouts --- */

typedef struct cwbvh_build_params {
    int32_t num_threads;        /* 0 = all hardware threads */
    int32_t sah_bins;           /* binned SAH bins per axis (8..256) */
    float bvh2_node_cost;       /* SAH cost of a BVH2 node visit */
    float bvh2_tri_cost;        /* SAH cost of a triangle test (BVH2 leaves) */
    int32_t bvh2_max_leaf;      /* max triangles per BVH2 leaf (1..32) */
    float cwbvh_node_cost;      /* collapse DP: cost of an 8-wide node */
    float cwbvh_tri_cost;       /* collapse DP: cost of a triangle (default 0.6) */
    int32_t cwbvh_node_stride;  /* float4 per CWBVH node: 5 (80 B) or 8 (128 B) */
    int32_t max_depth;          /* BVH2 depth limit (<= CWBVH_BVH2_STACK_CAPACITY-2) */
    /* Spatial splits (SBVH): a node whose best object split leaves children
     * overlapping by more than split_alpha x the root's surface area also
     * tries splitting references at bin planes (0 = never, the default; 1e-5
     * is the usual setting; max 1).  A split triangle sits at several Woop
     * addresses: a closest hit's t, u, v and triangle (prim_idx) do not
     * depend on the traversal order, the address reported for it can. */
    float split_alpha;
    /* extra triangle references spatial splits may add, as a fraction of the
     * triangle count (0..4) */
    float split_budget;
} cwbvh_build_params;

int32_t cwbvh_build_params_default(cwbvh_build_params* params);

typedef struct cwbvh_layout cwbvh_layout;

typedef struct cwbvh_layout_view {
    int32_t kind;                 /* CWBVH_KIND_* */
    int32_t node_stride;          /* float4 per node */
    const cwbvh_float4* nodes;
    int64_t nodes_f4;             /* float4 count of the node array */
    const cwbvh_float4* woop;
    int64_t woop_f4;              /* float4 count of the Woop array */
    const int32_t* prim_idx;      /* one int per Woop float4 (-1 = not a row 0) */
    int64_t prim_idx_len;
    int32_t max_depth;            /* deepest node level (root = 1) */
    int32_t stack_need;           /* traversal stack entries needed */
    int64_t num_nodes;
    int64_t num_leaves;
    int64_t num_tri_refs;
    double sah_cost;
} cwbvh_layout_view;

/* Build both layouts from one SAH tree (either output may be NULL). */
int32_t cwbvh_build_layouts(const cwbvh_mesh* mesh, const cwbvh_build_params* params,
                            cwbvh_layout** bvh2_out, cwbvh_layout** cwbvh_out);
/* Wrap caller arrays (e.g. layouts written by another converter). */
int32_t cwbvh_layout_from_arrays(int32_t kind, int32_t node_stride, const cwbvh_float4* nodes,
                                 int64_t nodes_f4, const cwbvh_float4* woop, int64_t woop_f4,
                                 const int32_t* prim_idx, int64_t prim_idx_len,
                                 cwbvh_layout** out);
int32_t cwbvh_layout_get_view(const cwbvh_layout* layout, cwbvh_layout_view* view);
/* Layout cache: write a layout to a file / read it back.  The file holds a
 * header, the node, Woop and prim-index arrays and a 64-bit FNV-1a checksum;
 * loading checks size and checksum and validates the tree like
 * cwbvh_layout_from_arrays, so a scene can be reloaded instead of rebuilt. */
int32_t cwbvh_layout_save(const cwbvh_layout* layout, const char* path);
int32_t cwbvh_layout_load(const char* path, cwbvh_layout** out);
/* Re-lay a CWBVH node array to another node stride (5 <-> 8 float4). */
int32_t cwbvh_layout_restride(const cwbvh_layout* in, int32_t node_stride, cwbvh_layout** out);
void cwbvh_layout_free(cwbvh_layout* layout);

/* Woop transform of one triangle (3 float4 rows, -0.0 in row0.x made +0.0;
 * degenerate triangles get rows that never report a hit). */
int32_t cwbvh_woopify(const float* v0, const float* v1, const float* v2, cwbvh_float4* out3);

/* ------------------------------------------------- device plumbing --- */

typedef struct cwbvh_device_info {
    char name[64];
    char arch[32];
    int32_t compute_units;
    int32_t wave_size;
    int64_t lds_per_block;
    int64_t l2_bytes;
    int64_t global_mem_bytes;
    int32_t clock_khz;
    int32_t mem_clock_khz;
    int32_t mem_bus_width;
} cwbvh_device_info;

int32_t cwbvh_device_count(int32_t* count);
int32_t cwbvh_get_device_info(int32_t device, cwbvh_device_info* info);
int32_t cwbvh_malloc(void** dptr, int64_t bytes);
int32_t cwbvh_free(void* dptr);
int32_t cwbvh_memcpy_htod(void* dst, const void* src, int64_t bytes);
int32_t cwbvh_memcpy_dtoh(void* dst, const void* src, int64_t bytes);
int32_t cwbvh_memcpy_dtod(void* dst, const void* src, int64_t bytes);
int32_t cwbvh_memset(void* dptr, int32_t value, int64_t bytes);
int32_t cwbvh_stream_create(void** stream);
int32_t cwbvh_stream_destroy(void* stream);
int32_t cwbvh_stream_synchronize(void* stream);
int32_t cwbvh_device_synchronize(void);
int32_t cwbvh_event_create(void** event);
int32_t cwbvh_event_destroy(void* event);
int32_t cwbvh_event_record(void* event, void* stream);
int32_t cwbvh_event_synchronize(void* event);
int32_t cwbvh_event_elapsed_ms(void* start, void* stop, float* ms);
/* Self-check of the ray setup's reciprocal (no counterpart in the reference):
 * over all 2^32 binary32 patterns x, d_counts[0] = patterns with |x| in
 * [2^-125, 2^125] whose v_rcp_f32 + one Newton step differs from the IEEE
 * quotient 1.0f / x (the kernels rely on 0), d_counts[1] = patterns in that
 * range (4,194,304,002), d_counts[2] = differing patterns outside it.
 * d_counts: 3 device uint64, cleared by the call; asynchronous on `stream`. */
int32_t cwbvh_rcp_check(uint64_t* d_counts, void* stream);
/* Calibration kernel: dst[i] = src[i] for n uint4 words (HBM copy roofline). */
int32_t cwbvh_copy_u4(void* dst, const void* src, int64_t n_u4, void* stream);
/* Calibration kernel for the VALU issue ceiling: `blocks` blocks of 256
 * threads run `iters` iterations of a fixed vector body on independent
 * register chains and store one float per thread to d_out (blocks * 256
 * floats).  mode 0: FP32 fma; 1: 32-bit integer add/xor/min3/max3; 2: the
 * CWBVH node-visit slab test (two 4-child quantized tests per iteration).
 * One launch's duration against its SQ_INSTS_VALU count is the issue rate of
 * that instruction mix with every SIMD busy. */
int32_t cwbvh_valu_probe(int32_t mode, int64_t iters, int32_t blocks, float* d_out, void* stream);

/* ------------------------------------------------- device scenes --- */

typedef struct cwbvh_scene cwbvh_scene;

/* Node layout of an uploaded CWBVH scene (cwbvh_scene_create_ex). */
#define CWBVH_SCENE_STRIDE_AUTO 0 /* the layout's stride, but a scene of 80 B nodes whose node
                                   * and Woop arrays exceed 256 MiB gets 128 B nodes */
#define CWBVH_SCENE_STRIDE_KEEP 1 /* the layout's stride */
#define CWBVH_SCENE_STRIDE_128 2  /* 128 B nodes */

/* Upload a host layout (owning device copy), CWBVH_SCENE_STRIDE_AUTO. */
int32_t cwbvh_scene_create(const cwbvh_layout* layout, void* stream, cwbvh_scene** out);
/* The same with a chosen node layout.  A 128 B node holds the 80 B node's
 * bytes and 48 B of zeros, so hits are bit-identical for every choice; a
 * scene that streams from HBM (larger than the 256 MB Infinity Cache) reads
 * one cache line per node instead of two for most nodes.  BVH2 layouts
 * ignore stride_mode. */
int32_t cwbvh_scene_create_ex(const cwbvh_layout* layout, void* stream, int32_t stride_mode,
                              cwbvh_scene** out);
/* float4 per node of the scene's device node array (BVH2 4, CWBVH 5 or 8). */
int32_t cwbvh_scene_node_stride(const cwbvh_scene* scene, int32_t* stride);
/* Wrap device arrays without taking ownership (bind semantics). prim_idx may
 * be NULL; stack_need = traversal stack entries the tree needs (0 = unknown).
 * The root node is read back once for the scene bounds. */
int32_t cwbvh_scene_wrap_device(int32_t kind, int32_t node_stride, const void* d_nodes,
                                const void* d_woop, const int32_t* d_prim_idx,
                                int32_t stack_need, cwbvh_scene** out);
int32_t cwbvh_scene_device_bytes(const cwbvh_scene* scene, int64_t* bytes);
int32_t cwbvh_scene_kind(const cwbvh_scene* scene, int32_t* kind);
void cwbvh_scene_destroy(cwbvh_scene* scene);

/* ------------------------------------------------------------ trace --- */

typedef struct cwbvh_trace_config {
    int32_t block_size;      /* 64, 128 or 256 threads */
    int32_t lds_depth;       /* stack entries per lane kept in LDS (1..stack capacity) */
    int32_t grid_blocks;     /* persistent grid (0 = fill the device) */
    int32_t fetch_chunk;     /* rays claimed per wave-level atomic (persistent grids; 0 = automatic) */
    int32_t refill_idle;     /* leave the inner loop once idle lanes exceed this ... */
    int32_t refill_budget;   /* ... for this many accumulated lane-iterations */
    int32_t postpone_pct;    /* 0 = off; loop_mode 0: postpone a triangle group when fewer than this %
                                of the lanes are active; loop_mode 2: hold triangle tests back until
                                this % of the active lanes hold triangles and no node to visit;
                                not for loop_mode 1 */
    int32_t prefetch;        /* 0 or 1: issue the next node's loads before the triangle loop */
    int32_t work_slices;     /* 1..8 contiguous ray ranges, one drained first per XCD (persistent grids) */
    int32_t persistent;      /* 1 = persistent grid with work fetching, 0 = one ray per thread,
                                2 = persistent from 786,432 (3 x 2^18) rays up.  CWBVH default 2.  BVH2 default 0
                                (the reference's kernel); its persistent grid hands rays to idle lanes
                                between while-while rounds and uses grid_blocks, fetch_chunk,
                                refill_idle (>= 0), refill_budget (>= 1) and work_slices.  BVH2
                                configurations keep loop_mode, prefetch, postpone_pct and occupancy 0 */
    int32_t loop_mode;       /* CWBVH inner-loop schedule: 0 = node visit + its whole triangle group
                                per iteration (reference order), 1 = node visit or one triangle test
                                per iteration, 2 = as 1, and lanes holding triangles also visit a
                                node, parking its triangle group on the stack; 3 (per-wave triangle
                                queue) was taken out and is refused with CWBVH_ERR_INVALID_ARGUMENT */
    int32_t occupancy;       /* register budget of the single-scene loop_mode 1 kernels (no prefetch,
                                no stats; rays from a buffer or from surfels): 0 = automatic
                                (7 waves per SIMD for launches of 2^22 rays or more), 1 = unconstrained,
                                7 = 7 waves per SIMD (72 VGPRs) */
    int32_t ray_order;       /* 0 = trace rays in the given order; 1 = a launch of 2^20 rays or more
                                first sorts them by direction octant and origin Morton code (as
                                cwbvh_morton_order, in the stream's workspace) and traces them in
                                that order, in pieces of at most 2^26 rays.  Hits still go to each ray's own slot, so results do
                                not change.  Applies to cwbvh_trace; ignored when an order is given,
                                by surfel traces, and by batches and the server (batches have their
                                own Morton pre-pass, the `morton` argument) */
} cwbvh_trace_config;

int32_t cwbvh_trace_config_default(int32_t kind, cwbvh_trace_config* cfg);

/* The configuration cwbvh_trace uses when cfg is NULL: the kind's default,
 * except for scenes larger than the 256 MB Infinity Cache, whose traversal is
 * bound by HBM traffic.  Such a scene gets the Morton ray pre-pass
 * (ray_order 1, in bounds read from its root node) and, for CWBVH, the
 * interleaved schedule with node prefetch (loop_mode 1, prefetch 1); one
 * whose root holds no child box gets the grouped schedule (loop_mode 0).
 * A large BVH2 scene gets the persistent grid (persistent 2, refill_idle 2,
 * refill_budget 8, work_slices 8).
 * When the pre-pass cannot allocate its workspace, the launch traces the rays
 * in the given order.  A CWBVH tree whose stack need exceeds the default LDS
 * depth gets lds_depth = its stack need, so its whole stack stays in LDS. */
int32_t cwbvh_trace_config_for_scene(const cwbvh_scene* scene, cwbvh_trace_config* cfg);

/* Traverse n rays.  d_counters (nullable, CWBVH_TRACE_STATS only) receives
 * uint64[2] = {visited nodes, tested triangles}, accumulated.  Ray and hit
 * buffers must be 16-byte aligned (CWBVH_ERR_INVALID_ARGUMENT otherwise). */
int32_t cwbvh_trace(const cwbvh_scene* scene, const cwbvh_ray* d_rays, cwbvh_hit* d_hits,
                    int64_t n, uint32_t flags, const cwbvh_trace_config* cfg,
                    uint64_t* d_counters, void* stream);
/* Same, with a ray permutation: ray i of the launch is d_rays[d_order[i]] and
 * its hit goes to d_hits[d_order[i]].  An entry outside [0, n) traces ray i
 * into d_hits[i] instead, so no access leaves the buffers. */
int32_t cwbvh_trace_ordered(const cwbvh_scene* scene, const cwbvh_ray* d_rays,
                            const int32_t* d_order, cwbvh_hit* d_hits, int64_t n,
                            uint32_t flags, const cwbvh_trace_config* cfg,
                            uint64_t* d_counters, void* stream);

/* Device generator of cwbvh_rays_ao_from_samples (same rays, bit for bit):
 * surfel arrays in device memory, rays written to d_out on `stream`. */
int32_t cwbvh_rays_ao_from_samples_device(const cwbvh_float4* d_positions, const cwbvh_float4* d_normals,
                                          const float* d_radius, int64_t num_texels, int32_t rays_per_texel,
                                          float tmin, float tmax, uint64_t seed, cwbvh_ray* d_out, void* stream);
/* AO rays generated inside the CWBVH traversal kernel (no ray buffer): ray
 * r = i * rays_per_texel + k is the ray cwbvh_rays_ao_from_samples makes for
 * surfel i, sample k.  Output is either one hit per ray in d_hits (as
 * cwbvh_trace writes them for the generated rays) or, with d_hits NULL, the
 * number of rays of each surfel that hit something added to d_texel_hits[i]
 * (uint32 per surfel; the caller clears it).  Flags as cwbvh_trace (PRIM_ID
 * only with d_hits).  Node prefetch is not used: cfg must have prefetch 0,
 * and with cfg NULL the scene's default is taken without it. */
int32_t cwbvh_trace_surfels(const cwbvh_scene* scene, const cwbvh_float4* d_positions,
                            const cwbvh_float4* d_normals, const float* d_radius, int64_t num_texels,
                            int32_t rays_per_texel, float tmin, float tmax, uint64_t seed, uint32_t flags,
                            cwbvh_hit* d_hits, uint32_t* d_texel_hits, const cwbvh_trace_config* cfg,
                            uint64_t* d_counters, void* stream);

/* 1 when a traversal on this stream overflowed its stack since the last
 * call (the flag is cleared); the affected rays may have missed hits. */
int32_t cwbvh_stack_overflowed(void* stream, int32_t* flag);

/* Diagnostic: loop-occupancy counters accumulated by CWBVH launches with
 * CWBVH_TRACE_STATS on this stream since the last call (then cleared):
 * out[6] = {iteration slots, active lane-iterations, node-code slots,
 * node steps, triangle-code slots, triangle tests}.  A slot is one lane of a
 * wave pass through that code (64 per pass), so uses / slots is the SIMD lane
 * utilization of each part of the inner loop. */
int32_t cwbvh_loop_profile(void* stream, uint64_t* out6);

/* Scene bounds (lo[3], hi[3]): the union of the root node's child boxes, read
 * at creation (the unit cube when the root holds no child box). */
int32_t cwbvh_scene_bounds(const cwbvh_scene* scene, float* lo3, float* hi3);

/* Ray order for coherent traversal: d_order[i] = index of the i-th ray when
 * rays are sorted stably by a 24-bit key, the octant of their direction (3
 * bits, sign of x highest) above the 21-bit Morton code of their origin
 * quantized to 7 bits per axis inside [lo, hi] (q = trunc(clamp((o - lo) *
 * (128 / (hi - lo)), 0, 127)) in float; x bit highest in each triple).
 * d_order holds n int32.  Stream-ordered (the call does not wait).  The
 * keys and sort temporaries (about 16 B per ray) live in a workspace kept per
 * device and stream and grown on demand, so repeated calls do not allocate;
 * cwbvh_stream_destroy releases a stream's workspace. */
int32_t cwbvh_morton_order(const cwbvh_ray* d_rays, int64_t n, const float* lo3, const float* hi3,
                           int32_t* d_order, void* stream);
/* Largest sort workspace one stream may hold, in bytes (0 = no limit, the
 * default).  Workspaces above a new limit are released.  A call that would
 * need more fails with CWBVH_ERR_OUT_OF_MEMORY; a cwbvh_trace with ray_order 1
 * then traces its rays in the given order instead, as when the device has no
 * room for the workspace. */
int32_t cwbvh_set_workspace_limit(int64_t bytes);

/* ------------------------------------------- multi-scene batches --- */

typedef struct cwbvh_batch_item {
    const cwbvh_scene* scene;    /* CWBVH scene (node stride 5 or 8) */
    const cwbvh_ray* d_rays;
    cwbvh_hit* d_hits;
    int64_t n;
} cwbvh_batch_item;

typedef struct cwbvh_batch cwbvh_batch;

/* Plan one persistent launch that traverses every item: a device job table
 * (one 64-byte record per item) and private fetch counters.  morton != 0 adds
 * a per-launch pre-pass that orders every item's rays as cwbvh_morton_order
 * does (key = item index : ray key, one radix sort over all rays). */
int32_t cwbvh_batch_create(const cwbvh_batch_item* items, int32_t count, uint32_t flags,
                           const cwbvh_trace_config* cfg, int32_t morton, cwbvh_batch** out);
/* Morton pre-pass (if enabled) + the persistent traversal, on `stream`. */
int32_t cwbvh_batch_launch(cwbvh_batch* batch, void* stream);
/* Capture the traversal launch into a hipGraph (instantiated once); the
 * Morton pre-pass, when enabled, stays a stream-ordered step in front of it. */
int32_t cwbvh_batch_capture(cwbvh_batch* batch, void* stream);
/* Morton pre-pass (if enabled) + replay of the captured graph. */
int32_t cwbvh_batch_replay(cwbvh_batch* batch, void* stream);
int32_t cwbvh_batch_info(const cwbvh_batch* batch, int64_t* total_rays, int32_t* num_jobs,
                         int64_t* device_bytes);
void cwbvh_batch_destroy(cwbvh_batch* batch);

/* ------------------------------------------------------------ serving --- */

/* Pinned host memory (for request rays / hits that the server copies). */
int32_t cwbvh_host_alloc(void** ptr, int64_t bytes);
int32_t cwbvh_host_free(void* ptr);

typedef struct cwbvh_server cwbvh_server;

typedef struct cwbvh_server_config {
    int64_t memory_budget;        /* device bytes for in-flight ray + hit buffers (48 B / ray) */
    int32_t max_inflight;         /* batches on the device at once (one stream each), 1..16 */
    int32_t max_batch_requests;   /* requests per batch launch */
    int64_t max_batch_rays;       /* rays per batch launch */
    int32_t serial;               /* 1: one request at a time with a plain per-scene launch */
    uint32_t flags;               /* CWBVH_TRACE_CLOSEST or CWBVH_TRACE_ANY (+ PRIM_ID) */
    cwbvh_trace_config trace;     /* kernel configuration; trace.lds_depth 0 (the default) =
                                     each launch keeps the deepest stack need of its trees in
                                     LDS (10 when a tree's need is unknown) */
} cwbvh_server_config;

int32_t cwbvh_server_config_default(cwbvh_server_config* cfg);

/* Most scenes one server holds (at creation and added later). */
#define CWBVH_SERVER_MAX_SCENES 4096
/* Resident scenes (CWBVH, not owned; they must outlive the server). */
int32_t cwbvh_server_create(const cwbvh_scene* const* scenes, int32_t num_scenes,
                            const cwbvh_server_config* cfg, cwbvh_server** out);
/* Queue one request: n host rays against scenes[scene]; the hits are copied
 * to host_hits when the request completes.  Host buffers should be pinned
 * (cwbvh_host_alloc) and must stay valid until completion.  Requests larger
 * than the memory budget fail with CWBVH_ERR_BUDGET. */
int32_t cwbvh_server_submit(cwbvh_server* s, int32_t scene, const cwbvh_ray* host_rays, int64_t n,
                            cwbvh_hit* host_hits, uint64_t* request_id);
/* Admit another resident scene while the server runs (CWBVH, not owned; it
 * must outlive the server).  *index receives its scene index, the next free
 * one; requests may name it as soon as the call returns.  The scene's build
 * and upload happen before this call, so they bound how fast a new scene
 * becomes servable, not the latency of the requests already admitted. */
int32_t cwbvh_server_add_scene(cwbvh_server* s, const cwbvh_scene* scene, int32_t* index);
/* Request kinds (cwbvh_server_submit, _submit_device, _submit_surfels). */
enum {
    CWBVH_REQUEST_HOST_RAYS = 0,    /* rays copied in, hits copied out (48 B per ray over PCIe) */
    CWBVH_REQUEST_DEVICE_RAYS = 1,  /* caller's device buffers used in place: no copies, no arena */
    CWBVH_REQUEST_SURFELS = 2       /* surfels in, per-surfel hit counts out (40 B per surfel) */
};
/* Queue a request on device-resident buffers: n rays at d_rays (16-byte
 * aligned device memory) traced against scenes[scene], hits written to
 * d_hits.  Nothing crosses PCIe and the memory budget is not charged; both
 * buffers must stay valid and untouched until completion.  Requests of this
 * kind join the same multi-scene launches as host-ray requests. */
int32_t cwbvh_server_submit_device(cwbvh_server* s, int32_t scene, const cwbvh_ray* d_rays, int64_t n,
                                   cwbvh_hit* d_hits, uint64_t* request_id);
/* An AO query from surfels (the reference's sample file contents,
 * src/main.cpp:52-64): num_texels host positions, normals and radii; the
 * server copies them into its arena (36 B per surfel, charged to the memory
 * budget with the 4 B count), generates rays_per_texel AO rays per surfel
 * inside the traversal kernel (as cwbvh_trace_surfels does, tmin / tmax /
 * seed as there) and writes to texel_hits[i] the number of surfel i's rays
 * that hit something.  Host arrays should be pinned and must stay valid until
 * completion.  A surfel request is one fused launch of its own on a free
 * slot; it overlaps with the other slots' launches. */
typedef struct cwbvh_surfel_request {
    const cwbvh_float4* positions;
    const cwbvh_float4* normals;
    const float* radius;
    int64_t num_texels;
    int32_t rays_per_texel;
    float tmin;
    float tmax;
    uint64_t seed;
    uint32_t* texel_hits;
} cwbvh_surfel_request;
int32_t cwbvh_server_submit_surfels(cwbvh_server* s, int32_t scene, const cwbvh_surfel_request* request,
                                    uint64_t* request_id);
/* Wait for a request (timeout_s < 0: no limit); *latency_ms = completion -
 * submission on the host clock. */
int32_t cwbvh_server_wait(cwbvh_server* s, uint64_t request_id, double timeout_s,
                          double* latency_ms);
/* Wait until every submitted request has completed.  After a device error
 * the server launches nothing more: every queued or in-flight request
 * completes with that status (wait and drain return it) and later
 * submissions are refused with it. */
int32_t cwbvh_server_drain(cwbvh_server* s);

typedef struct cwbvh_server_stats {
    int64_t completed;
    int64_t rays;
    int64_t batches;
    int32_t max_inflight_seen;    /* most batches on the device at once */
    int32_t max_batch_requests_seen;
    int64_t peak_bytes;           /* peak device bytes held by in-flight requests */
    double busy_s;                /* wall time with at least one batch in flight */
} cwbvh_server_stats;

int32_t cwbvh_server_get_stats(const cwbvh_server* s, cwbvh_server_stats* out);
void cwbvh_server_destroy(cwbvh_server* s);

/* Poisson load generator: num_requests requests at `rate` per second, rays per
 * request log-uniform in [min_rays, max_rays], scene uniform over the
 * server's scenes.  Request k's rays are a window of ray_pools[scene]
 * (pinned host arrays of pool_sizes[scene] rays); hits land in hit_pool
 * (pinned, hit_pool_size hits, used as a ring, so requests in flight together
 * may write the same slots: the generator measures load, its hits are not
 * results).  latencies_ms (nullable,
 * num_requests doubles) receives each request's latency in arrival order. */
typedef struct cwbvh_poisson_spec {
    double rate;
    int32_t num_requests;
    int64_t min_rays;
    int64_t max_rays;
    uint64_t seed;
} cwbvh_poisson_spec;

typedef struct cwbvh_poisson_result {
    double wall_s;
    double p50_ms;
    double p99_ms;
    double mean_ms;
    double max_ms;
    int64_t rays;
    double mrays_per_s;
} cwbvh_poisson_result;

int32_t cwbvh_server_run_poisson(cwbvh_server* s, const cwbvh_poisson_spec* spec,
                                 const cwbvh_ray* const* ray_pools, const int64_t* pool_sizes,
                                 cwbvh_hit* hit_pool, int64_t hit_pool_size,
                                 double* latencies_ms, cwbvh_poisson_result* out);
/* The same load on device-resident buffers (CWBVH_REQUEST_DEVICE_RAYS):
 * d_ray_pools[scene] and d_hit_pool are device arrays. */
int32_t cwbvh_server_run_poisson_device(cwbvh_server* s, const cwbvh_poisson_spec* spec,
                                        const cwbvh_ray* const* d_ray_pools, const int64_t* pool_sizes,
                                        cwbvh_hit* d_hit_pool, int64_t hit_pool_size, double* latencies_ms,
                                        cwbvh_poisson_result* out);
/* The same load as surfel requests (CWBVH_REQUEST_SURFELS): a request of r
 * rays (log-uniform as above) asks for r / rays_per_texel surfels, a window
 * of pools[scene] (pinned host arrays); counts land in count_pool (pinned,
 * count_pool_size entries, used as a ring).  Request k's seed is spec->seed
 * plus its window offset. */
typedef struct cwbvh_surfel_pool {
    const cwbvh_float4* positions;
    const cwbvh_float4* normals;
    const float* radius;
    int64_t num_texels;
} cwbvh_surfel_pool;
int32_t cwbvh_server_run_poisson_surfels(cwbvh_server* s, const cwbvh_poisson_spec* spec,
                                         const cwbvh_surfel_pool* pools, int32_t rays_per_texel, float tmin,
                                         float tmax, uint32_t* count_pool, int64_t count_pool_size,
                                         double* latencies_ms, cwbvh_poisson_result* out);

/* ---------------------------------------------------------- autotuner --- */

typedef struct cwbvh_tune_space {
    int32_t block_sizes[4];
    int32_t n_block;
    int32_t lds_depths[4];     /* 0 = the deepest stack need of the items' trees (the
                                  smallest depth at which the batch kernels keep the
                                  whole stack in LDS); 10 when a tree's need is unknown */
    int32_t n_lds;
    int32_t morton[2];         /* 0 = rays as given, 1 = Morton pre-pass */
    int32_t n_morton;
    int32_t prefetch[2];       /* node-prefetch distance 0 or 1 */
    int32_t n_prefetch;
    int32_t loop_modes[3];     /* cwbvh_trace_config.loop_mode values */
    int32_t n_loop_modes;      /* 0 = keep the default schedule */
} cwbvh_tune_space;

typedef struct cwbvh_tune_record {
    cwbvh_trace_config cfg;
    int32_t morton;
    float ms;                  /* median over reps: pre-pass + traversal */
} cwbvh_tune_record;

/* Time every point of `space` on the batch (median of `reps` launches after
 * one warm-up) and return the fastest.  records (nullable) receives up to
 * capacity entries; *n_records the number of points measured. */
int32_t cwbvh_autotune_batch(const cwbvh_batch_item* items, int32_t count, uint32_t flags,
                             const cwbvh_tune_space* space, int32_t reps,
                             cwbvh_tune_record* records, int32_t capacity, int32_t* n_records,
                             cwbvh_tune_record* best, void* stream);

/* ------------------------------------ reference entry points (drop-in) ---
 * Same names and argument meaning as the reference's
 * src/TraversalKernelCWBVH.h:3-11 and src/TraversalKernelBVH2.h:3-11:
 * bind stores device pointers (non-owning), trace is synchronous on the null
 * stream, writes one hit per ray and logs "<ms>MS, <rate>MRays/s (...)".
 * Errors are reported through cwbvh_last_status() instead of aborting. */
void rtBindCWBVHData(const cwbvh_float4* InBVHTreeNodes,
                     const cwbvh_float4* InTriangleWoopCoordinates,
                     const int* InMappingFromTriangleAddressToIndex);
void rtTraceCWBVH(cwbvh_ray* rayBuffer, cwbvh_hit* rayResultBuffer, int rayCount);
void rtBindBVH2Data(const cwbvh_float4* InBVHTreeNodes,
                    const cwbvh_float4* InTriangleWoopCoordinates,
                    const int* InMappingFromTriangleAddressToIndex);
void rtTraceBVH2(cwbvh_ray* rayBuffer, cwbvh_hit* rayResultBuffer, int rayCount);
/* Profiler scoping, the counterpart of the reference's cudaProfilerStart /
 * cudaProfilerStop (src/TraversalKernelCWBVH.cu:404,423, src/main.cpp:222,236):
 * resume / pause collection by a profiler that honours roctx (rocprofv3
 * --selected-regions --marker-trace).  rtTraceCWBVH / rtTraceBVH2 bracket
 * their launch with them, as the reference does, inside a roctx range named
 * after the call. */
int32_t cwbvh_profiler_start(void);
int32_t cwbvh_profiler_stop(void);
/* Status and kernel time (ms) of the last rt* call on this host thread. */
int32_t cwbvh_last_status(void);
float cwbvh_last_trace_ms(void);
/* 0 silences the timing line of rtTrace* (default 1). */
void cwbvh_set_log(int32_t enabled);

#ifdef __cplusplus
}
#endif

#endif /* CWBVH_AMD_H */
