/*
 * xenodon_amd C-ABI: MI355X (gfx950) volume ray traversal.
 *
 * Drop-in boundary for Xenodon's Vulkan compute dispatch and headless
 * backend.  Plain pointers and sizes only; every call returns an xd_status
 * (0 = XD_OK) and leaves a thread-local message for xd_last_error().  The C++
 * host layer rethrows it as the reference's Error so CLI text is unchanged
 * (src/core/Error.h:8-13, src/main.cpp:161-195).
 *
 * Threading: a context owns one HIP stream on one device and may be driven
 * by one host thread at a time (the reference renders from one thread and
 * waits on a fence per frame, src/backend/headless/HeadlessOutput.cpp:90-93).
 */
#ifndef XENODON_AMD_H
#define XENODON_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#define XD_API __attribute__((visibility("default")))
#else
#define XD_API
#endif

typedef enum {
    XD_OK = 0,
    XD_ERR_INVALID_ARGUMENT = 1,
    XD_ERR_DEVICE = 2,
    XD_ERR_OUT_OF_MEMORY = 3,
    XD_ERR_INCOMPATIBLE = 4, /* traversal vs volume kind, main_loop.cpp:109-126 */
    XD_ERR_FORMAT = 5,       /* malformed volume data */
    XD_ERR_IO = 6,
    XD_ERR_REJECTED = 7,     /* admission refused (serving) */
    XD_ERR_INTERNAL = 8
} xd_status;

/* Traversal ids in the order of SHADER_OPTIONS (src/main_loop.cpp:37-43). */
typedef enum {
    XD_DDA = 0,       /* resources/dda.comp        requires a grid volume */
    XD_SVO_NAIVE = 1, /* resources/svo_naive.comp  requires an SVO volume */
    XD_ESVO = 2,      /* resources/esvo.comp       requires an SVO volume */
    XD_SVO_DF = 3,    /* resources/svo_df.comp     requires an SVO volume */
    XD_SVO_ROPE = 4   /* resources/svo_rope.comp   SVO made by `convert --rope`;
                         on other trees it integrates the first leaf only,
                         exactly as the reference shader does */
} xd_traversal;

typedef enum { XD_VOLUME_GRID = 0, XD_VOLUME_SVO = 1 } xd_volume_kind;

/* Push constant Camera, 48 B (resources/common.glsl:6-10; host fill
 * src/render/Renderer.cpp:58-64).  translation is in unit-cube space already
 * divided by voxel_ratio; the fourth components are ignored. */
typedef struct {
    float forward[4];
    float up[4];
    float translation[4];
} xd_camera;

/* vk::Rect2D */
typedef struct {
    int32_t offset[2];
    uint32_t extent[2];
} xd_rect;

/* Uniform buffer (resources/common.glsl:12-33; host fill
 * src/render/Renderer.cpp:236-268, src/main_loop.cpp:197-201).
 * model_dim = grid dimensions, or the SVO side on all three axes. */
typedef struct {
    xd_rect output_region;
    xd_rect display_region;
    float voxel_ratio[4];
    uint32_t model_dim[4];
    float emission_coeff;
} xd_uniforms;

typedef struct xd_context xd_context;
typedef struct xd_volume xd_volume;
typedef struct xd_frame xd_frame;

typedef struct {
    int32_t kind;          /* xd_volume_kind */
    uint32_t max_depth;    /* SVO: deepest reachable node; grid: 0 */
    uint64_t dims[3];      /* grid dimensions, or side x3 for an SVO */
    uint64_t side;         /* SVO side (power of two); grid: max(dims) */
    uint64_t node_count;   /* SVO nodes; grid: 0 */
    uint64_t device_bytes; /* HBM held by this volume */
    uint32_t is_rope;      /* leaves carry ropes (all leaf children[6..7]==0 and
                              some rope != 0) */
    uint32_t depth_fields_exact; /* SVO: every internal node's depth field equals
                                    its depth in the tree (converter output) */
} xd_volume_info;

/* ---- errors / devices --------------------------------------------------- */

/* Message of the last failing call on this thread ("" if none). */
XD_API const char* xd_last_error(void);
XD_API int xd_device_count(int* count);
/* Name, PCI address (domain:bus:device.function, the fields of
 * VkPhysicalDevicePCIBusInfoPropertiesEXT that src/sysinfo.cpp:53-59 prints)
 * and memory size of a device (xenodon sysinfo, src/sysinfo.cpp:8-80).  Any
 * output pointer may be null. */
XD_API int xd_device_describe(int device_index, char* name, size_t name_len, uint32_t* pci_domain, uint32_t* pci_bus,
                              uint32_t* pci_device, uint32_t* pci_function, uint64_t* total_bytes);

/* ---- context (replaces Device/Queue/CommandPool, src/graphics/) ---------- */

/* device_index is the headless config's vkindex
 * (src/backend/headless/HeadlessConfig.cpp:5-16). */
XD_API int xd_open(int device_index, xd_context** out);
/* Waits for the context's work, then frees the context together with every
 * volume and frame still alive on it; their handles become invalid. */
XD_API int xd_close(xd_context* ctx);
/* Per-wavefront pixel tile of subsequent launches: tile_w = 8 (8x8 tile),
 * 16 (16x4) or 32 (32x2).  Results do not depend on it; speed does. */
XD_API int xd_context_set_tile(xd_context* ctx, uint32_t tile_w);
/* Named integer options of a context:
 *   "octree_layout"  0 (default) derive the tagged device layout for trees of
 *                    at most 2^31 nodes (31-bit indices above 2^30); 1 keep
 *                    the .svo node words as they are (the layout used above
 *                    2^31 nodes); 2 the tagged layout with 31-bit indices
 *                    whatever the size.  Applies to volumes created
 *                    afterwards; results are identical.
 *   "xcd_chunks"     plain launches (xd_render): 0 (default) leaves the
 *                    image's pixel blocks in dispatch order, so neighbouring
 *                    blocks run on different XCDs; N in [1, 64] cuts the
 *                    blocks into 8N runs dealt to the 8 XCDs in turn (N = 1:
 *                    one contiguous band per XCD).  Results are identical. */
XD_API int xd_context_set_option(xd_context* ctx, const char* name, int64_t value);

/* ---- volumes (replaces Dda/SvoRaytraceResources, src/render/) ------------ */

/* Copies an RGBA8 grid (x fastest, then y, then z; src/model/Grid.h:50-56)
 * into one HBM allocation (no 4 GiB cap).  Replaces
 * DdaRaytraceResources::DdaRaytraceResources (DdaRaytraceAlgorithm.cpp:15-97). */
XD_API int xd_volume_from_grid(xd_context* ctx, const void* rgba8, uint64_t dim_x, uint64_t dim_y, uint64_t dim_z,
                               xd_volume** out);
/* Copies node_count nodes of 10 little-endian u32 words each (children[8],
 * color, is_leaf_depth; src/model/Octree.h:35-45) into one HBM allocation.
 * Rejects trees with cycles or out-of-range children (XD_ERR_FORMAT).
 * Replaces SvoRaytraceResources::SvoRaytraceResources
 * (SvoRaytraceAlgorithm.cpp:12-48). */
XD_API int xd_volume_from_svo(xd_context* ctx, const void* nodes, uint64_t node_count, uint64_t side,
                              xd_volume** out);
XD_API int xd_volume_info_get(const xd_volume* vol, xd_volume_info* out);
/* HBM bytes a volume will hold once created (xd_volume_info.device_bytes),
 * computed on the host without touching the device: kind XD_VOLUME_GRID uses
 * the dimensions, XD_VOLUME_SVO the node count.  ctx may be NULL (default
 * octree layout).  Used to admit volumes by footprint before allocating. */
XD_API int xd_volume_footprint(const xd_context* ctx, int kind, uint64_t dim_x, uint64_t dim_y, uint64_t dim_z,
                               uint64_t node_count, uint64_t* bytes);
XD_API int xd_volume_free(xd_volume* vol);

/* ---- frames (replaces HeadlessOutput's render target) -------------------- */

XD_API int xd_frame_create(xd_context* ctx, uint32_t width, uint32_t height, xd_frame** out);
XD_API int xd_frame_free(xd_frame* frame);
/* Device address of the frame's RGBA8 pixels (row pitch = width). */
XD_API int xd_frame_device_ptr(xd_frame* frame, void** out);

/* ---- plain render: one launch per output per frame ----------------------- */

/* Renders output_region of the display into frame (whose size must be the
 * region's extent), waits for completion and returns the launch's own
 * duration in ms measured with HIP events around the launch, the semantics
 * of RenderStatsCollector (src/render/RenderStats.cpp:46-83).  Replaces
 * Renderer::render (src/render/Renderer.cpp:55-103) + the fence wait. */
XD_API int xd_render(xd_context* ctx, const xd_volume* vol, int traversal, const xd_camera* cam,
                     const xd_uniforms* uniforms, xd_frame* frame, double* dispatch_ms);

/* Strided copy of the frame into host memory (stride in pixels, 0 = width),
 * replacing HeadlessOutput::download (HeadlessOutput.cpp:95-136). */
XD_API int xd_frame_download(xd_frame* frame, uint32_t* dst, uint64_t stride_px);

/* Number of 32-bit volume words the reference algorithm reads for this
 * frame (the roofline's algorithmic bytes = 4*words + 4*pixels).  Runs an
 * instrumented variant of the kernel with every shortcut disabled. */
XD_API int xd_count_words(xd_context* ctx, const xd_volume* vol, int traversal, const xd_camera* cam,
                          const xd_uniforms* uniforms, uint64_t* words);

/* ---- batched form: one persistent launch over many views ---------------- */

/* One output of a batch: a volume, a traversal, the camera push constant and
 * the uniforms of that output (output_region.extent = the frame's size). */
typedef struct {
    const xd_volume* volume;
    int32_t traversal;
    xd_camera camera;
    xd_uniforms uniforms;
    xd_frame* frame;
} xd_view;

/* Launch shape of a batch: per traversal the wavefront tile (8 = 8x8,
 * 16 = 16x4, 32 = 32x2 pixels), the tiles a wavefront takes per queue grab
 * (1..8), the register budget of the persistent kernels (0 = compiler's
 * choice, 8 = capped for 8 waves per SIMD), and whether one persistent kernel
 * drains every traversal's tiles (class_kernels = 0) or each traversal gets
 * its own kernel, with registers and LDS of that traversal only, run in turn
 * (1; the DDA kernel then also loads the next steps' macro-cell words while
 * the current texels are gathered; a batch of one traversal always runs that
 * traversal's own kernel), and the
 * tile order: 0 = view by view, rows of tiles; 1 = views of the same volume,
 * traversal and size interleaved tile by tile, so that the same screen tile
 * of, e.g., consecutive camera frames is traced at the same time through the
 * same L2; 2, 3, 4 = view by view in squares of 8x8, 4x4, 16x16 tiles, so the
 * tiles in flight cover a compact region of the image (a narrow cone of
 * rays); 5 = interleaved as 1, in squares of 8x8 tiles.
 * Results never depend on it.  A new batch starts at 8x8 tiles for every
 * traversal, one tile per grab, 8 waves per SIMD, one kernel for every
 * traversal with mixed classes, squares of 8x8 tiles. */
typedef struct {
    uint32_t tile_w[5];
    uint32_t tiles_per_grab;
    uint32_t waves_per_simd;
    uint32_t class_kernels;
    uint32_t interleave;
    uint32_t mix_classes; /* all-traversal kernel: 1 = its waves start on different
                             traversals, so each CU runs a mix of them at once;
                             2 = as 1, and the XCDs are shared out between the
                             traversals in proportion to their work, so that
                             each XCD's L2 serves one or two traversals' data
                             (batches only; the served form reads 2 as 1) */
    uint32_t rays_per_lane[5]; /* per traversal, rays each lane traces per tile column
                                  (1, 2 or 4; 0 = 1): that many rows of wave
                                  tiles are queued together and taken in one
                                  grab, then traced in turn by the same
                                  wavefront (64 * rays_per_lane rays per grab).
                                  The served form reads it as 1 */
} xd_batch_tuning;

typedef struct xd_batch xd_batch;

XD_API int xd_batch_create(xd_context* ctx, xd_batch** out);
/* Sets the batch's views (at most 4096; volumes and frames of the same
 * context).  Uploads the view and tile tables; the captured hipGraph is reused
 * unless the launch shape changes. */
XD_API int xd_batch_set_views(xd_batch* batch, const xd_view* views, uint32_t count);
XD_API int xd_batch_set_tuning(xd_batch* batch, const xd_batch_tuning* tuning);
XD_API int xd_batch_get_tuning(const xd_batch* batch, xd_batch_tuning* tuning);
/* Coordinate search from the current tuning on the current views (best of
 * `trials` launches per setting, 0 = 3): the tile shape of each traversal on
 * that traversal's views alone, then class mixing, tile order, register
 * budget, per-class kernels and grab size on the whole batch.  A field
 * changes only for a setting at least 2 % faster than the one it has. */
XD_API int xd_batch_autotune(xd_batch* batch, uint32_t trials, xd_batch_tuning* chosen);
/* Replays the batch's graph and waits.  *ms is the HIP-event time of the
 * whole launch; *rays the number of pixels rendered. */
XD_API int xd_batch_render(xd_batch* batch, double* ms, uint64_t* rays);
XD_API int xd_batch_free(xd_batch* batch);
/* Allocation-free batched render.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Renders `count` views
 * (volumes and frames of ctx, as xd_batch_set_views takes them) in one
 * launch of the batched kernels under `tuning` (NULL = the defaults of a new
 * batch) and waits; *ms is the HIP-event time of the launch, *rays the pixels
 * rendered.  Frames are bit-identical to xd_batch_render of the same views.
 * The view and tile tables and the queue counters live in a static device
 * array of about 8.2 MB per GPU, held until the process ends: at most 256
 * views and 2^21 wave tiles (64 pixels each) per call, else
 * XD_ERR_INVALID_ARGUMENT.  It calls no hipMalloc or hipFree and captures no
 * graph.  Calls on one GPU share that array: calls that overlap on one GPU,
 * from any context or thread, race and must be serialised by the caller. */
XD_API int xd_batch_render_static(xd_context* ctx, const xd_view* views, uint32_t count,
                                  const xd_batch_tuning* tuning, double* ms, uint64_t* rays);

/* ---- served form: request loop over resident volumes -------------------
 *
 * No counterpart in the reference, which renders one camera script per
 * process.  A server owns volumes it admitted by HBM footprint and a worker
 * thread that folds queued render requests into the running work: every
 * scheduling step is one persistent-kernel launch over wave tiles drawn from
 * all admitted requests (a fair share each, oldest first), with up to
 * `launches_in_flight` launches queued on the device so that requests
 * arriving while others run join the next launch.  Each pixel is computed by
 * the same per-ray code as xd_render, so a result does not depend on the
 * requests it was batched with.  Completed frames are copied to the
 * request's host buffer by a second thread.  Times are ms on a steady clock
 * whose zero is the server's creation. */

typedef struct xd_server xd_server;

typedef struct {
    uint64_t hbm_budget_bytes;   /* resident volumes may hold at most this much;
                                    0 = 90% of the device memory free at creation */
    uint64_t frame_arena_bytes;  /* device memory for frames of admitted requests;
                                    0 = 512 MiB */
    uint32_t max_active;         /* requests admitted at once (1..1024; 0 = 256) */
    uint32_t tiles_per_launch;   /* wave tiles per scheduling step (0 = 65536) */
    uint32_t launches_in_flight; /* 1..4 (0 = 2) */
    uint32_t readback_threads;   /* host threads copying finished frames out (1..16; 0 = 4) */
    xd_batch_tuning tuning;      /* tile shapes etc.; zero-filled = defaults.  A
                                    step whose tiles are all of one traversal
                                    runs that traversal's own kernel */
    double latency_target_ms;    /* 0 = none (every request waits for its turn).
                                    With a target, a request coming up for
                                    admission whose arrival-to-done time,
                                    estimated from the rays the admitted
                                    requests still have to render, its own and
                                    the device's measured rate, exceeds the
                                    target is shed: it completes with status
                                    XD_ERR_REJECTED and no frame */
} xd_server_config;

typedef struct {
    uint32_t volume_id;          /* from xd_server_load_grid / _svo */
    int32_t traversal;
    xd_camera camera;
    uint32_t width, height;      /* the whole display; output region = display */
    float emission_coeff;
    float voxel_ratio[3];        /* 0 entries read as 1 */
    uint32_t* dst;               /* host pixels, width*height RGBA8 rows; NULL = discard */
    double release_ms;           /* arrival time on the server clock; the request
                                    is not started before it (0 = on submit) */
    uint64_t user_tag;
} xd_request;

typedef struct {
    uint64_t request_id;
    uint64_t user_tag;
    int32_t status;              /* XD_OK, or the error that ended the request */
    uint32_t launches;           /* scheduling steps that rendered its tiles */
    double arrival_ms;           /* release_ms, or the submit time */
    double start_ms;             /* first launch with its tiles was queued */
    double device_done_ms;       /* its last tile finished on the device */
    double done_ms;              /* frame in host memory (= device_done_ms if dst is NULL) */
    uint64_t rays;
} xd_completion;

typedef struct {
    uint64_t resident_bytes, budget_bytes;
    uint32_t volumes, pending, active, reserved;
    uint64_t completed, launches, tiles, rays;
    uint32_t max_requests_per_launch, max_launches_overlapped;
    double busy_ms;              /* sum of launch durations (HIP events) */
    uint64_t shed;               /* requests shed under latency_target_ms */
    double rays_per_ms;          /* measured device rate the admission uses */
} xd_server_stats;

XD_API int xd_server_create(xd_context* ctx, const xd_server_config* cfg, xd_server** out);
/* Load a volume if its footprint (xd_volume_footprint) fits the remaining
 * budget, else XD_ERR_REJECTED and nothing is allocated. */
XD_API int xd_server_load_grid(xd_server* srv, const void* rgba8, uint64_t dim_x, uint64_t dim_y, uint64_t dim_z,
                               uint32_t* volume_id);
XD_API int xd_server_load_svo(xd_server* srv, const void* nodes, uint64_t node_count, uint64_t side,
                              uint32_t* volume_id);
/* Waits until no queued or running request uses the volume, then frees it. */
XD_API int xd_server_unload(xd_server* srv, uint32_t volume_id);
/* Validates and queues a request (XD_ERR_INVALID_ARGUMENT for an unknown
 * volume or bad size, XD_ERR_INCOMPATIBLE for a traversal the volume cannot
 * run, XD_ERR_REJECTED if its frame can never fit the arena). */
XD_API int xd_server_submit(xd_server* srv, const xd_request* req, uint64_t* request_id);
/* Returns up to max finished requests in *n, waiting up to timeout_ms
 * (0 = do not wait, <0 = wait for at least one). */
XD_API int xd_server_poll(xd_server* srv, xd_completion* out, uint32_t max, uint32_t* n, int32_t timeout_ms);
XD_API int xd_server_stats_get(xd_server* srv, xd_server_stats* out);
/* Changes the latency target (xd_server_config.latency_target_ms) of a running
 * server; it applies from the next admission on. */
XD_API int xd_server_set_latency_target(xd_server* srv, double latency_target_ms);
/* Current time on the server clock (ms). */
XD_API double xd_server_now_ms(const xd_server* srv);
/* Finishes every queued request, stops the threads, frees the volumes. */
XD_API int xd_server_free(xd_server* srv);

/* ---- octree construction (`xenodon convert`, src/convert.cpp:13-122) ----- */

typedef enum { XD_HEURISTIC_CHAN_DIFF = 0, XD_HEURISTIC_STD_DEV = 1 } xd_heuristic;
typedef enum { XD_TREE_SPARSE = 0, XD_TREE_DAG = 1, XD_TREE_ROPE = 2 } xd_tree_type; /* Octree::Type */

/* ConstructionStats (OctreeConstruction.h:50-62) plus the result's shape. */
typedef struct {
    uint64_t total_leaves;
    uint64_t unique_leaves;
    uint64_t total_nodes;
    uint64_t depth;
    uint64_t side;
    uint64_t node_count;
} xd_build_stats;

/* build_octree(grid, stats, heuristic, type) (OctreeConstruction.h:227-237)
 * on an RGBA8 grid (x fastest).  The node buffer (node_count * 10 u32 words,
 * the .svo node layout) is bit-identical to the reference converter's; it is
 * malloc'd into *out_nodes and released with xd_free. */
XD_API int xd_build_octree(const void* rgba8, uint64_t dim_x, uint64_t dim_y, uint64_t dim_z, int heuristic,
                           int chan_diff, double std_dev, int tree_type, uint32_t** out_nodes, xd_build_stats* stats);
/* The same build on the GPU from a grid volume already in HBM (`xenodon
 * convert --gpu`), for both heuristics and sparse, DAG and rope trees
 * (src/convert.cpp:13-122).  The node buffer and stats are identical to
 * xd_build_octree's; the buffer is malloc'd into *out_nodes (xd_free). */
XD_API int xd_build_octree_device(xd_context* ctx, const xd_volume* grid, int heuristic, int chan_diff, double std_dev,
                                  int tree_type, uint32_t** out_nodes, xd_build_stats* stats);
XD_API void xd_free(void* p);

#ifdef __cplusplus
}
#endif

#endif
