/*
 * Oracle: CPU restatement of Xenodon's per-pixel volume ray traversal.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in xenodon_amd/.  Only tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py may load it; the product never links it.
 *
 * What it restates (reference files, read as text):
 *   resources/common.glsl   camera/ray/emission helpers
 *   resources/octree.glsl   Node layout (40 B, children[8], color, is_leaf_depth)
 *   resources/dda.comp      Amanatides-Woo DDA over the RGBA8 grid
 *   resources/svo_naive.comp, svo_df.comp, esvo.comp, svo_rope.comp
 *   src/model/OctreeConstruction.h, Grid.cpp, Octree.cpp  (converter, ropes)
 *
 * Every traversal exists twice: an FP32 version that follows the GLSL
 * operation order with IEEE-754 single precision, no contraction and
 * correctly rounded division/sqrt (build with -ffp-contract=off, no
 * fast-math), and an FP64 twin of the same algorithm used to adjudicate
 * disagreements.  Both can count the 32-bit volume words the reference
 * algorithm reads per ray (the algorithmic-bytes figure of the roofline).
 *
 * Parity status of each function is stated where it is defined; functions
 * without a closed-form, brute-force or invariant check say "parity unpinned".
 */
#ifndef XO_ORACLE_H
#define XO_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Traversal ids, in the order of SHADER_OPTIONS (src/main_loop.cpp:37-43). */
enum {
    XO_DDA = 0,
    XO_SVO_NAIVE = 1,
    XO_ESVO = 2,
    XO_SVO_DF = 3,
    XO_SVO_ROPE = 4
};

/* Push constant Camera (resources/common.glsl:6-10, src/render/Renderer.h:27-35).
 * translation is already divided by voxel_ratio (src/render/Renderer.cpp:62). */
typedef struct {
    float forward[4];
    float up[4];
    float translation[4];
} xo_camera;

/* Uniform buffer (resources/common.glsl:12-33). */
typedef struct {
    int32_t out_offset[2];
    uint32_t out_extent[2];
    int32_t disp_offset[2];
    uint32_t disp_extent[2];
    float voxel_ratio[4];
    uint32_t model_dim[4];
    float emission_coeff;
} xo_uniforms;

/* Volume: either an RGBA8 grid (x fastest, src/model/Grid.h:50-56) or a node
 * array of 10 u32 words per node (src/model/Octree.h:35-45). */
typedef struct {
    const uint8_t* grid;      /* RGBA8, dims[0]*dims[1]*dims[2]*4 bytes, or NULL */
    uint64_t dims[3];
    const uint32_t* nodes;    /* 10 words per node, or NULL */
    uint64_t node_count;
    uint64_t side;            /* SVO side (power of two) */
    uint32_t max_depth;       /* deepest node depth in the tree (0 for a leaf root) */
} xo_volume;

/* Iteration cap used by the oracle and the HIP kernels alike (a safety net
 * against non-terminating loops; never reached on well-formed input). */
uint64_t xo_iteration_cap(int traversal, const xo_volume* vol, const uint32_t model_dim[3]);

/*
 * Render one output region.
 *   precision:   32 (GLSL restatement) or 64 (FP64 twin)
 *   threads:     host threads (OpenMP); 0 = all
 *   out_rgba:    extent.x*extent.y*4 bytes, unorm8 RGBA (alpha 255), row-major
 *   out_linear:  optional extent.x*extent.y*3 floats, the colour before clamp
 *   out_words:   optional extent.x*extent.y u64, 32-bit volume words read per ray
 * Returns 0 on success, negative on invalid arguments.
 */
int xo_render(int traversal, const xo_volume* vol, const xo_camera* cam,
              const xo_uniforms* u, int precision, int threads,
              uint8_t* out_rgba, float* out_linear, uint64_t* out_words);

/* Trace a single ray given directly (ro, rd already in the traversal's own
 * space, before the emission coefficient): used by unit tests. */
int xo_trace_ray(int traversal, const xo_volume* vol, const double ro[3], const double rd[3],
                 int precision, double out_rgb[3], uint64_t* out_words);

/* The per-pixel ray direction of resources/common.glsl:45-56 (FP32). */
void xo_ray_dir(const xo_camera* cam, const xo_uniforms* u, uint32_t px, uint32_t py, float out_rd[3]);

/* ---- Converter (src/model/OctreeConstruction.h, src/model/Grid.cpp) ---- */

enum { XO_HEUR_CHAN_DIFF = 0, XO_HEUR_STD_DEV = 1 };
enum { XO_TREE_SPARSE = 0, XO_TREE_DAG = 1, XO_TREE_ROPE = 2 };

typedef struct {
    uint64_t total_leaves;
    uint64_t unique_leaves;
    uint64_t total_nodes;
    uint64_t depth;
    uint64_t side;
    uint64_t node_count;
} xo_build_stats;

/*
 * build_octree (OctreeConstruction.h:227-237).  Writes node_count*10 words
 * into *out_nodes (malloc'd; free with xo_free).  Returns 0 on success.
 */
int xo_build_octree(const uint8_t* grid, const uint64_t dims[3], int heuristic,
                    int chan_diff, double std_dev, int tree_type,
                    uint32_t** out_nodes, xo_build_stats* stats);

/* Octree::generate_ropes (Octree.cpp:181-201), in place. */
int xo_generate_ropes(uint32_t* nodes, uint64_t node_count, uint64_t side);

/* Octree::find (Octree.cpp:116-153): index of the node reached, or
 * UINT64_MAX when pos is outside the cube (the reference's nullptr). */
uint64_t xo_octree_find(const uint32_t* nodes, uint64_t side, const uint64_t pos[3], uint64_t max_depth);

/* Grid::vol_scan / stddev_scan (Grid.cpp:81-214). avg is packed RGBA8. */
void xo_vol_scan(const uint8_t* grid, const uint64_t dims[3], const uint64_t bmin[3], const uint64_t bmax[3],
                 uint32_t* avg, uint32_t* max_diff);
void xo_stddev_scan(const uint8_t* grid, const uint64_t dims[3], const uint64_t bmin[3], const uint64_t bmax[3],
                    uint32_t* avg, double* stddev);

void xo_free(void* p);

#ifdef __cplusplus
}
#endif

#endif
