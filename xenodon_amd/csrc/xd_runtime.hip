// C-ABI runtime: contexts, HBM-resident volumes, frames and plain rendering.
//
// Replaces the reference's Vulkan plumbing on the render path:
//   Renderer::render + RenderStatsCollector     -> xd_render (one launch, HIP events)
//   Dda/SvoRaytraceResources (staging uploads)  -> xd_volume_from_grid / _svo
//   HeadlessOutput::{synchronize, download}     -> xd_render's wait, xd_frame_download
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "xenodon.h"
#include "xd_internal.hpp"
#include "xd_runtime.hpp"

namespace {

thread_local std::string g_error;

}  // namespace

namespace xd {

int fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

int fail_hip(hipError_t e, const char* what) {
    return fail(e == hipErrorOutOfMemory ? XD_ERR_OUT_OF_MEMORY : XD_ERR_DEVICE, "%s: %s", what,
                hipGetErrorString(e));
}

void clear_error() { g_error.clear(); }

const char* traversal_name(int t) {
    static const char* names[] = {"dda", "svo-naive", "esvo", "svo-df", "svo-rope"};
    return (t >= 0 && t < 5) ? names[t] : "unknown";
}

int check_traversal(int traversal, const xd_volume* vol) {
    if (traversal < XD_DDA || traversal > XD_SVO_ROPE) return fail(XD_ERR_INVALID_ARGUMENT, "Invalid shader '%d'", traversal);
    const int need = traversal == XD_DDA ? XD_VOLUME_GRID : XD_VOLUME_SVO;
    if (vol->kind != need) {
        // wording of src/main_loop.cpp:120-125
        return fail(XD_ERR_INCOMPATIBLE, "Shader '%s' is incompatible with model type '%s' (requires '%s')",
                    traversal_name(traversal), vol->kind == XD_VOLUME_GRID ? "tiff" : "svo",
                    need == XD_VOLUME_GRID ? "tiff" : "svo");
    }
    return XD_OK;
}

uint64_t iteration_cap(int traversal, const xd_volume* vol, const uint32_t model_dim[3]) {
    // Same cap as the CPU restatement: it bounds every loop on well-formed
    // input with a wide margin and only stops runaway rays.
    if (traversal == XD_DDA) {
        uint64_t s = 0;
        for (int i = 0; i < 3; ++i) s += std::max<uint64_t>(vol->dims[i], model_dim[i]);
        return 2 * s + 16;
    }
    uint32_t depth = std::min<uint32_t>(vol->max_depth, 24);
    uint64_t side = std::max<uint64_t>(vol->side, 1ull << depth);
    return 64 * side + (1ull << 18);
}

void fill_params(RenderParams& p, const xd_volume* vol, int traversal, const xd_camera* cam, const xd_uniforms* u) {
    memset(&p, 0, sizeof(p));
    p.cam.forward = f3{cam->forward[0], cam->forward[1], cam->forward[2]};
    p.cam.up = f3{cam->up[0], cam->up[1], cam->up[2]};
    p.cam.translation = f3{cam->translation[0], cam->translation[1], cam->translation[2]};
    for (int i = 0; i < 2; ++i) {
        p.view.out_off[i] = u->output_region.offset[i];
        p.view.out_ext[i] = u->output_region.extent[i];
        p.view.disp_off[i] = u->display_region.offset[i];
        p.view.disp_ext[i] = u->display_region.extent[i];
    }
    p.view.ratio = f3{u->voxel_ratio[0], u->voxel_ratio[1], u->voxel_ratio[2]};
    for (int i = 0; i < 3; ++i) p.view.model_dim[i] = u->model_dim[i];
    p.view.emission = u->emission_coeff;
    p.cap = iteration_cap(traversal, vol, u->model_dim);
    if (vol->kind == XD_VOLUME_GRID) {
        p.bricks = vol->d_bricks;
        p.occ = vol->d_occ;
        p.macro = vol->d_macro;
        p.macro_shift = vol->macro_shift;
        p.macro_words = vol->macro_words;
        // 32-bit texel indices need the index below 2^32 and, for the 24-bit
        // multiplies of brick_texel32, nbz * nby and nbx below 2^24
        p.wide = (uint64_t)vol->nb[0] * vol->nb[1] * vol->nb[2] * 64 > 0xFFFFFFFFull ||
                         (uint64_t)vol->nb[1] * vol->nb[2] >= (1ull << 24) || vol->nb[0] >= (1u << 24)
                     ? 1u
                     : 0u;
        for (int i = 0; i < 3; ++i) {
            p.dims[i] = (uint32_t)vol->dims[i];
            p.nb[i] = vol->nb[i];
            p.mn[i] = vol->mn[i];
        }
    } else {
        p.nodes = vol->d_nodes;
        p.cw = vol->d_cw;
        p.meta = vol->d_meta;
        p.table = vol->d_table;
        p.rope_depth = vol->rope_depth;
        p.index_mask = vol->index_mask;
        p.empty_tag = vol->empty_tag;
        p.empty_start = vol->empty_start;
        p.root_tag = vol->root_tag;
        // levels 0..3 in LDS for svo-rope only: svo-naive starts every walk from
        // the start table, and esvo / svo-df run at full occupancy with their
        // stacks alone in LDS (3 % faster than with the top levels beside them)
        p.top_nodes = !vol->d_cw || traversal != XD_SVO_ROPE ? 0u : vol->top[3];
        p.stack_slots = vol->max_depth + 1;
        p.lds_slots = p.stack_slots;
        p.node_count = (uint32_t)vol->node_count;
        p.depth_exact = vol->depth_exact;
        p.leaf_base = vol->leaf_base;
    }
}

uint32_t macro_geometry(const uint64_t dims[3], uint32_t* shift, uint32_t mn[3]) {
    // macro cells: the smallest power-of-two side >= 8 voxels whose bitmask
    // fits the kernel's LDS staging area
    for (uint32_t s = 3;; ++s) {
        uint64_t cells = 1;
        for (int i = 0; i < 3; ++i) {
            mn[i] = (uint32_t)((dims[i] + (1ull << s) - 1) >> s);
            cells *= mn[i];
        }
        const uint64_t words = (cells + 31) / 32;
        if (words <= kMaxMacroWords) {
            *shift = s;
            return (uint32_t)words;
        }
    }
}

uint64_t grid_footprint(const uint64_t dims[3]) {
    uint64_t nbricks = 1;
    for (int i = 0; i < 3; ++i) nbricks *= (dims[i] + kBrick - 1) / kBrick;
    uint32_t shift, mn[3];
    const uint32_t words = macro_geometry(dims, &shift, mn);
    return nbricks * 64 * sizeof(uint32_t) + nbricks + (uint64_t)words * 4;
}

uint64_t svo_footprint(uint64_t node_count, bool raw) {
    if (raw || node_count > kMaxTaggedNodes) return node_count * kNodeWords * sizeof(uint32_t);
    return node_count * 8 * sizeof(uint32_t) + node_count * sizeof(uint2) + kTableBytes;
}

int check_view(const xd_uniforms* u) {
    if (u->output_region.extent[0] == 0 || u->output_region.extent[1] == 0)
        return fail(XD_ERR_INVALID_ARGUMENT, "output region is empty");
    if (u->display_region.extent[0] == 0 || u->display_region.extent[1] == 0)
        return fail(XD_ERR_INVALID_ARGUMENT, "display region is empty");
    return XD_OK;
}

}  // namespace xd

using namespace xd;

extern "C" {

const char* xd_last_error(void) { return g_error.c_str(); }

int xd_device_count(int* count) {
    if (!count) return fail(XD_ERR_INVALID_ARGUMENT, "count is null");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        *count = 0;
        return fail_hip(e, "hipGetDeviceCount");
    }
    *count = n;
    return XD_OK;
}

int xd_device_describe(int device_index, char* name, size_t name_len, uint32_t* pci_domain, uint32_t* pci_bus,
                       uint32_t* pci_device, uint32_t* pci_function, uint64_t* total_bytes) {
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, device_index);
    if (e != hipSuccess) return fail_hip(e, "hipGetDeviceProperties");
    if (name && name_len) {
        snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    // The properties hold domain, bus and device; the function number is only
    // in the bus-id string "dddd:bb:dd.f" (hex fields).
    unsigned dom = (unsigned)prop.pciDomainID, bus = (unsigned)prop.pciBusID, dev = (unsigned)prop.pciDeviceID, fn = 0;
    char busid[64] = {};
    if (hipDeviceGetPCIBusId(busid, (int)sizeof busid, device_index) == hipSuccess) {
        unsigned d2, b2, v2, f2;
        if (sscanf(busid, "%x:%x:%x.%x", &d2, &b2, &v2, &f2) == 4) {
            dom = d2;
            bus = b2;
            dev = v2;
            fn = f2;
        }
    }
    if (pci_domain) *pci_domain = dom;
    if (pci_bus) *pci_bus = bus;
    if (pci_device) *pci_device = dev;
    if (pci_function) *pci_function = fn;
    if (total_bytes) *total_bytes = (uint64_t)prop.totalGlobalMem;
    return XD_OK;
}

int xd_open(int device_index, xd_context** out) {
    if (!out) return fail(XD_ERR_INVALID_ARGUMENT, "out is null");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) return fail(XD_ERR_DEVICE, "no HIP device available");
    if (device_index < 0 || device_index >= n)
        return fail(XD_ERR_INVALID_ARGUMENT, "device index %d out of range (%d devices)", device_index, n);
    if ((e = hipSetDevice(device_index)) != hipSuccess) return fail_hip(e, "hipSetDevice");
    auto* ctx = new xd_context();
    ctx->device = device_index;
    ctx->tile_w = 8;
    if ((e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking)) != hipSuccess) {
        delete ctx;
        return fail_hip(e, "hipStreamCreate");
    }
    if ((e = hipEventCreate(&ctx->ev0)) != hipSuccess || (e = hipEventCreate(&ctx->ev1)) != hipSuccess ||
        (e = hipMalloc(&ctx->d_words, sizeof(unsigned long long))) != hipSuccess) {
        xd_close(ctx);
        return fail_hip(e, "context setup");
    }
    clear_error();
    *out = ctx;
    return XD_OK;
}

int xd_context_set_tile(xd_context* ctx, uint32_t tile_w) {
    if (!ctx) return fail(XD_ERR_INVALID_ARGUMENT, "null context");
    if (tile_w != 8 && tile_w != 16 && tile_w != 32)
        return fail(XD_ERR_INVALID_ARGUMENT, "tile width %u is not 8, 16 or 32", tile_w);
    ctx->tile_w = (int)tile_w;
    return XD_OK;
}

int xd_context_set_option(xd_context* ctx, const char* name, int64_t value) {
    if (!ctx || !name) return fail(XD_ERR_INVALID_ARGUMENT, "null argument");
    if (strcmp(name, "octree_layout") == 0) {
        if (value < 0 || value > 2) return fail(XD_ERR_INVALID_ARGUMENT, "octree_layout must be 0, 1 or 2");
        ctx->octree_layout = (int)value;
        return XD_OK;
    }
    if (strcmp(name, "tile_w") == 0) return xd_context_set_tile(ctx, (uint32_t)value);
    if (strcmp(name, "xcd_chunks") == 0) {
        if (value < 0 || value > 64) return fail(XD_ERR_INVALID_ARGUMENT, "xcd_chunks must be in [0, 64]");
        ctx->xcd_chunks = (uint32_t)value;
        return XD_OK;
    }
    return fail(XD_ERR_INVALID_ARGUMENT, "unknown option '%s'", name);
}

int xd_close(xd_context* ctx) {
    if (!ctx) return XD_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    // servers, batches, volumes and frames still alive are released with their context
    std::vector<xd_server*> srvs(ctx->servers.begin(), ctx->servers.end());
    for (auto* s : srvs) (void)xd_server_free(s);
    std::vector<xd_batch*> bats(ctx->batches.begin(), ctx->batches.end());
    for (auto* b : bats) xd_batch_free(b);
    std::vector<xd_volume*> vols(ctx->volumes.begin(), ctx->volumes.end());
    for (auto* v : vols) xd_volume_free(v);
    std::vector<xd_frame*> frs(ctx->frames.begin(), ctx->frames.end());
    for (auto* f : frs) xd_frame_free(f);
    if (ctx->d_words) (void)hipFree(ctx->d_words);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return XD_OK;
}

int xd_volume_from_grid(xd_context* ctx, const void* rgba8, uint64_t dim_x, uint64_t dim_y, uint64_t dim_z,
                        xd_volume** out) {
    if (!ctx || !rgba8 || !out) return fail(XD_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (dim_x == 0 || dim_y == 0 || dim_z == 0) return fail(XD_ERR_INVALID_ARGUMENT, "grid has a zero dimension");
    if (dim_x > (1u << 30) || dim_y > (1u << 30) || dim_z > (1u << 30))
        return fail(XD_ERR_INVALID_ARGUMENT, "grid dimension above 2^30");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return fail_hip(e, "hipSetDevice");
    auto* v = new xd_volume();
    v->ctx = ctx;
    v->kind = XD_VOLUME_GRID;
    v->dims[0] = dim_x;
    v->dims[1] = dim_y;
    v->dims[2] = dim_z;
    v->side = std::max({dim_x, dim_y, dim_z});
    for (int i = 0; i < 3; ++i) v->nb[i] = (uint32_t)((v->dims[i] + kBrick - 1) / kBrick);
    const uint64_t nbricks = (uint64_t)v->nb[0] * v->nb[1] * v->nb[2];
    const uint64_t brick_bytes = nbricks * 64 * sizeof(uint32_t);
    const uint32_t dims32[3] = {(uint32_t)dim_x, (uint32_t)dim_y, (uint32_t)dim_z};
    uint32_t* staging = nullptr;
    const uint64_t layer_texels = dim_x * dim_y;
    if ((e = hipMalloc(&v->d_bricks, brick_bytes)) != hipSuccess || (e = hipMalloc(&v->d_occ, nbricks)) != hipSuccess ||
        (e = hipMalloc(&staging, layer_texels * kBrick * sizeof(uint32_t))) != hipSuccess) {
        if (staging) (void)hipFree(staging);
        xd_volume_free(v);
        return fail_hip(e, "grid allocation");
    }
    v->bytes = brick_bytes + nbricks;
    const uint8_t* src = static_cast<const uint8_t*>(rgba8);
    for (uint32_t bz = 0; bz < v->nb[2]; ++bz) {
        const uint64_t z0 = (uint64_t)bz * kBrick;
        const uint32_t layers = (uint32_t)std::min<uint64_t>(kBrick, dim_z - z0);
        e = hipMemcpyAsync(staging, src + z0 * layer_texels * 4, layers * layer_texels * 4, hipMemcpyHostToDevice,
                           ctx->stream);
        if (e == hipSuccess) e = launch_brick_layer(staging, layers, bz, v->d_bricks, v->d_occ, dims32, v->nb, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
            (void)hipFree(staging);
            xd_volume_free(v);
            return fail_hip(e, "grid upload");
        }
    }
    (void)hipFree(staging);
    v->macro_words = macro_geometry(v->dims, &v->macro_shift, v->mn);
    if ((e = hipMalloc(&v->d_macro, (size_t)v->macro_words * 4)) != hipSuccess ||
        (e = launch_macro_mask(v->d_occ, v->nb, v->macro_shift, v->mn, v->d_macro, v->macro_words, ctx->stream)) !=
            hipSuccess ||
        (e = hipStreamSynchronize(ctx->stream)) != hipSuccess) {
        xd_volume_free(v);
        return fail_hip(e, "grid occupancy");
    }
    v->bytes += (uint64_t)v->macro_words * 4;
    ctx->volumes.insert(v);
    clear_error();
    *out = v;
    return XD_OK;
}

// Structural depth of the tree below the root, and format checks.
static int analyse_octree(const uint32_t* w, uint64_t n, uint32_t* max_depth, uint32_t* is_rope) {
    // one pass in slices over threads: the first bad node of the lowest slice
    // that has one is reported, as a sequential scan would
    const unsigned T = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(
        std::min(16u, std::max(1u, std::thread::hardware_concurrency())), n >> 20));
    std::vector<uint64_t> bad(T, UINT64_MAX);
    std::vector<uint8_t> unordered(T, 0), ropes(T, 0);
    {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < T; ++t)
            th.emplace_back([&, t] {
                const uint64_t lo = n * t / T, hi = n * (t + 1) / T;
                for (uint64_t i = lo; i < hi; ++i) {
                    const uint32_t* nd = w + i * kNodeWords;
                    const bool leaf = (nd[9] & kLeafMask) != 0;
                    bool ok = (nd[9] & kDepthMask) <= 64;
                    for (int k = 0; k < 8; ++k) {
                        ok = ok && nd[k] < n;
                        if (!leaf && nd[k] <= i) unordered[t] = 1;
                        if (leaf && k < 6 && nd[k] != 0) ropes[t] = 1;
                    }
                    if (!ok) {
                        bad[t] = i;
                        return;
                    }
                }
            });
        for (auto& x : th) x.join();
    }
    for (unsigned t = 0; t < T; ++t) {
        if (bad[t] == UINT64_MAX) continue;
        const uint64_t i = bad[t];
        const uint32_t* nd = w + i * kNodeWords;
        if ((nd[9] & kDepthMask) > 64)
            return fail(XD_ERR_FORMAT, "node %llu has depth field %u (above 64)", (unsigned long long)i,
                        nd[9] & kDepthMask);
        for (int k = 0; k < 8; ++k)
            if (nd[k] >= n)
                return fail(XD_ERR_FORMAT, "node %llu child %d = %u is out of range (%llu nodes)",
                            (unsigned long long)i, k, nd[k], (unsigned long long)n);
    }
    bool ordered = true, rope = false;
    for (unsigned t = 0; t < T; ++t) {
        ordered = ordered && !unordered[t];
        rope = rope || ropes[t];
    }
    *is_rope = rope ? 1u : 0u;
    std::vector<uint8_t> height(n, 0);
    if (ordered) {
        // children have larger indices than their parent (reverse post-order,
        // OctreeConstruction.h:92-112): one backward pass
        for (uint64_t i = n; i-- > 0;) {
            const uint32_t* nd = w + i * kNodeWords;
            if (nd[9] & kLeafMask) continue;
            uint32_t h = 0;
            for (int k = 0; k < 8; ++k) h = std::max<uint32_t>(h, height[nd[k]]);
            height[i] = (uint8_t)std::min<uint32_t>(h + 1, 255);
        }
    } else {
        // general DAG: iterative DFS with cycle detection
        std::vector<uint8_t> state(n, 0);  // 0 new, 1 on stack, 2 done
        std::vector<std::pair<uint64_t, int>> st;
        st.push_back({0, 0});
        state[0] = 1;
        while (!st.empty()) {
            auto& top = st.back();
            const uint32_t* nd = w + top.first * kNodeWords;
            if ((nd[9] & kLeafMask) || top.second == 8) {
                if (!(nd[9] & kLeafMask)) {
                    uint32_t h = 0;
                    for (int k = 0; k < 8; ++k) h = std::max<uint32_t>(h, height[nd[k]]);
                    height[top.first] = (uint8_t)std::min<uint32_t>(h + 1, 255);
                }
                state[top.first] = 2;
                st.pop_back();
                continue;
            }
            const uint32_t c = nd[top.second++];
            if (state[c] == 1) return fail(XD_ERR_FORMAT, "octree has a cycle through node %u", c);
            if (state[c] == 0) {
                state[c] = 1;
                st.push_back({c, 0});
            }
        }
    }
    *max_depth = height[0];
    if (*max_depth > 23) return fail(XD_ERR_FORMAT, "octree depth %u exceeds 23 levels", *max_depth);
    return XD_OK;
}

int xd_volume_from_svo(xd_context* ctx, const void* nodes, uint64_t node_count, uint64_t side, xd_volume** out) {
    if (!ctx || !nodes || !out) return fail(XD_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (node_count == 0) return fail(XD_ERR_FORMAT, "octree has no nodes");
    if (node_count > 0xFFFFFFFFull) return fail(XD_ERR_FORMAT, "octree has more than 2^32 nodes");
    const uint32_t* w = static_cast<const uint32_t*>(nodes);
    uint32_t max_depth = 0, is_rope = 0;
    int rc = analyse_octree(w, node_count, &max_depth, &is_rope);
    if (rc) return rc;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return fail_hip(e, "hipSetDevice");
    auto* v = new xd_volume();
    v->ctx = ctx;
    v->kind = XD_VOLUME_SVO;
    v->side = side;
    v->dims[0] = v->dims[1] = v->dims[2] = side;
    v->node_count = node_count;
    v->max_depth = max_depth;
    v->is_rope = is_rope;
    const uint64_t bytes = node_count * kNodeWords * sizeof(uint32_t);
    v->bytes = bytes;
    if (node_count <= kMaxTaggedNodes && ctx->octree_layout != 1) {
        // tagged, sibling-grouped layout, derived on the host
        // every word is written by the derivation: no zero-fill of tens of GB
        std::unique_ptr<uint32_t[]> cw_buf(new (std::nothrow) uint32_t[node_count * 8]);
        std::unique_ptr<uint32_t[]> meta_buf(new (std::nothrow) uint32_t[node_count * 2]);
        if (!cw_buf || !meta_buf) {
            xd_volume_free(v);
            return fail(XD_ERR_OUT_OF_MEMORY, "out of host memory deriving the octree layout");
        }
        uint32_t* const cw = cw_buf.get();
        uint32_t* const meta = meta_buf.get();
        const bool wide = node_count > kMaxNarrowTagNodes || ctx->octree_layout == 2;
        // large trees (the converter's, not DAGs) are numbered and checked in parallel
        const bool tree = node_count >= (1ull << 22) && octree_is_tree(w, node_count);
        build_tagged_layout(w, node_count, cw, meta, v->top, wide, tree, &v->empty_start, &v->leaf_base);
        v->depth_exact = octree_depths_exact(w, node_count, tree) ? 1u : 0u;
        v->rope_depth = !wide && pack_rope_depths(cw, meta, node_count) ? 1u : 0u;
        v->index_mask = v->rope_depth ? kRopeIndexMask : wide ? kWideIndexMask : kIndexMask;
        v->empty_tag = wide ? 0u : kTagEmpty;
        // the leaves' ropes as planes (the allocation keeps the grouped size,
        // which the footprint states before the tree is seen)
        const uint64_t cw_words = rope_plane_words(node_count, v->leaf_base);
        std::unique_ptr<uint32_t[]> planes(new (std::nothrow) uint32_t[cw_words]);
        if (!planes) {
            xd_volume_free(v);
            return fail(XD_ERR_OUT_OF_MEMORY, "out of host memory deriving the octree layout");
        }
        to_rope_planes(cw, node_count, v->leaf_base, planes.get());
        cw_buf.reset();
        // {color, is_leaf_depth} pairs as two planes, so that the colour reads
        // of neighbouring nodes share cache lines
        std::unique_ptr<uint32_t[]> meta_planes(new (std::nothrow) uint32_t[node_count * 2]);
        if (!meta_planes) {
            xd_volume_free(v);
            return fail(XD_ERR_OUT_OF_MEMORY, "out of host memory deriving the octree layout");
        }
        for (uint64_t i = 0; i < node_count; ++i) {
            meta_planes[i] = meta[i * 2];
            meta_planes[node_count + i] = meta[i * 2 + 1];
        }
        meta_buf.reset();
        if ((e = hipMalloc(&v->d_cw, node_count * 8 * sizeof(uint32_t))) != hipSuccess ||
            (e = hipMalloc(&v->d_meta, node_count * 2 * sizeof(uint32_t))) != hipSuccess ||
            (e = hipMemcpy(v->d_cw, planes.get(), cw_words * sizeof(uint32_t), hipMemcpyHostToDevice)) !=
                hipSuccess ||
            (e = hipMemcpy(v->d_meta, meta_planes.get(), node_count * 2 * sizeof(uint32_t), hipMemcpyHostToDevice)) !=
                hipSuccess) {
            xd_volume_free(v);
            return fail_hip(e, "octree upload");
        }
        const uint32_t root_ld = w[9], root_col = w[8];
        v->root_tag = (root_ld & kLeafMask) ? (kTagLeaf | ((root_col & 0xFFFFFFu) ? 0u : v->empty_tag)) : 0u;
        std::vector<uint2> table(kTableCells);
        build_start_table(planes.get(), v->root_tag, v->index_mask, table.data());  // internal nodes' words as in cw
        if ((e = hipMalloc(&v->d_table, kTableBytes)) != hipSuccess ||
            (e = hipMemcpy(v->d_table, table.data(), kTableBytes, hipMemcpyHostToDevice)) != hipSuccess) {
            xd_volume_free(v);
            return fail_hip(e, "octree upload");
        }
        v->bytes = svo_footprint(node_count, false);
    } else {
        if ((e = hipMalloc(&v->d_nodes, bytes)) != hipSuccess ||
            (e = hipMemcpy(v->d_nodes, nodes, bytes, hipMemcpyHostToDevice)) != hipSuccess) {
            xd_volume_free(v);
            return fail_hip(e, "octree upload");
        }
    }
    ctx->volumes.insert(v);
    clear_error();
    *out = v;
    return XD_OK;
}

int xd_volume_footprint(const xd_context* ctx, int kind, uint64_t dim_x, uint64_t dim_y, uint64_t dim_z,
                        uint64_t node_count, uint64_t* bytes) {
    if (!bytes) return fail(XD_ERR_INVALID_ARGUMENT, "null argument");
    *bytes = 0;
    if (kind == XD_VOLUME_GRID) {
        if (dim_x == 0 || dim_y == 0 || dim_z == 0) return fail(XD_ERR_INVALID_ARGUMENT, "grid has a zero dimension");
        if (dim_x > (1u << 30) || dim_y > (1u << 30) || dim_z > (1u << 30))
            return fail(XD_ERR_INVALID_ARGUMENT, "grid dimension above 2^30");
        const uint64_t dims[3] = {dim_x, dim_y, dim_z};
        *bytes = grid_footprint(dims);
    } else if (kind == XD_VOLUME_SVO) {
        if (node_count == 0 || node_count > 0xFFFFFFFFull)
            return fail(XD_ERR_FORMAT, "octree node count %llu is not in 1..2^32-1", (unsigned long long)node_count);
        *bytes = svo_footprint(node_count, ctx && ctx->octree_layout == 1);
    } else {
        return fail(XD_ERR_INVALID_ARGUMENT, "unknown volume kind %d", kind);
    }
    return XD_OK;
}

int xd_volume_info_get(const xd_volume* v, xd_volume_info* out) {
    if (!v || !out) return fail(XD_ERR_INVALID_ARGUMENT, "null argument");
    memset(out, 0, sizeof(*out));
    out->kind = v->kind;
    out->max_depth = v->max_depth;
    for (int i = 0; i < 3; ++i) out->dims[i] = v->dims[i];
    out->side = v->side;
    out->node_count = v->node_count;
    out->device_bytes = v->bytes;
    out->is_rope = v->is_rope;
    out->depth_fields_exact = v->depth_exact;
    return XD_OK;
}

int xd_volume_free(xd_volume* v) {
    if (!v) return XD_OK;
    (void)hipSetDevice(v->ctx->device);
    (void)hipStreamSynchronize(v->ctx->stream);
    if (v->d_bricks) (void)hipFree(v->d_bricks);
    if (v->d_occ) (void)hipFree(v->d_occ);
    if (v->d_macro) (void)hipFree(v->d_macro);
    if (v->d_nodes) (void)hipFree(v->d_nodes);
    if (v->d_cw) (void)hipFree(v->d_cw);
    if (v->d_meta) (void)hipFree(v->d_meta);
    if (v->d_table) (void)hipFree(v->d_table);
    v->ctx->volumes.erase(v);
    delete v;
    return XD_OK;
}

int xd_frame_create(xd_context* ctx, uint32_t width, uint32_t height, xd_frame** out) {
    if (!ctx || !out) return fail(XD_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (width == 0 || height == 0) return fail(XD_ERR_INVALID_ARGUMENT, "frame has a zero dimension");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return fail_hip(e, "hipSetDevice");
    auto* f = new xd_frame();
    f->ctx = ctx;
    f->w = width;
    f->h = height;
    if ((e = hipMalloc(&f->d_pixels, (uint64_t)width * height * 4)) != hipSuccess) {
        delete f;
        return fail_hip(e, "frame allocation");
    }
    ctx->frames.insert(f);
    *out = f;
    return XD_OK;
}

int xd_frame_free(xd_frame* f) {
    if (!f) return XD_OK;
    (void)hipSetDevice(f->ctx->device);
    (void)hipStreamSynchronize(f->ctx->stream);
    (void)hipFree(f->d_pixels);
    f->ctx->frames.erase(f);
    delete f;
    return XD_OK;
}

int xd_frame_device_ptr(xd_frame* f, void** out) {
    if (!f || !out) return fail(XD_ERR_INVALID_ARGUMENT, "null argument");
    *out = f->d_pixels;
    return XD_OK;
}

int xd_render(xd_context* ctx, const xd_volume* vol, int traversal, const xd_camera* cam, const xd_uniforms* u,
              xd_frame* frame, double* dispatch_ms) {
    if (!ctx || !vol || !cam || !u || !frame) return fail(XD_ERR_INVALID_ARGUMENT, "null argument");
    int rc = check_traversal(traversal, vol);
    if (rc) return rc;
    if ((rc = check_view(u))) return rc;
    if (frame->w != u->output_region.extent[0] || frame->h != u->output_region.extent[1])
        return fail(XD_ERR_INVALID_ARGUMENT, "frame is %ux%u but the output region is %ux%u", frame->w, frame->h,
                    u->output_region.extent[0], u->output_region.extent[1]);
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return fail_hip(e, "hipSetDevice");
    RenderParams p;
    fill_params(p, vol, traversal, cam, u);
    p.out = frame->d_pixels;
    p.out_pitch = frame->w;
    p.xcd_chunks = ctx->xcd_chunks;
    if ((e = hipEventRecord(ctx->ev0, ctx->stream)) != hipSuccess) return fail_hip(e, "hipEventRecord");
    if ((e = launch_traversal(traversal, ctx->tile_w, false, p, ctx->stream)) != hipSuccess)
        return fail_hip(e, "traversal launch");
    if ((e = hipEventRecord(ctx->ev1, ctx->stream)) != hipSuccess) return fail_hip(e, "hipEventRecord");
    if ((e = hipEventSynchronize(ctx->ev1)) != hipSuccess) return fail_hip(e, "traversal");
    if (dispatch_ms) {
        float ms = 0.0f;
        if ((e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1)) != hipSuccess) return fail_hip(e, "hipEventElapsedTime");
        *dispatch_ms = ms;
    }
    return XD_OK;
}

int xd_frame_download(xd_frame* f, uint32_t* dst, uint64_t stride_px) {
    if (!f || !dst) return fail(XD_ERR_INVALID_ARGUMENT, "null argument");
    if (stride_px == 0) stride_px = f->w;
    if (stride_px < f->w) return fail(XD_ERR_INVALID_ARGUMENT, "stride below frame width");
    hipError_t e = hipSetDevice(f->ctx->device);
    if (e == hipSuccess) e = hipStreamSynchronize(f->ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpy2D(dst, stride_px * 4, f->d_pixels, (size_t)f->w * 4, (size_t)f->w * 4, f->h,
                        hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail_hip(e, "frame download");
    return XD_OK;
}

int xd_count_words(xd_context* ctx, const xd_volume* vol, int traversal, const xd_camera* cam, const xd_uniforms* u,
                   uint64_t* words) {
    if (!ctx || !vol || !cam || !u || !words) return fail(XD_ERR_INVALID_ARGUMENT, "null argument");
    int rc = check_traversal(traversal, vol);
    if (rc) return rc;
    if ((rc = check_view(u))) return rc;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return fail_hip(e, "hipSetDevice");
    RenderParams p;
    fill_params(p, vol, traversal, cam, u);
    p.words = ctx->d_words;
    unsigned long long h = 0;
    if ((e = hipMemcpyAsync(ctx->d_words, &h, sizeof(h), hipMemcpyHostToDevice, ctx->stream)) != hipSuccess)
        return fail_hip(e, "counter reset");
    if ((e = launch_traversal(traversal, ctx->tile_w, true, p, ctx->stream)) != hipSuccess)
        return fail_hip(e, "counting launch");
    if ((e = hipMemcpyAsync(&h, ctx->d_words, sizeof(h), hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
        return fail_hip(e, "counter read");
    *words = h;
    return XD_OK;
}

}  // extern "C"
