// Batched form: one graph launch drains a global ray-tile queue that spans
// many views (any mix of sizes, volumes and traversals).
//
// Queue: every view is cut into wavefront tiles (8x8, 16x4 or 32x2 pixels,
// chosen per traversal); a tile is one u32 (view << 20 | tile).  Tiles are
// grouped by traversal class.  Either one persistent kernel drains every
// class (by default, its waves starting on different classes so that each CU
// runs latency-bound and issue-bound traversals side by side), or each class
// gets a kernel with the registers and LDS of its own traversal only.  Within
// a class the tiles are dealt into 8 ranges by the band of the view they lie
// in, one range per XCD (the hardware deals workgroups to the XCDs
// round-robin, so block b runs on XCD b % 8).  A wavefront takes tiles from
// its own XCD's range with one atomic per grab, and when that range is
// exhausted it moves on to the other ranges, so a kernel ends only when every
// tile of its classes is done and no CU idles while work is left.  Each tile
// is rendered by exactly the per-ray code of the plain kernels, so every
// pixel is identical to the plain form.
//
// The launch (counter reset + the class kernels) is captured once into a
// hipGraph and replayed for every batch; new cameras or views only rewrite
// the device-side view and tile tables.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <initializer_list>
#include <mutex>
#include <cstring>
#include <vector>

#include "xd_batch.hpp"
#include "xd_trace.hpp"

namespace xd {

__global__ void k_batch_reset(uint32_t* counters) {
    if (threadIdx.x < kClasses * kQueues) counters[threadIdx.x] = 0;
}

// PIPE: the DDA loads the next steps' macro words while it gathers the
// current ones (more registers; used by the DDA class kernel, whose register
// budget is its own, not by the all-class kernel).  For a DDA view,
// lds_mask says that `lds` holds this wave's copy of the view's macro
// bitmask: the loop then reads the mask from LDS, as the plain kernel does.
template <int TRAV, bool PIPE = (TRAV == XD_DDA)>
__device__ __forceinline__ void shade_tile(const BatchView& V, uint32_t tile, uint32_t lane, uint32_t* lds,
                                           bool lds_mask = false);

// Any traversal: the kernel for all classes at once dispatches on the view.
template <>
__device__ __forceinline__ void shade_tile<kClasses, false>(const BatchView& V, uint32_t tile, uint32_t lane,
                                                            uint32_t* lds, bool lds_mask) {
    switch (V.traversal) {
    case XD_DDA: shade_tile<XD_DDA, false>(V, tile, lane, lds, lds_mask); break;
    case XD_SVO_NAIVE: shade_tile<XD_SVO_NAIVE>(V, tile, lane, lds); break;
    case XD_ESVO: shade_tile<XD_ESVO>(V, tile, lane, lds); break;
    case XD_SVO_DF: shade_tile<XD_SVO_DF>(V, tile, lane, lds); break;
    default: shade_tile<XD_SVO_ROPE>(V, tile, lane, lds); break;
    }
}

template <int TRAV, bool PIPE>
__device__ __forceinline__ void shade_tile(const BatchView& V, uint32_t tile, uint32_t lane, uint32_t* lds,
                                           bool lds_mask) {
    const uint32_t sh = V.tw_shift;
    const uint32_t tx = tile % V.tiles_x, ty = tile / V.tiles_x;
    const uint32_t ix = (tx << sh) + (lane & ((1u << sh) - 1u));
    const uint32_t iy = ty * (64u >> sh) + (lane >> sh);
    const RenderParams& p = V.p;
    if (ix >= p.view.out_ext[0] || iy >= p.view.out_ext[1]) return;
    const f3 rd = ray_dir(p.cam, p.view, ix, iy);
    uint64_t words = 0;
    f3 c;
    float coeff;
    if constexpr (TRAV == XD_DDA) {
        const float side = max_elem(mk((float)p.dims[0], (float)p.dims[1], (float)p.dims[2]));
        const f3 ro = muls(p.cam.translation, side);
        coeff = emission_coeff(p.view, rd) / side;
        if (lds_mask)
            c = p.wide ? dda_trace<false, true, false>(p, lds, ro, rd, words)
                       : dda_trace<false, false, false>(p, lds, ro, rd, words);
        else
            c = p.wide ? dda_trace<false, true, PIPE>(p, p.macro, ro, rd, words)
                       : dda_trace<false, false, PIPE>(p, p.macro, ro, rd, words);
    } else {
        c = p.cw ? svo_pixel<TRAV, false, true, false>(p, rd, lds, nullptr, words)
                 : svo_pixel<TRAV, false, false, false>(p, rd, lds, nullptr, words);
        coeff = emission_coeff(p.view, rd);
    }
    p.out[(uint64_t)iy * p.out_pitch + ix] = pack_rgba(muls(c, coeff));
}

// The tiles of one traversal class: kQueues ranges (one per XCD) of
// qbeg/counters.  A wave takes tiles from its own XCD's range with one atomic
// per grab and moves on to the other ranges when that one is drained; every
// wave leaves once all ranges are drained.  grabs packs the tiles per grab of
// each class, 6 bits each (class c at bit 6c).
template <int TRAV>
__device__ __forceinline__ void class_body(const BatchView* __restrict__ views, const uint32_t* __restrict__ tiles,
                                           const uint32_t* __restrict__ qbeg, uint32_t* __restrict__ counters,
                                           uint32_t grabs, uint32_t mix, uint32_t* lds) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t home = blockIdx.x % kQueues;
    constexpr uint32_t kRanges = TRAV == (int)kClasses ? kClasses * kQueues : kQueues;
    // mixed: consecutive waves of a block start on different classes
    const uint32_t first = mix ? (blockIdx.x / kQueues * 4u + (threadIdx.x >> 6)) % kClasses : 0u;
    // The macro bitmask of the last DDA volume this wave copied to LDS
    // (wave-uniform).  Its region is the wave's own part of the block's LDS:
    // kWaveMacroWords words in the DDA class kernel, the wave's stack region
    // in the all-class kernel (whose esvo and svo-df tiles overwrite it).  A
    // mask larger than the region is read from device memory.
    const uint32_t* staged = nullptr;
    for (uint32_t qi = 0; qi < kRanges; ++qi) {
        // this XCD's range first (of every class, for the all-class kernel),
        // then the next XCD's
        const uint32_t qq = TRAV == (int)kClasses
                                ? ((qi + first) % kClasses) * kQueues + (home + qi / kClasses) % kQueues
                                : (home + qi) % kQueues;
        const uint32_t beg = qbeg[qq], end = qbeg[qq + 1];
        const uint32_t per_grab = (grabs >> (6u * (TRAV == (int)kClasses ? qq / kQueues : (uint32_t)TRAV))) & 63u;
        for (;;) {
            uint32_t t = 0;
            if (lane == 0) t = atomicAdd(&counters[qq], per_grab);
            // wave-uniform from here on: the view's parameters are read with
            // scalar loads into SGPRs instead of occupying VGPRs
            t = __builtin_amdgcn_readfirstlane(t);
            if (beg + t >= end) break;
            const uint32_t stop = min(end, beg + t + per_grab);
            for (uint32_t i = beg + t; i < stop; ++i) {
                const uint32_t e = __builtin_amdgcn_readfirstlane(tiles[i]);
                const BatchView& V = views[e >> 20];
                const uint32_t tile = e & (kMaxTilesPerView - 1u);
                if (TRAV == XD_DDA || (TRAV == (int)kClasses && V.traversal == XD_DDA)) {
                    const uint32_t region = TRAV == XD_DDA ? kWaveMacroWords : 2u * V.p.lds_slots * kWaveLanes;
                    uint32_t* const wave_macro = lds + (threadIdx.x >> 6) * region;
                    const uint32_t* m = V.p.macro;
                    const bool fits = V.p.macro_words <= region;
                    if (fits && m != staged) {
                        for (uint32_t w = lane; w < V.p.macro_words; w += kWaveLanes) wave_macro[w] = m[w];
                        // the copy's LDS writes complete before any lane reads
                        // another lane's words: s_waitcnt lgkmcnt(0) (0xC07F
                        // leaves the vmcnt and expcnt fields at their maxima)
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                        __builtin_amdgcn_s_waitcnt(0xC07F);
                        __builtin_amdgcn_wave_barrier();
                        staged = m;
                    }
                    shade_tile<TRAV>(V, tile, lane, wave_macro, fits);
                } else {
                    shade_tile<TRAV>(V, tile, lane, lds);
                    if (TRAV == (int)kClasses && (V.traversal == XD_ESVO || V.traversal == XD_SVO_DF)) staged = nullptr;
                }
            }
        }
    }
}

// One persistent kernel per traversal class, so each gets the register and
// LDS budget of its own traversal: the register budget left to the compiler
// ...
template <int TRAV>
__global__ __launch_bounds__(256) void k_batch(const BatchView* __restrict__ views, const uint32_t* __restrict__ tiles,
                                               const uint32_t* __restrict__ qbeg, uint32_t* __restrict__ counters,
                                               uint32_t grabs, uint32_t mix) {
    extern __shared__ uint32_t lds[];
    class_body<TRAV>(views, tiles, qbeg, counters, grabs, mix, lds);
}

// ... or capped for 8 waves per SIMD (registers beyond the cap spill).
template <int TRAV>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_batch_w8(
    const BatchView* __restrict__ views, const uint32_t* __restrict__ tiles, const uint32_t* __restrict__ qbeg,
    uint32_t* __restrict__ counters, uint32_t grabs, uint32_t mix) {
    extern __shared__ uint32_t lds[];
    class_body<TRAV>(views, tiles, qbeg, counters, grabs, mix, lds);
}

namespace {
using BatchKernel = void (*)(const BatchView*, const uint32_t*, const uint32_t*, uint32_t*, uint32_t, uint32_t);
const BatchKernel kKernels[2][kClasses + 1] = {
    {k_batch<0>, k_batch<1>, k_batch<2>, k_batch<3>, k_batch<4>, k_batch<kClasses>},
    {k_batch_w8<0>, k_batch_w8<1>, k_batch_w8<2>, k_batch_w8<3>, k_batch_w8<4>, k_batch_w8<kClasses>}};
}  // namespace

uint32_t tw_shift_of(uint32_t tw) { return tw == 32 ? 5u : (tw == 16 ? 4u : 3u); }

uint32_t pack_grabs(uint32_t tiles_per_grab, const uint32_t rays_per_lane[kClasses]) {
    const uint32_t g = std::max<uint32_t>(1, std::min<uint32_t>(tiles_per_grab, 8));
    uint32_t packed = 0;
    for (uint32_t c = 0; c < kClasses; ++c) {
        const uint32_t r = rays_per_lane && (rays_per_lane[c] == 2 || rays_per_lane[c] == 4) ? rays_per_lane[c] : 1u;
        packed |= (g * r) << (6 * c);
    }
    return packed;
}

void make_batch_view(const xd_volume* vol, int traversal, const xd_camera* cam, const xd_uniforms* u, uint32_t* out,
                     uint32_t pitch, uint32_t tile_w, uint32_t rays_per_lane, BatchView& B, uint64_t* ntiles) {
    memset(&B, 0, sizeof(B));
    fill_params(B.p, vol, traversal, cam, u);
    B.p.top_nodes = 0;  // no LDS top cache: views of one launch may use different volumes
    B.p.out = out;
    B.p.out_pitch = pitch;
    B.traversal = (uint32_t)traversal;
    if (tile_w != 8 && tile_w != 16 && tile_w != 32) tile_w = 8;
    B.tw_shift = tw_shift_of(tile_w);
    B.rpl = rays_per_lane == 2 || rays_per_lane == 4 ? rays_per_lane : 1u;
    const uint32_t th = 64u / tile_w;
    B.tiles_x = (u->output_region.extent[0] + tile_w - 1) / tile_w;
    *ntiles = (uint64_t)B.tiles_x * ((u->output_region.extent[1] + th - 1) / th);
}

uint32_t batch_view_slots(const BatchView& B) {
    return (B.traversal == XD_ESVO || B.traversal == XD_SVO_DF) ? B.p.stack_slots : 0u;
}

size_t batch_lds_bytes(int traversal, uint32_t slots) {
    if (traversal == XD_DDA) return (size_t)(kBlockThreads / kWaveLanes) * kWaveMacroWords * sizeof(uint32_t);
    // 4 waves x own stacks of p.lds_slots = max(slots, 1) entries (the
    // all-class kernel's DDA tiles use that region for their macro bitmask)
    if (traversal == (int)kClasses) slots = std::max<uint32_t>(slots, 1);
    return slots ? (size_t)2 * slots * kBlockThreads * sizeof(uint32_t) : 0;
}

int batch_grid(int device, int traversal, uint32_t slots, bool w8, uint32_t* grid) {
    hipError_t e;
    int cus = 0, per_cu = 0;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device)) != hipSuccess ||
        (e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu,
                                                          reinterpret_cast<const void*>(kKernels[w8][traversal]),
                                                          256, batch_lds_bytes(traversal, slots))) != hipSuccess)
        return fail_hip(e, "batch occupancy");
    uint32_t g = (uint32_t)std::max(1, cus * std::max(per_cu, 1));
    *grid = (g + kQueues - 1) / kQueues * kQueues;
    return XD_OK;
}

hipError_t launch_batch(const BatchView* views, const uint32_t* tiles, const uint32_t* qbeg, uint32_t* counters,
                        uint32_t grabs, const BatchClasses& cls, bool w8, bool mix, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, counters);
    for (int c = 0; c <= (int)kClasses; ++c) {
        if (!cls.grid[c]) continue;  // no tiles for this kernel
        const uint32_t base = c < (int)kClasses ? c * kQueues : 0;
        hipLaunchKernelGGL(kKernels[w8][c], dim3(cls.grid[c]), dim3(256), batch_lds_bytes(c, cls.slots[c]), s, views,
                           tiles, qbeg + base, counters + base, grabs, mix ? 1u : 0u);
    }
    return hipGetLastError();
}

}  // namespace xd

using namespace xd;

struct xd_batch {
    xd_context* ctx = nullptr;
    std::vector<xd_view> views;
    xd_batch_tuning tuning{};
    // device tables
    BatchView* d_views = nullptr;
    uint32_t* d_tiles = nullptr;
    uint32_t* d_qbeg = nullptr;
    uint32_t* d_counters = nullptr;
    uint32_t view_cap = 0, tile_cap = 0;
    uint64_t total_rays = 0;
    // captured launch
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    BatchClasses graph_cls{};  // per-class grids and LDS stack strides the graph was captured with
    uint32_t graph_grab = 0, graph_w8 = 0, graph_mix = 0;
    bool keep_mixed = false;  // autotuner: one traversal's views on the all-traversal kernel
};

namespace {

void drop_graph(xd_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

int ensure_capacity(xd_batch* b, uint32_t nviews, uint32_t ntiles) {
    hipError_t e;
    if (nviews > b->view_cap) {
        drop_graph(b);
        if (b->d_views) (void)hipFree(b->d_views);
        b->d_views = nullptr;
        const uint32_t cap = std::max<uint32_t>(nviews, 64);
        if ((e = hipMalloc(&b->d_views, sizeof(BatchView) * cap)) != hipSuccess) return fail_hip(e, "batch views");
        b->view_cap = cap;
    }
    if (ntiles > b->tile_cap) {
        drop_graph(b);
        if (b->d_tiles) (void)hipFree(b->d_tiles);
        b->d_tiles = nullptr;
        const uint32_t cap = std::max<uint32_t>(ntiles, 1u << 16);
        if ((e = hipMalloc(&b->d_tiles, sizeof(uint32_t) * cap)) != hipSuccess) return fail_hip(e, "batch tiles");
        b->tile_cap = cap;
    }
    return XD_OK;
}

// Relative cost of a ray per traversal, from the one-traversal batched rates
// on the 512x361x512 CT volume at 1920x1080 (dda 1604, svo-naive 1490, esvo
// 1227, svo-df 740, svo-rope 1514 Mrays/s).  Only the split of the XCDs
// between traversals depends on it; the tiles a wrong split leaves over are
// taken by the waves of the other XCDs.
constexpr double kRayCost[kClasses] = {1.0, 1.08, 1.31, 2.17, 1.06};

// XCD affinity: the eight XCDs are shared out between the traversal classes
// in proportion to each class's tiles x ray cost, as consecutive segments of
// one line of 8 XCD-widths, so that each XCD's L2 serves one or two
// traversals' data.  A class's tiles (its band ranges concatenated, so still
// in image order) are cut into runs, one per XCD its segment covers, in
// proportion to the cover.  The queue walk is unchanged: a wave drains its own
// XCD's ranges (of every class) first, then the next XCD's.
void deal_by_affinity(std::vector<std::vector<uint32_t>>& q) {
    double w[kClasses], total = 0;
    for (uint32_t c = 0; c < kClasses; ++c) {
        size_t n = 0;
        for (uint32_t b = 0; b < kQueues; ++b) n += q[c * kQueues + b].size();
        w[c] = (double)n * kRayCost[c];
        total += w[c];
    }
    if (total <= 0) return;
    double at = 0;  // start of class c's segment, in XCD widths
    for (uint32_t c = 0; c < kClasses; ++c) {
        std::vector<uint32_t> all;
        for (uint32_t b = 0; b < kQueues; ++b) {
            auto& v = q[c * kQueues + b];
            all.insert(all.end(), v.begin(), v.end());
            v.clear();
        }
        const double len = w[c] / total * kQueues;
        const double end = at + len;
        size_t taken = 0;
        for (uint32_t x = (uint32_t)at; x < kQueues && x < end && !all.empty(); ++x) {
            const double cover = std::min(end, (double)x + 1) - std::max(at, (double)x);
            const bool last = (double)x + 1 >= end - 1e-9 || x + 1 == kQueues;
            const size_t k = last ? all.size() - taken
                                  : std::min(all.size() - taken, (size_t)std::llround(cover / len * (double)all.size()));
            q[c * kQueues + x].assign(all.begin() + taken, all.begin() + taken + k);
            taken += k;
        }
        at = end;
    }
}

// The view and tile tables of a launch, built on the host.
struct HostTables {
    std::vector<BatchView> views;
    std::vector<uint32_t> tiles;
    uint32_t qbeg[kClasses * kQueues + 1];
    BatchClasses cls;  // grid[c] = 1 if kernel c has tiles; slots[c] its LDS stack stride
    uint64_t total_rays;
};

// Builds the view and tile tables of n views under a tuning.  Tiles are
// grouped by class, and within a class dealt into kQueues ranges by the band
// of their view they lie in.
// keep_mixed: a batch of one traversal still runs the all-traversal kernel
// when the tuning says so (the autotuner times one traversal's views on the
// kernel the whole batch will run).
int build_tables(const xd_view* views, uint32_t n, const xd_batch_tuning& tuning, HostTables& T,
                 bool keep_mixed = false) {
    T.views.assign(n, BatchView{});
    std::vector<BatchView>& bv = T.views;
    std::vector<std::vector<uint32_t>> q(kClasses * kQueues);
    struct Group {
        uint64_t key[6];
        std::vector<uint32_t> views;
        uint32_t tiles_x, tiles_y;  // tiles_y: groups of rays_per_lane rows
    };
    std::vector<Group> groups;
    BatchClasses cls{};
    T.total_rays = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const xd_view& v = views[i];
        BatchView& B = bv[i];
        uint64_t nt = 0;
        make_batch_view(v.volume, v.traversal, &v.camera, &v.uniforms, v.frame->d_pixels, v.frame->w,
                        tuning.tile_w[v.traversal], tuning.rays_per_lane[v.traversal], B, &nt);
        const uint32_t W = v.uniforms.output_region.extent[0], H = v.uniforms.output_region.extent[1];
        const uint32_t tiles_y = (uint32_t)(nt / B.tiles_x);
        // rays per lane: the rows of a wave tile column are taken rpl at a
        // time (one grab of rpl entries, traced in turn by the same lanes);
        // rows past the bottom of a view are queued as no-op entries so that
        // every grab stays aligned to a group of rpl rows
        const uint32_t rpl = B.rpl, big_y = (tiles_y + rpl - 1) / rpl;
        if ((uint64_t)big_y * rpl * B.tiles_x > kMaxTilesPerView)
            return fail(XD_ERR_INVALID_ARGUMENT, "view %u has %llu tiles (limit %u)", i,
                        (unsigned long long)big_y * rpl * B.tiles_x, kMaxTilesPerView);
        const uint32_t c = (uint32_t)v.traversal;
        cls.slots[c] = std::max(cls.slots[c], batch_view_slots(B));
        cls.grid[c] = 1;
        const uint32_t order = tuning.interleave;
        if (order == 1 || order == 5) {
            // views of one volume, traversal, size and tile shape form a group
            // whose tiles are dealt together below
            const uint64_t key[6] = {(uint64_t)(uintptr_t)v.volume, (uint64_t)c, W, H, B.tw_shift, B.rpl};
            size_t g = 0;
            while (g < groups.size() && memcmp(groups[g].key, key, sizeof key) != 0) ++g;
            if (g == groups.size()) groups.push_back({{key[0], key[1], key[2], key[3], key[4], key[5]}, {}, B.tiles_x, big_y});
            groups[g].views.push_back(i);
        } else {
            // squares of sq x sq groups of rows (sq = 1: plain rows),
            // row-major within each band's squares
            const uint32_t sq = order == 2 ? 8u : order == 3 ? 4u : order == 4 ? 16u : 1u;
            for (uint32_t ty0 = 0; ty0 < big_y; ty0 += sq) {
                const uint32_t band = (uint32_t)((uint64_t)ty0 * kQueues / big_y);
                const uint32_t ty1 = std::min(big_y, ty0 + sq);
                for (uint32_t tx0 = 0; tx0 < B.tiles_x; tx0 += sq)
                    for (uint32_t ty = ty0; ty < ty1; ++ty)
                        for (uint32_t tx = tx0; tx < std::min(B.tiles_x, tx0 + sq); ++tx)
                            for (uint32_t r = 0; r < rpl; ++r)
                                q[c * kQueues + band].push_back((i << 20) | ((ty * rpl + r) * B.tiles_x + tx));
            }
        }
        T.total_rays += (uint64_t)W * H;
    }
    // interleaved: the same tile of every view of a group, one after another
    // (order 5: in squares of 8 x 8 tiles)
    const uint32_t gsq = tuning.interleave == 5 ? 8u : 1u;
    for (const auto& G : groups)
        for (uint32_t ty0 = 0; ty0 < G.tiles_y; ty0 += gsq) {
            const uint32_t band = (uint32_t)((uint64_t)ty0 * kQueues / G.tiles_y);
            for (uint32_t tx0 = 0; tx0 < G.tiles_x; tx0 += gsq)
                for (uint32_t ty = ty0; ty < std::min(G.tiles_y, ty0 + gsq); ++ty)
                    for (uint32_t tx = tx0; tx < std::min(G.tiles_x, tx0 + gsq); ++tx)
                        for (uint32_t i : G.views)
                            for (uint32_t r = 0; r < (uint32_t)G.key[5]; ++r)
                                q[(uint32_t)G.key[1] * kQueues + band].push_back(
                                    (i << 20) | ((ty * (uint32_t)G.key[5] + r) * G.tiles_x + tx));
        }
    // A batch of one traversal runs that traversal's own kernel whatever the
    // tuning says: with one class there is nothing to mix, and the class
    // kernel has that traversal's register budget (and, for DDA, the loop
    // that loads the next steps' macro words during the current gathers).
    uint32_t present = 0;
    for (uint32_t c = 0; c < kClasses; ++c) present += cls.grid[c] ? 1u : 0u;
    const bool per_class = tuning.class_kernels || (present == 1 && !keep_mixed);
    // one LDS layout per kernel: every wave's stack region is sized for the
    // deepest tree among the views that kernel drains
    if (!per_class) {
        uint32_t all = 0, any = 0;
        for (uint32_t c = 0; c < kClasses; ++c) {
            all = std::max(all, cls.slots[c]);
            any |= cls.grid[c];
            cls.grid[c] = cls.slots[c] = 0;
        }
        cls.grid[kClasses] = any;
        cls.slots[kClasses] = all;
    }
    for (uint32_t i = 0; i < n; ++i)
        bv[i].p.lds_slots = std::max<uint32_t>(cls.slots[per_class ? bv[i].traversal : kClasses], 1);
    if (!per_class && tuning.mix_classes == 2) deal_by_affinity(q);
    uint64_t total = 0;
    for (uint32_t k = 0; k < kClasses * kQueues; ++k) {
        T.qbeg[k] = (uint32_t)total;
        total += q[k].size();
    }
    if (total > 0xFFFFFFFFull) return fail(XD_ERR_INVALID_ARGUMENT, "%llu wave tiles in one batch", (unsigned long long)total);
    T.qbeg[kClasses * kQueues] = (uint32_t)total;
    T.tiles.clear();
    T.tiles.reserve(total);
    for (auto& v : q) T.tiles.insert(T.tiles.end(), v.begin(), v.end());
    T.cls = cls;
    return XD_OK;
}

// Builds the current views' tables and uploads them into the batch's own
// device buffers; *need gets each kernel's LDS stack stride and, in grid[c],
// 1 if kernel c has tiles.
int upload_tables(xd_batch* b, BatchClasses* need) {
    HostTables T;
    const uint32_t n = (uint32_t)b->views.size();
    int rc = build_tables(b->views.data(), n, b->tuning, T, b->keep_mixed);
    if (rc) return rc;
    b->total_rays = T.total_rays;
    const uint32_t total = (uint32_t)T.tiles.size();
    if ((rc = ensure_capacity(b, n, std::max<uint32_t>(total, 1)))) return rc;
    hipError_t e;
    if ((e = hipMemcpy(b->d_views, T.views.data(), sizeof(BatchView) * n, hipMemcpyHostToDevice)) != hipSuccess ||
        (total && (e = hipMemcpy(b->d_tiles, T.tiles.data(), sizeof(uint32_t) * total, hipMemcpyHostToDevice)) != hipSuccess) ||
        (e = hipMemcpy(b->d_qbeg, T.qbeg, sizeof(T.qbeg), hipMemcpyHostToDevice)) != hipSuccess)
        return fail_hip(e, "batch tables");
    *need = T.cls;
    return XD_OK;
}

int capture(xd_batch* b, const BatchClasses& need) {
    drop_graph(b);
    hipError_t e;
    const bool w8 = b->tuning.waves_per_simd == 8;
    BatchClasses cls = need;
    for (int c = 0; c <= (int)kClasses; ++c) {
        if (!need.grid[c]) continue;
        int rc = batch_grid(b->ctx->device, c, need.slots[c], w8, &cls.grid[c]);
        if (rc) return rc;
    }
    hipStream_t s = b->ctx->stream;
    const uint32_t grab = pack_grabs(b->tuning.tiles_per_grab, b->tuning.rays_per_lane);
    if ((e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal)) != hipSuccess)
        return fail_hip(e, "hipStreamBeginCapture");
    const hipError_t le =
        launch_batch(b->d_views, b->d_tiles, b->d_qbeg, b->d_counters, grab, cls, w8, b->tuning.mix_classes != 0, s);
    if ((e = hipStreamEndCapture(s, &b->graph)) != hipSuccess) return fail_hip(e, "hipStreamEndCapture");
    if (le != hipSuccess) return fail_hip(le, "batch launch capture");
    if ((e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0)) != hipSuccess)
        return fail_hip(e, "hipGraphInstantiate");
    b->graph_cls = cls;
    b->graph_grab = grab;
    b->graph_w8 = w8;
    b->graph_mix = b->tuning.mix_classes;
    return XD_OK;
}

int prepare(xd_batch* b) {
    BatchClasses need{};
    int rc = upload_tables(b, &need);
    if (rc) return rc;
    const uint32_t grab = pack_grabs(b->tuning.tiles_per_grab, b->tuning.rays_per_lane);
    bool stale = !b->exec || grab != b->graph_grab || b->graph_w8 != (uint32_t)(b->tuning.waves_per_simd == 8) ||
                 b->graph_mix != b->tuning.mix_classes;
    for (int c = 0; c <= (int)kClasses; ++c)
        stale |= (need.grid[c] && !b->graph_cls.grid[c]) || need.slots[c] > b->graph_cls.slots[c];
    return stale ? capture(b, need) : XD_OK;
}

int run(xd_batch* b, double* ms) {
    hipStream_t s = b->ctx->stream;
    hipError_t e;
    if ((e = hipEventRecord(b->ctx->ev0, s)) != hipSuccess || (e = hipGraphLaunch(b->exec, s)) != hipSuccess ||
        (e = hipEventRecord(b->ctx->ev1, s)) != hipSuccess || (e = hipEventSynchronize(b->ctx->ev1)) != hipSuccess)
        return fail_hip(e, "batch render");
    if (ms) {
        float f = 0;
        if ((e = hipEventElapsedTime(&f, b->ctx->ev0, b->ctx->ev1)) != hipSuccess)
            return fail_hip(e, "hipEventElapsedTime");
        *ms = f;
    }
    return XD_OK;
}

// Every view's volume and frame belong to ctx, its traversal suits its
// volume, and its frame is the size of its output region.
int check_views(const xd_context* ctx, const xd_view* views, uint32_t count) {
    for (uint32_t i = 0; i < count; ++i) {
        const xd_view& v = views[i];
        if (!v.volume || !v.frame) return fail(XD_ERR_INVALID_ARGUMENT, "view %u: null volume or frame", i);
        if (v.volume->ctx != ctx || v.frame->ctx != ctx)
            return fail(XD_ERR_INVALID_ARGUMENT, "view %u: volume or frame belongs to another context", i);
        int rc = check_traversal(v.traversal, v.volume);
        if (rc) return rc;
        if ((rc = check_view(&v.uniforms))) return rc;
        if (v.frame->w != v.uniforms.output_region.extent[0] || v.frame->h != v.uniforms.output_region.extent[1])
            return fail(XD_ERR_INVALID_ARGUMENT, "view %u: frame is %ux%u but the output region is %ux%u", i,
                        v.frame->w, v.frame->h, v.uniforms.output_region.extent[0], v.uniforms.output_region.extent[1]);
    }
    return XD_OK;
}

// The fastest launch shape measured on the 1080p five-traversal batch: one
// kernel for every traversal, its waves starting on different traversals,
// 8 waves per SIMD, tiles in squares of 8.
xd_batch_tuning default_tuning() {
    xd_batch_tuning t{};
    for (int i = 0; i < 5; ++i) {
        t.tile_w[i] = 8;
        t.rays_per_lane[i] = 1;
    }
    t.tiles_per_grab = 1;
    t.interleave = 2;
    t.waves_per_simd = 8;
    t.mix_classes = 1;
    return t;
}

int check_tuning(const xd_batch_tuning* t) {
    for (int i = 0; i < 5; ++i)
        if (t->tile_w[i] != 8 && t->tile_w[i] != 16 && t->tile_w[i] != 32)
            return fail(XD_ERR_INVALID_ARGUMENT, "tile width %u is not 8, 16 or 32", t->tile_w[i]);
    if (t->tiles_per_grab < 1 || t->tiles_per_grab > 8)
        return fail(XD_ERR_INVALID_ARGUMENT, "tiles_per_grab must be 1..8");
    if (t->waves_per_simd != 0 && t->waves_per_simd != 8)
        return fail(XD_ERR_INVALID_ARGUMENT, "waves_per_simd must be 0 (compiler's choice) or 8");
    if (t->class_kernels > 1) return fail(XD_ERR_INVALID_ARGUMENT, "class_kernels must be 0 or 1");
    if (t->interleave > 5) return fail(XD_ERR_INVALID_ARGUMENT, "interleave (tile order) must be 0..5");
    if (t->mix_classes > 2) return fail(XD_ERR_INVALID_ARGUMENT, "mix_classes must be 0, 1 or 2");
    for (int i = 0; i < 5; ++i)
        if (t->rays_per_lane[i] > 4 || t->rays_per_lane[i] == 3)
            return fail(XD_ERR_INVALID_ARGUMENT, "rays_per_lane %u is not 1, 2 or 4 (0 = 1)", t->rays_per_lane[i]);
    return XD_OK;
}

}  // namespace

extern "C" {

int xd_batch_create(xd_context* ctx, xd_batch** out) {
    if (!ctx || !out) return fail(XD_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return fail_hip(e, "hipSetDevice");
    auto* b = new xd_batch();
    b->ctx = ctx;
    b->tuning = default_tuning();
    if ((e = hipMalloc(&b->d_qbeg, sizeof(uint32_t) * (kClasses * kQueues + 1))) != hipSuccess ||
        (e = hipMalloc(&b->d_counters, sizeof(uint32_t) * kClasses * kQueues)) != hipSuccess) {
        xd_batch_free(b);
        return fail_hip(e, "batch allocation");
    }
    ctx->batches.insert(b);
    *out = b;
    return XD_OK;
}

int xd_batch_set_views(xd_batch* b, const xd_view* views, uint32_t count) {
    if (!b || (!views && count)) return fail(XD_ERR_INVALID_ARGUMENT, "null argument");
    if (count > kMaxBatchViews) return fail(XD_ERR_INVALID_ARGUMENT, "%u views exceed the limit of %u", count, kMaxBatchViews);
    int rc = check_views(b->ctx, views, count);
    if (rc) return rc;
    hipError_t e = hipSetDevice(b->ctx->device);
    if (e != hipSuccess) return fail_hip(e, "hipSetDevice");
    b->views.assign(views, views + count);
    return prepare(b);
}

int xd_batch_set_tuning(xd_batch* b, const xd_batch_tuning* t) {
    if (!b || !t) return fail(XD_ERR_INVALID_ARGUMENT, "null argument");
    int rc = check_tuning(t);
    if (rc) return rc;
    b->tuning = *t;
    if (b->views.empty()) return XD_OK;
    hipError_t e = hipSetDevice(b->ctx->device);
    if (e != hipSuccess) return fail_hip(e, "hipSetDevice");
    return prepare(b);
}

int xd_batch_get_tuning(const xd_batch* b, xd_batch_tuning* t) {
    if (!b || !t) return fail(XD_ERR_INVALID_ARGUMENT, "null argument");
    *t = b->tuning;
    return XD_OK;
}

int xd_batch_render(xd_batch* b, double* ms, uint64_t* rays) {
    if (!b) return fail(XD_ERR_INVALID_ARGUMENT, "null argument");
    if (rays) *rays = b->total_rays;
    if (b->views.empty()) {
        if (ms) *ms = 0;
        return XD_OK;
    }
    hipError_t e = hipSetDevice(b->ctx->device);
    if (e != hipSuccess) return fail_hip(e, "hipSetDevice");
    return run(b, ms);
}

namespace {
// Best of `trials` timed launches of the batch's current tuning.
int time_tuning(xd_batch* b, uint32_t trials, double* best) {
    int rc = prepare(b);
    if (rc) return rc;
    *best = 1e30;
    for (uint32_t k = 0; k < trials; ++k) {
        double ms = 0;
        if ((rc = run(b, &ms))) return rc;
        *best = std::min(*best, ms);
    }
    return XD_OK;
}

// One coordinate of the search: times each candidate value of *field and
// keeps the one already set unless another is at least 2 % faster (launch
// times vary by about 1 % from run to run).
int sweep(xd_batch* b, uint32_t trials, uint32_t* field, std::initializer_list<uint32_t> values) {
    const uint32_t start = *field;
    double cur = 0;
    int rc = time_tuning(b, trials, &cur);
    if (rc) return rc;
    uint32_t pick = start;
    double pick_ms = cur;
    for (uint32_t v : values) {
        if (v == start) continue;
        *field = v;
        double ms = 0;
        if ((rc = time_tuning(b, trials, &ms))) return rc;
        if (ms < 0.98 * cur && ms < pick_ms) {
            pick = v;
            pick_ms = ms;
        }
    }
    *field = pick;
    return XD_OK;
}
}  // namespace

int xd_batch_autotune(xd_batch* b, uint32_t trials, xd_batch_tuning* chosen) {
    if (!b) return fail(XD_ERR_INVALID_ARGUMENT, "null argument");
    if (trials == 0) trials = 3;
    hipError_t e = hipSetDevice(b->ctx->device);
    if (e != hipSuccess) return fail_hip(e, "hipSetDevice");
    const std::vector<xd_view> all = b->views;
    const xd_batch_tuning start = b->tuning;
    int rc = XD_OK;
    // tile shape per traversal, timed on that traversal's views alone, on
    // the kernel the whole batch runs (a batch of one traversal would
    // otherwise run that traversal's own kernel)
    uint32_t present = 0;
    for (int t = 0; t < 5; ++t)
        for (const auto& v : all)
            if (v.traversal == t) {
                ++present;
                break;
            }
    b->keep_mixed = present > 1 && !b->tuning.class_kernels;
    for (int t = 0; t < 5 && !rc; ++t) {
        std::vector<xd_view> sub;
        for (const auto& v : all)
            if (v.traversal == t) sub.push_back(v);
        if (sub.empty()) continue;
        b->views = sub;
        rc = sweep(b, trials, &b->tuning.tile_w[t], {8u, 16u, 32u});
        if (!rc) rc = sweep(b, trials, &b->tuning.rays_per_lane[t], {1u, 2u, 4u});
    }
    b->keep_mixed = false;
    // then the launch shape, on the whole batch, one field at a time from
    // the starting tuning
    b->views = all;
    if (!rc) rc = sweep(b, trials, &b->tuning.mix_classes, {0u, 1u, 2u});
    if (!rc) rc = sweep(b, trials, &b->tuning.interleave, {0u, 1u, 2u, 3u, 4u, 5u});
    if (!rc) rc = sweep(b, trials, &b->tuning.waves_per_simd, {0u, 8u});
    if (!rc) rc = sweep(b, trials, &b->tuning.class_kernels, {0u, 1u});
    if (!rc) rc = sweep(b, trials, &b->tuning.tiles_per_grab, {1u, 2u, 4u});
    if (rc) {
        b->tuning = start;
        (void)prepare(b);
        return rc;
    }
    if ((rc = prepare(b))) return rc;
    if (chosen) *chosen = b->tuning;
    return XD_OK;
}

// ---- allocation-free batched render ---------------------------------------
//
// The view and tile tables and the queue counters of xd_batch_render_static
// live in one file-scope device array instead of memory the library
// allocates.  In host code `scratch` only names the symbol: its device address
// (one per GPU, each GPU holding its own copy) comes from hipGetSymbolAddress,
// taken once per device with that device current, and only that address is
// passed to copies and kernels.

}  // extern "C"

namespace xd {
namespace {

constexpr uint32_t kStaticViews = 256;
constexpr uint32_t kStaticTiles = 1u << 21;  // 2M wave tiles: 134M rays at 64 per tile
constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr size_t kScratchQbeg = align256(sizeof(BatchView) * kStaticViews);
constexpr size_t kScratchCounters = kScratchQbeg + align256(sizeof(uint32_t) * (kClasses * kQueues + 1));
constexpr size_t kScratchTiles = kScratchCounters + align256(sizeof(uint32_t) * kClasses * kQueues);
constexpr size_t kScratchBytes = kScratchTiles + sizeof(uint32_t) * kStaticTiles;
static_assert(alignof(BatchView) <= 256, "view table alignment");

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch[kMaxDevices] = {};

}  // namespace
}  // namespace xd

static __device__ __align__(256) uint8_t scratch[xd::kScratchBytes];

namespace xd {
namespace {

// Device address of this device's `scratch` (the calling thread has made
// `device` current).
int scratch_address(int device, uint8_t** out) {
    if (device < 0 || device >= kMaxDevices) return fail(XD_ERR_INVALID_ARGUMENT, "device %d out of range", device);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return fail_hip(e, "hipGetSymbolAddress(scratch)");
        g_scratch[device] = static_cast<uint8_t*>(ptr);
    }
    *out = g_scratch[device];
    return XD_OK;
}

}  // namespace
}  // namespace xd

extern "C" {

int xd_batch_render_static(xd_context* ctx, const xd_view* views, uint32_t count, const xd_batch_tuning* tuning,
                           double* ms, uint64_t* rays) {
    if (!ctx || (!views && count)) return fail(XD_ERR_INVALID_ARGUMENT, "null argument");
    if (count > kStaticViews)
        return fail(XD_ERR_INVALID_ARGUMENT, "%u views exceed the static limit of %u", count, kStaticViews);
    const xd_batch_tuning tun = tuning ? *tuning : default_tuning();
    int rc = check_tuning(&tun);
    if (rc) return rc;
    if ((rc = check_views(ctx, views, count))) return rc;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return fail_hip(e, "hipSetDevice");
    uint8_t* base = nullptr;
    if ((rc = scratch_address(ctx->device, &base))) return rc;
    HostTables T;
    if ((rc = build_tables(views, count, tun, T))) return rc;
    if (rays) *rays = T.total_rays;
    if (ms) *ms = 0;
    if (count == 0) return XD_OK;
    const uint32_t total = (uint32_t)T.tiles.size();
    if (total > kStaticTiles)
        return fail(XD_ERR_INVALID_ARGUMENT, "%u wave tiles exceed the static limit of %u", total, kStaticTiles);
    const bool w8 = tun.waves_per_simd == 8;
    BatchClasses cls = T.cls;
    for (int c = 0; c <= (int)kClasses; ++c)
        if (cls.grid[c] && (rc = batch_grid(ctx->device, c, cls.slots[c], w8, &cls.grid[c]))) return rc;
    BatchView* d_views = reinterpret_cast<BatchView*>(base);
    uint32_t* d_qbeg = reinterpret_cast<uint32_t*>(base + kScratchQbeg);
    uint32_t* d_counters = reinterpret_cast<uint32_t*>(base + kScratchCounters);
    uint32_t* d_tiles = reinterpret_cast<uint32_t*>(base + kScratchTiles);
    hipStream_t s = ctx->stream;
    // launch_batch clears the counters with a kernel before the class kernels
    if ((e = hipMemcpyAsync(d_views, T.views.data(), sizeof(BatchView) * count, hipMemcpyHostToDevice, s)) != hipSuccess ||
        (total && (e = hipMemcpyAsync(d_tiles, T.tiles.data(), sizeof(uint32_t) * total, hipMemcpyHostToDevice, s)) != hipSuccess) ||
        (e = hipMemcpyAsync(d_qbeg, T.qbeg, sizeof(T.qbeg), hipMemcpyHostToDevice, s)) != hipSuccess ||
        (e = hipEventRecord(ctx->ev0, s)) != hipSuccess ||
        (e = launch_batch(d_views, d_tiles, d_qbeg, d_counters, pack_grabs(tun.tiles_per_grab, tun.rays_per_lane), cls,
                          w8, tun.mix_classes != 0, s)) !=
            hipSuccess ||
        (e = hipEventRecord(ctx->ev1, s)) != hipSuccess || (e = hipEventSynchronize(ctx->ev1)) != hipSuccess) {
        (void)hipStreamSynchronize(s);  // no copy may still read the host tables when they go
        return fail_hip(e, "static batch render");
    }
    if (ms) {
        float f = 0;
        if ((e = hipEventElapsedTime(&f, ctx->ev0, ctx->ev1)) != hipSuccess) return fail_hip(e, "hipEventElapsedTime");
        *ms = f;
    }
    return XD_OK;
}

int xd_batch_free(xd_batch* b) {
    if (!b) return XD_OK;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    drop_graph(b);
    if (b->d_views) (void)hipFree(b->d_views);
    if (b->d_tiles) (void)hipFree(b->d_tiles);
    if (b->d_qbeg) (void)hipFree(b->d_qbeg);
    if (b->d_counters) (void)hipFree(b->d_counters);
    b->ctx->batches.erase(b);
    delete b;
    return XD_OK;
}

}  // extern "C"
