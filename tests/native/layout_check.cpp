// Host check of the derived octree device layout (xenodon_amd/csrc/xd_layout.cpp)
// against the .svo words it is derived from.  Usage:
//   layout_check <nodes.words> <wide 0|1> <points>
// Reads node_count * 10 little-endian u32 words and checks:
//   * the numbering: node 0 stays 0, internal nodes first (the top levels
//     ahead), then leaves with colour, then empty leaves; every child word
//     names the renumbered child (the walk a plain sequential numbering would
//     give, recomputed here), tags = the child's leaf / empty-leaf state;
//   * meta[new(i)] = {color, is_leaf_depth} of node i;
//   * octree_depths_exact: the subtree-parallel check equals the breadth-first
//     walk on trees;
//   * the start table: for `points` seeded points, the node find() reaches
//     from the root over the .svo words (svo_naive.comp:8-27 in FP32) equals
//     the node reached by resuming the walk over the tagged words from the
//     table entry of the point's level-kTableLevel cell, with the same offset and side;
//   * rope depths (when packed): bits 25..29 of every leaf child word hold the
//     depth field of the node it names;
//   * the device form (to_rope_planes): the words of the nodes before the first
//     leaf unchanged, and rope k of every leaf at its place in plane k.
// Prints "ok" and exits 0, or prints the first failures and exits 1.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "xd_runtime.hpp"

using namespace xd;

static int failures = 0;
#define CHECK(cond, ...)                          \
    do {                                          \
        if (!(cond) && failures++ < 10) {         \
            std::printf("FAIL %s: ", #cond);      \
            std::printf(__VA_ARGS__);             \
            std::printf("\n");                    \
        }                                         \
    } while (0)

// plain sequential numbering: top levels breadth-first, then sibling groups
// depth-first, then unreachable nodes
static std::vector<uint32_t> reference_ids(const uint32_t* w, uint64_t n) {
    std::vector<uint32_t> id(n, 0xFFFFFFFFu);
    uint32_t next = 1;
    id[0] = 0;
    std::vector<uint32_t> level{0};
    for (uint32_t L = 1; L < kTopLevels; ++L) {
        std::vector<uint32_t> nxt;
        for (uint32_t o : level) {
            const uint32_t* nd = w + (uint64_t)o * kNodeWords;
            if (nd[9] & kLeafMask) continue;
            for (int k = 0; k < 8; ++k)
                if (id[nd[k]] == 0xFFFFFFFFu) {
                    id[nd[k]] = next++;
                    nxt.push_back(nd[k]);
                }
        }
        level.swap(nxt);
    }
    std::vector<uint32_t> stack(level.rbegin(), level.rend());
    while (!stack.empty()) {
        const uint32_t* nd = w + (uint64_t)stack.back() * kNodeWords;
        stack.pop_back();
        if (nd[9] & kLeafMask) continue;
        uint32_t fresh[8];
        int nf = 0;
        for (int k = 0; k < 8; ++k)
            if (id[nd[k]] == 0xFFFFFFFFu) {
                id[nd[k]] = next++;
                fresh[nf++] = nd[k];
            }
        for (int k = nf - 1; k >= 0; --k) stack.push_back(fresh[k]);
    }
    for (uint64_t i = 0; i < n; ++i)
        if (id[i] == 0xFFFFFFFFu) id[i] = next++;
    return id;
}

// find() over the .svo words in FP32, as the shader walks (svo_naive.comp:8-27)
static uint32_t find_svo(const uint32_t* w, const float pos[3], float off[3], float* side, uint32_t* level) {
    float extent = 1.0f;
    uint32_t index = 0;
    off[0] = off[1] = off[2] = 0.0f;
    for (uint32_t L = 0;; ++L) {
        if ((w[(uint64_t)index * kNodeWords + 9] & kLeafMask) || L > kMaxLevels) {
            *side = extent;
            *level = L;
            return index;
        }
        extent = extent * 0.5f;
        uint32_t child = 0;
        for (int a = 0; a < 3; ++a) {
            const bool m = pos[a] >= off[a] + extent;
            child |= (m ? 1u : 0u) << (2 - a);
            off[a] = off[a] + (m ? 1.0f : 0.0f) * extent;
        }
        index = w[(uint64_t)index * kNodeWords + child];
    }
}

static uint32_t level_bits(float p) {
    if (!(p > 0.0f)) return 0u;
    if (!(p < 1.0f)) return (1u << 23) - 1u;
    return (uint32_t)(p * 8388608.0f);
}

// the device's find(): the table entry of the point's cell, then the walk
// over the tagged child words
static uint32_t find_tagged(const uint32_t* cw, const uint2* table, uint32_t mask, const float pos[3], float off[3],
                            float* side) {
    uint32_t b[3];
    for (int a = 0; a < 3; ++a) b[a] = level_bits(pos[a]);
    constexpr uint32_t sh = 23u - kTableLevel;
    const uint2 e = table[((b[2] >> sh) << (2 * kTableLevel)) | ((b[1] >> sh) << kTableLevel) | (b[0] >> sh)];
    uint32_t index = e.x & mask, tag = e.x, level = e.y;
    float extent = std::ldexp(1.0f, -(int)level);
    for (int a = 0; a < 3; ++a) off[a] = level == 0 ? 0.0f : extent * (float)(b[a] >> (23u - level));
    for (;; ++level) {
        if ((tag & kTagLeaf) || level > kMaxLevels) {
            *side = extent;
            return index;
        }
        extent = extent * 0.5f;
        uint32_t child = 0;
        for (int a = 0; a < 3; ++a) {
            const bool m = pos[a] >= off[a] + extent;
            child |= (m ? 1u : 0u) << (2 - a);
            off[a] = off[a] + (m ? 1.0f : 0.0f) * extent;
        }
        const uint32_t wd = cw[(uint64_t)index * 8 + child];
        index = wd & mask;
        tag = wd;
    }
}

int main(int argc, char** argv) {
    if (argc != 4) return 2;
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    std::vector<uint32_t> w;
    uint32_t buf[4096];
    size_t r;
    while ((r = std::fread(buf, 4, 4096, f)) > 0) w.insert(w.end(), buf, buf + r);
    std::fclose(f);
    const uint64_t n = w.size() / kNodeWords;
    const bool wide = std::atoi(argv[2]) != 0;
    const uint64_t points = std::strtoull(argv[3], nullptr, 10);

    const bool tree = n >= (1ull << 22) && octree_is_tree(w.data(), n);
    std::vector<uint32_t> cw(n * 8), meta(n * 2);
    uint32_t top[kTopLevels];
    uint32_t empty_start = 0, leaf_base = 0;
    build_tagged_layout(w.data(), n, cw.data(), meta.data(), top, wide, tree, &empty_start, &leaf_base);
    std::vector<uint32_t> id = reference_ids(w.data(), n);
    {
        // internal nodes (and the root), then leaves with colour, then empty
        // leaves, each in the order of the sequential numbering
        std::vector<uint32_t> order(n);
        for (uint64_t i = 0; i < n; ++i) order[id[i]] = (uint32_t)i;
        auto cls = [&](uint32_t i) {
            const uint32_t* nd = w.data() + (uint64_t)i * kNodeWords;
            if (i == 0 || !(nd[9] & kLeafMask)) return 0;
            return (nd[8] & 0xFFFFFFu) ? 1 : 2;
        };
        uint32_t count[3] = {0, 0, 0};
        for (uint64_t k = 0; k < n; ++k) ++count[cls(order[k])];
        const uint32_t kept = count[0] + count[1];
        CHECK(empty_start == (wide ? kept : 0xFFFFFFFFu), "empty_start %u (wide %d, %u non-empty)", empty_start,
              (int)wide, kept);
        CHECK(leaf_base == count[0], "leaf_base %u, %u internal nodes and root", leaf_base, count[0]);
        uint32_t at[3] = {0, count[0], kept};
        for (uint64_t k = 0; k < n; ++k) id[order[k]] = at[cls(order[k])]++;
    }
    const uint32_t empty_tag = wide ? 0u : kTagEmpty;
    for (uint64_t i = 0; i < n; ++i) {
        const uint32_t* nd = w.data() + i * kNodeWords;
        CHECK(meta[(uint64_t)id[i] * 2] == nd[8] && meta[(uint64_t)id[i] * 2 + 1] == nd[9], "meta of node %llu",
              (unsigned long long)i);
        for (int k = 0; k < 8; ++k) {
            const uint32_t* cn = w.data() + (uint64_t)nd[k] * kNodeWords;
            uint32_t tag = 0;
            if (cn[9] & kLeafMask) tag = kTagLeaf | ((cn[8] & 0xFFFFFFu) ? 0u : empty_tag);
            const uint32_t want = id[nd[k]] | tag;
            CHECK(cw[(uint64_t)id[i] * 8 + k] == want, "child word %d of node %llu: %08x, want %08x", k,
                  (unsigned long long)i, cw[(uint64_t)id[i] * 8 + k], want);
        }
    }
    CHECK(octree_depths_exact(w.data(), n, tree) == octree_depths_exact(w.data(), n, false), "depth check (tree %d)",
          (int)tree);

    // rope depths, as the volume upload packs them
    const bool packed = !wide && pack_rope_depths(cw.data(), meta.data(), n);
    const uint32_t mask = packed ? kRopeIndexMask : wide ? kWideIndexMask : kIndexMask;
    if (packed)
        for (uint64_t i = 0; i < n; ++i) {
            if (!(w[i * kNodeWords + 9] & kLeafMask)) continue;
            for (int k = 0; k < 8; ++k) {
                const uint32_t wd = cw[(uint64_t)id[i] * 8 + k];
                const uint32_t target = w[i * kNodeWords + k];
                CHECK(((wd >> kRopeDepthShift) & 31u) == (w[(uint64_t)target * kNodeWords + 9] & kDepthMask) &&
                          (wd & kRopeIndexMask) == id[target],
                      "rope %d of leaf %llu", k, (unsigned long long)i);
            }
        }

    // the device form: ropes of the leaves as planes
    {
        std::vector<uint32_t> planes(rope_plane_words(n, leaf_base));
        to_rope_planes(cw.data(), n, leaf_base, planes.data());
        CHECK(planes.size() == (uint64_t)leaf_base * 8 + 6 * (n - leaf_base), "plane words %zu", planes.size());
        CHECK(std::memcmp(planes.data(), cw.data(), (uint64_t)leaf_base * 8 * sizeof(uint32_t)) == 0,
              "words before the first leaf");
        for (uint64_t l = leaf_base; l < n; ++l)
            for (uint64_t k = 0; k < 6; ++k)
                CHECK(planes[(uint64_t)leaf_base * 8 + k * (n - leaf_base) + (l - leaf_base)] == cw[l * 8 + k],
                      "rope %llu of leaf %llu in its plane", (unsigned long long)k, (unsigned long long)l);
    }

    const uint32_t root_tag =
        (w[9] & kLeafMask) ? (kTagLeaf | ((w[8] & 0xFFFFFFu) ? 0u : empty_tag)) : 0u;
    std::vector<uint2> table(kTableCells);
    build_start_table(cw.data(), root_tag, mask, table.data());
    uint64_t s = 0x9E3779B97F4A7C15ull;
    auto next_float = [&s]() {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        return (float)((double)(s >> 11) / 9007199254740992.0 * 1.2 - 0.1);
    };
    for (uint64_t q = 0; q < points; ++q) {
        float pos[3] = {next_float(), next_float(), next_float()};
        if (q % 7 == 0)  // on a table cell boundary
            pos[q % 3] = std::ldexp((float)(s & ((1u << kTableLevel) - 1u)), -(int)kTableLevel);
        float o1[3], o2[3], s1, s2;
        uint32_t lvl;
        const uint32_t a = find_svo(w.data(), pos, o1, &s1, &lvl);
        const uint32_t b = find_tagged(cw.data(), table.data(), mask, pos, o2, &s2);
        CHECK(id[a] == b && s1 == s2 && std::memcmp(o1, o2, sizeof o1) == 0, "find at (%a, %a, %a): %u vs %u",
              pos[0], pos[1], pos[2], id[a], b);
    }
    if (failures) {
        std::printf("%d failures\n", failures);
        return 1;
    }
    std::printf("ok: %llu nodes, tree %d, wide %d, ropes packed %d\n", (unsigned long long)n, (int)tree, (int)wide,
                (int)packed);
    return 0;
}
