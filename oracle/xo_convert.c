/*
 * Oracle converter: restatement of `xenodon convert` (src/convert.cpp:13-122).
 * TEST INFRASTRUCTURE ONLY (see xo.h).
 *
 *   xo_vol_scan        Grid::vol_scan      (src/model/Grid.cpp:81-137)
 *   xo_stddev_scan     Grid::stddev_scan   (src/model/Grid.cpp:139-214)
 *   xo_build_octree    build_octree / detail::construct / OctreeBuilder
 *                      (src/model/OctreeConstruction.h:13-237)
 *   xo_octree_find     Octree::find        (src/model/Octree.cpp:116-153)
 *   xo_generate_ropes  Octree::generate_ropes (src/model/Octree.cpp:155-201)
 *
 * Parity: pinned by hand-derived trees (single voxel in 8^3: 25 sparse nodes,
 * 7 DAG nodes), lossless-reconstruction, DAG-expansion and rope-neighbour
 * invariants, and brute-force scans (tests/test_oracle_convert.py).
 */
#include "xo.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NODE_WORDS 10
#define LEAF 0x80000000u

void xo_vol_scan(const uint8_t* grid, const uint64_t dims[3], const uint64_t bmin_in[3], const uint64_t bmax_in[3],
                 uint32_t* avg, uint32_t* max_diff) {
    uint64_t bmin[3], bmax[3];
    for (int i = 0; i < 3; ++i) {
        bmin[i] = bmin_in[i] < dims[i] ? bmin_in[i] : dims[i];
        bmax[i] = bmax_in[i] < dims[i] ? bmax_in[i] : dims[i];
    }
    uint64_t n = (bmax[0] - bmin[0]) * (bmax[1] - bmin[1]) * (bmax[2] - bmin[2]);
    if (n == 0) {
        *avg = 0;
        *max_diff = 0;
        return;
    }
    uint8_t mn[4] = {255, 255, 255, 255}, mx[4] = {0, 0, 0, 0};
    uint64_t acc[4] = {0, 0, 0, 0};
    for (uint64_t z = bmin[2]; z < bmax[2]; ++z) {
        for (uint64_t y = bmin[1]; y < bmax[1]; ++y) {
            const uint8_t* row = grid + ((z * dims[1] + y) * dims[0]) * 4;
            for (uint64_t x = bmin[0]; x < bmax[0]; ++x) {
                const uint8_t* p = row + x * 4;
                for (int c = 0; c < 4; ++c) {
                    if (p[c] < mn[c])
                        mn[c] = p[c];
                    if (p[c] > mx[c])
                        mx[c] = p[c];
                    acc[c] += p[c];
                }
            }
        }
    }
    uint32_t md = 0;
    uint32_t a = 0;
    for (int c = 0; c < 4; ++c) {
        uint32_t d = (uint32_t)(mx[c] - mn[c]);
        if (d > md)
            md = d;
        a |= (uint32_t)(uint8_t)(acc[c] / n) << (8 * c);
    }
    *avg = a;
    *max_diff = md;
}

void xo_stddev_scan(const uint8_t* grid, const uint64_t dims[3], const uint64_t bmin_in[3], const uint64_t bmax_in[3],
                    uint32_t* avg_out, double* stddev_out) {
    uint64_t bmin[3], bmax[3];
    for (int i = 0; i < 3; ++i) {
        bmin[i] = bmin_in[i] < dims[i] ? bmin_in[i] : dims[i];
        bmax[i] = bmax_in[i] < dims[i] ? bmax_in[i] : dims[i];
    }
    uint64_t n = (bmax[0] - bmin[0]) * (bmax[1] - bmin[1]) * (bmax[2] - bmin[2]);
    if (n == 0) {
        *avg_out = 0;
        *stddev_out = 0;
        return;
    }
    uint64_t acc[4] = {0, 0, 0, 0};
    for (uint64_t z = bmin[2]; z < bmax[2]; ++z)
        for (uint64_t y = bmin[1]; y < bmax[1]; ++y)
            for (uint64_t x = bmin[0]; x < bmax[0]; ++x) {
                const uint8_t* p = grid + (((z * dims[1] + y) * dims[0]) + x) * 4;
                for (int c = 0; c < 4; ++c)
                    acc[c] += p[c];
            }
    double nd = (double)n;
    double avg[4];
    for (int c = 0; c < 4; ++c)
        avg[c] = (double)acc[c] / nd;
    double s = 0;
    for (uint64_t z = bmin[2]; z < bmax[2]; ++z)
        for (uint64_t y = bmin[1]; y < bmax[1]; ++y)
            for (uint64_t x = bmin[0]; x < bmax[0]; ++x) {
                const uint8_t* p = grid + (((z * dims[1] + y) * dims[0]) + x) * 4;
                double dr = (double)p[0] - avg[0];
                double dg = (double)p[1] - avg[1];
                double db = (double)p[2] - avg[2];
                double da = (double)p[3] - avg[3];
                s += dr * dr + dg * dg + db * db + da * da;
            }
    uint32_t a = 0;
    for (int c = 0; c < 4; ++c)
        a |= (uint32_t)(uint8_t)avg[c] << (8 * c);
    *avg_out = a;
    *stddev_out = sqrt(s / nd);
}

/* ---- OctreeBuilder with NoopCache / HashCache (OctreeConstruction.h:13-112) ---- */

typedef struct {
    uint32_t* nodes; /* NODE_WORDS per node, insertion order */
    uint64_t count, capacity;
    int dag;
    uint64_t* table; /* open addressing, holds index + 1 (0 = empty) */
    uint64_t table_cap;
    int failed;
} builder;

static uint64_t node_hash(const uint32_t* w) {
    uint64_t h = 1469598103934665603ull;
    for (int i = 0; i < NODE_WORDS; ++i) {
        h ^= w[i];
        h *= 1099511628211ull;
        h ^= h >> 29;
    }
    return h;
}

static int table_grow(builder* b) {
    uint64_t ncap = b->table_cap ? b->table_cap * 2 : 1024;
    uint64_t* nt = (uint64_t*)calloc(ncap, sizeof(uint64_t));
    if (!nt)
        return -1;
    for (uint64_t i = 0; i < b->table_cap; ++i) {
        uint64_t e = b->table[i];
        if (!e)
            continue;
        uint64_t h = node_hash(b->nodes + (e - 1) * NODE_WORDS) & (ncap - 1);
        while (nt[h])
            h = (h + 1) & (ncap - 1);
        nt[h] = e;
    }
    free(b->table);
    b->table = nt;
    b->table_cap = ncap;
    return 0;
}

/* Returns the index the node lives at; *inserted says whether it was new. */
static uint32_t builder_insert(builder* b, const uint32_t* node, int* inserted) {
    if (b->dag) {
        if ((b->count + 1) * 2 > b->table_cap && table_grow(b)) {
            b->failed = 1;
            return 0;
        }
        uint64_t h = node_hash(node) & (b->table_cap - 1);
        while (b->table[h]) {
            uint64_t e = b->table[h] - 1;
            if (memcmp(b->nodes + e * NODE_WORDS, node, NODE_WORDS * 4) == 0) {
                *inserted = 0;
                return (uint32_t)e;
            }
            h = (h + 1) & (b->table_cap - 1);
        }
        b->table[h] = b->count + 1;
    }
    if (b->count == b->capacity) {
        uint64_t ncap = b->capacity ? b->capacity * 2 : 4096;
        uint32_t* nn = (uint32_t*)realloc(b->nodes, ncap * NODE_WORDS * 4);
        if (!nn) {
            b->failed = 1;
            return 0;
        }
        b->nodes = nn;
        b->capacity = ncap;
    }
    memcpy(b->nodes + b->count * NODE_WORDS, node, NODE_WORDS * 4);
    *inserted = 1;
    return (uint32_t)b->count++;
}

typedef struct {
    const uint8_t* grid;
    const uint64_t* dims;
    int heuristic;
    int chan_diff;
    double std_dev;
    builder b;
    xo_build_stats* stats;
} ctx_t;

static uint32_t insert_counted(ctx_t* c, const uint32_t* node, int leaf) {
    int inserted = 0;
    uint32_t idx = builder_insert(&c->b, node, &inserted);
    c->stats->total_nodes++;
    if (leaf) {
        c->stats->total_leaves++;
        if (inserted)
            c->stats->unique_leaves++;
    }
    return idx;
}

/* detail::construct (OctreeConstruction.h:124-194) */
static uint32_t construct(ctx_t* c, const uint64_t off[3], uint64_t extent, uint64_t depth) {
    if (depth > c->stats->depth)
        c->stats->depth = depth;
    uint32_t node[NODE_WORDS];
    memset(node, 0, sizeof(node));
    int inside_any = off[0] < c->dims[0] && off[1] < c->dims[1] && off[2] < c->dims[2];
    if (!inside_any) {
        node[9] = LEAF | (uint32_t)depth;
        return insert_counted(c, node, 1);
    }
    int fully_inside = off[0] + extent <= c->dims[0] && off[1] + extent <= c->dims[1] && off[2] + extent <= c->dims[2];
    uint64_t bmax[3] = {off[0] + extent, off[1] + extent, off[2] + extent};
    uint32_t avg;
    int split;
    if (c->heuristic == XO_HEUR_CHAN_DIFF) {
        uint32_t md;
        xo_vol_scan(c->grid, c->dims, off, bmax, &avg, &md);
        split = md > (uint32_t)c->chan_diff;
    } else {
        double sd;
        xo_stddev_scan(c->grid, c->dims, off, bmax, &avg, &sd);
        split = sd > c->std_dev;
    }
    if ((!split && fully_inside) || extent == 1) {
        node[8] = avg;
        node[9] = LEAF | (uint32_t)depth;
        return insert_counted(c, node, 1);
    }
    uint64_t h = extent / 2;
    node[8] = avg;
    node[9] = (uint32_t)depth;
    int child = 0;
    for (int xi = 0; xi < 2; ++xi)
        for (int yi = 0; yi < 2; ++yi)
            for (int zi = 0; zi < 2; ++zi) {
                uint64_t co[3] = {off[0] + (xi ? h : 0), off[1] + (yi ? h : 0), off[2] + (zi ? h : 0)};
                node[child++] = construct(c, co, h, depth + 1);
            }
    return insert_counted(c, node, 0);
}

static uint64_t ceil_2pow(uint64_t x) {
    --x;
    x |= x >> 1;
    x |= x >> 2;
    x |= x >> 4;
    x |= x >> 8;
    x |= x >> 16;
    x |= x >> 32;
    return ++x;
}

int xo_build_octree(const uint8_t* grid, const uint64_t dims[3], int heuristic, int chan_diff, double std_dev,
                    int tree_type, uint32_t** out_nodes, xo_build_stats* stats) {
    if (!grid || !dims || !out_nodes || !stats)
        return -1;
    memset(stats, 0, sizeof(*stats));
    uint64_t side = ceil_2pow(dims[0]);
    if (ceil_2pow(dims[1]) > side)
        side = ceil_2pow(dims[1]);
    if (ceil_2pow(dims[2]) > side)
        side = ceil_2pow(dims[2]);
    ctx_t c;
    memset(&c, 0, sizeof(c));
    c.grid = grid;
    c.dims = dims;
    c.heuristic = heuristic;
    c.chan_diff = chan_diff;
    c.std_dev = std_dev;
    c.b.dag = tree_type == XO_TREE_DAG;
    c.stats = stats;
    uint64_t zero[3] = {0, 0, 0};
    construct(&c, zero, side, 0);
    free(c.b.table);
    if (c.b.failed) {
        free(c.b.nodes);
        return -2;
    }
    /* OctreeBuilder::build (OctreeConstruction.h:92-112): reverse, then remap. */
    uint64_t n = c.b.count;
    uint32_t* nodes = c.b.nodes;
    for (uint64_t i = 0; i < n / 2; ++i) {
        uint32_t tmp[NODE_WORDS];
        memcpy(tmp, nodes + i * NODE_WORDS, sizeof(tmp));
        memcpy(nodes + i * NODE_WORDS, nodes + (n - 1 - i) * NODE_WORDS, sizeof(tmp));
        memcpy(nodes + (n - 1 - i) * NODE_WORDS, tmp, sizeof(tmp));
    }
    uint32_t end = (uint32_t)(n - 1);
    for (uint64_t i = 0; i < n; ++i) {
        uint32_t* nd = nodes + i * NODE_WORDS;
        for (int k = 0; k < 8; ++k)
            nd[k] = (nd[9] & LEAF) ? 0u : end - nd[k];
    }
    stats->side = side;
    stats->node_count = n;
    if (tree_type == XO_TREE_ROPE)
        xo_generate_ropes(nodes, n, side);
    *out_nodes = nodes;
    return 0;
}

uint64_t xo_octree_find(const uint32_t* nodes, uint64_t side, const uint64_t pos[3], uint64_t max_depth) {
    uint64_t extent = side;
    if (pos[0] >= extent || pos[1] >= extent || pos[2] >= extent)
        return UINT64_MAX;
    uint64_t index = 0;
    uint64_t off[3] = {0, 0, 0};
    for (;;) {
        extent /= 2;
        if ((nodes[index * NODE_WORDS + 9] & LEAF) || extent == 0 || max_depth == 0)
            return index;
        uint32_t ci = 0;
        for (int a = 0; a < 3; ++a) {
            if (pos[a] >= off[a] + extent) {
                ci |= 4u >> a;
                off[a] += extent;
            }
        }
        index = nodes[index * NODE_WORDS + ci];
        --max_depth;
    }
}

static void rope_walk(uint32_t* nodes, uint64_t side, uint64_t index, const uint64_t pos[3], uint64_t extent,
                      uint64_t depth) {
    uint32_t* nd = nodes + index * NODE_WORDS;
    if (nd[9] & LEAF) {
        uint64_t r[6];
        for (int a = 0; a < 3; ++a) {
            uint64_t p[3] = {pos[0], pos[1], pos[2]};
            p[a] = pos[a] + extent;
            uint64_t f = xo_octree_find(nodes, side, p, depth);
            r[2 * a] = f == UINT64_MAX ? 0 : f;
            p[a] = pos[a] - extent; /* wraps below zero exactly like size_t */
            f = xo_octree_find(nodes, side, p, depth);
            r[2 * a + 1] = f == UINT64_MAX ? 0 : f;
        }
        for (int k = 0; k < 6; ++k)
            nd[k] = (uint32_t)r[k];
        return;
    }
    uint64_t h = extent / 2;
    int child = 0;
    for (int xi = 0; xi < 2; ++xi)
        for (int yi = 0; yi < 2; ++yi)
            for (int zi = 0; zi < 2; ++zi) {
                uint64_t cp[3] = {pos[0] + (xi ? h : 0), pos[1] + (yi ? h : 0), pos[2] + (zi ? h : 0)};
                rope_walk(nodes, side, nd[child++], cp, h, depth + 1);
            }
}

int xo_generate_ropes(uint32_t* nodes, uint64_t node_count, uint64_t side) {
    if (!nodes || node_count == 0)
        return -1;
    uint64_t zero[3] = {0, 0, 0};
    rope_walk(nodes, side, 0, zero, side, 0);
    return 0;
}
