// Host-side scene bookkeeping (see vgi_host_scene.h).
#include "vgi_host_scene.h"
#include "vgi_pool.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace vgi_host {

static inline size_t cell_index(int x, int y, int z, int sx, int sy) {
    return (size_t)x + (size_t)sx * ((size_t)y + (size_t)sy * (size_t)z);
}

// naadf.rs:117-148 (try_grow): grow `cur` one cell toward `direction` when the
// neighbour is inside, empty, and its box covers the slab the growth adds.
static inline void grow_toward(Extents &cur, int nx, int ny, int nz, int direction, int sx, int sy, int sz,
                               const uint8_t *blocked, const Extents *fields) {
    if (nx < 0 || ny < 0 || nz < 0 || nx >= sx || ny >= sy || nz >= sz) return;
    if (cur[direction] == 255) return;
    size_t n = cell_index(nx, ny, nz, sx, sy);
    if (blocked[n]) return;
    const Extents &o = fields[n];
    int opposite = direction ^ 1;
    for (int d = 0; d < 6; ++d)
        if (d != opposite && o[d] < cur[d]) return;
    cur[direction] += 1;
}

static const int STEP[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};

// The FIFO work list of world.rs:838-912 (Growth): a cell is queued at most
// once; a cell that grows re-queues itself and its six neighbours.
struct GrowthQueue {
    int sx, sy, sz;
    std::deque<uint32_t> fifo;
    std::vector<uint8_t> queued;
    GrowthQueue(int x, int y, int z) : sx(x), sy(y), sz(z), queued((size_t)x * y * z, 0) {}
    void push(uint32_t i) {
        if (!queued[i]) {
            queued[i] = 1;
            fifo.push_back(i);
        }
    }
    void touch(uint32_t i) {
        push(i);
        int x = (int)(i % (uint32_t)sx), y = (int)((i / (uint32_t)sx) % (uint32_t)sy), z = (int)(i / ((uint32_t)sx * sy));
        for (int d = 0; d < 6; ++d) {
            int nx = x + STEP[d][0], ny = y + STEP[d][1], nz = z + STEP[d][2];
            if (nx >= 0 && ny >= 0 && nz >= 0 && nx < sx && ny < sy && nz < sz) push((uint32_t)cell_index(nx, ny, nz, sx, sy));
        }
    }
    void drain(std::vector<Extents> &fields, const std::vector<uint8_t> &blocked) {
        while (!fifo.empty()) {
            uint32_t i = fifo.front();
            fifo.pop_front();
            queued[i] = 0;
            if (blocked[i]) continue;
            int x = (int)(i % (uint32_t)sx), y = (int)((i / (uint32_t)sx) % (uint32_t)sy), z = (int)(i / ((uint32_t)sx * sy));
            Extents cur = fields[i];
            for (int d = 0; d < 6; ++d)
                grow_toward(cur, x + STEP[d][0], y + STEP[d][1], z + STEP[d][2], d, sx, sy, sz, blocked.data(), fields.data());
            if (cur != fields[i]) {
                fields[i] = cur;
                touch(i);
            }
        }
    }
};

// world.rs:733-833 (cut_boxes): boxes holding a newly blocked cell shrink
// along the one direction keeping the most volume (or to their own cell).
static void shrink_boxes(std::vector<Extents> &fields, const std::vector<uint8_t> &blocked, int sx, int sy, int sz,
                         const std::vector<Int3> &newly, GrowthQueue &queue) {
    Int3 lo = newly[0], hi = newly[0];
    for (const Int3 &c : newly) {
        lo = Int3{std::min(lo.x, c.x), std::min(lo.y, c.y), std::min(lo.z, c.z)};
        hi = Int3{std::max(hi.x, c.x), std::max(hi.y, c.y), std::max(hi.z, c.z)};
    }
    int tx = hi.x - lo.x + 2, ty = hi.y - lo.y + 2, tz = hi.z - lo.z + 2;
    std::vector<uint32_t> sat((size_t)tx * ty * tz, 0);
    for (const Int3 &c : newly) sat[cell_index(c.x - lo.x + 1, c.y - lo.y + 1, c.z - lo.z + 1, tx, ty)] += 1;
    for (int z = 0; z < tz; ++z)
        for (int y = 0; y < ty; ++y)
            for (int x = 1; x < tx; ++x) sat[cell_index(x, y, z, tx, ty)] += sat[cell_index(x - 1, y, z, tx, ty)];
    for (int z = 0; z < tz; ++z)
        for (int y = 1; y < ty; ++y)
            for (int x = 0; x < tx; ++x) sat[cell_index(x, y, z, tx, ty)] += sat[cell_index(x, y - 1, z, tx, ty)];
    for (int z = 1; z < tz; ++z)
        for (int y = 0; y < ty; ++y)
            for (int x = 0; x < tx; ++x) sat[cell_index(x, y, z, tx, ty)] += sat[cell_index(x, y, z - 1, tx, ty)];
    auto inside = [&](int ax, int ay, int az, int bx, int by, int bz) -> uint32_t {
        ax = std::max(ax, lo.x); ay = std::max(ay, lo.y); az = std::max(az, lo.z);
        bx = std::min(bx, hi.x); by = std::min(by, hi.y); bz = std::min(bz, hi.z);
        if (ax > bx || ay > by || az > bz) return 0;
        ax -= lo.x; ay -= lo.y; az -= lo.z;
        bx = bx - lo.x + 1; by = by - lo.y + 1; bz = bz - lo.z + 1;
        auto at = [&](int x, int y, int z) { return (int64_t)sat[cell_index(x, y, z, tx, ty)]; };
        return (uint32_t)(at(bx, by, bz) - at(ax, by, bz) - at(bx, ay, bz) - at(bx, by, az) + at(ax, ay, bz) + at(ax, by, az) +
                          at(bx, ay, az) - at(ax, ay, az));
    };
    auto holds = [&](const Extents &f, int x, int y, int z) {
        return inside(x - f[0], y - f[2], z - f[4], x + f[1], y + f[3], z + f[5]);
    };
    const Extents zero{0, 0, 0, 0, 0, 0};
    for (int z = 0; z < sz; ++z)
        for (int y = 0; y < sy; ++y)
            for (int x = 0; x < sx; ++x) {
                size_t i = cell_index(x, y, z, sx, sy);
                if (blocked[i] || fields[i] == zero) continue;
                if (holds(fields[i], x, y, z) == 0) continue;
                bool have = false;
                uint64_t most = 0;
                Extents best = zero;
                for (int d = 0; d < 6; ++d) {
                    Extents f = fields[i];
                    f[d] = 0;
                    if (holds(f, x, y, z) != 0) continue;
                    // Largest clear extent by bisection: clear at 0, not at full.
                    uint8_t good = 0, bad = fields[i][d];
                    while (bad - good > 1) {
                        uint8_t mid = good + (bad - good) / 2;
                        f[d] = mid;
                        if (holds(f, x, y, z) == 0) good = mid; else bad = mid;
                    }
                    f[d] = good;
                    uint64_t volume = (uint64_t)(f[0] + f[1] + 1) * (f[2] + f[3] + 1) * (f[4] + f[5] + 1);
                    if (!have || volume > most) { have = true; most = volume; best = f; }
                }
                fields[i] = have ? best : zero;
                queue.push((uint32_t)i);
            }
}

void TopLevelField::update(Int3 mn, Int3 mx, const std::vector<Int3> &changed, const std::vector<uint8_t> &barrier) {
    int nx = mx.x - mn.x + 1, ny = mx.y - mn.y + 1, nz = mx.z - mn.z + 1;
    uint64_t cells = (uint64_t)nx * ny * nz;
    // world.rs:486-493: no field beyond 255 cells per axis or 4M cells.
    if (std::max(std::max(nx, ny), nz) > 255 || cells > (4ull << 20)) {
        minimum = mn; maximum = mx; sx = nx; sy = ny; sz = nz;
        enabled = false;
        fields.clear();
        blocked.clear();
        return;
    }
    GrowthQueue queue(nx, ny, nz);
    std::vector<Extents> nf;
    std::vector<uint8_t> nb;
    const Extents zero{0, 0, 0, 0, 0, 0};
    bool in_place = enabled && minimum == mn && sx == nx && sy == ny && sz == nz;
    if (in_place) {
        nf = fields;
        nb = blocked;
    } else {
        nf.assign(cells, zero);
        nb.assign(cells, 1);
        for (int z = mn.z; z <= mx.z; ++z)
            for (int y = mn.y; y <= mx.y; ++y)
                for (int x = mn.x; x <= mx.x; ++x) {
                    size_t i = cell_index(x - mn.x, y - mn.y, z - mn.z, nx, ny);
                    bool kept = enabled && x >= minimum.x && y >= minimum.y && z >= minimum.z && x <= maximum.x &&
                                y <= maximum.y && z <= maximum.z;
                    if (kept) {
                        size_t old = cell_index(x - minimum.x, y - minimum.y, z - minimum.z, sx, sy);
                        nb[i] = blocked[old];
                        // world.rs:715-722 (clip) to the new bounds.
                        Extents f = fields[old];
                        int below[3] = {x - mn.x, y - mn.y, z - mn.z}, above[3] = {mx.x - x, mx.y - y, mx.z - z};
                        for (int a = 0; a < 3; ++a) {
                            f[a * 2] = (uint8_t)std::min<int>(f[a * 2], std::min(below[a], 255));
                            f[a * 2 + 1] = (uint8_t)std::min<int>(f[a * 2 + 1], std::min(above[a], 255));
                        }
                        nf[i] = f;
                    } else {
                        nb[i] = barrier[i];
                        if (!nb[i]) queue.touch((uint32_t)i);
                    }
                }
    }
    std::vector<Int3> newly;
    for (const Int3 &c : changed) {
        if (c.x < mn.x || c.y < mn.y || c.z < mn.z || c.x > mx.x || c.y > mx.y || c.z > mx.z) continue;
        size_t i = cell_index(c.x - mn.x, c.y - mn.y, c.z - mn.z, nx, ny);
        uint8_t now = barrier[i];
        if (now == nb[i]) continue;
        nb[i] = now;
        if (now) {
            nf[i] = zero;
            newly.push_back(Int3{c.x - mn.x, c.y - mn.y, c.z - mn.z});
        } else {
            queue.touch((uint32_t)i);
        }
    }
    if (!newly.empty()) shrink_boxes(nf, nb, nx, ny, nz, newly, queue);
    queue.drain(nf, nb);
    minimum = mn; maximum = mx; sx = nx; sy = ny; sz = nz;
    enabled = true;
    fields.swap(nf);
    blocked.swap(nb);
}

Extents TopLevelField::at(Int3 c) const {
    if (!enabled) return Extents{0, 0, 0, 0, 0, 0};
    return fields[cell_index(c.x - minimum.x, c.y - minimum.y, c.z - minimum.z, sx, sy)];
}

// ---------------------------------------------------------------- emitters
static inline uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
static const int FACE_STEP[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};  // world.rs:50-57

static inline int floor_div(int a, int b) { int q = a / b; if ((a % b) != 0 && ((a < 0) != (b < 0))) --q; return q; }

uint32_t VoxelSource::material_at(int32_t x, int32_t y, int32_t z) const {
    int rx = floor_div(x, 16), ry = floor_div(y, 16), rz = floor_div(z, 16);
    if (rx < region_min.x || ry < region_min.y || rz < region_min.z || rx >= region_min.x + rsx || ry >= region_min.y + rsy ||
        rz >= region_min.z + rsz)
        return 0xffffu;
    if (!loaded[cell_index(rx - region_min.x, ry - region_min.y, rz - region_min.z, rsx, rsy)]) return 0xffffu;
    return grid[cell_index(x - origin.x, y - origin.y, z - origin.z, gx, gy)];
}

struct FaceRec {
    int32_t x, y, z;
    uint32_t face, material;
    float cumulative, probability;
};
struct TreeNode {
    float mn[3], mx[3], power;
    uint32_t normals, face, left, right, axis;
    int32_t pivot[3];
};

// emission.rs:206-216 at (u, v) = (0.5, 0.5), the face's centre.
static inline void face_centre(const FaceRec &f, float c[3]) {
    float n[3] = {(float)FACE_STEP[f.face][0], (float)FACE_STEP[f.face][1], (float)FACE_STEP[f.face][2]};
    uint32_t axis = f.face / 2;
    float t[3] = {axis == 0 ? 0.0f : 1.0f, axis == 0 ? 1.0f : 0.0f, 0.0f};
    float b[3] = {0.0f, axis == 2 ? 1.0f : 0.0f, axis == 2 ? 0.0f : 1.0f};
    float v[3] = {(float)f.x, (float)f.y, (float)f.z};
    for (int a = 0; a < 3; ++a) c[a] = v[a] + 0.5f + n[a] * 0.5f + t[a] * (0.5f - 0.5f) + b[a] * (0.5f - 0.5f);
}

static inline bool tree_less(const FaceRec &p, const FaceRec &q, uint32_t axis) {
    int32_t pv[3] = {p.x, p.y, p.z}, qv[3] = {q.x, q.y, q.z};
    for (uint32_t j = 0; j < 3; ++j) {
        uint32_t i = (axis + j) % 3;
        if (pv[i] != qv[i]) return pv[i] < qv[i];
    }
    return p.face < q.face;
}

// tree.rs:28-74: pre-order nodes; split at the median of the longest axis.
static uint32_t split_node(std::vector<TreeNode> &nodes, uint32_t *idx, size_t n, const std::vector<FaceRec> &faces) {
    uint32_t index = (uint32_t)nodes.size();
    TreeNode node{};
    const float inf = INFINITY;
    for (int a = 0; a < 3; ++a) { node.mn[a] = inf; node.mx[a] = -inf; }
    node.power = 0.0f;
    for (size_t k = 0; k < n; ++k) {
        const FaceRec &f = faces[idx[k]];
        float c[3];
        face_centre(f, c);
        for (int a = 0; a < 3; ++a) {
            float e = (1.0f - std::fabs((float)FACE_STEP[f.face][a])) * 0.5f;
            float lo = c[a] - e, hi = c[a] + e;
            node.mn[a] = lo < node.mn[a] ? lo : node.mn[a];
            node.mx[a] = node.mx[a] < hi ? hi : node.mx[a];
        }
        node.power = node.power + f.probability;
        node.normals |= 1u << f.face;
    }
    nodes.push_back(node);
    if (n == 1) {
        node.left = NONE;
        node.right = idx[0];
    } else {
        float e[3] = {node.mx[0] - node.mn[0], node.mx[1] - node.mn[1], node.mx[2] - node.mn[2]};
        node.axis = (e[0] >= e[1] && e[0] >= e[2]) ? 0 : (e[1] >= e[2] ? 1 : 2);
        uint32_t axis = node.axis;
        std::sort(idx, idx + n, [&](uint32_t a, uint32_t b) { return tree_less(faces[a], faces[b], axis); });
        size_t mid = n / 2;
        const FaceRec &p = faces[idx[mid]];
        node.pivot[0] = p.x; node.pivot[1] = p.y; node.pivot[2] = p.z;
        node.face = p.face;
        node.left = split_node(nodes, idx, mid, faces);
        node.right = split_node(nodes, idx + mid, n - mid, faces);
    }
    nodes[index] = node;
    return index;
}

void build_emitters(const VoxelSource &scene, const std::vector<Int3> &regions, const float *emission_rgb, EmitterTable &out) {
    out = EmitterTable();
    float weight[256];
    bool any = false;
    for (int m = 0; m < 256; ++m) {
        float r = emission_rgb[m * 3], g = emission_rgb[m * 3 + 1], b = emission_rgb[m * 3 + 2];
        float w = r < g ? g : r;
        weight[m] = w < b ? b : w;
        any = any || weight[m] > 0.0f;
    }
    std::vector<FaceRec> faces;
    double total = 0.0;
    if (any) {
        for (const Int3 &r : regions)
            for (int z = 0; z < 16; ++z)
                for (int y = 0; y < 16; ++y)
                    for (int x = 0; x < 16; ++x) {
                        int vx = r.x * 16 + x, vy = r.y * 16 + y, vz = r.z * 16 + z;
                        uint32_t m = scene.material_at(vx, vy, vz);
                        if (m >= AIR || weight[m] <= 0.0f) continue;
                        for (uint32_t f = 0; f < 6; ++f) {
                            uint32_t nm = scene.material_at(vx + FACE_STEP[f][0], vy + FACE_STEP[f][1], vz + FACE_STEP[f][2]);
                            if (nm != 0xffffu && nm != AIR) continue;
                            total += (double)weight[m];
                            faces.push_back(FaceRec{vx, vy, vz, f, m, (float)total, weight[m]});
                        }
                    }
    }
    std::vector<TreeNode> nodes;
    if (total != 0.0) {
        out.total_weight = (float)total;
        for (auto &f : faces) {
            f.cumulative = f.cumulative / out.total_weight;
            f.probability = f.probability / out.total_weight;
        }
        faces.back().cumulative = 1.0f;
        if (faces.size() >= 32) {
            std::vector<uint32_t> idx(faces.size());
            for (size_t i = 0; i < idx.size(); ++i) idx[i] = (uint32_t)i;
            nodes.reserve(faces.size() * 2);
            split_node(nodes, idx.data(), idx.size(), faces);
        }
    } else {
        faces.clear();
    }
    out.faces = (uint32_t)faces.size();
    out.nodes = (uint32_t)nodes.size();
    size_t records = std::max<size_t>(faces.size(), 1);
    out.words.assign(records * 8 + nodes.size() * 16, 0u);
    for (size_t i = 0; i < faces.size(); ++i) {
        const FaceRec &f = faces[i];
        uint32_t *w = &out.words[i * 8];
        w[0] = (uint32_t)f.x; w[1] = (uint32_t)f.y; w[2] = (uint32_t)f.z; w[3] = f.face;
        w[4] = fbits(f.cumulative); w[5] = fbits(f.probability); w[6] = fbits((float)f.material);
        w[7] = fbits((float)nodes.size());
    }
    // Nodes start after `faces` records (load_node offsets by extra.w * 2 uint4s).
    size_t base = faces.size() * 8;
    for (size_t i = 0; i < nodes.size(); ++i) {
        const TreeNode &n = nodes[i];
        uint32_t *w = &out.words[base + i * 16];
        w[0] = fbits(n.mn[0]); w[1] = fbits(n.mn[1]); w[2] = fbits(n.mn[2]); w[3] = fbits(n.power);
        w[4] = fbits(n.mx[0]); w[5] = fbits(n.mx[1]); w[6] = fbits(n.mx[2]); w[7] = n.normals;
        w[8] = (uint32_t)n.pivot[0]; w[9] = (uint32_t)n.pivot[1]; w[10] = (uint32_t)n.pivot[2]; w[11] = n.face;
        w[12] = n.left; w[13] = n.right; w[14] = n.axis; w[15] = 0u;
    }
    // The walk's view of each node (voxel_path.wgsl:108-131's per-node
    // terms): (mn + mx) * 0.5, len(mx - mn) * 0.5 with the dot product summed
    // x, y, z, and left.power / (left.power + right.power), each rounded as
    // binary32 exactly as the walk computes them.
    out.walk.assign(nodes.size() * 8, 0u);
    for (size_t i = 0; i < nodes.size(); ++i) {
        const TreeNode &n = nodes[i];
        uint32_t *w = &out.walk[i * 8];
        float e[3], c[3];
        for (int a = 0; a < 3; ++a) {
            c[a] = (n.mn[a] + n.mx[a]) * 0.5f;
            e[a] = n.mx[a] - n.mn[a];
        }
        float sq = e[0] * e[0] + e[1] * e[1];
        sq = sq + e[2] * e[2];
        float radius = std::sqrt(sq) * 0.5f;
        float fraction = 0.0f;
        if (n.left != NONE) fraction = nodes[n.left].power / (nodes[n.left].power + nodes[n.right].power);
        w[0] = fbits(c[0]); w[1] = fbits(c[1]); w[2] = fbits(c[2]); w[3] = fbits(radius);
        w[4] = fbits(n.power); w[5] = fbits(fraction); w[6] = n.left; w[7] = n.right | (n.normals << 26);
    }
}

// ---------------------------------------------------------------- face index
FaceSlots::FaceSlots() { entries.assign(1024, {0u, 0u, 0u, NONE}); }

long FaceSlots::find(int32_t x, int32_t y, int32_t z, uint32_t face) const {
    size_t mask = entries.size() - 1;
    size_t i = (hash_cell(x, y, z) ^ pcg(face + 1u)) & mask;
    for (;;) {
        const auto &e = entries[i];
        if (e[3] == NONE) return -1;
        if ((int32_t)e[0] == x && (int32_t)e[1] == y && (int32_t)e[2] == z && (e[3] >> 29) == face) return (long)i;
        i = (i + 1) & mask;
    }
}

static size_t place(std::vector<std::array<uint32_t, 4>> &entries, const std::array<uint32_t, 4> &e) {
    size_t mask = entries.size() - 1;
    size_t i = (hash_cell((int32_t)e[0], (int32_t)e[1], (int32_t)e[2]) ^ pcg((e[3] >> 29) + 1u)) & mask;
    while (entries[i][3] != NONE) i = (i + 1) & mask;
    entries[i] = e;
    return i;
}

void FaceSlots::insert(int32_t x, int32_t y, int32_t z, uint32_t face, uint32_t slot) {
    long f = find(x, y, z, face);
    if (f >= 0) {
        entries[f][3] = (face << 29) | slot;
        touched.push_back((uint32_t)f);
        return;
    }
    if ((live + 1) * 2 > entries.size()) {
        std::vector<std::array<uint32_t, 4>> old;
        old.swap(entries);
        entries.assign(old.size() * 2, {0u, 0u, 0u, NONE});
        for (auto &e : old)
            if (e[3] != NONE) place(entries, e);
        dirty = true;
    }
    touched.push_back((uint32_t)place(entries, {(uint32_t)x, (uint32_t)y, (uint32_t)z, (face << 29) | slot}));
    ++live;
}

// ---------------------------------------------------------------- edit retention
namespace {

// Bounds::distance_squared (geometry.rs:239-247): per-axis separation,
// saturated at 4096.
int64_t box_distance2(Int3 amin, Int3 amax, Int3 bmin, Int3 bmax) {
    int64_t d[3] = {std::max<int64_t>(std::max<int64_t>((int64_t)amin.x - bmax.x, (int64_t)bmin.x - amax.x), 0),
                    std::max<int64_t>(std::max<int64_t>((int64_t)amin.y - bmax.y, (int64_t)bmin.y - amax.y), 0),
                    std::max<int64_t>(std::max<int64_t>((int64_t)amin.z - bmax.z, (int64_t)bmin.z - amax.z), 0)};
    int64_t s = 0;
    for (int64_t v : d) {
        v = std::min<int64_t>(v, 4096);
        s += v * v;
    }
    return s;
}

int64_t floor_div64(int64_t a, int64_t b) {
    int64_t q = a / b;
    if ((a % b) != 0 && ((a < 0) != (b < 0))) --q;
    return q;
}

// geometry.rs:92-100
float emission_luminance(const float *emission_rgb, uint16_t material) {
    if (material >= AIR) return 0.0f;
    const float *e = emission_rgb + (size_t)material * 3;
    return 0.2126f * e[0] + 0.7152f * e[1] + 0.0722f * e[2];
}

bool box_less(const ChangeBox &a, const ChangeBox &b) {
    int32_t ka[6] = {a.min.x, a.min.y, a.min.z, a.max.x, a.max.y, a.max.z};
    int32_t kb[6] = {b.min.x, b.min.y, b.min.z, b.max.x, b.max.y, b.max.z};
    return std::lexicographical_compare(ka, ka + 6, kb, kb + 6);
}

bool box_equal(const ChangeBox &a, const ChangeBox &b) { return a.min == b.min && a.max == b.max; }

}  // namespace

bool PreviousScene::page_occupied(Int3 page) const {
    int x = page.x - page_min.x, y = page.y - page_min.y, z = page.z - page_min.z;
    if (x < 0 || y < 0 || z < 0 || x >= psx || y >= psy || z >= psz) return false;
    return page_solid[(size_t)x + (size_t)psx * ((size_t)y + (size_t)psy * (size_t)z)] != 0;
}

namespace {
// Whether a box lies within the influence radius of a page holding a solid
// region (from_local_bounds and from_resident_bounds select the same set).
bool within_reach(Int3 bmin, Int3 bmax, const PreviousScene &previous) {
    const int64_t r2 = (int64_t)INFLUENCE_RADIUS * INFLUENCE_RADIUS;
    int64_t lo[3] = {floor_div64((int64_t)bmin.x - INFLUENCE_RADIUS, 32), floor_div64((int64_t)bmin.y - INFLUENCE_RADIUS, 32),
                     floor_div64((int64_t)bmin.z - INFLUENCE_RADIUS, 32)};
    int64_t hi[3] = {floor_div64((int64_t)bmax.x + INFLUENCE_RADIUS, 32), floor_div64((int64_t)bmax.y + INFLUENCE_RADIUS, 32),
                     floor_div64((int64_t)bmax.z + INFLUENCE_RADIUS, 32)};
    // Clip the page range to the pages the scene holds.
    lo[0] = std::max<int64_t>(lo[0], previous.page_min.x); hi[0] = std::min<int64_t>(hi[0], previous.page_min.x + previous.psx - 1);
    lo[1] = std::max<int64_t>(lo[1], previous.page_min.y); hi[1] = std::min<int64_t>(hi[1], previous.page_min.y + previous.psy - 1);
    lo[2] = std::max<int64_t>(lo[2], previous.page_min.z); hi[2] = std::min<int64_t>(hi[2], previous.page_min.z + previous.psz - 1);
    // The page holding the box's minimum first: in a populated world it
    // usually decides at once.
    if (previous.page_occupied(Int3{(int32_t)floor_div64(bmin.x, 32), (int32_t)floor_div64(bmin.y, 32), (int32_t)floor_div64(bmin.z, 32)}))
        return true;
    for (int64_t z = lo[2]; z <= hi[2]; ++z)
        for (int64_t y = lo[1]; y <= hi[1]; ++y)
            for (int64_t x = lo[0]; x <= hi[0]; ++x) {
                Int3 page{(int32_t)x, (int32_t)y, (int32_t)z};
                Int3 pmin{page.x * 32, page.y * 32, page.z * 32}, pmax{pmin.x + 31, pmin.y + 31, pmin.z + 31};
                if (box_distance2(bmin, bmax, pmin, pmax) > r2) continue;
                if (previous.page_occupied(page)) return true;
            }
    return false;
}
}  // namespace

bool edit_event(const std::vector<RegionDelta> &deltas, const PreviousScene &previous, const float *emission_rgb,
                const std::vector<Int3> &edits, RetentionEvent &out, WorkPool *pool) {
    return scene_event(deltas, previous, emission_rgb, 0.0f, edits, out, pool);
}

bool scene_event(const std::vector<RegionDelta> &deltas, const PreviousScene &previous, const float *emission_rgb, float sky,
                 const std::vector<Int3> &edits, RetentionEvent &out, WorkPool *pool) {
    out.boxes.clear();
    // Each region's box, on the pool's threads when given: per-region slots,
    // then kept in region order.
    std::vector<ChangeBox> region_box(deltas.size());
    std::vector<uint8_t> has_box(deltas.size(), 0);
    auto scan = [&](size_t k) {
        const RegionDelta &d = deltas[k];
        Int3 origin{d.coord.x * 16, d.coord.y * 16, d.coord.z * 16};
        Int3 tmin{INT32_MAX, INT32_MAX, INT32_MAX}, tmax{INT32_MIN, INT32_MIN, INT32_MIN};
        uint32_t changed = 0;
        float flux = 0.0f;
        bool uniform_before = std::all_of(d.before.begin(), d.before.end(), [&](uint16_t v) { return v == d.before[0]; });
        bool uniform_after = std::all_of(d.after.begin(), d.after.end(), [&](uint16_t v) { return v == d.after[0]; });
        if (uniform_before && uniform_after) {
            // One material on each side: the region changes as one box of
            // 4096 voxels (geometry.rs:170-172).
            if (d.before[0] != d.after[0]) {
                tmin = origin;
                tmax = Int3{origin.x + 15, origin.y + 15, origin.z + 15};
                changed = 4096;
                flux = std::fabs(emission_luminance(emission_rgb, d.before[0]) - emission_luminance(emission_rgb, d.after[0])) * 6.0f * 4096.0f;
            }
        } else
        // Voxel by voxel in z, y, x order.
        for (int z = 0; z < 16; ++z)
            for (int y = 0; y < 16; ++y)
                for (int x = 0; x < 16; ++x) {
                    size_t i = (size_t)x + 16 * ((size_t)y + 16 * (size_t)z);
                    uint16_t a = d.before[i], b = d.after[i];
                    if (a == b) continue;
                    Int3 p{origin.x + x, origin.y + y, origin.z + z};
                    tmin = Int3{std::min(tmin.x, p.x), std::min(tmin.y, p.y), std::min(tmin.z, p.z)};
                    tmax = Int3{std::max(tmax.x, p.x), std::max(tmax.y, p.y), std::max(tmax.z, p.z)};
                    changed += 1;
                    flux += std::fabs(emission_luminance(emission_rgb, a) - emission_luminance(emission_rgb, b)) * 6.0f * 1.0f;
                }
        bool in_occupied = previous.any_solid && d.coord.x >= previous.solid_min.x && d.coord.y >= previous.solid_min.y &&
                           d.coord.z >= previous.solid_min.z && d.coord.x <= previous.solid_max.x && d.coord.y <= previous.solid_max.y &&
                           d.coord.z <= previous.solid_max.z;
        bool transport = d.solid_before || d.solid_after || in_occupied;
        // Residency changes knowledge of air as well as solids: a loaded or
        // unloaded region changes as its whole box (geometry.rs:181-189).
        bool residency = !d.loaded_before || !d.loaded_after;
        if (!transport || (!residency && changed == 0)) return;
        Int3 bmin = residency ? origin : tmin, bmax = residency ? Int3{origin.x + 15, origin.y + 15, origin.z + 15} : tmax;
        // Only changes within the influence radius of a page that held a
        // solid region can touch an existing receiver (from_local_bounds and
        // from_resident_bounds select the same set).
        if (!within_reach(bmin, bmax, previous)) return;
        ChangeBox box{bmin, bmax, flux, changed, 3u, false};
        if (!d.loaded_before) {
            // Load: the top solid row of the region (geometry.rs:192-204);
            // the flux per unit of height above a receiver is the sky's
            // over the region's width (publication.rs:438).
            box.kind = 2u;
            box.top = INT32_MIN;
            for (int y = 15; y >= 0 && box.top == INT32_MIN; --y)
                for (int z = 0; z < 16 && box.top == INT32_MIN; ++z)
                    for (int x = 0; x < 16; ++x)
                        if (d.after[(size_t)x + 16 * ((size_t)y + 16 * (size_t)z)] != AIR) {
                            box.top = origin.y + y;
                            break;
                        }
            box.flux = sky * 16.0f;
        } else if (!d.loaded_after) {
            box.kind = 1u;  // Unload: its samples become unknown, not contradicted
            box.flux = 0.0f;
        }
        region_box[k] = box;
        has_box[k] = 1;
    };
    if (pool)
        pool->run(deltas.size(), [&](size_t begin, size_t end) {
            for (size_t k = begin; k < end; ++k) scan(k);
        });
    else
        for (size_t k = 0; k < deltas.size(); ++k) scan(k);
    for (size_t k = 0; k < deltas.size(); ++k)
        if (has_box[k]) out.boxes.push_back(region_box[k]);
    if (out.boxes.empty()) return false;
    // Deterministic order and no duplicates (retention.rs:114-125). Two
    // regions never share a tight box, so the weights stay one per box.
    std::sort(out.boxes.begin(), out.boxes.end(), box_less);
    out.boxes.erase(std::unique(out.boxes.begin(), out.boxes.end(), box_equal), out.boxes.end());
    // finalize_change_flux: the GPU measures an edit while its one-voxel
    // scans fit the budget; later edits get unknown flux. The reference walks
    // a hash map here; this walks the sorted boxes, so the choice is
    // reproducible.
    uint64_t budget = FLUX_SCAN_BUDGET;
    for (ChangeBox &b : out.boxes) {
        if (b.kind != 3u) continue;  // loads and unloads keep their weight
        int64_t sx = std::max<int64_t>((int64_t)b.max.x - b.min.x + 3, 1), sy = std::max<int64_t>((int64_t)b.max.y - b.min.y + 3, 1),
                sz = std::max<int64_t>((int64_t)b.max.z - b.min.z + 3, 1);
        uint64_t positions = (uint64_t)(sx * sy * sz);
        if (positions <= budget) {
            budget -= positions;
            b.kind = 6u;
        } else {
            b.kind = 0u;
            b.flux = 0.0f;
        }
    }
    // Urgent: a box holding an edited voxel (retention.rs:151-158). A box
    // lies inside one region, so only that region's edits are tested.
    auto region_key = [](const Int3 &v) {
        return ((uint64_t)(uint32_t)(floor_div64(v.z, 16) + (1 << 20)) << 42) | ((uint64_t)(uint32_t)(floor_div64(v.y, 16) + (1 << 20)) << 21) |
               (uint64_t)(uint32_t)(floor_div64(v.x, 16) + (1 << 20));
    };
    std::vector<std::pair<uint64_t, Int3>> by_region;
    by_region.reserve(edits.size());
    for (const Int3 &v : edits) by_region.push_back({region_key(v), v});
    std::sort(by_region.begin(), by_region.end(), [](const auto &a, const auto &c) { return a.first < c.first; });
    for (ChangeBox &b : out.boxes) {
        uint64_t key = region_key(b.min);
        auto it = std::lower_bound(by_region.begin(), by_region.end(), key, [](const auto &a, uint64_t k) { return a.first < k; });
        for (; it != by_region.end() && it->first == key; ++it) {
            const Int3 &v = it->second;
            if (v.x >= b.min.x && v.y >= b.min.y && v.z >= b.min.z && v.x <= b.max.x && v.y <= b.max.y && v.z <= b.max.z) {
                b.urgent = true;
                break;
            }
        }
    }
    out.union_min = Int3{INT32_MAX, INT32_MAX, INT32_MAX};
    out.union_max = Int3{INT32_MIN, INT32_MIN, INT32_MIN};
    for (const ChangeBox &b : out.boxes) {
        out.union_min = Int3{std::min(out.union_min.x, b.min.x), std::min(out.union_min.y, b.min.y), std::min(out.union_min.z, b.min.z)};
        out.union_max = Int3{std::max(out.union_max.x, b.max.x), std::max(out.union_max.y, b.max.y), std::max(out.union_max.z, b.max.z)};
    }
    return true;
}

void encode_events(const std::vector<RetentionEvent> &events, EncodedEvents &out) {
    out.words.clear();
    out.measures.clear();
    out.events = (uint32_t)events.size();
    out.entries = 0;
    out.positions = 0;
    auto push = [](std::vector<uint32_t> &v, uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
        v.push_back(a); v.push_back(b); v.push_back(c); v.push_back(d);
    };
    uint32_t first = 0;
    for (const RetentionEvent &e : events) {
        push(out.words, (uint32_t)e.union_min.x, (uint32_t)e.union_min.y, (uint32_t)e.union_min.z, 0u);
        push(out.words, (uint32_t)e.union_max.x, (uint32_t)e.union_max.y, (uint32_t)e.union_max.z, 0u);
        uint32_t count = (uint32_t)e.boxes.size();
        push(out.words, 0u, 0u, first, count);
        first += count;
    }
    uint32_t record = out.events * 3u;
    for (const RetentionEvent &e : events)
        for (const ChangeBox &b : e.boxes) {
            // Edits and unknown changes carry the player-edit flag in bit 3;
            // loads carry their top there, saturated to 28 bits
            // (resident.rs:1319-1324, 1379-1382).
            uint32_t high = b.urgent ? 8u : 0u;
            if (b.kind == 1u) high = 0u;
            if (b.kind == 2u) high = (uint32_t)std::min(std::max(b.top, -(1 << 27)), (1 << 27) - 1) << 3;
            if (b.kind == 6u) {
                Int3 m{b.min.x - 1, b.min.y - 1, b.min.z - 1};
                uint32_t sx = (uint32_t)(b.max.x - b.min.x + 3), sy = (uint32_t)(b.max.y - b.min.y + 3), sz = (uint32_t)(b.max.z - b.min.z + 3);
                uint32_t count = sx * sy * sz;
                uint32_t flux_bits;
                std::memcpy(&flux_bits, &b.flux, 4);
                push(out.measures, record, out.positions, count, b.voxels);
                push(out.measures, (uint32_t)m.x, (uint32_t)m.y, (uint32_t)m.z, flux_bits);
                push(out.measures, sx, sy, sz, 0u);
                out.entries += 1;
                out.positions += count;
            }
            uint32_t flux_bits;
            std::memcpy(&flux_bits, &b.flux, 4);
            push(out.words, (uint32_t)b.min.x, (uint32_t)b.min.y, (uint32_t)b.min.z, flux_bits);
            push(out.words, (uint32_t)b.max.x, (uint32_t)b.max.y, (uint32_t)b.max.z, b.kind | high);
            record += 2u;
        }
}


void decode_sun_flips(const uint32_t *region, uint32_t records, std::vector<SunFlip> &out) {
    uint32_t reported = region[0];
    for (uint32_t k = 0; k < std::min(reported, records); ++k) {
        const uint32_t *r = region + 4u + 4u * k;
        Int3 v{(int32_t)r[0], (int32_t)r[1], (int32_t)r[2]};
        float flux;
        std::memcpy(&flux, &r[3], 4);
        out.push_back(SunFlip{v, v, flux});
    }
    if (reported > records) {
        const uint32_t *o = region + 4u + 4u * records;
        const int64_t bias = (int64_t)1 << 30;
        Int3 mn{(int32_t)(bias - (int64_t)o[0]), (int32_t)(bias - (int64_t)o[1]), (int32_t)(bias - (int64_t)o[2])};
        Int3 mx{(int32_t)((int64_t)o[3] - bias), (int32_t)((int64_t)o[4] - bias), (int32_t)((int64_t)o[5] - bias)};
        out.push_back(SunFlip{mn, mx, (float)o[6] / 64.0f});
    }
}

void SunChanges::add(const std::vector<SunFlip> &flips) {
    for (const SunFlip &f : flips) {
        auto key = std::make_tuple((int)floor_div64(f.min.z, 16), (int)floor_div64(f.min.y, 16), (int)floor_div64(f.min.x, 16));
        auto it = buckets.find(key);
        if (it == buckets.end()) {
            buckets.emplace(key, f);
            continue;
        }
        SunFlip &b = it->second;
        b.min = Int3{std::min(b.min.x, f.min.x), std::min(b.min.y, f.min.y), std::min(b.min.z, f.min.z)};
        b.max = Int3{std::max(b.max.x, f.max.x), std::max(b.max.y, f.max.y), std::max(b.max.z, f.max.z)};
        b.flux += f.flux;
    }
}

bool SunChanges::take(uint32_t now_ms, std::vector<SunFlip> &out) {
    out.clear();
    if (buckets.empty() || (published && now_ms - published_ms < 250u)) return false;
    published = true;
    published_ms = now_ms;
    for (auto &kv : buckets) out.push_back(kv.second);
    buckets.clear();
    return true;
}

bool light_event(const std::vector<SunFlip> &flips, const PreviousScene &scene, RetentionEvent &out) {
    out.boxes.clear();
    for (const SunFlip &f : flips)
        if (within_reach(f.min, f.max, scene)) out.boxes.push_back(ChangeBox{f.min, f.max, f.flux, 0u, 4u, false});
    if (out.boxes.empty()) return false;
    std::sort(out.boxes.begin(), out.boxes.end(), box_less);
    out.boxes.erase(std::unique(out.boxes.begin(), out.boxes.end(), box_equal), out.boxes.end());
    out.union_min = out.boxes[0].min;
    out.union_max = out.boxes[0].max;
    for (const ChangeBox &b : out.boxes) {
        out.union_min = Int3{std::min(out.union_min.x, b.min.x), std::min(out.union_min.y, b.min.y), std::min(out.union_min.z, b.min.z)};
        out.union_max = Int3{std::max(out.union_max.x, b.max.x), std::max(out.union_max.y, b.max.y), std::max(out.union_max.z, b.max.z)};
    }
    return true;
}

std::vector<Footprint> plan_footprints(std::vector<RetentionEvent> &events) {
    std::vector<Footprint> boxes;
    for (int pass = 0; pass < 2; ++pass) {
        bool edits = pass == 0;
        for (size_t e = 0; e < events.size(); ++e)
            for (const ChangeBox &b : events[e].boxes) {
                bool wanted = (b.kind == 3u || b.kind == 6u) ? edits : (b.kind == 0u ? !edits : false);
                if (wanted && boxes.size() < MAX_FOOTPRINTS) boxes.push_back(Footprint{b.min, b.max, 0});
            }
    }
    const Int3 lo{1 << 29, 1 << 29, 1 << 29}, hi{-(1 << 29), -(1 << 29), -(1 << 29)};
    for (Footprint &f : boxes) {
        RetentionEvent light;
        light.union_min = lo;
        light.union_max = hi;
        light.boxes.push_back(ChangeBox{lo, hi, 0.0f, 0u, 4u, false});
        events.push_back(light);
        f.event = events.size() - 1;
    }
    return boxes;
}

uint32_t encode_footprints(const std::vector<Footprint> &footprints, const std::vector<RetentionEvent> &events,
                           std::vector<int32_t> &words) {
    words.clear();
    std::vector<uint32_t> first_bound(events.size());
    uint32_t first = 0;
    for (size_t e = 0; e < events.size(); ++e) {
        first_bound[e] = first;
        first += (uint32_t)events[e].boxes.size();
    }
    const uint32_t records = (uint32_t)events.size() * 3u;
    auto push = [&](int32_t a, int32_t b, int32_t c, int32_t d) {
        words.push_back(a); words.push_back(b); words.push_back(c); words.push_back(d);
    };
    push((int32_t)footprints.size(), 0, 0, 0);
    uint32_t threads = 0;
    for (const Footprint &f : footprints) {
        float dx = (float)(f.max.x - f.min.x + 1), dy = (float)(f.max.y - f.min.y + 1), dz = (float)(f.max.z - f.min.z + 1);
        float length = std::sqrt(dx * dx + dy * dy + dz * dz);
        uint32_t grid = (uint32_t)std::min(std::max(std::ceil(length), 1.0f), 64.0f);
        push(f.min.x, f.min.y, f.min.z, (int32_t)(f.event * 3u));
        push(f.max.x, f.max.y, f.max.z, (int32_t)(records + first_bound[f.event] * 2u));
        push((int32_t)threads, (int32_t)grid, 0, 0);
        threads += grid * grid;
    }
    return threads;
}
}  // namespace vgi_host
