// Host-side scene bookkeeping (see vgi_host_scene.h).
#include "vgi_host_scene.h"

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

static void place(std::vector<std::array<uint32_t, 4>> &entries, const std::array<uint32_t, 4> &e) {
    size_t mask = entries.size() - 1;
    size_t i = (hash_cell((int32_t)e[0], (int32_t)e[1], (int32_t)e[2]) ^ pcg((e[3] >> 29) + 1u)) & mask;
    while (entries[i][3] != NONE) i = (i + 1) & mask;
    entries[i] = e;
}

void FaceSlots::insert(int32_t x, int32_t y, int32_t z, uint32_t face, uint32_t slot) {
    long f = find(x, y, z, face);
    dirty = true;
    if (f >= 0) {
        entries[f][3] = (face << 29) | slot;
        return;
    }
    if ((live + 1) * 2 > entries.size()) {
        std::vector<std::array<uint32_t, 4>> old;
        old.swap(entries);
        entries.assign(old.size() * 2, {0u, 0u, 0u, NONE});
        for (auto &e : old)
            if (e[3] != NONE) place(entries, e);
    }
    place(entries, {(uint32_t)x, (uint32_t)y, (uint32_t)z, (face << 29) | slot});
    ++live;
}

}  // namespace vgi_host
