// Seeded synthetic scenes for tests and benchmarks (not part of the
// reference seam). All randomness is the PCG hash (sampling.rs:5-9) of a
// seed and a counter, so every scene is reproducible from (workload, seed).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "vgi_workloads.h"

namespace {

constexpr uint16_t AIR = 256;

inline uint32_t pcg(uint32_t value) {
    uint32_t state = value * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
// A counter-based stream: the i-th draw of stream `seed`.
inline uint32_t draw(uint32_t seed, uint32_t i) { return pcg(seed ^ pcg(i + 0x9e3779b9u)); }
inline float unit(uint32_t seed, uint32_t i) { return (float)(draw(seed, i) >> 8) * (1.0f / 16777216.0f); }
inline size_t at(int x, int y, int z, int sx, int sy) { return (size_t)x + (size_t)sx * ((size_t)y + (size_t)sy * (size_t)z); }

// Value noise: PCG lattice values in [0, 1), smoothstep-trilinear blend.
inline float lattice(uint32_t seed, int x, int y, int z) {
    uint32_t h = pcg(pcg(pcg(pcg(seed) ^ (uint32_t)x) ^ (uint32_t)y) ^ (uint32_t)z);
    return (float)(h >> 8) * (1.0f / 16777216.0f);
}
inline float smooth(float t) { return t * t * (3.0f - 2.0f * t); }
float value_noise(uint32_t seed, float x, float y, float z) {
    int ix = (int)std::floor(x), iy = (int)std::floor(y), iz = (int)std::floor(z);
    float fx = smooth(x - ix), fy = smooth(y - iy), fz = smooth(z - iz);
    float c[2][2][2];
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) c[dz][dy][dx] = lattice(seed, ix + dx, iy + dy, iz + dz);
    float x00 = c[0][0][0] + (c[0][0][1] - c[0][0][0]) * fx, x01 = c[0][1][0] + (c[0][1][1] - c[0][1][0]) * fx;
    float x10 = c[1][0][0] + (c[1][0][1] - c[1][0][0]) * fx, x11 = c[1][1][0] + (c[1][1][1] - c[1][1][0]) * fx;
    float y0 = x00 + (x01 - x00) * fy, y1 = x10 + (x11 - x10) * fy;
    return y0 + (y1 - y0) * fz;
}
// Four octaves, base period 64, gain 0.5, lacunarity 2, normalised to [0, 1).
float fbm(uint32_t seed, int x, int y, int z, float period) {
    float sum = 0.0f, amp = 1.0f, norm = 0.0f, freq = 1.0f / period;
    for (int o = 0; o < 4; ++o) {
        sum += amp * value_noise(seed + (uint32_t)o * 7919u, x * freq, y * freq, z * freq);
        norm += amp;
        amp *= 0.5f;
        freq *= 2.0f;
    }
    return sum / norm;
}

template <class F>
void parallel_z(int nz, F fn) {
    unsigned n = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned w = 0; w < n; ++w)
        pool.emplace_back([&, w]() {
            for (int z = (int)w; z < nz; z += (int)n) fn(z);
        });
    for (auto &t : pool) t.join();
}

// Solid where the noise exceeds the threshold that leaves `solid` of a
// subsample of voxels solid (the bisected quantile).
void noise_fill(uint16_t *v, int n, uint32_t seed, float solid_fraction, std::vector<float> *values_out) {
    std::vector<float> values((size_t)n * n * n);
    parallel_z(n, [&](int z) {
        for (int y = 0; y < n; ++y)
            for (int x = 0; x < n; ++x) values[at(x, y, z, n, n)] = fbm(seed, x, y, z, 64.0f);
    });
    std::vector<float> sample;
    size_t stride = values.size() > (1u << 22) ? values.size() / (1u << 22) : 1;
    for (size_t i = 0; i < values.size(); i += stride) sample.push_back(values[i]);
    size_t k = (size_t)((1.0 - solid_fraction) * (double)(sample.size() - 1));
    std::nth_element(sample.begin(), sample.begin() + k, sample.end());
    float threshold = sample[k];
    for (size_t i = 0; i < values.size(); ++i) v[i] = values[i] > threshold ? 1 : AIR;
    if (values_out) values_out->swap(values);
}

// Materials 1..4 by 8^3 cell hash.
void assign_materials(uint16_t *v, int n, uint32_t seed) {
    parallel_z(n, [&](int z) {
        for (int y = 0; y < n; ++y)
            for (int x = 0; x < n; ++x) {
                size_t i = at(x, y, z, n, n);
                if (v[i] == AIR) continue;
                v[i] = (uint16_t)(1 + pcg(pcg(pcg(seed ^ 0x51u) ^ (uint32_t)(x >> 3)) ^ (uint32_t)(y >> 3) * 131u ^ (uint32_t)(z >> 3) * 7u) % 4);
            }
    });
}

bool exposed(const uint16_t *v, int n, int x, int y, int z) {
    static const int d[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    for (auto &o : d) {
        int a = x + o[0], b = y + o[1], c = z + o[2];
        if (a < 0 || b < 0 || c < 0 || a >= n || b >= n || c >= n) continue;
        if (v[at(a, b, c, n, n)] == AIR) return true;
    }
    return false;
}

// `count` emissive voxels on exposed solid voxels, drawn by rejection from the
// PCG stream. Emitter i takes material 5 + (i mod 128) (materials are u8).
constexpr int EMITTER_MATERIALS = 128;
void place_emitters(uint16_t *v, int n, uint32_t seed, int count, int y_max) {
    int placed = 0;
    for (uint32_t i = 0; placed < count && i < 100000000u; ++i) {
        int x = (int)(draw(seed ^ 0xe1u, i * 3) % (uint32_t)n), y = (int)(draw(seed ^ 0xe1u, i * 3 + 1) % (uint32_t)std::min(n, y_max)),
            z = (int)(draw(seed ^ 0xe1u, i * 3 + 2) % (uint32_t)n);
        size_t k = at(x, y, z, n, n);
        if (v[k] == AIR || v[k] >= 5) continue;
        if (!exposed(v, n, x, y, z)) continue;
        v[k] = (uint16_t)(5 + placed % EMITTER_MATERIALS);
        ++placed;
    }
}

void carve_sphere(uint16_t *v, int n, float cx, float cy, float cz, float r) {
    int x0 = std::max(0, (int)std::floor(cx - r)), x1 = std::min(n - 1, (int)std::ceil(cx + r));
    int y0 = std::max(0, (int)std::floor(cy - r)), y1 = std::min(n - 1, (int)std::ceil(cy + r));
    int z0 = std::max(0, (int)std::floor(cz - r)), z1 = std::min(n - 1, (int)std::ceil(cz + r));
    for (int z = z0; z <= z1; ++z)
        for (int y = y0; y <= y1; ++y)
            for (int x = x0; x <= x1; ++x) {
                float dx = x + 0.5f - cx, dy = y + 0.5f - cy, dz = z + 0.5f - cz;
                if (dx * dx + dy * dy + dz * dz <= r * r) v[at(x, y, z, n, n)] = AIR;
            }
}

int default_edge(int workload) {
    switch (workload) {
    case 1: return 64;
    case 2: return 256;
    case 3: return 64;
    case 4: return 512;
    case 5: return 512;
    case 6: return 32;
    default: return 0;
    }
}

}  // namespace

extern "C" {

int vgi_scene_generate(int workload, uint32_t seed, uint32_t edge, uint16_t *voxels, int32_t dims[3]) {
    int n = edge ? (int)edge : default_edge(workload);
    if (n <= 0) return -1;
    int padded = ((n + 31) / 32) * 32;
    if (dims) { dims[0] = dims[1] = dims[2] = padded; }
    if (!voxels) return 0;
    size_t total = (size_t)padded * padded * padded;
    std::fill_n(voxels, total, AIR);
    std::vector<uint16_t> v((size_t)n * n * n, AIR);
    switch (workload) {
    case 1: {  // hollow stone box, 4x4 window in the x = 0 wall, one emitter on the floor
        for (int z = 0; z < n; ++z)
            for (int y = 0; y < n; ++y)
                for (int x = 0; x < n; ++x) {
                    bool shell = x == 0 || y == 0 || z == 0 || x == n - 1 || y == n - 1 || z == n - 1;
                    bool window = x == 0 && y >= n / 2 - 2 && y < n / 2 + 2 && z >= n / 2 - 2 && z < n / 2 + 2;
                    if (shell && !window) v[at(x, y, z, n, n)] = 1;
                }
        v[at(n / 2, 1, n / 2, n, n)] = 2;
        break;
    }
    case 2:
    case 3: {  // noise terrain, ~35% solid, 4 albedos, 64 emitters (3: a batch member, 8 emitters)
        noise_fill(v.data(), n, seed, 0.35f, nullptr);
        assign_materials(v.data(), n, seed);
        place_emitters(v.data(), n, seed, workload == 2 ? 64 : 8, n);
        break;
    }
    case 4: {  // edit-stream terrain, ~30% solid, 16 emitters
        noise_fill(v.data(), n, seed, 0.30f, nullptr);
        assign_materials(v.data(), n, seed);
        place_emitters(v.data(), n, seed, 16, n);
        break;
    }
    case 5: {  // worm cave: rock below a sky band, 200 worm tunnels, 3 shafts, 300 emitters
        int ground = n - n / 16;
        noise_fill(v.data(), n, seed, 0.72f, nullptr);
        for (int z = 0; z < n; ++z)
            for (int y = ground; y < n; ++y)
                for (int x = 0; x < n; ++x) v[at(x, y, z, n, n)] = AIR;
        uint32_t s = seed ^ 0x3f3fu;
        uint32_t k = 0;
        for (int w = 0; w < 200; ++w) {
            float px = unit(s, k++) * n, py = unit(s, k++) * (ground - 16) + 8, pz = unit(s, k++) * n;
            float r = 3.0f + unit(s, k++) * 5.0f;
            float yaw = unit(s, k++) * 6.2831853f, pitch = (unit(s, k++) - 0.5f) * 0.6f;
            int length = 40 + (int)(unit(s, k++) * 120);
            for (int step = 0; step < length; ++step) {
                carve_sphere(v.data(), n, px, py, pz, r);
                yaw += (unit(s, k++) - 0.5f) * 0.5f;
                pitch = std::max(-0.5f, std::min(0.5f, pitch + (unit(s, k++) - 0.5f) * 0.2f));
                float c = std::cos(pitch);
                px += std::cos(yaw) * c * r * 0.6f;
                py += std::sin(pitch) * r * 0.6f;
                pz += std::sin(yaw) * c * r * 0.6f;
                py = std::max(r + 1.0f, std::min((float)ground - r - 1.0f, py));
                if (px < 0 || pz < 0 || px >= n || pz >= n) break;
            }
        }
        for (int o = 0; o < 3; ++o) {  // vertical shafts from the surface
            float cx = n * (0.25f + 0.25f * o), cz = n * (0.3f + 0.2f * (o % 2));
            for (int y = ground; y >= n / 4; y -= 3) carve_sphere(v.data(), n, cx, (float)y, cz, 6.0f);
        }
        assign_materials(v.data(), n, seed);
        place_emitters(v.data(), n, seed, 300, ground - 1);
        break;
    }
    case 6: {  // research.rs:71-89 furnace: shell 8..23 of material 4
        for (int z = 8; z < 24; ++z)
            for (int y = 8; y < 24; ++y)
                for (int x = 8; x < 24; ++x)
                    if (x == 8 || x == 23 || y == 8 || y == 23 || z == 8 || z == 23) v[at(x, y, z, n, n)] = 4;
        break;
    }
    default:
        return -1;
    }
    for (int z = 0; z < n; ++z)
        for (int y = 0; y < n; ++y) std::memcpy(&voxels[at(0, y, z, padded, padded)], &v[at(0, y, z, n, n)], (size_t)n * 2);
    return 0;
}

int vgi_scene_optics(int workload, uint32_t seed, float *diffuse, float *emission, vgi_light *lights, uint32_t *light_count,
                     float environment[3]) {
    if (!diffuse || !emission || !lights || !light_count || !environment) return -1;
    for (int m = 0; m < 256; ++m)
        for (int c = 0; c < 3; ++c) { diffuse[m * 3 + c] = 0.45f; emission[m * 3 + c] = 0.0f; }
    environment[0] = environment[1] = environment[2] = 0.0f;
    *light_count = 0;
    // The sun travels along (0.3, -0.8, 0.5); GpuLight stores the unit
    // direction toward it (sources.rs:130, gpu.rs:733) with w = -1.
    auto sun = [&](float radiance) {
        float d[3] = {-0.3f, 0.8f, -0.5f};
        float inv = 1.0f / std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        vgi_light l;
        std::memset(&l, 0, sizeof(l));
        l.position_range[0] = d[0] * inv; l.position_range[1] = d[1] * inv; l.position_range[2] = d[2] * inv;
        l.position_range[3] = -1.0f;
        l.energy[0] = l.energy[1] = l.energy[2] = radiance;
        lights[(*light_count)++] = l;
    };
    switch (workload) {
    case 1:
        for (int c = 0; c < 3; ++c) { diffuse[3 + c] = 0.5f; diffuse[6 + c] = 0.5f; }
        emission[6] = 8.0f; emission[7] = 4.0f; emission[8] = 1.0f;
        sun(3.0f);
        break;
    case 2:
    case 3:
    case 4:
    case 5: {
        int emitters = workload == 2 ? 64 : workload == 3 ? 8 : workload == 4 ? 16 : 300;
        emitters = std::min(emitters, EMITTER_MATERIALS);
        for (int m = 1; m <= 4; ++m)
            for (int c = 0; c < 3; ++c) diffuse[m * 3 + c] = 0.2f + 0.6f * unit(seed ^ 0xa1u, (uint32_t)(m * 3 + c));
        for (int e = 0; e < emitters; ++e) {
            float r = 2.0f + 8.0f * unit(seed ^ 0xb2u, (uint32_t)e);
            for (int c = 0; c < 3; ++c) { diffuse[(5 + e) * 3 + c] = 0.5f; emission[(5 + e) * 3 + c] = r; }
        }
        sun(3.0f);
        break;
    }
    case 6:
        diffuse[12] = 0.3f; diffuse[13] = 0.55f; diffuse[14] = 0.7f;
        emission[12] = 0.7f; emission[13] = 1.3f; emission[14] = 0.2f;
        break;
    default:
        return -1;
    }
    return 0;
}

int64_t vgi_scene_faces(const uint16_t *v, const int32_t dims[3], int interior_only, vgi_face_task *out, int64_t capacity) {
    static const int d[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    int sx = dims[0], sy = dims[1], sz = dims[2];
    int64_t count = 0;
    for (int z = 0; z < sz; ++z)
        for (int y = 0; y < sy; ++y)
            for (int x = 0; x < sx; ++x) {
                if (v[at(x, y, z, sx, sy)] == AIR) continue;
                for (uint32_t f = 0; f < 6; ++f) {
                    int a = x + d[f][0], b = y + d[f][1], c = z + d[f][2];
                    if (a < 0 || b < 0 || c < 0 || a >= sx || b >= sy || c >= sz) continue;
                    if (v[at(a, b, c, sx, sy)] != AIR) continue;
                    if (interior_only && (a < 1 || b < 1 || c < 1 || a > sx - 2 || b > sy - 2 || c > sz - 2)) continue;
                    if (out && count < capacity) {
                        vgi_face_task &t = out[count];
                        t.voxel[0] = x; t.voxel[1] = y; t.voxel[2] = z;
                        t.face = f;
                        t.sequence = 0;
                        t.padding[0] = t.padding[1] = t.padding[2] = 0;
                    }
                    ++count;
                }
            }
    return count;
}

int64_t vgi_scene_edits(uint32_t seed, uint32_t frame, uint32_t count, uint32_t edge, int32_t *xyz, uint16_t *values,
                        int64_t capacity) {
    int64_t n = 0;
    auto emit = [&](int x, int y, int z, uint16_t v) {
        if (xyz && n < capacity) {
            xyz[n * 3] = x; xyz[n * 3 + 1] = y; xyz[n * 3 + 2] = z;
            values[n] = v;
        }
        ++n;
    };
    uint32_t stream = pcg(seed ^ pcg(frame + 0x7f4a7c15u));
    for (uint32_t i = 0; i < count; ++i) {
        uint32_t kind = draw(stream, i * 6u) % 3u;
        int x = (int)(draw(stream, i * 6u + 1u) % edge), y = (int)(draw(stream, i * 6u + 2u) % edge),
            z = (int)(draw(stream, i * 6u + 3u) % edge);
        if (kind == 0u) {
            emit(x, y, z, (uint16_t)(1u + draw(stream, i * 6u + 4u) % 4u));
        } else if (kind == 1u) {
            emit(x, y, z, AIR);
        } else {
            float r = 1.0f + 4.0f * unit(stream, i * 6u + 5u);
            float cx = (float)x + 0.5f, cy = (float)y + 0.5f, cz = (float)z + 0.5f;
            int lo[3] = {(int)std::floor(cx - r), (int)std::floor(cy - r), (int)std::floor(cz - r)};
            int hi[3] = {(int)std::ceil(cx + r), (int)std::ceil(cy + r), (int)std::ceil(cz + r)};
            for (int c = 0; c < 3; ++c) { lo[c] = std::max(lo[c], 0); hi[c] = std::min(hi[c], (int)edge - 1); }
            for (int vz = lo[2]; vz <= hi[2]; ++vz)
                for (int vy = lo[1]; vy <= hi[1]; ++vy)
                    for (int vx = lo[0]; vx <= hi[0]; ++vx) {
                        float dx = vx + 0.5f - cx, dy = vy + 0.5f - cy, dz = vz + 0.5f - cz;
                        if (dx * dx + dy * dy + dz * dz <= r * r) emit(vx, vy, vz, AIR);
                    }
        }
    }
    return n;
}

}  // extern "C"

// ---------------------------------------------------------------- prepass
namespace {
struct Mat4 {
    double m[4][4];  // row-major
};
Mat4 mul(const Mat4 &a, const Mat4 &b) {
    Mat4 r{};
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            for (int k = 0; k < 4; ++k) r.m[i][j] += a.m[i][k] * b.m[k][j];
    return r;
}
// Gauss-Jordan inverse with partial pivoting (the matrices here are well
// conditioned perspective-times-rigid transforms).
Mat4 inverse(Mat4 a) {
    Mat4 r{};
    for (int i = 0; i < 4; ++i) r.m[i][i] = 1.0;
    for (int c = 0; c < 4; ++c) {
        int p = c;
        for (int i = c + 1; i < 4; ++i)
            if (std::fabs(a.m[i][c]) > std::fabs(a.m[p][c])) p = i;
        std::swap(a.m[p], a.m[c]);
        std::swap(r.m[p], r.m[c]);
        double d = a.m[c][c];
        for (int j = 0; j < 4; ++j) { a.m[c][j] /= d; r.m[c][j] /= d; }
        for (int i = 0; i < 4; ++i) {
            if (i == c) continue;
            double f = a.m[i][c];
            for (int j = 0; j < 4; ++j) { a.m[i][j] -= f * a.m[c][j]; r.m[i][j] -= f * r.m[c][j]; }
        }
    }
    return r;
}
}  // namespace

int64_t vgi_scene_prepass(const uint16_t *voxels, const int32_t dims[3], const float eye_f[3], const float target_f[3], float fov_y,
                          uint32_t width, uint32_t height, float *world_from_clip, float *depth, float *normals) {
    if (!voxels || !dims || !eye_f || !target_f || !world_from_clip || !depth || !normals || width == 0 || height == 0) return -1;
    const int sx = dims[0], sy = dims[1], sz = dims[2];
    double eye[3] = {eye_f[0], eye_f[1], eye_f[2]};
    double f[3] = {target_f[0] - eye[0], target_f[1] - eye[1], target_f[2] - eye[2]};
    double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    if (!(fl > 0.0)) return -1;
    for (double &v : f) v /= fl;
    double r[3] = {-f[2], 0.0, f[0]};  // f x (0, 1, 0)
    double rl = std::sqrt(r[0] * r[0] + r[2] * r[2]);
    if (!(rl > 1e-9)) return -1;
    r[0] /= rl; r[2] /= rl;
    double u[3] = {r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0]};
    Mat4 view{};
    for (int j = 0; j < 3; ++j) { view.m[0][j] = r[j]; view.m[1][j] = u[j]; view.m[2][j] = -f[j]; }
    for (int i = 0; i < 3; ++i) view.m[i][3] = -(view.m[i][0] * eye[0] + view.m[i][1] * eye[1] + view.m[i][2] * eye[2]);
    view.m[3][3] = 1.0;
    const double near = 0.1, t = 1.0 / std::tan(fov_y * 0.5), aspect = (double)width / (double)height;
    Mat4 proj{};
    proj.m[0][0] = t / aspect; proj.m[1][1] = t; proj.m[2][3] = near; proj.m[3][2] = -1.0;
    const Mat4 cfw = mul(proj, view), wfc = inverse(cfw);
    for (int c = 0; c < 4; ++c)
        for (int rr = 0; rr < 4; ++rr) world_from_clip[c * 4 + rr] = (float)wfc.m[rr][c];
    auto solid = [&](int x, int y, int z) { return voxels[at(x, y, z, sx, sy)] != AIR; };
    int64_t hits = 0;
    for (uint32_t py = 0; py < height; ++py) {
        for (uint32_t px = 0; px < width; ++px) {
            size_t i = (size_t)py * width + px;
            depth[i] = 0.0f;
            normals[i * 3] = normals[i * 3 + 1] = normals[i * 3 + 2] = 0.0f;
            double clip[4] = {(px + 0.5) / width * 2.0 - 1.0, 1.0 - (py + 0.5) / height * 2.0, 1.0, 1.0}, w[4] = {0, 0, 0, 0};
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b) w[a] += wfc.m[a][b] * clip[b];
            double d[3] = {w[0] / w[3] - eye[0], w[1] / w[3] - eye[1], w[2] / w[3] - eye[2]};
            double dl = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            for (double &v : d) v /= dl;
            // Clip the ray to the grid box, then walk voxels (Amanatides & Woo 1987).
            double t0 = 0.0, t1 = 1e30;
            const int size[3] = {sx, sy, sz};
            bool miss = false;
            for (int a = 0; a < 3; ++a) {
                if (std::fabs(d[a]) < 1e-12) {
                    if (eye[a] < 0.0 || eye[a] >= size[a]) miss = true;
                    continue;
                }
                double ta = (0.0 - eye[a]) / d[a], tb = (size[a] - eye[a]) / d[a];
                t0 = std::max(t0, std::min(ta, tb));
                t1 = std::min(t1, std::max(ta, tb));
            }
            if (miss || t0 >= t1) continue;
            double p[3];
            int v[3], step[3], axis = -1;
            double next[3], delta[3];
            for (int a = 0; a < 3; ++a) {
                p[a] = eye[a] + d[a] * t0;
                v[a] = std::min(std::max((int)std::floor(p[a]), 0), size[a] - 1);
                step[a] = d[a] > 0.0 ? 1 : -1;
                delta[a] = std::fabs(d[a]) < 1e-12 ? 1e30 : std::fabs(1.0 / d[a]);
                double boundary = d[a] > 0.0 ? v[a] + 1.0 : (double)v[a];
                next[a] = std::fabs(d[a]) < 1e-12 ? 1e30 : t0 + (boundary - p[a]) / d[a];
            }
            if (t0 > 0.0) {  // entered through the box face with the largest entry distance
                double best = -1.0;
                for (int a = 0; a < 3; ++a) {
                    if (std::fabs(d[a]) < 1e-12) continue;
                    double ta = ((d[a] > 0.0 ? 0.0 : size[a]) - eye[a]) / d[a];
                    if (ta > best) { best = ta; axis = a; }
                }
            }
            double tc = t0;
            while (true) {
                if (solid(v[0], v[1], v[2])) {
                    if (axis < 0) break;  // the camera is inside solid: no surface
                    double hit[3] = {eye[0] + d[0] * tc, eye[1] + d[1] * tc, eye[2] + d[2] * tc};
                    double c2 = 0.0, c3 = 0.0;
                    for (int b = 0; b < 3; ++b) { c2 += cfw.m[2][b] * hit[b]; c3 += cfw.m[3][b] * hit[b]; }
                    c2 += cfw.m[2][3];
                    c3 += cfw.m[3][3];
                    depth[i] = (float)(c2 / c3);
                    normals[i * 3 + axis] = step[axis] > 0 ? -1.0f : 1.0f;
                    ++hits;
                    break;
                }
                int a = next[0] < next[1] ? (next[0] < next[2] ? 0 : 2) : (next[1] < next[2] ? 1 : 2);
                tc = next[a];
                if (tc > t1) break;
                v[a] += step[a];
                if (v[a] < 0 || v[a] >= size[a]) break;
                next[a] += delta[a];
                axis = a;
            }
        }
    }
    return hits;
}

