// The C ABI of include/vgi.h: scene upload and edits (GPU NAADF build),
// optics, batches, the resident world cache and device-resident frames.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "vgi.h"
#include "vgi_host_scene.h"
#include "vgi_pool.h"
#include "vgi_kernels.h"
#include "vgi_mlp.h"
#include "vgi_wavefront.h"
#include "vgi_trace.h"
#include "vgi_view.h"

#include <cmath>
#include <numeric>
#include <random>

using vgi_host::Int3;

namespace {

thread_local std::string g_error;

// The device scratch of vgi_trace_faces_noalloc: face tasks, then records,
// for at most VGI_NOALLOC_MAX_FACES faces (128 bytes each: 32 MiB). Each GPU
// the library runs on holds its own copy until the process ends; host code
// uses only the address resolved per device (scratch_of), never the name.
constexpr size_t NOALLOC_TASKS = 0, NOALLOC_RECORDS = (size_t)VGI_NOALLOC_MAX_FACES * 32;
constexpr size_t NOALLOC_BYTES = (size_t)VGI_NOALLOC_MAX_FACES * 128;
static __device__ __align__(256) uint8_t scratch[NOALLOC_BYTES];
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
void *g_scratch[MAX_DEVICES] = {};

int fail(const std::string &message) {
    g_error = message;
    return -1;
}

#define HIP_OK(expr)                                                                          \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess) return fail(std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

// A device allocation that grows to a power of two (gpu.rs:254-276).
struct DeviceBuffer {
    void *ptr = nullptr;
    size_t bytes = 0;
    hipError_t reserve(size_t need) {
        if (need <= bytes && ptr) return hipSuccess;
        size_t cap = 256;
        while (cap < need) cap <<= 1;
        if (ptr) {
            hipError_t e = hipFree(ptr);
            if (e != hipSuccess) return e;
            ptr = nullptr;
        }
        bytes = 0;
        hipError_t e = hipMalloc(&ptr, cap);
        if (e == hipSuccess) bytes = cap;
        return e;
    }
    // Grow to `need` bytes keeping the first `keep` bytes: the copy follows
    // the stream's earlier work, and the old allocation is freed only once
    // the stream has passed it.
    hipError_t grow_keep(size_t need, size_t keep, hipStream_t stream) {
        if (need <= bytes && ptr) return hipSuccess;
        size_t cap = 256;
        while (cap < need) cap <<= 1;
        void *fresh = nullptr;
        hipError_t e = hipMalloc(&fresh, cap);
        if (e != hipSuccess) return e;
        keep = std::min(keep, bytes);
        if (ptr && keep) e = hipMemcpyAsync(fresh, ptr, keep, hipMemcpyDeviceToDevice, stream);
        if (e == hipSuccess && ptr) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) {
            (void)hipFree(fresh);
            return e;
        }
        if (ptr) e = hipFree(ptr);
        ptr = fresh;
        bytes = cap;
        return e;
    }
    hipError_t release() {
        hipError_t e = hipSuccess;
        if (ptr) e = hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
        return e;
    }
    template <class T>
    T *as() const {
        return reinterpret_cast<T *>(ptr);
    }
};

inline size_t lin3(int x, int y, int z, int sx, int sy) { return (size_t)x + (size_t)sx * ((size_t)y + (size_t)sy * (size_t)z); }
inline int floor_div(int a, int b) { int q = a / b; if ((a % b) != 0 && ((a < 0) != (b < 0))) --q; return q; }
inline int floor_mod(int a, int b) { int m = a % b; if (m < 0) m += b; return m; }
inline uint64_t page_key(int x, int y, int z) {
    return ((uint64_t)(uint32_t)(x + (1 << 20)) << 42) ^ ((uint64_t)(uint32_t)(y + (1 << 20)) << 21) ^ (uint64_t)(uint32_t)(z + (1 << 20));
}

struct Ticket {
    DeviceBuffer results, hits, counters;
    uint32_t count = 0, samples = 0;
    bool with_hits = false, with_counters = false;
    hipEvent_t done = nullptr;
};

}  // namespace

struct vgi_tracer {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[10] = {};
    float timings[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // GPU ms: prepare, trace, accumulate, build; host ms of the last edit: regions + GPU re-pack, top level, event, exposure
    // ---- scene (host authority for page contents, as the reference's RegionMap)
    std::unordered_map<uint64_t, Int3> pages;
    Int3 vmin{0, 0, 0};           // voxel origin of the dense box
    int gx = 0, gy = 0, gz = 0;   // dense box size in voxels
    std::vector<uint16_t> grid;   // host copy, x fastest
    Int3 rmin{0, 0, 0};           // region box (loaded bounds when non-empty)
    int rsx = 0, rsy = 0, rsz = 0;
    std::vector<uint8_t> loaded;  // per region of the box
    std::vector<uint8_t> solid;
    std::vector<uint64_t> masks;  // 4 words per region: materials present
    bool have_regions = false;
    Int3 bmin{0, 0, 0}, bmax{0, 0, 0};  // loaded region bounds
    vgi_host::TopLevelField top;
    vgi_host::EmitterTable emitters;
    bool emitters_dirty = true;
    // ---- device scene
    DeviceBuffer d_grid, d_headers, d_cells, d_info, d_list, d_top, d_totals, d_scatter_idx, d_scatter_val;
    uint32_t cells_used = 0;
    uint32_t cells_live = 0;
    // ---- optics
    float diffuse[768], emission[768];
    std::vector<float> lights;  // 12 floats per light
    uint32_t light_count = 0, directional = 0;
    DeviceBuffer d_materials, d_lights, d_emitters;
    vgi_transport transport{};
    // ---- resident cache
    bool resident = false;
    vgi_host::FaceSlots face_slots;
    std::vector<uint32_t> free_slots;
    uint32_t high_water = 0;
    DeviceBuffer d_index, d_state, d_checkpoint, d_meta;
    size_t slot_capacity = 0;
    // GPU work selection: the batch counter (28 bits, never 0 once a batch
    // ran), the scheduling settings record_stability and selection read, the
    // per-slot selection keys (persisting across batches) and the pick batch.
    uint32_t batch_id = 0;
    vgi_schedule schedule{};
    bool footprints = true;  // change footprints traced with each edit's events
    DeviceBuffer d_keys, d_selection, d_pick_tasks, d_pick_slots, d_pick_results, d_mirror, d_select_ranges;
    DeviceBuffer d_footprint_tasks, d_footprint_sums;
    // Sun corner words: one per resident slot (0 until traced), the last
    // call's flip region, and the flips binned until they publish.
    DeviceBuffer d_corners, d_sun_flux, d_sun_queue, d_sun_ranges, d_sun_region;
    std::vector<uint32_t> sun_region;
    vgi_host::SunChanges sun_changes;
    hipEvent_t sun_ev[2] = {};
    float sun_ms[2] = {0.0f, 0.0f};  // GPU passes, host binning and event
    uint32_t picks = 0, mirror_budget = 0;
    hipEvent_t sel_ev[4] = {};
    bool select_recorded = false;
    std::vector<uint32_t> last_slots;
    std::chrono::steady_clock::time_point origin = std::chrono::steady_clock::now();
    int64_t fixed_clock = -1;  // resident.rs:709-710 test clock; < 0: wall clock
    // ---- batches
    DeviceBuffer d_tasks, d_slots;
    std::map<uint64_t, std::unique_ptr<Ticket>> tickets;
    uint64_t next_ticket = 1;
    // ---- bound frames
    DeviceBuffer b_tasks, b_slots, b_results, b_counters, b_hits;
    uint32_t bound = 0;
    uint64_t bound_hit_words = 0;  // first-hit words of the last frame run with VGI_RECORD_HITS (0: none)
    // Changes of the bound set since the last frame (edits with bound faces
    // that follow the exposed surface): the last frame's records stay in the
    // rows of its own bound order, and reads map them to the faces bound
    // now. Each entry (index, source): the face at `source` took the place
    // of `index` and the set lost its last place; (NONE, 0): a face joined
    // at the end (no record of the last frame).
    std::vector<std::pair<uint32_t, uint32_t>> bound_moves;
    uint32_t frame_bound = 0;        // bound faces when the last frame ran
    bool bound_broken = false;       // an exposure update failed part way: rebind before the next frame
    float last_trace_ms = 0.0f;
    uint32_t frame_index = 0;
    bool frame_recorded = false;
    float last_frame_ms = 0.0f;
    // A bound face's direct terms (results words 16-23) depend only on the
    // geometry, the lights and the reach: every call that may change one of
    // them, or the bound set, advances direct_epoch, and a frame whose trace
    // wrote the terms records the epoch it wrote them at. While the two
    // agree, frames keep the terms instead of tracing them again.
    uint64_t direct_epoch = 1, direct_written = 0;
    // Pinned staging for the edit path's small uploads (an upload from
    // pinned memory returns at once; one from pageable memory waits for the
    // stream's earlier work). A bump region, started over only once the
    // stream has passed every copy issued from it (stage_done follows the
    // latest copy).
    uint8_t *stage = nullptr;
    size_t stage_bytes = 0, stage_used = 0;
    hipEvent_t stage_done = nullptr;
    bool stage_pending = false;
    uint32_t stage_wraps = 0, stage_grows = 0;  // times the region started over / was allocated larger
    std::vector<uint32_t> bound_material;  // material of each bound face's voxel
    // ---- learned radiance cache
    bool mlp_ready = false, learned = false;
    uint32_t mlp_encoding = 0u;  // vgi::MLP_ENCODING_*
    uint32_t incarnation = 0;    // the last slot allocation's incarnation (state word 28)
    // ---- view tier: two generations of per-pixel receivers, estimates,
    // owners and face hash; `view_next` is the generation the next frame
    // writes, the other holds the last frame.
    DeviceBuffer v_depth, v_normals, v_queue, v_receivers[2], v_estimates[2], v_owners[2], v_hash[2];
    DeviceBuffer v_pixels, v_counters;  // published records per pixel; far tier and presentation counters (16 words)
    uint32_t view_next = 0;
    uint32_t view_pixels[2] = {0u, 0u}, view_mask[2] = {0u, 0u};
    bool view_history = false;   // the last frame's generation may be reprojected
    // frame start, trace start, trace end, far tier start, frame end
    hipEvent_t view_ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    float view_display[3] = {0.031f, 1.0f, 0.19926f};  // noise threshold, spatial strength, temporal blend
    // The far tier: its settings, two generations of cells (one per view
    // generation), the material's output, the trace queue; the last frame's
    // camera (clip_from_world and frame number) for the next reprojection.
    vgi_far_settings far = {0.0f, 0u, 0u, 32.0f, 4u, 128u};
    DeviceBuffer f_cells[2], f_output, f_queue;
    uint32_t far_cx = 0u, far_cy = 0u, far_size = 0u;  // shape of the allocated cells (0: none)
    uint32_t far_last_cells = 0u;                      // cells of the last view frame (0: the tier was off)
    float far_prev_clip[16] = {};
    uint32_t far_prev_frame = 0u;
    bool far_prev = false;
    DeviceBuffer d_mlp_w, d_mlp_m, d_mlp_v, d_mlp_g, d_mlp_frag, d_mlp_loss;
    DeviceBuffer d_table, d_table_m, d_table_v, d_table_g;  // hash-grid encoding (FP32, Adam)
    std::vector<float> table_grad_last;                      // of the last vgi_mlp_train without update
    uint32_t mlp_step = 0, train_per_frame = 65536, train_steps = 4, train_cursor = 0;
    float mlp_lr = 1e-3f, last_loss = 0.0f;
    float table_lr = 1e-2f;  // the hash table's Adam rate
    uint32_t train_samples = 1;  // paths averaged into each training target
    // GPU timestamps of the last learned frame: inference (0, 1), and per
    // training step its MFMA passes (forward/backward, weight gradient) and
    // the whole step with Adam.
    hipEvent_t mlp_ev[2] = {};
    std::vector<hipEvent_t> mlp_step_ev;  // 3 per step: start, after the weight gradient, after the update
    uint32_t mlp_timed_steps = 0, mlp_timed_step_samples = 0, mlp_timed_queries = 0;
    bool mlp_timed = false;
    DeviceBuffer b_learned, b_queries, b_pred, d_perm, d_face_material;
    DeviceBuffer d_train_tasks, d_train_q, d_train_results, d_targets, d_acts, d_grads, d_train_pred, d_io, d_touched;
    // ---- wavefront path tracing of bound frames
    bool wavefront = false;
    bool dedup = false;  // the wavefront tracer's per-wavefront hit-face deduplication
    int cus = 0;
    DeviceBuffer d_order;      // bound faces in Morton order of their voxels
    bool has_order = false;
    DeviceBuffer wf_fields[vgi::WF_PATH_FIELDS], wf_queue, wf_hit0, wf_hit1, wf_vis, wf_control;
    hipStream_t side = nullptr;           // training paths beside a learned frame's trace
    hipEvent_t side_ev[2] = {};           // [tasks prepared, targets ready]
    // ---- edit retention: events of edits, applied with the next batch or frame
    std::vector<vgi_host::RetentionEvent> pending_events;
    DeviceBuffer d_events, d_measures, d_measure_max, d_material_light;
    uint32_t last_event_words = 0;
    bool material_light_dirty = true;
    // ---- bound faces that follow the exposed surface through edits
    bool follow_exposure = false;
    std::vector<vgi_face_task> bound_tasks;
    std::vector<uint32_t> bound_slots;
    std::vector<uint32_t> bound_of_slot;  // slot -> bound index or NONE
    std::vector<uint16_t> slot_material;  // material in each slot's identity word
    DeviceBuffer d_bound_index, d_bound_upd, d_bound_slot_upd, d_meta_writes, d_seed_writes;
};

namespace {

// ---------------------------------------------------------------- scene build
// The material presence mask and solidity of one region from the host grid.
void scan_region(vgi_tracer &t, int ri) {
    int rx = ri % t.rsx, ry = (ri / t.rsx) % t.rsy, rz = ri / (t.rsx * t.rsy);
    uint64_t *m = &t.masks[(size_t)ri * 4];
    m[0] = m[1] = m[2] = m[3] = 0;
    bool any = false;
    if (t.loaded[ri]) {
        int ox = (t.rmin.x + rx) * 16 - t.vmin.x, oy = (t.rmin.y + ry) * 16 - t.vmin.y, oz = (t.rmin.z + rz) * 16 - t.vmin.z;
        for (int z = 0; z < 16; ++z)
            for (int y = 0; y < 16; ++y) {
                const uint16_t *row = &t.grid[lin3(ox, oy + y, oz + z, t.gx, t.gy)];
                for (int x = 0; x < 16; ++x) {
                    uint16_t v = row[x];
                    if (v < vgi_host::AIR) {
                        any = true;
                        m[v >> 6] |= 1ull << (v & 63);
                    }
                }
            }
    }
    t.solid[ri] = any ? 1 : 0;
}

uint32_t *host_top_words(vgi_tracer &t, std::vector<uint32_t> &words) {
    size_t cells = (size_t)t.rsx * t.rsy * t.rsz;
    words.assign(cells * 2, 0u);
    for (int z = 0; z < t.rsz; ++z)
        for (int y = 0; y < t.rsy; ++y)
            for (int x = 0; x < t.rsx; ++x) {
                size_t i = lin3(x, y, z, t.rsx, t.rsy);
                if (!t.loaded[i]) continue;
                vgi_host::Extents e = t.top.at(Int3{t.rmin.x + x, t.rmin.y + y, t.rmin.z + z});
                // gpu_scene.rs:452-460: negative bytes, positive bytes (stride 0).
                words[i * 2] = (uint32_t)e[0] | ((uint32_t)e[2] << 8) | ((uint32_t)e[4] << 16);
                words[i * 2 + 1] = (uint32_t)e[1] | ((uint32_t)e[3] << 8) | ((uint32_t)e[5] << 16);
            }
    return words.data();
}

// Classify and pack the listed regions on the GPU (appending their tables and
// bricks to the cells buffer), then write their headers.
int gpu_pack(vgi_tracer &t, const std::vector<int32_t> &list4, bool restart) {
    uint32_t n = (uint32_t)(list4.size() / 4);
    if (n == 0) return 0;
    if (restart) t.cells_used = 0;
    HIP_OK(t.d_list.reserve(list4.size() * 4));
    HIP_OK(t.d_info.reserve((size_t)n * sizeof(vgi::RegionInfo)));
    HIP_OK(t.d_totals.reserve(16));
    HIP_OK(hipMemcpyAsync(t.d_list.ptr, list4.data(), list4.size() * 4, hipMemcpyHostToDevice, t.stream));
    vgi::launch_classify(t.d_grid.as<uint16_t>(), t.gx, t.gy, t.gz, t.vmin.x, t.vmin.y, t.vmin.z, t.d_list.as<int4>(), n,
                         t.d_info.as<vgi::RegionInfo>(), t.stream);
    HIP_OK(hipGetLastError());
    vgi::launch_scan(t.d_info.as<vgi::RegionInfo>(), n, t.cells_used, t.d_totals.as<uint32_t>(), t.stream);
    HIP_OK(hipGetLastError());
    uint32_t totals[2] = {0, 0};
    HIP_OK(hipMemcpyAsync(totals, t.d_totals.ptr, 8, hipMemcpyDeviceToHost, t.stream));
    HIP_OK(hipStreamSynchronize(t.stream));
    uint64_t need = (uint64_t)t.cells_used + (uint64_t)totals[0] * 64u + (uint64_t)totals[1] * 32u;
    if (need >= (1ull << 30)) return fail("packed scene exceeds 2^30 cell words (brick pointers are 30 bits)");
    if (need * 4 > t.d_cells.bytes) {
        // Grow, keeping the live words (tables of unchanged regions point into them).
        DeviceBuffer grown;
        HIP_OK(grown.reserve(need * 4 + (need * 4) / 2));
        if (t.cells_used) HIP_OK(hipMemcpyAsync(grown.ptr, t.d_cells.ptr, (size_t)t.cells_used * 4, hipMemcpyDeviceToDevice, t.stream));
        HIP_OK(hipStreamSynchronize(t.stream));
        HIP_OK(t.d_cells.release());
        t.d_cells = grown;
    }
    vgi::launch_pack(t.d_grid.as<uint16_t>(), t.gx, t.gy, t.gz, t.vmin.x, t.vmin.y, t.vmin.z, t.d_list.as<int4>(), n,
                     t.d_info.as<vgi::RegionInfo>(), t.d_cells.as<uint32_t>(), t.stream);
    HIP_OK(hipGetLastError());
    t.cells_used = (uint32_t)need;
    // Header words z, w come from the top level; written below for all regions.
    std::vector<uint32_t> zero((size_t)n * 2, 0u);
    HIP_OK(t.d_top.reserve(std::max<size_t>(zero.size() * 4, (size_t)t.rsx * t.rsy * t.rsz * 8)));
    HIP_OK(hipMemcpyAsync(t.d_top.ptr, zero.data(), zero.size() * 4, hipMemcpyHostToDevice, t.stream));
    vgi::launch_headers(t.d_list.as<int4>(), n, t.d_info.as<vgi::RegionInfo>(), t.d_top.as<uint32_t>(), t.d_headers.as<uint4>(),
                        t.rmin.x, t.rmin.y, t.rmin.z, t.rsx, t.rsy, t.stream);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(t.stream));
    return 0;
}

int upload_top(vgi_tracer &t) {
    std::vector<uint32_t> words;
    host_top_words(t, words);
    size_t cells = (size_t)t.rsx * t.rsy * t.rsz;
    HIP_OK(t.d_top.reserve(words.size() * 4));
    HIP_OK(hipMemcpyAsync(t.d_top.ptr, words.data(), words.size() * 4, hipMemcpyHostToDevice, t.stream));
    vgi::launch_top(t.d_top.as<uint32_t>(), (uint32_t)cells, t.d_headers.as<uint4>(), t.stream);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(t.stream));
    return 0;
}

std::vector<uint8_t> barrier_map(const vgi_tracer &t) {
    // world.rs:497: unloaded or solid regions are barriers.
    std::vector<uint8_t> b((size_t)t.rsx * t.rsy * t.rsz);
    for (size_t i = 0; i < b.size(); ++i) b[i] = (!t.loaded[i] || t.solid[i]) ? 1 : 0;
    return b;
}

std::vector<int32_t> region_list(const vgi_tracer &t, const std::vector<int> &indices) {
    std::vector<int32_t> list;
    list.reserve(indices.size() * 4);
    for (int i : indices) {
        int x = i % t.rsx, y = (i / t.rsx) % t.rsy, z = i / (t.rsx * t.rsy);
        list.push_back(t.rmin.x + x); list.push_back(t.rmin.y + y); list.push_back(t.rmin.z + z); list.push_back(0);
    }
    return list;
}

// Pages changed: re-lay the dense box if the page bounds moved, re-pack the
// affected regions on the GPU and update the top level for the regions whose
// barrier state flipped (TransportScene::with_page_contents, world.rs:982-1046).
vgi_host::PreviousScene scene_pages(const vgi_tracer &t, const std::vector<uint8_t> &solid, std::vector<uint8_t> &page_solid);
vgi_host::WorkPool &edit_pool();

// A region's 16^3 materials (x fastest) from the host grid; false (and air)
// when the region is not loaded.
bool region_materials(const vgi_tracer &t, Int3 r, std::vector<uint16_t> &out) {
    out.assign(4096, vgi_host::AIR);
    if (!t.have_regions) return false;
    int x = r.x - t.rmin.x, y = r.y - t.rmin.y, z = r.z - t.rmin.z;
    if (x < 0 || y < 0 || z < 0 || x >= t.rsx || y >= t.rsy || z >= t.rsz) return false;
    if (!t.loaded[lin3(x, y, z, t.rsx, t.rsy)]) return false;
    for (int lz = 0; lz < 16; ++lz)
        for (int ly = 0; ly < 16; ++ly)
            std::memcpy(&out[(size_t)16 * (ly + 16 * lz)], &t.grid[lin3(r.x * 16 - t.vmin.x, r.y * 16 - t.vmin.y + ly, r.z * 16 - t.vmin.z + lz, t.gx, t.gy)],
                        32);
    return true;
}

bool region_solid(const vgi_tracer &t, Int3 r) {
    if (!t.have_regions) return false;
    int x = r.x - t.rmin.x, y = r.y - t.rmin.y, z = r.z - t.rmin.z;
    if (x < 0 || y < 0 || z < 0 || x >= t.rsx || y >= t.rsy || z >= t.rsz) return false;
    size_t i = lin3(x, y, z, t.rsx, t.rsy);
    return t.loaded[i] && t.solid[i];
}

// The retention event of a page upload with the resident cache: every region
// of a loaded, replaced or removed page whose content or residency changed
// (world.rs:987-1022 changed regions) as SceneChanges::with_materials builds
// it (geometry.rs:110-229), with finalize_change_flux's load flux
// (publication.rs:423-454); queued for the next batch or frame.
float sky_luminance(const vgi_tracer &t) {
    auto lum = [](const float *c) { return 0.2126f * c[0] + 0.7152f * c[1] + 0.0722f * c[2]; };
    float sky = lum(t.transport.environment);
    float sun = 0.0f;
    for (uint32_t i = 0; i < t.light_count; ++i) {
        const float *l = &t.lights[(size_t)i * 12];
        if (l[3] < 0.0f) sun = std::max(sun, 0.5f * lum(&l[4]));
    }
    return std::max(sky, sun);
}

int apply_pages(vgi_tracer &t, const std::vector<std::pair<Int3, const uint16_t *>> &incoming, const std::vector<Int3> &removed) {
    // The scene before the upload, for the retention event: the touched
    // regions' contents and residency, and which pages held solid regions.
    bool events = t.resident && t.have_regions;
    std::vector<uint8_t> page_solid_before;
    vgi_host::PreviousScene previous{};
    std::map<std::tuple<int, int, int>, vgi_host::RegionDelta> deltas;  // z, y, x order
    if (events) {
        previous = scene_pages(t, t.solid, page_solid_before);
        auto capture = [&](Int3 page) {
            for (int o = 0; o < 8; ++o) {
                Int3 r{page.x * 2 + (o & 1), page.y * 2 + ((o >> 1) & 1), page.z * 2 + (o >> 2)};
                auto key = std::make_tuple(r.z, r.y, r.x);
                if (deltas.count(key)) continue;
                vgi_host::RegionDelta &d = deltas[key];
                d.coord = r;
                d.loaded_before = region_materials(t, r, d.before);
                d.solid_before = region_solid(t, r);
            }
        };
        for (auto &p : incoming) capture(p.first);
        for (auto &p : removed) capture(p);
    }
    auto finish_event = [&]() {
        if (!events) return;
        std::vector<vgi_host::RegionDelta> changed;
        for (auto &kv : deltas) {
            vgi_host::RegionDelta &d = kv.second;
            d.loaded_after = region_materials(t, d.coord, d.after);
            d.solid_after = region_solid(t, d.coord);
            if (!d.loaded_before && !d.loaded_after) continue;
            if (d.loaded_before && d.loaded_after && d.before == d.after) continue;  // the same region (world.rs:1008-1014)
            changed.push_back(std::move(d));
        }
        vgi_host::RetentionEvent ev;
        if (vgi_host::scene_event(changed, previous, t.emission, sky_luminance(t), {}, ev, &edit_pool()))
            t.pending_events.push_back(std::move(ev));
    };
    // Previous state per region coordinate (for the flip list).
    std::unordered_map<uint64_t, uint8_t> before;  // bit0 loaded, bit1 solid
    auto note = [&](int rx, int ry, int rz) {
        uint64_t k = page_key(rx, ry, rz);
        if (before.count(k)) return;
        uint8_t st = 0;
        if (t.have_regions && rx >= t.rmin.x && ry >= t.rmin.y && rz >= t.rmin.z && rx < t.rmin.x + t.rsx && ry < t.rmin.y + t.rsy &&
            rz < t.rmin.z + t.rsz) {
            size_t i = lin3(rx - t.rmin.x, ry - t.rmin.y, rz - t.rmin.z, t.rsx, t.rsy);
            st = (uint8_t)((t.loaded[i] ? 1 : 0) | (t.solid[i] ? 2 : 0));
        }
        before[k] = st;
    };
    std::vector<Int3> touched_pages;
    for (auto &p : incoming) touched_pages.push_back(p.first);
    for (auto &p : removed) touched_pages.push_back(p);
    for (const Int3 &p : touched_pages)
        for (int o = 0; o < 8; ++o) note(p.x * 2 + (o & 1), p.y * 2 + ((o >> 1) & 1), p.z * 2 + (o >> 2));
    // New page set and box.
    Int3 old_vmin = t.vmin;
    int ogx = t.gx, ogy = t.gy, ogz = t.gz;
    std::vector<Int3> old_pages;
    for (auto &kv : t.pages) old_pages.push_back(kv.second);
    for (const Int3 &p : removed) t.pages.erase(page_key(p.x, p.y, p.z));
    for (auto &p : incoming) t.pages[page_key(p.first.x, p.first.y, p.first.z)] = p.first;
    if (t.pages.empty()) {
        t.have_regions = false;
        t.rsx = t.rsy = t.rsz = 0;
        t.grid.clear();
        t.top = vgi_host::TopLevelField();
        t.emitters_dirty = true;
        finish_event();
        return 0;
    }
    Int3 pmn{INT32_MAX, INT32_MAX, INT32_MAX}, pmx{INT32_MIN, INT32_MIN, INT32_MIN};
    for (auto &kv : t.pages) {
        const Int3 &p = kv.second;
        pmn = Int3{std::min(pmn.x, p.x), std::min(pmn.y, p.y), std::min(pmn.z, p.z)};
        pmx = Int3{std::max(pmx.x, p.x), std::max(pmx.y, p.y), std::max(pmx.z, p.z)};
    }
    Int3 nvmin{pmn.x * 32, pmn.y * 32, pmn.z * 32};
    int ngx = (pmx.x - pmn.x + 1) * 32, ngy = (pmx.y - pmn.y + 1) * 32, ngz = (pmx.z - pmn.z + 1) * 32;
    bool relayout = !t.have_regions || !(nvmin == old_vmin) || ngx != ogx || ngy != ogy || ngz != ogz;
    if (relayout) {
        std::vector<uint16_t> grid((size_t)ngx * ngy * ngz, vgi_host::AIR);
        for (const Int3 &p : old_pages) {
            if (!t.pages.count(page_key(p.x, p.y, p.z))) continue;
            for (int z = 0; z < 32; ++z)
                for (int y = 0; y < 32; ++y) {
                    const uint16_t *src = &t.grid[lin3(p.x * 32 - old_vmin.x, p.y * 32 - old_vmin.y + y, p.z * 32 - old_vmin.z + z, ogx, ogy)];
                    uint16_t *dst = &grid[lin3(p.x * 32 - nvmin.x, p.y * 32 - nvmin.y + y, p.z * 32 - nvmin.z + z, ngx, ngy)];
                    std::memcpy(dst, src, 64);
                }
        }
        t.grid.swap(grid);
        t.vmin = nvmin; t.gx = ngx; t.gy = ngy; t.gz = ngz;
        t.rmin = Int3{pmn.x * 2, pmn.y * 2, pmn.z * 2};
        t.rsx = ngx / 16; t.rsy = ngy / 16; t.rsz = ngz / 16;
    }
    for (const Int3 &p : removed)
        for (int z = 0; z < 32; ++z)
            for (int y = 0; y < 32; ++y) {
                int px = p.x * 32 - t.vmin.x, py = p.y * 32 - t.vmin.y + y, pz = p.z * 32 - t.vmin.z + z;
                if (px < 0 || py < 0 || pz < 0 || px >= t.gx || py >= t.gy || pz >= t.gz) continue;
                std::fill_n(&t.grid[lin3(px, py, pz, t.gx, t.gy)], 32, vgi_host::AIR);
            }
    for (auto &p : incoming)
        for (int z = 0; z < 32; ++z)
            for (int y = 0; y < 32; ++y)
                std::memcpy(&t.grid[lin3(p.first.x * 32 - t.vmin.x, p.first.y * 32 - t.vmin.y + y, p.first.z * 32 - t.vmin.z + z, t.gx, t.gy)],
                            p.second + (size_t)32 * (y + 32 * z), 64);
    size_t rcells = (size_t)t.rsx * t.rsy * t.rsz;
    std::vector<int> repack;
    if (relayout) {
        t.loaded.assign(rcells, 0);
        t.solid.assign(rcells, 0);
        t.masks.assign(rcells * 4, 0);
        for (auto &kv : t.pages) {
            const Int3 &p = kv.second;
            for (int o = 0; o < 8; ++o)
                t.loaded[lin3(p.x * 2 + (o & 1) - t.rmin.x, p.y * 2 + ((o >> 1) & 1) - t.rmin.y, p.z * 2 + (o >> 2) - t.rmin.z, t.rsx, t.rsy)] = 1;
        }
        for (size_t i = 0; i < rcells; ++i) {
            scan_region(t, (int)i);
            if (t.loaded[i]) repack.push_back((int)i);
        }
    } else {
        for (const Int3 &p : touched_pages)
            for (int o = 0; o < 8; ++o) {
                int i = (int)lin3(p.x * 2 + (o & 1) - t.rmin.x, p.y * 2 + ((o >> 1) & 1) - t.rmin.y, p.z * 2 + (o >> 2) - t.rmin.z, t.rsx, t.rsy);
                t.loaded[i] = t.pages.count(page_key(p.x, p.y, p.z)) ? 1 : 0;
                scan_region(t, i);
                if (t.loaded[i]) repack.push_back(i);
            }
        std::sort(repack.begin(), repack.end());
        repack.erase(std::unique(repack.begin(), repack.end()), repack.end());
    }
    t.bmin = t.rmin;
    t.bmax = Int3{t.rmin.x + t.rsx - 1, t.rmin.y + t.rsy - 1, t.rmin.z + t.rsz - 1};
    t.have_regions = true;
    HIP_OK(t.d_grid.reserve(t.grid.size() * 2));
    HIP_OK(hipMemcpyAsync(t.d_grid.ptr, t.grid.data(), t.grid.size() * 2, hipMemcpyHostToDevice, t.stream));
    if (relayout) {
        HIP_OK(t.d_headers.reserve(rcells * 16));
        HIP_OK(hipMemsetAsync(t.d_headers.ptr, 0, rcells * 16, t.stream));
    } else {
        // Unloaded regions get kind 0 headers.
        for (const Int3 &p : removed)
            for (int o = 0; o < 8; ++o) {
                size_t i = lin3(p.x * 2 + (o & 1) - t.rmin.x, p.y * 2 + ((o >> 1) & 1) - t.rmin.y, p.z * 2 + (o >> 2) - t.rmin.z, t.rsx, t.rsy);
                if (t.loaded[i]) continue;
                HIP_OK(hipMemsetAsync(t.d_headers.as<uint4>() + i, 0, 16, t.stream));
            }
    }
    HIP_OK(hipEventRecord(t.ev[6], t.stream));
    if (gpu_pack(t, region_list(t, repack), relayout) != 0) return -1;
    HIP_OK(hipEventRecord(t.ev[7], t.stream));
    // Regions whose barrier state flipped, in z, y, x order (world.rs:595-601).
    std::vector<std::pair<std::tuple<int, int, int>, Int3>> flips;
    for (auto &kv : before) {
        (void)kv;
    }
    for (const Int3 &p : touched_pages)
        for (int o = 0; o < 8; ++o) {
            Int3 r{p.x * 2 + (o & 1), p.y * 2 + ((o >> 1) & 1), p.z * 2 + (o >> 2)};
            size_t i = lin3(r.x - t.rmin.x, r.y - t.rmin.y, r.z - t.rmin.z, t.rsx, t.rsy);
            uint8_t now = (uint8_t)((t.loaded[i] ? 1 : 0) | (t.solid[i] ? 2 : 0));
            if (now != before[page_key(r.x, r.y, r.z)]) flips.push_back({std::make_tuple(r.z, r.y, r.x), r});
        }
    std::sort(flips.begin(), flips.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
    flips.erase(std::unique(flips.begin(), flips.end(), [](const auto &a, const auto &b) { return a.first == b.first; }), flips.end());
    if (!flips.empty() || relayout) {
        std::vector<Int3> keys;
        for (auto &f : flips) keys.push_back(f.second);
        t.top.update(t.bmin, t.bmax, keys, barrier_map(t));
    }
    if (upload_top(t) != 0) return -1;
    HIP_OK(hipEventSynchronize(t.ev[7]));
    HIP_OK(hipEventElapsedTime(&t.timings[3], t.ev[6], t.ev[7]));
    t.emitters_dirty = true;
    finish_event();
    return 0;
}

// Voxel edits (pages.rs:498 page overlay): scatter into the GPU grid, re-pack
// the edited regions and update the top level if any region's emptiness
// flipped. With the resident cache, the edit's retention event is queued
// for the next batch (pages.rs:446-490), and bound faces that follow the
// exposed surface are added and removed.
int update_exposure(vgi_tracer &t, const std::vector<vgi_host::Int3> &changed, const std::vector<vgi_host::RegionDelta> &deltas);

// The solid regions of a scene as retention events need them: their bounding
// box and which 32^3 pages hold one (`solid`: one byte per region of the box).
// The edit path's host threads (up to 8 with the caller), shared by tracers.
vgi_host::WorkPool &edit_pool() {
    static vgi_host::WorkPool pool(std::min(std::max(std::thread::hardware_concurrency(), 1u), 8u) - 1u);
    return pool;
}

vgi_host::PreviousScene scene_pages(const vgi_tracer &t, const std::vector<uint8_t> &solid, std::vector<uint8_t> &page_solid) {
    vgi_host::PreviousScene prev;
    prev.any_solid = false;
    prev.solid_min = Int3{INT32_MAX, INT32_MAX, INT32_MAX};
    prev.solid_max = Int3{INT32_MIN, INT32_MIN, INT32_MIN};
    prev.page_min = Int3{t.rmin.x / 2, t.rmin.y / 2, t.rmin.z / 2};
    prev.psx = t.rsx / 2; prev.psy = t.rsy / 2; prev.psz = t.rsz / 2;
    page_solid.assign((size_t)prev.psx * prev.psy * prev.psz, 0);
    for (int z = 0; z < t.rsz; ++z)
        for (int y = 0; y < t.rsy; ++y)
            for (int x = 0; x < t.rsx; ++x) {
                if (!solid[lin3(x, y, z, t.rsx, t.rsy)]) continue;
                Int3 r{t.rmin.x + x, t.rmin.y + y, t.rmin.z + z};
                prev.any_solid = true;
                prev.solid_min = Int3{std::min(prev.solid_min.x, r.x), std::min(prev.solid_min.y, r.y), std::min(prev.solid_min.z, r.z)};
                prev.solid_max = Int3{std::max(prev.solid_max.x, r.x), std::max(prev.solid_max.y, r.y), std::max(prev.solid_max.z, r.z)};
                page_solid[lin3(x / 2, y / 2, z / 2, prev.psx, prev.psy)] = 1;
            }
    prev.page_solid = page_solid.data();
    return prev;
}

int staged_upload(vgi_tracer &t, void *dst, const void *src, size_t bytes);
void stage_recycle(vgi_tracer &t);

int apply_edits(vgi_tracer &t, const int32_t *xyz, const uint16_t *values, uint32_t count) {
    if (!t.have_regions) return 0;
    using clock = std::chrono::steady_clock;
    auto ms_since = [](clock::time_point a) { return std::chrono::duration<float, std::milli>(clock::now() - a).count(); };
    auto t_start = clock::now();
    t.timings[4] = t.timings[5] = t.timings[6] = t.timings[7] = 0.0f;
    stage_recycle(t);
    std::vector<uint32_t> accepted;
    std::vector<int> dirty;
    accepted.reserve(count);
    for (uint32_t e = 0; e < count; ++e) {
        int x = xyz[e * 3], y = xyz[e * 3 + 1], z = xyz[e * 3 + 2];
        int rx = floor_div(x, 16) - t.rmin.x, ry = floor_div(y, 16) - t.rmin.y, rz = floor_div(z, 16) - t.rmin.z;
        if (rx < 0 || ry < 0 || rz < 0 || rx >= t.rsx || ry >= t.rsy || rz >= t.rsz) continue;
        size_t ri = lin3(rx, ry, rz, t.rsx, t.rsy);
        if (!t.loaded[ri]) continue;  // edits outside loaded pages wait for the page (dropped here)
        accepted.push_back(e);
        dirty.push_back((int)ri);
    }
    if (accepted.empty()) return 0;
    std::sort(dirty.begin(), dirty.end());
    dirty.erase(std::unique(dirty.begin(), dirty.end()), dirty.end());
    auto region_origin = [&](int ri) {
        int x = ri % t.rsx, y = (ri / t.rsx) % t.rsy, z = ri / (t.rsx * t.rsy);
        return Int3{(t.rmin.x + x) * 16, (t.rmin.y + y) * 16, (t.rmin.z + z) * 16};
    };
    auto copy_region = [&](int ri, std::vector<uint16_t> &out) {
        Int3 o = region_origin(ri);
        out.resize(4096);
        for (int z = 0; z < 16; ++z)
            for (int y = 0; y < 16; ++y)
                std::memcpy(&out[(size_t)16 * (y + 16 * z)], &t.grid[lin3(o.x - t.vmin.x, o.y - t.vmin.y + y, o.z - t.vmin.z + z, t.gx, t.gy)], 32);
    };
    // The scene before the edit, for its retention event.
    bool events = t.resident;
    std::vector<vgi_host::RegionDelta> deltas;
    std::vector<uint8_t> solid_before;
    vgi_host::WorkPool &pool = edit_pool();
    if (events) {
        solid_before = t.solid;
        deltas.resize(dirty.size());
        pool.run(dirty.size(), [&](size_t begin, size_t end) {
            for (size_t k = begin; k < end; ++k) {
                int ri = dirty[k];
                Int3 o = region_origin(ri);
                deltas[k].coord = Int3{o.x / 16, o.y / 16, o.z / 16};
                copy_region(ri, deltas[k].before);
                deltas[k].solid_before = t.solid[ri] != 0;
            }
        });
    }
    // Edits apply in order (a later edit of a voxel wins); the GPU grid gets
    // each edited voxel's final value once, before the re-pack.
    std::vector<uint32_t> idx;
    std::vector<vgi_host::Int3> edited;
    idx.reserve(accepted.size());
    edited.reserve(accepted.size());
    for (uint32_t e : accepted) {
        int x = xyz[e * 3], y = xyz[e * 3 + 1], z = xyz[e * 3 + 2];
        size_t gi = lin3(x - t.vmin.x, y - t.vmin.y, z - t.vmin.z, t.gx, t.gy);
        t.grid[gi] = values[e];
        idx.push_back((uint32_t)gi);
        edited.push_back(Int3{x, y, z});
    }
    std::sort(idx.begin(), idx.end());
    idx.erase(std::unique(idx.begin(), idx.end()), idx.end());
    std::vector<uint16_t> val(idx.size());
    for (size_t k = 0; k < idx.size(); ++k) val[k] = t.grid[idx[k]];
    // Emitting materials near an edit change the emitter distribution.
    uint64_t emitting[4] = {0, 0, 0, 0};
    for (int m = 0; m < 256; ++m) {
        float w = std::max(std::max(t.emission[m * 3], t.emission[m * 3 + 1]), t.emission[m * 3 + 2]);
        if (w > 0.0f) emitting[m >> 6] |= 1ull << (m & 63);
    }
    auto emits = [&](int ri) {
        const uint64_t *m = &t.masks[(size_t)ri * 4];
        return (m[0] & emitting[0]) | (m[1] & emitting[1]) | (m[2] & emitting[2]) | (m[3] & emitting[3]);
    };
    // An edit next to an emitting region (before or after the edit) changes
    // the emitter distribution. Regions are rescanned in parallel (each
    // writes only its own masks and solidity); the emitter test reads the
    // neighbourhood's masks before and after, in passes of their own.
    const int nb[7][3] = {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    auto near_emitting = [&](int ri) {
        int x = ri % t.rsx, y = (ri / t.rsx) % t.rsy, z = ri / (t.rsx * t.rsy);
        for (auto &o : nb) {
            int nx = x + o[0], ny = y + o[1], nz = z + o[2];
            if (nx < 0 || ny < 0 || nz < 0 || nx >= t.rsx || ny >= t.rsy || nz >= t.rsz) continue;
            if (emits((int)lin3(nx, ny, nz, t.rsx, t.rsy))) return true;
        }
        return false;
    };
    std::vector<uint8_t> was(dirty.size()), near(dirty.size(), 0);
    pool.run(dirty.size(), [&](size_t begin, size_t end) {
        for (size_t k = begin; k < end; ++k) {
            was[k] = t.solid[dirty[k]];
            near[k] = near_emitting(dirty[k]) ? 1 : 0;
        }
    });
    pool.run(dirty.size(), [&](size_t begin, size_t end) {
        for (size_t k = begin; k < end; ++k) scan_region(t, dirty[k]);
    });
    pool.run(dirty.size(), [&](size_t begin, size_t end) {
        for (size_t k = begin; k < end; ++k)
            if (near_emitting(dirty[k])) near[k] = 1;
    });
    std::vector<std::pair<std::tuple<int, int, int>, Int3>> flips;
    for (size_t k = 0; k < dirty.size(); ++k) {
        int ri = dirty[k];
        int x = ri % t.rsx, y = (ri / t.rsx) % t.rsy, z = ri / (t.rsx * t.rsy);
        if (near[k]) t.emitters_dirty = true;
        if (was[k] != t.solid[ri]) flips.push_back({std::make_tuple(t.rmin.z + z, t.rmin.y + y, t.rmin.x + x), Int3{t.rmin.x + x, t.rmin.y + y, t.rmin.z + z}});
    }
    // The retention event and the exposure update need only the host's scene
    // and the stream-ordered uploads of their own buffers, so they run before
    // the first wait on the stream: a caller that edits while a frame is in
    // flight overlaps them with that frame.
    float regions_ms = ms_since(t_start);
    std::string exposure_error;
    if (events) {
        auto t_event = clock::now();
        // The retention event (SceneChanges::with_materials .. SpatialHistory::event).
        std::vector<std::vector<vgi_host::Int3>> changed_of(dirty.size());
        pool.run(dirty.size(), [&](size_t begin, size_t end) {
            for (size_t k = begin; k < end; ++k) {
                copy_region(dirty[k], deltas[k].after);
                deltas[k].solid_after = t.solid[dirty[k]] != 0;
                const vgi_host::RegionDelta &d = deltas[k];
                for (int i = 0; i < 4096; ++i)
                    if (d.before[i] != d.after[i])
                        changed_of[k].push_back(Int3{d.coord.x * 16 + (i & 15), d.coord.y * 16 + ((i >> 4) & 15), d.coord.z * 16 + (i >> 8)});
            }
        });
        std::vector<vgi_host::Int3> changed;
        for (const auto &c : changed_of) changed.insert(changed.end(), c.begin(), c.end());
        std::vector<uint8_t> page_solid;
        vgi_host::PreviousScene prev = scene_pages(t, solid_before, page_solid);
        vgi_host::RetentionEvent ev;
        if (vgi_host::edit_event(deltas, prev, t.emission, edited, ev, &pool)) t.pending_events.push_back(std::move(ev));
        t.timings[6] = ms_since(t_event);
        if (t.follow_exposure && !changed.empty()) {
            auto t_exposure = clock::now();
            if (update_exposure(t, changed, deltas) != 0) {
                // The scene edit itself still reaches the GPU below; the
                // bound set is left part way and frames refuse it until the
                // caller binds again.
                exposure_error = g_error;
                t.bound_broken = true;
            }
            t.timings[7] = ms_since(t_exposure);
        }
    }
    auto t_pack = clock::now();
    // Keep the cells buffer bounded: re-pack everything once appended tables
    // and bricks outgrow four times the last full pack.
    bool full = t.cells_used > 4u * std::max<uint32_t>(t.cells_live, 1u << 16);
    std::vector<int> repack = dirty;
    if (full) {
        repack.clear();
        for (size_t i = 0; i < t.loaded.size(); ++i)
            if (t.loaded[i]) repack.push_back((int)i);
    }
    // The GPU grid's copies of the edited voxels, read by the re-pack.
    HIP_OK(t.d_scatter_idx.reserve(idx.size() * 4));
    HIP_OK(t.d_scatter_val.reserve(val.size() * 2));
    if (staged_upload(t, t.d_scatter_idx.ptr, idx.data(), idx.size() * 4) != 0 ||
        staged_upload(t, t.d_scatter_val.ptr, val.data(), val.size() * 2) != 0)
        return -1;
    vgi::launch_scatter_voxels(t.d_grid.as<uint16_t>(), t.d_scatter_idx.as<uint32_t>(), t.d_scatter_val.as<uint16_t>(),
                               (uint32_t)idx.size(), t.stream);
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord(t.ev[6], t.stream));
    if (gpu_pack(t, region_list(t, repack), full) != 0) return -1;
    if (full) t.cells_live = t.cells_used;
    HIP_OK(hipEventRecord(t.ev[7], t.stream));
    t.timings[4] = regions_ms + ms_since(t_pack);
    auto t_top = clock::now();
    if (!flips.empty()) {
        std::sort(flips.begin(), flips.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
        std::vector<Int3> keys;
        for (auto &f : flips) keys.push_back(f.second);
        t.top.update(t.bmin, t.bmax, keys, barrier_map(t));
    }
    if (upload_top(t) != 0) return -1;
    HIP_OK(hipEventSynchronize(t.ev[7]));
    HIP_OK(hipEventElapsedTime(&t.timings[3], t.ev[6], t.ev[7]));
    t.timings[5] = ms_since(t_top);
    if (!exposure_error.empty())
        return fail("the edit was applied, but the bound faces could not follow it (" + exposure_error +
                    "): bind faces again before the next frame");
    return 0;
}

// ---------------------------------------------------------------- optics / emitters
int sync_emitters(vgi_tracer &t, bool no_alloc = false) {
    if (!t.emitters_dirty) return 0;
    vgi_host::VoxelSource src{t.grid.data(), t.vmin, t.gx, t.gy, t.gz, t.loaded.data(), t.rmin, t.rsx, t.rsy, t.rsz};
    uint64_t emitting[4] = {0, 0, 0, 0};
    for (int m = 0; m < 256; ++m) {
        float w = std::max(std::max(t.emission[m * 3], t.emission[m * 3 + 1]), t.emission[m * 3 + 2]);
        if (w > 0.0f) emitting[m >> 6] |= 1ull << (m & 63);
    }
    std::vector<Int3> regions;
    if (t.have_regions)
        for (int z = 0; z < t.rsz; ++z)
            for (int y = 0; y < t.rsy; ++y)
                for (int x = 0; x < t.rsx; ++x) {
                    size_t i = lin3(x, y, z, t.rsx, t.rsy);
                    const uint64_t *m = &t.masks[i * 4];
                    if (!t.loaded[i]) continue;
                    if ((m[0] & emitting[0]) | (m[1] & emitting[1]) | (m[2] & emitting[2]) | (m[3] & emitting[3]))
                        regions.push_back(Int3{t.rmin.x + x, t.rmin.y + y, t.rmin.z + z});
                }
    vgi_host::build_emitters(src, regions, t.emission, t.emitters);
    // The walk words hold a child's index in 26 bits.
    if (t.emitters.nodes >= (1u << 26)) return fail("more than 2^25 emissive faces: the light tree's node indices exceed 26 bits");
    size_t words = t.emitters.words.size(), walk = t.emitters.walk.size();
    if (no_alloc && ((words + walk) * 4 > t.d_emitters.bytes || t.d_emitters.ptr == nullptr))
        return fail("the emitter table outgrew its device buffer: one allocating call (vgi_submit_faces or vgi_run_frame) "
                    "must follow this scene or material change");
    // The emitter words, then the tree's walk words (Optics::tree).
    HIP_OK(t.d_emitters.reserve((words + walk) * 4));
    HIP_OK(hipMemcpyAsync(t.d_emitters.ptr, t.emitters.words.data(), words * 4, hipMemcpyHostToDevice, t.stream));
    if (walk)
        HIP_OK(hipMemcpyAsync(t.d_emitters.as<uint32_t>() + words, t.emitters.walk.data(), walk * 4, hipMemcpyHostToDevice,
                              t.stream));
    HIP_OK(hipStreamSynchronize(t.stream));
    t.emitters_dirty = false;
    return 0;
}

uint32_t lanes_for(const vgi_transport &tr) {
    if (tr.lanes) return tr.lanes;
    // voxel_trace.wgsl's lanes: 4 below 16 samples, else 8. When no lane
    // would hold more than one sample, one lane per face adds the samples in
    // the same order (the lane sums are the samples themselves) and wastes
    // no idle lanes.
    uint32_t reference = tr.samples < 16u ? 4u : 8u;
    return tr.samples <= reference ? 1u : reference;
}

void views(vgi_tracer &t, vgi::SceneView &s, vgi::Optics &o) {
    s.headers = t.d_headers.as<uint4>();
    s.cells = t.d_cells.as<uint32_t>();
    s.minx = t.bmin.x; s.miny = t.bmin.y; s.minz = t.bmin.z;
    s.maxx = t.bmax.x; s.maxy = t.bmax.y; s.maxz = t.bmax.z;
    s.sx = t.rsx; s.sy = t.rsy;
    o.materials = t.d_materials.as<float4>();
    o.lights = t.d_lights.as<float4>();
    o.emitters = t.d_emitters.as<uint4>();
    o.tree = t.d_emitters.as<uint4>() + t.emitters.words.size() / 4;
    o.light_count = t.light_count;
    o.directional = std::min(t.directional, t.light_count);
    o.emitter_faces = t.emitters.faces;
    o.total_weight = t.emitters.total_weight;
    o.has_tree = t.emitters.nodes > 0 ? 1u : 0u;
    o.env_r = t.transport.environment[0]; o.env_g = t.transport.environment[1]; o.env_b = t.transport.environment[2];
    o.max_distance = t.transport.max_distance > 0.0f ? t.transport.max_distance : 768.0f;
    o.reach = t.transport.reach;
    o.limited_reach = t.transport.limited_reach;
    o.bounces = t.transport.bounces;
    o.cache_index = t.d_index.as<uint4>();
    o.cache_state = t.d_state.as<uint32_t>();
    o.cache_mask = t.resident ? (uint32_t)(t.face_slots.entries.size() - 1) : 0u;
    o.cache_on = (t.resident && !t.transport.plain) ? 1u : 0u;
    o.direct_reuse = 0u;
}

// ---------------------------------------------------------------- staged uploads
// `bytes` from host `src` to device `dst` on the tracer's stream through the
// pinned staging region: the host may reuse `src` at once and does not wait.
const size_t STAGE_LIMIT = (size_t)64 << 20;  // largest pinned staging region per tracer

int staged_upload(vgi_tracer &t, void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return 0;
    size_t need = (bytes + 255) & ~(size_t)255;
    if (t.stage_used + need > t.stage_bytes) {
        bool had = t.stage_bytes > 0;
        if (t.stage_pending) HIP_OK(hipEventSynchronize(t.stage_done));
        t.stage_pending = false;
        t.stage_used = 0;
        if (need > t.stage_bytes) {
            // Grow by doubling from 256 KiB to what one edit needs; an upload
            // beyond the limit goes from the caller's memory and waits.
            if (need > STAGE_LIMIT) {
                HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, t.stream));
                HIP_OK(hipStreamSynchronize(t.stream));
                return 0;
            }
            if (t.stage) HIP_OK(hipHostFree(t.stage));
            t.stage = nullptr;
            t.stage_bytes = 0;
            size_t cap = (size_t)256 << 10;
            while (cap < need) cap <<= 1;
            HIP_OK(hipHostMalloc((void **)&t.stage, cap, hipHostMallocDefault));
            t.stage_bytes = cap;
            t.stage_grows += 1;
        } else if (had) {
            t.stage_wraps += 1;
        }
    }
    uint8_t *p = t.stage + t.stage_used;
    std::memcpy(p, src, bytes);
    t.stage_used += need;
    HIP_OK(hipMemcpyAsync(dst, p, bytes, hipMemcpyHostToDevice, t.stream));
    if (!t.stage_done) HIP_OK(hipEventCreateWithFlags(&t.stage_done, hipEventDisableTiming));
    HIP_OK(hipEventRecord(t.stage_done, t.stream));
    t.stage_pending = true;
    return 0;
}

// Start the staging region over if the stream has passed its copies.
void stage_recycle(vgi_tracer &t) {
    if (t.stage_pending && hipEventQuery(t.stage_done) != hipSuccess) return;
    t.stage_pending = false;
    t.stage_used = 0;
}

// ---------------------------------------------------------------- resident slots
// resident.rs:770-812 (allocate_slot): a free slot (LIFO) or the high-water mark.
int register_faces(vgi_tracer &t, const vgi_face_task *tasks, uint32_t count, std::vector<uint32_t> &slots) {
    slots.resize(count);
    uint32_t first_new = t.high_water;
    std::vector<uint32_t> identities, meta_writes;
    vgi_host::VoxelSource src{t.grid.data(), t.vmin, t.gx, t.gy, t.gz, t.loaded.data(), t.rmin, t.rsx, t.rsy, t.rsz};
    for (uint32_t i = 0; i < count; ++i) {
        const vgi_face_task &f = tasks[i];
        if (f.face == VGI_NO_FACE) { slots[i] = vgi_host::NONE; continue; }
        long at = t.face_slots.find(f.voxel[0], f.voxel[1], f.voxel[2], f.face);
        if (at >= 0) {
            uint32_t slot = t.face_slots.entries[at][3] & 0x1fffffffu;
            slots[i] = slot;
            // Material replacement (resident.rs:443-450): a relit face keeps
            // its slot; its identity word takes the voxel's new material.
            uint32_t m = t.have_regions ? src.material_at(f.voxel[0], f.voxel[1], f.voxel[2]) : 0xffffu;
            uint16_t material = (uint16_t)(m < vgi_host::AIR ? m : 0u);
            if (slot < t.slot_material.size() && t.slot_material[slot] != material) {
                t.slot_material[slot] = material;
                meta_writes.push_back(slot * 12u + 3u);
                meta_writes.push_back(f.face | 8u | ((uint32_t)material << 16));
            }
            continue;
        }
        uint32_t slot;
        if (!t.free_slots.empty()) { slot = t.free_slots.back(); t.free_slots.pop_back(); }
        else slot = t.high_water++;
        if (slot >= (1u << 29)) return fail("resident slot exceeds the face index");
        t.face_slots.insert(f.voxel[0], f.voxel[1], f.voxel[2], f.face, slot);
        uint32_t m = t.have_regions ? src.material_at(f.voxel[0], f.voxel[1], f.voxel[2]) : 0xffffu;
        uint32_t material = m < vgi_host::AIR ? m : 0u;
        identities.push_back(slot);
        identities.push_back((uint32_t)f.voxel[0]); identities.push_back((uint32_t)f.voxel[1]); identities.push_back((uint32_t)f.voxel[2]);
        identities.push_back(f.face | 8u | (material << 16));
        if (t.slot_material.size() <= slot) t.slot_material.resize((size_t)slot + 1, 0xffffu);
        t.slot_material[slot] = (uint16_t)material;
        slots[i] = slot;
    }
    // Grow the resident arrays, keeping existing slots.
    if (t.high_water > t.slot_capacity) {
        size_t cap = std::max<size_t>(1024, t.slot_capacity);
        while (cap < t.high_water) cap *= 2;
        DeviceBuffer ns, nc, nm;
        HIP_OK(ns.reserve(cap * 144));
        HIP_OK(nc.reserve(cap * 48));
        HIP_OK(nm.reserve(cap * 48));
        HIP_OK(hipMemsetAsync(ns.ptr, 0, cap * 144, t.stream));
        HIP_OK(hipMemsetAsync(nc.ptr, 0, cap * 48, t.stream));
        HIP_OK(hipMemsetAsync(nm.ptr, 0, cap * 48, t.stream));
        if (t.slot_capacity) {
            HIP_OK(hipMemcpyAsync(ns.ptr, t.d_state.ptr, t.slot_capacity * 144, hipMemcpyDeviceToDevice, t.stream));
            HIP_OK(hipMemcpyAsync(nc.ptr, t.d_checkpoint.ptr, t.slot_capacity * 48, hipMemcpyDeviceToDevice, t.stream));
            HIP_OK(hipMemcpyAsync(nm.ptr, t.d_meta.ptr, t.slot_capacity * 48, hipMemcpyDeviceToDevice, t.stream));
        }
        // Selection keys persist; slots never classified have none.
        DeviceBuffer nk;
        HIP_OK(nk.reserve(cap * 4));
        HIP_OK(hipMemsetAsync(nk.ptr, 0xff, cap * 4, t.stream));
        if (t.slot_capacity) HIP_OK(hipMemcpyAsync(nk.ptr, t.d_keys.ptr, t.slot_capacity * 4, hipMemcpyDeviceToDevice, t.stream));
        // Sun corner words persist; new slots have none (0).
        DeviceBuffer nw;
        HIP_OK(nw.reserve(cap * 4));
        HIP_OK(hipMemsetAsync(nw.ptr, 0, cap * 4, t.stream));
        if (t.slot_capacity) HIP_OK(hipMemcpyAsync(nw.ptr, t.d_corners.ptr, t.slot_capacity * 4, hipMemcpyDeviceToDevice, t.stream));
        HIP_OK(hipStreamSynchronize(t.stream));
        t.d_state.release(); t.d_checkpoint.release(); t.d_meta.release(); t.d_keys.release(); t.d_corners.release();
        t.d_state = ns; t.d_checkpoint = nc; t.d_meta = nm; t.d_keys = nk; t.d_corners = nw;
        t.slot_capacity = cap;
    }
    // Seed new slots (resident.rs:1426-1470 from an empty FaceWork: all zero)
    // and write their identity records. Slots come from the high-water mark
    // (nothing is freed by this build), so new slots form one range.
    if (!identities.empty()) {
        uint32_t lo = identities[0], hi = identities[0];
        for (size_t k = 0; k < identities.size(); k += 5) { lo = std::min(lo, identities[k]); hi = std::max(hi, identities[k]); }
        size_t n = (size_t)(hi - lo + 1);
        std::vector<uint32_t> meta(n * 12, 0u);
        for (size_t k = 0; k < identities.size(); k += 5) {
            uint32_t *m = &meta[(size_t)(identities[k] - lo) * 12];
            m[0] = identities[k + 1]; m[1] = identities[k + 2]; m[2] = identities[k + 3]; m[3] = identities[k + 4];
        }
        if (n * 5 != identities.size()) return fail("resident slots were not allocated as one range");
        HIP_OK(hipMemsetAsync(t.d_state.as<uint32_t>() + (size_t)lo * 36, 0, n * 144, t.stream));
        HIP_OK(hipMemsetAsync(t.d_checkpoint.as<uint32_t>() + (size_t)lo * 12, 0, n * 48, t.stream));
        HIP_OK(hipMemsetAsync(t.d_corners.as<uint32_t>() + lo, 0, n * 4, t.stream));
        if (staged_upload(t, t.d_meta.as<uint32_t>() + (size_t)lo * 12, meta.data(), n * 48) != 0) return -1;
        // Record 7's x: each allocation's incarnation (resident.rs:457-462),
        // in allocation order, which the view tier's exchange admits by.
        std::vector<uint32_t> seeds;
        seeds.reserve(n * 2);
        for (size_t k = 0; k < identities.size(); k += 5) {
            seeds.push_back(identities[k] * 36u + 28u);
            seeds.push_back(++t.incarnation);
        }
        HIP_OK(t.d_seed_writes.reserve(seeds.size() * 4));
        if (staged_upload(t, t.d_seed_writes.ptr, seeds.data(), seeds.size() * 4) != 0) return -1;
        vgi::launch_scatter_words(t.d_seed_writes.as<uint2>(), (uint32_t)(seeds.size() / 2), t.d_state.as<uint32_t>(), t.stream);
        HIP_OK(hipGetLastError());
    }
    (void)first_new;
    if (!meta_writes.empty()) {
        HIP_OK(t.d_meta_writes.reserve(meta_writes.size() * 4));
        if (staged_upload(t, t.d_meta_writes.ptr, meta_writes.data(), meta_writes.size() * 4) != 0) return -1;
        vgi::launch_scatter_words(t.d_meta_writes.as<uint2>(), (uint32_t)(meta_writes.size() / 2), t.d_meta.as<uint32_t>(), t.stream);
        HIP_OK(hipGetLastError());
    }
    if (t.face_slots.dirty) {
        size_t n = t.face_slots.entries.size();
        HIP_OK(t.d_index.reserve(n * 16));
        HIP_OK(hipMemcpyAsync(t.d_index.ptr, t.face_slots.entries.data(), n * 16, hipMemcpyHostToDevice, t.stream));
        HIP_OK(hipStreamSynchronize(t.stream));
        t.face_slots.dirty = false;
    } else if (!t.face_slots.touched.empty()) {
        // Only the written entries travel (4 words each).
        std::vector<uint32_t> &tw = t.face_slots.touched;
        std::sort(tw.begin(), tw.end());
        tw.erase(std::unique(tw.begin(), tw.end()), tw.end());
        std::vector<uint32_t> writes;
        writes.reserve(tw.size() * 8);
        for (uint32_t i : tw)
            for (int k = 0; k < 4; ++k) { writes.push_back(i * 4u + (uint32_t)k); writes.push_back(t.face_slots.entries[i][k]); }
        HIP_OK(t.d_meta_writes.reserve(writes.size() * 4));
        if (staged_upload(t, t.d_meta_writes.ptr, writes.data(), writes.size() * 4) != 0) return -1;
        vgi::launch_scatter_words(t.d_meta_writes.as<uint2>(), (uint32_t)(writes.size() / 2), t.d_index.as<uint32_t>(), t.stream);
        HIP_OK(hipGetLastError());
    }
    t.face_slots.touched.clear();
    return 0;
}

uint32_t clock_ms(const vgi_tracer &t) {
    if (t.fixed_clock >= 0) return (uint32_t)t.fixed_clock;
    return (uint32_t)std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - t.origin).count();
}

vgi::AccumulateParams accumulate_params(const vgi_tracer &t, uint32_t now) {
    vgi::AccumulateParams p;
    uint32_t limit = t.transport.history_limit;
    p.history_limit = limit == 0 ? 65536u : std::min(limit, 65536u);
    p.samples = t.transport.samples;
    p.clock_ms = now;
    p.recency = t.transport.history_recency >= 1.0f ? t.transport.history_recency : 1.0f;
    p.physical = std::max(t.transport.measured_retention, 0.0f);
    p.response = 0.075f;  // settings.rs:266 display_response 75 ms
    p.revision = t.schedule.transport_revision;
    p.stable_error = t.schedule.stable_error;
    p.stable_samples = t.schedule.stable_samples;
    p.picked = 0u;
    return p;
}

// dispatch.rs:126-131: batches that trace faces advance the batch counter
// (28 bits, never 0); GPU picks leave a host job's face alone for four.
uint32_t next_batch(vgi_tracer &t) {
    t.batch_id = (t.batch_id % 0x0fffffffu) + 1u;
    return t.batch_id;
}

vgi::MlpScene mlp_scene(const vgi_tracer &t) {
    vgi::MlpScene sc;
    sc.lo_x = (float)(t.bmin.x * 16);
    sc.lo_y = (float)(t.bmin.y * 16);
    sc.lo_z = (float)(t.bmin.z * 16);
    int extent = std::max(std::max(t.bmax.x - t.bmin.x + 1, t.bmax.y - t.bmin.y + 1), t.bmax.z - t.bmin.z + 1) * 16;
    sc.inv_extent = 1.0f / (float)std::max(extent, 1);
    sc.encoding = t.mlp_encoding;
    sc.materials = t.d_materials.as<float4>();
    sc.table = t.d_table.as<float2>();
    for (int l = 0; l < vgi::MLP_HASH_LEVELS; ++l) sc.res[l] = vgi::mlp_hash_res(l, std::max(extent, 1));
    return sc;
}

// One optimizer step on `count` samples (a multiple of 16) of queries/targets
// already on the device: fused forward/backward, weight-gradient GEMM, and
// with `update` an Adam step and a re-pack of the BF16 fragments.
int mlp_step(vgi_tracer &t, const vgi::MlpQuery *queries, const float4 *targets, uint32_t count, bool update,
             hipEvent_t after_gradients = nullptr) {
    size_t n = count;
    HIP_OK(t.d_acts.reserve(n * 288 * 2));
    HIP_OK(t.d_grads.reserve(n * 272 * 2));
    HIP_OK(t.d_train_pred.reserve(n * 12));
    HIP_OK(t.d_touched.reserve(n * vgi::MLP_TOUCHED * 4));
    HIP_OK(hipMemsetAsync(t.d_mlp_loss.ptr, 0, 4, t.stream));
    vgi::mlp_forward_backward(t.d_mlp_frag.as<uint16_t>(), mlp_scene(t), queries, reinterpret_cast<const float *>(targets), count,
                              t.d_acts.as<uint16_t>(), t.d_grads.as<uint16_t>(), t.d_mlp_loss.as<float>(), t.d_train_pred.as<float>(),
                              t.d_table_g.as<float>(), t.d_touched.as<uint32_t>(), t.stream);
    vgi::mlp_weight_grad(t.d_acts.as<uint16_t>(), t.d_grads.as<uint16_t>(), count, t.d_mlp_g.as<float>(), t.stream);
    if (after_gradients) HIP_OK(hipEventRecord(after_gradients, t.stream));
    if (update) {
        t.mlp_step += 1;
        vgi::mlp_adam(t.d_mlp_w.as<float>(), t.d_mlp_m.as<float>(), t.d_mlp_v.as<float>(), t.d_mlp_g.as<float>(), t.mlp_step,
                      t.mlp_lr, t.stream);
        if (t.mlp_encoding == vgi::MLP_ENCODING_HASH)
            vgi::mlp_table_adam(t.d_table.as<float>(), t.d_table_m.as<float>(), t.d_table_v.as<float>(), t.d_table_g.as<float>(),
                                t.d_touched.as<uint32_t>(), count, t.mlp_step, t.table_lr, t.stream);
        vgi::mlp_pack(t.d_mlp_w.as<float>(), t.d_mlp_frag.as<uint16_t>(), t.stream);
    }
    HIP_OK(hipGetLastError());
    return 0;
}

// The training set of a frame, on `stream`: a window of the face permutation
// traced with plain multi-bounce paths (one sample each) into targets. It
// needs only the scene and the bound tasks, not the network, so a learned
// frame traces it beside its own paths.
int trace_bound_faces(vgi_tracer &t, const vgi::SceneView &s, const vgi::Optics &o, const uint4 *tasks, uint32_t count,
                      uint32_t samples, uint32_t lanes, uint32_t sequence_base, float4 *results, uint32_t *hits,
                      unsigned long long *counters, bool full_counters, hipStream_t stream, const uint32_t *order,
                      float4 *learned = nullptr, vgi::MlpQuery *queries = nullptr);

int learned_targets(vgi_tracer &t, hipStream_t stream) {
    uint32_t total = (t.train_per_frame / 16u) * 16u;
    if (total == 0 || t.bound == 0) return 0;
    HIP_OK(t.d_train_tasks.reserve((size_t)total * 32));
    HIP_OK(t.d_train_q.reserve((size_t)total * 16));
    HIP_OK(t.d_train_results.reserve((size_t)total * 96));
    HIP_OK(t.d_targets.reserve((size_t)total * 16));
    vgi::launch_train_gather(t.b_tasks.as<uint4>(), t.d_perm.as<uint32_t>(), t.d_face_material.as<uint32_t>(), t.bound,
                             t.train_cursor, total, t.d_train_tasks.as<uint4>(), t.d_train_q.as<vgi::MlpQuery>(), stream);
    vgi::SceneView s;
    vgi::Optics o;
    views(t, s, o);
    o.cache_on = 0u;
    // The prediction replaces what a path gathers after its first hit beyond
    // that hit's direct sun, with one bounce of the budget spent: targets are
    // traced one bounce short.
    o.bounces = o.bounces > 1u ? o.bounces - 1u : 1u;
    // A target's paths on separate lanes (up to 16, a power of two dividing
    // its count): a training window is too small to fill the chip one face
    // per lane.
    uint32_t lanes = 1u;
    while (lanes < 16u && t.train_samples % (lanes * 2u) == 0u) lanes *= 2u;
    if (trace_bound_faces(t, s, o, t.d_train_tasks.as<uint4>(), total, t.train_samples, lanes, 0x40000000u + t.frame_index,
                          t.d_train_results.as<float4>(), nullptr, nullptr, false, stream, nullptr) != 0)
        return -1;
    vgi::launch_train_targets(t.d_train_results.as<float4>(), total, t.d_targets.as<float4>(), stream);
    HIP_OK(hipGetLastError());
    return 0;
}

// The frame's Adam steps on its training set (on the tracer's stream), each
// bracketed by events (timed = true: the learned frame's MFMA fraction).
int learned_steps(vgi_tracer &t, bool timed = false) {
    uint32_t total = (t.train_per_frame / 16u) * 16u;
    if (total == 0 || t.bound == 0) return 0;
    uint32_t per = std::max<uint32_t>(16u, (total / std::max<uint32_t>(t.train_steps, 1u) / 16u) * 16u);
    uint32_t steps = total / per;
    if (timed && t.mlp_step_ev.size() < (size_t)steps * 3) {
        size_t have = t.mlp_step_ev.size();
        t.mlp_step_ev.resize((size_t)steps * 3, nullptr);
        for (size_t k = have; k < t.mlp_step_ev.size(); ++k) HIP_OK(hipEventCreate(&t.mlp_step_ev[k]));
    }
    uint32_t k = 0;
    for (uint32_t off = 0; off + per <= total; off += per, ++k) {
        if (timed) HIP_OK(hipEventRecord(t.mlp_step_ev[k * 3], t.stream));
        if (mlp_step(t, t.d_train_q.as<vgi::MlpQuery>() + off, t.d_targets.as<float4>() + off, per, true, timed ? t.mlp_step_ev[k * 3 + 1] : nullptr) != 0)
            return -1;
        if (timed) HIP_OK(hipEventRecord(t.mlp_step_ev[k * 3 + 2], t.stream));
    }
    if (timed) {
        t.mlp_timed_steps = steps;
        t.mlp_timed_step_samples = per;
    }
    t.train_cursor = (uint32_t)(((uint64_t)t.train_cursor + total) % t.bound);
    return 0;
}

int learned_training(vgi_tracer &t) {
    if (learned_targets(t, t.stream) != 0) return -1;
    return learned_steps(t);
}

// Threads take the bound faces in Morton order of their voxels: each
// wavefront's faces -- and the scene their rays start in -- stay together
// in the caches. Records still go to each face's own index.
uint64_t spread_bits(uint32_t v) {
    uint64_t x = v & 0x1fffffu;
    x = (x | (x << 32)) & 0x1f00000000ffffull;
    x = (x | (x << 16)) & 0x1f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

// Morton keys of the tasks' voxels (relative to their minimum corner).
std::vector<uint64_t> morton_keys(const vgi_face_task *tasks, uint32_t count) {
    int32_t lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX};
    for (uint32_t i = 0; i < count; ++i)
        for (int a = 0; a < 3; ++a) lo[a] = std::min(lo[a], tasks[i].voxel[a]);
    std::vector<uint64_t> keys(count);
    for (uint32_t i = 0; i < count; ++i) {
        uint32_t c[3];
        for (int a = 0; a < 3; ++a) c[a] = (uint32_t)std::min<int64_t>((int64_t)tasks[i].voxel[a] - lo[a], 0x1fffff);
        keys[i] = spread_bits(c[0]) | (spread_bits(c[1]) << 1) | (spread_bits(c[2]) << 2);
    }
    return keys;
}

int compute_order(vgi_tracer &t, const vgi_face_task *tasks, uint32_t count) {
    t.has_order = false;
    if (count == 0) return 0;
    std::vector<uint64_t> morton = morton_keys(tasks, count);
    std::vector<std::pair<uint64_t, uint32_t>> keys(count);
    for (uint32_t i = 0; i < count; ++i) keys[i] = {morton[i], i};
    std::sort(keys.begin(), keys.end());
    std::vector<uint32_t> order(count);
    for (uint32_t i = 0; i < count; ++i) order[i] = keys[i].second;
    HIP_OK(t.d_order.reserve((size_t)count * 4));
    HIP_OK(hipMemcpyAsync(t.d_order.ptr, order.data(), (size_t)count * 4, hipMemcpyHostToDevice, t.stream));
    HIP_OK(hipStreamSynchronize(t.stream));
    t.has_order = true;
    return 0;
}

const uint32_t *order_of(const vgi_tracer &t) { return t.has_order ? t.d_order.as<uint32_t>() : nullptr; }

// The first `count` rows (of `row_bytes`) of a per-bound-face buffer of the
// last frame, for the faces bound now: while no edit has moved the bound
// set since that frame the rows are read as they are; otherwise the moves
// are replayed over the frame's row numbers, and a face that joined since
// reads zeros.
int read_bound_rows(const vgi_tracer &t, const void *rows, size_t row_bytes, uint32_t count, void *out) {
    if (t.bound_moves.empty()) {
        HIP_OK(hipMemcpy(out, rows, (size_t)count * row_bytes, hipMemcpyDeviceToHost));
        return 0;
    }
    std::vector<uint32_t> row_of(t.frame_bound);
    for (uint32_t i = 0; i < t.frame_bound; ++i) row_of[i] = i;
    for (const auto &m : t.bound_moves) {
        if (m.first == vgi_host::NONE) {
            row_of.push_back(vgi_host::NONE);
            continue;
        }
        if (m.second >= row_of.size() || m.first >= row_of.size()) return fail("bound-set moves out of range");
        row_of[m.first] = row_of[m.second];
        row_of.pop_back();
    }
    if (row_of.size() < count) return fail("bound-set moves disagree with the bound count");
    std::vector<uint8_t> frame((size_t)t.frame_bound * row_bytes);
    if (!frame.empty()) HIP_OK(hipMemcpy(frame.data(), rows, frame.size(), hipMemcpyDeviceToHost));
    uint8_t *dst = static_cast<uint8_t *>(out);
    for (uint32_t i = 0; i < count; ++i) {
        if (row_of[i] == vgi_host::NONE) std::memset(dst + (size_t)i * row_bytes, 0, row_bytes);
        else std::memcpy(dst + (size_t)i * row_bytes, frame.data() + (size_t)row_of[i] * row_bytes, row_bytes);
    }
    return 0;
}

// Bound faces of a frame (learned frames: with `learned` and `queries`).
int trace_bound_faces(vgi_tracer &t, const vgi::SceneView &s, const vgi::Optics &o, const uint4 *tasks, uint32_t count,
                      uint32_t samples, uint32_t lanes, uint32_t sequence_base, float4 *results, uint32_t *hits,
                      unsigned long long *counters, bool full_counters, hipStream_t stream, const uint32_t *order,
                      float4 *learned, vgi::MlpQuery *queries) {
    (void)t;
    if (learned) {
        vgi::launch_trace_learned(s, o, tasks, count, samples, lanes, sequence_base, results, counters, learned, queries, stream, order);
    } else {
        vgi::launch_trace(s, o, tasks, count, samples, lanes, sequence_base, results, hits, counters, full_counters, stream, order);
    }
    HIP_OK(hipGetLastError());
    return 0;
}

// A frame of bound faces by wavefront path tracing (same records as the
// megakernel): path state, ray queue and hit buffers sized to the frame.
int wavefront_trace(vgi_tracer &t, const vgi::SceneView &s, const vgi::Optics &o, const uint4 *tasks, uint32_t count,
                    uint32_t sequence_base, float4 *results, unsigned long long *rays) {
    uint32_t samples = t.transport.samples;
    size_t paths = (size_t)count * samples;
    if (paths == 0) return 0;
    if (paths * 8 >= (1ull << 32)) return fail("too many paths for one wavefront frame");
    for (auto &f : t.wf_fields) HIP_OK(f.reserve(paths * 16));
    HIP_OK(t.wf_queue.reserve(paths * (2 + o.light_count) * sizeof(vgi::WfRay)));
    HIP_OK(t.wf_hit0.reserve(paths * sizeof(vgi::WfHit)));
    HIP_OK(t.wf_hit1.reserve(paths * sizeof(vgi::WfHit)));
    HIP_OK(t.wf_vis.reserve(paths * 8));
    HIP_OK(t.wf_control.reserve(16));
    vgi::WfBuffers b;
    float4 **f[vgi::WF_PATH_FIELDS] = {&b.paths.throughput, &b.paths.energy, &b.paths.origin, &b.paths.direction, &b.paths.previous,
                                       &b.paths.nee0, &b.paths.pending_t, &b.paths.pending_e, &b.paths.pending_d, &b.paths.hit,
                                       &b.paths.candidate};
    for (uint32_t k = 0; k < vgi::WF_PATH_FIELDS; ++k) *f[k] = t.wf_fields[k].as<float4>();
    b.queue = t.wf_queue.as<vgi::WfRay>();
    b.hit0 = t.wf_hit0.as<vgi::WfHit>();
    b.hit1 = t.wf_hit1.as<vgi::WfHit>();
    b.vis = t.wf_vis.as<uint8_t>();
    b.control = t.wf_control.as<uint32_t>();
    b.order = order_of(t);
    uint32_t grid = (uint32_t)std::max(t.cus, 1) * (4u * VGI_TRACE_WAVES) / 4u;
    vgi::launch_wavefront(s, o, tasks, count, samples, lanes_for(t.transport), sequence_base, b, grid, results, rays, t.stream, t.dedup);
    HIP_OK(hipGetLastError());
    return 0;
}

// The queued retention events of edits, on the GPU before the batch's
// prepare (resident.rs:1005-1061): measure the edits' flux from the resident
// statistics, then retention over every slot.
int flush_events(vgi_tracer &t) {
    if (t.pending_events.empty() || !t.resident) {
        t.last_event_words = 0;  // this batch applied no retention event
        return 0;
    }
    // Change footprints reserve their Light events before the encoding
    // (gpu/resident.rs:606-626).
    std::vector<vgi_host::Footprint> footprints;
    if (t.footprints) footprints = vgi_host::plan_footprints(t.pending_events);
    vgi_host::EncodedEvents enc;
    vgi_host::encode_events(t.pending_events, enc);
    std::vector<int32_t> footprint_words;
    uint32_t footprint_threads = vgi_host::encode_footprints(footprints, t.pending_events, footprint_words);
    t.pending_events.clear();
    if (t.material_light_dirty) {
        // resident.rs:815-838: luminance of each material's albedo and emission.
        std::vector<float> light(512);
        for (int m = 0; m < 256; ++m) {
            const float *d = &t.diffuse[m * 3], *e = &t.emission[m * 3];
            light[m * 2] = 0.2126f * d[0] + 0.7152f * d[1] + 0.0722f * d[2];
            light[m * 2 + 1] = 0.2126f * e[0] + 0.7152f * e[1] + 0.0722f * e[2];
        }
        HIP_OK(t.d_material_light.reserve(light.size() * 4));
        HIP_OK(hipMemcpyAsync(t.d_material_light.ptr, light.data(), light.size() * 4, hipMemcpyHostToDevice, t.stream));
        t.material_light_dirty = false;
    }
    HIP_OK(t.d_events.reserve(enc.words.size() * 4));
    HIP_OK(hipMemcpyAsync(t.d_events.ptr, enc.words.data(), enc.words.size() * 4, hipMemcpyHostToDevice, t.stream));
    if (!footprints.empty()) {
        // Traced in the new scene, ahead of the measurement and the retention
        // that apply the events (gpu.rs:864-874, dispatch_footprints).
        std::vector<uint32_t> sums(footprints.size() * 8);
        for (size_t k = 0; k < footprints.size(); ++k) {
            uint32_t init[8] = {vgi::NONE, vgi::NONE, vgi::NONE, 0u, 0u, 0u, 0u, 0u};
            std::memcpy(&sums[k * 8], init, sizeof(init));
        }
        HIP_OK(t.d_footprint_tasks.reserve(footprint_words.size() * 4));
        HIP_OK(t.d_footprint_sums.reserve(sums.size() * 4));
        HIP_OK(hipMemcpyAsync(t.d_footprint_tasks.ptr, footprint_words.data(), footprint_words.size() * 4, hipMemcpyHostToDevice,
                              t.stream));
        HIP_OK(hipMemcpyAsync(t.d_footprint_sums.ptr, sums.data(), sums.size() * 4, hipMemcpyHostToDevice, t.stream));
        vgi::SceneView s;
        vgi::Optics o;
        views(t, s, o);
        vgi::launch_footprints(s, o, t.d_footprint_tasks.as<int4>(), (uint32_t)footprints.size(), footprint_threads,
                               t.d_footprint_sums.as<uint32_t>(), t.d_events.as<uint4>(), t.stream);
        HIP_OK(hipGetLastError());
    }
    if (enc.entries) {
        HIP_OK(t.d_measures.reserve(enc.measures.size() * 4));
        HIP_OK(t.d_measure_max.reserve((size_t)enc.entries * 4));
        HIP_OK(hipMemcpyAsync(t.d_measures.ptr, enc.measures.data(), enc.measures.size() * 4, hipMemcpyHostToDevice, t.stream));
        HIP_OK(hipMemsetAsync(t.d_measure_max.ptr, 0, (size_t)enc.entries * 4, t.stream));
        vgi::launch_measure_edits(t.d_measures.as<uint4>(), enc.entries, enc.positions, t.d_index.as<uint4>(),
                                  (uint32_t)(t.face_slots.entries.size() - 1), t.d_state.as<uint32_t>(), t.d_meta.as<uint32_t>(),
                                  t.d_material_light.as<float2>(), t.d_measure_max.as<uint32_t>(), t.d_events.as<uint4>(), t.stream);
        HIP_OK(hipGetLastError());
    }
    vgi::launch_apply_events(t.d_events.as<uint4>(), enc.events, t.high_water, t.d_state.as<uint32_t>(), t.d_checkpoint.as<uint32_t>(),
                             t.d_meta.as<uint32_t>(), std::max(t.transport.measured_retention, 0.0f), t.stream);
    HIP_OK(hipGetLastError());
    t.last_event_words = (uint32_t)enc.words.size();
    return 0;
}

// Each bound face's material (the learned cache encodes its albedo) and a
// fixed random order for choosing training faces, sorted in Morton order
// within runs of TRAIN_RUN faces: a frame's training window is a random
// subset whose paths still start near their neighbours on a wavefront.
constexpr uint32_t TRAIN_RUN = 8192;
int bind_materials(vgi_tracer &t, const vgi_face_task *tasks, uint32_t count) {
    vgi_host::VoxelSource src{t.grid.data(), t.vmin, t.gx, t.gy, t.gz, t.loaded.data(), t.rmin, t.rsx, t.rsy, t.rsz};
    t.bound_material.assign(count, 0u);
    for (uint32_t i = 0; i < count; ++i) {
        uint32_t m = t.have_regions ? src.material_at(tasks[i].voxel[0], tasks[i].voxel[1], tasks[i].voxel[2]) : 0xffffu;
        t.bound_material[i] = m < vgi_host::AIR ? m : 0u;
    }
    std::vector<uint32_t> perm(count);
    std::iota(perm.begin(), perm.end(), 0u);
    std::mt19937 rng(0x5eedu);
    std::shuffle(perm.begin(), perm.end(), rng);
    std::vector<uint64_t> morton = morton_keys(tasks, count);
    for (uint32_t b = 0; b < count; b += TRAIN_RUN)
        std::sort(perm.begin() + b, perm.begin() + std::min(count, b + TRAIN_RUN),
                  [&](uint32_t x, uint32_t y) { return morton[x] < morton[y]; });
    HIP_OK(t.d_face_material.reserve((size_t)std::max<uint32_t>(count, 1) * 4));
    HIP_OK(t.d_perm.reserve((size_t)std::max<uint32_t>(count, 1) * 4));
    HIP_OK(hipMemcpyAsync(t.d_face_material.ptr, t.bound_material.data(), (size_t)count * 4, hipMemcpyHostToDevice, t.stream));
    HIP_OK(hipMemcpyAsync(t.d_perm.ptr, perm.data(), (size_t)count * 4, hipMemcpyHostToDevice, t.stream));
    HIP_OK(hipStreamSynchronize(t.stream));
    t.train_cursor = 0;
    return 0;
}

// An exposed face: a solid voxel whose neighbour across the face is loaded air.
const int FACE_STEP[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
bool face_exposed(const vgi_host::VoxelSource &src, int x, int y, int z, uint32_t f) {
    if (src.material_at(x, y, z) >= vgi_host::AIR) return false;
    return src.material_at(x + FACE_STEP[f][0], y + FACE_STEP[f][1], z + FACE_STEP[f][2]) == vgi_host::AIR;
}

// Bound faces follow the exposed surface after an edit of the `changed`
// voxels: a face that lost its air neighbour leaves (the last bound face
// takes its place), a newly exposed face joins at the end (publication's
// rescan of the edited pages, pages.rs:451-458).
int update_exposure(vgi_tracer &t, const std::vector<vgi_host::Int3> &changed, const std::vector<vgi_host::RegionDelta> &deltas) {
    // The bound set equals the exposed set before the edit, so only faces
    // whose exposure flips (or whose voxel changed material) need work, and
    // they are found from the voxels' values before and after: a changed
    // voxel's own six faces, and its neighbours' faces towards it unless the
    // neighbour changed too (then they are its own faces) -- each face once.
    std::vector<int32_t> delta_of((size_t)t.rsx * t.rsy * t.rsz, -1);
    for (size_t k = 0; k < deltas.size(); ++k) {
        const Int3 &c = deltas[k].coord;
        delta_of[lin3(c.x - t.rmin.x, c.y - t.rmin.y, c.z - t.rmin.z, t.rsx, t.rsy)] = (int32_t)k;
    }
    // Material before and after (0xffff outside the loaded scene), and whether it changed.
    auto lookup = [&](int x, int y, int z, uint32_t &before, uint32_t &after) {
        int rx = floor_div(x, 16) - t.rmin.x, ry = floor_div(y, 16) - t.rmin.y, rz = floor_div(z, 16) - t.rmin.z;
        if (rx < 0 || ry < 0 || rz < 0 || rx >= t.rsx || ry >= t.rsy || rz >= t.rsz) { before = after = 0xffffu; return false; }
        size_t ri = lin3(rx, ry, rz, t.rsx, t.rsy);
        if (!t.loaded[ri]) { before = after = 0xffffu; return false; }
        after = t.grid[lin3(x - t.vmin.x, y - t.vmin.y, z - t.vmin.z, t.gx, t.gy)];
        int32_t k = delta_of[ri];
        if (k < 0) { before = after; return false; }
        before = deltas[k].before[(size_t)floor_mod(x, 16) + 16 * ((size_t)floor_mod(y, 16) + 16 * (size_t)floor_mod(z, 16))];
        return before != after;
    };
    // The visits only read the scene and the face index: contiguous ranges of
    // the changed voxels run on the pool, each with its own lists, joined in
    // range order.
    struct Found {
        std::vector<uint32_t> removals;
        std::vector<vgi_face_task> additions, refresh;
    };
    auto visit = [&](Found &out, int x, int y, int z, uint32_t f, bool own_changed) {
        std::vector<uint32_t> &removals = out.removals;
        std::vector<vgi_face_task> &additions = out.additions, &refresh = out.refresh;
        uint32_t vb, va, nb, na;
        lookup(x, y, z, vb, va);
        lookup(x + FACE_STEP[f][0], y + FACE_STEP[f][1], z + FACE_STEP[f][2], nb, na);
        bool was = vb < vgi_host::AIR && nb == vgi_host::AIR, now = va < vgi_host::AIR && na == vgi_host::AIR;
        if (!was && !now) return;
        vgi_face_task task{};
        task.voxel[0] = x; task.voxel[1] = y; task.voxel[2] = z;
        task.face = f;
        if (now && !was) { additions.push_back(task); return; }
        if (now && was) {
            if (own_changed) refresh.push_back(task);  // stays bound; its material changed
            return;
        }
        long at = t.face_slots.find(x, y, z, f);
        uint32_t slot = at >= 0 ? (t.face_slots.entries[at][3] & 0x1fffffffu) : vgi_host::NONE;
        uint32_t b = (slot != vgi_host::NONE && slot < t.bound_of_slot.size()) ? t.bound_of_slot[slot] : vgi_host::NONE;
        if (b != vgi_host::NONE) removals.push_back(b);
    };
    const size_t per = 2048;
    std::vector<Found> found((changed.size() + per - 1) / per);
    edit_pool().run(found.size(), [&](size_t begin, size_t end) {
        for (size_t chunk = begin; chunk < end; ++chunk)
            for (size_t i = chunk * per; i < std::min(changed.size(), chunk * per + per); ++i) {
                const Int3 &c = changed[i];
                for (uint32_t f = 0; f < 6; ++f) {
                    visit(found[chunk], c.x, c.y, c.z, f, true);
                    int nx = c.x + FACE_STEP[f][0], ny = c.y + FACE_STEP[f][1], nz = c.z + FACE_STEP[f][2];
                    uint32_t b, a;
                    if (!lookup(nx, ny, nz, b, a)) visit(found[chunk], nx, ny, nz, f ^ 1u, false);
                }
            }
    }, 1);
    std::vector<uint32_t> removals;
    std::vector<vgi_face_task> additions, refresh;
    for (const Found &fnd : found) {
        removals.insert(removals.end(), fnd.removals.begin(), fnd.removals.end());
        additions.insert(additions.end(), fnd.additions.begin(), fnd.additions.end());
        refresh.insert(refresh.end(), fnd.refresh.begin(), fnd.refresh.end());
    }
    std::vector<uint32_t> refreshed;
    if (register_faces(t, refresh.data(), (uint32_t)refresh.size(), refreshed) != 0) return -1;
    // Removals, highest index first, each filled by the current last face.
    std::sort(removals.rbegin(), removals.rend());
    std::vector<uint32_t> dirty_index;
    for (uint32_t b : removals) {
        uint32_t last = (uint32_t)t.bound_tasks.size() - 1;
        t.bound_moves.push_back({b, last});
        t.bound_of_slot[t.bound_slots[b]] = vgi_host::NONE;
        if (b != last) {
            t.bound_tasks[b] = t.bound_tasks[last];
            t.bound_slots[b] = t.bound_slots[last];
            t.bound_of_slot[t.bound_slots[b]] = b;
            dirty_index.push_back(b);
        }
        t.bound_tasks.pop_back();
        t.bound_slots.pop_back();
    }
    std::vector<uint32_t> new_slots;
    if (register_faces(t, additions.data(), (uint32_t)additions.size(), new_slots) != 0) return -1;
    for (size_t k = 0; k < additions.size(); ++k) {
        uint32_t b = (uint32_t)t.bound_tasks.size();
        t.bound_moves.push_back({vgi_host::NONE, 0u});
        t.bound_tasks.push_back(additions[k]);
        t.bound_slots.push_back(new_slots[k]);
        if (t.bound_of_slot.size() <= new_slots[k]) t.bound_of_slot.resize((size_t)new_slots[k] + 1, vgi_host::NONE);
        t.bound_of_slot[new_slots[k]] = b;
        dirty_index.push_back(b);
    }
    uint32_t count = (uint32_t)t.bound_tasks.size();
    std::sort(dirty_index.begin(), dirty_index.end());
    dirty_index.erase(std::unique(dirty_index.begin(), dirty_index.end()), dirty_index.end());
    dirty_index.erase(std::remove_if(dirty_index.begin(), dirty_index.end(), [&](uint32_t i) { return i >= count; }), dirty_index.end());
    size_t need_tasks = (size_t)std::max<uint32_t>(count, 1) * 32;
    size_t need_slots = (size_t)std::max<uint32_t>(count, 1) * 4;
    bool whole = need_tasks > t.b_tasks.bytes || need_slots > t.b_slots.bytes || t.b_slots.ptr == nullptr;
    if (whole) {
        // Grown past the allocation: upload the whole set again (after the
        // stream has passed the frame that may still read the old set).
        HIP_OK(hipStreamSynchronize(t.stream));
        HIP_OK(t.b_tasks.reserve(need_tasks));
        HIP_OK(t.b_slots.reserve((size_t)std::max<uint32_t>(count, 1) * 4));
        HIP_OK(hipMemcpyAsync(t.b_tasks.ptr, t.bound_tasks.data(), (size_t)count * 32, hipMemcpyHostToDevice, t.stream));
        HIP_OK(hipMemcpyAsync(t.b_slots.ptr, t.bound_slots.data(), (size_t)count * 4, hipMemcpyHostToDevice, t.stream));
    } else if (!dirty_index.empty()) {
        std::vector<vgi_face_task> up(dirty_index.size());
        std::vector<uint32_t> up_slots(dirty_index.size());
        for (size_t k = 0; k < dirty_index.size(); ++k) { up[k] = t.bound_tasks[dirty_index[k]]; up_slots[k] = t.bound_slots[dirty_index[k]]; }
        HIP_OK(t.d_bound_index.reserve(dirty_index.size() * 4));
        HIP_OK(t.d_bound_upd.reserve(up.size() * 32));
        HIP_OK(t.d_bound_slot_upd.reserve(up_slots.size() * 4));
        if (staged_upload(t, t.d_bound_index.ptr, dirty_index.data(), dirty_index.size() * 4) != 0 ||
            staged_upload(t, t.d_bound_upd.ptr, up.data(), up.size() * 32) != 0 ||
            staged_upload(t, t.d_bound_slot_upd.ptr, up_slots.data(), up_slots.size() * 4) != 0)
            return -1;
        vgi::launch_scatter_bound(t.d_bound_index.as<uint32_t>(), t.d_bound_upd.as<uint4>(), t.d_bound_slot_upd.as<uint32_t>(),
                                  (uint32_t)dirty_index.size(), t.b_tasks.as<uint4>(), t.b_slots.as<uint32_t>(), t.stream);
        HIP_OK(hipGetLastError());
    }
    HIP_OK(t.b_results.grow_keep((size_t)std::max<uint32_t>(count, 1) * 96, (size_t)t.frame_bound * 96, t.stream));
    // The whole set travels from pageable memory: wait for it.
    if (whole) HIP_OK(hipStreamSynchronize(t.stream));
    t.bound = count;
    t.has_order = false;  // the bound set changed: faces are taken in bound order
    if (t.mlp_ready && bind_materials(t, t.bound_tasks.data(), count) != 0) return -1;
    return 0;
}

}  // namespace

// ================================================================ C ABI
extern "C" {

const char *vgi_last_error(void) { return g_error.c_str(); }

int vgi_create(int device_ordinal, vgi_tracer **out) {
    if (!out) return fail("null output");
    int count = 0;
    HIP_OK(hipGetDeviceCount(&count));
    if (device_ordinal < 0 || device_ordinal >= count) return fail("no such HIP device");
    HIP_OK(hipSetDevice(device_ordinal));
    if (device_ordinal >= MAX_DEVICES) return fail("device ordinal beyond the library's device table");
    {
        // This device's copy of the no-allocation scratch, resolved once
        // with the device current.
        std::lock_guard<std::mutex> lock(g_scratch_mutex);
        if (!g_scratch[device_ordinal]) {
            void *ptr = nullptr;
            HIP_OK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
            if (!ptr) return fail("no device address for the scratch region");
            g_scratch[device_ordinal] = ptr;
        }
    }
    std::unique_ptr<vgi_tracer> t(new vgi_tracer());
    t->device = device_ordinal;
    HIP_OK(hipDeviceGetAttribute(&t->cus, hipDeviceAttributeMultiprocessorCount, device_ordinal));
    HIP_OK(hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking));
    for (auto &e : t->ev) HIP_OK(hipEventCreate(&e));
    for (int m = 0; m < 256; ++m)
        for (int c = 0; c < 3; ++c) { t->diffuse[m * 3 + c] = 0.45f; t->emission[m * 3 + c] = 0.0f; }  // world.rs:72-79
    t->transport.bounces = 3;
    t->transport.samples = 16;
    t->transport.max_distance = 768.0f;
    t->transport.reach = 768.0f;
    t->transport.history_recency = 1.0f;
    t->transport.measured_retention = 1.0f;
    if (vgi_set_materials(t.get(), t->diffuse, t->emission) != 0) return -1;
    if (vgi_set_lights(t.get(), nullptr, 0, 0) != 0) return -1;
    *out = t.release();
    return 0;
}

void vgi_destroy(vgi_tracer *t) {
    if (!t) return;
    hipSetDevice(t->device);
    hipStreamSynchronize(t->stream);
    for (auto &kv : t->tickets) {
        kv.second->results.release(); kv.second->hits.release(); kv.second->counters.release();
        if (kv.second->done) hipEventDestroy(kv.second->done);
    }
    DeviceBuffer *bufs[] = {&t->d_grid, &t->d_headers, &t->d_cells, &t->d_info, &t->d_list, &t->d_top, &t->d_totals,
                            &t->d_scatter_idx, &t->d_scatter_val, &t->d_materials, &t->d_lights, &t->d_emitters,
                            &t->d_index, &t->d_state, &t->d_checkpoint, &t->d_meta, &t->d_tasks, &t->d_slots,
                            &t->b_tasks, &t->b_slots, &t->b_results, &t->b_counters, &t->b_hits, &t->d_mlp_w, &t->d_mlp_m,
                            &t->d_mlp_v, &t->d_mlp_g, &t->d_mlp_frag, &t->d_mlp_loss, &t->d_table,
                            &t->d_table_m, &t->d_table_v, &t->d_table_g, &t->b_learned, &t->b_queries,
                            &t->b_pred, &t->d_perm, &t->d_face_material, &t->d_train_tasks, &t->d_train_q,
                            &t->d_train_results, &t->d_targets, &t->d_acts, &t->d_grads, &t->d_train_pred, &t->d_io, &t->d_touched,
                            &t->d_events, &t->d_measures, &t->d_measure_max, &t->d_material_light, &t->d_bound_index,
                            &t->d_bound_upd, &t->d_bound_slot_upd, &t->d_meta_writes, &t->d_seed_writes, &t->v_depth, &t->v_normals, &t->v_queue,
                            &t->v_receivers[0], &t->v_receivers[1], &t->v_estimates[0], &t->v_estimates[1], &t->v_owners[0],
                            &t->v_owners[1], &t->v_hash[0], &t->v_hash[1], &t->wf_queue, &t->wf_hit0, &t->wf_hit1,
                            &t->wf_vis, &t->wf_control, &t->d_order, &t->d_keys, &t->d_selection, &t->d_pick_tasks, &t->d_pick_slots,
                            &t->d_pick_results, &t->d_mirror, &t->d_select_ranges,
                            &t->d_footprint_tasks, &t->d_footprint_sums, &t->d_corners, &t->d_sun_flux, &t->d_sun_queue,
                            &t->d_sun_ranges, &t->d_sun_region, &t->v_pixels, &t->v_counters, &t->f_cells[0], &t->f_cells[1],
                            &t->f_output, &t->f_queue};
    for (auto &f : t->wf_fields) f.release();
    if (t->side) hipStreamSynchronize(t->side);
    for (DeviceBuffer *b : bufs) b->release();
    for (auto &e : t->ev) hipEventDestroy(e);
    for (auto &e : t->side_ev) if (e) hipEventDestroy(e);
    for (auto &e : t->sel_ev) if (e) hipEventDestroy(e);
    for (auto &e : t->sun_ev) if (e) hipEventDestroy(e);
    for (auto &e : t->mlp_ev) if (e) hipEventDestroy(e);
    for (auto &e : t->mlp_step_ev) if (e) hipEventDestroy(e);
    for (auto &e : t->view_ev) if (e) hipEventDestroy(e);
    if (t->stage_done) hipEventDestroy(t->stage_done);
    if (t->stage) hipHostFree(t->stage);
    if (t->side) hipStreamDestroy(t->side);
    hipStreamDestroy(t->stream);
    delete t;
}

int vgi_set_clock(vgi_tracer *t, int64_t milliseconds) {
    if (!t) return fail("null tracer");
    t->fixed_clock = milliseconds;
    return 0;
}

int vgi_enable_resident(vgi_tracer *t) {
    if (!t) return fail("null tracer");
    t->direct_epoch += 1;
    t->resident = true;
    return 0;
}

int vgi_upload_pages(vgi_tracer *t, const int32_t *page_xyz, uint32_t pages, const uint16_t *voxels, const int32_t *removed_xyz,
                     uint32_t removed) {
    if (!t) return fail("null tracer");
    if (pages && (!page_xyz || !voxels)) return fail("null page data");
    t->direct_epoch += 1;
    HIP_OK(hipSetDevice(t->device));
    std::vector<std::pair<Int3, const uint16_t *>> incoming;
    for (uint32_t i = 0; i < pages; ++i)
        incoming.push_back({Int3{page_xyz[i * 3], page_xyz[i * 3 + 1], page_xyz[i * 3 + 2]}, voxels + (size_t)i * 32768});
    std::vector<Int3> rem;
    for (uint32_t i = 0; i < removed; ++i) rem.push_back(Int3{removed_xyz[i * 3], removed_xyz[i * 3 + 1], removed_xyz[i * 3 + 2]});
    int rc = apply_pages(*t, incoming, rem);
    // Pages change residency, not just voxels: bound faces that follow the
    // exposed surface are found again over the new scene.
    if (rc == 0 && t->follow_exposure) rc = vgi_bind_exposed(t);
    if (rc == 0) t->cells_live = t->cells_used;
    return rc;
}

int vgi_edit_voxels(vgi_tracer *t, const int32_t *xyz, const uint16_t *values, uint32_t count) {
    if (!t) return fail("null tracer");
    if (count && (!xyz || !values)) return fail("null edit data");
    t->direct_epoch += 1;
    HIP_OK(hipSetDevice(t->device));
    return apply_edits(*t, xyz, values, count);
}

int vgi_set_materials(vgi_tracer *t, const float *diffuse_rgb, const float *emission_rgb) {
    if (!t || !diffuse_rgb || !emission_rgb) return fail("null materials");
    t->direct_epoch += 1;
    HIP_OK(hipSetDevice(t->device));
    std::vector<float> packed(256 * 8);
    for (int m = 0; m < 256; ++m) {
        float d[3], e[3];
        for (int c = 0; c < 3; ++c) {
            // world.rs:166-173 (bounded): diffuse to [0, 1], emission to [0, 1e12]; non-finite -> 0.
            float dv = diffuse_rgb[m * 3 + c], ev = emission_rgb[m * 3 + c];
            dv = std::isfinite(dv) ? std::min(std::max(dv, 0.0f), 1.0f) : 0.0f;
            ev = std::isfinite(ev) ? std::min(std::max(ev, 0.0f), 1e12f) : 0.0f;
            d[c] = dv; e[c] = ev;
        }
        float power = std::max(std::max(e[0], e[1]), e[2]);
        float rec[8] = {d[0], d[1], d[2], power, e[0], e[1], e[2], 0.0f};
        std::memcpy(&packed[m * 8], rec, sizeof(rec));
        for (int c = 0; c < 3; ++c) {
            if (t->emission[m * 3 + c] != e[c]) t->emitters_dirty = true;
            t->diffuse[m * 3 + c] = d[c];
            t->emission[m * 3 + c] = e[c];
        }
    }
    t->material_light_dirty = true;
    HIP_OK(t->d_materials.reserve(packed.size() * 4));
    HIP_OK(hipMemcpyAsync(t->d_materials.ptr, packed.data(), packed.size() * 4, hipMemcpyHostToDevice, t->stream));
    HIP_OK(hipStreamSynchronize(t->stream));
    return 0;
}

int vgi_set_lights(vgi_tracer *t, const vgi_light *lights, uint32_t count, uint32_t directional) {
    if (!t) return fail("null tracer");
    if (count && !lights) return fail("null lights");
    t->direct_epoch += 1;
    HIP_OK(hipSetDevice(t->device));
    t->lights.assign((size_t)std::max<uint32_t>(count, 1) * 12, 0.0f);
    for (uint32_t i = 0; i < count; ++i) std::memcpy(&t->lights[i * 12], &lights[i], 48);
    t->light_count = count;
    t->directional = directional;
    HIP_OK(t->d_lights.reserve(t->lights.size() * 4));
    HIP_OK(hipMemcpyAsync(t->d_lights.ptr, t->lights.data(), t->lights.size() * 4, hipMemcpyHostToDevice, t->stream));
    HIP_OK(hipStreamSynchronize(t->stream));
    return 0;
}

int vgi_set_transport(vgi_tracer *t, const vgi_transport *tr) {
    if (!t || !tr) return fail("null transport");
    if (tr->bounces < 1 || tr->bounces > 8) return fail("bounces must be 1..8");
    if (tr->samples < 1) return fail("samples must be >= 1");
    if (tr->lanes != 0 && (tr->lanes > 64 || (64 % tr->lanes) != 0)) return fail("lanes must divide 64");
    t->transport = *tr;
    if (!(t->transport.max_distance > 0.0f)) t->transport.max_distance = 768.0f;
    t->direct_epoch += 1;
    return 0;
}

int vgi_submit_faces(vgi_tracer *t, const vgi_face_task *tasks, uint32_t count, uint32_t flags, uint64_t *ticket) {
    if (!t || !ticket) return fail("null argument");
    if (count && !tasks) return fail("null tasks");
    if (!t->have_regions) return fail("no scene loaded");
    HIP_OK(hipSetDevice(t->device));
    if (sync_emitters(*t) != 0) return -1;
    std::unique_ptr<Ticket> k(new Ticket());
    k->count = count;
    k->samples = t->transport.samples;
    k->with_hits = (flags & VGI_RECORD_HITS) != 0;
    k->with_counters = (flags & VGI_COUNT_RAYS) != 0;
    HIP_OK(hipEventCreate(&k->done));
    HIP_OK(k->results.reserve((size_t)std::max<uint32_t>(count, 1) * 96));
    if (k->with_hits) HIP_OK(k->hits.reserve((size_t)std::max<uint32_t>(count, 1) * k->samples * 32));
    HIP_OK(k->counters.reserve(64));
    HIP_OK(hipMemsetAsync(k->counters.ptr, 0, 64, t->stream));
    HIP_OK(hipMemsetAsync(k->results.ptr, 0, (size_t)std::max<uint32_t>(count, 1) * 96, t->stream));
    HIP_OK(t->d_tasks.reserve((size_t)std::max<uint32_t>(count, 1) * 32));
    HIP_OK(hipMemcpyAsync(t->d_tasks.ptr, tasks, (size_t)count * 32, hipMemcpyHostToDevice, t->stream));
    vgi::SceneView s;
    vgi::Optics o;
    std::vector<uint32_t> slots;
    if (t->resident) {
        if (register_faces(*t, tasks, count, slots) != 0) return -1;
        HIP_OK(t->d_slots.reserve((size_t)std::max<uint32_t>(count, 1) * 4));
        HIP_OK(hipMemcpyAsync(t->d_slots.ptr, slots.data(), (size_t)count * 4, hipMemcpyHostToDevice, t->stream));
        t->last_slots = slots;
    }
    views(*t, s, o);
    uint32_t lanes = lanes_for(t->transport);
    if (flush_events(*t) != 0) return -1;
    HIP_OK(hipEventRecord(t->ev[0], t->stream));
    if (t->resident)
        vgi::launch_prepare(t->d_tasks.as<uint4>(), t->d_slots.as<uint32_t>(), count, t->d_state.as<uint32_t>(), t->d_meta.as<uint32_t>(),
                            std::max(t->transport.measured_retention, 0.0f), next_batch(*t), t->d_keys.as<uint32_t>(), t->stream);
    HIP_OK(hipEventRecord(t->ev[1], t->stream));
    if (trace_bound_faces(*t, s, o, t->d_tasks.as<uint4>(), count, t->transport.samples, lanes, 0u, k->results.as<float4>(),
                          k->with_hits ? k->hits.as<uint32_t>() : nullptr,
                          (k->with_counters || k->with_hits) ? k->counters.as<unsigned long long>() : nullptr, true, t->stream,
                          nullptr) != 0)
        return -1;
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord(t->ev[2], t->stream));
    if (t->resident)
        vgi::launch_accumulate(t->d_slots.as<uint32_t>(), count, k->results.as<float4>(), t->d_state.as<uint32_t>(),
                               t->d_checkpoint.as<uint32_t>(), t->d_meta.as<uint32_t>(), accumulate_params(*t, clock_ms(*t)), t->stream);
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord(t->ev[3], t->stream));
    HIP_OK(hipEventRecord(k->done, t->stream));
    uint64_t id = t->next_ticket++;
    t->tickets[id] = std::move(k);
    *ticket = id;
    return 0;
}

int vgi_trace_faces_noalloc(vgi_tracer *t, const vgi_face_task *tasks, uint32_t count, vgi_face_result *out) {
    if (!t || (count && (!tasks || !out))) return fail("null argument");
    if (count > VGI_NOALLOC_MAX_FACES)
        return fail("vgi_trace_faces_noalloc takes at most " + std::to_string(VGI_NOALLOC_MAX_FACES) + " faces per call (got " +
                    std::to_string(count) + ")");
    if (t->resident) return fail("vgi_trace_faces_noalloc traces without the resident cache (use vgi_submit_faces)");
    if (!t->have_regions) return fail("no scene loaded");
    if (count == 0) return 0;
    HIP_OK(hipSetDevice(t->device));
    void *base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mutex);
        base = g_scratch[t->device];
    }
    if (!base) return fail("the scratch region of this device was not resolved");
    if (sync_emitters(*t, true) != 0) return -1;
    uint8_t *scratch_ptr = static_cast<uint8_t *>(base);
    uint4 *d_tasks = reinterpret_cast<uint4 *>(scratch_ptr + NOALLOC_TASKS);
    float4 *d_records = reinterpret_cast<float4 *>(scratch_ptr + NOALLOC_RECORDS);
    HIP_OK(hipMemcpyAsync(d_tasks, tasks, (size_t)count * 32, hipMemcpyHostToDevice, t->stream));
    // The scratch keeps the last call's records: clear them (faces marked
    // VGI_NO_FACE are not written by the trace).
    vgi::launch_fill_words(reinterpret_cast<uint32_t *>(d_records), (uint64_t)count * 24, 0u, t->stream);
    vgi::SceneView s;
    vgi::Optics o;
    views(*t, s, o);
    vgi::launch_trace(s, o, d_tasks, count, t->transport.samples, lanes_for(t->transport), 0u, d_records, nullptr, nullptr, false,
                      t->stream, nullptr);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(out, d_records, (size_t)count * 96, hipMemcpyDeviceToHost, t->stream));
    HIP_OK(hipStreamSynchronize(t->stream));
    return 0;
}

int vgi_poll(vgi_tracer *t, uint64_t ticket, int *ready) {
    if (!t || !ready) return fail("null argument");
    auto it = t->tickets.find(ticket);
    if (it == t->tickets.end()) return fail("unknown ticket");
    hipError_t e = hipEventQuery(it->second->done);
    if (e == hipSuccess) { *ready = 1; return 0; }
    if (e == hipErrorNotReady) { *ready = 0; return 0; }
    return fail(std::string("hipEventQuery: ") + hipGetErrorString(e));
}

int vgi_read_faces(vgi_tracer *t, uint64_t ticket, vgi_face_result *out, uint32_t count) {
    if (!t || !out) return fail("null argument");
    auto it = t->tickets.find(ticket);
    if (it == t->tickets.end()) return fail("unknown ticket");
    Ticket &k = *it->second;
    if (count > k.count) return fail("more records than the ticket holds");
    HIP_OK(hipSetDevice(t->device));
    HIP_OK(hipEventSynchronize(k.done));
    HIP_OK(hipMemcpy(out, k.results.ptr, (size_t)count * 96, hipMemcpyDeviceToHost));
    hipEventElapsedTime(&t->timings[0], t->ev[0], t->ev[1]);
    hipEventElapsedTime(&t->timings[1], t->ev[1], t->ev[2]);
    hipEventElapsedTime(&t->timings[2], t->ev[2], t->ev[3]);
    // A ticket is complete once its faces are read unless it holds hits/counters.
    if (!k.with_hits && !k.with_counters) {
        k.results.release(); k.hits.release(); k.counters.release();
        hipEventDestroy(k.done);
        t->tickets.erase(it);
    }
    return 0;
}

int vgi_read_hits(vgi_tracer *t, uint64_t ticket, uint32_t *out, uint64_t words) {
    if (!t || !out) return fail("null argument");
    auto it = t->tickets.find(ticket);
    if (it == t->tickets.end()) return fail("unknown ticket");
    Ticket &k = *it->second;
    if (!k.with_hits) return fail("ticket has no hit records (VGI_RECORD_HITS)");
    uint64_t have = (uint64_t)k.count * k.samples * 8;
    if (words > have) return fail("more words than the ticket holds");
    HIP_OK(hipEventSynchronize(k.done));
    HIP_OK(hipMemcpy(out, k.hits.ptr, words * 4, hipMemcpyDeviceToHost));
    return 0;
}

int vgi_read_counters(vgi_tracer *t, uint64_t ticket, uint64_t *out, uint32_t count) {
    if (!t || !out) return fail("null argument");
    auto it = t->tickets.find(ticket);
    if (it == t->tickets.end()) return fail("unknown ticket");
    Ticket &k = *it->second;
    HIP_OK(hipEventSynchronize(k.done));
    uint64_t c[8];
    HIP_OK(hipMemcpy(c, k.counters.ptr, 64, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < count && i < 8; ++i) out[i] = c[i];
    return 0;
}

int vgi_timings(vgi_tracer *t, float *out, uint32_t count) {
    if (!t || !out) return fail("null argument");
    for (uint32_t i = 0; i < count && i < 8; ++i) out[i] = t->timings[i];
    return 0;
}

int vgi_read_slots(vgi_tracer *t, uint32_t *out, uint32_t count) {
    if (!t || !out) return fail("null argument");
    if (count > t->last_slots.size()) return fail("more slots than the last batch");
    std::memcpy(out, t->last_slots.data(), (size_t)count * 4);
    return 0;
}

int vgi_bind_faces(vgi_tracer *t, const vgi_face_task *tasks, uint32_t count) {
    if (!t || (count && !tasks)) return fail("null argument");
    t->direct_epoch += 1;
    HIP_OK(hipSetDevice(t->device));
    HIP_OK(t->b_tasks.reserve((size_t)std::max<uint32_t>(count, 1) * 32));
    HIP_OK(hipMemcpyAsync(t->b_tasks.ptr, tasks, (size_t)count * 32, hipMemcpyHostToDevice, t->stream));
    HIP_OK(t->b_results.reserve((size_t)std::max<uint32_t>(count, 1) * 96));
    HIP_OK(t->b_counters.reserve(64));
    HIP_OK(t->b_slots.reserve((size_t)std::max<uint32_t>(count, 1) * 4));
    t->follow_exposure = false;
    t->bound_broken = false;
    t->bound_moves.clear();
    t->frame_bound = 0;
    t->bound_hit_words = 0;
    t->bound_tasks.clear();
    t->bound_slots.clear();
    t->bound_of_slot.clear();
    if (t->resident) {
        std::vector<uint32_t> slots;
        if (register_faces(*t, tasks, count, slots) != 0) return -1;
        HIP_OK(hipMemcpyAsync(t->b_slots.ptr, slots.data(), (size_t)count * 4, hipMemcpyHostToDevice, t->stream));
        t->last_slots = slots;
    }
    if (bind_materials(*t, tasks, count) != 0) return -1;
    if (compute_order(*t, tasks, count) != 0) return -1;
    t->bound = count;
    return 0;
}

int vgi_bind_exposed(vgi_tracer *t) {
    if (!t) return fail("null tracer");
    if (!t->resident) return fail("bound faces follow edits through the resident cache: call vgi_enable_resident first");
    if (!t->have_regions) return fail("no scene loaded");
    HIP_OK(hipSetDevice(t->device));
    vgi_host::VoxelSource src{t->grid.data(), t->vmin, t->gx, t->gy, t->gz, t->loaded.data(), t->rmin, t->rsx, t->rsy, t->rsz};
    std::vector<vgi_face_task> faces;
    for (int z = 0; z < t->gz; ++z)
        for (int y = 0; y < t->gy; ++y)
            for (int x = 0; x < t->gx; ++x) {
                int vx = t->vmin.x + x, vy = t->vmin.y + y, vz = t->vmin.z + z;
                if (t->grid[lin3(x, y, z, t->gx, t->gy)] >= vgi_host::AIR) continue;
                for (uint32_t f = 0; f < 6; ++f) {
                    if (!face_exposed(src, vx, vy, vz, f)) continue;
                    vgi_face_task task{};
                    task.voxel[0] = vx; task.voxel[1] = vy; task.voxel[2] = vz;
                    task.face = f;
                    faces.push_back(task);
                }
            }
    if (faces.size() >= (1u << 31)) return fail("too many exposed faces");
    if (vgi_bind_faces(t, faces.data(), (uint32_t)faces.size()) != 0) return -1;
    t->bound_slots = t->last_slots;
    t->bound_tasks.swap(faces);
    t->bound_of_slot.assign(t->high_water, vgi_host::NONE);
    for (size_t i = 0; i < t->bound_slots.size(); ++i) t->bound_of_slot[t->bound_slots[i]] = (uint32_t)i;
    t->follow_exposure = true;
    return 0;
}

int vgi_read_bound_tasks(vgi_tracer *t, vgi_face_task *out, uint32_t count) {
    if (!t || (count && !out)) return fail("null argument");
    if (count > t->bound) return fail("more tasks than bound faces");
    HIP_OK(hipStreamSynchronize(t->stream));
    if (count) HIP_OK(hipMemcpy(out, t->b_tasks.ptr, (size_t)count * 32, hipMemcpyDeviceToHost));
    return 0;
}

uint32_t vgi_bound_count(vgi_tracer *t) { return t ? t->bound : 0u; }

int64_t vgi_read_events(vgi_tracer *t, uint32_t *out, uint32_t capacity) {
    if (!t) return fail("null tracer");
    if (out && capacity) {
        HIP_OK(hipStreamSynchronize(t->stream));
        uint32_t n = std::min(capacity, t->last_event_words);
        if (n) HIP_OK(hipMemcpy(out, t->d_events.ptr, (size_t)n * 4, hipMemcpyDeviceToHost));
    }
    return (int64_t)t->last_event_words;
}


int vgi_run_frame(vgi_tracer *t, uint32_t flags) {
    if (!t) return fail("null tracer");
    if (!t->have_regions) return fail("no scene loaded");
    HIP_OK(hipSetDevice(t->device));
    bool record_hits = (flags & VGI_RECORD_HITS) != 0;
    if (record_hits && t->learned)
        return fail("learned frames keep no first-hit records (their paths end in the network)");
    if (t->bound_broken) return fail("the bound faces did not follow the last edit: bind faces again");
    if (sync_emitters(*t) != 0) return -1;
    vgi::SceneView s;
    vgi::Optics o;
    views(*t, s, o);
    uint32_t count = t->bound;
    bool counting = (flags & (VGI_COUNT_RAYS | VGI_COUNT_GATHERS)) != 0;
    t->bound_hit_words = 0;
    t->bound_moves.clear();
    t->frame_bound = count;
    uint32_t *hits = nullptr;
    if (record_hits) {
        uint64_t words = (uint64_t)count * t->transport.samples * 8u;
        HIP_OK(t->b_hits.reserve(std::max<uint64_t>(words, 1) * 4));
        hits = t->b_hits.as<uint32_t>();
        t->bound_hit_words = words;
    }
    HIP_OK(hipEventRecord(t->ev[8], t->stream));
    if (flush_events(*t) != 0) return -1;
    o.direct_reuse = t->direct_written == t->direct_epoch ? 1u : 0u;
    if (counting) HIP_OK(hipMemsetAsync(t->b_counters.ptr, 0, 64, t->stream));
    if (t->resident)
        vgi::launch_prepare(t->b_tasks.as<uint4>(), t->b_slots.as<uint32_t>(), count, t->d_state.as<uint32_t>(), t->d_meta.as<uint32_t>(),
                            std::max(t->transport.measured_retention, 0.0f), next_batch(*t), t->d_keys.as<uint32_t>(), t->stream);
    HIP_OK(hipEventRecord(t->ev[4], t->stream));
    bool learned_frame = t->learned && t->transport.bounces >= 2u;
    if (t->learned) {
        // The frame's training paths run on a second stream beside the
        // trace (they need the prepared tasks, not the network).
        if (!t->side) {
            HIP_OK(hipStreamCreateWithFlags(&t->side, hipStreamNonBlocking));
            for (auto &e : t->side_ev) HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        }
        HIP_OK(hipEventRecord(t->side_ev[0], t->stream));
        HIP_OK(hipStreamWaitEvent(t->side, t->side_ev[0], 0));
        if (learned_targets(*t, t->side) != 0) return -1;
        HIP_OK(hipEventRecord(t->side_ev[1], t->side));
    }
    if (learned_frame) {
        // Paths stop at their first hit; the network answers every query.
        // (A one-bounce path has nothing after its first hit to predict.)
        size_t samples = (size_t)count * t->transport.samples;
        HIP_OK(t->b_learned.reserve(samples * 32));
        HIP_OK(t->b_queries.reserve(samples * 16));
        HIP_OK(t->b_pred.reserve(samples * 12));
        o.cache_on = 0u;
        if (trace_bound_faces(*t, s, o, t->b_tasks.as<uint4>(), count, t->transport.samples, lanes_for(t->transport), 0u,
                              t->b_results.as<float4>(), nullptr, counting ? t->b_counters.as<unsigned long long>() : nullptr, false,
                              t->stream, order_of(*t), t->b_learned.as<float4>(), t->b_queries.as<vgi::MlpQuery>()) != 0)
            return -1;
        if (!t->mlp_ev[0])
            for (auto &e : t->mlp_ev) HIP_OK(hipEventCreate(&e));
        HIP_OK(hipEventRecord(t->mlp_ev[0], t->stream));
        vgi::mlp_infer(t->d_mlp_frag.as<uint16_t>(), mlp_scene(*t), t->b_queries.as<vgi::MlpQuery>(), (uint32_t)samples,
                       t->b_pred.as<float>(), t->stream);
        HIP_OK(hipEventRecord(t->mlp_ev[1], t->stream));
        t->mlp_timed_queries = (uint32_t)samples;
        vgi::launch_learned_combine(t->b_learned.as<float4>(), t->b_pred.as<float>(), count, t->transport.samples,
                                    lanes_for(t->transport), t->b_tasks.as<uint4>(), t->b_results.as<float4>(), t->stream);
    } else if (t->wavefront && !record_hits && (flags & VGI_COUNT_GATHERS) == 0 && o.light_count <= vgi::WF_MAX_LIGHTS &&
               o.bounces <= 15u) {
        if (wavefront_trace(*t, s, o, t->b_tasks.as<uint4>(), count, 0u, t->b_results.as<float4>(),
                            counting ? t->b_counters.as<unsigned long long>() : nullptr) != 0)
            return -1;
    } else {
        if (trace_bound_faces(*t, s, o, t->b_tasks.as<uint4>(), count, t->transport.samples, lanes_for(t->transport), 0u,
                              t->b_results.as<float4>(), hits, counting ? t->b_counters.as<unsigned long long>() : nullptr,
                              (flags & VGI_COUNT_GATHERS) != 0, t->stream, order_of(*t)) != 0)
            return -1;
    }
    HIP_OK(hipGetLastError());
    t->direct_written = t->direct_epoch;
    HIP_OK(hipEventRecord(t->ev[5], t->stream));
    t->frame_recorded = true;

    if (t->resident)
        vgi::launch_accumulate(t->b_slots.as<uint32_t>(), count, t->b_results.as<float4>(), t->d_state.as<uint32_t>(),
                               t->d_checkpoint.as<uint32_t>(), t->d_meta.as<uint32_t>(), accumulate_params(*t, clock_ms(*t)), t->stream);
    HIP_OK(hipGetLastError());
    if (t->learned) {
        HIP_OK(hipStreamWaitEvent(t->stream, t->side_ev[1], 0));
        if (learned_steps(*t, learned_frame) != 0) return -1;
        t->mlp_timed = learned_frame;
    }
    HIP_OK(hipEventRecord(t->ev[9], t->stream));
    t->frame_index += 1;
    return 0;
}

int vgi_read_bound_hits(vgi_tracer *t, uint32_t *out, uint64_t words) {
    if (!t || (words && !out)) return fail("null argument");
    if (t->bound_hit_words == 0) return fail("the last frame kept no first-hit records (run it with VGI_RECORD_HITS)");
    uint64_t row_words = (uint64_t)t->transport.samples * 8u;
    if (row_words == 0 || words > (uint64_t)t->bound * row_words) return fail("more hit words than the bound faces' samples");
    if (words % row_words) return fail("hit words must cover whole faces (samples x 8 words each)");
    HIP_OK(hipSetDevice(t->device));
    HIP_OK(hipStreamSynchronize(t->stream));
    if (words && read_bound_rows(*t, t->b_hits.ptr, row_words * 4, (uint32_t)(words / row_words), out) != 0) return -1;
    return 0;
}

int vgi_converged(vgi_tracer *t, uint32_t launches, uint32_t first_sequence, double *out) {
    if (!t || !out) return fail("null argument");
    if (!t->have_regions) return fail("no scene loaded");
    if (t->bound == 0) return fail("no bound faces");
    HIP_OK(hipSetDevice(t->device));
    if (sync_emitters(*t) != 0) return -1;
    vgi::SceneView s;
    vgi::Optics o;
    views(*t, s, o);
    o.cache_on = 0u;  // the reference every estimator is measured against: plain multi-bounce paths
    uint32_t count = t->bound;
    DeviceBuffer sums, results;
    HIP_OK(sums.reserve((size_t)count * 64));
    HIP_OK(results.reserve((size_t)count * 96));
    HIP_OK(hipMemsetAsync(sums.ptr, 0, (size_t)count * 64, t->stream));
    for (uint32_t l = 0; l < launches; ++l) {
        if (trace_bound_faces(*t, s, o, t->b_tasks.as<uint4>(), count, t->transport.samples, lanes_for(t->transport),
                              first_sequence + l * t->transport.samples, results.as<float4>(), nullptr, nullptr, false, t->stream,
                              order_of(*t)) != 0)
            return -1;
        vgi::launch_fp64_accumulate(results.as<float4>(), count, sums.as<double>(), t->stream);
    }
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(t->stream));
    HIP_OK(hipMemcpy(out, sums.ptr, (size_t)count * 64, hipMemcpyDeviceToHost));
    sums.release();
    results.release();
    return 0;
}

int vgi_mlp_enable(vgi_tracer *t, uint32_t seed, uint32_t train_per_frame, uint32_t steps, float lr) {
    if (!t) return fail("null tracer");
    HIP_OK(hipSetDevice(t->device));
    HIP_OK(t->d_mlp_w.reserve(vgi::MLP_PARAMS * 4));
    HIP_OK(t->d_mlp_m.reserve(vgi::MLP_PARAMS * 4));
    HIP_OK(t->d_mlp_v.reserve(vgi::MLP_PARAMS * 4));
    HIP_OK(t->d_mlp_g.reserve(vgi::MLP_PARAMS * 4));
    HIP_OK(t->d_mlp_frag.reserve((size_t)vgi::MLP_FRAG_VALUES * 2));
    HIP_OK(t->d_mlp_loss.reserve(16));
    HIP_OK(t->d_table.reserve((size_t)vgi::MLP_TABLE * 4));
    HIP_OK(t->d_table_m.reserve((size_t)vgi::MLP_TABLE * 4));
    HIP_OK(t->d_table_v.reserve((size_t)vgi::MLP_TABLE * 4));
    HIP_OK(t->d_table_g.reserve((size_t)vgi::MLP_TABLE * 4));
    // He-uniform initialisation from the seed (fan-in of each layer).
    std::vector<float> w(vgi::MLP_PARAMS);
    std::mt19937 rng(seed);
    const int ins[5] = {vgi::MLP_IN, vgi::MLP_HIDDEN, vgi::MLP_HIDDEN, vgi::MLP_HIDDEN, vgi::MLP_HIDDEN};
    const int offs[6] = {vgi::MLP_W1, vgi::MLP_W2, vgi::MLP_W3, vgi::MLP_W4, vgi::MLP_W5, vgi::MLP_PARAMS};
    for (int l = 0; l < 5; ++l) {
        float a = std::sqrt(6.0f / (float)ins[l]);
        std::uniform_real_distribution<float> u(-a, a);
        for (int i = offs[l]; i < offs[l + 1]; ++i) w[i] = u(rng);
    }
    HIP_OK(hipMemcpyAsync(t->d_mlp_w.ptr, w.data(), w.size() * 4, hipMemcpyHostToDevice, t->stream));
    // Hash table entries uniform in +-1e-4 (the encoding's initialisation).
    std::vector<float> table(vgi::MLP_TABLE);
    std::uniform_real_distribution<float> small(-1e-4f, 1e-4f);
    for (float &x : table) x = small(rng);
    HIP_OK(hipMemcpyAsync(t->d_table.ptr, table.data(), table.size() * 4, hipMemcpyHostToDevice, t->stream));
    HIP_OK(hipMemsetAsync(t->d_table_m.ptr, 0, (size_t)vgi::MLP_TABLE * 4, t->stream));
    HIP_OK(hipMemsetAsync(t->d_table_v.ptr, 0, (size_t)vgi::MLP_TABLE * 4, t->stream));
    HIP_OK(hipMemsetAsync(t->d_table_g.ptr, 0, (size_t)vgi::MLP_TABLE * 4, t->stream));
    HIP_OK(hipMemsetAsync(t->d_mlp_m.ptr, 0, vgi::MLP_PARAMS * 4, t->stream));
    HIP_OK(hipMemsetAsync(t->d_mlp_v.ptr, 0, vgi::MLP_PARAMS * 4, t->stream));
    HIP_OK(hipMemsetAsync(t->d_mlp_g.ptr, 0, vgi::MLP_PARAMS * 4, t->stream));
    vgi::mlp_pack(t->d_mlp_w.as<float>(), t->d_mlp_frag.as<uint16_t>(), t->stream);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(t->stream));
    t->mlp_ready = true;
    t->learned = true;
    t->mlp_step = 0;
    t->train_per_frame = train_per_frame;
    t->train_steps = std::max<uint32_t>(steps, 1u);
    t->mlp_lr = lr;
    // The table learns ten times faster than the network by default (its
    // entries see few samples each).
    t->table_lr = 10.0f * lr;
    return 0;
}

int vgi_mlp_train_samples(vgi_tracer *t, uint32_t samples) {
    if (!t || !t->mlp_ready) return fail("the learned cache is not enabled");
    if (samples < 1 || samples > 64) return fail("training samples per target must be 1..64");
    t->train_samples = samples;
    return 0;
}

int vgi_mlp_encoding(vgi_tracer *t, uint32_t mode) {
    if (!t || !t->mlp_ready) return fail("the learned cache is not enabled");
    if (mode != vgi::MLP_ENCODING_HASH && mode != vgi::MLP_ENCODING_FREQUENCY) return fail("unknown encoding");
    t->mlp_encoding = mode;
    return 0;
}

int vgi_mlp_training(vgi_tracer *t, uint32_t train_per_frame, uint32_t steps, float lr) {
    if (!t || !t->mlp_ready) return fail("the learned cache is not enabled");
    if (!(lr >= 0.0f)) return fail("the learning rate is not a non-negative number");
    t->train_per_frame = train_per_frame;
    t->train_steps = std::max<uint32_t>(steps, 1u);
    t->mlp_lr = lr;
    t->table_lr = 10.0f * lr;
    return 0;
}

int vgi_mlp_table_lr(vgi_tracer *t, float lr) {
    if (!t || !t->mlp_ready) return fail("the learned cache is not enabled");
    if (!(lr >= 0.0f)) return fail("the table's learning rate is not a non-negative number");
    t->table_lr = lr;
    return 0;
}

int vgi_mlp_set_learned(vgi_tracer *t, int on) {
    if (!t) return fail("null tracer");
    if (on && !t->mlp_ready) return fail("the learned cache is not enabled (vgi_mlp_enable)");
    t->learned = on != 0;
    return 0;
}

int vgi_mlp_weights(vgi_tracer *t, float *out, const float *in) {
    if (!t || !t->mlp_ready) return fail("the learned cache is not enabled");
    HIP_OK(hipSetDevice(t->device));
    if (out) {
        HIP_OK(hipStreamSynchronize(t->stream));
        HIP_OK(hipMemcpy(out, t->d_mlp_w.ptr, vgi::MLP_PARAMS * 4, hipMemcpyDeviceToHost));
    }
    if (in) {
        HIP_OK(hipMemcpyAsync(t->d_mlp_w.ptr, in, vgi::MLP_PARAMS * 4, hipMemcpyHostToDevice, t->stream));
        vgi::mlp_pack(t->d_mlp_w.as<float>(), t->d_mlp_frag.as<uint16_t>(), t->stream);
        HIP_OK(hipStreamSynchronize(t->stream));
    }
    return 0;
}

int vgi_mlp_infer(vgi_tracer *t, const uint32_t *queries, uint32_t count, float *out) {
    if (!t || !t->mlp_ready) return fail("the learned cache is not enabled");
    if (count && (!queries || !out)) return fail("null argument");
    if (!t->have_regions) return fail("no scene loaded");
    HIP_OK(hipSetDevice(t->device));
    HIP_OK(t->d_io.reserve((size_t)std::max<uint32_t>(count, 1) * 28));
    char *base = t->d_io.as<char>();
    HIP_OK(hipMemcpyAsync(base, queries, (size_t)count * 16, hipMemcpyHostToDevice, t->stream));
    vgi::mlp_infer(t->d_mlp_frag.as<uint16_t>(), mlp_scene(*t), reinterpret_cast<vgi::MlpQuery *>(base), count,
                   reinterpret_cast<float *>(base + (size_t)count * 16), t->stream);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(t->stream));
    HIP_OK(hipMemcpy(out, base + (size_t)count * 16, (size_t)count * 12, hipMemcpyDeviceToHost));
    return 0;
}

int vgi_mlp_train(vgi_tracer *t, const uint32_t *queries, const float *targets, uint32_t count, int update, float *grad_out,
                  float *loss_out) {
    if (!t || !t->mlp_ready) return fail("the learned cache is not enabled");
    if (count == 0 || (count % 16u) != 0) return fail("training batches are a positive multiple of 16 samples");
    if (!queries || !targets) return fail("null argument");
    if (!t->have_regions) return fail("no scene loaded");
    HIP_OK(hipSetDevice(t->device));
    HIP_OK(t->d_io.reserve((size_t)count * 32));
    char *base = t->d_io.as<char>();
    HIP_OK(hipMemcpyAsync(base, queries, (size_t)count * 16, hipMemcpyHostToDevice, t->stream));
    HIP_OK(hipMemcpyAsync(base + (size_t)count * 16, targets, (size_t)count * 16, hipMemcpyHostToDevice, t->stream));
    if (mlp_step(*t, reinterpret_cast<vgi::MlpQuery *>(base), reinterpret_cast<const float4 *>(base + (size_t)count * 16), count,
                 update != 0) != 0)
        return -1;
    HIP_OK(hipStreamSynchronize(t->stream));
    if (loss_out) HIP_OK(hipMemcpy(loss_out, t->d_mlp_loss.ptr, 4, hipMemcpyDeviceToHost));
    if (!update) {
        if (grad_out) HIP_OK(hipMemcpy(grad_out, t->d_mlp_g.ptr, vgi::MLP_PARAMS * 4, hipMemcpyDeviceToHost));
        HIP_OK(hipMemset(t->d_mlp_g.ptr, 0, vgi::MLP_PARAMS * 4));
        t->table_grad_last.resize(vgi::MLP_TABLE);
        HIP_OK(hipMemcpy(t->table_grad_last.data(), t->d_table_g.ptr, (size_t)vgi::MLP_TABLE * 4, hipMemcpyDeviceToHost));
        HIP_OK(hipMemset(t->d_table_g.ptr, 0, (size_t)vgi::MLP_TABLE * 4));
    }
    return 0;
}

int vgi_mlp_table(vgi_tracer *t, float *out, const float *in, float *grad_out) {
    if (!t || !t->mlp_ready) return fail("the learned cache is not enabled");
    HIP_OK(hipSetDevice(t->device));
    if (out) {
        HIP_OK(hipStreamSynchronize(t->stream));
        HIP_OK(hipMemcpy(out, t->d_table.ptr, (size_t)vgi::MLP_TABLE * 4, hipMemcpyDeviceToHost));
    }
    if (in) {
        HIP_OK(hipMemcpyAsync(t->d_table.ptr, in, (size_t)vgi::MLP_TABLE * 4, hipMemcpyHostToDevice, t->stream));
        HIP_OK(hipStreamSynchronize(t->stream));
    }
    if (grad_out) {
        if (t->table_grad_last.size() != vgi::MLP_TABLE) return fail("no gradient yet: call vgi_mlp_train without update");
        std::memcpy(grad_out, t->table_grad_last.data(), (size_t)vgi::MLP_TABLE * 4);
    }
    return 0;
}

// Training-only frames: each traces one window of training faces (plain
// multi-bounce, one sample) and takes the frame's Adam steps, without tracing
// the bound faces -- the network's warm-up before it serves paths.
int vgi_mlp_pretrain(vgi_tracer *t, uint32_t frames) {
    if (!t || !t->mlp_ready) return fail("the learned cache is not enabled");
    if (!t->have_regions) return fail("no scene loaded");
    if (t->bound == 0) return fail("no bound faces");
    HIP_OK(hipSetDevice(t->device));
    if (sync_emitters(*t) != 0) return -1;
    for (uint32_t f = 0; f < frames; ++f) {
        if (learned_training(*t) != 0) return -1;
        t->frame_index += 1;
    }
    return 0;
}

int vgi_mlp_timings(vgi_tracer *t, float out[6]) {
    if (!t || !out) return fail("null argument");
    if (!t->mlp_timed) return fail("no learned frame has run");
    HIP_OK(hipSetDevice(t->device));
    HIP_OK(hipStreamSynchronize(t->stream));
    float infer = 0.0f, mfma = 0.0f, steps = 0.0f;
    HIP_OK(hipEventElapsedTime(&infer, t->mlp_ev[0], t->mlp_ev[1]));
    for (uint32_t k = 0; k < t->mlp_timed_steps; ++k) {
        float a = 0.0f, b = 0.0f;
        HIP_OK(hipEventElapsedTime(&a, t->mlp_step_ev[k * 3], t->mlp_step_ev[k * 3 + 1]));
        HIP_OK(hipEventElapsedTime(&b, t->mlp_step_ev[k * 3], t->mlp_step_ev[k * 3 + 2]));
        mfma += a;
        steps += b;
    }
    out[0] = infer;
    out[1] = mfma;
    out[2] = steps;
    out[3] = (float)t->mlp_timed_queries;
    out[4] = (float)t->mlp_timed_step_samples * (float)t->mlp_timed_steps;
    out[5] = (float)t->mlp_timed_steps;
    return 0;
}

float vgi_mlp_last_loss(vgi_tracer *t) {
    if (!t || !t->mlp_ready) return 0.0f;
    float l = 0.0f;
    hipStreamSynchronize(t->stream);
    hipMemcpy(&l, t->d_mlp_loss.ptr, 4, hipMemcpyDeviceToHost);
    return l;
}

int vgi_sync(vgi_tracer *t) {
    if (!t) return fail("null tracer");
    HIP_OK(hipStreamSynchronize(t->stream));
    // Only a frame records the trace events; timing them otherwise would
    // leave an error behind for the next call.
    if (t->frame_recorded) {
        HIP_OK(hipEventElapsedTime(&t->last_trace_ms, t->ev[4], t->ev[5]));
        HIP_OK(hipEventElapsedTime(&t->last_frame_ms, t->ev[8], t->ev[9]));
    }
    return 0;
}

float vgi_last_trace_ms(vgi_tracer *t) { return t ? t->last_trace_ms : 0.0f; }
float vgi_last_frame_ms(vgi_tracer *t) { return t ? t->last_frame_ms : 0.0f; }

int vgi_set_wavefront(vgi_tracer *t, int on) {
    if (!t) return fail("null tracer");
    t->wavefront = on != 0;
    return 0;
}

int vgi_set_dedup(vgi_tracer *t, int on) {
    if (!t) return fail("null tracer");
    t->dedup = on != 0;
    return 0;
}


int vgi_read_bound(vgi_tracer *t, vgi_face_result *out, uint32_t count, uint64_t *counters) {
    if (!t || (count && !out)) return fail("null argument");
    if (count > t->bound) return fail("more records than bound faces");
    HIP_OK(hipSetDevice(t->device));
    HIP_OK(hipStreamSynchronize(t->stream));
    if (count && read_bound_rows(*t, t->b_results.ptr, 96, count, out) != 0) return -1;
    if (counters) HIP_OK(hipMemcpy(counters, t->b_counters.ptr, 56, hipMemcpyDeviceToHost));
    return 0;
}

int vgi_apply_events(vgi_tracer *t, const uint32_t *events, uint32_t event_count, uint32_t words) {
    if (!t || (words && !events)) return fail("null argument");
    if (!t->resident) return fail("retention events need the resident cache");
    if ((uint64_t)event_count * 12 > words) return fail("event headers exceed the buffer");
    HIP_OK(hipSetDevice(t->device));
    DeviceBuffer ev;
    HIP_OK(ev.reserve((size_t)std::max<uint32_t>(words, 4) * 4));
    HIP_OK(hipMemcpyAsync(ev.ptr, events, (size_t)words * 4, hipMemcpyHostToDevice, t->stream));
    vgi::launch_apply_events(ev.as<uint4>(), event_count, t->high_water, t->d_state.as<uint32_t>(), t->d_checkpoint.as<uint32_t>(),
                             t->d_meta.as<uint32_t>(), std::max(t->transport.measured_retention, 0.0f), t->stream);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(t->stream));
    ev.release();
    return 0;
}

// Sun corner words of the bound faces (voxel_view_runtime.wgsl:168-268 for
// the world's faces): refresh, flips in bound order, then the host's
// SunChanges (sun_changes.rs) and, when due, a Light retention event applied
// with the next batch.
int vgi_sun_corners(vgi_tracer *t, uint32_t frame, uint32_t refresh_frames) {
    if (!t) return fail("null tracer");
    if (!t->resident) return fail("sun corner words are kept per resident slot: call vgi_enable_resident first");
    if (!t->have_regions) return fail("no scene loaded");
    HIP_OK(hipSetDevice(t->device));
    uint32_t count = t->bound;
    if (count && t->b_slots.bytes < (size_t)count * 4) return fail("bound faces have no resident slots: bind them with the cache on");
    HIP_OK(t->d_sun_flux.reserve((size_t)std::max<uint32_t>(count, 1) * 4));
    HIP_OK(t->d_sun_queue.reserve(((size_t)count + 1) * 4));
    HIP_OK(t->d_sun_ranges.reserve((size_t)10 * vgi::SUN_RANGES * 4));
    HIP_OK(t->d_sun_region.reserve((size_t)vgi::SUN_FLIP_WORDS * 4));
    if (!t->sun_ev[0]) {
        HIP_OK(hipEventCreate(&t->sun_ev[0]));
        HIP_OK(hipEventCreate(&t->sun_ev[1]));
    }
    vgi::SceneView s;
    vgi::Optics o;
    views(*t, s, o);
    HIP_OK(hipEventRecord(t->sun_ev[0], t->stream));
    vgi::launch_sun_corners(s, o, t->b_tasks.as<uint4>(), t->b_slots.as<uint32_t>(), count, t->d_meta.as<uint32_t>(),
                            t->d_corners.as<uint32_t>(), t->d_sun_flux.as<float>(), t->d_sun_queue.as<uint32_t>(),
                            (uint32_t)std::max(t->cus, 1) * 32u, frame, refresh_frames,
                            t->d_sun_ranges.as<uint32_t>(), t->d_sun_region.as<uint32_t>(), t->stream);
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord(t->sun_ev[1], t->stream));
    t->sun_region.resize(vgi::SUN_FLIP_WORDS);
    HIP_OK(hipMemcpyAsync(t->sun_region.data(), t->d_sun_region.ptr, (size_t)vgi::SUN_FLIP_WORDS * 4, hipMemcpyDeviceToHost, t->stream));
    HIP_OK(hipStreamSynchronize(t->stream));
    HIP_OK(hipEventElapsedTime(&t->sun_ms[0], t->sun_ev[0], t->sun_ev[1]));
    auto t_host = std::chrono::steady_clock::now();
    std::vector<vgi_host::SunFlip> flips;
    vgi_host::decode_sun_flips(t->sun_region.data(), vgi::SUN_FLIP_RECORDS, flips);
    t->sun_changes.add(flips);
    std::vector<vgi_host::SunFlip> due;
    if (t->sun_changes.take(clock_ms(*t), due)) {
        std::vector<uint8_t> page_solid;
        vgi_host::PreviousScene scene = scene_pages(*t, t->solid, page_solid);
        vgi_host::RetentionEvent ev;
        if (vgi_host::light_event(due, scene, ev)) t->pending_events.push_back(std::move(ev));
    }
    t->sun_ms[1] = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t_host).count();
    return 0;
}

int vgi_sun_timings(vgi_tracer *t, float *out, uint32_t n) {
    if (!t || (n && !out)) return fail("null argument");
    for (uint32_t i = 0; i < n && i < 2; ++i) out[i] = t->sun_ms[i];
    return 0;
}

int64_t vgi_read_sun_flips(vgi_tracer *t, uint32_t *out, uint32_t capacity) {
    if (!t) return fail("null tracer");
    if (out) std::memcpy(out, t->sun_region.data(), (size_t)std::min<size_t>(capacity, t->sun_region.size()) * 4);
    return (int64_t)t->sun_region.size();
}

int vgi_read_sun_corners(vgi_tracer *t, uint32_t first, uint32_t count, uint32_t *out) {
    if (!t || (count && !out)) return fail("null argument");
    if ((size_t)first + count > t->high_water) return fail("slot range beyond the resident slots");
    HIP_OK(hipSetDevice(t->device));
    HIP_OK(hipStreamSynchronize(t->stream));
    if (count) HIP_OK(hipMemcpy(out, t->d_corners.as<uint32_t>() + first, (size_t)count * 4, hipMemcpyDeviceToHost));
    return 0;
}

// ---------------------------------------------------------------- view tier
// render_view's passes for one camera (gpu/view/runtime.rs:1254-1266):
// discover, index, reproject, load_shared_faces, trace_visible_paths,
// reconstruct_view, store_shared_faces, publish_view, then the far tier's
// classify_pixels and trace_pixel_queue, on the tracer's stream after its
// earlier work.
int vgi_view_frame(vgi_tracer *t, const vgi_view_camera *camera, const float *depth, const float *normals) {
    if (!t || !camera || !depth || !normals) return fail("null argument");
    if (camera->width == 0 || camera->height == 0) return fail("empty prepass");
    const uint64_t pixels64 = (uint64_t)camera->width * camera->height;
    if (pixels64 > VGI_VIEW_MAX_PIXELS) return fail("prepass larger than VGI_VIEW_MAX_PIXELS");
    if (camera->samples > 1024u) return fail("more than 1024 paths per face per view frame");
    if (!t->have_regions) return fail("no scene loaded");
    HIP_OK(hipSetDevice(t->device));
    if (sync_emitters(*t) != 0) return -1;
    const uint32_t pixels = (uint32_t)pixels64;
    uint32_t capacity = 64;
    while (capacity < 2u * pixels) capacity <<= 1;
    const uint32_t g = t->view_next, pg = g ^ 1u;
    HIP_OK(t->v_depth.reserve((size_t)pixels * 4));
    HIP_OK(t->v_normals.reserve((size_t)pixels * 12));
    HIP_OK(t->v_queue.reserve(((size_t)pixels + 2) * 4));
    HIP_OK(t->v_receivers[g].reserve((size_t)pixels * sizeof(vgi::ViewReceiver)));
    HIP_OK(t->v_estimates[g].reserve((size_t)pixels * sizeof(vgi::ViewEstimate)));
    HIP_OK(t->v_owners[g].reserve((size_t)pixels * 4));
    HIP_OK(t->v_hash[g].reserve((size_t)capacity * 4));
    HIP_OK(t->v_pixels.reserve((size_t)pixels * sizeof(vgi::ViewPixel)));
    HIP_OK(t->v_counters.reserve(16 * 4));
    // The far tier's cell grid (settings.rs:141-147, gpu/view/runtime.rs:912-916).
    const vgi_far_settings &fs = t->far;
    const bool far_on = fs.voxel_pixels > 0.0f;
    if (far_on) {
        // The far tier measures faces against the camera's forward axis and
        // its pixels per unit (view.pixel_axis, gpu/view/runtime.rs:982-987).
        const float *fw = camera->forward;
        const float len2 = fw[0] * fw[0] + fw[1] * fw[1] + fw[2] * fw[2];
        if (!(camera->pixels_per_unit > 0.0f) || !std::isfinite(camera->pixels_per_unit))
            return fail("the far tier needs the camera's pixels per unit at view depth 1 (> 0)");
        if (!(len2 > 0.5f && len2 < 2.0f)) return fail("the far tier needs the camera's unit forward vector");
    }
    uint32_t size = fs.cell != 0u ? fs.cell : (uint32_t)std::lround((double)((float)camera->height / 480.0f));
    size = std::min(std::max(size, 1u), 4u);
    const uint32_t cells_x = (camera->width + size - 1u) / size, cells_y = (camera->height + size - 1u) / size;
    if (far_on && (cells_x != t->far_cx || cells_y != t->far_cy || size != t->far_size)) {
        // A new grid starts with empty cells (PixelState::new).
        const size_t cells = (size_t)cells_x * cells_y;
        for (auto &b : t->f_cells) {
            HIP_OK(b.reserve(cells * sizeof(vgi::PixelHistory)));
            vgi::launch_fill_words(b.as<uint32_t>(), cells * sizeof(vgi::PixelHistory) / 4, 0u, t->stream);
        }
        HIP_OK(t->f_output.reserve((16 + cells * 4) * 4));
        vgi::launch_fill_words(t->f_output.as<uint32_t>(), 16 + cells * 4, 0u, t->stream);
        HIP_OK(t->f_queue.reserve((1 + cells) * 4));
        t->far_cx = cells_x; t->far_cy = cells_y; t->far_size = size;
    }
    if (staged_upload(*t, t->v_depth.ptr, depth, (size_t)pixels * 4) != 0) return -1;
    if (staged_upload(*t, t->v_normals.ptr, normals, (size_t)pixels * 12) != 0) return -1;
    vgi::launch_fill_words(t->v_hash[g].as<uint32_t>(), capacity, 0u, t->stream);
    vgi::launch_fill_words(t->v_queue.as<uint32_t>(), 2, 0u, t->stream);
    vgi::launch_fill_words(t->v_counters.as<uint32_t>(), 16, 0u, t->stream);
    if (far_on) vgi::launch_fill_words(t->f_queue.as<uint32_t>(), 1, 0u, t->stream);
    vgi::SceneView s;
    vgi::Optics o;
    views(*t, s, o);
    // params.transport.y of the view scene (gpu/view.rs:104-117): some light
    // has custom visible rgb, which no host light draws.
    uint32_t custom = 0u;
    for (size_t l = 0; l + 12 <= t->lights.size(); l += 12)
        if (t->lights[l + 9] != 0.0f || t->lights[l + 10] != 0.0f || t->lights[l + 11] != 0.0f) custom = 1u;
    vgi::ViewParams p;
    std::memcpy(p.m, camera->world_from_clip, sizeof(p.m));
    p.cam[0] = camera->position[0]; p.cam[1] = camera->position[1]; p.cam[2] = camera->position[2];
    p.cam[3] = camera->inner_distance;
    p.width = camera->width; p.height = camera->height; p.pixels = pixels;
    p.mask = capacity - 1u;
    const bool history = t->view_history && t->view_pixels[pg] > 0u;
    p.prev_pixels = history ? t->view_pixels[pg] : 0u;
    p.prev_mask = history ? t->view_mask[pg] : 0u;
    p.samples = camera->samples;
    p.retention = std::max(camera->measured_retention, 0.0f);
    p.clock_ms = clock_ms(*t);
    p.exchange = t->resident && t->high_water > 0 ? 1u : 0u;
    std::memcpy(p.display, t->view_display, sizeof(p.display));
    p.custom = custom;
    vgi::FarParams f;
    // The previous camera's clip_from_world (this frame's when there is none).
    std::memcpy(f.clip, t->far_prev ? t->far_prev_clip : camera->clip_from_world, sizeof(f.clip));
    f.axis[0] = camera->forward[0]; f.axis[1] = camera->forward[1]; f.axis[2] = camera->forward[2];
    f.axis[3] = camera->pixels_per_unit;
    f.cells_x = cells_x; f.cells_y = cells_y; f.size = size;
    f.flags = (far_on ? 1u : 0u) | (fs.measure << 1u);
    f.threshold = fs.voxel_pixels;
    f.history = fs.history;
    f.rate = (float)fs.rate;
    f.previous = t->far_prev ? t->far_prev_frame : 0xffffffffu;
    f.frame = camera->frame;
    f.custom = custom;
    f.steps = fs.steps;
    f.cells = far_on ? cells_x * cells_y : 0u;
    const bool have_cells = t->far_size != 0u;
    vgi::FarBuffers fb;
    fb.previous = have_cells ? t->f_cells[pg].as<vgi::PixelHistory>() : nullptr;
    fb.cells = have_cells ? t->f_cells[g].as<vgi::PixelHistory>() : nullptr;
    fb.output = t->f_output.as<uint32_t>();
    fb.queue = t->f_queue.as<uint32_t>();
    vgi::SharedCache c;
    c.index = t->d_index.as<uint4>();
    c.state = t->d_state.as<uint32_t>();
    c.meta = t->d_meta.as<uint32_t>();
    c.mask = t->resident ? (uint32_t)(t->face_slots.entries.size() - 1) : 0u;
    if (!t->view_ev[0])
        for (auto &e : t->view_ev) HIP_OK(hipEventCreate(&e));
    HIP_OK(hipEventRecord(t->view_ev[0], t->stream));
    vgi::launch_view_frame(p, f, s, o, c, t->v_depth.as<float>(), t->v_normals.as<float>(), t->v_receivers[g].as<vgi::ViewReceiver>(),
                           t->v_estimates[g].as<vgi::ViewEstimate>(), t->v_hash[g].as<uint32_t>(), t->v_owners[g].as<uint32_t>(),
                           history ? t->v_receivers[pg].as<vgi::ViewReceiver>() : nullptr,
                           history ? t->v_estimates[pg].as<vgi::ViewEstimate>() : nullptr,
                           history ? t->v_hash[pg].as<uint32_t>() : nullptr, t->v_queue.as<uint32_t>(),
                           t->v_pixels.as<vgi::ViewPixel>(), t->v_counters.as<uint32_t>(), fb, t->stream, t->view_ev[1],
                           t->view_ev[2], t->view_ev[3]);
    HIP_OK(hipEventRecord(t->view_ev[4], t->stream));
    HIP_OK(hipGetLastError());
    t->view_pixels[g] = pixels;
    t->view_mask[g] = capacity - 1u;
    t->view_next = pg;
    t->view_history = true;
    t->far_last_cells = f.cells;
    std::memcpy(t->far_prev_clip, camera->clip_from_world, sizeof(t->far_prev_clip));
    t->far_prev_frame = camera->frame;
    t->far_prev = true;
    return 0;
}

int vgi_view_far(vgi_tracer *t, const vgi_far_settings *settings) {
    if (!t) return fail("null tracer");
    if (!settings) {
        t->far.voxel_pixels = 0.0f;
        return 0;
    }
    vgi_far_settings f = *settings;
    if (!(f.voxel_pixels >= 0.0f) || f.voxel_pixels > 4096.0f) return fail("voxel_pixels outside [0, 4096]");
    if (f.measure > 2u) return fail("measure is 0, 1 or 2");
    if (f.cell > 4u) return fail("cell is 0 (follow the height) or 1-4 pixels");
    if (!(f.history >= 1.0f && f.history <= 1024.0f)) return fail("history outside [1, 1024]");
    if (f.rate < 1u || f.rate > 8u) return fail("rate outside [1, 8]");
    if (f.steps < 1u || f.steps > 2048u) return fail("steps outside [1, 2048]");
    t->far = f;
    return 0;
}

int vgi_view_display(vgi_tracer *t, float noise_threshold, float spatial, float blend) {
    if (!t) return fail("null tracer");
    if (!(noise_threshold >= 0.0f && noise_threshold <= 1.0f)) return fail("noise threshold outside [0, 1]");
    if (!(spatial >= 0.0f && spatial <= 1.0f)) return fail("spatial strength outside [0, 1]");
    if (!(blend > 0.0f && blend <= 1.0f)) return fail("temporal blend outside (0, 1]");
    t->view_display[0] = noise_threshold;
    t->view_display[1] = spatial;
    t->view_display[2] = blend;
    return 0;
}

// The last view frame's generation, or -1 when there is none.
static int view_last(const vgi_tracer &t) {
    uint32_t g = t.view_next ^ 1u;
    return t.view_history && t.view_pixels[g] > 0u ? (int)g : -1;
}

int vgi_view_timings(vgi_tracer *t, float out[3]) {
    if (!t || !out) return fail("null argument");
    if (view_last(*t) < 0 || !t->view_ev[0]) return fail("no view frame has run");
    HIP_OK(hipSetDevice(t->device));
    HIP_OK(hipEventSynchronize(t->view_ev[4]));
    HIP_OK(hipEventElapsedTime(&out[0], t->view_ev[0], t->view_ev[4]));
    HIP_OK(hipEventElapsedTime(&out[1], t->view_ev[1], t->view_ev[2]));
    HIP_OK(hipEventElapsedTime(&out[2], t->view_ev[3], t->view_ev[4]));
    return 0;
}

int vgi_view_reset(vgi_tracer *t) {
    if (!t) return fail("null tracer");
    t->view_history = false;
    t->far_prev = false;
    return 0;
}

int vgi_view_read_pixels(vgi_tracer *t, uint32_t first, uint32_t count, vgi_view_pixel *out) {
    if (!t || (count && !out)) return fail("null argument");
    int g = view_last(*t);
    if (g < 0) return fail("no view frame has run");
    if ((uint64_t)first + count > t->view_pixels[g]) return fail("pixel range beyond the last view frame's prepass");
    if (count == 0) return 0;
    static_assert(sizeof(vgi_view_pixel) == sizeof(vgi::ViewPixel), "view pixel record layout");
    HIP_OK(hipSetDevice(t->device));
    HIP_OK(hipStreamSynchronize(t->stream));
    HIP_OK(hipMemcpy(out, t->v_pixels.as<vgi::ViewPixel>() + first, (size_t)count * sizeof(vgi::ViewPixel), hipMemcpyDeviceToHost));
    return 0;
}

int vgi_view_read_counters(vgi_tracer *t, uint32_t out[16]) {
    if (!t || !out) return fail("null argument");
    if (view_last(*t) < 0) return fail("no view frame has run");
    HIP_OK(hipSetDevice(t->device));
    HIP_OK(hipStreamSynchronize(t->stream));
    HIP_OK(hipMemcpy(out, t->v_counters.ptr, 16 * 4, hipMemcpyDeviceToHost));
    return 0;
}

int64_t vgi_view_read_far(vgi_tracer *t, uint32_t *output, uint64_t output_words, uint32_t *cells, uint64_t cell_words) {
    if (!t) return fail("null tracer");
    if (view_last(*t) < 0) return fail("no view frame has run");
    const uint64_t n = t->far_last_cells;
    if (n == 0) return fail("the far tier did not run in the last view frame");
    if (output && output_words < 16 + 4 * n) return fail("output holds fewer than 16 + 4 words per cell");
    if (cells && cell_words < 12 * n) return fail("cells hold fewer than 12 words per cell");
    HIP_OK(hipSetDevice(t->device));
    HIP_OK(hipStreamSynchronize(t->stream));
    if (output) HIP_OK(hipMemcpy(output, t->f_output.ptr, (16 + 4 * n) * 4, hipMemcpyDeviceToHost));
    if (cells) HIP_OK(hipMemcpy(cells, t->f_cells[view_last(*t)].ptr, 12 * n * 4, hipMemcpyDeviceToHost));
    return (int64_t)n;
}

int64_t vgi_view_read_owners(vgi_tracer *t, uint32_t *out, uint32_t pixels, uint32_t *overflow) {
    if (!t) return fail("null tracer");
    int g = view_last(*t);
    if (g < 0) return fail("no view frame has run");
    if (pixels && !out) return fail("null argument");
    if (pixels > t->view_pixels[g]) return fail("more pixels than the last view frame's prepass");
    HIP_OK(hipSetDevice(t->device));
    HIP_OK(hipStreamSynchronize(t->stream));
    uint32_t counts[2];
    HIP_OK(hipMemcpy(counts, t->v_queue.ptr, 8, hipMemcpyDeviceToHost));
    if (pixels) HIP_OK(hipMemcpy(out, t->v_owners[g].ptr, (size_t)pixels * 4, hipMemcpyDeviceToHost));
    if (overflow) *overflow = counts[1];
    return (int64_t)counts[0];
}

int vgi_view_read_faces(vgi_tracer *t, uint32_t first, uint32_t count, vgi_view_face *out) {
    if (!t || (count && !out)) return fail("null argument");
    int g = view_last(*t);
    if (g < 0) return fail("no view frame has run");
    if ((uint64_t)first + count > t->view_pixels[g]) return fail("pixel range beyond the last view frame's prepass");
    if (count == 0) return 0;
    HIP_OK(hipSetDevice(t->device));
    HIP_OK(hipStreamSynchronize(t->stream));
    std::vector<vgi::ViewReceiver> r(count);
    std::vector<vgi::ViewEstimate> e(count);
    HIP_OK(hipMemcpy(r.data(), t->v_receivers[g].as<vgi::ViewReceiver>() + first, (size_t)count * sizeof(vgi::ViewReceiver),
                     hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(e.data(), t->v_estimates[g].as<vgi::ViewEstimate>() + first, (size_t)count * sizeof(vgi::ViewEstimate),
                     hipMemcpyDeviceToHost));
    static_assert(sizeof(vgi_view_face) == sizeof(vgi::ViewReceiver) + sizeof(vgi::ViewEstimate), "view face record layout");
    for (uint32_t i = 0; i < count; ++i) {
        std::memcpy(&out[i], &r[i], sizeof(vgi::ViewReceiver));
        std::memcpy(reinterpret_cast<char *>(&out[i]) + sizeof(vgi::ViewReceiver), &e[i], sizeof(vgi::ViewEstimate));
    }
    return 0;
}

int vgi_set_footprints(vgi_tracer *t, uint32_t on) {
    if (!t) return fail("null tracer");
    t->footprints = on != 0u;
    return 0;
}

int vgi_set_schedule(vgi_tracer *t, const vgi_schedule *schedule) {
    if (!t || !schedule) return fail("null argument");
    t->schedule = *schedule;
    return 0;
}

// A resident batch of GPU picks (resident.rs:1093-1164 selection pass, then
// mark_changed, the trace and accumulate of the picked faces): after any
// queued retention, the selection passes pick up to `picks` faces among all
// resident slots, ranked by the schedule's focus and stability settings; the
// picks are traced and merged like host-requested faces, and their host
// mirror goes stale until classification emits their statistics.
int vgi_select_frame(vgi_tracer *t, uint32_t picks, uint32_t flags) {
    if (!t) return fail("null tracer");
    if (!t->resident) return fail("GPU selection picks among resident faces: call vgi_enable_resident first");
    if (!t->have_regions) return fail("no scene loaded");
    if (picks == 0) return fail("a pick batch needs a budget");
    HIP_OK(hipSetDevice(t->device));
    if (sync_emitters(*t) != 0) return -1;
    if (!t->sel_ev[0])
        for (auto &e : t->sel_ev) HIP_OK(hipEventCreate(&e));
    uint32_t batch = next_batch(*t);
    if (flush_events(*t) != 0) return -1;
    uint32_t mirror_budget = std::max<uint32_t>(t->schedule.mirror_budget, 1u);
    HIP_OK(t->d_selection.reserve(vgi::SEL_WORDS * 4));
    HIP_OK(t->d_pick_tasks.reserve((size_t)picks * 32));
    HIP_OK(t->d_pick_slots.reserve((size_t)picks * 4));
    HIP_OK(t->d_pick_results.reserve((size_t)picks * 96));
    HIP_OK(t->d_mirror.reserve((size_t)mirror_budget * 112));
    HIP_OK(t->d_select_ranges.reserve((size_t)vgi::SELECT_RANGES * 16 * 4));
    HIP_OK(t->b_counters.reserve(64));
    vgi::SelectParams p;
    p.fx = t->schedule.focus[0]; p.fy = t->schedule.focus[1]; p.fz = t->schedule.focus[2];
    p.budget = picks;
    p.batch = batch;
    p.now_ms = clock_ms(*t);
    p.revision = t->schedule.transport_revision;
    p.stable_samples = t->schedule.stable_samples;
    p.audit_ms = t->schedule.audit_ms;
    // resident.rs:685-690: each batch re-ranks one quarter of the slots.
    uint32_t chunk = (t->high_water + 3u) / 4u;
    uint32_t first = (batch % 4u) * chunk;
    p.first = std::min(first, t->high_water);
    p.end = std::min(first + chunk, t->high_water);
    p.slots = t->high_water;
    p.mirror_budget = mirror_budget;
    vgi::SceneView s;
    vgi::Optics o;
    views(*t, s, o);
    bool counting = (flags & (VGI_COUNT_RAYS | VGI_COUNT_GATHERS)) != 0;
    if (counting) HIP_OK(hipMemsetAsync(t->b_counters.ptr, 0, 64, t->stream));
    HIP_OK(hipEventRecord(t->sel_ev[0], t->stream));
    vgi::launch_select(p, t->d_meta.as<uint32_t>(), t->d_state.as<uint32_t>(), t->d_keys.as<uint32_t>(), t->d_selection.as<uint32_t>(),
                       t->d_mirror.as<uint4>(), t->d_pick_tasks.as<uint4>(), t->d_pick_slots.as<uint32_t>(),
                       t->d_select_ranges.as<uint32_t>(), (uint32_t)std::max(t->cus, 1) * 4u, t->stream);
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord(t->sel_ev[1], t->stream));
    // mark_changed for the picks (no host claim: batch 0).
    vgi::launch_prepare(t->d_pick_tasks.as<uint4>(), t->d_pick_slots.as<uint32_t>(), picks, t->d_state.as<uint32_t>(),
                        t->d_meta.as<uint32_t>(), std::max(t->transport.measured_retention, 0.0f), 0u, nullptr, t->stream);
    if (trace_bound_faces(*t, s, o, t->d_pick_tasks.as<uint4>(), picks, t->transport.samples, lanes_for(t->transport), 0u,
                          t->d_pick_results.as<float4>(), nullptr, counting ? t->b_counters.as<unsigned long long>() : nullptr,
                          (flags & VGI_COUNT_GATHERS) != 0, t->stream, nullptr) != 0)
        return -1;
    HIP_OK(hipEventRecord(t->sel_ev[2], t->stream));
    vgi::AccumulateParams ap = accumulate_params(*t, p.now_ms);
    ap.picked = 1u;
    vgi::launch_accumulate(t->d_pick_slots.as<uint32_t>(), picks, t->d_pick_results.as<float4>(), t->d_state.as<uint32_t>(),
                           t->d_checkpoint.as<uint32_t>(), t->d_meta.as<uint32_t>(), ap, t->stream);
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord(t->sel_ev[3], t->stream));
    t->picks = picks;
    t->mirror_budget = mirror_budget;
    t->select_recorded = true;
    return 0;
}

int vgi_select_timings(vgi_tracer *t, float *out) {
    if (!t || !out) return fail("null argument");
    if (!t->select_recorded) return fail("no pick batch yet");
    HIP_OK(hipEventSynchronize(t->sel_ev[3]));
    HIP_OK(hipEventElapsedTime(&out[0], t->sel_ev[0], t->sel_ev[1]));
    HIP_OK(hipEventElapsedTime(&out[1], t->sel_ev[1], t->sel_ev[2]));
    HIP_OK(hipEventElapsedTime(&out[2], t->sel_ev[2], t->sel_ev[3]));
    return 0;
}

int vgi_read_picks(vgi_tracer *t, uint32_t *slots, uint32_t capacity, uint32_t *count) {
    if (!t || !count || (capacity && !slots)) return fail("null argument");
    if (!t->select_recorded) return fail("no pick batch yet");
    HIP_OK(hipStreamSynchronize(t->stream));
    uint32_t n = std::min(capacity, t->picks);
    if (n) HIP_OK(hipMemcpy(slots, t->d_pick_slots.ptr, (size_t)n * 4, hipMemcpyDeviceToHost));
    *count = t->picks;
    return 0;
}

int vgi_read_selection(vgi_tracer *t, uint32_t *words, uint32_t count) {
    if (!t || !words) return fail("null argument");
    if (!t->select_recorded) return fail("no pick batch yet");
    if (count > vgi::SEL_WORDS) return fail("more words than the selection scratch holds");
    HIP_OK(hipStreamSynchronize(t->stream));
    HIP_OK(hipMemcpy(words, t->d_selection.ptr, (size_t)count * 4, hipMemcpyDeviceToHost));
    return 0;
}

int vgi_read_keys(vgi_tracer *t, uint32_t first, uint32_t count, uint32_t *keys) {
    if (!t || (count && !keys)) return fail("null argument");
    if ((size_t)first + count > t->high_water) return fail("slot range beyond the high-water mark");
    HIP_OK(hipStreamSynchronize(t->stream));
    if (count) HIP_OK(hipMemcpy(keys, t->d_keys.as<uint32_t>() + first, (size_t)count * 4, hipMemcpyDeviceToHost));
    return 0;
}

int vgi_read_mirror(vgi_tracer *t, uint32_t *words, uint32_t capacity, uint32_t *count) {
    if (!t || !count || (capacity && !words)) return fail("null argument");
    if (!t->select_recorded) return fail("no pick batch yet");
    HIP_OK(hipStreamSynchronize(t->stream));
    uint32_t emitted = 0;
    HIP_OK(hipMemcpy(&emitted, t->d_selection.as<uint32_t>() + vgi::SEL_MIRRORS, 4, hipMemcpyDeviceToHost));
    emitted = std::min(emitted, t->mirror_budget);
    uint32_t n = std::min(capacity, emitted);
    if (n) HIP_OK(hipMemcpy(words, t->d_mirror.ptr, (size_t)n * 112, hipMemcpyDeviceToHost));
    *count = emitted;
    return 0;
}

int vgi_cache_meta(vgi_tracer *t, uint32_t first, uint32_t count, uint32_t *out) {
    if (!t || !out) return fail("null argument");
    if ((size_t)first + count > t->high_water) return fail("slot range beyond the high-water mark");
    HIP_OK(hipStreamSynchronize(t->stream));
    HIP_OK(hipMemcpy(out, t->d_meta.as<uint32_t>() + (size_t)first * 12, (size_t)count * 48, hipMemcpyDeviceToHost));
    return 0;
}

int vgi_cache_state(vgi_tracer *t, uint32_t first, uint32_t count, uint32_t *out) {
    if (!t || !out) return fail("null argument");
    if ((size_t)first + count > t->high_water) return fail("slot range beyond the high-water mark");
    HIP_OK(hipStreamSynchronize(t->stream));
    HIP_OK(hipMemcpy(out, t->d_state.as<uint32_t>() + (size_t)first * 36, (size_t)count * 144, hipMemcpyDeviceToHost));
    return 0;
}

int vgi_probe_copy(vgi_tracer *t, uint64_t bytes, uint32_t repeats, float *ms) {
    if (!t) return fail("null tracer");
    HIP_OK(hipSetDevice(t->device));
    DeviceBuffer a, b;
    HIP_OK(a.reserve(bytes));
    HIP_OK(b.reserve(bytes));
    HIP_OK(hipMemsetAsync(a.ptr, 1, bytes, t->stream));
    HIP_OK(hipEventRecord(t->ev[6], t->stream));
    for (uint32_t r = 0; r < repeats; ++r) vgi::launch_copy(a.ptr, b.ptr, bytes, t->stream);
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord(t->ev[7], t->stream));
    HIP_OK(hipStreamSynchronize(t->stream));
    float total = 0.0f;
    hipEventElapsedTime(&total, t->ev[6], t->ev[7]);
    if (ms) *ms = repeats ? total / (float)repeats : 0.0f;
    a.release();
    b.release();
    return 0;
}

int vgi_scene_bounds(vgi_tracer *t, int32_t out[6]) {
    if (!t || !out) return fail("null argument");
    if (!t->have_regions) return 0;
    out[0] = t->bmin.x; out[1] = t->bmin.y; out[2] = t->bmin.z; out[3] = t->bmax.x; out[4] = t->bmax.y; out[5] = t->bmax.z;
    return 1;
}

int vgi_scene_region_canonical(vgi_tracer *t, int32_t rx, int32_t ry, int32_t rz, uint32_t *out) {
    if (!t || !out) return fail("null argument");
    std::memset(out, 0, (4 + 64 + 64 * 32) * 4);
    if (!t->have_regions || rx < t->bmin.x || ry < t->bmin.y || rz < t->bmin.z || rx > t->bmax.x || ry > t->bmax.y || rz > t->bmax.z)
        return 0;
    HIP_OK(hipStreamSynchronize(t->stream));
    size_t i = lin3(rx - t->rmin.x, ry - t->rmin.y, rz - t->rmin.z, t->rsx, t->rsy);
    uint32_t h[4];
    HIP_OK(hipMemcpy(h, t->d_headers.as<uint4>() + i, 16, hipMemcpyDeviceToHost));
    if (h[0] == 0) return 0;
    out[0] = h[0]; out[1] = h[0] == 3 ? 0u : h[1]; out[2] = h[2]; out[3] = h[3];
    if (h[0] != 3) return 1;
    uint32_t table[64];
    HIP_OK(hipMemcpy(table, t->d_cells.as<uint32_t>() + h[1], 256, hipMemcpyDeviceToHost));
    for (int b = 0; b < 64; ++b) {
        if ((table[b] & 3u) == 2u) {
            out[4 + b] = 2u;
            HIP_OK(hipMemcpy(&out[68 + b * 32], t->d_cells.as<uint32_t>() + (table[b] >> 2), 128, hipMemcpyDeviceToHost));
        } else {
            out[4 + b] = table[b];
        }
    }
    return 1;
}

int64_t vgi_scene_stats(vgi_tracer *t, int64_t out[8]) {
    if (!t || !out) return fail("null argument");
    int64_t loaded = 0, solid = 0;
    for (size_t i = 0; i < t->loaded.size(); ++i) { loaded += t->loaded[i]; solid += t->loaded[i] && t->solid[i]; }
    out[0] = loaded; out[1] = solid; out[2] = t->cells_used; out[3] = (int64_t)t->top.fields.size();
    out[4] = t->high_water; out[5] = (int64_t)t->emitters.faces; out[6] = (int64_t)t->emitters.nodes;
    out[7] = ((int64_t)t->stage_grows << 32) | t->stage_wraps;
    return loaded;
}

int64_t vgi_emitter_words(vgi_tracer *t, uint32_t *out, int64_t capacity) {
    if (!t) return fail("null tracer");
    if (sync_emitters(*t) != 0) return -1;
    int64_t n = (int64_t)t->emitters.words.size();
    if (out && capacity >= n) std::memcpy(out, t->emitters.words.data(), (size_t)n * 4);
    return n;
}

// ================================================================ batched frames
// Many tracers (scenes, or chunks of one scene bound to several tracers) in
// one persistent launch per frame, captured once into a HIP graph and
// replayed; workgroup size, lanes per face and queue tile come from
// vgi_batch_configure or the autotuner.
}  // extern "C"

struct vgi_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[2] = {};
    std::vector<vgi_tracer *> scenes;
    DeviceBuffer d_scenes, d_tiles, d_control, d_done, d_finish;
    uint32_t block = 256, lanes = 1, tile = 1024, grid = 0;
    uint32_t samples = 0, n_tiles = 0;
    uint32_t next_sequence = 0;
    bool sequence_dirty = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool dirty = true;
    float last_ms = 0.0f;
    int clock_khz = 0, cus = 0;
    std::vector<uint32_t> order;  // scenes by ascending face count
};

namespace {

void drop_graph(vgi_batch &b) {
    if (b.exec) hipGraphExecDestroy(b.exec);
    if (b.graph) hipGraphDestroy(b.graph);
    b.exec = nullptr;
    b.graph = nullptr;
}

// Lane counts that keep every face's sum in the reference's order: one
// lane, at least one lane per sample, or the reference's own count.
bool lanes_allowed(uint32_t lanes, uint32_t samples, uint32_t reference) {
    if (lanes != 1u && lanes != 2u && lanes != 4u && lanes != 8u) return false;
    return lanes == 1u || lanes >= samples || lanes == reference;
}

// Descriptors, tile queue and the captured frame graph.
int batch_prepare(vgi_batch &b) {
    if (!b.dirty && b.exec) return 0;
    drop_graph(b);
    uint32_t n = (uint32_t)b.scenes.size();
    if (n == 0) return fail("the batch holds no scenes");
    std::vector<vgi::BatchScene> desc(n);
    b.order.resize(n);
    std::iota(b.order.begin(), b.order.end(), 0u);
    std::stable_sort(b.order.begin(), b.order.end(), [&](uint32_t a, uint32_t c) { return b.scenes[a]->bound < b.scenes[c]->bound; });
    std::vector<vgi::BatchTile> tiles;
    for (uint32_t k : b.order) {
        vgi_tracer &t = *b.scenes[k];
        if (!t.have_regions) return fail("a batch scene has no scene loaded");
        if (sync_emitters(t) != 0) return -1;
        views(t, desc[k].scene, desc[k].optics);
        desc[k].tasks = t.b_tasks.as<uint4>();
        desc[k].results = t.b_results.as<float4>();
        desc[k].order = order_of(t);
        desc[k].count = t.bound;
        uint32_t first_tile = (uint32_t)tiles.size();
        for (uint32_t f = 0; f < t.bound; f += b.tile) tiles.push_back(vgi::BatchTile{k, f, std::min(b.tile, t.bound - f), 0u});
        desc[k].tiles = (uint32_t)tiles.size() - first_tile;
    }
    b.n_tiles = (uint32_t)tiles.size();
    HIP_OK(b.d_scenes.reserve(desc.size() * sizeof(vgi::BatchScene)));
    HIP_OK(b.d_tiles.reserve(std::max<size_t>(tiles.size(), 1) * sizeof(vgi::BatchTile)));
    HIP_OK(b.d_done.reserve((size_t)n * 4));
    HIP_OK(b.d_finish.reserve((size_t)n * 8));
    HIP_OK(b.d_control.reserve(sizeof(vgi::BatchControl)));
    HIP_OK(hipMemcpyAsync(b.d_scenes.ptr, desc.data(), desc.size() * sizeof(vgi::BatchScene), hipMemcpyHostToDevice, b.stream));
    if (!tiles.empty())
        HIP_OK(hipMemcpyAsync(b.d_tiles.ptr, tiles.data(), tiles.size() * sizeof(vgi::BatchTile), hipMemcpyHostToDevice, b.stream));
    HIP_OK(hipStreamSynchronize(b.stream));
    uint32_t grid = b.grid;
    if (grid == 0) grid = (uint32_t)std::max(1, b.cus) * (4u * VGI_TRACE_WAVES) / (b.block / 64u);
    hipGraph_t graph = nullptr;
    HIP_OK(hipStreamBeginCapture(b.stream, hipStreamCaptureModeThreadLocal));
    vgi::launch_batch_frame(b.d_scenes.as<vgi::BatchScene>(), n, b.d_tiles.as<vgi::BatchTile>(), b.n_tiles, b.lanes, b.block, grid,
                            b.d_control.as<vgi::BatchControl>(), b.d_done.as<uint32_t>(), b.d_finish.as<unsigned long long>(), b.stream);
    hipError_t launch = hipGetLastError();
    hipError_t end = hipStreamEndCapture(b.stream, &graph);
    if (launch != hipSuccess) { if (graph) hipGraphDestroy(graph); return fail(std::string("batch capture: ") + hipGetErrorString(launch)); }
    HIP_OK(end);
    b.graph = graph;
    HIP_OK(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
    b.dirty = false;
    return 0;
}

}  // namespace

extern "C" {

int vgi_batch_create(int device_ordinal, vgi_batch **out) {
    if (!out) return fail("null output");
    int count = 0;
    HIP_OK(hipGetDeviceCount(&count));
    if (device_ordinal < 0 || device_ordinal >= count) return fail("no such HIP device");
    HIP_OK(hipSetDevice(device_ordinal));
    std::unique_ptr<vgi_batch> b(new vgi_batch());
    b->device = device_ordinal;
    HIP_OK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    for (auto &e : b->ev) HIP_OK(hipEventCreate(&e));
    HIP_OK(hipDeviceGetAttribute(&b->clock_khz, hipDeviceAttributeWallClockRate, device_ordinal));
    HIP_OK(hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, device_ordinal));
    *out = b.release();
    return 0;
}

void vgi_batch_destroy(vgi_batch *b) {
    if (!b) return;
    hipSetDevice(b->device);
    hipStreamSynchronize(b->stream);
    drop_graph(*b);
    b->d_scenes.release(); b->d_tiles.release(); b->d_control.release(); b->d_done.release(); b->d_finish.release();
    for (auto &e : b->ev) if (e) hipEventDestroy(e);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
}

int vgi_batch_add(vgi_batch *b, vgi_tracer *t) {
    if (!b || !t) return fail("null argument");
    if (t->device != b->device) return fail("a batch scene lives on another device");
    if (t->bound == 0) return fail("bind the scene's faces before adding it to a batch");
    uint32_t samples = t->transport.samples;
    if (!b->scenes.empty() && samples != b->samples) return fail("every scene of a batch traces the same samples per face");
    if (!lanes_allowed(b->lanes, samples, lanes_for(t->transport))) b->lanes = 1u;
    b->samples = samples;
    b->scenes.push_back(t);
    b->dirty = true;
    return 0;
}

int vgi_batch_configure(vgi_batch *b, uint32_t block, uint32_t lanes, uint32_t tile, uint32_t grid) {
    if (!b) return fail("null batch");
    if (block != 64u && block != 128u && block != 256u) return fail("workgroup size must be 64, 128 or 256");
    if (tile == 0 || tile > (1u << 24)) return fail("queue tile must hold 1 to 2^24 faces");
    for (vgi_tracer *t : b->scenes)
        if (!lanes_allowed(lanes, t->transport.samples, lanes_for(t->transport)))
            return fail("lanes per face must be 1, at least the samples per face, or the reference's lane count");
    if (block % lanes) return fail("lanes per face must divide the workgroup");
    b->block = block; b->lanes = lanes; b->tile = tile; b->grid = grid;
    b->dirty = true;
    return 0;
}

// The next frame's sequence base (then advancing by the samples per face
// each frame, on the device).
int vgi_batch_set_sequence(vgi_batch *b, uint32_t base) {
    if (!b) return fail("null batch");
    b->next_sequence = base;
    b->sequence_dirty = true;
    return 0;
}

int vgi_batch_run_frame(vgi_batch *b) {
    if (!b) return fail("null batch");
    HIP_OK(hipSetDevice(b->device));
    if (batch_prepare(*b) != 0) return -1;
    if (b->sequence_dirty) {
        vgi::BatchControl c{};
        c.next_seq = b->next_sequence;
        c.samples = b->samples;
        HIP_OK(hipMemcpyAsync(b->d_control.ptr, &c, sizeof(c), hipMemcpyHostToDevice, b->stream));
        b->sequence_dirty = false;
    }
    for (vgi_tracer *t : b->scenes) {
        t->bound_moves.clear();
        t->frame_bound = t->bound;
        t->bound_hit_words = 0;
    }
    HIP_OK(hipEventRecord(b->ev[0], b->stream));
    HIP_OK(hipGraphLaunch(b->exec, b->stream));
    HIP_OK(hipEventRecord(b->ev[1], b->stream));
    return 0;
}

int vgi_batch_sync(vgi_batch *b) {
    if (!b) return fail("null batch");
    HIP_OK(hipStreamSynchronize(b->stream));
    HIP_OK(hipEventElapsedTime(&b->last_ms, b->ev[0], b->ev[1]));
    return 0;
}

float vgi_batch_last_ms(vgi_batch *b) { return b ? b->last_ms : 0.0f; }

// Per scene (in the order added): ms from the frame's start to its last
// tile's end; rays traced in the frame.
int vgi_batch_latency(vgi_batch *b, float *out_ms, uint32_t count, uint64_t *rays) {
    if (!b) return fail("null batch");
    if (count > b->scenes.size()) return fail("more scenes than the batch holds");
    HIP_OK(hipStreamSynchronize(b->stream));
    vgi::BatchControl c{};
    HIP_OK(hipMemcpy(&c, b->d_control.ptr, sizeof(c), hipMemcpyDeviceToHost));
    if (rays) *rays = c.rays;
    if (count && out_ms) {
        std::vector<unsigned long long> finish(count);
        HIP_OK(hipMemcpy(finish.data(), b->d_finish.ptr, (size_t)count * 8, hipMemcpyDeviceToHost));
        for (uint32_t k = 0; k < count; ++k)
            out_ms[k] = finish[k] >= c.start ? (float)((double)(finish[k] - c.start) / (double)std::max(b->clock_khz, 1)) : -1.0f;
    }
    return 0;
}

// Try every workgroup size, allowed lane count and queue tile on `frames`
// timed frames each (after one warm-up) and keep the fastest by median GPU
// time. out: block, lanes, tile, ms per frame (as float bits), candidates.
int vgi_batch_autotune(vgi_batch *b, uint32_t frames, uint32_t out[5]) {
    if (!b || !out) return fail("null argument");
    if (b->scenes.empty()) return fail("the batch holds no scenes");
    frames = std::max<uint32_t>(frames, 1);
    const uint32_t blocks[] = {64u, 128u, 256u};
    const uint32_t lane_options[] = {1u, 2u, 4u, 8u};
    const uint32_t tile_options[] = {128u, 512u, 2048u, 8192u};
    float best = 1e30f;
    uint32_t best_cfg[3] = {b->block, b->lanes, b->tile}, tried = 0;
    uint32_t saved_sequence = b->next_sequence;
    for (uint32_t block : blocks)
        for (uint32_t lanes : lane_options)
            for (uint32_t tile : tile_options) {
                bool ok = block % lanes == 0;
                for (vgi_tracer *t : b->scenes) ok = ok && lanes_allowed(lanes, t->transport.samples, lanes_for(t->transport));
                if (!ok) continue;
                if (vgi_batch_configure(b, block, lanes, tile, 0) != 0) return -1;
                std::vector<float> ms;
                for (uint32_t f = 0; f <= frames; ++f) {
                    if (vgi_batch_run_frame(b) != 0 || vgi_batch_sync(b) != 0) return -1;
                    if (f > 0) ms.push_back(b->last_ms);
                }
                std::sort(ms.begin(), ms.end());
                float med = ms[ms.size() / 2];
                ++tried;
                if (med < best) { best = med; best_cfg[0] = block; best_cfg[1] = lanes; best_cfg[2] = tile; }
            }
    if (vgi_batch_configure(b, best_cfg[0], best_cfg[1], best_cfg[2], 0) != 0) return -1;
    vgi_batch_set_sequence(b, saved_sequence);
    out[0] = best_cfg[0]; out[1] = best_cfg[1]; out[2] = best_cfg[2];
    std::memcpy(&out[3], &best, 4);
    out[4] = tried;
    return 0;
}

}  // extern "C"
