// Internal host-side declarations shared by the C ABI and the kernel units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "cantucci_hip.h"
#include "de_device.h"

namespace cth {

// Work counters: a ring of {next, done} pairs, zeroed at context creation.
// Each launch takes the next pair, and the launch's last wave zeroes it again
// (work_loop.h), so launches on different streams never share a pair unless
// more than COUNTER_RING launches are in flight at once.  Launches recorded
// into a graph through cth_graph_begin take their pairs from the graph's own
// pool instead (CapturePool), since a graph replays with the pairs it was
// captured with for as long as it lives.
constexpr uint32_t COUNTER_RING = 4096;

// Counter pairs owned by the graph being captured on this thread: launches
// recorded on `stream` take pairs from `pairs` (zeroed before the capture
// began), cycling through them.  Launches of one graph on its one stream run
// one after another, and each leaves its pair at zero, so a pair can serve
// several of them.
constexpr uint32_t GRAPH_PAIRS = 64;
struct CapturePool {
    hipStream_t stream = nullptr;
    uint32_t* pairs = nullptr;  // 2 * GRAPH_PAIRS, device
    uint32_t pos = 0;
};
CapturePool& capture_pool();  // this thread's (capi.hip)

struct LaunchShape {
    int block = 256;           // threads per workgroup (64, 128 or 256)
    int blocks_per_cu = 16;    // persistent workgroups per CU
    uint32_t chunk = 256;      // indices per chunk
    uint32_t refill_min = 32;  // idle lanes that trigger a refill (1..64)
    uint32_t dynamic = 0;      // 0: static interleaved chunks, 1: atomic claims
};

// Device scratch slots of a context, grown on demand and kept until the
// context is destroyed (calls on one context never overlap, and the users
// below finish with a stream synchronisation).
enum ScratchSlot {
    WS_NODES_A = 0,
    WS_NODES_B,
    WS_FLAGS,
    WS_OFFS,
    WS_SCAN,
    WS_COUNTS,
    WS_UOFF,
    WS_HOST_IN,   // staging of the host-buffer entry points
    WS_HOST_OUT,
    WS_SLOTS
};

struct Ctx {
    int device = 0;
    int num_cus = 0;
    int mode = CTH_MODE_PARITY;
    hipStream_t stream = nullptr;
    uint32_t* counters = nullptr;  // 2 * COUNTER_RING
    uint32_t ring_pos = 0;
    void* ws[WS_SLOTS] = {};
    size_t ws_bytes[WS_SLOTS] = {};
    uint64_t* pinned = nullptr;  // 8 host-pinned words for count read-backs
    unsigned long long* stats = nullptr;  // work counters of the refill kernels (cth_ctx_count_work)
    // Slot `s` with at least `bytes` bytes (a larger buffer replaces the old
    // one after a device-wide wait, so no launch can still be using it).
    int scratch(int s, size_t bytes, void** out) {
        if (bytes > ws_bytes[s]) {
            if (ws[s]) {
                (void)hipDeviceSynchronize();
                (void)hipFree(ws[s]);
                ws[s] = nullptr;
                ws_bytes[s] = 0;
            }
            size_t want = bytes + bytes / 4 + 256;
            if (hipMalloc(&ws[s], want) != hipSuccess) {
                ws[s] = nullptr;
                return CTH_E_NOMEM;
            }
            ws_bytes[s] = want;
        }
        *out = ws[s];
        return CTH_OK;
    }
    LaunchShape shape_de;
    LaunchShape shape_spans{256, 8, 1024, 32, 1};
    // The counter pair of a launch on stream `st`: the capturing graph's own
    // pairs while cth_graph_begin captures `st` on this thread, else the next
    // pair of the ring.
    uint32_t* take_counters(hipStream_t st) {
        CapturePool& cp = capture_pool();
        if (cp.pairs && cp.stream == st) return cp.pairs + 2 * (cp.pos++ % GRAPH_PAIRS);
        uint32_t i = __atomic_fetch_add(&ring_pos, 1u, __ATOMIC_RELAXED) % COUNTER_RING;
        return counters + 2 * i;
    }
    LaunchShape shape_trace{256, 16, 256, 32, 0};
    LaunchShape shape_adj{256, 16, 256, 32, 0};
    LaunchShape shape_jobs{256, 8, 256, 32, 0};
    LaunchShape shape_mesh{256, 8, 1024, 32, 1};
    LaunchShape* launch_shape(int kernel) {
        switch (kernel) {
            case CTH_KERNEL_DE: return &shape_de;
            case CTH_KERNEL_TRACE: return &shape_trace;
            case CTH_KERNEL_ADJOINT: return &shape_adj;
            case CTH_KERNEL_SPANS: return &shape_spans;
            case CTH_KERNEL_JOBS: return &shape_jobs;
            case CTH_KERNEL_MESH: return &shape_mesh;
            default: return nullptr;
        }
    }
};

// Status helpers (capi.hip)
int fail(int code, const std::string& msg);
int check_hip(hipError_t e, const char* what);

// Resolve a shape descriptor into kernel uniforms.  Returns CTH_OK or an error.
int resolve_mandel(const cth_shape& s, MandelUniform& u, int& family);

// Family codes used by the launchers.
enum Family { FAM_P8 = 0, FAM_TRIG = 1, FAM_SPHERE = 2 };

// Kernel launchers (kernels_de.hip).  All indices per launch are < 2^31.
int launch_de_points(Ctx& c, const cth_shape& s, const float* pts, const float* powers, uint32_t n, float* out,
                     uint32_t* iters, float* r_out, float* dr_out, hipStream_t st);
int launch_de_grid(Ctx& c, const cth_shape& s, const float* start, const float* end, uint32_t n, float divisor,
                   float* out, uint32_t* iters, hipStream_t st);
int launch_sample_spans(Ctx& c, const cth_shape& s, const float* spans, uint32_t nspans, uint32_t res,
                        float* dists, uint64_t* cross, uint32_t* iters, hipStream_t st);
int refine_levels(Ctx& c, const cth_shape& s, const float* root, uint32_t min_depth, uint32_t max_depth,
                  const float* half_diag, float* leaves, size_t cap, size_t* nleaves, uint64_t* level_counts,
                  hipStream_t st);
int launch_trace_camera(Ctx& c, const cth_shape& s, const cth_camera* cams, uint32_t ncams, uint32_t W, uint32_t H,
                        uint32_t max_steps, float eps_abs, float eps_rel, float t_max, float* depth,
                        uint32_t* steps, hipStream_t st);
int launch_trace_focii(Ctx& c, const cth_shape& s, const cth_ref_camera* cams, uint32_t ncams, uint32_t fp,
                       uint32_t max_steps, float eps, float* focii, float* depth, uint32_t* steps, float* dirs,
                       hipStream_t st);
int launch_trace_rays(Ctx& c, const cth_shape& s, const float* origins, int per_ray_origin, const float* dirs,
                      uint32_t n, uint32_t max_steps, float eps_abs, float eps_rel, float t_max, float* depth,
                      uint32_t* steps, float* hit_pos, hipStream_t st);
// Reverse-sweep VJP.  workspace NULL: stream-ordered allocations; else the
// caller's ws_bytes of device memory (at least vjp_workspace_bytes).
int launch_adjoint(Ctx& c, const cth_shape& s, const float* pts, const float* powers, uint32_t n,
                   const float* grad_d, float* d, float* grad_p, float* grad_power, float* normals, hipStream_t st,
                   void* workspace = nullptr, size_t ws_bytes = 0, float* grad_z0 = nullptr,
                   float* grad_c = nullptr);
// The allocation-free VJP's device scratch (kernels_adjoint.hip): its size,
// the per-device set-up (hipGetSymbolAddress with `device` current, done by
// cth_ctx_create) and the device address it recorded.
constexpr size_t kVjpScratchBytes = size_t(16) << 20;
int vjp_scratch_init(int device);
void* vjp_scratch(int device);
size_t vjp_workspace_bytes(uint32_t max_iters, size_t n);
int launch_tangent(Ctx& c, const cth_shape& s, const float* pts, const float* powers, uint32_t n,
                   const float* grad_d, float* d, float* grad_p, float* grad_power, float* normals, hipStream_t st);
int launch_fd_normals(Ctx& c, const cth_shape& s, const float* pts, const float* powers, uint32_t n, float h,
                      float* d, float* grad_p, float* normals, hipStream_t st);
struct JobSet;
int jobs_prepare(Ctx& c, int kind, float bailout, const cth_job* jobs, uint32_t njobs, JobSet** out);
void jobs_destroy(JobSet* js);
uint32_t jobs_total(const JobSet* js);
uint32_t jobs_offset(const JobSet* js, uint32_t i);
int jobs_run(Ctx& c, const JobSet& js, float* dists, float* normals, unsigned long long* done,
             unsigned long long* start, hipStream_t st);
int mesh_count(Ctx& c, const float* dists, uint32_t nleaves, uint32_t res, uint64_t* voff, uint64_t* ioff,
               uint64_t* cross, uint64_t totals[2], hipStream_t st);
int mesh_emit(Ctx& c, const cth_shape& s, const float* spans, const float* dists, uint32_t nleaves, uint32_t res,
              int normals, const uint64_t* voff, const uint64_t* ioff, uint64_t nverts, float* verts,
              uint32_t* idx, hipStream_t st);
int mesh_vertices(Ctx& c, const cth_shape& s, int normals, uint64_t nverts, float* verts, hipStream_t st);
int launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t st);
int launch_focus_split(Ctx& c, const float* leaves, uint32_t n, const float* focii, const float* depth, uint32_t nf,
                       const float eye[3], uint32_t* flags, hipStream_t st);
int split_leaves(Ctx& c, const float* leaves, uint32_t n, const uint32_t* flags, float* out, size_t cap,
                 size_t* nout, hipStream_t st);
int launch_contains_sphere(Ctx& c, const cth_shape& s, const float* pts, uint32_t n, uint8_t* out, hipStream_t st);
int launch_less_than_zero(Ctx& c, const float* d, uint32_t n, uint8_t* out, hipStream_t st);
int launch_dup_pairs(Ctx& c, const float* d, uint32_t n, float* pairs, hipStream_t st);

}  // namespace cth
