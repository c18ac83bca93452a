// C ABI: context management, argument validation and the Shape trait mirror.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "internal.h"

namespace cth {

static thread_local std::string g_last_error;
static thread_local CapturePool g_capture;

CapturePool& capture_pool() { return g_capture; }

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int check_hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return CTH_OK;
    return fail(e == hipErrorOutOfMemory ? CTH_E_NOMEM : CTH_E_DEVICE,
                std::string(what) + ": " + hipGetErrorString(e));
}

// `Mandelbulb::new` asserts max_iters >= 1 (src/shape/mandelbulb.rs:20); the
// power of `Mandelbulb<P>` is a u8 const generic (:13).
int resolve_mandel(const cth_shape& s, MandelUniform& u, int& family) {
    if (s.max_iters == 0) return fail(CTH_E_INVALID, "max_iters must be >= 1");
    u.max_iters = s.max_iters;
    u.bailout = s.bailout;
    u.power = s.power;
    if (s.kind == CTH_SHAPE_MANDELBULB) {
        float p = s.power;
        if (!(p >= 1.0f && p <= 255.0f) || std::floor(p) != p)
            return fail(CTH_E_INVALID, "Mandelbulb power must be an integer in [1, 255]");
        u.ipower = (int)p;
        u.integral = 1;
        family = (u.ipower == 8) ? FAM_P8 : FAM_TRIG;
        return CTH_OK;
    }
    if (s.kind == CTH_SHAPE_MANDELBULB_TRIG) {
        if (!std::isfinite(s.power)) return fail(CTH_E_INVALID, "power must be finite");
        u.ipower = 0;
        u.integral = 0;
        family = FAM_TRIG;
        return CTH_OK;
    }
    return fail(CTH_E_INVALID, "not a Mandelbulb shape");
}

static int check_shape(const cth_shape* s) {
    if (!s) return fail(CTH_E_INVALID, "null shape");
    if (s->kind == CTH_SHAPE_SPHERE) return CTH_OK;
    MandelUniform u;
    int fam;
    return resolve_mandel(*s, u, fam);
}

// `_dev` entry points run on the caller's stream; NULL is HIP's default
// (null) stream, as everywhere in HIP.  The context's own stream serves the
// host-pointer entry points.
static hipStream_t pick_stream(Ctx*, void* stream) { return (hipStream_t)stream; }

// RAII device buffer for the host-pointer entry points.
// Device staging for the host-buffer entry points: sub-buffers (256-byte
// aligned) of one of the context's scratch slots, so that a small query --
// one point per frame for the reference's fly camera -- allocates nothing.
template <size_t N>
int staging(Ctx& c, int slot, const size_t (&bytes)[N], void* (&out)[N]) {
    size_t off[N], total = 0;
    for (size_t i = 0; i < N; ++i) {
        off[i] = total;
        total += (bytes[i] + 255) & ~size_t(255);
    }
    void* base = nullptr;
    if (c.scratch(slot, total ? total : 256, &base)) return fail(CTH_E_NOMEM, "staging allocation failed");
    for (size_t i = 0; i < N; ++i) out[i] = static_cast<char*>(base) + off[i];
    return CTH_OK;
}

}  // namespace cth

using namespace cth;

struct cth_ctx : public Ctx {};

extern "C" {

const char* cth_last_error(void) { return g_last_error.c_str(); }

int cth_ctx_create(int device, cth_ctx** out) {
    if (!out) return fail(CTH_E_INVALID, "null output pointer");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) return fail(CTH_E_DEVICE, "no HIP device available");
    if (device < 0 || device >= count) return fail(CTH_E_INVALID, "device index out of range");
    hipDeviceProp_t prop;
    if (int rc = check_hip(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties")) return rc;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(CTH_E_DEVICE, std::string("library is built for gfx950, device is ") + prop.gcnArchName);
    if (int rc = check_hip(hipSetDevice(device), "hipSetDevice")) return rc;
    cth_ctx* c = new cth_ctx();
    c->device = device;
    c->num_cus = prop.multiProcessorCount;
    int rc = check_hip(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
    if (rc == CTH_OK) rc = check_hip(hipMalloc((void**)&c->counters, 2 * COUNTER_RING * sizeof(uint32_t)), "hipMalloc");
    if (rc == CTH_OK)
        rc = check_hip(hipMemset(c->counters, 0, 2 * COUNTER_RING * sizeof(uint32_t)), "hipMemset counters");
    if (rc == CTH_OK) rc = check_hip(hipHostMalloc((void**)&c->pinned, 8 * sizeof(uint64_t)), "hipHostMalloc");
    if (rc == CTH_OK) rc = vjp_scratch_init(device);
    if (rc == CTH_OK) rc = check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
    if (rc != CTH_OK) {
        cth_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return CTH_OK;
}

void cth_ctx_destroy(cth_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->counters) (void)hipFree(c->counters);
    for (void* p : c->ws)
        if (p) (void)hipFree(p);
    if (c->pinned) (void)hipHostFree(c->pinned);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int cth_ctx_set_mode(cth_ctx* c, int mode) {
    if (!c) return fail(CTH_E_INVALID, "null context");
    if (mode != CTH_MODE_PARITY && mode != CTH_MODE_FAST && mode != CTH_MODE_CONTRACT)
        return fail(CTH_E_INVALID, "unknown mode");
    c->mode = mode;
    return CTH_OK;
}

int cth_ctx_get_mode(const cth_ctx* c) { return c ? c->mode : -1; }

int cth_ctx_set_launch(cth_ctx* c, int kernel, const cth_launch* l) {
    if (!c || !l) return fail(CTH_E_INVALID, "null argument");
    LaunchShape* ls = c->launch_shape(kernel);
    if (!ls) return fail(CTH_E_INVALID, "unknown kernel family");
    if (!(l->block == 64 || l->block == 128 || l->block == 256)) return fail(CTH_E_INVALID, "block must be 64, 128 or 256");
    if (l->blocks_per_cu < 1 || l->blocks_per_cu > 32) return fail(CTH_E_INVALID, "blocks_per_cu must be in [1, 32]");
    if (l->chunk < 64 || l->chunk > 65536) return fail(CTH_E_INVALID, "chunk must be in [64, 65536]");
    if (l->refill_min < 1 || l->refill_min > 64) return fail(CTH_E_INVALID, "refill_min must be in [1, 64]");
    ls->block = l->block;
    ls->blocks_per_cu = l->blocks_per_cu;
    ls->chunk = l->chunk;
    ls->refill_min = l->refill_min;
    ls->dynamic = l->dynamic ? 1u : 0u;
    return CTH_OK;
}

int cth_ctx_get_launch(const cth_ctx* c, int kernel, cth_launch* l) {
    if (!c || !l) return fail(CTH_E_INVALID, "null argument");
    LaunchShape* ls = const_cast<cth_ctx*>(c)->launch_shape(kernel);
    if (!ls) return fail(CTH_E_INVALID, "unknown kernel family");
    l->block = ls->block;
    l->blocks_per_cu = ls->blocks_per_cu;
    l->chunk = ls->chunk;
    l->refill_min = ls->refill_min;
    l->dynamic = ls->dynamic;
    return CTH_OK;
}

int cth_ctx_count_work(cth_ctx* c, uint64_t* counters) {
    if (!c) return fail(CTH_E_INVALID, "null context");
    c->stats = reinterpret_cast<unsigned long long*>(counters);
    return CTH_OK;
}

void* cth_ctx_stream(cth_ctx* c) { return c ? (void*)c->stream : nullptr; }

int cth_ctx_num_cus(const cth_ctx* c) { return c ? c->num_cus : 0; }

int cth_synchronize(cth_ctx* c) {
    if (!c) return fail(CTH_E_INVALID, "null context");
    return check_hip(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
}

int cth_bounding_box(const cth_shape* s, cth_span* out) {
    if (!s || !out) return fail(CTH_E_INVALID, "null argument");
    if (s->kind == CTH_SHAPE_SPHERE) {
        // center + -off .. center + off  (src/shape/sphere.rs:27-30)
        for (int a = 0; a < 3; ++a) {
            out->start[a] = s->center[a] + -s->radius;
            out->end[a] = s->center[a] + s->radius;
        }
        return CTH_OK;
    }
    if (int rc = check_shape(s)) return rc;
    for (int a = 0; a < 3; ++a) {
        out->start[a] = -1.2f;
        out->end[a] = 1.2f;
    }
    return CTH_OK;
}

int cth_min_distance_dev(cth_ctx* c, const cth_shape* s, const float* pts, const float* powers, size_t n, float* out,
                         uint32_t* iters, void* stream) {
    if (!c) return fail(CTH_E_INVALID, "null context");
    if (int rc = check_shape(s)) return rc;
    if (n && (!pts || !out)) return fail(CTH_E_INVALID, "null buffer");
    hipStream_t st = pick_stream(c, stream);
    const size_t B = size_t(1) << 30;
    for (size_t off = 0; off < n; off += B) {
        uint32_t m = (uint32_t)((n - off) < B ? (n - off) : B);
        int rc = launch_de_points(*c, *s, pts + 3 * off, powers ? powers + off : nullptr, m, out + off,
                                  iters ? iters + off : nullptr, nullptr, nullptr, st);
        if (rc) return rc;
    }
    return CTH_OK;
}

int cth_de_state_dev(cth_ctx* c, const cth_shape* s, const float* pts, const float* powers, size_t n, float* r,
                     float* dr, uint32_t* iters, void* stream) {
    if (!c) return fail(CTH_E_INVALID, "null context");
    if (int rc = check_shape(s)) return rc;
    if (s->kind == CTH_SHAPE_SPHERE) return fail(CTH_E_UNSUPPORTED, "a sphere has no iteration state");
    if (n && (!pts || !r || !dr || !iters)) return fail(CTH_E_INVALID, "null buffer");
    if (n >= (size_t(1) << 31)) return fail(CTH_E_INVALID, "batch larger than 2^31");
    return launch_de_points(*c, *s, pts, powers, (uint32_t)n, nullptr, iters, r, dr, pick_stream(c, stream));
}

int cth_de_grid_dev(cth_ctx* c, const cth_shape* s, const float start[3], const float end[3], uint32_t n,
                    float divisor, float* out, uint32_t* iters, void* stream) {
    if (!c) return fail(CTH_E_INVALID, "null context");
    if (int rc = check_shape(s)) return rc;
    if (!start || !end || (n && !out)) return fail(CTH_E_INVALID, "null buffer");
    if (!(divisor > 0.0f)) return fail(CTH_E_INVALID, "divisor must be positive");
    return launch_de_grid(*c, *s, start, end, n, divisor, out, iters, pick_stream(c, stream));
}

int cth_sample_spans_dev(cth_ctx* c, const cth_shape* s, const float* spans, size_t nspans, uint32_t res,
                         float* dists, uint64_t* cross_bits, uint32_t* iters, void* stream) {
    if (!c) return fail(CTH_E_INVALID, "null context");
    if (int rc = check_shape(s)) return rc;
    if (nspans && (!spans || !dists)) return fail(CTH_E_INVALID, "null buffer");
    if (res == 0 || (res & (res - 1)) != 0) return fail(CTH_E_INVALID, "resolution must be a power of two");
    if (res > 512) return fail(CTH_E_INVALID, "resolution above 512");
    hipStream_t st = pick_stream(c, stream);
    size_t per = size_t(res + 1) * (res + 1) * (res + 1);
    size_t words = (size_t(res) * res * res + 63) / 64;
    size_t batch = ((size_t(1) << 30) / per);
    if (batch == 0) batch = 1;
    for (size_t off = 0; off < nspans; off += batch) {
        size_t m = (nspans - off) < batch ? (nspans - off) : batch;
        int rc = launch_sample_spans(*c, *s, spans + 6 * off, (uint32_t)m, res, dists + off * per,
                                     cross_bits ? cross_bits + off * words : nullptr,
                                     iters ? iters + off * per : nullptr, st);
        if (rc) return rc;
    }
    return CTH_OK;
}

int cth_refine_dev(cth_ctx* c, const cth_shape* s, const float root[6], uint32_t min_depth, uint32_t max_depth,
                   const float* half_diag, float* leaves, size_t cap, size_t* nleaves, uint64_t* level_counts,
                   void* stream) {
    if (!c) return fail(CTH_E_INVALID, "null context");
    if (int rc = check_shape(s)) return rc;
    if (!root || !half_diag || !nleaves) return fail(CTH_E_INVALID, "null argument");
    if (max_depth > 20) return fail(CTH_E_INVALID, "max_depth above 20");
    for (int a = 0; a < 3; ++a)
        if (!(root[a] < root[3 + a])) return fail(CTH_E_INVALID, "root span must have start < end");
    return refine_levels(*c, *s, root, min_depth, max_depth, half_diag, leaves, cap, nleaves, level_counts,
                         pick_stream(c, stream));
}

int cth_sphere_trace_dev(cth_ctx* c, const cth_shape* s, const float* origins, int per_ray_origin,
                         const float* dirs, size_t n, uint32_t max_steps, float eps_abs, float eps_rel, float t_max,
                         float* depth, uint32_t* steps, float* hit_pos, void* stream) {
    if (!c) return fail(CTH_E_INVALID, "null context");
    if (int rc = check_shape(s)) return rc;
    if (n && (!origins || !dirs || !depth || !steps)) return fail(CTH_E_INVALID, "null buffer");
    if (max_steps == 0) return fail(CTH_E_INVALID, "max_steps must be >= 1");
    if (n >= (size_t(1) << 31)) return fail(CTH_E_INVALID, "more than 2^31 rays");
    return launch_trace_rays(*c, *s, origins, per_ray_origin, dirs, (uint32_t)n, max_steps, eps_abs, eps_rel, t_max,
                             depth, steps, hit_pos, pick_stream(c, stream));
}

int cth_focus_split_dev(cth_ctx* c, const float* leaves, size_t n, const float* focii, const float* depth,
                        size_t nfocii, const float eye[3], uint32_t* flags, void* stream) {
    if (!c) return fail(CTH_E_INVALID, "null context");
    if (!eye || (n && (!leaves || !flags)) || (nfocii && !focii)) return fail(CTH_E_INVALID, "null buffer");
    if (n >= (size_t(1) << 31) || nfocii >= (size_t(1) << 31)) return fail(CTH_E_INVALID, "batch larger than 2^31");
    return launch_focus_split(*c, leaves, (uint32_t)n, focii, depth, (uint32_t)nfocii, eye, flags,
                              pick_stream(c, stream));
}

int cth_split_leaves_dev(cth_ctx* c, const float* leaves, size_t n, const uint32_t* flags, float* out, size_t cap,
                         size_t* nout, void* stream) {
    if (!c) return fail(CTH_E_INVALID, "null context");
    if (!nout || (n && (!leaves || !flags))) return fail(CTH_E_INVALID, "null buffer");
    if (n >= (size_t(1) << 29)) return fail(CTH_E_INVALID, "more than 2^29 leaves");
    return split_leaves(*c, leaves, (uint32_t)n, flags, out, cap, nout, pick_stream(c, stream));
}

int cth_sphere_trace(cth_ctx* c, const cth_shape* s, const float* origins, int per_ray_origin, const float* dirs,
                     size_t n, uint32_t max_steps, float eps_abs, float eps_rel, float t_max, float* depth,
                     uint32_t* steps, float* hit_pos) {
    if (!c) return fail(CTH_E_INVALID, "null context");
    if (int rc = check_shape(s)) return rc;
    if (n == 0) return CTH_OK;
    if (!origins || !dirs || !depth || !steps) return fail(CTH_E_INVALID, "null buffer");
    size_t no = per_ray_origin ? n : 1;
    void* b[5];
    int rc = staging(*c, WS_HOST_IN, {no * 12, n * 12, n * 4, n * 4, hit_pos ? n * 12 : 0}, b);
    if (rc) return rc;
    rc = check_hip(hipMemcpyAsync(b[0], origins, no * 12, hipMemcpyHostToDevice, c->stream), "H2D");
    if (!rc) rc = check_hip(hipMemcpyAsync(b[1], dirs, n * 12, hipMemcpyHostToDevice, c->stream), "H2D");
    if (!rc)
        rc = cth_sphere_trace_dev(c, s, (const float*)b[0], per_ray_origin, (const float*)b[1], n, max_steps,
                                  eps_abs, eps_rel, t_max, (float*)b[2], (uint32_t*)b[3],
                                  hit_pos ? (float*)b[4] : nullptr, c->stream);
    if (!rc) rc = check_hip(hipMemcpyAsync(depth, b[2], n * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
    if (!rc) rc = check_hip(hipMemcpyAsync(steps, b[3], n * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
    if (!rc && hit_pos) rc = check_hip(hipMemcpyAsync(hit_pos, b[4], n * 12, hipMemcpyDeviceToHost, c->stream), "D2H");
    int rs = check_hip(hipStreamSynchronize(c->stream), "sync");
    return rc ? rc : rs;
}

int cth_get_focii_dev(cth_ctx* c, const cth_shape* s, const cth_ref_camera* cams, uint32_t ncams,
                      uint32_t focus_points, uint32_t max_steps, float eps, float* focii, float* depth,
                      uint32_t* steps, float* dirs, void* stream) {
    if (!c) return fail(CTH_E_INVALID, "null context");
    if (int rc = check_shape(s)) return rc;
    if (focus_points > 255) return fail(CTH_E_INVALID, "focus_points must be in [0, 255]");
    if (max_steps == 0) return fail(CTH_E_INVALID, "max_steps must be >= 1");
    if (focus_points == 0 || ncams == 0) return CTH_OK;
    if (!cams || !focii || !depth || !steps) return fail(CTH_E_INVALID, "null buffer");
    return launch_trace_focii(*c, *s, cams, ncams, focus_points, max_steps, eps, focii, depth, steps, dirs,
                              pick_stream(c, stream));
}

int cth_get_focii(cth_ctx* c, const cth_shape* s, const cth_ref_camera* cam, uint32_t focus_points, float* focii,
                  size_t* nhits) {
    if (!c) return fail(CTH_E_INVALID, "null context");
    if (int rc = check_shape(s)) return rc;
    if (!cam || !nhits) return fail(CTH_E_INVALID, "null buffer");
    if (focus_points > 255) return fail(CTH_E_INVALID, "focus_points must be in [0, 255]");
    *nhits = 0;
    if (focus_points == 0) return CTH_OK;  // the reference's iter::square(0) is empty
    if (!focii) return fail(CTH_E_INVALID, "null buffer");
    const size_t n = size_t(focus_points) * focus_points;
    void* b[4];
    int rc = staging(*c, WS_HOST_IN, {sizeof(cth_ref_camera), n * 12, n * 4, n * 4}, b);
    if (rc) return rc;
    rc = check_hip(hipMemcpyAsync(b[0], cam, sizeof(cth_ref_camera), hipMemcpyHostToDevice, c->stream), "H2D");
    // get_focii's constants (src/mesh/mod.rs:206-207): EPSILON 1e-6, MAX_ITERS 100
    if (!rc)
        rc = cth_get_focii_dev(c, s, (const cth_ref_camera*)b[0], 1, focus_points, 100, 0.000001f, (float*)b[1],
                               (float*)b[2], (uint32_t*)b[3], nullptr, c->stream);
    std::vector<float> pos(3 * n), depth(n);
    if (!rc) rc = check_hip(hipMemcpyAsync(pos.data(), b[1], n * 12, hipMemcpyDeviceToHost, c->stream), "D2H");
    if (!rc) rc = check_hip(hipMemcpyAsync(depth.data(), b[2], n * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
    int rs = check_hip(hipStreamSynchronize(c->stream), "sync");
    if (rc || rs) return rc ? rc : rs;
    // filter_map: the rays that hit, in ray order (src/mesh/mod.rs:229-242)
    size_t k = 0;
    for (size_t i = 0; i < n; ++i) {
        if (depth[i] == __builtin_inff()) continue;
        focii[3 * k] = pos[3 * i];
        focii[3 * k + 1] = pos[3 * i + 1];
        focii[3 * k + 2] = pos[3 * i + 2];
        ++k;
    }
    *nhits = k;
    return CTH_OK;
}

int cth_render_depth_dev(cth_ctx* c, const cth_shape* s, const cth_camera* cams, uint32_t ncams, uint32_t width,
                         uint32_t height, uint32_t max_steps, float eps_abs, float eps_rel, float t_max, float* depth,
                         uint32_t* steps, void* stream) {
    if (!c) return fail(CTH_E_INVALID, "null context");
    if (int rc = check_shape(s)) return rc;
    if (ncams && width && height && (!cams || !depth || !steps)) return fail(CTH_E_INVALID, "null buffer");
    if (max_steps == 0) return fail(CTH_E_INVALID, "max_steps must be >= 1");
    if ((uint64_t)width * height == 0 || ncams == 0) return CTH_OK;
    return launch_trace_camera(*c, *s, cams, ncams, width, height, max_steps, eps_abs, eps_rel, t_max, depth, steps,
                               pick_stream(c, stream));
}

int cth_de_vjp_dev(cth_ctx* c, const cth_shape* s, const float* pts, const float* powers, size_t n,
                   const float* grad_d, float* d, float* grad_p, float* grad_power, void* stream) {
    if (!c) return fail(CTH_E_INVALID, "null context");
    if (int rc = check_shape(s)) return rc;
    if (n && !pts) return fail(CTH_E_INVALID, "null buffer");
    if (n >= (size_t(1) << 31)) return fail(CTH_E_INVALID, "batch larger than 2^31");
    hipStream_t st = pick_stream(c, stream);
    float* gp = grad_power;
    if (grad_power && s->kind != CTH_SHAPE_MANDELBULB_TRIG) {
        int rc = launch_fill_u32(reinterpret_cast<uint32_t*>(grad_power), n, 0u, st);
        if (rc) return rc;
        gp = nullptr;
    }
    return launch_adjoint(*c, *s, pts, powers, (uint32_t)n, grad_d, d, grad_p, gp, nullptr, st);
}

int cth_de_vjp_split_dev(cth_ctx* c, const cth_shape* s, const float* pts, const float* powers, size_t n,
                         const float* grad_d, float* d, float* grad_z0, float* grad_c, float* grad_power,
                         void* stream) {
    if (!c) return fail(CTH_E_INVALID, "null context");
    if (int rc = check_shape(s)) return rc;
    if (n && !pts) return fail(CTH_E_INVALID, "null buffer");
    if (n >= (size_t(1) << 31)) return fail(CTH_E_INVALID, "batch larger than 2^31");
    hipStream_t st = pick_stream(c, stream);
    float* gp = grad_power;
    if (grad_power && s->kind != CTH_SHAPE_MANDELBULB_TRIG) {
        int rc = launch_fill_u32(reinterpret_cast<uint32_t*>(grad_power), n, 0u, st);
        if (rc) return rc;
        gp = nullptr;
    }
    return launch_adjoint(*c, *s, pts, powers, (uint32_t)n, grad_d, d, nullptr, gp, nullptr, st, nullptr, 0, grad_z0,
                          grad_c);
}

int cth_de_vjp_noalloc_dev(cth_ctx* c, const cth_shape* s, const float* pts, const float* powers, size_t n,
                           const float* grad_d, float* d, float* grad_p, float* grad_power, void* stream) {
    if (!c) return fail(CTH_E_INVALID, "null context");
    if (int rc = check_shape(s)) return rc;
    if (n && !pts) return fail(CTH_E_INVALID, "null buffer");
    if (s->kind == CTH_SHAPE_SPHERE) return fail(CTH_E_UNSUPPORTED, "use the closed form for a sphere");
    const size_t need = vjp_workspace_bytes(s->max_iters, n);
    if (need > kVjpScratchBytes)
        return fail(CTH_E_INVALID, "n * ((max_iters + 1) * 16 + 16) bytes exceeds the 16 MiB scratch of "
                                   "cth_de_vjp_noalloc_dev (" + std::to_string(need) + " bytes needed)");
    void* ws = vjp_scratch(c->device);
    if (!ws) return fail(CTH_E_DEVICE, "the device's scratch is not set up");
    hipStream_t st = pick_stream(c, stream);
    float* gp = grad_power;
    if (grad_power && s->kind != CTH_SHAPE_MANDELBULB_TRIG) {
        int rc = launch_fill_u32(reinterpret_cast<uint32_t*>(grad_power), n, 0u, st);
        if (rc) return rc;
        gp = nullptr;
    }
    return launch_adjoint(*c, *s, pts, powers, (uint32_t)n, grad_d, d, grad_p, gp, nullptr, st, ws,
                          kVjpScratchBytes);
}

int cth_normals_dev(cth_ctx* c, const cth_shape* s, const float* pts, const float* powers, size_t n, float* d,
                    float* normals, void* stream) {
    if (!c) return fail(CTH_E_INVALID, "null context");
    if (int rc = check_shape(s)) return rc;
    if (n && (!pts || !normals)) return fail(CTH_E_INVALID, "null buffer");
    if (n >= (size_t(1) << 31)) return fail(CTH_E_INVALID, "batch larger than 2^31");
    return launch_tangent(*c, *s, pts, powers, (uint32_t)n, nullptr, d, nullptr, nullptr, normals,
                          pick_stream(c, stream));
}

int cth_de_grad_dev(cth_ctx* c, const cth_shape* s, const float* pts, const float* powers, size_t n,
                    const float* grad_d, float* d, float* grad_p, float* grad_power, void* stream) {
    if (!c) return fail(CTH_E_INVALID, "null context");
    if (int rc = check_shape(s)) return rc;
    if (n && !pts) return fail(CTH_E_INVALID, "null buffer");
    if (n >= (size_t(1) << 31)) return fail(CTH_E_INVALID, "batch larger than 2^31");
    hipStream_t st = pick_stream(c, stream);
    float* gp = grad_power;
    if (grad_power && s->kind != CTH_SHAPE_MANDELBULB_TRIG) {
        int rc = launch_fill_u32(reinterpret_cast<uint32_t*>(grad_power), n, 0u, st);
        if (rc) return rc;
        gp = nullptr;
    }
    return launch_tangent(*c, *s, pts, powers, (uint32_t)n, grad_d, d, grad_p, gp, nullptr, st);
}

int cth_normals_fd_dev(cth_ctx* c, const cth_shape* s, const float* pts, const float* powers, size_t n, float h,
                       float* d, float* normals, float* grad_p, void* stream) {
    if (!c) return fail(CTH_E_INVALID, "null context");
    if (int rc = check_shape(s)) return rc;
    if (n && !pts) return fail(CTH_E_INVALID, "null buffer");
    if (n >= (size_t(1) << 31)) return fail(CTH_E_INVALID, "batch larger than 2^31");
    if (!(h > 0.0f)) return fail(CTH_E_INVALID, "h must be positive");
    return launch_fd_normals(*c, *s, pts, powers, (uint32_t)n, h, d, grad_p, normals, pick_stream(c, stream));
}

struct cth_jobset {
    cth::JobSet* js;
};

int cth_jobs_create(cth_ctx* c, int kind, float bailout, const cth_job* jobs, uint32_t njobs, cth_jobset** out) {
    if (!c || !jobs || !out) return fail(CTH_E_INVALID, "null argument");
    *out = nullptr;
    cth::JobSet* js = nullptr;
    int rc = jobs_prepare(*c, kind, bailout, jobs, njobs, &js);
    if (rc) return rc;
    *out = new cth_jobset{js};
    return CTH_OK;
}

void cth_jobs_destroy(cth_jobset* j) {
    if (!j) return;
    jobs_destroy(j->js);
    delete j;
}

size_t cth_jobs_total(const cth_jobset* j) { return j ? jobs_total(j->js) : 0; }
size_t cth_jobs_offset(const cth_jobset* j, uint32_t i) { return j ? jobs_offset(j->js, i) : 0; }

int cth_jobs_run_dev(cth_ctx* c, const cth_jobset* j, float* dists, float* normals, uint64_t* done_ticks,
                     uint64_t* start_tick, void* stream) {
    if (!c || !j || !dists || !done_ticks || !start_tick) return fail(CTH_E_INVALID, "null argument");
    return jobs_run(*c, *j->js, dists, normals, (unsigned long long*)done_ticks, (unsigned long long*)start_tick,
                    pick_stream(c, stream));
}

int cth_wall_clock_khz(const cth_ctx* c) {
    if (!c) return 0;
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeWallClockRate, c->device) != hipSuccess) return 0;
    return v;
}

struct cth_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint32_t* pairs = nullptr;  // the graph's own work-counter pairs
};

// The graph's counter pairs are allocated and zeroed before the capture
// starts (no allocation may happen inside it) and kept until the graph is
// destroyed, so its replays never share a pair with other launches.
int cth_graph_begin(void* stream) {
    if (!stream) return fail(CTH_E_INVALID, "graph capture needs a non-null stream");
    if (g_capture.pairs) return fail(CTH_E_INVALID, "a capture is already in progress on this thread");
    uint32_t* pairs = nullptr;
    const size_t bytes = 2 * GRAPH_PAIRS * sizeof(uint32_t);
    int rc = check_hip(hipMalloc((void**)&pairs, bytes), "hipMalloc graph counters");
    if (!rc) rc = check_hip(hipMemsetAsync(pairs, 0, bytes, (hipStream_t)stream), "zero graph counters");
    if (!rc) rc = check_hip(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
    if (!rc) rc = check_hip(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal), "begin capture");
    if (rc) {
        if (pairs) (void)hipFree(pairs);
        return rc;
    }
    g_capture.stream = (hipStream_t)stream;
    g_capture.pairs = pairs;
    g_capture.pos = 0;
    return CTH_OK;
}

int cth_graph_end(void* stream, cth_graph** out) {
    if (!stream || !out) return fail(CTH_E_INVALID, "null argument");
    *out = nullptr;
    cth_graph* g = new cth_graph();
    if (g_capture.pairs && g_capture.stream == (hipStream_t)stream) {
        g->pairs = g_capture.pairs;
        g_capture = CapturePool{};
    }
    int rc = check_hip(hipStreamEndCapture((hipStream_t)stream, &g->graph), "end capture");
    if (!rc) rc = check_hip(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0), "graph instantiate");
    if (rc) {
        cth_graph_destroy(g);
        return rc;
    }
    *out = g;
    return CTH_OK;
}

int cth_graph_launch(cth_graph* g, void* stream) {
    if (!g || !g->exec) return fail(CTH_E_INVALID, "null graph");
    return check_hip(hipGraphLaunch(g->exec, (hipStream_t)stream), "graph launch");
}

void cth_graph_destroy(cth_graph* g) {
    if (!g) return;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->pairs) (void)hipFree(g->pairs);
    delete g;
}

int cth_sample_spans(cth_ctx* c, const cth_shape* s, const cth_span* spans, size_t nspans, uint32_t res,
                     float* dists, uint64_t* cross_bits) {
    if (!c) return fail(CTH_E_INVALID, "null context");
    if (int rc = check_shape(s)) return rc;
    if (nspans == 0) return CTH_OK;
    if (!spans || !dists) return fail(CTH_E_INVALID, "null buffer");
    if (res == 0 || (res & (res - 1)) != 0) return fail(CTH_E_INVALID, "resolution must be a power of two");
    if (res > 512) return fail(CTH_E_INVALID, "resolution above 512");
    size_t per = size_t(res + 1) * (res + 1) * (res + 1);
    size_t words = (size_t(res) * res * res + 63) / 64;
    void* b[3];
    int rc = staging(*c, WS_HOST_IN,
                     {nspans * sizeof(cth_span), nspans * per * sizeof(float),
                      cross_bits ? nspans * words * sizeof(uint64_t) : 0},
                     b);
    if (rc) return rc;
    rc = check_hip(hipMemcpyAsync(b[0], spans, nspans * sizeof(cth_span), hipMemcpyHostToDevice, c->stream), "H2D");
    if (!rc)
        rc = cth_sample_spans_dev(c, s, (const float*)b[0], nspans, res, (float*)b[1],
                                  cross_bits ? (uint64_t*)b[2] : nullptr, nullptr, c->stream);
    if (!rc)
        rc = check_hip(hipMemcpyAsync(dists, b[1], nspans * per * sizeof(float), hipMemcpyDeviceToHost, c->stream),
                       "D2H");
    if (!rc && cross_bits)
        rc = check_hip(hipMemcpyAsync(cross_bits, b[2], nspans * words * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                      c->stream),
                       "D2H");
    int rs = check_hip(hipStreamSynchronize(c->stream), "sync");
    return rc ? rc : rs;
}

int cth_mesh_count_dev(cth_ctx* c, const float* dists, size_t nspans, uint32_t res, uint64_t* voff, uint64_t* ioff,
                       uint64_t* cross_bits, uint64_t* nverts, uint64_t* nidx, void* stream) {
    if (!c) return fail(CTH_E_INVALID, "null context");
    if (!voff || !ioff || !nverts || !nidx || (nspans && !dists)) return fail(CTH_E_INVALID, "null buffer");
    if (nspans >= (size_t(1) << 31)) return fail(CTH_E_INVALID, "more than 2^31 spans");
    uint64_t tot[2] = {0, 0};
    int rc = mesh_count(*c, dists, (uint32_t)nspans, res, voff, ioff, cross_bits, tot, pick_stream(c, stream));
    *nverts = tot[0];
    *nidx = tot[1];
    return rc;
}

int cth_mesh_emit_dev(cth_ctx* c, const cth_shape* s, const float* spans, const float* dists, size_t nspans,
                      uint32_t res, int normals, const uint64_t* voff, const uint64_t* ioff, uint64_t nverts,
                      cth_vertex* vertices, uint32_t* indices, void* stream) {
    if (!c) return fail(CTH_E_INVALID, "null context");
    if (int rc = check_shape(s)) return rc;
    if (nspans && (!spans || !dists || !voff || !ioff || !indices || (nverts && !vertices)))
        return fail(CTH_E_INVALID, "null buffer");
    if (nspans >= (size_t(1) << 31)) return fail(CTH_E_INVALID, "more than 2^31 spans");
    return mesh_emit(*c, *s, spans, dists, (uint32_t)nspans, res, normals, voff, ioff, nverts,
                     reinterpret_cast<float*>(vertices), indices, pick_stream(c, stream));
}

int cth_mesh_vertices_dev(cth_ctx* c, const cth_shape* s, cth_vertex* vertices, uint64_t nverts, int normals,
                          void* stream) {
    if (!c) return fail(CTH_E_INVALID, "null context");
    if (int rc = check_shape(s)) return rc;
    if (nverts && !vertices) return fail(CTH_E_INVALID, "null buffer");
    return mesh_vertices(*c, *s, normals, nverts, reinterpret_cast<float*>(vertices), pick_stream(c, stream));
}

int cth_mesh_span(cth_ctx* c, const cth_shape* s, const cth_span* span, uint32_t res, int normals,
                  cth_vertex* vertices, size_t vert_cap, uint32_t* indices, size_t idx_cap, size_t* nverts,
                  size_t* nidx) {
    if (!c) return fail(CTH_E_INVALID, "null context");
    if (int rc = check_shape(s)) return rc;
    if (!span || !nverts || !nidx) return fail(CTH_E_INVALID, "null buffer");
    for (int a = 0; a < 3; ++a)
        if (!(span->start[a] < span->end[a])) return fail(CTH_E_INVALID, "span start must be below end on every axis");
    if (res == 0 || (res & (res - 1)) != 0 || res > 64)
        return fail(CTH_E_INVALID, "mesh resolution must be a power of two in [1, 64]");
    if (normals != CTH_NORMALS_FD6 && normals != CTH_NORMALS_TANGENT && normals != CTH_NORMALS_DEFERRED)
        return fail(CTH_E_INVALID, "unknown normals mode");
    if (normals == CTH_NORMALS_TANGENT && s->kind == CTH_SHAPE_SPHERE)
        return fail(CTH_E_UNSUPPORTED, "tangent normals need a Mandelbulb kind");
    const size_t per = size_t(res + 1) * (res + 1) * (res + 1);
    void* b[3];
    int rc = staging(*c, WS_HOST_IN, {sizeof(cth_span), per * sizeof(float), 4 * sizeof(uint64_t)}, b);
    if (rc) return rc;
    rc = check_hip(hipMemcpyAsync(b[0], span, sizeof(cth_span), hipMemcpyHostToDevice, c->stream), "H2D");
    if (!rc) rc = cth_sample_spans_dev(c, s, (const float*)b[0], 1, res, (float*)b[1], nullptr, nullptr, c->stream);
    uint64_t nv = 0, ni = 0;
    uint64_t* off = (uint64_t*)b[2];
    if (!rc) rc = cth_mesh_count_dev(c, (const float*)b[1], 1, res, off, off + 2, nullptr, &nv, &ni, c->stream);
    if (rc) return rc;
    *nverts = (size_t)nv;
    *nidx = (size_t)ni;
    if (!vertices && !indices) return CTH_OK;
    if ((vertices && vert_cap < nv) || (indices && idx_cap < ni) || (nv && !vertices) || (ni && !indices))
        return fail(CTH_E_INVALID, "vertex or index buffer too small for the mesh");
    void* o[2];
    rc = staging(*c, WS_HOST_OUT, {nv * sizeof(cth_vertex), ni * sizeof(uint32_t) + 4}, o);
    if (!rc)
        rc = cth_mesh_emit_dev(c, s, (const float*)b[0], (const float*)b[1], 1, res, normals, off, off + 2, nv,
                               (cth_vertex*)o[0], (uint32_t*)o[1], c->stream);
    if (!rc && nv)
        rc = check_hip(hipMemcpyAsync(vertices, o[0], nv * sizeof(cth_vertex), hipMemcpyDeviceToHost, c->stream),
                       "D2H");
    if (!rc && ni)
        rc = check_hip(hipMemcpyAsync(indices, o[1], ni * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream), "D2H");
    int rs = check_hip(hipStreamSynchronize(c->stream), "sync");
    return rc ? rc : rs;
}

int cth_mesh_spans(cth_ctx* c, const cth_shape* s, const cth_span* spans, size_t nspans, uint32_t res, int normals,
                   cth_vertex* vertices, size_t vert_cap, uint32_t* indices, size_t idx_cap, uint64_t* voff,
                   uint64_t* ioff, size_t* nverts, size_t* nidx) {
    if (!c) return fail(CTH_E_INVALID, "null context");
    if (int rc = check_shape(s)) return rc;
    if (!nverts || !nidx || (nspans && !spans)) return fail(CTH_E_INVALID, "null buffer");
    if (res == 0 || (res & (res - 1)) != 0 || res > 64)
        return fail(CTH_E_INVALID, "mesh resolution must be a power of two in [1, 64]");
    if (normals != CTH_NORMALS_FD6 && normals != CTH_NORMALS_TANGENT && normals != CTH_NORMALS_DEFERRED)
        return fail(CTH_E_INVALID, "unknown normals mode");
    if (normals == CTH_NORMALS_TANGENT && s->kind == CTH_SHAPE_SPHERE)
        return fail(CTH_E_UNSUPPORTED, "tangent normals need a Mandelbulb kind");
    if (nspans >= (size_t(1) << 31)) return fail(CTH_E_INVALID, "more than 2^31 spans");
    for (size_t i = 0; i < nspans; ++i)
        for (int a = 0; a < 3; ++a)
            if (!(spans[i].start[a] < spans[i].end[a]))
                return fail(CTH_E_INVALID, "span start must be below end on every axis");
    *nverts = 0;
    *nidx = 0;
    if (nspans == 0) {
        if (voff) voff[0] = 0;
        if (ioff) ioff[0] = 0;
        return CTH_OK;
    }
    const size_t per = size_t(res + 1) * (res + 1) * (res + 1);
    void* b[4];
    int rc = staging(*c, WS_HOST_IN,
                     {nspans * sizeof(cth_span), nspans * per * sizeof(float), (nspans + 1) * sizeof(uint64_t),
                      (nspans + 1) * sizeof(uint64_t)},
                     b);
    if (rc) return rc;
    uint64_t* voff_d = (uint64_t*)b[2];
    uint64_t* ioff_d = (uint64_t*)b[3];
    rc = check_hip(hipMemcpyAsync(b[0], spans, nspans * sizeof(cth_span), hipMemcpyHostToDevice, c->stream), "H2D");
    if (!rc)
        rc = cth_sample_spans_dev(c, s, (const float*)b[0], nspans, res, (float*)b[1], nullptr, nullptr, c->stream);
    uint64_t nv = 0, ni = 0;
    if (!rc) rc = cth_mesh_count_dev(c, (const float*)b[1], nspans, res, voff_d, ioff_d, nullptr, &nv, &ni, c->stream);
    if (rc) return rc;
    *nverts = (size_t)nv;
    *nidx = (size_t)ni;
    if (voff)
        rc = check_hip(hipMemcpyAsync(voff, voff_d, (nspans + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream),
                       "D2H");
    if (!rc && ioff)
        rc = check_hip(hipMemcpyAsync(ioff, ioff_d, (nspans + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream),
                       "D2H");
    if (!rc && (vertices || indices)) {
        if ((vertices && vert_cap < nv) || (indices && idx_cap < ni) || (nv && !vertices) || (ni && !indices))
            rc = fail(CTH_E_INVALID, "vertex or index buffer too small for the meshes");
        void* o[2];
        if (!rc) rc = staging(*c, WS_HOST_OUT, {nv * sizeof(cth_vertex), ni * sizeof(uint32_t) + 4}, o);
        if (!rc)
            rc = cth_mesh_emit_dev(c, s, (const float*)b[0], (const float*)b[1], nspans, res, normals, voff_d, ioff_d,
                                   nv, (cth_vertex*)o[0], (uint32_t*)o[1], c->stream);
        if (!rc && nv)
            rc = check_hip(hipMemcpyAsync(vertices, o[0], nv * sizeof(cth_vertex), hipMemcpyDeviceToHost, c->stream),
                           "D2H");
        if (!rc && ni)
            rc = check_hip(hipMemcpyAsync(indices, o[1], ni * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream),
                           "D2H");
    }
    int rs = check_hip(hipStreamSynchronize(c->stream), "sync");
    return rc ? rc : rs;
}

int cth_focus_split(cth_ctx* c, const cth_span* leaves, size_t n, const cth_point3* focii, size_t nfocii,
                    const float eye[3], uint32_t* flags) {
    if (!c) return fail(CTH_E_INVALID, "null context");
    if (!eye || (n && (!leaves || !flags)) || (nfocii && !focii)) return fail(CTH_E_INVALID, "null buffer");
    if (n == 0) return CTH_OK;
    if (n >= (size_t(1) << 31) || nfocii >= (size_t(1) << 31)) return fail(CTH_E_INVALID, "batch larger than 2^31");
    void* b[3];
    int rc = staging(*c, WS_HOST_IN, {n * sizeof(cth_span), nfocii * sizeof(cth_point3), n * sizeof(uint32_t)}, b);
    if (rc) return rc;
    rc = check_hip(hipMemcpyAsync(b[0], leaves, n * sizeof(cth_span), hipMemcpyHostToDevice, c->stream), "H2D");
    if (!rc && nfocii)
        rc = check_hip(hipMemcpyAsync(b[1], focii, nfocii * sizeof(cth_point3), hipMemcpyHostToDevice, c->stream),
                       "H2D");
    if (!rc)
        rc = cth_focus_split_dev(c, (const float*)b[0], n, (const float*)b[1], nullptr, nfocii, eye, (uint32_t*)b[2],
                                 c->stream);
    if (!rc) rc = check_hip(hipMemcpyAsync(flags, b[2], n * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream), "D2H");
    int rs = check_hip(hipStreamSynchronize(c->stream), "sync");
    return rc ? rc : rs;
}

// Host-pointer Shape-trait queries: copy in, run, copy out.
enum Query { Q_MIN, Q_MAX, Q_BOUNDED, Q_CONTAINS };

static int host_query(cth_ctx* c, const cth_shape* s, const cth_point3* pts, size_t n, void* out, Query q) {
    if (!c) return fail(CTH_E_INVALID, "null context");
    if (int rc = check_shape(s)) return rc;
    if ((q == Q_MAX || q == Q_BOUNDED) && s->kind != CTH_SHAPE_SPHERE)
        return fail(CTH_E_UNSUPPORTED, "this shape has no upper distance bound (max_distance_from is None)");
    if (n == 0) return CTH_OK;
    if (!pts || !out) return fail(CTH_E_INVALID, "null buffer");
    if (n >= (size_t(1) << 31)) return fail(CTH_E_INVALID, "batch larger than 2^31");
    uint32_t m = (uint32_t)n;
    size_t out_bytes = (q == Q_BOUNDED) ? n * 8 : (q == Q_CONTAINS ? n : n * 4);
    void* b[3];  // points, distances, derived output
    int rc = staging(*c, WS_HOST_IN, {n * 12, n * 4, (q == Q_BOUNDED || q == Q_CONTAINS) ? out_bytes : 0}, b);
    if (rc) return rc;
    rc = check_hip(hipMemcpyAsync(b[0], pts, n * 12, hipMemcpyHostToDevice, c->stream), "H2D");
    if (!rc) {
        if (q == Q_CONTAINS && s->kind == CTH_SHAPE_SPHERE) {
            rc = launch_contains_sphere(*c, *s, (const float*)b[0], m, (uint8_t*)b[2], c->stream);
        } else {
            rc = launch_de_points(*c, *s, (const float*)b[0], nullptr, m, (float*)b[1], nullptr, nullptr, nullptr,
                                  c->stream);
            if (!rc && q == Q_CONTAINS) rc = launch_less_than_zero(*c, (const float*)b[1], m, (uint8_t*)b[2], c->stream);
            if (!rc && q == Q_BOUNDED) rc = launch_dup_pairs(*c, (const float*)b[1], m, (float*)b[2], c->stream);
        }
    }
    const void* src = (q == Q_MIN || q == Q_MAX) ? b[1] : b[2];
    if (!rc) rc = check_hip(hipMemcpyAsync(out, src, out_bytes, hipMemcpyDeviceToHost, c->stream), "D2H");
    int rs = check_hip(hipStreamSynchronize(c->stream), "sync");
    return rc ? rc : rs;
}

int cth_min_distance_batch(cth_ctx* c, const cth_shape* s, const cth_point3* pts, size_t n, float* out) {
    return host_query(c, s, pts, n, out, Q_MIN);
}

int cth_max_distance_batch(cth_ctx* c, const cth_shape* s, const cth_point3* pts, size_t n, float* out) {
    return host_query(c, s, pts, n, out, Q_MAX);
}

int cth_bounded_distance_batch(cth_ctx* c, const cth_shape* s, const cth_point3* pts, size_t n, float* out_pairs) {
    return host_query(c, s, pts, n, out_pairs, Q_BOUNDED);
}

int cth_contains_batch(cth_ctx* c, const cth_shape* s, const cth_point3* pts, size_t n, uint8_t* out) {
    return host_query(c, s, pts, n, out, Q_CONTAINS);
}

int cth_de_vjp_batch(cth_ctx* c, const cth_shape* s, const cth_point3* pts, const float* powers, size_t n,
                     const float* grad_d, float* d, float* grad_p, float* grad_power, int sweep) {
    if (!c) return fail(CTH_E_INVALID, "null context");
    if (int rc = check_shape(s)) return rc;
    if (sweep != CTH_SWEEP_FORWARD && sweep != CTH_SWEEP_REVERSE)
        return fail(CTH_E_INVALID, "sweep must be CTH_SWEEP_FORWARD or CTH_SWEEP_REVERSE");
    if (n == 0) return CTH_OK;
    if (!pts) return fail(CTH_E_INVALID, "null buffer");
    if (n >= (size_t(1) << 31)) return fail(CTH_E_INVALID, "batch larger than 2^31");
    void* b[6];  // points, powers, grad_d, d, grad_p, grad_power
    int rc = staging(*c, WS_HOST_IN, {n * 12, powers ? n * 4 : 0, grad_d ? n * 4 : 0, n * 4, n * 12, n * 4}, b);
    if (rc) return rc;
    rc = check_hip(hipMemcpyAsync(b[0], pts, n * 12, hipMemcpyHostToDevice, c->stream), "H2D");
    if (!rc && powers) rc = check_hip(hipMemcpyAsync(b[1], powers, n * 4, hipMemcpyHostToDevice, c->stream), "H2D");
    if (!rc && grad_d) rc = check_hip(hipMemcpyAsync(b[2], grad_d, n * 4, hipMemcpyHostToDevice, c->stream), "H2D");
    if (!rc) {
        const float* dp = (const float*)b[0];
        const float* dpw = powers ? (const float*)b[1] : nullptr;
        const float* dg = grad_d ? (const float*)b[2] : nullptr;
        float* gpw = grad_power ? (float*)b[5] : nullptr;
        rc = sweep == CTH_SWEEP_FORWARD
                 ? cth_de_grad_dev(c, s, dp, dpw, n, dg, (float*)b[3], (float*)b[4], gpw, c->stream)
                 : cth_de_vjp_dev(c, s, dp, dpw, n, dg, (float*)b[3], (float*)b[4], gpw, c->stream);
    }
    if (!rc && d) rc = check_hip(hipMemcpyAsync(d, b[3], n * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
    if (!rc && grad_p) rc = check_hip(hipMemcpyAsync(grad_p, b[4], n * 12, hipMemcpyDeviceToHost, c->stream), "D2H");
    if (!rc && grad_power)
        rc = check_hip(hipMemcpyAsync(grad_power, b[5], n * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
    int rs = check_hip(hipStreamSynchronize(c->stream), "sync");
    return rc ? rc : rs;
}

}  // extern "C"
