// Differentiated distance estimation: hand-written adjoint (reverse sweep)
// and the 6-tap central-difference baseline it replaces.
//
// The reference computes vertex normals with six extra distance evaluations
// (src/mesh/buffer.rs:256-266).  Here the gradient of
//   d = 0.5 ln(r_f) r_f / dr_f,   dr_{k+1} = P r_k^(P-1) dr_k + 1,
//   z_{k+1} = F(z_k; P) + c,      z_0 = c = p
// with respect to p (and the power P) comes from one reverse sweep:
//
//   kernel 1 (k_adj_forward): the forward iteration, recomputed per lane in
//     parity arithmetic, records each point's orbit (z_k, dr_k) in a
//     workspace (point-major, so a lane's own records are contiguous) plus the
//     loop end state (K, r_f, dr_f) and the distance;
//   kernel 2 (k_adj_reverse): seeds r_f_bar, dr_f_bar from grad_d and runs the
//     reverse sweep k = K-1 .. 0 through the spherical power map (trig family)
//     or the closed-form P8 polynomial, accumulating c_bar and P_bar.
//
// Both kernels use the lane-refill schedule of work_loop.h: one forward body
// or one reverse step per lane per tick, finished lanes take new points, so
// neither phase pays for the spread of iteration counts.
#include <hip/hip_runtime.h>

#include <mutex>

#include "internal.h"
#include "launch.h"
#include "point_sources.h"
#include "tangent.h"
#include "work_loop.h"

namespace cth {

struct AdjRecord {
    uint32_t k;   // loop bodies executed
    float r, dr;  // end state entering the distance formula
    float pad;
};

// ---------------------------------------------------------------------------
// Kernel 1: forward with orbit recording
// ---------------------------------------------------------------------------
template <class Ev>
__global__ __launch_bounds__(256) void k_adj_forward(SrcPoints src, MandelUniform u, float4* __restrict__ orbit,
                                                     AdjRecord* __restrict__ rec, float* __restrict__ dout,
                                                     WorkQueue q) {
    Ev ev;
    bool live = false, hold = false;
    uint32_t idx = 0, k = 0;
    float r = 0.0f;
    uint32_t c_next = 0, c_end = 0, round = 0;
    bool exhausted = false;
    const size_t stride = (size_t)u.max_iters + 1;
    for (;;) {
        if ((uint32_t)__popcll(wave_mask(!live)) >= q.refill_min) {
            if (hold) {
                float d = ev.finish(r);
                if (dout) dout[idx] = d;
                rec[idx] = AdjRecord{k, r, ev.dr, 0.0f};
                // z_K (the state the bailout magnitude was taken of)
                orbit[(size_t)idx * stride + k] = make_float4(ev.x, ev.y, ev.z, ev.dr);
                hold = false;
            }
            while (!exhausted) {
                uint64_t need = wave_mask(!live);
                if (need == 0) break;
                if (c_next >= c_end) {
                    uint32_t b = next_chunk(q, round);
                    if (b >= q.total) {
                        exhausted = true;
                        break;
                    }
                    c_next = b;
                    c_end = (q.total - b < q.chunk) ? q.total : b + q.chunk;
                }
                uint32_t avail = c_end - c_next;
                uint32_t nneed = (uint32_t)__popcll(need);
                if (!live) {
                    uint32_t rank = lane_rank(need);
                    if (rank < avail) {
                        idx = c_next + rank;
                        float px, py, pz, pw;
                        src.get(idx, px, py, pz, pw);
                        ev.start(px, py, pz, pw, u);
                        k = 0;
                        r = ev.mag();
                        if (r > u.bailout) {
                            float d = ev.finish(r);
                            if (dout) dout[idx] = d;
                            rec[idx] = AdjRecord{0u, r, ev.dr, 0.0f};
                        } else {
                            live = true;
                        }
                    }
                }
                c_next += (nneed < avail) ? nneed : avail;
            }
            if (wave_mask(live) == 0) break;
        }
        if (live) orbit[(size_t)idx * stride + k] = make_float4(ev.x, ev.y, ev.z, ev.dr);
        ev.body(r, u, live);
        k += live ? 1u : 0u;
        bool cap = live && (k == u.max_iters);
        float rn = ev.mag();
        bool bail = live && !cap && (rn > u.bailout);
        if (live && !cap) r = rn;
        if (cap || bail) {
            live = false;
            hold = true;
        }
    }
    if (q.dynamic) wave_retire(q);
}

// ---------------------------------------------------------------------------
// Reverse steps.  Inputs: z = z_k, dr = dr_k, the adjoints zb (of z_{k+1}),
// drb (of dr_{k+1}) and rb_extra (an adjoint reaching r_k directly: the
// distance formula's at the iteration cap).  Outputs: zb (of z_k), drb (of
// dr_k); cb += zb_in (z_{k+1} = F(z_k) + c); Pb += dP.
// ---------------------------------------------------------------------------

// Trig family: F = R u(Theta, Phi), R = r^P, Theta = P acos(z/r),
// Phi = P atan2(y, x), u = (sin Theta cos Phi, sin Theta sin Phi, cos Theta)
// (src/shape/mandelbulb.rs:128-146).
__device__ __forceinline__ void rev_step_trig(float x, float y, float z, float dr, float P, float rb_extra, float& zbx,
                                              float& zby, float& zbz, float& drb, float& cbx, float& cby, float& cbz,
                                              float& Pb) {
    cbx += zbx;
    cby += zby;
    cbz += zbz;
    float rho2 = __builtin_fmaf(x, x, y * y);
    float r2 = __builtin_fmaf(z, z, rho2);
    float r = __builtin_sqrtf(r2);
    float rho = __builtin_sqrtf(rho2);
    float lnr = __logf(r);
    float th = acosf(z / r);
    float ph = atan2f(y, x);
    float Th = th * P, Ph = ph * P;
    float sT, cT, sP, cP;
    __sincosf(Th, &sT, &cT);
    __sincosf(Ph, &sP, &cP);
    float R = __expf(P * lnr);     // r^P
    float Rm1 = R / r;             // r^(P-1)
    // F_bar . u, F_bar . du/dTheta, F_bar . du/dPhi
    float gu = zbx * (sT * cP) + zby * (sT * sP) + zbz * cT;
    float gT = zbx * (cT * cP) + zby * (cT * sP) - zbz * sT;
    float gP = sP * (-zbx * sT) + cP * (zby * sT);
    float rb = P * Rm1 * gu + rb_extra;
    float thb = R * P * gT;
    float phb = R * P * gP;
    Pb += R * (lnr * gu + th * gT + ph * gP);
    // dr_{k+1} = P r^(P-1) dr_k + 1
    rb += drb * P * (P - 1.0f) * (Rm1 / r) * dr;
    Pb += drb * Rm1 * dr * (1.0f + P * lnr);
    drb = drb * P * Rm1;
    // chain through r, theta, phi of z_k
    float ir = 1.0f / r, ir2 = ir * ir;
    float k1 = thb * z * ir2 / rho;  // theta_x = x z / (r^2 rho)
    float ip = 1.0f / rho2;
    zbx = rb * x * ir + k1 * x - phb * y * ip;
    zby = rb * y * ir + k1 * y + phb * x * ip;
    zbz = rb * z * ir - thb * rho * ir2;
}

// P8 polynomial (src/shape/mandelbulb.rs:148-200), reverse-mode by hand.
__device__ __forceinline__ void rev_step_p8(float x, float y, float z, float dr, float rb_extra, float& zbx,
                                            float& zby, float& zbz, float& drb, float& cbx, float& cby, float& cbz) {
    cbx += zbx;
    cby += zby;
    cbz += zbz;
    const float Fx = zbx, Fy = zby, Fz = zbz;
    float x2 = x * x, x4 = x2 * x2, x6 = x4 * x2, x8 = x4 * x4;
    float y2 = y * y, y4 = y2 * y2, y6 = y4 * y2, y8 = y4 * y4;
    float z2 = z * z, z4 = z2 * z2, z6 = z4 * z2, z8 = z4 * z4;
    float w2 = x2 + y2, w4 = w2 * w2, w6 = w2 * w4, w8 = w4 * w4;
    float num = ((z8 - 28.0f * z6 * w2) + 70.0f * z4 * w4) - 28.0f * z2 * w6;
    float iw8 = 1.0f / w8;
    float a = 1.0f + num * iw8;
    float X = (((x8 - 28.0f * x6 * y2) + 70.0f * x4 * y4) - 28.0f * x2 * y6) - y8;
    float Y = ((x6 - 7.0f * x4 * y2) + 7.0f * x2 * y4) - y6;
    float sw = __builtin_sqrtf(w2);
    float D = z2 - w2;
    float E = (z4 - 6.0f * z2 * w2) + w4;
    // F_x = a X, F_y = 8 a x y Y, F_z = 8 z sw D E
    float ab = Fx * X + Fy * 8.0f * x * y * Y;
    float Xb = Fx * a;
    float Yb = Fy * 8.0f * a * x * y;
    float xb = Fy * 8.0f * a * y * Y;
    float yb = Fy * 8.0f * a * x * Y;
    float zb = Fz * 8.0f * sw * D * E;
    float swb = Fz * 8.0f * z * D * E;
    float Db = Fz * 8.0f * z * sw * E;
    float Eb = Fz * 8.0f * z * sw * D;
    float w2b = (sw > 0.0f) ? swb * 0.5f / sw : 0.0f;
    float z2b = Db, z4b = Eb, w4b = Eb;
    w2b -= Db;
    z2b += -6.0f * w2 * Eb;
    w2b += -6.0f * z2 * Eb;
    // a = 1 + num / w8
    float numb = ab * iw8;
    float w8b = -ab * num * iw8 * iw8;
    float z8b = numb;
    float z6b = -28.0f * w2 * numb;
    w2b += -28.0f * z6 * numb;
    z4b += 70.0f * w4 * numb;
    w4b += 70.0f * z4 * numb;
    z2b += -28.0f * w6 * numb;
    float w6b = -28.0f * z2 * numb;
    // X and Y
    float x8b = Xb;
    float x6b = -28.0f * y2 * Xb + Yb;
    float y2b = -28.0f * x6 * Xb - 7.0f * x4 * Yb;
    float x4b = 70.0f * y4 * Xb - 7.0f * y2 * Yb;
    float y4b = 70.0f * x4 * Xb + 7.0f * x2 * Yb;
    float x2b = -28.0f * y6 * Xb + 7.0f * y4 * Yb;
    float y6b = -28.0f * x2 * Xb - Yb;
    float y8b = -Xb;
    // w powers: w8 = w4^2, w6 = w2 w4, w4 = w2^2, w2 = x2 + y2
    w4b += 2.0f * w4 * w8b + w2 * w6b;
    w2b += w4 * w6b + 2.0f * w2 * w4b;
    x2b += w2b;
    y2b += w2b;
    // x powers: x8 = x4^2, x6 = x4 x2, x4 = x2^2, x2 = x^2 (same for y, z)
    x4b += 2.0f * x4 * x8b + x2 * x6b;
    x2b += x4 * x6b + 2.0f * x2 * x4b;
    xb += 2.0f * x * x2b;
    y4b += 2.0f * y4 * y8b + y2 * y6b;
    y2b += y4 * y6b + 2.0f * y2 * y4b;
    yb += 2.0f * y * y2b;
    z4b += 2.0f * z4 * z8b + z2 * z6b;
    z2b += z4 * z6b + 2.0f * z2 * z4b;
    zb += 2.0f * z * z2b;
    // dr_{k+1} = 8 r^7 dr_k + 1 and r = |z_k|
    float r = __builtin_sqrtf(__builtin_fmaf(z, z, w2));
    float r2 = r * r, r6 = r2 * r2 * r2;
    float rb = drb * 56.0f * r6 * dr + rb_extra;
    drb = drb * 8.0f * r6 * r;
    float ir = (r > 0.0f) ? 1.0f / r : 0.0f;
    zbx = xb + rb * x * ir;
    zby = yb + rb * y * ir;
    zbz = zb + rb * z * ir;
}

// ---------------------------------------------------------------------------
// Kernel 2: reverse sweep with lane refill
// ---------------------------------------------------------------------------
struct AdjOut {
    const float* grad_d;  // nullable: 1
    float* grad_p;        // n x 3
    float* grad_power;    // nullable
    float* normals;       // nullable: grad_p / |grad_p| instead of grad_p
    float* grad_z0;       // nullable: the part of grad_p through the orbit's start z0
    float* grad_c;        // nullable: the part through the constant c of every body
};

// grad_p = (gx, gy, gz) = z_bar_0 + c_bar (z_0 = c = p), and its two parts.
__device__ __forceinline__ void adj_write(const AdjOut& out, uint32_t idx, float gx, float gy, float gz, float zx,
                                          float zy, float zz, float cx, float cy, float cz, float Pb) {
    if (out.grad_z0) {
        out.grad_z0[3 * (size_t)idx] = zx;
        out.grad_z0[3 * (size_t)idx + 1] = zy;
        out.grad_z0[3 * (size_t)idx + 2] = zz;
    }
    if (out.grad_c) {
        out.grad_c[3 * (size_t)idx] = cx;
        out.grad_c[3 * (size_t)idx + 1] = cy;
        out.grad_c[3 * (size_t)idx + 2] = cz;
    }
    if (out.normals) {
        // cgmath normalize: v * (1 / |v|)
        float m = __builtin_sqrtf(__builtin_fmaf(gx, gx, __builtin_fmaf(gy, gy, gz * gz)));
        float im = 1.0f / m;
        out.normals[3 * (size_t)idx] = gx * im;
        out.normals[3 * (size_t)idx + 1] = gy * im;
        out.normals[3 * (size_t)idx + 2] = gz * im;
    }
    if (out.grad_p) {
        out.grad_p[3 * (size_t)idx] = gx;
        out.grad_p[3 * (size_t)idx + 1] = gy;
        out.grad_p[3 * (size_t)idx + 2] = gz;
    }
    if (out.grad_power) out.grad_power[idx] = Pb;
}

template <bool TRIG>
__global__ __launch_bounds__(256) void k_adj_reverse(SrcPoints src, MandelUniform u, const float4* __restrict__ orbit,
                                                     const AdjRecord* __restrict__ rec, AdjOut out, WorkQueue q) {
    bool live = false;
    uint32_t idx = 0;
    int j = 0;  // next orbit index of the reverse sweep
    float P = 0.0f, px = 0.0f, py = 0.0f, pz = 0.0f;
    float zbx = 0.0f, zby = 0.0f, zbz = 0.0f, drb = 0.0f, cbx = 0.0f, cby = 0.0f, cbz = 0.0f, Pb = 0.0f;
    float rb_extra = 0.0f;
    uint32_t c_next = 0, c_end = 0, round = 0;
    bool exhausted = false;
    const size_t stride = (size_t)u.max_iters + 1;
    for (;;) {
        if ((uint32_t)__popcll(wave_mask(!live)) >= q.refill_min) {
            while (!exhausted) {
                uint64_t need = wave_mask(!live);
                if (need == 0) break;
                if (c_next >= c_end) {
                    uint32_t b = next_chunk(q, round);
                    if (b >= q.total) {
                        exhausted = true;
                        break;
                    }
                    c_next = b;
                    c_end = (q.total - b < q.chunk) ? q.total : b + q.chunk;
                }
                uint32_t avail = c_end - c_next;
                uint32_t nneed = (uint32_t)__popcll(need);
                if (!live) {
                    uint32_t rank = lane_rank(need);
                    if (rank < avail) {
                        idx = c_next + rank;
                        src.get(idx, px, py, pz, P);
                        AdjRecord rc = rec[idx];
                        float g = out.grad_d ? out.grad_d[idx] : 1.0f;
                        // seeds of d = 0.5 ln(r_f) r_f / dr_f
                        float rf = rc.r, drf = rc.dr;
                        float lnr = __logf(rf);
                        float rbar = g * 0.5f * (lnr + 1.0f) / drf;
                        drb = -g * (0.5f * lnr * rf / drf) / drf;
                        cbx = cby = cbz = Pb = 0.0f;
                        rb_extra = 0.0f;
                        int K = (int)rc.k;
                        float ir = 1.0f / rf;
                        if (K == 0) {
                            // r_f = |z_0|: no loop body, so c does not reach d
                            const float gx = rbar * px * ir, gy = rbar * py * ir, gz = rbar * pz * ir;
                            adj_write(out, idx, gx, gy, gz, gx, gy, gz, 0.0f, 0.0f, 0.0f, 0.0f);
                        } else {
                            if ((uint32_t)K == u.max_iters) {
                                // iteration cap: r_f = |z_{K-1}|, z_K unused
                                zbx = zby = zbz = 0.0f;
                                rb_extra = rbar;
                            } else {
                                // bailout: r_f = |z_K|
                                float4 zk = orbit[(size_t)idx * stride + K];
                                zbx = rbar * zk.x * ir;
                                zby = rbar * zk.y * ir;
                                zbz = rbar * zk.z * ir;
                            }
                            j = K - 1;
                            live = true;
                        }
                    }
                }
                c_next += (nneed < avail) ? nneed : avail;
            }
            if (wave_mask(live) == 0) break;
        }
        if (live) {
            float4 s = orbit[(size_t)idx * stride + j];
            if (TRIG)
                rev_step_trig(s.x, s.y, s.z, s.w, P, rb_extra, zbx, zby, zbz, drb, cbx, cby, cbz, Pb);
            else
                rev_step_p8(s.x, s.y, s.z, s.w, rb_extra, zbx, zby, zbz, drb, cbx, cby, cbz);
            rb_extra = 0.0f;
            if (--j < 0) {
                adj_write(out, idx, cbx + zbx, cby + zby, cbz + zbz, zbx, zby, zbz, cbx, cby, cbz, Pb);
                live = false;
            }
        }
    }
    if (q.dynamic) wave_retire(q);
}

// ---------------------------------------------------------------------------
// 6-tap central differences (the reference's normal stencil): virtual point
// 7 i + t is p_i (t = 0) or p_i +- h e_a (t = 1 + 2a, 2 + 2a), written as
// (p_b + u_b * h, ...) with u = e_a exactly like `p + Vector3::unit_x() * delta.x`
// (src/mesh/buffer.rs:259-264).
// ---------------------------------------------------------------------------
struct SrcFD {
    const float* pts;
    const float* powers;
    float power, h;
    __device__ __forceinline__ void get(uint32_t v, float& x, float& y, float& z, float& p) const {
        uint32_t i = v / 7u, t = v - 7u * i;
        float c[3] = {pts[3 * (size_t)i], pts[3 * (size_t)i + 1], pts[3 * (size_t)i + 2]};
        if (t > 0) {
            int a = (int)(t - 1) >> 1;
            float hh = (t & 1u) ? h : -h;
#pragma unroll
            for (int b = 0; b < 3; ++b) c[b] = c[b] + ((a == b) ? 1.0f : 0.0f) * hh;
        }
        x = c[0];
        y = c[1];
        z = c[2];
        p = powers ? powers[i] : power;
    }
};

__global__ void k_fd_combine(const float* __restrict__ taps, uint32_t n, float* __restrict__ normals,
                             float* __restrict__ grad, float* __restrict__ dout, float inv2h) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* t = taps + 7 * (size_t)i;
    float g[3] = {t[1] - t[2], t[3] - t[4], t[5] - t[6]};
    if (dout) dout[i] = t[0];
    if (normals) {
        // cgmath normalize of the difference vector (src/mesh/buffer.rs:258-265)
        float m = sqrt_rn((g[0] * g[0] + g[1] * g[1]) + g[2] * g[2]);
        float im = div_rn(1.0f, m);
        for (int a = 0; a < 3; ++a) normals[3 * (size_t)i + a] = g[a] * im;
    }
    if (grad)
        for (int a = 0; a < 3; ++a) grad[3 * (size_t)i + a] = g[a] * inv2h;
}

static inline uint32_t cdiv_adj(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// ---------------------------------------------------------------------------
// Scratch of the allocation-free VJP (cth_de_vjp_noalloc_dev): the orbit
// records and end states of one call, in a file-scope device array.  Every
// GPU the library runs on holds its own copy from the code object's load
// until the process ends.  The host refers to a device's copy only through
// the address hipGetSymbolAddress gives with that device current (taken once
// per device at context creation); the array's name is a host-side stand-in.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[kVjpScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

int vjp_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(CTH_E_DEVICE, "device index above 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return CTH_OK;
    void* ptr = nullptr;
    int rc = check_hip(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (rc) return rc;
    g_scratch_ptr[device] = ptr;
    return CTH_OK;
}

void* vjp_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[device];
}

size_t vjp_workspace_bytes(uint32_t max_iters, size_t n) {
    const size_t stride = (size_t)max_iters + 1;
    const size_t orbit = (n * stride * sizeof(float4) + 255) & ~size_t(255);
    return orbit + n * sizeof(AdjRecord);
}

int launch_adjoint(Ctx& c, const cth_shape& s, const float* pts, const float* powers, uint32_t n,
                   const float* grad_d, float* d, float* grad_p, float* grad_power, float* normals, hipStream_t st,
                   void* workspace, size_t ws_bytes, float* grad_z0, float* grad_c) {
    if (n == 0) return CTH_OK;
    if (s.kind == CTH_SHAPE_SPHERE) return fail(CTH_E_UNSUPPORTED, "use the closed form for a sphere");
    MandelUniform u;
    int family = 0;
    int rc = resolve_mandel(s, u, family);
    if (rc) return rc;
    const size_t stride = (size_t)u.max_iters + 1;
    if (stride > 4096) return fail(CTH_E_INVALID, "max_iters above 4095 for the adjoint");
    size_t per_point = stride * sizeof(float4) + sizeof(AdjRecord);
    float4* orbit = nullptr;
    AdjRecord* rec = nullptr;
    size_t batch = 0;
    const bool own = workspace == nullptr;
    if (own) {
        // Workspace in batches of at most ~1 GiB of orbit records.
        batch = std::max<size_t>(1, std::min<size_t>(n, (size_t(1) << 30) / per_point));
        rc = check_hip(hipMallocAsync((void**)&orbit, batch * stride * sizeof(float4), st), "hipMallocAsync orbit");
        if (!rc) rc = check_hip(hipMallocAsync((void**)&rec, batch * sizeof(AdjRecord), st), "hipMallocAsync records");
    } else {
        // The caller's workspace holds the whole call: orbit records, then
        // (256-byte aligned) the end states.  The forward kernel writes every
        // record the reverse kernel reads, so nothing needs clearing.
        if (vjp_workspace_bytes(u.max_iters, n) > ws_bytes) return fail(CTH_E_INVALID, "workspace too small");
        batch = n;
        orbit = static_cast<float4*>(workspace);
        rec = reinterpret_cast<AdjRecord*>(static_cast<char*>(workspace) +
                                           ((n * stride * sizeof(float4) + 255) & ~size_t(255)));
    }
    for (size_t off = 0; off < n && !rc; off += batch) {
        uint32_t m = (uint32_t)std::min<size_t>(batch, n - off);
        SrcPoints src{pts + 3 * off, powers ? powers + off : nullptr, s.power};
        // kernel 1: parity forward with orbit recording
        uint32_t blocks = 0;
        WorkQueue q = make_queue(c, m, c.shape_adj, blocks, st);
        if (family == FAM_P8)
            hipLaunchKernelGGL((k_adj_forward<P8Exact>), dim3(blocks), dim3(c.shape_adj.block), 0, st, src, u, orbit, rec,
                               d ? d + off : nullptr, q);
        else
            hipLaunchKernelGGL((k_adj_forward<TrigExact>), dim3(blocks), dim3(c.shape_adj.block), 0, st, src, u, orbit,
                               rec, d ? d + off : nullptr, q);
        rc = check_hip(hipGetLastError(), "adjoint forward launch");
        if (rc) break;
        // kernel 2: reverse sweep
        AdjOut out{grad_d ? grad_d + off : nullptr,         grad_p ? grad_p + 3 * off : nullptr,
                   grad_power ? grad_power + off : nullptr, normals ? normals + 3 * off : nullptr,
                   grad_z0 ? grad_z0 + 3 * off : nullptr,   grad_c ? grad_c + 3 * off : nullptr};
        WorkQueue q2 = make_queue(c, m, c.shape_adj, blocks, st);
        if (family == FAM_P8)
            hipLaunchKernelGGL((k_adj_reverse<false>), dim3(blocks), dim3(c.shape_adj.block), 0, st, src, u, orbit, rec,
                               out, q2);
        else
            hipLaunchKernelGGL((k_adj_reverse<true>), dim3(blocks), dim3(c.shape_adj.block), 0, st, src, u, orbit, rec,
                               out, q2);
        rc = check_hip(hipGetLastError(), "adjoint reverse launch");
    }
    if (own && orbit) (void)hipFreeAsync(orbit, st);
    if (own && rec) (void)hipFreeAsync(rec, st);
    return rc;
}

// Forward-mode normals/gradients: the tangent evaluators in the lane-refill
// loop; one launch, no workspace.
int launch_tangent(Ctx& c, const cth_shape& s, const float* pts, const float* powers, uint32_t n,
                   const float* grad_d, float* d, float* grad_p, float* grad_power, float* normals, hipStream_t st) {
    if (n == 0) return CTH_OK;
    if (s.kind == CTH_SHAPE_SPHERE) return fail(CTH_E_UNSUPPORTED, "use the closed form for a sphere");
    MandelUniform u;
    int family = 0;
    int rc = resolve_mandel(s, u, family);
    if (rc) return rc;
    SrcPoints src{pts, powers, s.power};
    SinkGrad sink{d, normals, grad_p, grad_power, grad_d, u.max_iters};
    uint32_t blocks = 0;
    WorkQueue q = make_queue(c, n, c.shape_adj, blocks, st);
    if (family == FAM_P8)
        hipLaunchKernelGGL((k_refill<P8Tangent, SrcPoints, SinkGrad>), dim3(blocks), dim3(c.shape_adj.block), 0, st, src,
                           u, sink, q);
    else if (grad_power)  // the power column only when its gradient is asked for
        hipLaunchKernelGGL((k_refill<TrigTangent, SrcPoints, SinkGrad>), dim3(blocks), dim3(c.shape_adj.block), 0, st,
                           src, u, sink, q);
    else
        hipLaunchKernelGGL((k_refill<TrigNormal, SrcPoints, SinkGrad>), dim3(blocks), dim3(c.shape_adj.block), 0, st,
                           src, u, sink, q);
    return check_hip(hipGetLastError(), "tangent launch");
}

int launch_fd_normals(Ctx& c, const cth_shape& s, const float* pts, const float* powers, uint32_t n, float h,
                      float* d, float* grad_p, float* normals, hipStream_t st) {
    if (n == 0) return CTH_OK;
    if ((uint64_t)n * 7 >= (1ull << 31)) return fail(CTH_E_INVALID, "more than 2^31 / 7 points");
    float* taps = nullptr;
    int rc = check_hip(hipMallocAsync((void**)&taps, (size_t)n * 7 * sizeof(float), st), "hipMallocAsync taps");
    if (rc) return rc;
    SrcFD src{pts, powers, s.power, h};
    if (s.kind == CTH_SHAPE_SPHERE) {
        hipLaunchKernelGGL((k_sphere<SrcFD>), dim3(cdiv_adj((uint64_t)n * 7, 256)), dim3(256), 0, st, src, n * 7,
                           s.center[0], s.center[1], s.center[2], s.radius, taps, (uint32_t*)nullptr);
        rc = check_hip(hipGetLastError(), "fd sphere launch");
    } else {
        MandelUniform u;
        int family = 0;
        rc = resolve_mandel(s, u, family);
        if (!rc) rc = run_refill(c, family, u, src, SinkDist{taps, nullptr}, n * 7, st, c.shape_de);
    }
    if (!rc) {
        hipLaunchKernelGGL(k_fd_combine, dim3(cdiv_adj(n, 256)), dim3(256), 0, st, taps, n, normals, grad_p, d,
                           0.5f / h);
        rc = check_hip(hipGetLastError(), "fd combine launch");
    }
    (void)hipFreeAsync(taps, st);
    return rc;
}

}  // namespace cth
