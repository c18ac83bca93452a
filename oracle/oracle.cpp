// Oracle: a CPU restatement of cantucci's Mandelbulb distance-estimation path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, its C ABI
// or the Python host package) links, loads or calls this file.  It is used by
// `tests/` (as the checker), by `__graft_entry__.smoke()` (as the checker) and
// by the `cpu_baseline` leg of `bench.py` (the reference CPU path, timed).
//
// Every function restates the reference Rust source op for op, in IEEE f32 with
// round-to-nearest and *no* FMA contraction (rustc never contracts).  Build with
// `-ffp-contract=off -fno-fast-math` (see oracle/Makefile); x86-64 SSE gives the
// same default-NaN (0xFFC00000, sign bit set) that the reference produces.
// Transcendentals are glibc's (`logf`, `acosf`, `atan2f`, `sinf`, `cosf`,
// `powf`), which is what Rust std links on x86_64-linux-gnu.  One switch,
// `orc_set_ln_mode(1)`, replaces the final formula's `logf` with the f64 log
// rounded to f32 (the correctly rounded f32 ln outside hard cases, inputs
// whose ln lies within 2^-18 ulps of a rounding boundary); the trace entry
// `orc_sphere_trace_marked` then marks the rays whose march met a hard case.
//
// Citations are path:line in the cantucci source tree.
//
// Pins (checked in tests/test_oracle.py):
//   * DE, P8 polynomial      -- 0-iteration closed form, 1-iteration closed form,
//                               y/z reflection invariance (bitwise), FP64 latitude
//                               trig form on the y=0 plane, FP64 agreement.
//   * DE, z-axis branch      -- 1-D recurrence z <- |z|^8 + pz (brute force);
//                               origin gives a sign-negative NaN.
//   * DE, NaN inputs         -- NaN coordinates of one sign give a NaN
//                               distance of that sign (IEEE 754 propagation;
//                               P8, trig and sphere); with both signs the
//                               first NaN coordinate's (x86 keeps the first
//                               operand's NaN).
//   * DE, trig family        -- FP64 textbook formula, y-reflection equivariance.
//   * powi ladder            -- exact integer powers of small integers.
//   * Sphere                 -- closed forms.
//   * span sampling / sign   -- sphere sign field; dyadic grid coordinates exact.
//   * surface nets           -- sphere: watertight, every edge shared by 2 quads.
//   * vertex weights         -- linear corner fields: each vertex is the mean of
//                               the zero crossings mirrored about their edge
//                               midpoints; the +0 / -0 special cases.
//   * quad winding           -- sphere: every directed edge once, every
//                               triangle's right-hand normal outward.
//   * 6-tap stencil          -- sphere and zero-iteration closed forms.
//   * pinhole rays           -- image orientation and field of view.
//   * get_focii rays         -- exact-arithmetic near-plane geometry within the
//                               f32 cancellation bound; sphere hits analytic.
//   * sphere tracing         -- sphere: analytic ray/sphere hit.
//   * dual-number gradients  -- FP64 central differences; sphere closed form;
//                               0-iteration closed form.
//   * dd/dz0 vs dd/dc        -- central differences of an FP64 orbit with z0
//                               and c apart; one-body closed form (bailout at
//                               |z1|: dd/dc = 0.5 (ln r1 + 1) z1 / (r1 dr1));
//                               dd/dc = 0 at the cap of one body and at 0 bodies.
//   * ln mode 1, marked trace -- 50-digit decimal ln outside hard cases; a
//                               march restated step by step marks the same rays.
// Parity unpinned (no reference test or closed form fixes them; stated):
//   * LLVM's lowering order of `powi` (the ladder below) -- at most 1 ulp/op.
//   * cgmath's `normalize`, `lerp` operation order (vertex normals/positions).
//   * cgmath's general 4x4 `invert` (get_focii uses the closed-form inverse).

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

extern "C" {

// Same layout as `cth_shape` in include/cantucci_hip.h (kept in sync by a test).
struct orc_shape {
    int32_t kind;        // 0 Mandelbulb<P> (P8 polynomial when P == 8), 1 trig real power, 2 sphere
    float power;
    uint32_t max_iters;
    float bailout;
    float center[3];
    float radius;
};

}  // extern "C"

namespace {

enum { K_MANDELBULB = 0, K_MANDELBULB_TRIG = 1, K_SPHERE = 2 };

// ---------------------------------------------------------------------------
// Scalar helpers
// ---------------------------------------------------------------------------

// `Vec3::magnitude` (src/shape/mandelbulb.rs:412-418): DPPS with mask 0x71
// sums lanes as (x*x + y*y) + (z*z + 0), then `sqrtss`.
template <class T>
inline T magnitude(T x, T y, T z) {
    T s = (x * x + y * y) + z * z;
    return std::sqrt(s);
}

// `f32::powi` with a constant exponent.  LLVM expands `llvm.powi` into the
// binary-exponentiation ladder of SelectionDAG::ExpandPowI (the same order as
// compiler-rt's __powisf2): the running power x^(2^k) is multiplied into the
// result at every set bit k of n, lowest bit first.  n = 7 gives (x*x^2)*x^4,
// n = 8 gives ((x^2)^2)^2.  (Lowering order: parity unpinned, see header.)
template <class T>
inline T powi_ladder(T x, int n) {
    if (n == 0) return T(1);
    unsigned v = (unsigned)(n < 0 ? -n : n);
    T cur = x, res = T(0);
    bool have = false;
    while (v) {
        if (v & 1u) {
            res = have ? res * cur : cur;
            have = true;
        }
        v >>= 1;
        if (v) cur = cur * cur;
    }
    return n < 0 ? T(1) / res : res;
}

inline float lib_acos(float v) { return acosf(v); }
inline double lib_acos(double v) { return acos(v); }
inline float lib_atan2(float y, float x) { return atan2f(y, x); }
inline double lib_atan2(double y, double x) { return atan2(y, x); }
inline float lib_sin(float v) { return sinf(v); }
inline double lib_sin(double v) { return sin(v); }
inline float lib_cos(float v) { return cosf(v); }
inline double lib_cos(double v) { return cos(v); }
inline float lib_pow(float a, float b) { return powf(a, b); }
inline double lib_pow(double a, double b) { return pow(a, b); }
// 0: glibc's logf (the reference).  1: the f64 log rounded to f32, with
// t_ln_hard set when the f64 value lies within 2^-18 f32 ulps of a midpoint
// between two f32 values, where an f64 evaluation with a relative error
// below 2^-43 may round either way.
std::atomic<int> g_ln_mode{0};
thread_local bool t_ln_hard = false;

inline float cr_log(float v) {
    const double y = std::log((double)v);
    const float f = (float)y;
    if (std::isfinite(y) && (double)f != y) {
        const float nb = std::nextafterf(f, y > (double)f ? INFINITY : -INFINITY);
        const double mid = ((double)f + (double)nb) * 0.5;
        const double gap = std::fabs((double)nb - (double)f);
        if (std::fabs(y - mid) < 0x1p-18 * gap) t_ln_hard = true;
    }
    return f;
}
inline float lib_log(float v) { return g_ln_mode.load(std::memory_order_relaxed) == 1 ? cr_log(v) : logf(v); }
inline double lib_log(double v) { return log(v); }

// `Vec3::is_on_z_axis` (src/shape/mandelbulb.rs:421-426): PTEST of the x and y
// lanes against 0x7FFFFFFF, i.e. both are +0 or -0.
inline bool on_z_axis(float x, float y) {
    uint32_t bx, by;
    std::memcpy(&bx, &x, 4);
    std::memcpy(&by, &y, 4);
    return ((bx | by) & 0x7FFFFFFFu) == 0;
}
inline bool on_z_axis(double x, double y) { return x == 0.0 && y == 0.0; }

// ---------------------------------------------------------------------------
// The rotation ("triplex power") maps
// ---------------------------------------------------------------------------

// `rotate_inner_p8_scalar` (src/shape/mandelbulb.rs:148-200), association
// exactly as written in Rust (left to right; `28.0 * z6 * rxy2` is
// `(28.0*z6)*rxy2`).  Note the `- y8` term of x' (:187).
template <class T>
inline void rotate_p8_poly(T x, T y, T z, T& ox, T& oy, T& oz) {
    T x2 = x * x, x4 = x2 * x2, x6 = x4 * x2, x8 = x4 * x4;
    T y2 = y * y, y4 = y2 * y2, y6 = y4 * y2, y8 = y4 * y4;
    T z2 = z * z, z4 = z2 * z2, z6 = z4 * z2, z8 = z4 * z4;
    T rxy2 = x2 + y2;
    T rxy4 = rxy2 * rxy2;
    T rxy6 = rxy2 * rxy4;
    T rxy8 = rxy4 * rxy4;
    T a = T(1) + ((((z8 - (T(28) * z6) * rxy2) + (T(70) * z4) * rxy4) - (T(28) * z2) * rxy6) / rxy8);
    ox = a * ((((x8 - (T(28) * x6) * y2) + (T(70) * x4) * y4) - (T(28) * x2) * y6) - y8);
    oy = (((T(8) * a) * x) * y) * (((x6 - (T(7) * x4) * y2) + (T(7) * x2) * y4) - y6);
    oz = (((T(8) * z) * std::sqrt(rxy2)) * (z2 - rxy2)) * ((z4 - (T(6) * z2) * rxy2) + rxy4);
}

// Power law used by the trig family: integer P uses the `powi` ladder (the
// reference's const-generic `u8` power), real P uses `powf` (the real-power
// extension; same map, src/shape/mandelbulb.frag:1-39 uses `pow`).
template <class T>
inline T ipow_or_pow(T r, T P, int Pi, bool integral) {
    return integral ? powi_ladder(r, Pi) : lib_pow(r, P);
}

// `rotate_on_z_axis` (src/shape/mandelbulb.rs:114-126).
template <class T>
inline void rotate_z_axis(T x, T y, T z, T P, int Pi, bool integral, T& ox, T& oy, T& oz) {
    T old_radius = magnitude(x, y, z);
    T theta = lib_acos(z / old_radius);
    T new_radius = ipow_or_pow(old_radius, P, Pi, integral);
    theta = theta * P;
    ox = T(0);
    oy = T(0);
    oz = new_radius * lib_cos(theta);
}

// `rotate_inner_px_generic` (src/shape/mandelbulb.rs:128-146).
template <class T>
inline void rotate_trig(T x, T y, T z, T P, int Pi, bool integral, T& ox, T& oy, T& oz) {
    T old_radius = magnitude(x, y, z);
    T theta = lib_acos(z / old_radius);
    T phi = lib_atan2(y, x);
    T new_radius = ipow_or_pow(old_radius, P, Pi, integral);
    theta = theta * P;
    phi = phi * P;
    T vx = lib_sin(theta) * lib_cos(phi);
    T vy = lib_sin(phi) * lib_sin(theta);
    T vz = lib_cos(theta);
    ox = vx * new_radius;
    oy = vy * new_radius;
    oz = vz * new_radius;
}

struct MandelParams {
    bool real_power;     // kind 1: the power is a real parameter (d/dpower defined)
    int family;          // 0: P8 polynomial, 1: trig
    float P;             // power as f32
    int Pi;              // integer power (when integral)
    bool integral;
    uint32_t max_iters;
    float bailout;
};

// Resolve the reference's dispatch (src/shape/mandelbulb.rs:97-112): the
// z-axis branch first, then P == 8 -> polynomial, else the trig form.
inline MandelParams mandel_params(const orc_shape& s, float power) {
    MandelParams m;
    m.max_iters = s.max_iters;
    m.bailout = s.bailout;
    m.P = power;
    m.real_power = (s.kind == K_MANDELBULB_TRIG);
    if (s.kind == K_MANDELBULB) {
        m.Pi = (int)power;
        m.integral = true;
        m.family = (m.Pi == 8) ? 0 : 1;
    } else {
        float fl = std::floor(power);
        m.integral = (fl == power && std::fabs(power) < 64.0f);
        m.Pi = m.integral ? (int)power : 0;
        m.family = 1;
    }
    return m;
}

template <class T>
struct DeResult {
    T d;
    T r;
    T dr;
    uint32_t iters;
    bool bailed;
};

// `<Mandelbulb<P> as Shape>::min_distance_from` (src/shape/mandelbulb.rs:59-79).
template <class T>
inline DeResult<T> mandel_de(T px, T py, T pz, const MandelParams& m) {
    const T P = T(m.P);
    T x = px, y = py, z = pz;
    T dr = T(1);
    T r = T(0);
    uint32_t k = 0;
    bool bailed = false;
    for (uint64_t i = 0; i < m.max_iters; ++i) {
        r = magnitude(x, y, z);
        if (r > T(m.bailout)) {
            bailed = true;
            break;
        }
        // dr = r.powi(P - 1) * (P as f32) * dr + 1.0     (:73)
        T rp = ipow_or_pow(r, P - T(1), m.Pi - 1, m.integral);
        dr = ((rp * P) * dr) + T(1);
        T nx, ny, nz;
        if (on_z_axis(x, y)) {
            rotate_z_axis(x, y, z, P, m.Pi, m.integral, nx, ny, nz);
        } else if (m.family == 0) {
            rotate_p8_poly(x, y, z, nx, ny, nz);
        } else {
            rotate_trig(x, y, z, P, m.Pi, m.integral, nx, ny, nz);
        }
        x = nx + px;
        y = ny + py;
        z = nz + pz;
        ++k;
    }
    // let ln_r = r.ln() * r;  0.5 * ln_r / dr      (:77-78)
    T ln_r = lib_log(r) * r;
    DeResult<T> out;
    out.d = (T(0.5) * ln_r) / dr;
    out.r = r;
    out.dr = dr;
    out.iters = k;
    out.bailed = bailed;
    return out;
}

// `Sphere::min_distance_from` (src/shape/sphere.rs:33-35): cgmath
// `(center - p).magnitude() - radius`, magnitude = sqrt((x*x + y*y) + z*z).
template <class T>
inline T sphere_de(T px, T py, T pz, const orc_shape& s) {
    T dx = T(s.center[0]) - px, dy = T(s.center[1]) - py, dz = T(s.center[2]) - pz;
    return std::sqrt((dx * dx + dy * dy) + dz * dz) - T(s.radius);
}

inline float shape_de(const orc_shape& s, float px, float py, float pz, uint32_t* iters = nullptr) {
    if (s.kind == K_SPHERE) {
        if (iters) *iters = 0;
        return sphere_de(px, py, pz, s);
    }
    MandelParams m = mandel_params(s, s.power);
    DeResult<float> r = mandel_de(px, py, pz, m);
    if (iters) *iters = r.iters;
    return r.d;
}

// ---------------------------------------------------------------------------
// Forward-mode dual numbers (FP64) for the gradient oracle
// ---------------------------------------------------------------------------

// Tangent slots: d/dz0 (x, y, z), d/dpower, d/dc (x, y, z).  The point p is
// both the orbit's start z0 and the constant c added in every body
// (src/shape/mandelbulb.rs:60-74), so d/dp = d/dz0 + d/dc.
constexpr int NG = 7;

struct Dual {
    double v;
    double g[NG];
};

inline Dual dconst(double v) {
    Dual a;
    a.v = v;
    for (int i = 0; i < NG; ++i) a.g[i] = 0.0;
    return a;
}
inline Dual operator+(const Dual& a, const Dual& b) {
    Dual c;
    c.v = a.v + b.v;
    for (int i = 0; i < NG; ++i) c.g[i] = a.g[i] + b.g[i];
    return c;
}
inline Dual operator-(const Dual& a, const Dual& b) {
    Dual c;
    c.v = a.v - b.v;
    for (int i = 0; i < NG; ++i) c.g[i] = a.g[i] - b.g[i];
    return c;
}
inline Dual operator*(const Dual& a, const Dual& b) {
    Dual c;
    c.v = a.v * b.v;
    for (int i = 0; i < NG; ++i) c.g[i] = a.g[i] * b.v + a.v * b.g[i];
    return c;
}
inline Dual operator/(const Dual& a, const Dual& b) {
    Dual c;
    c.v = a.v / b.v;
    for (int i = 0; i < NG; ++i) c.g[i] = (a.g[i] * b.v - a.v * b.g[i]) / (b.v * b.v);
    return c;
}
inline Dual chain(const Dual& a, double v, double dv) {
    Dual c;
    c.v = v;
    for (int i = 0; i < NG; ++i) c.g[i] = dv * a.g[i];
    return c;
}
inline Dual dsqrt(const Dual& a) {
    double s = std::sqrt(a.v);
    return chain(a, s, 0.5 / s);
}
inline Dual dlog(const Dual& a) { return chain(a, std::log(a.v), 1.0 / a.v); }
inline Dual dacos(const Dual& a) { return chain(a, std::acos(a.v), -1.0 / std::sqrt(1.0 - a.v * a.v)); }
inline Dual dsin(const Dual& a) { return chain(a, std::sin(a.v), std::cos(a.v)); }
inline Dual dcos(const Dual& a) { return chain(a, std::cos(a.v), -std::sin(a.v)); }
inline Dual datan2(const Dual& y, const Dual& x) {
    Dual c;
    c.v = std::atan2(y.v, x.v);
    double den = x.v * x.v + y.v * y.v;
    for (int i = 0; i < NG; ++i) c.g[i] = (x.v * y.g[i] - y.v * x.g[i]) / den;
    return c;
}
// a^b with both operands dual: d(a^b) = a^b (b' ln a + b a'/a).
inline Dual dpow(const Dual& a, const Dual& b) {
    Dual c;
    c.v = std::pow(a.v, b.v);
    double la = std::log(a.v);
    for (int i = 0; i < NG; ++i) c.g[i] = c.v * (b.g[i] * la + b.v * a.g[i] / a.v);
    return c;
}
inline Dual dpowi(const Dual& a, int n) {
    // integer power with the same ladder; exact derivative n a^(n-1) a'
    double v = powi_ladder(a.v, n);
    double dv = (n == 0) ? 0.0 : double(n) * powi_ladder(a.v, n - 1);
    return chain(a, v, dv);
}
inline Dual dmag(const Dual& x, const Dual& y, const Dual& z) { return dsqrt((x * x + y * y) + z * z); }

inline Dual dpow_or_powi(const Dual& r, const Dual& P, int Pi, bool integral) {
    return integral ? dpowi(r, Pi) : dpow(r, P);
}

// The reference DE evaluated in FP64 forward mode; P carries d/dpower when the
// family is trig with real power.  Returns d with gradient (dd/dz0, dd/dP,
// dd/dc) in the slots above.
inline Dual mandel_de_dual(double px, double py, double pz, const MandelParams& m, uint32_t* iters) {
    Dual Px = dconst(px), Py = dconst(py), Pz = dconst(pz);  // c
    Px.g[4] = 1.0;
    Py.g[5] = 1.0;
    Pz.g[6] = 1.0;
    Dual Zx = dconst(px), Zy = dconst(py), Zz = dconst(pz);  // z0
    Zx.g[0] = 1.0;
    Zy.g[1] = 1.0;
    Zz.g[2] = 1.0;
    Dual P = dconst(m.P);
    if (m.real_power) P.g[3] = 1.0;
    // The real-power kind is differentiated as powf throughout (an integral
    // power only changes the f32 forward's lowering, not the function).
    const bool integral = m.integral && !m.real_power;
    Dual one = dconst(1.0);
    Dual x = Zx, y = Zy, z = Zz, dr = one, r = dconst(0.0);
    uint32_t k = 0;
    for (uint64_t i = 0; i < m.max_iters; ++i) {
        r = dmag(x, y, z);
        if (r.v > double(m.bailout)) break;
        dr = (dpow_or_powi(r, P - one, m.Pi - 1, integral) * P) * dr + one;
        Dual nx, ny, nz;
        if (x.v == 0.0 && y.v == 0.0) {
            Dual th = dacos(z / r);
            Dual R = dpow_or_powi(r, P, m.Pi, integral);
            nx = dconst(0.0);
            ny = dconst(0.0);
            nz = R * dcos(th * P);
        } else if (m.family == 0) {
            Dual x2 = x * x, x4 = x2 * x2, x6 = x4 * x2, x8 = x4 * x4;
            Dual y2 = y * y, y4 = y2 * y2, y6 = y4 * y2, y8 = y4 * y4;
            Dual z2 = z * z, z4 = z2 * z2, z6 = z4 * z2, z8 = z4 * z4;
            Dual w2 = x2 + y2, w4 = w2 * w2, w6 = w2 * w4, w8 = w4 * w4;
            Dual c28 = dconst(28), c70 = dconst(70), c7 = dconst(7), c6 = dconst(6), c8 = dconst(8);
            Dual a = one + ((((z8 - (c28 * z6) * w2) + (c70 * z4) * w4) - (c28 * z2) * w6) / w8);
            nx = a * ((((x8 - (c28 * x6) * y2) + (c70 * x4) * y4) - (c28 * x2) * y6) - y8);
            ny = (((c8 * a) * x) * y) * (((x6 - (c7 * x4) * y2) + (c7 * x2) * y4) - y6);
            nz = (((c8 * z) * dsqrt(w2)) * (z2 - w2)) * ((z4 - (c6 * z2) * w2) + w4);
        } else {
            Dual th = dacos(z / r);
            Dual ph = datan2(y, x);
            Dual R = dpow_or_powi(r, P, m.Pi, integral);
            th = th * P;
            ph = ph * P;
            nx = (dsin(th) * dcos(ph)) * R;
            ny = (dsin(ph) * dsin(th)) * R;
            nz = dcos(th) * R;
        }
        x = nx + Px;
        y = ny + Py;
        z = nz + Pz;
        ++k;
    }
    if (iters) *iters = k;
    return (dconst(0.5) * (dlog(r) * r)) / dr;
}

// ---------------------------------------------------------------------------
// Octree spans, grid sampling, sign classification, surface nets
// ---------------------------------------------------------------------------

struct V3 {
    float x, y, z;
};

// Naive-surface-nets span adjustment (src/mesh/buffer.rs:64-67):
//   overflow = (end - start) / res;  start' = start + -overflow;  end' = end + overflow
inline void adjust_span(const float* s, const float* e, uint32_t res, float* s2, float* e2) {
    float fres = float(res);
    for (int a = 0; a < 3; ++a) {
        float of = (e[a] - s[a]) / fres;
        s2[a] = s[a] + -of;
        e2[a] = e[a] + of;
    }
}

// Corner position of a GridTable fill (src/mesh/buffer.rs:78-80):
//   v = (x, y, z) / res  (component-wise f32 division);  p = start + across * v
inline float grid_coord(float s, float across, uint32_t i, float fres) {
    float v = float(i) / fres;
    return s + across * v;
}

// `is_sign_positive`: the IEEE sign bit (src/mesh/buffer.rs:131-135).
inline bool sign_positive(float d) {
    uint32_t b;
    std::memcpy(&b, &d, 4);
    return (b >> 31) == 0;
}

struct Vertex {
    float position[3];
    float normal[3];
    float distance_from_surface;
};

// Steps 2 and 3 of `naive_surface_nets` (src/mesh/buffer.rs:97-372) given the
// step-1 corner distances `dists` ((res+1)^3, GridTable order) of the span.
void mesh_from_dists(const orc_shape& shape, const float* span_start, const float* span_end, uint32_t res,
                     const float* dists, std::vector<Vertex>& verts, std::vector<uint32_t>& indices) {
    const uint32_t n = res + 1;
    float s[3], e[3];
    adjust_span(span_start, span_end, res, s, e);
    const float fres = float(res);
    auto D = [&](uint32_t x, uint32_t y, uint32_t z) { return dists[(size_t(x) * n + y) * n + z]; };

    // Step 2 (src/mesh/buffer.rs:97-275).
    float step[3] = {(e[0] - s[0]) / fres, (e[1] - s[1]) / fres, (e[2] - s[2]) / fres};
    const float off[8][3] = {
        {0.0f, 0.0f, 0.0f},          {0.0f, 0.0f, step[2]},          {0.0f, step[1], 0.0f},
        {0.0f, step[1], step[2]},    {step[0], 0.0f, 0.0f},          {step[0], 0.0f, step[2]},
        {step[0], step[1], 0.0f},    {step[0], step[1], step[2]},
    };
    static const uint8_t EDGES[12][2] = {{0, 4}, {1, 5}, {2, 6}, {3, 7}, {0, 2}, {1, 3},
                                         {4, 6}, {5, 7}, {0, 1}, {2, 3}, {4, 5}, {6, 7}};
    std::vector<uint32_t> points(size_t(res) * res * res, 0xFFFFFFFFu);
    verts.clear();
    const float delta[3] = {(0.7f * (e[0] - s[0])) / fres, (0.7f * (e[1] - s[1])) / fres,
                            (0.7f * (e[2] - s[2])) / fres};
    for (uint32_t x = 0; x < res; ++x)
        for (uint32_t y = 0; y < res; ++y)
            for (uint32_t z = 0; z < res; ++z) {
                float dc[8] = {D(x, y, z),         D(x, y, z + 1),     D(x, y + 1, z),     D(x, y + 1, z + 1),
                               D(x + 1, y, z),     D(x + 1, y, z + 1), D(x + 1, y + 1, z), D(x + 1, y + 1, z + 1)};
                bool first = sign_positive(dc[0]);
                bool all_same = true;
                for (int c = 1; c < 8; ++c)
                    if (sign_positive(dc[c]) != first) {
                        all_same = false;
                        break;
                    }
                if (all_same) continue;  // u32::MAX sentinel (:143-147)
                float p0[3] = {s[0] + float(x) * step[0], s[1] + float(y) * step[1], s[2] + float(z) * step[2]};
                int count = 0;
                float sum[3] = {0.0f, 0.0f, 0.0f};
                for (int ei = 0; ei < 12; ++ei) {
                    int from = EDGES[ei][0], to = EDGES[ei][1];
                    if (sign_positive(dc[from]) == sign_positive(dc[to])) continue;
                    float d_from, d_to;
                    if (dc[from] < 0.0f) {
                        d_from = dc[from];
                        d_to = dc[to];
                    } else {
                        d_from = -dc[from];
                        d_to = -dc[to];
                    }
                    float w;
                    if (d_to == d_from) {
                        w = 0.5f;
                    } else {
                        float dl = d_to - d_from;
                        w = (d_from + dl) / dl;
                    }
                    // lerp(a, b, t) = a * (1 - t) + b * t   (src/math.rs:45-49)
                    for (int a = 0; a < 3; ++a) {
                        float pa = p0[a] + off[from][a];
                        float pb = p0[a] + off[to][a];
                        float l = pa * (1.0f - w) + pb * w;
                        sum[a] = sum[a] + l;
                    }
                    ++count;
                }
                float fc = float(count);
                float p[3] = {0.0f + sum[0] / fc, 0.0f + sum[1] / fc, 0.0f + sum[2] / fc};
                // `Point3::origin() + v` adds to +0.0: normalises a -0.0 sum.
                float dist_p = shape_de(shape, p[0], p[1], p[2]);
                float g[3];
                for (int a = 0; a < 3; ++a) {
                    // p + unit_a * delta_a  and  p + unit_a * -delta_a (:259-264)
                    float pp[3], pm[3];
                    for (int b = 0; b < 3; ++b) {
                        float u = (a == b) ? 1.0f : 0.0f;
                        pp[b] = p[b] + u * delta[a];
                        pm[b] = p[b] + u * -delta[a];
                    }
                    g[a] = shape_de(shape, pp[0], pp[1], pp[2]) - shape_de(shape, pm[0], pm[1], pm[2]);
                }
                // cgmath normalize: v * (1 / |v|)
                float mag = std::sqrt((g[0] * g[0] + g[1] * g[1]) + g[2] * g[2]);
                float inv = 1.0f / mag;
                Vertex vx;
                for (int a = 0; a < 3; ++a) {
                    vx.position[a] = p[a];
                    vx.normal[a] = g[a] * inv;
                }
                vx.distance_from_surface = dist_p;
                verts.push_back(vx);
                points[(size_t(x) * res + y) * res + z] = uint32_t(verts.size() - 1);
            }
    auto PT = [&](uint32_t x, uint32_t y, uint32_t z) { return points[(size_t(x) * res + y) * res + z]; };

    // Step 3 (src/mesh/buffer.rs:288-372).
    indices.clear();
    for (uint32_t x = 0; x < res; ++x)
        for (uint32_t y = 0; y < res; ++y)
            for (uint32_t z = 0; z < res; ++z) {
                bool base = sign_positive(D(x, y, z));
                bool neg = D(x, y, z) < 0.0f;
                if (y > 0 && z > 0 && base != sign_positive(D(x + 1, y, z))) {
                    uint32_t v0 = PT(x, y - 1, z - 1), v1 = PT(x, y - 1, z), v2 = PT(x, y, z - 1), v3 = PT(x, y, z);
                    uint32_t q[6] = {v0, v2, v1, v1, v2, v3};
                    uint32_t r[6] = {v0, v1, v2, v1, v3, v2};
                    indices.insert(indices.end(), neg ? q : r, (neg ? q : r) + 6);
                }
                if (x > 0 && z > 0 && base != sign_positive(D(x, y + 1, z))) {
                    uint32_t v0 = PT(x - 1, y, z - 1), v1 = PT(x - 1, y, z), v2 = PT(x, y, z - 1), v3 = PT(x, y, z);
                    uint32_t q[6] = {v0, v1, v2, v1, v3, v2};
                    uint32_t r[6] = {v0, v2, v1, v1, v2, v3};
                    indices.insert(indices.end(), neg ? q : r, (neg ? q : r) + 6);
                }
                if (x > 0 && y > 0 && base != sign_positive(D(x, y, z + 1))) {
                    uint32_t v0 = PT(x - 1, y - 1, z), v1 = PT(x - 1, y, z), v2 = PT(x, y - 1, z), v3 = PT(x, y, z);
                    uint32_t q[6] = {v0, v2, v1, v1, v2, v3};
                    uint32_t r[6] = {v0, v1, v2, v1, v3, v2};
                    indices.insert(indices.end(), neg ? q : r, (neg ? q : r) + 6);
                }
            }
}

// Step 1 of `naive_surface_nets` (src/mesh/buffer.rs:63-83): corner DE grid
// over the adjusted span, cube() order, z fastest (src/util/iter.rs:49-64).
void sample_span(const orc_shape& shape, const float* span_start, const float* span_end, uint32_t res,
                 std::vector<float>& dists) {
    const uint32_t n = res + 1;
    float s[3], e[3];
    adjust_span(span_start, span_end, res, s, e);
    const float fres = float(res);
    float across[3] = {e[0] - s[0], e[1] - s[1], e[2] - s[2]};
    dists.assign(size_t(n) * n * n, 0.0f);
    for (uint32_t x = 0; x < n; ++x)
        for (uint32_t y = 0; y < n; ++y)
            for (uint32_t z = 0; z < n; ++z) {
                float px = grid_coord(s[0], across[0], x, fres);
                float py = grid_coord(s[1], across[1], y, fres);
                float pz = grid_coord(s[2], across[2], z, fres);
                dists[(size_t(x) * n + y) * n + z] = shape_de(shape, px, py, pz);
            }
}

// The reference's whole `naive_surface_nets` (src/mesh/buffer.rs:58-391).
void surface_nets(const orc_shape& shape, const float* span_start, const float* span_end, uint32_t res,
                  std::vector<float>& dists, std::vector<Vertex>& verts, std::vector<uint32_t>& indices) {
    sample_span(shape, span_start, span_end, res, dists);
    mesh_from_dists(shape, span_start, span_end, res, dists.data(), verts, indices);
}

// `create_spans` (src/octree/mod.rs:315-329): center = start + (end - start) / 2;
// child i has bit2 = +x, bit1 = +y, bit0 = +z half (src/octree/mod.rs:128-138).
inline void create_spans(const float* s, const float* e, float out[8][6]) {
    float c[3];
    for (int a = 0; a < 3; ++a) c[a] = s[a] + (e[a] - s[a]) / 2.0f;
    for (int i = 0; i < 8; ++i) {
        int bx = (i >> 2) & 1, by = (i >> 1) & 1, bz = i & 1;
        out[i][0] = bx ? c[0] : s[0];
        out[i][3] = bx ? e[0] : c[0];
        out[i][1] = by ? c[1] : s[1];
        out[i][4] = by ? e[1] : c[1];
        out[i][2] = bz ? c[2] : s[2];
        out[i][5] = bz ? e[2] : c[2];
    }
}

// A counter-based parallel-for over [0, n) with `threads` workers (the
// reference's `num_cpus::get()` pool, src/mesh/mod.rs:60-63).
template <class F>
void parallel_for(size_t n, int threads, size_t chunk, F f) {
    if (threads <= 1 || n <= chunk) {
        for (size_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::atomic<size_t> next(0);
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&]() {
            for (;;) {
                size_t b = next.fetch_add(chunk);
                if (b >= n) break;
                size_t e = std::min(n, b + chunk);
                for (size_t i = b; i < e; ++i) f(i);
            }
        });
    for (auto& th : pool) th.join();
}

}  // namespace

// ===========================================================================
// C entry points (ctypes)
// ===========================================================================
extern "C" {

int orc_shape_size(void) { return (int)sizeof(orc_shape); }

// DE of explicit f32 points (AoS xyz).  Optional outputs: iteration counts,
// final r and dr (the loop state the kernel must reproduce bit for bit).
// `powers` (nullable) gives a per-point power for the trig family.
void orc_de_points_f32(const orc_shape* shape, const float* pts, const float* powers, size_t n, float* d,
                       uint32_t* iters, float* r_out, float* dr_out, int threads) {
    const orc_shape s = *shape;
    parallel_for(n, threads, 256, [&](size_t i) {
        float px = pts[3 * i], py = pts[3 * i + 1], pz = pts[3 * i + 2];
        if (s.kind == K_SPHERE) {
            d[i] = sphere_de(px, py, pz, s);
            if (iters) iters[i] = 0;
            return;
        }
        MandelParams m = mandel_params(s, powers ? powers[i] : s.power);
        DeResult<float> res = mandel_de(px, py, pz, m);
        d[i] = res.d;
        if (iters) iters[i] = res.iters;
        if (r_out) r_out[i] = res.r;
        if (dr_out) dr_out[i] = res.dr;
    });
}

// The same map evaluated in FP64 (the accuracy oracle): inputs are f64.
void orc_de_points_f64(const orc_shape* shape, const double* pts, const float* powers, size_t n, double* d,
                       uint32_t* iters, int threads) {
    const orc_shape s = *shape;
    parallel_for(n, threads, 256, [&](size_t i) {
        double px = pts[3 * i], py = pts[3 * i + 1], pz = pts[3 * i + 2];
        if (s.kind == K_SPHERE) {
            d[i] = sphere_de(px, py, pz, s);
            if (iters) iters[i] = 0;
            return;
        }
        MandelParams m = mandel_params(s, powers ? powers[i] : s.power);
        DeResult<double> res = mandel_de<double>(px, py, pz, m);
        d[i] = res.d;
        if (iters) iters[i] = res.iters;
    });
}

// FP64 dual-number value and gradient: grad is n x 4 (dd/dpx, dd/dpy, dd/dpz,
// dd/dpower; dd/dpower is 0 unless the family is trig with real power).
void orc_de_grad_f64(const orc_shape* shape, const double* pts, const float* powers, size_t n, double* d,
                     double* grad, uint32_t* iters, int threads) {
    const orc_shape s = *shape;
    parallel_for(n, threads, 64, [&](size_t i) {
        double px = pts[3 * i], py = pts[3 * i + 1], pz = pts[3 * i + 2];
        if (s.kind == K_SPHERE) {
            double dx = px - s.center[0], dy = py - s.center[1], dz = pz - s.center[2];
            double m = std::sqrt(dx * dx + dy * dy + dz * dz);
            d[i] = m - s.radius;
            grad[4 * i] = dx / m;
            grad[4 * i + 1] = dy / m;
            grad[4 * i + 2] = dz / m;
            grad[4 * i + 3] = 0.0;
            if (iters) iters[i] = 0;
            return;
        }
        MandelParams m = mandel_params(s, powers ? powers[i] : s.power);
        uint32_t it = 0;
        Dual r = mandel_de_dual(px, py, pz, m, &it);
        d[i] = r.v;
        for (int k = 0; k < 3; ++k) grad[4 * i + k] = r.g[k] + r.g[4 + k];
        grad[4 * i + 3] = r.g[3];
        if (iters) iters[i] = it;
    });
}

// The gradient with the point's two roles apart: grad[n x 7] = dd/dz0 (the
// orbit's start), dd/dpower, dd/dc (the constant of every body).  A sphere
// has no orbit: its gradient goes to dd/dz0 and dd/dc is 0.
void orc_de_grad_split_f64(const orc_shape* shape, const double* pts, const float* powers, size_t n, double* d,
                           double* grad, uint32_t* iters, int threads) {
    const orc_shape s = *shape;
    std::vector<double> g4(4 * n);
    if (s.kind == K_SPHERE) {
        orc_de_grad_f64(shape, pts, powers, n, d, g4.data(), iters, threads);
        for (size_t i = 0; i < n; ++i) {
            for (int k = 0; k < 4; ++k) grad[7 * i + k] = g4[4 * i + k];
            for (int k = 4; k < 7; ++k) grad[7 * i + k] = 0.0;
        }
        return;
    }
    parallel_for(n, threads, 64, [&](size_t i) {
        MandelParams m = mandel_params(s, powers ? powers[i] : s.power);
        uint32_t it = 0;
        Dual r = mandel_de_dual(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], m, &it);
        d[i] = r.v;
        for (int k = 0; k < 7; ++k) grad[7 * i + k] = r.g[k];
        if (iters) iters[i] = it;
    });
}

// Regular grid (the 64^3 grid workload): p = start + (end - start) * (i / divisor), GridTable order
// (x slowest, z fastest).  The reference's own form of a grid coordinate.
void orc_de_grid_f32(const orc_shape* shape, const float* start, const float* end, uint32_t n, float divisor,
                     float* d, uint32_t* iters, int threads) {
    const orc_shape s = *shape;
    float across[3] = {end[0] - start[0], end[1] - start[1], end[2] - start[2]};
    size_t total = size_t(n) * n * n;
    parallel_for(total, threads, 4096, [&](size_t idx) {
        uint32_t k = uint32_t(idx % n), j = uint32_t((idx / n) % n), i = uint32_t(idx / (size_t(n) * n));
        float px = grid_coord(start[0], across[0], i, divisor);
        float py = grid_coord(start[1], across[1], j, divisor);
        float pz = grid_coord(start[2], across[2], k, divisor);
        uint32_t it = 0;
        d[idx] = shape_de(s, px, py, pz, &it);
        if (iters) iters[idx] = it;
    });
}

// Span sampling + sign classification (the octree workload): for each span (start xyz, end xyz),
// the (res+1)^3 corner distances of naive surface nets step 1 and one crossing
// bit per cell (bit = cell linear index (x*res + y)*res + z, 64 bits a word).
void orc_sample_spans(const orc_shape* shape, const float* spans, size_t nspans, uint32_t res, float* dists,
                      uint64_t* cross_bits, uint32_t* iters, int threads) {
    const orc_shape s = *shape;
    const uint32_t n = res + 1;
    const size_t per = size_t(n) * n * n;
    const size_t cells = size_t(res) * res * res;
    const size_t words = (cells + 63) / 64;
    parallel_for(nspans, threads, 1, [&](size_t li) {
        float a[3], b[3];
        adjust_span(spans + 6 * li, spans + 6 * li + 3, res, a, b);
        float across[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        float fres = float(res);
        float* D = dists + li * per;
        for (uint32_t x = 0; x < n; ++x)
            for (uint32_t y = 0; y < n; ++y)
                for (uint32_t z = 0; z < n; ++z) {
                    size_t idx = (size_t(x) * n + y) * n + z;
                    uint32_t it = 0;
                    D[idx] = shape_de(s, grid_coord(a[0], across[0], x, fres), grid_coord(a[1], across[1], y, fres),
                                      grid_coord(a[2], across[2], z, fres), &it);
                    if (iters) iters[li * per + idx] = it;
                }
        if (!cross_bits) return;
        uint64_t* W = cross_bits + li * words;
        for (size_t w = 0; w < words; ++w) W[w] = 0;
        for (uint32_t x = 0; x < res; ++x)
            for (uint32_t y = 0; y < res; ++y)
                for (uint32_t z = 0; z < res; ++z) {
                    bool first = sign_positive(D[(size_t(x) * n + y) * n + z]);
                    bool cross = false;
                    for (int c = 1; c < 8; ++c) {
                        uint32_t cx = x + ((c >> 2) & 1), cy = y + ((c >> 1) & 1), cz = z + (c & 1);
                        if (sign_positive(D[(size_t(cx) * n + cy) * n + cz]) != first) cross = true;
                    }
                    size_t cell = (size_t(x) * res + y) * res + z;
                    if (cross) W[cell / 64] |= (uint64_t(1) << (cell % 64));
                }
    });
}

// Full naive surface nets for one span (src/mesh/buffer.rs:30-391).  Returns
// the vertex and index counts; buffers may be null to query sizes.
// verts: 7 floats per vertex (the `Vertex` layout, src/mesh/mod.rs:255-261).
void orc_surface_nets(const orc_shape* shape, const float* span, uint32_t res, float* dists_out, float* verts_out,
                      size_t verts_cap, uint32_t* idx_out, size_t idx_cap, size_t* nverts, size_t* nidx) {
    std::vector<float> dists;
    std::vector<Vertex> verts;
    std::vector<uint32_t> indices;
    surface_nets(*shape, span, span + 3, res, dists, verts, indices);
    if (dists_out) std::memcpy(dists_out, dists.data(), dists.size() * sizeof(float));
    *nverts = verts.size();
    *nidx = indices.size();
    if (verts_out && verts_cap >= verts.size()) std::memcpy(verts_out, verts.data(), verts.size() * sizeof(Vertex));
    if (idx_out && idx_cap >= indices.size()) std::memcpy(idx_out, indices.data(), indices.size() * 4);
}

// Surface nets over a batch of spans, threaded over spans (the reference's
// per-leaf jobs, src/mesh/mod.rs:129-161).  `dists_in` (nullable) holds each
// span's (res+1)^3 step-1 distances; when null they are computed.  Fills the
// per-span exclusive offsets voff/ioff (n+1 each; vertex records of 7 floats,
// u32 indices local to the span) and, when verts/idx are non-null and their
// capacities suffice, the concatenated vertices and indices.
void orc_mesh_batch(const orc_shape* shape, const float* spans, size_t nspans, uint32_t res, const float* dists_in,
                    int threads, uint64_t* voff, uint64_t* ioff, float* verts, size_t vcap, uint32_t* idx,
                    size_t icap) {
    std::vector<std::vector<Vertex>> V(nspans);
    std::vector<std::vector<uint32_t>> I(nspans);
    const size_t per = size_t(res + 1) * (res + 1) * (res + 1);
    parallel_for(nspans, threads, 1, [&](size_t li) {
        const float* sp = spans + 6 * li;
        if (dists_in) {
            mesh_from_dists(*shape, sp, sp + 3, res, dists_in + li * per, V[li], I[li]);
        } else {
            std::vector<float> d;
            surface_nets(*shape, sp, sp + 3, res, d, V[li], I[li]);
        }
    });
    voff[0] = 0;
    ioff[0] = 0;
    for (size_t i = 0; i < nspans; ++i) {
        voff[i + 1] = voff[i] + V[i].size();
        ioff[i + 1] = ioff[i] + I[i].size();
    }
    if (verts && vcap >= voff[nspans])
        for (size_t i = 0; i < nspans; ++i)
            if (!V[i].empty()) std::memcpy(verts + 7 * voff[i], V[i].data(), V[i].size() * sizeof(Vertex));
    if (idx && icap >= ioff[nspans])
        for (size_t i = 0; i < nspans; ++i)
            if (!I[i].empty()) std::memcpy(idx + ioff[i], I[i].data(), I[i].size() * 4);
}

// ShapeMesh::update's split rule (src/mesh/mod.rs:91-104) over a flat leaf
// set: for each focus point (skipping rays whose depth is not finite, as
// get_focii drops misses, :229-242) the leaf around it -- the leaf whose
// half-open span contains it (Octree::leaf_around_mut + SpanExt::contains,
// src/octree/mod.rs:25-31,75-89) -- splits when
// camera.position.distance(focus) < 2 * |span.end.x - span.start.x|.
void orc_focus_split(const float* leaves, size_t n, const float* focii, const float* depth, size_t nf,
                     const float* eye, uint32_t* flags) {
    for (size_t i = 0; i < n; ++i) flags[i] = 0;
    for (size_t f = 0; f < nf; ++f) {
        if (depth && !std::isfinite(depth[f])) continue;
        const float* p = focii + 3 * f;
        for (size_t i = 0; i < n; ++i) {
            const float* s = leaves + 6 * i;
            const bool in = s[0] <= p[0] && s[1] <= p[1] && s[2] <= p[2] && p[0] < s[3] && p[1] < s[4] && p[2] < s[5];
            if (!in) continue;
            const float dx = eye[0] - p[0], dy = eye[1] - p[1], dz = eye[2] - p[2];
            const float dist = std::sqrt((dx * dx + dy * dy) + dz * dz);
            const float threshold = 2.0f * std::fabs(s[3] - s[0]);
            if (dist < threshold) flags[i] = 1;
            break;  // the leaf around the point is unique
        }
    }
}

void orc_create_spans(const float* span, float* out48) {
    float o[8][6];
    create_spans(span, span + 3, o);
    std::memcpy(out48, o, sizeof(o));
}

// Level-synchronous refinement (the octree workload's policy; the reference's own policy is
// camera driven, src/mesh/mod.rs:91-104): nodes of depth < min_depth always
// split (the reference pre-splits two levels, src/mesh/mod.rs:52-56); deeper
// nodes split when |DE(center)| < half_diag[depth] or DE is NaN.  Children are
// emitted in `create_spans` order, parents in level order.  Returns the number
// of leaves at max_depth that pass the same test (the surface leaves), written
// to `leaves` (6 floats each) when it fits `cap`.  level_counts[d] = nodes
// evaluated at depth d.
size_t orc_refine(const orc_shape* shape, const float* root, uint32_t min_depth, uint32_t max_depth,
                  const float* half_diag, float* leaves, size_t cap, uint64_t* level_counts, int threads) {
    const orc_shape s = *shape;
    std::vector<float> cur(root, root + 6), next;
    size_t out = 0;
    for (uint32_t depth = 0; depth <= max_depth; ++depth) {
        size_t nn = cur.size() / 6;
        if (level_counts) level_counts[depth] = nn;
        std::vector<uint8_t> split(nn, 0);
        parallel_for(nn, threads, 64, [&](size_t i) {
            if (depth < min_depth) {
                split[i] = 1;  // forced split: no distance evaluated
                return;
            }
            const float* sp = &cur[6 * i];
            float c[3];
            for (int a = 0; a < 3; ++a) c[a] = sp[a] + (sp[3 + a] - sp[a]) / 2.0f;
            float dd = shape_de(s, c[0], c[1], c[2]);
            split[i] = std::isnan(dd) || (std::fabs(dd) < half_diag[depth]);
        });
        if (depth == max_depth) {
            for (size_t i = 0; i < nn; ++i)
                if (split[i]) {
                    if (leaves && out < cap) std::memcpy(leaves + 6 * out, &cur[6 * i], 6 * sizeof(float));
                    ++out;
                }
            break;
        }
        next.clear();
        for (size_t i = 0; i < nn; ++i)
            if (split[i]) {
                float o[8][6];
                create_spans(&cur[6 * i], &cur[6 * i + 3], o);
                for (int c = 0; c < 8; ++c) next.insert(next.end(), o[c], o[c] + 6);
            }
        cur.swap(next);
    }
    return out;
}

// Primary-ray directions for a pinhole camera (f32, this op order):
//   u = ((2 (px + 0.5)) / W - 1) * tan_half * aspect
//   v = (1 - (2 (py + 0.5)) / H) * tan_half
//   d = (fwd + right * u) + up * v ;  d = d * (1 / |d|)
// pose = {eye[3], right[3], up[3], fwd[3], tan_half, aspect}.  The reference's
// focus rays (src/mesh/mod.rs:209-231) are the near-plane form of this.
void orc_ray_dirs(const float* pose, uint32_t W, uint32_t H, uint32_t row0, uint32_t rows, float* dirs) {
    const float* right = pose + 3;
    const float* up = pose + 6;
    const float* fwd = pose + 9;
    float th = pose[12], asp = pose[13];
    for (uint32_t yy = 0; yy < rows; ++yy)
        for (uint32_t x = 0; x < W; ++x) {
            uint32_t y = row0 + yy;
            float u = (((2.0f * (float(x) + 0.5f)) / float(W)) - 1.0f) * th * asp;
            float v = (1.0f - ((2.0f * (float(y) + 0.5f)) / float(H))) * th;
            float d[3];
            for (int a = 0; a < 3; ++a) d[a] = (fwd[a] + right[a] * u) + up[a] * v;
            float m = std::sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
            float inv = 1.0f / m;
            size_t o = 3 * (size_t(yy) * W + x);
            for (int a = 0; a < 3; ++a) dirs[o + a] = d[a] * inv;
        }
}

// The focus rays of `ShapeMesh::get_focii` (src/mesh/mod.rs:205-231) for the
// reference camera cam = {position[3], direction[3], fov, aspect, near, far}
// (src/camera.rs:16-20,114-127; `direction` unit length, as Camera::new
// stores it).  Ray (x, y) of iter::square (x outer, src/util/iter.rs:55-87)
// has index x * fp + y.  Every step is the f32 operation of the reference:
//   view = look_at_rh(pos, pos + dir, (0, 0, 1))      (src/camera.rs:69-75)
//     f = normalize((pos + dir) - pos); s = normalize(cross(f, up));
//     u = cross(s, f)  (cgmath's look_to_rh)
//   height = (2 near) tan(fov 0.5); width = height aspect  (:186-189)
//   top_left = (-width / 2, -height / 2, -near), bottom_right likewise (:90-103)
//   size = (width / fp, height / fp); center_diff = (br - tl) / (2 fp)
//   center = (top_left + (x size_h, y size_v, 0)) + center_diff
//   p = inv_view * (center, 1);  dir = normalize(p - pos)
// The inverse of the view matrix is taken in closed form: columns s, u, -f
// and pos, since the view transform is a rotation and a translation;
// cgmath's `invert` is the general cofactor inverse, whose entries can
// differ in the last bits (parity unpinned: cgmath is not vendored).  The
// product follows cgmath's column order ((c0 x + c1 y) + c2 z) + c3 w, and
// its w is 1, so from_homogeneous leaves p unchanged.  Writes fp^2 unit
// directions (3 floats each) to `dirs`.
void orc_focus_dirs(const float* cam, uint32_t fp, float* dirs) {
    const float* pos = cam;
    const float* dir = cam + 3;
    const float fov = cam[6], aspect = cam[7], near_plane = cam[8];
    auto normalize = [](float v[3]) {
        // cgmath InnerSpace::normalize: v * (1 / |v|), |v| = sqrt((x x + y y) + z z)
        float m = std::sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
        float inv = 1.0f / m;
        for (int a = 0; a < 3; ++a) v[a] = v[a] * inv;
    };
    auto cross = [](const float a[3], const float b[3], float o[3]) {
        o[0] = a[1] * b[2] - a[2] * b[1];
        o[1] = a[2] * b[0] - a[0] * b[2];
        o[2] = a[0] * b[1] - a[1] * b[0];
    };
    const float up[3] = {0.0f, 0.0f, 1.0f};
    float f[3], s[3], u[3];
    for (int a = 0; a < 3; ++a) f[a] = (pos[a] + dir[a]) - pos[a];
    normalize(f);
    cross(f, up, s);
    normalize(s);
    cross(s, f, u);
    const float height = (2.0f * near_plane) * std::tan(fov * 0.5f);
    const float width = height * aspect;
    const float tl[3] = {-width / 2.0f, -height / 2.0f, -near_plane};
    const float br[3] = {width / 2.0f, height / 2.0f, -near_plane};
    const float ffp = float(fp);
    const float size_h = width / ffp, size_v = height / ffp;
    float cd[3];
    for (int a = 0; a < 3; ++a) cd[a] = (br[a] - tl[a]) / (2.0f * ffp);
    for (uint32_t x = 0; x < fp; ++x)
        for (uint32_t y = 0; y < fp; ++y) {
            const float off[3] = {float(x) * size_h, float(y) * size_v, 0.0f};
            float c[3];
            for (int a = 0; a < 3; ++a) c[a] = (tl[a] + off[a]) + cd[a];
            float d[3];
            for (int a = 0; a < 3; ++a) {
                float p = ((s[a] * c[0] + u[a] * c[1]) + (-f[a]) * c[2]) + pos[a] * 1.0f;
                d[a] = p - pos[a];
            }
            normalize(d);
            float* o = dirs + 3 * (size_t(x) * fp + y);
            for (int a = 0; a < 3; ++a) o[a] = d[a];
        }
}

// Sphere tracing (the `get_focii` marcher, src/mesh/mod.rs:229-241):
//   d = DE(pos); pos += dir * d; t += d;
//   hit  if d < eps_abs + eps_rel * t  -> depth = t   (reference: eps_abs 1e-6, eps_rel 0)
//   miss if t > t_max                  -> depth = +inf (reference: no escape)
// after max_steps without either -> depth = +inf, steps = max_steps.
// hit_pos (nullable) receives the position after the last step.
// orc_sphere_trace_marked also sets hard[i] (nullable) to 1 when ray i's
// march took the final ln of a hard case (ln mode 1 only), else 0.
void orc_sphere_trace_marked(const orc_shape* shape, const float* origins, const float* dirs, size_t n,
                             int origin_stride, uint32_t max_steps, float eps_abs, float eps_rel, float t_max,
                             float* depth, uint32_t* steps, float* hit_pos, uint8_t* hard, int threads) {
    const orc_shape s = *shape;
    parallel_for(n, threads, 64, [&](size_t i) {
        t_ln_hard = false;
        const float* o = origins + (origin_stride ? 3 * i : 0);
        float px = o[0], py = o[1], pz = o[2];
        float dx = dirs[3 * i], dy = dirs[3 * i + 1], dz = dirs[3 * i + 2];
        float t = 0.0f;
        float out = INFINITY;
        uint32_t k = 0;
        while (k < max_steps) {
            float d = shape_de(s, px, py, pz);
            px = px + dx * d;
            py = py + dy * d;
            pz = pz + dz * d;
            t = t + d;
            ++k;
            if (d < eps_abs + eps_rel * t) {
                out = t;
                break;
            }
            if (t > t_max) break;
        }
        depth[i] = out;
        steps[i] = k;
        if (hit_pos) {
            hit_pos[3 * i] = px;
            hit_pos[3 * i + 1] = py;
            hit_pos[3 * i + 2] = pz;
        }
        if (hard) hard[i] = t_ln_hard ? 1 : 0;
    });
}

void orc_sphere_trace(const orc_shape* shape, const float* origins, const float* dirs, size_t n, int origin_stride,
                      uint32_t max_steps, float eps_abs, float eps_rel, float t_max, float* depth, uint32_t* steps,
                      float* hit_pos, int threads) {
    orc_sphere_trace_marked(shape, origins, dirs, n, origin_stride, max_steps, eps_abs, eps_rel, t_max, depth, steps,
                            hit_pos, nullptr, threads);
}

// Final-formula ln: 0 glibc's logf (default), 1 the f64 log rounded to f32.
void orc_set_ln_mode(int mode) { g_ln_mode.store(mode == 1 ? 1 : 0); }

// 6-tap central-difference gradient, f32 (the reference's normal stencil,
// src/mesh/buffer.rs:256-266, with a caller-given step h per axis): returns
// the un-normalised differences DE(p + h e_a) - DE(p - h e_a).
void orc_fd_diff(const orc_shape* shape, const float* pts, const float* powers, size_t n, float h, float* diff,
                 int threads) {
    const orc_shape s = *shape;
    parallel_for(n, threads, 64, [&](size_t i) {
        const float* p = pts + 3 * i;
        MandelParams m = mandel_params(s, powers ? powers[i] : s.power);
        for (int a = 0; a < 3; ++a) {
            float pp[3], pm[3];
            for (int b = 0; b < 3; ++b) {
                float u = (a == b) ? 1.0f : 0.0f;
                pp[b] = p[b] + u * h;
                pm[b] = p[b] + u * -h;
            }
            float dp, dm;
            if (s.kind == K_SPHERE) {
                dp = sphere_de(pp[0], pp[1], pp[2], s);
                dm = sphere_de(pm[0], pm[1], pm[2], s);
            } else {
                dp = mandel_de(pp[0], pp[1], pp[2], m).d;
                dm = mandel_de(pm[0], pm[1], pm[2], m).d;
            }
            diff[3 * i + a] = dp - dm;
        }
    });
}

// powi ladder exposed for its own pin.
float orc_powi_f32(float x, int n) { return powi_ladder(x, n); }

}  // extern "C"
