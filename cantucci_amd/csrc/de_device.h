// Device-side Mandelbulb / sphere distance estimators for gfx950.
//
// Two arithmetic modes share one per-lane interface (start / mag / body /
// finish), so the lane-refill work loop in work_loop.h can drive any of them:
//
//   * Exact evaluators (parity mode) reproduce the reference's f32 operation
//     sequence (src/shape/mandelbulb.rs:59-79, 96-200): no FMA contraction
//     (the library is compiled with -ffp-contract=off), IEEE-correct division
//     and square root, denormals kept, the same association order.  The loop
//     state (r, dr, iteration count) is therefore bit-identical to the
//     reference; only the final `ln` (correctly rounded here, glibc's logf in
//     the reference) may differ by an ulp.  NaN results are canonicalised to
//     the NaN x86 would give (canon_nan), so that `is_sign_positive`
//     classifies them as the reference does.
//
//   * Fast evaluators (perf mode) compute the same map with explicit FMAs,
//     shared subexpressions, hardware reciprocal / square root and the
//     squared-radius bailout test.  They are not bit-exact (the chaotic tail of
//     non-escaping points moves by more than 1e-5); their error is measured
//     and reported, never used for parity claims.
//
// Packed-FP32 instructions are disabled for the whole library at compile time
// (-Xclang -target-feature -Xclang -packed-fp32-ops).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "cantucci_hip.h"
#include "log_table.h"

namespace cth {

// ---------------------------------------------------------------------------
// Bit-level helpers
// ---------------------------------------------------------------------------

// Lanes of the wave (active ones only) where `p` holds.  The builtin keeps
// the predicate in its scalar lane mask; HIP's __ballot first widens it to a
// vector register and compares it again, two VALU instructions per call.
__device__ __forceinline__ uint64_t wave_mask(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// `Vec3::is_on_z_axis` (src/shape/mandelbulb.rs:421-426): x and y both +-0.
__device__ __forceinline__ bool on_z_axis(float x, float y) {
    return ((__float_as_uint(x) | __float_as_uint(y)) & 0x7fffffffu) == 0u;
}

// IEEE-correct f32 square root and division.  The library is compiled with
// -fhip-fp32-correctly-rounded-divide-sqrt, under which `/` and
// __builtin_sqrtf lower to correctly rounded sequences.  (HIP's `__fsqrt_rn`
// is the native approximate square root unless OCML_BASIC_ROUNDED_OPERATIONS
// is defined, so it is not used.)
__device__ __forceinline__ float sqrt_rn(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ float div_rn(float a, float b) { return a / b; }

// a / b without operand scaling or special-case handling: the reciprocal
// estimate, one Newton step on it, the quotient, and one remainder
// correction of the quotient (Markstein's final step, "IA-64 and elementary
// functions", 2000): 6 instructions instead of the 11 of the compiler's
// correctly rounded sequence.  Where no operand, reciprocal, quotient or
// remainder over- or underflows, every step scales exactly with powers of
// two, so a check over all pairs of significands covers every such pair:
// tests/hip/div_check.hip finds the result equal to `a / b` on all 2^46 of
// them (and on 2^32 random pairs over the P8 body's operand range).  A second
// remainder correction, as in the compiler's sequence, changes nothing there;
// without the Newton step 26,621 pairs differ.  Callers guarantee the range.
__device__ __forceinline__ float div_rn_unscaled(float a, float b) {
    float y = __builtin_amdgcn_rcpf(b);
    const float e = __builtin_fmaf(-b, y, 1.0f);
    y = __builtin_fmaf(e, y, y);
    const float q = a * y;
    const float r = __builtin_fmaf(-b, q, a);
    return __builtin_fmaf(r, y, q);
}

// The correctly rounded square root of x in [2^-96, FLT_MAX]: the hardware
// reciprocal square root estimate y, g = x y, h = y / 2, and one Markstein
// correction g + (x - g^2) h ("IA-64 and elementary functions", 2000).  Over
// that range no step over- or underflows; tests/hip/sqrt_exhaustive.hip finds
// the result equal to __builtin_sqrtf on every input (through sqrt_rn_wave).
// Infinite input gives NaN, so callers keep it out.  4 instructions and the
// estimate, where the compiler's sequence has 8 and the estimate.
__device__ __forceinline__ float sqrt_rn_unscaled(float x) {
    const float y = __builtin_amdgcn_rsqf(x);
    const float g = x * y, h = 0.5f * y;
    const float r = __builtin_fmaf(-g, g, x);
    return __builtin_fmaf(r, h, g);
}
// sqrt_rn_unscaled when every lane of the wave is in its range (false for
// NaN), else the full sequence.
__device__ __forceinline__ float sqrt_rn_wave(float x) {
    if (__builtin_expect(wave_mask(!(x >= 0x1p-96f && x <= 0x1.fffffep127f)) == 0, 1)) return sqrt_rn_unscaled(x);
    return __builtin_sqrtf(x);
}

// NaN sign of the reference.  x86 SSE arithmetic creates the default NaN
// 0xFFC00000 (sign set) and propagates NaN operands unchanged (where two NaN
// operands meet, the first source operand's); glibc's acosf returns +NaN
// (0x7FC00000) for arguments outside [-1, 1].  An orbit's NaNs have one
// source:
//  * a NaN input coordinate.  The first magnitude (x*x + y*y) + z*z meets the
//    coordinates in x, y, z order, so r carries the first NaN coordinate's
//    sign, and every later NaN of the orbit is a propagation of r or of that
//    coordinate (the orbit adds the point back each body);
//  * acosf(z / r) = acosf(+-inf), which happens only when r underflows to 0
//    with z != 0 and the input has no NaN: +NaN.
// A NaN distance carries the orbit's sign when r or dr is NaN at the end;
// otherwise it was created by the final formula itself (ln 0 * 0) and is the
// default NaN.  `is_sign_positive` then classifies these NaNs exactly as the
// reference does.  (A point with a NaN coordinate and also an infinite or
// overflowing one is not covered: there the rotation creates a default NaN
// (0 * inf, inf - inf) that meets the input's, and which one survives
// depends on the compiled operand order.)
__device__ __forceinline__ float canon_nan(float d, bool pos) {
    return (d != d) ? __uint_as_float(pos ? 0x7fc00000u : 0xffc00000u) : d;
}
__device__ __forceinline__ float canon_nan(float d) { return canon_nan(d, false); }
// The orbit's NaN sign at the start: that of the first NaN coordinate in
// x, y, z order, positive only if there is one and its sign bit is clear.
__device__ __forceinline__ bool input_pos_nan(float x, float y, float z) {
    const float f = (x != x) ? x : ((y != y) ? y : z);
    return f != f && !__builtin_signbit(f);
}
// ln for the final formula, correctly rounded to f32 (up to a ~2^-18 chance
// per input of a double-rounding miss).  glibc's logf, which the reference
// calls, returns the correctly rounded value for 99.6% of f32 inputs in
// [0.25, 64), so this agrees with the reference far more often than a
// faithfully rounded f32 logf.  Method: Tang's table-driven logarithm in f64
// (ACM TOMS 16(4), 1990): x = 2^e m, m in [1, 2); interval i = the top five
// mantissa bits; ln x = e ln2 + logc_i + log1p(m invc_i - 1) with
// |m invc_i - 1| < 1/32 and a degree-8 Taylor polynomial (relative error
// < 2^-43).  Near 1 (x in [1 - 1/32, 1 + 1/32)) log1p(x - 1) is used directly so
// the result keeps its relative accuracy.  Zero, subnormal, negative, inf and
// NaN inputs take ocml's f64 log.
__device__ __forceinline__ float log_cr(float x) {
    uint32_t ix = __float_as_uint(x);
    if (ix - 0x00800000u >= 0x7f800000u - 0x00800000u) return (float)log((double)x);
    int e = (int)(ix >> 23) - 127;
    uint32_t mb = ix & 0x7fffffu;
    int i = (int)(mb >> 18);
    double r, base;
    if ((e == 0 && i == 0) || (e == -1 && i == 31)) {
        r = (double)x - 1.0;
        base = 0.0;
    } else {
        double m = (double)__uint_as_float(mb | 0x3f800000u);
        r = __builtin_fma(m, kLogInvC[i], -1.0);
        base = __builtin_fma((double)e, 0x1.62e42fefa39efp-1, kLogC[i]);
    }
    double p = -0.125;
    p = __builtin_fma(p, r, 1.0 / 7.0);
    p = __builtin_fma(p, r, -1.0 / 6.0);
    p = __builtin_fma(p, r, 0.2);
    p = __builtin_fma(p, r, -0.25);
    p = __builtin_fma(p, r, 1.0 / 3.0);
    p = __builtin_fma(p, r, -0.5);
    p = __builtin_fma(p, r, 1.0);
    return (float)__builtin_fma(p, r, base);
}

__device__ __forceinline__ float finish_exact(float r, float dr, bool orbit_pos_nan) {
    float d = div_rn(0.5f * (log_cr(r) * r), dr);
    return canon_nan(d, orbit_pos_nan && (r != r || dr != dr));
}

// `f32::powi` with the LLVM ExpandPowI ladder: multiply the running square
// into the result at each set bit of |n|, lowest first; a negative n takes
// the reciprocal of the ladder's result.
__device__ __forceinline__ float powi_ladder(float x, int n) {
    if (n == 0) return 1.0f;
    float cur = x, res = 1.0f;
    bool have = false;
    unsigned v = n < 0 ? 0u - (unsigned)n : (unsigned)n;
    while (v) {
        if (v & 1u) {
            res = have ? res * cur : cur;
            have = true;
        }
        v >>= 1;
        if (v) cur = cur * cur;
    }
    return n < 0 ? div_rn(1.0f, res) : res;
}

// ---------------------------------------------------------------------------
// Evaluator parameters (uniform across a launch)
// ---------------------------------------------------------------------------
struct MandelUniform {
    uint32_t max_iters;
    float bailout;
    float power;    // used when no per-point power is given
    int ipower;     // integer power when `integral`
    int integral;   // 1: every point uses ipower (reference Mandelbulb<P>);
                    // 0: real power per point (powf unless integral and < 64)
    float center[3];  // sphere evaluator only
    float radius;
};

// ---------------------------------------------------------------------------
// Exact P8 polynomial evaluator (reference `Mandelbulb<8>`)
// ---------------------------------------------------------------------------
struct P8Exact {
    static constexpr int kWavesPerSimd = 8;  // fits 64 VGPRs in the refill kernel
    // The evaluator has last_dr(): the work loops end a point that reaches
    // the iteration cap after max_iters - 1 bodies and apply the last body's
    // dr update alone (see last_dr).
    static constexpr bool kLastDr = true;
    // The work loops may use guard(), mag_tick() and body_guarded().
    static constexpr bool kCarriedGuard = true;
    float px, py, pz, x, y, z, dr;
    float x2, y2, w2, z2;  // squares of the last magnitude, reused by rotate
    bool pos_nan;          // the orbit's NaNs are positive (canon_nan)
    bool fast = false;     // guard(): the next body may take the unscaled sequences

    __device__ __forceinline__ void start(float ax, float ay, float az, float, const MandelUniform&) {
        px = x = ax;
        py = y = ay;
        pz = z = az;
        dr = 1.0f;
        pos_nan = input_pos_nan(ax, ay, az);
    }
    // r = sqrt((x*x + y*y) + z*z) (:412-418).  x*x, y*y, x*x + y*y are the same
    // IEEE operations as the rotate's x2, y2, rxy2 (:153-168), so they are
    // computed once.
    __device__ __forceinline__ float mag() {
        x2 = x * x;
        y2 = y * y;
        w2 = x2 + y2;
        z2 = z * z;
        return sqrt_rn_wave(w2 + z2);
    }
    // The wave test of body() (see there) on the current squares, kept for the
    // next body_guarded().  The work loops call it after starting points.
    __device__ __forceinline__ void guard(const MandelUniform& u) {
        fast = u.bailout <= 16.0f && wave_mask(!(w2 >= 0x1p-15f)) == 0;
    }
    // mag() after a loop body, with the wave test of the next body made here
    // and also used for this square root: m = w2 + z2 >= w2 >= 2^-15, and a
    // point that was live before the body (r <= bailout <= 16) has every
    // coordinate of its rotation below 2^46 (each is a degree-8 form in x, y,
    // z whose denominator w8 cancels against the numerator's w^8 factor), so
    // m < 2^94: the range of sqrt_rn_unscaled.  Lanes without a live point
    // may be outside it; their results are dropped.
    __device__ __forceinline__ float mag_tick(const MandelUniform& u) {
        x2 = x * x;
        y2 = y * y;
        w2 = x2 + y2;
        z2 = z * z;
        guard(u);
        const float m = w2 + z2;
        return fast ? sqrt_rn_unscaled(m) : sqrt_rn_wave(m);
    }
    __device__ __forceinline__ void body(float r, const MandelUniform& u, bool live) {
        body_with(r, live, u.bailout <= 16.0f && wave_mask(!(w2 >= 0x1p-15f)) == 0);
    }
    // body() with the wave test made by the last guard() or mag_tick().
    __device__ __forceinline__ void body_guarded(float r, const MandelUniform&, bool live) { body_with(r, live, fast); }
    // One loop body for every lane of the wave, committed only where `live`
    // (lanes without a live point compute on stale state and discard it).
    // The z-axis branch is rare, so it runs behind a wave-uniform test.
    __device__ __forceinline__ void body_with(float r, bool live, bool unscaled) {
        // dr = r.powi(7) * 8.0 * dr + 1.0      (:73)
        float r2 = r * r;
        float r4 = r2 * r2;
        float r7 = (r * r2) * r4;
        float ndr = ((r7 * 8.0f) * dr) + 1.0f;
        // rotate_inner_p8_scalar (:148-200)
        float x4 = x2 * x2, x6 = x4 * x2, x8 = x4 * x4;
        float y4 = y2 * y2, y6 = y4 * y2, y8 = y4 * y4;
        float z4 = z2 * z2, z6 = z4 * z2, z8 = z4 * z4;
        float w4 = w2 * w2, w6 = w2 * w4, w8 = w4 * w4;
        float num = ((z8 - (28.0f * z6) * w2) + (70.0f * z4) * w4) - (28.0f * z2) * w6;
        // A live lane has r <= bailout (the loop stops at r > bailout), so
        // with bailout <= 16: w8 <= 2^32 and |num| < 127 * 2^32.  If every
        // lane also has w2 >= 2^-15 (false for NaN), then w8 >= 2^-60 and no
        // step of the unscaled division over- or underflows unless
        // |num / w8| < 2^-25, where 1 + num / w8 rounds to 1 either way; a is
        // then the same float as with the full sequence.  The same test
        // keeps a live lane's w2 in the square root's range [2^-96, 256]
        // (lanes without a live point may be outside: their results are
        // dropped), and it rules out the z axis (x = y = 0 gives w2 = 0).
        // `unscaled` is that test: bailout <= 16 and every lane's w2 >= 2^-15.
        float q, sw2;
        if (unscaled) {
            q = div_rn_unscaled(num, w8);
            sw2 = sqrt_rn_unscaled(w2);
        } else {
            q = div_rn(num, w8);
            sw2 = sqrt_rn_wave(w2);
        }
        float a = 1.0f + q;
        float nx = a * ((((x8 - (28.0f * x6) * y2) + (70.0f * x4) * y4) - (28.0f * x2) * y6) - y8);
        float ny = (((8.0f * a) * x) * y) * (((x6 - (7.0f * x4) * y2) + (7.0f * x2) * y4) - y6);
        float nz = (((8.0f * z) * sw2) * (z2 - w2)) * ((z4 - (6.0f * z2) * w2) + w4);
        // On the z axis (only reachable with the scaled sequence).  The wave
        // test looks at every lane (a compare's lane mask is free; folding
        // `live` in first would cost two VALU instructions); a lane without a
        // live point that happens to sit on the axis only sends the wave
        // through the branch, where it does nothing.
        if (!unscaled) {
            const bool axis = on_z_axis(x, y);
            if (wave_mask(axis)) {
                if (live && axis) {
                    // rotate_on_z_axis::<8> (:114-126)
                    float c = div_rn(z, r);
                    if (__builtin_isinf(c)) pos_nan = true;
                    float th = acosf(c);
                    float r8 = r4 * r4;
                    nx = 0.0f;
                    ny = 0.0f;
                    nz = r8 * cosf(th * 8.0f);
                }
            }
        }
        if (live) {
            dr = ndr;
            x = nx + px;
            y = ny + py;
            z = nz + pz;
        }
    }
    // The dr update of a body, alone: dr = r.powi(7) * 8.0 * dr + 1.0 (:73).
    // When a point's max_iters-th body is its last (it did not bail before),
    // only this part of it is live: the rotated point it computes is never
    // read (:65-78, the loop ends and r is not recomputed).
    __device__ __forceinline__ void last_dr(float r) {
        float r2 = r * r;
        float r4 = r2 * r2;
        float r7 = (r * r2) * r4;
        dr = ((r7 * 8.0f) * dr) + 1.0f;
    }
    // 0.5 * (r.ln() * r) / dr   (:77-78)
    __device__ __forceinline__ float finish(float r) const {
        return finish_exact(r, dr, pos_nan);
    }
};

// ---------------------------------------------------------------------------
// P8 with FMA contraction (CTH_MODE_CONTRACT): the parity evaluator's
// operations, except that the compiler may fuse a product and a following
// sum into one FMA, as a contracting compiler would for the reference; the
// division and the square roots stay correctly rounded and the guards are
// the same.  Not bit-exact: each fused operation skips one rounding, and
// the chaotic tail of non-escaping orbits moves.
// ---------------------------------------------------------------------------
struct P8Contract : P8Exact {
    __device__ __forceinline__ void last_dr(float r) {
#pragma clang fp contract(fast)
        float r2 = r * r;
        float r4 = r2 * r2;
        float r7 = (r * r2) * r4;
        dr = ((r7 * 8.0f) * dr) + 1.0f;
    }
    __device__ __forceinline__ float mag() {
#pragma clang fp contract(fast)
        x2 = x * x;
        y2 = y * y;
        w2 = x2 + y2;
        z2 = z * z;
        return sqrt_rn_wave(w2 + z2);
    }
    // P8Exact::mag_tick with contraction (the same ranges hold).
    __device__ __forceinline__ float mag_tick(const MandelUniform& u) {
#pragma clang fp contract(fast)
        x2 = x * x;
        y2 = y * y;
        w2 = x2 + y2;
        z2 = z * z;
        guard(u);
        const float m = w2 + z2;
        return fast ? sqrt_rn_unscaled(m) : sqrt_rn_wave(m);
    }
    __device__ __forceinline__ void body(float r, const MandelUniform& u, bool live) {
        body_with(r, live, u.bailout <= 16.0f && wave_mask(!(w2 >= 0x1p-15f)) == 0);
    }
    __device__ __forceinline__ void body_guarded(float r, const MandelUniform&, bool live) { body_with(r, live, fast); }
    __device__ __forceinline__ void body_with(float r, bool live, bool unscaled) {
#pragma clang fp contract(fast)
        float r2 = r * r;
        float r4 = r2 * r2;
        float r7 = (r * r2) * r4;
        float ndr = ((r7 * 8.0f) * dr) + 1.0f;
        float x4 = x2 * x2, x6 = x4 * x2, x8 = x4 * x4;
        float y4 = y2 * y2, y6 = y4 * y2, y8 = y4 * y4;
        float z4 = z2 * z2, z6 = z4 * z2, z8 = z4 * z4;
        float w4 = w2 * w2, w6 = w2 * w4, w8 = w4 * w4;
        float num = ((z8 - (28.0f * z6) * w2) + (70.0f * z4) * w4) - (28.0f * z2) * w6;
        // The ranges of P8Exact::body: the guard depends on w2 only.
        float q, sw2;
        if (unscaled) {
            q = div_rn_unscaled(num, w8);
            sw2 = sqrt_rn_unscaled(w2);
        } else {
            q = div_rn(num, w8);
            sw2 = sqrt_rn_wave(w2);
        }
        float a = 1.0f + q;
        float nx = a * ((((x8 - (28.0f * x6) * y2) + (70.0f * x4) * y4) - (28.0f * x2) * y6) - y8);
        float ny = (((8.0f * a) * x) * y) * (((x6 - (7.0f * x4) * y2) + (7.0f * x2) * y4) - y6);
        float nz = (((8.0f * z) * sw2) * (z2 - w2)) * ((z4 - (6.0f * z2) * w2) + w4);
        if (!unscaled) {
            const bool axis = on_z_axis(x, y);
            if (wave_mask(axis)) {
                if (live && axis) {
                    float c = div_rn(z, r);
                    if (__builtin_isinf(c)) pos_nan = true;
                    float th = acosf(c);
                    float r8 = r4 * r4;
                    nx = 0.0f;
                    ny = 0.0f;
                    nz = r8 * cosf(th * 8.0f);
                }
            }
        }
        if (live) {
            dr = ndr;
            x = nx + px;
            y = ny + py;
            z = nz + pz;
        }
    }
};

// ---------------------------------------------------------------------------
// Exact trig evaluator (reference `Mandelbulb<P != 8>`, and the real-power
// extension with powf).  Transcendentals are ocml's.
// ---------------------------------------------------------------------------
struct TrigExact {
    static constexpr int kWavesPerSimd = 1;  // no constraint (64 VGPRs would spill)
    static constexpr bool kLastDr = true;
    float px, py, pz, x, y, z, dr, P;
    int Pi;       // integer power of the powi ladder (when `ladder`)
    bool ladder;  // the powi ladder (integral power) rather than powf
    bool pos_nan;

    // Integral powers use the powi ladder (the reference's const-generic u8
    // power); the real-power kind uses powf unless the point's power happens
    // to be an integer of magnitude below 64, zero and negative ones included
    // (the oracle applies the same rule).
    __device__ __forceinline__ void start(float ax, float ay, float az, float power, const MandelUniform& u) {
        px = x = ax;
        py = y = ay;
        pz = z = az;
        dr = 1.0f;
        P = power;
        pos_nan = input_pos_nan(ax, ay, az);
        if (u.integral) {
            Pi = u.ipower;
            ladder = true;
        } else {
            ladder = floorf(power) == power && fabsf(power) < 64.0f;
            Pi = ladder ? (int)power : 0;
        }
    }
    __device__ __forceinline__ float mag() const { return sqrt_rn_wave((x * x + y * y) + z * z); }

    // One loop body's intermediate values, kept so that the tangent
    // evaluator (tangent.h) differentiates the same quantities without
    // evaluating the transcendentals a second time.
    struct Terms {
        float rp, R;      // r^(P-1), r^P
        float th, ph;     // theta = acos(z / r), phi = atan2(y, x) (before scaling by P)
        float st, ct;     // sin(P theta), cos(P theta)
        float sp, cp;     // sin(P phi), cos(P phi)
        float c;          // z / r
        float nx, ny, nz, ndr;
    };

    __device__ __forceinline__ void eval(float r, Terms& t) const {
        // dr = r.powi(P - 1) * (P as f32) * dr + 1.0   (:73)
        t.rp = ladder ? powi_ladder(r, Pi - 1) : powf(r, P - 1.0f);
        t.ndr = ((t.rp * P) * dr) + 1.0f;
        // rotate_inner_px_generic (:128-146); on the z-axis rotate_on_z_axis
        // (:114-126) keeps only (0, 0, R cos(P theta)), the same z product.
        t.R = ladder ? powi_ladder(r, Pi) : powf(r, P);
        t.c = div_rn(z, r);
        t.th = acosf(t.c);
        t.ph = atan2f(y, x);
        const float th = t.th * P;
        const float ph = t.ph * P;
        // one shared argument reduction per angle; the values are sinf's and
        // cosf's (identical outputs, 2-3% faster on the trig workloads)
        sincosf(th, &t.st, &t.ct);
        sincosf(ph, &t.sp, &t.cp);
        t.nx = (t.st * t.cp) * t.R;
        t.ny = (t.sp * t.st) * t.R;
        t.nz = t.ct * t.R;
        if (on_z_axis(x, y)) {
            t.nx = 0.0f;
            t.ny = 0.0f;
        }
    }

    __device__ __forceinline__ void commit(const Terms& t, bool live) {
        if (live) {
            if (__builtin_isinf(t.c)) pos_nan = true;
            dr = t.ndr;
            x = t.nx + px;
            y = t.ny + py;
            z = t.nz + pz;
        }
    }

    __device__ __forceinline__ void body(float r, const MandelUniform&, bool live) {
        Terms t;
        eval(r, t);
        commit(t, live);
    }
    // The last body's dr update alone (P8Exact::last_dr).
    __device__ __forceinline__ void last_dr(float r) {
        const float rp = ladder ? powi_ladder(r, Pi - 1) : powf(r, P - 1.0f);
        dr = ((rp * P) * dr) + 1.0f;
    }
    __device__ __forceinline__ float finish(float r) const {
        return finish_exact(r, dr, pos_nan);
    }
};

// ---------------------------------------------------------------------------
// Fast P8 evaluator (perf mode).  Same map; FMAs, reciprocal, hardware sqrt.
// ---------------------------------------------------------------------------
struct P8Fast {
    static constexpr int kWavesPerSimd = 8;
    static constexpr bool kLastDr = true;
    float px, py, pz, x, y, z, dr;
    float x2, y2, w2, z2;

    __device__ __forceinline__ void start(float ax, float ay, float az, float, const MandelUniform&) {
        px = x = ax;
        py = y = ay;
        pz = z = az;
        dr = 1.0f;
    }
    __device__ __forceinline__ float mag() {
        x2 = x * x;
        y2 = y * y;
        w2 = x2 + y2;
        z2 = z * z;
        return __builtin_amdgcn_sqrtf(w2 + z2);
    }
    __device__ __forceinline__ void body(float r, const MandelUniform&, bool live) {
        float r2 = r * r;
        float r4 = r2 * r2;
        float r7 = (r * r2) * r4;
        float ndr = __builtin_fmaf(r7 * 8.0f, dr, 1.0f);
        float x4 = x2 * x2, x6 = x4 * x2, x8 = x4 * x4;
        float y4 = y2 * y2, y6 = y4 * y2, y8 = y4 * y4;
        float z4 = z2 * z2, z6 = z4 * z2, z8 = z4 * z4;
        float w4 = w2 * w2, w6 = w2 * w4, w8 = w4 * w4;
        float num = __builtin_fmaf(-28.0f * z2, w6, __builtin_fmaf(70.0f * z4, w4, __builtin_fmaf(-28.0f * z6, w2, z8)));
        float a = __builtin_fmaf(num, __builtin_amdgcn_rcpf(w8), 1.0f);
        float px8 = __builtin_fmaf(-28.0f * x2, y6, __builtin_fmaf(70.0f * x4, y4, __builtin_fmaf(-28.0f * x6, y2, x8))) - y8;
        float py6 = __builtin_fmaf(7.0f * x2, y4, __builtin_fmaf(-7.0f * x4, y2, x6)) - y6;
        float pz4 = __builtin_fmaf(-6.0f * z2, w2, z4) + w4;
        float nx = a * px8;
        float ny = ((8.0f * a) * (x * y)) * py6;
        float nz = ((8.0f * z) * __builtin_amdgcn_sqrtf(w2)) * ((z2 - w2) * pz4);
        if (on_z_axis(x, y)) {
            // z on the axis: cos(8 acos(+-1)) == 1 and z / r == +-1 unless r == 0.
            nx = 0.0f;
            ny = 0.0f;
            nz = (z == 0.0f) ? __uint_as_float(0xffc00000u) : r4 * r4;
        }
        if (live) {
            dr = ndr;
            x = nx + px;
            y = ny + py;
            z = nz + pz;
        }
    }
    __device__ __forceinline__ void last_dr(float r) {
        float r2 = r * r;
        float r4 = r2 * r2;
        float r7 = (r * r2) * r4;
        dr = __builtin_fmaf(r7 * 8.0f, dr, 1.0f);
    }
    __device__ __forceinline__ float finish(float r) const {
        return canon_nan((0.5f * (__logf(r) * r)) * __builtin_amdgcn_rcpf(dr));
    }
};

// ---------------------------------------------------------------------------
// Sphere (src/shape/sphere.rs:33-35): |center - p| - radius.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float sphere_de(float px, float py, float pz, const float* c, float radius) {
    float dx = c[0] - px, dy = c[1] - py, dz = c[2] - pz;
    return canon_nan(sqrt_rn((dx * dx + dy * dy) + dz * dz) - radius, input_pos_nan(px, py, pz));
}

// The sphere as an evaluator of the lane-refill loops: it has no loop body,
// so mag() reports +inf (an immediate "bailout") and finish() returns the
// exact distance.  The centre and radius travel in the uniforms' spare fields.
struct SphereEv {
    static constexpr int kWavesPerSimd = 8;
    static constexpr bool kLastDr = false;
    float px, py, pz, dr;
    float c[3], radius;
    __device__ __forceinline__ void start(float ax, float ay, float az, float, const MandelUniform& u) {
        px = ax;
        py = ay;
        pz = az;
        dr = 1.0f;
        c[0] = u.center[0];
        c[1] = u.center[1];
        c[2] = u.center[2];
        radius = u.radius;
    }
    __device__ __forceinline__ float mag() const { return __builtin_inff(); }
    __device__ __forceinline__ void body(float, const MandelUniform&, bool) {}
    __device__ __forceinline__ float finish(float) const { return sphere_de(px, py, pz, c, radius); }
};

// The tick of a work loop through an evaluator.  Evaluators with
// kCarriedGuard make the next body's wave test in the magnitude after a body
// (mag_tick) and in guard(), which the loops call after starting points,
// instead of in the body (body_guarded); the others keep mag() and body().
template <class E, class = void>
struct CarriesGuard : std::false_type {};
template <class E>
struct CarriesGuard<E, std::void_t<decltype(E::kCarriedGuard)>> : std::bool_constant<E::kCarriedGuard> {};

template <class E>
__device__ __forceinline__ void tick_guard(E& ev, const MandelUniform& u) {
    if constexpr (CarriesGuard<E>::value) ev.guard(u);
}
template <class E>
__device__ __forceinline__ void tick_body(E& ev, float r, const MandelUniform& u, bool live) {
    if constexpr (CarriesGuard<E>::value)
        ev.body_guarded(r, u, live);
    else
        ev.body(r, u, live);
}
template <class E>
__device__ __forceinline__ float tick_mag(E& ev, const MandelUniform& u) {
    if constexpr (CarriesGuard<E>::value)
        return ev.mag_tick(u);
    else
        return ev.mag();
}

// Calls f(Ev{}) with the P8 evaluator of an arithmetic mode.
template <class F>
__host__ __forceinline__ void with_p8(int mode, F&& f) {
    if (mode == CTH_MODE_FAST)
        f(P8Fast{});
    else if (mode == CTH_MODE_CONTRACT)
        f(P8Contract{});
    else
        f(P8Exact{});
}

}  // namespace cth
