// Forward-mode (tangent-linear) evaluators: the parity forward iteration
// plus the Jacobian of the orbit with respect to the point p (and the power
// P of the real-power trig map), carried along each loop body.
//
// For a scalar distance of 3-4 inputs this needs no orbit storage, so it runs
// inside the lane-refill loop unchanged (the reverse sweep of
// kernels_adjoint.hip needs the orbit).  Per body, with T_q = dz/dq:
//   dr_{k+1} = P r^(P-1) dr_k + 1   ->  t_dr' = P(P-1) r^(P-2) dr (dr/dq) + P r^(P-1) t_dr [+ dP term]
//   z_{k+1} = F(z_k) + p            ->  T' = J_F T + [I | 0] [+ dF/dP e_P]
// and at the end
//   dd/dq = 0.5 ((ln r + 1) dr_f (dr_f/dq... r) - ln r r t_dr) / dr^2.
// The forward value is the parity evaluator's; the derivative arithmetic is
// ordinary f32 with FMAs (its accuracy target is angular, not bitwise).
#pragma once

#include "de_device.h"

namespace cth {

// Gradient of d = 0.5 ln(r) r / dr with respect to q, given dr/dq and ddr/dq.
__device__ __forceinline__ float grad_final(float r, float dr, float rq, float drq) {
    float lnr = __logf(r);
    return 0.5f * ((lnr + 1.0f) * rq * dr - lnr * r * drq) / (dr * dr);
}

// ---------------------------------------------------------------------------
// P8 polynomial: Jacobian from the complex form of the map
//   c = x + i y, w = |c|, q = w + i z, a = Re(q^8) / w^8
//   F = (a (Re c^8 - 2 y^8), a Im c^8, -Im q^8)
// (equal to rotate_inner_p8_scalar, src/shape/mandelbulb.rs:148-200).
// ---------------------------------------------------------------------------
struct P8Tangent : P8Exact {
    static constexpr int kWavesPerSimd = 1;  // the tangent state does not fit 64 VGPRs
    static constexpr bool kLastDr = false;   // the work loops run its last body in full
    static constexpr bool kCarriedGuard = false;  // its body() also carries the tangents
    float T[3][3];    // dz/dp
    float tdr[3];     // d dr / dp
    float rq_prev[3]; // dr/dp of the magnitude entering the last body

    __device__ __forceinline__ void start(float ax, float ay, float az, float pw, const MandelUniform& u) {
        P8Exact::start(ax, ay, az, pw, u);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
#pragma unroll
            for (int j = 0; j < 3; ++j) T[i][j] = (i == j) ? 1.0f : 0.0f;
            tdr[i] = 0.0f;
            rq_prev[i] = 0.0f;
        }
    }

    __device__ __forceinline__ void body(float r, const MandelUniform& u, bool live) {
        const float zx = x, zy = y, zz = z;
        // dr/dq = (z . T_q) / r
        float ir = 1.0f / r;
        float rq[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) rq[q] = (zx * T[0][q] + zy * T[1][q] + zz * T[2][q]) * ir;
        float r2 = r * r, r6 = r2 * r2 * r2, r7 = r6 * r;
        float c56 = 56.0f * r6 * dr, c8 = 8.0f * r7;
        float ntdr[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) ntdr[q] = c56 * rq[q] + c8 * tdr[q];
        // Jacobian of F at z
        float w2 = zx * zx + zy * zy;
        float w = __builtin_sqrtf(w2);
        float iw = 1.0f / w;
        // c^7 and c^8
        float c2r = zx * zx - zy * zy, c2i = 2.0f * zx * zy;
        float c4r = c2r * c2r - c2i * c2i, c4i = 2.0f * c2r * c2i;
        float c6r = c4r * c2r - c4i * c2i, c6i = c4r * c2i + c4i * c2r;
        float c7r = c6r * zx - c6i * zy, c7i = c6r * zy + c6i * zx;
        float c8i = c7r * zy + c7i * zx;
        float c8r = c7r * zx - c7i * zy;
        // q^7 and q^8, q = w + i z
        float q2r = w2 - zz * zz, q2i = 2.0f * w * zz;
        float q4r = q2r * q2r - q2i * q2i, q4i = 2.0f * q2r * q2i;
        float q6r = q4r * q2r - q4i * q2i, q6i = q4r * q2i + q4i * q2r;
        float q7r = q6r * w - q6i * zz, q7i = q6r * zz + q6i * w;
        float q8r = q7r * w - q7i * zz;
        float w4 = w2 * w2, iw8 = 1.0f / (w4 * w4);
        float a = q8r * iw8;
        float y2 = zy * zy, y4 = y2 * y2, y7 = y4 * y2 * zy, y8 = y4 * y4;
        float Xp = c8r - 2.0f * y8;
        // da/dx = A x, da/dy = A y, da/dz = -8 Im(q^7) / w^8
        float A = 8.0f * (q7r * w - q8r) * iw8 * iw * iw;
        float daz = -8.0f * q7i * iw8;
        float J[3][3];
        J[0][0] = A * zx * Xp + a * 8.0f * c7r;
        J[0][1] = A * zy * Xp + a * (-8.0f * c7i - 16.0f * y7);
        J[0][2] = daz * Xp;
        J[1][0] = A * zx * c8i + a * 8.0f * c7i;
        J[1][1] = A * zy * c8i + a * 8.0f * c7r;
        J[1][2] = daz * c8i;
        J[2][0] = -8.0f * q7i * zx * iw;
        J[2][1] = -8.0f * q7i * zy * iw;
        J[2][2] = -8.0f * q7r;
        float NT[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int q = 0; q < 3; ++q)
                NT[i][q] = J[i][0] * T[0][q] + J[i][1] * T[1][q] + J[i][2] * T[2][q] + ((i == q) ? 1.0f : 0.0f);
        P8Exact::body(r, u, live);
        if (live) {
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                rq_prev[q] = rq[q];
                tdr[q] = ntdr[q];
#pragma unroll
                for (int i = 0; i < 3; ++i) T[i][q] = NT[i][q];
            }
        }
    }

    // dd/dp; `capped`: r is the magnitude before the last body.
    __device__ __forceinline__ void grad(float r, bool capped, float g[4]) const {
        float rq[3];
        if (capped) {
#pragma unroll
            for (int q = 0; q < 3; ++q) rq[q] = rq_prev[q];
        } else {
            float ir = 1.0f / r;
#pragma unroll
            for (int q = 0; q < 3; ++q) rq[q] = (x * T[0][q] + y * T[1][q] + z * T[2][q]) * ir;
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) g[q] = grad_final(r, dr, rq[q], tdr[q]);
        g[3] = 0.0f;
    }
};

// ---------------------------------------------------------------------------
// Trig map: F = R u(Theta, Phi), R = r^P, Theta = P theta, Phi = P phi,
// u = (sin Theta cos Phi, sin Theta sin Phi, cos Theta),
// theta = acos(z / r), phi = atan2(y, x) (src/shape/mandelbulb.rs:128-146).
// NQ = 4 carries the d/dP column as well (real-power kind); NQ = 3 only the
// point derivatives (normals).
// ---------------------------------------------------------------------------
template <int NQ>
struct TrigTangentN : TrigExact {
    static_assert(NQ == 3 || NQ == 4, "3 point columns, optionally the power");
    static constexpr int kWavesPerSimd = 1;
    static constexpr bool kLastDr = false;  // the work loops run its last body in full
    float T[3][NQ];
    float tdr[NQ];
    float rq_prev[NQ];
    bool dP;  // differentiate with respect to the power

    __device__ __forceinline__ void start(float ax, float ay, float az, float pw, const MandelUniform& u) {
        TrigExact::start(ax, ay, az, pw, u);
        dP = (NQ == 4) && !u.integral;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < NQ; ++j) T[i][j] = (i == j) ? 1.0f : 0.0f;
#pragma unroll
        for (int j = 0; j < NQ; ++j) {
            tdr[j] = 0.0f;
            rq_prev[j] = 0.0f;
        }
    }

    __device__ __forceinline__ void body(float r, const MandelUniform&, bool live) {
        const float zx = x, zy = y, zz = z;
        // The forward body's own terms: R = r^P, r^(P-1), theta, phi and the
        // sines/cosines of P theta, P phi.
        Terms t;
        TrigExact::eval(r, t);
        float ir = 1.0f / r;
        float rho2 = zx * zx + zy * zy;
        float rho = __builtin_sqrtf(rho2);
        const float R = t.R, Rm1 = t.rp;
        const float sT = t.st, cT = t.ct, sP = t.sp, cP = t.cp;
        // gradients of r, theta, phi with respect to z
        float gr[3] = {zx * ir, zy * ir, zz * ir};
        float kth = zz * ir * ir / rho;
        float gth[3] = {kth * zx, kth * zy, -rho * ir * ir};
        float irho2 = 1.0f / rho2;
        float gph[3] = {-zy * irho2, zx * irho2, 0.0f};
        float u0[3] = {sT * cP, sT * sP, cT};
        float uT[3] = {cT * cP, cT * sP, -sT};
        float uF[3] = {-sT * sP, sT * cP, 0.0f};
        float PR = P * R;
        float cr = P * Rm1;
        float ntdr[NQ], NT[3][NQ], rq[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            float rq_ = gr[0] * T[0][q] + gr[1] * T[1][q] + gr[2] * T[2][q];
            float tq = gth[0] * T[0][q] + gth[1] * T[1][q] + gth[2] * T[2][q];
            float pq = gph[0] * T[0][q] + gph[1] * T[1][q];
            rq[q] = rq_;
            // dr_{k+1} = P r^(P-1) dr + 1
            ntdr[q] = P * (P - 1.0f) * Rm1 * ir * dr * rq_ + cr * tdr[q];
#pragma unroll
            for (int i = 0; i < 3; ++i) NT[i][q] = cr * u0[i] * rq_ + PR * (uT[i] * tq + uF[i] * pq);
        }
        if constexpr (NQ == 4) {
            if (dP) {
                // explicit power dependence of F and of the dr update
                const float lnr = __logf(r);
#pragma unroll
                for (int i = 0; i < 3; ++i) NT[i][NQ - 1] += R * (lnr * u0[i] + t.th * uT[i] + t.ph * uF[i]);
                ntdr[NQ - 1] += Rm1 * dr * (1.0f + P * lnr);
            }
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) NT[i][i] += 1.0f;  // z_{k+1} = F + p
        TrigExact::commit(t, live);
        if (live) {
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                rq_prev[q] = rq[q];
                tdr[q] = ntdr[q];
#pragma unroll
                for (int i = 0; i < 3; ++i) T[i][q] = NT[i][q];
            }
        }
    }

    // dd/dp (and dd/dP in g[3] when NQ == 4 and the power is real).
    __device__ __forceinline__ void grad(float r, bool capped, float g[4]) const {
        float rq[NQ];
        if (capped) {
#pragma unroll
            for (int q = 0; q < NQ; ++q) rq[q] = rq_prev[q];
        } else {
            float ir = 1.0f / r;
#pragma unroll
            for (int q = 0; q < NQ; ++q) rq[q] = (x * T[0][q] + y * T[1][q] + z * T[2][q]) * ir;
        }
#pragma unroll
        for (int q = 0; q < NQ; ++q) g[q] = grad_final(r, dr, rq[q], tdr[q]);
        if (NQ == 3 || !dP) g[3] = 0.0f;
    }
};

using TrigTangent = TrigTangentN<4>;  // gradients with respect to p and the power
using TrigNormal = TrigTangentN<3>;   // gradients with respect to p only

}  // namespace cth
