// Naive surface nets over a batch of sampled octree leaves: steps 2 and 3 of
// MeshBuffer::naive_surface_nets (src/mesh/buffer.rs:97-372), given the
// step-1 corner distances that cth_sample_spans_dev wrote.
//
// Two passes over the leaves, one workgroup per leaf:
//   count: per leaf, the number of crossing cells (one vertex each) and of
//          emitted quads (6 indices each); an exclusive scan turns the counts
//          into the leaves' offsets in the concatenated outputs.
//   emit:  per leaf, the crossing mask of every cell (one ballot per 64
//          cells, kept in LDS with its running popcount), so that the vertex
//          index of any cell -- `points[(x, y, z)]` of the reference -- is a
//          popcount; then the cells in cube() order, tile by tile: vertex
//          positions (edge crossings in EDGES order, inverted weights,
//          centroid) and the quads of the +x/+y/+z edges of each corner.
// The vertex metadata -- DE at the vertex and its normal -- are distance
// evaluations again, so they run in the lane-refill loop afterwards: either
// the reference's 6-tap central difference (7 evaluations per vertex,
// :254-266) or the forward-mode gradient (one tangent evaluation).
//
// Vertices use the reference's `Vertex` layout (src/mesh/mod.rs:255-261):
// position[3], normal[3], distance_from_surface, 28 bytes.  Indices are u32
// local to the leaf, exactly as each leaf's MeshBuffer holds them.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <type_traits>

#include "internal.h"
#include "launch.h"
#include "point_sources.h"
#include "tangent.h"
#include "work_loop.h"

namespace cth {

namespace {

constexpr uint32_t kMeshBlock = 256;

__device__ __forceinline__ uint32_t sign_bit(float d) { return __float_as_uint(d) >> 31; }

// The eight corner distances of cell (x, y, z) in the reference's corner
// order (src/mesh/buffer.rs:116-125) and their IEEE sign bits (bit k =
// corner k is negative, `!is_sign_positive`).
__device__ __forceinline__ uint32_t load_cell(const float* __restrict__ D, uint32_t n1, uint32_t x, uint32_t y,
                                              uint32_t z, float d[8]) {
    const size_t sy = n1, sx = (size_t)n1 * n1;
    const size_t b = (size_t)x * sx + (size_t)y * sy + z;
    d[0] = D[b];
    d[1] = D[b + 1];
    d[2] = D[b + sy];
    d[3] = D[b + sy + 1];
    d[4] = D[b + sx];
    d[5] = D[b + sx + 1];
    d[6] = D[b + sx + sy];
    d[7] = D[b + sx + sy + 1];
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) m |= sign_bit(d[k]) << k;
    return m;
}

__device__ __forceinline__ bool crossing(uint32_t m) { return m != 0u && m != 0xFFu; }

// Quads of corner (x, y, z) (src/mesh/buffer.rs:299-363): bit 0 the +x edge
// (y, z > 0), bit 1 +y (x, z > 0), bit 2 +z (x, y > 0); an edge emits when
// the signs of its two corners differ (corners 0-4, 0-2, 0-1 of the cell).
__device__ __forceinline__ uint32_t face_bits(uint32_t m, uint32_t x, uint32_t y, uint32_t z) {
    uint32_t b0 = m & 1u;
    uint32_t fx = (y > 0 && z > 0 && b0 != ((m >> 4) & 1u)) ? 1u : 0u;
    uint32_t fy = (x > 0 && z > 0 && b0 != ((m >> 2) & 1u)) ? 2u : 0u;
    uint32_t fz = (x > 0 && y > 0 && b0 != ((m >> 1) & 1u)) ? 4u : 0u;
    return fx | fy | fz;
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const uint32_t lane = __lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(v, d, 64);
        if (lane >= (uint32_t)d) v += t;
    }
    return v;
}

// Exclusive scan over the workgroup (all threads must call it); `total`
// receives the sum.  `wsum` holds one slot per wave.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* wsum, uint32_t& total) {
    const uint32_t inc = wave_incl_scan(v);
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    if (lane == 63u) wsum[wave] = inc;
    __syncthreads();
    uint32_t off = 0, tot = 0;
    const uint32_t nw = blockDim.x >> 6;
    for (uint32_t w = 0; w < nw; ++w) {
        uint32_t s = wsum[w];
        off += (w < wave) ? s : 0u;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return off + inc - v;
}

// The wave's LDS writes become visible to its other lanes.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, uint32_t& total) {
    const uint32_t inc = wave_incl_scan(v);
    total = (uint32_t)__shfl((int)inc, 63, 64);
    return inc - v;
}

__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t* wsum) {
    uint32_t t;
    (void)block_excl_scan(v, wsum, t);
    return t;
}

struct MeshDims {
    uint32_t res, lg, n1, cells;
    float fres;
};

__device__ __forceinline__ void cell_xyz(const MeshDims& m, uint32_t c, uint32_t& x, uint32_t& y, uint32_t& z) {
    z = c & (m.res - 1u);
    y = (c >> m.lg) & (m.res - 1u);
    x = c >> (2u * m.lg);
}

// ---------------------------------------------------------------------------
// Plane units for the large resolutions (res 32 and 64, the reference's own
// RESOLUTION): unit u = leaf * res + x holds the res^2 cells (x, *, *) of a
// leaf, so a single leaf still spreads over res workgroups.  The units are
// counted and scanned in cube() order; a unit's vertices then start at its
// scanned offset, and cells of plane x - 1 (the only other plane its quads
// touch) are numbered from the previous unit's offset.
// ---------------------------------------------------------------------------

// Per unit: vertex count and index count (6 per quad); slot `units` of both
// arrays is zeroed so that an exclusive scan over units + 1 entries ends
// with the totals.  With `cross` non-null it also writes the leaf's
// crossing words (bit c of word c / 64 = cell c crosses), the layout
// cth_sample_spans_dev uses (res^2 is a multiple of 256, so every wave's
// 64 cells are one word).
__global__ __launch_bounds__(kMeshBlock) void k_mesh_count_p(const float* __restrict__ dists, MeshDims m,
                                                             uint32_t units, uint64_t* __restrict__ vcnt,
                                                             uint64_t* __restrict__ icnt, uint64_t* __restrict__ cross) {
    __shared__ uint32_t wsum[kMeshBlock / 64];
    const uint32_t u = blockIdx.x;
    const uint32_t leaf = u >> m.lg, x = u & (m.res - 1u);
    const float* D = dists + (size_t)leaf * m.n1 * m.n1 * m.n1;
    const uint32_t pc = m.res * m.res;
    const uint32_t nwords = m.cells >> 6;
    uint32_t nv = 0, nf = 0;
    for (uint32_t j = threadIdx.x; j < pc; j += blockDim.x) {
        const uint32_t y = j >> m.lg, z = j & (m.res - 1u);
        float d[8];
        const uint32_t mk = load_cell(D, m.n1, x, y, z, d);
        const bool cr = crossing(mk);
        nv += cr ? 1u : 0u;
        nf += (uint32_t)__popc(face_bits(mk, x, y, z));
        const uint64_t bits = wave_mask(cr);
        if (cross && (threadIdx.x & 63u) == 0) cross[(size_t)leaf * nwords + ((x * pc + j) >> 6)] = bits;
    }
    nv = block_sum(nv, wsum);
    nf = block_sum(nf, wsum);
    if (threadIdx.x == 0) {
        vcnt[u] = nv;
        icnt[u] = 6ull * nf;
        if (u == 0) {
            vcnt[units] = 0;
            icnt[units] = 0;
        }
    }
}

// Per-leaf offsets from the scanned unit offsets: leaf i starts at unit i * res.
__global__ void k_leaf_offsets(const uint64_t* __restrict__ uv, const uint64_t* __restrict__ ui, uint32_t nleaves,
                               uint32_t res, uint64_t* __restrict__ voff, uint64_t* __restrict__ ioff) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > nleaves) return;
    voff[i] = uv[(size_t)i * res];
    ioff[i] = ui[(size_t)i * res];
}

// Vertex position of a crossing cell (src/mesh/buffer.rs:149-250), in the
// reference's f32 operation order.  s1: adjusted span start; step: cell size.
__device__ __forceinline__ void place_vertex(const float d[8], uint32_t x, uint32_t y, uint32_t z, const float s1[3],
                                             const float step[3], float p[3]) {
    const float p0[3] = {s1[0] + (float)x * step[0], s1[1] + (float)y * step[1], s1[2] + (float)z * step[2]};
    // EDGES (src/mesh/buffer.rs:155-176), encoded as from | to << 4.
    const uint8_t edges[12] = {0x40, 0x51, 0x62, 0x73, 0x20, 0x31, 0x64, 0x75, 0x10, 0x32, 0x54, 0x76};
    float sum[3] = {0.0f, 0.0f, 0.0f};
    uint32_t count = 0;
#pragma unroll
    for (int ei = 0; ei < 12; ++ei) {
        const uint32_t from = edges[ei] & 7u, to = edges[ei] >> 4;
        if (sign_bit(d[from]) == sign_bit(d[to])) continue;
        float d_from, d_to;
        if (d[from] < 0.0f) {
            d_from = d[from];
            d_to = d[to];
        } else {
            d_from = -d[from];
            d_to = -d[to];
        }
        float w;
        if (d_to == d_from) {
            w = 0.5f;
        } else {
            float dl = d_to - d_from;
            w = div_rn(d_from + dl, dl);
        }
        // lerp(p0 + off[from], p0 + off[to], w) = a * (1 - w) + b * w (src/math.rs:45-49);
        // corner k has +x for bit 2, +y for bit 1, +z for bit 0.
        const float om = 1.0f - w;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const uint32_t bit = 2u - (uint32_t)a;
            float pa = p0[a] + (((from >> bit) & 1u) ? step[a] : 0.0f);
            float pb = p0[a] + (((to >> bit) & 1u) ? step[a] : 0.0f);
            sum[a] = sum[a] + (pa * om + pb * w);
        }
        ++count;
    }
    const float fc = (float)count;
    // Point3::origin() + sum / count
#pragma unroll
    for (int a = 0; a < 3; ++a) p[a] = 0.0f + div_rn(sum[a], fc);
}

// Adjusted span (src/mesh/buffer.rs:64-67), cell step (:101) and the 6-tap
// half-step 0.7 * (e' - s') / res (:257) of a leaf.
__device__ __forceinline__ void leaf_frame(const float* sp, float fres, float s1[3], float step[3], float delta[3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float s0 = sp[a], e0 = sp[3 + a];
        const float of = div_rn(e0 - s0, fres);
        const float sa = s0 + -of;
        const float ea = e0 + of;
        s1[a] = sa;
        step[a] = div_rn(ea - sa, fres);
        delta[a] = div_rn(0.7f * (ea - sa), fres);
    }
}

__device__ __forceinline__ void put_vertex(float* v, const float p[3], const float delta[3]) {
    v[0] = p[0];
    v[1] = p[1];
    v[2] = p[2];
    v[3] = delta[0];
    v[4] = delta[1];
    v[5] = delta[2];
    v[6] = 0.0f;
}

// The quads of corner (x, y, z) (src/mesh/buffer.rs:299-363), in the
// reference's order (+x, +y, +z edge) and winding (`dists[(x,y,z)] < 0.0`
// picks the clockwise pair).  pt(x, y, z) is the cell's vertex number.
template <class PT>
__device__ __forceinline__ void emit_quads(uint32_t* out, uint32_t fb, bool neg, uint32_t x, uint32_t y, uint32_t z,
                                           const PT& pt) {
    const uint32_t v3 = pt(x, y, z);
    if (fb & 1u) {  // +x edge
        const uint32_t v0 = pt(x, y - 1, z - 1), v1 = pt(x, y - 1, z), v2 = pt(x, y, z - 1);
        const uint32_t q[6] = {v0, v2, v1, v1, v2, v3}, r[6] = {v0, v1, v2, v1, v3, v2};
#pragma unroll
        for (int k = 0; k < 6; ++k) out[k] = neg ? q[k] : r[k];
        out += 6;
    }
    if (fb & 2u) {  // +y edge
        const uint32_t v0 = pt(x - 1, y, z - 1), v1 = pt(x - 1, y, z), v2 = pt(x, y, z - 1);
        const uint32_t q[6] = {v0, v1, v2, v1, v3, v2}, r[6] = {v0, v2, v1, v1, v2, v3};
#pragma unroll
        for (int k = 0; k < 6; ++k) out[k] = neg ? q[k] : r[k];
        out += 6;
    }
    if (fb & 4u) {  // +z edge
        const uint32_t v0 = pt(x - 1, y - 1, z), v1 = pt(x - 1, y, z), v2 = pt(x, y - 1, z);
        const uint32_t q[6] = {v0, v2, v1, v1, v2, v3}, r[6] = {v0, v1, v2, v1, v3, v2};
#pragma unroll
        for (int k = 0; k < 6; ++k) out[k] = neg ? q[k] : r[k];
    }
}

// One workgroup per plane unit.  LDS: the crossing words of planes x and
// x - 1 and their exclusive popcount prefixes (res^2 / 64 words each).  The
// vertex records temporarily hold the leaf's 6-tap half-steps in their normal
// slots (src/mesh/buffer.rs:257: 0.7 * (e - s) / res per axis).
__global__ __launch_bounds__(kMeshBlock) void k_mesh_emit_p(
    const float* __restrict__ spans, const float* __restrict__ dists, MeshDims m, const uint64_t* __restrict__ uv,
    const uint64_t* __restrict__ ui, const uint64_t* __restrict__ voff, const uint64_t* __restrict__ ioff,
    float* __restrict__ verts, uint32_t* __restrict__ idx) {
    extern __shared__ uint64_t lds_words[];
    __shared__ uint32_t wsum[kMeshBlock / 64];
    const uint32_t u = blockIdx.x;
    const uint32_t leaf = u >> m.lg, x = u & (m.res - 1u);
    const uint32_t pc = m.res * m.res, pw = pc >> 6;
    uint64_t* wc = lds_words;   // plane x
    uint64_t* wp = wc + pw;     // plane x - 1
    uint32_t* prc = reinterpret_cast<uint32_t*>(wp + pw);
    uint32_t* prp = prc + pw;
    const float* D = dists + (size_t)leaf * m.n1 * m.n1 * m.n1;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;

    const uint32_t nplanes = (x > 0) ? 2u : 1u;
    for (uint32_t w = wave; w < nplanes * pw; w += nwaves) {
        const uint32_t px = (w < pw) ? x : x - 1u, ww = (w < pw) ? w : w - pw;
        const uint32_t j = (ww << 6) | lane;
        float d[8];
        const bool cr = crossing(load_cell(D, m.n1, px, j >> m.lg, j & (m.res - 1u), d));
        const uint64_t bits = wave_mask(cr);
        if (lane == 0) ((w < pw) ? wc : wp)[ww] = bits;
    }
    __syncthreads();
    if (wave < nplanes) {  // pw <= 64: one wave scans one plane's words
        const uint64_t* words = (wave == 0) ? wc : wp;
        uint32_t* pre = (wave == 0) ? prc : prp;
        const uint32_t v = (lane < pw) ? (uint32_t)__popcll(words[lane]) : 0u;
        uint32_t tot;
        const uint32_t ex = wave_excl_scan(v, tot);
        if (lane < pw) pre[lane] = ex;
    }
    __syncthreads();

    const size_t u0 = (size_t)leaf * m.res;  // the leaf's first unit
    const uint32_t vb_x = (uint32_t)(uv[u] - uv[u0]);
    const uint32_t vb_p = (x > 0) ? (uint32_t)(uv[u - 1] - uv[u0]) : 0u;
    auto pt = [&](uint32_t px, uint32_t py, uint32_t pz) {
        const bool cur = (px == x);
        const uint32_t j = (py << m.lg) | pz, w = j >> 6, b = j & 63u;
        const uint64_t word = cur ? wc[w] : wp[w];
        const uint64_t below = (b == 0) ? 0ull : (word & ((~0ull) >> (64u - b)));
        return (cur ? vb_x + prc[w] : vb_p + prp[w]) + (uint32_t)__popcll(below);
    };
    float s1[3], step[3], delta[3];
    leaf_frame(spans + 6 * (size_t)leaf, m.fres, s1, step, delta);
    float* vout = verts + 7 * voff[leaf];
    uint32_t* iout = idx + ioff[leaf] + (ui[u] - ui[u0]);

    uint32_t fcarry = 0;
    for (uint32_t base = 0; base < pc; base += blockDim.x) {
        const uint32_t j = base + threadIdx.x;  // pc is a multiple of the block
        const uint32_t y = j >> m.lg, z = j & (m.res - 1u);
        float d[8];
        const uint32_t mk = load_cell(D, m.n1, x, y, z, d);
        const uint32_t fb = face_bits(mk, x, y, z);
        uint32_t tile_faces;
        const uint32_t frank = block_excl_scan((uint32_t)__popc(fb), wsum, tile_faces);
        if (crossing(mk)) {
            float p[3];
            place_vertex(d, x, y, z, s1, step, p);
            put_vertex(vout + 7 * (size_t)pt(x, y, z), p, delta);
        }
        if (fb) emit_quads(iout + 6ull * (fcarry + frank), fb, d[0] < 0.0f, x, y, z, pt);
        fcarry += tile_faces;
    }
}

// ---------------------------------------------------------------------------
// Wave-per-leaf kernels for res <= 16 (the octree workload meshes 8^3 cells
// per leaf): one wave owns a leaf, each lane a run of whole cell rows
// (x, y, z = 0..res-1), which are contiguous in cube() order.  A lane
// streams the four corner columns of its row along z (each a contiguous run
// of res + 1 floats), so every corner load serves two cells, and no
// workgroup barrier is needed: the leaf's row bookkeeping lives in the
// wave's own slice of LDS.  The grid is persistent over leaves.
// ---------------------------------------------------------------------------


// Calls f(z, d[8], mask) for the cells (x, y, 0..RES-1) of the leaf's
// corner grid D ((RES+1)^3, GridTable order).
template <int RES, class F>
__device__ __forceinline__ void row_cells(const float* __restrict__ D, uint32_t x, uint32_t y, F&& f) {
    constexpr uint32_t N1 = RES + 1;
    const float* c00 = D + ((size_t)x * N1 + y) * N1;
    const float* c01 = c00 + N1;
    const float* c10 = c00 + N1 * N1;
    const float* c11 = c10 + N1;
    float p0 = c00[0], p1 = c01[0], p2 = c10[0], p3 = c11[0];
#pragma unroll 8
    for (uint32_t z = 0; z < (uint32_t)RES; ++z) {
        const float q0 = c00[z + 1], q1 = c01[z + 1], q2 = c10[z + 1], q3 = c11[z + 1];
        const float d[8] = {p0, q0, p1, q1, p2, q2, p3, q3};
        uint32_t mk = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) mk |= sign_bit(d[k]) << k;
        f(z, d, mk);
        p0 = q0;
        p1 = q1;
        p2 = q2;
        p3 = q3;
    }
}

template <int RES>
struct WaveLeaf {
    static constexpr uint32_t N1 = RES + 1;
    static constexpr uint32_t ROWS = RES * RES;
    static constexpr uint32_t RPL = (ROWS + 63) / 64;  // rows per lane
    static constexpr uint32_t CELLS = RES * RES * RES;
    static constexpr uint32_t NWORDS = (CELLS + 63) / 64;
    static constexpr uint32_t RPW = 64 / RES;  // rows per crossing word
    static constexpr size_t CORNERS = (size_t)N1 * N1 * N1;
    // Resolutions up to 8 stage the leaf's corner distances in the wave's LDS
    // (729 floats at res 8): one round of coalesced loads per leaf, after
    // which the row pass and the vertex pass read LDS instead of re-reading
    // global memory with strided and dependent loads.
    static constexpr bool STAGE = RES <= 8;
    static constexpr uint32_t STAGED = STAGE ? (uint32_t)((CORNERS + 63) / 64 * 64) : 1;
};

// Bit fields of a leaf's corner grid (res <= 8), built with one wave ballot
// per 64 corners: bit i of `sgn` = IEEE sign bit of corner i (the
// reference's !is_sign_positive), bit i of `neg` = corner i < 0.0.  Words
// are 32-bit with two zero words of padding, so any 32-bit window can be read
// with a funnel shift.  With these, one lane classifies its whole row of
// cells with a handful of bit operations instead of per-cell compares.
template <int RES>
struct LeafFields {
    static constexpr uint32_t WORDS = WaveLeaf<RES>::STAGED / 32 + 2;
    uint32_t sgn[WORDS];
    uint32_t neg[WORDS];
};

// A leaf's corner distances, one coalesced round of loads (lane-strided).
template <int RES>
struct LeafRegs {
    static constexpr uint32_t NJ = WaveLeaf<RES>::STAGED / 64;
    float v[NJ];
    __device__ __forceinline__ void load(const float* __restrict__ D, uint32_t lane) {
#pragma unroll
        for (uint32_t j = 0; j < NJ; ++j) {
            const uint32_t i = j * 64 + lane;
            v[j] = i < WaveLeaf<RES>::CORNERS ? D[i] : 0.0f;
        }
    }
};

// Optionally stages the floats in S, and writes the sign fields (the `< 0`
// field only when NEG).
template <int RES, bool NEG>
__device__ __forceinline__ void leaf_fields(const LeafRegs<RES>& R, float* S, LeafFields<RES>& F, uint32_t lane) {
    using L = WaveLeaf<RES>;
    constexpr uint32_t NJ = L::STAGED / 64;
    const float* v = R.v;
    if (S) {
#pragma unroll
        for (uint32_t j = 0; j < NJ; ++j) S[j * 64 + lane] = v[j];
    }
#pragma unroll
    for (uint32_t j = 0; j < NJ; ++j) {
        const uint64_t ms = wave_mask((int)__float_as_uint(v[j]) < 0);
        // lane 0 stores the low word, lane 1 the high word
        if (lane < 2) F.sgn[2 * j + lane] = (uint32_t)(ms >> (32 * lane));
        if constexpr (NEG) {
            const uint64_t mn = wave_mask(v[j] < 0.0f);
            if (lane < 2) F.neg[2 * j + lane] = (uint32_t)(mn >> (32 * lane));
        }
    }
    if (lane < 2) {
        F.sgn[2 * NJ + lane] = 0u;
        if constexpr (NEG) F.neg[2 * NJ + lane] = 0u;
    }
    wave_sync();
}

// Bits [o, o + 32) of a field.
__device__ __forceinline__ uint32_t field_window(const uint32_t* f, uint32_t o) {
    const uint32_t w = o >> 5;
    return __builtin_amdgcn_alignbit(f[w + 1], f[w], o & 31u);
}

// Row (x, y) of a leaf from its sign field: crossing bits, and the +x / +y / +z
// quad bits of its cells (src/mesh/buffer.rs:113-147, 299-350).  A, B, C, D are
// the signs of the corner rows (x, y), (x, y + 1), (x + 1, y), (x + 1, y + 1);
// bit z of `a & a >> 1` is "corners z and z + 1 both negative".
template <int RES>
struct RowBits {
    uint32_t cross, fx, fy, fz;
    __device__ __forceinline__ RowBits(const uint32_t* sgn, uint32_t x, uint32_t y) {
        constexpr uint32_t N1 = RES + 1;
        constexpr uint32_t cm = (1u << RES) - 1u;
        const uint32_t o = (x * N1 + y) * N1;
        const uint32_t a = field_window(sgn, o), b = field_window(sgn, o + N1);
        const uint32_t c = field_window(sgn, o + N1 * N1), d = field_window(sgn, o + N1 * N1 + N1);
        const uint32_t all = a & (a >> 1) & b & (b >> 1) & c & (c >> 1) & d & (d >> 1);
        const uint32_t any = a | (a >> 1) | b | (b >> 1) | c | (c >> 1) | d | (d >> 1);
        cross = any & ~all & cm;                      // not all 8 signs equal
        const uint32_t zpos = cm & ~1u;               // z > 0
        fx = (y > 0) ? ((a ^ c) & zpos) : 0u;         // corner 0 vs 4, needs y, z > 0
        fy = (x > 0) ? ((a ^ b) & zpos) : 0u;         // corner 0 vs 2, needs x, z > 0
        fz = (x > 0 && y > 0) ? ((a ^ (a >> 1)) & cm) : 0u;  // corner 0 vs 1, needs x, y > 0
    }
};

template <int RES>
__global__ __launch_bounds__(kMeshBlock) void k_mesh_count_w(const float* __restrict__ dists, uint32_t nleaves,
                                                             uint64_t* __restrict__ vcnt, uint64_t* __restrict__ icnt,
                                                             uint64_t* __restrict__ cross) {
    using L = WaveLeaf<RES>;
    __shared__ uint32_t rowbits_all[kMeshBlock / 64][L::ROWS];
    __shared__ LeafFields<RES> fields_all[kMeshBlock / 64];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t* rowbits = rowbits_all[wave];
    LeafFields<RES>& F = fields_all[wave];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        vcnt[nleaves] = 0;
        icnt[nleaves] = 0;
    }
    const uint32_t wstride = gridDim.x * (blockDim.x >> 6);
    // The next leaf's corners are loaded while this one is classified.
    LeafRegs<RES> next;
    const uint32_t first = blockIdx.x * (blockDim.x >> 6) + wave;
    if constexpr (L::STAGE) {
        if (first < nleaves) next.load(dists + (size_t)first * L::CORNERS, lane);
    }
    for (uint32_t leaf = first; leaf < nleaves; leaf += wstride) {
        const float* D = dists + (size_t)leaf * L::CORNERS;
        if constexpr (L::STAGE) {
            const LeafRegs<RES> cur = next;
            if (leaf + wstride < nleaves) next.load(dists + (size_t)(leaf + wstride) * L::CORNERS, lane);
            leaf_fields<RES, false>(cur, nullptr, F, lane);
        }
        uint32_t nv = 0, nf = 0;
        for (uint32_t k = 0; k < L::RPL; ++k) {
            const uint32_t row = lane * L::RPL + k;
            if (row < L::ROWS) {
                const uint32_t x = row / RES, y = row % RES;
                uint32_t bits = 0;
                if constexpr (L::STAGE) {
                    const RowBits<RES> rb(F.sgn, x, y);
                    bits = rb.cross;
                    nf += (uint32_t)(__popc(rb.fx) + __popc(rb.fy) + __popc(rb.fz));
                } else {
                    row_cells<RES>(D, x, y, [&](uint32_t z, const float*, uint32_t mk) {
                        const bool cr = crossing(mk);
                        bits |= (cr ? 1u : 0u) << z;
                        nf += (uint32_t)__popc(face_bits(mk, x, y, z));
                    });
                }
                nv += (uint32_t)__popc(bits);
                rowbits[row] = bits;
            }
        }
        uint32_t tv, tf;
        (void)wave_excl_scan(nv, tv);
        (void)wave_excl_scan(nf, tf);
        if (lane == 0) {
            vcnt[leaf] = tv;
            icnt[leaf] = 6ull * tf;
        }
        if (cross) {
            wave_sync();
            for (uint32_t w = lane; w < L::NWORDS; w += 64) {
                uint64_t word = 0;
                for (uint32_t j = 0; j < L::RPW; ++j) {
                    const uint32_t r = w * L::RPW + j;
                    if (r < L::ROWS) word |= (uint64_t)rowbits[r] << (j * RES);
                }
                cross[(size_t)leaf * L::NWORDS + w] = word;
            }
        }
        wave_sync();  // rowbits is rewritten for the next leaf
    }
}

template <int RES>
__global__ __launch_bounds__(kMeshBlock) void k_mesh_emit_w(const float* __restrict__ spans,
                                                            const float* __restrict__ dists, uint32_t nleaves,
                                                            const uint64_t* __restrict__ voff,
                                                            const uint64_t* __restrict__ ioff,
                                                            float* __restrict__ verts, uint32_t* __restrict__ idx) {
    using L = WaveLeaf<RES>;
    static_assert(RES <= 16, "cell numbers and face bits are packed into 16-bit LDS lists");
    using RowFaces = std::conditional_t<(3 * RES <= 32), uint32_t, uint64_t>;
    // Per wave: row bookkeeping and the leaf's compacted vertex and quad lists.
    struct Lists {
        uint32_t rb[L::ROWS];     // crossing bits of the row's cells
        uint32_t rp[L::ROWS];     // vertices before the row
        uint32_t rn[L::ROWS];     // bit z: dists[(x, y, z)] < 0.0
        // face bits of cell z: at bits 3z..3z+2, or (res <= 8) +x at bit z,
        // +y at bit 8 + z, +z at bit 16 + z
        RowFaces rf[L::ROWS];
        uint16_t vlist[L::CELLS];  // cell of each vertex, in order
        uint16_t flist[L::CELLS];  // cell | face bits << 12, cells with faces in order
        float staged[L::STAGED];   // the leaf's corner distances (res <= 8)
        LeafFields<RES> fields;    // their sign fields (res <= 8)
    };
    __shared__ Lists lists_all[kMeshBlock / 64];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    Lists& W = lists_all[wave];
    const uint32_t wstride = gridDim.x * (blockDim.x >> 6);
    // The next leaf's corners are loaded while this one is processed.
    LeafRegs<RES> next;
    const uint32_t first = blockIdx.x * (blockDim.x >> 6) + wave;
    if constexpr (L::STAGE) {
        if (first < nleaves) next.load(dists + (size_t)first * L::CORNERS, lane);
    }
    for (uint32_t leaf = first; leaf < nleaves; leaf += wstride) {
        const float* D = dists + (size_t)leaf * L::CORNERS;
        if constexpr (L::STAGE) {
            const LeafRegs<RES> cur = next;
            if (leaf + wstride < nleaves) next.load(dists + (size_t)(leaf + wstride) * L::CORNERS, lane);
            leaf_fields<RES, true>(cur, W.staged, W.fields, lane);
            D = W.staged;
        }
        // Pass over the lane's rows: crossing, sign and face bits.
        uint32_t lv = 0, lf = 0, lc = 0;
        for (uint32_t k = 0; k < L::RPL; ++k) {
            const uint32_t row = lane * L::RPL + k;
            if (row < L::ROWS) {
                const uint32_t x = row / RES, y = row % RES;
                uint32_t bits = 0, neg = 0;
                RowFaces fw = 0;
                if constexpr (L::STAGE) {
                    const RowBits<RES> rb(W.fields.sgn, x, y);
                    bits = rb.cross;
                    neg = field_window(W.fields.neg, (x * L::N1 + y) * L::N1) & ((1u << RES) - 1u);
                    fw = rb.fx | (rb.fy << 8) | (rb.fz << 16);
                    lf += (uint32_t)(__popc(rb.fx) + __popc(rb.fy) + __popc(rb.fz));
                    lc += (uint32_t)__popc(rb.fx | rb.fy | rb.fz);
                } else {
                    row_cells<RES>(D, x, y, [&](uint32_t z, const float* d, uint32_t mk) {
                        bits |= (crossing(mk) ? 1u : 0u) << z;
                        neg |= (d[0] < 0.0f ? 1u : 0u) << z;
                        const uint32_t fb = face_bits(mk, x, y, z);
                        fw |= (RowFaces)fb << (3u * z);
                        lf += (uint32_t)__popc(fb);
                        lc += fb ? 1u : 0u;
                    });
                }
                W.rb[row] = bits;
                W.rn[row] = neg;
                W.rf[row] = fw;
                lv += (uint32_t)__popc(bits);
            }
        }
        uint32_t tv, tf, tc;
        uint32_t vex = wave_excl_scan(lv, tv);
        (void)wave_excl_scan(lf, tf);
        uint32_t cex = wave_excl_scan(lc, tc);
        // Compacted lists, in cube() cell order.
        for (uint32_t k = 0; k < L::RPL; ++k) {
            const uint32_t row = lane * L::RPL + k;
            if (row < L::ROWS) {
                W.rp[row] = vex;
                for (uint32_t bits = W.rb[row]; bits; bits &= bits - 1u)
                    W.vlist[vex++] = (uint16_t)(row * RES + (uint32_t)__builtin_ctz(bits));
                const RowFaces fw = W.rf[row];
                for (uint32_t z = 0; z < (uint32_t)RES; ++z) {
                    uint32_t fb;
                    if constexpr (L::STAGE)
                        fb = ((uint32_t)(fw >> z) & 1u) | ((uint32_t)(fw >> (z + 7u)) & 2u) |
                             ((uint32_t)(fw >> (z + 14u)) & 4u);
                    else
                        fb = (uint32_t)(fw >> (3u * z)) & 7u;
                    if (fb) W.flist[cex++] = (uint16_t)((row * RES + z) | (fb << 12));
                }
            }
        }
        wave_sync();
        float s1[3], step[3], delta[3];
        leaf_frame(spans + 6 * (size_t)leaf, (float)RES, s1, step, delta);
        const uint64_t vbase = voff[leaf];
        // Vertex placement, one vertex per lane per round.
        for (uint32_t j = lane; j < tv; j += 64) {
            const uint32_t cell = W.vlist[j];
            const uint32_t x = cell / (RES * RES), y = (cell / RES) % RES, z = cell % RES;
            float d[8];
            (void)load_cell(D, L::N1, x, y, z, d);
            float p[3];
            place_vertex(d, x, y, z, s1, step, p);
            put_vertex(verts + 7 * (vbase + j), p, delta);
        }
        // Quads, one corner per lane per round; a round's first quad numbers
        // come from a wave scan of its corners' face counts.
        uint32_t* iout = idx + ioff[leaf];
        auto pt = [&](uint32_t px, uint32_t py, uint32_t pz) {
            const uint32_t r = px * RES + py;
            return W.rp[r] + (uint32_t)__popc(W.rb[r] & ((1u << pz) - 1u));
        };
        uint32_t qbase = 0;
        for (uint32_t j0 = 0; j0 < tc; j0 += 64) {
            const uint32_t j = j0 + lane;
            const uint32_t e = j < tc ? (uint32_t)W.flist[j] : 0u;
            const uint32_t cell = e & 0xFFFu, fb = (e >> 12) & 7u;
            uint32_t round_quads;
            const uint32_t q = qbase + wave_excl_scan((uint32_t)__popc(fb), round_quads);
            if (j < tc) {
                const uint32_t x = cell / (RES * RES), y = (cell / RES) % RES, z = cell % RES;
                const bool neg = (W.rn[x * RES + y] >> z) & 1u;
                emit_quads(iout + 6ull * q, fb, neg, x, y, z, pt);
            }
            qbase += round_quads;
        }
        wave_sync();  // the lists are rewritten for the next leaf
    }
}

// 6-tap stencil points of vertex v = i / 7 (src/mesh/buffer.rs:254-264):
// tap 0 = p, tap 1 + 2a = p + unit_a * delta_a, tap 2 + 2a = p + unit_a * -delta_a.
struct SrcMeshTaps {
    const float* verts;  // 7 floats per vertex; slots 3..5 hold delta
    __device__ __forceinline__ void get(uint32_t i, float& x, float& y, float& z, float& pw) const {
        const uint32_t v = i / 7u, t = i - 7u * v;
        const float* r = verts + 7 * (size_t)v;
        float c[3] = {r[0], r[1], r[2]};
        if (t > 0) {
            const int a = (int)(t - 1) >> 1;
            const float h = (t & 1u) ? r[3 + a] : -r[3 + a];
#pragma unroll
            for (int b = 0; b < 3; ++b) c[b] = c[b] + ((a == b) ? 1.0f : 0.0f) * h;
        }
        x = c[0];
        y = c[1];
        z = c[2];
        pw = 0.0f;
    }
};

struct SrcMeshVerts {
    const float* verts;
    float power;
    __device__ __forceinline__ void get(uint32_t i, float& x, float& y, float& z, float& pw) const {
        const float* r = verts + 7 * (size_t)i;
        x = r[0];
        y = r[1];
        z = r[2];
        pw = power;
    }
};

// Forward-mode normal and distance straight into the vertex record.
struct SinkMeshVerts {
    float* verts;
    uint32_t max_iters;
    __device__ __forceinline__ bool wants_k() const { return true; }
    template <class E>
    __device__ __forceinline__ void put(uint32_t i, float dv, uint32_t k, float r, float, const E& ev) const {
        float g[4];
        ev.grad(r, k == max_iters, g);
        const float mg = sqrt_rn((g[0] * g[0] + g[1] * g[1]) + g[2] * g[2]);
        const float im = div_rn(1.0f, mg);
        float* v = verts + 7 * (size_t)i;
        v[3] = g[0] * im;
        v[4] = g[1] * im;
        v[5] = g[2] * im;
        v[6] = dv;
    }
};

// normal = normalize((DE(p+dx) - DE(p-dx), ...)) = g * (1 / |g|) (cgmath
// normalize); distance_from_surface = DE(p) (src/mesh/buffer.rs:254-272).
__device__ __forceinline__ void put_fd_normal(float* v, const float t[7]) {
    const float g[3] = {t[1] - t[2], t[3] - t[4], t[5] - t[6]};
    const float mg = sqrt_rn((g[0] * g[0] + g[1] * g[1]) + g[2] * g[2]);
    const float im = div_rn(1.0f, mg);
    v[3] = g[0] * im;
    v[4] = g[1] * im;
    v[5] = g[2] * im;
    v[6] = t[0];
}

// The 6-tap stencil with the normal formed in registers: a lane evaluates
// the 7 taps of one vertex one after another -- each tap restarts the lane
// at the wave's batched refill, exactly like a new point of the lane-refill
// loop (work_loop.h) -- keeps the tap distances and writes the normal and
// distance_from_surface itself.  The arithmetic is that of SrcMeshTaps +
// k_mesh_fd_combine; no tap buffer.  Index space: vertices.
//
// A lane's vertex record (position, half-steps) and its tap distances are
// only touched at the batched refill, so they live in the lane's own LDS
// column rather than in 13 vector registers.  The kernel then fits in 72
// VGPRs without spilling, and 7 waves share each SIMD (5 with the record in
// registers); asking for 8 forces spills and is slower.
template <class Ev, bool kCount = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) void k_mesh_taps(float* __restrict__ verts,
                                                                                          MandelUniform u, WorkQueue q) {
    // slots 0..2: vertex position, 3..5: half-steps, 6 + t: distance of tap t
    __shared__ float lane_slots[13][256];
    float* const slot = &lane_slots[0][threadIdx.x];  // slot j at slot[256 * j]
    Ev ev;
    // Lane state as in refill_loop (work_loop.h): live / holding lanes are
    // wave-uniform lane masks; `has_v` (the lane owns a vertex whose taps are
    // not all evaluated) is only read at the batched refill.
    const uint64_t wave = wave_mask(true);
    uint64_t live_m = 0, hold_m = 0;
    uint64_t last_m = 0;  // held taps whose last body's dr update is due (Ev::kLastDr, as in refill_loop)
    const bool cap0 = Ev::kLastDr && u.max_iters == 1u;  // at the cap before any full body
    bool has_v = false;
    uint32_t vid = 0, t = 0;
    uint32_t tick = 0, t0 = 0;  // wave's body count; tick at which the lane's tap started
    float r = 0.0f;
    uint32_t c_next = 0, c_end = 0, round = 0;
    bool exhausted = false;
    WorkTally<kCount> tally;

    for (;;) {
        if ((uint32_t)__popcll(wave & ~live_m) >= q.refill_min) {
            if (__builtin_amdgcn_inverse_ballot_w64(hold_m)) {
                if constexpr (Ev::kLastDr) {
                    if (__builtin_amdgcn_inverse_ballot_w64(last_m)) ev.last_dr(r);
                }
                slot[256 * (6 + t)] = ev.finish(r);
                ++t;
            }
            hold_m = 0;
            last_m = 0;
            for (;;) {
                // every lane here is idle; a lane with all 7 taps writes its
                // vertex: the whole 28-byte record, position unchanged.  A
                // 16-byte write of normal and distance alone dirties every
                // 32-byte sector the record touches anyway (measured: 28 B
                // per vertex written back either way), so the record is
                // written whole and in one place.
                if (has_v && t == 7u && !__builtin_amdgcn_inverse_ballot_w64(live_m)) {
                    float tv[7];
#pragma unroll
                    for (int j = 0; j < 7; ++j) tv[j] = slot[256 * (6 + j)];
                    float* rec = verts + 7 * (size_t)vid;
                    rec[0] = slot[0];
                    rec[1] = slot[256];
                    rec[2] = slot[512];
                    put_fd_normal(rec, tv);
                    has_v = false;
                }
                const uint64_t need = wave_mask(!has_v) & ~live_m;
                if (need != 0 && !exhausted) {
                    if (c_next >= c_end) {
                        const uint32_t b = next_chunk(q, round);
                        if (b >= q.total) {
                            exhausted = true;
                        } else {
                            c_next = b;
                            c_end = (q.total - b < q.chunk) ? q.total : b + q.chunk;
                        }
                    }
                    if (!exhausted) {
                        const uint32_t avail = c_end - c_next;
                        const uint32_t nneed = (uint32_t)__popcll(need);
                        const uint32_t rank = lane_rank(need);
                        if (__builtin_amdgcn_inverse_ballot_w64(need) && rank < avail) {
                            vid = c_next + rank;
                            const float* rec = verts + 7 * (size_t)vid;
#pragma unroll
                            for (int j = 0; j < 6; ++j) slot[256 * j] = rec[j];
                            t = 0;
                            has_v = true;
                        }
                        c_next += (nneed < avail) ? nneed : avail;
                    }
                }
                const uint64_t start_m = wave_mask(has_v) & ~live_m;
                if (start_m == 0) {
                    if (exhausted || (wave_mask(!has_v) & ~live_m) == 0) break;
                    continue;
                }
                bool now = false;  // the tap needs no full body
                if (__builtin_amdgcn_inverse_ballot_w64(start_m)) {
                    // tap t: p, or p + unit_a * (+-delta_a) (src/mesh/buffer.rs:259-264)
                    float c[3] = {slot[0], slot[256], slot[512]};
                    if (t > 0) {
                        const int a = (int)(t - 1) >> 1;
                        const float da = slot[256 * (3 + a)];
                        const float h = (t & 1u) ? da : -da;
#pragma unroll
                        for (int b = 0; b < 3; ++b) c[b] = c[b] + ((a == b) ? 1.0f : 0.0f) * h;
                    }
                    ev.start(c[0], c[1], c[2], 0.0f, u);
                    t0 = tick;
                    r = ev.mag();
                    now = r > u.bailout || cap0;
                    if (now) {
                        if constexpr (Ev::kLastDr) {
                            if (!(r > u.bailout)) ev.last_dr(r);
                        }
                        slot[256 * (6 + t)] = ev.finish(r);
                        ++t;
                    }
                }
                const uint64_t zero_m = wave_mask(r > u.bailout) & start_m;
                const uint64_t now_m = wave_mask(now) & start_m;
                live_m |= start_m & ~now_m;
                tally.eval(start_m);
                tally.bail(zero_m);
                tally.body(now_m & ~zero_m);
            }
            if (live_m == 0) break;
            tick_guard(ev, u);  // the started taps' squares
        }
        const bool live = __builtin_amdgcn_inverse_ballot_w64(live_m);
        if constexpr (Ev::kLastDr) {
            // as in refill_loop: a tap at the cap leaves after max_iters - 1
            // bodies and its last body's dr update is applied when it is held
            tick_body(ev, r, u, live);
            ++tick;
            tally.body(live_m);
            const float rn = tick_mag(ev, u);
            const uint64_t bail_m = wave_mask(rn > u.bailout) & live_m;
            tally.bail(bail_m);
            // Every lane takes rn: a lane without a live point ran the body
            // without committing it, so rn is its last magnitude again, bit
            // for bit (a held point's r, the one its finish needs).
            r = rn;
            const uint64_t cap_m = wave_mask(t0 == tick + 1u - u.max_iters) & live_m & ~bail_m;
            tally.body(cap_m);
            last_m |= cap_m;
            live_m &= ~(cap_m | bail_m);
            hold_m |= cap_m | bail_m;
        } else {
            const uint64_t cap_m = wave_mask(t0 == tick + 1u - u.max_iters) & live_m;  // the max_iters-th body
            tick_body(ev, r, u, live);
            ++tick;
            tally.body(live_m);
            const float rn = tick_mag(ev, u);
            const uint64_t bail_m = wave_mask(rn > u.bailout) & live_m & ~cap_m;
            tally.bail(bail_m);
            // at the cap r stays the magnitude before the last body (src/shape/mandelbulb.rs:65-78)
            if (__builtin_amdgcn_inverse_ballot_w64(live_m & ~cap_m)) r = rn;
            live_m &= ~(cap_m | bail_m);
            hold_m |= cap_m | bail_m;
        }
    }
    tally.flush(q.stats);
    if (q.dynamic) wave_retire(q);
}

__global__ void k_mesh_fd_combine(const float* __restrict__ taps, uint32_t n, float* __restrict__ verts) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* t = taps + 7 * (size_t)i;
    const float tv[7] = {t[0], t[1], t[2], t[3], t[4], t[5], t[6]};
    put_fd_normal(verts + 7 * (size_t)i, tv);
}

inline uint32_t cdiv_u32(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

int mesh_dims(uint32_t res, MeshDims& m) {
    if (res == 0 || (res & (res - 1)) != 0 || res > 64)
        return fail(CTH_E_INVALID, "mesh resolution must be a power of two in [1, 64]");
    m.res = res;
    m.lg = (uint32_t)__builtin_ctz(res);
    m.n1 = res + 1;
    m.cells = res * res * res;
    m.fres = (float)res;
    return CTH_OK;
}

// Persistent grid of wave-per-leaf kernels: as many workgroups per CU as
// can be resident at once (at most 8).  Leaves are dealt to waves with a
// fixed stride, so a workgroup that is not resident from the start would run
// its whole share after the others: the grid must not exceed residency.
template <class K>
inline uint32_t wave_grid(const Ctx& c, uint32_t nleaves, K kernel) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kMeshBlock, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    const uint64_t want = (uint64_t)c.num_cus * (uint64_t)std::min(per_cu, 8);
    const uint64_t need = ((uint64_t)nleaves + 3u) / 4u;
    return (uint32_t)std::max<uint64_t>(1, std::min(want, need));
}

#define LAUNCH_COUNT_W(R)                                                                                    \
    hipLaunchKernelGGL(k_mesh_count_w<R>, dim3(wave_grid(c, nleaves, k_mesh_count_w<R>)), dim3(kMeshBlock), 0, \
                       st, dists, nleaves, counts, counts + cnt, cross)
#define LAUNCH_EMIT_W(R)                                                                                    \
    hipLaunchKernelGGL(k_mesh_emit_w<R>, dim3(wave_grid(c, nleaves, k_mesh_emit_w<R>)), dim3(kMeshBlock), 0, \
                       st, spans, dists, nleaves, voff, ioff, verts, idx)

}  // namespace

// Exclusive scans of the vertex and index counts (2 * cnt entries, vertex
// counts first) into voff / ioff.
static int scan_pair(Ctx& c, uint64_t* counts, size_t cnt, uint64_t* voff, uint64_t* ioff, hipStream_t st) {
    size_t tmp_bytes = 0;
    void* tmp = nullptr;
    int rc = check_hip(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, counts, voff, (int)cnt, st), "scan size");
    if (!rc && c.scratch(WS_SCAN, tmp_bytes, &tmp)) rc = fail(CTH_E_NOMEM, "scan scratch allocation failed");
    if (!rc) rc = check_hip(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, counts, voff, (int)cnt, st), "scan");
    if (!rc)
        rc = check_hip(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, counts + cnt, ioff, (int)cnt, st), "scan");
    return rc;
}

// Plane-unit counts and their exclusive scans, in the context's scratch:
// *uoff holds units + 1 vertex offsets followed by units + 1 index offsets.
static int plane_offsets(Ctx& c, const float* dists, const MeshDims& m, uint32_t nleaves, uint64_t* cross,
                         uint64_t** uoff, hipStream_t st) {
    const size_t units = (size_t)nleaves * m.res;
    if (units >= (size_t(1) << 31)) return fail(CTH_E_INVALID, "more than 2^31 mesh plane units");
    const size_t cnt = units + 1;
    uint64_t* counts = nullptr;
    if (c.scratch(WS_COUNTS, 2 * cnt * sizeof(uint64_t), (void**)&counts) ||
        c.scratch(WS_UOFF, 2 * cnt * sizeof(uint64_t), (void**)uoff))
        return fail(CTH_E_NOMEM, "mesh scratch allocation failed");
    hipLaunchKernelGGL(k_mesh_count_p, dim3((uint32_t)units), dim3(kMeshBlock), 0, st, dists, m, (uint32_t)units,
                       counts, counts + cnt, cross);
    int rc = check_hip(hipGetLastError(), "mesh plane count launch");
    if (!rc) rc = scan_pair(c, counts, cnt, *uoff, *uoff + cnt, st);
    return rc;
}

int mesh_count(Ctx& c, const float* dists, uint32_t nleaves, uint32_t res, uint64_t* voff, uint64_t* ioff,
               uint64_t* cross, uint64_t totals[2], hipStream_t st) {
    MeshDims m;
    if (int rc = mesh_dims(res, m)) return rc;
    totals[0] = totals[1] = 0;
    if (nleaves == 0) {
        const uint64_t zero[2] = {0, 0};
        int rc = check_hip(hipMemcpyAsync(voff, zero, 8, hipMemcpyHostToDevice, st), "H2D offsets");
        if (!rc) rc = check_hip(hipMemcpyAsync(ioff, zero + 1, 8, hipMemcpyHostToDevice, st), "H2D offsets");
        if (!rc) rc = check_hip(hipStreamSynchronize(st), "mesh count sync");
        return rc;
    }
    int rc = CTH_OK;
    if (res <= 16) {
        // wave per leaf: the counts are per leaf already
        const size_t cnt = (size_t)nleaves + 1;
        uint64_t* counts = nullptr;
        if (c.scratch(WS_COUNTS, 2 * cnt * sizeof(uint64_t), (void**)&counts))
            return fail(CTH_E_NOMEM, "mesh count scratch allocation failed");
        switch (res) {
            case 1: LAUNCH_COUNT_W(1); break;
            case 2: LAUNCH_COUNT_W(2); break;
            case 4: LAUNCH_COUNT_W(4); break;
            case 8: LAUNCH_COUNT_W(8); break;
            default: LAUNCH_COUNT_W(16); break;
        }
        rc = check_hip(hipGetLastError(), "mesh count launch");
        if (!rc) rc = scan_pair(c, counts, cnt, voff, ioff, st);
    } else {
        // plane units, then the leaves' offsets from the units'
        uint64_t* uoff = nullptr;
        rc = plane_offsets(c, dists, m, nleaves, cross, &uoff, st);
        if (!rc) {
            const size_t units = (size_t)nleaves * res;
            hipLaunchKernelGGL(k_leaf_offsets, dim3(cdiv_u32((uint64_t)nleaves + 1, 256)), dim3(256), 0, st, uoff,
                               uoff + units + 1, nleaves, res, voff, ioff);
            rc = check_hip(hipGetLastError(), "leaf offsets launch");
        }
    }
    if (!rc) rc = check_hip(hipMemcpyAsync(c.pinned, voff + nleaves, 8, hipMemcpyDeviceToHost, st), "D2H totals");
    if (!rc) rc = check_hip(hipMemcpyAsync(c.pinned + 1, ioff + nleaves, 8, hipMemcpyDeviceToHost, st), "D2H totals");
    int rs = check_hip(hipStreamSynchronize(st), "mesh count sync");
    if (!rc && !rs) {
        totals[0] = c.pinned[0];
        totals[1] = c.pinned[1];
    }
    return rc ? rc : rs;
}

int mesh_emit(Ctx& c, const cth_shape& s, const float* spans, const float* dists, uint32_t nleaves, uint32_t res,
              int normals, const uint64_t* voff, const uint64_t* ioff, uint64_t nverts, float* verts,
              uint32_t* idx, hipStream_t st) {
    MeshDims m;
    if (int rc = mesh_dims(res, m)) return rc;
    if (normals != CTH_NORMALS_FD6 && normals != CTH_NORMALS_TANGENT && normals != CTH_NORMALS_DEFERRED)
        return fail(CTH_E_INVALID, "unknown normals mode");
    if (normals == CTH_NORMALS_TANGENT && s.kind == CTH_SHAPE_SPHERE)
        return fail(CTH_E_UNSUPPORTED, "tangent normals need a Mandelbulb kind");
    if (nleaves == 0) return CTH_OK;
    if (res <= 16) {
        switch (res) {
            case 1: LAUNCH_EMIT_W(1); break;
            case 2: LAUNCH_EMIT_W(2); break;
            case 4: LAUNCH_EMIT_W(4); break;
            case 8: LAUNCH_EMIT_W(8); break;
            default: LAUNCH_EMIT_W(16); break;
        }
    } else {
        // the plane units' offsets (recounted: a count call leaves only the leaves')
        uint64_t* uoff = nullptr;
        if (int rc = plane_offsets(c, dists, m, nleaves, nullptr, &uoff, st)) return rc;
        const size_t units = (size_t)nleaves * res;
        const size_t lds = 2 * (size_t)(m.res * m.res / 64) * (sizeof(uint64_t) + sizeof(uint32_t));
        hipLaunchKernelGGL(k_mesh_emit_p, dim3((uint32_t)units), dim3(kMeshBlock), lds, st, spans, dists, m, uoff,
                           uoff + units + 1, voff, ioff, verts, idx);
    }
    int rc = check_hip(hipGetLastError(), "mesh emit launch");
    if (rc || nverts == 0 || normals == CTH_NORMALS_DEFERRED) return rc;
    return mesh_vertices(c, s, normals, nverts, verts, st);
}

// The vertex pass: distance and normal of records that hold their position
// and, in the normal slots, their leaf's 6-tap half-steps.
int mesh_vertices(Ctx& c, const cth_shape& s, int normals, uint64_t nverts, float* verts, hipStream_t st) {
    if (normals != CTH_NORMALS_FD6 && normals != CTH_NORMALS_TANGENT)
        return fail(CTH_E_INVALID, "the vertex pass computes FD6 or TANGENT normals");
    if (normals == CTH_NORMALS_TANGENT && s.kind == CTH_SHAPE_SPHERE)
        return fail(CTH_E_UNSUPPORTED, "tangent normals need a Mandelbulb kind");
    if (nverts == 0) return CTH_OK;
    int rc = CTH_OK;
    MandelUniform u{};
    int family = FAM_SPHERE;
    if (s.kind != CTH_SHAPE_SPHERE) {
        rc = resolve_mandel(s, u, family);
        if (rc) return rc;
    }
    if (normals == CTH_NORMALS_TANGENT) {
        if (nverts >= (1ull << 31)) return fail(CTH_E_INVALID, "more than 2^31 vertices");
        uint32_t blocks = 0;
        WorkQueue q = make_queue(c, (uint32_t)nverts, c.shape_mesh, blocks, st);
        SrcMeshVerts src{verts, s.power};
        SinkMeshVerts sink{verts, u.max_iters};
        if (family == FAM_P8)
            hipLaunchKernelGGL((k_refill<P8Tangent, SrcMeshVerts, SinkMeshVerts>), dim3(blocks),
                               dim3(c.shape_mesh.block), 0, st, src, u, sink, q);
        else
            hipLaunchKernelGGL((k_refill<TrigNormal, SrcMeshVerts, SinkMeshVerts>), dim3(blocks),
                               dim3(c.shape_mesh.block), 0, st, src, u, sink, q);
        return check_hip(hipGetLastError(), "mesh tangent normals launch");
    }

    // 6-tap central differences, 7 evaluations per vertex.  Mandelbulb kinds:
    // one launch, the lane forms the normal from its own taps.
    if (family != FAM_SPHERE) {
        if (nverts >= (1ull << 31)) return fail(CTH_E_INVALID, "more than 2^31 vertices");
        uint32_t blocks = 0;
        WorkQueue q = make_queue(c, (uint32_t)nverts, c.shape_mesh, blocks, st);
        dim3 grid(blocks), block(c.shape_mesh.block);
        if (family == FAM_P8) {
            with_p8(c.mode, [&](auto ev) {
                using Ev = decltype(ev);
                if (q.stats)
                    hipLaunchKernelGGL((k_mesh_taps<Ev, true>), grid, block, 0, st, verts, u, q);
                else
                    hipLaunchKernelGGL((k_mesh_taps<Ev>), grid, block, 0, st, verts, u, q);
            });
        } else if (q.stats) {
            hipLaunchKernelGGL((k_mesh_taps<TrigExact, true>), grid, block, 0, st, verts, u, q);
        } else {
            hipLaunchKernelGGL((k_mesh_taps<TrigExact>), grid, block, 0, st, verts, u, q);
        }
        return check_hip(hipGetLastError(), "mesh taps launch");
    }
    // Sphere: one thread per tap, then the combine, in batches whose index
    // space stays below 2^31.
    const uint64_t batch = std::min<uint64_t>(nverts, ((1ull << 31) - 1) / 7);
    float* taps = nullptr;
    rc = check_hip(hipMallocAsync((void**)&taps, batch * 7 * sizeof(float), st), "hipMallocAsync taps");
    for (uint64_t off = 0; off < nverts && !rc; off += batch) {
        const uint32_t nb = (uint32_t)std::min<uint64_t>(batch, nverts - off);
        float* vb = verts + 7 * off;
        SrcMeshTaps src{vb};
        hipLaunchKernelGGL((k_sphere<SrcMeshTaps>), dim3(cdiv_u32((uint64_t)nb * 7, 256)), dim3(256), 0, st, src,
                           nb * 7, s.center[0], s.center[1], s.center[2], s.radius, taps, nullptr);
        rc = check_hip(hipGetLastError(), "mesh taps launch");
        if (rc) break;
        hipLaunchKernelGGL(k_mesh_fd_combine, dim3(cdiv_u32(nb, 256)), dim3(256), 0, st, taps, nb, vb);
        rc = check_hip(hipGetLastError(), "mesh normals launch");
    }
    if (taps) (void)hipFreeAsync(taps, st);
    return rc;
}

}  // namespace cth
