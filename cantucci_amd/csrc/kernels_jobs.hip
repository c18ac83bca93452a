// Mixed-size batched jobs in one persistent launch.
//
// The reference meshes each octree leaf as an independent thread-pool job
// (src/mesh/mod.rs:129-161).  Here a batch of independent grid jobs -- each
// with its own size n^3, power and iteration cap -- is one index space: the
// jobs' grids are concatenated, every lane of the persistent lane-refill loop
// finds its point's job by a binary search of the job offsets (kept in LDS),
// and evaluates distance plus the forward-mode (tangent) gradient for the
// unit normal.  A job's completion time is the latest wall-clock stamp of its
// points' write-back (one atomicMax per job per wave refill), so per-job
// latency comes out of the same launch.  The launch (plus its small reset
// kernel) can be captured in a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "internal.h"
#include "launch.h"
#include "point_sources.h"
#include "tangent.h"
#include "work_loop.h"

namespace cth {

struct JobDev {
    float start[3];
    float across[3];
    float divisor, inv;  // inv != 0 when divisor is a power of two
    float power;
    uint32_t max_iters;
    FastDiv n, n2;
    uint32_t offset;
};

constexpr int kMaxJobs = 1024;

__global__ void k_jobs_reset(unsigned long long* done, uint32_t njobs, unsigned long long* start) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < njobs) done[i] = 0ull;
    if (i == 0) *start = ~0ull;
}

template <class Ev>
__global__ __launch_bounds__(256) void k_jobs(const JobDev* __restrict__ jobs, uint32_t njobs, MandelUniform u,
                                              float* __restrict__ dists, float* __restrict__ normals,
                                              unsigned long long* __restrict__ done, unsigned long long* start,
                                              WorkQueue q) {
    __shared__ uint32_t offs[kMaxJobs + 1];
    for (uint32_t i = threadIdx.x; i <= njobs; i += blockDim.x) offs[i] = (i < njobs) ? jobs[i].offset : q.total;
    __syncthreads();
    if ((threadIdx.x & 63u) == 0) atomicMin(start, (unsigned long long)wall_clock64());

    Ev ev;
    // live / holding lanes as wave-uniform lane masks (see refill_loop)
    const uint64_t wave = wave_mask(true);
    uint64_t live_m = 0, hold_m = 0;
    uint64_t capd_m = 0;  // held lanes whose point stopped at its job's iteration cap
    uint32_t idx = 0, job = 0, kcap = 1;
    // The wave's body count, and per lane the tick of its point's last body
    // (the job's cap counted from the tick at which the lane took the point).
    uint32_t tick = 0, tcap = 0;
    float r = 0.0f;
    uint32_t c_next = 0, c_end = 0, round = 0;
    bool exhausted = false;
    SinkGrad sink{dists, normals, nullptr, nullptr, nullptr, 0u};

    for (;;) {
        if ((uint32_t)__popcll(wave & ~live_m) >= q.refill_min) {
            // write back held points, then stamp their jobs (one atomic per job per wave)
            // SinkGrad reads k only as k == max_iters (the point stopped at the cap).
            if (__builtin_amdgcn_inverse_ballot_w64(hold_m)) {
                sink.max_iters = kcap;
                sink.put(idx, ev.finish(r), __builtin_amdgcn_inverse_ballot_w64(capd_m) ? kcap : 0u, r, ev.dr, ev);
            }
            uint64_t pend = hold_m;
            if (pend) {
                unsigned long long now = wall_clock64();
                while (pend) {
                    int lead = __ffsll((unsigned long long)pend) - 1;
                    uint32_t jl = (uint32_t)__builtin_amdgcn_readlane((int)job, lead);
                    uint64_t same = wave_mask(job == jl) & hold_m;
                    if ((int)__lane_id() == lead) atomicMax(done + jl, now);
                    pend &= ~same;
                }
            }
            hold_m = 0;
            capd_m = 0;
            while (!exhausted) {
                const uint64_t need = wave & ~live_m;
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
                const uint32_t rank = lane_rank(need);
                const bool take = __builtin_amdgcn_inverse_ballot_w64(need) && rank < avail;
                {
                    if (take) {
                        idx = c_next + rank;
                        // job of idx: the last offset <= idx
                        uint32_t lo = 0, hi = njobs;
                        while (hi - lo > 1) {
                            uint32_t mid = (lo + hi) >> 1;
                            if (offs[mid] <= idx) lo = mid; else hi = mid;
                        }
                        job = lo;
                        const JobDev& J = jobs[job];
                        uint32_t local = idx - J.offset;
                        uint32_t i = fdiv(local, J.n2);
                        uint32_t rem = local - i * J.n2.d;
                        uint32_t j = fdiv(rem, J.n);
                        uint32_t kk = rem - j * J.n.d;
                        uint32_t ijk[3] = {i, j, kk};
                        float c[3];
#pragma unroll
                        for (int a = 0; a < 3; ++a) {
                            float v = (J.inv != 0.0f) ? (float)ijk[a] * J.inv : div_rn((float)ijk[a], J.divisor);
                            c[a] = J.start[a] + J.across[a] * v;
                        }
                        kcap = J.max_iters;
                        tcap = tick + kcap - 1u;
                        ev.start(c[0], c[1], c[2], J.power, u);
                        r = ev.mag();
                        if (r > u.bailout) {
                            sink.max_iters = kcap;
                            sink.put(idx, ev.finish(r), 0u, r, ev.dr, ev);
                        }
                    }
                }
                live_m |= wave_mask(take) & ~wave_mask(r > u.bailout);
                c_next += (nneed < avail) ? nneed : avail;
            }
            if (live_m == 0) break;
        }
        const bool live = __builtin_amdgcn_inverse_ballot_w64(live_m);
        const uint64_t cap_m = wave_mask(tcap == tick) & live_m;
        ev.body(r, u, live);
        ++tick;
        const float rn = ev.mag();
        const uint64_t bail_m = wave_mask(rn > u.bailout) & live_m & ~cap_m;
        if (__builtin_amdgcn_inverse_ballot_w64(live_m & ~cap_m)) r = rn;
        live_m &= ~(cap_m | bail_m);
        hold_m |= cap_m | bail_m;
        capd_m |= cap_m;
    }
    if (q.dynamic) wave_retire(q);
}

struct JobSet {
    int kind;
    float bailout;
    uint32_t njobs = 0;
    uint32_t total = 0;
    JobDev* dev = nullptr;
    std::vector<uint32_t> offsets;
};

static bool pow2_inv(float v, float& inv) {
    int e = 0;
    float m = frexpf(v, &e);
    inv = (v > 0.0f && m == 0.5f) ? ldexpf(1.0f, 1 - e) : 0.0f;
    return inv != 0.0f;
}

int jobs_prepare(Ctx&, int kind, float bailout, const cth_job* jobs, uint32_t njobs, JobSet** out) {
    if (njobs == 0 || njobs > kMaxJobs) return fail(CTH_E_INVALID, "job count must be in [1, 1024]");
    if (kind != CTH_SHAPE_MANDELBULB && kind != CTH_SHAPE_MANDELBULB_TRIG)
        return fail(CTH_E_INVALID, "jobs need a Mandelbulb kind");
    std::vector<JobDev> h(njobs);
    uint64_t off = 0;
    JobSet* js = new JobSet();
    js->kind = kind;
    js->bailout = bailout;
    for (uint32_t i = 0; i < njobs; ++i) {
        const cth_job& j = jobs[i];
        if (j.n == 0 || j.max_iters == 0 || !(j.divisor > 0.0f)) {
            delete js;
            return fail(CTH_E_INVALID, "job with n == 0, max_iters == 0 or divisor <= 0");
        }
        if (kind == CTH_SHAPE_MANDELBULB && (j.power != 8.0f)) {
            delete js;
            return fail(CTH_E_INVALID, "CTH_SHAPE_MANDELBULB jobs run the P8 polynomial: power must be 8");
        }
        JobDev& d = h[i];
        for (int a = 0; a < 3; ++a) {
            d.start[a] = j.start[a];
            d.across[a] = j.end[a] - j.start[a];
        }
        d.divisor = j.divisor;
        pow2_inv(j.divisor, d.inv);
        d.power = j.power;
        d.max_iters = j.max_iters;
        d.n = make_fastdiv(j.n);
        d.n2 = make_fastdiv(j.n * j.n);
        d.offset = (uint32_t)off;
        js->offsets.push_back((uint32_t)off);
        off += (uint64_t)j.n * j.n * j.n;
        if (off >= (1ull << 31)) {
            delete js;
            return fail(CTH_E_INVALID, "jobs exceed 2^31 points");
        }
    }
    js->njobs = njobs;
    js->total = (uint32_t)off;
    int rc = check_hip(hipMalloc((void**)&js->dev, njobs * sizeof(JobDev)), "hipMalloc jobs");
    if (!rc) rc = check_hip(hipMemcpy(js->dev, h.data(), njobs * sizeof(JobDev), hipMemcpyHostToDevice), "H2D jobs");
    if (rc) {
        if (js->dev) (void)hipFree(js->dev);
        delete js;
        return rc;
    }
    *out = js;
    return CTH_OK;
}

void jobs_destroy(JobSet* js) {
    if (!js) return;
    if (js->dev) (void)hipFree(js->dev);
    delete js;
}

uint32_t jobs_total(const JobSet* js) { return js ? js->total : 0; }
uint32_t jobs_offset(const JobSet* js, uint32_t i) { return (js && i < js->njobs) ? js->offsets[i] : 0; }

int jobs_run(Ctx& c, const JobSet& js, float* dists, float* normals, unsigned long long* done,
             unsigned long long* start, hipStream_t st) {
    MandelUniform u{};
    u.max_iters = 0;  // per job
    u.bailout = js.bailout;
    u.power = 8.0f;
    u.integral = (js.kind == CTH_SHAPE_MANDELBULB) ? 1 : 0;
    u.ipower = 8;
    hipLaunchKernelGGL(k_jobs_reset, dim3((js.njobs + 255) / 256), dim3(256), 0, st, done, js.njobs, start);
    int rc = check_hip(hipGetLastError(), "jobs reset launch");
    if (rc) return rc;
    uint32_t blocks = 0;
    WorkQueue q = make_queue(c, js.total, c.shape_jobs, blocks, st);
    if (js.kind == CTH_SHAPE_MANDELBULB)
        hipLaunchKernelGGL((k_jobs<P8Tangent>), dim3(blocks), dim3(c.shape_jobs.block), 0, st, js.dev, js.njobs, u,
                           dists, normals, done, start, q);
    else
        hipLaunchKernelGGL((k_jobs<TrigNormal>), dim3(blocks), dim3(c.shape_jobs.block), 0, st, js.dev, js.njobs, u,
                           dists, normals, done, start, q);
    return check_hip(hipGetLastError(), "jobs launch");
}

}  // namespace cth
