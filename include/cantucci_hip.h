/*
 * cantucci_hip.h -- C ABI of the MI355X (gfx950) Mandelbulb distance-estimation
 * library.
 *
 * The entry points replace cantucci's `Shape` trait surface and the hot loops
 * of its callers (citations are path:line in the cantucci source tree):
 *
 *   Shape::min_distance_from / batch_min_distance_from  src/shape/mod.rs:37,89
 *   Shape::max_distance_from / batch_max_distance_from  src/shape/mod.rs:62-64,96
 *   Shape::bounded_distance_from / batch_bounded_...    src/shape/mod.rs:73-75,103
 *   Shape::contains                                     src/shape/mod.rs:78-80
 *   Shape::bounding_box                                 src/shape/mod.rs:39
 *   MeshBuffer step 1 (corner sampling) + sign test     src/mesh/buffer.rs:63-147
 *   MeshBuffer steps 2-3 (vertices, normals, quads)     src/mesh/buffer.rs:97-372
 *   MeshBuffer::generate_for_box                        src/mesh/buffer.rs:30-42
 *   ShapeMesh::get_focii (sphere tracing)               src/mesh/mod.rs:205-243
 *   ShapeMesh::update split rule, Octree::leaf_around  src/mesh/mod.rs:91-104, src/octree/mod.rs:75-89
 *
 * Conventions
 *   - Plain C types only.  Host-pointer entry points are synchronous (they use
 *     the context's own stream); `_dev` entry points take device pointers and
 *     are asynchronous on `stream` (a hipStream_t passed as void*; NULL is
 *     HIP's default stream).
 *   - Every call returns a cth_status; on failure cth_last_error() returns a
 *     thread-local message.  Nothing aborts the process.  Where the reference
 *     panics (e.g. `batch_max_distance_from` on a Mandelbulb,
 *     src/shape/util.rs:8) the call returns CTH_E_UNSUPPORTED.
 *   - The caller owns every buffer.  The library never frees caller memory.
 *   - A context owns one HIP stream and small device work counters; calls on
 *     one context must not overlap in time (use one context per host thread,
 *     the analogue of the reference sharing `Arc<dyn Shape>` across its pool).
 *   - There is no CPU fallback: without a usable gfx950 device,
 *     cth_ctx_create fails with CTH_E_DEVICE.
 */
#ifndef CANTUCCI_HIP_H
#define CANTUCCI_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum cth_status {
    CTH_OK = 0,
    CTH_E_INVALID = 1,     /* bad argument (null pointer, max_iters == 0, ...) */
    CTH_E_UNSUPPORTED = 2, /* the reference would panic / has no such bound */
    CTH_E_DEVICE = 3,      /* HIP runtime error */
    CTH_E_NOMEM = 4        /* device allocation failed */
} cth_status;

/* Shape kinds.
 *   CTH_SHAPE_MANDELBULB       reference `Mandelbulb<P>` (src/shape/mandelbulb.rs:13):
 *                              `power` must be an integer in [1, 255]; P == 8 uses
 *                              the closed-form polynomial (:148-200), any other P
 *                              the trigonometric map with integer powers (:128-146).
 *   CTH_SHAPE_MANDELBULB_TRIG  the trigonometric map with a real power (powf); the
 *                              per-point power arrays of the gradient entry points
 *                              apply to this kind.
 *   CTH_SHAPE_SPHERE           reference `Sphere` (src/shape/sphere.rs:6-52). */
enum {
    CTH_SHAPE_MANDELBULB = 0,
    CTH_SHAPE_MANDELBULB_TRIG = 1,
    CTH_SHAPE_SPHERE = 2
};

/* Arithmetic modes of a context (they select the P8 evaluator; the trig map
 * and the sphere always run their parity arithmetic).
 *   CTH_MODE_PARITY    the reference's f32 operation sequence (no FMA, IEEE
 *                      div/sqrt): loop state bit-identical, final ln within an ulp.
 *   CTH_MODE_FAST      FMAs, hardware reciprocal/sqrt; not bit-exact.
 *   CTH_MODE_CONTRACT  the parity sequence with FMA contraction (a product
 *                      and a following sum fused), IEEE div/sqrt kept: what a
 *                      contracting compiler would make of the reference. */
enum { CTH_MODE_PARITY = 0, CTH_MODE_FAST = 1, CTH_MODE_CONTRACT = 2 };

/* Shape descriptor (the reference's type-level parameters made runtime):
 * `Mandelbulb::new(max_iters, bailout)` (src/shape/mandelbulb.rs:19-26, asserts
 * max_iters >= 1) and `Sphere::new(center, radius)` (src/shape/sphere.rs:13-19). */
typedef struct cth_shape {
    int32_t kind;
    float power;
    uint32_t max_iters;
    float bailout;
    float center[3];
    float radius;
} cth_shape;

/* cgmath `Point3<f32>`: 12-byte AoS. */
typedef struct cth_point3 {
    float x, y, z;
} cth_point3;

/* octree `Span = Range<Point3<f32>>` (src/octree/mod.rs:13). */
typedef struct cth_span {
    float start[3];
    float end[3];
} cth_span;

/* Pinhole camera: eye position, orthonormal right/up/forward basis (the
 * look_at_rh frame of src/camera.rs:69-75), tan(vertical fov / 2) and
 * width / height (src/camera.rs:114-144,186-189). */
typedef struct cth_camera {
    float eye[3];
    float right[3];
    float up[3];
    float fwd[3];
    float tan_half_fov;
    float aspect;
} cth_camera;

/* The reference's camera (src/camera.rs:16-20 `Camera`, :114-127
 * `Projection`): position, unit view direction (Camera::new normalises it),
 * vertical field of view in radians, aspect ratio width / height, near and
 * far plane.  The up vector is always (0, 0, 1) (src/camera.rs:69-75). */
typedef struct cth_ref_camera {
    float position[3];
    float direction[3];
    float fov;
    float aspect;
    float near_plane;
    float far_plane;
} cth_ref_camera;

typedef struct cth_ctx cth_ctx;

/* ---- context ------------------------------------------------------------ */
int cth_ctx_create(int device, cth_ctx** out);
void cth_ctx_destroy(cth_ctx* ctx);
const char* cth_last_error(void);
int cth_ctx_set_mode(cth_ctx* ctx, int mode);
int cth_ctx_get_mode(const cth_ctx* ctx);
void* cth_ctx_stream(cth_ctx* ctx);
int cth_synchronize(cth_ctx* ctx);
/* Number of compute units of the context's device. */
int cth_ctx_num_cus(const cth_ctx* ctx);

/* Launch shape of a kernel family (the autotuner's knobs).  block is 64, 128
 * or 256 threads; blocks_per_cu persistent workgroups per CU (1..32); chunk
 * indices claimed per wave from the work counter (64..65536); refill_min the
 * number of idle lanes (1..64) at which a wave writes finished points and
 * takes new ones; dynamic = 1 claims chunks from an atomic counter, 0 walks
 * chunks w, w+W, w+2W, ... statically (W waves).  No counterpart in the
 * reference (its work is scheduled by a CPU thread pool,
 * src/mesh/mod.rs:60-63,141-148). */
typedef struct cth_launch {
    int32_t block;
    int32_t blocks_per_cu;
    uint32_t chunk;
    uint32_t refill_min;
    uint32_t dynamic;
} cth_launch;

/* Kernel families: point/grid distances, sphere tracing, gradients (VJP,
 * tangent, normals), octree span sampling, batched jobs, and the mesh
 * vertex evaluations (6-tap stencils or tangent normals). */
enum {
    CTH_KERNEL_DE = 0,
    CTH_KERNEL_TRACE = 1,
    CTH_KERNEL_ADJOINT = 2,
    CTH_KERNEL_SPANS = 3,
    CTH_KERNEL_JOBS = 4,
    CTH_KERNEL_MESH = 5
};

int cth_ctx_set_launch(cth_ctx* ctx, int kernel, const cth_launch* launch);
int cth_ctx_get_launch(const cth_ctx* ctx, int kernel, cth_launch* launch);

/* Work counters for measurement: while `counters` (device, 3 x uint64) is
 * set, every lane-refill launch of the context (distances, grids, span
 * sampling, tracing, the mesh's vertex taps) adds to counters[0] the loop
 * bodies it ran, to counters[1] its distance evaluations that ended by
 * bailout (the final magnitude test passed), to counters[2] its distance
 * evaluations.  With the per-evaluation operation count of the map they give
 * a launch's algorithmic FLOP.  NULL stops the counting.  The caller zeroes
 * the counters.  While counting, the launches run counting instantiations of
 * their kernels, which are slower (about 11% for span sampling): time launches
 * with counting off.  No counterpart in the reference. */
int cth_ctx_count_work(cth_ctx* ctx, uint64_t* counters);

/* ---- Shape trait mirror (host buffers, synchronous) ------------------------ */

/* Shape::bounding_box (src/shape/mod.rs:39): Mandelbulb [-1.2,1.2]^3
 * (src/shape/mandelbulb.rs:53-57), Sphere center +- radius (src/shape/sphere.rs:27-30). */
int cth_bounding_box(const cth_shape* shape, cth_span* out);

/* Shape::batch_min_distance_from (src/shape/mod.rs:89, src/shape/util.rs:3-5). */
int cth_min_distance_batch(cth_ctx* ctx, const cth_shape* shape, const cth_point3* pts, size_t n, float* out);

/* Shape::batch_max_distance_from (src/shape/mod.rs:96, src/shape/util.rs:7-9):
 * Sphere returns its exact distance (src/shape/sphere.rs:37-39); Mandelbulb has
 * no upper bound (src/shape/mod.rs:62-64) -> CTH_E_UNSUPPORTED. */
int cth_max_distance_batch(cth_ctx* ctx, const cth_shape* shape, const cth_point3* pts, size_t n, float* out);

/* Shape::batch_bounded_distance_from (src/shape/mod.rs:103, src/shape/util.rs:11-16):
 * out holds (min, max) pairs; CTH_E_UNSUPPORTED for Mandelbulb. */
int cth_bounded_distance_batch(cth_ctx* ctx, const cth_shape* shape, const cth_point3* pts, size_t n,
                               float* out_pairs);

/* Shape::contains (src/shape/mod.rs:78-80: DE < 0; Sphere overrides with
 * |c - p|^2 <= r^2, src/shape/sphere.rs:23-26).  One byte (0/1) per point. */
int cth_contains_batch(cth_ctx* ctx, const cth_shape* shape, const cth_point3* pts, size_t n, uint8_t* out);

/* The backward of a differentiable distance op whose forward is
 * cth_min_distance_batch (host buffers, synchronous), for a host-side op with
 * an autograd function's contract (the forward keeps its inputs, the backward
 * maps the output's gradient to the inputs').  The reference has no derivative: this differentiates its min_distance_from
 * (src/shape/mandelbulb.rs:59-79).  With sweep CTH_SWEEP_FORWARD the results
 * are those of cth_de_grad_dev, with CTH_SWEEP_REVERSE those of
 * cth_de_vjp_dev, bit for bit.  powers (per-point powers of
 * CTH_SHAPE_MANDELBULB_TRIG) and grad_d (NULL: ones) are nullable inputs;
 * d, grad_p (n x 3) and grad_power are nullable outputs. */
enum { CTH_SWEEP_FORWARD = 0, CTH_SWEEP_REVERSE = 1 };
int cth_de_vjp_batch(cth_ctx* ctx, const cth_shape* shape, const cth_point3* pts, const float* powers, size_t n,
                     const float* grad_d, float* d, float* grad_p, float* grad_power, int sweep);

/* ---- device-pointer entry points (asynchronous on `stream`) ------------------ */

/* Distances of n AoS points (3 floats each).  `powers` (nullable) gives a
 * per-point power for CTH_SHAPE_MANDELBULB_TRIG.  `iters` (nullable) receives
 * the number of loop bodies each point executed (src/shape/mandelbulb.rs:65-75). */
int cth_min_distance_dev(cth_ctx* ctx, const cth_shape* shape, const float* pts, const float* powers, size_t n,
                         float* out, uint32_t* iters, void* stream);

/* The loop state the distance is formed from: final r, dr and iteration count
 * (src/shape/mandelbulb.rs:61-78).  Used to check bit-exactness of the loop. */
int cth_de_state_dev(cth_ctx* ctx, const cth_shape* shape, const float* pts, const float* powers, size_t n,
                     float* r, float* dr, uint32_t* iters, void* stream);

/* Regular n^3 grid in GridTable order (x slowest, z fastest; src/util/grid.rs:22-51):
 *   p(i,j,k) = start + (end - start) * ((i, j, k) / divisor)
 * which is the reference's corner formula (src/mesh/buffer.rs:78-80) with
 * `divisor` = resolution.  `iters` nullable. */
int cth_de_grid_dev(cth_ctx* ctx, const cth_shape* shape, const float start[3], const float end[3], uint32_t n,
                    float divisor, float* out, uint32_t* iters, void* stream);

/* Octree-leaf span sampling + sign classification: naive surface nets step 1
 * and its crossing test (src/mesh/buffer.rs:63-83, 113-147).  For each span the
 * adjusted span s' = s - (e - s)/res, e' = e + (e - s)/res is sampled at
 * (res+1)^3 corners (GridTable order) into dists[span*(res+1)^3 ...].  If
 * cross_bits is non-null, bit c of span word block w = ceil(res^3/64) is set
 * when cell c = (x*res + y)*res + z has corners of both IEEE signs.  `res`
 * must be a power of two (src/mesh/buffer.rs:38-39).  `spans` holds 6 floats
 * per span (start xyz, end xyz).  `iters` (nullable) receives each corner's
 * loop-body count, laid out like `dists`. */
int cth_sample_spans_dev(cth_ctx* ctx, const cth_shape* shape, const float* spans, size_t nspans, uint32_t res,
                         float* dists, uint64_t* cross_bits, uint32_t* iters, void* stream);

/* Level-synchronous octree refinement (the octree-adaptive workload's policy;
 * the reference splits leaves near camera focus points, src/mesh/mod.rs:91-104,
 * after pre-splitting two levels, :52-56).  Starting from `root` (6 floats),
 * every node of depth d < min_depth splits; deeper nodes split when
 * |DE(centre)| < half_diag[d] or the distance is NaN; centre and children follow
 * SpanExt::center and create_spans (src/octree/mod.rs:21-23, 315-329).  Nodes of
 * max_depth that pass the test are the surface leaves: their count goes to
 * *nleaves and, if `leaves` (device, 6 floats each) is non-null, the first
 * `cap` of them in breadth-first create_spans order.  half_diag and
 * level_counts (nullable; nodes per depth) are host arrays of max_depth + 1.
 * Returns after the leaf count is known (kernels run on `stream`). */
int cth_refine_dev(cth_ctx* ctx, const cth_shape* shape, const float root[6], uint32_t min_depth, uint32_t max_depth,
                   const float* half_diag, float* leaves, size_t cap, size_t* nleaves, uint64_t* level_counts,
                   void* stream);

/* The reference's own split rule (ShapeMesh::update, src/mesh/mod.rs:91-104)
 * over a set of leaves that tiles an octree (n spans, 6 floats each, device):
 * flags[i] = 1 when some focus point (nfocii x 3 floats, device) lies in leaf
 * i's half-open span (SpanExt::contains, src/octree/mod.rs:25-31) and the
 * eye is nearer to it than 2 |end.x - start.x|.  depth (nullable, nfocii)
 * skips the focus points whose depth is not finite -- the rays get_focii
 * drops (src/mesh/mod.rs:229-242); hit_pos and depth of
 * cth_sphere_trace_dev can be passed straight in. */
int cth_focus_split_dev(cth_ctx* ctx, const float* leaves, size_t n, const float* focii, const float* depth,
                        size_t nfocii, const float eye[3], uint32_t* flags, void* stream);

/* Replace every flagged leaf by its 8 children (create_spans order,
 * src/octree/mod.rs:315-329) in place: out holds, for each leaf in order,
 * either the leaf or its children.  *nout receives the span count; out
 * (device, cap spans) may be NULL to query it.  Returns after *nout is known. */
int cth_split_leaves_dev(cth_ctx* ctx, const float* leaves, size_t n, const uint32_t* flags, float* out, size_t cap,
                         size_t* nout, void* stream);

/* Sphere tracing of explicit rays: the marcher of ShapeMesh::get_focii
 * (src/mesh/mod.rs:229-241):  d = DE(pos); pos += dir * d; t += d; a ray hits
 * when d < eps_abs + eps_rel * t (the reference: eps_abs = 1e-6, eps_rel = 0),
 * escapes when t > t_max (the reference has none: +inf) and stops after
 * max_steps (the reference: 100).  depth = t on a hit, +inf otherwise; steps =
 * distance evaluations; hit_pos (nullable, 3 floats per ray) = the position
 * after the last step (get_focii returns it).  origins holds one point, or one
 * per ray when per_ray_origin != 0. */
int cth_sphere_trace_dev(cth_ctx* ctx, const cth_shape* shape, const float* origins, int per_ray_origin,
                         const float* dirs, size_t n, uint32_t max_steps, float eps_abs, float eps_rel, float t_max,
                         float* depth, uint32_t* steps, float* hit_pos, void* stream);

/* Host-buffer variant of cth_sphere_trace_dev (synchronous). */
int cth_sphere_trace(cth_ctx* ctx, const cth_shape* shape, const float* origins, int per_ray_origin,
                     const float* dirs, size_t n, uint32_t max_steps, float eps_abs, float eps_rel, float t_max,
                     float* depth, uint32_t* steps, float* hit_pos);

/* ShapeMesh::get_focii (src/mesh/mod.rs:205-243) for ncams cameras (device
 * array): focus_points^2 rays per camera, built on the device from the
 * reference camera exactly as get_focii builds them -- the near-plane
 * bounding box and cell centres (src/camera.rs:90-103,186-189), mapped to
 * world space by the inverse of look_at_rh (src/camera.rs:69-80; taken in
 * closed form, see below), direction normalize(p - position) -- and marched
 * from the camera position: d = DE(pos); pos += dir * d; a hit when d < eps,
 * at most max_steps evaluations (the reference: eps 1e-6, 100 steps, no
 * escape).  Ray r = (cam * fp + x) * fp + y for the cell (x, y) of
 * iter::square (x outer, src/util/iter.rs:55-87).  focii (3 floats per ray)
 * receives the position after the last step (get_focii returns it for hits),
 * depth the distance marched on a hit and +inf on a miss (get_focii drops
 * those rays), steps the evaluations; dirs (nullable, 3 floats per ray) the
 * ray directions.  focus_points in [0, 255] (the reference's u8; 0 gives
 * no rays, as the reference's get_focii returns no points).
 * The view matrix is a rotation plus a translation, so its inverse is taken
 * as the matrix with columns s, u, -f and position; cgmath inverts the
 * general 4x4 matrix by cofactors, which can differ in the entries' last
 * bits.  Because p - position cancels at the reference's near plane (4e-5,
 * src/app.rs:98), such a difference can turn a direction by up to about
 * ulp(|position|) / near_plane radians. */
int cth_get_focii_dev(cth_ctx* ctx, const cth_shape* shape, const cth_ref_camera* cams, uint32_t ncams,
                      uint32_t focus_points, uint32_t max_steps, float eps, float* focii, float* depth,
                      uint32_t* steps, float* dirs, void* stream);

/* ShapeMesh::get_focii for one camera (host buffers, synchronous), with the
 * reference's constants (eps 1e-6, 100 steps): the hit positions in ray
 * order, as the reference's filter_map collects them.  focii must hold
 * focus_points^2 points (3 floats each; may be null when focus_points is
 * 0); *nhits receives their number. */
int cth_get_focii(cth_ctx* ctx, const cth_shape* shape, const cth_ref_camera* cam, uint32_t focus_points,
                  float* focii, size_t* nhits);

/* Primary-ray depth images: ncams cameras (device array) x height x width
 * pixel-centre rays, generated on the device in f32 as
 *   u = ((2 (x + 0.5)) / W - 1) * tan_half_fov * aspect,  v = (1 - (2 (y + 0.5)) / H) * tan_half_fov,
 *   dir = normalize((fwd + right * u) + up * v),  origin = eye,
 * and traced like cth_sphere_trace_dev.  Output index = (cam * H + y) * W + x. */
int cth_render_depth_dev(cth_ctx* ctx, const cth_shape* shape, const cth_camera* cams, uint32_t ncams,
                         uint32_t width, uint32_t height, uint32_t max_steps, float eps_abs, float eps_rel,
                         float t_max, float* depth, uint32_t* steps, void* stream);

/* Differentiable distance: the vector-Jacobian product of min_distance_from
 * (the backward of an autograd function whose forward is
 * cth_min_distance_dev).  The forward iteration is recomputed per lane in
 * parity arithmetic (its distance goes to `d`, nullable), then a reverse sweep
 * through the spherical power map (trig family) or the P8 polynomial
 * (src/shape/mandelbulb.rs:59-79,128-200) gives
 *   grad_p[i] = grad_d[i] * dd/dp   (n x 3, nullable)
 *   grad_power[i] = grad_d[i] * dd/dpower  (nullable; CTH_SHAPE_MANDELBULB_TRIG
 *                                           only, 0 for the other kinds)
 * grad_d NULL means all ones.  The derivative ignores the measure-zero
 * switches of the loop (bailout, iteration cap).  Uses a device workspace of
 * at most ~1 GiB (stream-ordered allocation). */
int cth_de_vjp_dev(cth_ctx* ctx, const cth_shape* shape, const float* pts, const float* powers, size_t n,
                   const float* grad_d, float* d, float* grad_p, float* grad_power, void* stream);

/* cth_de_vjp_dev with the point's two roles apart.  The reference's orbit
 * starts at z_0 = p and adds c = p in every body (src/shape/mandelbulb.rs:
 * 60-74), so dd/dp = dd/dz0 + dd/dc; for fitting the additive constant as a
 * shape parameter the parts are wanted alone:
 *   grad_z0[i] = grad_d[i] * dd/dz0   (n x 3, nullable)
 *   grad_c[i]  = grad_d[i] * dd/dc    (n x 3, nullable; 0 when no loop body
 *                                      reaches the distance)
 * grad_z0 + grad_c, summed in f32, is cth_de_vjp_dev's grad_p bit for bit (up to
 * the sign of a zero where no loop body runs).
 * d, grad_power, grad_d and the workspace as for cth_de_vjp_dev. */
int cth_de_vjp_split_dev(cth_ctx* ctx, const cth_shape* shape, const float* pts, const float* powers, size_t n,
                         const float* grad_d, float* d, float* grad_z0, float* grad_c, float* grad_power,
                         void* stream);

/* cth_de_vjp_dev without any allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid it).  The orbit
 * workspace is a 16 MiB device array of the library, one per GPU, held from
 * the library's load until the process ends; the call allocates and frees
 * nothing.  A call needs n * ((max_iters + 1) * 16 + 16) bytes of it, plus
 * at most 255 bytes of alignment; larger calls fail with CTH_E_INVALID
 * (15 iterations: up to 61,680 points per call).  Results are those of
 * cth_de_vjp_dev, bit for bit.  Every call on one GPU uses the same array, so
 * calls that overlap in time on one GPU (other streams, other contexts,
 * other threads) race and must be ordered by the caller. */
int cth_de_vjp_noalloc_dev(cth_ctx* ctx, const cth_shape* shape, const float* pts, const float* powers, size_t n,
                           const float* grad_d, float* d, float* grad_p, float* grad_power, void* stream);

/* The same product as cth_de_vjp_dev, computed in forward (tangent-linear)
 * mode: the Jacobian of the orbit with respect to p (and the power) is carried
 * through every loop body, so no orbit workspace is needed.  One launch. */
int cth_de_grad_dev(cth_ctx* ctx, const cth_shape* shape, const float* pts, const float* powers, size_t n,
                    const float* grad_d, float* d, float* grad_p, float* grad_power, void* stream);

/* Surface normals from one fused launch (parity forward + tangent Jacobian):
 * normals[i] = grad / |grad| (n x 3); the replacement for the 6-tap difference
 * normal of src/mesh/buffer.rs:256-266.  d (nullable) receives the distances. */
int cth_normals_dev(cth_ctx* ctx, const cth_shape* shape, const float* pts, const float* powers, size_t n, float* d,
                    float* normals, void* stream);

/* The reference's 6-tap central-difference normal (src/mesh/buffer.rs:256-266)
 * with step h on every axis: 7 distance evaluations per point (p and p +- h e_a,
 * formed as p + unit_a * (+-h)); normals = normalize(differences) (nullable),
 * grad_p = differences / (2h) (nullable), d = DE(p) (nullable). */
int cth_normals_fd_dev(cth_ctx* ctx, const cth_shape* shape, const float* pts, const float* powers, size_t n, float h,
                       float* d, float* normals, float* grad_p, void* stream);

/* ---- batched jobs and graphs -------------------------------------------------- */

/* One independent grid job of a batch: an n^3 grid over [start, end] in
 * GridTable order, p = start + (end - start) * (i / divisor), with its own
 * power and iteration cap.  The batch generalises the reference's per-leaf
 * thread-pool jobs (src/mesh/mod.rs:129-161). */
typedef struct cth_job {
    uint32_t n;
    float power;
    uint32_t max_iters;
    float start[3];
    float end[3];
    float divisor;
} cth_job;

typedef struct cth_jobset cth_jobset;

/* Upload a batch (1..1024 jobs, < 2^31 points).  kind CTH_SHAPE_MANDELBULB_TRIG
 * (real powers) or CTH_SHAPE_MANDELBULB (every power 8: the P8 polynomial);
 * bailout is shared. */
int cth_jobs_create(cth_ctx* ctx, int kind, float bailout, const cth_job* jobs, uint32_t njobs, cth_jobset** out);
void cth_jobs_destroy(cth_jobset* jobs);
/* Total points, and the first index of job i in the concatenated outputs. */
size_t cth_jobs_total(const cth_jobset* jobs);
size_t cth_jobs_offset(const cth_jobset* jobs, uint32_t i);

/* Run every job in one persistent launch: dists (total floats), normals
 * (nullable, 3 x total; forward-mode gradient, unit length), done_ticks
 * (device, one per job, zeroed by the call: wall-clock stamp of the job's
 * last write-back of a point that ran at least one loop body; points beyond
 * the bailout radius are written when they are taken and not stamped, so a
 * job made only of such points keeps 0), start_tick (device, one: earliest
 * wave start).  Capture-safe. */
int cth_jobs_run_dev(cth_ctx* ctx, const cth_jobset* jobs, float* dists, float* normals, uint64_t* done_ticks,
                     uint64_t* start_tick, void* stream);

/* Wall-clock counter rate of the device (kHz), for done_ticks. */
int cth_wall_clock_khz(const cth_ctx* ctx);

/* Stream capture into a HIP graph.  Between begin and end on a non-null
 * stream, the capture-safe `_dev` entry points (distance, grid, span
 * sampling, tracing, focus rays, tangent gradients/normals, jobs) only record
 * their launches; the graph then replays them with one launch call.  The
 * graph owns the work counters of the launches recorded on `stream` by this
 * thread (allocated by cth_graph_begin, before the capture starts, and freed
 * by cth_graph_destroy), so replays never share counters with eager launches
 * on other streams.  One capture per thread at a time. */
typedef struct cth_graph cth_graph;
int cth_graph_begin(void* stream);
int cth_graph_end(void* stream, cth_graph** out);
int cth_graph_launch(cth_graph* graph, void* stream);
void cth_graph_destroy(cth_graph* graph);

/* ---- surface nets (mesh extraction) ---------------------------------------------- */

/* Vertex normals of the mesh entry points.
 *   CTH_NORMALS_FD6      the reference's 6-tap central difference with half-step
 *                        0.7 * (e' - s') / res per axis, normalised
 *                        (src/mesh/buffer.rs:256-266): 7 distance evaluations
 *                        per vertex, distance_from_surface = DE(p) (:254).
 *   CTH_NORMALS_TANGENT  the forward-mode gradient of DE at the vertex,
 *                        normalised (one tangent evaluation per vertex; the
 *                        distance is the parity DE(p)).  Mandelbulb kinds only.
 *   CTH_NORMALS_DEFERRED no vertex evaluation yet: each record holds its
 *                        position, in the normal slots its leaf's 6-tap
 *                        half-steps, and a distance of 0, for
 *                        cth_mesh_vertices_dev. */
enum { CTH_NORMALS_FD6 = 0, CTH_NORMALS_TANGENT = 1, CTH_NORMALS_DEFERRED = 2 };

/* The reference's `Vertex` (src/mesh/mod.rs:255-261, #[repr(C)]), 28 bytes. */
typedef struct cth_vertex {
    float position[3];
    float normal[3];
    float distance_from_surface;
} cth_vertex;

/* Surface nets, pass 1, over a batch of leaves whose step-1 corner distances
 * `dists` (nspans x (res+1)^3, GridTable order, as cth_sample_spans_dev writes
 * them) are on the device.  Per leaf: one vertex per cell whose corners have
 * both IEEE signs (src/mesh/buffer.rs:127-147), six indices per corner edge
 * whose end signs differ (:299-363).  vert_offsets / idx_offsets (device,
 * nspans + 1 each) receive the exclusive prefix sums; *nverts / *nidx (host)
 * the totals.  cross_bits (nullable, device) receives the same crossing words
 * as cth_sample_spans_dev.  res: power of two in [1, 64].  Returns after the
 * totals are known (kernels run on `stream`). */
int cth_mesh_count_dev(cth_ctx* ctx, const float* dists, size_t nspans, uint32_t res, uint64_t* vert_offsets,
                       uint64_t* idx_offsets, uint64_t* cross_bits, uint64_t* nverts, uint64_t* nidx, void* stream);

/* Surface nets, pass 2: leaf i's vertices go to vertices[vert_offsets[i] ...]
 * in cube() cell order (x slowest, z fastest; src/mesh/buffer.rs:113), its
 * indices to indices[idx_offsets[i] ...], local to the leaf (0-based, as each
 * leaf's MeshBuffer holds them), in the reference's emission order and
 * winding (:288-372).  Vertex position = centroid of the edge crossings in
 * EDGES order with the reference's distance weights (:149-250).  `nverts` is
 * the total from pass 1.  spans: 6 floats per leaf (the unadjusted span, as
 * given to cth_sample_spans_dev).  Uses a stream-ordered workspace of 28 bytes
 * per vertex for CTH_NORMALS_FD6. */
int cth_mesh_emit_dev(cth_ctx* ctx, const cth_shape* shape, const float* spans, const float* dists, size_t nspans,
                      uint32_t res, int normals, const uint64_t* vert_offsets, const uint64_t* idx_offsets,
                      uint64_t nverts, cth_vertex* vertices, uint32_t* indices, void* stream);

/* Surface nets, the vertex pass on its own: distance_from_surface and normal
 * (CTH_NORMALS_FD6 or CTH_NORMALS_TANGENT, src/mesh/buffer.rs:254-272) of
 * nverts records that cth_mesh_emit_dev wrote with CTH_NORMALS_DEFERRED.
 * cth_mesh_emit_dev with FD6 or TANGENT is the deferred emit followed by this
 * pass on the same stream; split, the pass can be timed or scheduled alone. */
int cth_mesh_vertices_dev(cth_ctx* ctx, const cth_shape* shape, cth_vertex* vertices, uint64_t nverts, int normals,
                          void* stream);

/* MeshBuffer::generate_for_box for one span (host buffers, synchronous):
 * steps 1-3 on the device.  *nverts / *nidx receive the mesh size; vertices /
 * indices may be NULL to query it, and when non-NULL must hold vert_cap /
 * idx_cap entries of at least that size (else CTH_E_INVALID).  Like the
 * reference (src/mesh/buffer.rs:35-39), start < end on every axis and a
 * power-of-two res are required. */
int cth_mesh_span(cth_ctx* ctx, const cth_shape* shape, const cth_span* span, uint32_t res, int normals,
                  cth_vertex* vertices, size_t vert_cap, uint32_t* indices, size_t idx_cap, size_t* nverts,
                  size_t* nidx);

/* MeshBuffer::generate_for_box for a batch of spans (host buffers,
 * synchronous): the mesh jobs that ShapeMesh::update hands its thread pool,
 * one per empty leaf at resolution 64 (src/mesh/mod.rs:129-161), in one pass
 * on the device.  The meshes are concatenated in span order, each span's
 * indices local to its own vertices; voff / ioff (nspans + 1 entries each,
 * nullable) receive the offsets of each span's vertices and indices, and
 * *nverts / *nidx the totals.  vertices / indices may be NULL to query the
 * sizes; when non-NULL they must hold vert_cap / idx_cap entries of at least
 * those sizes (else CTH_E_INVALID).  Each span needs start < end on every
 * axis; res is a power of two in [1, 64]. */
int cth_mesh_spans(cth_ctx* ctx, const cth_shape* shape, const cth_span* spans, size_t nspans, uint32_t res,
                   int normals, cth_vertex* vertices, size_t vert_cap, uint32_t* indices, size_t idx_cap,
                   uint64_t* voff, uint64_t* ioff, size_t* nverts, size_t* nidx);

/* Host-buffer variant of cth_focus_split_dev (synchronous): flags[i] = 1
 * when one of the nfocii focus points lies in leaf i and the eye is nearer
 * to it than twice the leaf's width (src/mesh/mod.rs:91-104). */
int cth_focus_split(cth_ctx* ctx, const cth_span* leaves, size_t n, const cth_point3* focii, size_t nfocii,
                    const float eye[3], uint32_t* flags);

/* Host-buffer variant of cth_sample_spans_dev (synchronous). */
int cth_sample_spans(cth_ctx* ctx, const cth_shape* shape, const cth_span* spans, size_t nspans, uint32_t res,
                     float* dists, uint64_t* cross_bits);

#ifdef __cplusplus
}
#endif

#endif /* CANTUCCI_HIP_H */
