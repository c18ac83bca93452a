// C++17 host mirror of cantucci's `Shape` trait, mesh-buffer entry point and
// focus-point marcher, over the C ABI of libcantucci_hip.so
// (include/cantucci_hip.h).
//
// The reference's interface (src/shape/mod.rs:24-105, src/shape/mandelbulb.rs:12-33,
// src/shape/sphere.rs:6-52, src/mesh/buffer.rs:23-42, src/camera.rs:14-34,108-144,
// src/mesh/mod.rs:205-243) is kept method for method,
// with the same names, argument meaning and results.  Where the reference
// panics (an `assert!`, or `unwrap` on a missing upper bound) this layer throws
// cantucci::Error carrying the library status; nothing aborts.  Every query
// runs on the GPU: there is no CPU fallback, and constructing a Context
// without a usable gfx950 device throws.
//
// Header-only; link with -lcantucci_hip.
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "cantucci_hip.h"

namespace cantucci {

// A failed library call.  `status` is a cth_status.
class Error : public std::runtime_error {
  public:
    Error(int status, const std::string& what) : std::runtime_error(what), status(status) {}
    int status;
};

inline void check(int rc) {
    if (rc != CTH_OK) {
        const char* m = cth_last_error();
        throw Error(rc, m ? m : "cantucci_hip error");
    }
}

// cgmath `Point3<f32>` (12 bytes, the layout of cth_point3).
struct Point3 {
    float x, y, z;
};
static_assert(sizeof(Point3) == sizeof(cth_point3), "Point3 must be 12 bytes");

// octree `Span = Range<Point3<f32>>` (src/octree/mod.rs:13).
struct Span {
    Point3 start, end;
};
static_assert(sizeof(Span) == sizeof(cth_span), "Span must be 24 bytes");

// A library context: one HIP stream on one device.  Calls on one context
// must not overlap in time; use one context per host thread (the reference
// shares one `Arc<dyn Shape>` across its pool threads, src/mesh/mod.rs:61-62).
class Context {
  public:
    explicit Context(int device = 0, int mode = CTH_MODE_PARITY) {
        check(cth_ctx_create(device, &ctx_));
        if (mode != CTH_MODE_PARITY) {
            int rc = cth_ctx_set_mode(ctx_, mode);
            if (rc != CTH_OK) {
                cth_ctx_destroy(ctx_);
                check(rc);
            }
        }
    }
    ~Context() { cth_ctx_destroy(ctx_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    cth_ctx* get() const { return ctx_; }

  private:
    cth_ctx* ctx_ = nullptr;
};

// The `Shape` trait (src/shape/mod.rs:24-105).  Shapes are plain descriptors
// evaluated through a context; they are cheap to copy.
class Shape {
  public:
    virtual ~Shape() = default;

    // fn min_distance_from(&self, p: Point3<f32>) -> f32          (:37)
    float min_distance_from(Point3 p) const { return batch_min_distance_from(std::vector<Point3>{p})[0]; }

    // fn bounding_box(&self) -> Range<Point3<f32>>                (:39)
    Span bounding_box() const {
        cth_span s;
        check(cth_bounding_box(&desc_, &s));
        return Span{{s.start[0], s.start[1], s.start[2]}, {s.end[0], s.end[1], s.end[2]}};
    }

    // fn max_distance_from(&self, _p) -> Option<f32> { None }      (:62-64)
    virtual std::optional<float> max_distance_from(Point3) const { return std::nullopt; }

    // fn bounded_distance_from(&self, p) -> (f32, Option<f32>)     (:73-75)
    std::pair<float, std::optional<float>> bounded_distance_from(Point3 p) const {
        return {min_distance_from(p), max_distance_from(p)};
    }

    // fn contains(&self, p) -> bool                                (:78-80)
    virtual bool contains(Point3 p) const { return min_distance_from(p) < 0.0f; }

    // fn batch_min_distance_from(&self, &[Point3<f32>]) -> Vec<f32> (:89)
    std::vector<float> batch_min_distance_from(const std::vector<Point3>& points) const {
        std::vector<float> out(points.size());
        check(cth_min_distance_batch(ctx_->get(), &desc_, reinterpret_cast<const cth_point3*>(points.data()),
                                     points.size(), out.data()));
        return out;
    }

    // fn batch_max_distance_from(&self, &[Point3<f32>]) -> Vec<f32> (:96)
    // Throws (CTH_E_UNSUPPORTED) where the reference panics: a shape with no
    // upper bound.
    std::vector<float> batch_max_distance_from(const std::vector<Point3>& points) const {
        std::vector<float> out(points.size());
        check(cth_max_distance_batch(ctx_->get(), &desc_, reinterpret_cast<const cth_point3*>(points.data()),
                                     points.size(), out.data()));
        return out;
    }

    // fn batch_bounded_distance_from(&self, ...) -> Vec<(f32, f32)>  (:103)
    std::vector<std::pair<float, float>> batch_bounded_distance_from(const std::vector<Point3>& points) const {
        std::vector<float> flat(2 * points.size());
        check(cth_bounded_distance_batch(ctx_->get(), &desc_, reinterpret_cast<const cth_point3*>(points.data()),
                                         points.size(), flat.data()));
        std::vector<std::pair<float, float>> out(points.size());
        for (size_t i = 0; i < points.size(); ++i) out[i] = {flat[2 * i], flat[2 * i + 1]};
        return out;
    }

    // Batched `contains` (one byte per point on the device, here a bool).
    std::vector<bool> batch_contains(const std::vector<Point3>& points) const {
        std::vector<uint8_t> flags(points.size());
        check(cth_contains_batch(ctx_->get(), &desc_, reinterpret_cast<const cth_point3*>(points.data()),
                                 points.size(), flags.data()));
        return std::vector<bool>(flags.begin(), flags.end());
    }

    const cth_shape& desc() const { return desc_; }
    Context& context() const { return *ctx_; }

  protected:
    Shape(Context& ctx, const cth_shape& desc) : ctx_(&ctx), desc_(desc) {}

    Context* ctx_;
    cth_shape desc_;
};

namespace detail {
inline cth_shape mandel_desc(int kind, float power, uint64_t max_iters, float bailout) {
    // assert!(max_iters >= 1)  (src/shape/mandelbulb.rs:20)
    if (max_iters < 1) throw Error(CTH_E_INVALID, "Mandelbulb: max_iters must be >= 1");
    if (max_iters > UINT32_MAX) throw Error(CTH_E_INVALID, "Mandelbulb: max_iters above 2^32 - 1");
    cth_shape d{};
    d.kind = kind;
    d.power = power;
    d.max_iters = (uint32_t)max_iters;
    d.bailout = bailout;
    return d;
}
}  // namespace detail

// `Mandelbulb<const P: u8>` (src/shape/mandelbulb.rs:12-33): P == 8 runs the
// reference's closed-form polynomial, any other P its trigonometric map.
template <uint8_t P>
class Mandelbulb : public Shape {
  public:
    // Mandelbulb::new(max_iters, bailout)  (:19-26)
    Mandelbulb(Context& ctx, uint64_t max_iters, float bailout)
        : Shape(ctx, detail::mandel_desc(CTH_SHAPE_MANDELBULB, (float)P, max_iters, bailout)) {}

    // Mandelbulb::<8>::classic(max_iters, bailout)  (:28-33)
    static Mandelbulb classic(Context& ctx, uint64_t max_iters, float bailout) {
        static_assert(P == 8, "classic() exists for the power-8 Mandelbulb only");
        return Mandelbulb(ctx, max_iters, bailout);
    }
};

// The trigonometric Mandelbulb with a real power (powf): the shape-parameter
// extension of `rotate_inner_px_generic` (src/shape/mandelbulb.rs:128-146).
class MandelbulbTrig : public Shape {
  public:
    MandelbulbTrig(Context& ctx, uint64_t max_iters, float bailout, float power)
        : Shape(ctx, detail::mandel_desc(CTH_SHAPE_MANDELBULB_TRIG, power, max_iters, bailout)) {}
};

// `Sphere` (src/shape/sphere.rs:6-52).
class Sphere : public Shape {
  public:
    Sphere(Context& ctx, Point3 center, float radius) : Shape(ctx, sphere_desc(center, radius)) {}

    // fn max_distance_from(&self, p) -> Option<f32> { Some(min) }  (:37-39)
    std::optional<float> max_distance_from(Point3 p) const override { return min_distance_from(p); }

    // fn contains: |c - p|^2 <= r^2  (:23-26)
    bool contains(Point3 p) const override { return batch_contains(std::vector<Point3>{p})[0]; }

  private:
    static cth_shape sphere_desc(Point3 c, float r) {
        cth_shape d{};
        d.kind = CTH_SHAPE_SPHERE;
        d.max_iters = 1;
        d.center[0] = c.x;
        d.center[1] = c.y;
        d.center[2] = c.z;
        d.radius = r;
        return d;
    }
};

// A differentiable distance op with the contract of torch.autograd.Function:
// forward(points[, powers]) returns Shape::batch_min_distance_from and keeps
// its inputs (save_for_backward); backward(grad_d) returns the gradients of
// sum(grad_d * d) with respect to the points and the per-point powers.  The
// reference has no derivative; this differentiates its min_distance_from
// (src/shape/mandelbulb.rs:59-79).  The sweep is the forward-mode Jacobian
// (the faster one) or the reverse sweep; both are the device kernels'.  The
// op refers to its shape, which must outlive it.
class DistanceOp {
  public:
    struct Grads {
        std::vector<Point3> points;  // dd/dp * grad_d
        std::vector<float> powers;   // dd/dpower * grad_d (0 except for MandelbulbTrig)
    };

    explicit DistanceOp(const Shape& shape, int sweep = CTH_SWEEP_FORWARD) : shape_(shape), sweep_(sweep) {}

    std::vector<float> forward(const std::vector<Point3>& points, const std::vector<float>& powers = {}) {
        if (!powers.empty() && powers.size() != points.size())
            throw Error(CTH_E_INVALID, "DistanceOp: one power per point");
        points_ = points;
        powers_ = powers;
        if (powers.empty()) return shape_.batch_min_distance_from(points);
        // per-point powers: the distances of the gradient kernel, which equal
        // the distance kernel's bit for bit
        std::vector<float> d(points.size());
        check(cth_de_vjp_batch(shape_.context().get(), &shape_.desc(),
                               reinterpret_cast<const cth_point3*>(points.data()), powers.data(), points.size(),
                               nullptr, d.data(), nullptr, nullptr, sweep_));
        return d;
    }

    Grads backward(const std::vector<float>& grad_d) const {
        if (grad_d.size() != points_.size()) throw Error(CTH_E_INVALID, "DistanceOp: one grad_d per point");
        Grads g{std::vector<Point3>(points_.size()), std::vector<float>(points_.size())};
        check(cth_de_vjp_batch(shape_.context().get(), &shape_.desc(),
                               reinterpret_cast<const cth_point3*>(points_.data()),
                               powers_.empty() ? nullptr : powers_.data(), points_.size(), grad_d.data(), nullptr,
                               reinterpret_cast<float*>(g.points.data()), g.powers.data(), sweep_));
        return g;
    }

  private:
    const Shape& shape_;
    int sweep_;
    std::vector<Point3> points_;
    std::vector<float> powers_;
};

// `Vertex` (src/mesh/mod.rs:255-261).
struct Vertex {
    float position[3];
    float normal[3];
    float distance_from_surface;
};
static_assert(sizeof(Vertex) == sizeof(cth_vertex), "Vertex must be 28 bytes");

// The counts of `Timings` (src/mesh/buffer.rs:375-381).
struct Timings {
    uint32_t vertices;
    uint32_t faces;
};

// `MeshBuffer` (src/mesh/buffer.rs:23-27).
struct MeshBuffer {
    std::vector<Vertex> vertices;
    std::vector<uint32_t> indices;

    // MeshBuffer::generate_for_box(span, shape, resolution)  (:30-42):
    // naive surface nets on the device.  `normals` selects the reference's
    // 6-tap difference (CTH_NORMALS_FD6) or the forward-mode gradient
    // (CTH_NORMALS_TANGENT).  Throws where the reference asserts.
    static std::pair<MeshBuffer, Timings> generate_for_box(const Span& span, const Shape& shape, uint32_t resolution,
                                                           int normals = CTH_NORMALS_FD6) {
        cth_span s;
        s.start[0] = span.start.x;
        s.start[1] = span.start.y;
        s.start[2] = span.start.z;
        s.end[0] = span.end.x;
        s.end[1] = span.end.y;
        s.end[2] = span.end.z;
        size_t nv = 0, ni = 0;
        cth_ctx* c = shape.context().get();
        check(cth_mesh_span(c, &shape.desc(), &s, resolution, normals, nullptr, 0, nullptr, 0, &nv, &ni));
        MeshBuffer buf;
        buf.vertices.resize(nv);
        buf.indices.resize(ni);
        if (nv || ni)
            check(cth_mesh_span(c, &shape.desc(), &s, resolution, normals,
                                reinterpret_cast<cth_vertex*>(buf.vertices.data()), nv, buf.indices.data(), ni, &nv,
                                &ni));
        return {std::move(buf), Timings{(uint32_t)nv, (uint32_t)(ni / 6)}};
    }
};

// cgmath `Vector3<f32>`.
struct Vector3 {
    float x, y, z;
};

// `Projection` (src/camera.rs:108-127): vertical field of view (radians),
// aspect ratio, near and far plane.
struct Projection {
    float fov;
    float aspect_ratio;
    float near_plane;
    float far_plane;

    // Projection::new(fov, near..far, (w, h))  (:135-144): aspect w / h;
    // throws where the reference asserts a non-zero screen dimension.
    static Projection make(float fov, float near_plane, float far_plane, uint32_t w, uint32_t h) {
        if (w == 0 || h == 0) throw Error(CTH_E_INVALID, "screen dimension must not be zero");
        return Projection{fov, (float)w / (float)h, near_plane, far_plane};
    }
};

// `Camera` (src/camera.rs:14-34): position, unit view direction and
// projection; the up vector is always +z.
class Camera {
  public:
    // Camera::new(pos, dir, proj)  (:26-34): throws where the reference
    // asserts dir != 0, and keeps dir.normalize() = dir * (1 / |dir|).
    Camera(Point3 position, Vector3 direction, Projection projection) : position(position), projection(projection) {
        if (direction.x == 0.0f && direction.y == 0.0f && direction.z == 0.0f)
            throw Error(CTH_E_INVALID, "camera direction must not be zero");
        const float m = std::sqrt((direction.x * direction.x + direction.y * direction.y) + direction.z * direction.z);
        const float inv = 1.0f / m;
        this->direction = {direction.x * inv, direction.y * inv, direction.z * inv};
    }
    Point3 position;
    Vector3 direction;
    Projection projection;
};

// `ShapeMesh::get_focii(&self, camera, focus_points)` (src/mesh/mod.rs:205-243)
// for the mesh's shape: the points where the focus_points^2 rays through
// the near-plane cell centres hit it (100 steps, a hit when the distance is
// below 1e-6), in ray order; the rays are built and marched on the device.
inline std::vector<Point3> get_focii(const Shape& shape, const Camera& camera, uint8_t focus_points) {
    cth_ref_camera c;
    c.position[0] = camera.position.x;
    c.position[1] = camera.position.y;
    c.position[2] = camera.position.z;
    c.direction[0] = camera.direction.x;
    c.direction[1] = camera.direction.y;
    c.direction[2] = camera.direction.z;
    c.fov = camera.projection.fov;
    c.aspect = camera.projection.aspect_ratio;
    c.near_plane = camera.projection.near_plane;
    c.far_plane = camera.projection.far_plane;
    std::vector<Point3> out((size_t)focus_points * focus_points);
    size_t hits = 0;
    check(cth_get_focii(shape.context().get(), &shape.desc(), &c, focus_points,
                        out.empty() ? nullptr : reinterpret_cast<float*>(out.data()), &hits));
    out.resize(hits);
    return out;
}

// `create_spans` (src/octree/mod.rs:315-329): the 8 children of a span,
// child bit 2 = upper x half, bit 1 = y, bit 0 = z, around the centre
// start + (end - start) / 2 (SpanExt::center, :21-23), so that the children
// tile the parent without gaps.
inline std::vector<Span> create_spans(const Span& p) {
    const Point3 c{p.start.x + (p.end.x - p.start.x) / 2.0f, p.start.y + (p.end.y - p.start.y) / 2.0f,
                   p.start.z + (p.end.z - p.start.z) / 2.0f};
    std::vector<Span> out(8);
    for (int i = 0; i < 8; ++i) {
        const bool hx = i & 4, hy = i & 2, hz = i & 1;
        out[i].start = {hx ? c.x : p.start.x, hy ? c.y : p.start.y, hz ? c.z : p.start.z};
        out[i].end = {hx ? p.end.x : c.x, hy ? p.end.y : c.y, hz ? p.end.z : c.z};
    }
    return out;
}

// `ShapeMesh` (src/mesh/mod.rs:22-178), the mesh manager of the reference's
// app, without its rendering half (the wgpu pipeline and per-leaf views):
// the leaves of an octree over the shape's bounding box, each holding the
// MeshBuffer of its span at resolution 64 once meshed.  The reference meshes
// the empty leaves on a CPU thread pool and picks the buffers up in later
// frames; update() meshes all of them in one batch on the device
// (cth_mesh_spans) before it returns.  Leaves are kept in depth-first order,
// a split leaf replaced by its 8 children in create_spans order.  The
// ShapeMesh keeps its own copy of the shape's descriptor (the reference
// holds an Arc<dyn Shape>); the shape's Context must outlive it.
class ShapeMesh {
  public:
    static constexpr uint32_t RESOLUTION = 64;  // (:131)
    static constexpr uint8_t FOCUS_POINTS = 5;  // (:86)

    struct Leaf {
        Span span;
        std::optional<MeshBuffer> mesh;  // MeshStatus::Ready; empty until meshed
    };

    // ShapeMesh::new (:45-74): the bounding box, split twice: 64 leaves.
    explicit ShapeMesh(const Shape& shape) : shape_(shape) {
        for (const Span& child : create_spans(shape.bounding_box()))
            for (const Span& leaf : create_spans(child)) leaves_.push_back(Leaf{leaf, std::nullopt});
    }

    // ShapeMesh::update (:82-170): every meshed leaf that holds one of the
    // camera's focus points, with the camera nearer to that point than twice
    // the leaf's width, is split (its mesh dropped); then every empty leaf is
    // meshed.  The split test runs on the device (cth_focus_split).
    void update(const Camera& camera) {
        const std::vector<Point3> focii = get_focii(camera, FOCUS_POINTS);
        std::vector<cth_span> spans(leaves_.size());
        for (size_t i = 0; i < leaves_.size(); ++i) spans[i] = to_cth(leaves_[i].span);
        std::vector<uint32_t> flags(leaves_.size(), 0u);
        const float eye[3] = {camera.position.x, camera.position.y, camera.position.z};
        check(cth_focus_split(shape_.context().get(), spans.data(), spans.size(),
                              reinterpret_cast<const cth_point3*>(focii.data()), focii.size(), eye, flags.data()));
        std::vector<Leaf> next;
        for (size_t i = 0; i < leaves_.size(); ++i) {
            if (flags[i] && leaves_[i].mesh) {  // only leaves showing a mesh are split (:94)
                for (const Span& child : create_spans(leaves_[i].span)) next.push_back(Leaf{child, std::nullopt});
            } else {
                next.push_back(std::move(leaves_[i]));
            }
        }
        leaves_ = std::move(next);
        mesh_empty_leaves();
    }

    // ShapeMesh::get_focii (:205-243).
    std::vector<Point3> get_focii(const Camera& camera, uint8_t focus_points) const {
        return cantucci::get_focii(shape_, camera, focus_points);
    }

    const std::vector<Leaf>& leaves() const { return leaves_; }

    // Totals of the meshes made so far (the reference's batch_timings counts, :38-41).
    Timings batch_timings{0, 0};

  private:
    static cth_span to_cth(const Span& s) {
        cth_span c;
        c.start[0] = s.start.x;
        c.start[1] = s.start.y;
        c.start[2] = s.start.z;
        c.end[0] = s.end.x;
        c.end[1] = s.end.y;
        c.end[2] = s.end.z;
        return c;
    }

    void mesh_empty_leaves() {
        std::vector<size_t> todo;
        std::vector<cth_span> spans;
        for (size_t i = 0; i < leaves_.size(); ++i)
            if (!leaves_[i].mesh) {
                todo.push_back(i);
                spans.push_back(to_cth(leaves_[i].span));
            }
        if (todo.empty()) return;
        cth_ctx* c = shape_.context().get();
        size_t nv = 0, ni = 0;
        check(cth_mesh_spans(c, &shape_.desc(), spans.data(), spans.size(), RESOLUTION, CTH_NORMALS_FD6, nullptr, 0,
                             nullptr, 0, nullptr, nullptr, &nv, &ni));
        std::vector<Vertex> verts(nv);
        std::vector<uint32_t> idx(ni);
        std::vector<uint64_t> voff(spans.size() + 1), ioff(spans.size() + 1);
        check(cth_mesh_spans(c, &shape_.desc(), spans.data(), spans.size(), RESOLUTION, CTH_NORMALS_FD6,
                             reinterpret_cast<cth_vertex*>(verts.data()), nv, idx.data(), ni, voff.data(), ioff.data(),
                             &nv, &ni));
        for (size_t k = 0; k < todo.size(); ++k) {
            MeshBuffer buf;
            buf.vertices.assign(verts.begin() + voff[k], verts.begin() + voff[k + 1]);
            buf.indices.assign(idx.begin() + ioff[k], idx.begin() + ioff[k + 1]);
            batch_timings.vertices += (uint32_t)buf.vertices.size();
            batch_timings.faces += (uint32_t)(buf.indices.size() / 6);
            leaves_[todo[k]].mesh = std::move(buf);
        }
    }

    Shape shape_;  // the descriptor and context of the caller's shape
    std::vector<Leaf> leaves_;
};

}  // namespace cantucci
