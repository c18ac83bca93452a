// Parity of the C++ host mirror (cantucci_amd/host/cantucci.hpp) against the
// oracle, through the trait methods a caller of the reference uses:
// Shape::{min,max,bounded}_distance_from and their batch forms, contains,
// bounding_box, Mandelbulb::new/classic asserts, MeshBuffer::generate_for_box,
// ShapeMesh::get_focii with Camera::new / Projection::new, and ShapeMesh::new /
// update (split rule, create_spans, the leaves' meshes).
// Needs a gfx950 device.  Prints one line per check; exits non-zero on the
// first failure.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "cantucci.hpp"

// The oracle's C exports (test infrastructure; oracle/oracle.cpp).  The
// shape descriptor has cth_shape's layout.
extern "C" {
void orc_de_points_f32(const cth_shape* shape, const float* pts, const float* powers, size_t n, float* d,
                       uint32_t* iters, float* r_out, float* dr_out, int threads);
void orc_surface_nets(const cth_shape* shape, const float* span, uint32_t res, float* dists_out, float* verts_out,
                      size_t verts_cap, uint32_t* idx_out, size_t idx_cap, size_t* nverts, size_t* nidx);
void orc_focus_dirs(const float* cam, uint32_t fp, float* dirs);
void orc_focus_split(const float* leaves, size_t n, const float* focii, const float* depth, size_t nf,
                     const float* eye, uint32_t* flags);
void orc_create_spans(const float* span, float* out48);
void orc_mesh_batch(const cth_shape* shape, const float* spans, size_t nspans, uint32_t res, const float* dists_in,
                    int threads, uint64_t* voff, uint64_t* ioff, float* verts, size_t vcap, uint32_t* idx,
                    size_t icap);
void orc_sphere_trace(const cth_shape* shape, const float* origins, const float* dirs, size_t n, int origin_stride,
                      uint32_t max_steps, float eps_abs, float eps_rel, float t_max, float* depth, uint32_t* steps,
                      float* hit_pos, int threads);
void orc_de_grad_f64(const cth_shape* shape, const double* pts, const float* powers, size_t n, double* d,
                     double* grad, uint32_t* iters, int threads);
}

namespace {

int failures = 0;

void expect(bool ok, const char* what) {
    std::printf("%s %s\n", ok ? "ok  " : "FAIL", what);
    std::fflush(stdout);
    if (!ok) ++failures;
}

uint32_t bits(float f) {
    uint32_t b;
    std::memcpy(&b, &f, 4);
    return b;
}

int64_t ulp_distance(float a, float b) {
    auto key = [](float f) {
        int32_t i;
        std::memcpy(&i, &f, 4);
        return i < 0 ? (int64_t)INT32_MIN - i : (int64_t)i;
    };
    int64_t d = key(a) - key(b);
    return d < 0 ? -d : d;
}

std::vector<float> oracle_de(const cantucci::Shape& s, const std::vector<cantucci::Point3>& pts) {
    std::vector<float> d(pts.size());
    orc_de_points_f32(&s.desc(), reinterpret_cast<const float*>(pts.data()), nullptr, pts.size(), d.data(), nullptr,
                      nullptr, nullptr, 4);
    return d;
}

template <class F>
bool throws_status(int status, F&& f) {
    try {
        f();
    } catch (const cantucci::Error& e) {
        return e.status == status;
    }
    return false;
}

}  // namespace

int run() {
    using namespace cantucci;
    Context ctx(0);

    // BENCH_POINTS-like near-surface points plus a seeded cloud.
    std::mt19937 rng(7);
    std::uniform_real_distribution<float> U(-1.5f, 1.5f);
    std::vector<Point3> pts(20000);
    for (auto& p : pts) p = {U(rng), U(rng), U(rng)};
    pts.push_back({0.0f, 0.0f, 0.0f});   // origin: NaN with the x86 sign
    pts.push_back({0.0f, -0.0f, 0.5f});  // z axis

    // Mandelbulb<8>::classic(15, 2.0): the P8 polynomial
    auto m8 = Mandelbulb<8>::classic(ctx, 15, 2.0f);
    {
        auto d = m8.batch_min_distance_from(pts);
        auto o = oracle_de(m8, pts);
        size_t same = 0;
        int64_t worst = 0;
        bool nan_ok = true;
        for (size_t i = 0; i < pts.size(); ++i) {
            if (std::isnan(o[i])) {
                nan_ok &= std::isnan(d[i]) && (bits(d[i]) >> 31) == (bits(o[i]) >> 31);
                ++same;
                continue;
            }
            same += bits(d[i]) == bits(o[i]);
            worst = std::max(worst, ulp_distance(d[i], o[i]));
        }
        std::printf("P8: %zu of %zu bit-identical, worst %lld ulp\n", same, pts.size(), (long long)worst);
        expect(same >= pts.size() * 99 / 100 && worst <= 4 && nan_ok, "Mandelbulb<8> batch_min_distance_from");
        expect(bits(m8.min_distance_from(pts[5])) == bits(d[5]), "min_distance_from == batch element");
        auto c = m8.batch_contains(pts);
        bool cok = true;
        for (size_t i = 0; i < pts.size(); ++i) cok &= c[i] == (d[i] < 0.0f);
        expect(cok && m8.contains(pts[5]) == (d[5] < 0.0f), "contains == (d < 0)");
        auto bb = m8.bounding_box();
        expect(bb.start.x == -1.2f && bb.end.z == 1.2f, "bounding_box [-1.2, 1.2]^3");
        expect(!m8.max_distance_from(pts[0]).has_value(), "max_distance_from is None");
        expect(!m8.bounded_distance_from(pts[0]).second.has_value(), "bounded_distance_from upper is None");
        expect(throws_status(CTH_E_UNSUPPORTED, [&] { m8.batch_max_distance_from(pts); }),
               "batch_max_distance_from refuses (the reference panics)");
        expect(throws_status(CTH_E_UNSUPPORTED, [&] { m8.batch_bounded_distance_from(pts); }),
               "batch_bounded_distance_from refuses (the reference panics)");
        expect(throws_status(CTH_E_INVALID, [&] { Mandelbulb<8>::classic(ctx, 0, 2.0f); }),
               "Mandelbulb::new asserts max_iters >= 1");
    }

    // Mandelbulb<3>: the trigonometric map at an integer power
    {
        Mandelbulb<3> m3(ctx, 10, 2.0f);
        auto d = m3.batch_min_distance_from(pts);
        auto o = oracle_de(m3, pts);
        size_t close = 0, n = 0;
        for (size_t i = 0; i < pts.size(); ++i) {
            if (!std::isfinite(o[i])) continue;
            ++n;
            double rel = std::fabs((double)d[i] - o[i]) / std::max(std::fabs((double)o[i]), 1e-30);
            close += rel <= 1e-5;
        }
        std::printf("P3 trig: %zu of %zu within 1e-5 relative\n", close, n);
        expect(close >= n * 97 / 100, "Mandelbulb<3> batch_min_distance_from");
    }

    // Sphere: closed form, upper bound equals the lower one
    {
        Sphere sp(ctx, {0.1f, -0.2f, 0.3f}, 0.7f);
        auto d = sp.batch_min_distance_from(pts);
        auto o = oracle_de(sp, pts);
        bool same = true;
        for (size_t i = 0; i < pts.size(); ++i) same &= bits(d[i]) == bits(o[i]);
        expect(same, "Sphere batch_min_distance_from bit-identical");
        auto mx = sp.batch_max_distance_from(pts);
        auto bd = sp.batch_bounded_distance_from(pts);
        bool ok = true;
        for (size_t i = 0; i < pts.size(); ++i)
            ok &= bits(mx[i]) == bits(d[i]) && bits(bd[i].first) == bits(d[i]) && bits(bd[i].second) == bits(d[i]);
        expect(ok && sp.max_distance_from(pts[3]).value() == d[3], "Sphere max/bounded distance");
        auto bb = sp.bounding_box();
        expect(bb.start.x == 0.1f + -0.7f && bb.end.y == -0.2f + 0.7f, "Sphere bounding_box");
        auto c = sp.batch_contains(pts);
        bool cok = true;
        for (size_t i = 0; i < pts.size(); ++i) {
            float dx = 0.1f - pts[i].x, dy = -0.2f - pts[i].y, dz = 0.3f - pts[i].z;
            cok &= c[i] == (((dx * dx + dy * dy) + dz * dz) <= 0.7f * 0.7f);
        }
        expect(cok, "Sphere contains by squared norm");
    }

    // MeshBuffer::generate_for_box at the reference's resolution (64)
    {
        Sphere sp(ctx, {0.05f, -0.02f, 0.01f}, 0.55f);
        Span box{{-1.0f, -1.0f, -1.0f}, {1.0f, 1.0f, 1.0f}};
        auto [buf, tm] = MeshBuffer::generate_for_box(box, sp, 64);
        const float sp6[6] = {-1.0f, -1.0f, -1.0f, 1.0f, 1.0f, 1.0f};
        size_t nv = 0, ni = 0;
        std::vector<float> dists(65 * 65 * 65);
        orc_surface_nets(&sp.desc(), sp6, 64, dists.data(), nullptr, 0, nullptr, 0, &nv, &ni);
        std::vector<float> v(7 * nv);
        std::vector<uint32_t> ix(ni);
        orc_surface_nets(&sp.desc(), sp6, 64, dists.data(), v.data(), nv, ix.data(), ni, &nv, &ni);
        bool same = buf.vertices.size() == nv && buf.indices == ix && tm.faces * 6 == ni;
        if (same) same = std::memcmp(buf.vertices.data(), v.data(), 28 * nv) == 0;
        std::printf("sphere mesh: %zu vertices, %zu indices\n", nv, ni);
        expect(same && nv > 1000, "MeshBuffer::generate_for_box (sphere) bit-identical");

        auto m = Mandelbulb<8>::classic(ctx, 6, 2.5f);
        Span b2{{-0.6f, -0.6f, 0.7f}, {-0.3f, -0.3f, 1.0f}};
        auto [mb, mt] = MeshBuffer::generate_for_box(b2, m, 64);
        const float m6[6] = {-0.6f, -0.6f, 0.7f, -0.3f, -0.3f, 1.0f};
        orc_surface_nets(&m.desc(), m6, 64, dists.data(), nullptr, 0, nullptr, 0, &nv, &ni);
        std::vector<uint32_t> ix2(ni);
        std::vector<float> v2(7 * nv);
        orc_surface_nets(&m.desc(), m6, 64, dists.data(), v2.data(), nv, ix2.data(), ni, &nv, &ni);
        std::printf("mandelbulb mesh: %zu vertices, %zu indices\n", nv, ni);
        expect(mb.vertices.size() == nv && mb.indices == ix2 && mt.vertices == nv,
               "MeshBuffer::generate_for_box (Mandelbulb<8>) topology identical");
        expect(throws_status(CTH_E_INVALID, [&] { MeshBuffer::generate_for_box(b2, m, 12); }),
               "generate_for_box asserts a power-of-two resolution");
        Span flat{{0.0f, 0.0f, 0.0f}, {0.0f, 1.0f, 1.0f}};
        expect(throws_status(CTH_E_INVALID, [&] { MeshBuffer::generate_for_box(flat, m, 8); }),
               "generate_for_box asserts start < end");
    }

    // ShapeMesh::get_focii with the app's shape and near plane (src/app.rs:98,105)
    {
        auto m = Mandelbulb<8>::classic(ctx, 6, 2.5f);
        Camera cam({-2.2f, 0.4f, 0.3f}, {1.0f, -0.15f, -0.1f}, Projection::make(1.0f, 0.00004f, 10.0f, 1600, 900));
        auto f = get_focii(m, cam, 5);
        const float c10[10] = {cam.position.x,  cam.position.y,  cam.position.z,
                               cam.direction.x, cam.direction.y, cam.direction.z,
                               cam.projection.fov, cam.projection.aspect_ratio,
                               cam.projection.near_plane, cam.projection.far_plane};
        std::vector<float> dirs(75), depth(25), pos(75);
        std::vector<uint32_t> steps(25);
        orc_focus_dirs(c10, 5, dirs.data());
        orc_sphere_trace(&m.desc(), c10, dirs.data(), 25, 0, 100, 1e-6f, 0.0f, INFINITY, depth.data(), steps.data(),
                         pos.data(), 1);
        std::vector<Point3> want;
        for (int i = 0; i < 25; ++i)
            if (depth[i] != INFINITY) want.push_back({pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]});
        bool same = f.size() == want.size() && !f.empty();
        if (same) same = std::memcmp(f.data(), want.data(), 12 * f.size()) == 0;
        std::printf("get_focii: %zu of 25 rays hit\n", f.size());
        expect(same, "get_focii (Mandelbulb<8>::classic(6, 2.5)) bit-identical");
        expect(get_focii(m, cam, 0).empty(), "get_focii with no focus points is empty");
        expect(throws_status(CTH_E_INVALID, [&] { Camera(cam.position, {0.0f, 0.0f, 0.0f}, cam.projection); }),
               "Camera::new asserts a non-zero direction");
        expect(throws_status(CTH_E_INVALID, [&] { Projection::make(1.0f, 0.1f, 10.0f, 0, 900); }),
               "Projection::new asserts a non-zero screen dimension");
    }

    // ShapeMesh::new and two updates, against the oracle's split rule,
    // create_spans and surface nets
    {
        auto m = Mandelbulb<8>::classic(ctx, 6, 2.5f);
        ShapeMesh sm(m);
        expect(sm.leaves().size() == 64, "ShapeMesh::new splits the bounding box twice (64 leaves)");
        Camera cam({-1.6f, 0.2f, 0.1f}, {1.0f, -0.1f, -0.05f}, Projection::make(1.0f, 0.00004f, 10.0f, 1600, 900));
        // the oracle's leaves: the bounding box split twice in create_spans order
        std::vector<float> want;
        {
            const float root[6] = {-1.2f, -1.2f, -1.2f, 1.2f, 1.2f, 1.2f};
            float l1[48], l2[48];
            orc_create_spans(root, l1);
            for (int i = 0; i < 8; ++i) {
                orc_create_spans(l1 + 6 * i, l2);
                want.insert(want.end(), l2, l2 + 48);
            }
        }
        auto same_spans = [&](const std::vector<float>& w) {
            if (w.size() != 6 * sm.leaves().size()) return false;
            for (size_t i = 0; i < sm.leaves().size(); ++i)
                if (std::memcmp(&sm.leaves()[i].span, w.data() + 6 * i, 24) != 0) return false;
            return true;
        };
        expect(same_spans(want), "ShapeMesh::new leaves equal the oracle's create_spans");
        // meshes of leaves against the oracle's surface nets at resolution 64
        // on the device's corner distances (the oracle's own differ by the
        // final ln's last bits): topology and positions bit for bit
        auto meshes_match = [&](const std::vector<size_t>& which) {
            std::vector<float> dists(65 * 65 * 65);
            for (size_t i : which) {
                const auto& leaf = sm.leaves()[i];
                if (!leaf.mesh) return false;
                const float sp6[6] = {leaf.span.start.x, leaf.span.start.y, leaf.span.start.z,
                                      leaf.span.end.x,   leaf.span.end.y,   leaf.span.end.z};
                check(cth_sample_spans(ctx.get(), &m.desc(), reinterpret_cast<const cth_span*>(sp6), 1, 64,
                                       dists.data(), nullptr));
                uint64_t vo[2], io[2];
                orc_mesh_batch(&m.desc(), sp6, 1, 64, dists.data(), 4, vo, io, nullptr, 0, nullptr, 0);
                const size_t nv = vo[1], ni = io[1];
                std::vector<float> v(7 * nv);
                std::vector<uint32_t> ix(ni);
                orc_mesh_batch(&m.desc(), sp6, 1, 64, dists.data(), 4, vo, io, v.data(), nv, ix.data(), ni);
                if (leaf.mesh->vertices.size() != nv || leaf.mesh->indices != ix) return false;
                for (size_t k = 0; k < nv; ++k)
                    if (std::memcmp(leaf.mesh->vertices[k].position, v.data() + 7 * k, 12) != 0) return false;
            }
            return true;
        };
        sm.update(cam);  // nothing is meshed yet, so nothing splits; the 64 leaves are meshed
        std::vector<size_t> all64(64);
        for (size_t i = 0; i < 64; ++i) all64[i] = i;
        expect(same_spans(want) && meshes_match(all64), "update 1: the 64 leaves meshed as the oracle meshes them");
        // the oracle's split rule on the focus points
        const float c10[10] = {cam.position.x,  cam.position.y,  cam.position.z,
                               cam.direction.x, cam.direction.y, cam.direction.z,
                               cam.projection.fov, cam.projection.aspect_ratio,
                               cam.projection.near_plane, cam.projection.far_plane};
        std::vector<float> dirs(75), depth(25), pos(75);
        std::vector<uint32_t> steps(25), flags(64);
        orc_focus_dirs(c10, 5, dirs.data());
        orc_sphere_trace(&m.desc(), c10, dirs.data(), 25, 0, 100, 1e-6f, 0.0f, INFINITY, depth.data(), steps.data(),
                         pos.data(), 1);
        orc_focus_split(want.data(), 64, pos.data(), depth.data(), 25, c10, flags.data());
        std::vector<float> want2;
        std::vector<size_t> children;  // positions of the new leaves
        size_t nsplit = 0;
        for (size_t i = 0; i < 64; ++i) {
            if (flags[i]) {
                float ch[48];
                orc_create_spans(want.data() + 6 * i, ch);
                for (int k = 0; k < 8; ++k) children.push_back(want2.size() / 6 + k);
                want2.insert(want2.end(), ch, ch + 48);
                ++nsplit;
            } else {
                want2.insert(want2.end(), want.begin() + 6 * i, want.begin() + 6 * i + 6);
            }
        }
        sm.update(cam);
        std::printf("ShapeMesh update 2: %zu leaves split, %zu leaves, %u vertices meshed in all\n", nsplit,
                    sm.leaves().size(), sm.batch_timings.vertices);
        expect(nsplit > 0 && same_spans(want2), "update 2: the split leaves equal the oracle's split rule");
        bool all_meshed = true;
        for (const auto& leaf : sm.leaves()) all_meshed &= leaf.mesh.has_value();
        expect(all_meshed && meshes_match(children), "update 2: the new leaves meshed as the oracle meshes them");
    }

    // DistanceOp: forward / backward over cth_de_vjp_batch, against the
    // oracle's FP64 dual-number gradient (shell 0.9 < |p| < 1.3).
    {
        std::mt19937 rg(11);
        std::normal_distribution<float> N(0.0f, 1.0f);
        std::uniform_real_distribution<float> R(0.9f, 1.3f), W(6.0f, 10.0f);
        std::vector<Point3> sp(4000);
        std::vector<float> pw(sp.size());
        for (size_t i = 0; i < sp.size(); ++i) {
            float v[3] = {N(rg), N(rg), N(rg)};
            float m = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]), r = R(rg);
            sp[i] = {v[0] / m * r, v[1] / m * r, v[2] / m * r};
            pw[i] = W(rg);
        }
        std::vector<double> p64(3 * sp.size());
        for (size_t i = 0; i < sp.size(); ++i) {
            p64[3 * i] = sp[i].x;
            p64[3 * i + 1] = sp[i].y;
            p64[3 * i + 2] = sp[i].z;
        }
        auto angle = [](Point3 a, const double* b) {
            double dot = a.x * b[0] + a.y * b[1] + a.z * b[2];
            double na = std::sqrt((double)a.x * a.x + (double)a.y * a.y + (double)a.z * a.z);
            double nb = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
            double c = dot / (na * nb);
            return std::acos(std::max(-1.0, std::min(1.0, c)));
        };
        // the share of points within 1e-3 rad, and the median angle, over
        // points whose f32 and FP64 orbits take the same number of bodies
        auto score = [&](const std::vector<Point3>& g, const std::vector<double>& g64, const std::vector<uint32_t>& it64,
                         const std::vector<uint32_t>& it32, double& within, double& median) {
            std::vector<double> a;
            for (size_t i = 0; i < g.size(); ++i)
                if (it64[i] == it32[i]) a.push_back(angle(g[i], &g64[4 * i]));
            std::sort(a.begin(), a.end());
            within = a.empty() ? 0.0 : (double)(std::upper_bound(a.begin(), a.end(), 1e-3) - a.begin()) / a.size();
            median = a.empty() ? 1.0 : a[a.size() / 2];
        };
        for (int sweep : {CTH_SWEEP_FORWARD, CTH_SWEEP_REVERSE}) {
            const char* name = sweep == CTH_SWEEP_FORWARD ? "forward sweep" : "reverse sweep";
            DistanceOp op(m8, sweep);
            auto d = op.forward(sp);
            auto ref = m8.batch_min_distance_from(sp);
            bool same = true;
            for (size_t i = 0; i < sp.size(); ++i) same &= bits(d[i]) == bits(ref[i]);
            std::vector<float> ones(sp.size(), 1.0f), twos(sp.size(), 2.0f);
            auto g1 = op.backward(ones);
            auto g2 = op.backward(twos);
            bool lin = true;
            for (size_t i = 0; i < sp.size(); ++i)
                lin &= bits(g2.points[i].x) == bits(2.0f * g1.points[i].x) &&
                       bits(g2.points[i].y) == bits(2.0f * g1.points[i].y) &&
                       bits(g2.points[i].z) == bits(2.0f * g1.points[i].z) && g1.powers[i] == 0.0f;
            std::vector<double> d64(sp.size()), g64(4 * sp.size());
            std::vector<uint32_t> it64(sp.size()), it32(sp.size());
            orc_de_grad_f64(&m8.desc(), p64.data(), nullptr, sp.size(), d64.data(), g64.data(), it64.data(), 4);
            orc_de_points_f32(&m8.desc(), reinterpret_cast<const float*>(sp.data()), nullptr, sp.size(), d.data(),
                              it32.data(), nullptr, nullptr, 4);
            double within, median;
            score(g1.points, g64, it64, it32, within, median);
            std::printf("DistanceOp P8 %s: %.4f within 1e-3 rad of the FP64 gradient, median %.2e rad\n", name,
                        within, median);
            expect(same, "DistanceOp::forward == batch_min_distance_from");
            expect(lin, "DistanceOp::backward is linear in grad_d (bit for bit), no power gradient for P8");
            expect(within > 0.98 && median < 1e-5, "DistanceOp::backward vs the FP64 dual gradient (P8)");
        }
        {
            MandelbulbTrig mt(ctx, 15, 2.0f, 8.0f);
            DistanceOp op(mt);
            auto d = op.forward(sp, pw);
            std::vector<float> ones(sp.size(), 1.0f);
            auto g = op.backward(ones);
            std::vector<double> d64(sp.size()), g64(4 * sp.size());
            std::vector<uint32_t> it64(sp.size()), it32(sp.size());
            std::vector<float> d32(sp.size());
            orc_de_grad_f64(&mt.desc(), p64.data(), pw.data(), sp.size(), d64.data(), g64.data(), it64.data(), 4);
            orc_de_points_f32(&mt.desc(), reinterpret_cast<const float*>(sp.data()), pw.data(), sp.size(),
                              d32.data(), it32.data(), nullptr, nullptr, 4);
            double within, median;
            score(g.points, g64, it64, it32, within, median);
            std::vector<double> rel;
            size_t close = 0;
            for (size_t i = 0; i < sp.size(); ++i) {
                if (it64[i] != it32[i]) continue;
                rel.push_back(std::fabs(g.powers[i] - g64[4 * i + 3]) / std::max(std::fabs(g64[4 * i + 3]), 1e-12));
                close += std::fabs((double)d[i] - d64[i]) <= 1e-4 * std::fabs(d64[i]);
            }
            std::sort(rel.begin(), rel.end());
            double rmed = rel.empty() ? 1.0 : rel[rel.size() / 2];
            std::printf("DistanceOp trig, per-point powers: %.4f within 1e-3 rad, median %.2e rad; "
                        "d/dpower median rel %.2e; %zu of %zu distances within 1e-4 of FP64\n",
                        within, median, rmed, close, rel.size());
            expect(within > 0.98 && median < 1e-5 && rmed < 1e-4 && close >= rel.size() * 90 / 100,
                   "DistanceOp with per-point powers vs the FP64 dual gradient (trig)");
            expect(throws_status(CTH_E_INVALID, [&] { op.backward(std::vector<float>(3, 1.0f)); }),
                   "DistanceOp::backward refuses a grad_d of the wrong length");
        }
        {
            Sphere sph(ctx, {0.0f, 0.0f, 0.0f}, 1.0f);
            DistanceOp op(sph);
            op.forward(sp);
            expect(throws_status(CTH_E_UNSUPPORTED, [&] { op.backward(std::vector<float>(sp.size(), 1.0f)); }),
                   "DistanceOp on a sphere refuses (its gradient is the closed form)");
        }
    }

    std::printf("%s: %d failure(s)\n", failures ? "FAILED" : "PASSED", failures);
    return failures ? 1 : 0;
}

int main() {
    try {
        return run();
    } catch (const cantucci::Error& e) {
        // e.g. no gfx950 device: the library refuses to run (no CPU fallback)
        std::printf("cantucci error %d: %s\n", e.status, e.what());
        return 2;
    }
}
