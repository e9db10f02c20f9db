// sanitize/driver.cpp — test infrastructure: drives the oracle's C ABI under
// AddressSanitizer and UndefinedBehaviorSanitizer (built by
// build_lib.build_oracle_asan, run by tests/test_oracle_asan.py).  Every
// constraint family, the three collider shapes, attachments, the press,
// reset/launch/rigid velocity, deformed shapes, self-collision within and
// across objects, both Gauss-Seidel colourings, the volume-cluster rules
// (k-ring with farthest-point sampling and greedy cover, explicit clusters),
// shape clusters and the argument-checking error paths, on small scenes.
// Exits non-zero on the first unexpected return code; the sanitizers abort on
// the first memory or undefined-behaviour error.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "xpbd_capi.h"

namespace {

int g_fail = 0;

void expect(bool ok, const char* what)
{
    if (!ok) {
        std::fprintf(stderr, "unexpected result: %s (%s)\n", what, xpbd_last_error());
        g_fail = 1;
    }
}

// Jittered lattice of nx*ny*nz points, spacing h, base offset; edges to the
// six axis neighbours and one face diagonal (triangles for the angle pass).
struct Scene {
    std::vector<float> pos, scales, rots, opacity, rest, stiff;
    std::vector<int> edges;
    int n = 0;
};

void add_lattice(Scene& s, int nx, int ny, int nz, float h, float ox, float oy, float oz, unsigned seed)
{
    const int base = s.n;
    unsigned st = seed;
    auto rnd = [&]() {
        st = st * 1664525u + 1013904223u;
        return ((st >> 8) & 0xffff) / 65536.0f - 0.5f;
    };
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                s.pos.push_back(ox + h * (x + 0.2f * rnd()));
                s.pos.push_back(oy + h * (y + 0.2f * rnd()));
                s.pos.push_back(oz + h * (z + 0.2f * rnd()));
                for (int k = 0; k < 3; ++k) s.scales.push_back(0.3f * h * (1.0f + 0.2f * rnd()));
                const float q[4] = {1.0f + 0.1f * rnd(), 0.1f * rnd(), 0.1f * rnd(), 0.1f * rnd()};
                const float l = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
                for (float v : q) s.rots.push_back(v / l);
                s.opacity.push_back(0.9f);
            }
    auto id = [&](int x, int y, int z) { return base + (z * ny + y) * nx + x; };
    auto edge = [&](int a, int b) {
        s.edges.push_back(a);
        s.edges.push_back(b);
        const float dx = s.pos[3 * a] - s.pos[3 * b], dy = s.pos[3 * a + 1] - s.pos[3 * b + 1],
                    dz = s.pos[3 * a + 2] - s.pos[3 * b + 2];
        s.rest.push_back(std::sqrt(dx * dx + dy * dy + dz * dz));
        s.stiff.push_back(1.0f);
    };
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                if (x + 1 < nx) edge(id(x, y, z), id(x + 1, y, z));
                if (y + 1 < ny) edge(id(x, y, z), id(x, y + 1, z));
                if (z + 1 < nz) edge(id(x, y, z), id(x, y, z + 1));
                if (x + 1 < nx && y + 1 < ny) edge(id(x, y, z), id(x + 1, y + 1, z));
            }
    s.n += nx * ny * nz;
}

int create(const Scene& s)
{
    return xpbd_create(s.n, s.pos.data(), s.scales.data(), s.rots.data(), s.opacity.data(), (int)s.rest.size(),
                       s.edges.data(), s.rest.data(), s.stiff.data());
}

void colliders(float t)
{
    const int types[3] = {0, 1, 2};
    float poses[36] = {0};
    for (int c = 0; c < 3; ++c) {
        float* P = poses + 12 * c;
        const float a = 0.3f * (c + 1) + t;
        P[0] = std::cos(a); P[1] = -std::sin(a); P[3] = std::sin(a); P[4] = std::cos(a); P[8] = 1.0f;
        P[9] = 0.02f * c + 0.01f * t; P[10] = 0.03f; P[11] = 0.02f + 0.01f * c;
    }
    const float dims[9] = {0.02f, 0.0f, 0.0f, 0.015f, 0.01f, 0.02f, 0.01f, 0.02f, 0.0f};
    xpbd_set_colliders(3, types, poses, dims, 0.001f, 0.3f);
}

void full_scene()
{
    Scene s;
    add_lattice(s, 9, 7, 6, 0.01f, 0.0f, 0.0f, 0.005f, 7);
    expect(create(s) == 0, "create");
    xpbd_set_solver(4, 1.0f / 600, 0.6f, 0.01f, 1.0f, 1.0f, 1e-6f, 1e-5f, 0.5f, 1e-4f, 0.5f);
    xpbd_set_constraints(1, 1, 1, 1);
    xpbd_set_volume(1e-6f, 3, 24, 2);   // farthest-point sampling, greedy cover
    xpbd_set_ground(1, 0.0f, 0.0f, 1.0f, 0.0f, 0.5f, 0.1f, 0.0f, 0.0f, 9.81f);
    xpbd_set_object_shape(0.3f);
    xpbd_set_self_collision(1, 1.5f, 2.0f, 1);
    xpbd_set_shape_robust(1);
    expect(xpbd_set_contact_shape(1, 0.2f, 0.01f) == 0, "contact shape");
    colliders(0.0f);
    std::vector<float> out(3 * s.n), scales(3 * s.n), quats(4 * s.n), vel(3 * s.n);
    for (int f = 0; f < 6; ++f) {
        colliders(0.01f * f);
        for (int k = 0; k < 5; ++k) expect(xpbd_step() == 0, "step");
        if (f == 1) {
            const int idx[4] = {0, 5, 17, 40};
            const float tgt[12] = {0, 0, 0.05f, 0.05f, 0, 0.05f, 0.02f, 0.02f, 0.06f, 0.04f, 0.04f, 0.07f};
            expect(xpbd_set_attached(4, idx, tgt, 1.0f) == 0, "attach");
        }
        if (f == 3) expect(xpbd_set_attached(0, nullptr, nullptr, 0.0f) == 0, "detach");
        expect(xpbd_compute_shapes(1e-3f, scales.data(), quats.data(), s.n) >= 0, "compute shapes");
    }
    expect(xpbd_positions_host(out.data(), s.n) == 0, "positions");
    int hits[32];
    float push[96], cen[96];
    expect(xpbd_collider_stats(hits, push, cen, 32) >= 0, "collider stats");
    int sc[3];
    xpbd_self_collision_stats(sc);
    float jvw = 0.0f, jstd = 0.0f;
    xpbd_volume_stats(1, &jvw, &jstd);
    std::printf("full scene: %d nodes, self-collision candidates %d active %d overflow %d, J %.4f\n", s.n, sc[0],
                sc[1], sc[2], jvw);
    const float dv[3] = {0.1f, 0.0f, 0.0f}, dw[3] = {0.0f, 0.0f, 1.0f}, c[3] = {0.04f, 0.03f, 0.03f};
    expect(xpbd_add_rigid_velocity(dv, dw, c) == 0, "rigid velocity");
    expect(xpbd_step() == 0, "step after rigid velocity");
    // oracle extensions: state, lists, topology, colourings, clusters
    expect(xpbd_ext_get_state(out.data(), vel.data(), s.n) == 0, "get state");
    expect(xpbd_ext_set_state(out.data(), vel.data(), s.n) == 0, "set state");
    std::vector<unsigned int> keys(s.n);
    std::vector<int> sorted(s.n), counts(s.n), lists(32 * (size_t)s.n);
    expect(xpbd_ext_self_collision_lists(keys.data(), sorted.data(), counts.data(), lists.data(), s.n) == 0,
           "self-collision lists");
    std::vector<int> lead(s.n), mc(s.n), rc(s.n), mt(s.n);
    std::vector<float> dr(s.n);
    expect(xpbd_ext_volume_topology(lead.data(), mc.data(), rc.data(), mt.data(), dr.data(), s.n) == 0, "topology");
    std::vector<int> mem(64 * (size_t)s.n);
    expect(xpbd_ext_volume_members(mem.data(), (int)mem.size()) >= 0, "members");
    expect(xpbd_ext_set_gauss_seidel(1, 1) == 0, "gauss-seidel");
    for (int k = 0; k < 3; ++k) expect(xpbd_step() == 0, "gauss-seidel step");
    int gsInfo[8];
    expect(xpbd_ext_gs_info(gsInfo) == 0, "gs info");
    expect(xpbd_ext_set_gauss_seidel(0, 0) == 0, "jacobi again");
    std::vector<int> clusterOf(s.n);
    for (int i = 0; i < s.n; ++i) clusterOf[i] = i / 24;
    expect(xpbd_ext_set_shape_clusters(clusterOf.data(), s.n) == 0, "shape clusters");
    expect(xpbd_ext_set_self_collision_grid(0.012f, 1024) == 0, "grid override");
    for (int k = 0; k < 3; ++k) expect(xpbd_step() == 0, "step with clusters and a small table");
    // press cycle, reset, launch
    expect(xpbd_press_start(1, 0.15f, 0.3f, 0) == 0, "press start");
    for (int k = 0; k < 10; ++k) expect(xpbd_step() == 0, "press step");
    float lo, hi, cur, mx;
    expect(xpbd_press_state(&lo, &hi, &cur, &mx) == 1, "press state");
    xpbd_press_stop();
    expect(xpbd_reset() == 0, "reset");
    const float lin[3] = {0.0f, 0.0f, 1.0f}, ang[3] = {0.0f, 1.0f, 0.0f}, tilt[3] = {0.1f, 0.0f, 0.0f};
    expect(xpbd_launch(lin, ang, tilt) == 0, "launch");
    for (int k = 0; k < 5; ++k) expect(xpbd_step() == 0, "step after launch");
    xpbd_destroy();
}

void two_objects_explicit_clusters()
{
    Scene s;
    add_lattice(s, 6, 6, 4, 0.01f, 0.0f, 0.0f, 0.005f, 11);
    add_lattice(s, 6, 6, 4, 0.01f, 0.01f, 0.005f, 0.03f, 13);
    expect(create(s) == 0, "create two objects");
    xpbd_set_solver(2, 1.0f / 600, 0.6f, 0.01f, 1.0f, 1.0f, 1e-6f, 1e-5f, 0.5f, 1e-4f, 0.5f);
    xpbd_set_constraints(1, 0, 0, 1);
    xpbd_set_volume(1e-6f, 1, 64, 1);
    xpbd_set_ground(1, 0.0f, 0.0f, 1.0f, 0.0f, 0.5f, 0.0f, 0.0f, 0.0f, 9.81f);
    xpbd_set_self_collision(1, 1.5f, 2.0f, 0);
    xpbd_set_object_shape(0.2f);
    for (int k = 0; k < 10; ++k) expect(xpbd_step() == 0, "step (k = 1 clusters)");
    // node + 3 next nodes as explicit clusters
    std::vector<int> off(s.n + 1), memb;
    for (int i = 0; i < s.n; ++i) {
        off[i] = (int)memb.size();
        for (int d = 1; d <= 3; ++d) memb.push_back((i + d) % s.n);
    }
    off[s.n] = (int)memb.size();
    expect(xpbd_ext_set_volume_clusters(off.data(), memb.data(), 4) == 0, "explicit clusters");
    for (int k = 0; k < 10; ++k) expect(xpbd_step() == 0, "step (explicit clusters)");
    xpbd_destroy();
}

void error_paths()
{
    expect(xpbd_step() != 0, "step before create fails");
    float p[3] = {0, 0, 0};
    expect(xpbd_create(0, p, p, p, p, 0, nullptr, nullptr, nullptr) != 0, "create with n = 0 fails");
    expect(xpbd_create(1, nullptr, p, p, p, 0, nullptr, nullptr, nullptr) != 0, "create with null positions fails");
    Scene s;
    add_lattice(s, 3, 3, 2, 0.01f, 0.0f, 0.0f, 0.0f, 3);
    expect(create(s) == 0, "create small");
    std::vector<float> out(3 * s.n);
    const int idx[1] = {0};
    expect(xpbd_set_attached(1, idx, p, 1.0f) != 0, "attach before the first step fails");
    expect(xpbd_positions_host(out.data(), s.n) != 0, "positions before the first step fail");
    expect(xpbd_step() == 0, "step small");
    expect(xpbd_positions_host(out.data(), s.n + 1) != 0, "positions with the wrong size fail");
    const int bad[2] = {s.n + 5, -1};   // skipped, as in the reference
    const float bp[6] = {0, 0, 0, 0, 0, 0};
    expect(xpbd_set_attached(2, bad, bp, 1.0f) == 0, "attach out of range is skipped");
    expect(xpbd_step() == 0, "step with skipped attachments");
    xpbd_set_solver(-3, -1.0f, 5.0f, 2.0f, -1.0f, -1.0f, -1.0f, -1.0f, 9.0f, -1.0f, 9.0f);   // clamped
    xpbd_set_volume(-1.0f, -2, -5, -7);
    expect(xpbd_step() == 0, "step with clamped settings");
    xpbd_destroy();
}

}  // namespace

int main()
{
    error_paths();
    full_scene();
    two_objects_explicit_clusters();
    if (g_fail) return 1;
    std::printf("oracle sanitizer run: ok\n");
    return 0;
}
