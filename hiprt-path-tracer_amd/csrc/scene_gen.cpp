// scene_gen.cpp -- seeded procedural scenes standing in for the ASSIMP
// SceneParser (src/Scene/SceneParser.h:109-170), which is out of scope.
//
// Scene randomness is PCG32 (O'Neill 2014, pcg32 XSH-RR 64/32) with
// state seed 0x2026120200000000 + config and stream = config. Geometry is
// built in double and stored as binary32; the arrays are the exact bits both
// the GPU and the CPU oracle consume.
//
//   config 1: Cornell box [-1,1]^3 of 12 quads (24 tris), 2 emissive ceiling
//             quads (4 tris), 20 non-overlapping icospheres (subdivision 4,
//             5120 tris each). 102,428 triangles. Lambertian (specular 0,
//             metallic 0, albedo U[0.1,0.9]^3) or GGX metal (metallic 1,
//             roughness U[0.05,1], base colour U[0.1,0.9]^3) by coin flip.
//   config 2: 512x512-quad terrain (524,288 tris) displaced by 6-octave fBm
//             of value noise, 400 icospheres (subdivision 3, 1280 tris) with
//             random scale/rotation/translation, 4096 emissive triangles of
//             power log-uniform over [1, 1000]. 1,040,384 triangles. Principled
//             parameters uniform per object.
//   config 5: 8 fBm grids (320x320 quads each), 3000 instances of
//             icospheres of 1280 / 5120 / 20480 tris, 16,384 emitters; every
//             Principled layer (coat, thin film, dispersion) per object.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <new>
#include <vector>

#include "../../include/mipt.h"

namespace mipt {
namespace {

struct Pcg32 {
    uint64_t state = 0, inc = 0;
    Pcg32(uint64_t initstate, uint64_t initseq)
    {
        state = 0u;
        inc = (initseq << 1u) | 1u;
        next();
        state += initstate;
        next();
    }
    uint32_t next()
    {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
    }
    // Uniform double in [0, 1) from 32 bits.
    double uniform() { return next() * (1.0 / 4294967296.0); }
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
};

struct V3 {
    double x, y, z;
};
V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
V3 normalize(V3 a) { return (1.0 / std::sqrt(dot(a, a))) * a; }

struct Mesh {
    std::vector<V3> v;
    std::vector<int32_t> tri; // 3 per triangle
};

Mesh icosphere(int subdiv)
{
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    Mesh m;
    m.v = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
           {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (auto& p : m.v) p = normalize(p);
    m.tri = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
             3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    for (int s = 0; s < subdiv; s++)
    {
        std::map<std::pair<int32_t, int32_t>, int32_t> mid;
        auto midpoint = [&](int32_t a, int32_t b) {
            auto key = a < b ? std::make_pair(a, b) : std::make_pair(b, a);
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            int32_t id = (int32_t)m.v.size();
            m.v.push_back(normalize(0.5 * (m.v[a] + m.v[b])));
            mid[key] = id;
            return id;
        };
        std::vector<int32_t> nt;
        nt.reserve(m.tri.size() * 4);
        for (size_t f = 0; f < m.tri.size(); f += 3)
        {
            int32_t a = m.tri[f], b = m.tri[f + 1], c = m.tri[f + 2];
            int32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            int32_t add[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
            nt.insert(nt.end(), add, add + 12);
        }
        m.tri.swap(nt);
    }
    return m;
}

struct Builder {
    std::vector<float> verts;
    std::vector<int32_t> idx;
    std::vector<int32_t> mat_idx;
    std::vector<mipt_material> mats;
    std::vector<int32_t> emissive;

    int32_t add_vertex(V3 p)
    {
        verts.push_back((float)p.x);
        verts.push_back((float)p.y);
        verts.push_back((float)p.z);
        return (int32_t)(verts.size() / 3 - 1);
    }
    void add_tri(int32_t a, int32_t b, int32_t c, int32_t mat)
    {
        idx.push_back(a);
        idx.push_back(b);
        idx.push_back(c);
        mat_idx.push_back(mat);
        const mipt_material& m = mats[mat];
        if (m.emission[0] != 0.0f || m.emission[1] != 0.0f || m.emission[2] != 0.0f)
            emissive.push_back((int32_t)mat_idx.size() - 1);
    }
    // Quad p0 p1 p2 p3 (counter-clockwise) as triangles (p0,p1,p2), (p0,p2,p3).
    void add_quad(V3 p0, V3 p1, V3 p2, V3 p3, int32_t mat)
    {
        int32_t a = add_vertex(p0), b = add_vertex(p1), c = add_vertex(p2), d = add_vertex(p3);
        add_tri(a, b, c, mat);
        add_tri(a, c, d, mat);
    }
    int32_t add_material(const mipt_material& m)
    {
        mats.push_back(m);
        return (int32_t)mats.size() - 1;
    }
    void add_mesh(const Mesh& mesh, const double R[3][3], double scale, V3 t, int32_t mat)
    {
        int32_t base = (int32_t)(verts.size() / 3);
        for (const V3& p : mesh.v)
        {
            V3 r = {R[0][0] * p.x + R[0][1] * p.y + R[0][2] * p.z, R[1][0] * p.x + R[1][1] * p.y + R[1][2] * p.z,
                    R[2][0] * p.x + R[2][1] * p.y + R[2][2] * p.z};
            add_vertex(t + scale * r);
        }
        for (size_t f = 0; f < mesh.tri.size(); f += 3)
            add_tri(base + mesh.tri[f], base + mesh.tri[f + 1], base + mesh.tri[f + 2], mat);
    }
};

mipt_material lambertian(double r, double g, double b)
{
    mipt_material m;
    mipt_material_default(&m);
    m.base_color[0] = (float)r;
    m.base_color[1] = (float)g;
    m.base_color[2] = (float)b;
    m.specular = 0.0f;
    m.metallic = 0.0f;
    m.roughness = 1.0f;
    return m;
}

// Row-major look-at and OpenGL perspective (glm::lookAt / glm::perspective
// conventions: camera looks down -z, NDC z in [-1, 1]); inverses in double.
mipt_camera make_camera(V3 eye, V3 target, V3 up, double vfov, int32_t w, int32_t h)
{
    mipt_camera cam;
    std::memset(&cam, 0, sizeof(cam));
    V3 f = normalize(target - eye);
    V3 s = normalize(cross(f, up));
    V3 u = cross(s, f);
    double view[4][4] = {{s.x, s.y, s.z, -dot(s, eye)},
                         {u.x, u.y, u.z, -dot(u, eye)},
                         {-f.x, -f.y, -f.z, dot(f, eye)},
                         {0, 0, 0, 1}};
    double iview[4][4] = {{s.x, u.x, -f.x, eye.x}, {s.y, u.y, -f.y, eye.y}, {s.z, u.z, -f.z, eye.z}, {0, 0, 0, 1}};
    const double aspect = (double)w / (double)h, n = 0.1, fa = 1000.0;
    const double ft = 1.0 / std::tan(vfov / 2.0);
    double proj[4][4] = {{ft / aspect, 0, 0, 0},
                         {0, ft, 0, 0},
                         {0, 0, (fa + n) / (n - fa), 2.0 * fa * n / (n - fa)},
                         {0, 0, -1, 0}};
    const double A = proj[2][2], B = proj[2][3];
    double iproj[4][4] = {{aspect / ft, 0, 0, 0}, {0, 1.0 / ft, 0, 0}, {0, 0, 0, -1}, {0, 0, 1.0 / B, A / B}};
    double vp[4][4];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
        {
            double acc = 0;
            for (int k = 0; k < 4; k++) acc += proj[i][k] * view[k][j];
            vp[i][j] = acc;
        }
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
        {
            cam.inverse_view[i][j] = (float)iview[i][j];
            cam.inverse_projection[i][j] = (float)iproj[i][j];
            cam.view_projection[i][j] = (float)vp[i][j];
        }
    cam.position[0] = (float)eye.x;
    cam.position[1] = (float)eye.y;
    cam.position[2] = (float)eye.z;
    cam.vertical_fov = (float)vfov;
    cam.sensor_width = w;
    cam.sensor_height = h;
    cam.do_jittering = 1;
    return cam;
}

void random_rotation(Pcg32& rng, double R[3][3])
{
    // Shoemake's uniform random unit quaternion.
    double u1 = rng.uniform(), u2 = rng.uniform(), u3 = rng.uniform();
    double a = std::sqrt(1 - u1), b = std::sqrt(u1);
    double qx = a * std::sin(2 * M_PI * u2), qy = a * std::cos(2 * M_PI * u2);
    double qz = b * std::sin(2 * M_PI * u3), qw = b * std::cos(2 * M_PI * u3);
    R[0][0] = 1 - 2 * (qy * qy + qz * qz); R[0][1] = 2 * (qx * qy - qz * qw); R[0][2] = 2 * (qx * qz + qy * qw);
    R[1][0] = 2 * (qx * qy + qz * qw); R[1][1] = 1 - 2 * (qx * qx + qz * qz); R[1][2] = 2 * (qy * qz - qx * qw);
    R[2][0] = 2 * (qx * qz - qy * qw); R[2][1] = 2 * (qy * qz + qx * qw); R[2][2] = 1 - 2 * (qx * qx + qy * qy);
}

void scene_config1(Builder& B, mipt_camera& cam, int32_t w, int32_t h)
{
    Pcg32 rng(0x2026120200000000ULL + 1, 1);
    const int32_t white = B.add_material(lambertian(0.73, 0.73, 0.73));
    const int32_t red = B.add_material(lambertian(0.63, 0.065, 0.05));
    const int32_t green = B.add_material(lambertian(0.14, 0.45, 0.091));
    // Six faces of [-1,1]^3, each split into two quads along x (or z).
    auto face = [&](V3 o, V3 e0, V3 e1, int32_t mat) {
        V3 half = 0.5 * e0;
        B.add_quad(o, o + half, o + half + e1, o + e1, mat);
        B.add_quad(o + half, o + e0, o + e0 + e1, o + half + e1, mat);
    };
    face({-1, -1, -1}, {2, 0, 0}, {0, 0, 2}, white);  // floor y=-1
    face({-1, 1, 1}, {2, 0, 0}, {0, 0, -2}, white);   // ceiling y=+1
    face({-1, -1, -1}, {2, 0, 0}, {0, 2, 0}, white);  // back z=-1
    face({1, -1, 1}, {-2, 0, 0}, {0, 2, 0}, white);   // front z=+1
    face({-1, -1, 1}, {0, 0, -2}, {0, 2, 0}, red);    // left x=-1
    face({1, -1, -1}, {0, 0, 2}, {0, 2, 0}, green);   // right x=+1
    // Two emissive quads just under the ceiling, facing down.
    for (int q = 0; q < 2; q++)
    {
        mipt_material m = lambertian(0.0, 0.0, 0.0);
        for (int c = 0; c < 3; c++) m.emission[c] = (float)(15.0 + rng.uniform(-1.0, 1.0));
        int32_t mat = B.add_material(m);
        double x0 = q == 0 ? -0.5 : 0.1;
        // Wound so that cross(p1 - p0, p2 - p0) points down (-y), into the
        // box: back-facing lights receive no light samples.
        V3 p0 = {x0, 0.999, -0.3}, p1 = {x0 + 0.4, 0.999, -0.3}, p2 = {x0 + 0.4, 0.999, 0.3}, p3 = {x0, 0.999, 0.3};
        B.add_quad(p0, p1, p2, p3, mat);
    }
    const Mesh sphere = icosphere(4);
    std::vector<V3> centres;
    std::vector<double> radii;
    const double I[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    while (centres.size() < 20)
    {
        double r = rng.uniform(0.08, 0.2);
        V3 c = {rng.uniform(-1 + r + 0.01, 1 - r - 0.01), rng.uniform(-1 + r + 0.01, 0.999 - r - 0.01),
                rng.uniform(-1 + r + 0.01, 1 - r - 0.01)};
        bool ok = true;
        for (size_t k = 0; k < centres.size() && ok; k++)
        {
            V3 dv = c - centres[k];
            ok = std::sqrt(dot(dv, dv)) > r + radii[k] + 0.01;
        }
        if (!ok)
            continue;
        centres.push_back(c);
        radii.push_back(r);
        mipt_material m;
        if (rng.uniform() < 0.5)
            m = lambertian(rng.uniform(0.1, 0.9), rng.uniform(0.1, 0.9), rng.uniform(0.1, 0.9));
        else
        {
            mipt_material_default(&m);
            m.metallic = 1.0f;
            m.roughness = (float)rng.uniform(0.05, 1.0);
            for (int k = 0; k < 3; k++) m.base_color[k] = (float)rng.uniform(0.1, 0.9);
        }
        B.add_mesh(sphere, I, r, c, B.add_material(m));
    }
    cam = make_camera({0, 0, 0.98}, {0, 0, -1}, {0, 1, 0}, 60.0 * M_PI / 180.0, w, h);
}

// Value noise on an integer lattice hashed with a per-octave seed.
double lattice(int64_t ix, int64_t iz, uint32_t seed)
{
    uint64_t hsh = (uint64_t)(ix * 73856093LL) ^ (uint64_t)(iz * 19349663LL) ^ ((uint64_t)seed * 83492791ULL);
    hsh ^= hsh >> 33;
    hsh *= 0xff51afd7ed558ccdULL;
    hsh ^= hsh >> 33;
    hsh *= 0xc4ceb9fe1a85ec53ULL;
    hsh ^= hsh >> 33;
    return (double)(hsh >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

double value_noise(double x, double z, uint32_t seed)
{
    double fx = std::floor(x), fz = std::floor(z);
    int64_t ix = (int64_t)fx, iz = (int64_t)fz;
    double tx = x - fx, tz = z - fz;
    double sx = tx * tx * (3 - 2 * tx), sz = tz * tz * (3 - 2 * tz);
    double a = lattice(ix, iz, seed), b = lattice(ix + 1, iz, seed);
    double c = lattice(ix, iz + 1, seed), d = lattice(ix + 1, iz + 1, seed);
    return (a + (b - a) * sx) * (1 - sz) + (c + (d - c) * sx) * sz;
}

double fbm(double x, double z, const uint32_t seeds[6])
{
    double sum = 0, amp = 2.0, freq = 1.0 / 16.0;
    for (int o = 0; o < 6; o++)
    {
        sum += amp * value_noise(x * freq, z * freq, seeds[o]);
        amp *= 0.5;
        freq *= 2.0;
    }
    return sum;
}

mipt_material principled_random(Pcg32& rng)
{
    mipt_material m;
    mipt_material_default(&m);
    for (int k = 0; k < 3; k++) m.base_color[k] = (float)rng.uniform();
    m.roughness = (float)rng.uniform(0.05, 1.0);
    m.metallic = (float)rng.uniform();
    m.specular = (float)rng.uniform();
    m.specular_tint = (float)rng.uniform();
    m.coat = (float)rng.uniform();
    m.coat_roughness = (float)rng.uniform();
    m.sheen = (float)rng.uniform();
    m.sheen_roughness = (float)rng.uniform(0.05, 1.0);
    m.anisotropy = (float)rng.uniform();
    m.ior = (float)rng.uniform(1.1, 2.0);
    m.specular_transmission = (float)rng.uniform();
    return m;
}

void scene_config2(Builder& B, mipt_camera& cam, int32_t w, int32_t h)
{
    Pcg32 rng(0x2026120200000000ULL + 2, 2);
    uint32_t seeds[6];
    for (int o = 0; o < 6; o++) seeds[o] = rng.next();
    // Terrain: (N+1)^2 vertices over [-32, 32]^2.
    const int N = 512;
    const double ext = 32.0;
    const int32_t terrain_mat = B.add_material(principled_random(rng));
    const int32_t base = (int32_t)(B.verts.size() / 3);
    for (int j = 0; j <= N; j++)
        for (int i = 0; i <= N; i++)
        {
            double x = -ext + 2 * ext * i / N, z = -ext + 2 * ext * j / N;
            B.add_vertex({x, fbm(x, z, seeds), z});
        }
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++)
        {
            int32_t a = base + j * (N + 1) + i, b = a + 1, c = a + (N + 1), d = c + 1;
            B.add_tri(a, c, b, terrain_mat);
            B.add_tri(b, c, d, terrain_mat);
        }
    // 400 spheres.
    const Mesh sphere = icosphere(3);
    for (int s = 0; s < 400; s++)
    {
        double R[3][3];
        random_rotation(rng, R);
        double scale = rng.uniform(0.3, 1.5);
        double x = rng.uniform(-28, 28), z = rng.uniform(-28, 28);
        double y = fbm(x, z, seeds) + scale * rng.uniform(0.5, 3.0);
        B.add_mesh(sphere, R, scale, {x, y, z}, B.add_material(principled_random(rng)));
    }
    // 4096 emissive triangles above the scene, facing roughly down.
    for (int e = 0; e < 4096; e++)
    {
        V3 c = {rng.uniform(-28, 28), rng.uniform(6, 10), rng.uniform(-28, 28)};
        double size = rng.uniform(0.2, 0.6);
        double phi = rng.uniform(0, 2 * M_PI);
        V3 p0 = c + size * V3{std::cos(phi), 0, std::sin(phi)};
        V3 p1 = c + size * V3{std::cos(phi + 2.0944), rng.uniform(-0.2, 0.2), std::sin(phi + 2.0944)};
        V3 p2 = c + size * V3{std::cos(phi + 4.1888), rng.uniform(-0.2, 0.2), std::sin(phi + 4.1888)};
        double area = 0.5 * std::sqrt(dot(cross(p1 - p0, p2 - p0), cross(p1 - p0, p2 - p0)));
        double power = std::pow(10.0, rng.uniform(0.0, 3.0));
        mipt_material m = lambertian(0, 0, 0);
        V3 tint = {rng.uniform(0.5, 1.0), rng.uniform(0.5, 1.0), rng.uniform(0.5, 1.0)};
        double L = power / (area > 1e-12 ? area : 1e-12);
        m.emission[0] = (float)(L * tint.x);
        m.emission[1] = (float)(L * tint.y);
        m.emission[2] = (float)(L * tint.z);
        int32_t mat = B.add_material(m);
        int32_t a = B.add_vertex(p0), b = B.add_vertex(p1), d = B.add_vertex(p2);
        B.add_tri(a, b, d, mat);
    }
    cam = make_camera({0, 14, 16}, {0, 0, 0}, {0, 1, 0}, 45.0 * M_PI / 180.0, w, h);
}

// Every layer of the Principled BSDF drawn per object, thin film and
// dispersion included.
mipt_material principled_full(Pcg32& rng)
{
    mipt_material m = principled_random(rng);
    m.thin_film = (float)rng.uniform();
    m.thin_film_thickness = (float)rng.uniform(100.0, 900.0);
    m.thin_film_ior = (float)rng.uniform(1.2, 2.0);
    m.dispersion_scale = (float)rng.uniform();
    m.dispersion_abbe_number = (float)rng.uniform(20.0, 60.0);
    return m;
}

// Full stack at scale: 8 displaced grids of 320 x 320 quads (1,638,400
// triangles), 3000 instances of three base meshes (icospheres of 1280, 5120
// and 20480 triangles, drawn with probabilities 0.70 / 0.27 / 0.03, random
// rotation and scale) and 16,384 emissive triangles of power log-uniform
// over three decades: about 10M triangles. Instances are flattened into the
// one triangle list the single-level BVH takes.
void scene_config5(Builder& B, mipt_camera& cam, int32_t w, int32_t h)
{
    Pcg32 rng(0x2026120200000000ULL + 5, 5);
    const int N = 320;
    const double ext = 24.0; // each grid spans 48 x 48, the 8 grids a 4 x 2 layout
    uint32_t seeds[8][6];
    for (int g = 0; g < 8; g++)
        for (int o = 0; o < 6; o++) seeds[g][o] = rng.next();
    for (int g = 0; g < 8; g++)
    {
        const double ox = -72.0 + 48.0 * (g % 4), oz = -24.0 + 48.0 * (g / 4);
        const int32_t mat = B.add_material(principled_full(rng));
        const int32_t base = (int32_t)(B.verts.size() / 3);
        for (int j = 0; j <= N; j++)
            for (int i = 0; i <= N; i++)
            {
                const double x = -ext + 2 * ext * i / N, z = -ext + 2 * ext * j / N;
                B.add_vertex({ox + x, fbm(x, z, seeds[g]), oz + z});
            }
        for (int j = 0; j < N; j++)
            for (int i = 0; i < N; i++)
            {
                const int32_t a = base + j * (N + 1) + i, b = a + 1, c = a + (N + 1), d = c + 1;
                B.add_tri(a, c, b, mat);
                B.add_tri(b, c, d, mat);
            }
    }
    const Mesh meshes[3] = {icosphere(3), icosphere(4), icosphere(5)};
    for (int s = 0; s < 3000; s++)
    {
        const double u = rng.uniform();
        const Mesh& mesh = meshes[u < 0.70 ? 0 : (u < 0.97 ? 1 : 2)];
        double R[3][3];
        random_rotation(rng, R);
        const double scale = rng.uniform(0.3, 1.8);
        const int g = (int)rng.uniform(0, 8);
        const double x = rng.uniform(-22, 22), z = rng.uniform(-22, 22);
        const double y = fbm(x, z, seeds[g]) + scale * rng.uniform(0.5, 3.0);
        B.add_mesh(mesh, R, scale, {-72.0 + 48.0 * (g % 4) + x, y, -24.0 + 48.0 * (g / 4) + z},
                   B.add_material(principled_full(rng)));
    }
    for (int e = 0; e < 16384; e++)
    {
        const V3 c = {rng.uniform(-94, 94), rng.uniform(8, 14), rng.uniform(-46, 46)};
        const double size = rng.uniform(0.2, 0.6);
        const double phi = rng.uniform(0, 2 * M_PI);
        const V3 p0 = c + size * V3{std::cos(phi), 0, std::sin(phi)};
        const V3 p1 = c + size * V3{std::cos(phi + 2.0944), rng.uniform(-0.2, 0.2), std::sin(phi + 2.0944)};
        const V3 p2 = c + size * V3{std::cos(phi + 4.1888), rng.uniform(-0.2, 0.2), std::sin(phi + 4.1888)};
        const double area = 0.5 * std::sqrt(dot(cross(p1 - p0, p2 - p0), cross(p1 - p0, p2 - p0)));
        const double power = std::pow(10.0, rng.uniform(0.0, 3.0));
        mipt_material m = lambertian(0, 0, 0);
        const V3 tint = {rng.uniform(0.5, 1.0), rng.uniform(0.5, 1.0), rng.uniform(0.5, 1.0)};
        const double L = power / (area > 1e-12 ? area : 1e-12);
        m.emission[0] = (float)(L * tint.x);
        m.emission[1] = (float)(L * tint.y);
        m.emission[2] = (float)(L * tint.z);
        const int32_t mat = B.add_material(m);
        const int32_t a = B.add_vertex(p0), b = B.add_vertex(p1), d = B.add_vertex(p2);
        B.add_tri(a, b, d, mat);
    }
    cam = make_camera({0, 26, 58}, {0, 0, 0}, {0, 1, 0}, 50.0 * M_PI / 180.0, w, h);
}

template <class T> T* dup(const std::vector<T>& v)
{
    T* p = new T[v.size() > 0 ? v.size() : 1];
    if (!v.empty())
        std::memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}

} // namespace

// The look-at camera of the scene loaders (glTF ingest).
mipt_camera look_at_camera(const double eye[3], const double target[3], const double up[3], double vfov, int32_t w,
                           int32_t h)
{
    return make_camera({eye[0], eye[1], eye[2]}, {target[0], target[1], target[2]}, {up[0], up[1], up[2]}, vfov, w,
                       h);
}

} // namespace mipt

extern "C" int mipt_scene_generate(int32_t config_id, int32_t width, int32_t height, mipt_generated_scene** out)
{
    if (!out || width <= 0 || height <= 0)
        return MIPT_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    try
    {
        mipt::Builder B;
        mipt_camera cam;
        if (config_id == 1)
            mipt::scene_config1(B, cam, width, height);
        else if (config_id == 2)
            mipt::scene_config2(B, cam, width, height);
        else if (config_id == 5)
            mipt::scene_config5(B, cam, width, height);
        else
            return MIPT_ERR_INVALID_ARGUMENT;
        auto* g = new mipt_generated_scene;
        std::memset(g, 0, sizeof(*g));
        g->desc.triangles_vertex_indices = mipt::dup(B.idx);
        g->desc.triangle_count = (int32_t)(B.idx.size() / 3);
        g->desc.vertices_positions = mipt::dup(B.verts);
        g->desc.vertex_count = (int32_t)(B.verts.size() / 3);
        g->desc.has_vertex_normals = nullptr;
        g->desc.vertex_normals = nullptr;
        g->desc.material_indices = mipt::dup(B.mat_idx);
        g->desc.materials = mipt::dup(B.mats);
        g->desc.material_count = (int32_t)B.mats.size();
        g->desc.emissive_triangles_primitive_indices = mipt::dup(B.emissive);
        g->desc.emissive_triangle_count = (int32_t)B.emissive.size();
        g->camera = cam;
        g->width = width;
        g->height = height;
        *out = g;
        return MIPT_OK;
    }
    catch (const std::bad_alloc&)
    {
        return MIPT_ERR_OUT_OF_MEMORY;
    }
}

extern "C" int mipt_views_generate(int32_t count, mipt_view* out)
{
    if (count < 0 || (count > 0 && !out))
        return MIPT_ERR_INVALID_ARGUMENT;
    mipt::Pcg32 rng(0x2026120200000000ULL + 3, 3);
    for (int32_t v = 0; v < count; v++)
    {
        // height log-uniform in [128, 1024], aspect uniform in [1, 16/9]
        const double h = 128.0 * std::exp(rng.uniform() * std::log(8.0));
        const double aspect = 1.0 + rng.uniform() * (16.0 / 9.0 - 1.0);
        const int32_t H = (int32_t)std::lround(h);
        const int32_t W = (int32_t)std::lround(h * aspect);
        // inside the box, in front of the spheres, looking at the back wall
        const mipt::V3 eye{rng.uniform(-0.6, 0.6), rng.uniform(-0.6, 0.6), rng.uniform(0.5, 0.98)};
        const mipt::V3 target{rng.uniform(-0.4, 0.4), rng.uniform(-0.4, 0.4), -1.0};
        const double vfov = rng.uniform(40.0, 70.0) * M_PI / 180.0;
        out[v].camera = mipt::make_camera(eye, target, {0, 1, 0}, vfov, W, H);
        out[v].width = W;
        out[v].height = H;
    }
    return MIPT_OK;
}

extern "C" int mipt_envmap_generate(int32_t width, int32_t height, float* out_rgb)
{
    if (width <= 0 || height <= 0 || !out_rgb || (int64_t)width * height > (1ll << 26))
        return MIPT_ERR_INVALID_ARGUMENT;
    mipt::Pcg32 rng(0x2026120200000000ULL + 5, 5);
    // the sun: 35 degrees above the horizon, azimuth 60 degrees
    const double elev = 35.0 * M_PI / 180.0, azim = 60.0 * M_PI / 180.0;
    const mipt::V3 sun{std::cos(elev) * std::cos(azim), std::sin(elev), std::cos(elev) * std::sin(azim)};
    const double sun_cos = std::cos(1.5 * M_PI / 180.0);
    for (int32_t y = 0; y < height; y++)
    {
        // the direction of the texel's centre in the sampler's convention:
        // theta from straight down, phi around Y
        const double theta = (y + 0.5) / height * M_PI;
        for (int32_t x = 0; x < width; x++)
        {
            const double phi = (x + 0.5) / width * 2.0 * M_PI;
            const mipt::V3 d{-std::sin(theta) * std::cos(phi), -std::cos(theta), -std::sin(theta) * std::sin(phi)};
            double r, g, b;
            if (d.y > 0.0)
            {
                // zenith blue to a pale horizon
                const double t = std::pow(1.0 - d.y, 4.0);
                r = 0.25 + 0.65 * t;
                g = 0.45 + 0.45 * t;
                b = 0.90 + 0.10 * t;
            }
            else
            {
                r = 0.20;
                g = 0.17;
                b = 0.14;
            }
            // lognormal texel factor exp(0.35 N), N by Box-Muller
            const double u1 = 1.0 - rng.uniform(), u2 = rng.uniform();
            const double n01 = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
            const double k = std::exp(0.35 * n01);
            r *= k;
            g *= k;
            b *= k;
            if (mipt::dot(d, sun) >= sun_cos)
            {
                r = 2000.0;
                g = 1800.0;
                b = 1500.0;
            }
            float* o = out_rgb + 3 * ((size_t)y * (size_t)width + (size_t)x);
            o[0] = (float)r;
            o[1] = (float)g;
            o[2] = (float)b;
        }
    }
    return MIPT_OK;
}

extern "C" void mipt_scene_free(mipt_generated_scene* g)
{
    if (!g)
        return;
    delete[] g->desc.triangles_vertex_indices;
    delete[] g->desc.vertices_positions;
    delete[] g->desc.has_vertex_normals;
    delete[] g->desc.vertex_normals;
    delete[] g->desc.material_indices;
    delete[] g->desc.materials;
    delete[] g->desc.emissive_triangles_primitive_indices;
    delete g;
}
