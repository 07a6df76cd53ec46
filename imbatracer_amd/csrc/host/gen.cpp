// Seeded synthetic scenes and camera rays for the measured configurations.
#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <utility>
#include <vector>

#include "imba_hip.h"
#include "rng.h"
#include "status.h"
#include "vec.h"

using namespace imba;

namespace {

// Row-major 4x4 helpers with the reference's summation order (core/float4x4.h:114-135).
Mat4 mul(const Mat4& a, const Mat4& b) {
    Mat4 r;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j] + a.m[i][3] * b.m[3][j];
    return r;
}

Mat4 diag(float x, float y, float z, float w) {
    Mat4 r{};
    r.m[0][0] = x; r.m[1][1] = y; r.m[2][2] = z; r.m[3][3] = w;
    return r;
}

Mat4 translation(float x, float y, float z) {
    Mat4 r = diag(1, 1, 1, 1);
    r.m[0][3] = x; r.m[1][3] = y; r.m[2][3] = z;
    return r;
}

const float kPi = 3.14159265359f;  // core/common.h:10

// Projection of a camera looking down -z (core/float4x4.h:32-46).
Mat4 perspective(float fov, float aspect, float znear, float zfar) {
    const float f_h = 1.0f / std::tan(fov * kPi / 360.0f);
    const float f_v = aspect / std::tan(fov * kPi / 360.0f);
    const float d = 1.0f / (znear - zfar);
    Mat4 r{};
    r.m[0][0] = f_h;
    r.m[1][1] = -f_v;
    r.m[2][2] = (znear + zfar) * d;
    r.m[2][3] = 2.0f * znear * zfar * d;
    r.m[3][2] = -1.0f;
    return r;
}

}  // namespace

extern "C" {

int imba_gen_random_triangles(uint64_t seed, int n, float* vertices, uint32_t* indices) {
    IMBA_CHECK_ARG(vertices && indices && n >= 0, "bad arguments");
    Mwc64x rng(seed);
    for (int t = 0; t < n; t++) {
        float c[3];
        for (int k = 0; k < 3; k++) c[k] = rng.uniform();
        for (int v = 0; v < 3; v++) {
            float* out = vertices + 4 * (3 * (size_t)t + v);
            for (int k = 0; k < 3; k++) out[k] = c[k] + rng.uniform(-0.01f, 0.01f);
            out[3] = 1.0f;
        }
        indices[4 * (size_t)t + 0] = 3 * t + 0;
        indices[4 * (size_t)t + 1] = 3 * t + 1;
        indices[4 * (size_t)t + 2] = 3 * t + 2;
        indices[4 * (size_t)t + 3] = 0;
    }
    return IMBA_OK;
}

int imba_gen_icosphere_counts(int level, int* vertex_count, int* tri_count) {
    IMBA_CHECK_ARG(level >= 0 && level <= 8 && vertex_count && tri_count, "bad arguments");
    int f = 20;
    for (int l = 0; l < level; l++) f *= 4;
    *tri_count = f;
    *vertex_count = f / 2 + 2;  // Euler: V - E + F = 2, E = 3F/2
    return IMBA_OK;
}

// Icosahedron with 12 vertices (0, +-1, +-phi) and cyclic permutations; each level splits
// every triangle into four at the edge midpoints, which are pushed onto the unit sphere.
int imba_gen_icosphere(int level, float radius, float cx, float cy, float cz, uint32_t material,
                       float* vertices, uint32_t* indices) {
    IMBA_CHECK_ARG(level >= 0 && level <= 8 && vertices && indices, "bad arguments");
    const double phi = (1.0 + std::sqrt(5.0)) / 2.0;
    std::vector<std::array<double, 3>> v = {
        {-1, phi, 0}, {1, phi, 0}, {-1, -phi, 0}, {1, -phi, 0},
        {0, -1, phi}, {0, 1, phi}, {0, -1, -phi}, {0, 1, -phi},
        {phi, 0, -1}, {phi, 0, 1}, {-phi, 0, -1}, {-phi, 0, 1}};
    for (auto& p : v) {
        const double l = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        for (auto& c : p) c /= l;
    }
    std::vector<std::array<uint32_t, 3>> f = {
        {0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11},
        {1, 5, 9}, {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
        {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8}, {3, 8, 9},
        {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    for (int l = 0; l < level; l++) {
        std::map<std::pair<uint32_t, uint32_t>, uint32_t> mid;
        auto midpoint = [&](uint32_t a, uint32_t b) {
            const auto key = std::make_pair(std::min(a, b), std::max(a, b));
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            std::array<double, 3> p;
            for (int k = 0; k < 3; k++) p[k] = 0.5 * (v[a][k] + v[b][k]);
            const double len = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
            for (auto& c : p) c /= len;
            v.push_back(p);
            const uint32_t id = (uint32_t)v.size() - 1;
            mid[key] = id;
            return id;
        };
        std::vector<std::array<uint32_t, 3>> nf;
        nf.reserve(f.size() * 4);
        for (const auto& t : f) {
            const uint32_t a = midpoint(t[0], t[1]), b = midpoint(t[1], t[2]), c = midpoint(t[2], t[0]);
            nf.push_back({t[0], a, c});
            nf.push_back({t[1], b, a});
            nf.push_back({t[2], c, b});
            nf.push_back({a, b, c});
        }
        f.swap(nf);
    }
    for (size_t i = 0; i < v.size(); i++) {
        vertices[4 * i + 0] = (float)(cx + radius * v[i][0]);
        vertices[4 * i + 1] = (float)(cy + radius * v[i][1]);
        vertices[4 * i + 2] = (float)(cz + radius * v[i][2]);
        vertices[4 * i + 3] = 1.0f;
    }
    for (size_t i = 0; i < f.size(); i++) {
        indices[4 * i + 0] = f[i][0];
        indices[4 * i + 1] = f[i][1];
        indices[4 * i + 2] = f[i][2];
        indices[4 * i + 3] = material;
    }
    return IMBA_OK;
}

// PerspectiveCamera(w, h, fov) + move(pos, dir, up) + generate_ray(x + 0.5, y + 0.5)
// (render/ray_gen/camera.h:19-79), aspect 1, near 0.1, far 10000.
int imba_gen_camera_matrix(const float* pos3, const float* dir3, const float* up3, float fov_deg,
                           int width, int height, float* out16) {
    IMBA_CHECK_ARG(pos3 && dir3 && up3 && out16 && width > 0 && height > 0, "bad arguments");
    const float3 pos(pos3[0], pos3[1], pos3[2]);
    const float3 dir = normalize(float3(dir3[0], dir3[1], dir3[2]));
    const float3 right = normalize(cross(dir, float3(up3[0], up3[1], up3[2])));
    const float3 up = cross(dir, right);
    const float3 nright(-right.x, -right.y, -right.z), ndir(-dir.x, -dir.y, -dir.z);
    const float3 lp(dot(up, pos), dot(nright, pos), dot(ndir, pos));
    Mat4 w2c{};
    const float3 rows[3] = {up, nright, ndir};
    const float offs[3] = {-lp.x, -lp.y, -lp.z};
    for (int r = 0; r < 3; r++) {
        w2c.m[r][0] = rows[r].x; w2c.m[r][1] = rows[r].y; w2c.m[r][2] = rows[r].z; w2c.m[r][3] = offs[r];
    }
    w2c.m[3][3] = 1.0f;
    const Mat4 w2s = mul(perspective(fov_deg, 1.0f, 0.1f, 10000.0f), w2c);
    const Mat4 s2w = invert(w2s);
    const Mat4 r2w = mul(mul(s2w, translation(-1.0f, -1.0f, 0.0f)),
                         diag(2.0f / (float)height, 2.0f / (float)width, 0.0f, 1.0f));
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) out16[4 * r + c] = r2w.m[r][c];
    return IMBA_OK;
}

int imba_gen_camera_frame(const float* pos3, const float* dir3, const float* up3, float fov_deg, int width,
                          int height, float* out16, float* forward3, float* image_plane_dist) {
    IMBA_CHECK_ARG(pos3 && dir3 && up3 && out16 && forward3 && image_plane_dist && width > 0 && height > 0,
                   "bad arguments");
    const float3 pos(pos3[0], pos3[1], pos3[2]);
    const float3 dir = normalize(float3(dir3[0], dir3[1], dir3[2]));
    const float3 right = normalize(cross(dir, float3(up3[0], up3[1], up3[2])));
    const float3 up = cross(dir, right);
    const float3 nright(-right.x, -right.y, -right.z), ndir(-dir.x, -dir.y, -dir.z);
    const float3 lp(dot(up, pos), dot(nright, pos), dot(ndir, pos));
    Mat4 w2c{};
    const float3 rows[3] = {up, nright, ndir};
    const float offs[3] = {-lp.x, -lp.y, -lp.z};
    for (int r = 0; r < 3; r++) {
        w2c.m[r][0] = rows[r].x; w2c.m[r][1] = rows[r].y; w2c.m[r][2] = rows[r].z; w2c.m[r][3] = offs[r];
    }
    w2c.m[3][3] = 1.0f;
    const Mat4 w2s = mul(perspective(fov_deg, 1.0f, 0.1f, 10000.0f), w2c);
    // world_to_raster_ = scale(h / 2, w / 2, 0, 1) * translate(1, 1, 0) * world_to_screen
    const Mat4 w2r = mul(mul(diag((float)height * 0.5f, (float)width * 0.5f, 0.0f, 1.0f),
                             translation(1.0f, 1.0f, 0.0f)), w2s);
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) out16[4 * r + c] = w2r.m[r][c];
    forward3[0] = dir.x;
    forward3[1] = dir.y;
    forward3[2] = dir.z;
    const float tan_half = std::tan(fov_deg * kPi / 360.0f);
    *image_plane_dist = (float)width / (2.0f * tan_half);
    return IMBA_OK;
}

int imba_gen_camera_rays(const float* pos3, const float* dir3, const float* up3, float fov_deg,
                         int width, int height, ImbaRay* rays) {
    IMBA_CHECK_ARG(pos3 && dir3 && up3 && rays && width > 0 && height > 0, "bad arguments");
    Mat4 r2w;
    int rc = imba_gen_camera_matrix(pos3, dir3, up3, fov_deg, width, height, &r2w.m[0][0]);
    if (rc) return rc;
    const float3 pos(pos3[0], pos3[1], pos3[2]);
    for (int y = 0; y < height; y++) {
        for (int x = 0; x < width; x++) {
            const float px = (float)x + 0.5f, py = (float)y + 0.5f;
            const float in[4] = {py, px, 0.0f, 1.0f};
            float t[4];
            for (int r = 0; r < 4; r++)
                t[r] = r2w.m[r][0] * in[0] + r2w.m[r][1] * in[1] + r2w.m[r][2] * in[2] + r2w.m[r][3] * in[3];
            const float3 w(t[0] / t[3], t[1] / t[3], t[2] / t[3]);
            const float3 d = normalize(w - pos);
            ImbaRay& ray = rays[(size_t)y * width + x];
            ray.org = {pos.x, pos.y, pos.z, 0.0f};
            ray.dir = {d.x, d.y, d.z, FLT_MAX};
        }
    }
    return IMBA_OK;
}

}  // extern "C"
