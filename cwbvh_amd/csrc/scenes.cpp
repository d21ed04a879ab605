// Seeded scene and ray generators (host).
//
// Replaces the reference's input stage: mesh/surfel files read in
// src/main.cpp:31-64, the AO ray kernel src/ValidationKernels.cu:41-101 and the
// known-answer scene src/ValidationKernels.cu:6-39.  Every generator is a pure
// function of its arguments and seed (counter-based SplitMix64 streams), so
// the same rays come out on any host with the same libm.
#include <algorithm>
#include <array>
#include <map>
#include <unordered_map>

#include "aogen.h"
#include "common.h"

using namespace cw;

namespace {

constexpr double kPi = 3.14159265358979323846;

cwbvh_mesh* new_mesh() { return new cwbvh_mesh(); }

// Shirley-Chiu concentric square-to-disk map, lifted to the cosine-weighted
// hemisphere (Malley's method).
V3 cosine_hemisphere(float u1, float u2) {
    double a = 2.0 * u1 - 1.0, b = 2.0 * u2 - 1.0;
    double r, phi;
    if (a == 0.0 && b == 0.0) {
        r = 0.0;
        phi = 0.0;
    } else if (a * a > b * b) {
        r = a;
        phi = (kPi / 4.0) * (b / a);
    } else {
        r = b;
        phi = (kPi / 2.0) - (kPi / 4.0) * (a / b);
    }
    double x = r * std::cos(phi), y = r * std::sin(phi);
    double z = std::sqrt(std::max(0.0, 1.0 - x * x - y * y));
    return V3{(float)x, (float)y, (float)z};
}

// Orthonormal basis around unit n (Duff et al. 2017, "Building an
// Orthonormal Basis, Revisited").
void onb(V3 n, V3& t, V3& b) {
    float sign = std::copysign(1.0f, n.z);
    float a = -1.0f / (sign + n.z);
    float c = n.x * n.y * a;
    t = V3{1.0f + sign * n.x * n.x * a, sign * c, -sign * n.x};
    b = V3{c, sign + n.y * n.y * a, -n.y};
}

V3 uniform_sphere(float u1, float u2) {
    double z = 1.0 - 2.0 * u1;
    double r = std::sqrt(std::max(0.0, 1.0 - z * z));
    double phi = 2.0 * kPi * u2;
    return V3{(float)(r * std::cos(phi)), (float)(r * std::sin(phi)), (float)z};
}

V3 tri_normal(V3 a, V3 b, V3 c) { return normalize(cross(b - a, c - a)); }

// Area-uniform surface sampler.
struct AreaSampler {
    const cwbvh_mesh* m;
    std::vector<double> cdf;  // inclusive prefix of triangle areas
    explicit AreaSampler(const cwbvh_mesh* mesh) : m(mesh) {
        int64_t nt = m->nt();
        cdf.resize(nt);
        double acc = 0.0;
        for (int64_t t = 0; t < nt; ++t) {
            V3 a, b, c;
            m->tri(t, a, b, c);
            acc += 0.5 * (double)length(cross(b - a, c - a));
            cdf[t] = acc;
        }
    }
    double total() const { return cdf.empty() ? 0.0 : cdf.back(); }
    // returns triangle index and point
    int64_t sample(Rng& rng, V3& p, V3& n) const {
        double u = rng.uniform_d() * total();
        int64_t t = std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin();
        if (t >= (int64_t)cdf.size()) t = (int64_t)cdf.size() - 1;
        V3 a, b, c;
        m->tri(t, a, b, c);
        float r1 = std::sqrt(rng.uniform()), r2 = rng.uniform();
        float w0 = 1.0f - r1, w1 = r1 * (1.0f - r2), w2 = r1 * r2;
        p = a * w0 + b * w1 + c * w2;
        n = tri_normal(a, b, c);
        return t;
    }
};

void put_ray(cwbvh_ray& r, V3 o, float tmin, V3 d, float tmax) {
    r.origin_tmin = cwbvh_float4{o.x, o.y, o.z, tmin};
    r.dir_tmax = cwbvh_float4{d.x, d.y, d.z, tmax};
}

// Uniform random rotation (Shoemake's quaternion method) as a 3x3 matrix.
void random_rotation(Rng& rng, float R[9]) {
    double u1 = rng.uniform_d(), u2 = rng.uniform_d(), u3 = rng.uniform_d();
    double a = std::sqrt(1 - u1), b = std::sqrt(u1);
    double qx = a * std::sin(2 * kPi * u2), qy = a * std::cos(2 * kPi * u2);
    double qz = b * std::sin(2 * kPi * u3), qw = b * std::cos(2 * kPi * u3);
    R[0] = (float)(1 - 2 * (qy * qy + qz * qz));
    R[1] = (float)(2 * (qx * qy - qz * qw));
    R[2] = (float)(2 * (qx * qz + qy * qw));
    R[3] = (float)(2 * (qx * qy + qz * qw));
    R[4] = (float)(1 - 2 * (qx * qx + qz * qz));
    R[5] = (float)(2 * (qy * qz - qx * qw));
    R[6] = (float)(2 * (qx * qz - qy * qw));
    R[7] = (float)(2 * (qy * qz + qx * qw));
    R[8] = (float)(1 - 2 * (qx * qx + qy * qy));
}

// Unit geodesic sphere of the given frequency (20*f^2 triangles), vertices
// deduplicated by their integer barycentric key on the icosahedron.
void geodesic_sphere(int f, std::vector<V3>& verts, std::vector<std::array<int, 3>>& tris) {
    const double phi = (1.0 + std::sqrt(5.0)) / 2.0;
    const double ico[12][3] = {{-1, phi, 0}, {1, phi, 0},  {-1, -phi, 0}, {1, -phi, 0},
                               {0, -1, phi}, {0, 1, phi},  {0, -1, -phi}, {0, 1, -phi},
                               {phi, 0, -1}, {phi, 0, 1},  {-phi, 0, -1}, {-phi, 0, 1}};
    int faces[20][3] = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11},
                        {1, 5, 9},  {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                        {3, 9, 4},  {3, 4, 2},  {3, 2, 6},   {3, 6, 8},  {3, 8, 9},
                        {4, 9, 5},  {2, 4, 11}, {6, 2, 10},  {8, 6, 7},  {9, 8, 1}};
    // orient every face outward
    for (auto& fc : faces) {
        double A[3], B[3], C[3];
        for (int k = 0; k < 3; ++k) { A[k] = ico[fc[0]][k]; B[k] = ico[fc[1]][k]; C[k] = ico[fc[2]][k]; }
        double e1[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
        double e2[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
        double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        if (n[0] * (A[0] + B[0] + C[0]) + n[1] * (A[1] + B[1] + C[1]) + n[2] * (A[2] + B[2] + C[2]) < 0)
            std::swap(fc[1], fc[2]);
    }
    std::map<std::vector<int>, int> key_to_vertex;
    auto vertex_of = [&](int ia, int wa, int ib, int wb, int ic, int wc) {
        std::vector<std::pair<int, int>> parts;
        if (wa) parts.push_back({ia, wa});
        if (wb) parts.push_back({ib, wb});
        if (wc) parts.push_back({ic, wc});
        std::sort(parts.begin(), parts.end());
        std::vector<int> key;
        for (auto& p : parts) { key.push_back(p.first); key.push_back(p.second); }
        auto it = key_to_vertex.find(key);
        if (it != key_to_vertex.end()) return it->second;
        double P[3] = {0, 0, 0};
        for (auto& p : parts)
            for (int k = 0; k < 3; ++k) P[k] += p.second * ico[p.first][k];
        double l = std::sqrt(P[0] * P[0] + P[1] * P[1] + P[2] * P[2]);
        int id = (int)verts.size();
        verts.push_back(V3{(float)(P[0] / l), (float)(P[1] / l), (float)(P[2] / l)});
        key_to_vertex.emplace(key, id);
        return id;
    };
    for (auto& fc : faces) {
        // grid point (i, j): weight i on B, j on C, f-i-j on A
        std::vector<std::vector<int>> id(f + 1);
        for (int i = 0; i <= f; ++i) {
            id[i].resize(f + 1 - i);
            for (int j = 0; j + i <= f; ++j) id[i][j] = vertex_of(fc[0], f - i - j, fc[1], i, fc[2], j);
        }
        for (int i = 0; i < f; ++i)
            for (int j = 0; j + i < f; ++j) {
                tris.push_back({id[i][j], id[i + 1][j], id[i][j + 1]});
                if (i + j + 1 < f) tris.push_back({id[i + 1][j], id[i + 1][j + 1], id[i][j + 1]});
            }
    }
}

}  // namespace

extern "C" {

int32_t cwbvh_mesh_from_arrays(const float* vertices, int64_t num_vertices, const int32_t* indices,
                               int64_t num_triangles, cwbvh_mesh** out) {
    CW_CHECK_ARG(out != nullptr, "out is NULL");
    CW_CHECK_ARG(num_vertices >= 0 && num_triangles >= 0, "negative count");
    CW_CHECK_ARG(num_vertices == 0 || vertices != nullptr, "vertices is NULL");
    CW_CHECK_ARG(num_triangles == 0 || indices != nullptr, "indices is NULL");
    for (int64_t i = 0; i < 3 * num_triangles; ++i)
        CW_CHECK_ARG(indices[i] >= 0 && indices[i] < num_vertices, "vertex index out of range");
    cwbvh_mesh* m = new_mesh();
    m->vertices.assign(vertices, vertices + 3 * num_vertices);
    m->indices.assign(indices, indices + 3 * num_triangles);
    *out = m;
    return CWBVH_OK;
}

int32_t cwbvh_mesh_view(const cwbvh_mesh* mesh, const float** vertices, int64_t* num_vertices,
                        const int32_t** indices, int64_t* num_triangles) {
    CW_CHECK_ARG(mesh != nullptr, "mesh is NULL");
    if (vertices) *vertices = mesh->vertices.data();
    if (num_vertices) *num_vertices = mesh->nv();
    if (indices) *indices = mesh->indices.data();
    if (num_triangles) *num_triangles = mesh->nt();
    return CWBVH_OK;
}

int32_t cwbvh_mesh_bounds(const cwbvh_mesh* mesh, float* lo3, float* hi3) {
    CW_CHECK_ARG(mesh != nullptr && lo3 != nullptr && hi3 != nullptr, "NULL argument");
    Box b;
    for (int64_t t = 0; t < mesh->nt(); ++t)
        for (int k = 0; k < 3; ++k) b.grow(mesh->vert(mesh->indices[3 * t + k]));
    if (b.empty()) b.lo = b.hi = V3{0, 0, 0};
    lo3[0] = b.lo.x; lo3[1] = b.lo.y; lo3[2] = b.lo.z;
    hi3[0] = b.hi.x; hi3[1] = b.hi.y; hi3[2] = b.hi.z;
    return CWBVH_OK;
}

void cwbvh_mesh_free(cwbvh_mesh* mesh) { delete mesh; }

int32_t cwbvh_mesh_heightfield(int32_t nx, int32_t ny, int32_t octaves, float amp0, uint64_t seed,
                               cwbvh_mesh** out) {
    CW_CHECK_ARG(out != nullptr, "out is NULL");
    CW_CHECK_ARG(nx >= 1 && ny >= 1 && octaves >= 0 && octaves <= 30, "bad heightfield size");
    Rng rng(seed, 1);
    std::vector<double> th(octaves), ph(octaves);
    for (int k = 0; k < octaves; ++k) {
        th[k] = 2.0 * kPi * rng.uniform_d();
        ph[k] = 2.0 * kPi * rng.uniform_d();
    }
    cwbvh_mesh* m = new_mesh();
    m->vertices.resize(3 * (size_t)(nx + 1) * (ny + 1));
    for (int j = 0; j <= ny; ++j)
        for (int i = 0; i <= nx; ++i) {
            double x = (double)i / nx, y = (double)j / ny, h = 0.0;
            for (int k = 0; k < octaves; ++k) {
                double f = std::ldexp(1.0, k);
                h += amp0 * std::ldexp(1.0, -k) * std::sin(2.0 * kPi * f * (x * std::cos(th[k]) + y * std::sin(th[k])) + ph[k]);
            }
            size_t v = (size_t)j * (nx + 1) + i;
            m->vertices[3 * v + 0] = (float)x;
            m->vertices[3 * v + 1] = (float)y;
            m->vertices[3 * v + 2] = (float)h;
        }
    m->indices.reserve(6 * (size_t)nx * ny);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            int32_t a = j * (nx + 1) + i, b = a + 1, c = a + (nx + 1) + 1, d = a + (nx + 1);
            int32_t t0[3] = {a, b, c}, t1[3] = {a, c, d};
            m->indices.insert(m->indices.end(), t0, t0 + 3);
            m->indices.insert(m->indices.end(), t1, t1 + 3);
        }
    *out = m;
    return CWBVH_OK;
}

int32_t cwbvh_mesh_icosphere_grid(int32_t gx, int32_t gy, int32_t gz, int32_t frequency, float spacing,
                                  float scale_min, float scale_max, float ground_margin, uint64_t seed,
                                  cwbvh_mesh** out) {
    CW_CHECK_ARG(out != nullptr, "out is NULL");
    CW_CHECK_ARG(gx >= 1 && gy >= 1 && gz >= 1 && frequency >= 1 && frequency <= 256, "bad grid");
    CW_CHECK_ARG(spacing > 0 && scale_min > 0 && scale_max >= scale_min, "bad scale");
    std::vector<V3> sv;
    std::vector<std::array<int, 3>> st;
    geodesic_sphere(frequency, sv, st);
    cwbvh_mesh* m = new_mesh();
    int64_t ns = (int64_t)gx * gy * gz;
    m->vertices.reserve((size_t)(ns * sv.size() * 3 + 12));
    m->indices.reserve((size_t)(ns * st.size() * 3 + 6));
    Rng rng(seed, 2);
    for (int k = 0; k < gz; ++k)
        for (int j = 0; j < gy; ++j)
            for (int i = 0; i < gx; ++i) {
                float R[9];
                random_rotation(rng, R);
                float s = scale_min + (scale_max - scale_min) * rng.uniform();
                V3 c{i * spacing, j * spacing, k * spacing};
                int32_t base = (int32_t)(m->vertices.size() / 3);
                for (const V3& p : sv) {
                    V3 q{R[0] * p.x + R[1] * p.y + R[2] * p.z, R[3] * p.x + R[4] * p.y + R[5] * p.z,
                         R[6] * p.x + R[7] * p.y + R[8] * p.z};
                    V3 w = c + q * s;
                    m->vertices.push_back(w.x);
                    m->vertices.push_back(w.y);
                    m->vertices.push_back(w.z);
                }
                for (auto& t : st) {
                    m->indices.push_back(base + t[0]);
                    m->indices.push_back(base + t[1]);
                    m->indices.push_back(base + t[2]);
                }
            }
    if (ground_margin >= 0.f) {
        // square below the lattice, extending ground_margin lattice extents
        // beyond it on every side
        float z = -0.5f * spacing - scale_max;
        float ex = (gx - 1) * spacing, ey = (gy - 1) * spacing;
        float x0 = -ground_margin * ex - spacing, y0 = -ground_margin * ey - spacing;
        float x1 = (1.0f + ground_margin) * ex + spacing, y1 = (1.0f + ground_margin) * ey + spacing;
        int32_t base = (int32_t)(m->vertices.size() / 3);
        float q[12] = {x0, y0, z, x1, y0, z, x1, y1, z, x0, y1, z};
        m->vertices.insert(m->vertices.end(), q, q + 12);
        int32_t t[6] = {base, base + 1, base + 2, base, base + 2, base + 3};
        m->indices.insert(m->indices.end(), t, t + 6);
    }
    *out = m;
    return CWBVH_OK;
}

int32_t cwbvh_mesh_cluster_soup(int64_t num_triangles, int32_t num_clusters, float sigma_min, float sigma_max,
                                float edge_mu, float edge_sigma, uint64_t seed, cwbvh_mesh** out) {
    CW_CHECK_ARG(out != nullptr, "out is NULL");
    CW_CHECK_ARG(num_triangles >= 1 && num_triangles <= (int64_t)700000000 && num_clusters >= 1, "bad soup size");
    CW_CHECK_ARG(sigma_min > 0 && sigma_max >= sigma_min, "bad sigma");
    std::vector<V3> center(num_clusters);
    std::vector<float> sigma(num_clusters);
    Rng rng(seed, 3);
    for (int c = 0; c < num_clusters; ++c) {
        center[c] = V3{rng.uniform(), rng.uniform(), rng.uniform()};
        double ls = std::log((double)sigma_min) + (std::log((double)sigma_max) - std::log((double)sigma_min)) * rng.uniform_d();
        sigma[c] = (float)std::exp(ls);
    }
    cwbvh_mesh* m = new_mesh();
    m->vertices.resize((size_t)num_triangles * 9);
    m->indices.resize((size_t)num_triangles * 3);
    int threads = num_threads_or_default(0);
    parallel_for(num_triangles, threads, [&](int64_t b, int64_t e) {
        for (int64_t t = b; t < e; ++t) {
            Rng r(seed ^ 0x5eed5eedULL, (uint64_t)t);
            int c = (int)(t % num_clusters);
            V3 p{center[c].x + sigma[c] * r.normal(), center[c].y + sigma[c] * r.normal(),
                 center[c].z + sigma[c] * r.normal()};
            p = vmax(V3{0, 0, 0}, vmin(V3{1, 1, 1}, p));
            float l1 = (float)std::exp(edge_mu + edge_sigma * r.normal());
            float l2 = (float)std::exp(edge_mu + edge_sigma * r.normal());
            double ang = kPi / 6 + (2 * kPi / 3) * r.uniform_d();
            float R[9];
            random_rotation(r, R);
            V3 ex{R[0], R[3], R[6]}, ey{R[1], R[4], R[7]};
            V3 e1 = ex * l1;
            V3 e2 = (ex * (float)std::cos(ang) + ey * (float)std::sin(ang)) * l2;
            V3 v0 = p - (e1 + e2) * (1.0f / 3.0f);
            V3 v1 = v0 + e1, v2 = v0 + e2;
            float* dst = &m->vertices[(size_t)t * 9];
            dst[0] = v0.x; dst[1] = v0.y; dst[2] = v0.z;
            dst[3] = v1.x; dst[4] = v1.y; dst[5] = v1.z;
            dst[6] = v2.x; dst[7] = v2.y; dst[8] = v2.z;
            m->indices[(size_t)t * 3 + 0] = (int32_t)(3 * t);
            m->indices[(size_t)t * 3 + 1] = (int32_t)(3 * t + 1);
            m->indices[(size_t)t * 3 + 2] = (int32_t)(3 * t + 2);
        }
    });
    *out = m;
    return CWBVH_OK;
}

int32_t cwbvh_mesh_validation_planes(int32_t count, cwbvh_mesh** out) {
    CW_CHECK_ARG(out != nullptr, "out is NULL");
    CW_CHECK_ARG(count >= 1 && count <= (1 << 24), "bad count");
    cwbvh_mesh* m = new_mesh();
    m->vertices.resize((size_t)count * 9);
    m->indices.resize((size_t)count * 3);
    for (int32_t i = 0; i < count; ++i) {
        float z = (float)i;
        float v[9] = {-1.1f, -1.1f, z, 1.1f, -1.1f, z, 1.1f, 1.1f, z};
        std::memcpy(&m->vertices[(size_t)i * 9], v, sizeof(v));
        m->indices[(size_t)i * 3 + 0] = 3 * i;
        m->indices[(size_t)i * 3 + 1] = 3 * i + 1;
        m->indices[(size_t)i * 3 + 2] = 3 * i + 2;
    }
    *out = m;
    return CWBVH_OK;
}

int32_t cwbvh_rays_ao_surface(const cwbvh_mesh* mesh, int64_t num_points, int32_t rays_per_point, float offset,
                              float tmin, float tmax, uint64_t seed, cwbvh_ray* out) {
    CW_CHECK_ARG(mesh != nullptr && out != nullptr, "NULL argument");
    CW_CHECK_ARG(num_points >= 0 && rays_per_point >= 1, "bad ray count");
    CW_CHECK_ARG(mesh->nt() > 0, "empty mesh");
    AreaSampler sampler(mesh);
    parallel_for(num_points, num_threads_or_default(0), [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) {
            Rng rp(seed, (uint64_t)i);
            V3 p, n;
            sampler.sample(rp, p, n);
            V3 o = p + n * offset, t, bt;
            onb(n, t, bt);
            for (int k = 0; k < rays_per_point; ++k) {
                Rng rr(seed ^ 0xa0a0a0a0a0a0a0a0ULL, (uint64_t)(i * rays_per_point + k));
                float u1 = rr.uniform(), u2 = rr.uniform();
                V3 l = cosine_hemisphere(u1, u2);
                V3 d = normalize(t * l.x + bt * l.y + n * l.z);
                put_ray(out[i * rays_per_point + k], o, tmin, d, tmax);
            }
        }
    });
    return CWBVH_OK;
}

int32_t cwbvh_rays_ao_from_samples(const cwbvh_float4* positions, const cwbvh_float4* normals, const float* radius,
                                   int64_t num_texels, int32_t rays_per_texel, float tmin, float tmax, uint64_t seed,
                                   cwbvh_ray* out) {
    CW_CHECK_ARG(num_texels >= 0 && rays_per_texel >= 1, "bad ray count");
    CW_CHECK_ARG(num_texels == 0 || (positions && normals && radius && out), "NULL argument");
    // the same arithmetic as the device generator (aogen.h), so host and
    // device rays are bit-identical
    parallel_for(num_texels, num_threads_or_default(0), [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) {
            const cwao::F3 p{positions[i].x, positions[i].y, positions[i].z};
            const cwao::F3 n{normals[i].x, normals[i].y, normals[i].z};
            for (int32_t k = 0; k < rays_per_texel; ++k) {
                const int64_t r = i * rays_per_texel + k;
                cwao::F3 o, d;
                cwao::surfel_ray(p, n, radius[i], seed, (uint64_t)r, o, d);
                put_ray(out[r], V3{o.x, o.y, o.z}, tmin, V3{d.x, d.y, d.z}, tmax);
            }
        }
    });
    return CWBVH_OK;
}

int32_t cwbvh_rays_pinhole(int32_t width, int32_t height, const float* eye3, const float* target3, const float* up3,
                           float fov_deg, float tmin, float tmax, cwbvh_ray* out) {
    CW_CHECK_ARG(out != nullptr && eye3 && target3 && up3, "NULL argument");
    CW_CHECK_ARG(width >= 1 && height >= 1 && fov_deg > 0 && fov_deg < 180, "bad camera");
    V3 eye{eye3[0], eye3[1], eye3[2]}, tgt{target3[0], target3[1], target3[2]}, up{up3[0], up3[1], up3[2]};
    V3 fwd = normalize(tgt - eye);
    V3 right = normalize(cross(fwd, up));
    V3 vup = cross(right, fwd);
    float th = (float)std::tan(0.5 * fov_deg * kPi / 180.0);
    float aspect = (float)width / (float)height;
    parallel_for((int64_t)width * height, num_threads_or_default(0), [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) {
            int64_t y = i / width, x = i % width;
            float sx = (2.0f * ((float)x + 0.5f) / (float)width - 1.0f) * th * aspect;
            float sy = (1.0f - 2.0f * ((float)y + 0.5f) / (float)height) * th;
            V3 d = normalize(fwd + right * sx + vup * sy);
            put_ray(out[i], eye, tmin, d, tmax);
        }
    });
    return CWBVH_OK;
}

int32_t cwbvh_rays_ao_from_hits(const cwbvh_mesh* mesh, const cwbvh_ray* primary, const cwbvh_hit* hits,
                                const int32_t* prim_ids, int64_t n, int32_t rays_per_hit, float offset, float tmax,
                                uint64_t seed, cwbvh_ray* out, int64_t* out_count) {
    CW_CHECK_ARG(mesh && primary && hits && prim_ids && out && out_count, "NULL argument");
    CW_CHECK_ARG(n >= 0 && rays_per_hit >= 1, "bad count");
    // exclusive prefix over hits so that the output is dense and ordered
    std::vector<int64_t> slot(n + 1, 0);
    for (int64_t i = 0; i < n; ++i) {
        int32_t p = prim_ids[i];
        CW_CHECK_ARG(p < (int32_t)mesh->nt(), "prim id out of range");
        slot[i + 1] = slot[i] + (p >= 0 ? 1 : 0);
    }
    parallel_for(n, num_threads_or_default(0), [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) {
            int32_t p = prim_ids[i];
            if (p < 0) continue;
            const cwbvh_ray& r = primary[i];
            V3 o{r.origin_tmin.x, r.origin_tmin.y, r.origin_tmin.z};
            V3 d{r.dir_tmax.x, r.dir_tmax.y, r.dir_tmax.z};
            float t = hits[i].t_triId_u_v.x;
            V3 a, bb, c;
            mesh->tri(p, a, bb, c);
            V3 nrm = tri_normal(a, bb, c);
            if (dot(nrm, d) > 0) nrm = nrm * -1.0f;
            V3 q = o + d * t + nrm * offset, tt, bt;
            onb(nrm, tt, bt);
            for (int k = 0; k < rays_per_hit; ++k) {
                Rng rr(seed, (uint64_t)(i * rays_per_hit + k));
                float u1 = rr.uniform(), u2 = rr.uniform();
                V3 l = cosine_hemisphere(u1, u2);
                V3 dir = normalize(tt * l.x + bt * l.y + nrm * l.z);
                put_ray(out[slot[i] * rays_per_hit + k], q, 0.0f, dir, tmax);
            }
        }
    });
    *out_count = slot[n] * rays_per_hit;
    return CWBVH_OK;
}

int32_t cwbvh_rays_sphere_from_surface(const cwbvh_mesh* mesh, int64_t num_rays, float offset, float tmin,
                                       float tmax, uint64_t seed, cwbvh_ray* out) {
    CW_CHECK_ARG(mesh != nullptr && out != nullptr, "NULL argument");
    CW_CHECK_ARG(num_rays >= 0 && mesh->nt() > 0, "bad arguments");
    AreaSampler sampler(mesh);
    parallel_for(num_rays, num_threads_or_default(0), [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) {
            Rng r(seed, (uint64_t)i);
            V3 p, n;
            sampler.sample(r, p, n);
            float u1 = r.uniform(), u2 = r.uniform();
            V3 d = uniform_sphere(u1, u2);
            if (dot(d, n) < 0) n = n * -1.0f;
            put_ray(out[i], p + n * offset, tmin, d, tmax);
        }
    });
    return CWBVH_OK;
}

int32_t cwbvh_rays_validation(int64_t count, int32_t num_triangles, cwbvh_ray* out) {
    CW_CHECK_ARG(out != nullptr && count >= 0 && num_triangles >= 1, "bad arguments");
    V3 d = normalize(V3{-0.01f, 0.5f, 1.0f});
    for (int64_t r = 0; r < count; ++r) {
        int64_t k = (int64_t)(((uint64_t)r * 12345678ULL) % (uint64_t)num_triangles);
        put_ray(out[r], V3{0.5f, -0.5f, (float)k - 0.5f}, 0.0f, d, 1e20f);
    }
    return CWBVH_OK;
}

int32_t cwbvh_rays_shuffle(cwbvh_ray* rays, int64_t n, uint64_t seed) {
    CW_CHECK_ARG(rays != nullptr || n == 0, "rays is NULL");
    Rng r(seed, 7);
    for (int64_t i = n - 1; i > 0; --i) {
        int64_t j = (int64_t)(r.next() % (uint64_t)(i + 1));
        std::swap(rays[i], rays[j]);
    }
    return CWBVH_OK;
}

}  // extern "C"
