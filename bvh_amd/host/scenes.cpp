// scenes.cpp — seeded synthetic meshes and ray sets (include/bvh_scenes.h).
// Compiled with -ffp-contract=off so every platform produces identical bits.
#include "bvh_scenes.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace {

inline uint64_t splitmix64(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (auto& w : s) w = splitmix64(x);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    float u01() { return (float)(next() >> 40) * (1.0f / 16777216.0f); }
    double ud() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    // Standard normal by Box-Muller in double.
    double gauss() {
        double u1 = 1.0 - ud();  // (0, 1]
        double u2 = ud();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
    void unit(double out[3]) {
        double z = 2.0 * ud() - 1.0;
        double phi = 6.283185307179586 * ud();
        double r = std::sqrt(std::fmax(0.0, 1.0 - z * z));
        out[0] = r * std::cos(phi);
        out[1] = r * std::sin(phi);
        out[2] = z;
    }
    uint32_t below(uint32_t m) { return (uint32_t)((next() >> 32) % m); }
};

inline void put(float* const* soa, uint32_t i, int vtx, float x, float y, float z) {
    soa[3 * vtx + 0][i] = x;
    soa[3 * vtx + 1][i] = y;
    soa[3 * vtx + 2][i] = z;
}

void gen_uniform(Rng& g, uint32_t n, float* const* soa) {
    for (uint32_t i = 0; i < n; ++i) {
        float cx = g.u01(), cy = g.u01(), cz = g.u01();
        for (int v = 0; v < 3; ++v) {
            float ox = g.u01() * 0.01f - 0.005f;
            float oy = g.u01() * 0.01f - 0.005f;
            float oz = g.u01() * 0.01f - 0.005f;
            put(soa, i, v, cx + ox, cy + oy, cz + oz);
        }
    }
}

inline float dotf(const float* a, const float* b) {
    float s = a[0] * b[0];
    s = s + a[1] * b[1];
    s = s + a[2] * b[2];
    return s;
}
inline void crossf(const float* a, const float* b, float* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
// d_normalize (render_engine.cu:130-135).
inline void normalizef(float* v) {
    float len = std::sqrt(dotf(v, v));
    if (len < 1e-12f) { v[0] = v[1] = v[2] = 0.0f; return; }
    float inv = 1.0f / len;
    v[0] = v[0] * inv;
    v[1] = v[1] * inv;
    v[2] = v[2] * inv;
}

}  // namespace

extern "C" {

uint64_t bvh_fnv1a64(const void* data, size_t bytes, uint64_t h) {
    const unsigned char* p = (const unsigned char*)data;
    for (size_t i = 0; i < bytes; ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

int bvh_rng_u64(uint64_t seed, uint32_t n, uint64_t* out) {
    if (!out) return 1;
    Rng g(seed);
    for (uint32_t i = 0; i < n; ++i) out[i] = g.next();
    return 0;
}

int bvh_scene_uniform(uint64_t seed, uint32_t n, float* const* soa) {
    if (!soa) return 1;
    Rng g(seed);
    gen_uniform(g, n, soa);
    return 0;
}

int bvh_scene_blobs(uint64_t seed, uint32_t n, float* const* soa) {
    if (!soa) return 1;
    Rng g(seed);
    double centres[64][3];
    for (auto& c : centres)
        for (double& x : c) x = (double)g.u01();
    const uint32_t n_blob = (uint32_t)((uint64_t)n * 95 / 100);
    for (uint32_t i = 0; i < n; ++i) {
        if (i < n_blob) {
            uint32_t b = g.below(64);
            float c[3];
            for (int k = 0; k < 3; ++k) c[k] = (float)(centres[b][k] + 0.02 * g.gauss());
            for (int v = 0; v < 3; ++v) {
                float ox = g.u01() * 0.006f - 0.003f;
                float oy = g.u01() * 0.006f - 0.003f;
                float oz = g.u01() * 0.006f - 0.003f;
                put(soa, i, v, c[0] + ox, c[1] + oy, c[2] + oz);
            }
        } else {
            double c[3] = {(double)g.u01(), (double)g.u01(), (double)g.u01()};
            double e1[3], e2[3];
            g.unit(e1);
            g.unit(e2);
            put(soa, i, 0, (float)c[0], (float)c[1], (float)c[2]);
            put(soa, i, 1, (float)(c[0] + 0.1 * e1[0]), (float)(c[1] + 0.1 * e1[1]), (float)(c[2] + 0.1 * e1[2]));
            put(soa, i, 2, (float)(c[0] + 0.1 * e2[0]), (float)(c[1] + 0.1 * e2[1]), (float)(c[2] + 0.1 * e2[2]));
        }
    }
    // Fisher-Yates shuffle of whole triangles.
    for (uint32_t i = n; i > 1; --i) {
        uint32_t j = (uint32_t)(g.next() % i);
        uint32_t a = i - 1;
        if (j != a)
            for (int k = 0; k < 9; ++k) {
                float t = soa[k][a];
                soa[k][a] = soa[k][j];
                soa[k][j] = t;
            }
    }
    return 0;
}

int bvh_scene_strands(uint64_t seed, uint32_t strands, uint32_t steps, float* const* soa) {
    if (!soa) return 1;
    Rng g(seed);
    uint32_t i = 0;
    for (uint32_t s = 0; s < strands; ++s) {
        double p[3] = {(double)g.u01(), (double)g.u01(), (double)g.u01()};
        for (uint32_t k = 0; k < steps; ++k) {
            double d[3];
            g.unit(d);
            double q[3];
            for (int a = 0; a < 3; ++a) {
                q[a] = p[a] + 0.01 * d[a];
                if (q[a] < 0.0) q[a] = -q[a];
                if (q[a] > 1.0) q[a] = 2.0 - q[a];
            }
            double seg[3] = {q[0] - p[0], q[1] - p[1], q[2] - p[2]};
            double ax[3] = {1.0, 0.0, 0.0};
            double sl = std::sqrt(seg[0] * seg[0] + seg[1] * seg[1] + seg[2] * seg[2]);
            if (sl > 0 && std::fabs(seg[0]) / sl > 0.9) { ax[0] = 0.0; ax[1] = 1.0; }
            double w[3] = {seg[1] * ax[2] - seg[2] * ax[1], seg[2] * ax[0] - seg[0] * ax[2],
                           seg[0] * ax[1] - seg[1] * ax[0]};
            double wl = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
            if (wl > 0) for (double& x : w) x /= wl;
            put(soa, i, 0, (float)p[0], (float)p[1], (float)p[2]);
            put(soa, i, 1, (float)q[0], (float)q[1], (float)q[2]);
            put(soa, i, 2, (float)(p[0] + 0.0005 * w[0]), (float)(p[1] + 0.0005 * w[1]), (float)(p[2] + 0.0005 * w[2]));
            ++i;
            for (int a = 0; a < 3; ++a) p[a] = q[a];
        }
    }
    return 0;
}

int bvh_scene_batch_counts(uint64_t seed, uint32_t meshes, uint32_t lo, uint32_t hi, uint32_t* counts) {
    if (!counts || lo == 0 || hi < lo) return 1;
    Rng g(seed);
    const double la = std::log((double)lo), lr = std::log((double)hi / (double)lo);
    for (uint32_t i = 0; i < meshes; ++i) {
        double c = std::floor(std::exp(la + g.ud() * lr));
        if (c < lo) c = lo;
        if (c > hi) c = hi;
        counts[i] = (uint32_t)c;
    }
    return 0;
}

int bvh_scene_batch_mesh(uint64_t seed, uint32_t index, uint32_t n, float* const* soa) {
    if (!soa) return 1;
    // Scale and offset of mesh `index` come from a stream keyed by (seed, index).
    Rng t(seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull + index);
    float scale = 0.1f + t.u01() * 9.9f;
    float off[3];
    for (float& o : off) o = t.u01() * 200.0f - 100.0f;
    Rng g(seed ^ (uint64_t)index);
    gen_uniform(g, n, soa);
    for (uint32_t i = 0; i < n; ++i)
        for (int v = 0; v < 3; ++v)
            for (int a = 0; a < 3; ++a) {
                float* p = &soa[3 * v + a][i];
                *p = *p * scale + off[a];
            }
    return 0;
}

int bvh_rays_camera(const float* eye, const float* look_at, const float* up, float fov_y_deg, uint32_t width,
                    uint32_t height, float* rays) {
    if (!eye || !look_at || !up || !rays || width == 0 || height == 0) return 1;
    float fwd[3] = {look_at[0] - eye[0], look_at[1] - eye[1], look_at[2] - eye[2]};
    normalizef(fwd);
    float right[3];
    crossf(fwd, up, right);
    normalizef(right);
    float camUp[3];
    crossf(right, fwd, camUp);
    float fovRad = fov_y_deg * 3.14159265358979f / 180.0f;
    float halfH = std::tan(fovRad * 0.5f);
    float aspect = (float)width / (float)height;
    float halfW = halfH * aspect;
    for (uint32_t y = 0; y < height; ++y) {
        for (uint32_t x = 0; x < width; ++x) {
            float u = (2.0f * ((float)x + 0.5f) / (float)width - 1.0f) * halfW;
            float v = (2.0f * ((float)y + 0.5f) / (float)height - 1.0f) * halfH;
            float d[3];
            for (int a = 0; a < 3; ++a) {
                float t = fwd[a] + u * right[a];
                d[a] = t - v * camUp[a];
            }
            normalizef(d);
            float* r = rays + 6 * ((size_t)y * width + x);
            r[0] = eye[0]; r[1] = eye[1]; r[2] = eye[2];
            r[3] = d[0]; r[4] = d[1]; r[5] = d[2];
        }
    }
    return 0;
}

int bvh_rays_sphere_camera(const float* box, float fov_y_deg, uint32_t width, uint32_t height, float* rays) {
    if (!box) return 1;
    float c[3], e[3];
    for (int a = 0; a < 3; ++a) {
        c[a] = (box[a] + box[3 + a]) * 0.5f;
        e[a] = box[3 + a] - box[a];
    }
    float radius = std::sqrt(dotf(e, e)) * 0.5f;
    float dir[3] = {0.6f, 0.3f, 0.7f};
    normalizef(dir);
    float eye[3] = {c[0] + radius * dir[0], c[1] + radius * dir[1], c[2] + radius * dir[2]};
    float up[3] = {0.0f, 1.0f, 0.0f};
    return bvh_rays_camera(eye, c, up, fov_y_deg, width, height, rays);
}

int bvh_rays_sphere_to_box(uint64_t seed, const float* box, uint32_t n, float* rays) {
    if (!box || !rays) return 1;
    Rng g(seed);
    double c[3], e[3];
    for (int a = 0; a < 3; ++a) {
        c[a] = 0.5 * ((double)box[a] + (double)box[3 + a]);
        e[a] = (double)box[3 + a] - (double)box[a];
    }
    double radius = 0.5 * std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
    for (uint32_t i = 0; i < n; ++i) {
        double u[3];
        g.unit(u);
        double o[3] = {c[0] + radius * u[0], c[1] + radius * u[1], c[2] + radius * u[2]};
        double t[3];
        for (int a = 0; a < 3; ++a) t[a] = (double)box[a] + g.ud() * e[a];
        double d[3] = {t[0] - o[0], t[1] - o[1], t[2] - o[2]};
        double dl = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (dl <= 0) { d[0] = 1; d[1] = 0; d[2] = 0; dl = 1; }
        float* r = rays + 6 * (size_t)i;
        for (int a = 0; a < 3; ++a) { r[a] = (float)o[a]; r[3 + a] = (float)(d[a] / dl); }
    }
    return 0;
}

int bvh_rays_uniform_cube(uint64_t seed, uint32_t n, float* rays) {
    if (!rays) return 1;
    Rng g(seed);
    for (uint32_t i = 0; i < n; ++i) {
        float* r = rays + 6 * (size_t)i;
        r[0] = g.u01(); r[1] = g.u01(); r[2] = g.u01();
        double d[3];
        g.unit(d);
        r[3] = (float)d[0]; r[4] = (float)d[1]; r[5] = (float)d[2];
    }
    return 0;
}

}  // extern "C"
