// render.cpp — the renderer API of include/render_engine.h over bvh_trace.
//
// Rays are generated on the host (bvh_rays_camera, the reference's pinhole
// model), traced on the GPU with per-ray counters, and shaded on the host.
// The heatmap uses the published polynomial fit of Google's Turbo colormap.
#include "../../include/render_engine.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <stdexcept>

#include "../../include/bvh_hip.h"
#include "../../include/bvh_scenes.h"

namespace {

constexpr float kBg[3] = {0.15f, 0.15f, 0.18f};

float len3(float x, float y, float z) { return std::sqrt(x * x + y * y + z * z); }

// Turbo colormap, polynomial approximation (Google AI blog, 2019).
void turbo(float t, float& r, float& g, float& b) {
    t = std::min(1.0f, std::max(0.0f, t));
    const double x = t, x2 = x * x, x3 = x2 * x, x4 = x3 * x, x5 = x4 * x;
    r = (float)(0.13572138 + 4.61539260 * x - 42.66032258 * x2 + 132.13108234 * x3 - 152.94239396 * x4 +
                59.28637943 * x5);
    g = (float)(0.09140261 + 2.19418839 * x + 4.84296658 * x2 - 14.18503333 * x3 + 4.27729857 * x4 +
                2.82956604 * x5);
    b = (float)(0.10667330 + 12.64194608 * x - 60.58204836 * x2 + 110.36276771 * x3 - 89.90310912 * x4 +
                27.34824973 * x5);
    r = std::min(1.0f, std::max(0.0f, r));
    g = std::min(1.0f, std::max(0.0f, g));
    b = std::min(1.0f, std::max(0.0f, b));
}

std::vector<float> floats(const std::string& s) {
    std::vector<float> v;
    std::string cur;
    for (char c : s) {
        if (c == ',') {
            if (!cur.empty()) v.push_back(std::stof(cur));
            cur.clear();
        } else {
            cur += c;
        }
    }
    if (!cur.empty()) v.push_back(std::stof(cur));
    return v;
}

bool write_ppm(const std::string& file, const std::vector<float>& rgb, int w, int h) {
    std::ofstream f(file, std::ios::binary);
    if (!f) {
        std::cerr << "Error: Cannot open " << file << " for writing\n";
        return false;
    }
    f << "P6\n" << w << " " << h << "\n255\n";
    std::vector<unsigned char> px(rgb.size());
    for (size_t i = 0; i < rgb.size(); ++i)
        px[i] = (unsigned char)(std::min(1.0f, std::max(0.0f, rgb[i])) * 255.0f + 0.5f);
    f.write(reinterpret_cast<const char*>(px.data()), (std::streamsize)px.size());
    return true;
}

}  // namespace

Camera autoCamera(const std::vector<BVHNode>& nodes) {
    Camera cam;
    cam.fovY = 60.0f;
    cam.up = float3_cw(0, 1, 0);
    if (nodes.empty()) return cam;
    const AABB_cw& b = nodes[0].bbox;
    float3_cw c((b.min.x + b.max.x) * 0.5f, (b.min.y + b.max.y) * 0.5f, (b.min.z + b.max.z) * 0.5f);
    float diag = len3(b.max.x - b.min.x, b.max.y - b.min.y, b.max.z - b.min.z);
    if (diag < 1e-6f) diag = 2.0f;
    const float fov = cam.fovY * 3.14159265358979f / 180.0f;
    const float dist = (diag * 0.5f) / std::tan(fov * 0.5f) * 1.2f;
    cam.eye = float3_cw(c.x + dist * 0.6f, c.y + dist * 0.3f, c.z + dist * 0.7f);
    cam.lookAt = c;
    return cam;
}

Camera parseCameraString(const std::string& cameraStr, const std::string& upStr, float fov,
                         const std::vector<BVHNode>& nodes) {
    Camera cam = autoCamera(nodes);
    if (!cameraStr.empty()) {
        auto v = floats(cameraStr);
        if (v.size() >= 6) {
            cam.eye = float3_cw(v[0], v[1], v[2]);
            cam.lookAt = float3_cw(v[3], v[4], v[5]);
        } else {
            std::cerr << "Warning: --camera expects 6 values (ex,ey,ez,lx,ly,lz). Using auto-camera.\n";
        }
    }
    if (!upStr.empty()) {
        auto v = floats(upStr);
        if (v.size() >= 3) cam.up = float3_cw(v[0], v[1], v[2]);
    }
    if (fov > 0.0f) cam.fovY = fov;
    return cam;
}

ShadingMode parseShadingMode(const std::string& s) {
    if (s == "normal") return ShadingMode::NORMAL;
    if (s == "depth") return ShadingMode::DEPTH;
    if (s == "diffuse") return ShadingMode::DIFFUSE;
    if (s == "heatmap") return ShadingMode::HEATMAP;
    std::cerr << "Warning: Unknown shading mode '" << s << "'. Using 'normal'.\n";
    return ShadingMode::NORMAL;
}

RenderStats renderImage(const std::vector<BVHNode>& nodes, const TriangleMesh& mesh, int width, int height,
                        const Camera& camera, ShadingMode mode, const std::string& outputFile,
                        float fixedHeatmapMax) {
    RenderStats stats = {};
    stats.width = width;
    stats.height = height;
    stats.totalRays = (uint64_t)width * (uint64_t)height;
    if (nodes.empty() || mesh.size() == 0 || width <= 0 || height <= 0) {
        std::cerr << "Error: Empty BVH or mesh, cannot render\n";
        return stats;
    }
    const size_t np = (size_t)width * height;
    std::vector<float> rays(np * 6);
    const float eye[3] = {camera.eye.x, camera.eye.y, camera.eye.z};
    const float at[3] = {camera.lookAt.x, camera.lookAt.y, camera.lookAt.z};
    const float up[3] = {camera.up.x, camera.up.y, camera.up.z};
    bvh_rays_camera(eye, at, up, camera.fovY, (uint32_t)width, (uint32_t)height, rays.data());

    bvh_ctx_t ctx = nullptr;
    if (bvh_ctx_create(0, &ctx) != BVH_OK) throw std::runtime_error("renderImage: no HIP device");
    const float* soa[9] = {mesh.v0x.data(), mesh.v0y.data(), mesh.v0z.data(), mesh.v1x.data(), mesh.v1y.data(),
                           mesh.v1z.data(), mesh.v2x.data(), mesh.v2y.data(), mesh.v2z.data()};
    bvh_mesh_t dm = nullptr;
    bvh_tree_t tree = nullptr;
    std::vector<float> t(np);
    std::vector<int32_t> prim(np), cnt(np * 3);
    float ms = 0.0f;
    int st = bvh_mesh_upload(ctx, soa, (uint32_t)mesh.size(), &dm);
    if (st == BVH_OK) st = bvh_tree_create(ctx, &tree);
    if (st == BVH_OK) st = bvh_tree_from_nodes(ctx, nodes.data(), (uint32_t)nodes.size(), tree);
    if (st == BVH_OK)
        st = bvh_trace_host(ctx, tree, dm, rays.data(), (uint32_t)np, t.data(), prim.data(), cnt.data(), &ms);
    std::string err = st == BVH_OK ? "" : bvh_last_error(ctx);
    if (tree) bvh_tree_free(ctx, tree);
    if (dm) bvh_mesh_free(ctx, dm);
    bvh_ctx_destroy(ctx);
    if (st != BVH_OK) throw std::runtime_error("renderImage: " + err);
    stats.renderTimeMs = ms;

    const AABB_cw& rb = nodes[0].bbox;
    float diag = len3(rb.max.x - rb.min.x, rb.max.y - rb.min.y, rb.max.z - rb.min.z);
    if (diag < 1e-6f) diag = 1.0f;
    float light[3] = {0.5f, 0.8f, 0.6f};
    const float ll = len3(light[0], light[1], light[2]);
    for (float& x : light) x /= ll;

    double sn = 0, sa = 0, stt = 0;
    int mx = 0;
    uint64_t counted = 0;
    std::vector<int> visits;
    visits.reserve(np);
    for (size_t i = 0; i < np; ++i) {
        if (cnt[3 * i] > 0) {
            sn += cnt[3 * i];
            sa += cnt[3 * i + 1];
            stt += cnt[3 * i + 2];
            mx = std::max(mx, cnt[3 * i]);
            ++counted;
            visits.push_back(cnt[3 * i]);
        }
    }
    if (counted) {
        stats.avgNodesVisited = (float)(sn / counted);
        stats.avgAABBTests = (float)(sa / counted);
        stats.avgTriTests = (float)(stt / counted);
    }
    stats.maxNodesVisited = (float)mx;

    float hmax = fixedHeatmapMax;
    if (mode == ShadingMode::HEATMAP && hmax <= 0.0f) {
        hmax = 128.0f;
        if (!visits.empty()) {
            std::sort(visits.begin(), visits.end());
            size_t k = std::min(visits.size() - 1, (size_t)(visits.size() * 0.999));
            hmax = std::max(1.0f, (float)visits[k]);
        }
    }
    std::vector<float> rgb(np * 3);
    for (size_t i = 0; i < np; ++i) {
        float r = kBg[0], g = kBg[1], b = kBg[2];
        if (mode == ShadingMode::HEATMAP) {
            if (cnt[3 * i] > 0) turbo((float)cnt[3 * i] / hmax, r, g, b);
        } else if (prim[i] >= 0) {
            const size_t p = (size_t)prim[i];
            float e1[3] = {mesh.v1x[p] - mesh.v0x[p], mesh.v1y[p] - mesh.v0y[p], mesh.v1z[p] - mesh.v0z[p]};
            float e2[3] = {mesh.v2x[p] - mesh.v0x[p], mesh.v2y[p] - mesh.v0y[p], mesh.v2z[p] - mesh.v0z[p]};
            float n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
            float nl = len3(n[0], n[1], n[2]);
            if (nl > 1e-12f)
                for (float& x : n) x /= nl;
            if (mode == ShadingMode::NORMAL) {
                r = (n[0] + 1.0f) * 0.5f;
                g = (n[1] + 1.0f) * 0.5f;
                b = (n[2] + 1.0f) * 0.5f;
            } else if (mode == ShadingMode::DEPTH) {
                float v = 1.0f - std::min(1.0f, std::max(0.0f, t[i] / diag));
                r = g = b = v;
            } else {
                const float* d = &rays[6 * i + 3];
                if (n[0] * d[0] + n[1] * d[1] + n[2] * d[2] > 0.0f)
                    for (float& x : n) x = -x;
                float diff = std::max(0.0f, n[0] * light[0] + n[1] * light[1] + n[2] * light[2]);
                float shade = std::min(1.0f, 0.15f + diff * 0.85f);
                r = shade * (0.7f + 0.3f * std::fabs(n[0]));
                g = shade * (0.7f + 0.3f * std::fabs(n[1]));
                b = shade * (0.7f + 0.3f * std::fabs(n[2]));
            }
        }
        rgb[3 * i] = r;
        rgb[3 * i + 1] = g;
        rgb[3 * i + 2] = b;
    }
    if (!outputFile.empty() && write_ppm(outputFile, rgb, width, height))
        std::cout << "  Saved: " << outputFile << " (" << width << "x" << height << ")\n";
    return stats;
}

void printRenderStats(const std::string& name, const RenderStats& s) {
    std::cout << "  --- Render: " << name << " ---\n";
    std::cout << "    Resolution:         " << s.width << " x " << s.height << "\n";
    std::cout << "    Render Time:        " << std::fixed << std::setprecision(3) << s.renderTimeMs << " ms\n";
    std::cout << "    Total Rays:         " << s.totalRays << "\n";
    if (s.renderTimeMs > 0.0f)
        std::cout << "    Rays/sec:           " << std::fixed << std::setprecision(2)
                  << (double)s.totalRays / (s.renderTimeMs / 1000.0) / 1e6 << " M\n";
    std::cout << "    Avg Nodes Visited:  " << std::fixed << std::setprecision(1) << s.avgNodesVisited << "\n";
    std::cout << "    Avg AABB Tests:     " << std::fixed << std::setprecision(1) << s.avgAABBTests << "\n";
    std::cout << "    Avg Tri Tests:      " << std::fixed << std::setprecision(1) << s.avgTriTests << "\n";
    std::cout << "    Max Nodes Visited:  " << std::fixed << std::setprecision(0) << s.maxNodesVisited << "\n";
}
