// render.cpp — the renderer API of include/render_engine.h over bvh_trace.
//
// Rays are generated on the host (bvh_rays_camera, the reference's pinhole
// model), traced on the GPU with per-ray counters, and shaded on the host.
// Shading follows the reference kernel and host post-process
// (render_engine.cu:362-446, :541-593) operation for operation.
#include "../../include/render_engine.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <stdexcept>

#include "../../include/bvh_hip.h"
#include "../../include/bvh_render.h"
#include "../../include/bvh_scenes.h"

namespace {

constexpr float kBg[3] = {0.15f, 0.15f, 0.18f};

float len3(float x, float y, float z) { return std::sqrt(x * x + y * y + z * z); }

// Google's Turbo colormap (A. Mikhailov, Google Research, 2019; published
// 256-entry sRGB table, Apache-2.0), the table the reference renderer
// interpolates (render_engine.cu:38-108, :519-535).  Stored in units of 1e-5:
// every published value has five decimals, and the correctly rounded
// quotient v / 100000 is the same float as the five-decimal literal.
const int32_t kTurbo1e5[256 * 3] = {
    18995, 7176, 23217, 19483, 8339, 26149, 19956, 9498, 29024, 20415, 10652, 31844, 20860, 11802, 34607, 21291, 12947, 37314, 21708, 14087, 39964, 22111, 15223, 42558,
    22500, 16354, 45096, 22875, 17481, 47578, 23236, 18603, 50004, 23582, 19720, 52373, 23915, 20833, 54686, 24234, 21941, 56942, 24539, 23044, 59142, 24830, 24143, 61286,
    25107, 25237, 63374, 25369, 26327, 65406, 25618, 27412, 67381, 25853, 28492, 69300, 26074, 29568, 71162, 26280, 30639, 72968, 26473, 31706, 74718, 26652, 32768, 76412,
    26816, 33825, 78050, 26967, 34878, 79631, 27103, 35926, 81156, 27226, 36970, 82624, 27334, 38008, 84037, 27429, 39043, 85393, 27509, 40072, 86692, 27576, 41097, 87936,
    27628, 42118, 89123, 27667, 43134, 90254, 27691, 44145, 91328, 27701, 45152, 92347, 27698, 46153, 93309, 27680, 47151, 94214, 27648, 48144, 95064, 27603, 49132, 95857,
    27543, 50115, 96594, 27469, 51094, 97275, 27381, 52069, 97899, 27273, 53040, 98461, 27106, 54015, 98930, 26878, 54995, 99303, 26592, 55979, 99583, 26252, 56967, 99773,
    25862, 57958, 99876, 25425, 58950, 99896, 24946, 59943, 99835, 24427, 60937, 99697, 23874, 61931, 99485, 23288, 62923, 99202, 22676, 63913, 98851, 22039, 64901, 98436,
    21382, 65886, 97959, 20708, 66866, 97423, 20021, 67842, 96833, 19326, 68812, 96190, 18625, 69775, 95498, 17923, 70732, 94761, 17223, 71680, 93981, 16529, 72620, 93161,
    15844, 73551, 92305, 15173, 74472, 91416, 14519, 75381, 90496, 13886, 76279, 89550, 13278, 77165, 88580, 12698, 78037, 87590, 12151, 78896, 86581, 11639, 79740, 85559,
    11167, 80569, 84525, 10738, 81381, 83484, 10357, 82177, 82437, 10026, 82955, 81389, 9750, 83714, 80342, 9532, 84455, 79299, 9377, 85175, 78264, 9287, 85875, 77240,
    9267, 86554, 76230, 9320, 87211, 75237, 9451, 87844, 74265, 9662, 88454, 73316, 9958, 89040, 72393, 10342, 89600, 71500, 10815, 90142, 70599, 11374, 90673, 69651,
    12014, 91193, 68660, 12733, 91701, 67627, 13526, 92197, 66556, 14391, 92680, 65448, 15323, 93151, 64308, 16319, 93609, 63137, 17377, 94053, 61938, 18491, 94484, 60713,
    19659, 94901, 59466, 20877, 95304, 58199, 22142, 95692, 56914, 23449, 96065, 55614, 24797, 96423, 54303, 26180, 96765, 52981, 27597, 97092, 51653, 29042, 97403, 50321,
    30513, 97697, 48987, 32006, 97974, 47654, 33517, 98234, 46325, 35043, 98477, 45002, 36581, 98702, 43688, 38127, 98909, 42386, 39678, 99098, 41098, 41229, 99268, 39826,
    42778, 99419, 38575, 44321, 99551, 37345, 45854, 99663, 36140, 47375, 99755, 34963, 48879, 99828, 33816, 50362, 99879, 32701, 51822, 99910, 31622, 53255, 99919, 30581,
    54658, 99907, 29581, 56026, 99873, 28623, 57357, 99817, 27712, 58646, 99739, 26849, 59891, 99638, 26038, 61088, 99514, 25280, 62233, 99366, 24579, 63323, 99195, 23937,
    64362, 98999, 23356, 65394, 98775, 22835, 66428, 98524, 22370, 67462, 98246, 21960, 68494, 97941, 21602, 69525, 97610, 21294, 70553, 97255, 21032, 71577, 96875, 20815,
    72596, 96470, 20640, 73610, 96043, 20504, 74617, 95593, 20406, 75617, 95121, 20343, 76608, 94627, 20311, 77591, 94113, 20310, 78563, 93579, 20336, 79524, 93025, 20386,
    80473, 92452, 20459, 81410, 91861, 20552, 82333, 91253, 20663, 83241, 90627, 20788, 84133, 89986, 20926, 85010, 89328, 21074, 85868, 88655, 21230, 86709, 87968, 21391,
    87530, 87267, 21555, 88331, 86553, 21719, 89112, 85826, 21880, 89870, 85087, 22038, 90605, 84337, 22188, 91317, 83576, 22328, 92004, 82806, 22456, 92666, 82025, 22570,
    93301, 81236, 22667, 93909, 80439, 22744, 94489, 79634, 22800, 95039, 78823, 22831, 95560, 78005, 22836, 96049, 77181, 22811, 96507, 76352, 22754, 96931, 75519, 22663,
    97323, 74682, 22536, 97679, 73842, 22369, 98000, 73000, 22161, 98289, 72140, 21918, 98549, 71250, 21650, 98781, 70330, 21358, 98986, 69382, 21043, 99163, 68408, 20706,
    99314, 67408, 20348, 99438, 66386, 19971, 99535, 65341, 19577, 99607, 64277, 19165, 99654, 63193, 18738, 99675, 62093, 18297, 99672, 60977, 17842, 99644, 59846, 17376,
    99593, 58703, 16899, 99517, 57549, 16412, 99419, 56386, 15918, 99297, 55214, 15417, 99153, 54036, 14910, 98987, 52854, 14398, 98799, 51667, 13883, 98590, 50479, 13367,
    98360, 49291, 12849, 98108, 48104, 12332, 97837, 46920, 11817, 97545, 45740, 11305, 97234, 44565, 10797, 96904, 43399, 10294, 96555, 42241, 9798, 96187, 41093, 9310,
    95801, 39958, 8831, 95398, 38836, 8362, 94977, 37729, 7905, 94538, 36638, 7461, 94084, 35566, 7031, 93612, 34513, 6616, 93125, 33482, 6218, 92623, 32473, 5837,
    92105, 31489, 5475, 91572, 30530, 5134, 91024, 29599, 4814, 90463, 28696, 4516, 89888, 27824, 4243, 89298, 26981, 3993, 88691, 26152, 3753, 88066, 25334, 3521,
    87422, 24526, 3297, 86760, 23730, 3082, 86079, 22945, 2875, 85380, 22170, 2677, 84662, 21407, 2487, 83926, 20654, 2305, 83172, 19912, 2131, 82399, 19182, 1966,
    81608, 18462, 1809, 80799, 17753, 1660, 79971, 17055, 1520, 79125, 16368, 1387, 78260, 15693, 1264, 77377, 15028, 1148, 76476, 14374, 1041, 75556, 13731, 942,
    74617, 13098, 851, 73661, 12477, 769, 72686, 11867, 695, 71692, 11268, 629, 70680, 10680, 571, 69650, 10102, 522, 68602, 9536, 481, 67535, 8980, 449,
    66449, 8436, 424, 65345, 7902, 408, 64223, 7380, 401, 63082, 6868, 401, 61923, 6367, 410, 60746, 5878, 427, 59550, 5399, 453, 58336, 4931, 486,
    57103, 4474, 529, 55852, 4028, 579, 54583, 3593, 638, 53295, 3169, 705, 51989, 2756, 780, 50664, 2354, 863, 49321, 1963, 955, 47960, 1583, 1055
};

struct TurboTable {
    float v[256][3];
    TurboTable() {
        for (int i = 0; i < 256; ++i)
            for (int c = 0; c < 3; ++c) v[i][c] = (float)kTurbo1e5[3 * i + c] / 100000.0f;
    }
};
const TurboTable& turbo_table() {
    static const TurboTable t;
    return t;
}

// getTurboColor (render_engine.cu:519-535): clamp, scale to the table,
// linear interpolation between neighbouring entries.
void turbo(float t, float& r, float& g, float& b) {
    const TurboTable& T = turbo_table();
    t = std::max(0.0f, std::min(1.0f, t));
    const float scaled = t * 255.0f;
    const int i0 = (int)scaled;
    const int i1 = std::min(i0 + 1, 255);
    const float frac = scaled - (float)i0;
    r = T.v[i0][0] * (1.0f - frac) + T.v[i1][0] * frac;
    g = T.v[i0][1] * (1.0f - frac) + T.v[i1][1] * frac;
    b = T.v[i0][2] * (1.0f - frac) + T.v[i1][2] * frac;
}

// d_normalize (render_engine.cu:130-135): multiply by the reciprocal length;
// a vector shorter than 1e-12 becomes zero.
void normalize3(float* v) {
    const float len = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (len < 1e-12f) {
        v[0] = v[1] = v[2] = 0.0f;
        return;
    }
    const float inv = 1.0f / len;
    v[0] = v[0] * inv;
    v[1] = v[1] * inv;
    v[2] = v[2] * inv;
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
    normalize3(light);

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
            std::cout << "    Heatmap normalization: p99.9=" << hmax << " (actual max=" << visits.back() << ")\n";
        }
    }
    std::vector<float> rgb(np * 3);
    for (size_t i = 0; i < np; ++i) {
        float r = kBg[0], g = kBg[1], b = kBg[2];
        if (mode == ShadingMode::HEATMAP) {
            if (cnt[3 * i] > 0) {
                float tt = (float)cnt[3 * i] / hmax;
                tt = std::max(0.0f, std::min(1.0f, tt));
                turbo(tt, r, g, b);
            }
        } else if (prim[i] >= 0) {
            const size_t p = (size_t)prim[i];
            float e1[3] = {mesh.v1x[p] - mesh.v0x[p], mesh.v1y[p] - mesh.v0y[p], mesh.v1z[p] - mesh.v0z[p]};
            float e2[3] = {mesh.v2x[p] - mesh.v0x[p], mesh.v2y[p] - mesh.v0y[p], mesh.v2z[p] - mesh.v0z[p]};
            float n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
            normalize3(n);
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

extern "C" int bvh_render_ppm(const void* nodes, uint32_t num_nodes, const float* const* soa, uint32_t ntri,
                              int32_t width, int32_t height, const float* camera, int32_t mode,
                              float fixed_heatmap_max, const char* path, float* stats) {
    if (!nodes || num_nodes == 0 || !soa || ntri == 0 || !camera || width <= 0 || height <= 0 || mode < 0 ||
        mode > 3)
        return BVH_E_ARG;
    try {
        std::vector<BVHNode> nv((const BVHNode*)nodes, (const BVHNode*)nodes + num_nodes);
        TriangleMesh m;
        m.v0x.assign(soa[0], soa[0] + ntri); m.v0y.assign(soa[1], soa[1] + ntri); m.v0z.assign(soa[2], soa[2] + ntri);
        m.v1x.assign(soa[3], soa[3] + ntri); m.v1y.assign(soa[4], soa[4] + ntri); m.v1z.assign(soa[5], soa[5] + ntri);
        m.v2x.assign(soa[6], soa[6] + ntri); m.v2y.assign(soa[7], soa[7] + ntri); m.v2z.assign(soa[8], soa[8] + ntri);
        Camera cam(float3_cw(camera[0], camera[1], camera[2]), float3_cw(camera[3], camera[4], camera[5]),
                   float3_cw(camera[6], camera[7], camera[8]), camera[9]);
        RenderStats rs = renderImage(nv, m, width, height, cam, (ShadingMode)mode, path ? path : "",
                                     fixed_heatmap_max);
        if (stats) {
            stats[0] = rs.renderTimeMs;
            stats[1] = rs.avgNodesVisited;
            stats[2] = rs.avgAABBTests;
            stats[3] = rs.avgTriTests;
            stats[4] = rs.maxNodesVisited;
        }
        return BVH_OK;
    } catch (const std::bad_alloc&) {
        return BVH_E_NOMEM;
    } catch (const std::exception&) {
        return BVH_E_STATE;
    }
}
