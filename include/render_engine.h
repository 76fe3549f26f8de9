#pragma once
// Renderer API, declaration-compatible with the reference include/render_engine.h
// (ShadingMode, Camera, RenderStats, autoCamera, parseCameraString,
// parseShadingMode, renderImage, printRenderStats), implemented over the
// engine's bvh_trace (bvh_amd/host/render.cpp).

#include "common.h"
#include "bvh_node.h"
#include "mesh.h"

#include <cstdint>
#include <string>
#include <vector>

enum class ShadingMode {
    NORMAL,   // surface normal mapped to RGB
    DEPTH,    // grayscale depth
    DIFFUSE,  // directional light + ambient
    HEATMAP   // nodes visited per ray
};

struct Camera {
    float3_cw eye;
    float3_cw lookAt;
    float3_cw up;
    float fovY;  // vertical field of view, degrees

    Camera() : eye(0, 0, 5), lookAt(0, 0, 0), up(0, 1, 0), fovY(60.0f) {}
    Camera(float3_cw e, float3_cw l, float3_cw u, float fov) : eye(e), lookAt(l), up(u), fovY(fov) {}
};

struct RenderStats {
    float renderTimeMs;
    int width;
    int height;
    uint64_t totalRays;
    float avgNodesVisited;
    float avgAABBTests;
    float avgTriTests;
    float maxNodesVisited;
};

// Camera framing the root box from a 3/4 view.
Camera autoCamera(const std::vector<BVHNode>& nodes);

// "ex,ey,ez,lx,ly,lz" (+ optional "ux,uy,uz" and fov); falls back to autoCamera.
Camera parseCameraString(const std::string& cameraStr, const std::string& upStr, float fov,
                         const std::vector<BVHNode>& nodes);

ShadingMode parseShadingMode(const std::string& str);

// Traces one primary ray per pixel on the GPU, shades on the host, writes a
// binary PPM when outputFile is not empty, and returns traversal statistics.
RenderStats renderImage(const std::vector<BVHNode>& nodes, const TriangleMesh& mesh, int width, int height,
                        const Camera& camera, ShadingMode mode, const std::string& outputFile,
                        float fixedHeatmapMax = 0.0f);

void printRenderStats(const std::string& algorithmName, const RenderStats& stats);
