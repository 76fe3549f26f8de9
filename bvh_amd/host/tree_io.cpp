// tree_io.cpp — tree files and the reference's VizNode export (include/bvh_io.h).
#include "bvh_io.h"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <new>
#include <vector>

namespace {

constexpr char kMagic[8] = {'B', 'V', 'H', 'A', 'M', 'D', '0', '1'};
constexpr uint32_t kVersion = 1;
constexpr size_t kNodeBytes = 48;

struct FileCloser {
    void operator()(FILE* f) const {
        if (f) std::fclose(f);
    }
};
using File = std::unique_ptr<FILE, FileCloser>;

// The reference's VizNode (src/main.cpp:24-29).
struct VizNode {
    float mn[3];
    float mx[3];
    int32_t left;
    int32_t right;
};
static_assert(sizeof(VizNode) == 32, "VizNode is 32 bytes");

// Bytes left in the file from the current position (-1 when unknown).
long long bytes_left(FILE* f) {
    const long here = std::ftell(f);
    if (here < 0 || std::fseek(f, 0, SEEK_END) != 0) return -1;
    const long end = std::ftell(f);
    if (std::fseek(f, here, SEEK_SET) != 0 || end < here) return -1;
    return (long long)(end - here);
}

}  // namespace

extern "C" {

int bvh_tree_validate(const void* nodes, uint32_t num_nodes, const uint32_t* indices, uint32_t num_prims) {
    if (!nodes || num_nodes == 0) return BVH_E_ARG;
    if (num_prims && (uint64_t)num_nodes != 2ull * num_prims - 1) return BVH_E_STATE;
    const unsigned char* p = static_cast<const unsigned char*>(nodes);
    uint64_t leaves = 0;
    for (uint32_t i = 0; i < num_nodes; ++i) {
        int32_t l, r;
        std::memcpy(&l, p + (size_t)i * kNodeBytes + 32, 4);
        std::memcpy(&r, p + (size_t)i * kNodeBytes + 36, 4);
        if (l < 0) {  // leaf: primitive in the low 31 bits
            ++leaves;
            if (num_prims && (uint32_t)(l & 0x7FFFFFFF) >= num_prims) return BVH_E_STATE;
        } else if ((uint32_t)l >= num_nodes || r < 0 || (uint32_t)r >= num_nodes || l == (int32_t)i ||
                   r == (int32_t)i) {
            return BVH_E_STATE;
        }
    }
    if (leaves != (uint64_t)(num_nodes / 2) + 1) return BVH_E_STATE;  // a binary tree has n leaves, n-1 internal
    if (indices)
        for (uint32_t i = 0; i < num_prims; ++i)
            if (indices[i] >= num_prims) return BVH_E_STATE;
    return BVH_OK;
}


int bvh_tree_file_write(const char* path, const void* nodes, uint32_t num_nodes, const uint32_t* indices,
                        uint32_t num_prims) {
    if (!path || !nodes || num_nodes == 0 || (num_prims && !indices)) return BVH_E_ARG;
    File f(std::fopen(path, "wb"));
    if (!f) return BVH_E_ARG;
    const uint32_t head[4] = {kVersion, num_nodes, num_prims, 0u};
    bool ok = std::fwrite(kMagic, 1, 8, f.get()) == 8 && std::fwrite(head, 4, 4, f.get()) == 4 &&
              std::fwrite(nodes, kNodeBytes, num_nodes, f.get()) == num_nodes &&
              (num_prims == 0 || std::fwrite(indices, 4, num_prims, f.get()) == num_prims);
    ok = ok && std::fflush(f.get()) == 0;
    return ok ? BVH_OK : BVH_E_STATE;
}

int bvh_tree_file_read(const char* path, void* nodes, uint32_t* num_nodes, uint32_t* indices, uint32_t* num_prims) {
    if (!path || !num_nodes || !num_prims) return BVH_E_ARG;
    File f(std::fopen(path, "rb"));
    if (!f) return BVH_E_ARG;
    char magic[8];
    uint32_t head[4];
    if (std::fread(magic, 1, 8, f.get()) != 8 || std::memcmp(magic, kMagic, 8) != 0) return BVH_E_STATE;
    if (std::fread(head, 4, 4, f.get()) != 4 || head[0] != kVersion || head[1] == 0) return BVH_E_STATE;
    // The header's counts must describe exactly the bytes that follow (a
    // truncated or foreign file is rejected before anything is allocated).
    const long long left = bytes_left(f.get());
    if (left < 0 || (unsigned long long)left != (unsigned long long)head[1] * kNodeBytes + 4ull * head[2])
        return BVH_E_STATE;
    if (!nodes) {  // sizes only
        *num_nodes = head[1];
        *num_prims = head[2];
        return BVH_OK;
    }
    if (*num_nodes < head[1] || (head[2] && (!indices || *num_prims < head[2]))) return BVH_E_ARG;
    if (std::fread(nodes, kNodeBytes, head[1], f.get()) != head[1]) return BVH_E_STATE;
    if (head[2] && std::fread(indices, 4, head[2], f.get()) != head[2]) return BVH_E_STATE;
    *num_nodes = head[1];
    *num_prims = head[2];
    return BVH_OK;
}

int bvh_tree_save(bvh_ctx_t ctx, bvh_tree_t tree, const char* path) {
    if (!ctx || !tree || !path) return BVH_E_ARG;
    uint32_t nn = 0, np = 0;
    int st = bvh_tree_info(tree, &nn, &np);
    if (st != BVH_OK) return st;
    if (nn == 0) return BVH_E_STATE;
    try {
        std::vector<unsigned char> nodes((size_t)nn * kNodeBytes);
        std::vector<uint32_t> idx(np);
        if ((st = bvh_tree_download(ctx, tree, nodes.data(), np ? idx.data() : nullptr)) != BVH_OK) return st;
        return bvh_tree_file_write(path, nodes.data(), nn, idx.data(), np);
    } catch (const std::bad_alloc&) {
        return BVH_E_NOMEM;
    }
}

int bvh_tree_load(bvh_ctx_t ctx, const char* path, bvh_tree_t tree) {
    if (!ctx || !tree || !path) return BVH_E_ARG;
    uint32_t nn = 0, np = 0;
    int st = bvh_tree_file_read(path, nullptr, &nn, nullptr, &np);  // counts checked against the file size
    if (st != BVH_OK) return st;
    try {
        std::vector<unsigned char> nodes((size_t)nn * kNodeBytes);
        std::vector<uint32_t> idx(np);
        if ((st = bvh_tree_file_read(path, nodes.data(), &nn, idx.data(), &np)) != BVH_OK) return st;
        if ((st = bvh_tree_validate(nodes.data(), nn, np ? idx.data() : nullptr, np)) != BVH_OK) return st;
        if ((st = bvh_tree_from_nodes(ctx, nodes.data(), nn, tree)) != BVH_OK) return st;
        if (np) st = bvh_tree_set_indices(ctx, tree, idx.data(), np);
        return st;
    } catch (const std::bad_alloc&) {
        return BVH_E_NOMEM;
    }
}

// exportBVHToOBJ (reference src/main.cpp:66-99): every box as eight "v"
// lines and six "l" polylines, written with the stream's default float
// format.  leaves_only keeps the leaf boxes; a leaf is a node whose left
// child has the leaf bit (the reference tests the node index against the
// LBVH layout, leaves last, which selects the wrong nodes for PLOC's leaves-
// first layout).
int bvh_nodes_export_obj(const void* nodes, uint32_t num_nodes, const char* path, int32_t leaves_only) {
    if (!nodes || num_nodes == 0 || !path) return BVH_E_ARG;
    std::ofstream file(path);
    if (!file) return BVH_E_ARG;
    const unsigned char* p = static_cast<const unsigned char*>(nodes);
    long long v_offset = 1;
    for (uint32_t i = 0; i < num_nodes; ++i) {
        const unsigned char* rec = p + (size_t)i * kNodeBytes;
        float mn[3], mx[3];
        int32_t left;
        std::memcpy(mn, rec + 0, 12);
        std::memcpy(mx, rec + 16, 12);
        std::memcpy(&left, rec + 32, 4);
        if (leaves_only && left >= 0) continue;
        const float xs[2] = {mn[0], mx[0]}, ys[2] = {mn[1], mx[1]}, zs[2] = {mn[2], mx[2]};
        // corners: bottom face (z = min) counter-clockwise, then the top face
        static const int kCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                          {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
        for (const auto& c : kCorner) file << "v " << xs[c[0]] << " " << ys[c[1]] << " " << zs[c[2]] << "\n";
        const long long o = v_offset;
        file << "l " << o << " " << o + 1 << " " << o + 2 << " " << o + 3 << " " << o << "\n";
        file << "l " << o + 4 << " " << o + 5 << " " << o + 6 << " " << o + 7 << " " << o + 4 << "\n";
        for (int e = 0; e < 4; ++e) file << "l " << o + e << " " << o + 4 + e << "\n";
        v_offset += 8;
    }
    file.flush();
    return file ? BVH_OK : BVH_E_STATE;
}

int bvh_tree_export_obj(bvh_ctx_t ctx, bvh_tree_t tree, const char* path, int32_t leaves_only) {
    if (!ctx || !tree || !path) return BVH_E_ARG;
    uint32_t nn = 0, np = 0;
    int st = bvh_tree_info(tree, &nn, &np);
    if (st != BVH_OK) return st;
    if (nn == 0) return BVH_E_STATE;
    std::vector<unsigned char> raw;
    try {
        raw.resize((size_t)nn * kNodeBytes);
    } catch (const std::bad_alloc&) {
        return BVH_E_NOMEM;
    }
    if ((st = bvh_tree_download(ctx, tree, raw.data(), nullptr)) != BVH_OK) return st;
    return bvh_nodes_export_obj(raw.data(), nn, path, leaves_only);
}

int bvh_tree_export_viz(bvh_ctx_t ctx, bvh_tree_t tree, const char* path) {
    if (!ctx || !tree || !path) return BVH_E_ARG;
    uint32_t nn = 0, np = 0;
    int st = bvh_tree_info(tree, &nn, &np);
    if (st != BVH_OK) return st;
    if (nn == 0) return BVH_E_STATE;
    std::vector<unsigned char> raw;
    std::vector<VizNode> out;
    try {
        raw.resize((size_t)nn * kNodeBytes);
        out.resize(nn);
    } catch (const std::bad_alloc&) {
        return BVH_E_NOMEM;
    }
    if ((st = bvh_tree_download(ctx, tree, raw.data(), nullptr)) != BVH_OK) return st;
    for (uint32_t i = 0; i < nn; ++i) {
        const unsigned char* p = raw.data() + (size_t)i * kNodeBytes;
        std::memcpy(out[i].mn, p + 0, 12);   // bbox.min (x, y, z, pad)
        std::memcpy(out[i].mx, p + 16, 12);  // bbox.max
        std::memcpy(&out[i].left, p + 32, 4);
        std::memcpy(&out[i].right, p + 36, 4);
    }
    File f(std::fopen(path, "wb"));
    if (!f) return BVH_E_ARG;
    if (std::fwrite(out.data(), sizeof(VizNode), nn, f.get()) != nn) return BVH_E_STATE;
    return std::fflush(f.get()) == 0 ? BVH_OK : BVH_E_STATE;
}

}  // extern "C"
