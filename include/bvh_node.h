#pragma once
// Unified BVH node, byte-compatible with the reference include/bvh_node.h:8-31:
// a 32-byte box, two int32 child references and 8 bytes of padding (48 B,
// 16-byte aligned).  A leaf has bit 31 of leftChild set; its low 31 bits are
// the original triangle index and rightChild is -1.  Internal nodes store
// plain node indices in both fields.

#include "common.h"
#include <cstdint>

struct BVHNode {
    AABB_cw bbox;
    int32_t leftChild;
    int32_t rightChild;

    BVH_HD bool isLeaf() const { return (leftChild < 0) || (leftChild & 0x80000000); }
    BVH_HD uint32_t getPrimitiveIndex() const { return leftChild & 0x7FFFFFFF; }
};

static_assert(sizeof(BVHNode) == 48, "BVHNode must stay 48 bytes");
