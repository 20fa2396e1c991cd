// Packing of a forest's split nodes into 128-byte blocks for prediction (host only, no HIP).
//
// A block holds a subtree of up to three split levels at heap positions 0..6 (position p's children at
// 2p+1 and 2p+2) as 32 words: slots 0..6 are node records of 4 words {threshold bits, offsets
// o1x o1y o2x o2y, regions r1x r1y r2x r2y, type c1 c2 -} (signed bytes), slot 7 is the tail {first
// child block, first leaf, leaf mask, child-block mask | first exit position << 16}.  A child position q
// below the first exit position (7, or 1 / 3 for a tree's top block of one or two levels) holds a split
// of the block; at or past it, a split child starts another block.  Bit q of the leaf mask marks a leaf
// child, numbered from the first leaf in position order; bit q of the child-block mask a child block,
// numbered from the first child block in position order.  Each tree's blocks are breadth-first from its
// root block, and a block's child blocks and leaves are consecutive.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/curfil_hip.h"

namespace cf {

constexpr int kBlockWords = 32;   // 128 bytes

// columns of cf_tree_node::feature
enum FeatColumn { FC_TYPE = 0, FC_O1X, FC_O1Y, FC_O2X, FC_O2Y, FC_R1X, FC_R1Y, FC_R2X, FC_R2Y, FC_C1, FC_C2 };

struct PackedForest {
    std::vector<uint32_t> blocks;   // [blocks][kBlockWords]
    std::vector<int32_t> roots;     // root block of each tree, or -(leaf index)-1 for a one-node tree
    std::vector<double> post;       // [leaves][rowStride(C)], a block's leaves together
};

// doubles per posterior row: C rounded up to even (16-byte rows for paired loads), the padding zero
constexpr int rowStride(int C) { return (C + 1) & ~1; }

// trees[t] holds sizes[t] nodes in creation order (node 0 the root, children after their parents) and
// post[t] their normalised histograms [nodes][C].  CF_OK, or an error code with `error` set.
int packForest(const std::vector<const cf_tree_node*>& trees, const std::vector<int>& sizes,
        const std::vector<const double*>& post, int C, PackedForest& out, std::string& error);

}  // namespace cf
