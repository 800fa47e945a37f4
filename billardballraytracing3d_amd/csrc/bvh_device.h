// bvh_device.h -- the median-split BVH (BVH::build, include/BVH.hpp:136-189)
// built on the device, for scene handles whose build runs on the GPU and for
// bbrt_build_bvh_gpu.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "scene_host.h"

namespace bbrt {

// A device-built forest: one tree per part of the object array.
struct DeviceTree {
    BvhNode* bvh = nullptr;       // [n_nodes], the parts' trees one after another, each in the reference's
                                  // depth-first layout with global node, leaf and object indices
    Bvh4Node* bvh4 = nullptr;     // [n_nodes] 4-wide traversal nodes (only when widths are given)
    int32_t* leaf = nullptr;      // [n_objects] object indices in leaf order
    double* leaf_box = nullptr;   // [n_objects][6] FP64 box of the leaf whose objects start there
    int64_t n_nodes = 0;
    std::vector<int32_t> roots;   // per part: its root node
};

// Builds, for each part i (objects begin[i] .. begin[i+1]-1, every part
// non-empty), the tree BVH(objects, leaf_threshold) builds from those objects
// in index order: the same nodes, boxes, split axes, child links and leaf
// order, bit for bit.  d_box [n][6] (min xyz, max xyz) and d_cent [n][3] are
// device arrays.  widen (host, per part, may be NULL) asks for the 4-wide
// traversal nodes with the FP32 boxes widened by widen[i] (build_bvh4).
// Device arrays are allocated with hipMalloc and appended to `allocs` (their
// bytes to `bytes`); the work runs on `st`, which is synchronised on return.
void build_bvh_device(const double* d_box, const double* d_cent, const std::vector<int64_t>& begin,
                      int leaf_threshold, const double* widen, hipStream_t st, std::vector<void*>& allocs,
                      int64_t& bytes, DeviceTree& out);

}  // namespace bbrt
