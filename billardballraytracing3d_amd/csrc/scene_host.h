// scene_host.h -- host-side scene representation shared by the scene builder
// (scene_build.cpp) and the device runtime (bbrt_hip.hip).
#pragma once

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/bbrt.h"

namespace bbrt {

constexpr int kCoarseN = 100;                         // SurfacePatch.hpp:117
constexpr int kCells = kCoarseN - 1;                  // 99 cells per side
constexpr int kTileN = 33, kBlockN = 11, kSuperN = 3; // culling pyramid: 3x3 cells/tile, 3x3 tiles/block,
constexpr int kBlocksPerSuper = 4;                    // 4x4 blocks per super (last row/col: 3)
// The pyramid is stored children-contiguous so that one level of one parent is
// a single aligned chunk: a group of 9 boxes takes 56 floats (54 + 2 pad), a
// group of 16 boxes 96 floats.  Per patch: supers [56], blocks [9][96] (slot
// 4*bi'+bj' within the super, empty boxes where the grid ends), tiles
// [121][56] (by block bi*11+bj, slot 3*ti'+tj'), cells [1089][56] (by tile
// ti*33+tj, slot 3*i'+j').  A box is lo xyz, hi xyz; an empty box is
// (+inf, -inf) and never passes the slab test.
constexpr int kG9 = 56, kG16 = 96;
constexpr size_t kSuperFloats = kG9;
constexpr size_t kBlockFloats = 9 * kG16;
constexpr size_t kTileFloats = (size_t)kBlockN * kBlockN * kG9;
constexpr size_t kCellFloats = (size_t)kTileN * kTileN * kG9;
constexpr int kMaxNodes1D = 16;                       // largest imax / jmax the kernels accept
// Compact scenes keep each cell's box only as 6 bytes relative to its tile's box
// (quantized outward with one step of slack, decoded in FP64 and rounded
// outward), a superset of the exact FP32 cell box: by tile, 9 cells x 6 bytes + 2 pad.
constexpr int kQCellBytes = 56;
constexpr size_t kQCellPatchBytes = (size_t)kTileN * kTileN * kQCellBytes;

enum : int8_t { kEscaped = BBRT_ESCAPED, kTrapped = BBRT_TRAPPED, kError = BBRT_ERROR };
enum : int {
    kErrNone = BBRT_ERR_NONE, kErrSingular = BBRT_ERR_SINGULAR, kErrNoConv = BBRT_ERR_NOCONV,
    kErrDomain = BBRT_ERR_DOMAIN, kErrBackface = BBRT_ERR_BACKFACE, kErrZeroDir = BBRT_ERR_ZERO_DIR,
    kErrStack = BBRT_ERR_STACK
};
enum : int {
    kCtrRays = BBRT_CTR_RAYS, kCtrBounces = BBRT_CTR_BOUNCES, kCtrBvh = BBRT_CTR_BVH_NODE_TESTS,
    kCtrCand = BBRT_CTR_CANDIDATES, kCtrCull = BBRT_CTR_CULL_BOX_TESTS, kCtrMt = BBRT_CTR_MT_TESTS,
    kCtrNewtonCalls = BBRT_CTR_NEWTON_CALLS, kCtrNewtonIters = BBRT_CTR_NEWTON_ITERS,
    kCtrNewtonFlops = BBRT_CTR_NEWTON_FLOPS, kCtrPolishFail = BBRT_CTR_POLISH_FAIL, kCtrLaunches = BBRT_CTR_LAUNCHES
};

constexpr int kRowStride = kMaxNodes1D + 1;           // node-set table row: first-form row values, then scale

// Per-patch record, 48 bytes.
struct PatchMeta {
    int32_t imax, jmax;
    int64_t node_off;    // into nodes[]: u nodes, u weights, v nodes, v weights
    int64_t field_off;   // into fields[]: 12 * f_off[k]
    double max_sag;
    int32_t cull;        // 1 when every coarse value is finite (culling pyramid usable)
    int32_t nodeset;     // compact scenes: (u node set << 16) | v node set, into the node-set tables
    double pad1;
};

// Per sub-scene of a scene handle (a batch holds many), 32 bytes.
struct SceneInfo {
    int32_t root;          // BVH root node (global node index)
    int32_t patch_begin;   // first global patch index
    double diag;           // scene diagonal (RayTracingDriver.hpp:386-396)
    double cull_limit;     // culling is used while |origin|_inf <= cull_limit
    double cull_widen;     // absolute widening of the FP32 culling boxes (compact scenes rebuild cell boxes)
};

// BVH node, 64 bytes, the reference's depth-first layout (BVH.hpp:56-62).
struct BvhNode {
    double lo[3], hi[3];
    int32_t right, start, count, axis;
};

// 4-wide traversal node over the binary BVH, stored at the index of binary node
// B: its slots are B's grandchildren (slots 0-1 under B's left child, 2-3 under
// its right child; a leaf child takes one slot of its pair), so one load and
// four box tests replace two levels of the binary walk.  The boxes are FP32,
// rounded outward and widened, so a slot whose binary node passes the FP64 ray
// test also passes the FP32 one; leaves are tested again in FP64 before use.
// 128 bytes: [axis][slot] lo and hi, then the slots' binary node indices.
struct Bvh4Node {
    float lo[3][4], hi[3][4];
    int32_t ref[4];   // >= 0: internal binary node (a 4-wide node), <= -2: leaf, -2 - ref = start << 3 | count
                      // (its objects leaf[start .. start + count)), -1: empty
    int32_t axes;     // split axes that fix the depth-first order: B | left child << 2 | right child << 4
    int32_t pad[3];
};
static_assert(sizeof(Bvh4Node) == 128, "Bvh4Node is two cache lines");

struct HostScene {
    int64_t n = 0;
    std::vector<PatchMeta> meta;
    std::vector<double> nodes;      // per patch: un[imax] uw[imax] vn[jmax] vw[jmax]
    std::vector<double> fields;     // 12 * total nodes
    std::vector<double> cxyz;       // [n][100*100][3]
    std::vector<double> cell;       // [n][99*99][4] = sagitta, normal xyz
    std::vector<float> box_cell;    // [n][kCellFloats]
    std::vector<float> box_tile;    // [n][kTileFloats]
    std::vector<float> box_block;   // [n][kBlockFloats]
    std::vector<float> box_super;   // [n][kSuperFloats]
    std::vector<float> box_patch;   // [n][8]: union of the super boxes (+-inf when the patch is not culled)
    std::vector<double> pbox;       // [n][6] inflated patch boxes
    std::vector<double> pcent;      // [n][3]
    std::vector<BvhNode> bvh;
    std::vector<int32_t> leaf;
    double diag = 0.0, max_patch_sag = 0.0;
    double scale = 1.0;             // max(1, |coordinate|) over all patch boxes
    double cull_widen = 0.0;        // absolute widening of the FP32 culling boxes
    // Compact scenes keep neither the coarse grid, the cell padding nor the cell
    // boxes (cxyz, cell, box_cell stay empty): the kernels evaluate a cell from
    // the nodal fields and these per-node-set tables.
    bool compact = false;
    std::vector<double> ff_rows;    // [set][kCoarseN][kRowStride]: first-form rows at the coarse values, scale
    std::vector<double> mid_basis;  // [set][kCells][kMaxNodes1D]: second-form basis at the cell midpoints
    std::vector<uint8_t> qbox_cell; // [n][kQCellPatchBytes]: quantized cell boxes, by tile
    int n_nodesets = 0;
};

#if defined(__HIPCC__)
#define BBRT_HOST_DEVICE __host__ __device__
#else
#define BBRT_HOST_DEVICE
#endif

// Quantization of a cell box b within its tile box t (both lo xyz, hi xyz) to 6
// bytes; the host and the device build use this same code.
BBRT_HOST_DEVICE inline void quantize_cell_box(const float* t, const float* b, uint8_t* q) {
    for (int a = 0; a < 3; ++a) {
        const double lo = t[a], ext = (double)t[3 + a] - (double)t[a];
        int qlo = 0, qhi = 255;
        if (ext > 0.0 && ext < 1e300) {
            const double fl = ::floor(((double)b[a] - lo) / ext * 255.0) - 1.0;
            const double fh = ::ceil(((double)b[3 + a] - lo) / ext * 255.0) + 1.0;
            qlo = fl < 0.0 ? 0 : (fl > 255.0 ? 255 : (int)fl);
            qhi = fh < 0.0 ? 0 : (fh > 255.0 ? 255 : (int)fh);
        }
        q[a] = (uint8_t)qlo;
        q[3 + a] = (uint8_t)qhi;
    }
}

// Builds everything above from a patch set; throws std::runtime_error on bad input.
// deferred (compact scenes only): stop after the patch records, node-set
// tables, centroids and uninflated node boxes; the device computes each
// patch's max sagitta, finiteness and longest cell diagonal, then
// finish_host_scene completes the host side (with the BVH unless build_tree is
// false: the device then builds it) and the device builds the pyramid.
void build_host_scene(const bbrt_patch_set& ps, HostScene& out, int threads, bool compact = false,
                      bool deferred = false);
void finish_host_scene(HostScene& S, const std::vector<double>& max_sag, const std::vector<int32_t>& finite,
                       const std::vector<double>& hmax, bool build_tree = true);

// The 4-wide nodes of every binary node at even depth below each root (roots[i],
// its FP32 widening widen[i]); out has one entry per binary node.
void build_bvh4(const std::vector<BvhNode>& bvh, const std::vector<int32_t>& roots, const std::vector<double>& widen,
                std::vector<Bvh4Node>& out);

// BVH::build (BVH.hpp:136-189) over objects with boxes [n][6] and centroids [n][3].
void build_bvh(const std::vector<double>& box, const std::vector<double>& cent, int leaf_threshold,
               std::vector<BvhNode>& nodes, std::vector<int32_t>& leaf);

}  // namespace bbrt
