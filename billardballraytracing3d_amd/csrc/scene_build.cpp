// scene_build.cpp -- host-side scene construction for the bounce-loop library.
//
// Per patch this computes what SurfacePatch's constructor computes
// (include/ScatObjects/SurfacePatch.hpp:125-131): barycentric weights
// (:335-336, PatchInterpolation.hpp:25-52), centroid and node box (:380-407),
// the 100x100 coarse grid by first-form barycentric interpolation (:416-436,
// PatchInterpolation.hpp:54-159), per-cell sagitta and normal and the
// inflated box (:446-492).  Then the BVH exactly as BVH::build (BVH.hpp:136-189)
// and the scene diagonal (RayTracingDriver.hpp:386-396).
//
// On top of that it builds the FP32 culling pyramid the kernels use to skip
// coarse cells a ray cannot touch: per cell the box of all 12 corners of the
// three offset copies (sign 0, +s, -s, SurfacePatch.hpp:287-294), widened by
// an absolute margin and rounded outward to float; then unions over 3x3 cells
// (tiles), 3x3 tiles (blocks) and 4x4 blocks (supers).
//
// Floating point: compiled with -ffp-contract=off and no -march, so every
// expression rounds exactly as written; the results are bit-identical to the
// reference's formulas in the same evaluation order.  The two matrix products
// of the first-form interpolation sum their inner index upwards from 0.0.
#include "scene_host.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <limits>
#include <map>
#include <string>
#include <unordered_map>
#include <stdexcept>
#include <thread>

namespace bbrt {
namespace {

constexpr double kOnNodeTol = 1.0e-12;   // EQUAL_TOL, PatchInterpolation.hpp:21

// std::min / std::max as the reference calls them: (b < a) ? b : a and (a < b) ? b : a.
inline double lesser(double a, double b) { return (b < a) ? b : a; }
inline double greater(double a, double b) { return (a < b) ? b : a; }

struct P3 { double x, y, z; };
inline P3 unit(P3 v) {
    const double len = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    return { v.x / len, v.y / len, v.z / len };
}

// One axis of the first-form interpolation operator at `count` points:
// rows[i][k] = w_k / (x_i - node_k) and scale[i] = prod_k (x_i - node_k), or
// a one-hot row with scale 1 when x_i is within 1e-12 of a node.
void first_form_rows(const double* nodes, const double* w, int n, const double* x, int count,
                     std::vector<double>& rows, std::vector<double>& scale) {
    rows.assign((size_t)count * n, 0.0);
    scale.assign(count, 1.0);
    for (int i = 0; i < count; ++i) {
        double* r = &rows[(size_t)i * n];
        double prod = 1.0;
        int hot = -1;
        for (int k = 0; k < n; ++k) {
            const double diff = x[i] - nodes[k];
            if (std::fabs(diff) < kOnNodeTol) { hot = k; break; }
            r[k] = w[k] / diff;
            prod *= diff;
        }
        if (hot >= 0) {
            for (int k = 0; k < n; ++k) r[k] = 0.0;
            r[hot] = 1.0;
            scale[i] = 1.0;
        } else {
            scale[i] = prod;
        }
    }
}

// Second-form basis at one point (SurfacePatch.hpp:518-538).
void basis_at(const double* nodes, const double* w, int n, double x, double* b) {
    for (int i = 0; i < n; ++i) b[i] = 0.0;
    for (int i = 0; i < n; ++i)
        if (std::fabs(x - nodes[i]) < kOnNodeTol) { b[i] = 1.0; return; }
    double denom = 0.0;
    for (int i = 0; i < n; ++i) {
        const double li = w[i] / (x - nodes[i]);
        b[i] = li;
        denom += li;
    }
    for (int i = 0; i < n; ++i) b[i] /= denom;
}

// SurfacePatch.hpp:542-555
double field_at(const double* lu, const double* lv, const double* f, int imax, int jmax) {
    double sum = 0.0;
    for (int i = 0; i < imax; ++i) {
        const double li = lu[i];
        if (li == 0.0) continue;
        double row = 0.0;
        for (int j = 0; j < jmax; ++j) row += lv[j] * f[(size_t)i * jmax + j];
        sum += li * row;
    }
    return sum;
}

float round_down(double v) {
    float f = (float)v;
    if ((double)f > v) f = std::nextafter(f, -std::numeric_limits<float>::infinity());
    return f;
}
float round_up(double v) {
    float f = (float)v;
    if ((double)f < v) f = std::nextafter(f, std::numeric_limits<float>::infinity());
    return f;
}

struct PatchWork {
    const bbrt_patch_set* ps;
    HostScene* S;
    std::vector<double> hmax;   // per patch: longest coarse cell diagonal
};

// Per-thread coarse grid and cell padding of one patch, for compact scenes
// (which keep neither).
double* tls_grid() {
    thread_local std::vector<double> g;
    if (g.empty()) g.resize((size_t)kCoarseN * kCoarseN * 3);
    return g.data();
}
double* tls_cells() {
    thread_local std::vector<double> c;
    if (c.empty()) c.resize((size_t)kCells * kCells * 4);
    return c.data();
}

// Coarse grid of patch k into G (SurfacePatch.hpp:416-436).
void compute_grid(const HostScene& S, int64_t k, double* G) {
    const PatchMeta& M = S.meta[k];
    const int im = M.imax, jm = M.jmax;
    const double* un = &S.nodes[M.node_off];
    const double* uw = un + im;
    const double* vn = uw + im;
    const double* vw = vn + jm;
    const double* F = &S.fields[M.field_off];
    const size_t nn = (size_t)im * jm;
    double cu[kCoarseN];
    for (int i = 0; i < kCoarseN; ++i) cu[i] = -1.0 + 2.0 * static_cast<double>(i) / (kCoarseN - 1);
    std::vector<double> ru, su, rv, sv;
    first_form_rows(un, uw, im, cu, kCoarseN, ru, su);
    first_form_rows(vn, vw, jm, cu, kCoarseN, rv, sv);
    std::vector<double> mid((size_t)kCoarseN * jm);
    for (int c = 0; c < 3; ++c) {
        const double* f = F + c * nn;
        for (int i = 0; i < kCoarseN; ++i)
            for (int m = 0; m < jm; ++m) {
                double acc = 0.0;
                for (int q = 0; q < im; ++q) acc += ru[(size_t)i * im + q] * f[(size_t)q * jm + m];
                mid[(size_t)i * jm + m] = acc;
            }
        for (int i = 0; i < kCoarseN; ++i)
            for (int j = 0; j < kCoarseN; ++j) {
                double acc = 0.0;
                for (int m = 0; m < jm; ++m) acc += mid[(size_t)i * jm + m] * rv[(size_t)j * jm + m];
                G[((size_t)i * kCoarseN + j) * 3 + c] = acc * (su[i] * sv[j]);
            }
    }
}

// Per-cell sagitta and unit normal of patch k into C (SurfacePatch.hpp:446-492).
void compute_cells(const HostScene& S, int64_t k, const double* G, double* C) {
    const PatchMeta& M = S.meta[k];
    const int im = M.imax, jm = M.jmax;
    const double* un = &S.nodes[M.node_off];
    const double* uw = un + im;
    const double* vn = uw + im;
    const double* vw = vn + jm;
    const double* F = &S.fields[M.field_off];
    const size_t nn = (size_t)im * jm;
    double cu[kCoarseN];
    for (int i = 0; i < kCoarseN; ++i) cu[i] = -1.0 + 2.0 * static_cast<double>(i) / (kCoarseN - 1);
    double lu[kMaxNodes1D], lv[kMaxNodes1D];
    for (int i = 0; i < kCells; ++i)
        for (int j = 0; j < kCells; ++j) {
            const double um = 0.5 * (cu[i] + cu[i + 1]);
            const double vm = 0.5 * (cu[j] + cu[j + 1]);
            basis_at(un, uw, im, um, lu);
            basis_at(vn, vw, jm, vm, lv);
            const P3 tm = { field_at(lu, lv, F, im, jm), field_at(lu, lv, F + nn, im, jm),
                            field_at(lu, lv, F + 2 * nn, im, jm) };
            const P3 nm = unit({ field_at(lu, lv, F + 9 * nn, im, jm), field_at(lu, lv, F + 10 * nn, im, jm),
                                 field_at(lu, lv, F + 11 * nn, im, jm) });
            const double* p00 = &G[((size_t)i * kCoarseN + j) * 3];
            const double* p10 = &G[((size_t)(i + 1) * kCoarseN + j) * 3];
            const double* p01 = &G[((size_t)i * kCoarseN + j + 1) * 3];
            const double* p11 = &G[((size_t)(i + 1) * kCoarseN + j + 1) * 3];
            const P3 bm = { 0.25 * (p00[0] + p10[0] + p01[0] + p11[0]), 0.25 * (p00[1] + p10[1] + p01[1] + p11[1]),
                            0.25 * (p00[2] + p10[2] + p01[2] + p11[2]) };
            const P3 d = { tm.x - bm.x, tm.y - bm.y, tm.z - bm.z };
            const double s = std::sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
            double* c = &C[((size_t)i * kCells + j) * 4];
            c[0] = s; c[1] = nm.x; c[2] = nm.y; c[3] = nm.z;
        }
}

// Node-set tables of a compact scene: for each distinct node vector, the
// first-form rows (values then scale) at the 100 coarse values and the
// second-form basis at the 99 cell midpoints -- exactly what compute_grid and
// compute_cells evaluate per patch, shared by all patches with those nodes.
struct NodeSetTable {
    HostScene& S;
    std::map<std::vector<double>, int> index;
    int get(const double* x, const double* w, int n) {
        std::vector<double> key(x, x + n);
        auto it = index.find(key);
        if (it != index.end()) return it->second;
        const int id = (int)index.size();
        if (id >= (1 << 15)) throw std::runtime_error("compact scene: more than 32768 distinct node sets");
        index.emplace(std::move(key), id);
        double cu[kCoarseN];
        for (int i = 0; i < kCoarseN; ++i) cu[i] = -1.0 + 2.0 * static_cast<double>(i) / (kCoarseN - 1);
        std::vector<double> rows, scale;
        first_form_rows(x, w, n, cu, kCoarseN, rows, scale);
        const size_t r0 = S.ff_rows.size();
        S.ff_rows.resize(r0 + (size_t)kCoarseN * kRowStride, 0.0);
        for (int i = 0; i < kCoarseN; ++i) {
            for (int q = 0; q < n; ++q) S.ff_rows[r0 + (size_t)i * kRowStride + q] = rows[(size_t)i * n + q];
            S.ff_rows[r0 + (size_t)i * kRowStride + kMaxNodes1D] = scale[i];
        }
        const size_t m0 = S.mid_basis.size();
        S.mid_basis.resize(m0 + (size_t)kCells * kMaxNodes1D, 0.0);
        for (int i = 0; i < kCells; ++i) {
            const double um = 0.5 * (cu[i] + cu[i + 1]);
            basis_at(x, w, n, um, &S.mid_basis[m0 + (size_t)i * kMaxNodes1D]);
        }
        S.n_nodesets = (int)index.size();
        return id;
    }
};

// Centroid and node box of patch k (SurfacePatch.hpp:380-407); the box goes to
// pbox uninflated.
void centroid_box(HostScene& S, int64_t k, double lo[3], double hi[3]) {
    const PatchMeta& M = S.meta[k];
    const double* F = &S.fields[M.field_off];
    const size_t nn = (size_t)M.imax * M.jmax;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int a = 0; a < 3; ++a) { lo[a] = std::numeric_limits<double>::infinity(); hi[a] = -lo[a]; }
    for (size_t q = 0; q < nn; ++q) {
        const double px = F[q], py = F[nn + q], pz = F[2 * nn + q];
        sx += px; sy += py; sz += pz;
        lo[0] = lesser(lo[0], px); lo[1] = lesser(lo[1], py); lo[2] = lesser(lo[2], pz);
        hi[0] = greater(hi[0], px); hi[1] = greater(hi[1], py); hi[2] = greater(hi[2], pz);
    }
    const double inv_n = 1.0 / static_cast<double>(nn);
    double* cent = &S.pcent[3 * k];
    cent[0] = sx * inv_n; cent[1] = sy * inv_n; cent[2] = sz * inv_n;
    double* b = &S.pbox[6 * k];
    for (int a = 0; a < 3; ++a) { b[a] = lo[a]; b[3 + a] = hi[a]; }
}

void prepare_patch(PatchWork& W, int64_t k) {
    HostScene& S = *W.S;
    double lo[3], hi[3];
    centroid_box(S, k, lo, hi);

    // coarse grid (SurfacePatch.hpp:416-436)
    double* G = S.compact ? tls_grid() : &S.cxyz[(size_t)k * kCoarseN * kCoarseN * 3];
    compute_grid(S, k, G);

    // cell padding (SurfacePatch.hpp:446-492)
    double* C = S.compact ? tls_cells() : &S.cell[(size_t)k * kCells * kCells * 4];
    compute_cells(S, k, G, C);
    double ms = 0.0;
    bool finite = true;
    for (size_t q = 0; q < (size_t)kCells * kCells; ++q) {
        const double* c = &C[q * 4];
        ms = greater(ms, c[0]);
        if (!(std::isfinite(c[0]) && std::isfinite(c[1]) && std::isfinite(c[2]) && std::isfinite(c[3]))) finite = false;
    }
    for (size_t q = 0; q < (size_t)kCoarseN * kCoarseN * 3; ++q)
        if (!std::isfinite(G[q])) finite = false;
    double* b = &S.pbox[6 * k];
    b[0] = lo[0] - ms; b[1] = lo[1] - ms; b[2] = lo[2] - ms;
    b[3] = hi[0] + ms; b[4] = hi[1] + ms; b[5] = hi[2] + ms;
    S.meta[k].max_sag = ms;
    S.meta[k].cull = finite ? 1 : 0;

    // longest cell diagonal, for the culling margin
    double h = 0.0;
    if (finite)
        for (int i = 0; i < kCells; ++i)
            for (int j = 0; j < kCells; ++j) {
                const double* p00 = &G[((size_t)i * kCoarseN + j) * 3];
                const double* p11 = &G[((size_t)(i + 1) * kCoarseN + j + 1) * 3];
                const double* p10 = &G[((size_t)(i + 1) * kCoarseN + j) * 3];
                const double* p01 = &G[((size_t)i * kCoarseN + j + 1) * 3];
                double a = 0.0, e = 0.0;
                for (int c = 0; c < 3; ++c) { a += (p11[c] - p00[c]) * (p11[c] - p00[c]); e += (p10[c] - p01[c]) * (p10[c] - p01[c]); }
                h = std::max(h, std::sqrt(std::max(a, e)));
            }
    W.hmax[k] = h;
}

// FP32 culling pyramid of one patch, widened by `widen`, in the grouped layout
// described in scene_host.h.
void build_pyramid(HostScene& S, int64_t k, double widen) {
    const double* G;
    const double* C;
    float* gc;
    std::vector<float> gc_local;
    if (S.compact) {   // recompute the grid and the cell padding; the cell boxes are not kept
        double* g = tls_grid();
        double* cp = tls_cells();
        compute_grid(S, k, g);
        compute_cells(S, k, g, cp);
        G = g;
        C = cp;
        gc_local.assign(kCellFloats, 0.0f);
        gc = gc_local.data();
    } else {
        G = &S.cxyz[(size_t)k * kCoarseN * kCoarseN * 3];
        C = &S.cell[(size_t)k * kCells * kCells * 4];
        gc = &S.box_cell[(size_t)k * kCellFloats];
    }
    float* gt = &S.box_tile[(size_t)k * kTileFloats];
    float* gb = &S.box_block[(size_t)k * kBlockFloats];
    float* gs = &S.box_super[(size_t)k * kSuperFloats];
    const float inf = std::numeric_limits<float>::infinity();
    auto empty = [&](float* b) { b[0] = b[1] = b[2] = inf; b[3] = b[4] = b[5] = -inf; };
    auto merge = [&](float* dst, const float* src) {
        for (int a = 0; a < 3; ++a) { dst[a] = std::min(dst[a], src[a]); dst[3 + a] = std::max(dst[3 + a], src[3 + a]); }
    };
    // every box slot starts empty, padding floats zero
    for (size_t q = 0; q < kTileFloats; ++q) gt[q] = 0.0f;
    for (size_t q = 0; q < kBlockFloats; ++q) gb[q] = 0.0f;
    for (size_t q = 0; q < kSuperFloats; ++q) gs[q] = 0.0f;
    for (size_t q = 0; q < kCellFloats; ++q) gc[q] = 0.0f;
    auto empty_group = [&](float* g, int nbox) { for (int q = 0; q < nbox; ++q) empty(g + 6 * q); };
    for (int t = 0; t < kTileN * kTileN; ++t) empty_group(gc + (size_t)t * kG9, 9);
    for (int b = 0; b < kBlockN * kBlockN; ++b) empty_group(gt + (size_t)b * kG9, 9);
    for (int s = 0; s < kSuperN * kSuperN; ++s) empty_group(gb + (size_t)s * kG16, 16);
    empty_group(gs, 9);
    float* gp = &S.box_patch[(size_t)k * 8];
    for (int a = 0; a < 3; ++a) { gp[a] = -inf; gp[3 + a] = inf; }   // full box: always entered
    gp[6] = gp[7] = 0.0f;
    if (!S.meta[k].cull) return;   // never consulted: such a patch is tested exhaustively

    std::vector<float> cellb((size_t)kCells * kCells * 6), tileb((size_t)kTileN * kTileN * 6),
        blockb((size_t)kBlockN * kBlockN * 6);
    for (int i = 0; i < kCells; ++i)
        for (int j = 0; j < kCells; ++j) {
            const double* c = &C[((size_t)i * kCells + j) * 4];
            const double* corner[4] = { &G[((size_t)i * kCoarseN + j) * 3], &G[((size_t)(i + 1) * kCoarseN + j) * 3],
                                        &G[((size_t)i * kCoarseN + j + 1) * 3],
                                        &G[((size_t)(i + 1) * kCoarseN + j + 1) * 3] };
            double lo[3] = { std::numeric_limits<double>::infinity(), std::numeric_limits<double>::infinity(),
                             std::numeric_limits<double>::infinity() };
            double hi[3] = { -lo[0], -lo[1], -lo[2] };
            const double signs[3] = { 0.0, 1.0, -1.0 };
            for (double sg : signs) {
                const double d = sg * c[0];
                for (const double* p : corner) {
                    // exactly the offset corner the stage-1 test builds (SurfacePatch.hpp:268-270)
                    const double q[3] = { p[0] + c[1] * d, p[1] + c[2] * d, p[2] + c[3] * d };
                    for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], q[a]); hi[a] = std::max(hi[a], q[a]); }
                }
            }
            float* b = &cellb[((size_t)i * kCells + j) * 6];
            for (int a = 0; a < 3; ++a) { b[a] = round_down(lo[a] - widen); b[3 + a] = round_up(hi[a] + widen); }
            // grouped: tile (i/3, j/3), slot 3*(i%3) + j%3
            float* g = gc + (size_t)((i / 3) * kTileN + j / 3) * kG9 + 6 * (3 * (i % 3) + j % 3);
            for (int a = 0; a < 6; ++a) g[a] = b[a];
        }
    for (int ti = 0; ti < kTileN; ++ti)
        for (int tj = 0; tj < kTileN; ++tj) {
            float* b = &tileb[((size_t)ti * kTileN + tj) * 6];
            empty(b);
            for (int i = 3 * ti; i < 3 * ti + 3; ++i)
                for (int j = 3 * tj; j < 3 * tj + 3; ++j) merge(b, &cellb[((size_t)i * kCells + j) * 6]);
            float* g = gt + (size_t)((ti / 3) * kBlockN + tj / 3) * kG9 + 6 * (3 * (ti % 3) + tj % 3);
            for (int a = 0; a < 6; ++a) g[a] = b[a];
            if (S.compact) {
                uint8_t* q = &S.qbox_cell[(size_t)k * kQCellPatchBytes + ((size_t)ti * kTileN + tj) * kQCellBytes];
                for (int c = 0; c < 9; ++c)
                    quantize_cell_box(b, &cellb[((size_t)(3 * ti + c / 3) * kCells + 3 * tj + c % 3) * 6], q + 6 * c);
            }
        }
    for (int bi = 0; bi < kBlockN; ++bi)
        for (int bj = 0; bj < kBlockN; ++bj) {
            float* b = &blockb[((size_t)bi * kBlockN + bj) * 6];
            empty(b);
            for (int ti = 3 * bi; ti < 3 * bi + 3; ++ti)
                for (int tj = 3 * bj; tj < 3 * bj + 3; ++tj) merge(b, &tileb[((size_t)ti * kTileN + tj) * 6]);
            const int si = bi / kBlocksPerSuper, sj = bj / kBlocksPerSuper;
            float* g = gb + (size_t)(si * kSuperN + sj) * kG16 + 6 * (kBlocksPerSuper * (bi % kBlocksPerSuper) + bj % kBlocksPerSuper);
            for (int a = 0; a < 6; ++a) g[a] = b[a];
        }
    empty(gp);
    for (int si = 0; si < kSuperN; ++si)
        for (int sj = 0; sj < kSuperN; ++sj) {
            float* b = gs + 6 * (si * kSuperN + sj);
            empty(b);
            for (int bi = kBlocksPerSuper * si; bi < std::min(kBlockN, kBlocksPerSuper * (si + 1)); ++bi)
                for (int bj = kBlocksPerSuper * sj; bj < std::min(kBlockN, kBlocksPerSuper * (sj + 1)); ++bj)
                    merge(b, &blockb[((size_t)bi * kBlockN + bj) * 6]);
            merge(gp, b);
        }
}

// BVH::build (BVH.hpp:136-189) over patch indices.
struct TreeBuilder {
    const std::vector<double>& box;    // [n][6]
    const std::vector<double>& cent;   // [n][3]
    int leaf_threshold;
    std::vector<BvhNode>& nodes;
    std::vector<int32_t>& leaf;

    int32_t build(std::vector<int32_t> objs) {
        const int32_t me = (int32_t)nodes.size();
        nodes.emplace_back();
        double lo[3], hi[3];
        for (int a = 0; a < 3; ++a) { lo[a] = box[6 * objs.front() + a]; hi[a] = box[6 * objs.front() + 3 + a]; }
        for (size_t i = 1; i < objs.size(); ++i)
            for (int a = 0; a < 3; ++a) {
                lo[a] = lesser(lo[a], box[6 * objs[i] + a]);
                hi[a] = greater(hi[a], box[6 * objs[i] + 3 + a]);
            }
        if ((int)objs.size() <= leaf_threshold) {
            BvhNode& nd = nodes[me];
            for (int a = 0; a < 3; ++a) { nd.lo[a] = lo[a]; nd.hi[a] = hi[a]; }
            nd.right = -1; nd.start = (int32_t)leaf.size(); nd.count = (int32_t)objs.size(); nd.axis = 0;
            for (int32_t o : objs) leaf.push_back(o);
            return me;
        }
        double cmin[3], cmax[3];
        for (int a = 0; a < 3; ++a) cmin[a] = cmax[a] = cent[3 * objs.front() + a];
        for (int32_t o : objs)
            for (int a = 0; a < 3; ++a) {
                cmin[a] = lesser(cmin[a], cent[3 * o + a]);
                cmax[a] = greater(cmax[a], cent[3 * o + a]);
            }
        double spread[3] = { cmax[0] - cmin[0], cmax[1] - cmin[1], cmax[2] - cmin[2] };
        const int axis = (int)(std::max_element(spread, spread + 3) - spread);
        auto mid = objs.begin() + objs.size() / 2;
        const std::vector<double>& c = cent;
        std::nth_element(objs.begin(), mid, objs.end(),
                         [axis, &c](int32_t a, int32_t b) { return c[3 * a + axis] < c[3 * b + axis]; });
        std::vector<int32_t> left(objs.begin(), mid), right(mid, objs.end());
        objs.clear(); objs.shrink_to_fit();
        build(std::move(left));
        const int32_t r = build(std::move(right));
        BvhNode& nd = nodes[me];
        for (int a = 0; a < 3; ++a) { nd.lo[a] = lo[a]; nd.hi[a] = hi[a]; }
        nd.right = r; nd.start = 0; nd.count = 0; nd.axis = axis;
        return me;
    }
};

template <class F>
void parallel_for(int64_t n, int threads, F&& fn) {
    if (threads <= 1 || n < 64) { for (int64_t k = 0; k < n; ++k) fn(k); return; }
    std::atomic<int64_t> next{ 0 };
    std::vector<std::thread> pool;
    std::exception_ptr err = nullptr;
    std::atomic<bool> failed{ false };
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&]() {
            try {
                for (;;) {
                    const int64_t k = next.fetch_add(16);
                    if (k >= n || failed.load()) break;
                    for (int64_t q = k; q < std::min(n, k + 16); ++q) fn(q);
                }
            } catch (...) {
                if (!failed.exchange(true)) err = std::current_exception();
            }
        });
    for (auto& th : pool) th.join();
    if (err) std::rethrow_exception(err);
}


// Scene diagonal, largest sagitta, scale and culling margin from the inflated
// patch boxes and each patch's longest coarse-cell diagonal.
void scene_constants(HostScene& S, const std::vector<double>& patch_hmax) {
    const int64_t n = S.n;
    // scene diagonal and largest sagitta (RayTracingDriver.hpp:386-396)
    double lo[3] = { S.pbox[0], S.pbox[1], S.pbox[2] }, hi[3] = { S.pbox[3], S.pbox[4], S.pbox[5] };
    double msag = 0.0, scale = 1.0, hmax = 0.0;
    for (int64_t k = 0; k < n; ++k) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = lesser(lo[a], S.pbox[6 * k + a]);
            hi[a] = greater(hi[a], S.pbox[6 * k + 3 + a]);
        }
        msag = greater(msag, S.meta[k].max_sag);
        if (S.meta[k].cull) {
            for (int a = 0; a < 6; ++a) scale = std::max(scale, std::fabs(S.pbox[6 * k + a]));
            hmax = std::max(hmax, patch_hmax[k]);
        }
    }
    const double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    S.diag = std::sqrt(dx * dx + dy * dy + dz * dz);
    S.max_patch_sag = msag;
    S.scale = scale;
    // Margin of the FP32 culling boxes.  It must exceed (a) the FP32 slab-test
    // error for rays starting within 8*scale of the origin, about
    // 4 * 2^-24 * 20 * scale, and (b) how far outside its triangle a hit that
    // Moller-Trumbore accepts can lie, at most about 2^-52 * |o - v0| * h^2 / 1e-9
    // for cell diagonal h (the 1e-9 is the determinant guard).  1e-5 * scale *
    // max(1, h^2) exceeds both by more than an order of magnitude.
    S.cull_widen = 1e-5 * scale * std::max(1.0, hmax * hmax);
}

}  // namespace

void finish_host_scene(HostScene& S, const std::vector<double>& max_sag, const std::vector<int32_t>& finite,
                       const std::vector<double>& hmax, bool build_tree) {
    for (int64_t k = 0; k < S.n; ++k) {
        const double ms = max_sag[k];
        double* b = &S.pbox[6 * k];
        b[0] = b[0] - ms; b[1] = b[1] - ms; b[2] = b[2] - ms;
        b[3] = b[3] + ms; b[4] = b[4] + ms; b[5] = b[5] + ms;
        S.meta[k].max_sag = ms;
        S.meta[k].cull = finite[k] ? 1 : 0;
    }
    scene_constants(S, hmax);
    if (build_tree) build_bvh(S.pbox, S.pcent, 4, S.bvh, S.leaf);
}

void build_host_scene(const bbrt_patch_set& ps, HostScene& S, int threads, bool compact, bool deferred) {
    const int64_t n = ps.n_patches;
    if (n <= 0) throw std::runtime_error("bbrt_scene_create: no patches");
    if (!ps.imax || !ps.jmax || !ps.u_off || !ps.v_off || !ps.f_off || !ps.unodes || !ps.vnodes || !ps.fields)
        throw std::runtime_error("bbrt_scene_create: null array in patch set");
    S.n = n;
    S.meta.resize(n);
    int64_t node_total = 0;
    // patches with the same node vectors (bit for bit) share one table of nodes and barycentric weights, so
    // the lanes of a batched Newton read the same few cache lines
    std::unordered_map<std::string, int64_t> node_tables;
    std::string key;
    for (int64_t k = 0; k < n; ++k) {
        const int im = ps.imax[k], jm = ps.jmax[k];
        if (im < 2 || jm < 2 || im > kMaxNodes1D || jm > kMaxNodes1D)
            throw std::runtime_error("bbrt_scene_create: patch " + std::to_string(k) + " has imax/jmax outside [2, 16]");
        if (ps.u_off[k + 1] - ps.u_off[k] != im || ps.v_off[k + 1] - ps.v_off[k] != jm ||
            ps.f_off[k + 1] - ps.f_off[k] != (int64_t)im * jm)
            throw std::runtime_error("bbrt_scene_create: inconsistent offsets at patch " + std::to_string(k));
        PatchMeta& M = S.meta[k];
        M.imax = im; M.jmax = jm;
        key.assign(reinterpret_cast<const char*>(&im), sizeof(im));
        key.append(reinterpret_cast<const char*>(&jm), sizeof(jm));
        key.append(reinterpret_cast<const char*>(ps.unodes + ps.u_off[k]), sizeof(double) * (size_t)im);
        key.append(reinterpret_cast<const char*>(ps.vnodes + ps.v_off[k]), sizeof(double) * (size_t)jm);
        const auto ins = node_tables.emplace(key, node_total);
        M.node_off = ins.first->second;
        if (ins.second) node_total += 2 * (im + jm);
        M.field_off = 12 * ps.f_off[k];
        M.max_sag = 0.0; M.cull = 0; M.nodeset = 0; M.pad1 = 0.0;
    }
    S.nodes.resize(node_total);
    const int64_t field_total = 12 * ps.f_off[n];
    S.fields.assign(ps.fields, ps.fields + field_total);
    for (int64_t k = 0; k < n; ++k) {
        const PatchMeta& M = S.meta[k];
        double* dst = &S.nodes[M.node_off];
        const double* u = ps.unodes + ps.u_off[k];
        const double* v = ps.vnodes + ps.v_off[k];
        auto weights = [](const double* x, int m, double* w) {   // PatchInterpolation.hpp:25-52
            for (int j = 0; j < m; ++j) {
                double p = 1.0;
                for (int i = 0; i < m; ++i)
                    if (i != j) p *= (x[j] - x[i]);
                w[j] = 1.0 / p;
            }
        };
        for (int i = 0; i < M.imax; ++i) dst[i] = u[i];
        weights(u, M.imax, dst + M.imax);
        for (int j = 0; j < M.jmax; ++j) dst[2 * M.imax + j] = v[j];
        weights(v, M.jmax, dst + 2 * M.imax + M.jmax);
    }
    S.compact = compact;
    if (compact) {
        NodeSetTable T{ S, {} };
        for (int64_t k = 0; k < n; ++k) {
            PatchMeta& M = S.meta[k];
            const double* un = &S.nodes[M.node_off];
            const double* vn = un + 2 * M.imax;
            const int us = T.get(un, un + M.imax, M.imax), vs = T.get(vn, vn + M.jmax, M.jmax);
            M.nodeset = (us << 16) | vs;
        }
    } else {
        S.cxyz.assign((size_t)n * kCoarseN * kCoarseN * 3, 0.0);
        S.cell.assign((size_t)n * kCells * kCells * 4, 0.0);
    }
    S.pbox.assign((size_t)n * 6, 0.0);
    S.pcent.assign((size_t)n * 3, 0.0);
    if (deferred) {   // the device computes the per-patch statistics and the pyramid (finish_host_scene)
        parallel_for(n, threads, [&](int64_t k) { double lo[3], hi[3]; centroid_box(S, k, lo, hi); });
        return;
    }
    PatchWork W{ &ps, &S, std::vector<double>(n, 0.0) };
    parallel_for(n, threads, [&](int64_t k) { prepare_patch(W, k); });
    scene_constants(S, W.hmax);

    if (!compact) S.box_cell.assign((size_t)n * kCellFloats, 0.0f);
    else S.qbox_cell.assign((size_t)n * kQCellPatchBytes, 0);
    S.box_tile.assign((size_t)n * kTileFloats, 0.0f);
    S.box_block.assign((size_t)n * kBlockFloats, 0.0f);
    S.box_super.assign((size_t)n * kSuperFloats, 0.0f);
    S.box_patch.assign((size_t)n * 8, 0.0f);
    parallel_for(n, threads, [&](int64_t k) { build_pyramid(S, k, S.cull_widen); });

    // BVH (RayTracingDriver.hpp:375-384 -> BVH(objects), leaf threshold 4)
    build_bvh(S.pbox, S.pcent, 4, S.bvh, S.leaf);
}

void build_bvh(const std::vector<double>& box, const std::vector<double>& cent, int leaf_threshold,
               std::vector<BvhNode>& nodes, std::vector<int32_t>& leaf) {
    const int64_t n = (int64_t)cent.size() / 3;
    nodes.clear(); leaf.clear();
    nodes.reserve(2 * n); leaf.reserve(n);
    std::vector<int32_t> objs(n);
    for (int64_t k = 0; k < n; ++k) objs[k] = (int32_t)k;
    TreeBuilder tb{ box, cent, leaf_threshold, nodes, leaf };
    tb.build(std::move(objs));
}

void build_bvh4(const std::vector<BvhNode>& bvh, const std::vector<int32_t>& roots, const std::vector<double>& widen,
                std::vector<Bvh4Node>& out) {
    out.assign(bvh.size(), Bvh4Node{});
    const float inf = std::numeric_limits<float>::infinity();
    for (size_t i = 0; i < roots.size(); ++i) {
        const double w = widen[i];
        std::vector<int32_t> todo{ roots[i] };
        while (!todo.empty()) {
            const int32_t b = todo.back();
            todo.pop_back();
            Bvh4Node& q = out[b];
            for (int s = 0; s < 4; ++s) {
                q.ref[s] = -1;
                for (int a = 0; a < 3; ++a) { q.lo[a][s] = inf; q.hi[a][s] = -inf; }
            }
            q.axes = 0;
            auto put = [&](int s, int32_t k) {
                const BvhNode& nd = bvh[k];
                for (int a = 0; a < 3; ++a) { q.lo[a][s] = round_down(nd.lo[a] - w); q.hi[a][s] = round_up(nd.hi[a] + w); }
                if (nd.right >= 0) { q.ref[s] = k; todo.push_back(k); }
                else q.ref[s] = -2 - ((nd.start << 3) | nd.count);
            };
            const BvhNode& B = bvh[b];
            if (B.right < 0) {          // a leaf root: one slot
                put(0, b);
                continue;
            }
            const int32_t kids[2] = { b + 1, B.right };
            q.axes = B.axis;
            for (int h = 0; h < 2; ++h) {
                const BvhNode& X = bvh[kids[h]];
                if (X.right < 0) {
                    put(2 * h, kids[h]);
                } else {
                    q.axes |= X.axis << (2 + 2 * h);
                    put(2 * h, kids[h] + 1);
                    put(2 * h + 1, X.right);
                }
            }
        }
    }
}

}  // namespace bbrt
