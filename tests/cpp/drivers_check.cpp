// drivers_check.cpp -- runs the C++ drop-in drivers (include/bbrt/drivers.hpp)
// on a patch directory and writes their outputs as raw little-endian arrays,
// for tests/test_gpu_drivers.py to compare against the CPU oracle.
//
//   drivers_check trace <dir> <rays.bin> <n> <max_bounces> <newton_mode> <out>
//   drivers_check plane <dir> c0x c0y c0z c1x c1y c1z dx dy dz nx ny max_bounces count_direct <out>
//   drivers_check bmap  <dir> <prefix> <first> <last> <n> <theta_n> <phi_n> <cone_deg> <max_bounces> <capture> <out>
//   drivers_check cache <dir> <cache_dir>   (host only: the binary patch cache round trip and its staleness check)
//   drivers_check kat                       (the known-answer scene of the reference's Triangle test: four
//                                            triangles, four rays, BVH with leaf threshold 1 against brute force)
//   drivers_check objects <dir> <rays.bin> <n> <newton_mode> <out>
//                                           (the patch files as standalone SurfacePatch objects: BVH::intersect,
//                                            brute_force_intersect and SurfacePatch::intersect of patch 0 per ray,
//                                            and patch 0's coarse grid)
//
// Every run writes <out>.zone: int64 zoneID of each patch in the driver's
// patch order (lexicographic file order).
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "bbrt/drivers.hpp"

using namespace bbrt_cpp;

template <class T>
static void dump(const std::string& path, const std::vector<T>& v) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
}

static void dump_zones(const RayTracingDriver& d, const std::string& out) {
    std::vector<int64_t> z;
    for (const auto& p : d.patches()) z.push_back(p->data().zoneID);
    dump(out + ".zone", z);
}

static void dump_paths(const RayTracingDriver& d, const std::vector<RayTracingDriver::RayPath>& paths, int mb,
                       const std::string& out) {
    const size_t n = paths.size();
    const int K = mb + 2;
    std::vector<int8_t> status(n);
    std::vector<int32_t> bounces(n), hits(n * (size_t)std::max(1, mb), -1);
    std::vector<double> pts(n * (size_t)K * 3, std::nan(""));
    std::string errs;
    for (size_t i = 0; i < n; ++i) {
        const auto& p = paths[i];
        status[i] = (int8_t)p.status;
        bounces[i] = p.bounce_count;
        for (size_t k = 0; k < p.hit_objects.size(); ++k) hits[i * std::max(1, mb) + k] = (int32_t)d.patch_index_of(p.hit_objects[k]);
        for (size_t k = 0; k < p.points.size() && (int)k < K; ++k) {
            pts[(i * K + k) * 3] = p.points[k].x; pts[(i * K + k) * 3 + 1] = p.points[k].y; pts[(i * K + k) * 3 + 2] = p.points[k].z;
        }
        errs += p.error_message + "\n";
    }
    dump(out + ".status", status);
    dump(out + ".bounces", bounces);
    dump(out + ".hits", hits);
    dump(out + ".pts", pts);
    std::ofstream(out + ".err") << errs;
}

int main(int argc, char** argv) {
    try {
        if (argc < 2) throw std::invalid_argument("usage: drivers_check trace|plane|bmap ...");
        const std::string cmd = argv[1];
        if (cmd == "cache" && argc == 4) {
            namespace fs = std::filesystem;
            set_patch_cache_directory(argv[3]);
            std::vector<std::string> files;
            for (const auto& e : fs::directory_iterator(argv[2]))
                if (e.is_regular_file() && e.path().extension() == ".txt") files.push_back(e.path().string());
            std::sort(files.begin(), files.end());
            bool c1 = true, c2 = false, c3 = true;
            const auto a = detail::load_patch_list(files, &c1);   // parsed, cache written
            const auto b = detail::load_patch_list(files, &c2);   // from the cache
            auto same = [](const SurfacePatch::Data& x, const SurfacePatch::Data& y) {
                auto eq = [](const auto& u, const auto& v) {
                    return u.size() == v.size() &&
                           std::memcmp(u.data(), v.data(), u.size() * sizeof(u[0])) == 0;
                };
                return x.imax == y.imax && x.jmax == y.jmax && x.kmax == y.kmax && x.zoneID == y.zoneID &&
                       eq(x.uNodes, y.uNodes) && eq(x.vNodes, y.vNodes) && eq(x.uWeights, y.uWeights) &&
                       eq(x.vWeights, y.vWeights) && eq(x.x, y.x) && eq(x.y, y.y) && eq(x.z, y.z) &&
                       eq(x.mask, y.mask) && eq(x.dxdu, y.dxdu) && eq(x.dydu, y.dydu) && eq(x.dzdu, y.dzdu) &&
                       eq(x.dxdv, y.dxdv) && eq(x.dydv, y.dydv) && eq(x.dzdv, y.dzdv) && eq(x.dS, y.dS) &&
                       eq(x.nuX, y.nuX) && eq(x.nuY, y.nuY) && eq(x.nuZ, y.nuZ) && x.edge_flags == y.edge_flags;
            };
            bool equal = a.size() == b.size() && !a.empty();
            for (size_t k = 0; equal && k < a.size(); ++k) {
                equal = same(a[k]->data(), b[k]->data()) && a[k]->centroid().x == b[k]->centroid().x &&
                        a[k]->centroid().y == b[k]->centroid().y && a[k]->centroid().z == b[k]->centroid().z;
            }
            { std::ofstream(files[0], std::ios::app) << "\n"; }   // the first file changes: the cache is stale
            const auto c = detail::load_patch_list(files, &c3);
            std::cout << "cache patches=" << a.size() << " first_from_cache=" << c1 << " second_from_cache=" << c2
                      << " after_change_from_cache=" << c3 << " equal=" << equal << std::endl;
            return (!c1 && c2 && !c3 && equal && c.size() == a.size()) ? 0 : 1;
        }
        if (cmd == "kat" && argc == 2) {
            // The scene of the reference's tests/Triangle_Intersect.cpp: spread-out triangles, so the tree splits.
            std::vector<Triangle> tris;
            tris.emplace_back(Vec3{ 0, 0, 0 }, Vec3{ 1, 0, 0 }, Vec3{ 0, 1, 0 });
            tris.emplace_back(Vec3{ 10, 0, 0 }, Vec3{ 11, 0, 0 }, Vec3{ 10, 1, 0 });
            tris.emplace_back(Vec3{ 0, 0, 5 }, Vec3{ 1, 0, 5 }, Vec3{ 0, 1, 5 });
            tris.emplace_back(Vec3{ -5, 0, 0 }, Vec3{ -4, 0, 0 }, Vec3{ -5, 1, 0 });
            std::vector<const ScatObject*> objects;
            for (auto& t : tris) objects.push_back(&t);
            BVH bvh(objects, 1);
            const Ray rays[4] = { { { 0.2, 0.2, -1 }, { 0, 0, 1 } }, { { 10.2, 0.2, -1 }, { 0, 0, 1 } },
                                  { { 100, 100, -1 }, { 0, 0, 1 } }, { { 0.2, 0.2, 10 }, { 0, 0, -1 } } };
            const double want[4] = { 1.0, 1.0, -1.0, 5.0 };   // -1: no hit
            int failures = 0;
            for (int r = 0; r < 4; ++r) {
                const auto a = bvh.intersect(rays[r]);
                const auto b = brute_force_intersect(rays[r], objects);
                const bool agree = a.has_value() == b.has_value() && (!a || std::fabs(a->t - b->t) < 1e-9);
                const bool right = want[r] < 0 ? !a : (a && std::fabs(a->t - want[r]) < 1e-9);
                std::cout << "ray " << r << ": " << (agree && right ? "PASS" : "FAIL") << " bvh t=" << (a ? a->t : -1)
                          << " brute t=" << (b ? b->t : -1) << std::endl;
                if (!(agree && right)) ++failures;
                if (a && a->object != objects[r == 3 ? 2 : (r == 1 ? 1 : 0)]) ++failures;
            }
            std::cout << (failures ? "Some tests FAILED." : "All tests passed.") << std::endl;
            return failures;
        }
        if (cmd == "objects" && argc == 7) {
            namespace fs = std::filesystem;
            std::vector<std::string> files;
            for (const auto& e : fs::directory_iterator(argv[2]))
                if (e.is_regular_file() && e.path().extension() == ".txt") files.push_back(e.path().string());
            std::sort(files.begin(), files.end());
            std::vector<std::unique_ptr<SurfacePatch>> patches;
            std::vector<const ScatObject*> objects;
            const int mode = std::stoi(argv[5]);
            for (const auto& f : files) {
                patches.push_back(std::make_unique<SurfacePatch>(f));
                patches.back()->newton_mode = mode;
                objects.push_back(patches.back().get());
            }
            BVH bvh(objects);
            bvh.newton_mode = mode;
            const size_t n = std::stoul(argv[4]);
            std::vector<double> rays(6 * n);
            std::ifstream(argv[3], std::ios::binary).read(reinterpret_cast<char*>(rays.data()),
                                                         (std::streamsize)(rays.size() * sizeof(double)));
            // per ray and query (0 BVH, 1 brute force, 2 patch 0 alone): object (-1 none, -2 error), t, point, normal
            std::vector<int32_t> obj(3 * n, -1);
            std::vector<double> val(3 * n * 7, 0.0);
            std::string errs;
            auto idx = [&](const ScatObject* o) {
                for (size_t k = 0; k < objects.size(); ++k) if (objects[k] == o) return (int32_t)k;
                return (int32_t)-1;
            };
            for (size_t i = 0; i < n; ++i) {
                const Ray ray{ { rays[6 * i], rays[6 * i + 1], rays[6 * i + 2] },
                               { rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5] } };
                for (int q = 0; q < 3; ++q) {
                    try {
                        std::optional<Hit> h = q == 0 ? bvh.intersect(ray)
                                               : (q == 1 ? brute_force_intersect(ray, objects) : patches[0]->intersect(ray));
                        if (h) {
                            obj[3 * i + q] = q == 0 ? idx(h->object) : (q == 1 ? -3 : 0);
                            const double v[7] = { h->t, h->point.x, h->point.y, h->point.z, h->normal.x, h->normal.y,
                                                  h->normal.z };
                            std::copy(v, v + 7, val.begin() + 7 * (3 * i + q));
                        }
                    } catch (const std::runtime_error& e) {
                        obj[3 * i + q] = -2;
                        errs += std::to_string(i) + " " + std::to_string(q) + " " + e.what() + "\n";
                    }
                }
            }
            // brute_force_intersect returns the hit, not the object (BruteForce.hpp:7-15): mark hits by -3
            dump(std::string(argv[6]) + ".obj", obj);
            dump(std::string(argv[6]) + ".val", val);
            std::ofstream(std::string(argv[6]) + ".err") << errs;
            const auto& g = patches[0]->coarse_grid();
            std::vector<double> cg;
            for (size_t k = 0; k < g.x.size(); ++k) { cg.push_back(g.x[k]); cg.push_back(g.y[k]); cg.push_back(g.z[k]); }
            for (size_t k = 0; k < g.cell_sagitta.size(); ++k) {
                cg.push_back(g.cell_sagitta[k]);
                cg.push_back(g.cell_normal[k].x); cg.push_back(g.cell_normal[k].y); cg.push_back(g.cell_normal[k].z);
            }
            dump(std::string(argv[6]) + ".coarse", cg);
            std::vector<double> meta;
            for (const auto& p : patches) {
                const AABB& b = p->bounding_box();
                const double v[7] = { b.min.x, b.min.y, b.min.z, b.max.x, b.max.y, b.max.z, p->max_sagitta() };
                meta.insert(meta.end(), v, v + 7);
            }
            dump(std::string(argv[6]) + ".meta", meta);
            std::vector<int32_t> leaf;
            for (const auto* o : bvh.leaf_objects()) leaf.push_back(idx(o));
            dump(std::string(argv[6]) + ".leaf", leaf);
            return 0;
        }
        if (cmd == "trace" && argc == 8) {
            RayTracingDriver drv(argv[2]);
            drv.newton_mode = std::stoi(argv[6]);
            const size_t n = std::stoul(argv[4]);
            std::vector<double> raw(6 * n);
            std::ifstream f(argv[3], std::ios::binary);
            f.read(reinterpret_cast<char*>(raw.data()), (std::streamsize)(raw.size() * sizeof(double)));
            std::vector<Vec3> src(n), dir(n);
            for (size_t i = 0; i < n; ++i) {
                src[i] = { raw[6 * i], raw[6 * i + 1], raw[6 * i + 2] };
                dir[i] = { raw[6 * i + 3], raw[6 * i + 4], raw[6 * i + 5] };
            }
            drv.set_rays(src, dir);
            const int mb = std::stoi(argv[5]);
            drv.run(mb);
            dump_zones(drv, argv[7]);
            dump_paths(drv, drv.paths(), mb, argv[7]);
            // one ray again through trace_ray: must equal the batch result
            if (n > 0) {
                auto one = drv.trace_ray(src[n / 2], dir[n / 2], mb);
                const auto& ref = drv.paths()[n / 2];
                if (one.bounce_count != ref.bounce_count || one.status != ref.status) throw std::runtime_error("trace_ray != run()");
            }
            drv.write_vtk(std::string(argv[7]) + ".vtk");
        } else if (cmd == "plane" && argc == 17) {
            PlaneSourceDriver drv(argv[2]);
            auto a = [&](int i) { return std::stod(argv[i]); };
            drv.run({ a(3), a(4), a(5) }, { a(6), a(7), a(8) }, { a(9), a(10), a(11) }, std::stoi(argv[12]),
                    std::stoi(argv[13]), std::stoi(argv[14]), std::stoi(argv[15]) != 0);
            const std::string out = argv[16];
            dump_zones(drv.driver(), out);
            std::vector<int64_t> hc(drv.hit_counts().begin(), drv.hit_counts().end());
            dump(out + ".hitcount", hc);
            dump_paths(drv.driver(), drv.paths(), std::stoi(argv[14]), out);
            drv.write_vtk(out + ".vtk");
        } else if (cmd == "bmap" && argc == 13) {
            BounceMapDriver drv(argv[2], argv[3], std::stoi(argv[4]), std::stoi(argv[5]));
            drv.keep_launch_rays = true;   // the launch rays are compared with the oracle's
            drv.run(std::stoi(argv[6]), std::stoi(argv[7]), std::stoi(argv[8]), std::stod(argv[9]), std::stoi(argv[10]),
                    std::stoi(argv[11]) != 0);
            const std::string out = argv[12];
            dump_zones(drv.driver(), out);
            const auto& L = drv.launches();
            std::vector<double> lo, ld;
            for (size_t i = 0; i < L.origins.size(); ++i) {
                lo.push_back(L.origins[i].x); lo.push_back(L.origins[i].y); lo.push_back(L.origins[i].z);
                ld.push_back(L.dirs[i].x); ld.push_back(L.dirs[i].y); ld.push_back(L.dirs[i].z);
            }
            dump(out + ".launch_o", lo);
            dump(out + ".launch_d", ld);
            std::vector<int32_t> mb, ec, bb;
            for (size_t p = 0; p < drv.max_bounces().size(); ++p)
                for (size_t k = 0; k < drv.max_bounces()[p].size(); ++k) {
                    mb.push_back(drv.max_bounces()[p][k]);
                    ec.push_back(drv.error_counts()[p][k]);
                    bb.push_back(drv.best_paths().empty() ? -1 : drv.best_paths()[p][k].bounce_count);
                }
            dump(out + ".maxb", mb);
            dump(out + ".errc", ec);
            dump(out + ".bestb", bb);
            drv.write_vtk(out + ".vtk");
            if (std::stoi(argv[11]) != 0) drv.write_best_ray_vtks(out + "_best");
        } else {
            throw std::invalid_argument("bad arguments for " + cmd);
        }
    } catch (const std::exception& e) {
        std::cerr << "drivers_check: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
