// drivers_check.cpp -- runs the C++ drop-in drivers (include/bbrt/drivers.hpp)
// on a patch directory and writes their outputs as raw little-endian arrays,
// for tests/test_gpu_drivers.py to compare against the CPU oracle.
//
//   drivers_check trace <dir> <rays.bin> <n> <max_bounces> <newton_mode> <out>
//   drivers_check plane <dir> c0x c0y c0z c1x c1y c1z dx dy dz nx ny max_bounces count_direct <out>
//   drivers_check bmap  <dir> <prefix> <first> <last> <n> <theta_n> <phi_n> <cone_deg> <max_bounces> <capture> <out>
//   drivers_check cache <dir> <cache_dir>   (host only: the binary patch cache round trip and its staleness check)
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
