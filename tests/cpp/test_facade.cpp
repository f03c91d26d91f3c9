// The reference's own unit tests (test/MinMaxTest.cpp, test/CompressedShadowTest.cpp),
// re-expressed against the C++ facade include/cpvs_amd.hpp.  Reads its cases from a text file
// written by tests/test_gpu_facade.py from tests/golden/reference_tests.json:
//   image <name> <N> <N*N floats>
//   minmax <name> <level> <x> <y> <min|max> <expected> <tolerance>
//   levels <name> <expected numLevels>
//   traverse <name> <x> <y> <z> <tryLeafmasks> <expected>
//   sweep <name> <x0> <x1> <y0> <y1> <z> <expected>
// Exit status 0 when every case passes.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "cpvs_amd.hpp"

using namespace cpvs_amd;

struct V3 {
    float x, y, z;
};

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s cases.txt\n", argv[0]);
        return 2;
    }
    std::ifstream in(argv[1]);
    std::map<std::string, std::unique_ptr<ImageF>> images;
    std::map<std::string, std::unique_ptr<MinMaxHierarchy>> pyramids;
    std::map<std::string, std::unique_ptr<CompressedShadow>> shadows;
    auto pyr = [&](const std::string& n) -> MinMaxHierarchy& {
        auto& p = pyramids[n];
        if (!p) p.reset(new MinMaxHierarchy(*images.at(n)));
        return *p;
    };
    auto dag = [&](const std::string& n) -> CompressedShadow& {
        auto& s = shadows[n];
        if (!s) s = CompressedShadow::create(pyr(n));
        return *s;
    };
    // The same map built without leafmasks (the reference compiled without LEAFMASKS); its
    // traverse(ndc, false) walks to level 0 and must give the same visibilities.
    std::map<std::string, std::unique_ptr<CompressedShadow>> shadows_off;
    auto dag_off = [&](const std::string& n) -> CompressedShadow& {
        auto& s = shadows_off[n];
        if (!s) {
            Context::instance().setLeafmasks(false);
            try {
                s = CompressedShadow::create(pyr(n));
            } catch (...) {
                Context::instance().setLeafmasks(true);
                throw;
            }
            Context::instance().setLeafmasks(true);
        }
        return *s;
    };
    int fails = 0, cases = 0;
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        std::string kind, name;
        ss >> kind >> name;
        try {
            if (kind == "image") {
                size_t N;
                ss >> N;
                std::unique_ptr<ImageF> img(new ImageF(N, N, 1));
                for (size_t i = 0; i < N * N; ++i) ss >> img->data()[i];
                images[name] = std::move(img);
            } else if (kind == "levels") {
                int exp;
                ss >> exp;
                ++cases;
                if (pyr(name).getNumLevels() != exp) {
                    std::printf("FAIL levels %s\n", name.c_str());
                    ++fails;
                }
            } else if (kind == "minmax") {
                size_t L, x, y;
                std::string which;
                double exp, tol;
                ss >> L >> x >> y >> which >> exp >> tol;
                ++cases;
                const float v = which == "min" ? pyr(name).getMin(L, x, y) : pyr(name).getMax(L, x, y);
                const bool ok = tol == 0 ? v == (float)exp : std::fabs(v - exp) < tol;
                if (!ok) {
                    std::printf("FAIL minmax %s %zu %zu %zu %s: %g\n", name.c_str(), L, x, y, which.c_str(), v);
                    ++fails;
                }
            } else if (kind == "traverse") {
                V3 p;
                int tryLeaf, exp;
                ss >> p.x >> p.y >> p.z >> tryLeaf >> exp;
                ++cases;
                if (dag(name).traverse(p, tryLeaf != 0) != exp) {
                    std::printf("FAIL traverse %s (%g,%g,%g)\n", name.c_str(), p.x, p.y, p.z);
                    ++fails;
                }
            } else if (kind == "sweep") {
                int x0, x1, y0, y1, exp;
                float z;
                ss >> x0 >> x1 >> y0 >> y1 >> z >> exp;
                for (int y = y0; y < y1; ++y)
                    for (int x = x0; x < x1; ++x) {
                        ++cases;
                        // convertToNdc(vec3(x/32, y/32, z)) = v * 2 - 1 (test/CompressedShadowTest.cpp:46-48)
                        V3 p{(x / 32.0f) * 2.0f + -1.0f, (y / 32.0f) * 2.0f + -1.0f, z * 2.0f + -1.0f};
                        if (dag(name).traverse(p) != exp) {
                            std::printf("FAIL sweep %s %d %d\n", name.c_str(), x, y);
                            ++fails;
                        }
                        ++cases;
                        if (dag_off(name).traverse(p, false) != exp) {
                            std::printf("FAIL sweep without leafmasks %s %d %d\n", name.c_str(), x, y);
                            ++fails;
                        }
                    }
            }
        } catch (const Error& e) {
            std::printf("ERROR %s %s: %s\n", kind.c_str(), name.c_str(), e.what());
            ++fails;
        }
    }
    // Container round trip through the facade: total visibility and DAG words survive the move.
    try {
        auto& d = dag("depths32x32");
        const size_t words = d.getDAG().size();
        std::unique_ptr<CompressedShadow> copy = CompressedShadow::create(pyr("depths32x32"));
        CompressedShadowContainer c(std::move(copy));
        c.setFilterSize(3);
        // get(x, y, z) is a CompressedShadow of the container's (src/CompressedShadowContainer.h:41-46).
        const CompressedShadow* in = c.get(0, 0, 0);
        ++cases;
        if (!in || in->getDAG() != d.getDAG() || in->getNumLevels() != d.getNumLevels() ||
            in->getTotalVisibility() != d.getTotalVisibility()) {
            std::printf("FAIL container get\n");
            ++fails;
        }
        // The lookup top table (no counterpart in the reference): same visibilities with and
        // without it, over a grid of positions in and around the unit cube.
        const int G = 24, n = G * G * G;
        std::vector<float> pos(4 * (size_t)n);
        for (int i = 0; i < n; ++i) {
            pos[4 * i + 0] = -1.1f + 2.2f * (float)(i % G) / (G - 1);
            pos[4 * i + 1] = -1.1f + 2.2f * (float)((i / G) % G) / (G - 1);
            pos[4 * i + 2] = -1.1f + 2.2f * (float)(i / (G * G)) / (G - 1);
            pos[4 * i + 3] = 1.0f;
        }
        cpvs_ctx* cx = Context::instance().get();
        void *dpos = nullptr, *dvis = nullptr;
        check(cpvs_device_alloc(cx, 16ull * n, &dpos), "cpvs_device_alloc");
        check(cpvs_device_alloc(cx, (uint64_t)n, &dvis), "cpvs_device_alloc");
        check(cpvs_memcpy_h2d(cx, dpos, pos.data(), 16ull * n), "cpvs_memcpy_h2d");
        std::vector<uint8_t> v0(n), v2(n);
        c.setTopLevels(0);
        c.copyToGPU();
        c.evaluate(static_cast<const float*>(dpos), 4, n, nullptr, static_cast<uint8_t*>(dvis));
        check(cpvs_memcpy_d2h(cx, v0.data(), dvis, n), "cpvs_memcpy_d2h");
        c.setTopLevels(2);  // 32^2: two inner levels above the last
        c.copyToGPU();
        c.evaluate(static_cast<const float*>(dpos), 4, n, nullptr, static_cast<uint8_t*>(dvis));
        check(cpvs_memcpy_d2h(cx, v2.data(), dvis, n), "cpvs_memcpy_d2h");
        cpvs_device_free(cx, dpos);
        cpvs_device_free(cx, dvis);
        ++cases;
        if (c.topLevels() != 2 || v0 != v2) {
            std::printf("FAIL container top table (levels %d)\n", c.topLevels());
            ++fails;
        }
        c.moveToGPU();
        ++cases;
        if (c.get(0, 0, 0) != nullptr || words == 0) {
            std::printf("FAIL container state\n");
            ++fails;
        }
    } catch (const Error& e) {
        std::printf("ERROR container: %s\n", e.what());
        ++fails;
    }
    // DAG file round trip through the facade.
    try {
        auto& d = dag("depths32x32");
        const char* tmp = std::getenv("TMPDIR");
        const std::string path = std::string(tmp ? tmp : "/tmp") + "/cpvs_facade_dag.bin";
        d.save(path);
        std::unique_ptr<CompressedShadow> back = CompressedShadow::load(path);
        std::remove(path.c_str());
        ++cases;
        if (back->getDAG() != d.getDAG() || back->getNumLevels() != d.getNumLevels() ||
            back->getTotalVisibility() != d.getTotalVisibility()) {
            std::printf("FAIL dag file round trip\n");
            ++fails;
        }
    } catch (const Error& e) {
        std::printf("ERROR dag file: %s\n", e.what());
        ++fails;
    }
    std::printf("%d cases, %d failures\n", cases, fails);
    return fails ? 1 : 0;
}
