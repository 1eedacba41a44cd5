// Host-side checks of libdwg_hip (no GPU needed), built with AddressSanitizer
// and UndefinedBehaviorSanitizer on the host code: the .xyz reader on
// malformed files, the voxel grid on degenerate boxes, the STL/.xyz writers,
// and the C ABI's argument checks (null handles, bad sizes) that must return
// a status instead of crashing.  Prints "host checks ok" and exits 0 when
// every check holds.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../dwg_amd/csrc/turbo_dwg.h"

static int failures = 0;
#define CHECK(c)                                                  \
    do {                                                          \
        if (!(c)) {                                               \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            ++failures;                                           \
        }                                                         \
    } while (0)

static std::string write(const std::string& dir, const char* name, const char* text) {
    const std::string p = dir + "/" + name;
    std::ofstream(p) << text;
    return p;
}

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : "/tmp";
    std::vector<float> P, N;
    // .xyz reader: 6 columns, 3 columns (normals 0), blank and garbage lines skipped
    CHECK(dwg_host::read_xyz(write(dir, "six.xyz", "0 0 0 1 0 0\n1 2 3 0 1 0\n"), P, N));
    CHECK(P.size() == 6 && N.size() == 6 && P[5] == 3.0f && N[4] == 1.0f);
    CHECK(dwg_host::read_xyz(write(dir, "three.xyz", "1 2 3\n4 5 6\n"), P, N));
    CHECK(P.size() == 6 && N.size() == 6 && N[0] == 0.0f);
    CHECK(dwg_host::read_xyz(write(dir, "mixed.xyz", "\n# comment\n1 2 3 4 5 6\nnot numbers\n7 8\n9 10 11\n"), P, N));
    CHECK(P.size() == 6);
    CHECK(dwg_host::read_xyz(write(dir, "empty.xyz", ""), P, N) && P.empty());
    CHECK(!dwg_host::read_xyz(dir + "/missing.xyz", P, N));
    // voxel grid: a flat box, a single point (offset only), a tiny resolution
    const double lo[3] = {0, 0, 0}, hi[3] = {1, 0.5, 0};
    dwg_host::Grid g = dwg_host::voxel_grid_bounds(lo, hi, 0.05, 32);
    CHECK(g.res[0] == 32 && g.res[1] > 2 && g.res[2] >= 3 && g.h > 0);
    g = dwg_host::voxel_grid_bounds(lo, lo, 0.05, 16);
    CHECK(g.res[0] == 16 && g.res[1] == 16 && g.res[2] == 16 && std::isfinite(g.h));
    bool threw = false;
    try {
        dwg_host::voxel_grid_bounds(lo, hi, 0.05, 3);
    } catch (const std::exception&) {
        threw = true;
    }
    CHECK(threw);
    threw = false;
    try {
        dwg_host::voxel_grid(nullptr, 0, 0.05, 16);
    } catch (const std::exception&) {
        threw = true;
    }
    CHECK(threw);
    // writers: a degenerate triangle gets a zero normal, not NaN
    const float tri[18] = {0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    dwg_host::save_binary_stl(dir + "/t.stl", tri, 2);
    std::ifstream f(dir + "/t.stl", std::ios::binary);
    std::vector<char> b((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    CHECK(b.size() == 84 + 2 * 50);
    float n2[3];
    std::memcpy(n2, b.data() + 84 + 50, 12);
    CHECK(n2[0] == 0.f && n2[1] == 0.f && n2[2] == 0.f);
    dwg_host::save_xyz(dir + "/t.xyz", tri, tri + 9, 3);
    CHECK(dwg_host::read_xyz(dir + "/t.xyz", P, N) && P.size() == 9);
    threw = false;
    try {
        dwg_host::save_binary_stl(dir + "/no/such/dir/t.stl", tri, 1);
    } catch (const std::exception&) {
        threw = true;
    }
    CHECK(threw);
    // C ABI argument checks without a device
    dwg_ctx* ctx = nullptr;
    CHECK(dwg_ctx_create(0, nullptr, nullptr) == DWG_ERR_ARG);
    CHECK(dwg_launch(nullptr, nullptr, nullptr, 0, nullptr, nullptr, nullptr) == DWG_ERR_ARG);
    CHECK(dwg_octree_info(nullptr, nullptr, nullptr, nullptr) == DWG_ERR_ARG);
    CHECK(dwg_gwn_direct(nullptr, nullptr, nullptr, nullptr, nullptr, 0, nullptr, 0, 0.f, 0, nullptr, nullptr) ==
          DWG_ERR_ARG);
    CHECK(dwg_knn_query(nullptr, nullptr, nullptr, 0, 10, nullptr, nullptr, nullptr, nullptr) == DWG_ERR_ARG);
    CHECK(std::strcmp(dwg_ctx_error(nullptr), "null context") == 0);
    dwg_options o;
    dwg_options_default(&o);
    CHECK(o.total_iter == 40 && o.grid_reso == 400 && o.final_reso == 500 && o.beta > 2.29f && o.residual_tol == 0.f);
    dwg_free(nullptr);
    dwg_ctx_destroy(nullptr);
    (void)ctx;
    if (failures) return 1;
    std::printf("host checks ok\n");
    return 0;
}
