// DWG_HIP: the reference's DWG_CUDA command line (src/main.cpp:17-166) on
// libdwg_hip.  Same flags, defaults, output naming and --help text; extra
// flags for the MI355X build are listed after the reference ones.
#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/dwg_hip.h"
#include "turbo_dwg.h"

static void usage() {
    std::cout << "--in_path   The absolute path of the input point cloud\n"
                 "--in_name   The file name of the input point cloud (only supports .xyz currently)\n"
                 "--out_path   The absolute path of the output mesh/point cloud with normals\n"
                 "--output_mode   Ouput mode: 1 Mesh(.stl) (default)   2 Point cloud with normals(.xyz)   3 Both 1 and 2\n"
                 "--in_normal   Initialization mode: 0 Random   1 Input   3 Gauss map (default)\n"
                 "--lambda   Screening Coefficient\n"
                 "--max_depth   The max depth of octree\n"
                 "--use_radii   Turn on screening mode or not (0 turn off   1 turn on)\n"
                 "--kw   The number of nearest neighbors for the iteration (default 10)\n"
                 "--final_kw   The number of nearest neighbors for the final output, usually be same to kw (default 10)\n"
                 "--grid_reso   The resolution of Marching Cubes for the iteration (e.g. 200 for <10w size; 450 for >1000w size)\n"
                 "--final_reso   The resolution of Marching Cubes for the final output (e.g. 260 for <10w size; 700 for >1000w size)\n"
                 "--total_iter   The number of iteration\n"
                 "--test_iter   Output the result each [test_iter] iterations (it is turned off by default)\n"
                 "--com_As   Computing area weight or not (0 turn off   1 turn on)\n"
                 "--use_leaves_generate   Recommend to open it (0 turn off   1 turn on), when the point cloud noise is high\n"
                 "MI355X build:\n"
                 "--in_normal   also 2 Random (PCA normals for the areas with --com_As 1)   4 Learned MLP (needs --mlp)   5 PCA (k=10)\n"
                 "--source_mode   0 octree-leaf centres as sources (default)   1 the input points\n"
                 "--leaf_size   points per octree leaf in the winding-number sums (default 1 = octreeDWG)\n"
                 "--beta   Barnes-Hut accuracy (default 2.3; <= 0 direct sums)\n"
                 "--kahan   1: compensated accumulation\n"
                 "--bf16_far   1: bf16 far-field moments\n"
                 "--near_fp64   1: point terms inside the screening radius in FP64\n"
                 "--seed   seed of the random initial normals\n"
                 "--residual_tol   > 0: stop after an even iteration count whose convergence residual is below it\n"
                 "--mlp   learned-initialisation weights (.bin) for --in_normal 4\n"
                 "--verbose   1: per-iteration log\n";
}

int main(int argc, char* argv[]) {
    if (argc < 2) {
        usage();
        return 1;
    }
    if (std::strcmp(argv[1], "--help") == 0) {
        usage();
        return 0;
    }
    std::string in_path = " ", in_name = " ", out_path = " ", mlp_path;
    dwg_options opt;
    dwg_options_default(&opt);
    if ((argc - 1) % 2 != 0) {
        std::cerr << "option " << argv[argc - 1] << " has no value\n";
        return 1;
    }
    for (int i = 1; i + 1 < argc; i += 2) {
        const std::string k = argv[i];
        const char* v = argv[i + 1];
        try {
            if (k == "--in_path") in_path = v;
            else if (k == "--in_name") in_name = v;
            else if (k == "--out_path") out_path = v;
            else if (k == "--in_normal") opt.input_normal = std::stoi(v);
            else if (k == "--grid_reso") opt.grid_reso = std::stoi(v);
            else if (k == "--final_reso") opt.final_reso = std::stoi(v);
            else if (k == "--max_depth") opt.max_depth = std::stoi(v);
            else if (k == "--kw") opt.kw = std::stoi(v);
            else if (k == "--final_kw") opt.final_kw = std::stoi(v);
            else if (k == "--normal_kw") opt.normal_kw = std::stoi(v);
            else if (k == "--lambda") opt.sigma = std::stof(v);
            else if (k == "--min_num") opt.min_num = std::stoi(v);
            else if (k == "--total_iter") opt.total_iter = std::stoi(v);
            else if (k == "--output_mode") opt.output_mode = std::stoi(v);
            else if (k == "--use_leaves_generate") opt.use_leaves_generate = std::stoi(v);
            else if (k == "--use_radii") opt.use_radii = std::stoi(v);
            else if (k == "--com_As") opt.com_As = std::stoi(v);
            else if (k == "--As_thres") opt.As_threshold = std::stof(v);
            else if (k == "--offset_1") opt.offset_1 = std::stof(v);
            else if (k == "--offset_2") opt.offset_2 = std::stof(v);
            else if (k == "--test_iter") opt.test_iter = std::stoi(v);
            else if (k == "--source_mode") opt.source_mode = std::stoi(v);
            else if (k == "--leaf_size") opt.leaf_size = std::stoi(v);
            else if (k == "--beta") opt.beta = std::stof(v);
            else if (k == "--kahan") opt.precision = std::stoi(v) ? (opt.precision | DWG_KAHAN) : (opt.precision & ~DWG_KAHAN);
            else if (k == "--bf16_far") opt.precision = std::stoi(v) ? (opt.precision | DWG_FAR_BF16) : (opt.precision & ~DWG_FAR_BF16);
            else if (k == "--near_fp64") opt.precision = std::stoi(v) ? (opt.precision | DWG_NEAR_FP64) : (opt.precision & ~DWG_NEAR_FP64);
            else if (k == "--seed") opt.seed = std::stoull(v);
            else if (k == "--residual_tol") opt.residual_tol = std::stof(v);
            else if (k == "--mlp") mlp_path = v;
            else if (k == "--verbose") opt.verbose = std::stoi(v);
        } catch (const std::exception&) {
            std::cerr << "bad value '" << v << "' for option " << k << "\n";
            return 1;
        }
    }
    if (opt.total_iter < 1) {
        std::cerr << "--total_iter must be at least 1\n";
        return 1;
    }
    // output naming (src/main.cpp:151-156)
    const std::string full_in = in_path + "/" + in_name;
    const std::string pure = in_name.size() > 4 ? in_name.substr(0, in_name.size() - 4) : in_name;
    const std::string mesh_name = pure + "_md_" + std::to_string(opt.max_depth) + "_gr_" + std::to_string(opt.grid_reso) +
                                  "_iter_" + std::to_string(opt.total_iter);
    const std::string out_base = out_path + "/" + mesh_name;
    if (opt.com_As == 1 && opt.input_normal == 0) opt.input_normal = 2;  // src/main.cpp:158-161

    std::vector<float> pts, nrm;
    if (!dwg_host::read_xyz(full_in, pts, nrm) || pts.empty()) {
        std::cerr << "cannot read point cloud " << full_in << "\n";
        return 2;
    }
    const int64_t n = int64_t(pts.size() / 3);
    std::cout << "point cloud cnt = " << n << "\n";
    dwg_ctx* ctx = nullptr;
    if (dwg_ctx_create(0, nullptr, &ctx) != DWG_OK) {
        std::cerr << "no HIP device\n";
        return 3;
    }
    dwg_mlp* mlp = nullptr;
    if (!mlp_path.empty()) {
        if (dwg_mlp_load(ctx, mlp_path.c_str(), &mlp) != DWG_OK) {
            std::cerr << "cannot load MLP weights " << mlp_path << ": " << dwg_ctx_error(ctx) << "\n";
            dwg_ctx_destroy(ctx);
            return 4;
        }
        opt.mlp = mlp;
    }
    dwg_result res;
    std::memset(&res, 0, sizeof(res));
    const int rc = dwg_launch(ctx, pts.data(), opt.input_normal == 1 ? nrm.data() : nullptr, n, &opt, out_base.c_str(),
                              &res);
    if (rc != DWG_OK) {
        std::cerr << "DWG failed: " << dwg_ctx_error(ctx) << "\n";
    } else {
        std::cout << "Num triangles: " << res.tri_count << "\n";
        std::printf("Prepare Time = %.3f s\nIter Time = %.3f s\nFinal Time = %.3f s\nTotal Time (Wall time) = %.3f s\n",
                    res.time_setup_s, res.time_iter_s, res.time_final_s, res.time_total_s);
    }
    if (mlp) dwg_mlp_destroy(mlp);
    dwg_ctx_destroy(ctx);
    return rc == DWG_OK ? 0 : 5;
}
