// poisson_recon: command-line front end over the C ABI; the drop-in for the
// reference's main() (main.cu:3247-4573), whose input/output paths and depth
// are compile-time constants (main.cu:69,3251-3252).
//
//   poisson_recon --in points.ply --out mesh.ply [--depth 8] [--tol 1e-6]
//                 [--maxIters 2000] [--refresh 20] [--device 0] [--json]
//                 [--no-subdivide] [--weld] [--edgeCollapse R | --triangleCollapse R]
//
// The mesh holds the leaf-level batch and the coarse-leaf subdivision batches,
// as the reference's main() always writes (main.cu:3799-4561);
// --no-subdivide keeps only the leaf-level batch (main.cu:3793), and
// --subdivide is accepted for older scripts (it is the default).  --weld
// merges the vertices the batches duplicate (prg_weld_mesh, which the
// reference does not do).  --edgeCollapse R / --triangleCollapse R decimate
// the mesh afterwards (prg_decimate_mesh: EdgeCollapse / TriangleCollapse of
// Geometry.inl:50-254, which the reference's main() does not call).
//
// Input formats by extension (.ply, .bnpts, ASCII otherwise, main.cu:517-520);
// output is an ASCII PLY triangle mesh (PlyWriteTriangles, plyfile.cu:2769).
// Stage times are printed with the reference's stage names.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/prg.h"

namespace {

void usage() {
    std::fprintf(stderr,
                 "usage: poisson_recon --in <points> --out <mesh.ply> [--depth D] [--tol T] [--maxIters N]\n"
                 "                     [--refresh K] [--device I] [--json] [--no-subdivide] [--weld]\n"
                 "                     [--edgeCollapse R | --triangleCollapse R]\n");
}

const char* kStageNames[PRG_STAGE_COUNT] = {
    "Read + normalise", "GPU NodeArray build", "Inner product tables", "Compute Vector Field",
    "Compute divergence", "GPU Laplacian Iteration", "GPU calculate isoValue", "Build Vertex/Edge arrays",
    "Marching cubes evaluation", "Coarse-leaf subdivision", "Copy mesh to host"};

}  // namespace

int main(int argc, char** argv) {
    std::string in, out;
    prg_config cfg;
    prg_config_default(&cfg);
    bool json = false, subdivide = true, weld = false;
    int decimate_mode = -1;
    float decimate_ratio = 0.0f;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        auto next = [&]() -> const char* {
            if (i + 1 >= argc) {
                usage();
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "--in") in = next();
        else if (a == "--out") out = next();
        else if (a == "--depth") cfg.max_depth = std::atoi(next());
        else if (a == "--tol") cfg.cg_tol = float(std::atof(next()));
        else if (a == "--maxIters") cfg.cg_max_iters = std::atoi(next());
        else if (a == "--refresh") cfg.cg_refresh = std::atoi(next());
        else if (a == "--device") cfg.device = std::atoi(next());
        else if (a == "--json") json = true;
        else if (a == "--subdivide") subdivide = true;
        else if (a == "--no-subdivide") subdivide = false;
        else if (a == "--weld") weld = true;
        else if (a == "--edgeCollapse" || a == "--triangleCollapse") {
            decimate_mode = a == "--edgeCollapse" ? 0 : 1;
            decimate_ratio = float(std::atof(next()));
            if (!(decimate_ratio > 0.0f)) {
                usage();
                return 2;
            }
        }
        else {
            usage();
            return 2;
        }
    }
    if (in.empty() || out.empty()) {
        usage();
        return 2;
    }
    float *xyz = nullptr, *nxyz = nullptr;
    int64_t n = 0;
    prg_status st = prg_read_points(in.c_str(), &xyz, &nxyz, &n);
    if (st != PRG_OK) {
        std::fprintf(stderr, "cannot read %s (status %d)\n", in.c_str(), int(st));
        return 1;
    }
    prg_ctx* ctx = nullptr;
    st = prg_create(&ctx, &cfg);
    if (st != PRG_OK) {
        std::fprintf(stderr, "prg_create failed (status %d)\n", int(st));
        prg_points_release(xyz, nxyz);
        return 1;
    }
    prg_mesh mesh{};
    prg_solve_stats stats[11] = {};
    st = prg_set_points(ctx, xyz, nxyz, n, 0);
    prg_points_release(xyz, nxyz);
    if (st == PRG_OK) st = prg_reconstruct(ctx, subdivide ? 1 : 0, &mesh, stats);
    if (st != PRG_OK && st != PRG_E_NOT_CONVERGED) {
        std::fprintf(stderr, "reconstruction failed (status %d): %s\n", int(st), prg_last_error(ctx));
        prg_destroy(ctx);
        return 1;
    }
    if (st == PRG_E_NOT_CONVERGED) std::fprintf(stderr, "warning: %s\n", prg_last_error(ctx));
    int64_t merged = 0;
    if (weld && prg_weld_mesh(ctx, &mesh, &merged) != PRG_OK) {
        std::fprintf(stderr, "welding failed: %s\n", prg_last_error(ctx));
        prg_mesh_release(&mesh);
        prg_destroy(ctx);
        return 1;
    }
    if (decimate_mode >= 0 && prg_decimate_mesh(&mesh, decimate_mode, decimate_ratio) != PRG_OK) {
        std::fprintf(stderr, "decimation failed\n");
        prg_mesh_release(&mesh);
        prg_destroy(ctx);
        return 1;
    }
    prg_status ws = prg_write_mesh_ply(out.c_str(), &mesh);
    prg_timings t{};
    prg_stage_times(ctx, &t);
    double total = 0.0;
    for (int s = 0; s < PRG_STAGE_COUNT; s++) total += t.ms[s];
    if (json) {
        std::printf("{\"points\": %lld, \"depth\": %d, \"vertices\": %lld, \"triangles\": %lld, \"stage_ms\": {",
                    (long long)n, cfg.max_depth, (long long)mesh.n_vertices, (long long)mesh.n_triangles);
        for (int s = 0; s < PRG_STAGE_COUNT; s++) std::printf("%s\"%d\": %.3f", s ? ", " : "", s, t.ms[s]);
        std::printf("}, \"cg_iters\": [");
        for (int d = 0; d <= cfg.max_depth; d++) std::printf("%s%d", d ? ", " : "", stats[d].iters);
        std::printf("]}\n");
    } else {
        std::printf("Total points number:%lld\n", (long long)n);
        for (int s = 0; s < PRG_STAGE_COUNT; s++) std::printf("%s takes:%.3fms\n", kStageNames[s], t.ms[s]);
        for (int d = 0; d <= cfg.max_depth; d++)
            std::printf("Depth %d: %lld nodes, %d CG iterations, relative residual %.2e\n", d,
                        (long long)stats[d].nodes, stats[d].iters, stats[d].true_rel_residual);
        std::printf("Vertices:%lld Triangles:%lld Total:%.3fms\n", (long long)mesh.n_vertices,
                    (long long)mesh.n_triangles, total);
        if (weld) std::printf("Welded vertices:%lld\n", (long long)merged);
    }
    prg_mesh_release(&mesh);
    prg_destroy(ctx);
    if (ws != PRG_OK) {
        std::fprintf(stderr, "cannot write %s\n", out.c_str());
        return 1;
    }
    return 0;
}
