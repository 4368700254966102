// ref_loop.cpp -- test driver for the reference-typed C++ surface of
// include/hyperflow.hpp.  TEST INFRASTRUCTURE: built by the package build into
// tests/bin/, run by tests/test_gpu_cpp_api.py.
//
// It builds a case with the JSONParser factories and then runs the step body
// of the reference's Simulation::solve (simulation.cpp:85-110, without the
// t_end/frame clamps and output) through the signatures a caller of the
// reference uses: Scheme::operator() bound into a std::function<Vec3D()>,
// ODESolver::operator()(func, MeshBlock::get_fields(), dt) and
// MeshBlock::set_fields(Vec3D).
//
//   ref_loop case.json nsteps out.bin [fields_in.bin]
//
// fields_in.bin (optional): every block's padded fields [4][ny+2g][nx+2g],
// set with MeshBlock::set_fields(Vec3D) after the initial condition.
// out.bin: the layout oracle/ref_harness writes (little endian):
//   int32 nblocks; per block int32 nx, ny, g; per block double[4][ny+2g][nx+2g];
//   int32 0 (TimeStep calls are not recorded); int32 nsteps; double[nsteps]
//   step dts; double elapsed seconds.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <memory>
#include <vector>

#include "../../include/hyperflow.hpp"

using namespace HyperFlow;

namespace {
template <typename T> void put(std::FILE* f, T v) { std::fwrite(&v, sizeof v, 1, f); }
}  // namespace

int main(int argc, char** argv)
{
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s case.json nsteps out.bin [fields_in.bin]\n", argv[0]);
        return 2;
    }
    try {
        JSONParser parser = JSONParser::from_file(argv[1]);
        const int nsteps = std::atoi(argv[2]);
        std::shared_ptr<Model> model = parser.generate_model();
        std::shared_ptr<Mesh> mesh = parser.generate_mesh();
        std::shared_ptr<InitialCondition> ic = parser.generate_initial_condition();
        std::shared_ptr<RiemannSolver> riemann = parser.generate_riemann_solver();
        std::shared_ptr<TimeStep> time_step = parser.generate_time_step();
        std::shared_ptr<Scheme> scheme = parser.generate_spatial_scheme(model, riemann, time_step, mesh);
        std::shared_ptr<ODESolver> ode_solver = parser.generate_ode_solver();
        if (!model || !mesh || !riemann || !time_step || !scheme || !ode_solver) {
            std::fprintf(stderr, "unsupported case\n");
            return 3;
        }
        if (ic) mesh->initialise_field_values(ic);
        mesh->attach(*scheme, *model, *time_step, parser.get_device_options());
        std::vector<std::shared_ptr<MeshBlock> >& mesh_blocks = mesh->get_mesh_blocks();
        if (argc > 4) {
            std::FILE* fin = std::fopen(argv[4], "rb");
            if (!fin) {
                std::perror("fields_in");
                return 4;
            }
            for (auto& mb : mesh_blocks) {
                const unsigned tx = mb->get_total_x_cells(), ty = mb->get_total_y_cells();
                std::vector<double> buf((size_t)4 * tx * ty);
                if (std::fread(buf.data(), sizeof(double), buf.size(), fin) != buf.size()) {
                    std::fprintf(stderr, "short fields file\n");
                    return 4;
                }
                Vec3D f(tx, Vec2D(ty, Vec1D(4)));
                for (unsigned i = 0; i < tx; i++)
                    for (unsigned j = 0; j < ty; j++)
                        for (unsigned v = 0; v < 4; v++) f[i][j][v] = buf[((size_t)v * ty + j) * tx + i];
                mb->set_fields(f);
            }
            std::fclose(fin);
        }

        std::vector<double> step_dts;
        const auto t0 = std::chrono::steady_clock::now();
        for (int n = 0; n < nsteps; n++) {
            mesh->apply_bcs();
            const double dt = scheme->calculate_time_step(mesh);
            step_dts.push_back(dt);
            for (unsigned int b = 0; b < mesh_blocks.size(); b++) {
                auto evolve_func = std::bind(&Scheme::operator(), scheme, mesh_blocks[b]);
                Vec3D new_fields = ode_solver->operator()(evolve_func, mesh_blocks[b]->get_fields(), dt);
                mesh_blocks[b]->set_fields(new_fields);
            }
        }
        const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

        std::FILE* fo = std::fopen(argv[3], "wb");
        if (!fo) {
            std::perror("out");
            return 5;
        }
        put<int>(fo, (int)mesh_blocks.size());
        for (auto& mb : mesh_blocks) {
            put<int>(fo, (int)mb->get_x_cells());
            put<int>(fo, (int)mb->get_y_cells());
            put<int>(fo, (int)mb->get_ghost_cells());
        }
        for (auto& mb : mesh_blocks) {
            const unsigned tx = mb->get_total_x_cells(), ty = mb->get_total_y_cells();
            const Vec3D& f = mb->get_fields();
            std::vector<double> buf((size_t)4 * tx * ty);
            for (unsigned i = 0; i < tx; i++)
                for (unsigned j = 0; j < ty; j++)
                    for (unsigned v = 0; v < 4; v++) buf[((size_t)v * ty + j) * tx + i] = f[i][j][v];
            std::fwrite(buf.data(), sizeof(double), buf.size(), fo);
        }
        put<int>(fo, 0);
        put<int>(fo, (int)step_dts.size());
        std::fwrite(step_dts.data(), sizeof(double), step_dts.size(), fo);
        put<double>(fo, elapsed);
        std::fclose(fo);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "ref_loop: %s\n", e.what());
        return 1;
    }
    return 0;
}
