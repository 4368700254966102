/*
 * ref_harness.cpp -- drives the reference's own classes to produce golden
 * vectors.  TEST INFRASTRUCTURE ONLY: linked with the reference sources
 * compiled where they lie (see oracle/Makefile); the binary lands in
 * oracle/_ref/ and is only run by tests/golden/make_golden.py,
 * tests/test_oracle_vs_ref.py and bench.py's cpu_baseline leg.
 *
 * usage: ref_harness <case.json> <nsteps> <out.bin> [fields_in.bin]
 *
 * The case is built with the reference's JSONParser factories
 * (src/fvm_2d_cart/simulation/parser/parser.cpp:104-488) and initialised as
 * Simulation's constructor does (simulation.cpp:43).  A fields file, when
 * given, replaces every block's padded fields (as MeshBlock::set_fields).
 * The step loop is Simulation::solve's body (simulation.cpp:85-110) without
 * the t_end/frame clamps and without output.  A TimeStep subclass records
 * every call (step and MUSCL predictor calls).
 *
 * out.bin (little endian):
 *   int32 nblocks; per block int32 nx, ny, g;
 *   per block double[4][ny+2g][nx+2g] fields after nsteps ([var][y][x]);
 *   int32 ncalls; double[ncalls] every TimeStep result;
 *   int32 nsteps; double[nsteps] step dts;
 *   double elapsed seconds of the step loop.
 */
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <functional>
#include <iostream>
#include <vector>

#include "simulation/parser/parser.h"

using namespace HyperFlow;

namespace {

class RecordingTimeStep : public TimeStep {
public:
    RecordingTimeStep(const std::shared_ptr<Model>& model, double cfl) : TimeStep(model, cfl) {}
    double operator()(const std::shared_ptr<Mesh>& mesh) override
    {
        double dt = TimeStep::operator()(mesh);
        calls.push_back(dt);
        return dt;
    }
    std::vector<double> calls;
};

template <typename T> void put(std::FILE* f, T v) { std::fwrite(&v, sizeof v, 1, f); }

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s case.json nsteps out.bin [fields_in.bin]\n", argv[0]);
        return 2;
    }
    json config;
    {
        std::ifstream jf(argv[1]);
        jf >> config;
    }
    const int nsteps = std::atoi(argv[2]);

    JSONParser parser(config);
    std::shared_ptr<Model> model = parser.generate_model();
    std::shared_ptr<Mesh> mesh = parser.generate_mesh();
    std::shared_ptr<InitialCondition> ic = parser.generate_initial_condition();
    std::shared_ptr<RiemannSolver> riemann = parser.generate_riemann_solver();
    auto ts = std::make_shared<RecordingTimeStep>(model, parser.get_cfl());
    std::shared_ptr<Scheme> scheme = parser.generate_spatial_scheme(model, riemann, ts, mesh);
    std::shared_ptr<ODESolver> ode = parser.generate_ode_solver();
    if (!model || !mesh || !riemann || !scheme || !ode) {
        std::fprintf(stderr, "unsupported case\n");
        return 3;
    }
    if (ic) mesh->initialise_field_values(ic);

    std::vector<std::shared_ptr<MeshBlock> > blocks = mesh->get_mesh_blocks();
    if (argc > 4) {
        std::FILE* fin = std::fopen(argv[4], "rb");
        if (!fin) { std::perror("fields_in"); return 4; }
        for (auto& mb : blocks) {
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
    auto t0 = std::chrono::steady_clock::now();
    for (int n = 0; n < nsteps; n++) {
        mesh->apply_bcs();
        double dt = scheme->calculate_time_step(mesh);
        step_dts.push_back(dt);
        for (auto& mb : blocks) {
            auto evolve = std::bind(&Scheme::operator(), scheme, mb);
            Vec3D nf = ode->operator()(evolve, mb->get_fields(), dt);
            mb->set_fields(nf);
        }
    }
    double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

    std::FILE* fo = std::fopen(argv[3], "wb");
    if (!fo) { std::perror("out"); return 5; }
    put<int>(fo, (int)blocks.size());
    for (auto& mb : blocks) {
        put<int>(fo, (int)mb->get_x_cells());
        put<int>(fo, (int)mb->get_y_cells());
        put<int>(fo, (int)mb->get_ghost_cells());
    }
    for (auto& mb : blocks) {
        const unsigned tx = mb->get_total_x_cells(), ty = mb->get_total_y_cells();
        const Vec3D& f = mb->get_fields();
        std::vector<double> buf((size_t)4 * tx * ty);
        for (unsigned i = 0; i < tx; i++)
            for (unsigned j = 0; j < ty; j++)
                for (unsigned v = 0; v < 4; v++) buf[((size_t)v * ty + j) * tx + i] = f[i][j][v];
        std::fwrite(buf.data(), sizeof(double), buf.size(), fo);
    }
    put<int>(fo, (int)ts->calls.size());
    std::fwrite(ts->calls.data(), sizeof(double), ts->calls.size(), fo);
    put<int>(fo, (int)step_dts.size());
    std::fwrite(step_dts.data(), sizeof(double), step_dts.size(), fo);
    put<double>(fo, elapsed);
    std::fclose(fo);
    return 0;
}
