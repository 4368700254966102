// ref_uns_harness.cpp -- drives the reference's unstructured solver
// (src/fvm_2d_uns) step by step, for the oracle's golden vectors.  Test
// infrastructure only; built by oracle/Makefile into oracle/_ref/.
//
//   ref_uns_harness case.json nsteps out.bin [fields.bin]
//
// The case is the reference's own JSON format ("mesh.file" names the .hmsh
// mesh).  The objects are made by the reference's JSONParser factories and
// stepped like Simulation::solve (src/fvm_2d_uns/simulation/simulation/
// simulation.cpp:85-118) without the t_end / frame clamps or output.  The
// mesh loader's progress lines go to /dev/null.  With fields.bin (f64
// [ncells][4], conserved) the cells start from those values instead of the
// initial condition.
//
// out.bin: int32 ncells, int32 nsteps;
//          int32 [ncells*4][3] cell-edge records (cell, neighbour or boundary code, edge);
//          f64 [ncells][2] centroids; f64 [ncells] volumes;
//          f64 [ncells][4] fields after nsteps; f64 [nsteps] dts; f64 elapsed seconds.
#include <chrono>
#include <cstdio>
#include <fstream>
#include <functional>
#include <iostream>
#include <memory>

#include <nlohmann/json.hpp>

#include "simulation/parser/parser.h"

using json = nlohmann::json;
using namespace HyperFlow;

int main(int argc, char* argv[])
{
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s case.json nsteps out.bin [fields.bin]\n", argv[0]);
        return 2;
    }
    std::ifstream jf(argv[1]);
    json config;
    jf >> config;
    const int nsteps = std::atoi(argv[2]);

    std::streambuf* saved = std::cout.rdbuf();
    std::ofstream devnull("/dev/null");
    std::cout.rdbuf(devnull.rdbuf());

    JSONParser parser(config);
    std::shared_ptr<Model> model = parser.generate_model();
    std::shared_ptr<Mesh> mesh = parser.generate_mesh();
    std::shared_ptr<InitialCondition> initcon = parser.generate_initial_condition();
    std::shared_ptr<RiemannSolver> riemann = parser.generate_riemann_solver();
    std::shared_ptr<TimeStep> time_step = parser.generate_time_step();
    std::shared_ptr<Scheme> scheme = parser.generate_spatial_scheme(model, riemann, time_step, mesh);
    std::shared_ptr<ODESolver> ode = parser.generate_ode_solver();
    mesh->initialise_field_values(initcon);
    if (argc > 4) {
        std::ifstream ff(argv[4], std::ios::binary);
        Vec2D start(mesh->get_cells().size(), Vec1D(4));
        for (auto& u : start)
            for (int k = 0; k < 4; k++) ff.read(reinterpret_cast<char*>(&u[k]), 8);
        if (!ff) {
            std::cout.rdbuf(saved);
            std::fprintf(stderr, "%s: short fields file\n", argv[4]);
            return 1;
        }
        mesh->set_cells_flow_values(start);
    }

    std::vector<double> dts;
    const auto t0 = std::chrono::steady_clock::now();
    for (int n = 0; n < nsteps; n++) {
        const double dt = (*time_step)(mesh);
        auto evolve = std::bind(&Scheme::operator(), scheme);
        Vec2D upd = (*ode)(evolve, mesh->get_cells_flow_values(), dt);
        mesh->set_cells_flow_values(upd);
        dts.push_back(dt);
    }
    const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::cout.rdbuf(saved);

    std::FILE* f = std::fopen(argv[3], "wb");
    if (!f) return 1;
    const int32_t nc = (int32_t)mesh->get_cells().size(), ns = nsteps;
    std::fwrite(&nc, 4, 1, f);
    std::fwrite(&ns, 4, 1, f);
    for (const auto& r : mesh->get_cell_edge_indices())
        for (int k = 0; k < 3; k++) {
            const int32_t v = r[k];
            std::fwrite(&v, 4, 1, f);
        }
    for (auto& c : mesh->get_cells()) {
        const double xy[2] = {c.get_centroid()[0], c.get_centroid()[1]};
        std::fwrite(xy, 8, 2, f);
    }
    for (auto& c : mesh->get_cells()) {
        const double v = c.get_volume();
        std::fwrite(&v, 8, 1, f);
    }
    for (auto& c : mesh->get_cells()) {
        const Vec1D& u = c.get_flow_values();
        for (int k = 0; k < 4; k++) {
            const double x = u[k];
            std::fwrite(&x, 8, 1, f);
        }
    }
    if (!dts.empty()) std::fwrite(dts.data(), 8, dts.size(), f);
    std::fwrite(&elapsed, 8, 1, f);
    std::fclose(f);
    return 0;
}
