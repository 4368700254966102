// hf_cli.cpp -- `hyperflow <case.json>`: runs a HyperFlow case on the GPU.
//
// Same flow as the reference's main (src/fvm_2d_cart/main.cpp:24-87): parse the
// case, build every object with the JSONParser factories, construct the
// Simulation and solve, writing results_%06d.csv frames.
//
//   hyperflow case.json [--output-dir DIR] [--quiet]
#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>

#include "../../../include/hyperflow.hpp"
#include "hf_json.hpp"

using namespace HyperFlow;

int main(int argc, char* argv[])
{
    std::string path, outdir = ".";
    bool quiet = false;
    for (int i = 1; i < argc; i++) {
        if (!std::strcmp(argv[i], "--output-dir") && i + 1 < argc) outdir = argv[++i];
        else if (!std::strcmp(argv[i], "--quiet")) quiet = true;
        else if (path.empty()) path = argv[i];
        else {
            std::fprintf(stderr, "usage: %s case.json [--output-dir DIR] [--quiet]\n", argv[0]);
            return 2;
        }
    }
    if (path.empty()) {
        std::fprintf(stderr, "usage: %s case.json [--output-dir DIR] [--quiet]\n", argv[0]);
        return 2;
    }
    try {
        JSONParser parser = JSONParser::from_file(path);
        std::shared_ptr<Model> model = parser.generate_model();
        std::shared_ptr<DataOutput> data_output = parser.generate_data_output();
        std::shared_ptr<Mesh> mesh = parser.generate_mesh();
        std::shared_ptr<InitialCondition> initcon = parser.generate_initial_condition();
        std::shared_ptr<RiemannSolver> riemann = parser.generate_riemann_solver();
        std::shared_ptr<TimeStep> time_step = parser.generate_time_step();
        std::shared_ptr<Scheme> scheme = parser.generate_spatial_scheme(model, riemann, time_step, mesh);
        std::shared_ptr<ODESolver> ode_solver = parser.generate_ode_solver();
        if (data_output) data_output->set_directory(outdir);
        Simulation sim(model, data_output, mesh, initcon, ode_solver, scheme, riemann, time_step, parser.get_cfl(),
                       parser.get_t_start(), parser.get_t_end(), parser.get_anim_duration(), parser.get_anim_fps(),
                       parser.get_device_options());
        sim.set_quiet(quiet);
        const unsigned steps = sim.solve();
        if (quiet) std::cout << "steps " << steps << " t " << sim.time() << std::endl;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "hyperflow: %s\n", e.what());
        return 1;
    }
    return 0;
}
