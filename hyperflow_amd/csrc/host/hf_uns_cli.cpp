// hf_uns_cli.cpp -- bin/hyperflow_uns: runs an unstructured case file like
// the reference's src/fvm_2d_uns program (main.cpp:24-88), on the GPU.
//
//   hyperflow_uns case.json [--output-dir DIR] [--quiet]
//
// Reads the .hmsh mesh named by mesh.file (LoadStructuredMesh,
// mesh/mesh_structured/mesh_structured.cpp:23-122; the path is taken as given,
// then relative to the case file), sets the initial condition at the cell
// centroids (parser.cpp:231-291, mesh/mesh/mesh.cpp:60-68), and steps like
// Simulation::solve (simulation/simulation/simulation.cpp:51-119): t_end and
// frame clamps on the TimeStep result, frame 0 before the first update,
// results_%06d.csv frames (simulation/output/euler_data_output/
// euler_data_output.cpp:37-80).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../../include/hyperflow.hpp"
#include "../../../include/hyperflow_uns.h"
#include "hf_json.hpp"

using namespace HyperFlow;

namespace {

struct HmshMesh {
    std::vector<double> vertices;   // [n][4][2]
    std::vector<int32_t> codes;     // [n][4]
};

HmshMesh read_hmsh(const std::string& path)
{
    std::ifstream f(path);
    if (!f) throw std::runtime_error("Unable to open HyperFlow mesh file '" + path + "'");
    std::string line;
    std::getline(f, line);   // header
    std::getline(f, line);
    const unsigned long cells = std::stoul(line, nullptr, 0);
    std::cout << "CELL COUNT: " << cells << std::endl;
    HmshMesh m;
    while (std::getline(f, line)) {
        if (line == "EDGES") break;
        std::istringstream is(line);
        std::string tok;
        int k = 0;
        while (is >> tok && k < 8) {
            m.vertices.push_back(std::stod(tok));
            k++;
        }
        if (k != 8) throw std::runtime_error("mesh file '" + path + "': a cell line needs 8 coordinates");
    }
    std::getline(f, line);
    const unsigned long edges = std::stoul(line, nullptr, 0);
    std::cout << "EDGE COUNT: " << edges << std::endl;
    while (std::getline(f, line)) m.codes.push_back(std::stoi(line, nullptr, 0));
    if (m.codes.size() != m.vertices.size() / 2)
        throw std::runtime_error("mesh file '" + path + "': " + std::to_string(m.codes.size()) + " edge codes for " +
                                 std::to_string(m.vertices.size() / 8) + " cells");
    return m;
}

void check(int rc)
{
    if (rc != HF_OK) throw std::runtime_error(hf_last_error());
}

void write_frame(const std::string& dir, unsigned idx, double t, const hf_json::Value& cfg,
                 const EulerIdealGasModel& model, const std::vector<double>& cen, const std::vector<double>& f, int n)
{
    char name[32];
    std::snprintf(name, sizeof name, "results_%06d.csv", idx);
    std::ofstream out(dir + "/" + name);
    const hf_json::Value& d = cfg["domain"];
    out << "X_LEFT=" << d["x_left"].num() << std::endl;
    out << "X_RIGHT=" << d["x_right"].num() << std::endl;
    out << "Y_BOTTOM=" << d["y_bottom"].num() << std::endl;
    out << "Y_TOP=" << d["y_top"].num() << std::endl;
    out << "T_START=" << cfg["simulation"]["t_start"].num() << std::endl;
    out << "T_END=" << cfg["simulation"]["t_end"].num() << std::endl;
    out << "T=" << t << std::endl << std::endl;
    out << "x,y,rho,rho_u,rho_v,E,u,v,p" << std::endl;
    for (int c = 0; c < n; c++) {
        const Vec1D cons = {f[(size_t)c], f[(size_t)n + c], f[(size_t)2 * n + c], f[(size_t)3 * n + c]};
        const Vec1D prim = model.cons_to_prim(cons);
        out << cen[2 * c] << "," << cen[2 * c + 1] << "," << cons[0] << "," << cons[1] << "," << cons[2] << ","
            << cons[3] << "," << prim[1] << "," << prim[2] << "," << prim[3] << std::endl;
    }
}

Vec1D numbers(const hf_json::Value& v)
{
    Vec1D out;
    for (size_t i = 0; i < v.size(); i++) out.push_back(v[i].num());
    return out;
}

}  // namespace

int main(int argc, char* argv[])
{
    std::string path, outdir = ".";
    bool quiet = false, bad = false;
    for (int i = 1; i < argc; i++) {
        if (!std::strcmp(argv[i], "--output-dir") && i + 1 < argc) outdir = argv[++i];
        else if (!std::strcmp(argv[i], "--quiet")) quiet = true;
        else if (path.empty()) path = argv[i];
        else bad = true;
    }
    if (path.empty() || bad) {
        std::fprintf(stderr, "usage: %s case.json [--output-dir DIR] [--quiet]\n", argv[0]);
        return 2;
    }
    hf_uns* u = nullptr;
    try {
        const hf_json::Value cfg = hf_json::Value::load(path);
        if (cfg["model"]["model"].str() != "EulerIdealGas")
            throw std::runtime_error("Model type '" + cfg["model"]["model"].str() + "' not supported");
        if (cfg["mesh"]["type"].str() != "hyperflow_structured")
            throw std::runtime_error("Mesh type '" + cfg["mesh"]["type"].str() + "' not supported");
        if (cfg["scheme"]["type"].str() != "Godunov")
            throw std::runtime_error("Scheme type '" + cfg["scheme"]["type"].str() + "' not supported");
        if (cfg["scheme"]["riemann"].str() != "HLLCEuler")
            throw std::runtime_error("Riemann solver type '" + cfg["scheme"]["riemann"].str() + "' not supported");
        if (cfg["scheme"]["ode"].str() != "ForwardEuler")
            throw std::runtime_error("ODE Solver type '" + cfg["scheme"]["ode"].str() + "' not supported");
        const double gamma = cfg["model"]["gamma"].num();
        const double cfl = cfg["simulation"]["cfl"].num();
        const double t_start = cfg["simulation"]["t_start"].num(), t_end = cfg["simulation"]["t_end"].num();
        const double duration = cfg["simulation"]["animation"]["duration"].num();
        const double fps = cfg["simulation"]["animation"]["fps"].num();
        EulerIdealGasModel model(gamma);

        std::string mesh_path = cfg["mesh"]["file"].str();
        if (!std::ifstream(mesh_path)) {
            const size_t slash = path.find_last_of('/');
            if (slash != std::string::npos) mesh_path = path.substr(0, slash + 1) + mesh_path;
        }
        const HmshMesh mesh = read_hmsh(mesh_path);
        const int n = (int)(mesh.vertices.size() / 8);

        const Vec1D inlet = model.prim_to_cons(numbers(cfg["mesh"]["boundaries"]["inlet"]));
        hf_uns_desc d;
        std::memset(&d, 0, sizeof d);
        d.ncells = n;
        d.vertices = mesh.vertices.data();
        d.edge_types = mesh.codes.data();
        d.gamma = gamma;
        d.cfl = cfl;
        for (int k = 0; k < 4; k++) d.inlet[k] = inlet.at(k);
        check(hf_uns_create(&d, &u));
        std::cout << "COMPLETED MESH" << std::endl;

        std::vector<double> cen((size_t)2 * n), vol(n), f((size_t)4 * n);
        check(hf_uns_geometry(u, cen.data(), vol.data()));
        const hf_json::Value& ic = cfg["initial_condition"];
        std::unique_ptr<InitialCondition> init;
        if (ic["type"].str() == "Constant") {
            init.reset(new ConstantInitialCondition(model.prim_to_cons(numbers(ic["init_state"]))));
        } else if (ic["type"].str() == "SphericalRiemannProblem") {
            init.reset(new SphericalRiemannProblemInitialCondition(
                ic["radius"].num(), ic["x_origin"].num(), ic["y_origin"].num(),
                model.prim_to_cons(numbers(ic["init_sphere_state"])), model.prim_to_cons(numbers(ic["init_ext_state"]))));
        } else {
            throw std::runtime_error("Initial condition '" + ic["type"].str() + "' not supported");
        }
        for (int c = 0; c < n; c++) {
            const Vec1D v = (*init)(cen[2 * c], cen[2 * c + 1]);
            for (int k = 0; k < 4; k++) f[(size_t)k * n + c] = v[k];
        }
        check(hf_uns_upload(u, f.data(), f.size()));

        // Simulation::solve (simulation.cpp:51-119)
        const auto start = std::chrono::steady_clock::now();
        const double interval = t_end / (duration * fps);
        double t = t_start, dt = 0.0, cur_frame_out = 0.0;
        unsigned frame_idx = 0, n_step = 1;
        auto frame = [&](double tt) {
            check(hf_uns_download(u, f.data(), f.size()));
            write_frame(outdir, frame_idx, tt, cfg, model, cen, f, n);
        };
        const unsigned max_steps = 1000000000u;   // Simulation::MAX_STEPS (simulation.h:50)
        for (; n_step < max_steps; n_step++) {
            if (!quiet) {
                const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
                std::cout << "Steps: " << n_step << ", T: " << t << ", dt: " << dt << ", complete: "
                          << t * 100.0 / t_end << "%, elapsed: " << elapsed << "s" << std::endl;
            }
            if (t == t_end) break;
            check(hf_uns_time_step(u, &dt));
            if (t + dt > t_end) dt = t_end - t;
            if (cur_frame_out == 0.0) {
                frame(t);
                frame_idx++;
            }
            if ((t + dt > cur_frame_out) && (cur_frame_out > 0.0)) dt = cur_frame_out - t;
            check(hf_uns_advance(u, dt));
            if (std::fabs(t - cur_frame_out) < 1e-12) {
                frame(t);
                cur_frame_out += interval;
                frame_idx++;
            }
            t += dt;
        }
        frame(t);
        if (quiet) std::cout << "steps " << (n_step - 1) << " t " << t << std::endl;
        hf_uns_destroy(u);
    } catch (const std::exception& e) {
        if (u) hf_uns_destroy(u);
        std::fprintf(stderr, "hyperflow_uns: %s\n", e.what());
        return 1;
    }
    return 0;
}
