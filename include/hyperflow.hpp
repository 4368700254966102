// hyperflow.hpp -- C++ drop-in surface of the MI355X block-Cartesian Euler step.
//
// The classes keep the names, constructor arguments, factory keys and call
// sequence of HyperFlow's src/fvm_2d_cart so that a main program written
// against the reference (main.cpp:24-87) compiles against this header with
// the include paths changed.  The numerics run on the GPU through the C ABI
// of hyperflow.h; these classes hold descriptions and host copies only.
//
// Differences a user of the reference sees:
//  * Vec1D is std::vector<double> and Vec3D nested std::vectors, with the
//    reference's [x][y][var] indexing; MeshBlock::get_fields/set_fields,
//    Scheme::operator() and ODESolver::operator() take and return Vec3D as
//    the reference's do (slow adapters: every call copies the block between
//    host and device).  The fast paths exchange flat [var][y][x] arrays
//    (x fastest, the C ABI's layout): MeshBlock::get_field_array /
//    set_field_array, Scheme::residual and ODESolver::advance;
//  * a Mesh is bound to a device context when a Simulation is constructed
//    (or by Mesh::attach); Mesh::apply_bcs, TimeStep and Scheme calls before
//    that throw std::logic_error;
//  * factories print the reference's message and return nullptr for unknown
//    type strings, as parser.cpp does; malformed JSON throws hf_json::Error
//    where nlohmann::json would throw.
#pragma once

#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "hyperflow.h"

namespace hf_json {
class Value;
}

namespace HyperFlow {

typedef std::vector<double> Vec1D;
typedef std::vector<Vec1D> Vec2D;
typedef std::vector<Vec2D> Vec3D;   // fields[x][y][var] (src/share/tensor/tensor.h:114)

class Device;

// ------------------------------------------------------------------ model --
// model/model/model.h:21-37, model/euler_ideal_gas/euler_ideal_gas.h:13-73
class Model {
public:
    virtual ~Model() {}
    virtual Vec1D cons_to_prim(const Vec1D& cons) const = 0;
    virtual Vec1D prim_to_cons(const Vec1D& prim) const = 0;
    virtual double get_gamma() const = 0;
};

class EulerIdealGasModel : public Model {
public:
    explicit EulerIdealGasModel(double gamma);
    double get_gamma() const override { return gamma; }
    Vec1D cons_to_prim(const Vec1D& cons) const override;   // euler_ideal_gas.cpp:76-92
    Vec1D prim_to_cons(const Vec1D& prim) const override;   // euler_ideal_gas.cpp:96-109

private:
    double gamma;
};

// ------------------------------------------------------ initial conditions --
// simulation/initcon/*: the value of each conserved variable at a centroid.
class InitialCondition {
public:
    virtual ~InitialCondition() {}
    Vec1D operator()(double x, double y) const;             // initcon.cpp:32-41
    virtual double field_func_dim(unsigned dim, double x, double y) const = 0;
};

class ConstantInitialCondition : public InitialCondition {
public:
    explicit ConstantInitialCondition(const Vec1D& cons_state);
    double field_func_dim(unsigned dim, double x, double y) const override;

private:
    Vec1D state;
};

class DoubleRiemannProblemInitialCondition : public InitialCondition {
public:
    DoubleRiemannProblemInitialCondition(double x_lm_interface, double x_mr_interface, const Vec1D& left,
                                         const Vec1D& middle, const Vec1D& right);
    double field_func_dim(unsigned dim, double x, double y) const override;

private:
    double x_lm, x_mr;
    Vec1D l, m, r;
};

class QuadRiemannProblemInitialCondition : public InitialCondition {
public:
    QuadRiemannProblemInitialCondition(double x_interface, double y_interface, const Vec1D& nw, const Vec1D& ne,
                                       const Vec1D& sw, const Vec1D& se);
    double field_func_dim(unsigned dim, double x, double y) const override;

private:
    double xi, yi;
    Vec1D nw, ne, sw, se;
};

class SphericalRiemannProblemInitialCondition : public InitialCondition {
public:
    SphericalRiemannProblemInitialCondition(double radius, double x_origin, double y_origin, const Vec1D& inside,
                                            const Vec1D& outside);
    double field_func_dim(unsigned dim, double x, double y) const override;

private:
    double radius, x0, y0;
    Vec1D in, ext;
};

// -------------------------------------------------------------------- mesh --
// mesh/mesh_block/mesh_block.h:9-164
class MeshBlock {
public:
    MeshBlock(const Vec1D& extents, unsigned x_cells, unsigned y_cells, unsigned ghost_cells,
              unsigned dimension = 4);

    unsigned get_x_cells() const { return x_cells; }
    unsigned get_y_cells() const { return y_cells; }
    unsigned get_ghost_cells() const { return ghost_cells; }
    unsigned get_total_x_cells() const { return x_cells + 2 * ghost_cells; }
    unsigned get_total_y_cells() const { return y_cells + 2 * ghost_cells; }
    unsigned get_dimension() const { return dimension; }
    double get_dx() const { return dx; }
    double get_dy() const { return dy; }
    double get_x_left() const { return extents[0]; }
    double get_x_right() const { return extents[1]; }
    double get_y_bottom() const { return extents[2]; }
    double get_y_top() const { return extents[3]; }
    Vec1D get_centroid(unsigned i, unsigned j) const;        // mesh_block.cpp:239-257

    // MeshBlock::get_fields / set_fields (mesh_block.cpp:155-164): the padded
    // conserved fields as fields[x][y][var]; reads / writes the device state
    // when bound.
    const Vec3D& get_fields() const;
    void set_fields(const Vec3D& fields);
    // The same fields as one flat array [var][y][x] (x fastest).
    const std::vector<double>& get_field_array() const;
    void set_field_array(const std::vector<double>& fields);
    // fields[x][y][var] of the reference
    double get_flow_value(unsigned i, unsigned j, unsigned var) const;
    // interior cells from an initial condition (mesh_block.cpp:215-224)
    void initialise_field_values(const InitialCondition& ic);

private:
    friend class Mesh;
    friend class Device;
    friend class GodunovScheme;
    friend class MUSCLHancockScheme;
    Vec1D extents;
    unsigned x_cells, y_cells, ghost_cells, dimension;
    double dx, dy;
    mutable std::vector<double> fields;
    mutable Vec3D fields3;                // the Vec3D view, rebuilt when stale
    mutable bool host_current = true;
    mutable bool fields3_current = false;
    Device* device = nullptr;
    int index = -1;
};

// scheme/boundary/*: descriptions consumed by the device ghost fill.
class BoundaryCondition {
public:
    virtual ~BoundaryCondition() {}
    virtual int kind() const = 0;                           // HF_TRANSMISSIVE ..
    virtual int neighbour() const { return -1; }
    virtual Vec1D state() const { return Vec1D(4, 0.0); }
};

class TransmissiveBoundaryCondition : public BoundaryCondition {
public:
    int kind() const override { return HF_TRANSMISSIVE; }
};

class ReflectiveBoundaryCondition : public BoundaryCondition {
public:
    int kind() const override { return HF_REFLECTIVE; }
};

class DirichletBoundaryCondition : public BoundaryCondition {
public:
    explicit DirichletBoundaryCondition(const Vec1D& cons_state) : cons(cons_state) {}
    int kind() const override { return HF_DIRICHLET; }
    Vec1D state() const override { return cons; }

private:
    Vec1D cons;
};

class InternalBoundaryCondition : public BoundaryCondition {
public:
    explicit InternalBoundaryCondition(int external_block) : ext(external_block) {}
    int kind() const override { return HF_INTERNAL; }
    int neighbour() const override { return ext; }

private:
    int ext;
};

class Scheme;
class TimeStep;

struct DeviceOptions {
    int precision = HF_FP64;
    int dt_mode = HF_DT_COMPAT;
    int tile_rows = 0;   // work plan of the update launches (hf_case_desc); 0 = library default
    int tail_rows = 0;
    int fix_rows = 0;
};

// mesh/mesh/mesh.h:11-44
class Mesh {
public:
    Mesh(const std::vector<std::shared_ptr<MeshBlock> >& blocks,
         const std::vector<std::vector<std::shared_ptr<BoundaryCondition> > >& block_bcs);
    ~Mesh();
    std::vector<std::shared_ptr<MeshBlock> >& get_mesh_blocks() { return blocks; }
    const std::vector<std::vector<std::shared_ptr<BoundaryCondition> > >& get_block_bcs() const { return bcs; }
    void apply_bcs();                                        // mesh.cpp:31-39
    void initialise_field_values(const std::shared_ptr<InitialCondition>& ic);
    // Binds the mesh to a device context for the given scheme and model.
    void attach(const Scheme& scheme, const Model& model, const TimeStep& time_step,
                const DeviceOptions& options = DeviceOptions());
    Device* device() const { return dev.get(); }

private:
    std::vector<std::shared_ptr<MeshBlock> > blocks;
    std::vector<std::vector<std::shared_ptr<BoundaryCondition> > > bcs;
    std::shared_ptr<Device> dev;
};

// ------------------------------------------------------------------ scheme --
class RiemannSolver {                                       // scheme/riemann/riemann/riemann.h:14-27
public:
    virtual ~RiemannSolver() {}
};

class HLLCEulerRiemannSolver : public RiemannSolver {      // riemann_hllc.cpp:13-18
public:
    explicit HLLCEulerRiemannSolver(const std::shared_ptr<EulerIdealGasModel>& model) : model(model) {}

private:
    std::shared_ptr<EulerIdealGasModel> model;
};

class Limiter {                                             // scheme/limiter/limiter/limiter.h
public:
    Limiter(double omega, double beta_im_half, double beta_ip_half, double tol);
    virtual ~Limiter() {}
    virtual int kind() const = 0;                            // HF_MINBEE | HF_VANLEER
};

class MinBeeLimiter : public Limiter {
public:
    MinBeeLimiter(double omega, double beta_im_half, double beta_ip_half, double tol)
        : Limiter(omega, beta_im_half, beta_ip_half, tol) {}
    int kind() const override { return HF_MINBEE; }
};

class VanLeerLimiter : public Limiter {
public:
    VanLeerLimiter(double omega, double beta_im_half, double beta_ip_half, double tol)
        : Limiter(omega, beta_im_half, beta_ip_half, tol) {}
    int kind() const override { return HF_VANLEER; }
};

// scheme/timestep/timestep.h:13-56 -- stateful: every call is counted.
class TimeStep {
public:
    TimeStep(const std::shared_ptr<Model>& model, double cfl);
    virtual ~TimeStep() {}
    virtual double operator()(const std::shared_ptr<Mesh>& mesh);
    double get_cfl() const { return cfl; }

private:
    std::shared_ptr<Model> model;
    double cfl;
};

// scheme/scheme/scheme.h:9-24
class Scheme {
public:
    virtual ~Scheme() {}
    // Scheme::operator() (scheme.h:20): the spatial update L of a padded
    // block, [x][y][var], zero in the ghosts
    virtual Vec3D operator()(const std::shared_ptr<MeshBlock>& block);
    // the same as a flat [var][y][x] array
    virtual Vec1D residual(const std::shared_ptr<MeshBlock>& block) = 0;
    virtual double calculate_time_step(const std::shared_ptr<Mesh>& mesh) = 0;
    virtual int scheme_kind() const = 0;                     // HF_GODUNOV | HF_MUSCL_HANCOCK
    virtual int limiter_kind() const { return HF_MINBEE; }
};

class GodunovScheme : public Scheme {                       // scheme/godunov/godunov.h
public:
    GodunovScheme(const std::shared_ptr<Model>& model, const std::shared_ptr<RiemannSolver>& riemann,
                  const std::shared_ptr<TimeStep>& time_step);
    Vec1D residual(const std::shared_ptr<MeshBlock>& block) override;
    double calculate_time_step(const std::shared_ptr<Mesh>& mesh) override;
    int scheme_kind() const override { return HF_GODUNOV; }
    static const unsigned ghost_cells = 1;                   // godunov.cpp:42

private:
    std::shared_ptr<Model> model;
    std::shared_ptr<RiemannSolver> riemann;
    std::shared_ptr<TimeStep> time_step;
};

class MUSCLHancockScheme : public Scheme {                  // scheme/muscl_hancock/muscl_hancock.h
public:
    MUSCLHancockScheme(const std::shared_ptr<Model>& model, const std::shared_ptr<RiemannSolver>& riemann,
                       const std::shared_ptr<TimeStep>& time_step, const std::shared_ptr<Limiter>& limiter,
                       const std::shared_ptr<Mesh>& mesh);
    Vec1D residual(const std::shared_ptr<MeshBlock>& block) override;
    double calculate_time_step(const std::shared_ptr<Mesh>& mesh) override;
    int scheme_kind() const override { return HF_MUSCL_HANCOCK; }
    int limiter_kind() const override { return limiter->kind(); }
    static const unsigned ghost_cells = 2;                   // muscl_hancock.cpp:64

private:
    std::shared_ptr<Model> model;
    std::shared_ptr<RiemannSolver> riemann;
    std::shared_ptr<TimeStep> time_step;
    std::shared_ptr<Limiter> limiter;
    std::shared_ptr<Mesh> mesh;
};

// scheme/ode/*
class ODESolver {
public:
    virtual ~ODESolver() {}
    // ODESolver::operator() (ode.h:22-24): one step of a block's fields from
    // the spatial update func() returns
    virtual Vec3D operator()(const std::function<Vec3D()>& func, const Vec3D& val_to_step, double dt) = 0;
    // the per-block loop of Simulation::solve (simulation.cpp:106-110) on the
    // device, every block of the mesh in order
    virtual void advance(const std::shared_ptr<Mesh>& mesh, double dt) = 0;
};

class ForwardEulerODESolver : public ODESolver {            // forward_euler.cpp:15-33
public:
    // (dt * L) + U element by element (forward_euler.cpp:27-28)
    Vec3D operator()(const std::function<Vec3D()>& func, const Vec3D& val_to_step, double dt) override;
    void advance(const std::shared_ptr<Mesh>& mesh, double dt) override;
};

// ------------------------------------------------------------------ output --
class DataOutput {                                          // simulation/output/data_output
public:
    DataOutput(const std::shared_ptr<Model>& model, double x_left, double x_right, double y_bottom, double y_top,
               double t_start, double t_end);
    virtual ~DataOutput() {}
    virtual void to_file(const std::shared_ptr<Mesh>& mesh, unsigned n, double t) = 0;
    void set_directory(const std::string& dir) { directory = dir; }

protected:
    std::shared_ptr<Model> model;
    double x_left, x_right, y_bottom, y_top, t_start, t_end;
    std::string directory = ".";
};

// results_%06d.csv: 7 metadata lines, a blank line, the header
// x,y,rho,rho_u,rho_v,E,u,v,p and one line per interior cell
// (euler_data_output.cpp:37-84), default 6-significant-digit formatting.
class EulerIdealGasDataOutput : public DataOutput {
public:
    using DataOutput::DataOutput;
    void to_file(const std::shared_ptr<Mesh>& mesh, unsigned n, double t) override;
};

// -------------------------------------------------------------- simulation --
class Simulation {                                          // simulation/simulation/simulation.h
public:
    Simulation(const std::shared_ptr<Model>& model, const std::shared_ptr<DataOutput>& data_output,
               const std::shared_ptr<Mesh>& mesh, const std::shared_ptr<InitialCondition>& initial_condition,
               const std::shared_ptr<ODESolver>& ode_solver, const std::shared_ptr<Scheme>& scheme,
               const std::shared_ptr<RiemannSolver>& riemann, const std::shared_ptr<TimeStep>& time_step,
               double cfl, double t_start, double t_end, double anim_duration, double anim_fps,
               const DeviceOptions& options = DeviceOptions());
    // simulation.cpp:51-124: apply_bcs, counted dt, t_end and frame clamps,
    // per-block forward Euler, output at frame times; returns the steps taken
    unsigned solve();
    void set_quiet(bool q) { quiet = q; }
    double time() const { return t_now; }

private:
    std::shared_ptr<Model> model;
    std::shared_ptr<DataOutput> data_output;
    std::shared_ptr<Mesh> mesh;
    std::shared_ptr<ODESolver> ode_solver;
    std::shared_ptr<Scheme> scheme;
    std::shared_ptr<TimeStep> time_step;
    double cfl, t_start, t_end, anim_duration, anim_fps;
    double t_now = 0.0;
    bool quiet = false;
    static const unsigned MAX_STEPS = 1000000000;          // simulation.h:50
};

// ------------------------------------------------------------------ parser --
// simulation/parser/parser.h: the string-keyed factories of the JSON case.
// Build-only options live under an optional "hyperflow_amd" object:
// {"precision": "fp64"|"fp32", "dt_mode": "compat"|"uniform", "tile_rows": n, "tail_rows": n,
//  "fix_rows": n}.
class JSONParser {
public:
    explicit JSONParser(const hf_json::Value& config);
    static JSONParser from_file(const std::string& path);
    ~JSONParser();

    double get_t_start() const { return t_start; }
    double get_t_end() const { return t_end; }
    double get_anim_duration() const { return anim_duration; }
    double get_anim_fps() const { return anim_fps; }
    double get_x_left() const { return x_left; }
    double get_x_right() const { return x_right; }
    double get_y_bottom() const { return y_bottom; }
    double get_y_top() const { return y_top; }
    double get_cfl() const { return cfl; }
    DeviceOptions get_device_options() const { return options; }

    std::shared_ptr<Model> generate_model();
    std::shared_ptr<DataOutput> generate_data_output();
    std::shared_ptr<Mesh> generate_mesh();
    std::shared_ptr<InitialCondition> generate_initial_condition();
    std::shared_ptr<RiemannSolver> generate_riemann_solver();
    std::shared_ptr<TimeStep> generate_time_step();
    std::shared_ptr<Scheme> generate_spatial_scheme(const std::shared_ptr<Model>& model,
                                                    const std::shared_ptr<RiemannSolver>& riemann,
                                                    const std::shared_ptr<TimeStep>& time_step,
                                                    const std::shared_ptr<Mesh>& mesh);
    std::shared_ptr<ODESolver> generate_ode_solver();

private:
    std::shared_ptr<hf_json::Value> config;
    std::shared_ptr<Model> model;
    double t_start, t_end, anim_duration, anim_fps, x_left, x_right, y_bottom, y_top, cfl;
    DeviceOptions options;
    std::shared_ptr<MeshBlock> generate_mesh_block(const hf_json::Value& sub);
    std::shared_ptr<BoundaryCondition> generate_boundary_condition(const hf_json::Value& bc);
};

}  // namespace HyperFlow
