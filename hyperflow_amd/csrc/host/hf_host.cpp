// hf_host.cpp -- the C++ drop-in classes (include/hyperflow.hpp) over the C ABI.
//
// Paths in comments are relative to the reference checkout, under
// src/fvm_2d_cart/ unless stated otherwise.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <stdexcept>

#include "../../../include/hyperflow.hpp"
#include "hf_json.hpp"

namespace HyperFlow {

namespace {

void check(int rc, const char* what)
{
    if (rc != HF_OK) throw std::runtime_error(std::string(what) + ": " + hf_last_error());
}

}  // namespace

// The C-ABI context a Mesh is bound to.
class Device {
public:
    Device(Mesh& mesh, const Scheme& scheme, const Model& model, const TimeStep& ts, const DeviceOptions& opt)
        : mesh(mesh)
    {
        const auto& blocks = mesh.get_mesh_blocks();
        const auto& bcs = mesh.get_block_bcs();
        const unsigned g = scheme.scheme_kind() == HF_MUSCL_HANCOCK ? 2u : 1u;
        std::vector<hf_block_desc> d(blocks.size());
        for (size_t b = 0; b < blocks.size(); b++) {
            const MeshBlock& B = *blocks[b];
            if (B.get_ghost_cells() != g)
                throw std::logic_error("mesh block ghost cells do not match the scheme");
            hf_block_desc& D = d[b];
            D.x_left = B.get_x_left();
            D.x_right = B.get_x_right();
            D.y_bottom = B.get_y_bottom();
            D.y_top = B.get_y_top();
            D.nx = (int32_t)B.get_x_cells();
            D.ny = (int32_t)B.get_y_cells();
            D.problem = 0;
            D.reserved = 0;
            if (bcs[b].size() != 4) throw std::logic_error("every block needs four boundary conditions");
            for (int s = 0; s < 4; s++) {
                if (!bcs[b][s]) throw std::logic_error("null boundary condition (unrecognised type in the case)");
                D.bc[s] = bcs[b][s]->kind();
                D.neighbour[s] = bcs[b][s]->neighbour();
                const Vec1D st = bcs[b][s]->state();
                for (int v = 0; v < 4; v++) D.state[s][v] = st[v];
            }
        }
        hf_case_desc c;
        c.scheme = scheme.scheme_kind();
        c.limiter = scheme.limiter_kind();
        c.precision = opt.precision;
        c.dt_mode = opt.dt_mode;
        c.gamma = model.get_gamma();
        c.cfl = ts.get_cfl();
        c.nblocks = (int32_t)d.size();
        c.tile_rows = opt.tile_rows;
        c.tail_rows = opt.tail_rows;
        c.fix_rows = opt.fix_rows;
        c.blocks = d.data();
        check(hf_create(&c, &ctx), "hf_create");
        for (size_t b = 0; b < blocks.size(); b++) {
            blocks[b]->device = this;
            blocks[b]->index = (int)b;
            upload((int)b, blocks[b]->fields);
        }
    }
    ~Device()
    {
        for (auto& b : mesh.get_mesh_blocks()) {
            b->get_field_array();            // keep the last state on the host
            b->device = nullptr;
        }
        hf_destroy(ctx);
    }

    void upload(int b, const std::vector<double>& f) { check(hf_upload(ctx, b, f.data(), f.size()), "hf_upload"); }
    void download(int b, std::vector<double>& f) { check(hf_download(ctx, b, f.data(), f.size()), "hf_download"); }
    void stale()
    {
        for (auto& b : mesh.get_mesh_blocks()) b->host_current = false;
    }
    void apply_bcs()
    {
        check(hf_apply_bcs(ctx), "hf_apply_bcs");
        stale();
    }
    double time_step()
    {
        double dt = 0.0;
        check(hf_time_step(ctx, 0, &dt), "hf_time_step");
        return dt;
    }
    Vec1D residual(int b, size_t n)
    {
        Vec1D L(n);
        check(hf_scheme_residual(ctx, b, L.data(), n), "hf_scheme_residual");
        return L;
    }
    void advance(double dt)
    {
        check(hf_advance(ctx, &dt), "hf_advance");
        stale();
    }

    hf_ctx* ctx = nullptr;
    Mesh& mesh;
};

// ------------------------------------------------------------------ model --
EulerIdealGasModel::EulerIdealGasModel(double g) : gamma(g) {}

Vec1D EulerIdealGasModel::cons_to_prim(const Vec1D& c) const
{
    const double rho = c[0], u = c[1] / rho, v = c[2] / rho;
    return Vec1D{rho, u, v, (gamma - 1.0) * (c[3] - 0.5 * rho * (u * u + v * v))};
}

Vec1D EulerIdealGasModel::prim_to_cons(const Vec1D& w) const
{
    const double rho = w[0], u = w[1], v = w[2], p = w[3];
    return Vec1D{rho, rho * u, rho * v, p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)};
}

// ------------------------------------------------------ initial conditions --
Vec1D InitialCondition::operator()(double x, double y) const
{
    Vec1D out;
    for (unsigned d = 0; d < 4; d++) out.push_back(field_func_dim(d, x, y));
    return out;
}

ConstantInitialCondition::ConstantInitialCondition(const Vec1D& s) : state(s) {}
double ConstantInitialCondition::field_func_dim(unsigned d, double, double) const { return state[d]; }

DoubleRiemannProblemInitialCondition::DoubleRiemannProblemInitialCondition(double a, double b, const Vec1D& l_,
                                                                           const Vec1D& m_, const Vec1D& r_)
    : x_lm(a), x_mr(b), l(l_), m(m_), r(r_) {}
// doublerp.cpp:42-55
double DoubleRiemannProblemInitialCondition::field_func_dim(unsigned d, double x, double) const
{
    if (x <= x_lm) return l[d];
    if ((x > x_lm) && (x <= x_mr)) return m[d];
    return r[d];
}

QuadRiemannProblemInitialCondition::QuadRiemannProblemInitialCondition(double x, double y, const Vec1D& nw_,
                                                                       const Vec1D& ne_, const Vec1D& sw_,
                                                                       const Vec1D& se_)
    : xi(x), yi(y), nw(nw_), ne(ne_), sw(sw_), se(se_) {}
// quadrp.cpp:49-64
double QuadRiemannProblemInitialCondition::field_func_dim(unsigned d, double x, double y) const
{
    if ((x <= xi) && (y <= yi)) return sw[d];
    if ((x <= xi) && (y > yi)) return nw[d];
    if ((x > xi) && (y <= yi)) return se[d];
    return ne[d];
}

SphericalRiemannProblemInitialCondition::SphericalRiemannProblemInitialCondition(double rad, double x, double y,
                                                                                 const Vec1D& i, const Vec1D& e)
    : radius(rad), x0(x), y0(y), in(i), ext(e) {}
// sphericalrp.cpp:62-72
double SphericalRiemannProblemInitialCondition::field_func_dim(unsigned d, double x, double y) const
{
    if (((x - x0) * (x - x0) + (y - y0) * (y - y0)) <= (radius * radius)) return in[d];
    return ext[d];
}

// -------------------------------------------------------------------- mesh --
MeshBlock::MeshBlock(const Vec1D& ext, unsigned nx, unsigned ny, unsigned g, unsigned dim)
    : extents(ext), x_cells(nx), y_cells(ny), ghost_cells(g), dimension(dim)
{
    if (ext.size() != 4) throw std::invalid_argument("MeshBlock extents need 4 values");
    dx = (get_x_right() - get_x_left()) / x_cells;           // mesh_block.cpp:227-236
    dy = (get_y_top() - get_y_bottom()) / y_cells;
    fields.assign((size_t)4 * get_total_x_cells() * get_total_y_cells(), 0.0);
}

Vec1D MeshBlock::get_centroid(unsigned i, unsigned j) const
{
    const unsigned tx = get_total_x_cells(), ty = get_total_y_cells();
    if (i > ghost_cells - 1 && i < tx - ghost_cells && j > ghost_cells - 1 && j < ty - ghost_cells) {
        const double x0 = get_x_left() - (static_cast<double>(ghost_cells) - 0.5) * dx;
        const double y0 = get_y_bottom() - (static_cast<double>(ghost_cells) - 0.5) * dy;
        return Vec1D{x0 + static_cast<double>(i) * dx, y0 + static_cast<double>(j) * dy};
    }
    return Vec1D{0.0, 0.0};
}

namespace {

// fields[x][y][var] (the reference's Vec3D) from a flat [var][y][x] array
Vec3D to_vec3d(const std::vector<double>& f, unsigned tx, unsigned ty)
{
    Vec3D out(tx, Vec2D(ty, Vec1D(4)));
    for (unsigned i = 0; i < tx; i++)
        for (unsigned j = 0; j < ty; j++)
            for (unsigned v = 0; v < 4; v++) out[i][j][v] = f[((size_t)v * ty + j) * tx + i];
    return out;
}

std::vector<double> from_vec3d(const Vec3D& F, unsigned tx, unsigned ty, const char* what)
{
    if (F.size() != tx) throw std::invalid_argument(std::string(what) + ": expected nx+2g columns");
    std::vector<double> f((size_t)4 * tx * ty);
    for (unsigned i = 0; i < tx; i++) {
        if (F[i].size() != ty) throw std::invalid_argument(std::string(what) + ": expected ny+2g rows");
        for (unsigned j = 0; j < ty; j++) {
            if (F[i][j].size() != 4) throw std::invalid_argument(std::string(what) + ": expected 4 variables");
            for (unsigned v = 0; v < 4; v++) f[((size_t)v * ty + j) * tx + i] = F[i][j][v];
        }
    }
    return f;
}

}  // namespace

const std::vector<double>& MeshBlock::get_field_array() const
{
    if (device && !host_current) {
        device->download(index, fields);
        host_current = true;
        fields3_current = false;
    }
    return fields;
}

void MeshBlock::set_field_array(const std::vector<double>& f)
{
    if (f.size() != fields.size()) throw std::invalid_argument("set_fields: size must be 4*(nx+2g)*(ny+2g)");
    fields = f;
    host_current = true;
    fields3_current = false;
    if (device) device->upload(index, fields);
}

const Vec3D& MeshBlock::get_fields() const
{
    const std::vector<double>& f = get_field_array();
    if (!fields3_current) {
        fields3 = to_vec3d(f, get_total_x_cells(), get_total_y_cells());
        fields3_current = true;
    }
    return fields3;
}

void MeshBlock::set_fields(const Vec3D& F)
{
    set_field_array(from_vec3d(F, get_total_x_cells(), get_total_y_cells(), "set_fields"));
    fields3 = F;
    fields3_current = true;
}

double MeshBlock::get_flow_value(unsigned i, unsigned j, unsigned var) const
{
    const std::vector<double>& f = get_field_array();
    return f[((size_t)var * get_total_y_cells() + j) * get_total_x_cells() + i];
}

void MeshBlock::initialise_field_values(const InitialCondition& ic)
{
    std::vector<double> f = get_field_array();
    const unsigned tx = get_total_x_cells(), ty = get_total_y_cells();
    for (unsigned i = ghost_cells; i < tx - ghost_cells; i++)
        for (unsigned j = ghost_cells; j < ty - ghost_cells; j++) {
            const Vec1D c = get_centroid(i, j);
            const Vec1D v = ic(c[0], c[1]);
            for (unsigned k = 0; k < 4; k++) f[((size_t)k * ty + j) * tx + i] = v[k];
        }
    set_field_array(f);
}

Mesh::Mesh(const std::vector<std::shared_ptr<MeshBlock> >& b,
           const std::vector<std::vector<std::shared_ptr<BoundaryCondition> > >& c)
    : blocks(b), bcs(c) {}

Mesh::~Mesh() { dev.reset(); }

void Mesh::apply_bcs()
{
    if (!dev) throw std::logic_error("Mesh is not bound to a device: construct a Simulation or call Mesh::attach");
    dev->apply_bcs();
}

void Mesh::initialise_field_values(const std::shared_ptr<InitialCondition>& ic)
{
    for (auto& b : blocks) b->initialise_field_values(*ic);   // mesh.cpp:42-47
}

void Mesh::attach(const Scheme& scheme, const Model& model, const TimeStep& ts, const DeviceOptions& opt)
{
    dev.reset();
    dev = std::make_shared<Device>(*this, scheme, model, ts, opt);
}

// ------------------------------------------------------------------ scheme --
Limiter::Limiter(double omega, double bim, double bip, double tol)
{
    if (omega != 0.0 || bim != 1.0 || bip != 1.0 || tol != 1e-6)
        throw std::invalid_argument("limiters support the parser's parameters (0, 1, 1, 1e-6) only");
}

TimeStep::TimeStep(const std::shared_ptr<Model>& m, double c) : model(m), cfl(c) {}

double TimeStep::operator()(const std::shared_ptr<Mesh>& mesh)
{
    if (!mesh->device()) throw std::logic_error("TimeStep: mesh is not bound to a device");
    return mesh->device()->time_step();
}

GodunovScheme::GodunovScheme(const std::shared_ptr<Model>& m, const std::shared_ptr<RiemannSolver>& r,
                             const std::shared_ptr<TimeStep>& t)
    : model(m), riemann(r), time_step(t) {}

Vec3D Scheme::operator()(const std::shared_ptr<MeshBlock>& block)
{
    return to_vec3d(residual(block), block->get_total_x_cells(), block->get_total_y_cells());
}

Vec1D GodunovScheme::residual(const std::shared_ptr<MeshBlock>& block)
{
    if (!block->device) throw std::logic_error("Scheme: block is not bound to a device");
    return block->device->residual(block->index, block->fields.size());
}

double GodunovScheme::calculate_time_step(const std::shared_ptr<Mesh>& mesh) { return (*time_step)(mesh); }

MUSCLHancockScheme::MUSCLHancockScheme(const std::shared_ptr<Model>& m, const std::shared_ptr<RiemannSolver>& r,
                                       const std::shared_ptr<TimeStep>& t, const std::shared_ptr<Limiter>& l,
                                       const std::shared_ptr<Mesh>& msh)
    : model(m), riemann(r), time_step(t), limiter(l), mesh(msh) {}

Vec1D MUSCLHancockScheme::residual(const std::shared_ptr<MeshBlock>& block)
{
    if (!block->device) throw std::logic_error("Scheme: block is not bound to a device");
    return block->device->residual(block->index, block->fields.size());
}

double MUSCLHancockScheme::calculate_time_step(const std::shared_ptr<Mesh>& m) { return (*time_step)(m); }

Vec3D ForwardEulerODESolver::operator()(const std::function<Vec3D()>& func, const Vec3D& U, double dt)
{
    const Vec3D L = func();
    if (L.size() != U.size()) throw std::invalid_argument("ODESolver: update and fields differ in shape");
    Vec3D out(U.size());
    for (size_t i = 0; i < U.size(); i++) {
        if (L[i].size() != U[i].size()) throw std::invalid_argument("ODESolver: update and fields differ in shape");
        out[i].resize(U[i].size());
        for (size_t j = 0; j < U[i].size(); j++) {
            if (L[i][j].size() != U[i][j].size()) throw std::invalid_argument("ODESolver: update and fields differ in shape");
            out[i][j].resize(U[i][j].size());
            for (size_t v = 0; v < U[i][j].size(); v++) out[i][j][v] = (dt * L[i][j][v]) + U[i][j][v];
        }
    }
    return out;
}

void ForwardEulerODESolver::advance(const std::shared_ptr<Mesh>& mesh, double dt)
{
    if (!mesh->device()) throw std::logic_error("ODESolver: mesh is not bound to a device");
    mesh->device()->advance(dt);
}

// ------------------------------------------------------------------ output --
DataOutput::DataOutput(const std::shared_ptr<Model>& m, double xl, double xr, double yb, double yt, double ts,
                       double te)
    : model(m), x_left(xl), x_right(xr), y_bottom(yb), y_top(yt), t_start(ts), t_end(te) {}

void EulerIdealGasDataOutput::to_file(const std::shared_ptr<Mesh>& mesh, unsigned n, double t)
{
    char name[32];
    std::snprintf(name, sizeof name, "results_%06d.csv", n);
    std::ofstream out(directory + "/" + name);
    out << "X_LEFT=" << x_left << std::endl;
    out << "X_RIGHT=" << x_right << std::endl;
    out << "Y_BOTTOM=" << y_bottom << std::endl;
    out << "Y_TOP=" << y_top << std::endl;
    out << "T_START=" << t_start << std::endl;
    out << "T_END=" << t_end << std::endl;
    out << "T=" << t << std::endl << std::endl;
    out << "x,y,rho,rho_u,rho_v,E,u,v,p" << std::endl;
    for (auto& blk : mesh->get_mesh_blocks()) {
        const MeshBlock& B = *blk;
        const unsigned g = B.get_ghost_cells(), tx = B.get_total_x_cells(), ty = B.get_total_y_cells();
        const std::vector<double>& f = B.get_field_array();
        for (unsigned i = g; i < tx - g; i++)
            for (unsigned j = g; j < ty - g; j++) {
                const Vec1D c = B.get_centroid(i, j);
                Vec1D cons(4);
                for (unsigned k = 0; k < 4; k++) cons[k] = f[((size_t)k * ty + j) * tx + i];
                const Vec1D prim = model->cons_to_prim(cons);
                out << c[0] << "," << c[1] << "," << cons[0] << "," << cons[1] << "," << cons[2] << "," << cons[3]
                    << "," << prim[1] << "," << prim[2] << "," << prim[3] << std::endl;
            }
    }
}

// -------------------------------------------------------------- simulation --
Simulation::Simulation(const std::shared_ptr<Model>& m, const std::shared_ptr<DataOutput>& out,
                       const std::shared_ptr<Mesh>& msh, const std::shared_ptr<InitialCondition>& ic,
                       const std::shared_ptr<ODESolver>& ode, const std::shared_ptr<Scheme>& sch,
                       const std::shared_ptr<RiemannSolver>&, const std::shared_ptr<TimeStep>& ts, double c,
                       double t0, double t1, double dur, double fps, const DeviceOptions& opt)
    : model(m), data_output(out), mesh(msh), ode_solver(ode), scheme(sch), time_step(ts), cfl(c), t_start(t0),
      t_end(t1), anim_duration(dur), anim_fps(fps)
{
    if (!model || !mesh || !ode_solver || !scheme || !time_step)
        throw std::invalid_argument("Simulation: a required component is null");
    if (ic) mesh->initialise_field_values(ic);              // simulation.cpp:43
    mesh->attach(*scheme, *model, *time_step, opt);
}

unsigned Simulation::solve()
{
    double t = t_start;
    double dt = 0.0;
    const auto start = std::chrono::steady_clock::now();
    const double interval = t_end / (anim_duration * anim_fps);
    double cur_frame_out = 0.0;
    unsigned frame_idx = 0;
    unsigned n = 1;
    for (; n < MAX_STEPS; n++) {
        if (!quiet) {
            const double elapsed =
                std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
            std::cout << "Steps: " << n << ", T: " << t << ", dt: " << dt << ", complete: " << t * 100.0 / t_end
                      << "%, elapsed: " << elapsed << "s" << std::endl;
        }
        if (t == t_end) break;
        mesh->apply_bcs();
        dt = scheme->calculate_time_step(mesh);
        if (t + dt > t_end) dt = t_end - t;
        if (cur_frame_out == 0.0) {
            if (data_output) data_output->to_file(mesh, frame_idx, t);
            frame_idx++;
        }
        if ((t + dt > cur_frame_out) && (cur_frame_out > 0.0)) dt = cur_frame_out - t;
        ode_solver->advance(mesh, dt);
        if (std::fabs(t - cur_frame_out) < 1e-12) {
            if (data_output) data_output->to_file(mesh, frame_idx, t);
            cur_frame_out += interval;
            frame_idx++;
        }
        t += dt;
        t_now = t;
    }
    if (data_output) data_output->to_file(mesh, frame_idx, t);
    t_now = t;
    return n - 1;
}

// ------------------------------------------------------------------ parser --
JSONParser::JSONParser(const hf_json::Value& cfg) : config(std::make_shared<hf_json::Value>(cfg))
{
    const hf_json::Value& c = *config;
    t_start = c["simulation"]["t_start"].num();              // parser.cpp:13-43
    t_end = c["simulation"]["t_end"].num();
    anim_duration = c["simulation"]["animation"]["duration"].num();
    anim_fps = c["simulation"]["animation"]["fps"].num();
    x_left = c["domain"]["x_left"].num();
    x_right = c["domain"]["x_right"].num();
    y_bottom = c["domain"]["y_bottom"].num();
    y_top = c["domain"]["y_top"].num();
    cfl = c["simulation"]["cfl"].num();
    if (c.has("hyperflow_amd")) {
        const hf_json::Value& o = c["hyperflow_amd"];
        if (o.has("precision")) options.precision = o["precision"].str() == "fp32" ? HF_FP32 : HF_FP64;
        if (o.has("dt_mode")) options.dt_mode = o["dt_mode"].str() == "uniform" ? HF_DT_UNIFORM : HF_DT_COMPAT;
        if (o.has("tile_rows")) options.tile_rows = o["tile_rows"].integer();
        if (o.has("tail_rows")) options.tail_rows = o["tail_rows"].integer();
        if (o.has("fix_rows")) options.fix_rows = o["fix_rows"].integer();
    }
}

JSONParser JSONParser::from_file(const std::string& path) { return JSONParser(hf_json::Value::load(path)); }

JSONParser::~JSONParser() {}

std::shared_ptr<Model> JSONParser::generate_model()
{
    const std::string t = (*config)["model"]["model"].str();
    if (t == "EulerIdealGas") {
        model = std::make_shared<EulerIdealGasModel>((*config)["model"]["gamma"].num());
    } else {
        std::cout << "Model type '" << t << "' not supported. Returning null model." << std::endl;
    }
    return model;
}

std::shared_ptr<DataOutput> JSONParser::generate_data_output()
{
    const std::string t = (*config)["model"]["model"].str();
    if (t == "EulerIdealGas")
        return std::make_shared<EulerIdealGasDataOutput>(model, x_left, x_right, y_bottom, y_top, t_start, t_end);
    std::cout << "Model type '" << t << "' not supported. Returning null data output." << std::endl;
    return nullptr;
}

std::shared_ptr<MeshBlock> JSONParser::generate_mesh_block(const hf_json::Value& sub)
{
    const hf_json::Value& e = sub["extents"];
    Vec1D extents{e["x_left"].num(), e["x_right"].num(), e["y_bottom"].num(), e["y_top"].num()};
    unsigned ghost = 1;                                      // parser.cpp:274-288
    const std::string scheme = (*config)["scheme"]["type"].str();
    if (scheme == "Godunov") ghost = 1;
    else if (scheme == "MUSCLHancock") ghost = 2;
    else
        std::cout << "Scheme type '" << scheme << "' not supported. Number of ghost cells (" << ghost
                  << ") may be incorrect." << std::endl;
    return std::make_shared<MeshBlock>(extents, (unsigned)sub["x_cells"].integer(), (unsigned)sub["y_cells"].integer(),
                                       ghost, 4);
}

namespace {
// The reference dereferences the model returned by generate_model even when
// it is null; here that is an error with a message instead of a crash.
const Model& need(const std::shared_ptr<Model>& m)
{
    if (!m) throw std::runtime_error("no model: the case's model type is not supported");
    return *m;
}
}  // namespace

std::shared_ptr<BoundaryCondition> JSONParser::generate_boundary_condition(const hf_json::Value& bc)
{
    const std::string t = bc["type"].str();                 // parser.cpp:432-466
    if (t == "Dirichlet") {
        Vec1D st = bc["flow_state"].numbers();
        if (bc["variables"].str() == "Primitive") st = need(model).prim_to_cons(st);
        return std::make_shared<DirichletBoundaryCondition>(st);
    }
    if (t == "Reflective") return std::make_shared<ReflectiveBoundaryCondition>();
    if (t == "Transmissive") return std::make_shared<TransmissiveBoundaryCondition>();
    if (t == "Internal") return std::make_shared<InternalBoundaryCondition>(bc["block"].integer());
    std::cout << "Boundary condition '" << t << "' not recognised." << std::endl;
    return nullptr;
}

std::shared_ptr<Mesh> JSONParser::generate_mesh()
{
    const hf_json::Value& blocks = (*config)["mesh"]["blocks"];
    std::vector<std::shared_ptr<MeshBlock> > mbs;
    for (size_t b = 0; b < blocks.size(); b++) mbs.push_back(generate_mesh_block(blocks[b]));
    std::vector<std::vector<std::shared_ptr<BoundaryCondition> > > bcs;
    for (size_t b = 0; b < blocks.size(); b++) {
        const hf_json::Value& bd = blocks[b]["boundaries"];
        bcs.push_back({generate_boundary_condition(bd["bottom"]), generate_boundary_condition(bd["right"]),
                       generate_boundary_condition(bd["top"]), generate_boundary_condition(bd["left"])});
    }
    return std::make_shared<Mesh>(mbs, bcs);
}

std::shared_ptr<InitialCondition> JSONParser::generate_initial_condition()
{
    const hf_json::Value& ic = (*config)["initial_condition"];
    const std::string t = ic["type"].str();                 // parser.cpp:166-187
    const bool prim = ic["variables"].str() == "Primitive";
    auto conv = [&](const char* key, bool force) {
        Vec1D s = ic[key].numbers();
        return (prim || force) ? need(model).prim_to_cons(s) : s;
    };
    if (t == "Constant") return std::make_shared<ConstantInitialCondition>(conv("init_state", true));
    if (t == "DoubleRiemannProblem")
        return std::make_shared<DoubleRiemannProblemInitialCondition>(
            ic["x_lm_interface"].num(), ic["x_mr_interface"].num(), conv("init_left_state", false),
            conv("init_middle_state", false), conv("init_right_state", false));
    if (t == "QuadRiemannProblem")
        return std::make_shared<QuadRiemannProblemInitialCondition>(
            ic["x_interface"].num(), ic["y_interface"].num(), conv("init_nw_state", false),
            conv("init_ne_state", false), conv("init_sw_state", false), conv("init_se_state", false));
    if (t == "SphericalRiemannProblem")
        return std::make_shared<SphericalRiemannProblemInitialCondition>(
            ic["radius"].num(), ic["x_origin"].num(), ic["y_origin"].num(), conv("init_sphere_state", true),
            conv("init_ext_state", true));
    std::cout << "Initial condition '" << t << "' not supported. Returning null initial condition." << std::endl;
    return nullptr;
}

std::shared_ptr<RiemannSolver> JSONParser::generate_riemann_solver()
{
    const std::string m = (*config)["model"]["model"].str();
    const std::string r = (*config)["scheme"]["riemann"].str();
    if (m == "EulerIdealGas" && r == "HLLCEuler")
        return std::make_shared<HLLCEulerRiemannSolver>(std::static_pointer_cast<EulerIdealGasModel>(model));
    std::cout << "Riemann solver type '" << r << "' not supported for model '" << m
              << "'. Returning null Riemann solver." << std::endl;
    return nullptr;
}

std::shared_ptr<TimeStep> JSONParser::generate_time_step()
{
    need(model);
    return std::make_shared<TimeStep>(model, cfl);
}

std::shared_ptr<Scheme> JSONParser::generate_spatial_scheme(const std::shared_ptr<Model>& m,
                                                            const std::shared_ptr<RiemannSolver>& r,
                                                            const std::shared_ptr<TimeStep>& ts,
                                                            const std::shared_ptr<Mesh>& mesh)
{
    const std::string t = (*config)["scheme"]["type"].str();  // parser.cpp:215-245
    if (t == "Godunov") return std::make_shared<GodunovScheme>(m, r, ts);
    if (t == "MUSCLHancock") {
        const std::string l = (*config)["scheme"]["limiter"].str();
        std::shared_ptr<Limiter> lim;
        if (l == "MinBee") lim = std::make_shared<MinBeeLimiter>(0.0, 1.0, 1.0, 1e-6);
        else if (l == "VanLeer") lim = std::make_shared<VanLeerLimiter>(0.0, 1.0, 1.0, 1e-6);
        else {
            std::cout << "Limiter type '" << l << "' not supported. Exiting." << std::endl;
            return nullptr;
        }
        return std::make_shared<MUSCLHancockScheme>(m, r, ts, lim, mesh);
    }
    std::cout << "Scheme type '" << t << "' not supported. Returning null scheme." << std::endl;
    return nullptr;
}

std::shared_ptr<ODESolver> JSONParser::generate_ode_solver()
{
    const std::string t = (*config)["scheme"]["ode"].str();
    if (t == "ForwardEuler") return std::make_shared<ForwardEulerODESolver>();
    std::cout << "ODE Solver type '" << t << "' not supported. Returning null ODE solver." << std::endl;
    return nullptr;
}

}  // namespace HyperFlow
