/*
 * hf_oracle.h -- CPU oracle for the block-Cartesian Euler step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a clean-room FP64 restatement of
 * HyperFlow's src/fvm_2d_cart time step, used as the checker for the HIP
 * path.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it; the product library never links or calls it.
 *
 * Layout: every block is stored as four planes [var][y][x] of the padded
 * block (x fastest), var = (rho, rho*u, rho*v, E).  The reference stores
 * fields[x][y][var] (src/fvm_2d_cart/mesh/mesh_block/mesh_block.cpp:35); the
 * arithmetic is identical, only the addressing differs.
 */
#ifndef HF_ORACLE_H
#define HF_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { HFO_TRANSMISSIVE = 0, HFO_REFLECTIVE = 1, HFO_DIRICHLET = 2, HFO_INTERNAL = 3 };
enum { HFO_GODUNOV = 0, HFO_MUSCL = 1 };
enum { HFO_MINBEE = 0, HFO_VANLEER = 1 };

typedef struct hfo_mesh hfo_mesh;

/* Block descriptors.  Per block b:
 *   idesc[10*b + 0..1] = nx, ny (interior cells)
 *   idesc[10*b + 2..5] = bc kind for sides (bottom, right, top, left)
 *   idesc[10*b + 6..9] = neighbour block index for Internal sides (else -1)
 *   ddesc[20*b + 0..3] = x_left, x_right, y_bottom, y_top
 *   ddesc[20*b + 4 + 4*s + k] = conserved Dirichlet state of side s
 * Side order follows Mesh::apply_bcs (src/fvm_2d_cart/mesh/mesh/mesh.cpp:33-38). */
hfo_mesh* hfo_create(int nblocks, const int* idesc, const double* ddesc,
                     int scheme, int limiter, double gamma, double cfl);
void hfo_destroy(hfo_mesh* m);
int hfo_ghost(const hfo_mesh* m);
void hfo_set_fields(hfo_mesh* m, int b, const double* U);
void hfo_get_fields(const hfo_mesh* m, int b, double* U);
void hfo_set_threads(int n);

/* Mesh::apply_bcs (mesh.cpp:31-39) with the four BC classes. */
void hfo_apply_bcs(hfo_mesh* m);
/* TimeStep::operator() (timestep.cpp:29-68): stateful call counter. */
double hfo_time_step(hfo_mesh* m);
unsigned hfo_calls(const hfo_mesh* m);
/* Scheme::operator()(block): padded spatial update L, zero in ghosts
 * (godunov.cpp:27-33, muscl_hancock.cpp:41-55).  MUSCL calls the TimeStep. */
void hfo_scheme_residual(hfo_mesh* m, int b, double* L);
/* One Simulation::solve step body without clamps or output
 * (simulation.cpp:85-110): apply_bcs, dt, sequential per-block ForwardEuler. */
double hfo_step(hfo_mesh* m);
/* nsteps of hfo_step; dt_steps[n] receives each step's dt.  Returns the number
 * of TimeStep calls logged into dt_calls (at most max_calls). */
int hfo_run(hfo_mesh* m, int nsteps, double* dt_steps, double* dt_calls, int max_calls);
/* The log of every TimeStep call made so far (step and predictor calls). */
int hfo_dt_log(const hfo_mesh* m, double* out, int max);

/* ---- unstructured quadrilateral-cell Godunov solver (src/fvm_2d_uns) ---- */
typedef struct hfo_uns hfo_uns;
/* vertices [n][4][2] (counter-clockwise), codes [n][4] (0 interior, -1
 * reflective, -2 transmissive, -3 inlet); inlet: conserved state. */
hfo_uns* hfo_uns_create(int ncells, const double* vertices, const int* codes, double gamma, double cfl,
                        const double* inlet);
void hfo_uns_destroy(hfo_uns* m);
/* Neighbour cell or boundary code of every cell edge [n][4]. */
void hfo_uns_connectivity(const hfo_uns* m, int* out);
void hfo_uns_geometry(const hfo_uns* m, double* centroids, double* volumes);
void hfo_uns_set_fields(hfo_uns* m, const double* U);   /* [n][4] */
void hfo_uns_get_fields(const hfo_uns* m, double* U);
double hfo_uns_time_step(hfo_uns* m);                   /* counted TimeStep call */
void hfo_uns_advance(hfo_uns* m, double dt);            /* ForwardEuler(GodunovScheme) */

/* Point-wise pieces, for unit tests. */
void hfo_cons_to_prim(double gamma, const double* U, double* W);
void hfo_prim_to_cons(double gamma, const double* W, double* U);
void hfo_cons_flux(double gamma, const double* U, int dir, double* F);
void hfo_hllc(double gamma, const double* UL, const double* UR, int dir, double* F);
double hfo_limiter(int kind, double r);

#ifdef __cplusplus
}
#endif
#endif
