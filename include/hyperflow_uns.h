/*
 * hyperflow_uns.h -- C ABI of the MI355X unstructured quadrilateral-cell
 * Godunov solver: HyperFlow's src/fvm_2d_uns (rotated HLLC fluxes, one
 * TimeStep per step, forward Euler).  Each entry point names the reference
 * interface it replaces (paths relative to the reference checkout).
 *
 * Cells are quadrilaterals given by four vertices, counter-clockwise, in the
 * order of the .hmsh mesh file; cell edge k runs from vertex k to vertex
 * k+1 (mod 4).  Fields are conserved variables (rho, rho*u, rho*v, E) laid
 * out [4][ncells].  FP64 results are bitwise equal to the reference's.
 * Conventions as in hyperflow.h (status codes, hf_last_error, one context
 * per thread, work in order on the context's own HIP stream).
 */
#ifndef HYPERFLOW_UNS_H
#define HYPERFLOW_UNS_H

#include "hyperflow.h"

#ifdef __cplusplus
extern "C" {
#endif

/* boundary codes of src/fvm_2d_uns/scheme/boundary/boundary.h:12 */
enum { HF_UNS_INTERIOR = 0, HF_UNS_REFLECTIVE = -1, HF_UNS_TRANSMISSIVE = -2, HF_UNS_INLET = -3 };

typedef struct {
    int32_t ncells;
    int32_t pad;
    const double* vertices;     /* [ncells][4][2] */
    const int32_t* edge_types;  /* [ncells][4]: the .hmsh EDGES codes */
    double gamma;               /* EulerIdealGasModel(gamma) */
    double cfl;                 /* TimeStep(model, cfl) */
    double inlet[4];            /* conserved inlet state (parser.cpp:211-217) */
} hf_uns_desc;

typedef struct hf_uns hf_uns;

/* Mesh + geometry + connectivity (LoadStructuredMesh::load_structured_mesh,
 * mesh/mesh_structured/mesh_structured.cpp:23-122: an edge's neighbour is the
 * last cell holding the same edge in the opposite direction, else the edge's
 * code) and the Godunov scheme with its boundary conditions
 * (parser.cpp:206-230). */
HF_API int hf_uns_create(const hf_uns_desc* desc, hf_uns** out);
HF_API void hf_uns_destroy(hf_uns* u);
HF_API int hf_uns_cells(const hf_uns* u);
/* Neighbour cell or boundary code of every cell edge, [ncells][4]
 * (Mesh::get_cell_edge_indices, mesh/mesh/mesh.cpp:54-57). */
HF_API int hf_uns_connectivity(const hf_uns* u, int32_t* out);
/* Polygon::get_centroid / get_volume (geometry/polygon/polygon.cpp:162-210). */
HF_API int hf_uns_geometry(const hf_uns* u, double* centroids, double* volumes);
/* Mesh::set_cells_flow_values / get_cells_flow_values (mesh.cpp:27-44). */
HF_API int hf_uns_upload(hf_uns* u, const double* fields, size_t n);
HF_API int hf_uns_download(hf_uns* u, double* fields, size_t n);
/* TimeStep::operator()(mesh) (scheme/timestep/timestep.cpp:29-62): one
 * counted call over every cell. */
HF_API int hf_uns_time_step(hf_uns* u, double* dt);
/* ForwardEulerODESolver(GodunovScheme) with the given dt
 * (godunov.cpp:31-77, forward_euler.cpp:16-30). */
HF_API int hf_uns_advance(hf_uns* u, double dt);
/* nsteps of Simulation::solve's step body (simulation.cpp:85-118, without
 * clamps or output) with the dt kept on the device; asynchronous.  One launch
 * per step; flags: 0, or HF_UNS_RUN_GRAPH to capture the loop in a HIP graph
 * (cached per step count and starting state, replayed on later calls; the
 * TimeStep call counter then lives on the device). */
enum { HF_UNS_RUN_GRAPH = 1 };
HF_API int hf_uns_run(hf_uns* u, int nsteps, int flags);
HF_API int hf_uns_synchronize(hf_uns* u);
/* Every TimeStep result so far, in call order. */
HF_API int hf_uns_dt_history(hf_uns* u, double* out, int max, int* count);
/* Diagnostics, no counterpart in the reference: HIP events around every
 * step launch while enabled; stats return launches and summed milliseconds
 * since the last call, then reset. */
HF_API int hf_uns_profile(hf_uns* u, int enable);
/* Bytes one step reads and writes per cell, counting each array once:
 * state in and out, neighbour indices, edge lengths, rotation terms (2 per
 * edge when the back rotation mirrors the forward one on every edge, else
 * 4), 1/V.  No counterpart in the reference (roofline accounting). */
HF_API int hf_uns_bytes_per_cell(const hf_uns* u);
HF_API int hf_uns_kernel_stats(hf_uns* u, int64_t* launches, double* total_ms);

#ifdef __cplusplus
}
#endif
#endif
