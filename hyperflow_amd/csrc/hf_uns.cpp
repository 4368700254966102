// hf_uns.cpp -- C ABI of the unstructured quadrilateral-cell Godunov solver
// (include/hyperflow_uns.h): geometry and connectivity on the host, in the
// reference's evaluation order, then the device-resident step loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <unordered_map>
#include <map>
#include <tuple>
#include <vector>

#include "../../include/hyperflow_uns.h"
#include "hf_context.h"
#include "hf_uns.h"

struct hf_uns {
    int n = 0;
    double gamma = 1.4, cfl = 0.5, inlet[4] = {0, 0, 0, 0};
    std::vector<int32_t> nbr;        // [n][4]
    std::vector<double> cen, vol;    // [n][2], [n]
    hipStream_t stream = nullptr;
    int32_t* d_nbr = nullptr;
    double* d_len = nullptr;
    double* d_cs = nullptr;
    double* d_iv = nullptr;
    double* buf[2] = {nullptr, nullptr};
    int cur = 0;
    unsigned long long* d_slots = nullptr;   // [3][hfu::SLOT_WORDS]
    int slot = -1;                           // slot holding buf[cur]'s minimum, or -1
    double* d_hist = nullptr;
    int hist_cap = 1 << 20;
    int64_t calls = 0;
    hfu::UnsMesh mesh;
    // graph-captured loops: the device's TimeStep call counter (two words,
    // alternating by parity), its pinned staging word, and the graphs by
    // (steps, slot, buffer parity, call-count parity)
    unsigned long long* d_kcall = nullptr;
    unsigned long long* h_kcall = nullptr;
    hipEvent_t kcall_copied = nullptr;
    std::map<std::tuple<int, int, int, int>, std::pair<hipGraph_t, hipGraphExec_t>> graphs;
    bool profile = false;                    // HIP events around every step launch
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pending, ev_free;
    int64_t prof_launches = 0;
    double prof_ms = 0.0;
};

namespace {

int fail(int code, const std::string& msg) { return hf_fail(code, msg); }

inline uint64_t bits(double x)
{
    uint64_t b;
    std::memcpy(&b, &x, sizeof b);
    return b;
}

// Edge(start, end) (geometry/edge/edge.cpp:12-238): length, and the cosine
// and sine of the outward normal angle and of its negative
// (rotate_flow_values, model/euler_ideal_gas/euler_ideal_gas.cpp:152-164).
void edge_geometry(const double* a, const double* b, double& len, double cs[4])
{
    const double px = b[0] - a[0], py = b[1] - a[1];
    len = std::sqrt((0.0 + px * px) + py * py);
    double n0 = py, n1 = -1.0 * px;
    const double nl = std::sqrt(n0 * n0 + n1 * n1);
    n0 /= nl;
    n1 /= nl;
    const double deg = std::atan2(n1, n0) * 180.0 / M_PI;
    const double r = deg * (M_PI / 180.0);
    const double rb = (-1.0 * deg) * (M_PI / 180.0);
    // cos and sin of one angle as the reference's g++ -O3 build evaluates
    // them: one glibc sincos call (separate cos and sin calls differ from it
    // in the last bit for some angles)
    ::sincos(r, &cs[1], &cs[0]);
    ::sincos(rb, &cs[3], &cs[2]);
}

// Edge::opposite_edges_equal (edge.cpp:47-66), tolerance 1e-12
bool opposite(const double* a0, const double* a1, const double* b0, const double* b1)
{
    const double tol = 1e-12;
    return std::fabs(b0[0] - a1[0]) < tol && std::fabs(b0[1] - a1[1]) < tol && std::fabs(b1[0] - a0[0]) < tol &&
           std::fabs(b1[1] - a0[1]) < tol;
}

// The reference scans every cell edge for each edge (O(n^2)); the same
// relation through a hash of edge start points on a grid coarser than the
// tolerance, searching the neighbouring grid cells, keeping the last match
// in (cell, edge) order.  Non-finite coordinates are rejected by the caller.
void connect(int n, const double* V, const int32_t* codes, std::vector<int32_t>& nbr)
{
    // grid spacing: at least the tolerance (a match lies in the same or an
    // adjacent grid cell), and coarse enough that x / h fits an integer
    double big = 0.0;
    for (size_t i = 0; i < (size_t)8 * n; i++) big = std::max(big, std::fabs(V[i]));
    const double h = std::max(1e-9, big * 0x1p-50);
    auto key = [&](double x, double y) {
        const long long i = (long long)std::floor(x / h), j = (long long)std::floor(y / h);
        return (unsigned long long)(i * 1000003LL) ^ (unsigned long long)(j * 998244353LL);
    };
    auto cell_of = [&](double x) { return (long long)std::floor(x / h); };
    std::unordered_multimap<unsigned long long, int> starts;
    starts.reserve((size_t)n * 4 * 2);
    for (int d = 0; d < n; d++)
        for (int f = 0; f < 4; f++) {
            const double* p = V + 8 * d + 2 * f;
            starts.emplace(key(p[0], p[1]), 4 * d + f);
        }
    nbr.assign((size_t)n * 4, 0);
    for (int c = 0; c < n; c++)
        for (int e = 0; e < 4; e++) {
            const double* a0 = V + 8 * c + 2 * e;
            const double* a1 = V + 8 * c + 2 * ((e + 1) % 4);
            int best = -1;
            const long long ci = cell_of(a1[0]), cj = cell_of(a1[1]);
            for (long long di = -1; di <= 1; di++)
                for (long long dj = -1; dj <= 1; dj++) {
                    const unsigned long long k = (unsigned long long)((ci + di) * 1000003LL) ^
                                                 (unsigned long long)((cj + dj) * 998244353LL);
                    auto range = starts.equal_range(k);
                    for (auto it = range.first; it != range.second; ++it) {
                        const int id = it->second, d = id / 4, f = id % 4;
                        if (cell_of(V[8 * d + 2 * f]) != ci + di || cell_of(V[8 * d + 2 * f + 1]) != cj + dj)
                            continue;   // hash collision of another grid cell
                        if (opposite(a0, a1, V + 8 * d + 2 * f, V + 8 * d + 2 * ((f + 1) % 4)) && id > best) best = id;
                    }
                }
            nbr[4 * c + e] = best >= 0 ? best / 4 : codes[4 * c + e];
        }
}

int ensure_slot(hf_uns* u)
{
    if (u->slot >= 0) return HF_OK;
    HF_HIP(hfu::launch_uns_scan(u->mesh, u->buf[u->cur], u->d_slots, 0, u->stream));
    u->slot = 0;
    return HF_OK;
}

// one step launch, its dispatch stamped by a pair of events when profiling
hipError_t launch_step(hf_uns* u, const hfu::UnsStep& s)
{
    if (!u->profile) return hfu::launch_uns_step(u->mesh, s, u->stream);
    std::pair<hipEvent_t, hipEvent_t> e;
    if (u->ev_free.empty()) {
        hipError_t r = hipEventCreate(&e.first);
        if (r == hipSuccess) r = hipEventCreate(&e.second);
        if (r != hipSuccess) return r;
    } else {
        e = u->ev_free.back();
        u->ev_free.pop_back();
    }
    const hipError_t r = hfu::launch_uns_step(u->mesh, s, u->stream, e.first, e.second);
    u->ev_pending.push_back(e);
    return r;
}

hfu::UnsStep step_args(hf_uns* u)
{
    hfu::UnsStep s;
    std::memset(&s, 0, sizeof s);
    s.A = u->buf[u->cur];
    s.B = u->buf[u->cur ^ 1];
    s.slots = u->d_slots;
    s.hist = u->d_hist;
    s.hist_cap = u->hist_cap;
    return s;
}

}  // namespace

extern "C" {

int hf_uns_create(const hf_uns_desc* d, hf_uns** out)
{
    if (!d || !out) return fail(HF_EINVAL, "hf_uns_create: null argument");
    *out = nullptr;
    if (d->ncells < 1 || !d->vertices || !d->edge_types) return fail(HF_EINVAL, "hf_uns_create: no cells");
    if (d->ncells > (1 << 28)) return fail(HF_EINVAL, "hf_uns_create: more than 2^28 cells");
    if (!std::isfinite(d->gamma) || !std::isfinite(d->cfl) || d->cfl <= 0.0)
        return fail(HF_EINVAL, "hf_uns_create: gamma and cfl must be finite, cfl > 0");
    const int n = d->ncells;
    for (size_t i = 0; i < (size_t)8 * n; i++)
        if (!std::isfinite(d->vertices[i])) return fail(HF_EINVAL, "hf_uns_create: non-finite vertex coordinate");
    for (int i = 0; i < 4 * n; i++) {
        const int t = d->edge_types[i];
        if (t != HF_UNS_INTERIOR && t != HF_UNS_REFLECTIVE && t != HF_UNS_TRANSMISSIVE && t != HF_UNS_INLET)
            return fail(HF_EINVAL, "hf_uns_create: edge " + std::to_string(i) + ": boundary code " +
                                       std::to_string(t) + " not supported");
    }
    std::unique_ptr<hf_uns> u(new hf_uns());
    u->n = n;
    u->gamma = d->gamma;
    u->cfl = d->cfl;
    for (int k = 0; k < 4; k++) u->inlet[k] = d->inlet[k];

    const double* V = d->vertices;
    std::vector<double> len((size_t)4 * n), cs((size_t)16 * n), iv(n);
    bool mirrored = true;   // every edge's back rotation is (cos, -sin) of the forward one, bit for bit
    u->cen.resize((size_t)2 * n);
    u->vol.resize(n);
    for (int c = 0; c < n; c++) {
        const double* q = V + 8 * c;
        // Polygon::calculate_volume / calculate_centroid (polygon.cpp:162-210)
        double vol = 0.0;
        for (int i = 0; i < 4; i++) {
            const int j = (i + 1) % 4;
            vol += (q[2 * i] * q[2 * j + 1] - q[2 * j] * q[2 * i + 1]);
        }
        vol *= 0.5;
        double cx = 0.0, cy = 0.0;
        for (int i = 0; i < 4; i++) {
            const int j = (i + 1) % 4;
            const double second = (q[2 * i] * q[2 * j + 1] - q[2 * j] * q[2 * i + 1]);
            cx += ((q[2 * i] + q[2 * j]) * second);
            cy += ((q[2 * i + 1] + q[2 * j + 1]) * second);
        }
        u->vol[c] = vol;
        u->cen[2 * c] = cx / (6.0 * vol);
        u->cen[2 * c + 1] = cy / (6.0 * vol);
        iv[c] = 1.0 / vol;   // godunov.cpp:71
        // the incircle radius (shortest edge, polygon.cpp:79-91) is taken from
        // these lengths on the device
        for (int e = 0; e < 4; e++) {
            double l, r[4];
            edge_geometry(q + 2 * e, q + 2 * ((e + 1) % 4), l, r);
            len[(size_t)e * n + c] = l;
            for (int k = 0; k < 4; k++) cs[(size_t)(4 * e + k) * n + c] = r[k];
            mirrored = mirrored && bits(r[2]) == bits(r[0]) && bits(r[3]) == bits(-r[1]);
        }
    }
    connect(n, V, d->edge_types, u->nbr);
    std::vector<int32_t> nbr_em((size_t)4 * n);
    for (int c = 0; c < n; c++)
        for (int e = 0; e < 4; e++) nbr_em[(size_t)e * n + c] = u->nbr[4 * c + e];

    hf_uns* U = u.get();
#define HF_HIPU(call)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess) {                                                                        \
            hf_uns_destroy(u.release());                                                               \
            return fail(HF_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));                  \
        }                                                                                              \
    } while (0)
    HF_HIPU(hipStreamCreateWithFlags(&U->stream, hipStreamNonBlocking));
    HF_HIPU(hipMalloc(&U->d_nbr, sizeof(int32_t) * 4 * n));
    HF_HIPU(hipMalloc(&U->d_len, sizeof(double) * 4 * n));
    // the device keeps 2 rotation terms per edge when the back rotation
    // mirrors the forward one (glibc's sin is odd and cos even), else all 4
    const int terms = mirrored ? 2 : 4;
    std::vector<double> cst((size_t)4 * terms * n);
    for (int e = 0; e < 4; e++)
        for (int k = 0; k < terms; k++)
            std::memcpy(&cst[(size_t)(terms * e + k) * n], &cs[(size_t)(4 * e + k) * n], sizeof(double) * n);
    HF_HIPU(hipMalloc(&U->d_cs, sizeof(double) * cst.size()));
    HF_HIPU(hipMalloc(&U->d_iv, sizeof(double) * n));
    HF_HIPU(hipMalloc(&U->buf[0], sizeof(double) * 4 * n));
    HF_HIPU(hipMalloc(&U->buf[1], sizeof(double) * 4 * n));
    HF_HIPU(hipMalloc(&U->d_slots, sizeof(unsigned long long) * 3 * hfu::SLOT_WORDS));
    HF_HIPU(hipMemset(U->d_slots, 0, sizeof(unsigned long long) * 3 * hfu::SLOT_WORDS));
    HF_HIPU(hipMalloc(&U->d_hist, sizeof(double) * U->hist_cap));
    HF_HIPU(hipMalloc(&U->d_kcall, sizeof(unsigned long long) * 2));
    HF_HIPU(hipHostMalloc((void**)&U->h_kcall, sizeof(unsigned long long), hipHostMallocDefault));
    HF_HIPU(hipEventCreateWithFlags(&U->kcall_copied, hipEventDisableTiming));
    HF_HIPU(hipMemcpy(U->d_nbr, nbr_em.data(), sizeof(int32_t) * 4 * n, hipMemcpyHostToDevice));
    HF_HIPU(hipMemcpy(U->d_len, len.data(), sizeof(double) * 4 * n, hipMemcpyHostToDevice));
    HF_HIPU(hipMemcpy(U->d_cs, cst.data(), sizeof(double) * cst.size(), hipMemcpyHostToDevice));
    HF_HIPU(hipMemcpy(U->d_iv, iv.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    HF_HIPU(hipMemset(U->buf[0], 0, sizeof(double) * 4 * n));
    HF_HIPU(hipMemset(U->buf[1], 0, sizeof(double) * 4 * n));
    HF_HIPU(hipMemset(U->d_hist, 0, sizeof(double) * U->hist_cap));
#undef HF_HIPU
    hfu::UnsMesh& m = U->mesh;
    std::memset(&m, 0, sizeof m);
    m.n = n;
    m.nbr = U->d_nbr;
    m.len = U->d_len;
    m.cs = U->d_cs;
    m.inv_vol = U->d_iv;
    m.cs_terms = terms;
    m.gamma = U->gamma;
    m.cfl = U->cfl;
    for (int k = 0; k < 4; k++) m.inlet[k] = U->inlet[k];
    *out = u.release();
    return HF_OK;
}

void hf_uns_destroy(hf_uns* u)
{
    if (!u) return;
    if (hf_process_exiting()) {   // host memory only once the process is exiting
        delete u;
        return;
    }
    if (u->stream) (void)hipStreamSynchronize(u->stream);
    for (auto& kv : u->graphs) {
        (void)hipGraphExecDestroy(kv.second.second);
        (void)hipGraphDestroy(kv.second.first);
    }
    for (auto* list : {&u->ev_pending, &u->ev_free})
        for (auto& e : *list) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
    if (u->kcall_copied) (void)hipEventDestroy(u->kcall_copied);
    if (u->h_kcall) (void)hipHostFree(u->h_kcall);
    for (void* p : {(void*)u->d_nbr, (void*)u->d_len, (void*)u->d_cs, (void*)u->d_iv, (void*)u->buf[0],
                    (void*)u->buf[1], (void*)u->d_slots, (void*)u->d_hist, (void*)u->d_kcall})
        if (p) (void)hipFree(p);
    if (u->stream) (void)hipStreamDestroy(u->stream);
    delete u;
}

int hf_uns_cells(const hf_uns* u) { return u ? u->n : 0; }

int hf_uns_bytes_per_cell(const hf_uns* u)
{
    if (!u) return fail(HF_EINVAL, "hf_uns_bytes_per_cell: null context");
    return 2 * 4 * 8 + 4 * 4 + 4 * 8 + 4 * u->mesh.cs_terms * 8 + 8;
}

int hf_uns_connectivity(const hf_uns* u, int32_t* out)
{
    if (!u || !out) return fail(HF_EINVAL, "hf_uns_connectivity: null argument");
    std::memcpy(out, u->nbr.data(), sizeof(int32_t) * 4 * u->n);
    return HF_OK;
}

int hf_uns_geometry(const hf_uns* u, double* centroids, double* volumes)
{
    if (!u) return fail(HF_EINVAL, "hf_uns_geometry: null context");
    if (centroids) std::memcpy(centroids, u->cen.data(), sizeof(double) * 2 * u->n);
    if (volumes) std::memcpy(volumes, u->vol.data(), sizeof(double) * u->n);
    return HF_OK;
}

int hf_uns_upload(hf_uns* u, const double* f, size_t n)
{
    if (!u || !f) return fail(HF_EINVAL, "hf_uns_upload: null argument");
    if (n != (size_t)4 * u->n) return fail(HF_EINVAL, "hf_uns_upload: expected 4*ncells values");
    HF_HIP(hipStreamSynchronize(u->stream));
    HF_HIP(hipMemcpy(u->buf[u->cur], f, sizeof(double) * n, hipMemcpyHostToDevice));
    u->slot = -1;
    return HF_OK;
}

int hf_uns_download(hf_uns* u, double* f, size_t n)
{
    if (!u || !f) return fail(HF_EINVAL, "hf_uns_download: null argument");
    if (n != (size_t)4 * u->n) return fail(HF_EINVAL, "hf_uns_download: expected 4*ncells values");
    HF_HIP(hipStreamSynchronize(u->stream));
    HF_HIP(hipMemcpy(f, u->buf[u->cur], sizeof(double) * n, hipMemcpyDeviceToHost));
    return HF_OK;
}

int hf_uns_time_step(hf_uns* u, double* dt)
{
    if (!u || !dt) return fail(HF_EINVAL, "hf_uns_time_step: null argument");
    int rc = ensure_slot(u);
    if (rc) return rc;
    std::vector<unsigned long long> w(hfu::SLOT_WORDS);
    HF_HIP(hipMemcpyAsync(w.data(), u->d_slots + (size_t)u->slot * hfu::SLOT_WORDS, sizeof(unsigned long long) * w.size(),
                          hipMemcpyDeviceToHost, u->stream));
    HF_HIP(hipStreamSynchronize(u->stream));
    unsigned long long b = w[0];
    for (int k = 1; k < hfu::SLOT_WAYS; k++) b = std::min(b, w[(size_t)k * hfu::SLOT_STRIDE]);
    double m;
    std::memcpy(&m, &b, sizeof m);
    const int64_t k = ++u->calls;
    const double cf = (k <= 10) ? 0.2 * u->cfl : u->cfl;   // timestep.cpp:57-61
    *dt = cf * m;
    if (k <= u->hist_cap) HF_HIP(hipMemcpy(u->d_hist + (k - 1), dt, sizeof(double), hipMemcpyHostToDevice));
    return HF_OK;
}

int hf_uns_advance(hf_uns* u, double dt)
{
    if (!u) return fail(HF_EINVAL, "hf_uns_advance: null context");
    hfu::UnsStep s = step_args(u);
    s.use_fixed = 1;
    s.fixed_dt = dt;
    s.rd = 0;
    s.wr = -1;
    s.rs = -1;
    HF_HIP(launch_step(u, s));
    u->cur ^= 1;
    u->slot = -1;
    return HF_OK;
}

int hf_uns_run(hf_uns* u, int nsteps, int flags)
{
    if (!u) return fail(HF_EINVAL, "hf_uns_run: null context");
    if (flags & ~HF_UNS_RUN_GRAPH) return fail(HF_EINVAL, "hf_uns_run: unknown flags");
    if (nsteps < 0) return fail(HF_EINVAL, "hf_uns_run: nsteps must be >= 0");
    if (nsteps == 0) return HF_OK;
    int rc = ensure_slot(u);
    if (rc) return rc;
    // slots rotate: step reads rd, writes (rd+1)%3 (reset by the previous
    // step or here), resets (rd+2)%3 for the step after
    static const std::vector<unsigned long long> empty(hfu::SLOT_WORDS, hfu::SLOT_EMPTY);
    HF_HIP(hipMemcpyAsync(u->d_slots + (size_t)((u->slot + 1) % 3) * hfu::SLOT_WORDS, empty.data(),
                          sizeof(unsigned long long) * empty.size(), hipMemcpyHostToDevice, u->stream));
    if (flags & HF_UNS_RUN_GRAPH) {
        // the call counter goes to the device (word calls % 2), outside the
        // graph; every launch argument of the loop then depends only on the
        // slot, the buffer parity and the call-count parity at its start
        const int par = (int)(u->calls & 1);
        HF_HIP(hipEventSynchronize(u->kcall_copied));   // the previous copy has read the staging word
        *u->h_kcall = (unsigned long long)u->calls;
        HF_HIP(hipMemcpyAsync(u->d_kcall + par, u->h_kcall, sizeof(unsigned long long), hipMemcpyHostToDevice,
                              u->stream));
        HF_HIP(hipEventRecord(u->kcall_copied, u->stream));
        const auto key = std::make_tuple(nsteps, u->slot, u->cur, par);
        auto it = u->graphs.find(key);
        if (it == u->graphs.end()) {
            hipGraph_t g = nullptr;
            hipGraphExec_t ge = nullptr;
            HF_HIP(hipStreamBeginCapture(u->stream, hipStreamCaptureModeThreadLocal));
            int cur = u->cur, slot = u->slot;
            hipError_t le = hipSuccess;
            for (int i = 0; i < nsteps && le == hipSuccess; i++) {
                hfu::UnsStep s = step_args(u);
                s.A = u->buf[cur];
                s.B = u->buf[cur ^ 1];
                s.rd = slot;
                s.wr = (slot + 1) % 3;
                s.rs = (slot + 2) % 3;
                s.kdev = u->d_kcall;
                s.kpar = (par + i) & 1;
                le = hfu::launch_uns_step(u->mesh, s, u->stream);
                cur ^= 1;
                slot = s.wr;
            }
            const hipError_t ce = hipStreamEndCapture(u->stream, &g);
            if (le != hipSuccess || ce != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                HF_HIP(le != hipSuccess ? le : ce);
            }
            HF_HIP(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
            it = u->graphs.emplace(key, std::make_pair(g, ge)).first;
        }
        HF_HIP(hipGraphLaunch(it->second.second, u->stream));
        u->calls += nsteps;
        u->cur ^= (nsteps & 1);
        u->slot = (u->slot + nsteps) % 3;
        return HF_OK;
    }
    for (int i = 0; i < nsteps; i++) {
        hfu::UnsStep s = step_args(u);
        s.rd = u->slot;
        s.wr = (u->slot + 1) % 3;
        s.rs = (u->slot + 2) % 3;
        s.kcall = ++u->calls;
        HF_HIP(launch_step(u, s));
        u->cur ^= 1;
        u->slot = s.wr;
    }
    return HF_OK;
}

int hf_uns_synchronize(hf_uns* u)
{
    if (!u) return fail(HF_EINVAL, "null context");
    HF_HIP(hipStreamSynchronize(u->stream));
    return HF_OK;
}

int hf_uns_dt_history(hf_uns* u, double* out, int max, int* count)
{
    if (!u) return fail(HF_EINVAL, "null context");
    const int64_t n = std::min<int64_t>(u->calls, u->hist_cap);
    if (count) *count = (int)n;
    if (out && max > 0) {
        HF_HIP(hipStreamSynchronize(u->stream));
        HF_HIP(hipMemcpy(out, u->d_hist, sizeof(double) * std::min<int64_t>(n, max), hipMemcpyDeviceToHost));
    }
    return HF_OK;
}

int hf_uns_profile(hf_uns* u, int enable)
{
    if (!u) return fail(HF_EINVAL, "hf_uns_profile: null context");
    u->profile = enable != 0;
    return HF_OK;
}

int hf_uns_kernel_stats(hf_uns* u, int64_t* launches, double* total_ms)
{
    if (!u) return fail(HF_EINVAL, "hf_uns_kernel_stats: null context");
    HF_HIP(hipStreamSynchronize(u->stream));
    for (auto& e : u->ev_pending) {
        float ms = 0.f;
        HF_HIP(hipEventElapsedTime(&ms, e.first, e.second));
        u->prof_ms += ms;
        u->prof_launches++;
        u->ev_free.push_back(e);
    }
    u->ev_pending.clear();
    if (launches) *launches = u->prof_launches;
    if (total_ms) *total_ms = u->prof_ms;
    u->prof_launches = 0;
    u->prof_ms = 0.0;
    return HF_OK;
}

}  // extern "C"
