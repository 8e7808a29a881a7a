// fta_batch.cpp -- batched form: several independent meshes, each with its
// own tracers, advanced together by ONE hipGraph launch per step.
//
// Each mesh lives in its own Engine (its own stream and device arrays). A
// step of the whole batch is captured once: a fork event on the batch
// stream releases every engine's stream, each engine enqueues its step (all
// its tracers, batched by scheme), and join events bring the branches back.
// The instantiated graph therefore runs the meshes concurrently -- small
// meshes fill the GPU next to large ones -- and replaying it costs one launch
// per step. No reference counterpart (the reference advances one mesh per
// MPI rank).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/fesom_tracer_amd.h"
#include "fta_engine.h"
#include "fta_mesh.h"

using namespace fta;

namespace {

struct BatchMesh {
    HostMesh mesh;
    std::unique_ptr<Engine> eng;
    double dt = 0.0;
    int ntr = 0;
    double Lx = 0.0, Ly = 0.0;
    std::vector<double> uv, w;  // the velocity the engine holds (w_e = w)
};

struct Batch {
    std::vector<std::unique_ptr<BatchMesh>> meshes;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> joins;
    bool valid = false;

    void drop_graph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        valid = false;
    }
    ~Batch()
    {
        drop_graph();
        meshes.clear();
        for (hipEvent_t e : joins) (void)hipEventDestroy(e);
        if (fork) (void)hipEventDestroy(fork);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

std::unique_ptr<Batch> g_batch;

int bfail(const std::string &msg)
{
    set_error(msg);
    std::fprintf(stderr, "### error: %s\n", msg.c_str());
    return -1;
}

#define BCHECK(call)                                                                   \
    do {                                                                               \
        hipError_t _e = (call);                                                        \
        if (_e != hipSuccess) throw std::runtime_error(std::string(#call) + ": " + hipGetErrorString(_e)); \
    } while (0)

uint64_t splitmix(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double unit(uint64_t &s) { return (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

// tracer k of a mesh (column-major (nl-1, nod2D)): 0 = T (the reference's
// analytic-dwarf field), 1 = S (34 + blob), then passive tracers cycling
// Gaussian blob / 0-1 step / white noise, seeded per mesh and tracer
void fill_tracer(const BatchMesh &b, int k, uint64_t seed, std::vector<double> &v)
{
    const HostMesh &m = b.mesh;
    const int L = m.nl - 1, N = m.nod2D;
    v.assign((size_t)L * N, 0.0);
    uint64_t s = seed * 1000003ull + (uint64_t)k * 7919ull + 1ull;
    const double Lx = b.Lx, Ly = b.Ly, pi = 3.14159265358979323846;
    auto zn = [&](int l) { return L > 1 ? (double)l / (double)(L - 1) : 0.0; };
    auto pd = [](double d, double P) { return d - P * std::round(d / P); };
    const int kind = k < 2 ? k : 2 + (k - 2) % 3;
    double xc = 0, yc = 0, w = 1, x0 = 0, y0 = 0;
    if (kind == 1 || kind == 2) {
        xc = (0.2 + 0.6 * unit(s)) * Lx;
        yc = (0.2 + 0.6 * unit(s)) * Ly;
        w = (0.08 + 0.08 * unit(s)) * Lx;
    } else if (kind == 3) {
        x0 = (0.25 + 0.25 * unit(s)) * Lx;
        y0 = (0.25 + 0.25 * unit(s)) * Ly;
    }
    for (int n = 0; n < N; ++n) {
        const double x = m.coord[2 * n], y = m.coord[2 * n + 1];
        double col = 0.0;
        if (kind == 0) col = 15.0 + 5.0 * std::sin(2.0 * pi * x / Lx) * std::cos(2.0 * pi * y / Ly);
        if (kind == 1 || kind == 2) {
            const double dx = pd(x - xc, Lx), dy = pd(y - yc, Ly);
            col = std::exp(-(dx * dx + dy * dy) / (w * w));
        }
        if (kind == 3) col = (x >= x0 && x < x0 + 0.4 * Lx && y >= y0 && y < y0 + 0.4 * Ly) ? 1.0 : 0.0;
        for (int l = 0; l < L; ++l) {
            double val = col;
            if (kind == 0) val = col - 10.0 * zn(l);
            if (kind == 1) val = 34.0 + col * (1.0 - zn(l));
            if (kind == 4) val = unit(s);
            v[(size_t)n * L + l] = val;
        }
    }
}

BatchMesh *get_mesh(int i)
{
    if (!g_batch || i < 0 || i >= (int)g_batch->meshes.size()) return nullptr;
    return g_batch->meshes[i].get();
}

}  // namespace

extern "C" {

int tracer_batch_create(void)
{
    try {
        g_batch.reset(new Batch);
        BCHECK(hipStreamCreateWithFlags(&g_batch->stream, hipStreamNonBlocking));
        BCHECK(hipEventCreateWithFlags(&g_batch->fork, hipEventDisableTiming));
    } catch (const std::exception &e) {
        g_batch.reset();
        return bfail(e.what());
    }
    return 0;
}

int tracer_batch_destroy(void)
{
    g_batch.reset();
    return 0;
}

int tracer_batch_add_periodic_mesh(int nx, int ny, int nl, double Lx, double Ly, double dz, int lev_lo, int lev_hi,
                                   unsigned long long seed, int ntracers, int precision, const char *hor,
                                   const char *ver, const char *lim, double cfl, double *dt_out)
{
    if (!g_batch) return bfail("tracer_batch_create first");
    try {
        std::unique_ptr<BatchMesh> b(new BatchMesh);
        generate_periodic(b->mesh, nx, ny, Lx, Ly, 1);
        set_vertical_grid(b->mesh, nl, dz, 0.0);
        set_levels_random(b->mesh, lev_lo, lev_hi, seed);
        derive_all_device(b->mesh);  // the batch runs on the GPU anyway
        b->Lx = Lx;
        b->Ly = Ly;
        std::vector<double> &uv = b->uv, &w = b->w;
        b->dt = make_velocity(b->mesh, 0.1, 0.05, 0.2, 0.5, cfl, uv, w);
        const int prec = precision == 32 ? PREC_FP32 : (precision == 16 ? PREC_FP32_H16 : PREC_FP64);
        b->eng.reset(new Engine(b->mesh, ntracers, prec));
        b->eng->set_velocity(uv.data(), w.data(), w.data());
        TracerSpec sp;
        const std::string h = hor ? hor : "MUSCL", v = ver ? ver : "QR4C", l = lim ? lim : "FCT";
        sp.hor = h == "UPW1" ? HOR_UPW1 : (h == "MFCT" ? HOR_MFCT : HOR_MUSCL);
        sp.ver = v == "UPW1" ? VER_UPW1 : (v == "PPM" ? VER_PPM : (v == "CDIFF" ? VER_CDIFF : VER_QR4C));
        sp.lim = l == "NONE" ? LIM_NONE : (l == "LEARNED" ? LIM_LEARNED : LIM_FCT);
        b->ntr = ntracers;
        std::vector<double> vals;
        for (int t = 0; t < ntracers; ++t) {
            b->eng->set_spec(t, sp);
            fill_tracer(*b, t, seed, vals);
            b->eng->set_values(t, vals.data());
        }
        if (dt_out) *dt_out = b->dt;
        g_batch->meshes.push_back(std::move(b));
        g_batch->drop_graph();
        return (int)g_batch->meshes.size() - 1;
    } catch (const std::exception &e) {
        return bfail(e.what());
    }
}

int tracer_batch_run(int nsteps, int use_graph)
{
    if (!g_batch || g_batch->meshes.empty()) return bfail("empty batch");
    Batch &B = *g_batch;
    try {
        if (!use_graph) {
            for (int s = 0; s < nsteps; ++s)
                for (auto &b : B.meshes) b->eng->enqueue_step(b->dt, 0, b->ntr);
            for (auto &b : B.meshes) b->eng->synchronize();
            return 0;
        }
        if (!B.valid) {
            B.drop_graph();
            while (B.joins.size() < B.meshes.size()) {
                hipEvent_t e;
                BCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
                B.joins.push_back(e);
            }
            BCHECK(hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal));
            BCHECK(hipEventRecord(B.fork, B.stream));
            for (size_t i = 0; i < B.meshes.size(); ++i) {
                BatchMesh &b = *B.meshes[i];
                BCHECK(hipStreamWaitEvent(b.eng->stream(), B.fork, 0));
                b.eng->enqueue_step(b.dt, 0, b.ntr);
                BCHECK(hipEventRecord(B.joins[i], b.eng->stream()));
                BCHECK(hipStreamWaitEvent(B.stream, B.joins[i], 0));
            }
            BCHECK(hipStreamEndCapture(B.stream, &B.graph));
            BCHECK(hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0));
            B.valid = true;
        }
        for (int s = 0; s < nsteps; ++s) BCHECK(hipGraphLaunch(B.exec, B.stream));
        BCHECK(hipStreamSynchronize(B.stream));
    } catch (const std::exception &e) {
        return bfail(e.what());
    }
    return 0;
}

int tracer_batch_time(int nsteps, double *ms_out)
{
    if (!g_batch || g_batch->meshes.empty()) return bfail("empty batch");
    Batch &B = *g_batch;
    if (tracer_batch_run(0, 1)) return -1;  // capture + instantiate outside the timed region
    try {
        hipEvent_t a, b;
        BCHECK(hipEventCreate(&a));
        BCHECK(hipEventCreate(&b));
        BCHECK(hipEventRecord(a, B.stream));
        for (int s = 0; s < nsteps; ++s) BCHECK(hipGraphLaunch(B.exec, B.stream));
        BCHECK(hipEventRecord(b, B.stream));
        BCHECK(hipEventSynchronize(b));
        float ms = 0.0f;
        BCHECK(hipEventElapsedTime(&ms, a, b));
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        *ms_out = ms;
    } catch (const std::exception &e) {
        return bfail(e.what());
    }
    return 0;
}

int tracer_batch_autotune(int reps, double *table, int max_rows)
{
    if (!g_batch || g_batch->meshes.empty()) return bfail("empty batch");
    Batch &B = *g_batch;
    try {
        // each mesh's engine times its own schedules alone (its stream, its
        // tracers restored afterwards) and keeps the fastest
        B.drop_graph();
        int rows = 0;
        for (size_t i = 0; i < B.meshes.size(); ++i) {
            BatchMesh &b = *B.meshes[i];
            b.eng->autotune(b.dt, reps);
            int w = 0, nt = 0;
            b.eng->schedule(w, nt);
            if (table && rows < max_rows) {
                table[2 * rows] = (double)w;
                table[2 * rows + 1] = (double)nt;
            }
            ++rows;
        }
        return rows;
    } catch (const std::exception &e) {
        return bfail(e.what());
    }
}

int tracer_batch_set_schedule(int waves, int max_nt)
{
    if (!g_batch || g_batch->meshes.empty()) return bfail("empty batch");
    try {
        g_batch->drop_graph();
        for (auto &b : g_batch->meshes) b->eng->set_schedule(waves, max_nt);
    } catch (const std::exception &e) {
        return bfail(e.what());
    }
    return 0;
}

int tracer_batch_mesh_info(int mesh, int *nl, int *nod2D, int *ntracers, long long *active_cells, double *dt)
{
    BatchMesh *b = get_mesh(mesh);
    if (!b) return bfail("invalid batch mesh index");
    *nl = b->mesh.nl;
    *nod2D = b->mesh.nod2D;
    *ntracers = b->ntr;
    *active_cells = (long long)active_node_cells(b->mesh);
    *dt = b->dt;
    return 0;
}

int tracer_batch_get_values(int mesh, int tracer_id, double *values)
{
    BatchMesh *b = get_mesh(mesh);
    if (!b) return bfail("invalid batch mesh index");
    if (tracer_id < 1 || tracer_id > b->ntr) return bfail("invalid tracer_id");
    try { b->eng->get_values(tracer_id - 1, values); } catch (const std::exception &e) { return bfail(e.what()); }
    return 0;
}

int tracer_batch_get_mesh_array(int mesh, const char *name, void *out, long long count, long long *size, int *kind_out)
{
    BatchMesh *b = get_mesh(mesh);
    if (!b) return bfail("invalid batch mesh index");
    const std::string n = name ? name : "";
    const void *p = nullptr;
    long long cnt = 0;
    int kind = 1;
    if (n == "uv") { p = b->uv.data(); cnt = (long long)b->uv.size(); }
    else if (n == "w" || n == "w_e") { p = b->w.data(); cnt = (long long)b->w.size(); }
    else if (!named_mesh_array(b->mesh, n, p, cnt, kind)) return bfail("unknown mesh array " + n);
    if (size) *size = cnt;
    if (kind_out) *kind_out = kind;
    if (!out) return 0;
    if (count < cnt) return bfail("tracer_batch_get_mesh_array: buffer too small");
    std::memcpy(out, p, (size_t)cnt * (kind ? sizeof(double) : sizeof(int)));
    return 0;
}

int tracer_batch_diagnostics(int mesh, int tracer_id, double *out)
{
    BatchMesh *b = get_mesh(mesh);
    if (!b) return bfail("invalid batch mesh index");
    if (tracer_id < 1 || tracer_id > b->ntr) return bfail("invalid tracer_id");
    try {
        Diag d = b->eng->diagnostics(tracer_id - 1);
        out[0] = d.vmin; out[1] = d.vmax; out[2] = d.sum;
        out[3] = d.amin; out[4] = d.amax; out[5] = d.mass; out[6] = d.nan_count;
    } catch (const std::exception &e) {
        return bfail(e.what());
    }
    return 0;
}

}  // extern "C"
