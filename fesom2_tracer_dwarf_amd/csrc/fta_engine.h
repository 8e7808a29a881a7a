// fta_engine.h -- device engine for FESOM2's tracer-advection step on one
// MI355X (gfx950). Owns the device copy of the mesh, the velocity fields,
// the tracers and the work arrays, and launches the step on its own stream.
//
// Device layout: every level-indexed array stores only the wet levels of
// each column, columns packed whole, in mesh order (nodes, elements and
// edges follow a space-filling curve when the mesh was generated with one),
// into 128-cell chunks; a wavefront handles one chunk, two cells per lane, so
// a sweep over consecutive columns streams through contiguous memory
// (requires nl <= 128 level interfaces).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <string>
#include <vector>

#include "fta_mesh.h"

namespace fta {

enum Hor { HOR_UPW1 = 0, HOR_MUSCL = 1, HOR_MFCT = 2 };
enum Ver { VER_UPW1 = 0, VER_QR4C = 1, VER_PPM = 2, VER_CDIFF = 3 };
// LIM_LEARNED: a small network predicts each horizontal edge-level's limiter
// coefficient; it is applied in flux form and Zalesak's coefficient replaces
// it on every edge of a node whose bounds the prediction would break
enum Lim { LIM_NONE = 0, LIM_FCT = 1, LIM_LEARNED = 2 };
// FP64: everything in double (parity path). FP32: fluxes, gradients and the
// limiter in float, mesh and velocity stored as float, node tracer, LO
// solution and divergence accumulation kept in double. FP32_H16: FP32 with
// the elemental/nodal gradients stored as half.
enum Prec { PREC_FP64 = 0, PREC_FP32 = 1, PREC_FP32_H16 = 2 };
// how a step ends at each node layer: the tracer update of the analytic
// driver (del_ttf zeroed every step), the same with del_ttf carried over
// between steps as the reference's C interface does, or the tendencies
// dttf_h / dttf_v returned with the tracer unchanged (do_oce_adv_tra's own
// contract)
enum Out { OUT_UPDATE = 0, OUT_ACCUM = 1, OUT_TEND = 2 };

struct TracerSpec {
    int hor = HOR_UPW1, ver = VER_UPW1, lim = LIM_NONE;
    double ph = 1.0, pv = 1.0;  // tra_adv_ph / tra_adv_pv
};

struct Diag {
    double vmin, vmax, sum;          // over all (nl-1) x nod2D values, as tracer_get_stats
    double amin, amax, mass;         // over active cells; mass = sum T*hnode_new*areasvol
    double nan_count;
};

class Engine {
public:
    Engine(const HostMesh &m, int ntracers, int precision);
    ~Engine();
    Engine(const Engine &) = delete;
    Engine &operator=(const Engine &) = delete;

    int ntracers() const { return ntr_; }
    int precision() const { return prec_; }

    // (2,nl-1,elem2D), (nl,nod2D), (nl,nod2D) unpadded double arrays
    void set_velocity(const double *uv, const double *w, const double *we);
    void set_values(int tr, const double *v);    // (nl-1,nod2D)
    void set_valuesAB(int tr, const double *v);  // (nl-1,nod2D)
    void get_values(int tr, double *v);
    void set_spec(int tr, const TracerSpec &s);
    const TracerSpec &spec(int tr) const { return spec_.at(tr); }

    // del_ttf carried over between steps (one array per tracer, zeroed when
    // switched on), as the reference's tracer_advect_step never zeroes it
    void set_accumulate(bool on);
    bool accumulate() const;
    // one do_oce_adv_tra of tracer tr: its horizontal and vertical tendencies
    // are ADDED into the host arrays dttf_h, dttf_v ((nl-1, nod2D), wet layers)
    // and the tracer is left unchanged
    void tendencies(int tr, double dt, double *dttf_h, double *dttf_v);
    // the same into DEVICE arrays dttf_h, dttf_v ((nl-1, nod2D)), with the two
    // packed tendency arrays in the caller's device scratch (2 * node cells
    // doubles at most scratch_doubles, else an error); allocates nothing
    void tendencies_device(int tr, double dt, double *dttf_h, double *dttf_v, double *scratch,
                           size_t scratch_doubles);

    // refresh_ab: valuesAB := values at the start of each step (AB1)
    void set_refresh_ab(bool r) { refresh_ab_ = r; graph_valid_ = false; }
    static constexpr int kMaxBatch = 4;  // tracers per launch (and per work-array set)

    // enqueue one step for tracers [first, first+count) on the stream
    void enqueue_step(double dt, int first, int count);
    // nsteps steps of all tracers, replayed from a captured hipGraph
    void run_steps(int nsteps, double dt, bool use_graph);
    void synchronize();
    hipStream_t stream() const { return stream_; }

    Diag diagnostics(int tr);
    // copy a device work array out in the unpadded reference layout (the
    // work arrays are shared by the tracer batches of a step: a tracer's
    // values are those of the last step of its batch)
    // names: fct_LO, adv_flux_hor, adv_flux_ver, fct_plus, fct_minus, gnod
    int get_work(const std::string &name, int tr, double *out, size_t count);
    size_t work_count(const std::string &name) const;

    // tracer tr of this engine against the same tracer of ref (same mesh,
    // any precision), over the wet layers: out = {max|a-b|, ||a-b||_2,
    // max|a-b|/max|b|, ||a-b||_2/||b||_2}
    void difference(const Engine &ref, int tr, double out[4]);

    // ---- learned limiter (network 10-64-64-1, parameters flattened as
    // W1[f][h], b1, W2[h1][h2], b2, w3, b3: mlp_param_count() floats)
    static int mlp_param_count();
    void mlp_set(const float *params);
    void mlp_get(float *params);
    // nsteps FCT steps of all FCT/LEARNED tracers, appending (features,
    // Zalesak coefficient) rows of every edge level to the training set
    // (capacity max_rows); returns the rows held
    long long learn_collect(int nsteps, double dt, long long max_rows);
    void learn_clear();
    // Adam on the training set, minibatches of batch rows; mean loss per epoch
    std::vector<double> learn_train(int epochs, long long batch, double lr, unsigned seed);
    // the network on given rows of features (rows x 10 floats)
    void mlp_forward(const float *X, long long rows, float *out);
    // gradient of sum((sigmoid(z) - y)^2) * scale over the rows (no update)
    void mlp_gradient(const float *X, const float *y, long long rows, float scale, float *grad, double *loss);
    // fraction of wet layers of tracer tr whose Zalesak fallback was taken in the last step
    double learn_fallback(int tr);

    // nodal mean gradients over each node's fan (default; meshes whose nodes
    // all have one) or element by element (the same values to rounding)
    void set_gradient_fan(bool on);

    // launch schedule of the chunk kernels: waves per workgroup (1, 2, 4, 8)
    // and most tracers per launch (1, 2, 4); results do not depend on it
    void set_schedule(int waves, int max_nt);
    void schedule(int &waves, int &max_nt) const;
    // time every schedule on reps graph-replayed steps of the current state,
    // keep the fastest, restore the tracers; rows {waves, max_nt, ms/step}
    std::vector<std::array<double, 3>> autotune(double dt, int reps);

    // per-kernel timing of one step with hip events on the engine stream
    // (names and milliseconds, averaged over reps)
    void time_kernels(double dt, int reps, std::vector<std::string> &names, std::vector<double> &ms);

    struct Impl;  // device buffers (defined in fta_engine.hip)

private:
    Impl *p_;
    int N_, T_, E_, nl_, S_, ntr_, prec_;
    std::vector<TracerSpec> spec_;
    bool refresh_ab_ = true;
    bool graph_valid_ = false;
    double graph_dt_ = 0.0;
    hipGraphExec_t graph_exec_ = nullptr;
    hipGraph_t graph_ = nullptr;
    hipStream_t stream_ = nullptr;
};

// read+write rate (GB/s) of the best of a few launch shapes of a plain
// 16-byte streaming copy of `bytes` (hand-written kernel, fta_stream.hip);
// variant receives {workgroups, words in flight per lane, nontemporal}
double hbm_copy_rate(long long bytes, int reps, int variant[3]);

// this GPU's copy of the allocation-free entry point's device scratch
// (fta_noalloc.hip) and its size in doubles
double *noalloc_scratch(size_t *doubles);

const char *last_error();
void set_error(const std::string &s);

}  // namespace fta
