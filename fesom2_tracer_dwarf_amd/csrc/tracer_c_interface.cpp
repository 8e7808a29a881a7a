// tracer_c_interface.cpp -- the library's C ABI (declared in
// include/fesom_tracer_amd.h). One module-level instance, as the reference
// module keeps its global_mesh/global_tracers/global_dyn singletons
// (lib/tracer_c_interface.F90:28-31); not thread-safe.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/fesom_tracer_amd.h"
#include "fta_engine.h"
#include "fta_hostcopy.h"
#include "fta_mesh.h"
#include "fta_restart.h"

using namespace fta;

namespace {

struct State {
    bool initialized = false;
    int mpi_comm = 0;
    int nl = 0, nod2D = 0, edge2D = 0, elem2D = 0, myDim = 0, eDim = 0;
    HostMesh mesh;
    bool mesh_ready = false;
    int ntr = 0, ab_order = 1;
    std::vector<std::vector<double>> values, valuesAB;
    std::vector<TracerSpec> specs;
    std::vector<double> uv, w, we;
    std::unique_ptr<Engine> eng;
    int precision = 64, refresh = 1, use_graph = 1;
    std::vector<float> mlp;  // learned-limiter parameters carried across engine rebuilds
    int tune_waves = 4, tune_nt = 4;  // launch schedule carried across engine rebuilds
    FesomFileOptions file_opt;         // mesh-file reader settings (force_rotation)
    std::vector<FILE *> frames;        // open scalar-output files, one per tracer
    bool mesh_building = false;        // between tracer_mesh_begin and tracer_mesh_end
    std::vector<std::string> mesh_given;  // arrays set since tracer_mesh_begin
    int fan_fault = -1;                // first node without an element fan (caller-supplied meshes)
    bool accumulate = false;           // del_ttf carried over between steps
    bool grad_fan = true;              // nodal gradients over the node fans
};

State &S()
{
    static State s;
    return s;
}

int fail(const std::string &msg, int code = -1)
{
    set_error(msg);
    std::fprintf(stderr, "### error: %s\n", msg.c_str());
    return code;
}

bool dir_exists(const std::string &p)
{
    struct stat st;
    return ::stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

int prec_code(int p) { return p == 32 ? PREC_FP32 : (p == 16 ? PREC_FP32_H16 : PREC_FP64); }

void drop_engine() { S().eng.reset(); }

// make sure host-side arrays exist with the current mesh dimensions
void ensure_host_arrays()
{
    State &s = S();
    const size_t L = s.mesh.nl - 1;
    if (s.uv.size() != 2 * L * s.mesh.elem2D) s.uv.assign(2 * L * s.mesh.elem2D, 0.0);
    if (s.w.size() != (size_t)s.mesh.nl * s.mesh.nod2D) s.w.assign((size_t)s.mesh.nl * s.mesh.nod2D, 0.0);
    if (s.we.size() != (size_t)s.mesh.nl * s.mesh.nod2D) s.we.assign((size_t)s.mesh.nl * s.mesh.nod2D, 0.0);
}

// create the GPU engine from the host state (mesh, velocity, tracers)
int ensure_engine()
{
    State &s = S();
    if (s.eng) return 0;
    if (!s.mesh_ready) return fail("mesh topology missing: generate or load a mesh first");
    if (s.ntr < 1) return fail("no tracers allocated");
    try {
        ensure_host_arrays();
        s.eng.reset(new Engine(s.mesh, s.ntr, prec_code(s.precision)));
        s.eng->set_velocity(s.uv.data(), s.w.data(), s.we.data());
        s.eng->set_refresh_ab(s.refresh != 0);
        if (!s.mlp.empty()) s.eng->mlp_set(s.mlp.data());
        s.eng->set_schedule(s.tune_waves, s.tune_nt);
        if (s.accumulate) s.eng->set_accumulate(true);
        if (!s.grad_fan) s.eng->set_gradient_fan(false);
        for (int t = 0; t < s.ntr; ++t) {
            s.eng->set_spec(t, s.specs[t]);
            s.eng->set_values(t, s.values[t].data());
            if (!s.refresh) s.eng->set_valuesAB(t, s.valuesAB[t].data());
        }
    } catch (const std::exception &e) {
        s.eng.reset();
        return fail(e.what());
    }
    return 0;
}

// upload the host velocity to a live engine (no-op without one)
int push_velocity()
{
    State &s = S();
    if (!s.eng) return 0;
    try {
        s.eng->set_velocity(s.uv.data(), s.w.data(), s.we.data());
    } catch (const std::exception &e) {
        return fail(e.what());
    }
    return 0;
}

void adopt_mesh_dims()
{
    State &s = S();
    s.nl = s.mesh.nl;
    s.nod2D = s.mesh.nod2D;
    s.edge2D = s.mesh.edge2D;
    s.elem2D = s.mesh.elem2D;
    s.myDim = s.mesh.nod2D;
    s.eDim = 0;
}

// check a QR4C/PPM column needs >= 2 layers (the reference reads outside
// the column otherwise, lib/oce_adv_tra_ver.F90:405-411, :508-510)
int check_vertical_scheme(int ver)
{
    State &s = S();
    if (!s.mesh_ready || (ver != VER_QR4C && ver != VER_PPM)) return 0;
    for (int n = 0; n < s.mesh.nod2D; ++n)
        if (s.mesh.nlevels_nod2D[n] - s.mesh.ulevels_nod2D[n] < 2)
            return fail("QR4C/PPM need at least two layers in every column");
    return 0;
}

// first node whose nod_in_elem2D is not one fan (-1: every node has one)
int first_fan_fault(const HostMesh &m)
{
    std::vector<int> ring, fan;
    for (int n = 0; n < m.nod2D; ++n)
        if (!node_fan(m, n, ring, fan)) return n;
    return -1;
}

// take over a restart's mesh, and with tracers its tracer state; the
// velocity of its dynamics (zero where absent)
int adopt_restart(RestartState &rs, bool tracers)
{
    State &s = S();
    drop_engine();
    s.mesh = std::move(rs.mesh);
    s.fan_fault = first_fan_fault(s.mesh);
    s.mesh_building = false;
    s.mesh_ready = true;
    s.initialized = true;
    adopt_mesh_dims();
    const size_t L = s.mesh.nl - 1, N = s.mesh.nod2D;
    s.ntr = tracers ? (int)rs.tracers.size() : 0;
    s.values.assign(s.ntr, std::vector<double>());
    s.valuesAB.assign(s.ntr, std::vector<double>());
    s.specs.assign(s.ntr, TracerSpec());
    for (int t = 0; t < s.ntr; ++t) {
        s.values[t] = std::move(rs.tracers[t].values);
        s.valuesAB[t] = std::move(rs.tracers[t].valuesAB);
        if (s.values[t].size() != L * N || s.valuesAB[t].size() != L * N) return fail("restart tracer size mismatch");
    }
    s.uv = std::move(rs.uv);
    s.w = std::move(rs.w);
    s.we = std::move(rs.w_e);
    ensure_host_arrays();
    for (int t = 0; t < s.ntr; ++t) {
        const RestartTracer &r = rs.tracers[t];
        if (tracer_set_schemes(t + 1, r.hor.c_str(), r.ver.c_str(), r.lim.c_str(), r.ph, r.pv)) return -1;
    }
    return 0;
}

}  // namespace

extern "C" {

int tracer_init_mpi(int mpi_comm)
{
    S().mpi_comm = mpi_comm;
    return 0;
}

int tracer_finalize_mpi(void) { return 0; }

int tracer_init(const char *restart_dir)
{
    State &s = S();
    const std::string p = std::string(restart_dir ? restart_dir : "") + "/np1";
    if (!dir_exists(p)) return fail("could not find restart directory: " + p);
    RestartState rs;
    try {
        read_restart(p, rs, true, true);
    } catch (const std::exception &e) {
        return fail(std::string("tracer_init: ") + e.what());
    }
    if (rs.tracers.empty()) return fail("tracer_init: the restart holds no tracers");
    if (adopt_restart(rs, true)) return -1;
    std::printf("Tracer initialized from: %s\n", p.c_str());
    return 0;
}

int tracer_advect_step(double dt)
{
    State &s = S();
    if (!s.initialized) return fail("tracer not initialized. Call tracer_init first.");
    if (ensure_engine()) return -1;
    try {
        s.eng->enqueue_step(dt, 0, 1);
        s.eng->synchronize();
    } catch (const std::exception &e) {
        return fail(e.what());
    }
    return 0;
}

int tracer_advect_steps(int nsteps, double dt)
{
    State &s = S();
    if (!s.initialized) return fail("tracer not initialized.");
    if (nsteps < 0) return fail("nsteps must be >= 0");
    if (ensure_engine()) return -1;
    try {
        s.eng->run_steps(nsteps, dt, s.use_graph != 0);
        s.eng->synchronize();
    } catch (const std::exception &e) {
        return fail(e.what());
    }
    return 0;
}

int tracer_do_oce_adv_tra(int tracer_id, double dt, double *dttf_h, double *dttf_v, int nl, int nod2D)
{
    State &s = S();
    if (!s.initialized) return fail("tracer not initialized.");
    if (tracer_id < 1 || tracer_id > s.ntr) return fail("invalid tracer_id");
    if (nl != s.nl || nod2D != s.nod2D) return fail("tracer_do_oce_adv_tra: size mismatch");
    if (!dttf_h || !dttf_v) return fail("tracer_do_oce_adv_tra: null tendency array");
    if (ensure_engine()) return -1;
    try {
        s.eng->tendencies(tracer_id - 1, dt, dttf_h, dttf_v);
    } catch (const std::exception &e) {
        return fail(e.what());
    }
    return 0;
}

int tracer_do_oce_adv_tra_noalloc(int tracer_id, double dt, double *dttf_h, double *dttf_v, int nl, int nod2D)
{
    State &s = S();
    if (!s.initialized) return fail("tracer not initialized.");
    if (tracer_id < 1 || tracer_id > s.ntr) return fail("invalid tracer_id");
    if (nl != s.nl || nod2D != s.nod2D) return fail("tracer_do_oce_adv_tra_noalloc: size mismatch");
    if (!dttf_h || !dttf_v) return fail("tracer_do_oce_adv_tra_noalloc: null tendency array");
    // building the device state allocates: it must exist already
    if (!s.eng) return fail("tracer_do_oce_adv_tra_noalloc: no device state yet (run tracer_init_for_advection, "
                            "a step or tracer_do_oce_adv_tra first)");
    try {
        size_t n = 0;
        double *scratch = noalloc_scratch(&n);
        s.eng->tendencies_device(tracer_id - 1, dt, dttf_h, dttf_v, scratch, n);
    } catch (const std::exception &e) {
        return fail(std::string("tracer_do_oce_adv_tra_noalloc: ") + e.what());
    }
    return 0;
}

int tracer_get_stats(double *tmin, double *tmax, double *tsum)
{
    State &s = S();
    if (!s.initialized || s.ntr < 1) {
        if (tmin) *tmin = 0.0;
        if (tmax) *tmax = 0.0;
        if (tsum) *tsum = 0.0;
        return fail("tracer not initialized.");
    }
    double d[7];
    if (tracer_get_diagnostics(1, d)) return -1;
    *tmin = d[0];
    *tmax = d[1];
    // the reference's sum(values) propagates a NaN (the diagnostics' sum skips them)
    *tsum = d[6] > 0.0 ? std::nan("") : d[2];
    return 0;
}

int tracer_get_size(int *nz, int *nn)
{
    State &s = S();
    if (!s.initialized) {
        *nz = 0;
        *nn = 0;
        return fail("tracer not initialized.");
    }
    *nz = s.nl - 1;
    *nn = s.myDim;
    return 0;
}

int tracer_get_values_id(int tracer_id, double *values, int nz, int nn)
{
    State &s = S();
    if (!s.initialized) return fail("tracer not initialized.");
    if (tracer_id < 1 || tracer_id > s.ntr) return fail("invalid tracer_id");
    if (nz != s.nl - 1 || nn > s.nod2D || nn < 0) return fail("tracer_get_values: size mismatch");
    if (s.eng && nn == s.nod2D) {
        // the whole array: straight from the device into the caller's buffer
        try {
            s.eng->get_values(tracer_id - 1, values);
        } catch (const std::exception &e) {
            return fail(e.what());
        }
        return 0;
    }
    const std::vector<double> *src = &s.values[tracer_id - 1];
    std::vector<double> tmp;
    if (s.eng) {
        tmp.resize((size_t)(s.nl - 1) * s.nod2D);
        try {
            s.eng->get_values(tracer_id - 1, tmp.data());
        } catch (const std::exception &e) {
            return fail(e.what());
        }
        src = &tmp;
    }
    copy_host(values, src->data(), sizeof(double) * (size_t)nz * nn);
    return 0;
}

int tracer_get_values(double *values, int nz, int nn) { return tracer_get_values_id(1, values, nz, nn); }

// ---- mesh and scalar output for the reference's plotting scripts
// (lib/mesh_output.F90; read by visualization/plot_scalars.py:21-80)
namespace {

bool make_dir(const std::string &p) { return ::mkdir(p.c_str(), 0755) == 0 || dir_exists(p); }

bool write_raw(const std::string &path, const void *data, size_t bytes)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(data, 1, bytes, f) == bytes;
    return std::fclose(f) == 0 && ok;
}

std::string frame_name(int t)
{
    // the reference names tracer 1 and 2 (mesh_output.F90:109-110); further
    // tracers get their number
    return t == 0 ? "temperature" : (t == 1 ? "salinity" : "tracer_" + std::to_string(t + 1));
}

}  // namespace

int tracer_write_mesh(const char *output_dir)
{
    State &s = S();
    if (!s.mesh_ready) return fail("tracer_write_mesh: no mesh");
    const std::string d = output_dir ? output_dir : "";
    if (d.empty() || !make_dir(d)) return fail("tracer_write_mesh: cannot create " + d);
    const HostMesh &m = s.mesh;
    {
        std::ofstream f(d + "/mesh_info.txt");
        f << "nod2D=" << m.nod2D << "\nelem2D=" << m.elem2D << "\nnl=" << m.nl << "\n";
        if (!f) return fail("tracer_write_mesh: cannot write mesh_info.txt");
    }
    std::vector<double> xy(m.coord_nod2D.size());
    for (size_t i = 0; i < xy.size(); ++i) xy[i] = m.coord_nod2D[i] * kREarth;  // metres, (2, nod2D)
    std::vector<int32_t> el(m.elem2D_nodes.begin(), m.elem2D_nodes.end());     // 0-based, (3, elem2D)
    if (!write_raw(d + "/coord_nod2D.bin", xy.data(), xy.size() * sizeof(double)) ||
        !write_raw(d + "/elem2D_nodes.bin", el.data(), el.size() * sizeof(int32_t)) ||
        !write_raw(d + "/zbar.bin", m.zbar.data(), m.zbar.size() * sizeof(double)) ||
        !write_raw(d + "/depth.bin", m.depth.data(), m.depth.size() * sizeof(double)))
        return fail("tracer_write_mesh: cannot write the binary files in " + d);
    return 0;
}

int tracer_output_close(void)
{
    State &s = S();
    int bad = 0;
    for (FILE *f : s.frames) bad |= std::fclose(f) != 0;
    s.frames.clear();
    return bad ? fail("tracer_output_close: write error") : 0;
}

int tracer_output_open(const char *output_dir, int nsteps)
{
    State &s = S();
    if (!s.initialized || s.ntr < 1) return fail("tracer_output_open: no tracers");
    tracer_output_close();
    const std::string d = output_dir ? output_dir : "";
    if (d.empty() || !make_dir(d)) return fail("tracer_output_open: cannot create " + d);
    std::ofstream f(d + "/scalar_info.txt");
    f << "nod2D=" << s.nod2D << "\nnz=" << s.nl - 1 << "\nnsteps=" << nsteps << "\nnum_tracers=" << s.ntr
      << "\nwp=8\n";
    for (int t = 0; t < s.ntr; ++t) f << "tracer_" << t + 1 << "=" << frame_name(t) << "\n";
    if (!f) return fail("tracer_output_open: cannot write scalar_info.txt");
    for (int t = 0; t < s.ntr; ++t) {
        FILE *o = std::fopen((d + "/" + frame_name(t) + ".bin").c_str(), "wb");
        if (!o) {
            tracer_output_close();
            return fail("tracer_output_open: cannot create the frame files in " + d);
        }
        s.frames.push_back(o);
    }
    return 0;
}

int tracer_output_step(void)
{
    State &s = S();
    if (s.frames.size() != (size_t)s.ntr) return fail("tracer_output_step: output not open");
    const size_t n = (size_t)(s.nl - 1) * s.nod2D;
    std::vector<double> v(n);
    for (int t = 0; t < s.ntr; ++t) {
        if (tracer_get_values_id(t + 1, v.data(), s.nl - 1, s.nod2D) != 0) return -1;
        if (std::fwrite(v.data(), sizeof(double), n, s.frames[t]) != n) return fail("tracer_output_step: write error");
    }
    return 0;
}

int tracer_finalize(const char *output_dir)
{
    State &s = S();
    if (!s.initialized) return fail("tracer not initialized.");
    if (!s.mesh_ready) return fail("tracer_finalize: no mesh to write");
    const std::string out = output_dir ? output_dir : "";
    const std::string p = out + "/np1";
    try {
        std::vector<RestartTracer> trs(s.ntr);
        for (int t = 0; t < s.ntr; ++t) {
            RestartTracer &r = trs[t];
            r.values = s.values[t];
            if (s.eng) s.eng->get_values(t, r.values.data());
            r.valuesAB = s.refresh ? r.values : s.valuesAB[t];
            const TracerSpec &sp = s.specs[t];
            r.hor = sp.hor == HOR_MUSCL ? "MUSCL" : (sp.hor == HOR_MFCT ? "MFCT" : "UPW1");
            r.ver = sp.ver == VER_QR4C ? "QR4C" : (sp.ver == VER_PPM ? "PPM" : (sp.ver == VER_CDIFF ? "CDIFF" : "UPW1"));
            r.lim = sp.lim == LIM_FCT ? "FCT" : (sp.lim == LIM_LEARNED ? "LEARNED" : "NONE");
            r.ph = sp.ph;
            r.pv = sp.pv;
            r.id = t + 1;
        }
        write_restart(p, out + "/meta.time", s.mesh, trs, s.uv, s.w, s.we);
    } catch (const std::exception &e) {
        return fail(std::string("tracer_finalize: ") + e.what());
    }
    std::printf("Restart files written to: %s\n", p.c_str());
    s.initialized = false;
    return 0;
}

int tracer_restart_file_copy(const char *kind, const char *in_path, const char *out_path)
{
    try {
        const std::string k = kind ? kind : "";
        write_restart_file(out_path ? out_path : "", k, read_restart_file(in_path ? in_path : "", k));
    } catch (const std::exception &e) {
        return fail(std::string("tracer_restart_file_copy: ") + e.what());
    }
    return 0;
}

int tracer_run_workflow(const char *restart_dir, int nsteps, double dt)
{
    int st = tracer_init(restart_dir);
    if (st) return st;
    for (int i = 1; i <= nsteps; ++i) {
        st = tracer_advect_step(dt);
        if (st) return st;
        double tmin, tmax, tsum;
        st = tracer_get_stats(&tmin, &tmax, &tsum);
        if (st) return st;
        std::printf("Step %3d: min/max/sum = %15.6E%15.6E%15.6E\n", i, tmin, tmax, tsum);
    }
    return tracer_finalize(restart_dir);
}

int tracer_cleanup(void)
{
    State &s = S();
    drop_engine();
    s.initialized = false;
    return 0;
}

int tracer_set_mesh_dims(int nl, int nod2D, int edge2D, int elem2D)
{
    State &s = S();
    s.nl = nl;
    s.nod2D = nod2D;
    s.edge2D = edge2D;
    s.elem2D = elem2D;
    return 0;
}

int tracer_set_levels(const int *nlevels, const int *ulevels, int nod2D)
{
    State &s = S();
    if (s.mesh_ready && nod2D == s.mesh.nod2D) {
        // node levels must stay consistent with the element levels (FESOM's
        // nlevels_nod2D = max over adjacent elements); accept only that
        for (int n = 0; n < nod2D; ++n)
            if (nlevels[n] != s.mesh.nlevels_nod2D[n] || ulevels[n] != s.mesh.ulevels_nod2D[n])
                return fail("tracer_set_levels: levels disagree with the mesh's element levels");
    }
    return 0;
}

int tracer_set_thickness(const double *hnode, int nl, int nod2D)
{
    State &s = S();
    if (!s.mesh_ready || nl != s.mesh.nl || nod2D != s.mesh.nod2D) return fail("tracer_set_thickness: mesh mismatch");
    std::memcpy(s.mesh.hnode_new.data(), hnode, sizeof(double) * (size_t)(nl - 1) * nod2D);
    drop_engine();
    return 0;
}

int tracer_set_partition(int myDim_nod2D, int eDim_nod2D)
{
    State &s = S();
    if (eDim_nod2D != 0) return fail("one GPU, one partition: eDim_nod2D must be 0");
    s.myDim = myDim_nod2D;
    s.eDim = eDim_nod2D;
    return 0;
}

int tracer_allocate_tracers(int num_tracers, int nl, int nod2D, int AB_order)
{
    State &s = S();
    if (num_tracers < 1) return fail("num_tracers must be >= 1");
    if (s.mesh_ready && (nl != s.mesh.nl || nod2D != s.mesh.nod2D)) return fail("tracer_allocate_tracers: mesh mismatch");
    s.ntr = num_tracers;
    s.ab_order = AB_order;
    s.values.assign(num_tracers, std::vector<double>((size_t)(nl - 1) * nod2D, 0.0));
    s.valuesAB = s.values;
    s.specs.assign(num_tracers, TracerSpec());
    drop_engine();
    return 0;
}

int tracer_set_values(int tracer_id, const double *values, int nl, int nod2D)
{
    State &s = S();
    if (tracer_id < 1 || tracer_id > s.ntr) {
        return fail("invalid tracer_id " + std::to_string(tracer_id) + ", num_tracers=" + std::to_string(s.ntr));
    }
    const size_t n = (size_t)(nl - 1) * nod2D;
    if (s.values[tracer_id - 1].size() != n) return fail("tracer_set_values: size mismatch");
    copy_host(s.values[tracer_id - 1].data(), values, sizeof(double) * n);
    copy_host(s.valuesAB[tracer_id - 1].data(), values, sizeof(double) * n);
    if (s.eng) {
        try {
            s.eng->set_values(tracer_id - 1, values);
            if (!s.refresh) s.eng->set_valuesAB(tracer_id - 1, values);
        } catch (const std::exception &e) {
            return fail(e.what());
        }
    }
    return 0;
}

int tracer_set_values_ab(int tracer_id, const double *values, int nl, int nod2D)
{
    State &s = S();
    if (tracer_id < 1 || tracer_id > s.ntr) return fail("invalid tracer_id");
    const size_t n = (size_t)(nl - 1) * nod2D;
    if (s.valuesAB[tracer_id - 1].size() != n) return fail("tracer_set_values_ab: size mismatch");
    copy_host(s.valuesAB[tracer_id - 1].data(), values, sizeof(double) * n);
    if (s.eng) {
        try {
            s.eng->set_valuesAB(tracer_id - 1, values);
        } catch (const std::exception &e) {
            return fail(e.what());
        }
    }
    return 0;
}

int tracer_set_velocity_zero(int nl, int nod2D, int edge2D)
{
    State &s = S();
    (void)edge2D;
    if (!s.mesh_ready || nl != s.mesh.nl || nod2D != s.mesh.nod2D) return fail("tracer_set_velocity_zero: mesh mismatch");
    s.uv.assign(2 * (size_t)(nl - 1) * s.mesh.elem2D, 0.0);
    s.w.assign((size_t)nl * nod2D, 0.0);
    s.we.assign((size_t)nl * nod2D, 0.0);
    if (push_velocity()) return -1;
    return 0;
}

int tracer_set_velocity(const double *uv, const double *w, int nl, int nod2D, int edge2D)
{
    State &s = S();
    if (!s.mesh_ready || nl != s.mesh.nl || nod2D != s.mesh.nod2D) return fail("tracer_set_velocity: mesh mismatch");
    if (edge2D < s.mesh.elem2D) return fail("tracer_set_velocity: buffer smaller than vel(2,nl-1,elem2D)");
    const size_t nv = 2 * (size_t)(nl - 1) * s.mesh.elem2D;
    s.uv.assign(uv, uv + nv);
    s.w.assign(w, w + (size_t)nl * nod2D);
    s.we.assign((size_t)nl * nod2D, 0.0);
    if (push_velocity()) return -1;
    return 0;
}

int tracer_set_velocity_elem(const double *uv, const double *w, const double *w_e, int nl, int nod2D, int elem2D)
{
    State &s = S();
    if (!s.mesh_ready || nl != s.mesh.nl || nod2D != s.mesh.nod2D || elem2D != s.mesh.elem2D)
        return fail("tracer_set_velocity_elem: mesh mismatch");
    s.uv.assign(uv, uv + 2 * (size_t)(nl - 1) * elem2D);
    s.w.assign(w, w + (size_t)nl * nod2D);
    s.we.assign(w_e, w_e + (size_t)nl * nod2D);
    if (push_velocity()) return -1;
    return 0;
}

int tracer_make_velocity(double u_bg, double v_bg, double cell_amp, double shear, double cfl, double *dt_out)
{
    State &s = S();
    if (!s.mesh_ready) return fail("mesh topology missing");
    double dt = make_velocity(s.mesh, u_bg, v_bg, cell_amp, shear, cfl, s.uv, s.w);
    s.we = s.w;
    if (dt_out) *dt_out = dt;
    if (push_velocity()) return -1;
    return 0;
}

int tracer_init_complete(void)
{
    S().initialized = true;
    return 0;
}

int tracer_load_mesh_partition(const char *restart_dir)
{
    const std::string p = std::string(restart_dir ? restart_dir : "") + "/np1";
    if (!dir_exists(p)) return fail("could not find mesh partition directory: " + p);
    RestartState rs;
    try {
        read_restart(p, rs, false, true);
    } catch (const std::exception &e) {
        return fail(std::string("tracer_load_mesh_partition: ") + e.what());
    }
    if (adopt_restart(rs, false)) return -1;
    const HostMesh &m = S().mesh;
    std::printf("Loaded mesh partition from: %s\n", p.c_str());
    std::printf("  Mesh: nl=%8d, nod2D=%8d, edge2D=%8d\n", m.nl, m.nod2D, m.edge2D);
    std::printf("  Partition: myDim=%8d, eDim=%8d\n", m.nod2D, 0);
    return 0;
}

int tracer_load_mesh_from_files(const char *mesh_path)
{
    State &s = S();
    std::string p = mesh_path ? mesh_path : "";
    if (!dir_exists(p)) return fail("could not find mesh directory: " + p);
    try {
        drop_engine();
        HostMesh m;
        read_fesom_mesh(m, p, s.file_opt);
        s.mesh = std::move(m);
        s.fan_fault = -1;
        s.mesh_building = false;
    } catch (const std::exception &e) {
        s.mesh_ready = false;
        return fail(std::string("tracer_load_mesh_from_files: ") + e.what());
    }
    s.mesh_ready = true;
    s.initialized = true;
    adopt_mesh_dims();
    s.ntr = 0;
    s.values.clear();
    s.valuesAB.clear();
    s.specs.clear();
    s.uv.clear();
    s.w.clear();
    s.we.clear();
    ensure_host_arrays();
    std::printf("Mesh loaded from files: nl=%8d, nod2D=%8d, edge2D=%8d\n", s.mesh.nl, s.mesh.nod2D, s.mesh.edge2D);
    return 0;
}

int tracer_init_ale_arrays(void)
{
    State &s = S();
    if (!s.initialized) return fail("tracer_init_ale_arrays called before initialization");
    if (!s.mesh_ready) return fail("mesh topology missing");
    HostMesh &m = s.mesh;
    for (int n = 0; n < m.nod2D; ++n)
        for (int k = 0; k < m.nl - 1; ++k) {
            m.hnode[(size_t)n * (m.nl - 1) + k] = m.zbar[k] - m.zbar[k + 1];
            m.hnode_new[(size_t)n * (m.nl - 1) + k] = m.hnode[(size_t)n * (m.nl - 1) + k];
        }
    for (int el = 0; el < m.elem2D; ++el)
        for (int k = 0; k < m.nl - 1; ++k) m.helem[(size_t)el * (m.nl - 1) + k] = m.zbar[k] - m.zbar[k + 1];
    drop_engine();
    return 0;
}

int tracer_init_for_advection(void)
{
    State &s = S();
    if (!s.initialized) return fail("tracer_init_for_advection called before mesh loaded");
    if (s.ntr < 1) tracer_allocate_tracers(1, s.mesh.nl, s.mesh.nod2D, 1);
    return ensure_engine();
}

int tracer_generate_mesh(int nx, int ny, int periodic, double Lx, double Ly, int nl, double dz, double stretch,
                         int lev_lo, int lev_hi, unsigned long long seed, int sfc)
{
    State &s = S();
    try {
        drop_engine();
        HostMesh m;
        if (periodic) generate_periodic(m, nx, ny, Lx, Ly, sfc ? 1 : 0);
        else generate_closed(m, nx, ny, Lx, Ly, sfc ? 1 : 0);
        set_vertical_grid(m, nl, dz, stretch);
        set_levels_random(m, lev_lo, lev_hi, seed);
        if (s.file_opt.device_geometry) derive_all_device(m);
        else derive_all(m);
        s.mesh = std::move(m);
        s.fan_fault = -1;
        s.mesh_building = false;
    } catch (const std::exception &e) {
        s.mesh_ready = false;
        return fail(e.what());
    }
    s.mesh_ready = true;
    s.initialized = true;
    adopt_mesh_dims();
    s.ntr = 0;
    s.values.clear();
    s.valuesAB.clear();
    s.specs.clear();
    s.uv.clear();
    s.w.clear();
    s.we.clear();
    ensure_host_arrays();
    return 0;
}

namespace {
// A caller-supplied mesh array by name: the destination vector and the
// entry count the dimensions fix (names and layouts as tracer_get_mesh_array)
struct Slot { std::vector<int> *i; std::vector<double> *d; long long n; };

bool mesh_slot(HostMesh &m, const std::string &name, Slot &r)
{
    const long long N = m.nod2D, T = m.elem2D, E = m.edge2D, L = m.nl, A = m.maxAdj;
#define SI(NAME, VEC, CNT) if (name == NAME) { r = {&VEC, nullptr, CNT}; return true; }
#define SD(NAME, VEC, CNT) if (name == NAME) { r = {nullptr, &VEC, CNT}; return true; }
    SI("edges", m.edges, 2 * E) SI("edge_tri", m.edge_tri, 2 * E) SI("elem2D_nodes", m.elem2D_nodes, 3 * T)
    SI("nlevels", m.nlevels, T) SI("ulevels", m.ulevels, T) SI("nlevels_nod2D", m.nlevels_nod2D, N)
    SI("ulevels_nod2D", m.ulevels_nod2D, N) SI("nlevels_nod2D_min", m.nlevels_nod2D_min, N)
    SI("ulevels_nod2D_max", m.ulevels_nod2D_max, N) SI("nod_in_elem2D_num", m.nod_in_elem2D_num, N)
    SI("nod_in_elem2D", m.nod_in_elem2D, A * N) SI("edge_up_dn_tri", m.edge_up_dn_tri, 2 * E)
    SI("nboundary_lay", m.nboundary_lay, N)
    SD("coord_nod2D", m.coord_nod2D, 2 * N) SD("zbar", m.zbar, L) SD("Z", m.Z, L - 1) SD("depth", m.depth, N)
    SD("elem_cos", m.elem_cos, T) SD("elem_area", m.elem_area, T) SD("gradient_sca", m.gradient_sca, 6 * T)
    SD("edge_dxdy", m.edge_dxdy, 2 * E) SD("edge_cross_dxdy", m.edge_cross_dxdy, 4 * E)
    SD("area", m.area, L * N) SD("areasvol", m.areasvol, L * N) SD("hnode", m.hnode, (L - 1) * N)
    SD("hnode_new", m.hnode_new, (L - 1) * N) SD("helem", m.helem, (L - 1) * T) SD("Z_3d_n", m.Z_3d_n, (L - 1) * N)
    SD("zbar_3d_n", m.zbar_3d_n, L * N)
#undef SI
#undef SD
    return false;
}

bool given(const std::string &name)
{
    const auto &g = S().mesh_given;
    return std::find(g.begin(), g.end(), name) != g.end();
}
}  // namespace

int tracer_mesh_begin(int nl, int nod2D, int elem2D, int edge2D, int maxAdj)
{
    State &s = S();
    if (nl < 3 || nod2D < 1 || elem2D < 1 || edge2D < 1 || maxAdj < 1)
        return fail("tracer_mesh_begin: dimensions must be positive and nl >= 3");
    drop_engine();
    HostMesh m;
    m.nl = nl;
    m.nod2D = nod2D;
    m.elem2D = elem2D;
    m.edge2D = edge2D;
    m.maxAdj = maxAdj;
    s.mesh = std::move(m);
    s.mesh_ready = false;
    s.mesh_building = true;
    s.mesh_given.clear();
    s.fan_fault = -1;
    return 0;
}

int tracer_set_mesh_array(const char *name, const void *data, long long count)
{
    State &s = S();
    const std::string n = name ? name : "";
    if (!s.mesh_building) return fail("tracer_set_mesh_array: call tracer_mesh_begin first");
    // a refused array ends the mesh: tracer_mesh_begin starts over
    Slot r;
    if (!mesh_slot(s.mesh, n, r)) {
        s.mesh_building = false;
        return fail("tracer_set_mesh_array: unknown mesh array " + n);
    }
    if (count != r.n || !data) {
        s.mesh_building = false;
        return fail("tracer_set_mesh_array: " + n + " takes " + std::to_string(r.n) + " entries, got " +
                    (data ? std::to_string(count) : std::string("none")));
    }
    if (r.i) r.i->assign((const int *)data, (const int *)data + count);
    else r.d->assign((const double *)data, (const double *)data + count);
    if (!given(n)) s.mesh_given.push_back(n);
    return 0;
}

int tracer_mesh_end(void)
{
    State &s = S();
    if (!s.mesh_building) return fail("tracer_mesh_end: call tracer_mesh_begin first");
    HostMesh &m = s.mesh;
    try {
        complete_mesh(m, [](const std::string &n) { return given(n); });
    } catch (const std::exception &e) {
        s.mesh_building = false;
        return fail(std::string("tracer_mesh_end: ") + e.what());
    }
    const int N = m.nod2D;
    std::vector<int> ring, fan;
    s.fan_fault = -1;
    for (int n = 0; n < N && s.fan_fault < 0; ++n)
        if (!node_fan(m, n, ring, fan)) s.fan_fault = n;
    s.mesh_building = false;
    s.mesh_given.clear();
    s.mesh_ready = true;
    s.initialized = true;
    adopt_mesh_dims();
    s.ntr = 0;
    s.values.clear();
    s.valuesAB.clear();
    s.specs.clear();
    s.uv.clear();
    s.w.clear();
    s.we.clear();
    ensure_host_arrays();
    return 0;
}

int tracer_set_schemes(int tracer_id, const char *hor, const char *ver, const char *lim, double ph, double pv)
{
    State &s = S();
    if (tracer_id < 1 || tracer_id > s.ntr) return fail("invalid tracer_id");
    TracerSpec sp;
    std::string h = hor ? hor : "", v = ver ? ver : "", l = lim ? lim : "";
    if (h == "UPW1") sp.hor = HOR_UPW1;
    else if (h == "MUSCL") sp.hor = HOR_MUSCL;
    else if (h == "MFCT") sp.hor = HOR_MFCT;
    else return fail("Unknown horizontal advection type " + h);
    if (v == "UPW1") sp.ver = VER_UPW1;
    else if (v == "QR4C") sp.ver = VER_QR4C;
    else if (v == "PPM") sp.ver = VER_PPM;
    else if (v == "CDIFF") sp.ver = VER_CDIFF;
    else return fail("Unknown vertical advection type " + v);
    if (l == "NONE") sp.lim = LIM_NONE;
    else if (l == "FCT") sp.lim = LIM_FCT;
    else if (l == "LEARNED") sp.lim = LIM_LEARNED;
    else return fail("Unknown limiter type " + l);
    if (check_vertical_scheme(sp.ver)) return -1;
    sp.ph = ph;
    sp.pv = pv;
    if ((sp.lim == LIM_FCT || sp.lim == LIM_LEARNED) && s.fan_fault >= 0)
        return fail("FCT and LEARNED need the elements around every node to form one fan; node " +
                    std::to_string(s.fan_fault + 1) + " does not");
    s.specs[tracer_id - 1] = sp;
    if (s.eng) {
        try {
            s.eng->set_spec(tracer_id - 1, sp);
        } catch (const std::exception &e) {
            return fail(e.what());
        }
    }
    return 0;
}

int tracer_set_option(const char *name, int value)
{
    State &s = S();
    std::string n = name ? name : "";
    if (n == "force_rotation") {  // mesh-file reader: rotate nod2d.out coordinates (default 1)
        s.file_opt.force_rotation = value != 0;
        return 0;
    }
    if (n == "device_geometry") {  // mesh geometry and up/down-wind search as GPU kernels (default 0)
        s.file_opt.device_geometry = value != 0;
        return 0;
    }
    if (n == "precision") {
        if (value != 64 && value != 32 && value != 16) return fail("precision must be 64, 32 or 16");
        if (s.eng) {
            // carry the current tracer state over to the new engine
            try {
                for (int t = 0; t < s.ntr; ++t) s.eng->get_values(t, s.values[t].data());
            } catch (const std::exception &e) {
                return fail(e.what());
            }
        }
        s.precision = value;
        drop_engine();
        return 0;
    }
    if (n == "refresh_valuesAB") {
        s.refresh = value ? 1 : 0;
        if (s.eng) {
            try {
                if (!value)
                    for (int t = 0; t < s.ntr; ++t) s.eng->set_valuesAB(t, s.valuesAB[t].data());
                s.eng->set_refresh_ab(value != 0);
            } catch (const std::exception &e) {
                return fail(e.what());
            }
        }
        return 0;
    }
    if (n == "block_waves" || n == "batch_tracers") {
        const int w = n == "block_waves" ? value : s.tune_waves, b = n == "batch_tracers" ? value : s.tune_nt;
        if (w != 1 && w != 2 && w != 4 && w != 8) return fail("block_waves must be 1, 2, 4 or 8");
        if (b != 1 && b != 2 && b != 4) return fail("batch_tracers must be 1, 2 or 4");
        s.tune_waves = w;
        s.tune_nt = b;
        if (s.eng) s.eng->set_schedule(w, b);
        return 0;
    }
    if (n == "use_graph") {
        s.use_graph = value ? 1 : 0;
        return 0;
    }
    if (n == "gradient_fan") {  // nodal mean gradients over the node fans (1, default) or element by element
        s.grad_fan = value != 0;
        if (s.eng) s.eng->set_gradient_fan(s.grad_fan);
        return 0;
    }
    if (n == "accumulate_del_ttf") {  // del_ttf carried over between steps (reference C interface, Q6)
        s.accumulate = value != 0;
        if (s.eng) {
            try {
                s.eng->set_accumulate(s.accumulate);
            } catch (const std::exception &e) {
                return fail(e.what());
            }
        }
        return 0;
    }
    return fail("unknown option " + n);
}

int tracer_get_diagnostics(int tracer_id, double *out)
{
    State &s = S();
    if (!s.initialized) return fail("tracer not initialized.");
    if (tracer_id < 1 || tracer_id > s.ntr) return fail("invalid tracer_id");
    if (ensure_engine()) return -1;
    try {
        Diag d = s.eng->diagnostics(tracer_id - 1);
        out[0] = d.vmin; out[1] = d.vmax; out[2] = d.sum;
        out[3] = d.amin; out[4] = d.amax; out[5] = d.mass; out[6] = d.nan_count;
    } catch (const std::exception &e) {
        return fail(e.what());
    }
    return 0;
}

namespace {
struct ArrRef { const void *p; long long n; int kind; };

bool mesh_array(const std::string &name, ArrRef &r)
{
    State &s = S();
    HostMesh &m = s.mesh;
    static std::vector<int> dims;
    if (name == "dims") {
        dims = {m.nl, m.nod2D, m.elem2D, m.edge2D, m.maxAdj};
        r = {dims.data(), (long long)dims.size(), 0};
        return true;
    }
    if (name == "uv") { r = {s.uv.data(), (long long)s.uv.size(), 1}; return true; }
    if (name == "w") { r = {s.w.data(), (long long)s.w.size(), 1}; return true; }
    if (name == "w_e") { r = {s.we.data(), (long long)s.we.size(), 1}; return true; }
    return named_mesh_array(m, name, r.p, r.n, r.kind);
}
}  // namespace

int tracer_mesh_array_info(const char *name, long long *count, int *kind)
{
    if (!S().mesh_ready) return fail("mesh topology missing");
    ArrRef r;
    if (!mesh_array(name ? name : "", r)) return fail(std::string("unknown mesh array ") + (name ? name : ""));
    *count = r.n;
    *kind = r.kind;
    return 0;
}

int tracer_get_mesh_array(const char *name, void *out, long long count)
{
    if (!S().mesh_ready) return fail("mesh topology missing");
    ArrRef r;
    if (!mesh_array(name ? name : "", r)) return fail(std::string("unknown mesh array ") + (name ? name : ""));
    if (count < r.n) return fail("tracer_get_mesh_array: buffer too small");
    std::memcpy(out, r.p, (size_t)r.n * (r.kind ? sizeof(double) : sizeof(int)));
    return 0;
}

int tracer_get_work_array(int tracer_id, const char *name, double *out, long long count)
{
    State &s = S();
    if (!s.eng) return fail("no device state yet");
    if (tracer_id < 1 || tracer_id > s.ntr) return fail("invalid tracer_id");
    try {
        if (s.eng->get_work(name ? name : "", tracer_id - 1, out, (size_t)count))
            return fail("unknown work array or short buffer");
    } catch (const std::exception &e) {
        return fail(e.what());
    }
    return 0;
}

int tracer_time_steps(int nsteps, double dt, double *ms_out)
{
    State &s = S();
    if (!s.initialized) return fail("tracer not initialized.");
    if (ensure_engine()) return -1;
    try {
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return fail("hipEventCreate failed");
        if (s.use_graph) s.eng->run_steps(0, dt, true);  // instantiate outside the timed region
        hipEventRecord(a, s.eng->stream());
        s.eng->run_steps(nsteps, dt, s.use_graph != 0);
        hipEventRecord(b, s.eng->stream());
        hipEventSynchronize(b);
        float ms = 0.0f;
        hipEventElapsedTime(&ms, a, b);
        hipEventDestroy(a);
        hipEventDestroy(b);
        *ms_out = ms;
    } catch (const std::exception &e) {
        return fail(e.what());
    }
    return 0;
}

int tracer_time_kernels(double dt, int reps, char *names, int names_len, double *ms, int max_kernels)
{
    State &s = S();
    if (!s.initialized) return fail("tracer not initialized.");
    if (ensure_engine()) return -1;
    try {
        std::vector<std::string> nm;
        std::vector<double> t;
        s.eng->time_kernels(dt, reps, nm, t);
        std::string joined;
        for (size_t i = 0; i < nm.size(); ++i) {
            if (i) joined += ";";
            joined += nm[i];
            if ((int)i < max_kernels) ms[i] = t[i];
        }
        std::snprintf(names, (size_t)names_len, "%s", joined.c_str());
        return (int)nm.size();
    } catch (const std::exception &e) {
        return fail(e.what());
    }
}

int tracer_precision_error(int ref_precision, int nsteps, double dt, double *out)
{
    State &s = S();
    if (!s.initialized) return fail("tracer not initialized.");
    if (ref_precision != 64 && ref_precision != 32 && ref_precision != 16)
        return fail("precision must be 64, 32 or 16");
    if (nsteps < 0) return fail("nsteps must be >= 0");
    if (ensure_engine()) return -1;
    try {
        // a second engine at the reference precision, started from the current state
        Engine ref(s.mesh, s.ntr, prec_code(ref_precision));
        ref.set_velocity(s.uv.data(), s.w.data(), s.we.data());
        ref.set_refresh_ab(s.refresh != 0);
        std::vector<double> v((size_t)(s.mesh.nl - 1) * s.mesh.nod2D);
        for (int t = 0; t < s.ntr; ++t) {
            ref.set_spec(t, s.specs[t]);
            s.eng->get_values(t, v.data());
            ref.set_values(t, v.data());
            if (!s.refresh) ref.set_valuesAB(t, s.valuesAB[t].data());
        }
        for (int k = 0; k < nsteps; ++k) {
            s.eng->run_steps(1, dt, s.use_graph != 0);
            ref.run_steps(1, dt, s.use_graph != 0);
            for (int t = 0; t < s.ntr; ++t) s.eng->difference(ref, t, out + 4 * ((size_t)k * s.ntr + t));
        }
        s.eng->synchronize();
    } catch (const std::exception &e) {
        return fail(e.what());
    }
    return 0;
}

int tracer_mlp_param_count(void) { return Engine::mlp_param_count(); }

int tracer_mlp_set_params(const float *params, int n)
{
    State &s = S();
    if (n != Engine::mlp_param_count()) return fail("learned limiter needs " +
                                                    std::to_string(Engine::mlp_param_count()) + " parameters");
    s.mlp.assign(params, params + n);
    if (s.eng) {
        try { s.eng->mlp_set(params); } catch (const std::exception &e) { return fail(e.what()); }
    }
    return 0;
}

int tracer_mlp_get_params(float *params, int n)
{
    State &s = S();
    if (n != Engine::mlp_param_count()) return fail("learned limiter has " +
                                                    std::to_string(Engine::mlp_param_count()) + " parameters");
    if (!s.initialized) return fail("tracer not initialized.");
    if (ensure_engine()) return -1;
    try { s.eng->mlp_get(params); } catch (const std::exception &e) { return fail(e.what()); }
    return 0;
}

int tracer_learn_collect(int nsteps, double dt, long long max_rows, long long *rows_out)
{
    State &s = S();
    if (!s.initialized) return fail("tracer not initialized.");
    if (nsteps < 0 || max_rows < 0) return fail("nsteps and max_rows must be >= 0");
    if (ensure_engine()) return -1;
    try { *rows_out = s.eng->learn_collect(nsteps, dt, max_rows); } catch (const std::exception &e) { return fail(e.what()); }
    return 0;
}

int tracer_learn_clear(void)
{
    State &s = S();
    if (s.eng) s.eng->learn_clear();
    return 0;
}

int tracer_learn_train(int epochs, long long batch, double lr, unsigned seed, double *losses)
{
    State &s = S();
    if (!s.initialized) return fail("tracer not initialized.");
    if (ensure_engine()) return -1;
    try {
        std::vector<double> l = s.eng->learn_train(epochs, batch, lr, seed);
        for (size_t i = 0; i < l.size(); ++i) losses[i] = l[i];
        s.mlp.assign(Engine::mlp_param_count(), 0.0f);
        s.eng->mlp_get(s.mlp.data());
    } catch (const std::exception &e) {
        return fail(e.what());
    }
    return 0;
}

int tracer_mlp_forward(const float *X, long long rows, float *out)
{
    State &s = S();
    if (!s.initialized) return fail("tracer not initialized.");
    if (ensure_engine()) return -1;
    try { s.eng->mlp_forward(X, rows, out); } catch (const std::exception &e) { return fail(e.what()); }
    return 0;
}

int tracer_mlp_gradient(const float *X, const float *y, long long rows, float scale, float *grad, double *loss)
{
    State &s = S();
    if (!s.initialized) return fail("tracer not initialized.");
    if (ensure_engine()) return -1;
    try { s.eng->mlp_gradient(X, y, rows, scale, grad, loss); } catch (const std::exception &e) { return fail(e.what()); }
    return 0;
}

int tracer_learn_fallback(int tracer_id, double *fraction)
{
    State &s = S();
    if (!s.initialized) return fail("tracer not initialized.");
    if (tracer_id < 1 || tracer_id > s.ntr) return fail("invalid tracer_id");
    if (ensure_engine()) return -1;
    try { *fraction = s.eng->learn_fallback(tracer_id - 1); } catch (const std::exception &e) { return fail(e.what()); }
    return 0;
}

int tracer_autotune(double dt, int reps, double *table, int max_rows)
{
    State &s = S();
    if (!s.initialized) return fail("tracer not initialized.");
    if (ensure_engine()) return -1;
    try {
        auto t = s.eng->autotune(dt, reps);
        for (size_t i = 0; i < t.size() && (int)i < max_rows; ++i)
            for (int j = 0; j < 3; ++j) table[3 * i + j] = t[i][j];
        s.eng->schedule(s.tune_waves, s.tune_nt);
        return (int)t.size();
    } catch (const std::exception &e) {
        return fail(e.what());
    }
}

int tracer_hbm_copy_rate(long long bytes, int reps, double *gbs, int *variant)
{
    try {
        int v[3] = {0, 0, 0};
        *gbs = hbm_copy_rate(bytes, reps, v);
        if (variant) for (int i = 0; i < 3; ++i) variant[i] = v[i];
    } catch (const std::exception &e) {
        return fail(std::string("tracer_hbm_copy_rate: ") + e.what());
    }
    return 0;
}

const char *tracer_last_error(void) { return last_error(); }

}  // extern "C"
