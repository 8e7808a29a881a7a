// fta_restart.cpp -- FESOM derived-type binary restarts on the host (format
// described in fta_restart.h). The item tables below restate the order of
// the reference's user-defined I/O procedures; the codec walks a table
// either reading into a name -> item map or writing from one.
#include "fta_restart.h"

#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <stdexcept>

namespace fta {

namespace {

constexpr size_t kMaxSubrecord = 2147483639;  // gfortran's largest subrecord

std::runtime_error err(const std::string &m) { return std::runtime_error(m); }

// ---- one Fortran sequential unformatted record per file -------------------

std::vector<unsigned char> read_record(const std::string &path)
{
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) throw err("cannot open " + path);
    std::vector<unsigned char> out;
    auto fail = [&](const std::string &m) {
        std::fclose(f);
        return err(path + ": " + m);
    };
    for (;;) {
        int32_t lead, trail;
        if (std::fread(&lead, 4, 1, f) != 1) throw fail("truncated record marker");
        const uint64_t n = lead < 0 ? (uint64_t)(-(int64_t)lead) : (uint64_t)lead;
        const size_t old = out.size();
        out.resize(old + n);
        if (n && std::fread(out.data() + old, 1, n, f) != n) throw fail("truncated record");
        if (std::fread(&trail, 4, 1, f) != 1) throw fail("truncated record marker");
        const uint64_t m = trail < 0 ? (uint64_t)(-(int64_t)trail) : (uint64_t)trail;
        if (m != n) throw fail("leading and trailing record markers disagree");
        if (lead >= 0) break;
    }
    if (std::fgetc(f) != EOF) throw fail("more than one record");
    std::fclose(f);
    return out;
}

// leading marker negative when another subrecord follows, trailing marker
// negative on every subrecord after the first
void write_record(const std::string &path, const std::vector<unsigned char> &payload)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) throw err("cannot create " + path + ": " + std::strerror(errno));
    size_t off = 0;
    bool first = true, ok = true;
    do {
        const size_t n = std::min(kMaxSubrecord, payload.size() - off);
        const bool more = off + n < payload.size();
        const int32_t lead = more ? -(int32_t)n : (int32_t)n;
        const int32_t trail = first ? (int32_t)n : -(int32_t)n;
        ok = ok && std::fwrite(&lead, 4, 1, f) == 1;
        ok = ok && (n == 0 || std::fwrite(payload.data() + off, 1, n, f) == n);
        ok = ok && std::fwrite(&trail, 4, 1, f) == 1;
        off += n;
        first = false;
    } while (off < payload.size());
    if (std::fclose(f) != 0 || !ok) throw err("write failed: " + path);
}

// ---- the item codec ---------------------------------------------------------
// kinds: 'i' integer, 'r' real(8), 'l' logical, 'c' character(20),
//        'I' / 'R' allocatable integer / real array of the given rank,
//        'S' integer array of fixed length (extent always written)
class Codec {
public:
    Codec(RRecord &rec, const std::vector<unsigned char> &in, const std::string &what)
        : rec_(rec), in_(&in), what_(what) {}
    Codec(RRecord &rec, const std::string &what) : rec_(rec), what_(what) {}

    void item(const std::string &name, char kind, int rank = 0)
    {
        if (in_) get(name, kind, rank);
        else put(name, kind, rank);
    }
    int integer(const std::string &name) const
    {
        auto it = rec_.find(name);
        return it == rec_.end() || it->second.i.empty() ? 0 : it->second.i[0];
    }
    bool done() const { return !in_ || pos_ == in_->size(); }
    size_t pos() const { return pos_; }
    std::vector<unsigned char> out;

private:
    RRecord &rec_;
    const std::vector<unsigned char> *in_ = nullptr;
    std::string what_;
    size_t pos_ = 0;

    void need(size_t n, const std::string &name)
    {
        if (n > in_->size() - pos_)
            throw err(what_ + ": record ends inside item " + name + " (a different build or layout)");
    }
    template <class T>
    T take(const std::string &name)
    {
        need(sizeof(T), name);
        T v;
        std::memcpy(&v, in_->data() + pos_, sizeof(T));
        pos_ += sizeof(T);
        return v;
    }
    template <class T>
    void emit(const T &v)
    {
        const unsigned char *p = reinterpret_cast<const unsigned char *>(&v);
        out.insert(out.end(), p, p + sizeof(T));
    }

    void get(const std::string &name, char kind, int rank)
    {
        RItem it;
        switch (kind) {
        case 'i': it.i = {take<int32_t>(name)}; break;
        case 'l': it.i = {take<int32_t>(name) != 0 ? 1 : 0}; break;
        case 'r': it.r = {take<double>(name)}; break;
        case 'c': {
            need(20, name);
            std::string s(reinterpret_cast<const char *>(in_->data() + pos_), 20);
            pos_ += 20;
            s.erase(s.find_last_not_of(' ') + 1);
            it.s = s;
            break;
        }
        case 'S': {
            const int32_t n = take<int32_t>(name);
            if (n < 0) throw err(what_ + ": negative extent of " + name);
            need(4 * (size_t)n, name);
            it.dims = {n};
            it.i.resize(n);
            std::memcpy(it.i.data(), in_->data() + pos_, 4 * (size_t)n);
            pos_ += 4 * (size_t)n;
            break;
        }
        default: {  // 'I' / 'R'
            uint64_t total = 1;
            bool empty = false;
            for (int d = 0; d < rank; ++d) {
                const int32_t e = take<int32_t>(name);
                if (e < 0) throw err(what_ + ": negative extent of " + name);
                it.dims.push_back(e);
                empty = empty || e == 0;
                total *= (uint64_t)e;
                if (total > in_->size()) throw err(what_ + ": extents of " + name + " exceed the record");
            }
            if (empty) break;  // unallocated: the reader skips the data
            const size_t w = kind == 'I' ? 4 : 8;
            need(total * w, name);
            if (kind == 'I') {
                it.i.resize(total);
                std::memcpy(it.i.data(), in_->data() + pos_, total * w);
            } else {
                it.r.resize(total);
                std::memcpy(it.r.data(), in_->data() + pos_, total * w);
            }
            pos_ += total * w;
        }
        }
        rec_[name] = std::move(it);
    }

    void put(const std::string &name, char kind, int rank)
    {
        auto f = rec_.find(name);
        const RItem *it = f == rec_.end() ? nullptr : &f->second;
        switch (kind) {
        case 'i':
        case 'l': emit<int32_t>(it && !it->i.empty() ? it->i[0] : 0); break;
        case 'r': emit<double>(it && !it->r.empty() ? it->r[0] : 0.0); break;
        case 'c': {
            std::string s = it ? it->s : std::string();
            s.resize(20, ' ');
            out.insert(out.end(), s.begin(), s.end());
            break;
        }
        case 'S': {
            std::vector<int> v = it ? it->i : std::vector<int>();
            v.resize((size_t)rank, 0);  // fixed length
            emit<int32_t>(rank);
            for (int x : v) emit<int32_t>(x);
            break;
        }
        default: {
            if (!it || (int)it->dims.size() != rank) {  // unallocated
                for (int d = 0; d < rank; ++d) emit<int32_t>(0);
                break;
            }
            const size_t n = kind == 'I' ? it->i.size() : it->r.size();
            uint64_t total = 1;
            for (int e : it->dims) total *= (uint64_t)e;
            if (total != n) throw err(what_ + ": extents of " + name + " do not match its data");
            for (int e : it->dims) emit<int32_t>(e);
            const unsigned char *p = kind == 'I' ? reinterpret_cast<const unsigned char *>(it->i.data())
                                                 : reinterpret_cast<const unsigned char *>(it->r.data());
            out.insert(out.end(), p, p + n * (kind == 'I' ? 4 : 8));
        }
        }
    }
};

// ---- item tables ------------------------------------------------------------

struct Spec {
    const char *name;
    char kind;
    int rank;
};

template <size_t K>
void run(Codec &c, const Spec (&t)[K])
{
    for (const Spec &x : t) c.item(x.name, x.kind, x.rank);
}

// write_t_mesh (lib/MOD_MESH.F90:185-279)
void mesh_items(Codec &c)
{
    c.item("nod2D", 'i');
    c.item("ocean_area", 'r');
    c.item("ocean_areawithcav", 'r');
    c.item("edge2D", 'i');
    c.item("edge2D_in", 'i');
    c.item("elem2D", 'i');
    static const Spec a1[] = {
        {"elem2D_nodes", 'I', 2}, {"edges", 'I', 2}, {"edge_tri", 'I', 2},
        {"elem_edges", 'I', 2}, {"elem_area", 'R', 1}, {"edge_dxdy", 'R', 2},
        {"edge_cross_dxdy", 'R', 2}, {"elem_cos", 'R', 1}, {"metric_factor", 'R', 1},
        {"elem_neighbors", 'I', 2}, {"nod_in_elem2D", 'I', 2}, {"x_corners", 'R', 2},
        {"y_corners", 'R', 2}, {"nod_in_elem2D_num", 'I', 1}, {"depth", 'R', 1},
        {"gradient_vec", 'R', 2}, {"gradient_sca", 'R', 2}, {"bc_index_nod2D", 'I', 1}};
    run(c, a1);
    c.item("nl", 'i');
    static const Spec a2[] = {
        {"zbar", 'R', 1}, {"Z", 'R', 1}, {"elem_depth", 'R', 1}, {"ulevels", 'I', 1},
        {"ulevels_nod2D", 'I', 1}, {"ulevels_nod2D_max", 'I', 1}, {"nlevels", 'I', 1},
        {"nlevels_nod2D", 'I', 1}, {"nlevels_nod2D_min", 'I', 1}, {"area", 'R', 2},
        {"area_inv", 'R', 2}, {"areasvol", 'R', 2}, {"areasvol_inv", 'R', 2},
        {"mesh_resolution", 'R', 1}, {"cavity_flag_n", 'I', 1}, {"cavity_flag_e", 'I', 1},
        {"cavity_depth", 'R', 1}, {"cavity_nrst_cavlpnt_xyz", 'R', 2}};
    run(c, a2);
    c.item("ssh_stiff%dim", 'i');
    c.item("ssh_stiff%nza", 'i');
    static const Spec a3[] = {
        {"ssh_stiff%rowptr", 'I', 1}, {"ssh_stiff%colind", 'I', 1}, {"ssh_stiff%values", 'R', 1},
        {"ssh_stiff%colind_loc", 'I', 1}, {"ssh_stiff%rowptr_loc", 'I', 1},
        {"lump2d_south", 'R', 1}, {"lump2d_north", 'R', 1}, {"ind_south", 'I', 1},
        {"ind_north", 'I', 1}};
    run(c, a3);
    c.item("nn_size", 'i');
    static const Spec a4[] = {
        {"nn_num", 'I', 1}, {"nn_pos", 'I', 2}, {"hnode", 'R', 2}, {"hnode_new", 'R', 2},
        {"zbar_3d_n", 'R', 2}, {"Z_3d_n", 'R', 2}, {"Z_3d_n_ib", 'R', 2}, {"helem", 'R', 2},
        {"bottom_elem_thickness", 'R', 1}, {"bottom_node_thickness", 'R', 1}, {"dhe", 'R', 1},
        {"hbar", 'R', 1}, {"hbar_old", 'R', 1}, {"zbar_n_bot", 'R', 1},
        {"zbar_e_bot", 'R', 1}, {"zbar_n_srf", 'R', 1}, {"zbar_e_srf", 'R', 1},
        {"coriolis", 'R', 1}, {"coriolis_node", 'R', 1}};
    run(c, a4);
}

// WRITE_T_COM_STRUCT (lib/MOD_PARTIT.F90:114-132); fixed lengths from
// MAX_NEIGHBOR_PARTITIONS = 32 (rptr has one more)
void com_items(Codec &c, const std::string &p)
{
    c.item(p + "rPEnum", 'i');
    c.item(p + "rPE", 'S', 32);
    c.item(p + "rptr", 'S', 33);
    c.item(p + "rlist", 'I', 1);
    c.item(p + "sPEnum", 'i');
    c.item(p + "sPE", 'S', 32);
    c.item(p + "sptr", 'S', 32);
    c.item(p + "slist", 'I', 1);
    c.item(p + "nreq", 'i');
}

// WRITE_T_PARTIT (lib/MOD_PARTIT.F90:155-182)
void partit_items(Codec &c)
{
    com_items(c, "com_nod2D%");
    com_items(c, "com_elem2D%");
    com_items(c, "com_elem2D_full%");
    c.item("npes", 'i');
    c.item("mype", 'i');
    c.item("maxPEnum", 'i');
    c.item("part", 'I', 1);
    c.item("myDim_nod2D", 'i');
    c.item("eDim_nod2D", 'i');
    c.item("myList_nod2D", 'I', 1);
    c.item("myDim_elem2D", 'i');
    c.item("eDim_elem2D", 'i');
    c.item("eXDim_elem2D", 'i');
    c.item("myList_elem2D", 'I', 1);
    c.item("myDim_edge2D", 'i');
    c.item("eDim_edge2D", 'i');
    c.item("myList_edge2D", 'I', 1);
    c.item("pe_status", 'i');
}

std::string tracer_prefix(int k) { return "data(" + std::to_string(k) + ")%"; }

// WRITE_T_TRACER, _DATA, _WORK (lib/MOD_TRACER.F90:104-239), double-precision build
void tracer_items(Codec &c)
{
    c.item("num_tracers", 'i');
    const int n = c.integer("num_tracers");
    if (n < 0 || n > 100000) throw err("t_tracer: implausible num_tracers " + std::to_string(n));
    for (int k = 1; k <= n; ++k) {
        const std::string p = tracer_prefix(k);
        c.item(p + "values", 'R', 2);
        c.item(p + "valuesold", 'R', 3);
        c.item(p + "valuesAB", 'R', 2);
        c.item(p + "smooth_bh_tra", 'l');
        c.item(p + "gamma0_tra", 'r');
        c.item(p + "gamma1_tra", 'r');
        c.item(p + "gamma2_tra", 'r');
        c.item(p + "i_vert_diff", 'l');
        c.item(p + "tra_adv_hor", 'c');
        c.item(p + "tra_adv_ver", 'c');
        c.item(p + "tra_adv_lim", 'c');
        c.item(p + "tra_adv_ph", 'r');
        c.item(p + "tra_adv_pv", 'r');
        c.item(p + "ID", 'i');
    }
    static const std::pair<const char *, int> w[] = {
        {"del_ttf", 2}, {"del_ttf_advhoriz", 2}, {"del_ttf_advvert", 2}, {"dvd_trflx_hor", 3},
        {"dvd_trflx_ver", 3}, {"fct_LO", 2}, {"adv_flux_hor", 2}, {"adv_flux_ver", 2}, {"fct_ttf_max", 2},
        {"fct_ttf_min", 2}, {"fct_plus", 2}, {"fct_minus", 2}, {"edge_up_dn_grad", 3}};
    for (auto &q : w) c.item(std::string("work%") + q.first, 'R', q.second);
    c.item("work%nboundary_lay", 'I', 1);
    c.item("work%edge_up_dn_tri", 'I', 2);
}

// WRITE_T_DYN, _SOLVERINFO, _DYN_WORK (lib/MOD_DYN.F90:183-290), with
// Fer_GM off (its default, lib/oce_modules.F90:46)
void dyn_items(Codec &c)
{
    c.item("opt_visc", 'i');
    c.item("visc_gamma0", 'r');
    c.item("visc_gamma1", 'r');
    c.item("visc_gamma2", 'r');
    c.item("visc_easybsreturn", 'r');
    c.item("use_ivertvisc", 'l');
    c.item("momadv_opt", 'i');
    c.item("use_freeslip", 'l');
    c.item("use_wsplit", 'l');
    c.item("wsplit_maxcfl", 'r');
    c.item("use_ssh_se_subcycl", 'l');
    for (const char *s : {"ident", "maxiter", "restart", "fillin", "lutype"}) c.item(std::string("solverinfo%") + s, 'i');
    c.item("solverinfo%droptol", 'r');
    c.item("solverinfo%soltol", 'r');
    for (const char *s : {"rr", "zz", "pp", "App"}) c.item(std::string("solverinfo%") + s, 'R', 1);
    c.item("work%uvnode_rhs", 'R', 3);
    for (const char *s : {"u_c", "v_c", "u_b", "v_b"}) c.item(std::string("work%") + s, 'R', 2);
    c.item("uv", 'R', 3);
    c.item("uv_rhs", 'R', 3);
    c.item("uv_rhsAB", 'R', 4);
    c.item("uvnode", 'R', 3);
    for (const char *s : {"w", "w_e", "w_i", "cfl_z"}) c.item(s, 'R', 2);
    if (c.integer("use_ssh_se_subcycl")) {
        c.item("se_uvh", 'R', 3);
        for (const char *s : {"se_uvBT_rhs", "se_uvBT_4AB", "se_uvBT", "se_uvBT_theta", "se_uvBT_mean",
                              "se_uvBT_12", "se_uvBT_stab_hvisc"})
            c.item(s, 'R', 2);
        c.item("se_uvBT_stab_bdrag", 'R', 1);
    }
}

void items(Codec &c, const std::string &kind)
{
    if (kind == "t_mesh") mesh_items(c);
    else if (kind == "t_partit") partit_items(c);
    else if (kind == "t_tracer") tracer_items(c);
    else if (kind == "t_dynamics") dyn_items(c);
    else throw err("unknown restart kind " + kind);
}

// ---- mapping to the library's state ----------------------------------------

const RItem *find(const RRecord &r, const std::string &name)
{
    auto it = r.find(name);
    return it == r.end() ? nullptr : &it->second;
}

bool has(const RRecord &r, const std::string &name)
{
    const RItem *it = find(r, name);
    return it && (!it->i.empty() || !it->r.empty());
}

void expect_dims(const RItem &it, std::initializer_list<long long> want, const std::string &name)
{
    std::vector<long long> d(it.dims.begin(), it.dims.end());
    if (d != std::vector<long long>(want)) {
        std::string w;
        for (long long x : want) w += (w.empty() ? "" : ",") + std::to_string(x);
        throw err("restart array " + name + ": extents differ from (" + w + ")");
    }
}

std::vector<int> ints(const RRecord &r, const std::string &name, std::initializer_list<long long> dims, int shift)
{
    const RItem *it = find(r, name);
    if (!it || it->i.empty()) throw err("restart array " + name + " is missing");
    expect_dims(*it, dims, name);
    std::vector<int> v = it->i;
    for (int &x : v) x += shift;
    return v;
}

std::vector<double> reals(const RRecord &r, const std::string &name, std::initializer_list<long long> dims)
{
    const RItem *it = find(r, name);
    if (!it || it->r.empty()) throw err("restart array " + name + " is missing");
    expect_dims(*it, dims, name);
    return it->r;
}

RItem arr_i(std::vector<int> dims, std::vector<int> v, int shift = 0)
{
    RItem it;
    it.dims = std::move(dims);
    for (int &x : v) x += shift;
    it.i = std::move(v);
    return it;
}
RItem arr_r(std::vector<int> dims, std::vector<double> v)
{
    RItem it;
    it.dims = std::move(dims);
    it.r = std::move(v);
    return it;
}
RItem sc_i(int v)
{
    RItem it;
    it.i = {v};
    return it;
}
RItem sc_r(double v)
{
    RItem it;
    it.r = {v};
    return it;
}
RItem sc_s(const std::string &s)
{
    RItem it;
    it.s = s;
    return it;
}

std::string rank_file(const std::string &npdir, const char *kind) { return npdir + "/" + kind + ".0"; }

}  // namespace

RRecord read_restart_file(const std::string &path, const std::string &kind)
{
    const std::vector<unsigned char> payload = read_record(path);
    RRecord rec;
    Codec c(rec, payload, path);
    items(c, kind);
    if (!c.done())
        throw err(path + ": " + std::to_string(payload.size() - c.pos()) +
                  " bytes left after the last item (a different build or layout)");
    return rec;
}

void write_restart_file(const std::string &path, const std::string &kind, const RRecord &rec)
{
    RRecord copy = rec;  // the codec reads scalars it has written (counts, flags)
    Codec c(copy, path);
    items(c, kind);
    write_record(path, c.out);
}

void read_restart(const std::string &npdir, RestartState &st, bool with_tracers, bool with_dynamics)
{
    const RRecord part = read_restart_file(rank_file(npdir, "t_partit"), "t_partit");
    st.npes = find(part, "npes") ? find(part, "npes")->i[0] : 1;
    st.mype = find(part, "mype") ? find(part, "mype")->i[0] : 0;
    const int eDim = find(part, "eDim_nod2D") ? find(part, "eDim_nod2D")->i[0] : 0;
    if (st.npes != 1 || eDim != 0)
        throw err(npdir + ": a restart of " + std::to_string(st.npes) +
                  " ranks; the library reads one-rank restarts (npes = 1, no halo)");

    const RRecord mr = read_restart_file(rank_file(npdir, "t_mesh"), "t_mesh");
    HostMesh m;
    auto sc = [&](const char *n) {
        const RItem *it = find(mr, n);
        if (!it || it->i.empty()) throw err("t_mesh: " + std::string(n) + " missing");
        return it->i[0];
    };
    const long long N = m.nod2D = sc("nod2D");
    const long long T = m.elem2D = sc("elem2D");
    const long long E = m.edge2D = sc("edge2D");
    const long long L = m.nl = sc("nl");
    if (N < 1 || T < 1 || E < 1 || L < 3) throw err("t_mesh: implausible dimensions");
    const RItem *nie = find(mr, "nod_in_elem2D");
    if (!nie || nie->dims.size() != 2 || nie->dims[1] != N) throw err("t_mesh: nod_in_elem2D missing or not (maxAdj, nod2D)");
    const long long A = m.maxAdj = nie->dims[0];
    // indices are 1-based in the file; 0 ("none") becomes -1
    m.elem2D_nodes = ints(mr, "elem2D_nodes", {3, T}, -1);
    m.edges = ints(mr, "edges", {2, E}, -1);
    m.edge_tri = ints(mr, "edge_tri", {2, E}, -1);
    m.nod_in_elem2D = ints(mr, "nod_in_elem2D", {A, N}, -1);
    m.nod_in_elem2D_num = ints(mr, "nod_in_elem2D_num", {N}, 0);
    m.elem_area = reals(mr, "elem_area", {T});
    m.edge_dxdy = reals(mr, "edge_dxdy", {2, E});
    m.edge_cross_dxdy = reals(mr, "edge_cross_dxdy", {4, E});
    m.elem_cos = reals(mr, "elem_cos", {T});
    m.gradient_sca = reals(mr, "gradient_sca", {6, T});
    m.zbar = reals(mr, "zbar", {L});
    m.nlevels = ints(mr, "nlevels", {T}, 0);
    m.nlevels_nod2D = ints(mr, "nlevels_nod2D", {N}, 0);
    m.area = reals(mr, "area", {L, N});
    m.hnode = reals(mr, "hnode", {L - 1, N});
    m.helem = reals(mr, "helem", {L - 1, T});
    std::vector<std::string> given = {"edges", "edge_tri", "elem2D_nodes", "nlevels", "nlevels_nod2D",
                                      "nod_in_elem2D_num", "nod_in_elem2D", "zbar", "elem_cos", "elem_area",
                                      "gradient_sca", "edge_dxdy", "edge_cross_dxdy", "area", "hnode", "helem"};
    auto opt_i = [&](const char *n, std::vector<int> &dst, std::initializer_list<long long> d) {
        if (has(mr, n)) { dst = ints(mr, n, d, 0); given.push_back(n); }
    };
    auto opt_r = [&](const char *n, std::vector<double> &dst, std::initializer_list<long long> d) {
        if (has(mr, n)) { dst = reals(mr, n, d); given.push_back(n); }
    };
    opt_i("ulevels", m.ulevels, {T});
    opt_i("ulevels_nod2D", m.ulevels_nod2D, {N});
    opt_i("ulevels_nod2D_max", m.ulevels_nod2D_max, {N});
    opt_i("nlevels_nod2D_min", m.nlevels_nod2D_min, {N});
    opt_r("Z", m.Z, {L - 1});
    opt_r("depth", m.depth, {N});
    opt_r("areasvol", m.areasvol, {L, N});
    opt_r("hnode_new", m.hnode_new, {L - 1, N});
    opt_r("zbar_3d_n", m.zbar_3d_n, {L, N});
    opt_r("Z_3d_n", m.Z_3d_n, {L - 1, N});

    RRecord tr;
    if (with_tracers) tr = read_restart_file(rank_file(npdir, "t_tracer"), "t_tracer");
    if (has(tr, "work%nboundary_lay")) {
        m.nboundary_lay = ints(tr, "work%nboundary_lay", {N}, 0);
        given.push_back("nboundary_lay");
    }
    if (has(tr, "work%edge_up_dn_tri")) {
        m.edge_up_dn_tri = ints(tr, "work%edge_up_dn_tri", {2, E}, -1);
        given.push_back("edge_up_dn_tri");
    }
    complete_mesh(m, [&](const std::string &n) { return std::find(given.begin(), given.end(), n) != given.end(); });
    st.mesh = std::move(m);

    st.tracers.clear();
    if (with_tracers) {
        const int n = find(tr, "num_tracers") ? find(tr, "num_tracers")->i[0] : 0;
        for (int k = 1; k <= n; ++k) {
            const std::string p = tracer_prefix(k);
            RestartTracer t;
            t.values = reals(tr, p + "values", {L - 1, N});
            t.valuesAB = has(tr, p + "valuesAB") ? reals(tr, p + "valuesAB", {L - 1, N}) : t.values;
            t.hor = find(tr, p + "tra_adv_hor")->s;
            t.ver = find(tr, p + "tra_adv_ver")->s;
            t.lim = find(tr, p + "tra_adv_lim")->s;
            t.ph = find(tr, p + "tra_adv_ph")->r[0];
            t.pv = find(tr, p + "tra_adv_pv")->r[0];
            t.id = find(tr, p + "ID")->i[0];
            st.tracers.push_back(std::move(t));
        }
    }
    st.uv.clear();
    st.w.clear();
    st.w_e.clear();
    if (with_dynamics) {
        const RRecord dr = read_restart_file(rank_file(npdir, "t_dynamics"), "t_dynamics");
        if (has(dr, "uv")) st.uv = reals(dr, "uv", {2, L - 1, T});
        if (has(dr, "w")) st.w = reals(dr, "w", {L, N});
        if (has(dr, "w_e")) st.w_e = reals(dr, "w_e", {L, N});
    }
}

void write_restart(const std::string &npdir, const std::string &meta_path, const HostMesh &m,
                   const std::vector<RestartTracer> &tracers, const std::vector<double> &uv,
                   const std::vector<double> &w, const std::vector<double> &w_e)
{
    const int N = m.nod2D, T = m.elem2D, E = m.edge2D, L = m.nl, A = m.maxAdj;
    {
        struct stat sb;
        const std::string parent = npdir.substr(0, npdir.find_last_of('/'));
        if (!parent.empty() && ::stat(parent.c_str(), &sb) != 0) ::mkdir(parent.c_str(), 0755);
        if (::stat(npdir.c_str(), &sb) != 0 && ::mkdir(npdir.c_str(), 0755) != 0)
            throw err("cannot create " + npdir + ": " + std::strerror(errno));
    }
    // fields outside the tracer path are written unallocated, scalars at
    // FESOM's defaults
    RRecord mr;
    double ocean = 0.0;
    for (int n = 0; n < N; ++n) ocean += m.area[(size_t)n * L];
    int inner = 0;
    for (int e = 0; e < E; ++e) inner += m.edge_tri[2 * e + 1] >= 0 ? 1 : 0;
    mr["nod2D"] = sc_i(N);
    mr["ocean_area"] = sc_r(ocean);
    mr["ocean_areawithcav"] = sc_r(ocean);
    mr["edge2D"] = sc_i(E);
    mr["edge2D_in"] = sc_i(inner);
    mr["elem2D"] = sc_i(T);
    mr["nl"] = sc_i(L);
    mr["elem2D_nodes"] = arr_i({3, T}, m.elem2D_nodes, 1);
    mr["edges"] = arr_i({2, E}, m.edges, 1);
    mr["edge_tri"] = arr_i({2, E}, m.edge_tri, 1);
    {
        std::vector<int> nie = m.nod_in_elem2D;
        for (int n = 0; n < N; ++n)
            for (int j = m.nod_in_elem2D_num[n]; j < A; ++j) nie[(size_t)n * A + j] = -1;
        mr["nod_in_elem2D"] = arr_i({A, N}, nie, 1);
    }
    mr["nod_in_elem2D_num"] = arr_i({N}, m.nod_in_elem2D_num);
    mr["elem_area"] = arr_r({T}, m.elem_area);
    mr["edge_dxdy"] = arr_r({2, E}, m.edge_dxdy);
    mr["edge_cross_dxdy"] = arr_r({4, E}, m.edge_cross_dxdy);
    mr["elem_cos"] = arr_r({T}, m.elem_cos);
    mr["depth"] = arr_r({N}, m.depth);
    mr["gradient_sca"] = arr_r({6, T}, m.gradient_sca);
    mr["zbar"] = arr_r({L}, m.zbar);
    mr["Z"] = arr_r({L - 1}, m.Z);
    mr["ulevels"] = arr_i({T}, m.ulevels);
    mr["ulevels_nod2D"] = arr_i({N}, m.ulevels_nod2D);
    mr["ulevels_nod2D_max"] = arr_i({N}, m.ulevels_nod2D_max);
    mr["nlevels"] = arr_i({T}, m.nlevels);
    mr["nlevels_nod2D"] = arr_i({N}, m.nlevels_nod2D);
    mr["nlevels_nod2D_min"] = arr_i({N}, m.nlevels_nod2D_min);
    mr["area"] = arr_r({L, N}, m.area);
    mr["areasvol"] = arr_r({L, N}, m.areasvol);
    mr["hnode"] = arr_r({L - 1, N}, m.hnode);
    mr["hnode_new"] = arr_r({L - 1, N}, m.hnode_new);
    mr["zbar_3d_n"] = arr_r({L, N}, m.zbar_3d_n);
    mr["Z_3d_n"] = arr_r({L - 1, N}, m.Z_3d_n);
    mr["helem"] = arr_r({L - 1, T}, m.helem);
    write_restart_file(rank_file(npdir, "t_mesh"), "t_mesh", mr);

    RRecord pr;
    for (const char *c : {"com_nod2D%", "com_elem2D%", "com_elem2D_full%"}) {
        pr[std::string(c) + "rptr"] = arr_i({33}, std::vector<int>(33, 0));
        pr[std::string(c) + "rptr"].i[0] = 1;  // an empty exchange list starts at 1
        pr[std::string(c) + "sptr"] = arr_i({32}, std::vector<int>(32, 0));
        pr[std::string(c) + "sptr"].i[0] = 1;
    }
    auto iota1 = [](int n) {
        std::vector<int> v(n);
        for (int k = 0; k < n; ++k) v[k] = k + 1;
        return v;
    };
    pr["npes"] = sc_i(1);
    pr["mype"] = sc_i(0);
    pr["maxPEnum"] = sc_i(100);
    pr["part"] = arr_i({N}, std::vector<int>(N, 0));
    pr["myDim_nod2D"] = sc_i(N);
    pr["myList_nod2D"] = arr_i({N}, iota1(N));
    pr["myDim_elem2D"] = sc_i(T);
    pr["myList_elem2D"] = arr_i({T}, iota1(T));
    pr["myDim_edge2D"] = sc_i(E);
    pr["myList_edge2D"] = arr_i({E}, iota1(E));
    write_restart_file(rank_file(npdir, "t_partit"), "t_partit", pr);

    RRecord tr;
    tr["num_tracers"] = sc_i((int)tracers.size());
    for (size_t k = 0; k < tracers.size(); ++k) {
        const RestartTracer &t = tracers[k];
        const std::string p = tracer_prefix((int)k + 1);
        tr[p + "values"] = arr_r({L - 1, N}, t.values);
        tr[p + "valuesAB"] = arr_r({L - 1, N}, t.valuesAB);
        tr[p + "tra_adv_hor"] = sc_s(t.hor);
        tr[p + "tra_adv_ver"] = sc_s(t.ver);
        tr[p + "tra_adv_lim"] = sc_s(t.lim);
        tr[p + "tra_adv_ph"] = sc_r(t.ph);
        tr[p + "tra_adv_pv"] = sc_r(t.pv);
        tr[p + "ID"] = sc_i(t.id);
    }
    tr["work%nboundary_lay"] = arr_i({N}, m.nboundary_lay);
    tr["work%edge_up_dn_tri"] = arr_i({2, E}, m.edge_up_dn_tri, 1);
    write_restart_file(rank_file(npdir, "t_tracer"), "t_tracer", tr);

    RRecord dr;
    dr["opt_visc"] = sc_i(5);
    dr["visc_gamma0"] = sc_r(0.03);
    dr["visc_gamma1"] = sc_r(0.1);
    dr["visc_gamma2"] = sc_r(0.285);
    dr["visc_easybsreturn"] = sc_r(1.5);
    dr["use_ivertvisc"] = sc_i(1);
    dr["momadv_opt"] = sc_i(2);
    dr["wsplit_maxcfl"] = sc_r(1.0);
    dr["solverinfo%ident"] = sc_i(1);
    dr["solverinfo%maxiter"] = sc_i(2000);
    dr["solverinfo%restart"] = sc_i(15);
    dr["solverinfo%fillin"] = sc_i(3);
    dr["solverinfo%lutype"] = sc_i(2);
    dr["solverinfo%droptol"] = sc_r(1e-8);
    dr["solverinfo%soltol"] = sc_r(1e-5);
    if (!uv.empty()) dr["uv"] = arr_r({2, L - 1, T}, uv);
    if (!w.empty()) dr["w"] = arr_r({L, N}, w);
    if (!w_e.empty()) dr["w_e"] = arr_r({L, N}, w_e);
    write_restart_file(rank_file(npdir, "t_dynamics"), "t_dynamics", dr);

    // meta file as write_all_bin_restarts leaves it (time stamp 0 0 0)
    FILE *f = std::fopen(meta_path.c_str(), "w");
    if (!f) throw err("cannot create " + meta_path);
    std::fprintf(f, "0\n0\n! year: 0\n!   t_mesh\n!   t_partit\n!   t_tracer\n!   t_dynamics\n");
    std::fclose(f);
}

}  // namespace fta
