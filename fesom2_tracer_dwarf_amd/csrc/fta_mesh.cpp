// fta_mesh.cpp -- host-side mesh generation and the setup-time derivations
// the tracer-advection step reads (geometry, levels, adjacency, MUSCL
// up/down-wind triangles). Setup only; nothing here runs inside a step.
#include "fta_mesh.h"

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <numeric>
#include <stdexcept>
#include <unordered_map>

namespace fta {

namespace {

uint64_t splitmix64(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// position of grid point (x, y) along the Hilbert curve that fills the
// square of side 2^bits: quadrant by quadrant from the top bit down, each
// quadrant's sub-square turned so that the curve enters and leaves it next
// to its neighbours (unlike a Morton curve, no jumps between quadrants)
uint64_t hilbert2(uint32_t x, uint32_t y, int bits)
{
    const uint32_t n = 1u << bits;
    uint64_t d = 0;
    for (uint32_t s = n >> 1; s > 0; s >>= 1) {
        const uint32_t rx = (x & s) ? 1u : 0u, ry = (y & s) ? 1u : 0u;
        d += (uint64_t)s * s * ((3u * rx) ^ ry);
        if (ry == 0) {  // lower quadrants: reflect (right one) and transpose
            if (rx == 1) {
                x = n - 1 - x;
                y = n - 1 - y;
            }
            std::swap(x, y);
        }
    }
    return d;
}

// bits of a square side that covers an nx x ny grid
int side_bits(int nx, int ny)
{
    int b = 0;
    while ((1 << b) < std::max(nx, ny)) ++b;
    return b;
}

// minimum-image trim of a coordinate difference (trim_cyclic,
// lib/gen_modules_config.F90:207-223, generalised to both axes)
inline double trim(double d, double cyc)
{
    if (cyc > 0.0) {
        if (d > cyc / 2.0) d -= cyc;
        if (d < -cyc / 2.0) d += cyc;
    }
    return d;
}

// Edges from clockwise elements, in order of first appearance. A clockwise
// element lies to the right of each of its directed sides a->b, i.e. to the
// left of b->a, so the first element met on a side becomes edge_tri(1) with
// edges = (b, a): FESOM's convention (lib/MOD_MESH.F90:35-37).
void build_edges(HostMesh &m)
{
    std::unordered_map<uint64_t, int> idx;
    idx.reserve((size_t)m.elem2D * 2);
    m.edges.clear();
    m.edge_tri.clear();
    for (int el = 0; el < m.elem2D; ++el) {
        for (int s = 0; s < 3; ++s) {
            int a = m.elem2D_nodes[3 * el + s], b = m.elem2D_nodes[3 * el + (s + 1) % 3];
            uint64_t key = ((uint64_t)std::min(a, b) << 32) | (uint32_t)std::max(a, b);
            auto it = idx.find(key);
            if (it == idx.end()) {
                int e = (int)(m.edges.size() / 2);
                idx.emplace(key, e);
                m.edges.push_back(b);
                m.edges.push_back(a);
                m.edge_tri.push_back(el);
                m.edge_tri.push_back(-1);
            } else {
                int e = it->second;
                if (m.edge_tri[2 * e + 1] != -1)
                    throw std::runtime_error("mesh: edge shared by more than two elements");
                m.edge_tri[2 * e + 1] = el;
            }
        }
    }
    m.edge2D = (int)(m.edges.size() / 2);
}

void build_nod_in_elem(HostMesh &m)
{
    // element loop in index order, appended per node (find_neighbors order)
    m.nod_in_elem2D_num.assign(m.nod2D, 0);
    for (int el = 0; el < m.elem2D; ++el)
        for (int s = 0; s < 3; ++s) m.nod_in_elem2D_num[m.elem2D_nodes[3 * el + s]]++;
    m.maxAdj = *std::max_element(m.nod_in_elem2D_num.begin(), m.nod_in_elem2D_num.end());
    m.nod_in_elem2D.assign((size_t)m.maxAdj * m.nod2D, -1);
    std::vector<int> cnt(m.nod2D, 0);
    for (int el = 0; el < m.elem2D; ++el)
        for (int s = 0; s < 3; ++s) {
            int n = m.elem2D_nodes[3 * el + s];
            m.nod_in_elem2D[(size_t)n * m.maxAdj + cnt[n]++] = el;
        }
}

void finish_topology(HostMesh &m)
{
    m.coord_nod2D.resize(2 * (size_t)m.nod2D);
    for (size_t i = 0; i < m.coord_nod2D.size(); ++i) m.coord_nod2D[i] = m.coord[i] / kREarth;
    // clockwise check (test_tri, lib/oce_mesh.F90:1565-1610)
    for (int el = 0; el < m.elem2D; ++el) {
        int *en = &m.elem2D_nodes[3 * el];
        double bx = trim(m.coord[2 * en[1]] - m.coord[2 * en[0]], m.cyc_x);
        double by = trim(m.coord[2 * en[1] + 1] - m.coord[2 * en[0] + 1], m.cyc_y);
        double cx = trim(m.coord[2 * en[2]] - m.coord[2 * en[0]], m.cyc_x);
        double cy = trim(m.coord[2 * en[2] + 1] - m.coord[2 * en[0] + 1], m.cyc_y);
        if (bx * cy - by * cx > 0.0) std::swap(en[1], en[2]);
    }
    build_edges(m);
    build_nod_in_elem(m);
}

void rect_nodes(HostMesh &m, int nx, int ny, double Lx, double Ly, int order, bool periodic,
                std::vector<int> &node_id)
{
    const double dx = periodic ? Lx / nx : Lx / (nx - 1);
    const double dy = periodic ? Ly / ny : Ly / (ny - 1);
    m.nod2D = nx * ny;
    std::vector<int> perm(m.nod2D);
    std::iota(perm.begin(), perm.end(), 0);
    if (order == 1) {
        std::vector<uint64_t> key(m.nod2D);
        const int bits = side_bits(nx, ny);
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) key[j * nx + i] = hilbert2(i, j, bits);
        std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return key[a] < key[b]; });
    }
    node_id.assign(m.nod2D, 0);
    for (int r = 0; r < m.nod2D; ++r) node_id[perm[r]] = r;
    m.coord.assign(2 * (size_t)m.nod2D, 0.0);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            int n = node_id[j * nx + i];
            m.coord[2 * n] = i * dx;
            m.coord[2 * n + 1] = j * dy;
        }
}

void rect_elements(HostMesh &m, int ncx, int ncy, int nx, int order, bool periodic,
                   const std::vector<int> &node_id)
{
    int ny_nodes = m.nod2D / nx;
    auto nid = [&](int i, int j) {
        if (periodic) { i = (i + nx) % nx; j = (j + ny_nodes) % ny_nodes; }
        return node_id[j * nx + i];
    };
    std::vector<int> cells(ncx * ncy);
    std::iota(cells.begin(), cells.end(), 0);
    if (order == 1) {
        std::vector<uint64_t> key(cells.size());
        const int bits = side_bits(ncx, ncy);
        for (int j = 0; j < ncy; ++j)
            for (int i = 0; i < ncx; ++i) key[j * ncx + i] = hilbert2(i, j, bits);
        std::stable_sort(cells.begin(), cells.end(), [&](int a, int b) { return key[a] < key[b]; });
    }
    m.elem2D = 2 * ncx * ncy;
    m.elem2D_nodes.assign(3 * (size_t)m.elem2D, 0);
    int el = 0;
    for (int c : cells) {
        int i = c % ncx, j = c / ncx;
        // A: (i,j),(i,j+1),(i+1,j+1)   B: (i,j),(i+1,j+1),(i+1,j)   both clockwise
        int A[3] = {nid(i, j), nid(i, j + 1), nid(i + 1, j + 1)};
        int B[3] = {nid(i, j), nid(i + 1, j + 1), nid(i + 1, j)};
        for (int s = 0; s < 3; ++s) m.elem2D_nodes[3 * el + s] = A[s];
        ++el;
        for (int s = 0; s < 3; ++s) m.elem2D_nodes[3 * el + s] = B[s];
        ++el;
    }
}

}  // namespace

void generate_periodic(HostMesh &m, int nx, int ny, double Lx, double Ly, int order)
{
    if (nx < 3 || ny < 3) throw std::runtime_error("periodic mesh needs nx, ny >= 3");
    m = HostMesh();
    m.cartesian = true;
    m.cyc_x = Lx;
    m.cyc_y = Ly;
    std::vector<int> node_id;
    rect_nodes(m, nx, ny, Lx, Ly, order, true, node_id);
    rect_elements(m, nx, ny, nx, order, true, node_id);
    finish_topology(m);
}

void generate_closed(HostMesh &m, int nx, int ny, double Lx, double Ly, int order)
{
    if (nx < 2 || ny < 2) throw std::runtime_error("closed mesh needs nx, ny >= 2");
    m = HostMesh();
    m.cartesian = true;
    std::vector<int> node_id;
    rect_nodes(m, nx, ny, Lx, Ly, order, false, node_id);
    rect_elements(m, nx - 1, ny - 1, nx, order, false, node_id);
    finish_topology(m);
}

void set_vertical_grid(HostMesh &m, int nl, double dz, double stretch)
{
    if (nl < 3) throw std::runtime_error("need nl >= 3 level interfaces");
    m.nl = nl;
    m.zbar.assign(nl, 0.0);
    m.Z.assign(nl - 1, 0.0);
    for (int k = 1; k < nl; ++k) {
        double t = (nl > 2) ? (double)(k - 1) / (double)(nl - 2) : 0.0;
        m.zbar[k] = m.zbar[k - 1] - dz * (1.0 + stretch * t);
    }
    for (int k = 0; k < nl - 1; ++k) m.Z[k] = 0.5 * (m.zbar[k] + m.zbar[k + 1]);
}

static void node_levels_from_elements(HostMesh &m)
{
    m.ulevels.assign(m.elem2D, 1);
    m.nlevels_nod2D.assign(m.nod2D, 0);
    m.ulevels_nod2D.assign(m.nod2D, 1);
    for (int n = 0; n < m.nod2D; ++n) {
        int mx = 0;
        for (int k = 0; k < m.nod_in_elem2D_num[n]; ++k)
            mx = std::max(mx, m.nlevels[m.nod_in_elem2D[(size_t)n * m.maxAdj + k]]);
        m.nlevels_nod2D[n] = mx;
    }
}

void set_levels_random(HostMesh &m, int lo, int hi, uint64_t seed)
{
    if (m.nl == 0) throw std::runtime_error("set the vertical grid first");
    lo = std::max(2, std::min(lo, m.nl));
    hi = std::max(lo, std::min(hi, m.nl));
    std::vector<int> raw(m.nod2D);
    uint64_t s = seed;
    for (int n = 0; n < m.nod2D; ++n) raw[n] = lo + (int)(splitmix64(s) % (uint64_t)(hi - lo + 1));
    m.nlevels.assign(m.elem2D, 0);
    for (int el = 0; el < m.elem2D; ++el) {
        const int *en = &m.elem2D_nodes[3 * el];
        m.nlevels[el] = std::min(raw[en[0]], std::min(raw[en[1]], raw[en[2]]));
    }
    node_levels_from_elements(m);
}

void set_levels_from_elements(HostMesh &m, const int *nlevels_elem)
{
    m.nlevels.assign(nlevels_elem, nlevels_elem + m.elem2D);
    node_levels_from_elements(m);
}

// ---------------------------------------------------------------------------
// FESOM mesh files
// ---------------------------------------------------------------------------
void rotation_matrix(double al, double be, double ga, double M[9])
{
    // set_mesh_transform_matrix, lib/gen_modules_rotate_grid.F90:33-56
    M[0] = std::cos(ga) * std::cos(al) - std::sin(ga) * std::cos(be) * std::sin(al);
    M[1] = std::cos(ga) * std::sin(al) + std::sin(ga) * std::cos(be) * std::cos(al);
    M[2] = std::sin(ga) * std::sin(be);
    M[3] = -std::sin(ga) * std::cos(al) - std::cos(ga) * std::cos(be) * std::sin(al);
    M[4] = -std::sin(ga) * std::sin(al) + std::cos(ga) * std::cos(be) * std::cos(al);
    M[5] = std::cos(ga) * std::sin(be);
    M[6] = std::sin(be) * std::sin(al);
    M[7] = -std::sin(be) * std::cos(al);
    M[8] = std::cos(be);
}

void geo_to_rot(const double M[9], double glon, double glat, double &rlon, double &rlat)
{
    // g2r, lib/gen_modules_rotate_grid.F90:92-119: rotated = M * geographic
    const double xg = std::cos(glat) * std::cos(glon), yg = std::cos(glat) * std::sin(glon), zg = std::sin(glat);
    const double xr = M[0] * xg + M[1] * yg + M[2] * zg;
    const double yr = M[3] * xg + M[4] * yg + M[5] * zg;
    const double zr = M[6] * xg + M[7] * yg + M[8] * zg;
    rlat = std::asin(zr);
    rlon = (yr == 0.0 && xr == 0.0) ? 0.0 : std::atan2(yr, xr);
}

void rot_to_geo(const double M[9], double rlon, double rlat, double &glon, double &glat)
{
    // r2g, lib/gen_modules_rotate_grid.F90:60-88: geographic = M^T * rotated
    const double xr = std::cos(rlat) * std::cos(rlon), yr = std::cos(rlat) * std::sin(rlon), zr = std::sin(rlat);
    const double xg = M[0] * xr + M[3] * yr + M[6] * zr;
    const double yg = M[1] * xr + M[4] * yr + M[7] * zr;
    const double zg = M[2] * xr + M[5] * yr + M[8] * zr;
    glat = std::asin(zg);
    glon = (yg == 0.0 && xg == 0.0) ? 0.0 : std::atan2(yg, xg);
}

namespace {

// whitespace-separated numbers of a text file, read as Fortran list-directed
// input reads them
class Tokens {
  public:
    Tokens(const std::string &path, bool required) : path_(path)
    {
        FILE *f = std::fopen(path.c_str(), "rb");
        if (!f) {
            if (required) throw std::runtime_error("mesh file not found: " + path);
            return;
        }
        std::fseek(f, 0, SEEK_END);
        const long n = std::ftell(f);
        std::fseek(f, 0, SEEK_SET);
        buf_.resize(n > 0 ? (size_t)n : 0);
        const size_t got = n > 0 ? std::fread(&buf_[0], 1, (size_t)n, f) : 0;
        std::fclose(f);
        buf_.resize(got);
        ok_ = true;
    }
    bool ok() const { return ok_; }
    double real()
    {
        skip();
        if (pos_ >= buf_.size()) throw std::runtime_error("unexpected end of " + path_);
        const char *b = buf_.c_str() + pos_;
        char *e = nullptr;
        errno = 0;
        const double v = std::strtod(b, &e);
        if (e == b || errno == ERANGE) throw std::runtime_error("bad number in " + path_ + " near byte " + std::to_string(pos_));
        pos_ += (size_t)(e - b);
        return v;
    }
    int integer()
    {
        const double v = real();
        if (v != std::floor(v) || std::fabs(v) > 2.0e9) throw std::runtime_error("expected an integer in " + path_);
        return (int)v;
    }

  private:
    void skip()
    {
        while (pos_ < buf_.size() && (buf_[pos_] == ' ' || buf_[pos_] == '\t' || buf_[pos_] == '\n' ||
                                      buf_[pos_] == '\r' || buf_[pos_] == ','))
            ++pos_;
    }
    std::string path_, buf_;
    size_t pos_ = 0;
    bool ok_ = false;
};

}  // namespace

void read_fesom_mesh(HostMesh &m, const std::string &dir_in, const FesomFileOptions &opt)
{
    const double PI = 3.14159265358979323846, rad = PI / 180.0;
    const std::string dir = (!dir_in.empty() && dir_in.back() != '/') ? dir_in + "/" : dir_in;
    m = HostMesh();
    m.cartesian = false;
    m.cyc_x = opt.cyclic_length_deg * rad * kREarth;
    m.cyc_y = 0.0;

    // nodes (read_mesh :220-313): index, lon, lat, boundary flag; degrees,
    // rotated to the model grid
    {
        Tokens t(dir + "nod2d.out", true);
        m.nod2D = t.integer();
        if (m.nod2D < 3) throw std::runtime_error("nod2d.out: bad node count");
        double M[9];
        rotation_matrix(opt.alpha_deg * rad, opt.beta_deg * rad, opt.gamma_deg * rad, M);
        m.coord_nod2D.assign(2 * (size_t)m.nod2D, 0.0);
        for (int n = 0; n < m.nod2D; ++n) {
            const int id = t.integer();
            if (id != n + 1) throw std::runtime_error("nod2d.out: node " + std::to_string(n + 1) + " out of order");
            double x = t.real() * rad, y = t.real() * rad;
            (void)t.integer();
            if (opt.force_rotation) {
                const double gx = x, gy = y;
                geo_to_rot(M, gx, gy, x, y);
            }
            m.coord_nod2D[2 * n] = x;
            m.coord_nod2D[2 * n + 1] = y;
        }
        m.coord.resize(m.coord_nod2D.size());
        for (size_t i = 0; i < m.coord.size(); ++i) m.coord[i] = m.coord_nod2D[i] * kREarth;
    }

    // elements (read_mesh :366-434), made clockwise (test_tri)
    {
        Tokens t(dir + "elem2d.out", true);
        m.elem2D = t.integer();
        if (m.elem2D < 1) throw std::runtime_error("elem2d.out: bad element count");
        m.elem2D_nodes.assign(3 * (size_t)m.elem2D, 0);
        for (size_t i = 0; i < m.elem2D_nodes.size(); ++i) {
            const int v = t.integer();
            if (v < 1 || v > m.nod2D) throw std::runtime_error("elem2d.out: node index out of range");
            m.elem2D_nodes[i] = v - 1;
        }
        const double C = m.cyc_x / kREarth;
        for (int el = 0; el < m.elem2D; ++el) {
            int *en = &m.elem2D_nodes[3 * el];
            const double bx = trim(m.coord_nod2D[2 * en[1]] - m.coord_nod2D[2 * en[0]], C);
            const double by = m.coord_nod2D[2 * en[1] + 1] - m.coord_nod2D[2 * en[0] + 1];
            const double cx = trim(m.coord_nod2D[2 * en[2]] - m.coord_nod2D[2 * en[0]], C);
            const double cy = m.coord_nod2D[2 * en[2] + 1] - m.coord_nod2D[2 * en[0] + 1];
            if (bx * cy - by * cx > 0.0) std::swap(en[1], en[2]);
        }
    }

    // vertical grid (read_mesh :440-468): nl, zbar (made negative), Z
    {
        Tokens t(dir + "aux3d.out", true);
        m.nl = t.integer();
        if (m.nl < 3) throw std::runtime_error("aux3d.out: number of levels is less than 3");
        m.zbar.assign(m.nl, 0.0);
        for (int k = 0; k < m.nl; ++k) m.zbar[k] = t.real();
        if (m.zbar[1] > 0.0)
            for (double &z : m.zbar) z = -z;
        m.Z.assign(m.nl - 1, 0.0);
        for (int k = 0; k < m.nl - 1; ++k) m.Z[k] = 0.5 * (m.zbar[k] + m.zbar[k + 1]);
        // node depths (read_mesh :606-646): made negative, no shallower than
        // zbar(thers_zbar_lev = 5) (lib/gen_modules_config.F90:90)
        m.depth.assign(m.nod2D, 0.0);
        const double shallow = m.zbar[std::min(4, m.nl - 1)];
        for (int n = 0; n < m.nod2D; ++n) {
            double x = t.real();
            if (x > 0.0) x = -x;
            if (x > shallow) x = shallow;
            m.depth[n] = x;
        }
    }

    // edges (load_edges :1617-1790): edges.out, edge_tri.out (<= 0: none)
    {
        Tokens tn(dir + "edgenum.out", false), te(dir + "edges.out", false), tt(dir + "edge_tri.out", false);
        if (tn.ok() && te.ok() && tt.ok()) {
            m.edge2D = tn.integer();
            if (m.edge2D < 1) throw std::runtime_error("edgenum.out: bad edge count");
            m.edges.assign(2 * (size_t)m.edge2D, 0);
            m.edge_tri.assign(2 * (size_t)m.edge2D, -1);
            for (size_t i = 0; i < m.edges.size(); ++i) {
                const int v = te.integer();
                if (v < 1 || v > m.nod2D) throw std::runtime_error("edges.out: node index out of range");
                m.edges[i] = v - 1;
                const int w = tt.integer();
                if (w > m.elem2D) throw std::runtime_error("edge_tri.out: element index out of range");
                m.edge_tri[i] = w > 0 ? w - 1 : -1;
            }
            for (int e = 0; e < m.edge2D; ++e)
                if (m.edge_tri[2 * e] < 0) throw std::runtime_error("edge_tri.out: edge without a first element");
        } else {
            build_edges(m);
        }
    }
    build_nod_in_elem(m);

    // levels (find_levels :807-1001): elvls.out per element, nlvls.out per node
    {
        Tokens te(dir + "elvls.out", true);
        m.nlevels.assign(m.elem2D, 0);
        for (int el = 0; el < m.elem2D; ++el) {
            m.nlevels[el] = te.integer();
            if (m.nlevels[el] < 2 || m.nlevels[el] > m.nl) throw std::runtime_error("elvls.out: level out of range");
        }
        node_levels_from_elements(m);
        Tokens tn(dir + "nlvls.out", false);
        if (tn.ok())
            for (int n = 0; n < m.nod2D; ++n)
                if (tn.integer() != m.nlevels_nod2D[n])
                    throw std::runtime_error("nlvls.out: node " + std::to_string(n + 1) +
                                             " is not the deepest of its elements");
    }
    if (opt.device_geometry) derive_all_device(m);
    else derive_all(m);
}

// Geometry of a planar mesh in metres (the generator's space): orientation
// check, mesh_areas and mesh_auxiliary_arrays with elem_cos = 1
static void cartesian_geometry(HostMesh &m)
{
    const int N = m.nod2D, T = m.elem2D, E = m.edge2D, nl = m.nl;
    // edge orientation check: edge_tri(1) must lie left of edges(1)->edges(2)
    auto centroid = [&](int el, double &cx, double &cy) {
        const int *en = &m.elem2D_nodes[3 * el];
        double x0 = m.coord[2 * en[0]], y0 = m.coord[2 * en[0] + 1];
        cx = x0 + (trim(m.coord[2 * en[1]] - x0, m.cyc_x) + trim(m.coord[2 * en[2]] - x0, m.cyc_x)) / 3.0;
        cy = y0 + (trim(m.coord[2 * en[1] + 1] - y0, m.cyc_y) + trim(m.coord[2 * en[2] + 1] - y0, m.cyc_y)) / 3.0;
    };
    for (int e = 0; e < E; ++e) {
        int n1 = m.edges[2 * e], n2 = m.edges[2 * e + 1];
        double ex = trim(m.coord[2 * n2] - m.coord[2 * n1], m.cyc_x);
        double ey = trim(m.coord[2 * n2 + 1] - m.coord[2 * n1 + 1], m.cyc_y);
        double cx, cy;
        centroid(m.edge_tri[2 * e], cx, cy);
        double rx = trim(cx - m.coord[2 * n1], m.cyc_x), ry = trim(cy - m.coord[2 * n1 + 1], m.cyc_y);
        if (ex * ry - ey * rx <= 0.0) throw std::runtime_error("mesh: edge_tri(1) is not left of its edge");
    }

    // mesh_areas (lib/oce_mesh.F90:2043-2309), cartesian, no cavity
    m.elem_area.assign(T, 0.0);
    m.elem_cos.assign(T, 1.0);
    for (int el = 0; el < T; ++el) {
        const int *en = &m.elem2D_nodes[3 * el];
        double ax = trim(m.coord[2 * en[1]] - m.coord[2 * en[0]], m.cyc_x);
        double ay = trim(m.coord[2 * en[1] + 1] - m.coord[2 * en[0] + 1], m.cyc_y);
        double bx = trim(m.coord[2 * en[2]] - m.coord[2 * en[0]], m.cyc_x);
        double by = trim(m.coord[2 * en[2] + 1] - m.coord[2 * en[0] + 1], m.cyc_y);
        m.elem_area[el] = 0.5 * std::fabs(ax * by - bx * ay);
    }
    m.area.assign((size_t)nl * N, 0.0);
    m.areasvol.assign((size_t)nl * N, 0.0);
    for (int n = 0; n < N; ++n)
        for (int j = 0; j < m.nod_in_elem2D_num[n]; ++j) {
            int el = m.nod_in_elem2D[(size_t)n * m.maxAdj + j];
            for (int nz = m.ulevels[el]; nz <= m.nlevels[el] - 1; ++nz)
                m.area[(size_t)n * nl + nz - 1] += m.elem_area[el] / 3.0;
        }
    for (int n = 0; n < N; ++n)
        for (int nz = m.ulevels_nod2D[n]; nz <= m.nlevels_nod2D[n] - 1; ++nz)
            m.areasvol[(size_t)n * nl + nz - 1] = m.area[(size_t)n * nl + nz - 1];

    // mesh_auxiliary_arrays (lib/oce_mesh.F90:2313-2534), cartesian
    m.edge_dxdy.assign(2 * (size_t)E, 0.0);
    m.edge_cross_dxdy.assign(4 * (size_t)E, 0.0);
    for (int e = 0; e < E; ++e) {
        int n1 = m.edges[2 * e], n2 = m.edges[2 * e + 1];
        double ex = trim(m.coord[2 * n2] - m.coord[2 * n1], m.cyc_x);
        double ey = trim(m.coord[2 * n2 + 1] - m.coord[2 * n1 + 1], m.cyc_y);
        m.edge_dxdy[2 * e] = ex / kREarth;
        m.edge_dxdy[2 * e + 1] = ey / kREarth;
        double mx = m.coord[2 * n1] + 0.5 * ex, my = m.coord[2 * n1 + 1] + 0.5 * ey;
        for (int s = 0; s < 2; ++s) {
            int el = m.edge_tri[2 * e + s];
            if (el < 0) continue;
            double cx, cy;
            centroid(el, cx, cy);
            m.edge_cross_dxdy[4 * e + 2 * s] = trim(cx - mx, m.cyc_x) * m.elem_cos[el];
            m.edge_cross_dxdy[4 * e + 2 * s + 1] = trim(cy - my, m.cyc_y);
        }
    }
    m.gradient_sca.assign(6 * (size_t)T, 0.0);
    for (int el = 0; el < T; ++el) {
        const int *en = &m.elem2D_nodes[3 * el];
        double dX31 = trim(m.coord[2 * en[2]] - m.coord[2 * en[0]], m.cyc_x) * m.elem_cos[el];
        double dX21 = trim(m.coord[2 * en[1]] - m.coord[2 * en[0]], m.cyc_x) * m.elem_cos[el];
        double dY31 = trim(m.coord[2 * en[2] + 1] - m.coord[2 * en[0] + 1], m.cyc_y);
        double dY21 = trim(m.coord[2 * en[1] + 1] - m.coord[2 * en[0] + 1], m.cyc_y);
        double dfactor = -0.5 / m.elem_area[el];
        double *g = &m.gradient_sca[6 * el];
        g[0] = (-dY31 + dY21) * dfactor;
        g[1] = dY31 * dfactor;
        g[2] = -dY21 * dfactor;
        g[3] = (dX31 - dX21) * dfactor;
        g[4] = -dX31 * dfactor;
        g[5] = dX21 * dfactor;
    }

}

// Geometry of a mesh on the sphere, coord_nod2D in radians, x cyclic with
// C = cyc_x / r_earth (0: not cyclic):
//   elem_center / edge_center       lib/oce_mesh.F90:2001-2041
//   mesh_areas                      lib/oce_mesh.F90:2043-2309 (no cavity)
//   mesh_auxiliary_arrays           lib/oce_mesh.F90:2313-2534
// elem_area and area accumulate in radians^2 and are scaled by r_earth^2
// afterwards; elem_cos is the cosine of the element centre's latitude and
// scales x differences inside an element; gradient_sca carries r_earth.
static void spherical_geometry(HostMesh &m)
{
    const int N = m.nod2D, T = m.elem2D, E = m.edge2D, nl = m.nl;
    const double C = m.cyc_x / kREarth;
    const double *X = m.coord_nod2D.data();
    auto elem_center = [&](int el, double &cx, double &cy) {
        const int *en = &m.elem2D_nodes[3 * el];
        double ax[3] = {X[2 * en[0]], X[2 * en[1]], X[2 * en[2]]};
        const double amin = std::min(ax[0], std::min(ax[1], ax[2]));
        if (C > 0.0)
            for (double &a : ax) {
                if (a - amin >= C / 2.0) a -= C;
                if (a - amin < -C / 2.0) a += C;
            }
        cx = (ax[0] + ax[1] + ax[2]) / 3.0;
        cy = (X[2 * en[0] + 1] + X[2 * en[1] + 1] + X[2 * en[2] + 1]) / 3.0;
    };
    auto edge_center = [&](int n1, int n2, double &x, double &y) {
        double a = X[2 * n1], b = X[2 * n2];
        if (C > 0.0) {
            if (a - b > C / 2.0) a -= C;
            if (a - b < -C / 2.0) b -= C;
        }
        x = 0.5 * (a + b);
        y = 0.5 * (X[2 * n1 + 1] + X[2 * n2 + 1]);
    };

    // edge orientation check: edge_tri(1) must lie left of edges(1)->edges(2)
    for (int e = 0; e < E; ++e) {
        const int n1 = m.edges[2 * e], n2 = m.edges[2 * e + 1];
        const double ex = trim(X[2 * n2] - X[2 * n1], C), ey = X[2 * n2 + 1] - X[2 * n1 + 1];
        double cx, cy;
        elem_center(m.edge_tri[2 * e], cx, cy);
        const double rx = trim(cx - X[2 * n1], C), ry = cy - X[2 * n1 + 1];
        if (ex * ry - ey * rx <= 0.0) throw std::runtime_error("mesh: edge_tri(1) is not left of its edge");
    }

    // mesh_areas
    m.elem_area.assign(T, 0.0);
    for (int el = 0; el < T; ++el) {
        const int *en = &m.elem2D_nodes[3 * el];
        const double ay = std::cos((X[2 * en[0] + 1] + X[2 * en[1] + 1] + X[2 * en[2] + 1]) / 3.0);
        double a0 = trim(X[2 * en[1]] - X[2 * en[0]], C), a1 = X[2 * en[1] + 1] - X[2 * en[0] + 1];
        double b0 = trim(X[2 * en[2]] - X[2 * en[0]], C), b1 = X[2 * en[2] + 1] - X[2 * en[0] + 1];
        a0 *= ay;
        b0 *= ay;
        m.elem_area[el] = 0.5 * std::fabs(a0 * b1 - b0 * a1);
    }
    m.area.assign((size_t)nl * N, 0.0);
    m.areasvol.assign((size_t)nl * N, 0.0);
    for (int n = 0; n < N; ++n)
        for (int j = 0; j < m.nod_in_elem2D_num[n]; ++j) {
            const int el = m.nod_in_elem2D[(size_t)n * m.maxAdj + j];
            for (int nz = m.ulevels[el]; nz <= m.nlevels[el] - 1; ++nz)
                m.area[(size_t)n * nl + nz - 1] += m.elem_area[el] / 3.0;
        }
    for (int n = 0; n < N; ++n)
        for (int nz = m.ulevels_nod2D[n]; nz <= m.nlevels_nod2D[n] - 1; ++nz)
            m.areasvol[(size_t)n * nl + nz - 1] = m.area[(size_t)n * nl + nz - 1];
    for (double &a : m.elem_area) a = a * kREarth * kREarth;
    for (double &a : m.area) a = a * kREarth * kREarth;
    for (double &a : m.areasvol) a = a * kREarth * kREarth;

    // mesh_auxiliary_arrays
    std::vector<double> center_x(T), center_y(T);
    m.elem_cos.assign(T, 1.0);
    for (int el = 0; el < T; ++el) {
        elem_center(el, center_x[el], center_y[el]);
        m.elem_cos[el] = std::cos(center_y[el]);
    }
    m.edge_dxdy.assign(2 * (size_t)E, 0.0);
    m.edge_cross_dxdy.assign(4 * (size_t)E, 0.0);
    for (int e = 0; e < E; ++e) {
        const int n1 = m.edges[2 * e], n2 = m.edges[2 * e + 1];
        m.edge_dxdy[2 * e] = trim(X[2 * n2] - X[2 * n1], C);
        m.edge_dxdy[2 * e + 1] = X[2 * n2 + 1] - X[2 * n1 + 1];
        double ax, ay;
        edge_center(n1, n2, ax, ay);
        for (int side = 0; side < 2; ++side) {
            const int el = m.edge_tri[2 * e + side];
            if (el < 0) continue;
            double b0 = trim(center_x[el] - ax, C), b1 = center_y[el] - ay;
            b0 = b0 * m.elem_cos[el];
            m.edge_cross_dxdy[4 * e + 2 * side] = b0 * kREarth;
            m.edge_cross_dxdy[4 * e + 2 * side + 1] = b1 * kREarth;
        }
    }
    m.gradient_sca.assign(6 * (size_t)T, 0.0);
    for (int el = 0; el < T; ++el) {
        const int *en = &m.elem2D_nodes[3 * el];
        const double dX31 = m.elem_cos[el] * trim(X[2 * en[2]] - X[2 * en[0]], C);
        const double dX21 = m.elem_cos[el] * trim(X[2 * en[1]] - X[2 * en[0]], C);
        const double dY31 = X[2 * en[2] + 1] - X[2 * en[0] + 1];
        const double dY21 = X[2 * en[1] + 1] - X[2 * en[0] + 1];
        const double dfactor = -0.5 * kREarth / m.elem_area[el];
        double *g = &m.gradient_sca[6 * el];
        g[0] = (-dY31 + dY21) * dfactor;
        g[1] = dY31 * dfactor;
        g[2] = -dY21 * dfactor;
        g[3] = (dX31 - dX21) * dfactor;
        g[4] = -dX31 * dfactor;
        g[5] = dX21 * dfactor;
    }
}

void derive_all(HostMesh &m)
{
    check_derivable(m);

    derive_levels_min(m);
    if (m.cartesian) cartesian_geometry(m);
    else spherical_geometry(m);
    derive_columns(m);
    derive_up_down(m);
    derive_boundary_lay(m);
    derive_node_edges(m);
}

void check_derivable(const HostMesh &m)
{
    if (m.nl < 3) throw std::runtime_error("vertical grid not set");
    if ((int)m.nlevels.size() != m.elem2D) throw std::runtime_error("levels not set");
}

void derive_columns(HostMesh &m)
{
    const int N = m.nod2D, T = m.elem2D, nl = m.nl;
    if ((int)m.depth.size() != N) {  // generated meshes: the bottom of each node's deepest layer
        m.depth.assign(N, 0.0);
        for (int n = 0; n < N; ++n) m.depth[n] = m.zbar[m.nlevels_nod2D[n] - 1];
    }

    // ALE thicknesses and 3-D depth arrays (src/fesom_analytic.F90:226-244)
    m.hnode.assign((size_t)(nl - 1) * N, 0.0);
    m.hnode_new.assign((size_t)(nl - 1) * N, 0.0);
    m.helem.assign((size_t)(nl - 1) * T, 0.0);
    m.Z_3d_n.assign((size_t)(nl - 1) * N, 0.0);
    m.zbar_3d_n.assign((size_t)nl * N, 0.0);
    for (int n = 0; n < N; ++n) {
        for (int k = 0; k < nl - 1; ++k) {
            m.hnode[(size_t)n * (nl - 1) + k] = m.zbar[k] - m.zbar[k + 1];
            m.hnode_new[(size_t)n * (nl - 1) + k] = m.hnode[(size_t)n * (nl - 1) + k];
            m.Z_3d_n[(size_t)n * (nl - 1) + k] = m.Z[k];
        }
        for (int k = 0; k < nl; ++k) m.zbar_3d_n[(size_t)n * nl + k] = m.zbar[k];
    }
    for (int el = 0; el < T; ++el)
        for (int k = 0; k < nl - 1; ++k) m.helem[(size_t)el * (nl - 1) + k] = m.zbar[k] - m.zbar[k + 1];
}

void derive_up_down(HostMesh &m)
{
    const int E = m.edge2D;
    // find_up_downwind_triangles (lib/oce_muscl_adv.F90:166-357) on the
    // radian coordinates; both axes trimmed; the last matching element wins
    {
        const double cx = m.cyc_x / kREarth, cy = m.cyc_y / kREarth;
        const double *X = m.coord_nod2D.data();
        m.edge_up_dn_tri.assign(2 * (size_t)E, -1);
        for (int n = 0; n < E; ++n) {
            int ed[2] = {m.edges[2 * n], m.edges[2 * n + 1]};
            double x0 = trim(X[2 * ed[1]] - X[2 * ed[0]], cx), x1 = trim(X[2 * ed[1] + 1] - X[2 * ed[0] + 1], cy);
            for (int side = 0; side < 2; ++side) {
                int node = ed[side];
                double xs0 = side == 0 ? -x0 : x0, xs1 = side == 0 ? -x1 : x1;
                for (int k = 0; k < m.nod_in_elem2D_num[node]; ++k) {
                    int elem = m.nod_in_elem2D[(size_t)node * m.maxAdj + k];
                    const int *en = &m.elem2D_nodes[3 * elem];
                    int i0, i1, i2;
                    if (en[0] == node) { i0 = 0; i1 = 1; i2 = 2; }
                    else if (en[1] == node) { i0 = 1; i1 = 0; i2 = 2; }
                    else { i0 = 2; i1 = 0; i2 = 1; }
                    double b0 = trim(X[2 * en[i1]] - X[2 * en[i0]], cx);
                    double b1 = trim(X[2 * en[i1] + 1] - X[2 * en[i0] + 1], cy);
                    double c0 = trim(X[2 * en[i2]] - X[2 * en[i0]], cx);
                    double c1 = trim(X[2 * en[i2] + 1] - X[2 * en[i0] + 1], cy);
                    double cr = c0 * c0 + c1 * c1;
                    double bx = (b0 * c0 + b1 * c1) / cr;
                    double by = (-b0 * c1 + b1 * c0) / cr;
                    double xx = (xs0 * c0 + xs1 * c1) / cr;
                    double xy = (-xs0 * c1 + xs1 * c0) / cr;
                    double ab = std::atan2(by, bx), ax = std::atan2(xy, xx);
                    if ((ab > 0.0) && (ax > 0.0) && (ax < ab)) { m.edge_up_dn_tri[2 * n + side] = elem; continue; }
                    if ((ab < 0.0) && (ax < 0.0) && (ax > ab)) { m.edge_up_dn_tri[2 * n + side] = elem; continue; }
                    if ((ab == ax) || (ax == 0.0)) { m.edge_up_dn_tri[2 * n + side] = elem; continue; }
                }
            }
        }
    }
}

void derive_levels_min(HostMesh &m)
{
    // find_levels_min_e2n (lib/oce_mesh.F90:1522-1560)
    const int N = m.nod2D;
    m.nlevels_nod2D_min.assign(N, 0);
    m.ulevels_nod2D_max.assign(N, 0);
    for (int n = 0; n < N; ++n) {
        int mn = 1 << 30, mx = 0;
        for (int k = 0; k < m.nod_in_elem2D_num[n]; ++k) {
            int el = m.nod_in_elem2D[(size_t)n * m.maxAdj + k];
            mn = std::min(mn, m.nlevels[el]);
            mx = std::max(mx, m.ulevels[el]);
        }
        m.nlevels_nod2D_min[n] = mn;
        m.ulevels_nod2D_max[n] = mx;
    }
}

void derive_boundary_lay(HostMesh &m)
{
    // nboundary_lay (lib/oce_muscl_adv.F90:68-104)
    m.nboundary_lay.assign(m.nod2D, m.nl - 1);
    for (int e = 0; e < m.edge2D; ++e) {
        int n1 = m.edges[2 * e], n2 = m.edges[2 * e + 1];
        int t1 = m.edge_tri[2 * e], t2 = m.edge_tri[2 * e + 1];
        if (t1 < 0 || t2 < 0) {
            m.nboundary_lay[n1] = 0;
            m.nboundary_lay[n2] = 0;
        } else {
            int mn = std::min(m.nlevels[t1], m.nlevels[t2]) - 1;
            m.nboundary_lay[n1] = std::min(m.nboundary_lay[n1], mn);
            m.nboundary_lay[n2] = std::min(m.nboundary_lay[n2], mn);
        }
    }
}

void derive_node_edges(HostMesh &m)
{
    // node -> incident edges, sorted by edge index (the order of the
    // reference's edge-loop scatters)
    const int N = m.nod2D, E = m.edge2D;
    m.n2e_ptr.assign(N + 1, 0);
    for (int e = 0; e < E; ++e) {
        m.n2e_ptr[m.edges[2 * e] + 1]++;
        m.n2e_ptr[m.edges[2 * e + 1] + 1]++;
    }
    for (int n = 0; n < N; ++n) {
        const int deg = std::max(m.n2e_ptr[n + 1], m.nod_in_elem2D_num[n]);
        if (deg > kMaxNodeDegree)
            throw std::runtime_error("node " + std::to_string(n + 1) + " has " + std::to_string(deg) +
                                     " edges or elements; at most " + std::to_string(kMaxNodeDegree) +
                                     " are supported");
        m.n2e_ptr[n + 1] += m.n2e_ptr[n];
    }
    m.n2e.assign(m.n2e_ptr[N], 0);
    std::vector<int> fill(m.n2e_ptr.begin(), m.n2e_ptr.end() - 1);
    for (int e = 0; e < E; ++e) {
        m.n2e[fill[m.edges[2 * e]]++] = e + 1;
        m.n2e[fill[m.edges[2 * e + 1]]++] = -(e + 1);
    }
}

bool node_fan(const HostMesh &m, int n, std::vector<int> &ring, std::vector<int> &fan)
{
    const int ne = m.nod_in_elem2D_num[n];
    std::vector<int> ea(ne), eb(ne), el(ne);
    for (int j = 0; j < ne; ++j) {
        el[j] = m.nod_in_elem2D[(size_t)n * m.maxAdj + j];
        int o[2], q = 0;
        for (int v = 0; v < 3; ++v) {
            const int node = m.elem2D_nodes[3 * el[j] + v];
            if (node != n && q < 2) o[q++] = node;
        }
        if (q != 2) return false;  // degenerate element, or n is not one of its nodes
        ea[j] = o[0];
        eb[j] = o[1];
    }
    // start at a ring node met by one element only (open fan), else anywhere
    int start = ne ? ea[0] : -1;
    for (int j = 0; j < ne; ++j)
        for (int s = 0; s < 2; ++s) {
            const int v = s ? eb[j] : ea[j];
            int cnt = 0;
            for (int i = 0; i < ne; ++i) cnt += (ea[i] == v) + (eb[i] == v);
            if (cnt == 1) start = v;
        }
    ring.assign(1, start);
    fan.clear();
    std::vector<char> used(ne, 0);
    int cur = start;
    for (int it = 0; it < ne; ++it) {
        int pick = -1, nxt = -1;
        for (int i = 0; i < ne && pick < 0; ++i) {
            if (used[i]) continue;
            if (ea[i] == cur) { pick = i; nxt = eb[i]; }
            else if (eb[i] == cur) { pick = i; nxt = ea[i]; }
        }
        if (pick < 0) return false;
        used[pick] = 1;
        fan.push_back(el[pick]);
        cur = nxt;
        if (nxt == start && it == ne - 1) break;  // closed fan
        ring.push_back(nxt);
    }
    if ((int)fan.size() != ne) return false;
    if ((int)ring.size() > (int)fan.size()) fan.push_back(-1);  // open fan: no element after the last ring node
    return true;
}

void complete_mesh(HostMesh &m, const std::function<bool(const std::string &)> &given)
{
    static const char *required[] = {"edges", "edge_tri", "elem2D_nodes", "nlevels", "nlevels_nod2D",
                                     "nod_in_elem2D_num", "nod_in_elem2D", "zbar", "elem_cos", "elem_area",
                                     "gradient_sca", "edge_dxdy", "edge_cross_dxdy", "area", "hnode", "helem"};
    for (const char *r : required)
        if (!given(r)) throw std::runtime_error(std::string("mesh array ") + r + " was not set");
    const int N = m.nod2D, T = m.elem2D, E = m.edge2D, nl = m.nl;
    if (nl < 3 || N < 1 || T < 1 || E < 1) throw std::runtime_error("mesh dimensions must be positive, nl >= 3");
    if ((int)m.zbar.size() != nl) throw std::runtime_error("mesh array zbar: wrong size");
    // defaults for what FESOM derives from the arrays above: no cavities,
    // areasvol = area, time-level thickness = hnode, 3-D depths from zbar
    if (!given("ulevels")) m.ulevels.assign(T, 1);
    if (!given("ulevels_nod2D")) m.ulevels_nod2D.assign(N, 1);
    if (!given("edge_up_dn_tri")) m.edge_up_dn_tri.assign(2 * (size_t)E, -1);
    if (!given("areasvol")) m.areasvol = m.area;
    if (!given("hnode_new")) m.hnode_new = m.hnode;
    if (!given("coord_nod2D")) m.coord_nod2D.assign(2 * (size_t)N, 0.0);
    if (!given("Z")) {
        m.Z.resize(nl - 1);
        for (int k = 0; k < nl - 1; ++k) m.Z[k] = 0.5 * (m.zbar[k] + m.zbar[k + 1]);
    }
    if (!given("Z_3d_n")) {
        m.Z_3d_n.resize((size_t)(nl - 1) * N);
        for (int n = 0; n < N; ++n) std::copy(m.Z.begin(), m.Z.end(), m.Z_3d_n.begin() + (size_t)n * (nl - 1));
    }
    if (!given("zbar_3d_n")) {
        m.zbar_3d_n.resize((size_t)nl * N);
        for (int n = 0; n < N; ++n) std::copy(m.zbar.begin(), m.zbar.end(), m.zbar_3d_n.begin() + (size_t)n * nl);
    }
    // placeholders that pass the range checks, replaced below by derivation
    // from the (then validated) topology
    if (!given("nlevels_nod2D_min")) m.nlevels_nod2D_min = m.nlevels_nod2D;
    if (!given("ulevels_nod2D_max")) m.ulevels_nod2D_max.assign(N, 1);
    if (!given("nboundary_lay")) m.nboundary_lay.assign(N, nl - 1);
    if (!given("depth")) m.depth.assign(N, 0.0);
    m.n2e_ptr.assign(N + 1, 0);
    check_mesh(m);
    if (!given("nlevels_nod2D_min") || !given("ulevels_nod2D_max")) {
        std::vector<int> keep_min = m.nlevels_nod2D_min, keep_max = m.ulevels_nod2D_max;
        derive_levels_min(m);
        if (given("nlevels_nod2D_min")) m.nlevels_nod2D_min = keep_min;
        if (given("ulevels_nod2D_max")) m.ulevels_nod2D_max = keep_max;
    }
    if (!given("nboundary_lay")) derive_boundary_lay(m);
    if (!given("depth"))
        for (int n = 0; n < N; ++n) m.depth[n] = m.zbar[m.nlevels_nod2D[n] - 1];
    derive_node_edges(m);
    m.coord.resize(2 * (size_t)N);
    for (size_t i = 0; i < m.coord.size(); ++i) m.coord[i] = m.coord_nod2D[i] * kREarth;
    check_mesh(m);
}

void check_mesh(const HostMesh &m)
{
    const int N = m.nod2D, T = m.elem2D, E = m.edge2D, nl = m.nl, A = m.maxAdj;
    auto bad = [](const std::string &what, long long i) {
        throw std::runtime_error("mesh array " + what + ": bad entry at " + std::to_string(i));
    };
    auto size_is = [](const char *what, size_t got, long long want) {
        if ((long long)got != want)
            throw std::runtime_error(std::string("mesh array ") + what + ": " + std::to_string(got) +
                                     " entries, expected " + std::to_string(want));
    };
    if (nl < 3 || N < 1 || T < 1 || E < 1 || A < 1) throw std::runtime_error("mesh dimensions must be positive, nl >= 3");
    size_is("edges", m.edges.size(), 2ll * E);
    size_is("edge_tri", m.edge_tri.size(), 2ll * E);
    size_is("elem2D_nodes", m.elem2D_nodes.size(), 3ll * T);
    size_is("nlevels", m.nlevels.size(), T);
    size_is("ulevels", m.ulevels.size(), T);
    size_is("nlevels_nod2D", m.nlevels_nod2D.size(), N);
    size_is("ulevels_nod2D", m.ulevels_nod2D.size(), N);
    size_is("nlevels_nod2D_min", m.nlevels_nod2D_min.size(), N);
    size_is("ulevels_nod2D_max", m.ulevels_nod2D_max.size(), N);
    size_is("nod_in_elem2D_num", m.nod_in_elem2D_num.size(), N);
    size_is("nod_in_elem2D", m.nod_in_elem2D.size(), (long long)A * N);
    size_is("edge_up_dn_tri", m.edge_up_dn_tri.size(), 2ll * E);
    size_is("nboundary_lay", m.nboundary_lay.size(), N);
    for (long long i = 0; i < 2ll * E; ++i)
        if (m.edges[i] < 0 || m.edges[i] >= N) bad("edges", i);
    for (int e = 0; e < E; ++e) {
        if (m.edges[2 * e] == m.edges[2 * e + 1]) bad("edges", 2ll * e);
        if (m.edge_tri[2 * e] < 0 || m.edge_tri[2 * e] >= T) bad("edge_tri", 2ll * e);
        if (m.edge_tri[2 * e + 1] < -1 || m.edge_tri[2 * e + 1] >= T) bad("edge_tri", 2ll * e + 1);
        for (int s = 0; s < 2; ++s)
            if (m.edge_up_dn_tri[2 * e + s] < -1 || m.edge_up_dn_tri[2 * e + s] >= T) bad("edge_up_dn_tri", 2ll * e + s);
    }
    for (int el = 0; el < T; ++el) {
        if (m.ulevels[el] < 1 || m.nlevels[el] <= m.ulevels[el] || m.nlevels[el] > nl) bad("nlevels/ulevels", el);
        for (int v = 0; v < 3; ++v) {
            const int n = m.elem2D_nodes[3 * el + v];
            if (n < 0 || n >= N) bad("elem2D_nodes", 3ll * el + v);
            // an element's layers lie inside each of its nodes' columns
            // (FESOM's node levels are the max over the adjacent elements)
            if (m.nlevels[el] > m.nlevels_nod2D[n] || m.ulevels[el] < m.ulevels_nod2D[n]) bad("nlevels_nod2D", n);
        }
    }
    for (int n = 0; n < N; ++n) {
        if (m.ulevels_nod2D[n] < 1 || m.nlevels_nod2D[n] <= m.ulevels_nod2D[n] || m.nlevels_nod2D[n] > nl)
            bad("nlevels_nod2D/ulevels_nod2D", n);
        if (m.nlevels_nod2D_min[n] < 1 || m.nlevels_nod2D_min[n] > m.nlevels_nod2D[n]) bad("nlevels_nod2D_min", n);
        if (m.nboundary_lay[n] < 0 || m.nboundary_lay[n] > nl - 1) bad("nboundary_lay", n);
        const int k = m.nod_in_elem2D_num[n];
        if (k < 0 || k > A) bad("nod_in_elem2D_num", n);
        for (int j = 0; j < k; ++j) {
            const int el = m.nod_in_elem2D[(size_t)n * A + j];
            if (el < 0 || el >= T) bad("nod_in_elem2D", (long long)n * A + j);
            // the nodal gradient reads the element's layers in the node's column
            if (m.nlevels[el] > m.nlevels_nod2D[n]) bad("nod_in_elem2D", (long long)n * A + j);
        }
    }
    for (int e = 0; e < E; ++e)
        for (int s = 0; s < 2; ++s) {
            const int el = m.edge_tri[2 * e + s];
            if (el < 0) continue;
            for (int q = 0; q < 2; ++q)
                if (m.nlevels[el] > m.nlevels_nod2D[m.edges[2 * e + q]]) bad("edge_tri", 2ll * e + s);
        }
    size_is("zbar", m.zbar.size(), nl);
    size_is("Z", m.Z.size(), nl - 1);
    size_is("depth", m.depth.size(), N);
    size_is("coord_nod2D", m.coord_nod2D.size(), 2ll * N);
    size_is("elem_cos", m.elem_cos.size(), T);
    size_is("elem_area", m.elem_area.size(), T);
    size_is("gradient_sca", m.gradient_sca.size(), 6ll * T);
    size_is("edge_dxdy", m.edge_dxdy.size(), 2ll * E);
    size_is("edge_cross_dxdy", m.edge_cross_dxdy.size(), 4ll * E);
    size_is("area", m.area.size(), (long long)nl * N);
    size_is("areasvol", m.areasvol.size(), (long long)nl * N);
    size_is("hnode", m.hnode.size(), (long long)(nl - 1) * N);
    size_is("hnode_new", m.hnode_new.size(), (long long)(nl - 1) * N);
    size_is("helem", m.helem.size(), (long long)(nl - 1) * T);
    size_is("Z_3d_n", m.Z_3d_n.size(), (long long)(nl - 1) * N);
    size_is("zbar_3d_n", m.zbar_3d_n.size(), (long long)nl * N);
    size_is("n2e_ptr", m.n2e_ptr.size(), N + 1ll);
}

double make_velocity(const HostMesh &m, double u_bg, double v_bg, double cell_amp, double shear,
                     double cfl, std::vector<double> &uv, std::vector<double> &w)
{
    const int N = m.nod2D, T = m.elem2D, E = m.edge2D, nl = m.nl, L = nl - 1;
    const double PI = 3.14159265358979323846;
    const double Lx = m.cyc_x > 0 ? m.cyc_x : 1.0, Ly = m.cyc_y > 0 ? m.cyc_y : 1.0;
    // reference column depth: mean over elements
    double Href = 0.0;
    std::vector<double> H(T, 0.0);
    for (int el = 0; el < T; ++el) {
        for (int k = 0; k < m.nlevels[el] - 1; ++k) H[el] += m.helem[(size_t)el * L + k];
        Href += H[el];
    }
    Href /= T;
    // transport streamfunction: rotation cells; closed meshes get psi = 0 on the rim
    std::vector<double> psi(N, 0.0);
    if (m.cartesian) {
        double xm = 0, ym = 0;
        for (int q = 0; q < N; ++q) { xm = std::max(xm, m.coord[2 * q]); ym = std::max(ym, m.coord[2 * q + 1]); }
        for (int n = 0; n < N; ++n) {
            double x = m.coord[2 * n], y = m.coord[2 * n + 1];
            double sx, sy;
            if (m.cyc_x > 0) { sx = std::sin(2.0 * PI * x / Lx); sy = std::sin(2.0 * PI * y / Ly); }
            else { sx = std::sin(PI * x / xm); sy = std::sin(PI * y / ym); }
            psi[n] = cell_amp * Href * (Lx / (2.0 * PI)) * sx * sy;
        }
    } else {
        // sphere: a zonal current u_bg*cos(lat) plus wave-2 cells, tapered to
        // psi = 0 on every coast (graph distance 0..3 from a boundary node) so
        // that no transport crosses a coast and the column divergence vanishes
        std::vector<int> hop(N, 1 << 20), queue;
        for (int e = 0; e < E; ++e)
            if (m.edge_tri[2 * e + 1] < 0)
                for (int q = 0; q < 2; ++q)
                    if (hop[m.edges[2 * e + q]] != 0) {
                        hop[m.edges[2 * e + q]] = 0;
                        queue.push_back(m.edges[2 * e + q]);
                    }
        for (size_t h = 0; h < queue.size(); ++h) {
            const int n = queue[h];
            for (int i = m.n2e_ptr[n]; i < m.n2e_ptr[n + 1]; ++i) {
                const int e = std::abs(m.n2e[i]) - 1;
                const int o = m.edges[2 * e] == n ? m.edges[2 * e + 1] : m.edges[2 * e];
                if (hop[o] > hop[n] + 1) {
                    hop[o] = hop[n] + 1;
                    queue.push_back(o);
                }
            }
        }
        for (int n = 0; n < N; ++n) {
            const double lon = m.coord_nod2D[2 * n], lat = m.coord_nod2D[2 * n + 1];
            const double taper = std::min(1.0, hop[n] / 3.0);
            psi[n] = taper * Href * kREarth *
                     (-u_bg * std::sin(lat) + cell_amp * std::cos(lat) * std::cos(lat) * std::sin(2.0 * lon));
        }
    }
    uv.assign(2 * (size_t)L * T, 0.0);
    double umax = 0.0;
    for (int el = 0; el < T; ++el) {
        const int *en = &m.elem2D_nodes[3 * el];
        const double *g = &m.gradient_sca[6 * el];
        double px = g[0] * psi[en[0]] + g[1] * psi[en[1]] + g[2] * psi[en[2]];
        double py = g[3] * psi[en[0]] + g[4] * psi[en[1]] + g[5] * psi[en[2]];
        double bg = (m.cartesian && m.cyc_x > 0) ? Href : 0.0;  // a uniform current only on periodic planes
        double tx = u_bg * bg - py, ty = v_bg * bg + px;
        int nk = m.nlevels[el] - 1;
        // vertical shear profile with zero thickness-weighted mean
        std::vector<double> c(nk, 0.0);
        double depth = 0.0, mean = 0.0;
        for (int k = 0; k < nk; ++k) {
            double h = m.helem[(size_t)el * L + k];
            c[k] = std::cos(PI * (depth + 0.5 * h) / H[el]);
            depth += h;
            mean += h * c[k];
        }
        mean /= H[el];
        for (int k = 0; k < nk; ++k) {
            double f = (1.0 + shear * (c[k] - mean)) / H[el];
            uv[((size_t)el * L + k) * 2] = tx * f;
            uv[((size_t)el * L + k) * 2 + 1] = ty * f;
            umax = std::max(umax, std::hypot(tx * f, ty * f));
        }
    }
    // W from discrete continuity of the flux-form update (T == 1 is preserved)
    std::vector<double> inflow((size_t)L * N, 0.0);
    for (int e = 0; e < E; ++e) {
        int n1 = m.edges[2 * e], n2 = m.edges[2 * e + 1];
        int el1 = m.edge_tri[2 * e], el2 = m.edge_tri[2 * e + 1];
        int nl1 = m.nlevels[el1] - 1, nl2 = el2 >= 0 ? m.nlevels[el2] - 1 : 0;
        int nl12 = std::min(nl1, nl2);
        const double *cd = &m.edge_cross_dxdy[4 * e];
        for (int nz = 1; nz <= std::max(nl1, nl2); ++nz) {
            double vf = 0.0;
            if (nz <= nl1) {
                const double *u = &uv[((size_t)el1 * L + nz - 1) * 2];
                vf += (-u[1] * cd[0] + u[0] * cd[1]) * m.helem[(size_t)el1 * L + nz - 1];
            }
            if (el2 >= 0 && (nz <= nl12 || nz > nl1)) {
                const double *u = &uv[((size_t)el2 * L + nz - 1) * 2];
                vf += (u[1] * cd[2] - u[0] * cd[3]) * m.helem[(size_t)el2 * L + nz - 1];
            }
            inflow[(size_t)n1 * L + nz - 1] -= vf;
            inflow[(size_t)n2 * L + nz - 1] += vf;
        }
    }
    w.assign((size_t)nl * N, 0.0);
    for (int n = 0; n < N; ++n) {
        int nzmax = m.nlevels_nod2D[n];
        double *wc = &w[(size_t)n * nl];
        const double *ar = &m.area[(size_t)n * nl];
        wc[nzmax - 1] = 0.0;
        for (int nz = nzmax - 1; nz >= 1; --nz)
            wc[nz - 1] = (wc[nz] * ar[nz] + inflow[(size_t)n * L + nz - 1]) / ar[nz - 1];
        wc[0] = 0.0;  // rigid lid: the column-integrated divergence vanishes
    }
    double dxmin = 1e300;
    for (int e = 0; e < E; ++e) {
        const double c = m.elem_cos[m.edge_tri[2 * e]];  // 1 on planes
        dxmin = std::min(dxmin, std::hypot(m.edge_dxdy[2 * e] * c, m.edge_dxdy[2 * e + 1]) * kREarth);
    }
    return umax > 0.0 ? cfl * dxmin / umax : 0.0;
}

int64_t active_node_cells(const HostMesh &m)
{
    int64_t c = 0;
    for (int n = 0; n < m.nod2D; ++n) c += m.nlevels_nod2D[n] - m.ulevels_nod2D[n];
    return c;
}

int64_t active_elem_cells(const HostMesh &m)
{
    int64_t c = 0;
    for (int el = 0; el < m.elem2D; ++el) c += m.nlevels[el] - m.ulevels[el];
    return c;
}

bool named_mesh_array(const HostMesh &m, const std::string &name, const void *&data, long long &count, int &kind)
{
#define IA(NAME, VEC) if (name == NAME) { data = VEC.data(); count = (long long)VEC.size(); kind = 0; return true; }
#define DA(NAME, VEC) if (name == NAME) { data = VEC.data(); count = (long long)VEC.size(); kind = 1; return true; }
    IA("edges", m.edges) IA("edge_tri", m.edge_tri) IA("elem2D_nodes", m.elem2D_nodes)
    IA("nlevels", m.nlevels) IA("ulevels", m.ulevels) IA("nlevels_nod2D", m.nlevels_nod2D)
    IA("ulevels_nod2D", m.ulevels_nod2D) IA("nlevels_nod2D_min", m.nlevels_nod2D_min)
    IA("ulevels_nod2D_max", m.ulevels_nod2D_max) IA("nod_in_elem2D_num", m.nod_in_elem2D_num)
    IA("nod_in_elem2D", m.nod_in_elem2D) IA("edge_up_dn_tri", m.edge_up_dn_tri)
    IA("nboundary_lay", m.nboundary_lay)
    DA("coord", m.coord) DA("coord_nod2D", m.coord_nod2D) DA("zbar", m.zbar) DA("Z", m.Z) DA("depth", m.depth)
    DA("elem_cos", m.elem_cos) DA("elem_area", m.elem_area) DA("gradient_sca", m.gradient_sca)
    DA("edge_dxdy", m.edge_dxdy) DA("edge_cross_dxdy", m.edge_cross_dxdy) DA("area", m.area)
    DA("areasvol", m.areasvol) DA("hnode", m.hnode) DA("hnode_new", m.hnode_new) DA("helem", m.helem)
    DA("Z_3d_n", m.Z_3d_n) DA("zbar_3d_n", m.zbar_3d_n)
#undef IA
#undef DA
    return false;
}

}  // namespace fta
