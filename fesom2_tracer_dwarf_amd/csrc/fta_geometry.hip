// fta_geometry.hip -- mesh geometry as GPU kernels (setup time, one thread
// per element, edge or node):
//   mesh_areas                  lib/oce_mesh.F90:2043-2309 (no cavity)
//   mesh_auxiliary_arrays       lib/oce_mesh.F90:2313-2534
// Each kernel evaluates the host derivation's expressions in the same order
// (no contraction), so planar meshes come out bit for bit as on the host; on
// the sphere the device cos can differ from the host's in the last bit.
// find_up_downwind_triangles stays on the host: on regular meshes the edge
// direction is often collinear with an element side, and which neighbour wins
// such a tie depends on the last bit of atan2, so a device atan2 would pick
// different, equally valid, triangles than the host's.
#include <hip/hip_runtime.h>

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "fta_mesh.h"

namespace fta {

namespace {

#define GEO_OK(call)                                                                        \
    do {                                                                                    \
        hipError_t _e = (call);                                                             \
        if (_e != hipSuccess)                                                               \
            throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(_e) +    \
                                     " at " #call);                                         \
    } while (0)

constexpr int kGeoBlock = 256;

__device__ __forceinline__ double trim_d(double d, double cyc)
{
    if (cyc > 0.0) {
        if (d > cyc / 2.0) d -= cyc;
        if (d < -cyc / 2.0) d += cyc;
    }
    return d;
}

struct GeoMesh {
    const double *X;      // (2,N): metres (planar) or radians (sphere)
    const int *en;        // (3,T)
    const int *edges;     // (2,E)
    const int *etri;      // (2,E)
    const int *nie;       // (maxAdj,N)
    const int *nien;      // (N)
    const int *nlev, *ulev;    // (T)
    const int *nlevn, *ulevn;  // (N)
    int N, T, E, nl, maxAdj;
    double cx, cy;        // cyclic lengths in X's units (0: not cyclic); the sphere trims x only
};

// element centre: planar -- node 1 plus the mean of the trimmed offsets;
// sphere -- elem_center (lib/oce_mesh.F90:2001-2021), x shifted next to the
// smallest longitude
template <bool SPH>
__device__ __forceinline__ void centre(const GeoMesh &g, int el, double &cx, double &cy)
{
    const int a = g.en[3 * el], b = g.en[3 * el + 1], c = g.en[3 * el + 2];
    if constexpr (!SPH) {
        const double x0 = g.X[2 * a], y0 = g.X[2 * a + 1];
        cx = x0 + (trim_d(g.X[2 * b] - x0, g.cx) + trim_d(g.X[2 * c] - x0, g.cx)) / 3.0;
        cy = y0 + (trim_d(g.X[2 * b + 1] - y0, g.cy) + trim_d(g.X[2 * c + 1] - y0, g.cy)) / 3.0;
    } else {
        double ax[3] = {g.X[2 * a], g.X[2 * b], g.X[2 * c]};
        const double amin = fmin(ax[0], fmin(ax[1], ax[2]));
        if (g.cx > 0.0)
            for (int i = 0; i < 3; ++i) {
                if (ax[i] - amin >= g.cx / 2.0) ax[i] -= g.cx;
                if (ax[i] - amin < -g.cx / 2.0) ax[i] += g.cx;
            }
        cx = (ax[0] + ax[1] + ax[2]) / 3.0;
        cy = (g.X[2 * a + 1] + g.X[2 * b + 1] + g.X[2 * c + 1]) / 3.0;
    }
}

// elem_area (planar: metres^2; sphere: radians^2, scaled later), elem_cos,
// centre and gradient_sca (which needs the scaled area)
template <bool SPH>
__global__ __launch_bounds__(kGeoBlock) void k_geo_elem(GeoMesh g, double *__restrict__ elem_area,
                                                    double *__restrict__ elem_cos, double *__restrict__ ctr,
                                                    double *__restrict__ grad)
{
    const int el = blockIdx.x * kGeoBlock + threadIdx.x;
    if (el >= g.T) return;
    const int a = g.en[3 * el], b = g.en[3 * el + 1], c = g.en[3 * el + 2];
    double area, ecos;
    if constexpr (!SPH) {
        const double ax = trim_d(g.X[2 * b] - g.X[2 * a], g.cx), ay = trim_d(g.X[2 * b + 1] - g.X[2 * a + 1], g.cy);
        const double bx = trim_d(g.X[2 * c] - g.X[2 * a], g.cx), by = trim_d(g.X[2 * c + 1] - g.X[2 * a + 1], g.cy);
        area = 0.5 * fabs(ax * by - bx * ay);
        ecos = 1.0;
    } else {
        const double cy = cos((g.X[2 * a + 1] + g.X[2 * b + 1] + g.X[2 * c + 1]) / 3.0);
        double a0 = trim_d(g.X[2 * b] - g.X[2 * a], g.cx), a1 = g.X[2 * b + 1] - g.X[2 * a + 1];
        double b0 = trim_d(g.X[2 * c] - g.X[2 * a], g.cx), b1 = g.X[2 * c + 1] - g.X[2 * a + 1];
        a0 *= cy;
        b0 *= cy;
        area = 0.5 * fabs(a0 * b1 - b0 * a1);
        ecos = cy;  // cos of the centre latitude: the same argument as above
    }
    double px, py;
    centre<SPH>(g, el, px, py);
    elem_area[el] = area;
    elem_cos[el] = ecos;
    ctr[2 * el] = px;
    ctr[2 * el + 1] = py;
    double dX31, dX21, dY31, dY21, dfactor;
    if constexpr (!SPH) {
        dX31 = trim_d(g.X[2 * c] - g.X[2 * a], g.cx) * ecos;
        dX21 = trim_d(g.X[2 * b] - g.X[2 * a], g.cx) * ecos;
        dY31 = trim_d(g.X[2 * c + 1] - g.X[2 * a + 1], g.cy);
        dY21 = trim_d(g.X[2 * b + 1] - g.X[2 * a + 1], g.cy);
        dfactor = -0.5 / area;
    } else {
        dX31 = ecos * trim_d(g.X[2 * c] - g.X[2 * a], g.cx);
        dX21 = ecos * trim_d(g.X[2 * b] - g.X[2 * a], g.cx);
        dY31 = g.X[2 * c + 1] - g.X[2 * a + 1];
        dY21 = g.X[2 * b + 1] - g.X[2 * a + 1];
        dfactor = -0.5 * kREarth / (area * kREarth * kREarth);
    }
    double *o = grad + 6 * (size_t)el;
    o[0] = (-dY31 + dY21) * dfactor;
    o[1] = dY31 * dfactor;
    o[2] = -dY21 * dfactor;
    o[3] = (dX31 - dX21) * dfactor;
    o[4] = -dX31 * dfactor;
    o[5] = dX21 * dfactor;
}

// edge_dxdy, edge_cross_dxdy and the orientation check (edge_tri(1) left of
// edges(1)->edges(2)); bad[e] = 1 flags a violation
template <bool SPH>
__global__ __launch_bounds__(kGeoBlock) void k_geo_edge(GeoMesh g, const double *__restrict__ elem_cos,
                                                    const double *__restrict__ ctr, double *__restrict__ dxdy,
                                                    double *__restrict__ cross, unsigned char *__restrict__ bad)
{
    const int e = blockIdx.x * kGeoBlock + threadIdx.x;
    if (e >= g.E) return;
    const int n1 = g.edges[2 * e], n2 = g.edges[2 * e + 1];
    const double *X = g.X;
    const int t1 = g.etri[2 * e];
    double ex, ey, rx, ry;
    if constexpr (!SPH) {
        ex = trim_d(X[2 * n2] - X[2 * n1], g.cx);
        ey = trim_d(X[2 * n2 + 1] - X[2 * n1 + 1], g.cy);
        rx = trim_d(ctr[2 * t1] - X[2 * n1], g.cx);
        ry = trim_d(ctr[2 * t1 + 1] - X[2 * n1 + 1], g.cy);
    } else {
        ex = trim_d(X[2 * n2] - X[2 * n1], g.cx);
        ey = X[2 * n2 + 1] - X[2 * n1 + 1];
        rx = trim_d(ctr[2 * t1] - X[2 * n1], g.cx);
        ry = ctr[2 * t1 + 1] - X[2 * n1 + 1];
    }
    bad[e] = (ex * ry - ey * rx <= 0.0) ? 1 : 0;
    double mx, my;
    if constexpr (!SPH) {
        dxdy[2 * e] = ex / kREarth;
        dxdy[2 * e + 1] = ey / kREarth;
        mx = X[2 * n1] + 0.5 * ex;
        my = X[2 * n1 + 1] + 0.5 * ey;
    } else {
        dxdy[2 * e] = ex;
        dxdy[2 * e + 1] = ey;
        // edge_center (lib/oce_mesh.F90:2023-2041)
        double a = X[2 * n1], b = X[2 * n2];
        if (g.cx > 0.0) {
            if (a - b > g.cx / 2.0) a -= g.cx;
            if (a - b < -g.cx / 2.0) b -= g.cx;
        }
        mx = 0.5 * (a + b);
        my = 0.5 * (X[2 * n1 + 1] + X[2 * n2 + 1]);
    }
    for (int s = 0; s < 2; ++s) {
        const int el = g.etri[2 * e + s];
        double c0 = 0.0, c1 = 0.0;
        if (el >= 0) {
            if constexpr (!SPH) {
                c0 = trim_d(ctr[2 * el] - mx, g.cx) * elem_cos[el];
                c1 = trim_d(ctr[2 * el + 1] - my, g.cy);
            } else {
                const double b0 = trim_d(ctr[2 * el] - mx, g.cx) * elem_cos[el];
                c0 = b0 * kREarth;
                c1 = (ctr[2 * el + 1] - my) * kREarth;
            }
        }
        cross[4 * e + 2 * s] = c0;
        cross[4 * e + 2 * s + 1] = c1;
    }
}

// area(nz, n): elem_area/3 summed over the node's elements that have layer
// nz, in nod_in_elem2D order (the host's order, so the sums agree bit for
// bit); areasvol = area on the node's own layers; on the sphere all three are
// scaled by r_earth^2 afterwards (elem_area too). One thread per (node,
// level), consecutive threads on consecutive levels of a column.
template <bool SPH>
__global__ __launch_bounds__(kGeoBlock) void k_geo_node(GeoMesh g, const double *__restrict__ elem_area,
                                                    double *__restrict__ area, double *__restrict__ areasvol)
{
    const long long i = (long long)blockIdx.x * kGeoBlock + threadIdx.x;
    if (i >= (long long)g.N * g.nl) return;
    const int n = (int)(i / g.nl), nz = (int)(i % g.nl) + 1;
    double x = 0.0;
    for (int j = 0; j < g.nien[n]; ++j) {
        const int el = g.nie[(size_t)n * g.maxAdj + j];
        if (nz >= g.ulev[el] && nz <= g.nlev[el] - 1) x += elem_area[el] / 3.0;
    }
    double y = (nz >= g.ulevn[n] && nz <= g.nlevn[n] - 1) ? x : 0.0;
    if constexpr (SPH) {
        x = x * kREarth * kREarth;
        y = y * kREarth * kREarth;
    }
    area[i] = x;
    areasvol[i] = y;
}

__global__ __launch_bounds__(kGeoBlock) void k_geo_scale_area(double *__restrict__ elem_area, int T)
{
    const int el = blockIdx.x * kGeoBlock + threadIdx.x;
    if (el < T) elem_area[el] = elem_area[el] * kREarth * kREarth;
}

template <class T>
T *upload(const std::vector<T> &h, std::vector<void *> &owned)
{
    T *d = nullptr;
    GEO_OK(hipMalloc(&d, std::max<size_t>(1, h.size()) * sizeof(T)));
    owned.push_back(d);
    if (!h.empty()) GEO_OK(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

template <class T>
T *scratch(size_t n, std::vector<void *> &owned)
{
    T *d = nullptr;
    GEO_OK(hipMalloc(&d, std::max<size_t>(1, n) * sizeof(T)));
    owned.push_back(d);
    return d;
}

template <class T>
void download(std::vector<T> &h, const T *d, size_t n)
{
    h.resize(n);
    if (n) GEO_OK(hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost));
}

// index ranges every kernel above relies on, checked before any launch
void check_indices(const HostMesh &m)
{
    const int N = m.nod2D, T = m.elem2D, E = m.edge2D;
    auto bad = [](const char *what) { throw std::runtime_error(std::string("device geometry: bad ") + what); };
    if (N < 1 || T < 1 || E < 1) bad("mesh dimensions");
    if ((int)m.elem2D_nodes.size() != 3 * T || (int)m.edges.size() != 2 * E || (int)m.edge_tri.size() != 2 * E ||
        (int)m.nod_in_elem2D_num.size() != N || (long long)m.nod_in_elem2D.size() != (long long)m.maxAdj * N ||
        (int)m.ulevels.size() != T || (int)m.nlevels_nod2D.size() != N || (int)m.ulevels_nod2D.size() != N ||
        (int)m.coord_nod2D.size() != 2 * N || (m.cartesian && (int)m.coord.size() != 2 * N))
        bad("array sizes");
    for (int v : m.elem2D_nodes) if (v < 0 || v >= N) bad("elem2D_nodes");
    for (int v : m.edges) if (v < 0 || v >= N) bad("edges");
    for (int e = 0; e < E; ++e)
        if (m.edge_tri[2 * e] < 0 || m.edge_tri[2 * e] >= T || m.edge_tri[2 * e + 1] < -1 ||
            m.edge_tri[2 * e + 1] >= T)
            bad("edge_tri");
    for (int n = 0; n < N; ++n) {
        if (m.nod_in_elem2D_num[n] < 0 || m.nod_in_elem2D_num[n] > m.maxAdj) bad("nod_in_elem2D_num");
        for (int j = 0; j < m.nod_in_elem2D_num[n]; ++j) {
            const int el = m.nod_in_elem2D[(size_t)n * m.maxAdj + j];
            if (el < 0 || el >= T) bad("nod_in_elem2D");
        }
    }
    for (int el = 0; el < T; ++el)
        if (m.ulevels[el] < 1 || m.nlevels[el] > m.nl) bad("levels");
}

template <bool SPH>
void run(HostMesh &m)
{
    const int N = m.nod2D, T = m.elem2D, E = m.edge2D, nl = m.nl;
    std::vector<void *> owned;
    try {
        GeoMesh g{};
        g.X = upload(SPH ? m.coord_nod2D : m.coord, owned);
        g.en = upload(m.elem2D_nodes, owned);
        g.edges = upload(m.edges, owned);
        g.etri = upload(m.edge_tri, owned);
        g.nie = upload(m.nod_in_elem2D, owned);
        g.nien = upload(m.nod_in_elem2D_num, owned);
        g.nlev = upload(m.nlevels, owned);
        g.ulev = upload(m.ulevels, owned);
        g.nlevn = upload(m.nlevels_nod2D, owned);
        g.ulevn = upload(m.ulevels_nod2D, owned);
        g.N = N; g.T = T; g.E = E; g.nl = nl; g.maxAdj = m.maxAdj;
        g.cx = SPH ? m.cyc_x / kREarth : m.cyc_x;
        g.cy = SPH ? 0.0 : m.cyc_y;
        double *ea = scratch<double>(T, owned), *ec = scratch<double>(T, owned);
        double *ct = scratch<double>(2 * (size_t)T, owned), *gs = scratch<double>(6 * (size_t)T, owned);
        double *dx = scratch<double>(2 * (size_t)E, owned), *cr = scratch<double>(4 * (size_t)E, owned);
        double *ar = scratch<double>((size_t)nl * N, owned), *av = scratch<double>((size_t)nl * N, owned);
        unsigned char *bad = scratch<unsigned char>(E, owned);
        auto blocks = [](int n) { return dim3((unsigned)((n + kGeoBlock - 1) / kGeoBlock)); };
        k_geo_elem<SPH><<<blocks(T), kGeoBlock>>>(g, ea, ec, ct, gs);
        GEO_OK(hipGetLastError());
        k_geo_edge<SPH><<<blocks(E), kGeoBlock>>>(g, ec, ct, dx, cr, bad);
        GEO_OK(hipGetLastError());
        k_geo_node<SPH><<<dim3((unsigned)(((long long)N * nl + kGeoBlock - 1) / kGeoBlock)), kGeoBlock>>>(g, ea, ar, av);
        GEO_OK(hipGetLastError());
        if (SPH) {
            k_geo_scale_area<<<blocks(T), kGeoBlock>>>(ea, T);
            GEO_OK(hipGetLastError());
        }
        GEO_OK(hipDeviceSynchronize());
        std::vector<unsigned char> flags;
        download(flags, bad, E);
        for (unsigned char f : flags)
            if (f) throw std::runtime_error("mesh: edge_tri(1) is not left of its edge");
        download(m.elem_area, ea, T);
        download(m.elem_cos, ec, T);
        download(m.gradient_sca, gs, 6 * (size_t)T);
        download(m.edge_dxdy, dx, 2 * (size_t)E);
        download(m.edge_cross_dxdy, cr, 4 * (size_t)E);
        download(m.area, ar, (size_t)nl * N);
        download(m.areasvol, av, (size_t)nl * N);
    } catch (...) {
        for (void *p : owned) (void)hipFree(p);
        throw;
    }
    for (void *p : owned) (void)hipFree(p);
}

}  // namespace

void derive_all_device(HostMesh &m)
{
    check_derivable(m);
    derive_levels_min(m);
    check_indices(m);
    if (m.cartesian) run<false>(m);
    else run<true>(m);
    derive_columns(m);
    derive_up_down(m);
    derive_boundary_lay(m);
    derive_node_edges(m);
}

}  // namespace fta
