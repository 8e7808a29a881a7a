// fta_mesh.h -- host-side mesh, geometry and configuration for the tracer
// advection step. Arrays use the Fortran data model of FESOM2 (column-major,
// no padding) with 0-based entity indices and -1 for "no element"; level
// numbers (nlevels, ulevels, nboundary_lay) keep FESOM's 1-based values.
//
// Reference fields mirrored: T_MESH (lib/MOD_MESH.F90:24-179) for the
// members the hot path reads, and the work arrays edge_up_dn_tri and
// nboundary_lay (lib/MOD_TRACER.F90:30-62).
#pragma once

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

namespace fta {

constexpr double kREarth = 6367500.0;  // lib/oce_modules.F90:24

struct HostMesh {
    int nl = 0;        // number of level interfaces (layers = nl-1)
    int nod2D = 0, elem2D = 0, edge2D = 0, maxAdj = 0;
    bool cartesian = true;
    double cyc_x = 0.0, cyc_y = 0.0;  // periodic lengths in metres (0 = closed); spherical: cyclic_length * r_earth

    std::vector<double> coord;        // (2,N) metres: generator space; spherical meshes: coord_nod2D * r_earth
    std::vector<double> coord_nod2D;  // (2,N) radians, as FESOM stores it
    std::vector<int> elem2D_nodes;    // (3,T), clockwise
    std::vector<int> edges;           // (2,E)
    std::vector<int> edge_tri;        // (2,E), edge_tri(1) left of edges(1)->edges(2)

    std::vector<int> nlevels, ulevels;                 // (T)
    std::vector<int> nlevels_nod2D, ulevels_nod2D;     // (N)
    std::vector<int> nlevels_nod2D_min, ulevels_nod2D_max;
    std::vector<int> nod_in_elem2D_num;                // (N)
    std::vector<int> nod_in_elem2D;                    // (maxAdj,N)

    std::vector<double> zbar, Z;                       // (nl), (nl-1)
    std::vector<double> depth;                         // (N) bottom depth, negative (mesh files: aux3d.out)
    std::vector<double> elem_cos, elem_area;           // (T)
    std::vector<double> gradient_sca;                  // (6,T)
    std::vector<double> edge_dxdy;                     // (2,E) radians
    std::vector<double> edge_cross_dxdy;               // (4,E) metres
    std::vector<double> area, areasvol;                // (nl,N)
    std::vector<double> hnode, hnode_new;              // (nl-1,N)
    std::vector<double> helem;                         // (nl-1,T)
    std::vector<double> Z_3d_n;                        // (nl-1,N)
    std::vector<double> zbar_3d_n;                     // (nl,N)
    std::vector<int> edge_up_dn_tri;                   // (2,E)
    std::vector<int> nboundary_lay;                    // (N)

    // node -> incident edges, sorted by edge index; entry = +(e+1) when the
    // node is edges(1,e), -(e+1) when it is edges(2,e)
    std::vector<int> n2e_ptr, n2e;
};

// Doubly periodic nx x ny node rectangle of Lx x Ly metres split into right
// triangles (a re-design of generate_analytic_mesh, lib/analytic_mesh.F90:114,
// with FESOM's edge orientation). order: 0 = row-major, 1 = Hilbert curve.
void generate_periodic(HostMesh &m, int nx, int ny, double Lx, double Ly, int order);

// Closed (non-periodic) rectangle: boundary edges have edge_tri(2) = -1.
void generate_closed(HostMesh &m, int nx, int ny, double Lx, double Ly, int order);

// FESOM mesh files in directory dir (read_mesh, find_levels, test_tri,
// load_edges: lib/oce_mesh.F90:98-806,807-1001,1565-1615,1617-1849) on one
// rank: nod2d.out (lon/lat in degrees), elem2d.out, aux3d.out (nl, zbar;
// the depth column is not needed once the levels are read), elvls.out and
// nlvls.out, and edges.out / edge_tri.out / edgenum.out when present (built
// from the elements otherwise). Node coordinates are rotated to the model
// grid when force_rotation is set (Euler angles in degrees, converted to
// radians as FESOM's namelist reader does); the x axis is cyclic with
// cyclic_length_deg. The geometry is then spherical (cartesian = false).
struct FesomFileOptions {
    bool force_rotation = true;
    double alpha_deg = 50.0, beta_deg = 15.0, gamma_deg = -90.0;  // lib/gen_modules_config.F90:85-88
    double cyclic_length_deg = 360.0;                             // lib/gen_modules_config.F90:83
    bool device_geometry = false;  // derive_all_device instead of derive_all
};
void read_fesom_mesh(HostMesh &m, const std::string &dir, const FesomFileOptions &opt);

// Euler rotation of lib/gen_modules_rotate_grid.F90 (set_mesh_transform_matrix,
// g2r, r2g), angles in radians; M is the 3x3 r2g matrix, row-major M[i][j] =
// r2g_matrix(i+1, j+1).
void rotation_matrix(double alpha, double beta, double gamma, double M[9]);
void geo_to_rot(const double M[9], double glon, double glat, double &rlon, double &rlat);
void rot_to_geo(const double M[9], double rlon, double rlat, double &glon, double &glat);

// Vertical grid: zbar(1)=0, layer k thickness dz*(1+stretch*(k-1)/(nl-2)).
void set_vertical_grid(HostMesh &m, int nl, double dz, double stretch);

// Bottom levels: raw node levels drawn uniformly from [lo, hi] (seeded), element
// levels = min over its 3 nodes, node levels = max over adjacent elements
// (FESOM's find_levels invariant). lo=hi=nl gives a flat bottom.
void set_levels_random(HostMesh &m, int lo, int hi, uint64_t seed);
// Element levels given, node levels derived (max over adjacent elements).
void set_levels_from_elements(HostMesh &m, const int *nlevels_elem);

// Everything derived from coordinates, topology and levels: nod_in_elem2D,
// level min/max, mesh_areas, mesh_auxiliary_arrays, ALE thicknesses,
// up/down-wind triangles, nboundary_lay and the node->edge lists.
void derive_all(HostMesh &m);
// The same derivation with the geometry (mesh_areas, mesh_auxiliary_arrays)
// run as GPU kernels on the current device; the integer, column and
// up/down-wind steps stay on the host. Results equal derive_all's bit for
// bit on planar meshes; on the sphere elem_cos and what depends on it may
// differ in the last bits (device cos).
void derive_all_device(HostMesh &m);

// Stages of derive_all: vertical grid and levels present; depth, ALE
// thicknesses and 3-D depths from zbar; find_up_downwind_triangles.
void check_derivable(const HostMesh &m);
void derive_columns(HostMesh &m);
void derive_up_down(HostMesh &m);
// Pieces of derive_all that need only topology and levels, for meshes whose
// geometry a caller supplies: nlevels_nod2D_min / ulevels_nod2D_max,
// nboundary_lay, and the node->edge lists.
void derive_levels_min(HostMesh &m);
void derive_boundary_lay(HostMesh &m);
// Also refuses a node with more than kMaxNodeDegree edges or elements (the
// widest per-node record the GPU kernels take).
constexpr int kMaxNodeDegree = 16;
void derive_node_edges(HostMesh &m);

// The elements around node n in fan order: ring[j], ring[(j+1) mod R] and the
// node form element fan[j] (fan[j] = -1: open end of a boundary fan). False
// when nod_in_elem2D of n does not form one fan (e.g. an element list that
// misses the elements across a periodic seam).
bool node_fan(const HostMesh &m, int n, std::vector<int> &ring, std::vector<int> &fan);

// A mesh whose arrays a caller (or a restart file) supplied: given(name)
// says which of the arrays named as in tracer_get_mesh_array were set. The
// required ones must be; the rest get FESOM's derivation or default (no
// cavities, areasvol = area, hnode_new = hnode, 3-D depths from zbar,
// edge_up_dn_tri -1). Validates with check_mesh, before and after deriving.
void complete_mesh(HostMesh &m, const std::function<bool(const std::string &)> &given);

// Sizes, index ranges and the level nesting the device path relies on
// (element layers inside each of its nodes' columns); throws with the array
// and entry at fault.
void check_mesh(const HostMesh &m);

// Analytic divergence-free velocity on elements, (2,nl-1,T), plus W (nl,N)
// from discrete continuity. Returns the dt giving the requested CFL.
double make_velocity(const HostMesh &m, double u_bg, double v_bg, double cell_amp,
                     double shear, double cfl, std::vector<double> &uv, std::vector<double> &w);

// counts of active node-cells and element-cells
// a named array of m, by the names and layouts of tracer_get_mesh_array
// (coord, edges, nlevels_nod2D, area, ...): its data, entry count and kind
// (0 int32, 1 float64); false for a name m does not hold
bool named_mesh_array(const HostMesh &m, const std::string &name, const void *&data, long long &count, int &kind);

int64_t active_node_cells(const HostMesh &m);
int64_t active_elem_cells(const HostMesh &m);

}  // namespace fta
