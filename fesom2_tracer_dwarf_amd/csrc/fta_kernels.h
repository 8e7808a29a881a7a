// fta_kernels.h -- gfx950 device kernels of the tracer-advection step.
//
// Mapping: one 64-lane wavefront per column (node, edge or element), lane k
// handles level k+1 (layers 1..nl-1, interfaces 1..nl). Column indices are
// made wave-uniform with readfirstlane so that all per-column mesh data
// (indices, level counts, geometry) is fetched with scalar loads, and blocks
// are remapped so that each of the 8 XCDs sweeps one contiguous range of
// columns (its own L2 then holds the neighbourhood the gathers revisit).
//
// Scatter-free: every node-side sum over incident edges is a gather over a
// node->edge list sorted by edge index, which is the order the reference's
// serial edge loops accumulate in -- deterministic, no atomics.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

namespace fta {

constexpr int kWave = 64;
constexpr int kWavesPerBlock = 4;
constexpr int kBlock = kWave * kWavesPerBlock;
constexpr double kRE = 6367500.0;  // r_earth, lib/oce_modules.F90:24

template <class R>
struct DMesh {
    int N, T, E, nl, S;
    const int *en;          // (3,T) element nodes
    const int2 *edges;      // (E) node pairs
    const int2 *etri;       // (E) element pairs, .y = -1 on boundary
    const int2 *updn;       // (E) up/down-wind triangles
    const int *nlev_e;      // nlevels (T)
    const int *nlev_n;      // nlevels_nod2D (N)
    const int *nlev_n_min;  // nlevels_nod2D_min (N)
    const int *nbl;         // nboundary_lay (N)
    const int *edge_nlmax;  // max(nlevels(el1),nlevels(el2))-1 (E)
    const int *n2e_ptr, *n2e;  // node -> signed edge (+(e+1): node is edges(1))
    const int *n2t_ptr, *n2t;  // node -> elements (nod_in_elem2D order)
    const R *ecross;        // (4,E) edge_cross_dxdy
    const R *edxdy;         // (2,E) edge_dxdy
    const R *edge_a;        // (E) r_earth*elem_cos averaged over the edge's elements
    const R *earea;         // (T) elem_area
    const R *gsca;          // (6,T) gradient_sca
    const R *area, *areasvol, *hnode, *hnode_new, *Zn, *zbarn;  // (S,N)
    const R *helem;         // (S,T)
};

__device__ __forceinline__ int wave_column()
{
    // hardware dispatches workgroup b to XCD b % 8; give XCD x the
    // contiguous block range [x*per, (x+1)*per) (gridDim.x is a multiple of 8)
    const unsigned nb = gridDim.x, b = blockIdx.x, per = nb >> 3;
    const unsigned lb = (b & 7u) * per + (b >> 3);
    int col = (int)(lb * kWavesPerBlock + (threadIdx.x >> 6));
    return __builtin_amdgcn_readfirstlane(col);
}

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

// Fortran max/min as the reference compiles them (a > b ? a : b)
template <class T> __device__ __forceinline__ T fmx(T a, T b) { return a > b ? a : b; }
template <class T> __device__ __forceinline__ T fmn(T a, T b) { return a < b ? a : b; }

template <class T> __device__ __forceinline__ T shdn(T v) { return __shfl_down(v, 1, kWave); }
template <class T> __device__ __forceinline__ T shup(T v) { return __shfl_up(v, 1, kWave); }

template <class G> struct GradIO;
template <> struct GradIO<double> {
    __device__ static double ld(const double *p, double) { return *p; }
    __device__ static void st(double *p, double v, double) { *p = v; }
};
template <> struct GradIO<float> {
    __device__ static float ld(const float *p, float) { return *p; }
    __device__ static void st(float *p, float v, float) { *p = v; }
};
// half storage holds gradient * gscale (a mesh length scale), so the stored
// values are O(tracer increments per edge) instead of subnormal
template <> struct GradIO<__half> {
    __device__ static float ld(const __half *p, float inv_scale) { return __half2float(*p) * inv_scale; }
    __device__ static void st(__half *p, float v, float scale) { *p = __float2half(v * scale); }
};

}  // namespace fta
