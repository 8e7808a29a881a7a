// fta_kernels.h -- gfx950 device data structures of the tracer-advection step.
//
// Mapping: one 64-lane wavefront per column (node or edge), lane k handles
// level k+1 (layers 1..nl-1, interfaces 1..nl). Column indices are made
// wave-uniform with readfirstlane so that all per-column metadata comes in
// through scalar loads of small fixed-width records (one s_load per record,
// no chains of dependent index loads), and blocks are remapped so that each
// of the 8 XCDs sweeps one contiguous range of columns (its own L2 then holds
// the neighbourhood that the gathers revisit).
//
// Tracers that share their advection schemes go through the kernels
// together (NT per launch): the mesh records and velocity are fetched once
// per column for all of them, and each wave keeps NT times more independent
// loads in flight.
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
constexpr int kDeg = 8;  // padded node degree (incident edges / elements)
constexpr double kRE = 6367500.0;  // r_earth, lib/oce_modules.F90:24

// Compact layout: every level-indexed array stores only the wet levels of
// each column, columns back to back: a node's layers 1..nlev-1 start at its
// layer offset, its interfaces 1..nlev at its interface offset (= layer
// offset + node index), an element's layers at its element offset, an edge's
// levels 1..nlmax at its edge offset. The records below carry these offsets
// directly, so a gather never waits on a separate offset load.

// per-edge record (64 bytes): layer offsets of both nodes and both elements,
// the up/down-wind triangles, packed level counts
// nl = nl1 | nl2 << 8 | nzmax_e << 16 | nlmax << 24, where nzmax_e =
// min(nlevels_nod2D_min(n1), nlevels_nod2D_min(n2)) bounds the levels shared
// by all elements around both nodes (fill_up_dn_grad), and
// nb = nboundary_lay(n1) | nboundary_lay(n2) << 16
struct alignas(64) EdgeMeta {
    int o1, o2;   // node layer offsets
    int t1, t2;   // element layer offsets (t2 = -1 on a boundary edge)
    int up, dn;   // up/down-wind triangle indices (-1 = none)
    int nl, nb;
    int oe;       // this edge's offset in the edge arrays
    int el2;      // second element index (-1 = boundary)
    int pad[6];
};

// per-element record: layer offsets of its three nodes and nlevels
struct alignas(16) ElemMeta {
    int o0, o1, o2, nlev;
};

// per-node record: nlevels_nod2D, layer offset, interface offset
struct alignas(16) NodeRec {
    int nlev, off, offi, pad;
};

// incident edges of a node padded to kDeg, sorted by edge index:
// oe = edge offset (-1 = empty); info = nlmax | 256 if the node is edges(1)
// | (layers of the other node) << 16; oo = layer offset of the edge's other
// node (the node's 1-ring neighbours, slot j)
struct alignas(32) NodeEdgeRec {
    int oe[kDeg];
    int info[kDeg];
    int oo[kDeg];
};
// elements around a node (nod_in_elem2D order), padded to kDeg: element
// index (-1 = empty) and sl = s0 | s1 << 4 | s2 << 8 | nlevels << 16, where
// s0..s2 locate the element's nodes 1..3 in the node's neighbourhood:
// slot j < kDeg = the other node of incident edge j, slot 15 = the node itself
struct alignas(32) NodeElems {
    int t[kDeg];
    int sl[kDeg];
};
constexpr int kSelf = 15;

template <class R>
struct DMesh {
    int N, T, E, nl;
    long long CN, CI, CT, CE;   // cell counts: node layers, node interfaces, element layers, edge levels
    const NodeRec *nrec;        // (N)
    const EdgeMeta *emeta;      // (E)
    const R *egeo;              // (8,E): dx1 dy1 dx2 dy2 (edge_cross_dxdy) ex ey (edge_dxdy) a
    const ElemMeta *tmeta;      // (T)
    const R *tgeo;              // (8,T): gradient_sca(1:6), elem_area, 0
    const NodeEdgeRec *nedge;   // (N)
    const NodeElems *nelem;     // (N)
    const R *area, *zbarn;      // interface arrays (CI)
    const R *areasvol, *hnode, *hnode_new, *Zn;  // layer arrays (CN)
    const R *helem;             // element layer array (CT)
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
