// fta_kernels.h -- gfx950 device data structures of the tracer-advection step.
//
// Cell-packed mapping. Every level-indexed array (node layers and
// interfaces, element layers, edge levels) stores the wet levels of each
// column back to back, each column padded to an even length and columns
// packed whole, in mesh order (nodes and edges with a short look-ahead that
// fills chunk tails), into chunks of kChunk = 128 cells; a column never
// straddles two chunks. One 64-lane wavefront processes one chunk and
// each lane two consecutive cells of one column (levels k0 = 2p, k0+1), so
// every tracer/flux/geometry access of a lane is one 16-byte load for FP64
// data (8 bytes for FP32). On gfx950 a vector-memory instruction of 64- or
// 128-bit lanes occupies the texture path for the same time whatever its
// active-lane count, so filling all lanes with wet cells and moving two
// cells per lane is what sets the throughput of these gather-heavy kernels.
//
// A per-slot word (slot = cell pair) names the lane's column and pair index;
// the column's records (incident edges, 1-ring, elements) are fixed-width
// and read by vector loads. Vertical neighbours of a cell are in the same
// lane or in lane +-1 (the whole column is in the wave): cross-lane moves.
//
// Tracers that share their advection schemes go through the kernels
// together (NT per launch): the mesh records and velocity are fetched once
// per cell pair for all of them.
//
// Scatter-free: every node-side sum over incident edges is a gather over a
// node->edge list sorted by edge index, which is the order the reference's
// serial edge loops accumulate in -- deterministic, no atomics.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

namespace fta {

constexpr int kWave = 64;
constexpr int kWavesPerBlock = 4;                  // default launch shape (and the fixed one of LDS kernels)
constexpr int kBlock = kWave * kWavesPerBlock;
constexpr int kMaxBlock = 512;                     // chunk kernels take 64..512 threads (autotuned)
constexpr int kChunk = 2 * kWave;   // cells per wave chunk
constexpr int kDeg = 8;             // node record width (incident edges / elements / 1-ring) of ordinary meshes
constexpr int kDegWide = 16;        // the width used when some node has more than kDeg
constexpr int kColBits = 25;        // slot word: column | pair index << kColBits (pairs 0..63, sign bit clear)
constexpr int kMaxLevels = 2 * kWave;  // level interfaces a column may have: a whole column in one chunk
constexpr double kRE = 6367500.0;   // r_earth, lib/oce_modules.F90:24

// per-edge record: layer offsets of both nodes, of both elements (t2 = -1 on
// a boundary edge) and of the up/down-wind triangles (-1 = none);
// nl = nl1 | nl2 << 8 | nzmax_e << 16 | nlmax << 24 where nl1/nl2 are the
// elements' layer counts, nzmax_e = min(nlevels_nod2D_min(n1),
// nlevels_nod2D_min(n2)) bounds the levels shared by all elements around
// both nodes (fill_up_dn_grad) and nlmax = max(nl1, nl2);
// nb = nboundary_lay(n1) | nboundary_lay(n2) << 16
struct alignas(32) EdgeRec {
    int o1, o2, t1, t2, tu, td, nl, nb;
};

// per-element record: layer offsets of its three nodes and nlevels
struct alignas(16) ElemRec {
    int o0, o1, o2, nlev;
};

// Node records come in two widths D (kDeg, or kDegWide for meshes with a node
// above kDeg); the node-side kernels are instantiated per width.
//
// node -> incident edges (sorted by edge index), D slots:
// oe = edge offset (-1 = empty), oo = layer offset of the edge's other node,
// info = nlmax (7 bits) | 128 if the node is edges(1) | (layers of the other node) << 8
template <int D>
struct alignas(16) NodeEdges {
    int oe[D];
    int oo[D];
};
template <int D>
struct alignas(16) NodeEdgeInfo {
    unsigned short info[D];
};

// node -> 1-ring in cyclic (fan) order, for the Zalesak bounds: element j of
// the fan is (node, ring j, ring (j+1) mod R); oo = ring node layer offsets
// (-1 = empty), lr = ring node layer counts, le = layer count of fan element
// j (0 = no element: the open end of a boundary fan)
template <int D>
struct alignas(16) NodeRing {
    int oo[D];
    unsigned char lr[D];
    unsigned char le[D];
};

// node -> elements (nod_in_elem2D order) for the nodal mean gradient, formed
// from the node values: ei = element index (-1 = empty), oa / ob = layer
// offsets of the element's other two nodes in its elem2D_nodes order,
// le = element layer count, pos = this node's place (0..2) in elem2D_nodes
template <int D>
struct alignas(16) NodeElems {
    int ei[D];
    int oa[D];
    int ob[D];
    unsigned char le[D];
    unsigned char pos[D];
};

// 16-byte units of one edge's packed record: EdgeRec (2), egeo (8 R),
// eud (1), egud (4 R)
template <class R>
constexpr int kEdgeUnits = 3 + (12 * (int)sizeof(R)) / 16;
// edges a wave stages into LDS (chunks with more fall back to global loads)
constexpr int kStageEdges = 16;
// nodes a wave stages into LDS (their NodeEdges + NodeEdgeInfo)
constexpr int kStageNodes = 16;
// bytes of nodal-gradient fan weights a wave stages into LDS
constexpr int kStageFanBytes = 4096;
template <int D>
constexpr int kNodeEdgeUnits = (int)(sizeof(NodeEdges<D>) + sizeof(NodeEdgeInfo<D>)) / 16;

template <class R>
struct DMesh {
    int N, T, E, nl;
    int nch_n, nch_t, nch_e;      // wave chunks of the node / element / edge layouts
    long long CN, CT, CE;         // padded cell counts (nch * kChunk)
    // per cell pair: column | pair << kColBits, -1 = pad. A column is named by
    // its place in the layout (the order its columns were packed in), and the
    // per-column records below are stored in that order, so that the columns
    // of one chunk are consecutive records
    const int *slot_n, *slot_t, *slot_e;
    const int *nlev;              // (N) nlevels_nod2D, node layout order
    const int *ncol;              // (N) node layout column -> the caller's node index
    const int *nlev_c;            // (N) nlevels_nod2D, the caller's node order
    const EdgeRec *erec;          // (E) edge layout order
    const R *egeo;                // (8,E): dx1 dy1 dx2 dy2 (edge_cross_dxdy) ex ey (edge_dxdy) a
    const ElemRec *trec;          // (T)
    const R *tgeo;                // (8,T): gradient_sca(1:6), elem_area, 0
    // MUSCL on the levels where the up/down-wind triangles give the gradient:
    // eud = layer offsets of the up-wind triangle's two nodes a, b other than
    // edges(1) and of the down-wind triangle's two nodes c, d other than
    // edges(2); egud = (4,E): weights with which ex*a*d/dx + ey*r_earth*d/dy of
    // the up-wind triangle's gradient is w0 (T_a - T_1) + w1 (T_b - T_1) and of
    // the down-wind one w2 (T_c - T_2) + w3 (T_d - T_2) (a triangle's gradient
    // coefficients sum to zero, so the edge's own node enters through the
    // differences)
    const int4 *eud;              // (E)
    const R *egud;                // (4,E)
    // the same per-edge records packed edge by edge in 16-byte units (erec,
    // egeo, eud, egud; kEdgeUnits<R> per edge), so that a wave can stage its
    // chunk's edges into LDS with a few cooperative loads
    const uint4 *epack;           // (kEdgeUnits<R>, E)
    // the nodal mean gradient over a node's fan: slot j (the element between
    // ring nodes j and j+1) holds A*c of ring j and ring j+1 for d/dx, the same
    // for d/dy, and A (elem_area), with c the element's gradient coefficients,
    // so that its gradient times A is the weighted differences to this node
    const R *ngw;                 // (8*D,N), unused slots 0
    int deg;                      // node record width D (kDeg or kDegWide)
    const void *nedge;            // (N) NodeEdges<D>
    const void *ninfo;            // (N) NodeEdgeInfo<D>
    const void *nring;            // (N) NodeRing<D>
    const void *nelem;            // (N) NodeElems<D>
    // node layout (interfaces 1..nlev and layers 1..nlev-1 share a column)
    const R *area, *zbarn;        // interface arrays
    const R *areasvol, *hnode, *hnode_new, *Zn;  // layer arrays
    const R *helem;               // element layout
    // areasvol, hnode, hnode_new in FP64 whatever R is: the LO solution and the
    // flux-divergence accumulation into the node tracer use them, so the
    // tracer mass is conserved to FP64 rounding in every precision form
    const double *av_d, *h_d, *hn_d;
};

template <int D, class R>
__device__ __forceinline__ const NodeEdges<D> *node_edges(const DMesh<R> &m)
{
    return static_cast<const NodeEdges<D> *>(m.nedge);
}
template <int D, class R>
__device__ __forceinline__ const NodeEdgeInfo<D> *node_info(const DMesh<R> &m)
{
    return static_cast<const NodeEdgeInfo<D> *>(m.ninfo);
}
template <int D, class R>
__device__ __forceinline__ const NodeRing<D> *node_ring(const DMesh<R> &m)
{
    return static_cast<const NodeRing<D> *>(m.nring);
}
template <int D, class R>
__device__ __forceinline__ const NodeElems<D> *node_elems(const DMesh<R> &m)
{
    return static_cast<const NodeElems<D> *>(m.nelem);
}

__device__ __forceinline__ int wave_chunk(bool rev = false)
{
    // hardware dispatches workgroup b to XCD b % 8; give XCD x the
    // contiguous block range [x*per, (x+1)*per) (gridDim.x is a multiple of 8),
    // swept from its end when rev is set. The kernels of a step alternate
    // (k_gnod_fan, k_node_lo, k_node_update and k_node_nolim sweep from the
    // end, k_edge and k_node_limit from the start), so that each kernel
    // begins on the chunks its predecessor touched last, whose lines the
    // memory-side cache in front of HBM (256 MB) still holds
    const unsigned nb = gridDim.x, b = blockIdx.x, per = nb >> 3, i = b >> 3;
    const unsigned lb = (b & 7u) * per + (rev ? per - 1u - i : i);
    int ch = (int)(lb * (blockDim.x >> 6) + (threadIdx.x >> 6));
    return __builtin_amdgcn_readfirstlane(ch);
}

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

// this lane's cell pair
struct Pair {
    int col;      // column index (valid when live)
    int k0;       // 0-based level of the first cell
    long long c;  // layout index of the first cell
    bool live;
};

__device__ __forceinline__ Pair lane_pair(const int *slot, int chunk)
{
    const long long s = (long long)chunk * kWave + lane_id();
    const int w = slot[s];
    Pair p;
    p.live = w >= 0;
    p.col = w & ((1 << kColBits) - 1);
    p.k0 = p.live ? 2 * (int)((unsigned)w >> kColBits) : 0;
    p.c = 2 * s;
    return p;
}

// the columns of a chunk are consecutive: [c0, c0 + ncol). A wave stages
// per-column records into its LDS slice only when it can see that range
// whole (lane 0 live, every live lane inside it) and ncol <= cap; the result
// is wave-uniform
struct StageRange {
    int c0, ncol;
    bool staged;
};

__device__ __forceinline__ StageRange stage_range(const Pair &p, int cap)
{
    const unsigned long long lv = __ballot(p.live);
    StageRange r;
    r.c0 = __builtin_amdgcn_readfirstlane(p.col);
    const int cmax = __builtin_amdgcn_readfirstlane(__shfl(p.col, __popcll(lv) - 1));
    r.ncol = cmax - r.c0 + 1;
    bool ok = (lv & 1ull) && cmax >= r.c0 && r.ncol <= cap;
    ok = ok && __ballot(p.live && (p.col < r.c0 || p.col > cmax)) == 0ull;
    r.staged = ok;
    return r;
}

// copies the packed records (U 16-byte units per column) of the range into
// the wave's slice with cooperative loads; call stage_sync() before reading
template <int U>
__device__ __forceinline__ void stage_copy(const uint4 *__restrict__ pack, uint4 *slice, const StageRange &r)
{
    const uint4 *src = pack + (size_t)r.c0 * U;
    const int nu = r.ncol * U;
    for (int u = (int)(threadIdx.x & 63); u < nu; u += 64) slice[u] = src[u];
}

__device__ __forceinline__ void stage_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Fortran max/min as the reference compiles them (a > b ? a : b)
// max / min on the hardware instructions (v_max_f64, v_min_f32, ...); they
// differ from a compare-and-select only in the sign of an exact zero and for
// NaN, which no limiter sum or bound of this path can tell apart
__device__ __forceinline__ double fmx(double a, double b) { return __builtin_fmax(a, b); }
__device__ __forceinline__ double fmn(double a, double b) { return __builtin_fmin(a, b); }
__device__ __forceinline__ float fmx(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ __forceinline__ float fmn(float a, float b) { return __builtin_fminf(a, b); }

// NaN test on the bit pattern: device code is compiled assuming no NaN
// arithmetic (so that max/min need no operand canonicalisation), under which
// x != x may fold to false
__host__ __device__ __forceinline__ bool is_nan_bits(double x)
{
    unsigned long long u;
    __builtin_memcpy(&u, &x, 8);
    return (u & 0x7FFFFFFFFFFFFFFFull) > 0x7FF0000000000000ull;
}

// lane l receives lane l-1's / l+1's value (own value at the wave ends)
template <class T> __device__ __forceinline__ T from_prev(T v) { return __shfl_up(v, 1, kWave); }
template <class T> __device__ __forceinline__ T from_next(T v) { return __shfl_down(v, 1, kWave); }
// an interface quantity of cell k0+2 taken from the next lane (nx), or zero
// where that is the column's bottom interface (k0 + 2 == L layers): node
// columns keep no cell for it, and its flux is zero in every scheme
template <class T> __device__ __forceinline__ T below(T nx, int k0, int L) { return k0 + 2 < L ? nx : (T)0; }

// two consecutive cells
template <class T>
struct V2 {
    T x, y;
};

__device__ __forceinline__ V2<double> ld2(const double *p)
{
    const double2 v = *reinterpret_cast<const double2 *>(p);
    return {v.x, v.y};
}
__device__ __forceinline__ V2<float> ld2(const float *p)
{
    const float2 v = *reinterpret_cast<const float2 *>(p);
    return {v.x, v.y};
}
// a pair that no other lane of the launch reads (the own layers of a node in
// the node kernels), loaded with the nontemporal hint so that it does not
// keep L2 lines from the gathered data (FP64 step -0.4%, FP32 -0.8%)
typedef double ld_d2 __attribute__((ext_vector_type(2)));
typedef float ld_f2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ V2<double> ld2_nt(const double *p)
{
    const ld_d2 v = __builtin_nontemporal_load(reinterpret_cast<const ld_d2 *>(p));
    return {v.x, v.y};
}
__device__ __forceinline__ V2<float> ld2_nt(const float *p)
{
    const ld_f2 v = __builtin_nontemporal_load(reinterpret_cast<const ld_f2 *>(p));
    return {v.x, v.y};
}
// Pair stores carry the nontemporal hint: the lanes of a column write whole
// lines, which the step reads again only in a later kernel, so the writes
// need not hold L2 lines that the kernel's gathers use (FP64 step -1.6%,
// FP32 -0.5%). The hint marks lines for early eviction and changes no value
// a later read sees. The exception is an FP32 store that covers every other 8
// bytes of an interleaved array (adf_at's tracer pairs, R+ alone of an R+/R-
// group), which keeps ordinary stores through st2_mixed (the hint made k_edge
// 13% slower there)
typedef double st_d2 __attribute__((ext_vector_type(2)));
typedef float st_f2 __attribute__((ext_vector_type(2)));
typedef float st_f4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void st2(double *p, double a, double b)
{
    const st_d2 v = {a, b};
    __builtin_nontemporal_store(v, reinterpret_cast<st_d2 *>(p));
}
__device__ __forceinline__ void st2(float *p, float a, float b)
{
    const st_f2 v = {a, b};
    __builtin_nontemporal_store(v, reinterpret_cast<st_f2 *>(p));
}
__device__ __forceinline__ void st4(float *p, float a, float b, float c, float d)
{
    const st_f4 v = {a, b, c, d};
    __builtin_nontemporal_store(v, reinterpret_cast<st_f4 *>(p));
}
// a pair of an interleaved array (FP32: every other 8 bytes)
__device__ __forceinline__ void st2_mixed(double *p, double a, double b) { st2(p, a, b); }
__device__ __forceinline__ void st2_mixed(float *p, float a, float b)
{
    *reinterpret_cast<float2 *>(p) = make_float2(a, b);
}

// the pair at i of array b, or `known` (the pair at i of array a) when b is
// a's storage (the host shares one copy between arrays that are equal); the
// test is on kernel arguments, so the branch is uniform
template <class T>
__device__ __forceinline__ V2<T> ld_same(const T *b, const T *a, size_t i, V2<T> known)
{
    if (b == a) return known;
    return ld2(b + i);
}

// Zalesak's R+/R- of the cell pair at i (i even, tracer offset included).
// FP64: two arrays, each pair one 16-byte access. FP32: one array holding
// (R+ k, R+ k+1, R- k, R- k+1) per pair, so that both are one 16-byte access
// (the array is twice as long per tracer; Rm is not used).
__device__ __forceinline__ void ld_rpm(const double *Rp, const double *Rm, size_t i, V2<double> &rp,
                                       V2<double> &rm)
{
    rp = ld2(Rp + i);
    rm = ld2(Rm + i);
}
__device__ __forceinline__ void ld_rpm(const float *Rp, const float *, size_t i, V2<float> &rp, V2<float> &rm)
{
    const float4 v = *reinterpret_cast<const float4 *>(Rp + 2 * i);
    rp = {v.x, v.y};
    rm = {v.z, v.w};
}
__device__ __forceinline__ void st_rpm(double *Rp, double *Rm, size_t i, double p0, double p1, double m0, double m1)
{
    st2(Rp + i, p0, p1);
    st2(Rm + i, m0, m1);
}
__device__ __forceinline__ void st_rpm(float *Rp, float *, size_t i, float p0, float p1, float m0, float m1)
{
    st4(Rp + 2 * i, p0, p1, m0, m1);
}
// The learned limiter's bound-check flag of a node level rides on the sign
// bit of its R+ (R+ >= 0, so the bit is otherwise clear): with_flag sets it,
// has_flag reads it, and fabs gives the value back.
__device__ __forceinline__ float with_flag(float r, bool f)
{
    return __uint_as_float(__float_as_uint(r) | (f ? 0x80000000u : 0u));
}
__device__ __forceinline__ double with_flag(double r, bool f)
{
    return __longlong_as_double(__double_as_longlong(r) | (f ? (long long)(1ull << 63) : 0ll));
}
__device__ __forceinline__ bool has_flag(float r) { return (__float_as_uint(r) >> 31) != 0u; }
__device__ __forceinline__ bool has_flag(double r) { return __double_as_longlong(r) < 0; }

// HO-LO (or HO) edge fluxes of work slot t at the cell pair i (i even).
// FP64: slot t is the block [t*tse, (t+1)*tse). FP32: slots 2s and 2s+1 share
// the block of 2*tse floats at 2s*tse, interleaved per cell pair (slot 2s's
// k, k+1, then slot 2s+1's), so that both slots' pairs are one 16-byte access.
template <class R>
constexpr size_t kAdfStride = sizeof(R) == 4 ? 2 : 1;
template <class R>
__host__ __device__ __forceinline__ size_t adf_slot(size_t t, size_t tse)
{
    return sizeof(R) == 4 ? (t >> 1) * 2 * tse + 2 * (t & 1) : t * tse;
}
// (i must start a pair; its second cell follows at + 1)
template <class R>
__host__ __device__ __forceinline__ size_t adf_at(size_t t, size_t tse, size_t i)
{
    return adf_slot<R>(t, tse) + kAdfStride<R> * i;
}
// slots t and t + 1 (t even) of the cell pair i
__device__ __forceinline__ void ld_adf2(const double *a, size_t t, size_t tse, size_t i, V2<double> &x,
                                        V2<double> &y)
{
    x = ld2(a + t * tse + i);
    y = ld2(a + (t + 1) * tse + i);
}
__device__ __forceinline__ void ld_adf2(const float *a, size_t t, size_t tse, size_t i, V2<float> &x, V2<float> &y)
{
    const float4 v = *reinterpret_cast<const float4 *>(a + (t >> 1) * 2 * tse + 2 * i);
    x = {v.x, v.y};
    y = {v.z, v.w};
}

// elements of R+/R- storage per cell (see ld_rpm)
template <class R>
constexpr size_t kRpmPerCell = sizeof(R) == 4 ? 2 : 1;

// two cells x two components (x0 y0 x1 y1), gradients and velocity
template <class T>
struct V4 {
    T x0, y0, x1, y1;
};
__device__ __forceinline__ V4<double> ld4(const double *p)
{
    const double2 a = reinterpret_cast<const double2 *>(p)[0];
    const double2 b = reinterpret_cast<const double2 *>(p)[1];
    return {a.x, a.y, b.x, b.y};
}
__device__ __forceinline__ V4<float> ld4(const float *p)
{
    const float4 a = *reinterpret_cast<const float4 *>(p);
    return {a.x, a.y, a.z, a.w};
}

// gradient storage: double, float, or half holding gradient * gscale (a
// mesh length scale, so the stored values are O(tracer increments per edge)
// instead of subnormal)
template <class G> struct GradIO;
template <> struct GradIO<double> {
    __device__ static V4<double> ld(const double *p, double) { return ld4(p); }
    __device__ static void st(double *p, double a, double b, double c, double d, double)
    {
        st2(p, a, b);
        st2(p + 2, c, d);
    }
};
template <> struct GradIO<float> {
    __device__ static V4<float> ld(const float *p, float) { return ld4(p); }
    __device__ static void st(float *p, float a, float b, float c, float d, float)
    {
        st4(p, a, b, c, d);
    }
};
template <> struct GradIO<__half> {
    __device__ static V4<float> ld(const __half *p, float inv)
    {
        const uint2 u = *reinterpret_cast<const uint2 *>(p);
        const __half2 a = *reinterpret_cast<const __half2 *>(&u.x);
        const __half2 b = *reinterpret_cast<const __half2 *>(&u.y);
        const float2 fa = __half22float2(a), fb = __half22float2(b);
        return {fa.x * inv, fa.y * inv, fb.x * inv, fb.y * inv};
    }
    __device__ static void st(__half *p, float a, float b, float c, float d, float scale)
    {
        const __half2 h0 = __floats2half2_rn(a * scale, b * scale);
        const __half2 h1 = __floats2half2_rn(c * scale, d * scale);
        uint2 u;
        u.x = *reinterpret_cast<const unsigned *>(&h0);
        u.y = *reinterpret_cast<const unsigned *>(&h1);
        *reinterpret_cast<uint2 *>(p) = u;
    }
};

}  // namespace fta
