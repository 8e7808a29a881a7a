// fta_engine.hip -- the tracer-advection step on gfx950 and its host driver.
//
// One step (reference: do_oce_adv_tra, lib/oce_adv_tra_driver.F90:23-294,
// plus the caller's update, src/fesom_analytic.F90:405-437) for a batch of
// NT tracers that share their schemes is five launches:
//   k_gnod        per node-level: area-weighted mean of the elemental
//                 gradients of valuesAB around the node (MUSCL/MFCT)
//   k_edge        per edge-level: LO upwind flux and HO-LO antidiffusive flux
//   k_node_lo     per node-level: LO solution, vertical HO-LO flux
//   k_node_limit  per node-level: P+/P- sums, Zalesak bounds (a1-a3), R+/R- (b2)
//   k_node_update per node-level: limited fluxes (b3), divergence, T update
// (without FCT the last three collapse into k_node_nolim). Every kernel runs
// one wave per 128-cell chunk of its layout, two cells per lane.
#include "fta_engine.h"
#include "fta_hostcopy.h"
#include "fta_kernels.h"
#include "fta_mlp.h"
#include "fta_vertical.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <initializer_list>
#include <limits>
#include <memory>
#include <stdexcept>
#include <type_traits>

namespace fta {

static std::string g_last_error;
const char *last_error() { return g_last_error.c_str(); }
void set_error(const std::string &s) { g_last_error = s; }

#define HIP_OK(call)                                                                          \
    do {                                                                                      \
        hipError_t _e = (call);                                                               \
        if (_e != hipSuccess)                                                                 \
            throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(_e) +      \
                                     " at " #call);                                           \
    } while (0)

// node->edge slot info: nlmax (7 bits) | 128 if the node is edges(1) | other node's layers << 8
__device__ __forceinline__ int info_nlmax(unsigned i) { return (int)(i & 127u); }
__device__ __forceinline__ bool info_pos(unsigned i) { return (i & 128u) != 0; }

// node -> edge records from their 16-byte units (global memory or LDS)
template <int D>
__device__ __forceinline__ void decode_node_edges(const int4 *qe, const uint4 *qi, int (&oe)[D], unsigned (&info)[D])
{
#pragma unroll
    for (int h = 0; h < D / 4; ++h) {
        const int4 a = qe[h];
        oe[4 * h] = a.x; oe[4 * h + 1] = a.y; oe[4 * h + 2] = a.z; oe[4 * h + 3] = a.w;
    }
#pragma unroll
    for (int h = 0; h < D / 8; ++h) {
        const uint4 v = qi[h];
        const unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            info[8 * h + 2 * j] = w[j] & 0xFFFFu;
            info[8 * h + 2 * j + 1] = w[j] >> 16;
        }
    }
}

template <int D>
__device__ __forceinline__ void decode_other_nodes(const int4 *qo, int (&oo)[D])
{
#pragma unroll
    for (int h = 0; h < D / 4; ++h) {
        const int4 a = qo[h];
        oo[4 * h] = a.x; oo[4 * h + 1] = a.y; oo[4 * h + 2] = a.z; oo[4 * h + 3] = a.w;
    }
}

// the chunk's node -> edge records (NodeEdges, then NodeEdgeInfo, per node)
// staged into the wave's LDS slice when the chunk has at most kStageNodes
// nodes; the wave must call this before any lane leaves
template <int D, class R>
__device__ __forceinline__ StageRange stage_node_edges(const DMesh<R> &m, const Pair &p, uint4 *slice)
{
    constexpr int UE = (int)sizeof(NodeEdges<D>) / 16, UI = (int)sizeof(NodeEdgeInfo<D>) / 16;
    const StageRange sr = stage_range(p, kStageNodes);
    if (sr.staged) {
        stage_copy<UE>(static_cast<const uint4 *>(m.nedge), slice, sr);
        stage_copy<UI>(static_cast<const uint4 *>(m.ninfo), slice + kStageNodes * UE, sr);
        stage_sync();
    }
    return sr;
}

// incident edges of node n (edge-index order) and, with OO, their other nodes
template <int D, bool OO, class R>
__device__ __forceinline__ void load_node_edges(const DMesh<R> &m, const StageRange &sr, const uint4 *slice, int n,
                                                bool live, int (&oe)[D], unsigned (&info)[D], int (&oo)[D])
{
    constexpr int UE = (int)sizeof(NodeEdges<D>) / 16, UI = (int)sizeof(NodeEdgeInfo<D>) / 16;
    if (live) {
        if (sr.staged) {
            const uint4 *q = slice + (n - sr.c0) * UE;
            decode_node_edges<D>(reinterpret_cast<const int4 *>(q), slice + kStageNodes * UE + (n - sr.c0) * UI, oe,
                                 info);
            if constexpr (OO) decode_other_nodes<D>(reinterpret_cast<const int4 *>(q) + D / 4, oo);
        } else {
            const int4 *q = reinterpret_cast<const int4 *>(&node_edges<D>(m)[n]);
            decode_node_edges<D>(q, reinterpret_cast<const uint4 *>(&node_info<D>(m)[n]), oe, info);
            if constexpr (OO) decode_other_nodes<D>(q + D / 4, oo);
        }
    } else {
#pragma unroll
        for (int j = 0; j < D; ++j) { oe[j] = -1; info[j] = 0; oo[j] = 0; }
    }
}

// records without staging (kernels that do not stage); the other nodes are
// decoded too and dropped: leaving them out makes the compiler keep one slot
// of the partly written arrays in scratch memory
template <int D, class R>
__device__ __forceinline__ void load_node_edges(const DMesh<R> &m, int n, bool live, int (&oe)[D],
                                                unsigned (&info)[D])
{
    int oo[D];
    const StageRange none{0, 0, false};
    load_node_edges<D, true>(m, none, nullptr, n, live, oe, info, oo);
}

// ===========================================================================
// MUSCL prerequisites (fill_up_dn_grad, lib/oce_muscl_adv.F90:361-530): the
// area-weighted mean (:397-407) over the node's elements that have the layer
// of their elemental gradients of valuesAB (tracer_gradient_elements:
// gradient_sca(1:3)*T(nodes), gradient_sca(4:6)*T(nodes), node order 1,2,3),
// each element's gradient formed here from the values at its three nodes and
// never stored: the sums run in the same order with the same operands as when
// the elemental gradients were a separate array, so the result is the same
// ===========================================================================
template <class R>
__device__ __forceinline__ void ld_tgeo(const R *g, R (&c)[6], R &area)
{
    const V4<R> a = ld4(g), b = ld4(g + 4);
    c[0] = a.x0; c[1] = a.y0; c[2] = a.x1; c[3] = a.y1; c[4] = b.x0; c[5] = b.y0;
    area = b.x1;
}

template <class R, class G, int NT, int D>
__global__ __launch_bounds__(kMaxBlock) void k_gnod(DMesh<R> m, const double *__restrict__ ab, size_t tsn,
                                                 G *__restrict__ gnod, size_t tsg, float gscale)
{
    const int ch = wave_chunk(true);  // from the end of each XCD's range (wave_chunk)
    if (ch >= m.nch_n) return;
    const Pair p = lane_pair(m.slot_n, ch);
    if (!p.live) return;
    const int L = m.nlev[p.col] - 1;
    if (p.k0 >= L) return;  // (every live pair starts on a layer; kept as a guard)
    const bool v1 = p.k0 + 1 < L;
    int ei[D], oa[D], ob[D], le[D], pos[D];
    {
        const NodeElems<D> *rec = &node_elems<D>(m)[p.col];
        const int4 *q = reinterpret_cast<const int4 *>(rec->ei);
#pragma unroll
        for (int h = 0; h < D / 4; ++h) {
            const int4 a = q[h], b = q[D / 4 + h], c = q[D / 2 + h];
            ei[4 * h] = a.x; ei[4 * h + 1] = a.y; ei[4 * h + 2] = a.z; ei[4 * h + 3] = a.w;
            oa[4 * h] = b.x; oa[4 * h + 1] = b.y; oa[4 * h + 2] = b.z; oa[4 * h + 3] = b.w;
            ob[4 * h] = c.x; ob[4 * h + 1] = c.y; ob[4 * h + 2] = c.z; ob[4 * h + 3] = c.w;
        }
        // le[0..D) then pos[0..D), four bytes to a word
        const uint4 *lw = reinterpret_cast<const uint4 *>(rec->le);
        if constexpr (D == 8) {
            const uint4 u = lw[0];
            const unsigned w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                le[j] = (int)((w[j >> 2] >> (8 * (j & 3))) & 255u);
                pos[j] = (int)((w[2 + (j >> 2)] >> (8 * (j & 3))) & 255u);
            }
        } else {
            const uint4 u0 = lw[0], u1 = lw[1];
            const unsigned w[8] = {u0.x, u0.y, u0.z, u0.w, u1.x, u1.y, u1.z, u1.w};
#pragma unroll
            for (int j = 0; j < D; ++j) {
                le[j] = (int)((w[j >> 2] >> (8 * (j & 3))) & 255u);
                pos[j] = (int)((w[(D >> 2) + (j >> 2)] >> (8 * (j & 3))) & 255u);
            }
        }
    }
    V2<double> Tn[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) Tn[t] = ld2(ab + t * tsn + p.c);
    R tvol0 = 0, tvol1 = 0;
    R tx0[NT], ty0[NT], tx1[NT], ty1[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) { tx0[t] = 0; ty0[t] = 0; tx1[t] = 0; ty1[t] = 0; }
#pragma unroll
    for (int j = 0; j < D; ++j) {
        if (ei[j] < 0) break;
        if (p.k0 >= le[j]) continue;
        const bool a1 = p.k0 + 1 < le[j];
        R g[6], a;
        ld_tgeo(m.tgeo + 8 * (size_t)ei[j], g, a);
        tvol0 = tvol0 + a;
        tvol1 = a1 ? tvol1 + a : tvol1;
        const int ps = pos[j];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const double *A = ab + t * tsn + p.k0;
            const V2<double> Ta = ld2(A + oa[j]), Tb = ld2(A + ob[j]);
            // the element's node values in its elem2D_nodes order (this node at ps)
            const V2<double> T0 = ps == 0 ? Tn[t] : Ta;
            const V2<double> T1 = ps == 0 ? Ta : (ps == 1 ? Tn[t] : Tb);
            const V2<double> T2 = ps == 2 ? Tn[t] : Tb;
            const R gx0 = g[0] * (R)T0.x + g[1] * (R)T1.x + g[2] * (R)T2.x;
            const R gy0 = g[3] * (R)T0.x + g[4] * (R)T1.x + g[5] * (R)T2.x;
            const R gx1 = g[0] * (R)T0.y + g[1] * (R)T1.y + g[2] * (R)T2.y;
            const R gy1 = g[3] * (R)T0.y + g[4] * (R)T1.y + g[5] * (R)T2.y;
            tx0[t] = tx0[t] + gx0 * a;
            ty0[t] = ty0[t] + gy0 * a;
            const R x1 = tx1[t] + gx1 * a, y1 = ty1[t] + gy1 * a;
            tx1[t] = a1 ? x1 : tx1[t];
            ty1[t] = a1 ? y1 : ty1[t];
        }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t)
        GradIO<G>::st(gnod + t * tsg + 2 * p.c, tx0[t] / tvol0, ty0[t] / tvol0, v1 ? tx1[t] / tvol1 : (R)0,
                      v1 ? ty1[t] / tvol1 : (R)0, gscale);
}

// the same mean over the node's fan (ring node j and j+1 bound element j):
// every ring value is loaded once, and each element's gradient times its area
// is the weighted sum of its two ring values' differences to this node's
// value (a triangle's gradient coefficients sum to zero). Sums in fan order;
// the result equals the element-by-element form to rounding.
//
// The chunk's fan weights (8*D values per column) are staged into the wave's
// LDS slice by cooperative 16-byte loads when the chunk has at most
// kStageFanBytes / (column's bytes) columns: every lane of a column reads the
// same weights, which from global memory cost one vector-memory instruction
// per lane and slot.
template <class R, int D>
struct FanW {
    static constexpr int kUnits = 8 * D * (int)sizeof(R) / 16;     // 16-byte units per column
    static constexpr int kCols = kStageFanBytes / (kUnits * 16);   // columns a wave stages
};

template <class R>
__device__ __forceinline__ void fan_weights(const R *w, R &ax, R &bx, R &ay, R &by, R &a)
{
    const V2<R> wa = ld2(w), wb = ld2(w + 2), wc = ld2(w + 4);
    ax = wa.x; bx = wa.y; ay = wb.x; by = wb.y; a = wc.x;
}

__device__ __forceinline__ void fan_weights_lds(const uint4 *q, double &ax, double &bx, double &ay, double &by,
                                                double &a)
{
    const uint4 u0 = q[0], u1 = q[1], u2 = q[2];
    double v[6];
    __builtin_memcpy(&v[0], &u0, 16);
    __builtin_memcpy(&v[2], &u1, 16);
    __builtin_memcpy(&v[4], &u2, 16);
    ax = v[0]; bx = v[1]; ay = v[2]; by = v[3]; a = v[4];
}

__device__ __forceinline__ void fan_weights_lds(const uint4 *q, float &ax, float &bx, float &ay, float &by, float &a)
{
    const uint4 u0 = q[0], u1 = q[1];
    float v[8];
    __builtin_memcpy(&v[0], &u0, 16);
    __builtin_memcpy(&v[4], &u1, 16);
    ax = v[0]; bx = v[1]; ay = v[2]; by = v[3]; a = v[4];
}

template <class R, class G, int NT, int D>
__global__ __launch_bounds__(kMaxBlock) void k_gnod_fan(DMesh<R> m, const double *__restrict__ ab, size_t tsn,
                                                     G *__restrict__ gnod, size_t tsg, float gscale)
{
    const int ch = wave_chunk(true);  // from the end of each XCD's range (wave_chunk)
    if (ch >= m.nch_n) return;
    const Pair p = lane_pair(m.slot_n, ch);
    constexpr int WU = FanW<R, D>::kUnits, WC = FanW<R, D>::kCols;
    extern __shared__ uint4 stage_all[];
    uint4 *stage = stage_all + (threadIdx.x >> 6) * (WC * WU);
    const StageRange sr = stage_range(p, WC);
    if (sr.staged) {
        stage_copy<WU>(reinterpret_cast<const uint4 *>(m.ngw), stage, sr);
        stage_sync();
    }
    if (!p.live) return;
    const int L = m.nlev[p.col] - 1, k0 = p.k0;
    if (k0 >= L) return;  // (every live pair starts on a layer; kept as a guard)
    const bool v1 = k0 + 1 < L;
    const NodeRing<D> rg = node_ring<D>(m)[p.col];
    int nr = 0;
#pragma unroll
    for (int j = 0; j < D; ++j) nr += rg.oo[j] >= 0 ? 1 : 0;
    V2<double> Tn[NT], first[NT], prev[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        Tn[t] = ld2(ab + t * tsn + p.c);
        first[t] = prev[t] = V2<double>{0.0, 0.0};
    }
    R tvol0 = 0, tvol1 = 0;
    R tx0[NT], ty0[NT], tx1[NT], ty1[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) { tx0[t] = 0; ty0[t] = 0; tx1[t] = 0; ty1[t] = 0; }
    const R *W = m.ngw + (size_t)8 * D * p.col;
    const uint4 *WS = stage + (p.col - sr.c0) * WU;
#pragma unroll
    for (int j = 0; j <= D; ++j) {
        V2<double> cur[NT];
        if (j < D && rg.oo[j] >= 0) {
            // a ring node shallower than this pair borders only elements that
            // are dry here (masked below): its column's first pair stands in
            const int kk = k0 < (int)rg.lr[j] ? k0 : 0;
#pragma unroll
            for (int t = 0; t < NT; ++t) cur[t] = ld2(ab + t * tsn + rg.oo[j] + kk);
        } else {
#pragma unroll
            for (int t = 0; t < NT; ++t) cur[t] = V2<double>{0.0, 0.0};
        }
        if (j == 0) {
#pragma unroll
            for (int t = 0; t < NT; ++t) first[t] = cur[t];
        } else if (rg.le[j - 1] != 0) {
            const int le = rg.le[j - 1];
            const bool m0 = k0 < le, m1 = k0 + 1 < le;
            const bool wrap = j >= nr;
            R wax, wbx, way, wby, a;
            if (sr.staged) fan_weights_lds(WS + (j - 1) * (8 * (int)sizeof(R) / 16), wax, wbx, way, wby, a);
            else fan_weights(W + 8 * (j - 1), wax, wbx, way, wby, a);
            tvol0 = m0 ? tvol0 + a : tvol0;
            tvol1 = m1 ? tvol1 + a : tvol1;
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const V2<double> b = wrap ? first[t] : cur[t];
                const R da0 = (R)(prev[t].x - Tn[t].x), db0 = (R)(b.x - Tn[t].x);
                const R da1 = (R)(prev[t].y - Tn[t].y), db1 = (R)(b.y - Tn[t].y);
                const R gx0 = wax * da0 + wbx * db0, gy0 = way * da0 + wby * db0;
                const R gx1 = wax * da1 + wbx * db1, gy1 = way * da1 + wby * db1;
                tx0[t] = m0 ? tx0[t] + gx0 : tx0[t];
                ty0[t] = m0 ? ty0[t] + gy0 : ty0[t];
                tx1[t] = m1 ? tx1[t] + gx1 : tx1[t];
                ty1[t] = m1 ? ty1[t] + gy1 : ty1[t];
            }
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) prev[t] = cur[t];
    }
#pragma unroll
    for (int t = 0; t < NT; ++t)
        GradIO<G>::st(gnod + t * tsg + 2 * p.c, tx0[t] / tvol0, ty0[t] / tvol0, v1 ? tx1[t] / tvol1 : (R)0,
                      v1 ? ty1[t] / tvol1 : (R)0, gscale);
}

// ===========================================================================
// Horizontal edge fluxes (lib/oce_adv_tra_hor.F90): UPW1 (:64-261),
// MUSCL (:265-550), MFCT (:554-846). Segments C (both elements), D (el1
// only) and E (el2 only); no cavities, so segments A/B are empty.
// FCT: writes HO-LO (HO from ttfAB, LO upwind from ttf) per level and tracer,
// and the tracer-independent transport vflux per level, from which the node
// side re-forms the LO fluxes. No limiter: writes the HO flux. Gradients follow fill_up_dn_grad: the
// up/down-wind triangle on levels shared by all elements around both nodes,
// nodal means below.
// ===========================================================================
// LO upwind edge flux (adv_tra_hor_upw1, lib/oce_adv_tra_hor.F90:64-261) from
// the edge transport v and the tracer at edges(1), edges(2)
__device__ __forceinline__ double lo_flux(double v, double T1, double T2)
{
    return -0.5 * (T1 * (v + fabs(v)) + T2 * (v - fabs(v))) - 0.0;
}

template <class R, class G, int HOR, bool FCT, int NT, bool SAME>
__global__ __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 && NT <= 2 ? 7 : 1))) void k_edge(DMesh<R> m, const R *__restrict__ vel,
                                                 const double *__restrict__ ttf,
                                                 const double *__restrict__ ttfAB, size_t tsn,
                                                 const G *__restrict__ gnod, size_t tsg,
                                                 R num_ord, float ginv,
                                                 R *__restrict__ vfl, R *__restrict__ flux, size_t tse)
{
    const int ch = wave_chunk();
    if (ch >= m.nch_e) return;
    const Pair p = lane_pair(m.slot_e, ch);
    // the chunk's edges are consecutive columns; when there are at most
    // kStageEdges of them the wave loads their packed records into its LDS
    // slice with cooperative 16-byte loads (one instruction moves 64 units)
    // and every lane reads its edge's record from there
    constexpr int U = kEdgeUnits<R>;
    extern __shared__ uint4 stage_all[];
    uint4 *stage = stage_all + (threadIdx.x >> 6) * (kStageEdges * U);
    const StageRange sr = stage_range(p, kStageEdges);
    if (sr.staged) {
        stage_copy<U>(m.epack, stage, sr);
        stage_sync();
    }
    if (!p.live) return;
    uint4 rec[U];
    if (sr.staged) {
        const uint4 *q = stage + (p.col - sr.c0) * U;
#pragma unroll
        for (int u = 0; u < U; ++u) rec[u] = q[u];
    } else {
        const uint4 *q = m.epack + (size_t)p.col * U;
#pragma unroll
        for (int u = 0; u < U; ++u) rec[u] = q[u];
    }
    EdgeRec er;
    R geo[8], gud[4];
    int4 eud4;
    __builtin_memcpy(&er, &rec[0], 32);
    __builtin_memcpy(geo, &rec[2], 8 * sizeof(R));
    __builtin_memcpy(&eud4, reinterpret_cast<const char *>(rec) + 32 + 8 * sizeof(R), 16);
    __builtin_memcpy(gud, reinterpret_cast<const char *>(rec) + 48 + 8 * sizeof(R), 4 * sizeof(R));
    const int nl1 = er.nl & 255, nl2 = (er.nl >> 8) & 255, nzmax_e = (er.nl >> 16) & 255;
    const int nlmax = (int)((unsigned)er.nl >> 24);
    const int nl12 = min(nl1, nl2);
    const int k0 = p.k0, nz0 = k0 + 1, nz1 = k0 + 2;
    const bool valid1 = nz1 <= nlmax;
    const R dx1 = geo[0], dy1 = geo[1], dx2 = geo[2], dy2 = geo[3];
    const bool el2 = er.t2 >= 0;
    const bool h10 = nz0 <= nl1, h11 = valid1 && nz1 <= nl1;
    const bool h20 = el2 && (nz0 <= nl12 || nz0 > nl1);
    const bool h21 = el2 && valid1 && (nz1 <= nl12 || nz1 > nl1);
    R vf10 = 0, vf11 = 0, vf20 = 0, vf21 = 0;
    if (h10) {
        const V4<R> u = ld4(vel + 2 * ((size_t)er.t1 + k0));
        const V2<R> h = ld2(m.helem + er.t1 + k0);
        vf10 = (-u.y0 * dx1 + u.x0 * dy1) * h.x;
        vf11 = (-u.y1 * dx1 + u.x1 * dy1) * h.y;
    }
    if (h20) {
        const V4<R> u = ld4(vel + 2 * ((size_t)er.t2 + k0));
        const V2<R> h = ld2(m.helem + er.t2 + k0);
        vf20 = (u.y0 * dx2 - u.x0 * dy2) * h.x;
        vf21 = (u.y1 * dx2 - u.x1 * dy2) * h.y;
    }
    const R vflux0 = (h10 && h20) ? vf10 + vf20 : (h10 ? vf10 : vf20);
    const R vflux1 = (h11 && h21) ? vf11 + vf21 : (h11 ? vf11 : (h21 ? vf21 : (R)0));

    // MUSCL/MFCT: reconstruction geometry, gradient source per level
    R ex = 0, ey = 0, a = 0, clo10 = 1, clo20 = 1, clo11 = 1, clo21 = 1;
    bool sh0 = false, sh1 = false;
    int4 ud{0, 0, 0, 0};
    V4<R> wud{0, 0, 0, 0};
    if constexpr (HOR != HOR_UPW1) {
        ud = eud4;
        wud = V4<R>{gud[0], gud[1], gud[2], gud[3]};
        ex = geo[4]; ey = geo[5]; a = geo[6];
        if constexpr (HOR == HOR_MUSCL) {
            const int nb1 = er.nb & 0xFFFF, nb2 = er.nb >> 16;
            clo10 = (nb1 - nz0) >= 0 ? (R)1 : (R)0;
            clo20 = (nb2 - nz0) >= 0 ? (R)1 : (R)0;
            clo11 = (nb1 - nz1) >= 0 ? (R)1 : (R)0;
            clo21 = (nb2 - nz1) >= 0 ? (R)1 : (R)0;
        }
        const bool updn = er.tu >= 0 && er.td >= 0;
        sh0 = updn && nz0 <= nzmax_e - 1;
        sh1 = updn && nz1 <= nzmax_e - 1;
    }
    const size_t o = (size_t)p.c;
    R q0 = 0, q1 = 0;  // FP32: the even tracer's pair, stored with the odd one's
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const size_t tn = t * tsn;
        // the LO flux in FP64 whatever R is (it must see the FP64 tracer
        // unrounded for the LO solution to stay monotone); HO in R
        double lo0 = 0.0, lo1 = 0.0;
        V2<double> T1{0.0, 0.0}, T2{0.0, 0.0};
        if constexpr (FCT || SAME) {
            T1 = ld2(ttf + tn + er.o1 + k0);
            T2 = ld2(ttf + tn + er.o2 + k0);
        }
        if constexpr (FCT) {
            lo0 = lo_flux((double)vflux0, T1.x, T2.x);
            if (valid1) lo1 = lo_flux((double)vflux1, T1.y, T2.y);
            if (t == 0) st2(vfl + o, vflux0, valid1 ? vflux1 : (R)0);
        }
        V2<double> B1 = T1, B2 = T2;
        if constexpr (!SAME) {
            B1 = ld2(ttfAB + tn + er.o1 + k0);
            B2 = ld2(ttfAB + tn + er.o2 + k0);
        }
        const R A10 = (R)B1.x, A20 = (R)B2.x, A11 = (R)B1.y, A21 = (R)B2.y;
        R ho0, ho1;
        if constexpr (HOR == HOR_UPW1) {
            ho0 = (R)(-0.5) * (A10 * (vflux0 + fabs(vflux0)) + A20 * (vflux0 - fabs(vflux0)));
            ho1 = (R)(-0.5) * (A11 * (vflux1 + fabs(vflux1)) + A21 * (vflux1 - fabs(vflux1)));
        } else {
            // the gradient along the edge, ex*a*d/dx + ey*r_earth*d/dy: of the
            // up/down-wind triangle on the levels all elements around both
            // nodes share (sum w T over its nodes), of the nodal means below
            R du0 = 0, dd0 = 0, du1 = 0, dd1 = 0;
            if (sh0) {
                const double *A = (SAME ? ttf : ttfAB) + tn + k0;
                const V2<double> ua = ld2(A + ud.x), ub = ld2(A + ud.y), da = ld2(A + ud.z), db = ld2(A + ud.w);
                du0 = wud.x0 * (R)(ua.x - B1.x) + wud.y0 * (R)(ub.x - B1.x);
                dd0 = wud.x1 * (R)(da.x - B2.x) + wud.y1 * (R)(db.x - B2.x);
                du1 = wud.x0 * (R)(ua.y - B1.y) + wud.y0 * (R)(ub.y - B1.y);
                dd1 = wud.x1 * (R)(da.y - B2.y) + wud.y1 * (R)(db.y - B2.y);
            }
            V4<R> n1{0, 0, 0, 0}, n2{0, 0, 0, 0};
            if (!sh0 || (valid1 && !sh1)) {
                const auto u = GradIO<G>::ld(gnod + t * tsg + 2 * ((size_t)er.o1 + k0), ginv);
                const auto d = GradIO<G>::ld(gnod + t * tsg + 2 * ((size_t)er.o2 + k0), ginv);
                n1 = {(R)u.x0, (R)u.y0, (R)u.x1, (R)u.y1};
                n2 = {(R)d.x0, (R)d.y0, (R)d.x1, (R)d.y1};
            }
            const R dot10 = sh0 ? du0 : ex * a * n1.x0 + ey * (R)kRE * n1.y0;
            const R dot20 = sh0 ? dd0 : ex * a * n2.x0 + ey * (R)kRE * n2.y0;
            const R dot11 = sh1 ? du1 : ex * a * n1.x1 + ey * (R)kRE * n1.y1;
            const R dot21 = sh1 ? dd1 : ex * a * n2.x1 + ey * (R)kRE * n2.y1;
            {
                const R Tmean2 = A20 - ((R)2 * (A20 - A10) + dot20) / (R)6 * clo20;
                const R Tmean1 = A10 + ((R)2 * (A20 - A10) + dot10) / (R)6 * clo10;
                const R cHO = (vflux0 + fabs(vflux0)) * Tmean1 + (vflux0 - fabs(vflux0)) * Tmean2;
                ho0 = (R)(-0.5) * ((R)1 - num_ord) * cHO - vflux0 * num_ord * (R)0.5 * (Tmean1 + Tmean2);
            }
            {
                const R Tmean2 = A21 - ((R)2 * (A21 - A11) + dot21) / (R)6 * clo21;
                const R Tmean1 = A11 + ((R)2 * (A21 - A11) + dot11) / (R)6 * clo11;
                const R cHO = (vflux1 + fabs(vflux1)) * Tmean1 + (vflux1 - fabs(vflux1)) * Tmean2;
                ho1 = (R)(-0.5) * ((R)1 - num_ord) * cHO - vflux1 * num_ord * (R)0.5 * (Tmean1 + Tmean2);
            }
        }
        const R v0 = (R)((double)ho0 - lo0), v1 = valid1 ? (R)((double)ho1 - lo1) : (R)0;
        if constexpr (sizeof(R) == 4 && NT % 2 == 0) {
            // the two slots of a tracer pair share 16 bytes per cell pair: one store
            if (t & 1) st4(reinterpret_cast<float *>(flux + adf_at<R>(t - 1, tse, o)), q0, q1, v0, v1);
            else { q0 = v0; q1 = v1; }
        } else {
            st2_mixed(flux + adf_at<R>(t, tse, o), v0, v1);
        }
    }
}

// the geometry a node pair's vertical scheme reads (all lanes must call)
template <class R, int VER>
__device__ __forceinline__ VCol2<R> node_vcol(const DMesh<R> &m, const Pair &p, int nlev, const R *__restrict__ Wv)
{
    VCol2<R> c{};
    c.nlev = nlev;
    c.k0 = p.k0;
    const size_t i = (size_t)p.c;
    V2<R> W{0, 0}, ar{0, 0}, zb{0, 0}, Z{0, 0}, h{0, 0}, hn{0, 0};
    if (p.live) {
        W = ld2(Wv + i);
        ar = ld2(m.area + i);
        if constexpr (VER == VER_QR4C) { zb = ld2(m.zbarn + i); Z = ld2(m.Zn + i); }
        if constexpr (VER == VER_PPM) { h = ld2(m.hnode + i); hn = ld_same(m.hnode_new, m.hnode, i, h); }
    }
    c.W = W;
    c.area = ar;
    if constexpr (VER == VER_QR4C) { c.zbar = zb; c.Z = window(Z); }
    if constexpr (VER == VER_PPM) { c.h = window(h); c.hn = window(hn); }
    return c;
}

// ===========================================================================
// LO solution (lib/oce_adv_tra_driver.F90:95-211) and HO-LO vertical flux
// (:257-283). The antidiffusive sums P+/P- that the reference forms next
// (lib/oce_adv_tra_fct.F90:232-346) are formed in k_node_limit, which reads
// the HO-LO fluxes straight away instead of a stored copy of the sums
// ===========================================================================
template <class R, int VER, int NT, bool SAME, int D>
__global__ __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 && NT <= 2 && VER != VER_PPM ? 4 : 1))) void k_node_lo(DMesh<R> m, const R *__restrict__ Wv,
                                                    const R *__restrict__ WEv,
                                                    const double *__restrict__ ttf,
                                                    const double *__restrict__ ttfAB, size_t tsn,
                                                    const R *__restrict__ vfl,
                                                    double dt, R num_ord,
                                                    double *__restrict__ LO, R *__restrict__ adf_v)
{
    const int ch = wave_chunk(true);  // from the end of each XCD's range (wave_chunk)
    if (ch >= m.nch_n) return;
    const Pair p = lane_pair(m.slot_n, ch);
    const int nlev = p.live ? m.nlev[p.col] : 0;
    const int L = nlev - 1, k0 = p.k0;
    const bool lay0 = k0 < L, lay1 = k0 + 1 < L;
    const bool if0 = p.live && k0 < nlev, if1 = p.live && k0 + 1 < nlev;
    const size_t c = (size_t)p.c;
    int oe[D], oo[D];
    unsigned info[D];
    extern __shared__ uint4 stage_all[];
    uint4 *slice = stage_all + (threadIdx.x >> 6) * (kStageNodes * kNodeEdgeUnits<D>);
    const StageRange sr = stage_node_edges<D>(m, p, slice);
    load_node_edges<D, true>(m, sr, slice, p.col, p.live && lay0, oe, info, oo);
    // the LO vertical flux in FP64 (see lo_flux)
    VCol2<double> lc{};
    lc.nlev = nlev;
    lc.k0 = k0;
    if (p.live) {
        const V2<R> we = ld2(WEv + (size_t)c), ar = ld2(m.area + c);
        lc.W = {(double)we.x, (double)we.y};
        lc.area = {(double)ar.x, (double)ar.y};
    }
    // tracers in groups of TG, each group start to end (vertical fluxes, then
    // the edge sums), so that only one group's state is live at a time
    constexpr int TG = NT < 2 ? NT : 2;  // tracers whose gathers share a round trip
#pragma unroll
    for (int t0 = 0; t0 < NT; t0 += TG) {
        // vertical fluxes of the group's tracers first: the column
        // reconstruction's state (windows of Z, zbar, W) is dead before the
        // edge loop starts (every lane takes part: the windows cross lanes)
        V2<double> T[TG];
        double fl0[TG], fl1[TG], fl2[TG];
        {
            VCol2<R> hc = node_vcol<R, VER>(m, p, nlev, Wv);
#pragma unroll
            for (int u = 0; u < TG; ++u) {
                const size_t tn = (t0 + u) * tsn;
                T[u] = V2<double>{0.0, 0.0};
                if (p.live) T[u] = ld2(ttf + tn + c);
                lc.A = tracer_window<double>(T[u], k0, nlev);
                // LO vertical flux (UPW1 with w_e and ttf) and HO-LO with w and ttfAB
                fl0[u] = vflux_upw1<0>(lc, 0.0);
                fl1[u] = vflux_upw1<1>(lc, 0.0);
                fl2[u] = below(from_next(fl0[u]), k0, L);
                if constexpr (SAME) {
                    hc.A = tracer_window<R>(T[u], k0, nlev);
                } else {
                    V2<double> B{0.0, 0.0};
                    if (p.live) B = ld2(ttfAB + tn + c);
                    hc.A = tracer_window<R>(B, k0, nlev);
                }
                R ad0, ad1;
                ver_flux<R, VER>(hc, num_ord, (R)dt, (R)fl0[u], (R)fl1[u], ad0, ad1);  // HO - LO
                if (p.live) st2(adf_v + tn + c, if0 ? ad0 : (R)0, if1 ? ad1 : (R)0);
            }
        }
        if (!lay0) continue;
        // LO horizontal sums over incident edges, edge-index order; edge slots
        // outer and the group's tracers inner, so that one memory round trip
        // per edge carries its transport and the other node's value of TG tracers
        double slo0[TG], slo1[TG];
#pragma unroll
        for (int u = 0; u < TG; ++u) { slo0[u] = 0.0; slo1[u] = 0.0; }
#pragma unroll
        for (int j = 0; j < D; ++j) {
            const int nlm = info_nlmax(info[j]);
            if (oe[j] < 0 || k0 >= nlm) continue;
            const bool on1 = lay1 && k0 + 1 < nlm;
            const bool pos = info_pos(info[j]);
            // LO upwind flux of the edge (adv_tra_hor_upw1, lib/oce_adv_tra_hor.F90:64-261)
            // from its transport and the tracer at edges(1), edges(2);
            // seen from edges(2) the flux enters with the opposite sign:
            // -lo_flux(v, To, T) == lo_flux(-v, T, To) exactly (sign flips and a
            // commuted two-term sum), so one sign on the transport orients it
            const double sg = pos ? 1.0 : -1.0;
            const V2<R> v = ld2(vfl + (size_t)oe[j] + k0);
            V2<double> To[TG];
#pragma unroll
            for (int u = 0; u < TG; ++u) To[u] = ld2(ttf + (t0 + u) * tsn + (size_t)oo[j] + k0);
#pragma unroll
            for (int u = 0; u < TG; ++u) {
                slo0[u] = slo0[u] + lo_flux((double)v.x * sg, T[u].x, To[u].x);
                // second cell by selects, so that each pair stays one 16-byte load
                const double s1 = slo1[u] + lo_flux((double)v.y * sg, T[u].y, To[u].y);
                slo1[u] = on1 ? s1 : slo1[u];
            }
        }
        // FP64 geometry (the LO solution feeds the conservative update); without
        // cavities areasvol equals area on every wet layer
        V2<double> av;
        if constexpr (std::is_same<R, double>::value) av = lc.area;
        else av = ld2(m.av_d + c);
        const V2<double> h = ld2(m.h_d + c), hn = ld_same(m.hn_d, m.h_d, c, h);
#pragma unroll
        for (int u = 0; u < TG; ++u) {
            // flux differences formed in FP64: each interface flux then leaves one
            // layer and enters the next with exactly opposite sign in every precision
            const double LO0 =
                (T[u].x * (double)h.x + (slo0[u] + (fl0[u] - fl1[u])) * dt / (double)av.x) / (double)hn.x;
            // second layer branch-free (selected away below a one-layer pair)
            const double L1 =
                (T[u].y * (double)h.y + (slo1[u] + (fl1[u] - fl2[u])) * dt / (double)av.y) / (double)hn.y;
            st2(LO + (t0 + u) * tsn + c, LO0, lay1 ? L1 : 0.0);
        }
    }
}

// The raw antidiffusive sums P+/P- (lib/oce_adv_tra_fct.F90:232-346: the
// vertical HO-LO fluxes of the layer's two interfaces, then the horizontal
// ones of the incident edges in edge-index order, each oriented into the
// node), Zalesak bounds a1-a3 and limiting factors b2 (:94-217,361-374).
// a1 (max/min of LO and T) is formed once per 1-ring node and a2/a3 as one
// running max/min over the ring.
template <class R, int NT, bool STOREQ, int D>
__global__ __launch_bounds__(kMaxBlock) void k_node_limit(DMesh<R> m, const double *__restrict__ ttf,
                                                       const double *__restrict__ LO, size_t tsn,
                                                       const R *__restrict__ adf_h, size_t tse,
                                                       const R *__restrict__ adf_v,
                                                       double dt, R *__restrict__ Rp,
                                                       R *__restrict__ Rm, R *__restrict__ Qp_out,
                                                       R *__restrict__ Qm_out, R *__restrict__ Pp_out,
                                                       R *__restrict__ Pm_out)
{
    const int ch = wave_chunk();
    if (ch >= m.nch_n) return;
    const Pair p = lane_pair(m.slot_n, ch);
    const int nlev = p.live ? m.nlev[p.col] : 0;
    const int L = nlev - 1, k0 = p.k0;
    const bool lay0 = k0 < L, lay1 = k0 + 1 < L;
    const size_t c = (size_t)p.c;
    // bounds in FP64 whatever R is: they are selections among FP64 tracer
    // values, so no new extrema appear beyond the flux arithmetic's rounding
    const double bignumber = 1000.0;
    constexpr int TG = NT < 2 ? NT : 2;  // tracers whose gathers share a round trip
    static_assert(NT % TG == 0, "tracer batches are 1, 2 or 4");

    // the chunk's node->edge records go to LDS first (every lane of the wave
    // takes part); then per group of TG tracers the bounds, the P+/P- sums
    // and R+/R-, so that only one group's state is live at a time
    extern __shared__ uint4 stage_all[];
    uint4 *slice = stage_all + (threadIdx.x >> 6) * (kStageNodes * kNodeEdgeUnits<D>);
    const StageRange sr = stage_node_edges<D>(m, p, slice);

    NodeRing<D> rg;
    int lw[D];  // layers of the deeper element next to ring node j
    bool dry0 = false, dry1 = false;
    {
        // a2/a3 as one max/min per level: element e's bound is the max (min)
        // of max(LO, T) (min(LO, T)) over its three nodes when it has the
        // level, -bignumber (+bignumber) when it does not; the node's bound is
        // the max (min) over its elements. The node itself and every ring
        // node next to an element that has the level enter it, plus the
        // padding value when some element lacks the level (max and min are
        // exact, so this equals the element-by-element form bit for bit).
        // Fan element j lies between ring nodes j and j + 1 (ring node 0 where
        // the fan closes, j + 1 == nr); le[j] = 0 where a boundary fan is open
        if (lay0) rg = node_ring<D>(m)[p.col];
        else {
#pragma unroll
            for (int j = 0; j < D; ++j) { rg.oo[j] = -1; rg.lr[j] = 0; rg.le[j] = 0; }
        }
        int nr = 0, le_last = 0, lemin = 255;
#pragma unroll
        for (int j = 0; j < D; ++j) nr += rg.oo[j] >= 0 ? 1 : 0;
#pragma unroll
        for (int j = 0; j < D; ++j) {
            le_last = j == nr - 1 ? (int)rg.le[j] : le_last;
            lemin = rg.le[j] != 0 ? min(lemin, (int)rg.le[j]) : lemin;
        }
#pragma unroll
        for (int j = 0; j < D; ++j) lw[j] = rg.oo[j] >= 0 ? max(j == 0 ? le_last : (int)rg.le[j - 1], (int)rg.le[j]) : 0;
        dry0 = k0 >= lemin;
        dry1 = k0 + 1 >= lemin;
    }
    const R flux_eps = (R)1e-16;
#pragma unroll
    for (int t0 = 0; t0 < NT; t0 += TG) {
        // bounds: ring nodes outer, the group's tracers inner, one memory
        // round trip per ring node carrying the LO and tracer pairs of TG tracers
        double tvmax0[TG], tvmin0[TG], tvmax1[TG], tvmin1[TG];
#pragma unroll
        for (int u = 0; u < TG; ++u) { tvmax0[u] = 0; tvmin0[u] = 0; tvmax1[u] = 0; tvmin1[u] = 0; }
        if (lay0) {
            double vmax0[TG], vmin0[TG], vmax1[TG], vmin1[TG];
#pragma unroll
            for (int u = 0; u < TG; ++u) {
                const size_t tn = (t0 + u) * tsn;
                const V2<double> ls = ld2(LO + tn + c), ts = ld2(ttf + tn + c);
                vmax0[u] = fmx(ls.x, ts.x); vmin0[u] = fmn(ls.x, ts.x);
                vmax1[u] = fmx(ls.y, ts.y); vmin1[u] = fmn(ls.y, ts.y);
            }
#pragma unroll
            for (int j = 0; j < D; ++j) {
                if (k0 >= lw[j]) continue;
                const bool on1 = k0 + 1 < lw[j];
                V2<double> l[TG], v[TG];
#pragma unroll
                for (int u = 0; u < TG; ++u) {
                    const size_t o = (t0 + u) * tsn + (size_t)rg.oo[j] + k0;
                    l[u] = ld2(LO + o);
                    v[u] = ld2(ttf + o);
                }
#pragma unroll
                for (int u = 0; u < TG; ++u) {
                    vmax0[u] = fmx(vmax0[u], fmx(l[u].x, v[u].x));
                    vmin0[u] = fmn(vmin0[u], fmn(l[u].x, v[u].x));
                    const double x1 = fmx(vmax1[u], fmx(l[u].y, v[u].y));
                    const double n1 = fmn(vmin1[u], fmn(l[u].y, v[u].y));
                    vmax1[u] = on1 ? x1 : vmax1[u];
                    vmin1[u] = on1 ? n1 : vmin1[u];
                }
            }
#pragma unroll
            for (int u = 0; u < TG; ++u) {
                tvmax0[u] = dry0 ? fmx(vmax0[u], -bignumber) : vmax0[u];
                tvmin0[u] = dry0 ? fmn(vmin0[u], bignumber) : vmin0[u];
                tvmax1[u] = lay1 ? (dry1 ? fmx(vmax1[u], -bignumber) : vmax1[u]) : 0.0;
                tvmin1[u] = lay1 ? (dry1 ? fmn(vmin1[u], bignumber) : vmin1[u]) : 0.0;
            }
        }
        // the node->edge records (from the LDS slice when staged), loaded
        // here so that they are not live during the bounds
        int oe[D], oo[D];
        unsigned info[D];
        load_node_edges<D, true>(m, sr, slice, p.col, lay0, oe, info, oo);
        V2<R> av{1, 1};
        if (lay0) av = ld2_nt(m.areasvol + c);
        // P+/P-: the vertical HO-LO fluxes of interfaces k0+1, k0+2 (layer k0)
        // and k0+2, k0+3 (layer k0+1; the last one in the next lane: where it
        // is used the layer below exists, so that lane holds the same column),
        // then the incident edges in edge-index order, each oriented into the
        // node; edge slots outer and the group's tracers inner, so that one
        // memory round trip per edge carries the flux pairs of TG tracers
        R pp0[TG], pm0[TG], pp1[TG], pm1[TG];
#pragma unroll
        for (int u = 0; u < TG; ++u) {
            V2<R> ad{0, 0};
            if (p.live) ad = ld2_nt(adf_v + (t0 + u) * tsn + c);
            const R ad0 = ad.x, ad1 = ad.y, ad2 = below(from_next(ad0), k0, L);
            pp0[u] = (R)0 + (fmx((R)0, ad0) + fmx((R)0, -ad1));
            pm0[u] = (R)0 + (fmn((R)0, ad0) + fmn((R)0, -ad1));
            pp1[u] = (R)0 + (fmx((R)0, ad1) + fmx((R)0, -ad2));
            pm1[u] = (R)0 + (fmn((R)0, ad1) + fmn((R)0, -ad2));
        }
#pragma unroll
        for (int j = 0; j < D; ++j) {
            const int nlm = info_nlmax(info[j]);
            if (oe[j] < 0 || k0 >= nlm) continue;
            const bool on1 = lay1 && k0 + 1 < nlm;
            const bool pos = info_pos(info[j]);
            V2<R> a[TG];
            if constexpr (TG == 2) ld_adf2(adf_h, t0, tse, (size_t)oe[j] + k0, a[0], a[1]);
            else a[0] = ld2(adf_h + adf_at<R>(t0, tse, (size_t)oe[j] + k0));
#pragma unroll
            for (int u = 0; u < TG; ++u) {
                const R ax = pos ? a[u].x : -a[u].x;
                pp0[u] = pp0[u] + fmx((R)0, ax);
                pm0[u] = pm0[u] + fmn((R)0, ax);
                // second cell by selects, so that each pair stays one 16-byte load
                const R ay = pos ? a[u].y : -a[u].y;
                const R q1 = pp1[u] + fmx((R)0, ay), r1 = pm1[u] + fmn((R)0, ay);
                pp1[u] = on1 ? q1 : pp1[u];
                pm1[u] = on1 ? r1 : pm1[u];
            }
        }
#pragma unroll
        for (int u = 0; u < TG; ++u) {
            const int t = t0 + u;
            const double upx0 = from_prev(tvmax1[u]), upn0 = from_prev(tvmin1[u]);
            const double dnx1 = from_next(tvmax0[u]), dnn1 = from_next(tvmin0[u]);
            if (!lay0) continue;
            const size_t tn = t * tsn;
            const V2<double> lo = ld2(LO + tn + c);
            const V2<R> P{pp0[u], lay1 ? pp1[u] : (R)0}, M{pm0[u], lay1 ? pm1[u] : (R)0};
            R rp0, rm0, rp1 = 0, rm1 = 0;
            double qp0 = 0.0, qm0 = 0.0, qp1 = 0.0, qm1 = 0.0;
            {
                const int nz = k0 + 1;
                double qp, qm;
                if (nz == 1 || nz == L) {
                    qp = tvmax0[u] - lo.x;
                    qm = tvmin0[u] - lo.x;
                } else {
                    qp = fmx(fmx(upx0, tvmax0[u]), tvmax1[u]) - lo.x;
                    qm = fmn(fmn(upn0, tvmin0[u]), tvmin1[u]) - lo.x;
                }
                R flux = P.x * (R)dt / av.x + flux_eps;
                rp0 = fmn((R)1, (R)qp / flux);
                flux = M.x * (R)dt / av.x - flux_eps;
                rm0 = fmn((R)1, (R)qm / flux);
                qp0 = qp;
                qm0 = qm;
            }
            if (lay1) {
                const int nz = k0 + 2;
                double qp, qm;
                if (nz == L) {
                    qp = tvmax1[u] - lo.y;
                    qm = tvmin1[u] - lo.y;
                } else {
                    qp = fmx(fmx(tvmax0[u], tvmax1[u]), dnx1) - lo.y;
                    qm = fmn(fmn(tvmin0[u], tvmin1[u]), dnn1) - lo.y;
                }
                R flux = P.y * (R)dt / av.y + flux_eps;
                rp1 = fmn((R)1, (R)qp / flux);
                flux = M.y * (R)dt / av.y - flux_eps;
                rm1 = fmn((R)1, (R)qm / flux);
                qp1 = qp;
                qm1 = qm;
            }
            st_rpm(Rp, Rm, tn + c, rp0, rp1, rm0, rm1);
            if constexpr (STOREQ) {  // Q+/Q- and P+/P- for the learned limiter's features and bound check
                st_rpm(Qp_out, Qm_out, tn + c, (R)qp0, (R)qp1, (R)qm0, (R)qm1);
                st_rpm(Pp_out, Pm_out, tn + c, P.x, P.y, M.x, M.y);
            }
        }
    }
}

// vertical limiting of interface nz (b3, lib/oce_adv_tra_fct.F90:401-428):
// rp/rm of the layer below the interface, rpu/rmu of the layer above
template <class R>
__device__ __forceinline__ R limit_vertical(int nz, bool layer, R fv, R rp, R rm, R rpu, R rmu)
{
    if (nz == 1) {
        R ae = (R)1;
        ae = fv >= (R)0 ? fmn(ae, rp) : fmn(ae, rm);
        return ae * fv;
    }
    if (layer) {
        R ae = (R)1;
        if (fv >= (R)0) { ae = fmn(ae, rmu); ae = fmn(ae, rp); }
        else            { ae = fmn(ae, rpu); ae = fmn(ae, rm); }
        return ae * fv;
    }
    return fv;
}

// limiter of one edge flux f seen from a node (b3, lib/oce_adv_tra_fct.F90:431-445):
// pos = the node is edges(1); rp/rm its R+/R-, rpo/rmo those of the other node
template <class R>
__device__ __forceinline__ R edge_alpha(bool pos, R f, R rp, R rm, R rpo, R rmo)
{
    const R rp1 = pos ? rp : rpo, rm1 = pos ? rm : rmo;
    const R rp2 = pos ? rpo : rp, rm2 = pos ? rmo : rm;
    R ae = (R)1;
    if (f >= (R)0) { ae = fmn(ae, rp1); ae = fmn(ae, rm2); }
    else           { ae = fmn(ae, rm1); ae = fmn(ae, rp2); }
    return ae;
}

// the end of a step at one node layer (OUT): del_ttf = (del + dttf_h) + dttf_v
// with del = 0 and values += del_ttf / hnode_new (OUT_UPDATE, the analytic
// driver, src/fesom_analytic.F90:405-437), the same with del carried over from
// the last step in acc_h (OUT_ACCUM, the reference C interface's del_ttf,
// lib/tracer_c_interface.F90:194-212), or dttf_h -> acc_h and dttf_v -> acc_v
// with the tracer left alone (OUT_TEND, what do_oce_adv_tra itself returns)
template <int OUT>
__device__ __forceinline__ void node_out(double *__restrict__ ttf, double *__restrict__ acc_h,
                                         double *__restrict__ acc_v, size_t i, bool lay1, V2<double> T,
                                         double sh0, double dv0, double sh1, double dv1, V2<double> hn)
{
    if constexpr (OUT == OUT_TEND) {
        st2(acc_h + i, 0.0 + sh0, lay1 ? 0.0 + sh1 : 0.0);
        st2(acc_v + i, dv0, lay1 ? dv1 : 0.0);
    } else {
        V2<double> d{0.0, 0.0};
        if constexpr (OUT == OUT_ACCUM) d = ld2(acc_h + i);
        const double del0 = (d.x + sh0) + dv0, del1 = (d.y + sh1) + dv1;
        if constexpr (OUT == OUT_ACCUM) st2(acc_h + i, del0, lay1 ? del1 : 0.0);
        st2(ttf + i, T.x + del0 / hn.x, lay1 ? T.y + del1 / hn.y : T.y);
    }
}

// limiting b3 (lib/oce_adv_tra_fct.F90:394-471), flux divergence
// (lib/oce_adv_tra_driver.F90:298-450) and the end of the step (node_out)
template <class R, int NT, bool LEARN, int D, int OUT>
__global__ __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 && NT <= 2 && !LEARN ? 5 : 1))) void k_node_update(DMesh<R> m, double *__restrict__ ttf,
                                                        const double *__restrict__ LO, size_t tsn,
                                                        const R *__restrict__ adf_h, size_t tse,
                                                        const R *__restrict__ adf_v,
                                                        const R *__restrict__ Rp,
                                                        const R *__restrict__ Rm, double dt,
                                                        const R *__restrict__ alpha,
                                                        double *__restrict__ acc_h, double *__restrict__ acc_v)
{
    const int ch = wave_chunk(true);  // from the end of each XCD's range (wave_chunk)
    if (ch >= m.nch_n) return;
    const Pair p = lane_pair(m.slot_n, ch);
    const int nlev = p.live ? m.nlev[p.col] : 0;
    const int L = nlev - 1, k0 = p.k0;
    const bool lay0 = k0 < L, lay1 = k0 + 1 < L;
    const bool if0 = p.live && k0 < nlev, if1 = p.live && k0 + 1 < nlev;
    const size_t c = (size_t)p.c;
    int oe[D], oo[D];
    unsigned info[D];
    extern __shared__ uint4 stage_all[];
    uint4 *slice = stage_all + (threadIdx.x >> 6) * (kStageNodes * kNodeEdgeUnits<D>);
    const StageRange sr = stage_node_edges<D>(m, p, slice);
    load_node_edges<D, true>(m, sr, slice, p.col, lay0, oe, info, oo);
    V2<double> av{1, 1}, h{0, 0}, hn{1, 1};  // FP64 geometry: conservative accumulation
    if (lay0) {
        av = ld2_nt(m.av_d + c);
        h = ld2_nt(m.h_d + c);
        hn = ld_same(m.hn_d, m.h_d, c, h);
    }

    constexpr int TG = NT < 2 ? NT : 2;  // tracers whose gathers share a round trip
#pragma unroll
    for (int t0 = 0; t0 < NT; t0 += TG) {
        V2<R> rp[TG], rm[TG];
        bool us0[TG], us1[TG];  // learned: this node level failed its bound check
        R fvl0[TG], fvl1[TG], fvl2[TG];
#pragma unroll
        for (int u = 0; u < TG; ++u) {
            const size_t tn = (t0 + u) * tsn;
            V2<R> fv{0, 0};
            rp[u] = V2<R>{0, 0};
            rm[u] = V2<R>{0, 0};
            if (if0) fv = ld2_nt(adf_v + tn + c);
            if (!if1) fv.y = 0;
            if (lay0) ld_rpm(Rp, Rm, tn + c, rp[u], rm[u]);
            us0[u] = LEARN && has_flag(rp[u].x);
            us1[u] = LEARN && has_flag(rp[u].y);
            if constexpr (LEARN) { rp[u].x = fabs(rp[u].x); rp[u].y = fabs(rp[u].y); }
            if (!lay1) { rp[u].y = 0; rm[u].y = 0; }
            const R rpu0 = from_prev(rp[u].y), rmu0 = from_prev(rm[u].y);
            fvl0[u] = limit_vertical(k0 + 1, lay0, fv.x, rp[u].x, rm[u].x, rpu0, rmu0);
            fvl1[u] = limit_vertical(k0 + 2, lay1, fv.y, rp[u].y, rm[u].y, rp[u].x, rm[u].x);
            fvl2[u] = below(from_next(fvl0[u]), k0, L);
        }
        if (!lay0) continue;

        // edge slots outer and the group's tracers inner: one memory round trip
        // per incident edge carries its flux and the other node's R+/R- of TG tracers
        double sh0[TG], sh1[TG];
#pragma unroll
        for (int u = 0; u < TG; ++u) { sh0[u] = 0.0; sh1[u] = 0.0; }
#pragma unroll
        for (int j = 0; j < D; ++j) {
            const int nlm = info_nlmax(info[j]);
            if (oe[j] < 0 || k0 >= nlm) continue;
            const bool on1 = lay1 && k0 + 1 < nlm;
            const bool pos = info_pos(info[j]);  // this node is edges(1)
            V2<R> f[TG], rpo[TG], rmo[TG], al[TG];
            if constexpr (TG == 2) ld_adf2(adf_h, t0, tse, (size_t)oe[j] + k0, f[0], f[1]);
            else f[0] = ld2(adf_h + adf_at<R>(t0, tse, (size_t)oe[j] + k0));
            if constexpr (LEARN) {
                if constexpr (TG == 2) ld_adf2(alpha, t0, tse, (size_t)oe[j] + k0, al[0], al[1]);
                else al[0] = ld2(alpha + adf_at<R>(t0, tse, (size_t)oe[j] + k0));
            }
#pragma unroll
            for (int u = 0; u < TG; ++u) ld_rpm(Rp, Rm, (t0 + u) * tsn + (size_t)oo[j] + k0, rpo[u], rmo[u]);
#pragma unroll
            for (int u = 0; u < TG; ++u) {
                bool uo0 = false, uo1 = false;  // learned: the other node level failed its bound check
                if constexpr (LEARN) {
                    uo0 = has_flag(rpo[u].x);
                    uo1 = has_flag(rpo[u].y);
                    rpo[u].x = fabs(rpo[u].x);
                    rpo[u].y = fabs(rpo[u].y);
                }
                // both cells evaluated unconditionally, the second one masked
                // (keeps every pair load one 16-byte access). Oriented as the flux
                // entering this node, fs = +-f, Zalesak's coefficient (b3) is
                // min(1, R+ here, R- there) for fs >= 0 and min(1, R- here, R+ there)
                // otherwise -- edge_alpha's choice for every f != 0, and for f == 0 the
                // product with the flux is zero either way
                const R fx = pos ? f[u].x : -f[u].x, fy = pos ? f[u].y : -f[u].y;
                const bool in0 = fx >= (R)0, in1 = fy >= (R)0;
                R a0 = fmn(fmn((R)1, in0 ? rp[u].x : rm[u].x), in0 ? rmo[u].x : rpo[u].x);
                R a1 = fmn(fmn((R)1, in1 ? rp[u].y : rm[u].y), in1 ? rmo[u].y : rpo[u].y);
                if constexpr (LEARN) {
                    // the network's coefficient, Zalesak's where either node failed its bound check
                    a0 = (us0[u] || uo0) ? fmn(al[u].x, a0) : al[u].x;
                    a1 = (us1[u] || uo1) ? fmn(al[u].y, a1) : al[u].y;
                }
                const double fl0 = (double)(a0 * fx);
                const double fl1 = (double)(a1 * fy);
                sh0[u] = sh0[u] + fl0;
                const double s1 = sh1[u] + fl1;
                sh1[u] = on1 ? s1 : sh1[u];
            }
        }
#pragma unroll
        for (int u = 0; u < TG; ++u) {
            const size_t tn = (t0 + u) * tsn;
            // the edge sum times dt/areasvol once (the reference scales each edge's
            // flux before summing; the difference is rounding)
            const double s0 = sh0[u] * dt / (double)av.x;
            const double s1 = sh1[u] * dt / (double)av.y;
            const V2<double> T = ld2_nt(ttf + tn + c), lo = ld2_nt(LO + tn + c);
            double dv0 = 0.0 - T.x * (double)h.x + lo.x * (double)hn.x;
            dv0 = dv0 + ((double)fvl0[u] - (double)fvl1[u]) * dt / (double)av.x;
            double dv1 = 0.0 - T.y * (double)h.y + lo.y * (double)hn.y;
            dv1 = dv1 + ((double)fvl1[u] - (double)fvl2[u]) * dt / (double)av.y;
            node_out<OUT>(ttf, acc_h, acc_v, tn + c, lay1, T, s0, dv0, s1, dv1, hn);
        }
    }
}

// lim = 'NONE': HO fluxes straight into the divergence (driver :291)
template <class R, int VER, int NT, int D, int OUT>
__global__ __launch_bounds__(kMaxBlock) void k_node_nolim(DMesh<R> m, const R *__restrict__ WEv,
                                                       double *__restrict__ ttf,
                                                       const double *__restrict__ ttfAB, size_t tsn,
                                                       const R *__restrict__ flux_h, size_t tse,
                                                       double dt, R num_ord,
                                                       double *__restrict__ acc_h, double *__restrict__ acc_v)
{
    const int ch = wave_chunk(true);  // from the end of each XCD's range (wave_chunk)
    if (ch >= m.nch_n) return;
    const Pair p = lane_pair(m.slot_n, ch);
    const int nlev = p.live ? m.nlev[p.col] : 0;
    const int L = nlev - 1, k0 = p.k0;
    const bool lay0 = k0 < L, lay1 = k0 + 1 < L;
    const size_t c = (size_t)p.c;
    int oe[D];
    unsigned info[D];
    load_node_edges<D>(m, p.col, lay0, oe, info);
    V2<double> av{1, 1}, hn{1, 1};  // FP64 geometry: conservative accumulation
    if (lay0) {
        av = ld2(m.av_d + c);
        hn = ld2(m.hn_d + c);
    }
    VCol2<R> hc = node_vcol<R, VER>(m, p, nlev, WEv);
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const size_t tn = t * tsn, te = t * tse;
        V2<double> B{0.0, 0.0};
        if (p.live) B = ld2(ttfAB + tn + c);
        hc.A = tracer_window<R>(B, k0, nlev);
        R fv0, fv1;
        ver_flux<R, VER>(hc, num_ord, (R)dt, (R)0, (R)0, fv0, fv1);
        const R fv2 = below(from_next(fv0), k0, L);
        if (!lay0) continue;
        double sh0 = 0.0, sh1 = 0.0;
#pragma unroll
        for (int j = 0; j < D; ++j) {
            const int nlm = info_nlmax(info[j]);
            if (oe[j] < 0 || k0 >= nlm) continue;
            const bool on1 = lay1 && k0 + 1 < nlm;
            const bool pos = info_pos(info[j]);
            const V2<R> f = ld2(flux_h + adf_at<R>(t, tse, (size_t)oe[j] + k0));
            const double f0 = (double)f.x, f1 = (double)f.y;
            sh0 = pos ? sh0 + f0 : sh0 - f0;
            const double s1 = pos ? sh1 + f1 : sh1 - f1;
            sh1 = on1 ? s1 : sh1;
        }
        sh0 = sh0 * dt / (double)av.x;
        sh1 = sh1 * dt / (double)av.y;
        const V2<double> T = ld2(ttf + tn + c);
        const double dv0 = 0.0 + ((double)fv0 - (double)fv1) * dt / (double)av.x;
        const double dv1 = 0.0 + ((double)fv1 - (double)fv2) * dt / (double)av.y;
        node_out<OUT>(ttf, acc_h, acc_v, tn + c, lay1, T, sh0, dv0, sh1, dv1, hn);
    }
}

// ===========================================================================
// Learned limiter. Features of a horizontal edge level (edges(1) = node 1):
//   x0, x1  LO and HO edge flux as a tracer change of node 1 per step,
//           flux*dt/(areasvol*hnode_new), over the local range s
//   x2, x3  tracer at node 1 / node 2 above the local minimum, over s
//   x4..x7  Zalesak headroom Q+ and -Q- of node 1 and node 2, over s
//   x8, x9  the antidiffusive sums P (Zalesak b1) on the receiving side of
//           node 1 and the giving side of node 2 for the flux's sign,
//           P*dt/areasvol over s (the reference's R = Q / (P*dt/areasvol))
// with the local range s = max(bounds) - min(bounds) of the two nodes; the
// target is Zalesak's coefficient of the edge level (b3). Values are clipped
// to [-8, 8] before the network sees them.
// ===========================================================================
template <class R>
struct LearnSrc {
    const double *T, *LO;  // node tracer and LO solution (this tracer)
    const R *adf;          // edge HO-LO flux (this tracer: cell i at adf[kAdfStride * i])
    // node headroom Q+/Q-, R+/R- and P+/P- (this tracer), each pair of
    // arrays stored as ld_rpm reads it
    const R *Qp, *Qm, *Rp, *Rm, *Pp, *Pm;
};

// the tracer-independent part of an edge level's features: the cell volumes
// of the end nodes as the scalings use them, and the edge transport
template <class R>
struct EdgeGeo {
    float vh1[2], v1[2], v2[2];  // (float)(areasvol * hnode_new) of node 1, (float)areasvol of nodes 1 and 2
    R vf[2];
};
template <class R>
__device__ __forceinline__ EdgeGeo<R> edge_geo(const DMesh<R> &m, const EdgeRec &er, int k0, size_t c,
                                               const R *vfl)
{
    const V2<double> av1 = ld2(m.av_d + er.o1 + k0), hn1 = ld2(m.hn_d + er.o1 + k0);
    const V2<double> av2 = ld2(m.av_d + er.o2 + k0);
    const V2<R> v = ld2(vfl + c);
    return EdgeGeo<R>{{(float)(av1.x * hn1.x), (float)(av1.y * hn1.y)},
                      {(float)av1.x, (float)av1.y},
                      {(float)av2.x, (float)av2.y},
                      {v.x, v.y}};
}

__device__ __forceinline__ float clip8(float v) { return fminf(8.0f, fmaxf(-8.0f, v)); }

// Zalesak's coefficient (b3) of the lane's two edge levels with HO-LO flux
// f: the training target
template <class R>
__device__ __forceinline__ V2<R> edge_zalesak(const EdgeRec &er, int k0, const LearnSrc<R> &S, V2<R> f)
{
    V2<R> rp1, rm1, rp2, rm2;
    ld_rpm(S.Rp, S.Rm, (size_t)er.o1 + k0, rp1, rm1);
    ld_rpm(S.Rp, S.Rm, (size_t)er.o2 + k0, rp2, rm2);
    return V2<R>{edge_alpha(true, f.x, rp1.x, rm1.x, rp2.x, rm2.x), edge_alpha(true, f.y, rp1.y, rm1.y, rp2.y, rm2.y)};
}

// features of the lane's two edge levels (a: their HO-LO flux).
// Differences of tracer values are formed in double (they may be small
// against the values); the scalings by the local range and the cell volume
// are float reciprocals: the features are clipped to [-8, 8] and rounded to
// bf16 afterwards.
template <class R>
__device__ __forceinline__ void edge_features(const EdgeRec &er, int k0, double dt, const EdgeGeo<R> &G,
                                              const LearnSrc<R> &S, V2<R> a, float (&x)[2][kMlpIn])
{
    const V2<double> T1 = ld2(S.T + er.o1 + k0), T2 = ld2(S.T + er.o2 + k0);
    const V2<double> L1 = ld2(S.LO + er.o1 + k0), L2 = ld2(S.LO + er.o2 + k0);
    V2<R> qp1, qm1, qp2, qm2, pp1, pm1, pp2, pm2;
    ld_rpm(S.Qp, S.Qm, (size_t)er.o1 + k0, qp1, qm1);
    ld_rpm(S.Qp, S.Qm, (size_t)er.o2 + k0, qp2, qm2);
    ld_rpm(S.Pp, S.Pm, (size_t)er.o1 + k0, pp1, pm1);
    ld_rpm(S.Pp, S.Pm, (size_t)er.o2 + k0, pp2, pm2);
    const double t1[2] = {T1.x, T1.y}, t2[2] = {T2.x, T2.y}, l1[2] = {L1.x, L1.y}, l2[2] = {L2.x, L2.y};
    const R Qp1[2] = {qp1.x, qp1.y}, Qm1[2] = {qm1.x, qm1.y}, Qp2[2] = {qp2.x, qp2.y}, Qm2[2] = {qm2.x, qm2.y};
    const R Pp1[2] = {pp1.x, pp1.y}, Pm1[2] = {pm1.x, pm1.y}, Pp2[2] = {pp2.x, pp2.y}, Pm2[2] = {pm2.x, pm2.y};
    const R A[2] = {a.x, a.y};
    const float fdt = (float)dt;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const double mx = fmax(l1[q] + (double)Qp1[q], l2[q] + (double)Qp2[q]);
        const double mn = fmin(l1[q] + (double)Qm1[q], l2[q] + (double)Qm2[q]);
        const float sc = mx - mn > 1e-30 ? (float)(mx - mn) : 1.0f;
        const float isc = __builtin_amdgcn_rcpf(sc);
        const double lo = lo_flux((double)G.vf[q], t1[q], t2[q]);
        const float toT = fdt * __builtin_amdgcn_rcpf(G.vh1[q] * sc);
        x[q][0] = clip8((float)lo * toT);
        x[q][1] = clip8((float)(lo + (double)A[q]) * toT);
        x[q][2] = clip8((float)(t1[q] - mn) * isc);
        x[q][3] = clip8((float)(t2[q] - mn) * isc);
        x[q][4] = clip8((float)Qp1[q] * isc);
        x[q][5] = clip8(-(float)Qm1[q] * isc);
        x[q][6] = clip8((float)Qp2[q] * isc);
        x[q][7] = clip8(-(float)Qm2[q] * isc);
        const bool in1 = A[q] >= (R)0;  // flux into node 1
        x[q][8] = clip8((float)(in1 ? Pp1[q] : -Pm1[q]) * (fdt * __builtin_amdgcn_rcpf(G.v1[q] * sc)));
        x[q][9] = clip8((float)(in1 ? -Pm2[q] : Pp2[q]) * (fdt * __builtin_amdgcn_rcpf(G.v2[q] * sc)));
    }
}

template <class R, int NT>
struct LearnArgs {
    const double *ttf, *LO;
    const R *adf, *vfl, *Qp, *Qm, *Rp, *Rm, *Pp, *Pm;
    size_t tsn, tse;
};

template <class R, int NT>
__device__ __forceinline__ LearnSrc<R> learn_src(const LearnArgs<R, NT> &a, int t)
{
    const size_t tq = t * a.tsn * kRpmPerCell<R>;
    return LearnSrc<R>{a.ttf + t * a.tsn, a.LO + t * a.tsn, a.adf + adf_slot<R>(t, a.tse),
                       a.Qp + tq, a.Qm + tq, a.Rp + tq, a.Rm + tq, a.Pp + tq, a.Pm + tq};
}

// training rows: features (bf16, 16 per row) and Zalesak target of every
// edge-level cell of every tracer; padding cells get target -1
template <class R, int NT>
__global__ __launch_bounds__(kMaxBlock) void k_learn_dump(DMesh<R> m, LearnArgs<R, NT> a, double dt,
                                                       unsigned short *__restrict__ X, float *__restrict__ Y,
                                                       long long row0)
{
    const int ch = wave_chunk();
    if (ch >= m.nch_e) return;
    const Pair p = lane_pair(m.slot_e, ch);
    EdgeRec er{};
    EdgeGeo<R> geo{};
    int nlmax = 0;
    if (p.live) {
        er = m.erec[p.col];
        nlmax = (int)((unsigned)er.nl >> 24);
        geo = edge_geo(m, er, p.k0, (size_t)p.c, a.vfl);
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        float x[2][kMlpIn] = {}, y[2] = {-1.0f, -1.0f};
        const LearnSrc<R> src = learn_src(a, t);
        if (p.live) {
            const V2<R> f = ld2(src.adf + kAdfStride<R> * p.c);
            edge_features(er, p.k0, dt, geo, src, f, x);
            const V2<R> az = edge_zalesak(er, p.k0, src, f);
            y[0] = (float)az.x;
            y[1] = (float)az.y;
        }
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const bool ok = p.live && p.k0 + q < nlmax;
            const long long row = row0 + (long long)t * m.CE + p.c + q;
            unsigned short b[kMlpInPad];
#pragma unroll
            for (int f = 0; f < kMlpInPad; ++f) b[f] = (ok && f < kMlpIn) ? f2bf(x[q][f]) : (unsigned short)0;
            uint4 *dst = reinterpret_cast<uint4 *>(X + row * kMlpInPad);
            unsigned u[8];
#pragma unroll
            for (int h = 0; h < 8; ++h) u[h] = (unsigned)b[2 * h] | ((unsigned)b[2 * h + 1] << 16);
            dst[0] = make_uint4(u[0], u[1], u[2], u[3]);
            dst[1] = make_uint4(u[4], u[5], u[6], u[7]);
            Y[row] = ok ? y[q] : -1.0f;
        }
    }
}

// the learned coefficient of every edge level: a persistent grid; each wave
// forms the feature rows (bf16) of the edge levels of its chunk that carry a
// nonzero HO-LO flux, packed to the front of its own LDS slice, runs them
// through the matrix cores 16 rows at a time and reads its outputs back; the
// other levels get 1. The lanes of one wave exchange data only with each
// other, so a wave-scope fence replaces block barriers and waves leave the
// loop independently.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// number of set bits of b below this lane
__device__ __forceinline__ int lanes_below(unsigned long long b)
{
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(b >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)b, 0u));
}

template <class R, int NT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_learn_alpha(DMesh<R> m, MlpDev w, LearnArgs<R, NT> a, double dt,
                                                        R *__restrict__ alpha, int iters)
{
    // one tracer staged at a time (18 KB of LDS per block); the weights held
    // in registers (80 VGPRs) leave room for three waves per SIMD
    __shared__ uint4 sX[kWavesPerBlock][kChunk][2];  // 16 bf16 features per row
    __shared__ float sY[kWavesPerBlock][kChunk];
    const int wv = (int)(threadIdx.x >> 6), l = lane_id(), i = l & 15, g = l >> 4;
    MlpInfer r;
    mlp_infer_load(w, r);
    for (int it = 0; it < iters; ++it) {
        const int ch = (it * (int)gridDim.x + (int)blockIdx.x) * kWavesPerBlock + wv;
        if (ch >= m.nch_e) break;  // wave-uniform
        const Pair p = lane_pair(m.slot_e, ch);
        EdgeRec er{};
        EdgeGeo<R> geo{};
        int nlmax = 0;
        if (p.live) {
            er = m.erec[p.col];
            nlmax = (int)((unsigned)er.nl >> 24);
            geo = edge_geo(m, er, p.k0, (size_t)p.c, a.vfl);
        }
#pragma unroll 1
        for (int t = 0; t < NT; ++t) {
            const LearnSrc<R> src = learn_src(a, t);
            V2<R> f{0, 0};
            if (p.live) f = ld2(src.adf + kAdfStride<R> * p.c);
            // the rows the network is asked about, packed in LDS lane by lane, first cell first:
            // edge levels with a nonzero HO-LO flux (where it is zero, every coefficient gives the
            // same limited flux, zero, and the level gets 1)
            const bool r0 = p.live && p.k0 < nlmax && f.x != (R)0;
            const bool r1 = p.live && p.k0 + 1 < nlmax && f.y != (R)0;
            const unsigned long long b0 = __ballot(r0), b1 = __ballot(r1);
            const int pos0 = lanes_below(b0) + lanes_below(b1), pos1 = pos0 + (r0 ? 1 : 0);
            const int ntile = (__popcll(b0) + __popcll(b1) + 15) >> 4;  // wave-uniform
            if (r0 || r1) {
                float x[2][kMlpIn];
                edge_features(er, p.k0, dt, geo, src, f, x);
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    unsigned u[8];
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        // v_cvt_pk_bf16_f32: two features per instruction, nearest even
                        const mlp_v2f v2 = {2 * e < kMlpIn ? x[q][2 * e] : 0.0f,
                                            2 * e + 1 < kMlpIn ? x[q][2 * e + 1] : 0.0f};
                        const unsigned b = __builtin_bit_cast(unsigned, __builtin_convertvector(v2, mlp_v2b));
                        u[e] = 2 * e < kMlpIn ? b : 0u;
                    }
                    if (q == 0 ? r0 : r1) {
                        const int row = q == 0 ? pos0 : pos1;
                        sX[wv][row][0] = make_uint4(u[0], u[1], u[2], u[3]);
                        sX[wv][row][1] = make_uint4(u[4], u[5], u[6], u[7]);
                    }
                }
            }
            wave_lds_sync();
            // rows past the last one in the last tile hold stale features;
            // their outputs are never read
            int tile = 0;
#pragma unroll 1
            for (; tile + 1 < ntile; tile += 2) {  // two tiles at a time (mlp_infer_tiles)
                const mlp_v4s xb[2] = {
                    ld_bf4(reinterpret_cast<const unsigned short *>(&sX[wv][16 * tile + i][0]) + 4 * g),
                    ld_bf4(reinterpret_cast<const unsigned short *>(&sX[wv][16 * tile + 16 + i][0]) + 4 * g)};
                float z[2];
                mlp_infer_tiles<2>(r, xb, z);
                if (g == 0) {
                    sY[wv][16 * tile + i] = mlp_sigmoid_fast(z[0]);
                    sY[wv][16 * tile + 16 + i] = mlp_sigmoid_fast(z[1]);
                }
            }
            if (tile < ntile) {
                const mlp_v4s xb = ld_bf4(reinterpret_cast<const unsigned short *>(&sX[wv][16 * tile + i][0]) + 4 * g);
                const float z = mlp_infer(r, xb);
                if (g == 0) sY[wv][16 * tile + i] = mlp_sigmoid_fast(z);
            }
            wave_lds_sync();
            const float y0 = r0 ? sY[wv][pos0] : 1.0f, y1 = r1 ? sY[wv][pos1] : 1.0f;
            if (p.live) st2_mixed(alpha + adf_at<R>(t, a.tse, p.c), (R)y0, p.k0 + 1 < nlmax ? (R)y1 : (R)0);
            wave_lds_sync();
        }
    }
}

// per node level: would the network's horizontal coefficients (vertical
// ones as Zalesak's) keep the node inside its bounds? Checked as Zalesak's
// own sufficient condition, positive and negative contributions apart:
// sum(max(0, c)) * dt / areasvol <= Q+ and sum(min(0, c)) * dt / areasvol >= Q-.
// The flag goes to the tracer's own record (unsafe) and onto the sign bit of
// the level's R+ (with_flag), where the update reads it with the R+/R- it
// gathers anyway.
template <class R, int NT, int D>
__global__ __launch_bounds__(kMaxBlock) void k_learn_check(DMesh<R> m, const R *__restrict__ adf_h,
                                                        const R *__restrict__ alpha, size_t tse,
                                                        const R *__restrict__ adf_v, R *__restrict__ Rp,
                                                        R *__restrict__ Rm, const R *__restrict__ Qp,
                                                        const R *__restrict__ Qm, size_t tsn, double dt,
                                                        R *__restrict__ unsafe)
{
    const int ch = wave_chunk();
    if (ch >= m.nch_n) return;
    const Pair p = lane_pair(m.slot_n, ch);
    const int nlev = p.live ? m.nlev[p.col] : 0;
    const int L = nlev - 1, k0 = p.k0;
    const bool lay0 = k0 < L, lay1 = k0 + 1 < L;
    const bool if0 = p.live && k0 < nlev, if1 = p.live && k0 + 1 < nlev;
    const size_t c = (size_t)p.c;
    int oe[D];
    unsigned info[D];
    load_node_edges<D>(m, p.col, lay0, oe, info);
    V2<R> av{1, 1};
    if (lay0) av = ld2(m.areasvol + c);
    constexpr int TG = NT < 2 ? NT : 2;  // tracers whose gathers share a round trip
#pragma unroll
    for (int t0 = 0; t0 < NT; t0 += TG) {
        V2<R> rp[TG], rm[TG];
        double sp0[TG], sn0[TG], sp1[TG], sn1[TG];
#pragma unroll
        for (int u = 0; u < TG; ++u) {
            const size_t tn = (t0 + u) * tsn;
            V2<R> fv{0, 0};
            rp[u] = V2<R>{0, 0};
            rm[u] = V2<R>{0, 0};
            if (if0) fv = ld2(adf_v + tn + c);
            if (!if1) fv.y = 0;
            if (lay0) ld_rpm(Rp, Rm, tn + c, rp[u], rm[u]);
            const R rpy = lay1 ? rp[u].y : (R)0, rmy = lay1 ? rm[u].y : (R)0;
            const R rpu0 = from_prev(rpy), rmu0 = from_prev(rmy);
            const R fvl0 = limit_vertical(k0 + 1, lay0, fv.x, rp[u].x, rm[u].x, rpu0, rmu0);
            const R fvl1 = limit_vertical(k0 + 2, lay1, fv.y, rpy, rmy, rp[u].x, rm[u].x);
            const R fvl2 = below(from_next(fvl0), k0, L);
            sp0[u] = fmax(0.0, (double)fvl0) + fmax(0.0, -(double)fvl1);
            sn0[u] = fmin(0.0, (double)fvl0) + fmin(0.0, -(double)fvl1);
            sp1[u] = fmax(0.0, (double)fvl1) + fmax(0.0, -(double)fvl2);
            sn1[u] = fmin(0.0, (double)fvl1) + fmin(0.0, -(double)fvl2);
        }
        if (!lay0) continue;
#pragma unroll
        for (int j = 0; j < D; ++j) {
            const int nlm = info_nlmax(info[j]);
            if (oe[j] < 0 || k0 >= nlm) continue;
            const bool on1 = lay1 && k0 + 1 < nlm;
            const double sg = info_pos(info[j]) ? 1.0 : -1.0;
            const size_t i = (size_t)oe[j] + k0;
            V2<R> f[TG], al[TG];
            if constexpr (TG == 2) {
                ld_adf2(adf_h, t0, tse, i, f[0], f[1]);
                ld_adf2(alpha, t0, tse, i, al[0], al[1]);
            } else {
                f[0] = ld2(adf_h + adf_at<R>(t0, tse, i));
                al[0] = ld2(alpha + adf_at<R>(t0, tse, i));
            }
#pragma unroll
            for (int u = 0; u < TG; ++u) {
                const double c0 = (double)(al[u].x * f[u].x) * sg;
                const double c1 = (double)(al[u].y * f[u].y) * sg;
                sp0[u] += fmax(0.0, c0);
                sn0[u] += fmin(0.0, c0);
                sp1[u] = on1 ? sp1[u] + fmax(0.0, c1) : sp1[u];
                sn1[u] = on1 ? sn1[u] + fmin(0.0, c1) : sn1[u];
            }
        }
        const double s0 = dt / (double)av.x, s1 = dt / (double)av.y;
#pragma unroll
        for (int u = 0; u < TG; ++u) {
            const size_t tn = (t0 + u) * tsn;
            V2<R> qp, qm;
            ld_rpm(Qp, Qm, tn + c, qp, qm);
            const bool bad0 = sp0[u] * s0 > (double)qp.x + 1e-7 * fabs((double)qp.x) ||
                              sn0[u] * s0 < (double)qm.x - 1e-7 * fabs((double)qm.x);
            const bool bad1 = lay1 && (sp1[u] * s1 > (double)qp.y + 1e-7 * fabs((double)qp.y) ||
                                       sn1[u] * s1 < (double)qm.y - 1e-7 * fabs((double)qm.y));
            st2(unsafe + tn + c, bad0 ? (R)1 : (R)0, bad1 ? (R)1 : (R)0);
            // R+ of the pair only (FP32: the first half of its 16-byte record), only where a flag is set
            if (bad0 || bad1)
                st2_mixed(Rp + kRpmPerCell<R> * (tn + c), with_flag(rp[u].x, bad0), with_flag(rp[u].y, bad1));
        }
    }
}

// wet layers of tracer 0 flagged unsafe (count) and wet layers (count)
template <class R>
__global__ __launch_bounds__(256) void k_count_flags(DMesh<R> m, const R *__restrict__ unsafe,
                                                     unsigned long long *__restrict__ out)
{
    unsigned long long nf = 0, nc = 0;
    const long long npair = m.CN / 2;
    for (long long s = (long long)blockIdx.x * 256 + threadIdx.x; s < npair; s += (long long)gridDim.x * 256) {
        const int w = m.slot_n[s];
        if (w < 0) continue;
        const int n = w & ((1 << kColBits) - 1), k0 = 2 * (int)((unsigned)w >> kColBits);
        const int L = m.nlev[n] - 1;
        for (int q = 0; q < 2 && k0 + q < L; ++q) {
            nc += 1;
            nf += unsafe[2 * s + q] != (R)0 ? 1 : 0;
        }
    }
    atomicAdd(out, nf);
    atomicAdd(out + 1, nc);
}

// ===========================================================================
// Diagnostics: deterministic two-pass block reduction in FP64
// ===========================================================================
struct DiagAcc { double vmin, vmax, sum, amin, amax, mass, nan; };

__device__ inline void diag_merge(DiagAcc &a, const DiagAcc &b)
{
    a.vmin = fmin(a.vmin, b.vmin); a.vmax = fmax(a.vmax, b.vmax); a.sum += b.sum;
    a.amin = fmin(a.amin, b.amin); a.amax = fmax(a.amax, b.amax); a.mass += b.mass; a.nan += b.nan;
}

// A tracer between the caller's dense (nl-1, nod2D) layout (levels fastest)
// and the chunk-packed node layers: one thread per cell pair of the layout.
// Packing writes every cell of the layout (0 outside the wet layers);
// unpacking writes only the wet layers, so the dense copy keeps the levels
// below the bottom as last set.
__global__ __launch_bounds__(256) void k_pack_layers(const int *__restrict__ slot, const int *__restrict__ nlev,
                                                     const int *__restrict__ ncol, long long npair,
                                                     const double *__restrict__ dense, int L, double *__restrict__ out)
{
    const long long s = (long long)blockIdx.x * 256 + threadIdx.x;
    if (s >= npair) return;
    const int w = slot[s];
    double a = 0.0, b = 0.0;
    if (w >= 0) {
        const int q = w & ((1 << kColBits) - 1), k0 = 2 * (int)((unsigned)w >> kColBits);
        const int Lc = nlev[q] - 1;
        const double *src = dense + (size_t)ncol[q] * L;
        if (k0 < Lc) a = src[k0];
        if (k0 + 1 < Lc) b = src[k0 + 1];
    }
    out[2 * s] = a;
    out[2 * s + 1] = b;
}

__global__ __launch_bounds__(256) void k_unpack_layers(const int *__restrict__ slot, const int *__restrict__ nlev,
                                                       const int *__restrict__ ncol, long long npair,
                                                       const double *__restrict__ in, int L, double *__restrict__ dense)
{
    const long long s = (long long)blockIdx.x * 256 + threadIdx.x;
    if (s >= npair) return;
    const int w = slot[s];
    if (w < 0) return;
    const int q = w & ((1 << kColBits) - 1), k0 = 2 * (int)((unsigned)w >> kColBits);
    const int Lc = nlev[q] - 1;
    double *dst = dense + (size_t)ncol[q] * L;
    if (k0 < Lc) dst[k0] = in[2 * s];
    if (k0 + 1 < Lc) dst[k0 + 1] = in[2 * s + 1];
}

// dense += the wet layers of a chunk-packed node array (the caller's tendency
// arrays accumulate what a step returns, as do_oce_adv_tra adds into them)
__global__ __launch_bounds__(256) void k_unpack_add_layers(const int *__restrict__ slot, const int *__restrict__ nlev,
                                                           const int *__restrict__ ncol, long long npair,
                                                           const double *__restrict__ in, int L,
                                                           double *__restrict__ dense)
{
    const long long s = (long long)blockIdx.x * 256 + threadIdx.x;
    if (s >= npair) return;
    const int w = slot[s];
    if (w < 0) return;
    const int q = w & ((1 << kColBits) - 1), k0 = 2 * (int)((unsigned)w >> kColBits);
    const int Lc = nlev[q] - 1;
    double *dst = dense + (size_t)ncol[q] * L;
    if (k0 < Lc) dst[k0] = dst[k0] + in[2 * s];
    if (k0 + 1 < Lc) dst[k0 + 1] = dst[k0 + 1] + in[2 * s + 1];
}

// reduction over the wet layers (node layout): min, max, sum, mass, NaN count
// plus, for min / max / sum / NaN count over all values, the levels below each
// column's bottom, read from the dense (nl-1, nod2D) copy (never stepped)
template <class R>
__global__ __launch_bounds__(256) void k_diag_partial(DMesh<R> m, const double *__restrict__ v,
                                                      const double *__restrict__ dense, int L,
                                                      DiagAcc *__restrict__ part)
{
    __shared__ DiagAcc sh[256];
    DiagAcc a{INFINITY, -INFINITY, 0.0, INFINITY, -INFINITY, 0.0, 0.0};
    const long long npair = m.CN / 2;
    const long long stride = (long long)gridDim.x * 256;
    for (long long s = (long long)blockIdx.x * 256 + threadIdx.x; s < npair; s += stride) {
        const int w = m.slot_n[s];
        if (w < 0) continue;
        const int n = w & ((1 << kColBits) - 1), k0 = 2 * (int)((unsigned)w >> kColBits);
        const int Lc = m.nlev[n] - 1;
        if (k0 >= Lc) continue;
        const V2<double> x2 = ld2(v + 2 * s), h2 = ld2(m.hn_d + 2 * s), a2 = ld2(m.av_d + 2 * s);
        const double xs[2] = {x2.x, x2.y}, hs[2] = {h2.x, h2.y}, as[2] = {a2.x, a2.y};
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            if (k0 + q >= Lc) break;
            const double x = xs[q];
            if (is_nan_bits(x)) { a.nan += 1.0; continue; }
            a.vmin = fmin(a.vmin, x); a.vmax = fmax(a.vmax, x); a.sum += x;
            a.mass += x * hs[q] * as[q];
        }
    }
    a.amin = a.vmin;
    a.amax = a.vmax;
    // the dense copy cell by cell (coalesced); the column from a reciprocal,
    // corrected by one either way
    const long long cells = (long long)m.N * L;
    const double invL = 1.0 / (double)L;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < cells; i += stride) {
        int n = (int)((double)i * invL);
        long long k = i - (long long)n * L;
        if (k < 0) { --n; k += L; }
        else if (k >= L) { ++n; k -= L; }
        if (k < m.nlev_c[n] - 1) continue;
        const double x = dense[i];
        if (is_nan_bits(x)) { a.nan += 1.0; continue; }
        a.vmin = fmin(a.vmin, x); a.vmax = fmax(a.vmax, x); a.sum += x;
    }
    sh[threadIdx.x] = a;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) diag_merge(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(256) void k_diag_final(const DiagAcc *__restrict__ part, int np, DiagAcc *__restrict__ out)
{
    __shared__ DiagAcc sh[256];
    DiagAcc a{INFINITY, -INFINITY, 0.0, INFINITY, -INFINITY, 0.0, 0.0};
    for (int i = threadIdx.x; i < np; i += 256) diag_merge(a, part[i]);
    sh[threadIdx.x] = a;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) diag_merge(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// difference of two engines' node tracer over the wet layers (same layout):
// max |a-b|, sum (a-b)^2, max |b|, sum b^2
struct ErrAcc { double dinf, d2, rinf, r2; };

__global__ __launch_bounds__(256) void k_err_partial(const int *__restrict__ slot, const int *__restrict__ nlev,
                                                     long long npair, const double *__restrict__ a,
                                                     const double *__restrict__ b, ErrAcc *__restrict__ part)
{
    __shared__ ErrAcc sh[256];
    ErrAcc e{0.0, 0.0, 0.0, 0.0};
    for (long long s = (long long)blockIdx.x * 256 + threadIdx.x; s < npair; s += (long long)gridDim.x * 256) {
        const int w = slot[s];
        if (w < 0) continue;
        const int n = w & ((1 << kColBits) - 1), k0 = 2 * (int)((unsigned)w >> kColBits);
        const int L = nlev[n] - 1;
        for (int q = 0; q < 2 && k0 + q < L; ++q) {
            const double d = a[2 * s + q] - b[2 * s + q], r = b[2 * s + q];
            e.dinf = fmax(e.dinf, fabs(d));
            e.d2 += d * d;
            e.rinf = fmax(e.rinf, fabs(r));
            e.r2 += r * r;
        }
    }
    sh[threadIdx.x] = e;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            ErrAcc &x = sh[threadIdx.x];
            const ErrAcc &y = sh[threadIdx.x + st];
            x.dinf = fmax(x.dinf, y.dinf); x.d2 += y.d2; x.rinf = fmax(x.rinf, y.rinf); x.r2 += y.r2;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// ===========================================================================
// Host side
// ===========================================================================
namespace {

template <class T>
T *dalloc(size_t n)
{
    void *p = nullptr;
    HIP_OK(hipMalloc(&p, (n ? n : 1) * sizeof(T)));
    return (T *)p;
}

template <class T>
T *dupload(const std::vector<T> &h)
{
    T *d = dalloc<T>(h.size());
    if (!h.empty()) HIP_OK(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

inline unsigned grid_for(int nchunks, int waves = kWavesPerBlock)
{
    unsigned nb = (unsigned)((nchunks + waves - 1) / waves);
    nb = nb ? nb : 1u;
    return (nb + 7u) & ~7u;  // multiple of 8 for the XCD remap
}

}  // namespace

// chunked layout of a set of columns: column c uses len[c] cells from off[c]
// (padded to an even count); columns are packed whole into kChunk-cell chunks,
// in the caller's order except that a column that no longer fits the current
// chunk may be passed over by up to `window` later ones that do (so chunks
// waste less of their tail while neighbours stay close), and a chunk takes at
// most `maxcols` columns (what a wave stages in LDS). order[q] is the column
// at place q; the slot words and the per-column records use q.
struct Layout {
    std::vector<long long> off;
    std::vector<int> len;
    std::vector<int> order;
    std::vector<int> slot;  // per cell pair: place | pair << kColBits, -1 = padding
    long long total = 0;    // cells, nch * kChunk
    int nch = 0;
    void build(const std::vector<int> &lens, int window = 0, int maxcols = kChunk)
    {
        const size_t n = lens.size();
        len = lens;
        off.assign(n, 0);
        order.clear();
        order.reserve(n);
        if (n >= (size_t)1 << kColBits) throw std::runtime_error("too many columns for the slot word");
        auto plen = [&](size_t c) { return lens[c] + (lens[c] & 1); };
        for (size_t c = 0; c < n; ++c)
            if (plen(c) > kChunk) throw std::runtime_error("column longer than a wave chunk");
        std::vector<char> placed(n, 0);
        long long chunk = 0;
        int pos = 0, ncols = 0;
        auto place = [&](size_t c) {
            off[c] = chunk * kChunk + pos;
            pos += plen(c);
            ++ncols;
            placed[c] = 1;
            order.push_back((int)c);
        };
        size_t head = 0;
        while (head < n) {
            if (placed[head]) { ++head; continue; }
            if (ncols < maxcols && pos + plen(head) <= kChunk) { place(head); continue; }
            // the next column in order does not fit: fill the tail from the window
            for (size_t j = head + 1; j < std::min(n, head + 1 + (size_t)window) && ncols < maxcols; ++j)
                if (!placed[j] && pos + plen(j) <= kChunk) place(j);
            ++chunk;
            pos = 0;
            ncols = 0;
        }
        nch = n == 0 ? 0 : (int)(pos > 0 ? chunk + 1 : chunk);
        total = (long long)nch * kChunk;
        slot.assign((size_t)total / 2, -1);
        for (size_t q = 0; q < n; ++q) {
            const int c = order[q], pairs = (lens[c] + 1) / 2;
            for (int k = 0; k < pairs; ++k) slot[(size_t)(off[c] / 2 + k)] = (int)q | (k << kColBits);
        }
    }
    // unpadded column-major (ld, ncol) with comps interleaved components ->
    // layout; the first len[c]-sub levels of each column are copied
    template <class T>
    std::vector<T> pack(const double *src, int ld, int sub, int comps = 1) const
    {
        std::vector<T> out((size_t)total * comps, (T)0);
        for (size_t c = 0; c < len.size(); ++c) {
            const int cnt = std::min(len[c] - sub, ld);
            for (int k = 0; k < cnt; ++k)
                for (int q = 0; q < comps; ++q)
                    out[((size_t)off[c] + k) * comps + q] = (T)src[((size_t)c * ld + k) * comps + q];
        }
        return out;
    }
};

struct Engine::Impl {
    std::vector<void *> allocs;
    template <class T> T *track(T *p) { allocs.push_back((void *)p); return p; }

    DMesh<double> m64{};
    DMesh<float> m32{};
    Layout ln, lt, le;  // node interfaces/layers (len nlev), element layers, edge levels
    // velocity
    void *vel = nullptr, *W = nullptr, *WE = nullptr;
    // tracers [ntr][CN]
    double *values = nullptr, *valuesAB = nullptr;
    // work, per tracer
    void *vfl = nullptr;                                   // [CE] edge transport (FCT)
    void *adf_h = nullptr;                                 // [ntr][CE]
    void *adf_v = nullptr;                                 // [ntr][CN] (interfaces)
    void *Pp = nullptr, *Pm = nullptr;                     // [ntr][CN]
    double *LO = nullptr;                                  // [ntr][CN]
    void *gnod = nullptr;                                  // [ntr][2*CN]
    DiagAcc *dpart = nullptr, *dout = nullptr;
    ErrAcc *epart = nullptr;
    // learned limiter: per-tracer Q+/Q-, coefficients and fallback flags
    // (allocated when a tracer first selects it), network parameters (FP32
    // master + Adam moments + bf16 operand copies), training rows
    void *Qp = nullptr, *Qm = nullptr, *Sp = nullptr, *Sm = nullptr, *alpha = nullptr, *unsafe = nullptr;
    float *mlp_p = nullptr, *mlp_m1 = nullptr, *mlp_m2 = nullptr, *mlp_g = nullptr, *mlp_loss = nullptr;
    unsigned short *w1t = nullptr, *w2t = nullptr, *w2 = nullptr, *w3 = nullptr;
    float *b1 = nullptr, *b2 = nullptr, *b3 = nullptr;
    unsigned short *dsX = nullptr;
    float *dsY = nullptr;
    long long ds_cap = 0, ds_rows = 0;
    int adam_t = 0;
    unsigned long long *cnt = nullptr;
    // set while collecting training rows: launch_batch dumps after the limiter
    struct Dump { unsigned short *X; float *Y; long long row; long long cap; } *dump = nullptr;
    int ndiag = 1024;
    float gscale = 1.0f, ginv = 1.0f;
    size_t tsn = 0, tse = 0;
    int wslots = 1;  // tracers the work arrays hold
    int fan_fault = -1;  // first node whose elements do not form one fan (-1: none)
    bool grad_fan = true;  // nodal gradients over the fan (when every node has one), else element by element
    bool we_is_w = false;  // w_e equals w (no implicit vertical split): one device copy serves both
    // how a step ends (OUT_UPDATE, OUT_ACCUM with del_ttf per tracer [ntr][CN],
    // OUT_TEND into tend_h / tend_v [CN]); dense_t stages a caller's tendency
    // array in its (nl-1, nod2D) layout
    int out_mode = OUT_UPDATE;
    double *del_ttf = nullptr, *tend_h = nullptr, *tend_v = nullptr, *dense_t = nullptr;
    int tune_waves = kWavesPerBlock;   // waves per block of the chunk kernels
    int tune_nt = 4;                   // most tracers per launch
    // each tracer in the caller's dense (nl-1, nod2D) layout, as last set and
    // then updated in the wet layers by get_values; the levels below the
    // bottom are never stepped, so get_values and the all-values statistics
    // read them here. dense_ab stages set_valuesAB uploads.
    double *dense = nullptr, *dense_ab = nullptr;
    // two pinned host buffers that stage host <-> device transfers in chunks,
    // the host copy of one chunk overlapping the DMA of the other; each holds
    // one tracer array (64 MB at most), so small meshes pin little
    static constexpr size_t kStageMax = (size_t)64 << 20;
    size_t stage_bytes = 0;
    void *stage[2] = {nullptr, nullptr};
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
};

// the per-node records (incident edges, 1-ring fan, elements) at width D, in
// the node layout's order (record q belongs to node ln.order[q])
template <int D, class R>
static void build_node_records(Engine::Impl &I, DMesh<R> &d, const HostMesh &m, const std::vector<int> &enlmax)
{
    const int N = m.nod2D;
    std::vector<NodeEdges<D>> ne(N);
    std::vector<NodeEdgeInfo<D>> ni(N);
    std::vector<NodeRing<D>> rg(N);
    std::vector<NodeElems<D>> nt(N);
    std::vector<R> gw((size_t)8 * D * N, (R)0);
    std::vector<int> ring, fan;
    for (int q = 0; q < N; ++q) {
        const int n = I.ln.order[q];
        const int deg = m.n2e_ptr[n + 1] - m.n2e_ptr[n];
        for (int j = 0; j < D; ++j) {
            ne[q].oe[j] = -1; ne[q].oo[j] = 0; ni[q].info[j] = 0;
            rg[q].oo[j] = -1; rg[q].lr[j] = 0; rg[q].le[j] = 0;
            nt[q].ei[j] = -1; nt[q].oa[j] = 0; nt[q].ob[j] = 0; nt[q].le[j] = 0; nt[q].pos[j] = 0;
        }
        for (int j = 0; j < deg; ++j) {
            const int se = m.n2e[m.n2e_ptr[n] + j];
            const int e = (se > 0 ? se : -se) - 1;
            const int other = se > 0 ? m.edges[2 * e + 1] : m.edges[2 * e];
            ne[q].oe[j] = (int)I.le.off[e];
            ne[q].oo[j] = (int)I.ln.off[other];
            ni[q].info[j] = (unsigned short)(enlmax[e] | (se > 0 ? 128 : 0) | ((m.nlevels_nod2D[other] - 1) << 8));
        }
        for (int j = 0; j < m.nod_in_elem2D_num[n]; ++j) {
            const int el = m.nod_in_elem2D[(size_t)n * m.maxAdj + j];
            const int *en = &m.elem2D_nodes[3 * (size_t)el];
            const int ps = en[0] == n ? 0 : (en[1] == n ? 1 : 2);
            if (en[ps] != n) throw std::runtime_error("nod_in_elem2D lists an element without the node");
            nt[q].ei[j] = el;
            nt[q].oa[j] = (int)I.ln.off[en[ps == 0 ? 1 : 0]];
            nt[q].ob[j] = (int)I.ln.off[en[ps == 2 ? 1 : 2]];
            nt[q].le[j] = (unsigned char)(m.nlevels[el] - 1);
            nt[q].pos[j] = (unsigned char)ps;
        }
        // the cluster bounds of FCT walk the fan; a node without one (a
        // caller-supplied element list that is not a fan) leaves its ring
        // empty and rules out the limiters that read it
        if (!node_fan(m, n, ring, fan)) {
            if (I.fan_fault < 0) I.fan_fault = n;
            continue;
        }
        if ((int)ring.size() > D) throw std::runtime_error("1-ring above " + std::to_string(D) + " nodes");
        for (size_t j = 0; j < ring.size(); ++j) {
            rg[q].oo[j] = (int)I.ln.off[ring[j]];
            rg[q].lr[j] = (unsigned char)(m.nlevels_nod2D[ring[j]] - 1);
            rg[q].le[j] = fan[j] >= 0 ? (unsigned char)(m.nlevels[fan[j]] - 1) : 0;
            // fan element j between ring j and ring j+1: its area-weighted
            // gradient coefficients of those two nodes
            const int el = fan[j];
            if (el < 0) continue;
            const int ra = ring[j], rb = ring[(j + 1) % ring.size()];
            const int *en = &m.elem2D_nodes[3 * (size_t)el];
            const int pa = en[0] == ra ? 0 : (en[1] == ra ? 1 : 2), pb = en[0] == rb ? 0 : (en[1] == rb ? 1 : 2);
            const double A = m.elem_area[el], *gs = &m.gradient_sca[6 * (size_t)el];
            R *w = &gw[(size_t)8 * D * q + 8 * j];
            w[0] = (R)(A * gs[pa]);
            w[1] = (R)(A * gs[pb]);
            w[2] = (R)(A * gs[3 + pa]);
            w[3] = (R)(A * gs[3 + pb]);
            w[4] = (R)A;
        }
    }
    d.deg = D;
    d.nedge = I.track(dupload(ne));
    d.ninfo = I.track(dupload(ni));
    d.nring = I.track(dupload(rg));
    d.nelem = I.track(dupload(nt));
    d.ngw = I.track(dupload(gw));
}

template <class R>
static void build_dmesh(Engine::Impl &I, DMesh<R> &d, const HostMesh &m)
{
    const int N = m.nod2D, T = m.elem2D, E = m.edge2D, nl = m.nl;
    d.N = N; d.T = T; d.E = E; d.nl = nl;
    d.nch_n = I.ln.nch; d.nch_t = I.lt.nch; d.nch_e = I.le.nch;
    d.CN = I.ln.total; d.CT = I.lt.total; d.CE = I.le.total;
    d.slot_n = I.track(dupload(I.ln.slot));
    d.slot_t = I.track(dupload(I.lt.slot));
    d.slot_e = I.track(dupload(I.le.slot));
    // per-column arrays in the layouts' orders (see Layout)
    {
        std::vector<int> nlq(N);
        for (int q = 0; q < N; ++q) nlq[q] = m.nlevels_nod2D[I.ln.order[q]];
        d.nlev = I.track(dupload(nlq));
        d.ncol = I.track(dupload(I.ln.order));
        d.nlev_c = I.track(dupload(m.nlevels_nod2D));
    }
    std::vector<EdgeRec> er(E);
    std::vector<int> enlmax(E);  // the caller's edge order
    std::vector<R> eg(8 * (size_t)E, (R)0), egud(4 * (size_t)E, (R)0);
    std::vector<int4> eud(E, make_int4(0, 0, 0, 0));
    for (int qe = 0; qe < E; ++qe) {
        const int e = I.le.order[qe];
        EdgeRec &r = er[qe];
        const int n1 = m.edges[2 * e], n2 = m.edges[2 * e + 1];
        const int el1 = m.edge_tri[2 * e], el2 = m.edge_tri[2 * e + 1];
        r.o1 = (int)I.ln.off[n1]; r.o2 = (int)I.ln.off[n2];
        r.t1 = (int)I.lt.off[el1]; r.t2 = el2 >= 0 ? (int)I.lt.off[el2] : -1;
        const int up = m.edge_up_dn_tri[2 * e], dn = m.edge_up_dn_tri[2 * e + 1];
        r.tu = up >= 0 ? (int)I.lt.off[up] : -1;
        r.td = dn >= 0 ? (int)I.lt.off[dn] : -1;
        const int nl1 = m.nlevels[el1] - 1, nl2 = el2 >= 0 ? m.nlevels[el2] - 1 : 0;
        const int nzmax_e = std::min(m.nlevels_nod2D_min[n1], m.nlevels_nod2D_min[n2]);
        const int nlmax = std::max(nl1, nl2);
        enlmax[e] = nlmax;
        r.nl = nl1 | (nl2 << 8) | (nzmax_e << 16) | (nlmax << 24);
        r.nb = (m.nboundary_lay[n1] & 0xFFFF) | ((m.nboundary_lay[n2] & 0xFFFF) << 16);
        double a = kRE * m.elem_cos[el1];
        if (el2 >= 0) a = 0.5 * (a + kRE * m.elem_cos[el2]);
        R *g = &eg[8 * (size_t)qe];
        for (int q = 0; q < 4; ++q) g[q] = (R)m.edge_cross_dxdy[4 * e + q];
        g[4] = (R)m.edge_dxdy[2 * e];
        g[5] = (R)m.edge_dxdy[2 * e + 1];
        g[6] = (R)a;
        // the up/down-wind triangles: ex*a*d/dx + ey*r_earth*d/dy of the triangle's
        // gradient as weights of its node values, the edge's own node first
        if (up >= 0 && dn >= 0) {
            const double ex = m.edge_dxdy[2 * e], ey = m.edge_dxdy[2 * e + 1];
            int off[4];
            double wt[4];
            for (int side = 0; side < 2; ++side) {
                const int tri = side == 0 ? up : dn, own = side == 0 ? n1 : n2;
                const int *en = &m.elem2D_nodes[3 * (size_t)tri];
                const int ps = en[0] == own ? 0 : (en[1] == own ? 1 : 2);
                if (en[ps] != own)
                    throw std::runtime_error("edge " + std::to_string(e + 1) + ": its " +
                                             (side == 0 ? "up" : "down") + "-wind triangle lacks the edge's node");
                const int ia = ps == 0 ? 1 : 0, ib = ps == 2 ? 1 : 2;
                const double *gs = &m.gradient_sca[6 * (size_t)tri];
                auto w = [&](int i) { return ex * a * gs[i] + ey * kRE * gs[3 + i]; };
                wt[2 * side] = w(ia);
                wt[2 * side + 1] = w(ib);
                off[2 * side] = (int)I.ln.off[en[ia]];
                off[2 * side + 1] = (int)I.ln.off[en[ib]];
            }
            eud[qe] = make_int4(off[0], off[1], off[2], off[3]);
            for (int q = 0; q < 4; ++q) egud[4 * (size_t)qe + q] = (R)wt[q];
        }
    }
    d.erec = I.track(dupload(er));
    d.egeo = I.track(dupload(eg));
    d.eud = I.track(dupload(eud));
    d.egud = I.track(dupload(egud));
    {
        constexpr int U = kEdgeUnits<R>;
        static_assert(sizeof(EdgeRec) == 32, "EdgeRec is two 16-byte units");
        std::vector<uint4> pk((size_t)U * E);
        for (int qe = 0; qe < E; ++qe) {
            char *q = reinterpret_cast<char *>(&pk[(size_t)U * qe]);
            std::memcpy(q, &er[qe], 32);
            std::memcpy(q + 32, &eg[8 * (size_t)qe], 8 * sizeof(R));
            std::memcpy(q + 32 + 8 * sizeof(R), &eud[qe], 16);
            std::memcpy(q + 48 + 8 * sizeof(R), &egud[4 * (size_t)qe], 4 * sizeof(R));
        }
        d.epack = I.track(dupload(pk));
    }
    std::vector<ElemRec> tr(T);
    std::vector<R> tg(8 * (size_t)T, (R)0);
    for (int el = 0; el < T; ++el) {
        const int *en = &m.elem2D_nodes[3 * el];
        tr[el] = ElemRec{(int)I.ln.off[en[0]], (int)I.ln.off[en[1]], (int)I.ln.off[en[2]], m.nlevels[el]};
        for (int q = 0; q < 6; ++q) tg[8 * (size_t)el + q] = (R)m.gradient_sca[6 * el + q];
        tg[8 * (size_t)el + 6] = (R)m.elem_area[el];
    }
    d.trec = I.track(dupload(tr));
    d.tgeo = I.track(dupload(tg));
    // node record width: kDeg unless some node has more incident edges or
    // elements (its 1-ring is never larger than its edge count)
    int need = 0;
    for (int n = 0; n < N; ++n)
        need = std::max(need, std::max(m.n2e_ptr[n + 1] - m.n2e_ptr[n], m.nod_in_elem2D_num[n]));
    static_assert(kDegWide == kMaxNodeDegree, "the host mesh check and the widest node record must agree");
    if (need > kDegWide) throw std::runtime_error("node degree above " + std::to_string(kDegWide) + " is not supported");
    if (need > kDeg) build_node_records<kDegWide>(I, d, m, enlmax);
    else build_node_records<kDeg>(I, d, m, enlmax);
    d.area = I.track(dupload(I.ln.pack<R>(m.area.data(), nl, 0)));
    d.zbarn = I.track(dupload(I.ln.pack<R>(m.zbar_3d_n.data(), nl, 0)));
    d.areasvol = I.track(dupload(I.ln.pack<R>(m.areasvol.data(), nl, 0)));
    d.hnode = I.track(dupload(I.ln.pack<R>(m.hnode.data(), nl - 1, 0)));
    // layers that keep their thickness over the step (hnode_new == hnode, no
    // tracer_set_thickness): one device copy, which the node kernels then
    // read once (ld_same)
    const bool same_h = m.hnode_new == m.hnode;
    d.hnode_new = same_h ? d.hnode : I.track(dupload(I.ln.pack<R>(m.hnode_new.data(), nl - 1, 0)));
    d.Zn = I.track(dupload(I.ln.pack<R>(m.Z_3d_n.data(), nl - 1, 0)));
    d.helem = I.track(dupload(I.lt.pack<R>(m.helem.data(), nl - 1, 0)));
    if constexpr (std::is_same<R, double>::value) {
        d.av_d = d.areasvol; d.h_d = d.hnode; d.hn_d = d.hnode_new;
    } else {
        d.av_d = I.track(dupload(I.ln.pack<double>(m.areasvol.data(), nl, 0)));
        d.h_d = I.track(dupload(I.ln.pack<double>(m.hnode.data(), nl - 1, 0)));
        d.hn_d = same_h ? d.h_d : I.track(dupload(I.ln.pack<double>(m.hnode_new.data(), nl - 1, 0)));
    }
}

Engine::Engine(const HostMesh &m, int ntracers, int precision)
    : p_(new Impl), N_(m.nod2D), T_(m.elem2D), E_(m.edge2D), nl_(m.nl), ntr_(ntracers), prec_(precision)
{
    // a column's interfaces fill at most one 128-cell chunk (two per lane of one wave)
    if (nl_ > kMaxLevels) throw std::runtime_error("the device path supports nl <= 128 level interfaces");
    if (ntracers < 1) throw std::runtime_error("need at least one tracer");
    if (precision < PREC_FP64 || precision > PREC_FP32_H16) throw std::runtime_error("unknown precision");
    for (int u : m.ulevels) if (u != 1) throw std::runtime_error("cavities (ulevels > 1) are not supported");
    S_ = 0;
    spec_.assign(ntr_, TracerSpec());
    Impl &I = *p_;
    {
        std::vector<int> a(N_), c(T_), d(E_);
        // node columns hold the layers; interfaces 1..L share their cells, and
        // the bottom interface (zero flux in every scheme) has none
        for (int n = 0; n < N_; ++n) a[n] = m.nlevels_nod2D[n] - 1;
        for (int el = 0; el < T_; ++el) c[el] = m.nlevels[el] - 1;
        for (int e = 0; e < E_; ++e) {
            int el1 = m.edge_tri[2 * e], el2 = m.edge_tri[2 * e + 1];
            d[e] = std::max(m.nlevels[el1] - 1, el2 >= 0 ? m.nlevels[el2] - 1 : 0);
        }
        // node and edge columns may be passed over by up to 128 later ones that
        // fill a chunk's tail (a chunk then wastes about 1-2% instead of 11-13%
        // of its cells on the benchmark's random bottoms). Where the columns are
        // long enough that a full chunk holds few of them, a chunk also stays
        // within what a wave stages in LDS: 8 node columns (FP64 fan weights) or
        // 16 (node->edge records), 16 edge columns; shallower meshes, whose
        // chunks hold more columns than that anyway, keep chunks full instead
        constexpr int kPackWindow = 128;
        auto col_cap = [](const std::vector<int> &lens, std::initializer_list<int> caps) {
            double cells = 0.0;
            for (int l : lens) cells += l + (l & 1);
            const double mean = lens.empty() ? (double)kChunk : cells / (double)lens.size();
            for (int cap : caps)  // the tightest cap whose columns still fill a chunk
                if (mean * cap >= kChunk) return cap;
            return kChunk;
        };
        I.ln.build(a, kPackWindow, col_cap(a, {FanW<double, kDeg>::kCols, kStageNodes}));
        I.lt.build(c);
        I.le.build(d, kPackWindow, col_cap(d, {kStageEdges}));
        if (I.le.total >= (1ll << 31) || I.ln.total >= (1ll << 31) || I.lt.total >= (1ll << 31))
            throw std::runtime_error("mesh too large for 32-bit column offsets");
    }
    HIP_OK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    const size_t rs = prec_ == PREC_FP64 ? sizeof(double) : sizeof(float);
    const size_t gs = prec_ == PREC_FP64 ? sizeof(double) : (prec_ == PREC_FP32 ? sizeof(float) : sizeof(__half));
    if (prec_ == PREC_FP64) build_dmesh(I, I.m64, m);
    else build_dmesh(I, I.m32, m);
    I.tsn = (size_t)I.ln.total;
    I.tse = (size_t)I.le.total;
    // work arrays hold one tracer batch (at most 4 tracers); batches run in
    // stream order, so every batch reuses them
    I.wslots = std::min(ntr_, kMaxBatch);
    auto zalloc = [&](size_t bytes) {
        void *p = I.track((void *)dalloc<char>(bytes));
        HIP_OK(hipMemset(p, 0, bytes));
        return p;
    };
    I.vel = zalloc(2 * (size_t)I.lt.total * rs);
    I.W = zalloc(I.tsn * rs);
    I.WE = zalloc(I.tsn * rs);
    I.values = (double *)zalloc(I.tsn * ntr_ * sizeof(double));
    // valuesAB only exists once a caller sets it (the default refresh mode reads values)
    I.vfl = zalloc(I.tse * rs);
    // HO-LO edge fluxes: FP32 interleaves work slots in pairs (adf_at), so an odd count is rounded up
    I.adf_h = zalloc(I.tse * (rs == 8 ? I.wslots : (I.wslots + 1) / 2 * 2) * rs);
    I.adf_v = zalloc(I.tsn * I.wslots * rs);
    // R+/R-: two arrays in FP64, one interleaved array in FP32 (ld_rpm)
    I.Pp = zalloc(I.tsn * I.wslots * 2 * rs);
    I.Pm = rs == 8 ? (char *)I.Pp + I.tsn * I.wslots * rs : I.Pp;
    I.LO = (double *)zalloc(I.tsn * I.wslots * sizeof(double));
    I.gnod = zalloc(2 * I.tsn * I.wslots * gs);
    I.dpart = I.track(dalloc<DiagAcc>(I.ndiag));
    I.dout = I.track(dalloc<DiagAcc>(1));
    I.epart = I.track(dalloc<ErrAcc>(I.ndiag));
    // network parameters: FP32 master, Adam moments, gradient, bf16 operand copies
    I.mlp_p = I.track(dalloc<float>(kMlpParams));
    I.mlp_m1 = (float *)zalloc(kMlpParams * sizeof(float));
    I.mlp_m2 = (float *)zalloc(kMlpParams * sizeof(float));
    I.mlp_g = (float *)zalloc(kMlpParams * sizeof(float));
    I.mlp_loss = (float *)zalloc(sizeof(float));
    I.w1t = (unsigned short *)zalloc(kMlpHid * kMlpInPad * 2);
    I.w2t = (unsigned short *)zalloc(kMlpHid * kMlpHid * 2);
    I.w2 = (unsigned short *)zalloc(kMlpHid * kMlpHid * 2);
    I.w3 = (unsigned short *)zalloc(kMlpHid * 2);
    I.b1 = (float *)zalloc(kMlpHid * sizeof(float));
    I.b2 = (float *)zalloc(kMlpHid * sizeof(float));
    I.b3 = (float *)zalloc(sizeof(float));
    I.cnt = (unsigned long long *)zalloc(2 * sizeof(unsigned long long));
    {
        // untrained network: uniform Glorot weights from a fixed seed, zero
        // biases, output bias 2 (coefficient ~0.88)
        std::vector<float> p(kMlpParams, 0.0f);
        uint64_t st = 0x9E3779B97F4A7C15ull;
        auto uni = [&]() {
            st = st * 6364136223846793005ull + 1442695040888963407ull;
            return (double)(st >> 11) * (1.0 / 9007199254740992.0);
        };
        auto fill = [&](int off, int n, double lim) {
            for (int k = 0; k < n; ++k) p[off + k] = (float)((2.0 * uni() - 1.0) * lim);
        };
        fill(kMlpOffW1, kMlpIn * kMlpHid, std::sqrt(6.0 / (kMlpIn + kMlpHid)));
        fill(kMlpOffW2, kMlpHid * kMlpHid, std::sqrt(6.0 / (2 * kMlpHid)));
        fill(kMlpOffW3, kMlpHid, std::sqrt(6.0 / (kMlpHid + 1)));
        p[kMlpOffB3] = 2.0f;
        mlp_set(p.data());
    }
    I.dense = (double *)zalloc((size_t)(nl_ - 1) * N_ * ntr_ * sizeof(double));
    // FP16 gradient scale: mean edge length (gradient * length = tracer increment)
    double len = 0.0;
    for (int e = 0; e < E_; ++e) len += std::hypot(m.edge_dxdy[2 * e], m.edge_dxdy[2 * e + 1]) * kRE;
    len = E_ ? len / E_ : 1.0;
    if (prec_ == PREC_FP32_H16) { I.gscale = (float)len; I.ginv = (float)(1.0 / len); }
}

Engine::~Engine()
{
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    for (void *p : p_->allocs) (void)hipFree(p);
    if (p_->dsX) { (void)hipFree(p_->dsX); (void)hipFree(p_->dsY); }
    for (int b = 0; b < 2; ++b) {
        if (p_->stage_ev[b]) (void)hipEventDestroy(p_->stage_ev[b]);
        if (p_->stage[b]) (void)hipHostFree(p_->stage[b]);
    }
    if (stream_) (void)hipStreamDestroy(stream_);
    delete p_;
}

// Q+/Q- and P+/P- of every work slot for the learned limiter (features and
// bound check), each pair stored as R+/R- are (ld_rpm): two arrays in FP64,
// one interleaved array of twice the length in FP32
static void alloc_learn_bounds(Engine::Impl &I, size_t rs)
{
    const size_t n = I.tsn * I.wslots;
    for (void **q : {&I.Qp, &I.Sp}) {
        *q = I.track((void *)dalloc<char>(2 * n * rs));
        HIP_OK(hipMemset(*q, 0, 2 * n * rs));
    }
    I.Qm = rs == 8 ? (char *)I.Qp + n * rs : I.Qp;
    I.Sm = rs == 8 ? (char *)I.Sp + n * rs : I.Sp;
}

void Engine::set_spec(int tr, const TracerSpec &sp)
{
    Impl &I = *p_;
    if ((sp.lim == LIM_FCT || sp.lim == LIM_LEARNED) && I.fan_fault >= 0)
        throw std::runtime_error("FCT and LEARNED need the elements around every node to form one fan; node " +
                                 std::to_string(I.fan_fault + 1) + " does not");
    spec_.at(tr) = sp;
    graph_valid_ = false;
    if (sp.lim == LIM_LEARNED && !I.alpha) {
        const size_t rs = prec_ == PREC_FP64 ? sizeof(double) : sizeof(float);
        auto zalloc = [&](size_t bytes) {
            void *p = I.track((void *)dalloc<char>(bytes));
            HIP_OK(hipMemset(p, 0, bytes));
            return p;
        };
        // coefficients of the edge levels, stored as the HO-LO fluxes are (adf_at)
        I.alpha = zalloc(I.tse * (rs == 8 ? I.wslots : (I.wslots + 1) / 2 * 2) * rs);
        // the fallback flags stay per tracer (tracer_learn_fallback reads them after the step)
        I.unsafe = zalloc(I.tsn * ntr_ * rs);
        if (!I.Qp) alloc_learn_bounds(I, rs);
    }
}

void Engine::set_velocity(const double *uv, const double *w, const double *we)
{
    Impl &I = *p_;
    auto put = [&](void *dst, const Layout &lay, const double *src, int L, int comps) {
        if (prec_ == PREC_FP64) {
            auto v = lay.pack<double>(src, L, 0, comps);
            HIP_OK(hipMemcpy(dst, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
        } else {
            auto v = lay.pack<float>(src, L, 0, comps);
            HIP_OK(hipMemcpy(dst, v.data(), v.size() * sizeof(float), hipMemcpyHostToDevice));
        }
    };
    put(I.vel, I.lt, uv, nl_ - 1, 2);
    put(I.W, I.ln, w, nl_, 1);
    const bool same = w == we || std::memcmp(w, we, sizeof(double) * (size_t)nl_ * N_) == 0;
    if (!same) put(I.WE, I.ln, we, nl_, 1);
    if (same != I.we_is_w) graph_valid_ = false;  // the captured launches name the other array
    I.we_is_w = same;
}

// the node layout's pair slots, column heights and caller node of each
// column (the same in both precisions)
static void node_slots(const Engine::Impl &I, int prec, const int *&slot, const int *&nlev, const int *&ncol,
                       long long &npair)
{
    if (prec == PREC_FP64) { slot = I.m64.slot_n; nlev = I.m64.nlev; ncol = I.m64.ncol; npair = I.m64.CN / 2; }
    else { slot = I.m32.slot_n; nlev = I.m32.nlev; ncol = I.m32.ncol; npair = I.m32.CN / 2; }
}

static void ensure_staging(Engine::Impl &I, hipStream_t st, size_t array_bytes)
{
    if (!I.stage_bytes)
        I.stage_bytes = std::min(Engine::Impl::kStageMax, std::max<size_t>(4096, (array_bytes + 4095) & ~(size_t)4095));
    for (int b = 0; b < 2; ++b) {
        if (I.stage[b]) continue;
        HIP_OK(hipHostMalloc(&I.stage[b], I.stage_bytes, hipHostMallocDefault));
        HIP_OK(hipEventCreateWithFlags(&I.stage_ev[b], hipEventDisableTiming));
        HIP_OK(hipEventRecord(I.stage_ev[b], st));
    }
}

// host (pageable) -> device through the pinned buffers, in stream order on st
static void h2d_staged(Engine::Impl &I, hipStream_t st, void *dst, const void *src, size_t bytes)
{
    ensure_staging(I, st, bytes);
    const size_t B = I.stage_bytes;
    int b = 0;
    for (size_t o = 0; o < bytes; o += B, b ^= 1) {
        const size_t n = std::min(B, bytes - o);
        HIP_OK(hipEventSynchronize(I.stage_ev[b]));  // the DMA that last read this buffer is done
        copy_host(I.stage[b], (const char *)src + o, n);
        HIP_OK(hipMemcpyAsync((char *)dst + o, I.stage[b], n, hipMemcpyHostToDevice, st));
        HIP_OK(hipEventRecord(I.stage_ev[b], st));
    }
    HIP_OK(hipStreamSynchronize(st));
}

// device -> host (pageable) through the pinned buffers, after the work on st
static void d2h_staged(Engine::Impl &I, hipStream_t st, void *dst, const void *src, size_t bytes)
{
    ensure_staging(I, st, bytes);
    const size_t B = I.stage_bytes;
    const size_t nchunk = (bytes + B - 1) / B;
    auto issue = [&](size_t i) {
        const size_t o = i * B, n = std::min(B, bytes - o);
        HIP_OK(hipMemcpyAsync(I.stage[i & 1], (const char *)src + o, n, hipMemcpyDeviceToHost, st));
        HIP_OK(hipEventRecord(I.stage_ev[i & 1], st));
    };
    if (nchunk) issue(0);
    for (size_t i = 0; i < nchunk; ++i) {
        // the other buffer was last read by the host copy of chunk i-1, which has returned
        if (i + 1 < nchunk) issue(i + 1);
        HIP_OK(hipEventSynchronize(I.stage_ev[i & 1]));
        const size_t o = i * B, n = std::min(B, bytes - o);
        copy_host((char *)dst + o, I.stage[i & 1], n);
    }
    HIP_OK(hipStreamSynchronize(st));
}

// dense (nl-1, nod2D) host array -> device dense copy -> chunk layout at dst
static void upload_layers(Engine::Impl &I, int prec, hipStream_t st, int N, int L, const double *v, double *dense,
                          double *dst)
{
    const int *slot, *nlev, *ncol;
    long long npair;
    node_slots(I, prec, slot, nlev, ncol, npair);
    HIP_OK(hipStreamSynchronize(st));  // no step or earlier transfer still reads dense or dst
    h2d_staged(I, st, dense, v, (size_t)L * N * sizeof(double));
    hipLaunchKernelGGL(k_pack_layers, dim3((unsigned)((npair + 255) / 256)), dim3(256), 0, st, slot, nlev, ncol,
                       npair, dense, L, dst);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(st));
}

void Engine::set_values(int tr, const double *v)
{
    if (tr < 0 || tr >= ntr_) throw std::runtime_error("tracer index out of range");
    Impl &I = *p_;
    const size_t cells = (size_t)(nl_ - 1) * N_;
    upload_layers(I, prec_, stream_, N_, nl_ - 1, v, I.dense + tr * cells, I.values + tr * I.tsn);
}

void Engine::set_valuesAB(int tr, const double *v)
{
    if (tr < 0 || tr >= ntr_) throw std::runtime_error("tracer index out of range");
    Impl &I = *p_;
    if (!I.valuesAB) {
        HIP_OK(hipStreamSynchronize(stream_));
        I.valuesAB = I.track(dalloc<double>(I.tsn * ntr_));
        HIP_OK(hipMemcpy(I.valuesAB, I.values, I.tsn * ntr_ * sizeof(double), hipMemcpyDeviceToDevice));
    }
    if (!I.dense_ab) I.dense_ab = I.track(dalloc<double>((size_t)(nl_ - 1) * N_));
    upload_layers(I, prec_, stream_, N_, nl_ - 1, v, I.dense_ab, I.valuesAB + tr * I.tsn);
}

void Engine::get_values(int tr, double *v)
{
    if (tr < 0 || tr >= ntr_) throw std::runtime_error("tracer index out of range");
    Impl &I = *p_;
    const int *slot, *nlev, *ncol;
    long long npair;
    node_slots(I, prec_, slot, nlev, ncol, npair);
    const int L = nl_ - 1;
    double *dense = I.dense + tr * (size_t)L * N_;
    hipLaunchKernelGGL(k_unpack_layers, dim3((unsigned)((npair + 255) / 256)), dim3(256), 0, stream_, slot, nlev,
                       ncol, npair, I.values + tr * I.tsn, L, dense);
    HIP_OK(hipGetLastError());
    d2h_staged(I, stream_, v, dense, (size_t)L * N_ * sizeof(double));
}

void Engine::set_accumulate(bool on)
{
    Impl &I = *p_;
    HIP_OK(hipStreamSynchronize(stream_));
    if (on) {
        if (!I.del_ttf) I.del_ttf = I.track(dalloc<double>(I.tsn * ntr_));
        HIP_OK(hipMemset(I.del_ttf, 0, I.tsn * ntr_ * sizeof(double)));
    }
    I.out_mode = on ? OUT_ACCUM : OUT_UPDATE;
    graph_valid_ = false;
}

bool Engine::accumulate() const { return p_->out_mode == OUT_ACCUM; }

void Engine::tendencies(int tr, double dt, double *dttf_h, double *dttf_v)
{
    if (tr < 0 || tr >= ntr_) throw std::runtime_error("tracer index out of range");
    Impl &I = *p_;
    if (spec_[tr].lim == LIM_LEARNED) throw std::runtime_error("tendencies are not returned for the LEARNED limiter");
    const size_t L = nl_ - 1;
    if (!I.tend_h) {
        I.tend_h = I.track(dalloc<double>(I.tsn));
        I.tend_v = I.track(dalloc<double>(I.tsn));
    }
    if (!I.dense_t) I.dense_t = I.track(dalloc<double>(L * N_));
    const int saved = I.out_mode;
    I.out_mode = OUT_TEND;
    try {
        enqueue_step(dt, tr, 1);
    } catch (...) {
        I.out_mode = saved;
        throw;
    }
    I.out_mode = saved;
    const int *slot, *nlev, *ncol;
    long long npair;
    node_slots(I, prec_, slot, nlev, ncol, npair);
    for (int q = 0; q < 2; ++q) {
        double *host = q == 0 ? dttf_h : dttf_v;
        h2d_staged(I, stream_, I.dense_t, host, L * N_ * sizeof(double));
        hipLaunchKernelGGL(k_unpack_add_layers, dim3((unsigned)((npair + 255) / 256)), dim3(256), 0, stream_, slot,
                           nlev, ncol, npair, q == 0 ? I.tend_h : I.tend_v, (int)L, I.dense_t);
        HIP_OK(hipGetLastError());
        d2h_staged(I, stream_, host, I.dense_t, L * N_ * sizeof(double));
    }
}

void Engine::tendencies_device(int tr, double dt, double *dttf_h, double *dttf_v, double *scratch,
                               size_t scratch_doubles)
{
    if (tr < 0 || tr >= ntr_) throw std::runtime_error("tracer index out of range");
    Impl &I = *p_;
    if (spec_[tr].lim == LIM_LEARNED) throw std::runtime_error("tendencies are not returned for the LEARNED limiter");
    if (2 * I.tsn > scratch_doubles)
        throw std::runtime_error("the mesh needs " + std::to_string(2 * I.tsn * sizeof(double) >> 20) +
                                 " MiB of tendency scratch, more than the " +
                                 std::to_string(scratch_doubles * sizeof(double) >> 20) + " MiB available");
    // the step writes every node layer it covers into the two scratch halves
    // (nothing needs to start at zero); the engine's own tendency arrays are
    // left as they are
    double *own_h = I.tend_h, *own_v = I.tend_v;
    const int saved = I.out_mode;
    I.tend_h = scratch;
    I.tend_v = scratch + I.tsn;
    I.out_mode = OUT_TEND;
    try {
        enqueue_step(dt, tr, 1);
    } catch (...) {
        I.tend_h = own_h; I.tend_v = own_v; I.out_mode = saved;
        throw;
    }
    I.tend_h = own_h; I.tend_v = own_v; I.out_mode = saved;
    const int *slot, *nlev, *ncol;
    long long npair;
    node_slots(I, prec_, slot, nlev, ncol, npair);
    for (int q = 0; q < 2; ++q)
        hipLaunchKernelGGL(k_unpack_add_layers, dim3((unsigned)((npair + 255) / 256)), dim3(256), 0, stream_, slot,
                           nlev, ncol, npair, scratch + (size_t)q * I.tsn, nl_ - 1, q == 0 ? dttf_h : dttf_v);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(stream_));
}

struct LaunchHook {
    std::vector<std::pair<std::string, hipEvent_t>> *ev = nullptr;
    void mark(const char *name, hipStream_t st)
    {
        if (!ev) return;
        hipEvent_t e;
        HIP_OK(hipEventCreate(&e));
        HIP_OK(hipEventRecord(e, st));
        ev->emplace_back(name, e);
    }
};

static MlpDev mlp_dev(const Engine::Impl &I)
{
    return MlpDev{I.w1t, I.w2t, I.w2, I.w3, I.b1, I.b2, I.b3};
}

// one step of the NT tracers [t0, t0+NT), all with schemes sp, on a mesh whose
// node records are D wide
template <class R, class G, int NT, int D>
static void launch_batch(Engine::Impl &I, const DMesh<R> &m, const TracerSpec &sp, double dt, int t0, bool refresh,
                         hipStream_t st, LaunchHook &hk)
{
    const R *vel = (const R *)I.vel, *W = (const R *)I.W, *WE = (const R *)(I.we_is_w ? I.W : I.WE);
    const size_t tsn = I.tsn, tse = I.tse, tsg = 2 * I.tsn;
    double *vals = I.values + t0 * tsn;
    if (!refresh && !I.valuesAB) throw std::runtime_error("valuesAB was never set");
    const double *ab = refresh ? vals : I.valuesAB + t0 * tsn;
    R *vfl = (R *)I.vfl, *adf_h = (R *)I.adf_h;
    R *adf_v = (R *)I.adf_v, *Pp = (R *)I.Pp, *Pm = (R *)I.Pm;
    double *LO = I.LO;
    G *gnod = (G *)I.gnod;
    const int wv = I.tune_waves;
    const dim3 blk(kWave * wv), ge(grid_for(m.nch_e, wv)), gn(grid_for(m.nch_n, wv));
    const bool learned = sp.lim == LIM_LEARNED;
    const bool fct = sp.lim == LIM_FCT || learned;
    Engine::Impl::Dump *dump = I.dump;
    const R ph = (R)sp.ph, pv = (R)sp.pv;
    // how the step ends: update the tracer, update it carrying del_ttf over
    // from the last step (one array per tracer), or only return the tendencies
    const int out = I.out_mode;
    if (out != OUT_UPDATE && learned) throw std::runtime_error("the LEARNED limiter updates the tracer only");
    if (out == OUT_TEND && NT != 1) throw std::runtime_error("tendencies are returned one tracer at a time");
    double *acc_h = out == OUT_ACCUM ? I.del_ttf + t0 * tsn : (out == OUT_TEND ? I.tend_h : nullptr);
    double *acc_v = out == OUT_TEND ? I.tend_v : nullptr;
    if (out != OUT_UPDATE && !acc_h) throw std::runtime_error("accumulation arrays not allocated");
    if (sp.hor == HOR_MUSCL || sp.hor == HOR_MFCT) {
        using FW = FanW<R, D>;
        const size_t flds = (size_t)wv * FW::kCols * FW::kUnits * sizeof(uint4);  // k_gnod_fan's LDS slices
        if (I.grad_fan && I.fan_fault < 0)
            hipLaunchKernelGGL((k_gnod_fan<R, G, NT, D>), gn, blk, flds, st, m, ab, tsn, gnod, tsg, I.gscale);
        else
            hipLaunchKernelGGL((k_gnod<R, G, NT, D>), gn, blk, 0, st, m, ab, tsn, gnod, tsg, I.gscale);
        hk.mark("k_gnod", st);
    }
    const size_t elds = (size_t)wv * kStageEdges * kEdgeUnits<R> * sizeof(uint4);  // k_edge's LDS slices
    const size_t nlds = (size_t)wv * kStageNodes * kNodeEdgeUnits<D> * sizeof(uint4);  // node kernels' LDS slices
#define EDGE(HOR, F)                                                                                        \
    do {                                                                                                    \
        if (refresh)                                                                                        \
            hipLaunchKernelGGL((k_edge<R, G, HOR, F, NT, true>), ge, blk, elds, st, m, vel, vals, ab, tsn,          \
                               gnod, tsg, ph, I.ginv, vfl, adf_h, tse);                                 \
        else                                                                                                \
            hipLaunchKernelGGL((k_edge<R, G, HOR, F, NT, false>), ge, blk, elds, st, m, vel, vals, ab, tsn,         \
                               gnod, tsg, ph, I.ginv, vfl, adf_h, tse);                                 \
    } while (0)
    if (fct) {
        if (sp.hor == HOR_MUSCL) EDGE(HOR_MUSCL, true);
        else if (sp.hor == HOR_MFCT) EDGE(HOR_MFCT, true);
        else EDGE(HOR_UPW1, true);
    } else {
        if (sp.hor == HOR_MUSCL) EDGE(HOR_MUSCL, false);
        else if (sp.hor == HOR_MFCT) EDGE(HOR_MFCT, false);
        else EDGE(HOR_UPW1, false);
    }
#undef EDGE
    hk.mark("k_edge", st);
    if (fct) {
#define NLO(V)                                                                                              \
    do {                                                                                                    \
        if (refresh)                                                                                        \
            hipLaunchKernelGGL((k_node_lo<R, V, NT, true, D>), gn, blk, nlds, st, m, W, WE, vals, ab, tsn, vfl, \
                               dt, pv, LO, adf_v);                                                          \
        else                                                                                                \
            hipLaunchKernelGGL((k_node_lo<R, V, NT, false, D>), gn, blk, nlds, st, m, W, WE, vals, ab, tsn, vfl, \
                               dt, pv, LO, adf_v);                                                          \
    } while (0)
        switch (sp.ver) {
        case VER_QR4C: NLO(VER_QR4C); break;
        case VER_PPM: NLO(VER_PPM); break;
        case VER_CDIFF: NLO(VER_CDIFF); break;
        default: NLO(VER_UPW1); break;
        }
#undef NLO
        hk.mark("k_node_lo", st);
        R *Qp = learned || dump ? (R *)I.Qp : nullptr;
        R *Qm = learned || dump ? (R *)I.Qm : nullptr;
        R *Sp = learned || dump ? (R *)I.Sp : nullptr;
        R *Sm = learned || dump ? (R *)I.Sm : nullptr;
        // R+/R- go to the arrays named Pp/Pm (fct_plus/fct_minus)
        if (learned || dump)
            hipLaunchKernelGGL((k_node_limit<R, NT, true, D>), gn, blk, nlds, st, m, vals, LO, tsn, adf_h, tse, adf_v,
                               dt, Pp, Pm, Qp, Qm, Sp, Sm);
        else
            hipLaunchKernelGGL((k_node_limit<R, NT, false, D>), gn, blk, nlds, st, m, vals, LO, tsn, adf_h, tse, adf_v,
                               dt, Pp, Pm, Qp, Qm, Sp, Sm);
        hk.mark("k_node_limit", st);
        const LearnArgs<R, NT> la{vals, LO, adf_h, vfl, Qp, Qm, Pp, Pm, Sp, Sm, tsn, tse};
        R *alpha = learned ? (R *)I.alpha : nullptr, *unsafe = learned ? (R *)I.unsafe + t0 * tsn : nullptr;
        if (dump) {
            if (dump->row + (long long)NT * m.CE <= dump->cap) {
                hipLaunchKernelGGL((k_learn_dump<R, NT>), ge, blk, 0, st, m, la, dt, dump->X, dump->Y, dump->row);
                dump->row += (long long)NT * m.CE;
            }
            hipLaunchKernelGGL((k_node_update<R, NT, false, D, OUT_UPDATE>), gn, blk, nlds, st, m, vals, LO, tsn, adf_h,
                               tse, adf_v, Pp, Pm, dt, alpha, nullptr, nullptr);
        } else if (learned) {
            const MlpDev w = mlp_dev(I);
            const int gb = std::min((m.nch_e + kWavesPerBlock - 1) / kWavesPerBlock, 2048);
            const int iters = (m.nch_e + gb * kWavesPerBlock - 1) / (gb * kWavesPerBlock);
            hipLaunchKernelGGL((k_learn_alpha<R, NT>), dim3(gb), dim3(kBlock), 0, st, m, w, la, dt, alpha, iters);
            hk.mark("k_learn_alpha", st);
            hipLaunchKernelGGL((k_learn_check<R, NT, D>), gn, blk, 0, st, m, adf_h, alpha, tse, adf_v, Pp, Pm, Qp, Qm,
                               tsn, dt, unsafe);
            hk.mark("k_learn_check", st);
            hipLaunchKernelGGL((k_node_update<R, NT, true, D, OUT_UPDATE>), gn, blk, nlds, st, m, vals, LO, tsn, adf_h,
                               tse, adf_v, Pp, Pm, dt, alpha, nullptr, nullptr);
        } else if (out == OUT_ACCUM) {
            hipLaunchKernelGGL((k_node_update<R, NT, false, D, OUT_ACCUM>), gn, blk, nlds, st, m, vals, LO, tsn, adf_h,
                               tse, adf_v, Pp, Pm, dt, alpha, acc_h, nullptr);
        } else if (out == OUT_TEND) {
            if constexpr (NT == 1)
                hipLaunchKernelGGL((k_node_update<R, 1, false, D, OUT_TEND>), gn, blk, nlds, st, m, vals, LO, tsn,
                                   adf_h, tse, adf_v, Pp, Pm, dt, alpha, acc_h, acc_v);
        } else {
            hipLaunchKernelGGL((k_node_update<R, NT, false, D, OUT_UPDATE>), gn, blk, nlds, st, m, vals, LO, tsn, adf_h,
                               tse, adf_v, Pp, Pm, dt, alpha, nullptr, nullptr);
        }
        hk.mark("k_node_update", st);
    } else {
#define NNL(V)                                                                                                      \
    do {                                                                                                            \
        if (out == OUT_ACCUM)                                                                                       \
            hipLaunchKernelGGL((k_node_nolim<R, V, NT, D, OUT_ACCUM>), gn, blk, 0, st, m, WE, vals, ab, tsn, adf_h, \
                               tse, dt, pv, acc_h, nullptr);                                                        \
        else if (out != OUT_TEND)                                                                                   \
            hipLaunchKernelGGL((k_node_nolim<R, V, NT, D, OUT_UPDATE>), gn, blk, 0, st, m, WE, vals, ab, tsn, adf_h, \
                               tse, dt, pv, nullptr, nullptr);                                                      \
        else if constexpr (NT == 1)                                                                                 \
            hipLaunchKernelGGL((k_node_nolim<R, V, 1, D, OUT_TEND>), gn, blk, 0, st, m, WE, vals, ab, tsn, adf_h,   \
                               tse, dt, pv, acc_h, acc_v);                                                          \
    } while (0)
        switch (sp.ver) {
        case VER_QR4C: NNL(VER_QR4C); break;
        case VER_PPM: NNL(VER_PPM); break;
        case VER_CDIFF: NNL(VER_CDIFF); break;
        default: NNL(VER_UPW1); break;
        }
#undef NNL
        hk.mark("k_node_nolim", st);
    }
    HIP_OK(hipGetLastError());
}

static bool same_spec(const TracerSpec &a, const TracerSpec &b)
{
    return a.hor == b.hor && a.ver == b.ver && a.lim == b.lim && a.ph == b.ph && a.pv == b.pv;
}

// tracers in the next launch of a run of identical schemes (4, 2 or 1, at most max_nt)
static int batch_size(int run, int max_nt)
{
    if (run >= 4 && max_nt >= 4) return 4;
    if (run >= 2 && max_nt >= 2) return 2;
    return 1;
}

// position of tracer tr in the work arrays: its index within the batch that
// launch_range puts it in
static int work_slot(const std::vector<TracerSpec> &spec, int tr, int max_nt)
{
    const int n = (int)spec.size();
    for (int t = 0; t < n;) {
        int run = 1;
        while (t + run < n && same_spec(spec[t], spec[t + run])) ++run;
        while (run > 0) {
            const int nb = batch_size(run, max_nt);
            if (tr >= t && tr < t + nb) return tr - t;
            t += nb;
            run -= nb;
        }
    }
    return 0;
}

template <class R, class G>
static void launch_range(Engine::Impl &I, const DMesh<R> &m, const std::vector<TracerSpec> &spec, double dt, int first,
                         int count, bool refresh, hipStream_t st, LaunchHook &hk)
{
    // runs of consecutive tracers with identical schemes, in batches of 4/2/1
    int t = first;
    const int end = first + count;
    while (t < end) {
        int run = 1;
        while (t + run < end && same_spec(spec[t], spec[t + run])) ++run;
        while (run > 0) {
            hk.mark("start", st);
            const int nb = batch_size(run, I.tune_nt);
            if (m.deg == kDegWide) {
                if (nb == 4) launch_batch<R, G, 4, kDegWide>(I, m, spec[t], dt, t, refresh, st, hk);
                else if (nb == 2) launch_batch<R, G, 2, kDegWide>(I, m, spec[t], dt, t, refresh, st, hk);
                else launch_batch<R, G, 1, kDegWide>(I, m, spec[t], dt, t, refresh, st, hk);
            } else {
                if (nb == 4) launch_batch<R, G, 4, kDeg>(I, m, spec[t], dt, t, refresh, st, hk);
                else if (nb == 2) launch_batch<R, G, 2, kDeg>(I, m, spec[t], dt, t, refresh, st, hk);
                else launch_batch<R, G, 1, kDeg>(I, m, spec[t], dt, t, refresh, st, hk);
            }
            t += nb;
            run -= nb;
        }
    }
}

static void enqueue(Engine::Impl &I, int prec, const std::vector<TracerSpec> &spec, double dt, int first, int count,
                    bool refresh, hipStream_t st, LaunchHook hk = LaunchHook())
{
    if (prec == PREC_FP64) launch_range<double, double>(I, I.m64, spec, dt, first, count, refresh, st, hk);
    else if (prec == PREC_FP32) launch_range<float, float>(I, I.m32, spec, dt, first, count, refresh, st, hk);
    else launch_range<float, __half>(I, I.m32, spec, dt, first, count, refresh, st, hk);
}

void Engine::enqueue_step(double dt, int first, int count)
{
    if (first < 0 || count < 0 || first + count > ntr_) throw std::runtime_error("tracer range out of bounds");
    enqueue(*p_, prec_, spec_, dt, first, count, refresh_ab_, stream_);
}

void Engine::run_steps(int nsteps, double dt, bool use_graph)
{
    if (!use_graph) {
        for (int s = 0; s < nsteps; ++s) enqueue_step(dt, 0, ntr_);
        return;
    }
    if (!graph_valid_ || graph_dt_ != dt) {
        if (graph_exec_) { HIP_OK(hipGraphExecDestroy(graph_exec_)); graph_exec_ = nullptr; }
        if (graph_) { HIP_OK(hipGraphDestroy(graph_)); graph_ = nullptr; }
        HIP_OK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        enqueue_step(dt, 0, ntr_);
        HIP_OK(hipStreamEndCapture(stream_, &graph_));
        HIP_OK(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
        graph_valid_ = true;
        graph_dt_ = dt;
    }
    for (int s = 0; s < nsteps; ++s) HIP_OK(hipGraphLaunch(graph_exec_, stream_));
}

int Engine::mlp_param_count() { return kMlpParams; }

static void mlp_pack(Engine::Impl &I, hipStream_t st)
{
    hipLaunchKernelGGL(k_mlp_pack, dim3((kMlpHid * kMlpHid + 255) / 256), dim3(256), 0, st, I.mlp_p,
                       MlpMut{I.w1t, I.w2t, I.w2, I.w3, I.b1, I.b2, I.b3});
    HIP_OK(hipGetLastError());
}

void Engine::mlp_set(const float *params)
{
    Impl &I = *p_;
    HIP_OK(hipMemcpy(I.mlp_p, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice));
    HIP_OK(hipMemset(I.mlp_m1, 0, kMlpParams * sizeof(float)));
    HIP_OK(hipMemset(I.mlp_m2, 0, kMlpParams * sizeof(float)));
    I.adam_t = 0;
    mlp_pack(I, stream_);
    HIP_OK(hipStreamSynchronize(stream_));
}

void Engine::mlp_get(float *params)
{
    Impl &I = *p_;
    HIP_OK(hipStreamSynchronize(stream_));
    HIP_OK(hipMemcpy(params, I.mlp_p, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
}

void Engine::learn_clear() { p_->ds_rows = 0; }

long long Engine::learn_collect(int nsteps, double dt, long long max_rows)
{
    Impl &I = *p_;
    if (max_rows > I.ds_cap) {
        unsigned short *X = nullptr;
        float *Y = nullptr;
        HIP_OK(hipMalloc(&X, (size_t)max_rows * kMlpInPad * sizeof(unsigned short)));
        HIP_OK(hipMalloc(&Y, (size_t)max_rows * sizeof(float)));
        if (I.ds_rows) {
            HIP_OK(hipMemcpy(X, I.dsX, (size_t)I.ds_rows * kMlpInPad * 2, hipMemcpyDeviceToDevice));
            HIP_OK(hipMemcpy(Y, I.dsY, (size_t)I.ds_rows * sizeof(float), hipMemcpyDeviceToDevice));
        }
        if (I.dsX) { HIP_OK(hipFree(I.dsX)); HIP_OK(hipFree(I.dsY)); }
        I.dsX = X;
        I.dsY = Y;
        I.ds_cap = max_rows;
    }
    if (!I.Qp) alloc_learn_bounds(I, prec_ == PREC_FP64 ? sizeof(double) : sizeof(float));
    Impl::Dump d{I.dsX, I.dsY, I.ds_rows, I.ds_cap};
    I.dump = &d;
    try {
        for (int s = 0; s < nsteps; ++s) enqueue(I, prec_, spec_, dt, 0, ntr_, refresh_ab_, stream_);
        HIP_OK(hipStreamSynchronize(stream_));
    } catch (...) {
        I.dump = nullptr;
        throw;
    }
    I.dump = nullptr;
    I.ds_rows = d.row;
    graph_valid_ = false;
    return I.ds_rows;
}

static void mlp_train_pass(Engine::Impl &I, const unsigned short *X, const float *Y, long long rows, float scale,
                           hipStream_t st)
{
    const long long tiles = (rows + 15) / 16;
    // few blocks, many tiles each: every block adds its gradient with atomics once
    const int grid = (int)std::max(1LL, std::min((tiles + 3) / 4, 256LL));
    const int iters = (int)((tiles + 4LL * grid - 1) / (4LL * grid));
    hipLaunchKernelGGL(k_mlp_train, dim3(grid), dim3(256), 0, st, mlp_dev(I), X, Y, rows, iters, scale, I.mlp_g,
                       I.mlp_loss);
    HIP_OK(hipGetLastError());
}

std::vector<double> Engine::learn_train(int epochs, long long batch, double lr, unsigned seed)
{
    Impl &I = *p_;
    std::vector<double> out;
    if (I.ds_rows == 0 || batch < 16) throw std::runtime_error("no training rows collected (or batch < 16)");
    const long long nb = (I.ds_rows + batch - 1) / batch;
    std::vector<long long> order(nb);
    uint64_t st = 0x2545F4914F6CDD1Dull ^ seed;
    for (int e = 0; e < epochs; ++e) {
        for (long long b = 0; b < nb; ++b) order[b] = b;
        for (long long b = nb - 1; b > 0; --b) {  // minibatches in a seeded random order
            st ^= st << 13; st ^= st >> 7; st ^= st << 17;
            std::swap(order[b], order[st % (uint64_t)(b + 1)]);
        }
        hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(256), 0, stream_, I.mlp_loss, 1LL, 0.0f);
        for (long long b : order) {
            const long long r0 = b * batch, rows = std::min(batch, I.ds_rows - r0);
            hipLaunchKernelGGL(k_fill_f32, dim3((kMlpParams + 255) / 256), dim3(256), 0, stream_, I.mlp_g,
                               (long long)kMlpParams, 0.0f);
            mlp_train_pass(I, I.dsX + r0 * kMlpInPad, I.dsY + r0, rows, 1.0f / (float)rows, stream_);
            I.adam_t += 1;
            const float c1 = 1.0f - std::pow(0.9f, (float)I.adam_t), c2 = 1.0f - std::pow(0.999f, (float)I.adam_t);
            hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, stream_, I.mlp_p, I.mlp_m1,
                               I.mlp_m2, I.mlp_g, (float)lr, c1, c2);
            mlp_pack(I, stream_);
        }
        float loss = 0.0f;
        HIP_OK(hipMemcpyAsync(&loss, I.mlp_loss, sizeof(float), hipMemcpyDeviceToHost, stream_));
        HIP_OK(hipStreamSynchronize(stream_));
        out.push_back((double)loss / (double)I.ds_rows);
    }
    graph_valid_ = false;
    return out;
}

static std::vector<unsigned short> rows_to_bf16(const float *X, long long rows)
{
    std::vector<unsigned short> b((size_t)rows * kMlpInPad, 0);
    for (long long r = 0; r < rows; ++r)
        for (int f = 0; f < kMlpIn; ++f) b[(size_t)r * kMlpInPad + f] = f2bf(X[(size_t)r * kMlpIn + f]);
    return b;
}

void Engine::mlp_forward(const float *X, long long rows, float *out)
{
    Impl &I = *p_;
    if (rows <= 0) return;
    auto b = rows_to_bf16(X, rows);
    unsigned short *dX = dupload(b);
    float *dO = dalloc<float>(rows);
    hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)((rows + 63) / 64)), dim3(256), 0, stream_, mlp_dev(I), dX, rows,
                       dO);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(stream_));
    HIP_OK(hipMemcpy(out, dO, rows * sizeof(float), hipMemcpyDeviceToHost));
    HIP_OK(hipFree(dX));
    HIP_OK(hipFree(dO));
}

void Engine::mlp_gradient(const float *X, const float *y, long long rows, float scale, float *grad, double *loss)
{
    Impl &I = *p_;
    auto b = rows_to_bf16(X, rows);
    unsigned short *dX = dupload(b);
    float *dY = dupload(std::vector<float>(y, y + rows));
    hipLaunchKernelGGL(k_fill_f32, dim3((kMlpParams + 255) / 256), dim3(256), 0, stream_, I.mlp_g,
                       (long long)kMlpParams, 0.0f);
    hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(256), 0, stream_, I.mlp_loss, 1LL, 0.0f);
    mlp_train_pass(I, dX, dY, rows, scale, stream_);
    HIP_OK(hipStreamSynchronize(stream_));
    HIP_OK(hipMemcpy(grad, I.mlp_g, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
    float l = 0.0f;
    HIP_OK(hipMemcpy(&l, I.mlp_loss, sizeof(float), hipMemcpyDeviceToHost));
    *loss = l;
    HIP_OK(hipFree(dX));
    HIP_OK(hipFree(dY));
}

double Engine::learn_fallback(int tr)
{
    Impl &I = *p_;
    if (!I.unsafe || tr < 0 || tr >= ntr_) return 0.0;
    HIP_OK(hipMemsetAsync(I.cnt, 0, 2 * sizeof(unsigned long long), stream_));
    const size_t rs = prec_ == PREC_FP64 ? sizeof(double) : sizeof(float);
    const char *u = (const char *)I.unsafe + (size_t)tr * I.tsn * rs;
    if (prec_ == PREC_FP64)
        hipLaunchKernelGGL((k_count_flags<double>), dim3(256), dim3(256), 0, stream_, I.m64, (const double *)u, I.cnt);
    else
        hipLaunchKernelGGL((k_count_flags<float>), dim3(256), dim3(256), 0, stream_, I.m32, (const float *)u, I.cnt);
    HIP_OK(hipGetLastError());
    unsigned long long h[2] = {0, 0};
    HIP_OK(hipMemcpyAsync(h, I.cnt, sizeof(h), hipMemcpyDeviceToHost, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
    return h[1] ? (double)h[0] / (double)h[1] : 0.0;
}

void Engine::set_gradient_fan(bool on)
{
    p_->grad_fan = on;
    graph_valid_ = false;
}

void Engine::set_schedule(int waves, int max_nt)
{
    if (waves != 1 && waves != 2 && waves != 4 && waves != 8) throw std::runtime_error("waves per block: 1, 2, 4 or 8");
    if (max_nt != 1 && max_nt != 2 && max_nt != 4) throw std::runtime_error("tracers per launch: 1, 2 or 4");
    p_->tune_waves = waves;
    p_->tune_nt = max_nt;
    graph_valid_ = false;
}

void Engine::schedule(int &waves, int &max_nt) const
{
    waves = p_->tune_waves;
    max_nt = p_->tune_nt;
}

std::vector<std::array<double, 3>> Engine::autotune(double dt, int reps)
{
    Impl &I = *p_;
    // the tracers are restored afterwards: the trial steps leave no trace
    const size_t nv = I.tsn * ntr_;
    double *save = dalloc<double>(nv), *save_ab = I.valuesAB ? dalloc<double>(nv) : nullptr;
    HIP_OK(hipStreamSynchronize(stream_));
    HIP_OK(hipMemcpy(save, I.values, nv * sizeof(double), hipMemcpyDeviceToDevice));
    if (save_ab) HIP_OK(hipMemcpy(save_ab, I.valuesAB, nv * sizeof(double), hipMemcpyDeviceToDevice));
    const int w0 = I.tune_waves, n0 = I.tune_nt;
    std::vector<std::array<double, 3>> table;
    double best = 1e300;
    int bw = w0, bn = n0;
    hipEvent_t a, b;
    HIP_OK(hipEventCreate(&a));
    HIP_OK(hipEventCreate(&b));
    try {
        for (int nt : {4, 2, 1}) {
            if (nt > std::max(1, std::min(ntr_, kMaxBatch))) continue;
            for (int w : {2, 4, 8}) {
                set_schedule(w, nt);
                run_steps(1, dt, true);  // capture + warm-up
                HIP_OK(hipEventRecord(a, stream_));
                run_steps(reps, dt, true);
                HIP_OK(hipEventRecord(b, stream_));
                HIP_OK(hipEventSynchronize(b));
                float ms = 0.0f;
                HIP_OK(hipEventElapsedTime(&ms, a, b));
                const double per = (double)ms / std::max(reps, 1);
                table.push_back({(double)w, (double)nt, per});
                if (per < best) { best = per; bw = w; bn = nt; }
            }
        }
    } catch (...) {
        set_schedule(w0, n0);
        (void)hipEventDestroy(a); (void)hipEventDestroy(b);
        (void)hipFree(save); if (save_ab) (void)hipFree(save_ab);
        throw;
    }
    HIP_OK(hipEventDestroy(a));
    HIP_OK(hipEventDestroy(b));
    set_schedule(bw, bn);
    HIP_OK(hipStreamSynchronize(stream_));
    HIP_OK(hipMemcpy(I.values, save, nv * sizeof(double), hipMemcpyDeviceToDevice));
    if (save_ab) HIP_OK(hipMemcpy(I.valuesAB, save_ab, nv * sizeof(double), hipMemcpyDeviceToDevice));
    HIP_OK(hipFree(save));
    if (save_ab) HIP_OK(hipFree(save_ab));
    return table;
}

void Engine::synchronize() { HIP_OK(hipStreamSynchronize(stream_)); }

Diag Engine::diagnostics(int tr)
{
    Impl &I = *p_;
    if (tr < 0 || tr >= ntr_) throw std::runtime_error("tracer index out of range");
    const double *v = I.values + tr * I.tsn;
    const int L = nl_ - 1;
    const double *dense = I.dense + tr * (size_t)L * N_;
    if (prec_ == PREC_FP64)
        hipLaunchKernelGGL((k_diag_partial<double>), dim3(I.ndiag), dim3(256), 0, stream_, I.m64, v, dense, L,
                           I.dpart);
    else
        hipLaunchKernelGGL((k_diag_partial<float>), dim3(I.ndiag), dim3(256), 0, stream_, I.m32, v, dense, L,
                           I.dpart);
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(256), 0, stream_, I.dpart, I.ndiag, I.dout);
    HIP_OK(hipGetLastError());
    DiagAcc a;
    HIP_OK(hipMemcpyAsync(&a, I.dout, sizeof(a), hipMemcpyDeviceToHost, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
    return Diag{a.vmin, a.vmax, a.sum, a.amin, a.amax, a.mass, a.nan};
}

void Engine::difference(const Engine &ref, int tr, double out[4])
{
    Impl &I = *p_;
    const Impl &J = *ref.p_;
    if (tr < 0 || tr >= ntr_ || tr >= ref.ntr_) throw std::runtime_error("tracer index out of range");
    if (I.ln.total != J.ln.total || I.ln.off != J.ln.off) throw std::runtime_error("engines on different meshes");
    HIP_OK(hipStreamSynchronize(ref.stream_));
    const int *slot = prec_ == PREC_FP64 ? I.m64.slot_n : I.m32.slot_n;
    const int *nlev = prec_ == PREC_FP64 ? I.m64.nlev : I.m32.nlev;
    hipLaunchKernelGGL(k_err_partial, dim3(I.ndiag), dim3(256), 0, stream_, slot, nlev, (long long)(I.ln.total / 2),
                       I.values + tr * I.tsn, J.values + tr * J.tsn, I.epart);
    HIP_OK(hipGetLastError());
    std::vector<ErrAcc> h(I.ndiag);
    HIP_OK(hipMemcpyAsync(h.data(), I.epart, h.size() * sizeof(ErrAcc), hipMemcpyDeviceToHost, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
    ErrAcc e{0.0, 0.0, 0.0, 0.0};
    for (const ErrAcc &x : h) {
        e.dinf = std::max(e.dinf, x.dinf); e.d2 += x.d2; e.rinf = std::max(e.rinf, x.rinf); e.r2 += x.r2;
    }
    out[0] = e.dinf;
    out[1] = std::sqrt(e.d2);
    out[2] = e.rinf > 0.0 ? e.dinf / e.rinf : 0.0;
    out[3] = e.r2 > 0.0 ? std::sqrt(e.d2 / e.r2) : 0.0;
}

size_t Engine::work_count(const std::string &name) const
{
    const size_t L = nl_ - 1;
    if (name == "fct_LO" || name == "fct_plus" || name == "fct_minus") return L * N_;
    if (name == "adv_flux_hor") return L * E_;
    if (name == "adv_flux_ver") return (size_t)nl_ * N_;
    if (name == "gnod") return 2 * L * N_;
    return 0;
}

int Engine::get_work(const std::string &name, int tr, double *out, size_t count)
{
    Impl &I = *p_;
    const size_t need = work_count(name);
    if (need == 0 || count < need || tr < 0 || tr >= ntr_) return -1;
    HIP_OK(hipStreamSynchronize(stream_));
    const char *src = nullptr;
    const Layout *lay = &I.ln;
    int L = nl_ - 1, comps = 1;
    size_t es = prec_ == PREC_FP64 ? 8 : 4;
    if (name == "fct_LO") { src = (const char *)I.LO; es = 8; }
    else if (name == "fct_plus") src = (const char *)I.Pp;
    else if (name == "fct_minus") src = (const char *)I.Pm;
    else if (name == "adv_flux_hor") { src = (const char *)I.adf_h; lay = &I.le; }
    else if (name == "adv_flux_ver") { src = (const char *)I.adf_v; L = nl_; }  // the bottom interface stays 0
    else if (name == "gnod") {
        src = (const char *)I.gnod; comps = 2;
        es = prec_ == PREC_FP64 ? 8 : (prec_ == PREC_FP32 ? 4 : 2);
    }
    const int slot = work_slot(spec_, tr, p_->tune_nt);
    const size_t per = (size_t)lay->total * comps;
    // FP32 HO-LO fluxes sit interleaved with the neighbouring work slot's
    // (adf_at): fetch the pair's block and take this slot's values
    if (es == 4 && name == "adv_flux_hor") {
        const size_t tse = (size_t)lay->total;
        std::vector<float> pr(2 * tse);
        HIP_OK(hipMemcpy(pr.data(), (const char *)I.adf_h + (size_t)(slot >> 1) * 2 * tse * es, pr.size() * es,
                         hipMemcpyDeviceToHost));
        std::fill(out, out + need, 0.0);
        for (size_t c = 0; c < lay->len.size(); ++c)
            for (int k = 0; k < std::min(lay->len[c], L); ++k) {
                const size_t i = (size_t)lay->off[c] + k;
                out[(size_t)c * L + k] = pr[adf_at<float>((size_t)(slot & 1), tse, i & ~(size_t)1) + (i & 1)];
            }
        return 0;
    }
    // FP32 R+/R- sit interleaved per cell pair (ld_rpm): fetch the slot's
    // pairs and take every other two values
    const bool rpm32 = es == 4 && (name == "fct_plus" || name == "fct_minus");
    if (rpm32) {
        std::vector<float> pr(2 * per);
        HIP_OK(hipMemcpy(pr.data(), (const char *)I.Pp + slot * 2 * per * es, pr.size() * es,
                         hipMemcpyDeviceToHost));
        const size_t half = name == "fct_plus" ? 0 : 2;
        std::fill(out, out + need, 0.0);
        for (size_t c = 0; c < lay->len.size(); ++c)
            for (int k = 0; k < std::min(lay->len[c], L); ++k) {
                const size_t i = (size_t)lay->off[c] + k;
                // fabs: the learned limiter's bound-check flags ride on R+'s sign bit
                out[(size_t)c * L + k] = std::fabs(pr[2 * (i & ~(size_t)1) + half + (i & 1)]);
            }
        return 0;
    }
    std::vector<char> buf(per * es);
    HIP_OK(hipMemcpy(buf.data(), src + slot * per * es, buf.size(), hipMemcpyDeviceToHost));
    std::fill(out, out + need, 0.0);
    for (size_t c = 0; c < lay->len.size(); ++c)
        for (int k = 0; k < std::min(lay->len[c], L); ++k)
            for (int q = 0; q < comps; ++q) {
                size_t i = ((size_t)lay->off[c] + k) * comps + q;
                double v;
                if (es == 8) v = ((const double *)buf.data())[i];
                else if (es == 4) v = ((const float *)buf.data())[i];
                else v = (double)__half2float(((const __half *)buf.data())[i]) * (double)I.ginv;
                out[((size_t)c * L + k) * comps + q] = name == "fct_plus" ? std::fabs(v) : v;
            }
    return 0;
}

void Engine::time_kernels(double dt, int reps, std::vector<std::string> &names, std::vector<double> &ms)
{
    // one event after every launch of a step on the engine stream; the time
    // of a kernel is the gap to the previous event (launches are serialised
    // on the one stream, so the gap is that kernel's execution)
    names.clear();
    ms.clear();
    for (int r = 0; r < reps; ++r) {
        std::vector<std::pair<std::string, hipEvent_t>> ev;
        LaunchHook hk;
        hk.ev = &ev;
        enqueue(*p_, prec_, spec_, dt, 0, ntr_, refresh_ab_, stream_, hk);
        HIP_OK(hipStreamSynchronize(stream_));
        for (size_t i = 1; i < ev.size(); ++i) {
            if (ev[i].first == "start") continue;
            float e = 0.0f;
            HIP_OK(hipEventElapsedTime(&e, ev[i - 1].second, ev[i].second));
            size_t j = 0;
            while (j < names.size() && names[j] != ev[i].first) ++j;
            if (j == names.size()) { names.push_back(ev[i].first); ms.push_back(0.0); }
            ms[j] += e;
        }
        for (auto &p : ev) HIP_OK(hipEventDestroy(p.second));
    }
    // per step (all tracers), averaged over reps
    for (double &v : ms) v /= (double)reps;
}

}  // namespace fta
