// fta_engine.hip -- the tracer-advection step on gfx950 and its host driver.
//
// One step (reference: do_oce_adv_tra, lib/oce_adv_tra_driver.F90:23-294,
// plus the caller's update, src/fesom_analytic.F90:405-437) for a batch of
// NT tracers that share their schemes is five launches:
//   k_grad        nodal mean of the elemental gradients of valuesAB (MUSCL/MFCT)
//   k_edge        per edge-level: LO upwind flux and HO-LO antidiffusive flux
//   k_node_lo     per node-level: LO solution, vertical HO-LO flux, P+/P- sums
//   k_node_limit  per node-level: Zalesak bounds (a1-a3) and R+/R- (b2)
//   k_node_update per node-level: limited fluxes (b3), divergence, T update
// (without FCT the last three collapse into k_node_nolim). Elemental
// gradients are recomputed from their three nodes where needed, never stored.
#include "fta_engine.h"
#include "fta_kernels.h"
#include "fta_vertical.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>

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

// the elemental gradient of tracer A at layer k (tracer_gradient_elements:
// gradient_sca(1:3)*T(nodes) and gradient_sca(4:6)*T(nodes), node order 1,2,3)
template <class R>
__device__ __forceinline__ void elem_gradient(const ElemMeta &em, const R *g, const double *A, int k, R &gx, R &gy)
{
    const R A0 = (R)A[em.o0 + k], A1 = (R)A[em.o1 + k], A2 = (R)A[em.o2 + k];
    gx = g[0] * A0 + g[1] * A1 + g[2] * A2;
    gy = g[3] * A0 + g[4] * A1 + g[5] * A2;
}

// ===========================================================================
// MUSCL prerequisite: area-weighted mean over the node's elements that have
// layer nz of their elemental gradients (fill_up_dn_grad, lib/oce_muscl_adv.F90:397-407)
// ===========================================================================
// value of neighbourhood slot s (wave-uniform): incident-edge neighbour s, or the node itself
template <class R>
__device__ __forceinline__ R pick(const R (&nb)[kDeg], R self, int s)
{
    R v = self;  // branch-free select chain on the uniform slot
#pragma unroll
    for (int j = 0; j < kDeg; ++j) v = (s == j) ? nb[j] : v;
    return v;
}

// the 1-ring neighbours' values of array A at this level (each distinct
// column once; a neighbour shallower than this level is not read)
template <class R>
__device__ __forceinline__ void load_ring(const NodeEdgeRec &ne, const double *A, int k, R (&nb)[kDeg])
{
    const int nz = k + 1;
#pragma unroll
    for (int j = 0; j < kDeg; ++j) {
        const bool ok = ne.oe[j] >= 0 && nz <= ((ne.info[j] >> 16) & 255);
        nb[j] = ok ? (R)A[(size_t)ne.oo[j] + k] : (R)0;
    }
}

template <class R, class G, int NT>
__global__ __launch_bounds__(kBlock) void k_grad(DMesh<R> m, const double *__restrict__ ab, size_t tsn,
                                                 G *__restrict__ gnod, float gscale)
{
    const int n = wave_column();
    if (n >= m.N) return;
    const int k = lane_id(), nz = k + 1;
    const NodeRec nr = m.nrec[n];
    if (nz > nr.nlev - 1) return;
    const NodeElems ne = m.nelem[n];
    R tvol = 0, tx[NT], ty[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) { tx[t] = 0; ty[t] = 0; }
#pragma unroll
    for (int j = 0; j < kDeg; ++j) {
        const int el = ne.t[j];
        if (el < 0) break;
        const ElemMeta em = m.tmeta[el];
        if (em.nlev - 1 < nz) continue;
        const R *g = m.tgeo + 8 * (size_t)el;
        const R a = g[6];
        tvol = tvol + a;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            R gx, gy;
            elem_gradient(em, g, ab + t * tsn, k, gx, gy);
            tx[t] = tx[t] + gx * a;
            ty[t] = ty[t] + gy * a;
        }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        G *o = gnod + 2 * (t * tsn + nr.off + k);
        GradIO<G>::st(o, tx[t] / tvol, gscale);
        GradIO<G>::st(o + 1, ty[t] / tvol, gscale);
    }
}

// ===========================================================================
// Horizontal edge fluxes (lib/oce_adv_tra_hor.F90): UPW1 (:64-261),
// MUSCL (:265-550), MFCT (:554-846). Segments C (both elements), D (el1
// only) and E (el2 only); no cavities, so segments A/B are empty.
// FCT: writes the LO flux (from ttf) and HO-LO (HO from ttfAB) per level.
// No limiter: writes the HO flux. Gradients follow fill_up_dn_grad
// (lib/oce_muscl_adv.F90:361-530): up/down-wind triangle on levels shared by
// all elements around both nodes, nodal means below.
// ===========================================================================
template <class R, class G, int HOR, bool FCT, int NT, bool SAME>
__global__ __launch_bounds__(kBlock) void k_edge(DMesh<R> m, const R *__restrict__ vel,
                                                 const double *__restrict__ ttf,
                                                 const double *__restrict__ ttfAB, size_t tsn,
                                                 const G *__restrict__ gnod,
                                                 R num_ord, float ginv,
                                                 R *__restrict__ flux_lo, R *__restrict__ flux, size_t tse)
{
    const int e = wave_column();
    if (e >= m.E) return;
    const int k = lane_id(), nz = k + 1;
    const EdgeMeta em = m.emeta[e];
    const int nl1 = em.nl & 255, nl2 = (em.nl >> 8) & 255, nzmax_e = (em.nl >> 16) & 255;
    const int nlmax = (int)((unsigned)em.nl >> 24);
    const int nl12 = min(nl1, nl2);
    if (nz > nlmax) return;
    const R *geo = m.egeo + 8 * (size_t)e;
    const R dx1 = geo[0], dy1 = geo[1], dx2 = geo[2], dy2 = geo[3];
    const bool has1 = nz <= nl1;
    const bool has2 = em.el2 >= 0 && (nz <= nl12 || nz > nl1);
    R vf1 = 0, vf2 = 0;
    if (has1) {
        const R *u = vel + 2 * ((size_t)em.t1 + k);
        vf1 = (-u[1] * dx1 + u[0] * dy1) * m.helem[em.t1 + k];
    }
    if (has2) {
        const R *u = vel + 2 * ((size_t)em.t2 + k);
        vf2 = (u[1] * dx2 - u[0] * dy2) * m.helem[em.t2 + k];
    }
    const R vflux = (has1 && has2) ? vf1 + vf2 : (has1 ? vf1 : vf2);
    const size_t o = (size_t)em.oe + k;
    const size_t o1 = (size_t)em.o1 + k, o2 = (size_t)em.o2 + k;

    // MUSCL/MFCT: reconstruction geometry, gradient source per level
    R ex = 0, ey = 0, a = 0, clo1 = 1, clo2 = 1;
    bool shared = false;
    ElemMeta tu{}, td{};
    const R *gu = nullptr, *gd = nullptr;
    if constexpr (HOR != HOR_UPW1) {
        ex = geo[4]; ey = geo[5]; a = geo[6];
        if constexpr (HOR == HOR_MUSCL) {
            clo1 = ((em.nb & 0xFFFF) - nz) >= 0 ? (R)1 : (R)0;
            clo2 = ((em.nb >> 16) - nz) >= 0 ? (R)1 : (R)0;
        }
        const bool updn = em.up >= 0 && em.dn >= 0;
        shared = updn && nz <= nzmax_e - 1;
        const int up = updn ? em.up : 0, dn = updn ? em.dn : 0;
        tu = m.tmeta[up];
        td = m.tmeta[dn];
        gu = m.tgeo + 8 * (size_t)up;
        gd = m.tgeo + 8 * (size_t)dn;
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        R lo = 0;
        if constexpr (FCT) {
            const R T1 = (R)ttf[t * tsn + o1], T2 = (R)ttf[t * tsn + o2];
            lo = (R)(-0.5) * (T1 * (vflux + fabs(vflux)) + T2 * (vflux - fabs(vflux))) - (R)0;
            flux_lo[t * tse + o] = lo;
        }
        const double *AB = ttfAB + t * tsn;
        R A1, A2;
        if constexpr (FCT && SAME) {
            A1 = (R)ttf[t * tsn + o1];  // the same loads as T1, T2 above
            A2 = (R)ttf[t * tsn + o2];
        } else {
            A1 = (R)AB[o1];
            A2 = (R)AB[o2];
        }
        R ho;
        if constexpr (HOR == HOR_UPW1) {
            ho = (R)(-0.5) * (A1 * (vflux + fabs(vflux)) + A2 * (vflux - fabs(vflux)));
        } else {
            R g1x, g1y, g2x, g2y;
            if (shared) {
                elem_gradient(tu, gu, AB, k, g1x, g1y);
                elem_gradient(td, gd, AB, k, g2x, g2y);
            } else {
                const G *p1 = gnod + 2 * (t * tsn + o1), *p2 = gnod + 2 * (t * tsn + o2);
                g1x = (R)GradIO<G>::ld(p1, ginv); g1y = (R)GradIO<G>::ld(p1 + 1, ginv);
                g2x = (R)GradIO<G>::ld(p2, ginv); g2y = (R)GradIO<G>::ld(p2 + 1, ginv);
            }
            const R Tmean2 = A2 - ((R)2 * (A2 - A1) + ex * a * g2x + ey * (R)kRE * g2y) / (R)6 * clo2;
            const R Tmean1 = A1 + ((R)2 * (A2 - A1) + ex * a * g1x + ey * (R)kRE * g1y) / (R)6 * clo1;
            const R cHO = (vflux + fabs(vflux)) * Tmean1 + (vflux - fabs(vflux)) * Tmean2;
            ho = (R)(-0.5) * ((R)1 - num_ord) * cHO - vflux * num_ord * (R)0.5 * (Tmean1 + Tmean2);
        }
        flux[t * tse + o] = ho - lo;
    }
}

// decoded incident-edge list of a node, valid for this lane's layer
struct NodeEdgeList {
    int oe[kDeg];
    int oo[kDeg];
    bool pos[kDeg];
    bool on[kDeg];
};

template <class R>
__device__ __forceinline__ void load_edge_list(const DMesh<R> &m, int n, int nz, bool layer, NodeEdgeList &L)
{
    const NodeEdgeRec ne = m.nedge[n];
#pragma unroll
    for (int j = 0; j < kDeg; ++j) {
        L.oe[j] = ne.oe[j];
        L.oo[j] = ne.oo[j];
        L.pos[j] = (ne.info[j] & 256) != 0;
        L.on[j] = ne.oe[j] >= 0 && layer && nz <= (ne.info[j] & 255);
    }
}

template <class R>
__device__ __forceinline__ VCol<R> vcol(const DMesh<R> &m, const NodeRec &nr, const double *A, const R *W, int nz)
{
    return VCol<R>{A + nr.off, W + nr.offi, m.area + nr.offi, m.Zn + nr.off, m.zbarn + nr.offi,
                   m.hnode + nr.off, m.hnode_new + nr.off, nr.nlev, nz};
}

// ===========================================================================
// LO solution (lib/oce_adv_tra_driver.F90:95-211), HO-LO vertical flux
// (:257-283), and the raw antidiffusive sums P+/P- (lib/oce_adv_tra_fct.F90:232-346)
// ===========================================================================
template <class R, int VER, int NT, bool SAME>
__global__ __launch_bounds__(kBlock) void k_node_lo(DMesh<R> m, const R *__restrict__ Wv,
                                                    const R *__restrict__ WEv,
                                                    const double *__restrict__ ttf,
                                                    const double *__restrict__ ttfAB, size_t tsn, size_t tsi,
                                                    const R *__restrict__ flux_lo,
                                                    const R *__restrict__ adf_h, size_t tse,
                                                    double dt, R num_ord,
                                                    double *__restrict__ LO, R *__restrict__ adf_v,
                                                    R *__restrict__ Pp, R *__restrict__ Pm)
{
    const int n = wave_column();
    if (n >= m.N) return;
    const int k = lane_id(), nz = k + 1;
    const NodeRec nr = m.nrec[n];
    const int nlev = nr.nlev;
    const bool layer = nz <= nlev - 1;
    const size_t cl = (size_t)nr.off + k, ci = (size_t)nr.offi + k;
    NodeEdgeList L;
    load_edge_list(m, n, nz, layer, L);

#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const size_t tn = t * tsn, te = t * tse;
        // LO horizontal sum over incident edges, edge-index order
        double slo = 0.0;
        R fadf[kDeg];
#pragma unroll
        for (int j = 0; j < kDeg; ++j) {
            fadf[j] = 0;
            if (L.on[j]) {
                const size_t oe = te + (size_t)L.oe[j] + k;
                const double f = (double)flux_lo[oe];
                fadf[j] = adf_h[oe];
                slo = L.pos[j] ? slo + f : slo - f;
            }
        }
        // LO vertical flux (UPW1 with w_e and ttf) and HO-LO with w and ttfAB
        const VCol<R> lc = vcol(m, nr, ttf + tn, WEv, nz);
        const R fv_lo = vflux_upw1(lc, (R)0);
        const R fv_lo_next = shdn(fv_lo);
        const VCol<R> hc = vcol(m, nr, (SAME ? ttf : ttfAB) + tn, Wv, nz);
        const R adfv = ver_flux<R, VER>(hc, num_ord, (R)dt, fv_lo);
        const R adfv_next = shdn(adfv);
        if (nz <= nlev) adf_v[t * tsi + ci] = adfv;
        if (layer) {
            const double T = ttf[tn + cl];
            LO[tn + cl] = (T * (double)m.hnode[cl] +
                           (slo + (double)(fv_lo - fv_lo_next)) * dt / (double)m.areasvol[cl]) /
                          (double)m.hnode_new[cl];
            R pp = (R)0 + (fmx((R)0, adfv) + fmx((R)0, -adfv_next));
            R pm = (R)0 + (fmn((R)0, adfv) + fmn((R)0, -adfv_next));
#pragma unroll
            for (int j = 0; j < kDeg; ++j) {
                if (!L.on[j]) continue;
                const R a = fadf[j];
                if (L.pos[j]) { pp = pp + fmx((R)0, a); pm = pm + fmn((R)0, a); }
                else          { pp = pp + fmx((R)0, -a); pm = pm + fmn((R)0, -a); }
            }
            Pp[tn + cl] = pp;
            Pm[tn + cl] = pm;
        }
    }
}

// Zalesak bounds a1-a3 and limiting factors b2 (lib/oce_adv_tra_fct.F90:94-217,361-374);
// R+/R- overwrite P+/P- in place. a1 (max/min of LO and T) is formed once per
// distinct node of the 1-ring, a2 per element from its three nodes' a1.
template <class R, int NT>
__global__ __launch_bounds__(kBlock) void k_node_limit(DMesh<R> m, const double *__restrict__ ttf,
                                                       const double *__restrict__ LO, size_t tsn,
                                                       double dt, R *__restrict__ Pp_Rp,
                                                       R *__restrict__ Pm_Rm)
{
    const int n = wave_column();
    if (n >= m.N) return;
    const int k = lane_id(), nz = k + 1;
    const NodeRec nr = m.nrec[n];
    const int nlev = nr.nlev;
    const bool layer = nz <= nlev - 1;
    const size_t cl = (size_t)nr.off + k;
    const R bignumber = (R)1000;

    R tvmax[NT], tvmin[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) { tvmax[t] = 0; tvmin[t] = 0; }
    if (layer) {
        const NodeEdgeRec ne = m.nedge[n];
        const NodeElems nel = m.nelem[n];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const size_t tn = t * tsn;
            R lnb[kDeg], tnb[kDeg], mxn[kDeg], mnn[kDeg];
            load_ring<R>(ne, LO + tn, k, lnb);
            load_ring<R>(ne, ttf + tn, k, tnb);
            const R ls = (R)LO[tn + cl], ts = (R)ttf[tn + cl];
            const R mxs = fmx(ls, ts), mns = fmn(ls, ts);
#pragma unroll
            for (int j = 0; j < kDeg; ++j) { mxn[j] = fmx(lnb[j], tnb[j]); mnn[j] = fmn(lnb[j], tnb[j]); }
            R vmax = 0, vmin = 0;
#pragma unroll
            for (int j = 0; j < kDeg; ++j) {
                if (nel.t[j] < 0) break;
                const int sl = nel.sl[j];
                R emax = -bignumber, emin = bignumber;
                if (nz <= (sl >> 16) - 1) {
                    const int s0 = sl & 15, s1 = (sl >> 4) & 15, s2 = (sl >> 8) & 15;
                    emax = fmx(fmx(pick(mxn, mxs, s0), pick(mxn, mxs, s1)), pick(mxn, mxs, s2));
                    emin = fmn(fmn(pick(mnn, mns, s0), pick(mnn, mns, s1)), pick(mnn, mns, s2));
                }
                vmax = (j == 0) ? emax : fmx(vmax, emax);
                vmin = (j == 0) ? emin : fmn(vmin, emin);
            }
            tvmax[t] = vmax;
            tvmin[t] = vmin;
        }
    }
    const R flux_eps = (R)1e-16;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const R up_max = shup(tvmax[t]), dn_max = shdn(tvmax[t]);
        const R up_min = shup(tvmin[t]), dn_min = shdn(tvmin[t]);
        if (!layer) continue;
        const size_t tn = t * tsn;
        const R lo = (R)LO[tn + cl];
        R qp, qm;
        if (nz == 1 || nz == nlev - 1) {
            qp = tvmax[t] - lo;
            qm = tvmin[t] - lo;
        } else {
            qp = fmx(fmx(up_max, tvmax[t]), dn_max) - lo;
            qm = fmn(fmn(up_min, tvmin[t]), dn_min) - lo;
        }
        const R av = m.areasvol[cl];
        R flux = Pp_Rp[tn + cl] * (R)dt / av + flux_eps;
        const R rp = fmn((R)1, qp / flux);
        flux = Pm_Rm[tn + cl] * (R)dt / av - flux_eps;
        const R rm = fmn((R)1, qm / flux);
        Pp_Rp[tn + cl] = rp;
        Pm_Rm[tn + cl] = rm;
    }
}

// limiting b3 (lib/oce_adv_tra_fct.F90:394-471), flux divergence
// (lib/oce_adv_tra_driver.F90:298-450) and the update
// values += (dttf_h + dttf_v)/hnode_new (src/fesom_analytic.F90:419-437)
template <class R, int NT>
__global__ __launch_bounds__(kBlock) void k_node_update(DMesh<R> m, double *__restrict__ ttf,
                                                        const double *__restrict__ LO, size_t tsn, size_t tsi,
                                                        const R *__restrict__ adf_h, size_t tse,
                                                        const R *__restrict__ adf_v,
                                                        const R *__restrict__ Rp,
                                                        const R *__restrict__ Rm, double dt)
{
    const int n = wave_column();
    if (n >= m.N) return;
    const int k = lane_id(), nz = k + 1;
    const NodeRec nr = m.nrec[n];
    const int nlev = nr.nlev;
    const bool layer = nz <= nlev - 1;
    const size_t cl = (size_t)nr.off + k, ci = (size_t)nr.offi + k;
    NodeEdgeList L;
    load_edge_list(m, n, nz, layer, L);

#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const size_t tn = t * tsn, te = t * tse;
        R rp = 0, rm = 0, fv = 0;
        if (nz <= nlev) fv = adf_v[t * tsi + ci];
        if (layer) { rp = Rp[tn + cl]; rm = Rm[tn + cl]; }
        const R rp_up = shup(rp), rm_up = shup(rm);
        R fvl = fv;
        if (nz == 1) {
            R ae = (R)1;
            ae = fv >= (R)0 ? fmn(ae, rp) : fmn(ae, rm);
            fvl = ae * fv;
        } else if (layer) {
            R ae = (R)1;
            if (fv >= (R)0) { ae = fmn(ae, rm_up); ae = fmn(ae, rp); }
            else            { ae = fmn(ae, rp_up); ae = fmn(ae, rm); }
            fvl = ae * fv;
        }
        const R fvl_next = shdn(fvl);
        if (!layer) continue;

        const double av = (double)m.areasvol[cl];
        double sh = 0.0;
#pragma unroll
        for (int j = 0; j < kDeg; ++j) {
            if (!L.on[j]) continue;
            const R f = adf_h[te + (size_t)L.oe[j] + k];
            const size_t oo = tn + (size_t)L.oo[j] + k;
            const R rpo = Rp[oo], rmo = Rm[oo];
            // edge orientation: this node is edges(1) when pos
            const R rp1 = L.pos[j] ? rp : rpo, rm1 = L.pos[j] ? rm : rmo;
            const R rp2 = L.pos[j] ? rpo : rp, rm2 = L.pos[j] ? rmo : rm;
            R ae = (R)1;
            if (f >= (R)0) { ae = fmn(ae, rp1); ae = fmn(ae, rm2); }
            else           { ae = fmn(ae, rm1); ae = fmn(ae, rp2); }
            const double fl = (double)(ae * f);
            sh = L.pos[j] ? sh + fl * dt / av : sh - fl * dt / av;
        }
        const double T = ttf[tn + cl];
        double dv = 0.0 - T * (double)m.hnode[cl] + LO[tn + cl] * (double)m.hnode_new[cl];
        dv = dv + (double)(fvl - fvl_next) * dt / av;
        const double del = (0.0 + sh) + dv;
        ttf[tn + cl] = T + del / (double)m.hnode_new[cl];
    }
}

// lim = 'NONE': HO fluxes straight into the divergence (driver :291)
template <class R, int VER, int NT>
__global__ __launch_bounds__(kBlock) void k_node_nolim(DMesh<R> m, const R *__restrict__ WEv,
                                                       double *__restrict__ ttf,
                                                       const double *__restrict__ ttfAB, size_t tsn,
                                                       const R *__restrict__ flux_h, size_t tse,
                                                       double dt, R num_ord)
{
    const int n = wave_column();
    if (n >= m.N) return;
    const int k = lane_id(), nz = k + 1;
    const NodeRec nr = m.nrec[n];
    const int nlev = nr.nlev;
    const bool layer = nz <= nlev - 1;
    const size_t cl = (size_t)nr.off + k;
    NodeEdgeList L;
    load_edge_list(m, n, nz, layer, L);
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const size_t tn = t * tsn, te = t * tse;
        const VCol<R> hc = vcol(m, nr, ttfAB + tn, WEv, nz);
        const R fv = ver_flux<R, VER>(hc, num_ord, (R)dt, (R)0);
        const R fv_next = shdn(fv);
        if (!layer) continue;
        const double av = (double)m.areasvol[cl];
        double sh = 0.0;
#pragma unroll
        for (int j = 0; j < kDeg; ++j) {
            if (!L.on[j]) continue;
            const double f = (double)flux_h[te + (size_t)L.oe[j] + k];
            sh = L.pos[j] ? sh + f * dt / av : sh - f * dt / av;
        }
        const double dv = 0.0 + (double)(fv - fv_next) * dt / av;
        const double del = (0.0 + sh) + dv;
        const double T = ttf[tn + cl];
        ttf[tn + cl] = T + del / (double)m.hnode_new[cl];
    }
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

// reduction over the wet cells (compact layout): min, max, sum, mass, NaN count
template <class R>
__global__ __launch_bounds__(256) void k_diag_partial(DMesh<R> m, const double *__restrict__ v,
                                                      DiagAcc *__restrict__ part)
{
    __shared__ DiagAcc sh[256];
    DiagAcc a{INFINITY, -INFINITY, 0.0, INFINITY, -INFINITY, 0.0, 0.0};
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < m.CN; i += (long long)gridDim.x * 256) {
        const double x = v[i];
        if (x != x) { a.nan += 1.0; continue; }
        a.vmin = fmin(a.vmin, x); a.vmax = fmax(a.vmax, x); a.sum += x;
        a.mass += x * (double)m.hnode_new[i] * (double)m.areasvol[i];
    }
    a.amin = a.vmin;
    a.amax = a.vmax;
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

inline unsigned grid_for(int ncols)
{
    unsigned nb = (unsigned)((ncols + kWavesPerBlock - 1) / kWavesPerBlock);
    return (nb + 7u) & ~7u;  // multiple of 8 for the XCD remap
}

}  // namespace

// compact layout of a set of columns: column c holds len[c] levels at off[c]
struct Layout {
    std::vector<long long> off;
    std::vector<int> len;
    long long total = 0;
    void build(const std::vector<int> &lens)
    {
        len = lens;
        off.assign(lens.size() + 1, 0);
        for (size_t c = 0; c < lens.size(); ++c) off[c + 1] = off[c] + lens[c];
        total = off.back();
    }
    // unpadded column-major (L, ncol) with comps interleaved components -> compact
    template <class T>
    std::vector<T> pack(const double *src, int L, int comps = 1) const
    {
        std::vector<T> out((size_t)total * comps);
        for (size_t c = 0; c < len.size(); ++c)
            for (int k = 0; k < len[c]; ++k)
                for (int q = 0; q < comps; ++q)
                    out[((size_t)off[c] + k) * comps + q] = (T)src[((size_t)c * L + k) * comps + q];
        return out;
    }
};

struct Engine::Impl {
    std::vector<void *> allocs;
    template <class T> T *track(T *p) { allocs.push_back((void *)p); return p; }

    DMesh<double> m64{};
    DMesh<float> m32{};
    Layout ln, li, lt, le;  // node layers, node interfaces, element layers, edge levels
    // velocity
    void *vel = nullptr, *W = nullptr, *WE = nullptr;
    // tracers [ntr][CN]
    double *values = nullptr, *valuesAB = nullptr;
    // work, per tracer
    void *flux_lo = nullptr, *adf_h = nullptr;             // [ntr][CE]
    void *adf_v = nullptr;                                 // [ntr][CI]
    void *Pp = nullptr, *Pm = nullptr;                     // [ntr][CN]
    double *LO = nullptr;                                  // [ntr][CN]
    void *gnod = nullptr;                                  // [ntr][2*CN]
    DiagAcc *dpart = nullptr, *dout = nullptr;
    int ndiag = 1024;
    float gscale = 1.0f, ginv = 1.0f;
    size_t tsn = 0, tsi = 0, tse = 0;
    // host copy of each tracer as last set (levels below the bottom are never
    // touched by a step, so get_values and the all-values statistics read them here)
    std::vector<std::vector<double>> host_vals;
};

template <class R>
static void build_dmesh(Engine::Impl &I, DMesh<R> &d, const HostMesh &m)
{
    const int N = m.nod2D, T = m.elem2D, E = m.edge2D, nl = m.nl;
    d.N = N; d.T = T; d.E = E; d.nl = nl;
    d.CN = I.ln.total; d.CI = I.li.total; d.CT = I.lt.total; d.CE = I.le.total;
    std::vector<EdgeMeta> em(E);
    std::vector<R> eg(8 * (size_t)E, (R)0);
    for (int e = 0; e < E; ++e) {
        EdgeMeta &r = em[e];
        std::memset(&r, 0, sizeof(r));
        const int n1 = m.edges[2 * e], n2 = m.edges[2 * e + 1];
        const int el1 = m.edge_tri[2 * e], el2 = m.edge_tri[2 * e + 1];
        r.o1 = (int)I.ln.off[n1]; r.o2 = (int)I.ln.off[n2];
        r.t1 = (int)I.lt.off[el1]; r.t2 = el2 >= 0 ? (int)I.lt.off[el2] : -1;
        r.up = m.edge_up_dn_tri[2 * e]; r.dn = m.edge_up_dn_tri[2 * e + 1];
        const int nl1 = m.nlevels[el1] - 1, nl2 = el2 >= 0 ? m.nlevels[el2] - 1 : 0;
        const int nzmax_e = std::min(m.nlevels_nod2D_min[n1], m.nlevels_nod2D_min[n2]);
        const int nlmax = std::max(nl1, nl2);
        r.nl = nl1 | (nl2 << 8) | (nzmax_e << 16) | (nlmax << 24);
        r.nb = (m.nboundary_lay[n1] & 0xFFFF) | ((m.nboundary_lay[n2] & 0xFFFF) << 16);
        r.oe = (int)I.le.off[e];
        r.el2 = el2;
        double a = kRE * m.elem_cos[el1];
        if (el2 >= 0) a = 0.5 * (a + kRE * m.elem_cos[el2]);
        R *g = &eg[8 * (size_t)e];
        for (int q = 0; q < 4; ++q) g[q] = (R)m.edge_cross_dxdy[4 * e + q];
        g[4] = (R)m.edge_dxdy[2 * e];
        g[5] = (R)m.edge_dxdy[2 * e + 1];
        g[6] = (R)a;
    }
    d.emeta = I.track(dupload(em));
    d.egeo = I.track(dupload(eg));
    std::vector<ElemMeta> tm(T);
    std::vector<R> tg(8 * (size_t)T, (R)0);
    for (int el = 0; el < T; ++el) {
        const int *en = &m.elem2D_nodes[3 * el];
        tm[el] = ElemMeta{(int)I.ln.off[en[0]], (int)I.ln.off[en[1]], (int)I.ln.off[en[2]], m.nlevels[el]};
        for (int q = 0; q < 6; ++q) tg[8 * (size_t)el + q] = (R)m.gradient_sca[6 * el + q];
        tg[8 * (size_t)el + 6] = (R)m.elem_area[el];
    }
    d.tmeta = I.track(dupload(tm));
    d.tgeo = I.track(dupload(tg));
    std::vector<NodeRec> nr(N);
    std::vector<NodeEdgeRec> ne(N);
    std::vector<NodeElems> nt(N);
    for (int n = 0; n < N; ++n) {
        nr[n] = NodeRec{m.nlevels_nod2D[n], (int)I.ln.off[n], (int)I.li.off[n], 0};
        const int deg = m.n2e_ptr[n + 1] - m.n2e_ptr[n];
        if (deg > kDeg || m.nod_in_elem2D_num[n] > kDeg)
            throw std::runtime_error("node degree above " + std::to_string(kDeg) + " is not supported");
        int nbr[kDeg];
        for (int j = 0; j < kDeg; ++j) {
            ne[n].oe[j] = -1;
            ne[n].info[j] = 0;
            ne[n].oo[j] = 0;
            nt[n].t[j] = -1;
            nt[n].sl[j] = 0;
            nbr[j] = -1;
        }
        for (int j = 0; j < deg; ++j) {
            const int se = m.n2e[m.n2e_ptr[n] + j];
            const int e = (se > 0 ? se : -se) - 1;
            ne[n].oe[j] = (int)I.le.off[e];
            const int other = se > 0 ? m.edges[2 * e + 1] : m.edges[2 * e];
            ne[n].info[j] = ((unsigned)em[e].nl >> 24) | (se > 0 ? 256 : 0) | ((m.nlevels_nod2D[other] - 1) << 16);
            ne[n].oo[j] = (int)I.ln.off[other];
            nbr[j] = other;
        }
        for (int j = 0; j < m.nod_in_elem2D_num[n]; ++j) {
            const int el = m.nod_in_elem2D[(size_t)n * m.maxAdj + j];
            nt[n].t[j] = el;
            int slots = 0;
            for (int q = 0; q < 3; ++q) {
                const int node = m.elem2D_nodes[3 * el + q];
                int slot = -1;
                if (node == n) slot = kSelf;
                else
                    for (int jj = 0; jj < deg; ++jj)
                        if (nbr[jj] == node) slot = jj;
                if (slot < 0) throw std::runtime_error("element node is not a 1-ring neighbour");
                slots |= slot << (4 * q);
            }
            nt[n].sl[j] = slots | (m.nlevels[el] << 16);
        }
    }
    d.nrec = I.track(dupload(nr));
    d.nedge = I.track(dupload(ne));
    d.nelem = I.track(dupload(nt));
    d.area = I.track(dupload(I.li.pack<R>(m.area.data(), nl)));
    d.zbarn = I.track(dupload(I.li.pack<R>(m.zbar_3d_n.data(), nl)));
    d.areasvol = I.track(dupload(I.ln.pack<R>(m.areasvol.data(), nl)));
    d.hnode = I.track(dupload(I.ln.pack<R>(m.hnode.data(), nl - 1)));
    d.hnode_new = I.track(dupload(I.ln.pack<R>(m.hnode_new.data(), nl - 1)));
    d.Zn = I.track(dupload(I.ln.pack<R>(m.Z_3d_n.data(), nl - 1)));
    d.helem = I.track(dupload(I.lt.pack<R>(m.helem.data(), nl - 1)));
}

Engine::Engine(const HostMesh &m, int ntracers, int precision)
    : p_(new Impl), N_(m.nod2D), T_(m.elem2D), E_(m.edge2D), nl_(m.nl), ntr_(ntracers), prec_(precision)
{
    if (nl_ > kWave) throw std::runtime_error("the device path supports nl <= 64 level interfaces");
    if (ntracers < 1) throw std::runtime_error("need at least one tracer");
    if (precision < PREC_FP64 || precision > PREC_FP32_H16) throw std::runtime_error("unknown precision");
    for (int u : m.ulevels) if (u != 1) throw std::runtime_error("cavities (ulevels > 1) are not supported");
    S_ = 0;
    spec_.assign(ntr_, TracerSpec());
    Impl &I = *p_;
    {
        std::vector<int> a(N_), b(N_), c(T_), d(E_);
        for (int n = 0; n < N_; ++n) { a[n] = m.nlevels_nod2D[n] - 1; b[n] = m.nlevels_nod2D[n]; }
        for (int el = 0; el < T_; ++el) c[el] = m.nlevels[el] - 1;
        for (int e = 0; e < E_; ++e) {
            int el1 = m.edge_tri[2 * e], el2 = m.edge_tri[2 * e + 1];
            d[e] = std::max(m.nlevels[el1] - 1, el2 >= 0 ? m.nlevels[el2] - 1 : 0);
        }
        I.ln.build(a); I.li.build(b); I.lt.build(c); I.le.build(d);
        if (I.le.total >= (1ll << 31) || I.li.total >= (1ll << 31))
            throw std::runtime_error("mesh too large for 32-bit column offsets");
    }
    HIP_OK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    const size_t rs = prec_ == PREC_FP64 ? sizeof(double) : sizeof(float);
    const size_t gs = prec_ == PREC_FP64 ? sizeof(double) : (prec_ == PREC_FP32 ? sizeof(float) : sizeof(__half));
    if (prec_ == PREC_FP64) build_dmesh(I, I.m64, m);
    else build_dmesh(I, I.m32, m);
    I.tsn = (size_t)I.ln.total;
    I.tsi = (size_t)I.li.total;
    I.tse = (size_t)I.le.total;
    auto zalloc = [&](size_t bytes) {
        void *p = I.track((void *)dalloc<char>(bytes));
        HIP_OK(hipMemset(p, 0, bytes));
        return p;
    };
    I.vel = zalloc(2 * (size_t)I.lt.total * rs);
    I.W = zalloc(I.tsi * rs);
    I.WE = zalloc(I.tsi * rs);
    I.values = (double *)zalloc(I.tsn * ntr_ * sizeof(double));
    I.valuesAB = (double *)zalloc(I.tsn * ntr_ * sizeof(double));
    I.flux_lo = zalloc(I.tse * ntr_ * rs);
    I.adf_h = zalloc(I.tse * ntr_ * rs);
    I.adf_v = zalloc(I.tsi * ntr_ * rs);
    I.Pp = zalloc(I.tsn * ntr_ * rs);
    I.Pm = zalloc(I.tsn * ntr_ * rs);
    I.LO = (double *)zalloc(I.tsn * ntr_ * sizeof(double));
    I.gnod = zalloc(2 * I.tsn * ntr_ * gs);
    I.dpart = I.track(dalloc<DiagAcc>(I.ndiag));
    I.dout = I.track(dalloc<DiagAcc>(1));
    I.host_vals.assign(ntr_, std::vector<double>((size_t)(nl_ - 1) * N_, 0.0));
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
    if (stream_) (void)hipStreamDestroy(stream_);
    delete p_;
}

void Engine::set_velocity(const double *uv, const double *w, const double *we)
{
    Impl &I = *p_;
    auto put = [&](void *dst, const Layout &lay, const double *src, int L, int comps) {
        if (prec_ == PREC_FP64) {
            auto v = lay.pack<double>(src, L, comps);
            HIP_OK(hipMemcpy(dst, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
        } else {
            auto v = lay.pack<float>(src, L, comps);
            HIP_OK(hipMemcpy(dst, v.data(), v.size() * sizeof(float), hipMemcpyHostToDevice));
        }
    };
    put(I.vel, I.lt, uv, nl_ - 1, 2);
    put(I.W, I.li, w, nl_, 1);
    put(I.WE, I.li, we, nl_, 1);
}

void Engine::set_values(int tr, const double *v)
{
    if (tr < 0 || tr >= ntr_) throw std::runtime_error("tracer index out of range");
    Impl &I = *p_;
    std::memcpy(I.host_vals[tr].data(), v, sizeof(double) * I.host_vals[tr].size());
    auto a = I.ln.pack<double>(v, nl_ - 1);
    HIP_OK(hipMemcpy(I.values + tr * I.tsn, a.data(), a.size() * sizeof(double), hipMemcpyHostToDevice));
}

void Engine::set_valuesAB(int tr, const double *v)
{
    if (tr < 0 || tr >= ntr_) throw std::runtime_error("tracer index out of range");
    Impl &I = *p_;
    auto a = I.ln.pack<double>(v, nl_ - 1);
    HIP_OK(hipMemcpy(I.valuesAB + tr * I.tsn, a.data(), a.size() * sizeof(double), hipMemcpyHostToDevice));
}

void Engine::get_values(int tr, double *v)
{
    if (tr < 0 || tr >= ntr_) throw std::runtime_error("tracer index out of range");
    Impl &I = *p_;
    std::vector<double> a(I.tsn);
    HIP_OK(hipStreamSynchronize(stream_));
    HIP_OK(hipMemcpy(a.data(), I.values + tr * I.tsn, a.size() * sizeof(double), hipMemcpyDeviceToHost));
    const int L = nl_ - 1;
    std::memcpy(v, I.host_vals[tr].data(), sizeof(double) * I.host_vals[tr].size());
    for (int n = 0; n < N_; ++n)
        for (int k = 0; k < I.ln.len[n]; ++k) v[(size_t)n * L + k] = a[(size_t)I.ln.off[n] + k];
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

// one step of the NT tracers [t0, t0+NT), all with schemes sp
template <class R, class G, int NT>
static void launch_batch(Engine::Impl &I, const DMesh<R> &m, const TracerSpec &sp, double dt, int t0, bool refresh,
                         hipStream_t st, LaunchHook &hk)
{
    const R *vel = (const R *)I.vel, *W = (const R *)I.W, *WE = (const R *)I.WE;
    const size_t tsn = I.tsn, tsi = I.tsi, tse = I.tse;
    double *vals = I.values + t0 * tsn;
    const double *ab = refresh ? vals : I.valuesAB + t0 * tsn;
    R *flux_lo = (R *)I.flux_lo + t0 * tse, *adf_h = (R *)I.adf_h + t0 * tse;
    R *adf_v = (R *)I.adf_v + t0 * tsi, *Pp = (R *)I.Pp + t0 * tsn, *Pm = (R *)I.Pm + t0 * tsn;
    double *LO = I.LO + t0 * tsn;
    G *gnod = (G *)I.gnod + 2 * t0 * tsn;
    const dim3 blk(kBlock), ge(grid_for(m.E)), gn(grid_for(m.N));
    const bool fct = sp.lim == LIM_FCT;
    const R ph = (R)sp.ph, pv = (R)sp.pv;
    if (sp.hor == HOR_MUSCL || sp.hor == HOR_MFCT) {
        hipLaunchKernelGGL((k_grad<R, G, NT>), gn, blk, 0, st, m, ab, tsn, gnod, I.gscale);
        hk.mark("k_grad", st);
    }
#define EDGE(HOR, F)                                                                                        \
    do {                                                                                                    \
        if (refresh)                                                                                        \
            hipLaunchKernelGGL((k_edge<R, G, HOR, F, NT, true>), ge, blk, 0, st, m, vel, vals, ab, tsn, gnod, ph, \
                               I.ginv, flux_lo, adf_h, tse);                                                \
        else                                                                                                \
            hipLaunchKernelGGL((k_edge<R, G, HOR, F, NT, false>), ge, blk, 0, st, m, vel, vals, ab, tsn, gnod, ph, \
                               I.ginv, flux_lo, adf_h, tse);                                                \
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
            hipLaunchKernelGGL((k_node_lo<R, V, NT, true>), gn, blk, 0, st, m, W, WE, vals, ab, tsn, tsi, flux_lo, \
                               adf_h, tse, dt, pv, LO, adf_v, Pp, Pm);                                      \
        else                                                                                                \
            hipLaunchKernelGGL((k_node_lo<R, V, NT, false>), gn, blk, 0, st, m, W, WE, vals, ab, tsn, tsi, flux_lo, \
                               adf_h, tse, dt, pv, LO, adf_v, Pp, Pm);                                      \
    } while (0)
        switch (sp.ver) {
        case VER_QR4C: NLO(VER_QR4C); break;
        case VER_PPM: NLO(VER_PPM); break;
        case VER_CDIFF: NLO(VER_CDIFF); break;
        default: NLO(VER_UPW1); break;
        }
#undef NLO
        hk.mark("k_node_lo", st);
        hipLaunchKernelGGL((k_node_limit<R, NT>), gn, blk, 0, st, m, vals, LO, tsn, dt, Pp, Pm);
        hk.mark("k_node_limit", st);
        hipLaunchKernelGGL((k_node_update<R, NT>), gn, blk, 0, st, m, vals, LO, tsn, tsi, adf_h, tse, adf_v, Pp, Pm, dt);
        hk.mark("k_node_update", st);
    } else {
#define NNL(V) hipLaunchKernelGGL((k_node_nolim<R, V, NT>), gn, blk, 0, st, m, WE, vals, ab, tsn, adf_h, tse, dt, pv)
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
            if (run >= 4) { launch_batch<R, G, 4>(I, m, spec[t], dt, t, refresh, st, hk); t += 4; run -= 4; }
            else if (run >= 2) { launch_batch<R, G, 2>(I, m, spec[t], dt, t, refresh, st, hk); t += 2; run -= 2; }
            else { launch_batch<R, G, 1>(I, m, spec[t], dt, t, refresh, st, hk); t += 1; run -= 1; }
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

void Engine::synchronize() { HIP_OK(hipStreamSynchronize(stream_)); }

Diag Engine::diagnostics(int tr)
{
    Impl &I = *p_;
    if (tr < 0 || tr >= ntr_) throw std::runtime_error("tracer index out of range");
    const double *v = I.values + tr * I.tsn;
    if (prec_ == PREC_FP64)
        hipLaunchKernelGGL((k_diag_partial<double>), dim3(I.ndiag), dim3(256), 0, stream_, I.m64, v, I.dpart);
    else
        hipLaunchKernelGGL((k_diag_partial<float>), dim3(I.ndiag), dim3(256), 0, stream_, I.m32, v, I.dpart);
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(256), 0, stream_, I.dpart, I.ndiag, I.dout);
    HIP_OK(hipGetLastError());
    DiagAcc a;
    HIP_OK(hipMemcpyAsync(&a, I.dout, sizeof(a), hipMemcpyDeviceToHost, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
    // the levels below the bottom (never stepped) from the host copy
    Diag d{a.vmin, a.vmax, a.sum, a.amin, a.amax, a.mass, a.nan};
    const int L = nl_ - 1;
    const std::vector<double> &h = I.host_vals[tr];
    for (int n = 0; n < N_; ++n)
        for (int k = I.ln.len[n]; k < L; ++k) {
            const double x = h[(size_t)n * L + k];
            if (x != x) { d.nan_count += 1.0; continue; }
            d.vmin = std::min(d.vmin, x);
            d.vmax = std::max(d.vmax, x);
            d.sum += x;
        }
    return d;
}

size_t Engine::work_count(const std::string &name) const
{
    const size_t L = nl_ - 1;
    if (name == "fct_LO" || name == "fct_plus" || name == "fct_minus") return L * N_;
    if (name == "adv_flux_hor" || name == "adv_flux_lo") return L * E_;
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
    else if (name == "adv_flux_lo") { src = (const char *)I.flux_lo; lay = &I.le; }
    else if (name == "adv_flux_ver") { src = (const char *)I.adf_v; lay = &I.li; L = nl_; }
    else if (name == "gnod") {
        src = (const char *)I.gnod; comps = 2;
        es = prec_ == PREC_FP64 ? 8 : (prec_ == PREC_FP32 ? 4 : 2);
    }
    const size_t per = (size_t)lay->total * comps;
    std::vector<char> buf(per * es);
    HIP_OK(hipMemcpy(buf.data(), src + tr * per * es, buf.size(), hipMemcpyDeviceToHost));
    std::fill(out, out + need, 0.0);
    for (size_t c = 0; c < lay->len.size(); ++c)
        for (int k = 0; k < lay->len[c]; ++k)
            for (int q = 0; q < comps; ++q) {
                size_t i = ((size_t)lay->off[c] + k) * comps + q;
                double v;
                if (es == 8) v = ((const double *)buf.data())[i];
                else if (es == 4) v = ((const float *)buf.data())[i];
                else v = (double)__half2float(((const __half *)buf.data())[i]) * (double)I.ginv;
                out[((size_t)c * L + k) * comps + q] = v;
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
