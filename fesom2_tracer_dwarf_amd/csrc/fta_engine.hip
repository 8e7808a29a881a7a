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
__device__ __forceinline__ void elem_gradient(const ElemMeta &em, const R *g, const double *A, int S, int k,
                                              R &gx, R &gy)
{
    const R A0 = (R)A[(size_t)em.n0 * S + k], A1 = (R)A[(size_t)em.n1 * S + k], A2 = (R)A[(size_t)em.n2 * S + k];
    gx = g[0] * A0 + g[1] * A1 + g[2] * A2;
    gy = g[3] * A0 + g[4] * A1 + g[5] * A2;
}

// ===========================================================================
// MUSCL prerequisite: area-weighted mean over the node's elements that have
// layer nz of their elemental gradients (fill_up_dn_grad, lib/oce_muscl_adv.F90:397-407)
// ===========================================================================
template <class R, class G, int NT>
__global__ __launch_bounds__(kBlock) void k_grad(DMesh<R> m, const double *__restrict__ ab, size_t tsn,
                                                 G *__restrict__ gnod, size_t tsg, float gscale)
{
    const int n = wave_column();
    if (n >= m.N) return;
    const int k = lane_id(), nz = k + 1;
    if (nz > m.nlev_n[n] - 1) return;
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
            elem_gradient(em, g, ab + t * tsn, m.S, k, gx, gy);
            tx[t] = tx[t] + gx * a;
            ty[t] = ty[t] + gy * a;
        }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        G *o = gnod + t * tsg + ((size_t)n * m.S + k) * 2;
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
template <class R, class G, int HOR, bool FCT, int NT>
__global__ __launch_bounds__(kBlock) void k_edge(DMesh<R> m, const R *__restrict__ vel,
                                                 const double *__restrict__ ttf,
                                                 const double *__restrict__ ttfAB, size_t tsn,
                                                 const G *__restrict__ gnod, size_t tsg,
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
        const R *u = vel + ((size_t)em.el1 * m.S + k) * 2;
        vf1 = (-u[1] * dx1 + u[0] * dy1) * m.helem[(size_t)em.el1 * m.S + k];
    }
    if (has2) {
        const R *u = vel + ((size_t)em.el2 * m.S + k) * 2;
        vf2 = (u[1] * dx2 - u[0] * dy2) * m.helem[(size_t)em.el2 * m.S + k];
    }
    const R vflux = (has1 && has2) ? vf1 + vf2 : (has1 ? vf1 : vf2);
    const size_t o = (size_t)e * m.S + k;
    const size_t o1 = (size_t)em.n1 * m.S + k, o2 = (size_t)em.n2 * m.S + k;

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
        const R A1 = (R)AB[o1], A2 = (R)AB[o2];
        R ho;
        if constexpr (HOR == HOR_UPW1) {
            ho = (R)(-0.5) * (A1 * (vflux + fabs(vflux)) + A2 * (vflux - fabs(vflux)));
        } else {
            R g1x, g1y, g2x, g2y;
            if (shared) {
                elem_gradient(tu, gu, AB, m.S, k, g1x, g1y);
                elem_gradient(td, gd, AB, m.S, k, g2x, g2y);
            } else {
                const G *p1 = gnod + t * tsg + o1 * 2, *p2 = gnod + t * tsg + o2 * 2;
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
    int eid[kDeg];
    bool pos[kDeg];
    bool on[kDeg];
};

template <class R>
__device__ __forceinline__ void load_edge_list(const DMesh<R> &m, int n, int nz, bool layer, NodeEdgeList &L)
{
    const NodeEdges ne = m.nedge[n];
    const NodeEdgeLev nl = m.nedgel[n];
#pragma unroll
    for (int j = 0; j < kDeg; ++j) {
        const int se = ne.e[j];
        L.eid[j] = (se > 0 ? se : -se) - 1;
        L.pos[j] = se > 0;
        L.on[j] = se != 0 && layer && nz <= nl.nl[j];
    }
}

// ===========================================================================
// LO solution (lib/oce_adv_tra_driver.F90:95-211), HO-LO vertical flux
// (:257-283), and the raw antidiffusive sums P+/P- (lib/oce_adv_tra_fct.F90:232-346)
// ===========================================================================
template <class R, int VER, int NT>
__global__ __launch_bounds__(kBlock) void k_node_lo(DMesh<R> m, const R *__restrict__ Wv,
                                                    const R *__restrict__ WEv,
                                                    const double *__restrict__ ttf,
                                                    const double *__restrict__ ttfAB, size_t tsn,
                                                    const R *__restrict__ flux_lo,
                                                    const R *__restrict__ adf_h, size_t tse,
                                                    double dt, R num_ord,
                                                    double *__restrict__ LO, R *__restrict__ adf_v,
                                                    R *__restrict__ Pp, R *__restrict__ Pm)
{
    const int n = wave_column();
    if (n >= m.N) return;
    const int k = lane_id(), nz = k + 1;
    const int nlev = m.nlev_n[n];
    const size_t c0 = (size_t)n * m.S;
    const bool layer = nz <= nlev - 1;
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
                const size_t oe = te + (size_t)L.eid[j] * m.S + k;
                const double f = (double)flux_lo[oe];
                fadf[j] = adf_h[oe];
                slo = L.pos[j] ? slo + f : slo - f;
            }
        }
        // LO vertical flux (UPW1 with w_e and ttf) and HO-LO with w and ttfAB
        VCol<R> lc{ttf + tn + c0, WEv + c0, m.area + c0, m.Zn + c0, m.zbarn + c0, m.hnode + c0, m.hnode_new + c0, nlev, nz};
        const R fv_lo = vflux_upw1(lc, (R)0);
        const R fv_lo_next = shdn(fv_lo);
        VCol<R> hc{ttfAB + tn + c0, Wv + c0, m.area + c0, m.Zn + c0, m.zbarn + c0, m.hnode + c0, m.hnode_new + c0, nlev, nz};
        const R adfv = ver_flux<R, VER>(hc, num_ord, (R)dt, fv_lo);
        const R adfv_next = shdn(adfv);
        if (nz <= nlev) adf_v[tn + c0 + k] = adfv;
        if (layer) {
            const double T = ttf[tn + c0 + k];
            LO[tn + c0 + k] = (T * (double)m.hnode[c0 + k] +
                               (slo + (double)(fv_lo - fv_lo_next)) * dt / (double)m.areasvol[c0 + k]) /
                              (double)m.hnode_new[c0 + k];
            R pp = (R)0 + (fmx((R)0, adfv) + fmx((R)0, -adfv_next));
            R pm = (R)0 + (fmn((R)0, adfv) + fmn((R)0, -adfv_next));
#pragma unroll
            for (int j = 0; j < kDeg; ++j) {
                if (!L.on[j]) continue;
                const R a = fadf[j];
                if (L.pos[j]) { pp = pp + fmx((R)0, a); pm = pm + fmn((R)0, a); }
                else          { pp = pp + fmx((R)0, -a); pm = pm + fmn((R)0, -a); }
            }
            Pp[tn + c0 + k] = pp;
            Pm[tn + c0 + k] = pm;
        }
    }
}

// Zalesak bounds a1-a3 and limiting factors b2 (lib/oce_adv_tra_fct.F90:94-217,361-374);
// R+/R- overwrite P+/P- in place
template <class R, int NT>
__global__ __launch_bounds__(kBlock) void k_node_limit(DMesh<R> m, const double *__restrict__ ttf,
                                                       const double *__restrict__ LO, size_t tsn,
                                                       double dt, R *__restrict__ Pp_Rp,
                                                       R *__restrict__ Pm_Rm)
{
    const int n = wave_column();
    if (n >= m.N) return;
    const int k = lane_id(), nz = k + 1;
    const int nlev = m.nlev_n[n];
    const size_t c0 = (size_t)n * m.S;
    const bool layer = nz <= nlev - 1;
    const R bignumber = (R)1000;
    const NodeElems ne = m.nelem[n];

    R tvmax[NT], tvmin[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) { tvmax[t] = 0; tvmin[t] = 0; }
    if (layer) {
#pragma unroll
        for (int j = 0; j < kDeg; ++j) {
            const int el = ne.t[j];
            if (el < 0) break;
            const ElemMeta em = m.tmeta[el];
            const bool wet = nz <= em.nlev - 1;
            const size_t o0 = (size_t)em.n0 * m.S + k, o1 = (size_t)em.n1 * m.S + k, o2 = (size_t)em.n2 * m.S + k;
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                R emax = -bignumber, emin = bignumber;
                if (wet) {
                    const size_t tn = t * tsn;
                    const R l0 = (R)LO[tn + o0], l1 = (R)LO[tn + o1], l2 = (R)LO[tn + o2];
                    const R t0 = (R)ttf[tn + o0], t1 = (R)ttf[tn + o1], t2 = (R)ttf[tn + o2];
                    emax = fmx(fmx(fmx(l0, t0), fmx(l1, t1)), fmx(l2, t2));
                    emin = fmn(fmn(fmn(l0, t0), fmn(l1, t1)), fmn(l2, t2));
                }
                tvmax[t] = (j == 0) ? emax : fmx(tvmax[t], emax);
                tvmin[t] = (j == 0) ? emin : fmn(tvmin[t], emin);
            }
        }
    }
    const R flux_eps = (R)1e-16;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const R up_max = shup(tvmax[t]), dn_max = shdn(tvmax[t]);
        const R up_min = shup(tvmin[t]), dn_min = shdn(tvmin[t]);
        if (!layer) continue;
        const size_t tn = t * tsn;
        const R lo = (R)LO[tn + c0 + k];
        R qp, qm;
        if (nz == 1 || nz == nlev - 1) {
            qp = tvmax[t] - lo;
            qm = tvmin[t] - lo;
        } else {
            qp = fmx(fmx(up_max, tvmax[t]), dn_max) - lo;
            qm = fmn(fmn(up_min, tvmin[t]), dn_min) - lo;
        }
        const R av = m.areasvol[c0 + k];
        R flux = Pp_Rp[tn + c0 + k] * (R)dt / av + flux_eps;
        const R rp = fmn((R)1, qp / flux);
        flux = Pm_Rm[tn + c0 + k] * (R)dt / av - flux_eps;
        const R rm = fmn((R)1, qm / flux);
        Pp_Rp[tn + c0 + k] = rp;
        Pm_Rm[tn + c0 + k] = rm;
    }
}

// limiting b3 (lib/oce_adv_tra_fct.F90:394-471), flux divergence
// (lib/oce_adv_tra_driver.F90:298-450) and the update
// values += (dttf_h + dttf_v)/hnode_new (src/fesom_analytic.F90:419-437)
template <class R, int NT>
__global__ __launch_bounds__(kBlock) void k_node_update(DMesh<R> m, double *__restrict__ ttf,
                                                        const double *__restrict__ LO, size_t tsn,
                                                        const R *__restrict__ adf_h, size_t tse,
                                                        const R *__restrict__ adf_v,
                                                        const R *__restrict__ Rp,
                                                        const R *__restrict__ Rm,
                                                        const int *__restrict__ nother, double dt)
{
    const int n = wave_column();
    if (n >= m.N) return;
    const int k = lane_id(), nz = k + 1;
    const int nlev = m.nlev_n[n];
    const size_t c0 = (size_t)n * m.S;
    const bool layer = nz <= nlev - 1;
    NodeEdgeList L;
    load_edge_list(m, n, nz, layer, L);
    int oth[kDeg];
#pragma unroll
    for (int j = 0; j < kDeg; ++j) oth[j] = nother[(size_t)n * kDeg + j];

#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const size_t tn = t * tsn, te = t * tse;
        R rp = 0, rm = 0, fv = 0;
        if (nz <= nlev) fv = adf_v[tn + c0 + k];
        if (layer) { rp = Rp[tn + c0 + k]; rm = Rm[tn + c0 + k]; }
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

        const double av = (double)m.areasvol[c0 + k];
        double sh = 0.0;
#pragma unroll
        for (int j = 0; j < kDeg; ++j) {
            if (!L.on[j]) continue;
            const R f = adf_h[te + (size_t)L.eid[j] * m.S + k];
            const size_t oo = tn + (size_t)oth[j] * m.S + k;
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
        const double T = ttf[tn + c0 + k];
        double dv = 0.0 - T * (double)m.hnode[c0 + k] + LO[tn + c0 + k] * (double)m.hnode_new[c0 + k];
        dv = dv + (double)(fvl - fvl_next) * dt / av;
        const double del = (0.0 + sh) + dv;
        ttf[tn + c0 + k] = T + del / (double)m.hnode_new[c0 + k];
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
    const int nlev = m.nlev_n[n];
    const size_t c0 = (size_t)n * m.S;
    const bool layer = nz <= nlev - 1;
    NodeEdgeList L;
    load_edge_list(m, n, nz, layer, L);
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const size_t tn = t * tsn, te = t * tse;
        VCol<R> hc{ttfAB + tn + c0, WEv + c0, m.area + c0, m.Zn + c0, m.zbarn + c0, m.hnode + c0, m.hnode_new + c0, nlev, nz};
        const R fv = ver_flux<R, VER>(hc, num_ord, (R)dt, (R)0);
        const R fv_next = shdn(fv);
        if (!layer) continue;
        const double av = (double)m.areasvol[c0 + k];
        double sh = 0.0;
#pragma unroll
        for (int j = 0; j < kDeg; ++j) {
            if (!L.on[j]) continue;
            const double f = (double)flux_h[te + (size_t)L.eid[j] * m.S + k];
            sh = L.pos[j] ? sh + f * dt / av : sh - f * dt / av;
        }
        const double dv = 0.0 + (double)(fv - fv_next) * dt / av;
        const double del = (0.0 + sh) + dv;
        const double T = ttf[tn + c0 + k];
        ttf[tn + c0 + k] = T + del / (double)m.hnode_new[c0 + k];
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

template <class R>
__global__ __launch_bounds__(256) void k_diag_partial(DMesh<R> m, const double *__restrict__ v,
                                                      DiagAcc *__restrict__ part)
{
    __shared__ DiagAcc sh[256];
    DiagAcc a{INFINITY, -INFINITY, 0.0, INFINITY, -INFINITY, 0.0, 0.0};
    const int L = m.nl - 1;
    const long long total = (long long)m.N * L;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < total; i += (long long)gridDim.x * 256) {
        const int n = (int)(i / L), k = (int)(i % L);
        const double x = v[(size_t)n * m.S + k];
        if (x != x) { a.nan += 1.0; continue; }
        a.vmin = fmin(a.vmin, x); a.vmax = fmax(a.vmax, x); a.sum += x;
        if (k + 1 <= m.nlev_n[n] - 1) {
            a.amin = fmin(a.amin, x); a.amax = fmax(a.amax, x);
            a.mass += x * (double)m.hnode_new[(size_t)n * m.S + k] * (double)m.areasvol[(size_t)n * m.S + k];
        }
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

// unpadded (L, ncol) double -> padded (S, ncol) of type T (pads = 0)
template <class T>
std::vector<T> pad_cols(const double *src, int L, int ncol, int S, int comps = 1)
{
    std::vector<T> out((size_t)S * ncol * comps, (T)0);
    for (int c = 0; c < ncol; ++c)
        for (int k = 0; k < L; ++k)
            for (int q = 0; q < comps; ++q)
                out[((size_t)c * S + k) * comps + q] = (T)src[((size_t)c * L + k) * comps + q];
    return out;
}

inline unsigned grid_for(int ncols)
{
    unsigned nb = (unsigned)((ncols + kWavesPerBlock - 1) / kWavesPerBlock);
    return (nb + 7u) & ~7u;  // multiple of 8 for the XCD remap
}

}  // namespace

struct Engine::Impl {
    std::vector<void *> allocs;
    template <class T> T *track(T *p) { allocs.push_back((void *)p); return p; }

    DMesh<double> m64{};
    DMesh<float> m32{};
    int *nother = nullptr;  // (kDeg,N) other endpoint of each incident edge
    // velocity
    void *vel = nullptr, *W = nullptr, *WE = nullptr;
    // tracers [ntr][S*N]
    double *values = nullptr, *valuesAB = nullptr;
    // work, per tracer
    void *flux_lo = nullptr, *adf_h = nullptr;   // [ntr][S*E]
    void *adf_v = nullptr, *Pp = nullptr, *Pm = nullptr;  // [ntr][S*N]
    double *LO = nullptr;                        // [ntr][S*N]
    void *gnod = nullptr;                        // [ntr][2*S*N]
    DiagAcc *dpart = nullptr, *dout = nullptr;
    int ndiag = 1024;
    float gscale = 1.0f, ginv = 1.0f;
    size_t tsn = 0, tse = 0, tsg = 0;
};

template <class R>
static void build_dmesh(Engine::Impl &I, DMesh<R> &d, const HostMesh &m, int S)
{
    const int N = m.nod2D, T = m.elem2D, E = m.edge2D, nl = m.nl;
    d.N = N; d.T = T; d.E = E; d.nl = nl; d.S = S;
    std::vector<EdgeMeta> em(E);
    std::vector<R> eg(8 * (size_t)E, (R)0);
    for (int e = 0; e < E; ++e) {
        EdgeMeta &r = em[e];
        r.n1 = m.edges[2 * e]; r.n2 = m.edges[2 * e + 1];
        r.el1 = m.edge_tri[2 * e]; r.el2 = m.edge_tri[2 * e + 1];
        r.up = m.edge_up_dn_tri[2 * e]; r.dn = m.edge_up_dn_tri[2 * e + 1];
        int nl1 = m.nlevels[r.el1] - 1, nl2 = r.el2 >= 0 ? m.nlevels[r.el2] - 1 : 0;
        int nzmax_e = std::min(m.nlevels_nod2D_min[r.n1], m.nlevels_nod2D_min[r.n2]);
        int nlmax = std::max(nl1, nl2);
        r.nl = nl1 | (nl2 << 8) | (nzmax_e << 16) | (nlmax << 24);
        r.nb = (m.nboundary_lay[r.n1] & 0xFFFF) | ((m.nboundary_lay[r.n2] & 0xFFFF) << 16);
        double a = kRE * m.elem_cos[r.el1];
        if (r.el2 >= 0) a = 0.5 * (a + kRE * m.elem_cos[r.el2]);
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
        tm[el] = ElemMeta{m.elem2D_nodes[3 * el], m.elem2D_nodes[3 * el + 1], m.elem2D_nodes[3 * el + 2], m.nlevels[el]};
        for (int q = 0; q < 6; ++q) tg[8 * (size_t)el + q] = (R)m.gradient_sca[6 * el + q];
        tg[8 * (size_t)el + 6] = (R)m.elem_area[el];
    }
    d.tmeta = I.track(dupload(tm));
    d.tgeo = I.track(dupload(tg));
    d.nlev_n = I.track(dupload(m.nlevels_nod2D));
    std::vector<NodeEdges> ne(N);
    std::vector<NodeEdgeLev> nel(N);
    std::vector<NodeElems> nt(N);
    std::vector<int> oth((size_t)N * kDeg, 0);
    for (int n = 0; n < N; ++n) {
        const int deg = m.n2e_ptr[n + 1] - m.n2e_ptr[n];
        if (deg > kDeg || m.nod_in_elem2D_num[n] > kDeg)
            throw std::runtime_error("node degree above " + std::to_string(kDeg) + " is not supported");
        for (int j = 0; j < kDeg; ++j) {
            ne[n].e[j] = 0;
            nel[n].nl[j] = 0;
            nt[n].t[j] = -1;
        }
        for (int j = 0; j < deg; ++j) {
            int se = m.n2e[m.n2e_ptr[n] + j];
            int e = (se > 0 ? se : -se) - 1;
            ne[n].e[j] = se;
            nel[n].nl[j] = (int)((unsigned)em[e].nl >> 24);
            oth[(size_t)n * kDeg + j] = se > 0 ? em[e].n2 : em[e].n1;
        }
        for (int j = 0; j < m.nod_in_elem2D_num[n]; ++j) nt[n].t[j] = m.nod_in_elem2D[(size_t)n * m.maxAdj + j];
    }
    d.nedge = I.track(dupload(ne));
    d.nedgel = I.track(dupload(nel));
    d.nelem = I.track(dupload(nt));
    if (!I.nother) I.nother = I.track(dupload(oth));
    auto cv = [](const std::vector<double> &v) { return std::vector<R>(v.begin(), v.end()); };
    (void)cv;
    d.area = I.track(dupload(pad_cols<R>(m.area.data(), nl, N, S)));
    d.areasvol = I.track(dupload(pad_cols<R>(m.areasvol.data(), nl, N, S)));
    d.hnode = I.track(dupload(pad_cols<R>(m.hnode.data(), nl - 1, N, S)));
    d.hnode_new = I.track(dupload(pad_cols<R>(m.hnode_new.data(), nl - 1, N, S)));
    d.Zn = I.track(dupload(pad_cols<R>(m.Z_3d_n.data(), nl - 1, N, S)));
    d.zbarn = I.track(dupload(pad_cols<R>(m.zbar_3d_n.data(), nl, N, S)));
    d.helem = I.track(dupload(pad_cols<R>(m.helem.data(), nl - 1, T, S)));
}

Engine::Engine(const HostMesh &m, int ntracers, int precision)
    : p_(new Impl), N_(m.nod2D), T_(m.elem2D), E_(m.edge2D), nl_(m.nl), ntr_(ntracers), prec_(precision)
{
    if (nl_ > kWave) throw std::runtime_error("the device path supports nl <= 64 level interfaces");
    if (ntracers < 1) throw std::runtime_error("need at least one tracer");
    if (precision < PREC_FP64 || precision > PREC_FP32_H16) throw std::runtime_error("unknown precision");
    for (int u : m.ulevels) if (u != 1) throw std::runtime_error("cavities (ulevels > 1) are not supported");
    S_ = std::max(8, (nl_ + 7) / 8 * 8);
    spec_.assign(ntr_, TracerSpec());
    Impl &I = *p_;
    HIP_OK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    const size_t rs = prec_ == PREC_FP64 ? sizeof(double) : sizeof(float);
    const size_t gs = prec_ == PREC_FP64 ? sizeof(double) : (prec_ == PREC_FP32 ? sizeof(float) : sizeof(__half));
    if (prec_ == PREC_FP64) build_dmesh(I, I.m64, m, S_);
    else build_dmesh(I, I.m32, m, S_);
    I.tsn = (size_t)S_ * N_;
    I.tse = (size_t)S_ * E_;
    I.tsg = 2 * (size_t)S_ * N_;
    auto zalloc = [&](size_t bytes) {
        void *p = I.track((void *)dalloc<char>(bytes));
        HIP_OK(hipMemset(p, 0, bytes));
        return p;
    };
    I.vel = zalloc(2 * (size_t)S_ * T_ * rs);
    I.W = zalloc((size_t)S_ * N_ * rs);
    I.WE = zalloc((size_t)S_ * N_ * rs);
    I.values = (double *)zalloc(I.tsn * ntr_ * sizeof(double));
    I.valuesAB = (double *)zalloc(I.tsn * ntr_ * sizeof(double));
    I.flux_lo = zalloc(I.tse * ntr_ * rs);
    I.adf_h = zalloc(I.tse * ntr_ * rs);
    I.adf_v = zalloc(I.tsn * ntr_ * rs);
    I.Pp = zalloc(I.tsn * ntr_ * rs);
    I.Pm = zalloc(I.tsn * ntr_ * rs);
    I.LO = (double *)zalloc(I.tsn * ntr_ * sizeof(double));
    I.gnod = zalloc(I.tsg * ntr_ * gs);
    I.dpart = I.track(dalloc<DiagAcc>(I.ndiag));
    I.dout = I.track(dalloc<DiagAcc>(1));
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
    const int L = nl_ - 1;
    auto put = [&](void *dst, const double *src, int Lc, int ncol, int comps) {
        if (prec_ == PREC_FP64) {
            auto v = pad_cols<double>(src, Lc, ncol, S_, comps);
            HIP_OK(hipMemcpy(dst, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
        } else {
            auto v = pad_cols<float>(src, Lc, ncol, S_, comps);
            HIP_OK(hipMemcpy(dst, v.data(), v.size() * sizeof(float), hipMemcpyHostToDevice));
        }
    };
    put(I.vel, uv, L, T_, 2);
    put(I.W, w, nl_, N_, 1);
    put(I.WE, we, nl_, N_, 1);
}

void Engine::set_values(int tr, const double *v)
{
    if (tr < 0 || tr >= ntr_) throw std::runtime_error("tracer index out of range");
    auto a = pad_cols<double>(v, nl_ - 1, N_, S_);
    HIP_OK(hipMemcpy(p_->values + tr * p_->tsn, a.data(), a.size() * sizeof(double), hipMemcpyHostToDevice));
}

void Engine::set_valuesAB(int tr, const double *v)
{
    if (tr < 0 || tr >= ntr_) throw std::runtime_error("tracer index out of range");
    auto a = pad_cols<double>(v, nl_ - 1, N_, S_);
    HIP_OK(hipMemcpy(p_->valuesAB + tr * p_->tsn, a.data(), a.size() * sizeof(double), hipMemcpyHostToDevice));
}

void Engine::get_values(int tr, double *v)
{
    if (tr < 0 || tr >= ntr_) throw std::runtime_error("tracer index out of range");
    std::vector<double> a((size_t)S_ * N_);
    HIP_OK(hipStreamSynchronize(stream_));
    HIP_OK(hipMemcpy(a.data(), p_->values + tr * p_->tsn, a.size() * sizeof(double), hipMemcpyDeviceToHost));
    const int L = nl_ - 1;
    for (int n = 0; n < N_; ++n)
        for (int k = 0; k < L; ++k) v[(size_t)n * L + k] = a[(size_t)n * S_ + k];
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
    const size_t tsn = I.tsn, tse = I.tse, tsg = I.tsg;
    double *vals = I.values + t0 * tsn;
    const double *ab = refresh ? vals : I.valuesAB + t0 * tsn;
    R *flux_lo = (R *)I.flux_lo + t0 * tse, *adf_h = (R *)I.adf_h + t0 * tse;
    R *adf_v = (R *)I.adf_v + t0 * tsn, *Pp = (R *)I.Pp + t0 * tsn, *Pm = (R *)I.Pm + t0 * tsn;
    double *LO = I.LO + t0 * tsn;
    G *gnod = (G *)I.gnod + t0 * tsg;
    const dim3 blk(kBlock), ge(grid_for(m.E)), gn(grid_for(m.N));
    const bool fct = sp.lim == LIM_FCT;
    const R ph = (R)sp.ph, pv = (R)sp.pv;
    if (sp.hor == HOR_MUSCL || sp.hor == HOR_MFCT) {
        hipLaunchKernelGGL((k_grad<R, G, NT>), gn, blk, 0, st, m, ab, tsn, gnod, tsg, I.gscale);
        hk.mark("k_grad", st);
    }
#define EDGE(HOR, F) hipLaunchKernelGGL((k_edge<R, G, HOR, F, NT>), ge, blk, 0, st, m, vel, vals, ab, tsn, gnod, tsg, ph, I.ginv, flux_lo, adf_h, tse)
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
#define NLO(V) hipLaunchKernelGGL((k_node_lo<R, V, NT>), gn, blk, 0, st, m, W, WE, vals, ab, tsn, flux_lo, adf_h, tse, dt, pv, LO, adf_v, Pp, Pm)
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
        hipLaunchKernelGGL((k_node_update<R, NT>), gn, blk, 0, st, m, vals, LO, tsn, adf_h, tse, adf_v, Pp, Pm,
                           I.nother, dt);
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
    return Diag{a.vmin, a.vmax, a.sum, a.amin, a.amax, a.mass, a.nan};
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
    int L = nl_ - 1, ncol = N_, comps = 1;
    size_t es = prec_ == PREC_FP64 ? 8 : 4;
    size_t ts = I.tsn;
    if (name == "fct_LO") { src = (const char *)I.LO; es = 8; }
    else if (name == "fct_plus") src = (const char *)I.Pp;
    else if (name == "fct_minus") src = (const char *)I.Pm;
    else if (name == "adv_flux_hor") { src = (const char *)I.adf_h; ncol = E_; ts = I.tse; }
    else if (name == "adv_flux_lo") { src = (const char *)I.flux_lo; ncol = E_; ts = I.tse; }
    else if (name == "adv_flux_ver") { src = (const char *)I.adf_v; L = nl_; }
    else if (name == "gnod") {
        src = (const char *)I.gnod; comps = 2; ts = I.tsg;
        es = prec_ == PREC_FP64 ? 8 : (prec_ == PREC_FP32 ? 4 : 2);
    }
    std::vector<char> buf((size_t)S_ * ncol * comps * es);
    HIP_OK(hipMemcpy(buf.data(), src + tr * ts * es, buf.size(), hipMemcpyDeviceToHost));
    for (int c = 0; c < ncol; ++c)
        for (int k = 0; k < L; ++k)
            for (int q = 0; q < comps; ++q) {
                size_t i = ((size_t)c * S_ + k) * comps + q;
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
