// fta_engine.hip -- the tracer-advection step on gfx950 and its host driver.
//
// One step of one tracer (reference: do_oce_adv_tra, lib/oce_adv_tra_driver.F90:23-294,
// plus the caller's update, src/fesom_analytic.F90:405-437) is five launches:
//   k_elem_grad   elemental gradient of valuesAB (MUSCL/MFCT only)
//   k_node_grad   area-weighted nodal mean gradient (MUSCL/MFCT only)
//   k_edge_flux   per edge-level: LO upwind flux and HO-LO antidiffusive flux
//   k_node_lo     per node-level: LO solution, vertical HO-LO flux, P+/P- sums
//   k_node_limit  per node-level: Zalesak bounds (a1-a3) and R+/R- (b2)
//   k_node_update per node-level: limited fluxes (b3), divergence, T update
// (without FCT the last three collapse into k_node_nolim).
#include "fta_engine.h"
#include "fta_kernels.h"

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

// ===========================================================================
// Vertical fluxes at interface nz of a node column (lane = interface).
// A is the tracer the scheme reads, W the vertical velocity it uses and
// fin the flux already stored there (o_init_zero=.false. semantics:
// the routine returns new - fin). All lanes must call (PPM shuffles).
// ===========================================================================
template <class R>
struct VCol {
    const double *A;  // tracer column (double storage)
    const R *W, *area, *Z, *zbar, *h, *hn;
    int nlev;         // nlevels_nod2D: interfaces 1..nlev, layers 1..nlev-1
    int nz;           // this lane's interface / layer (1-based)
};

template <class R>
__device__ __forceinline__ R ldA(const VCol<R> &c, int nz)
{
    return (nz >= 1 && nz <= c.nlev - 1) ? (R)c.A[nz - 1] : (R)0;
}

// adv_tra_ver_upw1 (lib/oce_adv_tra_ver.F90:248-335)
template <class R>
__device__ __forceinline__ R vflux_upw1(const VCol<R> &c, R fin)
{
    const int nz = c.nz;
    if (nz == 1) return -c.W[0] * ldA(c, 1) * c.area[0] - fin;
    if (nz == c.nlev) return (R)0 - fin;
    if (nz < c.nlev) {
        R w = c.W[nz - 1];
        return (R)(-0.5) * (ldA(c, nz) * (w + fabs(w)) + ldA(c, nz - 1) * (w - fabs(w))) * c.area[nz - 1] - fin;
    }
    return fin;
}

// adv_tra_ver_qr4c (lib/oce_adv_tra_ver.F90:339-444) including the double
// write of interface nzmin+1 on two-layer columns
template <class R>
__device__ __forceinline__ R vflux_qr4c(const VCol<R> &c, R num_ord, R fin)
{
    const int nz = c.nz, nlev = c.nlev;
    if (nz == 1) return -ldA(c, 1) * c.W[0] * c.area[0] - fin;
    if (nz == nlev) return (R)0 - fin;
    if (nz > nlev) return fin;
    if (nz == 2 || nz == nlev - 1) {
        R cen = (R)(-0.5) * (ldA(c, nz - 1) + ldA(c, nz)) * c.W[nz - 1] * c.area[nz - 1];
        if (nz == 2 && nz == nlev - 1) return cen - (cen - fin);
        return cen - fin;
    }
    // 3 <= nz <= nlev-2
    R t0 = ldA(c, nz), tm1 = ldA(c, nz - 1), tm2 = ldA(c, nz - 2), tp1 = ldA(c, nz + 1);
    R Z0 = c.Z[nz - 1], Zm1 = c.Z[nz - 2], Zm2 = c.Z[nz - 3], Zp1 = c.Z[nz];
    R qc = (tm1 - t0) / (Zm1 - Z0);
    R qu = (t0 - tp1) / (Z0 - Zp1);
    R qd = (tm2 - tm1) / (Zm2 - Zm1);
    R zb = c.zbar[nz - 1];
    R Tmean1 = t0 + ((R)2 * qc + qu) * (zb - Z0) / (R)3;
    R Tmean2 = tm1 + ((R)2 * qc + qd) * (zb - Zm1) / (R)3;
    R w = c.W[nz - 1];
    R Tmean = (w + fabs(w)) * Tmean1 + (w - fabs(w)) * Tmean2;
    return ((R)(-0.5) * ((R)1 - num_ord) * Tmean - num_ord * ((R)0.5 * (Tmean1 + Tmean2)) * w) * c.area[nz - 1] - fin;
}

// adv_tra_ver_cdiff (lib/oce_adv_tra_ver.F90:648-712); interface nlev untouched
template <class R>
__device__ __forceinline__ R vflux_cdiff(const VCol<R> &c, R fin)
{
    const int nz = c.nz;
    if (nz == 1) return -c.W[0] * ldA(c, 1) * c.area[0] - fin;
    if (nz <= c.nlev - 1) {
        R tv = (R)0.5 * (ldA(c, nz - 1) + ldA(c, nz));
        return -tv * c.W[nz - 1] * c.area[nz - 1] - fin;
    }
    return fin;
}

// adv_tra_vert_ppm (lib/oce_adv_tra_ver.F90:448-644). Lane nz first builds
// the interface value tv(nz), then the limited parabola of layer nz; the
// flux through interface nz comes from its upwind layer (nz if W>0, nz-1 if
// W<0), which is exactly what the reference's per-layer writes produce
// (its CYCLE only skips layers that would write nothing).
template <class R>
__device__ __forceinline__ R vflux_ppm(const VCol<R> &c, R dt, R fin)
{
    const int nz = c.nz, nlev = c.nlev;
    const int nzmax = nlev;
    // --- interface values tv(nz)
    R tv = (R)0;
    if (nz == 1) tv = ldA(c, 1);
    else if (nz == nzmax) tv = ldA(c, nzmax - 1);
    else if (nz == nzmax - 1) {
        R sw = copysign((R)1, c.W[nzmax - 2]);
        tv = -ldA(c, nzmax - 2) * fmn(sw, (R)0) + ldA(c, nzmax - 1) * fmx(sw, (R)0);
    } else if (nz == 2) tv = (R)0.5 * (ldA(c, 1) + ldA(c, 2));
    else if (nz >= 3 && nz <= nzmax - 2) {
        const int j = nz - 1;  // reference loop index: tv(j+1)
        R dzjm1 = c.hn[j - 2], dzj = c.hn[j - 1], dzjp1 = c.hn[j], dzjp2 = c.hn[j + 1];
        R Tm1 = ldA(c, j - 1), T0 = ldA(c, j), Tp1 = ldA(c, j + 1), Tp2 = ldA(c, j + 2);
        R deltaj = dzj / (dzjm1 + dzj + dzjp1) *
                   (((R)2 * dzjm1 + dzj) / (dzjp1 + dzj) * (Tp1 - T0) + (dzj + (R)2 * dzjp1) / (dzjm1 + dzj) * (T0 - Tm1));
        R deltajp1 = dzjp1 / (dzj + dzjp1 + dzjp2) *
                     (((R)2 * dzj + dzjp1) / (dzjp2 + dzjp1) * (Tp2 - Tp1) + (dzjp1 + (R)2 * dzjp2) / (dzj + dzjp1) * (Tp1 - T0));
        if ((Tp1 - T0) * (T0 - Tm1) > (R)0)
            deltaj = fmn(fmn(fabs(deltaj), (R)2 * fabs(Tp1 - T0)), (R)2 * fabs(T0 - Tm1)) * copysign((R)1, deltaj);
        else
            deltaj = (R)0;
        if ((Tp2 - Tp1) * (Tp1 - T0) > (R)0)
            deltajp1 = fmn(fmn(fabs(deltajp1), (R)2 * fabs(Tp2 - Tp1)), (R)2 * fabs(Tp1 - T0)) * copysign((R)1, deltajp1);
        else
            deltajp1 = (R)0;
        tv = T0 + dzj / (dzj + dzjp1) * (Tp1 - T0) +
             (R)1 / (dzjm1 + dzj + dzjp1 + dzjp2) *
             ((R)2 * dzjp1 * dzj / (dzj + dzjp1) *
              ((dzjm1 + dzj) / ((R)2 * dzj + dzjp1) - (dzjp2 + dzjp1) / ((R)2 * dzjp1 + dzj)) * (Tp1 - T0) -
              dzj * (dzjm1 + dzj) / ((R)2 * dzj + dzjp1) * deltajp1 +
              dzjp1 * (dzjp1 + dzjp2) / (dzj + (R)2 * dzjp1) * deltaj);
    }
    // --- limited parabola of layer nz: aL = tv(nz), aR = tv(nz+1)
    R aL = tv, aR = shdn(tv);
    R a0 = ldA(c, nz);
    if ((aR - a0) * (a0 - aL) <= (R)0) { aL = a0; aR = a0; }
    if ((aR - aL) * (a0 - (R)0.5 * (aL + aR)) > (aR - aL) * (aR - aL) / (R)6) aL = (R)3 * a0 - (R)2 * aR;
    if ((aR - aL) * (a0 - (R)0.5 * (aR + aL)) < -(aR - aL) * (aR - aL) / (R)6) aR = (R)3 * a0 - (R)2 * aL;
    R aj = (R)6 * (a0 - (R)0.5 * (aL + aR));
    // parabola of layer nz-1 for downward flow through interface nz
    R aLu = shup(aL), aRu = shup(aR), aju = shup(aj);
    if (nz > nzmax) return fin;
    R tvert = (R)0;
    R w = c.W[nz - 1];
    if (nz <= nzmax - 1 && w > (R)0) {
        R x = fmn(w * dt / c.h[nz - 1], (R)1);
        tvert = (-aL - (R)0.5 * x * (aR - aL + ((R)1 - (R)2 / (R)3 * x) * aj));
        tvert = tvert * c.area[nz - 1] * w;
    }
    if (nz >= 2 && w < (R)0) {
        R x = fmn(-w * dt / c.h[nz - 2], (R)1);
        tvert = (-aRu + (R)0.5 * x * (aRu - aLu - ((R)1 - (R)2 / (R)3 * x) * aju));
        tvert = tvert * c.area[nz - 1] * w;
    }
    if (nz == 1) tvert = -tv * c.W[0] * c.area[0];  // surface: -tv(nzmin)*W*area
    if (nz == nzmax) tvert = (R)0;
    return tvert - fin;
}

template <class R, int VER>
__device__ __forceinline__ R ver_flux(const VCol<R> &c, R num_ord, R dt, R fin)
{
    if constexpr (VER == VER_QR4C) return vflux_qr4c(c, num_ord, fin);
    else if constexpr (VER == VER_PPM) return vflux_ppm(c, dt, fin);
    else if constexpr (VER == VER_CDIFF) return vflux_cdiff(c, fin);
    else return vflux_upw1(c, fin);
}

// ===========================================================================
// MUSCL prerequisites: elemental gradient and nodal mean gradient
// ===========================================================================
template <class R, class G>
__global__ __launch_bounds__(kBlock) void k_elem_grad(DMesh<R> m, const double *__restrict__ ttf,
                                                      G *__restrict__ trxy, float gscale)
{
    const int el = wave_column();
    if (el >= m.T) return;
    const int k = lane_id(), nz = k + 1;
    const int nlev = m.nlev_e[el];
    if (nz > nlev - 1) return;
    const int n0 = m.en[3 * el], n1 = m.en[3 * el + 1], n2 = m.en[3 * el + 2];
    const R *g = m.gsca + 6 * el;
    const R A0 = (R)ttf[(size_t)n0 * m.S + k], A1 = (R)ttf[(size_t)n1 * m.S + k], A2 = (R)ttf[(size_t)n2 * m.S + k];
    R gx = g[0] * A0 + g[1] * A1 + g[2] * A2;
    R gy = g[3] * A0 + g[4] * A1 + g[5] * A2;
    G *o = trxy + ((size_t)el * m.S + k) * 2;
    GradIO<G>::st(o, gx, gscale);
    GradIO<G>::st(o + 1, gy, gscale);
}

// area-weighted mean over the node's elements that have layer nz
// (inner loop of fill_up_dn_grad, lib/oce_muscl_adv.F90:397-407)
template <class R, class G>
__global__ __launch_bounds__(kBlock) void k_node_grad(DMesh<R> m, const G *__restrict__ trxy,
                                                      G *__restrict__ gnod, float gscale, float ginv)
{
    const int n = wave_column();
    if (n >= m.N) return;
    const int k = lane_id(), nz = k + 1;
    const int nlev = m.nlev_n[n];
    if (nz > nlev - 1) return;
    R tvol = 0, tx = 0, ty = 0;
    const int j0 = m.n2t_ptr[n], j1 = m.n2t_ptr[n + 1];
    for (int j = j0; j < j1; ++j) {
        const int el = m.n2t[j];
        if (m.nlev_e[el] - 1 < nz) continue;
        const R a = m.earea[el];
        const G *t = trxy + ((size_t)el * m.S + k) * 2;
        tvol = tvol + a;
        tx = tx + (R)GradIO<G>::ld(t, ginv) * a;
        ty = ty + (R)GradIO<G>::ld(t + 1, ginv) * a;
    }
    G *o = gnod + ((size_t)n * m.S + k) * 2;
    GradIO<G>::st(o, tx / tvol, gscale);
    GradIO<G>::st(o + 1, ty / tvol, gscale);
}

// ===========================================================================
// Horizontal edge fluxes (lib/oce_adv_tra_hor.F90): UPW1 (:64-261),
// MUSCL (:265-550), MFCT (:554-846). Segments C (both elements), D (el1
// only) and E (el2 only); no cavities, so segments A/B are empty.
// FCT: writes the LO flux (from ttf) and HO-LO (HO from ttfAB) per level.
// No limiter: writes the HO flux.
// ===========================================================================
template <class R, class G, int HOR, bool FCT>
__global__ __launch_bounds__(kBlock) void k_edge_flux(DMesh<R> m, const R *__restrict__ vel,
                                                      const double *__restrict__ ttf,
                                                      const double *__restrict__ ttfAB,
                                                      const G *__restrict__ trxy, const G *__restrict__ gnod,
                                                      R num_ord, float ginv,
                                                      R *__restrict__ flux_lo, R *__restrict__ flux)
{
    const int e = wave_column();
    if (e >= m.E) return;
    const int k = lane_id(), nz = k + 1;
    const int2 nn = m.edges[e], tt = m.etri[e];
    const int nl1 = m.nlev_e[tt.x] - 1;
    const int nl2 = tt.y >= 0 ? m.nlev_e[tt.y] - 1 : 0;
    const int nl12 = min(nl1, nl2), nlmax = max(nl1, nl2);
    if (nz > nlmax) return;
    const R dx1 = m.ecross[4 * e], dy1 = m.ecross[4 * e + 1];
    const R dx2 = m.ecross[4 * e + 2], dy2 = m.ecross[4 * e + 3];
    const bool has1 = nz <= nl1;
    const bool has2 = tt.y >= 0 && (nz <= nl12 || nz > nl1);
    R vf1 = 0, vf2 = 0;
    if (has1) {
        const R *u = vel + ((size_t)tt.x * m.S + k) * 2;
        vf1 = (-u[1] * dx1 + u[0] * dy1) * m.helem[(size_t)tt.x * m.S + k];
    }
    if (has2) {
        const R *u = vel + ((size_t)tt.y * m.S + k) * 2;
        vf2 = (u[1] * dx2 - u[0] * dy2) * m.helem[(size_t)tt.y * m.S + k];
    }
    const R vflux = (has1 && has2) ? vf1 + vf2 : (has1 ? vf1 : vf2);
    const size_t o = (size_t)e * m.S + k;

    R lo = 0;
    if constexpr (FCT) {
        const R T1 = (R)ttf[(size_t)nn.x * m.S + k], T2 = (R)ttf[(size_t)nn.y * m.S + k];
        lo = (R)(-0.5) * (T1 * (vflux + fabs(vflux)) + T2 * (vflux - fabs(vflux))) - (R)0;
        flux_lo[o] = lo;
    }
    const R A1 = (R)ttfAB[(size_t)nn.x * m.S + k], A2 = (R)ttfAB[(size_t)nn.y * m.S + k];
    R ho;
    if constexpr (HOR == HOR_UPW1) {
        ho = (R)(-0.5) * (A1 * (vflux + fabs(vflux)) + A2 * (vflux - fabs(vflux)));
    } else {
        // edge_up_dn_grad as fill_up_dn_grad builds it (lib/oce_muscl_adv.F90:361-530)
        const int2 ud = m.updn[e];
        const int nzmax_e = min(m.nlev_n_min[nn.x], m.nlev_n_min[nn.y]);
        R g1x, g1y, g2x, g2y;
        if (ud.x >= 0 && ud.y >= 0 && nz <= nzmax_e - 1) {
            const G *a = trxy + ((size_t)ud.x * m.S + k) * 2;
            const G *b = trxy + ((size_t)ud.y * m.S + k) * 2;
            g1x = (R)GradIO<G>::ld(a, ginv); g1y = (R)GradIO<G>::ld(a + 1, ginv);
            g2x = (R)GradIO<G>::ld(b, ginv); g2y = (R)GradIO<G>::ld(b + 1, ginv);
        } else {
            const G *a = gnod + ((size_t)nn.x * m.S + k) * 2;
            const G *b = gnod + ((size_t)nn.y * m.S + k) * 2;
            g1x = (R)GradIO<G>::ld(a, ginv); g1y = (R)GradIO<G>::ld(a + 1, ginv);
            g2x = (R)GradIO<G>::ld(b, ginv); g2y = (R)GradIO<G>::ld(b + 1, ginv);
        }
        R clo1 = 1, clo2 = 1;
        if constexpr (HOR == HOR_MUSCL) {
            clo1 = (m.nbl[nn.x] - nz) >= 0 ? (R)1 : (R)0;
            clo2 = (m.nbl[nn.y] - nz) >= 0 ? (R)1 : (R)0;
        }
        const R a = m.edge_a[e], ex = m.edxdy[2 * e], ey = m.edxdy[2 * e + 1];
        const R Tmean2 = A2 - ((R)2 * (A2 - A1) + ex * a * g2x + ey * (R)kRE * g2y) / (R)6 * clo2;
        const R Tmean1 = A1 + ((R)2 * (A2 - A1) + ex * a * g1x + ey * (R)kRE * g1y) / (R)6 * clo1;
        const R cHO = (vflux + fabs(vflux)) * Tmean1 + (vflux - fabs(vflux)) * Tmean2;
        ho = (R)(-0.5) * ((R)1 - num_ord) * cHO - vflux * num_ord * (R)0.5 * (Tmean1 + Tmean2);
    }
    flux[o] = ho - lo;
}

// ===========================================================================
// FCT node kernels
// ===========================================================================
// LO solution (lib/oce_adv_tra_driver.F90:95-211), HO-LO vertical flux
// (:257-283), and the raw antidiffusive sums P+/P- (lib/oce_adv_tra_fct.F90:232-346)
template <class R, int VER>
__global__ __launch_bounds__(kBlock) void k_node_lo(DMesh<R> m, const R *__restrict__ Wv,
                                                    const R *__restrict__ WEv,
                                                    const double *__restrict__ ttf,
                                                    const double *__restrict__ ttfAB,
                                                    const R *__restrict__ flux_lo,
                                                    const R *__restrict__ adf_h,
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

    // LO horizontal sum over incident edges, edge-index order
    double slo = 0.0;
    const int j0 = m.n2e_ptr[n], j1 = m.n2e_ptr[n + 1];
    if (layer)
        for (int j = j0; j < j1; ++j) {
            const int se = m.n2e[j];
            const int e = (se > 0 ? se : -se) - 1;
            if (nz > m.edge_nlmax[e]) continue;
            const double f = (double)flux_lo[(size_t)e * m.S + k];
            slo = se > 0 ? slo + f : slo - f;
        }

    // LO vertical flux (UPW1 with w_e and ttf) and HO-LO with w and ttfAB
    VCol<R> lc{ttf + c0, WEv + c0, m.area + c0, m.Zn + c0, m.zbarn + c0, m.hnode + c0, m.hnode_new + c0, nlev, nz};
    const R fv_lo = vflux_upw1(lc, (R)0);
    const R fv_lo_next = shdn(fv_lo);
    VCol<R> hc{ttfAB + c0, Wv + c0, m.area + c0, m.Zn + c0, m.zbarn + c0, m.hnode + c0, m.hnode_new + c0, nlev, nz};
    const R adfv = ver_flux<R, VER>(hc, num_ord, (R)dt, fv_lo);
    const R adfv_next = shdn(adfv);

    if (nz <= nlev) adf_v[c0 + k] = adfv;
    if (!layer) return;

    const double T = ttf[c0 + k];
    const double lo = (T * (double)m.hnode[c0 + k] +
                       (slo + (double)(fv_lo - fv_lo_next)) * dt / (double)m.areasvol[c0 + k]) /
                      (double)m.hnode_new[c0 + k];
    LO[c0 + k] = lo;

    R pp = (R)0 + (fmx((R)0, adfv) + fmx((R)0, -adfv_next));
    R pm = (R)0 + (fmn((R)0, adfv) + fmn((R)0, -adfv_next));
    for (int j = j0; j < j1; ++j) {
        const int se = m.n2e[j];
        const int e = (se > 0 ? se : -se) - 1;
        if (nz > m.edge_nlmax[e]) continue;
        const R a = adf_h[(size_t)e * m.S + k];
        if (se > 0) { pp = pp + fmx((R)0, a); pm = pm + fmn((R)0, a); }
        else        { pp = pp + fmx((R)0, -a); pm = pm + fmn((R)0, -a); }
    }
    Pp[c0 + k] = pp;
    Pm[c0 + k] = pm;
}

// Zalesak bounds a1-a3 and limiting factors b2 (lib/oce_adv_tra_fct.F90:94-217,361-374)
template <class R>
__global__ __launch_bounds__(kBlock) void k_node_limit(DMesh<R> m, const double *__restrict__ ttf,
                                                       const double *__restrict__ LO,
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

    R tvmax = 0, tvmin = 0;
    if (layer) {
        const int j0 = m.n2t_ptr[n], j1 = m.n2t_ptr[n + 1];
        for (int j = j0; j < j1; ++j) {
            const int el = m.n2t[j];
            R emax, emin;
            if (nz <= m.nlev_e[el] - 1) {
                R mx[3], mi[3];
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const size_t o = (size_t)m.en[3 * el + q] * m.S + k;
                    const R l = (R)LO[o], t = (R)ttf[o];
                    mx[q] = fmx(l, t);
                    mi[q] = fmn(l, t);
                }
                emax = fmx(fmx(mx[0], mx[1]), mx[2]);
                emin = fmn(fmn(mi[0], mi[1]), mi[2]);
            } else {
                emax = -bignumber;
                emin = bignumber;
            }
            tvmax = (j == j0) ? emax : fmx(tvmax, emax);
            tvmin = (j == j0) ? emin : fmn(tvmin, emin);
        }
    }
    const R up_max = shup(tvmax), dn_max = shdn(tvmax);
    const R up_min = shup(tvmin), dn_min = shdn(tvmin);
    if (!layer) return;
    const R lo = (R)LO[c0 + k];
    R qp, qm;
    if (nz == 1 || nz == nlev - 1) {
        qp = tvmax - lo;
        qm = tvmin - lo;
    } else {
        qp = fmx(fmx(up_max, tvmax), dn_max) - lo;
        qm = fmn(fmn(up_min, tvmin), dn_min) - lo;
    }
    const R flux_eps = (R)1e-16;
    const R av = m.areasvol[c0 + k];
    R flux = Pp_Rp[c0 + k] * (R)dt / av + flux_eps;
    const R rp = fmn((R)1, qp / flux);
    flux = Pm_Rm[c0 + k] * (R)dt / av - flux_eps;
    const R rm = fmn((R)1, qm / flux);
    Pp_Rp[c0 + k] = rp;
    Pm_Rm[c0 + k] = rm;
}

// limiting b3 (lib/oce_adv_tra_fct.F90:394-471), flux divergence
// (lib/oce_adv_tra_driver.F90:298-450) and the update
// values += (dttf_h + dttf_v)/hnode_new (src/fesom_analytic.F90:419-437)
template <class R>
__global__ __launch_bounds__(kBlock) void k_node_update(DMesh<R> m, double *__restrict__ ttf,
                                                        const double *__restrict__ LO,
                                                        const R *__restrict__ adf_h,
                                                        const R *__restrict__ adf_v,
                                                        const R *__restrict__ Rp,
                                                        const R *__restrict__ Rm, double dt)
{
    const int n = wave_column();
    if (n >= m.N) return;
    const int k = lane_id(), nz = k + 1;
    const int nlev = m.nlev_n[n];
    const size_t c0 = (size_t)n * m.S;
    const bool layer = nz <= nlev - 1;

    R rp = 0, rm = 0, fv = 0;
    if (nz <= nlev) fv = adf_v[c0 + k];
    if (layer) { rp = Rp[c0 + k]; rm = Rm[c0 + k]; }
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
    if (!layer) return;

    const double av = (double)m.areasvol[c0 + k];
    double sh = 0.0;
    const int j0 = m.n2e_ptr[n], j1 = m.n2e_ptr[n + 1];
    for (int j = j0; j < j1; ++j) {
        const int se = m.n2e[j];
        const int e = (se > 0 ? se : -se) - 1;
        if (nz > m.edge_nlmax[e]) continue;
        const int2 nn = m.edges[e];
        const R f = adf_h[(size_t)e * m.S + k];
        R ae = (R)1;
        if (f >= (R)0) {
            ae = fmn(ae, Rp[(size_t)nn.x * m.S + k]);
            ae = fmn(ae, Rm[(size_t)nn.y * m.S + k]);
        } else {
            ae = fmn(ae, Rm[(size_t)nn.x * m.S + k]);
            ae = fmn(ae, Rp[(size_t)nn.y * m.S + k]);
        }
        const double fl = (double)(ae * f);
        sh = se > 0 ? sh + fl * dt / av : sh - fl * dt / av;
    }
    const double T = ttf[c0 + k];
    double dv = 0.0 - T * (double)m.hnode[c0 + k] + LO[c0 + k] * (double)m.hnode_new[c0 + k];
    dv = dv + (double)(fvl - fvl_next) * dt / av;
    const double del = (0.0 + sh) + dv;
    ttf[c0 + k] = T + del / (double)m.hnode_new[c0 + k];
}

// lim = 'NONE': HO fluxes straight into the divergence (driver :291)
template <class R, int VER>
__global__ __launch_bounds__(kBlock) void k_node_nolim(DMesh<R> m, const R *__restrict__ WEv,
                                                       double *__restrict__ ttf,
                                                       const double *__restrict__ ttfAB,
                                                       const R *__restrict__ flux_h,
                                                       double dt, R num_ord)
{
    const int n = wave_column();
    if (n >= m.N) return;
    const int k = lane_id(), nz = k + 1;
    const int nlev = m.nlev_n[n];
    const size_t c0 = (size_t)n * m.S;
    const bool layer = nz <= nlev - 1;
    VCol<R> hc{ttfAB + c0, WEv + c0, m.area + c0, m.Zn + c0, m.zbarn + c0, m.hnode + c0, m.hnode_new + c0, nlev, nz};
    const R fv = ver_flux<R, VER>(hc, num_ord, (R)dt, (R)0);
    const R fv_next = shdn(fv);
    if (!layer) return;
    const double av = (double)m.areasvol[c0 + k];
    double sh = 0.0;
    const int j0 = m.n2e_ptr[n], j1 = m.n2e_ptr[n + 1];
    for (int j = j0; j < j1; ++j) {
        const int se = m.n2e[j];
        const int e = (se > 0 ? se : -se) - 1;
        if (nz > m.edge_nlmax[e]) continue;
        const double f = (double)flux_h[(size_t)e * m.S + k];
        sh = se > 0 ? sh + f * dt / av : sh - f * dt / av;
    }
    const double dv = 0.0 + (double)(fv - fv_next) * dt / av;
    const double del = (0.0 + sh) + dv;
    const double T = ttf[c0 + k];
    ttf[c0 + k] = T + del / (double)m.hnode_new[c0 + k];
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

template <class T>
std::vector<T> conv(const std::vector<double> &v)
{
    return std::vector<T>(v.begin(), v.end());
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

    // mesh (typed by precision)
    void *dm = nullptr;  // DMesh<R>* host copy
    DMesh<double> m64{};
    DMesh<float> m32{};
    // velocity
    void *vel = nullptr, *W = nullptr, *WE = nullptr;
    // tracers
    std::vector<double *> values, valuesAB;
    // work
    void *flux_lo = nullptr, *adf_h = nullptr, *adf_v = nullptr, *Pp = nullptr, *Pm = nullptr;
    double *LO = nullptr;
    void *trxy = nullptr, *gnod = nullptr;
    DiagAcc *dpart = nullptr, *dout = nullptr;
    int ndiag = 1024;
    float gscale = 1.0f, ginv = 1.0f;
    const HostMesh *host = nullptr;
};

template <class R>
static void build_dmesh(Engine::Impl &I, DMesh<R> &d, const HostMesh &m, int S)
{
    const int N = m.nod2D, T = m.elem2D, E = m.edge2D, nl = m.nl;
    d.N = N; d.T = T; d.E = E; d.nl = nl; d.S = S;
    d.en = I.track(dupload(m.elem2D_nodes));
    std::vector<int2> ed(E), et(E), ud(E);
    std::vector<int> nlmax(E), ea_dummy;
    std::vector<double> ea(E);
    for (int e = 0; e < E; ++e) {
        ed[e] = make_int2(m.edges[2 * e], m.edges[2 * e + 1]);
        et[e] = make_int2(m.edge_tri[2 * e], m.edge_tri[2 * e + 1]);
        ud[e] = make_int2(m.edge_up_dn_tri[2 * e], m.edge_up_dn_tri[2 * e + 1]);
        int nl1 = m.nlevels[et[e].x] - 1, nl2 = et[e].y >= 0 ? m.nlevels[et[e].y] - 1 : 0;
        nlmax[e] = std::max(nl1, nl2);
        double a = kRE * m.elem_cos[et[e].x];
        if (et[e].y >= 0) a = 0.5 * (a + kRE * m.elem_cos[et[e].y]);
        ea[e] = a;
    }
    d.edges = I.track(dupload(ed));
    d.etri = I.track(dupload(et));
    d.updn = I.track(dupload(ud));
    d.nlev_e = I.track(dupload(m.nlevels));
    d.nlev_n = I.track(dupload(m.nlevels_nod2D));
    d.nlev_n_min = I.track(dupload(m.nlevels_nod2D_min));
    d.nbl = I.track(dupload(m.nboundary_lay));
    d.edge_nlmax = I.track(dupload(nlmax));
    d.n2e_ptr = I.track(dupload(m.n2e_ptr));
    d.n2e = I.track(dupload(m.n2e));
    std::vector<int> tp(N + 1, 0), tl;
    tl.reserve((size_t)N * m.maxAdj);
    for (int n = 0; n < N; ++n) {
        for (int j = 0; j < m.nod_in_elem2D_num[n]; ++j) tl.push_back(m.nod_in_elem2D[(size_t)n * m.maxAdj + j]);
        tp[n + 1] = (int)tl.size();
    }
    d.n2t_ptr = I.track(dupload(tp));
    d.n2t = I.track(dupload(tl));
    d.ecross = I.track(dupload(conv<R>(m.edge_cross_dxdy)));
    d.edxdy = I.track(dupload(conv<R>(m.edge_dxdy)));
    d.edge_a = I.track(dupload(conv<R>(ea)));
    d.earea = I.track(dupload(conv<R>(m.elem_area)));
    d.gsca = I.track(dupload(conv<R>(m.gradient_sca)));
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
    I.host = &m;
    HIP_OK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    const size_t rs = prec_ == PREC_FP64 ? sizeof(double) : sizeof(float);
    const size_t gs = prec_ == PREC_FP64 ? sizeof(double) : (prec_ == PREC_FP32 ? sizeof(float) : sizeof(__half));
    if (prec_ == PREC_FP64) build_dmesh(I, I.m64, m, S_);
    else build_dmesh(I, I.m32, m, S_);
    auto ralloc = [&](size_t n) { return I.track((void *)dalloc<char>(n * rs)); };
    I.vel = ralloc(2 * (size_t)S_ * T_);
    I.W = ralloc((size_t)S_ * N_);
    I.WE = ralloc((size_t)S_ * N_);
    HIP_OK(hipMemset(I.vel, 0, 2 * (size_t)S_ * T_ * rs));
    HIP_OK(hipMemset(I.W, 0, (size_t)S_ * N_ * rs));
    HIP_OK(hipMemset(I.WE, 0, (size_t)S_ * N_ * rs));
    for (int t = 0; t < ntr_; ++t) {
        I.values.push_back(I.track(dalloc<double>((size_t)S_ * N_)));
        I.valuesAB.push_back(I.track(dalloc<double>((size_t)S_ * N_)));
        HIP_OK(hipMemset(I.values.back(), 0, (size_t)S_ * N_ * sizeof(double)));
        HIP_OK(hipMemset(I.valuesAB.back(), 0, (size_t)S_ * N_ * sizeof(double)));
    }
    I.flux_lo = ralloc((size_t)S_ * E_);
    I.adf_h = ralloc((size_t)S_ * E_);
    I.adf_v = ralloc((size_t)S_ * N_);
    I.Pp = ralloc((size_t)S_ * N_);
    I.Pm = ralloc((size_t)S_ * N_);
    I.LO = I.track(dalloc<double>((size_t)S_ * N_));
    I.trxy = I.track((void *)dalloc<char>(2 * (size_t)S_ * T_ * gs));
    I.gnod = I.track((void *)dalloc<char>(2 * (size_t)S_ * N_ * gs));
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
    if (prec_ == PREC_FP64) {
        auto v = pad_cols<double>(uv, L, T_, S_, 2);
        HIP_OK(hipMemcpy(I.vel, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
        auto a = pad_cols<double>(w, nl_, N_, S_);
        HIP_OK(hipMemcpy(I.W, a.data(), a.size() * sizeof(double), hipMemcpyHostToDevice));
        auto b = pad_cols<double>(we, nl_, N_, S_);
        HIP_OK(hipMemcpy(I.WE, b.data(), b.size() * sizeof(double), hipMemcpyHostToDevice));
    } else {
        auto v = pad_cols<float>(uv, L, T_, S_, 2);
        HIP_OK(hipMemcpy(I.vel, v.data(), v.size() * sizeof(float), hipMemcpyHostToDevice));
        auto a = pad_cols<float>(w, nl_, N_, S_);
        HIP_OK(hipMemcpy(I.W, a.data(), a.size() * sizeof(float), hipMemcpyHostToDevice));
        auto b = pad_cols<float>(we, nl_, N_, S_);
        HIP_OK(hipMemcpy(I.WE, b.data(), b.size() * sizeof(float), hipMemcpyHostToDevice));
    }
}

void Engine::set_values(int tr, const double *v)
{
    auto a = pad_cols<double>(v, nl_ - 1, N_, S_);
    HIP_OK(hipMemcpy(p_->values.at(tr), a.data(), a.size() * sizeof(double), hipMemcpyHostToDevice));
}

void Engine::set_valuesAB(int tr, const double *v)
{
    auto a = pad_cols<double>(v, nl_ - 1, N_, S_);
    HIP_OK(hipMemcpy(p_->valuesAB.at(tr), a.data(), a.size() * sizeof(double), hipMemcpyHostToDevice));
}

void Engine::get_values(int tr, double *v)
{
    std::vector<double> a((size_t)S_ * N_);
    HIP_OK(hipStreamSynchronize(stream_));
    HIP_OK(hipMemcpy(a.data(), p_->values.at(tr), a.size() * sizeof(double), hipMemcpyDeviceToHost));
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

template <class R, class G>
static void launch_step(Engine::Impl &I, const DMesh<R> &m, const TracerSpec &sp, double dt,
                        double *vals, const double *ab, hipStream_t st, LaunchHook hk = LaunchHook())
{
    hk.mark("start", st);
    const R *vel = (const R *)I.vel, *W = (const R *)I.W, *WE = (const R *)I.WE;
    R *flux_lo = (R *)I.flux_lo, *adf_h = (R *)I.adf_h, *adf_v = (R *)I.adf_v, *Pp = (R *)I.Pp, *Pm = (R *)I.Pm;
    G *trxy = (G *)I.trxy, *gnod = (G *)I.gnod;
    const dim3 blk(kBlock);
    const bool fct = sp.lim == LIM_FCT;
    if (sp.hor == HOR_MUSCL || sp.hor == HOR_MFCT) {
        hipLaunchKernelGGL((k_elem_grad<R, G>), dim3(grid_for(m.T)), blk, 0, st, m, ab, trxy, I.gscale);
        hk.mark("k_elem_grad", st);
        hipLaunchKernelGGL((k_node_grad<R, G>), dim3(grid_for(m.N)), blk, 0, st, m, trxy, gnod, I.gscale, I.ginv);
        hk.mark("k_node_grad", st);
    }
    const dim3 ge(grid_for(m.E)), gn(grid_for(m.N));
    const R ph = (R)sp.ph, pv = (R)sp.pv;
#define EDGE(HOR, F) hipLaunchKernelGGL((k_edge_flux<R, G, HOR, F>), ge, blk, 0, st, m, vel, vals, ab, trxy, gnod, ph, I.ginv, flux_lo, adf_h)
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
    hk.mark("k_edge_flux", st);
    if (fct) {
#define NLO(V) hipLaunchKernelGGL((k_node_lo<R, V>), gn, blk, 0, st, m, W, WE, vals, ab, flux_lo, adf_h, dt, pv, I.LO, adf_v, Pp, Pm)
        switch (sp.ver) {
        case VER_QR4C: NLO(VER_QR4C); break;
        case VER_PPM: NLO(VER_PPM); break;
        case VER_CDIFF: NLO(VER_CDIFF); break;
        default: NLO(VER_UPW1); break;
        }
#undef NLO
        hk.mark("k_node_lo", st);
        hipLaunchKernelGGL((k_node_limit<R>), gn, blk, 0, st, m, vals, I.LO, dt, Pp, Pm);
        hk.mark("k_node_limit", st);
        hipLaunchKernelGGL((k_node_update<R>), gn, blk, 0, st, m, vals, I.LO, adf_h, adf_v, Pp, Pm, dt);
        hk.mark("k_node_update", st);
    } else {
#define NNL(V) hipLaunchKernelGGL((k_node_nolim<R, V>), gn, blk, 0, st, m, WE, vals, ab, adf_h, dt, pv)
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

void Engine::enqueue_step(double dt, int first, int count)
{
    Impl &I = *p_;
    for (int t = first; t < first + count; ++t) {
        const TracerSpec &sp = spec_.at(t);
        double *vals = I.values[t];
        const double *ab = refresh_ab_ ? vals : I.valuesAB[t];
        if (prec_ == PREC_FP64) launch_step<double, double>(I, I.m64, sp, dt, vals, ab, stream_);
        else if (prec_ == PREC_FP32) launch_step<float, float>(I, I.m32, sp, dt, vals, ab, stream_);
        else launch_step<float, __half>(I, I.m32, sp, dt, vals, ab, stream_);
    }
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
    if (prec_ == PREC_FP64)
        hipLaunchKernelGGL((k_diag_partial<double>), dim3(I.ndiag), dim3(256), 0, stream_, I.m64, I.values.at(tr), I.dpart);
    else
        hipLaunchKernelGGL((k_diag_partial<float>), dim3(I.ndiag), dim3(256), 0, stream_, I.m32, I.values.at(tr), I.dpart);
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
    if (name == "adv_flux_hor") return L * E_;
    if (name == "adv_flux_lo") return L * E_;
    if (name == "adv_flux_ver") return (size_t)nl_ * N_;
    if (name == "tr_xy") return 2 * L * T_;
    if (name == "gnod") return 2 * L * N_;
    return 0;
}

int Engine::get_work(const std::string &name, int, double *out, size_t count)
{
    Impl &I = *p_;
    const size_t need = work_count(name);
    if (need == 0 || count < need) return -1;
    HIP_OK(hipStreamSynchronize(stream_));
    const void *src = nullptr;
    int L = nl_ - 1, ncol = N_, comps = 1;
    size_t es = prec_ == PREC_FP64 ? 8 : 4;
    if (name == "fct_LO") { src = I.LO; es = 8; }
    else if (name == "fct_plus") src = I.Pp;
    else if (name == "fct_minus") src = I.Pm;
    else if (name == "adv_flux_hor") { src = I.adf_h; ncol = E_; }
    else if (name == "adv_flux_lo") { src = I.flux_lo; ncol = E_; }
    else if (name == "adv_flux_ver") { src = I.adf_v; L = nl_; }
    else if (name == "tr_xy") { src = I.trxy; ncol = T_; comps = 2; es = prec_ == PREC_FP64 ? 8 : (prec_ == PREC_FP32 ? 4 : 2); }
    else if (name == "gnod") { src = I.gnod; comps = 2; es = prec_ == PREC_FP64 ? 8 : (prec_ == PREC_FP32 ? 4 : 2); }
    std::vector<char> buf((size_t)S_ * ncol * comps * es);
    HIP_OK(hipMemcpy(buf.data(), src, buf.size(), hipMemcpyDeviceToHost));
    for (int c = 0; c < ncol; ++c)
        for (int k = 0; k < L; ++k)
            for (int q = 0; q < comps; ++q) {
                size_t i = ((size_t)c * S_ + k) * comps + q;
                double v;
                if (es == 8) v = ((const double *)buf.data())[i];
                else if (es == 4) v = ((const float *)buf.data())[i];
                else {
                    v = (double)__half2float(((const __half *)buf.data())[i]) * (double)I.ginv;
                }
                out[((size_t)c * L + k) * comps + q] = v;
            }
    return 0;
}

void Engine::time_kernels(double dt, int reps, std::vector<std::string> &names, std::vector<double> &ms)
{
    // one event after every launch of a step on the engine stream; the time
    // of a kernel is the gap to the previous event (launches are serialised
    // on the one stream, so the gap is that kernel's execution)
    Impl &I = *p_;
    names.clear();
    ms.clear();
    for (int r = 0; r < reps; ++r) {
        for (int t = 0; t < ntr_; ++t) {
            std::vector<std::pair<std::string, hipEvent_t>> ev;
            LaunchHook hk;
            hk.ev = &ev;
            const TracerSpec &sp = spec_.at(t);
            double *vals = I.values[t];
            const double *ab = refresh_ab_ ? vals : I.valuesAB[t];
            if (prec_ == PREC_FP64) launch_step<double, double>(I, I.m64, sp, dt, vals, ab, stream_, hk);
            else if (prec_ == PREC_FP32) launch_step<float, float>(I, I.m32, sp, dt, vals, ab, stream_, hk);
            else launch_step<float, __half>(I, I.m32, sp, dt, vals, ab, stream_, hk);
            HIP_OK(hipStreamSynchronize(stream_));
            for (size_t i = 1; i < ev.size(); ++i) {
                float e = 0.0f;
                HIP_OK(hipEventElapsedTime(&e, ev[i - 1].second, ev[i].second));
                size_t j = 0;
                while (j < names.size() && names[j] != ev[i].first) ++j;
                if (j == names.size()) { names.push_back(ev[i].first); ms.push_back(0.0); }
                ms[j] += e;
            }
            for (auto &p : ev) HIP_OK(hipEventDestroy(p.second));
        }
    }
    for (double &v : ms) v /= (double)(reps * ntr_);
}

}  // namespace fta
