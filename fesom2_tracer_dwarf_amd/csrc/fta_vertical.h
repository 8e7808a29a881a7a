// fta_vertical.h -- vertical fluxes of a node column for a lane's two
// interfaces (device helpers of the node kernels).
//
// A lane holds interfaces nz = k0+1 and k0+2 of its column (k0 even). The
// stencils reach two levels up and down, which are the lane's own cells and
// those of lanes -1/+1 (a column never straddles a wave), so each array a
// scheme reads is widened into a window of six cells k0-2 .. k0+3 by
// cross-lane moves. Window cells outside the column belong to another column
// or to padding and are masked by the level logic exactly as the reference's
// loop bounds exclude them. A column's cells are its layers: the bottom
// interface nz = nlev (zero flux) has no cell, so what a scheme would read
// there is supplied by the level logic (ppm's tv(nzmax), the callers' below()).
#pragma once

#include "fta_engine.h"
#include "fta_kernels.h"

namespace fta {

template <class T>
struct W6 {
    T v[6];  // cells k0-2 .. k0+3
};

// all lanes of the wave must call (cross-lane moves)
template <class T>
__device__ __forceinline__ W6<T> window(const V2<T> &own)
{
    W6<T> w;
    w.v[0] = from_prev(own.x);
    w.v[1] = from_prev(own.y);
    w.v[2] = own.x;
    w.v[3] = own.y;
    w.v[4] = from_next(own.x);
    w.v[5] = from_next(own.y);
    return w;
}

// tracer window, zero outside layers 1..nlev-1 (the reference's ldA range)
template <class R>
__device__ __forceinline__ W6<R> tracer_window(const V2<double> &own, int k0, int nlev)
{
    W6<double> w = window(own);
    W6<R> r;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const int k = k0 - 2 + i;  // 0-based layer index
        r.v[i] = (k >= 0 && k <= nlev - 2) ? (R)w.v[i] : (R)0;
    }
    return r;
}

template <int Q, class T> __device__ __forceinline__ T cell(const V2<T> &v) { return Q == 0 ? v.x : v.y; }

// what one lane's column offers the vertical schemes
template <class R>
struct VCol2 {
    int nlev, k0;
    W6<R> A;          // tracer the scheme reads (masked to layers)
    V2<R> W, area;    // own interfaces
    W6<R> Z;          // layer depths (QR4C)
    V2<R> zbar;       // own interface depths (QR4C)
    W6<R> h;          // hnode (PPM: upwind layer thickness)
    W6<R> hn;         // hnode_new (PPM reconstruction)
};

// adv_tra_ver_upw1 (lib/oce_adv_tra_ver.F90:248-335) at interface k0+Q+1
template <int Q, class R>
__device__ __forceinline__ R vflux_upw1(const VCol2<R> &c, R fin)
{
    const int nz = c.k0 + Q + 1;
    const R a0 = c.A.v[2 + Q], am = c.A.v[1 + Q];
    if (nz == 1) return -cell<Q>(c.W) * a0 * cell<Q>(c.area) - fin;
    if (nz == c.nlev) return (R)0 - fin;
    if (nz < c.nlev) {
        const R w = cell<Q>(c.W);
        return (R)(-0.5) * (a0 * (w + fabs(w)) + am * (w - fabs(w))) * cell<Q>(c.area) - fin;
    }
    return fin;
}

// adv_tra_ver_qr4c (lib/oce_adv_tra_ver.F90:339-444) including the double
// write of interface nzmin+1 on two-layer columns. The three slopes of
// interface nz are qc(nz) = (T(nz-1) - T(nz)) / (Z(nz-1) - Z(nz)), qu = qc(nz+1)
// and qd = qc(nz-1): the same quotient of the same operands as the
// neighbouring interface's qc, so each lane divides once per interface and
// takes the others from its own second interface or the lanes beside it
template <int Q, class R>
__device__ __forceinline__ R vflux_qr4c(const VCol2<R> &c, R num_ord, R fin, R qd, R qc, R qu)
{
    const int nz = c.k0 + Q + 1, nlev = c.nlev;
    const R t0 = c.A.v[2 + Q], tm1 = c.A.v[1 + Q];
    const R w = cell<Q>(c.W), ar = cell<Q>(c.area);
    if (nz == 1) return -t0 * w * ar - fin;
    if (nz == nlev) return (R)0 - fin;
    if (nz > nlev) return fin;
    if (nz == 2 || nz == nlev - 1) {
        const R cen = (R)(-0.5) * (tm1 + t0) * w * ar;
        if (nz == 2 && nz == nlev - 1) return cen - (cen - fin);
        return cen - fin;
    }
    // 3 <= nz <= nlev-2
    const R Z0 = c.Z.v[2 + Q], Zm1 = c.Z.v[1 + Q];
    const R zb = cell<Q>(c.zbar);
    const R Tmean1 = t0 + ((R)2 * qc + qu) * (zb - Z0) / (R)3;
    const R Tmean2 = tm1 + ((R)2 * qc + qd) * (zb - Zm1) / (R)3;
    const R Tmean = (w + fabs(w)) * Tmean1 + (w - fabs(w)) * Tmean2;
    return ((R)(-0.5) * ((R)1 - num_ord) * Tmean - num_ord * ((R)0.5 * (Tmean1 + Tmean2)) * w) * ar - fin;
}

// both interfaces of the lane; all lanes must call (cross-lane moves)
template <class R>
__device__ __forceinline__ void vflux_qr4c2(const VCol2<R> &c, R num_ord, R fin0, R fin1, R &f0, R &f1)
{
    const R qc0 = (c.A.v[1] - c.A.v[2]) / (c.Z.v[1] - c.Z.v[2]);
    const R qc1 = (c.A.v[2] - c.A.v[3]) / (c.Z.v[2] - c.Z.v[3]);
    const R qd0 = from_prev(qc1);  // qc of interface k0 (the previous lane's second)
    const R qu1 = from_next(qc0);  // qc of interface k0+3 (the next lane's first)
    f0 = vflux_qr4c<0>(c, num_ord, fin0, qd0, qc0, qc1);
    f1 = vflux_qr4c<1>(c, num_ord, fin1, qc0, qc1, qu1);
}

// adv_tra_ver_cdiff (lib/oce_adv_tra_ver.F90:648-712); interface nlev untouched
template <int Q, class R>
__device__ __forceinline__ R vflux_cdiff(const VCol2<R> &c, R fin)
{
    const int nz = c.k0 + Q + 1;
    const R w = cell<Q>(c.W), ar = cell<Q>(c.area);
    if (nz == 1) return -w * c.A.v[2 + Q] * ar - fin;
    if (nz <= c.nlev - 1) {
        const R tv = (R)0.5 * (c.A.v[1 + Q] + c.A.v[2 + Q]);
        return -tv * w * ar - fin;
    }
    return fin;
}

// PPM interface value tv(nz) at interface k0+Q+1 (lib/oce_adv_tra_ver.F90:480-540)
template <int Q, class R>
__device__ __forceinline__ R ppm_tv(const VCol2<R> &c)
{
    const int nz = c.k0 + Q + 1, nzmax = c.nlev;
    const W6<R> &A = c.A;
    if (nz == 1) return A.v[2 + Q];                      // ttf(nzmin)
    if (nz == nzmax) return A.v[1 + Q];                  // ttf(nzmax-1)
    if (nz == nzmax - 1) {
        const R sw = copysign((R)1, cell<Q>(c.W));       // W(nzmax-1)
        return -A.v[1 + Q] * fmn(sw, (R)0) + A.v[2 + Q] * fmx(sw, (R)0);
    }
    if (nz == 2) return (R)0.5 * (A.v[1 + Q] + A.v[2 + Q]);
    if (nz >= 3 && nz <= nzmax - 2) {
        // reference loop index j = nz-1 builds tv(j+1) from layers j-1..j+2
        const R dzjm1 = c.hn.v[Q], dzj = c.hn.v[1 + Q], dzjp1 = c.hn.v[2 + Q], dzjp2 = c.hn.v[3 + Q];
        const R Tm1 = A.v[Q], T0 = A.v[1 + Q], Tp1 = A.v[2 + Q], Tp2 = A.v[3 + Q];
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
        return T0 + dzj / (dzj + dzjp1) * (Tp1 - T0) +
               (R)1 / (dzjm1 + dzj + dzjp1 + dzjp2) *
               ((R)2 * dzjp1 * dzj / (dzj + dzjp1) *
                ((dzjm1 + dzj) / ((R)2 * dzj + dzjp1) - (dzjp2 + dzjp1) / ((R)2 * dzjp1 + dzj)) * (Tp1 - T0) -
                dzj * (dzjm1 + dzj) / ((R)2 * dzj + dzjp1) * deltajp1 +
                dzjp1 * (dzjp1 + dzjp2) / (dzj + (R)2 * dzjp1) * deltaj);
    }
    return (R)0;
}

// limited parabola of layer nz from its interface values aL = tv(nz), aR = tv(nz+1)
template <class R>
__device__ __forceinline__ void ppm_parabola(R a0, R &aL, R &aR, R &aj)
{
    if ((aR - a0) * (a0 - aL) <= (R)0) { aL = a0; aR = a0; }
    if ((aR - aL) * (a0 - (R)0.5 * (aL + aR)) > (aR - aL) * (aR - aL) / (R)6) aL = (R)3 * a0 - (R)2 * aR;
    if ((aR - aL) * (a0 - (R)0.5 * (aR + aL)) < -(aR - aL) * (aR - aL) / (R)6) aR = (R)3 * a0 - (R)2 * aL;
    aj = (R)6 * (a0 - (R)0.5 * (aL + aR));
}

// PPM flux through interface nz from its upwind layer (nz if W>0, nz-1 if
// W<0), which is what the reference's per-layer writes produce (its CYCLE
// only skips layers that would write nothing)
template <int Q, class R>
__device__ __forceinline__ R ppm_flux(const VCol2<R> &c, R dt, R tv, R aL, R aR, R aj, R aLu, R aRu, R aju, R fin)
{
    const int nz = c.k0 + Q + 1, nzmax = c.nlev;
    if (nz > nzmax) return fin;
    R tvert = (R)0;
    const R w = cell<Q>(c.W), ar = cell<Q>(c.area);
    if (nz <= nzmax - 1 && w > (R)0) {
        const R x = fmn(w * dt / c.h.v[2 + Q], (R)1);
        tvert = (-aL - (R)0.5 * x * (aR - aL + ((R)1 - (R)2 / (R)3 * x) * aj));
        tvert = tvert * ar * w;
    }
    if (nz >= 2 && w < (R)0) {
        const R x = fmn(-w * dt / c.h.v[1 + Q], (R)1);
        tvert = (-aRu + (R)0.5 * x * (aRu - aLu - ((R)1 - (R)2 / (R)3 * x) * aju));
        tvert = tvert * ar * w;
    }
    if (nz == 1) tvert = -tv * w * ar;  // surface: -tv(nzmin)*W*area
    if (nz == nzmax) tvert = (R)0;
    return tvert - fin;
}

// adv_tra_vert_ppm (lib/oce_adv_tra_ver.F90:448-644) at both interfaces
template <class R>
__device__ __forceinline__ void vflux_ppm(const VCol2<R> &c, R dt, R fin0, R fin1, R &f0, R &f1)
{
    const R tv0 = ppm_tv<0>(c), tv1 = ppm_tv<1>(c);
    // tv(k0+3); at the bottom interface (no cell of its own) ttf(nzmax-1), the
    // lane's second layer
    const R tvn = from_next(tv0);
    const R tv2 = c.k0 + 3 == c.nlev ? c.A.v[3] : tvn;
    R aL0 = tv0, aR0 = tv1, aj0, aL1 = tv1, aR1 = tv2, aj1;
    ppm_parabola(c.A.v[2], aL0, aR0, aj0);
    ppm_parabola(c.A.v[3], aL1, aR1, aj1);
    // layer k0 (the one above interface k0+1) sits in the previous lane's second cell
    const R aLm = from_prev(aL1), aRm = from_prev(aR1), ajm = from_prev(aj1);
    f0 = ppm_flux<0>(c, dt, tv0, aL0, aR0, aj0, aLm, aRm, ajm, fin0);
    f1 = ppm_flux<1>(c, dt, tv1, aL1, aR1, aj1, aL0, aR0, aj0, fin1);
}

// both interfaces of the lane with scheme VER; all lanes must call
template <class R, int VER>
__device__ __forceinline__ void ver_flux(const VCol2<R> &c, R num_ord, R dt, R fin0, R fin1, R &f0, R &f1)
{
    if constexpr (VER == VER_QR4C) {
        vflux_qr4c2(c, num_ord, fin0, fin1, f0, f1);
    } else if constexpr (VER == VER_PPM) {
        vflux_ppm(c, dt, fin0, fin1, f0, f1);
    } else if constexpr (VER == VER_CDIFF) {
        f0 = vflux_cdiff<0>(c, fin0);
        f1 = vflux_cdiff<1>(c, fin1);
    } else {
        f0 = vflux_upw1<0>(c, fin0);
        f1 = vflux_upw1<1>(c, fin1);
    }
}

}  // namespace fta
