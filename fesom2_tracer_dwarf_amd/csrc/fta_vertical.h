// fta_vertical.h -- vertical fluxes of a node column, one lane per level
// interface (device helpers of the node kernels).
#pragma once

#include "fta_engine.h"
#include "fta_kernels.h"

namespace fta {

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

}  // namespace fta
