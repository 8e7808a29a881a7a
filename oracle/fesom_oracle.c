/*
 * fesom_oracle.c -- CPU FP64 restatement of FESOM2's do_oce_adv_tra.
 *
 * TEST INFRASTRUCTURE ONLY (the parity oracle). Only tests, the smoke check
 * and the benchmark's CPU-baseline leg call this code, to compare against or
 * to time. The product library never links or loads it.
 *
 * Every routine restates one reference routine loop for loop, in the same
 * loop order and the same operand order, so that FP64 results agree with the
 * reference to rounding. Level loops keep the reference's 1-based bounds;
 * the macros below map (level, index) to the column-major arrays.
 * Reference files are cited relative to the reference repository root.
 *
 * Built twice (oracle/Makefile): serially, and with -fopenmp for the
 * all-cores CPU baseline. The OpenMP build splits the loops whose iterations
 * are independent, and runs the three edge-to-node scatters (the LO sum, the
 * b1 sums of FCT, the divergence into dttf_h) as gathers over each node's
 * incident edges in increasing edge index: every node receives the same
 * additions in the same order as in the serial edge loop, so both builds give
 * identical bits.
 */
#include "fesom_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define R_EARTH 6367500.0 /* lib/oce_modules.F90:24 */

/* column-major accessors, level argument 1-based, index argument 0-based */
#define L1(a, nz, i) (a)[(size_t)(i) * (size_t)(m->nl - 1) + (size_t)((nz) - 1)]
#define LI(a, nz, i) (a)[(size_t)(i) * (size_t)(m->nl) + (size_t)((nz) - 1)]
#define VEL(c, nz, el) vel[((size_t)(el) * (size_t)(m->nl - 1) + (size_t)((nz) - 1)) * 2 + (size_t)((c) - 1)]
#define GRAD(c, nz, e) grad[((size_t)(e) * (size_t)(m->nl - 1) + (size_t)((nz) - 1)) * 4 + (size_t)((c) - 1)]
#define TRXY(c, nz, el) tr_xy[((size_t)(el) * (size_t)(m->nl - 1) + (size_t)((nz) - 1)) * 2 + (size_t)((c) - 1)]

static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline double dmin(double a, double b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }
static inline int imin(int a, int b) { return a < b ? a : b; }

#ifdef _OPENMP
#include <omp.h>
#define OMP_FOR _Pragma("omp parallel for schedule(static)")
#define OMP_PARALLEL _Pragma("omp parallel")
#define OMP_FOR_IN _Pragma("omp for schedule(static)")
#else
#define OMP_FOR
#define OMP_PARALLEL
#define OMP_FOR_IN
#endif

int oracle_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

#ifdef _OPENMP
/* node -> incident edges in increasing edge index, +(e+1) where the node is
 * edges(1), -(e+1) where it is edges(2); rebuilt when the mesh changes */
static struct { const int *edges; int N, E; int *ptr, *idx; } g_n2e;

static void n2e_lists(const oracle_mesh *m, const int **ptr, const int **idx)
{
    if (g_n2e.edges != m->edges || g_n2e.N != m->nod2D || g_n2e.E != m->edge2D || !g_n2e.ptr) {
        free(g_n2e.ptr);
        free(g_n2e.idx);
        int *p = (int *)calloc((size_t)m->nod2D + 1, sizeof(int));
        int *x = (int *)malloc(sizeof(int) * (size_t)(2 * m->edge2D + 1));
        for (int e = 0; e < m->edge2D; ++e) { p[m->edges[2 * e] + 1]++; p[m->edges[2 * e + 1] + 1]++; }
        for (int n = 0; n < m->nod2D; ++n) p[n + 1] += p[n];
        int *fill = (int *)malloc(sizeof(int) * (size_t)(m->nod2D + 1));
        memcpy(fill, p, sizeof(int) * (size_t)m->nod2D);
        for (int e = 0; e < m->edge2D; ++e) {
            x[fill[m->edges[2 * e]]++] = e + 1;
            x[fill[m->edges[2 * e + 1]]++] = -(e + 1);
        }
        free(fill);
        g_n2e.edges = m->edges; g_n2e.N = m->nod2D; g_n2e.E = m->edge2D; g_n2e.ptr = p; g_n2e.idx = x;
    }
    *ptr = g_n2e.ptr;
    *idx = g_n2e.idx;
}

/* the level range nu12..nl12 of an edge over which the scatters run */
static inline void edge_scatter_range(const oracle_mesh *m, int e, int *nu12, int *nl12)
{
    int el1 = m->edge_tri[2 * e], el2 = m->edge_tri[2 * e + 1];
    int nl1 = m->nlevels[el1] - 1, nu1 = m->ulevels[el1], nl2 = 0, nu2 = 0;
    if (el2 >= 0) { nl2 = m->nlevels[el2] - 1; nu2 = m->ulevels[el2]; }
    *nl12 = nl1 > nl2 ? nl1 : nl2;
    *nu12 = nu1;
    if (nu2 > 0) *nu12 = nu1 < nu2 ? nu1 : nu2;
}
#endif

/* ------------------------------------------------------------------------
 * Elemental tracer gradient tr_xy(1:2, nz, elem). The reference declares
 * tr_xy (lib/oce_modules.F90:212) but vendors no routine that fills it; this
 * is FESOM2's tracer_gradient_elements: the P1 gradient of the field with
 * the shape-function derivatives gradient_sca (lib/oce_mesh.F90:2508-2530),
 * summed in node order 1,2,3 over the element's wet levels.
 * ---------------------------------------------------------------------- */
void oracle_tracer_gradient_elements(const oracle_mesh *m, const double *ttf, double *tr_xy)
{
    OMP_FOR
    for (int elem = 0; elem < m->elem2D; ++elem) {
        const int *en = &m->elem2D_nodes[3 * elem];
        const double *g = &m->gradient_sca[6 * elem];
        int nzmin = m->ulevels[elem], nzmax = m->nlevels[elem];
        for (int nz = nzmin; nz <= nzmax - 1; ++nz) {
            TRXY(1, nz, elem) = g[0] * L1(ttf, nz, en[0]) + g[1] * L1(ttf, nz, en[1]) + g[2] * L1(ttf, nz, en[2]);
            TRXY(2, nz, elem) = g[3] * L1(ttf, nz, en[0]) + g[4] * L1(ttf, nz, en[1]) + g[5] * L1(ttf, nz, en[2]);
        }
    }
}

/* area-weighted mean of the elemental gradients around node n at level nz:
 * the inner loop of fill_up_dn_grad (lib/oce_muscl_adv.F90:397-407) */
static void node_mean_grad(const oracle_mesh *m, const double *tr_xy, int node, int nz,
                           double *gx, double *gy)
{
    double tvol = 0.0, tx = 0.0, ty = 0.0;
    for (int k = 0; k < m->nod_in_elem2D_num[node]; ++k) {
        int elem = m->nod_in_elem2D[(size_t)node * m->maxAdj + k];
        if (m->nlevels[elem] - 1 < nz || nz < m->ulevels[elem]) continue;
        tvol = tvol + m->elem_area[elem];
        tx = tx + TRXY(1, nz, elem) * m->elem_area[elem];
        ty = ty + TRXY(2, nz, elem) * m->elem_area[elem];
    }
    *gx = tx / tvol;
    *gy = ty / tvol;
}

/* fill_up_dn_grad: lib/oce_muscl_adv.F90:361-530 */
void oracle_fill_up_dn_grad(const oracle_mesh *m, const double *tr_xy, double *grad)
{
    OMP_FOR
    for (int edge = 0; edge < m->edge2D; ++edge) {
        double gx, gy;
        int n1 = m->edges[2 * edge], n2 = m->edges[2 * edge + 1];
        int up = m->edge_up_dn_tri[2 * edge], dn = m->edge_up_dn_tri[2 * edge + 1];
        if (up >= 0 && dn >= 0) {
            int nzmin = imax(m->ulevels_nod2D_max[n1], m->ulevels_nod2D_max[n2]);
            int nzmax = imin(m->nlevels_nod2D_min[n1], m->nlevels_nod2D_min[n2]);
            for (int nz = m->ulevels_nod2D[n1]; nz <= nzmin - 1; ++nz) {   /* :390-408 */
                node_mean_grad(m, tr_xy, n1, nz, &gx, &gy);
                GRAD(1, nz, edge) = gx; GRAD(3, nz, edge) = gy;
            }
            for (int nz = m->ulevels_nod2D[n2]; nz <= nzmin - 1; ++nz) {   /* :410-428 */
                node_mean_grad(m, tr_xy, n2, nz, &gx, &gy);
                GRAD(2, nz, edge) = gx; GRAD(4, nz, edge) = gy;
            }
            for (int nz = nzmin; nz <= nzmax - 1; ++nz) {                 /* :431-437 */
                GRAD(1, nz, edge) = TRXY(1, nz, up);
                GRAD(2, nz, edge) = TRXY(1, nz, dn);
                GRAD(3, nz, edge) = TRXY(2, nz, up);
                GRAD(4, nz, edge) = TRXY(2, nz, dn);
            }
            for (int nz = nzmax; nz <= m->nlevels_nod2D[n1] - 1; ++nz) {  /* :440-458 */
                node_mean_grad(m, tr_xy, n1, nz, &gx, &gy);
                GRAD(1, nz, edge) = gx; GRAD(3, nz, edge) = gy;
            }
            for (int nz = nzmax; nz <= m->nlevels_nod2D[n2] - 1; ++nz) {  /* :461-479 */
                node_mean_grad(m, tr_xy, n2, nz, &gx, &gy);
                GRAD(2, nz, edge) = gx; GRAD(4, nz, edge) = gy;
            }
        } else {                                                          /* :483-520 */
            for (int nz = m->ulevels_nod2D[n1]; nz <= m->nlevels_nod2D[n1] - 1; ++nz) {
                node_mean_grad(m, tr_xy, n1, nz, &gx, &gy);
                GRAD(1, nz, edge) = gx; GRAD(3, nz, edge) = gy;
            }
            for (int nz = m->ulevels_nod2D[n2]; nz <= m->nlevels_nod2D[n2] - 1; ++nz) {
                node_mean_grad(m, tr_xy, n2, nz, &gx, &gy);
                GRAD(2, nz, edge) = gx; GRAD(4, nz, edge) = gy;
            }
        }
    }
}

/* ------------------------------------------------------------------------
 * Horizontal fluxes. Segment structure (A) nu1..nu12-1, (B) nu2..nu12-1,
 * (C) nu12..nl12, (D) nl12+1..nl1, (E) nl12+1..nl2 with
 * nl12 = min(nl1,nl2), nu12 = max(nu1,nu2) (lib/oce_adv_tra_hor.F90:163-253).
 * ---------------------------------------------------------------------- */
typedef struct { int n1, n2, el1, el2, nl1, nu1, nl2, nu2, nl12, nu12; double dx1, dy1, dx2, dy2, a; } edge_ctx;

static void edge_setup(const oracle_mesh *m, int edge, edge_ctx *c)
{
    c->n1 = m->edges[2 * edge]; c->n2 = m->edges[2 * edge + 1];
    c->el1 = m->edge_tri[2 * edge]; c->el2 = m->edge_tri[2 * edge + 1];
    c->nl1 = m->nlevels[c->el1] - 1;
    c->nu1 = m->ulevels[c->el1];
    c->dx1 = m->edge_cross_dxdy[4 * edge + 0];
    c->dy1 = m->edge_cross_dxdy[4 * edge + 1];
    c->a = R_EARTH * m->elem_cos[c->el1];
    c->nl2 = 0; c->nu2 = 0; c->dx2 = 0.0; c->dy2 = 0.0;
    if (c->el2 >= 0) {
        c->dx2 = m->edge_cross_dxdy[4 * edge + 2];
        c->dy2 = m->edge_cross_dxdy[4 * edge + 3];
        c->nl2 = m->nlevels[c->el2] - 1;
        c->nu2 = m->ulevels[c->el2];
        c->a = 0.5 * (c->a + R_EARTH * m->elem_cos[c->el2]);
    }
    c->nl12 = imin(c->nl1, c->nl2);
    c->nu12 = imax(c->nu1, c->nu2);
}

/* volume flux of segment 1, 2 and both (lib/oce_adv_tra_hor.F90:174,194,215-216) */
#define VFLUX1(nz) ((-VEL(2, nz, c.el1) * c.dx1 + VEL(1, nz, c.el1) * c.dy1) * L1(m->helem, nz, c.el1))
#define VFLUX2(nz) ((VEL(2, nz, c.el2) * c.dx2 - VEL(1, nz, c.el2) * c.dy2) * L1(m->helem, nz, c.el2))

/* adv_tra_hor_upw1: lib/oce_adv_tra_hor.F90:64-261 */
void oracle_adv_tra_hor_upw1(const oracle_mesh *m, const double *vel, const double *ttf,
                             double *flux, int init_zero)
{
    if (init_zero) memset(flux, 0, sizeof(double) * (size_t)(m->nl - 1) * m->edge2D);
    OMP_FOR
    for (int edge = 0; edge < m->edge2D; ++edge) {
        edge_ctx c; edge_setup(m, edge, &c);
        double vflux;
#define UPW1_FLUX(nz) L1(flux, nz, edge) = -0.5 * (L1(ttf, nz, c.n1) * (vflux + fabs(vflux)) + \
                                                   L1(ttf, nz, c.n2) * (vflux - fabs(vflux))) - L1(flux, nz, edge)
        for (int nz = c.nu1; nz <= c.nu12 - 1; ++nz) { vflux = VFLUX1(nz); UPW1_FLUX(nz); }            /* (A) */
        if (c.nu2 > 0)
            for (int nz = c.nu2; nz <= c.nu12 - 1; ++nz) { vflux = VFLUX2(nz); UPW1_FLUX(nz); }        /* (B) */
        for (int nz = c.nu12; nz <= c.nl12; ++nz) { vflux = VFLUX1(nz) + VFLUX2(nz); UPW1_FLUX(nz); } /* (C) */
        for (int nz = c.nl12 + 1; nz <= c.nl1; ++nz) { vflux = VFLUX1(nz); UPW1_FLUX(nz); }          /* (D) */
        for (int nz = c.nl12 + 1; nz <= c.nl2; ++nz) { vflux = VFLUX2(nz); UPW1_FLUX(nz); }          /* (E) */
#undef UPW1_FLUX
    }
}

/* shared body of MUSCL (c_lo from nboundary_lay) and MFCT (c_lo = 1):
 * lib/oce_adv_tra_hor.F90:265-550 and :554-846 */
static void hor_muscl_common(const oracle_mesh *m, const double *vel, const double *ttf,
                             double num_ord, double *flux, const double *grad, int init_zero,
                             int use_clo)
{
    if (init_zero) memset(flux, 0, sizeof(double) * (size_t)(m->nl - 1) * m->edge2D);
    OMP_FOR
    for (int edge = 0; edge < m->edge2D; ++edge) {
        edge_ctx c; edge_setup(m, edge, &c);
        const double ex = m->edge_dxdy[2 * edge], ey = m->edge_dxdy[2 * edge + 1];
        double vflux, Tmean1, Tmean2, cHO, clo1 = 1.0, clo2 = 1.0;
#define MUSCL_BODY(nz, VF)                                                                        \
        do {                                                                                      \
            if (use_clo) {                                                                        \
                clo1 = (m->nboundary_lay[c.n1] - (nz)) >= 0 ? 1.0 : 0.0;                          \
                clo2 = (m->nboundary_lay[c.n2] - (nz)) >= 0 ? 1.0 : 0.0;                          \
            }                                                                                     \
            Tmean2 = L1(ttf, nz, c.n2) - (2.0 * (L1(ttf, nz, c.n2) - L1(ttf, nz, c.n1)) +         \
                     ex * c.a * GRAD(2, nz, edge) + ey * R_EARTH * GRAD(4, nz, edge)) / 6.0 * clo2; \
            Tmean1 = L1(ttf, nz, c.n1) + (2.0 * (L1(ttf, nz, c.n2) - L1(ttf, nz, c.n1)) +         \
                     ex * c.a * GRAD(1, nz, edge) + ey * R_EARTH * GRAD(3, nz, edge)) / 6.0 * clo1; \
            vflux = (VF);                                                                         \
            cHO = (vflux + fabs(vflux)) * Tmean1 + (vflux - fabs(vflux)) * Tmean2;                 \
            L1(flux, nz, edge) = -0.5 * (1.0 - num_ord) * cHO - vflux * num_ord * 0.5 * (Tmean1 + Tmean2) \
                                 - L1(flux, nz, edge);                                            \
        } while (0)
        for (int nz = c.nu1; nz <= c.nu12 - 1; ++nz) MUSCL_BODY(nz, VFLUX1(nz));                  /* (A) */
        if (c.nu2 > 0)
            for (int nz = c.nu2; nz <= c.nu12 - 1; ++nz) MUSCL_BODY(nz, VFLUX2(nz));              /* (B) */
        for (int nz = c.nu12; nz <= c.nl12; ++nz) MUSCL_BODY(nz, VFLUX1(nz) + VFLUX2(nz));        /* (C) */
        for (int nz = c.nl12 + 1; nz <= c.nl1; ++nz) MUSCL_BODY(nz, VFLUX1(nz));                 /* (D) */
        for (int nz = c.nl12 + 1; nz <= c.nl2; ++nz) MUSCL_BODY(nz, VFLUX2(nz));                 /* (E) */
#undef MUSCL_BODY
    }
}

/* c_lo = real(max(sign(1, nboundary_lay(n)-nz), 0)): 1 when nz <= nboundary_lay(n)
 * (lib/oce_adv_tra_hor.F90:364-365) */
void oracle_adv_tra_hor_muscl(const oracle_mesh *m, const double *vel, const double *ttf,
                              double num_ord, double *flux, const double *grad, int init_zero)
{
    hor_muscl_common(m, vel, ttf, num_ord, flux, grad, init_zero, 1);
}

void oracle_adv_tra_hor_mfct(const oracle_mesh *m, const double *vel, const double *ttf,
                             double num_ord, double *flux, const double *grad, int init_zero)
{
    hor_muscl_common(m, vel, ttf, num_ord, flux, grad, init_zero, 0);
}

/* ------------------------------------------------------------------------
 * Vertical fluxes, interface arrays flux(nl, nod2D)
 * ---------------------------------------------------------------------- */
static void zero_vflux(const oracle_mesh *m, double *flux)
{
    memset(flux, 0, sizeof(double) * (size_t)m->nl * m->nod2D);
}

/* adv_tra_ver_upw1: lib/oce_adv_tra_ver.F90:248-335 */
void oracle_adv_tra_ver_upw1(const oracle_mesh *m, const double *W, const double *ttf,
                             double *flux, int init_zero)
{
    if (init_zero) zero_vflux(m, flux);
    OMP_FOR
    for (int n = 0; n < m->nod2D; ++n) {
        int nzmax = m->nlevels_nod2D[n], nzmin = m->ulevels_nod2D[n];
        int nz = nzmin;
        LI(flux, nz, n) = -LI(W, nz, n) * L1(ttf, nz, n) * LI(m->area, nz, n) - LI(flux, nz, n);
        nz = nzmax;
        LI(flux, nz, n) = 0.0 - LI(flux, nz, n);
        for (nz = nzmin + 1; nz <= nzmax - 1; ++nz) {
            LI(flux, nz, n) = -0.5 * (L1(ttf, nz, n) * (LI(W, nz, n) + fabs(LI(W, nz, n))) +
                                      L1(ttf, nz - 1, n) * (LI(W, nz, n) - fabs(LI(W, nz, n)))) *
                              LI(m->area, nz, n) - LI(flux, nz, n);
        }
    }
}

/* adv_tra_ver_qr4c: lib/oce_adv_tra_ver.F90:339-444. On a two-layer column
 * interface nzmin+1 == nzmax-1 is written twice, as in the reference. */
void oracle_adv_tra_ver_qr4c(const oracle_mesh *m, const double *W, const double *ttf,
                             double num_ord, double *flux, int init_zero)
{
    if (init_zero) zero_vflux(m, flux);
    OMP_FOR
    for (int n = 0; n < m->nod2D; ++n) {
        int nzmax = m->nlevels_nod2D[n], nzmin = m->ulevels_nod2D[n];
        int nz = nzmin;
        LI(flux, nz, n) = -L1(ttf, nz, n) * LI(W, nz, n) * LI(m->area, nz, n) - LI(flux, nz, n);
        nz = nzmin + 1;
        LI(flux, nz, n) = -0.5 * (L1(ttf, nz - 1, n) + L1(ttf, nz, n)) * LI(W, nz, n) * LI(m->area, nz, n) - LI(flux, nz, n);
        nz = nzmax - 1;
        LI(flux, nz, n) = -0.5 * (L1(ttf, nz - 1, n) + L1(ttf, nz, n)) * LI(W, nz, n) * LI(m->area, nz, n) - LI(flux, nz, n);
        nz = nzmax;
        LI(flux, nz, n) = 0.0 - LI(flux, nz, n);
        for (nz = nzmin + 2; nz <= nzmax - 2; ++nz) {
            const double *Z = m->Z_3d_n, *zb = m->zbar_3d_n;
            double qc = (L1(ttf, nz - 1, n) - L1(ttf, nz, n)) / (L1(Z, nz - 1, n) - L1(Z, nz, n));
            double qu = (L1(ttf, nz, n) - L1(ttf, nz + 1, n)) / (L1(Z, nz, n) - L1(Z, nz + 1, n));
            double qd = (L1(ttf, nz - 2, n) - L1(ttf, nz - 1, n)) / (L1(Z, nz - 2, n) - L1(Z, nz - 1, n));
            double Tmean1 = L1(ttf, nz, n) + (2 * qc + qu) * (LI(zb, nz, n) - L1(Z, nz, n)) / 3.0;
            double Tmean2 = L1(ttf, nz - 1, n) + (2 * qc + qd) * (LI(zb, nz, n) - L1(Z, nz - 1, n)) / 3.0;
            double w = LI(W, nz, n);
            double Tmean = (w + fabs(w)) * Tmean1 + (w - fabs(w)) * Tmean2;
            LI(flux, nz, n) = (-0.5 * (1.0 - num_ord) * Tmean - num_ord * (0.5 * (Tmean1 + Tmean2)) * w) *
                              LI(m->area, nz, n) - LI(flux, nz, n);
        }
    }
}

/* adv_tra_vert_ppm: lib/oce_adv_tra_ver.F90:448-644 (sign() as copysign) */
void oracle_adv_tra_vert_ppm(const oracle_mesh *m, double dt, const double *W, const double *ttf,
                             double *flux, int init_zero)
{
    if (init_zero) zero_vflux(m, flux);
    OMP_PARALLEL
    {
    double *tv = (double *)malloc(sizeof(double) * (size_t)(m->nl + 1));
    double *tvert = (double *)malloc(sizeof(double) * (size_t)(m->nl + 1));
    OMP_FOR_IN
    for (int n = 0; n < m->nod2D; ++n) {
        int nzmax = m->nlevels_nod2D[n], nzmin = m->ulevels_nod2D[n];
        const double *h = m->hnode_new;
#define T_(k) L1(ttf, k, n)
        tv[nzmin] = T_(nzmin);
        tv[nzmin + 1] = 0.5 * (T_(nzmin) + T_(nzmin + 1));
        tv[nzmax - 1] = -T_(nzmax - 2) * dmin(copysign(1.0, LI(W, nzmax - 1, n)), 0.0) +
                        T_(nzmax - 1) * dmax(copysign(1.0, LI(W, nzmax - 1, n)), 0.0);
        tv[nzmax] = T_(nzmax - 1);
        for (int nz = nzmin + 1; nz <= nzmax - 3; ++nz) {
            double dzjm1 = L1(h, nz - 1, n), dzj = L1(h, nz, n), dzjp1 = L1(h, nz + 1, n), dzjp2 = L1(h, nz + 2, n);
            double deltaj = dzj / (dzjm1 + dzj + dzjp1) *
                            ((2.0 * dzjm1 + dzj) / (dzjp1 + dzj) * (T_(nz + 1) - T_(nz)) +
                             (dzj + 2.0 * dzjp1) / (dzjm1 + dzj) * (T_(nz) - T_(nz - 1)));
            double deltajp1 = dzjp1 / (dzj + dzjp1 + dzjp2) *
                              ((2.0 * dzj + dzjp1) / (dzjp2 + dzjp1) * (T_(nz + 2) - T_(nz + 1)) +
                               (dzjp1 + 2.0 * dzjp2) / (dzj + dzjp1) * (T_(nz + 1) - T_(nz)));
            if ((T_(nz + 1) - T_(nz)) * (T_(nz) - T_(nz - 1)) > 0.0)
                deltaj = dmin(dmin(fabs(deltaj), 2.0 * fabs(T_(nz + 1) - T_(nz))), 2.0 * fabs(T_(nz) - T_(nz - 1))) *
                         copysign(1.0, deltaj);
            else
                deltaj = 0.0;
            if ((T_(nz + 2) - T_(nz + 1)) * (T_(nz + 1) - T_(nz)) > 0.0)
                deltajp1 = dmin(dmin(fabs(deltajp1), 2.0 * fabs(T_(nz + 2) - T_(nz + 1))), 2.0 * fabs(T_(nz + 1) - T_(nz))) *
                           copysign(1.0, deltajp1);
            else
                deltajp1 = 0.0;
            tv[nz + 1] = T_(nz) + dzj / (dzj + dzjp1) * (T_(nz + 1) - T_(nz)) +
                         1.0 / (dzjm1 + dzj + dzjp1 + dzjp2) *
                         ((2.0 * dzjp1 * dzj) / (dzj + dzjp1) *
                          ((dzjm1 + dzj) / (2.0 * dzj + dzjp1) - (dzjp2 + dzjp1) / (2.0 * dzjp1 + dzj)) * (T_(nz + 1) - T_(nz)) -
                          dzj * (dzjm1 + dzj) / (2.0 * dzj + dzjp1) * deltajp1 +
                          dzjp1 * (dzjp1 + dzjp2) / (dzj + 2.0 * dzjp1) * deltaj);
        }
        for (int k = 1; k <= nzmax; ++k) tvert[k] = 0.0;
        for (int nz = nzmin; nz <= nzmax - 1; ++nz) {
            if ((LI(W, nz, n) <= 0.0) && (LI(W, nz + 1, n) >= 0.0)) continue;
            double aL = tv[nz], aR = tv[nz + 1];
            if ((aR - T_(nz)) * (T_(nz) - aL) <= 0.0) { aL = T_(nz); aR = T_(nz); }
            if ((aR - aL) * (T_(nz) - 0.5 * (aL + aR)) > (aR - aL) * (aR - aL) / 6.0) aL = 3.0 * T_(nz) - 2.0 * aR;
            if ((aR - aL) * (T_(nz) - 0.5 * (aR + aL)) < -(aR - aL) * (aR - aL) / 6.0) aR = 3.0 * T_(nz) - 2.0 * aL;
            double dzj = L1(m->hnode, nz, n);
            double aj = 6.0 * (T_(nz) - 0.5 * (aL + aR));
            if (LI(W, nz, n) > 0.0) {
                double x = dmin(LI(W, nz, n) * dt / dzj, 1.0);
                tvert[nz] = (-aL - 0.5 * x * (aR - aL + (1.0 - 2.0 / 3.0 * x) * aj));
                tvert[nz] = tvert[nz] * LI(m->area, nz, n) * LI(W, nz, n);
            }
            if (LI(W, nz + 1, n) < 0.0) {
                double x = dmin(-LI(W, nz + 1, n) * dt / dzj, 1.0);
                tvert[nz + 1] = (-aR + 0.5 * x * (aR - aL - (1.0 - 2.0 / 3.0 * x) * aj));
                tvert[nz + 1] = tvert[nz + 1] * LI(m->area, nz + 1, n) * LI(W, nz + 1, n);
            }
        }
        tvert[nzmin] = -tv[nzmin] * LI(W, nzmin, n) * LI(m->area, nzmin, n);
        tvert[nzmax] = 0.0;
        for (int k = nzmin; k <= nzmax; ++k) LI(flux, k, n) = tvert[k] - LI(flux, k, n);
#undef T_
    }
    free(tv);
    free(tvert);
    }
}

/* adv_tra_ver_cdiff: lib/oce_adv_tra_ver.F90:648-712 */
void oracle_adv_tra_ver_cdiff(const oracle_mesh *m, const double *W, const double *ttf,
                              double *flux, int init_zero)
{
    if (init_zero) zero_vflux(m, flux);
    OMP_PARALLEL
    {
    double *tvert = (double *)malloc(sizeof(double) * (size_t)(m->nl + 1));
    OMP_FOR_IN
    for (int n = 0; n < m->nod2D; ++n) {
        int nzmax = m->nlevels_nod2D[n] - 1, nzmin = m->ulevels_nod2D[n];
        tvert[nzmin] = -LI(W, nzmin, n) * L1(ttf, nzmin, n) * LI(m->area, nzmin, n);
        tvert[nzmax + 1] = 0.0;
        for (int nz = nzmin + 1; nz <= nzmax; ++nz) {
            double tv = 0.5 * (L1(ttf, nz - 1, n) + L1(ttf, nz, n));
            tvert[nz] = -tv * LI(W, nz, n) * LI(m->area, nz, n);
        }
        for (int k = nzmin; k <= nzmax; ++k) LI(flux, k, n) = tvert[k] - LI(flux, k, n);
    }
    free(tvert);
    }
}

/* ------------------------------------------------------------------------
 * oce_tra_adv_fct: lib/oce_adv_tra_fct.F90:53-483. AUX is edge_up_dn_grad
 * reused with element index, as the driver does (oce_adv_tra_driver.F90:288).
 * ---------------------------------------------------------------------- */
#define AUX(c, nz, el) aux[((size_t)(el) * (size_t)(m->nl - 1) + (size_t)((nz) - 1)) * 4 + (size_t)((c) - 1)]
void oracle_tra_adv_fct(const oracle_mesh *m, double dt, const double *ttf, const double *lo,
                        double *adf_h, double *adf_v, oracle_work *wk)
{
    const double flux_eps = 1e-16;   /* :80 (FP64 under -fdefault-real-8) */
    const double bignumber = 1e3;    /* :81 */
    double *tmax = wk->fct_ttf_max, *tmin = wk->fct_ttf_min;
    double *plus = wk->fct_plus, *minus = wk->fct_minus;
    double *aux = wk->edge_up_dn_grad;
    double *tvert_max = wk->tvert_max, *tvert_min = wk->tvert_min;

    /* a1 (:94-114) */
    OMP_FOR
    for (int n = 0; n < m->nod2D; ++n) {
        int nu1 = m->ulevels_nod2D[n], nl1 = m->nlevels_nod2D[n];
        for (int nz = nu1; nz <= nl1 - 1; ++nz) {
            L1(tmax, nz, n) = dmax(L1(lo, nz, n), L1(ttf, nz, n));
            L1(tmin, nz, n) = dmin(L1(lo, nz, n), L1(ttf, nz, n));
        }
    }
    /* a2 (:127-147) */
    OMP_FOR
    for (int elem = 0; elem < m->elem2D; ++elem) {
        const int *en = &m->elem2D_nodes[3 * elem];
        int nu1 = m->ulevels[elem], nl1 = m->nlevels[elem];
        for (int nz = nu1; nz <= nl1 - 1; ++nz) {
            AUX(1, nz, elem) = dmax(dmax(L1(tmax, nz, en[0]), L1(tmax, nz, en[1])), L1(tmax, nz, en[2]));
            AUX(2, nz, elem) = dmin(dmin(L1(tmin, nz, en[0]), L1(tmin, nz, en[1])), L1(tmin, nz, en[2]));
        }
        if (nl1 <= m->nl - 1)
            for (int nz = nl1; nz <= m->nl - 1; ++nz) { AUX(1, nz, elem) = -bignumber; AUX(2, nz, elem) = bignumber; }
    }
    /* a3 (:162-184) */
    OMP_FOR
    for (int n = 0; n < m->nod2D; ++n) {
        int nu1 = m->ulevels_nod2D[n], nl1 = m->nlevels_nod2D[n];
        const int *nie = &m->nod_in_elem2D[(size_t)n * m->maxAdj];
        for (int nz = nu1; nz <= nl1 - 1; ++nz) {
            L1(tvert_max, nz, n) = AUX(1, nz, nie[0]);
            L1(tvert_min, nz, n) = AUX(2, nz, nie[0]);
            for (int e = 1; e < m->nod_in_elem2D_num[n]; ++e) {
                L1(tvert_max, nz, n) = dmax(L1(tvert_max, nz, n), AUX(1, nz, nie[e]));
                L1(tvert_min, nz, n) = dmin(L1(tvert_min, nz, n), AUX(2, nz, nie[e]));
            }
        }
    }
    /* admissible increments (:194-217) */
    OMP_FOR
    for (int n = 0; n < m->nod2D; ++n) {
        int nu1 = m->ulevels_nod2D[n], nl1 = m->nlevels_nod2D[n];
        L1(tmax, nu1, n) = L1(tvert_max, nu1, n) - L1(lo, nu1, n);
        L1(tmin, nu1, n) = L1(tvert_min, nu1, n) - L1(lo, nu1, n);
        for (int nz = nu1 + 1; nz <= nl1 - 2; ++nz) {
            L1(tmax, nz, n) = dmax(dmax(L1(tvert_max, nz - 1, n), L1(tvert_max, nz, n)), L1(tvert_max, nz + 1, n)) - L1(lo, nz, n);
            L1(tmin, nz, n) = dmin(dmin(L1(tvert_min, nz - 1, n), L1(tvert_min, nz, n)), L1(tvert_min, nz + 1, n)) - L1(lo, nz, n);
        }
        int nz = nl1 - 1;
        L1(tmax, nz, n) = L1(tvert_max, nz, n) - L1(lo, nz, n);
        L1(tmin, nz, n) = L1(tvert_min, nz, n) - L1(lo, nz, n);
    }
    /* b1 (:232-346) */
    OMP_FOR
    for (int n = 0; n < m->nod2D; ++n) {
        int nu1 = m->ulevels_nod2D[n], nl1 = m->nlevels_nod2D[n];
        for (int nz = nu1; nz <= nl1 - 1; ++nz) { L1(plus, nz, n) = 0.0; L1(minus, nz, n) = 0.0; }
    }
    OMP_FOR
    for (int n = 0; n < m->nod2D; ++n) {
        int nu1 = m->ulevels_nod2D[n], nl1 = m->nlevels_nod2D[n];
        for (int nz = nu1; nz <= nl1 - 1; ++nz) {
            L1(plus, nz, n) = L1(plus, nz, n) + (dmax(0.0, LI(adf_v, nz, n)) + dmax(0.0, -LI(adf_v, nz + 1, n)));
            L1(minus, nz, n) = L1(minus, nz, n) + (dmin(0.0, LI(adf_v, nz, n)) + dmin(0.0, -LI(adf_v, nz + 1, n)));
        }
    }
#ifndef _OPENMP
    for (int edge = 0; edge < m->edge2D; ++edge) {
        int n1 = m->edges[2 * edge], n2 = m->edges[2 * edge + 1];
        int el1 = m->edge_tri[2 * edge], el2 = m->edge_tri[2 * edge + 1];
        int nl1 = m->nlevels[el1] - 1, nu1 = m->ulevels[el1], nl2 = 0, nu2 = 0;
        if (el2 >= 0) { nl2 = m->nlevels[el2] - 1; nu2 = m->ulevels[el2]; }
        int nl12 = imax(nl1, nl2), nu12 = nu1;
        if (nu2 > 0) nu12 = imin(nu1, nu2);
        for (int nz = nu12; nz <= nl12; ++nz) {
            L1(plus, nz, n1) = L1(plus, nz, n1) + dmax(0.0, L1(adf_h, nz, edge));
            L1(minus, nz, n1) = L1(minus, nz, n1) + dmin(0.0, L1(adf_h, nz, edge));
        }
        for (int nz = nu12; nz <= nl12; ++nz) {
            L1(plus, nz, n2) = L1(plus, nz, n2) + dmax(0.0, -L1(adf_h, nz, edge));
            L1(minus, nz, n2) = L1(minus, nz, n2) + dmin(0.0, -L1(adf_h, nz, edge));
        }
    }
#else
    {
        const int *ep, *ei;
        n2e_lists(m, &ep, &ei);
        OMP_FOR
        for (int n = 0; n < m->nod2D; ++n)
            for (int j = ep[n]; j < ep[n + 1]; ++j) {
                int edge = (ei[j] > 0 ? ei[j] : -ei[j]) - 1, nu12, nl12;
                double sg = ei[j] > 0 ? 1.0 : -1.0;
                edge_scatter_range(m, edge, &nu12, &nl12);
                for (int nz = nu12; nz <= nl12; ++nz) {
                    L1(plus, nz, n) = L1(plus, nz, n) + dmax(0.0, sg * L1(adf_h, nz, edge));
                    L1(minus, nz, n) = L1(minus, nz, n) + dmin(0.0, sg * L1(adf_h, nz, edge));
                }
            }
    }
#endif
    /* b2 (:361-374) */
    OMP_FOR
    for (int n = 0; n < m->nod2D; ++n) {
        int nu1 = m->ulevels_nod2D[n], nl1 = m->nlevels_nod2D[n];
        for (int nz = nu1; nz <= nl1 - 1; ++nz) {
            double flux = L1(plus, nz, n) * dt / LI(m->areasvol, nz, n) + flux_eps;
            L1(plus, nz, n) = dmin(1.0, L1(tmax, nz, n) / flux);
            flux = L1(minus, nz, n) * dt / LI(m->areasvol, nz, n) - flux_eps;
            L1(minus, nz, n) = dmin(1.0, L1(tmin, nz, n) / flux);
        }
    }
    /* b3 vertical (:394-427) */
    OMP_FOR
    for (int n = 0; n < m->nod2D; ++n) {
        int nu1 = m->ulevels_nod2D[n], nl1 = m->nlevels_nod2D[n];
        int nz = nu1;
        double ae = 1.0, flux = LI(adf_v, nz, n);
        if (flux >= 0.0) ae = dmin(ae, L1(plus, nz, n)); else ae = dmin(ae, L1(minus, nz, n));
        LI(adf_v, nz, n) = ae * LI(adf_v, nz, n);
        for (nz = nu1 + 1; nz <= nl1 - 1; ++nz) {
            ae = 1.0;
            flux = LI(adf_v, nz, n);
            if (flux >= 0.0) { ae = dmin(ae, L1(minus, nz - 1, n)); ae = dmin(ae, L1(plus, nz, n)); }
            else             { ae = dmin(ae, L1(plus, nz - 1, n));  ae = dmin(ae, L1(minus, nz, n)); }
            LI(adf_v, nz, n) = ae * LI(adf_v, nz, n);
        }
    }
    /* b3 horizontal (:437-471) */
    OMP_FOR
    for (int edge = 0; edge < m->edge2D; ++edge) {
        int n1 = m->edges[2 * edge], n2 = m->edges[2 * edge + 1];
        int el1 = m->edge_tri[2 * edge], el2 = m->edge_tri[2 * edge + 1];
        int nu1 = m->ulevels[el1], nl1 = m->nlevels[el1] - 1, nl2 = 0, nu2 = 0;
        if (el2 >= 0) { nu2 = m->ulevels[el2]; nl2 = m->nlevels[el2] - 1; }
        int nl12 = imax(nl1, nl2), nu12 = nu1;
        if (nu2 > 0) nu12 = imin(nu1, nu2);
        for (int nz = nu12; nz <= nl12; ++nz) {
            double ae = 1.0, flux = L1(adf_h, nz, edge);
            if (flux >= 0.0) { ae = dmin(ae, L1(plus, nz, n1)); ae = dmin(ae, L1(minus, nz, n2)); }
            else             { ae = dmin(ae, L1(minus, nz, n1)); ae = dmin(ae, L1(plus, nz, n2)); }
            L1(adf_h, nz, edge) = ae * L1(adf_h, nz, edge);
        }
    }
}
#undef AUX

/* oce_tra_adv_flux2dtracer: lib/oce_adv_tra_driver.F90:298-450 */
void oracle_tra_adv_flux2dtracer(const oracle_mesh *m, double dt, double *dttf_h, double *dttf_v,
                                 const double *flux_h, const double *flux_v, int use_lo,
                                 const double *ttf, const double *lo)
{
    if (use_lo) {
        OMP_FOR
        for (int n = 0; n < m->nod2D; ++n) {
            int nu1 = m->ulevels_nod2D[n], nl1 = m->nlevels_nod2D[n];
            for (int nz = nu1; nz <= nl1 - 1; ++nz)
                L1(dttf_v, nz, n) = L1(dttf_v, nz, n) - L1(ttf, nz, n) * L1(m->hnode, nz, n) +
                                    L1(lo, nz, n) * L1(m->hnode_new, nz, n);
        }
    }
    OMP_FOR
    for (int n = 0; n < m->nod2D; ++n) {
        int nu1 = m->ulevels_nod2D[n], nl1 = m->nlevels_nod2D[n];
        for (int nz = nu1; nz <= nl1 - 1; ++nz)
            L1(dttf_v, nz, n) = L1(dttf_v, nz, n) + (LI(flux_v, nz, n) - LI(flux_v, nz + 1, n)) * dt / LI(m->areasvol, nz, n);
    }
#ifndef _OPENMP
    for (int edge = 0; edge < m->edge2D; ++edge) {
        int n1 = m->edges[2 * edge], n2 = m->edges[2 * edge + 1];
        int el1 = m->edge_tri[2 * edge], el2 = m->edge_tri[2 * edge + 1];
        int nl1 = m->nlevels[el1] - 1, nu1 = m->ulevels[el1], nl2 = 0, nu2 = 0;
        if (el2 >= 0) { nl2 = m->nlevels[el2] - 1; nu2 = m->ulevels[el2]; }
        int nl12 = imax(nl1, nl2), nu12 = nu1;
        if (nu2 > 0) nu12 = imin(nu1, nu2);
        for (int nz = nu12; nz <= nl12; ++nz)
            L1(dttf_h, nz, n1) = L1(dttf_h, nz, n1) + L1(flux_h, nz, edge) * dt / LI(m->areasvol, nz, n1);
        for (int nz = nu12; nz <= nl12; ++nz)
            L1(dttf_h, nz, n2) = L1(dttf_h, nz, n2) - L1(flux_h, nz, edge) * dt / LI(m->areasvol, nz, n2);
    }
#else
    {
        const int *ep, *ei;
        n2e_lists(m, &ep, &ei);
        OMP_FOR
        for (int n = 0; n < m->nod2D; ++n)
            for (int j = ep[n]; j < ep[n + 1]; ++j) {
                int edge = (ei[j] > 0 ? ei[j] : -ei[j]) - 1, nu12, nl12;
                edge_scatter_range(m, edge, &nu12, &nl12);
                if (ei[j] > 0)
                    for (int nz = nu12; nz <= nl12; ++nz)
                        L1(dttf_h, nz, n) = L1(dttf_h, nz, n) + L1(flux_h, nz, edge) * dt / LI(m->areasvol, nz, n);
                else
                    for (int nz = nu12; nz <= nl12; ++nz)
                        L1(dttf_h, nz, n) = L1(dttf_h, nz, n) - L1(flux_h, nz, edge) * dt / LI(m->areasvol, nz, n);
            }
    }
#endif
}

/* do_oce_adv_tra: lib/oce_adv_tra_driver.F90:23-294 (use_wsplit = .false.,
 * lib/MOD_DYN.F90:128, so the implicit branch :221-232 is never taken) */
int oracle_do_oce_adv_tra(const oracle_mesh *m, double dt, const double *vel, const double *w,
                          const double *we, const double *ttf, const double *ttfAB,
                          int hor, int ver, int lim, double opth, double optv,
                          oracle_work *wk)
{
    double *fct_LO = wk->fct_LO, *fh = wk->adv_flux_hor, *fv = wk->adv_flux_ver;
    if (hor < OR_HOR_UPW1 || hor > OR_HOR_MFCT || ver < OR_VER_UPW1 || ver > OR_VER_CDIFF ||
        lim < OR_LIM_NONE || lim > OR_LIM_FCT)
        return -1;
    if (lim == OR_LIM_FCT) {
        oracle_adv_tra_hor_upw1(m, vel, ttf, fh, 1);                     /* :95 */
        memset(fct_LO, 0, sizeof(double) * (size_t)(m->nl - 1) * m->nod2D); /* :102-106 */
#ifndef _OPENMP
        for (int e = 0; e < m->edge2D; ++e) {                            /* :122-181 */
            int n1 = m->edges[2 * e], n2 = m->edges[2 * e + 1];
            int el1 = m->edge_tri[2 * e], el2 = m->edge_tri[2 * e + 1];
            int nl1 = m->nlevels[el1] - 1, nu1 = m->ulevels[el1], nl2 = 0, nu2 = 0;
            if (el2 >= 0) { nl2 = m->nlevels[el2] - 1; nu2 = m->ulevels[el2]; }
            int nl12 = imax(nl1, nl2), nu12 = nu1;
            if (nu2 > 0) nu12 = imin(nu1, nu2);
            for (int nz = nu12; nz <= nl12; ++nz) {
                L1(fct_LO, nz, n1) = L1(fct_LO, nz, n1) + L1(fh, nz, e);
                L1(fct_LO, nz, n2) = L1(fct_LO, nz, n2) - L1(fh, nz, e);
            }
        }
#else
        {
            const int *ep, *ei;
            n2e_lists(m, &ep, &ei);
            OMP_FOR
            for (int n = 0; n < m->nod2D; ++n)
                for (int j = ep[n]; j < ep[n + 1]; ++j) {
                    int e = (ei[j] > 0 ? ei[j] : -ei[j]) - 1, nu12, nl12;
                    edge_scatter_range(m, e, &nu12, &nl12);
                    if (ei[j] > 0)
                        for (int nz = nu12; nz <= nl12; ++nz) L1(fct_LO, nz, n) = L1(fct_LO, nz, n) + L1(fh, nz, e);
                    else
                        for (int nz = nu12; nz <= nl12; ++nz) L1(fct_LO, nz, n) = L1(fct_LO, nz, n) - L1(fh, nz, e);
                }
        }
#endif
        oracle_adv_tra_ver_upw1(m, we, ttf, fv, 1);                      /* :194 */
        OMP_FOR
        for (int n = 0; n < m->nod2D; ++n) {                             /* :202-211 */
            int nu1 = m->ulevels_nod2D[n], nl1 = m->nlevels_nod2D[n];
            for (int nz = nu1; nz <= nl1 - 1; ++nz)
                L1(fct_LO, nz, n) = (L1(ttf, nz, n) * L1(m->hnode, nz, n) +
                                     (L1(fct_LO, nz, n) + (LI(fv, nz, n) - LI(fv, nz + 1, n))) * dt /
                                     LI(m->areasvol, nz, n)) / L1(m->hnode_new, nz, n);
        }
    }
    int do_zero_flux = (lim == OR_LIM_FCT) ? 0 : 1;                     /* :239-240 */
    switch (hor) {                                                       /* :244-255 */
    case OR_HOR_MUSCL: oracle_adv_tra_hor_muscl(m, vel, ttfAB, opth, fh, wk->edge_up_dn_grad, do_zero_flux); break;
    case OR_HOR_MFCT:  oracle_adv_tra_hor_mfct(m, vel, ttfAB, opth, fh, wk->edge_up_dn_grad, do_zero_flux); break;
    default:           oracle_adv_tra_hor_upw1(m, vel, ttfAB, fh, do_zero_flux); break;
    }
    const double *pwvel = (lim == OR_LIM_FCT) ? w : we;                  /* :257-261 */
    switch (ver) {                                                       /* :266-282 */
    case OR_VER_QR4C:  oracle_adv_tra_ver_qr4c(m, pwvel, ttfAB, optv, fv, do_zero_flux); break;
    case OR_VER_CDIFF: oracle_adv_tra_ver_cdiff(m, pwvel, ttfAB, fv, do_zero_flux); break;
    case OR_VER_PPM:   oracle_adv_tra_vert_ppm(m, dt, pwvel, ttfAB, fv, do_zero_flux); break;
    default:           oracle_adv_tra_ver_upw1(m, pwvel, ttfAB, fv, do_zero_flux); break;
    }
    if (lim == OR_LIM_FCT) {                                             /* :286-292 */
        oracle_tra_adv_fct(m, dt, ttf, fct_LO, fh, fv, wk);
        oracle_tra_adv_flux2dtracer(m, dt, wk->dttf_h, wk->dttf_v, fh, fv, 1, ttf, fct_LO);
    } else {
        oracle_tra_adv_flux2dtracer(m, dt, wk->dttf_h, wk->dttf_v, fh, fv, 0, NULL, NULL);
    }
    return 0;
}

/* the analytic driver's per-tracer step body: src/fesom_analytic.F90:405-437,
 * preceded by FESOM2's MUSCL prerequisites (gradient of ttfAB, then
 * fill_up_dn_grad) when the horizontal scheme needs them */
int oracle_step_tracer(const oracle_mesh *m, double dt, const double *vel, const double *w,
                       const double *we, double *values, const double *valuesAB,
                       int hor, int ver, int lim, double opth, double optv,
                       oracle_work *wk)
{
    return oracle_step_tracer_ex(m, dt, vel, w, we, values, valuesAB, hor, ver, lim, opth, optv, wk, 0);
}

/* keep_del_ttf = 1: the reference C interface's tracer_advect_step, which
 * zeroes the two tendencies but never del_ttf (lib/tracer_c_interface.F90:
 * 176-212), so del_ttf grows from step to step */
int oracle_step_tracer_ex(const oracle_mesh *m, double dt, const double *vel, const double *w,
                          const double *we, double *values, const double *valuesAB,
                          int hor, int ver, int lim, double opth, double optv,
                          oracle_work *wk, int keep_del_ttf)
{
    size_t nn = (size_t)(m->nl - 1) * m->nod2D;
    memset(wk->dttf_h, 0, sizeof(double) * nn);
    memset(wk->dttf_v, 0, sizeof(double) * nn);
    if (!keep_del_ttf) memset(wk->del_ttf, 0, sizeof(double) * nn);
    if (hor == OR_HOR_MUSCL || hor == OR_HOR_MFCT) {
        oracle_tracer_gradient_elements(m, valuesAB, wk->tr_xy);
        oracle_fill_up_dn_grad(m, wk->tr_xy, wk->edge_up_dn_grad);
    }
    int st = oracle_do_oce_adv_tra(m, dt, vel, w, we, values, valuesAB, hor, ver, lim, opth, optv, wk);
    if (st) return st;
    OMP_FOR
    for (int k = 0; k < m->nod2D; ++k) {
        int nzmax = m->nlevels_nod2D[k] - 1, nzmin = m->ulevels_nod2D[k];
        for (int kz = nzmin; kz <= nzmax; ++kz)
            L1(wk->del_ttf, kz, k) = L1(wk->del_ttf, kz, k) + L1(wk->dttf_h, kz, k) + L1(wk->dttf_v, kz, k);
    }
    OMP_FOR
    for (int k = 0; k < m->nod2D; ++k) {
        int nzmax = m->nlevels_nod2D[k] - 1, nzmin = m->ulevels_nod2D[k];
        for (int kz = nzmin; kz <= nzmax; ++kz)
            L1(values, kz, k) = L1(values, kz, k) + L1(wk->del_ttf, kz, k) / L1(m->hnode_new, kz, k);
    }
    return 0;
}

/* ------------------------------------------------------------------------
 * Setup restatements
 * ---------------------------------------------------------------------- */
static inline double trim_cyc(double x, double cyc)
{
    if (cyc > 0.0) {
        if (x > cyc / 2.0) x = x - cyc;
        if (x < -cyc / 2.0) x = x + cyc;
    }
    return x;
}

/* find_up_downwind_triangles: lib/oce_muscl_adv.F90:166-357. The reference
 * trims x only (cyclic_length); a doubly periodic mesh also trims y. The
 * Fortran 'cycle' continues the element loop, so the LAST matching element
 * in nod_in_elem2D order is kept. */
void oracle_find_up_downwind_triangles(const oracle_mesh *m, double cx, double cy, int *tri)
{
    const double *X = m->coord_nod2D;
    for (int e = 0; e < 2 * m->edge2D; ++e) tri[e] = -1;
    for (int n = 0; n < m->edge2D; ++n) {
        int ed[2] = {m->edges[2 * n], m->edges[2 * n + 1]};
        double x[2] = {trim_cyc(X[2 * ed[1]] - X[2 * ed[0]], cx), trim_cyc(X[2 * ed[1] + 1] - X[2 * ed[0] + 1], cy)};
        for (int side = 0; side < 2; ++side) {
            int node = ed[side];
            double xs[2] = {side == 0 ? -x[0] : x[0], side == 0 ? -x[1] : x[1]};
            for (int k = 0; k < m->nod_in_elem2D_num[node]; ++k) {
                int elem = m->nod_in_elem2D[(size_t)node * m->maxAdj + k];
                const int *en = &m->elem2D_nodes[3 * elem];
                int i0, i1, i2;
                if (en[0] == node) { i0 = 0; i1 = 1; i2 = 2; }
                else if (en[1] == node) { i0 = 1; i1 = 0; i2 = 2; }
                else { i0 = 2; i1 = 0; i2 = 1; }
                double b[2], c[2];
                for (int d = 0; d < 2; ++d) {
                    b[d] = X[2 * en[i1] + d] - X[2 * en[i0] + d];
                    c[d] = X[2 * en[i2] + d] - X[2 * en[i0] + d];
                }
                b[0] = trim_cyc(b[0], cx); c[0] = trim_cyc(c[0], cx);
                b[1] = trim_cyc(b[1], cy); c[1] = trim_cyc(c[1], cy);
                double cr = c[0] * c[0] + c[1] * c[1];
                double bx = (b[0] * c[0] + b[1] * c[1]) / cr;
                double by = (-b[0] * c[1] + b[1] * c[0]) / cr;
                double xx = (xs[0] * c[0] + xs[1] * c[1]) / cr;
                double xy = (-xs[0] * c[1] + xs[1] * c[0]) / cr;
                double ab = atan2(by, bx), ax = atan2(xy, xx);
                if ((ab > 0.0) && (ax > 0.0) && (ax < ab)) { tri[2 * n + side] = elem; continue; }
                if ((ab < 0.0) && (ax < 0.0) && (ax > ab)) { tri[2 * n + side] = elem; continue; }
                if ((ab == ax) || (ax == 0.0)) { tri[2 * n + side] = elem; continue; }
            }
        }
    }
}

/* nboundary_lay: lib/oce_muscl_adv.F90:68-104 */
void oracle_nboundary_lay(const oracle_mesh *m, int *nbl)
{
    for (int n = 0; n < m->nod2D; ++n) nbl[n] = m->nl - 1;
    for (int n = 0; n < m->edge2D; ++n) {
        int n1 = m->edges[2 * n], n2 = m->edges[2 * n + 1];
        int t1 = m->edge_tri[2 * n], t2 = m->edge_tri[2 * n + 1];
        if (t1 < 0 || t2 < 0) {
            nbl[n1] = 0; nbl[n2] = 0;
        } else {
            int mn = imin(m->nlevels[t1], m->nlevels[t2]) - 1;
            nbl[n1] = imin(nbl[n1], mn);
            nbl[n2] = imin(nbl[n2], mn);
        }
    }
}
