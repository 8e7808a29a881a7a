/*
 * fesom_oracle.h -- CPU FP64 restatement of FESOM2's tracer-advection step.
 *
 * TEST INFRASTRUCTURE ONLY. This is the parity oracle: tests, the smoke
 * check and the CPU-baseline leg of the benchmark call it, to compare
 * against or to time. The product library never links, loads or calls it.
 *
 * Conventions (mirror the Fortran data model of the reference):
 *   - arrays are column-major exactly as declared in the Fortran:
 *       ttf(nl-1, nod2D), W(nl, nod2D), vel(2, nl-1, elem2D),
 *       adv_flux_hor(nl-1, edge2D), edge_up_dn_grad(4, nl-1, edge2D) ...
 *     with NO padding (leading dimension exactly nl-1 or nl);
 *   - node / element / edge indices are 0-based; a missing element
 *     (boundary edge, no up/down-wind triangle) is -1;
 *   - level numbers (nlevels, ulevels, nboundary_lay ...) keep FESOM's
 *     1-based values, and every loop runs over the same 1-based level
 *     range as the reference loop it restates.
 */
#ifndef FESOM_ORACLE_H
#define FESOM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int nl, nod2D, elem2D, edge2D, maxAdj;
    const int *edges;             /* (2, E) */
    const int *edge_tri;          /* (2, E), -1 = none */
    const int *elem2D_nodes;      /* (3, T) */
    const int *nlevels, *ulevels; /* (T) */
    const int *nlevels_nod2D, *ulevels_nod2D;         /* (N) */
    const int *nlevels_nod2D_min, *ulevels_nod2D_max; /* (N) */
    const int *nod_in_elem2D_num; /* (N) */
    const int *nod_in_elem2D;     /* (maxAdj, N) */
    const double *coord_nod2D;    /* (2, N) radians (setup restatements only) */
    const double *edge_dxdy;      /* (2, E) radians */
    const double *edge_cross_dxdy;/* (4, E) metres */
    const double *elem_cos;       /* (T) */
    const double *elem_area;      /* (T) m^2 */
    const double *gradient_sca;   /* (6, T) 1/m */
    const double *area, *areasvol;/* (nl, N) m^2 */
    const double *hnode, *hnode_new; /* (nl-1, N) */
    const double *helem;          /* (nl-1, T) */
    const double *Z_3d_n;         /* (nl-1, N) */
    const double *zbar_3d_n;      /* (nl, N) */
    const int *edge_up_dn_tri;    /* (2, E), -1 = none */
    const int *nboundary_lay;     /* (N) */
} oracle_mesh;

typedef struct {
    double *fct_LO;          /* (nl-1, N) */
    double *adv_flux_hor;    /* (nl-1, E) */
    double *adv_flux_ver;    /* (nl, N) */
    double *fct_ttf_max, *fct_ttf_min, *fct_plus, *fct_minus; /* (nl-1, N) */
    double *edge_up_dn_grad; /* (4, nl-1, E) -- also the FCT scratch AUX */
    double *tr_xy;           /* (2, nl-1, T) */
    double *dttf_h, *dttf_v; /* (nl-1, N) */
    double *del_ttf;         /* (nl-1, N) */
    double *tvert_max, *tvert_min; /* (nl-1, N) FCT automatic arrays */
} oracle_work;

/* scheme codes */
enum { OR_HOR_UPW1 = 0, OR_HOR_MUSCL = 1, OR_HOR_MFCT = 2 };
enum { OR_VER_UPW1 = 0, OR_VER_QR4C = 1, OR_VER_PPM = 2, OR_VER_CDIFF = 3 };
enum { OR_LIM_NONE = 0, OR_LIM_FCT = 1 };

/* individual routines (o_init_zero semantics as in the reference) */
void oracle_tracer_gradient_elements(const oracle_mesh *m, const double *ttf, double *tr_xy);
void oracle_fill_up_dn_grad(const oracle_mesh *m, const double *tr_xy, double *edge_up_dn_grad);
void oracle_adv_tra_hor_upw1(const oracle_mesh *m, const double *vel, const double *ttf,
                             double *flux, int init_zero);
void oracle_adv_tra_hor_muscl(const oracle_mesh *m, const double *vel, const double *ttf,
                              double num_ord, double *flux, const double *edge_up_dn_grad,
                              int init_zero);
void oracle_adv_tra_hor_mfct(const oracle_mesh *m, const double *vel, const double *ttf,
                             double num_ord, double *flux, const double *edge_up_dn_grad,
                             int init_zero);
void oracle_adv_tra_ver_upw1(const oracle_mesh *m, const double *w, const double *ttf,
                             double *flux, int init_zero);
void oracle_adv_tra_ver_qr4c(const oracle_mesh *m, const double *w, const double *ttf,
                             double num_ord, double *flux, int init_zero);
void oracle_adv_tra_vert_ppm(const oracle_mesh *m, double dt, const double *w, const double *ttf,
                             double *flux, int init_zero);
void oracle_adv_tra_ver_cdiff(const oracle_mesh *m, const double *w, const double *ttf,
                              double *flux, int init_zero);
void oracle_tra_adv_fct(const oracle_mesh *m, double dt, const double *ttf, const double *lo,
                        double *adf_h, double *adf_v, oracle_work *wk);
void oracle_tra_adv_flux2dtracer(const oracle_mesh *m, double dt, double *dttf_h, double *dttf_v,
                                 const double *flux_h, const double *flux_v, int use_lo,
                                 const double *ttf, const double *lo);

/* do_oce_adv_tra: one tracer, tendencies ACCUMULATED into wk->dttf_h/dttf_v */
int oracle_do_oce_adv_tra(const oracle_mesh *m, double dt, const double *vel, const double *w,
                          const double *we, const double *ttf, const double *ttfAB,
                          int hor, int ver, int lim, double opth, double optv,
                          oracle_work *wk);

/* one step of one tracer as the analytic driver does it: zero tendencies,
 * (MUSCL/MFCT prerequisites from ttfAB), do_oce_adv_tra, del_ttf, update.
 * values is updated in place. Returns 0, or -1 on an unknown scheme. */
int oracle_step_tracer(const oracle_mesh *m, double dt, const double *vel, const double *w,
                       const double *we, double *values, const double *valuesAB,
                       int hor, int ver, int lim, double opth, double optv,
                       oracle_work *wk);

/* the same, with del_ttf kept from the previous call when keep_del_ttf != 0
 * (the reference C interface's tracer_advect_step) */
int oracle_step_tracer_ex(const oracle_mesh *m, double dt, const double *vel, const double *w,
                          const double *we, double *values, const double *valuesAB,
                          int hor, int ver, int lim, double opth, double optv,
                          oracle_work *wk, int keep_del_ttf);

/* threads of the OpenMP build (n > 0 sets them); returns the count in use
 * (1 in the serial build) */
int oracle_set_threads(int n);

/* setup-time restatements */
void oracle_find_up_downwind_triangles(const oracle_mesh *m, double cyclic_x, double cyclic_y,
                                       int *edge_up_dn_tri);
void oracle_nboundary_lay(const oracle_mesh *m, int *nboundary_lay);

#ifdef __cplusplus
}
#endif
#endif
