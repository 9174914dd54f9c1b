/* CQBEM precompute oracle -- test infrastructure only (tests/, smoke(), bench.py cpu_baseline). */
#ifndef CQBEM_ORACLE_H
#define CQBEM_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double c1, c2, rho, nu; /* wave speeds, density, Poisson ratio (nu only for the collocation C matrix) */
    double dt, R;           /* time step and CQ contour radius */
    int L;                  /* number of CQ weights (time-history length + 1) */
    int bdf;                /* 1 = BDF1, 2 = BDF2 */
    int order, subdiv;      /* quadrature order (1..5) and uniform subdivision */
    int galerkin;           /* 0 = collocation, 1 = Galerkin */
    int mode;               /* 0 = fp64, 1 = reference precision typedefs, 2 = long double */
    int nthreads;           /* OpenMP threads, 0 = default */
} orc_params;

int orc_frequencies(int L, double R, double dt, int bdf, double *s_out);
int orc_rule_tri(int order, double *bary, double *w);
int orc_rule_tri_composite(int order, int sub, double *bary, double *w);
int orc_rule_gauss(int order, double *x, double *w);
int orc_kernel(int which, int mode, const double *x, const double *y, const double *n, double c1, double c2, double rho,
               double s_re, double s_im, double *out);
double orc_polar_inv_r(const double *y1, const double *y2, const double *y3, int order, int sub);
double orc_far_poly(const double *y1, const double *y2, const double *y3, const double *a, const double *b, int order,
                    int sub);
int orc_C_collocation(const double *V, int nv, const int *F, int nf, double nu, double *C);
int orc_C_collocation_split(const double *V, int nv, const int *F, int nf, const int *vmap, double nu, double *C);
int orc_C_galerkin(const double *V, int nv, const int *F, int nf, double *H0);
int orc_transform(const double *spec, long long M, int L, double R, double *W);
int orc_laplace(const double *V, int nv, const int *F, int nf, const int *vmap, const orc_params *p, double s_re,
                double s_im, int mask, int row0, int row1, double *U, double *P, double *Bt, double *Ba);
int orc_assemble(const double *V, int nv, const int *F, int nf, const int *vmap, const orc_params *p, int mask,
                 double *G, double *H, double *Bt, double *Ba, double *B6, double *spec_U, double *spec_P,
                 double *spec_Bt, double *spec_Ba);

#ifdef __cplusplus
}
#endif
#endif
