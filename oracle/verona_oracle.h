/*
 * verona_oracle.h -- C interface of the CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * A plain-C restatement of Verona's per-step algorithm (Wesenheit/Verona
 * 0.1.0, pure Julia) used to check the HIP library.  Only tests/, the smoke
 * check and bench.py's CPU-baseline leg load liboracle.so.
 */
#ifndef VERONA_ORACLE_H
#define VERONA_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_BC_NONE = 0, ORC_BC_PERIODIC = 1, ORC_BC_OUTFLOW = 2, ORC_BC_NOZZLE = 3 };

typedef struct {
    int ref;        /* 1: Verona semantics (HLL, midpoint RK2, BC callback at step start) */
    int rec;        /* 0 minmod, 1 PPM, 2 WENOZ */
    int riemann;    /* 0 HLL, 1 HLLC with HLL fallback */
    int integrator; /* 0 midpoint RK2, 1 SSP-RK3 */
    int n_iter;
    double tol;
    double gamma, floor_, dt, cfl; /* cfl > 0: dt from the CFL condition each step */
    double dx, dy, dz;
} orc_mode;

typedef struct {
    int type[6]; /* xlo, xhi, ylo, yhi, zlo, zhi */
    double dx, dy;
    double noz_radius, noz_rho, noz_u, noz_umag, noz_eps, noz_phi;
    const double *phi_seq; /* optional per-step nozzle angle */
} orc_bc3;

/* point functions exported for tests */
void orc_wenoz_v(const double *q5, long n, double *lo, double *hi);
void orc_ppm_v(const double *q5, long n, double *lo, double *hi);
void orc_minmod_v(const double *q3, long n, double *lo, double *hi);
void orc3_p2u_v(const double *P, long n, double g, double *U);
void orc3_p2f_v(const double *P, long n, double g, int dim, double *F);
void orc3_face_flux_v(const double *PL, const double *PR, long n, double g, int dim, int riemann,
                      double *F, double *Cmax, double *Cmin);
void orc1_face_flux_v(const double *PL, const double *PR, long n, double g, int riemann, double *F);

/* 3D */
void orc3_ptou(const double *P, double *U, int NX, int NY, int NZ, double g);
void orc3_limit(double *P, double floor_, int NX, int NY, int NZ);
void orc3_fluxes(const double *P, double g, double floor_, double *F, int dim, int NX, int NY,
                 int NZ, int rec, int riemann);
void orc3_update(const double *U, double *Ub, double dt, double dx, double dy, double dz,
                 const double *Fx, const double *Fy, const double *Fz, int NX, int NY, int NZ);
void orc3_f1dw(const double *Uc, double W, double g, double *f, double *fp);
int orc3_c2p_cell(const double *Uc, double *Pc, double g, int n_iter, double tol, int *iters);
long orc3_c2p(const double *U, double *P, int NX, int NY, int NZ, double g, int n_iter,
              double tol, long *method_counts);
long orc3_c2p_batch(const double *U, double *P, long n, double g, int n_iter, double tol,
                    int *method, int *iters);
void orc3_fill_periodic(double *P, int NX, int NY, int NZ, int mask);
void orc3_fill_outflow(double *P, int NX, int NY, int NZ, int lo_mask, int hi_mask);
void orc3_fill_nozzle(double *P, int NX, int NY, int NZ, double dx, double dy, double radius,
                      double rho, double u, double umag, double eps, double phi);
double orc3_max_rate(const double *P, int NX, int NY, int NZ, double g, double dx, double dy,
                     double dz);
long orc3_run(double *P, int NX, int NY, int NZ, const orc_mode *mode, const orc_bc3 *bc,
              int nsteps, double *t_out, double *dt_last);

/* 2D */
void orc2_ptou(const double *P, double *U, int NX, int NY, double g);
void orc2_limit(double *P, double floor_, int NX, int NY);
void orc2_fluxes(const double *P, double g, double floor_, double *F, int dim, int NX, int NY,
                 int rec, int riemann);
void orc2_update(const double *U, double *Ub, double dt, double dx, double dy, const double *Fx,
                 const double *Fy, int NX, int NY);
long orc2_c2p(const double *U, double *P, int NX, int NY, double g, int n_iter, double tol,
              long *n_singular);
void orc2_fill_periodic(double *P, int NX, int NY, int mask);
void orc2_fill_outflow(double *P, int NX, int NY, int lo_mask, int hi_mask);
double orc2_max_rate(const double *P, int NX, int NY, double g, double dx, double dy);
long orc2_run(double *P, int NX, int NY, const orc_mode *mode, const int *bc_type, int nsteps,
              double *t_out, double *dt_last);

/* 1D (SoA, no ghosts: rho[N], u[N], ux[N]) */
void orc1_f_ptou(const double *x, double g, double *f);
void orc1_jacobian(const double *x, double g, double *J);
void orc1_solve3(const double *A, const double *b, double *x);
void orc1_c2p(const double *U1, const double *U2, const double *U3, double *P1, double *P2,
              double *P3, long N, double g, int n_iter, double tol);
int orc1_run(double *P1, double *P2, double *P3, long N, double dt, double dx, double T,
             int max_steps, double g, int rec, int riemann, int n_iter, double tol,
             double *t_out);

#ifdef __cplusplus
}
#endif
#endif
