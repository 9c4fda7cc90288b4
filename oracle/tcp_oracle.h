/*
 * tcp_oracle.h -- CPU restatement (FP64) of the TCP_allpair force-and-integrate
 * path.  TEST INFRASTRUCTURE ONLY: the tests, the smoke check and the CPU
 * baseline leg of the benchmark call it as the checker.  The shipped library
 * never links it.
 *
 * Layout: particle arrays are AoS double[3*N] (x,y,z per particle), electrons
 * first (indices 0..ne-1) then ions, as the reference orders qel and qion
 * (main.f90:358-378).  Histograms mirror the reference's REAL8 arrays:
 * rdf[3*501] = RDF_ee, RDF_ei, RDF_ii and vdf[2*501] = VDF_e, VDF_i
 * (data.f90:99-101).  Return values: 0 ok, negative = error.
 */
#ifndef TCP_ORACLE_H
#define TCP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define TCPO_NBIN 501 /* Ndist+1, data.f90:16 */

enum { TCPO_NOTEMP = 0, TCPO_BEREND = 1, TCPO_ISOKIN = 2, TCPO_STOCH = 3 };
enum { TCPO_RECIP_NONE = 0, TCPO_RECIP_SPME = 1, TCPO_RECIP_EWALD = 2 };
/* regularisation options with no reference counterpart: Kelbg instead of the
 * Deutsch form, and the regularisation applied to electron-electron pairs */
enum { TCPO_KELBG = 1, TCPO_REG_EE = 2 };

typedef struct tcpo_params {
    int ne, ni;          /* NS%NTel, NS%NTion */
    int K[3];            /* param%K: SPME grid */
    int order;           /* B-spline order (4 = the reference's Force_SPME4) */
    int kmax;            /* direct Ewald k-sum extent for TCPO_RECIP_EWALD */
    int recip;           /* TCPO_RECIP_* */
    int flags;           /* TCPO_KELBG | TCPO_REG_EE, 0 = the reference's model */
    double box[3];       /* param%box */
    double q[2];         /* param%q */
    double xm[2];        /* param%xm */
    double rc;           /* param%rc */
    double rs;           /* param%rs: lambda of the e-i term */
    double alpha;        /* <= 0: 3.1/rc as force.f90:23 */
    double eps;          /* e^2/(4 pi eps0) in eV*A, data.f90:21 */
    double dx_ee, dx_ei, dx_ii; /* RDF bin widths */
    double dv_e, dv_i;   /* VDF bin widths, main.f90:286-287 */
} tcpo_params;

int tcpo_version(void);

/* EWALD_SELF (force.f90:4-49): alpha, rc2 and the self energy Uo (no eps). */
int tcpo_ewald_self(const tcpo_params* p, double* alpha, double* rc2, double* Uo);
/* PSI(K1,K2,K3) of EWALD_SELF's PME branch, Fortran order (i fastest); m=0 -> 0. */
int tcpo_psi(const tcpo_params* p, double alpha, double* psi);

/* Force_direct_self (force.f90:52-180) in the reference's loop order:
 * zeroes f, adds real-space forces, returns Ur (no eps), adds +2 per pair to rdf. */
int tcpo_force_direct_self(const tcpo_params* p, double alpha, double rc2,
                           const double* x, double* f, double* Ur, double* rdf);
/* Force_direct_ext (force.f90:183-343): own block (ne_o e, ni_o i) against an
 * external block xg (ne_g e, ni_g i); adds to f_own, writes gemit (zeroed first). */
int tcpo_force_direct_ext(const tcpo_params* p, double alpha, double rc2,
                          int ne_o, int ni_o, const double* x_own, double* f_own,
                          int ne_g, int ni_g, const double* xg, double* gemit,
                          double* Ur, double* rdf);
/* All ordered pairs, OpenMP over i, compensated per-particle sums.  Same maths
 * as tcpo_force_direct_self, a different summation order; for large N. */
int tcpo_force_allpair(const tcpo_params* p, double alpha, double rc2,
                       const double* x, double* f, double* Ur, double* rdf, int nthreads);

/* Rows of tcpo_force_allpair: for particle rows[k] (k < nrows) against every other
 * particle, f[3k..3k+2] = its real-space force and u[k] = the sum of its pair
 * energies (no eps; twice its share of Ur).  For checking a sample of a large system. */
int tcpo_force_rows(const tcpo_params* p, double alpha, double rc2, const double* x, int nrows,
                    const int* rows, double* f, double* u, int nthreads);

/* Force_SPME4 + SPME_SUM (force.f90:652-969) generalised to even order p:
 * writes the reciprocal forces into frec (overwrites) and Um (eps included). */
int tcpo_spme(const tcpo_params* p, double alpha, const double* x, double* frec, double* Um);
/* Order-p centred B-spline weights and derivatives for grid coordinate u:
 * w[i], dw[i] belong to grid point floor(u) - p/2 + 1 + i. */
int tcpo_bspline(int order, double u, double* w, double* dw);
/* |b(m)|^2 modulus of the order-p spline for m = 0..K-1 (force.f90:834-839 for p=4). */
int tcpo_bspline_modulus(int order, int K, double* bsq);
/* Direct Ewald reciprocal sum over |m_d| <= kmax, m != 0: frec (overwrite), Um. */
int tcpo_ewald_kspace(const tcpo_params* p, double alpha, const double* x, double* frec, double* Um);

/* Energy and its derivatives (no reference counterpart): U = eps (Ur + Uo) + Um,
 * dUdx = -F, dUdq[i] = dU/dq_i, *dUdl = dU/dlambda, dFdl = dF/dlambda (real-space
 * regularisation only; may be NULL).  qp: per-particle charges (NULL: species). */
int tcpo_adjoint(const tcpo_params* p, const double* x, const double* qp, double* U,
                 double* dUdx, double* dUdq, double* dUdl, double* dFdl, int nthreads);

/* REMOVE_RIGIDMTN (force.f90:974-1007). */
int tcpo_remove_rigidmtn(const tcpo_params* p, double* f);
/* VVerletNotemp1 / VVerletStoch1 (vverlet.f90:132-171 / 9-59). fric = param%alpha. */
int tcpo_vverlet1(const tcpo_params* p, int thermo, double dt, double fric,
                  double* x, double* v, const double* f);
/* VVerletNotemp2 / Berend2 / Isokin2 / Stoch2 (vverlet.f90:174,227,301,62).
 * For STOCH, eta[3*N] supplies the unit Gaussian draws (fluct, vverlet.f90:390)
 * so the update is deterministic; f then receives the noise like the reference. */
int tcpo_vverlet2(const tcpo_params* p, int thermo, double dt, double Te, double Ti,
                  double tau, double fric, double* v, double* f, const double* eta,
                  double* mv2el, double* mv2ion, double* vdf);

/* One MD step of main.f90:81-170 (kick-drift, reciprocal, direct, sum,
 * mean-force removal, kick+thermostat).  out[0..5] = Ur, Uo, Um, Epot (=(Ur+Uo)*eps),
 * mv2el, mv2ion.  fast != 0 uses tcpo_force_allpair. */
int tcpo_step(const tcpo_params* p, int thermo, double dt, double Te, double Ti,
              double tau, double fric, double* x, double* v, double* f,
              double* rdf, double* vdf, double* out, int fast, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
