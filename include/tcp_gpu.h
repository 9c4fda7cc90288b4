/*
 * tcp_gpu.h -- C ABI of the MI355X (gfx950) force-and-integrate path of the
 * TCP_allpair two-component plasma.
 *
 * Every entry point replaces one external Fortran subroutine that main.f90 and
 * init.f90 call (file:line of the reference given per function), so an
 * ISO_C_BINDING shim with the same subroutine names can bind them.  Plain
 * pointers and sizes only.  All functions return 0 on success and a negative
 * code on error; tcp_last_error() returns the message.  A call that takes or
 * returns host data is synchronous (results are in host memory, host inputs
 * may be reused, on return); tcp_vverlet1 and tcp_remove_rigidmtn, which only
 * change device state, queue their kernels on the context's stream and
 * return, and the next call that returns host data waits for them (a fault of
 * theirs is reported there).  A context owns all device memory for its
 * lifetime, is used by one host thread at a time and runs on its own HIP
 * streams (a second one, forked from and joined to the first within one
 * tcp_md_steps step, carries the reciprocal part of systems of 65536 and
 * more particles beside the force sum).
 *
 * Host arrays are AoS double[3*N] (x,y,z per particle) with the electrons first
 * and the ions after them, the order of qel(:) followed by qion(:)
 * (main.f90:358-378).  Histograms are REAL8 arrays laid out like
 * NS%RDF_ee/_ei/_ii and NS%VDF_e/_i (data.f90:99-101), 501 bins each.
 */
#ifndef TCP_GPU_H
#define TCP_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TCP_NBIN 501 /* Ndist+1, data.f90:16 */

/* arithmetic of the pair kernel */
enum { TCP_FP64 = 0, TCP_MIXED = 1 };
/* param%thermo (main.f90:88-153) */
enum { TCP_NOTEMP = 0, TCP_BEREND = 1, TCP_ISOKIN = 2, TCP_STOCH = 3 };
/* reciprocal-space method */
enum { TCP_RECIP_NONE = 0, TCP_RECIP_SPME = 1, TCP_RECIP_EWALD = 2 };
/* short-range regularisation (tcp_params.flags).  The reference adds
 * -exp(-r/lambda)/r to electron-ion pairs only (Deutsch form, force.f90:136-139).
 * TCP_KELBG replaces it by the Kelbg form -exp(-r^2/lambda^2)/r + sqrt(pi) erfc(r/lambda)/lambda;
 * TCP_REG_EE applies the regularisation to electron-electron pairs too. */
enum { TCP_KELBG = 1, TCP_REG_EE = 2 };

/* error codes */
enum {
    TCP_OK = 0,
    TCP_ERR_ARG = -1,      /* invalid argument or size */
    TCP_ERR_CUTOFF = -2,   /* rc > min(box)/2, " *** error at cut-off radius *** " (force.f90:19-22) */
    TCP_ERR_HIP = -3,      /* a HIP runtime call failed */
    TCP_ERR_UNSUPPORTED = -4
};

/* Parameters of one plasma: TYPE(PM) (data.f90:68-74) plus the NM counts and
 * bin widths the path reads.  Same layout as the oracle's tcpo_params. */
typedef struct tcp_params {
    int32_t ne, ni;          /* NS%NTel, NS%NTion */
    int32_t K[3];            /* param%K: SPME grid (powers of two up to 512) */
    int32_t order;           /* B-spline order, even, 4..12 (4 = Force_SPME4) */
    int32_t kmax;            /* direct Ewald k-sum extent for TCP_RECIP_EWALD */
    int32_t recip;           /* TCP_RECIP_* */
    int32_t flags;           /* TCP_KELBG | TCP_REG_EE; 0 = the reference's model */
    double box[3];           /* param%box (A) */
    double q[2];             /* param%q (e) */
    double xm[2];            /* param%xm (amu) */
    double rc;               /* param%rc (A) */
    double rs;               /* param%rs: lambda of the electron-ion term (A) */
    double alpha;            /* Ewald splitting; <= 0 selects 3.1/rc as force.f90:23.  TCP_MIXED
                              * contexts need alpha*rc <= 4.5 (TCP_ERR_UNSUPPORTED beyond) */
    double eps;              /* e^2/(4 pi eps0), 14.399644154020082 eV*A (data.f90:21) */
    double dx_ee, dx_ei, dx_ii; /* NS%dx_*: RDF bin widths */
    double dv_e, dv_i;       /* NS%dv_*: VDF bin widths (main.f90:286-287) */
} tcp_params;

typedef struct tcp_ctx tcp_ctx;

/* Library setup: a context for one plasma on the current HIP device.  No
 * counterpart in the reference (its state lives in qel/qion, main.f90:335). */
int tcp_create(tcp_ctx** out, const tcp_params* p, int precision);
int tcp_destroy(tcp_ctx* ctx);
/* Text of the last error of ctx (or of the last failed tcp_create if ctx is NULL). */
const char* tcp_last_error(const tcp_ctx* ctx);

/* EWALD_SELF (force.f90:4-49): checks rc <= min(box)/2, returns alpha
 * (3.1/rc unless set), rc2 and the self energy Uo (no eps, as sys%Uo).  If psi
 * is not NULL it receives PSI(K1,K2,K3) in Fortran order with PSI(1,1,1) = 0. */
int tcp_ewald_self(tcp_ctx* ctx, double* alpha, double* rc2, double* Uo, double* psi);

/* Copies of qel/qion%xx, %xv, %ff to / from the device (any pointer may be NULL). */
int tcp_set_state(tcp_ctx* ctx, const double* x, const double* v, const double* f);
int tcp_get_state(tcp_ctx* ctx, double* x, double* v, double* f);

/* Force_direct_self (force.f90:52-180) on the whole system: f := real-space
 * Ewald + electron-ion Deutsch forces; *Ur = sys%Ur (no eps); the RDF counts
 * (+2 per pair within rc) accumulate on the device, see tcp_get_histograms. */
int tcp_force_direct(tcp_ctx* ctx, double* Ur);

/* Force_direct_ext (force.f90:183-343): the own particles (the context's
 * state) against an external block xg of ne_g electrons then ni_g ions: f +=
 * cross forces, gemit[3*(ne_g+ni_g)] := reaction forces, *Ur := cross energy.
 * RDF counts accumulate as in the reference. */
int tcp_force_direct_ext(tcp_ctx* ctx, const double* xg, int ne_g, int ni_g, double* gemit,
                         double* Ur);

/* Force_SPME4 + SPME_SUM (force.f90:652-969): f += reciprocal forces,
 * *Um = sys%Um (eps included).  With recip = TCP_RECIP_EWALD the direct
 * k-space sum replaces the mesh. */
int tcp_force_recip(tcp_ctx* ctx, double* Um);

/* REMOVE_RIGIDMTN (force.f90:974-1007): f -= mean(f) over all particles. */
int tcp_remove_rigidmtn(tcp_ctx* ctx);

/* VVerletNotemp1 / VVerletStoch1 (vverlet.f90:132-171 / 9-59): half kick,
 * drift, DNINT wrap.  fric = param%alpha (STOCH only). */
int tcp_vverlet1(tcp_ctx* ctx, int thermo, double dt, double fric);

/* VVerletNotemp2 / Berend2 / Isokin2 / Stoch2 (vverlet.f90:174,227,301,62):
 * half kick and thermostat; *mv2el, *mv2ion = sys%mv2el, sys%mv2ion; VDF
 * counts accumulate on the device.  STOCH draws its Gaussian noise from a
 * counter-based generator keyed by (seed, step counter, particle, axis). */
int tcp_vverlet2(tcp_ctx* ctx, int thermo, double dt, double Te, double Ti, double tau,
                 double fric, uint64_t seed, double* mv2el, double* mv2ion);

/* VVerletBerend2 / VVerletIsokin2 (vverlet.f90:227, 301) split at their
 * MPI_ALLREDUCE over COMM1D (vverlet.f90:254, 278), for several direct ranks:
 * tcp_vverlet2_kick does the half kick and returns the local pre-rescale sums
 * sum m v^2 per species (vverlet.f90:246-253); the caller reduces them over
 * the ranks and forms lambda, and tcp_vverlet2_scale multiplies the velocities
 * by lambda per species, counts the VDF and returns the local post-rescale
 * sums sys%mv2el, sys%mv2ion (vverlet.f90:258-269). */
int tcp_vverlet2_kick(tcp_ctx* ctx, double dt, double* mv2el, double* mv2ion);
int tcp_vverlet2_scale(tcp_ctx* ctx, double lam_e, double lam_i, double* mv2el, double* mv2ion);

/* Adds the device RDF/VDF counts into rdf[3*501] and vdf[2*501] (either may be
 * NULL) and zeroes them on the device when reset != 0: the NS%RDF_* / NS%VDF_*
 * accumulation read by DUMP (result.f90:226-268). */
int tcp_get_histograms(tcp_ctx* ctx, double* rdf, double* vdf, int reset);

/* nsteps of the MD step of main.f90:87-153 entirely on the device: VVerlet*1,
 * reciprocal part, Force_direct_self, SPME_SUM, Epot = (Ur+Uo)*eps,
 * REMOVE_RIGIDMTN, VVerlet*2.  trace (may be NULL) receives 6 doubles per step:
 * Ur, Uo, Um, Epot, mv2el, mv2ion. */
int tcp_md_steps(tcp_ctx* ctx, int nsteps, int thermo, double dt, double Te, double Ti,
                 double tau, double fric, uint64_t seed, double* trace);

/* Sets lambda (param%rs) of the regularised pairs.  No counterpart in the
 * reference, which reads param%rs once from tcp.prm (main.f90:260). */
int tcp_set_lambda(tcp_ctx* ctx, double lambda);

/* Total potential energy and its derivatives at the current positions, for
 * fitting charges and lambda to reference forces or energies.  No counterpart
 * in the reference.  Per-particle charges q[N] (NULL: the species charges
 * param%q) replace the species charges in every term, including the self
 * energy -alpha/sqrt(pi) sum q_i^2 and the reciprocal part; always FP64.
 *   *U      = eps (Ur + Uo) + Um, sys%Epot (main.f90:138) + sys%Um
 *   dUdx[3N] = -F (real-space + reciprocal forces)
 *   dUdq[N]  = dU/dq_i
 *   *dUdl    = dU/dlambda
 *   dFdl[3N] = dF/dlambda (may be NULL); lambda enters only the regularised pairs.
 * Host arrays; the context's forces f are left unchanged. */
int tcp_energy_adjoint(tcp_ctx* ctx, const double* q, double* U, double* dUdx, double* dUdq,
                       double* dUdl, double* dFdl);

/* Device time in milliseconds of the pair kernel's last launch and of the
 * last tcp_md_steps call (HIP events on the context's stream). */
int tcp_last_timings(tcp_ctx* ctx, double* pair_ms, double* steps_ms);

/* ------------------------------------------------------ no allocation ---
 * Force_direct_self (force.f90:52-180) as tcp_force_direct computes it (same
 * kernels, block shape and summation order), with NO device allocation: no
 * context, and every buffer beyond the caller's inputs and outputs in one
 * static device array of the library (about 8.6 MB per GPU, held from the
 * library's first use on that GPU to the end of the process).  No counterpart
 * in the reference: this entry point is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops).
 *   p, precision: as tcp_create; at most TCP_NOALLOC_MAX_N particles
 *   x, y, z [N]:  device FP64 positions, electrons first (inputs)
 *   fx, fy, fz [N]: device FP64 real-space forces (outputs, overwritten)
 *   Ur:   device double, sys%Ur (no eps) (output)
 *   rdf:  device unsigned long long[3*512] or NULL: RDF counts (+2 per pair
 *         within rc) added to bins 0..500 of each 512-entry row (entries
 *         501..511 receive the counts past Ndist+1 that the reference drops)
 *   stream: a hipStream_t (NULL: the null stream)
 * Asynchronous: the work is queued on stream and the call returns without
 * waiting.  All calls on one GPU share that GPU's array, so calls that overlap
 * on one GPU (two streams, or two host threads) race; calls queued on one
 * stream are ordered.  Errors: tcp_last_error(NULL). */
#define TCP_NOALLOC_MAX_N 8192
int tcp_force_direct_noalloc(const tcp_params* p, int precision, const double* x, const double* y,
                             const double* z, double* fx, double* fy, double* fz, double* Ur,
                             unsigned long long* rdf, void* stream);

/* ---------------------------------------------------------------- batches ---
 * Many independent plasmas of different sizes stepped together, for ensembles
 * and parameter scans.  No counterpart in the reference, which runs one system
 * per job.  Each replica behaves like its own context (same state layout, same
 * histograms); one call steps all of them with one persistent pair launch per
 * step for all replicas and the step captured in a hipGraph.  Replicas must
 * share recip and flags, and kmax (TCP_RECIP_EWALD) or the spline order
 * (TCP_RECIP_SPME: grids may differ, every K[d] a multiple of 8, else
 * TCP_ERR_UNSUPPORTED).  Every
 * replica reproduces tcp_md_steps on its own context bit for bit: in FP64
 * always, in TCP_MIXED when the batch's shape (tcp_batch_set_shape) is the
 * context's own, 0 above about 46000 particles and 3 below. */
typedef struct tcp_batch tcp_batch;

int tcp_batch_create(tcp_batch** out, int nrep, const tcp_params* p /* [nrep] */, int precision);
int tcp_batch_destroy(tcp_batch* b);
const char* tcp_batch_last_error(const tcp_batch* b);
/* tcp_set_state / tcp_get_state / tcp_get_histograms of replica rep */
int tcp_batch_set_state(tcp_batch* b, int rep, const double* x, const double* v, const double* f);
int tcp_batch_get_state(tcp_batch* b, int rep, double* x, double* v, double* f);
int tcp_batch_get_histograms(tcp_batch* b, int rep, double* rdf, double* vdf, int reset);
/* tcp_md_steps for every replica; Te[nrep], Ti[nrep] (may be NULL for
 * TCP_NOTEMP); trace (may be NULL) receives [nsteps][nrep][6] doubles. */
int tcp_batch_md_steps(tcp_batch* b, int nsteps, int thermo, double dt, const double* Te,
                       const double* Ti, double tau, double fric, uint64_t seed, double* trace);
/* Block shape of the persistent pair kernel (TCP_MIXED: 0..5 = 8x8 tiles by
 * 4 waves (the single-context kernel's), 8x8 by 8, 16x16 by 8, 4x4 by 4,
 * 16x16 by 4, 8x8 by 2, a wave taking tiles/waves i tiles at a time;
 * TCP_FP64: 0 only). */
int tcp_batch_set_shape(tcp_batch* b, int shape);
/* Block shape (an index as above) of replica rep's pair work: the batch's shape in
 * TCP_MIXED; in TCP_FP64 the replica's own context's, 0 or, for small systems, 3
 * (4x4 tiles by 4 waves), so that it keeps its standalone summation order. */
int tcp_batch_replica_shape(tcp_batch* b, int rep, int* shape);
/* Times `reps` pair launches per shape on the current positions, keeps the
 * fastest; *best = its index, ms[6] (may be NULL) = time per launch and shape. */
int tcp_batch_autotune(tcp_batch* b, int reps, int* best, double* ms);
/* pair_ms: best autotuned pair launch; steps_ms: last tcp_batch_md_steps call. */
int tcp_batch_last_timings(tcp_batch* b, double* pair_ms, double* steps_ms);

#ifdef __cplusplus
}
#endif
#endif
