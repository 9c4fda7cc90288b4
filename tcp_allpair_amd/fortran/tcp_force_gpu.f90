! Drop-in replacement of the reference's force.f90 and vverlet.f90: the same
! external subroutines with the same argument lists, computed by libtcp_gpu
! through ISO_C_BINDING.  Link it with the unchanged data.f90, main.f90,
! init.f90 and result.f90.
!
! The host arrays qel/qion stay authoritative: each subroutine uploads what it
! reads and downloads what it writes, so Restart/DUMP/SAMPLING (result.f90) see
! exactly the state the reference would hold.  On one GPU the program runs as
! `mpirun -np 2`: rank 0 is the PME rank of data.f90:151 and does nothing here
! (the reference's LOCAL gather/scatter are dropped on both ranks alike), rank 1
! is the only direct rank and owns the GPU context with every particle.  With
! several direct ranks the real-space ring (Force_direct_ext, main.f90:107-130)
! runs on each rank's GPU; the reciprocal part then needs every position on one
! context, which this shim does not gather, so it stops with a message.
!
! Environment: TCP_PRECISION=MIXED selects the mixed-precision pair kernel
! (FP64 otherwise).

MODULE TCP_GPU_STATE
USE DATASTR
USE TCP_GPU_BIND
IMPLICIT NONE
INTEGER(C_INT64_T), SAVE :: stoch_seed = 20240607_C_INT64_T
CONTAINS

SUBROUTINE ensure_ctx(NS, param, single)
   TYPE(NM) :: NS
   TYPE(PM) :: param
   LOGICAL, INTENT(IN) :: single
   TYPE(tcp_params_t) :: p
   INTEGER(C_INT) :: rc
   IF (C_ASSOCIATED(tcp_ctx_h)) RETURN
   p%ne = NS%Nel
   p%ni = NS%Nion
   p%K = param%K
   p%order = 4                       ! Force_SPME4
   p%kmax = 0
   p%recip = TCP_RECIP_SPME
   IF (.NOT. single) p%recip = TCP_RECIP_NONE
   p%flags = 0                       ! the reference's model: Deutsch e-i only
   p%box = param%box
   p%q = param%q
   p%xm = param%xm
   p%rc = param%rc
   p%rs = param%rs
   p%alpha = 0.0D0                   ! 3.1/rc as force.f90:23
   p%eps = eps
   p%dx_ee = NS%dx_ee
   p%dx_ei = NS%dx_ei
   p%dx_ii = NS%dx_ii
   p%dv_e = NS%dv_e
   p%dv_i = NS%dv_i
   rc = tcp_create(tcp_ctx_h, p, tcp_precision())
   IF (rc == TCP_ERR_CUTOFF) STOP " *** error at cut-off radius *** "
   CALL tcp_check(rc, 'tcp_create')
   tcp_n = NS%Nel + NS%Nion
   tcp_ne = NS%Nel
   ALLOCATE(hx(3, tcp_n), hv(3, tcp_n), hf(3, tcp_n))
END SUBROUTINE ensure_ctx

SUBROUTINE put(NS, qel, qion, wx, wv, wf)
   TYPE(NM) :: NS
   TYPE(PT) :: qel(NS%Nel), qion(NS%Nion)
   LOGICAL, INTENT(IN) :: wx, wv, wf
   INTEGER :: i
   TYPE(C_PTR) :: px, pv, pf
   px = C_NULL_PTR; pv = C_NULL_PTR; pf = C_NULL_PTR
   DO i = 1, NS%Nel
      IF (wx) hx(:, i) = qel(i)%xx
      IF (wv) hv(:, i) = qel(i)%xv
      IF (wf) hf(:, i) = qel(i)%ff
   END DO
   DO i = 1, NS%Nion
      IF (wx) hx(:, NS%Nel + i) = qion(i)%xx
      IF (wv) hv(:, NS%Nel + i) = qion(i)%xv
      IF (wf) hf(:, NS%Nel + i) = qion(i)%ff
   END DO
   IF (wx) px = C_LOC(hx)
   IF (wv) pv = C_LOC(hv)
   IF (wf) pf = C_LOC(hf)
   CALL tcp_check(tcp_set_state(tcp_ctx_h, px, pv, pf), 'tcp_set_state')
END SUBROUTINE put

SUBROUTINE get(NS, qel, qion, wx, wv, wf)
   TYPE(NM) :: NS
   TYPE(PT) :: qel(NS%Nel), qion(NS%Nion)
   LOGICAL, INTENT(IN) :: wx, wv, wf
   INTEGER :: i
   TYPE(C_PTR) :: px, pv, pf
   px = C_NULL_PTR; pv = C_NULL_PTR; pf = C_NULL_PTR
   IF (wx) px = C_LOC(hx)
   IF (wv) pv = C_LOC(hv)
   IF (wf) pf = C_LOC(hf)
   CALL tcp_check(tcp_get_state(tcp_ctx_h, px, pv, pf), 'tcp_get_state')
   DO i = 1, NS%Nel
      IF (wx) qel(i)%xx = hx(:, i)
      IF (wv) qel(i)%xv = hv(:, i)
      IF (wf) qel(i)%ff = hf(:, i)
   END DO
   DO i = 1, NS%Nion
      IF (wx) qion(i)%xx = hx(:, NS%Nel + i)
      IF (wv) qion(i)%xv = hv(:, NS%Nel + i)
      IF (wf) qion(i)%ff = hf(:, NS%Nel + i)
   END DO
END SUBROUTINE get

SUBROUTINE add_rdf(NS)
   TYPE(NM) :: NS
   hrdf = 0.0D0
   CALL tcp_check(tcp_get_histograms(tcp_ctx_h, C_LOC(hrdf), C_NULL_PTR, 1_C_INT), 'tcp_get_histograms')
   NS%RDF_ee = NS%RDF_ee + hrdf(:, 1)
   NS%RDF_ei = NS%RDF_ei + hrdf(:, 2)
   NS%RDF_ii = NS%RDF_ii + hrdf(:, 3)
END SUBROUTINE add_rdf

SUBROUTINE add_vdf(NS)
   TYPE(NM) :: NS
   hvdf = 0.0D0
   CALL tcp_check(tcp_get_histograms(tcp_ctx_h, C_NULL_PTR, C_LOC(hvdf), 1_C_INT), 'tcp_get_histograms')
   NS%VDF_e = NS%VDF_e + hvdf(:, 1)
   NS%VDF_i = NS%VDF_i + hvdf(:, 2)
END SUBROUTINE add_vdf

SUBROUTINE vv2_common(NS, qel, qion, param, sys, dt, code)
   TYPE(NM) :: NS
   TYPE(PT) :: qel(NS%Nel), qion(NS%Nion)
   TYPE(PM) :: param
   TYPE(ST) :: sys
   REAL(KIND=DP) :: dt
   INTEGER(C_INT), INTENT(IN) :: code
   REAL(C_DOUBLE) :: a, b
   CALL put(NS, qel, qion, .FALSE., .TRUE., .TRUE.)
   stoch_seed = stoch_seed + 1
   CALL tcp_check(tcp_vverlet2(tcp_ctx_h, code, dt, param%Te, param%Ti, param%tau, param%alpha, &
        stoch_seed, a, b), 'tcp_vverlet2')
   sys%mv2el = a
   sys%mv2ion = b
   CALL get(NS, qel, qion, .FALSE., .TRUE., code == TCP_STOCH)
   CALL add_vdf(NS)
END SUBROUTINE vv2_common

END MODULE TCP_GPU_STATE

! ---------------------------------------------------------------- force.f90 --

SUBROUTINE EWALD_SELF(NS, param, sys, PSI, COMM)
USE DATASTR
USE TCP_GPU_STATE
IMPLICIT NONE
TYPE(NM):: NS
TYPE(PM):: param
TYPE(ST):: sys
TYPE(MP):: COMM
REAL(KIND=DP) :: PSI(param%K(1), param%K(2), param%K(3))
REAL(C_DOUBLE) :: a, rc2, uo
REAL(KIND=DP) :: rc_cut
rc_cut = MIN(MIN(param%box(1), param%box(2)), param%box(3))/2.0D0
IF (param%rc > rc_cut) THEN
   CALL MPI_Finalize(COMM%rc)
   STOP " *** error at cut-off radius *** "
END IF
param%a = 3.1D0/param%rc
param%rc2 = param%rc**2
PSI = 0.0D0
IF (COMM%TAG) THEN
   CALL ensure_ctx(NS, param, COMM%NUMPROC == 1)
   CALL tcp_check(tcp_ewald_self(tcp_ctx_h, a, rc2, uo, C_NULL_PTR), 'tcp_ewald_self')
   sys%Uo = uo
END IF
END SUBROUTINE EWALD_SELF

SUBROUTINE Force_direct_self(NS, qel, qion, param, sys)
USE DATASTR
USE TCP_GPU_STATE
IMPLICIT NONE
TYPE(NM):: NS
TYPE(PT):: qel(NS%Nel), qion(NS%Nion)
TYPE(PM):: param
TYPE(ST):: sys
REAL(C_DOUBLE) :: ur
CALL put(NS, qel, qion, .TRUE., .FALSE., .FALSE.)
CALL tcp_check(tcp_force_direct(tcp_ctx_h, ur), 'tcp_force_direct')
sys%Ur = ur
CALL get(NS, qel, qion, .FALSE., .FALSE., .TRUE.)
CALL add_rdf(NS)
END SUBROUTINE Force_direct_self

SUBROUTINE Force_direct_ext(NS, qel, qion, param, grec, gemit, Ur, Ne, Ni)
USE DATASTR
USE TCP_GPU_STATE
IMPLICIT NONE
TYPE(NM):: NS
TYPE(PT):: qel(NS%Nel), qion(NS%Nion)
TYPE(GH):: grec(NS%Ng), gemit(NS%Ng)
TYPE(PM):: param
REAL(KIND=DP):: Ur
INTEGER :: Ne, Ni
REAL(C_DOUBLE), ALLOCATABLE, TARGET :: xg(:,:), ge(:,:)
REAL(C_DOUBLE) :: u
INTEGER :: j
ALLOCATE(xg(3, MAX(Ne + Ni, 1)), ge(3, MAX(Ne + Ni, 1)))
DO j = 1, Ne + Ni
   xg(:, j) = grec(j)%xx
END DO
CALL put(NS, qel, qion, .TRUE., .FALSE., .TRUE.)
CALL tcp_check(tcp_force_direct_ext(tcp_ctx_h, C_LOC(xg), INT(Ne, C_INT), INT(Ni, C_INT), &
     C_LOC(ge), u), 'tcp_force_direct_ext')
Ur = u
DO j = 1, NS%Ng
   gemit(j)%xx = 0.0D0
END DO
DO j = 1, Ne + Ni
   gemit(j)%xx = ge(:, j)
END DO
CALL get(NS, qel, qion, .FALSE., .FALSE., .TRUE.)
CALL add_rdf(NS)
DEALLOCATE(xg, ge)
END SUBROUTINE Force_direct_ext

! ff += grec for the returned reaction block (host arithmetic, force.f90:358-369)
SUBROUTINE Force_direct_sum(NS, qel, qion, grec)
USE DATASTR
IMPLICIT NONE
TYPE(NM):: NS
TYPE(PT):: qel(NS%Nel), qion(NS%Nion)
TYPE(GH):: grec(NS%Ng)
INTEGER :: i
DO i = 1, NS%Nel
   qel(i)%ff = qel(i)%ff + grec(i)%xx
END DO
DO i = 1, NS%Nion
   qion(i)%ff = qion(i)%ff + grec(NS%Nel + i)%xx
END DO
END SUBROUTINE Force_direct_sum

! Positions reach the GPU in SPME_SUM; the LOCAL gather to the PME rank is not needed.
SUBROUTINE Force_SPME4(NS, param, sys, COMM, PSI, gsend, pme)
USE DATASTR
IMPLICIT NONE
TYPE(NM):: NS
TYPE(PM):: param
TYPE(ST):: sys
TYPE(MP):: COMM
TYPE(GH):: gsend(NS%Ng), pme(NS%NLpt)
REAL(KIND=DP) :: PSI(param%K(1), param%K(2), param%K(3))
IF (COMM%TAG .AND. COMM%NUMPROC > 1) &
     STOP " *** GPU shim: the reciprocal part needs a single direct rank (mpirun -np 2) *** "
END SUBROUTINE Force_SPME4

SUBROUTINE SPME_SUM(NS, qel, qion, sys, COMM, pme, grec)
USE DATASTR
USE TCP_GPU_STATE
IMPLICIT NONE
TYPE(NM):: NS
TYPE(PT):: qel(NS%Nel), qion(NS%Nion)
TYPE(ST):: sys
TYPE(MP):: COMM
TYPE(GH):: grec(NS%Ng), pme(NS%NLpt)
REAL(C_DOUBLE) :: um
IF (.NOT. COMM%TAG) RETURN
CALL put(NS, qel, qion, .TRUE., .FALSE., .TRUE.)
CALL tcp_check(tcp_force_recip(tcp_ctx_h, um), 'tcp_force_recip')
sys%Um = um
CALL get(NS, qel, qion, .FALSE., .FALSE., .TRUE.)
END SUBROUTINE SPME_SUM

SUBROUTINE REMOVE_RIGIDMTN(NS, qel, qion, COMM)
USE DATASTR
USE TCP_GPU_STATE
IMPLICIT NONE
INCLUDE 'mpif.h'
TYPE(NM):: NS
TYPE(PT):: qel(NS%Nel), qion(NS%Nion)
TYPE(MP):: COMM
REAL(KIND=DP) :: fl(3), fg(3)
INTEGER :: i, ierr
IF (COMM%NUMPROC == 1) THEN
   CALL put(NS, qel, qion, .FALSE., .FALSE., .TRUE.)
   CALL tcp_check(tcp_remove_rigidmtn(tcp_ctx_h), 'tcp_remove_rigidmtn')
   CALL get(NS, qel, qion, .FALSE., .FALSE., .TRUE.)
ELSE
   ! several direct ranks: the mean over all of them needs the COMM1D sum
   fl = 0.0D0
   DO i = 1, NS%Nel
      fl = fl + qel(i)%ff
   END DO
   DO i = 1, NS%Nion
      fl = fl + qion(i)%ff
   END DO
   CALL MPI_ALLREDUCE(fl, fg, 3, MPI_REAL8, MPI_SUM, COMM%COMM1D, ierr)
   fg = fg/DBLE(NS%NTpt)
   DO i = 1, NS%Nel
      qel(i)%ff = qel(i)%ff - fg
   END DO
   DO i = 1, NS%Nion
      qion(i)%ff = qion(i)%ff - fg
   END DO
END IF
END SUBROUTINE REMOVE_RIGIDMTN

! -------------------------------------------------------------- vverlet.f90 --

SUBROUTINE VVerletNotemp1(NS, qel, qion, param, dt, gsend)
USE DATASTR
USE TCP_GPU_STATE
IMPLICIT NONE
TYPE(NM)::NS
TYPE(PT)::qel(NS%Nel), qion(NS%Nion)
TYPE(GH)::gsend(NS%Ng)
TYPE(PM)::param
REAL(KIND=DP)::dt
INTEGER :: i
CALL put(NS, qel, qion, .TRUE., .TRUE., .TRUE.)
CALL tcp_check(tcp_vverlet1(tcp_ctx_h, TCP_NOTEMP, dt, 0.0D0), 'tcp_vverlet1')
CALL get(NS, qel, qion, .TRUE., .TRUE., .FALSE.)
DO i = 1, NS%Nel
   gsend(i)%xx = qel(i)%xx
END DO
DO i = 1, NS%Nion
   gsend(i + NS%Nel)%xx = qion(i)%xx
END DO
END SUBROUTINE VVerletNotemp1

SUBROUTINE VVerletStoch1(NS, qel, qion, param, dt, gsend)
USE DATASTR
USE TCP_GPU_STATE
IMPLICIT NONE
TYPE(NM)::NS
TYPE(PT)::qel(NS%Nel), qion(NS%Nion)
TYPE(GH)::gsend(NS%Ng)
TYPE(PM)::param
REAL(KIND=DP)::dt
INTEGER :: i
CALL put(NS, qel, qion, .TRUE., .TRUE., .TRUE.)
CALL tcp_check(tcp_vverlet1(tcp_ctx_h, TCP_STOCH, dt, param%alpha), 'tcp_vverlet1')
CALL get(NS, qel, qion, .TRUE., .TRUE., .FALSE.)
DO i = 1, NS%Nel
   gsend(i)%xx = qel(i)%xx
END DO
DO i = 1, NS%Nion
   gsend(i + NS%Nel)%xx = qion(i)%xx
END DO
END SUBROUTINE VVerletStoch1

SUBROUTINE VVerletNotemp2(NS, qel, qion, param, sys, dt)
USE DATASTR
USE TCP_GPU_STATE
IMPLICIT NONE
TYPE(NM)::NS
TYPE(PT)::qel(NS%Nel), qion(NS%Nion)
TYPE(PM)::param
TYPE(ST)::sys
REAL(KIND=DP)::dt
CALL vv2_common(NS, qel, qion, param, sys, dt, TCP_NOTEMP)
END SUBROUTINE VVerletNotemp2

SUBROUTINE VVerletStoch2(NS, qel, qion, param, sys, dt)
USE DATASTR
USE TCP_GPU_STATE
IMPLICIT NONE
TYPE(NM)::NS
TYPE(PT)::qel(NS%Nel), qion(NS%Nion)
TYPE(PM)::param
TYPE(ST)::sys
REAL(KIND=DP)::dt
CALL vv2_common(NS, qel, qion, param, sys, dt, TCP_STOCH)
END SUBROUTINE VVerletStoch2

SUBROUTINE VVerletBerend2(NS, qel, qion, param, sys, dt, COMM)
USE DATASTR
USE TCP_GPU_STATE
IMPLICIT NONE
TYPE(NM)::NS
TYPE(PT)::qel(NS%Nel), qion(NS%Nion)
TYPE(PM)::param
TYPE(ST)::sys
TYPE(MP)::COMM
REAL(KIND=DP)::dt
IF (COMM%NUMPROC > 1) STOP " *** GPU shim: BEREND needs a single direct rank *** "
CALL vv2_common(NS, qel, qion, param, sys, dt, TCP_BEREND)
END SUBROUTINE VVerletBerend2

SUBROUTINE VVerletIsokin2(NS, qel, qion, param, sys, dt, COMM)
USE DATASTR
USE TCP_GPU_STATE
IMPLICIT NONE
TYPE(NM)::NS
TYPE(PT)::qel(NS%Nel), qion(NS%Nion)
TYPE(PM)::param
TYPE(ST)::sys
TYPE(MP)::COMM
REAL(KIND=DP)::dt
IF (COMM%NUMPROC > 1) STOP " *** GPU shim: ISOKIN needs a single direct rank *** "
CALL vv2_common(NS, qel, qion, param, sys, dt, TCP_ISOKIN)
END SUBROUTINE VVerletIsokin2

! Unit-variance Gaussian times x (Box-Muller on RANDOM_NUMBER); kept for link
! compatibility with callers of vverlet.f90's fluct.
FUNCTION fluct(x)
USE DATASTR
IMPLICIT NONE
REAL(KIND=DP) :: fluct, x, u1, u2
CALL RANDOM_NUMBER(u1)
CALL RANDOM_NUMBER(u2)
u1 = 1.0D0 - u1
fluct = SQRT(-2.0D0*LOG(u1))*COS(2.0D0*pi*u2)*x
END FUNCTION fluct
