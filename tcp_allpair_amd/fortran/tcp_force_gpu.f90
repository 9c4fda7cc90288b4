! Drop-in replacement of the reference's force.f90 and vverlet.f90: the same
! external subroutines with the same argument lists, computed by libtcp_gpu
! through ISO_C_BINDING.  Link it with the unchanged data.f90, main.f90,
! init.f90 and result.f90.
!
! Transfers: the shim keeps a shadow of the device's x, v and f (hx, hv, hf).
! A subroutine uploads an array it reads only when the host copy differs from
! the shadow (a caller or another subroutine changed it), and downloads what it
! writes.  With one direct rank the step's results stay on the device from
! VVerlet*1 to VVerlet*2 (x and v of the kick-drift, the forces of
! Force_direct_self, SPME_SUM and REMOVE_RIGIDMTN, the RDF counts), and
! VVerlet*2 brings x, v, f and both histograms back in one call: the step of
! main.f90:87-153 reads none of them on the host in between (gsend feeds only
! the LOCAL gather of several direct ranks), and Restart/DUMP/SAMPLING
! (result.f90) run after VVerlet*2, so they see exactly the state the
! reference would hold.  Forces computed outside a step (Finit) stay on the
! device for the next VVerlet*1; their RDF counts reach NS%RDF_* at once, as
! the reference counts them before main.f90:63 clears them.
! TCP_DROPIN_SYNC=1 downloads after every subroutine instead.  With one direct rank
! (`mpirun -np 2`) rank 0, the PME rank of data.f90:151, does nothing (the
! LOCAL gather/scatter are dropped on both ranks alike) and rank 1 owns the GPU
! context with every particle, reciprocal part included.  With several direct
! ranks every rank keeps the reference's communication: the direct ranks run
! the real-space ring (Force_direct_ext, main.f90:107-130) on their own
! contexts, each PME rank receives its group's positions through the LOCAL
! gather (force.f90:679), the PME ranks of several groups exchange them, and
! each computes the reciprocal part on its own context; the kinetic sums of
! BEREND/ISOKIN are reduced over COMM1D between the kick and the rescale
! (vverlet.f90:254, 278).
!
! Environment: TCP_PRECISION=MIXED selects the mixed-precision pair kernel
! (FP64 otherwise).

MODULE TCP_GPU_STATE
USE DATASTR
USE TCP_GPU_BIND
IMPLICIT NONE
INTEGER(C_INT64_T), SAVE :: stoch_seed = 20240607_C_INT64_T
! rank of this direct rank in COMM1D: separates the ranks' noise streams
INTEGER(C_INT64_T), SAVE :: noise_rank = 0_C_INT64_T
! the device's x, v, f are newer than qel/qion and the shadows (download deferred)
LOGICAL, SAVE :: x_on_device = .FALSE., v_on_device = .FALSE., f_on_device = .FALSE.
! between VVerlet*1 and VVerlet*2 of a step; RDF counts of the step not yet added
LOGICAL, SAVE :: in_step = .FALSE., rdf_pending = .FALSE.
! 0: forces may stay on the device between subroutines of one step (one direct
! rank); 1: every subroutine downloads what it writes; -1: not yet read
INTEGER, SAVE :: sync_every = -1
CONTAINS

! Results stay on the device between the subroutines of a step: one direct rank
! and TCP_DROPIN_SYNC not set to 1.
LOGICAL FUNCTION defer_forces(NS)
   TYPE(NM) :: NS
   CHARACTER(LEN=8) :: val
   INTEGER :: ln, st
   IF (sync_every < 0) THEN
      sync_every = 0
      CALL GET_ENVIRONMENT_VARIABLE('TCP_DROPIN_SYNC', val, ln, st)
      IF (st == 0 .AND. ln >= 1) THEN
         IF (val(1:1) == '1') sync_every = 1
      END IF
   END IF
   defer_forces = sync_every == 0 .AND. direct_ranks(NS) == 1
END FUNCTION defer_forces

! direct ranks of all PME groups
INTEGER FUNCTION direct_ranks(NS)
   TYPE(NM) :: NS
   direct_ranks = NS%NCPUS - NS%Npe
END FUNCTION direct_ranks

! The rank's context: ne electrons and ni ions, reciprocal part on it or not.
SUBROUTINE ensure_ctx(NS, param, ne, ni, with_recip)
   TYPE(NM) :: NS
   TYPE(PM) :: param
   INTEGER, INTENT(IN) :: ne, ni
   LOGICAL, INTENT(IN) :: with_recip
   TYPE(tcp_params_t) :: p
   INTEGER(C_INT) :: rc
   IF (C_ASSOCIATED(tcp_ctx_h)) RETURN
   p%ne = ne
   p%ni = ni
   p%K = param%K
   p%order = 4                       ! Force_SPME4
   p%kmax = 0
   p%recip = TCP_RECIP_SPME
   IF (.NOT. with_recip) p%recip = TCP_RECIP_NONE
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
   tcp_n = ne + ni
   tcp_ne = ne
   ALLOCATE(hx(3, tcp_n), hv(3, tcp_n), hf(3, tcp_n))
   ! a new context holds zeros: so does its shadow
   hx = 0.0D0
   hv = 0.0D0
   hf = 0.0D0
END SUBROUTINE ensure_ctx

! Field k (1 = xx, 2 = xv, 3 = ff) of qel/qion into the shadow sh; true when
! any value differed in any bit (the device copy is then out of date).
LOGICAL FUNCTION to_shadow(NS, qel, qion, k, sh)
   TYPE(NM) :: NS
   TYPE(PT) :: qel(NS%Nel), qion(NS%Nion)
   INTEGER, INTENT(IN) :: k
   REAL(C_DOUBLE), INTENT(INOUT) :: sh(:,:)
   INTEGER :: i
   REAL(KIND=DP) :: t(3)
   to_shadow = .FALSE.
   DO i = 1, NS%Nel + NS%Nion
      IF (i <= NS%Nel) THEN
         SELECT CASE (k)
         CASE (1); t = qel(i)%xx
         CASE (2); t = qel(i)%xv
         CASE DEFAULT; t = qel(i)%ff
         END SELECT
      ELSE
         SELECT CASE (k)
         CASE (1); t = qion(i - NS%Nel)%xx
         CASE (2); t = qion(i - NS%Nel)%xv
         CASE DEFAULT; t = qion(i - NS%Nel)%ff
         END SELECT
      END IF
      ! bit patterns, so that -0 against +0 and NaNs count as changes too
      IF (ANY(TRANSFER(t, 0_C_INT64_T, 3) /= TRANSFER(sh(:, i), 0_C_INT64_T, 3))) THEN
         sh(:, i) = t
         to_shadow = .TRUE.
      END IF
   END DO
END FUNCTION to_shadow

! Uploads the requested arrays whose host copy differs from the device's (an
! array whose download is deferred is newer on the device and is not uploaded).
SUBROUTINE put(NS, qel, qion, wx, wv, wf)
   TYPE(NM) :: NS
   TYPE(PT) :: qel(NS%Nel), qion(NS%Nion)
   LOGICAL, INTENT(IN) :: wx, wv, wf
   TYPE(C_PTR) :: px, pv, pf
   px = C_NULL_PTR; pv = C_NULL_PTR; pf = C_NULL_PTR
   IF (wx .AND. .NOT. x_on_device) THEN
      IF (to_shadow(NS, qel, qion, 1, hx)) px = C_LOC(hx)
   END IF
   IF (wv .AND. .NOT. v_on_device) THEN
      IF (to_shadow(NS, qel, qion, 2, hv)) pv = C_LOC(hv)
   END IF
   IF (wf .AND. .NOT. f_on_device) THEN
      IF (to_shadow(NS, qel, qion, 3, hf)) pf = C_LOC(hf)
   END IF
   IF (C_ASSOCIATED(px) .OR. C_ASSOCIATED(pv) .OR. C_ASSOCIATED(pf)) &
        CALL tcp_check(tcp_set_state(tcp_ctx_h, px, pv, pf), 'tcp_set_state')
END SUBROUTINE put

! After a kernel changed f on the device: keep it there (one direct rank) or
! bring it back now.
SUBROUTINE forces_written(NS, qel, qion)
   TYPE(NM) :: NS
   TYPE(PT) :: qel(NS%Nel), qion(NS%Nion)
   IF (defer_forces(NS)) THEN
      f_on_device = .TRUE.
   ELSE
      CALL get(NS, qel, qion, .FALSE., .FALSE., .TRUE.)
   END IF
END SUBROUTINE forces_written

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
   IF (wx) x_on_device = .FALSE.
   IF (wv) v_on_device = .FALSE.
   IF (wf) f_on_device = .FALSE.
END SUBROUTINE get

SUBROUTINE add_rdf(NS)
   TYPE(NM) :: NS
   hrdf = 0.0D0
   CALL tcp_check(tcp_get_histograms(tcp_ctx_h, C_LOC(hrdf), C_NULL_PTR, 1_C_INT), 'tcp_get_histograms')
   NS%RDF_ee = NS%RDF_ee + hrdf(:, 1)
   NS%RDF_ei = NS%RDF_ei + hrdf(:, 2)
   NS%RDF_ii = NS%RDF_ii + hrdf(:, 3)
END SUBROUTINE add_rdf

SUBROUTINE add_histograms(NS)
   TYPE(NM) :: NS
   hrdf = 0.0D0
   hvdf = 0.0D0
   CALL tcp_check(tcp_get_histograms(tcp_ctx_h, C_LOC(hrdf), C_LOC(hvdf), 1_C_INT), &
        'tcp_get_histograms')
   NS%RDF_ee = NS%RDF_ee + hrdf(:, 1)
   NS%RDF_ei = NS%RDF_ei + hrdf(:, 2)
   NS%RDF_ii = NS%RDF_ii + hrdf(:, 3)
   NS%VDF_e = NS%VDF_e + hvdf(:, 1)
   NS%VDF_i = NS%VDF_i + hvdf(:, 2)
END SUBROUTINE add_histograms

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
   LOGICAL :: bx, bf
   CALL put(NS, qel, qion, .FALSE., .TRUE., .TRUE.)
   stoch_seed = stoch_seed + 1
   CALL tcp_check(tcp_vverlet2(tcp_ctx_h, code, dt, param%Te, param%Ti, param%tau, param%alpha, &
        stoch_seed + noise_rank*1099511627776_C_INT64_T, a, b), 'tcp_vverlet2')
   sys%mv2el = a
   sys%mv2ion = b
   ! the end of the step: the host holds x, v, f and the histograms again
   bx = x_on_device
   bf = code == TCP_STOCH .OR. f_on_device
   CALL get(NS, qel, qion, bx, .TRUE., bf)
   IF (rdf_pending) THEN
      CALL add_histograms(NS)
   ELSE
      CALL add_vdf(NS)
   END IF
   rdf_pending = .FALSE.
   in_step = .FALSE.
END SUBROUTINE vv2_common

! VVerletBerend2 / VVerletIsokin2 with several direct ranks: kick, the
! COMM1D sums of m v^2 per species, lambda as vverlet.f90:255-256, 279-280
! (BEREND) or 329-330, 353-354 (ISOKIN), rescale.
SUBROUTINE vv2_split(NS, qel, qion, param, sys, dt, COMM, berend)
   INCLUDE 'mpif.h'
   TYPE(NM) :: NS
   TYPE(PT) :: qel(NS%Nel), qion(NS%Nion)
   TYPE(PM) :: param
   TYPE(ST) :: sys
   TYPE(MP) :: COMM
   REAL(KIND=DP) :: dt
   LOGICAL, INTENT(IN) :: berend
   REAL(C_DOUBLE) :: a, b
   REAL(KIND=DP) :: mv2, T0, lam_e, lam_i
   INTEGER :: IERR
   CALL put(NS, qel, qion, .FALSE., .TRUE., .TRUE.)
   CALL tcp_check(tcp_vverlet2_kick(tcp_ctx_h, dt, a, b), 'tcp_vverlet2_kick')
   CALL MPI_ALLREDUCE(a, mv2, 1, MPI_REAL8, MPI_SUM, COMM%COMM1D, IERR)
   T0 = mv2/DBLE(3*NS%NTel)
   IF (berend) THEN
      lam_e = DSQRT(1.D0+ dt*(param%Te/T0 - 1.D0)/param%tau)
   ELSE
      lam_e = DSQRT(param%Te/T0)
   END IF
   CALL MPI_ALLREDUCE(b, mv2, 1, MPI_REAL8, MPI_SUM, COMM%COMM1D, IERR)
   T0 = mv2/DBLE(3*NS%NTion)
   IF (berend) THEN
      lam_i = DSQRT(1.D0+ dt*(param%Ti/T0 - 1.D0)/param%tau)
   ELSE
      lam_i = DSQRT(param%Ti/T0)
   END IF
   CALL tcp_check(tcp_vverlet2_scale(tcp_ctx_h, lam_e, lam_i, a, b), 'tcp_vverlet2_scale')
   sys%mv2el = a
   sys%mv2ion = b
   CALL get(NS, qel, qion, .FALSE., .TRUE., .FALSE.)
   CALL add_vdf(NS)
END SUBROUTINE vv2_split

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
   ! Uo from the rank's own particles, as force.f90:45-46
   CALL ensure_ctx(NS, param, NS%Nel, NS%Nion, direct_ranks(NS) == 1)
   noise_rank = INT(COMM%MYID2, C_INT64_T)
   CALL tcp_check(tcp_ewald_self(tcp_ctx_h, a, rc2, uo, C_NULL_PTR), 'tcp_ewald_self')
   sys%Uo = uo
ELSE IF (direct_ranks(NS) > 1) THEN
   ! the PME rank's context holds the group's particles for the reciprocal part
   CALL ensure_ctx(NS, param, NS%NTel, NS%NTion, .TRUE.)
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
CALL forces_written(NS, qel, qion)
IF (in_step .AND. defer_forces(NS)) THEN
   rdf_pending = .TRUE.
ELSE
   CALL add_rdf(NS)
END IF
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

! One direct rank: positions reach the GPU in SPME_SUM, no LOCAL gather.
! Several: the reference's LOCAL gather (force.f90:679-680), then the PME rank
! computes the reciprocal forces of the gathered particles on its context and
! leaves them in pme for SPME_SUM's scatter.  pme holds each direct rank's
! electrons then its ions (force.f90:694-697); the context holds all electrons
! first.  With several PME groups (data.f90:151) the mesh needs every group's
! charges (the reference sums the groups' grids over the PME ranks,
! force.f90:821-822): the PME ranks exchange their groups' positions over
! their COMM1D, groups in order, and each computes the whole system's
! reciprocal part and keeps its own group's forces.
SUBROUTINE Force_SPME4(NS, param, sys, COMM, PSI, gsend, pme)
USE DATASTR
USE TCP_GPU_STATE
IMPLICIT NONE
INCLUDE 'mpif.h'
TYPE(NM):: NS
TYPE(PM):: param
TYPE(ST):: sys
TYPE(MP):: COMM
TYPE(GH):: gsend(NS%Ng), pme(NS%NLpt)
REAL(KIND=DP) :: PSI(param%K(1), param%K(2), param%K(3))
REAL(C_DOUBLE) :: um
INTEGER :: l, x, m, ke, ki, IERR, ge, gi, e0, i0, g, mine
INTEGER, ALLOCATABLE :: cnt(:,:), rcv(:), dsp(:)
REAL(C_DOUBLE), ALLOCATABLE :: gxe(:,:), gxi(:,:)
IF (direct_ranks(NS) == 1) RETURN
CALL MPI_GATHERV(gsend, NS%CNT(COMM%LID+1), COMM%PTCL, pme, NS%CNT, &
     NS%DPL, COMM%PTCL, 0, COMM%LOCAL, IERR)
IF (COMM%TAG) RETURN
! the group's direct ranks: LOCAL ranks 1 .. Nlsize-1 (main.f90:322-333)
ge = SUM(NS%NGHB(1:COMM%Nlsize-1, 1))
gi = SUM(NS%NGHB(1:COMM%Nlsize-1, 2))
e0 = 0
i0 = 0
IF (NS%Npe > 1) THEN
   ALLOCATE(cnt(2, NS%Npe), rcv(NS%Npe), dsp(NS%Npe), gxe(3, MAX(ge, 1)), gxi(3, MAX(gi, 1)))
   CALL MPI_ALLGATHER((/ge, gi/), 2, MPI_INTEGER, cnt, 2, MPI_INTEGER, COMM%COMM1D, IERR)
   mine = COMM%MYID2 + 1
   DO g = 1, mine - 1
      e0 = e0 + cnt(1, g)
      i0 = i0 + cnt(2, g)
   END DO
   m = 0
   ke = 0
   ki = 0
   DO l = 1, COMM%Nlsize - 1
      DO x = 1, NS%NGHB(l,1)
         m = m + 1
         ke = ke + 1
         gxe(:, ke) = pme(m)%xx
      END DO
      DO x = 1, NS%NGHB(l,2)
         m = m + 1
         ki = ki + 1
         gxi(:, ki) = pme(m)%xx
      END DO
   END DO
   dsp(1) = 0
   DO g = 1, NS%Npe
      rcv(g) = 3*cnt(1, g)
      IF (g > 1) dsp(g) = dsp(g-1) + rcv(g-1)
   END DO
   CALL MPI_ALLGATHERV(gxe, 3*ge, MPI_REAL8, hx, rcv, dsp, MPI_REAL8, COMM%COMM1D, IERR)
   DO g = 1, NS%Npe
      rcv(g) = 3*cnt(2, g)
      IF (g > 1) dsp(g) = dsp(g-1) + rcv(g-1)
   END DO
   CALL MPI_ALLGATHERV(gxi, 3*gi, MPI_REAL8, hx(1, NS%NTel + 1), rcv, dsp, MPI_REAL8, &
        COMM%COMM1D, IERR)
   DEALLOCATE(cnt, rcv, dsp, gxe, gxi)
ELSE
   m = 0
   ke = 0
   ki = NS%NTel
   DO l = 1, COMM%Nlsize - 1
      DO x = 1, NS%NGHB(l,1)
         m = m + 1
         ke = ke + 1
         hx(:, ke) = pme(m)%xx
      END DO
      DO x = 1, NS%NGHB(l,2)
         m = m + 1
         ki = ki + 1
         hx(:, ki) = pme(m)%xx
      END DO
   END DO
END IF
hf = 0.0D0
CALL tcp_check(tcp_set_state(tcp_ctx_h, C_LOC(hx), C_NULL_PTR, C_LOC(hf)), 'tcp_set_state')
CALL tcp_check(tcp_force_recip(tcp_ctx_h, um), 'tcp_force_recip')
sys%Um = um
CALL tcp_check(tcp_get_state(tcp_ctx_h, C_NULL_PTR, C_NULL_PTR, C_LOC(hf)), 'tcp_get_state')
m = 0
ke = e0
ki = NS%NTel + i0
DO l = 1, COMM%Nlsize - 1
   DO x = 1, NS%NGHB(l,1)
      m = m + 1
      ke = ke + 1
      pme(m)%xx = hf(:, ke)
   END DO
   DO x = 1, NS%NGHB(l,2)
      m = m + 1
      ki = ki + 1
      pme(m)%xx = hf(:, ki)
   END DO
END DO
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
INTEGER :: i, k, IERR
INCLUDE 'mpif.h'
IF (direct_ranks(NS) == 1) THEN
   IF (.NOT. COMM%TAG) RETURN
   CALL put(NS, qel, qion, .TRUE., .FALSE., .TRUE.)
   CALL tcp_check(tcp_force_recip(tcp_ctx_h, um), 'tcp_force_recip')
   sys%Um = um
   CALL forces_written(NS, qel, qion)
   RETURN
END IF
! several direct ranks: the reference's broadcast and scatter (force.f90:950-965)
CALL MPI_BCAST(sys%Um, 1, MPI_REAL8, 0, COMM%LOCAL, IERR)
CALL MPI_SCATTERV(pme, NS%CNT, NS%DPL, COMM%PTCL, grec, NS%CNT(COMM%LID+1), &
     COMM%PTCL, 0, COMM%LOCAL, IERR)
IF (COMM%TAG) THEN
   DO i = 1, NS%Nel
      DO k = 1, 3
         qel(i)%ff(k) = qel(i)%ff(k) + grec(i)%xx(k)
      END DO
   END DO
   DO i = 1, NS%Nion
      DO k = 1, 3
         qion(i)%ff(k) = qion(i)%ff(k) + grec(i+NS%Nel)%xx(k)
      END DO
   END DO
END IF
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
   CALL forces_written(NS, qel, qion)
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
IF (defer_forces(NS)) THEN
   ! one direct rank: x and v come back with VVerlet*2; gsend has no reader
   x_on_device = .TRUE.
   v_on_device = .TRUE.
   in_step = .TRUE.
   RETURN
END IF
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
IF (defer_forces(NS)) THEN
   ! one direct rank: x and v come back with VVerlet*2; gsend has no reader
   x_on_device = .TRUE.
   v_on_device = .TRUE.
   in_step = .TRUE.
   RETURN
END IF
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
IF (direct_ranks(NS) > 1) THEN
   CALL vv2_split(NS, qel, qion, param, sys, dt, COMM, .TRUE.)
ELSE
   CALL vv2_common(NS, qel, qion, param, sys, dt, TCP_BEREND)
END IF
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
IF (direct_ranks(NS) > 1) THEN
   CALL vv2_split(NS, qel, qion, param, sys, dt, COMM, .FALSE.)
ELSE
   CALL vv2_common(NS, qel, qion, param, sys, dt, TCP_ISOKIN)
END IF
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
