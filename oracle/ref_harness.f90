! Test-only harness around the UNMODIFIED reference routines.
!
! This program is oracle infrastructure: it is compiled by oracle/Makefile
! together with data.f90, force.f90 and vverlet.f90 exactly as they lie in the
! reference tree, and it calls those routines on one MPI process
! (MPI_COMM_SELF) to write full-precision REAL8 results that the reference's
! own output writers (ES14.6 / ES15.8, result.f90:131,324) cannot carry.
!
! Force_SPME4 (force.f90:652) is NOT callable here: it needs FFTE's ZFFT3D,
! which the reference does not vendor (Makefile:13-14).  The executable is
! linked with zfft3d_ left unresolved and never calls it; the reciprocal part
! is pinned by mathematics in the tests instead.
!
! Input (stream, little endian):
!   int32  mode, Ne, Ni, K(3), nsteps, thermo, nown_e, nown_i
!   real64 box(3), q(2), xm(2), rc, rs, dx_ee, dx_ei, dx_ii, Te, Ti,
!          dt, tau, alpha
!   real64 x(3,N), v(3,N), f(3,N)            (electrons first, then ions)
! Output (stream): see write_state / the per-mode blocks below.
!
! mode 1: EWALD_SELF (direct PE) + Force_direct_self     (force.f90:4,52)
! mode 2: EWALD_SELF on a PME PE -> PSI(K1,K2,K3)        (force.f90:27-42)
! mode 3: own block vs external block: Force_direct_self on the own block,
!         then Force_direct_ext and Force_direct_sum     (force.f90:183,346)
! mode 4: REMOVE_RIGIDMTN                                 (force.f90:974)
! mode 5: VVerletNotemp1 / VVerletStoch1                  (vverlet.f90:132,9)
! mode 6: VVerletNotemp2 / VVerletBerend2 / VVerletIsokin2 (vverlet.f90:174,227,301)
! mode 7: nsteps of the main.f90:81-170 step without the reciprocal part:
!         VVerlet*1, Force_direct_self, REMOVE_RIGIDMTN, VVerlet*2
! mode 8: timing of Force_direct_ext for nown own particles against all N
! mode 9: the output writers of result.f90 on given data: DUMP (RDF###.dat,
!         VDF###.dat), SAMPLING (a row of ener00.dat) and Restart (REST##.xyz,
!         then the header of the next ener##.dat), written into the current
!         directory.  Extra input after f: int32 Ndump, Nloop, Nrest, id(N);
!         real64 tnow, mv2el, mv2ion, Epot, Um, RDF(501,3), VDF(501,2).
PROGRAM REF_HARNESS
USE DATASTR
IMPLICIT NONE
INCLUDE 'mpif.h'
TYPE(NM) :: NS
TYPE(PM) :: param
TYPE(ST) :: sys
TYPE(MP) :: COMM
TYPE(TM) :: ttime
TYPE(PT), ALLOCATABLE :: qel(:), qion(:)
TYPE(GH), ALLOCATABLE :: gsend(:), gemit(:), grec(:)
REAL(KIND=DP), ALLOCATABLE :: PSI(:,:,:), xin(:,:), vin(:,:), fin(:,:)
REAL(KIND=DP), ALLOCATABLE :: trace(:,:), hrdf(:,:), hvdf(:,:)
INTEGER, ALLOCATABLE :: ids(:)
INTEGER :: w_ndump, w_nloop, w_nrest
REAL(KIND=DP) :: w_tnow, w_mv2el, w_mv2ion, w_epot, w_um
INTEGER :: mode, Ne, Ni, K(3), nsteps, thermo, nown_e, nown_i, N, i, ierr
INTEGER :: istep, next_e, next_i, nrep
REAL(KIND=DP) :: dt, Ur, Ur_self, t0, t1
CHARACTER(len=512) :: fin_name, fout_name

CALL MPI_INIT(ierr)
CALL GET_COMMAND_ARGUMENT(1, fin_name)
CALL GET_COMMAND_ARGUMENT(2, fout_name)

OPEN(UNIT=21, FILE=TRIM(fin_name), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
READ(21) mode, Ne, Ni, K(1), K(2), K(3), nsteps, thermo, nown_e, nown_i
READ(21) param%box(1), param%box(2), param%box(3), param%q(1), param%q(2), &
     & param%xm(1), param%xm(2), param%rc, param%rs, NS%dx_ee, NS%dx_ei, &
     & NS%dx_ii, param%Te, param%Ti, dt, param%tau, param%alpha
N = Ne + Ni
ALLOCATE(xin(3,N), vin(3,N), fin(3,N))
READ(21) xin
READ(21) vin
READ(21) fin
IF (mode == 9) THEN
   ALLOCATE(ids(N), hrdf(Ndist+1, 3), hvdf(Ndist+1, 2))
   READ(21) w_ndump, w_nloop, w_nrest
   READ(21) ids
   READ(21) w_tnow, w_mv2el, w_mv2ion, w_epot, w_um
   READ(21) hrdf
   READ(21) hvdf
END IF
CLOSE(21)

param%K = K
NS%NTel = Ne
NS%NTion = Ni
NS%NTpt = N
NS%dv_e = param%Te*4.D0/DBLE(Ndist)      ! main.f90:286
NS%dv_i = param%Ti*4.D0/DBLE(Ndist)      ! main.f90:287
NS%RDF_ee = 0.D0; NS%RDF_ei = 0.D0; NS%RDF_ii = 0.D0
NS%VDF_e = 0.D0; NS%VDF_i = 0.D0
sys%Ur = 0.D0; sys%Uo = 0.D0; sys%Um = 0.D0; sys%Epot = 0.D0
sys%mv2el = 0.D0; sys%mv2ion = 0.D0
COMM%COMM1D = MPI_COMM_SELF
COMM%LOCAL = MPI_COMM_SELF
COMM%TAG = .TRUE.

IF (mode == 3 .OR. mode == 8) THEN
   ! own block = first nown_e electrons and first nown_i ions
   NS%Nel = nown_e
   NS%Nion = nown_i
ELSE
   NS%Nel = Ne
   NS%Nion = Ni
END IF
NS%Npt = NS%Nel + NS%Nion
NS%Ng = N
ALLOCATE(qel(MAX(NS%Nel,1)), qion(MAX(NS%Nion,1)))
ALLOCATE(gsend(N), gemit(N), grec(N))
DO i = 1, NS%Nel
   qel(i)%id = i
   qel(i)%xx = xin(:,i); qel(i)%xv = vin(:,i); qel(i)%ff = fin(:,i)
END DO
DO i = 1, NS%Nion
   qion(i)%id = i
   qion(i)%xx = xin(:,Ne+i); qion(i)%xv = vin(:,Ne+i); qion(i)%ff = fin(:,Ne+i)
END DO

! alpha and rc2 are set by EWALD_SELF on every PE (force.f90:23-24)
ALLOCATE(PSI(K(1),K(2),K(3)))
PSI = 0.D0
COMM%TAG = .TRUE.
CALL EWALD_SELF(NS, param, sys, PSI, COMM)

OPEN(UNIT=22, FILE=TRIM(fout_name), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
SELECT CASE (mode)
CASE (1)
   CALL Force_direct_self(NS, qel, qion, param, sys)
   sys%Epot = (sys%Ur + sys%Uo)*eps
   CALL write_state()
CASE (2)
   COMM%TAG = .FALSE.
   CALL EWALD_SELF(NS, param, sys, PSI, COMM)
   CALL write_state()
   WRITE(22) PSI
CASE (3)
   next_e = Ne - nown_e
   next_i = Ni - nown_i
   DO i = 1, next_e
      grec(i)%xx = xin(:, nown_e + i)
   END DO
   DO i = 1, next_i
      grec(next_e + i)%xx = xin(:, Ne + nown_i + i)
   END DO
   CALL Force_direct_self(NS, qel, qion, param, sys)
   Ur_self = sys%Ur
   CALL Force_direct_ext(NS, qel, qion, param, grec, gemit, Ur, next_e, next_i)
   sys%Ur = Ur
   CALL write_state()
   WRITE(22) Ur_self, Ur
   WRITE(22) (gemit(i)%xx, i = 1, next_e + next_i)
   ! Force_direct_sum adds a returned reaction block onto the own particles;
   ! feed it this block's own reaction array as the received one.
   DO i = 1, NS%Npt
      grec(i)%xx = DBLE(i)*1.0D-3
   END DO
   CALL Force_direct_sum(NS, qel, qion, grec)
   WRITE(22) (qel(i)%ff, i = 1, NS%Nel), (qion(i)%ff, i = 1, NS%Nion)
CASE (4)
   CALL REMOVE_RIGIDMTN(NS, qel, qion, COMM)
   CALL write_state()
CASE (5)
   IF (thermo == 3) THEN
      CALL VVerletStoch1(NS, qel, qion, param, dt, gsend)
   ELSE
      CALL VVerletNotemp1(NS, qel, qion, param, dt, gsend)
   END IF
   CALL write_state()
   WRITE(22) (gsend(i)%xx, i = 1, N)
CASE (6)
   CALL vv2()
   CALL write_state()
CASE (7)
   ALLOCATE(trace(4, nsteps))
   DO istep = 1, nsteps
      IF (thermo == 3) THEN
         CALL VVerletStoch1(NS, qel, qion, param, dt, gsend)
      ELSE
         CALL VVerletNotemp1(NS, qel, qion, param, dt, gsend)
      END IF
      CALL Force_direct_self(NS, qel, qion, param, sys)
      sys%Epot = (sys%Ur + sys%Uo)*eps
      CALL REMOVE_RIGIDMTN(NS, qel, qion, COMM)
      CALL vv2()
      trace(1, istep) = sys%Ur
      trace(2, istep) = sys%Epot
      trace(3, istep) = sys%mv2el
      trace(4, istep) = sys%mv2ion
   END DO
   CALL write_state()
   WRITE(22) trace
CASE (8)
   next_e = Ne
   next_i = Ni
   DO i = 1, N
      grec(i)%xx = xin(:, i)
   END DO
   nrep = MAX(nsteps, 1)
   CALL Force_direct_ext(NS, qel, qion, param, grec, gemit, Ur, next_e, next_i)
   t0 = MPI_Wtime()
   DO istep = 1, nrep
      CALL Force_direct_ext(NS, qel, qion, param, grec, gemit, Ur, next_e, next_i)
   END DO
   t1 = MPI_Wtime()
   WRITE(22) DBLE(NS%Npt)*DBLE(N)*DBLE(nrep), t1 - t0, Ur
CASE (9)
   COMM%NUMPROC = 1
   COMM%MYID2 = 0
   ALLOCATE(NS%NOPT(1, 2))
   NS%NOPT(1, 1) = Ne
   NS%NOPT(1, 2) = Ni
   DO i = 1, Ne
      qel(i)%id = ids(i)
   END DO
   DO i = 1, Ni
      qion(i)%id = ids(Ne + i)
   END DO
   ttime%Ndump = w_ndump
   ttime%Nloop = w_nloop
   ttime%Nrest = w_nrest
   ttime%tnow = w_tnow
   NS%RDF_ee = hrdf(:, 1)
   NS%RDF_ei = hrdf(:, 2)
   NS%RDF_ii = hrdf(:, 3)
   NS%VDF_e = hvdf(:, 1)
   NS%VDF_i = hvdf(:, 2)
   sys%mv2el = w_mv2el
   sys%mv2ion = w_mv2ion
   sys%Epot = w_epot
   sys%Um = w_um
   OPEN(UNIT=55, FILE="ener00.dat")
   CALL DUMP(NS, ttime, param, COMM)
   CALL SAMPLING(NS, ttime, sys, COMM, param)
   CALL Restart(NS, qel, qion, ttime, param, COMM)
   CLOSE(55)
CASE DEFAULT
   PRINT *, "unknown mode", mode
   STOP 2
END SELECT
CLOSE(22)
CALL MPI_FINALIZE(ierr)

CONTAINS

SUBROUTINE vv2()
   SELECT CASE (thermo)
   CASE (1)
      CALL VVerletBerend2(NS, qel, qion, param, sys, dt, COMM)
   CASE (2)
      CALL VVerletIsokin2(NS, qel, qion, param, sys, dt, COMM)
   CASE DEFAULT
      CALL VVerletNotemp2(NS, qel, qion, param, sys, dt)
   END SELECT
END SUBROUTINE vv2

SUBROUTINE write_state()
   INTEGER :: j
   WRITE(22) param%a, param%rc2, sys%Uo, sys%Ur, sys%Epot, sys%mv2el, sys%mv2ion
   WRITE(22) (qel(j)%xx, j = 1, NS%Nel), (qion(j)%xx, j = 1, NS%Nion)
   WRITE(22) (qel(j)%xv, j = 1, NS%Nel), (qion(j)%xv, j = 1, NS%Nion)
   WRITE(22) (qel(j)%ff, j = 1, NS%Nel), (qion(j)%ff, j = 1, NS%Nion)
   WRITE(22) NS%RDF_ee, NS%RDF_ei, NS%RDF_ii, NS%VDF_e, NS%VDF_i
END SUBROUTINE write_state

END PROGRAM REF_HARNESS
