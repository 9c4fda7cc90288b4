! ISO_C_BINDING interfaces of libtcp_gpu (include/tcp_gpu.h) and the host
! staging used by the drop-in subroutines in tcp_force_gpu.f90.
MODULE TCP_GPU_BIND
USE ISO_C_BINDING
IMPLICIT NONE

INTEGER(C_INT), PARAMETER :: TCP_FP64 = 0, TCP_MIXED = 1
INTEGER(C_INT), PARAMETER :: TCP_NOTEMP = 0, TCP_BEREND = 1, TCP_ISOKIN = 2, TCP_STOCH = 3
INTEGER(C_INT), PARAMETER :: TCP_RECIP_NONE = 0, TCP_RECIP_SPME = 1, TCP_RECIP_EWALD = 2
INTEGER(C_INT), PARAMETER :: TCP_ERR_CUTOFF = -2
INTEGER, PARAMETER :: TCP_NBIN = 501

TYPE, BIND(C) :: tcp_params_t
   INTEGER(C_INT32_T) :: ne, ni
   INTEGER(C_INT32_T) :: K(3)
   INTEGER(C_INT32_T) :: order, kmax, recip, flags
   REAL(C_DOUBLE) :: box(3), q(2), xm(2)
   REAL(C_DOUBLE) :: rc, rs, alpha, eps
   REAL(C_DOUBLE) :: dx_ee, dx_ei, dx_ii, dv_e, dv_i
END TYPE tcp_params_t

INTERFACE
   INTEGER(C_INT) FUNCTION tcp_create(ctx, p, precision) BIND(C, NAME='tcp_create')
      IMPORT :: C_INT, C_PTR, tcp_params_t
      TYPE(C_PTR) :: ctx
      TYPE(tcp_params_t), INTENT(IN) :: p
      INTEGER(C_INT), VALUE :: precision
   END FUNCTION
   INTEGER(C_INT) FUNCTION tcp_destroy(ctx) BIND(C, NAME='tcp_destroy')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
   END FUNCTION
   TYPE(C_PTR) FUNCTION tcp_last_error(ctx) BIND(C, NAME='tcp_last_error')
      IMPORT :: C_PTR
      TYPE(C_PTR), VALUE :: ctx
   END FUNCTION
   INTEGER(C_INT) FUNCTION tcp_ewald_self(ctx, alpha, rc2, uo, psi) BIND(C, NAME='tcp_ewald_self')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx, psi
      REAL(C_DOUBLE) :: alpha, rc2, uo
   END FUNCTION
   INTEGER(C_INT) FUNCTION tcp_set_state(ctx, x, v, f) BIND(C, NAME='tcp_set_state')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx, x, v, f
   END FUNCTION
   INTEGER(C_INT) FUNCTION tcp_get_state(ctx, x, v, f) BIND(C, NAME='tcp_get_state')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx, x, v, f
   END FUNCTION
   INTEGER(C_INT) FUNCTION tcp_force_direct(ctx, ur) BIND(C, NAME='tcp_force_direct')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE) :: ur
   END FUNCTION
   INTEGER(C_INT) FUNCTION tcp_force_direct_ext(ctx, xg, ne_g, ni_g, gemit, ur) &
        BIND(C, NAME='tcp_force_direct_ext')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx, xg, gemit
      INTEGER(C_INT), VALUE :: ne_g, ni_g
      REAL(C_DOUBLE) :: ur
   END FUNCTION
   INTEGER(C_INT) FUNCTION tcp_force_recip(ctx, um) BIND(C, NAME='tcp_force_recip')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE) :: um
   END FUNCTION
   INTEGER(C_INT) FUNCTION tcp_remove_rigidmtn(ctx) BIND(C, NAME='tcp_remove_rigidmtn')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
   END FUNCTION
   INTEGER(C_INT) FUNCTION tcp_vverlet1(ctx, thermo, dt, fric) BIND(C, NAME='tcp_vverlet1')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT), VALUE :: thermo
      REAL(C_DOUBLE), VALUE :: dt, fric
   END FUNCTION
   INTEGER(C_INT) FUNCTION tcp_vverlet2(ctx, thermo, dt, te, ti, tau, fric, seed, mv2el, mv2ion) &
        BIND(C, NAME='tcp_vverlet2')
      IMPORT :: C_INT, C_PTR, C_DOUBLE, C_INT64_T
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT), VALUE :: thermo
      REAL(C_DOUBLE), VALUE :: dt, te, ti, tau, fric
      INTEGER(C_INT64_T), VALUE :: seed
      REAL(C_DOUBLE) :: mv2el, mv2ion
   END FUNCTION
   INTEGER(C_INT) FUNCTION tcp_vverlet2_kick(ctx, dt, mv2el, mv2ion) BIND(C, NAME='tcp_vverlet2_kick')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), VALUE :: dt
      REAL(C_DOUBLE) :: mv2el, mv2ion
   END FUNCTION
   INTEGER(C_INT) FUNCTION tcp_vverlet2_scale(ctx, lam_e, lam_i, mv2el, mv2ion) &
        BIND(C, NAME='tcp_vverlet2_scale')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), VALUE :: lam_e, lam_i
      REAL(C_DOUBLE) :: mv2el, mv2ion
   END FUNCTION
   INTEGER(C_INT) FUNCTION tcp_get_histograms(ctx, rdf, vdf, reset) BIND(C, NAME='tcp_get_histograms')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx, rdf, vdf
      INTEGER(C_INT), VALUE :: reset
   END FUNCTION
END INTERFACE

TYPE(C_PTR), SAVE :: tcp_ctx_h = C_NULL_PTR
INTEGER, SAVE :: tcp_n = 0, tcp_ne = 0
REAL(C_DOUBLE), ALLOCATABLE, TARGET, SAVE :: hx(:,:), hv(:,:), hf(:,:)
REAL(C_DOUBLE), TARGET, SAVE :: hrdf(TCP_NBIN, 3), hvdf(TCP_NBIN, 2)

CONTAINS

SUBROUTINE tcp_check(rc, what)
   INTEGER(C_INT), INTENT(IN) :: rc
   CHARACTER(*), INTENT(IN) :: what
   CHARACTER(KIND=C_CHAR), POINTER :: msg(:)
   CHARACTER(LEN=256) :: text
   INTEGER :: i
   IF (rc == 0) RETURN
   CALL C_F_POINTER(tcp_last_error(tcp_ctx_h), msg, [256])
   text = ' '
   DO i = 1, 256
      IF (msg(i) == C_NULL_CHAR) EXIT
      text(i:i) = msg(i)
   END DO
   PRINT *, 'tcp_gpu: ', what, ' failed: ', TRIM(text)
   STOP " *** GPU force path error *** "
END SUBROUTINE tcp_check

INTEGER(C_INT) FUNCTION tcp_precision()
   CHARACTER(LEN=16) :: s
   INTEGER :: l, st
   CALL GET_ENVIRONMENT_VARIABLE('TCP_PRECISION', s, l, st)
   tcp_precision = TCP_FP64
   IF (st == 0 .AND. (s(1:5) == 'MIXED' .OR. s(1:5) == 'mixed')) tcp_precision = TCP_MIXED
END FUNCTION tcp_precision

INTEGER(C_INT) FUNCTION thermo_code(name)
   CHARACTER(LEN=*), INTENT(IN) :: name
   SELECT CASE (name(1:6))
   CASE ('BEREND')
      thermo_code = TCP_BEREND
   CASE ('ISOKIN')
      thermo_code = TCP_ISOKIN
   CASE ('STOCH ')
      thermo_code = TCP_STOCH
   CASE DEFAULT
      thermo_code = TCP_NOTEMP
   END SELECT
END FUNCTION thermo_code

END MODULE TCP_GPU_BIND
