! ref_harness.f90 -- C-callable entry points into the parts of the reference
! CRTM that build without netCDF (see oracle/Makefile).  Test infrastructure
! only: the tests call these through ctypes to check the build's own readers,
! file writers and numerical kernels against the reference routines.
!
! Wrapped reference routines:
!   Read_ODPS_Binary         src/Coefficients/TauCoeff/ODPS/ODPS_Binary_IO.f90:481
!   CRTM_GeometryInfo_Compute src/GeometryInfo/CRTM_GeometryInfo.f90:100
!   matinv, ASYMTX, Double_Gauss_Quadrature, Legendre
!                            src/CRTM_Utility/CRTM_Utility.f90:343-1094
MODULE ref_harness
  USE ISO_C_BINDING
  USE Type_Kinds, ONLY: fp
  USE ODPS_Define, ONLY: ODPS_type, Destroy_ODPS
  USE ODPS_Binary_IO, ONLY: Read_ODPS_Binary
  USE CRTM_GeometryInfo_Define, ONLY: CRTM_GeometryInfo_type, CRTM_GeometryInfo_SetValue, &
                                      CRTM_GeometryInfo_GetValue
  USE CRTM_Geometry_Define, ONLY: CRTM_Geometry_type, CRTM_Geometry_SetValue
  USE CRTM_GeometryInfo, ONLY: CRTM_GeometryInfo_Compute
  USE CRTM_Utility, ONLY: matinv, ASYMTX, Double_Gauss_Quadrature, Legendre
  IMPLICIT NONE
  TYPE(ODPS_type), SAVE :: tc
CONTAINS

  FUNCTION c2f(cstr, n) RESULT(s)
    CHARACTER(KIND=C_CHAR), INTENT(IN) :: cstr(*)
    INTEGER(C_INT), VALUE :: n
    CHARACTER(LEN=n) :: s
    INTEGER :: i
    DO i = 1, n
      s(i:i) = cstr(i)
    END DO
  END FUNCTION c2f

  ! Load an ODPS file with the reference reader; dims(1:9) =
  ! n_Layers, n_Components, n_Absorbers, n_Channels, n_Coeffs, n_OCoeffs,
  ! Group_Index, Sensor_Type, OComponent_Index
  INTEGER(C_INT) FUNCTION ref_odps_load(path, n, dims) BIND(C, NAME='ref_odps_load')
    CHARACTER(KIND=C_CHAR), INTENT(IN) :: path(*)
    INTEGER(C_INT), VALUE :: n
    INTEGER(C_INT), INTENT(OUT) :: dims(9)
    INTEGER :: err
    err = Destroy_ODPS(tc)
    err = Read_ODPS_Binary(c2f(path, n), tc, Quiet=1)
    ref_odps_load = err
    IF (err /= 0) RETURN
    dims = (/ tc%n_Layers, tc%n_Components, tc%n_Absorbers, tc%n_Channels, tc%n_Coeffs, &
              tc%n_OCoeffs, tc%Group_Index, tc%Sensor_Type, tc%OComponent_Index /)
  END FUNCTION ref_odps_load

  ! Copy the loaded arrays out (column-major, as stored)
  SUBROUTINE ref_odps_fetch(ref_lp, ref_p, ref_t, ref_a, min_a, max_a, npred, pos, c, &
                            osig, order, opidx, opos, oc, alpha3) BIND(C, NAME='ref_odps_fetch')
    REAL(C_DOUBLE), INTENT(OUT) :: ref_lp(*), ref_p(*), ref_t(*), ref_a(*), min_a(*), max_a(*)
    INTEGER(C_INT), INTENT(OUT) :: npred(*), pos(*)
    REAL(C_FLOAT),  INTENT(OUT) :: c(*)
    INTEGER(C_INT), INTENT(OUT) :: osig(*), order(*), opidx(*), opos(*)
    REAL(C_DOUBLE), INTENT(OUT) :: oc(*), alpha3(3)
    INTEGER :: Ko, J, Jm, L
    Ko = tc%n_Layers; J = tc%n_Components; Jm = tc%n_Absorbers; L = tc%n_Channels
    ref_lp(1:Ko+1) = tc%Ref_Level_Pressure
    ref_p(1:Ko) = tc%Ref_Pressure
    ref_t(1:Ko) = tc%Ref_Temperature
    ref_a(1:Ko*Jm) = RESHAPE(tc%Ref_Absorber, (/Ko*Jm/))
    min_a(1:Ko*Jm) = RESHAPE(tc%Min_Absorber, (/Ko*Jm/))
    max_a(1:Ko*Jm) = RESHAPE(tc%Max_Absorber, (/Ko*Jm/))
    npred(1:J*L) = RESHAPE(tc%n_Predictors, (/J*L/))
    pos(1:J*L) = RESHAPE(tc%Pos_Index, (/J*L/))
    IF (tc%n_Coeffs > 0) c(1:tc%n_Coeffs) = tc%C
    IF (tc%n_OCoeffs > 0) THEN
      osig(1:L) = tc%OSignificance
      order(1:L) = tc%Order
      opidx(1:7*L) = RESHAPE(tc%OP_Index, (/7*L/))
      opos(1:L) = tc%OPos_Index
      oc(1:tc%n_OCoeffs) = tc%OC
      alpha3 = (/ tc%Alpha, tc%Alpha_C1, tc%Alpha_C2 /)
    END IF
  END SUBROUTINE ref_odps_fetch

  ! Geometry: returns secant of sensor zenith, trans zenith radian, secant trans zenith
  SUBROUTINE ref_geometry(zenith, scan, out3) BIND(C, NAME='ref_geometry')
    REAL(C_DOUBLE), VALUE :: zenith, scan
    REAL(C_DOUBLE), INTENT(OUT) :: out3(3)
    TYPE(CRTM_Geometry_type) :: geo
    TYPE(CRTM_GeometryInfo_type) :: gi
    REAL(fp) :: a, b, c
    CALL CRTM_Geometry_SetValue(geo, Sensor_Zenith_Angle=zenith, Sensor_Scan_Angle=scan)
    CALL CRTM_GeometryInfo_SetValue(gi, Geometry=geo)
    CALL CRTM_GeometryInfo_Compute(gi)
    CALL CRTM_GeometryInfo_GetValue(gi, Secant_Sensor_Zenith=a, Trans_Zenith_Radian=b, &
                                    Secant_Trans_Zenith=c)
    out3 = (/ a, b, c /)
  END SUBROUTINE ref_geometry

  INTEGER(C_INT) FUNCTION ref_matinv(n, a, ainv) BIND(C, NAME='ref_matinv')
    INTEGER(C_INT), VALUE :: n
    REAL(C_DOUBLE), INTENT(IN) :: a(n, n)
    REAL(C_DOUBLE), INTENT(OUT) :: ainv(n, n)
    INTEGER :: err
    ainv = matinv(a, err)
    ref_matinv = err
  END FUNCTION ref_matinv

  SUBROUTINE ref_gauss(n, mu, w) BIND(C, NAME='ref_gauss')
    INTEGER(C_INT), VALUE :: n
    REAL(C_DOUBLE), INTENT(OUT) :: mu(n), w(n)
    CALL Double_Gauss_Quadrature(n, mu, w)
  END SUBROUTINE ref_gauss

  SUBROUTINE ref_legendre(nl, x, p) BIND(C, NAME='ref_legendre')
    INTEGER(C_INT), VALUE :: nl
    REAL(C_DOUBLE), VALUE :: x
    REAL(C_DOUBLE), INTENT(OUT) :: p(0:nl)
    CALL Legendre(nl, x, p)
  END SUBROUTINE ref_legendre

  INTEGER(C_INT) FUNCTION ref_asymtx(n, a, evec, eval) BIND(C, NAME='ref_asymtx')
    INTEGER(C_INT), VALUE :: n
    REAL(C_DOUBLE), INTENT(IN) :: a(n, n)
    REAL(C_DOUBLE), INTENT(OUT) :: evec(n, n), eval(n)
    REAL(fp) :: aa(n, n)
    INTEGER :: err
    aa = a
    CALL ASYMTX(aa, n, n, n, evec, eval, err)
    ref_asymtx = err
  END FUNCTION ref_asymtx

END MODULE ref_harness
