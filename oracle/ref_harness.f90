! ref_harness.f90 -- C-callable entry points into the parts of the reference
! CRTM that build without netCDF (see oracle/Makefile).  Test infrastructure
! only: the tests call these through ctypes to check the build's own readers,
! file writers and numerical kernels against the reference routines.
!
! Wrapped reference routines:
!   Read_ODPS_Binary         src/Coefficients/TauCoeff/ODPS/ODPS_Binary_IO.f90:481
!   CRTM_GeometryInfo_Compute src/GeometryInfo/CRTM_GeometryInfo.f90:100
!   matinv, ASYMTX, Double_Gauss_Quadrature, Legendre, Legendre_M
!                            src/CRTM_Utility/CRTM_Utility.f90:343-1094
!   CloudCoeff_Binary_ReadFile src/Coefficients/CloudCoeff/CloudCoeff_Binary_IO.f90:340
!   Find_Index, LPoly, Interp_1D/2D/3D
!                            src/Interpolation/CRTM_Interpolation.f90:225-860
!   CRTM_AtmOptics_Combine   src/AtmOptics/CRTM_AtmOptics.f90:369
!   AerosolCoeff_Binary_ReadFile, AerosolCoeff_Frequency, AerosolCoeff_typeID_to_index
!                            src/Coefficients/AerosolCoeff/AerosolCoeff_Binary_IO.f90:352,
!                            src/Coefficients/AerosolCoeff/AerosolCoeff_Define.f90:635-696
!   SpcCoeff_Binary_ReadFile src/Coefficients/SpcCoeff/SpcCoeff_Binary_IO.f90:325
!   MWwaterCoeff_ReadFile, Compute_FastemX(_AD)
!                            src/Coefficients/EmisCoeff/MW_Water/MWwaterCoeff_Define.f90:699,
!                            src/SfcOptics/MW_Water/FASTEM_MWSSEM/CRTM_FastemX.f90:302,803
!   IRwaterCoeff_ReadFile, CRTM_Compute_IRSSEM(_AD)
!                            src/Coefficients/EmisCoeff/IR_Water/IRwaterCoeff_Define.f90:805,
!                            src/SfcOptics/IR_Water/IRSSEM/CRTM_IRSSEM.f90:204,586
!   CRTM_Compute_Transmittance(_AD) src/AtmOptics/CRTM_AtmOptics.f90:210-368
MODULE ref_harness
  USE ISO_C_BINDING
  USE Type_Kinds, ONLY: fp
  USE ODPS_Define, ONLY: ODPS_type, Destroy_ODPS
  USE ODPS_Binary_IO, ONLY: Read_ODPS_Binary
  USE CRTM_GeometryInfo_Define, ONLY: CRTM_GeometryInfo_type, CRTM_GeometryInfo_SetValue, &
                                      CRTM_GeometryInfo_GetValue
  USE CRTM_Geometry_Define, ONLY: CRTM_Geometry_type, CRTM_Geometry_SetValue
  USE CRTM_GeometryInfo, ONLY: CRTM_GeometryInfo_Compute
  USE CRTM_Utility, ONLY: matinv, ASYMTX, Double_Gauss_Quadrature, Legendre, Legendre_M
  USE CloudCoeff_Define, ONLY: CloudCoeff_type, CloudCoeff_Destroy
  USE CloudCoeff_Binary_IO, ONLY: CloudCoeff_Binary_ReadFile
  USE CRTM_Interpolation, ONLY: Find_Index, LPoly, LPoly_type, Interp_1D, Interp_2D, Interp_3D
  USE CRTM_AtmOptics_Define, ONLY: CRTM_AtmOptics_type, CRTM_AtmOptics_Create
  USE AOvar_Define, ONLY: AOvar_type, AOvar_Create
  USE CRTM_AtmOptics, ONLY: CRTM_AtmOptics_Combine
  USE AerosolCoeff_Define, ONLY: AerosolCoeff_type, AerosolCoeff_Destroy, AerosolCoeff_Frequency, &
                                 AerosolCoeff_typeID_to_index
  USE AerosolCoeff_Binary_IO, ONLY: AerosolCoeff_Binary_ReadFile
  USE SpcCoeff_Define, ONLY: SpcCoeff_type, SpcCoeff_Destroy
  USE SpcCoeff_Binary_IO, ONLY: SpcCoeff_Binary_ReadFile
  USE ACCoeff_Define, ONLY: ACCoeff_Associated
  USE NLTECoeff_Define, ONLY: NLTECoeff_Associated
  USE CRTM_AtmOptics, ONLY: CRTM_Compute_Transmittance, CRTM_Compute_Transmittance_AD
  USE MWwaterCoeff_Define, ONLY: MWwaterCoeff_type, MWwaterCoeff_Destroy, MWwaterCoeff_ReadFile
  USE CRTM_FastemX, ONLY: FastemX_iVar_type => iVar_type, Compute_FastemX, Compute_FastemX_AD
  USE IRwaterCoeff_Define, ONLY: IRwaterCoeff_type, IRwaterCoeff_Destroy, IRwaterCoeff_ReadFile
  USE CRTM_IRSSEM, ONLY: IRSSEM_iVar_type => iVar_type, CRTM_Compute_IRSSEM, CRTM_Compute_IRSSEM_AD
  IMPLICIT NONE
  TYPE(ODPS_type), SAVE :: tc
  TYPE(CloudCoeff_type), SAVE :: ccf
  TYPE(AerosolCoeff_type), SAVE :: acf
  TYPE(SpcCoeff_type), SAVE :: scf
  TYPE(MWwaterCoeff_type), SAVE :: mwc
  TYPE(FastemX_iVar_type), SAVE :: fx_var
  TYPE(IRwaterCoeff_type), SAVE :: irc
  TYPE(IRSSEM_iVar_type), SAVE :: ir_var
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

  SUBROUTINE ref_legendre_m(nl, x, p) BIND(C, NAME='ref_legendre_m')
    INTEGER(C_INT), VALUE :: nl
    REAL(C_DOUBLE), VALUE :: x
    REAL(C_DOUBLE), INTENT(OUT) :: p(0:nl)
    CALL Legendre_M(0, nl, x, p)
  END SUBROUTINE ref_legendre_m

  ! Load a CloudCoeff file with the reference reader; dims(1:9) =
  ! n_MW_Frequencies, n_MW_Radii, n_IR_Frequencies, n_IR_Radii, n_Temperatures,
  ! n_MW_Densities, n_IR_Densities, n_Legendre_Terms, n_Phase_Elements
  INTEGER(C_INT) FUNCTION ref_cloudcoeff_load(path, n, dims) BIND(C, NAME='ref_cloudcoeff_load')
    CHARACTER(KIND=C_CHAR), INTENT(IN) :: path(*)
    INTEGER(C_INT), VALUE :: n
    INTEGER(C_INT), INTENT(OUT) :: dims(9)
    CALL CloudCoeff_Destroy(ccf)
    ref_cloudcoeff_load = CloudCoeff_Binary_ReadFile(c2f(path, n), ccf, Quiet=.TRUE.)
    IF (ref_cloudcoeff_load /= 0) RETURN
    dims = (/ ccf%n_MW_Frequencies, ccf%n_MW_Radii, ccf%n_IR_Frequencies, ccf%n_IR_Radii, &
              ccf%n_Temperatures, ccf%n_MW_Densities, ccf%n_IR_Densities, ccf%n_Legendre_Terms, &
              ccf%n_Phase_Elements /)
  END FUNCTION ref_cloudcoeff_load

  ! Copy the loaded CloudCoeff arrays out (column-major, as stored)
  SUBROUTINE ref_cloudcoeff_fetch(fmw, fir, rmw, rir, t, dmw, dir, kel, wl, pl, kes, ws, ps, kei, wi, pi) &
      BIND(C, NAME='ref_cloudcoeff_fetch')
    REAL(C_DOUBLE), INTENT(OUT) :: fmw(*), fir(*), rmw(*), rir(*), t(*), dmw(*), dir(*)
    REAL(C_DOUBLE), INTENT(OUT) :: kel(*), wl(*), pl(*), kes(*), ws(*), ps(*), kei(*), wi(*), pi(*)
    fmw(1:SIZE(ccf%Frequency_MW)) = ccf%Frequency_MW
    fir(1:SIZE(ccf%Frequency_IR)) = ccf%Frequency_IR
    rmw(1:SIZE(ccf%Reff_MW)) = ccf%Reff_MW
    rir(1:SIZE(ccf%Reff_IR)) = ccf%Reff_IR
    t(1:SIZE(ccf%Temperature)) = ccf%Temperature
    dmw(1:SIZE(ccf%Density_MW)) = ccf%Density_MW
    dir(1:SIZE(ccf%Density_IR)) = ccf%Density_IR
    kel(1:SIZE(ccf%ke_L_MW)) = RESHAPE(ccf%ke_L_MW, (/SIZE(ccf%ke_L_MW)/))
    wl(1:SIZE(ccf%w_L_MW)) = RESHAPE(ccf%w_L_MW, (/SIZE(ccf%w_L_MW)/))
    pl(1:SIZE(ccf%pcoeff_L_MW)) = RESHAPE(ccf%pcoeff_L_MW, (/SIZE(ccf%pcoeff_L_MW)/))
    kes(1:SIZE(ccf%ke_S_MW)) = RESHAPE(ccf%ke_S_MW, (/SIZE(ccf%ke_S_MW)/))
    ws(1:SIZE(ccf%w_S_MW)) = RESHAPE(ccf%w_S_MW, (/SIZE(ccf%w_S_MW)/))
    ps(1:SIZE(ccf%pcoeff_S_MW)) = RESHAPE(ccf%pcoeff_S_MW, (/SIZE(ccf%pcoeff_S_MW)/))
    kei(1:SIZE(ccf%ke_IR)) = RESHAPE(ccf%ke_IR, (/SIZE(ccf%ke_IR)/))
    wi(1:SIZE(ccf%w_IR)) = RESHAPE(ccf%w_IR, (/SIZE(ccf%w_IR)/))
    pi(1:SIZE(ccf%pcoeff_IR)) = RESHAPE(ccf%pcoeff_IR, (/SIZE(ccf%pcoeff_IR)/))
  END SUBROUTINE ref_cloudcoeff_fetch

  ! Table interpolation as CRTM_CloudScatter does it: clamp, Find_Index,
  ! LPoly per axis, Interp_nD over the 4^nd sub-table.  nd = 1..3; x packs
  ! the axes (n(1), n(2), n(3) points), z the full column-major table.
  SUBROUTINE ref_interp(nd, n, x, xi, z, out) BIND(C, NAME='ref_interp')
    INTEGER(C_INT), VALUE :: nd
    INTEGER(C_INT), INTENT(IN) :: n(3)
    REAL(C_DOUBLE), INTENT(IN) :: x(*), xi(3), z(*)
    REAL(C_DOUBLE), INTENT(OUT) :: out
    TYPE(LPoly_type) :: lp(3)
    INTEGER :: i1(3), i2(3), d, o, a, b, c
    REAL(fp) :: xv, sub1(4), sub2(4,4), sub3(4,4,4)
    LOGICAL :: oob
    o = 0
    DO d = 1, nd
      xv = MAX(MIN(x(o+n(d)), xi(d)), x(o+1))
      CALL Find_Index(x(o+1:o+n(d)), xv, i1(d), i2(d), oob)
      CALL LPoly(x(o+i1(d):o+i2(d)), xv, lp(d))
      o = o + n(d)
    END DO
    SELECT CASE (nd)
      CASE (1)
        sub1 = z(i1(1):i2(1))
        CALL Interp_1D(sub1, lp(1), out)
      CASE (2)
        DO b = 1, 4
          DO a = 1, 4
            sub2(a,b) = z((i1(1)+a-1) + n(1)*(i1(2)+b-2))
          END DO
        END DO
        CALL Interp_2D(sub2, lp(1), lp(2), out)
      CASE (3)
        DO c = 1, 4
          DO b = 1, 4
            DO a = 1, 4
              sub3(a,b,c) = z((i1(1)+a-1) + n(1)*((i1(2)+b-2) + n(2)*(i1(3)+c-2)))
            END DO
          END DO
        END DO
        CALL Interp_3D(sub3, lp(1), lp(2), lp(3), out)
    END SELECT
  END SUBROUTINE ref_interp

  ! CRTM_AtmOptics_Combine on K layers: od, ssa in/out, pc(0:nl, K) in/out, sod out
  SUBROUTINE ref_combine(nl, k, od, ssa, pc, sod) BIND(C, NAME='ref_combine')
    INTEGER(C_INT), VALUE :: nl, k
    REAL(C_DOUBLE), INTENT(INOUT) :: od(k), ssa(k), pc(0:nl, k)
    REAL(C_DOUBLE), INTENT(OUT) :: sod
    TYPE(CRTM_AtmOptics_type) :: ao
    TYPE(AOvar_type) :: aov
    CALL CRTM_AtmOptics_Create(ao, k, nl, 1)
    CALL AOvar_Create(aov, k)
    ao%Optical_Depth = od
    ao%Single_Scatter_Albedo = ssa
    ao%Phase_Coefficient = 0.0_fp
    ao%Phase_Coefficient(0:nl, 1, 1:k) = pc
    ao%Backscat_Coefficient = 0.0_fp
    CALL CRTM_AtmOptics_Combine(ao, aov)
    od = ao%Optical_Depth
    ssa = ao%Single_Scatter_Albedo
    pc = ao%Phase_Coefficient(0:nl, 1, 1:k)
    sod = ao%Scattering_Optical_Depth
  END SUBROUTINE ref_combine

  ! AerosolCoeff_Binary_ReadFile with the 'CRTM' scheme, then AerosolCoeff_Frequency;
  ! dims = n_Wavelengths, n_Radii, n_Sigma, n_Types, n_RH, n_RH_Radii, n_Legendre_Terms,
  ! n_Phase_Elements, Release, Version
  INTEGER(C_INT) FUNCTION ref_aerosolcoeff_load(path, n, dims) BIND(C, NAME='ref_aerosolcoeff_load')
    CHARACTER(KIND=C_CHAR), INTENT(IN) :: path(*)
    INTEGER(C_INT), VALUE :: n
    INTEGER(C_INT), INTENT(OUT) :: dims(10)
    CALL AerosolCoeff_Destroy(acf)
    ref_aerosolcoeff_load = AerosolCoeff_Binary_ReadFile('CRTM', c2f(path, n), acf, Quiet=.TRUE.)
    IF (ref_aerosolcoeff_load /= 0) RETURN
    CALL AerosolCoeff_Frequency(acf)
    dims = (/ acf%n_Wavelengths, acf%n_Radii, acf%n_Sigma, acf%n_Types, acf%n_RH, acf%n_RH_Radii, &
              acf%n_Legendre_Terms, acf%n_Phase_Elements, acf%Release, acf%Version /)
  END FUNCTION ref_aerosolcoeff_load

  ! Copy the loaded AerosolCoeff arrays out (column-major, as stored)
  SUBROUTINE ref_aerosolcoeff_fetch(tid, wl, fr, reff, rsig, rh, ke, w, g, kb, pc) &
      BIND(C, NAME='ref_aerosolcoeff_fetch')
    INTEGER(C_INT), INTENT(OUT) :: tid(*)
    REAL(C_DOUBLE), INTENT(OUT) :: wl(*), fr(*), reff(*), rsig(*), rh(*), ke(*), w(*), g(*), kb(*), pc(*)
    tid(1:SIZE(acf%Type)) = acf%Type
    wl(1:SIZE(acf%Wavelength)) = acf%Wavelength
    fr(1:SIZE(acf%Frequency)) = acf%Frequency
    reff(1:SIZE(acf%Reff)) = RESHAPE(acf%Reff, (/SIZE(acf%Reff)/))
    rsig(1:SIZE(acf%Rsig)) = RESHAPE(acf%Rsig, (/SIZE(acf%Rsig)/))
    rh(1:SIZE(acf%RH)) = acf%RH
    ke(1:SIZE(acf%ke)) = RESHAPE(acf%ke, (/SIZE(acf%ke)/))
    w(1:SIZE(acf%w)) = RESHAPE(acf%w, (/SIZE(acf%w)/))
    g(1:SIZE(acf%g)) = RESHAPE(acf%g, (/SIZE(acf%g)/))
    kb(1:SIZE(acf%kb)) = RESHAPE(acf%kb, (/SIZE(acf%kb)/))
    pc(1:SIZE(acf%pcoeff)) = RESHAPE(acf%pcoeff, (/SIZE(acf%pcoeff)/))
  END SUBROUTINE ref_aerosolcoeff_fetch

  ! AerosolCoeff_typeID_to_index of the loaded table (1-based; -1 when absent)
  INTEGER(C_INT) FUNCTION ref_aerosol_type_index(id) BIND(C, NAME='ref_aerosol_type_index')
    INTEGER(C_INT), VALUE :: id
    ref_aerosol_type_index = AerosolCoeff_typeID_to_index(acf, INT(id))
  END FUNCTION ref_aerosol_type_index

  ! SpcCoeff_Binary_ReadFile; dims = Release, Version, n_Channels, Sensor_Type,
  ! WMO_Satellite_ID, WMO_Sensor_ID
  INTEGER(C_INT) FUNCTION ref_spccoeff_load(path, n, dims) BIND(C, NAME='ref_spccoeff_load')
    CHARACTER(KIND=C_CHAR), INTENT(IN) :: path(*)
    INTEGER(C_INT), VALUE :: n
    INTEGER(C_INT), INTENT(OUT) :: dims(6)
    CALL SpcCoeff_Destroy(scf)
    ref_spccoeff_load = SpcCoeff_Binary_ReadFile(c2f(path, n), scf, Quiet=.TRUE.)
    IF (ref_spccoeff_load /= 0) RETURN
    dims = (/ scf%Release, scf%Version, scf%n_Channels, scf%Sensor_Type, scf%WMO_Satellite_ID, &
              scf%WMO_Sensor_ID /)
  END FUNCTION ref_spccoeff_load

  ! Copy the loaded SpcCoeff arrays out; sid receives the 20-character Sensor_Id
  SUBROUTINE ref_spccoeff_fetch(sid, chan, pol, pang, flag, f, nu, c1, c2, bc1, bc2, cosmic, solar) &
      BIND(C, NAME='ref_spccoeff_fetch')
    CHARACTER(KIND=C_CHAR), INTENT(OUT) :: sid(20)
    INTEGER(C_INT), INTENT(OUT) :: chan(*), pol(*), flag(*)
    REAL(C_DOUBLE), INTENT(OUT) :: pang(*), f(*), nu(*), c1(*), c2(*), bc1(*), bc2(*), cosmic(*), solar(*)
    INTEGER :: i, L
    L = scf%n_Channels
    DO i = 1, 20
      sid(i) = scf%Sensor_Id(i:i)
    END DO
    chan(1:L) = scf%Sensor_Channel
    pol(1:L) = scf%Polarization
    pang(1:L) = scf%PolAngle
    flag(1:L) = scf%Channel_Flag
    f(1:L) = scf%Frequency
    nu(1:L) = scf%Wavenumber
    c1(1:L) = scf%Planck_C1
    c2(1:L) = scf%Planck_C2
    bc1(1:L) = scf%Band_C1
    bc2(1:L) = scf%Band_C2
    cosmic(1:L) = scf%Cosmic_Background_Radiance
    solar(1:L) = scf%Solar_Irradiance
  END SUBROUTINE ref_spccoeff_fetch

  ! Antenna-correction block of the loaded SpcCoeff (SpcCoeff%AC, ACCoeff_type):
  ! A_earth, A_space, A_platform (n_FOVs, n_Channels) column-major; returns
  ! n_FOVs, 0 when the file has no such block
  INTEGER(C_INT) FUNCTION ref_spccoeff_ac(ae, asp, ap) BIND(C, NAME='ref_spccoeff_ac')
    REAL(C_DOUBLE), INTENT(OUT) :: ae(*), asp(*), ap(*)
    INTEGER :: n
    ref_spccoeff_ac = 0
    IF (.NOT. ACCoeff_Associated(scf%AC)) RETURN
    n = SIZE(scf%AC%A_earth)
    ae(1:n) = RESHAPE(scf%AC%A_earth, (/n/))
    asp(1:n) = RESHAPE(scf%AC%A_space, (/n/))
    ap(1:n) = RESHAPE(scf%AC%A_platform, (/n/))
    ref_spccoeff_ac = scf%AC%n_FOVs
  END FUNCTION ref_spccoeff_ac

  ! Non-LTE block of the loaded SpcCoeff (SpcCoeff%NC, NLTECoeff_type): dims =
  ! n_Predictors, n_Sensor_Angles, n_Solar_Angles, n_NLTE_Channels; lev =
  ! Upper_Plevel(2), Lower_Plevel(2); tm = Min_Tm(2), Max_Tm(2), Mean_Tm(2);
  ! returns 1, 0 when the file has no such block
  INTEGER(C_INT) FUNCTION ref_spccoeff_nc(dims, lev, tm, ssen, ssol, cidx, isnlte, c) &
      BIND(C, NAME='ref_spccoeff_nc')
    INTEGER(C_INT), INTENT(OUT) :: dims(4), cidx(*), isnlte(*)
    REAL(C_DOUBLE), INTENT(OUT) :: lev(4), tm(6), ssen(*), ssol(*), c(*)
    INTEGER :: n
    ref_spccoeff_nc = 0
    IF (.NOT. NLTECoeff_Associated(scf%NC)) RETURN
    dims = (/ scf%NC%n_Predictors, scf%NC%n_Sensor_Angles, scf%NC%n_Solar_Angles, scf%NC%n_NLTE_Channels /)
    lev = (/ scf%NC%Upper_Plevel, scf%NC%Lower_Plevel /)
    tm = (/ scf%NC%Min_Tm, scf%NC%Max_Tm, scf%NC%Mean_Tm /)
    ssen(1:dims(2)) = scf%NC%Secant_Sensor_Zenith
    ssol(1:dims(3)) = scf%NC%Secant_Solar_Zenith
    cidx(1:scf%NC%n_Channels) = scf%NC%C_Index
    isnlte(1:scf%NC%n_Channels) = MERGE(1, 0, scf%NC%Is_NLTE_Channel)
    n = SIZE(scf%NC%C)
    c(1:n) = RESHAPE(scf%NC%C, (/n/))
    ref_spccoeff_nc = 1
  END FUNCTION ref_spccoeff_nc

  ! MWwaterCoeff_ReadFile (FASTEM coefficient file)
  INTEGER(C_INT) FUNCTION ref_mwwater_load(path, n) BIND(C, NAME='ref_mwwater_load')
    CHARACTER(KIND=C_CHAR), INTENT(IN) :: path(*)
    INTEGER(C_INT), VALUE :: n
    CALL MWwaterCoeff_Destroy(mwc)
    ref_mwwater_load = MWwaterCoeff_ReadFile(mwc, c2f(path, n), Quiet=.TRUE.)
  END FUNCTION ref_mwwater_load

  ! Compute_FastemX with both optional inputs given (an azimuth beyond +-360
  ! or a transmittance outside (0,1) switches that part off, as in the
  ! reference); the internal variables are kept for ref_fastemx_ad
  SUBROUTINE ref_fastemx(freq, n_angles, zenith, t, s, u, az, tr, emis, refl) BIND(C, NAME='ref_fastemx')
    REAL(C_DOUBLE), VALUE :: freq, zenith, t, s, u, az, tr
    INTEGER(C_INT), VALUE :: n_angles
    REAL(C_DOUBLE), INTENT(OUT) :: emis(4), refl(4)
    REAL(fp) :: e(4), r(4)
    CALL Compute_FastemX(mwc, freq, INT(n_angles), zenith, t, s, u, fx_var, e, r, &
                         Azimuth_Angle=az, Transmittance=tr)
    emis = e
    refl = r
  END SUBROUTINE ref_fastemx

  ! Compute_FastemX_AD of the last ref_fastemx call; out5 = adjoints of
  ! temperature, salinity, wind speed, azimuth angle, transmittance
  SUBROUTINE ref_fastemx_ad(emis_ad, refl_ad, out5) BIND(C, NAME='ref_fastemx_ad')
    REAL(C_DOUBLE), INTENT(IN) :: emis_ad(4), refl_ad(4)
    REAL(C_DOUBLE), INTENT(OUT) :: out5(5)
    REAL(fp) :: e(4), r(4), ta, sa, ua, aa, tra
    e = emis_ad
    r = refl_ad
    ta = 0.0_fp; sa = 0.0_fp; ua = 0.0_fp; aa = 0.0_fp; tra = 0.0_fp
    CALL Compute_FastemX_AD(mwc, e, r, fx_var, ta, sa, ua, Azimuth_Angle_AD=aa, Transmittance_AD=tra)
    out5 = (/ ta, sa, ua, aa, tra /)
  END SUBROUTINE ref_fastemx_ad

  ! IRwaterCoeff_ReadFile; dims = n_Angles, n_Frequencies, n_Wind_Speeds, n_Temperature
  INTEGER(C_INT) FUNCTION ref_irwater_load(path, n, dims) BIND(C, NAME='ref_irwater_load')
    CHARACTER(KIND=C_CHAR), INTENT(IN) :: path(*)
    INTEGER(C_INT), VALUE :: n
    INTEGER(C_INT), INTENT(OUT) :: dims(4)
    CALL IRwaterCoeff_Destroy(irc)
    ref_irwater_load = IRwaterCoeff_ReadFile(irc, c2f(path, n), Quiet=.TRUE.)
    IF (ref_irwater_load /= 0) RETURN
    dims = (/ irc%n_Angles, irc%n_Frequencies, irc%n_Wind_Speeds, irc%n_Temperature /)
  END FUNCTION ref_irwater_load

  ! CRTM_Compute_IRSSEM at n angles (degrees)
  INTEGER(C_INT) FUNCTION ref_irssem(t, u, f, n, angles, emis) BIND(C, NAME='ref_irssem')
    REAL(C_DOUBLE), VALUE :: t, u, f
    INTEGER(C_INT), VALUE :: n
    REAL(C_DOUBLE), INTENT(IN) :: angles(n)
    REAL(C_DOUBLE), INTENT(OUT) :: emis(n)
    REAL(fp) :: a(n), e(n)
    a = angles
    ref_irssem = CRTM_Compute_IRSSEM(irc, t, u, f, a, ir_var, e)
    emis = e
  END FUNCTION ref_irssem

  ! CRTM_Compute_IRSSEM_AD of the last ref_irssem call; out2 = wind speed and
  ! water temperature adjoints
  INTEGER(C_INT) FUNCTION ref_irssem_ad(n, emis_ad, out2) BIND(C, NAME='ref_irssem_ad')
    INTEGER(C_INT), VALUE :: n
    REAL(C_DOUBLE), INTENT(IN) :: emis_ad(n)
    REAL(C_DOUBLE), INTENT(OUT) :: out2(2)
    REAL(fp) :: e(n), ua, ta
    e = emis_ad
    ua = 0.0_fp
    ta = 0.0_fp
    ref_irssem_ad = CRTM_Compute_IRSSEM_AD(irc, e, ir_var, ua, ta)
    out2 = (/ ua, ta /)
  END FUNCTION ref_irssem_ad

  ! CRTM_Compute_Transmittance and its adjoint on k layers of optical depth
  SUBROUTINE ref_transmittance(k, od, tr_ad, tr, od_ad) BIND(C, NAME='ref_transmittance')
    INTEGER(C_INT), VALUE :: k
    REAL(C_DOUBLE), INTENT(IN) :: od(k)
    REAL(C_DOUBLE), VALUE :: tr_ad
    REAL(C_DOUBLE), INTENT(OUT) :: tr, od_ad(k)
    TYPE(CRTM_AtmOptics_type) :: ao, ao_ad
    REAL(fp) :: t, ta
    CALL CRTM_AtmOptics_Create(ao, k, 1, 1)
    CALL CRTM_AtmOptics_Create(ao_ad, k, 1, 1)
    ao%Optical_Depth = od
    ao_ad%Optical_Depth = 0.0_fp
    CALL CRTM_Compute_Transmittance(ao, t)
    tr = t
    ta = tr_ad
    CALL CRTM_Compute_Transmittance_AD(ao, ta, ao_ad)
    od_ad = ao_ad%Optical_Depth
  END SUBROUTINE ref_transmittance

  ! Geometry files through the reference writer and reader
  ! (CRTM_Geometry_WriteFile / CRTM_Geometry_ReadFile).  Per profile:
  ! ints(1:4) = iFOV, Year, Month, Day; reals(1:9) = Longitude, Latitude,
  ! Surface_Altitude, Sensor_Scan_Angle, Sensor_Zenith_Angle,
  ! Sensor_Azimuth_Angle, Source_Zenith_Angle, Source_Azimuth_Angle,
  ! Flux_Zenith_Angle.
  INTEGER(C_INT) FUNCTION ref_geometry_write(path, n, m, ints, reals) BIND(C, NAME='ref_geometry_write')
    USE CRTM_Geometry_Define, ONLY: CRTM_Geometry_WriteFile
    CHARACTER(KIND=C_CHAR), INTENT(IN) :: path(*)
    INTEGER(C_INT), VALUE :: n, m
    INTEGER(C_INT), INTENT(IN) :: ints(4, m)
    REAL(C_DOUBLE), INTENT(IN) :: reals(9, m)
    TYPE(CRTM_Geometry_type) :: g(m)
    INTEGER :: i
    DO i = 1, m
      g(i)%iFOV = ints(1, i); g(i)%Year = ints(2, i); g(i)%Month = ints(3, i); g(i)%Day = ints(4, i)
      g(i)%Longitude = reals(1, i); g(i)%Latitude = reals(2, i); g(i)%Surface_Altitude = reals(3, i)
      g(i)%Sensor_Scan_Angle = reals(4, i); g(i)%Sensor_Zenith_Angle = reals(5, i)
      g(i)%Sensor_Azimuth_Angle = reals(6, i); g(i)%Source_Zenith_Angle = reals(7, i)
      g(i)%Source_Azimuth_Angle = reals(8, i); g(i)%Flux_Zenith_Angle = reals(9, i)
    END DO
    ref_geometry_write = CRTM_Geometry_WriteFile(c2f(path, n), g, Quiet=.TRUE.)
  END FUNCTION ref_geometry_write

  ! Reads up to mcap profiles; mout = the file's profile count
  INTEGER(C_INT) FUNCTION ref_geometry_read(path, n, mcap, mout, ints, reals) BIND(C, NAME='ref_geometry_read')
    USE CRTM_Geometry_Define, ONLY: CRTM_Geometry_ReadFile
    CHARACTER(KIND=C_CHAR), INTENT(IN) :: path(*)
    INTEGER(C_INT), VALUE :: n, mcap
    INTEGER(C_INT), INTENT(OUT) :: mout
    INTEGER(C_INT), INTENT(OUT) :: ints(4, mcap)
    REAL(C_DOUBLE), INTENT(OUT) :: reals(9, mcap)
    TYPE(CRTM_Geometry_type), ALLOCATABLE :: g(:)
    INTEGER :: i
    ref_geometry_read = CRTM_Geometry_ReadFile(c2f(path, n), g, Quiet=.TRUE.)
    mout = 0
    IF (ref_geometry_read /= 0) RETURN
    mout = SIZE(g)
    DO i = 1, MIN(mout, mcap)
      ints(:, i) = (/ g(i)%iFOV, g(i)%Year, g(i)%Month, g(i)%Day /)
      reals(:, i) = (/ g(i)%Longitude, g(i)%Latitude, g(i)%Surface_Altitude, g(i)%Sensor_Scan_Angle, &
                       g(i)%Sensor_Zenith_Angle, g(i)%Sensor_Azimuth_Angle, g(i)%Source_Zenith_Angle, &
                       g(i)%Source_Azimuth_Angle, g(i)%Flux_Zenith_Angle /)
    END DO
  END FUNCTION ref_geometry_read

  ! Cloud files through CRTM_Cloud_WriteFile / CRTM_Cloud_ReadFile: nc clouds
  ! of k layers each; reff, rvar, wc (k, nc)
  INTEGER(C_INT) FUNCTION ref_cloud_write(path, n, nc, k, types, reff, rvar, wc) BIND(C, NAME='ref_cloud_write')
    USE CRTM_Cloud_Define, ONLY: CRTM_Cloud_type, CRTM_Cloud_Create, CRTM_Cloud_WriteFile
    CHARACTER(KIND=C_CHAR), INTENT(IN) :: path(*)
    INTEGER(C_INT), VALUE :: n, nc, k
    INTEGER(C_INT), INTENT(IN) :: types(nc)
    REAL(C_DOUBLE), INTENT(IN) :: reff(k, nc), rvar(k, nc), wc(k, nc)
    TYPE(CRTM_Cloud_type) :: c(nc)
    INTEGER :: i
    DO i = 1, nc
      CALL CRTM_Cloud_Create(c(i), k)
      c(i)%Type = types(i)
      c(i)%Effective_Radius = reff(:, i)
      c(i)%Effective_Variance = rvar(:, i)
      c(i)%Water_Content = wc(:, i)
    END DO
    ref_cloud_write = CRTM_Cloud_WriteFile(c2f(path, n), c, Quiet=.TRUE.)
  END FUNCTION ref_cloud_write

  INTEGER(C_INT) FUNCTION ref_cloud_read(path, n, nccap, kcap, ncout, kout, types, reff, rvar, wc) &
      BIND(C, NAME='ref_cloud_read')
    USE CRTM_Cloud_Define, ONLY: CRTM_Cloud_type, CRTM_Cloud_ReadFile
    CHARACTER(KIND=C_CHAR), INTENT(IN) :: path(*)
    INTEGER(C_INT), VALUE :: n, nccap, kcap
    INTEGER(C_INT), INTENT(OUT) :: ncout, kout
    INTEGER(C_INT), INTENT(OUT) :: types(nccap)
    REAL(C_DOUBLE), INTENT(OUT) :: reff(kcap, nccap), rvar(kcap, nccap), wc(kcap, nccap)
    TYPE(CRTM_Cloud_type), ALLOCATABLE :: c(:)
    INTEGER :: i, kk
    ref_cloud_read = CRTM_Cloud_ReadFile(c2f(path, n), c, Quiet=.TRUE.)
    ncout = 0; kout = 0
    IF (ref_cloud_read /= 0) RETURN
    ncout = SIZE(c)
    DO i = 1, MIN(ncout, nccap)
      kk = MIN(c(i)%n_Layers, kcap)
      kout = c(i)%n_Layers
      types(i) = c(i)%Type
      reff(1:kk, i) = c(i)%Effective_Radius(1:kk)
      rvar(1:kk, i) = c(i)%Effective_Variance(1:kk)
      wc(1:kk, i) = c(i)%Water_Content(1:kk)
    END DO
  END FUNCTION ref_cloud_read

END MODULE ref_harness
