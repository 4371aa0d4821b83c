! crtm_mi355x_module.f90 -- Fortran host interface of the MI355X CRTM hot path.
!
! A thin ISO_C_BINDING layer over libcrtm_mi355x.so (include/crtm_mi355x.h)
! that keeps the names, argument order and result codes of CRTM's public
! API for the forward / K-matrix path (clear and cloudy); crtm_module.f90
! re-exports it as MODULE CRTM_Module (src/CRTM_Module.F90:7-38), so a caller
! written against CRTM_Module compiles unchanged for that path:
!
!   CRTM_Init(Sensor_ID, ChannelInfo, File_Path)          src/CRTM_LifeCycle.f90:520
!   CRTM_Forward(Atm, Sfc, Geo, ChannelInfo, RTS, Opt)    src/CRTM_Forward_Module.f90:229-245
!   CRTM_K_Matrix(Atm, Sfc, RTS_K, Geo, ChannelInfo, Atm_K, Sfc_K, RTS, Opt)
!                                                         src/CRTM_K_Matrix_Module.f90:300-325
!   CRTM_Destroy(ChannelInfo)                             src/CRTM_LifeCycle.f90:1131
!   CRTM_Atmosphere_Create(Atm, n_Layers, n_Absorbers, n_Clouds, n_Aerosols)
!                                                         src/Atmosphere/CRTM_Atmosphere_Define.f90:612
!
! Derived types carry the components of the reference types this path
! reads or writes (src/Atmosphere/CRTM_Atmosphere_Define.f90:433-464,
! src/Surface/CRTM_Surface_Define.f90:204-244,
! src/GeometryInfo/CRTM_Geometry_Define.f90:99-121,
! src/RTSolution/CRTM_RTSolution_Define.f90:201-245,
! src/Options/CRTM_Options_Define.f90:122-179).  As in the reference,
! Atmosphere_K / Surface_K are accumulated into (callers zero them first).
! Without Options%Use_Emissivity the surface optics of water surfaces are
! computed (FASTEM6, IR sea-surface emissivity table; Surface%Wind_Speed,
! %Wind_Direction, %Salinity, Geometry%Sensor_Azimuth_Angle), and Surface_K
! carries their sensitivities.  Components of the reference types that feed
! models outside this path (land, snow and ice surface models, NLTE,
! antenna correction, solar geometry) are present so that callers compile.  Helpers a CRTM caller uses around the calls (CRTM_Version,
! Display_Message, CRTM_Geometry_SetValue, CRTM_ChannelInfo_n_Channels,
! CRTM_*_Associated / _Destroy / _Inspect) are provided with the reference's
! names and arguments.
MODULE CRTM_MI355X_Module
  USE ISO_C_BINDING
  IMPLICIT NONE
  PRIVATE

  INTEGER, PARAMETER, PUBLIC :: fp = C_DOUBLE
  REAL(fp), PARAMETER, PUBLIC :: ZERO = 0.0_fp, ONE = 1.0_fp
  ! src/Utility/Message_Handler.f90:27-31
  INTEGER, PARAMETER, PUBLIC :: SUCCESS = 0, INFORMATION = 1, WARNING = 2, FAILURE = 3, UNDEFINED = 4
  ! src/Coefficients/SensorInfo/SensorInfo_Parameters.f90 (sensor types)
  INTEGER, PARAMETER, PUBLIC :: INVALID_SENSOR = 0, MICROWAVE_SENSOR = 1, INFRARED_SENSOR = 2, &
                                VISIBLE_SENSOR = 3, ULTRAVIOLET_SENSOR = 4
  INTEGER, PARAMETER, PUBLIC :: H2O_ID = 1, CO2_ID = 2, O3_ID = 3, N2O_ID = 4, CO_ID = 5, CH4_ID = 6
  INTEGER, PARAMETER, PUBLIC :: VOLUME_MIXING_RATIO_UNITS = 1, NUMBER_DENSITY_UNITS = 2, &
                                MASS_MIXING_RATIO_UNITS = 3, MASS_DENSITY_UNITS = 4, PARTIAL_PRESSURE_UNITS = 5
  ! src/Atmosphere/CRTM_Atmosphere_Define.f90:404-411
  INTEGER, PARAMETER, PUBLIC :: INVALID_MODEL = 0, TROPICAL = 1, MIDLATITUDE_SUMMER = 2, MIDLATITUDE_WINTER = 3, &
                                SUBARCTIC_SUMMER = 4, SUBARCTIC_WINTER = 5, US_STANDARD_ATMOSPHERE = 6
  ! src/CRTM_Parameters.f90:286-294
  INTEGER, PARAMETER, PUBLIC :: INVALID_AEROSOL = 0, DUST_AEROSOL = 1, SEASALT_SSAM_AEROSOL = 2, &
                                SEASALT_SSCM1_AEROSOL = 3, SEASALT_SSCM2_AEROSOL = 4, SEASALT_SSCM3_AEROSOL = 5, &
                                ORGANIC_CARBON_AEROSOL = 6, BLACK_CARBON_AEROSOL = 7, SULFATE_AEROSOL = 8
  ! src/CRTM_Parameters.f90:247-250: the ADA solver is the one this path implements
  INTEGER, PARAMETER, PUBLIC :: RT_ADA = 56, RT_AMOM = 56, RT_SOI = 168
  INTEGER, PARAMETER, PUBLIC :: CRTM_PREC_FP64 = 0, CRTM_PREC_FP32_SPLIT = 1

  TYPE, PUBLIC :: CRTM_ChannelInfo_type
    CHARACTER(LEN=20) :: Sensor_ID = ''
    INTEGER :: Sensor_Type = 0
    INTEGER :: n_Channels = 0
    INTEGER, ALLOCATABLE :: Sensor_Channel(:)
    INTEGER :: Channel_Offset = 0
    TYPE(C_PTR) :: ctx = C_NULL_PTR      ! shared by all sensors of one CRTM_Init call
  END TYPE

  ! CRTM_Cloud_type (src/Atmosphere/Cloud/CRTM_Cloud_Define.f90:295-310)
  TYPE, PUBLIC :: CRTM_Cloud_type
    INTEGER :: n_Layers = 0
    INTEGER :: Type = 0                            ! WATER_CLOUD .. HAIL_CLOUD
    REAL(fp), ALLOCATABLE :: Effective_Radius(:)   ! microns
    REAL(fp), ALLOCATABLE :: Water_Content(:)      ! kg/m^2
  END TYPE
  INTEGER, PARAMETER, PUBLIC :: WATER_CLOUD = 1, ICE_CLOUD = 2, RAIN_CLOUD = 3, SNOW_CLOUD = 4, &
                                GRAUPEL_CLOUD = 5, HAIL_CLOUD = 6

  ! CRTM_Aerosol_type (src/Atmosphere/Aerosol/CRTM_Aerosol_Define.f90:119-133)
  TYPE, PUBLIC :: CRTM_Aerosol_type
    INTEGER :: n_Layers = 0
    INTEGER :: Type = 0                            ! aerosol type id of the AerosolCoeff tables
    REAL(fp), ALLOCATABLE :: Effective_Radius(:)   ! microns
    REAL(fp), ALLOCATABLE :: Concentration(:)      ! kg/m^2
  END TYPE

  TYPE, PUBLIC :: CRTM_Atmosphere_type
    INTEGER :: n_Layers = 0, n_Absorbers = 0, n_Clouds = 0, n_Aerosols = 0
    INTEGER :: Climatology = 6                     ! US_STANDARD_ATMOSPHERE
    INTEGER, ALLOCATABLE :: Absorber_ID(:), Absorber_Units(:)
    REAL(fp), ALLOCATABLE :: Level_Pressure(:)     ! (0:n_Layers)
    REAL(fp), ALLOCATABLE :: Pressure(:), Temperature(:)
    REAL(fp), ALLOCATABLE :: Absorber(:, :)        ! (n_Layers, n_Absorbers)
    REAL(fp), ALLOCATABLE :: Cloud_Fraction(:)     ! (n_Layers)
    TYPE(CRTM_Cloud_type), ALLOCATABLE :: Cloud(:) ! (n_Clouds)
    TYPE(CRTM_Aerosol_type), ALLOCATABLE :: Aerosol(:) ! (n_Aerosols)
  END TYPE

  ! src/Surface/CRTM_Surface_Define.f90:204-244: the coverages and
  ! temperatures enter every call; Wind_Speed, Wind_Direction and Salinity
  ! the computed optics of water surfaces (Options%Use_Emissivity = .FALSE.);
  ! the land, snow and ice type and state components feed surface models
  ! this path does not have (those surfaces need Options%Emissivity)
  TYPE, PUBLIC :: CRTM_Surface_type
    REAL(fp) :: Land_Coverage = 0, Water_Coverage = 0, Snow_Coverage = 0, Ice_Coverage = 0
    REAL(fp) :: Land_Temperature = 283, Water_Temperature = 283, Snow_Temperature = 263, Ice_Temperature = 263
    INTEGER :: Land_Type = 0, Water_Type = 1, Snow_Type = 0, Ice_Type = 0
    INTEGER :: Soil_Type = 0, Vegetation_Type = 0
    REAL(fp) :: Soil_Moisture_Content = 0.05_fp, Canopy_Water_Content = 0.05_fp, Vegetation_Fraction = 0.3_fp
    REAL(fp) :: Soil_Temperature = 283, Lai = 3.5_fp
    REAL(fp) :: Wind_Speed = 5, Wind_Direction = 0, Salinity = 33
    REAL(fp) :: Snow_Depth = 50, Snow_Density = 0.2_fp, Snow_Grain_Size = 2
    REAL(fp) :: Ice_Thickness = 10, Ice_Density = 0.9_fp, Ice_Roughness = 0
  END TYPE

  ! src/GeometryInfo/CRTM_Geometry_Define.f90:99-121
  TYPE, PUBLIC :: CRTM_Geometry_type
    INTEGER :: iFOV = 0
    REAL(fp) :: Longitude = 0, Latitude = 0, Surface_Altitude = 0   ! degrees, degrees, metres
    REAL(fp) :: Sensor_Scan_Angle = 0, Sensor_Zenith_Angle = 0, Sensor_Azimuth_Angle = 999.9_fp
    REAL(fp) :: Source_Zenith_Angle = 100, Source_Azimuth_Angle = 0, Flux_Zenith_Angle = 53.130102_fp
    INTEGER :: Year = 2001, Month = 1, Day = 1
  END TYPE

  ! src/RTSolution/CRTM_RTSolution_Define.f90:201-245 (the components this path produces)
  TYPE, PUBLIC :: CRTM_RTSolution_type
    CHARACTER(LEN=20) :: Sensor_Id = ''
    INTEGER :: WMO_Satellite_Id = 1023, WMO_Sensor_Id = 2047, Sensor_Channel = 0
    REAL(fp) :: Radiance = 0, Brightness_Temperature = 0
    REAL(fp) :: Up_Radiance = 0, Down_Radiance = 0, Surface_Planck_Radiance = 0, Surface_Emissivity = 0
    REAL(fp) :: Surface_Reflectivity = 0
    REAL(fp) :: R_Clear = 0, Tb_Clear = 0, Total_Cloud_Cover = 0
  END TYPE

  ! src/Options/CRTM_Options_Define.f90:122-179 (the options this path honours)
  TYPE, PUBLIC :: CRTM_Options_type
    LOGICAL :: Check_Input = .TRUE.
    LOGICAL :: Use_Emissivity = .FALSE.
    REAL(fp), ALLOCATABLE :: Emissivity(:)         ! (n_Channels)
    LOGICAL :: Use_N_Streams = .FALSE.
    INTEGER :: n_Streams = 0
    LOGICAL :: Include_Scattering = .TRUE.
    INTEGER :: Overlap_Id = 4                      ! average overlap (CRTM_CloudCover_Define.f90:79)
    INTEGER :: RT_Algorithm_Id = RT_ADA
    LOGICAL :: Skip_Profile = .FALSE.
    ! src/Options/CRTM_Options_Define.f90:133,136 (one setting per call here)
    LOGICAL :: Use_Antenna_Correction = .FALSE.
    LOGICAL :: Apply_NLTE_Correction = .TRUE.
    INTEGER :: Precision = CRTM_PREC_FP64          ! extension: arithmetic of the GPU path
  END TYPE

  PUBLIC :: CRTM_Init, CRTM_Destroy, CRTM_Forward, CRTM_K_Matrix, CRTM_Adjoint, CRTM_Tangent_Linear
  PUBLIC :: CRTM_Atmosphere_Create, CRTM_Atmosphere_Zero, CRTM_Surface_Zero
  PUBLIC :: CRTM_Atmosphere_Associated, CRTM_Atmosphere_Destroy, CRTM_Atmosphere_Inspect
  PUBLIC :: CRTM_Surface_Inspect, CRTM_RTSolution_Inspect, CRTM_Geometry_SetValue
  PUBLIC :: CRTM_ChannelInfo_n_Channels, CRTM_Version, Display_Message, Program_Message

  ! ---- C-ABI mirrors (include/crtm_mi355x.h) ----
  TYPE, BIND(C) :: c_atmosphere
    INTEGER(C_INT) :: n_profiles, max_layers
    INTEGER(C_INT64_T) :: total_layers
    TYPE(C_PTR) :: n_layers, layer_offset
    INTEGER(C_INT) :: n_absorbers
    TYPE(C_PTR) :: absorber_id, absorber_units, level_pressure, pressure, temperature, absorber
    TYPE(C_PTR) :: climatology
    INTEGER(C_INT) :: n_clouds
    TYPE(C_PTR) :: cloud_type, cloud_water_content, cloud_effective_radius, cloud_fraction
    INTEGER(C_INT) :: n_aerosols
    TYPE(C_PTR) :: aerosol_type, aerosol_concentration, aerosol_effective_radius
  END TYPE
  TYPE, BIND(C) :: c_surface
    TYPE(C_PTR) :: land_coverage, water_coverage, snow_coverage, ice_coverage, &
                   land_temperature, water_temperature, snow_temperature, ice_temperature, &
                   wind_speed, wind_direction, salinity
  END TYPE
  TYPE, BIND(C) :: c_geometry
    TYPE(C_PTR) :: sensor_zenith_angle, sensor_scan_angle, source_zenith_angle, surface_altitude, &
                   sensor_azimuth_angle, ifov
  END TYPE
  TYPE, BIND(C) :: c_options
    INTEGER(C_INT) :: use_emissivity
    TYPE(C_PTR) :: emissivity
    INTEGER(C_INT) :: precision, memory, jacobian_f32, check_input
    INTEGER(C_INT) :: n_streams, overlap_id
    TYPE(C_PTR) :: n_streams_profile
    INTEGER(C_INT) :: no_scattering
    TYPE(C_PTR) :: skip_profile
    INTEGER(C_INT) :: use_antenna_correction, no_nlte_correction
  END TYPE
  TYPE, BIND(C) :: c_rtsolution
    TYPE(C_PTR) :: radiance, brightness_temperature, up_radiance, down_radiance, &
                   surface_planck_radiance, surface_emissivity, status, r_clear, tb_clear, &
                   total_cloud_cover, surface_reflectivity
  END TYPE
  TYPE, BIND(C) :: c_ad_output
    TYPE(C_PTR) :: brightness_temperature_ad, radiance_ad, surface_emissivity_ad, temperature_ad, absorber_ad, &
                   land_temperature_ad, water_temperature_ad, snow_temperature_ad, ice_temperature_ad, &
                   cloud_water_content_ad, cloud_effective_radius_ad, cloud_fraction_ad, &
                   aerosol_concentration_ad, aerosol_effective_radius_ad, &
                   salinity_ad, wind_speed_ad, wind_direction_ad
  END TYPE
  TYPE, BIND(C) :: c_k_output
    TYPE(C_PTR) :: brightness_temperature_k, radiance_k, surface_emissivity_k, temperature_k, &
                   absorber_k, land_temperature_k, water_temperature_k, snow_temperature_k, ice_temperature_k, &
                   cloud_water_content_k, cloud_effective_radius_k, cloud_fraction_k, &
                   aerosol_concentration_k, aerosol_effective_radius_k, &
                   salinity_k, wind_speed_k, wind_direction_k
  END TYPE

  INTERFACE
    INTEGER(C_INT) FUNCTION c_crtm_init(n, ids, path, ctx) BIND(C, NAME='crtm_init')
      IMPORT
      INTEGER(C_INT), VALUE :: n
      TYPE(C_PTR), INTENT(IN) :: ids(*)
      CHARACTER(KIND=C_CHAR), INTENT(IN) :: path(*)
      TYPE(C_PTR), INTENT(OUT) :: ctx
    END FUNCTION
    INTEGER(C_INT) FUNCTION c_crtm_destroy(ctx) BIND(C, NAME='crtm_destroy')
      IMPORT
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    INTEGER(C_INT) FUNCTION c_crtm_n_sensors(ctx) BIND(C, NAME='crtm_n_sensors')
      IMPORT
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    INTEGER(C_INT) FUNCTION c_crtm_channel_info(ctx, s, n, stype, chans, sid) BIND(C, NAME='crtm_channel_info')
      IMPORT
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT), VALUE :: s
      TYPE(C_PTR), VALUE :: n, stype, chans, sid
    END FUNCTION
    INTEGER(C_INT) FUNCTION c_crtm_forward(ctx, atm, sfc, geo, opt, rts) BIND(C, NAME='crtm_forward')
      IMPORT
      TYPE(C_PTR), VALUE :: ctx
      TYPE(c_atmosphere), INTENT(IN) :: atm
      TYPE(c_surface), INTENT(IN) :: sfc
      TYPE(c_geometry), INTENT(IN) :: geo
      TYPE(c_options), INTENT(IN) :: opt
      TYPE(c_rtsolution), INTENT(INOUT) :: rts
    END FUNCTION
    INTEGER(C_INT) FUNCTION c_crtm_k_matrix(ctx, atm, sfc, geo, opt, rts, k) BIND(C, NAME='crtm_k_matrix')
      IMPORT
      TYPE(C_PTR), VALUE :: ctx
      TYPE(c_atmosphere), INTENT(IN) :: atm
      TYPE(c_surface), INTENT(IN) :: sfc
      TYPE(c_geometry), INTENT(IN) :: geo
      TYPE(c_options), INTENT(IN) :: opt
      TYPE(c_rtsolution), INTENT(INOUT) :: rts
      TYPE(c_k_output), INTENT(INOUT) :: k
    END FUNCTION
    INTEGER(C_INT) FUNCTION c_crtm_adjoint(ctx, atm, sfc, geo, opt, rts, ad) BIND(C, NAME='crtm_adjoint')
      IMPORT
      TYPE(C_PTR), VALUE :: ctx
      TYPE(c_atmosphere), INTENT(IN) :: atm
      TYPE(c_surface), INTENT(IN) :: sfc
      TYPE(c_geometry), INTENT(IN) :: geo
      TYPE(c_options), INTENT(IN) :: opt
      TYPE(c_rtsolution), INTENT(INOUT) :: rts
      TYPE(c_ad_output), INTENT(INOUT) :: ad
    END FUNCTION
    FUNCTION c_crtm_last_error(ctx) RESULT(p) BIND(C, NAME='crtm_last_error')
      IMPORT
      TYPE(C_PTR), VALUE :: ctx
      TYPE(C_PTR) :: p
    END FUNCTION
  END INTERFACE

  ! Flat host arrays of one call (kept alive while the C-ABI reads them)
  TYPE :: batch_type
    INTEGER(C_INT), ALLOCATABLE :: nl(:), ids(:), units(:), clim(:)
    INTEGER(C_INT64_T), ALLOCATABLE :: off(:)
    REAL(C_DOUBLE), ALLOCATABLE :: lp(:), p(:), t(:), a(:), cov(:, :), st(:, :), zen(:), scan(:), src(:), emis(:)
    REAL(C_DOUBLE), ALLOCATABLE :: alt(:), ws(:), wd(:), sal(:), azi(:)
    INTEGER(C_INT), ALLOCATABLE :: skip(:), fov(:)
    REAL(C_DOUBLE), ALLOCATABLE :: rad(:), bt(:), up(:), dn(:), sp(:), se(:), sr(:), rcl(:), tbc(:), tcc(:)
    INTEGER(C_INT), ALLOCATABLE :: status(:), ctype(:)
    REAL(C_DOUBLE), ALLOCATABLE :: cwc(:), creff(:), cfrac(:)
    INTEGER(C_INT), ALLOCATABLE :: atype(:)
    REAL(C_DOUBLE), ALLOCATABLE :: aconc(:), areff(:)
    INTEGER(C_INT), ALLOCATABLE :: nsp(:)
    TYPE(c_atmosphere) :: atm
    TYPE(c_surface) :: sfc
    TYPE(c_geometry) :: geo
    TYPE(c_options) :: opt
    TYPE(c_rtsolution) :: rts
  END TYPE

  ! K-matrix outputs of one profile chunk, flat in the C layout
  TYPE :: k_buffers
    TYPE(batch_type) :: b
    INTEGER :: na = 0, nc = 0, nae = 0
    REAL(C_DOUBLE), ALLOCATABLE :: btk(:), radk(:), emk(:), tk(:), ak(:), stk(:, :), swk(:, :)
    REAL(C_DOUBLE), ALLOCATABLE :: cwk(:), crk(:), cfk(:), ack(:), ark(:)
  END TYPE

CONTAINS

  FUNCTION CRTM_Init(Sensor_ID, ChannelInfo, File_Path) RESULT(Error_Status)
    CHARACTER(*), INTENT(IN) :: Sensor_ID(:)
    TYPE(CRTM_ChannelInfo_type), INTENT(INOUT) :: ChannelInfo(:)
    CHARACTER(*), OPTIONAL, INTENT(IN) :: File_Path
    INTEGER :: Error_Status
    CHARACTER(KIND=C_CHAR, LEN=64), TARGET :: cid(SIZE(Sensor_ID))
    TYPE(C_PTR) :: pid(SIZE(Sensor_ID)), ctx
    CHARACTER(KIND=C_CHAR, LEN=1024) :: cpath
    INTEGER(C_INT), TARGET :: n, st
    CHARACTER(KIND=C_CHAR, LEN=32), TARGET :: sid
    INTEGER(C_INT), ALLOCATABLE, TARGET :: chans(:)
    INTEGER :: i, off
    IF (SIZE(ChannelInfo) < SIZE(Sensor_ID)) THEN
      Error_Status = FAILURE
      RETURN
    END IF
    DO i = 1, SIZE(Sensor_ID)
      cid(i) = TRIM(Sensor_ID(i)) // C_NULL_CHAR
      pid(i) = C_LOC(cid(i))
    END DO
    cpath = '.' // C_NULL_CHAR
    IF (PRESENT(File_Path)) cpath = TRIM(File_Path) // C_NULL_CHAR
    Error_Status = c_crtm_init(INT(SIZE(Sensor_ID), C_INT), pid, cpath, ctx)
    IF (Error_Status /= SUCCESS) RETURN
    off = 0
    DO i = 1, SIZE(Sensor_ID)
      Error_Status = c_crtm_channel_info(ctx, INT(i - 1, C_INT), C_LOC(n), C_LOC(st), C_NULL_PTR, C_NULL_PTR)
      IF (Error_Status /= SUCCESS) RETURN
      ChannelInfo(i)%n_Channels = n
      ChannelInfo(i)%Sensor_Type = st
      IF (ALLOCATED(ChannelInfo(i)%Sensor_Channel)) DEALLOCATE(ChannelInfo(i)%Sensor_Channel)
      ALLOCATE(ChannelInfo(i)%Sensor_Channel(n))
      IF (ALLOCATED(chans)) DEALLOCATE(chans)
      ALLOCATE(chans(n))
      Error_Status = c_crtm_channel_info(ctx, INT(i - 1, C_INT), C_NULL_PTR, C_NULL_PTR, C_LOC(chans), C_LOC(sid))
      ChannelInfo(i)%Sensor_Channel = chans
      ChannelInfo(i)%Sensor_ID = c_to_f(sid)
      ChannelInfo(i)%Channel_Offset = off
      ChannelInfo(i)%ctx = ctx
      off = off + n
    END DO
  END FUNCTION CRTM_Init

  FUNCTION CRTM_Destroy(ChannelInfo) RESULT(Error_Status)
    TYPE(CRTM_ChannelInfo_type), INTENT(INOUT) :: ChannelInfo(:)
    INTEGER :: Error_Status
    Error_Status = SUCCESS
    IF (SIZE(ChannelInfo) < 1) RETURN
    IF (C_ASSOCIATED(ChannelInfo(1)%ctx)) Error_Status = c_crtm_destroy(ChannelInfo(1)%ctx)
    ChannelInfo(:)%ctx = C_NULL_PTR
  END FUNCTION CRTM_Destroy

  SUBROUTINE CRTM_Atmosphere_Create(Atm, n_Layers, n_Absorbers, n_Clouds, n_Aerosols)
    TYPE(CRTM_Atmosphere_type), INTENT(INOUT) :: Atm(:)
    INTEGER, INTENT(IN) :: n_Layers, n_Absorbers, n_Clouds, n_Aerosols
    INTEGER :: m, ic
    DO m = 1, SIZE(Atm)
      Atm(m)%n_Layers = n_Layers
      Atm(m)%n_Absorbers = n_Absorbers
      Atm(m)%n_Clouds = n_Clouds
      Atm(m)%n_Aerosols = n_Aerosols
      IF (ALLOCATED(Atm(m)%Level_Pressure)) DEALLOCATE(Atm(m)%Level_Pressure, Atm(m)%Pressure, &
          Atm(m)%Temperature, Atm(m)%Absorber, Atm(m)%Absorber_ID, Atm(m)%Absorber_Units)
      ALLOCATE(Atm(m)%Level_Pressure(0:n_Layers), Atm(m)%Pressure(n_Layers), Atm(m)%Temperature(n_Layers), &
               Atm(m)%Absorber(n_Layers, n_Absorbers), Atm(m)%Absorber_ID(n_Absorbers), &
               Atm(m)%Absorber_Units(n_Absorbers))
      Atm(m)%Level_Pressure = 0; Atm(m)%Pressure = 0; Atm(m)%Temperature = 0; Atm(m)%Absorber = 0
      Atm(m)%Absorber_ID = 0; Atm(m)%Absorber_Units = 0
      IF (ALLOCATED(Atm(m)%Cloud)) DEALLOCATE(Atm(m)%Cloud)
      IF (ALLOCATED(Atm(m)%Cloud_Fraction)) DEALLOCATE(Atm(m)%Cloud_Fraction)
      ALLOCATE(Atm(m)%Cloud(n_Clouds), Atm(m)%Cloud_Fraction(n_Layers))
      Atm(m)%Cloud_Fraction = 0
      DO ic = 1, n_Clouds
        Atm(m)%Cloud(ic)%n_Layers = n_Layers
        ALLOCATE(Atm(m)%Cloud(ic)%Effective_Radius(n_Layers), Atm(m)%Cloud(ic)%Water_Content(n_Layers))
        Atm(m)%Cloud(ic)%Effective_Radius = 0
        Atm(m)%Cloud(ic)%Water_Content = 0
      END DO
      IF (ALLOCATED(Atm(m)%Aerosol)) DEALLOCATE(Atm(m)%Aerosol)
      ALLOCATE(Atm(m)%Aerosol(n_Aerosols))
      DO ic = 1, n_Aerosols
        Atm(m)%Aerosol(ic)%n_Layers = n_Layers
        ALLOCATE(Atm(m)%Aerosol(ic)%Effective_Radius(n_Layers), Atm(m)%Aerosol(ic)%Concentration(n_Layers))
        Atm(m)%Aerosol(ic)%Effective_Radius = 0
        Atm(m)%Aerosol(ic)%Concentration = 0
      END DO
    END DO
  END SUBROUTINE CRTM_Atmosphere_Create

  SUBROUTINE CRTM_Atmosphere_Zero(Atm)
    TYPE(CRTM_Atmosphere_type), INTENT(INOUT) :: Atm(:, :)
    INTEGER :: i, j, n
    DO j = 1, SIZE(Atm, 2)
      DO i = 1, SIZE(Atm, 1)
        IF (ALLOCATED(Atm(i, j)%Temperature)) Atm(i, j)%Temperature = 0
        IF (ALLOCATED(Atm(i, j)%Absorber)) Atm(i, j)%Absorber = 0
        IF (ALLOCATED(Atm(i, j)%Cloud_Fraction)) Atm(i, j)%Cloud_Fraction = 0
        IF (ALLOCATED(Atm(i, j)%Cloud)) THEN
          DO n = 1, SIZE(Atm(i, j)%Cloud)
            Atm(i, j)%Cloud(n)%Water_Content = 0
            Atm(i, j)%Cloud(n)%Effective_Radius = 0
          END DO
        END IF
        IF (ALLOCATED(Atm(i, j)%Aerosol)) THEN
          DO n = 1, SIZE(Atm(i, j)%Aerosol)
            Atm(i, j)%Aerosol(n)%Concentration = 0
            Atm(i, j)%Aerosol(n)%Effective_Radius = 0
          END DO
        END IF
      END DO
    END DO
  END SUBROUTINE CRTM_Atmosphere_Zero

  ! CRTM_Surface_Zero (src/Surface/CRTM_Surface_Define.f90:461-479, with
  ! CRTM_WaterSurface_Zero :1904-1911): zero the coverages, temperatures,
  ! salinity and wind of a Surface_K array before CRTM_K_Matrix
  SUBROUTINE CRTM_Surface_Zero(Sfc)
    TYPE(CRTM_Surface_type), INTENT(INOUT) :: Sfc(:, :)
    Sfc(:, :)%Land_Coverage = 0; Sfc(:, :)%Water_Coverage = 0
    Sfc(:, :)%Snow_Coverage = 0; Sfc(:, :)%Ice_Coverage = 0
    Sfc(:, :)%Land_Temperature = 0; Sfc(:, :)%Water_Temperature = 0
    Sfc(:, :)%Snow_Temperature = 0; Sfc(:, :)%Ice_Temperature = 0
    Sfc(:, :)%Wind_Speed = 0; Sfc(:, :)%Wind_Direction = 0; Sfc(:, :)%Salinity = 0
  END SUBROUTINE CRTM_Surface_Zero

  FUNCTION CRTM_Forward(Atmosphere, Surface, Geometry, ChannelInfo, RTSolution, Options) RESULT(Error_Status)
    TYPE(CRTM_Atmosphere_type), INTENT(IN) :: Atmosphere(:)
    TYPE(CRTM_Surface_type), INTENT(IN) :: Surface(:)
    TYPE(CRTM_Geometry_type), INTENT(IN) :: Geometry(:)
    TYPE(CRTM_ChannelInfo_type), INTENT(IN) :: ChannelInfo(:)
    TYPE(CRTM_RTSolution_type), INTENT(INOUT) :: RTSolution(:, :)
    TYPE(CRTM_Options_type), OPTIONAL, INTENT(IN) :: Options(:)
    INTEGER :: Error_Status
    TYPE(batch_type), TARGET :: b
    INTEGER :: L, M, mm, ic
    Error_Status = pack_batch(Atmosphere, Surface, Geometry, ChannelInfo, RTSolution, Options, b, L, M)
    IF (Error_Status /= SUCCESS) RETURN
    Error_Status = c_crtm_forward(ChannelInfo(1)%ctx, b%atm, b%sfc, b%geo, b%opt, b%rts)
    IF (Error_Status /= SUCCESS) THEN
      CALL report(ChannelInfo(1)%ctx, 'CRTM_Forward')
      RETURN
    END IF
    DO mm = 1, M
      DO ic = 1, L
        CALL unpack_rts(b, (mm - 1) * L + ic, RTSolution(ic, mm))
      END DO
    END DO
    CALL label_rts(ChannelInfo, RTSolution)
  END FUNCTION CRTM_Forward

  FUNCTION CRTM_K_Matrix(Atmosphere, Surface, RTSolution_K, Geometry, ChannelInfo, Atmosphere_K, Surface_K, &
                         RTSolution, Options) RESULT(Error_Status)
    TYPE(CRTM_Atmosphere_type), INTENT(IN) :: Atmosphere(:)
    TYPE(CRTM_Surface_type), INTENT(IN) :: Surface(:)
    TYPE(CRTM_RTSolution_type), INTENT(INOUT) :: RTSolution_K(:, :)
    TYPE(CRTM_Geometry_type), INTENT(IN) :: Geometry(:)
    TYPE(CRTM_ChannelInfo_type), INTENT(IN) :: ChannelInfo(:)
    TYPE(CRTM_Atmosphere_type), INTENT(INOUT) :: Atmosphere_K(:, :)
    TYPE(CRTM_Surface_type), INTENT(INOUT) :: Surface_K(:, :)
    TYPE(CRTM_RTSolution_type), INTENT(INOUT) :: RTSolution(:, :)
    TYPE(CRTM_Options_type), OPTIONAL, INTENT(IN) :: Options(:)
    INTEGER :: Error_Status
    TYPE(batch_type), TARGET :: b
    TYPE(c_k_output) :: k
    REAL(C_DOUBLE), ALLOCATABLE, TARGET :: btk(:), radk(:), emk(:), tk(:), ak(:), stk(:, :), swk(:, :)
    REAL(C_DOUBLE), ALLOCATABLE, TARGET :: cwk(:), crk(:), cfk(:), ack(:), ark(:)
    INTEGER :: L, M, mm, ic, j, kk, NT, na, K0, nc, nae
    Error_Status = pack_batch(Atmosphere, Surface, Geometry, ChannelInfo, RTSolution, Options, b, L, M)
    IF (Error_Status /= SUCCESS) RETURN
    NT = INT(b%off(M + 1))
    na = Atmosphere(1)%n_Absorbers
    ALLOCATE(btk(L * M), radk(L * M), emk(L * M), tk(NT * L), ak(na * NT * L), stk(L * M, 4), swk(L * M, 3))
    DO mm = 1, M
      DO ic = 1, L
        btk((mm - 1) * L + ic) = RTSolution_K(ic, mm)%Brightness_Temperature
        radk((mm - 1) * L + ic) = RTSolution_K(ic, mm)%Radiance
      END DO
    END DO
    k%brightness_temperature_k = C_LOC(btk)
    k%radiance_k = C_LOC(radk)
    k%surface_emissivity_k = C_LOC(emk)
    k%temperature_k = C_LOC(tk)
    k%absorber_k = C_LOC(ak)
    k%land_temperature_k = C_LOC(stk(1, 1))
    k%water_temperature_k = C_LOC(stk(1, 2))
    k%snow_temperature_k = C_LOC(stk(1, 3))
    k%ice_temperature_k = C_LOC(stk(1, 4))
    k%salinity_k = C_LOC(swk(1, 1))
    k%wind_speed_k = C_LOC(swk(1, 2))
    k%wind_direction_k = C_LOC(swk(1, 3))
    nc = Atmosphere(1)%n_Clouds
    ALLOCATE(cwk(MAX(nc, 1) * NT * L), crk(MAX(nc, 1) * NT * L), cfk(NT * L))
    k%cloud_water_content_k = C_LOC(cwk)
    k%cloud_effective_radius_k = C_LOC(crk)
    k%cloud_fraction_k = C_LOC(cfk)
    nae = Atmosphere(1)%n_Aerosols
    ALLOCATE(ack(MAX(nae, 1) * NT * L), ark(MAX(nae, 1) * NT * L))
    k%aerosol_concentration_k = C_LOC(ack)
    k%aerosol_effective_radius_k = C_LOC(ark)
    Error_Status = c_crtm_k_matrix(ChannelInfo(1)%ctx, b%atm, b%sfc, b%geo, b%opt, b%rts, k)
    IF (Error_Status /= SUCCESS) THEN
      CALL report(ChannelInfo(1)%ctx, 'CRTM_K_Matrix')
      RETURN
    END IF
    DO mm = 1, M
      K0 = INT(b%off(mm))
      DO ic = 1, L
        CALL unpack_rts(b, (mm - 1) * L + ic, RTSolution(ic, mm))
        RTSolution_K(ic, mm)%Radiance = radk((mm - 1) * L + ic)
        RTSolution_K(ic, mm)%Surface_Emissivity = RTSolution_K(ic, mm)%Surface_Emissivity + emk((mm - 1) * L + ic)
        DO kk = 1, Atmosphere(mm)%n_Layers
          Atmosphere_K(ic, mm)%Temperature(kk) = Atmosphere_K(ic, mm)%Temperature(kk) + tk((K0 + kk - 1) * L + ic)
          DO j = 1, na
            Atmosphere_K(ic, mm)%Absorber(kk, j) = Atmosphere_K(ic, mm)%Absorber(kk, j) + &
                ak(((j - 1) * NT + K0 + kk - 1) * L + ic)
          END DO
          ! Atmosphere_K%Cloud(j)%Water_Content / %Effective_Radius, %Cloud_Fraction
          IF (nc > 0 .AND. ALLOCATED(Atmosphere_K(ic, mm)%Cloud)) THEN
            DO j = 1, nc
              Atmosphere_K(ic, mm)%Cloud(j)%Water_Content(kk) = Atmosphere_K(ic, mm)%Cloud(j)%Water_Content(kk) + &
                  cwk(((j - 1) * NT + K0 + kk - 1) * L + ic)
              Atmosphere_K(ic, mm)%Cloud(j)%Effective_Radius(kk) = &
                  Atmosphere_K(ic, mm)%Cloud(j)%Effective_Radius(kk) + crk(((j - 1) * NT + K0 + kk - 1) * L + ic)
            END DO
            Atmosphere_K(ic, mm)%Cloud_Fraction(kk) = Atmosphere_K(ic, mm)%Cloud_Fraction(kk) + &
                cfk((K0 + kk - 1) * L + ic)
          END IF
          ! Atmosphere_K%Aerosol(j)%Concentration / %Effective_Radius
          IF (nae > 0 .AND. ALLOCATED(Atmosphere_K(ic, mm)%Aerosol)) THEN
            DO j = 1, nae
              Atmosphere_K(ic, mm)%Aerosol(j)%Concentration(kk) = &
                  Atmosphere_K(ic, mm)%Aerosol(j)%Concentration(kk) + ack(((j - 1) * NT + K0 + kk - 1) * L + ic)
              Atmosphere_K(ic, mm)%Aerosol(j)%Effective_Radius(kk) = &
                  Atmosphere_K(ic, mm)%Aerosol(j)%Effective_Radius(kk) + ark(((j - 1) * NT + K0 + kk - 1) * L + ic)
            END DO
          END IF
        END DO
        Surface_K(ic, mm)%Land_Temperature = Surface_K(ic, mm)%Land_Temperature + stk((mm - 1) * L + ic, 1)
        Surface_K(ic, mm)%Water_Temperature = Surface_K(ic, mm)%Water_Temperature + stk((mm - 1) * L + ic, 2)
        Surface_K(ic, mm)%Snow_Temperature = Surface_K(ic, mm)%Snow_Temperature + stk((mm - 1) * L + ic, 3)
        Surface_K(ic, mm)%Ice_Temperature = Surface_K(ic, mm)%Ice_Temperature + stk((mm - 1) * L + ic, 4)
        Surface_K(ic, mm)%Salinity = Surface_K(ic, mm)%Salinity + swk((mm - 1) * L + ic, 1)
        Surface_K(ic, mm)%Wind_Speed = Surface_K(ic, mm)%Wind_Speed + swk((mm - 1) * L + ic, 2)
        Surface_K(ic, mm)%Wind_Direction = Surface_K(ic, mm)%Wind_Direction + swk((mm - 1) * L + ic, 3)
      END DO
    END DO
    CALL label_rts(ChannelInfo, RTSolution)
  END FUNCTION CRTM_K_Matrix

  ! CRTM_Adjoint (src/CRTM_Adjoint_Module.f90:299-325) through crtm_adjoint:
  ! the channel-summed adjoint of RTSolution_AD%Brightness_Temperature /
  ! %Radiance, summed over channels in the kernels (no per-channel Jacobian
  ! is formed).  Atmosphere_AD and Surface_AD are accumulated into; the
  ! emissivity adjoint is per channel, added to RTSolution_AD%Surface_Emissivity.
  FUNCTION CRTM_Adjoint(Atmosphere, Surface, RTSolution_AD, Geometry, ChannelInfo, Atmosphere_AD, Surface_AD, &
                        RTSolution, Options) RESULT(Error_Status)
    TYPE(CRTM_Atmosphere_type), INTENT(IN) :: Atmosphere(:)
    TYPE(CRTM_Surface_type), INTENT(IN) :: Surface(:)
    TYPE(CRTM_RTSolution_type), INTENT(INOUT) :: RTSolution_AD(:, :)
    TYPE(CRTM_Geometry_type), INTENT(IN) :: Geometry(:)
    TYPE(CRTM_ChannelInfo_type), INTENT(IN) :: ChannelInfo(:)
    TYPE(CRTM_Atmosphere_type), INTENT(INOUT) :: Atmosphere_AD(:)
    TYPE(CRTM_Surface_type), INTENT(INOUT) :: Surface_AD(:)
    TYPE(CRTM_RTSolution_type), INTENT(INOUT) :: RTSolution(:, :)
    TYPE(CRTM_Options_type), OPTIONAL, INTENT(IN) :: Options(:)
    INTEGER :: Error_Status
    TYPE(batch_type), TARGET :: b
    TYPE(c_ad_output) :: ad
    REAL(C_DOUBLE), ALLOCATABLE, TARGET :: btad(:), radad(:), emad(:), tad(:), aad(:), stad(:, :), swad(:, :)
    REAL(C_DOUBLE), ALLOCATABLE, TARGET :: cwad(:), crad(:), cfad(:), acad(:), arad(:)
    INTEGER :: L, M, mm, ic, kk, j, K0, NT, na, nc, nae
    Error_Status = pack_batch(Atmosphere, Surface, Geometry, ChannelInfo, RTSolution, Options, b, L, M)
    IF (Error_Status /= SUCCESS) RETURN
    NT = INT(b%off(M + 1))
    na = Atmosphere(1)%n_Absorbers
    nc = Atmosphere(1)%n_Clouds
    nae = Atmosphere(1)%n_Aerosols
    ALLOCATE(btad(L * M), radad(L * M), emad(L * M), tad(NT), aad(na * NT), stad(M, 4), swad(M, 3))
    swad = 0
    ALLOCATE(cwad(MAX(nc, 1) * NT), crad(MAX(nc, 1) * NT), cfad(NT), acad(MAX(nae, 1) * NT), arad(MAX(nae, 1) * NT))
    tad = 0; aad = 0; stad = 0; emad = 0; cwad = 0; crad = 0; cfad = 0; acad = 0; arad = 0
    DO mm = 1, M
      DO ic = 1, L
        btad((mm - 1) * L + ic) = RTSolution_AD(ic, mm)%Brightness_Temperature
        radad((mm - 1) * L + ic) = RTSolution_AD(ic, mm)%Radiance
      END DO
    END DO
    ad%brightness_temperature_ad = C_LOC(btad)
    ad%radiance_ad = C_LOC(radad)
    ad%surface_emissivity_ad = C_LOC(emad)
    ad%temperature_ad = C_LOC(tad)
    ad%absorber_ad = C_LOC(aad)
    ad%land_temperature_ad = C_LOC(stad(1, 1))
    ad%water_temperature_ad = C_LOC(stad(1, 2))
    ad%snow_temperature_ad = C_LOC(stad(1, 3))
    ad%ice_temperature_ad = C_LOC(stad(1, 4))
    ad%salinity_ad = C_LOC(swad(1, 1))
    ad%wind_speed_ad = C_LOC(swad(1, 2))
    ad%wind_direction_ad = C_LOC(swad(1, 3))
    ad%cloud_water_content_ad = C_LOC(cwad)
    ad%cloud_effective_radius_ad = C_LOC(crad)
    ad%cloud_fraction_ad = C_LOC(cfad)
    ad%aerosol_concentration_ad = C_LOC(acad)
    ad%aerosol_effective_radius_ad = C_LOC(arad)
    Error_Status = c_crtm_adjoint(ChannelInfo(1)%ctx, b%atm, b%sfc, b%geo, b%opt, b%rts, ad)
    IF (Error_Status /= SUCCESS) THEN
      CALL report(ChannelInfo(1)%ctx, 'CRTM_Adjoint')
      RETURN
    END IF
    DO mm = 1, M
      K0 = INT(b%off(mm))
      DO ic = 1, L
        CALL unpack_rts(b, (mm - 1) * L + ic, RTSolution(ic, mm))
        RTSolution_AD(ic, mm)%Radiance = radad((mm - 1) * L + ic)
        RTSolution_AD(ic, mm)%Surface_Emissivity = RTSolution_AD(ic, mm)%Surface_Emissivity + emad((mm - 1) * L + ic)
      END DO
      Surface_AD(mm)%Land_Temperature = Surface_AD(mm)%Land_Temperature + stad(mm, 1)
      Surface_AD(mm)%Water_Temperature = Surface_AD(mm)%Water_Temperature + stad(mm, 2)
      Surface_AD(mm)%Snow_Temperature = Surface_AD(mm)%Snow_Temperature + stad(mm, 3)
      Surface_AD(mm)%Ice_Temperature = Surface_AD(mm)%Ice_Temperature + stad(mm, 4)
      Surface_AD(mm)%Salinity = Surface_AD(mm)%Salinity + swad(mm, 1)
      Surface_AD(mm)%Wind_Speed = Surface_AD(mm)%Wind_Speed + swad(mm, 2)
      Surface_AD(mm)%Wind_Direction = Surface_AD(mm)%Wind_Direction + swad(mm, 3)
      DO kk = 1, Atmosphere(mm)%n_Layers
        Atmosphere_AD(mm)%Temperature(kk) = Atmosphere_AD(mm)%Temperature(kk) + tad(K0 + kk)
        DO j = 1, na
          Atmosphere_AD(mm)%Absorber(kk, j) = Atmosphere_AD(mm)%Absorber(kk, j) + aad((j - 1) * NT + K0 + kk)
        END DO
        IF (nc > 0 .AND. ALLOCATED(Atmosphere_AD(mm)%Cloud)) THEN
          DO j = 1, nc
            Atmosphere_AD(mm)%Cloud(j)%Water_Content(kk) = Atmosphere_AD(mm)%Cloud(j)%Water_Content(kk) + &
                cwad((j - 1) * NT + K0 + kk)
            Atmosphere_AD(mm)%Cloud(j)%Effective_Radius(kk) = Atmosphere_AD(mm)%Cloud(j)%Effective_Radius(kk) + &
                crad((j - 1) * NT + K0 + kk)
          END DO
          Atmosphere_AD(mm)%Cloud_Fraction(kk) = Atmosphere_AD(mm)%Cloud_Fraction(kk) + cfad(K0 + kk)
        END IF
        IF (nae > 0 .AND. ALLOCATED(Atmosphere_AD(mm)%Aerosol)) THEN
          DO j = 1, nae
            Atmosphere_AD(mm)%Aerosol(j)%Concentration(kk) = Atmosphere_AD(mm)%Aerosol(j)%Concentration(kk) + &
                acad((j - 1) * NT + K0 + kk)
            Atmosphere_AD(mm)%Aerosol(j)%Effective_Radius(kk) = Atmosphere_AD(mm)%Aerosol(j)%Effective_Radius(kk) + &
                arad((j - 1) * NT + K0 + kk)
          END DO
        END IF
      END DO
    END DO
    CALL label_rts(ChannelInfo, RTSolution)
  END FUNCTION CRTM_Adjoint

  ! CRTM_Tangent_Linear (src/CRTM_Tangent_Linear_Module.f90:271-296):
  ! RTSolution_TL%Brightness_Temperature = K . (Atmosphere_TL, Surface_TL),
  ! with the K-matrix (BT seeded with 1) run on profile chunks.  Radiance_TL
  ! is not formed (the brightness temperature is this path's observable).
  FUNCTION CRTM_Tangent_Linear(Atmosphere, Surface, Atmosphere_TL, Surface_TL, Geometry, ChannelInfo, RTSolution, &
                               RTSolution_TL, Options) RESULT(Error_Status)
    TYPE(CRTM_Atmosphere_type), INTENT(IN) :: Atmosphere(:)
    TYPE(CRTM_Surface_type), INTENT(IN) :: Surface(:)
    TYPE(CRTM_Atmosphere_type), INTENT(IN) :: Atmosphere_TL(:)
    TYPE(CRTM_Surface_type), INTENT(IN) :: Surface_TL(:)
    TYPE(CRTM_Geometry_type), INTENT(IN) :: Geometry(:)
    TYPE(CRTM_ChannelInfo_type), INTENT(IN) :: ChannelInfo(:)
    TYPE(CRTM_RTSolution_type), INTENT(INOUT) :: RTSolution(:, :)
    TYPE(CRTM_RTSolution_type), INTENT(INOUT) :: RTSolution_TL(:, :)
    TYPE(CRTM_Options_type), OPTIONAL, INTENT(IN) :: Options(:)
    INTEGER :: Error_Status
    TYPE(k_buffers) :: kb
    TYPE(CRTM_RTSolution_type), ALLOCATABLE :: seed(:, :)
    TYPE(CRTM_Options_type), ALLOCATABLE :: od(:)
    INTEGER :: M, L, m0, m1, mm, ic, kk, j, K0, NT, nl
    REAL(fp) :: s
    IF (PRESENT(Options)) THEN
      od = Options
    ELSE   ! the reference defaults
      ALLOCATE(od(SIZE(Atmosphere)))
    END IF
    M = SIZE(Atmosphere)
    ALLOCATE(seed(SIZE(RTSolution, 1), M))
    seed(:, :)%Brightness_Temperature = 1.0_fp
    seed(:, :)%Radiance = 0.0_fp
    m0 = 1
    DO WHILE (m0 <= M)
      m1 = chunk_end(Atmosphere, ChannelInfo, m0)
      Error_Status = run_k_chunk(Atmosphere(m0:m1), Surface(m0:m1), Geometry(m0:m1), ChannelInfo, &
                                 RTSolution(:, m0:m1), od, m0, m1, seed(:, m0:m1), kb, L, NT)
      IF (Error_Status /= SUCCESS) RETURN
      DO mm = m0, m1
        K0 = INT(kb%b%off(mm - m0 + 1))
        nl = Atmosphere(mm)%n_Layers
        DO ic = 1, L
          s = kb%stk((mm - m0) * L + ic, 1) * Surface_TL(mm)%Land_Temperature + &
              kb%stk((mm - m0) * L + ic, 2) * Surface_TL(mm)%Water_Temperature + &
              kb%stk((mm - m0) * L + ic, 3) * Surface_TL(mm)%Snow_Temperature + &
              kb%stk((mm - m0) * L + ic, 4) * Surface_TL(mm)%Ice_Temperature + &
              kb%swk((mm - m0) * L + ic, 1) * Surface_TL(mm)%Salinity + &
              kb%swk((mm - m0) * L + ic, 2) * Surface_TL(mm)%Wind_Speed + &
              kb%swk((mm - m0) * L + ic, 3) * Surface_TL(mm)%Wind_Direction
          DO kk = 1, nl
            s = s + kb%tk((K0 + kk - 1) * L + ic) * Atmosphere_TL(mm)%Temperature(kk)
            DO j = 1, kb%na
              s = s + kb%ak(((j - 1) * NT + K0 + kk - 1) * L + ic) * Atmosphere_TL(mm)%Absorber(kk, j)
            END DO
            IF (kb%nc > 0 .AND. ALLOCATED(Atmosphere_TL(mm)%Cloud)) THEN
              DO j = 1, kb%nc
                s = s + kb%cwk(((j - 1) * NT + K0 + kk - 1) * L + ic) * Atmosphere_TL(mm)%Cloud(j)%Water_Content(kk) &
                      + kb%crk(((j - 1) * NT + K0 + kk - 1) * L + ic) * Atmosphere_TL(mm)%Cloud(j)%Effective_Radius(kk)
              END DO
              IF (ALLOCATED(Atmosphere_TL(mm)%Cloud_Fraction)) &
                s = s + kb%cfk((K0 + kk - 1) * L + ic) * Atmosphere_TL(mm)%Cloud_Fraction(kk)
            END IF
            IF (kb%nae > 0 .AND. ALLOCATED(Atmosphere_TL(mm)%Aerosol)) THEN
              DO j = 1, kb%nae
                s = s + kb%ack(((j - 1) * NT + K0 + kk - 1) * L + ic) * Atmosphere_TL(mm)%Aerosol(j)%Concentration(kk) &
                      + kb%ark(((j - 1) * NT + K0 + kk - 1) * L + ic) * Atmosphere_TL(mm)%Aerosol(j)%Effective_Radius(kk)
              END DO
            END IF
          END DO
          RTSolution_TL(ic, mm)%Brightness_Temperature = s
        END DO
      END DO
      m0 = m1 + 1
    END DO
  END FUNCTION CRTM_Tangent_Linear

  ! last profile of the chunk starting at m0: at most K_CHUNK_BYTES of
  ! per-channel Jacobians (and at least one profile)
  FUNCTION chunk_end(Atm, ChannelInfo, m0) RESULT(m1)
    TYPE(CRTM_Atmosphere_type), INTENT(IN) :: Atm(:)
    TYPE(CRTM_ChannelInfo_type), INTENT(IN) :: ChannelInfo(:)
    INTEGER, INTENT(IN) :: m0
    INTEGER :: m1, L, fields
    INTEGER(C_INT64_T) :: bytes
    INTEGER(C_INT64_T), PARAMETER :: K_CHUNK_BYTES = 2_C_INT64_T * 1024 * 1024 * 1024
    L = SUM(ChannelInfo(:)%n_Channels)
    fields = 1 + Atm(m0)%n_Absorbers + 3 * Atm(m0)%n_Clouds + 2 * Atm(m0)%n_Aerosols
    m1 = m0
    bytes = INT(Atm(m0)%n_Layers, C_INT64_T) * L * fields * 8
    DO WHILE (m1 < SIZE(Atm))
      bytes = bytes + INT(Atm(m1 + 1)%n_Layers, C_INT64_T) * L * fields * 8
      IF (bytes > K_CHUNK_BYTES) EXIT
      m1 = m1 + 1
    END DO
  END FUNCTION chunk_end

  ! K-matrix of profiles m0..m1 (their sections passed in) with RTSolution_K
  ! from Seed, into the flat buffers of kb; forward results into RTS
  FUNCTION run_k_chunk(Atm, Sfc, Geo, ChannelInfo, RTS, Opt, m0, m1, Seed, kb, L, NT) RESULT(Error_Status)
    TYPE(CRTM_Atmosphere_type), INTENT(IN) :: Atm(:)
    TYPE(CRTM_Surface_type), INTENT(IN) :: Sfc(:)
    TYPE(CRTM_Geometry_type), INTENT(IN) :: Geo(:)
    TYPE(CRTM_ChannelInfo_type), INTENT(IN) :: ChannelInfo(:)
    TYPE(CRTM_RTSolution_type), INTENT(INOUT) :: RTS(:, :)
    TYPE(CRTM_Options_type), INTENT(IN) :: Opt(:)
    INTEGER, INTENT(IN) :: m0, m1
    TYPE(CRTM_RTSolution_type), INTENT(IN) :: Seed(:, :)
    TYPE(k_buffers), TARGET, INTENT(INOUT) :: kb
    INTEGER, INTENT(OUT) :: L, NT
    INTEGER :: Error_Status, M, mm, ic
    TYPE(c_k_output) :: k
    IF (SIZE(Opt) >= m1) THEN
      Error_Status = pack_batch(Atm, Sfc, Geo, ChannelInfo, RTS, Opt(m0:m1), kb%b, L, M)
    ELSE
      Error_Status = pack_batch(Atm, Sfc, Geo, ChannelInfo, RTS, Opt(1:1), kb%b, L, M)
    END IF
    IF (Error_Status /= SUCCESS) RETURN
    NT = INT(kb%b%off(M + 1))
    kb%na = Atm(1)%n_Absorbers
    kb%nc = Atm(1)%n_Clouds
    kb%nae = Atm(1)%n_Aerosols
    IF (ALLOCATED(kb%btk)) DEALLOCATE(kb%btk, kb%radk, kb%emk, kb%tk, kb%ak, kb%stk, kb%swk, kb%cwk, kb%crk, &
                                      kb%cfk, kb%ack, kb%ark)
    ALLOCATE(kb%btk(L * M), kb%radk(L * M), kb%emk(L * M), kb%tk(NT * L), kb%ak(kb%na * NT * L), kb%stk(L * M, 4), &
             kb%swk(L * M, 3))
    ALLOCATE(kb%cwk(MAX(kb%nc, 1) * NT * L), kb%crk(MAX(kb%nc, 1) * NT * L), kb%cfk(NT * L))
    ALLOCATE(kb%ack(MAX(kb%nae, 1) * NT * L), kb%ark(MAX(kb%nae, 1) * NT * L))
    DO mm = 1, M
      DO ic = 1, L
        kb%btk((mm - 1) * L + ic) = Seed(ic, mm)%Brightness_Temperature
        kb%radk((mm - 1) * L + ic) = Seed(ic, mm)%Radiance
      END DO
    END DO
    k%brightness_temperature_k = C_LOC(kb%btk)
    k%radiance_k = C_LOC(kb%radk)
    k%surface_emissivity_k = C_LOC(kb%emk)
    k%temperature_k = C_LOC(kb%tk)
    k%absorber_k = C_LOC(kb%ak)
    k%land_temperature_k = C_LOC(kb%stk(1, 1))
    k%water_temperature_k = C_LOC(kb%stk(1, 2))
    k%snow_temperature_k = C_LOC(kb%stk(1, 3))
    k%ice_temperature_k = C_LOC(kb%stk(1, 4))
    k%salinity_k = C_LOC(kb%swk(1, 1))
    k%wind_speed_k = C_LOC(kb%swk(1, 2))
    k%wind_direction_k = C_LOC(kb%swk(1, 3))
    k%cloud_water_content_k = C_LOC(kb%cwk)
    k%cloud_effective_radius_k = C_LOC(kb%crk)
    k%cloud_fraction_k = C_LOC(kb%cfk)
    k%aerosol_concentration_k = C_LOC(kb%ack)
    k%aerosol_effective_radius_k = C_LOC(kb%ark)
    Error_Status = c_crtm_k_matrix(ChannelInfo(1)%ctx, kb%b%atm, kb%b%sfc, kb%b%geo, kb%b%opt, kb%b%rts, k)
    IF (Error_Status /= SUCCESS) THEN
      CALL report(ChannelInfo(1)%ctx, 'CRTM_Tangent_Linear')
      RETURN
    END IF
    DO mm = 1, M
      DO ic = 1, L
        CALL unpack_rts(kb%b, (mm - 1) * L + ic, RTS(ic, mm))
      END DO
    END DO
  END FUNCTION run_k_chunk

  ! ---- companions of the calls (names and arguments of the reference's) ----

  ! CRTM_Version (src/CRTM_Version.inc): identifies this implementation
  SUBROUTINE CRTM_Version(Id)
    CHARACTER(*), INTENT(OUT) :: Id
    Id = 'CRTM v3 interface, MI355X forward / K-matrix path (libcrtm_mi355x)'
  END SUBROUTINE CRTM_Version

  ! Display_Message (src/Utility/Message_Handler.f90:110-160): one line
  ! "routine(STATE) : message" on standard output or appended to Message_Log
  SUBROUTINE Display_Message(Routine_Name, Message, Error_State, Message_Log)
    CHARACTER(*), INTENT(IN) :: Routine_Name, Message
    INTEGER, INTENT(IN) :: Error_State
    CHARACTER(*), OPTIONAL, INTENT(IN) :: Message_Log
    CHARACTER(LEN=11), PARAMETER :: STATE(0:4) = (/ 'SUCCESS    ', 'INFORMATION', 'WARNING    ', 'FAILURE    ', &
                                                   'UNDEFINED  ' /)
    CHARACTER(*), PARAMETER :: FMT = '( 1x, a, "(", a, ") : ", a )'
    INTEGER :: st, u, ios
    st = Error_State
    IF (st < 0 .OR. st > 4) st = UNDEFINED
    IF (PRESENT(Message_Log)) THEN
      OPEN(NEWUNIT=u, FILE=TRIM(Message_Log), POSITION='APPEND', ACTION='WRITE', IOSTAT=ios)
      IF (ios == 0) THEN
        WRITE(u, FMT) TRIM(Routine_Name), TRIM(STATE(st)), TRIM(Message)
        CLOSE(u)
        RETURN
      END IF
    END IF
    WRITE(*, FMT) TRIM(Routine_Name), TRIM(STATE(st)), TRIM(Message)
  END SUBROUTINE Display_Message

  ! Program_Message (src/Utility/Message_Handler.f90:48-100): program banner
  SUBROUTINE Program_Message(Name, Description, Revision)
    CHARACTER(*), INTENT(IN) :: Name, Description, Revision
    CHARACTER(*), PARAMETER :: BAR = '**********************************************************'
    WRITE(*, '(/5x, a)') BAR
    WRITE(*, '(5x, a)') TRIM(ADJUSTL(Name))
    WRITE(*, '(/5x, a)') TRIM(Description)
    WRITE(*, '(/5x, a)') TRIM(Revision)
    WRITE(*, '(5x, a/)') BAR
  END SUBROUTINE Program_Message

  ! CRTM_ChannelInfo_n_Channels (src/ChannelInfo/CRTM_ChannelInfo_Define.f90)
  ELEMENTAL FUNCTION CRTM_ChannelInfo_n_Channels(ChannelInfo) RESULT(n)
    TYPE(CRTM_ChannelInfo_type), INTENT(IN) :: ChannelInfo
    INTEGER :: n
    n = ChannelInfo%n_Channels
  END FUNCTION CRTM_ChannelInfo_n_Channels

  ! CRTM_Atmosphere_Associated / _Destroy (src/Atmosphere/CRTM_Atmosphere_Define.f90)
  ELEMENTAL FUNCTION CRTM_Atmosphere_Associated(Atm) RESULT(yes)
    TYPE(CRTM_Atmosphere_type), INTENT(IN) :: Atm
    LOGICAL :: yes
    yes = ALLOCATED(Atm%Level_Pressure) .AND. ALLOCATED(Atm%Temperature) .AND. ALLOCATED(Atm%Absorber)
  END FUNCTION CRTM_Atmosphere_Associated

  ELEMENTAL SUBROUTINE CRTM_Atmosphere_Destroy(Atm)
    TYPE(CRTM_Atmosphere_type), INTENT(OUT) :: Atm   ! INTENT(OUT) deallocates every component
    Atm%n_Layers = 0
  END SUBROUTINE CRTM_Atmosphere_Destroy

  ! CRTM_Geometry_SetValue (src/GeometryInfo/CRTM_Geometry_Define.f90:365-411)
  ELEMENTAL SUBROUTINE CRTM_Geometry_SetValue(geo, iFOV, Longitude, Latitude, Surface_Altitude, Sensor_Scan_Angle, &
                                              Sensor_Zenith_Angle, Sensor_Azimuth_Angle, Source_Zenith_Angle, &
                                              Source_Azimuth_Angle, Flux_Zenith_Angle, Year, Month, Day)
    TYPE(CRTM_Geometry_type), INTENT(INOUT) :: geo
    INTEGER, OPTIONAL, INTENT(IN) :: iFOV, Year, Month, Day
    REAL(fp), OPTIONAL, INTENT(IN) :: Longitude, Latitude, Surface_Altitude, Sensor_Scan_Angle, &
                                      Sensor_Zenith_Angle, Sensor_Azimuth_Angle, Source_Zenith_Angle, &
                                      Source_Azimuth_Angle, Flux_Zenith_Angle
    IF (PRESENT(iFOV)) geo%iFOV = iFOV
    IF (PRESENT(Longitude)) geo%Longitude = Longitude
    IF (PRESENT(Latitude)) geo%Latitude = Latitude
    IF (PRESENT(Surface_Altitude)) geo%Surface_Altitude = Surface_Altitude
    IF (PRESENT(Sensor_Scan_Angle)) geo%Sensor_Scan_Angle = Sensor_Scan_Angle
    IF (PRESENT(Sensor_Zenith_Angle)) geo%Sensor_Zenith_Angle = Sensor_Zenith_Angle
    IF (PRESENT(Sensor_Azimuth_Angle)) geo%Sensor_Azimuth_Angle = Sensor_Azimuth_Angle
    IF (PRESENT(Source_Zenith_Angle)) geo%Source_Zenith_Angle = Source_Zenith_Angle
    IF (PRESENT(Source_Azimuth_Angle)) geo%Source_Azimuth_Angle = Source_Azimuth_Angle
    IF (PRESENT(Flux_Zenith_Angle)) geo%Flux_Zenith_Angle = Flux_Zenith_Angle
    IF (PRESENT(Year)) geo%Year = Year
    IF (PRESENT(Month)) geo%Month = Month
    IF (PRESENT(Day)) geo%Day = Day
  END SUBROUTINE CRTM_Geometry_SetValue

  ! CRTM_RTSolution_Inspect / CRTM_Surface_Inspect / CRTM_Atmosphere_Inspect:
  ! the components of this path, one per line
  SUBROUTINE CRTM_RTSolution_Inspect(RTS)
    TYPE(CRTM_RTSolution_type), INTENT(IN) :: RTS
    WRITE(*, '(1x, "RTSolution OBJECT")')
    WRITE(*, '(3x, "Sensor Id                     : ", a)') TRIM(RTS%Sensor_Id)
    WRITE(*, '(3x, "Channel                       : ", i0)') RTS%Sensor_Channel
    WRITE(*, '(3x, "Surface Emissivity            : ", es22.15)') RTS%Surface_Emissivity
    WRITE(*, '(3x, "Up Radiance                   : ", es22.15)') RTS%Up_Radiance
    WRITE(*, '(3x, "Down Radiance                 : ", es22.15)') RTS%Down_Radiance
    WRITE(*, '(3x, "Surface Planck Radiance       : ", es22.15)') RTS%Surface_Planck_Radiance
    WRITE(*, '(3x, "Total Cloud Cover             : ", es22.15)') RTS%Total_Cloud_Cover
    WRITE(*, '(3x, "Radiance (clear)              : ", es22.15)') RTS%R_Clear
    WRITE(*, '(3x, "Brightness Temperature (clear): ", es22.15)') RTS%Tb_Clear
    WRITE(*, '(3x, "Radiance                      : ", es22.15)') RTS%Radiance
    WRITE(*, '(3x, "Brightness Temperature        : ", es22.15)') RTS%Brightness_Temperature
  END SUBROUTINE CRTM_RTSolution_Inspect

  SUBROUTINE CRTM_Surface_Inspect(Sfc)
    TYPE(CRTM_Surface_type), INTENT(IN) :: Sfc
    WRITE(*, '(1x, "Surface OBJECT")')
    WRITE(*, '(3x, "Land, water, snow, ice coverage   : ", 4es13.6)') Sfc%Land_Coverage, Sfc%Water_Coverage, &
        Sfc%Snow_Coverage, Sfc%Ice_Coverage
    WRITE(*, '(3x, "Land, water, snow, ice temperature: ", 4es13.6)') Sfc%Land_Temperature, Sfc%Water_Temperature, &
        Sfc%Snow_Temperature, Sfc%Ice_Temperature
  END SUBROUTINE CRTM_Surface_Inspect

  SUBROUTINE CRTM_Atmosphere_Inspect(Atm)
    TYPE(CRTM_Atmosphere_type), INTENT(IN) :: Atm
    INTEGER :: j
    WRITE(*, '(1x, "ATMOSPHERE OBJECT")')
    WRITE(*, '(3x, "Dimensions (layers, absorbers, clouds, aerosols): ", 4(1x, i0))') Atm%n_Layers, &
        Atm%n_Absorbers, Atm%n_Clouds, Atm%n_Aerosols
    IF (.NOT. ALLOCATED(Atm%Temperature)) RETURN
    WRITE(*, '(3x, "Temperature:")')
    WRITE(*, '(5(1x, es13.6))') Atm%Temperature
    DO j = 1, Atm%n_Absorbers
      WRITE(*, '(3x, "Absorber #", i0, ":")') j
      WRITE(*, '(5(1x, es13.6))') Atm%Absorber(:, j)
    END DO
  END SUBROUTINE CRTM_Atmosphere_Inspect

  ! Sensor_Id / Sensor_Channel of RTSolution(l, :) from the channel order of
  ! the CRTM_Init call
  SUBROUTINE label_rts(ChannelInfo, RTS)
    TYPE(CRTM_ChannelInfo_type), INTENT(IN) :: ChannelInfo(:)
    TYPE(CRTM_RTSolution_type), INTENT(INOUT) :: RTS(:, :)
    INTEGER :: n, l, ln
    ln = 0
    DO n = 1, SIZE(ChannelInfo)
      DO l = 1, ChannelInfo(n)%n_Channels
        ln = ln + 1
        IF (ln > SIZE(RTS, 1)) RETURN
        RTS(ln, :)%Sensor_Id = ChannelInfo(n)%Sensor_ID
        RTS(ln, :)%Sensor_Channel = ChannelInfo(n)%Sensor_Channel(l)
      END DO
    END DO
  END SUBROUTINE label_rts

  ! ---- helpers ----
  FUNCTION pack_batch(Atm, Sfc, Geo, ChannelInfo, RTS, Opt_in, b, L, M) RESULT(Error_Status)
    TYPE(CRTM_Atmosphere_type), INTENT(IN) :: Atm(:)
    TYPE(CRTM_Surface_type), INTENT(IN) :: Sfc(:)
    TYPE(CRTM_Geometry_type), INTENT(IN) :: Geo(:)
    TYPE(CRTM_ChannelInfo_type), INTENT(IN) :: ChannelInfo(:)
    TYPE(CRTM_RTSolution_type), INTENT(IN) :: RTS(:, :)
    TYPE(CRTM_Options_type), OPTIONAL, INTENT(IN) :: Opt_in(:)
    TYPE(batch_type), TARGET, INTENT(INOUT) :: b
    INTEGER, INTENT(OUT) :: L, M
    INTEGER :: Error_Status, mm, NT, na, K0, K, j, nc, nae
    TYPE(CRTM_Options_type), ALLOCATABLE :: Opt(:)
    Error_Status = FAILURE
    M = SIZE(Atm)
    L = SUM(ChannelInfo(:)%n_Channels)
    IF (SIZE(Sfc) /= M .OR. SIZE(Geo) /= M .OR. SIZE(RTS, 2) < M .OR. SIZE(RTS, 1) < L) RETURN
    ! no Options: the reference defaults (computed surface optics)
    IF (PRESENT(Opt_in)) THEN
      IF (SIZE(Opt_in) /= M) RETURN
      Opt = Opt_in
    ELSE
      ALLOCATE(Opt(M))
    END IF
    na = Atm(1)%n_Absorbers
    ALLOCATE(b%nl(M), b%off(M + 1), b%clim(M))
    b%off(1) = 0
    DO mm = 1, M
      IF (Atm(mm)%n_Absorbers /= na) RETURN
      b%nl(mm) = Atm(mm)%n_Layers
      b%clim(mm) = Atm(mm)%Climatology
      b%off(mm + 1) = b%off(mm) + Atm(mm)%n_Layers
    END DO
    NT = INT(b%off(M + 1))
    ALLOCATE(b%ids(na), b%units(na), b%lp(NT + M), b%p(NT), b%t(NT), b%a(na * NT), b%cov(M, 4), b%st(M, 4), &
             b%zen(M), b%scan(M), b%src(M), b%alt(M), b%skip(M), b%emis(L * M), b%rad(L * M), b%bt(L * M), b%up(L * M), &
             b%dn(L * M), b%sp(L * M), b%se(L * M), b%status(M), b%rcl(L * M), b%tbc(L * M), b%tcc(L * M), &
             b%sr(L * M), b%ws(M), b%wd(M), b%sal(M), b%azi(M), b%fov(M))
    b%emis = 0
    nc = Atm(1)%n_Clouds
    ALLOCATE(b%ctype(MAX(nc, 1) * M), b%cwc(MAX(nc, 1) * NT), b%creff(MAX(nc, 1) * NT), b%cfrac(NT))
    b%ctype = 0; b%cwc = 0; b%creff = 0; b%cfrac = 0
    nae = Atm(1)%n_Aerosols
    ALLOCATE(b%atype(MAX(nae, 1) * M), b%aconc(MAX(nae, 1) * NT), b%areff(MAX(nae, 1) * NT))
    b%atype = 0; b%aconc = 0; b%areff = 0
    b%ids = Atm(1)%Absorber_ID
    b%units = Atm(1)%Absorber_Units
    DO mm = 1, M
      K0 = INT(b%off(mm))
      K = Atm(mm)%n_Layers
      b%lp(K0 + mm:K0 + mm + K) = Atm(mm)%Level_Pressure(0:K)
      b%p(K0 + 1:K0 + K) = Atm(mm)%Pressure(1:K)
      b%t(K0 + 1:K0 + K) = Atm(mm)%Temperature(1:K)
      DO j = 1, na
        b%a((j - 1) * NT + K0 + 1:(j - 1) * NT + K0 + K) = Atm(mm)%Absorber(1:K, j)
      END DO
      b%cov(mm, :) = (/ Sfc(mm)%Land_Coverage, Sfc(mm)%Water_Coverage, Sfc(mm)%Snow_Coverage, Sfc(mm)%Ice_Coverage /)
      b%st(mm, :) = (/ Sfc(mm)%Land_Temperature, Sfc(mm)%Water_Temperature, Sfc(mm)%Snow_Temperature, &
                       Sfc(mm)%Ice_Temperature /)
      b%zen(mm) = Geo(mm)%Sensor_Zenith_Angle
      b%scan(mm) = Geo(mm)%Sensor_Scan_Angle
      b%src(mm) = Geo(mm)%Source_Zenith_Angle
      b%alt(mm) = Geo(mm)%Surface_Altitude
      b%ws(mm) = Sfc(mm)%Wind_Speed
      b%wd(mm) = Sfc(mm)%Wind_Direction
      b%sal(mm) = Sfc(mm)%Salinity
      b%azi(mm) = Geo(mm)%Sensor_Azimuth_Angle
      b%fov(mm) = Geo(mm)%iFOV
      b%skip(mm) = MERGE(1, 0, Opt(mm)%Skip_Profile)
      IF (Opt(mm)%Skip_Profile) CYCLE                ! no inputs read (src/CRTM_Forward_Module.f90:536)
      IF (Opt(mm)%RT_Algorithm_Id /= RT_ADA) THEN
        CALL Display_Message('CRTM', 'only the ADA solver (RT_ADA) is implemented', FAILURE)
        RETURN
      END IF
      IF (Opt(mm)%Include_Scattering .NEQV. Opt(1)%Include_Scattering) THEN
        CALL Display_Message('CRTM', 'Include_Scattering must be the same for all profiles of a call', FAILURE)
        RETURN
      END IF
      ! Use_Emissivity is one switch per call here: user emissivities for every
      ! profile, or the computed (water) surface optics for every profile
      IF (Opt(mm)%Use_Emissivity .NEQV. Opt(1)%Use_Emissivity) THEN
        CALL Display_Message('CRTM', 'Use_Emissivity must be the same for all profiles of a call', FAILURE)
        RETURN
      END IF
      IF ((Opt(mm)%Use_Antenna_Correction .NEQV. Opt(1)%Use_Antenna_Correction) .OR. &
          (Opt(mm)%Apply_NLTE_Correction .NEQV. Opt(1)%Apply_NLTE_Correction)) THEN
        CALL Display_Message('CRTM', 'Use_Antenna_Correction and Apply_NLTE_Correction must be the same for all '// &
                             'profiles of a call', FAILURE)
        RETURN
      END IF
      IF (Opt(mm)%Use_Emissivity) THEN
        IF (SIZE(Opt(mm)%Emissivity) < L) RETURN
        b%emis((mm - 1) * L + 1:mm * L) = Opt(mm)%Emissivity(1:L)
      END IF
      IF (Atm(mm)%n_Clouds /= nc) RETURN
      IF (nc > 0) THEN
        b%cfrac(K0 + 1:K0 + K) = Atm(mm)%Cloud_Fraction(1:K)
        DO j = 1, nc
          b%ctype((mm - 1) * nc + j) = Atm(mm)%Cloud(j)%Type
          b%cwc((j - 1) * NT + K0 + 1:(j - 1) * NT + K0 + K) = Atm(mm)%Cloud(j)%Water_Content(1:K)
          b%creff((j - 1) * NT + K0 + 1:(j - 1) * NT + K0 + K) = Atm(mm)%Cloud(j)%Effective_Radius(1:K)
        END DO
      END IF
      IF (Atm(mm)%n_Aerosols /= nae) RETURN
      DO j = 1, nae
        b%atype((mm - 1) * nae + j) = Atm(mm)%Aerosol(j)%Type
        b%aconc((j - 1) * NT + K0 + 1:(j - 1) * NT + K0 + K) = Atm(mm)%Aerosol(j)%Concentration(1:K)
        b%areff((j - 1) * NT + K0 + 1:(j - 1) * NT + K0 + K) = Atm(mm)%Aerosol(j)%Effective_Radius(1:K)
      END DO
    END DO
    b%atm%n_profiles = M
    b%atm%max_layers = MAXVAL(b%nl)
    b%atm%total_layers = NT
    b%atm%n_layers = C_LOC(b%nl)
    b%atm%layer_offset = C_LOC(b%off)
    b%atm%n_absorbers = na
    b%atm%absorber_id = C_LOC(b%ids)
    b%atm%absorber_units = C_LOC(b%units)
    b%atm%level_pressure = C_LOC(b%lp)
    b%atm%pressure = C_LOC(b%p)
    b%atm%temperature = C_LOC(b%t)
    b%atm%absorber = C_LOC(b%a)
    b%atm%climatology = C_LOC(b%clim)
    b%atm%n_clouds = nc
    b%atm%cloud_type = C_LOC(b%ctype)
    b%atm%cloud_water_content = C_LOC(b%cwc)
    b%atm%cloud_effective_radius = C_LOC(b%creff)
    b%atm%cloud_fraction = C_LOC(b%cfrac)
    b%atm%n_aerosols = nae
    b%atm%aerosol_type = C_LOC(b%atype)
    b%atm%aerosol_concentration = C_LOC(b%aconc)
    b%atm%aerosol_effective_radius = C_LOC(b%areff)
    b%sfc%land_coverage = C_LOC(b%cov(1, 1))
    b%sfc%water_coverage = C_LOC(b%cov(1, 2))
    b%sfc%snow_coverage = C_LOC(b%cov(1, 3))
    b%sfc%ice_coverage = C_LOC(b%cov(1, 4))
    b%sfc%land_temperature = C_LOC(b%st(1, 1))
    b%sfc%water_temperature = C_LOC(b%st(1, 2))
    b%sfc%snow_temperature = C_LOC(b%st(1, 3))
    b%sfc%ice_temperature = C_LOC(b%st(1, 4))
    b%geo%sensor_zenith_angle = C_LOC(b%zen)
    b%geo%sensor_scan_angle = C_LOC(b%scan)
    b%geo%source_zenith_angle = C_LOC(b%src)
    b%geo%surface_altitude = C_LOC(b%alt)
    b%geo%sensor_azimuth_angle = C_LOC(b%azi)
    b%geo%ifov = C_LOC(b%fov)
    b%sfc%wind_speed = C_LOC(b%ws)
    b%sfc%wind_direction = C_LOC(b%wd)
    b%sfc%salinity = C_LOC(b%sal)
    b%opt%use_emissivity = MERGE(1, 0, Opt(1)%Use_Emissivity)
    b%opt%emissivity = C_LOC(b%emis)
    b%opt%precision = Opt(1)%Precision
    b%opt%memory = 0
    b%opt%jacobian_f32 = 0
    b%opt%check_input = MERGE(1, 0, Opt(1)%Check_Input)
    b%opt%n_streams = MERGE(Opt(1)%n_Streams, 0, Opt(1)%Use_N_Streams)
    b%opt%overlap_id = Opt(1)%Overlap_Id
    b%opt%n_streams_profile = C_NULL_PTR
    b%opt%no_scattering = MERGE(0, 1, Opt(1)%Include_Scattering)
    b%opt%skip_profile = C_NULL_PTR
    IF (ANY(b%skip /= 0)) b%opt%skip_profile = C_LOC(b%skip)
    b%opt%use_antenna_correction = MERGE(1, 0, Opt(1)%Use_Antenna_Correction)
    b%opt%no_nlte_correction = MERGE(0, 1, Opt(1)%Apply_NLTE_Correction)
    IF (SIZE(Opt) == M .AND. M > 1) THEN   ! Options(m)%n_Streams per profile
      ALLOCATE(b%nsp(M))
      DO mm = 1, M
        b%nsp(mm) = MERGE(Opt(mm)%n_Streams, 0, Opt(mm)%Use_N_Streams)
      END DO
      b%opt%n_streams_profile = C_LOC(b%nsp)
    END IF
    b%rts%radiance = C_LOC(b%rad)
    b%rts%brightness_temperature = C_LOC(b%bt)
    b%rts%up_radiance = C_LOC(b%up)
    b%rts%down_radiance = C_LOC(b%dn)
    b%rts%surface_planck_radiance = C_LOC(b%sp)
    b%rts%surface_emissivity = C_LOC(b%se)
    b%rts%status = C_LOC(b%status)
    b%rts%r_clear = C_LOC(b%rcl)
    b%rts%tb_clear = C_LOC(b%tbc)
    b%rts%total_cloud_cover = C_LOC(b%tcc)
    b%rts%surface_reflectivity = C_LOC(b%sr)
    Error_Status = SUCCESS
  END FUNCTION pack_batch

  SUBROUTINE unpack_rts(b, i, r)
    TYPE(batch_type), INTENT(IN) :: b
    INTEGER, INTENT(IN) :: i
    TYPE(CRTM_RTSolution_type), INTENT(INOUT) :: r
    r%Radiance = b%rad(i)
    r%Brightness_Temperature = b%bt(i)
    r%Up_Radiance = b%up(i)
    r%Down_Radiance = b%dn(i)
    r%Surface_Planck_Radiance = b%sp(i)
    r%Surface_Emissivity = b%se(i)
    r%Surface_Reflectivity = b%sr(i)
    r%R_Clear = b%rcl(i)
    r%Tb_Clear = b%tbc(i)
    r%Total_Cloud_Cover = b%tcc(i)
  END SUBROUTINE unpack_rts

  SUBROUTINE report(ctx, who)
    TYPE(C_PTR), INTENT(IN) :: ctx
    CHARACTER(*), INTENT(IN) :: who
    CHARACTER(KIND=C_CHAR), POINTER :: msg(:)
    INTEGER :: n
    CALL C_F_POINTER(c_crtm_last_error(ctx), msg, (/ 512 /))
    n = 0
    DO WHILE (n < 512)
      IF (msg(n + 1) == C_NULL_CHAR) EXIT
      n = n + 1
    END DO
    WRITE(*, '(A, ": ", 512A)') who, msg(1:n)
  END SUBROUTINE report

  FUNCTION c_to_f(s) RESULT(f)
    CHARACTER(KIND=C_CHAR, LEN=*), INTENT(IN) :: s
    CHARACTER(LEN=20) :: f
    INTEGER :: i
    f = ''
    DO i = 1, MIN(20, LEN(s))
      IF (s(i:i) == C_NULL_CHAR) EXIT
      f(i:i) = s(i:i)
    END DO
  END FUNCTION c_to_f

END MODULE CRTM_MI355X_Module
