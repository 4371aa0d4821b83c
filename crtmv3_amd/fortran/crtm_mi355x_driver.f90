! crtm_mi355x_driver.f90 -- Fortran caller written the way a CRTM application
! is: USE CRTM_Module, then CRTM_Init, CRTM_Forward,
! CRTM_K_Matrix with zeroed Atmosphere_K / Surface_K, CRTM_Adjoint with
! zeroed Atmosphere_AD / Surface_AD, CRTM_Tangent_Linear, CRTM_Destroy.
! Reads a batch from a text file and writes BT, T_K, absorber_K, water
! temperature K and emissivity K as text, and the adjoint (seeded with
! BT_AD(l) = 0.5 + l/L) and tangent linear (T_TL = 1 K, absorber_TL = 1 %,
! water temperature_TL = 0.5 K) to <output.txt>.ad, and a forward call with
! Options%Skip_Profile on the last profile to <output.txt>.skip, and a
! K-matrix with the surface optics computed (Options%Use_Emissivity =
! .FALSE.; wind speed 4 + m m/s, wind direction 30 m degrees, salinity 34,
! sensor azimuth 100 degrees) to <output.txt>.sfc (used by the Fortran tests).
!
!   crtm_mi355x_driver <input.txt> <output.txt>
PROGRAM crtm_mi355x_driver
  USE CRTM_Module
  IMPLICIT NONE
  CHARACTER(LEN=1024) :: fin, fout, cdir
  CHARACTER(LEN=20) :: sid(1)
  TYPE(CRTM_ChannelInfo_type) :: chinfo(1)
  TYPE(CRTM_Atmosphere_type), ALLOCATABLE :: atm(:), atm_k(:, :)
  TYPE(CRTM_Surface_type), ALLOCATABLE :: sfc(:), sfc_k(:, :)
  TYPE(CRTM_Geometry_type), ALLOCATABLE :: geo(:)
  TYPE(CRTM_RTSolution_type), ALLOCATABLE :: rts(:, :), rts_k(:, :), rts_f(:, :)
  TYPE(CRTM_Options_type), ALLOCATABLE :: opt(:)
  TYPE(CRTM_Atmosphere_type), ALLOCATABLE :: atm_ad(:, :), atm_tl(:, :)
  TYPE(CRTM_Surface_type), ALLOCATABLE :: sfc_ad(:, :), sfc_tl(:)
  TYPE(CRTM_RTSolution_type), ALLOCATABLE :: rts_ad(:, :), rts_tl(:, :)
  INTEGER :: M, na, prec, m1, K, L, j, l1, k1, err, u
  REAL(fp) :: zen, ts

  CALL GET_COMMAND_ARGUMENT(1, fin)
  CALL GET_COMMAND_ARGUMENT(2, fout)
  OPEN(NEWUNIT=u, FILE=TRIM(fin), STATUS='OLD', ACTION='READ')
  READ(u, '(A)') sid(1)
  READ(u, '(A)') cdir
  READ(u, *) M, na, prec
  err = CRTM_Init(sid, chinfo, File_Path=TRIM(cdir))
  IF (err /= SUCCESS) STOP 'CRTM_Init failed'
  L = chinfo(1)%n_Channels
  ALLOCATE(atm(M), sfc(M), geo(M), opt(M), rts(L, M), rts_k(L, M), rts_f(L, M), atm_k(L, M), sfc_k(L, M))
  DO m1 = 1, M
    READ(u, *) K, zen, ts
    CALL CRTM_Atmosphere_Create(atm(m1:m1), K, na, 0, 0)
    atm(m1)%Absorber_ID = (/ (MERGE(H2O_ID, O3_ID, j == 1), j = 1, na) /)
    atm(m1)%Absorber_Units = (/ (MERGE(MASS_MIXING_RATIO_UNITS, VOLUME_MIXING_RATIO_UNITS, j == 1), j = 1, na) /)
    READ(u, *) atm(m1)%Level_Pressure(0:K)
    READ(u, *) atm(m1)%Pressure(1:K)
    READ(u, *) atm(m1)%Temperature(1:K)
    DO j = 1, na
      READ(u, *) atm(m1)%Absorber(1:K, j)
    END DO
    sfc(m1)%Water_Coverage = 1.0_fp
    sfc(m1)%Water_Temperature = ts
    geo(m1)%Sensor_Zenith_Angle = zen
    opt(m1)%Use_Emissivity = .TRUE.
    opt(m1)%Precision = prec
    ALLOCATE(opt(m1)%Emissivity(L))
    READ(u, *) opt(m1)%Emissivity(1:L)
    DO l1 = 1, L
      CALL CRTM_Atmosphere_Create(atm_k(l1:l1, m1), K, na, 0, 0)
    END DO
  END DO
  CLOSE(u)

  err = CRTM_Forward(atm, sfc, geo, chinfo, rts_f, opt)
  IF (err /= SUCCESS) STOP 'CRTM_Forward failed'
  CALL CRTM_Atmosphere_Zero(atm_k)
  CALL CRTM_Surface_Zero(sfc_k)
  rts_k(:, :)%Brightness_Temperature = 1.0_fp
  rts_k(:, :)%Radiance = 0.0_fp
  err = CRTM_K_Matrix(atm, sfc, rts_k, geo, chinfo, atm_k, sfc_k, rts, opt)
  IF (err /= SUCCESS) STOP 'CRTM_K_Matrix failed'

  ! adjoint and tangent linear of the same batch
  ALLOCATE(atm_ad(1, M), atm_tl(1, M), sfc_ad(1, M), sfc_tl(M), rts_ad(L, M), rts_tl(L, M))
  DO m1 = 1, M
    K = atm(m1)%n_Layers
    CALL CRTM_Atmosphere_Create(atm_ad(1:1, m1), K, na, 0, 0)
    CALL CRTM_Atmosphere_Create(atm_tl(1:1, m1), K, na, 0, 0)
    atm_tl(1, m1)%Temperature = 1.0_fp
    atm_tl(1, m1)%Absorber = 0.01_fp * atm(m1)%Absorber
    sfc_tl(m1)%Water_Temperature = 0.5_fp
    DO l1 = 1, L
      rts_ad(l1, m1)%Brightness_Temperature = 0.5_fp + REAL(l1, fp) / REAL(L, fp)
      rts_ad(l1, m1)%Radiance = 0.0_fp
    END DO
  END DO
  CALL CRTM_Atmosphere_Zero(atm_ad)
  CALL CRTM_Surface_Zero(sfc_ad)
  err = CRTM_Adjoint(atm, sfc, rts_ad, geo, chinfo, atm_ad(1, :), sfc_ad(1, :), rts, opt)
  IF (err /= SUCCESS) STOP 'CRTM_Adjoint failed'
  err = CRTM_Tangent_Linear(atm, sfc, atm_tl(1, :), sfc_tl, geo, chinfo, rts, rts_tl, opt)
  IF (err /= SUCCESS) STOP 'CRTM_Tangent_Linear failed'
  OPEN(NEWUNIT=u, FILE=TRIM(fout)//'.ad', STATUS='REPLACE', ACTION='WRITE')
  DO m1 = 1, M
    K = atm(m1)%n_Layers
    WRITE(u, '(*(ES25.16E3))') atm_ad(1, m1)%Temperature(1:K)
    DO j = 1, na
      WRITE(u, '(*(ES25.16E3))') atm_ad(1, m1)%Absorber(1:K, j)
    END DO
    WRITE(u, '(*(ES25.16E3))') sfc_ad(1, m1)%Water_Temperature
    WRITE(u, '(*(ES25.16E3))') (rts_tl(l1, m1)%Brightness_Temperature, l1 = 1, L)
  END DO
  CLOSE(u)

  OPEN(NEWUNIT=u, FILE=TRIM(fout), STATUS='REPLACE', ACTION='WRITE')
  WRITE(u, '(I8, I8)') M, L
  DO m1 = 1, M
    K = atm(m1)%n_Layers
    WRITE(u, '(*(ES25.16E3))') (rts_f(l1, m1)%Brightness_Temperature, l1 = 1, L)
    WRITE(u, '(*(ES25.16E3))') (rts(l1, m1)%Brightness_Temperature, l1 = 1, L)
    WRITE(u, '(*(ES25.16E3))') (sfc_k(l1, m1)%Water_Temperature, l1 = 1, L)
    WRITE(u, '(*(ES25.16E3))') (rts_k(l1, m1)%Surface_Emissivity, l1 = 1, L)
    DO k1 = 1, K
      WRITE(u, '(*(ES25.16E3))') (atm_k(l1, m1)%Temperature(k1), l1 = 1, L)
    END DO
    DO j = 1, na
      DO k1 = 1, K
        WRITE(u, '(*(ES25.16E3))') (atm_k(l1, m1)%Absorber(k1, j), l1 = 1, L)
      END DO
    END DO
  END DO
  CLOSE(u)

  ! Options%Skip_Profile on the last profile: its RTSolution is zero and the
  ! others are unchanged; RTSolution carries the sensor channel numbers
  opt(M)%Skip_Profile = .TRUE.
  err = CRTM_Forward(atm, sfc, geo, chinfo, rts_tl, opt)
  IF (err /= SUCCESS) STOP 'CRTM_Forward with Skip_Profile failed'
  OPEN(NEWUNIT=u, FILE=TRIM(fout)//'.skip', STATUS='REPLACE', ACTION='WRITE')
  WRITE(u, '(*(I8))') (rts_tl(l1, 1)%Sensor_Channel, l1 = 1, L)
  DO m1 = 1, M
    WRITE(u, '(*(ES25.16E3))') (rts_tl(l1, m1)%Brightness_Temperature, l1 = 1, L)
  END DO
  CLOSE(u)

  ! computed water surface optics: emissivity / reflectivity and Surface_K
  ! from the wind, salinity and view azimuth
  DO m1 = 1, M
    opt(m1)%Skip_Profile = .FALSE.
    opt(m1)%Use_Emissivity = .FALSE.
    sfc(m1)%Wind_Speed = 4.0_fp + REAL(m1, fp)
    sfc(m1)%Wind_Direction = 30.0_fp * REAL(m1, fp)
    sfc(m1)%Salinity = 34.0_fp
    geo(m1)%Sensor_Azimuth_Angle = 100.0_fp
  END DO
  CALL CRTM_Atmosphere_Zero(atm_k)
  CALL CRTM_Surface_Zero(sfc_k)
  rts_k(:, :)%Brightness_Temperature = 1.0_fp
  rts_k(:, :)%Radiance = 0.0_fp
  err = CRTM_K_Matrix(atm, sfc, rts_k, geo, chinfo, atm_k, sfc_k, rts, opt)
  IF (err /= SUCCESS) STOP 'CRTM_K_Matrix with computed surface optics failed'
  OPEN(NEWUNIT=u, FILE=TRIM(fout)//'.sfc', STATUS='REPLACE', ACTION='WRITE')
  DO m1 = 1, M
    K = atm(m1)%n_Layers
    WRITE(u, '(*(ES25.16E3))') (rts(l1, m1)%Brightness_Temperature, l1 = 1, L)
    WRITE(u, '(*(ES25.16E3))') (rts(l1, m1)%Surface_Emissivity, l1 = 1, L)
    WRITE(u, '(*(ES25.16E3))') (rts(l1, m1)%Surface_Reflectivity, l1 = 1, L)
    WRITE(u, '(*(ES25.16E3))') (sfc_k(l1, m1)%Water_Temperature, l1 = 1, L)
    WRITE(u, '(*(ES25.16E3))') (sfc_k(l1, m1)%Wind_Speed, l1 = 1, L)
    DO k1 = 1, K
      WRITE(u, '(*(ES25.16E3))') (atm_k(l1, m1)%Temperature(k1), l1 = 1, L)
    END DO
  END DO
  CLOSE(u)
  err = CRTM_Destroy(chinfo)
END PROGRAM crtm_mi355x_driver
