! crtm_module.f90 -- MODULE CRTM_Module, the module name CRTM applications
! USE (src/CRTM_Module.F90:7-38): it re-exports CRTM_MI355X_Module, the
! ISO_C_BINDING layer over libcrtm_mi355x.so, so that a caller written as
!
!   USE CRTM_Module
!   Error_Status = CRTM_Init(Sensor_Id, ChannelInfo, File_Path=...)
!   Error_Status = CRTM_K_Matrix(Atm, Sfc, RTSolution_K, Geometry, ChannelInfo, &
!                                Atmosphere_K, Surface_K, RTSolution, Options=Opt)
!
! compiles against this library without a change to its source.
MODULE CRTM_Module
  USE CRTM_MI355X_Module
  IMPLICIT NONE
  PUBLIC
END MODULE CRTM_Module
