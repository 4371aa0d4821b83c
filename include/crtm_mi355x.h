/*
 * crtm_mi355x.h -- C-ABI of the MI355X implementation of CRTM's forward and
 * K-matrix hot path.  Plain pointers and sizes only; every entry point
 * returns a CRTM status code.  This is the boundary that the Fortran
 * ISO_C_BINDING layer (crtmv3_amd/fortran/crtm_mi355x_module.f90, re-exported
 * under the reference's module name by crtmv3_amd/fortran/crtm_module.f90) and
 * the Python ctypes binding (crtmv3_amd/api.py) bind.
 *
 * Layout conventions
 *   - M profiles, each with K_m layers (top-down, index 0 = top of atmosphere)
 *     and K_m+1 levels.  layer_offset[m] = sum_{i<m} K_i (length M+1).
 *     Per-layer arrays have layer_offset[M] elements, per-level arrays
 *     layer_offset[M] + M (level block of profile m starts at
 *     layer_offset[m] + m).  absorber is [n_absorbers][layer_offset[M]].
 *   - L = total channels over all sensors of the context (ChannelInfo order).
 *     Per-channel outputs are [M][L].
 *   - Jacobians (K-matrix) for profile m are stored channel-fastest,
 *     [K_m][L] per variable; the block of profile m starts at
 *     layer_offset[m] * L.  absorber_k is [n_absorbers] such blocks.
 *   - memory = CRTM_MEM_HOST: all batch pointers are host pointers
 *     (the library stages them through its device pool);
 *     memory = CRTM_MEM_DEVICE: all batch pointers (except the absorber id
 *     and unit lists) are device pointers on the context's GPU; with
 *     max_layers and total_layers given the call is asynchronous on the
 *     context stream (crtm_synchronize waits) and issues no host sync.
 */
#ifndef CRTM_MI355X_H
#define CRTM_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: src/Utility/Message_Handler.f90:27-31 */
#define CRTM_SUCCESS     0
#define CRTM_INFORMATION 1
#define CRTM_WARNING     2
#define CRTM_FAILURE     3

/* Sensor types: src/InstrumentInfo/SensorInfo/SensorInfo_Parameters.f90:74-78 */
#define CRTM_INVALID_SENSOR     0
#define CRTM_MICROWAVE_SENSOR   1
#define CRTM_INFRARED_SENSOR    2
#define CRTM_VISIBLE_SENSOR     3
#define CRTM_ULTRAVIOLET_SENSOR 4

/* Arithmetic of the call */
#define CRTM_PREC_FP64       0  /* everything in FP64 */
#define CRTM_PREC_FP32_SPLIT 1  /* FP32 absorption + RT, FP64 Planck / BT inversion */

#define CRTM_MEM_HOST   0
#define CRTM_MEM_DEVICE 1

#define CRTM_MAX_N_LAYERS 200   /* src/CRTM_Parameters.f90:64 */

typedef struct crtm_context crtm_context;

/* ---- Atmosphere (src/Atmosphere/CRTM_Atmosphere_Define.f90:433-464) ---- */
typedef struct {
  int n_profiles;
  int max_layers;                  /* host scalar: max K_m (0 = derive from n_layers) */
  int64_t total_layers;            /* host scalar: layer_offset[M] (0 = derive) */
  const int* n_layers;             /* [M] */
  const int64_t* layer_offset;     /* [M+1] */
  int n_absorbers;
  const int* absorber_id;          /* [n_absorbers], host array in both memory modes */
  const int* absorber_units;       /* [n_absorbers], host array in both memory modes */
  const double* level_pressure;    /* hPa, [layer_offset[M]+M] */
  const double* pressure;          /* hPa, [layer_offset[M]] */
  const double* temperature;       /* K */
  const double* absorber;          /* [n_absorbers][layer_offset[M]] */
  const int* climatology;          /* [M] model climatology 1..6 used to extend a profile
                                      whose top is below TOA_PRESSURE (CRTM_Atmosphere_AddLayers,
                                      src/Atmosphere/CRTM_Atmosphere.f90:248-400); NULL = US standard */
  /* Clouds (Atmosphere%Cloud(n), src/Atmosphere/Cloud/CRTM_Cloud_Define.f90:295-310, and
   * Atmosphere%Cloud_Fraction): n_clouds slots per profile; 0 = clear-sky call. */
  int n_clouds;
  const int* cloud_type;               /* [M][n_clouds] CRTM cloud type 1..6
                                          (WATER, ICE, RAIN, SNOW, GRAUPEL, HAIL; 0 = empty slot) */
  const double* cloud_water_content;   /* kg/m^2, [n_clouds][layer_offset[M]] */
  const double* cloud_effective_radius;/* microns, [n_clouds][layer_offset[M]] */
  const double* cloud_fraction;        /* [layer_offset[M]]; clouds of layers with a fraction below
                                          1e-12 are ignored (src/CRTM_Forward_Module.f90:368-378) */
  /* Aerosols (Atmosphere%Aerosol(n), src/Atmosphere/Aerosol/CRTM_Aerosol_Define.f90:119-133):
   * n_aerosols slots per profile; infrared sensors only (CRTM_AerosolScatter.f90:200). */
  int n_aerosols;
  const int* aerosol_type;                 /* [M][n_aerosols] type id of the AerosolCoeff tables
                                              (0 = empty slot) */
  const double* aerosol_concentration;     /* kg/m^2, [n_aerosols][layer_offset[M]] */
  const double* aerosol_effective_radius;  /* microns, [n_aerosols][layer_offset[M]] */
} crtm_atmosphere;

/* ---- Surface (src/Surface/CRTM_Surface_Define.f90:204-244): [M] each ---- */
typedef struct {
  const double* land_coverage;
  const double* water_coverage;
  const double* snow_coverage;
  const double* ice_coverage;
  const double* land_temperature;
  const double* water_temperature;
  const double* snow_temperature;
  const double* ice_temperature;
  /* Water surface inputs of the computed emissivity (use_emissivity = 0),
   * [M] each, NULL = the reference defaults (src/Surface/CRTM_Surface_Define.f90:183-185) */
  const double* wind_speed;       /* m/s, default 5 */
  const double* wind_direction;   /* degrees, default 0 */
  const double* salinity;         /* ppt, default 33 */
} crtm_surface;

/* ---- Geometry (src/GeometryInfo/CRTM_Geometry_Define.f90:99-121): [M] ---- */
typedef struct {
  const double* sensor_zenith_angle;   /* degrees */
  const double* sensor_scan_angle;     /* degrees, may be NULL */
  const double* source_zenith_angle;   /* degrees, may be NULL (100 = no source) */
  const double* surface_altitude;      /* metres, may be NULL */
  const double* sensor_azimuth_angle;  /* degrees, may be NULL (999.9: no azimuth dependence of the
                                          FASTEM emissivity; src/GeometryInfo/CRTM_Geometry_Define.f90:111) */
  const int* ifov;                     /* Geometry%iFOV, the field-of-view index of the antenna
                                          correction (1..n_FOVs of the sensor's ACCoeff data; 0: no
                                          correction), may be NULL (all 0) */
} crtm_geometry;

/* ---- Options (src/Options/CRTM_Options_Define.f90:122-179) ---- */
typedef struct {
  int use_emissivity;            /* Options%Use_Emissivity; 0: the surface optics are computed
                                    (CRTM_Compute_SfcOptics, src/SfcOptics/CRTM_SfcOptics.f90:454):
                                    FASTEM6 for microwave and the IR sea-surface emissivity table
                                    for infrared sensors, for water surfaces (land, snow and ice
                                    coverage 0) in clear-sky calls; needs
                                    FASTEM6.MWwater.EmisCoeff.bin / Nalli.IRwater.EmisCoeff.bin in
                                    the crtm_init path */
  const double* emissivity;      /* [M][L] when use_emissivity */
  int precision;                 /* CRTM_PREC_* */
  int memory;                    /* CRTM_MEM_* */
  int jacobian_f32;              /* K outputs are float (1) or double (0) */
  int check_input;               /* Options%Check_Input */
  int n_streams;                 /* Options%n_Streams with Use_N_Streams (2, 4, 6, 8, 16);
                                    0 = chosen per channel from the Mie parameter
                                    (CRTM_Compute_nStreams, src/RTSolution/CRTM_RTSolution.f90:1204) */
  int overlap_id;                /* Options%Overlap_Id 1..5 (maximum, random, maximum-random,
                                    average, overcast); 0 = default (average) */
  const int* n_streams_profile;  /* [M] or NULL: Options(m)%n_Streams per profile (the
                                    reference's Options is per profile), overriding
                                    n_streams; same values.  Host memory is read
                                    directly in either memory mode; device memory is
                                    copied back with one synchronisation.  Under HIP
                                    graph capture a cloudy call takes them in device
                                    memory only, reads nothing back, launches the
                                    kernels of every stream count and fails a profile
                                    whose count is not 0, 2, 4, 6, 8 or 16 (status 8) */
  int no_scattering;             /* nonzero: Options%Include_Scattering = .FALSE.: clouds and
                                    aerosols absorb only (extinction ke (1 - w), no phase
                                    function) and the radiance is the emission solution
                                    (src/AtmScatter/CRTM_CloudScatter.f90:986-997, 1372-1407;
                                    src/RTSolution/Common_RTSolution.f90:263).  0 keeps the
                                    reference default .TRUE., so zero-initialised options
                                    behave as the reference's defaults */
  const int* skip_profile;       /* [M] or NULL: Options(m)%Skip_Profile.  A nonzero entry
                                    skips profile m: its RTSolution is zero, its Jacobians
                                    are zero, radiance_k is left as given, and it neither
                                    fails input checks nor fails the call
                                    (src/CRTM_Forward_Module.f90:498, 536;
                                    src/CRTM_K_Matrix_Module.f90:594, 631).  Host memory in
                                    host-memory calls; device memory in device-memory calls
                                    (host memory there is staged, outside graph capture) */
  int use_antenna_correction;    /* nonzero: Options%Use_Antenna_Correction = .TRUE. (default
                                    .FALSE.): the brightness temperature of a sensor with antenna
                                    correction data in its SpcCoeff file becomes the antenna
                                    temperature (A_earth + A_platform) Tb + A_space TSPACE at the
                                    profile's Geometry%iFOV (src/AntennaCorrection/
                                    CRTM_AntennaCorrection.f90:118-140); clear-sky calls */
  int no_nlte_correction;        /* nonzero: Options%Apply_NLTE_Correction = .FALSE.; 0 keeps the
                                    reference default .TRUE.: sensors with non-LTE correction data
                                    correct the radiance of their NLTE channels where the source
                                    (sun) is above the horizon (src/NLTE/NLTE_Correction/
                                    CRTM_NLTECorrection.f90); clear-sky calls */
} crtm_options;

/* ---- RTSolution (src/RTSolution/CRTM_RTSolution_Define.f90:201-245) ---- */
typedef struct {
  double* radiance;                /* [M][L] */
  double* brightness_temperature;  /* [M][L] */
  double* up_radiance;             /* [M][L], optional */
  double* down_radiance;           /* [M][L], optional */
  double* surface_planck_radiance; /* [M][L], optional */
  double* surface_emissivity;      /* [M][L], optional */
  int* status;                     /* [M] per-profile status, optional */
  double* r_clear;                 /* RTSolution%R_Clear  [M][L], optional */
  double* tb_clear;                /* RTSolution%Tb_Clear [M][L], optional */
  double* total_cloud_cover;       /* RTSolution%Total_Cloud_Cover [M][L], optional */
  double* surface_reflectivity;    /* RTSolution%Surface_Reflectivity [M][L], optional */
} crtm_rtsolution;

/* ---- K-matrix adjoint inputs/outputs ---- */
typedef struct {
  /* RTSolution_K inputs ([M][L]); radiance_k is overwritten with the total
   * radiance adjoint (Pre_Process_RTSolution_K, src/CRTM_K_Matrix_Module.f90:1725) */
  const double* brightness_temperature_k;
  double* radiance_k;
  /* RTSolution_K%Surface_Emissivity output [M][L], optional */
  double* surface_emissivity_k;
  /* Atmosphere_K: layout above; element type per options.jacobian_f32.
   * Overwritten (the reference accumulates into caller-zeroed arrays; the
   * Fortran/Python layers zero-then-add to keep that contract). */
  void* temperature_k;
  void* absorber_k;
  /* Surface_K temperatures [M][L] each, optional */
  double* land_temperature_k;
  double* water_temperature_k;
  double* snow_temperature_k;
  double* ice_temperature_k;
  /* Atmosphere_K%Cloud(n)%Water_Content and %Effective_Radius
   * [n_clouds][NT][L], and Atmosphere_K%Cloud_Fraction [NT][L]
   * (CRTM_Compute_CloudScatter_AD, src/AtmScatter/CRTM_CloudScatter.f90:737-907;
   * Compute_CloudCover_AD, src/Atmosphere/Cloud/CRTM_CloudCover_Define.f90).
   * Element type per options.jacobian_f32; required when the atmosphere has
   * clouds, ignored otherwise; overwritten. */
  void* cloud_water_content_k;
  void* cloud_effective_radius_k;
  void* cloud_fraction_k;
  /* Atmosphere_K%Aerosol(n)%Concentration and %Effective_Radius [n_aerosols][NT][L]
   * (CRTM_Compute_AerosolScatter_AD, src/AtmScatter/CRTM_AerosolScatter.f90:614-751);
   * required when the atmosphere has aerosols; overwritten. */
  void* aerosol_concentration_k;
  void* aerosol_effective_radius_k;
  /* Surface_K%Salinity, %Wind_Speed, %Wind_Direction [M][L], optional: the
   * computed optics' sensitivities (CRTM_Compute_SfcOptics_AD,
   * src/SfcOptics/CRTM_SfcOptics.f90:1857; zero with user emissivity) */
  double* salinity_k;
  double* wind_speed_k;
  double* wind_direction_k;
} crtm_k_output;

/* CRTM_Init (src/CRTM_LifeCycle.f90:520): loads <id>.SpcCoeff.bin and
 * <id>.TauCoeff.bin (ODPS) for each sensor from file_path, and CloudCoeff.bin
 * (release 3, binary) and AerosolCoeff.bin (release 4, binary, 'CRTM' scheme)
 * when present there, uploads the coefficients to the
 * current GPU and creates the context stream.  Cloudy calls need the
 * CloudCoeff data (src/CRTM_Forward_Module.f90:379-386). */
int crtm_init(int n_sensors, const char* const* sensor_ids, const char* file_path,
              crtm_context** ctx);

/* CRTM_Destroy (src/CRTM_LifeCycle.f90:1131) */
int crtm_destroy(crtm_context* ctx);

/* ChannelInfo (src/ChannelInfo/CRTM_ChannelInfo_Define.f90:70-85) */
int crtm_n_sensors(const crtm_context* ctx);
int crtm_n_channels_total(const crtm_context* ctx);
int crtm_channel_info(const crtm_context* ctx, int sensor_index, int* n_channels, int* sensor_type,
                      int* sensor_channel /* [n_channels] or NULL */,
                      char* sensor_id /* >= 21 bytes or NULL */);

/* CRTM_Forward (src/CRTM_Forward_Module.f90:229-245).  With clouds the
 * radiances follow the all-sky path: cloud optics, AtmOptics combination and
 * the ADA scattering solver (or Emission when no layer scatters), combined
 * with the clear column by the total cloud cover when any cloudy layer has a
 * cloud fraction above 1e-6.  A cloudy device-memory call synchronises once
 * with the host (the number of cloudy layers sizes the layer work). */
int crtm_forward(crtm_context* ctx, const crtm_atmosphere* atm, const crtm_surface* sfc,
                 const crtm_geometry* geo, const crtm_options* opt, crtm_rtsolution* rts);

/* CRTM_K_Matrix (src/CRTM_K_Matrix_Module.f90:300-325) */
int crtm_k_matrix(crtm_context* ctx, const crtm_atmosphere* atm, const crtm_surface* sfc,
                  const crtm_geometry* geo, const crtm_options* opt, crtm_rtsolution* rts,
                  crtm_k_output* k);

/* ---- Adjoint (CRTM_Adjoint) inputs and outputs ----
 * Atmosphere_AD and Surface_AD are FP64 channel sums, accumulated into the
 * caller's arrays as the reference's INTENT(IN OUT) arguments are (callers
 * zero them first).  Layout: caller's layers, [NT] per variable, absorbers
 * and clouds [n][NT]; surface temperatures [M]. */
typedef struct {
  const double* brightness_temperature_ad;  /* RTSolution_AD%Brightness_Temperature [M][L] */
  double* radiance_ad;                      /* RTSolution_AD%Radiance [M][L], optional; overwritten with the
                                               total radiance adjoint as radiance_k is */
  double* surface_emissivity_ad;            /* [M][L] per-channel emissivity adjoint, optional, overwritten */
  double* temperature_ad;                   /* Atmosphere_AD%Temperature [NT] */
  double* absorber_ad;                      /* Atmosphere_AD%Absorber [n_absorbers][NT] */
  double* land_temperature_ad;              /* Surface_AD temperatures [M] each, optional */
  double* water_temperature_ad;
  double* snow_temperature_ad;
  double* ice_temperature_ad;
  double* cloud_water_content_ad;           /* Atmosphere_AD%Cloud(n)%Water_Content [n_clouds][NT] */
  double* cloud_effective_radius_ad;        /* Atmosphere_AD%Cloud(n)%Effective_Radius [n_clouds][NT] */
  double* cloud_fraction_ad;                /* Atmosphere_AD%Cloud_Fraction [NT] */
  double* aerosol_concentration_ad;         /* Atmosphere_AD%Aerosol(n)%Concentration [n_aerosols][NT] */
  double* aerosol_effective_radius_ad;      /* Atmosphere_AD%Aerosol(n)%Effective_Radius [n_aerosols][NT] */
  double* salinity_ad;                      /* Surface_AD%Salinity, %Wind_Speed, %Wind_Direction [M],
                                               optional (computed surface optics) */
  double* wind_speed_ad;
  double* wind_direction_ad;
} crtm_ad_output;

/* CRTM_Adjoint (src/CRTM_Adjoint_Module.f90:299-325): the forward RTSolution
 * and the channel-summed adjoint of RTSolution_AD.  The kernels reduce over
 * each 64-channel block before the per-profile rows are touched (FP64 atomic
 * adds across blocks, so the last bits depend on the order the blocks
 * finish); no per-channel Jacobian is stored.  Cloud and aerosol arrays are
 * required when the atmosphere carries them. */
int crtm_adjoint(crtm_context* ctx, const crtm_atmosphere* atm, const crtm_surface* sfc,
                 const crtm_geometry* geo, const crtm_options* opt, crtm_rtsolution* rts,
                 crtm_ad_output* ad);

/* CRTM_K_Matrix without device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory during a call (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results as crtm_k_matrix, restricted to
 * clear-sky, device-memory calls (options.memory = CRTM_MEM_DEVICE) with
 * atm->max_layers and atm->total_layers given.  The call's work space is
 * carved from a fixed 128 MiB device array that each GPU holds from library
 * load until the process ends; a batch needing more (above about 930 FP32-split
 * or 520 FP64 profiles of 100 layers for an ODPS group-1 sensor) is
 * rejected with CRTM_FAILURE before any work is issued.  The caller allocates
 * inputs and outputs.  The library allocates nothing during the call: no
 * hipMalloc, hipFree or allocating library routine.  All no-allocation calls on one GPU share its array, so
 * overlapping calls on one GPU (on different streams or host threads) race;
 * the array keeps the last call's contents.  Kernel timing
 * (crtm_profile_enable) must be off. */
int crtm_k_matrix_noalloc(crtm_context* ctx, const crtm_atmosphere* atm, const crtm_surface* sfc,
                          const crtm_geometry* geo, const crtm_options* opt, crtm_rtsolution* rts,
                          crtm_k_output* k);

/* Issue the context's work on `stream` (a hipStream_t of the context's
 * device, e.g. the caller's framework stream; NULL is the null stream)
 * instead of the context's own stream; own_stream != 0 restores the own
 * stream and ignores `stream`.  Device-memory calls are then
 * ordered with the caller's other work on that stream.  The reference has
 * no streams (its calls are synchronous); this is the MI355X analogue of
 * the OpenMP execution-context choice in src/CRTM_Forward_Module.f90:337-357. */
int crtm_set_stream(crtm_context* ctx, void* stream, int own_stream);

/* Scattering-kernel layout: angle counts >= min_angles run on the wavefront
 * kernels (one 64-lane wavefront per layer or profile-channel, stream
 * matrices on the matrix cores), smaller ones on the per-lane kernels (one
 * lane per layer or profile-channel, matrices in registers).  min_angles is
 * 6 (8 streams with the sensor angle appended, and 16 streams, on the
 * wavefront kernels) or 9 (the default: only 16 streams).  Results agree to
 * rounding either way.  No counterpart in the reference, whose solver has one
 * serial form (src/RTSolution/ADA/ADA_Module.f90). */
int crtm_set_wavefront_angles(crtm_context* ctx, int min_angles);

/* ---- Tile shapes of the fused clear-sky forward / K-matrix kernel ----
 * No counterpart in the reference (its channel loop has no GPU tiles).  A
 * workgroup of that kernel is one 64-channel block x P profiles (one wave per
 * profile); CRTM_N_TILES tile indices, index 1 the default, with P =
 * crtm_tile_profiles(precision, tile) (FP32-split 4 / 8 / 16, FP64 2 / 4 / 8).
 * Results do not depend on the tile (each profile-channel runs the same
 * arithmetic in one lane); only the speed does. */
#define CRTM_N_TILES 3
int crtm_tile_profiles(int precision, int tile);
/* tile used for sensor s (index in the crtm_init list) at a precision */
int crtm_set_tile(crtm_context* ctx, int sensor, int precision, int tile);
int crtm_get_tile(const crtm_context* ctx, int sensor, int precision, int* tile);
/* Autotuner: runs crtm_k_matrix with the caller's arguments `repeats` times
 * per tile (after one untimed run each), times the fused kernel of each
 * sensor with HIP events on the context stream, and keeps for every sensor
 * the tile with the shortest mean time at options->precision.  ms
 * [n_sensors][CRTM_N_TILES] (may be NULL) receives the mean kernel times in
 * milliseconds, best [n_sensors] (may be NULL) the chosen tile indices.  The
 * outputs hold a final run with the chosen tiles.  Synchronizes the context
 * stream (not capturable). */
int crtm_autotune_k_matrix(crtm_context* ctx, const crtm_atmosphere* atm, const crtm_surface* sfc,
                           const crtm_geometry* geo, const crtm_options* opt, crtm_rtsolution* rts,
                           crtm_k_output* k, int repeats, double* ms, int* best);

/* Scattering-layer capacity for calls with clouds or aerosols.  0 (the
 * default): each such call counts its scattering layers on the device and
 * reads the count back once to size the scattering launches.  max_layers > 0:
 * the launches and their work space are sized for max_layers scattering
 * layers of the batch, nothing is read back, and a device-memory call can be
 * captured in a HIP graph (after one eager call of the same shape has sized
 * the work space, with one n_streams for the batch).  A profile whose
 * scattering layers do not fit fails (status FAILURE, NaN radiances).  No
 * counterpart in the reference, which allocates per profile on the host
 * (src/CRTM_Forward_Module.f90:515-663). */
int crtm_set_cloudy_layer_capacity(crtm_context* ctx, int64_t max_layers);

/* Wait for the context stream (device-memory calls are asynchronous). */
int crtm_synchronize(crtm_context* ctx);

/* Last error message of the context (static storage, NUL-terminated). */
const char* crtm_last_error(const crtm_context* ctx);

/* Kernel timing, the analogue of the forward driver's compile-time timing
 * hook (src/CRTM_Forward_Module.f90:262-282): when enabled, every kernel
 * launch is bracketed by HIP events on the context stream.  query returns
 * the summed milliseconds of the preparation and forward/K kernels since the
 * last enable/query and resets them. */
int crtm_profile_enable(crtm_context* ctx, int on);
int crtm_profile_query(crtm_context* ctx, double* prep_ms, double* main_ms, int* n_launches);
/* Per kind: [0] preparation, [1] forward/K-matrix, [2] all-sky scattering
 * kernels (cloud optics, layer R/T, adding sweep), [3] surface optics;
 * n_kinds <= 4. */
int crtm_profile_query_kinds(crtm_context* ctx, double* ms, int* n_launches, int n_kinds);

#ifdef __cplusplus
}
#endif
#endif /* CRTM_MI355X_H */
