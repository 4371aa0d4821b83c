/* crtm_oracle.h -- FP64 CPU restatement of CRTM's per-profile x per-channel
 * hot path.  TEST INFRASTRUCTURE ONLY: the oracle is the checker for the
 * HIP implementation in crtmv3_amd/csrc; it is never linked into the
 * product.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg load it.
 *
 * Array conventions: Fortran column-major arrays from the coefficient files
 * are passed unchanged (e.g. n_predictors(J,L) -> index j + J*l).  Profile
 * layers are top-down, levels 0..K.  absorber[j*K + k].
 */
#ifndef CRTM_ORACLE_H
#define CRTM_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  /* ODPS TauCoeff (src/Coefficients/TauCoeff/ODPS/ODPS_Define.f90:109-205) */
  int group, n_layers, n_components, n_absorbers, n_channels;
  const int* absorber_id;            /* [Jm] */
  const double* ref_level_pressure;  /* [Ko+1] */
  const double* ref_pressure;        /* [Ko] */
  const double* ref_temperature;     /* [Ko] */
  const double* ref_absorber;        /* (Ko,Jm) column-major */
  const double* min_absorber;
  const double* max_absorber;
  const int* n_predictors;           /* (J,L) */
  const int* pos_index;              /* (J,L), 1-based */
  const float* c;
  int n_ocoeffs;
  const int* osignificance;          /* [L] */
  const int* order;                  /* [L] */
  const int* op_index;               /* (0:6,L) */
  const int* opos_index;             /* [L], 1-based */
  const double* oc;
  double alpha, alpha_c1, alpha_c2;
  int ocomponent_index;
  /* SpcCoeff (src/Coefficients/SpcCoeff/SpcCoeff_Define.f90:163-189) */
  int sensor_type;
  const double* planck_c1;
  const double* planck_c2;
  const double* band_c1;
  const double* band_c2;
  const double* cosmic_background;
  const double* frequency;           /* GHz [L] (microwave cloud tables) */
  const double* wavenumber;          /* cm^-1 [L] (infrared tables, stream choice) */
  /* antenna correction (src/Coefficients/ACCoeff/ACCoeff_Define.f90:78-96), (n_fov, L)
   * column-major; ac_nfov = 0: absent */
  int ac_nfov;
  const double *ac_earth, *ac_space, *ac_platform;
  /* non-LTE correction (src/Coefficients/NLTECoeff/NLTECoeff_Define.f90:87-121); nl_nsen = 0: absent */
  int nl_nsen, nl_nsol, nl_nch;
  const double* nl_sec_sensor;       /* [nsen] */
  const double* nl_sec_solar;        /* [nsol] */
  const int* nl_c_index;             /* [L] 1-based, 0 = not an NLTE channel */
  const double* nl_c;                /* (3, nsen, nsol, nch) column-major */
  double nl_upper[2], nl_lower[2], nl_min_tm[2], nl_max_tm[2], nl_mean_tm[2];
} or_sensor;

typedef struct {
  int n_layers, n_absorbers;
  const int* absorber_id;       /* [nabs] */
  const double* level_pressure; /* [K+1] */
  const double* pressure;       /* [K] */
  const double* temperature;    /* [K] */
  const double* absorber;       /* [nabs*K] */
  double coverage[4];           /* land, water, snow, ice */
  double sfc_temperature[4];
  double sensor_zenith, sensor_scan, surface_altitude;
  int climatology;              /* 1..6 (src/Atmosphere/CRTM_Atmosphere_Define.f90:406-411), 0 = US standard */
  double source_zenith;         /* degrees (non-LTE correction; >= 90: no source) */
  int ifov;                     /* Geometry%iFOV of the antenna correction (0: none) */
  int use_antenna_correction;   /* Options%Use_Antenna_Correction */
  int no_nlte_correction;       /* Options%Apply_NLTE_Correction = .FALSE. */
} or_profile;

typedef struct {
  double* radiance;       /* [L] */
  double* bt;             /* [L] */
  double* tau;            /* [L*K] nadir layer OD, optional */
  double* up_radiance;    /* [L] optional */
  double* down_radiance;  /* [L] optional */
  /* Jacobians (K-matrix), optional */
  double* t_k;            /* [L*K] */
  double* abs_k;          /* [L*nabs*K] */
  double* sfc_t_k;        /* [L*4] land, water, snow, ice */
  double* emis_k;         /* [L] */
  double* radiance_k;     /* [L] radiance adjoint after the BT pre-process */
} or_out;

/* Forward model for one profile, all channels. */
int oracle_forward(const or_sensor* s, const or_profile* p, const double* emissivity, or_out* o);

/* K-matrix by the hand-written reverse-mode adjoint (CRTM_K_Matrix analogue).
 * bt_k, rad_k: RTSolution_K Brightness_Temperature / Radiance inputs [L]. */
int oracle_k_matrix(const or_sensor* s, const or_profile* p, const double* emissivity,
                    const double* bt_k, const double* rad_k, or_out* o);

/* K-matrix by complex-step forward-mode differentiation of the forward
 * restatement: exact derivatives, independent of the hand-written adjoint. */
int oracle_k_matrix_cstep(const or_sensor* s, const or_profile* p, const double* emissivity,
                          const double* bt_k, const double* rad_k, or_out* o);

/* Batched K-matrix over M profiles with OpenMP (CPU baseline); outputs are
 * concatenated per profile with the same per-profile layout as above. */
int oracle_k_matrix_batch(const or_sensor* s, int n_profiles, const or_profile* p,
                          const double* emissivity, const double* bt_k, const double* rad_k,
                          or_out* o, int n_threads);

/* ---- Surface optics of a water surface (computed emissivity) ---- */
typedef struct {
  /* FASTEM FitCoeff arrays, Fortran order (src/Coefficients/EmisCoeff/MW_Water/MWwaterCoeff_Define.f90:78-92):
   * fc[2], fr[6], rc(3,7,2), az(6,6,2), ssc[8], lsc(3,6,2) */
  const double *fc, *fr, *rc, *az, *ssc, *lsc;
  int version;
} or_mwwater;

typedef struct {
  /* IR sea-surface emissivity table (src/Coefficients/EmisCoeff/IR_Water/IRwaterCoeff_Define.f90:84-106) */
  int n_angle, n_frequency, n_wind, n_temperature, nalli2;
  const double *angle, *frequency, *wind_speed, *temperature;
  const double* emissivity;   /* (angle, frequency, wind, temperature) column-major */
} or_irwater;

typedef struct {
  const or_mwwater* mw;       /* microwave sensors */
  const or_irwater* ir;       /* infrared sensors */
  const int* polarization;    /* [L] SpcCoeff%Polarization */
  const double* pol_angle;    /* [L] SpcCoeff%PolAngle, degrees */
  double salinity, wind_speed, wind_direction, sensor_azimuth;
} or_surface_optics;

/* Compute_FastemX at one angle; in5 = temperature, salinity, wind speed,
 * relative azimuth, transmittance; d_emis/d_refl [4][5] the derivatives of
 * the four Stokes components with respect to in5 (complex step). */
void oracle_fastemx(const or_mwwater* c, double frequency, int n_angles, double zenith, const double* in5,
                    double* emis, double* refl, double* d_emis, double* d_refl);
/* CRTM_Compute_IRSSEM at one angle; d2 = d/d(wind speed), d/d(temperature) */
void oracle_irssem(const or_irwater* c, double temperature, double wind, double frequency, double angle,
                   double* emis, double* d2);
/* Stokes weights (V, H) of a polarisation flag; returns nonzero for an unknown flag */
int oracle_pol_weights(int pol, double angle, double distance_ratio, double pol_angle, double scan_rad,
                       double* w2);
/* Clear-sky forward / K-matrix with the surface optics computed (Water_Coverage
 * only); emis/refl [L] the emissivity and reflectivity at the sensor angle,
 * sfc_k [L][3] the salinity, wind speed and wind direction Jacobians (the water
 * temperature one is in sfc_t_k). */
int oracle_forward_sfc(const or_sensor* s, const or_profile* p, const or_surface_optics* so, or_out* o,
                       double* emis, double* refl);
int oracle_k_matrix_sfc(const or_sensor* s, const or_profile* p, const or_surface_optics* so,
                        const double* bt_k, const double* rad_k, or_out* o, double* sfc_k);
int oracle_k_matrix_sfc_cstep(const or_sensor* s, const or_profile* p, const or_surface_optics* so,
                              const double* bt_k, const double* rad_k, or_out* o, double* sfc_k);

/* The six model climatologies on the 101 model levels
 * (src/Atmosphere/CRTM_Model_Profiles.f90): p[101] hPa, t[6][101] K,
 * a[6 models][6 absorbers in id order H2O,CO2,O3,N2O,CO,CH4][101].  Must be
 * set before profiles whose top is below TOA_PRESSURE are computed. */
void oracle_set_climatology(const double* p, const double* t, const double* a);

/* CRTM_Atmosphere_AddLayers (src/Atmosphere/CRTM_Atmosphere.f90:248-400):
 * number of layers added above the profile (0 when its top is at or above
 * TOA_PRESSURE); when the out arrays are non-NULL (sized K+n+1 / K+n /
 * nabs*(K+n)) the extended profile is written to them. */
int oracle_add_layers(const or_profile* p, double* level_pressure, double* pressure, double* temperature,
                      double* absorber);

/* ---------------- all-sky (clouds, ADA) ---------------- */

/* CloudCoeff release 3 (src/Coefficients/CloudCoeff/CloudCoeff_Define.f90:95-154);
 * arrays column-major as read from the file. */
typedef struct {
  int n_mw_frequencies, n_mw_radii, n_ir_frequencies, n_ir_radii, n_temperatures;
  int n_mw_densities, n_ir_densities, n_legendre_terms, n_phase_elements;
  const double *frequency_mw, *frequency_ir, *reff_mw, *reff_ir, *temperature, *density_mw;
  const double *ke_l_mw, *w_l_mw, *pcoeff_l_mw;  /* (f,r,T), (f,r,T,0:nL,nP) */
  const double *ke_s_mw, *w_s_mw, *pcoeff_s_mw;  /* (f,r,D), (f,r,D,0:nL,nP) */
  const double *ke_ir, *w_ir, *pcoeff_ir;        /* (f,r,Dir), (f,r,Dir,1:nL,nP) */
  const struct or_aerosolcoeff* aerosol;         /* AerosolCoeff tables, NULL when not loaded */
} or_cloudcoeff;

/* AerosolCoeff release 4, 'CRTM' scheme (src/Coefficients/AerosolCoeff/
 * AerosolCoeff_Define.f90:100-128); arrays column-major as read from the file. */
typedef struct or_aerosolcoeff {
  int n_wavelengths, n_radii, n_sigma, n_types, n_rh, n_legendre_terms, n_phase_elements;
  const int* type_id;        /* [n_types] aerosol type ids of the table */
  const double* frequency;   /* [n_wavelengths] cm^-1, ascending (10^4 / wavelength) */
  const double* reff;        /* (n_radii, n_types) microns */
  const double *ke, *w;      /* (n_wavelengths, n_rh, n_radii, n_sigma, n_types) */
  const double* pcoeff;      /* (n_wavelengths, n_rh, n_radii, n_sigma, n_types, 0:nL, nP) */
} or_aerosolcoeff;

typedef struct {
  int n_clouds;
  const int* type;        /* [n_clouds], CRTM cloud type 1..6 */
  const double* wc;       /* [n_clouds*K] water content, kg/m^2 */
  const double* reff;     /* [n_clouds*K] effective radius, microns */
  const double* fraction; /* [K] layer cloud fraction (NULL: 0) */
  int n_aerosols;                 /* Atmosphere%Aerosol(:) slots */
  const int* aerosol_type;        /* [n_aerosols] type id (0: empty slot) */
  const double* aerosol_conc;     /* [n_aerosols*K] concentration, kg/m^2 */
  const double* aerosol_reff;     /* [n_aerosols*K] effective radius, microns */
} or_clouds;

typedef struct {
  int n_streams;  /* Options%n_Streams (0: chosen from the Mie parameter) */
  int overlap;    /* Options%Overlap_Id 1..5 (0: default, average) */
  int no_scattering; /* nonzero: Options%Include_Scattering = .FALSE. (absorption only) */
} or_scatter_opts;

typedef struct {
  double* radiance;          /* [L] */
  double* bt;                /* [L] */
  double* r_clear;           /* [L] optional */
  double* tb_clear;          /* [L] optional */
  double* total_cloud_cover; /* [L] optional */
  double* sod;               /* [L] optional, scattering optical depth */
  int* n_streams;            /* [L] optional, Legendre terms of the scattering RT (0: Emission) */
} or_allsky_out;

typedef struct {
  double* t_k;      /* [L][K] */
  double* abs_k;    /* [L][nabs][K] */
  double* wc_k;     /* [L][n_clouds][K] Atmosphere_K%Cloud(n)%Water_Content */
  double* reff_k;   /* [L][n_clouds][K] Atmosphere_K%Cloud(n)%Effective_Radius */
  double* cf_k;     /* [L][K] Atmosphere_K%Cloud_Fraction */
  double* sfc_t_k;  /* [L][4] */
  double* emis_k;   /* [L] */
  double* aconc_k;  /* [L][n_aerosols][K] Atmosphere_K%Aerosol(n)%Concentration */
  double* areff_k;  /* [L][n_aerosols][K] Atmosphere_K%Aerosol(n)%Effective_Radius */
} or_allsky_k;

/* CRTM_Forward with clouds for one profile, all channels. */
int oracle_forward_allsky(const or_sensor* s, const or_cloudcoeff* cc, const or_profile* p, const or_clouds* c,
                          const double* emissivity, const or_scatter_opts* opt, or_allsky_out* o);

/* CRTM_K_Matrix with clouds by complex-step differentiation (exact derivatives). */
int oracle_k_matrix_allsky(const or_sensor* s, const or_cloudcoeff* cc, const or_profile* p, const or_clouds* c,
                           const double* emissivity, const or_scatter_opts* opt, const double* bt_k,
                           const double* rad_k, or_allsky_out* fo, or_allsky_k* ko);
/* The same K-matrix by the hand-written adjoint (CRTM_K_Matrix's AD chain). */
int oracle_k_matrix_allsky_ad(const or_sensor* s, const or_cloudcoeff* cc, const or_profile* p, const or_clouds* c,
                              const double* emissivity, const or_scatter_opts* opt, const double* bt_k,
                              const double* rad_k, or_allsky_out* fo, or_allsky_k* ko);
int oracle_k_matrix_allsky_ad_batch(const or_sensor* s, const or_cloudcoeff* cc, int n_profiles, const or_profile* p,
                                    const or_clouds* c, const double* emissivity, const or_scatter_opts* opt,
                                    const double* bt_k, double* bt, double* t_k, double* wc_k,
                                    const long long* offset, int nc_stride, int n_threads);
int oracle_forward_allsky_batch(const or_sensor* s, const or_cloudcoeff* cc, int n_profiles, const or_profile* p,
                                const or_clouds* c, const double* emissivity, const or_scatter_opts* opt,
                                double* radiance, double* bt, int n_threads);

/* White-box pieces of the all-sky path. */
int oracle_aerosol_optics(const or_aerosolcoeff* ac, double wavenumber, int type_id, double reff, int n_legendre,
                          double* ke, double* w, double* pcoeff);
int oracle_cloud_optics(const or_cloudcoeff* cc, int microwave, double freq, int type, double reff,
                        double temperature, int n_legendre, double* ke, double* w, double* pcoeff);
void oracle_lpoly_interp(const double* x, int n, double xi, const double* z, double* out);
void oracle_gauss(int n, double* mu, double* w);
int oracle_phase_matrix(int n_streams, double cos_sensor, int n_legendre, const double* pl, double* mu,
                        double* wt, double* pff, double* pbb);
int oracle_matrix_exp(int n, double tau, const double* A, const double* B, double* refl, double* trans);
double oracle_cloud_cover(int K, int n_clouds, double* wc, const double* cf, int overlap);
double oracle_combine(int n_legendre, int K, double* od, double* ssa, double* pc);
int oracle_ada(int n_streams, double cos_sensor, int n_legendre, int K, const double* tau, const double* ssa,
               const double* pl, const double* planck, double planck_sfc, double emissivity, double cosmic,
               double* radiance);
int oracle_amom_layer(int n_streams, double cos_sensor, int n_legendre, const double* pl, double tau, double ssa,
                      double planck, double* refl, double* trans, double* src);

/* Intermediate quantities for white-box tests. */
int oracle_layer_avg(int n_odps, const double* ref_lnp, int n_user, const double* user_lnp,
                     double* weights /* [n_odps*n_user] (user fastest) */, int* index /* [2*n_odps] */);
int oracle_predictors(const or_sensor* s, const or_profile* p, double* x /* [Ko*J*18] */,
                      double* secant /* [Ko] */, double* t_odps, double* a_odps);

#ifdef __cplusplus
}
#endif
#endif
