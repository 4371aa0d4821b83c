// sfc_optics.h -- computed surface optics of a water surface: coefficient
// tables on the device, the per-(profile, channel) planes the optics kernels
// write, and the transmittance-dependent part the clear-sky forward/K kernel
// evaluates at the surface.
//
// Split of the work.  FASTEM-X (src/SfcOptics/MW_Water/FASTEM_MWSSEM/
// CRTM_FastemX.f90:302-464) depends on the column transmittance
// T = exp(-sum of the nadir layer optical depths) (CRTM_Compute_Transmittance,
// src/AtmOptics/CRTM_AtmOptics.f90:210-218) only through the reflection
// correction R_mod = (1 - T^rough) / (1 - T), rough = 1 + A0 + A1 x + A2 x^2,
// x = ln(-ln(T) cos_z).  The optics kernels (sfc_optics.hip) evaluate
// everything else once per (profile, channel) with forward-mode derivatives
// in water temperature, salinity, wind speed and relative azimuth; the fused
// clear-sky kernel, which knows T only after its layer loop, adds the
// reflection correction, and in the K-matrix chains its adjoint back through
// the planes and into every layer's optical depth
// (CRTM_Compute_Transmittance_AD, :314-368).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace crtm {

// FASTEM coefficient set (src/Coefficients/EmisCoeff/MW_Water/MWwaterCoeff_Define.f90:78-92),
// Fortran order: fc[2], fr[6], rc(3,7,2), az(6,6,2), ssc[8], lsc(3,6,2)
struct DevMWwater {
  const double* fc;
  const double* fr;
  const double* rc;
  const double* az;
  const double* ssc;
  const double* lsc;
};

// IR sea-surface emissivity table (src/Coefficients/EmisCoeff/IR_Water/IRwaterCoeff_Define.f90:84-106)
struct DevIRwater {
  int na, nf, nv, nt;
  int has_t;                 // 'Nalli2': interpolated in temperature too
  const double* angle;       // degrees
  const double* frequency;   // cm^-1
  const double* wind;
  const double* temp;
  const double* emis;        // (angle, frequency, wind, temperature) column-major
};

// Surface inputs of computed optics, [M] each (null: the reference defaults)
struct DevSfcIn {
  const double* salinity;        // Surface%Salinity, default 33
  const double* wind_speed;      // Surface%Wind_Speed, default 5 m/s
  const double* wind_direction;  // Surface%Wind_Direction, default 0
  const double* azimuth;         // Geometry%Sensor_Azimuth_Angle, default 999.9 (azimuth model off)
  const double* scan;            // Geometry%Sensor_Scan_Angle, default 0
};
constexpr double DEFAULT_SALINITY = 33.0;        // src/Surface/CRTM_Surface_Define.f90:185
constexpr double DEFAULT_WIND_SPEED = 5.0;       // :183
constexpr double DEFAULT_WIND_DIRECTION = 0.0;   // :184
constexpr double INVALID_AZIMUTH = 999.9;        // src/GeometryInfo/CRTM_Geometry_Define.f90:111

// Surface-optics mode of a sensor in a call
enum SfcMode { SFC_USER = 0, SFC_MW = 1, SFC_IR = 2 };

// Planes [plane][M][L of the sensor] written by the optics kernels.
// Microwave: Stokes emissivities e_v, e_h; sensor weights w_v, w_h (the
// polarisation combination times Water_Coverage); their derivatives in
// (water temperature, salinity, wind speed, azimuth); the reflection
// correction's polynomial A0..A2 per polarisation and d/d(wind speed); cos_z.
enum MWPlane {
  MP_EV = 0, MP_EH, MP_WV, MP_WH,
  MP_DEV = 4,        // 4 planes
  MP_DEH = 8,        // 4 planes
  MP_AV = 12,        // 3 planes
  MP_AH = 15,        // 3 planes
  MP_DAV = 18,       // 3 planes, d/d(wind speed)
  MP_DAH = 21,       // 3 planes
  MP_COSZ = 24,
  MP_N = 25
};
// Infrared: emissivity and reflectivity (Water_Coverage e, Water_Coverage
// (1 - e)) and Water_Coverage de/d(water temperature), de/d(wind speed)
enum IRPlane { IP_E = 0, IP_R, IP_DT, IP_DU, IP_N };

struct SfcPlanes {
  int mode;            // SfcMode
  const double* p;     // planes of the sensor
  int64_t stride;      // M * L of the sensor
};

// Forward state of the surface step kept for its adjoint
struct SfcLin {
  double rm[2];        // R_mod per polarisation
  double drm_dt[2];    // dR_mod/dT (T = transmittance), through rough as well
  double drm_dr[2];    // dR_mod/d(rough)
  double x;            // ln(-ln(T) cos_z)
  bool rc;             // reflection correction applied (0 < T < 1)
};

// Surface emissivity and reflectivity at the sensor angle from the planes
// and the column transmittance trans (element i = m * L + l).
__device__ inline void sfc_forward(const SfcPlanes& sp, int64_t i, double trans, double& eps, double& rho,
                                   SfcLin& z) {
  const double* P = sp.p + i;
  const int64_t S = sp.stride;
  if (sp.mode == SFC_IR) {
    eps = P[IP_E * S];
    rho = P[IP_R * S];
    z.rc = false;
    return;
  }
  const double ev = P[MP_EV * S], eh = P[MP_EH * S], wv = P[MP_WV * S], wh = P[MP_WH * S];
  z.rc = trans > 0.0 && trans < 1.0;
  z.rm[0] = z.rm[1] = 1.0;
  z.drm_dt[0] = z.drm_dt[1] = z.drm_dr[0] = z.drm_dr[1] = 0.0;
  z.x = 0.0;
  if (z.rc) {
    // Reflection_Correction (src/SfcOptics/MW_Water/FASTEM_MWSSEM/Reflection_Correction_Module.f90)
    const double lt = log(trans);
    const double od = -lt * P[MP_COSZ * S];
    const double x = log(od);
    z.x = x;
    const double dxdt = 1.0 / (trans * lt);          // d ln(-ln(T) c) / dT
    for (int q = 0; q < 2; ++q) {
      const double* A = P + (q ? MP_AH : MP_AV) * S;
      const double a1 = A[S], a2 = A[2 * S];
      const double rough = 1.0 + A[0] + x * a1 + x * x * a2;
      const double tp = pow(trans, rough);
      const double rm = (1.0 - tp) / (1.0 - trans);
      z.rm[q] = rm;
      z.drm_dr[q] = -tp * lt / (1.0 - trans);
      const double drm_dt_r = (rm - rough * pow(trans, rough - 1.0)) / (1.0 - trans);
      z.drm_dt[q] = drm_dt_r + z.drm_dr[q] * (a1 + 2.0 * x * a2) * dxdt;
    }
  }
  eps = wv * ev + wh * eh;
  rho = wv * z.rm[0] * (1.0 - ev) + wh * z.rm[1] * (1.0 - eh);
}

// Adjoint of sfc_forward (with CRTM_Compute_SfcOptics_AD's chain into the
// surface inputs): eps_ad, rho_ad -> transmittance adjoint and g[4] =
// (water temperature, salinity, wind speed, wind direction) adjoints.
__device__ inline void sfc_adjoint(const SfcPlanes& sp, int64_t i, const SfcLin& z, double eps_ad, double rho_ad,
                                   double& tr_ad, double g[4]) {
  const double* P = sp.p + i;
  const int64_t S = sp.stride;
  tr_ad = 0.0;
  if (sp.mode == SFC_IR) {
    // eps = c e, rho = c (1 - e): e_ad = c (eps_ad - rho_ad), folded into the planes
    const double e_ad = eps_ad - rho_ad;
    g[0] = P[IP_DT * S] * e_ad;
    g[1] = 0.0;
    g[2] = P[IP_DU * S] * e_ad;
    g[3] = 0.0;
    return;
  }
  const double ev = P[MP_EV * S], eh = P[MP_EH * S], wv = P[MP_WV * S], wh = P[MP_WH * S];
  const double e_ad[2] = {wv * (eps_ad - z.rm[0] * rho_ad), wh * (eps_ad - z.rm[1] * rho_ad)};
  const double rm_ad[2] = {wv * (1.0 - ev) * rho_ad, wh * (1.0 - eh) * rho_ad};
  for (int j = 0; j < 4; ++j) g[j] = e_ad[0] * P[(MP_DEV + j) * S] + e_ad[1] * P[(MP_DEH + j) * S];
  if (z.rc)
    for (int q = 0; q < 2; ++q) {
      tr_ad += z.drm_dt[q] * rm_ad[q];
      const double* dA = P + (q ? MP_DAH : MP_DAV) * S;
      g[2] += z.drm_dr[q] * (dA[0] + z.x * dA[S] + z.x * z.x * dA[2 * S]) * rm_ad[q];
    }
}

}  // namespace crtm
