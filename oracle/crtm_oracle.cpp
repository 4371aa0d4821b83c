// crtm_oracle.cpp -- FP64 CPU restatement of CRTM's clear-sky per-profile x
// per-channel path: Map_Input / LayerAvg / Geopotential_Height, ODPS
// predictors (groups 1-3) with OPTRAN water-vapour-line predictors, ODPS
// absorption with level-to-user-level path interpolation, Planck, the
// Emission RT solver and brightness temperature; plus the K-matrix.
//
// TEST INFRASTRUCTURE ONLY (the checker, never the product).
//
// The forward restatement is a template over the scalar type so it runs in
// double (forward model) and in std::complex<double> (complex-step
// differentiation: d f/d x = Im f(x + i h)/h exactly, no cancellation).  The
// complex-step Jacobian is independent of the hand-written adjoint in
// oracle_k_matrix and pins it.  The reference's AD quirks are reproduced in
// the complex-step path where the reference AD is not the derivative of its
// forward:
//   * absorber Min/Max clamps are transparent to the AD (derivative passes
//     straight through) and only a final value <= EPSILON kills it
//     (src/AtmAbsorption/ODPS/ODPS_CoordinateMapping.f90:264-270 vs :516-518)
#include "crtm_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstring>
#include <stdexcept>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

typedef std::complex<double> cplx;

constexpr double TSPACE = 2.7253;   // K, src/CRTM_Parameters.f90:173
inline double re(double x) { return x; }
inline double re(const cplx& z) { return z.real(); }
inline double im(double) { return 0.0; }
inline double im(const cplx& z) { return z.imag(); }
template <class R> inline R absr(const R& x) { return re(x) < 0.0 ? R(-x) : x; }

// src/CRTM_Parameters.f90, src/Utility/Profile_Utility/Profile_Utility_Parameters.f90,
// src/Utility/Fundamental_Constants.f90
const double PI = 3.141592653589793238462643383279;
const double D2R = PI / 180.0;
const double EARTH_RADIUS = 6370.0;                 // CRTM_Parameters.f90:184
const double MAX_TRANS_ZENITH = 63.6122;            // CRTM_Parameters.f90:205
const double MW_H2O = 18.01528, MW_DRYAIR = 28.9648;
const double EPS_MW = MW_H2O / MW_DRYAIR;
const double R_DRYAIR = 8.314472 / (MW_DRYAIR * 1.0e-3);
const double G0 = 9.80665;
const double CV = (1.0 / EPS_MW - 1.0) / 1000.0;    // ODPS_CoordinateMapping.f90:50
const double CCH = 0.001 * R_DRYAIR / G0;           // ODPS_CoordinateMapping.f90:52
const double SMALLDIFF = 1.0e-20;
const double MAX_OD = 20.0;                         // ODPS_AtmAbsorption.f90:57
const double LIMIT_EXP = 20.0, LIMIT_LOG = 4.8e+08; // CRTM_Parameters.f90:150-151
const double RECIP_GRAV = 1.0 / 980.665;            // CRTM_Parameters.f90:164
const double MIN_ABSORBER = 1.0e-76;                // CRTM_Parameters.f90:140
const double TOLERANCE = 2.220446049250313e-16;     // EPSILON(1.0_fp)
const int MAXP = 18;                                // max predictors per component
const int MAX_OPTRAN_ORDER = 10;
const int N_OPTRAN_PRED = 14;

const int N_ABS_G[4] = {0, 6, 3, 1};
const int H2O_ID = 1;

// ---------------------------------------------------------------------------
// LayerAvg (ODPS_CoordinateMapping.f90:1108-1320): lnP layer-averaging
// weights of user layers for each ODPS layer.  1-based restatement.
// pz[(ki-1)*kn2 + (j-1)], idx[2*(ki-1) + 0/1] (1-based user layer range)
// ---------------------------------------------------------------------------
void layer_avg(const double* px1, int kn1, const double* px2, int kn2, double* pz, int* idx) {
  auto P1 = [&](int i) { return px1[i - 1]; };
  auto P2 = [&](int i) { return px2[i - 1]; };
  auto PZ = [&](int j, int ki) -> double& { return pz[(size_t)(ki - 1) * kn2 + (j - 1)]; };
  int istart = 1;
  for (int ki = 1; ki <= kn1; ++ki) {
    double z2 = P1(ki);
    double z1 = (ki == 1) ? 2.0 * z2 - P1(ki + 1) : P1(ki - 1);
    double z3 = (ki == kn1) ? 2.0 * z2 - z1 : P1(ki + 1);
    if (z3 > P2(kn2)) z3 = P2(kn2);
    int iskip = 0;
    if (z2 >= P2(kn2)) {
      z3 = P2(kn2);
      z2 = P2(kn2);
      iskip = 1;
    }
    for (int j = 1; j <= kn2; ++j) PZ(j, ki) = 0.0;
    int ic = 0;
    int j;
    bool jumped = false;
    for (j = istart; j <= kn2 - 1; ++j) {
      if (P2(j) > z3) { jumped = true; break; }
      if (P2(j) <= z2 && P2(j + 1) > z1) {
        int itop = 0, ibot = 0;
        double y1, y2;
        if (z1 < z3) {
          y1 = z1; if (P2(j) > z1) { y1 = P2(j); itop = 1; }
          y2 = z2; if (P2(j + 1) < z2) { y2 = P2(j + 1); ibot = 1; }
        } else {
          y1 = z2; if (P2(j) > z2) { y1 = P2(j); itop = 1; }
          y2 = z1; if (P2(j + 1) < z1) { y2 = P2(j + 1); ibot = 1; }
        }
        double dy = y2 - y1, dz = z1 - z2;
        if (std::fabs(dz) < SMALLDIFF) return;
        double dzd = 1.0 / dz;
        double zw1 = (z1 - y1) * dzd * dy;
        double zw2 = (z1 - y2) * dzd * dy;
        double dx = P2(j + 1) - P2(j);
        if (std::fabs(dx) < SMALLDIFF) return;
        double dxd = 1.0 / dx;
        double d = (P2(j + 1) - z2) * dxd;
        if (z1 < z3 && ibot == 0) {
          zw1 = zw1 + zw2 * d;
          zw2 = zw2 * (1.0 - d);
        } else if (z1 > z3 && itop == 0) {
          zw2 = zw2 + zw1 * (1.0 - d);
          zw1 = zw1 * d;
        }
        PZ(j, ki) += zw1;
        PZ(j + 1, ki) += zw2;
        ic = 1;
      }
      if (P2(j) < z3 && P2(j + 1) >= z2 && iskip == 0) {
        int itop = 0, ibot = 0;
        double y1, y2;
        if (z3 < z1) {
          y1 = z3; if (P2(j) > z3) { y1 = P2(j); itop = 1; }
          y2 = z2; if (P2(j + 1) < z2) { y2 = P2(j + 1); ibot = 1; }
        } else {
          y1 = z2; if (P2(j) > z2) { y1 = P2(j); itop = 1; }
          y2 = z3; if (P2(j + 1) < z3) { y2 = P2(j + 1); ibot = 1; }
        }
        double dy = y2 - y1, dz = z3 - z2;
        if (std::fabs(dz) < SMALLDIFF) return;
        if (dz == 0.0) dz = 1.0;
        double dzd = 1.0 / dz;
        double zw1 = (z3 - y1) * dzd * dy;
        double zw2 = (z3 - y2) * dzd * dy;
        double dx = P2(j + 1) - P2(j);
        if (std::fabs(dx) < SMALLDIFF) return;
        double dxd = 1.0 / dx;
        double d = (P2(j + 1) - z2) * dxd;
        if (z3 < z1 && ibot == 0) {
          zw1 = zw1 + zw2 * d;
          zw2 = zw2 * (1.0 - d);
        } else if (z3 > z1 && itop == 0) {
          zw2 = zw2 + zw1 * (1.0 - d);
          zw1 = zw1 * d;
        }
        PZ(j, ki) += zw1;
        PZ(j + 1, ki) += zw2;
        ic = 1;
      }
    }
    if (!jumped) j = kn2;
    if (ic == 0) PZ(j, ki) = 1.0;
    int i1 = istart;
    for (int ii = istart; ii <= kn2; ++ii)
      if (PZ(ii, ki) != 0.0) { i1 = ii; break; }
    istart = i1;
    int i2 = kn2;
    for (int ii = i1 + 1; ii <= kn2; ++ii)
      if (PZ(ii, ki) == 0.0) { i2 = ii - 1; break; }
    double zsum = 0.0;
    for (int ii = i1; ii <= i2; ++ii) zsum += PZ(ii, ki);
    for (int ii = i1; ii <= i2; ++ii) PZ(ii, ki) = PZ(ii, ki) / zsum;
    idx[2 * (ki - 1)] = i1;
    idx[2 * (ki - 1) + 1] = i2;
  }
}

// Compute_Interp_Index (ODPS_CoordinateMapping.f90:963-1012); 1-based results.
void interp_index(const double* x, int nx, const double* u, int nu, int* idx) {
  int k1 = nu + 1;
  for (int iu = 1; iu <= nu; ++iu) {
    if (u[iu - 1] < x[0]) {
      idx[2 * (iu - 1)] = 1;
      idx[2 * (iu - 1) + 1] = 1;
    } else {
      k1 = iu;
      break;
    }
  }
  int k2 = 0;
  for (int iu = nu; iu >= k1; --iu) {
    if (u[iu - 1] > x[nx - 1]) {
      idx[2 * (iu - 1)] = nx;
      idx[2 * (iu - 1) + 1] = nx;
    } else {
      k2 = iu;
      break;
    }
  }
  int j = 1;
  for (int iu = k1; iu <= k2; ++iu) {
    for (int ix = j; ix <= nx - 1; ++ix) {
      if (u[iu - 1] >= x[ix - 1] && u[iu - 1] <= x[ix]) {
        idx[2 * (iu - 1)] = ix;
        idx[2 * (iu - 1) + 1] = ix + 1;
        j = ix;
        break;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Per-profile state on the ODPS grid (Map_Input + predictors)
// ---------------------------------------------------------------------------
template <class R>
struct ProfileState {
  int K = 0, Ko = 0, J = 0, Jm = 0;
  std::vector<double> ulnp_lev, rlnp_lev;  // [K+1], [Ko+1]
  std::vector<double> acc_w;               // [Ko*K]
  std::vector<int> acc_idx;                // [2*Ko]
  std::vector<int> idx_map;                // [Jm] user absorber (0-based) or -1
  int h2o_idx = 1;                         // 1-based ODPS absorber index (Map_Input quirk)
  std::vector<R> t, a;                     // [Ko], [Ko*Jm] (a[j*Ko+k])
  std::vector<double> sec;                 // [Ko]
  double sec_sfc = 1.0, cos_sensor = 1.0;
  std::vector<R> x;                        // [Ko*J*MAXP]
  bool optran = false;
  std::vector<R> ox, ap, da;               // [Ko*14], [Ko*10], [Ko]
  std::vector<int> u_idx;                  // [2*(K+1)] ODPS level brackets of user levels
  // saved intermediates for the adjoint
  std::vector<double> pdp, tz_ref, tzp_ref, gaz_ref, gazp_ref, gatzp_ref;
  std::vector<R> tz, tzp, gazp, gatzp;     // [Ko], [Ko], [Ko*Jm], [Ko*Jm]
  std::vector<R> int_vapor, ave_a, s_t, s_p, inverse, ap1;  // OPTRAN intermediates [Ko]
  // non-LTE correction predictor (NLTE_Predictor_type): mean layer temperatures
  // as predictors, their level ranges (k1, k2], bilinear angle weights
  bool nl_active = false, nl_ctm = true;
  int nl_k1[2] = {0, 0}, nl_k2[2] = {0, 0}, nl_isen = 0, nl_isol = 0;
  R nl_pred[3] = {R(0.0), R(0.0), R(0.0)};
  double nl_w[4] = {0.0, 0.0, 0.0, 0.0};
};

// Compute_NLTE_Predictor (src/NLTE/NLTE_Correction/CRTM_NLTECorrection.f90:95-260)
template <class R>
void nlte_predictor(const or_sensor* s, const or_profile* p, const R* t, double sec_sensor, ProfileState<R>& st) {
  if (p->no_nlte_correction || s->nl_nsen == 0) return;
  const double cosv = std::cos(D2R * p->source_zenith);
  const double sec_src = cosv != 0.0 ? 1.0 / cosv : (cosv < 0.0 ? -1e308 : 1e308);
  if (sec_src < 0.0) return;
  const int K = p->n_layers;
  const double* lp = p->level_pressure;
  int k1 = 1 << 30, k2 = 1 << 30;
  for (int i = 0; i < 2; ++i) {
    R tm;
    if (s->nl_upper[i] < lp[0] || s->nl_lower[i] > lp[K]) {
      tm = R(s->nl_mean_tm[i]);
      st.nl_ctm = false;
    } else {
      for (int k = 1; k <= K; ++k)
        if (lp[k] > s->nl_upper[i]) { k1 = k - 1; break; }
      for (int k = k1; k <= K; ++k)
        if (lp[k] > s->nl_lower[i]) { k2 = k; break; }
      tm = R(0.0);
      for (int k = k1 + 1; k <= k2; ++k) tm = tm + t[k - 1];
      tm = tm / double(k2 - k1);
      st.nl_k1[i] = k1;
      st.nl_k2[i] = k2;
      st.nl_ctm = true;
      if (re(tm) < s->nl_min_tm[i]) { tm = R(s->nl_min_tm[i]); st.nl_ctm = false; }
      else if (re(tm) > s->nl_max_tm[i]) { tm = R(s->nl_max_tm[i]); st.nl_ctm = false; }
    }
    st.nl_pred[1 + i] = tm;
  }
  st.nl_pred[0] = R(1.0);
  const int ns = s->nl_nsen, no = s->nl_nsol;
  const double* X = s->nl_sec_sensor;
  const double* Y = s->nl_sec_solar;
  const double x = std::min(std::max(sec_sensor, X[0]), X[ns - 1]);
  const double y = std::min(std::max(sec_src, Y[0]), Y[no - 1]);
  int isen = 0, isol = 0;
  for (int i = 0; i < ns - 1; ++i) if (x >= X[i] && x <= X[i + 1]) { isen = i; break; }
  for (int i = 0; i < no - 1; ++i) if (y >= Y[i] && y <= Y[i + 1]) { isol = i; break; }
  const double xx1 = (x - X[isen + 1]) / (X[isen] - X[isen + 1]), xx2 = (x - X[isen]) / (X[isen + 1] - X[isen]);
  const double yy1 = (y - Y[isol + 1]) / (Y[isol] - Y[isol + 1]), yy2 = (y - Y[isol]) / (Y[isol + 1] - Y[isol]);
  st.nl_w[0] = xx1 * yy1; st.nl_w[1] = xx1 * yy2; st.nl_w[2] = xx2 * yy1; st.nl_w[3] = xx2 * yy2;
  st.nl_isen = isen;
  st.nl_isol = isol;
  st.nl_active = true;
}

// the channel's NLTE coefficients at the profile's angles (Compute_NLTE_Correction :535-540);
// false when the correction does not apply to channel l
template <class R>
bool nlte_coeffs(const or_sensor* s, const ProfileState<R>& st, int l, double c[3]) {
  if (!st.nl_active || s->nl_c_index[l] <= 0) return false;
  const int ich = s->nl_c_index[l] - 1, ns = s->nl_nsen, no = s->nl_nsol, a = st.nl_isen, b = st.nl_isol;
  auto C = [&](int i, int is, int io) { return s->nl_c[i + 3 * (is + ns * (io + no * ich))]; };
  for (int i = 0; i < 3; ++i)
    c[i] = st.nl_w[0] * C(i, a, b) + st.nl_w[1] * C(i, a, b + 1) + st.nl_w[2] * C(i, a + 1, b) +
           st.nl_w[3] * C(i, a + 1, b + 1);
  return true;
}

// (A_earth + A_platform, A_space) of channel l at the profile's field of view; false without the correction
inline bool antenna_coeffs(const or_sensor* s, const or_profile* p, int l, double& aep, double& as) {
  if (!p->use_antenna_correction || s->ac_nfov == 0 || p->ifov <= 0) return false;
  const size_t i = (size_t)(p->ifov - 1) + (size_t)s->ac_nfov * l;
  aep = s->ac_earth[i] + s->ac_platform[i];
  as = s->ac_space[i];
  return true;
}

inline int XI(int k, int j, int i, int J) { return (k * J + j) * MAXP + i; }

template <class R>
void map_and_predict(const or_sensor* s, const or_profile* p, const R* t_user, const R* a_user,
                     ProfileState<R>& st) {
  const int K = p->n_layers, Ko = s->n_layers, J = s->n_components, Jm = s->n_absorbers;
  st.K = K; st.Ko = Ko; st.J = J; st.Jm = Jm;
  // Geometry (src/GeometryInfo/CRTM_GeometryInfo.f90:100-143)
  double zen_rad = D2R * p->sensor_zenith;
  double sec_sensor = 1.0 / std::cos(zen_rad);
  double trans_rad, sec_trans;
  if (p->sensor_zenith > MAX_TRANS_ZENITH) {
    trans_rad = D2R * MAX_TRANS_ZENITH;
    sec_trans = 1.0 / std::cos(trans_rad);
  } else {
    trans_rad = zen_rad;
    sec_trans = sec_sensor;
  }
  st.sec_sfc = sec_trans;
  nlte_predictor(s, p, t_user, sec_sensor, st);
  st.cos_sensor = 1.0 / sec_sensor;   // Common_RTSolution.f90:224
  // Map_Input (ODPS_CoordinateMapping.f90:151-319)
  std::vector<double> ref_lnp(Ko), user_lnp(K);
  st.ulnp_lev.assign(K + 1, 0.0);
  st.rlnp_lev.assign(Ko + 1, 0.0);
  for (int k = 0; k < Ko; ++k) ref_lnp[k] = std::log(s->ref_pressure[k]);
  for (int k = 0; k < K; ++k) user_lnp[k] = std::log(p->pressure[k]);
  for (int k = 0; k <= Ko; ++k) st.rlnp_lev[k] = std::log(s->ref_level_pressure[k]);
  st.ulnp_lev[0] = (p->level_pressure[0] <= 0.0) ? std::log(p->level_pressure[1] / 2.0)
                                                 : std::log(p->level_pressure[0]);
  for (int k = 1; k <= K; ++k) st.ulnp_lev[k] = std::log(p->level_pressure[k]);
  int sfc_idx = Ko;
  double sfc_frac = 0.0;
  if (s->ref_level_pressure[Ko] > p->level_pressure[K]) {
    for (int k = Ko; k >= 0; --k) {
      if (s->ref_level_pressure[k] < p->level_pressure[K]) {
        sfc_idx = k + 1;
        sfc_frac = (st.rlnp_lev[sfc_idx] - st.ulnp_lev[K]) / (st.rlnp_lev[sfc_idx] - st.rlnp_lev[k]);
        break;
      }
    }
  }
  st.acc_w.assign((size_t)Ko * K, 0.0);
  st.acc_idx.assign(2 * Ko, 1);
  layer_avg(ref_lnp.data(), Ko, user_lnp.data(), K, st.acc_w.data(), st.acc_idx.data());
  st.t.assign(Ko, R(0.0));
  for (int k = 0; k < Ko; ++k) {
    R sum(0.0);
    for (int u = st.acc_idx[2 * k]; u <= st.acc_idx[2 * k + 1]; ++u)
      sum += st.acc_w[(size_t)k * K + (u - 1)] * t_user[u - 1];
    st.t[k] = sum;
  }
  st.a.assign((size_t)Ko * Jm, R(0.0));
  st.idx_map.assign(Jm, -1);
  for (int j = 0; j < Jm; ++j) {
    for (int jj = 0; jj < p->n_absorbers; ++jj)
      if (p->absorber_id[jj] == s->absorber_id[j]) { st.idx_map[j] = jj; break; }
    st.h2o_idx = 1;                                 // reset every pass, as the reference does
    if (s->absorber_id[j] == H2O_ID) st.h2o_idx = j + 1;
    if (st.idx_map[j] >= 0) {
      const R* au = a_user + (size_t)st.idx_map[j] * K;
      for (int k = 0; k < Ko; ++k) {
        R sum(0.0);
        for (int u = st.acc_idx[2 * k]; u <= st.acc_idx[2 * k + 1]; ++u)
          sum += st.acc_w[(size_t)k * K + (u - 1)] * au[u - 1];
        double v = re(sum), v2 = v;
        double mn = s->min_absorber[(size_t)j * Ko + k], mx = s->max_absorber[(size_t)j * Ko + k];
        if (v2 <= TOLERANCE) v2 = TOLERANCE;
        if (v2 < mn) v2 = mn;
        if (v2 > mx) v2 = mx;
        R val = sum + (v2 - v);          // value clamped, derivative straight through
        if (v2 <= TOLERANCE) val = R(v2);
        st.a[(size_t)j * Ko + k] = val;
      }
    } else {
      for (int k = 0; k < Ko; ++k) st.a[(size_t)j * Ko + k] = R(s->ref_absorber[(size_t)j * Ko + k]);
    }
  }
  // Geopotential_Height of the reference profile (ODPS_CoordinateMapping.f90:607-634)
  std::vector<double> z(Ko + 1);
  const double* wv = s->ref_absorber + (size_t)(st.h2o_idx - 1) * Ko;
  z[Ko] = 0.0;
  for (int k = Ko; k >= 1; --k) {
    double tv = s->ref_temperature[k - 1] * (1.0 + CV * wv[k - 1]);
    double h = CCH * tv;
    z[k - 1] = z[k] + h * std::log(s->ref_level_pressure[k] / s->ref_level_pressure[k - 1]);
  }
  double z_off;
  if (s->ref_level_pressure[Ko] >= p->level_pressure[K])
    z_off = -(z[sfc_idx] + sfc_frac * (z[sfc_idx - 1] - z[sfc_idx])) + p->surface_altitude;
  else
    z_off = CCH * s->ref_temperature[Ko - 1] * std::log(1013.0 / s->ref_level_pressure[Ko]);
  for (int k = 0; k <= Ko; ++k) z[k] += z_off;
  double sv = (EARTH_RADIUS + p->surface_altitude) * std::sin(trans_rad);
  st.sec.assign(Ko, 0.0);
  for (int k = 1; k <= Ko; ++k) {
    double sa = sv / (EARTH_RADIUS + z[k]);
    st.sec[k - 1] = 1.0 / std::sqrt(1.0 - sa * sa);
  }
  // ODPS_Compute_Predictor (ODPS_Predictor.f90:645-759): integrated ratios
  const int g = s->group, na = N_ABS_G[g];
  const double* rt = s->ref_temperature;
  const double* ra = s->ref_absorber;
  const double* rlp = s->ref_level_pressure;
  st.pdp.assign(Ko, 0.0); st.tz_ref.assign(Ko, 0.0); st.tzp_ref.assign(Ko, 0.0);
  st.gaz_ref.assign((size_t)Ko * Jm, 0.0); st.gazp_ref.assign((size_t)Ko * Jm, 0.0);
  st.gatzp_ref.assign((size_t)Ko * Jm, 0.0);
  st.tz.assign(Ko, R(0.0)); st.tzp.assign(Ko, R(0.0));
  st.gazp.assign((size_t)Ko * Jm, R(0.0)); st.gatzp.assign((size_t)Ko * Jm, R(0.0));
  {
    double tz_ref = 0.0, tzp_ref = 0.0;
    R tz_sum(0.0), tzp_sum(0.0);
    std::vector<double> gaz_ref(na, 0.0), gazp_ref(na, 0.0), gatzp_ref(na, 0.0);
    std::vector<R> gaz_sum(na, R(0.0)), gazp_sum(na, R(0.0)), gatzp_sum(na, R(0.0));
    for (int k = 1; k <= Ko; ++k) {
      double pdp = (k == 1) ? rlp[0] * (rlp[1] - rlp[0]) : rlp[k] * (rlp[k] - rlp[k - 1]);
      st.pdp[k - 1] = pdp;
      tz_ref += rt[k - 1];
      tz_sum += st.t[k - 1];
      st.tz[k - 1] = tz_sum / tz_ref;
      tzp_ref += pdp * rt[k - 1];
      tzp_sum += pdp * st.t[k - 1];
      st.tzp[k - 1] = tzp_sum / tzp_ref;
      st.tz_ref[k - 1] = tz_ref;
      st.tzp_ref[k - 1] = tzp_ref;
      for (int j = 0; j < na; ++j) {
        double rak = ra[(size_t)j * Ko + k - 1];
        const R& ak = st.a[(size_t)j * Ko + k - 1];
        gaz_ref[j] += rak;
        gaz_sum[j] += ak;
        gazp_ref[j] += pdp * rak;
        gazp_sum[j] += pdp * ak;
        st.gazp[(size_t)j * Ko + k - 1] = gazp_sum[j] / gazp_ref[j];
        gatzp_ref[j] += pdp * rak * rt[k - 1];
        gatzp_sum[j] += pdp * ak * st.t[k - 1];
        st.gatzp[(size_t)j * Ko + k - 1] = gatzp_sum[j] / gatzp_ref[j];
        st.gaz_ref[(size_t)j * Ko + k - 1] = gaz_ref[j];
        st.gazp_ref[(size_t)j * Ko + k - 1] = gazp_ref[j];
        st.gatzp_ref[(size_t)j * Ko + k - 1] = gatzp_ref[j];
      }
    }
  }
  st.x.assign((size_t)Ko * J * MAXP, R(0.0));
  auto X = [&](int k, int i, int j) -> R& { return st.x[XI(k, j - 1, i - 1, J)]; };
  for (int k = 0; k < Ko; ++k) {
    R sec(st.sec[k]);
    R dt = st.t[k] - rt[k];
    R T = st.t[k] / rt[k];
    R T2 = T * T;
    R dt2 = dt * absr(dt);
    auto A = [&](int j) { return st.a[(size_t)j * Ko + k] / ra[(size_t)j * Ko + k]; };
    auto GAzp = [&](int j) { return st.gazp[(size_t)j * Ko + k]; };
    auto GATzp = [&](int j) { return st.gatzp[(size_t)j * Ko + k]; };
    R h2o = A(0);
    R h2o_a = sec * h2o;
    R h2o_r = std::sqrt(h2o_a);
    R h2o_s = h2o_a * h2o_a;
    R h2o_r4 = std::sqrt(h2o_r);
    R h2o_dtzp = h2o / GATzp(0);
    if (g == 1 || g == 2) {
      // ODPS_Compute_Predictor_IR (ODPS_Predictor.f90:765-1024)
      R o3 = A(1), co2 = A(2);
      R o3_a = sec * o3, o3_r = std::sqrt(o3_a);
      X(k, 1, 1) = sec; X(k, 2, 1) = sec * T; X(k, 3, 1) = sec * T2; X(k, 4, 1) = T;
      X(k, 5, 1) = sec * sec; X(k, 6, 1) = T2; X(k, 7, 1) = st.tz[k];
      X(k, 1, 3) = h2o_a / T; X(k, 2, 3) = h2o_a / T * h2o; X(k, 3, 3) = h2o_a / T2 * h2o / T2;
      X(k, 4, 3) = h2o_a / T2; X(k, 5, 3) = h2o_a / T2 * h2o; X(k, 6, 3) = h2o_a / (T2 * T2);
      X(k, 7, 3) = h2o_a;
      R o3p = GAzp(1);
      X(k, 1, 4) = o3_a; X(k, 2, 4) = o3_a * dt; X(k, 3, 4) = o3_a * o3 * o3p; X(k, 4, 4) = o3_a * o3_a;
      X(k, 5, 4) = o3_a * o3p; X(k, 6, 4) = o3_a * std::sqrt(sec * o3p); X(k, 7, 4) = o3_r * dt;
      X(k, 8, 4) = o3_r; X(k, 9, 4) = o3_r * o3 / o3p; X(k, 10, 4) = sec * o3p;
      X(k, 11, 4) = (sec * o3p) * (sec * o3p);
      R tzp = st.tzp[k];
      X(k, 1, 5) = sec * T; X(k, 2, 5) = sec * T2; X(k, 3, 5) = T; X(k, 4, 5) = T2; X(k, 5, 5) = sec;
      X(k, 6, 5) = sec * co2; X(k, 7, 5) = sec * tzp; X(k, 8, 5) = (sec * GAzp(2)) * (sec * GAzp(2));
      X(k, 9, 5) = tzp * tzp * tzp; X(k, 10, 5) = sec * tzp * std::sqrt(T);
      X(k, 1, 2) = h2o_a; X(k, 2, 2) = h2o_a * dt; X(k, 3, 2) = h2o_s; X(k, 4, 2) = h2o_a * dt2;
      X(k, 5, 2) = h2o_r4; X(k, 6, 2) = h2o_s * h2o_a; X(k, 7, 2) = h2o_r; X(k, 8, 2) = h2o_r * dt;
      X(k, 9, 2) = h2o_s * h2o_s; X(k, 10, 2) = h2o_dtzp; X(k, 11, 2) = h2o_r * h2o_dtzp;
      X(k, 12, 2) = (sec * GAzp(0)) * (sec * GAzp(0)); X(k, 13, 2) = sec * GAzp(0); X(k, 14, 2) = sec;
      X(k, 15, 2) = sec * co2;
      if (g == 1) {
        R co = A(4), n2o = A(3), ch4 = A(5);
        R n2o_a = sec * n2o, n2o_r = std::sqrt(n2o_a), n2o_s = n2o_a * n2o_a;
        R co_a = sec * co, co_r = std::sqrt(co_a), co_s = co_a * co_a;
        R co_acodcozp = co_a * co / GAzp(4);
        R ch4_a = sec * ch4, ch4_r = std::sqrt(ch4_a);
        R ch4_ach4zp = sec * GAzp(5);
        X(k, 11, 5) = co_a;
        X(k, 16, 2) = ch4_a; X(k, 17, 2) = ch4_a * ch4_a * dt; X(k, 18, 2) = co_a;
        X(k, 1, 7) = co_a; X(k, 2, 7) = co_a * dt; X(k, 3, 7) = std::sqrt(co_r); X(k, 4, 7) = co_r * dt;
        X(k, 5, 7) = co_s; X(k, 6, 7) = co_r; X(k, 7, 7) = co_a * dt2; X(k, 8, 7) = co_acodcozp;
        X(k, 9, 7) = co_acodcozp / co_r; X(k, 10, 7) = co_acodcozp * std::sqrt(GAzp(4));
        X(k, 1, 8) = ch4_a * dt; X(k, 2, 8) = ch4_r; X(k, 3, 8) = ch4_a * ch4_a; X(k, 4, 8) = ch4_a;
        X(k, 5, 8) = ch4 * dt; X(k, 6, 8) = ch4_ach4zp; X(k, 7, 8) = ch4_ach4zp * ch4_ach4zp;
        X(k, 8, 8) = std::sqrt(ch4_r); X(k, 9, 8) = GATzp(5); X(k, 10, 8) = sec * GATzp(5);
        X(k, 11, 8) = ch4_r * ch4 / GAzp(5);
        X(k, 1, 6) = n2o_a * dt; X(k, 2, 6) = n2o_r; X(k, 3, 6) = n2o * dt;
        X(k, 4, 6) = std::pow(n2o_a, 0.25); X(k, 5, 6) = n2o_a; X(k, 6, 6) = sec * GAzp(3);
        X(k, 7, 6) = sec * GATzp(3); X(k, 8, 6) = n2o_s; X(k, 9, 6) = GATzp(3);
        X(k, 10, 6) = n2o_r * n2o / GAzp(3); X(k, 11, 6) = ch4_a; X(k, 12, 6) = ch4_a * GAzp(5);
        X(k, 13, 6) = co_a; X(k, 14, 6) = co_a * sec * GAzp(4);
      }
    } else {
      // ODPS_Compute_Predictor_MW (ODPS_Predictor.f90:1026-1091)
      X(k, 1, 1) = sec; X(k, 2, 1) = sec * T; X(k, 3, 1) = sec * T2; X(k, 4, 1) = T;
      X(k, 5, 1) = sec * sec; X(k, 6, 1) = T2; X(k, 7, 1) = st.tz[k];
      X(k, 1, 2) = h2o_a / T; X(k, 2, 2) = h2o_a / T * h2o; X(k, 3, 2) = h2o_a / T2 * h2o / T2;
      X(k, 4, 2) = h2o_a / T2; X(k, 5, 2) = h2o_a / T2 * h2o; X(k, 6, 2) = h2o_a / (T2 * T2);
      X(k, 7, 2) = h2o_a; X(k, 8, 2) = h2o_a * dt; X(k, 9, 2) = (sec * GAzp(0)) * (sec * GAzp(0));
      X(k, 10, 2) = sec * GAzp(0); X(k, 11, 2) = sec; X(k, 12, 2) = h2o_s * h2o_a;
      X(k, 13, 2) = h2o_s * h2o_s; X(k, 14, 2) = h2o_dtzp;
    }
  }
  // ODPS_Compute_Predictor_ODAS (ODPS_Predictor.f90:2716-2816)
  st.optran = s->n_ocoeffs > 0;
  if (st.optran) {
    st.ox.assign((size_t)Ko * N_OPTRAN_PRED, R(0.0));
    st.ap.assign((size_t)Ko * MAX_OPTRAN_ORDER, R(0.0));
    st.da.assign(Ko, R(0.0));
    st.int_vapor.assign(Ko, R(0.0)); st.ave_a.assign(Ko, R(0.0)); st.s_t.assign(Ko, R(0.0));
    st.s_p.assign(Ko, R(0.0)); st.inverse.assign(Ko, R(0.0)); st.ap1.assign(Ko, R(0.0));
    const R* vap = &st.a[(size_t)(st.h2o_idx - 1) * Ko];
    const double* pr = s->ref_pressure;
    auto OX = [&](int k, int i) -> R& { return st.ox[(size_t)(k - 1) * N_OPTRAN_PRED + (i - 1)]; };
    for (int k = 1; k <= Ko; ++k) {
      R tk = st.t[k - 1];
      R t2 = tk * tk;
      double pk = pr[k - 1], p2 = pk * pk;
      OX(k, 1) = tk; OX(k, 2) = R(pk); OX(k, 3) = t2; OX(k, 4) = R(p2); OX(k, 5) = tk * pk;
      OX(k, 6) = t2 * pk; OX(k, 7) = tk * p2; OX(k, 8) = t2 * p2; OX(k, 9) = R(std::pow(pk, 0.25));
      OX(k, 10) = vap[k - 1]; OX(k, 11) = vap[k - 1] / t2; OX(k, 14) = R(st.sec[k - 1]);
    }
    R iv_prev(0.0), s_t(0.0), s_p(0.0), xlt_prev(0.0), xlp_prev(0.0);
    for (int k = 1; k <= Ko; ++k) {
      double dpong = RECIP_GRAV * (rlp[k] - rlp[k - 1]);
      R dab = dpong * vap[k - 1] * st.sec[k - 1];
      R iv = iv_prev + dab;
      R ave = 0.5 * (iv_prev + iv);
      st.da[k - 1] = dab;
      s_t += st.t[k - 1] * dab;
      s_p += pr[k - 1] * dab;
      R inv = (re(iv) > MIN_ABSORBER) ? R(1.0) / iv : R(0.0);
      R xlt = 0.5 * s_t * inv, xlp = 0.5 * s_p * inv;
      OX(k, 12) = xlt + xlt_prev;
      OX(k, 13) = xlp + xlp_prev;
      R ap1 = std::log((ave - s->alpha_c2) / s->alpha_c1) / s->alpha;
      st.ap[(size_t)(k - 1) * MAX_OPTRAN_ORDER] = ap1;
      for (int i = 2; i <= MAX_OPTRAN_ORDER; ++i)
        st.ap[(size_t)(k - 1) * MAX_OPTRAN_ORDER + i - 1] = st.ap[(size_t)(k - 1) * MAX_OPTRAN_ORDER + i - 2] * ap1;
      st.int_vapor[k - 1] = iv; st.ave_a[k - 1] = ave; st.s_t[k - 1] = s_t; st.s_p[k - 1] = s_p;
      st.inverse[k - 1] = inv; st.ap1[k - 1] = ap1;
      iv_prev = iv; xlt_prev = xlt; xlp_prev = xlp;
    }
  }
  // Compute_Interp_Index(Ref_Level_LnPressure, User_Level_LnPressure)
  st.u_idx.assign(2 * (K + 1), 1);
  interp_index(st.rlnp_lev.data(), Ko + 1, st.ulnp_lev.data(), K + 1, st.u_idx.data());
}

// ---------------------------------------------------------------------------
// Per-channel forward: absorption + Emission RT + BT
// ---------------------------------------------------------------------------
template <class R>
struct ChannelFwd {
  std::vector<R> od, path, upath, tau, B, td, tu, rd, ru;
  R bs, radiance, bt, up_radiance;
  R eps, rho, trans;           // surface emissivity, reflectivity, column transmittance
  std::vector<R> ln_chi, chi;  // OPTRAN
  std::vector<R> b;            // OPTRAN b(k, 0:np)
  bool optran_ch = false;
};

template <class R>
void optran_od(const or_sensor* s, int l, const ProfileState<R>& st, ChannelFwd<R>& cf) {
  const int Ko = st.Ko;
  int np = s->op_index[7 * l];
  if (np <= 0) return;
  int js = s->opos_index[l];
  int no = s->order[l];
  cf.b.assign((size_t)Ko * (np + 1), R(0.0));
  cf.ln_chi.assign(Ko, R(0.0));
  cf.chi.assign(Ko, R(0.0));
  for (int i = 0; i <= np; ++i) {
    int jj = js + i * (no + 1);
    const double* coeff = s->oc + (jj - 1);
    for (int k = 0; k < Ko; ++k) {
      R bb(coeff[0]);
      for (int j = 1; j <= no; ++j) bb += coeff[j] * st.ap[(size_t)k * MAX_OPTRAN_ORDER + j - 1];
      cf.b[(size_t)k * (np + 1) + i] = bb;
    }
  }
  for (int k = 0; k < Ko; ++k) {
    R lc = cf.b[(size_t)k * (np + 1)];
    for (int i = 1; i <= np; ++i) {
      int ii = s->op_index[7 * l + i];
      lc += cf.b[(size_t)k * (np + 1) + i] * st.ox[(size_t)k * N_OPTRAN_PRED + ii - 1];
    }
    cf.ln_chi[k] = lc;
    R c;
    if (re(lc) > LIMIT_EXP) c = R(LIMIT_LOG);
    else if (re(lc) < -LIMIT_EXP) c = R(0.0);
    else c = std::exp(lc);
    cf.chi[k] = c;
    cf.od[k] += c * st.da[k];
  }
}

#include "crtm_oracle_sfc.inc"

// Surface inputs of computed optics for one channel (the optics need the
// column transmittance, known only inside channel_forward)
template <class R>
struct SfcArgs {
  SfcPoint sp;
  R tw, sal, wind, az;
};

template <class R>
void channel_forward(const or_sensor* s, const or_profile* p, const R* t_user, const R& tsfc,
                     const R& eps, int l, const ProfileState<R>& st, ChannelFwd<R>& cf,
                     const SfcArgs<R>* sa = nullptr) {
  const int K = st.K, Ko = st.Ko, J = st.J;
  // ODPS_Compute_AtmAbsorption (ODPS_AtmAbsorption.f90:133-256)
  cf.od.assign(Ko, R(0.0));
  cf.optran_ch = false;
  for (int j = 0; j < J; ++j) {
    int np = s->n_predictors[j + J * l];
    if (np <= 0) continue;
    bool optran = st.optran && (j + 1 == s->ocomponent_index) && s->osignificance[l] == 1;
    if (optran) {
      cf.optran_ch = true;
      optran_od(s, l, st, cf);
    } else {
      int j0 = s->pos_index[j + J * l];
      for (int i = 1; i <= np; ++i) {
        int js = j0 + (i - 1) * Ko - 1;
        for (int k = 1; k <= Ko; ++k)
          cf.od[k - 1] += (double)s->c[js + k - 1] * st.x[XI(k - 1, j, i - 1, J)];
      }
    }
  }
  cf.path.assign(Ko + 1, R(0.0));
  for (int k = 1; k <= Ko; ++k) {
    R odt = cf.od[k - 1];
    if (re(odt) < 0.0) odt = R(0.0);
    else if (re(odt) > MAX_OD) odt = R(MAX_OD);
    cf.path[k] = cf.path[k - 1] + odt;
  }
  // Interpolate_Profile (ODPS_CoordinateMapping.f90:673-690) to user levels
  cf.upath.assign(K + 1, R(0.0));
  for (int i = 0; i <= K; ++i) {
    int k1 = st.u_idx[2 * i], k2 = st.u_idx[2 * i + 1];
    if (k1 == k2) {
      cf.upath[i] = cf.path[k1 - 1];
    } else {
      double x1 = st.rlnp_lev[k1 - 1], x2 = st.rlnp_lev[k2 - 1], u = st.ulnp_lev[i];
      cf.upath[i] = cf.path[k1 - 1] + (cf.path[k2 - 1] - cf.path[k1 - 1]) * (u - x1) / (x2 - x1);
    }
  }
  cf.tau.assign(K, R(0.0));
  for (int k = 1; k <= K; ++k) cf.tau[k - 1] = (cf.upath[k] - cf.upath[k - 1]) / st.sec_sfc;
  // Planck (src/Source_Functions/CRTM_Planck_Functions.f90:116-136)
  double c1 = s->planck_c1[l], c2 = s->planck_c2[l], bc1 = s->band_c1[l], bc2 = s->band_c2[l];
  cf.B.assign(K, R(0.0));
  for (int k = 0; k < K; ++k) cf.B[k] = c1 / (std::exp(c2 / (bc1 + bc2 * t_user[k])) - 1.0);
  cf.bs = c1 / (std::exp(c2 / (bc1 + bc2 * tsfc)) - 1.0);
  // CRTM_Emission (src/RTSolution/Emission/Emission_Module.f90:54-174), specular surface
  double u = st.cos_sensor;
  double sec_down = 1.0 / u;
  cf.td.assign(K, R(0.0)); cf.tu.assign(K, R(0.0));
  cf.rd.assign(K + 1, R(0.0)); cf.ru.assign(K + 1, R(0.0));
  cf.rd[0] = R(s->cosmic_background[l]);
  for (int k = 1; k <= K; ++k) {
    cf.td[k - 1] = std::exp(-cf.tau[k - 1] * sec_down);
    cf.rd[k] = cf.rd[k - 1] * cf.td[k - 1] + cf.B[k - 1] * (1.0 - cf.td[k - 1]);
    cf.tu[k - 1] = std::exp(-cf.tau[k - 1] / u);
  }
  if (sa) {
    // CRTM_Compute_Transmittance (CRTM_AtmOptics.f90:210-218) feeds the
    // FASTEM reflection correction
    R od_sum(0.0);
    for (int k = 0; k < K; ++k) od_sum = od_sum + cf.tau[k];
    cf.trans = std::exp(-od_sum);
    sfc_eval(sa->sp, sa->tw, sa->sal, sa->wind, sa->az, cf.trans, cf.eps, cf.rho);
  } else {
    cf.eps = eps;
    cf.rho = 1.0 - eps;
  }
  cf.ru[K] = cf.eps * cf.bs + cf.rho * cf.rd[K];
  cf.up_radiance = R(0.0);
  for (int k = K; k >= 1; --k) {
    R src = cf.B[k - 1] * (1.0 - cf.tu[k - 1]);
    cf.ru[k - 1] = cf.ru[k] * cf.tu[k - 1] + src;
    cf.up_radiance = cf.up_radiance * cf.tu[k - 1] + src;
  }
  cf.radiance = cf.ru[0];
  // Post_Process_RTSolution (src/CRTM_Forward_Module.f90:1166-1199): the
  // non-LTE correction of the radiance, the Planck temperature
  // (CRTM_Planck_Functions.f90:470-492), the antenna correction
  // (CRTM_Compute_AntCorr, src/AntennaCorrection/CRTM_AntennaCorrection.f90:118-140)
  double nc[3];
  if (nlte_coeffs(s, st, l, nc))
    for (int i = 0; i < 3; ++i) cf.radiance = cf.radiance + nc[i] * st.nl_pred[i];
  cf.bt = (c2 / std::log(c1 / cf.radiance + 1.0) - bc1) / bc2;
  double aep, as;
  if (antenna_coeffs(s, p, l, aep, as)) cf.bt = aep * cf.bt + as * TSPACE;
}

double sfc_temperature(const or_profile* p) {
  // CRTM_Compute_SurfaceT (src/SfcOptics/CRTM_SfcOptics.f90:204-217)
  return p->coverage[0] * p->sfc_temperature[0] + p->coverage[1] * p->sfc_temperature[1] +
         p->coverage[2] * p->sfc_temperature[2] + p->coverage[3] * p->sfc_temperature[3];
}

// The computed-optics inputs of profile p, channel l (SfcOptics%Angle of the
// sensor, src/RTSolution/Common_RTSolution.f90:290-352).  False for an
// unknown polarisation flag.
template <class R>
bool make_sfc_args(const or_sensor* s, const or_profile* p, const or_surface_optics* so, int l, double cos_sensor,
                   SfcArgs<R>& sa) {
  SfcPoint& sp = sa.sp;
  sp.so = so;
  sp.microwave = s->sensor_type == 1;
  sp.f = sp.microwave ? s->frequency[l] : s->wavenumber[l];
  sp.angle = std::acos(cos_sensor) / D2R;
  sp.coverage = p->coverage[1];
  sa.tw = R(p->sfc_temperature[1]);
  sa.sal = R(so->salinity);
  sa.wind = R(so->wind_speed);
  sa.az = R(so->wind_direction - so->sensor_azimuth);
  if (!sp.microwave) return true;
  return pol_weights(so->polarization[l], sp.angle, distance_ratio(p->sensor_zenith, p->sensor_scan),
                     so->pol_angle[l], D2R * p->sensor_scan, sp.wv, sp.wh);
}

// Radiance adjoint of BT (CRTM_Planck_Temperature_AD, CRTM_Planck_Functions.f90:697-723)
double planck_t_ad_factor(const or_sensor* s, int l, double radiance) {
  double c1 = s->planck_c1[l], c2 = s->planck_c2[l], bc2 = s->band_c2[l];
  double arg = c1 / radiance + 1.0;
  double lg = std::log(arg);
  return c1 * c2 / (bc2 * arg * (radiance * lg) * (radiance * lg));
}

// dB/dT (CRTM_Planck_Radiance_AD, CRTM_Planck_Functions.f90:358-388)
double planck_r_ad_factor(const or_sensor* s, int l, double t) {
  double c1 = s->planck_c1[l], c2 = s->planck_c2[l], bc1 = s->band_c1[l], bc2 = s->band_c2[l];
  double te = bc1 + bc2 * t;
  double ex = std::exp(c2 / te);
  double d = te * (ex - 1.0);
  return c1 * c2 * ex * bc2 / (d * d);
}

}  // namespace

// ===========================================================================
// CRTM_Atmosphere_AddLayers (src/Atmosphere/CRTM_Atmosphere.f90:248-400)
// ===========================================================================
namespace {
const double TOA_PRESSURE = 0.005;             // src/CRTM_Parameters.f90:157
const double MINIMUM_ABSORBER_AMOUNT = 1.0e-76; // src/CRTM_Parameters.f90:140
const int N_CLIM_LEVELS = 101;
struct Climatology {
  bool set = false;
  double p[N_CLIM_LEVELS];
  double t[6][N_CLIM_LEVELS];
  double a[6][6][N_CLIM_LEVELS];   // [model][absorber id - 1][level]
} g_clim;

// Profile extended to TOA_PRESSURE with climatology layers on top; owns its arrays.
struct ExtProfile {
  or_profile p;
  int n = 0;                       // layers added on top
  std::vector<double> lp, pr, t, a;
};

// Extra_Layers: first model level (counting from index 1 of the 0-based
// level array, as the reference's loop does) at or below the profile top.
int extra_layers(double ptop) {
  for (int n = 1; n < N_CLIM_LEVELS; ++n)
    if (g_clim.p[n] >= ptop) return n;
  return 0;
}

void extend_profile(const or_profile* in, ExtProfile& e) {
  e.p = *in;
  e.n = 0;
  const int K = in->n_layers, na = in->n_absorbers;
  if (in->level_pressure[0] <= TOA_PRESSURE) return;
  if (!g_clim.set) throw std::runtime_error("oracle_set_climatology not called");
  const int n = extra_layers(in->level_pressure[0]);
  if (n < 1) throw std::runtime_error("Error determining extra layer count");
  const int model = (in->climatology >= 1 && in->climatology <= 6) ? in->climatology - 1 : 5;
  // model levels 0..n, level n replaced by the profile top (Interp_LPoly / Interp_Linear)
  std::vector<double> pl(n + 1), tl(n + 1), al((size_t)na * (n + 1));
  for (int k = 0; k <= n; ++k) {
    pl[k] = g_clim.p[k];
    tl[k] = g_clim.t[model][k];
    for (int j = 0; j < na; ++j) al[(size_t)j * (n + 1) + k] = g_clim.a[model][in->absorber_id[j] - 1][k];
  }
  const double ptop = in->level_pressure[0];
  const double lpoly = (std::log(ptop) - std::log(pl[n - 1])) / (std::log(pl[n]) - std::log(pl[n - 1]));
  pl[n] = ptop;
  tl[n] = (tl[n] - tl[n - 1]) * lpoly + tl[n - 1];
  for (int j = 0; j < na; ++j) {
    double* x = &al[(size_t)j * (n + 1)];
    x[n] = (x[n] - x[n - 1]) * lpoly + x[n - 1];
  }
  // layer means (Layer_P: log-mean pressure; Layer_X: arithmetic mean)
  std::vector<double> pn(n), tn(n), an((size_t)na * n);
  for (int k = 1; k <= n; ++k) {
    pn[k - 1] = (pl[k] - pl[k - 1]) / std::log(pl[k] / pl[k - 1]);
    tn[k - 1] = 0.5 * (tl[k - 1] + tl[k]);
    for (int j = 0; j < na; ++j) an[(size_t)j * n + k - 1] = 0.5 * (al[(size_t)j * (n + 1) + k - 1] + al[(size_t)j * (n + 1) + k]);
  }
  // Shift_Profile: move the added layers so the lowest one continues the
  // profile's top two layers (extrapolated in log pressure)
  const double elpoly = (std::log(pn[n - 1]) - std::log(in->pressure[0])) / (std::log(in->pressure[1]) - std::log(in->pressure[0]));
  auto shift = [&](double* x, const double* xtoa) {
    const double xint = (xtoa[1] - xtoa[0]) * elpoly + xtoa[0];
    const double dx = xint - x[n - 1];
    for (int k = 0; k < n; ++k) x[k] += dx;
  };
  shift(tn.data(), in->temperature);
  for (int j = 0; j < na; ++j) {
    shift(&an[(size_t)j * n], in->absorber + (size_t)j * K);
    for (int k = 0; k < n; ++k)
      if (an[(size_t)j * n + k] < 0.0) an[(size_t)j * n + k] = MINIMUM_ABSORBER_AMOUNT;
  }
  const int Kt = K + n;
  e.n = n;
  e.lp.resize(Kt + 1);
  e.pr.resize(Kt);
  e.t.resize(Kt);
  e.a.resize((size_t)na * Kt);
  for (int k = 0; k <= n; ++k) e.lp[k] = pl[k];
  for (int k = 1; k <= K; ++k) e.lp[n + k] = in->level_pressure[k];
  for (int k = 0; k < n; ++k) {
    e.pr[k] = pn[k];
    e.t[k] = tn[k];
  }
  for (int k = 0; k < K; ++k) {
    e.pr[n + k] = in->pressure[k];
    e.t[n + k] = in->temperature[k];
  }
  for (int j = 0; j < na; ++j) {
    for (int k = 0; k < n; ++k) e.a[(size_t)j * Kt + k] = an[(size_t)j * n + k];
    for (int k = 0; k < K; ++k) e.a[(size_t)j * Kt + n + k] = in->absorber[(size_t)j * K + k];
  }
  e.p.n_layers = Kt;
  e.p.level_pressure = e.lp.data();
  e.p.pressure = e.pr.data();
  e.p.temperature = e.t.data();
  e.p.absorber = e.a.data();
}

// Outputs on the extended profile, then the user layers' part copied out
// (CRTM_Atmosphere_AddLayers_AD, CRTM_Atmosphere.f90:569-659, keeps only the
// sensitivities of the original layers).
struct ExtOut {
  or_out o;
  std::vector<double> tau, t_k, abs_k;
  void init(const or_out* user, int L, int Kt, int na) {
    o = *user;
    if (user->tau) { tau.assign((size_t)L * Kt, 0.0); o.tau = tau.data(); }
    if (user->t_k) { t_k.assign((size_t)L * Kt, 0.0); o.t_k = t_k.data(); }
    if (user->abs_k) { abs_k.assign((size_t)L * na * Kt, 0.0); o.abs_k = abs_k.data(); }
  }
  void copy_back(const or_out* user, int L, int K, int n, int na) const {
    const int Kt = K + n;
    for (int l = 0; l < L; ++l) {
      if (user->tau) std::memcpy(user->tau + (size_t)l * K, &tau[(size_t)l * Kt + n], sizeof(double) * K);
      if (user->t_k) std::memcpy(user->t_k + (size_t)l * K, &t_k[(size_t)l * Kt + n], sizeof(double) * K);
      if (user->abs_k)
        for (int j = 0; j < na; ++j)
          std::memcpy(user->abs_k + ((size_t)l * na + j) * K, &abs_k[((size_t)l * na + j) * Kt + n], sizeof(double) * K);
    }
  }
};
}  // namespace

// ===========================================================================
// C API
// ===========================================================================
static int forward_ext(const or_sensor* s, const or_profile* p, const double* emissivity, or_out* o,
                       const or_surface_optics* so = nullptr, double* emis = nullptr, double* refl = nullptr) {
  const int K = p->n_layers, L = s->n_channels;
  ProfileState<double> st;
  map_and_predict<double>(s, p, p->temperature, p->absorber, st);
  double tsfc = sfc_temperature(p);
  ChannelFwd<double> cf;
  for (int l = 0; l < L; ++l) {
    SfcArgs<double> sa;
    if (so && !make_sfc_args(s, p, so, l, st.cos_sensor, sa)) return 3;
    channel_forward<double>(s, p, p->temperature, tsfc, so ? 0.0 : emissivity[l], l, st, cf, so ? &sa : nullptr);
    if (emis) emis[l] = cf.eps;
    if (refl) refl[l] = cf.rho;
    o->radiance[l] = cf.radiance;
    o->bt[l] = cf.bt;
    if (o->tau) std::memcpy(o->tau + (size_t)l * K, cf.tau.data(), sizeof(double) * K);
    if (o->up_radiance) o->up_radiance[l] = cf.up_radiance;
    if (o->down_radiance) o->down_radiance[l] = cf.rd[K];
  }
  return 0;
}

static int cstep_ext(const or_sensor* s, const or_profile* p, const double* emissivity, const double* bt_k,
                     const double* rad_k, or_out* o, const or_surface_optics* so = nullptr,
                     double* sfc_k = nullptr) {
  const int K = p->n_layers, L = s->n_channels, na = p->n_absorbers;
  if (int r = forward_ext(s, p, emissivity, o, so)) return r;
  const double h = 1.0e-30;
  // directions: layer temperatures, absorbers, surface temperature, then
  // either the user emissivities or (computed optics) salinity, wind speed
  // and wind direction; the surface temperature direction perturbs the
  // water temperature of the optics too
  const int nd0 = K * (1 + na) + 1;
  const int ndir = nd0 + (so ? 3 : 1);
  std::vector<cplx> tu(K), au((size_t)na * K), eps(L);
  for (int d = 0; d < ndir; ++d) {
    for (int k = 0; k < K; ++k) tu[k] = cplx(p->temperature[k], 0.0);
    for (int i = 0; i < na * K; ++i) au[i] = cplx(p->absorber[i], 0.0);
    cplx tsfc(sfc_temperature(p), 0.0);
    for (int l = 0; l < L; ++l) eps[l] = cplx(so ? 0.0 : emissivity[l], 0.0);
    cplx dtw(0.0), dsal(0.0), dwind(0.0), daz(0.0);
    if (d < K) tu[d] += cplx(0.0, h);
    else if (d < K * (1 + na)) au[d - K] += cplx(0.0, h);
    else if (d == K * (1 + na)) {
      if (so) { tsfc += cplx(0.0, h * p->coverage[1]); dtw = cplx(0.0, h); }
      else tsfc += cplx(0.0, h);
    } else if (!so) for (int l = 0; l < L; ++l) eps[l] += cplx(0.0, h);
    else if (d == nd0) dsal = cplx(0.0, h);
    else if (d == nd0 + 1) dwind = cplx(0.0, h);
    else daz = cplx(0.0, h);
    ProfileState<cplx> st;
    map_and_predict<cplx>(s, p, tu.data(), au.data(), st);
    ChannelFwd<cplx> cf;
    for (int l = 0; l < L; ++l) {
      SfcArgs<cplx> sa;
      if (so) {
        make_sfc_args(s, p, so, l, st.cos_sensor, sa);
        sa.tw += dtw; sa.sal += dsal; sa.wind += dwind; sa.az += daz;
      }
      channel_forward<cplx>(s, p, tu.data(), tsfc, eps[l], l, st, cf, so ? &sa : nullptr);
      double dbt = im(cf.bt) / h, drad = im(cf.radiance) / h;
      double g = dbt * bt_k[l] + drad * rad_k[l];
      if (d < K) {
        if (o->t_k) o->t_k[(size_t)l * K + d] = g;
      } else if (d < K * (1 + na)) {
        if (o->abs_k) o->abs_k[(size_t)l * na * K + (d - K)] = g;
      } else if (d == K * (1 + na)) {
        if (o->sfc_t_k)
          for (int c = 0; c < 4; ++c) o->sfc_t_k[(size_t)l * 4 + c] = so ? (c == 1 ? g : 0.0) : p->coverage[c] * g;
      } else if (!so) {
        if (o->emis_k) o->emis_k[l] = g;
      } else {
        if (sfc_k) sfc_k[(size_t)l * 3 + (d - nd0)] = g;
        if (o->emis_k) o->emis_k[l] = 0.0;
      }
    }
  }
  if (o->radiance_k)
    for (int l = 0; l < L; ++l)
      o->radiance_k[l] = rad_k[l] + planck_t_ad_factor(s, l, o->radiance[l]) * bt_k[l];
  return 0;
}

extern "C" int oracle_layer_avg(int n_odps, const double* ref_lnp, int n_user, const double* user_lnp,
                                double* weights, int* index) {
  layer_avg(ref_lnp, n_odps, user_lnp, n_user, weights, index);
  return 0;
}

extern "C" int oracle_predictors(const or_sensor* s, const or_profile* p, double* x, double* secant,
                                 double* t_odps, double* a_odps) {
  ProfileState<double> st;
  map_and_predict<double>(s, p, p->temperature, p->absorber, st);
  if (x) std::memcpy(x, st.x.data(), sizeof(double) * st.x.size());
  if (secant) std::memcpy(secant, st.sec.data(), sizeof(double) * st.sec.size());
  if (t_odps) std::memcpy(t_odps, st.t.data(), sizeof(double) * st.t.size());
  if (a_odps) std::memcpy(a_odps, st.a.data(), sizeof(double) * st.a.size());
  return 0;
}

// ===========================================================================
// Reverse-mode K-matrix (CRTM_K_Matrix channel loop, src/CRTM_K_Matrix_Module.f90:984-1631)
// ===========================================================================
namespace {

struct AdjWork {
  std::vector<double> tau_ad, b_ad, upath_ad, path_ad, od_ad, x_ad, t_ad, a_ad;
  std::vector<double> ox_ad, ap_ad, da_ad, tz_ad, tzp_ad, gazp_ad, gatzp_ad;
};

// ODPS_Compute_Predictor_AD per layer (ODPS_Predictor.f90:1894-2641): from
// X_AD(k,:,:) accumulate direct T/absorber adjoints and the adjoints of the
// integrated ratios Tz, Tzp, GAzp(j), GATzp(j) at layer k.
void predictor_layer_ad(const or_sensor* s, const ProfileState<double>& st, int k, const double* xad,
                        AdjWork& w) {
  const int g = s->group, J = st.J, Ko = st.Ko;
  const double* rt = s->ref_temperature;
  const double* ra = s->ref_absorber;
  auto XA = [&](int i, int j) { return xad[(j - 1) * MAXP + (i - 1)]; };
  double sec = st.sec[k];
  double dt = st.t[k] - rt[k];
  double T = st.t[k] / rt[k], T2 = T * T;
  double adt = std::fabs(dt), dt2 = dt * adt;
  auto A = [&](int j) { return st.a[(size_t)j * Ko + k] / ra[(size_t)j * Ko + k]; };
  auto GAzp = [&](int j) { return st.gazp[(size_t)j * Ko + k]; };
  auto GATzp = [&](int j) { return st.gatzp[(size_t)j * Ko + k]; };
  double h2o = A(0), h2o_a = sec * h2o, h2o_r = std::sqrt(h2o_a), h2o_s = h2o_a * h2o_a;
  double h2o_r4 = std::sqrt(h2o_r), gt0 = GATzp(0), h2o_dtzp = h2o / gt0;
  // adjoint accumulators of the intermediates
  double T_ad = 0, T2_ad = 0, dt_ad = 0, dt2_ad = 0;
  double h2o_ad = 0, h2o_a_ad = 0, h2o_s_ad = 0, h2o_r_ad = 0, h2o_r4_ad = 0, h2o_dtzp_ad = 0;
  double* gazp_ad = &w.gazp_ad[0];
  double* gatzp_ad = &w.gatzp_ad[0];
  auto GZ = [&](int j) -> double& { return gazp_ad[(size_t)j * Ko + k]; };
  auto GT = [&](int j) -> double& { return gatzp_ad[(size_t)j * Ko + k]; };
  double tz_ad = 0, tzp_ad = 0;
  std::vector<double> a_ad(st.Jm, 0.0);   // adjoint of absorber ratios A(j)
  // dry (component 1): sec, sec*T, sec*T2, T, sec^2, T2, Tz
  T_ad += XA(2, 1) * sec + XA(4, 1);
  T2_ad += XA(3, 1) * sec + XA(6, 1);
  tz_ad += XA(7, 1);
  if (g == 1 || g == 2) {
    double o3 = A(1), co2 = A(2), o3_a = sec * o3, o3_r = std::sqrt(o3_a), o3p = GAzp(1);
    // water-vapour continuum (3)
    {
      double x1 = XA(1, 3), x2 = XA(2, 3), x3 = XA(3, 3), x4 = XA(4, 3), x5 = XA(5, 3), x6 = XA(6, 3),
             x7 = XA(7, 3);
      double iT = 1.0 / T, iT2 = 1.0 / T2, iT4 = iT2 * iT2;
      h2o_a_ad += x1 * iT + x2 * h2o * iT + x3 * h2o * iT4 + x4 * iT2 + x5 * h2o * iT2 + x6 * iT4 + x7;
      h2o_ad += (x2 * iT + x3 * iT4 + x5 * iT2) * h2o_a;
      T_ad -= (x1 + x2 * h2o) * h2o_a * iT * iT;
      T2_ad -= (2.0 * x3 * h2o * iT4 * iT2 + x4 * iT4 + x5 * h2o * iT4 + 2.0 * x6 * iT4 * iT2) * h2o_a;
    }
    // ozone (4)
    {
      double x[12];
      for (int i = 1; i <= 11; ++i) x[i] = XA(i, 4);
      double sp = std::sqrt(sec * o3p);
      double o3_a_ad = x[1] + x[2] * dt + x[3] * o3 * o3p + 2.0 * x[4] * o3_a + x[5] * o3p + x[6] * sp;
      dt_ad += x[2] * o3_a + x[7] * o3_r;
      double o3_r_ad = x[7] * dt + x[8] + x[9] * o3 / o3p;
      double o3_ad = x[3] * o3_a * o3p + x[9] * o3_r / o3p;
      GZ(1) += x[3] * o3_a * o3 + x[5] * o3_a + x[6] * 0.5 * o3_a * std::sqrt(sec / o3p) -
               x[9] * o3_r * o3 / (o3p * o3p) + x[10] * sec + x[11] * 2.0 * sec * sec * o3p;
      o3_a_ad += o3_r_ad * 0.5 / o3_r;
      o3_ad += o3_a_ad * sec;
      a_ad[1] += o3_ad;
    }
    // CO2 (5)
    {
      double tzp = st.tzp[k], sT = std::sqrt(T);
      T_ad += XA(1, 5) * sec + XA(3, 5) + XA(10, 5) * sec * tzp * 0.5 / sT;
      T2_ad += XA(2, 5) * sec + XA(4, 5);
      a_ad[2] += XA(6, 5) * sec;
      tzp_ad += XA(7, 5) * sec + XA(9, 5) * 3.0 * tzp * tzp + XA(10, 5) * sec * sT;
      GZ(2) += XA(8, 5) * 2.0 * sec * sec * GAzp(2);
    }
    // water-vapour lines (2)
    {
      h2o_a_ad += XA(1, 2) + XA(2, 2) * dt + XA(4, 2) * dt2 + XA(6, 2) * h2o_s;
      dt_ad += XA(2, 2) * h2o_a + XA(8, 2) * h2o_r;
      h2o_s_ad += XA(3, 2) + XA(6, 2) * h2o_a + XA(9, 2) * 2.0 * h2o_s;
      dt2_ad += XA(4, 2) * h2o_a;
      h2o_r4_ad += XA(5, 2);
      h2o_r_ad += XA(7, 2) + XA(8, 2) * dt + XA(11, 2) * h2o_dtzp;
      h2o_dtzp_ad += XA(10, 2) + XA(11, 2) * h2o_r;
      GZ(0) += XA(12, 2) * 2.0 * sec * sec * GAzp(0) + XA(13, 2) * sec;
      a_ad[2] += XA(15, 2) * sec;
    }
    if (g == 1) {
      double co = A(4), n2o = A(3), ch4 = A(5);
      double n2o_a = sec * n2o, n2o_r = std::sqrt(n2o_a);
      double co_a = sec * co, co_r = std::sqrt(co_a);
      double cozp = GAzp(4), co_q = co_a * co / cozp;
      double ch4_a = sec * ch4, ch4_r = std::sqrt(ch4_a), ch4zp = GAzp(5), ch4_q = sec * ch4zp;
      double n2ozp = GAzp(3);
      double co_a_ad = 0, co_r_ad = 0, co_s_ad = 0, co_q_ad = 0, co_ad = 0;
      double ch4_a_ad = 0, ch4_r_ad = 0, ch4_q_ad = 0, ch4_ad = 0;
      double n2o_a_ad = 0, n2o_r_ad = 0, n2o_s_ad = 0, n2o_ad = 0;
      co_a_ad += XA(18, 2) + XA(11, 5);
      ch4_a_ad += XA(16, 2) + XA(17, 2) * 2.0 * ch4_a * dt;
      dt_ad += XA(17, 2) * ch4_a * ch4_a;
      // CO (7)
      co_a_ad += XA(1, 7) + XA(2, 7) * dt + XA(7, 7) * dt2;
      dt_ad += XA(2, 7) * co_a + XA(4, 7) * co_r;
      co_r_ad += XA(3, 7) * 0.5 / std::sqrt(co_r) + XA(4, 7) * dt + XA(6, 7) - XA(9, 7) * co_q / (co_r * co_r);
      co_s_ad += XA(5, 7);
      dt2_ad += XA(7, 7) * co_a;
      co_q_ad += XA(8, 7) + XA(9, 7) / co_r + XA(10, 7) * std::sqrt(cozp);
      GZ(4) += XA(10, 7) * 0.5 * co_q / std::sqrt(cozp);
      // CH4 (8)
      ch4_a_ad += XA(1, 8) * dt + XA(3, 8) * 2.0 * ch4_a + XA(4, 8);
      dt_ad += XA(1, 8) * ch4_a + XA(5, 8) * ch4;
      ch4_r_ad += XA(2, 8) + XA(8, 8) * 0.5 / std::sqrt(ch4_r) + XA(11, 8) * ch4 / ch4zp;
      ch4_ad += XA(5, 8) * dt + XA(11, 8) * ch4_r / ch4zp;
      ch4_q_ad += XA(6, 8) + XA(7, 8) * 2.0 * ch4_q;
      GT(5) += XA(9, 8) + XA(10, 8) * sec;
      GZ(5) -= XA(11, 8) * ch4_r * ch4 / (ch4zp * ch4zp);
      // N2O (6)
      n2o_a_ad += XA(1, 6) * dt + XA(4, 6) * 0.25 * std::pow(n2o_a, -0.75) + XA(5, 6);
      dt_ad += XA(1, 6) * n2o_a + XA(3, 6) * n2o;
      n2o_r_ad += XA(2, 6) + XA(10, 6) * n2o / n2ozp;
      n2o_ad += XA(3, 6) * dt + XA(10, 6) * n2o_r / n2ozp;
      GZ(3) += XA(6, 6) * sec - XA(10, 6) * n2o_r * n2o / (n2ozp * n2ozp);
      GT(3) += XA(7, 6) * sec + XA(9, 6);
      n2o_s_ad += XA(8, 6);
      ch4_a_ad += XA(11, 6) + XA(12, 6) * ch4zp;
      GZ(5) += XA(12, 6) * ch4_a;
      co_a_ad += XA(13, 6) + XA(14, 6) * sec * cozp;
      GZ(4) += XA(14, 6) * co_a * sec;
      // chain the intermediates back to the ratios
      GZ(5) += sec * ch4_q_ad;
      ch4_a_ad += 0.5 / ch4_r * ch4_r_ad;
      ch4_ad += sec * ch4_a_ad;
      GZ(4) -= co_q_ad * co_a * co / (cozp * cozp);
      co_a_ad += co_q_ad * co / cozp + co_s_ad * 2.0 * co_a + co_r_ad * 0.5 / co_r;
      co_ad += co_q_ad * co_a / cozp + co_a_ad * sec;
      n2o_a_ad += n2o_r_ad * 0.5 / n2o_r + n2o_s_ad * 2.0 * n2o_a;
      n2o_ad += n2o_a_ad * sec;
      a_ad[5] += ch4_ad;
      a_ad[3] += n2o_ad;
      a_ad[4] += co_ad;
    }
  } else {
    // MW wet component (2)
    double iT = 1.0 / T, iT2 = 1.0 / T2, iT4 = iT2 * iT2;
    double x[15];
    for (int i = 1; i <= 14; ++i) x[i] = XA(i, 2);
    h2o_a_ad += x[1] * iT + x[2] * h2o * iT + x[3] * h2o * iT4 + x[4] * iT2 + x[5] * h2o * iT2 + x[6] * iT4 +
                x[7] + x[8] * dt + x[12] * h2o_s;
    h2o_ad += (x[2] * iT + x[3] * iT4 + x[5] * iT2) * h2o_a;
    T_ad -= (x[1] + x[2] * h2o) * h2o_a * iT * iT;
    T2_ad -= (2.0 * x[3] * h2o * iT4 * iT2 + x[4] * iT4 + x[5] * h2o * iT4 + 2.0 * x[6] * iT4 * iT2) * h2o_a;
    dt_ad += x[8] * h2o_a;
    GZ(0) += x[9] * 2.0 * sec * sec * GAzp(0) + x[10] * sec;
    h2o_s_ad += x[12] * h2o_a + x[13] * 2.0 * h2o_s;
    h2o_dtzp_ad += x[14];
  }
  // water vapour chain
  GT(0) -= h2o_dtzp_ad * h2o / (gt0 * gt0);
  h2o_r_ad += h2o_r4_ad * 0.5 / h2o_r4;
  h2o_a_ad += h2o_s_ad * 2.0 * h2o_a + h2o_r_ad * 0.5 / h2o_r;
  h2o_ad += h2o_a_ad * sec + h2o_dtzp_ad / gt0;
  a_ad[0] += h2o_ad;
  // temperature chain: DT2 = DT*|DT|, T2 = T*T, T = t/tref, DT = t - tref
  dt_ad += dt2_ad * 2.0 * adt;
  T_ad += T2_ad * 2.0 * T;
  w.t_ad[k] += T_ad / rt[k] + dt_ad;
  for (int j = 0; j < N_ABS_G[g]; ++j) w.a_ad[(size_t)j * Ko + k] += a_ad[j] / ra[(size_t)j * Ko + k];
  w.tz_ad[k] += tz_ad;
  w.tzp_ad[k] += tzp_ad;
}

// Adjoint seeds of one channel from the all-sky path (cloudy columns): the
// clear-sky Emission adjoint is run with rad_ad_clear (the clear share of
// the radiance adjoint under fractional cover, else 0) and the layer and
// surface adjoints of the scattering RT are added on top.
struct ChannelSeeds {
  bool on = false;
  double radiance = 0.0, bt = 0.0, rad_k = 0.0;   // all-sky radiance, BT, processed radiance adjoint
  double rad_ad_clear = 0.0;
  std::vector<double> tau_ad, b_ad, t_ad;          // [K] nadir OD, layer Planck and direct T adjoints
  double emis_ad = 0.0, bs_ad = 0.0;               // surface emissivity and Planck adjoints
};

void k_matrix_profile(const or_sensor* s, const or_profile* p, const double* emissivity,
                      const double* bt_k, const double* rad_k, or_out* o,
                      const std::vector<ChannelSeeds>* seeds = nullptr,
                      const or_surface_optics* so = nullptr, double* sfc_k = nullptr) {
  const int K = p->n_layers, L = s->n_channels, na = p->n_absorbers;
  ProfileState<double> st;
  map_and_predict<double>(s, p, p->temperature, p->absorber, st);
  const int Ko = st.Ko, J = st.J, Jm = st.Jm;
  double tsfc = sfc_temperature(p);
  ChannelFwd<double> cf;
  AdjWork w;
  for (int l = 0; l < L; ++l) {
    SfcArgs<double> sa;
    if (so) make_sfc_args(s, p, so, l, st.cos_sensor, sa);
    channel_forward<double>(s, p, p->temperature, tsfc, so ? 0.0 : emissivity[l], l, st, cf, so ? &sa : nullptr);
    const ChannelSeeds* sd = seeds ? &(*seeds)[l] : nullptr;
    const bool ext = sd && sd->on;
    o->radiance[l] = ext ? sd->radiance : cf.radiance;
    o->bt[l] = ext ? sd->bt : cf.bt;
    if (o->tau) std::memcpy(o->tau + (size_t)l * K, cf.tau.data(), sizeof(double) * K);
    if (o->up_radiance) o->up_radiance[l] = cf.up_radiance;
    if (o->down_radiance) o->down_radiance[l] = cf.rd[K];
    // Pre_Process_RTSolution_K (src/CRTM_K_Matrix_Module.f90:1725-1761)
    // CRTM_Compute_AntCorr_AD scales the BT adjoint; the Planck temperature
    // adjoint is taken at the NLTE-corrected radiance
    double aep = 1.0, as_ = 0.0;
    antenna_coeffs(s, p, l, aep, as_);
    double rad_ad = ext ? sd->rad_ad_clear : rad_k[l] + planck_t_ad_factor(s, l, cf.radiance) * aep * bt_k[l];
    if (o->radiance_k) o->radiance_k[l] = ext ? sd->rad_k : rad_ad;
    // CRTM_Emission_AD (Emission_Module.f90:280-395)
    double u = st.cos_sensor, sec_down = 1.0 / u, eps = cf.eps, rho = cf.rho;
    w.tau_ad.assign(K, 0.0);
    w.b_ad.assign(K, 0.0);
    double up_ad = rad_ad;
    for (int k = 1; k <= K; ++k) {
      double tu = cf.tu[k - 1];
      double a_ad = tu * (cf.ru[k] - cf.B[k - 1]) * up_ad;
      w.b_ad[k - 1] += up_ad * (1.0 - tu);
      up_ad *= tu;
      w.tau_ad[k - 1] -= a_ad / u;
    }
    double eps_ad = up_ad * cf.bs;
    double bs_ad = eps * up_ad;
    double r_ad = up_ad * cf.rd[K];
    double down_ad = rho * up_ad;
    for (int k = K; k >= 1; --k) {
      double td = cf.td[k - 1];
      double a_ad = td * (cf.rd[k - 1] - cf.B[k - 1]) * down_ad;
      w.b_ad[k - 1] += down_ad * (1.0 - td);
      down_ad *= td;
      w.tau_ad[k - 1] -= a_ad * sec_down;
    }
    if (ext) {
      for (int k = 0; k < K; ++k) {
        w.tau_ad[k] += sd->tau_ad[k];
        w.b_ad[k] += sd->b_ad[k];
      }
      bs_ad += sd->bs_ad;
      eps_ad += sd->emis_ad;
    }
    // computed optics: CRTM_Compute_SfcOptics_AD through the emissivity and
    // reflectivity adjoints (derivatives of the optics by complex step), then
    // CRTM_Compute_Transmittance_AD (CRTM_AtmOptics.f90:314-368)
    double sg[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (so) {
      const double h = 1.0e-30;
      for (int d = 0; d < 5; ++d) {
        cplx in[5] = {cplx(sa.tw), cplx(sa.sal), cplx(sa.wind), cplx(sa.az), cplx(cf.trans)};
        in[d] += cplx(0.0, h);
        cplx e, r;
        SfcPoint spc = sa.sp;
        sfc_eval<cplx>(spc, in[0], in[1], in[2], in[3], in[4], e, r);
        sg[d] = im(e) / h * eps_ad + im(r) / h * r_ad;
      }
      for (int k = 0; k < K; ++k) w.tau_ad[k] -= cf.trans * sg[4];
      if (sfc_k)
        for (int d = 0; d < 3; ++d) sfc_k[(size_t)l * 3 + d] = sg[1 + d];
    }
    // Planck AD (Assign_Common_Output_AD, Common_RTSolution.f90:1680-1700)
    double* tk = o->t_k ? o->t_k + (size_t)l * K : nullptr;
    double* ak = o->abs_k ? o->abs_k + (size_t)l * na * K : nullptr;
    if (tk)
      for (int k = 0; k < K; ++k)
        tk[k] = planck_r_ad_factor(s, l, p->temperature[k]) * w.b_ad[k] + (ext ? sd->t_ad[k] : 0.0);
    // Compute_NLTE_Correction_AD / Compute_NLTE_Predictor_AD
    // (CRTM_NLTECorrection.f90:396-446, 704-744): the mean temperature's
    // adjoint spread over the levels (k1, k2] of each layer
    double nc[3];
    if (tk && !ext && st.nl_ctm && nlte_coeffs(s, st, l, nc))
      for (int i = 0; i < 2; ++i) {
        const int k1 = st.nl_k1[i], k2 = st.nl_k2[i];
        if (k2 <= k1) continue;
        const double g = nc[1 + i] * rad_ad / double(k2 - k1);
        for (int k = k1; k < k2; ++k) tk[k] += g;
      }
    if (ak) for (int i = 0; i < na * K; ++i) ak[i] = 0.0;
    double sfct_ad = planck_r_ad_factor(s, l, tsfc) * bs_ad;
    if (o->sfc_t_k)
      for (int c = 0; c < 4; ++c) o->sfc_t_k[(size_t)l * 4 + c] = p->coverage[c] * sfct_ad + (c == 1 ? sg[0] : 0.0);
    if (o->emis_k) o->emis_k[l] = so ? 0.0 : eps_ad - r_ad;
    // ODPS_Compute_AtmAbsorption_AD (ODPS_AtmAbsorption.f90:487-600)
    w.upath_ad.assign(K + 1, 0.0);
    for (int k = K; k >= 1; --k) {
      w.upath_ad[k] += w.tau_ad[k - 1] / st.sec_sfc;
      w.upath_ad[k - 1] -= w.tau_ad[k - 1] / st.sec_sfc;
    }
    w.path_ad.assign(Ko + 1, 0.0);
    for (int i = K; i >= 0; --i) {
      int k1 = st.u_idx[2 * i], k2 = st.u_idx[2 * i + 1];
      if (k1 == k2) {
        w.path_ad[k1 - 1] += w.upath_ad[i];
      } else {
        double x1 = st.rlnp_lev[k1 - 1], x2 = st.rlnp_lev[k2 - 1];
        double fac = (st.ulnp_lev[i] - x1) / (x2 - x1);
        w.path_ad[k1 - 1] += (1.0 - fac) * w.upath_ad[i];
        w.path_ad[k2 - 1] += fac * w.upath_ad[i];
      }
    }
    w.od_ad.assign(Ko, 0.0);
    for (int k = Ko; k >= 1; --k) {
      w.path_ad[k - 1] += w.path_ad[k];
      double od = cf.od[k - 1];
      w.od_ad[k - 1] = (od < 0.0 || od > MAX_OD) ? 0.0 : w.path_ad[k];
    }
    w.x_ad.assign((size_t)Ko * J * MAXP, 0.0);
    w.ox_ad.assign((size_t)Ko * N_OPTRAN_PRED, 0.0);
    w.ap_ad.assign((size_t)Ko * MAX_OPTRAN_ORDER, 0.0);
    w.da_ad.assign(Ko, 0.0);
    for (int j = 0; j < J; ++j) {
      int np = s->n_predictors[j + J * l];
      if (np <= 0) continue;
      bool optran = st.optran && (j + 1 == s->ocomponent_index) && s->osignificance[l] == 1;
      if (optran) {
        // Add_OPTRAN_wloOD_AD (ODPS_AtmAbsorption.f90:911-984)
        int nop = s->op_index[7 * l];
        if (nop <= 0) continue;
        int js = s->opos_index[l], no = s->order[l];
        for (int k = 0; k < Ko; ++k) {
          double chi_ad = w.od_ad[k] * st.da[k];
          w.da_ad[k] += w.od_ad[k] * cf.chi[k];
          double lc = cf.ln_chi[k];
          double lc_ad = (lc > LIMIT_EXP || lc < -LIMIT_EXP) ? 0.0 : cf.chi[k] * chi_ad;
          for (int i = 0; i <= nop; ++i) {
            double b_ad = lc_ad;
            if (i > 0) {
              int ii = s->op_index[7 * l + i];
              b_ad = lc_ad * st.ox[(size_t)k * N_OPTRAN_PRED + ii - 1];
              w.ox_ad[(size_t)k * N_OPTRAN_PRED + ii - 1] += lc_ad * cf.b[(size_t)k * (nop + 1) + i];
            }
            const double* coeff = s->oc + (js + i * (no + 1) - 1);
            for (int o2 = 1; o2 <= no; ++o2) w.ap_ad[(size_t)k * MAX_OPTRAN_ORDER + o2 - 1] += coeff[o2] * b_ad;
          }
        }
      } else {
        int j0 = s->pos_index[j + J * l];
        for (int i = 1; i <= np; ++i) {
          int js = j0 + (i - 1) * Ko - 1;
          for (int k = 1; k <= Ko; ++k)
            w.x_ad[XI(k - 1, j, i - 1, J)] += (double)s->c[js + k - 1] * w.od_ad[k - 1];
        }
      }
    }
    // ODPS_Assemble_Predictors_AD (ODPS_Predictor.f90:514-570)
    w.t_ad.assign(Ko, 0.0);
    w.a_ad.assign((size_t)Ko * Jm, 0.0);
    if (st.optran) {
      // ODPS_Compute_Predictor_ODAS_AD (ODPS_Predictor.f90:3092-3231)
      const double* pr = s->ref_pressure;
      double* vap_ad = &w.a_ad[(size_t)(st.h2o_idx - 1) * Ko];
      const double* vap = &st.a[(size_t)(st.h2o_idx - 1) * Ko];
      double iv_next_ad = 0.0, s_t_ad = 0.0, s_p_ad = 0.0, xlt_ad_carry = 0.0, xlp_ad_carry = 0.0;
      for (int k = Ko; k >= 1; --k) {
        int q = k - 1;
        double iv_ad = iv_next_ad;  // IV(k) feeds IV(k+1) and AveA(k+1)
        double ap1 = st.ap1[q], ap1_ad = 0.0, pw = 1.0;
        for (int i = 1; i <= MAX_OPTRAN_ORDER; ++i) {
          ap1_ad += i * pw * w.ap_ad[(size_t)q * MAX_OPTRAN_ORDER + i - 1];
          pw *= ap1;
        }
        double ave_ad = ap1_ad / (s->alpha * (st.ave_a[q] - s->alpha_c2));
        double xlt_ad = xlt_ad_carry + w.ox_ad[(size_t)q * N_OPTRAN_PRED + 11];
        double xlp_ad = xlp_ad_carry + w.ox_ad[(size_t)q * N_OPTRAN_PRED + 12];
        xlt_ad_carry = w.ox_ad[(size_t)q * N_OPTRAN_PRED + 11];
        xlp_ad_carry = w.ox_ad[(size_t)q * N_OPTRAN_PRED + 12];
        s_t_ad += 0.5 * st.inverse[q] * xlt_ad;
        s_p_ad += 0.5 * st.inverse[q] * xlp_ad;
        double inv_ad = 0.5 * st.s_t[q] * xlt_ad + 0.5 * st.s_p[q] * xlp_ad;
        if (st.int_vapor[q] > MIN_ABSORBER) iv_ad -= inv_ad / (st.int_vapor[q] * st.int_vapor[q]);
        double da_ad = w.da_ad[q] + pr[q] * s_p_ad + st.t[q] * s_t_ad;
        w.t_ad[q] += st.da[q] * s_t_ad;
        iv_ad += 0.5 * ave_ad;
        double ivprev_ad = 0.5 * ave_ad + iv_ad;
        da_ad += iv_ad;
        double dpong = RECIP_GRAV * (s->ref_level_pressure[k] - s->ref_level_pressure[k - 1]);
        vap_ad[q] += dpong * st.sec[q] * da_ad;
        iv_next_ad = ivprev_ad;
        // direct OX terms
        double t = st.t[q], t2 = t * t, pk = pr[q], p2 = pk * pk;
        const double* oa = &w.ox_ad[(size_t)q * N_OPTRAN_PRED];
        w.t_ad[q] += oa[0] + pk * oa[4] + p2 * oa[6] +
                     2.0 * t * (oa[2] + pk * oa[5] + p2 * oa[7] - vap[q] / (t2 * t2) * oa[10]);
        vap_ad[q] += oa[9] + oa[10] / t2;
      }
    }
    w.tz_ad.assign(Ko, 0.0);
    w.tzp_ad.assign(Ko, 0.0);
    w.gazp_ad.assign((size_t)Ko * Jm, 0.0);
    w.gatzp_ad.assign((size_t)Ko * Jm, 0.0);
    for (int k = 0; k < Ko; ++k) predictor_layer_ad(s, st, k, &w.x_ad[(size_t)k * J * MAXP], w);
    // adjoint of the integrated ratios (reverse cumulative sums)
    {
      double tz_sum_ad = 0.0, tzp_sum_ad = 0.0;
      std::vector<double> gzs(Jm, 0.0), gts(Jm, 0.0);
      for (int k = Ko - 1; k >= 0; --k) {
        for (int j = 0; j < N_ABS_G[s->group]; ++j) {
          size_t jk = (size_t)j * Ko + k;
          gts[j] += w.gatzp_ad[jk] / st.gatzp_ref[jk];
          gzs[j] += w.gazp_ad[jk] / st.gazp_ref[jk];
          w.t_ad[k] += gts[j] * st.pdp[k] * st.a[jk];
          w.a_ad[jk] += gzs[j] * st.pdp[k] + gts[j] * st.pdp[k] * st.t[k];
        }
        tzp_sum_ad += w.tzp_ad[k] / st.tzp_ref[k];
        tz_sum_ad += w.tz_ad[k] / st.tz_ref[k];
        w.t_ad[k] += tz_sum_ad + st.pdp[k] * tzp_sum_ad;
      }
    }
    // Map_Input_AD (ODPS_CoordinateMapping.f90:480-547)
    for (int j = 0; j < Jm; ++j) {
      if (st.idx_map[j] < 0) continue;
      double* dst = ak ? ak + (size_t)st.idx_map[j] * K : nullptr;
      for (int k = 0; k < Ko; ++k) {
        double aad = (st.a[(size_t)j * Ko + k] <= TOLERANCE) ? 0.0 : w.a_ad[(size_t)j * Ko + k];
        if (dst)
          for (int uu = st.acc_idx[2 * k]; uu <= st.acc_idx[2 * k + 1]; ++uu)
            dst[uu - 1] += st.acc_w[(size_t)k * K + uu - 1] * aad;
      }
    }
    if (tk)
      for (int k = 0; k < Ko; ++k)
        for (int uu = st.acc_idx[2 * k]; uu <= st.acc_idx[2 * k + 1]; ++uu)
          tk[uu - 1] += st.acc_w[(size_t)k * K + uu - 1] * w.t_ad[k];
  }
}

}  // namespace

namespace {
// Run `fn` on the profile extended to TOA_PRESSURE and copy the user layers' part out.
template <class F>
int with_added_layers(const or_sensor* s, const or_profile* p, or_out* o, F fn) {
  ExtProfile e;
  extend_profile(p, e);
  if (e.n == 0) return fn(p, o);
  ExtOut eo;
  eo.init(o, s->n_channels, e.p.n_layers, p->n_absorbers);
  int r = fn(&e.p, &eo.o);
  eo.copy_back(o, s->n_channels, p->n_layers, e.n, p->n_absorbers);
  return r;
}
}  // namespace

extern "C" void oracle_set_climatology(const double* p, const double* t, const double* a) {
  std::memcpy(g_clim.p, p, sizeof(g_clim.p));
  std::memcpy(g_clim.t, t, sizeof(g_clim.t));
  std::memcpy(g_clim.a, a, sizeof(g_clim.a));
  g_clim.set = true;
}

extern "C" int oracle_add_layers(const or_profile* p, double* level_pressure, double* pressure,
                                 double* temperature, double* absorber) {
  ExtProfile e;
  extend_profile(p, e);
  if (level_pressure) {
    const int Kt = e.p.n_layers;
    std::memcpy(level_pressure, e.p.level_pressure, sizeof(double) * (Kt + 1));
    std::memcpy(pressure, e.p.pressure, sizeof(double) * Kt);
    std::memcpy(temperature, e.p.temperature, sizeof(double) * Kt);
    std::memcpy(absorber, e.p.absorber, sizeof(double) * Kt * p->n_absorbers);
  }
  return e.n;
}

extern "C" int oracle_forward(const or_sensor* s, const or_profile* p, const double* emissivity, or_out* o) {
  return with_added_layers(s, p, o, [&](const or_profile* q, or_out* oo) { return forward_ext(s, q, emissivity, oo); });
}

extern "C" int oracle_k_matrix_cstep(const or_sensor* s, const or_profile* p, const double* emissivity,
                                     const double* bt_k, const double* rad_k, or_out* o) {
  return with_added_layers(s, p, o,
                           [&](const or_profile* q, or_out* oo) { return cstep_ext(s, q, emissivity, bt_k, rad_k, oo); });
}

extern "C" int oracle_k_matrix(const or_sensor* s, const or_profile* p, const double* emissivity,
                               const double* bt_k, const double* rad_k, or_out* o) {
  return with_added_layers(s, p, o, [&](const or_profile* q, or_out* oo) {
    k_matrix_profile(s, q, emissivity, bt_k, rad_k, oo);
    return 0;
  });
}

extern "C" int oracle_forward_sfc(const or_sensor* s, const or_profile* p, const or_surface_optics* so, or_out* o,
                                  double* emis, double* refl) {
  return with_added_layers(s, p, o,
                           [&](const or_profile* q, or_out* oo) { return forward_ext(s, q, nullptr, oo, so, emis, refl); });
}

extern "C" int oracle_k_matrix_sfc(const or_sensor* s, const or_profile* p, const or_surface_optics* so,
                                   const double* bt_k, const double* rad_k, or_out* o, double* sfc_k) {
  return with_added_layers(s, p, o, [&](const or_profile* q, or_out* oo) {
    ProfileState<double> st;
    map_and_predict<double>(s, q, q->temperature, q->absorber, st);
    SfcArgs<double> sa;
    for (int l = 0; l < s->n_channels; ++l)
      if (!make_sfc_args(s, q, so, l, st.cos_sensor, sa)) return 3;
    k_matrix_profile(s, q, nullptr, bt_k, rad_k, oo, nullptr, so, sfc_k);
    return 0;
  });
}

extern "C" int oracle_k_matrix_sfc_cstep(const or_sensor* s, const or_profile* p, const or_surface_optics* so,
                                         const double* bt_k, const double* rad_k, or_out* o, double* sfc_k) {
  return with_added_layers(s, p, o, [&](const or_profile* q, or_out* oo) {
    return cstep_ext(s, q, nullptr, bt_k, rad_k, oo, so, sfc_k);
  });
}

extern "C" void oracle_fastemx(const or_mwwater* c, double frequency, int n_angles, double zenith, const double* in5,
                               double* emis, double* refl, double* d_emis, double* d_refl) {
  fastemx<double>(c, frequency, n_angles, zenith, in5[0], in5[1], in5[2], in5[3], in5[4], emis, refl);
  const double h = 1.0e-30;
  for (int d = 0; d < 5; ++d) {
    cplx in[5];
    for (int i = 0; i < 5; ++i) in[i] = cplx(in5[i], i == d ? h : 0.0);
    cplx e[4], r[4];
    fastemx<cplx>(c, frequency, n_angles, zenith, in[0], in[1], in[2], in[3], in[4], e, r);
    for (int j = 0; j < 4; ++j) {
      d_emis[j * 5 + d] = im(e[j]) / h;
      d_refl[j * 5 + d] = im(r[j]) / h;
    }
  }
}

extern "C" void oracle_irssem(const or_irwater* c, double temperature, double wind, double frequency, double angle,
                              double* emis, double* d2) {
  *emis = irssem<double>(c, temperature, wind, frequency, angle);
  const double h = 1.0e-30;
  d2[0] = im(irssem<cplx>(c, cplx(temperature), cplx(wind, h), frequency, angle)) / h;
  d2[1] = im(irssem<cplx>(c, cplx(temperature, h), cplx(wind), frequency, angle)) / h;
}

extern "C" int oracle_pol_weights(int pol, double angle, double distance_ratio_, double pol_angle, double scan_rad,
                                  double* w2) {
  return pol_weights(pol, angle, distance_ratio_, pol_angle, scan_rad, w2[0], w2[1]) ? 0 : 3;
}

extern "C" int oracle_k_matrix_batch(const or_sensor* s, int n_profiles, const or_profile* p,
                                     const double* emissivity, const double* bt_k, const double* rad_k,
                                     or_out* o, int n_threads) {
  const int L = s->n_channels;
  std::vector<size_t> off(n_profiles + 1, 0);
  for (int m = 0; m < n_profiles; ++m) off[m + 1] = off[m] + (size_t)p[m].n_layers;
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int m = 0; m < n_profiles; ++m) {
    const int na = p[m].n_absorbers;
    or_out om;
    std::memset(&om, 0, sizeof(om));
    om.radiance = o->radiance + (size_t)m * L;
    om.bt = o->bt + (size_t)m * L;
    if (o->t_k) om.t_k = o->t_k + off[m] * L;
    if (o->abs_k) om.abs_k = o->abs_k + off[m] * L * na;
    if (o->sfc_t_k) om.sfc_t_k = o->sfc_t_k + (size_t)m * L * 4;
    if (o->emis_k) om.emis_k = o->emis_k + (size_t)m * L;
    if (o->radiance_k) om.radiance_k = o->radiance_k + (size_t)m * L;
    oracle_k_matrix(s, &p[m], emissivity + (size_t)m * L, bt_k + (size_t)m * L, rad_k + (size_t)m * L, &om);
  }
  return 0;
}

#include "crtm_oracle_allsky.inc"
#include "crtm_oracle_allsky_ad.inc"
