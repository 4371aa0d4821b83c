// allsky_ad.hip -- all-sky K-matrix: hand-written adjoint of the scattering path.
//
//   allsky_sweep_ad_kernel  one thread per (profile, channel) of the cloudy
//                           profiles: the forward adding sweep (ADA, or
//                           Emission when no layer scatters) storing each
//                           layer's incoming state (R_up, I_up) on a per-thread
//                           tape, the fractional-cover combination and BT, then
//                           the reverse sweep top-down: CRTM_ADA_AD
//                           (src/RTSolution/ADA/ADA_Module.f90:1547-1756) or
//                           CRTM_Emission_AD (src/RTSolution/Emission/
//                           Emission_Module.f90:280-395).  Writes the layer
//                           R/T/source adjoints of scattering layers, the optical
//                           depth adjoint of the other layers, the Planck and
//                           surface adjoints and the cover adjoint.
//   cloud_rt_ad_kernel      one thread per scattering (cloudy layer, channel):
//                           CRTM_AMOM_layer_AD with Matrix_Exp_AD
//                           (ADA_Module.f90:963-1109, 1760-2081) and the phase
//                           matrix + Normalize_Phase adjoint
//                           (src/RTSolution/Common_RTSolution.f90:2548-2879,
//                           3044-3116).  The layer is recomputed from its optics;
//                           doubling steps are recomputed from the series start
//                           instead of being stored.
//   cloud_optics_ad_kernel  one thread per (cloudy layer, channel):
//                           CRTM_AtmOptics_Combine_AD (src/AtmOptics/
//                           CRTM_AtmOptics.f90:628-784) and CRTM_Compute_CloudScatter_AD
//                           with the table-interpolation derivatives
//                           (src/AtmScatter/CRTM_CloudScatter.f90:737-907,
//                           1134-1258, 1660-1937; src/Interpolation/
//                           CRTM_Interpolation.f90:535-668): gas optical depth,
//                           water content, effective radius and temperature adjoints.
//   cloud_cover_ad_kernel   one thread per (profile, channel) of fractional
//                           profiles: Compute_CloudCover_AD with the water-content
//                           rescaling (src/Atmosphere/Cloud/CRTM_CloudCover_Define.f90).
//
// The gas-absorption, Planck and surface parts are completed by the clear-sky
// K-matrix kernel, which adds these seeds to its own Emission adjoint.
#include "allsky_common.h"

namespace crtm {

namespace {

// ------------------------------------------------ small-matrix helpers
template <class T, int N>
__device__ inline void mul_tn(const T (&a)[N][N], const T (&b)[N][N], T (&c)[N][N]) {   // c = a^T b
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      T s = T(0);
#pragma unroll
      for (int k = 0; k < N; ++k) s += a[k][i] * b[k][j];
      c[i][j] = s;
    }
}
template <class T, int N>
__device__ inline void mul_nt(const T (&a)[N][N], const T (&b)[N][N], T (&c)[N][N]) {   // c = a b^T
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      T s = T(0);
#pragma unroll
      for (int k = 0; k < N; ++k) s += a[i][k] * b[j][k];
      c[i][j] = s;
    }
}
template <class T, int N>
__device__ inline void add_mul_nt(const T (&a)[N][N], const T (&b)[N][N], T (&c)[N][N]) {   // c += a b^T
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      T s = c[i][j];
#pragma unroll
      for (int k = 0; k < N; ++k) s += a[i][k] * b[j][k];
      c[i][j] = s;
    }
}
template <class T, int N>
__device__ inline void add_mul_tn(const T (&a)[N][N], const T (&b)[N][N], T (&c)[N][N]) {   // c += a^T b
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      T s = c[i][j];
#pragma unroll
      for (int k = 0; k < N; ++k) s += a[k][i] * b[k][j];
      c[i][j] = s;
    }
}
template <class T, int N>
__device__ inline T dot(const T (&a)[N][N], const T (&b)[N][N]) {
  T s = T(0);
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) s += a[i][j] * b[i][j];
  return s;
}

// Number of doublings of Matrix_Exp, as in the forward (allsky_common.h)
template <class T, int N>
__device__ inline int doublings(const T (&amb)[N][N], const T (&apb)[N][N], T tau) {
  T mh = T(0);
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      T s = T(0);
#pragma unroll
      for (int k = 0; k < N; ++k) s += amb[i][k] * apb[k][j];
      mh = fmax(mh, fabs(s));
    }
  const double max_h = sqrt((double)mh) * (double)tau;
  const double lv = max_h > 0.0 ? log(max_h / 0.5) / log(2.0) : -2.0;
  return lv <= -1.0 ? 0 : (lv < 60.0 ? (int)trunc(lv) + 1 : 60);
}

// Series start of Matrix_Exp (the forward's arithmetic): R0, T0 of the layer
// at optical depth t = tau / 2^nd.
template <class T, int N>
__device__ void mexp_start(const T (&apb)[N][N], const T (&amb)[N][N], T t, T (&R0)[N][N], T (&T0)[N][N]) {
  T hh[N][N], ff[N][N], u[N][N], v[N][N], w[N][N];
  mul(amb, apb, hh);
  mul(apb, amb, ff);
  const T t2 = t * t;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      hh[i][j] *= t2;
      ff[i][j] *= t2;
    }
  series(hh, u, 0);
  series(hh, v, 1);
  mul(apb, v, w);
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      const T s1 = u[i][j], s3 = t * w[i][j];
      u[i][j] = s1 - s3;
      v[i][j] = s1 + s3;
    }
  series(ff, hh, 0);
  series(ff, w, 1);
  mul(amb, w, ff);
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      const T s2 = hh[i][j], s4 = t * ff[i][j];
      u[i][j] += s2 - s4;
      v[i][j] -= s2 + s4;
    }
  inv_inplace(u);
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) T0[i][j] = T(2) * u[i][j];
  mul(v, u, R0);
}

// One doubling step in place: R' = R + T (I - R R)^-1 R T, T' = T (I - R R)^-1 T
template <class T, int N>
__device__ inline void double_step(T (&R)[N][N], T (&Tr)[N][N]) {
  T u[N][N], v[N][N], w[N][N];
  mul(R, R, u);
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) u[i][j] = (i == j ? T(1) : T(0)) - u[i][j];
  inv_inplace(u);
  mul(Tr, u, v);
  mul(v, R, u);
  mul(u, Tr, w);
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) R[i][j] += w[i][j];
  mul_left_inplace(v, Tr);
}

// Reverse of one doubling step entered with (R, T); Ra, Ta in: adjoints of
// the step's outputs, out: adjoints of its inputs.
template <class T, int N>
__device__ void double_step_ad(const T (&R)[N][N], const T (&Tr)[N][N], T (&Ra)[N][N], T (&Ta)[N][N]) {
  T mi[N][N], t1[N][N], x[N][N], t1a[N][N], nR[N][N];
  mul(R, R, mi);
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) mi[i][j] = (i == j ? T(1) : T(0)) - mi[i][j];
  inv_inplace(mi);
  mul(Tr, mi, t1);
  // t1_ad = Ra (R T)^T + Ta T^T
  mul(R, Tr, x);
  mul_nt(Ra, x, t1a);
  add_mul_nt(Ta, Tr, t1a);
  // nR = Ra + t1^T Ra T^T
  mul_tn(t1, Ra, x);
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) nR[i][j] = Ra[i][j];
  add_mul_nt(x, Tr, nR);
  // nT = (t1 R)^T Ra + t1^T Ta + t1_ad mi^T
  T nT[N][N];
  mul(t1, R, x);
  mul_tn(x, Ra, nT);
  add_mul_tn(t1, Ta, nT);
  add_mul_nt(t1a, mi, nT);
  // mi_ad = T^T t1_ad;  rr_ad = mi^T mi_ad mi^T;  nR += rr_ad R^T + R^T rr_ad
  mul_tn(Tr, t1a, x);        // mi_ad
  mul_tn(mi, x, t1);         // mi^T mi_ad
  mul_nt(t1, mi, x);         // rr_ad
  add_mul_nt(x, R, nR);
  add_mul_tn(R, x, nR);
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      Ra[i][j] = nR[i][j];
      Ta[i][j] = nT[i][j];
    }
}

// Adjoint of the pair of Horner series S_e = sum h^n / (2n)!, S_o = sum h^n / (2n+1)!
// (n = 0..MAX_N_AMOM) that share h: h_ad = sum_n sum_{a+b=n-1} (h^T)^a Q_n (h^T)^b,
// Q_n = Xe / (2n)! + Xo / (2n+1)!, by two interleaved Horner recursions.
template <class T, int N>
__device__ void series_pair_ad(const T (&h)[N][N], const T (&Xe)[N][N], const T (&Xo)[N][N], T (&Z)[N][N]) {
  T ce[MAX_N_AMOM + 1], co[MAX_N_AMOM + 1];
  ce[0] = co[0] = T(1);
#pragma unroll
  for (int n = 1; n <= MAX_N_AMOM; ++n) {
    ce[n] = ce[n - 1] / T((2 * n) * (2 * n - 1));
    co[n] = co[n - 1] / T((2 * n) * (2 * n + 1));
  }
  T Y[N][N], x[N][N];
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) Z[i][j] = Y[i][j] = ce[MAX_N_AMOM] * Xe[i][j] + co[MAX_N_AMOM] * Xo[i][j];
  for (int a = MAX_N_AMOM - 2; a >= 0; --a) {
    mul_nt(Y, h, x);   // Y h^T
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) Y[i][j] = ce[a + 1] * Xe[i][j] + co[a + 1] * Xo[i][j] + x[i][j];
    mul_tn(h, Z, x);   // h^T Z
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) Z[i][j] = Y[i][j] + x[i][j];
  }
}

// Reverse of the series start: R0_ad, T0_ad -> apb_ad, amb_ad, t_ad.
template <class T, int N>
__device__ void mexp_start_ad(const T (&apb)[N][N], const T (&amb)[N][N], T t, const T (&R0a)[N][N],
                              const T (&T0a)[N][N], T (&apb_ad)[N][N], T (&amb_ad)[N][N], T& t_ad) {
  T hh[N][N], ff[N][N];
  mul(amb, apb, hh);
  mul(apb, amb, ff);
  const T t2 = t * t;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      hh[i][j] *= t2;
      ff[i][j] *= t2;
    }
  // S = s1 + s2 - t P3 - t P4, Sr = s1 - s2 + t P3 - t P4 with P3 = apb s3, P4 = amb s4
  T Sad[N][N], Srad[N][N];
  {
    T S[N][N], Sr[N][N], x[N][N], y[N][N];
    series(hh, S, 0);              // s1
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) Sr[i][j] = S[i][j];
    series(ff, x, 0);              // s2
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) {
        S[i][j] += x[i][j];
        Sr[i][j] -= x[i][j];
      }
    series(hh, x, 1);              // s3
    mul(apb, x, y);                // P3
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) {
        S[i][j] -= t * y[i][j];
        Sr[i][j] += t * y[i][j];
      }
    series(ff, x, 1);              // s4
    mul(amb, x, y);                // P4
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) {
        S[i][j] -= t * y[i][j];
        Sr[i][j] -= t * y[i][j];
      }
    inv_inplace(S);                // Si
    // Si_ad = 2 T0_ad + Sr^T R0_ad;  Sr_ad = R0_ad Si^T;  S_ad = -Si^T Si_ad Si^T
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) x[i][j] = T(2) * T0a[i][j];
    add_mul_tn(Sr, R0a, x);
    mul_nt(R0a, S, Srad);
    mul_tn(S, x, y);
    mul_nt(y, S, Sad);
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) Sad[i][j] = -Sad[i][j];
  }
  // s1_ad = S_ad + Sr_ad, s2_ad = S_ad - Sr_ad; (t P3)_ad = Sr_ad - S_ad, (t P4)_ad = -S_ad - Sr_ad
  T c3[N][N], c4[N][N];
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      const T a = Sad[i][j], b = Srad[i][j];
      c3[i][j] = b - a;
      c4[i][j] = -a - b;
      Sad[i][j] = a + b;    // s1_ad
      Srad[i][j] = a - b;   // s2_ad
    }
  T x[N][N], s34[N][N];
  // P3 = apb s3: t_ad += <P3, c3>; apb_ad = t c3 s3^T; s3_ad = t apb^T c3
  series(hh, s34, 1);
  mul(apb, s34, x);
  t_ad += dot(x, c3);
  mul_nt(c3, s34, apb_ad);
  mul_tn(apb, c3, x);
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      apb_ad[i][j] *= t;
      c3[i][j] = t * x[i][j];   // s3_ad
    }
  // P4 = amb s4
  series(ff, s34, 1);
  mul(amb, s34, x);
  t_ad += dot(x, c4);
  mul_nt(c4, s34, amb_ad);
  mul_tn(amb, c4, x);
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      amb_ad[i][j] *= t;
      c4[i][j] = t * x[i][j];   // s4_ad
    }
  // series adjoints -> hh_ad, ff_ad; hh = HH t^2, HH = amb apb; ff = FF t^2, FF = apb amb
  series_pair_ad(hh, Sad, c3, x);     // x = hh_ad
  t_ad += T(2) / t * dot(hh, x);
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) x[i][j] *= t2;   // HH_ad
  add_mul_nt(x, apb, amb_ad);
  add_mul_tn(amb, x, apb_ad);
  series_pair_ad(ff, Srad, c4, x);    // x = ff_ad
  t_ad += T(2) / t * dot(ff, x);
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) x[i][j] *= t2;   // FF_ad
  add_mul_nt(x, amb, apb_ad);
  add_mul_tn(apb, x, amb_ad);
}

// ------------------------------------------------ table derivatives
// 4-point weights of one axis and their derivatives with respect to the
// abscissa (LPoly_AD); zero derivative outside the table, where the point is
// clamped.
struct W4D {
  W4 v, d;
};
__device__ inline W4D weights_d(const double* x, int n, double xi) {
  const bool clamped = xi > x[n - 1] || xi < x[0];
  const LP p = lpoly(x, n, xi);
  xi = fmax(fmin(xi, x[n - 1]), x[0]);
  const double q0 = x[p.i1], q1 = x[p.i1 + 1], q2 = x[p.i1 + 2], q3 = x[p.i1 + 3];
  const double d0 = xi - q0, d1 = xi - q1, d2 = xi - q2, d3 = xi - q3;
  const double dl0 = (d1 + d2) / ((q0 - q1) * (q0 - q2));
  const double dl1 = (d0 + d2) / (-(q0 - q1) * (q1 - q2));
  const double dl2 = (d0 + d1) / ((q0 - q2) * (q1 - q2));
  const double dr0 = (d2 + d3) / ((q1 - q2) * (q1 - q3));
  const double dr1 = (d1 + d3) / (-(q1 - q2) * (q2 - q3));
  const double dr2 = (d1 + d2) / ((q1 - q3) * (q2 - q3));
  const double dwr = (xi < q1 || xi > q2) ? 0.0 : 1.0 / (q2 - q1);
  W4D r;
  r.v = W4{p.i1, {p.wl * p.l0, p.wl * p.l1 + p.wr * p.r0, p.wl * p.l2 + p.wr * p.r1, p.wr * p.r2}};
  r.d.i1 = p.i1;
  r.d.w[0] = -dwr * p.l0 + p.wl * dl0;
  r.d.w[1] = -dwr * p.l1 + p.wl * dl1 + dwr * p.r0 + p.wr * dr0;
  r.d.w[2] = -dwr * p.l2 + p.wl * dl2 + dwr * p.r1 + p.wr * dr1;
  r.d.w[3] = dwr * p.r2 + p.wr * dr2;
  if (clamped)
    for (int a = 0; a < 4; ++a) r.d.w[a] = 0.0;
  return r;
}

// One cloud of one layer: Get_Cloud_Opt_MW/IR values after the
// CRTM_Compute_CloudScatter clean-up, their radius (r) and temperature (t)
// derivatives, and the phase-coefficient table addressing.
struct CloudLU {
  double ke, w, ke_r, ke_t, w_r, w_t;
  bool ke_on, w_on, pc_on;     // which derivatives survive the clean-up
  const double* pct;
  int64_t sl, s2, s3;
  int dims;
  W4 u, v, t, dv, dt;
};

// Options%Include_Scattering = .FALSE.: the extinction is the absorption
// ke (1 - w), taken before the clean-up, with its derivatives; nothing
// scatters (Get_Cloud_Opt_IR/MW and their adjoints,
// src/AtmScatter/CRTM_CloudScatter.f90:994-997, 1213-1221, 1815-1822;
// Get_Aerosol_Opt, CRTM_AerosolScatter.f90:845-848, 1260-1268).  Types
// without an albedo table carry w = 0 and keep their ke.
__device__ inline void absorption_only(CloudLU& c) {
  const double a = 1.0 - c.w;
  c.ke_r = c.ke_r * a - c.ke * c.w_r;
  c.ke_t = c.ke_t * a - c.ke * c.w_t;
  c.ke *= a;
  c.w = 0.0;
  c.w_r = 0.0;
  c.w_t = 0.0;
  c.pct = nullptr;
}

__device__ CloudLU cloud_lookup(const DevCloudCoeff& cc, bool mw, double freq, int type, double reff, double temp,
                                int loff, bool absorb) {
  CloudLU c{};
  c.w = 0.0;
  if (mw) {
    const int nF = cc.nfm, nR = cc.nrm;
    c.u = weights(cc.f_mw, nF, freq);
    if (type == WATER_CLOUD || type == RAIN_CLOUD) {
      const int64_t sT = (int64_t)nF * nR;
      const W4D td = weights_d(cc.temp, cc.nt, temp);
      c.t = td.v;
      c.dt = td.d;
      if (type == WATER_CLOUD) {
        const double* z = cc.ke_l + c.u.i1 + sT * c.t.i1;
        c.ke = tab2(z, sT, c.u, c.t);
        c.ke_t = tab2(z, sT, c.u, c.dt);
      } else {
        const W4D rd = weights_d(cc.r_mw, nR, reff);
        c.v = rd.v;
        c.dv = rd.d;
        const int64_t o = c.u.i1 + nF * c.v.i1 + sT * c.t.i1;
        c.ke = tab3(cc.ke_l + o, nF, sT, c.u, c.v, c.t);
        c.ke_r = tab3(cc.ke_l + o, nF, sT, c.u, c.dv, c.t);
        c.ke_t = tab3(cc.ke_l + o, nF, sT, c.u, c.v, c.dt);
        c.w = tab3(cc.w_l + o, nF, sT, c.u, c.v, c.t);
        c.w_r = tab3(cc.w_l + o, nF, sT, c.u, c.dv, c.t);
        c.w_t = tab3(cc.w_l + o, nF, sT, c.u, c.v, c.dt);
        c.pct = cc.pc_l + o + sT * cc.nt * loff;
        c.sl = sT * cc.nt;
        c.s2 = nF;
        c.s3 = sT;
        c.dims = 3;
      }
    } else {
      const int kd = type == SNOW_CLOUD ? 0 : type == GRAUPEL_CLOUD ? 1 : 2;
      const int64_t sD = (int64_t)nF * nR;
      if (type == ICE_CLOUD) {
        c.ke = tab1(cc.ke_s + c.u.i1 + sD * kd, c.u);
      } else {
        const W4D rd = weights_d(cc.r_mw, nR, reff);
        c.v = rd.v;
        c.dv = rd.d;
        const int64_t o = c.u.i1 + nF * c.v.i1 + sD * kd;
        c.ke = tab2(cc.ke_s + o, nF, c.u, c.v);
        c.ke_r = tab2(cc.ke_s + o, nF, c.u, c.dv);
        c.w = tab2(cc.w_s + o, nF, c.u, c.v);
        c.w_r = tab2(cc.w_s + o, nF, c.u, c.dv);
        c.pct = cc.pc_s + o + sD * cc.ndm * loff;
        c.sl = sD * cc.ndm;
        c.s2 = nF;
        c.dims = 2;
      }
    }
  } else {
    const int nF = cc.nfi, nR = cc.nri;
    c.u = weights(cc.f_ir, nF, freq);
    const W4D rd = weights_d(cc.r_ir, nR, reff);
    c.v = rd.v;
    c.dv = rd.d;
    const int k = (type == WATER_CLOUD || type == RAIN_CLOUD) ? 0 : type == SNOW_CLOUD ? 1
                  : type == GRAUPEL_CLOUD ? 2 : 3;
    const int64_t sD = (int64_t)nF * nR;
    const int64_t o = c.u.i1 + nF * c.v.i1 + sD * k;
    c.ke = tab2(cc.ke_i + o, nF, c.u, c.v);
    c.ke_r = tab2(cc.ke_i + o, nF, c.u, c.dv);
    c.w = tab2(cc.w_i + o, nF, c.u, c.v);
    c.w_r = tab2(cc.w_i + o, nF, c.u, c.dv);
    c.pct = cc.pc_i + o + sD * cc.ndi * (loff - 1);
    c.sl = sD * cc.ndi;
    c.s2 = nF;
    c.dims = 2;
  }
  if (absorb) absorption_only(c);
  c.ke_on = c.ke > 0.0;
  c.w_on = c.ke_on && c.w > 0.0 && c.w < 1.0;
  c.pc_on = c.ke_on && c.w > 0.0 && c.pct != nullptr;
  if (!c.ke_on) { c.ke = 0.0; c.w = 0.0; }
  if (c.w <= 0.0) c.w = 0.0;
  if (c.w >= 1.0) c.w = 1.0;
  return c;
}
// One aerosol (Get_Aerosol_Opt, 'CRTM' scheme) with its radius derivatives,
// in the CloudLU form (no temperature dependence).
__device__ CloudLU aerosol_lookup(const DevAerosolCoeff& ac, double wavenumber, int k, double reff, int aoff,
                                  bool absorb) {
  CloudLU c{};
  c.u = weights(ac.freq, ac.nw, wavenumber);
  const W4D rd = weights_d(ac.reff + (int64_t)k * ac.nr, ac.nr, reff);
  c.v = rd.v;
  c.dv = rd.d;
  const int64_t sR = (int64_t)ac.nw * ac.nrh, sT = sR * ac.nr * ac.ns, sL = sT * ac.nt;
  const int64_t o = c.u.i1 + sR * c.v.i1 + sT * k;
  c.ke = tab2(ac.ke + o, sR, c.u, c.v);
  c.ke_r = tab2(ac.ke + o, sR, c.u, c.dv);
  c.w = tab2(ac.w + o, sR, c.u, c.v);
  c.w_r = tab2(ac.w + o, sR, c.u, c.dv);
  c.pct = ac.pc + o + sL * aoff;
  c.sl = sL;
  c.s2 = sR;
  c.dims = 2;
  if (absorb) absorption_only(c);
  c.ke_on = c.ke > 0.0;
  c.w_on = c.ke_on && c.w > 0.0 && c.w < 1.0;
  c.pc_on = c.ke_on && c.w > 0.0;
  if (!c.ke_on) { c.ke = 0.0; c.w = 0.0; }
  if (c.w <= 0.0) c.w = 0.0;
  if (c.w >= 1.0) c.w = 1.0;
  return c;
}

// phase coefficient l of a looked-up cloud and its r / t derivatives
__device__ inline double pc_val(const CloudLU& c, int l) {
  const double* z = c.pct + c.sl * l;
  return c.dims == 3 ? tab3(z, c.s2, c.s3, c.u, c.v, c.t) : tab2(z, c.s2, c.u, c.v);
}
__device__ inline double pc_dr(const CloudLU& c, int l) {
  const double* z = c.pct + c.sl * l;
  return c.dims == 3 ? tab3(z, c.s2, c.s3, c.u, c.dv, c.t) : tab2(z, c.s2, c.u, c.dv);
}
__device__ inline double pc_dt(const CloudLU& c, int l) {
  const double* z = c.pct + c.sl * l;
  return c.dims == 3 ? tab3(z, c.s2, c.s3, c.u, c.v, c.dt) : 0.0;
}

// Phase matrices of a layer in the forward's form (fs = Pff + Pbb,
// fd = Pff - Pbb after Normalize_Phase, full), with the quantities the
// adjoint needs: raw clipped sums, row scales, row norms, the sensor-column
// divisor and the clip masks.
template <class T, int NZ>
struct PhaseFwd {
  T rs[NZ][NZ], rd[NZ][NZ];     // raw fs, fd (upper triangle)
  T sc[NZ], nf[NZ];
  T nfl;
  uint64_t clip_f, clip_b;      // bit i*NZ+j: Pff / Pbb clipped to the floor
};

template <class T, int NZ>
__device__ void phase_forward(int ns, int nl, const T (&plr)[NL_AS - 1], const T (&leg)[NZ][NL_AS - 1],
                              const T (&wt)[NZ], PhaseFwd<T, NZ>& pf, T (&fs)[NZ][NZ], T (&fd)[NZ][NZ]) {
  pf.clip_f = pf.clip_b = 0;
#pragma unroll
  for (int i = 0; i < NZ; ++i)
#pragma unroll
    for (int j = i; j < NZ; ++j) {
      T e = T(0), o = T(0);
#pragma unroll
      for (int q = 0; q < NL_AS - 1; ++q) {
        if (q >= nl) break;
        const T tq = plr[q] * leg[i][q] * leg[j][q];
        if (q & 1) o += tq; else e += tq;
      }
      T ff = e + o, bb = e - o;
      if (ff < T(0)) { ff = T(PHASE_THRESHOLD); pf.clip_f |= 1ull << (i * NZ + j); }
      if (bb < T(0)) { bb = T(PHASE_THRESHOLD); pf.clip_b |= 1ull << (i * NZ + j); }
      pf.rs[i][j] = fs[i][j] = ff + bb;
      pf.rd[i][j] = fd[i][j] = ff - bb;
    }
  T sum_prev = T(0);
#pragma unroll
  for (int i = 0; i < NZ; ++i) {
    pf.sc[i] = T(1);
    pf.nf[i] = T(1);
    if (i < ns) {
      T nf = T(0);
#pragma unroll
      for (int j = i; j < NZ; ++j) nf += fs[i][j] * wt[j];
      const T scale = (T(1) - sum_prev) / nf;
      pf.nf[i] = nf;
      pf.sc[i] = scale;
#pragma unroll
      for (int j = i; j < NZ; ++j) {
        fs[i][j] *= scale;
        fd[i][j] *= scale;
      }
      T si = T(0);
      if (i < NZ - 1) {
#pragma unroll
        for (int j = 0; j <= i; ++j) si += fs[j][i + 1] * wt[j];
      }
      sum_prev = si;
    }
  }
  pf.nfl = sum_prev;
  if (ns < NZ) {
    const T inf = T(1) / sum_prev;
#pragma unroll
    for (int j = 0; j < NZ; ++j) {
      fs[j][NZ - 1] *= inf;
      fd[j][NZ - 1] *= inf;
    }
  }
#pragma unroll
  for (int i = 0; i < NZ; ++i)
#pragma unroll
    for (int j = 0; j < i; ++j) {
      fs[i][j] = fs[j][i];
      fd[i][j] = fd[j][i];
    }
}

// Adjoint of phase_forward: fa, da (adjoints of the final fs, fd; consumed)
// -> pl_ad[q], q < nl.  fs holds the forward's final matrices.
template <class T, int NZ>
__device__ void phase_ad(int ns, int nl, const T (&leg)[NZ][NL_AS - 1], const T (&wt)[NZ],
                         const PhaseFwd<T, NZ>& pf, const T (&fs)[NZ][NZ], const T (&fd)[NZ][NZ], T (&fa)[NZ][NZ],
                         T (&da)[NZ][NZ], T (&pl_ad)[NL_AS - 1]) {
  // the lower triangle is a copy of the upper one
#pragma unroll
  for (int i = 0; i < NZ; ++i)
#pragma unroll
    for (int j = 0; j < i; ++j) {
      fa[j][i] += fa[i][j];
      da[j][i] += da[i][j];
    }
  T S_ad[NZ];
#pragma unroll
  for (int i = 0; i < NZ; ++i) S_ad[i] = T(0);
  if (ns < NZ) {
    T nfl_ad = T(0);
    const T inf = T(1) / pf.nfl;
#pragma unroll
    for (int j = 0; j < NZ; ++j) {
      nfl_ad -= (fa[j][NZ - 1] * fs[j][NZ - 1] + da[j][NZ - 1] * fd[j][NZ - 1]) * inf;
      fa[j][NZ - 1] *= inf;
      da[j][NZ - 1] *= inf;
    }
#pragma unroll
    for (int i = 0; i < NZ; ++i)
      if (i == ns - 1) S_ad[i] += nfl_ad;
  }
  // rows in reverse: fs_i,j = raw_i,j * sc_i (j >= i), sc_i = (1 - S_{i-1}) / nf_i,
  // nf_i = sum_{j>=i} raw_fs_i,j w_j, S_i = sum_{j<=i} fs_j,i+1 w_j.  The raw
  // adjoints overwrite fa / da in place.
#pragma unroll
  for (int i = NZ - 1; i >= 0; --i) {
    if (i < ns) {
      if (i < NZ - 1) {
#pragma unroll
        for (int j = 0; j <= i; ++j) fa[j][i + 1] += S_ad[i] * wt[j];
      }
      T sc_ad = T(0);
#pragma unroll
      for (int j = i; j < NZ; ++j) sc_ad += fa[i][j] * pf.rs[i][j] + da[i][j] * pf.rd[i][j];
      const T nf_ad = -sc_ad * pf.sc[i] / pf.nf[i];
#pragma unroll
      for (int j = 0; j < NZ; ++j)
        if (i >= 1 && j == i - 1) S_ad[j] -= sc_ad / pf.nf[i];
#pragma unroll
      for (int j = i; j < NZ; ++j) {
        fa[i][j] = fa[i][j] * pf.sc[i] + nf_ad * wt[j];
        da[i][j] = da[i][j] * pf.sc[i];
      }
    }
  }
  // raw: fs = Pff + Pbb, fd = Pff - Pbb with Pff = e + o, Pbb = e - o (clipped)
#pragma unroll
  for (int i = 0; i < NZ; ++i)
#pragma unroll
    for (int j = i; j < NZ; ++j) {
      if (i >= ns && j != i) continue;   // sensor row: only its diagonal is raw
      T ffa = fa[i][j] + da[i][j], bba = fa[i][j] - da[i][j];
      if (pf.clip_f >> (i * NZ + j) & 1) ffa = T(0);
      if (pf.clip_b >> (i * NZ + j) & 1) bba = T(0);
      const T ea = ffa + bba, oa = ffa - bba;
#pragma unroll
      for (int q = 0; q < NL_AS - 1; ++q) {
        if (q >= nl) break;
        pl_ad[q] += ((q & 1) ? oa : ea) * leg[i][q] * leg[j][q];
      }
    }
}
}  // namespace

// =========================================================================
// forward + reverse adding sweep: one thread per (profile, channel)
// =========================================================================
template <class RT, int NZ>
__global__ void __launch_bounds__(256) allsky_sweep_ad_kernel(DevSensor s, DevBatch be, DevClouds cl, CloudWork cw,
                                                               DevCloudCoeff cc, LayerScratch<RT> sc,
                                                               LayerScratch<RT> sa, const RT* tau_g,
                                                               const double* geo, DevOut o, AllskyOut ao, RT* tape,
                                                               int64_t tape_lanes, int te) {
  const int64_t n = (int64_t)be.M * s.L;
  const int64_t lane0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;   // this thread's tape column
  auto tp = [&](int k, int e) -> RT& { return tape[((int64_t)k * te + e) * tape_lanes + lane0]; };
  RT* tau_s = reinterpret_cast<RT*>(o.ks.tau);
  RT* b_s = reinterpret_cast<RT*>(o.ks.b);
  RT* t_s = reinterpret_cast<RT*>(o.ks.t);
  for (int64_t t = lane0; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const int m = (int)(t / s.L), l = (int)(t - (int64_t)m * s.L);
    if (cw.cl_count[m] == 0 || cw.status[m] != 0 || cw.aer_err[m] != 0) continue;
    const double cosu = geo[(int64_t)m * 4 + 1];
    const int nl = legendre_terms(profile_streams(cl, m), cw.max_reff[m], s.wavenumber[l]);
    if (nl > NL_AS - 1) continue;
    const AngleSet as = angle_set(cc.gauss, (nl + 2) / 2, cosu);
    if (as.nz != NZ) continue;
    const int K = be.n_layers[m];
    const int64_t le = be.loff[m];
    const int lg = s.chan_offset + l;
    const int64_t ol = (int64_t)m * be.Ltot + lg;
    const int64_t sl = (int64_t)m * s.L + l;
    const int64_t st = sc.n_items;
    bool scat = false;
    for (int k = 0; k < K; ++k) {
      const int c = cw.cl_slot[le + k];
      if (c >= 0 && (double)sc.v[st * sc.oSsa() + (int64_t)c * s.L + l] > BS_THRESHOLD) scat = true;
    }
    const RT c1 = (RT)s.c1[l], c2 = (RT)s.c2[l], bc1 = (RT)s.bc1[l], bc2 = (RT)s.bc2[l];
    RT tes;
    const RT Bs = planck(c1, c2, bc1, bc2, geo[(int64_t)m * 4 + 2], tes);
    const RT eps = (RT)be.emissivity[ol];
    const RT cosmic = (RT)s.cosmic[l];
    auto layer_tau = [&](int k) -> RT {
      const int c = cw.cl_slot[le + k];
      return c >= 0 ? sc.v[st * sc.oTau() + (int64_t)c * s.L + l] : tau_g[(le + k) * s.L + l];
    };
    auto planck_k = [&](int k) -> RT {
      RT te_;
      return planck(c1, c2, bc1, bc2, be.t[le + k], te_);
    };
    // layer optical-depth adjoint: the cloud-optics adjoint of a cloudy layer,
    // else the gas optical depth seed
    auto put_tau_ad = [&](int k, RT v) {
      const int c = cw.cl_slot[le + k];
      if (c >= 0) {
        const int64_t it = (int64_t)c * s.L + l;
        sa.v[st * sa.oTau() + it] = v;
        sa.v[st * sa.oSsa() + it] = RT(0);
#pragma unroll
        for (int q = 0; q < NL_AS - 1; ++q) sa.v[st * (sa.oP() + q) + it] = RT(0);
      } else {
        tau_s[(le + k) * s.L + l] = v;
      }
      t_s[(le + k) * s.L + l] = RT(0);
    };
    const RT iu = (RT)(1.0 / cosu);
    // ---------------- forward ----------------
    double r_cloudy;
    double up_rad = 0.0, down_rad = 0.0, sfc_b = 0.0;
    RT rdK = RT(0);
    const int ns = as.ns;
    const double* gmu = cc.gauss + ns * 20;
    RT imu[NZ];
#pragma unroll
    for (int i = 0; i < NZ; ++i) imu[i] = (RT)(1.0 / (i < ns ? gmu[i] : cosu));
    if (!scat) {
      // Emission with the cloud absorption; tape: downwelling radiance at the
      // top of each layer (slot 0), upwelling at its bottom (slot 1)
      RT rd = cosmic;
      for (int k = 0; k < K; ++k) {
        const RT tr = rexp<RT>(-layer_tau(k) * iu);
        tp(k, 0) = rd;
        rd = rd * tr + planck_k(k) * (RT(1) - tr);
      }
      rdK = rd;
      RT ru = eps * Bs + (RT(1) - eps) * rd;
      RT uacc = RT(0);
      for (int k = K - 1; k >= 0; --k) {
        const RT tr = rexp<RT>(-layer_tau(k) * iu);
        const RT src = planck_k(k) * (RT(1) - tr);
        tp(k, 1) = ru;
        ru = ru * tr + src;
        uacc = uacc * tr + src;
      }
      r_cloudy = (double)ru;
      up_rad = (double)uacc;
      down_rad = (double)rd;
      sfc_b = (double)Bs;
    } else {
      RT rup[NZ][NZ], iup[NZ];
#pragma unroll
      for (int i = 0; i < NZ; ++i) {
#pragma unroll
        for (int j = 0; j < NZ; ++j) rup[i][j] = i == j ? RT(1) - eps : RT(0);
        iup[i] = eps * Bs;
      }
      for (int k = K - 1; k >= 0; --k) {
#pragma unroll
        for (int i = 0; i < NZ; ++i) {
#pragma unroll
          for (int j = 0; j < NZ; ++j) tp(k, i * NZ + j) = rup[i][j];
          tp(k, NZ * NZ + i) = iup[i];
        }
        const RT B = planck_k(k);
        const int c = cw.cl_slot[le + k];
        const int64_t it = (int64_t)c * s.L + l;
        if (c >= 0 && (double)sc.v[st * sc.oSsa() + it] > BS_THRESHOLD) {
          RT rl[NZ][NZ], tl[NZ][NZ], src[NZ];
#pragma unroll
          for (int i = 0; i < NZ; ++i) {
#pragma unroll
            for (int j = 0; j < NZ; ++j) {
              rl[i][j] = sc.v[st * (sc.oR() + i * sc.nzc + j) + it];
              tl[i][j] = sc.v[st * (sc.oT() + i * sc.nzc + j) + it];
            }
            src[i] = sc.v[st * (sc.oE() + i) + it] * B;
          }
          double g[NZ][NZ];
#pragma unroll
          for (int i = 0; i < NZ; ++i)
#pragma unroll
            for (int j = 0; j < NZ; ++j) {
              double sum = 0.0;
#pragma unroll
              for (int q = 0; q < NZ; ++q) sum += (double)rup[i][q] * (double)rl[q][j];
              g[i][j] = (i == j ? 1.0 : 0.0) - sum;
            }
          inv_inplace(g);
          RT gt[NZ][NZ];
#pragma unroll
          for (int i = 0; i < NZ; ++i)
#pragma unroll
            for (int j = 0; j < NZ; ++j) {
              double sum = 0.0;
#pragma unroll
              for (int q = 0; q < NZ; ++q) sum += (double)tl[i][q] * g[q][j];
              gt[i][j] = (RT)sum;
            }
          RT v[NZ];
#pragma unroll
          for (int i = 0; i < NZ; ++i) {
            RT sum = iup[i];
#pragma unroll
            for (int q = 0; q < NZ; ++q) sum += rup[i][q] * src[q];
            v[i] = sum;
          }
#pragma unroll
          for (int i = 0; i < NZ; ++i) {
            RT sum = src[i];
#pragma unroll
            for (int q = 0; q < NZ; ++q) sum += gt[i][q] * v[q];
            iup[i] = sum;
          }
          mul_left_inplace(rup, tl);
          mul_left_inplace(gt, tl);
#pragma unroll
          for (int i = 0; i < NZ; ++i)
#pragma unroll
            for (int j = 0; j < NZ; ++j) rup[i][j] = rl[i][j] + tl[i][j];
        } else {
          const RT tau = layer_tau(k);
          RT tr[NZ], sr[NZ];
#pragma unroll
          for (int i = 0; i < NZ; ++i) {
            tr[i] = rexp<RT>(-tau * imu[i]);
            sr[i] = B * (RT(1) - tr[i]);
          }
          RT nv[NZ];
#pragma unroll
          for (int i = 0; i < NZ; ++i) {
            RT sum = iup[i];
#pragma unroll
            for (int q = 0; q < NZ; ++q) sum += rup[i][q] * sr[q];
            nv[i] = sr[i] + tr[i] * sum;
          }
#pragma unroll
          for (int i = 0; i < NZ; ++i) {
            iup[i] = nv[i];
#pragma unroll
            for (int j = 0; j < NZ; ++j) rup[i][j] = tr[i] * rup[i][j] * tr[j];
          }
        }
      }
      RT top = RT(0), rs = RT(0);
#pragma unroll
      for (int i = 0; i < NZ; ++i)
        if (i == as.isat) {
          top = iup[i];
#pragma unroll
          for (int j = 0; j < NZ; ++j) rs += rup[i][j];
        }
      r_cloudy = (double)(top + rs * cosmic);
    }
    const double rclr = o.radiance[ol];
    const bool frac = cw.frac[m] != 0;
    const double tcc = cw.tcc[m];
    const double R = frac ? (1.0 - tcc) * rclr + tcc * r_cloudy : r_cloudy;
    const double c1d = s.c1[l], c2d = s.c2[l], bc1d = s.bc1[l], bc2d = s.bc2[l];
    const double btv = (c2d / log(c1d / R + 1.0) - bc1d) / bc2d;
    if (ao.r_clear) ao.r_clear[ol] = frac ? rclr : R;
    if (ao.tb_clear) ao.tb_clear[ol] = frac ? (c2d / log(c1d / rclr + 1.0) - bc1d) / bc2d : btv;
    if (ao.tcc) ao.tcc[ol] = tcc;
    o.radiance[ol] = R;
    o.bt[ol] = btv;
    if (o.up_radiance) o.up_radiance[ol] = up_rad;
    if (o.down_radiance) o.down_radiance[ol] = down_rad;
    if (o.sfc_planck) o.sfc_planck[ol] = sfc_b;
    if (o.sfc_emis) o.sfc_emis[ol] = be.emissivity[ol];
    // ---------------- adjoint seeds (Pre_Process_RTSolution_K) ----------------
    const double arg = c1d / R + 1.0;
    const double lga = log(arg);
    const double F = c1d * c2d / (bc2d * arg * (R * lga) * (R * lga));
    const double R_ad = (o.rad_k ? o.rad_k[ol] : 0.0) + F * (o.bt_k ? o.bt_k[ol] : 0.0);
    if (o.rad_k) o.rad_k[ol] = R_ad;
    o.ks.rad[sl] = frac ? (1.0 - tcc) * R_ad : 0.0;
    o.ks.tcc_ad[sl] = frac ? (r_cloudy - rclr) * R_ad : 0.0;
    const RT rc_ad = (RT)(frac ? tcc * R_ad : R_ad);
    RT eps_ad = RT(0), bs_ad = RT(0);
    // ---------------- reverse ----------------
    if (!scat) {
      // upwelling part, top-down; partial B and tau adjoints parked on the tape
      RT up_ad = rc_ad;
      for (int k = 0; k < K; ++k) {
        const RT tr = rexp<RT>(-layer_tau(k) * iu);
        const RT B = planck_k(k);
        const RT t_ad = up_ad * (tp(k, 1) - B);
        tp(k, 2) = up_ad * (RT(1) - tr);
        up_ad *= tr;
        tp(k, 3) = -t_ad * tr * iu;
      }
      eps_ad = up_ad * (Bs - rdK);
      bs_ad = up_ad * eps;
      RT dn_ad = (RT(1) - eps) * up_ad;
      for (int k = K - 1; k >= 0; --k) {
        const RT tr = rexp<RT>(-layer_tau(k) * iu);
        const RT B = planck_k(k);
        const RT t_ad = dn_ad * (tp(k, 0) - B);
        const RT b_ad = tp(k, 2) + dn_ad * (RT(1) - tr);
        dn_ad *= tr;
        const RT tau_ad = tp(k, 3) - t_ad * tr * iu;
        b_s[(le + k) * s.L + l] = b_ad;
        put_tau_ad(k, tau_ad);
      }
    } else {
      RT rup_ad[NZ][NZ], iup_ad[NZ];
#pragma unroll
      for (int i = 0; i < NZ; ++i) {
        iup_ad[i] = i == as.isat ? rc_ad : RT(0);
#pragma unroll
        for (int j = 0; j < NZ; ++j) rup_ad[i][j] = i == as.isat ? rc_ad * cosmic : RT(0);
      }
      for (int k = 0; k < K; ++k) {
        RT rup[NZ][NZ], iup[NZ];
#pragma unroll
        for (int i = 0; i < NZ; ++i) {
#pragma unroll
          for (int j = 0; j < NZ; ++j) rup[i][j] = tp(k, i * NZ + j);
          iup[i] = tp(k, NZ * NZ + i);
        }
        const RT B = planck_k(k);
        const int c = cw.cl_slot[le + k];
        const int64_t it = (int64_t)c * s.L + l;
        if (c >= 0 && (double)sc.v[st * sc.oSsa() + it] > BS_THRESHOLD) {
          RT rl[NZ][NZ], tl[NZ][NZ], e[NZ], src[NZ];
#pragma unroll
          for (int i = 0; i < NZ; ++i) {
#pragma unroll
            for (int j = 0; j < NZ; ++j) {
              rl[i][j] = sc.v[st * (sc.oR() + i * sc.nzc + j) + it];
              tl[i][j] = sc.v[st * (sc.oT() + i * sc.nzc + j) + it];
            }
            e[i] = sc.v[st * (sc.oE() + i) + it];
            src[i] = e[i] * B;
          }
          double g[NZ][NZ];
#pragma unroll
          for (int i = 0; i < NZ; ++i)
#pragma unroll
            for (int j = 0; j < NZ; ++j) {
              double sum = 0.0;
#pragma unroll
              for (int q = 0; q < NZ; ++q) sum += (double)rup[i][q] * (double)rl[q][j];
              g[i][j] = (i == j ? 1.0 : 0.0) - sum;
            }
          inv_inplace(g);
          RT gt[NZ][NZ], v[NZ];
#pragma unroll
          for (int i = 0; i < NZ; ++i)
#pragma unroll
            for (int j = 0; j < NZ; ++j) {
              double sum = 0.0;
#pragma unroll
              for (int q = 0; q < NZ; ++q) sum += (double)tl[i][q] * g[q][j];
              gt[i][j] = (RT)sum;
            }
#pragma unroll
          for (int i = 0; i < NZ; ++i) {
            RT sum = iup[i];
#pragma unroll
            for (int q = 0; q < NZ; ++q) sum += rup[i][q] * src[q];
            v[i] = sum;
          }
          // I_up' = src + gt v;  R_up' = rl + gt (rup tl)
          RT gt_ad[NZ][NZ], x[NZ][NZ], y[NZ][NZ];
#pragma unroll
          for (int i = 0; i < NZ; ++i)
#pragma unroll
            for (int j = 0; j < NZ; ++j) gt_ad[i][j] = iup_ad[i] * v[j];
          mul(rup, tl, x);                    // rt = rup tl
          add_mul_nt(rup_ad, x, gt_ad);       // gt_ad += R'_ad rt^T
          mul_tn(gt, rup_ad, x);              // rt_ad = gt^T R'_ad
          RT tl_ad[NZ][NZ], nrup_ad[NZ][NZ];
          mul_tn(rup, x, tl_ad);              // tl_ad = rup^T rt_ad
          mul_nt(x, tl, nrup_ad);             // rup_ad = rt_ad tl^T
          RT v_ad[NZ], src_ad[NZ];
#pragma unroll
          for (int j = 0; j < NZ; ++j) {
            RT sum = RT(0);
#pragma unroll
            for (int i = 0; i < NZ; ++i) sum += gt[i][j] * iup_ad[i];
            v_ad[j] = sum;
          }
#pragma unroll
          for (int i = 0; i < NZ; ++i)
#pragma unroll
            for (int j = 0; j < NZ; ++j) nrup_ad[i][j] += v_ad[i] * src[j];
#pragma unroll
          for (int j = 0; j < NZ; ++j) {
            RT sum = iup_ad[j];
#pragma unroll
            for (int i = 0; i < NZ; ++i) sum += rup[i][j] * v_ad[i];
            src_ad[j] = sum;
          }
          // gt = tl g:  tl_ad += gt_ad g^T;  g_ad = tl^T gt_ad;  P_ad = g^T g_ad g^T (P = rup rl)
          double gad[NZ][NZ], pad[NZ][NZ];
#pragma unroll
          for (int i = 0; i < NZ; ++i)
#pragma unroll
            for (int j = 0; j < NZ; ++j) {
              double sum = 0.0, sum2 = 0.0;
#pragma unroll
              for (int q = 0; q < NZ; ++q) {
                sum += (double)gt_ad[i][q] * g[j][q];
                sum2 += (double)tl[q][i] * (double)gt_ad[q][j];
              }
              tl_ad[i][j] += (RT)sum;
              gad[i][j] = sum2;
            }
          mul_tn(g, gad, pad);
          mul_nt(pad, g, gad);
#pragma unroll
          for (int i = 0; i < NZ; ++i)
#pragma unroll
            for (int j = 0; j < NZ; ++j) y[i][j] = (RT)gad[i][j];   // P_ad
          add_mul_nt(y, rl, nrup_ad);         // rup_ad += P_ad rl^T
          add_mul_tn(rup, y, rup_ad);         // rl_ad = R'_ad + rup^T P_ad
          RT b_ad = RT(0);
#pragma unroll
          for (int i = 0; i < NZ; ++i) {
            b_ad += src_ad[i] * e[i];
#pragma unroll
            for (int j = 0; j < NZ; ++j) {
              sa.v[st * (sa.oR() + i * sa.nzc + j) + it] = rup_ad[i][j];
              sa.v[st * (sa.oT() + i * sa.nzc + j) + it] = tl_ad[i][j];
            }
            sa.v[st * (sa.oE() + i) + it] = src_ad[i] * B;
          }
          b_s[(le + k) * s.L + l] = b_ad;
          sa.v[st * sa.oTau() + it] = RT(0);
          sa.v[st * sa.oSsa() + it] = RT(0);
#pragma unroll
          for (int q = 0; q < NL_AS - 1; ++q) sa.v[st * (sa.oP() + q) + it] = RT(0);
          t_s[(le + k) * s.L + l] = RT(0);
#pragma unroll
          for (int i = 0; i < NZ; ++i) {
            iup_ad[i] = v_ad[i];
#pragma unroll
            for (int j = 0; j < NZ; ++j) rup_ad[i][j] = nrup_ad[i][j];
          }
        } else {
          const RT tau = layer_tau(k);
          RT tr[NZ], sr[NZ], w_[NZ];
#pragma unroll
          for (int i = 0; i < NZ; ++i) {
            tr[i] = rexp<RT>(-tau * imu[i]);
            sr[i] = B * (RT(1) - tr[i]);
          }
#pragma unroll
          for (int i = 0; i < NZ; ++i) {
            RT sum = iup[i];
#pragma unroll
            for (int q = 0; q < NZ; ++q) sum += rup[i][q] * sr[q];
            w_[i] = sum;
          }
          RT s_ad[NZ], t_ad[NZ], w_ad[NZ];
#pragma unroll
          for (int i = 0; i < NZ; ++i) {
            s_ad[i] = iup_ad[i];
            t_ad[i] = iup_ad[i] * w_[i];
            w_ad[i] = iup_ad[i] * tr[i];
          }
#pragma unroll
          for (int i = 0; i < NZ; ++i)
#pragma unroll
            for (int j = 0; j < NZ; ++j) {
              const RT ra = rup_ad[i][j];
              s_ad[j] += w_ad[i] * rup[i][j];
              t_ad[i] += ra * rup[i][j] * tr[j];
              t_ad[j] += ra * tr[i] * rup[i][j];
              rup_ad[i][j] = w_ad[i] * sr[j] + ra * tr[i] * tr[j];
            }
          RT b_ad = RT(0), tau_ad = RT(0);
#pragma unroll
          for (int i = 0; i < NZ; ++i) {
            b_ad += s_ad[i] * (RT(1) - tr[i]);
            tau_ad -= (t_ad[i] - s_ad[i] * B) * tr[i] * imu[i];
            iup_ad[i] = w_ad[i];
          }
          b_s[(le + k) * s.L + l] = b_ad;
          put_tau_ad(k, tau_ad);
        }
      }
#pragma unroll
      for (int i = 0; i < NZ; ++i) {
        eps_ad += Bs * iup_ad[i] - rup_ad[i][i];
        bs_ad += eps * iup_ad[i];
      }
    }
    o.ks.emis[sl] = (double)eps_ad;
    o.ks.bs[sl] = (double)bs_ad;
  }
}

// =========================================================================
// layer R/T adjoint: one thread per scattering (cloudy layer, channel)
// =========================================================================
template <class RT, int NZ>
__global__ void __launch_bounds__(256) cloud_rt_ad_kernel(DevSensor s, CloudWork cw, DevClouds cl, DevCloudCoeff cc,
                                                           LayerScratch<RT> sc, LayerScratch<RT> sa,
                                                           const double* geo, int n_items) {
  const int n_live = live_items(cw, s.L, n_items);
  for (int item = blockIdx.x * blockDim.x + threadIdx.x; item < n_live; item += gridDim.x * blockDim.x) {
    const int c = item / s.L, l = item - c * s.L;
    const int m = cw.cl_desc[c].x;
    if (cw.status[m] != 0 || cw.aer_err[m] != 0) continue;
    const int nl = legendre_terms(profile_streams(cl, m), cw.max_reff[m], s.wavenumber[l]);
    if (nl > NL_AS - 1) continue;
    const double cosu = geo[(int64_t)m * 4 + 1];
    const AngleSet as = angle_set(cc.gauss, (nl + 2) / 2, cosu);
    if (as.nz != NZ) continue;
    const RT* fw = sc.v + item;
    RT* ad = sa.v + item;
    const int64_t st = sc.n_items;
    const RT om = fw[st * sc.oSsa()];
    if (!((double)om > BS_THRESHOLD)) continue;
    const RT tau = fw[st * sc.oTau()];
    RT plr[NL_AS - 1];
#pragma unroll
    for (int q = 0; q < NL_AS - 1; ++q) plr[q] = fw[st * (sc.oP() + q)];
    const int ns = as.ns;
    const double* gmu = cc.gauss + ns * 20;
    RT mu[NZ], wt[NZ], leg[NZ][NL_AS - 1];
#pragma unroll
    for (int i = 0; i < NZ; ++i) {
      const bool node = i < ns;
      mu[i] = node ? (RT)gmu[i] : (RT)cosu;
      wt[i] = node ? (RT)gmu[10 + i] : RT(0);
      leg[i][0] = RT(1);
      leg[i][1] = mu[i];
#pragma unroll
      for (int q = 1; q < NL_AS - 2; ++q)
        leg[i][q + 1] = (RT(2 * q + 1) * mu[i] * leg[i][q] - RT(q) * leg[i][q - 1]) / RT(q + 1);
    }
    // adjoints of R, T and e = 1 - Thermal_C
    RT Ra[NZ][NZ], Ta[NZ][NZ], ea[NZ];
#pragma unroll
    for (int i = 0; i < NZ; ++i) {
#pragma unroll
      for (int j = 0; j < NZ; ++j) {
        Ra[i][j] = ad[st * (sa.oR() + i * sa.nzc + j)];
        Ta[i][j] = ad[st * (sa.oT() + i * sa.nzc + j)];
      }
      ea[i] = ad[st * (sa.oE() + i)];
    }
    PhaseFwd<RT, NZ> pf;
    RT fs[NZ][NZ], fd[NZ][NZ];
    phase_forward<RT, NZ>(ns, nl, plr, leg, wt, pf, fs, fd);
    RT tau_ad = RT(0), om_ad = RT(0), pl_ad[NL_AS - 1];
#pragma unroll
    for (int q = 0; q < NL_AS - 1; ++q) pl_ad[q] = RT(0);
    RT fa[NZ][NZ], da[NZ][NZ];
    if ((double)tau < DELTA_OPTICAL_DEPTH) {
      // R = c Pbb w, T = c Pff w + I (1 - tau/mu), c = tau om / mu_i; e = 1 - sum_j (R + T)
      const RT sv = tau * om;
      RT s_ad = RT(0);
#pragma unroll
      for (int i = 0; i < NZ; ++i) {
        const RT cst = sv / mu[i];
        RT c_ad = RT(0);
#pragma unroll
        for (int j = 0; j < NZ; ++j) {
          const RT ra = Ra[i][j] - ea[i], ta = Ta[i][j] - ea[i];
          const RT pbb = RT(0.5) * (fs[i][j] - fd[i][j]), pff = RT(0.5) * (fs[i][j] + fd[i][j]);
          const RT bba = ra * cst * wt[j], ffa = ta * cst * wt[j];
          fa[i][j] = RT(0.5) * (ffa + bba);
          da[i][j] = RT(0.5) * (ffa - bba);
          c_ad += (ra * pbb + ta * pff) * wt[j];
          if (i == j) tau_ad -= ta / mu[i];
        }
        s_ad += c_ad / mu[i];
      }
      tau_ad += s_ad * om;
      om_ad += s_ad * tau;
    } else {
      // e_i = 1 - sum_{j<ns} (T_ij + R_ij) [- T_nn for the sensor row]
#pragma unroll
      for (int i = 0; i < NZ; ++i) {
#pragma unroll
        for (int j = 0; j < NZ; ++j)
          if (j < ns) {
            Ra[i][j] -= ea[i];
            Ta[i][j] -= ea[i];
          }
        if (i == NZ - 1 && NZ == ns + 1) Ta[NZ - 1][NZ - 1] -= ea[i];
      }
      const RT sv = om < RT(MAX_ALBEDO) ? om : RT(MAX_ALBEDO);
      RT apb[NZ][NZ], amb[NZ][NZ];
#pragma unroll
      for (int i = 0; i < NZ; ++i) {
        const RT cst = sv / mu[i], dg = RT(1) / mu[i];
#pragma unroll
        for (int j = 0; j < NZ; ++j) {
          apb[i][j] = cst * fs[i][j] * wt[j] - (i == j ? dg : RT(0));
          amb[i][j] = cst * fd[i][j] * wt[j] - (i == j ? dg : RT(0));
        }
      }
      const int nd = doublings(amb, apb, tau);
      const RT t = (RT)((double)tau / ldexp(1.0, nd));
      if (nd > 0) {
        RT R0[NZ][NZ], T0[NZ][NZ];
        mexp_start(apb, amb, t, R0, T0);
        for (int d = nd - 1; d >= 0; --d) {
          RT Rd[NZ][NZ], Td[NZ][NZ];
#pragma unroll
          for (int i = 0; i < NZ; ++i)
#pragma unroll
            for (int j = 0; j < NZ; ++j) {
              Rd[i][j] = R0[i][j];
              Td[i][j] = T0[i][j];
            }
          for (int q = 0; q < d; ++q) double_step(Rd, Td);
          double_step_ad(Rd, Td, Ra, Ta);
        }
      }
      RT t_ad = RT(0);
      mexp_start_ad(apb, amb, t, Ra, Ta, fa, da, t_ad);   // fa = apb_ad, da = amb_ad
      tau_ad += t_ad / (RT)ldexp(1.0, nd);
      // A +- B = c fs|fd w - I / mu with c = s / mu_i
      RT s_ad = RT(0);
#pragma unroll
      for (int i = 0; i < NZ; ++i) {
        const RT cst = sv / mu[i];
#pragma unroll
        for (int j = 0; j < NZ; ++j) {
          s_ad += (fa[i][j] * fs[i][j] + da[i][j] * fd[i][j]) * wt[j] / mu[i];
          fa[i][j] *= cst * wt[j];
          da[i][j] *= cst * wt[j];
        }
      }
      if (om < RT(MAX_ALBEDO)) om_ad += s_ad;
    }
    phase_ad<RT, NZ>(ns, nl, leg, wt, pf, fs, fd, fa, da, pl_ad);
    ad[st * sa.oTau()] = tau_ad;
    ad[st * sa.oSsa()] = om_ad;
#pragma unroll
    for (int q = 0; q < NL_AS - 1; ++q) ad[st * (sa.oP() + q)] = q < nl ? pl_ad[q] : RT(0);
  }
}

// =========================================================================
// optics adjoint: one thread per (cloudy layer, channel)
// =========================================================================
template <class RT, class KT>
__global__ void __launch_bounds__(256) cloud_optics_ad_kernel(DevSensor s, DevBatch be, DevClouds cl, CloudWork cw,
                                                               DevCloudCoeff cc, LayerScratch<RT> sa,
                                                               const RT* tau_g, DevOut o, CloudKOut ko,
                                                               int n_items) {
  const bool mw = s.sensor_type == 1;
  KT* wck = reinterpret_cast<KT*>(ko.wc_k);
  KT* rek = reinterpret_cast<KT*>(ko.reff_k);
  const int n_live = live_items(cw, s.L, n_items);
  for (int item = blockIdx.x * blockDim.x + threadIdx.x; item < n_live; item += gridDim.x * blockDim.x) {
    const int c = item / s.L, l = item - c * s.L;
    const int2 d = cw.cl_desc[c];
    const int m = d.x, k = d.y;
    if (cw.status[m] != 0 || cw.aer_err[m] != 0) continue;
    const int nl = legendre_terms(profile_streams(cl, m), cw.max_reff[m], s.wavenumber[l]);
    if (nl > NL_W - 1) continue;
    const int64_t ik = be.loff[m] + k;
    const int loff = legendre_offset(nl);
    const double freq = mw ? s.frequency[l] : s.wavenumber[l];
    const double temp = be.t[ik];
    // forward accumulation (CRTM_Compute_CloudScatter)
    double od0 = (double)tau_g[ik * s.L + l], ssa0 = 0.0, pl0[NL_W];
#pragma unroll
    for (int q = 0; q < NL_W; ++q) pl0[q] = 0.0;
#pragma unroll 1
    for (int n = 0; n < cl.nc; ++n) {
      const double wc = cw.wc[n * be.NT + ik];
      if (!(wc > WATER_CONTENT_THRESHOLD)) continue;
      const CloudLU lu = cloud_lookup(cc, mw, freq, cl.type[(int64_t)m * cl.nc + n], cw.reff[n * be.NT + ik], temp,
                                      loff, cl.no_scat != 0);
      od0 += lu.ke * wc;
      const double bs = wc * lu.ke * lu.w;
      ssa0 += bs;
      if (bs != 0.0 && lu.pct) {
#pragma unroll 1
        for (int q = 1; q <= nl; ++q) {
          const double pv = pc_val(lu, q);
#pragma unroll
          for (int qq = 1; qq < NL_W; ++qq)
            if (qq == q) pl0[qq] += bs * pv;
        }
      }
    }
    const int ku = k - be.n_extra[m];
    const bool aer = !mw && cl.na > 0 && ku >= 0 && aerosol_offset(nl) >= 0;
    const int aoff = aer ? aerosol_offset(nl) : 0;
    const int64_t iu = be.out_loff[m] + ku;
    if (aer) {
#pragma unroll 1
      for (int n = 0; n < cl.na; ++n) {
        const int ty = cl.atype[(int64_t)m * cl.na + n];
        const double conc = cl.aconc[n * be.out_NT + iu];
        const int ka = ty != 0 ? aerosol_index(cc.aero, ty) : -1;
        if (ka < 0 || !(conc > AEROSOL_CONTENT_THRESHOLD)) continue;
        const CloudLU lu = aerosol_lookup(cc.aero, s.wavenumber[l], ka, cl.areff[n * be.out_NT + iu], aoff,
                                                 cl.no_scat != 0);
        od0 += lu.ke * conc;
        const double bs = conc * lu.ke * lu.w;
        ssa0 += bs;
        if (bs != 0.0) {
#pragma unroll 1
          for (int q = 1; q <= nl; ++q) {
            const double pv = pc_val(lu, q);
#pragma unroll
            for (int qq = 1; qq < NL_W; ++qq)
              if (qq == q) pl0[qq] += bs * pv;
          }
        }
      }
    }
    // CRTM_AtmOptics_Combine_AD
    const int64_t st = sa.n_items;
    const double tau_ad = (double)sa.v[st * sa.oTau() + item];
    const double ssa_ad = (double)sa.v[st * sa.oSsa() + item];
    double od_ad, ssa_pre_ad, plpre_ad[NL_W];
#pragma unroll
    for (int q = 0; q < NL_W; ++q) plpre_ad[q] = 0.0;
    if (ssa0 > BS_THRESHOLD) {
      double plnl = 0.0;
#pragma unroll
      for (int q = 1; q < NL_W; ++q)
        if (q == nl) plnl = pl0[q];
      const double w = ssa0 / od0, f = plnl / ssa0, den = 1.0 - f * w;
      const double f_ad = -w * od0 * tau_ad + ssa_ad * w * (w - 1.0) / (den * den);
      const double w_ad = -f * od0 * tau_ad + ssa_ad * (1.0 - f) / (den * den);
      od_ad = den * tau_ad;
      ssa_pre_ad = 0.0;
#pragma unroll
      for (int q = 1; q < NL_W; ++q) {
        double pa = q < nl ? (double)sa.v[st * (sa.oP() + q) + item] : 0.0;
        if (q == nl) pa = f_ad;
        if (q > nl) pa = 0.0;
        plpre_ad[q] = pa / ssa0;
        ssa_pre_ad -= pa * pl0[q] / (ssa0 * ssa0);
      }
      ssa_pre_ad += w_ad / od0;
      od_ad -= w_ad * ssa0 / (od0 * od0);
    } else {
      od_ad = tau_ad;
      ssa_pre_ad = ssa_ad;
    }
    reinterpret_cast<RT*>(o.ks.tau)[ik * s.L + l] = (RT)od_ad;
    // CRTM_Compute_CloudScatter_AD, per cloud
    const int nx = be.n_extra[m];
    const int64_t row = be.out_loff[m] + (k - nx);
    const int lg = s.chan_offset + l;
    double t_sum = 0.0;
#pragma unroll 1
    for (int n = 0; n < cl.nc; ++n) {
      const double wc = cw.wc[n * be.NT + ik];
      if (!(wc > WATER_CONTENT_THRESHOLD)) continue;
      const CloudLU lu = cloud_lookup(cc, mw, freq, cl.type[(int64_t)m * cl.nc + n], cw.reff[n * be.NT + ik], temp,
                                      loff, cl.no_scat != 0);
      const double bs = wc * lu.ke * lu.w;
      double s0 = 0.0, sr = 0.0, stt = 0.0;
      if (lu.pc_on) {
#pragma unroll 1
        for (int q = 1; q <= nl; ++q) {
          double pa = 0.0;
#pragma unroll
          for (int qq = 1; qq < NL_W; ++qq)
            if (qq == q) pa = plpre_ad[qq];
          s0 += pa * pc_val(lu, q);
          sr += pa * pc_dr(lu, q);
          stt += pa * pc_dt(lu, q);
        }
      }
      const double bs_ad = ssa_pre_ad + s0;
      const double ke_ad = od_ad * wc + bs_ad * wc * lu.w;
      const double wc_ad = od_ad * lu.ke + bs_ad * lu.ke * lu.w;
      const double w_ad = bs_ad * wc * lu.ke;
      double r_ad = bs * sr, t_ad = bs * stt;
      if (lu.ke_on) {
        r_ad += ke_ad * lu.ke_r;
        t_ad += ke_ad * lu.ke_t;
      }
      if (lu.w_on) {
        r_ad += w_ad * lu.w_r;
        t_ad += w_ad * lu.w_t;
      }
      if (ko.reduce) {
        // adjoint: channel sums [nc][out_NT]; the cover rescaling of the
        // water content (wc / tcc) is applied here, and its share of the
        // cover adjoint goes to the (profile, channel) seed that
        // cloud_cover_ad_kernel reads after this kernel
        const int64_t ri = (int64_t)n * be.out_NT + row;
        double wa = wc_ad;
        if (cw.frac[m] && cw.tcc[m] > MIN_COVERAGE_THRESHOLD) {
          unsafeAtomicAdd(o.ks.tcc_ad + (int64_t)m * s.L + l, -wc_ad * wc / cw.tcc[m]);
          wa = wc_ad / cw.tcc[m];
        }
        unsafeAtomicAdd(reinterpret_cast<double*>(ko.wc_k) + ri, wa);
        unsafeAtomicAdd(reinterpret_cast<double*>(ko.reff_k) + ri, r_ad);
      } else {
        const int64_t oi = ((int64_t)n * be.out_NT + row) * be.Ltot + lg;
        wck[oi] = (KT)wc_ad;
        rek[oi] = (KT)r_ad;
      }
      t_sum += t_ad;
    }
    // CRTM_Compute_AerosolScatter_AD (src/AtmScatter/CRTM_AerosolScatter.f90:614-751)
    if (aer) {
      KT* ack = reinterpret_cast<KT*>(ko.aconc_k);
      KT* ark = reinterpret_cast<KT*>(ko.areff_k);
#pragma unroll 1
      for (int n = 0; n < cl.na; ++n) {
        const int ty = cl.atype[(int64_t)m * cl.na + n];
        const double conc = cl.aconc[n * be.out_NT + iu];
        const int ka = ty != 0 ? aerosol_index(cc.aero, ty) : -1;
        if (ka < 0 || !(conc > AEROSOL_CONTENT_THRESHOLD)) continue;
        const CloudLU lu = aerosol_lookup(cc.aero, s.wavenumber[l], ka, cl.areff[n * be.out_NT + iu], aoff,
                                                 cl.no_scat != 0);
        const double bs = conc * lu.ke * lu.w;
        double s0 = 0.0, sr = 0.0;
        if (lu.pc_on) {
#pragma unroll 1
          for (int q = 1; q <= nl; ++q) {
            double pa = 0.0;
#pragma unroll
            for (int qq = 1; qq < NL_W; ++qq)
              if (qq == q) pa = plpre_ad[qq];
            s0 += pa * pc_val(lu, q);
            sr += pa * pc_dr(lu, q);
          }
        }
        const double bs_ad = ssa_pre_ad + s0;
        const double ke_ad = od_ad * conc + bs_ad * conc * lu.w;
        const double c_ad = od_ad * lu.ke + bs_ad * lu.ke * lu.w;
        const double w_ad = bs_ad * conc * lu.ke;
        double r_ad = bs * sr;
        if (lu.ke_on) r_ad += ke_ad * lu.ke_r;
        if (lu.w_on) r_ad += w_ad * lu.w_r;
        if (ko.reduce) {
          const int64_t ri = (int64_t)n * be.out_NT + row;
          unsafeAtomicAdd(reinterpret_cast<double*>(ko.aconc_k) + ri, c_ad);
          unsafeAtomicAdd(reinterpret_cast<double*>(ko.areff_k) + ri, r_ad);
        } else {
          const int64_t oi = ((int64_t)n * be.out_NT + row) * be.Ltot + lg;
          ack[oi] = (KT)c_ad;
          ark[oi] = (KT)r_ad;
        }
      }
    }
    reinterpret_cast<RT*>(o.ks.t)[ik * s.L + l] = (RT)t_sum;
  }
}

// =========================================================================
// cloud-cover adjoint: one thread per (profile, channel) of fractional profiles
// =========================================================================
template <class KT>
__global__ void __launch_bounds__(256) cloud_cover_ad_kernel(DevSensor s, DevBatch be, DevClouds cl, CloudWork cw,
                                                              DevOut o, CloudKOut ko) {
  KT* wck = reinterpret_cast<KT*>(ko.wc_k);
  KT* cfk = reinterpret_cast<KT*>(ko.cf_k);
  const int64_t n = (int64_t)be.M * s.L;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const int m = (int)(t / s.L), l = (int)(t - (int64_t)m * s.L);
    if (cw.cl_count[m] == 0 || cw.status[m] != 0 || cw.aer_err[m] != 0 || !cw.frac[m]) continue;
    const int K = be.n_layers[m], nx = be.n_extra[m], nc = cl.nc;
    const int64_t le = be.loff[m], lu = be.out_loff[m];
    const int64_t NTe = be.NT, NTu = be.out_NT, Lt = be.Ltot;
    const int lg = s.chan_offset + l;
    const double tcc = cw.tcc[m];
    const bool scaled = tcc > MIN_COVERAGE_THRESHOLD;
    auto cf_at = [&](int kk) -> double {   // cleaned fraction, extended layer index
      const int ku = kk - nx;
      if (ku < 0 || !cl.cf) return 0.0;
      const double f = cl.cf[lu + ku];
      return f < MIN_COVERAGE_THRESHOLD ? 0.0 : f;
    };
    auto wcu = [&](int nn, int ku) -> double {   // caller's water content after the clean-up
      const int ty = cl.type[(int64_t)m * nc + nn];
      if (ty < WATER_CLOUD || ty > HAIL_CLOUD || !cl.cf || cl.cf[lu + ku] < MIN_COVERAGE_THRESHOLD) return 0.0;
      return cl.wc[nn * NTu + lu + ku];
    };
    // wc_scaled = wc / tcc (adjoint calls: applied by cloud_optics_ad_kernel)
    double tot_ad = o.ks.tcc_ad[(int64_t)m * s.L + l];
    for (int nn = 0; nn < nc && !ko.reduce; ++nn)
      for (int ku = 0; ku < K - nx; ++ku) {
        const double wcs = cw.wc[nn * NTe + le + ku + nx];
        if (!(wcs > WATER_CONTENT_THRESHOLD)) continue;
        const int64_t oi = ((int64_t)nn * NTu + lu + ku) * Lt + lg;
        const double a = (double)wck[oi];
        if (scaled) {
          tot_ad -= a * wcs / tcc;
          wck[oi] = (KT)(a / tcc);
        }
      }
    auto add_cf = [&](int kk, double v) {
      const int ku = kk - nx;
      if (ku < 0) return;
      if (ko.reduce) {
        unsafeAtomicAdd(reinterpret_cast<double*>(cfk) + lu + ku, v);
        return;
      }
      KT* p = cfk + (lu + ku) * Lt + lg;
      *p = (KT)((double)*p + v);
    };
    switch (cl.overlap) {
      case 1: {   // maximum: the first layer holding the largest fraction
        int arg = 0;
        double cov = cf_at(0);
        for (int kk = 1; kk < K; ++kk)
          if (cf_at(kk) > cov) {
            cov = cf_at(kk);
            arg = kk;
          }
        add_cf(arg, tot_ad);
        break;
      }
      case 2: {   // random
        for (int kk = 0; kk < K; ++kk) {
          if (!(cf_at(kk) > MIN_COVERAGE_THRESHOLD)) continue;
          double prod = 1.0;
          for (int j = 0; j < K; ++j)
            if (j != kk && cf_at(j) > MIN_COVERAGE_THRESHOLD) prod *= 1.0 - cf_at(j);
          add_cf(kk, tot_ad * prod);
        }
        break;
      }
      case 3: {   // maximum-random: reverse of the product recursion, earlier products recomputed
        auto prod_to = [&](int j) -> double {
          double p = 1.0 - cf_at(0);
          for (int kk = 1; kk <= j; ++kk) {
            const double fp = cf_at(kk - 1), f = cf_at(kk);
            const double mc = 1.0 - fmax(fp, f);
            if (p == 0.0 && 1.0 - fp == 0.0) p = mc;
            else p = p * (mc / (1.0 - fp));
          }
          return p;
        };
        double p_ad = -tot_ad;
        for (int kk = K - 1; kk >= 1 && p_ad != 0.0; --kk) {
          const double pp = prod_to(kk - 1);
          const double fp = cf_at(kk - 1), f = cf_at(kk);
          const int am = fp > f ? kk - 1 : kk;
          const double mc = 1.0 - fmax(fp, f);
          if (pp == 0.0 && 1.0 - fp == 0.0) {
            add_cf(am, -p_ad);
            p_ad = 0.0;
            break;
          }
          const double dd = 1.0 - fp;
          add_cf(am, -p_ad * pp / dd);
          add_cf(kk - 1, p_ad * pp * mc / (dd * dd));
          p_ad = p_ad * mc / dd;
        }
        add_cf(0, -p_ad);
        break;
      }
      case 5:     // overcast: constant cover
        break;
      default: {  // average
        double ws = 0.0;
        for (int ku = 0; ku < K - nx; ++ku)
          for (int nn = 0; nn < nc; ++nn) {
            const double w = wcu(nn, ku);
            if (w > WATER_CONTENT_THRESHOLD) ws += w;
          }
        if (ws > 0.0)
          for (int ku = 0; ku < K - nx; ++ku) {
            double lwc = 0.0;
            for (int nn = 0; nn < nc; ++nn) {
              const double w = wcu(nn, ku);
              if (w > WATER_CONTENT_THRESHOLD) {
                lwc += w;
                const double add = tot_ad * (cf_at(ku + nx) - tcc) / ws;
                if (ko.reduce) {
                  unsafeAtomicAdd(reinterpret_cast<double*>(wck) + (int64_t)nn * NTu + lu + ku, add);
                } else {
                  const int64_t oi = ((int64_t)nn * NTu + lu + ku) * Lt + lg;
                  wck[oi] = (KT)((double)wck[oi] + add);
                }
              }
            }
            add_cf(ku + nx, tot_ad * lwc / ws);
          }
      }
    }
  }
}

// ------------------------------------------------------------------ launchers
template <class RT>
void launch_cloud_optics(const DevSensor&, const DevBatch&, const DevClouds&, const CloudWork&, const DevCloudCoeff&,
                         const LayerScratch<RT>&, const RT*, int, hipStream_t);
template <class RT>
void launch_cloud_rt(int nz, const DevSensor&, const CloudWork&, const DevClouds&, const DevCloudCoeff&,
                     const LayerScratch<RT>&, const double*, int, hipStream_t);

// wavefront kernels for 9 and 10 angles (allsky_wide.hip, allsky_wide_ad.hip)
template <class RT>
void launch_cloud_rt_wide(int nz, const DevSensor&, const CloudWork&, const DevClouds&, const DevCloudCoeff&,
                          const LayerScratch<RT>&, const double*, int, hipStream_t);
template <class RT>
void launch_allsky_ad_wide(int nz, const DevSensor&, const DevBatch&, const DevClouds&, const CloudWork&,
                           const DevCloudCoeff&, const LayerScratch<RT>&, const LayerScratch<RT>&, const RT*,
                           const double*, int, const DevOut&, const AllskyOut&, RT*, int, RT*, hipStream_t);

static int blocks_for(int64_t n, int cap) {
  const int64_t want = (n + 255) / 256;
  return (int)(want < cap ? (want > 0 ? want : 1) : cap);
}

int allsky_sweep_ad_blocks(int64_t n_lanes) { return blocks_for(n_lanes, 512); }

template <class RT, class KT>
void launch_allsky_k(const DevSensor& s, const DevBatch& be, const DevClouds& cl, const CloudWork& cw,
                     const DevCloudCoeff& cc, const LayerScratch<RT>& sc, const LayerScratch<RT>& sa,
                     const RT* tau_g, const double* geo, int n_items, const int* nz_list, int n_nz, const DevOut& o,
                     const AllskyOut& ao, const CloudKOut& ko, RT* tape, int64_t tape_lanes, int te, RT* wtape,
                     int kmax, RT* dtape, hipStream_t st) {
  const int lb = blocks_for(n_items, 8192);
  const int sb = allsky_sweep_ad_blocks((int64_t)be.M * s.L);
  if ((int64_t)sb * 256 != tape_lanes) return;   // the host sizes the tape for this grid
  launch_cloud_optics<RT>(s, be, cl, cw, cc, sc, tau_g, n_items, st);
  for (int q = 0; q < n_nz; ++q) {
    if (nz_list[q] > sc.nzc) {   // wavefront kernels
      launch_cloud_rt_wide<RT>(nz_list[q], s, cw, cl, cc, sc, geo, n_items, st);
      launch_allsky_ad_wide<RT>(nz_list[q], s, be, cl, cw, cc, sc, sa, tau_g, geo, n_items, o, ao, wtape, kmax, dtape,
                                st);
      continue;
    }
    launch_cloud_rt<RT>(nz_list[q], s, cw, cl, cc, sc, geo, n_items, st);
    switch (nz_list[q]) {
#define CRTM_NZ_CASE(N)                                                                                        \
  case N:                                                                                                      \
    hipLaunchKernelGGL((allsky_sweep_ad_kernel<RT, N>), dim3(sb), dim3(256), 0, st, s, be, cl, cw, cc, sc, sa,  \
                       tau_g, geo, o, ao, tape, tape_lanes, te);                                               \
    if (n_items > 0)                                                                                           \
      hipLaunchKernelGGL((cloud_rt_ad_kernel<RT, N>), dim3(lb), dim3(256), 0, st, s, cw, cl, cc, sc, sa, geo,   \
                         n_items);                                                                             \
    break;
      CRTM_NZ_CASE(2)
      CRTM_NZ_CASE(3)
      CRTM_NZ_CASE(4)
      CRTM_NZ_CASE(5)
      CRTM_NZ_CASE(6)
#undef CRTM_NZ_CASE
      default:
        break;
    }
  }
  if (n_items > 0)
    hipLaunchKernelGGL((cloud_optics_ad_kernel<RT, KT>), dim3(lb), dim3(256), 0, st, s, be, cl, cw, cc, sa, tau_g,
                       o, ko, n_items);
  hipLaunchKernelGGL((cloud_cover_ad_kernel<KT>), dim3(blocks_for((int64_t)be.M * s.L, 8192)), dim3(256), 0, st, s,
                     be, cl, cw, o, ko);
}

#define CRTM_INST_K(RT, KT)                                                                                   \
  template void launch_allsky_k<RT, KT>(const DevSensor&, const DevBatch&, const DevClouds&, const CloudWork&, \
                                        const DevCloudCoeff&, const LayerScratch<RT>&, const LayerScratch<RT>&, \
                                        const RT*, const double*, int, const int*, int, const DevOut&,        \
                                        const AllskyOut&, const CloudKOut&, RT*, int64_t, int, RT*, int, RT*,   \
                                        hipStream_t);
CRTM_INST_K(float, float)
CRTM_INST_K(float, double)
CRTM_INST_K(double, float)
CRTM_INST_K(double, double)
#undef CRTM_INST_K

}  // namespace crtm
