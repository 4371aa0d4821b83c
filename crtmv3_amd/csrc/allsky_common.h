// allsky_common.h -- device helpers shared by the all-sky forward kernels
// (allsky.hip) and their adjoints (allsky_ad.hip): thresholds, stream choice,
// angle sets, CloudCoeff table interpolation, per-lane small-matrix algebra
// and Matrix_Exp.
#pragma once
#include "crtm_device.h"

namespace crtm {

namespace {
constexpr double WATER_CONTENT_THRESHOLD = 1.0e-6;  // src/CRTM_Parameters.f90:256
constexpr double AEROSOL_CONTENT_THRESHOLD = 1.0e-9;  // src/CRTM_Parameters.f90:255
constexpr double MIN_COVERAGE_THRESHOLD = 1.0e-12;  // src/CRTM_Parameters.f90:257
constexpr double MIN_COVERAGE_FLAG = 1.0e-6;        // src/Atmosphere/CRTM_Atmosphere.f90:164
constexpr double BS_THRESHOLD = 1.0e-10;            // src/CRTM_Parameters.f90:260
constexpr double ANGLE_THRESHOLD = 1.0e-7;          // src/RTSolution/RTV_Define.f90:70
constexpr double PHASE_THRESHOLD = 1.0e-7;          // RTV_Define.f90:74
constexpr double DELTA_OPTICAL_DEPTH = 1.0e-8;      // RTV_Define.f90:76
constexpr double MAX_ALBEDO = 0.999999;             // RTV_Define.f90:77
constexpr int MAX_N_AMOM = 5;                       // RTV_Define.f90:88
enum { WATER_CLOUD = 1, ICE_CLOUD, RAIN_CLOUD, SNOW_CLOUD, GRAUPEL_CLOUD, HAIL_CLOUD };

__device__ inline int legendre_offset(int nl) {
  return nl == 6 ? 5 : nl == 8 ? 12 : nl == 16 ? 21 : 0;
}

// n_Legendre_Terms of a (profile, channel): Options%n_Streams, or
// CRTM_Compute_nStreams (src/RTSolution/CRTM_RTSolution.f90:1204-1283)
__device__ inline int legendre_terms(int user, double max_reff, double wavenumber) {
  if (user > 0) return user;
  const double mie = 2.0 * M_PI * max_reff * wavenumber / 10000.0;
  return mie < 0.01 ? 2 : (mie < 1.0 ? 4 : 6);
}

// Options(m)%n_Streams of profile m
__device__ inline int profile_streams(const DevClouds& cl, int m) { return cl.ns_prof ? cl.ns_prof[m] : cl.n_streams; }

// Angle set of a (profile, channel): ns Gauss nodes plus the sensor angle
// unless it coincides with a node (Common_RTSolution.f90:255-281)
struct AngleSet {
  int ns, nz, isat;
};
__device__ inline AngleSet angle_set(const double* gauss, int ns, double cosu) {
  AngleSet a{ns, ns + 1, ns};
  const double* mu = gauss + ns * 20;
  for (int i = 0; i < ns; ++i)
    if (fabs(mu[i] - cosu) < ANGLE_THRESHOLD) {
      a.nz = ns;
      a.isat = i;
    }
  return a;
}

// ------------------------------------------------------------------ tables
// Four-point piecewise-quadratic Lagrange weights (Find_Random_Index + LPoly,
// src/Interpolation/CRTM_Interpolation.f90:785-805, 832-860, 1079-1196)
struct LP {
  int i1;
  double l0, l1, l2, r0, r1, r2, wl, wr;
};
__device__ inline LP lpoly(const double* x, int n, double xi) {
  xi = fmax(fmin(xi, x[n - 1]), x[0]);
  int k = 0;
  while (k < n && !(xi <= x[k])) ++k;
  LP p;
  p.i1 = min(max(k - 2, 0), n - 4);
  const double q0 = x[p.i1], q1 = x[p.i1 + 1], q2 = x[p.i1 + 2], q3 = x[p.i1 + 3];
  const double d0 = xi - q0, d1 = xi - q1, d2 = xi - q2, d3 = xi - q3;
  p.l0 = d1 * d2 / ((q0 - q1) * (q0 - q2));
  p.l1 = d0 * d2 / (-(q0 - q1) * (q1 - q2));
  p.l2 = d0 * d1 / ((q0 - q2) * (q1 - q2));
  p.r0 = d2 * d3 / ((q1 - q2) * (q1 - q3));
  p.r1 = d1 * d3 / (-(q1 - q2) * (q2 - q3));
  p.r2 = d1 * d2 / ((q1 - q3) * (q2 - q3));
  if (xi < q1) {
    p.wl = 1.0; p.wr = 0.0;
  } else if (xi > q2) {
    p.wl = 0.0; p.wr = 1.0;
  } else {
    p.wr = d1 / (q2 - q1);
    p.wl = 1.0 - p.wr;
  }
  return p;
}
// Combined 4-point weights of one axis: Interp_1D's blend of the left and
// right quadratics written as one weight per table point.
struct W4 {
  int i1;
  double w[4];
};
__device__ inline W4 weights(const double* x, int n, double xi) {
  const LP p = lpoly(x, n, xi);
  return W4{p.i1, {p.wl * p.l0, p.wl * p.l1 + p.wr * p.r0, p.wl * p.l2 + p.wr * p.r1, p.wr * p.r2}};
}
// Interp_1D/2D/3D (CRTM_Interpolation.f90:225-275) on the 4^nd sub-table at z
// with strides s1..s3; equal to the nested 1-D interpolations up to rounding.
__device__ inline double tab1(const double* z, const W4& u) {
  double r = 0.0;
#pragma unroll
  for (int a = 0; a < 4; ++a) r += u.w[a] * z[a];
  return r;
}
__device__ inline double tab2(const double* z, int64_t s2, const W4& u, const W4& v) {
  double r = 0.0;
#pragma unroll
  for (int b = 0; b < 4; ++b) r += v.w[b] * tab1(z + b * s2, u);
  return r;
}
__device__ inline double tab3(const double* z, int64_t s2, int64_t s3, const W4& u, const W4& v, const W4& w) {
  double r = 0.0;
#pragma unroll
  for (int c = 0; c < 4; ++c) r += w.w[c] * tab2(z + c * s3, s2, u, v);
  return r;
}

// Get_Cloud_Opt_MW / Get_Cloud_Opt_IR for one cloud of one layer, with the
// CRTM_Compute_CloudScatter clean-up and accumulation (ke <= 0, w <= 0,
// w >= 1) into od, bs and pl[1..nl].  absorb (Options%Include_Scattering =
// .FALSE.): the extinction becomes the absorption ke (1 - w) before the
// clean-up and nothing is scattered (src/AtmScatter/CRTM_CloudScatter.f90:
// 986-997, 1372-1386, 1395-1407).
__device__ void add_cloud(const DevCloudCoeff& cc, bool mw, double freq, int type, double reff, double temp,
                          double wc, int nl, int loff, bool absorb, double& od, double& bs_sum, double* pl) {
  double ke, w = 0.0;
  const double* pct = nullptr;   // phase-coefficient table of slot l at pct + l * sl
  int64_t sl = 0, s2 = 0, s3 = 0;
  W4 u{}, v{}, t{};
  int dims = 0;
  if (mw) {
    const int nF = cc.nfm, nR = cc.nrm;
    u = weights(cc.f_mw, nF, freq);
    if (type == WATER_CLOUD || type == RAIN_CLOUD) {
      const int64_t sT = (int64_t)nF * nR;
      t = weights(cc.temp, cc.nt, temp);
      if (type == WATER_CLOUD) {   // radius index 1, absorption only
        ke = tab2(cc.ke_l + u.i1 + sT * t.i1, sT, u, t);
      } else {
        v = weights(cc.r_mw, nR, reff);
        const int64_t o = u.i1 + nF * v.i1 + sT * t.i1;
        ke = tab3(cc.ke_l + o, nF, sT, u, v, t);
        w = tab3(cc.w_l + o, nF, sT, u, v, t);
        pct = cc.pc_l + o + sT * cc.nt * loff;
        sl = sT * cc.nt;
        s2 = nF;
        s3 = sT;
        dims = 3;
      }
    } else {
      const int kd = type == SNOW_CLOUD ? 0 : type == GRAUPEL_CLOUD ? 1 : 2;   // density index 1, 2, 3
      const int64_t sD = (int64_t)nF * nR;
      if (type == ICE_CLOUD) {   // radius index 1, absorption only
        ke = tab1(cc.ke_s + u.i1 + sD * kd, u);
      } else {
        v = weights(cc.r_mw, nR, reff);
        const int64_t o = u.i1 + nF * v.i1 + sD * kd;
        ke = tab2(cc.ke_s + o, nF, u, v);
        w = tab2(cc.w_s + o, nF, u, v);
        pct = cc.pc_s + o + sD * cc.ndm * loff;
        sl = sD * cc.ndm;
        s2 = nF;
        dims = 2;
      }
    }
  } else {
    const int nF = cc.nfi, nR = cc.nri;
    u = weights(cc.f_ir, nF, freq);
    v = weights(cc.r_ir, nR, reff);
    const int k = (type == WATER_CLOUD || type == RAIN_CLOUD) ? 0 : type == SNOW_CLOUD ? 1
                  : type == GRAUPEL_CLOUD ? 2 : 3;
    const int64_t sD = (int64_t)nF * nR;
    const int64_t o = u.i1 + nF * v.i1 + sD * k;
    ke = tab2(cc.ke_i + o, nF, u, v);
    w = tab2(cc.w_i + o, nF, u, v);
    pct = cc.pc_i + o + sD * cc.ndi * (loff - 1);   // Legendre dimension starts at 1
    sl = sD * cc.ndi;
    s2 = nF;
    dims = 2;
  }
  if (absorb) {   // types without an albedo table keep w = 0 and their ke
    ke *= 1.0 - w;
    w = 0.0;
    pct = nullptr;
  }
  if (ke <= 0.0) { ke = 0.0; w = 0.0; }
  if (w <= 0.0) w = 0.0;
  if (w >= 1.0) w = 1.0;
  od += ke * wc;
  const double bs = wc * ke * w;
  bs_sum += bs;
  if (bs == 0.0 || !pct) return;
#pragma unroll 1
  for (int l = 1; l <= nl; ++l) {
    const double* z = pct + sl * l;
    pl[l] += bs * (dims == 3 ? tab3(z, s2, s3, u, v, t) : tab2(z, s2, u, v));
  }
}

// Legendre offset of the aerosol tables per stream set
// (src/AtmScatter/CRTM_AerosolScatter.f90:208-225); -1: no aerosol phase
// function at this stream count (two streams)
__device__ inline int aerosol_offset(int nl) {
  return nl == 4 ? 0 : nl == 6 ? 5 : nl == 8 ? 12 : nl == 16 ? 21 : -1;
}
// AerosolCoeff_typeID_to_index (0-based; -1 when the table lacks the type)
__device__ inline int aerosol_index(const DevAerosolCoeff& ac, int type) {
  int k = -1;
  for (int q = 0; q < ac.nt; ++q)
    if (ac.type_id[q] == type) k = q;
  return k;
}

// Get_Aerosol_Opt, 'CRTM' scheme (CRTM_AerosolScatter.f90:772-840): 2-D
// lookup over frequency and the type's radius grid at the first RH and
// size-variance entries, with the CRTM_Compute_AerosolScatter clean-up and
// accumulation into od, bs and pl[1..nl]; absorb as for add_cloud
// (CRTM_AerosolScatter.f90:838-848).
__device__ void add_aerosol(const DevAerosolCoeff& ac, double wavenumber, int k, double reff, double conc, int nl,
                            int aoff, bool absorb, double& od, double& bs_sum, double* pl) {
  const W4 u = weights(ac.freq, ac.nw, wavenumber);
  const W4 v = weights(ac.reff + (int64_t)k * ac.nr, ac.nr, reff);
  const int64_t sR = (int64_t)ac.nw * ac.nrh, sT = sR * ac.nr * ac.ns, sL = sT * ac.nt;
  const int64_t o = u.i1 + sR * v.i1 + sT * k;
  double ke = tab2(ac.ke + o, sR, u, v);
  double w = tab2(ac.w + o, sR, u, v);
  if (absorb) {
    ke *= 1.0 - w;
    w = 0.0;
  }
  if (ke <= 0.0) { ke = 0.0; w = 0.0; }
  if (w <= 0.0) w = 0.0;
  if (w >= 1.0) w = 1.0;
  od += ke * conc;
  const double bs = conc * ke * w;
  bs_sum += bs;
  if (bs == 0.0) return;
#pragma unroll 1
  for (int l = 1; l <= nl; ++l) pl[l] += bs * tab2(ac.pc + o + sL * (l + aoff), sR, u, v);
}

// ----------------------------------------------------------- small matrices
// In-place Gauss-Jordan inverse without pivoting: every matrix inverted here
// (the series denominator and I - R R) is diagonally dominant; the result
// equals matinv's (CRTM_Utility.f90:381-435) up to rounding.
template <class T, int N>
__device__ inline void inv_inplace(T (&b)[N][N]) {
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const T d = T(1) / b[k][k];
    T tmp[N];
#pragma unroll
    for (int i = 0; i < N; ++i) tmp[i] = b[i][k];
#pragma unroll
    for (int j = 0; j < N; ++j) {
      const T c = b[k][j] * d;
#pragma unroll
      for (int i = 0; i < N; ++i) b[i][j] -= tmp[i] * c;
      b[k][j] = c;
    }
#pragma unroll
    for (int i = 0; i < N; ++i) b[i][k] = -tmp[i] * d;
    b[k][k] = d;
  }
}

// b <- a * b, column by column (column j of the product needs only column j of b)
template <class T, int N>
__device__ inline void mul_left_inplace(const T (&a)[N][N], T (&b)[N][N]) {
#pragma unroll
  for (int j = 0; j < N; ++j) {
    T col[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      T s = T(0);
#pragma unroll
      for (int k = 0; k < N; ++k) s += a[i][k] * b[k][j];
      col[i] = s;
    }
#pragma unroll
    for (int i = 0; i < N; ++i) b[i][j] = col[i];
  }
}

template <class T, int N>
__device__ inline void mul(const T (&a)[N][N], const T (&b)[N][N], T (&c)[N][N]) {
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      T s = T(0);
#pragma unroll
      for (int k = 0; k < N; ++k) s += a[i][k] * b[k][j];
      c[i][j] = s;
    }
}

// Horner form of the truncated series in Matrix_Exp (ADA_Module.f90:787-804):
// s <- I + h s / (2i (2i + odd)), i = MAX_N_AMOM .. 1, starting from s = I
template <class T, int N>
__device__ inline void series(const T (&h)[N][N], T (&s)[N][N], int odd) {
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) s[i][j] = T(i == j ? 1 : 0);
  for (int it = MAX_N_AMOM; it >= 1; --it) {
    const T f = T(1) / T((2 * it) * (2 * it + (odd ? 1 : -1)));
    mul_left_inplace(h, s);
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) s[i][j] = (i == j ? T(1) : T(0)) + s[i][j] * f;
  }
}

// Matrix_Exp (ADA_Module.f90:746-853) for A + B = apb, A - B = amb: layer R
// and T of optical depth tau.  apb and amb are consumed.
template <class T, int N>
__device__ void matrix_exp(T (&apb)[N][N], T (&amb)[N][N], T tau, T (&R)[N][N], T (&Tr)[N][N]) {
  T hh[N][N], ff[N][N];
  mul(amb, apb, hh);
  mul(apb, amb, ff);
  T mh = T(0);
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) mh = fmax(mh, fabs(hh[i][j]));
  const double max_h = sqrt((double)mh) * (double)tau;
  // INT(log(max_h/0.5)/log(2)) + 1, INT() truncating toward zero, floored at 0
  // (and bounded, so a non-finite optical depth cannot run away)
  const double lv = max_h > 0.0 ? log(max_h / 0.5) / log(2.0) : -2.0;
  int nd = lv <= -1.0 ? 0 : (lv < 60.0 ? (int)trunc(lv) + 1 : 60);
  const T t = (T)((double)tau / ldexp(1.0, nd));
  const T t2 = t * t;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      hh[i][j] *= t2;
      ff[i][j] *= t2;
    }
  // S1, S3 from HH; S3 <- t (A+B) S3; keep U = S1 - S3, V = S1 + S3
  T u[N][N], v[N][N];
  series(hh, u, 0);
  series(hh, v, 1);
  mul_left_inplace(apb, v);
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      const T s1 = u[i][j], s3 = t * v[i][j];
      u[i][j] = s1 - s3;
      v[i][j] = s1 + s3;
    }
  // S2, S4 from FF; S4 <- t (A-B) S4;  S = U + S2 - S4, Sr = V - S2 - S4
  series(ff, hh, 0);          // hh now holds S2
  series(ff, apb, 1);         // apb now holds S4
  mul_left_inplace(amb, apb);
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      const T s2 = hh[i][j], s4 = t * apb[i][j];
      u[i][j] += s2 - s4;
      v[i][j] -= s2 + s4;
    }
  // trans = 2 S^-1, refl = Sr S^-1
  inv_inplace(u);
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) Tr[i][j] = T(2) * u[i][j];
  mul(v, u, R);
  // doubling: R' = R + T (I - R R)^-1 R T, T' = T (I - R R)^-1 T
  for (int d = 0; d < nd; ++d) {
    mul(R, R, u);
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) u[i][j] = (i == j ? T(1) : T(0)) - u[i][j];
    inv_inplace(u);
    mul(Tr, u, v);            // term1 = T (I - RR)^-1
    mul(v, R, u);             // term2 = term1 R
    mul(u, Tr, hh);           // term2 T
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) R[i][j] += hh[i][j];
    mul_left_inplace(v, Tr);  // T' = term1 T
  }
}
}  // namespace

}  // namespace crtm
