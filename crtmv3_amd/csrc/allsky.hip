// allsky.hip -- all-sky (cloudy) forward path.
//
//   cloud_prep_kernel    one thread per profile: clouds removed where the
//                        layer cloud fraction is below 1e-12
//                        (src/CRTM_Forward_Module.f90:368-378), cloud arrays
//                        moved to the AddLayers layout, coverage type
//                        (src/Atmosphere/CRTM_Atmosphere.f90:158-203), total
//                        cloud cover and water-content scaling
//                        (src/Atmosphere/Cloud/CRTM_CloudCover_Define.f90:447-642),
//                        list of cloudy layers.
//   cloud_layer_kernel   one thread per (cloudy layer, channel): cloud optics
//                        from the CloudCoeff tables (src/AtmScatter/CRTM_CloudScatter.f90:
//                        221-399, 927-999, 1269-1416), AtmOptics combination
//                        with the gas OD (src/AtmOptics/CRTM_AtmOptics.f90:369-475),
//                        phase matrix and Normalize_Phase
//                        (src/RTSolution/Common_RTSolution.f90:1857-1920, 2885-2979),
//                        layer reflection / transmission by Matrix_Exp and
//                        doubling and the thermal emissivity
//                        (src/RTSolution/ADA/ADA_Module.f90:552-680, 746-853).
//   allsky_sweep_kernel  one thread per (profile, channel) of cloudy profiles:
//                        ADA adding from the surface up (ADA_Module.f90:97-258)
//                        or Emission when no layer scatters, fractional-cover
//                        combination (src/CRTM_Forward_Module.f90:1086-1120), BT.
//
// The layer kernel is the arithmetic: ~60 nz^3 FMAs per (layer, channel) in
// registers, templated on the angle count nz so every matrix index is a
// compile-time constant.  Matrices are row-major [i][j] per lane.
#include "crtm_device.h"

namespace crtm {

namespace {
constexpr double WATER_CONTENT_THRESHOLD = 1.0e-6;  // src/CRTM_Parameters.f90:256
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
// w >= 1) into od, bs and pl[1..nl].
__device__ void add_cloud(const DevCloudCoeff& cc, bool mw, double freq, int type, double reff, double temp,
                          double wc, int nl, int loff, double& od, double& bs_sum, double* pl) {
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

// =========================================================================
// cloud preparation: one thread per profile
// =========================================================================
__global__ void __launch_bounds__(64) cloud_prep_kernel(DevBatch be, DevClouds cl, CloudWork cw) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= be.M) return;
  const int K = be.n_layers[m];
  const int nx = be.n_extra[m];
  const int64_t le = be.loff[m], lu = be.out_loff[m];
  const int64_t NTe = be.NT, NTu = be.out_NT;
  const int nc = cl.nc;
  auto cf_at = [&](int kk) -> double {   // layer fraction after the clean-up, extended layout
    const int ku = kk - nx;
    if (ku < 0 || !cl.cf) return 0.0;
    const double f = cl.cf[lu + ku];
    return f < MIN_COVERAGE_THRESHOLD ? 0.0 : f;
  };
  double max_reff = 0.0;
  bool fractional = false;
  for (int n = 0; n < nc; ++n) {
    const int ty = cl.type[(int64_t)m * nc + n];
    const bool typed = ty >= WATER_CLOUD && ty <= HAIL_CLOUD;   // other values: empty slot
    for (int kk = 0; kk < K; ++kk) {
      const int ku = kk - nx;
      double w = 0.0, r = 0.0;
      if (typed && ku >= 0 && cl.cf && cl.cf[lu + ku] >= MIN_COVERAGE_THRESHOLD) {
        w = cl.wc[n * NTu + lu + ku];
        r = cl.reff[n * NTu + lu + ku];
      }
      cw.wc[n * NTe + le + kk] = w;
      cw.reff[n * NTe + le + kk] = r;
      max_reff = fmax(max_reff, r);
      if (w > WATER_CONTENT_THRESHOLD && cf_at(kk) > MIN_COVERAGE_FLAG) fractional = true;
    }
  }
  double tcc = 0.0;
  if (fractional) {
    switch (cl.overlap) {
      case 1: {   // maximum
        double cov = cf_at(0);
        for (int kk = 1; kk < K; ++kk) cov = fmax(cov, cf_at(kk));
        tcc = cov;
        break;
      }
      case 2: {   // random
        double prod = 1.0;
        for (int kk = 0; kk < K; ++kk) {
          const double f = cf_at(kk);
          if (f > MIN_COVERAGE_THRESHOLD) prod *= 1.0 - f;
        }
        tcc = 1.0 - prod;
        break;
      }
      case 3: {   // maximum-random
        double prod = 1.0 - cf_at(0);
        for (int kk = 1; kk < K; ++kk) {
          const double fp = cf_at(kk - 1), f = cf_at(kk);
          const double maxcov = 1.0 - fmax(fp, f);
          if (prod == 0.0 && 1.0 - fp == 0.0) prod = maxcov;
          else prod = prod * (maxcov / (1.0 - fp));
        }
        tcc = 1.0 - prod;
        break;
      }
      case 5:     // overcast
        tcc = 1.0;
        break;
      default: {  // average (the default overlap)
        double ws = 0.0, cws = 0.0, cc = 0.0;
        for (int kk = 0; kk < K; ++kk) {
          double lwc = 0.0;
          for (int n = 0; n < nc; ++n) {
            const double w = cw.wc[n * NTe + le + kk];
            if (w > WATER_CONTENT_THRESHOLD) lwc += w;
          }
          ws += lwc;
          cws += cf_at(kk) * lwc;
          if (ws > 0.0) cc = cws / ws;
        }
        tcc = cc;
      }
    }
    if (tcc > MIN_COVERAGE_THRESHOLD)
      for (int n = 0; n < nc; ++n)
        for (int kk = 0; kk < K; ++kk) cw.wc[n * NTe + le + kk] /= tcc;
  }
  int count = 0;
  for (int kk = 0; kk < K; ++kk) {
    bool cloudy = false;
    for (int n = 0; n < nc; ++n) cloudy |= cw.wc[n * NTe + le + kk] > WATER_CONTENT_THRESHOLD;
    count += cloudy;
  }
  const int base = count ? atomicAdd(cw.total, count) : 0;
  int j = 0;
  for (int kk = 0; kk < K; ++kk) {
    bool cloudy = false;
    for (int n = 0; n < nc; ++n) cloudy |= cw.wc[n * NTe + le + kk] > WATER_CONTENT_THRESHOLD;
    int slot = -1;
    if (cloudy && base + j < cw.cap) {
      slot = base + j;
      cw.cl_desc[slot] = make_int2(m, kk);
    }
    j += cloudy;
    cw.cl_slot[le + kk] = slot;
  }
  cw.cl_count[m] = count;
  cw.tcc[m] = fractional ? tcc : 0.0;
  cw.frac[m] = fractional;
  cw.max_reff[m] = max_reff;
}

// =========================================================================
// layer optics: one thread per (cloudy layer, channel)
// =========================================================================
template <class RT>
__global__ void __launch_bounds__(256) cloud_optics_kernel(DevSensor s, DevBatch be, DevClouds cl, CloudWork cw,
                                                            DevCloudCoeff cc, LayerScratch<RT> sc, const RT* tau_g,
                                                            int n_items) {
  const bool mw = s.sensor_type == 1;
  for (int item = blockIdx.x * blockDim.x + threadIdx.x; item < n_items; item += gridDim.x * blockDim.x) {
    const int c = item / s.L, l = item - c * s.L;
    const int2 d = cw.cl_desc[c];
    const int m = d.x, k = d.y;
    if (cw.status[m] != 0) continue;
    const int nl = legendre_terms(cl.n_streams, cw.max_reff[m], s.wavenumber[l]);
    if (nl > NL_AS - 1) continue;   // stream counts above the built tables are rejected by the host
    const int64_t ik = be.loff[m] + k;
    // CRTM_Compute_CloudScatter: accumulate over the clouds of the layer
    double od = (double)tau_g[ik * s.L + l], ssa = 0.0, pl[NL_AS];
#pragma unroll
    for (int q = 0; q < NL_AS; ++q) pl[q] = 0.0;
    const int loff = legendre_offset(nl);
    const double freq = mw ? s.frequency[l] : s.wavenumber[l];
#pragma unroll 1
    for (int n = 0; n < cl.nc; ++n) {
      const double wc = cw.wc[n * be.NT + ik];
      if (wc > WATER_CONTENT_THRESHOLD)
        add_cloud(cc, mw, freq, cl.type[(int64_t)m * cl.nc + n], cw.reff[n * be.NT + ik], be.t[ik], wc, nl, loff,
                  od, ssa, pl);
    }
    // CRTM_AtmOptics_Combine: P_l /= ssa, P_0 = 1/2, delta truncation f = P_nl
    if (ssa > BS_THRESHOLD) {
      const double w = ssa / od;
      const double iss = 1.0 / ssa;
#pragma unroll
      for (int q = 1; q < NL_AS; ++q) pl[q] *= iss;
      pl[0] = 0.5;
      double f = 0.0;
#pragma unroll
      for (int q = 1; q < NL_AS; ++q)
        if (q == nl) f = pl[q];
      od = (1.0 - f * w) * od;
      ssa = (1.0 - f) * w / (1.0 - f * w);
    }
    RT* out = sc.v + item;
    const int64_t st = sc.n_items;
    out[st * sc.oTau()] = (RT)od;
    out[st * sc.oSsa()] = (RT)ssa;
#pragma unroll
    for (int q = 0; q < NL_AS - 1; ++q) out[st * (sc.oP() + q)] = (RT)(q < nl ? pl[q] : 0.0);
  }
}

// =========================================================================
// layer reflection / transmission: one thread per (cloudy layer, channel)
// of angle count NZ
// =========================================================================
template <class RT, int NZ>
__global__ void __launch_bounds__(256) cloud_rt_kernel(DevSensor s, CloudWork cw, DevClouds cl, DevCloudCoeff cc,
                                                        LayerScratch<RT> sc, const double* geo, int n_items) {
  for (int item = blockIdx.x * blockDim.x + threadIdx.x; item < n_items; item += gridDim.x * blockDim.x) {
    const int c = item / s.L, l = item - c * s.L;
    const int m = cw.cl_desc[c].x;
    if (cw.status[m] != 0) continue;
    const int nl = legendre_terms(cl.n_streams, cw.max_reff[m], s.wavenumber[l]);
    if (nl > NL_AS - 1) continue;
    const double cosu = geo[(int64_t)m * 4 + 1];
    const AngleSet as = angle_set(cc.gauss, (nl + 2) / 2, cosu);
    if (as.nz != NZ) continue;
    RT* out = sc.v + item;
    const int64_t st = sc.n_items;
    const RT om = out[st * sc.oSsa()];
    if (!((double)om > BS_THRESHOLD)) continue;
    const RT tau = out[st * sc.oTau()];
    RT plr[NL_AS - 1];
#pragma unroll
    for (int q = 0; q < NL_AS - 1; ++q) plr[q] = out[st * (sc.oP() + q)];

    // ---- phase matrix (n_Stokes = 1) and Normalize_Phase ----
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
    RT fs[NZ][NZ], fd[NZ][NZ];   // ff + bb and ff - bb, upper triangle first
#pragma unroll
    for (int i = 0; i < NZ; ++i)
#pragma unroll
      for (int j = i; j < NZ; ++j) {
        RT e = RT(0), o = RT(0);
#pragma unroll
        for (int q = 0; q < NL_AS - 1; ++q) {
          if (q >= nl) break;
          const RT t = plr[q] * leg[i][q] * leg[j][q];
          if (q & 1) o += t; else e += t;
        }
        RT ff = e + o, bb = e - o;
        if (ff < RT(0)) ff = RT(PHASE_THRESHOLD);
        if (bb < RT(0)) bb = RT(PHASE_THRESHOLD);
        fs[i][j] = ff + bb;
        fd[i][j] = ff - bb;
      }
    RT sum_prev = RT(0);
#pragma unroll
    for (int i = 0; i < NZ; ++i) {
      if (i < ns) {
        RT nf = RT(0);
#pragma unroll
        for (int j = i; j < NZ; ++j) nf += fs[i][j] * wt[j];
        const RT scale = (RT(1) - sum_prev) / nf;
#pragma unroll
        for (int j = i; j < NZ; ++j) {
          fs[i][j] *= scale;
          fd[i][j] *= scale;
        }
        RT si = RT(0);
        if (i < NZ - 1) {
#pragma unroll
          for (int j = 0; j <= i; ++j) si += fs[j][i + 1] * wt[j];
        }
        sum_prev = si;
      }
    }
    if (ns < NZ) {
      const RT inf = RT(1) / sum_prev;
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

    // ---- CRTM_AMOM_layer ----
    RT R[NZ][NZ], Tm[NZ][NZ];
    RT e[NZ];
    if ((double)tau < DELTA_OPTICAL_DEPTH) {   // single-scattering initialisation
      const RT sv = tau * om;
#pragma unroll
      for (int i = 0; i < NZ; ++i) {
        const RT cst = sv / mu[i];
        RT thermal = RT(0);
#pragma unroll
        for (int j = 0; j < NZ; ++j) {
          R[i][j] = cst * RT(0.5) * (fs[i][j] - fd[i][j]) * wt[j];
          Tm[i][j] = cst * RT(0.5) * (fs[i][j] + fd[i][j]) * wt[j] + (i == j ? RT(1) - tau / mu[i] : RT(0));
          thermal += R[i][j] + Tm[i][j];
        }
        e[i] = RT(1) - thermal;
      }
    } else {
      const RT sv = om < RT(MAX_ALBEDO) ? om : RT(MAX_ALBEDO);
#pragma unroll
      for (int i = 0; i < NZ; ++i) {
        const RT cst = sv / mu[i], dg = RT(1) / mu[i];
#pragma unroll
        for (int j = 0; j < NZ; ++j) {
          fs[i][j] = cst * fs[i][j] * wt[j] - (i == j ? dg : RT(0));   // A + B
          fd[i][j] = cst * fd[i][j] * wt[j] - (i == j ? dg : RT(0));   // A - B
        }
      }
      matrix_exp(fs, fd, tau, R, Tm);
#pragma unroll
      for (int i = 0; i < NZ; ++i) {
        RT thermal = RT(0);
#pragma unroll
        for (int j = 0; j < NZ; ++j)
          if (j < ns) thermal += Tm[i][j] + R[i][j];
        if (i == NZ - 1 && NZ == ns + 1) thermal += Tm[NZ - 1][NZ - 1];
        e[i] = RT(1) - thermal;
      }
    }
#pragma unroll
    for (int i = 0; i < NZ; ++i) {
#pragma unroll
      for (int j = 0; j < NZ; ++j) {
        out[st * (sc.oR() + i * sc.nzc + j)] = R[i][j];
        out[st * (sc.oT() + i * sc.nzc + j)] = Tm[i][j];
      }
      out[st * (sc.oE() + i)] = e[i];
    }
  }
}

// =========================================================================
// adding sweep: one thread per (profile, channel) of the cloudy profiles
// =========================================================================
template <class RT, int NZ>
__global__ void __launch_bounds__(256) allsky_sweep_kernel(DevSensor s, DevBatch be, DevClouds cl, CloudWork cw,
                                                            DevCloudCoeff cc, LayerScratch<RT> sc, const RT* tau_g,
                                                            const double* geo, DevOut o, AllskyOut ao) {
  const int64_t n = (int64_t)be.M * s.L;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const int m = (int)(t / s.L), l = (int)(t - (int64_t)m * s.L);
    if (cw.cl_count[m] == 0 || cw.status[m] != 0) continue;
    const double cosu = geo[(int64_t)m * 4 + 1];
    const int nl = legendre_terms(cl.n_streams, cw.max_reff[m], s.wavenumber[l]);
    if (nl > NL_AS - 1) continue;
    const AngleSet as = angle_set(cc.gauss, (nl + 2) / 2, cosu);
    if (as.nz != NZ) continue;
    const int K = be.n_layers[m];
    const int64_t le = be.loff[m];
    const int lg = s.chan_offset + l;
    const int64_t ol = (int64_t)m * be.Ltot + lg;
    const int64_t st = sc.n_items;
    // Scattering_RT: some layer of the column scatters
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
    double r_cloudy;
    double up_rad = 0.0, down_rad = 0.0, sfc_b = 0.0;
    if (!scat) {
      // CRTM_Emission with the cloud absorption in the layer ODs
      const RT iu = (RT)(1.0 / cosu);
      RT rd = cosmic, uacc = RT(0), tc = RT(1);
      for (int k = 0; k < K; ++k) {
        RT te;
        const RT B = planck(c1, c2, bc1, bc2, be.t[le + k], te);
        const RT tr = rexp<RT>(-layer_tau(k) * iu);
        rd = rd * tr + B * (RT(1) - tr);
        uacc += B * (RT(1) - tr) * tc;
        tc *= tr;
      }
      r_cloudy = (double)(uacc + tc * (eps * Bs + (RT(1) - eps) * rd));
      up_rad = (double)uacc;
      down_rad = (double)rd;
      sfc_b = (double)Bs;
    } else {
      const int ns = as.ns;
      const double* gmu = cc.gauss + ns * 20;
      RT imu[NZ];   // Gauss nodes, then the sensor angle unless it coincides with a node
#pragma unroll
      for (int i = 0; i < NZ; ++i) imu[i] = (RT)(1.0 / (i < ns ? gmu[i] : cosu));
      RT rup[NZ][NZ], iup[NZ];
#pragma unroll
      for (int i = 0; i < NZ; ++i) {
#pragma unroll
        for (int j = 0; j < NZ; ++j) rup[i][j] = i == j ? RT(1) - eps : RT(0);
        iup[i] = eps * Bs;
      }
      for (int k = K - 1; k >= 0; --k) {
        RT te;
        const RT B = planck(c1, c2, bc1, bc2, be.t[le + k], te);
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
          // Gamma = (I - R_up R_layer)^-1 and T Gamma in FP64
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
          mul_left_inplace(rup, tl);          // tl <- R_up T_layer
          mul_left_inplace(gt, tl);           // tl <- T Gamma R_up T
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
      // cosmic background reflected at the top; radiance at the sensor angle
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
  }
}

// r_clear / tb_clear / cover of the profiles without cloudy layers
__global__ void allsky_clear_copy_kernel(DevSensor s, int M, int Ltot, const int* cl_count, DevOut o, AllskyOut ao) {
  const int64_t n = (int64_t)M * s.L;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const int m = (int)(t / s.L), l = (int)(t - (int64_t)m * s.L);
    if (cl_count && cl_count[m] > 0) continue;
    const int64_t ol = (int64_t)m * Ltot + s.chan_offset + l;
    if (ao.r_clear) ao.r_clear[ol] = o.radiance[ol];
    if (ao.tb_clear) ao.tb_clear[ol] = o.bt[ol];
    if (ao.tcc) ao.tcc[ol] = 0.0;
  }
}

// ------------------------------------------------------------------ launchers
void launch_cloud_prep(const DevBatch& be, const DevClouds& cl, const CloudWork& cw, hipStream_t st) {
  hipLaunchKernelGGL(cloud_prep_kernel, dim3((be.M + 63) / 64), dim3(64), 0, st, be, cl, cw);
}

static int persistent_blocks(int64_t n) {
  const int64_t want = (n + 255) / 256;
  return (int)(want < 8192 ? (want > 0 ? want : 1) : 8192);
}

template <class RT>
void launch_allsky(const DevSensor& s, const DevBatch& be, const DevClouds& cl, const CloudWork& cw,
                   const DevCloudCoeff& cc, const LayerScratch<RT>& sc, const RT* tau_g, const double* geo,
                   int n_items, const int* nz_list, int n_nz, const DevOut& o, const AllskyOut& ao,
                   hipStream_t st) {
  const int lb = persistent_blocks(n_items);
  const int sb = persistent_blocks((int64_t)be.M * s.L);
  if (n_items > 0)
    hipLaunchKernelGGL((cloud_optics_kernel<RT>), dim3(lb), dim3(256), 0, st, s, be, cl, cw, cc, sc, tau_g,
                       n_items);
  for (int q = 0; q < n_nz; ++q) {
    switch (nz_list[q]) {
#define CRTM_NZ_CASE(N)                                                                                     \
  case N:                                                                                                   \
    if (n_items > 0)                                                                                        \
      hipLaunchKernelGGL((cloud_rt_kernel<RT, N>), dim3(lb), dim3(256), 0, st, s, cw, cl, cc, sc, geo,     \
                         n_items);                                                                          \
    hipLaunchKernelGGL((allsky_sweep_kernel<RT, N>), dim3(sb), dim3(256), 0, st, s, be, cl, cw, cc, sc,    \
                       tau_g, geo, o, ao);                                                                  \
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
  if (ao.r_clear || ao.tb_clear || ao.tcc)
    hipLaunchKernelGGL(allsky_clear_copy_kernel, dim3(sb), dim3(256), 0, st, s, be.M, be.Ltot, cw.cl_count, o, ao);
}

void launch_clear_copy(const DevSensor& s, int M, int Ltot, const DevOut& o, const AllskyOut& ao, hipStream_t st) {
  if (!(ao.r_clear || ao.tb_clear || ao.tcc)) return;
  hipLaunchKernelGGL(allsky_clear_copy_kernel, dim3(persistent_blocks((int64_t)M * s.L)), dim3(256), 0, st, s, M,
                     Ltot, (const int*)nullptr, o, ao);
}

template void launch_allsky<float>(const DevSensor&, const DevBatch&, const DevClouds&, const CloudWork&,
                                   const DevCloudCoeff&, const LayerScratch<float>&, const float*, const double*,
                                   int, const int*, int, const DevOut&, const AllskyOut&, hipStream_t);
template void launch_allsky<double>(const DevSensor&, const DevBatch&, const DevClouds&, const CloudWork&,
                                    const DevCloudCoeff&, const LayerScratch<double>&, const double*,
                                    const double*, int, const int*, int, const DevOut&, const AllskyOut&,
                                    hipStream_t);

}  // namespace crtm
