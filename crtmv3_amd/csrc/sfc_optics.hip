// sfc_optics.hip -- surface optics kernels of a water surface: FASTEM-X with
// the FASTEM6 azimuth model for microwave sensors and the IR sea-surface
// emissivity table for infrared sensors.  One lane per channel, one wave per
// (profile, 64-channel block); each lane writes the planes of sfc_optics.h
// for its (profile, channel).  Derivatives are carried forward-mode (Dv: a
// value and its partials in water temperature, salinity, wind speed and
// relative azimuth), so the K-matrix needs no second pass through the model.
//
//   FASTEM-X      src/SfcOptics/MW_Water/FASTEM_MWSSEM/CRTM_FastemX.f90:302-464
//   permittivity  src/SfcOptics/MW_Water/Ocean_Permittivity/Liu.f90:213-285
//   Fresnel       src/SfcOptics/MW_Water/Fresnel/Fresnel.f90
//   foam, small/large-scale and reflection corrections, slope variance, FASTEM6
//   azimuth       src/SfcOptics/MW_Water/FASTEM_MWSSEM/*.f90
//   polarisation  src/SfcOptics/CRTM_SfcOptics.f90:641-802
//   IR table      src/SfcOptics/IR_Water/IRSSEM/CRTM_IRSSEM.f90:204-334, 586-689
//
// The work is tiny next to the layer loop of the fused kernel (tens of
// flops per table point or model term per channel); FP64 throughout.
#include "crtm_device.h"
#include "sfc_optics.h"

namespace crtm {
namespace {

constexpr double PI_D = 3.141592653589793238462643383279;
constexpr double D2R_D = PI_D / 180.0;

// value and partials in (water temperature, salinity, wind speed, azimuth)
struct Dv {
  double v, d[4];
};
__device__ inline Dv cst(double x) { return Dv{x, {0.0, 0.0, 0.0, 0.0}}; }
__device__ inline Dv var(double x, int j) {
  Dv r = cst(x);
  r.d[j] = 1.0;
  return r;
}
__device__ inline Dv operator+(const Dv& a, const Dv& b) {
  return Dv{a.v + b.v, {a.d[0] + b.d[0], a.d[1] + b.d[1], a.d[2] + b.d[2], a.d[3] + b.d[3]}};
}
__device__ inline Dv operator-(const Dv& a, const Dv& b) {
  return Dv{a.v - b.v, {a.d[0] - b.d[0], a.d[1] - b.d[1], a.d[2] - b.d[2], a.d[3] - b.d[3]}};
}
__device__ inline Dv operator-(const Dv& a) { return Dv{-a.v, {-a.d[0], -a.d[1], -a.d[2], -a.d[3]}}; }
__device__ inline Dv operator*(const Dv& a, const Dv& b) {
  Dv r;
  r.v = a.v * b.v;
#pragma unroll
  for (int j = 0; j < 4; ++j) r.d[j] = a.d[j] * b.v + a.v * b.d[j];
  return r;
}
__device__ inline Dv scale(const Dv& a, double s) { return Dv{a.v * s, {a.d[0] * s, a.d[1] * s, a.d[2] * s, a.d[3] * s}}; }
__device__ inline Dv operator*(const Dv& a, double s) { return scale(a, s); }
__device__ inline Dv operator*(double s, const Dv& a) { return scale(a, s); }
__device__ inline Dv operator+(const Dv& a, double s) { Dv r = a; r.v += s; return r; }
__device__ inline Dv operator+(double s, const Dv& a) { Dv r = a; r.v += s; return r; }
__device__ inline Dv operator-(const Dv& a, double s) { Dv r = a; r.v -= s; return r; }
__device__ inline Dv operator-(double s, const Dv& a) { return (-a) + s; }
// f(a) with f'(a) = df
__device__ inline Dv chain(const Dv& a, double f, double df) { return Dv{f, {a.d[0] * df, a.d[1] * df, a.d[2] * df, a.d[3] * df}}; }
__device__ inline Dv operator/(const Dv& a, const Dv& b) {
  const double ib = 1.0 / b.v, q = a.v * ib;
  Dv r;
  r.v = q;
#pragma unroll
  for (int j = 0; j < 4; ++j) r.d[j] = (a.d[j] - q * b.d[j]) * ib;
  return r;
}
__device__ inline Dv operator/(const Dv& a, double s) { return scale(a, 1.0 / s); }
__device__ inline Dv dexp(const Dv& a) { const double e = exp(a.v); return chain(a, e, e); }
__device__ inline Dv dlog(const Dv& a) { return chain(a, log(a.v), 1.0 / a.v); }
__device__ inline Dv dsqrt(const Dv& a) { const double s = sqrt(a.v); return chain(a, s, 0.5 / s); }
__device__ inline Dv dcos(const Dv& a) { return chain(a, cos(a.v), -sin(a.v)); }
__device__ inline Dv dpow(const Dv& a, double p) {   // a^p, a >= 0
  const double f = pow(a.v, p);
  return chain(a, f, a.v > 0.0 ? p * pow(a.v, p - 1.0) : 0.0);
}
// Step (src/Utility/Hyperbolic_Step/Hyperbolic_Step.f90:78-82); its adjoint
// keeps no sensitivity beyond |x| = 70
__device__ inline Dv hstep(const Dv& x) {
  const double t = tanh(x.v);
  const double dg = fabs(x.v) < 70.0 ? 0.5 / (cosh(x.v) * cosh(x.v)) : 0.0;
  return chain(x, 0.5 * (1.0 + t), dg);
}

// complex numbers over Dv
struct Cv {
  Dv a, b;
};
__device__ inline Cv operator+(const Cv& x, const Cv& y) { return Cv{x.a + y.a, x.b + y.b}; }
__device__ inline Cv operator-(const Cv& x, const Cv& y) { return Cv{x.a - y.a, x.b - y.b}; }
__device__ inline Cv operator*(const Cv& x, const Cv& y) { return Cv{x.a * y.a - x.b * y.b, x.a * y.b + x.b * y.a}; }
__device__ inline Cv operator/(const Cv& x, const Cv& y) {
  const Dv d = y.a * y.a + y.b * y.b;
  return Cv{(x.a * y.a + x.b * y.b) / d, (x.b * y.a - x.a * y.b) / d};
}
__device__ inline Dv abs2(const Cv& z) { return z.a * z.a + z.b * z.b; }
// principal square root
__device__ inline Cv csqrt(const Cv& z) {
  const Dv m = dsqrt(z.a * z.a + z.b * z.b);
  if (z.a.v >= 0.0) {
    const Dv t = dsqrt((m + z.a) * 0.5);
    return Cv{t, z.b / (2.0 * t)};
  }
  Dv t = dsqrt((m - z.a) * 0.5);
  if (z.b.v < 0.0) t = -t;
  return Cv{z.b / (2.0 * t), t};
}

// Liu et al. (2011) double-Debye sea-water permittivity, e' - i e''
// (Liu.f90:213-285; the fit's coefficient tables, Liu.f90:60-110)
__device__ Cv ocean_permittivity(const Dv& temperature, const Dv& salinity, double f) {
  const double E0 = 1.0 / (PI_D * 4.0e-7 * 2.99792458e+08 * 2.99792458e+08);   // Fundamental_Constants.f90:87-95
  const double TAU0 = 2.0 * PI_D * E0 * 1.0e9;
  const Dv t = temperature - 273.15, s = salinity;
  const Dv einf = 3.8 + t * 2.48033e-02;
  Dv es = 87.9181727 + t * (-4.031592248e-01 + t * (9.493088010e-04 + t * -1.930858348e-06));
  Dv e1 = 5.723 + t * (2.2379e-02 + t * -7.1237e-04);
  Dv tau1 = 1.124465e-01 + t * (-3.9815727e-03 + t * (8.113381e-05 + t * -7.1824242e-07));
  Dv tau2 = 3.049979018e-03 + t * (-3.010041629e-05 + t * (4.811910733e-06 + t * -4.259775841e-08));
  Dv ie = cst(0.0);
  if (s.v > 0.0) {
    const Dv delta = 25.0 - t;
    const Dv beta = 2.033e-02 + delta * (1.266e-04 + delta * 2.464e-06) +
                    (-1.849e-05 + delta * (2.551e-07 + delta * -2.551e-08)) * s;
    const Dv alpha25 = s * (1.82521e-01 + s * (-1.46192e-03 + s * (2.09324e-05 + s * -1.28205e-07)));
    const Dv alpha = alpha25 * dexp(-(delta * beta));
    ie = -(alpha / (f * TAU0));
    es = es * (1.0 + s * (-2.697e-03 + s * -7.3e-06 + t * -8.9e-06));
    e1 = e1 * (1.0 + s * (-6.28908e-03 + s * 1.76032e-04 + t * -9.22144e-05));
    tau1 = tau1 * (1.0 + s * (-2.39357e-03 + t * (3.1353e-05 + t * -2.52477e-07)));
    tau2 = tau2 * (1.0 + s * (1.49e-01 + t * -8.8e-04 + (s * s) * -1.05e-04));
  }
  const Dv f1 = tau1 * f, f2 = tau2 * f;
  const Dv del1 = es - e1, del2 = e1 - einf;
  const Dv d1 = 1.0 + f1 * f1, d2 = 1.0 + f2 * f2;
  const Dv er = einf + del1 / d1 + del2 / d2;
  const Dv ei = (-ie) + del1 * f1 / d1 + del2 * f2 / d2;
  return Cv{er, -ei};
}

// FASTEM-X at one angle (n_Angles = 1: cos_z_mod = cos_z).  ev, eh: Stokes
// emissivities with the azimuth term; A[q][0..2]: reflection-correction
// polynomial of polarisation q (value and d/d wind speed in A[q][k].d[2]).
__device__ void fastemx(const DevMWwater& c, double f, double zenith, const Dv& tw, const Dv& sal, const Dv& wind,
                        const Dv& az, double cos_z, Dv& ev, Dv& eh, Dv A[2][3]) {
  const Cv perm = ocean_permittivity(tw, sal, f);
  // Fresnel_Reflectivity
  const Cv cz{cst(cos_z), cst(0.0)};
  const Cv z1 = csqrt(perm - Cv{cst(1.0), cst(0.0)} + Cv{cst(cos_z * cos_z), cst(0.0)});
  const Cv z2 = perm * cz;
  const Dv rh_f = abs2((cz - z1) / (cz + z1));
  const Dv rv_f = abs2((z2 - z1) / (z2 + z1));
  // Foam_Reflectivity (zenith in degrees) and Foam_Coverage
  const double* fr = c.fr;
  const double ff = fr[4] * exp(fr[5] * f);
  const double rv_foam = (1.0 - fr[0]) * ff;
  const double rh_foam = (1.0 - (1.0 + zenith * (fr[1] + zenith * (fr[2] + zenith * fr[3]))) * fr[0]) * ff;
  const Dv foam = wind.v < 0.0 ? cst(0.0) : dpow(wind, c.fc[1]) * c.fc[0];
  // Large_Scale_Correction (sec_z capped at 2)
  const double sec = fmin(1.0 / cos_z, 2.0);
  Dv rl[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    double z[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      const double* k = c.lsc + 3 * (i + 6 * q);
      z[i] = k[0] + f * (k[1] + f * k[2]);
    }
    rl[q] = wind * (z[3] + z[5] * sec) + wind * wind * z[4] + (z[0] + z[1] * sec + z[2] * sec * sec);
  }
  // Small_Scale_Correction (frequency and wind clamped; no sensitivity when clamped)
  const double fq = fmin(fmax(f, 1.4), 200.0), f2 = fq * fq;
  const Dv w = wind.v < 0.3 ? cst(0.3) : wind.v > 35.0 ? cst(35.0) : wind;
  const Dv w2 = w * w;
  const double* sc = c.ssc;
  const Dv y = w * (sc[0] * fq + sc[1] * f2 + sc[6]) + w2 * (sc[2] * fq + sc[3] * f2 + sc[4] / fq + sc[5] / f2 + sc[7]);
  const Dv fsmall = dexp(y * -(cos_z * cos_z));
  const Dv rv = rv_f * fsmall - rl[0];
  const Dv rh = rh_f * fsmall - rl[1];
  ev = 1.0 - (1.0 - foam) * rv - foam * rv_foam;
  eh = 1.0 - (1.0 - foam) * rh - foam * rh_foam;
  // Azimuth_Emissivity_F6, switched off by an azimuth beyond +-360 degrees
  if (fabs(az.v) <= 360.0) {
    const double FIT_F[6] = {6.925, 10.65, 18.7, 23.8, 36.5, 89.0};
    const double zr = zenith / 55.2;
    const double zr2 = zr * zr, zr4 = zr2 * zr2;
    const Dv w18 = wind.v > 18.0 ? cst(18.0) : wind;
    const Dv w15 = wind.v > 15.0 ? cst(15.0) : wind;
    const double lf = 1.0 - log10(30.0 / fmin(f, 37.0));
    const Dv phi = az * D2R_D;
    const Dv c1 = dcos(phi), c2 = dcos(phi * 2.0);
    // the fit frequencies bracketing f and the weight of the upper one
    int j0, j1;
    double lp;
    if (f < FIT_F[0]) { j0 = j1 = 0; lp = 0.0; }
    else if (f > FIT_F[5]) { j0 = j1 = 5; lp = 0.0; }
    else {
      j0 = 0;
      while (j0 < 4 && f >= FIT_F[j0 + 1]) ++j0;
      j1 = j0 + 1;
      lp = (f - FIT_F[j0]) / (FIT_F[j1] - FIT_F[j0]);
    }
    const Dv a2s2_0 = (w15 * w15 - w15 * w15 * w15 / 22.5) * (1.0 / 55.5556 * (2.0 / 290.0) * lf);
    const double* a = c.az;   // C(6,6,2)
    Dv comp[2][2];
#pragma unroll
    for (int jj = 0; jj < 2; ++jj) {
      const int j = jj ? j1 : j0;
      auto C = [&](int i, int q) { return a[(i - 1) + 6 * (j + 6 * q)]; };
      const Dv a1v = dexp(w18 * w18 * -C(5, 0)) - 1.0;
      const Dv A1v = a1v * (w18 * (C(2, 0) + w18 * (C(3, 0) + w18 * C(4, 0)))) * C(1, 0);
      const Dv A2v = w18 * C(6, 0);
      const Dv A1h = w18 * C(1, 1);
      const Dv a2h = dexp(w18 * w18 * -C(6, 1)) - 1.0;
      const Dv A2h = a2h * (w18 * (C(3, 1) + w18 * (C(4, 1) + w18 * C(5, 1)))) * C(2, 1);
      const Dv a1s1 = (A1v + A1h) * 0.5, a1s2 = A1v - A1h;
      const Dv a2s1 = (A2v + A2h) * 0.5, a2s2 = A2v - A2h;
      const Dv a1s1_t = a1s1 * zr2, a2s1_t = a2s1 * zr2, a1s2_t = a1s2 * zr;
      const Dv a2s2_t = a2s2_0 + (a2s2 - a2s2_0) * zr4;
      comp[jj][0] = (a1s1_t * 2.0 + a1s2_t) * 0.5 * c1 + (a2s1_t * 2.0 + a2s2_t) * 0.5 * c2;
      comp[jj][1] = (a1s1_t * 2.0 - a1s2_t) * 0.5 * c1 + (a2s1_t * 2.0 - a2s2_t) * 0.5 * c2;
    }
    ev = ev + (j0 == j1 ? comp[0][0] : comp[1][0] * lp + comp[0][0] * (1.0 - lp));
    eh = eh + (j0 == j1 ? comp[0][1] : comp[1][1] * lp + comp[0][1] * (1.0 - lp));
  }
  // Reflection_Correction polynomial: rough = 1 + sum_i zx_i (C1 + x C2 + x^2 C3)
  // with the slope variance (Compute_Slope_Variance) in zx
  const double fterm = 0.02 * f + 0.3;
  const Dv cv = (0.0030 + wind * 0.00512) * 1.0;
  const Dv g1 = hstep((cv * fterm - cv) * 1.0e6);
  const Dv var1 = cv * g1 + cv * fterm * (1.0 - g1);
  const Dv var = var1 * hstep(cv * fterm * 1.0e6);
  const double isec = 1.0 / cos_z;
  const Dv zx[7] = {cst(1.0), var, var * isec, cst(isec), var * isec * (var * isec), cst(isec * isec), var * var};
#pragma unroll
  for (int q = 0; q < 2; ++q)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      Dv s = cst(0.0);
#pragma unroll
      for (int i = 0; i < 7; ++i) s = s + zx[i] * c.rc[k + 3 * (i + 7 * q)];
      A[q][k] = s;
    }
}

// Stokes weights of the channel's polarisation (CRTM_SfcOptics.f90:641-802);
// false for a flag the reference rejects
__device__ inline bool pol_weights(int pol, double angle, double dratio, double pol_angle, double scan_rad,
                                   double& wv, double& wh) {
  switch (pol) {
    case 1: wv = 0.5; wh = 0.5; return true;                        // intensity
    case 2: wv = 0.5; wh = -0.5; return true;                       // second Stokes component
    case 3: case 4: wv = 0.0; wh = 0.0; return true;                // third, fourth (zero for water)
    case 5: case 7: case 8: case 11: case 12: wv = 1.0; wh = 0.0; return true;
    case 6: wv = 0.0; wh = 1.0; return true;
    case 9: case 10: {                                              // quasi-V / quasi-H
      const double s = dratio * sin(D2R_D * angle), s2 = s * s;
      if (pol == 9) { wv = 1.0 - s2; wh = s2; } else { wv = s2; wh = 1.0 - s2; }
      return true;
    }
    case 13: {                                                      // constant mixing angle
      const double s = dratio * sin(D2R_D * pol_angle);
      wv = s * s;
      wh = 1.0 - s * s;
      return true;
    }
    case 14: {                                                      // polarisation rotation angle
      const double tf = D2R_D * pol_angle, sp = sin(scan_rad), cp = cos(scan_rad), st = sin(tf);
      const double den = sqrt(sp * sp + st * st * (1.0 - cp * cp));
      const double ph = sp * (cp + st * (1.0 - cp)) / den;
      const double pv = -(sp * sp - st * (1.0 - cp) * cp) / den;
      const double s = sin(atan(-pv / ph));
      wv = s * s;
      wh = 1.0 - s * s;
      return true;
    }
    default: wv = 0.0; wh = 0.0; return false;
  }
}

struct ProfSfc {
  double tw, cov, sal, wind, az, zen_user, scan, cos_sensor;
};
__device__ inline ProfSfc prof_sfc(const DevBatch& b, const DevSfcIn& in, int m) {
  ProfSfc p;
  p.tw = b.sfc_t[1][m];
  p.cov = b.cov[1][m];
  p.sal = in.salinity ? in.salinity[m] : DEFAULT_SALINITY;
  p.wind = in.wind_speed ? in.wind_speed[m] : DEFAULT_WIND_SPEED;
  const double wd = in.wind_direction ? in.wind_direction[m] : DEFAULT_WIND_DIRECTION;
  p.az = wd - (in.azimuth ? in.azimuth[m] : INVALID_AZIMUTH);
  p.zen_user = b.zenith[m];
  p.scan = in.scan ? in.scan[m] : 0.0;
  // Secant_Sensor_Zenith and its cosine (CRTM_GeometryInfo.f90:104-107,
  // Common_RTSolution.f90:224)
  p.cos_sensor = 1.0 / (1.0 / cos(D2R_D * p.zen_user));
  return p;
}

// a surface other than water is outside the computed optics: flag the
// profile (its outputs become NaN and the call fails)
__device__ inline void flag_surface(const DevBatch& b, int m, int l, int* err) {
  if (l == 0 && (b.cov[0][m] > 0.0 || b.cov[2][m] > 0.0 || b.cov[3][m] > 0.0)) err[m] = 1;
}

__global__ __launch_bounds__(256) void sfc_mw_kernel(DevSensor s, DevBatch b, DevSfcIn in, DevMWwater c,
                                                     double* planes, int64_t stride, int* err) {
  const int nb = (s.L + WAVE - 1) / WAVE;
  const int64_t item = (int64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
  if (item >= (int64_t)b.M * nb) return;
  const int m = (int)(item / nb);
  const int l = (int)(item % nb) * WAVE + (threadIdx.x & (WAVE - 1));
  if (l >= s.L) return;
  if (b.skip && b.skip[m]) return;
  flag_surface(b, m, l, err);
  const ProfSfc p = prof_sfc(b, in, m);
  // SfcOptics%Angle of the sensor (Common_RTSolution.f90:290-352) and the
  // model's own cosine of it
  const double angle = acos(p.cos_sensor) / D2R_D;
  const double cos_z = cos(angle * D2R_D);
  const double dratio = fabs(p.zen_user) > 1.0 ? fabs(sin(D2R_D * p.scan) / sin(D2R_D * p.zen_user))
                                               : 6370.0 / (6370.0 + 800.0);   // CRTM_Parameters.f90:184-185
  double wv, wh;
  if (!pol_weights(s.polarization[l], angle, dratio, s.pol_angle[l], D2R_D * p.scan, wv, wh)) wv = wh = NAN;
  Dv ev, eh, A[2][3];
  fastemx(c, s.frequency[l], angle, var(p.tw, 0), var(p.sal, 1), var(p.wind, 2), var(p.az, 3), cos_z, ev, eh, A);
  double* P = planes + (int64_t)m * s.L + l;
  const int64_t S = stride;
  P[MP_EV * S] = ev.v;
  P[MP_EH * S] = eh.v;
  P[MP_WV * S] = wv * p.cov;
  P[MP_WH * S] = wh * p.cov;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    P[(MP_DEV + j) * S] = ev.d[j];
    P[(MP_DEH + j) * S] = eh.d[j];
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    P[(MP_AV + k) * S] = A[0][k].v;
    P[(MP_AH + k) * S] = A[1][k].v;
    P[(MP_DAV + k) * S] = A[0][k].d[2];
    P[(MP_DAH + k) * S] = A[1][k].d[2];
  }
  P[MP_COSZ * S] = cos_z;
}

// Four-point piecewise-quadratic Lagrange weights of one table axis
// (Find_Random_Index + LPoly, src/Interpolation/CRTM_Interpolation.f90:785-860)
// with their derivative in the abscissa
struct W4d {
  int i1;
  double w[4], dw[4];
};
__device__ inline W4d weights_d(const double* x, int n, double xi) {
  xi = fmax(fmin(xi, x[n - 1]), x[0]);
  int k = 0;
  while (k < n && !(xi <= x[k])) ++k;
  W4d r;
  r.i1 = min(max(k - 2, 0), n - 4);
  const double q0 = x[r.i1], q1 = x[r.i1 + 1], q2 = x[r.i1 + 2], q3 = x[r.i1 + 3];
  const double d0 = xi - q0, d1 = xi - q1, d2 = xi - q2, d3 = xi - q3;
  const double a0 = 1.0 / ((q0 - q1) * (q0 - q2)), a1 = 1.0 / (-(q0 - q1) * (q1 - q2)), a2 = 1.0 / ((q0 - q2) * (q1 - q2));
  const double b0 = 1.0 / ((q1 - q2) * (q1 - q3)), b1 = 1.0 / (-(q1 - q2) * (q2 - q3)), b2 = 1.0 / ((q1 - q3) * (q2 - q3));
  const double l0 = d1 * d2 * a0, l1 = d0 * d2 * a1, l2 = d0 * d1 * a2;
  const double r0 = d2 * d3 * b0, r1 = d1 * d3 * b1, r2 = d1 * d2 * b2;
  const double dl0 = (d1 + d2) * a0, dl1 = (d0 + d2) * a1, dl2 = (d0 + d1) * a2;
  const double dr0 = (d2 + d3) * b0, dr1 = (d1 + d3) * b1, dr2 = (d1 + d2) * b2;
  double wl, wr, dwl, dwr;
  if (xi < q1) { wl = 1.0; wr = 0.0; dwl = dwr = 0.0; }
  else if (xi > q2) { wl = 0.0; wr = 1.0; dwl = dwr = 0.0; }
  else { wr = d1 / (q2 - q1); wl = 1.0 - wr; dwr = 1.0 / (q2 - q1); dwl = -dwr; }
  r.w[0] = wl * l0;
  r.w[1] = wl * l1 + wr * r0;
  r.w[2] = wl * l2 + wr * r1;
  r.w[3] = wr * r2;
  r.dw[0] = dwl * l0 + wl * dl0;
  r.dw[1] = dwl * l1 + wl * dl1 + dwr * r0 + wr * dr0;
  r.dw[2] = dwl * l2 + wl * dl2 + dwr * r1 + wr * dr1;
  r.dw[3] = dwr * r2 + wr * dr2;
  return r;
}

__global__ __launch_bounds__(256) void sfc_ir_kernel(DevSensor s, DevBatch b, DevSfcIn in, DevIRwater c,
                                                     double* planes, int64_t stride, int* err) {
  const int nb = (s.L + WAVE - 1) / WAVE;
  const int64_t item = (int64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
  if (item >= (int64_t)b.M * nb) return;
  const int m = (int)(item / nb);
  const int l = (int)(item % nb) * WAVE + (threadIdx.x & (WAVE - 1));
  if (l >= s.L) return;
  if (b.skip && b.skip[m]) return;
  flag_surface(b, m, l, err);
  const ProfSfc p = prof_sfc(b, in, m);
  const double angle = fabs(acos(p.cos_sensor) / D2R_D);
  // a wind speed (or, with a temperature axis, a water temperature) outside
  // the table leaves the reference's adjoint with no sensitivity at all
  // (CRTM_IRSSEM.f90:640-641)
  const bool v_out = p.wind < c.wind[0] || p.wind > c.wind[c.nv - 1];
  const bool t_out = c.has_t && (p.tw < c.temp[0] || p.tw > c.temp[c.nt - 1]);
  const W4d wa = weights_d(c.angle, c.na, angle);
  const W4d wf = weights_d(c.frequency, c.nf, s.wavenumber[l]);
  const W4d wu = weights_d(c.wind, c.nv, p.wind);
  W4d wt{};
  if (c.has_t) wt = weights_d(c.temp, c.nt, p.tw);
  const int64_t s2 = c.na, s3 = s2 * c.nf, s4 = s3 * c.nv;
  double e = 0.0, du = 0.0, dt = 0.0;
  const int nd = c.has_t ? 4 : 1;
  for (int d = 0; d < nd; ++d) {
    const double wtd = c.has_t ? wt.w[d] : 1.0, dwtd = c.has_t ? wt.dw[d] : 0.0;
    const double* zt = c.emis + (c.has_t ? (wt.i1 + d) * s4 : 0);
    for (int cc = 0; cc < 4; ++cc) {
      const double* zv = zt + (wu.i1 + cc) * s3;
      double ev = 0.0;
      for (int bb = 0; bb < 4; ++bb) {
        const double* zf = zv + (wf.i1 + bb) * s2 + wa.i1;
        const double ea = wa.w[0] * zf[0] + wa.w[1] * zf[1] + wa.w[2] * zf[2] + wa.w[3] * zf[3];
        ev += wf.w[bb] * ea;
      }
      e += wtd * wu.w[cc] * ev;
      du += wtd * wu.dw[cc] * ev;
      dt += dwtd * wu.w[cc] * ev;
    }
  }
  if (v_out || t_out) du = dt = 0.0;
  double* P = planes + (int64_t)m * s.L + l;
  P[IP_E * stride] = p.cov * e;
  P[IP_R * stride] = p.cov * (1.0 - e);
  P[IP_DT * stride] = p.cov * dt;
  P[IP_DU * stride] = p.cov * du;
}

}  // namespace

void launch_sfc_optics(const DevSensor& s, const DevBatch& b, const DevSfcIn& in, int mode, const DevMWwater& mw,
                       const DevIRwater& ir, double* planes, int64_t stride, int* err, hipStream_t st) {
  const int nb = (s.L + WAVE - 1) / WAVE;
  const int64_t items = (int64_t)b.M * nb;
  if (items == 0) return;
  const int per_block = 4;
  const unsigned grid = (unsigned)((items + per_block - 1) / per_block);
  if (mode == SFC_MW)
    hipLaunchKernelGGL(sfc_mw_kernel, dim3(grid), dim3(per_block * WAVE), 0, st, s, b, in, mw, planes, stride, err);
  else
    hipLaunchKernelGGL(sfc_ir_kernel, dim3(grid), dim3(per_block * WAVE), 0, st, s, b, in, ir, planes, stride, err);
}

}  // namespace crtm
