// odps_prep.hip -- per-profile preparation on the GPU: geometry, Map_Input
// (LayerAvg weights, absorber clamps, height-dependent secants), the ODPS
// integrated ratios and predictors X, the OPTRAN water-vapour-line
// predictors, the ODPS-level -> user-level interpolation brackets, and the
// channel-independent partial-derivative table the adjoint kernel reads as
// wave-uniform scalars.
//
// One workgroup per profile.  Sequential recurrences (LayerAvg's running
// search start, the integrated-ratio prefix sums, the OPTRAN column
// integrals and the level bracket merge) run in single lanes; everything
// per-layer runs across the workgroup.  Arithmetic is FP64 throughout; the
// results are stored in the precision RT of the fused kernel that reads them.
//
// Reference: src/AtmAbsorption/ODPS/ODPS_CoordinateMapping.f90:151-319,
// :607-634, :963-1012, :1108-1320; src/AtmAbsorption/ODPS/ODPS_Predictor.f90:
// 290-357, 645-1095, 2716-2816; src/GeometryInfo/CRTM_GeometryInfo.f90:100-143.
#include <cfloat>
#include <climits>

#include "crtm_device.h"
#include "pred_adjoint.h"

namespace crtm {

namespace {
constexpr double PI = 3.141592653589793238462643383279;
constexpr double D2R = PI / 180.0;
constexpr double EARTH_RADIUS = 6370.0;
constexpr double MAX_TRANS_ZENITH = 63.6122;
constexpr double MW_H2O = 18.01528, MW_DRYAIR = 28.9648;
constexpr double R_DRYAIR = 8.314472 / (MW_DRYAIR * 1.0e-3);
constexpr double CCH = 0.001 * R_DRYAIR / 9.80665;
constexpr double SMALLDIFF = 1.0e-20;
constexpr double RECIP_GRAV = 1.0 / 980.665;
constexpr double MIN_ABSORBER = 1.0e-76;
constexpr double TOLERANCE = 2.220446049250313e-16;
constexpr int PREP_THREADS = 128;
}  // namespace

// Status codes written per profile
enum { ST_OK = 0, ST_LAYERAVG = 1, ST_NZ = 2, ST_LAYERS = 3, ST_BRACKET = 4, ST_FOV = 9 };

// ---------------------------------------------------------------------------
// CRTM_Atmosphere_AddLayers (src/Atmosphere/CRTM_Atmosphere.f90:248-400):
// a profile whose top level is below TOA_PRESSURE gets the model
// climatology's layers above it (Extra_Layers :1043-1050, Interp_LPoly /
// Interp_Linear :1070-1084, Layer_P / Layer_X :1054-1066), shifted so the
// lowest added layer continues the profile's top two layers (Shift_Profile
// :1088-1098), and negative absorber amounts raised to
// MINIMUM_ABSORBER_AMOUNT.  One 64-thread workgroup per profile: lane 0
// builds the added layers, all lanes copy the extended profile.
// ---------------------------------------------------------------------------
constexpr double TOA_PRESSURE = 0.005;   // src/CRTM_Parameters.f90:157

__global__ void __launch_bounds__(64) add_layers_kernel(DevBatch in, ExtBatch e, const double* clim,
                                                        int ids0, int ids1, int ids2, int ids3, int ids4, int ids5) {
  const int m = blockIdx.x, tid = threadIdx.x;
  const int ids[NA_MAX] = {ids0, ids1, ids2, ids3, ids4, ids5};
  const int K = (in.skip && in.skip[m]) ? 0 : in.n_layers[m];   // a skipped profile is not extended
  const int64_t lo = in.loff[m];
  const double* lp = in.level_p + lo + m;
  const int na = in.na;
  __shared__ double xlp[N_EXTRA_MAX + 1], xp[N_EXTRA_MAX], xt[N_EXTRA_MAX], xa[NA_MAX][N_EXTRA_MAX];
  __shared__ int sn;
  if (tid == 0) {
    int n = 0;
    const double ptop = lp[0];
    if (ptop > TOA_PRESSURE) {
      n = -1;                                   // no model level below the top: reference FAILURE
      for (int q = 1; q < N_MODEL_LEVELS; ++q)
        if (clim[q] >= ptop) {
          n = q;
          break;
        }
      for (int j = 0; j < na; ++j)
        if (ids[j] < 1 || ids[j] > 6) n = -1;   // no climatology for this absorber
      if (K < 2) n = -1;                        // Shift_Profile needs two layers
    }
    if (n > 0) {
      const int c = in.climatology ? in.climatology[m] : 6;
      const int model = (c >= 1 && c <= 6) ? c - 1 : 5;
      const double* P = clim;
      const double* T = clim + CLIM_T0 + model * N_MODEL_LEVELS;
      const double lpoly = (log(ptop) - log(P[n - 1])) / (log(P[n]) - log(P[n - 1]));
      for (int q = 0; q < n; ++q) xlp[q] = P[q];
      xlp[n] = ptop;
      auto lev = [&](const double* X, int q) { return q < n ? X[q] : (X[n] - X[n - 1]) * lpoly + X[n - 1]; };
      for (int k = 0; k < n; ++k) {
        xp[k] = (xlp[k + 1] - xlp[k]) / log(xlp[k + 1] / xlp[k]);
        xt[k] = 0.5 * (lev(T, k) + lev(T, k + 1));
      }
      const double elpoly = (log(xp[n - 1]) - log(in.p[lo])) / (log(in.p[lo + 1]) - log(in.p[lo]));
      auto shift = [&](double* x, const double* xu) {
        const double dx = (xu[1] - xu[0]) * elpoly + xu[0] - x[n - 1];
        for (int k = 0; k < n; ++k) x[k] += dx;
      };
      shift(xt, in.t + lo);
      for (int j = 0; j < na; ++j) {
        const double* A = clim + CLIM_A0 + (model * 6 + ids[j] - 1) * N_MODEL_LEVELS;
        for (int k = 0; k < n; ++k) xa[j][k] = 0.5 * (lev(A, k) + lev(A, k + 1));
        shift(xa[j], in.a + (size_t)j * in.NT + lo);
        for (int k = 0; k < n; ++k)
          if (xa[j][k] < 0.0) xa[j][k] = MIN_ABSORBER;
      }
    }
    sn = n;
  }
  __syncthreads();
  const int n = sn < 0 ? 0 : sn;
  const int64_t eo = lo + (int64_t)m * N_EXTRA_MAX;
  if (tid == 0) {
    e.n_layers[m] = sn < 0 ? 0 : K + n;         // 0 layers: the preparation kernel fails the profile
    e.n_extra[m] = n;
    e.loff[m] = eo;
    if (m == in.M - 1) e.loff[in.M] = in.loff[in.M] + (int64_t)in.M * N_EXTRA_MAX;
  }
  const int Kt = K + n;
  for (int q = tid; q <= Kt; q += 64) e.level_p[eo + m + q] = q <= n ? (q < n ? xlp[q] : lp[0]) : lp[q - n];
  for (int q = tid; q < Kt; q += 64) {
    const bool add = q < n;
    e.p[eo + q] = add ? xp[q] : in.p[lo + q - n];
    e.t[eo + q] = add ? xt[q] : in.t[lo + q - n];
    for (int j = 0; j < na; ++j) e.a[(size_t)j * e.NT + eo + q] = add ? xa[j][q] : in.a[(size_t)j * in.NT + lo + q - n];
  }
}

void launch_add_layers(const DevBatch& b, const ExtBatch& e, const double* clim, const int* ids, hipStream_t stream) {
  hipLaunchKernelGGL(add_layers_kernel, dim3(b.M), dim3(64), 0, stream, b, e, clim, ids[0], ids[1], ids[2], ids[3],
                     ids[4], ids[5]);
}

// ---------------------------------------------------------------------------
// Layer mapping, one lane per profile: LayerAvg (ODPS_CoordinateMapping.f90:
// 1108-1320), Compute_Interp_Index (:963-1012) and the Map_Input_AD window
// tables of the K-matrix kernel.  These are sequential recurrences (LayerAvg's
// running search start, the bracket merge), so each profile's chain runs in
// its own lane and the batch's chains run side by side; per-profile columns
// live in HBM laid out [row][profile] so lanes stepping through the same rows
// coalesce.
// ---------------------------------------------------------------------------
template <class RT>
__global__ void __launch_bounds__(64) layer_map_kernel(DevSensor s, DevBatch bt, ProfWork<RT> w) {
  __shared__ double rlnp[KO_MAX], rlnp_lev[KO_MAX + 1];
  const int Ko = s.Ko;
  for (int k = threadIdx.x; k < Ko && k < KO_MAX; k += 64) rlnp[k] = log(s.ref_p[k]);
  for (int k = threadIdx.x; k <= Ko && k <= KO_MAX; k += 64) rlnp_lev[k] = log(s.ref_lp[k]);
  __syncthreads();
  const int m = blockIdx.x * 64 + threadIdx.x;
  if (m >= bt.M) return;
  if (bt.skip && bt.skip[m]) {   // Options(m)%Skip_Profile: no kernel computes the profile
    w.status[m] = ST_SKIP;
    return;
  }
  const int K = bt.n_layers[m];
  if (K < 1 || K > K_MAX || K > bt.max_k || Ko > KO_MAX) {
    w.status[m] = ST_LAYERS;
    return;
  }
  // the antenna correction indexes its coefficients by Geometry%iFOV
  if (bt.ac_on && s.ac_nfov > 0 && bt.ifov && (bt.ifov[m] < 0 || bt.ifov[m] > s.ac_nfov)) {
    w.status[m] = ST_FOV;
    return;
  }
  const int64_t M = bt.M, loff = bt.loff[m];
  const double* lp_u = bt.level_p + loff + m;   // [K+1]
  const double* p_u = bt.p + loff;
  double* ul = w.lm + m;                                   // user layer lnP, row k
  double* ulv = w.lm + (int64_t)bt.max_k * M + m;          // user level lnP, row k (0..K)
  double* pz = w.lm + (int64_t)(2 * bt.max_k + 1) * M + m; // LayerAvg accumulation column
  for (int k = 0; k < K; ++k) {
    ul[k * M] = log(p_u[k]);
    pz[k * M] = 0.0;
  }
  for (int k = 0; k <= K; ++k) ulv[k * M] = (k == 0 && lp_u[0] <= 0.0) ? log(lp_u[1] / 2.0) : log(lp_u[k]);
  int* acc = w.acc_i + (size_t)m * Ko * 3;
  double* aw64 = w.accw64 + (size_t)m * NZ_MAX;
  RT* aw = w.acc_w + (size_t)m * NZ_MAX;

  // ---- LayerAvg: weights of user layers in each ODPS layer ----
  const int kn1 = Ko, kn2 = K;
  auto P1 = [&](int i) { return rlnp[i - 1]; };
  auto P2 = [&](int i) { return ul[(int64_t)(i - 1) * M]; };
  auto PZ = [&](int i) -> double& { return pz[(int64_t)(i - 1) * M]; };
  int istart = 1;
  int st = ST_OK;
  int nz = 0;
  const double p2n = P2(kn2);
  for (int ki = 1; ki <= kn1 && st == ST_OK; ++ki) {
    double z2 = P1(ki);
    double z1 = (ki == 1) ? 2.0 * z2 - P1(ki + 1) : P1(ki - 1);
    double z3 = (ki == kn1) ? 2.0 * z2 - z1 : P1(ki + 1);
    if (z3 > p2n) z3 = p2n;
    int iskip = 0;
    if (z2 >= p2n) { z3 = p2n; z2 = p2n; iskip = 1; }
    int ic = 0, j, lo = kn2 + 1, hi = 0;
    bool jumped = false;
    double pj1 = P2(istart);
    for (j = istart; j <= kn2 - 1; ++j) {
      const double pj = pj1;   // P2(j)
      if (pj > z3) { jumped = true; break; }
      pj1 = P2(j + 1);
      for (int side = 0; side < 2; ++side) {
        // side 0: the upper half-layer (toward z1); side 1: the lower half (toward z3)
        double ze = side == 0 ? z1 : z3;
        bool take = side == 0 ? (pj <= z2 && pj1 > z1) : (pj < z3 && pj1 >= z2 && iskip == 0);
        if (!take) continue;
        double zo = side == 0 ? z3 : z1;
        int itop = 0, ibot = 0;
        double y1, y2;
        if (ze < zo) {
          y1 = ze; if (pj > ze) { y1 = pj; itop = 1; }
          y2 = z2; if (pj1 < z2) { y2 = pj1; ibot = 1; }
        } else {
          y1 = z2; if (pj > z2) { y1 = pj; itop = 1; }
          y2 = ze; if (pj1 < ze) { y2 = pj1; ibot = 1; }
        }
        double dy = y2 - y1, dz = ze - z2;
        if (fabs(dz) < SMALLDIFF) { st = ST_LAYERAVG; break; }
        double dzd = 1.0 / dz;
        double zw1 = (ze - y1) * dzd * dy;
        double zw2 = (ze - y2) * dzd * dy;
        double dx = pj1 - pj;
        if (fabs(dx) < SMALLDIFF) { st = ST_LAYERAVG; break; }
        double d = (pj1 - z2) * (1.0 / dx);
        if (ze < zo && ibot == 0) {
          zw1 = zw1 + zw2 * d;
          zw2 = zw2 * (1.0 - d);
        } else if (ze > zo && itop == 0) {
          zw2 = zw2 + zw1 * (1.0 - d);
          zw1 = zw1 * d;
        }
        PZ(j) += zw1;
        PZ(j + 1) += zw2;
        lo = min(lo, j);
        hi = max(hi, j + 1);
        ic = 1;
      }
      if (st != ST_OK) break;
    }
    if (st != ST_OK) break;
    if (!jumped) j = kn2;
    if (ic == 0) { PZ(j) = 1.0; lo = min(lo, j); hi = max(hi, j); }
    int i1 = istart;
    for (int ii = istart; ii <= hi; ++ii)
      if (PZ(ii) != 0.0) { i1 = ii; break; }
    istart = i1;
    int i2 = kn2;
    for (int ii = i1 + 1; ii <= kn2; ++ii)
      if (ii > hi || PZ(ii) == 0.0) { i2 = ii - 1; break; }
    int n = i2 - i1 + 1;
    if (nz + n > NZ_MAX) { st = ST_NZ; break; }
    double zsum = 0.0;
    for (int ii = i1; ii <= i2; ++ii) zsum += PZ(ii);
    for (int ii = i1; ii <= i2; ++ii) {
      const double v = PZ(ii) / zsum;
      aw64[nz + ii - i1] = v;
      aw[nz + ii - i1] = (RT)v;
    }
    acc[3 * (ki - 1)] = i1 - 1;
    acc[3 * (ki - 1) + 1] = n;
    acc[3 * (ki - 1) + 2] = nz;
    nz += n;
    for (int ii = lo; ii <= hi; ++ii) PZ(ii) = 0.0;
  }
  if (st != ST_OK) {
    w.status[m] = st;
    return;
  }

  // ---- ODPS level brackets of the user levels (Compute_Interp_Index) ----
  {
    const int nx = Ko + 1, nu = K + 1;
    int* ui = w.uidx + (size_t)m * (bt.max_k + 1);
    RT* uf = w.ufrac + (size_t)m * (bt.max_k + 1);
    auto UL = [&](int i) { return ulv[(int64_t)i * M]; };
    int k1 = nu + 1;
    for (int iu = 1; iu <= nu; ++iu) {
      if (UL(iu - 1) < rlnp_lev[0]) { ui[iu - 1] = 0; uf[iu - 1] = RT(-1); }
      else { k1 = iu; break; }
    }
    int k2 = 0;
    for (int iu = nu; iu >= k1; --iu) {
      if (UL(iu - 1) > rlnp_lev[nx - 1]) { ui[iu - 1] = nx - 1; uf[iu - 1] = RT(-1); }
      else { k2 = iu; break; }
    }
    int j = 1;
    for (int iu = k1; iu <= k2; ++iu) {
      bool found = false;
      const double u = UL(iu - 1);
      for (int ix = j; ix <= nx - 1; ++ix) {
        if (u >= rlnp_lev[ix - 1] && u <= rlnp_lev[ix]) {
          ui[iu - 1] = ix - 1;
          uf[iu - 1] = (RT)((u - rlnp_lev[ix - 1]) / (rlnp_lev[ix] - rlnp_lev[ix - 1]));
          j = ix;
          found = true;
          break;
        }
      }
      if (!found) st = ST_BRACKET;
    }
  }
  // widest run of ODPS layers whose averaging supports share one user
  // layer: the K-matrix kernel folds Map_Input_AD into its sweep with a
  // register window when this fits the window
  int span = 0;
  {
    // user layer iu is complete after ODPS layer kmin(iu), the first whose
    // support reaches it; its weights from layers kmin(iu)+c, c < MAP_WIN
    RT* uw = w.uwin + (size_t)m * bt.max_k * MAP_WIN;
    int kmin = 0, kmax = -1;
    for (int iu = 0; iu < K; ++iu) {
      while (kmax + 1 < Ko && acc[3 * (kmax + 1)] <= iu) ++kmax;
      while (kmin < Ko && acc[3 * kmin] + acc[3 * kmin + 1] - 1 < iu) ++kmin;
      span = max(span, kmax - kmin + 1);
      for (int c = 0; c < MAP_WIN; ++c) {
        const int kk = kmin + c;
        double wv = 0.0;
        if (kk < Ko && kk <= kmax && iu >= acc[3 * kk] && iu < acc[3 * kk] + acc[3 * kk + 1])
          wv = aw64[acc[3 * kk + 2] + iu - acc[3 * kk]];
        uw[(size_t)iu * MAP_WIN + c] = (RT)wv;
      }
    }
    // rows completed at ODPS layer k: (i2(k-1), i2(k)], i2 = last user layer of the support
    int* fr = w.frange + (size_t)m * Ko * 2;
    for (int k = 0; k < Ko; ++k) {
      const int i2 = acc[3 * k] + acc[3 * k + 1] - 1;
      const int lo = k > 0 ? acc[3 * (k - 1)] + acc[3 * (k - 1) + 1] : 0;
      if (i2 < lo - 1) span = 1 << 20;   // supports not ordered: the kernel takes the general path
      fr[2 * k] = lo;
      fr[2 * k + 1] = i2;
    }
  }
  w.geo[(size_t)m * 4 + 3] = (double)span;
  w.status[m] = st;
}

// Component of predictor slot s of group G.
template <int G> constexpr int comp_of(int s) {
  int j = 0;
  while (j + 1 < GT<G>::NCOMP && GT<G>::XOFF[j + 1] <= s) ++j;
  return j;
}

// Sensitivity-table entries of slot S onwards: the predictor adjoint chain
// (comp_adjoint, finish_layer, final_quantities) run on the unit predictor
// adjoint of each slot, which the compiler folds to the slot's own terms.
template <int G, int S>
__device__ inline void sens_slot_row(const double* U, double* R) {
  if constexpr (S < GT<G>::NX) {
    constexpr unsigned msk = SensPattern<G>::tab.mask[S];
    if constexpr (msk != 0) {
      constexpr int J = comp_of<G>(S), I = S - GT<G>::XOFF[J];
      double xa[18] = {};
      xa[I] = 1.0;
      LayerAdj<double> a{};
      comp_adjoint<G, J, double>(xa, U, a);
      finish_layer<G, double>(U, a);
      double q[QIdx<G>::NQ];
      final_quantities<G, double>(a, U, q);
      int e = SensPattern<G>::tab.off[S];
#pragma unroll
      for (int c = 0; c < QIdx<G>::NQ; ++c)
        if ((msk >> c) & 1) R[e++] = q[c];
    }
    sens_slot_row<G, S + 1>(U, R);
  }
}

// ---------------------------------------------------------------------------
// Per-layer preparation, one workgroup per profile (after layer_map_kernel):
// geometry, mapped T / absorbers with the Min/Max clamps, secants, integrated
// ratios, predictors, the adjoint table and the OPTRAN table.
// ---------------------------------------------------------------------------
// Compute_NLTE_Predictor (src/NLTE/NLTE_Correction/CRTM_NLTECorrection.f90:95-260)
// for one profile: mean temperatures of the two NLTE layers (the profile's
// layers between each pair of pressure levels, clamped to the coefficient
// file's range), the bilinear weights in the sensor- and solar-zenith secants
// (clamped to the tables).  The predictor stays inactive without NLTE data,
// without the correction, or with the source below the horizon (secant < 0).
// compute_tm is one flag for both layers, left by the last layer, as in the
// reference.
__device__ void nlte_predictor(const DevSensor& s, const DevBatch& bt, int m, const double* lp, const double* t,
                               int K, double sec_sensor, double* rec) {
  for (int i = 0; i < NLTE_REC; ++i) rec[i] = 0.0;
  if (!bt.nlte_on || s.nl_nsen == 0) return;
  const double szen = bt.source_zenith ? bt.source_zenith[m] : 100.0;
  const double cosv = cos(D2R * szen);
  const double sec_src = cosv != 0.0 ? 1.0 / cosv : (cosv < 0.0 ? -DBL_MAX : DBL_MAX);
  if (sec_src < 0.0) return;
  bool ctm = true;
  int k1 = INT_MAX, k2 = INT_MAX;
  for (int i = 0; i < 2; ++i) {
    double tm;
    if (s.nl_up[i] < lp[0] || s.nl_lo[i] > lp[K]) {
      tm = s.nl_mean[i];
      ctm = false;
    } else {
      for (int k = 1; k <= K; ++k)
        if (lp[k] > s.nl_up[i]) {
          k1 = k - 1;
          break;
        }
      for (int k = k1; k <= K; ++k)
        if (lp[k] > s.nl_lo[i]) {
          k2 = k;
          break;
        }
      tm = 0.0;
      for (int k = k1 + 1; k <= k2; ++k) tm += t[k - 1];
      tm /= (double)(k2 - k1);
      rec[NL_K1 + i] = k1;
      rec[NL_K2 + i] = k2;
      ctm = true;
      if (tm < s.nl_min[i]) {
        tm = s.nl_min[i];
        ctm = false;
      } else if (tm > s.nl_max[i]) {
        tm = s.nl_max[i];
        ctm = false;
      }
    }
    rec[NL_PRED + 1 + i] = tm;
  }
  rec[NL_PRED] = 1.0;
  const int ns = s.nl_nsen, no = s.nl_nsol;
  const double x = fmin(fmax(sec_sensor, s.nl_ssen[0]), s.nl_ssen[ns - 1]);
  const double y = fmin(fmax(sec_src, s.nl_ssol[0]), s.nl_ssol[no - 1]);
  int isen = 0, isol = 0;
  for (int i = 0; i < ns - 1; ++i)
    if (x >= s.nl_ssen[i] && x <= s.nl_ssen[i + 1]) {
      isen = i;
      break;
    }
  for (int i = 0; i < no - 1; ++i)
    if (y >= s.nl_ssol[i] && y <= s.nl_ssol[i + 1]) {
      isol = i;
      break;
    }
  const double x1 = s.nl_ssen[isen], x2 = s.nl_ssen[isen + 1], y1 = s.nl_ssol[isol], y2 = s.nl_ssol[isol + 1];
  const double xx1 = (x - x2) / (x1 - x2), xx2 = (x - x1) / (x2 - x1);
  const double yy1 = (y - y2) / (y1 - y2), yy2 = (y - y1) / (y2 - y1);
  rec[NL_WGT + 0] = xx1 * yy1;
  rec[NL_WGT + 1] = xx1 * yy2;
  rec[NL_WGT + 2] = xx2 * yy1;
  rec[NL_WGT + 3] = xx2 * yy2;
  rec[NL_ISEN] = isen;
  rec[NL_ISOL] = isol;
  rec[NL_CTM] = ctm ? 1.0 : 0.0;
  rec[NL_ACTIVE] = 1.0;
}

template <int G, class RT>
__global__ void __launch_bounds__(PREP_THREADS) odps_prep_kernel(DevSensor s, DevBatch bt, ProfWork<RT> w) {
  using T = GT<G>;
  constexpr int NA = T::NABS;
  constexpr int OPT = G == 3 ? 1 : KO_MAX;   // OPTRAN tables only for the infrared groups
  const int m = blockIdx.x;
  const int tid = threadIdx.x;
  const int Ko = s.Ko;
  const int K = bt.n_layers[m];
  const int64_t loff = bt.loff[m];
  const double* lp_u = bt.level_p + loff + m;   // [K+1]
  const double* t_u = bt.t + loff;

  __shared__ double rlnp_lev[KO_MAX + 1];
  __shared__ int acc_i1[KO_MAX], acc_n[KO_MAX], acc_off[KO_MAX];
  __shared__ double t_o[KO_MAX];
  __shared__ double a_o[NA][KO_MAX];
  __shared__ double sec_o[KO_MAX];
  __shared__ double tz[KO_MAX], tzp[KO_MAX], gazp[NA][KO_MAX], gatzp[NA][KO_MAX];
  __shared__ double iv_s[OPT], ave_s[OPT], st_s[OPT], sp_s[OPT], xlt_s[OPT], xlp_s[OPT];
  __shared__ double geo[4];

  // status of the layer mapping (uniform over the workgroup)
  if (w.status[m] != ST_OK) return;
  const double* aw64 = w.accw64 + (size_t)m * NZ_MAX;
  for (int k = tid; k < Ko; k += PREP_THREADS) {
    const int* a = w.acc_i + ((size_t)m * Ko + k) * 3;
    acc_i1[k] = a[0];
    acc_n[k] = a[1];
    acc_off[k] = a[2];
  }
  for (int k = tid; k <= Ko; k += PREP_THREADS) rlnp_lev[k] = log(s.ref_lp[k]);
  if (tid == 0) {
    // geometry (CRTM_GeometryInfo_Compute) and the surface temperature
    double zen = bt.zenith[m];
    double zen_rad = D2R * zen;
    double sec_sensor = 1.0 / cos(zen_rad);
    double trans_rad = zen > MAX_TRANS_ZENITH ? D2R * MAX_TRANS_ZENITH : zen_rad;
    double sec_trans = zen > MAX_TRANS_ZENITH ? 1.0 / cos(trans_rad) : sec_sensor;
    geo[0] = sec_trans;
    geo[1] = 1.0 / sec_sensor;
    double ts = 0.0;
    for (int c = 0; c < 4; ++c) ts += bt.cov[c][m] * bt.sfc_t[c][m];
    geo[2] = ts;
    geo[3] = trans_rad;
    double* g = w.geo + (size_t)m * 4;
    g[0] = geo[0];
    g[1] = geo[1];
    g[2] = geo[2];
    if (w.nlte) nlte_predictor(s, bt, m, lp_u, t_u, K, sec_sensor, w.nlte + (size_t)m * NLTE_REC);
  }
  __syncthreads();

  // ---- mapped temperature / absorbers on ODPS layers, secants ----
  const double alt = bt.altitude ? bt.altitude[m] : 0.0;
  double z_off = 0.0;
  {
    // surface bracket (Map_Input, ODPS_CoordinateMapping.f90:228-243)
    const double psfc = lp_u[K];
    int sfc_idx = Ko;
    double frac = 0.0;
    if (s.ref_lp[Ko] > psfc) {
      for (int k = Ko; k >= 0; --k)
        if (s.ref_lp[k] < psfc) {
          sfc_idx = k + 1;
          frac = (rlnp_lev[sfc_idx] - log(lp_u[K])) / (rlnp_lev[sfc_idx] - rlnp_lev[k]);
          break;
        }
    }
    if (s.ref_lp[Ko] >= psfc)
      z_off = -(s.z_ref[sfc_idx] + frac * (s.z_ref[sfc_idx - 1] - s.z_ref[sfc_idx])) + alt;
    else
      z_off = CCH * s.ref_t[Ko - 1] * log(1013.0 / s.ref_lp[Ko]);
  }
  const double sv = (EARTH_RADIUS + alt) * sin(geo[3]);
  for (int k = tid; k < Ko; k += PREP_THREADS) {
    const int i1 = acc_i1[k], n = acc_n[k];
    double sum = 0.0;
    for (int q = 0; q < n; ++q) sum += aw64[acc_off[k] + q] * t_u[i1 + q];
    t_o[k] = sum;
    for (int j = 0; j < T::NABS; ++j) {
      const int ju = bt.jmap[j];
      double v;
      if (ju >= 0) {
        const double* au = bt.a + (size_t)ju * bt.NT + loff;
        double sm = 0.0;
        for (int q = 0; q < n; ++q) sm += aw64[acc_off[k] + q] * au[i1 + q];
        v = sm;
        if (v <= TOLERANCE) v = TOLERANCE;
        const double mn = s.min_a[(size_t)j * Ko + k], mx = s.max_a[(size_t)j * Ko + k];
        if (v < mn) v = mn;
        if (v > mx) v = mx;
      } else {
        v = s.ref_a[(size_t)j * Ko + k];
      }
      a_o[j][k] = v;
    }
    double sa = sv / (EARTH_RADIUS + s.z_ref[k + 1] + z_off);
    sec_o[k] = 1.0 / sqrt(1.0 - sa * sa);
  }
  __syncthreads();

  // ---- integrated ratios: one prefix sum per lane ----
  {
    const int nser = 2 + 2 * T::NABS;
    if (tid < nser) {
      double ref = 0.0, sum = 0.0;
      for (int k = 0; k < Ko; ++k) {
        const double pdp = s.pdp[k];
        if (tid == 0) {
          ref += s.ref_t[k]; sum += t_o[k]; tz[k] = sum / ref;
        } else if (tid == 1) {
          ref += pdp * s.ref_t[k]; sum += pdp * t_o[k]; tzp[k] = sum / ref;
        } else if (tid < 2 + T::NABS) {
          const int j = tid - 2;
          ref += pdp * s.ref_a[(size_t)j * Ko + k]; sum += pdp * a_o[j][k]; gazp[j][k] = sum / ref;
        } else {
          const int j = tid - 2 - T::NABS;
          ref += pdp * s.ref_a[(size_t)j * Ko + k] * s.ref_t[k];
          sum += pdp * a_o[j][k] * t_o[k];
          gatzp[j][k] = sum / ref;
        }
      }
    } else if (tid == nser && s.optran) {
      // OPTRAN column integrals (ODPS_Predictor.f90:2760-2800)
      const double* vap = a_o[s.h2o_idx - 1];
      double ivp = 0.0, stt = 0.0, spp = 0.0;
      for (int k = 0; k < Ko; ++k) {
        const double dab = RECIP_GRAV * (s.ref_lp[k + 1] - s.ref_lp[k]) * vap[k] * sec_o[k];
        const double iv = ivp + dab;
        stt += t_o[k] * dab;
        spp += s.ref_p[k] * dab;
        const double inv = iv > MIN_ABSORBER ? 1.0 / iv : 0.0;
        iv_s[k] = iv;
        ave_s[k] = 0.5 * (ivp + iv);
        st_s[k] = stt;
        sp_s[k] = spp;
        xlt_s[k] = 0.5 * stt * inv;
        xlp_s[k] = 0.5 * spp * inv;
        ivp = iv;
      }
    }
  }
  __syncthreads();

  // ---- predictors, adjoint table, OPTRAN table (one layer per lane) ----
  for (int k = tid; k < Ko; k += PREP_THREADS) {
    double X[T::NX];
    double U[NU];
    for (int i = 0; i < T::NX; ++i) X[i] = 0.0;
    for (int i = 0; i < NU; ++i) U[i] = 0.0;
    const double rt = s.ref_t[k];
    auto A = [&](int j) { return a_o[j][k] / s.ref_a[(size_t)j * Ko + k]; };
    const double sec = sec_o[k];
    const double dt = t_o[k] - rt, Tr = t_o[k] / rt, T2 = Tr * Tr;
    const double adt = fabs(dt), dt2 = dt * adt;
    const double h2o = A(0), h2o_a = sec * h2o, h2o_r = sqrt(h2o_a), h2o_s = h2o_a * h2o_a;
    const double h2o_r4 = sqrt(h2o_r), gt0 = gatzp[0][k], h2o_dtzp = h2o / gt0;
    U[U_SEC] = sec; U[U_T] = Tr; U[U_IT] = 1.0 / Tr; U[U_IT2] = 1.0 / T2; U[U_IT4] = 1.0 / (T2 * T2);
    U[U_DT] = dt; U[U_ADT] = adt; U[U_DT2] = dt2;
    U[U_H2O] = h2o; U[U_H2OA] = h2o_a; U[U_H2OR] = h2o_r; U[U_H2OS] = h2o_s; U[U_H2OR4] = h2o_r4;
    U[U_GT0] = gt0; U[U_H2ODTZP] = h2o_dtzp; U[U_H2OP] = gazp[0][k];
    auto XS = [&](int j, int i, double v) { X[T::XOFF[j - 1] + i - 1] = v; };
    XS(1, 1, sec); XS(1, 2, sec * Tr); XS(1, 3, sec * T2); XS(1, 4, Tr); XS(1, 5, sec * sec);
    XS(1, 6, T2); XS(1, 7, tz[k]);
    if constexpr (G == 1 || G == 2) {
      const double o3 = A(1), co2 = A(2), o3_a = sec * o3, o3_r = sqrt(o3_a), o3p = gazp[1][k];
      const double tzpk = tzp[k], co2p = gazp[2][k];
      U[U_O3] = o3; U[U_O3A] = o3_a; U[U_O3R] = o3_r; U[U_O3P] = o3p; U[U_SP] = sqrt(sec * o3p);
      U[U_SQSO3P] = sqrt(sec / o3p); U[U_TZP] = tzpk; U[U_ST] = sqrt(Tr); U[U_CO2P] = co2p;
      XS(3, 1, h2o_a / Tr); XS(3, 2, h2o_a / Tr * h2o); XS(3, 3, h2o_a / T2 * h2o / T2);
      XS(3, 4, h2o_a / T2); XS(3, 5, h2o_a / T2 * h2o); XS(3, 6, h2o_a / (T2 * T2)); XS(3, 7, h2o_a);
      XS(4, 1, o3_a); XS(4, 2, o3_a * dt); XS(4, 3, o3_a * o3 * o3p); XS(4, 4, o3_a * o3_a);
      XS(4, 5, o3_a * o3p); XS(4, 6, o3_a * sqrt(sec * o3p)); XS(4, 7, o3_r * dt); XS(4, 8, o3_r);
      XS(4, 9, o3_r * o3 / o3p); XS(4, 10, sec * o3p); XS(4, 11, (sec * o3p) * (sec * o3p));
      XS(5, 1, sec * Tr); XS(5, 2, sec * T2); XS(5, 3, Tr); XS(5, 4, T2); XS(5, 5, sec);
      XS(5, 6, sec * co2); XS(5, 7, sec * tzpk); XS(5, 8, (sec * co2p) * (sec * co2p));
      XS(5, 9, tzpk * tzpk * tzpk); XS(5, 10, sec * tzpk * sqrt(Tr));
      const double h2op = gazp[0][k];
      XS(2, 1, h2o_a); XS(2, 2, h2o_a * dt); XS(2, 3, h2o_s); XS(2, 4, h2o_a * dt2); XS(2, 5, h2o_r4);
      XS(2, 6, h2o_s * h2o_a); XS(2, 7, h2o_r); XS(2, 8, h2o_r * dt); XS(2, 9, h2o_s * h2o_s);
      XS(2, 10, h2o_dtzp); XS(2, 11, h2o_r * h2o_dtzp); XS(2, 12, (sec * h2op) * (sec * h2op));
      XS(2, 13, sec * h2op); XS(2, 14, sec); XS(2, 15, sec * co2);
      if constexpr (G == 1) {
        const double co = A(4), n2o = A(3), ch4 = A(5);
        const double n2o_a = sec * n2o, n2o_r = sqrt(n2o_a), n2o_s = n2o_a * n2o_a;
        const double co_a = sec * co, co_r = sqrt(co_a), co_s = co_a * co_a, cozp = gazp[4][k];
        const double co_q = co_a * co / cozp;
        const double ch4_a = sec * ch4, ch4_r = sqrt(ch4_a), ch4zp = gazp[5][k], ch4_q = sec * ch4zp;
        const double n2ozp = gazp[3][k];
        U[U_CO] = co; U[U_N2O] = n2o; U[U_CH4] = ch4; U[U_N2OA] = n2o_a; U[U_N2OR] = n2o_r;
        U[U_COA] = co_a; U[U_COR] = co_r; U[U_COZP] = cozp; U[U_COQ] = co_q; U[U_CH4A] = ch4_a;
        U[U_CH4R] = ch4_r; U[U_CH4ZP] = ch4zp; U[U_CH4Q] = ch4_q; U[U_N2OZP] = n2ozp;
        U[U_SQCOR] = sqrt(co_r); U[U_SQCH4R] = sqrt(ch4_r); U[U_N2OPOW] = pow(n2o_a, -0.75);
        U[U_SQCOZP] = sqrt(cozp);
        XS(5, 11, co_a);
        XS(2, 16, ch4_a); XS(2, 17, ch4_a * ch4_a * dt); XS(2, 18, co_a);
        XS(7, 1, co_a); XS(7, 2, co_a * dt); XS(7, 3, sqrt(co_r)); XS(7, 4, co_r * dt); XS(7, 5, co_s);
        XS(7, 6, co_r); XS(7, 7, co_a * dt2); XS(7, 8, co_q); XS(7, 9, co_q / co_r);
        XS(7, 10, co_q * sqrt(cozp));
        XS(8, 1, ch4_a * dt); XS(8, 2, ch4_r); XS(8, 3, ch4_a * ch4_a); XS(8, 4, ch4_a); XS(8, 5, ch4 * dt);
        XS(8, 6, ch4_q); XS(8, 7, ch4_q * ch4_q); XS(8, 8, sqrt(ch4_r)); XS(8, 9, gatzp[5][k]);
        XS(8, 10, sec * gatzp[5][k]); XS(8, 11, ch4_r * ch4 / ch4zp);
        XS(6, 1, n2o_a * dt); XS(6, 2, n2o_r); XS(6, 3, n2o * dt); XS(6, 4, pow(n2o_a, 0.25));
        XS(6, 5, n2o_a); XS(6, 6, sec * n2ozp); XS(6, 7, sec * gatzp[3][k]); XS(6, 8, n2o_s);
        XS(6, 9, gatzp[3][k]); XS(6, 10, n2o_r * n2o / n2ozp); XS(6, 11, ch4_a); XS(6, 12, ch4_a * ch4zp);
        XS(6, 13, co_a); XS(6, 14, co_a * sec * cozp);
      }
    } else {
      const double h2op = gazp[0][k];
      XS(2, 1, h2o_a / Tr); XS(2, 2, h2o_a / Tr * h2o); XS(2, 3, h2o_a / T2 * h2o / T2);
      XS(2, 4, h2o_a / T2); XS(2, 5, h2o_a / T2 * h2o); XS(2, 6, h2o_a / (T2 * T2)); XS(2, 7, h2o_a);
      XS(2, 8, h2o_a * dt); XS(2, 9, (sec * h2op) * (sec * h2op)); XS(2, 10, sec * h2op);
      XS(2, 11, sec); XS(2, 12, h2o_s * h2o_a); XS(2, 13, h2o_s * h2o_s); XS(2, 14, h2o_dtzp);
    }
    auto inv = [](double v) { return v != 0.0 ? 1.0 / v : 0.0; };
    U[U_IO3P] = inv(U[U_O3P]);
    U[U_IN2OZP] = inv(U[U_N2OZP]);
    U[U_ICOR] = inv(U[U_COR]);
    U[U_ISQCOR] = inv(U[U_SQCOR]);
    U[U_ISQCOZP] = inv(U[U_SQCOZP]);
    U[U_ISQCH4R] = inv(U[U_SQCH4R]);
    U[U_ICH4ZP] = inv(U[U_CH4ZP]);
    U[U_IST] = inv(U[U_ST]);
    U[U_ICH4R] = inv(U[U_CH4R]);
    U[U_ICOZP] = inv(U[U_COZP]);
    U[U_IN2OR] = inv(U[U_N2OR]);
    U[U_IO3R] = inv(U[U_O3R]);
    U[U_IGT0] = inv(gt0);
    U[U_IH2OR4] = inv(h2o_r4);
    U[U_IH2OR] = inv(h2o_r);
    U[U_IRT] = 1.0 / rt;
    U[U_PDP] = s.pdp[k];
    U[U_ITZREF] = 1.0 / s.tz_ref[k];
    U[U_ITZPREF] = 1.0 / s.tzp_ref[k];
    U[U_TODPS] = t_o[k];
    for (int j = 0; j < T::NABS; ++j) {
      U[U_IRA0 + j] = 1.0 / s.ref_a[(size_t)j * Ko + k];
      U[U_IGAZP0 + j] = 1.0 / s.gazp_ref[(size_t)j * Ko + k];
      U[U_IGATZP0 + j] = 1.0 / s.gatzp_ref[(size_t)j * Ko + k];
      U[U_AODPS0 + j] = a_o[j][k];
      // Map_Input_AD: adjoint passes unless unmapped or the final value <= EPSILON
      U[U_MASK0 + j] = (bt.jmap[j] >= 0 && a_o[j][k] > TOLERANCE) ? 1.0 : 0.0;
    }
    RT* xo = w.x + ((size_t)m * Ko + k) * T::NX;
    for (int i = 0; i < T::NX; ++i) xo[i] = (RT)X[i];
    // adjoint row: sensitivities of the final quantities to each predictor,
    // then the cumulative-chain factors and the Map_Input_AD masks
    using RR = RRow<G>;
    double R[RR::NR];
    sens_slot_row<G, 0>(U, R);
    R[RR::PDP] = U[U_PDP];
    R[RR::TODPS] = U[U_TODPS];
    for (int j = 0; j < T::NABS; ++j) {
      R[RR::AODPS0 + j] = U[U_AODPS0 + j];
      R[RR::MASK0 + j] = U[U_MASK0 + j];
    }
    RT* uo = w.u + ((size_t)m * Ko + k) * RR::NR;
    for (int i = 0; i < RR::NR; ++i) uo[i] = (RT)R[i];
    if (s.optran) {
      double O[NO];
      const double tk = t_o[k], t2 = tk * tk, pk = s.ref_p[k], p2 = pk * pk;
      const double vap = a_o[s.h2o_idx - 1][k];
      O[O_X0 + 0] = 1.0;
      O[O_X0 + 1] = tk; O[O_X0 + 2] = pk; O[O_X0 + 3] = t2; O[O_X0 + 4] = p2; O[O_X0 + 5] = tk * pk;
      O[O_X0 + 6] = t2 * pk; O[O_X0 + 7] = tk * p2; O[O_X0 + 8] = t2 * p2; O[O_X0 + 9] = pow(pk, 0.25);
      O[O_X0 + 10] = vap; O[O_X0 + 11] = vap / t2;
      O[O_X0 + 12] = xlt_s[k] + (k > 0 ? xlt_s[k - 1] : 0.0);
      O[O_X0 + 13] = xlp_s[k] + (k > 0 ? xlp_s[k - 1] : 0.0);
      O[O_X0 + 14] = sec;
      const double ap1 = log((ave_s[k] - s.alpha_c2) / s.alpha_c1) / s.alpha;
      O[O_AP0] = 1.0;
      for (int i = 1; i <= MAX_OPTRAN_ORDER; ++i) O[O_AP0 + i] = O[O_AP0 + i - 1] * ap1;
      for (int p = 0; p < OPZ; ++p) O[O_DX_DT0 + p] = O[O_DX_DV0 + p] = 0.0;
      O[O_DX_DT0 + 1] = 1.0;
      O[O_DX_DT0 + 3] = 2.0 * tk;
      O[O_DX_DT0 + 5] = pk;
      O[O_DX_DT0 + 6] = 2.0 * tk * pk;
      O[O_DX_DT0 + 7] = p2;
      O[O_DX_DT0 + 8] = 2.0 * tk * p2;
      O[O_DX_DT0 + 11] = -2.0 * vap / (t2 * tk);
      O[O_DX_DV0 + 10] = 1.0;
      O[O_DX_DV0 + 11] = 1.0 / t2;
      const double dpong = RECIP_GRAV * (s.ref_lp[k + 1] - s.ref_lp[k]);
      O[O_DA] = dpong * vap * sec;
      O[O_IDEN] = 1.0 / (s.alpha * (ave_s[k] - s.alpha_c2));
      O[O_INV] = iv_s[k] > MIN_ABSORBER ? 1.0 / iv_s[k] : 0.0;
      O[O_ST] = st_s[k];
      O[O_SP] = sp_s[k];
      O[O_IIV2] = iv_s[k] > MIN_ABSORBER ? 1.0 / (iv_s[k] * iv_s[k]) : 0.0;
      O[O_DPSEC] = dpong * sec;
      O[O_PR] = pk;
      O[O_T] = tk;
      RT* oo = w.o + ((size_t)m * Ko + k) * NO;
      for (int i = 0; i < NO; ++i) oo[i] = (RT)O[i];
    }
  }
}

template <int G, class RT>
void launch_prep(const DevSensor& s, const DevBatch& b, const ProfWork<RT>& w, hipStream_t stream) {
  hipLaunchKernelGGL((layer_map_kernel<RT>), dim3((b.M + 63) / 64), dim3(64), 0, stream, s, b, w);
  hipLaunchKernelGGL((odps_prep_kernel<G, RT>), dim3(b.M), dim3(PREP_THREADS), 0, stream, s, b, w);
}

template void launch_prep<1, float>(const DevSensor&, const DevBatch&, const ProfWork<float>&, hipStream_t);
template void launch_prep<2, float>(const DevSensor&, const DevBatch&, const ProfWork<float>&, hipStream_t);
template void launch_prep<3, float>(const DevSensor&, const DevBatch&, const ProfWork<float>&, hipStream_t);
template void launch_prep<1, double>(const DevSensor&, const DevBatch&, const ProfWork<double>&, hipStream_t);
template void launch_prep<2, double>(const DevSensor&, const DevBatch&, const ProfWork<double>&, hipStream_t);
template void launch_prep<3, double>(const DevSensor&, const DevBatch&, const ProfWork<double>&, hipStream_t);

}  // namespace crtm
