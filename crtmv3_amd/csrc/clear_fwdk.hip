// clear_fwdk.hip -- fused clear-sky forward model and K-matrix adjoint.
//
// Mapping: one lane per channel (64-channel blocks), one wave per profile.
// Coefficients of a block are laid out [layer][slot][lane] so every load is
// a coalesced 256-byte line; predictors and the adjoint partial-derivative
// table of the wave's profile are wave-uniform and read as scalars.
//
// Per (profile, channel) the kernel runs:
//   forward:  ODPS regression + OPTRAN WLO  -> clamped level-to-space path
//             (ODPS_AtmAbsorption.f90:133-256, :664-727)
//             -> path interpolated to user levels, nadir layer OD
//             -> Planck (FP64), Emission RT (Emission_Module.f90:54-174)
//             -> BT (FP64, CRTM_Planck_Functions.f90:470-492)
//   adjoint:  Pre_Process_RTSolution_K (CRTM_K_Matrix_Module.f90:1725-1761),
//             CRTM_Emission_AD (Emission_Module.f90:280-395), Planck AD,
//             ODPS_Compute_AtmAbsorption_AD (ODPS_AtmAbsorption.f90:487-600),
//             Add_OPTRAN_wloOD_AD (:911-984), ODPS_Compute_Predictor_AD
//             (ODPS_Predictor.f90:1783-2643), ODPS_Compute_Predictor_ODAS_AD
//             (:3092-3231), Map_Input_AD (ODPS_CoordinateMapping.f90:480-547)
// all in one pass over the profile's layers, with nothing per-layer stored
// in HBM.  RT is the arithmetic type of absorption and RT (float for the
// split path, double for the plain FP64 path); Planck and the BT inversion
// are always FP64.
#include "crtm_device.h"

namespace crtm {

namespace {
constexpr double MAX_OD = 20.0;
constexpr double LIMIT_EXP = 20.0, LIMIT_LOG = 4.8e+08;

template <class RT> __device__ inline RT rexp(RT x);
template <> __device__ inline float rexp<float>(float x) { return expf(x); }
template <> __device__ inline double rexp<double>(double x) { return exp(x); }

template <class RT, class KT>
__device__ inline void store_k(void* base, int64_t idx, RT v) {
  reinterpret_cast<KT*>(base)[idx] = (KT)v;
}
}  // namespace

// Adjoint accumulators of the per-layer predictor intermediates
template <class RT> struct LayerAdj {
  RT T, T2, dt, dt2, h2o, h2o_a, h2o_s, h2o_r, h2o_r4, h2o_q, tz, tzp;
  RT gz[6], gt[6], ra[6];
  RT o3, o3_a, o3_r;
  RT co, co_a, co_r, co_s, co_q, ch4, ch4_a, ch4_r, ch4_q, n2o, n2o_a, n2o_r, n2o_s;
};

// Adjoint of component j's predictors at one layer: xa[i] = X_AD(k, i, j).
// Partial derivatives use the wave-uniform table u (ODPS_Predictor.f90:1894-2641).
template <int G, int J, class RT>
__device__ inline void comp_adjoint(const RT* xa, const RT* u, LayerAdj<RT>& a) {
  const RT sec = u[U_SEC];
  if constexpr (J == 0) {
    a.T += xa[1] * sec + xa[3];
    a.T2 += xa[2] * sec + xa[5];
    a.tz += xa[6];
  } else if constexpr (G == 3 && J == 1) {
    const RT iT = u[U_IT], iT2 = u[U_IT2], iT4 = u[U_IT4], h2o = u[U_H2O], h2o_a = u[U_H2OA];
    a.h2o_a += xa[0] * iT + xa[1] * h2o * iT + xa[2] * h2o * iT4 + xa[3] * iT2 + xa[4] * h2o * iT2 +
               xa[5] * iT4 + xa[6] + xa[7] * u[U_DT] + xa[11] * u[U_H2OS];
    a.h2o += (xa[1] * iT + xa[2] * iT4 + xa[4] * iT2) * h2o_a;
    a.T -= (xa[0] + xa[1] * h2o) * h2o_a * iT * iT;
    a.T2 -= (RT(2) * xa[2] * h2o * iT4 * iT2 + xa[3] * iT4 + xa[4] * h2o * iT4 + RT(2) * xa[5] * iT4 * iT2) * h2o_a;
    a.dt += xa[7] * h2o_a;
    a.gz[0] += xa[8] * RT(2) * sec * sec * u[U_H2OP] + xa[9] * sec;
    a.h2o_s += xa[11] * h2o_a + xa[12] * RT(2) * u[U_H2OS];
    a.h2o_q += xa[13];
  } else if constexpr (J == 1) {  // IR water-vapour lines
    const RT h2o_a = u[U_H2OA], dt = u[U_DT], h2o_r = u[U_H2OR], h2o_s = u[U_H2OS];
    a.h2o_a += xa[0] + xa[1] * dt + xa[3] * u[U_DT2] + xa[5] * h2o_s;
    a.dt += xa[1] * h2o_a + xa[7] * h2o_r;
    a.h2o_s += xa[2] + xa[5] * h2o_a + xa[8] * RT(2) * h2o_s;
    a.dt2 += xa[3] * h2o_a;
    a.h2o_r4 += xa[4];
    a.h2o_r += xa[6] + xa[7] * dt + xa[10] * u[U_H2ODTZP];
    a.h2o_q += xa[9] + xa[10] * h2o_r;
    a.gz[0] += xa[11] * RT(2) * sec * sec * u[U_H2OP] + xa[12] * sec;
    a.ra[2] += xa[14] * sec;
    if constexpr (G == 1) {
      a.ch4_a += xa[15] + xa[16] * RT(2) * u[U_CH4A] * dt;
      a.dt += xa[16] * u[U_CH4A] * u[U_CH4A];
      a.co_a += xa[17];
    }
  } else if constexpr (J == 2) {  // IR water-vapour continuum
    const RT iT = u[U_IT], iT2 = u[U_IT2], iT4 = u[U_IT4], h2o = u[U_H2O], h2o_a = u[U_H2OA];
    a.h2o_a += xa[0] * iT + xa[1] * h2o * iT + xa[2] * h2o * iT4 + xa[3] * iT2 + xa[4] * h2o * iT2 +
               xa[5] * iT4 + xa[6];
    a.h2o += (xa[1] * iT + xa[2] * iT4 + xa[4] * iT2) * h2o_a;
    a.T -= (xa[0] + xa[1] * h2o) * h2o_a * iT * iT;
    a.T2 -= (RT(2) * xa[2] * h2o * iT4 * iT2 + xa[3] * iT4 + xa[4] * h2o * iT4 + RT(2) * xa[5] * iT4 * iT2) * h2o_a;
  } else if constexpr (J == 3) {  // ozone
    const RT o3 = u[U_O3], o3_a = u[U_O3A], o3_r = u[U_O3R], o3p = u[U_O3P], dt = u[U_DT];
    a.o3_a += xa[0] + xa[1] * dt + xa[2] * o3 * o3p + RT(2) * xa[3] * o3_a + xa[4] * o3p + xa[5] * u[U_SP];
    a.dt += xa[1] * o3_a + xa[6] * o3_r;
    a.o3_r += xa[6] * dt + xa[7] + xa[8] * o3 / o3p;
    a.o3 += xa[2] * o3_a * o3p + xa[8] * o3_r / o3p;
    a.gz[1] += xa[2] * o3_a * o3 + xa[4] * o3_a + xa[5] * RT(0.5) * o3_a * u[U_SQSO3P] -
               xa[8] * o3_r * o3 / (o3p * o3p) + xa[9] * sec + xa[10] * RT(2) * sec * sec * o3p;
  } else if constexpr (J == 4) {  // CO2
    const RT tzp = u[U_TZP], sT = u[U_ST];
    a.T += xa[0] * sec + xa[2] + xa[9] * sec * tzp * RT(0.5) / sT;
    a.T2 += xa[1] * sec + xa[3];
    a.ra[2] += xa[5] * sec;
    a.tzp += xa[6] * sec + xa[8] * RT(3) * tzp * tzp + xa[9] * sec * sT;
    a.gz[2] += xa[7] * RT(2) * sec * sec * u[U_CO2P];
    if constexpr (G == 1) a.co_a += xa[10];
  } else if constexpr (J == 5) {  // N2O
    const RT dt = u[U_DT], n2o_a = u[U_N2OA], n2o = u[U_N2O], n2o_r = u[U_N2OR], n2ozp = u[U_N2OZP];
    a.n2o_a += xa[0] * dt + xa[3] * RT(0.25) * u[U_N2OPOW] + xa[4];
    a.dt += xa[0] * n2o_a + xa[2] * n2o;
    a.n2o_r += xa[1] + xa[9] * n2o / n2ozp;
    a.n2o += xa[2] * dt + xa[9] * n2o_r / n2ozp;
    a.gz[3] += xa[5] * sec - xa[9] * n2o_r * n2o / (n2ozp * n2ozp);
    a.gt[3] += xa[6] * sec + xa[8];
    a.n2o_s += xa[7];
    a.ch4_a += xa[10] + xa[11] * u[U_CH4ZP];
    a.gz[5] += xa[11] * u[U_CH4A];
    a.co_a += xa[12] + xa[13] * sec * u[U_COZP];
    a.gz[4] += xa[13] * u[U_COA] * sec;
  } else if constexpr (J == 6) {  // CO
    const RT dt = u[U_DT], co_a = u[U_COA], co_r = u[U_COR], co_q = u[U_COQ];
    a.co_a += xa[0] + xa[1] * dt + xa[6] * u[U_DT2];
    a.dt += xa[1] * co_a + xa[3] * co_r;
    a.co_r += xa[2] * RT(0.5) / u[U_SQCOR] + xa[3] * dt + xa[5] - xa[8] * co_q / (co_r * co_r);
    a.co_s += xa[4];
    a.dt2 += xa[6] * co_a;
    a.co_q += xa[7] + xa[8] / co_r + xa[9] * u[U_SQCOZP];
    a.gz[4] += xa[9] * RT(0.5) * co_q / u[U_SQCOZP];
  } else if constexpr (J == 7) {  // CH4
    const RT dt = u[U_DT], ch4_a = u[U_CH4A], ch4 = u[U_CH4], ch4_r = u[U_CH4R], ch4zp = u[U_CH4ZP];
    a.ch4_a += xa[0] * dt + xa[2] * RT(2) * ch4_a + xa[3];
    a.dt += xa[0] * ch4_a + xa[4] * ch4;
    a.ch4_r += xa[1] + xa[7] * RT(0.5) / u[U_SQCH4R] + xa[10] * ch4 / ch4zp;
    a.ch4 += xa[4] * dt + xa[10] * ch4_r / ch4zp;
    a.ch4_q += xa[5] + xa[6] * RT(2) * u[U_CH4Q];
    a.gt[5] += xa[8] + xa[9] * sec;
    a.gz[5] -= xa[10] * ch4_r * ch4 / (ch4zp * ch4zp);
  }
}

// Close the layer: chain the intermediates back to T, the absorber ratios
// and the integrated ratios (tail of ODPS_Compute_Predictor_AD).
template <int G, class RT>
__device__ inline void finish_layer(const RT* u, LayerAdj<RT>& a) {
  const RT sec = u[U_SEC];
  if constexpr (G == 1) {
    a.gz[5] += sec * a.ch4_q;
    a.ch4_a += RT(0.5) / u[U_CH4R] * a.ch4_r;
    a.ch4 += sec * a.ch4_a;
    const RT co_a = u[U_COA], co = u[U_CO], cozp = u[U_COZP];
    a.gz[4] -= a.co_q * co_a * co / (cozp * cozp);
    a.co_a += a.co_q * co / cozp + a.co_s * RT(2) * co_a + a.co_r * RT(0.5) / u[U_COR];
    a.co += a.co_q * co_a / cozp + a.co_a * sec;
    a.n2o_a += a.n2o_r * RT(0.5) / u[U_N2OR] + a.n2o_s * RT(2) * u[U_N2OA];
    a.n2o += a.n2o_a * sec;
    a.ra[5] += a.ch4;
    a.ra[3] += a.n2o;
    a.ra[4] += a.co;
  }
  if constexpr (G == 1 || G == 2) {
    a.o3_a += a.o3_r * RT(0.5) / u[U_O3R];
    a.o3 += a.o3_a * sec;
    a.ra[1] += a.o3;
  }
  const RT gt0 = u[U_GT0];
  a.gt[0] -= a.h2o_q * u[U_H2O] / (gt0 * gt0);
  a.h2o_r += a.h2o_r4 * RT(0.5) / u[U_H2OR4];
  a.h2o_a += a.h2o_s * RT(2) * u[U_H2OA] + a.h2o_r * RT(0.5) / u[U_H2OR];
  a.h2o += a.h2o_a * sec + a.h2o_q / gt0;
  a.ra[0] += a.h2o;
  a.dt += a.dt2 * RT(2) * u[U_ADT];
  a.T += a.T2 * RT(2) * u[U_T];
}

template <int G, class RT, int J>
__device__ inline void comp_layer_ad(const float* ck, const int* npb, const int* coff, RT od_ad, const RT* uk,
                                     LayerAdj<RT>& a) {
  using Tg = GT<G>;
  if constexpr (J < Tg::NCOMP) {
    const int np = npb[J];
    if (np > 0) {
      RT xa[Tg::NP[J]];
      const float* cj = ck + coff[J] * WAVE;
#pragma unroll
      for (int i = 0; i < Tg::NP[J]; ++i) xa[i] = (i < np) ? (RT)cj[i * WAVE] * od_ad : RT(0);
      comp_adjoint<G, J, RT>(xa, uk, a);
    }
    comp_layer_ad<G, RT, J + 1>(ck, npb, coff, od_ad, uk, a);
  }
}

template <int G, class RT, int J>
__device__ inline RT comp_layer_od(const float* ck, const int* npb, const int* coff, const RT* xk) {
  using Tg = GT<G>;
  if constexpr (J < Tg::NCOMP) {
    RT od = RT(0);
    const int np = npb[J];
    if (np > 0) {
      const float* cj = ck + coff[J] * WAVE;
      const RT* xj = xk + Tg::XOFF[J];
#pragma unroll
      for (int i = 0; i < Tg::NP[J]; ++i)
        if (i < np) od += (RT)cj[i * WAVE] * xj[i];
    }
    return od + comp_layer_od<G, RT, J + 1>(ck, npb, coff, xk);
  } else {
    return RT(0);
  }
}

template <int G, class RT, class KT, bool DOK>
__global__ void __launch_bounds__(WAVE) clear_fwdk_kernel(DevSensor s, DevBatch bt, ProfWork<RT> w, DevOut o) {
  using Tg = GT<G>;
  const int nb = s.n_blocks;
  const int wg = blockIdx.x;
  const int b = wg % nb;
  const int m = wg / nb;
  const int lane = threadIdx.x;
  const int l = b * WAVE + lane;
  const bool valid = l < s.L;
  const int lc = valid ? l : s.L - 1;
  const int lg = s.chan_offset + lc;
  const int64_t ol = (int64_t)m * bt.Ltot + lg;
  if (w.status[m] != 0) {
    if (valid) {
      o.radiance[ol] = __builtin_nan("");
      o.bt[ol] = __builtin_nan("");
    }
    return;
  }
  const int K = bt.n_layers[m];
  const int64_t loff = bt.loff[m];
  const int Ko = s.Ko;
  const int* meta = s.blk_meta + b * BM_N;
  const int nsl = meta[BM_NSL];
  int npb[Tg::NCOMP], coff[Tg::NCOMP];
#pragma unroll
  for (int j = 0; j < Tg::NCOMP; ++j) {
    npb[j] = meta[BM_NP0 + j];
    coff[j] = meta[BM_OFF0 + j];
  }
  const float* cb = s.cblk + s.blk_base[b] + lane;
  const int64_t ob = s.optran ? s.oblk_base[b] : -1;
  const bool blk_opt = ob >= 0;
  const bool lane_opt = blk_opt && s.lane_optran[b * WAVE + lane] != 0;
  const double* ocb = blk_opt ? s.ocd + ob + lane : nullptr;
  const int opmask = meta[BM_OPMASK], omaxo = meta[BM_OMAXO];
  const RT* X = w.x + (size_t)m * Ko * Tg::NX;
  const RT* Ut = w.u + (size_t)m * Ko * NU;
  const RT* Ot = w.o + (size_t)m * Ko * NO;

  // ---------------- forward: ODPS absorption ----------------
  RT path[KO_MAX + 1];
  RT chi[KO_MAX];
  uint64_t odclamp[2] = {0, 0}, lcclamp[2] = {0, 0};
  path[0] = RT(0);
  for (int k = 0; k < Ko; ++k) {
    const float* ck = cb + (int64_t)k * nsl * WAVE;
    RT od = comp_layer_od<G, RT, 0>(ck, npb, coff, X + (size_t)k * Tg::NX);
    if (blk_opt) {
      const RT* ok = Ot + (size_t)k * NO;
      RT lnchi = RT(0);
#pragma unroll
      for (int p = 0; p < OPZ; ++p) {
        if ((opmask >> p) & 1) {
          RT bp = RT(0);
#pragma unroll
          for (int q = 0; q < OPO; ++q)
            if (q <= omaxo) bp += (RT)ocb[(p * OPO + q) * WAVE] * ok[O_AP0 + q];
          lnchi += bp * ok[O_X0 + p];
        }
      }
      if (lane_opt) {
        RT c;
        if (lnchi > RT(LIMIT_EXP)) {
          c = RT(LIMIT_LOG);
          lcclamp[k >> 6] |= 1ull << (k & 63);
        } else if (lnchi < RT(-LIMIT_EXP)) {
          c = RT(0);
          lcclamp[k >> 6] |= 1ull << (k & 63);
        } else {
          c = rexp<RT>(lnchi);
        }
        chi[k] = c;
        od += c * ok[O_DA];
      }
    }
    RT odc = od;
    if (od < RT(0)) {
      odc = RT(0);
      odclamp[k >> 6] |= 1ull << (k & 63);
    } else if (od > RT(MAX_OD)) {
      odc = RT(MAX_OD);
      odclamp[k >> 6] |= 1ull << (k & 63);
    }
    path[k + 1] = path[k] + odc;
  }

  // ---------------- forward: user layers, Planck, Emission ----------------
  const double* geo = w.geo + (size_t)m * 4;
  const double sec_sfc = geo[0], cosu = geo[1], tsfc = geo[2];
  const RT rsec = (RT)sec_sfc;
  const RT u = (RT)cosu;
  const int* ui = w.uidx + (size_t)m * (bt.max_k + 1);
  const RT* uf = w.ufrac + (size_t)m * (bt.max_k + 1);
  auto upath = [&](int i) -> RT {
    const int k1 = ui[i];
    const RT f = uf[i];
    return f < RT(0) ? path[k1] : path[k1] + (path[k1 + 1] - path[k1]) * f;
  };
  const double c1 = s.c1[lc], c2 = s.c2[lc], bc1 = s.bc1[lc], bc2 = s.bc2[lc];
  const double* tuser = bt.t + loff;
  RT tr[K_MAX], Bk[K_MAX], rd[K_MAX + 1], ru[K_MAX + 1];
  rd[0] = (RT)s.cosmic[lc];
  RT upp = upath(0);
  for (int k = 0; k < K; ++k) {
    const RT up = upath(k + 1);
    const RT tau = (up - upp) / rsec;
    upp = up;
    const RT t = rexp<RT>(-tau / u);
    const double te = bc1 + bc2 * tuser[k];
    const double B = c1 / (exp(c2 / te) - 1.0);
    tr[k] = t;
    Bk[k] = (RT)B;
    rd[k + 1] = rd[k] * t + (RT)B * (RT(1) - t);
  }
  const double tes = bc1 + bc2 * tsfc;
  const double Bs_d = c1 / (exp(c2 / tes) - 1.0);
  const RT Bs = (RT)Bs_d;
  const double eps_d = bt.emissivity[(size_t)m * bt.Ltot + lg];
  const RT eps = (RT)eps_d;
  ru[K] = eps * Bs + (RT(1) - eps) * rd[K];
  RT upr = RT(0);
  for (int k = K - 1; k >= 0; --k) {
    const RT src = Bk[k] * (RT(1) - tr[k]);
    ru[k] = ru[k + 1] * tr[k] + src;
    upr = upr * tr[k] + src;
  }
  const double R = (double)ru[0];
  const double bt_v = (c2 / log(c1 / R + 1.0) - bc1) / bc2;
  if (valid) {
    o.radiance[ol] = R;
    o.bt[ol] = bt_v;
    if (o.up_radiance) o.up_radiance[ol] = (double)upr;
    if (o.down_radiance) o.down_radiance[ol] = (double)rd[K];
    if (o.sfc_planck) o.sfc_planck[ol] = Bs_d;
    if (o.sfc_emis) o.sfc_emis[ol] = eps_d;
  }
  if constexpr (!DOK) return;

  // ---------------- adjoint: BT -> radiance, Emission_AD ----------------
  const double arg = c1 / R + 1.0;
  const double lga = log(arg);
  const double F = c1 * c2 / (bc2 * arg * (R * lga) * (R * lga));
  const double rad_ad_d = (o.rad_k ? o.rad_k[ol] : 0.0) + F * (o.bt_k ? o.bt_k[ol] : 0.0);
  if (valid && o.rad_k) o.rad_k[ol] = rad_ad_d;
  const RT rad_ad = (RT)rad_ad_d;
  RT tau_ad[K_MAX], b_ad[K_MAX];
  RT upad = rad_ad;
  for (int k = 0; k < K; ++k) {
    const RT t = tr[k];
    const RT a = t * (ru[k + 1] - Bk[k]) * upad;
    b_ad[k] = upad * (RT(1) - t);
    upad *= t;
    tau_ad[k] = -a / u;
  }
  const RT eps_ad = upad * Bs, bs_ad = eps * upad, r_ad = upad * rd[K];
  RT dnad = (RT(1) - eps) * upad;
  if (valid) {
    if (o.emis_k) o.emis_k[ol] = (double)(eps_ad - r_ad);
    const double dbsdt = c2 * bc2 * Bs_d * (Bs_d + c1) / (c1 * tes * tes);
    const double sad = dbsdt * (double)bs_ad;
    for (int c = 0; c < 4; ++c)
      if (o.sfc_t_k[c]) o.sfc_t_k[c][ol] = bt.cov[c][m] * sad;
  }
  RT path_ad[KO_MAX + 1];
  for (int k = 0; k <= Ko; ++k) path_ad[k] = RT(0);
  auto scatter = [&](int i, RT v) {
    const int k1 = ui[i];
    const RT f = uf[i];
    if (f < RT(0)) {
      path_ad[k1] += v;
    } else {
      path_ad[k1] += (RT(1) - f) * v;
      path_ad[k1 + 1] += f * v;
    }
  };
  RT tk_acc[K_MAX];
  RT carry = RT(0);
  for (int k = K - 1; k >= 0; --k) {
    const RT t = tr[k];
    const RT a = t * (rd[k] - Bk[k]) * dnad;
    const RT bad = b_ad[k] + dnad * (RT(1) - t);
    dnad *= t;
    const RT tad = tau_ad[k] - a / u;
    const double te = bc1 + bc2 * tuser[k];
    const double Bd = (double)Bk[k];
    tk_acc[k] = (RT)(c2 * bc2 * Bd * (Bd + c1) / (c1 * te * te)) * bad;
    const RT gq = tad / rsec;
    scatter(k + 1, carry + gq);
    carry = -gq;
  }
  scatter(0, carry);

  // ---------------- adjoint: ODPS regression, predictors, Map_Input ----------------
  const int na = bt.na;
  RT ak_acc[NA_MAX][K_MAX];
  for (int j = 0; j < na; ++j)
    for (int k = 0; k < K; ++k) ak_acc[j][k] = RT(0);
  RT S = RT(0), tzs = RT(0), tzps = RT(0);
  RT gzs[Tg::NABS], gts[Tg::NABS];
#pragma unroll
  for (int j = 0; j < Tg::NABS; ++j) gzs[j] = gts[j] = RT(0);
  RT ivn = RT(0), sta = RT(0), spa = RT(0), xltc = RT(0), xlpc = RT(0);
  const int h2o_j = s.h2o_idx - 1;
  const int* acc_i = w.acc_i + (size_t)m * Ko * 2;
  const RT* acc_w = w.acc_w + (size_t)m * Ko * MAXW;
  for (int k = Ko - 1; k >= 0; --k) {
    S += path_ad[k + 1];
    const RT od_ad = ((odclamp[k >> 6] >> (k & 63)) & 1) ? RT(0) : S;
    const RT* uk = Ut + (size_t)k * NU;
    LayerAdj<RT> a{};
    const float* ck = cb + (int64_t)k * nsl * WAVE;
    comp_layer_ad<G, RT, 0>(ck, npb, coff, od_ad, uk, a);
    RT t_ad = RT(0), vap_ad = RT(0);
    if (blk_opt) {
      const RT* ok = Ot + (size_t)k * NO;
      const RT chi_ad = od_ad * ok[O_DA];
      const RT da_ad = lane_opt ? od_ad * chi[k] : RT(0);
      const bool lcc = (lcclamp[k >> 6] >> (k & 63)) & 1;
      const RT lc_ad = (lane_opt && !lcc) ? chi[k] * chi_ad : RT(0);
      RT ox_ad[OPZ], ap_ad[OPO];
#pragma unroll
      for (int p = 0; p < OPZ; ++p) ox_ad[p] = RT(0);
#pragma unroll
      for (int q = 0; q < OPO; ++q) ap_ad[q] = RT(0);
#pragma unroll
      for (int p = 0; p < OPZ; ++p) {
        if ((opmask >> p) & 1) {
          RT bp = RT(0);
#pragma unroll
          for (int q = 0; q < OPO; ++q)
            if (q <= omaxo) bp += (RT)ocb[(p * OPO + q) * WAVE] * ok[O_AP0 + q];
          ox_ad[p] += lc_ad * bp;
          const RT bad = lc_ad * ok[O_X0 + p];
#pragma unroll
          for (int q = 0; q < OPO; ++q)
            if (q <= omaxo) ap_ad[q] += (RT)ocb[(p * OPO + q) * WAVE] * bad;
        }
      }
      // ODPS_Compute_Predictor_ODAS_AD at layer k (reverse running state)
      RT iv_ad = ivn;
      RT ap1_ad = RT(0);
#pragma unroll
      for (int q = 1; q < OPO; ++q) ap1_ad += RT(q) * ok[O_AP0 + q - 1] * ap_ad[q];
      const RT ave_ad = ap1_ad * ok[O_IDEN];
      const RT xlt_ad = xltc + ox_ad[12], xlp_ad = xlpc + ox_ad[13];
      xltc = ox_ad[12];
      xlpc = ox_ad[13];
      sta += RT(0.5) * ok[O_INV] * xlt_ad;
      spa += RT(0.5) * ok[O_INV] * xlp_ad;
      const RT inv_ad = RT(0.5) * ok[O_ST] * xlt_ad + RT(0.5) * ok[O_SP] * xlp_ad;
      iv_ad -= inv_ad * ok[O_IIV2];
      RT dab_ad = da_ad + ok[O_PR] * spa + ok[O_T] * sta;
      t_ad += ok[O_DA] * sta;
      iv_ad += RT(0.5) * ave_ad;
      ivn = RT(0.5) * ave_ad + iv_ad;
      dab_ad += iv_ad;
      vap_ad += ok[O_DPSEC] * dab_ad;
      const RT tk = ok[O_T], pk = ok[O_PR];
      t_ad += ox_ad[1] + pk * ox_ad[5] + pk * pk * ox_ad[7] +
              RT(2) * tk * (ox_ad[3] + pk * ox_ad[6] + pk * pk * ox_ad[8] - ok[O_VT4] * ox_ad[11]);
      vap_ad += ox_ad[10] + ox_ad[11] / (tk * tk);
    }
    finish_layer<G, RT>(uk, a);
    t_ad += a.T * uk[U_IRT] + a.dt;
    RT a_ad[Tg::NABS];
#pragma unroll
    for (int j = 0; j < Tg::NABS; ++j) a_ad[j] = a.ra[j] * uk[U_IRA0 + j];
    if (s.optran) a_ad[h2o_j] += vap_ad;
    const RT pdp = uk[U_PDP];
#pragma unroll
    for (int j = 0; j < Tg::NABS; ++j) {
      gts[j] += a.gt[j] * uk[U_IGATZP0 + j];
      gzs[j] += a.gz[j] * uk[U_IGAZP0 + j];
      t_ad += gts[j] * pdp * uk[U_AODPS0 + j];
      a_ad[j] += gzs[j] * pdp + gts[j] * pdp * uk[U_TODPS];
    }
    tzps += a.tzp * uk[U_ITZPREF];
    tzs += a.tz * uk[U_ITZREF];
    t_ad += tzs + pdp * tzps;
    // Map_Input_AD: scatter onto the user layers averaged into this ODPS layer
    const int i1 = acc_i[2 * k], n = acc_i[2 * k + 1];
    const RT* aw = acc_w + (size_t)k * MAXW;
    for (int q = 0; q < n; ++q) tk_acc[i1 + q] += aw[q] * t_ad;
#pragma unroll
    for (int j = 0; j < Tg::NABS; ++j) {
      const int ju = bt.jmap[j];
      if (ju >= 0 && uk[U_MASK0 + j] != RT(0))
        for (int q = 0; q < n; ++q) ak_acc[ju][i1 + q] += aw[q] * a_ad[j];
    }
  }
  if (valid) {
    const int64_t Lt = bt.Ltot;
    for (int k = 0; k < K; ++k) store_k<RT, KT>(o.t_k, (loff + k) * Lt + lg, tk_acc[k]);
    for (int j = 0; j < na; ++j)
      for (int k = 0; k < K; ++k)
        store_k<RT, KT>(o.a_k, ((int64_t)j * bt.NT + loff + k) * Lt + lg, ak_acc[j][k]);
  }
}

template <int G, class RT, class KT, bool DOK>
void launch_fwdk(const DevSensor& s, const DevBatch& b, const ProfWork<RT>& w, const DevOut& o,
                 hipStream_t stream) {
  const int64_t n = (int64_t)s.n_blocks * b.M;
  hipLaunchKernelGGL((clear_fwdk_kernel<G, RT, KT, DOK>), dim3((unsigned)n), dim3(WAVE), 0, stream, s, b, w, o);
}

#define INST(G, RT, KT, DOK) \
  template void launch_fwdk<G, RT, KT, DOK>(const DevSensor&, const DevBatch&, const ProfWork<RT>&, \
                                            const DevOut&, hipStream_t);
#define INST_G(G)                                                                        \
  INST(G, float, float, true) INST(G, float, double, true) INST(G, float, float, false)  \
  INST(G, double, double, true) INST(G, double, float, true) INST(G, double, double, false)
INST_G(1)
INST_G(2)
INST_G(3)

}  // namespace crtm
