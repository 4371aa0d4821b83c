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
    const RT io3p = u[U_IO3P];
    a.o3_r += xa[6] * dt + xa[7] + xa[8] * o3 * io3p;
    a.o3 += xa[2] * o3_a * o3p + xa[8] * o3_r * io3p;
    a.gz[1] += xa[2] * o3_a * o3 + xa[4] * o3_a + xa[5] * RT(0.5) * o3_a * u[U_SQSO3P] -
               xa[8] * o3_r * o3 * io3p * io3p + xa[9] * sec + xa[10] * RT(2) * sec * sec * o3p;
  } else if constexpr (J == 4) {  // CO2
    const RT tzp = u[U_TZP], sT = u[U_ST];
    a.T += xa[0] * sec + xa[2] + xa[9] * sec * tzp * RT(0.5) * u[U_IST];
    a.T2 += xa[1] * sec + xa[3];
    a.ra[2] += xa[5] * sec;
    a.tzp += xa[6] * sec + xa[8] * RT(3) * tzp * tzp + xa[9] * sec * sT;
    a.gz[2] += xa[7] * RT(2) * sec * sec * u[U_CO2P];
    if constexpr (G == 1) a.co_a += xa[10];
  } else if constexpr (J == 5) {  // N2O
    const RT dt = u[U_DT], n2o_a = u[U_N2OA], n2o = u[U_N2O], n2o_r = u[U_N2OR], n2ozp = u[U_N2OZP];
    a.n2o_a += xa[0] * dt + xa[3] * RT(0.25) * u[U_N2OPOW] + xa[4];
    a.dt += xa[0] * n2o_a + xa[2] * n2o;
    const RT in2ozp = u[U_IN2OZP];
    a.n2o_r += xa[1] + xa[9] * n2o * in2ozp;
    a.n2o += xa[2] * dt + xa[9] * n2o_r * in2ozp;
    a.gz[3] += xa[5] * sec - xa[9] * n2o_r * n2o * in2ozp * in2ozp;
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
    const RT icor = u[U_ICOR];
    a.co_r += xa[2] * RT(0.5) * u[U_ISQCOR] + xa[3] * dt + xa[5] - xa[8] * co_q * icor * icor;
    a.co_s += xa[4];
    a.dt2 += xa[6] * co_a;
    a.co_q += xa[7] + xa[8] * icor + xa[9] * u[U_SQCOZP];
    a.gz[4] += xa[9] * RT(0.5) * co_q * u[U_ISQCOZP];
  } else if constexpr (J == 7) {  // CH4
    const RT dt = u[U_DT], ch4_a = u[U_CH4A], ch4 = u[U_CH4], ch4_r = u[U_CH4R], ch4zp = u[U_CH4ZP];
    a.ch4_a += xa[0] * dt + xa[2] * RT(2) * ch4_a + xa[3];
    a.dt += xa[0] * ch4_a + xa[4] * ch4;
    const RT ich4zp = u[U_ICH4ZP];
    a.ch4_r += xa[1] + xa[7] * RT(0.5) * u[U_ISQCH4R] + xa[10] * ch4 * ich4zp;
    a.ch4 += xa[4] * dt + xa[10] * ch4_r * ich4zp;
    a.ch4_q += xa[5] + xa[6] * RT(2) * u[U_CH4Q];
    a.gt[5] += xa[8] + xa[9] * sec;
    a.gz[5] -= xa[10] * ch4_r * ch4 * ich4zp * ich4zp;
  }
}

// Close the layer: chain the intermediates back to T, the absorber ratios
// and the integrated ratios (tail of ODPS_Compute_Predictor_AD).
template <int G, class RT>
__device__ inline void finish_layer(const RT* u, LayerAdj<RT>& a) {
  const RT sec = u[U_SEC];
  if constexpr (G == 1) {
    a.gz[5] += sec * a.ch4_q;
    a.ch4_a += RT(0.5) * u[U_ICH4R] * a.ch4_r;
    a.ch4 += sec * a.ch4_a;
    const RT co_a = u[U_COA], co = u[U_CO], cozp = u[U_COZP];
    const RT icozp = u[U_ICOZP];
    a.gz[4] -= a.co_q * co_a * co * icozp * icozp;
    a.co_a += a.co_q * co * icozp + a.co_s * RT(2) * co_a + a.co_r * RT(0.5) * u[U_ICOR];
    a.co += a.co_q * co_a * icozp + a.co_a * sec;
    a.n2o_a += a.n2o_r * RT(0.5) * u[U_IN2OR] + a.n2o_s * RT(2) * u[U_N2OA];
    a.n2o += a.n2o_a * sec;
    a.ra[5] += a.ch4;
    a.ra[3] += a.n2o;
    a.ra[4] += a.co;
  }
  if constexpr (G == 1 || G == 2) {
    a.o3_a += a.o3_r * RT(0.5) * u[U_IO3R];
    a.o3 += a.o3_a * sec;
    a.ra[1] += a.o3;
  }
  const RT igt0 = u[U_IGT0];
  a.gt[0] -= a.h2o_q * u[U_H2O] * igt0 * igt0;
  a.h2o_r += a.h2o_r4 * RT(0.5) * u[U_IH2OR4];
  a.h2o_a += a.h2o_s * RT(2) * u[U_H2OA] + a.h2o_r * RT(0.5) * u[U_IH2OR];
  a.h2o += a.h2o_a * sec + a.h2o_q * igt0;
  a.ra[0] += a.h2o;
  a.dt += a.dt2 * RT(2) * u[U_ADT];
  a.T += a.T2 * RT(2) * u[U_T];
}

// The layer's predictor adjoint is linear in X_AD(k, :, :) = C(:, :, k) * OD_AD(k)
// with wave-uniform partials, so it runs on the coefficients alone and the
// caller scales the layer's results by OD_AD(k) once.
template <int G, class RT, int J>
__device__ inline void comp_layer_ad(const float* ck, const int* npb, const RT* uk, LayerAdj<RT>& a) {
  using Tg = GT<G>;
  if constexpr (J < Tg::NCOMP) {
    if (npb[J] > 0) {   // block-uniform: no channel of the block uses component J
      RT xa[Tg::NP[J]];
      const float* cj = ck + Tg::XOFF[J] * WAVE;
      // slots beyond a channel's predictor count hold zeros
#pragma unroll
      for (int i = 0; i < Tg::NP[J]; ++i) xa[i] = (RT)cj[i * WAVE];
      comp_adjoint<G, J, RT>(xa, uk, a);
    }
    comp_layer_ad<G, RT, J + 1>(ck, npb, uk, a);
  }
}

template <int G, class RT, int J>
__device__ inline RT comp_layer_od(const float* ck, const int* npb, const RT* xk) {
  using Tg = GT<G>;
  if constexpr (J < Tg::NCOMP) {
    RT od = RT(0);
    const int np = npb[J];
    if (np > 0) {
      const float* cj = ck + Tg::XOFF[J] * WAVE;
      const RT* xj = xk + Tg::XOFF[J];
#pragma unroll
      for (int i = 0; i < Tg::NP[J]; ++i) od += (RT)cj[i * WAVE] * xj[i];
    }
    return od + comp_layer_od<G, RT, J + 1>(ck, npb, xk);
  } else {
    return RT(0);
  }
}

// Workgroup shape per arithmetic type, from measurements at the benchmark
// shape (2211 channels x 1024 profiles): FP32 runs best with 16 profiles per
// workgroup at 4 waves per SIMD (<= 128 VGPRs); FP64 needs more registers and
// runs best with 8 profiles, a single LDS buffer and no wave-count target.
//   PW   profiles (waves) per workgroup, all on one 64-channel block
//   NBUF LDS buffers per staged layer (2: the next layer's stores overlap
//        the current layer's reads; 1: half the LDS, one more barrier)
//   WPE  waves per SIMD the register allocation targets (0: compiler's choice)
template <class RT> struct KCfg { static constexpr int PW = 16, NBUF = 2, WPE = 4; };
template <> struct KCfg<double> { static constexpr int PW = 8, NBUF = 1, WPE = 0; };
// LDS row of one wave: predictors X or adjoint table U, followed by the OPTRAN table O
constexpr int TABW = (NX_MAX > NU ? NX_MAX : NU) + NO;
constexpr int RQ = (TABW + WAVE - 1) / WAVE;   // row elements loaded per lane
constexpr int OC_LDS_BYTES = 16384;            // OPTRAN coefficients staged when they fit
constexpr int MW = MAP_WIN;                    // Map_Input_AD register window (ODPS layers)
constexpr int NAW = 2;                         // user absorbers the window carries
constexpr int PF = 4;                          // user-layer rows prefetched one ODPS layer ahead

// OPTRAN ln(chi) of one lane at one layer (Add_OPTRAN_wloOD,
// ODPS_AtmAbsorption.f90:664-727): b_p = sum_q coef(p,q) Ap(q), lnchi = sum_p b_p OX(p)
template <class RT, class OC>
__device__ inline RT optran_lnchi(const OC& oc, const int* meta, const RT* ok) {
  const int nrow = meta[BM_ONROW], nq = meta[BM_OMAXO] + 1;
  RT lnchi = RT(0);
  for (int r = 0; r < nrow; ++r) {
    const int p = meta[BM_OROW0 + r];
    RT bp = RT(0);
    for (int q = 0; q < nq; ++q) bp += oc(r * nq + q) * ok[O_AP0 + q];
    lnchi += bp * ok[O_X0 + p];
  }
  return lnchi;
}

template <int G, class RT, class KT, bool DOK>
__global__ void __attribute__((amdgpu_flat_work_group_size(KCfg<RT>::PW * WAVE, KCfg<RT>::PW * WAVE),
                               amdgpu_waves_per_eu(KCfg<RT>::WPE > 0 ? KCfg<RT>::WPE : 1)))
clear_fwdk_kernel(DevSensor s, DevBatch bt, ProfWork<RT> w, DevOut o) {
  using Tg = GT<G>;
  constexpr int PW = KCfg<RT>::PW, NBUF = KCfg<RT>::NBUF, TPB = PW * WAVE;
  // Workgroup = one 64-channel block x PW profiles (one wave each).  The
  // block's coefficients of the current layer are staged once in LDS and
  // read by all PW waves, so each coefficient leaves L2 once per PW
  // profiles; each wave stages its own profile's uniform row.
  // XCD-aware order: workgroups are dispatched round-robin over the 8 XCDs,
  // so XCD x receives ids x, x+8, ...; it gets the contiguous range
  // [x*chunk, (x+1)*chunk) of the block-major (block, profile group) order,
  // so the workgroups resident on one XCD share coefficient blocks in its L2.
  const int ng = (bt.M + PW - 1) / PW;
  const int64_t total = (int64_t)s.n_blocks * ng;
  const int64_t chunk = (total + 7) / 8;
  const int64_t tix = (int64_t)(blockIdx.x & 7) * chunk + (blockIdx.x >> 3);
  if (tix >= total) return;                 // uniform over the workgroup
  const int b = s.blk_order[(int)(tix / ng)];
  const int tid = threadIdx.x;
  const int wv = tid >> 6;
  const int lane = tid & (WAVE - 1);
  const int m_raw = (int)(tix % ng) * PW + wv;
  const bool in_batch = m_raw < bt.M;
  const int m = in_batch ? m_raw : bt.M - 1;
  const bool ok_prof = in_batch && w.status[m] == 0 && !(o.fail && o.fail[m]);   // wave-uniform
  const int l = b * WAVE + lane;
  const bool lane_in = l < s.L;
  const bool valid = ok_prof && lane_in;
  const int lc = lane_in ? l : s.L - 1;
  const int lg = s.chan_offset + lc;
  const int64_t ol = (int64_t)m * bt.Ltot + lg;
  if (in_batch && !ok_prof && lane_in) {
    // a failed profile: NaN radiances, and NaN Jacobian rows so that no
    // gradient built from them passes silently
    o.radiance[ol] = __builtin_nan("");
    o.bt[ol] = __builtin_nan("");
    if constexpr (DOK) {
      const int64_t r0 = bt.out_loff[m], r1 = bt.out_loff[m + 1];
      const int64_t astr_f = bt.out_NT * bt.Ltot;
      for (int64_t r = r0; r < r1; ++r) {
        reinterpret_cast<KT*>(o.t_k)[r * bt.Ltot + lg] = KT(__builtin_nan(""));
        for (int ja = 0; ja < bt.na; ++ja)
          reinterpret_cast<KT*>(o.a_k)[ja * astr_f + r * bt.Ltot + lg] = KT(__builtin_nan(""));
      }
    }
  }
  const int K = ok_prof ? bt.n_layers[m] : 0;
  const int64_t loff = bt.loff[m];
  const int Ko = s.Ko;
  const int* meta = s.blk_meta + b * BM_N;
  int npb[Tg::NCOMP];
#pragma unroll
  for (int j = 0; j < Tg::NCOMP; ++j) npb[j] = meta[BM_NP0 + j];
  const int64_t ob = s.optran ? s.oblk_base[b] : -1;
  const bool blk_opt = ob >= 0;
  const bool lane_opt = blk_opt && s.lane_optran[b * WAVE + lane] != 0;
  const RT* X = w.x + (size_t)m * Ko * Tg::NX;
  const RT* Ut = w.u + (size_t)m * Ko * NU;
  const RT* Ot = w.o + (size_t)m * Ko * NO;
  const int optw = s.optran ? NO : 0;

  __shared__ float cs[NBUF][Tg::NX * WAVE];       // coefficients [slot][lane] of the staged layer(s)
  __shared__ RT tab[PW][NBUF][TABW];              // per-wave uniform rows
  constexpr int OCAP = OC_LDS_BYTES / (WAVE * (int)sizeof(RT));
  __shared__ RT ocs[OCAP * WAVE];                 // OPTRAN coefficients [row*order][lane]

  // ---- coefficient staging: layer k's [NX][64] floats, float4 per thread ----
  constexpr int CQ = (Tg::NX * WAVE / 4 + TPB - 1) / TPB;
  const float4* cg = reinterpret_cast<const float4*>(s.cblk + s.blk_base[b]);
  constexpr int n4 = Tg::NX * (WAVE / 4);
  float4 creg[CQ];
  auto cload = [&](int k) {
#pragma unroll
    for (int q = 0; q < CQ; ++q) {
      const int e = tid + q * TPB;
      if (e < n4) creg[q] = cg[(int64_t)k * n4 + e];
    }
  };
  auto cstore = [&](int buf) {
#pragma unroll
    for (int q = 0; q < CQ; ++q) {
      const int e = tid + q * TPB;
      if (e < n4) reinterpret_cast<float4*>(cs[buf])[e] = creg[q];
    }
  };
  // ---- per-wave row staging: element t of layer k is [X or U row | O row] ----
  RT rreg[RQ];
  auto rload = [&](const RT* base, int nbase, int k) {
#pragma unroll
    for (int q = 0; q < RQ; ++q) {
      const int t = lane + q * WAVE;
      RT v = RT(0);
      if (ok_prof) {
        if (t < nbase) v = base[(size_t)k * nbase + t];
        else if (t < nbase + optw) v = Ot[(size_t)k * NO + (t - nbase)];
      }
      rreg[q] = v;
    }
  };
  auto rstore = [&](int buf) {
#pragma unroll
    for (int q = 0; q < RQ; ++q) {
      const int t = lane + q * WAVE;
      if (t < TABW) tab[wv][buf][t] = rreg[q];
    }
  };
  // ---- OPTRAN coefficients: LDS copy when the block's table fits ----
  const int onq = blk_opt ? meta[BM_ONROW] * (meta[BM_OMAXO] + 1) : 0;
  const bool oc_lds = onq <= OCAP;
  const double* ocg = blk_opt ? s.ocd + ob : nullptr;
  if (blk_opt && oc_lds)
    for (int e = tid; e < onq * WAVE; e += TPB) ocs[e] = (RT)ocg[e];
  // two accessors, chosen once per block, so the inner loops carry no select
  auto oc_l = [&](int i) -> RT { return ocs[i * WAVE + lane]; };
  auto oc_g = [&](int i) -> RT { return (RT)ocg[i * WAVE + lane]; };

  // ------------- forward (1): ODPS layer OD and level-to-space path -------------
  RT path[KO_MAX + 1];
  uint64_t odclamp[2] = {0, 0}, lcclamp[2] = {0, 0};
  path[0] = RT(0);
  cload(0);
  rload(X, Tg::NX, 0);
  cstore(0);
  rstore(0);
  __syncthreads();
  for (int k = 0; k < Ko; ++k) {
    if (k + 1 < Ko) {
      cload(k + 1);
      rload(X, Tg::NX, k + 1);
    }
    const RT* xk = tab[wv][k % NBUF];
    const float* ck = cs[k % NBUF] + lane;
    RT od = comp_layer_od<G, RT, 0>(ck, npb, xk);
    if (blk_opt) {
      const RT* okr = xk + Tg::NX;
      const RT lnchi = oc_lds ? optran_lnchi<RT>(oc_l, meta, okr) : optran_lnchi<RT>(oc_g, meta, okr);
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
        od += c * okr[O_DA];
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
    if constexpr (NBUF == 1) __syncthreads();   // every wave is done with the buffer
    if (k + 1 < Ko) {
      cstore((k + 1) % NBUF);
      rstore((k + 1) % NBUF);
    }
    __syncthreads();
  }

  // ------------- forward (2): user layers, Planck, Emission (top-down) -------------
  const double* geo = w.geo + (size_t)m * 4;
  const double sec_sfc = geo[0], cosu = geo[1], tsfc = geo[2];
  const RT isec = (RT)(1.0 / sec_sfc);
  const RT iu = (RT)(1.0 / cosu);
  const int* ui = w.uidx + (size_t)m * (bt.max_k + 1);
  const RT* uf = w.ufrac + (size_t)m * (bt.max_k + 1);
  auto upath = [&](int i) -> RT {
    const int k1 = ui[i];
    const RT f = uf[i];
    return f < RT(0) ? path[k1] : path[k1] + (path[k1 + 1] - path[k1]) * f;
  };
  const RT c1 = (RT)s.c1[lc], c2 = (RT)s.c2[lc], bc1 = (RT)s.bc1[lc], bc2 = (RT)s.bc2[lc];
  const RT ic1 = (RT)(1.0 / s.c1[lc]);
  const double* tuser = bt.t + loff;
  RT rd = (RT)s.cosmic[lc];
  RT uacc = RT(0), tc = RT(1);
  const RT up0 = ok_prof ? upath(0) : RT(0);
  RT upp = up0;
  // all-sky calls keep the gas OD of cloudy profiles for the scattering RT
  RT* tau_out = (o.tau_g && valid && o.tau_mask[m] > 0) ? reinterpret_cast<RT*>(o.tau_g) + loff * s.L + l : nullptr;
  for (int k = 0; k < K; ++k) {
    const RT up = upath(k + 1);
    const RT tau = (up - upp) * isec;
    if (tau_out) tau_out[(int64_t)k * s.L] = tau;
    upp = up;
    const RT t = rexp<RT>(-tau * iu);
    RT te;
    const RT B = planck(c1, c2, bc1, bc2, tuser[k], te);
    rd = rd * t + B * (RT(1) - t);
    uacc += B * (RT(1) - t) * tc;
    tc *= t;
  }
  RT tes;
  const RT Bs = planck(c1, c2, bc1, bc2, tsfc, tes);
  const double eps_d = bt.emissivity[(size_t)m * bt.Ltot + lg];
  const RT eps = (RT)eps_d, rho = RT(1) - eps;
  const RT sfc_src = eps * Bs + rho * rd;
  const double R = (double)(uacc + tc * sfc_src);
  // CRTM_Planck_Temperature in FP64 (CRTM_Planck_Functions.f90:470-492)
  const double c1d = s.c1[lc], c2d = s.c2[lc], bc1d = s.bc1[lc], bc2d = s.bc2[lc];
  const double bt_v = (c2d / log(c1d / R + 1.0) - bc1d) / bc2d;
  // profiles with cloudy layers in an all-sky K-matrix call: the scattering
  // path has written the outputs and hands over adjoint seeds
  const bool seeded = DOK && o.ks.mask && ok_prof && o.ks.mask[m] > 0;   // wave-uniform
  const int64_t sl = (int64_t)m * s.L + lc;
  if (valid && !seeded) {
    o.radiance[ol] = R;
    o.bt[ol] = bt_v;
    if (o.up_radiance) o.up_radiance[ol] = (double)uacc;
    if (o.down_radiance) o.down_radiance[ol] = (double)rd;
    if (o.sfc_planck) o.sfc_planck[ol] = (double)Bs;
    if (o.sfc_emis) o.sfc_emis[ol] = eps_d;
  }
  if constexpr (!DOK) return;

  // ------------- adjoint seeds: Pre_Process_RTSolution_K and the surface terms -------------
  const double arg = c1d / R + 1.0;
  const double lga = log(arg);
  const double F = c1d * c2d / (bc2d * arg * (R * lga) * (R * lga));
  double rad_ad_d = 0.0;
  if (seeded) rad_ad_d = valid ? o.ks.rad[sl] : 0.0;
  else rad_ad_d = valid ? (o.rad_k ? o.rad_k[ol] : 0.0) + F * (o.bt_k ? o.bt_k[ol] : 0.0) : 0.0;
  if (valid && o.rad_k && !seeded) o.rad_k[ol] = rad_ad_d;
  const RT rad_ad = (RT)rad_ad_d;
  const RT upad_sfc = rad_ad * tc;           // adjoint of the surface-leaving radiance
  if (valid) {
    const double e_seed = seeded ? o.ks.emis[sl] : 0.0, bs_seed = seeded ? o.ks.bs[sl] : 0.0;
    if (o.emis_k) o.emis_k[ol] = (double)(upad_sfc * Bs - upad_sfc * rd) + e_seed;
    const RT dbsdt = c2 * bc2 * Bs * (Bs + c1) * ic1 / (tes * tes);
    const double sad = (double)(dbsdt * (eps * upad_sfc + (RT)bs_seed));
    for (int c = 0; c < 4; ++c)
      if (o.sfc_t_k[c]) o.sfc_t_k[c][ol] = bt.cov[c][m] * sad;
  }

  // ------------- adjoint sweep, bottom-up: user levels interleaved with ODPS layers -------------
  // Emission_AD in carried form: ru = upwelling radiance at the bottom of the
  // current layer, dn = adjoint of the downwelling radiance there, gd = rd*dn
  // (so the downwelling profile never has to be stored or inverted).
  RT ru = sfc_src;
  RT dn = rho * upad_sfc;
  RT gd = rd * dn;
  // Path adjoint per ODPS level, written once per level: user levels are
  // stepped bottom-up with non-increasing brackets ui, so their additions
  // go to a two-level window (lo = level cur, hi = level cur+1) that is
  // flushed as it slides up.
  RT pad[KO_MAX + 2];
  RT pa_hi = RT(0), pa_lo = RT(0);
  int pcur = Ko;
  const int64_t Lt = bt.Ltot;
  const int na = bt.na;
  // Jacobian rows of the caller's layers; the nx layers AddLayers put on
  // top carry no output (CRTM_Atmosphere_AddLayers_AD keeps only the
  // original layers, src/Atmosphere/CRTM_Atmosphere.f90:569-659)
  const int nx = ok_prof ? bt.n_extra[m] : 0;
  KT* tko = reinterpret_cast<KT*>(o.t_k) + (bt.out_loff[m] - nx) * Lt + lg;
  KT* ako = reinterpret_cast<KT*>(o.a_k) + (bt.out_loff[m] - nx) * Lt + lg;
  const int64_t astr = bt.out_NT * Lt;
  // Map_Input_AD in the sweep: user layer i is complete once the lowest ODPS
  // layer whose support holds it has been processed; with at most MW such
  // layers per user layer, their adjoints sit in a register window.  Wider
  // supports (very fine user grids) or more absorbers take the gather path.
  const bool winmap = ok_prof && geo[3] <= (double)MW && na <= NAW;
  // The Emission_AD pass stores each user layer's Planck part of T_K in the
  // output rows (and, for the gather path, zero absorber rows to add onto);
  // Map_Input_AD then completes them, writing the absorber rows of the
  // register-window path once.
  RT carry = RT(0);
  RT up_next = ok_prof ? upath(K) : RT(0);   // path at user level iu_ of the next step
  auto ustep = [&](int iu_) {
    RT v;
    if (iu_ >= 1) {
      const int k = iu_ - 1;   // user layer above level iu_
      const RT upk1 = up_next;
      const RT upk = upath(k);
      up_next = upk;
      const RT t = rexp<RT>(-((upk1 - upk) * isec) * iu);
      const RT tabove = rexp<RT>(-((upk - up0) * isec) * iu);
      RT te;
      const RT B = planck(c1, c2, bc1, bc2, tuser[k], te);
      const RT upad = rad_ad * tabove;
      const RT a_up = t * (ru - B) * upad;
      const RT a_dn = gd - B * dn;
      RT b_ad = (upad + dn) * (RT(1) - t);
      RT tau_ad = -(a_up + a_dn) * iu;
      RT t_seed = RT(0);
      if (seeded) {
        const int64_t is = (loff + k) * s.L + lc;
        b_ad += reinterpret_cast<const RT*>(o.ks.b)[is];
        tau_ad += reinterpret_cast<const RT*>(o.ks.tau)[is];
        t_seed = reinterpret_cast<const RT*>(o.ks.t)[is];
      }
      gd -= B * (RT(1) - t) * dn;
      dn *= t;
      ru = ru * t + B * (RT(1) - t);
      const RT dbdt = c2 * bc2 * B * (B + c1) * ic1 / (te * te);
      if (valid && k >= nx) {
        tko[(int64_t)k * Lt] = (KT)(dbdt * b_ad + t_seed);
        if (!winmap)
          for (int ja = 0; ja < na; ++ja) ako[ja * astr + (int64_t)k * Lt] = KT(0);
      }
      const RT gq = tau_ad * isec;
      v = carry + gq;
      carry = -gq;
    } else {
      v = carry;
    }
    const int k1 = ui[iu_];
    const RT f = uf[iu_];
    while (pcur > k1) {
      pad[pcur + 1] = pa_hi;
      pa_hi = pa_lo;
      pa_lo = RT(0);
      --pcur;
    }
    pa_lo += f < RT(0) ? v : (RT(1) - f) * v;   // level k1
    pa_hi += f < RT(0) ? RT(0) : f * v;          // level k1+1
  };
  // CRTM_Emission_AD over all user levels first (no barrier inside, so the
  // waves' serial chains overlap), then the ODPS sweep reads pad[].
  if (ok_prof) {
    for (int i = K; i >= 0; --i) ustep(i);
    while (pcur >= 0) {
      pad[pcur + 1] = pa_hi;
      pa_hi = pa_lo;
      pa_lo = RT(0);
      --pcur;
    }
  }
  RT pad_next = ok_prof ? pad[Ko] : RT(0);
  const int* acc_i = w.acc_i + (size_t)m * Ko * 3;
  const RT* acc_w = w.acc_w + (size_t)m * NZ_MAX;
  RT tw[MW], aw[MW][NAW];
#pragma unroll
  for (int q = 0; q < MW; ++q) {
    tw[q] = RT(0);
#pragma unroll
    for (int ju = 0; ju < NAW; ++ju) aw[q][ju] = RT(0);
  }
  // The Planck part of the rows finalized at ODPS layer k is loaded while
  // layer k+1 is processed (pfc / pfn: current and next layer's rows from
  // pf_lo / pf_lo_n), so the read does not stall the synchronized waves.
  RT pfc[PF], pfn[PF];
  int pf_lo = 0, pf_lo_n = 0;
  const int* frange = w.frange + (size_t)m * Ko * 2;
  const RT* uwin = w.uwin + (size_t)m * bt.max_k * MW;
  auto final_lo = [&](int k) { return frange[2 * k]; };
  auto prefetch = [&](int k) {   // rows finalized at ODPS layer k, into pfn
    pf_lo_n = final_lo(k);
#pragma unroll
    for (int q = 0; q < PF; ++q) {
      const int iu = pf_lo_n + q;
      pfn[q] = (valid && iu >= nx && iu < K) ? (RT)tko[(int64_t)iu * Lt] : RT(0);
    }
  };
  auto finish_user_layer = [&](int iu) {   // completed at the ODPS layer in window slot 0
    if (!valid || iu < nx) return;
    RT st;                                         // Planck part from the Emission_AD pass
    if ((unsigned)(iu - pf_lo) < (unsigned)PF) {
      st = pfc[0];
#pragma unroll
      for (int q = 1; q < PF; ++q)
        if (iu - pf_lo == q) st = pfc[q];
    } else {
      st = (RT)tko[(int64_t)iu * Lt];
    }
    const RT* wr = uwin + (size_t)iu * MW;         // weights from ODPS layers in slots 0..MW-1
    RT sa[NAW];
#pragma unroll
    for (int ju = 0; ju < NAW; ++ju) sa[ju] = RT(0);
#pragma unroll
    for (int q = 0; q < MW; ++q) {
      const RT wk = wr[q];
      st += wk * tw[q];
#pragma unroll
      for (int ju = 0; ju < NAW; ++ju) sa[ju] += wk * aw[q][ju];
    }
    tko[(int64_t)iu * Lt] = (KT)st;
#pragma unroll
    for (int ju = 0; ju < NAW; ++ju)
      if (ju < na) ako[ju * astr + (int64_t)iu * Lt] = (KT)sa[ju];
  };
  if (winmap) {   // user layers below every ODPS support take only their Planck part
    const int ilast = acc_i[3 * (Ko - 1)] + acc_i[3 * (Ko - 1) + 1] - 1;
    pf_lo = K;
    for (int iu = ilast + 1; iu < K; ++iu) finish_user_layer(iu);   // window empty: Planck part only
    prefetch(Ko - 1);
  }

  RT S = RT(0), tzs = RT(0), tzps = RT(0);
  RT gzs[Tg::NABS], gts[Tg::NABS];
#pragma unroll
  for (int j = 0; j < Tg::NABS; ++j) gzs[j] = gts[j] = RT(0);
  RT ivn = RT(0), sta = RT(0), spa = RT(0), xltc = RT(0), xlpc = RT(0);
  const int h2o_j = s.h2o_idx - 1;
  // the forward loop ended on a barrier, so both LDS buffers are free
  cload(Ko - 1);
  rload(Ut, NU, Ko - 1);
  cstore((Ko - 1) % NBUF);
  rstore((Ko - 1) % NBUF);
  __syncthreads();
  for (int k = Ko - 1; k >= 0; --k) {
    if (k > 0) {
      cload(k - 1);
      rload(Ut, NU, k - 1);
    }
    if (winmap) {
#pragma unroll
      for (int q = 0; q < PF; ++q) pfc[q] = pfn[q];
      pf_lo = pf_lo_n;
      if (k > 0) prefetch(k - 1);
    }
    if (ok_prof) {
      S += pad_next;
      pad_next = k > 0 ? pad[k] : RT(0);
      const RT od_ad = ((odclamp[k >> 6] >> (k & 63)) & 1) ? RT(0) : S;
      const RT* uk = tab[wv][k % NBUF];
      LayerAdj<RT> a{};
      const float* ck = cs[k % NBUF] + lane;
      comp_layer_ad<G, RT, 0>(ck, npb, uk, a);
      RT t_ad = RT(0), vap_ad = RT(0);
      if (blk_opt) {
        // Add_OPTRAN_wloOD_AD (ODPS_AtmAbsorption.f90:911-984) with chi recomputed
        const RT* okr = uk + NU;
        const RT lnchi = oc_lds ? optran_lnchi<RT>(oc_l, meta, okr) : optran_lnchi<RT>(oc_g, meta, okr);
        const bool lcc = (lcclamp[k >> 6] >> (k & 63)) & 1;
        const RT chi = lane_opt ? (lcc ? (lnchi > RT(0) ? RT(LIMIT_LOG) : RT(0)) : rexp<RT>(lnchi)) : RT(0);
        const RT da_ad = lane_opt ? od_ad * chi : RT(0);
        const RT lc_ad = (lane_opt && !lcc) ? chi * od_ad * okr[O_DA] : RT(0);
        const int nrow = meta[BM_ONROW], nq = meta[BM_OMAXO] + 1;
        RT ap1_ad = RT(0), e12 = RT(0), e13 = RT(0);
        auto rows_ad = [&](const auto& oc) {
          for (int r = 0; r < nrow; ++r) {
            const int p = meta[BM_OROW0 + r];
            RT bp = RT(0), dbp = RT(0);
            for (int q = 0; q < nq; ++q) {
              const RT cq = oc(r * nq + q);
              bp += cq * okr[O_AP0 + q];
              if (q > 0) dbp += RT(q) * cq * okr[O_AP0 + q - 1];
            }
            const RT xa = lc_ad * bp;          // adjoint of OX(p)
            ap1_ad += lc_ad * okr[O_X0 + p] * dbp;
            t_ad += xa * okr[O_DX_DT0 + p];
            vap_ad += xa * okr[O_DX_DV0 + p];
            if (p == 12) e12 += xa;
            if (p == 13) e13 += xa;
          }
        };
        if (oc_lds) rows_ad(oc_l);
        else rows_ad(oc_g);
        // ODPS_Compute_Predictor_ODAS_AD at layer k (ODPS_Predictor.f90:3092-3231)
        const RT ave_ad = ap1_ad * okr[O_IDEN];
        const RT xlt_ad = xltc + e12, xlp_ad = xlpc + e13;
        xltc = e12;
        xlpc = e13;
        sta += RT(0.5) * okr[O_INV] * xlt_ad;
        spa += RT(0.5) * okr[O_INV] * xlp_ad;
        const RT inv_ad = RT(0.5) * (okr[O_ST] * xlt_ad + okr[O_SP] * xlp_ad);
        const RT iv_ad = ivn - inv_ad * okr[O_IIV2] + RT(0.5) * ave_ad;
        ivn = RT(0.5) * ave_ad + iv_ad;
        const RT dab_ad = da_ad + okr[O_PR] * spa + okr[O_T] * sta + iv_ad;
        t_ad += okr[O_DA] * sta;
        vap_ad += okr[O_DPSEC] * dab_ad;
      }
      finish_layer<G, RT>(uk, a);
      t_ad += (a.T * uk[U_IRT] + a.dt) * od_ad;
      RT a_ad[Tg::NABS];
#pragma unroll
      for (int j = 0; j < Tg::NABS; ++j) a_ad[j] = a.ra[j] * od_ad * uk[U_IRA0 + j];
      if (s.optran) a_ad[h2o_j] += vap_ad;
      const RT pdp = uk[U_PDP];
#pragma unroll
      for (int j = 0; j < Tg::NABS; ++j) {
        gts[j] += a.gt[j] * od_ad * uk[U_IGATZP0 + j];
        gzs[j] += a.gz[j] * od_ad * uk[U_IGAZP0 + j];
        t_ad += gts[j] * pdp * uk[U_AODPS0 + j];
        a_ad[j] += gzs[j] * pdp + gts[j] * pdp * uk[U_TODPS];
      }
      tzps += a.tzp * od_ad * uk[U_ITZPREF];
      tzs += a.tz * od_ad * uk[U_ITZREF];
      t_ad += tzs + pdp * tzps;
      // Map_Input_AD (ODPS_CoordinateMapping.f90:480-547): add onto the user
      // layers averaged into this ODPS layer (their Planck part is already stored)
      if (winmap) {
#pragma unroll
        for (int q = MW - 1; q > 0; --q) {
          tw[q] = tw[q - 1];
#pragma unroll
          for (int ju = 0; ju < NAW; ++ju) aw[q][ju] = aw[q - 1][ju];
        }
        tw[0] = t_ad;
#pragma unroll
        for (int ju = 0; ju < NAW; ++ju) aw[0][ju] = RT(0);
#pragma unroll
        for (int j = 0; j < Tg::NABS; ++j) {
          const int ju = bt.jmap[j];
#pragma unroll
          for (int q = 0; q < NAW; ++q)
            if (ju == q && uk[U_MASK0 + j] != RT(0)) aw[0][q] = a_ad[j];
        }
        // user layers (i2(k-1), i2(k)] are complete
        const int ihi = frange[2 * k + 1];
        for (int iu = pf_lo; iu <= ihi; ++iu) finish_user_layer(iu);
      } else if (valid) {
        // wide supports or more absorbers: add onto the output rows.  Each
        // lane owns its channel's column, so the adds need no atomicity; the
        // returnless atomic add retires without a load round trip (same-lane,
        // same-address operations stay in program order).
        const int i1 = acc_i[3 * k], n = acc_i[3 * k + 1];
        const RT* awk = acc_w + acc_i[3 * k + 2];
        for (int q = 0; q < n; ++q)
          if (i1 + q >= nx) unsafeAtomicAdd(tko + (int64_t)(i1 + q) * Lt, (KT)(awk[q] * t_ad));
#pragma unroll
        for (int j = 0; j < Tg::NABS; ++j) {
          const int ju = bt.jmap[j];
          if (ju >= 0 && uk[U_MASK0 + j] != RT(0))
            for (int q = 0; q < n; ++q)
              if (i1 + q >= nx) unsafeAtomicAdd(ako + ju * astr + (int64_t)(i1 + q) * Lt, (KT)(awk[q] * a_ad[j]));
        }
      }
    }
    if constexpr (NBUF == 1) __syncthreads();
    if (k > 0) {
      cstore((k - 1) % NBUF);
      rstore((k - 1) % NBUF);
    }
    __syncthreads();
  }
}

template <int G, class RT, class KT, bool DOK>
void launch_fwdk(const DevSensor& s, const DevBatch& b, const ProfWork<RT>& w, const DevOut& o,
                 hipStream_t stream) {
  constexpr int PW = KCfg<RT>::PW;
  const int64_t ng = (b.M + PW - 1) / PW;
  const int64_t n = 8 * ((s.n_blocks * ng + 7) / 8);
  hipLaunchKernelGGL((clear_fwdk_kernel<G, RT, KT, DOK>), dim3((unsigned)n), dim3(PW * WAVE), 0, stream, s, b, w,
                     o);
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
