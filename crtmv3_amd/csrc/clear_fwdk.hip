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
//             -> Planck, Emission RT (Emission_Module.f90:54-174)
//             -> BT (FP64, CRTM_Planck_Functions.f90:470-492)
//   adjoint:  Pre_Process_RTSolution_K (CRTM_K_Matrix_Module.f90:1725-1761),
//             CRTM_Emission_AD (Emission_Module.f90:280-395), Planck AD,
//             ODPS_Compute_AtmAbsorption_AD (ODPS_AtmAbsorption.f90:487-600),
//             Add_OPTRAN_wloOD_AD (:911-984), ODPS_Compute_Predictor_AD
//             (ODPS_Predictor.f90:1783-2643), ODPS_Compute_Predictor_ODAS_AD
//             (:3092-3231), Map_Input_AD (ODPS_CoordinateMapping.f90:480-547)
// all in one pass over the profile's layers, with nothing per-layer stored
// in HBM.  RT is the arithmetic type of absorption, Planck and RT (float for
// the split path, double for the plain FP64 path); the radiance is
// accumulated into FP64 and the BT inversion is always FP64.
#include "crtm_device.h"
#include "pred_adjoint.h"

namespace crtm {

namespace {
constexpr double MAX_OD = 20.0;
constexpr double LIMIT_EXP = 20.0, LIMIT_LOG = 4.8e+08;

template <class RT, class KT>
__device__ inline void store_k(void* base, int64_t idx, RT v) {
  reinterpret_cast<KT*>(base)[idx] = (KT)v;
}
}  // namespace

// The layer's predictor adjoint is linear in X_AD(k, :, :) = C(:, :, k) * OD_AD(k)
// with per-layer factors, so the final quantities are y[q] = sum_s C(s) D(s, q)
// with D the profile's sensitivity table (prepared per layer) and C the
// channel's coefficients; the caller scales y by OD_AD(k) once.  Only the
// structurally nonzero (s, q) pairs are stored and visited, as compile-time
// indices, so y lives in registers.
template <int G, class RT, int S, int E, int Q>
__device__ inline void sens_q(RT c, const RT* dr, RT* y) {
  if constexpr (Q < QIdx<G>::NQ) {
    constexpr bool on = (SensPattern<G>::tab.mask[S] >> Q) & 1u;
    if constexpr (on) y[Q] += c * dr[E];
    sens_q<G, RT, S, E + (on ? 1 : 0), Q + 1>(c, dr, y);
  }
}
// Slots are taken in chunks of SCH: the chunk's coefficients are read from
// LDS first and the scheduling barrier keeps the reads together, so one
// wait covers several reads instead of one read pair per wait.
constexpr int SCH = 4;
template <int G, class RT, int S, int I, int N>
__device__ inline void sens_load(const float* ck, RT* c) {
  if constexpr (I < N) {
    if constexpr (SensPattern<G>::tab.mask[S + I] != 0u) c[I] = (RT)ck[(S + I) * WAVE];
    sens_load<G, RT, S, I + 1, N>(ck, c);
  }
}
template <int G, class RT, int S, int I, int N>
__device__ inline void sens_use(const RT* c, const RT* dr, RT* y) {
  if constexpr (I < N) {
    if constexpr (SensPattern<G>::tab.mask[S + I] != 0u)
      sens_q<G, RT, S + I, SensPattern<G>::tab.off[S + I], 0>(c[I], dr, y);
    sens_use<G, RT, S, I + 1, N>(c, dr, y);
  }
}
template <int G, class RT, int S, int SEND>
__device__ inline void sens_slots(const float* ck, const RT* dr, RT* y) {
  if constexpr (S < SEND) {
    constexpr int N = SEND - S < SCH ? SEND - S : SCH;
    constexpr int E0 = SensPattern<G>::tab.off[S], NE = SensPattern<G>::tab.off[S + N] - E0;
    RT c[SCH];
    RT dv[NE > 0 ? NE : 1];
    sens_load<G, RT, S, 0, N>(ck, c);
#pragma unroll
    for (int e = 0; e < NE; ++e) dv[e] = dr[E0 + e];
    __builtin_amdgcn_sched_barrier(0);
    sens_use<G, RT, S, 0, N>(c, dv - E0, y);
    sens_slots<G, RT, S + N, SEND>(ck, dr, y);
  }
}
template <int G, class RT, int J>
__device__ inline void sens_layer(const float* ck, const int* npb, const RT* dr, RT* y) {
  using Tg = GT<G>;
  if constexpr (J < Tg::NCOMP) {
    if (npb[J] > 0)   // block-uniform: some channel of the block uses component J
      sens_slots<G, RT, Tg::XOFF[J], Tg::XOFF[J] + Tg::NP[J]>(ck, dr, y);
    sens_layer<G, RT, J + 1>(ck, npb, dr, y);
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

// Workgroup shape: one 64-channel block x PW profiles (one wave each).  PW is
// a tile-shape choice per sensor and precision (crtm_set_tile /
// crtm_autotune pick it at run time from TileSet<RT>); the default,
// measured at the CrIS-like benchmark shape, runs two workgroups per CU so
// that one workgroup's layer barrier overlaps the other's arithmetic.
//   WPE  waves per SIMD the register allocation targets (0: compiler's choice):
//        FP32 keeps <= 128 VGPRs (4 waves per SIMD); FP64 needs more registers.
template <class RT> struct KCfg { static constexpr int WPE = 4; };
template <> struct KCfg<double> { static constexpr int WPE = 0; };
constexpr int NBUF = 2;
// LDS row of one wave: predictors X or the adjoint row, followed by the OPTRAN table O
constexpr int TABW = (NX_MAX > NR_MAX ? NX_MAX : NR_MAX) + NO;
constexpr int OC_LDS_BYTES = 16384;            // OPTRAN coefficients staged when they fit
constexpr int MW = MAP_WIN;                    // Map_Input_AD register window (ODPS layers)
constexpr int NAW = 2;                         // user absorbers the window carries
constexpr int PF = 4;                          // user-layer rows prefetched one ODPS layer ahead

// OPTRAN ln(chi) of one lane at one layer (Add_OPTRAN_wloOD,
// ODPS_AtmAbsorption.f90:664-727): b_p = sum_q coef(p,q) Ap(q), lnchi = sum_p b_p OX(p)
template <class RT, class OC, class MP>
__device__ inline RT optran_lnchi(const OC& oc, MP meta, const RT* ok) {
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

// RED (adjoint): the Jacobian outputs are channel sums, CRTM_Adjoint's
// Atmosphere_AD / Surface_AD seeded by RTSolution_AD (src/CRTM_Adjoint_Module.f90:
// 1378-1396): each wave sums its 64 channels and adds the sum to the profile's
// row (FP64 atomics across channel blocks), so no per-channel Jacobian is
// stored; KT is double.
template <int G, class RT, class KT, bool DOK, bool RED, int PW>
__global__ void __attribute__((amdgpu_flat_work_group_size(PW * WAVE, PW * WAVE),
                               amdgpu_waves_per_eu(KCfg<RT>::WPE > 0 ? KCfg<RT>::WPE : 1)))
clear_fwdk_kernel(DevSensor s, DevBatch bt, ProfWork<RT> w, DevOut o) {
  static_assert(!RED || (DOK && sizeof(KT) == 8), "the adjoint accumulates FP64 channel sums");
  using Tg = GT<G>;
  constexpr int TPB = PW * WAVE;
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
  // the wave index is wave-uniform: from an SGPR, every per-profile address
  // below is uniform and its loads become scalar loads
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
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
    // gradient built from them passes silently.  A skipped profile
    // (Options(m)%Skip_Profile): the zeroed RTSolution and untouched (zero)
    // Jacobians of the reference (src/CRTM_K_Matrix_Module.f90:594, 631)
    const bool skipped = w.status[m] == ST_SKIP;
    const double fill = skipped ? 0.0 : __builtin_nan("");
    o.radiance[ol] = fill;
    o.bt[ol] = fill;
    if (skipped) {
      if (o.up_radiance) o.up_radiance[ol] = 0.0;
      if (o.down_radiance) o.down_radiance[ol] = 0.0;
      if (o.sfc_planck) o.sfc_planck[ol] = 0.0;
      if (o.sfc_emis) o.sfc_emis[ol] = 0.0;
      if (o.sfc_refl) o.sfc_refl[ol] = 0.0;
    }
    if constexpr (RED) {   // a skipped profile adds nothing; a failed one's rows become NaN
      if (!skipped && lane == 0 && b == 0)
        for (int64_t r = bt.out_loff[m]; r < bt.out_loff[m + 1]; ++r) {
          reinterpret_cast<double*>(o.t_k)[r] = fill;
          for (int ja = 0; ja < bt.na; ++ja) reinterpret_cast<double*>(o.a_k)[ja * bt.out_NT + r] = fill;
        }
      if (skipped && o.emis_k) o.emis_k[ol] = 0.0;
    } else if constexpr (DOK) {
      const int64_t r0 = bt.out_loff[m], r1 = bt.out_loff[m + 1];
      const int64_t astr_f = bt.out_NT * bt.Ltot;
      for (int64_t r = r0; r < r1; ++r) {
        reinterpret_cast<KT*>(o.t_k)[r * bt.Ltot + lg] = KT(fill);
        for (int ja = 0; ja < bt.na; ++ja)
          reinterpret_cast<KT*>(o.a_k)[ja * astr_f + r * bt.Ltot + lg] = KT(fill);
      }
      if (skipped) {
        if (o.emis_k) o.emis_k[ol] = 0.0;
        for (int c = 0; c < 4; ++c)
          if (o.sfc_t_k[c]) o.sfc_t_k[c][ol] = 0.0;
        for (int j = 0; j < 3; ++j)
          if (o.sfc_k[j]) o.sfc_k[j][ol] = 0.0;
      }
    }
  }
  const int K = ok_prof ? bt.n_layers[m] : 0;
  const int64_t loff = bt.loff[m];
  const int Ko = s.Ko;
  const cst_ptr<int> meta = as_const(s.blk_meta + b * BM_N);
  int npb[Tg::NCOMP];
#pragma unroll
  for (int j = 0; j < Tg::NCOMP; ++j) npb[j] = meta[BM_NP0 + j];
  const int64_t ob = s.optran ? s.oblk_base[b] : -1;
  const bool blk_opt = ob >= 0;
  const bool lane_opt = blk_opt && s.lane_optran[b * WAVE + lane] != 0;
  const RT* X = w.x + (size_t)m * Ko * Tg::NX;
  using RR = RRow<G>;
  using Qi = QIdx<G>;
  const RT* Ut = w.u + (size_t)m * Ko * RR::NR;
  const RT* Ot = w.o + (size_t)m * Ko * NO;
  const int optw = s.optran ? NO : 0;

  __shared__ float cs[NBUF][Tg::NX * WAVE];       // coefficients [slot][lane] of the staged layer(s)
  __shared__ RT tab[PW][NBUF][TABW];              // per-wave uniform rows
  constexpr int OCAP = OC_LDS_BYTES / (WAVE * (int)sizeof(RT));
  __shared__ RT ocs[OCAP * WAVE];                 // OPTRAN coefficients [row*order][lane]

  // ---- staging of layer k straight into LDS (global_load_lds): lane i of a
  // wave's load lands at base + size*i, so neither the block's coefficient
  // tile nor the wave's rows pass through VGPRs.  Each wave waits for its own
  // loads before the barrier that publishes the buffer (stage_wait). ----
  using lds_ptr = __attribute__((address_space(3))) void*;
  const float4* cg = reinterpret_cast<const float4*>(s.cblk + s.blk_base[b]);
  constexpr int n4 = Tg::NX * (WAVE / 4);          // float4 of one layer's coefficients
  constexpr int NCH = (n4 + WAVE - 1) / WAVE;      // 1 KB pieces, spread over the waves
  auto cload = [&](int k, int buf) {
    for (int c = wv; c < NCH; c += PW) {           // wave-uniform
      const int e = c * WAVE + lane;
      if (e < n4)
        __builtin_amdgcn_global_load_lds(cg + (int64_t)k * n4 + e,
                                         (lds_ptr)(reinterpret_cast<float4*>(cs[buf]) + c * WAVE), 16, 0, 0);
    }
  };
  // element t of the row staged for layer k: [X or adjoint row | O row], as dwords
  constexpr int DW = (int)(sizeof(RT) / 4);
  auto rload = [&](const RT* base, int nbase, int k, int buf) {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(base + (size_t)k * nbase);
    uint32_t* dst = reinterpret_cast<uint32_t*>(tab[wv][buf]);
    const int nb = nbase * DW;
    for (int c = 0; c * WAVE < nb; ++c)
      if (c * WAVE + lane < nb)
        __builtin_amdgcn_global_load_lds(src + c * WAVE + lane, (lds_ptr)(dst + c * WAVE), 4, 0, 0);
    if (optw) {
      const uint32_t* so = reinterpret_cast<const uint32_t*>(Ot + (size_t)k * NO);
      constexpr int no = NO * DW;
      for (int c = 0; c * WAVE < no; ++c)
        if (c * WAVE + lane < no)
          __builtin_amdgcn_global_load_lds(so + c * WAVE + lane, (lds_ptr)(dst + nb + c * WAVE), 4, 0, 0);
    }
  };
  // all of this wave's memory operations done (the staged loads land in LDS)
  auto stage_wait = [] { __builtin_amdgcn_s_waitcnt(0); };
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
  // Path at the user levels, interpolated (Interpolate_Profile) as soon as the
  // two ODPS levels that bracket a level are known; the ODPS-level path itself
  // is only carried from one layer to the next.
  RT upth[K_MAX + 1];
  RT pk = RT(0);
  const cst_ptr<int> ui = as_const(w.uidx + (size_t)m * (bt.max_k + 1));
  const cst_ptr<RT> uf = as_const(w.ufrac + (size_t)m * (bt.max_k + 1));
  const int nup = ok_prof ? bt.n_layers[m] + 1 : 0;
  int iup = 0;
  int ui_n = nup > 0 ? ui[0] : Ko + 1;
  // clamp bitmasks of layers 0..63 and 64..127, kept in registers (a
  // dynamically indexed array would live in scratch memory)
  uint64_t odc_lo = 0, odc_hi = 0, lcc_lo = 0, lcc_hi = 0;
  auto setbit = [](uint64_t& lo, uint64_t& hi, int k) {
    if (k < 64) lo |= 1ull << k;
    else hi |= 1ull << (k - 64);
  };
  auto getbit = [](uint64_t lo, uint64_t hi, int k) -> bool {
    return ((k < 64 ? lo : hi) >> (k & 63)) & 1;
  };
  cload(0, 0);
  rload(X, Tg::NX, 0, 0);
  stage_wait();
  __syncthreads();
  for (int k = 0; k < Ko; ++k) {
    if (k + 1 < Ko) {
      cload(k + 1, (k + 1) % NBUF);
      rload(X, Tg::NX, k + 1, (k + 1) % NBUF);
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
          setbit(lcc_lo, lcc_hi, k);
        } else if (lnchi < RT(-LIMIT_EXP)) {
          c = RT(0);
          setbit(lcc_lo, lcc_hi, k);
        } else {
          c = rexp<RT>(lnchi);
        }
        od += c * okr[O_DA];
      }
    }
    RT odc = od;
    if (od < RT(0)) {
      odc = RT(0);
      setbit(odc_lo, odc_hi, k);
    } else if (od > RT(MAX_OD)) {
      odc = RT(MAX_OD);
      setbit(odc_lo, odc_hi, k);
    }
    const RT pk1 = pk + odc;
    while (ui_n == k) {   // user levels bracketed by ODPS levels k and k+1 (brackets never decrease)
      const RT f = uf[iup];
      upth[iup] = f < RT(0) ? pk : pk + (pk1 - pk) * f;
      ++iup;
      ui_n = iup < nup ? ui[iup] : Ko + 1;
    }
    pk = pk1;
    stage_wait();
    __syncthreads();
  }

  for (; iup < nup; ++iup) upth[iup] = pk;   // below the lowest ODPS level
  // ------------- forward (2): user layers, Planck, Emission (top-down) -------------
  const cst_ptr<double> geo = as_const(w.geo + (size_t)m * 4);
  const double sec_sfc = geo[0], cosu = geo[1], tsfc = geo[2];
  const RT isec = (RT)(1.0 / sec_sfc);
  const RT iu = (RT)(1.0 / cosu);
  auto upath = [&](int i) -> RT { return upth[i]; };
  const RT c1 = (RT)s.c1[lc], c2 = (RT)s.c2[lc], bc1 = (RT)s.bc1[lc], bc2 = (RT)s.bc2[lc];
  const RT ic1 = (RT)(1.0 / s.c1[lc]);
  const cst_ptr<double> tuser = as_const(bt.t + loff);
  RT rd = (RT)s.cosmic[lc];
  RT uacc = RT(0), tc = RT(1);
  const RT up0 = ok_prof ? upath(0) : RT(0);
  RT upp = up0;
  // all-sky calls keep the gas OD of cloudy profiles for the scattering RT
  RT* tau_out = (o.tau_g && valid && o.tau_mask[m] > 0) ? reinterpret_cast<RT*>(o.tau_g) + loff * s.L + l : nullptr;
  double odsum = 0.0;   // column nadir optical depth (CRTM_Compute_Transmittance)
  for (int k = 0; k < K; ++k) {
    const RT up = upath(k + 1);
    const RT tau = (up - upp) * isec;
    odsum += (double)tau;
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
  const int64_t sl = (int64_t)m * s.L + lc;
  // surface emissivity and reflectivity: the user's (reflectivity 1 - e), or
  // the computed optics with the reflection correction at the column
  // transmittance (sfc_optics.h)
  const bool sfc_comp = o.sp.mode != SFC_USER;
  double eps_d, rho_d = 0.0, trans = 1.0;
  SfcLin zl;
  if (!sfc_comp) {
    eps_d = bt.emissivity[(size_t)m * bt.Ltot + lg];
  } else {
    trans = exp(-odsum);
    sfc_forward(o.sp, sl, trans, eps_d, rho_d, zl);
  }
  const RT eps = (RT)eps_d, rho = sfc_comp ? (RT)rho_d : RT(1) - eps;
  const RT sfc_src = eps * Bs + rho * rd;
  // Post_Process_RTSolution (src/CRTM_Forward_Module.f90:1166-1199): the
  // non-LTE correction of the radiance (Compute_NLTE_Correction,
  // src/NLTE/NLTE_Correction/CRTM_NLTECorrection.f90:503-547), the FP64
  // Planck temperature (CRTM_Planck_Functions.f90:470-492), the antenna
  // correction of the brightness temperature (CRTM_Compute_AntCorr,
  // src/AntennaCorrection/CRTM_AntennaCorrection.f90:118-140)
  double R = (double)(uacc + tc * sfc_src);
  const cst_ptr<double> nlr = (w.nlte && ok_prof) ? as_const(w.nlte + (size_t)m * NLTE_REC) : nullptr;
  const bool nl_on = nlr && nlr[NL_ACTIVE] != 0.0 && s.nl_cidx[lc] > 0;
  double nlc[3] = {0.0, 0.0, 0.0};   // the channel's coefficients at the profile's angles
  if (nl_on) {
    const int ich = s.nl_cidx[lc] - 1, is = (int)nlr[NL_ISEN], io = (int)nlr[NL_ISOL];
    const int ns = s.nl_nsen, no = s.nl_nsol;
    auto C = [&](int i, int a, int b) { return s.nl_c[i + 3 * (a + ns * (b + no * ich))]; };
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      nlc[i] = nlr[NL_WGT + 0] * C(i, is, io) + nlr[NL_WGT + 1] * C(i, is, io + 1) + nlr[NL_WGT + 2] * C(i, is + 1, io) +
               nlr[NL_WGT + 3] * C(i, is + 1, io + 1);
      R += nlc[i] * nlr[NL_PRED + i];
    }
  }
  const double c1d = s.c1[lc], c2d = s.c2[lc], bc1d = s.bc1[lc], bc2d = s.bc2[lc];
  double bt_v = (c2d / log(c1d / R + 1.0) - bc1d) / bc2d;
  double a_ep = 1.0;   // d(antenna temperature) / d(brightness temperature)
  if (bt.ac_on && s.ac_nfov > 0 && bt.ifov && ok_prof && bt.ifov[m] > 0) {
    const int64_t fa = (int64_t)(bt.ifov[m] - 1) * s.L + lc;
    a_ep = s.ac_ep[fa];
    bt_v = a_ep * bt_v + s.ac_sp[fa] * TSPACE;
  }
  // profiles with cloudy layers in an all-sky K-matrix call: the scattering
  // path has written the outputs and hands over adjoint seeds
  const bool seeded = DOK && o.ks.mask && ok_prof && o.ks.mask[m] > 0;   // wave-uniform
  if (valid && !seeded) {
    o.radiance[ol] = R;
    o.bt[ol] = bt_v;
    if (o.up_radiance) o.up_radiance[ol] = (double)uacc;
    if (o.down_radiance) o.down_radiance[ol] = (double)rd;
    if (o.sfc_planck) o.sfc_planck[ol] = (double)Bs;
    if (o.sfc_emis) o.sfc_emis[ol] = eps_d;
    if (o.sfc_refl) o.sfc_refl[ol] = sfc_comp ? rho_d : (double)rho;
  }
  if constexpr (!DOK) return;

  // ------------- adjoint seeds: Pre_Process_RTSolution_K and the surface terms -------------
  const double arg = c1d / R + 1.0;
  const double lga = log(arg);
  const double F = c1d * c2d / (bc2d * arg * (R * lga) * (R * lga));
  double rad_ad_d = 0.0;
  if (seeded) rad_ad_d = valid ? o.ks.rad[sl] : 0.0;
  else rad_ad_d = valid ? (o.rad_k ? o.rad_k[ol] : 0.0) + F * a_ep * (o.bt_k ? o.bt_k[ol] : 0.0) : 0.0;
  if (valid && o.rad_k && !seeded) o.rad_k[ol] = rad_ad_d;
  const RT rad_ad = (RT)rad_ad_d;
  // Compute_NLTE_Correction_AD and Compute_NLTE_Predictor_AD (CRTM_NLTECorrection.f90:396-446,
  // 704-744): each mean-temperature layer's predictor adjoint spread evenly over its
  // levels (k1, k2], added to those layers' temperature adjoint in the Emission_AD pass
  RT nl_g[2] = {RT(0), RT(0)};
  int nl_k1[2] = {0, 0}, nl_k2[2] = {0, 0};
  if (nl_on && nlr[NL_CTM] != 0.0) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      nl_k1[i] = (int)nlr[NL_K1 + i];
      nl_k2[i] = (int)nlr[NL_K2 + i];
      if (nl_k2[i] > nl_k1[i]) nl_g[i] = (RT)(nlc[1 + i] * rad_ad_d / (double)(nl_k2[i] - nl_k1[i]));
    }
  }
  auto nl_t = [&](int k) -> RT {   // the NLTE part of user layer k's temperature adjoint
    return (k >= nl_k1[0] && k < nl_k2[0] ? nl_g[0] : RT(0)) + (k >= nl_k1[1] && k < nl_k2[1] ? nl_g[1] : RT(0));
  };
  const RT upad_sfc = rad_ad * tc;           // adjoint of the surface-leaving radiance
  // computed optics: CRTM_Compute_SfcOptics_AD from the emissivity and
  // reflectivity adjoints (sg: water temperature, salinity, wind speed, wind
  // direction), and CRTM_Compute_Transmittance_AD: every layer's optical
  // depth adjoint takes -T * T_AD (tau_sfc_ad, added in the sweep below)
  double sg[4] = {0.0, 0.0, 0.0, 0.0};
  RT tau_sfc_ad = RT(0);
  if (sfc_comp && valid) {
    double tr_ad;
    sfc_adjoint(o.sp, sl, zl, (double)(upad_sfc * Bs), (double)(upad_sfc * rd), tr_ad, sg);
    tau_sfc_ad = (RT)(-trans * tr_ad);
  }
  if constexpr (RED) {   // surface temperature adjoints [M]: channel sums
    if (ok_prof) {
      const double bs_seed = (seeded && valid) ? o.ks.bs[sl] : 0.0;
      const RT dbsdt = c2 * bc2 * Bs * (Bs + c1) * ic1 * rrcp<RT>(tes * tes);
      const double sad = wave_sum(valid ? (double)(dbsdt * (eps * upad_sfc + (RT)bs_seed)) : 0.0);
      if (lane == 0)
        for (int c = 0; c < 4; ++c)
          if (o.sfc_t_k[c]) unsafeAtomicAdd(o.sfc_t_k[c] + m, bt.cov[c][m] * sad);
      if (sfc_comp) {   // wave-uniform
        const double s0 = wave_sum(sg[0]);
        double sj[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) sj[j] = wave_sum(sg[1 + j]);
        if (lane == 0) {
          if (o.sfc_t_k[1]) unsafeAtomicAdd(o.sfc_t_k[1] + m, s0);
          for (int j = 0; j < 3; ++j)
            if (o.sfc_k[j]) unsafeAtomicAdd(o.sfc_k[j] + m, sj[j]);
        }
      }
    }
    if (valid && o.emis_k) {
      const double e_seed = seeded ? o.ks.emis[sl] : 0.0;
      o.emis_k[ol] = sfc_comp ? 0.0 : (double)(upad_sfc * Bs - upad_sfc * rd) + e_seed;
    }
  } else if (valid) {
    const double e_seed = seeded ? o.ks.emis[sl] : 0.0, bs_seed = seeded ? o.ks.bs[sl] : 0.0;
    if (o.emis_k) o.emis_k[ol] = sfc_comp ? 0.0 : (double)(upad_sfc * Bs - upad_sfc * rd) + e_seed;
    const RT dbsdt = c2 * bc2 * Bs * (Bs + c1) * ic1 * rrcp<RT>(tes * tes);
    const double sad = (double)(dbsdt * (eps * upad_sfc + (RT)bs_seed));
    for (int c = 0; c < 4; ++c)
      if (o.sfc_t_k[c]) o.sfc_t_k[c][ol] = bt.cov[c][m] * sad + (c == 1 ? sg[0] : 0.0);
    for (int j = 0; j < 3; ++j)
      if (o.sfc_k[j]) o.sfc_k[j][ol] = sg[1 + j];
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
  // adjoint rows [out_NT] (RED): row of user layer k at tad[k]
  double* tad = reinterpret_cast<double*>(o.t_k) + (bt.out_loff[m] - nx);
  double* aad = reinterpret_cast<double*>(o.a_k) + (bt.out_loff[m] - nx);
  // wave sum of a lane value (channels past L give 0), added to row p by lane 0
  auto red_add = [&](double* p, RT v) {
    const double sv = wave_sum(lane_in ? (double)v : 0.0);
    if (lane == 0) unsafeAtomicAdd(p, sv);
  };
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
      RT tau_ad = -(a_up + a_dn) * iu + tau_sfc_ad;
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
      const RT dbdt = c2 * bc2 * B * (B + c1) * ic1 * rrcp<RT>(te * te);
      if constexpr (RED) {
        if (k >= nx) red_add(tad + k, dbdt * b_ad + t_seed + nl_t(k));   // wave-uniform: every lane takes part
      } else if (valid && k >= nx) {
        tko[(int64_t)k * Lt] = (KT)(dbdt * b_ad + t_seed + nl_t(k));
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
  const cst_ptr<int> acc_i = as_const(w.acc_i + (size_t)m * Ko * 3);
  const cst_ptr<RT> acc_w = as_const(w.acc_w + (size_t)m * NZ_MAX);
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
  const cst_ptr<int> frange = as_const(w.frange + (size_t)m * Ko * 2);
  const cst_ptr<RT> uwin = as_const(w.uwin + (size_t)m * bt.max_k * MW);
  auto final_lo = [&](int k) { return frange[2 * k]; };
  auto prefetch = [&](int k) {   // rows finalized at ODPS layer k, into pfn
    pf_lo_n = final_lo(k);
#pragma unroll
    for (int q = 0; q < PF; ++q) {
      const int iu = pf_lo_n + q;
      pfn[q] = (!RED && valid && iu >= nx && iu < K) ? (RT)tko[(int64_t)iu * Lt] : RT(0);
    }
  };
  auto finish_user_layer = [&](int iu) {   // completed at the ODPS layer in window slot 0
    if constexpr (RED) {   // the Planck part was summed in the Emission_AD pass
      if (iu < nx) return;                         // wave-uniform
      const cst_ptr<RT> wr = uwin + (size_t)iu * MW;
      RT st = RT(0), sa[NAW];
#pragma unroll
      for (int ju = 0; ju < NAW; ++ju) sa[ju] = RT(0);
#pragma unroll
      for (int q = 0; q < MW; ++q) {
        const RT wk = wr[q];
        st += wk * tw[q];
#pragma unroll
        for (int ju = 0; ju < NAW; ++ju) sa[ju] += wk * aw[q][ju];
      }
      red_add(tad + iu, st);
#pragma unroll
      for (int ju = 0; ju < NAW; ++ju)
        if (ju < na) red_add(aad + ju * bt.out_NT + iu, sa[ju]);
      return;
    }
    if (!valid || iu < nx) return;
    RT st;                                         // Planck part from the Emission_AD pass
    if ((unsigned)(iu - pf_lo) < (unsigned)PF) {
      const int d = iu - pf_lo;
      st = pfc[0];
#pragma unroll
      for (int q = 1; q < PF; ++q) st = d == q ? pfc[q] : st;   // value selects keep pfc in registers
    } else {
      st = (RT)tko[(int64_t)iu * Lt];
    }
    const cst_ptr<RT> wr = uwin + (size_t)iu * MW;   // weights from ODPS layers in slots 0..MW-1
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
  cload(Ko - 1, (Ko - 1) % NBUF);
  rload(Ut, RR::NR, Ko - 1, (Ko - 1) % NBUF);
  stage_wait();
  __syncthreads();
  for (int k = Ko - 1; k >= 0; --k) {
    if (k > 0) {
      cload(k - 1, (k - 1) % NBUF);
      rload(Ut, RR::NR, k - 1, (k - 1) % NBUF);
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
      const RT od_ad = getbit(odc_lo, odc_hi, k) ? RT(0) : S;
      const RT* uk = tab[wv][k % NBUF];
      const float* ck = cs[k % NBUF] + lane;
      RT y[Qi::NQ];
#pragma unroll
      for (int q = 0; q < Qi::NQ; ++q) y[q] = RT(0);
      sens_layer<G, RT, 0>(ck, npb, uk, y);
      RT t_ad = y[Qi::TT] * od_ad, vap_ad = RT(0);
      if (blk_opt) {
        // Add_OPTRAN_wloOD_AD (ODPS_AtmAbsorption.f90:911-984) with chi recomputed
        const RT* okr = uk + RR::NR;
        const RT lnchi = oc_lds ? optran_lnchi<RT>(oc_l, meta, okr) : optran_lnchi<RT>(oc_g, meta, okr);
        const bool lcc = getbit(lcc_lo, lcc_hi, k);
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
            e12 += p == 12 ? xa : RT(0);   // selects, not a pointer to one of the two
            e13 += p == 13 ? xa : RT(0);
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
      RT a_ad[Tg::NABS];
#pragma unroll
      for (int j = 0; j < Tg::NABS; ++j) a_ad[j] = y[Qi::RA0 + j] * od_ad;
      if (s.optran) {   // compile-time indices keep a_ad in registers
#pragma unroll
        for (int j = 0; j < Tg::NABS; ++j)
          if (j == h2o_j) a_ad[j] += vap_ad;
      }
      // the integrated ratios accumulate from the top, so their adjoints
      // carry down the sweep (ODPS_Predictor.f90:707-759 reversed)
      const RT pdp = uk[RR::PDP];
#pragma unroll
      for (int j = 0; j < Tg::NABS; ++j) {
        gts[j] += y[Qi::GT0 + j] * od_ad;
        gzs[j] += y[Qi::GZ0 + j] * od_ad;
        t_ad += gts[j] * pdp * uk[RR::AODPS0 + j];
        a_ad[j] += gzs[j] * pdp + gts[j] * pdp * uk[RR::TODPS];
      }
      tzps += y[Qi::TZP] * od_ad;
      tzs += y[Qi::TZ] * od_ad;
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
            if (ju == q && uk[RR::MASK0 + j] != RT(0)) aw[0][q] = a_ad[j];
        }
        // user layers (i2(k-1), i2(k)] are complete
        const int ihi = frange[2 * k + 1];
        for (int iu = pf_lo; iu <= ihi; ++iu) finish_user_layer(iu);
      } else if (RED) {
        // wide supports or more absorbers, adjoint: channel sums per row
        const int i1 = acc_i[3 * k], n = acc_i[3 * k + 1];
        const cst_ptr<RT> awk = acc_w + acc_i[3 * k + 2];
        for (int q = 0; q < n; ++q)
          if (i1 + q >= nx) red_add(tad + i1 + q, awk[q] * t_ad);
#pragma unroll
        for (int j = 0; j < Tg::NABS; ++j) {
          const int ju = bt.jmap[j];
          if (ju >= 0 && uk[RR::MASK0 + j] != RT(0))
            for (int q = 0; q < n; ++q)
              if (i1 + q >= nx) red_add(aad + ju * bt.out_NT + i1 + q, awk[q] * a_ad[j]);
        }
      } else if (valid) {
        // wide supports or more absorbers: add onto the output rows.  Each
        // lane owns its channel's column, so the adds need no atomicity; the
        // returnless atomic add retires without a load round trip (same-lane,
        // same-address operations stay in program order).
        const int i1 = acc_i[3 * k], n = acc_i[3 * k + 1];
        const cst_ptr<RT> awk = acc_w + acc_i[3 * k + 2];
        for (int q = 0; q < n; ++q)
          if (i1 + q >= nx) unsafeAtomicAdd(tko + (int64_t)(i1 + q) * Lt, (KT)(awk[q] * t_ad));
#pragma unroll
        for (int j = 0; j < Tg::NABS; ++j) {
          const int ju = bt.jmap[j];
          if (ju >= 0 && uk[RR::MASK0 + j] != RT(0))
            for (int q = 0; q < n; ++q)
              if (i1 + q >= nx) unsafeAtomicAdd(ako + ju * astr + (int64_t)(i1 + q) * Lt, (KT)(awk[q] * a_ad[j]));
        }
      }
    }
    stage_wait();
    __syncthreads();
  }
}

template <int G, class RT, class KT, bool DOK, bool RED, int PW>
void launch_fwdk_pw(const DevSensor& s, const DevBatch& b, const ProfWork<RT>& w, const DevOut& o,
                    hipStream_t stream) {
  const int64_t ng = (b.M + PW - 1) / PW;
  const int64_t n = 8 * ((s.n_blocks * ng + 7) / 8);
  hipLaunchKernelGGL((clear_fwdk_kernel<G, RT, KT, DOK, RED, PW>), dim3((unsigned)n), dim3(PW * WAVE), 0, stream,
                     s, b, w, o);
}

// The three tile sets are instantiated in three translation units compiled in
// parallel: this file (index 1, the default), clear_fwdk_pw0.hip and
// clear_fwdk_pw2.hip, each defining CRTM_FWDK_TILE before including this one.
#ifndef CRTM_FWDK_TILE
#define CRTM_FWDK_TILE 1
#define CRTM_FWDK_DISPATCH 1
#endif

#define INST(G, RT, KT, DOK, RED) \
  template void launch_fwdk_pw<G, RT, KT, DOK, RED, TileSet<RT>::pw[CRTM_FWDK_TILE]>( \
      const DevSensor&, const DevBatch&, const ProfWork<RT>&, const DevOut&, hipStream_t);
#define INST_G(G)                                                                                     \
  INST(G, float, float, true, false) INST(G, float, double, true, false)                              \
  INST(G, float, float, false, false) INST(G, double, double, true, false)                            \
  INST(G, double, float, true, false) INST(G, double, double, false, false)                           \
  INST(G, float, double, true, true) INST(G, double, double, true, true)
INST_G(1)
INST_G(2)
INST_G(3)

#ifdef CRTM_FWDK_DISPATCH
// the other two tile sets come from their own translation units
#define EXT(G, RT, KT, DOK, RED, T) \
  extern template void launch_fwdk_pw<G, RT, KT, DOK, RED, TileSet<RT>::pw[T]>( \
      const DevSensor&, const DevBatch&, const ProfWork<RT>&, const DevOut&, hipStream_t);
#define EXT_G(G, T)                                                                                   \
  EXT(G, float, float, true, false, T) EXT(G, float, double, true, false, T)                          \
  EXT(G, float, float, false, false, T) EXT(G, double, double, true, false, T)                        \
  EXT(G, double, float, true, false, T) EXT(G, double, double, false, false, T)                       \
  EXT(G, float, double, true, true, T) EXT(G, double, double, true, true, T)
EXT_G(1, 0) EXT_G(2, 0) EXT_G(3, 0)
EXT_G(1, 2) EXT_G(2, 2) EXT_G(3, 2)
// tile index -> instantiation (any other index: the default)
template <int G, class RT, class KT, bool DOK, bool RED>
void launch_fwdk(const DevSensor& s, const DevBatch& b, const ProfWork<RT>& w, const DevOut& o, hipStream_t stream,
                 int tile) {
  switch (tile) {
    case 0: launch_fwdk_pw<G, RT, KT, DOK, RED, TileSet<RT>::pw[0]>(s, b, w, o, stream); break;
    case 2: launch_fwdk_pw<G, RT, KT, DOK, RED, TileSet<RT>::pw[2]>(s, b, w, o, stream); break;
    default: launch_fwdk_pw<G, RT, KT, DOK, RED, TileSet<RT>::pw[1]>(s, b, w, o, stream); break;
  }
}
#define INSTD(G, RT, KT, DOK, RED) \
  template void launch_fwdk<G, RT, KT, DOK, RED>(const DevSensor&, const DevBatch&, const ProfWork<RT>&, \
                                                 const DevOut&, hipStream_t, int);
#define INSTD_G(G)                                                                                    \
  INSTD(G, float, float, true, false) INSTD(G, float, double, true, false)                            \
  INSTD(G, float, float, false, false) INSTD(G, double, double, true, false)                          \
  INSTD(G, double, float, true, false) INSTD(G, double, double, false, false)                         \
  INSTD(G, float, double, true, true) INSTD(G, double, double, true, true)
INSTD_G(1)
INSTD_G(2)
INSTD_G(3)
#endif

}  // namespace crtm
