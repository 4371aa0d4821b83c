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
#include "allsky_common.h"

namespace crtm {


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
  // aerosols (Atmosphere%Aerosol): every typed slot's radius enters the
  // stream choice (CRTM_Compute_nStreams); layers holding aerosol are
  // scattering layers like cloudy ones
  auto aerosol_layer = [&](int kk) -> bool {
    const int ku = kk - nx;
    if (ku < 0) return false;
    for (int n = 0; n < cl.na; ++n)
      if (cl.atype[(int64_t)m * cl.na + n] != 0 && cl.aconc[n * NTu + lu + ku] > AEROSOL_CONTENT_THRESHOLD)
        return true;
    return false;
  };
  for (int n = 0; n < cl.na; ++n)
    if (cl.atype[(int64_t)m * cl.na + n] != 0)
      for (int ku = 0; ku < K - nx; ++ku) max_reff = fmax(max_reff, cl.areff[n * NTu + lu + ku]);
  // a skipped profile (Options(m)%Skip_Profile) takes no scattering slots
  const bool skip = be.skip && be.skip[m];
  if (skip) fractional = false;
  int count = 0;
  for (int kk = 0; kk < K; ++kk) {
    bool cloudy = false;
    for (int n = 0; n < nc; ++n) cloudy |= cw.wc[n * NTe + le + kk] > WATER_CONTENT_THRESHOLD;
    const bool scat = !skip && (cloudy || aerosol_layer(kk));
    count += scat;
    cw.cl_slot[le + kk] = scat ? 0 : -1;   // slot numbers follow in cloud_slot_kernel
  }
  if (count) atomicMax(cw.total + 1, K);
  cw.cl_count[m] = count;
  cw.tcc[m] = fractional ? tcc : 0.0;
  cw.frac[m] = fractional;
  cw.max_reff[m] = max_reff;
}

// Exclusive prefix sum of the per-profile scattering-layer counts, one
// workgroup: slots are numbered in profile order, so which profiles overflow
// a layer capacity is the same on every call and every graph replay.
constexpr int SCAN_THREADS = 1024;
__global__ void __launch_bounds__(SCAN_THREADS) cloud_scan_kernel(int M, CloudWork cw) {
  __shared__ int part[SCAN_THREADS];
  int carry = 0;
  for (int m0 = 0; m0 < M; m0 += SCAN_THREADS) {
    const int m = m0 + (int)threadIdx.x;
    const int v = m < M ? cw.cl_count[m] : 0;
    part[threadIdx.x] = v;
    __syncthreads();
    for (int d = 1; d < SCAN_THREADS; d <<= 1) {   // Hillis-Steele inclusive scan
      const int add = threadIdx.x >= (unsigned)d ? part[threadIdx.x - d] : 0;
      __syncthreads();
      part[threadIdx.x] += add;
      __syncthreads();
    }
    if (m < M) cw.cl_base[m] = carry + part[threadIdx.x] - v;
    carry += part[SCAN_THREADS - 1];
    __syncthreads();   // part[] is rewritten by the next chunk
  }
  if (threadIdx.x == 0) cw.total[0] = carry;
}

// Slot numbers of each profile's scattering layers (after the scan); a
// profile whose layers do not all fit the slot capacity fails, and so does
// one whose per-profile stream count (device memory, not checked on the
// host in a captured call) is not one the kernels take.
__global__ void __launch_bounds__(64) cloud_slot_kernel(DevBatch be, DevClouds cl, CloudWork cw) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= be.M) return;
  const int K = be.n_layers[m], count = cw.cl_count[m], base = cw.cl_base[m];
  const int64_t le = be.loff[m];
  int j = 0;
  for (int kk = 0; kk < K && j < count; ++kk) {
    if (cw.cl_slot[le + kk] < 0) continue;
    int slot = -1;
    if (base + j < cw.cap) {
      slot = base + j;
      cw.cl_desc[slot] = make_int2(m, kk);
    }
    ++j;
    cw.cl_slot[le + kk] = slot;
  }
  int err = (count > 0 && base + count > cw.cap) ? 2 : 0;
  if (cl.ns_prof) {
    const int n = cl.ns_prof[m];
    if (!(n == 0 || n == 2 || n == 4 || n == 6 || n == 8 || n == 16)) err = 3;
  }
  cw.aer_err[m] = err;
}

// =========================================================================
// layer optics: one thread per (cloudy layer, channel)
// =========================================================================
template <class RT>
__global__ void __launch_bounds__(256) cloud_optics_kernel(DevSensor s, DevBatch be, DevClouds cl, CloudWork cw,
                                                            DevCloudCoeff cc, LayerScratch<RT> sc, const RT* tau_g,
                                                            int n_items) {
  const bool mw = s.sensor_type == 1;
  const int n_live = live_items(cw, s.L, n_items);
  for (int item = blockIdx.x * blockDim.x + threadIdx.x; item < n_live; item += gridDim.x * blockDim.x) {
    const int c = item / s.L, l = item - c * s.L;
    const int2 d = cw.cl_desc[c];
    const int m = d.x, k = d.y;
    if (cw.status[m] != 0) continue;
    const int nl = legendre_terms(profile_streams(cl, m), cw.max_reff[m], s.wavenumber[l]);
    if (nl > NL_W - 1) continue;   // stream counts above the built tables are rejected by the host
    const int64_t ik = be.loff[m] + k;
    // CRTM_Compute_CloudScatter: accumulate over the clouds of the layer
    double od = (double)tau_g[ik * s.L + l], ssa = 0.0, pl[NL_W];
#pragma unroll
    for (int q = 0; q < NL_W; ++q) pl[q] = 0.0;
    const int loff = legendre_offset(nl);
    const double freq = mw ? s.frequency[l] : s.wavenumber[l];
#pragma unroll 1
    for (int n = 0; n < cl.nc; ++n) {
      const double wc = cw.wc[n * be.NT + ik];
      if (wc > WATER_CONTENT_THRESHOLD)
        add_cloud(cc, mw, freq, cl.type[(int64_t)m * cl.nc + n], cw.reff[n * be.NT + ik], be.t[ik], wc, nl, loff,
                  cl.no_scat != 0, od, ssa, pl);
    }
    // CRTM_Compute_AerosolScatter (src/AtmScatter/CRTM_AerosolScatter.f90:170-317): infrared only
    const int ku = k - be.n_extra[m];
    if (!mw && cl.na > 0 && ku >= 0) {
      const int aoff = aerosol_offset(nl);
      const int64_t iu = be.out_loff[m] + ku;
#pragma unroll 1
      for (int n = 0; n < cl.na; ++n) {
        const int ty = cl.atype[(int64_t)m * cl.na + n];
        const double conc = cl.aconc[n * be.out_NT + iu];
        if (ty == 0 || !(conc > AEROSOL_CONTENT_THRESHOLD)) continue;
        const int ka = aerosol_index(cc.aero, ty);
        if (aoff < 0 || ka < 0) continue;   // failed profile (aerosol_check_kernel)
        add_aerosol(cc.aero, s.wavenumber[l], ka, cl.areff[n * be.out_NT + iu], conc, nl, aoff, cl.no_scat != 0, od,
                    ssa, pl);
      }
    }
    // CRTM_AtmOptics_Combine: P_l /= ssa, P_0 = 1/2, delta truncation f = P_nl
    if (ssa > BS_THRESHOLD) {
      const double w = ssa / od;
      const double iss = 1.0 / ssa;
#pragma unroll
      for (int q = 1; q < NL_W; ++q) pl[q] *= iss;
      pl[0] = 0.5;
      double f = 0.0;
#pragma unroll
      for (int q = 1; q < NL_W; ++q)
        if (q == nl) f = pl[q];
      od = (1.0 - f * w) * od;
      ssa = (1.0 - f) * w / (1.0 - f * w);
    }
    RT* out = sc.v + item;
    const int64_t st = sc.n_items;
    out[st * sc.oTau()] = (RT)od;
    out[st * sc.oSsa()] = (RT)ssa;
#pragma unroll
    for (int q = 0; q < NL_W - 1; ++q)
      if (q < sc.npc) out[st * (sc.oP() + q)] = (RT)(q < nl ? pl[q] : 0.0);
  }
}

// =========================================================================
// layer reflection / transmission: one thread per (cloudy layer, channel)
// of angle count NZ
// =========================================================================
template <class RT, int NZ>
__global__ void __launch_bounds__(256) cloud_rt_kernel(DevSensor s, CloudWork cw, DevClouds cl, DevCloudCoeff cc,
                                                        LayerScratch<RT> sc, const double* geo, int n_items) {
  const int n_live = live_items(cw, s.L, n_items);
  for (int item = blockIdx.x * blockDim.x + threadIdx.x; item < n_live; item += gridDim.x * blockDim.x) {
    const int c = item / s.L, l = item - c * s.L;
    const int m = cw.cl_desc[c].x;
    if (cw.status[m] != 0) continue;
    const int nl = legendre_terms(profile_streams(cl, m), cw.max_reff[m], s.wavenumber[l]);
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

// Infrared channels whose stream count has no aerosol phase function (two
// streams) fail the profile when the atmosphere carries aerosol slots, as
// CRTM_Compute_AerosolScatter does (CRTM_AerosolScatter.f90:208-225); so do
// aerosol types the AerosolCoeff tables lack (AerosolCoeff_typeID_to_index).
__global__ void aerosol_check_kernel(DevSensor s, int M, DevClouds cl, CloudWork cw, DevAerosolCoeff ac) {
  const int64_t n = (int64_t)M * s.L;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const int m = (int)(t / s.L), l = (int)(t - (int64_t)m * s.L);
    const int nl = legendre_terms(profile_streams(cl, m), cw.max_reff[m], s.wavenumber[l]);
    // the profile's n_Aerosols counts its typed slots (type 0: unused slot)
    bool typed = false, bad = false;
    for (int q = 0; q < cl.na; ++q) {
      const int ty = cl.atype[(int64_t)m * cl.na + q];
      typed |= ty != 0;
      if (ty != 0 && aerosol_index(ac, ty) < 0) bad = true;
    }
    if (typed && aerosol_offset(nl) < 0) bad = true;
    if (bad) atomicMax(cw.aer_err + m, 1);   // a capacity overflow (2) keeps its code
  }
}
// Call-level profile status, merged after each sensor: the preparation
// status of the sensor (layer mapping) and the profile failures of the
// scattering path (aerosol check, layer capacity).  The status only rises,
// so a failure of one sensor fails the profile for the whole call, as a
// channel error fails the profile in the reference
// (src/CRTM_Forward_Module.f90:412-421).
__global__ void status_merge_kernel(int M, const int* prep_status, const int* aer_err, int* call_status) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  if (prep_status[m] == ST_SKIP) return;   // a skipped profile does not fail the call
  int st = call_status[m];
  if (prep_status[m] > st) st = prep_status[m];
  if (aer_err && aer_err[m]) {
    // 5: aerosol failure, 6: layer capacity exceeded, 8: invalid stream count
    const int a = aer_err[m] == 2 ? 6 : aer_err[m] == 3 ? 8 : 5;
    if (a > st) st = a;
  }
  call_status[m] = st;
}
// zero fill of n words (4-byte aligned), a kernel so that capturable calls
// record no memset nodes: 16-byte stores from the first 16-byte boundary,
// single words before it and after the last full quad
__global__ void zero_fill_kernel(uint32_t* p, int64_t n_words) {
  const int64_t lead = (int64_t)((16 - ((uintptr_t)p & 15)) & 15) / 4;   // words before the boundary
  const int64_t head = lead < n_words ? lead : n_words;
  const int64_t nq = (n_words - head) / 4;
  uint4* q = reinterpret_cast<uint4*>(p + head);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  for (int64_t t = t0; t < nq; t += stride) q[t] = make_uint4(0u, 0u, 0u, 0u);
  if (t0 < head) p[t0] = 0u;
  const int64_t tail = head + 4 * nq;
  if (t0 < n_words - tail) p[tail + t0] = 0u;
}

// ------------------------------------------------------------------ launchers
void launch_cloud_prep(const DevBatch& be, const DevClouds& cl, const CloudWork& cw, hipStream_t st) {
  hipLaunchKernelGGL(cloud_prep_kernel, dim3((be.M + 63) / 64), dim3(64), 0, st, be, cl, cw);
  hipLaunchKernelGGL(cloud_scan_kernel, dim3(1), dim3(SCAN_THREADS), 0, st, be.M, cw);
  hipLaunchKernelGGL(cloud_slot_kernel, dim3((be.M + 63) / 64), dim3(64), 0, st, be, cl, cw);
}

void launch_aerosol_check(const DevSensor& s, int M, const DevClouds& cl, const CloudWork& cw,
                          const DevCloudCoeff& cc, hipStream_t st) {
  if (s.sensor_type == 1 || cl.na == 0) return;   // aerosols scatter in the infrared only
  const int64_t n = (int64_t)M * s.L;
  hipLaunchKernelGGL(aerosol_check_kernel, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 8192)), dim3(256), 0,
                     st, s, M, cl, cw, cc.aero);
}

void launch_zero(void* p, size_t bytes, hipStream_t st) {
  const int64_t n = (int64_t)(bytes / 4);
  if (n > 0)
    hipLaunchKernelGGL(zero_fill_kernel, dim3((unsigned)std::min<int64_t>((n / 4 + 255) / 256 + 1, 16384)), dim3(256),
                       0, st, static_cast<uint32_t*>(p), n);
}

void launch_status_merge(int M, const int* prep_status, const int* aer_err, int* call_status, hipStream_t st) {
  hipLaunchKernelGGL(status_merge_kernel, dim3((M + 255) / 256), dim3(256), 0, st, M, prep_status, aer_err,
                     call_status);
}

// wavefront kernels for 9 and 10 angles (allsky_wide.hip)
template <class RT>
void launch_cloud_rt_wide(int nz, const DevSensor& s, const CloudWork& cw, const DevClouds& cl,
                          const DevCloudCoeff& cc, const LayerScratch<RT>& sc, const double* geo, int n_items,
                          hipStream_t st);
template <class RT>
void launch_sweep_wide(int nz, const DevSensor& s, const DevBatch& be, const DevClouds& cl, const CloudWork& cw,
                       const DevCloudCoeff& cc, const LayerScratch<RT>& sc, const RT* tau_g, const double* geo,
                       const DevOut& o, const AllskyOut& ao, hipStream_t st);

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
    if (nz_list[q] > sc.nzc) {   // wavefront kernels
      launch_cloud_rt_wide<RT>(nz_list[q], s, cw, cl, cc, sc, geo, n_items, st);
      launch_sweep_wide<RT>(nz_list[q], s, be, cl, cw, cc, sc, tau_g, geo, o, ao, st);
      continue;
    }
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

// forward optics and layer R/T kernels, for the K-matrix path in allsky_ad.hip
template <class RT>
void launch_cloud_optics(const DevSensor& s, const DevBatch& be, const DevClouds& cl, const CloudWork& cw,
                         const DevCloudCoeff& cc, const LayerScratch<RT>& sc, const RT* tau_g, int n_items,
                         hipStream_t st) {
  if (n_items > 0)
    hipLaunchKernelGGL((cloud_optics_kernel<RT>), dim3(persistent_blocks(n_items)), dim3(256), 0, st, s, be, cl, cw,
                       cc, sc, tau_g, n_items);
}
template <class RT>
void launch_cloud_rt(int nz, const DevSensor& s, const CloudWork& cw, const DevClouds& cl, const DevCloudCoeff& cc,
                     const LayerScratch<RT>& sc, const double* geo, int n_items, hipStream_t st) {
  if (n_items <= 0) return;
  const int lb = persistent_blocks(n_items);
  switch (nz) {
#define CRTM_NZ_CASE(N)                                                                                         \
  case N:                                                                                                       \
    hipLaunchKernelGGL((cloud_rt_kernel<RT, N>), dim3(lb), dim3(256), 0, st, s, cw, cl, cc, sc, geo, n_items);   \
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
template void launch_cloud_optics<float>(const DevSensor&, const DevBatch&, const DevClouds&, const CloudWork&,
                                         const DevCloudCoeff&, const LayerScratch<float>&, const float*, int,
                                         hipStream_t);
template void launch_cloud_optics<double>(const DevSensor&, const DevBatch&, const DevClouds&, const CloudWork&,
                                          const DevCloudCoeff&, const LayerScratch<double>&, const double*, int,
                                          hipStream_t);
template void launch_cloud_rt<float>(int, const DevSensor&, const CloudWork&, const DevClouds&, const DevCloudCoeff&,
                                     const LayerScratch<float>&, const double*, int, hipStream_t);
template void launch_cloud_rt<double>(int, const DevSensor&, const CloudWork&, const DevClouds&, const DevCloudCoeff&,
                                      const LayerScratch<double>&, const double*, int, hipStream_t);

// R_Clear = Radiance etc. for every profile (cl_count null) or for those without cloudy layers
void launch_clear_copy(const DevSensor& s, int M, int Ltot, const int* cl_count, const DevOut& o, const AllskyOut& ao,
                       hipStream_t st) {
  if (!(ao.r_clear || ao.tb_clear || ao.tcc)) return;
  hipLaunchKernelGGL(allsky_clear_copy_kernel, dim3(persistent_blocks((int64_t)M * s.L)), dim3(256), 0, st, s, M,
                     Ltot, cl_count, o, ao);
}

template void launch_allsky<float>(const DevSensor&, const DevBatch&, const DevClouds&, const CloudWork&,
                                   const DevCloudCoeff&, const LayerScratch<float>&, const float*, const double*,
                                   int, const int*, int, const DevOut&, const AllskyOut&, hipStream_t);
template void launch_allsky<double>(const DevSensor&, const DevBatch&, const DevClouds&, const CloudWork&,
                                    const DevCloudCoeff&, const LayerScratch<double>&, const double*,
                                    const double*, int, const int*, int, const DevOut&, const AllskyOut&,
                                    hipStream_t);

}  // namespace crtm
