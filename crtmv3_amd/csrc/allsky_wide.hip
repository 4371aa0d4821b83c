// allsky_wide.hip -- all-sky forward kernels at 16 streams (9 or 10 angles):
// one wavefront per work unit, matrices spread over its 64 lanes and
// multiplied on the matrix cores (allsky_wide.h).
//
//   cloud_rt_wide_kernel     one wavefront per (scattering layer, channel):
//                            phase matrix + Normalize_Phase
//                            (src/RTSolution/Common_RTSolution.f90:1857-1920,
//                            2885-2979) and CRTM_AMOM_layer with Matrix_Exp
//                            and doubling (src/RTSolution/ADA/ADA_Module.f90:
//                            552-680, 746-853).
//   allsky_sweep_wide_kernel one wavefront per (profile, channel): ADA adding
//                            from the surface up (ADA_Module.f90:97-258) or
//                            Emission when no layer scatters, fractional cover
//                            combination (src/CRTM_Forward_Module.f90:1086-1120), BT.
//
// The arithmetic is that of the per-lane kernels in allsky.hip; only the
// summation order of the matrix products differs.
#include "allsky_wide.h"

namespace crtm {

// =========================================================================
// layer reflection / transmission: one wavefront per (scattering layer, channel)
// =========================================================================
template <class RT, int NZ>
__global__ void __launch_bounds__(64) cloud_rt_wide_kernel(DevSensor s, CloudWork cw, DevClouds cl, DevCloudCoeff cc,
                                                           LayerScratch<RT> sc, const double* geo, int n_items) {
  __shared__ RT buf[16 * WLD];
  const int n_live = live_items(cw, s.L, n_items);
  for (int item = blockIdx.x; item < n_live; item += gridDim.x) {
    const int c = item / s.L, l = item - c * s.L;
    const int m = cw.cl_desc[c].x;
    if (cw.status[m] != 0) continue;
    const int nl = legendre_terms(profile_streams(cl, m), cw.max_reff[m], s.wavenumber[l]);
    if (nl > NL_W - 1) continue;
    const double cosu = geo[(int64_t)m * 4 + 1];
    const AngleSet as = angle_set(cc.gauss, (nl + 2) / 2, cosu);
    if (as.nz != NZ) continue;
    const int64_t st = sc.n_items;
    const RT om = sc.v[st * sc.oSsa() + item];
    if (!((double)om > BS_THRESHOLD)) continue;
    const RT tau = sc.v[st * sc.oTau() + item];
    RT plr[NL_W - 1];
#pragma unroll
    for (int q = 0; q < NL_W - 1; ++q) plr[q] = q < sc.npc ? sc.v[st * (sc.oP() + q) + item] : RT(0);
    const WAngles<RT> a = wangles<RT, NZ>(cc.gauss, as.ns, cosu);
    WPhase<RT, NZ> pf;
    WM<RT> fs, fd, R, Tm;
    WCol<RT> e;
    wphase_forward<RT, NZ>(as.ns, nl, plr, a, buf, pf, fs, fd);
    wamom_layer<RT, NZ>(as.ns, tau, om, a, fs, fd, buf, R, Tm, e);
    RT* out = sc.w + (int64_t)item * sc.w_stride();
    wstore(out, R, NZ);
    wstore(out + NZ * NZ, Tm, NZ);
    wstore_col(out + 2 * NZ * NZ, e, NZ);
  }
}

// =========================================================================
// adding sweep: one wavefront per (profile, channel) of the cloudy profiles
// =========================================================================
template <class RT, int NZ>
__global__ void __launch_bounds__(64) allsky_sweep_wide_kernel(DevSensor s, DevBatch be, DevClouds cl, CloudWork cw,
                                                               DevCloudCoeff cc, LayerScratch<RT> sc, const RT* tau_g,
                                                               const double* geo, DevOut o, AllskyOut ao) {
  __shared__ double buf[16 * WLD];
  RT* rbuf = reinterpret_cast<RT*>(buf);
  const int j = wj(), g = wg();
  const int64_t n = (int64_t)be.M * s.L;
  for (int64_t t = blockIdx.x; t < n; t += gridDim.x) {
    const int m = (int)(t / s.L), l = (int)(t - (int64_t)m * s.L);
    if (cw.cl_count[m] == 0 || cw.status[m] != 0 || cw.aer_err[m] != 0) continue;
    const double cosu = geo[(int64_t)m * 4 + 1];
    const int nl = legendre_terms(profile_streams(cl, m), cw.max_reff[m], s.wavenumber[l]);
    if (nl > NL_W - 1) continue;
    const AngleSet as = angle_set(cc.gauss, (nl + 2) / 2, cosu);
    if (as.nz != NZ) continue;
    const int K = be.n_layers[m];
    const int64_t le = be.loff[m];
    const int lg = s.chan_offset + l;
    const int64_t ol = (int64_t)m * be.Ltot + lg;
    const int64_t st = sc.n_items;
    // Scattering_RT: some layer of the column scatters
    bool mine = false;
    for (int k = threadIdx.x; k < K; k += 64) {
      const int c = cw.cl_slot[le + k];
      if (c >= 0 && (double)sc.v[st * sc.oSsa() + (int64_t)c * s.L + l] > BS_THRESHOLD) mine = true;
    }
    const bool scat = __ballot(mine) != 0;
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
      const WAngles<RT> a = wangles<RT, NZ>(cc.gauss, as.ns, cosu);
      WM<RT> rup;
      WCol<RT> iup;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = WL<RT>::row(g, r);
        rup.v[r] = i == j && i < NZ ? RT(1) - eps : RT(0);
        iup.v[r] = i < NZ ? eps * Bs : RT(0);
      }
      for (int k = K - 1; k >= 0; --k) {
        RT te;
        const RT B = planck(c1, c2, bc1, bc2, be.t[le + k], te);
        const int c = cw.cl_slot[le + k];
        const int64_t it = (int64_t)c * s.L + l;
        if (c >= 0 && (double)sc.v[st * sc.oSsa() + it] > BS_THRESHOLD) {
          const RT* lay = sc.w + it * sc.w_stride();
          const WM<RT> rl = wload<RT>(lay, NZ), tl = wload<RT>(lay + NZ * NZ, NZ);
          WCol<RT> src = wload_col<RT>(lay + 2 * NZ * NZ, NZ);
#pragma unroll
          for (int r = 0; r < 4; ++r) src.v[r] *= B;
          // Gamma = (I - R_up R_layer)^-1 and T Gamma in FP64
          WM<double> gm = wadd(weye<double>(NZ), wnn(wconv<double>(rup, buf), wconv<double>(rl, buf), buf), -1.0);
          winv(gm, NZ);
          const WM<RT> gt = wconv<RT>(wnn(wconv<double>(tl, buf), gm, buf), buf);
          WCol<RT> v = wmv(rup, wc2r(src));
#pragma unroll
          for (int r = 0; r < 4; ++r) v.v[r] += iup.v[r];
          const WCol<RT> gv = wmv(gt, wc2r(v));
#pragma unroll
          for (int r = 0; r < 4; ++r) iup.v[r] = src.v[r] + gv.v[r];
          rup = wadd(rl, wnn(gt, wnn(rup, tl, rbuf), rbuf));
        } else {
          const RT tau = layer_tau(k);
          const RT tr_j = j < NZ ? rexp<RT>(-tau * a.imu_j) : RT(0);
          const RT sr_j = j < NZ ? B * (RT(1) - tr_j) : RT(0);
          const WCol<RT> w = wmv(rup, sr_j);
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const bool in = WL<RT>::row(g, r) < NZ;
            const RT tr = in ? rexp<RT>(-tau * a.imu_r[r]) : RT(0);
            const RT sr = in ? B * (RT(1) - tr) : RT(0);
            iup.v[r] = sr + tr * (iup.v[r] + w.v[r]);
            rup.v[r] = tr * rup.v[r] * tr_j;
          }
        }
      }
      // cosmic background reflected at the top; radiance at the sensor angle
      const int is = as.isat;
      const RT top = wat(iup, is);
      const RT rs = __shfl(wsum16(wsel(rup.v, WL<RT>::reg(is))), 16 * WL<RT>::grp(is));
      r_cloudy = (double)(top + rs * cosmic);
    }
    if (threadIdx.x != 0) continue;
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

// ------------------------------------------------------------------ launchers
static int wide_blocks(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(n, 16384)); }

// angle counts of the wavefront kernels: 8 streams (6) and 16 streams (9, 10)
#define CRTM_WIDE_NZ(X) X(6) X(9) X(10)

template <class RT>
void launch_cloud_rt_wide(int nz, const DevSensor& s, const CloudWork& cw, const DevClouds& cl,
                          const DevCloudCoeff& cc, const LayerScratch<RT>& sc, const double* geo, int n_items,
                          hipStream_t st) {
  if (n_items <= 0) return;
  switch (nz) {
#define CRTM_CASE(N)                                                                                             \
  case N:                                                                                                        \
    hipLaunchKernelGGL((cloud_rt_wide_kernel<RT, N>), dim3(wide_blocks(n_items)), dim3(64), 0, st, s, cw, cl, cc, \
                       sc, geo, n_items);                                                                        \
    break;
    CRTM_WIDE_NZ(CRTM_CASE)
#undef CRTM_CASE
    default:
      break;
  }
}

template <class RT>
void launch_sweep_wide(int nz, const DevSensor& s, const DevBatch& be, const DevClouds& cl, const CloudWork& cw,
                       const DevCloudCoeff& cc, const LayerScratch<RT>& sc, const RT* tau_g, const double* geo,
                       const DevOut& o, const AllskyOut& ao, hipStream_t st) {
  const int64_t n = (int64_t)be.M * s.L;
  switch (nz) {
#define CRTM_CASE(N)                                                                                              \
  case N:                                                                                                         \
    hipLaunchKernelGGL((allsky_sweep_wide_kernel<RT, N>), dim3(wide_blocks(n)), dim3(64), 0, st, s, be, cl, cw, cc, \
                       sc, tau_g, geo, o, ao);                                                                    \
    break;
    CRTM_WIDE_NZ(CRTM_CASE)
#undef CRTM_CASE
    default:
      break;
  }
}

template void launch_cloud_rt_wide<float>(int, const DevSensor&, const CloudWork&, const DevClouds&,
                                          const DevCloudCoeff&, const LayerScratch<float>&, const double*, int,
                                          hipStream_t);
template void launch_cloud_rt_wide<double>(int, const DevSensor&, const CloudWork&, const DevClouds&,
                                           const DevCloudCoeff&, const LayerScratch<double>&, const double*, int,
                                           hipStream_t);
template void launch_sweep_wide<float>(int, const DevSensor&, const DevBatch&, const DevClouds&, const CloudWork&,
                                       const DevCloudCoeff&, const LayerScratch<float>&, const float*, const double*,
                                       const DevOut&, const AllskyOut&, hipStream_t);
template void launch_sweep_wide<double>(int, const DevSensor&, const DevBatch&, const DevClouds&, const CloudWork&,
                                        const DevCloudCoeff&, const LayerScratch<double>&, const double*,
                                        const double*, const DevOut&, const AllskyOut&, hipStream_t);

}  // namespace crtm
