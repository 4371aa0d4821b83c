// allsky_wide_ad.hip -- all-sky K-matrix at 16 streams (9 or 10 angles):
// the hand-written adjoint of the wavefront kernels of allsky_wide.hip, one
// wavefront per work unit with the stream matrices on the matrix cores.
//
//   allsky_sweep_ad_wide_kernel  one wavefront per (profile, channel): the
//                                forward adding sweep storing each layer's
//                                incoming (R_up, I_up) on a per-wavefront tape,
//                                the fractional-cover combination and BT, then
//                                CRTM_ADA_AD (src/RTSolution/ADA/ADA_Module.f90:
//                                1547-1756) or CRTM_Emission_AD (src/RTSolution/
//                                Emission/Emission_Module.f90:280-395).
//   cloud_rt_ad_wide_kernel      one wavefront per scattering (layer, channel):
//                                CRTM_AMOM_layer_AD with Matrix_Exp_AD
//                                (ADA_Module.f90:963-1109, 1760-2081) and the
//                                phase matrix + Normalize_Phase adjoint
//                                (src/RTSolution/Common_RTSolution.f90:2548-2879,
//                                3044-3116); the states entering the doubling
//                                steps come from a per-wavefront tape.
//
// The reverse recurrences are those of the per-lane adjoints in allsky_ad.hip;
// the optics, cover and clear-sky adjoints that follow are shared with them.
#include "allsky_wide.h"

namespace crtm {

namespace {

// Reverse of one doubling step entered with (R, T); Ra, Ta in: adjoints of the
// step's outputs, out: adjoints of its inputs.
template <class T>
__device__ void wdouble_step_ad(const WM<T>& R, const WM<T>& Tr, WM<T>& Ra, WM<T>& Ta, int n, T* buf) {
  WM<T> mi = wadd(weye<T>(n), wnn(R, R, buf), T(-1));
  winv(mi, n);
  const WM<T> t1 = wnn(Tr, mi, buf);
  const WM<T> TrT = wtr(Tr, buf), miT = wtr(mi, buf);
  // t1_ad = Ra (R T)^T + Ta T^T
  const WM<T> x0 = wnn(R, Tr, buf);
  const WM<T> t1a = wtn(wtr(Ta, buf), TrT, wtn(wtr(Ra, buf), wtr(x0, buf)));
  // nR = Ra + t1^T Ra T^T
  WM<T> nR = wadd(Ra, wtn(wtr(wtn(t1, Ra), buf), TrT));
  // nT = (t1 R)^T Ra + t1^T Ta + t1_ad mi^T
  const WM<T> nT = wtn(wtr(t1a, buf), miT, wtn(t1, Ta, wtn(wnn(t1, R, buf), Ra)));
  // mi_ad = T^T t1_ad;  rr_ad = mi^T mi_ad mi^T;  nR += rr_ad R^T + R^T rr_ad
  const WM<T> rr = wtn(wtr(wtn(mi, wtn(Tr, t1a)), buf), miT);
  nR = wtn(R, rr, wtn(wtr(rr, buf), wtr(R, buf), nR));
  Ra = nR;
  Ta = nT;
}

// Adjoint of the pair of Horner series that share h (allsky_ad.hip
// series_pair_ad): Z = sum_n sum_{a+b=n-1} (h^T)^a Q_n (h^T)^b,
// Q_n = Xe / (2n)! + Xo / (2n+1)!  (hT = h^T)
template <class T>
__device__ WM<T> wseries_pair_ad(const WM<T>& h, const WM<T>& hT, const WM<T>& Xe, const WM<T>& Xo, T* buf) {
  T ce[MAX_N_AMOM + 1], co[MAX_N_AMOM + 1];
  ce[0] = co[0] = T(1);
#pragma unroll
  for (int n = 1; n <= MAX_N_AMOM; ++n) {
    ce[n] = ce[n - 1] / T((2 * n) * (2 * n - 1));
    co[n] = co[n - 1] / T((2 * n) * (2 * n + 1));
  }
  WM<T> Y, Z;
#pragma unroll
  for (int r = 0; r < 4; ++r) Z.v[r] = Y.v[r] = ce[MAX_N_AMOM] * Xe.v[r] + co[MAX_N_AMOM] * Xo.v[r];
  for (int a = MAX_N_AMOM - 2; a >= 0; --a) {
    const WM<T> x = wtn(wtr(Y, buf), hT);   // Y h^T
#pragma unroll
    for (int r = 0; r < 4; ++r) Y.v[r] = ce[a + 1] * Xe.v[r] + co[a + 1] * Xo.v[r] + x.v[r];
    Z = wtn(h, Z, Y);                        // Y + h^T Z
  }
  return Z;
}

// Reverse of the series start: R0_ad, T0_ad -> apb_ad, amb_ad, t_ad
template <class T>
__device__ void wmexp_start_ad(const WM<T>& apb, const WM<T>& amb, T t, const WM<T>& R0a, const WM<T>& T0a, int n,
                               T* buf, WM<T>& apb_ad, WM<T>& amb_ad, T& t_ad) {
  const T t2 = t * t;
  WM<T> hh = wnn(amb, apb, buf), ff = wnn(apb, amb, buf);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    hh.v[r] *= t2;
    ff.v[r] *= t2;
  }
  const WM<T> hT = wtr(hh, buf), fT = wtr(ff, buf);
  // S = s1 + s2 - t P3 - t P4, Sr = s1 - s2 + t P3 - t P4, P3 = apb s3, P4 = amb s4
  WM<T> Sad, Srad;
  {
    WM<T> S = wseries(hT, 0, n), Sr = S;
    const WM<T> s2 = wseries(fT, 0, n);
    const WM<T> p3 = wnn(apb, wseries(hT, 1, n), buf);
    const WM<T> p4 = wnn(amb, wseries(fT, 1, n), buf);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      S.v[r] += s2.v[r];
      Sr.v[r] -= s2.v[r];
      S.v[r] -= t * p3.v[r];
      Sr.v[r] += t * p3.v[r];
      S.v[r] -= t * p4.v[r];
      Sr.v[r] -= t * p4.v[r];
    }
    winv(S, n);                                   // Si
    // Si_ad = 2 T0_ad + Sr^T R0_ad;  Sr_ad = R0_ad Si^T;  S_ad = -Si^T Si_ad Si^T
    WM<T> x;
#pragma unroll
    for (int r = 0; r < 4; ++r) x.v[r] = T(2) * T0a.v[r];
    x = wtn(Sr, R0a, x);
    const WM<T> SiT = wtr(S, buf);
    Srad = wtn(wtr(R0a, buf), SiT);
    Sad = wtn(wtr(wtn(S, x), buf), SiT);
#pragma unroll
    for (int r = 0; r < 4; ++r) Sad.v[r] = -Sad.v[r];
  }
  WM<T> c3, c4;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const T a = Sad.v[r], b = Srad.v[r];
    c3.v[r] = b - a;
    c4.v[r] = -a - b;
    Sad.v[r] = a + b;    // s1_ad
    Srad.v[r] = a - b;   // s2_ad
  }
  // P3 = apb s3: t_ad += <P3, c3>; apb_ad = t c3 s3^T; s3_ad = t apb^T c3
  {
    const WM<T> s3 = wseries(hT, 1, n);
    t_ad += wdot(wnn(apb, s3, buf), c3);
    apb_ad = wtn(wtr(c3, buf), wtr(s3, buf));
    const WM<T> x = wtn(apb, c3);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      apb_ad.v[r] *= t;
      c3.v[r] = t * x.v[r];
    }
  }
  {
    const WM<T> s4 = wseries(fT, 1, n);
    t_ad += wdot(wnn(amb, s4, buf), c4);
    amb_ad = wtn(wtr(c4, buf), wtr(s4, buf));
    const WM<T> x = wtn(amb, c4);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      amb_ad.v[r] *= t;
      c4.v[r] = t * x.v[r];
    }
  }
  // series adjoints -> hh_ad, ff_ad; hh = HH t^2 with HH = amb apb, ff = FF t^2 with FF = apb amb
  const WM<T> apbT = wtr(apb, buf), ambT = wtr(amb, buf);
  {
    WM<T> x = wseries_pair_ad(hh, hT, Sad, c3, buf);
    t_ad += T(2) / t * wdot(hh, x);
#pragma unroll
    for (int r = 0; r < 4; ++r) x.v[r] *= t2;
    amb_ad = wtn(wtr(x, buf), apbT, amb_ad);   // += HH_ad apb^T
    apb_ad = wtn(amb, x, apb_ad);              // += amb^T HH_ad
  }
  {
    WM<T> x = wseries_pair_ad(ff, fT, Srad, c4, buf);
    t_ad += T(2) / t * wdot(ff, x);
#pragma unroll
    for (int r = 0; r < 4; ++r) x.v[r] *= t2;
    apb_ad = wtn(wtr(x, buf), ambT, apb_ad);   // += FF_ad amb^T
    amb_ad = wtn(apb, x, amb_ad);              // += apb^T FF_ad
  }
}

// Adjoint of wphase_forward: fa, da (adjoints of the final fs, fd; consumed)
// -> pl_ad[q], q < nl
template <class RT, int NZ>
__device__ void wphase_ad(int ns, int nl, const WAngles<RT>& a, const WPhase<RT, NZ>& pf, const WM<RT>& fs,
                          const WM<RT>& fd, WM<RT>& fa, WM<RT>& da, RT* buf, RT (&pl_ad)[NL_W - 1]) {
  const int j = wj(), g = wg();
  // the lower triangle is a copy of the upper one
  {
    const WM<RT> ft = wtr(fa, buf), dt = wtr(da, buf);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = WL<RT>::row(g, r);
      fa.v[r] = i < j ? fa.v[r] + ft.v[r] : i == j ? fa.v[r] : RT(0);
      da.v[r] = i < j ? da.v[r] + dt.v[r] : i == j ? da.v[r] : RT(0);
    }
  }
  RT S_ad[NZ];
#pragma unroll
  for (int i = 0; i < NZ; ++i) S_ad[i] = RT(0);
  if (ns < NZ) {
    const RT inf = RT(1) / pf.nfl;
    RT x = RT(0);
    if (j == NZ - 1) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        x += fa.v[r] * fs.v[r] + da.v[r] * fd.v[r];
        fa.v[r] *= inf;
        da.v[r] *= inf;
      }
    }
    const RT nfl_ad = -__shfl(wsum4(x), NZ - 1) * inf;
#pragma unroll
    for (int i = 0; i < NZ; ++i)
      if (i == ns - 1) S_ad[i] += nfl_ad;
  }
  // rows in reverse: fs_ij = raw_ij sc_i (j >= i), sc_i = (1 - S_{i-1}) / nf_i,
  // nf_i = sum_{j>=i} raw_fs_ij w_j, S_i = sum_{j<=i} fs_j,i+1 w_j
#pragma unroll
  for (int i = NZ - 1; i >= 0; --i) {
    if (i < ns) {
      if (i < NZ - 1 && j == i + 1) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (WL<RT>::row(g, r) <= i) fa.v[r] += S_ad[i] * a.wt_r[r];
      }
      const int gi = WL<RT>::grp(i), ri = WL<RT>::reg(i);
      const bool mine = g == gi && j >= i;
      const RT part = mine ? fa.v[ri] * pf.rs.v[ri] + da.v[ri] * pf.rd.v[ri] : RT(0);
      const RT sc_ad = __shfl(wsum16(part), 16 * gi);
      const RT nf_ad = -sc_ad * pf.sc[i] / pf.nf[i];
#pragma unroll
      for (int q = 0; q < NZ; ++q)
        if (i >= 1 && q == i - 1) S_ad[q] -= sc_ad / pf.nf[i];
      if (mine) {
        fa.v[ri] = fa.v[ri] * pf.sc[i] + nf_ad * a.wt_j;
        da.v[ri] = da.v[ri] * pf.sc[i];
      }
    }
  }
  // raw: fs = Pff + Pbb, fd = Pff - Pbb with Pff = e + o, Pbb = e - o (clipped)
  RT ea[4], oa[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = WL<RT>::row(g, r);
    const bool raw = i <= j && i < NZ && j < NZ && !(i >= ns && j != i);
    RT ffa = fa.v[r] + da.v[r], bba = fa.v[r] - da.v[r];
    if (pf.clip >> r & 1) ffa = RT(0);
    if (pf.clip >> (4 + r) & 1) bba = RT(0);
    ea[r] = raw ? ffa + bba : RT(0);
    oa[r] = raw ? ffa - bba : RT(0);
  }
  RT lj0 = RT(1), lj1 = a.mu_j, li0[4], li1[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    li0[r] = RT(1);
    li1[r] = a.mu_r[r];
  }
#pragma unroll
  for (int q = 0; q < NL_W - 1; ++q) {
    if (q >= nl) break;
    const RT ljq = q == 0 ? lj0 : lj1;
    RT part = RT(0);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const RT liq = q == 0 ? li0[r] : li1[r];
      part += ((q & 1) ? oa[r] : ea[r]) * liq * ljq;
    }
    pl_ad[q] += wsum64(part);
    if (q >= 1) {
      lj1 = (RT(2 * q + 1) * a.mu_j * lj1 - RT(q) * lj0) / RT(q + 1);
      lj0 = ljq;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const RT nx = (RT(2 * q + 1) * a.mu_r[r] * li1[r] - RT(q) * li0[r]) / RT(q + 1);
        li0[r] = li1[r];
        li1[r] = nx;
      }
    }
  }
}
}  // namespace

// =========================================================================
// forward + reverse adding sweep: one wavefront per (profile, channel)
// =========================================================================
// Tape of one wavefront: per layer 8 slots of 64 values ([k][slot][lane]):
// slots 0-3 the lane's R_up elements, 4-7 its I_up elements (Emission path:
// slots 0-3 hold the scalar tape of every lane).
constexpr int WTAPE_SLOTS = 8;

template <class RT, int NZ>
__global__ void __launch_bounds__(64) allsky_sweep_ad_wide_kernel(DevSensor s, DevBatch be, DevClouds cl,
                                                                  CloudWork cw, DevCloudCoeff cc, LayerScratch<RT> sc,
                                                                  LayerScratch<RT> sa, const RT* tau_g,
                                                                  const double* geo, DevOut o, AllskyOut ao, RT* tape,
                                                                  int kmax) {
  __shared__ double buf[16 * WLD];
  RT* rbuf = reinterpret_cast<RT*>(buf);
  const int j = wj(), g = wg();
  const int lane = threadIdx.x;
  const int64_t n = (int64_t)be.M * s.L;
  RT* wt_base = tape + (int64_t)blockIdx.x * kmax * WTAPE_SLOTS * 64;
  auto tp = [&](int k, int slot) -> RT& { return wt_base[((int64_t)k * WTAPE_SLOTS + slot) * 64 + lane]; };
  RT* tau_s = reinterpret_cast<RT*>(o.ks.tau);
  RT* b_s = reinterpret_cast<RT*>(o.ks.b);
  RT* t_s = reinterpret_cast<RT*>(o.ks.t);
  const bool lead = lane == 0;
  for (int64_t t = blockIdx.x; t < n; t += gridDim.x) {
    const int m = (int)(t / s.L), l = (int)(t - (int64_t)m * s.L);
    if (cw.cl_count[m] == 0 || cw.status[m] != 0 || cw.aer_err[m] != 0) continue;
    const double cosu = geo[(int64_t)m * 4 + 1];
    const int nl = legendre_terms(profile_streams(cl, m), cw.max_reff[m], s.wavenumber[l]);
    if (nl > NL_W - 1) continue;
    const AngleSet as = angle_set(cc.gauss, (nl + 2) / 2, cosu);
    if (as.nz != NZ) continue;
    const int K = be.n_layers[m];
    if (K > kmax) continue;   // the host sizes the tape for the largest cloudy profile
    const int64_t le = be.loff[m];
    const int lg = s.chan_offset + l;
    const int64_t ol = (int64_t)m * be.Ltot + lg;
    const int64_t sl = (int64_t)m * s.L + l;
    const int64_t st = sc.n_items;
    bool mine = false;
    for (int k = lane; k < K; k += 64) {
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
    auto planck_k = [&](int k) -> RT {
      RT te_;
      return planck(c1, c2, bc1, bc2, be.t[le + k], te_);
    };
    // optical-depth adjoint of layer k: the cloud-optics adjoint of a cloudy
    // layer, else the gas optical depth seed
    auto put_tau_ad = [&](int k, RT v) {
      if (!lead) return;
      const int c = cw.cl_slot[le + k];
      if (c >= 0) {
        const int64_t it = (int64_t)c * s.L + l;
        sa.v[st * sa.oTau() + it] = v;
        sa.v[st * sa.oSsa() + it] = RT(0);
#pragma unroll
        for (int q = 0; q < NL_W - 1; ++q) sa.v[st * (sa.oP() + q) + it] = RT(0);
      } else {
        tau_s[(le + k) * s.L + l] = v;
      }
      t_s[(le + k) * s.L + l] = RT(0);
    };
    const RT iu = (RT)(1.0 / cosu);
    const WAngles<RT> a = wangles<RT, NZ>(cc.gauss, as.ns, cosu);
    // ---------------- forward ----------------
    double r_cloudy;
    double up_rad = 0.0, down_rad = 0.0, sfc_b = 0.0;
    RT rdK = RT(0);
    if (!scat) {
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
      WM<RT> rup;
      WCol<RT> iup;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = WL<RT>::row(g, r);
        rup.v[r] = i == j && i < NZ ? RT(1) - eps : RT(0);
        iup.v[r] = i < NZ ? eps * Bs : RT(0);
      }
      for (int k = K - 1; k >= 0; --k) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          tp(k, r) = rup.v[r];
          tp(k, 4 + r) = iup.v[r];
        }
        const RT B = planck_k(k);
        const int c = cw.cl_slot[le + k];
        const int64_t it = (int64_t)c * s.L + l;
        if (c >= 0 && (double)sc.v[st * sc.oSsa() + it] > BS_THRESHOLD) {
          const RT* lay = sc.w + it * sc.w_stride();
          const WM<RT> rl = wload<RT>(lay, NZ), tl = wload<RT>(lay + NZ * NZ, NZ);
          WCol<RT> src = wload_col<RT>(lay + 2 * NZ * NZ, NZ);
#pragma unroll
          for (int r = 0; r < 4; ++r) src.v[r] *= B;
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
      const int is = as.isat;
      const RT top = wat(iup, is);
      const RT rs = __shfl(wsum16(wsel(rup.v, WL<RT>::reg(is))), 16 * WL<RT>::grp(is));
      r_cloudy = (double)(top + rs * cosmic);
    }
    const double rclr = o.radiance[ol];
    const bool frac = cw.frac[m] != 0;
    const double tcc = cw.tcc[m];
    const double R = frac ? (1.0 - tcc) * rclr + tcc * r_cloudy : r_cloudy;
    const double c1d = s.c1[l], c2d = s.c2[l], bc1d = s.bc1[l], bc2d = s.bc2[l];
    const double btv = (c2d / log(c1d / R + 1.0) - bc1d) / bc2d;
    // ---------------- adjoint seeds (Pre_Process_RTSolution_K) ----------------
    const double arg = c1d / R + 1.0;
    const double lga = log(arg);
    const double F = c1d * c2d / (bc2d * arg * (R * lga) * (R * lga));
    const double R_ad = (o.rad_k ? o.rad_k[ol] : 0.0) + F * (o.bt_k ? o.bt_k[ol] : 0.0);
    if (lead) {
      if (ao.r_clear) ao.r_clear[ol] = frac ? rclr : R;
      if (ao.tb_clear) ao.tb_clear[ol] = frac ? (c2d / log(c1d / rclr + 1.0) - bc1d) / bc2d : btv;
      if (ao.tcc) ao.tcc[ol] = tcc;
      o.radiance[ol] = R;
      o.bt[ol] = btv;
      if (o.up_radiance) o.up_radiance[ol] = up_rad;
      if (o.down_radiance) o.down_radiance[ol] = down_rad;
      if (o.sfc_planck) o.sfc_planck[ol] = sfc_b;
      if (o.sfc_emis) o.sfc_emis[ol] = be.emissivity[ol];
      if (o.rad_k) o.rad_k[ol] = R_ad;
      o.ks.rad[sl] = frac ? (1.0 - tcc) * R_ad : 0.0;
      o.ks.tcc_ad[sl] = frac ? (r_cloudy - rclr) * R_ad : 0.0;
    }
    const RT rc_ad = (RT)(frac ? tcc * R_ad : R_ad);
    RT eps_ad = RT(0), bs_ad = RT(0);
    // ---------------- reverse ----------------
    if (!scat) {
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
        if (lead) b_s[(le + k) * s.L + l] = b_ad;
        put_tau_ad(k, tau_ad);
      }
    } else {
      const int is = as.isat;
      WM<RT> rup_ad;
      WCol<RT> iup_ad;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const bool row = WL<RT>::row(g, r) == is;
        iup_ad.v[r] = row ? rc_ad : RT(0);
        rup_ad.v[r] = row && j < NZ ? rc_ad * cosmic : RT(0);
      }
      for (int k = 0; k < K; ++k) {
        WM<RT> rup;
        WCol<RT> iup;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          rup.v[r] = tp(k, r);
          iup.v[r] = tp(k, 4 + r);
        }
        const RT B = planck_k(k);
        const int c = cw.cl_slot[le + k];
        const int64_t it = (int64_t)c * s.L + l;
        if (c >= 0 && (double)sc.v[st * sc.oSsa() + it] > BS_THRESHOLD) {
          const RT* lay = sc.w + it * sc.w_stride();
          const WM<RT> rl = wload<RT>(lay, NZ), tl = wload<RT>(lay + NZ * NZ, NZ);
          const WCol<RT> e = wload_col<RT>(lay + 2 * NZ * NZ, NZ);
          WCol<RT> src;
#pragma unroll
          for (int r = 0; r < 4; ++r) src.v[r] = e.v[r] * B;
          WM<double> gm = wadd(weye<double>(NZ), wnn(wconv<double>(rup, buf), wconv<double>(rl, buf), buf), -1.0);
          winv(gm, NZ);
          const WM<double> tlD = wconv<double>(tl, buf);
          const WM<RT> gt = wconv<RT>(wnn(tlD, gm, buf), buf);
          WCol<RT> v = wmv(rup, wc2r(src));
#pragma unroll
          for (int r = 0; r < 4; ++r) v.v[r] += iup.v[r];
          // I_up' = src + gt v;  R_up' = rl + gt (rup tl)
          const RT vj = wc2r(v);
          WM<RT> gt_ad;
#pragma unroll
          for (int r = 0; r < 4; ++r) gt_ad.v[r] = iup_ad.v[r] * vj;
          const WM<RT> x = wnn(rup, tl, rbuf);                            // rt = rup tl
          gt_ad = wtn(wtr(rup_ad, rbuf), wtr(x, rbuf), gt_ad);            // += R'_ad rt^T
          const WM<RT> x2 = wtn(gt, rup_ad);                              // rt_ad = gt^T R'_ad
          WM<RT> tl_ad = wtn(rup, x2);                                    // tl_ad = rup^T rt_ad
          WM<RT> nrup_ad = wtn(wtr(x2, rbuf), wtr(tl, rbuf));             // rup_ad = rt_ad tl^T
          const RT v_ad = wmtv(gt, iup_ad);                               // gt^T iup_ad (row vector)
          const WCol<RT> v_adc = wr2c(v_ad);
          const RT srcj = wc2r(src);
#pragma unroll
          for (int r = 0; r < 4; ++r) nrup_ad.v[r] += v_adc.v[r] * srcj;
          const RT src_ad = wc2r(iup_ad) + wmtv(rup, v_adc);              // row vector
          // gt = tl g:  tl_ad += gt_ad g^T;  g_ad = tl^T gt_ad;  P_ad = g^T g_ad g^T (P = rup rl)
          const WM<double> gadD = wconv<double>(gt_ad, buf);
          const WM<double> gT = wtr(gm, buf);
          tl_ad = wadd(tl_ad, wconv<RT>(wtn(wtr(gadD, buf), gT), buf));
          const WM<double> pad = wtn(wtr(wtn(gm, wtn(tlD, gadD)), buf), gT);
          const WM<RT> y = wconv<RT>(pad, buf);                           // P_ad
          nrup_ad = wtn(wtr(y, rbuf), wtr(rl, rbuf), nrup_ad);            // rup_ad += P_ad rl^T
          rup_ad = wtn(rup, y, rup_ad);                                   // rl_ad = R'_ad + rup^T P_ad
          const RT ej = wc2r(e);
          const RT b_ad = wsum16(src_ad * ej);
          RT* lad = sa.w + it * sa.w_stride();
          wstore(lad, rup_ad, NZ);
          wstore(lad + NZ * NZ, tl_ad, NZ);
          WCol<RT> eadc = wr2c(src_ad);
#pragma unroll
          for (int r = 0; r < 4; ++r) eadc.v[r] *= B;
          wstore_col(lad + 2 * NZ * NZ, eadc, NZ);
          if (lead) {
            b_s[(le + k) * s.L + l] = b_ad;
            sa.v[st * sa.oTau() + it] = RT(0);
            sa.v[st * sa.oSsa() + it] = RT(0);
#pragma unroll
            for (int q = 0; q < NL_W - 1; ++q) sa.v[st * (sa.oP() + q) + it] = RT(0);
            t_s[(le + k) * s.L + l] = RT(0);
          }
          iup_ad = v_adc;
          rup_ad = nrup_ad;
        } else {
          const RT tau = layer_tau(k);
          const RT tr_j = j < NZ ? rexp<RT>(-tau * a.imu_j) : RT(0);
          const RT sr_j = j < NZ ? B * (RT(1) - tr_j) : RT(0);
          RT tr[4], sr[4];
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const bool in = WL<RT>::row(g, r) < NZ;
            tr[r] = in ? rexp<RT>(-tau * a.imu_r[r]) : RT(0);
            sr[r] = in ? B * (RT(1) - tr[r]) : RT(0);
          }
          const WCol<RT> w = wmv(rup, sr_j);
          WCol<RT> w_ad, t_adc;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            w_ad.v[r] = iup_ad.v[r] * tr[r];
            t_adc.v[r] = iup_ad.v[r] * (iup.v[r] + w.v[r]);
          }
          // s_ad[j] = iup_ad[j] + sum_i w_ad[i] rup[i][j]
          const RT s_ad = wc2r(iup_ad) + wmtv(rup, w_ad);
          // t_ad[i] += sum_j ra rup[i][j] tr[j]  (column)  and  t_ad[j] += sum_i ra tr[i] rup[i][j]  (row)
          RT colpart = RT(0);
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const RT ra = rup_ad.v[r];
            t_adc.v[r] += wsum16(ra * rup.v[r] * tr_j);
            colpart += ra * tr[r] * rup.v[r];
            rup_ad.v[r] = w_ad.v[r] * sr_j + ra * tr[r] * tr_j;
          }
          const RT t_adr = wsum4(colpart);
          // b_ad = sum_i s_ad[i] (1 - tr[i]);  tau_ad = -sum_i (t_ad[i] - s_ad[i] B) tr[i] imu[i]
          const RT b_ad = wsum16(j < NZ ? s_ad * (RT(1) - tr_j) : RT(0));
          RT tc = RT(0);
#pragma unroll
          for (int r = 0; r < 4; ++r) tc += t_adc.v[r] * tr[r] * a.imu_r[r];
          tc = wsum4(tc);   // column terms, replicated over j
          const RT tj = wsum16((t_adr - s_ad * B) * tr_j * a.imu_j);
          const RT tau_ad = -(tc + tj);
          if (lead) b_s[(le + k) * s.L + l] = b_ad;
          put_tau_ad(k, tau_ad);
          iup_ad = w_ad;
        }
      }
      RT dg = RT(0), su = RT(0);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (WL<RT>::row(g, r) == j) dg += rup_ad.v[r];
        su += iup_ad.v[r];
      }
      dg = wsum64(dg);
      su = wsum4(su);
      eps_ad += Bs * su - dg;
      bs_ad += eps * su;
    }
    if (lead) {
      o.ks.emis[sl] = (double)eps_ad;
      o.ks.bs[sl] = (double)bs_ad;
    }
  }
}

// =========================================================================
// layer R/T adjoint: one wavefront per scattering (layer, channel)
// =========================================================================
// Doubling tape of one wavefront: the state (R, T) entering doubling step d,
// d < WDCAP, as [d][matrix][r][lane]; steps beyond the cap are recomputed
// from the series start.
constexpr int WDCAP = 24;

template <class RT, int NZ>
__global__ void __launch_bounds__(64) cloud_rt_ad_wide_kernel(DevSensor s, CloudWork cw, DevClouds cl,
                                                              DevCloudCoeff cc, LayerScratch<RT> sc,
                                                              LayerScratch<RT> sa, const double* geo, int n_items,
                                                              RT* dtape) {
  __shared__ RT buf[16 * WLD];
  RT* dt = dtape + (int64_t)blockIdx.x * WDCAP * 2 * 4 * 64 + threadIdx.x;
  auto put = [&](int d, const WM<RT>& R_, const WM<RT>& T_) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      dt[((d * 2 + 0) * 4 + r) * 64] = R_.v[r];
      dt[((d * 2 + 1) * 4 + r) * 64] = T_.v[r];
    }
  };
  auto get = [&](int d, WM<RT>& R_, WM<RT>& T_) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      R_.v[r] = dt[((d * 2 + 0) * 4 + r) * 64];
      T_.v[r] = dt[((d * 2 + 1) * 4 + r) * 64];
    }
  };
  const int j = wj(), g = wg();
  const int n_live = live_items(cw, s.L, n_items);
  for (int item = blockIdx.x; item < n_live; item += gridDim.x) {
    const int c = item / s.L, l = item - c * s.L;
    const int m = cw.cl_desc[c].x;
    if (cw.status[m] != 0 || cw.aer_err[m] != 0) continue;
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
    for (int q = 0; q < NL_W - 1; ++q) plr[q] = sc.v[st * (sc.oP() + q) + item];
    const int ns = as.ns;
    const WAngles<RT> a = wangles<RT, NZ>(cc.gauss, ns, cosu);
    const RT* lad = sa.w + (int64_t)item * sa.w_stride();
    WM<RT> Ra = wload<RT>(lad, NZ), Ta = wload<RT>(lad + NZ * NZ, NZ);
    const WCol<RT> ea = wload_col<RT>(lad + 2 * NZ * NZ, NZ);
    WPhase<RT, NZ> pf;
    WM<RT> fs, fd;
    wphase_forward<RT, NZ>(ns, nl, plr, a, buf, pf, fs, fd);
    RT tau_ad = RT(0), om_ad = RT(0), pl_ad[NL_W - 1];
#pragma unroll
    for (int q = 0; q < NL_W - 1; ++q) pl_ad[q] = RT(0);
    WM<RT> fa, da;
    if ((double)tau < DELTA_OPTICAL_DEPTH) {
      // R = c Pbb w, T = c Pff w + I (1 - tau/mu), c = tau om / mu_i; e = 1 - sum_j (R + T)
      const RT sv = tau * om;
      RT s_part = RT(0), tdiag = RT(0);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = WL<RT>::row(g, r);
        const RT cst = sv / a.mu_r[r];
        const RT ra = Ra.v[r] - ea.v[r], ta = Ta.v[r] - ea.v[r];
        const RT pbb = RT(0.5) * (fs.v[r] - fd.v[r]), pff = RT(0.5) * (fs.v[r] + fd.v[r]);
        const bool in = i < NZ && j < NZ;
        const RT bba = in ? ra * cst * a.wt_j : RT(0), ffa = in ? ta * cst * a.wt_j : RT(0);
        fa.v[r] = RT(0.5) * (ffa + bba);
        da.v[r] = RT(0.5) * (ffa - bba);
        const RT c_ad = wsum16(in ? (ra * pbb + ta * pff) * a.wt_j : RT(0));
        if (j == 0 && i < NZ) s_part += c_ad / a.mu_r[r];
        if (i == j && i < NZ) tdiag -= ta / a.mu_r[r];
      }
      const RT s_ad = wsum64(s_part);
      tau_ad += wsum64(tdiag) + s_ad * om;
      om_ad += s_ad * tau;
    } else {
      // e_i = 1 - sum_{j<ns} (T_ij + R_ij) [- T_nn for the sensor row]
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = WL<RT>::row(g, r);
        if (j < ns) {
          Ra.v[r] -= ea.v[r];
          Ta.v[r] -= ea.v[r];
        }
        if (i == NZ - 1 && j == NZ - 1 && NZ == ns + 1) Ta.v[r] -= ea.v[r];
      }
      const RT sv = om < RT(MAX_ALBEDO) ? om : RT(MAX_ALBEDO);
      WM<RT> apb, amb;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = WL<RT>::row(g, r);
        const RT cst = sv / a.mu_r[r], dg = i == j && i < NZ ? RT(1) / a.mu_r[r] : RT(0);
        apb.v[r] = cst * fs.v[r] * a.wt_j - dg;
        amb.v[r] = cst * fd.v[r] * a.wt_j - dg;
      }
      const int nd = wdoublings(wnn(amb, apb, buf), tau);
      const RT t = (RT)((double)tau / ldexp(1.0, nd));
      if (nd > 0) {
        WM<RT> R0, T0;
        wmexp_start(apb, amb, t, NZ, buf, R0, T0);
        {   // tape the states entering the first WDCAP steps
          WM<RT> Rd = R0, Td = T0;
          const int nt = nd < WDCAP ? nd : WDCAP;
          for (int d = 0; d < nt; ++d) {
            put(d, Rd, Td);
            if (d + 1 < nt) wdouble_step(Rd, Td, NZ, buf);
          }
        }
        for (int d = nd - 1; d >= 0; --d) {
          WM<RT> Rd, Td;
          if (d < WDCAP) {
            get(d, Rd, Td);
          } else {
            get(WDCAP - 1, Rd, Td);
            for (int q = WDCAP - 1; q < d; ++q) wdouble_step(Rd, Td, NZ, buf);
          }
          wdouble_step_ad(Rd, Td, Ra, Ta, NZ, buf);
        }
      }
      RT t_ad = RT(0);
      wmexp_start_ad(apb, amb, t, Ra, Ta, NZ, buf, fa, da, t_ad);   // fa = apb_ad, da = amb_ad
      tau_ad += t_ad / (RT)ldexp(1.0, nd);
      // A +- B = c fs|fd w - I / mu with c = s / mu_i
      RT s_part = RT(0);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const RT cst = sv / a.mu_r[r];
        s_part += (fa.v[r] * fs.v[r] + da.v[r] * fd.v[r]) * a.wt_j / a.mu_r[r];
        fa.v[r] *= cst * a.wt_j;
        da.v[r] *= cst * a.wt_j;
      }
      const RT s_ad = wsum64(s_part);
      if (om < RT(MAX_ALBEDO)) om_ad += s_ad;
    }
    wphase_ad<RT, NZ>(ns, nl, a, pf, fs, fd, fa, da, buf, pl_ad);
    if (threadIdx.x == 0) {
      RT* ad = sa.v + item;
      ad[st * sa.oTau()] = tau_ad;
      ad[st * sa.oSsa()] = om_ad;
#pragma unroll
      for (int q = 0; q < NL_W - 1; ++q) ad[st * (sa.oP() + q)] = q < nl ? pl_ad[q] : RT(0);
    }
  }
}

// ------------------------------------------------------------------ launchers
static int wide_ad_blocks(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(n, 8192)); }

int allsky_sweep_ad_wide_blocks(int64_t n_pc) { return wide_ad_blocks(n_pc); }
// elements of the doubling tape of the layer adjoint for n_items items
int64_t wide_doubling_tape_elems(int64_t n_items) { return (int64_t)wide_ad_blocks(n_items) * WDCAP * 2 * 4 * 64; }

template <class RT>
void launch_allsky_ad_wide(int nz, const DevSensor& s, const DevBatch& be, const DevClouds& cl, const CloudWork& cw,
                           const DevCloudCoeff& cc, const LayerScratch<RT>& sc, const LayerScratch<RT>& sa,
                           const RT* tau_g, const double* geo, int n_items, const DevOut& o, const AllskyOut& ao,
                           RT* tape, int kmax, RT* dtape, hipStream_t st) {
  const int sb = wide_ad_blocks((int64_t)be.M * s.L);
  const int lb = wide_ad_blocks(n_items);
  switch (nz) {
#define CRTM_CASE(N)                                                                                               \
  case N:                                                                                                          \
    hipLaunchKernelGGL((allsky_sweep_ad_wide_kernel<RT, N>), dim3(sb), dim3(64), 0, st, s, be, cl, cw, cc, sc, sa, \
                       tau_g, geo, o, ao, tape, kmax);                                                             \
    if (n_items > 0)                                                                                               \
      hipLaunchKernelGGL((cloud_rt_ad_wide_kernel<RT, N>), dim3(lb), dim3(64), 0, st, s, cw, cl, cc, sc, sa, geo,  \
                         n_items, dtape);                                                                          \
    break;
    CRTM_CASE(6)
    CRTM_CASE(9)
    CRTM_CASE(10)
#undef CRTM_CASE
    default:
      break;
  }
}

template void launch_allsky_ad_wide<float>(int, const DevSensor&, const DevBatch&, const DevClouds&,
                                           const CloudWork&, const DevCloudCoeff&, const LayerScratch<float>&,
                                           const LayerScratch<float>&, const float*, const double*, int,
                                           const DevOut&, const AllskyOut&, float*, int, float*, hipStream_t);
template void launch_allsky_ad_wide<double>(int, const DevSensor&, const DevBatch&, const DevClouds&,
                                            const CloudWork&, const DevCloudCoeff&, const LayerScratch<double>&,
                                            const LayerScratch<double>&, const double*, const double*, int,
                                            const DevOut&, const AllskyOut&, double*, int, double*, hipStream_t);

}  // namespace crtm
