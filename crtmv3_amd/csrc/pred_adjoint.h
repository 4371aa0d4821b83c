// pred_adjoint.h -- adjoint of the ODPS predictors of one layer
// (ODPS_Compute_Predictor_AD, src/AtmAbsorption/ODPS/ODPS_Predictor.f90:
// 1783-2643): the predictor adjoints X_AD(k, i, j) chained back to the
// layer's temperature, absorber ratios and integrated ratios.  Host- and
// device-callable and templated on the scalar type, so the same code runs on
// the GPU and, on a dependence-tracking type, at compile time to find which
// predictor reaches which final quantity.
#pragma once
#include "crtm_device.h"

#if defined(__HIPCC__)
#define CRTM_HD __host__ __device__
#else
#define CRTM_HD
#endif

namespace crtm {

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
CRTM_HD constexpr void comp_adjoint(const RT* xa, const RT* u, LayerAdj<RT>& a) {
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
CRTM_HD constexpr void finish_layer(const RT* u, LayerAdj<RT>& a) {
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


// Final per-layer quantities the sweep needs from the predictor adjoint, all
// linear in X_AD with per-layer factors: the local temperature adjoint
// (T and dT terms), then per absorber the ratio, GATzp and GAzp adjoints,
// then the Tzp and Tz adjoints.
template <int G> struct QIdx {
  static constexpr int NA = GT<G>::NABS;
  static constexpr int TT = 0, RA0 = 1, GT0 = 1 + NA, GZ0 = 1 + 2 * NA, TZP = 1 + 3 * NA, TZ = 2 + 3 * NA,
                       NQ = 3 + 3 * NA;
};

template <int G, class RT>
CRTM_HD constexpr void final_quantities(const LayerAdj<RT>& a, const RT* u, RT* q) {
  using Q = QIdx<G>;
  q[Q::TT] = a.T * u[U_IRT] + a.dt;
  for (int j = 0; j < Q::NA; ++j) {
    q[Q::RA0 + j] = a.ra[j] * u[U_IRA0 + j];
    q[Q::GT0 + j] = a.gt[j] * u[U_IGATZP0 + j];
    q[Q::GZ0 + j] = a.gz[j] * u[U_IGAZP0 + j];
  }
  q[Q::TZP] = a.tzp * u[U_ITZPREF];
  q[Q::TZ] = a.tz * u[U_ITZREF];
}

// Structural dependence: nonzero or not, with the arithmetic of a ring where
// every table value is nonzero.  Running the adjoint above on it with a unit
// predictor adjoint tells which final quantities that predictor reaches.
struct DepFlag {
  bool nz = false;
  constexpr DepFlag() = default;
  constexpr DepFlag(double v) : nz(v != 0.0) {}
  static constexpr DepFlag on() { DepFlag d; d.nz = true; return d; }
};
constexpr DepFlag operator+(DepFlag a, DepFlag b) { DepFlag d; d.nz = a.nz || b.nz; return d; }
constexpr DepFlag operator-(DepFlag a, DepFlag b) { return a + b; }
constexpr DepFlag operator*(DepFlag a, DepFlag b) { DepFlag d; d.nz = a.nz && b.nz; return d; }
constexpr DepFlag operator-(DepFlag a) { return a; }
constexpr DepFlag& operator+=(DepFlag& a, DepFlag b) { a = a + b; return a; }
constexpr DepFlag& operator-=(DepFlag& a, DepFlag b) { a = a + b; return a; }

// Sensitivity table of one ODPS layer: for predictor slot s (component J,
// predictor i at s = XOFF[J] + i), the values dq/dX(s) of the final
// quantities q it reaches, stored at off[s] ... off[s+1]-1 in increasing q.
// The channel kernel's predictor adjoint is then y[q] = sum_s C(s) dq/dX(s):
// one FMA per stored value instead of the chain through the intermediates.
template <int G> struct SensPattern {
  static constexpr int NX = GT<G>::NX, NQ = QIdx<G>::NQ;
  struct Tab {
    unsigned mask[NX];
    int off[NX + 1];
  };
  template <int J> static constexpr void fill(Tab& t) {
    if constexpr (J < GT<G>::NCOMP) {
      for (int i = 0; i < GT<G>::NP[J]; ++i) {
        DepFlag xa[18] = {};
        DepFlag u[NU] = {};
        for (int k = 0; k < NU; ++k) u[k] = DepFlag::on();
        xa[i] = DepFlag::on();
        LayerAdj<DepFlag> a{};
        comp_adjoint<G, J, DepFlag>(xa, u, a);
        finish_layer<G, DepFlag>(u, a);
        DepFlag q[NQ] = {};
        final_quantities<G, DepFlag>(a, u, q);
        unsigned m = 0;
        for (int c = 0; c < NQ; ++c)
          if (q[c].nz) m |= 1u << c;
        t.mask[GT<G>::XOFF[J] + i] = m;
      }
      fill<J + 1>(t);
    }
  }
  static constexpr Tab make() {
    Tab t{};
    fill<0>(t);
    t.off[0] = 0;
    for (int s = 0; s < NX; ++s) {
      int n = 0;
      for (int c = 0; c < NQ; ++c) n += (t.mask[s] >> c) & 1;
      t.off[s + 1] = t.off[s] + n;
    }
    return t;
  }
  static constexpr Tab tab = make();
  static constexpr int ND = tab.off[NX];
};

// Per-layer adjoint row of the K-matrix kernel: the sensitivity table, then
// the factors of the cumulative integrated-ratio chain and the Map_Input_AD
// masks (ODPS_Predictor.f90:707-759 reversed; ODPS_CoordinateMapping.f90:517).
template <int G> struct RRow {
  static constexpr int ND = SensPattern<G>::ND, NA = GT<G>::NABS;
  static constexpr int PDP = ND, AODPS0 = ND + 1, TODPS = ND + 1 + NA, MASK0 = ND + 2 + NA, NR = ND + 2 + 2 * NA;
};
constexpr int NR_MAX = 144;
static_assert(RRow<1>::NR <= NR_MAX && RRow<2>::NR <= NR_MAX && RRow<3>::NR <= NR_MAX, "adjoint row");

}  // namespace crtm
