// allsky_wide.h -- wavefront-wide small-matrix algebra for the scattering
// kernels at 16 streams (9 or 10 angles), where a per-lane matrix no longer
// fits in registers.
//
// One 64-lane wavefront owns one 16x16 matrix X, zero beyond the angle count,
// as four values per lane: lane j + 16 g holds X[row(g, r)][j] in v[r], the
// accumulator layout of the 16x16x4 MFMA instructions: row(g, r) = 4g + r for
// FP32 (v_mfma_f32_16x16x4f32) and g + 4r for FP64 (v_mfma_f64_16x16x4f64).
// Both make P^T Q one chain of four MFMAs with no data movement: in step s
// lane i + 16 kk feeds P[row(kk, s)][i] (= (P^T)[i][k]) as the A operand and
// lane j + 16 kk feeds Q[row(kk, s)][j] as the B operand, k = row(kk, s).
// Other products transpose an operand through LDS first; so does a change of
// precision, which changes the row order.
//
// Vectors come in two layouts: a row vector x[j] is one value per lane
// (replicated over g); a column vector x[row(g, r)] is four values per lane
// (replicated over j).  X y takes a row vector and gives a column vector;
// X^T y takes a column vector and gives a row vector.
//
// Kernels using this header run one wavefront per workgroup (64 threads), so
// __syncthreads() only orders the wavefront's own LDS accesses.
#pragma once
#include "allsky_common.h"

namespace crtm {
namespace {

typedef float wf4 __attribute__((ext_vector_type(4)));
typedef double wd4 __attribute__((ext_vector_type(4)));

constexpr int WLD = 17;   // LDS row stride of the transpose buffer (elements)

template <class T> struct WM { T v[4]; };    // 16x16 matrix, v[r] = X[row(g, r)][j]
template <class T> struct WCol { T v[4]; };  // column vector, v[r] = x[row(g, r)]

__device__ inline int wj() { return threadIdx.x & 15; }
__device__ inline int wg() { return (threadIdx.x >> 4) & 3; }

// row order of the accumulator layout: row(g, r), and the group / register
// holding row i
template <class T> struct WL;
template <> struct WL<float> {
  __device__ static constexpr int row(int g, int r) { return 4 * g + r; }
  __device__ static constexpr int grp(int i) { return i >> 2; }
  __device__ static constexpr int reg(int i) { return i & 3; }
};
template <> struct WL<double> {
  __device__ static constexpr int row(int g, int r) { return g + 4 * r; }
  __device__ static constexpr int grp(int i) { return i & 3; }
  __device__ static constexpr int reg(int i) { return i >> 2; }
};
// row of the lane's element r
template <class T> __device__ inline int wrow(int r) { return WL<T>::row(wg(), r); }

template <class T> __device__ inline WM<T> wzero() {
  WM<T> x;
#pragma unroll
  for (int r = 0; r < 4; ++r) x.v[r] = T(0);
  return x;
}
// identity on the leading n x n block
template <class T> __device__ inline WM<T> weye(int n) {
  WM<T> x;
  const int j = wj();
#pragma unroll
  for (int r = 0; r < 4; ++r) x.v[r] = (wrow<T>(r) == j && j < n) ? T(1) : T(0);
  return x;
}

// ------------------------------------------------------------- products
// C + P^T Q on the matrix cores
__device__ inline WM<float> wtn(const WM<float>& P, const WM<float>& Q, const WM<float>& C) {
  wf4 c = {C.v[0], C.v[1], C.v[2], C.v[3]};
#pragma unroll
  for (int s = 0; s < 4; ++s) c = __builtin_amdgcn_mfma_f32_16x16x4f32(P.v[s], Q.v[s], c, 0, 0, 0);
  WM<float> x;
#pragma unroll
  for (int r = 0; r < 4; ++r) x.v[r] = c[r];
  return x;
}
__device__ inline WM<double> wtn(const WM<double>& P, const WM<double>& Q, const WM<double>& C) {
  wd4 c = {C.v[0], C.v[1], C.v[2], C.v[3]};
#pragma unroll
  for (int s = 0; s < 4; ++s) c = __builtin_amdgcn_mfma_f64_16x16x4f64(P.v[s], Q.v[s], c, 0, 0, 0);
  WM<double> x;
#pragma unroll
  for (int r = 0; r < 4; ++r) x.v[r] = c[r];
  return x;
}
template <class T> __device__ inline WM<T> wtn(const WM<T>& P, const WM<T>& Q) { return wtn(P, Q, wzero<T>()); }

// X^T through the wavefront's LDS buffer (16 x WLD elements of T)
template <class T> __device__ inline WM<T> wtr(const WM<T>& X, T* buf) {
  const int j = wj();
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 4; ++r) buf[wrow<T>(r) * WLD + j] = X.v[r];
  __syncthreads();
  WM<T> y;
#pragma unroll
  for (int r = 0; r < 4; ++r) y.v[r] = buf[j * WLD + wrow<T>(r)];
  return y;
}
// the same matrix in another precision (and row order), through LDS
template <class D, class S> __device__ inline WM<D> wconv(const WM<S>& X, double* buf) {
  const int j = wj();
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 4; ++r) buf[wrow<S>(r) * WLD + j] = (double)X.v[r];
  __syncthreads();
  WM<D> y;
#pragma unroll
  for (int r = 0; r < 4; ++r) y.v[r] = (D)buf[wrow<D>(r) * WLD + j];
  return y;
}
template <class T> __device__ inline WM<T> wnn(const WM<T>& A, const WM<T>& B, T* buf) {   // A B
  return wtn(wtr(A, buf), B);
}
template <class T> __device__ inline WM<T> wnt(const WM<T>& A, const WM<T>& B, T* buf) {   // A B^T
  return wtn(wtr(A, buf), wtr(B, buf));
}
template <class T> __device__ inline WM<T> wadd(const WM<T>& a, const WM<T>& b, T sb = T(1)) {   // a + sb b
  WM<T> x;
#pragma unroll
  for (int r = 0; r < 4; ++r) x.v[r] = a.v[r] + sb * b.v[r];
  return x;
}

// ---------------------------------------------------- cross-lane moves
// DPP move within each 16-lane row (the lane's group g): CTRL is a DPP
// control word, e.g. quad_perm, row_half_mirror (0x141) or row_mirror (0x140).
// 64-bit values move as two 32-bit halves.
template <int CTRL> __device__ inline float wdpp(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xf, 0xf, false));
}
template <int CTRL> __device__ inline double wdpp(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
// element of lane `lane` (wavefront-uniform) to every lane
__device__ inline float wreadlane(float x, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), lane));
}
__device__ inline double wreadlane(double x, int lane) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)b, lane), hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// ------------------------------------------------------------ reductions
// Sum over j within the lane's group: the butterfly pairs lanes j^1, j^2, then
// the two quads of each half-row (half mirror) and the two half-rows (mirror),
// so every lane of a row adds the same operands in the same pairing and holds
// the identical sum.
template <class T> __device__ inline T wsum16(T x) {
  x += wdpp<0xB1>(x);    // quad_perm [1, 0, 3, 2]
  x += wdpp<0x4E>(x);    // quad_perm [2, 3, 0, 1]
  x += wdpp<0x141>(x);   // row_half_mirror
  x += wdpp<0x140>(x);   // row_mirror
  return x;
}
// Sum over g at the lane's j: permlane16_swap pairs lanes l and l ^ 16,
// permlane32_swap lanes l and l ^ 32 (the two results of each swap are the
// pair's two values), with no trip through the LDS crossbar.
__device__ inline float wsum4(float x) {
  const auto a = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(x), false, false);
  x = __uint_as_float(a[0]) + __uint_as_float(a[1]);
  const auto b = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x), false, false);
  return __uint_as_float(b[0]) + __uint_as_float(b[1]);
}
__device__ inline double wsum4(double x) {
  auto pair16 = [](double v, double& u, double& w) {
    const long long bits = __double_as_longlong(v);
    const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)bits, (unsigned)bits, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(bits >> 32), (unsigned)(bits >> 32), false, false);
    u = __longlong_as_double(((long long)hi[0] << 32) | (long long)lo[0]);
    w = __longlong_as_double(((long long)hi[1] << 32) | (long long)lo[1]);
  };
  auto pair32 = [](double v, double& u, double& w) {
    const long long bits = __double_as_longlong(v);
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)bits, (unsigned)bits, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(bits >> 32), (unsigned)(bits >> 32), false, false);
    u = __longlong_as_double(((long long)hi[0] << 32) | (long long)lo[0]);
    w = __longlong_as_double(((long long)hi[1] << 32) | (long long)lo[1]);
  };
  double u, w;
  pair16(x, u, w);
  x = u + w;
  pair32(x, u, w);
  return u + w;
}
template <class T> __device__ inline T wmax64(T x) {
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) x = fmax(x, __shfl_xor(x, m));
  return x;
}
template <class T> __device__ inline T wsum64(T x) { return wsum4(wsum16(x)); }
// sum of all elements of a .* b
template <class T> __device__ inline T wdot(const WM<T>& a, const WM<T>& b) {
  T s = T(0);
#pragma unroll
  for (int r = 0; r < 4; ++r) s += a.v[r] * b.v[r];
  return wsum64(s);
}

template <class T> __device__ inline T wsel(const T (&v)[4], int r) {
  return r == 0 ? v[0] : r == 1 ? v[1] : r == 2 ? v[2] : v[3];
}
// X[i][k] for wavefront-uniform i, k
template <class T> __device__ inline T wat(const WM<T>& X, int i, int k) {
  return wreadlane(wsel(X.v, WL<T>::reg(i)), k + 16 * WL<T>::grp(i));
}
// element i of a column vector, i wavefront-uniform
template <class T> __device__ inline T wat(const WCol<T>& x, int i) {
  return wreadlane(wsel(x.v, WL<T>::reg(i)), 16 * WL<T>::grp(i));
}
// column vector -> row vector and back
template <class T> __device__ inline T wc2r(const WCol<T>& x) {
  const int j = wj();
  return __shfl(wsel(x.v, WL<T>::reg(j)), j + 16 * WL<T>::grp(j));
}
template <class T> __device__ inline WCol<T> wr2c(T x) {
  WCol<T> y;
#pragma unroll
  for (int r = 0; r < 4; ++r) y.v[r] = __shfl(x, wrow<T>(r));
  return y;
}
// X y (y a row vector) -> column vector
template <class T> __device__ inline WCol<T> wmv(const WM<T>& X, T y) {
  WCol<T> z;
#pragma unroll
  for (int r = 0; r < 4; ++r) z.v[r] = wsum16(X.v[r] * y);
  return z;
}
// X^T y (y a column vector) -> row vector
template <class T> __device__ inline T wmtv(const WM<T>& X, const WCol<T>& y) {
  T s = T(0);
#pragma unroll
  for (int r = 0; r < 4; ++r) s += X.v[r] * y.v[r];
  return wsum4(s);
}

// In-place Gauss-Jordan inverse of the leading n x n block without pivoting,
// the arithmetic of the per-lane inv_inplace (every matrix inverted here is
// diagonally dominant).  Row k and column k are broadcast by shuffles: a
// variant with DPP row broadcasts for column k matched this one bit for bit in
// isolation but gave wrong adjoints inside the 6-angle FP64 layer kernels, so
// the inverse keeps the shuffles.
template <class T> __device__ inline void winv(WM<T>& b, int n) {
  const int j = wj(), g = wg();
  for (int k = 0; k < n; ++k) {
    const T rowk = __shfl(wsel(b.v, WL<T>::reg(k)), j + 16 * WL<T>::grp(k));   // b[k][j]
    T colk[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) colk[r] = __shfl(b.v[r], k + 16 * g);   // b[row(g, r)][k]
    const T d = T(1) / __shfl(rowk, k);
    const T c = rowk * d;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = wrow<T>(r);
      if (i == k) b.v[r] = j == k ? d : c;
      else b.v[r] = j == k ? -colk[r] * d : b.v[r] - colk[r] * c;
    }
  }
}

// ---------------------------------------------------------- global memory
// n x n row-major block at p (elements beyond n read as zero)
template <class T> __device__ inline WM<T> wload(const T* p, int n) {
  WM<T> x;
  const int j = wj();
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = wrow<T>(r);
    x.v[r] = (i < n && j < n) ? p[i * n + j] : T(0);
  }
  return x;
}
template <class T> __device__ inline void wstore(T* p, const WM<T>& x, int n) {
  const int j = wj();
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = wrow<T>(r);
    if (i < n && j < n) p[i * n + j] = x.v[r];
  }
}
template <class T> __device__ inline WCol<T> wload_col(const T* p, int n) {
  WCol<T> x;
#pragma unroll
  for (int r = 0; r < 4; ++r) x.v[r] = wrow<T>(r) < n ? p[wrow<T>(r)] : T(0);
  return x;
}
template <class T> __device__ inline void wstore_col(T* p, const WCol<T>& x, int n) {
  const int j = wj();
#pragma unroll
  for (int r = 0; r < 4; ++r)
    if (j == r && wrow<T>(r) < n) p[wrow<T>(r)] = x.v[r];
}

// ------------------------------------------------------------- Matrix_Exp
// Horner series s = I + h s / (2i (2i + odd)), i = MAX_N_AMOM .. 1, on the
// leading n x n block (ht = h^T)
template <class T> __device__ inline WM<T> wseries(const WM<T>& ht, int odd, int n) {
  const WM<T> I = weye<T>(n);
  WM<T> s = I;
  for (int it = MAX_N_AMOM; it >= 1; --it) {
    const T f = T(1) / T((2 * it) * (2 * it + (odd ? 1 : -1)));
    s = wadd(I, wtn(ht, s), f);
  }
  return s;
}

// Doubling count of Matrix_Exp from max|amb apb| (ADA_Module.f90:776-782)
template <class T> __device__ inline int wdoublings(const WM<T>& hh, T tau) {
  T mh = T(0);
#pragma unroll
  for (int r = 0; r < 4; ++r) mh = fmax(mh, fabs(hh.v[r]));
  mh = wmax64(mh);
  const double max_h = sqrt((double)mh) * (double)tau;
  const double lv = max_h > 0.0 ? log(max_h / 0.5) / log(2.0) : -2.0;
  return lv <= -1.0 ? 0 : (lv < 60.0 ? (int)trunc(lv) + 1 : 60);
}

// Series start of Matrix_Exp (ADA_Module.f90:783-820) at optical depth t:
// R0 = Sr S^-1, T0 = 2 S^-1
template <class T> __device__ inline void wmexp_start(const WM<T>& apb, const WM<T>& amb, T t, int n, T* buf,
                                                      WM<T>& R0, WM<T>& T0) {
  const T t2 = t * t;
  WM<T> hh = wnn(amb, apb, buf), ff = wnn(apb, amb, buf);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    hh.v[r] *= t2;
    ff.v[r] *= t2;
  }
  const WM<T> ht = wtr(hh, buf), ft = wtr(ff, buf);
  WM<T> u = wseries(ht, 0, n);                              // S1
  WM<T> v = wtn(wtr(apb, buf), wseries(ht, 1, n));          // (A+B) S3
  const WM<T> s2 = wseries(ft, 0, n);
  const WM<T> s4 = wtn(wtr(amb, buf), wseries(ft, 1, n));   // (A-B) S4
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const T s1 = u.v[r], s3 = t * v.v[r];
    u.v[r] = s1 - s3 + (s2.v[r] - t * s4.v[r]);
    v.v[r] = s1 + s3 - (s2.v[r] + t * s4.v[r]);
  }
  winv(u, n);
#pragma unroll
  for (int r = 0; r < 4; ++r) T0.v[r] = T(2) * u.v[r];
  R0 = wnn(v, u, buf);
}

// One doubling step: R' = R + T (I - R R)^-1 R T, T' = T (I - R R)^-1 T
template <class T> __device__ inline void wdouble_step(WM<T>& R, WM<T>& Tr, int n, T* buf) {
  WM<T> u = wadd(weye<T>(n), wnn(R, R, buf), T(-1));
  winv(u, n);
  const WM<T> v = wnn(Tr, u, buf);
  const WM<T> vt = wtr(v, buf);
  R = wadd(R, wnn(wtn(vt, R), Tr, buf));
  Tr = wtn(vt, Tr);
}

// ------------------------------------------------------ layer of the ADA
// Angles of the lane's column j and rows 4g + r: Gauss nodes, then the sensor
// angle; padding beyond NZ has mu = 1 and weight 0.
template <class RT> struct WAngles {
  RT mu_j, wt_j, imu_j;
  RT mu_r[4], wt_r[4], imu_r[4];
};
template <class RT, int NZ>
__device__ inline WAngles<RT> wangles(const double* gauss, int ns, double cosu) {
  const double* gmu = gauss + ns * 20;
  auto mu = [&](int i) -> RT { return i < ns ? (RT)gmu[i] : i < NZ ? (RT)cosu : RT(1); };
  auto wt = [&](int i) -> RT { return i < ns ? (RT)gmu[10 + i] : RT(0); };
  auto imu = [&](int i) -> RT { return i < NZ ? (RT)(1.0 / (i < ns ? gmu[i] : cosu)) : RT(0); };
  WAngles<RT> a;
  const int j = wj(), g = wg();
  a.mu_j = mu(j);
  a.wt_j = wt(j);
  a.imu_j = imu(j);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    a.mu_r[r] = mu(WL<RT>::row(g, r));
    a.wt_r[r] = wt(WL<RT>::row(g, r));
    a.imu_r[r] = imu(WL<RT>::row(g, r));
  }
  return a;
}

// Phase matrices fs = Pff + Pbb, fd = Pff - Pbb after Normalize_Phase (full,
// symmetric), with what the adjoint needs.
template <class RT, int NZ> struct WPhase {
  WM<RT> rs, rd;        // raw clipped sums (upper triangle)
  RT sc[NZ], nf[NZ];    // row scales and norms
  RT nfl;               // sensor-column divisor
  unsigned clip;        // bit r: Pff of the lane's element r clipped; bit 4 + r: Pbb
};

template <class RT, int NZ>
__device__ void wphase_forward(int ns, int nl, const RT (&plr)[NL_W - 1], const WAngles<RT>& a, RT* buf,
                               WPhase<RT, NZ>& pf, WM<RT>& fs, WM<RT>& fd) {
  const int j = wj(), g = wg();
  // Legendre polynomials of the lane's column and rows by the recurrence of
  // the per-lane kernels
  RT e[4], o[4], lj0 = RT(1), lj1 = a.mu_j, li0[4], li1[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    e[r] = o[r] = RT(0);
    li0[r] = RT(1);
    li1[r] = a.mu_r[r];
  }
#pragma unroll
  for (int q = 0; q < NL_W - 1; ++q) {
    if (q >= nl) break;
    const RT pq = plr[q];
    const RT ljq = q == 0 ? lj0 : lj1;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const RT liq = q == 0 ? li0[r] : li1[r];
      const RT tq = pq * liq * ljq;
      if (q & 1) o[r] += tq; else e[r] += tq;
    }
    if (q >= 1) {   // advance P_q -> P_{q+1}
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
  pf.clip = 0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = WL<RT>::row(g, r);
    RT ff = e[r] + o[r], bb = e[r] - o[r];
    if (ff < RT(0)) { ff = RT(PHASE_THRESHOLD); pf.clip |= 1u << r; }
    if (bb < RT(0)) { bb = RT(PHASE_THRESHOLD); pf.clip |= 16u << r; }
    const bool up = i < NZ && j < NZ && j >= i;
    if (!up) pf.clip &= ~(17u << r);
    fs.v[r] = up ? ff + bb : RT(0);
    fd.v[r] = up ? ff - bb : RT(0);
  }
  pf.rs = fs;
  pf.rd = fd;
  RT sum_prev = RT(0);
#pragma unroll
  for (int i = 0; i < NZ; ++i) {
    pf.sc[i] = RT(1);
    pf.nf[i] = RT(1);
    if (i < ns) {
      const int gi = WL<RT>::grp(i), ri = WL<RT>::reg(i);
      const bool mine = g == gi && j >= i;
      const RT nf = __shfl(wsum16(mine ? fs.v[ri] * a.wt_j : RT(0)), 16 * gi);
      const RT scale = (RT(1) - sum_prev) / nf;
      pf.nf[i] = nf;
      pf.sc[i] = scale;
      if (mine) {
        fs.v[ri] *= scale;
        fd.v[ri] *= scale;
      }
      RT si = RT(0);
      if (i < NZ - 1) {
        RT y = RT(0);
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (WL<RT>::row(g, r) <= i) y += fs.v[r] * a.wt_r[r];
        si = __shfl(wsum4(j == i + 1 ? y : RT(0)), i + 1);
      }
      sum_prev = si;
    }
  }
  pf.nfl = sum_prev;
  if (ns < NZ) {
    const RT inf = RT(1) / sum_prev;
    if (j == NZ - 1) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        fs.v[r] *= inf;
        fd.v[r] *= inf;
      }
    }
  }
  const WM<RT> ts = wtr(fs, buf), td = wtr(fd, buf);
#pragma unroll
  for (int r = 0; r < 4; ++r)
    if (WL<RT>::row(g, r) > j) {
      fs.v[r] = ts.v[r];
      fd.v[r] = td.v[r];
    }
}

// Layer R, T and e = 1 - Thermal_C from the phase matrices (CRTM_AMOM_layer)
template <class RT, int NZ>
__device__ void wamom_layer(int ns, RT tau, RT om, const WAngles<RT>& a, const WM<RT>& fs, const WM<RT>& fd, RT* buf,
                            WM<RT>& R, WM<RT>& Tm, WCol<RT>& e) {
  const int j = wj(), g = wg();
  if ((double)tau < DELTA_OPTICAL_DEPTH) {   // single-scattering initialisation
    const RT sv = tau * om;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = WL<RT>::row(g, r);
      const RT cst = sv / a.mu_r[r];
      R.v[r] = cst * RT(0.5) * (fs.v[r] - fd.v[r]) * a.wt_j;
      Tm.v[r] = cst * RT(0.5) * (fs.v[r] + fd.v[r]) * a.wt_j + (i == j && i < NZ ? RT(1) - tau / a.mu_r[r] : RT(0));
      const RT th = wsum16(R.v[r] + Tm.v[r]);
      e.v[r] = i < NZ ? RT(1) - th : RT(0);
    }
    return;
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
  wmexp_start(apb, amb, t, NZ, buf, R, Tm);
  for (int d = 0; d < nd; ++d) wdouble_step(R, Tm, NZ, buf);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = WL<RT>::row(g, r);
    RT x = j < ns ? Tm.v[r] + R.v[r] : RT(0);
    if (i == NZ - 1 && j == NZ - 1 && NZ == ns + 1) x += Tm.v[r];
    const RT th = wsum16(x);
    e.v[r] = i < NZ ? RT(1) - th : RT(0);
  }
}

}  // namespace
}  // namespace crtm
