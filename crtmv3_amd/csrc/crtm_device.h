// crtm_device.h -- device-side data structures and ODPS group tables shared by
// the profile-preparation and the fused forward/K-matrix kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sfc_optics.h"

namespace crtm {

constexpr int WAVE = 64;           // CDNA wavefront width: one channel per lane
constexpr int KO_MAX = 128;        // ODPS internal layers supported
constexpr int MAX_N_LAYERS = 200;  // user layers (src/CRTM_Parameters.f90:64)
constexpr int N_MODEL_LEVELS = 101; // climatology levels (src/Atmosphere/CRTM_Model_Profiles.f90)
constexpr int N_EXTRA_MAX = N_MODEL_LEVELS - 1;   // layers AddLayers can put on top
constexpr int K_MAX = MAX_N_LAYERS + N_EXTRA_MAX; // layers of an extended profile
constexpr int NZ_MAX = 2048;       // LayerAvg weights stored per profile (all ODPS layers)
constexpr int MAP_WIN = 4;         // ODPS layers per user layer the K-matrix kernel's window holds
constexpr int NA_MAX = 6;          // user absorbers carried through the adjoint
constexpr int MAX_OPTRAN_ORDER = 10;
constexpr int N_OPTRAN_PRED = 14;
constexpr int OPZ = 15;            // OPTRAN predictor rows incl. constant row 0
constexpr int OPO = 11;            // OPTRAN polynomial orders 0..10
// Preparation status of a profile skipped by Options(m)%Skip_Profile: no
// kernel computes it, its outputs are zero and it does not fail the call
constexpr int ST_SKIP = 7;

// Per-profile non-LTE correction predictor (NLTE_Predictor_type,
// src/NLTE/NLTE_Predictor/NLTE_Predictor_Define.f90:72-94) as the
// preparation kernel leaves it for the fused kernel, FP64:
//   active, compute_tm, k1[2], k2[2] (0-based level indices of the mean-
//   temperature layers), predictors [3] (1, Tm1, Tm2), bilinear weights
//   w11, w12, w21, w22, isen, isol (0-based lower corners)
enum NLteRec { NL_ACTIVE = 0, NL_CTM, NL_K1, NL_K2 = NL_K1 + 2, NL_PRED = NL_K2 + 2, NL_WGT = NL_PRED + 3,
               NL_ISEN = NL_WGT + 4, NL_ISOL, NLTE_REC = 16 };
constexpr double TSPACE = 2.7253;   // K, src/CRTM_Parameters.f90:173

// Tile shapes of the fused clear-sky kernel: profiles (waves) per workgroup,
// each workgroup one 64-channel block.  Index 1 is the default.
constexpr int N_TILES = 3;
template <class RT> struct TileSet { static constexpr int pw[N_TILES] = {4, 8, 16}; };
template <> struct TileSet<double> { static constexpr int pw[N_TILES] = {2, 4, 8}; };

// Read-only view of data no kernel writes while it runs, through the
// constant address space: loads at wave-uniform addresses become scalar
// loads (scalar cache, counted apart from the vector-memory counter the
// staged layer loads use), the others stay vector loads.
template <class T> using cst_ptr = const __attribute__((address_space(4))) T*;
template <class T> __device__ inline cst_ptr<T> as_const(const T* p) { return (cst_ptr<T>)p; }

template <class RT> __device__ inline RT rexp(RT x);
template <> __device__ inline float rexp<float>(float x) { return __expf(x); }
template <> __device__ inline double rexp<double>(double x) { return exp(x); }

// Reciprocal: the hardware's 1-ulp v_rcp_f32 in FP32 (the FP32 path's error
// budget is far above an ulp), a correctly rounded division in FP64.
template <class RT> __device__ inline RT rrcp(RT x);
template <> __device__ inline float rrcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template <> __device__ inline double rrcp<double>(double x) { return 1.0 / x; }

// Planck radiance B = c1 / (exp(c2/Te) - 1), Te = bc1 + bc2*T
// (CRTM_Planck_Functions.f90:116-136).  In FP32, expm1 keeps the microwave
// case (c2/Te ~ 0.01) accurate; where c2/Te > 0.5 (infrared) exp(x) - 1
// loses at most a factor 2.5 to cancellation and the shorter expf serves.
__device__ inline float planck(float c1, float c2, float bc1, float bc2, double t, float& te) {
  te = bc1 + bc2 * (float)t;
  const float x = c2 * __builtin_amdgcn_rcpf(te);
  const float em1 = x > 0.5f ? expf(x) - 1.0f : expm1f(x);
  return c1 * __builtin_amdgcn_rcpf(em1);
}
__device__ inline double planck(double c1, double c2, double bc1, double bc2, double t, double& te) {
  te = bc1 + bc2 * t;
  return c1 / (exp(c2 / te) - 1.0);
}

// ODPS predictor groups (src/AtmAbsorption/ODPS/ODPS_Predictor.f90:85-117)
template <int G> struct GT;
template <> struct GT<1> {
  static constexpr int NCOMP = 8, NABS = 6, NX = 89;
  static constexpr int NP[8] = {7, 18, 7, 11, 11, 14, 10, 11};
  static constexpr int XOFF[8] = {0, 7, 25, 32, 43, 54, 68, 78};
};
template <> struct GT<2> {
  static constexpr int NCOMP = 5, NABS = 3, NX = 50;
  static constexpr int NP[5] = {7, 15, 7, 11, 10};
  static constexpr int XOFF[5] = {0, 7, 22, 29, 40};
};
template <> struct GT<3> {
  static constexpr int NCOMP = 2, NABS = 1, NX = 21;
  static constexpr int NP[2] = {7, 14};
  static constexpr int XOFF[2] = {0, 7};
};
constexpr int NX_MAX = 89;
constexpr int NCOMP_MAX = 8;

// Per-layer uniform table used by the hand-written predictor adjoint.
// Values are channel-independent and read as wave-uniform scalars.
enum UIdx {
  U_SEC = 0, U_T, U_IT, U_IT2, U_IT4, U_DT, U_ADT, U_DT2,
  U_H2O, U_H2OA, U_H2OR, U_H2OS, U_H2OR4, U_GT0, U_H2ODTZP, U_H2OP,
  U_O3, U_O3A, U_O3R, U_O3P, U_SP, U_SQSO3P, U_TZP, U_ST, U_CO2P,
  U_CO, U_N2O, U_CH4, U_N2OA, U_N2OR, U_COA, U_COR, U_COZP, U_COQ, U_CH4A, U_CH4R,
  U_CH4ZP, U_CH4Q, U_N2OZP, U_SQCOR, U_SQCH4R, U_N2OPOW, U_SQCOZP,
  U_IRT, U_PDP, U_ITZREF, U_ITZPREF, U_TODPS,
  // reciprocals so the adjoint chain needs no divisions
  U_IO3P, U_IN2OZP, U_ICOR, U_ISQCOR, U_ISQCOZP, U_ISQCH4R, U_ICH4ZP, U_IST, U_ICH4R, U_ICOZP,
  U_IN2OR, U_IO3R, U_IGT0, U_IH2OR4, U_IH2OR,
  U_IRA0,                 // 6 entries: 1/ref_absorber(j) (0 where the AD is killed)
  U_IGAZP0 = U_IRA0 + 6,  // 6 entries: 1/GAzp_ref(j)
  U_IGATZP0 = U_IGAZP0 + 6,
  U_AODPS0 = U_IGATZP0 + 6,
  U_MASK0 = U_AODPS0 + 6, // 6 entries: 1 if the absorber adjoint passes Map_Input_AD
  NU = U_MASK0 + 6
};

// Per-layer uniform OPTRAN table (ODPS_Compute_Predictor_ODAS and its AD)
enum OIdx {
  O_X0 = 0,               // OX(0..14), OX(0) = 1
  O_AP0 = O_X0 + OPZ,     // Ap(0..10), Ap(0) = 1
  O_DX_DT0 = O_AP0 + OPO, // dOX(p)/dT
  O_DX_DV0 = O_DX_DT0 + OPZ, // dOX(p)/dVapour
  O_DA = O_DX_DV0 + OPZ,
  O_IDEN, O_INV, O_ST, O_SP, O_IIV2, O_DPSEC, O_PR, O_T,
  NO = O_T + 1
};

// Sensor resident in HBM (built by crtm_init from the coefficient files)
struct DevSensor {
  int group, Ko, J, Jm, L;
  int n_blocks;
  int chan_offset;          // first channel of this sensor in the context channel list
  int h2o_idx;              // 1-based ODPS absorber index used for heights (reference quirk)
  int optran;               // OPTRAN block present
  int ocomp;                // OComponent_Index (1-based)
  double alpha, alpha_c1, alpha_c2;
  const double* ref_lp;     // [Ko+1]
  const double* ref_p;      // [Ko]
  const double* ref_t;      // [Ko]
  const double* ref_a;      // [Jm][Ko]
  const double* min_a;      // [Jm][Ko]
  const double* max_a;      // [Jm][Ko]
  const double* z_ref;      // [Ko+1] geopotential heights of the reference profile (Z(Ko)=0)
  const double* pdp;        // [Ko]
  const double* tz_ref;     // [Ko]
  const double* tzp_ref;    // [Ko]
  const double* gazp_ref;   // [Jm][Ko]
  const double* gatzp_ref;  // [Jm][Ko]
  const int* abs_id;        // [Jm]
  // coefficients: per 64-channel block, [k][slot][lane]
  const int64_t* blk_base;  // [n_blocks] float offset of the block
  const int* blk_meta;      // [n_blocks][BM_N]
  const float* cblk;
  // OPTRAN per block: [row][order][lane] doubles over the block's active
  // predictor rows (blk_meta BM_OROW0..) and orders 0..BM_OMAXO
  const int64_t* oblk_base; // [n_blocks], -1 when the block has no OPTRAN lane
  const double* ocd;
  const int* lane_optran;   // [n_blocks*64] 1 if the lane's WLO OD comes from OPTRAN
  const int* blk_order;     // [n_blocks] block processed at each position of the workgroup order
  // per channel (SpcCoeff)
  const double* c1;
  const double* c2;
  const double* bc1;
  const double* bc2;
  const double* cosmic;
  const double* frequency;  // GHz (microwave cloud tables)
  const double* wavenumber; // cm^-1 (infrared cloud tables, stream choice)
  int sensor_type;          // 1 microwave, 2 infrared, ...
  const int* polarization;  // SpcCoeff%Polarization
  const double* pol_angle;  // SpcCoeff%PolAngle, degrees
  // antenna correction (SpcCoeff%AC, src/Coefficients/ACCoeff/ACCoeff_Define.f90:78-96):
  // per field of view f and channel l, [f][l]: A_earth + A_platform, and A_space
  int ac_nfov;              // 0: the sensor has no antenna correction data
  const double* ac_ep;
  const double* ac_sp;
  // non-LTE correction (SpcCoeff%NC, src/Coefficients/NLTECoeff/NLTECoeff_Define.f90:87-121)
  int nl_nsen, nl_nsol, nl_nch;  // nl_nsen = 0: no NLTE correction data
  const double* nl_ssen;    // [nsen] secants of the sensor zenith angle
  const double* nl_ssol;    // [nsol] secants of the solar zenith angle
  const int* nl_cidx;       // [L] 1-based index of the channel's coefficients, 0: not an NLTE channel
  const double* nl_c;       // [nch][nsol][nsen][3] (the reference's C(3, nsen, nsol, nch), first index fastest)
  double nl_up[2], nl_lo[2], nl_min[2], nl_max[2], nl_mean[2];
};

// blk_meta layout
enum BMeta { BM_NSL = 0, BM_NP0 = 1, BM_OFF0 = BM_NP0 + NCOMP_MAX, BM_ONROW = BM_OFF0 + NCOMP_MAX,
             BM_OMAXO, BM_OROW0, BM_N = BM_OROW0 + OPZ };

// Batch of profiles (device pointers)
struct DevBatch {
  int M;
  int na;                   // user absorbers
  int max_k;
  const int* n_layers;      // [M]
  const int64_t* loff;      // [M+1]
  const double* level_p;
  const double* p;
  const double* t;
  const double* a;          // [na][NT]
  int64_t NT;               // total layers
  const double* cov[4];     // [M] each: land, water, snow, ice coverage
  const double* sfc_t[4];   // [M] each: land, water, snow, ice temperature
  const double* zenith;     // [M]
  const double* altitude;   // [M] or null
  const double* emissivity; // [M][Ltot]
  int Ltot;
  int jmap[NA_MAX];         // ODPS absorber j -> user absorber index (or -1)
  // AddLayers: profile arrays above describe the extended profiles; the
  // Jacobian outputs keep the caller's layout
  const int* climatology;   // [M] or null (input of the extension)
  const int* n_extra;       // [M] layers added on top of profile m
  const int64_t* out_loff;  // [M+1] caller's layer offsets (Jacobian rows)
  int64_t out_NT;           // caller's total layers (absorber Jacobian stride)
  const int* skip;          // [M] or null: Options(m)%Skip_Profile
  DevSfcIn sfc_in;          // computed surface optics: salinity, wind, azimuths
  // post-processing (Post_Process_RTSolution, src/CRTM_Forward_Module.f90:1166-1199)
  const double* source_zenith;  // [M] degrees, or null (100: no source)
  const int* ifov;              // [M] Geometry%iFOV (0: no antenna correction), or null
  int ac_on;                    // Options%Use_Antenna_Correction
  int nlte_on;                  // Options%Apply_NLTE_Correction
};

// Extended profiles written by the AddLayers kernel: profile m occupies
// layers [loff[m], loff[m] + n_layers[m]) with loff[m] = caller's offset +
// m * N_EXTRA_MAX (levels: loff[m] + m), the added layers first.
struct ExtBatch {
  int* n_layers;      // [M]
  int64_t* loff;      // [M+1]
  int* n_extra;       // [M]
  double* level_p;    // [NT + M]
  double* p;          // [NT]
  double* t;          // [NT]
  double* a;          // [na][NT]
  int64_t NT;         // caller's total layers + M * N_EXTRA_MAX
};

// Model climatology table on the device: level pressure [101], temperature
// [6][101], absorbers [6 models][6 ids: H2O, CO2, O3, N2O, CO, CH4][101]
constexpr int CLIM_T0 = N_MODEL_LEVELS;
constexpr int CLIM_A0 = CLIM_T0 + 6 * N_MODEL_LEVELS;
constexpr int CLIM_SIZE = CLIM_A0 + 36 * N_MODEL_LEVELS;

// Per-profile work produced by the preparation kernel, precision RT
template <class RT> struct ProfWork {
  RT* x;          // [M][KO_MAX? Ko][NX]
  RT* u;          // [M][Ko][NU]
  RT* o;          // [M][Ko][NO]
  int* acc_i;     // [M][Ko][3]: first user layer (0-based), count, offset into acc_w
  RT* acc_w;      // [M][NZ_MAX]: weights of ODPS layer k at [offset, offset + count)
  int* uidx;      // [M][max_k+1]: lower ODPS level k1 (0-based); k2 = k1+1 unless flagged
  RT* ufrac;      // [M][max_k+1]: interpolation fraction, or -1 when k1 == k2
  double* geo;    // [M][4]: sec_sfc, cos_sensor, sfc_T, widest Map_Input span (ODPS layers per user layer)
  RT* uwin;       // [M][max_k][MAP_WIN]: Map_Input weights of user layer i from ODPS layers kmin(i)+c
  int* frange;    // [M][Ko][2]: user layers whose last Map_Input contribution comes from ODPS layer k
  int* status;    // [M]
  double* accw64; // [M][NZ_MAX]: the LayerAvg weights in FP64 (mapping of T and absorbers)
  double* nlte;   // [M][NLTE_REC]: the profile's non-LTE correction predictor (NLteRec)
  double* lm;     // [3 * max_k + 1][M]: per-profile columns of the layer-mapping kernel
                  // (user layer lnP, user level lnP, LayerAvg accumulation column)
};

// All-sky K-matrix: adjoints the scattering path hands to the clear-sky
// K-matrix kernel for the profiles with cloudy layers (mask > 0).  Layer
// arrays are [NT (extended layout)][L of the sensor] in the call's precision,
// per-channel arrays [M][L of the sensor].
struct KSeeds {
  const int* mask;   // [M] cloudy-layer count of the profile
  void* tau;         // adjoint of the nadir gas optical depth of the layer
  void* b;           // adjoint of the layer Planck radiance
  void* t;           // direct temperature adjoint (temperature-dependent cloud tables)
  double* rad;       // adjoint of the clear-column radiance (fractional cover), else 0
  double* emis;      // surface emissivity adjoint of the cloudy RT
  double* bs;        // surface Planck radiance adjoint of the cloudy RT
  double* tcc_ad;    // adjoint of the total cloud cover
};

// Sum over the 64 lanes of a wavefront; every lane of the wave must be active
template <class T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
  return v;
}

struct DevOut {
  double* radiance;        // [M][Ltot]
  double* bt;
  double* up_radiance;
  double* down_radiance;
  double* sfc_planck;
  double* sfc_emis;
  // K-matrix
  const double* bt_k;      // [M][Ltot]
  double* rad_k;           // [M][Ltot] in/out
  double* emis_k;
  double* sfc_t_k[4];
  void* t_k;               // [NT][Ltot]; adjoint calls: double [NT], channel sums
  void* a_k;               // [na][NT][Ltot]; adjoint calls: double [na][NT]
                           // (adjoint calls: sfc_t_k[c] are double [M], channel sums)
  int k_f32;
  // all-sky: nadir gas OD per layer of the profiles with clouds,
  // [NT (extended layout)][L of the sensor], precision of the call
  void* tau_g;
  const int* tau_mask;     // [M]: write the profile's ODs when > 0
  KSeeds ks;               // all-sky K-matrix seeds (ks.mask null: none)
  const int* fail;         // [M] or null: nonzero fails the profile (scattering-path
                           // failures found before the clear-sky kernels run)
  // computed surface optics (sp.mode != SFC_USER): the sensor's planes,
  // RTSolution%Surface_Reflectivity, and the salinity / wind speed / wind
  // direction Jacobians ([M][Ltot]; adjoint calls: [M] channel sums)
  SfcPlanes sp;
  double* sfc_refl;
  double* sfc_k[3];
};

// Cloud Jacobians of the K-matrix, caller's layer layout, element type KT:
// water content and effective radius [nc][out_NT][Ltot], fraction [out_NT][Ltot].
// reduce (adjoint calls): double channel sums [nc][out_NT] / [out_NT],
// accumulated into with atomics
struct CloudKOut {
  void* wc_k;
  void* reff_k;
  void* cf_k;
  void* aconc_k;   // [na][out_NT][Ltot]
  void* areff_k;
  int reduce;
};

// ---------------------------------------------------------------------------
// All-sky (clouds, ADA)
// ---------------------------------------------------------------------------
constexpr int NZ_AS = 6;           // largest angle count of the per-lane kernels (8 streams)
constexpr int NL_AS = 9;           // Legendre terms 0..8 of the per-lane kernels
constexpr int NZ_W = 10;           // largest angle count of the wavefront kernels (16 streams)
constexpr int NL_W = 17;           // Legendre terms 0..16 of the optics and wavefront kernels

// AerosolCoeff tables, 'CRTM' scheme (src/Coefficients/AerosolCoeff/
// AerosolCoeff_Define.f90:100-128), column-major as in the file
struct DevAerosolCoeff {
  int nw, nr, ns, nt, nrh, nlf;   // wavelengths, radii, size variances, types, RH entries, Legendre terms
  const int* type_id;   // [nt]
  const double* freq;   // [nw] cm^-1 ascending
  const double* reff;   // (nr, nt)
  const double* ke;     // (nw, nrh, nr, ns, nt)
  const double* w;
  const double* pc;     // (nw, nrh, nr, ns, nt, 0:nlf, nP)
};

// CloudCoeff tables (src/Coefficients/CloudCoeff/CloudCoeff_Define.f90:95-154),
// column-major as in the file, phase element 1 only
struct DevCloudCoeff {
  int nfm, nrm, nt, ndm, nfi, nri, ndi, nlf;   // nlf: Legendre terms of the file
  const double* f_mw;
  const double* r_mw;
  const double* temp;
  const double* f_ir;
  const double* r_ir;
  const double* ke_l;   // (f, r, T)
  const double* w_l;
  const double* pc_l;   // (f, r, T, 0:nlf)
  const double* ke_s;   // (f, r, D)
  const double* w_s;
  const double* pc_s;   // (f, r, D, 0:nlf)
  const double* ke_i;   // (f, r, D+1)
  const double* w_i;
  const double* pc_i;   // (f, r, D+1, 1:nlf)
  const double* gauss;  // [10][2][10]: nodes and weights on [0,1] for n = 0..9 streams
  DevAerosolCoeff aero; // aerosol tables (aero.nt = 0: not loaded)
};

// Cloud and aerosol inputs of a batch (caller's layer layout)
struct DevClouds {
  int nc;               // cloud slots per profile
  const int* type;      // [M][nc]
  const double* wc;     // [nc][out_NT] kg/m^2
  const double* reff;   // [nc][out_NT] microns
  const double* cf;     // [out_NT] or null
  int n_streams;        // Options%n_Streams (0: chosen from the Mie parameter)
  const int* ns_prof;   // [M] per-profile Options(m)%n_Streams, or null (n_streams for all)
  unsigned nz_mask;     // host side: bit z set when angle count z occurs in the call
  int overlap;          // Options%Overlap_Id 1..5
  int na;               // aerosol slots per profile
  const int* atype;     // [M][na] aerosol type id (0: empty slot)
  const double* aconc;  // [na][out_NT] kg/m^2
  const double* areff;  // [na][out_NT] microns
  int no_scat;          // Options%Include_Scattering = .FALSE.: absorption only
};

// Per-call cloud work (extended layer layout)
struct CloudWork {
  double* wc;           // [nc][NT] cleaned, cover-scaled water content
  double* reff;         // [nc][NT]
  int* cl_count;        // [M] cloudy layers of profile m
  int* cl_base;         // [M] first slot of profile m (exclusive prefix sum of cl_count)
  int* cl_slot;         // [NT] cloudy-layer id of (m, k), -1 for clear layers
  int2* cl_desc;        // [cap] (m, k) of cloudy layer c
  int* total;           // [2] cloudy layers in the batch, largest extended layer count of a cloudy profile
  double* tcc;          // [M] total cloud cover of fractional profiles (else 0)
  int* frac;            // [M] 1: fractional coverage
  double* max_reff;     // [M] largest effective radius (stream choice)
  int* aer_err;         // [M] 1: an infrared channel's stream count has no aerosol phase function,
                        //     2: the profile's scattering layers exceed the layer capacity
  int cap;              // cloudy-layer slots
  int capmode;          // 1: launches are sized for cap, not for the counted layers
  const int* status;    // [M] profile status of the preparation kernel (0 = good)
};

// Items (cloudy layer x channel) of a scattering launch that hold a layer.  A
// launch sized for the slot capacity (capturable calls) stops at the count
// the preparation kernel left on the device; otherwise the two agree.
__device__ inline int live_items(const CloudWork& cw, int L, int n_items) {
  const int n = cw.total[0] < cw.cap ? cw.total[0] : cw.cap;
  const int64_t live = (int64_t)n * L;
  return live < n_items ? (int)live : n_items;
}

// Per-(cloudy layer, channel) layer results, structure of arrays
// [element][item], item = c * L + l
template <class RT> struct LayerScratch {
  RT* v;
  int64_t n_items;
  int nzc;              // matrix size the per-lane slots are laid out for (0: none)
  int npc;              // phase coefficients carried: NL_AS - 1 or NL_W - 1
  // element offsets: R [nzc*nzc], T [nzc*nzc], e = 1 - Thermal_C [nzc],
  // delta-scaled tau and ssa, phase coefficients P_0..P_{npc-1}
  __host__ __device__ int oR() const { return 0; }
  __host__ __device__ int oT() const { return nzc * nzc; }
  __host__ __device__ int oE() const { return 2 * nzc * nzc; }
  __host__ __device__ int oTau() const { return 2 * nzc * nzc + nzc; }
  __host__ __device__ int oSsa() const { return oTau() + 1; }
  __host__ __device__ int oP() const { return oTau() + 2; }
  __host__ __device__ int n_elem() const { return oP() + npc; }
  // wavefront kernels: per item R [nzw*nzw], T [nzw*nzw] row-major and
  // e [nzw], item-major (one item's block contiguous)
  RT* w;
  int nzw;              // angle count of the wavefront slots (0: none)
  __host__ __device__ int64_t w_stride() const { return 2 * (int64_t)nzw * nzw + nzw; }
};

struct AllskyOut {
  double* r_clear;      // [M][Ltot] optional
  double* tb_clear;
  double* tcc;
};

}  // namespace crtm
