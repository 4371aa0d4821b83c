// crtm_abi.cpp -- host side of the MI355X CRTM hot path: coefficient file
// loaders (CRTM binary formats), the device coefficient layout, the device
// memory pool, and the C-ABI declared in include/crtm_mi355x.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstdint>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/crtm_mi355x.h"
#include "crtm_device.h"
#include "pred_adjoint.h"
#include "model_profiles.inc"   // CLIM_TABLE[CLIM_SIZE], generated from the model climatologies

namespace crtm {
void launch_add_layers(const DevBatch&, const ExtBatch&, const double* clim, const int* ids, hipStream_t);
template <int G, class RT>
void launch_prep(const DevSensor&, const DevBatch&, const ProfWork<RT>&, hipStream_t);
template <int G, class RT, class KT, bool DOK, bool RED>
void launch_fwdk(const DevSensor&, const DevBatch&, const ProfWork<RT>&, const DevOut&, hipStream_t, int tile);
void launch_cloud_prep(const DevBatch&, const DevClouds&, const CloudWork&, hipStream_t);
template <class RT>
void launch_allsky(const DevSensor&, const DevBatch&, const DevClouds&, const CloudWork&, const DevCloudCoeff&,
                   const LayerScratch<RT>&, const RT*, const double*, int, const int*, int, const DevOut&,
                   const AllskyOut&, hipStream_t);
void launch_clear_copy(const DevSensor&, int M, int Ltot, const int* cl_count, const DevOut&, const AllskyOut&,
                       hipStream_t);
template <class RT, class KT>
void launch_allsky_k(const DevSensor&, const DevBatch&, const DevClouds&, const CloudWork&, const DevCloudCoeff&,
                     const LayerScratch<RT>&, const LayerScratch<RT>&, const RT*, const double*, int, const int*, int,
                     const DevOut&, const AllskyOut&, const CloudKOut&, RT*, int64_t, int, RT*, int, RT*,
                     hipStream_t);
int allsky_sweep_ad_blocks(int64_t n_lanes);
int allsky_sweep_ad_wide_blocks(int64_t n_pc);
int64_t wide_doubling_tape_elems(int64_t n_items);
void launch_aerosol_check(const DevSensor&, int M, const DevClouds&, const CloudWork&, const DevCloudCoeff&,
                          hipStream_t);
void launch_status_merge(int M, const int* prep_status, const int* aer_err, int* call_status, hipStream_t st);
void launch_sfc_optics(const DevSensor& s, const DevBatch& b, const DevSfcIn& in, int mode, const DevMWwater& mw,
                       const DevIRwater& ir, double* planes, int64_t stride, int* err, hipStream_t st);
void launch_zero(void* p, size_t bytes, hipStream_t st);
}  // namespace crtm

using namespace crtm;

namespace {

struct CrtmError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

std::string fmt(const char* f, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, f);
  vsnprintf(buf, sizeof buf, f, ap);
  va_end(ap);
  return buf;
}

#define HIPCHECK(x)                                                                   \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess) throw CrtmError(fmt("%s: %s", #x, hipGetErrorString(e_)));  \
  } while (0)

// ---------------------------------------------------------------------------
// Fortran sequential-unformatted record reader (magic number first record,
// src/Utility/Binary_File_Utility.f90:56, 235)
// ---------------------------------------------------------------------------
class Records {
 public:
  explicit Records(const std::string& path) : path_(path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw CrtmError("File " + path + " not found.");
    buf_.assign(std::istreambuf_iterator<char>(f), {});
    auto r = next();
    int32_t magic;
    if (r.size() != 4) throw CrtmError("Error reading magic number from " + path);
    std::memcpy(&magic, r.data(), 4);
    if (magic != 123456789) throw CrtmError("Invalid magic number in " + path);
  }
  bool more() const { return pos_ < buf_.size(); }
  std::vector<char> next() {
    if (pos_ + 4 > buf_.size()) throw CrtmError("Unexpected end of file " + path_);
    int32_t n;
    std::memcpy(&n, &buf_[pos_], 4);
    if (n < 0 || pos_ + 8 + (size_t)n > buf_.size()) throw CrtmError("Corrupt record in " + path_);
    std::vector<char> r(buf_.begin() + pos_ + 4, buf_.begin() + pos_ + 4 + n);
    int32_t n2;
    std::memcpy(&n2, &buf_[pos_ + 4 + n], 4);
    if (n2 != n) throw CrtmError("Corrupt record framing in " + path_);
    pos_ += 8 + n;
    return r;
  }

 private:
  std::string path_;
  std::vector<char> buf_;
  size_t pos_ = 0;
};

class Cursor {
 public:
  explicit Cursor(std::vector<char> r) : r_(std::move(r)) {}
  template <class T> T get() {
    need(sizeof(T));
    T v;
    std::memcpy(&v, &r_[p_], sizeof(T));
    p_ += sizeof(T);
    return v;
  }
  template <class T> std::vector<T> arr(size_t n) {
    need(n * sizeof(T));
    std::vector<T> v(n);
    if (n) std::memcpy(v.data(), &r_[p_], n * sizeof(T));
    p_ += n * sizeof(T);
    return v;
  }
  std::string str(size_t n) {
    need(n);
    std::string s(&r_[p_], n);
    p_ += n;
    while (!s.empty() && (s.back() == ' ' || s.back() == '\0')) s.pop_back();
    return s;
  }
  void done(const char* what) {
    if (p_ != r_.size()) throw CrtmError(fmt("record length mismatch reading %s", what));
  }

 private:
  void need(size_t n) {
    if (p_ + n > r_.size()) throw CrtmError("record too short");
  }
  std::vector<char> r_;
  size_t p_ = 0;
};

struct SpcCoeffH {
  std::string sensor_id;
  int sensor_type = 0, n_channels = 0, wmo_sat = 0, wmo_sen = 0;
  std::vector<int> sensor_channel, polarization;
  std::vector<double> pol_angle, frequency, wavenumber, c1, c2, bc1, bc2, cosmic, solar;
  // antenna correction (ACCoeff, [fov + n_fov * channel]); n_fov = 0: absent
  int ac_nfov = 0;
  std::vector<double> ac_e, ac_s, ac_p;
  // non-LTE correction (NLTECoeff); nl_nsen = 0: absent
  int nl_npred = 0, nl_nsen = 0, nl_nsol = 0, nl_nch = 0;
  std::vector<double> nl_up, nl_lo, nl_min, nl_max, nl_mean, nl_ssen, nl_ssol, nl_c;
  std::vector<int> nl_cidx;
};

// the sensor identification record of an ACCoeff / NLTECoeff block must
// repeat the SpcCoeff's (SpcCoeff_Binary_IO.f90:486-495, 516-525)
void check_block_sensor(const SpcCoeffH& sc, const std::string& id, int wsat, int wsen, const std::vector<int>& ch,
                        const char* what) {
  if (id != sc.sensor_id || wsat != sc.wmo_sat || wsen != sc.wmo_sen || ch != sc.sensor_channel)
    throw CrtmError(fmt("%s sensor information is inconsistent with SpcCoeff", what));
}

// ACCoeff_Binary_ReadFile (src/Coefficients/ACCoeff/ACCoeff_Binary_IO.f90), as
// embedded in a SpcCoeff file
void read_ac_block(Records& rec, SpcCoeffH& sc) {
  Cursor r0(rec.next());
  const int rel = r0.get<int32_t>();
  r0.get<int32_t>();
  r0.done("ACCoeff release");
  if (rel != 1) throw CrtmError(fmt("ACCoeff Release check failed (release %d)", rel));
  Cursor r1(rec.next());
  const int n = r1.get<int32_t>(), L = r1.get<int32_t>();
  r1.done("ACCoeff dimensions");
  if (n < 1 || L != sc.n_channels) throw CrtmError("ACCoeff dimensions inconsistent with SpcCoeff");
  Cursor r2(rec.next());
  const std::string id = r2.str(20);
  const int wsat = r2.get<int32_t>(), wsen = r2.get<int32_t>();
  r2.done("ACCoeff sensor ids");
  Cursor r3(rec.next());
  const std::vector<int> ch = r3.arr<int32_t>(L);
  r3.done("ACCoeff sensor channels");
  check_block_sensor(sc, id, wsat, wsen, ch, "Antenna correction");
  Cursor r4(rec.next());
  sc.ac_e = r4.arr<double>((size_t)n * L);
  sc.ac_s = r4.arr<double>((size_t)n * L);
  sc.ac_p = r4.arr<double>((size_t)n * L);
  r4.done("ACCoeff coefficients");
  sc.ac_nfov = n;
}

// NLTECoeff_Binary_ReadFile (src/Coefficients/NLTECoeff/NLTECoeff_Binary_IO.f90),
// as embedded in a SpcCoeff file
void read_nc_block(Records& rec, SpcCoeffH& sc) {
  Cursor r0(rec.next());
  const int rel = r0.get<int32_t>();
  r0.get<int32_t>();
  r0.done("NLTECoeff release");
  if (rel != 2) throw CrtmError(fmt("NLTECoeff Release check failed (release %d)", rel));
  Cursor r1(rec.next());
  const int n1 = r1.get<int32_t>(), n2 = r1.get<int32_t>(), n3 = r1.get<int32_t>(), n4 = r1.get<int32_t>(),
            L = r1.get<int32_t>();
  r1.done("NLTECoeff dimensions");
  if (n1 != 3 || n2 < 2 || n3 < 2 || n4 < 0 || L != sc.n_channels)
    throw CrtmError("NLTECoeff dimensions inconsistent (3 predictors, at least two angles per axis)");
  Cursor r2(rec.next());
  const std::string id = r2.str(20);
  const int wsat = r2.get<int32_t>(), wsen = r2.get<int32_t>();
  const std::vector<int> ch = r2.arr<int32_t>(L);
  r2.done("NLTECoeff sensor ids");
  check_block_sensor(sc, id, wsat, wsen, ch, "non-LTE correction");
  Cursor r3(rec.next());
  sc.nl_up = r3.arr<double>(2);
  sc.nl_lo = r3.arr<double>(2);
  r3.done("NLTECoeff pressure levels");
  Cursor r4(rec.next());
  sc.nl_min = r4.arr<double>(2);
  sc.nl_max = r4.arr<double>(2);
  sc.nl_mean = r4.arr<double>(2);
  r4.done("NLTECoeff temperatures");
  Cursor r5(rec.next());
  r5.arr<int32_t>(n4);   // NLTE_Channel
  r5.done("NLTECoeff channels");
  Cursor r6(rec.next());
  sc.nl_ssen = r6.arr<double>(n2);
  sc.nl_ssol = r6.arr<double>(n3);
  r6.done("NLTECoeff secants");
  Cursor r7(rec.next());
  sc.nl_cidx = r7.arr<int32_t>(L);
  sc.nl_c = r7.arr<double>((size_t)n1 * n2 * n3 * n4);
  r7.done("NLTECoeff coefficients");
  for (int c : sc.nl_cidx)
    if (c < 0 || c > n4) throw CrtmError("NLTECoeff C_Index out of range");
  sc.nl_npred = n1;
  sc.nl_nsen = n2;
  sc.nl_nsol = n3;
  sc.nl_nch = n4;
}

// SpcCoeff_Binary_IO.f90:374-470
SpcCoeffH read_spccoeff(const std::string& path) {
  Records rec(path);
  SpcCoeffH sc;
  Cursor r0(rec.next());
  int rel = r0.get<int32_t>(), ver = r0.get<int32_t>();
  r0.done("SpcCoeff release");
  if (rel != 8) throw CrtmError(fmt("SpcCoeff Release check failed for %s (release %d)", path.c_str(), rel));
  Cursor r1(rec.next());
  const int n = r1.get<int32_t>();
  r1.done("SpcCoeff dimensions");
  sc.n_channels = n;
  Cursor r2(rec.next());
  sc.sensor_id = r2.str(20);
  sc.sensor_type = r2.get<int32_t>();
  sc.wmo_sat = r2.get<int32_t>();
  sc.wmo_sen = r2.get<int32_t>();
  r2.done("SpcCoeff sensor ids");
  if (sc.sensor_type > 100) sc.sensor_type -= 100;   // ACTIVE_SENSOR offset
  Cursor r3(rec.next());
  sc.sensor_channel = r3.arr<int32_t>(n);
  sc.polarization = r3.arr<int32_t>(n);
  sc.pol_angle = ver > 2 ? r3.arr<double>(n) : std::vector<double>(n, 0.0);   // PolAngle from version 3
  r3.arr<int32_t>(n);                    // channel flag
  sc.frequency = r3.arr<double>(n);
  sc.wavenumber = r3.arr<double>(n);
  sc.c1 = r3.arr<double>(n);
  sc.c2 = r3.arr<double>(n);
  sc.bc1 = r3.arr<double>(n);
  sc.bc2 = r3.arr<double>(n);
  sc.cosmic = r3.arr<double>(n);
  sc.solar = r3.arr<double>(n);
  r3.done("SpcCoeff channel data");
  // antenna- and NLTE-correction blocks behind their DATA_PRESENT flags
  // (SpcCoeff_Binary_IO.f90:466-527)
  if (rec.more()) {
    Cursor fa(rec.next());
    const int ac_present = fa.get<int32_t>();
    fa.done("antenna correction data indicator");
    if (ac_present == 1) read_ac_block(rec, sc);
    Cursor fn(rec.next());
    const int nc_present = fn.get<int32_t>();
    fn.done("NLTE correction data indicator");
    if (nc_present == 1) read_nc_block(rec, sc);
  }
  return sc;
}

struct OdpsH {
  int Ko = 0, J = 0, Jm = 0, L = 0, n_coeffs = 0, n_opindex = 0, n_ocoeffs = 0, group = 0, sensor_type = 0;
  std::string sensor_id;
  std::vector<int> sensor_channel, component_id, absorber_id, n_pred, pos;
  std::vector<double> ref_lp, ref_p, ref_t, ref_a, min_a, max_a;
  std::vector<float> c;
  std::vector<int> osig, order, op_index, opos;
  std::vector<double> oc;
  double alpha = 0, alpha_c1 = 0, alpha_c2 = 0;
  int ocomp = -1;
};

// Read_ODPS_Data, ODPS_Binary_IO.f90:617-800
OdpsH read_odps(const std::string& path) {
  Records rec(path);
  OdpsH t;
  {
    Cursor r(rec.next());
    int rel = r.get<int32_t>();
    r.get<int32_t>();
    r.done("ODPS release");
    if (rel != 2) throw CrtmError(fmt("ODPS Release check failed for %s (release %d)", path.c_str(), rel));
  }
  {
    Cursor r(rec.next());
    int alg = r.get<int32_t>();
    r.done("ODPS algorithm");
    if (alg != 2) throw CrtmError(fmt("%s is not an ODPS TauCoeff file (algorithm %d)", path.c_str(), alg));
  }
  {
    Cursor r(rec.next());
    t.Ko = r.get<int32_t>();
    t.J = r.get<int32_t>();
    t.Jm = r.get<int32_t>();
    t.L = r.get<int32_t>();
    t.n_coeffs = r.get<int32_t>();
    t.n_opindex = r.get<int32_t>();
    t.n_ocoeffs = r.get<int32_t>();
    r.done("ODPS dimensions");
  }
  {
    Cursor r(rec.next());
    t.group = r.get<int32_t>();
    r.done("ODPS group");
  }
  {
    Cursor r(rec.next());
    t.sensor_id = r.str(20);
    r.get<int32_t>();
    r.get<int32_t>();
    t.sensor_type = r.get<int32_t>();
    r.done("ODPS sensor info");
  }
  {
    Cursor r(rec.next());
    t.sensor_channel = r.arr<int32_t>(t.L);
    r.done("ODPS channels");
  }
  {
    Cursor r(rec.next());
    t.component_id = r.arr<int32_t>(t.J);
    r.done("ODPS component ids");
  }
  {
    Cursor r(rec.next());
    t.absorber_id = r.arr<int32_t>(t.Jm);
    r.done("ODPS absorber ids");
  }
  {
    Cursor r(rec.next());
    t.ref_lp = r.arr<double>(t.Ko + 1);
    t.ref_p = r.arr<double>(t.Ko);
    t.ref_t = r.arr<double>(t.Ko);
    t.ref_a = r.arr<double>((size_t)t.Ko * t.Jm);
    t.min_a = r.arr<double>((size_t)t.Ko * t.Jm);
    t.max_a = r.arr<double>((size_t)t.Ko * t.Jm);
    r.done("ODPS reference profiles");
  }
  {
    Cursor r(rec.next());
    t.n_pred = r.arr<int32_t>((size_t)t.J * t.L);
    t.pos = r.arr<int32_t>((size_t)t.J * t.L);
    r.done("ODPS n_Predictors/Pos_Index");
  }
  if (t.n_coeffs > 0) {
    Cursor r(rec.next());
    t.c = r.arr<float>(t.n_coeffs);
    r.done("ODPS coefficients");
  }
  if (t.n_ocoeffs > 0) {
    Cursor r(rec.next());
    t.osig = r.arr<int32_t>(t.L);
    t.order = r.arr<int32_t>(t.L);
    t.op_index = r.arr<int32_t>((size_t)(t.n_opindex + 1) * t.L);
    t.opos = r.arr<int32_t>(t.L);
    t.oc = r.arr<double>(t.n_ocoeffs);
    t.alpha = r.get<double>();
    t.alpha_c1 = r.get<double>();
    t.alpha_c2 = r.get<double>();
    t.ocomp = r.get<int32_t>();
    r.done("ODPS OPTRAN block");
  }
  return t;
}

// ---------------------------------------------------------------------------
// Device memory
// ---------------------------------------------------------------------------
struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  void reserve(size_t bytes) {
    if (bytes <= n) return;
    if (p) HIPCHECK(hipFree(p));
    p = nullptr;
    n = 0;
    HIPCHECK(hipMalloc(&p, bytes));
    n = bytes;
  }
  template <class T> T* as() const { return reinterpret_cast<T*>(p); }
};

template <class T> void upload(DevBuf& b, const std::vector<T>& v) {
  b.reserve(std::max<size_t>(v.size() * sizeof(T), 16));
  if (!v.empty()) HIPCHECK(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
}

struct Sensor {
  OdpsH tc;
  SpcCoeffH sc;
  DevSensor d{};
  int n_blocks = 0;
  int index = 0;               // position in the context's sensor list
  int tile[2] = {1, 1};        // tile index of the fused clear-sky kernel: [FP64, FP32-split]
  template <class RT> int tile_of() const { return tile[sizeof(RT) == 4 ? 1 : 0]; }
  std::vector<std::unique_ptr<DevBuf>> bufs;
  template <class T> const T* put(const std::vector<T>& v) {
    bufs.emplace_back(new DevBuf);
    upload(*bufs.back(), v);
    return bufs.back()->as<T>();
  }
};

const double R_DRYAIR_H = 8.314472 / (28.9648 * 1.0e-3);
const double CCH_H = 0.001 * R_DRYAIR_H / 9.80665;
const double CV_H = (1.0 / (18.01528 / 28.9648) - 1.0) / 1000.0;

void build_sensor(Sensor& S, int chan_offset) {
  const OdpsH& t = S.tc;
  const SpcCoeffH& sc = S.sc;
  if (t.L != sc.n_channels)
    throw CrtmError(fmt("%s: TauCoeff has %d channels, SpcCoeff %d", t.sensor_id.c_str(), t.L, sc.n_channels));
  if (t.group < 1 || t.group > 3) throw CrtmError(fmt("unsupported ODPS group %d", t.group));
  static const int NCOMP[4] = {0, 8, 5, 2}, NABS[4] = {0, 6, 3, 1};
  static const int NPMAX[4][8] = {{0}, {7, 18, 7, 11, 11, 14, 10, 11}, {7, 15, 7, 11, 10}, {7, 14}};
  if (t.J != NCOMP[t.group] || t.Jm != NABS[t.group])
    throw CrtmError(fmt("ODPS group %d expects %d components / %d absorbers, file has %d / %d", t.group,
                        NCOMP[t.group], NABS[t.group], t.J, t.Jm));
  if (t.Ko > KO_MAX) throw CrtmError(fmt("ODPS layers %d exceed %d", t.Ko, KO_MAX));
  const int Ko = t.Ko, J = t.J, Jm = t.Jm, L = t.L;
  DevSensor& d = S.d;
  d.group = t.group;
  d.Ko = Ko;
  d.J = J;
  d.Jm = Jm;
  d.L = L;
  d.chan_offset = chan_offset;
  // Map_Input resets the H2O index on every absorber pass (ODPS_CoordinateMapping.f90:253-258)
  int h2o_idx = 1;
  for (int j = 0; j < Jm; ++j) {
    h2o_idx = 1;
    if (t.absorber_id[j] == 1) h2o_idx = j + 1;
  }
  d.h2o_idx = h2o_idx;
  // Geopotential_Height of the reference profile (ODPS_CoordinateMapping.f90:607-634)
  std::vector<double> z(Ko + 1);
  z[Ko] = 0.0;
  for (int k = Ko; k >= 1; --k) {
    double tv = t.ref_t[k - 1] * (1.0 + CV_H * t.ref_a[(size_t)(h2o_idx - 1) * Ko + k - 1]);
    z[k - 1] = z[k] + CCH_H * tv * std::log(t.ref_lp[k] / t.ref_lp[k - 1]);
  }
  // integrated reference sums (ODPS_Predictor.f90:707-759)
  std::vector<double> pdp(Ko), tzr(Ko), tzpr(Ko), gazpr((size_t)Jm * Ko), gatzpr((size_t)Jm * Ko);
  {
    double a = 0, b = 0;
    std::vector<double> g1(Jm, 0.0), g2(Jm, 0.0);
    for (int k = 1; k <= Ko; ++k) {
      double p = (k == 1) ? t.ref_lp[0] * (t.ref_lp[1] - t.ref_lp[0]) : t.ref_lp[k] * (t.ref_lp[k] - t.ref_lp[k - 1]);
      pdp[k - 1] = p;
      a += t.ref_t[k - 1];
      b += p * t.ref_t[k - 1];
      tzr[k - 1] = a;
      tzpr[k - 1] = b;
      for (int j = 0; j < Jm; ++j) {
        double ra = t.ref_a[(size_t)j * Ko + k - 1];
        g1[j] += p * ra;
        g2[j] += p * ra * t.ref_t[k - 1];
        gazpr[(size_t)j * Ko + k - 1] = g1[j];
        gatzpr[(size_t)j * Ko + k - 1] = g2[j];
      }
    }
  }
  // coefficient block layout: [block][k][slot][lane] with the group's fixed
  // slot positions (component j at XOFF[j]); slots a channel does not use
  // hold zeros, so the kernel multiplies them without masks
  int xoff[NCOMP_MAX] = {0};
  for (int j = 1; j < J; ++j) xoff[j] = xoff[j - 1] + NPMAX[t.group][j - 1];
  const int nslots = xoff[J - 1] + NPMAX[t.group][J - 1];
  const int nb = (L + WAVE - 1) / WAVE;
  S.n_blocks = nb;
  d.n_blocks = nb;
  const bool optran = t.n_ocoeffs > 0;
  const int ocj = t.ocomp - 1;
  std::vector<int> lane_opt((size_t)nb * WAVE, 0);
  if (optran)
    for (int l = 0; l < L; ++l)
      lane_opt[l] = (t.osig[l] == 1 && ocj >= 0 && ocj < J && t.n_pred[ocj + (size_t)J * l] > 0 &&
                     t.op_index[(size_t)(t.n_opindex + 1) * l] > 0);
  std::vector<int> meta((size_t)nb * BM_N, 0);
  std::vector<int64_t> base(nb), obase(nb, -1);
  int64_t total = 0, ototal = 0;
  for (int b = 0; b < nb; ++b) {
    int* mt = &meta[(size_t)b * BM_N];
    int nsl = 0;
    for (int j = 0; j < J; ++j) {
      int mx = 0;
      for (int l = b * WAVE; l < std::min(L, (b + 1) * WAVE); ++l) {
        int np = t.n_pred[j + (size_t)J * l];
        if (lane_opt[l] && j == ocj) np = 0;
        if (np > NPMAX[t.group][j])
          throw CrtmError(fmt("channel %d component %d has %d predictors (max %d)", l + 1, j + 1, np,
                              NPMAX[t.group][j]));
        mx = std::max(mx, np);
      }
      mt[BM_NP0 + j] = mx;
      mt[BM_OFF0 + j] = xoff[j];
    }
    nsl = nslots;
    mt[BM_NSL] = nsl;
    base[b] = total;
    total += (int64_t)Ko * nsl * WAVE;
    bool any = false;
    int omax = 0, pmask = 0;
    for (int l = b * WAVE; l < std::min(L, (b + 1) * WAVE); ++l)
      if (lane_opt[l]) {
        any = true;
        omax = std::max(omax, t.order[l]);
        pmask |= 1;
        const int np = t.op_index[(size_t)(t.n_opindex + 1) * l];
        for (int i = 1; i <= np; ++i) {
          const int p = t.op_index[(size_t)(t.n_opindex + 1) * l + i];
          if (p < 1 || p > N_OPTRAN_PRED) throw CrtmError(fmt("channel %d: OPTRAN predictor index %d", l + 1, p));
          pmask |= 1 << p;
        }
      }
    if (any) {
      if (omax > MAX_OPTRAN_ORDER) throw CrtmError("OPTRAN order exceeds 10");
      int nrow = 0;
      for (int p = 0; p < OPZ; ++p)
        if (pmask >> p & 1) mt[BM_OROW0 + nrow++] = p;
      mt[BM_ONROW] = nrow;
      mt[BM_OMAXO] = omax;
      obase[b] = ototal;
      ototal += (int64_t)nrow * (omax + 1) * WAVE;
    }
  }
  // padded so a wave may read a component's full predictor range past the
  // last slot of the last block
  std::vector<float> cblk(total + (int64_t)NX_MAX * WAVE, 0.0f);
  for (int b = 0; b < nb; ++b) {
    const int* mt = &meta[(size_t)b * BM_N];
    const int nsl = mt[BM_NSL];
    for (int l = b * WAVE; l < std::min(L, (b + 1) * WAVE); ++l) {
      const int lane = l - b * WAVE;
      for (int j = 0; j < J; ++j) {
        if (lane_opt[l] && j == ocj) continue;
        const int np = t.n_pred[j + (size_t)J * l];
        const int64_t p0 = t.pos[j + (size_t)J * l] - 1;   // 1-based Pos_Index
        for (int i = 0; i < np; ++i)
          for (int k = 0; k < Ko; ++k) {
            const int64_t src = p0 + (int64_t)i * Ko + k;
            if (src < 0 || src >= t.n_coeffs) throw CrtmError("Pos_Index out of range");
            cblk[base[b] + ((int64_t)k * nsl + mt[BM_OFF0 + j] + i) * WAVE + lane] = t.c[src];
          }
      }
    }
  }
  std::vector<double> ocd(std::max<int64_t>(ototal, 1), 0.0);
  if (optran)
    for (int l = 0; l < L; ++l) {
      if (!lane_opt[l]) continue;
      const int b = l / WAVE, lane = l % WAVE;
      const int* mt = &meta[(size_t)b * BM_N];
      const int nq = mt[BM_OMAXO] + 1;
      const int np = t.op_index[(size_t)(t.n_opindex + 1) * l];
      const int no = t.order[l];
      for (int i = 0; i <= np; ++i) {
        const int p = i == 0 ? 0 : t.op_index[(size_t)(t.n_opindex + 1) * l + i];
        int r = 0;
        while (mt[BM_OROW0 + r] != p) ++r;
        const int64_t js = t.opos[l] - 1 + (int64_t)i * (no + 1);
        for (int q = 0; q <= no; ++q) {
          if (js + q < 0 || js + q >= t.n_ocoeffs) throw CrtmError("OPos_Index out of range");
          ocd[obase[b] + ((int64_t)r * nq + q) * WAVE + lane] += t.oc[js + q];
        }
      }
    }
  d.optran = optran ? 1 : 0;
  d.ocomp = t.ocomp;
  d.alpha = t.alpha;
  d.alpha_c1 = t.alpha_c1;
  d.alpha_c2 = t.alpha_c2;
  d.ref_lp = S.put(t.ref_lp);
  d.ref_p = S.put(t.ref_p);
  d.ref_t = S.put(t.ref_t);
  d.ref_a = S.put(t.ref_a);
  d.min_a = S.put(t.min_a);
  d.max_a = S.put(t.max_a);
  d.z_ref = S.put(z);
  d.pdp = S.put(pdp);
  d.tz_ref = S.put(tzr);
  d.tzp_ref = S.put(tzpr);
  d.gazp_ref = S.put(gazpr);
  d.gatzp_ref = S.put(gatzpr);
  d.abs_id = S.put(t.absorber_id);
  d.blk_base = S.put(base);
  d.blk_meta = S.put(meta);
  d.cblk = S.put(cblk);
  d.oblk_base = S.put(obase);
  d.ocd = S.put(ocd);
  d.lane_optran = S.put(lane_opt);
  // Workgroup order over blocks: heavy and light blocks alternate, so every
  // XCD's contiguous range of the order (the kernel's XCD mapping) carries
  // a similar share of the work.  Cost ~ predictor slots in use, with the
  // OPTRAN blocks' extra regression and adjoint work.
  {
    std::vector<std::pair<double, int>> cost(nb);
    for (int b = 0; b < nb; ++b) {
      const int* mt = &meta[(size_t)b * BM_N];
      double c = 20.0;
      for (int j = 0; j < J; ++j) c += mt[BM_NP0 + j];
      if (obase[b] >= 0) c *= 1.5;
      cost[b] = {c, b};
    }
    std::sort(cost.begin(), cost.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
    std::vector<int> order;
    for (int lo = 0, hi = nb - 1; lo <= hi; ++lo, --hi) {
      order.push_back(cost[lo].second);
      if (lo != hi) order.push_back(cost[hi].second);
    }
    d.blk_order = S.put(order);
  }
  d.c1 = S.put(sc.c1);
  d.c2 = S.put(sc.c2);
  d.bc1 = S.put(sc.bc1);
  d.bc2 = S.put(sc.bc2);
  d.cosmic = S.put(sc.cosmic);
  d.frequency = S.put(sc.frequency);
  d.wavenumber = S.put(sc.wavenumber);
  d.sensor_type = sc.sensor_type;
  d.polarization = S.put(sc.polarization);
  d.pol_angle = S.put(sc.pol_angle);
  // antenna correction: [fov][channel] A_earth + A_platform and A_space
  d.ac_nfov = sc.ac_nfov;
  if (sc.ac_nfov > 0) {
    const int n = sc.ac_nfov, L = sc.n_channels;
    std::vector<double> ep((size_t)n * L), sp((size_t)n * L);
    for (int l = 0; l < L; ++l)
      for (int f = 0; f < n; ++f) {
        const size_t src = f + (size_t)n * l, dst = (size_t)f * L + l;
        ep[dst] = sc.ac_e[src] + sc.ac_p[src];
        sp[dst] = sc.ac_s[src];
      }
    d.ac_ep = S.put(ep);
    d.ac_sp = S.put(sp);
  }
  d.nl_nsen = sc.nl_nsen;
  d.nl_nsol = sc.nl_nsol;
  d.nl_nch = sc.nl_nch;
  if (sc.nl_nsen > 0) {
    d.nl_ssen = S.put(sc.nl_ssen);
    d.nl_ssol = S.put(sc.nl_ssol);
    d.nl_cidx = S.put(sc.nl_cidx);
    d.nl_c = S.put(sc.nl_c.empty() ? std::vector<double>(1, 0.0) : sc.nl_c);
    for (int i = 0; i < 2; ++i) {
      d.nl_up[i] = sc.nl_up[i];
      d.nl_lo[i] = sc.nl_lo[i];
      d.nl_min[i] = sc.nl_min[i];
      d.nl_max[i] = sc.nl_max[i];
      d.nl_mean[i] = sc.nl_mean[i];
    }
  }
}

// Global attributes of the binary coefficient files: per attribute a length
// record and, for a non-zero length, the text (ReadGAtts_Binary_File,
// src/Utility/Binary_File_Utility.f90)
void skip_gatts(Records& rec) {
  for (int i = 0; i < 5; ++i) {
    Cursor c(rec.next());
    const int32_t n = c.get<int32_t>();
    c.done("global attribute length");
    if (n > 0) rec.next();
  }
}

// FASTEM coefficient file (MWwaterCoeff_ReadFile, src/Coefficients/EmisCoeff/
// MW_Water/MWwaterCoeff_Define.f90:699-922): release/version, attributes, the
// FC, FR, RC, AZ, SSC, LSC FitCoeff blocks (FitCoeff_ReadFile.inc), LUT flag
struct MWwaterH {
  int version = 0;
  std::vector<double> fit[6];
};
MWwaterH read_mwwater(const std::string& path) {
  Records rec(path);
  MWwaterH h;
  Cursor r0(rec.next());
  const int rel = r0.get<int32_t>();
  h.version = r0.get<int32_t>();
  r0.done("MWwaterCoeff release");
  if (rel != 1) throw CrtmError(fmt("MWwaterCoeff Release check failed for %s (release %d)", path.c_str(), rel));
  skip_gatts(rec);
  // expected shapes: fc >= 2, fr >= 6, rc (3,7,2), az (6,6,2), ssc >= 8, lsc (3,6,2)
  static const int rank[6] = {1, 1, 3, 3, 1, 3};
  static const int need[6][3] = {{2, 0, 0}, {6, 0, 0}, {3, 7, 2}, {6, 6, 2}, {8, 0, 0}, {3, 6, 2}};
  for (int b = 0; b < 6; ++b) {
    Cursor c0(rec.next());
    const int frel = c0.get<int32_t>();
    c0.get<int32_t>();
    c0.done("FitCoeff release");
    if (frel != 1) throw CrtmError(fmt("FitCoeff Release check failed in %s", path.c_str()));
    Cursor c1(rec.next());
    const int nd = c1.get<int32_t>();
    c1.done("FitCoeff rank");
    if (nd != rank[b]) throw CrtmError(fmt("%s: FitCoeff block %d has rank %d", path.c_str(), b + 1, nd));
    Cursor c2(rec.next());
    std::vector<int32_t> dims = c2.arr<int32_t>(nd);
    c2.done("FitCoeff dimensions");
    size_t n = 1;
    for (int k = 0; k < nd; ++k) {
      n *= (size_t)dims[k];
      if ((nd == 1 && dims[k] < need[b][0]) || (nd == 3 && dims[k] != need[b][k]))
        throw CrtmError(fmt("%s: FitCoeff block %d has unexpected dimensions", path.c_str(), b + 1));
    }
    skip_gatts(rec);
    Cursor c3(rec.next());
    h.fit[b] = c3.arr<double>(n);
    c3.done("FitCoeff coefficients");
  }
  Cursor r9(rec.next());
  if (r9.get<int32_t>() != 0) throw CrtmError(path + ": FASTEM look-up tables are not supported");
  if (h.version != 6) throw CrtmError(fmt("%s: MWwaterCoeff version %d; the FASTEM6 model (version 6) is supported",
                                          path.c_str(), h.version));
  return h;
}

// IR sea-surface emissivity table (IRwaterCoeff_ReadFile,
// src/Coefficients/EmisCoeff/IR_Water/IRwaterCoeff_Define.f90:805-984)
struct IRwaterH {
  int na = 0, nf = 0, nv = 0, nt = 0;
  std::string name;
  std::vector<double> angle, freq, wind, temp, emis;
};
IRwaterH read_irwater(const std::string& path) {
  Records rec(path);
  IRwaterH h;
  Cursor r0(rec.next());
  const int rel = r0.get<int32_t>();
  r0.get<int32_t>();
  r0.done("IRwaterCoeff release");
  if (rel != 4) throw CrtmError(fmt("IRwaterCoeff Release check failed for %s (release %d)", path.c_str(), rel));
  Cursor r1(rec.next());
  h.na = r1.get<int32_t>(); h.nf = r1.get<int32_t>(); h.nv = r1.get<int32_t>(); h.nt = r1.get<int32_t>();
  r1.done("IRwaterCoeff dimensions");
  if (h.na < 4 || h.nf < 4 || h.nv < 4 || h.nt < 1)
    throw CrtmError(path + ": the IR emissivity table needs at least 4 points per interpolated axis");
  skip_gatts(rec);
  Cursor r2(rec.next());
  h.name = r2.str(80);
  r2.done("IRwaterCoeff classification");
  if (h.name == "Nalli2" && h.nt < 4) throw CrtmError(path + ": a Nalli2 table needs at least 4 temperatures");
  Cursor r3(rec.next());
  h.angle = r3.arr<double>(h.na);
  h.freq = r3.arr<double>(h.nf);
  h.wind = r3.arr<double>(h.nv);
  h.temp = r3.arr<double>(h.nt);
  r3.done("IRwaterCoeff dimension vectors");
  Cursor r4(rec.next());
  h.emis = r4.arr<double>((size_t)h.na * h.nf * h.nv * h.nt);
  r4.done("IRwaterCoeff emissivity");
  return h;
}

// CloudCoeff release 3 (CloudCoeff_Binary_IO.f90:367-446): dimensions, axes,
// MW liquid, MW solid, IR tables.  Only phase element 1 is kept.
struct CloudCoeffH {
  int nfm = 0, nrm = 0, nfi = 0, nri = 0, nt = 0, ndm = 0, nl = 0, np = 0;
  std::vector<double> f_mw, f_ir, r_mw, r_ir, temp, dens;
  std::vector<double> ke_l, w_l, pc_l, ke_s, w_s, pc_s, ke_i, w_i, pc_i;
};

CloudCoeffH read_cloudcoeff(const std::string& path) {
  Records rec(path);
  CloudCoeffH c;
  Cursor r0(rec.next());
  const int rel = r0.get<int32_t>();
  r0.get<int32_t>();
  r0.done("CloudCoeff release");
  if (rel < 3) throw CrtmError(fmt("CloudCoeff Release check failed for %s (release %d)", path.c_str(), rel));
  Cursor r1(rec.next());
  c.nfm = r1.get<int32_t>(); c.nrm = r1.get<int32_t>(); c.nfi = r1.get<int32_t>(); c.nri = r1.get<int32_t>();
  c.nt = r1.get<int32_t>(); c.ndm = r1.get<int32_t>(); c.nl = r1.get<int32_t>(); c.np = r1.get<int32_t>();
  r1.done("CloudCoeff dimensions");
  if (c.nfm < 4 || c.nrm < 4 || c.nfi < 4 || c.nri < 4 || c.nt < 4 || c.ndm < 3 || c.np < 1)
    throw CrtmError("CloudCoeff tables need at least 4 points per interpolation axis");
  if (c.nl < 37) throw CrtmError(fmt("CloudCoeff has %d Legendre terms; the stream sets need 37", c.nl));
  const size_t ndi = c.ndm + 1;
  Cursor r2(rec.next());
  c.f_mw = r2.arr<double>(c.nfm); c.f_ir = r2.arr<double>(c.nfi); c.r_mw = r2.arr<double>(c.nrm);
  c.r_ir = r2.arr<double>(c.nri); c.temp = r2.arr<double>(c.nt); c.dens = r2.arr<double>(c.ndm);
  r2.done("CloudCoeff axes");
  const size_t nL = (size_t)c.nfm * c.nrm * c.nt, nS = (size_t)c.nfm * c.nrm * c.ndm, nI = (size_t)c.nfi * c.nri * ndi;
  Cursor r3(rec.next());
  c.ke_l = r3.arr<double>(nL); c.w_l = r3.arr<double>(nL); r3.arr<double>(nL);
  c.pc_l = r3.arr<double>(nL * (c.nl + 1) * c.np);
  r3.done("CloudCoeff microwave liquid");
  Cursor r4(rec.next());
  c.ke_s = r4.arr<double>(nS); c.w_s = r4.arr<double>(nS); r4.arr<double>(nS);
  c.pc_s = r4.arr<double>(nS * (c.nl + 1) * c.np);
  r4.done("CloudCoeff microwave solid");
  Cursor r5(rec.next());
  c.ke_i = r5.arr<double>(nI); c.w_i = r5.arr<double>(nI); r5.arr<double>(nI);
  c.pc_i = r5.arr<double>(nI * c.nl * c.np);
  r5.done("CloudCoeff infrared");
  // phase element 1 only: the leading block of each pcoeff array
  c.pc_l.resize(nL * (c.nl + 1));
  c.pc_s.resize(nS * (c.nl + 1));
  c.pc_i.resize(nI * c.nl);
  return c;
}

// AerosolCoeff, binary, release 4, 'CRTM' scheme: record order of
// AerosolCoeff_Binary_ReadFile (src/Coefficients/AerosolCoeff/AerosolCoeff_Binary_IO.f90:396-520);
// frequencies from AerosolCoeff_Frequency (10^4 / wavelength).
struct AerosolCoeffH {
  int nw = 0, nr = 0, ns = 0, nt = 0, nrh = 0, nl = 0, np = 0;
  std::vector<int> type_id;
  std::vector<double> freq, reff, ke, w, pc;
};

AerosolCoeffH read_aerosolcoeff(const std::string& path) {
  Records rec(path);
  AerosolCoeffH a;
  Cursor r0(rec.next());
  const int rel = r0.get<int32_t>();
  const int ver = r0.get<int32_t>();
  r0.done("AerosolCoeff release");
  if (rel != 4) throw CrtmError(fmt("AerosolCoeff Release check failed for %s (release %d)", path.c_str(), rel));
  Cursor r1(rec.next());
  a.nw = r1.get<int32_t>(); a.nr = r1.get<int32_t>(); a.ns = r1.get<int32_t>(); a.nt = r1.get<int32_t>();
  a.nrh = r1.get<int32_t>(); a.nl = r1.get<int32_t>(); a.np = r1.get<int32_t>();
  r1.done("AerosolCoeff dimensions");
  if (a.nw < 4 || a.nr < 4 || a.ns < 1 || a.nt < 1 || a.nrh < 1 || a.np < 1)
    throw CrtmError("AerosolCoeff tables need at least 4 points per interpolation axis");
  if (a.nl < 37) throw CrtmError(fmt("AerosolCoeff has %d Legendre terms; the stream sets need 37", a.nl));
  Cursor r2(rec.next());
  const int nrhr = r2.get<int32_t>();   // n_RH_Radii ('CRTM' scheme)
  r2.done("AerosolCoeff n_RH_Radii");
  Cursor r3(rec.next());
  const int sl = r3.get<int32_t>();
  r3.done("AerosolCoeff data source length");
  rec.next();                           // data source string
  Cursor r5(rec.next());
  a.type_id = r5.arr<int32_t>(a.nt);
  r5.done("AerosolCoeff type ids");
  Cursor r6(rec.next());
  if (r6.get<int32_t>() != sl) throw CrtmError("AerosolCoeff type name string length mismatch");
  rec.next();                           // type names
  Cursor r8(rec.next());
  const std::vector<double> wl = r8.arr<double>(a.nw);
  a.reff = r8.arr<double>((size_t)a.nr * a.nt);
  r8.arr<double>((size_t)a.ns * a.nt);
  r8.arr<double>(nrhr);
  r8.done("AerosolCoeff axes");
  a.freq.resize(a.nw);
  for (int i = 0; i < a.nw; ++i) a.freq[i] = 10000.0 / wl[i];
  for (int i = 1; i < a.nw; ++i)
    if (!(a.freq[i] > a.freq[i - 1])) throw CrtmError("AerosolCoeff wavelengths must be descending");
  const size_t n5 = (size_t)a.nw * a.nrh * a.nr * a.ns * a.nt;
  Cursor r9(rec.next());
  a.ke = r9.arr<double>(n5);
  a.w = r9.arr<double>(n5);
  r9.arr<double>(n5);                  // g
  if (ver > 300) r9.arr<double>(n5);   // kb
  a.pc = r9.arr<double>(n5 * (a.nl + 1) * a.np);
  r9.done("AerosolCoeff tables");
  a.pc.resize(n5 * (a.nl + 1));        // phase element 1
  return a;
}

// Gauss-Legendre nodes and weights on [0,1], ascending, for 1..9 points
// (Double_Gauss_Quadrature, src/CRTM_Utility/CRTM_Utility.f90:343-378);
// layout [n][2][10].
std::vector<double> gauss_table() {
  std::vector<double> g(10 * 20, 0.0);
  for (int n = 1; n <= 9; ++n) {
    double* mu = &g[n * 20];
    double* w = mu + 10;
    for (int j = 1; j <= (n + 1) / 2; ++j) {
      double x = std::cos(M_PI * (j - 0.25) / (n + 0.5)), dp = 0.0;
      for (int it = 0; it < 10; ++it) {
        double p1 = 1.0, p = x;
        for (int l = 2; l <= n; ++l) {
          const double p2 = p1;
          p1 = p;
          p = ((2 * l - 1) * x * p1 - (l - 1) * p2) / l;
        }
        dp = n * (x * p - p1) / (x * x - 1.0);
        const double xp = x;
        x = xp - p / dp;
        if (std::fabs(x - xp) <= 3.0e-14) break;
      }
      mu[j - 1] = (1.0 - x) / 2.0;
      mu[n - j] = (1.0 + x) / 2.0;
      w[j - 1] = w[n - j] = 1.0 / ((1.0 - x * x) * dp * dp);
    }
  }
  return g;
}

}  // namespace

// ---------------------------------------------------------------------------
// Scratch array of the no-allocation entry point (crtm_k_matrix_noalloc).
// Each GPU the library runs on holds its own copy from module load until the
// process ends.  Host code never uses the array's name as an address: the
// device address comes from hipGetSymbolAddress, once per device, at
// crtm_init with that device current.  Every no-allocation call on a device
// shares that device's copy; overlapping calls on one device race.
// ---------------------------------------------------------------------------
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(128) << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static char* g_scratch_ptr[MAX_DEVICES] = {};

// Device address of the current device's scratch array (set-up, once per device).
static void bind_scratch(int device) {
  if (device < 0 || device >= MAX_DEVICES) return;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void* ptr = nullptr;
  const hipError_t rc = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (rc != hipSuccess || !ptr) {
    std::fprintf(stderr, "CRTM_MI355X: hipGetSymbolAddress(scratch) on device %d: %s\n", device,
                 hipGetErrorString(rc));
    return;
  }
  g_scratch_ptr[device] = static_cast<char*>(ptr);
}

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
struct crtm_context {
  int device = 0;
  hipStream_t own_stream = nullptr;  // created by crtm_init
  hipStream_t stream = nullptr;      // where work is issued: own_stream or the caller's (crtm_set_stream)
  std::vector<std::unique_ptr<Sensor>> sensors;
  int n_channels = 0;
  std::string last_error;
  // work pool (grown on demand, reused across calls)
  DevBuf wx, wu, wo, wacci, waccw, wuidx, wufrac, wgeo, wstatus, wuwin, wfrange, waccw64, wlm, wnlte;
  DevBuf stage;  // host-mode staging of inputs/outputs
  DevBuf clim;   // model climatology table (AddLayers)
  DevBuf xnl, xloff, xnx, xlp, xp, xt, xa;   // profiles extended by AddLayers
  // clouds: CloudCoeff tables (loaded at init when present) and per-call work
  bool cloud_loaded = false;
  bool aerosol_loaded = false;
  std::vector<int> aerosol_types;   // type ids of the AerosolCoeff tables
  std::vector<std::unique_ptr<DevBuf>> cloud_bufs;
  DevCloudCoeff cc{};
  DevBuf cwc, creff, ccount, cbase, cslot, cdesc, ctotal, ctcc, cfrac, cmaxr, ctau, cscr, caerr, cwide;
  // call-level profile status: the per-sensor preparation status and the
  // scattering-path failures merged after each sensor (only ever raised)
  DevBuf cstatus;
  int* call_status = nullptr;   // the current call's call-level status (pool or scratch array)
  int* prep_status = nullptr;   // preparation status of the sensor being run
  // crtm_k_matrix_noalloc: work space is carved from the device's scratch
  // array (arena non-null) instead of the pool
  char* arena = nullptr;
  size_t arena_cap = 0, arena_used = 0;
  // all-sky K-matrix: layer adjoints, per-thread adding tape, seeds for the clear-sky K kernel
  DevBuf cscr_ad, cwide_ad, ctape, ctape_w, ctape_d, ks_tau, ks_b, ks_t, ks_ch;
  // smallest angle count run by the wavefront (matrix-core) kernels
  // (crtm_set_wavefront_angles)
  int wide_min = 9;
  // scattering-layer slots of capturable cloudy calls (crtm_set_cloudy_layer_capacity);
  // 0: the count is read back and the launches are sized for it
  int64_t cloud_cap = 0;
  // device copy of per-profile stream counts given in host memory to a device-memory call
  DevBuf nsprof;
  DevBuf skipbuf;      // Skip_Profile flags staged from host memory in device-memory calls
  // surface emissivity models (loaded at init when their files are present)
  // and the per-call optics planes / unsupported-surface flags
  bool mw_loaded = false, ir_loaded = false;
  DevMWwater mw{};
  DevIRwater ir{};
  std::vector<std::unique_ptr<DevBuf>> sfc_bufs;
  DevBuf sfcplanes, sfcerr;
  // kernel timing (crtm_profile_enable): HIP events on the context stream
  bool profiling = false;
  std::vector<hipEvent_t> ev;      // pairs (start, stop) per recorded launch
  std::vector<int> ev_kind;        // 0 = preparation kernel, 1 = forward/K kernel, 2 = scattering kernels
  std::vector<int> ev_sensor;      // sensor of each recorded launch (-1: not sensor specific)
  size_t ev_used = 0;
  hipEvent_t mark() {
    if (ev_used == ev.size()) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) throw std::runtime_error("hipEventCreate failed");
      ev.push_back(e);
    }
    hipEvent_t e = ev[ev_used++];
    if (hipEventRecord(e, stream) != hipSuccess) throw std::runtime_error("hipEventRecord failed");
    return e;
  }
};

namespace {

int fail(crtm_context* ctx, const std::string& msg) {
  if (ctx) ctx->last_error = msg;
  std::fprintf(stderr, "CRTM_MI355X: %s\n", msg.c_str());
  return CRTM_FAILURE;
}

// Work space of a call: a pool buffer grown on demand, or in a no-allocation
// call the next 256-byte-aligned piece of the device's scratch array.
void* take(crtm_context* c, DevBuf& b, size_t bytes) {
  if (!c->arena) {
    b.reserve(bytes);
    return b.p;
  }
  const size_t at = (c->arena_used + 255) & ~size_t(255);
  if (at + bytes > c->arena_cap)
    throw CrtmError(fmt("work space of %zu bytes beyond the %zu-byte scratch array", at + bytes, c->arena_cap));
  c->arena_used = at + bytes;
  return c->arena + at;
}

// Bytes of each per-profile work array of one sensor's preparation and fused
// kernels (ProfWork), in the order make_work takes them.
constexpr int N_WORK = 14;
template <class RT>
std::array<size_t, N_WORK> work_sizes(int M, int Ko, int nx, int max_k) {
  const size_t m = (size_t)M, ko = (size_t)Ko, mk = (size_t)max_k;
  return {m * ko * nx * sizeof(RT),       m * ko * NR_MAX * sizeof(RT),    m * ko * NO * sizeof(RT),
          m * ko * 3 * sizeof(int),       m * NZ_MAX * sizeof(RT),         m * (mk + 1) * sizeof(int),
          m * (mk + 1) * sizeof(RT),      m * 4 * sizeof(double),          m * mk * MAP_WIN * sizeof(RT),
          m * ko * 2 * sizeof(int),       m * sizeof(int),                 m * NZ_MAX * sizeof(double),
          (3 * mk + 1) * m * sizeof(double), m * NLTE_REC * sizeof(double)};
}

template <class RT>
ProfWork<RT> make_work(crtm_context* c, int M, int Ko, int nx, int max_k) {
  const std::array<size_t, N_WORK> sz = work_sizes<RT>(M, Ko, nx, max_k);
  DevBuf* bufs[N_WORK] = {&c->wx,    &c->wu,   &c->wo,     &c->wacci,   &c->waccw,   &c->wuidx, &c->wufrac,
                          &c->wgeo,  &c->wuwin, &c->wfrange, &c->wstatus, &c->waccw64, &c->wlm, &c->wnlte};
  void* p[N_WORK];
  for (int i = 0; i < N_WORK; ++i) p[i] = take(c, *bufs[i], sz[i]);
  ProfWork<RT> w{};
  w.x = static_cast<RT*>(p[0]);
  w.u = static_cast<RT*>(p[1]);
  w.o = static_cast<RT*>(p[2]);
  w.acc_i = static_cast<int*>(p[3]);
  w.acc_w = static_cast<RT*>(p[4]);
  w.uidx = static_cast<int*>(p[5]);
  w.ufrac = static_cast<RT*>(p[6]);
  w.geo = static_cast<double*>(p[7]);
  w.uwin = static_cast<RT*>(p[8]);
  w.frange = static_cast<int*>(p[9]);
  w.status = static_cast<int*>(p[10]);
  w.accw64 = static_cast<double*>(p[11]);
  w.lm = static_cast<double*>(p[12]);
  w.nlte = static_cast<double*>(p[13]);
  c->prep_status = w.status;
  return w;
}

// the batch as the sensor's kernels see it: ODPS absorber j -> user absorber
DevBatch sensor_batch(const Sensor& S, const DevBatch& b) {
  DevBatch bb = b;
  for (int j = 0; j < NA_MAX; ++j) bb.jmap[j] = -1;
  for (int j = 0; j < S.d.Jm; ++j)
    for (int ju = 0; ju < b.na; ++ju)
      if (S.tc.absorber_id[j] == b.jmap[ju]) {   // b.jmap carries the user absorber ids here
        bb.jmap[j] = ju;
        break;
      }
  return bb;
}

// the fused clear-sky forward / K-matrix kernel alone (preparation already run)
template <class RT, class KT, bool DOK, bool RED = false>
void run_fwdk(crtm_context* c, Sensor& S, const DevBatch& b, const ProfWork<RT>& w, const DevOut& o) {
  const DevBatch bb = sensor_batch(S, b);
  if (c->profiling) {
    c->mark();
    c->ev_kind.push_back(1);
    c->ev_sensor.push_back(S.index);
  }
  const int tile = S.tile_of<RT>();
  switch (S.d.group) {
    case 1: launch_fwdk<1, RT, KT, DOK, RED>(S.d, bb, w, o, c->stream, tile); break;
    case 2: launch_fwdk<2, RT, KT, DOK, RED>(S.d, bb, w, o, c->stream, tile); break;
    case 3: launch_fwdk<3, RT, KT, DOK, RED>(S.d, bb, w, o, c->stream, tile); break;
  }
  if (c->profiling) c->mark();
  HIPCHECK(hipGetLastError());
}

template <class RT, class KT, bool DOK, bool RED = false>
ProfWork<RT> run_sensor(crtm_context* c, Sensor& S, const DevBatch& b, const DevOut& o) {
  const int nx = S.d.group == 1 ? 89 : S.d.group == 2 ? 50 : 21;
  ProfWork<RT> w = make_work<RT>(c, b.M, S.d.Ko, nx, b.max_k);
  DevBatch bb = b;
  for (int j = 0; j < NA_MAX; ++j) bb.jmap[j] = -1;
  for (int j = 0; j < S.d.Jm; ++j)
    for (int ju = 0; ju < b.na; ++ju)
      if (S.tc.absorber_id[j] == b.jmap[ju]) {   // b.jmap carries the user absorber ids here
        bb.jmap[j] = ju;
        break;
      }
  auto timed = [&](int kind, auto&& fn) {
    if (c->profiling) {
      c->mark();
      c->ev_kind.push_back(kind);
      c->ev_sensor.push_back(S.index);
    }
    fn();
    if (c->profiling) c->mark();
  };
  const int tile = S.tile_of<RT>();
  switch (S.d.group) {
    case 1:
      timed(0, [&] { launch_prep<1, RT>(S.d, bb, w, c->stream); });
      timed(1, [&] { launch_fwdk<1, RT, KT, DOK, RED>(S.d, bb, w, o, c->stream, tile); });
      break;
    case 2:
      timed(0, [&] { launch_prep<2, RT>(S.d, bb, w, c->stream); });
      timed(1, [&] { launch_fwdk<2, RT, KT, DOK, RED>(S.d, bb, w, o, c->stream, tile); });
      break;
    case 3:
      timed(0, [&] { launch_prep<3, RT>(S.d, bb, w, c->stream); });
      timed(1, [&] { launch_fwdk<3, RT, KT, DOK, RED>(S.d, bb, w, o, c->stream, tile); });
      break;
  }
  HIPCHECK(hipGetLastError());
  return w;
}

// Cloud preparation for the whole batch (sensor independent); returns the
// number of cloudy layers the scattering launches are sized for: the counted
// layers (one host synchronisation), or with a layer capacity set the
// capacity itself (no synchronisation; the kernels stop at the device-side
// count, and profiles that do not fit fail).
int prepare_clouds(crtm_context* c, const DevBatch& be, const DevClouds& cl, CloudWork& cw, int* max_k_cloudy) {
  const int M = be.M;
  const int64_t NTe = be.NT;
  c->cwc.reserve(std::max<size_t>((size_t)cl.nc * NTe * sizeof(double), 16));
  c->creff.reserve(std::max<size_t>((size_t)cl.nc * NTe * sizeof(double), 16));
  c->ccount.reserve(M * sizeof(int));
  c->cbase.reserve(M * sizeof(int));
  c->cslot.reserve(NTe * sizeof(int));
  c->cdesc.reserve(std::max<int64_t>(be.out_NT, 1) * sizeof(int2));
  c->ctotal.reserve(2 * sizeof(int));
  c->ctcc.reserve(M * sizeof(double));
  c->cfrac.reserve(M * sizeof(int));
  c->cmaxr.reserve(M * sizeof(double));
  c->caerr.reserve(M * sizeof(int));
  cw.wc = c->cwc.as<double>();
  cw.reff = c->creff.as<double>();
  cw.cl_count = c->ccount.as<int>();
  cw.cl_base = c->cbase.as<int>();
  cw.cl_slot = c->cslot.as<int>();
  cw.cl_desc = c->cdesc.as<int2>();
  cw.total = c->ctotal.as<int>();
  cw.tcc = c->ctcc.as<double>();
  cw.frac = c->cfrac.as<int>();
  cw.max_reff = c->cmaxr.as<double>();
  cw.aer_err = c->caerr.as<int>();
  cw.capmode = c->cloud_cap > 0;
  cw.cap = (int)std::min<int64_t>(cw.capmode ? std::min<int64_t>(c->cloud_cap, be.out_NT) : be.out_NT, INT32_MAX);
  launch_zero(cw.total, 2 * sizeof(int), c->stream);
  launch_cloud_prep(be, cl, cw, c->stream);
  // aerosol failures of every infrared sensor before any sensor runs, so the
  // profile fails on all sensors of the context whatever their order
  for (auto& S : c->sensors) launch_aerosol_check(S->d, M, cl, cw, c->cc, c->stream);
  HIPCHECK(hipGetLastError());
  if (cw.capmode) {
    if (max_k_cloudy) *max_k_cloudy = be.max_k;
    return cw.cap;
  }
  int tot[2] = {0, 0};
  HIPCHECK(hipMemcpyAsync(tot, cw.total, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  if (max_k_cloudy) *max_k_cloudy = tot[1];
  return std::min(tot[0], cw.cap);
}

// Angle counts the layer and sweep kernels are launched for: the stream
// choice gives n quadrature angles plus the sensor angle (n alone when the
// sensor angle coincides with a node).
unsigned angle_mask(int n_streams) {
  if (n_streams > 0) {
    const int ns = (n_streams + 2) / 2;
    return 3u << ns;
  }
  return 0xFu << 2;   // Mie-parameter choice: 2, 4 or 6 streams
}
std::vector<int> angle_counts(unsigned mask) {
  std::vector<int> v;
  for (int z = 1; z < 32; ++z)
    if (mask >> z & 1) v.push_back(z);
  return v;
}
// A caller's array in host memory (pageable or pinned) is read directly; in
// device memory it is copied back with one synchronisation.
bool host_resident(const void* p) {
  hipPointerAttribute_t a{};
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();   // pageable host memory is unknown to the runtime
    return true;
  }
  return a.type == hipMemoryTypeHost || a.type == hipMemoryTypeUnregistered;
}
bool valid_streams(int n) { return n == 0 || n == 2 || n == 4 || n == 6 || n == 8 || n == 16; }

// Layer scratch for a set of angle counts: per-lane slots (SoA) for counts
// below wide_min, item-major wavefront slots for the others; the optics part
// (tau, ssa, phase coefficients) is always SoA.  The kernels take an angle
// count to the wavefront kernels when it exceeds sc.nzc.
template <class RT>
LayerScratch<RT> layer_scratch(const std::vector<int>& nz, int64_t n_items, DevBuf& soa, DevBuf& wide,
                               int wide_min) {
  LayerScratch<RT> sc{};
  for (int z : nz) {
    if (z < wide_min) sc.nzc = std::max(sc.nzc, z);
    else sc.nzw = std::max(sc.nzw, z);
  }
  sc.npc = sc.nzw > 0 ? NL_W - 1 : NL_AS - 1;
  sc.n_items = std::max<int64_t>(n_items, 1);
  soa.reserve((size_t)sc.n_elem() * sc.n_items * sizeof(RT));
  sc.v = soa.as<RT>();
  if (sc.nzw > 0) {
    wide.reserve((size_t)sc.w_stride() * sc.n_items * sizeof(RT));
    sc.w = wide.as<RT>();
  }
  return sc;
}

template <class RT>
void run_allsky(crtm_context* c, Sensor& S, const DevBatch& be, const DevClouds& cl, const CloudWork& cw0,
                int total_cl, const ProfWork<RT>& w, const DevOut& o, const AllskyOut& ao) {
  CloudWork cw = cw0;
  cw.status = w.status;
  const int64_t items = (int64_t)total_cl * S.d.L;
  if (items > INT32_MAX) throw CrtmError(fmt("%lld cloudy layer-channels exceed one launch", (long long)items));
  const std::vector<int> nz = angle_counts(cl.nz_mask);
  const LayerScratch<RT> sc = layer_scratch<RT>(nz, items, c->cscr, c->cwide, c->wide_min);
  auto timed = [&](auto&& fn) {
    if (c->profiling) {
      c->mark();
      c->ev_kind.push_back(2);
    c->ev_sensor.push_back(-1);
    }
    fn();
    if (c->profiling) c->mark();
  };
  timed([&] {
    launch_allsky<RT>(S.d, be, cl, cw, c->cc, sc, reinterpret_cast<const RT*>(o.tau_g), w.geo, (int)items,
                      nz.data(), (int)nz.size(), o, ao, c->stream);
  });
  HIPCHECK(hipGetLastError());
}

// Host-mode staging: copy every input array into one device block.
struct Stager {
  std::vector<std::pair<const void*, size_t>> ins;
  std::vector<std::pair<void*, size_t>> outs;
  std::vector<size_t> in_off, out_off;
  size_t total = 0;
  size_t add(size_t bytes) {
    size_t o = total;
    total += (bytes + 255) & ~size_t(255);
    return o;
  }
};

// Bytes of the arrays of the batch extended by AddLayers, in add_layers' order.
constexpr int N_EXT = 7;
std::array<size_t, N_EXT> ext_sizes(int M, int64_t NT, int na) {
  const size_t NTe = (size_t)(NT + (int64_t)M * N_EXTRA_MAX);
  return {M * sizeof(int),          (M + 1) * sizeof(int64_t), M * sizeof(int),  (NTe + M) * sizeof(double),
          NTe * sizeof(double),     NTe * sizeof(double),      std::max(na, 1) * NTe * sizeof(double)};
}

// CRTM_Atmosphere_AddLayers for the whole batch: the returned batch
// describes the extended profiles; Jacobian outputs keep the caller's layout.
DevBatch add_layers(crtm_context* c, const DevBatch& b, const int* user_ids) {
  const int M = b.M;
  const int64_t NTe = b.NT + (int64_t)M * N_EXTRA_MAX;
  const std::array<size_t, N_EXT> sz = ext_sizes(M, b.NT, b.na);
  DevBuf* bufs[N_EXT] = {&c->xnl, &c->xloff, &c->xnx, &c->xlp, &c->xp, &c->xt, &c->xa};
  void* p[N_EXT];
  for (int i = 0; i < N_EXT; ++i) p[i] = take(c, *bufs[i], sz[i]);
  ExtBatch e{static_cast<int*>(p[0]),    static_cast<int64_t*>(p[1]), static_cast<int*>(p[2]),
             static_cast<double*>(p[3]), static_cast<double*>(p[4]),  static_cast<double*>(p[5]),
             static_cast<double*>(p[6]), NTe};
  launch_add_layers(b, e, c->clim.as<double>(), user_ids, c->stream);
  HIPCHECK(hipGetLastError());
  DevBatch be = b;
  be.n_layers = e.n_layers;
  be.loff = e.loff;
  be.level_p = e.level_p;
  be.p = e.p;
  be.t = e.t;
  be.a = e.a;
  be.NT = NTe;
  be.max_k = b.max_k + N_EXTRA_MAX;
  be.n_extra = e.n_extra;
  be.out_loff = b.loff;
  be.out_NT = b.NT;
  return be;
}

// All sensors of the context: clear-sky kernels, then the scattering path
// for the cloudy profiles.
struct CloudCall {
  bool on = false;
  DevClouds cl{};
  CloudWork cw{};
  int total = 0;
  int max_k = 0;       // largest (extended) layer count of a profile with cloudy layers
  AllskyOut ao{};
  CloudKOut ko{};      // cloud Jacobians (K-matrix calls)
};

// All-sky K-matrix for one sensor: clear-sky forward (radiances and gas
// optical depths), the scattering forward + adjoint producing the cloud
// Jacobians and the seeds, then the clear-sky K-matrix kernel completing the
// gas, Planck and surface adjoints with those seeds.
template <class RT, class KT, bool RED>
void run_sensor_k_allsky(crtm_context* c, Sensor& S, const DevBatch& be, const DevOut& o, const CloudCall& cc) {
  const int L = S.d.L;
  DevOut o1 = o;
  c->ctau.reserve(std::max<size_t>((size_t)be.NT * L * sizeof(RT), 16));
  o1.tau_g = c->ctau.p;
  o1.tau_mask = cc.cw.cl_count;
  o1.ks = KSeeds{};
  ProfWork<RT> w = run_sensor<RT, RT, false>(c, S, be, o1);
  CloudWork cw = cc.cw;
  cw.status = w.status;
  const int64_t items = (int64_t)cc.total * L;
  if (items > INT32_MAX) throw CrtmError(fmt("%lld cloudy layer-channels exceed one launch", (long long)items));
  const std::vector<int> nz = angle_counts(cc.cl.nz_mask);
  const LayerScratch<RT> sc = layer_scratch<RT>(nz, items, c->cscr, c->cwide, c->wide_min);
  const LayerScratch<RT> sa = layer_scratch<RT>(nz, items, c->cscr_ad, c->cwide_ad, c->wide_min);
  const int64_t ML = (int64_t)be.M * L;
  const int64_t tape_lanes = (int64_t)allsky_sweep_ad_blocks(ML) * 256;
  const int te = sc.nzc * sc.nzc + sc.nzc;
  c->ctape.reserve(std::max<size_t>((size_t)tape_lanes * std::max(cc.max_k, 1) * te * sizeof(RT), 16));
  // wavefront adding adjoint (16 streams): 8 x 64 values per layer and wavefront
  const int kmax = std::max(cc.max_k, 1);
  if (sc.nzw > 0) {
    c->ctape_w.reserve((size_t)allsky_sweep_ad_wide_blocks(ML) * kmax * 8 * 64 * sizeof(RT));
    c->ctape_d.reserve((size_t)wide_doubling_tape_elems(std::max<int64_t>(items, 1)) * sizeof(RT));
  }
  c->ks_tau.reserve(std::max<size_t>((size_t)be.NT * L * sizeof(RT), 16));
  c->ks_b.reserve(std::max<size_t>((size_t)be.NT * L * sizeof(RT), 16));
  c->ks_t.reserve(std::max<size_t>((size_t)be.NT * L * sizeof(RT), 16));
  c->ks_ch.reserve((size_t)4 * ML * sizeof(double));
  DevOut o2 = o;
  o2.tau_g = nullptr;
  o2.tau_mask = nullptr;
  o2.ks.mask = cc.cw.cl_count;
  o2.ks.tau = c->ks_tau.p;
  o2.ks.b = c->ks_b.p;
  o2.ks.t = c->ks_t.p;
  o2.ks.rad = c->ks_ch.as<double>();
  o2.ks.emis = o2.ks.rad + ML;
  o2.ks.bs = o2.ks.emis + ML;
  o2.ks.tcc_ad = o2.ks.bs + ML;
  if (c->profiling) {
    c->mark();
    c->ev_kind.push_back(2);
    c->ev_sensor.push_back(-1);
  }
  launch_allsky_k<RT, KT>(S.d, be, cc.cl, cw, c->cc, sc, sa, reinterpret_cast<const RT*>(o1.tau_g), w.geo,
                          (int)items, nz.data(), (int)nz.size(), o2, cc.ao, cc.ko, c->ctape.as<RT>(), tape_lanes, te,
                          c->ctape_w.as<RT>(), kmax, c->ctape_d.as<RT>(), c->stream);
  launch_clear_copy(S.d, be.M, be.Ltot, cc.cw.cl_count, o1, cc.ao, c->stream);   // profiles without cloudy layers
  if (c->profiling) c->mark();
  HIPCHECK(hipGetLastError());
  run_fwdk<RT, KT, true, RED>(c, S, be, w, o2);
}

template <class RT, class KT, bool DOK, bool RED = false>
void run_sensors(crtm_context* c, const DevBatch& be, DevOut o, const CloudCall& cc) {
  c->call_status = static_cast<int*>(take(c, c->cstatus, (size_t)be.M * sizeof(int)));
  launch_zero(c->call_status, (size_t)be.M * sizeof(int), c->stream);
  const int* aer_err = cc.on ? cc.cw.aer_err : nullptr;
  // computed surface optics (clear-sky calls only): the optics kernels flag
  // profiles with a surface other than water, which then fail
  const bool sfc_comp = o.sp.mode != SFC_USER;
  int* sfc_err = nullptr;
  if (sfc_comp) {
    sfc_err = static_cast<int*>(take(c, c->sfcerr, (size_t)be.M * sizeof(int)));
    launch_zero(sfc_err, (size_t)be.M * sizeof(int), c->stream);
    aer_err = sfc_err;
  }
  o.fail = aer_err;
  auto merge = [&] {
    launch_status_merge(be.M, c->prep_status, aer_err, c->call_status, c->stream);
    HIPCHECK(hipGetLastError());
  };
  for (auto& S : c->sensors) {
    // the sensor's work space is dead once its kernels are enqueued (stream
    // order), so the next sensor reuses it
    const size_t mark = c->arena_used;
    if (DOK && cc.on) {
      run_sensor_k_allsky<RT, KT, RED>(c, *S, be, o, cc);
      merge();
      c->arena_used = mark;
      continue;
    }
    if (cc.on) {
      c->ctau.reserve(std::max<size_t>((size_t)be.NT * S->d.L * sizeof(RT), 16));
      o.tau_g = c->ctau.p;
      o.tau_mask = cc.cw.cl_count;
    }
    if (sfc_comp) {
      const int mode = S->d.sensor_type == CRTM_MICROWAVE_SENSOR ? SFC_MW : SFC_IR;
      const int64_t stride = (int64_t)be.M * S->d.L;
      double* planes = static_cast<double*>(
          take(c, c->sfcplanes, (size_t)stride * (mode == SFC_MW ? MP_N : IP_N) * sizeof(double)));
      if (c->profiling) {
        c->mark();
        c->ev_kind.push_back(3);
        c->ev_sensor.push_back(-1);
      }
      launch_sfc_optics(S->d, be, be.sfc_in, mode, c->mw, c->ir, planes, stride, sfc_err, c->stream);
      if (c->profiling) c->mark();
      HIPCHECK(hipGetLastError());
      o.sp = SfcPlanes{mode, planes, stride};
    }
    ProfWork<RT> w = run_sensor<RT, KT, DOK, RED>(c, *S, be, o);
    if (cc.on) run_allsky<RT>(c, *S, be, cc.cl, cc.cw, cc.total, w, o, cc.ao);
    else launch_clear_copy(S->d, be.M, be.Ltot, nullptr, o, cc.ao, c->stream);   // R_Clear = Radiance without clouds
    merge();
    c->arena_used = mark;
  }
}

void dispatch(crtm_context* c, const DevBatch& be, const DevOut& o, const CloudCall& cc, bool split, bool kf32,
              bool dok, bool adj = false) {
  if (adj) {   // CRTM_Adjoint: FP64 channel sums
    if (split) run_sensors<float, double, true, true>(c, be, o, cc);
    else run_sensors<double, double, true, true>(c, be, o, cc);
  } else if (dok) {
    if (split) { if (kf32) run_sensors<float, float, true>(c, be, o, cc); else run_sensors<float, double, true>(c, be, o, cc); }
    else { if (kf32) run_sensors<double, float, true>(c, be, o, cc); else run_sensors<double, double, true>(c, be, o, cc); }
  } else {
    if (split) run_sensors<float, float, false>(c, be, o, cc); else run_sensors<double, double, false>(c, be, o, cc);
  }
}

// cloud Jacobians are written only where a cloud is present
void zero_cloud_k(crtm_context* ctx, const CloudKOut& ko, int nc, int na, int64_t NT, int L, size_t szK) {
  const size_t n = (size_t)NT * L * szK;
  if (nc > 0) {
    launch_zero(ko.wc_k, nc * n, ctx->stream);
    launch_zero(ko.reff_k, nc * n, ctx->stream);
    launch_zero(ko.cf_k, n, ctx->stream);
  }
  if (na > 0) {
    launch_zero(ko.aconc_k, na * n, ctx->stream);
    launch_zero(ko.areff_k, na * n, ctx->stream);
  }
  HIPCHECK(hipGetLastError());
}

// Scratch bytes a clear-sky call takes, in the order the call carves them:
// the AddLayers batch, the call-level status, then the largest sensor's work
// (sensors run one after another on the same space).
size_t noalloc_bytes(const crtm_context* c, int M, int max_k, int64_t NT, int na, bool split) {
  size_t used = 0;
  auto add = [&](size_t& u, size_t b) { u = ((u + 255) & ~size_t(255)) + b; };
  for (size_t b : ext_sizes(M, NT, na)) add(used, b);
  add(used, (size_t)M * sizeof(int));
  size_t worst = used;
  const int max_ke = max_k + N_EXTRA_MAX;
  for (const auto& S : c->sensors) {
    const int nx = S->d.group == 1 ? 89 : S->d.group == 2 ? 50 : 21;
    size_t u = used;
    if (split)
      for (size_t b : work_sizes<float>(M, S->d.Ko, nx, max_ke)) add(u, b);
    else
      for (size_t b : work_sizes<double>(M, S->d.Ko, nx, max_ke)) add(u, b);
    worst = std::max(worst, u);
  }
  return worst;
}

// k: K-matrix outputs, or with adj the adjoint outputs in the same fields
// (FP64 channel sums: [NT] / [n][NT] rows, [M] surface temperatures,
// accumulated into)
int do_call(crtm_context* ctx, const crtm_atmosphere* atm, const crtm_surface* sfc, const crtm_geometry* geo,
            const crtm_options* opt, crtm_rtsolution* rts, crtm_k_output* k, bool adj = false) {
  if (!ctx) return fail(nullptr, "null context");
  if (!atm || !sfc || !geo || !opt || !rts) return fail(ctx, "null argument");
  // Options%Use_Emissivity = .FALSE.: the surface optics are computed
  // (water surfaces, clear-sky calls)
  const bool sfc_comp = !opt->use_emissivity;
  if (!sfc_comp && !opt->emissivity) return fail(ctx, "Options%Emissivity required with Use_Emissivity");
  if (sfc_comp) {
    if (atm->n_clouds > 0 || atm->n_aerosols > 0)
      return fail(ctx, "computed surface optics are available for clear-sky calls; give the emissivity "
                       "(Options%Use_Emissivity) for calls with clouds or aerosols");
    for (auto& S : ctx->sensors) {
      const int st = S->sc.sensor_type;
      if (st == CRTM_MICROWAVE_SENSOR) {
        if (!ctx->mw_loaded)
          return fail(ctx, "The MW water emissivity coefficients must be loaded (FASTEM6.MWwater.EmisCoeff.bin in "
                           "the crtm_init path) to compute microwave surface optics");
        for (int l = 0; l < S->sc.n_channels; ++l)
          if (S->sc.polarization[l] < 1 || S->sc.polarization[l] > 14)
            return fail(ctx, fmt("Unrecognised polarization flag for microwave channel index %d", l + 1));
      } else if (st == CRTM_INFRARED_SENSOR) {
        if (!ctx->ir_loaded)
          return fail(ctx, "The IR water emissivity coefficients must be loaded (Nalli.IRwater.EmisCoeff.bin in "
                           "the crtm_init path) to compute infrared surface optics");
      } else {
        return fail(ctx, fmt("%s: computed surface optics cover microwave and infrared sensors",
                             S->sc.sensor_id.c_str()));
      }
    }
  }
  const int M = atm->n_profiles;
  if (M <= 0) return CRTM_SUCCESS;
  if (atm->n_absorbers < 1 || atm->n_absorbers > NA_MAX)
    return fail(ctx, fmt("n_absorbers %d outside 1..%d", atm->n_absorbers, NA_MAX));
  if (!rts->radiance || !rts->brightness_temperature) return fail(ctx, "RTSolution radiance/BT arrays required");
  // under HIP graph capture only device-memory calls are valid, cloudy ones
  // with a layer capacity and one stream count for the batch: host staging,
  // the cloud-layer count and per-profile stream counts in device memory each
  // synchronise with the host
  hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
  HIPCHECK(hipStreamIsCapturing(ctx->stream, &capture));
  if (capture != hipStreamCaptureStatusNone) {
    const bool cloudy = atm->n_clouds > 0 || atm->n_aerosols > 0;
    if (opt->memory == CRTM_MEM_HOST)
      return fail(ctx, "graph capture: host-memory calls synchronise with the host; use device memory");
    if (cloudy && ctx->cloud_cap <= 0)
      return fail(ctx, "graph capture: a cloudy call reads its cloud-layer count back unless a layer capacity is "
                       "set (crtm_set_cloudy_layer_capacity)");
    if (cloudy && opt->n_streams_profile && host_resident(opt->n_streams_profile))
      return fail(ctx, "graph capture: per-profile stream counts in host memory are staged through the host; give "
                       "them in device memory or one n_streams for the batch");
  }
  CloudCall cc;
  cc.on = atm->n_clouds > 0 || atm->n_aerosols > 0;
  // the antenna and non-LTE corrections are applied by the clear-sky kernel
  // only: a cloudy call where either would change a result is refused
  if (cc.on) {
    for (const auto& S : ctx->sensors) {
      if (opt->use_antenna_correction && S->sc.ac_nfov > 0 && geo->ifov)
        return fail(ctx, "the antenna correction is applied in clear-sky calls only; a call with clouds or aerosols "
                         "needs use_antenna_correction = 0 or no ifov");
      if (!opt->no_nlte_correction && S->sc.nl_nsen > 0 && geo->source_zenith_angle)
        return fail(ctx, "the non-LTE correction is applied in clear-sky calls only; a call with clouds or aerosols "
                         "needs no_nlte_correction = 1 or no source zenith angles");
    }
  }
  if (cc.on) {
    if (atm->n_clouds > 0 && !ctx->cloud_loaded)
      return fail(ctx, "The CloudCoeff data must be loaded (CloudCoeff.bin in the crtm_init path) for cloudy profiles");
    if (atm->n_aerosols > 0 && !ctx->aerosol_loaded)
      return fail(ctx, "The AerosolCoeff data must be loaded (AerosolCoeff.bin in the crtm_init path) for aerosol "
                       "profiles");
    if (atm->n_clouds > 0 && (!atm->cloud_type || !atm->cloud_water_content || !atm->cloud_effective_radius))
      return fail(ctx, "cloud type, water content and effective radius arrays required");
    if (atm->n_aerosols > 0 &&
        (!atm->aerosol_type || !atm->aerosol_concentration || !atm->aerosol_effective_radius))
      return fail(ctx, "aerosol type, concentration and effective radius arrays required");
    if (atm->n_clouds < 0 || atm->n_aerosols < 0) return fail(ctx, "negative cloud or aerosol count");
    if (!valid_streams(opt->n_streams))
      return fail(ctx, fmt("n_streams %d: the scattering kernels take 2, 4, 6, 8 or 16 streams (0 = automatic)",
                           opt->n_streams));
    // per-profile stream counts (Options(m)%n_Streams): a host copy gives the
    // angle counts to launch for
    cc.cl.nz_mask = angle_mask(opt->n_streams);
    if (opt->n_streams_profile && capture != hipStreamCaptureStatusNone) {
      // captured call: the counts stay on the device; every angle count a
      // valid count can give is launched (kernels skip the items of other
      // counts) and a profile with an invalid count fails in the cloud
      // preparation
      cc.cl.nz_mask = 0;
      for (int n : {0, 2, 4, 6, 8, 16}) cc.cl.nz_mask |= angle_mask(n);
    } else if (opt->n_streams_profile) {
      std::vector<int> nsp(M);
      if (opt->memory == CRTM_MEM_HOST || host_resident(opt->n_streams_profile)) {
        std::memcpy(nsp.data(), opt->n_streams_profile, M * sizeof(int));
      } else {
        HIPCHECK(hipMemcpyAsync(nsp.data(), opt->n_streams_profile, M * sizeof(int), hipMemcpyDeviceToHost,
                                ctx->stream));
        HIPCHECK(hipStreamSynchronize(ctx->stream));
      }
      cc.cl.nz_mask = 0;
      for (int m = 0; m < M; ++m) {
        if (!valid_streams(nsp[m]))
          return fail(ctx, fmt("profile #%d: n_streams %d: the scattering kernels take 2, 4, 6, 8 or 16 streams "
                               "(0 = automatic)", m + 1, nsp[m]));
        cc.cl.nz_mask |= angle_mask(nsp[m]);
      }
    }
    if (opt->overlap_id < 0 || opt->overlap_id > 5) return fail(ctx, "Invalid overlap assumption");
    if (k && atm->n_clouds > 0 &&
        (!k->cloud_water_content_k || !k->cloud_effective_radius_k || !k->cloud_fraction_k))
      return fail(ctx, "cloud Jacobian arrays (water content, effective radius, fraction) required");
    if (k && atm->n_aerosols > 0 && (!k->aerosol_concentration_k || !k->aerosol_effective_radius_k))
      return fail(ctx, "aerosol Jacobian arrays (concentration, effective radius) required");
    cc.cl.nc = atm->n_clouds;
    cc.cl.na = atm->n_aerosols;
    cc.cl.n_streams = opt->n_streams;
    cc.cl.overlap = opt->overlap_id > 0 ? opt->overlap_id : 4;
    cc.cl.no_scat = opt->no_scattering != 0;
  }
  const bool host = opt->memory == CRTM_MEM_HOST;
  const bool split = opt->precision == CRTM_PREC_FP32_SPLIT;
  const bool kf32 = !adj && opt->jacobian_f32 != 0;
  const int L = ctx->n_channels;
  HIPCHECK(hipSetDevice(ctx->device));
  // layer bookkeeping: host copies of n_layers / offsets (host mode), or the
  // caller's scalars (device mode, no synchronisation)
  std::vector<int> nl;
  std::vector<int64_t> off;
  int max_k = atm->max_layers;
  int64_t NT = atm->total_layers;
  if (host || max_k <= 0 || NT <= 0) {
    nl.resize(M);
    off.resize(M + 1);
    if (host) {
      std::memcpy(nl.data(), atm->n_layers, M * sizeof(int));
      std::memcpy(off.data(), atm->layer_offset, (M + 1) * sizeof(int64_t));
    } else {
      HIPCHECK(hipMemcpyAsync(nl.data(), atm->n_layers, M * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
      HIPCHECK(hipMemcpyAsync(off.data(), atm->layer_offset, (M + 1) * sizeof(int64_t), hipMemcpyDeviceToHost,
                              ctx->stream));
      HIPCHECK(hipStreamSynchronize(ctx->stream));
    }
    max_k = 0;
    for (int m = 0; m < M; ++m) {
      if (nl[m] < 1 || nl[m] > CRTM_MAX_N_LAYERS)
        return fail(ctx, fmt("profile %d: invalid n_layers %d", m + 1, nl[m]));
      if (off[m + 1] - off[m] != nl[m]) return fail(ctx, fmt("profile %d: layer_offset inconsistent", m + 1));
      max_k = std::max(max_k, nl[m]);
    }
    NT = off[M];
  }
  if (max_k < 1 || max_k > CRTM_MAX_N_LAYERS) return fail(ctx, fmt("invalid max_layers %d", max_k));
  const int na = atm->n_absorbers;
  std::vector<int> abs_id(atm->absorber_id, atm->absorber_id + na);
  abs_id.resize(NA_MAX, 0);
  if (opt->check_input && host) {
    for (int m = 0; m < M; ++m) {
      if (opt->skip_profile && opt->skip_profile[m]) continue;   // checked after the skip in the reference
      const double* lp = atm->level_pressure + off[m] + m;
      for (int kk = 0; kk < nl[m]; ++kk)
        if (!(lp[kk + 1] > lp[kk]) || !(atm->temperature[off[m] + kk] > 0.0))
          return fail(ctx, fmt("Input data check failed for profile #%d", m + 1));
    }
  }
  if (sfc_comp && host)
    for (int m = 0; m < M; ++m) {
      if (opt->skip_profile && opt->skip_profile[m]) continue;
      if (sfc->land_coverage[m] > 0.0 || sfc->snow_coverage[m] > 0.0 || sfc->ice_coverage[m] > 0.0)
        return fail(ctx, fmt("profile #%d: computed surface optics cover water surfaces (land, snow and ice "
                             "coverage must be zero)", m + 1));
    }
  const size_t szK = kf32 ? 4 : 8;
  if (rts->status && host)
    for (int m = 0; m < M; ++m) rts->status[m] = CRTM_SUCCESS;
  DevBatch b{};
  b.M = M;
  b.na = na;
  b.ac_on = opt->use_antenna_correction != 0;
  b.nlte_on = opt->no_nlte_correction == 0;
  b.max_k = max_k;
  b.NT = NT;
  b.Ltot = L;
  for (int j = 0; j < NA_MAX; ++j) b.jmap[j] = j < na ? abs_id[j] : -1;
  DevOut o{};
  o.k_f32 = kf32;
  o.sp.mode = sfc_comp ? SFC_MW : SFC_USER;   // per sensor in run_sensors
  const bool dok = k != nullptr;
  if (host) {
    // stage inputs
    struct In { const void* h; size_t n; const void** d; };
    std::vector<In> ins;
    auto in = [&](const void* h, size_t n, const void** d) { if (h) ins.push_back({h, n, d}); else *d = nullptr; };
    in(atm->n_layers, M * sizeof(int), (const void**)&b.n_layers);
    in(atm->layer_offset, (M + 1) * sizeof(int64_t), (const void**)&b.loff);
    in(atm->level_pressure, (NT + M) * sizeof(double), (const void**)&b.level_p);
    in(atm->pressure, NT * sizeof(double), (const void**)&b.p);
    in(atm->temperature, NT * sizeof(double), (const void**)&b.t);
    in(atm->absorber, na * NT * sizeof(double), (const void**)&b.a);
    in(atm->climatology, M * sizeof(int), (const void**)&b.climatology);
    in(opt->skip_profile, M * sizeof(int), (const void**)&b.skip);
    if (cc.on) {
      const size_t nc = atm->n_clouds;
      in(atm->cloud_type, M * nc * sizeof(int), (const void**)&cc.cl.type);
      in(atm->cloud_water_content, nc * NT * sizeof(double), (const void**)&cc.cl.wc);
      in(atm->cloud_effective_radius, nc * NT * sizeof(double), (const void**)&cc.cl.reff);
      in(atm->cloud_fraction, NT * sizeof(double), (const void**)&cc.cl.cf);
      in(opt->n_streams_profile, M * sizeof(int), (const void**)&cc.cl.ns_prof);
      for (size_t i = 0; i < (size_t)M * nc; ++i)
        if (atm->cloud_type[i] < 0 || atm->cloud_type[i] > 6)
          return fail(ctx, fmt("profile #%d: invalid cloud type %d", (int)(i / nc) + 1, atm->cloud_type[i]));
      const size_t na_ = atm->n_aerosols;
      in(atm->aerosol_type, M * na_ * sizeof(int), (const void**)&cc.cl.atype);
      in(atm->aerosol_concentration, na_ * NT * sizeof(double), (const void**)&cc.cl.aconc);
      in(atm->aerosol_effective_radius, na_ * NT * sizeof(double), (const void**)&cc.cl.areff);
      for (size_t i = 0; i < (size_t)M * na_; ++i) {
        const int t = atm->aerosol_type[i];
        if (t != 0 && std::find(ctx->aerosol_types.begin(), ctx->aerosol_types.end(), t) == ctx->aerosol_types.end())
          return fail(ctx, fmt("Invalid aerosol type ID %d, not exist in specified aerosol coefficient file", t));
      }
    }
    const double* cov[4] = {sfc->land_coverage, sfc->water_coverage, sfc->snow_coverage, sfc->ice_coverage};
    const double* st[4] = {sfc->land_temperature, sfc->water_temperature, sfc->snow_temperature, sfc->ice_temperature};
    for (int c = 0; c < 4; ++c) {
      in(cov[c], M * sizeof(double), (const void**)&b.cov[c]);
      in(st[c], M * sizeof(double), (const void**)&b.sfc_t[c]);
    }
    in(geo->sensor_zenith_angle, M * sizeof(double), (const void**)&b.zenith);
    in(geo->surface_altitude, M * sizeof(double), (const void**)&b.altitude);
    in(geo->source_zenith_angle, M * sizeof(double), (const void**)&b.source_zenith);
    in(geo->ifov, M * sizeof(int), (const void**)&b.ifov);
    in(sfc_comp ? nullptr : opt->emissivity, (size_t)M * L * sizeof(double), (const void**)&b.emissivity);
    if (sfc_comp) {
      in(sfc->salinity, M * sizeof(double), (const void**)&b.sfc_in.salinity);
      in(sfc->wind_speed, M * sizeof(double), (const void**)&b.sfc_in.wind_speed);
      in(sfc->wind_direction, M * sizeof(double), (const void**)&b.sfc_in.wind_direction);
      in(geo->sensor_azimuth_angle, M * sizeof(double), (const void**)&b.sfc_in.azimuth);
      in(geo->sensor_scan_angle, M * sizeof(double), (const void**)&b.sfc_in.scan);
    }
    if (dok) {
      in(k->brightness_temperature_k, (size_t)M * L * sizeof(double), (const void**)&o.bt_k);
      in(k->radiance_k, (size_t)M * L * sizeof(double), (const void**)&o.rad_k);
    }
    // acc: an adjoint output, added onto the caller's array after the call
    struct Out { void* h; size_t n; void** d; bool acc; };
    std::vector<Out> outs;
    auto out = [&](void* h, size_t n, void** d, bool acc = false) {
      if (h) outs.push_back({h, n, d, acc}); else *d = nullptr;
    };
    const size_t rowsL = adj ? 1 : (size_t)L;   // adjoint rows are channel sums
    const size_t szO = adj ? sizeof(double) : szK;
    const size_t ML = (size_t)M * L * sizeof(double);
    out(rts->radiance, ML, (void**)&o.radiance);
    out(rts->brightness_temperature, ML, (void**)&o.bt);
    out(rts->up_radiance, ML, (void**)&o.up_radiance);
    out(rts->down_radiance, ML, (void**)&o.down_radiance);
    out(rts->surface_planck_radiance, ML, (void**)&o.sfc_planck);
    out(rts->surface_emissivity, ML, (void**)&o.sfc_emis);
    out(rts->surface_reflectivity, ML, (void**)&o.sfc_refl);
    out(rts->r_clear, ML, (void**)&cc.ao.r_clear);
    out(rts->tb_clear, ML, (void**)&cc.ao.tb_clear);
    out(rts->total_cloud_cover, ML, (void**)&cc.ao.tcc);
    if (dok) {
      out(k->radiance_k, ML, (void**)&o.rad_k);
      out(k->surface_emissivity_k, ML, (void**)&o.emis_k);
      double* stk[4] = {k->land_temperature_k, k->water_temperature_k, k->snow_temperature_k, k->ice_temperature_k};
      for (int c = 0; c < 4; ++c) out(stk[c], adj ? M * sizeof(double) : ML, (void**)&o.sfc_t_k[c], adj);
      double* sk3[3] = {k->salinity_k, k->wind_speed_k, k->wind_direction_k};
      for (int j = 0; j < 3; ++j) out(sk3[j], adj ? M * sizeof(double) : ML, (void**)&o.sfc_k[j], adj);
      out(k->temperature_k, (size_t)NT * rowsL * szO, (void**)&o.t_k, adj);
      out(k->absorber_k, (size_t)na * NT * rowsL * szO, (void**)&o.a_k, adj);
      if (cc.on) {
        const size_t nc = atm->n_clouds, na_ = atm->n_aerosols;
        if (nc > 0) {
          out(k->cloud_water_content_k, nc * NT * rowsL * szO, &cc.ko.wc_k, adj);
          out(k->cloud_effective_radius_k, nc * NT * rowsL * szO, &cc.ko.reff_k, adj);
          out(k->cloud_fraction_k, (size_t)NT * rowsL * szO, &cc.ko.cf_k, adj);
        }
        if (na_ > 0) {
          out(k->aerosol_concentration_k, na_ * NT * rowsL * szO, &cc.ko.aconc_k, adj);
          out(k->aerosol_effective_radius_k, na_ * NT * rowsL * szO, &cc.ko.areff_k, adj);
        }
      }
    }
    // rad_k is both input and output: give it one device slot
    size_t total = 0;
    std::vector<size_t> ioff, ooff;
    for (auto& x : ins) { ioff.push_back(total); total += (x.n + 255) & ~size_t(255); }
    for (auto& x : outs) {
      if (dok && x.h == (void*)k->radiance_k) { ooff.push_back(SIZE_MAX); continue; }
      ooff.push_back(total);
      total += (x.n + 255) & ~size_t(255);
    }
    ctx->stage.reserve(std::max<size_t>(total, 256));
    char* base = ctx->stage.as<char>();
    for (size_t i = 0; i < ins.size(); ++i) {
      HIPCHECK(hipMemcpyAsync(base + ioff[i], ins[i].h, ins[i].n, hipMemcpyHostToDevice, ctx->stream));
      *ins[i].d = base + ioff[i];
    }
    for (size_t i = 0; i < outs.size(); ++i) {
      if (ooff[i] == SIZE_MAX) { ooff[i] = (char*)o.rad_k - base; continue; }
      *outs[i].d = base + ooff[i];
    }
    if (dok && !k->radiance_k) o.rad_k = nullptr;
    if (adj) {   // the device sums start from zero; the host adds them to the caller's arrays
      for (size_t i = 0; i < outs.size(); ++i)
        if (outs[i].acc) launch_zero(base + ooff[i], outs[i].n, ctx->stream);
    } else if (dok && cc.on) {
      zero_cloud_k(ctx, cc.ko, atm->n_clouds, atm->n_aerosols, NT, L, szK);
    }
    cc.ko.reduce = adj;
    const DevBatch be = add_layers(ctx, b, abs_id.data());
    if (cc.on) cc.total = prepare_clouds(ctx, be, cc.cl, cc.cw, &cc.max_k);
    dispatch(ctx, be, o, cc, split, kf32, dok, adj);
    // per-profile failures: the outputs of the good profiles are written and
    // the call returns FAILURE (src/CRTM_Forward_Module.f90:412-421)
    std::vector<int> pst(M);
    HIPCHECK(hipMemcpyAsync(pst.data(), ctx->call_status, M * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    std::vector<std::vector<double>> sums(outs.size());
    for (size_t i = 0; i < outs.size(); ++i) {
      void* dst = outs[i].h;
      if (outs[i].acc) {
        sums[i].resize(outs[i].n / sizeof(double));
        dst = sums[i].data();
      }
      HIPCHECK(hipMemcpyAsync(dst, base + ooff[i], outs[i].n, hipMemcpyDeviceToHost, ctx->stream));
    }
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    for (size_t i = 0; i < outs.size(); ++i)
      if (outs[i].acc) {
        double* h = static_cast<double*>(outs[i].h);
        for (size_t j = 0; j < sums[i].size(); ++j) h[j] += sums[i][j];
      }
    int first_bad = -1;
    for (int m = 0; m < M; ++m) {
      if (rts->status) rts->status[m] = std::max(rts->status[m], pst[m] ? CRTM_FAILURE : CRTM_SUCCESS);
      if (pst[m] && first_bad < 0) first_bad = m;
    }
    if (first_bad >= 0) return fail(ctx, fmt("profile #%d failed", first_bad + 1));
    return CRTM_SUCCESS;
  }
  // device mode: pointers are used as given, asynchronously on the context stream
  b.n_layers = atm->n_layers;
  b.loff = atm->layer_offset;
  b.level_p = atm->level_pressure;
  b.p = atm->pressure;
  b.t = atm->temperature;
  b.a = atm->absorber;
  b.climatology = atm->climatology;
  b.skip = opt->skip_profile;
  if (b.skip && host_resident(b.skip)) {   // stage host flags for the kernels
    if (capture != hipStreamCaptureStatusNone)
      return fail(ctx, "graph capture: skip_profile must be device memory in a captured call");
    int* d = static_cast<int*>(take(ctx, ctx->skipbuf, (size_t)M * sizeof(int)));
    HIPCHECK(hipMemcpyAsync(d, b.skip, (size_t)M * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    b.skip = d;
  }
  b.cov[0] = sfc->land_coverage; b.cov[1] = sfc->water_coverage; b.cov[2] = sfc->snow_coverage; b.cov[3] = sfc->ice_coverage;
  b.sfc_t[0] = sfc->land_temperature; b.sfc_t[1] = sfc->water_temperature;
  b.sfc_t[2] = sfc->snow_temperature; b.sfc_t[3] = sfc->ice_temperature;
  for (int c = 0; c < 4; ++c)
    if (!b.cov[c] || !b.sfc_t[c]) return fail(ctx, "surface coverage/temperature arrays required");
  b.zenith = geo->sensor_zenith_angle;
  b.altitude = geo->surface_altitude;
  b.source_zenith = geo->source_zenith_angle;
  b.ifov = geo->ifov;
  b.emissivity = sfc_comp ? nullptr : opt->emissivity;
  if (sfc_comp) {
    b.sfc_in.salinity = sfc->salinity;
    b.sfc_in.wind_speed = sfc->wind_speed;
    b.sfc_in.wind_direction = sfc->wind_direction;
    b.sfc_in.azimuth = geo->sensor_azimuth_angle;
    b.sfc_in.scan = geo->sensor_scan_angle;
  }
  o.radiance = rts->radiance;
  o.bt = rts->brightness_temperature;
  o.up_radiance = rts->up_radiance;
  o.down_radiance = rts->down_radiance;
  o.sfc_planck = rts->surface_planck_radiance;
  o.sfc_emis = rts->surface_emissivity;
  o.sfc_refl = rts->surface_reflectivity;
  if (dok) {
    o.sfc_k[0] = k->salinity_k;
    o.sfc_k[1] = k->wind_speed_k;
    o.sfc_k[2] = k->wind_direction_k;
    o.bt_k = k->brightness_temperature_k;
    o.rad_k = k->radiance_k;
    o.emis_k = k->surface_emissivity_k;
    o.sfc_t_k[0] = k->land_temperature_k;
    o.sfc_t_k[1] = k->water_temperature_k;
    o.sfc_t_k[2] = k->snow_temperature_k;
    o.sfc_t_k[3] = k->ice_temperature_k;
    o.t_k = k->temperature_k;
    o.a_k = k->absorber_k;
    if (!o.t_k || !o.a_k) return fail(ctx, "K-matrix temperature/absorber arrays required");
  }
  cc.ao.r_clear = rts->r_clear;
  cc.ao.tb_clear = rts->tb_clear;
  cc.ao.tcc = rts->total_cloud_cover;
  if (cc.on) {
    cc.cl.type = atm->cloud_type;
    cc.cl.wc = atm->cloud_water_content;
    cc.cl.reff = atm->cloud_effective_radius;
    cc.cl.cf = atm->cloud_fraction;
    cc.cl.atype = atm->aerosol_type;
    cc.cl.aconc = atm->aerosol_concentration;
    cc.cl.areff = atm->aerosol_effective_radius;
    cc.cl.ns_prof = opt->n_streams_profile;
    if (opt->n_streams_profile && host_resident(opt->n_streams_profile)) {   // stage host values for the kernels
      ctx->nsprof.reserve((size_t)M * sizeof(int));
      HIPCHECK(hipMemcpyAsync(ctx->nsprof.p, opt->n_streams_profile, (size_t)M * sizeof(int), hipMemcpyHostToDevice,
                              ctx->stream));
      cc.cl.ns_prof = ctx->nsprof.as<int>();
    }
    if (dok) {
      cc.ko.wc_k = k->cloud_water_content_k;
      cc.ko.reff_k = k->cloud_effective_radius_k;
      cc.ko.cf_k = k->cloud_fraction_k;
      cc.ko.aconc_k = k->aerosol_concentration_k;
      cc.ko.areff_k = k->aerosol_effective_radius_k;
      if (!adj) zero_cloud_k(ctx, cc.ko, atm->n_clouds, atm->n_aerosols, NT, L, szK);
      cc.ko.reduce = adj;
    }
  }
  const DevBatch be = add_layers(ctx, b, abs_id.data());
  if (cc.on) cc.total = prepare_clouds(ctx, be, cc.cl, cc.cw, &cc.max_k);
  dispatch(ctx, be, o, cc, split, kf32, dok, adj);
  if (rts->status)
    HIPCHECK(hipMemcpyAsync(rts->status, ctx->call_status, M * sizeof(int), hipMemcpyDeviceToDevice, ctx->stream));
  return CRTM_SUCCESS;
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

int crtm_init(int n_sensors, const char* const* sensor_ids, const char* file_path, crtm_context** out) {
  if (!out) return CRTM_FAILURE;
  *out = nullptr;
  std::unique_ptr<crtm_context> ctx(new crtm_context);
  try {
    HIPCHECK(hipGetDevice(&ctx->device));
    bind_scratch(ctx->device);   // the device is current here
    HIPCHECK(hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking));
    ctx->stream = ctx->own_stream;
    std::string dir = file_path ? file_path : ".";
    if (!dir.empty() && dir.back() != '/') dir += '/';
    int off = 0;
    for (int i = 0; i < n_sensors; ++i) {
      std::unique_ptr<Sensor> S(new Sensor);
      S->sc = read_spccoeff(dir + sensor_ids[i] + ".SpcCoeff.bin");
      S->tc = read_odps(dir + sensor_ids[i] + ".TauCoeff.bin");
      build_sensor(*S, off);
      S->index = i;
      off += S->tc.L;
      ctx->sensors.push_back(std::move(S));
    }
    ctx->n_channels = off;
    upload(ctx->clim, std::vector<double>(CLIM_TABLE, CLIM_TABLE + CLIM_SIZE));
    const std::string ccpath = dir + "CloudCoeff.bin";
    if (std::ifstream(ccpath).good()) {
      const CloudCoeffH h = read_cloudcoeff(ccpath);
      auto put = [&](const std::vector<double>& v) {
        ctx->cloud_bufs.emplace_back(new DevBuf);
        upload(*ctx->cloud_bufs.back(), v);
        return ctx->cloud_bufs.back()->as<double>();
      };
      DevCloudCoeff& c = ctx->cc;
      c.nfm = h.nfm; c.nrm = h.nrm; c.nt = h.nt; c.ndm = h.ndm; c.nfi = h.nfi; c.nri = h.nri; c.ndi = h.ndm + 1;
      c.nlf = h.nl;
      c.f_mw = put(h.f_mw); c.r_mw = put(h.r_mw); c.temp = put(h.temp); c.f_ir = put(h.f_ir); c.r_ir = put(h.r_ir);
      c.ke_l = put(h.ke_l); c.w_l = put(h.w_l); c.pc_l = put(h.pc_l);
      c.ke_s = put(h.ke_s); c.w_s = put(h.w_s); c.pc_s = put(h.pc_s);
      c.ke_i = put(h.ke_i); c.w_i = put(h.w_i); c.pc_i = put(h.pc_i);
      ctx->cloud_loaded = true;
    }
    {
      ctx->cloud_bufs.emplace_back(new DevBuf);
      upload(*ctx->cloud_bufs.back(), gauss_table());
      ctx->cc.gauss = ctx->cloud_bufs.back()->as<double>();
    }
    // surface emissivity models under the reference's default file names
    // (src/CRTM_LifeCycle.f90:682, 691)
    auto putd = [&](const std::vector<double>& v) {
      ctx->sfc_bufs.emplace_back(new DevBuf);
      upload(*ctx->sfc_bufs.back(), v);
      return ctx->sfc_bufs.back()->as<double>();
    };
    const std::string mwpath = dir + "FASTEM6.MWwater.EmisCoeff.bin";
    if (std::ifstream(mwpath).good()) {
      const MWwaterH h = read_mwwater(mwpath);
      DevMWwater& c = ctx->mw;
      c.fc = putd(h.fit[0]); c.fr = putd(h.fit[1]); c.rc = putd(h.fit[2]);
      c.az = putd(h.fit[3]); c.ssc = putd(h.fit[4]); c.lsc = putd(h.fit[5]);
      ctx->mw_loaded = true;
    }
    const std::string irpath = dir + "Nalli.IRwater.EmisCoeff.bin";
    if (std::ifstream(irpath).good()) {
      const IRwaterH h = read_irwater(irpath);
      DevIRwater& c = ctx->ir;
      c.na = h.na; c.nf = h.nf; c.nv = h.nv; c.nt = h.nt;
      c.has_t = h.name == "Nalli2";
      c.angle = putd(h.angle); c.frequency = putd(h.freq); c.wind = putd(h.wind); c.temp = putd(h.temp);
      c.emis = putd(h.emis);
      ctx->ir_loaded = true;
    }
    const std::string acpath = dir + "AerosolCoeff.bin";
    if (std::ifstream(acpath).good()) {
      const AerosolCoeffH h = read_aerosolcoeff(acpath);
      auto put = [&](const std::vector<double>& v) {
        ctx->cloud_bufs.emplace_back(new DevBuf);
        upload(*ctx->cloud_bufs.back(), v);
        return ctx->cloud_bufs.back()->as<double>();
      };
      DevAerosolCoeff& a = ctx->cc.aero;
      a.nw = h.nw; a.nr = h.nr; a.ns = h.ns; a.nt = h.nt; a.nrh = h.nrh; a.nlf = h.nl;
      ctx->cloud_bufs.emplace_back(new DevBuf);
      upload(*ctx->cloud_bufs.back(), h.type_id);
      a.type_id = ctx->cloud_bufs.back()->as<int>();
      a.freq = put(h.freq); a.reff = put(h.reff); a.ke = put(h.ke); a.w = put(h.w); a.pc = put(h.pc);
      ctx->aerosol_types = h.type_id;
      ctx->aerosol_loaded = true;
    }
    HIPCHECK(hipDeviceSynchronize());
  } catch (const std::exception& e) {
    std::fprintf(stderr, "CRTM_MI355X crtm_init: %s\n", e.what());
    return CRTM_FAILURE;
  }
  *out = ctx.release();
  return CRTM_SUCCESS;
}

int crtm_destroy(crtm_context* ctx) {
  if (!ctx) return CRTM_FAILURE;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  ctx->sensors.clear();
  for (hipEvent_t e : ctx->ev) (void)hipEventDestroy(e);
  if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
  delete ctx;
  return CRTM_SUCCESS;
}

int crtm_n_sensors(const crtm_context* ctx) { return ctx ? (int)ctx->sensors.size() : 0; }
int crtm_n_channels_total(const crtm_context* ctx) { return ctx ? ctx->n_channels : 0; }

int crtm_channel_info(const crtm_context* ctx, int s, int* n_channels, int* sensor_type, int* sensor_channel,
                      char* sensor_id) {
  if (!ctx || s < 0 || s >= (int)ctx->sensors.size()) return CRTM_FAILURE;
  const Sensor& S = *ctx->sensors[s];
  if (n_channels) *n_channels = S.tc.L;
  if (sensor_type) *sensor_type = S.sc.sensor_type;
  if (sensor_channel) std::copy(S.sc.sensor_channel.begin(), S.sc.sensor_channel.end(), sensor_channel);
  if (sensor_id) {
    std::strncpy(sensor_id, S.sc.sensor_id.c_str(), 20);
    sensor_id[20] = '\0';
  }
  return CRTM_SUCCESS;
}

int crtm_forward(crtm_context* ctx, const crtm_atmosphere* atm, const crtm_surface* sfc, const crtm_geometry* geo,
                 const crtm_options* opt, crtm_rtsolution* rts) {
  try {
    return do_call(ctx, atm, sfc, geo, opt, rts, nullptr);
  } catch (const std::exception& e) {
    return fail(ctx, e.what());
  }
}

int crtm_k_matrix(crtm_context* ctx, const crtm_atmosphere* atm, const crtm_surface* sfc, const crtm_geometry* geo,
                  const crtm_options* opt, crtm_rtsolution* rts, crtm_k_output* k) {
  if (!k) return fail(ctx, "null K-matrix output");
  try {
    return do_call(ctx, atm, sfc, geo, opt, rts, k);
  } catch (const std::exception& e) {
    return fail(ctx, e.what());
  }
}

int crtm_adjoint(crtm_context* ctx, const crtm_atmosphere* atm, const crtm_surface* sfc, const crtm_geometry* geo,
                 const crtm_options* opt, crtm_rtsolution* rts, crtm_ad_output* ad) {
  if (!ctx) return fail(nullptr, "null context");
  if (!ad) return fail(ctx, "null adjoint output");
  if (!ad->temperature_ad || !ad->absorber_ad) return fail(ctx, "adjoint temperature/absorber arrays required");
  crtm_k_output k{};   // the same kernels with channel-summed outputs
  k.brightness_temperature_k = ad->brightness_temperature_ad;
  k.radiance_k = ad->radiance_ad;
  k.surface_emissivity_k = ad->surface_emissivity_ad;
  k.salinity_k = ad->salinity_ad;
  k.wind_speed_k = ad->wind_speed_ad;
  k.wind_direction_k = ad->wind_direction_ad;
  k.temperature_k = ad->temperature_ad;
  k.absorber_k = ad->absorber_ad;
  k.land_temperature_k = ad->land_temperature_ad;
  k.water_temperature_k = ad->water_temperature_ad;
  k.snow_temperature_k = ad->snow_temperature_ad;
  k.ice_temperature_k = ad->ice_temperature_ad;
  k.cloud_water_content_k = ad->cloud_water_content_ad;
  k.cloud_effective_radius_k = ad->cloud_effective_radius_ad;
  k.cloud_fraction_k = ad->cloud_fraction_ad;
  k.aerosol_concentration_k = ad->aerosol_concentration_ad;
  k.aerosol_effective_radius_k = ad->aerosol_effective_radius_ad;
  try {
    return do_call(ctx, atm, sfc, geo, opt, rts, &k, true);
  } catch (const std::exception& e) {
    return fail(ctx, e.what());
  }
}

int crtm_k_matrix_noalloc(crtm_context* ctx, const crtm_atmosphere* atm, const crtm_surface* sfc,
                          const crtm_geometry* geo, const crtm_options* opt, crtm_rtsolution* rts,
                          crtm_k_output* k) {
  if (!ctx) return fail(nullptr, "null context");
  if (!atm || !sfc || !geo || !opt || !rts || !k) return fail(ctx, "crtm_k_matrix_noalloc: null argument");
  if (opt->memory != CRTM_MEM_DEVICE)
    return fail(ctx, "crtm_k_matrix_noalloc: device-memory calls only (host calls stage through allocated memory)");
  if (atm->n_clouds > 0 || atm->n_aerosols > 0)
    return fail(ctx, "crtm_k_matrix_noalloc: clear-sky calls only (the scattering path sizes its work space from "
                     "the batch's cloudy layers)");
  if (atm->max_layers <= 0 || atm->max_layers > CRTM_MAX_N_LAYERS || atm->total_layers <= 0)
    return fail(ctx, "crtm_k_matrix_noalloc: give max_layers (1..200) and total_layers");
  if (ctx->profiling)
    return fail(ctx, "crtm_k_matrix_noalloc: kernel timing creates HIP events; disable crtm_profile_enable");
  if (opt->skip_profile && host_resident(opt->skip_profile))
    return fail(ctx, "crtm_k_matrix_noalloc: skip_profile must be device memory");
  const int dev = ctx->device;
  char* base = (dev >= 0 && dev < MAX_DEVICES) ? g_scratch_ptr[dev] : nullptr;
  if (!base) return fail(ctx, fmt("crtm_k_matrix_noalloc: no scratch array address on device %d", dev));
  const int M = atm->n_profiles;
  if (M <= 0) return CRTM_SUCCESS;
  const size_t need = noalloc_bytes(ctx, M, atm->max_layers, atm->total_layers, atm->n_absorbers,
                                    opt->precision == CRTM_PREC_FP32_SPLIT);
  if (need > NOALLOC_SCRATCH_BYTES)
    return fail(ctx, fmt("crtm_k_matrix_noalloc: %d profiles of up to %d layers need %.1f MiB of work space, more "
                         "than the %zu MiB scratch array; split the batch",
                         M, atm->max_layers, need / 1048576.0, NOALLOC_SCRATCH_BYTES >> 20));
  ctx->arena = base;
  ctx->arena_cap = NOALLOC_SCRATCH_BYTES;
  ctx->arena_used = 0;
  int st;
  try {
    st = do_call(ctx, atm, sfc, geo, opt, rts, k);
  } catch (const std::exception& e) {
    st = fail(ctx, e.what());
  }
  ctx->arena = nullptr;
  ctx->arena_cap = 0;
  ctx->arena_used = 0;
  return st;
}

int crtm_tile_profiles(int precision, int tile) {
  if (tile < 0 || tile >= N_TILES) return 0;
  return precision == CRTM_PREC_FP64 ? TileSet<double>::pw[tile] : TileSet<float>::pw[tile];
}

int crtm_set_tile(crtm_context* ctx, int sensor, int precision, int tile) {
  if (!ctx) return CRTM_FAILURE;
  if (sensor < 0 || sensor >= (int)ctx->sensors.size()) return fail(ctx, fmt("sensor index %d", sensor));
  if (precision != CRTM_PREC_FP64 && precision != CRTM_PREC_FP32_SPLIT)
    return fail(ctx, fmt("precision %d", precision));
  if (tile < 0 || tile >= N_TILES) return fail(ctx, fmt("tile index %d: 0..%d", tile, N_TILES - 1));
  ctx->sensors[sensor]->tile[precision == CRTM_PREC_FP64 ? 0 : 1] = tile;
  return CRTM_SUCCESS;
}

int crtm_get_tile(const crtm_context* ctx, int sensor, int precision, int* tile) {
  if (!ctx || !tile || sensor < 0 || sensor >= (int)ctx->sensors.size()) return CRTM_FAILURE;
  if (precision != CRTM_PREC_FP64 && precision != CRTM_PREC_FP32_SPLIT) return CRTM_FAILURE;
  *tile = ctx->sensors[sensor]->tile[precision == CRTM_PREC_FP64 ? 0 : 1];
  return CRTM_SUCCESS;
}

int crtm_autotune_k_matrix(crtm_context* ctx, const crtm_atmosphere* atm, const crtm_surface* sfc,
                           const crtm_geometry* geo, const crtm_options* opt, crtm_rtsolution* rts,
                           crtm_k_output* k, int repeats, double* ms, int* best) {
  if (!ctx) return fail(nullptr, "null context");
  if (!opt) return fail(ctx, "null options");
  if (repeats < 1 || repeats > 100) return fail(ctx, fmt("autotune repeats %d: 1..100", repeats));
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(ctx->stream, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone)
    return fail(ctx, "crtm_autotune_k_matrix cannot run under graph capture (it synchronizes to read timers)");
  const int ns = (int)ctx->sensors.size();
  const int pi = opt->precision == CRTM_PREC_FP64 ? 0 : 1;
  std::vector<int> saved(ns);
  for (int i = 0; i < ns; ++i) saved[i] = ctx->sensors[i]->tile[pi];
  const bool was_profiling = ctx->profiling;
  std::vector<double> t((size_t)ns * N_TILES, 0.0);
  std::vector<int> cnt((size_t)ns * N_TILES, 0);
  auto restore = [&] {
    for (int i = 0; i < ns; ++i) ctx->sensors[i]->tile[pi] = saved[i];
    ctx->profiling = was_profiling;
    ctx->ev_used = 0;
    ctx->ev_kind.clear();
    ctx->ev_sensor.clear();
  };
  for (int tl = 0; tl < N_TILES; ++tl) {
    for (int i = 0; i < ns; ++i) ctx->sensors[i]->tile[pi] = tl;
    for (int r = 0; r <= repeats; ++r) {   // the first run of each tile is not timed
      if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
        restore();
        return fail(ctx, "stream synchronize failed");
      }
      ctx->profiling = r > 0;
      ctx->ev_used = 0;
      ctx->ev_kind.clear();
      ctx->ev_sensor.clear();
      const int rc = crtm_k_matrix(ctx, atm, sfc, geo, opt, rts, k);
      if (rc != CRTM_SUCCESS) {
        restore();
        return rc;
      }
      if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
        restore();
        return fail(ctx, "stream synchronize failed");
      }
      for (size_t e = 0; r > 0 && e < ctx->ev_kind.size(); ++e) {
        if (ctx->ev_kind[e] != 1 || ctx->ev_sensor[e] < 0) continue;
        float el = 0.f;
        if (hipEventElapsedTime(&el, ctx->ev[2 * e], ctx->ev[2 * e + 1]) != hipSuccess) {
          restore();
          return fail(ctx, "hipEventElapsedTime failed");
        }
        t[(size_t)ctx->ev_sensor[e] * N_TILES + tl] += el;
        cnt[(size_t)ctx->ev_sensor[e] * N_TILES + tl] += 1;
      }
    }
  }
  restore();
  for (int i = 0; i < ns; ++i) {
    int b = saved[i];
    double bt = -1.0;
    for (int tl = 0; tl < N_TILES; ++tl) {
      const size_t q = (size_t)i * N_TILES + tl;
      const double v = cnt[q] ? t[q] / cnt[q] : -1.0;
      if (ms) ms[q] = v;
      if (v >= 0.0 && (bt < 0.0 || v < bt)) {
        bt = v;
        b = tl;
      }
    }
    ctx->sensors[i]->tile[pi] = b;
    if (best) best[i] = b;
  }
  // leave the outputs computed with the chosen tiles
  return crtm_k_matrix(ctx, atm, sfc, geo, opt, rts, k);
}

int crtm_set_wavefront_angles(crtm_context* ctx, int min_angles) {
  if (!ctx) return CRTM_FAILURE;
  if (min_angles != 6 && min_angles != 9)
    return fail(ctx, fmt("wavefront angle threshold %d: 6 or 9", min_angles));
  ctx->wide_min = min_angles;
  return CRTM_SUCCESS;
}

int crtm_set_cloudy_layer_capacity(crtm_context* ctx, int64_t max_layers) {
  if (!ctx) return CRTM_FAILURE;
  if (max_layers < 0 || max_layers > INT32_MAX)
    return fail(ctx, fmt("cloudy layer capacity %lld: 0 (off) .. %d", (long long)max_layers, INT32_MAX));
  ctx->cloud_cap = max_layers;
  return CRTM_SUCCESS;
}

int crtm_set_stream(crtm_context* ctx, void* stream, int own_stream) {
  if (!ctx) return CRTM_FAILURE;
  hipStream_t next = own_stream ? ctx->own_stream : (hipStream_t)stream;
  if (next == ctx->stream) return CRTM_SUCCESS;
  // work already issued on the previous stream completes before the switch
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return fail(ctx, hipGetErrorString(e));
  ctx->stream = next;
  ctx->ev_used = 0;
  ctx->ev_kind.clear();
  ctx->ev_sensor.clear();
  return CRTM_SUCCESS;
}

int crtm_synchronize(crtm_context* ctx) {
  if (!ctx) return CRTM_FAILURE;
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return fail(ctx, hipGetErrorString(e));
  return CRTM_SUCCESS;
}

const char* crtm_last_error(const crtm_context* ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

int crtm_profile_enable(crtm_context* ctx, int on) {
  if (!ctx) return CRTM_FAILURE;
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail(ctx, "stream synchronize failed");
  ctx->profiling = on != 0;
  ctx->ev_used = 0;
  ctx->ev_kind.clear();
  ctx->ev_sensor.clear();
  return CRTM_SUCCESS;
}

int crtm_profile_query_kinds(crtm_context* ctx, double* ms, int* n_launches, int n_kinds) {
  if (!ctx) return CRTM_FAILURE;
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail(ctx, "stream synchronize failed");
  double t[4] = {0.0, 0.0, 0.0, 0.0};
  int n[4] = {0, 0, 0, 0};
  for (size_t i = 0; i < ctx->ev_kind.size(); ++i) {
    float e = 0.f;
    if (hipEventElapsedTime(&e, ctx->ev[2 * i], ctx->ev[2 * i + 1]) != hipSuccess)
      return fail(ctx, "hipEventElapsedTime failed");
    t[ctx->ev_kind[i]] += e;
    n[ctx->ev_kind[i]] += 1;
  }
  for (int q = 0; q < n_kinds && q < 4; ++q) {
    if (ms) ms[q] = t[q];
    if (n_launches) n_launches[q] = n[q];
  }
  ctx->ev_used = 0;
  ctx->ev_kind.clear();
  ctx->ev_sensor.clear();
  return CRTM_SUCCESS;
}

int crtm_profile_query(crtm_context* ctx, double* prep_ms, double* main_ms, int* n_launches) {
  double t[3];
  int n[3];
  const int rc = crtm_profile_query_kinds(ctx, t, n, 3);
  if (rc != CRTM_SUCCESS) return rc;
  if (prep_ms) *prep_ms = t[0];
  if (main_ms) *main_ms = t[1];
  if (n_launches) *n_launches = n[1];
  return CRTM_SUCCESS;
}

}  // extern "C"
