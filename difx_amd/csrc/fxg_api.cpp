// Host layer of libfxg: the C ABI of include/fxg.h.
//
// fxg_create plays the part of Configuration (validation, derived quantities, the result
// layout of populateResultLengths, configuration.cpp:2348-2630) and of the Core constructor
// (device-resident buffers, core.cpp:23-154).  fxg_process_subint_device plays
// Core::processdata (core.cpp:657-1129): per subint it runs the setup kernel (delay
// interpolators and frame validity), then per chunk of FFTs the fused F-engine and the
// X-engine, then the accumulation into the slot's results (core.cpp:1166-1376, 1605-1938).
#include "fxg_internal.h"
#include "../../include/fxg.h"

#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <numeric>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

namespace {

constexpr double TWO_PI = 6.283185307179586476925286766559;  // architecture.h.in:430
constexpr double TINY = 0.000000001;                          // mode.cpp:27
constexpr int FLAGS_PER_INT = 30;                             // mpifxcorr.h:29
constexpr int INVALID_SUBINT = -99999999;                     // mode.h:222
constexpr int RING = 4;                                       // RECEIVE_RING_LENGTH, core.cpp:156

// Two freqs covering the same sky range (lower band edge = edge - bandwidth for LSB) with the
// same bandwidth, channels and averaging (configuration.cpp:3108-3135)
bool same_sky_band(const fxg_freq& a, const fxg_freq& b) {
  const double la = a.freq_mhz - (a.lsb ? a.bw_mhz : 0.0), lb = b.freq_mhz - (b.lsb ? b.bw_mhz : 0.0);
  return std::fabs(la - lb) < TINY && std::fabs(a.bw_mhz - b.bw_mhz) < TINY && a.nchan == b.nchan &&
         a.chanavg == b.chanavg;
}

// configuration.cpp:38-51
int calcstridelength(int n) {
  if (n <= 16) return n;
  for (int g = (int)std::sqrt(n + 0.1); g > 0; --g)
    if (n % g == 0) return n / g;
  return n;
}

bool is_pow2(long long v) { return v > 0 && (v & (v - 1)) == 0; }

long long gcd_ll(long long a, long long b) {
  while (b) { const long long t = a % b; a = b; b = t; }
  return a;
}

struct Slot {
  hipStream_t copy = nullptr;
  hipEvent_t copied = nullptr, done = nullptr;
  std::vector<uint8_t*> d_data;
  std::vector<long long> cap;
  float2* d_results = nullptr;
  float* h_results = nullptr;     // pinned
  bool busy = false;
};

struct Staging {
  DevDsRun* h_run = nullptr;      // pinned
  int32_t* h_flags = nullptr;     // pinned
  double* h_uvdd = nullptr;       // pinned: uv-shift delays [interval][baseline][phase centre]
  float* h_uvdec = nullptr;       // pinned: shift decorrelation sums of the subint
  double* d_uvdd = nullptr;
  float* d_uvdec = nullptr;
  hipEvent_t used = nullptr;
  bool pending = false;
};

}  // namespace

struct fxg_ctx {
  std::string err;
  int device = 0;
  bool dry = false;               // size accounting only: no device allocation or upload
  long long dev_bytes = 0;        // device memory held by this context
  // configuration copy
  std::vector<fxg_freq> freqs;
  std::vector<fxg_datastream> dss;
  std::vector<fxg_baseline> bls;
  std::vector<std::vector<double>> models;
  std::vector<std::vector<double>> pcmodels;   // [ds] phase centres 1..P: [P][nsamples][POLY]
  long long subint_ns = 0;
  int guard_ns = 0, rotorder = 1, nbuffered = 10, crossautos = 1, xengine = 0;
  double model_inc = 120.0, scan_start = 0.0;
  // derived
  int nds = 0, nfreqs = 0, nbl = 0;
  int B = 0, flaglen = 0;
  // FFT-length classes: datastreams sharing (L, sampling) run in one F-engine launch
  struct FClass {
    int L = 0, cplx = 0;
    std::vector<int> dsl;        // datastream indices
    long long tw_off = 0;        // offset of this L's twiddle tables in d_twiddle
    long long dsl_off = 0;       // offset of dsl in d_dslist
    int phcap = 0;               // phasor-table entries: max over dsl of S + NS and S + NFS
    bool big = false;            // L > FXG_MAX_L: four-step F-engine
    long long twL_off = 0;       // big: exp(-2 pi i j / L), j < L
    int wgx = 1;                 // big: workgroups per datastream (scratch rows)
  };
  float2* d_bigscratch = nullptr;
  int bigL = 0;
  std::vector<FClass> fclasses;
  int* d_dslist = nullptr;
  // F-engine launch form (per FFT length, or one persistent launch per sampling kind)
  struct FChoice {
    int mode = FXG_F_PER_LENGTH;
    int ntb = 0;                 // workgroup size; 0 = 256 below L = 4096, 512 from there
    int grid_mult = 1;           // persistent: workgroups per CU
  };
  FChoice fsel;
  int cfg_fmode = 0, cfg_fthreads = 0, cfg_autotune = 0;
  std::vector<std::pair<std::string, double>> tune_log;   // candidate, ms per subint (< 0: failed)
  std::vector<FItem> items[2];   // persistent work items: [0] real, [1] complex sampling
  FItem* d_items[2] = {nullptr, nullptr};
  std::vector<IClass> iclass[2];  // their FFT lengths, longest first, with item ranges
  IClass* d_iclass[2] = {nullptr, nullptr};
  int* d_counter = nullptr;
  int ncu = 256;
  std::vector<long long> databytes;
  int bpa = 1;
  long long coreresultlength = 0;
  std::vector<int> boff, bwoff, acoff, acwoff;
  // multi-phase-centre uv shift (core.cpp:1605-1938): P > 1 phase centres
  int nphase = 1, nthreads = 1;
  long long maxns_xc = 0;
  int xms = 0, rs = 0;             // xmac and rotate stride lengths (configuration.cpp:2653-2734)
  std::vector<int> sdoff;          // [freq][baseline] cf32 offset of the shift-decorrelation words
  struct Interval { int i0, n; double nsoff, nswidth; };
  std::vector<Interval> ivals;     // uv-shift intervals of a subint
  std::vector<UvFreq> uvfreq;      // per freq: rotation constants
  UvFreq* d_uvfreq = nullptr;
  std::vector<int> uvdec_idx;      // float index in results of each shift-decorrelation sum
  int* d_uvdec_idx = nullptr;
  std::vector<int> used, equiv;
  std::vector<int> aclabel;        // [freq] freq index of its autocorrelation records
  // device tables
  std::vector<DevDs> h_ds;
  DevDs* d_ds = nullptr;
  DevDsRun* d_run = nullptr;
  double* d_model = nullptr;
  float* d_chanfreq = nullptr;
  uint8_t* d_fvalid = nullptr;
  int2* d_fzone = nullptr;                 // per frame: decoding samples of a partial frame
  long long fvalid_total = 0;
  int maxframes = 0;
  float2* d_twiddle = nullptr;
  FftDelay* d_params = nullptr;
  float* d_weights = nullptr;
  int32_t* d_flags = nullptr;
  // chunk buffers
  int chunkcap = 0;
  float2* d_spec32 = nullptr;
  uint32_t* d_spec16 = nullptr;
  long long spec_elems = 0;
  // plain X
  std::vector<DevProd> prods;
  DevProd* d_prods = nullptr;
  // blocked plain X-engine: groups of rows, their pair tables and work items; the products it
  // does not take (rows of groups above 64, repeated row pairs) stay with the per-product kernel
  std::vector<XPGroup> xp_groups;
  std::vector<XPRow> xp_rows;
  std::vector<int> xp_pairs;
  std::vector<XPItem> xp_items;
  std::vector<DevProd> prods_rest;
  XPGroup* d_xp_groups = nullptr;
  XPRow* d_xp_rows = nullptr;
  int* d_xp_pairs = nullptr;
  XPItem* d_xp_items = nullptr;
  DevProd* d_prods_rest = nullptr;
  float2* d_xacc = nullptr;
  long long xacc_len = 0;
  int maxchan = 0;
  // MFMA X
  std::vector<DevXGroup> groups;
  std::vector<std::vector<std::pair<int, int>>> group_rows;   // (ds, band) per row
  std::map<std::pair<int, int>, std::pair<int, int>> rowof;   // (ds, recband) -> (group, row)
  DevXGroup* d_groups = nullptr;
  float* d_vacc = nullptr;
  long long vacc_len = 0;
  float* d_xscale = nullptr;
  long long xscale_len = 0;
  // accumulation entries
  std::vector<FinPlain> finp;
  std::vector<FinMfma> finm;
  std::vector<FinWeight> finw;
  FinPlain* d_finp = nullptr;
  FinMfma* d_finm = nullptr;
  FinWeight* d_finw = nullptr;
  int maxoutp = 0, maxoutm = 0;
  // amplitude scaling (Visibility::writedata)
  std::vector<NormSeg> normseg;
  std::vector<NormCal> normcal;     // + one zero entry at the end (bands without a mean)
  std::vector<int> normcal_of;      // [ds][total band] -> normcal index (zero entry if none)
  std::vector<int> acw_of;          // [ds][total band] -> float index of its parallel autocorr weight
  int maxbands = 0;
  NormSeg* d_normseg = nullptr;
  NormCal* d_normcal = nullptr;
  float* d_calmean = nullptr;
  double* d_tsys = nullptr;
  float* d_dpct = nullptr;
  // spectral kurtosis sums (off until fxg_set_kurtosis)
  bool kurt_on = false;
  std::vector<KurtBand> kurt_tab;
  std::vector<int> kurt_index;     // [ds] first kurt_tab entry of the datastream
  KurtBand* d_kurt_tab = nullptr;
  float2* d_kurt_acc = nullptr;
  long long kurt_len = 0;          // float2 sums
  int kurt_maxC = 0;
  // phase-cal extraction (off until fxg_set_pcal)
  struct PcalTones {
    int ds, band;
    long long offset_hz;           // baseband frequency of tone 0 (-1: no tone in the band)
    long long interval_hz, fs_hz;
    int ntones;
    int tone0;                     // first entry of the band's tones in the device tone buffer
  };
  std::vector<long long> pcal_interval, pcal_base;   // [ds]; interval 0 = off
  std::vector<PcalTones> pcal_tones;
  struct PcalBins { int M; int bins_off; };
  std::vector<PcalBins> pcal_tab;  // bands with tones (bins on the device)
  std::vector<int> pcal_entry;     // pcal_tones index -> pcal_tab index or -1
  std::vector<PcalGroup> pcal_groups;
  PcalGroup* d_pcal_groups = nullptr;
  float* d_pcal_part = nullptr;    // per-workgroup partial bins
  float* d_pcal_bins = nullptr;
  std::vector<std::array<int, 3>> pcal_runs;   // {form, first group, groups}
  std::vector<PcalToneDev> pcal_tdev;           // every tone an extractor reads
  PcalToneDev* d_pcal_tdev = nullptr;
  float2* d_pcal_tonebuf = nullptr;             // the last subint's tones of every band
  int pcal_ntone_total = 0;
  std::vector<long long> cfg_pcal_interval, cfg_pcal_base;   // [ds] of the configuration
  std::vector<int> pcaloff;                     // [ds] cf32 offset of the results' phase-cal block
  long long pcal_part_len = 0;
  int pcal_max_gsize = 0, pcal_nsplit = 1;
  // streams, staging, ring
  hipStream_t stream = nullptr;
  Staging stage[2];
  int stage_next = 0;
  Slot slots[RING];
  int last_chunk0 = 0, last_nffts = 0;
  std::mutex mu;
  // optional launch timing: (category, start, stop) event triples
  bool timing = false;
  std::vector<std::tuple<int, hipEvent_t, hipEvent_t>> tev;
  std::vector<hipEvent_t> evpool;
};

namespace {

// The message of this thread's last failed create (the context it belonged to is gone).
std::string& create_error() {
  static thread_local std::string msg;
  return msg;
}

int fail(fxg_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(call)                                                                        \
  do {                                                                                      \
    hipError_t _e = (call);                                                                 \
    if (_e != hipSuccess)                                                                   \
      return fail(ctx, FXG_E_HIP, std::string(#call) + ": " + hipGetErrorString(_e));       \
  } while (0)

// Every device allocation of a context goes through here, so that a dry run (fxg_footprint)
// reports exactly what fxg_create would hold.
int dev_malloc(fxg_ctx* ctx, void** p, long long bytes) {
  bytes = (std::max<long long>(bytes, 16) + 255) / 256 * 256;
  ctx->dev_bytes += bytes;
  if (ctx->dry) { *p = nullptr; return FXG_OK; }
  HIPCHK(hipMalloc(p, (size_t)bytes));
  return FXG_OK;
}

int dev_zero(fxg_ctx* ctx, void* p, long long bytes) {
  if (ctx->dry) return FXG_OK;
  HIPCHK(hipMemset(p, 0, (size_t)bytes));
  return FXG_OK;
}

template <typename T>
int dev_upload(fxg_ctx* ctx, T** dptr, const std::vector<T>& h) {
  if (int rc = dev_malloc(ctx, (void**)dptr, (long long)(h.size() * sizeof(T)))) return rc;
  if (!ctx->dry && !h.empty()) HIPCHK(hipMemcpy(*dptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return FXG_OK;
}

int validate_and_derive(fxg_ctx* ctx) {
  if (ctx->rotorder < 0 || ctx->rotorder > 2)
    return fail(ctx, FXG_E_INVALID, "fringe rotation order must be 0, 1 or 2");
  if (ctx->nds < 1) return fail(ctx, FXG_E_INVALID, "no datastreams");
  if (ctx->nfreqs < 1) return fail(ctx, FXG_E_INVALID, "no freqs");
  // every table index is checked before any lookup through it
  for (int d = 0; d < ctx->nds; ++d) {
    const fxg_datastream& ds = ctx->dss[d];
    if (ds.nrecfreqs < 1 || ds.nrecfreqs > FXG_MAX_FREQS || ds.nrecbands < 1 || ds.nrecbands > FXG_MAX_BANDS)
      return fail(ctx, FXG_E_INVALID, "bad recorded freq/band count");
    if (ds.nzoomfreqs < 0 || ds.nzoomfreqs > FXG_MAX_FREQS || ds.nzoombands < 0 || ds.nzoombands > FXG_MAX_BANDS ||
        ds.nrecbands + ds.nzoombands > FXG_MAX_BANDS)
      return fail(ctx, FXG_E_INVALID, "bad zoom freq/band count");
    for (int k = 0; k < ds.nrecfreqs; ++k)
      if (ds.recfreq_index[k] < 0 || ds.recfreq_index[k] >= ctx->nfreqs)
        return fail(ctx, FXG_E_INVALID, "recorded freq index out of range");
    if (ds.phasecal_interval_hz < 0) return fail(ctx, FXG_E_INVALID, "negative phase-cal interval");
    for (int k = 0; k < ds.nzoomfreqs; ++k)
      if (ds.zoomfreq_index[k] < 0 || ds.zoomfreq_index[k] >= ctx->nfreqs)
        return fail(ctx, FXG_E_INVALID, "zoom freq index out of range");
    for (int b = 0; b < ds.nrecbands; ++b)
      if (ds.recband_localfreq[b] < 0 || ds.recband_localfreq[b] >= ds.nrecfreqs)
        return fail(ctx, FXG_E_INVALID, "band local freq out of range");
    for (int z = 0; z < ds.nzoombands; ++z)
      if (ds.zoomband_localfreq[z] < 0 || ds.zoomband_localfreq[z] >= ds.nzoomfreqs)
        return fail(ctx, FXG_E_INVALID, "zoom band freq out of range");
    for (int k = 0; k < ds.nrecfreqs; ++k)
      if (!std::isfinite(ds.phase_offset_deg[k]) || !std::isfinite(ds.clock_offset_us[k]) ||
          !std::isfinite(ds.clock_offset_delta_us[k]) || !std::isfinite(ds.lo_offset_hz[k]))
        return fail(ctx, FXG_E_INVALID, "non-finite per-freq offset");
  }
  // blockspersend from the first datastream's first recorded freq (configuration.cpp:1843-1852)
  {
    const fxg_freq& f = ctx->freqs[ctx->dss[0].recfreq_index[0]];
    const double ffttime = 1000.0 * f.nchan / f.bw_mhz;
    const double bd = ctx->subint_ns / ffttime;
    ctx->B = (int)(bd + 0.5);
    if (std::fabs(bd - ctx->B) > TINY)
      return fail(ctx, FXG_E_INVALID, "subint is not an integer number of FFTs");
  }
  ctx->flaglen = (ctx->B + FLAGS_PER_INT - 1) / FLAGS_PER_INT;
  for (int d = 0; d < ctx->nds; ++d) {
    const fxg_datastream& ds = ctx->dss[d];
    if (ds.format != FXG_FMT_VDIF && ds.format != FXG_FMT_MARK5B)
      return fail(ctx, FXG_E_INVALID, "unsupported data format");
    if (ds.complex_sampling < 0 || ds.complex_sampling > 2)
      return fail(ctx, FXG_E_INVALID, "complex_sampling must be 0 (real), 1 (complex) or 2 (double sideband)");
    if (!((ds.nbits == 2 && !ds.complex_sampling) || ds.nbits == 8))
      return fail(ctx, FXG_E_INVALID, "only 2-bit real and 8-bit real/complex sampling are supported");
    if (ds.complex_sampling && ds.format != FXG_FMT_VDIF)
      return fail(ctx, FXG_E_INVALID, "complex sampling requires VDIF");
    if (ds.complex_sampling && ctx->rotorder == 0)       // mode.cpp:1130-1133
      return fail(ctx, FXG_E_INVALID, "post-F (order 0) fringe rotation of complex data is not supported");
    if (ds.nrecfreqs < 1 || ds.nrecfreqs > FXG_MAX_FREQS || ds.nrecbands < 1 || ds.nrecbands > FXG_MAX_BANDS)
      return fail(ctx, FXG_E_INVALID, "bad recorded freq/band count");
    if (ds.nbits == 2 && !is_pow2(ds.nrecbands))
      return fail(ctx, FXG_E_INVALID, "2-bit data needs 1, 2, 4, 8 or 16 bands per sample word");
    if (ds.nbits == 2 && ds.nrecbands > 16)
      return fail(ctx, FXG_E_INVALID, "at most 16 2-bit bands per sample word");
    const fxg_freq& f0 = ctx->freqs[ds.recfreq_index[0]];
    for (int k = 0; k < ds.nrecfreqs; ++k) {
      int fi = ds.recfreq_index[k];
      if (fi < 0 || fi >= ctx->nfreqs) return fail(ctx, FXG_E_INVALID, "recorded freq index out of range");
      const fxg_freq& f = ctx->freqs[fi];
      if (f.nchan != f0.nchan || std::fabs(f.bw_mhz - f0.bw_mhz) > TINY || f.chanavg != f0.chanavg)
        return fail(ctx, FXG_E_INVALID, "a datastream's recorded freqs must share nchan, bandwidth, averaging");
      if (ds.complex_sampling == 2 && f.nchan % 16)
        return fail(ctx, FXG_E_INVALID, "double-sideband data needs nchan to be a multiple of 16");
      if (f.chanavg < 1 || f.nchan % f.chanavg)
        return fail(ctx, FXG_E_INVALID, "channel averaging must divide nchan");
    }
    const int L = ds.complex_sampling ? f0.nchan : 2 * f0.nchan;
    if (!is_pow2(L) || L < 256 || L > FXG_MAX_BIG_L)
      return fail(ctx, FXG_E_INVALID, "FFT length must be a power of two in [256, 32768]");
    // every datastream: a whole number of nanoseconds per FFT and of FFTs per subint
    // (configuration.cpp:3008-3028)
    {
      const double ffttime = 1000.0 * f0.nchan / f0.bw_mhz;
      const double nffts = ctx->subint_ns / ffttime;
      if (std::fabs(ffttime - (int)(ffttime + 0.5)) > 1e-8)
        return fail(ctx, FXG_E_INVALID, "FFT time is not a whole number of nanoseconds");
      if (std::fabs(nffts - (int)(nffts + 0.5)) > TINY)
        return fail(ctx, FXG_E_INVALID, "subint is not an integer number of FFTs for a datastream");
    }
    int S = ds.arraystridelen ? ds.arraystridelen : calcstridelength(f0.nchan);
    const int C = f0.nchan;
    const int phcap = std::min(L / 4, 512);        // LDS phasor tables of the F-engine
    if (!is_pow2(S) || (L / 8) % S || S + L / S > phcap || S + C / S > phcap)
      return fail(ctx, FXG_E_INVALID, "unsupported array stride length");
    const int pay = ds.framebytes - ds.headerbytes;
    const int bits = ds.nbits * (ds.complex_sampling ? 2 : 1) * ds.nrecbands;
    if (pay <= 16 || pay % 8 || (pay * 8) % bits || ds.framebytes % 4 || ds.headerbytes % 4)
      return fail(ctx, FXG_E_INVALID, "frame size inconsistent with the sample format");
    if ((long long)pay * 8 / bits > (1LL << 20))      // the F-engine's frame index arithmetic
      return fail(ctx, FXG_E_INVALID, "more than 2^20 samples per frame");
    if (L > FXG_MAX_L && pay * 8 / bits < L / 2048)   // four-step: R consecutive samples cross
      return fail(ctx, FXG_E_INVALID, "frames too short for the FFT length");   // one frame edge at most
    int nb_per_freq[FXG_MAX_FREQS] = {0};
    for (int b = 0; b < ds.nrecbands; ++b) {
      int lf = ds.recband_localfreq[b];
      if (lf < 0 || lf >= ds.nrecfreqs) return fail(ctx, FXG_E_INVALID, "band local freq out of range");
      if (++nb_per_freq[lf] > 4) return fail(ctx, FXG_E_INVALID, "at most 4 bands per recorded freq");
    }
    for (int z = 0; z < ds.nzoombands; ++z) {
      int lz = ds.zoomband_localfreq[z];
      if (lz < 0 || lz >= ds.nzoomfreqs) return fail(ctx, FXG_E_INVALID, "zoom band freq out of range");
      const fxg_freq& zf = ctx->freqs[ds.zoomfreq_index[lz]];
      int parent = ds.zoomfreq_parent_localfreq[lz];
      if (parent < 0 || parent >= ds.nrecfreqs) return fail(ctx, FXG_E_INVALID, "zoom parent out of range");
      if (ds.zoomfreq_chanoffset[lz] < 0 || ds.zoomfreq_chanoffset[lz] + zf.nchan > C)
        return fail(ctx, FXG_E_INVALID, "zoom window outside its parent band");
      if (zf.chanavg != f0.chanavg || ds.zoomfreq_chanoffset[lz] % f0.chanavg)
        return fail(ctx, FXG_E_INVALID, "zoom averaging must match its parent's");
      bool found = false;
      for (int b = 0; b < ds.nrecbands; ++b)
        if (ds.recband_localfreq[b] == parent && ds.recband_pol[b] == ds.zoomband_pol[z]) found = true;
      if (!found) return fail(ctx, FXG_E_INVALID, "zoom band without a parent band");
    }
    if (ds.model_order < 0 || ds.model_order >= FXG_MAX_POLY || ds.model_nsamples < 1 || !ds.model_delay_us)
      return fail(ctx, FXG_E_INVALID, "bad delay model");
  }
  for (int j = 0; j < ctx->nbl; ++j) {
    const fxg_baseline& b = ctx->bls[j];
    if (b.ds1 < 0 || b.ds1 >= ctx->nds || b.ds2 < 0 || b.ds2 >= ctx->nds)
      return fail(ctx, FXG_E_INVALID, "baseline datastream out of range");
    if (b.nfreqs < 0 || b.nfreqs > FXG_MAX_FREQS) return fail(ctx, FXG_E_INVALID, "bad baseline freq count");
    for (int k = 0; k < b.nfreqs; ++k) {
      if (b.freq_index[k] < 0 || b.freq_index[k] >= ctx->nfreqs)
        return fail(ctx, FXG_E_INVALID, "baseline freq out of range");
      if (b.npolprod[k] < 1 || b.npolprod[k] > FXG_MAX_POLPROD)
        return fail(ctx, FXG_E_INVALID, "bad polarisation product count");
      const fxg_freq& f = ctx->freqs[b.freq_index[k]];
      for (int p = 0; p < b.npolprod[k]; ++p) {
        const int bands[2] = {b.ds1band[k][p], b.ds2band[k][p]};
        const int dsi[2] = {b.ds1, b.ds2};
        for (int s = 0; s < 2; ++s) {
          const fxg_datastream& ds = ctx->dss[dsi[s]];
          int bi = bands[s];
          if (bi < 0 || bi >= ds.nrecbands + ds.nzoombands)
            return fail(ctx, FXG_E_INVALID, "baseline band index out of range");
          int fi = bi < ds.nrecbands ? ds.recfreq_index[ds.recband_localfreq[bi]]
                                     : ds.zoomfreq_index[ds.zoomband_localfreq[bi - ds.nrecbands]];
          // USB x LSB (configuration.cpp:3125-3135): the second datastream's recorded band may
          // be the opposite sideband of the same sky band, real or single-sideband complex data,
          // against a recorded band of the first
          const bool mixed = s == 1 && bi < ds.nrecbands && bands[0] < ctx->dss[b.ds1].nrecbands &&
                             ds.complex_sampling != 2 && ctx->dss[b.ds1].complex_sampling != 2 &&
                             same_sky_band(ctx->freqs[fi], f) && ctx->freqs[fi].lsb != f.lsb;
          if (fi != b.freq_index[k] && !mixed)
            return fail(ctx, FXG_E_INVALID, "baseline products must correlate bands of the baseline freq "
                                            "(or, on the second datastream, its opposite sideband)");
        }
      }
      if (f.nchan % f.chanavg) return fail(ctx, FXG_E_INVALID, "channel averaging must divide nchan");
    }
  }
  if (ctx->xengine != FXG_X_FP32)
    for (int d = 0; d < ctx->nds; ++d) {
      const fxg_datastream& ds = ctx->dss[d];
      if (ctx->freqs[ds.recfreq_index[0]].nchan % 8)
        return fail(ctx, FXG_E_INVALID, "the MFMA X-engine needs nchan to be a multiple of 8");
      for (int z = 0; z < ds.nzoomfreqs; ++z)
        if (ctx->freqs[ds.zoomfreq_index[z]].nchan % 8)
          return fail(ctx, FXG_E_INVALID, "the MFMA X-engine needs zoom nchan to be a multiple of 8");
    }
  if (ctx->xengine < FXG_X_FP32 || ctx->xengine > FXG_X_BF16)
    return fail(ctx, FXG_E_INVALID, "unknown X-engine");
  if (ctx->cfg_fmode < FXG_F_AUTO || ctx->cfg_fmode > FXG_F_PERSISTENT)
    return fail(ctx, FXG_E_INVALID, "unknown F-engine launch form");
  if (ctx->cfg_fthreads != 0 && ctx->cfg_fthreads != 128 && ctx->cfg_fthreads != 256 && ctx->cfg_fthreads != 512)
    return fail(ctx, FXG_E_INVALID, "F-engine workgroup size must be 128, 256 or 512");
  if (ctx->cfg_fthreads == 128 && ctx->cfg_fmode == FXG_F_PERSISTENT)
    return fail(ctx, FXG_E_INVALID, "the persistent F-engine launch runs 256 or 512 threads");
  if (ctx->nphase > FXG_MAX_PHASECENTRES)
    return fail(ctx, FXG_E_INVALID, "too many phase centres");
  if (ctx->nphase > 1) {
    for (int d = 0; d < ctx->nds; ++d)
      if (!ctx->dss[d].phasecentre_delay_us)
        return fail(ctx, FXG_E_INVALID, "phase-centre delay polynomials missing");
    if (ctx->nbuffered < 1) return fail(ctx, FXG_E_INVALID, "numBufferedFFTs must be positive");
    if (ctx->maxns_xc < 0) return fail(ctx, FXG_E_INVALID, "negative uv-shift interval");
  }
  return FXG_OK;
}

// Configuration::getDataBytes (configuration.cpp:794-820)
long long get_data_bytes(const fxg_ctx* ctx, int d) {
  const fxg_datastream& ds = ctx->dss[d];
  const fxg_freq& f = ctx->freqs[ds.recfreq_index[0]];
  long long valid = ((long long)ctx->B * ds.nrecbands * 2 * ds.nbits * f.nchan) / 8;
  valid += ((long long)(int)(ctx->guard_ns / (1000.0 / (f.bw_mhz * 2.0)) + 1.0) * ds.nrecbands * ds.nbits) / 8;
  const int pay = ds.framebytes - ds.headerbytes;
  long long nframes = valid / pay + 2;
  if (ds.format == FXG_FMT_MARK5B) nframes += nframes % 2;
  return nframes * ds.framebytes;
}

int freq_of_band(const fxg_datastream& ds, int bi) {
  return bi < ds.nrecbands ? ds.recfreq_index[ds.recband_localfreq[bi]]
                           : ds.zoomfreq_index[ds.zoomband_localfreq[bi - ds.nrecbands]];
}

// record band (parent for zooms) and channel offset of band bi
void band_source(const fxg_datastream& ds, int bi, int* rec, int* choff) {
  if (bi < ds.nrecbands) { *rec = bi; *choff = 0; return; }
  int z = bi - ds.nrecbands;
  int lz = ds.zoomband_localfreq[z];
  int parent = ds.zoomfreq_parent_localfreq[lz];
  for (int b = 0; b < ds.nrecbands; ++b)
    if (ds.recband_localfreq[b] == parent && ds.recband_pol[b] == ds.zoomband_pol[z]) { *rec = b; break; }
  *choff = ds.zoomfreq_chanoffset[lz];
}

// Configuration::populateResultLengths for one config, one phase centre, no pulsar bins
// (configuration.cpp:2348-2630) with the used/equivalent freq flags of :2270-2320.
// The tones of one recorded freq (Configuration's phase-cal table, configuration.cpp:1768-1835,
// for a whole-Hz interval): the first tone inside the band away from the LO edge, their count,
// and the baseband offset of that tone.
static void pcal_tones_of(double lofreq, double bw, bool lsb, long long interval, long long base, int* ntones,
                          long long* offset) {
  const double stepMHz = interval * 1e-6;
  long harmonic = (long)((lofreq - base * 1e-6) / stepMHz);
  double tonefreq = harmonic * stepMHz + base * 1e-6;
  int n = 0;
  if (lsb) {
    if (tonefreq >= lofreq) { tonefreq -= stepMHz; harmonic -= 1; }
    if (tonefreq >= lofreq - bw)
      while (tonefreq - n * stepMHz >= lofreq - bw) n++;
    *offset = n > 0 ? (long long)(1e6 * lofreq) - (long long)harmonic * interval : -1;
  } else {
    if (tonefreq <= lofreq) { tonefreq += stepMHz; harmonic += 1; }
    if (tonefreq <= lofreq + bw)
      while (tonefreq + n * stepMHz <= lofreq + bw) n++;
    *offset = n > 0 ? (long long)harmonic * interval - (long long)(1e6 * lofreq) : -1;
  }
  *ntones = n > 0 ? n : 1;                         // one placeholder tone when none fits
}

// Configuration::getOppositeSidebandFreqIndex (configuration.cpp:774-791): the last freq of the
// other sideband with the same bandwidth and channels whose band covers the same sky range
int opposite_sideband_freq(const fxg_ctx* ctx, int fi) {
  const fxg_freq& f1 = ctx->freqs[fi];
  int r = -1;
  for (int i = 0; i < ctx->nfreqs; ++i) {
    const fxg_freq& f = ctx->freqs[i];
    if (f.bw_mhz == f1.bw_mhz && f.lsb != f1.lsb && f.nchan == f1.nchan && f.chanavg == f1.chanavg) {
      if (f.lsb && f1.freq_mhz + f1.bw_mhz == f.freq_mhz) r = i;
      if (!f.lsb && f.freq_mhz + f.bw_mhz == f1.freq_mhz) r = i;
    }
  }
  return r;
}

// The freq index under which each freq's autocorrelations are written: with USB x LSB
// products anywhere, an LSB freq correlated against a USB one is written as its opposite
// sideband freq (visibility.cpp:912-925; correlatedagainstupper from the first product of
// each baseline freq, configuration.cpp:1131-1158)
void autocorr_labels(fxg_ctx* ctx) {
  const int nf = ctx->nfreqs;
  ctx->aclabel.resize(nf);
  for (int i = 0; i < nf; ++i) ctx->aclabel[i] = i;
  std::vector<char> against_upper(nf, 0);
  bool any = false;
  for (const auto& b : ctx->bls)
    for (int k = 0; k < b.nfreqs; ++k) {
      const int f1 = freq_of_band(ctx->dss[b.ds1], b.ds1band[k][0]);
      const int f2 = freq_of_band(ctx->dss[b.ds2], b.ds2band[k][0]);
      if (ctx->freqs[f1].lsb != ctx->freqs[f2].lsb) {
        any = true;
        against_upper[ctx->freqs[f1].lsb ? f1 : f2] = 1;
      }
    }
  if (!any) return;
  for (int i = 0; i < nf; ++i)
    if (ctx->freqs[i].lsb && against_upper[i]) {
      const int o = opposite_sideband_freq(ctx, i);
      if (o >= 0) ctx->aclabel[i] = o;
    }
}

void build_layout(fxg_ctx* ctx) {
  const int nf = ctx->nfreqs, nb = ctx->nbl;
  autocorr_labels(ctx);
  ctx->used.assign(nf, 0);
  ctx->equiv.assign(nf, 0);
  int maxprod = 0;
  for (const auto& b : ctx->bls)
    for (int k = 0; k < b.nfreqs; ++k) {
      ctx->used[b.freq_index[k]] = 1;
      maxprod = std::max(maxprod, b.npolprod[k]);
    }
  for (int j = 0; j < nf; ++j) {
    if (ctx->used[j]) continue;
    for (int k = 0; k < nf; ++k) {
      const fxg_freq &fj = ctx->freqs[j], &fk = ctx->freqs[k];
      double bwdiff = fj.bw_mhz - fk.bw_mhz, fdiff = fj.freq_mhz - fk.freq_mhz;
      if (fj.lsb) fdiff -= fj.bw_mhz;
      if (fk.lsb) fdiff += fk.bw_mhz;
      if (std::fabs(bwdiff) < TINY && std::fabs(fdiff) < TINY && fj.nchan == fk.nchan &&
          fj.chanavg == fk.chanavg && ctx->used[k])
        ctx->equiv[j] = 1;
    }
  }
  ctx->bpa = maxprod > 2 ? 2 : 1;
  long long idx = 0;
  ctx->boff.assign((size_t)nf * nb, -1);
  ctx->bwoff.assign((size_t)nf * nb, -1);
  for (int f = 0; f < nf; ++f) {
    if (!ctx->used[f]) continue;
    const fxg_freq& fr = ctx->freqs[f];
    for (int j = 0; j < nb; ++j)
      for (int k = 0; k < ctx->bls[j].nfreqs; ++k)
        if (ctx->bls[j].freq_index[k] == f) {
          ctx->boff[(size_t)f * nb + j] = (int)idx;
          idx += (long long)ctx->nphase * ctx->bls[j].npolprod[k] * ((fr.nchan + fr.chanavg - 1) / fr.chanavg);
        }
  }
  for (int f = 0; f < nf; ++f) {
    if (!ctx->used[f]) continue;
    for (int j = 0; j < nb; ++j)
      for (int k = 0; k < ctx->bls[j].nfreqs; ++k)
        if (ctx->bls[j].freq_index[k] == f) {
          ctx->bwoff[(size_t)f * nb + j] = (int)idx;
          idx += ctx->bls[j].npolprod[k] / 2 + ctx->bls[j].npolprod[k] % 2;
        }
  }
  ctx->sdoff.assign((size_t)nf * nb, -1);        // shift decorrelation factors (configuration.cpp:2549-2566)
  for (int f = 0; f < nf && ctx->nphase > 1; ++f) {
    if (!ctx->used[f]) continue;
    for (int j = 0; j < nb; ++j)
      for (int k = 0; k < ctx->bls[j].nfreqs; ++k)
        if (ctx->bls[j].freq_index[k] == f) {
          ctx->sdoff[(size_t)f * nb + j] = (int)idx;
          idx += ctx->nphase / 2 + ctx->nphase % 2;
        }
  }
  ctx->acoff.assign(ctx->nds, 0);
  ctx->acwoff.assign(ctx->nds, 0);
  for (int d = 0; d < ctx->nds; ++d) {
    const fxg_datastream& ds = ctx->dss[d];
    ctx->acoff[d] = (int)idx;
    for (int b = 0; b < ds.nrecbands + ds.nzoombands; ++b) {
      int fi = freq_of_band(ds, b);
      if (ctx->used[fi] || ctx->equiv[fi])
        idx += ctx->bpa * ctx->freqs[fi].nchan / ctx->freqs[fi].chanavg;
    }
  }
  for (int d = 0; d < ctx->nds; ++d) {
    const fxg_datastream& ds = ctx->dss[d];
    ctx->acwoff[d] = (int)idx;
    int toadd = 0;
    for (int b = 0; b < ds.nrecbands + ds.nzoombands; ++b) {
      int fi = freq_of_band(ds, b);
      if (ctx->used[fi] || ctx->equiv[fi]) toadd += ctx->bpa;
    }
    toadd = (toadd + 1) / 2;
    idx += std::max(toadd, 1);
  }
  // the phase-cal tones of datastreams with a phase-cal interval, per recorded freq its pols
  // times its tones (configuration.cpp:2611-2619)
  ctx->pcaloff.assign(ctx->nds, 0);
  for (int d = 0; d < ctx->nds; ++d) {
    const fxg_datastream& ds = ctx->dss[d];
    ctx->pcaloff[d] = (int)idx;
    if (ctx->cfg_pcal_interval[d] <= 0) continue;
    for (int j = 0; j < ds.nrecfreqs; ++j) {
      int npol = 0;
      for (int b = 0; b < ds.nrecbands; ++b) npol += ds.recband_localfreq[b] == j;
      const fxg_freq& f = ctx->freqs[ds.recfreq_index[j]];
      int nt;
      long long off;
      pcal_tones_of(f.freq_mhz, f.bw_mhz, f.lsb != 0, ctx->cfg_pcal_interval[d], ctx->cfg_pcal_base[d], &nt, &off);
      idx += (long long)npol * nt;
    }
  }
  ctx->coreresultlength = idx;
}

// For recorded band b of datastream ds: the partner band for cross-polarisation autos
// (first two matching bands of its freq, mode.cpp:1320-1352), or -1.
int crosspol_partner(const fxg_datastream& ds, int b);

// zoom bandwidth over its parent's, in the reference's float steps (visibility.cpp:525-530)
static float zoom_ratio(const fxg_ctx* ctx, const fxg_datastream& s, int band) {
  const int lz = s.zoomband_localfreq[band - s.nrecbands];
  float m = 1.0f;
  m = (float)(m * ctx->freqs[s.zoomfreq_index[lz]].bw_mhz);
  m = (float)(m / ctx->freqs[s.recfreq_index[s.zoomfreq_parent_localfreq[lz]]].bw_mhz);
  return m;
}

// Configuration::getDMatchingBand (configuration.cpp:885-904) for recorded bands: the first
// other band of the same local freq; -1 for zoom bands here
static int matching_band(const fxg_datastream& s, int band) {
  if (band >= s.nrecbands) return -1;
  for (int i = 0; i < s.nrecbands; ++i)
    if (i != band && s.recband_localfreq[i] == s.recband_localfreq[band]) return i;
  return -1;
}

// The segments of Visibility::writedata's scaling in the results layout: cross-correlations
// per (freq, baseline, product), autocorrelations per (datastream, product, band) when they
// are written, and the parallel-autocorrelation band means the Tsys calibration uses.
int build_norm(fxg_ctx* ctx) {
  const int nds = ctx->nds, nb = ctx->nbl, B = ctx->B;
  ctx->normseg.clear();
  ctx->normcal.clear();
  ctx->maxbands = 1;
  for (int d = 0; d < nds; ++d)
    ctx->maxbands = std::max(ctx->maxbands, ctx->dss[d].nrecbands + ctx->dss[d].nzoombands);
  const int MB = ctx->maxbands;
  ctx->normcal_of.assign((size_t)nds * MB, -1);
  ctx->acw_of.assign((size_t)nds * MB, -1);
  for (int d = 0; d < nds; ++d) {                  // j = 0 autocorrelation bands
    const fxg_datastream& s = ctx->dss[d];
    long long idx = ctx->acoff[d];
    int widx = ctx->acwoff[d] * 2;
    for (int k = 0; k < s.nrecbands + s.nzoombands; ++k) {
      const int fi = freq_of_band(s, k);
      if (!(ctx->used[fi] || ctx->equiv[fi])) continue;
      const int fc = ctx->freqs[fi].nchan / ctx->freqs[fi].chanavg;
      ctx->normcal_of[(size_t)d * MB + k] = (int)ctx->normcal.size();
      ctx->normcal.push_back(NormCal{(int)idx, fc});
      ctx->acw_of[(size_t)d * MB + k] = widx++;
      idx += fc;
    }
  }
  const int zero = (int)ctx->normcal.size();
  ctx->normcal.push_back(NormCal{0, 0});           // never computed: stays 0
  for (auto& c : ctx->normcal_of) if (c < 0) c = zero;
  for (int f = 0; f < ctx->nfreqs; ++f) {          // cross-correlations (visibility.cpp:486-586)
    if (!ctx->used[f]) continue;
    const fxg_freq& fr = ctx->freqs[f];
    const int fc = (fr.nchan + fr.chanavg - 1) / fr.chanavg;
    for (int j = 0; j < nb; ++j) {
      const fxg_baseline& b = ctx->bls[j];
      for (int k = 0; k < b.nfreqs; ++k) {
        if (b.freq_index[k] != f) continue;
        for (int p = 0; p < b.npolprod[k]; ++p)
        for (int pc = 0; pc < ctx->nphase; ++pc) {
          NormSeg g{};
          g.kind = 0;
          g.off = ctx->boff[(size_t)f * nb + j] + (pc * b.npolprod[k] + p) * fc;
          g.decf = ctx->nphase > 1 ? ctx->sdoff[(size_t)f * nb + j] * 2 + pc : -1;
          g.nout = fc;
          g.wf = ctx->bwoff[(size_t)f * nb + j] * 2 + p;
          g.ds1 = b.ds1; g.ds2 = b.ds2;
          const int b1 = b.ds1band[k][p], b2 = b.ds2band[k][p];
          const fxg_datastream& s1 = ctx->dss[b.ds1];
          const fxg_datastream& s2 = ctx->dss[b.ds2];
          g.cal1 = ctx->normcal_of[(size_t)b.ds1 * MB + b1];
          g.cal2 = ctx->normcal_of[(size_t)b.ds2 * MB + b2];
          g.aw1 = ctx->acw_of[(size_t)b.ds1 * MB + b1];   // -1: weight 0
          g.aw2 = ctx->acw_of[(size_t)b.ds2 * MB + b2];
          g.zs1 = b1 >= s1.nrecbands ? std::sqrt(zoom_ratio(ctx, s1, b1)) : 1.0f;
          g.zs2 = b2 >= s2.nrecbands ? std::sqrt(zoom_ratio(ctx, s2, b2)) : 1.0f;
          g.chanfactor = (float)(B * 2 * fc * fr.chanavg);
          ctx->normseg.push_back(g);
        }
      }
    }
  }
  if (ctx->crossautos) {                           // autocorrelations (visibility.cpp:588-646)
    for (int d = 0; d < nds; ++d) {
      const fxg_datastream& s = ctx->dss[d];
      long long idx = ctx->acoff[d];
      int widx = ctx->acwoff[d] * 2;
      for (int w = 0; w < ctx->bpa; ++w) {
        for (int k = 0; k < s.nrecbands + s.nzoombands; ++k) {
          const int fi = freq_of_band(s, k);
          if (!(ctx->used[fi] || ctx->equiv[fi])) continue;
          const fxg_freq& fr = ctx->freqs[fi];
          const int fc = fr.nchan / fr.chanavg;
          NormSeg g{};
          g.kind = 1;
          g.decf = -1;
          g.off = (int)idx;
          g.nout = fc;
          g.wf = widx;
          g.ds1 = g.ds2 = d;
          g.cal1 = ctx->normcal_of[(size_t)d * MB + k];
          const int mk = w == 0 ? k : matching_band(s, k);
          g.cal2 = mk >= 0 ? ctx->normcal_of[(size_t)d * MB + mk] : zero;
          g.zs1 = k >= s.nrecbands ? zoom_ratio(ctx, s, k) : 1.0f;
          g.zs2 = 1.0f;
          g.chanfactor = (float)(B * 2 * fc * fr.chanavg);
          ctx->normseg.push_back(g);
          idx += fc;
          widx++;
        }
      }
    }
  }
  for (int d = 0; d < nds; ++d) {                  // phase-cal tones (visibility.cpp:650-675)
    if (ctx->cfg_pcal_interval[d] <= 0) continue;
    const fxg_datastream& s = ctx->dss[d];
    int idx = ctx->pcaloff[d];
    for (int j = 0; j < s.nrecbands; ++j) {
      const fxg_freq& fr = ctx->freqs[s.recfreq_index[s.recband_localfreq[j]]];
      int nt;
      long long off;
      pcal_tones_of(fr.freq_mhz, fr.bw_mhz, fr.lsb != 0, ctx->cfg_pcal_interval[d], ctx->cfg_pcal_base[d], &nt, &off);
      NormSeg g{};
      g.kind = 2;
      g.decf = -1;
      g.off = idx;
      g.nout = nt;
      g.wf = ctx->acw_of[(size_t)d * MB + j];     // the band's parallel autocorrelation weight, -1: none
      g.pcz = s.nzoombands > 0;                    // numtotalbands > numrecordedbands
      g.ds1 = g.ds2 = d;
      g.cal1 = g.cal2 = zero;
      g.aw1 = g.aw2 = -1;
      g.zs1 = g.zs2 = 1.0f;
      g.chanfactor = (float)(B * 2 * (fr.nchan / fr.chanavg) * fr.chanavg);
      ctx->normseg.push_back(g);
      idx += nt;
    }
  }
  if (int rc = dev_upload(ctx, &ctx->d_normseg, ctx->normseg)) return rc;
  if (int rc = dev_upload(ctx, &ctx->d_normcal, ctx->normcal)) return rc;
  if (int rc = dev_malloc(ctx, (void**)&ctx->d_calmean, sizeof(float) * ctx->normcal.size())) return rc;
  if (int rc = dev_zero(ctx, ctx->d_calmean, sizeof(float) * ctx->normcal.size())) return rc;
  if (int rc = dev_malloc(ctx, (void**)&ctx->d_tsys, sizeof(double) * nds)) return rc;
  if (int rc = dev_malloc(ctx, (void**)&ctx->d_dpct, sizeof(float) * nds)) return rc;
  return FXG_OK;
}

int crosspol_partner(const fxg_datastream& ds, int b) {
  int lf = ds.recband_localfreq[b];
  int idx[2] = {-1, -1}, n = 0;
  for (int k = 0; k < ds.nrecbands && n < 2; ++k)
    if (ds.recband_localfreq[k] == lf) idx[n++] = k;
  if (n < 2) return -1;
  if (b == idx[0]) return idx[1];
  if (b == idx[1]) return idx[0];
  return -1;
}

int build_tables(fxg_ctx* ctx) {
  const int nds = ctx->nds;
  ctx->databytes.resize(nds);
  // models
  std::vector<double> model_all;
  std::vector<float> chanfreq;
  ctx->h_ds.assign(nds, DevDs());
  ctx->maxframes = 0;
  long long fvo = 0;
  for (int d = 0; d < nds; ++d) {
    const fxg_datastream& s = ctx->dss[d];
    DevDs& x = ctx->h_ds[d];
    std::memset(&x, 0, sizeof(x));
    const fxg_freq& f0 = ctx->freqs[s.recfreq_index[0]];
    x.format = s.format;
    x.nbits = s.nbits;
    x.iscomplex = s.complex_sampling ? 1 : 0;
    x.dsb = s.complex_sampling == 2;
    x.xsc = 1.0f;
    x.framebytes = s.framebytes;
    x.headerbytes = s.headerbytes;
    x.nrecbands = s.nrecbands;
    x.nrecfreqs = s.nrecfreqs;
    x.C = f0.nchan;
    x.L = s.complex_sampling ? f0.nchan : 2 * f0.nchan;
    x.S = s.arraystridelen ? s.arraystridelen : calcstridelength(f0.nchan);
    x.NS = x.L / x.S;
    x.NFS = x.C / x.S;
    const int nsb = s.nbits * (s.complex_sampling ? 2 : 1);
    x.spf = (s.framebytes - s.headerbytes) * 8 / (nsb * s.nrecbands);
    // bytes-per-block ratio (mode.cpp:124-139); samplesperblock is 1 for Mk5Mode
    x.bpb_num = (s.nrecbands * nsb) / 8;
    x.bpb_den = 1;
    if (x.bpb_num == 0) { x.bpb_num = 1; x.bpb_den = 8 / (s.nrecbands * nsb); }
    x.sampletime_us = s.complex_sampling ? 1.0 / f0.bw_mhz : 1.0 / (2.0 * f0.bw_mhz);
    x.bw_mhz = f0.bw_mhz;
    x.intclockseconds = (int)std::floor(s.clock_coeff0_us / 1000000.0 + 0.5);
    x.model_order = s.model_order;
    x.model_nsamples = s.model_nsamples;
    x.model_off = (int)(model_all.size() / FXG_MAX_POLY);
    for (int k = 0; k < s.model_nsamples * FXG_MAX_POLY; ++k) model_all.push_back(s.model_delay_us[k]);
    // fractional LO frequency (mode.cpp:83-101); for double-sideband data the LO is the band
    // centre, the edge moved by BW/2 towards the band
    x.fractionallo = 0;
    for (int k = 0; k < s.nrecfreqs; ++k) {
      const fxg_freq& f = ctx->freqs[s.recfreq_index[k]];
      double lo = f.freq_mhz;
      if (s.complex_sampling == 2) lo += f.lsb ? -f.bw_mhz / 2.0 : f.bw_mhz / 2.0;
      if (std::fabs(lo - (int)(lo + 0.5)) > TINY) x.fractionallo = 1;
    }
    for (int k = 0; k < s.nrecfreqs; ++k) {
      const fxg_freq& f = ctx->freqs[s.recfreq_index[k]];
      DevRecFreq& r = x.rf[k];
      r.lofreq = f.freq_mhz;
      if (s.complex_sampling && f.lsb) r.lofreq = -r.lofreq;     // mode.cpp:875-877
      if (s.complex_sampling == 2) r.lofreq += f.bw_mhz / 2.0;   // mode.cpp:878-879
      r.lo_offset_hz = s.lo_offset_hz[k];
      r.looff = (r.lo_offset_hz > 0.0 || r.lo_offset_hz < 0.0);
      double fl = std::fabs(r.lo_offset_hz) - (int)std::fabs(r.lo_offset_hz);
      r.isfraclooffset = r.looff && fl > TINY;
      if (r.lo_offset_hz < 0) fl = -fl;
      r.fraclooffset = r.looff ? fl : 0.0;
      r.clockoff = s.clock_offset_us[k];
      r.halfdelta = s.clock_offset_delta_us[k] / 2;
      r.lsb = f.lsb;
      r.nbands = 0;
      for (int b = 0; b < s.nrecbands; ++b)
        if (s.recband_localfreq[b] == k) r.band[r.nbands++] = b;
      // the per-pol fractional-sample phasors: R takes fracsamprotatorA (+delta/2), the other
      // pols fracsamprotatorB (-delta/2) (mode.cpp:1076-1118, 1257-1261); with two or more
      // bands the second pol (the first non-R band of the first two) turns by phasecorrA =
      // exp(+i pi phi / 360), phi the pol phase offset in degrees, computed in double and
      // rounded to float (mode.cpp:421-447, 1318-1330)
      int lcp = -1;
      if (r.nbands > 1) lcp = s.recband_pol[r.band[0]] == 'R' ? r.band[1] : r.band[0];
      const double degphase = -s.phase_offset_deg[k] / 2;
      const float2 pca = make_float2((float)std::cos(degphase * M_PI / 180.0),
                                     (float)-std::sin(degphase * M_PI / 180.0));
      for (int kb = 0; kb < r.nbands; ++kb) {
        const int b = r.band[kb];
        r.bhd[kb] = s.recband_pol[b] == 'R' ? r.halfdelta : -r.halfdelta;
        r.pcre[kb] = b == lcp ? pca.x : 1.0f;
        r.pcim[kb] = b == lcp ? pca.y : 0.0f;
        r.newfrac[kb] = kb > 0 && (r.bhd[kb] != r.bhd[kb - 1] || r.pcre[kb] != r.pcre[kb - 1] ||
                                   r.pcim[kb] != r.pcim[kb - 1]);
      }
    }
    // channel-frequency slopes (mode.cpp:303-324), computed as the reference does
    x.chanfreq_off = (int)chanfreq.size();
    const double bw = f0.bw_mhz;
    const int C = x.C, S = x.S, NFS = x.NFS;
    for (int i = 0; i < S; ++i) chanfreq.push_back((float)((TWO_PI * (i) * bw) / C));
    for (int i = 0; i < NFS; ++i) chanfreq.push_back((float)((TWO_PI * i * S * bw) / C));
    for (int i = 0; i < NFS; ++i) chanfreq.push_back((float)((-TWO_PI * ((NFS - i) * S) * bw) / C));
    for (int i = 0; i < NFS; ++i) chanfreq.push_back((float)((TWO_PI * i * S * bw) / C - TWO_PI * bw / 2.0));
    ctx->databytes[d] = get_data_bytes(ctx, d);
    const int cap_frames = (int)((ctx->databytes[d] + 4LL * s.framebytes) / s.framebytes);
    x.fvalid_off = (int)fvo;
    fvo += cap_frames;
    ctx->maxframes = std::max(ctx->maxframes, cap_frames);
  }
  ctx->fvalid_total = fvo;
  int rc;
  if ((rc = dev_upload(ctx, &ctx->d_model, model_all))) return rc;
  if ((rc = dev_upload(ctx, &ctx->d_chanfreq, chanfreq))) return rc;
  if (int rc_ = dev_malloc(ctx, (void**)&ctx->d_fvalid, std::max<long long>(fvo, 16))) return rc_;
  if (int rc_ = dev_zero(ctx, ctx->d_fvalid, std::max<long long>(fvo, 16))) return rc_;
  if (int rc_ = dev_malloc(ctx, (void**)&ctx->d_fzone, sizeof(int2) * std::max<long long>(fvo, 16))) return rc_;
  if (int rc_ = dev_zero(ctx, ctx->d_fzone, sizeof(int2) * std::max<long long>(fvo, 16))) return rc_;
  // FFT-length classes, in order of first appearance
  ctx->fclasses.clear();
  for (int d = 0; d < nds; ++d) {
    const int L = ctx->h_ds[d].L, cplx = ctx->h_ds[d].iscomplex;
    size_t k = 0;
    while (k < ctx->fclasses.size() && !(ctx->fclasses[k].L == L && ctx->fclasses[k].cplx == cplx)) ++k;
    if (k == ctx->fclasses.size()) {
      fxg_ctx::FClass c;
      c.L = L;
      c.cplx = cplx;
      c.big = L > FXG_MAX_L;
      ctx->fclasses.push_back(c);
    }
    ctx->fclasses[k].dsl.push_back(d);
    const DevDs& x = ctx->h_ds[d];
    const int need = (std::max(x.S + x.NS, x.S + x.NFS) + 7) / 8 * 8;
    ctx->fclasses[k].phcap = std::max(ctx->fclasses[k].phcap, need);
  }
  // FFT twiddles of the Stockham passes after the first, per pass (R, Ns): entries
  // exp(-2 pi i k r / (Ns R)) for r = 1..R-1, k = 0..Ns-1, k fastest; computed in FP64.
  // One table set per FFT length.
  {
    std::vector<float2> tw;
    std::vector<int> dsl;
    std::map<int, long long> tw_of_L;
    for (auto& c : ctx->fclasses) {
      c.dsl_off = (long long)dsl.size();
      dsl.insert(dsl.end(), c.dsl.begin(), c.dsl.end());
      if (c.big) {                 // sub-FFT tables of 2048 points plus the L-point twiddles
        c.twL_off = (long long)tw.size();
        for (int j = 0; j < c.L; ++j) {
          const double a = -TWO_PI * (double)j / (double)c.L;
          tw.push_back(make_float2((float)std::cos(a), (float)std::sin(a)));
        }
      }
      const int Lt = c.big ? 2048 : c.L;
      if (tw_of_L.count(Lt)) { c.tw_off = tw_of_L[Lt]; continue; }
      c.tw_off = (long long)tw.size();
      tw_of_L[Lt] = c.tw_off;
      int log = 0;
      while ((1 << log) < Lt) ++log;
      const int n8 = log / 3, rem = log % 3;
      auto add_pass = [&](int R, int Ns) {
        for (int r = 1; r < R; ++r)
          for (int k = 0; k < Ns; ++k) {
            const double a = -TWO_PI * (double)k * (double)r / ((double)Ns * (double)R);
            tw.push_back(make_float2((float)std::cos(a), (float)std::sin(a)));
          }
      };
      int ns = 8;
      for (int p = 1; p < n8; ++p) { add_pass(8, ns); ns *= 8; }
      if (rem) add_pass(1 << rem, ns);
    }
    if ((rc = dev_upload(ctx, &ctx->d_twiddle, tw))) return rc;
    if ((rc = dev_upload(ctx, &ctx->d_dslist, dsl))) return rc;
  }
  if (int rc_ = dev_malloc(ctx, (void**)&ctx->d_run, sizeof(DevDsRun) * nds)) return rc_;
  if (int rc_ = dev_malloc(ctx, (void**)&ctx->d_weights, sizeof(float) * (size_t)nds * ctx->B)) return rc_;
  if (int rc_ = dev_zero(ctx, ctx->d_weights, sizeof(float) * (size_t)nds * ctx->B)) return rc_;
  if (int rc_ = dev_malloc(ctx, (void**)&ctx->d_flags, sizeof(int32_t) * (size_t)nds * ctx->flaglen * 2)) return rc_;
  return FXG_OK;
}

long long round_up(long long v, long long m) { return (v + m - 1) / m * m; }

// Channel tile width of the 16-bit spectra layout the F-engine writes and the MFMA X-engine
// reads (log2): 16 channels where the group has one 32-row tile and a multiple of 16 channels
// (one FFT then writes 64-byte runs), else 8.
#ifndef FXG_XTW_DEFAULT
#define FXG_XTW_DEFAULT 4
#endif
int tile_width_log(const DevXGroup& g, int want) {
  return (want == 4 && g.ntile == 1 && g.C % 16 == 0) ? 4 : 3;
}

// The blocked plain X-engine's tables.  A row is a spectra band or zoom window with its channel
// count, (datastream, band, channel offset, channels); products join rows into groups.  Each
// group of at most 64 rows gets a pair table (the accumulator of the product of rows (i, j), or
// -1) and work items of 8 x 8 row blocks over 64-channel tiles.  Products of groups above 64
// rows or whose rows' datastreams differ in FFT stride, a second product of the same row pair,
// and every product when the FP32 spectra reach 4 GiB (the kernel's offsets are 32-bit) stay
// with the per-product kernel.
void build_xplain_blocks(fxg_ctx* ctx) {
  const auto& P = ctx->prods;
  ctx->xp_groups.clear(); ctx->xp_rows.clear(); ctx->xp_pairs.clear(); ctx->xp_items.clear();
  ctx->prods_rest.clear();
  std::map<std::array<int, 4>, int> rid;
  std::vector<std::array<int, 4>> rkey;
  auto row = [&](int d, int b, int co, int nc) {
    const std::array<int, 4> k = {d, b, co, nc};
    auto f = rid.find(k);
    if (f != rid.end()) return f->second;
    const int r = (int)rkey.size();
    rid[k] = r;
    rkey.push_back(k);
    return r;
  };
  std::vector<int> r1(P.size()), r2(P.size());
  for (size_t i = 0; i < P.size(); ++i) {
    r1[i] = row(P[i].ds1, P[i].band1, P[i].choff1, P[i].nchan);
    r2[i] = row(P[i].ds2, P[i].band2, P[i].choff2, P[i].nchan);
  }
  std::vector<int> par(rkey.size());
  for (size_t r = 0; r < par.size(); ++r) par[r] = (int)r;
  auto root = [&](int x) { while (par[x] != x) x = par[x] = par[par[x]]; return x; };
  for (size_t i = 0; i < P.size(); ++i) {
    const int a = root(r1[i]), b = root(r2[i]);
    if (a != b) par[std::max(a, b)] = std::min(a, b);
  }
  std::map<int, int> gid;
  std::vector<std::vector<int>> grows;
  std::vector<int> local(rkey.size());
  for (size_t r = 0; r < rkey.size(); ++r) {
    const int g = root((int)r);
    if (!gid.count(g)) { gid[g] = (int)grows.size(); grows.emplace_back(); }
    local[r] = (int)grows[gid[g]].size();
    grows[gid[g]].push_back((int)r);
  }
  if (8LL * ctx->spec_elems >= (1LL << 32)) {
    ctx->prods_rest = P;
    return;
  }
  std::vector<int> gslot(grows.size(), -1);
  for (size_t g = 0; g < grows.size(); ++g) {
    const int n = (int)grows[g].size();
    if (n > 64) continue;
    bool onestride = true;
    for (int r : grows[g]) onestride = onestride && ctx->h_ds[rkey[r][0]].C == ctx->h_ds[rkey[grows[g][0]][0]].C;
    if (!onestride) continue;
    XPGroup G;
    G.nrows = n;
    G.C = rkey[grows[g][0]][3];
    G.stride = ctx->h_ds[rkey[grows[g][0]][0]].C;
    G.pad = 0;
    G.row0 = (int)ctx->xp_rows.size();
    G.pair0 = (int)ctx->xp_pairs.size();
    for (int r : grows[g]) {
      const DevDs& x = ctx->h_ds[rkey[r][0]];
      XPRow R;
      R.off = x.spec_off + (long long)rkey[r][1] * ctx->chunkcap * x.C + rkey[r][2];
      R.stride = x.C;
      R.pad = 0;
      ctx->xp_rows.push_back(R);
    }
    ctx->xp_pairs.resize(ctx->xp_pairs.size() + (size_t)n * n, -1);
    gslot[g] = (int)ctx->xp_groups.size();
    ctx->xp_groups.push_back(G);
  }
  for (size_t i = 0; i < P.size(); ++i) {
    const int gs = gslot[gid[root(r1[i])]];
    if (gs < 0) { ctx->prods_rest.push_back(P[i]); continue; }
    const XPGroup& G = ctx->xp_groups[gs];
    int& slot = ctx->xp_pairs[G.pair0 + (size_t)local[r1[i]] * G.nrows + local[r2[i]]];
    if (slot >= 0) { ctx->prods_rest.push_back(P[i]); continue; }
    slot = P[i].acc_off;
  }
  // Workgroups are dispatched to the eight XCDs in turn by workgroup index, and each XCD has
  // its own L2: the blocks of one (group, channel tile), which read the same rows, go to one
  // XCD (unit u to XCD u mod 8, its blocks at indices x, x + 8, ...); the shorter XCD lists
  // are padded with empty items (group -1).
  constexpr int NXCD = 8;
  std::vector<std::vector<XPItem>> per(NXCD);
  int unit = 0;
  for (size_t gi = 0; gi < ctx->xp_groups.size(); ++gi) {
    const XPGroup& G = ctx->xp_groups[gi];
    const int nbk = (G.nrows + 7) / 8;
    for (int c0 = 0; c0 < G.C; c0 += 64, ++unit)
      for (int rb = 0; rb < nbk; ++rb)
        for (int cb = rb; cb < nbk; ++cb) {
          XPItem it;
          it.group = (int)gi; it.rb = rb; it.cb = cb; it.c0 = c0;
          per[unit % NXCD].push_back(it);
        }
  }
  size_t len = 0;
  for (const auto& v : per) len = std::max(len, v.size());
  for (size_t k = 0; k < len; ++k)
    for (int x = 0; x < NXCD; ++x) {
      XPItem it;
      it.group = -1; it.rb = 0; it.cb = 0; it.c0 = 0;
      ctx->xp_items.push_back(k < per[x].size() ? per[x][k] : it);
    }
}

int build_engines(fxg_ctx* ctx, int chunk_req) {
  const int nds = ctx->nds, nb = ctx->nbl;
  const bool mfma = ctx->xengine != FXG_X_FP32;
  // ---- MFMA groups: one per recorded freq table entry, rows = station bands ----
  if (mfma) {
    // freqs that meet in USB x LSB products share a group (one contraction over the rows of
    // both sidebands), keyed by the smallest freq index among them
    std::vector<int> gkey(ctx->nfreqs);
    for (int i = 0; i < ctx->nfreqs; ++i) gkey[i] = i;
    auto root = [&](int i) { while (gkey[i] != i) i = gkey[i]; return i; };
    for (const auto& bl : ctx->bls)
      for (int k = 0; k < bl.nfreqs; ++k)
        for (int p = 0; p < bl.npolprod[k]; ++p) {
          const int a = root(freq_of_band(ctx->dss[bl.ds1], bl.ds1band[k][p]));
          const int c = root(freq_of_band(ctx->dss[bl.ds2], bl.ds2band[k][p]));
          if (a != c) gkey[std::max(a, c)] = std::min(a, c);
        }
    std::map<int, int> gof;
    for (int d = 0; d < nds; ++d) {
      const fxg_datastream& s = ctx->dss[d];
      for (int b = 0; b < s.nrecbands; ++b) {
        int fi = root(s.recfreq_index[s.recband_localfreq[b]]);
        if (!gof.count(fi)) {
          gof[fi] = (int)ctx->groups.size();
          DevXGroup g;
          std::memset(&g, 0, sizeof(g));
          g.C = ctx->freqs[fi].nchan;
          ctx->groups.push_back(g);
          ctx->group_rows.emplace_back();
        }
        int g = gof[fi];
        ctx->rowof[{d, b}] = {g, (int)ctx->group_rows[g].size()};
        ctx->group_rows[g].push_back({d, b});
        ctx->h_ds[d].band_xfreq[b] = g;
        ctx->h_ds[d].band_xrow[b] = (int)ctx->group_rows[g].size() - 1;
      }
    }
    // zoom bands correlated on baselines: rows of their zoom freq's group, written by the
    // F-engine from the parent band's channel window
    for (int j = 0; j < nb; ++j) {
      const fxg_baseline& bl = ctx->bls[j];
      for (int k = 0; k < bl.nfreqs; ++k)
        for (int p = 0; p < bl.npolprod[k]; ++p)
          for (int side = 0; side < 2; ++side) {
            const int d = side ? bl.ds2 : bl.ds1;
            const int bi = side ? bl.ds2band[k][p] : bl.ds1band[k][p];
            const fxg_datastream& s = ctx->dss[d];
            if (bi < s.nrecbands || ctx->rowof.count({d, bi})) continue;
            const int fi = root(freq_of_band(s, bi));
            if (!gof.count(fi)) {
              gof[fi] = (int)ctx->groups.size();
              DevXGroup g;
              std::memset(&g, 0, sizeof(g));
              g.C = ctx->freqs[fi].nchan;
              ctx->groups.push_back(g);
              ctx->group_rows.emplace_back();
            }
            const int g = gof[fi];
            int rec = 0, choff = 0;
            band_source(s, bi, &rec, &choff);
            DevDs& x = ctx->h_ds[d];
            if (x.nzout >= FXG_MAX_ZOUT)
              return fail(ctx, FXG_E_INVALID, "too many correlated zoom bands in one datastream");
            x.zout_band[x.nzout] = rec;
            x.zout_group[x.nzout] = g;
            x.zout_row[x.nzout] = (int)ctx->group_rows[g].size();
            x.zout_c0[x.nzout] = choff;
            x.zout_n[x.nzout] = ctx->freqs[fi].nchan;
            x.nzout++;
            ctx->rowof[{d, bi}] = {g, (int)ctx->group_rows[g].size()};
            ctx->group_rows[g].push_back({d, bi});
          }
    }
    for (size_t g = 0; g < ctx->groups.size(); ++g) {
      int nin = (int)ctx->group_rows[g].size();
      if (nin > 64) return fail(ctx, FXG_E_INVALID, "MFMA X-engine supports at most 64 inputs per freq");
      ctx->groups[g].nin = nin;
      ctx->groups[g].ntile = (nin + 31) / 32;
      ctx->groups[g].twlog = tile_width_log(ctx->groups[g], FXG_XTW_DEFAULT);
    }
  }
  // ---- chunk capacity: spectra buffer within a budget ----
  long long per_fft = 0;
  if (mfma) {
    for (auto& g : ctx->groups) per_fft += (long long)g.ntile * 32 * g.C * 4;   // bytes
  } else {
    for (int d = 0; d < nds; ++d) per_fft += (long long)ctx->h_ds[d].nrecbands * ctx->h_ds[d].C * 8;
  }
  // (FP32 spectra stay below 4 GiB: the blocked plain X-engine addresses them with 32-bit offsets)
  const long long budget = mfma ? (4LL << 30) : (4LL << 30) - (256LL << 20);
  long long cap = chunk_req > 0 ? chunk_req : std::max<long long>(64, budget / std::max<long long>(per_fft, 1));
  cap = std::min<long long>(cap, round_up(ctx->B, FXG_XPAD));
  cap = round_up(cap, FXG_XPAD);
  ctx->chunkcap = (int)cap;
  if (int rc_ = dev_malloc(ctx, (void**)&ctx->d_params, sizeof(FftDelay) * (size_t)nds * cap)) return rc_;
  if (mfma) {
    long long off = 0, aoff = 0, soff = 0;
    for (auto& g : ctx->groups) {
      g.spec_off = off;
      g.rowstride = cap * g.C + 32;       // one 128-byte line of padding per row
      off += (long long)g.ntile * 32 * g.rowstride;
      g.acc_off = aoff;
      aoff += (long long)g.C * (g.ntile * (g.ntile + 1) / 2) * 2 * 1024;
      g.scale_off = soff;
      soff += (long long)g.ntile * 32;                 // one scale per input row
    }
    ctx->spec_elems = off;
    ctx->vacc_len = aoff;
    ctx->xscale_len = soff;
    if (int rc_ = dev_malloc(ctx, (void**)&ctx->d_spec16, sizeof(uint32_t) * std::max<long long>(off, 4))) return rc_;
    if (int rc_ = dev_zero(ctx, ctx->d_spec16, sizeof(uint32_t) * std::max<long long>(off, 4))) return rc_;
    if (int rc_ = dev_malloc(ctx, (void**)&ctx->d_vacc, sizeof(float) * std::max<long long>(aoff, 4))) return rc_;
    // nominal per-input scale 1/sqrt(L <x^2>) of the input's sample levels, which keeps the
    // 16-bit spectra near unit magnitude; the accumulation divides it out again
    const double p_hi = 0.5 * std::erfc(0.95 / std::sqrt(2.0));
    std::vector<float> xs((size_t)std::max<long long>(soff, 1), 0.0f);
    for (size_t g = 0; g < ctx->groups.size(); ++g)
      for (size_t r = 0; r < ctx->group_rows[g].size(); ++r) {
        const int d = ctx->group_rows[g][r].first;
        const DevDs& x = ctx->h_ds[d];
        const double xsq = x.iscomplex || x.nbits == 8 ? 2.0 * (20.0 / 3.3) * (20.0 / 3.3)
                                                       : 2 * p_hi * 3.3359 * 3.3359 + (1 - 2 * p_hi);
        const float sc = (float)(1.0 / std::sqrt((double)x.L * xsq));
        xs[(size_t)(ctx->groups[g].scale_off + (long long)r)] = sc;
        ctx->h_ds[d].xsc = sc;                  // the F-engine folds it into its phasors
      }
    if (int rc = dev_upload(ctx, &ctx->d_xscale, xs)) return rc;
    if (int rc = dev_upload(ctx, &ctx->d_groups, ctx->groups)) return rc;
  } else {
    long long off = 0;
    for (int d = 0; d < nds; ++d) {
      ctx->h_ds[d].spec_off = off;
      off += (long long)ctx->h_ds[d].nrecbands * cap * ctx->h_ds[d].C;
    }
    ctx->spec_elems = off;
    if (int rc_ = dev_malloc(ctx, (void**)&ctx->d_spec32, sizeof(float2) * std::max<long long>(off, 2))) return rc_;
    if (int rc_ = dev_zero(ctx, ctx->d_spec32, sizeof(float2) * std::max<long long>(off, 2))) return rc_;
  }
  // ---- products and accumulation entries ----
  const int bpa = ctx->bpa;
  if (!mfma) {
    long long acc = 0;
    auto add_prod = [&](int d1, int b1, int c1, int d2, int b2, int c2, int nchan) {
      DevProd p;
      p.ds1 = d1; p.band1 = b1; p.choff1 = c1;
      p.ds2 = d2; p.band2 = b2; p.choff2 = c2;
      p.nchan = nchan;
      p.acc_off = (int)acc;
      acc += nchan;
      ctx->prods.push_back(p);
      ctx->maxchan = std::max(ctx->maxchan, nchan);
      return (long long)p.acc_off;
    };
    for (int j = 0; j < nb; ++j) {
      const fxg_baseline& b = ctx->bls[j];
      for (int k = 0; k < b.nfreqs; ++k) {
        const int f = b.freq_index[k];
        const fxg_freq& fr = ctx->freqs[f];
        const int nout = (fr.nchan + fr.chanavg - 1) / fr.chanavg;
        for (int p = 0; p < b.npolprod[k]; ++p) {
          int r1, c1, r2, c2;
          band_source(ctx->dss[b.ds1], b.ds1band[k][p], &r1, &c1);
          band_source(ctx->dss[b.ds2], b.ds2band[k][p], &r2, &c2);
          long long src = add_prod(b.ds1, r1, c1, b.ds2, r2, c2, fr.nchan);
          FinPlain e;
          e.src = src;
          e.dst = ctx->boff[(size_t)f * nb + j] + p * nout;
          e.nout = nout;
          e.avg = fr.chanavg;
          e.bl = j;
          e.fq = f;
          e.pcstride = b.npolprod[k] * nout;
          ctx->finp.push_back(e);
          ctx->maxoutp = std::max(ctx->maxoutp, nout);
        }
      }
    }
    // autocorrelations (mode.cpp:1272, 1333-1339) and their placement (core.cpp:1258-1304)
    for (int d = 0; d < nds; ++d) {
      const fxg_datastream& s = ctx->dss[d];
      const int C = ctx->h_ds[d].C;
      std::vector<long long> aacc(s.nrecbands), xacc(s.nrecbands, -1);
      for (int b = 0; b < s.nrecbands; ++b) aacc[b] = add_prod(d, b, 0, d, b, 0, C);
      if (ctx->crossautos)
        for (int b = 0; b < s.nrecbands; ++b) {
          int pb = crosspol_partner(s, b);
          if (pb >= 0) xacc[b] = add_prod(d, b, 0, d, pb, 0, C);
        }
      const int cavg = ctx->freqs[s.recfreq_index[0]].chanavg;
      long long idx = ctx->acoff[d];
      for (int w = 0; w < bpa; ++w) {
        if (w == 1 && !ctx->crossautos) break;
        for (int b = 0; b < s.nrecbands + s.nzoombands; ++b) {
          const int fi = freq_of_band(s, b);
          if (!(ctx->used[fi] || ctx->equiv[fi])) continue;
          const int nout = ctx->freqs[fi].nchan / ctx->freqs[fi].chanavg;
          int rec, choff;
          band_source(s, b, &rec, &choff);
          long long src = w == 0 ? aacc[rec] : xacc[rec];
          if (src >= 0) {
            FinPlain e;
            e.src = src + choff;
            e.dst = (int)idx;
            e.nout = nout;
            e.avg = cavg;
            e.bl = -1;
            e.fq = 0;
            e.pcstride = 0;
            ctx->finp.push_back(e);
            ctx->maxoutp = std::max(ctx->maxoutp, nout);
          }
          idx += nout;
        }
      }
    }
    ctx->xacc_len = acc;
    if (int rc_ = dev_malloc(ctx, (void**)&ctx->d_xacc, sizeof(float2) * std::max<long long>(acc, 2))) return rc_;
    if (int rc = dev_upload(ctx, &ctx->d_prods, ctx->prods)) return rc;
    if (int rc = dev_upload(ctx, &ctx->d_finp, ctx->finp)) return rc;
    build_xplain_blocks(ctx);
    if (!ctx->xp_items.empty()) {
      if (int rc = dev_upload(ctx, &ctx->d_xp_groups, ctx->xp_groups)) return rc;
      if (int rc = dev_upload(ctx, &ctx->d_xp_rows, ctx->xp_rows)) return rc;
      if (int rc = dev_upload(ctx, &ctx->d_xp_pairs, ctx->xp_pairs)) return rc;
      if (int rc = dev_upload(ctx, &ctx->d_xp_items, ctx->xp_items)) return rc;
    }
    if (!ctx->prods_rest.empty())
      if (int rc = dev_upload(ctx, &ctx->d_prods_rest, ctx->prods_rest)) return rc;
  } else {
    for (int j = 0; j < nb; ++j) {
      const fxg_baseline& b = ctx->bls[j];
      for (int k = 0; k < b.nfreqs; ++k) {
        const int f = b.freq_index[k];
        const fxg_freq& fr = ctx->freqs[f];
        const int nout = (fr.nchan + fr.chanavg - 1) / fr.chanavg;
        for (int p = 0; p < b.npolprod[k]; ++p) {
          auto g1 = ctx->rowof[{b.ds1, b.ds1band[k][p]}];
          auto g2 = ctx->rowof[{b.ds2, b.ds2band[k][p]}];
          if (g1.first != g2.first) return fail(ctx, FXG_E_INVALID, "product rows in different MFMA groups");
          FinMfma e;
          e.group = g1.first;
          e.row1 = g1.second;
          e.row2 = g2.second;
          e.chan0 = 0;
          e.dst = ctx->boff[(size_t)f * nb + j] + p * nout;
          e.nout = nout;
          e.avg = fr.chanavg;
          e.isauto = 0;
          e.bl = j;
          e.fq = f;
          e.pcstride = b.npolprod[k] * nout;
          ctx->finm.push_back(e);
          ctx->maxoutm = std::max(ctx->maxoutm, nout);
        }
      }
    }
    // autocorrelations: recorded bands from the diagonal of their group, zoom bands from
    // the channel window of their parent band (mode.cpp:184-196, 1272, 1333-1339)
    for (int d = 0; d < nds; ++d) {
      const fxg_datastream& s = ctx->dss[d];
      const int cavg = ctx->freqs[s.recfreq_index[0]].chanavg;
      long long idx = ctx->acoff[d];
      for (int w = 0; w < bpa; ++w) {
        if (w == 1 && !ctx->crossautos) break;
        for (int b = 0; b < s.nrecbands + s.nzoombands; ++b) {
          const int fi = freq_of_band(s, b);
          if (!(ctx->used[fi] || ctx->equiv[fi])) continue;
          const int nout = ctx->freqs[fi].nchan / ctx->freqs[fi].chanavg;
          int rec = 0, choff = 0;
          band_source(s, b, &rec, &choff);
          int other = w == 0 ? rec : crosspol_partner(s, rec);
          if (other >= 0) {
            auto g1 = ctx->rowof[{d, rec}];
            auto g2 = ctx->rowof[{d, other}];
            FinMfma e;
            e.group = g1.first;
            e.row1 = g1.second;
            e.row2 = g2.second;
            e.chan0 = choff;
            e.dst = (int)idx;
            e.nout = nout;
            e.avg = cavg;
            e.isauto = (w == 0);
            e.bl = -1;
            e.fq = 0;
            e.pcstride = 0;
            ctx->finm.push_back(e);
            ctx->maxoutm = std::max(ctx->maxoutm, nout);
          }
          idx += nout;
        }
      }
    }
    if (int rc = dev_upload(ctx, &ctx->d_finm, ctx->finm)) return rc;
  }
  // weights (core.cpp:1005-1052, 1306-1370)
  for (int f = 0; f < ctx->nfreqs; ++f) {
    if (!ctx->used[f]) continue;
    for (int j = 0; j < nb; ++j) {
      const fxg_baseline& b = ctx->bls[j];
      for (int k = 0; k < b.nfreqs; ++k)
        if (b.freq_index[k] == f)
          for (int p = 0; p < b.npolprod[k]; ++p) {
            FinWeight w;
            w.ds1 = b.ds1; w.ds2 = b.ds2;
            w.dstf = ctx->bwoff[(size_t)f * nb + j] * 2 + p;
            ctx->finw.push_back(w);
          }
    }
  }
  for (int d = 0; d < nds; ++d) {
    const fxg_datastream& s = ctx->dss[d];
    int widx = ctx->acwoff[d] * 2;
    for (int w = 0; w < bpa; ++w) {
      if (w == 1 && !ctx->crossautos) break;
      for (int b = 0; b < s.nrecbands + s.nzoombands; ++b) {
        const int fi = freq_of_band(s, b);
        if (!(ctx->used[fi] || ctx->equiv[fi])) continue;
        int rec, choff;
        band_source(s, b, &rec, &choff);
        if (w == 0 || crosspol_partner(s, rec) >= 0) {
          FinWeight e;
          e.ds1 = d; e.ds2 = -1; e.dstf = widx;
          ctx->finw.push_back(e);
        }
        widx++;
      }
    }
  }
  if (int rc = dev_upload(ctx, &ctx->d_finw, ctx->finw)) return rc;
  if (int rc = build_norm(ctx)) return rc;
  if (int rc = dev_upload(ctx, &ctx->d_ds, ctx->h_ds)) return rc;
  return FXG_OK;
}

hipEvent_t pool_event(fxg_ctx* ctx) {
  if (!ctx->evpool.empty()) {
    hipEvent_t e = ctx->evpool.back();
    ctx->evpool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Brackets one launch with timing events when timing is enabled.
struct TimedLaunch {
  fxg_ctx* ctx;
  int cat;
  hipStream_t st;
  hipEvent_t a = nullptr, b = nullptr;
  TimedLaunch(fxg_ctx* c, int category, hipStream_t s) : ctx(c), cat(category), st(s) {
    if (!ctx->timing) return;
    a = pool_event(ctx);
    b = pool_event(ctx);
    if (a) (void)hipEventRecord(a, st);
  }
  ~TimedLaunch() {
    if (!ctx->timing || !a || !b) return;
    (void)hipEventRecord(b, st);
    ctx->tev.emplace_back(cat, a, b);
  }
};

int ntb_for(const fxg_ctx* ctx, int L) {
  if (ctx->fsel.ntb) return ctx->fsel.ntb;
  return L >= 4096 ? 512 : 256;
}

int log2i(int v) {
  int l = 0;
  while ((1 << l) < v) ++l;
  return l;
}

// Work items of the persistent launch: every datastream's chunk in `parts` parts, longest FFT
// first (the items of one length run back to back, so twiddles reload rarely).
// Configuration::calcgoodxmacstridelength (configuration.cpp:2653-2698): the divisor of the gcd
// of the baseline freqs' channel counts closest to 150 (the larger one on a tie).
int good_xmac_stride(const fxg_ctx* ctx) {
  long long g = 0;
  for (const auto& b : ctx->bls)
    for (int k = 0; k < b.nfreqs; ++k) g = std::gcd(g, (long long)ctx->freqs[b.freq_index[k]].nchan);
  int best = 0;
  double err = 1.0e12;
  for (long long guess = g; guess > 0; --guess) {
    if (g % guess) continue;
    const double d = std::fabs((double)guess - 150.0);
    if (d < err) { err = d; best = (int)guess; }
  }
  return best;
}

// Multi-phase-centre uv shift: stride lengths, per-freq rotation constants, the intervals of a
// subint (core.cpp:761-776, 984-991, 1055-1057 per process thread, core.cpp:2158-2166) and the
// places of the shift-decorrelation sums.
int build_uvshift(fxg_ctx* ctx) {
  if (ctx->nphase <= 1) return FXG_OK;
  if (ctx->xms <= 0) ctx->xms = good_xmac_stride(ctx);
  if (ctx->xms <= 0) return fail(ctx, FXG_E_INVALID, "no xmac stride length");
  ctx->rs = calcstridelength(ctx->xms);
  for (const auto& b : ctx->bls)
    for (int k = 0; k < b.nfreqs; ++k)
      if (ctx->freqs[b.freq_index[k]].nchan % ctx->xms)
        return fail(ctx, FXG_E_INVALID, "xmac stride length must divide every baseline freq's nchan");
  ctx->uvfreq.assign(ctx->nfreqs, UvFreq{});
  for (int f = 0; f < ctx->nfreqs; ++f) {
    const fxg_freq& fr = ctx->freqs[f];
    UvFreq& q = ctx->uvfreq[f];
    q.lofreq = fr.freq_mhz;
    q.chbw = fr.bw_mhz / (double)fr.nchan;
    q.stepbw = ctx->rs * q.chbw;
    q.lsb = fr.lsb;
    q.rs = ctx->rs;
    q.nsteps = (fr.nchan / ctx->xms) * (ctx->xms / ctx->rs);
  }
  const int B = ctx->B;
  const double blockns = (double)ctx->subint_ns / (double)B;
  int maxxc = (int)((double)ctx->maxns_xc / blockns);
  maxxc -= maxxc % ctx->nbuffered;
  if (maxxc == 0) maxxc = ctx->nbuffered;
  ctx->ivals.clear();
  int start = 0, nblk = 0;
  for (int t = 0; t < ctx->nthreads; ++t) {
    start += nblk;
    nblk = B / ctx->nthreads + (t < B % ctx->nthreads ? 1 : 0);
    for (int k = 0, sh = 0; k < nblk; k += maxxc, ++sh) {
      const int n = std::min(maxxc, nblk - k);
      ctx->ivals.push_back({start + k, n, (start + (double)sh * maxxc + n / 2.0) * blockns, n * blockns});
    }
  }
  const int nb = ctx->nbl;
  ctx->uvdec_idx.clear();
  for (int f = 0; f < ctx->nfreqs; ++f) {
    if (!ctx->used[f]) continue;
    for (int j = 0; j < nb; ++j)
      for (int k = 0; k < ctx->bls[j].nfreqs; ++k)
        if (ctx->bls[j].freq_index[k] == f)
          for (int p = 0; p < ctx->nphase; ++p) ctx->uvdec_idx.push_back(ctx->sdoff[(size_t)f * nb + j] * 2 + p);
  }
  int rc;
  if ((rc = dev_upload(ctx, &ctx->d_uvfreq, ctx->uvfreq))) return rc;
  if ((rc = dev_upload(ctx, &ctx->d_uvdec_idx, ctx->uvdec_idx))) return rc;
  const long long ndd = (long long)ctx->ivals.size() * nb * ctx->nphase;
  for (auto& sg : ctx->stage) {
    if ((rc = dev_malloc(ctx, (void**)&sg.d_uvdd, sizeof(double) * std::max<long long>(ndd, 1)))) return rc;
    if ((rc = dev_malloc(ctx, (void**)&sg.d_uvdec, sizeof(float) * std::max<size_t>(ctx->uvdec_idx.size(), 1)))) return rc;
  }
  return FXG_OK;
}

// Model::calculateDelayInterpolator of order 1 (model.cpp:219-283) on the host, for a model
// table [nsamples][FXG_MAX_POLY] whose polynomials start at the scan start: (slope per
// increment, offset); false outside the model.
bool host_linear_interp(const double* m, int order, int nsamples, double inc, double t, double span, int ninc,
                        double out[2]) {
  int ss = (int)(t / inc);
  if (ss == nsamples && (t - ss * inc) < 1e-6) --ss;
  if (ss < 0 || ss >= nsamples) return false;
  const double* c = m + (size_t)ss * FXG_MAX_POLY;
  auto ev = [&](double dt) {
    double tp = 1.0, v = 0.0;
    for (int i = 0; i <= order; ++i) { v += tp * c[i]; tp *= dt; }
    return v;
  };
  const double d1 = ev(t + span / 2.0 - ss * inc);
  const double d0 = ev(t - ss * inc);
  const double d2 = ev(t + span - ss * inc);
  out[0] = (d2 - d0) / ninc;
  out[1] = d0 + (d1 - (out[0] * ninc / 2.0 + d0)) / 3.0;
  return true;
}

// Per subint: the applied delay of every (interval, baseline, phase centre) and the shift
// decorrelation sums (core.cpp:1636-1670, 1887-1922), into the staging buffers.
void uvshift_tables(fxg_ctx* ctx, const int32_t offsets[3], double* dd, float* dec) {
  const int P = ctx->nphase, nb = ctx->nbl, nds = ctx->nds;
  const int ni = (int)ctx->ivals.size();
  std::vector<double> pt((size_t)nds * 2), pc((size_t)nds * P * 2);
  std::vector<char> ok(nds);
  std::vector<double> ddr((size_t)ni * nb * P);        // rate differences
  for (int iv = 0; iv < ni; ++iv) {
    const double t = offsets[1] + (offsets[2] + ctx->ivals[iv].nsoff) / 1000000000.0;
    for (int d = 0; d < nds; ++d) {
      const fxg_datastream& s = ctx->dss[d];
      ok[d] = host_linear_interp(ctx->models[d].data(), s.model_order, s.model_nsamples, ctx->model_inc, t,
                                 0.000001, 1, &pt[2 * d]);
      for (int p = 0; p < P && ok[d]; ++p)
        ok[d] = host_linear_interp(ctx->pcmodels[d].data() + (size_t)p * s.model_nsamples * FXG_MAX_POLY,
                                   s.model_order, s.model_nsamples, ctx->model_inc, t, 0.000001, 1,
                                   &pc[((size_t)d * P + p) * 2]);
    }
    for (int j = 0; j < nb; ++j) {
      const int a = ctx->bls[j].ds1, b = ctx->bls[j].ds2;
      for (int p = 0; p < P; ++p) {
        double v = 0.0, r = 0.0;                      // outside the model: no shift
        if (ok[a] && ok[b]) {
          const double* p1 = &pc[((size_t)a * P + p) * 2];
          const double* p2 = &pc[((size_t)b * P + p) * 2];
          double a1 = p1[1] - pt[2 * a + 1], a2 = p2[1] - pt[2 * b + 1];
          a1 += a1 * pt[2 * a];
          a2 += a2 * pt[2 * b];
          v = a2 - a1;
          r = p2[0] + pt[2 * a] - (p1[0] + pt[2 * b]);
        }
        dd[((size_t)iv * nb + j) * P + p] = v;
        ddr[((size_t)iv * nb + j) * P + p] = r;
      }
    }
  }
  size_t k = 0;
  for (int f = 0; f < ctx->nfreqs; ++f) {
    if (!ctx->used[f]) continue;
    const fxg_freq& fr = ctx->freqs[f];
    const double window = fr.nchan / fr.bw_mhz;     // delaywindow (core.cpp:1621)
    for (int j = 0; j < nb; ++j)
      for (int kk = 0; kk < ctx->bls[j].nfreqs; ++kk)
        if (ctx->bls[j].freq_index[kk] == f)
          for (int p = 0; p < P; ++p) {
            float acc = 0.0f;
            for (int iv = 0; iv < ni; ++iv) {
              const double r = ddr[((size_t)iv * nb + j) * P + p], v = dd[((size_t)iv * nb + j) * P + p];
              double ts = 1.0, dl = 1.0;
              if (std::fabs(r) > 1e-18) {
                const double m = TWO_PI * r * (ctx->ivals[iv].nswidth / 1000.0) * fr.freq_mhz;
                ts = std::sin(m / 2.0) / (m / 2.0);
                if (ts < 0.0) ts = 0.0;
              }
              if (std::fabs(v) > 1e-18) {
                dl = 1.0 - std::fabs(v / window);
                if (dl < 0.0) dl = 0.0;
              }
              acc += ctx->ivals[iv].nswidth * ts * dl;
            }
            dec[k++] = acc;
          }
  }
}

int build_items(fxg_ctx* ctx) {
  int nds_kind[2] = {0, 0};
  for (const auto& fc : ctx->fclasses) nds_kind[fc.cplx] += (int)fc.dsl.size();
  for (int kind = 0; kind < 2; ++kind) {
    ctx->items[kind].clear();
    if (!nds_kind[kind]) continue;
    const int parts = std::max(1, std::min(64, (4 * ctx->ncu + nds_kind[kind] - 1) / nds_kind[kind]));
    std::vector<const fxg_ctx::FClass*> cls;
    for (const auto& fc : ctx->fclasses) if (fc.cplx == kind) cls.push_back(&fc);
    std::sort(cls.begin(), cls.end(), [](const fxg_ctx::FClass* a, const fxg_ctx::FClass* b) { return a->L > b->L; });
    ctx->iclass[kind].clear();
    for (const auto* fc : cls) {
      if (ctx->iclass[kind].empty() || ctx->iclass[kind].back().L != fc->L) {
        IClass ic;
        ic.L = fc->L;
        ic.log2L = log2i(fc->L);
        ic.begin = (int)ctx->items[kind].size();
        ic.count = 0;
        ctx->iclass[kind].push_back(ic);
      }
      for (int d : fc->dsl)
        for (int pt = 0; pt < parts; ++pt) {
          FItem it;
          it.L = fc->L;
          it.log2L = log2i(fc->L);
          it.d = d;
          it.part = pt;
          it.nparts = parts;
          ctx->items[kind].push_back(it);
          ctx->iclass[kind].back().count++;
        }
    }
    if ((int)ctx->iclass[kind].size() > FXG_MAX_ICLASS) return fail(ctx, FXG_E_INVALID, "too many FFT lengths");
    if (int rc = dev_upload(ctx, &ctx->d_items[kind], ctx->items[kind])) return rc;
    if (int rc = dev_upload(ctx, &ctx->d_iclass[kind], ctx->iclass[kind])) return rc;
  }
  // four-step classes: scratch rows of L cf32 per workgroup, shared by the classes (launched
  // one after another on the context's stream)
  long long rows = 0;
  int bigL = 0;
  for (auto& fc : ctx->fclasses) {
    if (!fc.big) continue;
    const int ncl = (int)fc.dsl.size();
    fc.wgx = std::max(1, std::min(ctx->chunkcap, (3 * ctx->ncu + ncl - 1) / ncl));   // 3 per CU
    rows = std::max(rows, (long long)fc.wgx * ncl * fc.L);
    bigL = std::max(bigL, fc.L);
  }
  // the persistent launch: up to 2 workgroups per CU of up to 512 threads (2 rows each)
  if (bigL) rows = std::max(rows, 4LL * ctx->ncu * bigL);
  ctx->bigL = bigL;
  if (rows > 0)
    if (int rc = dev_malloc(ctx, (void**)&ctx->d_bigscratch, (long long)sizeof(float2) * rows)) return rc;
  return dev_malloc(ctx, (void**)&ctx->d_counter, sizeof(int) * FXG_MAX_ICLASS);
}

// The F-engine of one chunk in the selected launch form.
int launch_fengines(fxg_ctx* ctx, FengineArgs& fa, hipStream_t st) {
  const int nunits = fa.nffts_pad / 8;
  for (const auto& fc : ctx->fclasses) {
    if (!fc.big || ctx->fsel.mode == FXG_F_PERSISTENT) continue;
    fa.twiddle = ctx->d_twiddle + fc.tw_off;
    fa.dslist = ctx->d_dslist + fc.dsl_off;
    fa.phcap = fc.phcap;
    if (fxg_launch_fbig(fa, fc.L, fc.cplx, (int)fc.dsl.size(), fc.wgx, ctx->d_bigscratch,
                        ctx->d_twiddle + fc.twL_off, st))
      return fail(ctx, FXG_E_HIP, "four-step F-engine launch");
  }
  if (ctx->fsel.mode == FXG_F_PERSISTENT) {
    for (int kind = 0; kind < 2; ++kind) {
      if (ctx->items[kind].empty()) continue;
      int phmax = 0;
      bool anybig = false;
      for (const auto& fc : ctx->fclasses) {
        if (fc.cplx != kind) continue;
        fa.twiddle = ctx->d_twiddle + fc.tw_off;
        fa.dslist = ctx->d_dslist + fc.dsl_off;
        fa.phcap = fc.phcap;
        if (fxg_launch_fftparams(fa, (int)fc.dsl.size(), st)) return fail(ctx, FXG_E_HIP, "FFT parameter launch");
        fa.tw_by_log[log2i(fc.L) - 8] = ctx->d_twiddle + (fc.big ? fc.twL_off : fc.tw_off);
        if (fc.big) { fa.tw2048 = ctx->d_twiddle + fc.tw_off; anybig = true; }
        phmax = std::max(phmax, fc.phcap);
      }
      fa.bigscratch = ctx->d_bigscratch;
      fa.bigL = ctx->bigL;
      const int ntb = ctx->fsel.ntb ? ctx->fsel.ntb : 512;      // 512 measured faster on uneven bands
      size_t lds = 0;
      for (const auto& fc : ctx->fclasses)
        if (fc.cplx == kind && !fc.big) lds = std::max(lds, fxg_fengine_lds(fc.L, ntb, phmax, ctx->rotorder));
      if (anybig) lds = std::max(lds, fxg_fbig_lds(ntb, phmax));
      fa.phcap = phmax;
      const int nitems = (int)ctx->items[kind].size();
      const int nwg = std::max(1, std::min(nitems, ctx->ncu * ctx->fsel.grid_mult));
      if (fxg_launch_fpersist(fa, ctx->d_items[kind], ctx->d_iclass[kind], (int)ctx->iclass[kind].size(),
                              ctx->d_counter, nwg, ntb, kind, lds, st))
        return fail(ctx, FXG_E_HIP, "persistent F-engine launch");
    }
    return FXG_OK;
  }
  for (const auto& fc : ctx->fclasses) {
    if (fc.big) continue;
    const int ncl = (int)fc.dsl.size();
    // enough workgroups to fill the CUs several times over
    const long long wgs_per_ds = std::max(1LL, (long long)(8 * ctx->ncu + ncl - 1) / ncl);
    fa.units_per_wg = (int)std::max(1LL, std::min<long long>(8, nunits / wgs_per_ds));
    fa.twiddle = ctx->d_twiddle + fc.tw_off;
    fa.dslist = ctx->d_dslist + fc.dsl_off;
    fa.phcap = fc.phcap;
    if (fxg_launch_fengine(fa, fc.L, fc.cplx, ncl, ntb_for(ctx, fc.L), st))
      return fail(ctx, FXG_E_HIP, "F-engine launch");
  }
  return FXG_OK;
}

int run_subint(fxg_ctx* ctx, const int32_t offsets[3], const uint8_t* const* d_data, const int32_t* nbytes,
               const int32_t* const* control, float2* d_results, hipStream_t st) {
  const int nds = ctx->nds;
  for (int d = 0; d < nds; ++d) {
    if (!control || !control[d]) return fail(ctx, FXG_E_INVALID, "missing control buffer");
    if (nbytes[d] > 1 && !d_data[d]) return fail(ctx, FXG_E_INVALID, "missing data buffer");
    if (nbytes[d] < 0) return fail(ctx, FXG_E_INVALID, "negative data length");
    if ((long long)(nbytes[d] / ctx->dss[d].framebytes) * ctx->h_ds[d].spf >= (1LL << 31))
      return fail(ctx, FXG_E_INVALID, "slot buffer holds more than 2^31 samples");
    const long long capframes = (ctx->databytes[d] + 4LL * ctx->dss[d].framebytes) / ctx->dss[d].framebytes;
    if (nbytes[d] / ctx->dss[d].framebytes > capframes)
      return fail(ctx, FXG_E_INVALID, "data buffer larger than the slot capacity");
  }
  // staging of the per-subint datastream state and valid flags (double buffered)
  Staging& sg = ctx->stage[ctx->stage_next];
  ctx->stage_next ^= 1;
  if (sg.pending) {
    HIPCHK(hipEventSynchronize(sg.used));
    sg.pending = false;
  }
  int maxframes = 0;
  int32_t* d_flags = ctx->d_flags + (size_t)(&sg - ctx->stage) * nds * ctx->flaglen;
  for (int d = 0; d < nds; ++d) {
    DevDsRun& r = sg.h_run[d];
    std::memset(&r, 0, sizeof(r));
    r.scan = control[d][0];
    r.datasec = control[d][1];
    r.datans = control[d][2];
    r.datalen = nbytes[d];
    r.data = d_data[d];
    r.nframes = nbytes[d] / ctx->dss[d].framebytes;
    r.valid = (nbytes[d] > 1) && (r.datasec > INVALID_SUBINT) ? 1 : 0;
    r.vflags = d_flags + (size_t)d * ctx->flaglen;
    std::memcpy(sg.h_flags + (size_t)d * ctx->flaglen, control[d] + 3, sizeof(int32_t) * ctx->flaglen);
    maxframes = std::max(maxframes, r.nframes);
  }
  HIPCHK(hipMemcpyAsync(ctx->d_run, sg.h_run, sizeof(DevDsRun) * nds, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d_flags, sg.h_flags, sizeof(int32_t) * nds * ctx->flaglen, hipMemcpyHostToDevice, st));
  const bool uvs = ctx->nphase > 1;
  if (uvs) {                                       // uv-shift delays and decorrelation of this subint
    uvshift_tables(ctx, offsets, sg.h_uvdd, sg.h_uvdec);
    HIPCHK(hipMemcpyAsync(sg.d_uvdd, sg.h_uvdd, sizeof(double) * ctx->ivals.size() * ctx->nbl * ctx->nphase,
                          hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(sg.d_uvdec, sg.h_uvdec, sizeof(float) * ctx->uvdec_idx.size(), hipMemcpyHostToDevice, st));
  }
  HIPCHK(hipEventRecord(sg.used, st));
  sg.pending = true;

  const bool mfma = ctx->xengine != FXG_X_FP32;
  if (fxg_launch_zero(d_results, ctx->coreresultlength * 8, st)) return fail(ctx, FXG_E_HIP, "zero results");
  if (mfma) {
    if (fxg_launch_zero(ctx->d_vacc, ctx->vacc_len * 4, st)) return fail(ctx, FXG_E_HIP, "zero acc");
  } else {
    if (fxg_launch_zero(ctx->d_xacc, ctx->xacc_len * 8, st)) return fail(ctx, FXG_E_HIP, "zero acc");
  }
  SetupArgs sa;
  sa.ds = ctx->d_ds;
  sa.run = ctx->d_run;
  sa.model = ctx->d_model;
  sa.fvalid = ctx->d_fvalid;
  sa.fzone = ctx->d_fzone;
  sa.nds = nds;
  sa.offscan = offsets[0];
  sa.offsec = offsets[1];
  sa.offns = offsets[2];
  sa.model_inc_secs = ctx->model_inc;
  sa.B = ctx->B;
  {
    TimedLaunch tl(ctx, 2, st);
    if (fxg_launch_setup(sa, std::max(maxframes, 1), st)) return fail(ctx, FXG_E_HIP, "setup launch");
  }

  FengineArgs fa;
  std::memset(&fa, 0, sizeof(fa));
  fa.ds = ctx->d_ds;
  fa.run = ctx->d_run;
  fa.twiddle = ctx->d_twiddle;
  fa.params = ctx->d_params;
  fa.weights = ctx->d_weights;
  fa.B = ctx->B;
  fa.offsec = offsets[1];
  fa.offns = offsets[2];
  fa.scan_start_sec = ctx->scan_start;
  fa.out_mode = ctx->xengine;
  fa.spec32 = ctx->d_spec32;
  fa.spec16 = ctx->d_spec16;
  fa.xgroups = ctx->d_groups;
  fa.xscale = ctx->d_xscale;
  fa.chanfreq = ctx->d_chanfreq;
  fa.fvalid = ctx->d_fvalid;
  fa.fzone = ctx->d_fzone;
  fa.chunkcap = ctx->chunkcap;
  fa.rotorder = ctx->rotorder;
  if (ctx->kurt_on && fxg_launch_zero(ctx->d_kurt_acc, sizeof(float2) * ctx->kurt_len, st))
    return fail(ctx, FXG_E_HIP, "kurtosis zero launch");
  PcalArgs pa{};
  pa.ds = ctx->d_ds;
  pa.run = ctx->d_run;
  pa.params = ctx->d_params;
  pa.fvalid = ctx->d_fvalid;
  pa.fzone = ctx->d_fzone;
  pa.groups = ctx->d_pcal_groups;
  pa.part = ctx->d_pcal_part;
  pa.bins = ctx->d_pcal_bins;
  pa.chunkcap = ctx->chunkcap;
  pa.nsplit = ctx->pcal_nsplit;
  const int npg = (int)ctx->pcal_groups.size();
  if (npg && fxg_launch_zero(ctx->d_pcal_part, sizeof(float) * ctx->pcal_part_len, st))
    return fail(ctx, FXG_E_HIP, "phase-cal zero launch");
  // The subint in chunks of at most chunkcap FFTs; with several phase centres each uv-shift
  // interval separately, its products shifted into every phase centre's block at its end
  // (core.cpp:984-991, 1055-1057) and the accumulators cleared for the next interval.
  std::vector<std::array<int, 3>> pieces;          // {first FFT, FFTs, interval ending here or -1}
  if (!uvs) {
    for (int c0 = 0; c0 < ctx->B; c0 += ctx->chunkcap) pieces.push_back({c0, std::min(ctx->chunkcap, ctx->B - c0), -1});
  } else {
    for (size_t iv = 0; iv < ctx->ivals.size(); ++iv) {
      const int e = ctx->ivals[iv].i0 + ctx->ivals[iv].n;
      for (int c0 = ctx->ivals[iv].i0; c0 < e; c0 += ctx->chunkcap) {
        const int n = std::min(ctx->chunkcap, e - c0);
        pieces.push_back({c0, n, c0 + n == e ? (int)iv : -1});
      }
    }
  }
  UvArgs uv{nullptr, ctx->d_uvfreq, uvs ? ctx->nphase : 1};
  auto accumulate = [&](const UvArgs& u) -> int {
    TimedLaunch tl(ctx, 2, st);
    if (mfma) {
      if (fxg_launch_fin_mfma(ctx->d_finm, (int)ctx->finm.size(), ctx->maxoutm, ctx->d_groups, ctx->d_vacc,
                              ctx->d_xscale, d_results, u, st))
        return fail(ctx, FXG_E_HIP, "accumulate launch");
    } else {
      if (fxg_launch_fin_plain(ctx->d_finp, (int)ctx->finp.size(), ctx->maxoutp, ctx->d_xacc, d_results, u, st))
        return fail(ctx, FXG_E_HIP, "accumulate launch");
    }
    return FXG_OK;
  };
  for (const auto& piece : pieces) {
    const int c0 = piece[0], n = piece[1];
    const int npad = (int)round_up(n, FXG_XPAD);
    fa.chunk0 = c0;
    fa.nffts = n;
    fa.nffts_pad = npad;
    {
      TimedLaunch tl(ctx, 0, st);
      if (int rc = launch_fengines(ctx, fa, st)) return rc;
    }
    if (npg) {                                     // while the chunk's per-FFT delays are current
      pa.nffts = n;
      TimedLaunch tp(ctx, 2, st);
      for (const auto& r : ctx->pcal_runs) {
        PcalArgs pr = pa;
        pr.groups = ctx->d_pcal_groups + r[1];
        if (fxg_launch_pcal(pr, r[0], r[2], ctx->pcal_max_gsize, st))
          return fail(ctx, FXG_E_HIP, "phase-cal launch");
      }
    }
    if (mfma) {
      XMfmaArgs xa;
      xa.groups = ctx->d_groups;
      xa.spec16 = ctx->d_spec16;
      xa.acc = ctx->d_vacc;
      xa.nffts_pad = npad;
      xa.chunkcap = ctx->chunkcap;
      xa.mode = ctx->xengine;
      TimedLaunch tl(ctx, 1, st);
      // one launch per run of groups sharing tile count and channel count
      size_t g0 = 0;
      while (g0 < ctx->groups.size()) {
        size_t g1 = g0 + 1;
        while (g1 < ctx->groups.size() && ctx->groups[g1].ntile == ctx->groups[g0].ntile &&
               ctx->groups[g1].C == ctx->groups[g0].C && ctx->groups[g1].twlog == ctx->groups[g0].twlog)
          ++g1;
        if (fxg_launch_xmfma(xa, (int)(g1 - g0), (int)g0, ctx->groups[g0].ntile, ctx->groups[g0].C,
                             ctx->groups[g0].twlog, st))
          return fail(ctx, FXG_E_HIP, "MFMA X-engine launch");
        g0 = g1;
      }
    } else {
      TimedLaunch tl(ctx, 1, st);
      if (!ctx->xp_items.empty()) {
        XPBlkArgs ba;
        ba.items = ctx->d_xp_items;
        ba.groups = ctx->d_xp_groups;
        ba.rows = ctx->d_xp_rows;
        ba.pairs = ctx->d_xp_pairs;
        ba.spec32 = ctx->d_spec32;
        ba.acc = ctx->d_xacc;
        ba.nffts = n;
        if (fxg_launch_xplain_blk(ba, (int)ctx->xp_items.size(), st)) return fail(ctx, FXG_E_HIP, "X-engine launch");
      }
      if (!ctx->prods_rest.empty()) {
        XPlainArgs xa;
        xa.prods = ctx->d_prods_rest;
        xa.ds = ctx->d_ds;
        xa.spec32 = ctx->d_spec32;
        xa.acc = ctx->d_xacc;
        xa.nffts = n;
        xa.chunkcap = ctx->chunkcap;
        if (fxg_launch_xplain(xa, (int)ctx->prods_rest.size(), ctx->maxchan, st))
          return fail(ctx, FXG_E_HIP, "X-engine launch");
      }
    }
    if (ctx->kurt_on) {                            // before the next chunk overwrites the spectra
      KurtArgs ka;
      ka.tab = ctx->d_kurt_tab;
      ka.spec32 = ctx->d_spec32;
      ka.spec16 = ctx->d_spec16;
      ka.xscale = ctx->d_xscale;
      ka.acc = ctx->d_kurt_acc;
      ka.nffts = n;
      ka.chunkcap = ctx->chunkcap;
      ka.mode = ctx->xengine;
      TimedLaunch tk(ctx, 2, st);
      int mintw = 4;
      for (const auto& kb : ctx->kurt_tab) mintw = std::min(mintw, kb.twlog);
      if (fxg_launch_kurt(ka, (int)ctx->kurt_tab.size(), ctx->kurt_maxC, mintw, st))
        return fail(ctx, FXG_E_HIP, "kurtosis launch");
    }
    ctx->last_chunk0 = c0;
    ctx->last_nffts = n;
    if (piece[2] >= 0) {                           // end of a uv-shift interval
      uv.dd = sg.d_uvdd + (size_t)piece[2] * ctx->nbl * ctx->nphase;
      if (int rc = accumulate(uv)) return rc;
      if (mfma ? fxg_launch_zero(ctx->d_vacc, ctx->vacc_len * 4, st) : fxg_launch_zero(ctx->d_xacc, ctx->xacc_len * 8, st))
        return fail(ctx, FXG_E_HIP, "zero acc");
    }
  }
  if (npg) {
    TimedLaunch tp(ctx, 2, st);
    if (fxg_launch_pcal_reduce(pa, npg, ctx->pcal_max_gsize, st))
      return fail(ctx, FXG_E_HIP, "phase-cal reduce launch");
    // the tones (getFinalPCal), also added into the results' phase-cal block
    if (fxg_launch_pcal_tones(ctx->d_pcal_tdev, (int)ctx->pcal_tdev.size(), ctx->d_pcal_bins, ctx->d_pcal_tonebuf,
                              d_results, st))
      return fail(ctx, FXG_E_HIP, "phase-cal tone launch");
  }
  if (!uvs)
    if (int rc = accumulate(uv)) return rc;
  TimedLaunch tl(ctx, 2, st);
  if (fxg_launch_fin_weights(ctx->d_finw, (int)ctx->finw.size(), ctx->d_weights, ctx->B, (float*)d_results, st))
    return fail(ctx, FXG_E_HIP, "weights launch");
  if (uvs && fxg_launch_add_at(ctx->d_uvdec_idx, sg.d_uvdec, (int)ctx->uvdec_idx.size(), (float*)d_results, st))
    return fail(ctx, FXG_E_HIP, "shift decorrelation launch");
  return FXG_OK;
}

// Autotuner: times one synthetic subint (zeroed slot buffers: the F-engine's control flow and
// arithmetic do not depend on the data, only its loads of valid frames do) in every F-engine
// launch form and workgroup size, and keeps the fastest.  The forms give bit-identical
// results, so the choice only affects speed.
int autotune(fxg_ctx* ctx) {
  std::vector<std::pair<std::string, fxg_ctx::FChoice>> cand;
  auto add = [&](const char* name, int mode, int ntb, int mult) {
    if (ctx->cfg_fmode != FXG_F_AUTO && ctx->cfg_fmode != mode) return;
    if (ctx->cfg_fthreads && ntb && ctx->cfg_fthreads != ntb) return;
    fxg_ctx::FChoice c;
    c.mode = mode;
    c.ntb = ntb;
    c.grid_mult = mult;
    cand.emplace_back(name, c);
  };
  add("per-length, 256/512 threads by length", FXG_F_PER_LENGTH, 0, 1);
  add("per-length, 256 threads", FXG_F_PER_LENGTH, 256, 1);
  add("per-length, 512 threads", FXG_F_PER_LENGTH, 512, 1);
  add("per-length, 128 threads", FXG_F_PER_LENGTH, 128, 1);
  add("persistent, 256 threads, 1 workgroup/CU", FXG_F_PERSISTENT, 256, 1);
  add("persistent, 256 threads, 2 workgroups/CU", FXG_F_PERSISTENT, 256, 2);
  add("persistent, 512 threads, 1 workgroup/CU", FXG_F_PERSISTENT, 512, 1);
  if (cand.empty()) return FXG_OK;
  // synthetic slot: zeroed data, every FFT flagged valid, subint at the model's start
  const int nds = ctx->nds;
  std::vector<uint8_t*> bufs(nds, nullptr);
  std::vector<std::vector<int32_t>> ctl(nds, std::vector<int32_t>(3 + ctx->flaglen, 0));
  std::vector<const uint8_t*> ptr(nds);
  std::vector<const int32_t*> cptr(nds);
  std::vector<int32_t> nb(nds);
  float2* d_res = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = FXG_OK;
  for (int d = 0; d < nds && rc == FXG_OK; ++d) {
    nb[d] = (int32_t)ctx->databytes[d];
    if (hipMalloc((void**)&bufs[d], (size_t)nb[d]) != hipSuccess || hipMemset(bufs[d], 0, (size_t)nb[d]) != hipSuccess)
      rc = fail(ctx, FXG_E_NOMEM, "autotune buffers");
    for (int k = 0; k < ctx->flaglen; ++k) ctl[d][3 + k] = (1 << FLAGS_PER_INT) - 1;
    ptr[d] = bufs[d];
    cptr[d] = ctl[d].data();
  }
  if (rc == FXG_OK && (hipMalloc((void**)&d_res, sizeof(float2) * std::max<long long>(ctx->coreresultlength, 1)) != hipSuccess ||
                       hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = fail(ctx, FXG_E_NOMEM, "autotune buffers");
  const int32_t offs[3] = {0, 0, 0};
  double best = 1e30;
  fxg_ctx::FChoice bestc = ctx->fsel;
  const bool timing = ctx->timing;
  ctx->timing = false;
  for (auto& c : cand) {
    if (rc != FXG_OK) break;
    ctx->fsel = c.second;
    double ms = 1e30;
    bool ok = run_subint(ctx, offs, ptr.data(), nb.data(), cptr.data(), d_res, ctx->stream) == FXG_OK;   // warm
    for (int rep = 0; ok && rep < 2; ++rep) {
      ok = hipEventRecord(e0, ctx->stream) == hipSuccess &&
           run_subint(ctx, offs, ptr.data(), nb.data(), cptr.data(), d_res, ctx->stream) == FXG_OK &&
           hipEventRecord(e1, ctx->stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
      float t = 0.0f;
      if (ok && hipEventElapsedTime(&t, e0, e1) == hipSuccess) ms = std::min(ms, (double)t);
    }
    if (!ok) {
      (void)hipGetLastError();
      (void)hipStreamSynchronize(ctx->stream);
      ctx->err.clear();
    }
    ctx->tune_log.emplace_back(c.first, ok ? ms : -1.0);
    if (ok && ms < best) { best = ms; bestc = c.second; }
  }
  ctx->fsel = bestc;
  if (best >= 1e30 && rc == FXG_OK) rc = fail(ctx, FXG_E_HIP, "autotune: no F-engine launch form ran");
  // the MFMA X-engine's channel tile (8 or 16 channels per workgroup, which is also the tile
  // width of the spectra layout the F-engine writes) for groups of one 32-row tile, timed with
  // the chosen F-engine form
  bool xtile = false;
  for (const auto& g : ctx->groups) xtile |= tile_width_log(g, 4) == 4;
  if (rc == FXG_OK && xtile && ctx->xengine != FXG_X_FP32 && !ctx->d_kurt_tab) {
    double xbest = 1e30;
    int xw = 3;
    for (int want : {3, 4}) {
      for (auto& g : ctx->groups) g.twlog = tile_width_log(g, want);
      if (hipMemcpy(ctx->d_groups, ctx->groups.data(), sizeof(DevXGroup) * ctx->groups.size(),
                    hipMemcpyHostToDevice) != hipSuccess) { rc = fail(ctx, FXG_E_HIP, "autotune: group upload"); break; }
      double ms = 1e30;
      bool ok = run_subint(ctx, offs, ptr.data(), nb.data(), cptr.data(), d_res, ctx->stream) == FXG_OK;
      for (int rep = 0; ok && rep < 2; ++rep) {
        ok = hipEventRecord(e0, ctx->stream) == hipSuccess &&
             run_subint(ctx, offs, ptr.data(), nb.data(), cptr.data(), d_res, ctx->stream) == FXG_OK &&
             hipEventRecord(e1, ctx->stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
        float t = 0.0f;
        if (ok && hipEventElapsedTime(&t, e0, e1) == hipSuccess) ms = std::min(ms, (double)t);
      }
      if (!ok) {
        (void)hipGetLastError();
        (void)hipStreamSynchronize(ctx->stream);
        ctx->err.clear();
      }
      ctx->tune_log.emplace_back(want == 4 ? "MFMA X-engine, 16-channel tiles (16 channels per workgroup)"
                                           : "MFMA X-engine, 8-channel tiles (8 channels per workgroup)",
                                 ok ? ms : -1.0);
      if (ok && ms < xbest) { xbest = ms; xw = want; }
    }
    for (auto& g : ctx->groups) g.twlog = tile_width_log(g, xw);
    if (rc == FXG_OK && hipMemcpy(ctx->d_groups, ctx->groups.data(), sizeof(DevXGroup) * ctx->groups.size(),
                                  hipMemcpyHostToDevice) != hipSuccess)
      rc = fail(ctx, FXG_E_HIP, "autotune: group upload");
  }
  ctx->timing = timing;
  (void)hipStreamSynchronize(ctx->stream);
  for (auto b : bufs) if (b) (void)hipFree(b);
  if (d_res) (void)hipFree(d_res);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  return rc;
}

void free_ctx(fxg_ctx* ctx) {
  auto f = [](void* p) { if (p) (void)hipFree(p); };
  f(ctx->d_ds); f(ctx->d_run); f(ctx->d_model); f(ctx->d_chanfreq); f(ctx->d_fvalid); f(ctx->d_fzone);
  f(ctx->d_twiddle); f(ctx->d_dslist); f(ctx->d_params); f(ctx->d_items[0]); f(ctx->d_items[1]);
  f(ctx->d_iclass[0]); f(ctx->d_iclass[1]);
  f(ctx->d_counter); f(ctx->d_bigscratch); f(ctx->d_weights); f(ctx->d_flags); f(ctx->d_spec32); f(ctx->d_spec16);
  f(ctx->d_prods); f(ctx->d_xp_groups); f(ctx->d_xp_rows); f(ctx->d_xp_pairs); f(ctx->d_xp_items);
  f(ctx->d_prods_rest); f(ctx->d_xacc); f(ctx->d_groups); f(ctx->d_vacc); f(ctx->d_xscale);
  f(ctx->d_finp); f(ctx->d_finm); f(ctx->d_finw);
  f(ctx->d_normseg); f(ctx->d_normcal); f(ctx->d_calmean); f(ctx->d_tsys); f(ctx->d_dpct);
  f(ctx->d_kurt_tab); f(ctx->d_kurt_acc); f(ctx->d_pcal_groups); f(ctx->d_pcal_part); f(ctx->d_pcal_bins); f(ctx->d_pcal_tdev); f(ctx->d_pcal_tonebuf);
  f(ctx->d_uvfreq); f(ctx->d_uvdec_idx);
  for (auto& sg : ctx->stage) {
    f(sg.d_uvdd); f(sg.d_uvdec);
    if (sg.h_run) (void)hipHostFree(sg.h_run);
    if (sg.h_flags) (void)hipHostFree(sg.h_flags);
    if (sg.h_uvdd) (void)hipHostFree(sg.h_uvdd);
    if (sg.h_uvdec) (void)hipHostFree(sg.h_uvdec);
    if (sg.used) (void)hipEventDestroy(sg.used);
  }
  for (auto& s : ctx->slots) {
    for (auto p : s.d_data) f(p);
    f(s.d_results);
    if (s.h_results) (void)hipHostFree(s.h_results);
    if (s.copy) (void)hipStreamDestroy(s.copy);
    if (s.copied) (void)hipEventDestroy(s.copied);
    if (s.done) (void)hipEventDestroy(s.done);
  }
  for (auto& t : ctx->tev) {
    (void)hipEventDestroy(std::get<1>(t));
    (void)hipEventDestroy(std::get<2>(t));
  }
  for (auto e : ctx->evpool) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
}

int ensure_slots(fxg_ctx* ctx) {
  for (auto& s : ctx->slots) {
    if (s.copy) continue;
    HIPCHK(hipStreamCreateWithFlags(&s.copy, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&s.copied, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
    s.d_data.assign(ctx->nds, nullptr);
    s.cap.assign(ctx->nds, 0);
    for (int d = 0; d < ctx->nds; ++d) {
      s.cap[d] = ctx->databytes[d] + 4LL * ctx->dss[d].framebytes;
      if (int rc_ = dev_malloc(ctx, (void**)&s.d_data[d], s.cap[d])) return rc_;
    }
    if (int rc_ = dev_malloc(ctx, (void**)&s.d_results, sizeof(float2) * std::max<long long>(ctx->coreresultlength, 1))) return rc_;
    HIPCHK(hipHostMalloc((void**)&s.h_results, sizeof(float2) * std::max<long long>(ctx->coreresultlength, 1), 0));
  }
  return FXG_OK;
}

}  // namespace

extern "C" {

const char* fxg_strerror(int code) {
  switch (code) {
    case FXG_OK: return "ok";
    case FXG_E_INVALID: return "invalid argument or unsupported configuration";
    case FXG_E_HIP: return "HIP runtime error";
    case FXG_E_NOMEM: return "out of memory";
    case FXG_E_STATE: return "call out of order";
    case FXG_E_NODEVICE: return "no usable gfx950 device";
  }
  return "unknown error";
}

const char* fxg_last_error(const fxg_ctx* ctx) {
  if (ctx) return ctx->err.c_str();
  return create_error().empty() ? "null context" : create_error().c_str();
}

}  // extern "C"

namespace {

// fxg_create, or with dry set a pass that validates and sizes everything without touching
// the device (fxg_footprint).
// Phase-cal extraction of datastream ds with a whole-Hz interval (0: off), and the tables of
// every datastream with it on: per recorded band the tone table (Configuration,
// configuration.cpp:1765-1835), the rational frequency of each tone the reference's extractor
// reads (PCal::getNew and the extractors' getFinalPCal, pcal.cpp:191-274), the fold period,
// the groups of sample fields the device folds together and the device tone table.
int pcal_configure(fxg_ctx* ctx, int ds, long long interval_hz, long long base_hz) {
  const fxg_datastream& xs = ctx->dss[ds];
  if (interval_hz > 0) {
    if (xs.nbits != 2 && xs.nbits != 8)
      return fail(ctx, FXG_E_INVALID, "phase-cal extraction needs 2-bit or 8-bit samples");
    if (xs.complex_sampling && xs.nbits != 8)
      return fail(ctx, FXG_E_INVALID, "phase-cal extraction of complex data needs 8-bit samples");
    if (xs.framebytes % 4 || xs.headerbytes % 4)
      return fail(ctx, FXG_E_INVALID, "phase-cal extraction needs 4-byte aligned frame payloads");
  }
  ctx->pcal_interval.resize(ctx->nds, 0);
  ctx->pcal_base.resize(ctx->nds, 0);
  const long long prev_interval = ctx->pcal_interval[ds], prev_base = ctx->pcal_base[ds];
  ctx->pcal_interval[ds] = interval_hz;
  ctx->pcal_base[ds] = base_hz;
  auto refuse = [&](const char* why) {             // keep the previous setting and tables
    ctx->pcal_interval[ds] = prev_interval;
    ctx->pcal_base[ds] = prev_base;
    return fail(ctx, FXG_E_INVALID, why);
  };
  std::vector<fxg_ctx::PcalTones> tones;
  std::vector<fxg_ctx::PcalBins> tab;
  std::vector<int> entry;
  std::vector<PcalGroup> groups;
  std::vector<PcalToneDev> tdev;
  long long off = 0;
  int ntone_total = 0;
  for (int d = 0; d < ctx->nds; ++d) {
    if (ctx->pcal_interval[d] <= 0) continue;
    const fxg_datastream& x = ctx->dss[d];
    const bool cplx = x.complex_sampling != 0, dsb = x.complex_sampling == 2;
    const int nf = cplx ? 2 : 1;                  // sample fields per band
    int res = ctx->cfg_pcal_interval[d] > 0 ? ctx->pcaloff[d] : -1;   // Core::copyPCalTones order
    for (int b = 0; b < x.nrecbands; ++b) {
      const fxg_freq& f = ctx->freqs[x.recfreq_index[x.recband_localfreq[b]]];
      fxg_ctx::PcalTones t{};
      t.ds = d;
      t.band = b;
      t.interval_hz = ctx->pcal_interval[d];
      t.fs_hz = (long long)std::llround((cplx ? 1e6 : 2e6) * f.bw_mhz);
      pcal_tones_of(f.freq_mhz, f.bw_mhz, f.lsb != 0, t.interval_hz, ctx->pcal_base[d], &t.ntones, &t.offset_hz);
      t.tone0 = ntone_total;
      ntone_total += t.ntones;
      // the tones the extractor reads, as reduced fractions num/den of cycles per sample
      std::vector<std::pair<long long, long long>> fr((size_t)t.ntones, {-1, -1});
      auto red = [](long long num, long long den) {
        num %= den;
        const long long g = gcd_ll(num, den);
        return std::make_pair(num / g, den / g);
      };
      const long long fs = t.fs_hz, iv = t.interval_hz, of = t.offset_hz;
      if (of >= 0) {
        if (!cplx) {                              // trivial, implicit-shift and shifting extractors
          const long long first = of == 0 ? iv : of;
          for (int k = 0; k < t.ntones; ++k) fr[k] = red(first + k * iv, fs);
        } else {
          const long long no = fs / gcd_ll(fs, of), np = fs / gcd_ll(fs, iv);
          const bool implicit = no % np == 0;     // PCalExtractorImplicitShiftComplex, else Shifting
          long long n = implicit ? no : np;       // DFT length: doubled to 1 MHz resolution (mode.cpp:402)
          while (fs / (2 * n) > 1000000) n *= 2;
          const long long step = (long long)std::floor((double)n * (double)iv / (double)fs);
          const long long obin = implicit ? (long long)std::floor((double)n * (double)of / (double)fs) : 0;
          for (int k = 0; k < t.ntones; ++k) {
            const long long bin = obin + (long long)k * step;
            if (bin >= n) break;                  // the extractor stops there: the tone stays zero
            const long long bb = dsb ? n - 1 - bin : bin;   // DSB: the DFT output reversed
            fr[k] = implicit ? red(bb, n) : red(of * n + bb * fs, fs * n);
          }
        }
      }
      long long M = 0;                            // fold period
      if (!cplx) {
        if (of >= 0) M = fs / std::max<long long>(gcd_ll(fs, gcd_ll(of, iv)), 1);   // offset 0: gcd = interval
      } else {
        for (const auto& q : fr)
          if (q.second > 0) {
            M = M == 0 ? q.second : M / gcd_ll(M, q.second) * q.second;
            if (M > 16384) break;
          }
      }
      int e = -1;
      if (M > 0) {
        if (M > 16384) return refuse("phase-cal tone period longer than 16384 samples");
        for (int fd = 0; fd < nf; ++fd) {         // fields of the band: real (and imaginary)
          const int field = b * nf + fd;
          // extend the open group (same datastream, the previous field, same M, room in LDS)
          PcalGroup* gr = groups.empty() ? nullptr : &groups.back();
          if (!(gr && gr->ds == d && gr->band0 + gr->nb == field && gr->M == M && gr->nb < PCAL_MAXNB &&
                (long long)(gr->nb + 1) * M <= PCAL_MAX_GSIZE)) {
            PcalGroup ng{};
            ng.ds = d;
            ng.band0 = field;
            ng.nb = 0;
            ng.M = (int)M;
            ng.bins_off = (int)off;
            ng.fps = x.nrecbands * nf;
            groups.push_back(ng);
            gr = &groups.back();
          }
          gr->nb += 1;
          if (fd == 0) {
            e = (int)tab.size();
            tab.push_back({(int)M, (int)off});
          }
          off += M;
        }
        for (int k = 0; k < t.ntones; ++k)
          if (fr[k].second > 0)
            tdev.push_back(PcalToneDev{tab[e].bins_off, (int)M, cplx ? 1 : 0, (int)fr[k].first, (int)fr[k].second,
                                       t.tone0 + k, res >= 0 ? res + k : -1});
      }
      if (res >= 0) res += t.ntones;
      entry.push_back(e);
      tones.push_back(t);
    }
  }
  // kernel form of each group, and the groups ordered by form (one launch per form)
  for (auto& g : groups) {
    const fxg_datastream& x = ctx->dss[g.ds];
    g.form = (256 % g.M == 0 ? PCAL_FORM_REG : 0) | (x.nbits == 8 ? PCAL_FORM_8BIT : 0) |
             (g.fps * x.nbits == 32 && g.band0 == 0 && g.nb == g.fps ? PCAL_FORM_WHOLE : 0);
  }
  std::stable_sort(groups.begin(), groups.end(),
                   [](const PcalGroup& p, const PcalGroup& q) { return p.form < q.form; });
  std::vector<std::array<int, 3>> runs;
  for (int i = 0; i < (int)groups.size(); ++i) {
    if (runs.empty() || runs.back()[0] != groups[i].form) runs.push_back({groups[i].form, i, 0});
    runs.back()[2] += 1;
  }
  // workgroups per group: about eight per CU over all groups, at most one per FFT of a chunk
  const int ng = (int)groups.size();
  const int nsplit = ng ? std::max(1, std::min(ctx->chunkcap, (8 * ctx->ncu + ng - 1) / ng)) : 1;
  long long part = 0;
  int max_gsize = 0;
  for (auto& g : groups) {
    g.part_off = part;
    part += (long long)nsplit * g.nb * g.M;
    max_gsize = std::max(max_gsize, g.nb * g.M);
  }
  if (ctx->d_pcal_groups) { (void)hipFree(ctx->d_pcal_groups); ctx->d_pcal_groups = nullptr; }
  if (ctx->d_pcal_part) { (void)hipFree(ctx->d_pcal_part); ctx->d_pcal_part = nullptr; }
  if (ctx->d_pcal_bins) { (void)hipFree(ctx->d_pcal_bins); ctx->d_pcal_bins = nullptr; }
  if (ctx->d_pcal_tdev) { (void)hipFree(ctx->d_pcal_tdev); ctx->d_pcal_tdev = nullptr; }
  if (ctx->d_pcal_tonebuf) { (void)hipFree(ctx->d_pcal_tonebuf); ctx->d_pcal_tonebuf = nullptr; }
  ctx->pcal_tones = tones;
  ctx->pcal_tab = tab;
  ctx->pcal_entry = entry;
  ctx->pcal_groups = groups;
  ctx->pcal_nsplit = nsplit;
  ctx->pcal_runs = runs;
  ctx->pcal_part_len = part;
  ctx->pcal_max_gsize = max_gsize;
  ctx->pcal_tdev = tdev;
  ctx->pcal_ntone_total = ntone_total;
  if (!ctx->dry && ntone_total > 0) {             // a host-only context keeps the tables only
    if (int rc = dev_malloc(ctx, (void**)&ctx->d_pcal_tonebuf, sizeof(float2) * ntone_total)) return rc;
    if (int rc = dev_zero(ctx, ctx->d_pcal_tonebuf, sizeof(float2) * ntone_total)) return rc;
  }
  if (!groups.empty() && !ctx->dry) {
    if (int rc = dev_upload(ctx, &ctx->d_pcal_groups, ctx->pcal_groups)) return rc;
    if (int rc = dev_malloc(ctx, (void**)&ctx->d_pcal_part, sizeof(float) * part)) return rc;
    if (int rc = dev_malloc(ctx, (void**)&ctx->d_pcal_bins, sizeof(float) * off)) return rc;
    if (int rc = dev_zero(ctx, ctx->d_pcal_bins, sizeof(float) * off)) return rc;
    if (int rc = dev_upload(ctx, &ctx->d_pcal_tdev, ctx->pcal_tdev)) return rc;
  }
  return FXG_OK;
}

int create_impl(const fxg_config* cfg, fxg_ctx** out, bool dry) {
  create_error().clear();
  if (!cfg || !out) {
    create_error() = "null configuration or output pointer";
    return FXG_E_INVALID;
  }
  *out = nullptr;
  if (cfg->nfreqs < 1 || !cfg->freqs || cfg->ndatastreams < 1 || !cfg->datastreams ||
      cfg->nbaselines < 0 || (cfg->nbaselines > 0 && !cfg->baselines)) {
    create_error() = "empty or missing freq, datastream or baseline table";
    return FXG_E_INVALID;
  }
  if (!dry) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
      create_error() = "no HIP device";
      return FXG_E_NODEVICE;
    }
  }
  fxg_ctx* ctx = new fxg_ctx();
  ctx->dry = dry;
  if (!dry) (void)hipGetDevice(&ctx->device);
  ctx->nfreqs = cfg->nfreqs;
  ctx->nds = cfg->ndatastreams;
  ctx->nbl = cfg->nbaselines;
  ctx->freqs.assign(cfg->freqs, cfg->freqs + cfg->nfreqs);
  ctx->dss.assign(cfg->datastreams, cfg->datastreams + cfg->ndatastreams);
  if (cfg->nbaselines > 0) ctx->bls.assign(cfg->baselines, cfg->baselines + cfg->nbaselines);
  ctx->models.resize(ctx->nds);
  for (int d = 0; d < ctx->nds; ++d) {
    const fxg_datastream& s = ctx->dss[d];
    if (s.model_delay_us && s.model_nsamples > 0)
      ctx->models[d].assign(s.model_delay_us, s.model_delay_us + (size_t)s.model_nsamples * FXG_MAX_POLY);
    ctx->dss[d].model_delay_us = ctx->models[d].empty() ? nullptr : ctx->models[d].data();
  }
  ctx->cfg_pcal_interval.assign(ctx->nds, 0);
  ctx->cfg_pcal_base.assign(ctx->nds, 0);
  for (int d = 0; d < ctx->nds; ++d) {
    ctx->cfg_pcal_interval[d] = ctx->dss[d].phasecal_interval_hz;
    ctx->cfg_pcal_base[d] = ctx->dss[d].phasecal_base_hz;
  }
  ctx->xms = cfg->xmac_stride_len;
  ctx->nphase = cfg->nphasecentres > 1 ? cfg->nphasecentres : 1;
  ctx->nthreads = cfg->process_threads > 0 ? cfg->process_threads : 1;
  ctx->maxns_xc = cfg->max_ns_between_xc_avg;
  ctx->pcmodels.resize(ctx->nds);
  for (int d = 0; d < ctx->nds; ++d) {
    const fxg_datastream& s = ctx->dss[d];
    if (ctx->nphase > 1 && cfg->nphasecentres <= FXG_MAX_PHASECENTRES && s.phasecentre_delay_us &&
        s.model_nsamples > 0)
      ctx->pcmodels[d].assign(s.phasecentre_delay_us,
                              s.phasecentre_delay_us + (size_t)ctx->nphase * s.model_nsamples * FXG_MAX_POLY);
    ctx->dss[d].phasecentre_delay_us = ctx->pcmodels[d].empty() ? nullptr : ctx->pcmodels[d].data();
  }
  ctx->subint_ns = cfg->subint_ns;
  ctx->guard_ns = cfg->guard_ns;
  ctx->rotorder = cfg->fringe_rot_order;
  ctx->nbuffered = cfg->num_buffered_ffts;
  ctx->crossautos = cfg->write_cross_autocorrs;
  ctx->model_inc = cfg->model_inc_secs;
  ctx->scan_start = cfg->scan_start_sec;
  ctx->xengine = cfg->xengine;
  ctx->cfg_fmode = cfg->fengine_mode;
  ctx->cfg_fthreads = cfg->fengine_threads;
  ctx->cfg_autotune = cfg->autotune;
  ctx->fsel.mode = cfg->fengine_mode == FXG_F_PERSISTENT ? FXG_F_PERSISTENT : FXG_F_PER_LENGTH;
  ctx->fsel.ntb = cfg->fengine_threads;
  if (!dry) (void)hipDeviceGetAttribute(&ctx->ncu, hipDeviceAttributeMultiprocessorCount, ctx->device);
  int rc = validate_and_derive(ctx);
  if (rc == FXG_OK) {
    build_layout(ctx);
    if (!dry && hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess)
      rc = fail(ctx, FXG_E_HIP, "stream create");
  }
  if (rc == FXG_OK) rc = build_tables(ctx);
  if (rc == FXG_OK) rc = build_engines(ctx, cfg->chunk_ffts);
  if (rc == FXG_OK) rc = build_items(ctx);
  if (rc == FXG_OK) rc = build_uvshift(ctx);
  for (int d = 0; rc == FXG_OK && d < ctx->nds; ++d)    // the configuration's phase-cal extraction
    if (ctx->cfg_pcal_interval[d] > 0) rc = pcal_configure(ctx, d, ctx->cfg_pcal_interval[d], ctx->cfg_pcal_base[d]);
  if (rc == FXG_OK && !dry) {
    const size_t ndd = std::max<size_t>(ctx->ivals.size() * ctx->nbl * ctx->nphase, 1);
    const size_t ndec = std::max<size_t>(ctx->uvdec_idx.size(), 1);
    for (auto& sg : ctx->stage) {
      if (hipHostMalloc((void**)&sg.h_run, sizeof(DevDsRun) * ctx->nds, 0) != hipSuccess ||
          hipHostMalloc((void**)&sg.h_flags, sizeof(int32_t) * ctx->nds * ctx->flaglen + 16, 0) != hipSuccess ||
          hipHostMalloc((void**)&sg.h_uvdd, sizeof(double) * ndd, 0) != hipSuccess ||
          hipHostMalloc((void**)&sg.h_uvdec, sizeof(float) * ndec, 0) != hipSuccess ||
          hipEventCreateWithFlags(&sg.used, hipEventDisableTiming) != hipSuccess) {
        rc = fail(ctx, FXG_E_HIP, "staging allocation");
        break;
      }
    }
  }
  if (rc == FXG_OK && !dry && ctx->cfg_autotune) rc = autotune(ctx);
  if (rc != FXG_OK) {
    create_error() = ctx->err;                   // fxg_last_error(NULL) returns it
    free_ctx(ctx);
    delete ctx;
    return rc;
  }
  *out = ctx;
  return FXG_OK;
}

}  // namespace

extern "C" {

int fxg_create(const fxg_config* cfg, fxg_ctx** out) { return create_impl(cfg, out, false); }

int fxg_create_layout(const fxg_config* cfg, fxg_ctx** out) { return create_impl(cfg, out, true); }

int fxg_footprint(const fxg_config* cfg, int64_t* device_bytes, int64_t* coreresultlength, int64_t* slot_bytes) {
  fxg_ctx* ctx = nullptr;
  int rc = create_impl(cfg, &ctx, true);
  if (rc != FXG_OK) return rc;
  if (device_bytes) *device_bytes = ctx->dev_bytes;
  if (coreresultlength) *coreresultlength = ctx->coreresultlength;
  if (slot_bytes) {
    long long sb = 0;
    for (int d = 0; d < ctx->nds; ++d) sb += ctx->databytes[d] + 4LL * ctx->dss[d].framebytes;
    *slot_bytes = sb;
  }
  delete ctx;
  return FXG_OK;
}

int fxg_input_samples(const fxg_ctx* ctx, int64_t* samples) {
  if (!ctx || !samples) return FXG_E_INVALID;
  long long n = 0;
  for (int d = 0; d < ctx->nds; ++d) n += (long long)ctx->h_ds[d].nrecbands * ctx->B * ctx->h_ds[d].L;
  *samples = n;
  return FXG_OK;
}

int fxg_device_bytes(const fxg_ctx* ctx, int64_t* bytes) {
  if (!ctx || !bytes) return FXG_E_INVALID;
  *bytes = ctx->dev_bytes;
  return FXG_OK;
}

int fxg_destroy(fxg_ctx* ctx) {
  if (!ctx) return FXG_E_INVALID;
  if (!ctx->dry) {
    (void)hipDeviceSynchronize();
    free_ctx(ctx);
  }
  delete ctx;
  return FXG_OK;
}

int fxg_get_layout(const fxg_ctx* ctx, fxg_layout* out) {
  if (!ctx || !out) return FXG_E_INVALID;
  out->coreresultlength = ctx->coreresultlength;
  out->blockspersend = ctx->B;
  out->nbaselinefreqs = ctx->nfreqs * ctx->nbl;
  long long m = 0;
  for (auto v : ctx->databytes) m = std::max(m, v);
  out->max_databytes = m;
  return FXG_OK;
}

int fxg_data_bytes(const fxg_ctx* ctx, int32_t ds, int64_t* bytes) {
  if (!ctx || !bytes || ds < 0 || ds >= ctx->nds) return FXG_E_INVALID;
  *bytes = ctx->databytes[ds];
  return FXG_OK;
}

// Configuration::getFramesPerSecond (configuration.cpp:654-668) as the framed datastreams
// hold it, in an int (vdiffile.cpp:350, mark5bfile.cpp:123), decimation 1
static int frames_per_second(const fxg_ctx* ctx, int ds) {
  const fxg_datastream& d = ctx->dss[ds];
  const double samplerate = 2.0e6 * ctx->freqs[d.recfreq_index[0]].bw_mhz;
  const int payload = d.framebytes - d.headerbytes;
  if (payload < 1) return 0;
  return (int)(samplerate * d.nrecbands * d.nbits / (8.0 * payload));
}

// DataStream::updateConfig's per-segment constants (datastream.cpp:745-763), with the
// bytes-per-sample fraction of Configuration (configuration.cpp:1656-1664, decimation 1).
static int segment_params(const fxg_ctx* ctx, int32_t ds, int32_t numdatasegments, int32_t bufferbytes,
                          fxg_segment* seg) {
  if (ds < 0 || ds >= ctx->nds || numdatasegments < 1 || bufferbytes < numdatasegments) return FXG_E_INVALID;
  const fxg_datastream& d = ctx->dss[ds];
  const fxg_freq& f = ctx->freqs[d.recfreq_index[0]];
  const bool cplx = d.complex_sampling != 0;
  seg->sendbytes = int((ctx->databytes[ds] * (ctx->subint_ns + ctx->guard_ns)) / ctx->subint_ns);
  seg->blockspersend = ctx->B;
  const int samplebits = d.nbits * (cplx ? 2 : 1);
  seg->bytespersamplenum = (d.nrecbands * samplebits) / 8;
  if (seg->bytespersamplenum == 0) {
    seg->bytespersamplenum = 1;
    seg->bytespersampledenom = 8 / (d.nrecbands * samplebits);
  } else {
    seg->bytespersampledenom = 1;
  }
  seg->numchannels = f.nchan;
  seg->complex_sampling = cplx ? 1 : 0;
  seg->sampletimens = 500.0 / f.bw_mhz;
  if (cplx) seg->sampletimens *= 2;
  const double tiny50 = 50 * 0.000000001f;          // 50 * Mode::TINY (a float, mode.cpp:27)
  seg->bytesbetweenintegerns = 0;
  double nsaccumulate = 0.0;
  int guard = 0;
  do {
    nsaccumulate += seg->bytespersampledenom * seg->sampletimens;
    seg->bytesbetweenintegerns += seg->bytespersamplenum;
    if (++guard > 1000000) return FXG_E_INVALID;    // no whole-nanosecond sample boundary
  } while (!(std::fabs(nsaccumulate - int(nsaccumulate + 0.5)) < tiny50));
  seg->nsinc = int((seg->sampletimens * (bufferbytes / numdatasegments) * seg->bytespersampledenom) /
                   seg->bytespersamplenum + 0.5);
  // framed streams count whole frames (vdiffile.cpp:484-491, mark5bfile.cpp:258-266)
  const int fb = d.framebytes, fps = frames_per_second(ctx, ds);
  if (fps < 1) return FXG_E_INVALID;
  seg->nsinc = int(((bufferbytes / numdatasegments) / fb) * (1000000000.0 / double(fps)) + 0.5);
  seg->sendbytes = int((((double)seg->sendbytes) * ((double)ctx->subint_ns)) / (ctx->subint_ns + ctx->guard_ns) + 0.5);
  return FXG_OK;
}

// (validates the configuration as fxg_create does, without a device)
int fxg_buffer_params(const fxg_config* cfg, int32_t ds, int32_t numdatasegments, int32_t bufferbytes,
                      fxg_databuffer* buf) {
  if (!cfg || !buf || numdatasegments < 1 || bufferbytes < numdatasegments) return FXG_E_INVALID;
  fxg_ctx* ctx = nullptr;
  int rc = create_impl(cfg, &ctx, true);
  if (rc != FXG_OK) return rc;
  if (ds < 0 || ds >= ctx->nds) {
    delete ctx;
    return FXG_E_INVALID;
  }
  const fxg_datastream& d = ctx->dss[ds];
  buf->numdatasegments = numdatasegments;
  buf->bufferbytes = bufferbytes;
  buf->readbytes = bufferbytes / numdatasegments;
  buf->intclockseconds = (int)std::floor(d.clock_coeff0_us / 1000000.0 + 0.5);
  buf->framing = d.format == FXG_FMT_MARK5B ? FXG_FRAMING_MARK5B : FXG_FRAMING_VDIF;
  buf->framebytes = d.framebytes;
  buf->payloadbytes = d.framebytes - d.headerbytes;
  buf->framespersecond = frames_per_second(ctx, ds);
  // the fewest frames that span whole nanoseconds (vdifmux.c:140)
  buf->framegranularity = buf->framespersecond > 0
      ? (int)(buf->framespersecond / gcd_ll(buf->framespersecond, 1000000000LL)) : 0;
  delete ctx;
  return buf->framespersecond > 0 ? FXG_OK : FXG_E_INVALID;
}

int fxg_segment_params(const fxg_config* cfg, int32_t ds, int32_t numdatasegments, int32_t bufferbytes,
                       fxg_segment* seg) {
  if (!cfg || !seg) return FXG_E_INVALID;
  fxg_ctx* ctx = nullptr;
  int rc = create_impl(cfg, &ctx, true);
  if (rc != FXG_OK) return rc;
  rc = segment_params(ctx, ds, numdatasegments, bufferbytes, seg);
  delete ctx;
  return rc;
}

// ---- Visibility::writedata scaling and the DIFX record writer -----------------------------

int fxg_normalise_device(fxg_ctx* ctx, float* d_results, double meansubints, const double* tsys, void* stream) {
  if (!ctx || !d_results || !(meansubints > 0)) return FXG_E_INVALID;
  if (ctx->dry) return FXG_E_STATE;
  std::lock_guard<std::mutex> lk(ctx->mu);
  static const float dpct_table[8] = {0.63662f, 0.88f, 0.94f, 0.96f, 0.98f, 0.99f, 0.996f, 0.998f};  // mode.cpp:1571
  const int nds = ctx->nds;
  std::vector<double> t(nds, 0.0);
  std::vector<float> dp(nds, 0.0f);
  bool anyt = false;
  for (int d = 0; d < nds; ++d) {
    if (tsys) t[d] = tsys[d];
    anyt = anyt || t[d] > 0.0;
    const int nb = ctx->dss[d].nbits;
    dp[d] = (nb >= 1 && nb <= 8) ? dpct_table[nb - 1] : 0.0f;
  }
  hipStream_t st = stream ? reinterpret_cast<hipStream_t>(stream) : ctx->stream;
  HIPCHK(hipMemcpyAsync(ctx->d_tsys, t.data(), sizeof(double) * nds, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(ctx->d_dpct, dp.data(), sizeof(float) * nds, hipMemcpyHostToDevice, st));
  NormArgs a;
  a.res = d_results;
  a.seg = ctx->d_normseg;
  a.cal = ctx->d_normcal;
  a.calmean = ctx->d_calmean;
  a.tsys = ctx->d_tsys;
  a.dpct = ctx->d_dpct;
  a.msi = meansubints;
  a.fftsperint = meansubints * ctx->B;
  a.inttime_s = meansubints * (double)ctx->subint_ns / 1e9;
  const int ncal = anyt ? (int)ctx->normcal.size() - 1 : 0;   // (not the final zero entry)
  if (fxg_launch_normalise(a, (int)ctx->normseg.size(), ncal, st))
    return fail(ctx, FXG_E_HIP, "normalisation launch failed");
  HIPCHK(hipStreamSynchronize(st));                // the host arrays above are copied from
  return FXG_OK;
}

static char opposite_pol(char p) {                   // Configuration::getOppositePol
  return p == 'R' ? 'L' : p == 'L' ? 'R' : p == 'X' ? 'Y' : 'X';
}

static char band_pol(const fxg_datastream& s, int b) {
  return (char)(b < s.nrecbands ? s.recband_pol[b] : s.zoomband_pol[b - s.nrecbands]);
}

// Visibility::appendSWINHeaderBuffered (visibility.cpp:1205-1234): 74 bytes, little-endian
static void put_header(uint8_t* o, int32_t bl, int32_t mjd, double sec, int32_t cfgi, int32_t src, int32_t freq,
                       char p0, char p1, int32_t bin, double weight, const double* uvw) {
  const uint32_t sync = 0xFF00FF00u;              // SYNC_WORD (visibility.h:125)
  const int32_t version = 1;                       // BINARY_HEADER_VERSION
  std::memcpy(o + 0, &sync, 4);
  std::memcpy(o + 4, &version, 4);
  std::memcpy(o + 8, &bl, 4);
  std::memcpy(o + 12, &mjd, 4);
  std::memcpy(o + 16, &sec, 8);
  std::memcpy(o + 24, &cfgi, 4);
  std::memcpy(o + 28, &src, 4);
  std::memcpy(o + 32, &freq, 4);
  o[36] = (uint8_t)p0;
  o[37] = (uint8_t)p1;
  std::memcpy(o + 38, &bin, 4);
  std::memcpy(o + 42, &weight, 8);
  std::memcpy(o + 50, uvw, 24);
}

int fxg_difx_records(const fxg_ctx* ctx, const float* results, const fxg_difx_meta* m, uint8_t* out,
                     int64_t capacity, int64_t* nbytes, int32_t* nrecords) {
  if (!ctx || !results || !m || !nbytes || !(m->fftsperintegration > 0)) return FXG_E_INVALID;
  const int nb = ctx->nbl;
  if (nb > 0 && (!m->baselinenumber || !m->uvw)) return FXG_E_INVALID;
  if (ctx->crossautos && !m->telescopeindex) return FXG_E_INVALID;
  const int P = ctx->nphase, pc = ctx->nphase > 1 ? m->phasecentre : 0;
  if (pc < 0 || pc >= P) return FXG_E_INVALID;
  const double inttime_s = m->fftsperintegration / ctx->B * (double)ctx->subint_ns / 1e9;
  int64_t pos = 0;
  int32_t nrec = 0;
  auto emit = [&](int32_t bl, int32_t freq, char p0, char p1, double w, const double* uvw, const float* data, int fc) {
    const int64_t len = 74 + 8LL * fc;
    if (out) {
      if (pos + len > capacity) return false;
      put_header(out + pos, bl, m->mjd, m->seconds, m->configindex, m->sourceindex, freq, p0, p1, 0, w, uvw);
      std::memcpy(out + pos + 74, data, 8 * (size_t)fc);
    }
    pos += len;
    nrec++;
    return true;
  };
  // cross-correlations (visibility.cpp:815-877): baseline, then output freq, then product
  for (int i = 0; i < nb; ++i) {
    const fxg_baseline& b = ctx->bls[i];
    for (int f = 0; f < ctx->nfreqs; ++f) {
      int k = -1;
      for (int q = 0; q < b.nfreqs; ++q) if (b.freq_index[q] == f) { k = q; break; }
      if (k < 0) continue;
      const fxg_freq& fr = ctx->freqs[f];
      const int fc = fr.nchan / fr.chanavg;
      const int core = ctx->boff[(size_t)f * nb + i];
      for (int p = 0; p < b.npolprod[k]; ++p) {
        const float w = (float)(results[ctx->bwoff[(size_t)f * nb + i] * 2 + p] / m->fftsperintegration);
        if (!(w > 0.0f)) continue;
        float cw = w;                                // visibility.cpp:869-873
        if (P > 1)
          cw = w * (float)(results[ctx->sdoff[(size_t)f * nb + i] * 2 + pc] / (1000000000.0 * inttime_s));
        if (!emit(m->baselinenumber[i], f, band_pol(ctx->dss[b.ds1], b.ds1band[k][p]),
                  band_pol(ctx->dss[b.ds2], b.ds2band[k][p]), (double)cw, m->uvw + 3 * i,
                  results + 2LL * (core + (long long)(pc * b.npolprod[k] + p) * fc), fc))
          return FXG_E_INVALID;
      }
    }
  }
  // autocorrelations (visibility.cpp:889-931), in the first phase centre's file
  if (ctx->crossautos && pc == 0) {
    const double zero[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < ctx->nds; ++d) {
      const fxg_datastream& s = ctx->dss[d];
      const int32_t bl = 257 * (m->telescopeindex[d] + 1);
      long long idx = ctx->acoff[d];
      int widx = ctx->acwoff[d] * 2;
      for (int j = 0; j < ctx->bpa; ++j) {
        for (int k = 0; k < s.nrecbands + s.nzoombands; ++k) {
          const int fi = ctx->aclabel[freq_of_band(s, k)];
          if (!(ctx->used[fi] || ctx->equiv[fi])) continue;
          const int fc = ctx->freqs[fi].nchan / ctx->freqs[fi].chanavg;
          const float w = (float)(results[widx++] / m->fftsperintegration);
          if (w > 0.0f) {
            const char p0 = band_pol(s, k);
            if (!emit(bl, fi, p0, j == 0 ? p0 : opposite_pol(p0), (double)w, zero, results + 2 * idx, fc))
              return FXG_E_INVALID;
          }
          idx += fc;
        }
      }
    }
  }
  *nbytes = pos;
  if (nrecords) *nrecords = nrec;
  return FXG_OK;
}

// Core::averageAndSendAutocorrs' STA record data (core.cpp:1166-1258) for a recorded band,
// from the parallel autocorrelation and its weight in a results array
int fxg_sta_autocorr(const fxg_ctx* ctx, const float* results, int32_t ds, int32_t band, int32_t stadumpchannels,
                     double nswidth, float* out, int32_t* nout) {
  if (!ctx || !results || !out || !nout || ds < 0 || ds >= ctx->nds || stadumpchannels < 1) return FXG_E_INVALID;
  const fxg_datastream& s = ctx->dss[ds];
  if (band < 0 || band >= s.nrecbands) return FXG_E_INVALID;
  // datastreamsaveraged: every freq keeps at least stadumpchannels after averaging (core.cpp:1180-1186)
  int minpostav = 1 << 30;
  for (int f = 0; f < ctx->nfreqs; ++f)
    if (ctx->used[f]) minpostav = std::min(minpostav, ctx->freqs[f].nchan / ctx->freqs[f].chanavg);
  const bool averaged = minpostav >= stadumpchannels;
  // the parallel autocorrelation of the band and its weight in the results (recorded bands lead)
  long long idx = ctx->acoff[ds];
  int widx = ctx->acwoff[ds] * 2;
  for (int k = 0; k < band; ++k) {
    const int fi = freq_of_band(s, k);
    if (!(ctx->used[fi] || ctx->equiv[fi])) continue;
    idx += ctx->freqs[fi].nchan / ctx->freqs[fi].chanavg;
    widx++;
  }
  const int fi = freq_of_band(s, band);
  if (!(ctx->used[fi] || ctx->equiv[fi])) return FXG_E_INVALID;
  const fxg_freq& fr = ctx->freqs[fi];
  if (fr.chanavg > 1 && !averaged) return FXG_E_INVALID;   // needs the unaveraged spectrum
  int freqchannels = fr.nchan;
  const double stasamples = 0.001 * nswidth * 2 * fr.bw_mhz;
  const double minimumweight = 0.333 * stasamples / (2 * freqchannels);   // MINIMUM_FILTERBANK_WEIGHT
  const float weight = results[widx];
  if (weight < minimumweight) { *nout = 0; return FXG_OK; }              // no record for this band
  float renormvalue = (float)(1.0 / (2 * freqchannels * weight));
  if (averaged) {
    renormvalue /= fr.chanavg;
    freqchannels /= fr.chanavg;
  }
  const int nchan = std::min(stadumpchannels, freqchannels);
  const int cta = freqchannels / nchan;
  for (int k = 0; k < nchan; ++k) {
    float a = results[2 * (idx + (long long)k * cta)];
    for (int l = 1; l < cta; ++l) a += results[2 * (idx + (long long)k * cta + l)];
    out[k] = a * renormvalue;
  }
  *nout = nchan;
  return FXG_OK;
}

int fxg_result_offsets(const fxg_ctx* ctx, int32_t* baseline, int32_t* bweight, int32_t* autocorr,
                       int32_t* acweight) {
  if (!ctx) return FXG_E_INVALID;
  if (baseline) std::memcpy(baseline, ctx->boff.data(), sizeof(int32_t) * ctx->boff.size());
  if (bweight) std::memcpy(bweight, ctx->bwoff.data(), sizeof(int32_t) * ctx->bwoff.size());
  if (autocorr) std::memcpy(autocorr, ctx->acoff.data(), sizeof(int32_t) * ctx->nds);
  if (acweight) std::memcpy(acweight, ctx->acwoff.data(), sizeof(int32_t) * ctx->nds);
  return FXG_OK;
}

int fxg_process_subint_device(fxg_ctx* ctx, const int32_t offsets[3], const uint8_t* const* d_data,
                              const int32_t* nbytes, const int32_t* const* control, float* d_results,
                              void* stream) {
  if (ctx && ctx->dry) return FXG_E_STATE;          // host-only context
  if (!ctx || !offsets || !d_data || !nbytes || !control || !d_results) return FXG_E_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mu);
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  return run_subint(ctx, offsets, d_data, nbytes, control, (float2*)d_results, st);
}

int fxg_ring_length(const fxg_ctx* ctx) { return ctx ? RING : 0; }

int fxg_submit_subint(fxg_ctx* ctx, const int32_t offsets[3], const uint8_t* const* data,
                      const int32_t* nbytes, const int32_t* const* control, fxg_handle* handle) {
  if (ctx && ctx->dry) return FXG_E_STATE;          // host-only context
  if (!ctx || !offsets || !data || !nbytes || !control || !handle) return FXG_E_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mu);
  int rc = ensure_slots(ctx);
  if (rc) return rc;
  int k = -1;
  for (int i = 0; i < RING; ++i)
    if (!ctx->slots[i].busy) { k = i; break; }
  if (k < 0) return fail(ctx, FXG_E_STATE, "all ring slots are in use");
  Slot& s = ctx->slots[k];
  std::vector<const uint8_t*> dd(ctx->nds);
  for (int d = 0; d < ctx->nds; ++d) {
    if (nbytes[d] > s.cap[d]) return fail(ctx, FXG_E_INVALID, "data buffer larger than the slot capacity");
    if (nbytes[d] > 0) HIPCHK(hipMemcpyAsync(s.d_data[d], data[d], nbytes[d], hipMemcpyHostToDevice, s.copy));
    dd[d] = s.d_data[d];
  }
  HIPCHK(hipEventRecord(s.copied, s.copy));
  HIPCHK(hipStreamWaitEvent(ctx->stream, s.copied, 0));
  rc = run_subint(ctx, offsets, dd.data(), nbytes, control, s.d_results, ctx->stream);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(s.h_results, s.d_results, sizeof(float2) * ctx->coreresultlength,
                        hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipEventRecord(s.done, ctx->stream));
  s.busy = true;
  *handle = k;
  return FXG_OK;
}

int fxg_wait(fxg_ctx* ctx, fxg_handle handle, float* results) {
  if (ctx && ctx->dry) return FXG_E_STATE;          // host-only context
  if (!ctx || handle < 0 || handle >= RING) return FXG_E_INVALID;
  Slot& s = ctx->slots[handle];
  if (!s.busy) return fail(ctx, FXG_E_STATE, "wait on a slot that was not submitted");
  HIPCHK(hipEventSynchronize(s.done));
  if (results) std::memcpy(results, s.h_results, sizeof(float2) * ctx->coreresultlength);
  std::lock_guard<std::mutex> lk(ctx->mu);
  s.busy = false;
  return FXG_OK;
}

int fxg_process_subint(fxg_ctx* ctx, const int32_t offsets[3], const uint8_t* const* data,
                       const int32_t* nbytes, const int32_t* const* control, float* results) {
  if (ctx && ctx->dry) return FXG_E_STATE;          // host-only context
  fxg_handle h;
  int rc = fxg_submit_subint(ctx, offsets, data, nbytes, control, &h);
  if (rc) return rc;
  return fxg_wait(ctx, h, results);
}

int fxg_get_data_weights(const fxg_ctx* ctx, int32_t ds, float* out) {
  if (ctx && ctx->dry) return FXG_E_STATE;          // host-only context
  if (!ctx || !out || ds < 0 || ds >= ctx->nds) return FXG_E_INVALID;
  fxg_ctx* c = const_cast<fxg_ctx*>(ctx);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return FXG_E_HIP;
  if (hipMemcpy(out, ctx->d_weights + (size_t)ds * ctx->B, sizeof(float) * ctx->B, hipMemcpyDeviceToHost) != hipSuccess)
    return FXG_E_HIP;
  return FXG_OK;
}

int fxg_set_pcal(fxg_ctx* ctx, int32_t ds, int64_t interval_hz, int64_t base_hz) {
  if (!ctx || ds < 0 || ds >= ctx->nds || interval_hz < 0) return FXG_E_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (ctx->cfg_pcal_interval[ds] > 0 && (interval_hz != ctx->cfg_pcal_interval[ds] || base_hz != ctx->cfg_pcal_base[ds]))
    return fail(ctx, FXG_E_INVALID, "the configuration's phase-cal block fixes this datastream's phase-cal setting");
  return pcal_configure(ctx, ds, interval_hz, base_hz);
}

int fxg_pcal_info(const fxg_ctx* ctx, int32_t ds, int32_t band, int32_t* ntones, int64_t* offset_hz,
                  int32_t* period) {
  if (!ctx || !ntones || !offset_hz || !period || ds < 0 || ds >= ctx->nds) return FXG_E_INVALID;
  for (size_t i = 0; i < ctx->pcal_tones.size(); ++i) {
    const fxg_ctx::PcalTones& t = ctx->pcal_tones[i];
    if (t.ds != ds || t.band != band) continue;
    *ntones = t.ntones;
    *offset_hz = t.offset_hz;
    *period = ctx->pcal_entry[i] < 0 ? 0 : ctx->pcal_tab[ctx->pcal_entry[i]].M;
    return FXG_OK;
  }
  return FXG_E_STATE;                              // phase-cal not on for this band
}

// The extractors' getFinalPCal tones for the last subint, as k_pcal_tones formed them; a flagged
// offset (-1 for no tone in the band, or negative through base_hz) leaves them zero, as the
// reference's dummy extractor does.
int fxg_get_pcal(const fxg_ctx* ctx, int32_t ds, int32_t band, float* out, int32_t* ntones) {
  if (ctx && ctx->dry) return FXG_E_STATE;          // host-only context
  if (!ctx || !out || !ntones || ds < 0 || ds >= ctx->nds) return FXG_E_INVALID;
  int t = -1;
  for (size_t i = 0; i < ctx->pcal_tones.size(); ++i)
    if (ctx->pcal_tones[i].ds == ds && ctx->pcal_tones[i].band == band) t = (int)i;
  if (t < 0) return FXG_E_STATE;                   // phase-cal not on for this band
  const fxg_ctx::PcalTones& pt = ctx->pcal_tones[t];
  if (hipDeviceSynchronize() != hipSuccess) return FXG_E_HIP;
  if (hipMemcpy(out, ctx->d_pcal_tonebuf + pt.tone0, sizeof(float2) * pt.ntones, hipMemcpyDeviceToHost) != hipSuccess)
    return FXG_E_HIP;
  *ntones = pt.ntones;
  return FXG_OK;
}

int fxg_result_pcal_offsets(const fxg_ctx* ctx, int32_t* pcal) {
  if (!ctx || !pcal) return FXG_E_INVALID;
  for (int d = 0; d < ctx->nds; ++d) pcal[d] = ctx->pcaloff[d];
  return FXG_OK;
}

int fxg_set_kurtosis(fxg_ctx* ctx, int32_t enable) {
  if (ctx && ctx->dry) return FXG_E_STATE;          // host-only context
  if (!ctx) return FXG_E_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (enable && !ctx->d_kurt_tab) {
    ctx->kurt_tab.clear();
    ctx->kurt_index.assign(ctx->nds, 0);
    long long acc = 0;
    for (int d = 0; d < ctx->nds; ++d) {
      const DevDs& x = ctx->h_ds[d];
      ctx->kurt_index[d] = (int)ctx->kurt_tab.size();
      for (int b = 0; b < x.nrecbands; ++b) {
        KurtBand k{};
        k.C = x.C;
        if (ctx->xengine == FXG_X_FP32) {
          k.off = x.spec_off + (long long)b * ctx->chunkcap * x.C;
          k.scale_idx = 0;
        } else {
          const DevXGroup& g = ctx->groups[x.band_xfreq[b]];
          k.off = g.spec_off + (long long)x.band_xrow[b] * g.rowstride;
          k.scale_idx = g.scale_off + x.band_xrow[b];
          k.twlog = g.twlog;
        }
        k.acc_off = (int)acc;
        acc += x.C;
        ctx->kurt_maxC = std::max(ctx->kurt_maxC, x.C);
        ctx->kurt_tab.push_back(k);
      }
    }
    ctx->kurt_len = std::max<long long>(acc, 1);
    if (int rc = dev_upload(ctx, &ctx->d_kurt_tab, ctx->kurt_tab)) return rc;
    if (int rc = dev_malloc(ctx, (void**)&ctx->d_kurt_acc, sizeof(float2) * ctx->kurt_len)) return rc;
    if (int rc = dev_zero(ctx, ctx->d_kurt_acc, sizeof(float2) * ctx->kurt_len)) return rc;
  }
  ctx->kurt_on = enable != 0;
  return FXG_OK;
}

// Mode::calculateAndAverageKurtosis (mode.cpp:1405-1450) on the last subint's sums, with the
// subint's FFT blocks as numblocks, in float as the reference computes it
int fxg_get_kurtosis(const fxg_ctx* ctx, int32_t ds, int32_t band, int32_t maxchannels, float* out,
                     int32_t* nout) {
  if (ctx && ctx->dry) return FXG_E_STATE;          // host-only context
  if (!ctx || !out || !nout || ds < 0 || ds >= ctx->nds || maxchannels < 1) return FXG_E_INVALID;
  if (!ctx->kurt_on || !ctx->d_kurt_acc) return FXG_E_STATE;
  if (band < 0 || band >= ctx->h_ds[ds].nrecbands) return FXG_E_INVALID;
  const KurtBand& k = ctx->kurt_tab[ctx->kurt_index[ds] + band];
  const int C = k.C;
  std::vector<float2> v((size_t)C);
  if (hipDeviceSynchronize() != hipSuccess) return FXG_E_HIP;
  if (hipMemcpy(v.data(), ctx->d_kurt_acc + k.acc_off, sizeof(float2) * C, hipMemcpyDeviceToHost) != hipSuccess)
    return FXG_E_HIP;
  bool nonzero = false;
  for (int j = 0; j < C; ++j)
    if (v[j].x > 0.0f) { nonzero = true; break; }
  if (!nonzero) { *nout = 0; return FXG_OK; }     // the reference sends no record
  const int numblocks = ctx->B;
  std::vector<float> sk((size_t)C);
  const float nb = (float)numblocks;
  const float nbfac = (float)((float)numblocks / (numblocks - 1.0));
  for (int j = 0; j < C; ++j) {
    const float s1sq = v[j].x * v[j].x;
    const float s2 = v[j].y * nb;
    float x = s2 / s1sq;
    x = x + -1.0f;
    x = x * nbfac;
    x = x + -1.0f;
    sk[j] = x;
  }
  if (maxchannels < C) {
    const int kchanavg = C / maxchannels;
    const float inv = (float)(1.0 / ((float)kchanavg));
    for (int j = 0; j < C; ++j) sk[j] *= inv;
    for (int j = 0; j < maxchannels; ++j) {
      float a = sk[(size_t)j * kchanavg];
      for (int q = 1; q < kchanavg; ++q) a += sk[(size_t)j * kchanavg + q];
      out[j] = a;
    }
    *nout = maxchannels;
  } else {
    for (int j = 0; j < C; ++j) out[j] = sk[j];
    *nout = C;
  }
  return FXG_OK;
}

int fxg_get_spectrum(const fxg_ctx* ctx, int32_t ds, int32_t band, int32_t fft, float* out) {
  if (ctx && ctx->dry) return FXG_E_STATE;          // host-only context
  if (!ctx || !out || ds < 0 || ds >= ctx->nds) return FXG_E_INVALID;
  const DevDs& x = ctx->h_ds[ds];
  if (band < 0 || band >= x.nrecbands) return FXG_E_INVALID;
  const int lt = fft - ctx->last_chunk0;
  if (lt < 0 || lt >= ctx->last_nffts) return FXG_E_STATE;
  fxg_ctx* c = const_cast<fxg_ctx*>(ctx);
  if (hipDeviceSynchronize() != hipSuccess) return FXG_E_HIP;
  const int C = x.C;
  if (ctx->xengine == FXG_X_FP32) {
    const float2* src = ctx->d_spec32 + x.spec_off + ((long long)band * ctx->chunkcap + lt) * C;
    if (hipMemcpy(out, src, sizeof(float2) * C, hipMemcpyDeviceToHost) != hipSuccess) return FXG_E_HIP;
    return FXG_OK;
  }
  const DevXGroup& g = ctx->groups[x.band_xfreq[band]];
  const int row = x.band_xrow[band];
  std::vector<uint32_t> blk((size_t)C);
  float sc = 1.0f;
  const int tw = 1 << g.twlog;
  for (int ct = 0; ct < C / tw; ++ct) {
    const uint32_t* src = ctx->d_spec16 + g.spec_off + (long long)row * g.rowstride +
                          ((long long)ct * ctx->chunkcap + lt) * tw;
    if (hipMemcpy(blk.data() + tw * ct, src, sizeof(uint32_t) * tw, hipMemcpyDeviceToHost) != hipSuccess)
      return FXG_E_HIP;
  }
  if (hipMemcpy(&sc, ctx->d_xscale + g.scale_off + row, sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    return FXG_E_HIP;
  for (int ch = 0; ch < C; ++ch) {
    uint32_t u = blk[(size_t)ch];
    float re, im;
    if (ctx->xengine == FXG_X_FP16) {
      __half2 h;
      std::memcpy(&h, &u, 4);
      float2 f = __half22float2(h);
      re = f.x; im = f.y;
    } else {
      uint32_t lo = (u & 0xFFFFu) << 16, hi = u & 0xFFFF0000u;
      std::memcpy(&re, &lo, 4);
      std::memcpy(&im, &hi, 4);
    }
    out[2 * ch] = re / sc;
    out[2 * ch + 1] = im / sc;
  }
  (void)c;
  return FXG_OK;
}

int fxg_unpack_device(fxg_ctx* ctx, int32_t ds, const uint8_t* d_data, int32_t nbytes, int64_t startsample,
                      int32_t nsamples, float* d_out, int32_t* d_goodsamples, void* stream) {
  if (ctx && ctx->dry) return FXG_E_STATE;          // host-only context
  if (!ctx || ds < 0 || ds >= ctx->nds || !d_data || !d_out || !d_goodsamples || nsamples <= 0 || startsample < 0)
    return FXG_E_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mu);
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  const int nframes = nbytes / ctx->dss[ds].framebytes;
  const long long capframes = (ctx->databytes[ds] + 4LL * ctx->dss[ds].framebytes) / ctx->dss[ds].framebytes;
  if (nframes > capframes) return fail(ctx, FXG_E_INVALID, "buffer larger than the slot capacity");
  // frame validity through the setup kernel on a one-datastream run record
  DevDsRun r;
  std::memset(&r, 0, sizeof(r));
  r.data = d_data;
  r.datalen = nbytes;
  r.nframes = nframes;
  r.valid = 1;
  r.scan = 0;
  r.datasec = 0;
  DevDsRun none;
  std::memset(&none, 0, sizeof(none));          // other datastreams: no frames to scan
  std::vector<DevDsRun> runs(ctx->nds, none);
  runs[ds] = r;
  HIPCHK(hipMemcpyAsync(ctx->d_run, runs.data(), sizeof(DevDsRun) * ctx->nds, hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  SetupArgs sa;
  sa.ds = ctx->d_ds;
  sa.run = ctx->d_run;
  sa.model = ctx->d_model;
  sa.fvalid = ctx->d_fvalid;
  sa.fzone = ctx->d_fzone;
  sa.nds = ctx->nds;
  sa.offscan = 0; sa.offsec = 0; sa.offns = 0;
  sa.model_inc_secs = ctx->model_inc;
  sa.B = ctx->B;
  if (fxg_launch_setup(sa, std::max(nframes, 1), st)) return fail(ctx, FXG_E_HIP, "setup launch");
  HIPCHK(hipMemsetAsync(d_goodsamples, 0, sizeof(int32_t), st));
  UnpackArgs ua;
  ua.ds = ctx->d_ds;
  ua.d = ds;
  ua.data = d_data;
  ua.fv = ctx->d_fvalid + ctx->h_ds[ds].fvalid_off;
  ua.fz = ctx->d_fzone + ctx->h_ds[ds].fvalid_off;
  ua.nframes = nframes;
  ua.start = startsample;
  ua.nsamples = nsamples;
  ua.out = d_out;
  ua.good = d_goodsamples;
  if (fxg_launch_unpack(ua, st)) return fail(ctx, FXG_E_HIP, "unpack launch");
  HIPCHK(hipStreamSynchronize(st));
  return FXG_OK;
}

int fxg_delay_interpolator(fxg_ctx* ctx, int32_t ds, int32_t sec, int32_t ns, double* out) {
  if (ctx && ctx->dry) return FXG_E_STATE;          // host-only context
  if (!ctx || !out || ds < 0 || ds >= ctx->nds) return FXG_E_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mu);
  hipStream_t st = ctx->stream;
  DevDsRun r;
  std::memset(&r, 0, sizeof(r));
  r.valid = 1;
  r.nframes = 0;
  std::vector<DevDsRun> runs(ctx->nds, r);
  HIPCHK(hipMemcpyAsync(ctx->d_run, runs.data(), sizeof(DevDsRun) * ctx->nds, hipMemcpyHostToDevice, st));
  SetupArgs sa;
  sa.ds = ctx->d_ds;
  sa.run = ctx->d_run;
  sa.model = ctx->d_model;
  sa.fvalid = ctx->d_fvalid;
  sa.fzone = ctx->d_fzone;
  sa.nds = ctx->nds;
  sa.offscan = 0; sa.offsec = sec; sa.offns = ns;
  sa.model_inc_secs = ctx->model_inc;
  sa.B = ctx->B;
  if (fxg_launch_setup(sa, 1, st)) return fail(ctx, FXG_E_HIP, "setup launch");
  DevDsRun back;
  HIPCHK(hipMemcpyAsync(&back, ctx->d_run + ds, sizeof(DevDsRun), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (!back.valid) return fail(ctx, FXG_E_INVALID, "time outside the delay model");
  std::memcpy(out, back.interp, sizeof(double) * 3);
  return FXG_OK;
}

int fxg_autotune_report(const fxg_ctx* ctx, char* buf, int32_t len) {
  if (!ctx || !buf || len < 1) return FXG_E_INVALID;
  std::string r = "F-engine: ";
  r += ctx->fsel.mode == FXG_F_PERSISTENT ? "persistent" : "per-length";
  r += ", threads " + (ctx->fsel.ntb ? std::to_string(ctx->fsel.ntb) : std::string("by length"));
  if (ctx->fsel.mode == FXG_F_PERSISTENT) r += ", " + std::to_string(ctx->fsel.grid_mult) + " workgroup(s)/CU";
  for (const auto& e : ctx->tune_log) {
    char line[160];
    std::snprintf(line, sizeof(line), "\n  %s: %.3f ms per subint", e.first.c_str(), e.second);
    r += line;
  }
  std::strncpy(buf, r.c_str(), (size_t)len - 1);
  buf[len - 1] = 0;
  return FXG_OK;
}

int fxg_set_timing(fxg_ctx* ctx, int32_t enable) {
  if (ctx && ctx->dry) return FXG_E_STATE;          // host-only context
  if (!ctx) return FXG_E_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIPCHK(hipDeviceSynchronize());
  for (auto& t : ctx->tev) {
    ctx->evpool.push_back(std::get<1>(t));
    ctx->evpool.push_back(std::get<2>(t));
  }
  ctx->tev.clear();
  ctx->timing = enable != 0;
  return FXG_OK;
}

int fxg_get_timing(fxg_ctx* ctx, fxg_timing* out) {
  if (ctx && ctx->dry) return FXG_E_STATE;          // host-only context
  if (!ctx || !out) return FXG_E_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mu);
  std::memset(out, 0, sizeof(*out));
  for (auto& t : ctx->tev) {
    HIPCHK(hipEventSynchronize(std::get<2>(t)));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, std::get<1>(t), std::get<2>(t)));
    switch (std::get<0>(t)) {
      case 0: out->f_ms += ms; out->f_launches++; break;
      case 1: out->x_ms += ms; out->x_launches++; break;
      default: out->other_ms += ms; out->other_launches++; break;
    }
  }
  return FXG_OK;
}

}  // extern "C"
