// Internal tables shared by the host layer and the device kernels of libfxg.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/fxg.h"

#define FXG_TB 4            // consecutive FFTs per output time block (MFMA K packing)
#define FXG_XPAD 8          // FFT count padding of a device chunk (two time blocks)
#define FXG_MAX_ZOUT 16     // zoom-band rows written by the F-engine per datastream
#define FXG_MAX_L 8192      // longest FFT of the fused (single-workgroup LDS) F-engine
#define FXG_MAX_BIG_L 32768 // longest FFT of the four-step F-engine (fxg_fbig.hip)

// Per recorded freq of a datastream: what the rotator and the fractional-sample
// correction need (mode.cpp:853-931, 1030-1073).
struct DevRecFreq {
  double lofreq;          // MHz (band edge; negated for complex LSB)
  double lo_offset_hz;    // recordedfreqlooffsets
  double fraclooffset;    // signed fractional part of the LO offset
  double clockoff;        // recordedfreqclockoffsets (us)
  double halfdelta;       // recordedfreqclockoffsetsdelta / 2 (us)
  int    lsb;
  int    looff;           // LO offset non-zero
  int    isfraclooffset;
  int    nbands;          // recorded bands of this freq
  int    band[4];         // their band indices (sample-word order)
  int    pad;
  // per band of the freq: the clock-delta term of its fractional-sample phasor, +halfdelta
  // for R (fracsamprotatorA) and -halfdelta for the other pols (fracsamprotatorB,
  // mode.cpp:1076-1118, 1257-1261); the pol phase-offset phasor of the second pol
  // (phasecorrA, mode.cpp:421-447, 1318-1330), (1, 0) for the others; and whether band k's
  // phasor table differs from band k-1's
  double bhd[4];
  float  pcre[4], pcim[4];
  int    newfrac[4];
};

// Per datastream (everything the F-engine and the unpacker read).
struct DevDs {
  int format;             // FXG_FMT_*
  int nbits;
  int iscomplex;
  int dsb;                // complex double sideband (mode.cpp:41-49)
  int framebytes;
  int headerbytes;
  int spf;                // time samples per frame
  int nrecbands;
  int nrecfreqs;
  int L;                  // fftchannels
  int C;                  // recordedbandchannels
  int S;                  // arraystridelength
  int NS;                 // L / S  (rotator steps)
  int NFS;                // C / S  (fractional-sample steps)
  int bpb_num, bpb_den;   // bytes-per-block ratio of the bounds check (mode.cpp:124-139)
  int fractionallo;
  int intclockseconds;
  int model_order;
  int model_nsamples;
  int model_off;          // offset of this datastream's polynomials in the model table
  long long spec_off;     // FP32 spectra: float2 offset of (band 0, t 0) in the chunk buffer
  int chanfreq_off;       // offset of the channel-frequency tables in the chanfreq array
  int fvalid_off;         // offset of this datastream's frame-valid flags
  double sampletime_us;
  double bw_mhz;
  float xsc;              // 16-bit spectra: this datastream's input scale (1 for FP32 spectra)
  DevRecFreq rf[32];
  int band_xfreq[64];     // MFMA: X freq group of each recorded band
  int band_xrow[64];      // MFMA: input row of each recorded band within its group
  // MFMA: zoom bands correlated on baselines are extra rows of their zoom freq's group,
  // filled with the channel window [c0, c0 + n) of the parent band (mode.cpp:184-196)
  int nzout;
  int zout_band[FXG_MAX_ZOUT];   // parent recorded band
  int zout_group[FXG_MAX_ZOUT];
  int zout_row[FXG_MAX_ZOUT];
  int zout_c0[FXG_MAX_ZOUT];
  int zout_n[FXG_MAX_ZOUT];
};

// Per subint, per datastream runtime state written by the setup kernel.
struct DevDsRun {
  double interp[3];       // quadratic delay interpolator (FFT-index units)
  int valid;              // 0 = the whole subint is flagged for this datastream
  int scan, datasec, datans;
  int datalen;            // bytes
  int nframes;            // whole frames in the buffer
  const uint8_t* data;    // device pointer of the slot buffer
  const int32_t* vflags;  // device copy of the valid flags
};

// Plain FP32 X-engine product (core.cpp:963 and the autocorrelations of mode.cpp:1272,1334).
struct DevProd {
  int ds1, band1, choff1; // first operand: spectra band and channel offset (zoom windows)
  int ds2, band2, choff2; // second operand (conjugated)
  int nchan;              // channels of the product
  int acc_off;            // cf32 offset of its accumulator
};

// MFMA X-engine group: one recorded frequency with its inputs (station bands).
struct DevXGroup {
  int nin;                // inputs (rows)
  int ntile;              // 32-row tiles
  int C;                  // channels
  long long acc_off;      // float offset of its accumulator block [C][npairs][2][32][32]
  long long spec_off;     // element offset (32-bit units) of its spectra block
  long long rowstride;    // 32-bit units per input row: C x chunk FFTs plus one 128-byte line
  long long scale_off;    // float offset of its per-row scales [rows]
  int twlog;              // log2 of the channel tile width TW (8 or 16): spectra rows are
                          // [C/TW][t][TW] words, so one FFT writes TW channels contiguously
  int pad_;
};

// Frame flags of a slot buffer: 0 blanked, 1 every sample decodes, 2 only the time samples
// [zone.x, zone.y) decode (a Mark5B payload with the fill pattern at one end; the byte range
// of blanker_mark5.c:112-172 as samples).
#define FXG_FRAME_PARTIAL 2
__device__ __forceinline__ bool sample_ok(unsigned flag, const int2* __restrict__ zone, unsigned fr,
                                          unsigned within) {
  if (flag != FXG_FRAME_PARTIAL) return flag != 0u;
  const int2 z = zone[fr];
  return (int)within >= z.x && (int)within < z.y;
}

// Per-FFT delay chain of one datastream (mode.cpp:671-675, 743-757, 781-788).
struct FftDelay {
  int ok;               // 1 = processed, 0 = flagged (zero spectra, weight 0)
  int n0;               // nearestsample
  float fse;            // fracsampleerror (us)
  int intdelay;         // integerdelay (us)
  double a, b;          // interpolator across the FFT (FFT-fraction units); order 0: b is
                        // averagedelay - integerdelay
  double c;             // order 2: delay at the FFT start minus integerdelay
  float weight;
  int intwall;
  double fracwall;
};

// Kernel launch helpers implemented in the .hip files.
struct FengineArgs {
  const DevDs* ds;
  const DevDsRun* run;
  const float2* twiddle;  // per-pass FFT twiddle tables (see Plan<L> in fxg_fengine.hip)
  FftDelay* params;       // [nds][chunkcap] per-FFT delay chain of the current chunk
  float* weights;         // [nds][B]
  int B;                  // blockspersend
  int chunk0;             // first FFT index of this chunk
  int nffts;              // FFTs in this chunk
  int nffts_pad;          // padded to a multiple of FXG_XPAD
  int offsec, offns;      // subint offsets
  double scan_start_sec;
  int out_mode;           // 0 = FP32 [ds][band][t][C], 1 = FP16, 2 = BF16 MFMA layout
  float2* spec32;         // FP32 spectra
  uint32_t* spec16;       // half2/bf16x2 spectra, layout [group][row (rowstride)][C/TW][t][TW]
  const DevXGroup* xgroups;
  const float* xscale;    // per group [rows] scales applied before 16-bit conversion
  const float* chanfreq;  // per datastream: sub[S], then usb, lsb and dsb steps [NFS] (mode.cpp:303-324)
  const uint8_t* fvalid;  // frame flags
  const int2* fzone;      // per frame: the decoding samples of a partial frame
  int chunkcap;           // FFT capacity of the chunk buffers (multiple of FXG_XPAD)
  int units_per_wg;       // per-length launch: units of 8 FFTs per workgroup
  const int* dslist;      // datastreams of this launch, indexed by grid.y (one FFT-length class)
  int rotorder;           // fringe rotation order 0, 1 or 2
  int phcap;              // LDS entries of each phasor table (max S + NS, S + NFS of the launch)
  const float2* tw_by_log[8];  // persistent launch: twiddle tables of L = 256 << i (for
                               // L > 8192 the L-point exp(-2 pi i j / L) table)
  const float2* tw2048;        // persistent launch: 2048-point sub-FFT tables of the four-step F-engine
  float2* bigscratch;          // persistent launch: four-step scratch rows, NTB/256 per workgroup
  int bigL;                    // persistent launch: row length of bigscratch (largest four-step L)
};

// Work item of the persistent F-engine launch: part `part` of `nparts` of the chunk's FFT
// units of datastream d, FFT length L.
struct FItem {
  int L, log2L, d, part, nparts;
};

// Persistent launch: the FFT lengths, longest first, with their ranges of work items.
#define FXG_MAX_ICLASS 8
struct IClass {
  int L, log2L, begin, count;
};

// Setup for one subint (delay interpolators, frame validity).
struct SetupArgs {
  const DevDs* ds;
  DevDsRun* run;
  const double* model;        // [..][FXG_MAX_POLY]
  uint8_t* fvalid;
  int2* fzone;
  int nds;
  int offscan, offsec, offns;
  double model_inc_secs;
  int B;
};

// Stand-alone unpack of one datastream's buffer.
struct UnpackArgs {
  const DevDs* ds;
  int d;
  const uint8_t* data;
  const uint8_t* fv;
  const int2* fz;
  int nframes;
  long long start;
  int nsamples;
  float* out;
  int* good;
};

struct XPlainArgs {
  const DevProd* prods;
  const DevDs* ds;
  const float2* spec32;
  float2* acc;
  int nffts;
  int chunkcap;
};

// Blocked plain FP32 X-engine.  A group is a set of rows (spectra bands, or zoom windows of
// them, with one channel count) that products connect; a work item is one 8 x 8 block (row
// block rb <= column block cb) of the group's row x row matrix over 64 consecutive channels.
struct XPRow {
  long long off;          // cf32 offset in the FP32 spectra of (band, FFT 0, window start)
  int stride;             // cf32 entries between consecutive FFTs
  int pad;
};
struct XPGroup {
  int nrows, C, row0, pair0;   // pairs[pair0 + i*nrows + j]: accumulator of product (i, j) or -1
  int stride;                  // cf32 entries between consecutive FFTs, the same for every row
  int pad;
};
struct XPItem {
  int group, rb, cb, c0;
};
struct XPBlkArgs {
  const XPItem* items;
  const XPGroup* groups;
  const XPRow* rows;
  const int* pairs;
  const float2* spec32;
  float2* acc;
  int nffts;
};

struct XMfmaArgs {
  const DevXGroup* groups;
  const uint32_t* spec16;
  float* acc;
  int nffts_pad;          // multiple of 8
  int chunkcap;
  int mode;               // 1 = FP16, 2 = BF16
};

// Multi-phase-centre uv shift (core.cpp:1699-1808): the channel layout of a freq's rotator,
// sub phasors over rs channels and step phasors over strides of rs channels, with channel
// sky frequencies edge + c chbw (USB) or edge - (nchan - 1 - c) chbw (LSB).
struct UvFreq {
  double lofreq;          // band-edge sky frequency, MHz
  double chbw;            // channel width, MHz
  double stepbw;          // rs x chbw
  int lsb;
  int rs;                 // rotate stride length
  int nsteps;             // step phasors: (nchan / xmacstridelen) x (xmacstridelen / rs)
  int pad;
};

// One uv-shift interval: the applied delays of its baselines' phase centres.
struct UvArgs {
  const double* dd;       // [baseline][phase centre] differential delay (us); P <= 1: unused
  const UvFreq* fq;
  int P;
};

// Accumulation of a product accumulator into the results (core.cpp:1819-1853).
struct FinPlain {
  long long src;          // cf32 offset in the product accumulators
  int dst;                // cf32 offset in results (phase centre 0)
  int nout;               // output channels
  int avg;                // channels averaged per output channel
  int bl;                 // baseline of a cross product (-1: autocorrelation, not shifted)
  int fq;                 // freq of the product (uv-shift rotation constants)
  int pcstride;           // cf32 offset between the phase-centre blocks
};

struct FinMfma {
  int group;
  int row1, row2;         // V[row1][row2] (row1 is the unconjugated operand)
  int chan0;              // first channel of the group spectrum (zoom window)
  int dst;
  int nout;
  int avg;
  int isauto;             // parallel-hand autocorrelation: imaginary part forced to zero
  int bl, fq, pcstride;   // as FinPlain
};

struct FinWeight {
  int ds1, ds2;           // ds2 < 0: plain sum of ds1's weights
  int dstf;               // float index in results
};

int fxg_launch_setup(const SetupArgs& a, int maxframes, hipStream_t st);
int fxg_launch_unpack(const UnpackArgs& a, hipStream_t st);
int fxg_launch_fengine(const FengineArgs& a, int L, int iscomplex, int nds, int ntb, hipStream_t st);
int fxg_launch_fftparams(const FengineArgs& a, int nds, hipStream_t st);
int fxg_launch_xplain_blk(const XPBlkArgs& a, int nitems, hipStream_t st);
int fxg_launch_fpersist(const FengineArgs& a, const FItem* items, const IClass* cls, int ncls, int* counters,
                        int nwg, int ntb, int iscomplex, size_t lds, hipStream_t st);
size_t fxg_fengine_lds(int L, int ntb, int phcap, int order);
int fxg_launch_fbig(const FengineArgs& a, int L, int iscomplex, int nds, int wgx, float2* scratch,
                    const float2* twL, hipStream_t st);
size_t fxg_fbig_lds(int ntb, int phcap);
int fxg_launch_xplain(const XPlainArgs& a, int nprod, int maxchan, hipStream_t st);
int fxg_launch_xmfma(const XMfmaArgs& a, int ngroup_launch, int group0, int ntile, int C, int twlog, hipStream_t st);
int fxg_launch_fin_plain(const FinPlain* e, int n, int maxout, const float2* acc, float2* res, const UvArgs& uv,
                         hipStream_t st);
int fxg_launch_fin_mfma(const FinMfma* e, int n, int maxout, const DevXGroup* g, const float* acc,
                        const float* xscale, float2* res, const UvArgs& uv, hipStream_t st);
int fxg_launch_add_at(const int* idx, const float* val, int n, float* resf, hipStream_t st);
int fxg_launch_fin_weights(const FinWeight* e, int n, const float* w, int B, float* resf, hipStream_t st);
int fxg_launch_zero(void* p, long long bytes, hipStream_t st);
int fxg_launch_fill(float* p, long long n, float v, hipStream_t st);
int fxg_launch_add(float* dst, const float* src, long long n, hipStream_t st);

// Visibility::writedata's amplitude scaling (visibility.cpp:353-646), one segment per
// (baseline, freq, product) of cross-correlations and per (datastream, product, band) of
// autocorrelations.  Scales follow the reference's float / double order; zoom factors are
// precomputed on the host (cross: sqrt of the bandwidth ratio, autos: the ratio).
struct NormSeg {
  int off;            // cf32 offset of the first channel in the results
  int nout;           // channels
  int wf;             // float index of the segment's weight in the results
  int kind;           // 0 cross, 1 autocorrelation, 2 phase-cal tones
  int ds1, ds2;       // datastreams (autos: ds2 = ds1)
  int cal1, cal2;     // Tsys: band-mean indices (cross: parallel autos of each band; autos: k and
                      // its matching band); -1 when that datastream has no Tsys
  int aw1, aw2;       // Tsys, cross: float indices of the two bands' parallel autocorrelation weights
  float zs1, zs2;     // zoom factors (1 for recorded bands)
  float chanfactor;   // (float)(blockspersend * 2 * freqchannels * chanavg)
  int decf;           // several phase centres: float index of this segment's shift decorrelation
                      // sum, the scale divided by it over the integration time; -1 otherwise
  int pcz;            // kind 2 (phase-cal tones of a band): a zero autocorrelation weight counts
                      // as 1 (the datastream has zoom bands, visibility.cpp:665-667)
};
// one band mean of a parallel autocorrelation (the Tsys calibration's autocorrcalibs)
struct NormCal {
  int off, nout;
};
struct NormArgs {
  float* res;                 // results, cf32 as floats
  const NormSeg* seg;
  const NormCal* cal;
  float* calmean;             // [ncal] real parts of the band means
  const double* tsys;         // [nds]
  const float* dpct;          // [nds] decorrelation fraction of each datastream's bits
  double fftsperint, msi;
  double inttime_s;           // integration time: meansubints subints
};
int fxg_launch_normalise(const NormArgs& a, int nseg, int ncal, hipStream_t st);

// Spectral kurtosis sums (Mode::process, mode.cpp:1237-1253): per recorded band and channel,
// the sums over a subint's FFTs of |X|^2 and |X|^4, read from the spectra chunk the F-engine
// wrote.
struct KurtBand {
  long long off;      // FP32 spectra: float2 offset of (band, FFT 0); FP16/BF16: uint32 offset of the row
  int C;              // channels
  int scale_idx;      // FP16/BF16: xscale index of the row
  int acc_off;        // float2 offset of the band's (s1, s2) sums
  int twlog;          // FP16/BF16: log2 of the row's channel tile width
};
struct KurtArgs {
  const KurtBand* tab;
  const float2* spec32;
  const uint32_t* spec16;
  const float* xscale;
  float2* acc;
  int nffts, chunkcap, mode;   // mode: FXG_X_FP32, FXG_X_FP16 or FXG_X_BF16
};
int fxg_launch_kurt(const KurtArgs& a, int nband, int maxC, int mintwlog, hipStream_t st);

// Phase-cal folding (the PCal extractors fed by Mode::process, mode.cpp:758-772): the unpacked
// samples of each valid FFT, indexed by their sample number since the scan start, summed into
// the M bins of the band's tone period (M = fs / gcd(fs, tone offset, tone interval)).
// A group is a run of consecutive recorded bands of one datastream with the same M; band
// band0 + k has bins [bins_off + k M, bins_off + (k + 1) M).  Workgroup split s of group g adds
// into part[part_off + s nb M ...]; a reduce sums the nsplit slices into the bins.
constexpr int PCAL_MAXNB = 16;           // bands per group
constexpr int PCAL_MAX_GSIZE = 32768;    // nb M floats per group (LDS)
constexpr int PCAL_FORM_REG = 1;         // 256 is a multiple of M: register sums
constexpr int PCAL_FORM_WHOLE = 2;       // the group is every band of a one-time-sample word
constexpr int PCAL_FORM_8BIT = 4;        // 8-bit samples (else 2-bit)
// A run of consecutive sample fields of one datastream folded with one period M: the fields
// of a time sample are its recorded bands, two per band (real, imaginary) for complex data.
struct PcalGroup {
  int ds, band0, nb, M;                  // first field, fields, period
  int bins_off;
  int form;                              // PCAL_FORM_* bits
  long long part_off;
  int fps;                               // fields per time sample
};
// One extracted tone: sum over the M folded bins j (real, or real and imaginary fields) of
// bin[j] exp(-2 pi i ((num j) mod den) / den); written to the tone buffer at `out` and added
// to the results at cf32 index `res` (-1: no results block).
struct PcalToneDev {
  int bins_off, M, cplx, num, den, out, res;
};
struct PcalArgs {
  const DevDs* ds;
  const DevDsRun* run;
  const FftDelay* params;    // [nds][chunkcap] of the current chunk
  const uint8_t* fvalid;
  const int2* fzone;
  const PcalGroup* groups;
  float* part;
  float* bins;
  int nffts, chunkcap, nsplit;
};
size_t fxg_pcal_lds(int max_gsize);
int fxg_launch_pcal(const PcalArgs& a, int form, int ngroups, int max_gsize, hipStream_t st);
int fxg_launch_pcal_reduce(const PcalArgs& a, int ngroups, int max_gsize, hipStream_t st);
int fxg_launch_pcal_tones(const PcalToneDev* t, int n, const float* bins, float2* tones, float2* results,
                          hipStream_t st);
