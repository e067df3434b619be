/*
 * fxg.h -- C ABI of the MI355X-native FX hot path (per-subintegration F and X engines).
 *
 * The library replaces two C++ call surfaces of mpifxcorr:
 *   - the station-based Mode object (mpifxcorr/src/mode.h:71-222), built by
 *     Configuration::getMode (mpifxcorr/src/configuration.cpp:932-998), and
 *   - Core::processdata (mpifxcorr/src/core.cpp:657-1129) with the accumulation
 *     Core::uvshiftAndAverageBaselineFreq (core.cpp:1605-1938) and
 *     Core::averageAndSendAutocorrs (core.cpp:1166-1376).
 * One fxg_ctx corresponds to one Core process with its Modes for one configuration.
 * A slot in (offsets, data buffers, control buffers) produces a slot out: the cf32
 * results array in the order of Configuration::populateResultLengths
 * (configuration.cpp:2348-2630), unnormalised, exactly as Core sends it to FxManager.
 *
 * Conventions: every entry point returns an int status (FXG_OK or a negative
 * FXG_E* code) and never aborts; fxg_last_error() gives the detailed message.
 * Inputs are borrowed until the call that consumes them returns.  No torch or HIP
 * types appear in the signatures; streams are passed as void* (a hipStream_t).
 */
#ifndef FXG_H
#define FXG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FXG_MAX_FREQS   32  /* recorded or zoom freqs per datastream; freqs per baseline */
#define FXG_MAX_BANDS   64  /* recorded or zoom bands per datastream */
#define FXG_MAX_POLPROD 4   /* polarisation products per baseline freq */
#define FXG_MAX_POLY    6   /* delay polynomial coefficients per model interval (order <= 5) */

/* status codes */
#define FXG_OK            0
#define FXG_E_INVALID    -1  /* invalid argument or unsupported configuration */
#define FXG_E_HIP        -2  /* HIP runtime error */
#define FXG_E_NOMEM      -3  /* device or host allocation failed */
#define FXG_E_STATE      -4  /* call out of order (e.g. wait on an unknown handle) */
#define FXG_E_NODEVICE   -5  /* no usable gfx950 device */

/* data formats (Configuration::dataformat, configuration.h) that this path decodes */
#define FXG_FMT_VDIF    0   /* VDIF, 32-byte header, LUT of format_vdif.c:66-145 */
#define FXG_FMT_MARK5B  1   /* Mark5B, 16-byte header, LUT of mark5_format_mark5b.c:49-86 */

/* X-engine kinds */
#define FXG_X_FP32  0   /* plain FP32 complex multiply-accumulate per baseline product (core.cpp:963) */
#define FXG_X_FP16  1   /* per-channel Hermitian station x station contraction on MFMA, FP16 in, FP32 acc */
#define FXG_X_BF16  2   /* same, BF16 in */

/* One entry of the frequency table (Configuration::freqdata, configuration.h). */
typedef struct fxg_freq {
  double  freq_mhz;   /* band-edge sky frequency (the LO for USB/LSB real data), MHz */
  double  bw_mhz;     /* bandwidth, MHz */
  int32_t lsb;        /* 1 = lower sideband */
  int32_t nchan;      /* numchannels */
  int32_t chanavg;    /* channelstoaverage */
  int32_t reserved;
} fxg_freq;

/* One datastream (Configuration::datastreamdata, configuration.h) plus its delay model
 * (Model scan table for one scan and one phase centre, model.h). */
typedef struct fxg_datastream {
  int32_t format;             /* FXG_FMT_* */
  int32_t nbits;              /* bits per (real or imaginary) sample: 2 or 8 */
  int32_t complex_sampling;   /* 0 = real, 1 = complex, 2 = complex double sideband (mode.cpp:41-49) */
  int32_t framebytes;         /* frame size including header */
  int32_t headerbytes;        /* 32 (VDIF) or 16 (Mark5B) */
  int32_t arraystridelen;     /* ARRAY STRIDE LEN; 0 = automatic (configuration.cpp:38-51) */
  int32_t nrecfreqs;
  int32_t recfreq_index[FXG_MAX_FREQS];        /* into the freq table */
  double  clock_offset_us[FXG_MAX_FREQS];      /* recordedfreqclockoffsets */
  double  clock_offset_delta_us[FXG_MAX_FREQS];/* recordedfreqclockoffsetsdelta */
  double  phase_offset_deg[FXG_MAX_FREQS];     /* recordedfreqphaseoffset */
  double  lo_offset_hz[FXG_MAX_FREQS];         /* recordedfreqlooffsets */
  int32_t nrecbands;                           /* bands (channels) in the sample word */
  int32_t recband_localfreq[FXG_MAX_BANDS];    /* local recorded freq index of each band */
  int32_t recband_pol[FXG_MAX_BANDS];          /* ASCII 'R','L','X','Y' */
  int32_t nzoomfreqs;
  int32_t zoomfreq_index[FXG_MAX_FREQS];            /* into the freq table */
  int32_t zoomfreq_parent_localfreq[FXG_MAX_FREQS]; /* local recorded freq index of the parent */
  int32_t zoomfreq_chanoffset[FXG_MAX_FREQS];       /* first parent channel of the zoom */
  int32_t nzoombands;
  int32_t zoomband_localfreq[FXG_MAX_BANDS];   /* local zoom freq index of each zoom band */
  int32_t zoomband_pol[FXG_MAX_BANDS];
  double  clock_coeff0_us;    /* clock offset term 0; intclockseconds = floor(c0/1e6 + 0.5) (mode.cpp:69) */
  int32_t model_order;        /* delay polynomial order (<= 5) */
  int32_t model_nsamples;     /* number of model intervals of length model_inc_secs */
  const double* model_delay_us; /* [model_nsamples][FXG_MAX_POLY]; t in s from interval start:
                                   the phase centre with one phase centre, the pointing centre
                                   with several (Mode::setOffsets source index, mode.cpp:1518-1521) */
  const double* phasecentre_delay_us; /* [nphasecentres][model_nsamples][FXG_MAX_POLY]: phase
                                   centres 1..P (Model source indices 1..P); read only when the
                                   config has nphasecentres > 1, else may be NULL */
  int64_t phasecal_interval_hz; /* phasecalintervalhz: > 0 extracts the phase-cal tones of every
                                   recorded band each subint into the results' phase-cal block
                                   (Mode's extractors, Core::copyPCalTones, core.cpp:1131-1163);
                                   0 = off (fxg_set_pcal may still turn extraction on, without
                                   a results block) */
  int64_t phasecal_base_hz;     /* phasecalbasehz */
} fxg_datastream;

/* One baseline (Configuration::baselinedata, configuration.h): datastream pair and, per
 * baseline freq, the band index pairs of each polarisation product.  Band indices count
 * recorded bands first, then zoom bands (mode.cpp:164-196).  Both bands of a product belong
 * to the baseline freq, except for USB x LSB: the second datastream's band may be a recorded
 * band of the opposite sideband covering the same sky band (configuration.cpp:3125-3135). */
typedef struct fxg_baseline {
  int32_t ds1, ds2;                 /* ordered datastream indices */
  int32_t nfreqs;
  int32_t freq_index[FXG_MAX_FREQS];
  int32_t npolprod[FXG_MAX_FREQS];
  int32_t ds1band[FXG_MAX_FREQS][FXG_MAX_POLPROD];
  int32_t ds2band[FXG_MAX_FREQS][FXG_MAX_POLPROD];
} fxg_baseline;

/* Flattened Configuration for one config index. */
typedef struct fxg_config {
  int32_t nfreqs;
  const fxg_freq* freqs;
  int32_t ndatastreams;
  const fxg_datastream* datastreams;
  int32_t nbaselines;
  const fxg_baseline* baselines;
  int64_t subint_ns;          /* SUBINT NANOSECONDS */
  int32_t guard_ns;           /* GUARD NANOSECONDS */
  int32_t fringe_rot_order;   /* FRINGE ROTN ORDER: 0 (post-F), 1 (linear) or 2 (quadratic) */
  int32_t xmac_stride_len;    /* XMAC STRIDE LEN (the result layout does not depend on it) */
  int32_t num_buffered_ffts;  /* NUM BUFFERED FFTS */
  int32_t write_cross_autocorrs; /* cross-polarisation autocorrelations (conf.writeautocorrs) */
  double  model_inc_secs;     /* model interval length (Model::modelincsecs) */
  double  scan_start_sec;     /* Model::getScanStartSec of the scan, for wallclock LO-offset terms */
  int32_t xengine;            /* FXG_X_* */
  int32_t chunk_ffts;         /* FFTs per device spectra chunk; 0 = automatic */
  int32_t fengine_mode;       /* FXG_F_AUTO, FXG_F_PER_LENGTH or FXG_F_PERSISTENT */
  int32_t fengine_threads;    /* workgroup size 128, 256 or 512 (persistent: 256, 512); 0 = by FFT length (persistent: 512) */
  int32_t autotune;           /* 1: time the F-engine launch candidates and the MFMA X-engine's
                                 8- and 16-channel tiles at create and keep the fastest */
  int32_t nphasecentres;      /* Model::getNumPhaseCentres; 0 or 1 = one phase centre, up to
                                 FXG_MAX_PHASECENTRES: results per phase centre, uv-shifted
                                 (core.cpp:1605-1938) */
  int32_t process_threads;    /* Core process threads (their FFT-block ranges start the uv-shift
                                 intervals, core.cpp:2158-2166); 0 = 1 */
  int32_t reserved;
  int64_t max_ns_between_xc_avg; /* Model::getMaxNSBetweenXCAvg: uv-shift interval (core.cpp:771-776) */
} fxg_config;

#define FXG_MAX_PHASECENTRES 16

/* F-engine launch forms: one launch per FFT length (datastreams grouped by length), or every
 * length in one persistent launch whose workgroups take (length, datastream, part) work items. */
#define FXG_F_AUTO 0
#define FXG_F_PER_LENGTH 1
#define FXG_F_PERSISTENT 2

/* Result layout summary (Configuration::populateResultLengths and getDataBytes). */
typedef struct fxg_layout {
  int64_t coreresultlength;   /* cf32 elements of the results array */
  int32_t blockspersend;      /* FFTs per subint (configuration.cpp:1843-1852) */
  int32_t nbaselinefreqs;     /* entries of the per (freq, baseline) offset tables */
  int64_t max_databytes;      /* largest per-datastream slot data size in bytes */
} fxg_layout;

typedef struct fxg_ctx fxg_ctx;
typedef int32_t fxg_handle;

/* Create a context: validates the configuration, builds the result layout, allocates the
 * device-resident state.  Replaces Configuration::getMode + the Core constructor
 * (configuration.cpp:932-998, core.cpp:23-154). */
int fxg_create(const fxg_config* cfg, fxg_ctx** out);
/* A context without device state (no device needed): layout queries, fxg_difx_records;
 * the processing calls return FXG_E_STATE on it.  Released with fxg_destroy. */
int fxg_create_layout(const fxg_config* cfg, fxg_ctx** out);
int fxg_destroy(fxg_ctx* ctx);
const char* fxg_strerror(int code);
/* The detailed message of the context's last failure; with a null context, that of this
 * thread's last failed fxg_create / fxg_create_layout / fxg_footprint. */
const char* fxg_last_error(const fxg_ctx* ctx);

int fxg_get_layout(const fxg_ctx* ctx, fxg_layout* out);
/* Configuration::getDataBytes (configuration.cpp:794-820) for datastream ds. */
int fxg_data_bytes(const fxg_ctx* ctx, int32_t ds, int64_t* bytes);
/* Offsets (cf32 units) of populateResultLengths: baseline[freq*nbaselines+b] (or -1),
 * bweight[freq*nbaselines+b] (or -1), autocorr[ds], acweight[ds].  Any pointer may be null. */
int fxg_result_offsets(const fxg_ctx* ctx, int32_t* baseline, int32_t* bweight,
                       int32_t* autocorr, int32_t* acweight);
/* Configuration::getCoreResultPCalOffset (configuration.cpp:2611-2619) of every datastream, in
 * cf32 units: pcal[ds]; the block holds, for datastreams with a phase-cal interval in the
 * configuration, each recorded band's tones in band order (Core::copyPCalTones,
 * core.cpp:1131-1163). */
int fxg_result_pcal_offsets(const fxg_ctx* ctx, int32_t* pcal);

/* Core::processdata for one slot, host buffers (core.cpp:657-1129): offsets = {scan, sec, ns}
 * of the subint; data[ds] = the slot's data buffer of nbytes[ds] bytes; control[ds] =
 * {scan, sec, ns, validflags...} (datastream.cpp:380, 581-605).  results receives
 * coreresultlength cf32 values (interleaved re, im).  Synchronous. */
int fxg_process_subint(fxg_ctx* ctx, const int32_t offsets[3], const uint8_t* const* data,
                       const int32_t* nbytes, const int32_t* const* control, float* results);

/* Same, with the data already resident in device memory (d_data[ds]) and the results
 * written to device memory (d_results, coreresultlength cf32).  Work is enqueued on
 * `stream` (a hipStream_t, may be null for the context's own stream) and is complete when
 * the stream is.  Control words are read from host memory before the call returns. */
int fxg_process_subint_device(fxg_ctx* ctx, const int32_t offsets[3],
                              const uint8_t* const* d_data, const int32_t* nbytes,
                              const int32_t* const* control, float* d_results, void* stream);

/* Asynchronous slot ring (Core's RECEIVE_RING_LENGTH slots, core.cpp:156): submit copies the
 * host inputs into a free slot and enqueues the work; wait blocks until that slot is done
 * and copies its results out.  At most fxg_ring_length() submissions may be outstanding. */
int fxg_ring_length(const fxg_ctx* ctx);
int fxg_submit_subint(fxg_ctx* ctx, const int32_t offsets[3], const uint8_t* const* data,
                      const int32_t* nbytes, const int32_t* const* control, fxg_handle* handle);
int fxg_wait(fxg_ctx* ctx, fxg_handle handle, float* results);

/* Mode getters for the last processed subint (mode.h:138-203), for the FFT blocks of the last
 * device chunk: spectrum of band `band` at FFT `fft` (getFreqs, C cf32 values; zeros if the
 * block was flagged) and the per-FFT data weights (getDataWeight) of datastream ds. */
int fxg_get_spectrum(const fxg_ctx* ctx, int32_t ds, int32_t band, int32_t fft, float* out);
int fxg_get_data_weights(const fxg_ctx* ctx, int32_t ds, float* out /* [blockspersend] */);

/* Spectral kurtosis (Mode::setDumpKurtosis / calculateAndAverageKurtosis, mode.cpp:1237-1253,
 * 1405-1450; sent by Core::averageAndSendKurtosis, core.cpp:1378-1429).  When enabled, every
 * subint also sums |X|^2 and |X|^4 per recorded band and channel over its FFTs (one more pass
 * over each spectra chunk).  fxg_get_kurtosis gives the last subint's estimator
 * (numblocks s2 / s1^2 - 1) numblocks/(numblocks-1) - 1, numblocks = blockspersend, summed
 * over groups of C/maxchannels channels (and divided by the group size) when maxchannels < C;
 * *nout = the values written (0 when the band had no data, where the reference sends no
 * record). */
int fxg_set_kurtosis(fxg_ctx* ctx, int32_t enable);
/* The STA autocorrelation record data of a recorded band (Core::averageAndSendAutocorrs,
 * core.cpp:1166-1258) from a slot's results (host memory): the parallel autocorrelation's real
 * parts summed over groups of channels down to min(stadumpchannels, channels) and scaled by
 * 1/(2 nchan weight); nswidth is the record's span in ns (its minimum weight is 0.333 of a full
 * span's).  *nout = 0 when the weight is below it (the reference sends no record). */
int fxg_sta_autocorr(const fxg_ctx* ctx, const float* results, int32_t ds, int32_t band, int32_t stadumpchannels,
                     double nswidth, float* out, int32_t* nout);
int fxg_get_kurtosis(const fxg_ctx* ctx, int32_t ds, int32_t band, int32_t maxchannels, float* out,
                     int32_t* nout);

/* Phase-cal extraction of datastream ds (Mode's PCal extractors, mode.cpp:401-420, 758-772,
 * 1552-1568; tone table of Configuration, configuration.cpp:1765-1835), real-sampled (2- or
 * 8-bit) or complex-sampled (8-bit) data with a whole-Hz tone interval.  interval_hz = 0 turns
 * it off.  The tones of each recorded band follow Configuration: the first tone above (USB) or
 * below (LSB) the band edge, counted up to the far edge; offset = harmonic interval -
 * long(1e6 lofreq) (USB; mirrored for LSB; base_hz moves the comb but, as in the reference, not
 * the offset).  Tone k is sum over the valid FFTs' samples x[m] exp(-2 pi i f_k m), m the
 * sample number since the scan start, at the frequency f_k (cycles per sample) that the
 * reference's extractor reads: real data (offset + k interval) / fs, fs = 2 bw (the trivial
 * extractor skips DC: (k + 1) interval / fs for a zero offset); complex data (fs = bw) the DFT
 * bin of the implicit-shift or shifting complex extractor (pcal.cpp:808-1020, 1419-1590), whose
 * length doubles to 1 MHz resolution, reversed for double sideband, and zero for a bin past
 * the DFT.  The device folds the samples by m modulo the tone period M (a multiple of every
 * tone's denominator, at most 16384; real and imaginary parts separately) and forms the tones
 * after the subint.  fxg_get_pcal gives the last subint's tones as interleaved (re, im) floats;
 * *ntones = the count.  A flagged offset (a band with no tone: one placeholder tone; or an offset
 * made negative by base_hz) gives zero tones, as the reference's dummy extractor.  FXG_E_STATE
 * when phase-cal is off for the band.  A datastream whose configuration sets
 * phasecal_interval_hz has extraction on from fxg_create, its tones also added every subint
 * into the results' phase-cal block; fxg_set_pcal with another setting for it is refused
 * (FXG_E_INVALID), as the block's size depends on it. */
int fxg_set_pcal(fxg_ctx* ctx, int32_t ds, int64_t interval_hz, int64_t base_hz);
int fxg_get_pcal(const fxg_ctx* ctx, int32_t ds, int32_t band, float* out, int32_t* ntones);
/* The tone table entry of a recorded band (Configuration's getDRecordedFreqNumPCalTones and
 * getDRecordedFreqPCalOffsetsHz) and its tone period in samples (0 for a flagged offset).
 * fxg_set_pcal and this also work on a host-only context (fxg_create_layout), which builds
 * the tables without device memory. */
int fxg_pcal_info(const fxg_ctx* ctx, int32_t ds, int32_t band, int32_t* ntones, int64_t* offset_hz,
                  int32_t* period);

/* Mk5Mode::unpack (mk5mode.cpp:94-174) as a device call: decode nsamples time samples starting
 * at sample `startsample` of datastream ds from d_data (nbytes) into d_out[nrecbands][nsamples]
 * (f32, or interleaved cf32 for complex sampling), and the count of valid samples into
 * d_goodsamples[0] (int32).  Frames that are invalid or blanked decode as zeros. */
int fxg_unpack_device(fxg_ctx* ctx, int32_t ds, const uint8_t* d_data, int32_t nbytes,
                      int64_t startsample, int32_t nsamples, float* d_out,
                      int32_t* d_goodsamples, void* stream);

/* Model::calculateDelayInterpolator (model.cpp:219-289), order 2, as used by Mode::setOffsets
 * (mode.cpp:1523-1526): evaluated on the device in FP64.  out[3] = quadratic coefficients in
 * FFT-index units with 1e6*intclockseconds removed from out[2].  Returns FXG_E_INVALID if the
 * time is outside the model. */
int fxg_delay_interpolator(fxg_ctx* ctx, int32_t ds, int32_t sec, int32_t ns, double* out);

/* Kernel timing (no counterpart in the reference; for measurement harnesses).  When enabled,
 * HIP events bracket every F-engine, X-engine and accumulation launch on the stream they are
 * launched on; fxg_get_timing synchronizes and returns the summed durations since the last
 * enable. */
typedef struct fxg_timing {
  double  f_ms;   int32_t f_launches;    /* fused F-engine (unpack, rotate, FFT, post-F) */
  double  x_ms;   int32_t x_launches;    /* X-engine */
  double  other_ms; int32_t other_launches; /* setup, accumulation, weights */
} fxg_timing;
int fxg_set_timing(fxg_ctx* ctx, int32_t enable);
int fxg_get_timing(fxg_ctx* ctx, fxg_timing* out);

/* The F-engine launch form in use and, after an autotuned create, the candidates with their
 * measured subint times, as text (no counterpart in the reference). */
int fxg_autotune_report(const fxg_ctx* ctx, char* buf, int32_t len);

/* Sizing without a device (Core::getEstimatedBytes, core.cpp:2197-2210, and
 * Mode::getEstimatedBytes, mode.h:216): validates cfg and reports the device bytes
 * fxg_create would hold, coreresultlength, and the bytes of one slot's input data (all
 * datastreams).  Any output pointer may be null. */
int fxg_footprint(const fxg_config* cfg, int64_t* device_bytes, int64_t* coreresultlength,
                  int64_t* slot_bytes);
/* Input samples one subint processes: sum over datastreams of recorded bands x
 * blockspersend x fftchannels (a complex sample counts once). */
int fxg_input_samples(const fxg_ctx* ctx, int64_t* samples);
/* Device bytes held by a context. */
int fxg_device_bytes(const fxg_ctx* ctx, int64_t* bytes);

/* ---- VDIF thread corner turn ---------------------------------------------------------
 * Replaces VDIFMuxer (datamuxer.h:148-240): deinterlace (datamuxer.cpp:764-833) and
 * multiplex (datamuxer.cpp:670-757) with cornerturn_generic (datamuxer.cpp:243-266) for one
 * segment, on the device.  Input: single-channel VDIF frames of N threads in any order;
 * output: N-channel frames, the stream an fxg datastream of nrecbands = N decodes.  Slot s of
 * the segment is the frame whose count since the reference frame,
 *   ((mjd - ref_mjd) * 86400 + second - ref_second) * fps + frame - ref_frame,
 * equals start + s.  Per slot, in slot order, starting at the first slot with any thread
 * present (earlier ones are not output):
 *   all threads present   thread 0's header with invalid 0, log2(N) channels, the output
 *                         frame length and thread id 0, then the corner-turned payload
 *   some present          the first present thread's header with invalid 1, the output frame
 *                         length and thread id 0; the payload bytes are not written
 *   none present          the previous output header with invalid 1, frame number + 1
 *                         (wrapping to frame 0 of the next second at fps)
 * A later copy of a (slot, thread) in the input replaces an earlier one.  Frames of an
 * unmapped thread, before the segment or after its last slot are dropped and counted (the
 * reference keeps frames after the segment for the next one: pass them again). */
#define FXG_MUX_MAX_THREADS 32
typedef struct fxg_vdif_mux_params {
  int32_t inputframebytes;     /* one single-thread frame incl. its 32-byte header; multiple of 8 */
  int32_t nthreads;            /* N: a power of two <= FXG_MUX_MAX_THREADS */
  int32_t bitspersample;       /* 1, 2, 4, 8, 16 or 32, with N * bits <= 32 */
  int32_t framespersecond;     /* frames per second of one thread */
  int32_t threadmap[FXG_MUX_MAX_THREADS];  /* VDIF thread id of output channel j (the
                                              reference's threadindexmap) */
  int64_t ref_seconds;         /* reference frame: MJD * 86400 + seconds of the epoch */
  int32_t ref_frame;           /* reference frame: frame number within its second */
  int32_t reserved;
  int64_t start;               /* count of slot 0 since the reference frame */
} fxg_vdif_mux_params;
/* Device workspace for nslots slots of N threads (bytes). */
int64_t fxg_vdif_mux_workspace_bytes(int64_t nslots, int32_t nthreads);
/* Enqueue the corner turn of one segment of nslots slots on `stream` (0 = default).
 * d_in: in_bytes of input frames (8-byte aligned); d_out: out_bytes >= nslots * ((ifb - 32)
 * * N + 32) (16-byte aligned), written from its start; d_work: workspace_bytes.  Returns
 * at once; fxg_vdif_mux_result gives the outcome. */
int fxg_vdif_mux_device(const fxg_vdif_mux_params* p, const uint8_t* d_in, int64_t in_bytes,
                        uint8_t* d_out, int64_t out_bytes, int64_t nslots, void* d_work,
                        int64_t work_bytes, void* stream);
/* Synchronize `stream` and read the outcome of the last corner turn with this workspace: the
 * number of output frames (multiplex returns that times the output frame length) and the
 * dropped input frames {before the segment, after it, unmapped thread}.  FXG_E_INVALID if
 * an input frame's length field differed from inputframebytes (the reference aborts
 * there; the output is then undefined). */
int fxg_vdif_mux_result(const void* d_work, void* stream, int64_t* nout, int64_t* dropped /* [3] */);
/* No counterpart in the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight capture and serving loops, harnesses that forbid hipMalloc): the
 * corner turn of fxg_vdif_mux_device with its workspace in a library-owned static array of
 * FXG_MUX_STATIC_WORK_BYTES per GPU instead of d_work; segments whose workspace would be larger
 * (fxg_vdif_mux_workspace_bytes, about 4 (N + 1) bytes per slot) are refused with
 * FXG_E_INVALID.  Nothing is allocated.  The outcome goes to the caller's device buffer
 * d_outcome[5] (8-byte aligned): {output frames, dropped before the segment, after it,
 * unmapped thread, 1 if an input frame's length field differed}.  Every call on one GPU shares
 * that GPU's array (cleared by each call): calls that overlap on one GPU race. */
#define FXG_MUX_STATIC_WORK_BYTES (1 << 20)
int fxg_vdif_mux_device_static(const fxg_vdif_mux_params* p, const uint8_t* d_in, int64_t in_bytes,
                               uint8_t* d_out, int64_t out_bytes, int64_t nslots, int64_t* d_outcome,
                               void* stream);

/* ---- Slot start and valid flags: DataStream::calculateControlParams ------------------
 * Replaces DataStream::calculateControlParams (datastream.cpp:368-627), the coarse-delay
 * choice of where in the datastream's segmented read buffer a subint's data starts and which
 * of its FFT blocks have data, which fills the control words {scan, sec, ns, validflags...}
 * that fxg_process_subint takes.  Host code; the buffer ring, its readers and its locks stay
 * the caller's: when the slot lies beyond the current segment, `advance` is called before
 * each move to the next segment (the reference's waitForSendComplete, lock of segment
 * at + 2 and unlock of at, datastream.cpp:438-457). */
#define FXG_INVALID_SUBINT (-99999999)  /* Mode::INVALID_SUBINT (mode.h:222) */
#define FXG_FLAGS_PER_INT  30           /* FLAGS_PER_INT (mpifxcorr.h:29) */
/* One segment of the read buffer (DataStream::readinfo, datastream.h:87-113: the fields the
 * choice reads).  The reader sets scan .. readto as it fills the segment;
 * fxg_segment_params sets the rest from the configuration (DataStream::updateConfig,
 * datastream.cpp:745-763). */
typedef struct fxg_segment {
  int32_t scan, scanseconds, scanns;  /* time of the segment's first byte */
  int32_t validbytes;                 /* bytes read into the segment */
  int32_t readto;                     /* 1 once the segment has been read into */
  int32_t sendbytes;                  /* data bytes of one subint plus its guard */
  int32_t blockspersend;              /* FFT blocks per subint */
  int32_t bytespersamplenum, bytespersampledenom;
  int32_t bytesbetweenintegerns;      /* bytes between samples at whole nanoseconds */
  int32_t nsinc;                      /* nanoseconds one segment spans */
  int32_t numchannels;                /* channels of the first recorded freq */
  int32_t complex_sampling;           /* 1 if the sampling is complex */
  int32_t reserved;
  double  sampletimens;               /* nanoseconds per sample */
} fxg_segment;
/* Framed streams: the choice above counts payload bytes only; VDIFDataStream
 * (vdiffile.cpp:348-472) and Mark5BDataStream (mark5bfile.cpp:121-247) then step back to
 * the start of a frame (a multiple of framegranularity frames into the segment) and take
 * its time. */
#define FXG_FRAMING_NONE   0
#define FXG_FRAMING_VDIF   1
#define FXG_FRAMING_MARK5B 2
typedef struct fxg_databuffer {
  int32_t numdatasegments;
  int32_t bufferbytes;                /* whole ring */
  int32_t readbytes;                  /* bufferbytes / numdatasegments */
  int32_t keepreading;                /* the reader is still active */
  int32_t lastvalidsegment;           /* last segment read when the reader stopped */
  int32_t intclockseconds;            /* whole seconds of the clock offset (mode.cpp:69) */
  int32_t framing;                    /* FXG_FRAMING_* */
  int32_t framebytes, payloadbytes;   /* framed: bytes of a frame and of its payload */
  int32_t framespersecond;            /* framed: frames per second */
  int32_t framegranularity;           /* framed: frames per whole-nanosecond step */
  int32_t reserved;
  const fxg_segment* seg;             /* [numdatasegments] */
} fxg_databuffer;
typedef int (*fxg_advance_fn)(void* user, int32_t from_segment);
/* The config-derived fields of a segment for datastream ds (no device needed), with the
 * frame-based nsinc and sendbytes of VDIFDataStream::updateConfig (vdiffile.cpp:475-492) and
 * Mark5BDataStream::updateConfig (mark5bfile.cpp:249-267). */
int fxg_segment_params(const fxg_config* cfg, int32_t ds, int32_t numdatasegments, int32_t bufferbytes,
                       fxg_segment* seg);
/* The config-derived fields of a ring (numdatasegments .. intclockseconds and the framing of
 * the datastream's format); keepreading, lastvalidsegment and seg stay the caller's. */
int fxg_buffer_params(const fxg_config* cfg, int32_t ds, int32_t numdatasegments, int32_t bufferbytes,
                      fxg_databuffer* buf);
/* The choice for the subint at (scan, offsetsec, offsetns).  delayus1 / delayus2: the
 * datastream's delay in microseconds at the subint start and at the end of its data span
 * (Model::calculateDelayInterpolator with order 0; delay_ok = 0 if the model has no
 * interpolator there).  *atsegment is the caller's current segment, moved on as the
 * reference moves it.  control receives 3 + ceil(blockspersend / 30) words (an invalid
 * subint has control[1] = FXG_INVALID_SUBINT and zero flags); *bufferindex the byte offset of
 * the subint's data in the ring (0 when invalid).  `advance` returns 0 to move on (after the
 * segments have been refilled as needed) and nonzero to stop with FXG_E_STATE; with a null
 * `advance` the choice stops with FXG_E_STATE after one pass round the ring. */
int fxg_control_params(const fxg_databuffer* buf, int32_t* atsegment, int32_t scan, int32_t offsetsec,
                       int32_t offsetns, int32_t delay_ok, double delayus1, double delayus2,
                       fxg_advance_fn advance, void* user, int32_t* control, int32_t* bufferindex);

/* ---- Normalisation and the DIFX record writer ------------------------------------------
 * Visibility::writedata's amplitude scaling (visibility.cpp:353-646) and writeSWIN's binary
 * records (visibility.cpp:781-931, header appendSWINHeaderBuffered :1205-1234, SYNC_WORD
 * visibility.h:125), for one phase centre and no pulsar bins. */
/* Scale a results array (device memory; one slot, or the sum of meansubints slots) in place:
 * cross-correlations by 1/(weight * meansubints * blockspersend * 2 * nchan) (with the zoom
 * bandwidth factor), or, where both datastreams have tsys[ds] > 0, by
 * sqrt(Tsys1 Tsys2 / divisor) / weight from the band means of the parallel autocorrelations;
 * autocorrelations likewise when the configuration writes them.  tsys may be null (none).
 * Returns when the scaling is done. */
int fxg_normalise_device(fxg_ctx* ctx, float* d_results, double meansubints, const double* tsys,
                         void* stream);
typedef struct fxg_difx_meta {
  int32_t mjd;                     /* integration centre: MJD */
  int32_t configindex;
  int32_t sourceindex;             /* source index of the records (autocorrelations: the
                                      pointing centre's when there are several phase centres) */
  int32_t phasecentre;             /* several phase centres: the one whose file this is (0-based;
                                      writeSWIN writes one file per phase centre, visibility.cpp:
                                      848-885, the autocorrelations into the first); else 0 */
  double  seconds;                 /* integration centre: seconds of the day */
  double  fftsperintegration;      /* meansubints * blockspersend (record weights = weight / it) */
  const int32_t* baselinenumber;   /* [nbaselines] Configuration::getBNumber */
  const double* uvw;               /* [nbaselines][3] metres (Model::interpolateUVW) */
  const int32_t* telescopeindex;   /* [ndatastreams] for the autocorrelation baseline numbers */
} fxg_difx_meta;
/* The DIFX records of a normalised results array (host memory) into out[capacity]: cross-
 * correlations by baseline, freq, product; then autocorrelations when the configuration
 * writes them; a product of weight 0 is skipped.  With several phase centres the records of
 * meta->phasecentre's file, cross weights times that centre's shift decorrelation over the
 * integration (fftsperintegration / blockspersend subints).  out = null: size only. */
int fxg_difx_records(const fxg_ctx* ctx, const float* results, const fxg_difx_meta* meta, uint8_t* out,
                     int64_t capacity, int64_t* nbytes, int32_t* nrecords);

/* ---- Serving: concurrent correlation jobs on one device ------------------------------
 * No counterpart in the reference, which runs one job per mpifxcorr process; this replaces
 * the FxManager -> Core subint dispatch (fxmanager.cpp:331-336, core.cpp:205-330) and the
 * Visibility::addData integration (visibility.cpp:192-205) for many jobs at once.
 *
 * A server owns a scheduler thread.  Jobs are admitted in arrival order while the sum of
 * their device footprints stays within hbm_budget_bytes and at most max_jobs are active;
 * each admitted job gets its own context and stream, so the subints of different jobs
 * overlap on the device.  A job keeps its delay-model tables, per-subint delay state and
 * its integration accumulator (the sum of its subint results, unnormalised) in device
 * memory until it finishes.  A finished job's memory is released when a waiting job needs
 * it, which admits that job. */
typedef struct fxg_server fxg_server;
typedef struct fxg_server_config {
  int64_t hbm_budget_bytes;   /* admission limit; 0 = 85% of the device memory free at create */
  int32_t max_jobs;           /* jobs active at once (0 = 64) */
  int32_t inflight_per_job;   /* subints of one job enqueued ahead (0 = 2) */
} fxg_server_config;
typedef struct fxg_server_stats {
  int32_t jobs_added, jobs_admitted, jobs_finished, jobs_active;
  int32_t max_active;         /* most jobs active at once */
  int32_t subints_done;
  int64_t samples_done;       /* input samples of the finished subints */
  int64_t admitted_bytes, peak_admitted_bytes;
  double  first_launch_s, last_done_s;  /* seconds since server creation */
  double  latency_p50_ms, latency_p99_ms, latency_max_ms;  /* push -> results integrated */
} fxg_server_stats;
int fxg_server_create(const fxg_server_config* cfg, fxg_server** out);
int fxg_server_destroy(fxg_server* srv);
/* Queue a job of nsubints subints; returns at once with its id (admission happens later).
 * With data_on_device the job's subint data will be device pointers used in place; otherwise
 * host pointers, copied into per-job device input slots (counted in its footprint). */
int fxg_server_add_job(fxg_server* srv, const fxg_config* cfg, int32_t nsubints,
                       int32_t data_on_device, int32_t* job);
/* One subint of a job (arguments as fxg_process_subint).  The data pointers must stay valid
 * until fxg_server_wait_job returns; control words are copied during the call. */
int fxg_server_push_subint(fxg_server* srv, int32_t job, const int32_t offsets[3],
                           const uint8_t* const* data, const int32_t* nbytes,
                           const int32_t* const* control);
/* Block until all subints of the job are integrated; copies the integrated results
 * (coreresultlength cf32) and the job's p50 / p99 subint latency (ms). */
int fxg_server_wait_job(fxg_server* srv, int32_t job, float* integrated, double* p50_ms, double* p99_ms);
int fxg_server_get_stats(fxg_server* srv, fxg_server_stats* out);
/* Why a job failed (empty if it did not). */
int fxg_server_job_error(fxg_server* srv, int32_t job, char* buf, int32_t len);

#ifdef __cplusplus
}
#endif
#endif /* FXG_H */
