"""ctypes mirror of include/fxg.h (structure layouts and constants).

The structures mirror the flattened Configuration tables that the C ABI takes:
fxg_freq <- Configuration::freqdata, fxg_datastream <- datastreamdata plus the Model scan
table of one scan, fxg_baseline <- baselinedata, fxg_config <- configdata
(mpifxcorr/src/configuration.h:781-950).
"""
import ctypes as C

FXG_MAX_FREQS = 32
FXG_MAX_BANDS = 64
FXG_MAX_POLPROD = 4
FXG_MAX_POLY = 6

FXG_OK = 0
FXG_E_INVALID = -1
FXG_E_HIP = -2
FXG_E_NOMEM = -3
FXG_E_STATE = -4
FXG_E_NODEVICE = -5

FXG_FMT_VDIF = 0
FXG_FMT_MARK5B = 1

FXG_X_FP32 = 0
FXG_X_FP16 = 1
FXG_X_BF16 = 2

FXG_F_AUTO = 0
FXG_F_PER_LENGTH = 1
FXG_F_PERSISTENT = 2


class fxg_freq(C.Structure):
    _fields_ = [
        ("freq_mhz", C.c_double),
        ("bw_mhz", C.c_double),
        ("lsb", C.c_int32),
        ("nchan", C.c_int32),
        ("chanavg", C.c_int32),
        ("reserved", C.c_int32),
    ]


class fxg_datastream(C.Structure):
    _fields_ = [
        ("format", C.c_int32),
        ("nbits", C.c_int32),
        ("complex_sampling", C.c_int32),
        ("framebytes", C.c_int32),
        ("headerbytes", C.c_int32),
        ("arraystridelen", C.c_int32),
        ("nrecfreqs", C.c_int32),
        ("recfreq_index", C.c_int32 * FXG_MAX_FREQS),
        ("clock_offset_us", C.c_double * FXG_MAX_FREQS),
        ("clock_offset_delta_us", C.c_double * FXG_MAX_FREQS),
        ("phase_offset_deg", C.c_double * FXG_MAX_FREQS),
        ("lo_offset_hz", C.c_double * FXG_MAX_FREQS),
        ("nrecbands", C.c_int32),
        ("recband_localfreq", C.c_int32 * FXG_MAX_BANDS),
        ("recband_pol", C.c_int32 * FXG_MAX_BANDS),
        ("nzoomfreqs", C.c_int32),
        ("zoomfreq_index", C.c_int32 * FXG_MAX_FREQS),
        ("zoomfreq_parent_localfreq", C.c_int32 * FXG_MAX_FREQS),
        ("zoomfreq_chanoffset", C.c_int32 * FXG_MAX_FREQS),
        ("nzoombands", C.c_int32),
        ("zoomband_localfreq", C.c_int32 * FXG_MAX_BANDS),
        ("zoomband_pol", C.c_int32 * FXG_MAX_BANDS),
        ("clock_coeff0_us", C.c_double),
        ("model_order", C.c_int32),
        ("model_nsamples", C.c_int32),
        ("model_delay_us", C.POINTER(C.c_double)),
        ("phasecentre_delay_us", C.POINTER(C.c_double)),
        ("phasecal_interval_hz", C.c_int64),
        ("phasecal_base_hz", C.c_int64),
    ]


class fxg_baseline(C.Structure):
    _fields_ = [
        ("ds1", C.c_int32),
        ("ds2", C.c_int32),
        ("nfreqs", C.c_int32),
        ("freq_index", C.c_int32 * FXG_MAX_FREQS),
        ("npolprod", C.c_int32 * FXG_MAX_FREQS),
        ("ds1band", (C.c_int32 * FXG_MAX_POLPROD) * FXG_MAX_FREQS),
        ("ds2band", (C.c_int32 * FXG_MAX_POLPROD) * FXG_MAX_FREQS),
    ]


class fxg_config(C.Structure):
    _fields_ = [
        ("nfreqs", C.c_int32),
        ("freqs", C.POINTER(fxg_freq)),
        ("ndatastreams", C.c_int32),
        ("datastreams", C.POINTER(fxg_datastream)),
        ("nbaselines", C.c_int32),
        ("baselines", C.POINTER(fxg_baseline)),
        ("subint_ns", C.c_int64),
        ("guard_ns", C.c_int32),
        ("fringe_rot_order", C.c_int32),
        ("xmac_stride_len", C.c_int32),
        ("num_buffered_ffts", C.c_int32),
        ("write_cross_autocorrs", C.c_int32),
        ("model_inc_secs", C.c_double),
        ("scan_start_sec", C.c_double),
        ("xengine", C.c_int32),
        ("chunk_ffts", C.c_int32),
        ("fengine_mode", C.c_int32),
        ("fengine_threads", C.c_int32),
        ("autotune", C.c_int32),
        ("nphasecentres", C.c_int32),
        ("process_threads", C.c_int32),
        ("reserved", C.c_int32),
        ("max_ns_between_xc_avg", C.c_int64),
    ]


class fxg_layout(C.Structure):
    _fields_ = [
        ("coreresultlength", C.c_int64),
        ("blockspersend", C.c_int32),
        ("nbaselinefreqs", C.c_int32),
        ("max_databytes", C.c_int64),
    ]


class fxg_timing(C.Structure):
    _fields_ = [
        ("f_ms", C.c_double), ("f_launches", C.c_int32),
        ("x_ms", C.c_double), ("x_launches", C.c_int32),
        ("other_ms", C.c_double), ("other_launches", C.c_int32),
    ]


class fxg_server_config(C.Structure):
    _fields_ = [
        ("hbm_budget_bytes", C.c_int64),
        ("max_jobs", C.c_int32),
        ("inflight_per_job", C.c_int32),
    ]


class fxg_server_stats(C.Structure):
    _fields_ = [
        ("jobs_added", C.c_int32), ("jobs_admitted", C.c_int32), ("jobs_finished", C.c_int32),
        ("jobs_active", C.c_int32), ("max_active", C.c_int32), ("subints_done", C.c_int32),
        ("samples_done", C.c_int64), ("admitted_bytes", C.c_int64), ("peak_admitted_bytes", C.c_int64),
        ("first_launch_s", C.c_double), ("last_done_s", C.c_double),
        ("latency_p50_ms", C.c_double), ("latency_p99_ms", C.c_double), ("latency_max_ms", C.c_double),
    ]


FXG_MUX_MAX_THREADS = 32
FXG_MUX_STATIC_WORK_BYTES = 1 << 20


class fxg_vdif_mux_params(C.Structure):
    _fields_ = [
        ("inputframebytes", C.c_int32),
        ("nthreads", C.c_int32),
        ("bitspersample", C.c_int32),
        ("framespersecond", C.c_int32),
        ("threadmap", C.c_int32 * FXG_MUX_MAX_THREADS),
        ("ref_seconds", C.c_int64),
        ("ref_frame", C.c_int32),
        ("reserved", C.c_int32),
        ("start", C.c_int64),
    ]


FXG_INVALID_SUBINT = -99999999
FXG_FLAGS_PER_INT = 30


class fxg_segment(C.Structure):
    _fields_ = [
        ("scan", C.c_int32), ("scanseconds", C.c_int32), ("scanns", C.c_int32),
        ("validbytes", C.c_int32), ("readto", C.c_int32), ("sendbytes", C.c_int32),
        ("blockspersend", C.c_int32), ("bytespersamplenum", C.c_int32), ("bytespersampledenom", C.c_int32),
        ("bytesbetweenintegerns", C.c_int32), ("nsinc", C.c_int32), ("numchannels", C.c_int32),
        ("complex_sampling", C.c_int32), ("reserved", C.c_int32), ("sampletimens", C.c_double),
    ]


FXG_FRAMING_NONE = 0
FXG_FRAMING_VDIF = 1
FXG_FRAMING_MARK5B = 2


class fxg_databuffer(C.Structure):
    _fields_ = [
        ("numdatasegments", C.c_int32), ("bufferbytes", C.c_int32), ("readbytes", C.c_int32),
        ("keepreading", C.c_int32), ("lastvalidsegment", C.c_int32), ("intclockseconds", C.c_int32),
        ("framing", C.c_int32), ("framebytes", C.c_int32), ("payloadbytes", C.c_int32),
        ("framespersecond", C.c_int32), ("framegranularity", C.c_int32), ("reserved", C.c_int32),
        ("seg", C.POINTER(fxg_segment)),
    ]


FXG_ADVANCE_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int32)


class fxg_difx_meta(C.Structure):
    _fields_ = [
        ("mjd", C.c_int32), ("configindex", C.c_int32), ("sourceindex", C.c_int32), ("phasecentre", C.c_int32),
        ("seconds", C.c_double), ("fftsperintegration", C.c_double),
        ("baselinenumber", C.POINTER(C.c_int32)), ("uvw", C.POINTER(C.c_double)),
        ("telescopeindex", C.POINTER(C.c_int32)),
    ]


# exported symbols of libfxg.so (every function declared in include/fxg.h)
EXPORTED_SYMBOLS = [
    "fxg_create", "fxg_create_layout", "fxg_destroy", "fxg_strerror", "fxg_last_error", "fxg_get_layout",
    "fxg_data_bytes", "fxg_result_offsets", "fxg_result_pcal_offsets", "fxg_process_subint",
    "fxg_process_subint_device", "fxg_ring_length", "fxg_submit_subint", "fxg_wait",
    "fxg_get_spectrum", "fxg_get_data_weights", "fxg_unpack_device",
    "fxg_delay_interpolator", "fxg_set_timing", "fxg_get_timing", "fxg_autotune_report",
    "fxg_footprint", "fxg_input_samples", "fxg_device_bytes",
    "fxg_server_create", "fxg_server_destroy", "fxg_server_add_job", "fxg_server_push_subint",
    "fxg_server_wait_job", "fxg_server_get_stats", "fxg_server_job_error",
    "fxg_vdif_mux_workspace_bytes", "fxg_vdif_mux_device", "fxg_vdif_mux_result", "fxg_vdif_mux_device_static",
    "fxg_segment_params", "fxg_buffer_params", "fxg_control_params",
    "fxg_normalise_device", "fxg_difx_records", "fxg_set_kurtosis", "fxg_get_kurtosis",
    "fxg_sta_autocorr", "fxg_set_pcal", "fxg_get_pcal",
    "fxg_pcal_info",
]
