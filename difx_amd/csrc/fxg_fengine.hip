// F-engine of the FX hot path on CDNA4: unpack, delay / fringe rotation, LDS radix-8 FFT,
// sideband channel selection and fractional-sample correction, fused in one kernel.
//
// What one FFT block computes follows Mode::process (mpifxcorr/src/mode.cpp:634-1374)
// with Mk5Mode::unpack (mk5mode.cpp:94-174):
//   - the per-FFT delay chain in FP64 (mode.cpp:671-675, 743-757, 776-825),
//   - the rotator as S sub-phasors times NS step-phasors, FP64 phase wrapped to one turn,
//     FP32 sin/cos and complex products (order 1: mode.cpp:883-955; order 2 adds a
//     quadratic sub term raised to the step index, :956-1026; order 0 has none),
//   - the fractional-sample phasor from FP32 channel slopes (mode.cpp:303-324, 1030-1073),
//     carrying the post-F fringe phase for order 0 (mode.cpp:1040-1049),
//   - an unnormalised forward complex FFT of L = fftchannels points (mode.cpp:1174-1183),
//   - USB keeps channels [0, C), real LSB keeps [C, 2C) (mode.cpp:1185-1217); for order 0
//     the reference's real FFT with a conj-flip of bins 1..C (mode.cpp:1139-1156) gives the
//     same channels, conj(X[C-c]) = X[C+c] for real input.
// The FFT is a Stockham radix-8 (plus one radix-4 or radix-2 pass) exchange through LDS;
// each of the L/8 threads owns the 8 points t + s*L/8, so the first pass reads the decoded
// samples straight from registers and the last pass leaves channel t + s*L/8 in register s.
#include "fxg_fe_common.h"

// Per-FFT parameters of one chunk, one thread per (datastream, FFT): the FP64 delay chain
// and the data weight, computed once instead of by every thread of the F-engine.
__global__ void k_fftparams(FengineArgs args) {
  const int d = args.dslist[blockIdx.y];
  const int lt = blockIdx.x * blockDim.x + threadIdx.x;
  if (lt >= args.nffts_pad) return;
  const DevDs& ds = args.ds[d];
  const DevDsRun& run = args.run[d];
  FftDelay fd;
  fd.ok = 0; fd.weight = 0.0f; fd.n0 = 0; fd.fse = 0.0f; fd.a = 0; fd.b = 0; fd.intdelay = 0;
  fd.intwall = 0; fd.fracwall = 0;
  if (lt < args.nffts) {
    const int i = args.chunk0 + lt;
    fd = fft_delay(ds, run, args.fvalid + ds.fvalid_off, args.fzone + ds.fvalid_off, i, args.offsec, args.offns,
                   args.scan_start_sec, args.rotorder);
    args.weights[(long long)d * args.B + i] = fd.weight;
  }
  args.params[(long long)d * args.chunkcap + lt] = fd;
}

#ifndef FXG_F_MINWAVES
#define FXG_F_MINWAVES 1
#endif

// One FFT length: workgroup (x, y) = units [x*units_per_wg, ...) of datastream dslist[y].
template <int L, bool CPLX, int ORDER, int NTB>
__global__ void __launch_bounds__(NTB, FXG_F_MINWAVES) k_fengine(FengineArgs args) {
  extern __shared__ float2 fx_smem[];
  const int d = args.dslist[blockIdx.y];
  const int nunits = args.nffts_pad / 8;
  const int u0 = blockIdx.x * args.units_per_wg;
  const int u1 = min(nunits, u0 + args.units_per_wg);
  fe_range<L, CPLX, ORDER, NTB>(args, args.twiddle, d, u0, u1, fx_smem, true);
}

constexpr unsigned long long FILL64 = 0x1122334411223344ULL;   // blanker_mark5.c:36
constexpr uint32_t FILL32 = 0x11223344u;                          // blanker_mark5.c:35

// blanker_mark5's halving search over payload words [0, n) (blanker_mark5.c:43-111): probes
// start at n/2 and move by n/4, n/8, ..., 1, down after a probe whose word is (fill == want)
// and up otherwise; the lowest such probe, n - 1 if none, in bytes.
__device__ int zone_edge(const unsigned long long* w, int n, bool want) {
  int probe = n >> 1, hit = n - 1;
  for (int step = n >> 2; step > 0; step >>= 1) {
    const bool m = (w[probe] == FILL64) == want;
    hit = m ? probe : hit;
    probe += m ? -step : step;
  }
  return hit * 8;
}

// Zone 0 of a Mark5B payload (blanker_mark5.c:112-172) as the time samples [lo, hi) whose
// first byte (n * bits per sample) / 8 lies in the blanker's byte range
// (mark5_format_mark5b.c:241-243).  Returns the frame flag: 0, 1 or FXG_FRAME_PARTIAL.
__device__ int mark5_zone(const uint8_t* pay, const DevDs& ds, int2* zone) {
  const unsigned long long* w = reinterpret_cast<const unsigned long long*>(pay);
  const uint32_t* w32 = reinterpret_cast<const uint32_t*>(pay);
  const int e = min((ds.framebytes - ds.headerbytes) >> 3, 4096);   // 32768-byte zone
  const bool sok = w[0] != FILL64, eok = w[e - 1] != FILL64;
  if (sok && eok) return 1;
  if (!sok && !eok) return 0;
  long long sv = 0, ev = 1LL << 30;
  if (sok) {
    ev = zone_edge(w, e, true);
    if (w32[ev / 4 - 1] == FILL32) ev -= 4;      // a fill half word just before
  } else {
    sv = zone_edge(w, e, false);
    if (w32[sv / 4 - 2] == FILL32) sv += 4;      // the half word the reference tests
  }
  const long long bps = (long long)ds.nbits * (ds.iscomplex ? 2 : 1) * ds.nrecbands;
  const int lo = (int)min((8 * sv + bps - 1) / bps, (long long)ds.spf);
  const int hi = (int)min((8 * ev + bps - 1) / bps, (long long)ds.spf);
  *zone = make_int2(lo, hi);
  return hi <= lo ? 0 : (lo == 0 && hi == ds.spf) ? 1 : FXG_FRAME_PARTIAL;
}

// Setup for one subint: the delay interpolator of each datastream in FP64 on the device
// (Model::calculateDelayInterpolator, model.cpp:219-289, via Mode::setOffsets,
// mode.cpp:1502-1532) and the frame flags of each slot buffer.
__global__ void k_setup(SetupArgs a) {
  const int d = blockIdx.y;
  const DevDs& ds = a.ds[d];
  DevDsRun& run = a.run[d];
  // frame flags: VDIF word2 != 0 and invalid bit clear (format_vdif.c:7619-7660), fill pattern
  // at either end of the payload blanks the frame (blanker_mark5.c:272-310); Mark5B invalid
  // flag (mark5_format_mark5b.c:1723) and the fill-pattern zone of the payload.
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < run.nframes; f += gridDim.x * blockDim.x) {
    const uint8_t* fr = run.data + (long long)f * ds.framebytes;
    const uint8_t* pay = fr + ds.headerbytes;
    int ok;
    int2 zone = make_int2(0, 0);
    if (ds.format == FXG_FMT_VDIF) {
      const uint32_t w0 = *reinterpret_cast<const uint32_t*>(fr);
      const uint32_t w2 = *reinterpret_cast<const uint32_t*>(fr + 8);
      ok = (w2 != 0u) && (((w0 >> 31) & 1u) == 0u);
      const int pb = ds.framebytes - ds.headerbytes;
      const unsigned long long first = *reinterpret_cast<const unsigned long long*>(pay);
      const unsigned long long last = *reinterpret_cast<const unsigned long long*>(pay + pb - 8);
      if (first == FILL64 || last == FILL64) ok = 0;
    } else {
      ok = (fr[5] & 0x80) == 0 ? mark5_zone(pay, ds, &zone) : 0;
    }
    a.fvalid[ds.fvalid_off + f] = (uint8_t)ok;
    a.fzone[ds.fvalid_off + f] = zone;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    // Mode::setOffsets: scan mismatch or invalid data torch the subint (mode.cpp:1509-1516)
    if (run.datasec <= -99999999 || run.scan != a.offscan) run.valid = 0;
    const double offsettime = (double)a.offsec + (double)a.offns / 1000000000.0;
    const double timespan = (double)a.B * ds.C / ds.bw_mhz / 1e6;   // mode.cpp:1523-1524
    const double inc = a.model_inc_secs;
    int scansample = (int)(offsettime / inc);
    if (scansample == ds.model_nsamples && (offsettime - scansample * inc) < 1e-6) scansample--;
    if (scansample < 0 || scansample >= ds.model_nsamples) {
      run.valid = 0;                        // no interpolator: torch the subint (mode.cpp:1528-1531)
    } else {
      const double* c = a.model + (long long)(ds.model_off + scansample) * FXG_MAX_POLY;
      double smp[3];
      const double base = scansample * inc;
      const double dts[3] = {offsettime - base, offsettime + timespan / 2.0 - base,
                             offsettime + timespan - base};
      for (int k = 0; k < 3; ++k) {
        double tp = 1.0, acc = 0.0;
        for (int o = 0; o <= ds.model_order; ++o) { acc += tp * c[o]; tp *= dts[k]; }
        smp[k] = acc;
      }
      const double n = (double)a.B;
      run.interp[0] = (2.0 * smp[0] - 4.0 * smp[1] + 2.0 * smp[2]) / (n * n);
      run.interp[1] = (-3.0 * smp[0] + 4.0 * smp[1] - smp[2]) / n;
      run.interp[2] = smp[0] - 1000000.0 * ds.intclockseconds;
    }
  }
}

// Mk5Mode::unpack as a device call (bit-exact against the LUTs).
__global__ void k_unpack(UnpackArgs a) {
  const DevDs& ds = a.ds[a.d];
  const int nrec = ds.nrecbands;
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < a.nsamples; n += gridDim.x * blockDim.x) {
    const long long g = a.start + n;
    const int fr = (int)(g / ds.spf);
    const int within = (int)(g - (long long)fr * ds.spf);
    const int ok = (fr < a.nframes) && sample_ok(a.fv[fr], a.fz, fr, within);
    const long long fbase = (long long)fr * ds.framebytes + ds.headerbytes;
    if (ok) atomicAdd(a.good, 1);
    for (int b = 0; b < nrec; ++b) {
      if (ds.iscomplex) {
        float2 x = make_float2(0.f, 0.f);
        if (ok) {
          const uint8_t* p = a.data + fbase + ((long long)within * nrec + b) * 2;
          x = make_float2(lut8(p[0]), lut8(p[1]));
        }
        reinterpret_cast<float2*>(a.out)[(long long)b * a.nsamples + n] = x;
      } else {
        float x = 0.f;
        if (ok) {
          if (ds.nbits == 2) {
            const long long q = (long long)within * nrec + b;
            const uint8_t byte = a.data[fbase + (q >> 2)];
            x = lut2(ds.format, (byte >> (2 * (q & 3))) & 3u);
          } else {
            x = lut8(a.data[fbase + (long long)within * nrec + b]);
          }
        }
        a.out[(long long)b * a.nsamples + n] = x;
      }
    }
  }
}

// ---- host-side launchers -------------------------------------------------------------
int fxg_launch_setup(const SetupArgs& a, int maxframes, hipStream_t st) {
  dim3 grid((maxframes + 255) / 256 > 0 ? (maxframes + 255) / 256 : 1, a.nds);
  hipLaunchKernelGGL(k_setup, grid, dim3(256), 0, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int fxg_launch_unpack(const UnpackArgs& a, hipStream_t st) {
  int blocks = (a.nsamples + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_unpack, dim3(blocks), dim3(256), 0, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int L, bool CPLX, int ORDER, int NTB>
static int launch_k(const FengineArgs& a, dim3 grid, hipStream_t st) {
  static size_t raised = 0;
  const size_t lds = fe_lds_bytes<L, NTB>(a.phcap, ORDER);
  if (raise_lds(k_fengine<L, CPLX, ORDER, NTB>, lds, &raised)) return -1;
  hipLaunchKernelGGL((k_fengine<L, CPLX, ORDER, NTB>), grid, dim3(NTB), lds, st, a);
  return 0;
}

template <int L, bool CPLX, int NTB>
static int launch_f(const FengineArgs& a, int nds, hipStream_t st) {
  static_assert(L <= FXG_MAX_L, "FFT length beyond the fused F-engine");
  hipLaunchKernelGGL(k_fftparams, dim3((a.nffts_pad + 255) / 256, nds), dim3(256), 0, st, a);
  const int nunits = a.nffts_pad / 8;
  const int gx = (nunits + a.units_per_wg - 1) / a.units_per_wg;
  const dim3 grid(gx, nds);
  int rc;
  switch (a.rotorder) {
    case 1: rc = launch_k<L, CPLX, 1, NTB>(a, grid, st); break;
    case 2: rc = launch_k<L, CPLX, 2, NTB>(a, grid, st); break;
    case 0:
      if (CPLX) return -2;                   // post-F rotation of complex data (mode.cpp:1130-1133)
      rc = launch_k<L, false, 0, NTB>(a, grid, st);
      break;
    default: return -2;
  }
  if (rc) return rc;
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <bool CPLX, int NTB>
static int launch_len(const FengineArgs& a, int L, int nds, hipStream_t st) {
  switch (L) {
    case 256: return launch_f<256, CPLX, NTB>(a, nds, st);
    case 512: return launch_f<512, CPLX, NTB>(a, nds, st);
    case 1024: return launch_f<1024, CPLX, NTB>(a, nds, st);
    case 2048: return launch_f<2048, CPLX, NTB>(a, nds, st);
    case 4096: return launch_f<4096, CPLX, NTB>(a, nds, st);
    case 8192: return launch_f<8192, CPLX, NTB>(a, nds, st);
  }
  return -2;
}

int fxg_launch_fengine(const FengineArgs& a, int L, int iscomplex, int nds, int ntb, hipStream_t st) {
  if (ntb == 512) return iscomplex ? launch_len<true, 512>(a, L, nds, st) : launch_len<false, 512>(a, L, nds, st);
  if (ntb == 256) return iscomplex ? launch_len<true, 256>(a, L, nds, st) : launch_len<false, 256>(a, L, nds, st);
  if (ntb == 128) return iscomplex ? launch_len<true, 128>(a, L, nds, st) : launch_len<false, 128>(a, L, nds, st);
  return -2;
}

size_t fxg_fengine_lds(int L, int ntb, int phcap, int order) {
  if (ntb == 256) {
    switch (L) {
      case 256: return fe_lds_bytes<256, 256>(phcap, order);
      case 512: return fe_lds_bytes<512, 256>(phcap, order);
      case 1024: return fe_lds_bytes<1024, 256>(phcap, order);
      case 2048: return fe_lds_bytes<2048, 256>(phcap, order);
      case 4096: return fe_lds_bytes<4096, 256>(phcap, order);
      case 8192: return fe_lds_bytes<8192, 256>(phcap, order);
    }
  } else if (ntb == 128) {
    switch (L) {
      case 256: return fe_lds_bytes<256, 128>(phcap, order);
      case 512: return fe_lds_bytes<512, 128>(phcap, order);
      case 1024: return fe_lds_bytes<1024, 128>(phcap, order);
      case 2048: return fe_lds_bytes<2048, 128>(phcap, order);
      case 4096: return fe_lds_bytes<4096, 128>(phcap, order);
      case 8192: return fe_lds_bytes<8192, 128>(phcap, order);
    }
  } else if (ntb == 512) {
    switch (L) {
      case 256: return fe_lds_bytes<256, 512>(phcap, order);
      case 512: return fe_lds_bytes<512, 512>(phcap, order);
      case 1024: return fe_lds_bytes<1024, 512>(phcap, order);
      case 2048: return fe_lds_bytes<2048, 512>(phcap, order);
      case 4096: return fe_lds_bytes<4096, 512>(phcap, order);
      case 8192: return fe_lds_bytes<8192, 512>(phcap, order);
    }
  }
  return 0;
}

// Parameters of every FFT of the chunk for the datastreams of dslist (one launch per class
// list; the persistent launch calls it once per class before its own launch).
int fxg_launch_fftparams(const FengineArgs& a, int nds, hipStream_t st) {
  hipLaunchKernelGGL(k_fftparams, dim3((a.nffts_pad + 255) / 256, nds), dim3(256), 0, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

