// Phase-cal extraction on the device.  Mode::process hands each recorded band's unpacked
// samples (real, or complex for complex sampling) of every valid FFT to its PCal extractor,
// numbered from the scan start (mpifxcorr/src/mode.cpp:758-772); every extractor PCal::getNew
// picks (pcal.cpp:191-274) computes, per tone, sum x[m] exp(-2 pi i f m) over those samples at
// a rational frequency f = num / den cycles per sample that depends on the extractor (real:
// (offset + k interval) / fs; complex: the DFT bin the implicit-shift or shifting complex
// extractor reads, reversed for double sideband).  Here the samples are folded by their sample
// number modulo the tone period M (a multiple of every den; complex samples fold their real
// and imaginary fields separately), and k_pcal_tones forms each tone from the M bins.
//
// Work split: a workgroup takes a group of consecutive recorded bands of one datastream that
// share M, and a contiguous share of the chunk's FFTs.  Thread t takes time samples t, t + 256,
// ... of each FFT, issues the loads of eight of them before decoding any, and decodes all the
// group's bands of a time sample from one 32-bit load (the bands of a sample are adjacent
// fields of the payload).  When 256 is a multiple of M a thread's
// bin stays fixed through an FFT, so it sums in registers and adds to the LDS bins only when its
// bin changes; otherwise every sample is an LDS atomic add.  Each workgroup adds its LDS bins to
// its own slice of a partial array (no global atomics, a fixed summation order), and a reduce
// kernel sums the slices into the bands' bins after the subint's last chunk.
#include "fxg_fe_common.h"

namespace {

constexpr int PCAL_NT = 256;
constexpr int PCAL_UNR = 8;            // time samples per thread in flight (loads issued together)

// NBITS: sample width (2 or 8).  WHOLE: the group is every band of a sample word that holds
// exactly one time sample (nrec NBITS = 32), so field k of the word is band k.  REG: 256 is a
// multiple of M, so a thread's bin is fixed through an FFT and it sums in registers.
template <int NBITS, bool WHOLE, bool REG>
__global__ void __launch_bounds__(PCAL_NT) k_pcal(PcalArgs a) {
  constexpr int LGF = NBITS == 2 ? 4 : 2;          // log2 fields per 32-bit word
  constexpr unsigned FMASK = (1u << NBITS) - 1u;
  constexpr unsigned FSUB = (1u << LGF) - 1u;
  extern __shared__ float sm[];
  const PcalGroup g = a.groups[blockIdx.x];
  const DevDs& ds = a.ds[g.ds];
  const DevDsRun& run = a.run[g.ds];
  const int M = g.M;
  const int gsize = g.nb * M;
  float2* lutp = reinterpret_cast<float2*>(sm);    // 2-bit: levels of two fields (4-bit index)
  float* lut = sm + 32;                            // sample levels: 256 (8-bit) or 4 (2-bit)
  float* sb = sm + 288;                            // [nb][M]
  const int tid = threadIdx.x;
  for (int j = tid; j < gsize; j += PCAL_NT) sb[j] = 0.f;
  if (NBITS == 2) {
    if (tid < 4) lut[tid] = lut2(ds.format, tid);
    if (tid < 16) lutp[tid] = make_float2(lut2(ds.format, tid & 3), lut2(ds.format, tid >> 2));
  } else {
    lut[tid] = lut8(tid);
  }
  __syncthreads();
  const int per = (a.nffts + a.nsplit - 1) / a.nsplit;
  const int lt0 = blockIdx.y * per;
  const int lt1 = min(lt0 + per, a.nffts);
  if (run.valid) {
    // sample number of the buffer start since the scan start (mode.cpp:760)
    const long long s0 = (long long)(((double)run.datasec * 1e9 + (double)run.datans) / (ds.sampletime_us * 1e3) + 0.5);
    const uint8_t* __restrict__ fv = a.fvalid + ds.fvalid_off;
    const int2* __restrict__ fz = a.fzone + ds.fvalid_off;
    const uint8_t* __restrict__ data = run.data;
    const int L = ds.L, spf = ds.spf, nrec = g.fps, nb = g.nb;   // nrec: fields per time sample
    const int nsamp = run.nframes * spf;
    const int step = PCAL_NT % M;                  // bin advance per time step (0 when REG)
    float acc[PCAL_MAXNB];
#pragma unroll
    for (int k = 0; k < PCAL_MAXNB; ++k) acc[k] = 0.f;
    int accbin = -1;
    for (int lt = lt0; lt < lt1; ++lt) {
      const FftDelay fd = a.params[(long long)g.ds * a.chunkcap + lt];
      if (!fd.ok) continue;                        // uniform over the workgroup
      int t = fd.n0 + tid;                         // time sample of the buffer (n0 >= 0)
      int fr = t / spf;
      int within = t - fr * spf;
      // the sample number modulo M, kept in [0, M) also for samples before the scan start
      // (a buffer that starts before it has a negative datasec)
      int bin = (int)((s0 + t) % M);
      if (bin < 0) bin += M;
      if (REG && bin != accbin) {
        if (accbin >= 0) {
#pragma unroll
          for (int k = 0; k < PCAL_MAXNB; ++k)
            if (k < nb) { atomicAdd(&sb[k * M + accbin], acc[k]); acc[k] = 0.f; }
        }
        accbin = bin;
      }
      for (int nbase = 0; nbase < L; nbase += PCAL_NT * PCAL_UNR) {
        uint32_t w[PCAL_UNR];
        unsigned q0[PCAL_UNR];
        const uint8_t* pay[PCAL_UNR];
        bool ok[PCAL_UNR];
        int binu[PCAL_UNR];
        // addresses and loads of all the thread's time samples of this pass first
#pragma unroll
        for (int u = 0; u < PCAL_UNR; ++u) {
          const bool in = nbase + u * PCAL_NT + tid < L && t < nsamp;
          pay[u] = data + (long long)fr * ds.framebytes + ds.headerbytes;
          q0[u] = (unsigned)(within * nrec + g.band0);
          w[u] = in ? *reinterpret_cast<const uint32_t*>(pay[u] + 4u * (q0[u] >> LGF)) : 0u;
          ok[u] = in ? sample_ok(fv[fr], fz, fr, within) : false;
          binu[u] = bin;
          t += PCAL_NT;
          within += PCAL_NT;
          while (within >= spf) { within -= spf; ++fr; }
          if (!REG) { bin += step; if (bin >= M) bin -= M; }
        }
#pragma unroll
        for (int u = 0; u < PCAL_UNR; ++u) {
          if (!ok[u]) continue;
          if (WHOLE && NBITS == 2) {                 // two bands per table read
#pragma unroll
            for (int k = 0; k < 16; k += 2) {
              const float2 x = lutp[(w[u] >> (2 * k)) & 15u];
              if (REG) {
                acc[k] += x.x;
                acc[k + 1] += x.y;
              } else {
                atomicAdd(&sb[k * M + binu[u]], x.x);
                atomicAdd(&sb[(k + 1) * M + binu[u]], x.y);
              }
            }
          } else if (WHOLE) {
#pragma unroll
            for (int k = 0; k < (32 / NBITS); ++k) {
              const float x = lut[(w[u] >> (NBITS * k)) & FMASK];
              if (REG) acc[k] += x;
              else atomicAdd(&sb[k * M + binu[u]], x);
            }
          } else {
            unsigned wi = q0[u] >> LGF;
            uint32_t ww = w[u];
#pragma unroll
            for (int k = 0; k < PCAL_MAXNB; ++k) {
              if (k < nb) {
                const unsigned q = q0[u] + k;
                if ((q >> LGF) != wi) {            // the group's fields run into the next word
                  wi = q >> LGF;
                  ww = *reinterpret_cast<const uint32_t*>(pay[u] + 4u * wi);
                }
                const float x = lut[(ww >> (NBITS * (q & FSUB))) & FMASK];
                if (REG) acc[k] += x;
                else atomicAdd(&sb[k * M + binu[u]], x);
              }
            }
          }
        }
      }
    }
    if (REG && accbin >= 0) {
#pragma unroll
      for (int k = 0; k < PCAL_MAXNB; ++k)
        if (k < nb) atomicAdd(&sb[k * M + accbin], acc[k]);
    }
  }
  __syncthreads();
  float* part = a.part + g.part_off + (long long)blockIdx.y * gsize;
  for (int j = tid; j < gsize; j += PCAL_NT) part[j] += sb[j];
}

// bins of every band of group blockIdx.x = the sum of the group's nsplit partial slices; one
// 64-lane workgroup per 64 bins, so that many workgroups share the slices' reads
__global__ void __launch_bounds__(64) k_pcal_reduce(PcalArgs a) {
  const PcalGroup g = a.groups[blockIdx.x];
  const int gsize = g.nb * g.M;
  for (int j = blockIdx.y * 64 + threadIdx.x; j < gsize; j += gridDim.y * 64) {
    const float* src = a.part + g.part_off + j;
    double s = 0.0;
#pragma unroll 8
    for (int p = 0; p < a.nsplit; ++p) s += src[(long long)p * gsize];
    a.bins[g.bins_off + j] = (float)s;
  }
}

// One tone per workgroup: the M-point transform of its band's folded bins at the tone's
// rational frequency, in FP64 with the phase reduced exactly in integers.
__global__ void __launch_bounds__(256) k_pcal_tones(const PcalToneDev* tt, const float* bins, float2* tones,
                                                    float2* results) {
  const PcalToneDev t = tt[blockIdx.x];
  __shared__ double sre[256], sim[256];
  double re = 0.0, im = 0.0;
  const float* br = bins + t.bins_off;
  const float* bi = br + t.M;
  for (int j = threadIdx.x; j < t.M; j += 256) {
    const int r = (int)(((long long)t.num * j) % t.den);
    double sn, cs;
    sincospi(-2.0 * (double)r / (double)t.den, &sn, &cs);
    const double xr = br[j], xi = t.cplx ? (double)bi[j] : 0.0;
    re += xr * cs - xi * sn;
    im += xr * sn + xi * cs;
  }
  sre[threadIdx.x] = re;
  sim[threadIdx.x] = im;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) { sre[threadIdx.x] += sre[threadIdx.x + o]; sim[threadIdx.x] += sim[threadIdx.x + o]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const float2 v = make_float2((float)sre[0], (float)sim[0]);
    tones[t.out] = v;
    if (t.res >= 0) {                              // Core::copyPCalTones adds into the slot results
      float2 r = results[t.res];
      r.x += v.x;
      r.y += v.y;
      results[t.res] = r;
    }
  }
}

}  // namespace

size_t fxg_pcal_lds(int max_gsize) { return sizeof(float) * ((size_t)max_gsize + 288); }

template <int NBITS, bool WHOLE, bool REG>
int launch_pcal_form(const PcalArgs& a, int ngroups, size_t lds, hipStream_t st) {
  static size_t raised = 0;
  if (raise_lds(k_pcal<NBITS, WHOLE, REG>, lds, &raised)) return -1;
  hipLaunchKernelGGL((k_pcal<NBITS, WHOLE, REG>), dim3(ngroups, a.nsplit), dim3(PCAL_NT), lds, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// groups a.groups[0, ngroups) all of kernel form `form` (PCAL_FORM_* bits)
int fxg_launch_pcal(const PcalArgs& a, int form, int ngroups, int max_gsize, hipStream_t st) {
  if (ngroups <= 0 || a.nffts <= 0) return 0;
  const size_t lds = fxg_pcal_lds(max_gsize);
  switch (form) {
    case 0: return launch_pcal_form<2, false, false>(a, ngroups, lds, st);
    case PCAL_FORM_REG: return launch_pcal_form<2, false, true>(a, ngroups, lds, st);
    case PCAL_FORM_WHOLE: return launch_pcal_form<2, true, false>(a, ngroups, lds, st);
    case PCAL_FORM_WHOLE | PCAL_FORM_REG: return launch_pcal_form<2, true, true>(a, ngroups, lds, st);
    case PCAL_FORM_8BIT: return launch_pcal_form<8, false, false>(a, ngroups, lds, st);
    case PCAL_FORM_8BIT | PCAL_FORM_REG: return launch_pcal_form<8, false, true>(a, ngroups, lds, st);
    case PCAL_FORM_8BIT | PCAL_FORM_WHOLE: return launch_pcal_form<8, true, false>(a, ngroups, lds, st);
    case PCAL_FORM_8BIT | PCAL_FORM_WHOLE | PCAL_FORM_REG: return launch_pcal_form<8, true, true>(a, ngroups, lds, st);
    default: return -1;
  }
}

int fxg_launch_pcal_reduce(const PcalArgs& a, int ngroups, int max_gsize, hipStream_t st) {
  if (ngroups <= 0) return 0;
  const int gy = std::max(1, std::min(512, (max_gsize + 63) / 64));
  hipLaunchKernelGGL(k_pcal_reduce, dim3(ngroups, gy), dim3(64), 0, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int fxg_launch_pcal_tones(const PcalToneDev* t, int n, const float* bins, float2* tones, float2* results,
                          hipStream_t st) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_pcal_tones, dim3(n), dim3(256), 0, st, t, bins, tones, results);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
