// F-engine for FFT lengths beyond one workgroup's LDS: L = R * 2048 with R = 8 or 16
// (nchan 8192 / 16384 real, 16384 / 32768 complex).  What one FFT computes is what the fused
// F-engine computes (Mode::process, mode.cpp:634-1374: decode, rotator, L-point FFT, channel
// selection, fractional-sample correction); only the FFT is organised differently, as a
// four-step transform with n = R*m + r and k = k1 + 2048*k2:
//   1. for each r: the 2048-point FFT over m of the rotated samples x[R*m + r] (the LDS
//      Stockham FFT of the fused engine, 256 threads x 8 points), giving Y_r[k1];
//   2. Z_r[k1] = Y_r[k1] * exp(-2 pi i r k1 / L), staged in a per-workgroup scratch row
//      (L * 8 bytes, resident in L2 / MALL);
//   3. for each k1: X[k1 + 2048*k2] = sum_r Z_r[k1] exp(-2 pi i r k2 / R), an R-point DFT in
//      registers.
// Each thread ends with channels k1 = t + 256*s, consecutive across a wavefront, so the stores
// stay as coalesced as the fused engine's.
#include "fxg_fe_common.h"

namespace {

constexpr int BM = 2048;          // sub-FFT length
constexpr int BNT = BM / 8;       // threads of the workgroup

// 16-point forward DFT as 4 x 4: A[n2][k1] = DFT4 over n1 of x[4 n1 + n2], times
// exp(-2 pi i n2 k1 / 16), then DFT4 over n2 gives X[k1 + 4 k2].
__device__ __forceinline__ void dft16(float2* x) {
  // exp(-2 pi i m / 16) for m = 1, 2, 3, 4, 6, 9
  const float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f;
  const float c2 = 0.70710678118654752440f;
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) dft4(x[n2], x[4 + n2], x[8 + n2], x[12 + n2]);
  // x[4 k1 + n2] now holds A[n2][k1]; twiddles W16^(n2 k1)
  x[4 * 1 + 1] = cmul(x[4 * 1 + 1], make_float2(c1, -s1));    // W^1
  x[4 * 2 + 1] = cmul(x[4 * 2 + 1], make_float2(c2, -c2));    // W^2
  x[4 * 3 + 1] = cmul(x[4 * 3 + 1], make_float2(s1, -c1));    // W^3
  x[4 * 1 + 2] = cmul(x[4 * 1 + 2], make_float2(c2, -c2));    // W^2
  x[4 * 2 + 2] = mul_negi(x[4 * 2 + 2]);                      // W^4
  x[4 * 3 + 2] = cmul(x[4 * 3 + 2], make_float2(-c2, -c2));   // W^6
  x[4 * 1 + 3] = cmul(x[4 * 1 + 3], make_float2(s1, -c1));    // W^3
  x[4 * 2 + 3] = cmul(x[4 * 2 + 3], make_float2(-c2, -c2));   // W^6
  x[4 * 3 + 3] = cmul(x[4 * 3 + 3], make_float2(-c1, s1));    // W^9
  float2 y[16];
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    float2 a0 = x[4 * k1], a1 = x[4 * k1 + 1], a2 = x[4 * k1 + 2], a3 = x[4 * k1 + 3];
    dft4(a0, a1, a2, a3);
    y[k1] = a0; y[k1 + 4] = a1; y[k1 + 8] = a2; y[k1 + 12] = a3;
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) x[k] = y[k];
}

template <int R>
__device__ __forceinline__ void dftR(float2* x) {
  if (R == 8) dft8(x);
  else dft16(x);
}

}  // namespace

// One workgroup processes whole FFTs of datastream dslist[blockIdx.y], FFT lt = blockIdx.x,
// blockIdx.x + gridDim.x, ...; its scratch row is scratch[(y * gridDim.x + x) * L].
template <int L, bool CPLX, int ORDER>
__global__ void __launch_bounds__(BNT, 1) k_fbig(FengineArgs args, float2* __restrict__ scratch,
                                                  const float2* __restrict__ twL) {
  constexpr int R = L / BM;
  constexpr int KEEP = CPLX ? R : R / 2;            // k2 values kept per k1
  using P = Plan<BM>;
  extern __shared__ float2 fx_smem[];
  float2* sbuf = fx_smem;
  float2* stw = sbuf + (BM + BM / 8);
  float2* srot = stw + P::TWN;
  float2* sfrac = srot + args.phcap;
  float* squad = reinterpret_cast<float*>(sfrac + args.phcap);
  const int t = threadIdx.x;
  const int d = args.dslist[blockIdx.y];
  float2* __restrict__ scr = scratch + ((long long)blockIdx.y * gridDim.x + blockIdx.x) * L;

  for (int m = t; m < P::TWN; m += BNT) stw[m] = args.twiddle[m];

  const DevDs* __restrict__ dsp = args.ds + d;
  DsConst k;
  k.C = dsp->C; k.S = dsp->S; k.NS = dsp->NS; k.NFS = dsp->NFS; k.nrec = dsp->nrecbands;
  k.nrecfreqs = dsp->nrecfreqs; k.fmt = dsp->format; k.spf = dsp->spf;
  k.framebytes = dsp->framebytes; k.headerbytes = dsp->headerbytes; k.nbits = dsp->nbits;
  k.fractionallo = dsp->fractionallo; k.sampletime_us = dsp->sampletime_us; k.dsb = dsp->dsb;
  const long long spec_off32 = dsp->spec_off;
  const uint8_t* __restrict__ fv = args.fvalid + dsp->fvalid_off;
  const float* __restrict__ cf = args.chanfreq + dsp->chanfreq_off;
  const uint8_t* __restrict__ data = args.run[d].data;
  const int nframes = args.run[d].nframes;
  const FftDelay* __restrict__ params = args.params + (long long)d * args.chunkcap;
  const int C = k.C, S = k.S;
  const bool twobit = (k.nbits == 2);
  const int out_mode = args.out_mode;
  const int cap = args.chunkcap;

  for (int lt = blockIdx.x; lt < args.nffts_pad; lt += gridDim.x) {
    const FftDelay fd = params[lt];
    const bool fok_all = fd.ok != 0;               // flagged FFTs: zero spectra (mode.cpp:656-669)
    for (int f = 0; f < k.nrecfreqs; ++f) {
      const FreqConst rf = load_freq(dsp->rf[f]);
      if (rf.nbands == 0) continue;
      __syncthreads();                            // previous users of srot/sfrac are done
      if (ORDER != 0) build_rotator<ORDER>(srot, squad, k, rf, fd, L, t, BNT);
      build_frac<ORDER>(sfrac, k, rf, CPLX, cf, fd, t, BNT);
      __syncthreads();
      for (int kb = 0; kb < rf.nbands; ++kb) {
        const int band = dsp->rf[f].band[kb];
        for (int r = 0; r < R; ++r) {
          // samples n = R*(t + 256 s) + r of the FFT window
          float2 v[1][8];
          const int nb0 = R * t + r;
          const float2 sub = ORDER != 0 ? srot[nb0 & (S - 1)] : make_float2(1.0f, 0.0f);
          float2 qpow = make_float2(1.0f, 0.0f), qstep = make_float2(1.0f, 0.0f);
          if (ORDER == 2) {                      // quadratic sub term^j, j = n/S (mode.cpp:1017-1024)
            const float th = squad[nb0 & (S - 1)];
            float sn, cs;
            sincosf((float)(nb0 / S) * th, &sn, &cs);
            qpow = make_float2(cs, sn);
            sincosf((float)(R * BNT / S) * th, &sn, &cs);
            qstep = make_float2(cs, sn);
          }
          const unsigned g0 = (unsigned)fd.n0 + (unsigned)nb0;
          unsigned fr = g0 / (unsigned)k.spf;
          unsigned within = g0 - fr * (unsigned)k.spf;
#pragma unroll
          for (int s = 0; s < 8; ++s) {
            if (s > 0) {
              within += R * BNT;
              while (within >= (unsigned)k.spf) { within -= k.spf; ++fr; }
            }
            const int n = nb0 + s * R * BNT;
            const int st = n / S;
            float2 rot;
            if (ORDER == 1) {
              rot = st == 0 ? sub : cmul(sub, srot[S + st]);
            } else if (ORDER == 2) {
              const float2 pq = cmul(sub, qpow);
              rot = st == 0 ? pq : cmul(pq, srot[S + st]);
              qpow = cmul(qpow, qstep);
            } else {
              rot = sub;
            }
            const bool ok = fok_all && ((int)fr < nframes) && fv[fr];
            const unsigned fbase = fr * (unsigned)k.framebytes + (unsigned)k.headerbytes;
            if (!CPLX) {
              float x = 0.0f;
              if (ok) {
                if (twobit) {
                  const unsigned q = within * (unsigned)k.nrec + (unsigned)band;
                  const uint32_t w = *reinterpret_cast<const uint32_t*>(data + fbase + ((q >> 4) << 2));
                  x = lut2(k.fmt, (w >> (2 * (q & 15u))) & 3u);
                } else {
                  x = lut8(data[fbase + within * (unsigned)k.nrec + band]);
                }
              }
              v[0][s] = ORDER == 0 ? make_float2(x, 0.0f) : make_float2(x * rot.x, x * rot.y);
            } else {
              float2 x = make_float2(0.f, 0.f);
              if (ok) {
                const uint8_t* pp = data + fbase + (within * (unsigned)k.nrec + band) * 2u;
                x = make_float2(lut8(pp[0]), lut8(pp[1]));
              }
              v[0][s] = cmul(rot, x);
              if (rf.lsb) v[0][s].y = -v[0][s].y;   // conj input (see fe_range)
            }
          }
          __syncthreads();                        // sbuf free
          fft_stockham<BM, 1, BNT>(v, sbuf, stw, t);
#pragma unroll
          for (int s = 0; s < 8; ++s) {           // Z_r[k1] = Y_r[k1] W_L^(r k1)
            const int k1 = t + s * BNT;
            scr[r * BM + k1] = cmul(v[0][s], twL[r * k1]);
          }
        }
        __syncthreads();                          // every Z_r of this band is in the scratch row
        // output pointers of this band
        float2* __restrict__ o32 = nullptr;
        uint32_t* __restrict__ o16 = nullptr;
        const float* __restrict__ scp = nullptr;
        if (out_mode == 0) {
          o32 = args.spec32 + spec_off32 + ((long long)band * cap + lt) * C;
        } else {
          const int xg = dsp->band_xfreq[band];
          const int row = dsp->band_xrow[band];
          o16 = args.spec16 + args.xgroups[xg].spec_off + (long long)row * args.xgroups[xg].rowstride + lt * 8;
          scp = args.xscale + args.xgroups[xg].scale_off + (long long)row * C;
        }
#pragma unroll 1
        for (int s = 0; s < 8; ++s) {
          const int k1 = t + s * BNT;
          float2 z[R];
#pragma unroll
          for (int r = 0; r < R; ++r) z[r] = scr[r * BM + k1];
          dftR<R>(z);                             // z[k2] = X[k1 + 2048 k2]
#pragma unroll
          for (int j = 0; j < KEEP; ++j) {
            // channel c = k1 + 2048 j; real LSB keeps X[C + c]; double sideband swaps halves
            int k2 = j;
            if (!CPLX && rf.lsb) k2 = j + R / 2;
            if (CPLX && k.dsb) k2 = (j + R / 2) % R;
            const int c = k1 + BM * j;
            float2 X = z[0];
#pragma unroll
            for (int q = 1; q < R; ++q) X = (k2 == q) ? z[q] : X;
            if (CPLX && rf.lsb && !k.dsb && c == 0) X.y = -X.y;   // DC is not flipped
            const float2 frq = cmul(sfrac[c & (S - 1)], sfrac[S + c / S]);
            float2 F = cmul(X, frq);
            if (!fok_all) F = make_float2(0.0f, 0.0f);
            if (out_mode == 0) {
              o32[c] = F;
            } else {
              const float sc = scp[c];
              o16[(c >> 3) * cap * 8 + (c & 7)] = pack16(make_float2(F.x * sc, F.y * sc), out_mode);
              for (int zz = 0; zz < dsp->nzout; ++zz) {          // zoom windows of this band
                if (dsp->zout_band[zz] != band) continue;
                const int cz = c - dsp->zout_c0[zz];
                if (cz < 0 || cz >= dsp->zout_n[zz]) continue;
                const int xg = dsp->zout_group[zz], row = dsp->zout_row[zz];
                uint32_t* __restrict__ oz = args.spec16 + args.xgroups[xg].spec_off +
                                            (long long)row * args.xgroups[xg].rowstride + lt * 8;
                const float scz = args.xscale[args.xgroups[xg].scale_off + (long long)row * dsp->zout_n[zz] + cz];
                oz[(cz >> 3) * cap * 8 + (cz & 7)] = pack16(make_float2(F.x * scz, F.y * scz), out_mode);
              }
            }
          }
        }
        __syncthreads();                          // scratch row free for the next band
      }
    }
  }
}

namespace {

template <int L, bool CPLX, int ORDER>
int launch_big(const FengineArgs& a, dim3 grid, float2* scratch, const float2* twL, hipStream_t st) {
  static size_t raised = 0;
  const size_t lds = sizeof(float2) * ((size_t)(BM + BM / 8) + Plan<BM>::TWN + 2 * (size_t)a.phcap) +
                     sizeof(float) * (size_t)a.phcap;
  if (raise_lds(k_fbig<L, CPLX, ORDER>, lds, &raised)) return -1;
  hipLaunchKernelGGL((k_fbig<L, CPLX, ORDER>), grid, dim3(BNT), lds, st, a, scratch, twL);
  return 0;
}

template <int L, bool CPLX>
int launch_big_o(const FengineArgs& a, dim3 grid, float2* scratch, const float2* twL, hipStream_t st) {
  switch (a.rotorder) {
    case 1: return launch_big<L, CPLX, 1>(a, grid, scratch, twL, st);
    case 2: return launch_big<L, CPLX, 2>(a, grid, scratch, twL, st);
    case 0: return CPLX ? -2 : launch_big<L, false, 0>(a, grid, scratch, twL, st);
  }
  return -2;
}

}  // namespace

// FFT lengths 16384 and 32768: the per-FFT parameters, then the four-step F-engine with
// `wgx` workgroups per datastream (the scratch holds wgx * nds rows of L cf32).
int fxg_launch_fbig(const FengineArgs& a, int L, int iscomplex, int nds, int wgx, float2* scratch,
                    const float2* twL, hipStream_t st) {
  if (fxg_launch_fftparams(a, nds, st)) return -1;
  const dim3 grid(wgx, nds);
  int rc = -2;
  if (L == 16384) rc = iscomplex ? launch_big_o<16384, true>(a, grid, scratch, twL, st)
                                 : launch_big_o<16384, false>(a, grid, scratch, twL, st);
  else if (L == 32768) rc = iscomplex ? launch_big_o<32768, true>(a, grid, scratch, twL, st)
                                      : launch_big_o<32768, false>(a, grid, scratch, twL, st);
  if (rc) return rc;
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
