// Device code shared by the F-engine translation units (fxg_fengine.hip: one launch per FFT
// length; fxg_fpersist.hip: every length in one persistent launch).
#pragma once
#include "fxg_internal.h"
#include "../../include/fxg.h"
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

namespace {

// Read-only tables through the constant address space: with a uniform address the compiler
// loads them with scalar loads.  A vector-memory load between the bands would make its wait
// also wait for every store issued before it (the vector-memory counter counts both), so each
// band would stall on the previous band's spectra stores.
#define FXG_CONST __attribute__((address_space(4)))
typedef const FXG_CONST DevDs* CDevDs;
typedef const FXG_CONST DevXGroup* CDevXGroup;
// Sample data and frame flags as global-address-space pointers: their loads are then vector
// global loads (a flat load would also count against the LDS wait counter).
typedef const __attribute__((address_space(1))) uint8_t* GBytes;
typedef const __attribute__((address_space(1))) uint32_t* GWords;

constexpr double TWO_PI_D = 6.283185307179586476925286766559;   // architecture.h.in:430
constexpr float HIMAG = 3.3359f;                                 // mark5_stream.h:65
constexpr int FLAGS_PER_INT = 30;                                // mpifxcorr.h:29

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 mul_negi(float2 a) { return make_float2(a.y, -a.x); }   // a * (-i)

// 2-bit LUTs: VDIF {-H,-1,+1,+H}[field] (format_vdif.c:70); Mark5B {-H,+1,-1,+H}[field]
// where field = sign bit + 2*magnitude bit (mark5_format_mark5b.c:52-80).
__device__ __forceinline__ float lut2(int fmt, unsigned field) {
  const float v0 = -HIMAG, v3 = HIMAG;
  float v1 = (fmt == FXG_FMT_MARK5B) ? 1.0f : -1.0f;
  float v2 = -v1;
  return field == 0 ? v0 : field == 1 ? v1 : field == 2 ? v2 : v3;
}
// 8-bit LUT (b - 128)/3.3 computed in double and rounded to float (format_vdif.c:120).
__device__ __forceinline__ float lut8(unsigned b) { return (float)(((double)b - 128.0) / 3.3); }

// In-register DFTs (forward, sign -1), decimation in frequency.
__device__ __forceinline__ void dft2(float2& a, float2& b) {
  float2 t = a; a = cadd(t, b); b = csub(t, b);
}
__device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
  float2 b0 = cadd(a0, a2), b2 = csub(a0, a2);
  float2 b1 = cadd(a1, a3), b3 = mul_negi(csub(a1, a3));
  a0 = cadd(b0, b1); a2 = csub(b0, b1);
  a1 = cadd(b2, b3); a3 = csub(b2, b3);
}
__device__ __forceinline__ void dft8(float2* v) {
  const float c = 0.70710678118654752440f;
  float2 u0 = cadd(v[0], v[4]), u4 = csub(v[0], v[4]);
  float2 u1 = cadd(v[1], v[5]), u5 = csub(v[1], v[5]);
  float2 u2 = cadd(v[2], v[6]), u6 = csub(v[2], v[6]);
  float2 u3 = cadd(v[3], v[7]), u7 = csub(v[3], v[7]);
  u5 = make_float2(c * (u5.x + u5.y), c * (u5.y - u5.x));      // * W8^1
  u6 = mul_negi(u6);                                           // * W8^2
  u7 = make_float2(c * (u7.y - u7.x), -c * (u7.x + u7.y));     // * W8^3
  float2 w0 = cadd(u0, u2), w2 = csub(u0, u2);
  float2 w1 = cadd(u1, u3), w3 = mul_negi(csub(u1, u3));
  float2 w4 = cadd(u4, u6), w6 = csub(u4, u6);
  float2 w5 = cadd(u5, u7), w7 = mul_negi(csub(u5, u7));
  v[0] = cadd(w0, w1); v[4] = csub(w0, w1);
  v[2] = cadd(w2, w3); v[6] = csub(w2, w3);
  v[1] = cadd(w4, w5); v[5] = csub(w4, w5);
  v[3] = cadd(w6, w7); v[7] = csub(w6, w7);
}

// LDS exchange layout.  Each exchange of the Stockham FFT has its own layout, chosen by the
// span Ns of the pass that writes it, so that both sides of the exchange address the buffer as
// one per-thread base plus compile-time offsets (one address register per side instead of one
// per entry) and neither side has bank conflicts for 8-byte entries:
//   Ns = 1:  a thread writes 8 consecutive entries 8t..8t+7; one pad entry per 2^XPADB keeps
//            the lanes of a bank group apart (XPADB 4 for 32 banks, 5 for 64);
//   Ns = 8:  a thread writes 64c + k + 8r (c = t/8, k = t%8); eight pad entries per 64 place the
//            two (or four) c of a bank group in different halves (quarters) of the banks;
//   Ns >= 64: consecutive threads write consecutive entries; no padding.
// The reads at position t + s*NT follow the same layout.
#ifndef FXG_XPADB
#define FXG_XPADB 4
#endif
__host__ __device__ constexpr int xlay(int Ns, int e) {
  return Ns == 1 ? e + (e >> FXG_XPADB) : Ns == 8 ? e + 8 * (e >> 6) : e;
}
__host__ __device__ constexpr int sbuf_entries(int L) { return L + L / 8; }
// Stride in entries between the reads t + s*NT and t + (s+1)*NT of the layout of span Nw, or 0
// when the layout is not affine in s there (Nw = 8 with NT < 64).
__host__ __device__ constexpr int xread_stride(int Nw, int NT) {
  return Nw == 1 ? NT + (NT >> FXG_XPADB) : Nw == 8 ? (NT % 64 == 0 ? NT + NT / 8 : 0) : NT;
}
// Entry of position t + s*NT in the layout of span Nw.
template <int NT>
__device__ __forceinline__ int xread_idx(int Nw, int t, int s) {
  const int st = xread_stride(Nw, NT);
  return st ? xlay(Nw, t) + s * st : xlay(Nw, t + s * NT);
}
// Entry of the r-th output (base + r*Ns) of a thread whose first output is idxD (idxD % Ns = k,
// idxD = (t/Ns)*8*Ns + k), in the layout of span Ns: the base's layout plus r's offset.
__device__ __forceinline__ int xwrite_idx(int Ns, int idxD, int r) {
  return xlay(Ns, idxD) + r * (Ns == 1 ? 1 : Ns == 8 ? 8 : Ns);
}

// Which per-freq quantities fe_range keeps in registers across the bands of a freq (bit 0:
// the rotator of the thread's samples, bit 1: the fractional-sample phasors of its channels).
#ifndef FXG_F_HOIST
#define FXG_F_HOIST 3
#endif
// Waves per SIMD the persistent F-engine's register allocation is held to.
#ifndef FXG_P_MINWAVES
#define FXG_P_MINWAVES 1
#endif

template <int L>
struct Plan {
  static constexpr int LOG = (L == 256) ? 8 : (L == 512) ? 9 : (L == 1024) ? 10 : (L == 2048) ? 11
                           : (L == 4096) ? 12 : (L == 8192) ? 13 : 0;
  static constexpr int N8 = LOG / 3;      // radix-8 passes
  static constexpr int REM = LOG % 3;     // 0: none, 1: radix-2 pass, 2: radix-4 pass
  static constexpr int NT = L / 8;
  // Per-pass twiddle tables, laid out [pass][r-1][k] with k (the butterfly's position in its
  // span Ns) fastest, so lanes of a wavefront read consecutive entries.  Pass 0 needs none.
  static constexpr int pow8(int p) { int v = 1; for (int q = 0; q < p; ++q) v *= 8; return v; }
  static constexpr int off8(int p) { int o = 0; for (int q = 1; q < p; ++q) o += 7 * pow8(q); return o; }
  static constexpr int OFFREM = off8(N8);
  static constexpr int TWN = OFFREM + (REM ? ((1 << REM) - 1) * pow8(N8) : 0);
};

// Stockham FFT of L points over NT = L/8 butterfly positions, held V per thread: position
// vt = tl + u*TPF (u < V) holds v[u][s] = x[vt + s*NT] on entry and X[vt + s*NT] on exit.
// sbuf has sbuf_entries(L) entries (laid out by xlay); tw holds Plan<L>::TWN twiddles exp(-2 pi i k r / (Ns R))
// per pass.  Every thread of the workgroup calls it (the barriers are workgroup-wide); with
// several FFTs per workgroup each group passes its own sbuf.
template <int L, int V, int TPF>
__device__ __forceinline__ void fft_stockham(float2 (*v)[8], float2* sbuf, const float2* tw, int tl) {
  using P = Plan<L>;
  constexpr int NT = P::NT;
  static_assert(V * TPF == NT, "butterfly positions");
  int Ns = 1;
#pragma unroll
  for (int p = 0; p < P::N8; ++p) {
    if (p > 0) {
#pragma unroll
      for (int u = 0; u < V; ++u) {
        const int vt = tl + u * TPF;
#pragma unroll
        for (int s = 0; s < 8; ++s) v[u][s] = sbuf[xread_idx<NT>(Ns / 8, vt, s)];
        const int k = vt & (Ns - 1);
        const float2* twp = tw + P::off8(p) + k;
#pragma unroll
        for (int r = 1; r < 8; ++r) v[u][r] = cmul(v[u][r], twp[(r - 1) * Ns]);
      }
    }
#pragma unroll
    for (int u = 0; u < V; ++u) dft8(v[u]);
    const bool last = (P::REM == 0) && (p == P::N8 - 1);
    if (!last) {
      if (p > 0) __syncthreads();            // every thread has read its inputs (pass 0 reads
                                             // none; the caller's barrier freed sbuf)
#pragma unroll
      for (int u = 0; u < V; ++u) {
        const int vt = tl + u * TPF;
        const int k = vt & (Ns - 1);
        const int idxD = (vt / Ns) * Ns * 8 + k;
#pragma unroll
        for (int r = 0; r < 8; ++r) sbuf[xwrite_idx(Ns, idxD, r)] = v[u][r];
      }
      __syncthreads();
    }
    Ns *= 8;
  }
  if (P::REM == 2) {                         // radix-4 pass, two butterflies per position
#pragma unroll
    for (int u = 0; u < V; ++u) {
      const int vt = tl + u * TPF;
#pragma unroll
      for (int s = 0; s < 8; ++s) v[u][s] = sbuf[xread_idx<NT>(Ns / 8, vt, s)];
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const int j = vt + b * NT;
        const int k = j & (Ns - 1);
        const float2* twp = tw + P::OFFREM + k;
        float2 a0 = v[u][b], a1 = v[u][b + 2], a2 = v[u][b + 4], a3 = v[u][b + 6];
        a1 = cmul(a1, twp[0]);
        a2 = cmul(a2, twp[Ns]);
        a3 = cmul(a3, twp[2 * Ns]);
        dft4(a0, a1, a2, a3);
        v[u][b] = a0; v[u][b + 2] = a1; v[u][b + 4] = a2; v[u][b + 6] = a3;
      }
    }
  } else if (P::REM == 1) {                  // radix-2 pass, four butterflies per position
#pragma unroll
    for (int u = 0; u < V; ++u) {
      const int vt = tl + u * TPF;
#pragma unroll
      for (int s = 0; s < 8; ++s) v[u][s] = sbuf[xread_idx<NT>(Ns / 8, vt, s)];
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const int j = vt + m * NT;
        const int k = j & (Ns - 1);
        float2 a0 = v[u][m], a1 = cmul(v[u][m + 4], tw[P::OFFREM + k]);
        dft2(a0, a1);
        v[u][m] = a0; v[u][m + 4] = a1;
      }
    }
  }
}

// Valid samples of [g0, g0+n) given the frame flags (mk5mode.cpp:94-174: goodsamples is the
// count of decoded samples inside the FFT window).
__device__ int count_valid(const uint8_t* fv, const int2* fz, int nframes, int spf, long long g0, int n) {
  long long g1 = g0 + n;
  int f0 = (int)(g0 / spf), f1 = (int)((g1 - 1) / spf);
  int good = 0;
  for (int f = f0; f <= f1; ++f) {
    if (f >= nframes || !fv[f]) continue;
    long long a = (long long)f * spf, b = a + spf;
    if (fv[f] == FXG_FRAME_PARTIAL) { b = a + fz[f].y; a += fz[f].x; }
    long long lo = a > g0 ? a : g0, hi = b < g1 ? b : g1;
    if (hi > lo) good += (int)(hi - lo);
  }
  return good;
}

// Per-FFT delay chain (mode.cpp:671-675, 743-757, 776-825), all FP64.
//   order 0: integerdelay = (int)averagedelay, b = averagedelay - integerdelay (post-F phase)
//   order 1: a = d2 - d0, b = d0 + (d1 - (a/2 + d0))/3 - integerdelay
//   order 2: a = q0, b = q1 + 2 i q0, c = q2 + i q1 + i^2 q0 - integerdelay
__device__ FftDelay fft_delay(const DevDs& ds, const DevDsRun& run, const uint8_t* fv, const int2* fz, int i,
                              int offsec, int offns, double scan_start_sec, int order) {
  FftDelay r;
  r.ok = 0; r.weight = 0.0f; r.n0 = 0; r.fse = 0.0f; r.a = 0; r.b = 0; r.c = 0; r.intdelay = 0;
  r.intwall = 0; r.fracwall = 0;
  if (!run.valid || run.datalen <= 1) return r;
  if (((run.vflags[i / FLAGS_PER_INT] >> (i % FLAGS_PER_INT)) & 1) == 0) return r;
  const double* q = run.interp;
  const double dt = ds.sampletime_us;
  const int L = ds.L;
  const double fc = i + 0.5;
  const double avgdelay = q[0] * fc * fc + q[1] * fc + q[2];
  const double fftstart = i * L * dt;
  const double starttime = (offsec - run.datasec) * 1000000.0
                         + ((long long)offns - (long long)run.datans) / 1000.0 + fftstart - avgdelay;
  int n0 = (int)(starttime / dt + 0.5);
  if (n0 < -1 || (((long long)(n0 + L)) * ds.bpb_num) / ds.bpb_den > run.datalen) return r;
  if (n0 == -1) n0 = 0;
  const int good = count_valid(fv, fz, run.nframes, ds.spf, n0, L);
  const float w = (float)good / (float)L;
  if (!(w > 0.0f)) return r;
  r.ok = 1;
  r.n0 = n0;
  r.weight = w;
  r.fse = (float)(starttime - n0 * dt);
  const double di = (double)i;
  if (order == 1) {
    const double d0 = q[0] * di * di + q[1] * di + q[2];
    const double d1 = q[0] * (di + 0.5) * (di + 0.5) + q[1] * (di + 0.5) + q[2];
    const double d2 = q[0] * (di + 1) * (di + 1) + q[1] * (di + 1) + q[2];
    const double a = d2 - d0;
    double b = d0 + (d1 - (a * 0.5 + d0)) / 3.0;
    const int intdelay = (int)b;
    b -= intdelay;
    r.a = a; r.b = b; r.intdelay = intdelay;
  } else if (order == 2) {
    const double b = q[1] + di * q[0] * 2.0;
    double c = q[2] + di * q[1] + (double)((long long)i * i) * q[0];
    const int intdelay = (int)c;
    c -= intdelay;
    r.a = q[0]; r.b = b; r.c = c; r.intdelay = intdelay;
  } else {
    const int intdelay = (int)avgdelay;
    r.b = avgdelay - intdelay;
    r.intdelay = intdelay;
  }
  const double walltime = scan_start_sec + offsec + offns / 1.0e9 + fftstart / 1.0e6;
  r.intwall = (int)walltime;
  r.fracwall = walltime - r.intwall;
  return r;
}


// Per-datastream constants the F-engine reads, copied to registers once per workgroup.
struct DsConst {
  int C, S, NS, NFS, nrec, nrecfreqs, fmt, spf, framebytes, headerbytes, nbits, fractionallo, dsb;
  double sampletime_us;
};

// Per-freq constants (mode.cpp:853-931).
struct FreqConst {
  double lofreq, lo_offset_hz, fraclooffset, clockoff, halfdelta;
  int lsb, looff, isfraclooffset, nbands;
};

__device__ __forceinline__ FreqConst load_freq(const FXG_CONST DevRecFreq& rf) {
  FreqConst f;
  f.lofreq = rf.lofreq; f.lo_offset_hz = rf.lo_offset_hz; f.fraclooffset = rf.fraclooffset;
  f.clockoff = rf.clockoff; f.halfdelta = rf.halfdelta;
  f.lsb = rf.lsb; f.looff = rf.looff; f.isfraclooffset = rf.isfraclooffset; f.nbands = rf.nbands;
  return f;
}

// Fringe rotator sub-phasors (j < S) and step-phasors (S <= j < S+NS) for one freq
// (order 1: mode.cpp:883-955; order 2: mode.cpp:956-1026, whose quadratic term is the
// per-sub argument squad[j], applied as a power of the step index).  FP64 phase in turns,
// wrapped by truncation, FP32 sin/cos.
template <int ORDER>
__device__ void build_rotator(float2* srot, float* squad, const DsConst& k, const FreqConst& rf,
                              const FftDelay& fd, int L, int t, int nt) {
  const int S = k.S, NS = k.NS;
  const double lofreq = rf.lofreq;
  const double sx1 = (double)S / (double)L;                              // stepxoff[1]
  for (int j = t; j < S + NS; j += nt) {
    double ph;
    if (j < S) {
      const double xoff = (double)j / (double)L;                        // subxoff
      double xval;                                                       // subxval
      if (ORDER == 1) {
        xval = __dadd_rn(__dmul_rn(xoff, fd.a), fd.b);
      } else {
        xval = __dadd_rn(__dmul_rn(xoff, __dadd_rn(fd.b, __dmul_rn(fd.a, sx1))), fd.c);
        const double qx = __dmul_rn(xoff, __dmul_rn(__dmul_rn(2.0, fd.a), sx1));   // subquadxval
        const double qph = __dmul_rn(qx, lofreq);
        squad[j] = (float)(-TWO_PI_D * (qph - trunc(qph)));
      }
      ph = __dmul_rn(xval, lofreq);
      if (k.fractionallo)
        ph = __dadd_rn(ph, __dmul_rn(lofreq - (double)(int)lofreq, (double)fd.intdelay));
      if (rf.looff) {
        const double toff = j * k.sampletime_us / 1e6;                   // subtoff
        ph = __dadd_rn(ph, __dmul_rn(toff, -rf.lo_offset_hz));
        ph = __dadd_rn(ph, -rf.lo_offset_hz * fd.fracwall);
        if (rf.isfraclooffset) ph = __dadd_rn(ph, -rf.fraclooffset * fd.intwall);
      }
    } else {
      const int s = j - S;
      const double xoff = (double)(s * S) / (double)L;                  // stepxoff
      if (ORDER == 1)
        ph = __dmul_rn(__dmul_rn(xoff, fd.a), lofreq);
      else
        ph = __dmul_rn(__dadd_rn(__dmul_rn(xoff, fd.b), __dmul_rn(__dmul_rn(xoff, xoff), fd.a)), lofreq);
      if (rf.looff) {
        const double toff = s * S * k.sampletime_us / 1e6;               // steptoff
        ph = __dadd_rn(ph, __dmul_rn(toff, -rf.lo_offset_hz));
      }
    }
    const float arg = (float)(-TWO_PI_D * (ph - trunc(ph)));
    float sn, cs;
    sincosf(arg, &sn, &cs);
    srot[j] = make_float2(cs, sn);
  }
}

// Fractional-sample phasors of band kb of a freq: sub (j < S) and step (S <= j < S+NFS)
// (mode.cpp:1030-1122), with the band's clock-delta term hd (+delta/2 for R, -delta/2 for
// the other pols).  Order 0 adds the post-F fringe phase to the sub arguments
// (mode.cpp:1040-1049).  A pol phase offset pc (mode.cpp:1318-1330) multiplies the sub
// phasors, so every channel of the band turns by it.
template <int ORDER>
__device__ void build_frac(float2* sfrac, const DsConst& k, const FreqConst& rf, bool iscomplex,
                           const float* __restrict__ cf, const FftDelay& fd, double hd, float2 pc,
                           int t, int nt) {
  const int S = k.S, NFS = k.NFS;
  const bool turn = !(pc.x == 1.0f && pc.y == 0.0f);
  const float e = (float)((double)fd.fse - rf.clockoff + hd);
  const float* sub = cf;
  // double sideband: dsbstepchannelfreqs; any other LSB: lsbstepchannelfreqs (mode.cpp:838-857)
  const float* step = k.dsb ? cf + S + 2 * NFS : rf.lsb ? cf + S + NFS : cf + S;
  float prf = 0.0f;
  if (ORDER == 0) {
    const double lofreq = rf.lofreq;
    double pr = fd.b * lofreq;
    if (k.fractionallo) pr += fd.intdelay * (lofreq - (int)lofreq);
    pr -= ((double)fd.intwall + fd.fracwall) * rf.lo_offset_hz;
    prf = (float)(-TWO_PI_D * (pr - (int)pr));
  }
  for (int j = t; j < S + NFS; j += nt) {
    const float arg = (j < S) ? (ORDER == 0 ? sub[j] * e + prf : sub[j] * e) : step[j - S] * e;
    float sn, cs;
    sincosf(arg, &sn, &cs);
    float2 ph = make_float2(cs, sn);
    if (turn && j < S) ph = cmul(ph, pc);
    sfrac[j] = ph;
  }
}

// A complex value as two 16-bit floats in one word (MODE 1: FP16, 2: BF16), round to nearest.
template <int MODE>
__device__ __forceinline__ uint32_t pack16m(float2 v) {
  if (MODE == 1) {
    __half2 h = __floats2half2_rn(v.x, v.y);
    return *reinterpret_cast<uint32_t*>(&h);
  }
  __hip_bfloat162 h = __float22bfloat162_rn(v);
  return *reinterpret_cast<uint32_t*>(&h);
}
template <int N>
struct IntC { static constexpr int value = N; };



// LDS floats of the staged channel-frequency tables (S + 3 NFS <= 3 phcap).
__host__ __device__ constexpr size_t stage_lds_bytes(int phcap) { return sizeof(float) * 3 * (size_t)phcap; }

// Copy datastream d's channel-frequency tables (read per lane by build_frac) into LDS; all
// threads of the workgroup call it (the barriers are workgroup-wide).
__device__ void stage_cf(const FengineArgs& args, CDevDs ds, float* scf, int tid, int nt) {
  __syncthreads();                                  // earlier readers of the tables are done
  const float* cf = args.chanfreq + ds->chanfreq_off;
  const int ncf = ds->S + 3 * ds->NFS;
  for (int i = tid; i < ncf; i += nt) scf[i] = cf[i];
  __syncthreads();
}

// Geometry of FFT length L on a workgroup of NTB threads.  An FFT has NT = L/8 butterfly
// positions.  With NT < NTB the workgroup runs G = NTB/NT FFTs side by side, TPF = NT threads
// each (consecutive FFTs of one datastream); with NT > NTB each thread holds V = NT/NTB
// positions of the one FFT.
template <int L, int NTB>
struct Geo {
  static constexpr int NT = L / 8;
  static constexpr int TPF = NT < NTB ? NT : NTB;
  static constexpr int G = NTB / TPF;
  static constexpr int V = NT / TPF;
  static constexpr int SBUF = sbuf_entries(L); // exchange entries of one FFT
};

// LDS bytes of one (L, NTB) geometry: G exchange buffers, the twiddles, G rotator and G
// fractional-sample phasor tables of phcap entries, and for order 2 G quadratic-argument tables.
template <int L, int NTB>
size_t fe_lds_bytes(int phcap, int order) {
  using GE = Geo<L, NTB>;
  return sizeof(float2) * ((size_t)GE::G * GE::SBUF + Plan<L>::TWN + 2 * (size_t)GE::G * phcap) +
         sizeof(float) * ((size_t)GE::G * phcap + 264) +     // quadratic arguments, 8- and 2-bit levels
         stage_lds_bytes(phcap);
}

// The F-engine proper: FFT units [u0, u1) (8 consecutive FFTs each) of datastream d.
// Per FFT and recorded freq: the rotator and fractional-sample tables, then per band the
// decode + rotation, the FFT, channel selection, fractional-sample correction and the store
// (FP32 spectra, or scaled 16-bit MFMA operands including zoom rows).  The control flow is the
// same for every thread of the workgroup (flagged FFTs compute on zeros and store zeros), so
// the workgroup barriers are uniform with several FFTs in flight.  With load_tw the twiddles
// of L are (re)loaded into LDS first.
template <int L, bool CPLX, int ORDER, int NTB>
__device__ void fe_range(const FengineArgs& args, const float2* __restrict__ twiddle, int d, int u0, int u1,
                         float2* smem, bool load_tw) {
  using P = Plan<L>;
  using GE = Geo<L, NTB>;
  constexpr int NT = GE::NT, TPF = GE::TPF, G = GE::G, V = GE::V;
  const int t = threadIdx.x;
  const int g = t / TPF, tl = t - g * TPF;
  float2* sbuf = smem + g * GE::SBUF;
  float2* stw = smem + G * GE::SBUF;
  float2* srot = stw + P::TWN + g * args.phcap;
  float2* sfrac = stw + P::TWN + (G + g) * args.phcap;
  float* squad = reinterpret_cast<float*>(stw + P::TWN + 2 * G * args.phcap) + g * args.phcap;
  float* slut8 = reinterpret_cast<float*>(stw + P::TWN + 2 * G * args.phcap) + G * args.phcap;

  if (load_tw) {
    __syncthreads();                              // earlier users of the twiddle tables are done
    for (int m = t; m < P::TWN; m += NTB) stw[m] = twiddle[m];
    for (int m = t; m < 256; m += NTB) slut8[m] = lut8(m);      // 8-bit levels, once per workgroup
    for (int m = t; m < 8; m += NTB) slut8[256 + m] = lut2(m < 4 ? FXG_FMT_VDIF : FXG_FMT_MARK5B, m & 3);
  }
  const CDevDs dsp = (CDevDs)(args.ds + d);
  float* scf = slut8 + 264;                       // staged channel-frequency tables
  stage_cf(args, dsp, scf, t, NTB);

  DsConst k;
  k.C = dsp->C; k.S = dsp->S; k.NS = dsp->NS; k.NFS = dsp->NFS; k.nrec = dsp->nrecbands;
  k.nrecfreqs = dsp->nrecfreqs; k.fmt = dsp->format; k.spf = dsp->spf;
  k.framebytes = dsp->framebytes; k.headerbytes = dsp->headerbytes; k.nbits = dsp->nbits;
  k.fractionallo = dsp->fractionallo; k.sampletime_us = dsp->sampletime_us; k.dsb = dsp->dsb;
  const long long spec_off32 = dsp->spec_off;
  const GBytes fv = (GBytes)(args.fvalid + dsp->fvalid_off);
  const int2* __restrict__ fz = args.fzone + dsp->fvalid_off;
  const float* __restrict__ cf = scf;
  const GBytes data = (GBytes)args.run[d].data;
  const int nframes = args.run[d].nframes;
  const FftDelay* __restrict__ params = args.params + (long long)d * args.chunkcap;
  const int C = k.C, S = k.S;
  const bool twobit = (k.nbits == 2);
  const float xsc = dsp->xsc;
  const double inv_spf = 1.0 / (double)k.spf;
  const float* __restrict__ slev2 = slut8 + 256 + (k.fmt == FXG_FMT_MARK5B ? 4 : 0);   // 2-bit levels
  const int out_mode = args.out_mode;
  const int cap = args.chunkcap;

  for (int lt0 = u0 * 8; lt0 < u1 * 8; lt0 += G) {
    const int lt = lt0 + g;                        // this group's FFT within the chunk
    // with more groups than FFTs left (G = 16 > 8 FFTs per unit) the extra groups compute on
    // zeros, read no parameters and store nothing, keeping the barriers uniform
    const bool active = lt < u1 * 8;
    FftDelay fd;
    if (active) {
      fd = params[lt];
    } else {
      fd.ok = 0; fd.n0 = 0; fd.fse = 0.0f; fd.intdelay = 0; fd.a = 0; fd.b = 0; fd.c = 0;
      fd.weight = 0.0f; fd.intwall = 0; fd.fracwall = 0;
    }
    const bool fok_all = fd.ok != 0;               // flagged FFTs: zero spectra (mode.cpp:656-669)
    const float fs = fok_all ? xsc : 0.0f;         // output scale of this FFT's spectra

    // raw sample words of this thread's time samples, shared by all bands: for 2-bit data
    // the 32-bit payload word holding the sample's fields and the field offset in it; for
    // 8-bit data the byte offset of the sample
    uint32_t wv[V][8];
    uint32_t fok[V], qlo[V];
    // (all sixteen loads of a position are issued before any result is used: the frame index
    // is clamped to a frame of the buffer, and a sample past the data is masked afterwards)
    const bool anyload = fok_all && nframes > 0;
    // weight 1: every sample of the window decodes (the good-sample count equals L), so the
    // frame flags need not be read
    const bool whole = fok_all && fd.weight == 1.0f;
#pragma unroll
    for (int u = 0; u < V; ++u) {
      const unsigned g0 = (unsigned)fd.n0 + (unsigned)(tl + u * TPF);
      unsigned frs[8], wis[8];
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        // frame and sample within it: the quotient through the reciprocal of spf (exact after
        // the one correction: sample indices and spf stay below 2^31 and 2^20)
        const unsigned gs = g0 + (unsigned)(s * NT);
        unsigned fr = (unsigned)((double)gs * inv_spf);
        unsigned within = gs - fr * (unsigned)k.spf;
        if (within >= (unsigned)k.spf) { within -= k.spf; ++fr; }
        frs[s] = fr;
        wis[s] = within;
      }
      uint32_t fvb[8], w8[8];
#pragma unroll
      for (int s = 0; s < 8; ++s) { fvb[s] = 0u; w8[s] = 0u; }
      if (anyload) {
#pragma unroll
        for (int s = 0; s < 8; ++s) {
          const unsigned frc = min(frs[s], (unsigned)(nframes - 1));
          if (!whole) fvb[s] = fv[frc];
          if (twobit) {
            const unsigned q = wis[s] * (unsigned)k.nrec;   // first field of the sample
            w8[s] = *(GWords)(data + frc * (unsigned)k.framebytes + (unsigned)k.headerbytes + ((q >> 4) << 2));
          }
        }
      }
      fok[u] = whole ? 0xFFu : 0u;
      qlo[u] = 0u;
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        if (!whole) {
          const bool ok = frs[s] < (unsigned)nframes && sample_ok(fvb[s], fz, frs[s], wis[s]);
          fok[u] |= (ok ? 1u : 0u) << s;
        }
        const unsigned q = wis[s] * (unsigned)k.nrec;
        qlo[u] |= (q & 15u) << (4 * s);
        wv[u][s] = twobit ? w8[s]
                          : frs[s] * (unsigned)k.framebytes + (unsigned)k.headerbytes +
                                wis[s] * (unsigned)k.nrec * (CPLX ? 2u : 1u);
      }
    }
    // Wait for this FFT's loads here, once: left to its first use inside the freq and band
    // loops, the wait would run in every band and, since the vector-memory counter also counts
    // stores, wait for the previous band's spectra stores each time.
#pragma unroll
    for (int u = 0; u < V; ++u) {
#pragma unroll
      for (int s = 0; s < 8; ++s) asm volatile("" ::"v"(wv[u][s]));
    }
    asm volatile("" ::"v"(fd.a), "v"(fd.b), "v"(fd.c), "v"(fd.fse), "v"(fd.n0), "v"(fd.intdelay),
                 "v"(fd.intwall), "v"(fd.fracwall));

    for (int f = 0; f < k.nrecfreqs; ++f) {
      const FreqConst rf = load_freq(dsp->rf[f]);
      if (rf.nbands == 0) continue;
      __syncthreads();                            // previous users of srot/sfrac are done
      if (ORDER != 0) build_rotator<ORDER>(srot, squad, k, rf, fd, L, tl, TPF);
      build_frac<ORDER>(sfrac, k, rf, CPLX, cf, fd, dsp->rf[f].bhd[0], make_float2(dsp->rf[f].pcre[0], dsp->rf[f].pcim[0]), tl, TPF);
      __syncthreads();
      // rotator (orders 1, 2) of this thread's samples and fractional-sample phasors of its
      // channels: shared by the bands of the freq, so they can be computed once per freq
      // (FXG_F_HOIST) when a thread holds one butterfly position, at the cost of registers
      constexpr bool HOIST_ROT = (V == 1) && (FXG_F_HOIST & 1);
      constexpr bool HOIST = (V == 1) && (FXG_F_HOIST & 2);
      constexpr int NKEEP = CPLX ? 8 : 4;
      // Post-F rotation of real data (order 0): the FFT input is real, so two bands of the
      // freq go through one complex FFT as a + ib and are split after it,
      // A[j] = (Z[j] + conj Z[-j]) / 2 and B[j] = -i (Z[j] - conj Z[-j]) / 2 (the reference
      // takes one real-to-complex FFT per band, mode.cpp:1139-1156).  Bands pair (0, 1),
      // (2, 3) when their count is even and each pair shares its phasor table; the 1/2 rides
      // on the phasors.
      constexpr bool PAIRABLE = (ORDER == 0) && !CPLX;
      bool fpair = PAIRABLE && (rf.nbands % 2 == 0);
      if (PAIRABLE)
        for (int kb = 1; kb < rf.nbands; kb += 2)
          if (dsp->rf[f].newfrac[kb]) fpair = false;
      auto make_rot = [&](int u, float2* rot8) {
        const int vt = tl + u * TPF;
        const float2 sub = ORDER != 0 ? srot[vt & (S - 1)] : make_float2(1.0f, 0.0f);
        // order 2: the quadratic sub term to the power of the step index j = vt/S + s*NT/S
        // (mode.cpp:1017-1024), as A * D^s with A = e^{i j0 th}, D = e^{i (NT/S) th}
        float2 qpow = make_float2(1.0f, 0.0f), qstep = make_float2(1.0f, 0.0f);
        if (ORDER == 2) {
          const float th = squad[vt & (S - 1)];
          float sn, cs;
          sincosf((float)(vt / S) * th, &sn, &cs);
          qpow = make_float2(cs, sn);
          sincosf((float)(NT / S) * th, &sn, &cs);
          qstep = make_float2(cs, sn);
        }
        const int st0 = vt / S;                          // NT is a multiple of S (checked at set-up)
        // step 0's phasor is exp(-0 i) = (1, -0), so every step multiplies: no branch, and the
        // eight table reads are in flight together
#pragma unroll
        for (int s = 0; s < 8; ++s) {
          const int st = st0 + s * (NT / S);
          if (ORDER == 1) {
            rot8[s] = cmul(sub, srot[S + st]);
          } else if (ORDER == 2) {
            const float2 pq = cmul(sub, qpow);
            rot8[s] = cmul(pq, srot[S + st]);
            qpow = cmul(qpow, qstep);
          } else {
            rot8[s] = sub;                                   // no pre-F rotation
          }
        }
      };
      // the fractional-sample phasors times the output scale `fs` (the 16-bit input scale, 1
      // for FP32 spectra, 0 for a flagged FFT, whose spectra are zero: mode.cpp:656-669)
      auto make_frq = [&](int u, float2* frqk) {
        const int vt = tl + u * TPF;
        const float2 fsub0 = sfrac[vt & (S - 1)];         // the same sub-phasor for every m
        const float fsx = (PAIRABLE && fpair) ? 0.5f * fs : fs;
        const float2 fsub = make_float2(fsub0.x * fsx, fsub0.y * fsx);
        const int f0 = S + vt / S;
#pragma unroll
        for (int m = 0; m < NKEEP; ++m) frqk[m] = cmul(fsub, sfrac[f0 + m * (NT / S)]);
      };
      float2 rot_h[HOIST_ROT ? 8 : 1], frq_h[HOIST ? NKEEP : 1];
      if (HOIST_ROT) {
        make_rot(0, rot_h);
        // a sample of an invalid frame decodes to zero: a zero rotator gives x * 0 = 0
#pragma unroll
        for (int s = 0; s < 8; ++s)
          if (!((fok[0] >> s) & 1u)) rot_h[s] = make_float2(0.0f, 0.0f);
      }
      if (HOIST) make_frq(0, frq_h);
      for (int kb = 0; kb < rf.nbands; kb += (PAIRABLE && fpair) ? 2 : 1) {
        const int band = dsp->rf[f].band[kb];
        const int band2 = (PAIRABLE && fpair) ? dsp->rf[f].band[kb + 1] : band;
        if (kb > 0 && dsp->rf[f].newfrac[kb]) {  // another pol's phasor table (uniform branch)
          __syncthreads();                        // the previous band's phasor reads are done
          build_frac<ORDER>(sfrac, k, rf, CPLX, cf, fd, dsp->rf[f].bhd[kb], make_float2(dsp->rf[f].pcre[kb], dsp->rf[f].pcim[kb]), tl, TPF);
          __syncthreads();
          if (HOIST) make_frq(0, frq_h);
        }
        float2 v[V][8];
#pragma unroll
        for (int u = 0; u < V; ++u) {
          float2 rot_l[HOIST_ROT ? 1 : 8];
          if (!HOIST_ROT) make_rot(u, rot_l);
          if (!CPLX) {
            // the sample width is tested once per band, not per sample, so the eight decodes
            // are straight-line code
            float x[8], x2[8];
            if (twobit) {                            // levels from the LDS table of the format
#pragma unroll
              for (int s = 0; s < 8; ++s) {
                const int q = (int)((qlo[u] >> (4 * s)) & 15u);
                const float lv = slev2[(wv[u][s] >> (2 * (q + band))) & 3u];
                // invalid samples: zeroed through the rotator when it is kept per freq
                x[s] = (HOIST_ROT && ORDER != 0) || ((fok[u] >> s) & 1u) ? lv : 0.0f;
                x2[s] = 0.0f;
                if (PAIRABLE && fpair && ((fok[u] >> s) & 1u)) x2[s] = slev2[(wv[u][s] >> (2 * (q + band2))) & 3u];
              }
            } else {
#pragma unroll
              for (int s = 0; s < 8; ++s) {
                x[s] = 0.0f;
                x2[s] = 0.0f;
                if ((fok[u] >> s) & 1u) x[s] = slut8[data[wv[u][s] + band]];
                if (PAIRABLE && fpair && ((fok[u] >> s) & 1u)) x2[s] = slut8[data[wv[u][s] + band2]];
              }
            }
#pragma unroll
            for (int s = 0; s < 8; ++s) {
              const float2 rot = HOIST_ROT ? rot_h[s] : rot_l[HOIST_ROT ? 0 : s];
              v[u][s] = ORDER == 0 ? make_float2(x[s], x2[s])             // real input(s): a + ib
                                   : make_float2(x[s] * rot.x, x[s] * rot.y);  // (x + 0i) * rotator
            }
            continue;
          }
#pragma unroll
          for (int s = 0; s < 8; ++s) {             // complex 8-bit samples
            const float2 rot = HOIST_ROT ? rot_h[s] : rot_l[HOIST_ROT ? 0 : s];
            float2 x = make_float2(0.f, 0.f);
            if ((fok[u] >> s) & 1u) {
              const GBytes pp = data + wv[u][s] + 2 * band;
              x = make_float2(slut8[pp[0]], slut8[pp[1]]);
            }
            v[u][s] = cmul(rot, x);
            // complex LSB: the reference conj-flips the spectrum (mode.cpp:1186-1201);
            // FFT(conj y)[c] = conj(FFT(y)[-c]) gives the same channels without a flip
            if (rf.lsb) v[u][s].y = -v[u][s].y;
          }
        }
        __syncthreads();                          // sbuf free (previous band's FFT done)
        fft_stockham<L, V, TPF>(v, sbuf, stw, tl);
        // channel selection and fractional-sample correction
        float2 Fo[V][NKEEP];                      // output channels vt + m*NT
        float2 Fo2[PAIRABLE ? V : 1][PAIRABLE ? NKEEP : 1];   // the pair's second band
        if (PAIRABLE && fpair) {
          // Z[-j] of every kept channel j sits with another thread: through sbuf in natural
          // order (L <= sbuf entries)
          __syncthreads();                        // the FFT's last reads of sbuf are done
#pragma unroll
          for (int u = 0; u < V; ++u) {
            const int vt = tl + u * TPF;
#pragma unroll
            for (int s = 0; s < 8; ++s) sbuf[vt + s * NT] = v[u][s];
          }
          __syncthreads();
#pragma unroll
          for (int u = 0; u < V; ++u) {
            const int vt = tl + u * TPF;
            float2 frq_l[HOIST ? 1 : NKEEP];
            if (!HOIST) make_frq(u, frq_l);
#pragma unroll
            for (int m = 0; m < NKEEP; ++m) {
              // USB keeps X[vt + m*NT], real LSB X[C + vt + m*NT] (mode.cpp:1205)
              const bool upper = rf.lsb != 0;
              const int j = vt + m * NT + (upper ? 4 * NT : 0);
              const float2 frq = HOIST ? frq_h[m] : frq_l[HOIST ? 0 : m];
              const float2 zj = upper ? v[u][(m + 4) & 7] : v[u][m];
              const float2 zm = sbuf[(L - j) & (L - 1)];
              Fo[u][m] = cmul(make_float2(zj.x + zm.x, zj.y - zm.y), frq);
              Fo2[u][m] = cmul(make_float2(zj.y + zm.y, zm.x - zj.x), frq);   // -i (Z[j] - conj Z[-j])
            }
          }
        } else {
#pragma unroll
        for (int u = 0; u < V; ++u) {
          const int vt = tl + u * TPF;
          float2 frq_l[HOIST ? 1 : NKEEP];
          if (!HOIST) make_frq(u, frq_l);
#pragma unroll
          for (int m = 0; m < NKEEP; ++m) {
            // USB and complex keep X[vt + m*NT]; real LSB keeps X[C + vt + m*NT] (mode.cpp:1205);
            // double sideband swaps the halves (mode.cpp:1191-1192, 1211-1213).  Register
            // indices stay compile-time constants (a run-time index would move v to scratch).
            const bool upper = CPLX ? (k.dsb != 0) : (rf.lsb != 0);
            const int c = vt + m * NT;
            const float2 frq = HOIST ? frq_h[m] : frq_l[HOIST ? 0 : m];
            float2 X = upper ? v[u][(m + 4) & 7] : v[u][m];
            if (CPLX && rf.lsb && !k.dsb && c == 0) X.y = -X.y;   // DC is not flipped (mode.cpp:1199-1200)
            Fo[u][m] = cmul(X, frq);
          }
        }
        }
        if (!active) continue;
        // the stores: the output form is tested once per band, so each form's stores are
        // straight-line code.  Channel c = vt + m*NT; NT is a multiple of 8, so c >> 3 =
        // (vt >> 3) + m*NT/8 and c & 7 = vt & 7: per-thread bases with uniform strides.
        auto store_band = [&](const int band, const float2 (&Fo)[V][NKEEP]) {
        if (out_mode == 0) {
          float2* __restrict__ o32 = args.spec32 + spec_off32 + ((long long)band * cap + lt) * C;
#pragma unroll
          for (int u = 0; u < V; ++u) {
            float2* __restrict__ o32t = o32 + tl + u * TPF;
#pragma unroll
            for (int m = 0; m < NKEEP; ++m) o32t[m * NT] = Fo[u][m];
          }
          return;
        }
        auto store16 = [&](auto mode_c) {
          constexpr int OM = decltype(mode_c)::value;
          // channel c of FFT lt sits at (c / TW) cap TW + lt TW + c % TW of the row; NT is a
          // multiple of TW, so per thread a base and a uniform stride per m
          const CDevXGroup xgr = ((CDevXGroup)args.xgroups) + dsp->band_xfreq[band];
          const int twl = xgr->twlog, twm = (1 << twl) - 1;
          uint32_t* __restrict__ o16 = args.spec16 + xgr->spec_off + (long long)dsp->band_xrow[band] * xgr->rowstride +
                                       ((long long)lt << twl);
          const long long cs16 = ((long long)(NT >> twl) * cap) << twl;
#pragma unroll
          for (int u = 0; u < V; ++u) {
            const int vt = tl + u * TPF;
            uint32_t* __restrict__ o16t = o16 + (((long long)(vt >> twl) * cap) << twl) + (vt & twm);
#pragma unroll
            for (int m = 0; m < NKEEP; ++m) o16t[m * cs16] = pack16m<OM>(Fo[u][m]);
          }
          // zoom windows of this band: the same channels again, as rows of the zoom freq's group
          for (int z = 0; z < dsp->nzout; ++z) {
            if (dsp->zout_band[z] != band) continue;
            const int c0 = dsp->zout_c0[z], nz = dsp->zout_n[z];
            const CDevXGroup zg = ((CDevXGroup)args.xgroups) + dsp->zout_group[z];
            const int zl = zg->twlog, zm = (1 << zl) - 1;
            uint32_t* __restrict__ oz = args.spec16 + zg->spec_off + (long long)dsp->zout_row[z] * zg->rowstride +
                                        ((long long)lt << zl);
#pragma unroll
            for (int u = 0; u < V; ++u) {
#pragma unroll
              for (int m = 0; m < NKEEP; ++m) {
                const int cz = tl + u * TPF + m * NT - c0;
                if (cz >= 0 && cz < nz) oz[(((long long)(cz >> zl) * cap) << zl) + (cz & zm)] = pack16m<OM>(Fo[u][m]);
              }
            }
          }
        };
        if (out_mode == 1) store16(IntC<1>());
        else store16(IntC<2>());
        };
        store_band(band, Fo);
        if constexpr (PAIRABLE) {
          if (fpair) store_band(band2, Fo2);
        }
      }
    }
  }
}

// Raise a kernel's dynamic-LDS limit once where a launch needs more than the default.
template <typename K>
static int raise_lds(K kern, size_t lds, size_t* raised) {
  if (lds > 65536 && lds > *raised) {
    if (hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
      return -1;
    *raised = lds;
  }
  return 0;
}

}  // namespace

namespace {

// ---- four-step F-engine (FFT lengths 16384 and 32768) -------------------------------------

constexpr int BM = 2048;          // sub-FFT length
constexpr int BNT = BM / 8;       // threads of the workgroup
constexpr int BTWL = FXG_MAX_BIG_L / 64 + 64;   // LDS entries of the W_L factor tables

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

// Four-step FFTs of datastream d: FFTs lt0 + g, lt0 + g + step, ... below lt_end, where g is
// the thread's group of 256 (NTB / 256 groups per workgroup, each with its own exchange buffer,
// phasor tables and scratch row `scr + g * L`).  Twiddles of the 2048-point sub-FFTs are loaded
// into LDS first when load_tw is set.
template <int L, bool CPLX, int ORDER, int NTB>
__device__ void fb_range(const FengineArgs& args, const float2* __restrict__ tw2048, const float2* __restrict__ twL,
                         int d, int lt0, int lt_end, int step, float2* smem, float2* __restrict__ scr_base,
                         bool load_tw) {
  constexpr int R = L / BM;
  constexpr int KEEP = CPLX ? R : R / 2;            // k2 values kept per k1
  constexpr int G = NTB / BNT;
  using P = Plan<BM>;
  const int tid = threadIdx.x;
  const int g = tid / BNT, t = tid - g * BNT;
  float2* sbuf = smem + g * sbuf_entries(BM);
  float2* stw = smem + G * sbuf_entries(BM);
  float2* stl = stw + P::TWN;                     // W_L^(64 jh), jh < L/64, then W_L^jl, jl < 64
  float2* srot = stl + BTWL + g * args.phcap;
  float2* sfrac = stl + BTWL + (G + g) * args.phcap;
  float* squad = reinterpret_cast<float*>(stl + BTWL + 2 * G * args.phcap) + g * args.phcap;
  float* slut8 = reinterpret_cast<float*>(stl + BTWL + 2 * G * args.phcap) + G * args.phcap;
  float2* __restrict__ scr = scr_base + (long long)g * L;
  if (load_tw) {
    __syncthreads();                              // earlier users of the twiddle tables are done
    for (int m = tid; m < P::TWN; m += NTB) stw[m] = tw2048[m];
    for (int m = tid; m < L / 64; m += NTB) stl[m] = twL[64 * m];
    for (int m = tid; m < 64; m += NTB) stl[L / 64 + m] = twL[m];
    for (int m = tid; m < 256; m += NTB) slut8[m] = lut8(m);    // 8-bit levels, once per workgroup
    for (int m = tid; m < 8; m += NTB) slut8[256 + m] = lut2(m < 4 ? FXG_FMT_VDIF : FXG_FMT_MARK5B, m & 3);
  }
  const CDevDs dsp = (CDevDs)(args.ds + d);
  float* scf = slut8 + 264;                       // staged channel-frequency tables
  stage_cf(args, dsp, scf, tid, NTB);
  DsConst k;
  k.C = dsp->C; k.S = dsp->S; k.NS = dsp->NS; k.NFS = dsp->NFS; k.nrec = dsp->nrecbands;
  k.nrecfreqs = dsp->nrecfreqs; k.fmt = dsp->format; k.spf = dsp->spf;
  k.framebytes = dsp->framebytes; k.headerbytes = dsp->headerbytes; k.nbits = dsp->nbits;
  k.fractionallo = dsp->fractionallo; k.sampletime_us = dsp->sampletime_us; k.dsb = dsp->dsb;
  const long long spec_off32 = dsp->spec_off;
  const GBytes fv = (GBytes)(args.fvalid + dsp->fvalid_off);
  const int2* __restrict__ fz = args.fzone + dsp->fvalid_off;
  const float* __restrict__ cf = scf;
  const GBytes data = (GBytes)args.run[d].data;
  const int nframes = args.run[d].nframes;
  const FftDelay* __restrict__ params = args.params + (long long)d * args.chunkcap;
  const int C = k.C, S = k.S;
  const bool twobit = (k.nbits == 2);
  const float xsc = dsp->xsc;
  const double inv_spf = 1.0 / (double)k.spf;
  const float* __restrict__ slev2 = slut8 + 256 + (k.fmt == FXG_FMT_MARK5B ? 4 : 0);   // 2-bit levels
  const int out_mode = args.out_mode;
  const int cap = args.chunkcap;

  for (int ltb = lt0; ltb < lt_end; ltb += step * G) {
    const int lt = ltb + g * step;
    const bool active = lt < lt_end;              // a group past the end computes but stores nothing
    FftDelay fd;
    if (active) {
      fd = params[lt];
    } else {
      fd.ok = 0; fd.n0 = 0; fd.fse = 0.0f; fd.intdelay = 0; fd.a = 0; fd.b = 0; fd.c = 0;
      fd.weight = 0.0f; fd.intwall = 0; fd.fracwall = 0;
    }
    const bool fok_all = fd.ok != 0;               // flagged FFTs: zero spectra (mode.cpp:656-669)
    const float fs = fok_all ? xsc : 0.0f;         // output scale of this FFT's spectra
    for (int f = 0; f < k.nrecfreqs; ++f) {
      const FreqConst rf = load_freq(dsp->rf[f]);
      if (rf.nbands == 0) continue;
      __syncthreads();                            // previous users of srot/sfrac are done
      if (ORDER != 0) build_rotator<ORDER>(srot, squad, k, rf, fd, L, t, BNT);
      build_frac<ORDER>(sfrac, k, rf, CPLX, cf, fd, dsp->rf[f].bhd[0], make_float2(dsp->rf[f].pcre[0], dsp->rf[f].pcim[0]), t, BNT);
      __syncthreads();
      for (int kb = 0; kb < rf.nbands; ++kb) {
        const int band = dsp->rf[f].band[kb];
        if (kb > 0 && dsp->rf[f].newfrac[kb]) {  // another pol's phasor table (uniform branch;
          // the previous band ended on a barrier after its last phasor read)
          build_frac<ORDER>(sfrac, k, rf, CPLX, cf, fd, dsp->rf[f].bhd[kb], make_float2(dsp->rf[f].pcre[kb], dsp->rf[f].pcim[kb]), t, BNT);
          __syncthreads();
        }
        // 2-bit real data of an FFT whose samples are all valid (weight 1: every frame of the
        // window present and valid, mode.cpp:656-669 and the good-sample count of
        // mk5mode.cpp:94-174): each thread first gathers this band's fields of its R
        // consecutive samples R*m .. R*m + R - 1 of each of its eight m = t + 256 s into one
        // 2R-bit word.  The R loads of a word are issued back to back and neighbouring lanes
        // read neighbouring bytes, so the loads hit in L1; the sub-FFT loop then decodes from
        // registers.  Other FFTs decode with per-sample checks.
        const bool fast = !CPLX && twobit && fok_all && fd.weight == 1.0f;
        uint32_t pk[8];
        if (fast) {
          // the window start, opaque to the compiler in each band: otherwise it keeps the 8 R
          // band-independent address terms live across the band loop (and spills them)
          unsigned n0b = (unsigned)fd.n0;
          asm volatile("" : "+v"(n0b));
#pragma unroll
          for (int s = 0; s < 8; ++s) {
            const unsigned gs0 = n0b + (unsigned)(R * (t + s * BNT));
            unsigned fr0 = (unsigned)((double)gs0 * inv_spf);     // exact: see fe_range
            unsigned w0 = gs0 - fr0 * (unsigned)k.spf;
            if (w0 >= (unsigned)k.spf) { w0 -= k.spf; ++fr0; }
            uint32_t acc = 0u;
#pragma unroll
            for (int i = 0; i < R; ++i) {
              unsigned wi = w0 + i, fi = fr0;
              if (wi >= (unsigned)k.spf) { wi -= k.spf; ++fi; }   // at most one frame edge (spf >= R)
              const unsigned q = wi * (unsigned)k.nrec + (unsigned)band;
              const uint32_t w = *(GWords)(data + fi * (unsigned)k.framebytes + (unsigned)k.headerbytes + ((q >> 4) << 2));
              acc |= ((w >> (2 * (q & 15u))) & 3u) << (2 * i);
            }
            pk[s] = acc;
            __builtin_amdgcn_sched_barrier(0);    // R loads in flight at a time, not 8 R
          }
        }
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
          // the rotator of sample s (in increasing s: order 2 steps its quadratic term)
          auto rot_next = [&](int s) {
            const int st = (nb0 + s * R * BNT) / S;
            if (ORDER == 1) return cmul(sub, srot[S + st]);   // step 0's phasor is (1, -0)
            if (ORDER == 2) {
              const float2 pq = cmul(sub, qpow);
              qpow = cmul(qpow, qstep);
              return cmul(pq, srot[S + st]);
            }
            return sub;
          };
          if (fast) {
#pragma unroll
            for (int s = 0; s < 8; ++s) {
              const float2 rot = rot_next(s);
              const float x = slev2[(pk[s] >> (2 * r)) & 3u];
              v[0][s] = ORDER == 0 ? make_float2(x, 0.0f) : make_float2(x * rot.x, x * rot.y);
            }
          } else {
            const unsigned g0 = (unsigned)fd.n0 + (unsigned)nb0;
#pragma unroll
            for (int s = 0; s < 8; ++s) {
              const unsigned gs = g0 + (unsigned)(s * R * BNT);
              unsigned fr = (unsigned)((double)gs * inv_spf);
              unsigned within = gs - fr * (unsigned)k.spf;
              if (within >= (unsigned)k.spf) { within -= k.spf; ++fr; }
              const bool ok = fok_all && ((int)fr < nframes) && sample_ok(fv[fr], fz, fr, within);
              const unsigned fbase = fr * (unsigned)k.framebytes + (unsigned)k.headerbytes;
              const float2 rot = rot_next(s);
              if (!CPLX) {
                float x = 0.0f;
                if (ok) {
                  if (twobit) {
                    const unsigned q = within * (unsigned)k.nrec + (unsigned)band;
                    const uint32_t w = *(GWords)(data + fbase + ((q >> 4) << 2));
                    x = slev2[(w >> (2 * (q & 15u))) & 3u];
                  } else {
                    x = slut8[data[fbase + within * (unsigned)k.nrec + band]];
                  }
                }
                v[0][s] = ORDER == 0 ? make_float2(x, 0.0f) : make_float2(x * rot.x, x * rot.y);
              } else {
                float2 x = make_float2(0.f, 0.f);
                if (ok) {
                  const GBytes pp = data + fbase + (within * (unsigned)k.nrec + band) * 2u;
                  x = make_float2(slut8[pp[0]], slut8[pp[1]]);
                }
                v[0][s] = cmul(rot, x);
                if (rf.lsb) v[0][s].y = -v[0][s].y;   // conj input (see fe_range)
              }
            }
          }
          __syncthreads();                        // sbuf free
          fft_stockham<BM, 1, BNT>(v, sbuf, stw, t);
          // Z_r[k1] = Y_r[k1] W_L^(r k1), the twiddle as W_L^(64 jh) W_L^jl with r k1 = 64 jh + jl
#pragma unroll
          for (int s = 0; s < 8; ++s) {
            const int j = r * (t + s * BNT);
            v[0][s] = cmul(v[0][s], cmul(stl[j >> 6], stl[L / 64 + (j & 63)]));
          }
#pragma unroll
          for (int s = 0; s < 8; ++s) scr[r * BM + t + s * BNT] = v[0][s];
        }
        __syncthreads();                          // every Z_r of this band is in the scratch row
        // the second step and the stores; the output form is tested once per band
        auto finish = [&](auto mode_c) {
          constexpr int OM = decltype(mode_c)::value;   // 0: FP32, 1: FP16, 2: BF16
          float2* __restrict__ o32 = nullptr;
          uint32_t* __restrict__ o16 = nullptr;
          int twl = 3;
          if (OM == 0) {
            o32 = args.spec32 + spec_off32 + ((long long)band * cap + lt) * C;
          } else {
            const CDevXGroup xgr = ((CDevXGroup)args.xgroups) + dsp->band_xfreq[band];
            twl = xgr->twlog;
            o16 = args.spec16 + xgr->spec_off + (long long)dsp->band_xrow[band] * xgr->rowstride + ((long long)lt << twl);
          }
          auto load_z = [&](float2* z, int s) {
#pragma unroll
            for (int r = 0; r < R; ++r) z[r] = scr[r * BM + t + s * BNT];
          };
          auto emit = [&](float2* z, int s) {
            const int k1 = t + s * BNT;
            dftR<R>(z);                           // z[k2] = X[k1 + 2048 k2]
#pragma unroll
            for (int j = 0; j < KEEP; ++j) {
              // channel c = k1 + 2048 j; real LSB keeps X[C + c]; double sideband swaps halves
              // (compile-time register indices: one select between two candidates)
              const bool upper = CPLX ? (k.dsb != 0) : (rf.lsb != 0);
              const int c = k1 + BM * j;
              float2 X = upper ? z[(j + R / 2) % R] : z[j];
              if (CPLX && rf.lsb && !k.dsb && c == 0) X.y = -X.y;   // DC is not flipped
              // fractional-sample phasor times the output scale (0 for a flagged FFT)
              const float2 fsub0 = sfrac[c & (S - 1)];
              const float2 frq = cmul(make_float2(fsub0.x * fs, fsub0.y * fs), sfrac[S + c / S]);
              const float2 F = cmul(X, frq);
              if (!active) continue;
              if (OM == 0) {
                o32[c] = F;
              } else {
                o16[(((long long)(c >> twl) * cap) << twl) + (c & ((1 << twl) - 1))] = pack16m<OM>(F);
                for (int zz = 0; zz < dsp->nzout; ++zz) {          // zoom windows of this band
                  if (dsp->zout_band[zz] != band) continue;
                  const int cz = c - dsp->zout_c0[zz];
                  if (cz < 0 || cz >= dsp->zout_n[zz]) continue;
                  const CDevXGroup zg = ((CDevXGroup)args.xgroups) + dsp->zout_group[zz];
                  const int zl = zg->twlog;
                  uint32_t* __restrict__ oz = args.spec16 + zg->spec_off + (long long)dsp->zout_row[zz] * zg->rowstride +
                                              ((long long)lt << zl);
                  oz[(((long long)(cz >> zl) * cap) << zl) + (cz & ((1 << zl) - 1))] = pack16m<OM>(F);
                }
              }
            }
          };
          // (a second register buffer that loads block s + 1 before the stores of block s
          // measured no faster: it costs the third wave per SIMD)
#pragma unroll 1
          for (int s = 0; s < 8; ++s) {
            float2 z[R];
            load_z(z, s);
            emit(z, s);
          }
        };
        if (out_mode == 0) finish(IntC<0>());
        else if (out_mode == 1) finish(IntC<1>());
        else finish(IntC<2>());
        __syncthreads();                          // scratch row free for the next band
      }
    }
  }
}

// LDS bytes of the four-step F-engine with NTB threads (NTB / 256 FFTs in flight).
template <int NTB>
size_t fb_lds_bytes(int phcap) {
  constexpr int G = NTB / BNT;
  return sizeof(float2) * ((size_t)G * sbuf_entries(BM) + Plan<BM>::TWN + BTWL + 2 * (size_t)G * phcap) +
         sizeof(float) * ((size_t)G * phcap + 264) + stage_lds_bytes(phcap);
}

}  // namespace


