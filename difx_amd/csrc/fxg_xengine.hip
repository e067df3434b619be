// X-engines and accumulation of the FX hot path on CDNA4.
//
// Plain FP32 path: every product accumulates
//   acc += F1[t][c] * conj(F2[t][c])
// over the FFTs of the subint in FP32, the vectorAddProduct_cf32 of Core::processdata
// (mpifxcorr/src/core.cpp:884-979, :963) for the cross products and of Mode::process for the
// autocorrelations (mode.cpp:1272, 1333-1339).  The blocked kernel takes 8 x 8 blocks of the
// input x input matrix of the products' inputs (FFT order within each quarter of a chunk, the
// quarters in order); the per-product kernel (one thread per product and channel, FFT order)
// takes the products it leaves.
//
// MFMA path: per channel the visibilities of all inputs (station bands of one frequency)
// form the Hermitian matrix V = sum_t x_t x_t^H.  With x stored as interleaved (re, im)
// 16-bit pairs along time, the row of input n is A_n = (re_0, im_0, re_1, im_1, ...), so
//   Re V = A A^T            and   Im V = B A^T,  B_n = (im_0, -re_0, im_1, -re_1, ...)
// and both are 32x32x16 MFMA contractions whose A and B operands come from the same
// 16-byte load per lane.  FP32 accumulators live in registers for a whole chunk.
#include "fxg_internal.h"
#include <hip/hip_fp16.h>

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));


// Thread per (product, channel).  Parallel-hand autocorrelations keep a zero imaginary part,
// as F * conj(F) has in the reference's arithmetic.
__global__ void __launch_bounds__(256) k_xmac_f32(XPlainArgs a) {
  const DevProd p = a.prods[blockIdx.y];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= p.nchan) return;
  const DevDs& d1 = a.ds[p.ds1];
  const DevDs& d2 = a.ds[p.ds2];
  const int C1 = d1.C, C2 = d2.C;
  const float2* s1 = a.spec32 + d1.spec_off + (long long)p.band1 * a.chunkcap * C1 + p.choff1 + c;
  const float2* s2 = a.spec32 + d2.spec_off + (long long)p.band2 * a.chunkcap * C2 + p.choff2 + c;
  const bool isauto = (p.ds1 == p.ds2) && (p.band1 == p.band2) && (p.choff1 == p.choff2);
  float2 acc = a.acc[p.acc_off + c];
  int t = 0;
  for (; t + 4 <= a.nffts; t += 4) {
    float2 x[4], y[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) { x[u] = s1[(long long)(t + u) * C1]; y[u] = s2[(long long)(t + u) * C2]; }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      acc.x += x[u].x * y[u].x + x[u].y * y[u].y;
      acc.y += x[u].y * y[u].x - x[u].x * y[u].y;
    }
  }
  for (; t < a.nffts; ++t) {
    const float2 x = s1[(long long)t * C1], y = s2[(long long)t * C2];
    acc.x += x.x * y.x + x.y * y.y;
    acc.y += x.y * y.x - x.x * y.y;
  }
  if (isauto) acc.y = 0.0f;
  a.acc[p.acc_off + c] = acc;
}


// Blocked form of the plain FP32 X-engine: one workgroup per work item (an 8 x 8 block of a
// group's row x row matrix over 64 channels, one channel per lane).  Each of the four
// wavefronts sums X_i conj(X_j) over a quarter of the chunk's FFTs in FFT order, loading 16
// spectra values per FFT for 64 products; wave 0 then adds the four partial sums in wave order
// and adds the result to each product's accumulator: product (i, j) takes the block entry,
// product (j, i) its conjugate (off the diagonal blocks).  Autocorrelations (i = j) keep a zero
// imaginary part.  Rows past the group and channels past C are clamped for the loads and never
// written.  Items with group -1 are padding and do nothing.
__global__ void __launch_bounds__(256) k_xmac_f32_blk(XPBlkArgs a) {
  __shared__ float2 red[3][16][64];
  const XPItem it = a.items[blockIdx.x];
  if (it.group < 0) return;                      // padding item: the whole workgroup leaves
  const XPGroup g = a.groups[it.group];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // wave-uniform: scalar loop and pointers
  const int c = it.c0 + lane;
  const int cl = min(c, g.C - 1);
  // the row pointers are the same for every lane (scalar registers); each lane adds its channel
  const int t0 = (int)((long long)w * a.nffts / 4), t1 = (int)((long long)(w + 1) * a.nffts / 4);
  // 32-bit byte offsets from the spectra base (the host keeps the FP32 spectra below 4 GiB
  // for this kernel), so each load is the scalar base plus one offset register
  typedef float v2f __attribute__((ext_vector_type(2)));
  typedef const __attribute__((address_space(1))) v2f* GF2;
  typedef const __attribute__((address_space(1))) char* GB;
  // The rows of a group share one FFT stride, so the row offsets are the same for every lane
  // (scalar registers) and each load adds the lane's channel offset.
  const GB base = (GB)a.spec32;
  unsigned ox[8], oy[8];
  const unsigned st = 8u * (unsigned)g.stride;
  const unsigned co = (unsigned)cl * 8u;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    ox[i] = (unsigned)(8LL * (a.rows[g.row0 + min(it.rb * 8 + i, g.nrows - 1)].off + (long long)t0 * g.stride));
    oy[i] = (unsigned)(8LL * (a.rows[g.row0 + min(it.cb * 8 + i, g.nrows - 1)].off + (long long)t0 * g.stride));
  }
  float2 acc[8][8];
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[i][j] = make_float2(0.0f, 0.0f);
  for (int t = t0; t < t1; ++t) {
    float2 x[8], y[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const v2f vx = *(GF2)(base + (ox[i] + co));
      const v2f vy = *(GF2)(base + (oy[i] + co));
      x[i] = make_float2(vx.x, vx.y);
      y[i] = make_float2(vy.x, vy.y);
      ox[i] += st;
      oy[i] += st;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        acc[i][j].x = fmaf(x[i].x, y[j].x, acc[i][j].x);
        acc[i][j].x = fmaf(x[i].y, y[j].y, acc[i][j].x);
        acc[i][j].y = fmaf(x[i].y, y[j].x, acc[i][j].y);
        acc[i][j].y = fmaf(-x[i].x, y[j].y, acc[i][j].y);
      }
  }
  // the partial sums of waves 1-3 through LDS, 16 entries per round
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    __syncthreads();
    if (w > 0) {
#pragma unroll
      for (int k = 0; k < 16; ++k) red[w - 1][k][lane] = acc[(r * 16 + k) / 8][(r * 16 + k) % 8];
    }
    __syncthreads();
    if (w == 0) {
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        float2 v = acc[(r * 16 + k) / 8][(r * 16 + k) % 8];
#pragma unroll
        for (int q = 0; q < 3; ++q) { v.x += red[q][k][lane].x; v.y += red[q][k][lane].y; }
        acc[(r * 16 + k) / 8][(r * 16 + k) % 8] = v;
      }
    }
  }
  if (w != 0 || c >= g.C) return;
  const int n = g.nrows;
  const int* __restrict__ pairs = a.pairs + g.pair0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int ri = it.rb * 8 + i;
    if (ri >= n) break;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int cj = it.cb * 8 + j;
      if (cj >= n) break;
      const int p = pairs[ri * n + cj];
      if (p >= 0) {
        float2 v = acc[i][j];
        if (ri == cj) v.y = 0.0f;
        float2* d = a.acc + p + c;
        *d = make_float2(d->x + v.x, d->y + v.y);
      }
      if (it.rb != it.cb) {
        const int q = pairs[cj * n + ri];
        if (q >= 0) {
          float2* d = a.acc + q + c;
          *d = make_float2(d->x + acc[i][j].x, d->y - acc[i][j].y);
        }
      }
    }
  }
}

template <typename V8>
__device__ __forceinline__ floatx16 mfma32(V8 a, V8 b, floatx16 c);
template <>
__device__ __forceinline__ floatx16 mfma32<half8_t>(half8_t a, half8_t b, floatx16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}
template <>
__device__ __forceinline__ floatx16 mfma32<bf16x8_t>(bf16x8_t a, bf16x8_t b, floatx16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// (re, im) pairs -> (im, -re) pairs: flip the sign bit of the 16-bit value moved to the
// high half.  Same bit manipulation for FP16 and BF16 (sign is bit 15).
__device__ __forceinline__ uint4 rot_negi(uint4 u) {
  uint4 r;
  r.x = ((u.x >> 16) | (u.x << 16)) ^ 0x80000000u;
  r.y = ((u.y >> 16) | (u.y << 16)) ^ 0x80000000u;
  r.z = ((u.z >> 16) | (u.z << 16)) ^ 0x80000000u;
  r.w = ((u.w >> 16) | (u.w << 16)) ^ 0x80000000u;
  return r;
}

// One workgroup = one tile of CT channels (8 or 16, the tile width of the group's spectra
// layout) of one group; each wavefront owns CPW channels (2 for one 32-row tile, 1 for two).
// The F-engine writes the 16-bit spectra channel-tile-blocked ([row][C/CT][t][CT]), so per
// k-step (8 FFTs) every input row contributes one contiguous run of 32 CT bytes.  The workgroup
// streams those runs into LDS as [channel][row][8 FFTs], double buffered, so each lane reads its
// MFMA operand (one row, 4 consecutive FFTs = 8 halves) with one 16-byte LDS load.  Rows beyond
// the group's inputs stay zero in LDS and are never fetched.
template <int NTILE, int CT_>
struct XTile {
  static constexpr int CPW = (NTILE == 1) ? 2 : 1;      // channels per wavefront
  static constexpr int CT = CT_;                         // channels per workgroup
  static constexpr int NW = CT / CPW;                    // wavefronts per workgroup
  static constexpr int NTH = 64 * NW;
  static constexpr int R = 32 * NTILE;                   // operand rows
  static constexpr int CSTR = R * 8 + 8;                 // 32-bit words per channel in LDS (padded:
                                                         // conflict-free chunk stores)
  static constexpr int QPR = 2 * CT;                     // 16-byte chunks per row per k-step
  static constexpr int QMAX = (R * QPR + NTH - 1) / NTH; // chunks per thread per k-step
  static_assert(NTH <= 1024, "workgroup size");
};

template <typename V8, int NTILE, int CT_>
__global__ void __launch_bounds__((XTile<NTILE, CT_>::NTH)) k_xmac_mfma(XMfmaArgs a) {
  using T = XTile<NTILE, CT_>;
  constexpr int NPAIR = NTILE * (NTILE + 1) / 2;
  constexpr int CPW = T::CPW, CT = T::CT, CSTR = T::CSTR, QPR = T::QPR, QMAX = T::QMAX, NTH = T::NTH;
  __shared__ uint4 lds_raw[2][(CT * CSTR) / 4];
  uint32_t* lds0 = reinterpret_cast<uint32_t*>(lds_raw[0]);
  uint32_t* lds1 = reinterpret_cast<uint32_t*>(lds_raw[1]);
  const DevXGroup g = a.groups[blockIdx.y];
  const int ct = blockIdx.x;
  const int c0 = ct * CT;
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int row = lane & 31, half = lane >> 5;
  const int nq = g.nin * QPR;
  const uint32_t* gbase = a.spec16 + g.spec_off + (long long)ct * a.chunkcap * CT;

  for (int i = tid; i < 2 * (CT * CSTR) / 4; i += NTH)
    (&lds_raw[0][0])[i] = make_uint4(0u, 0u, 0u, 0u);

  floatx16 accr[CPW][NPAIR], acci[CPW][NPAIR];
#pragma unroll
  for (int cw = 0; cw < CPW; ++cw) {
    const int c = c0 + wave * CPW + cw;
    const float* accp = a.acc + g.acc_off + (long long)c * NPAIR * 2 * 1024;
#pragma unroll
    for (int p = 0; p < NPAIR; ++p) {
      const float* pr = accp + (p * 2 + 0) * 1024 + lane;    // immediate offsets r * 256 bytes
      const float* pi = accp + (p * 2 + 1) * 1024 + lane;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        accr[cw][p][r] = pr[r * 64];
        acci[cw][p][r] = pi[r * 64];
      }
    }
  }

  // chunk q of a k-step: row r = q / QPR, FFT t = (q % QPR) / (CT/4), channel quad
  // h = q % (CT/4).  The loads are unconditional (a chunk past the group's inputs re-reads
  // row 0 and is not stored), so the compiler can count them exactly in its waits.
  constexpr int QPC = CT / 4;                    // channel quads per FFT
  auto gload = [&](int k, uint4* b) {
#pragma unroll
    for (int i = 0; i < QMAX; ++i) {
      const int q0 = tid + NTH * i;
      const int q = q0 < nq ? q0 : (q0 % QPR);
      const int r = q / QPR, rem = q % QPR;
      b[i] = *reinterpret_cast<const uint4*>(gbase + (long long)r * g.rowstride + (long long)(8 * k) * CT + rem * 4);
    }
  };
  auto lstore = [&](uint32_t* L, const uint4* b) {
#pragma unroll
    for (int i = 0; i < QMAX; ++i) {
      const int q = tid + NTH * i;
      if (q < nq) {
        const int r = q / QPR, rem = q % QPR;
        const int t = rem / QPC, h = rem % QPC;
        uint32_t* d = L + (4 * h) * CSTR + r * 8 + t;
        d[0] = b[i].x; d[CSTR] = b[i].y; d[2 * CSTR] = b[i].z; d[3 * CSTR] = b[i].w;
      }
    }
  };
  auto mfma_step = [&](const uint32_t* L) {
#pragma unroll
      for (int cw = 0; cw < CPW; ++cw) {
        const uint32_t* cb = L + (wave * CPW + cw) * CSTR + 4 * half;
        uint4 x[NTILE];
  #pragma unroll
        for (int ti = 0; ti < NTILE; ++ti) x[ti] = *reinterpret_cast<const uint4*>(cb + (row + 32 * ti) * 8);
        int pi = 0;
  #pragma unroll
        for (int ti = 0; ti < NTILE; ++ti) {
          const uint4 xr = rot_negi(x[ti]);
  #pragma unroll
          for (int tj = 0; tj <= ti; ++tj) {
            const V8 av = __builtin_bit_cast(V8, x[ti]);
            const V8 bv = __builtin_bit_cast(V8, x[tj]);
            const V8 arv = __builtin_bit_cast(V8, xr);
            accr[cw][pi] = mfma32<V8>(av, bv, accr[cw][pi]);
            acci[cw][pi] = mfma32<V8>(arv, bv, acci[cw][pi]);
            ++pi;
          }
        }
      }
  };

  const int nk = a.nffts_pad / 8;
  // one k-step of global loads in flight over a double-buffered LDS stage (two k-steps ahead
  // in registers measured slower)
  uint4 buf[QMAX];
  __syncthreads();
  gload(0, buf);
  lstore(lds0, buf);
  __syncthreads();
  for (int k = 0; k < nk; ++k) {
    uint32_t* cur = (k & 1) ? lds1 : lds0;
    uint32_t* nxt = (k & 1) ? lds0 : lds1;
    gload(min(k + 1, nk - 1), buf);               // past the end: re-read, unused
    mfma_step(cur);
    if (k + 1 < nk) lstore(nxt, buf);
    __syncthreads();
  }

#pragma unroll
  for (int cw = 0; cw < CPW; ++cw) {
    const int c = c0 + wave * CPW + cw;
    float* accp = a.acc + g.acc_off + (long long)c * NPAIR * 2 * 1024;
#pragma unroll
    for (int p = 0; p < NPAIR; ++p) {
      float* pr = accp + (p * 2 + 0) * 1024 + lane;
      float* pi = accp + (p * 2 + 1) * 1024 + lane;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        pr[r * 64] = accr[cw][p][r];
        pi[r * 64] = acci[cw][p][r];
      }
    }
  }
}

// Register r of lane l of a 32x32 MFMA accumulator holds element
//   (i, j) = (8*(r/4) + 4*(l/32) + r%4, l%32).
__device__ __forceinline__ int mfma_idx(int i, int j) {
  const int r = (i >> 3) * 4 + (i & 3);
  const int l = ((i >> 2) & 1) * 32 + j;
  return r * 64 + l;
}

// ---- accumulation into the slot results (core.cpp:1819-1853, 1258-1376) ----------------

// The uv-shift rotation of channel ch of a freq to a phase centre at applied delay dd (us),
// in the reference's steps (core.cpp:1699-1766, 1789-1808): FP64 turns of the sub phasor
// (channel within its rotate stride, plus the band edge) and of the step phasor (the stride),
// wrapped, as FP32 arguments of FP32 sin/cos; the value times the sub phasor, then the step.
constexpr double UV_TWO_PI = 6.283185307179586476925286766559;   // architecture.h.in:430

__device__ __forceinline__ float2 uv_rotate(float2 v, double dd, const UvFreq& q, int ch) {
  const int c = ch % q.rs, j = ch / q.rs;
  double edge = dd * q.lofreq;
  edge -= floor(edge);
  const double subf = q.lsb ? -(double)(q.rs - (c + 1)) * q.chbw : (double)c * q.chbw;
  const double stepf = q.lsb ? -(double)(q.nsteps - (j + 1)) * q.stepbw : (double)j * q.stepbw;
  const double t1 = dd * subf + edge;
  const double t2 = dd * stepf;
  const float a1 = (float)((t1 - floor(t1)) * UV_TWO_PI);
  const float a2 = (float)((t2 - floor(t2)) * UV_TWO_PI);
  float s1, c1, s2, c2;
  sincosf(a1, &s1, &c1);
  sincosf(a2, &s2, &c2);
  const float2 r = make_float2(c1 * v.x - s1 * v.y, c1 * v.y + s1 * v.x);
  return make_float2(r.x * c2 - r.y * s2, r.x * s2 + r.y * c2);
}

__global__ void k_fin_plain(const FinPlain* e, const float2* acc, float2* res, UvArgs uv) {
  const FinPlain f = e[blockIdx.y];
  const int co = blockIdx.x * blockDim.x + threadIdx.x;
  if (co >= f.nout) return;
  if (uv.P > 1 && f.bl >= 0) {                    // rotated into every phase centre's block
    const UvFreq q = uv.fq[f.fq];
    for (int p = 0; p < uv.P; ++p) {
      const double dd = uv.dd[f.bl * uv.P + p];
      const bool turn = fabs(dd) > 1.0e-20;
      float2 s = make_float2(0.f, 0.f);
      for (int k = 0; k < f.avg; ++k) {
        const int ch = co * f.avg + k;
        float2 v = acc[f.src + ch];
        if (turn) v = uv_rotate(v, dd, q, ch);
        s.x += v.x; s.y += v.y;
      }
      if (f.avg > 1) { s.x /= (float)f.avg; s.y /= (float)f.avg; }
      float2* o = res + f.dst + (long long)p * f.pcstride + co;
      o->x += s.x; o->y += s.y;
    }
    return;
  }
  float2 s = make_float2(0.f, 0.f);
  for (int k = 0; k < f.avg; ++k) {
    const float2 v = acc[f.src + (long long)co * f.avg + k];
    s.x += v.x; s.y += v.y;
  }
  if (f.avg > 1) { s.x /= (float)f.avg; s.y /= (float)f.avg; }
  float2 o = res[f.dst + co];
  o.x += s.x; o.y += s.y;
  res[f.dst + co] = o;
}


__global__ void k_fin_mfma(const FinMfma* e, const DevXGroup* groups, const float* acc,
                           const float* xscale, float2* res, UvArgs uv) {
  const FinMfma f = e[blockIdx.y];
  const int co = blockIdx.x * blockDim.x + threadIdx.x;
  if (co >= f.nout) return;
  const DevXGroup g = groups[f.group];
  const int npair = g.ntile * (g.ntile + 1) / 2;
  int r1 = f.row1, r2 = f.row2;
  float sign = 1.0f;
  int ti = r1 >> 5, tj = r2 >> 5;
  if (ti < tj) { int tmp = r1; r1 = r2; r2 = tmp; ti = r1 >> 5; tj = r2 >> 5; sign = -1.0f; }
  const int pair = ti * (ti + 1) / 2 + tj;
  const int idx = mfma_idx(r1 & 31, r2 & 31);
  if (uv.P > 1 && f.bl >= 0) {                    // rotated into every phase centre's block
    const UvFreq q = uv.fq[f.fq];
    const float inv = 1.0f / (xscale[g.scale_off + f.row1] * xscale[g.scale_off + f.row2]);
    for (int p = 0; p < uv.P; ++p) {
      const double dd = uv.dd[f.bl * uv.P + p];
      const bool turn = fabs(dd) > 1.0e-20;
      float2 s = make_float2(0.f, 0.f);
      for (int k = 0; k < f.avg; ++k) {
        const int ch = co * f.avg + k;
        const float* base = acc + g.acc_off + ((long long)(f.chan0 + ch) * npair + pair) * 2 * 1024;
        float2 v = make_float2(base[idx] * inv, sign * base[1024 + idx] * inv);
        if (turn) v = uv_rotate(v, dd, q, ch);
        s.x += v.x; s.y += v.y;
      }
      if (f.avg > 1) { s.x /= (float)f.avg; s.y /= (float)f.avg; }
      float2* o = res + f.dst + (long long)p * f.pcstride + co;
      o->x += s.x; o->y += s.y;
    }
    return;
  }
  float2 s = make_float2(0.f, 0.f);
  for (int k = 0; k < f.avg; ++k) {
    const int c = f.chan0 + co * f.avg + k;
    const float* base = acc + g.acc_off + ((long long)c * npair + pair) * 2 * 1024;
    const float sc = xscale[g.scale_off + f.row1] * xscale[g.scale_off + f.row2];
    const float inv = 1.0f / sc;
    s.x += base[idx] * inv;
    s.y += sign * base[1024 + idx] * inv;
  }
  if (f.avg > 1) { s.x /= (float)f.avg; s.y /= (float)f.avg; }
  if (f.isauto) s.y = 0.0f;
  float2 o = res[f.dst + co];
  o.x += s.x; o.y += s.y;
  res[f.dst + co] = o;
}

// Baseline weights (core.cpp:1005-1052, 1066-1112) and autocorrelation weights
// (core.cpp:1306-1370): sequential FP32 sums over the FFTs of the subint.

// One workgroup per entry: a strided partial sum per thread, then a tree over the block.
__global__ void __launch_bounds__(256) k_fin_weights(const FinWeight* e, const float* w, int B, float* resf) {
  __shared__ float part[256];
  const FinWeight f = e[blockIdx.x];
  const float* w1 = w + (long long)f.ds1 * B;
  const float* w2 = f.ds2 >= 0 ? w + (long long)f.ds2 * B : nullptr;
  float s = 0.0f;
  for (int i = threadIdx.x; i < B; i += blockDim.x) s += w2 ? w1[i] * w2[i] : w1[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) resf[f.dstf] += part[0];
}

__global__ void k_zero(float* p, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = 0.0f;
}

// Visibility::addData (visibility.cpp:192-205): one subint's results into an integration.
__global__ void k_add(float* dst, const float* src, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    dst[i] += src[i];
}

// The shift decorrelation sums of a subint into their result words (core.cpp:1089-1104).
__global__ void k_add_at(const int* idx, const float* val, int n, float* resf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) resf[idx[i]] += val[i];
}

__global__ void k_fill(float* p, long long n, float v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

// ---- host-side launchers -------------------------------------------------------------
int fxg_launch_xplain_blk(const XPBlkArgs& a, int nitems, hipStream_t st) {
  if (nitems == 0) return 0;
  hipLaunchKernelGGL(k_xmac_f32_blk, dim3(nitems), dim3(256), 0, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int fxg_launch_xplain(const XPlainArgs& a, int nprod, int maxchan, hipStream_t st) {
  dim3 grid((maxchan + 255) / 256, nprod);
  hipLaunchKernelGGL(k_xmac_f32, grid, dim3(256), 0, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <typename V8, int NTILE, int CT>
static void launch_x(const XMfmaArgs& b, int C, int ngroups, hipStream_t st) {
  hipLaunchKernelGGL((k_xmac_mfma<V8, NTILE, CT>), dim3(C / CT, ngroups), dim3(XTile<NTILE, CT>::NTH), 0, st, b);
}

// groups [group0, group0 + ngroup_launch) share the tile count, channel count and tile width
int fxg_launch_xmfma(const XMfmaArgs& a, int ngroup_launch, int group0, int ntile, int C, int twlog, hipStream_t st) {
  XMfmaArgs b = a;
  b.groups = a.groups + group0;
  const int ct = 1 << twlog;
  if (C % ct || (ct != 8 && ct != 16) || (ntile != 1 && ntile != 2) || (ntile == 2 && ct != 8)) return -2;
  const int form = (a.mode == 1 ? 0 : 4) + (ntile == 1 ? 0 : 2) + (ct == 8 ? 0 : 1);
  switch (form) {
    case 0: launch_x<half8_t, 1, 8>(b, C, ngroup_launch, st); break;
    case 1: launch_x<half8_t, 1, 16>(b, C, ngroup_launch, st); break;
    case 2: launch_x<half8_t, 2, 8>(b, C, ngroup_launch, st); break;
    case 4: launch_x<bf16x8_t, 1, 8>(b, C, ngroup_launch, st); break;
    case 5: launch_x<bf16x8_t, 1, 16>(b, C, ngroup_launch, st); break;
    case 6: launch_x<bf16x8_t, 2, 8>(b, C, ngroup_launch, st); break;
    default: return -2;
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int fxg_launch_fin_plain(const FinPlain* e, int n, int maxout, const float2* acc, float2* res, const UvArgs& uv,
                         hipStream_t st) {
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_fin_plain, dim3((maxout + 255) / 256, n), dim3(256), 0, st, e, acc, res, uv);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int fxg_launch_add_at(const int* idx, const float* val, int n, float* resf, hipStream_t st) {
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_add_at, dim3((n + 255) / 256), dim3(256), 0, st, idx, val, n, resf);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int fxg_launch_fin_mfma(const FinMfma* e, int n, int maxout, const DevXGroup* g, const float* acc,
                        const float* xscale, float2* res, const UvArgs& uv, hipStream_t st) {
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_fin_mfma, dim3((maxout + 255) / 256, n), dim3(256), 0, st, e, g, acc, xscale, res, uv);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int fxg_launch_fin_weights(const FinWeight* e, int n, const float* w, int B, float* resf, hipStream_t st) {
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_fin_weights, dim3(n), dim3(256), 0, st, e, w, B, resf);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int fxg_launch_zero(void* p, long long bytes, hipStream_t st) {
  if (bytes <= 0) return 0;
  const long long n = bytes / 4;
  long long blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, st, (float*)p, n);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int fxg_launch_add(float* dst, const float* src, long long n, hipStream_t st) {
  if (n <= 0) return 0;
  long long blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_add, dim3((unsigned)blocks), dim3(256), 0, st, dst, src, n);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int fxg_launch_fill(float* p, long long n, float v, hipStream_t st) {
  if (n <= 0) return 0;
  long long blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, st, p, n, v);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
