// VDIF thread corner turn on the device: single-channel frames of N VDIF threads (any order)
// become N-channel frames, the step before the F-engine when a station records one VDIF
// thread per channel.  Same output as VDIFMuxer::deinterlace + multiplex + cornerturn_generic
// (mpifxcorr/src/datamuxer.cpp:764-833, 670-757, 243-266) for one segment, bit for bit.
//
// Kernels, all streaming through HBM once:
//   k_vmux_init    slot table (slots x threads: input frame index + 1, 0 = missing), counters
//   k_vmux_index   one lane per input frame: header fields -> (slot, thread); a later copy of a
//                  (slot, thread) replaces an earlier one (atomic max of the frame index); the
//                  slot's presence mask gets the thread's bit
//   k_vmux_frames  workgroups of one slot: the output header by the reference's rules (a slot
//                  with no thread looks back for the last one with any) and, for complete
//                  slots, the corner-turned payload: a lane reads 8 bytes of each thread's
//                  frame at one payload position and turns them into the 2N output words
//                  (8N bytes); for N >= 4 those pass through LDS so that the workgroup's
//                  stores leave as contiguous 16-byte pieces
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <algorithm>
#include "../../include/fxg.h"

namespace {

constexpr int MUXHDR = 32;                         // VDIF_HEADER_BYTES
constexpr int MUXNT = 256;                         // lanes per workgroup

struct MuxArgs {
  int ifb, n, bits, fps;
  int map[FXG_MUX_MAX_THREADS];
  long long ref_seconds, start;
  int ref_frame;
  long long nslots, nin;
};

// Workspace: 8 int64 counters -- [0] first slot with any thread present (nslots if none),
// [1] frames before the segment, [2] after it, [3] of unmapped threads, [4] of another frame
// length, [5] nslots -- then the slot table (int32 nslots x n) and the per-slot presence
// masks (uint32 nslots, bit l = thread l present).
struct MuxWork {
  unsigned long long* ctr;
  int* table;
  unsigned* mask;
};

MuxWork work_of(void* d_work, long long nslots, int n) {
  MuxWork w;
  w.ctr = reinterpret_cast<unsigned long long*>(d_work);
  w.table = reinterpret_cast<int*>(reinterpret_cast<uint8_t*>(d_work) + 64);
  w.mask = reinterpret_cast<unsigned*>(w.table + nslots * n);
  return w;
}

__device__ __forceinline__ int ymd2doy_d(int yr, int mo, int day) {     // dateutils.c:35-53
  const int L2 = yr / 4 - (yr + 7) / 4 - yr / 100 + (yr + 99) / 100 + yr / 400 - (yr + 399) / 400;
  const int m1[12] = {0, 31, 59, 90, 120, 151, 181, 212, 243, 273, 304, 334};
  const int m2[12] = {0, 31, 60, 91, 121, 152, 182, 213, 244, 274, 305, 335};
  int s = 0;
#pragma unroll
  for (int i = 0; i < 12; ++i)
    if (i == mo - 1) s = (L2 == -1) ? m2[i] : m1[i];
  return day + s;
}

// getVDIFEpochMJD (vdifio.c:127-131): ymd2mjd of 1 January / 1 July of 2000 + epoch/2
// (dateutils.c:55-63)
__device__ __forceinline__ long long epoch_mjd_d(int epoch) {
  const int yr = 2000 + epoch / 2, mo = (epoch % 2) * 6 + 1, yr1 = yr - 1;
  return ymd2doy_d(yr, mo, 1) - 678576 + 365 * yr1 + yr1 / 4 - yr1 / 100 + yr1 / 400;
}

__device__ __forceinline__ uint4 load_16(const uint8_t* p) {             // 8-byte aligned
  const uint2 a = *reinterpret_cast<const uint2*>(p);
  const uint2 b = *reinterpret_cast<const uint2*>(p + 8);
  return make_uint4(a.x, a.y, b.x, b.y);
}

__global__ void k_vmux_init(MuxWork w, long long ntable, long long nslots) {
  const long long i0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long step = (long long)gridDim.x * blockDim.x;
  for (long long i = i0; i < ntable; i += step) w.table[i] = 0;   // table and masks
  if (i0 < 8) w.ctr[i0] = (i0 == 0 || i0 == 5) ? (unsigned long long)nslots : 0ull;
}

// 256 lanes per workgroup; the first present slot is reduced in LDS first (one global atomic
// per workgroup instead of one per frame on a single address)
__global__ void __launch_bounds__(256) k_vmux_index(MuxArgs a, const uint8_t* __restrict__ in, MuxWork w) {
  __shared__ unsigned s_first;
  if (threadIdx.x == 0) s_first = 0xFFFFFFFFu;
  __syncthreads();
  const long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (f < a.nin) {
    const uint4 h = load_16(in + f * a.ifb);
    const int seconds = (int)(h.x & 0x3FFFFFFFu);
    const int frame = (int)(h.y & 0xFFFFFFu), epoch = (int)((h.y >> 24) & 0x3Fu);
    const int fbytes = (int)(h.z & 0xFFFFFFu) * 8;
    const int tid = (int)((h.w >> 16) & 0x3FFu);
    int ti = -1;
#pragma unroll
    for (int j = 0; j < FXG_MUX_MAX_THREADS; ++j)
      if (j < a.n && ti < 0 && a.map[j] == tid) ti = j;
    const long long fullsec = epoch_mjd_d(epoch) * 86400LL + seconds;
    const long long cur = (fullsec - a.ref_seconds) * a.fps + frame - a.ref_frame - a.start;
    if (fbytes != a.ifb) {                         // the reference aborts (datamuxer.cpp:784-787)
      atomicAdd(&w.ctr[4], 1ull);
    } else if (ti < 0) {
      atomicAdd(&w.ctr[3], 1ull);
    } else if (cur < 0) {
      atomicAdd(&w.ctr[1], 1ull);
    } else if (cur >= a.nslots) {
      atomicAdd(&w.ctr[2], 1ull);
    } else {
      atomicMax(&w.table[cur * a.n + ti], (int)(f + 1));
      atomicOr(&w.mask[cur], 1u << ti);
      atomicMin(&s_first, (unsigned)cur);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0 && s_first != 0xFFFFFFFFu) atomicMin(&w.ctr[0], (unsigned long long)s_first);
}

// The output header of a slot with at least one thread present (8 words), and whether all are.
__device__ void present_header(const MuxArgs& a, const uint8_t* in, const MuxWork& w, long long s, int ofb,
                               uint32_t* hw, bool* complete) {
  const unsigned m = w.mask[s];
  const bool all = m == (a.n == 32 ? 0xFFFFFFFFu : (1u << a.n) - 1u);
  const int first = __builtin_ctz(m);              // m != 0
  *complete = all;
  const long long fi = (long long)(w.table[s * a.n + (all ? 0 : first)] - 1);
  const uint8_t* src = in + fi * a.ifb;
  const uint4 h0 = load_16(src), h1 = load_16(src + 16);
  hw[0] = h0.x; hw[1] = h0.y; hw[2] = h0.z; hw[3] = h0.w;
  hw[4] = h1.x; hw[5] = h1.y; hw[6] = h1.z; hw[7] = h1.w;
  hw[0] = (hw[0] & 0x7FFFFFFFu) | (all ? 0u : 0x80000000u);           // setVDIFFrameInvalid
  if (all) {                                                           // setVDIFNumChannels
    int lg = 0;
    for (int n = a.n; n > 1; n /= 2) ++lg;
    hw[2] = (hw[2] & ~(0x1Fu << 24)) | ((uint32_t)lg << 24);
  }
  hw[2] = (hw[2] & 0xFF000000u) | ((uint32_t)(ofb / 8) & 0xFFFFFFu);  // setVDIFFrameBytes
  hw[3] &= ~(0x3FFu << 16);                                            // setVDIFThreadID(0)
}

// k frame-number increments of a header (datamuxer.cpp:720-725): the 24-bit frame number
// goes up by one each time and, when it reaches fps, restarts at 0 in the next second.
__device__ void advance_frames(uint32_t* hw, long long k, int fps) {
  long long fn = hw[1] & 0xFFFFFFu, carry = 0;
  if (fn >= fps) {                                 // above fps: counts up to the 24-bit wrap
    const long long d = 0x1000000LL - fn;
    if (k < d) { fn += k; k = 0; }
    else { fn = 0; k -= d; }
  }
  if (k > 0) { carry = (fn + k) / fps; fn = (fn + k) % fps; }
  hw[1] = (hw[1] & 0xFF000000u) | ((uint32_t)fn & 0xFFFFFFu);
  const uint32_t sec = (uint32_t)(((hw[0] & 0x3FFFFFFFu) + (unsigned long long)carry) & 0x3FFFFFFFu);
  hw[0] = (hw[0] & 0xC0000000u) | sec;             // setVDIFFrameSecond
}

constexpr uint32_t low_fields_mask(int b, int bits) {      // fields c with (c & b) == 0
  uint32_t m = 0;
  for (int c = 0; c < 32 / bits; ++c)
    if ((c & b) == 0) m |= (uint32_t)(((1ull << bits) - 1) << (c * bits));
  return m;
}

// cornerturn_generic (datamuxer.cpp:243-266) of one input word position: output word j holds,
// at field k*N + l, field j*spw + k of thread l's word (spw = 32 / (bits N)).  With spw = 1
// that is the transpose of an N x N matrix of bits-wide fields, done in log2(N) butterfly
// stages that swap the off-diagonal half-blocks of word pairs.
template <int N, int BITS>
__device__ __forceinline__ void turn_word(uint32_t (&w)[N]) {
  constexpr int SPW = (32 / BITS) / N;
  if constexpr (SPW == 1) {
#pragma unroll
    for (int b = N / 2; b >= 1; b /= 2) {
      const uint32_t m = low_fields_mask(b, BITS);
      const int sh = b * BITS;
#pragma unroll
      for (int r = 0; r < N; ++r) {
        if (r & b) continue;
        const uint32_t x = w[r], y = w[r + b];
        w[r] = (x & m) | ((y & m) << sh);
        w[r + b] = (y & ~m) | ((x >> sh) & m);
      }
    }
  } else {
    constexpr uint32_t M = (uint32_t)((1ull << BITS) - 1);
    uint32_t t[N];
#pragma unroll
    for (int l = 0; l < N; ++l) t[l] = w[l];
#pragma unroll
    for (int j = 0; j < N; ++j) {
      uint32_t o = 0;
#pragma unroll
      for (int k = 0; k < SPW; ++k)
#pragma unroll
        for (int l = 0; l < N; ++l) o |= ((t[l] >> ((j * SPW + k) * BITS)) & M) << ((k * N + l) * BITS);
      w[j] = o;
    }
  }
}

// LDS bytes of one lane's 2N turned words, padded by 16 bytes so that the rows of consecutive
// lanes start 4 banks apart
template <int N>
constexpr int row_bytes() { return 8 * N + 16; }

// grid (slots, chunks of MUXNT 8-byte payload positions)
template <int N, int BITS>
__global__ void __launch_bounds__(MUXNT) k_vmux_frames(MuxArgs a, const uint8_t* __restrict__ in,
                                                      uint8_t* __restrict__ out, MuxWork w) {
  const long long s = blockIdx.x;
  const long long first = (long long)w.ctr[0];
  const unsigned m = w.mask[s];
  if (s < first) return;                           // leading slots with no frame are not output
  const int ofb = (a.ifb - MUXHDR) * N + MUXHDR;
  uint8_t* o = out + (s - first) * ofb;
  constexpr unsigned FULL = N == 32 ? 0xFFFFFFFFu : (1u << N) - 1u;
  if (m != FULL) {
    // header only: the first present thread's header, or with no thread present
    // (datamuxer.cpp:716-726) that of the last slot before s with any thread present, which
    // exists because slot `first` has one; chunk 0's workgroup writes it
    if (blockIdx.y != 0) return;
    __shared__ uint32_t hw[8];
    __shared__ int s_p;
    long long p = s;
    if (m == 0) {
      if (threadIdx.x == 0) s_p = -1;
      __syncthreads();
      for (long long base = s - 1; base >= first; base -= MUXNT) {
        const long long c = base - threadIdx.x;
        if (c >= first && w.mask[c] != 0) atomicMax(&s_p, (int)c);
        __syncthreads();
        const bool found = s_p >= 0;
        __syncthreads();                           // every lane has read s_p
        if (found) break;
      }
      if (s_p < 0) return;                         // (unreachable: slot `first` has a thread)
      p = s_p;
    }
    if (threadIdx.x == 0) {
      bool all = false;
      present_header(a, in, w, p, ofb, hw, &all);
      if (p != s) {
        hw[0] |= 0x80000000u;
        advance_frames(hw, s - p, a.fps);
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) reinterpret_cast<uint32_t*>(o)[k] = hw[k];
    }
    return;
  }
  // every thread present: the payload is turned by every chunk's workgroup, the header
  // (thread 0's, made valid) is written by chunk 0's first lane once its loads are issued
  const uint8_t* src[N];
#pragma unroll
  for (int l = 0; l < N; ++l) {
    const int e = __builtin_amdgcn_readfirstlane(w.table[s * N + l]);
    src[l] = in + (long long)(e - 1) * a.ifb + MUXHDR;
  }
  const int npos = (a.ifb - MUXHDR) / 8;          // 8-byte payload positions
  constexpr int RB = row_bytes<N>();
  __shared__ __attribute__((aligned(16))) uint8_t stage[N >= 4 ? MUXNT * RB : 16];
  for (int qb = blockIdx.y * MUXNT; qb < npos; qb += gridDim.y * MUXNT) {
    const int q = qb + threadIdx.x;
    uint32_t w0[N], w1[N];
    if (q < npos) {
#pragma unroll
      for (int l = 0; l < N; ++l) {
        const uint2 v = *reinterpret_cast<const uint2*>(src[l] + 8 * q);
        w0[l] = v.x;
        w1[l] = v.y;
      }
      turn_word<N, BITS>(w0);
      turn_word<N, BITS>(w1);
    }
    // output words 2qN .. 2qN + 2N - 1: w0 then w1
    if constexpr (N == 1) {
      if (q < npos) reinterpret_cast<uint2*>(o + MUXHDR)[q] = make_uint2(w0[0], w1[0]);
    } else if constexpr (N == 2) {
      if (q < npos) reinterpret_cast<uint4*>(o + MUXHDR)[q] = make_uint4(w0[0], w0[1], w1[0], w1[1]);
    } else {
      if (q < npos) {
        uint4* row = reinterpret_cast<uint4*>(stage + threadIdx.x * RB);
#pragma unroll
        for (int j = 0; j < N; j += 4) {
          row[j / 4] = make_uint4(w0[j], w0[j + 1], w0[j + 2], w0[j + 3]);
          row[N / 4 + j / 4] = make_uint4(w1[j], w1[j + 1], w1[j + 2], w1[j + 3]);
        }
      }
      __syncthreads();
      // 16-byte piece c of this chunk's output: lane c / (N/2)'s row, piece c % (N/2)
      uint4* dst = reinterpret_cast<uint4*>(o + MUXHDR) + (long long)qb * (N / 2);
#pragma unroll
      for (int it = 0; it < N / 2; ++it) {
        const int c = it * MUXNT + threadIdx.x;
        const int r = c / (N / 2), k = c % (N / 2);
        if (qb + r < npos) dst[c] = *reinterpret_cast<const uint4*>(stage + r * RB + 16 * k);
      }
      __syncthreads();                             // the rows are rewritten next pass
    }
  }
  if (blockIdx.y == 0 && threadIdx.x == 0) {
    uint32_t hw[8];
    bool all = true;
    present_header(a, in, w, s, ofb, hw, &all);
#pragma unroll
    for (int k = 0; k < 8; ++k) reinterpret_cast<uint32_t*>(o)[k] = hw[k];
  }
}

template <int N, int BITS>
int launch_frames(const MuxArgs& a, const uint8_t* in, uint8_t* out, const MuxWork& w, hipStream_t st) {
  const int npos = (a.ifb - MUXHDR) / 8;
  const int nchunk = std::max(1, std::min(64, (npos + MUXNT - 1) / MUXNT));
  hipLaunchKernelGGL((k_vmux_frames<N, BITS>), dim3((unsigned)a.nslots, (unsigned)nchunk), dim3(MUXNT), 0, st,
                     a, in, out, w);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// (N, bits) with N * bits <= 32; -2 for any other pair
template <int N>
int launch_bits(const MuxArgs& a, const uint8_t* in, uint8_t* out, const MuxWork& w, hipStream_t st) {
  switch (a.bits) {
    case 1: return launch_frames<N, 1>(a, in, out, w, st);
    case 2: if constexpr (N <= 16) return launch_frames<N, 2>(a, in, out, w, st); else return -2;
    case 4: if constexpr (N <= 8) return launch_frames<N, 4>(a, in, out, w, st); else return -2;
    case 8: if constexpr (N <= 4) return launch_frames<N, 8>(a, in, out, w, st); else return -2;
    case 16: if constexpr (N <= 2) return launch_frames<N, 16>(a, in, out, w, st); else return -2;
    case 32: if constexpr (N <= 1) return launch_frames<N, 32>(a, in, out, w, st); else return -2;
    default: return -2;
  }
}

int dispatch_frames(const MuxArgs& a, const uint8_t* in, uint8_t* out, const MuxWork& w, hipStream_t st) {
  switch (a.n) {
    case 1: return launch_bits<1>(a, in, out, w, st);
    case 2: return launch_bits<2>(a, in, out, w, st);
    case 4: return launch_bits<4>(a, in, out, w, st);
    case 8: return launch_bits<8>(a, in, out, w, st);
    case 16: return launch_bits<16>(a, in, out, w, st);
    case 32: return launch_bits<32>(a, in, out, w, st);
    default: return -2;
  }
}

// The outcome of the last corner turn with workspace w, for a caller that reads it later from
// device memory: {output frames, dropped before, after, unmapped, length error}.
__global__ void k_vmux_outcome(MuxWork w, int64_t* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const unsigned long long c0 = w.ctr[0], c5 = w.ctr[5];
  out[0] = c0 < c5 ? (int64_t)(c5 - c0) : 0;
  out[1] = (int64_t)w.ctr[1];
  out[2] = (int64_t)w.ctr[2];
  out[3] = (int64_t)w.ctr[3];
  out[4] = w.ctr[4] ? 1 : 0;
}

// Workspace of the allocation-free corner turn: one array per GPU, held until the process
// ends.  Host code never uses this name as an address: it takes each device's address of the
// array from hipGetSymbolAddress, once per device.
static __device__ __align__(256) unsigned char scratch[FXG_MUX_STATIC_WORK_BYTES];

}  // namespace

#include <mutex>

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The current device's address of `scratch`, looked up on the first call on that device.
int scratch_of_current_device(void** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return FXG_E_HIP;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return FXG_E_HIP;
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return FXG_OK;
}
}  // namespace

extern "C" int64_t fxg_vdif_mux_workspace_bytes(int64_t nslots, int32_t nthreads) {
  if (nslots < 0 || nthreads < 1) return -1;
  return 64 + 4 * nslots * (int64_t)nthreads + 4 * nslots;
}

extern "C" int fxg_vdif_mux_device(const fxg_vdif_mux_params* p, const uint8_t* d_in, int64_t in_bytes,
                                   uint8_t* d_out, int64_t out_bytes, int64_t nslots, void* d_work,
                                   int64_t work_bytes, void* stream) {
  if (!p || !d_out || !d_work || nslots < 1 || in_bytes < 0 || (in_bytes > 0 && !d_in)) return FXG_E_INVALID;
  const int n = p->nthreads, bits = p->bitspersample;
  if (n < 1 || n > FXG_MUX_MAX_THREADS || (n & (n - 1))) return FXG_E_INVALID;
  if (bits < 1 || bits > 32 || (bits & (bits - 1)) || n * bits > 32) return FXG_E_INVALID;
  if (p->inputframebytes <= MUXHDR || p->inputframebytes % 8 || p->framespersecond < 1) return FXG_E_INVALID;
  if (p->inputframebytes > (1 << 27) / n) return FXG_E_INVALID;    // output length field: 24 bits of 8 bytes
  const int64_t ofb = (int64_t)(p->inputframebytes - MUXHDR) * n + MUXHDR;
  if (nslots > 0x7FFFFFFF || in_bytes / p->inputframebytes >= 0x7FFFFFFF) return FXG_E_INVALID;
  if (out_bytes < nslots * ofb || work_bytes < fxg_vdif_mux_workspace_bytes(nslots, n)) return FXG_E_INVALID;
  if ((reinterpret_cast<uintptr_t>(d_in) & 7) || (reinterpret_cast<uintptr_t>(d_out) & 15) ||
      (reinterpret_cast<uintptr_t>(d_work) & 7))
    return FXG_E_INVALID;
  MuxArgs a;
  a.ifb = p->inputframebytes; a.n = n; a.bits = bits; a.fps = p->framespersecond;
  for (int j = 0; j < FXG_MUX_MAX_THREADS; ++j) a.map[j] = j < n ? p->threadmap[j] : -1;
  a.ref_seconds = p->ref_seconds; a.ref_frame = p->ref_frame; a.start = p->start;
  a.nslots = nslots; a.nin = in_bytes / p->inputframebytes;
  const MuxWork w = work_of(d_work, nslots, n);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const long long ntable = nslots * (n + 1);       // slot table and presence masks
  const unsigned ib = (unsigned)std::min<long long>(4096, (ntable + 255) / 256 + 1);
  hipLaunchKernelGGL(k_vmux_init, dim3(ib), dim3(256), 0, st, w, ntable, nslots);
  if (a.nin > 0)
    hipLaunchKernelGGL(k_vmux_index, dim3((unsigned)((a.nin + 255) / 256)), dim3(256), 0, st, a, d_in, w);
  if (hipGetLastError() != hipSuccess) return FXG_E_HIP;
  const int r = dispatch_frames(a, d_in, d_out, w, st);
  return r == 0 ? FXG_OK : (r == -2 ? FXG_E_INVALID : FXG_E_HIP);
}

extern "C" int fxg_vdif_mux_result(const void* d_work, void* stream, int64_t* nout, int64_t* dropped) {
  if (!d_work) return FXG_E_INVALID;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  unsigned long long c[8];
  if (hipMemcpyAsync(c, d_work, sizeof(c), hipMemcpyDeviceToHost, st) != hipSuccess) return FXG_E_HIP;
  if (hipStreamSynchronize(st) != hipSuccess) return FXG_E_HIP;
  if (nout) *nout = c[0] < c[5] ? (int64_t)(c[5] - c[0]) : 0;   // slots from the first present one
  if (dropped) { dropped[0] = (int64_t)c[1]; dropped[1] = (int64_t)c[2]; dropped[2] = (int64_t)c[3]; }
  return c[4] ? FXG_E_INVALID : FXG_OK;
}

extern "C" int fxg_vdif_mux_device_static(const fxg_vdif_mux_params* p, const uint8_t* d_in, int64_t in_bytes,
                                          uint8_t* d_out, int64_t out_bytes, int64_t nslots, int64_t* d_outcome,
                                          void* stream) {
  if (!p || !d_outcome || (reinterpret_cast<uintptr_t>(d_outcome) & 7)) return FXG_E_INVALID;
  const int64_t need = fxg_vdif_mux_workspace_bytes(nslots, p->nthreads);
  if (need < 0 || need > FXG_MUX_STATIC_WORK_BYTES) return FXG_E_INVALID;   // beyond the static workspace
  void* work = nullptr;
  if (int rc = scratch_of_current_device(&work)) return rc;
  const int rc = fxg_vdif_mux_device(p, d_in, in_bytes, d_out, out_bytes, nslots, work, FXG_MUX_STATIC_WORK_BYTES,
                                     stream);
  if (rc != FXG_OK) return rc;
  hipLaunchKernelGGL(k_vmux_outcome, dim3(1), dim3(64), 0, reinterpret_cast<hipStream_t>(stream),
                     work_of(work, nslots, p->nthreads), d_outcome);
  return hipGetLastError() == hipSuccess ? FXG_OK : FXG_E_HIP;
}
