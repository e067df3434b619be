// Where a subint's data starts in a datastream's segmented read buffer, and which of its FFT
// blocks have data: DataStream::calculateControlParams (mpifxcorr/src/datastream.cpp:368-627)
// as a host function over a description of the buffer ring.  Integer and double arithmetic
// is kept in the reference's types and order (truncating casts, C remainders of negative
// values, the int byte offsets), because the result is a byte offset that both sides must
// agree on exactly.
#include <cmath>
#include <cstring>
#include <stdint.h>
#include "../../include/fxg.h"

namespace {

constexpr long long NS_PER_S = 1000000000LL;

struct Choice {
  const fxg_databuffer* buf;
  int32_t* control;
  int nflag;
  // an invalid subint: control[1] = INVALID_SUBINT, no valid flags, buffer index 0
  int invalid() {
    control[1] = FXG_INVALID_SUBINT;
    for (int k = 0; k < nflag; ++k) control[3 + k] = 0;
    return 0;
  }
};

// the byte offset of (offsetsec, firstoffsetns) in segment `at`, in double (datastream.cpp:490, 601)
double dindex(const fxg_databuffer* b, int at, int offsetsec, long long firstoffsetns) {
  const fxg_segment& g = b->seg[at];
  return at * b->readbytes + ((((offsetsec - g.scanseconds) * 1.0e9 + (firstoffsetns - (long long)g.scanns)) /
                               g.sampletimens) * g.bytespersamplenum) / g.bytespersampledenom;
}

}  // namespace

extern "C" int fxg_control_params(const fxg_databuffer* buf, int32_t* atsegment, int32_t scan, int32_t offsetsec,
                                  int32_t offsetns, int32_t delay_ok, double delayus1, double delayus2,
                                  fxg_advance_fn advance, void* user, int32_t* control, int32_t* bufferindex_out) {
  if (!buf || !atsegment || !control || !bufferindex_out || !buf->seg || buf->numdatasegments < 1 ||
      *atsegment < 0 || *atsegment >= buf->numdatasegments || buf->readbytes < 1)
    return FXG_E_INVALID;
  const int nseg = buf->numdatasegments;
  int at = *atsegment;
  const fxg_segment* seg = buf->seg;
  const int readbytes = buf->readbytes;
  if (buf->framing != FXG_FRAMING_NONE &&
      (buf->payloadbytes < 1 || buf->framebytes < buf->payloadbytes || buf->framespersecond < 1 ||
       buf->framegranularity < 1))
    return FXG_E_INVALID;
  for (int s = 0; s < nseg; ++s)
    if (seg[s].blockspersend < 1 || seg[s].bytespersamplenum < 1 || seg[s].bytespersampledenom < 1 ||
        seg[s].bytesbetweenintegerns < 1 || !(seg[s].sampletimens > 0))
      return FXG_E_INVALID;
  Choice c{buf, control, (seg[at].blockspersend + FXG_FLAGS_PER_INT - 1) / FXG_FLAGS_PER_INT};
  *bufferindex_out = 0;
  control[0] = scan;
  control[2] = 0;
  for (int k = 0; k < c.nflag; ++k) control[3 + k] = 0;

  // first and last sample times needed, in ns after offsetsec (datastream.cpp:380-409)
  if (!delay_ok) return c.invalid();
  const double d1 = delayus1 - buf->intclockseconds * 1000000;
  long long firstoffsetns = (long long)offsetns - static_cast<long long>(d1 * 1000);
  int64_t dataspanns = static_cast<int64_t>(seg[at].numchannels * seg[at].blockspersend * 2 * seg[at].sampletimens + 0.5);
  if (seg[at].complex_sampling) dataspanns /= 2;
  const double d2 = delayus2 - buf->intclockseconds * 1000000;
  long long lastoffsetns = (long long)offsetns + (long long)dataspanns + static_cast<long long>(-1000 * d2 + 0.5);
  for (int k = 0; k < 2; ++k) {
    if (lastoffsetns < 0) {
      offsetsec -= 1;
      firstoffsetns += NS_PER_S;
      lastoffsetns += NS_PER_S;
    }
  }

  // coarse checks, then move to the segment holding the start (datastream.cpp:413-457)
  if (scan < seg[at].scan) return c.invalid();
  if (!seg[at].readto) return c.invalid();
  if (scan == seg[at].scan && offsetsec < seg[at].scanseconds - 3) return c.invalid();
  long long nsdifference = (long long)(offsetsec - seg[at].scanseconds) * NS_PER_S + firstoffsetns -
                           static_cast<long long>(seg[at].scanns);
  long long validns = (static_cast<long long>(seg[at].validbytes) * static_cast<long long>(seg[at].nsinc)) / readbytes;
  int moves = 0;
  while ((scan > seg[at].scan || (scan == seg[at].scan && nsdifference >= validns)) &&
         (buf->keepreading || at != buf->lastvalidsegment)) {
    // the reference blocks here until its reader refills the ring; without a callback that
    // does the same, one pass round the ring is the end
    if (advance ? advance(user, at) != 0 : ++moves > nseg) return FXG_E_STATE;
    at = (at + 1) % nseg;
    *atsegment = at;
    nsdifference = (long long)(offsetsec - seg[at].scanseconds) * NS_PER_S + firstoffsetns -
                   static_cast<long long>(seg[at].scanns);
    validns = (static_cast<long long>(seg[at].validbytes) * static_cast<long long>(seg[at].nsinc)) / readbytes;
  }
  const fxg_segment& g = seg[at];
  const fxg_segment& nx = seg[(at + 1) % nseg];
  c.nflag = (g.blockspersend + FXG_FLAGS_PER_INT - 1) / FXG_FLAGS_PER_INT;
  for (int k = 0; k < c.nflag; ++k) control[3 + k] = 0;
  control[0] = scan;
  if (scan < g.scan || nsdifference < -g.nsinc) return c.invalid();
  if (scan > g.scan) return c.invalid();

  // far outside this segment when it is short or the next one is not contiguous
  // (datastream.cpp:481-508)
  const int fftchannels = g.complex_sampling ? g.numchannels : g.numchannels * 2;
  const int blockbytes = (fftchannels * g.bytespersamplenum) / g.bytespersampledenom;
  const double currentsegmenttime = g.scanseconds + g.scanns / 1.0e9;
  const double nextsegmenttime = nx.scanseconds + nx.scanns / 1.0e9;
  if (g.validbytes < readbytes || nx.scanseconds - g.scanseconds > 2 || nextsegmenttime < currentsegmenttime) {
    const double dbi = dindex(buf, at, offsetsec, firstoffsetns);
    const int segoffbytes = int(dbi - at * readbytes);
    if (nextsegmenttime < currentsegmenttime && segoffbytes > g.validbytes) return c.invalid();
    if (dbi < -blockbytes * g.blockspersend || dbi > buf->bufferbytes) return c.invalid();
  }

  // byte index of the start, on a 16-bit boundary (datastream.cpp:515-524)
  control[1] = g.scanseconds;
  int bufferindex = at * readbytes +
      (int(((long long)(offsetsec - g.scanseconds) * NS_PER_S + (firstoffsetns - static_cast<long long>(g.scanns))) /
           g.sampletimens + 0.5) * g.bytespersamplenum) / g.bytespersampledenom;
  if (bufferindex % 2 != 0) bufferindex--;
  if (bufferindex == buf->bufferbytes) {           // at the very end, and the next is not contiguous
    if (g.scan != nx.scan ||
        ((long long)(nx.scanseconds - g.scanseconds) * NS_PER_S + static_cast<long long>(nx.scanns) -
         static_cast<long long>(g.scanns) - static_cast<long long>(g.nsinc)) != 0)
      return c.invalid();
  }

  // skip the blocks that start before the segment (datastream.cpp:541-568)
  int count = 0;
  while (bufferindex < at * readbytes && count < g.blockspersend) {
    bufferindex += blockbytes;
    count++;
  }
  if (bufferindex < at * readbytes) return c.invalid();
  if (count > 0) {
    int tosubtract = (int(count * 1000.0 * (d2 - d1) / (g.sampletimens * g.blockspersend) + 0.5) *
                      g.bytespersamplenum) / g.bytespersampledenom;
    if (bufferindex - tosubtract < at * readbytes) {
      count++;
      bufferindex += blockbytes;
      if (count == g.blockspersend) return c.invalid();
      tosubtract = (int(count * 1000.0 * (d2 - d1) / (g.sampletimens * g.blockspersend) + 0.5) *
                    g.bytespersamplenum) / g.bytespersampledenom;
      bufferindex -= tosubtract;
    }
    bufferindex -= bufferindex % 2;
  }

  // back to a byte at a whole nanosecond; the start time (datastream.cpp:570-582)
  int segoffbytes = bufferindex - at * readbytes;
  bufferindex -= segoffbytes % g.bytesbetweenintegerns;
  segoffbytes -= segoffbytes % g.bytesbetweenintegerns;
  const long long segoffns = static_cast<long long>(
      double(segoffbytes * g.bytespersampledenom / g.bytespersamplenum) * g.sampletimens + 0.5);
  long long a = static_cast<long long>(control[1]) + segoffns / NS_PER_S;
  long long b = static_cast<long long>(g.scanns) + segoffns % NS_PER_S;
  a += b / NS_PER_S;
  b %= NS_PER_S;
  control[1] = (int32_t)a;
  control[2] = (int32_t)b;

  // valid flags (datastream.cpp:584-620)
  const bool nextcontig =
      g.validbytes == readbytes &&
      ((long long)(nx.scanseconds - g.scanseconds) * NS_PER_S + static_cast<long long>(nx.scanns) -
       static_cast<long long>(g.scanns)) == g.nsinc;
  if ((g.validbytes - segoffbytes >= g.sendbytes) ||
      (nextcontig && (g.validbytes - segoffbytes + nx.validbytes) > g.sendbytes)) {
    for (int i = count; i < g.blockspersend; ++i) control[3 + i / FXG_FLAGS_PER_INT] |= 1 << (i % FXG_FLAGS_PER_INT);
  } else {
    const double dbi = dindex(buf, at, offsetsec, firstoffsetns);
    if (dbi < -blockbytes * g.blockspersend || dbi > buf->bufferbytes) return c.invalid();
    for (int i = count; i < g.blockspersend; ++i) {
      if (bufferindex + i * blockbytes - at * readbytes < g.validbytes)
        control[3 + i / FXG_FLAGS_PER_INT] |= 1 << (i % FXG_FLAGS_PER_INT);
      else if (nextcontig && (g.validbytes - segoffbytes + nx.validbytes) > i * blockbytes)
        control[3 + i / FXG_FLAGS_PER_INT] |= 1 << (i % FXG_FLAGS_PER_INT);
    }
    // (the reference looks at the whole flag words only: a partial last word does not count)
    bool somegooddata = false;
    for (int i = 0; i < g.blockspersend / FXG_FLAGS_PER_INT; ++i)
      if (control[3 + i] > 0) somegooddata = true;
    if (!somegooddata) return c.invalid();
  }

  // framed streams: back to the start of a frame (vdiffile.cpp:418-467, mark5bfile.cpp:174-240;
  // the reader's configindex checks before it are the caller's, and an index past the ring,
  // where the reference aborts, is an invalid subint here)
  if (buf->framing != FXG_FRAMING_NONE) {
    const int vlbaoffset = bufferindex - at * readbytes;
    if (vlbaoffset < 0) return c.invalid();
    int framesin = vlbaoffset / buf->payloadbytes;
    if (framesin % buf->framegranularity != 0) {
      framesin -= framesin % buf->framegranularity;
      if (buf->framing == FXG_FRAMING_MARK5B && framesin < 0) framesin += buf->framegranularity;
    }
    long long segoffns;
    if (buf->framing == FXG_FRAMING_VDIF)
      segoffns = g.scanns + (long long)((1000000000.0 * framesin) / buf->framespersecond);
    else
      segoffns = g.scanns + static_cast<long long>((NS_PER_S * framesin) / buf->framespersecond);
    control[1] = g.scanseconds + (int)(segoffns / NS_PER_S);
    control[2] = (int)(segoffns % NS_PER_S);
    bufferindex = at * readbytes + framesin * buf->framebytes;
    if (bufferindex == buf->bufferbytes) {
      if (g.scan != nx.scan ||
          ((long long)(nx.scanseconds - g.scanseconds) * NS_PER_S + nx.scanns - g.scanns - g.nsinc) != 0)
        return c.invalid();
    }
    if (bufferindex > buf->bufferbytes) return c.invalid();
  }
  *bufferindex_out = bufferindex;
  return FXG_OK;
}
