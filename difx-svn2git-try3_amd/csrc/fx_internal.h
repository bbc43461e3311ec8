// fx_internal.h — structures shared by the host side (fx_api.cpp, fx_plan.cpp)
// and the device kernels (fx_station.hip, fx_xmac.hip) of libfxcorr_gpu.
#pragma once
#include <cstdint>
#include <string>
#include <vector>
#include "../../include/fxcorr_gpu.h"

namespace fxg {

constexpr double kTwoPi = 6.283185307179586476925286766559;  // architecture.h.in:435
constexpr double kTiny = 0.000000001;                         // mode.cpp:35
constexpr float kHiMag = 3.3359f;                              // mark5_stream.h:75
constexpr int kBig = 1 << 30;                                  // "valid to the end" byte index

// ---- per datastream constants (Mode / Mk5Mode construction) --------------
struct DevStation {
  int format, nbits, nchan, framebytes, headerbytes, databytes, framesamples;
  int granularity, groupbytes, groupsamples, quirk8, blanker;
  int C, L, s, nfrsteps, nfracsteps;
  int bpbnum, bpbden;
  int fractionalLo, deltapol;
  int band0;            // global index of recorded band 0 of this datastream
  int slog;             // log2(s) when s is a power of two, else -1
  int bpf;              // bytes per band per frame in the corner-turned streams
  int order;            // fringe rotation order 0 (post-F), 1 (linear), 2 (quadratic)
  int cplx;             // complex samples (real part in the low nbits of each unit)
  int unitbits;         // bits of one band's sample: nbits, or 2 nbits if complex
  int edv4;             // INTERLACEDVDIF: EDV4 per-band validity and weights
  int pad0;
  double sampletime, bw;
  float lut[4];         // 2-bit level table (VDIF or Mark5B order)
};

// ---- per recorded band (global band index) -------------------------------
struct DevBand {
  int station, band, localfreq, lsb;
  int useB;             // apply the "B" frac-sample rotator (delta pol offsets)
  int dsb;              // complex double sideband: halves swapped, frac steps offset by -BW/2
  double lofreq, clock, clockdelta, looffset;   // lofreq: the fringe rotator's LO
  long long specoff;    // element offset of this band in the spectra buffer
  float polre, polim;   // polarisation phase correction of the band's spectrum (1, 0: none)
};

// ---- per subintegration, per datastream ----------------------------------
struct DevStationSub {
  const unsigned char* data;
  long long nbytes;
  long long nframes;    // whole frames inside nbytes
  long long fbase;      // offset of this station in the frame table
  int datasec, datans, usable, nflags;
  int flagoff;          // offset of this station's valid flags
  int datasamples;      // samples from the second boundary to the data start (mode.cpp:1572)
  long long ctbase;     // byte offset of band 0's stream in the corner-turn buffer
  long long ctstride;   // bytes between consecutive bands' streams (multiple of 4)
  double interp[3];
  // the subintegration's time (procslots[].offsets, Model::getScanStartSec),
  // the same in every datastream's entry
  int offsec, offns, scanstart, pad;
  // INTERLACEDVDIF only (else null): per frame, the bands whose samples the
  // EDV4 validity mask blanks (bit b = band b); per (block, band), the band's
  // data weight (Mk5Mode perbandweights)
  unsigned long long* fblank;
  float* pbw;
};

// ---- per block, per datastream (Mode::process prologue result) -----------
struct DevBlock {
  int status, nearestsample, integerdelay, intwall;
  int allvalid;         // every sample of the block decodes from a valid group
  int pad;
  float weight, fse;
  long long validend;   // samples at or after this index decode as zero
  double a, b, fracwall;
  double c;             // order 2: constant term minus integerdelay; order 0: delay at the block centre
  double wall;          // walltimesecs of the block start
};

// ---- weight table entries ---------------------------------------------------
// Sum over the subintegration's blocks of w(d1, b1) * w(d2, b2) (a baseline
// product, core.cpp:1028-1037) or of w(d1, b1) (d2 < 0: an autocorrelation
// weight, mode.cpp:1300-1307), where w(d, b) is recorded band b's weight of
// an INTERLACEDVDIF datastream and the block's data weight otherwise (b < 0).
struct DevWTerm {
  int d1, b1, d2, b2;
};

// ---- result scatter segments -----------------------------------------------
enum SegKind { SEG_VIS = 0, SEG_FLOATS = 1 };
struct DevSeg {
  long long resoff;     // cf32 offset (SEG_VIS) or float offset (SEG_FLOATS)
  long long voff;       // element offset into V (SEG_VIS) or into the weight table
  int kind, nchan;
  int conj;             // take conj(V[b][a]) instead of V[a][b]
  int avg;              // channels averaged into one output channel (1: copy)
};

// ---- phase-cal extraction of one recorded band ----------------------------
// Samples of the block's L-sample window are summed into nbins bins by their
// absolute sample number (since the second boundary) modulo nbins, either as
// they are or (cplx) counter-rotated by the first tone's offset; the tones
// are DFT bins of the sums (pcal.cpp:388-410, 590-650, 1037-1062).
struct DevPcal {
  int gband;            // global recorded band
  int nbins;            // accumulation bins N
  int cplx;             // complex bins (the shifting extractors, complex samples)
  int rot;              // samples times the mixer e^{i dphi ((so + n) mod period)}
  int csamp;            // complex samples (unit = 2 x nbits, real part low)
  int period;           // rotator period in samples (rot)
  int ntones;
  int toneoff;          // first entry of the band's tone-bin table (-1 entries give 0)
  long long rotoff;     // float2 offset of the rotator table in the aux buffer (rot)
  long long twoff;      // float2 offset of e^{-2 pi i m / N}, m < N, in the aux buffer
  long long partoff;    // float offset of the band's partial sums [group][N (x2 if cplx)]
  long long resoff;     // cf32 offset of the band's tones in the result
};

// ---- the XMAC description of one frequency --------------------------------
struct XmacFreq {
  int freq;                        // freq table index
  int nchan;
  std::vector<long long> in_off;   // spectra element offset (band offset + channel offset)
  std::vector<int> in_station, in_band;  // datastream, recorded (parent) band
  std::vector<int> in_total;       // datastream total band index (rec or zoom)
  long long voff;                  // offset of V[nin][nin][nchan] in the V buffer
};

// ---- job plan (host side) --------------------------------------------------
struct Plan {
  fxg_plan_info info{};
  int nfreqs = 0, nds = 0, nbl = 0;
  std::vector<char> freqused, equivused;
  std::vector<std::vector<long long>> bloff, bwoff;    // [freq][baseline]
  std::vector<std::vector<int>> bllocal;               // [baseline][freq]
  std::vector<long long> acoff, acwoff;
  std::vector<int> arraystride;
  int xmacstride = 0, rotatestride = 0, bandsperac = 1;
  int C = 0, L = 0;                                    // of datastream 0
  bool uniform = true;                                 // every datastream has nchan == C
  int nbands_total = 0;                                // recorded bands over all datastreams
  std::vector<DevStation> st;
  std::vector<DevBand> bands;
  std::vector<XmacFreq> xf;                            // XMAC frequencies
  std::vector<int> xf_of_freq;                         // freq -> index in xf or -1
  long long vsize = 0;                                 // complex elements of V
  std::vector<DevSeg> segs;                            // result scatter table
  int maxseglen = 0;
  // weight table entries (float segments copy entry voff)
  std::vector<DevWTerm> wterms;
  bool any_edv4 = false;                               // an INTERLACEDVDIF datastream
  int chunk_blocks = 0;
  // phase cal: result offsets per datastream, tones and first-tone offsets
  // per (datastream, recorded freq) (empty: no extraction), the device tables
  std::vector<long long> pcaloff;
  std::vector<std::vector<int>> pcal_ntones, pcal_offhz;
  std::vector<DevPcal> pcal;
  std::vector<int> pcal_tonebin;
  std::vector<float> pcal_aux;                         // float2 pairs
  int pcal_groups = 0;                                 // block groups per band
  long long pcal_partfloats = 0;
  int pcal_maxtones = 0;
  // block-split XMAC: complex elements of the partial-matrix buffer
  long long vpart_elems = 0;
  // FFTs above 16384 points: blocks per sub-chunk and rotator-table entries per (band, block)
  int zcap = 0, rtabn = 0;
};

// bytes of one spectrum value (complex) entering the cross-multiply
inline int spectrum_bytes(int prec) { return prec == FXG_XMAC_FP32 ? 8 : prec == FXG_XMAC_FP8 ? 2 : 4; }

// the largest accumulation of one band's phase-cal bins (LDS floats)
constexpr int kPcalMaxFloats = 32768;

int build_plan(const fxg_job_desc* d, Plan* p, std::string* err);
// the library's error string (fxg_last_error) and code
int set_error(int code, const char* msg);
// stage 2 of VDIF multiplexing (fx_api.cpp, kernel in fx_mux.hip)
int ctx_mux_launch(fxg_ctx* c, const uint8_t* src, const uint8_t* src_device, long long srcsize, uint8_t* dest,
                   const uint32_t* hdr, const unsigned long long* mask, const long long* off, long long nframes,
                   int nthreads, int unitbits, int units, int outwords, int outframe, unsigned long long good,
                   int partial);
int64_t plan_device_bytes(const Plan& p, const fxg_job_desc* d);

}  // namespace fxg
