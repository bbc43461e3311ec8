// fx_kernels.h — kernel argument blocks and launchers of libfxcorr_gpu
// (implemented in fx_station.hip, fx_xmac.hip, fx_xmac_mfma.hip,
// fx_pcal.hip, fx_mux.hip and fx_kurtosis.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "fx_internal.h"

namespace fxg {

// LDS entries for an F-stage workgroup's rotator tables (sub, step,
// fractional sub and step, quadratic pieces) at FFT length L
constexpr int fxg_rotator_table_entries(int L) { return L >= 16384 ? 768 : 512; }
// the largest FFT length of the single-workgroup F stage; longer ones (to
// 65536) take the two-kernel path
constexpr int kFStageMaxL = 16384;

// Spectra layout.  Channel c of the spectrum of one (band, pass block) sits
// at that spectrum's base + cidx(c, gstride), the bases of consecutive blocks
// of a band bstride elements apart:
//   gstride = 16, bstride = C: [band][block][channel] (row-major);
//   gstride = 16 x pass blocks, bstride = 16: [band][channel / 16][block][16]
//   (channel-group-major: the 16 channels one matrix-core workgroup owns are
//   contiguous across blocks, so its loads are whole lines).
__host__ __device__ inline long long cidx(int c, long long gstride) { return (long long)(c >> 4) * gstride + (c & 15); }

struct FStageArgs {
  const DevStation* st;
  const DevBand* bands;
  const DevStationSub* ss;
  const DevBlock* blocks;
  const int2* ftab;
  const float* lut8;
  const float2* tw;           // e^{-2 pi i m/L}: m < 64, then 64 h for h < L/64; then e^{-2 pi i m/256}, m < 256
  const unsigned char* ct;    // corner-turned per-band sample streams
  const int* bandlist;        // global band per grid row (null: row = band)
  void* spectra;
  int prec;                   // FXG_XMAC_*
  float scale;                // applied before reduced-precision stores
  int nblocks_subint;
  int chunk0;                 // first block of the chunk
  // FFTs above 16384 points run zcap blocks of the chunk at a time:
  float2* zbuf;               // sub-FFT scratch [band][zcap][16][L/16]
  float2* rtab;               // rotator tables [band][zcap][rtabn] (sub, step, quad, frac sub, frac step)
  int zcap;                   // blocks per sub-chunk
  int rtabn;                  // table entries per (band, block)
  int zcb0;                   // first chunk block of this sub-chunk
  const signed char* cexp;    // fp8: power-of-two scale exponent per [global band][channel]
  long long gstride;          // spectra layout (cidx)
  int bstride;
};

struct XmacArgs {
  const void* spec;
  const long long* in_off;    // [nin] element offsets of the inputs' block 0, channel 0
  long long block_stride;     // elements between consecutive blocks of one input
  long long gstride;          // channel c of an input's block at + cidx(c, gstride)
  int nin, nchan, nblocks;
  int CT, TB, ntiles, G;      // channel tile, blocks per stage, tiles, 8-groups
  float2* V;                  // [nin][nin][nchan]
  int accumulate;
  // block split: grid row y takes blocks [y*split_blocks, (y+1)*split_blocks)
  // and writes V + y*vsplit (partial sums, reduced by launch_xmac_reduce)
  int split_blocks;
  long long vsplit;
  // fp8: scale exponents, and per input the offset of its channel 0 in them
  const signed char* cexp;
  const int* eoff;
};

struct PcalArgs {
  const DevStation* st;
  const DevBand* bands;
  const DevStationSub* ss;
  const DevBlock* blocks;
  const int2* ftab;
  const float* lut8;
  const unsigned char* ct;    // corner-turned per-band sample streams
  const DevPcal* pc;          // [bands with phase cal]
  const float2* aux;          // mixer and DFT tables
  const int* tonebin;         // DFT bin per tone, -1 = zero
  float* part;                // partial sums [band][group][bins]
  float* res;                 // the result buffer as floats
  int nblocks;                // blocks per subintegration
  int ngroups;                // block groups per band
};

struct KurtArgs {
  const void* spectra;        // [band][pass block][C] of the job's precision
  int prec;
  float invscale;             // 1 / the fp16 / bf16 spectrum scale
  const signed char* cexp;    // fp8: scale exponent per [band][channel]
  int nbands, C, nout;        // recorded bands, channels, record channels
  long long band_stride;      // spectrum elements between bands
  long long gstride;          // spectra layout (cidx)
  int bstride;
  int chunk0, nchunk;         // blocks of this pass
  int nblocks, nrec;          // blocks per subintegration, records (threads)
  float* s;                   // [record][s1, s2][band][C]
};
// s1/s2 of one pass's blocks; the records' sk into out [record][band][nout]
// and the nonzero flag of each (record, band)
hipError_t launch_kurt_accum(const KurtArgs& a, hipStream_t stream);
hipError_t launch_kurt_final(const KurtArgs& a, float* out, int* nzflag, hipStream_t stream);

struct MuxArgs {
  const unsigned char* src;     // input bytes on the device
  unsigned char* dest;          // output frames
  const long long* srcoff;      // [frame][thread] payload byte offsets into src
  const unsigned long long* mask;   // [frame] threads present
  const unsigned* headers;      // [frame][8] output header words
  int nframes, nthreads, unitbits, unitsper;   // unitsper: thread units per sample-time (power of two)
  int outwords;                 // payload words per output frame
  int framebytes;               // output frame bytes
  unsigned long long goodmask;
  int partial;                  // corner-turn the threads an incomplete frame has
};
hipError_t launch_vdifmux(const MuxArgs& a, hipStream_t stream);

// digests of frames [f0, f0 + nf) of a frame stream (device bytes) into out[nf],
// and the same digest of n host bytes
hipError_t launch_frame_digest(const unsigned char* data, long long f0, long long nf, int framebytes,
                               unsigned long long* out, hipStream_t stream);
unsigned long long host_frame_digest(const unsigned char* p, int n);
hipError_t launch_frames(const DevStation* st, const DevStationSub* ss, int2* ftab, int nst,
                         long long maxframes, hipStream_t stream);
hipError_t launch_cornerturn(const DevStation* st, const DevStationSub* ss, unsigned char* ct, int nst,
                             long long maxframes, int maxdatabytes, int fastW, long long maxsamples,
                             hipStream_t stream);
hipError_t launch_prologue(const DevStation* st, const DevStationSub* ss, const int* flags,
                           const int2* ftab, int nblocks, DevBlock* out, int nst, hipStream_t stream);
// cplx: complex samples (L = nchan, every output of the transform kept)
hipError_t launch_fstage(int L, int nbits, int cplx, const FStageArgs& a, int nbands, int nblocks, hipStream_t stream);
hipError_t launch_unpack(const DevStation* st, const DevStationSub* ss, const int2* ftab, const float* lut8,
                         long long first, int nsamp, int nchan, long long validend, float* out,
                         hipStream_t stream);
hipError_t launch_xmac_f32(const XmacArgs& a, int prec, hipStream_t stream);
bool xmac_mfma_supported(int nin, int nchan);
hipError_t launch_xmac_mfma(const XmacArgs& a, int prec, hipStream_t stream);
// the direct-to-LDS double-buffered variant; needs 16-byte aligned rows
// fp8 e4m3 spectra on the microscaled matrix cores (nin <= 64, nchan % 16 == 0)
hipError_t launch_xmac_fp8(const XmacArgs& a, hipStream_t stream);
bool xmac_mfma_streaming_ok(const XmacArgs& a, const long long* in_off_host);
// the streamed kernel also takes more than 64 inputs (chunks of 64 in pairs)
constexpr int kXmacMfmaMaxInputs = 512;
bool xmac_mfma_streaming_supported(int nin, int nchan);
hipError_t launch_xmac_mfma_streaming(const XmacArgs& a, int prec, const void* zero16, hipStream_t stream);
// FP32 spectra (row-major) on the FP32 matrix cores: nchan % 16 == 0, even
// input offsets and block stride
bool xmac_mfma_f32_ok(const XmacArgs& a, const long long* in_off_host);
hipError_t launch_xmac_mfma_f32(const XmacArgs& a, const void* zero16, hipStream_t stream);
// V[i] (+)= sum over splits of P[s][i], i < n, in split order
hipError_t launch_xmac_reduce(const float2* P, int nsplit, long long vsplit, float2* V, long long n, int accumulate,
                              hipStream_t stream);
hipError_t launch_weights(const DevStation* st, const DevStationSub* ss, const DevBlock* blocks, int nblocks,
                          const DevWTerm* terms, int nterms, float* w, hipStream_t stream);
hipError_t launch_pcal(const PcalArgs& a, int npcal, hipStream_t stream);
// p[0, n) = 0 with a kernel (graph-safe clearing)
hipError_t launch_zero(float* p, long long n, hipStream_t stream);
hipError_t launch_scatter(const DevSeg* segs, int nseg, int maxlen, const float2* V, const float* w,
                          float vscale, float* res, hipStream_t stream);

}  // namespace fxg
