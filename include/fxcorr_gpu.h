/*
 * fxcorr_gpu.h — C ABI of the MI355X FX station-processing + cross-multiply
 * library (libfxcorr_gpu.so).
 *
 * The library replaces, for one GPU, the hot path of DiFX mpifxcorr:
 *   - Mk5Mode::unpack + mark5access decoders  (mpifxcorr/src/mk5mode.cpp:104-177,
 *     libraries/mark5access/mark5access/format_vdif.c:768-2020)
 *   - Mode::process per FFT block             (mpifxcorr/src/mode.cpp:649-1397)
 *   - Core::processdata XMAC + weights        (mpifxcorr/src/core.cpp:667-1122)
 *   - uvshiftAndAverage / averageAndSendAutocorrs copy into the slot result
 *     buffer laid out by Configuration::populateResultLengths
 *                                              (core.cpp:1424-1867, 1159-1369;
 *                                               configuration.cpp:2212-2441)
 *
 * Conventions
 *   - Every function returns 0 on success and a negative FXG_E* code on error;
 *     fxg_last_error() returns a thread-local message for the last failure.
 *   - Handles are opaque.  Device memory is owned by the library.  Host pointers
 *     passed in are borrowed for the duration of the call only.
 *   - One host thread per job handle at a time.  Jobs on one context may be
 *     driven from different host threads.
 *   - All descriptor structs are plain fixed-size C structs (no nested
 *     pointers other than the documented top-level arrays) so that any FFI
 *     (ctypes, cgo, JNI) can build them.
 */
#ifndef FXCORR_GPU_H
#define FXCORR_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FXG_ABI_VERSION 3

/* ---- limits of the fixed-size descriptors ------------------------------ */
#define FXG_MAX_FREQS        64   /* frequency table entries per job          */
#define FXG_MAX_RECFREQS     32   /* recorded freqs per datastream            */
#define FXG_MAX_BANDS        64   /* recorded bands per datastream            */
#define FXG_MAX_ZOOMBANDS   128   /* zoom bands per datastream                */
#define FXG_MAX_BLFREQS      48   /* frequencies per baseline                 */
#define FXG_MAX_POLPRODUCTS   4   /* pol products per baseline frequency      */

/* ---- error codes ------------------------------------------------------- */
#define FXG_OK             0
#define FXG_EINVAL        -1   /* bad argument / descriptor                  */
#define FXG_ENOMEM        -2   /* device or host allocation failed           */
#define FXG_EHIP          -3   /* HIP runtime error                          */
#define FXG_EUNSUPPORTED  -4   /* valid for the reference, not built here    */
#define FXG_ESTATE        -5   /* call out of order (e.g. wait w/o submit)   */
#define FXG_ENODEV        -6   /* no usable gfx950 device                    */

/* ---- enums -------------------------------------------------------------- */
/* Data formats handled by the Mk5Mode path (configuration.cpp:991-1011). */
#define FXG_FORMAT_VDIF     0   /* VDIF, 32-byte (or 16-byte legacy) header */
#define FXG_FORMAT_MARK5B   1   /* Mark5B, 16-byte header, 10000-byte payload */
/* VDIF multiplexed from threads with per-thread validity (Configuration::
 * INTERLACEDVDIF, vdifmux with validity propagation): decoded as VDIF, then
 * the bands whose thread a frame's EDV4 validity mask marks absent are
 * blanked and every recorded band carries its own data weight
 * (mk5mode.cpp:78-90, 145-168; blank_vdif_EDV4, format_vdif.c:4699-4775);
 * real sampling, 32-byte headers. */
#define FXG_FORMAT_INTERLACEDVDIF 2

/* Sampling of a datastream (Configuration::datasampling / complextype). */
#define FXG_SAMPLING_REAL     0
#define FXG_SAMPLING_COMPLEX  1

/* Blanker selection (mark5access): the VDIF default blanks frames whose
 * payload starts or ends with the fill pattern (blanker_mark5.c:282-312);
 * network datastreams use blanker_none (mk5mode.cpp:64-65). */
#define FXG_BLANKER_FORMAT  0
#define FXG_BLANKER_NONE    1

/* Cross-multiply arithmetic. */
#define FXG_XMAC_FP32       0   /* plain drop-in: cf32 spectra, FP32 FMA     */
#define FXG_XMAC_FP16       1   /* fp16 spectra, MFMA, FP32 accumulation     */
#define FXG_XMAC_BF16       2   /* bf16 spectra, MFMA, FP32 accumulation     */
#define FXG_XMAC_FP8        3   /* fp8 e4m3 spectra with a power-of-two scale
                                   per (band, channel), microscaled MFMA, FP32
                                   accumulation                               */

/* Sentinel used by Mode for a torched subintegration (mode.h:232). */
#define FXG_INVALID_SUBINT  (-99999999)

/* Number of valid-flag bits per int32 (mpifxcorr.h:39). */
#define FXG_FLAGS_PER_INT   30

/* ---- frequency table entry (configuration.h freqdata) ------------------ */
typedef struct {
  double  freq_mhz;        /* band-edge sky frequency (= LO for fringe rotation) */
  double  bw_mhz;          /* bandwidth                                         */
  int32_t lsb;             /* 1 = lower sideband                                */
  int32_t nchan;           /* output channels (recordedbandchannels)            */
  int32_t chans_to_avg;    /* output channels averaged into one (divides nchan) */
  int32_t pad;
} fxg_freq;

/* ---- datastream (station) table entry (configuration.h datastreamdata) -- */
typedef struct {
  int32_t format;          /* FXG_FORMAT_*                                      */
  int32_t nbits;           /* 2 or 8 (VDIF, per component if complex); 2 (Mark5B) */
  int32_t framebytes;      /* header + payload bytes of one frame               */
  int32_t headerbytes;     /* 32 or 16 (VDIF legacy); 16 (Mark5B)               */
  int32_t blanker;         /* FXG_BLANKER_*                                     */
  int32_t nrecfreqs;       /* recorded frequencies                              */
  int32_t nrecbands;       /* recorded bands = channels in the stream           */
  int32_t nzoombands;      /* zoom bands (aliases of recorded band spectra)     */
  /* per recorded frequency (index = local freq index) */
  int32_t recfreq_index[FXG_MAX_RECFREQS];       /* freq table index          */
  double  recfreq_clockoffset_us[FXG_MAX_RECFREQS];
  double  recfreq_clockoffsetdelta_us[FXG_MAX_RECFREQS];
  /* polarisation phase offset: when any is set, each frequency with two or
   * more bands turns one of its first two (the second if the first is 'R',
   * else the first) by exp(i offset/2) after the fractional-sample
   * rotation (mode.cpp:435-460, 1341-1353) */
  double  recfreq_phaseoffset_deg[FXG_MAX_RECFREQS];
  double  recfreq_looffset_hz[FXG_MAX_RECFREQS];
  /* per recorded band: local freq index and polarisation ('R','L','X','Y') */
  int32_t recband_freq[FXG_MAX_BANDS];
  char    recband_pol[FXG_MAX_BANDS];
  /* per zoom band: freq table index, polarisation, parent recorded band and
   * channel offset into the parent spectrum (mode.cpp:194-206) */
  int32_t zoomband_freqindex[FXG_MAX_ZOOMBANDS];
  char    zoomband_pol[FXG_MAX_ZOOMBANDS];
  int32_t zoomband_parent[FXG_MAX_ZOOMBANDS];
  int32_t zoomband_choffset[FXG_MAX_ZOOMBANDS];
  /* multi-tone phase calibration (datastreamdata phasecalintervalmhz /
   * phasecalbasemhz, configuration.cpp:1615-1623): tones every interval MHz
   * of sky frequency, extracted per recorded band (Mode::process,
   * mode.cpp:774-788; PCal::getNew, pcal.cpp:139-203).  0 = no extraction. */
  double  pcal_interval_mhz;
  double  pcal_base_mhz;
  /* sampling (datastreamdata sampling / tcomplex, configuration.cpp:1540-1560):
   * FXG_SAMPLING_REAL, or FXG_SAMPLING_COMPLEX (VDIF complex samples: the
   * real part in the low nbits of each 2*nbits unit, the imaginary part in
   * the high nbits); complex_dsb = 1 for double-sideband complex data (the
   * LO at the band centre, COMPLEX_DSB). */
  int32_t sampling;
  int32_t complex_dsb;
} fxg_datastream;

/* ---- baseline table entry (configuration.h baselinedata) ---------------- */
/* Band indices are datastream "total" band indices: recorded bands first,
 * then zoom bands at nrecbands + z (configuration.cpp:1100-1124). */
typedef struct {
  int32_t ds1;             /* datastream index of the first antenna             */
  int32_t ds2;             /* datastream index of the second antenna            */
  int32_t nfreqs;
  int32_t freq_index[FXG_MAX_BLFREQS];           /* freq table index           */
  int32_t npolproducts[FXG_MAX_BLFREQS];
  int32_t band1[FXG_MAX_BLFREQS][FXG_MAX_POLPRODUCTS];
  int32_t band2[FXG_MAX_BLFREQS][FXG_MAX_POLPRODUCTS];
} fxg_baseline;

/* ---- job (one correlator configuration) -------------------------------- */
typedef struct {
  int32_t nfreqs;
  int32_t ndatastreams;
  int32_t nbaselines;
  int32_t blockspersend;   /* FFT blocks per subintegration                     */
  int32_t fringe_order;    /* 0 (post-F), 1 (linear, the vex2difx default), 2  */
  int32_t arraystridelen;  /* 0 = automatic (configuration.cpp:46-59)           */
  int32_t xmacstridelen;   /* 0 = automatic (configuration.cpp:2464-2511)       */
  int32_t crosspol_autocorrs; /* compute and output cross-pol autocorrelations  */
  int32_t xmac_precision;  /* FXG_XMAC_*                                        */
  int32_t max_chunk_blocks;/* blocks of spectra kept in HBM at once (0 = auto)  */
  const fxg_freq*       freqs;        /* [nfreqs]        */
  const fxg_datastream* datastreams;  /* [ndatastreams]  */
  const fxg_baseline*   baselines;    /* [nbaselines]    */
} fxg_job_desc;

/* ---- per-subintegration inputs (Core::processdata, core.cpp:705-730) ---- */
typedef struct {
  const uint8_t* data;     /* raw frames (host or device pointer, see call)     */
  int64_t nbytes;          /* datalengthbytes                                   */
  int32_t scan;            /* data time: controlbuffer[0..2]                    */
  int32_t sec;
  int32_t ns;
  int32_t nvalidflags;     /* ceil(blockspersend/30)                            */
  const int32_t* validflags;   /* host pointer, bit b%30 of word b/30 = block b valid */
  double interpolator[3];  /* quadratic delay model (µs vs block index), after
                              the intclockseconds subtraction (mode.cpp:1548-1549) */
} fxg_station_input;

typedef struct {
  int32_t scan;            /* procslots[].offsets[0..2]                         */
  int32_t sec;
  int32_t ns;
  int32_t scan_start_sec;  /* Model::getScanStartSec(scan, startmjd, startsec)  */
  int32_t model_ok;        /* 0 = no delay model found: subint torched          */
  int32_t pad;
} fxg_subint_desc;

/* ---- per-block station geometry (debug tap; Mode::process prologue) ----- */
typedef struct {
  int32_t status;          /* 0 ok, 1 flagged invalid, 2 out of range, 3 zero weight */
  int32_t nearestsample;
  int32_t integerdelay;
  float   weight;          /* dataweight                                        */
  float   fracsampleerror; /* µs                                                */
  int32_t intwalltime;
  /* delay model of the block by fringe rotation order (mode.cpp:789-838):
   *   order 1: a = slope over the block, b = delay at block start minus
   *            integerdelay, c = 0
   *   order 2: a, b, c = quadratic coefficients in block units, c minus
   *            integerdelay
   *   order 0: a = b = 0, c = delay at the block centre (averagedelay)       */
  double  a;
  double  b;
  double  fracwalltime;
  double  c;
} fxg_block_info;

/* ---- result layout (Configuration::populateResultLengths) --------------- */
typedef struct {
  int64_t result_length;        /* cf32 elements of the core result buffer     */
  int64_t device_bytes;         /* HBM footprint of one job (admission)         */
  int32_t arraystridelen;
  int32_t xmacstridelen;
  int32_t rotatestridelen;
  int32_t bandsperautocorr;
  int32_t fftchannels;          /* L = 2 * nchan for real sampling             */
  int32_t chunk_blocks;         /* blocks of spectra resident per pass          */
} fxg_plan_info;

/* ---- library ------------------------------------------------------------ */
int         fxg_abi_version(void);
const char* fxg_last_error(void);

/* ---- host-only planning: no GPU needed ---------------------------------- */
/* Validates the descriptor and computes the layout and HBM footprint. */
int fxg_plan(const fxg_job_desc* desc, fxg_plan_info* info);
/* Offsets into the cf32 result buffer (-1 if the entry does not exist):
 *   configuration.h getCoreResultBaselineOffset / BWeightOffset /
 *   AutocorrOffset / ACWeightOffset (configuration.cpp:2322-2433). */
int fxg_plan_baseline_offset(const fxg_job_desc* desc, int freq, int baseline, int64_t* off);
int fxg_plan_bweight_offset(const fxg_job_desc* desc, int freq, int baseline, int64_t* off);
int fxg_plan_autocorr_offset(const fxg_job_desc* desc, int datastream, int64_t* off);
int fxg_plan_acweight_offset(const fxg_job_desc* desc, int datastream, int64_t* off);
/* Phase-cal results of a datastream (Configuration::getCoreResultPCalOffset,
 * configuration.cpp:2422-2431): per recorded band in band order, the tones of
 * its frequency (getDRecordedFreqNumPCalTones, configuration.cpp:1771-1845).
 * A datastream without phase cal has no entries there. */
int fxg_plan_pcal_offset(const fxg_job_desc* desc, int datastream, int64_t* off);
/* Tones of a recorded frequency (local index) and the offset of the first
 * tone from the band edge in Hz (-1: no tone inside the band; then one zero
 * entry is kept, as the reference does).  *ntones = 0 without phase cal. */
int fxg_plan_pcal_tones(const fxg_job_desc* desc, int datastream, int recfreq, int32_t* ntones,
                        int32_t* offset_hz);

/* ---- context / device ---------------------------------------------------- */
typedef struct fxg_ctx fxg_ctx;
typedef struct fxg_job fxg_job;

int fxg_ctx_create(int device, fxg_ctx** ctx);
int fxg_ctx_destroy(fxg_ctx* ctx);
/* Free and total HBM of the context's device (admission by footprint). */
int fxg_ctx_memory(fxg_ctx* ctx, int64_t* free_bytes, int64_t* total_bytes);
/* Device allocations for callers that keep their raw data resident. */
int fxg_device_alloc(fxg_ctx* ctx, int64_t bytes, void** dptr);
int fxg_device_free(fxg_ctx* ctx, void* dptr);
int fxg_memcpy_h2d(fxg_ctx* ctx, void* dst, const void* src, int64_t bytes);
int fxg_memcpy_d2h(fxg_ctx* ctx, void* dst, const void* src, int64_t bytes);

/* ---- standalone unpack (Mk5Mode::unpack, mk5mode.cpp:104-177) ----------- */
/* Decodes nsamp sample-times starting at sample offset `sampleoffset` of the
 * frame stream `data` (device pointer, nbytes long) into out[nrecbands][nsamp]
 * (device pointer, band-major floats).  Returns the good-sample weight
 * goodsamples/nsamp in *weight, with the mark5access blanking rules. */
int fxg_unpack(fxg_ctx* ctx, const fxg_datastream* ds, const uint8_t* data,
               int64_t nbytes, int64_t sampleoffset, int32_t nsamp,
               float* out, float* weight);
/* fxg_unpack without any allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results as fxg_unpack; its frame table and station tables live in a
 * static 4 MiB device scratch held by each device until the process ends, so
 * the stream may hold at most (4 MiB - 512) / 8 = 524224 whole frames
 * (FXG_EINVAL beyond).  Every call on one device shares that scratch: calls
 * that overlap on one device (several threads, or several contexts on the
 * same device) race and must be serialised by the caller. */
int fxg_unpack_noalloc(fxg_ctx* ctx, const fxg_datastream* ds, const uint8_t* data,
                       int64_t nbytes, int64_t sampleoffset, int32_t nsamp,
                       float* out, float* weight);

/* ---- jobs ----------------------------------------------------------------- */
int fxg_job_create(fxg_ctx* ctx, const fxg_job_desc* desc, fxg_job** job);
int fxg_job_destroy(fxg_job* job);
int fxg_job_plan_info(const fxg_job* job, fxg_plan_info* info);

/* Process one subintegration: stations[ndatastreams].  If data_on_device is
 * nonzero, stations[i].data are device pointers that stay valid until the call
 * returns; otherwise they are host pointers copied in by the call (on the
 * job's copy stream, into one of two device buffer sets, with stage
 * profiling off; on the job's stream with it on).  results
 * receives result_length cf32 values (interleaved re,im floats), zeroed and
 * accumulated exactly as a Core slot (core.cpp:319, 1048-1108); if
 * results_on_device is nonzero it is a device pointer. */
int fxg_subint_process(fxg_job* job, const fxg_subint_desc* sub,
                       const fxg_station_input* stations, int data_on_device,
                       float* results, int results_on_device);

/* The same, asynchronous: enqueues the subintegration on the job's stream and
 * returns a ticket (core.cpp:539's per-slot call, split so that a request loop
 * can keep several slots of several jobs in flight).  Inputs and results must
 * stay valid until fxg_subint_wait returns for the ticket; with host results
 * (results_on_device == 0) the call may block on the copy.  Tickets of one job
 * complete in submission order.  Host raw data (pinned for an asynchronous
 * copy) of one submission is copied while the previous submission's kernels
 * run. */
int fxg_subint_submit(fxg_job* job, const fxg_subint_desc* sub,
                      const fxg_station_input* stations, int data_on_device,
                      float* results, int results_on_device, int64_t* ticket);
/* *done = 1 when the ticket's work has finished, else 0. */
int fxg_subint_query(fxg_job* job, int64_t ticket, int* done);
/* Blocks until the ticket (and every earlier one of the job) has finished. */
int fxg_subint_wait(fxg_job* job, int64_t ticket);

/* Mode::process for one block of one datastream (mode.cpp:649-1397): the
 * spectra of all its recorded bands (nrecbands x nchan cf32 as float pairs,
 * host pointer, band-major) and the block's geometry and data weight.  The
 * frame validity, corner turn and FP64 geometry of the whole subintegration
 * are computed on the first call and reused while sub/stations describe the
 * same subintegration (same pointers, lengths, times, delay models and valid
 * flags) and the frames the block reads hold the bytes they held then (a
 * 64-bit digest of each frame is compared; a refilled buffer is prepared
 * again). */
int fxg_station_process_block(fxg_job* job, const fxg_subint_desc* sub,
                              const fxg_station_input* stations, int data_on_device,
                              int datastream, int block, float* spectra,
                              fxg_block_info* info);

/* Per-stage device time, measured with HIP events on the job's stream
 * around every launch while profiling is on. */
#define FXG_STAGE_FRAMES    0   /* frame validity                              */
#define FXG_STAGE_PROLOGUE  1   /* FP64 block geometry                         */
#define FXG_STAGE_FSTAGE    2   /* unpack + rotate + FFT + frac rotation       */
#define FXG_STAGE_XMAC      3   /* cross-multiply-accumulate                    */
#define FXG_STAGE_WEIGHTS   4
#define FXG_STAGE_SCATTER   5   /* copy into the result layout                 */
#define FXG_STAGE_H2D       6   /* host -> device copies of raw data           */
#define FXG_STAGE_D2H       7   /* device -> host copy of results              */
#define FXG_STAGE_PCAL      8   /* phase-cal accumulation and tone DFT         */
#define FXG_STAGE_KURTOSIS  9   /* spectral kurtosis accumulation and records  */
#define FXG_NSTAGES        10
int fxg_job_set_profiling(fxg_job* job, int on);
/* ms[FXG_NSTAGES] accumulated milliseconds, launches[FXG_NSTAGES] counts;
 * reset != 0 clears the accumulators after reading. */
int fxg_job_stage_times(fxg_job* job, double* ms, int64_t* launches, int reset);

/* XMAC staging: 1 (default) streams the spectra from HBM into LDS with
 * direct loads where the layout allows (reduced precision: the ring of
 * stages; FP32: the FP32 matrix cores, 16-channel multiples); 0 stages them
 * through registers (reduced precision: the register-staged matrix-core
 * kernel up to 64 inputs; FP32: the VALU register tiles, xmac_ct/xmac_tb).
 * Above 64 inputs reduced precision always streams. */
int fxg_job_set_xmac_streaming(fxg_job* job, int on);

/* ---- VDIF thread multiplexing (libraries/vdifio/src/vdifmux.c) ----------
 * Multi-thread VDIF (one or more channels per thread) becomes single-thread
 * multi-channel VDIF that the station kernels read: configurevdifmux
 * (vdifmux.c:118-211, setvdifmuxinputchannels :213-253) and vdifmux
 * (:507-1044).  The frame search and ordering (stage 1) reads only headers
 * and runs on the host; the corner turn of the payloads (stage 2,
 * cornerturners.c) runs on the GPU into device memory.  Fanout 1, non-legacy
 * headers, real data.  An output frame missing a thread is marked invalid
 * and its payload zeroed; with FXG_MUX_PROPAGATEVALIDITY (what mpifxcorr's
 * datastreams use, vdiffile.cpp:536) the output headers are EDV4 (mask
 * length = threads, sync word 0xACABFEED) and a frame with at least one
 * thread stays valid with the threads present in its validity mask (header
 * bytes 24-31), the missing threads' units zero. */
#define FXG_MUX_MAX_THREADS        31
#define FXG_MUX_GOTOEND          0x01  /* VDIF_MUX_FLAG_GOTOEND            */
#define FXG_MUX_RESPECTGRANULARITY 0x02 /* VDIF_MUX_FLAG_RESPECTGRANULARITY */
#define FXG_MUX_ENABLEVALIDITY   0x04  /* VDIF_MUX_FLAG_ENABLEVALIDITY     */
#define FXG_MUX_PROPAGATEVALIDITY 0x40 /* VDIF_MUX_FLAG_PROPAGATEVALIDITY  */

typedef struct {
  int32_t input_frame_bytes;     /* single-thread frame, 32-byte header included   */
  int32_t frames_per_second;     /* per thread                                     */
  int32_t bits_per_sample;
  int32_t nthreads;
  int32_t thread_ids[FXG_MUX_MAX_THREADS];
  int32_t channels_per_thread;   /* 1 unless setvdifmuxinputchannels               */
  int32_t nsort;                 /* out-of-order frames tolerated                  */
  int32_t ngap;                  /* frame-number gap that ends the call            */
  int32_t flags;                 /* FXG_MUX_*                                      */
} fxg_vdifmux_desc;

/* struct vdif_mux_statistics of one call (vdifio.h:391-425) */
typedef struct {
  int64_t valid_frames, invalid_frames, discarded_frames, wrong_thread_frames;
  int64_t skipped_bytes, fill_bytes, duplicate_frames, bytes_processed;
  int64_t good_frames, partial_frames, filler_frames;
  int64_t src_used, dest_used, output_frame_bytes, output_frames, start_frame_number;
  int64_t output_frame_granularity, epoch;
} fxg_vdifmux_stats;

/* Output frame size and channel count (configurevdifmux). */
int fxg_vdifmux_output_shape(const fxg_vdifmux_desc* desc, int32_t* output_frame_bytes,
                             int32_t* output_channels);
/* vdifmux into dest_device (dest_bytes of device memory): src_host holds the
 * input for the header scan; src_device is the same bytes on the device, or
 * null to have them copied in.  start_frame < 0: from the lowest frame
 * found.  Returns the bytes of input consumed (>= 0) or an FXG_E* code. */
int64_t fxg_vdifmux(fxg_ctx* ctx, const fxg_vdifmux_desc* desc, const uint8_t* src_host,
                    const uint8_t* src_device, int64_t src_bytes, uint8_t* dest_device,
                    int64_t dest_bytes, int64_t start_frame, fxg_vdifmux_stats* stats);

/* ---- launch schedules and the autotuner --------------------------------
 * No counterpart in the reference (its FFT and XMAC are library calls with
 * fixed strategies).  A schedule picks, per job, the reduced-precision XMAC's
 * staging (through registers or streamed into LDS with direct loads), the
 * FP32 XMAC's tile (channels per workgroup, blocks per LDS stage) and the
 * blocks per F/X pass (spectra resident between the F stage and the XMAC).
 * Every schedule computes the same products; FP32 sums grouped differently
 * by a different pass length agree to rounding. */
typedef struct {
  int32_t xmac_streaming;  /* 0 register staging, 1 direct-to-LDS streaming    */
  int32_t xmac_ct;         /* FP32 VALU XMAC channels per workgroup, 0 = auto  */
  int32_t xmac_tb;         /* FP32 VALU XMAC blocks per LDS stage, 0 = auto    */
  int32_t chunk_blocks;    /* blocks per F/X pass, 0 = the plan's chunk_blocks */
} fxg_schedule;

typedef struct {
  fxg_schedule schedule;
  double  ms;              /* mean device time of the subintegration's kernels */
  double  rel_err;         /* ||result - first candidate's|| / ||first's||     */
  int32_t accepted;        /* rel_err within 1e-5                              */
  int32_t pad;
} fxg_tune_result;

int fxg_job_set_schedule(fxg_job* job, const fxg_schedule* schedule);
int fxg_job_get_schedule(const fxg_job* job, fxg_schedule* schedule);
/* Times the candidate schedules on one subintegration (reps timed runs each,
 * after one untimed), keeps the fastest accepted one on the job and reports
 * up to max_results candidates in order of trial. */
int fxg_job_autotune(fxg_job* job, const fxg_subint_desc* sub, const fxg_station_input* stations,
                     int data_on_device, int reps, fxg_tune_result* results, int max_results,
                     int* nresults);

/* 1: a subintegration is one graph launch (captured on the first call and
 * again whenever the launch shape changes: data sizes, result buffer);
 * 0 (default): the kernels are launched one by one.  Stage profiling
 * launches one by one either way.  captures counts the graphs captured. */
int fxg_job_set_graph(fxg_job* job, int on);
int fxg_job_graph_captures(const fxg_job* job, int64_t* captures);

/* ---- batches: one subintegration of several jobs in one graph launch ---
 * No counterpart in the reference (one configuration per Core slot): the
 * jobs' launch sequences are branches of one graph forked from the batch's
 * stream, so small jobs overlap on the device.  stations[k] and results[k]
 * are job k's station array and result buffer (layout of job k).  A job in
 * a batch must not be used elsewhere while fxg_batch_process runs; the call
 * returns when every result is complete. */
typedef struct fxg_batch fxg_batch;
int fxg_batch_create(fxg_ctx* ctx, fxg_job* const* jobs, int njobs, fxg_batch** batch);
int fxg_batch_destroy(fxg_batch* batch);
int fxg_batch_process(fxg_batch* batch, const fxg_subint_desc* subs,
                      const fxg_station_input* const* stations, int data_on_device,
                      float* const* results, int results_on_device);
int fxg_batch_graph_captures(const fxg_batch* batch, int64_t* captures);

/* FP8 jobs: the power-of-two scale of every (band, channel) is set so that
 * the channel's RMS spectrum magnitude maps to 16 in e4m3 (largest value
 * 448): from the quantiser's variance at job creation, and from measured
 * power by this call, which reads the autocorrelations of the last
 * subintegration the job processed.  The products are exact in the scale
 * (the matrix cores apply its inverse), so the scales change rounding only. */
int fxg_job_calibrate_scales(fxg_job* job);

/* ---- spectral kurtosis (Configuration::dumpKurtosis) -------------------
 * Mode::setDumpKurtosis / zeroKurtosis (mode.cpp:1496-1523), the per-block
 * accumulation s1 += |X|^2, s2 += |X|^4 of every recorded band's spectrum
 * (mode.cpp:1260-1277) in block order, and calculateAndAverageKurtosis
 * (:1426-1478) at the end of every subintegration:
 *     sk = (n/(n-1)) (n s2 / s1^2 - 1) - 1,  n = the thread's blocks,
 * summed over groups of nchan/maxchannels channels when maxchannels <
 * nchan.  nthreads records per subintegration, one per Core process thread
 * with the block ranges of Core::updateconfig (core.cpp:2083-2095).
 * maxchannels is Configuration::getSTADumpChannels (32 by default).
 * The spectra are read after the fractional-sample rotation (unit modulus,
 * so the same magnitudes to rounding) at the job's spectrum precision.
 * enable = 0 turns the accumulation off; the buffers are device memory of
 * 2 * nthreads * bands * nchan floats. */
int fxg_job_set_kurtosis(fxg_job* job, int enable, int maxchannels, int nthreads);
/* The record of (thread, datastream, recorded band) of the last
 * subintegration processed (Core::averageAndSendKurtosis, core.cpp:1371-
 * 1420): *nchan = min(maxchannels, nchan) values into out (maxout floats),
 * *valid = calculateAndAverageKurtosis's return for the datastream (nonzero
 * data in its last recorded band).  A band whose s1 is all zero reads zeros.
 * Synchronises with the job's stream. */
int fxg_job_get_kurtosis(fxg_job* job, int thread, int datastream, int band, float* out, int maxout,
                         int32_t* nchan, int32_t* valid);

/* Debug taps valid after fxg_subint_process with keep_taps set. */
int fxg_job_set_keep_taps(fxg_job* job, int keep);
int fxg_tap_block_info(fxg_job* job, int datastream, int block, fxg_block_info* out);
/* Data weight of every recorded band of one block (Mode::getDataWeight,
 * mode.h:168): INTERLACEDVDIF datastreams' per-band weights, else the
 * block's weight for every band; out[nrecbands]. */
int fxg_tap_band_weights(fxg_job* job, int datastream, int block, float* out);
/* Spectrum of one recorded band for one block (Mode::getFreqs, mode.h:199):
 * nchan cf32 values as FP32 (converted from the job's spectrum precision). */
int fxg_tap_spectrum(fxg_job* job, int datastream, int band, int block, float* out);

#ifdef __cplusplus
}
#endif
#endif /* FXCORR_GPU_H */
