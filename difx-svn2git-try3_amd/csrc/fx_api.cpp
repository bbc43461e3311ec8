// fx_api.cpp — the C ABI of libfxcorr_gpu (include/fxcorr_gpu.h): contexts,
// jobs and per-subintegration processing on one gfx950 device.
//
// One subintegration runs as (all on the job's stream):
//   k_frames    frame validity for every datastream
//   k_prologue  FP64 block geometry for every (datastream, block)
//   per chunk of blocks:
//     k_fstage  spectra of every recorded band of every datastream
//     k_xmac    accumulate V[a][b][c] for every XMAC frequency
//   k_weights   baseline / autocorrelation weights
//   k_scatter   V and weights into the core result layout
#include "fx_kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <deque>
#include <functional>
#include <vector>

using namespace fxg;

namespace {

thread_local std::string g_err;

int set_err(int code, const std::string& msg) { g_err = msg; return code; }

#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return set_err(FXG_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));     \
  } while (0)

// device buffer that grows on demand
struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  int ensure(size_t n) {
    if (n <= bytes) return 0;
    if (p) { hipFree(p); p = nullptr; bytes = 0; }
    if (n == 0) return 0;
    if (hipMalloc(&p, n) != hipSuccess) return set_err(FXG_ENOMEM, "hipMalloc of " + std::to_string(n) + " bytes failed");
    bytes = n;
    return 0;
  }
  void release() { if (p) hipFree(p); p = nullptr; bytes = 0; }
  template <typename T> T* as() const { return (T*)p; }
};

}  // namespace

// Device scratch of fxg_unpack_noalloc: [0, 256) the station constants,
// [256, 512) the station's subint entry, [512, kScratchBytes) the frame
// table (int2 per frame).  Each device holds its own copy until the process
// ends; host code reaches it only through the address hipGetSymbolAddress
// gives for that device (scratch_ptr), resolved once per device when a
// context is created on it.
constexpr long long kScratchBytes = 4LL << 20;
constexpr long long kScratchFrames = (kScratchBytes - 512) / (long long)sizeof(int2);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// the device's address of scratch (device current); 0 or an FXG_E* code
int resolve_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return set_err(FXG_ENODEV, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return 0;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return set_err(FXG_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[device] = p;
  return 0;
}

void* scratch_ptr(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return device >= 0 && device < kMaxDevices ? g_scratch_ptr[device] : nullptr;
}
}  // namespace

struct fxg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;   // for standalone calls
  DBuf lut8;
  DBuf mux_src, mux_hdr, mux_mask, mux_off;   // VDIF multiplexing: staged input and frame tables
};

struct fxg_job {
  fxg_ctx* ctx = nullptr;
  Plan plan;
  fxg_job_desc desc{};
  std::vector<fxg_freq> freqs;
  std::vector<fxg_datastream> dss;
  std::vector<fxg_baseline> bls;
  hipStream_t stream = nullptr;
  int keep_taps = 0;
  int have_subint = 0;
  float scale = 1.0f;             // spectrum scale for reduced precision
  // device tables
  DBuf d_st, d_bands, d_segs, d_wterms, d_inoff, d_tw, d_lut8;
  DBuf d_fblank, d_pbw;           // INTERLACEDVDIF: per-frame band blanking, per-(block, band) weights
  std::vector<long long> pbw_off; // per datastream: offset of its weights in d_pbw (-1: none)
  // per-subint
  DBuf d_ss, d_flags, d_ftab, d_blocks, d_weights, d_results, d_ct;
  DBuf d_bandlist;                // global bands grouped by sample size
  int nb2 = 0, nb8 = 0;           // bands of 2-bit and 8-bit datastreams
  std::vector<DBuf> d_raw;        // staged host data per datastream
  // host inputs of fxg_subint_process / fxg_subint_submit: two sets of raw
  // buffers filled on a copy stream, so that one subintegration's PCIe copy
  // runs under the previous one's kernels.  raw_free[k] is recorded on the
  // job's stream after the last kernel of the subintegration that read set
  // k; raw_copied[k] after set k's copies on the copy stream.
  std::vector<DBuf> d_rawp[2];
  hipStream_t copy_stream = nullptr;
  hipEvent_t raw_free[2] = {nullptr, nullptr}, raw_copied[2] = {nullptr, nullptr};
  bool raw_free_rec[2] = {false, false};
  int rawp_next = 0;
  // big buffers
  DBuf d_spec, d_V;
  DBuf d_z;                       // FFTs above 16384 points: the sub-FFT scratch
  DBuf d_rtab;                    // and the rotator tables of its blocks
  DBuf d_pcal, d_pcalaux, d_pcaltone, d_pcalpart;   // phase-cal tables and partial sums
  DBuf d_vpart;                   // block-split XMAC partial matrices (zeroed once)
  DBuf d_cexp, d_eoff;            // fp8: scale exponents [gband][channel], per-input offsets
  long long gstride = 16;         // spectra layout (cidx): 16 row-major, 16 x pass blocks grouped
  int bstride = 0;                // elements between the spectra of consecutive blocks of a band
  // spectral kurtosis: sums [record][s1, s2][gband][C], records [record][gband][nout], flags
  int kurt_on = 0, kurt_maxch = 32, kurt_nrec = 1, kurt_nout = 0, kurt_have = 0;
  DBuf d_kurt, d_kurtout, d_kurtnz;
  std::vector<long long> inoff_base;  // per XMAC freq: offset into d_inoff
  // tap bookkeeping: which chunk is resident in d_spec
  int tap_chunk0 = -1, tap_chunkn = 0;
  // the subintegration whose frames / corner turn / geometry are resident,
  // with the digest of every frame as it was prepared and the block table
  // (the per-block entry point re-prepares when a block's frames changed)
  std::vector<unsigned char> prepared_key;
  std::vector<unsigned long long> prepared_digest;
  std::vector<long long> prepared_fbase;
  std::vector<DevBlock> prepared_blocks;
  DBuf d_digest;
  // pinned host staging of the per-subint tables (station table, valid
  // flags): a ring, so that a copy still queued behind earlier work keeps
  // its source until the copy has run
  struct HostStage {
    DevStationSub* ss = nullptr;
    int* flags = nullptr;
    size_t ss_cap = 0, flags_cap = 0;
    hipEvent_t done = nullptr;
    bool queued = false;
  };
  static constexpr int kStages = 4;
  HostStage stage[kStages];
  int stage_next = 0;
  DBuf d_iota;                    // 0..nbands-1: band lists of single datastreams
  DBuf d_zero;                    // 64 zero bytes: the rows past the data for streaming loads
  int xmac_streaming = 1;         // direct-to-LDS MFMA XMAC where the layout allows
  int xmac_ct = 0, xmac_tb = 0;   // FP32 XMAC channel tile / LDS stage (0: automatic)
  int chunk_blocks = 0;           // blocks per F/X pass (<= the plan's; 0: the plan's)
  // the subintegration as one graph launch, captured per launch shape
  int use_graph = 0;
  hipGraphExec_t graph_exec = nullptr;
  std::vector<unsigned char> graph_key;
  long long graph_captures = 0;
  // asynchronous submissions: (ticket, completion event) in submission order
  std::deque<std::pair<int64_t, hipEvent_t>> pending;
  std::vector<hipEvent_t> free_events;
  int64_t next_ticket = 1;
  // stage profiling
  int profiling = 0;
  std::vector<hipEvent_t> evpool;
  std::vector<std::pair<int, int>> evused;   // (stage, pool index of the start event)
  double stage_ms[FXG_NSTAGES] = {0};
  long long stage_n[FXG_NSTAGES] = {0};
  std::vector<int> evfree;                     // free pairs of evpool (index of the start event)
  int ev_begin(int stage) {
    if (!profiling) return 0;
    if (evfree.empty()) {
      for (int k = 0; k < 2; k++) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return set_err(FXG_EHIP, "hipEventCreate");
        evpool.push_back(e);
      }
      evfree.push_back((int)evpool.size() - 2);
    }
    int idx = evfree.back();
    evfree.pop_back();
    evused.push_back({stage, idx});
    if (hipEventRecord(evpool[idx], stream) != hipSuccess) return set_err(FXG_EHIP, "hipEventRecord");
    return 0;
  }
  int ev_end() {
    if (!profiling) return 0;
    if (hipEventRecord(evpool[evused.back().second + 1], stream) != hipSuccess) return set_err(FXG_EHIP, "hipEventRecord");
    return 0;
  }
  // accumulate the stages whose events have completed; keep the rest
  void ev_collect() {
    std::vector<std::pair<int, int>> keep;
    for (auto& u : evused) {
      float ms = 0.f;
      hipError_t e = hipEventElapsedTime(&ms, evpool[u.second], evpool[u.second + 1]);
      if (e == hipErrorNotReady) { keep.push_back(u); continue; }
      if (e == hipSuccess) {
        stage_ms[u.first] += ms;
        stage_n[u.first] += 1;
      }
      evfree.push_back(u.second);
    }
    evused.swap(keep);
  }
};

namespace fxg {

int set_error(int code, const char* msg) { return set_err(code, msg); }

// stage 2 of VDIF multiplexing on the context's stream (the frame tables
// come from the host-side stage 1 in fx_mux.cpp)
int ctx_mux_launch(fxg_ctx* c, const uint8_t* src, const uint8_t* src_device, long long srcsize, uint8_t* dest,
                   const uint32_t* hdr, const unsigned long long* mask, const long long* off, long long nframes,
                   int nthreads, int unitbits, int units, int outwords, int outframe, unsigned long long good,
                   int partial) {
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  int rc;
  const uint8_t* dsrc = src_device;
  if (!dsrc) {
    if ((rc = c->mux_src.ensure((size_t)std::max(1LL, srcsize)))) return rc;
    HIPCHK(hipMemcpyAsync(c->mux_src.p, src, (size_t)srcsize, hipMemcpyHostToDevice, s));
    dsrc = c->mux_src.as<uint8_t>();
  }
  if ((rc = c->mux_hdr.ensure(sizeof(uint32_t) * 8 * (size_t)nframes)) ||
      (rc = c->mux_mask.ensure(sizeof(unsigned long long) * (size_t)nframes)) ||
      (rc = c->mux_off.ensure(sizeof(long long) * (size_t)nframes * nthreads)))
    return rc;
  HIPCHK(hipMemcpyAsync(c->mux_hdr.p, hdr, sizeof(uint32_t) * 8 * (size_t)nframes, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(c->mux_mask.p, mask, sizeof(unsigned long long) * (size_t)nframes, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(c->mux_off.p, off, sizeof(long long) * (size_t)nframes * nthreads, hipMemcpyHostToDevice, s));
  MuxArgs a{};
  a.src = dsrc; a.dest = dest; a.srcoff = c->mux_off.as<long long>();
  a.mask = c->mux_mask.as<unsigned long long>(); a.headers = c->mux_hdr.as<unsigned>();
  a.nframes = (int)nframes; a.nthreads = nthreads; a.unitbits = unitbits; a.unitsper = units;
  a.outwords = outwords; a.framebytes = outframe; a.goodmask = good; a.partial = partial;
  HIPCHK(launch_vdifmux(a, s));
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

}  // namespace fxg

extern "C" {

int fxg_abi_version(void) { return FXG_ABI_VERSION; }
const char* fxg_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------------------------------------
// planning (host only)
// ---------------------------------------------------------------------------
int fxg_plan(const fxg_job_desc* desc, fxg_plan_info* info) {
  if (!info) return set_err(FXG_EINVAL, "null info");
  Plan p;
  int rc = build_plan(desc, &p, &g_err);
  if (rc) return rc;
  *info = p.info;
  return 0;
}

static int plan_lookup(const fxg_job_desc* desc, Plan* p) { return build_plan(desc, p, &g_err); }

int fxg_plan_baseline_offset(const fxg_job_desc* desc, int freq, int baseline, int64_t* off) {
  Plan p; int rc = plan_lookup(desc, &p); if (rc) return rc;
  if (!off || freq < 0 || freq >= p.nfreqs || baseline < 0 || baseline >= p.nbl) return set_err(FXG_EINVAL, "index");
  *off = p.bloff[freq][baseline];
  return 0;
}
int fxg_plan_bweight_offset(const fxg_job_desc* desc, int freq, int baseline, int64_t* off) {
  Plan p; int rc = plan_lookup(desc, &p); if (rc) return rc;
  if (!off || freq < 0 || freq >= p.nfreqs || baseline < 0 || baseline >= p.nbl) return set_err(FXG_EINVAL, "index");
  *off = p.bwoff[freq][baseline];
  return 0;
}
int fxg_plan_autocorr_offset(const fxg_job_desc* desc, int ds, int64_t* off) {
  Plan p; int rc = plan_lookup(desc, &p); if (rc) return rc;
  if (!off || ds < 0 || ds >= p.nds) return set_err(FXG_EINVAL, "index");
  *off = p.acoff[ds];
  return 0;
}
int fxg_plan_acweight_offset(const fxg_job_desc* desc, int ds, int64_t* off) {
  Plan p; int rc = plan_lookup(desc, &p); if (rc) return rc;
  if (!off || ds < 0 || ds >= p.nds) return set_err(FXG_EINVAL, "index");
  *off = p.acwoff[ds];
  return 0;
}
int fxg_plan_pcal_offset(const fxg_job_desc* desc, int ds, int64_t* off) {
  Plan p; int rc = plan_lookup(desc, &p); if (rc) return rc;
  if (!off || ds < 0 || ds >= p.nds) return set_err(FXG_EINVAL, "index");
  *off = p.pcaloff[ds];
  return 0;
}
int fxg_plan_pcal_tones(const fxg_job_desc* desc, int ds, int recfreq, int32_t* ntones, int32_t* offset_hz) {
  Plan p; int rc = plan_lookup(desc, &p); if (rc) return rc;
  if (!ntones || !offset_hz || ds < 0 || ds >= p.nds) return set_err(FXG_EINVAL, "index");
  if (p.pcal_ntones[ds].empty()) { *ntones = 0; *offset_hz = -1; return 0; }
  if (recfreq < 0 || recfreq >= (int)p.pcal_ntones[ds].size()) return set_err(FXG_EINVAL, "index");
  *ntones = p.pcal_ntones[ds][recfreq];
  *offset_hz = p.pcal_offhz[ds][recfreq];
  return 0;
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
static void lut8_values(float* t) {
  for (int b = 0; b < 256; b++) t[b] = (b - 128) / 3.3;   // format_vdif.c:126
}

int fxg_ctx_create(int device, fxg_ctx** out) {
  if (!out) return set_err(FXG_EINVAL, "null out");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return set_err(FXG_ENODEV, "no HIP device");
  if (device < 0 || device >= n) return set_err(FXG_ENODEV, "device index out of range");
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, device));
  if (std::string(prop.gcnArchName).find("gfx950") == std::string::npos)
    return set_err(FXG_ENODEV, std::string("device is ") + prop.gcnArchName + ", this library is built for gfx950");
  HIPCHK(hipSetDevice(device));
  int src = resolve_scratch(device);
  if (src) return src;
  fxg_ctx* c = new fxg_ctx();
  c->device = device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return set_err(FXG_EHIP, "stream creation failed");
  }
  float lut[256];
  lut8_values(lut);
  if (c->lut8.ensure(sizeof(lut))) { hipStreamDestroy(c->stream); delete c; return FXG_ENOMEM; }
  HIPCHK(hipMemcpy(c->lut8.p, lut, sizeof(lut), hipMemcpyHostToDevice));
  *out = c;
  return 0;
}

int fxg_ctx_memory(fxg_ctx* c, int64_t* free_bytes, int64_t* total_bytes) {
  if (!c || !free_bytes || !total_bytes) return set_err(FXG_EINVAL, "null argument");
  HIPCHK(hipSetDevice(c->device));
  size_t fr = 0, tot = 0;
  HIPCHK(hipMemGetInfo(&fr, &tot));
  *free_bytes = (int64_t)fr;
  *total_bytes = (int64_t)tot;
  return 0;
}

int fxg_ctx_destroy(fxg_ctx* c) {
  if (!c) return 0;
  hipSetDevice(c->device);
  if (c->stream) hipStreamDestroy(c->stream);
  c->lut8.release();
  for (DBuf* b : {&c->mux_src, &c->mux_hdr, &c->mux_mask, &c->mux_off}) b->release();
  delete c;
  return 0;
}

int fxg_device_alloc(fxg_ctx* c, int64_t bytes, void** p) {
  if (!c || !p || bytes < 0) return set_err(FXG_EINVAL, "bad args");
  HIPCHK(hipSetDevice(c->device));
  if (hipMalloc(p, (size_t)std::max<int64_t>(bytes, 1)) != hipSuccess) return set_err(FXG_ENOMEM, "hipMalloc failed");
  return 0;
}
int fxg_device_free(fxg_ctx* c, void* p) {
  if (!c) return set_err(FXG_EINVAL, "null ctx");
  HIPCHK(hipSetDevice(c->device));
  if (p) HIPCHK(hipFree(p));
  return 0;
}
int fxg_memcpy_h2d(fxg_ctx* c, void* dst, const void* src, int64_t bytes) {
  if (!c || bytes < 0) return set_err(FXG_EINVAL, "bad args");
  HIPCHK(hipSetDevice(c->device));
  if (bytes) HIPCHK(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice));
  return 0;
}
int fxg_memcpy_d2h(fxg_ctx* c, void* dst, const void* src, int64_t bytes) {
  if (!c || bytes < 0) return set_err(FXG_EINVAL, "bad args");
  HIPCHK(hipSetDevice(c->device));
  if (bytes) HIPCHK(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost));
  return 0;
}

// ---------------------------------------------------------------------------
// standalone unpack: frames -> floats for one datastream
// ---------------------------------------------------------------------------
}  // extern "C"

// The decoder constants of a standalone unpack: a one-station, one-freq job
// description of the datastream (fxg_unpack and fxg_unpack_noalloc).
static int unpack_station(const fxg_datastream* ds, DevStation* S) {
  fxg_freq fq{};
  fq.freq_mhz = 1000.0; fq.bw_mhz = 16.0; fq.nchan = 64; fq.chans_to_avg = 1;
  fxg_datastream d = *ds;
  for (int i = 0; i < d.nrecfreqs && i < FXG_MAX_RECFREQS; i++) d.recfreq_index[i] = 0;
  fxg_job_desc jd{};
  jd.nfreqs = 1; jd.ndatastreams = 1; jd.nbaselines = 0; jd.blockspersend = 1; jd.fringe_order = 1;
  jd.freqs = &fq; jd.datastreams = &d; jd.baselines = nullptr;
  Plan p;
  int rc = build_plan(&jd, &p, &g_err);
  if (rc) return rc;
  *S = p.st[0];
  if (S->cplx) return set_err(FXG_EUNSUPPORTED, "fxg_unpack of complex samples (fxg_tap_spectrum gives their spectra)");
  return 0;
}

// Mk5Mode::unpack's window [ns, ns + nsamp) over the frame table: the first
// sample that decodes as zero for 8-bit streams that stop at a blank frame
// (validend), and the count of good samples.  frame(f) gives frame f's valid
// payload byte range (f < nframes).
template <typename FrameFn>
static void unpack_window(const DevStation& S, long long nframes, long long ns, int nsamp, FrameFn frame,
                          long long* validend_out, long long* good_out) {
  const long long tend = ns + nsamp;
  const long long unpackstart = ns - ns % S.granularity;
  long long validend = 0x7fffffffffffffffLL;
  if (S.quirk8) {
    for (long long f = unpackstart / S.framesamples; f <= (tend - 1) / S.framesamples; f++) {
      bool fv = false;
      if (f < nframes) { const int2 z = frame(f); fv = z.x < z.y; }
      if (!fv) { validend = std::max(unpackstart, f * (long long)S.framesamples); break; }
    }
  }
  long long good = 0;
  for (long long t = ns; t < std::min(tend, validend);) {
    long long f = t / S.framesamples;
    long long fe = std::min(std::min(tend, validend), (f + 1) * (long long)S.framesamples);
    if (f < nframes) {
      const int2 z = frame(f);
      if (z.x < z.y) {
        long long base = f * (long long)S.framesamples;
        long long g0 = ((long long)z.x + S.groupbytes - 1) / S.groupbytes;
        long long g1 = ((long long)z.y + S.groupbytes - 1) / S.groupbytes;
        long long a = std::max(t - base, g0 * S.groupsamples), b = std::min(fe - base, g1 * S.groupsamples);
        if (b > a) good += b - a;
      }
    }
    t = fe;
  }
  *validend_out = validend;
  *good_out = good;
}

extern "C" {

int fxg_unpack(fxg_ctx* c, const fxg_datastream* ds, const uint8_t* data, int64_t nbytes,
               int64_t sampleoffset, int32_t nsamp, float* out, float* weight) {
  if (!c || !ds || !data || !out || nsamp <= 0 || sampleoffset < 0 || nbytes < 0)
    return set_err(FXG_EINVAL, "bad args");
  DevStation S;
  int rc = unpack_station(ds, &S);
  if (rc) return rc;
  HIPCHK(hipSetDevice(c->device));
  DBuf d_st, d_ss, d_ftab;
  if ((rc = d_st.ensure(sizeof(DevStation))) || (rc = d_ss.ensure(sizeof(DevStationSub)))) return rc;
  DevStationSub Q{};
  Q.data = data; Q.nbytes = nbytes; Q.nframes = nbytes / S.framebytes; Q.fbase = 0;
  if ((rc = d_ftab.ensure(sizeof(int2) * std::max<long long>(1, Q.nframes)))) return rc;
  HIPCHK(hipMemcpyAsync(d_st.p, &S, sizeof(S), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_ss.p, &Q, sizeof(Q), hipMemcpyHostToDevice, c->stream));
  HIPCHK(launch_frames(d_st.as<DevStation>(), d_ss.as<DevStationSub>(), d_ftab.as<int2>(), 1, Q.nframes, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  // the window's validity is computed on the host from the frame table
  std::vector<int2> ftab((size_t)std::max<long long>(1, Q.nframes));
  if (Q.nframes) HIPCHK(hipMemcpy(ftab.data(), d_ftab.p, sizeof(int2) * Q.nframes, hipMemcpyDeviceToHost));
  long long validend = 0, good = 0;
  unpack_window(S, Q.nframes, sampleoffset, nsamp, [&](long long f) { return ftab[(size_t)f]; }, &validend, &good);
  HIPCHK(launch_unpack(d_st.as<DevStation>(), d_ss.as<DevStationSub>(), d_ftab.as<int2>(), c->lut8.as<float>(),
                       sampleoffset, nsamp, S.nchan, validend, out, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (weight) *weight = (float)good / (float)nsamp;
  return 0;
}

int fxg_unpack_noalloc(fxg_ctx* c, const fxg_datastream* ds, const uint8_t* data, int64_t nbytes,
                       int64_t sampleoffset, int32_t nsamp, float* out, float* weight) {
  if (!c || !ds || !data || !out || nsamp <= 0 || sampleoffset < 0 || nbytes < 0)
    return set_err(FXG_EINVAL, "bad args");
  DevStation S;
  int rc = unpack_station(ds, &S);
  if (rc) return rc;
  DevStationSub Q{};
  Q.data = data; Q.nbytes = nbytes; Q.nframes = nbytes / S.framebytes; Q.fbase = 0;
  if (Q.nframes > kScratchFrames)
    return set_err(FXG_EINVAL, "fxg_unpack_noalloc: " + std::to_string(Q.nframes) + " frames exceed the scratch's " +
                                   std::to_string(kScratchFrames) + " (use fxg_unpack)");
  HIPCHK(hipSetDevice(c->device));
  unsigned char* base = (unsigned char*)scratch_ptr(c->device);
  if (!base) return set_err(FXG_ESTATE, "scratch address of the device not resolved");
  DevStation* d_st = (DevStation*)base;
  DevStationSub* d_ss = (DevStationSub*)(base + 256);
  int2* d_ftab = (int2*)(base + 512);
  static_assert(sizeof(DevStation) <= 256 && sizeof(DevStationSub) <= 256, "scratch header layout");
  // (S and Q outlive the copies: the stream is synchronised below)
  HIPCHK(hipMemcpyAsync(d_st, &S, sizeof(S), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_ss, &Q, sizeof(Q), hipMemcpyHostToDevice, c->stream));
  HIPCHK(launch_frames(d_st, d_ss, d_ftab, 1, Q.nframes, c->stream));
  // the frame table entries the window reads, fetched in pieces into a
  // fixed host buffer (k_frames wrote every entry below Q.nframes)
  int2 piece[512];
  long long piece0 = -1;
  int err = 0;
  auto frame = [&](long long f) -> int2 {
    if (err) return make_int2(kBig, 0);
    if (piece0 < 0 || f < piece0 || f >= piece0 + 512) {
      piece0 = f;
      const long long n = std::min<long long>(512, Q.nframes - f);
      if (hipMemcpyAsync(piece, d_ftab + f, sizeof(int2) * (size_t)n, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
          hipStreamSynchronize(c->stream) != hipSuccess) { err = 1; return make_int2(kBig, 0); }
    }
    return piece[f - piece0];
  };
  long long validend = 0, good = 0;
  unpack_window(S, Q.nframes, sampleoffset, nsamp, frame, &validend, &good);
  if (err) return set_err(FXG_EHIP, "frame table read-back failed");
  HIPCHK(launch_unpack(d_st, d_ss, d_ftab, c->lut8.as<float>(), sampleoffset, nsamp, S.nchan, validend, out,
                       c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (weight) *weight = (float)good / (float)nsamp;
  return 0;
}

// ---------------------------------------------------------------------------
// jobs
// ---------------------------------------------------------------------------
static void job_free(fxg_job* j) {
  DBuf* bufs[] = {&j->d_st, &j->d_bands, &j->d_segs, &j->d_wterms, &j->d_fblank, &j->d_pbw, &j->d_inoff, &j->d_tw,
                  &j->d_lut8,
                  &j->d_ss, &j->d_flags, &j->d_ftab, &j->d_blocks, &j->d_weights, &j->d_results,
                  &j->d_spec, &j->d_V, &j->d_ct, &j->d_bandlist, &j->d_iota, &j->d_zero, &j->d_z, &j->d_rtab,
                  &j->d_pcal, &j->d_pcalaux, &j->d_pcaltone, &j->d_pcalpart, &j->d_vpart, &j->d_cexp,
                  &j->d_eoff, &j->d_kurt, &j->d_kurtout, &j->d_kurtnz};
  for (DBuf* b : bufs) b->release();
  for (DBuf& b : j->d_raw) b.release();
  if (j->copy_stream) { (void)hipStreamSynchronize(j->copy_stream); (void)hipStreamDestroy(j->copy_stream); j->copy_stream = nullptr; }
  for (int k = 0; k < 2; k++) {
    for (DBuf& b : j->d_rawp[k]) b.release();
    if (j->raw_free[k]) (void)hipEventDestroy(j->raw_free[k]);
    if (j->raw_copied[k]) (void)hipEventDestroy(j->raw_copied[k]);
    j->raw_free[k] = j->raw_copied[k] = nullptr;
  }
  j->d_digest.release();
  for (auto& hs : j->stage) {
    if (hs.done) (void)hipEventSynchronize(hs.done), (void)hipEventDestroy(hs.done);
    if (hs.ss) (void)hipHostFree(hs.ss);
    if (hs.flags) (void)hipHostFree(hs.flags);
    hs = fxg_job::HostStage{};
  }
  for (hipEvent_t e : j->evpool) hipEventDestroy(e);
  j->evpool.clear();
  for (auto& p : j->pending) hipEventDestroy(p.second);
  j->pending.clear();
  for (hipEvent_t e : j->free_events) hipEventDestroy(e);
  j->free_events.clear();
  if (j->graph_exec) (void)hipGraphExecDestroy(j->graph_exec);
  j->graph_exec = nullptr;
  if (j->stream) hipStreamDestroy(j->stream);
}

int fxg_job_create(fxg_ctx* c, const fxg_job_desc* desc, fxg_job** out) {
  if (!c || !desc || !out) return set_err(FXG_EINVAL, "null argument");
  HIPCHK(hipSetDevice(c->device));
  fxg_job* j = new fxg_job();
  j->ctx = c;
  int rc = build_plan(desc, &j->plan, &g_err);
  if (rc) { delete j; return rc; }
  Plan& P = j->plan;
  {
    // device-side limits of this build
    std::string why;
    if (!P.uniform) why = "all datastreams of a job must share nchan";
    else if (!(P.L >= 128 && P.L <= 65536 && (P.L & (P.L - 1)) == 0)) why = "FFT length must be a power of two in [128, 65536]";
    else if (P.st[0].cplx && P.L > kFStageMaxL) why = "complex sampling: at most 16384 channels";
    for (const DevStation& st : P.st) {
      const int tabs = fxg_rotator_table_entries(st.L);
      const int rotpts = (st.order == 2 ? 2 : 1) * st.s + st.L / st.s, fracpts = st.s + st.C / st.s;
      // one workgroup holds both sets up to kFStageMaxL; above, the two kernels hold one each
      const bool fits = st.L <= kFStageMaxL ? rotpts + fracpts <= tabs : (rotpts <= tabs && fracpts <= tabs);
      if (why.empty() && !fits) why = "array stride gives more rotator sub/step points than the F stage holds";
      if (why.empty() && st.L > kFStageMaxL && st.slog < 0) why = "FFTs above 16384 points need a power-of-two array stride";
    }
    if (!why.empty()) { delete j; return set_err(FXG_EUNSUPPORTED, why); }
  }
  // own copies of the descriptor tables
  j->desc = *desc;
  j->freqs.assign(desc->freqs, desc->freqs + desc->nfreqs);
  j->dss.assign(desc->datastreams, desc->datastreams + desc->ndatastreams);
  if (desc->nbaselines) j->bls.assign(desc->baselines, desc->baselines + desc->nbaselines);
  j->desc.freqs = j->freqs.data();
  j->desc.datastreams = j->dss.data();
  j->desc.baselines = j->bls.empty() ? nullptr : j->bls.data();
  auto fail = [&](int code) { job_free(j); delete j; return code; };
  if (hipStreamCreateWithFlags(&j->stream, hipStreamNonBlocking) != hipSuccess) return fail(set_err(FXG_EHIP, "stream"));
  // the host-input copy stream, created right after the job's own so that
  // the runtime's round of hardware queues puts the two on different queues
  // (a copy queued behind the job's kernels could not overlap them)
  if (hipStreamCreateWithFlags(&j->copy_stream, hipStreamNonBlocking) != hipSuccess)
    return fail(set_err(FXG_EHIP, "copy stream"));

  const int prec = desc->xmac_precision;
  // spectrum scale for reduced precision: 2^-round(log2(sqrt(L)))
  // (fp8: per-channel exponents instead, below)
  j->scale = (prec == FXG_XMAC_FP32 || prec == FXG_XMAC_FP8)
                 ? 1.0f : std::ldexp(1.0f, -(int)std::lround(0.5 * std::log2((double)P.L)));

  // spectra offsets: band g's chunk of spectra at g * chunk * C; within it
  // channel-group-major [C/16][chunk][16] when every cross-multiply of the
  // job runs on the matrix cores in 16-channel groups that start on a group
  // boundary (each workgroup then streams whole lines), else row-major
  const long long chunk = P.chunk_blocks;
  for (size_t g = 0; g < P.bands.size(); g++) P.bands[g].specoff = (long long)g * chunk * P.C;
  {
    // (FFTs above 16384 points write their channels from a separate pass
    // whose scattered 64-byte pieces cost more than the grouped reads save)
    bool grouped = prec != FXG_XMAC_FP32 && P.C % 16 == 0 && P.L <= kFStageMaxL;
    for (const XmacFreq& x : P.xf) {
      if (!xmac_mfma_streaming_supported((int)x.in_station.size(), x.nchan)) grouped = false;
      for (long long o : x.in_off) if (o % 16) grouped = false;     // zoom channel offsets
    }
    j->gstride = grouped ? 16 * chunk : 16;
    j->bstride = grouped ? 16 : P.C;
  }
  std::vector<long long> inoff;
  std::vector<int> eoff;            // fp8: [gband][channel] index of each input's channel 0
  for (XmacFreq& x : P.xf) {
    j->inoff_base.push_back((long long)inoff.size());
    for (size_t i = 0; i < x.in_station.size(); i++) {
      int gband = P.st[x.in_station[i]].band0 + x.in_band[i];
      eoff.push_back(gband * P.C + (int)x.in_off[i]);
      x.in_off[i] = P.bands[gband].specoff + cidx((int)x.in_off[i], j->gstride);
      inoff.push_back(x.in_off[i]);
    }
  }
  std::vector<DevWTerm> wt = P.wterms;
  if (wt.empty()) wt.push_back(DevWTerm{0, -1, -1, -1});
  // INTERLACEDVDIF datastreams: their per-(block, band) weights
  j->pbw_off.assign(P.nds, -1);
  long long npbw = 0;
  for (int d = 0; d < P.nds; d++)
    if (P.st[d].edv4) { j->pbw_off[d] = npbw; npbw += (long long)desc->blockspersend * P.st[d].nchan; }
  // split twiddle tables e^{-2 pi i m/L} = hi[m>>6] * lo[m&63], each entry
  // evaluated in double and rounded once: lo[64] then hi[L/64]
  // then e^{-2 pi i m/256}, m < 256: the second FFT pass's twiddles, read whole
  std::vector<float2> tw(64 + P.L / 64 + 256);
  for (int k = 0; k < 64; k++) {
    double ang = -kTwoPi * (double)k / (double)P.L;
    tw[k] = make_float2((float)std::cos(ang), (float)std::sin(ang));
  }
  for (int h = 0; h < P.L / 64; h++) {
    double ang = -kTwoPi * (double)(64 * h) / (double)P.L;
    tw[64 + h] = make_float2((float)std::cos(ang), (float)std::sin(ang));
  }
  for (int m = 0; m < 256; m++) {
    double ang = -kTwoPi * (double)m / 256.0;
    tw[64 + P.L / 64 + m] = make_float2((float)std::cos(ang), (float)std::sin(ang));
  }
  size_t specbytes = (size_t)P.nbands_total * chunk * P.C * spectrum_bytes(prec);
  const size_t zbytes = P.L > kFStageMaxL ? (size_t)P.nbands_total * P.zcap * P.L * sizeof(float2) : 0;
  const size_t rtbytes = P.L > kFStageMaxL ? (size_t)P.nbands_total * P.zcap * P.rtabn * sizeof(float2) : 0;
  if ((rc = j->d_st.ensure(sizeof(DevStation) * P.st.size())) ||
      (rc = j->d_bands.ensure(sizeof(DevBand) * P.bands.size())) ||
      (rc = j->d_segs.ensure(sizeof(DevSeg) * std::max<size_t>(1, P.segs.size()))) ||
      (rc = j->d_wterms.ensure(sizeof(DevWTerm) * wt.size())) ||
      (npbw > 0 && (rc = j->d_pbw.ensure(sizeof(float) * (size_t)npbw))) ||
      (rc = j->d_inoff.ensure(sizeof(long long) * std::max<size_t>(1, inoff.size()))) ||
      (rc = j->d_tw.ensure(sizeof(float2) * tw.size())) ||
      (rc = j->d_blocks.ensure(sizeof(DevBlock) * (size_t)P.nds * desc->blockspersend)) ||
      (rc = j->d_weights.ensure(sizeof(float) * wt.size())) ||
      (rc = j->d_results.ensure(sizeof(float) * 2 * (size_t)std::max<long long>(1, P.info.result_length))) ||
      (rc = j->d_spec.ensure(specbytes)) ||
      (zbytes && (rc = j->d_z.ensure(zbytes))) ||
      (rtbytes && (rc = j->d_rtab.ensure(rtbytes))) ||
      (rc = j->d_V.ensure(sizeof(float2) * (size_t)std::max<long long>(1, P.vsize))))
    return fail(rc);
  if (hipMemcpy(j->d_st.p, P.st.data(), sizeof(DevStation) * P.st.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(j->d_bands.p, P.bands.data(), sizeof(DevBand) * P.bands.size(), hipMemcpyHostToDevice) != hipSuccess ||
      (P.segs.size() && hipMemcpy(j->d_segs.p, P.segs.data(), sizeof(DevSeg) * P.segs.size(), hipMemcpyHostToDevice) != hipSuccess) ||
      hipMemcpy(j->d_wterms.p, wt.data(), sizeof(DevWTerm) * wt.size(), hipMemcpyHostToDevice) != hipSuccess ||
      (inoff.size() && hipMemcpy(j->d_inoff.p, inoff.data(), sizeof(long long) * inoff.size(), hipMemcpyHostToDevice) != hipSuccess) ||
      hipMemcpy(j->d_tw.p, tw.data(), sizeof(float2) * tw.size(), hipMemcpyHostToDevice) != hipSuccess)
    return fail(set_err(FXG_EHIP, "table upload failed"));
  // band lists per sample size (one F-stage launch per size)
  std::vector<int> bl2, bl8;
  for (size_t g = 0; g < P.bands.size(); g++) (P.st[P.bands[g].station].nbits == 8 ? bl8 : bl2).push_back((int)g);
  j->nb2 = (int)bl2.size(); j->nb8 = (int)bl8.size();
  std::vector<int> bl = bl2; bl.insert(bl.end(), bl8.begin(), bl8.end());
  if ((rc = j->d_bandlist.ensure(sizeof(int) * std::max<size_t>(1, bl.size())))) return fail(rc);
  if (hipMemcpy(j->d_bandlist.p, bl.data(), sizeof(int) * bl.size(), hipMemcpyHostToDevice) != hipSuccess)
    return fail(set_err(FXG_EHIP, "band list upload failed"));
  std::vector<int> iota(P.bands.size());
  for (size_t g = 0; g < iota.size(); g++) iota[g] = (int)g;
  if ((rc = j->d_iota.ensure(sizeof(int) * std::max<size_t>(1, iota.size())))) return fail(rc);
  if (hipMemcpy(j->d_iota.p, iota.data(), sizeof(int) * iota.size(), hipMemcpyHostToDevice) != hipSuccess)
    return fail(set_err(FXG_EHIP, "band index upload failed"));
  if ((rc = j->d_zero.ensure(64))) return fail(rc);
  if (hipMemset(j->d_zero.p, 0, 64) != hipSuccess) return fail(set_err(FXG_EHIP, "zero row"));
  if (prec == FXG_XMAC_FP8) {
    // default scales: RMS of an L-point transform of the quantised samples
    // (variance 4.31 for 2-bit levels at +-0.98 sigma thresholds; about 6 for
    // the 8-bit levels of unit-variance data) mapped to 16
    std::vector<signed char> ex((size_t)P.nbands_total * P.C);
    for (size_t g = 0; g < P.bands.size(); g++) {
      const double var = P.st[P.bands[g].station].nbits == 8 ? 6.0 : 4.31;
      const int e = (int)std::lround(std::log2(16.0 / std::sqrt((double)P.L * var)));
      std::fill(ex.begin() + g * P.C, ex.begin() + (g + 1) * P.C, (signed char)e);
    }
    if ((rc = j->d_cexp.ensure(ex.size())) || (rc = j->d_eoff.ensure(sizeof(int) * std::max<size_t>(1, eoff.size()))))
      return fail(rc);
    if (hipMemcpy(j->d_cexp.p, ex.data(), ex.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(j->d_eoff.p, eoff.data(), sizeof(int) * eoff.size(), hipMemcpyHostToDevice) != hipSuccess)
      return fail(set_err(FXG_EHIP, "scale table upload failed"));
  }
  if (P.vpart_elems > 0) {
    if ((rc = j->d_vpart.ensure(sizeof(float2) * (size_t)P.vpart_elems))) return fail(rc);
    if (hipMemset(j->d_vpart.p, 0, sizeof(float2) * (size_t)P.vpart_elems) != hipSuccess)
      return fail(set_err(FXG_EHIP, "partial matrix clear"));
  }
  if (!P.pcal.empty()) {
    if ((rc = j->d_pcal.ensure(sizeof(DevPcal) * P.pcal.size())) ||
        (rc = j->d_pcalaux.ensure(sizeof(float) * P.pcal_aux.size())) ||
        (rc = j->d_pcaltone.ensure(sizeof(int) * std::max<size_t>(1, P.pcal_tonebin.size()))) ||
        (rc = j->d_pcalpart.ensure(sizeof(float) * (size_t)P.pcal_partfloats)))
      return fail(rc);
    if (hipMemcpy(j->d_pcal.p, P.pcal.data(), sizeof(DevPcal) * P.pcal.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(j->d_pcalaux.p, P.pcal_aux.data(), sizeof(float) * P.pcal_aux.size(), hipMemcpyHostToDevice) != hipSuccess ||
        (P.pcal_tonebin.size() && hipMemcpy(j->d_pcaltone.p, P.pcal_tonebin.data(), sizeof(int) * P.pcal_tonebin.size(),
                                            hipMemcpyHostToDevice) != hipSuccess))
      return fail(set_err(FXG_EHIP, "phase-cal table upload failed"));
  }
  j->d_raw.resize(P.nds);
  *out = j;
  return 0;
}

int fxg_job_destroy(fxg_job* j) {
  if (!j) return 0;
  hipSetDevice(j->ctx->device);
  if (j->stream) hipStreamSynchronize(j->stream);
  job_free(j);
  delete j;
  return 0;
}

int fxg_job_plan_info(const fxg_job* j, fxg_plan_info* info) {
  if (!j || !info) return set_err(FXG_EINVAL, "null argument");
  *info = j->plan.info;
  return 0;
}

int fxg_job_set_xmac_streaming(fxg_job* j, int on) {   // one entry of the schedule
  if (!j) return set_err(FXG_EINVAL, "null job");
  j->xmac_streaming = on ? 1 : 0;
  return 0;
}

int fxg_job_calibrate_scales(fxg_job* j) {
  if (!j) return set_err(FXG_EINVAL, "null job");
  if (j->desc.xmac_precision != FXG_XMAC_FP8) return set_err(FXG_EINVAL, "not an fp8 job");
  if (!j->have_subint) return set_err(FXG_ESTATE, "no subintegration processed");
  HIPCHK(hipSetDevice(j->ctx->device));
  HIPCHK(hipStreamSynchronize(j->stream));
  const Plan& P = j->plan;
  std::vector<signed char> ex((size_t)P.nbands_total * P.C);
  HIPCHK(hipMemcpy(ex.data(), j->d_cexp.p, ex.size(), hipMemcpyDeviceToHost));
  const double nb = std::max(1, j->desc.blockspersend);
  std::vector<float2> row;
  for (const XmacFreq& x : P.xf) {
    const int nin = (int)x.in_station.size();
    row.resize(x.nchan);
    for (int a = 0; a < nin; a++) {
      const int ds = x.in_station[a];
      if (x.in_total[a] >= j->dss[ds].nrecbands) continue;      // zoom inputs: their parent's own frequency
      const size_t gband = (size_t)P.st[ds].band0 + x.in_band[a];
      HIPCHK(hipMemcpy(row.data(), j->d_V.as<float2>() + x.voff + ((long long)a * nin + a) * x.nchan,
                       sizeof(float2) * x.nchan, hipMemcpyDeviceToHost));
      for (int c = 0; c < x.nchan; c++) {
        const double pw = row[c].x / nb;                          // mean |X|^2 per block
        if (!(pw > 0) || !std::isfinite(pw)) continue;
        const long e = std::lround(std::log2(16.0 / std::sqrt(pw)));
        ex[gband * P.C + c] = (signed char)std::max(-100L, std::min(100L, e));
      }
    }
  }
  HIPCHK(hipMemcpy(j->d_cexp.p, ex.data(), ex.size(), hipMemcpyHostToDevice));
  return 0;
}

int fxg_job_set_graph(fxg_job* j, int on) {
  if (!j) return set_err(FXG_EINVAL, "null job");
  j->use_graph = on ? 1 : 0;
  return 0;
}

int fxg_job_graph_captures(const fxg_job* j, int64_t* captures) {
  if (!j || !captures) return set_err(FXG_EINVAL, "null argument");
  *captures = j->graph_captures;
  return 0;
}

int fxg_job_set_kurtosis(fxg_job* j, int enable, int maxchannels, int nthreads) {
  if (!j) return set_err(FXG_EINVAL, "null argument");
  if (!enable) {
    j->kurt_on = 0;
    j->kurt_have = 0;
    j->d_kurt.release(); j->d_kurtout.release(); j->d_kurtnz.release();
    return 0;
  }
  if (maxchannels < 1) return set_err(FXG_EINVAL, "kurtosis channels must be positive");
  if (nthreads < 1 || nthreads > 1024) return set_err(FXG_EINVAL, "kurtosis records (threads) must be in [1, 1024]");
  HIPCHK(hipSetDevice(j->ctx->device));
  HIPCHK(hipStreamSynchronize(j->stream));
  const Plan& P = j->plan;
  const long long nb = (long long)P.bands.size();
  const int nout = std::min(maxchannels, P.C);
  int rc;
  if ((rc = j->d_kurt.ensure(sizeof(float) * 2 * (size_t)(nthreads * nb * P.C))) ||
      (rc = j->d_kurtout.ensure(sizeof(float) * (size_t)(nthreads * nb * nout))) ||
      (rc = j->d_kurtnz.ensure(sizeof(int) * (size_t)(nthreads * nb)))) {
    j->kurt_on = 0;
    return rc;
  }
  j->kurt_on = 1;
  j->kurt_maxch = maxchannels;
  j->kurt_nrec = nthreads;
  j->kurt_nout = nout;
  j->kurt_have = 0;
  return 0;
}

int fxg_job_get_kurtosis(fxg_job* j, int thread, int ds, int band, float* out, int maxout, int32_t* nchan,
                         int32_t* valid) {
  if (!j || !out || !nchan || !valid) return set_err(FXG_EINVAL, "null argument");
  if (!j->kurt_on) return set_err(FXG_ESTATE, "kurtosis is off (fxg_job_set_kurtosis)");
  if (!j->kurt_have) return set_err(FXG_ESTATE, "no subintegration processed since kurtosis was enabled");
  const Plan& P = j->plan;
  if (thread < 0 || thread >= j->kurt_nrec) return set_err(FXG_EINVAL, "thread");
  if (ds < 0 || ds >= P.nds) return set_err(FXG_EINVAL, "datastream");
  if (band < 0 || band >= P.st[ds].nchan) return set_err(FXG_EINVAL, "recorded band");
  if (maxout < j->kurt_nout) return set_err(FXG_EINVAL, "output too short");
  HIPCHK(hipSetDevice(j->ctx->device));
  HIPCHK(hipStreamSynchronize(j->stream));
  const long long nb = (long long)P.bands.size();
  const long long g = P.st[ds].band0 + band;
  const long long glast = P.st[ds].band0 + P.st[ds].nchan - 1;
  HIPCHK(hipMemcpy(out, j->d_kurtout.as<float>() + ((long long)thread * nb + g) * j->kurt_nout,
                   sizeof(float) * (size_t)j->kurt_nout, hipMemcpyDeviceToHost));
  int nz = 0;
  HIPCHK(hipMemcpy(&nz, j->d_kurtnz.as<int>() + (long long)thread * nb + glast, sizeof(int), hipMemcpyDeviceToHost));
  *nchan = j->kurt_nout;
  *valid = nz;
  return 0;
}

int fxg_job_set_keep_taps(fxg_job* j, int keep) {
  if (!j) return set_err(FXG_EINVAL, "null job");
  j->keep_taps = keep;
  return 0;
}

// XMAC tiling for nin inputs: CT channels per workgroup, 8x8 tiles
// Grid rows for an XMAC launch of wgx channel groups over nb blocks: enough
// workgroups to cover the chip when the channel groups alone do not, at
// least 8 blocks per row, within the partial buffer (cap elements).
static int xmac_splits(int wgx, int nb, long long vsz, long long cap) {
  if (wgx >= 256 || nb < 16 || vsz <= 0) return 1;
  long long n = std::min<long long>((512 + wgx - 1) / wgx, nb / 8);
  n = std::min<long long>(n, cap / vsz);
  return (int)std::max<long long>(1, n);
}

// (ct_req, tb_req: a schedule's choice, used where it fits; 0 = automatic)
static void xmac_shape(int nin, int nchan, XmacArgs* a, int ct_req = 0, int tb_req = 0) {
  a->G = (nin + 7) / 8;
  a->ntiles = a->G * (a->G + 1) / 2;
  int ct = 16;
  while (ct > 1 && ct * a->ntiles > 576) ct >>= 1;
  while (ct > 1 && ct > nchan) ct >>= 1;
  if (ct_req > 0 && (ct_req & (ct_req - 1)) == 0 && ct_req * a->ntiles <= 576 && ct_req <= std::max(1, nchan)) ct = ct_req;
  a->CT = ct;
  int per_block = a->G * 8 * ct * (int)sizeof(float2);
  int tb = std::min(std::max(1, 65536 / per_block), 16);
  if (tb_req > 0 && tb_req * per_block <= 65536) tb = tb_req;
  a->TB = tb;
}

// What the launches of one subintegration depend on besides the job's fixed
// buffers: a graph captured for one shape replays for every subintegration of
// the same shape.
struct LaunchShape {
  long long maxframes = 0, maxsamples = 0;
  int maxdatabytes = 0;
  int fastW = 0;        // 2-bit register corner turn: W words per sample-time, else 0
};

// Raw data (host inputs), the per-datastream subint table and the valid
// flags of one subintegration, copied to the device on stream s.  With
// raw_set non-null, host inputs go into the next of the two pipelined raw
// sets on the copy stream (s waits for them); *raw_set receives the set, or
// -1 when nothing was copied that way.
static int upload_subint(fxg_job* j, const fxg_subint_desc* sub, const fxg_station_input* st,
                         int data_on_device, hipStream_t s, LaunchShape* sh, int* raw_set = nullptr) {
  Plan& P = j->plan;
  int rc;
  int set = -1;
  if (raw_set) {
    *raw_set = -1;
    if (!data_on_device) {
      set = j->rawp_next;
      for (int k = 0; k < 2; k++) {
        if (!j->raw_free[k]) HIPCHK(hipEventCreateWithFlags(&j->raw_free[k], hipEventDisableTiming));
        if (!j->raw_copied[k]) HIPCHK(hipEventCreateWithFlags(&j->raw_copied[k], hipEventDisableTiming));
      }
      j->d_rawp[set].resize(P.nds);
      // the set's previous reader must have finished before it is refilled
      if (j->raw_free_rec[set]) HIPCHK(hipStreamWaitEvent(j->copy_stream, j->raw_free[set], 0));
    }
  }
  std::vector<DevStationSub> ss(P.nds);
  std::vector<int> flags;
  long long fbase = 0, maxframes = 0, ctbytes = 0, maxsamples = 0;
  int maxdatabytes = 0;
  int fastW = -1;
  for (int d = 0; d < P.nds; d++) {
    const fxg_station_input& in = st[d];
    DevStationSub& q = ss[d];
    if (in.nbytes < 0 || (in.nbytes > 0 && !in.data)) return set_err(FXG_EINVAL, "station data");
    const unsigned char* dp = in.data;
    if (!data_on_device && in.nbytes > 0 && set >= 0) {
      DBuf& b = j->d_rawp[set][d];
      // a larger buffer replaces the set's: its last reader must be done
      if ((size_t)in.nbytes > b.bytes && j->raw_free_rec[set]) HIPCHK(hipEventSynchronize(j->raw_free[set]));
      if ((rc = b.ensure((size_t)in.nbytes))) return rc;
      HIPCHK(hipMemcpyAsync(b.p, in.data, (size_t)in.nbytes, hipMemcpyHostToDevice, j->copy_stream));
      dp = b.as<unsigned char>();
    } else if (!data_on_device && in.nbytes > 0) {
      if ((rc = j->d_raw[d].ensure((size_t)in.nbytes))) return rc;
      if ((rc = j->ev_begin(FXG_STAGE_H2D))) return rc;
      HIPCHK(hipMemcpyAsync(j->d_raw[d].p, in.data, (size_t)in.nbytes, hipMemcpyHostToDevice, s));
      if ((rc = j->ev_end())) return rc;
      dp = j->d_raw[d].as<unsigned char>();
    }
    q.data = dp;
    q.nbytes = in.nbytes;
    q.nframes = in.nbytes / P.st[d].framebytes;
    q.fbase = fbase;
    fbase += q.nframes;
    q.ctbase = ctbytes;
    q.ctstride = ((q.nframes * P.st[d].bpf + 3) / 4) * 4;
    ctbytes += q.ctstride * P.st[d].nchan;
    maxdatabytes = std::max(maxdatabytes, P.st[d].databytes);
    maxframes = std::max(maxframes, q.nframes);
    {
      const DevStation& S = P.st[d];
      const int W = S.nbits == 2 && !S.cplx && S.nchan % 16 == 0 ? S.nchan / 16 : 0;
      const bool ok = (W == 1 || W == 2 || W == 4) &&
                      (((uintptr_t)q.data | (uintptr_t)S.framebytes | (uintptr_t)S.headerbytes |
                        (uintptr_t)S.databytes) % (4 * W)) == 0;
      if (q.nframes > 0) fastW = (ok && (fastW < 0 || fastW == W)) ? W : 0;
      if (W) maxsamples = std::max(maxsamples, q.nframes * (S.databytes / (4 * W)));
    }
    q.datasec = in.sec; q.datans = in.ns;
    q.datasamples = static_cast<int>(in.ns / (P.st[d].sampletime * 1e3) + 0.5);   // mode.cpp:1572
    // Mode::setOffsets / setData guards (mode.cpp:1525-1573, 671)
    q.usable = (sub->model_ok && in.sec > FXG_INVALID_SUBINT && in.scan == sub->scan && in.nbytes > 1) ? 1 : 0;
    q.nflags = in.validflags ? in.nvalidflags : 0;
    q.flagoff = (int)flags.size();
    for (int k = 0; k < q.nflags; k++) flags.push_back(in.validflags[k]);
    for (int k = 0; k < 3; k++) q.interp[k] = in.interpolator[k];
    q.offsec = sub->sec; q.offns = sub->ns; q.scanstart = sub->scan_start_sec; q.pad = 0;
  }
  if (set >= 0) {
    HIPCHK(hipEventRecord(j->raw_copied[set], j->copy_stream));
    HIPCHK(hipStreamWaitEvent(s, j->raw_copied[set], 0));
    j->rawp_next ^= 1;
    *raw_set = set;
  }
  if ((rc = j->d_ss.ensure(sizeof(DevStationSub) * ss.size())) ||
      (rc = j->d_flags.ensure(sizeof(int) * std::max<size_t>(1, flags.size()))) ||
      (rc = j->d_ftab.ensure(sizeof(int2) * (size_t)std::max<long long>(1, fbase))) ||
      (rc = j->d_ct.ensure((size_t)std::max<long long>(4, ctbytes))) ||
      (P.any_edv4 && (rc = j->d_fblank.ensure(sizeof(unsigned long long) * (size_t)std::max<long long>(1, fbase)))))
    return rc;
  for (int d = 0; d < P.nds; d++) {
    DevStationSub& q = ss[d];
    q.fblank = P.st[d].edv4 ? j->d_fblank.as<unsigned long long>() + q.fbase : nullptr;
    q.pbw = P.st[d].edv4 ? j->d_pbw.as<float>() + j->pbw_off[d] : nullptr;
  }
  // the tables go through the next pinned staging slot; its previous copy
  // must have run before the slot is refilled
  fxg_job::HostStage& hs = j->stage[j->stage_next];
  j->stage_next = (j->stage_next + 1) % fxg_job::kStages;
  if (!hs.done) HIPCHK(hipEventCreateWithFlags(&hs.done, hipEventDisableTiming));
  if (hs.queued) { HIPCHK(hipEventSynchronize(hs.done)); hs.queued = false; }
  if (hs.ss_cap < ss.size()) {
    if (hs.ss) { HIPCHK(hipHostFree(hs.ss)); hs.ss = nullptr; hs.ss_cap = 0; }
    if (hipHostMalloc((void**)&hs.ss, sizeof(DevStationSub) * ss.size()) != hipSuccess)
      return set_err(FXG_ENOMEM, "pinned staging of the station table");
    hs.ss_cap = ss.size();
  }
  if (hs.flags_cap < flags.size()) {
    if (hs.flags) { HIPCHK(hipHostFree(hs.flags)); hs.flags = nullptr; hs.flags_cap = 0; }
    if (hipHostMalloc((void**)&hs.flags, sizeof(int) * flags.size()) != hipSuccess)
      return set_err(FXG_ENOMEM, "pinned staging of the valid flags");
    hs.flags_cap = flags.size();
  }
  std::memcpy(hs.ss, ss.data(), sizeof(DevStationSub) * ss.size());
  if (!flags.empty()) std::memcpy(hs.flags, flags.data(), sizeof(int) * flags.size());
  HIPCHK(hipMemcpyAsync(j->d_ss.p, hs.ss, sizeof(DevStationSub) * ss.size(), hipMemcpyHostToDevice, s));
  if (!flags.empty()) HIPCHK(hipMemcpyAsync(j->d_flags.p, hs.flags, sizeof(int) * flags.size(), hipMemcpyHostToDevice, s));
  HIPCHK(hipEventRecord(hs.done, s));
  hs.queued = true;
  sh->maxframes = maxframes;
  sh->maxsamples = maxsamples;
  sh->maxdatabytes = maxdatabytes;
  sh->fastW = fastW > 0 ? fastW : 0;
  return 0;
}

// Frame validity, corner turn and FP64 block geometry (everything before the
// F stage), launched on stream s.
static int launch_front(fxg_job* j, const LaunchShape& sh, hipStream_t s) {
  Plan& P = j->plan;
  int rc;
  if ((rc = j->ev_begin(FXG_STAGE_FRAMES))) return rc;
  HIPCHK(launch_frames(j->d_st.as<DevStation>(), j->d_ss.as<DevStationSub>(), j->d_ftab.as<int2>(), P.nds,
                       sh.maxframes, s));
  HIPCHK(launch_cornerturn(j->d_st.as<DevStation>(), j->d_ss.as<DevStationSub>(), j->d_ct.as<unsigned char>(),
                           P.nds, sh.maxframes, sh.maxdatabytes, sh.fastW, sh.maxsamples, s));
  if ((rc = j->ev_end())) return rc;
  if ((rc = j->ev_begin(FXG_STAGE_PROLOGUE))) return rc;
  HIPCHK(launch_prologue(j->d_st.as<DevStation>(), j->d_ss.as<DevStationSub>(), j->d_flags.as<int>(),
                         j->d_ftab.as<int2>(), j->desc.blockspersend, j->d_blocks.as<DevBlock>(), P.nds, s));
  if ((rc = j->ev_end())) return rc;
  return 0;
}

static int prepare_subint(fxg_job* j, const fxg_subint_desc* sub, const fxg_station_input* st,
                          int data_on_device) {
  LaunchShape sh;
  int rc = upload_subint(j, sub, st, data_on_device, j->stream, &sh);
  if (rc) return rc;
  return launch_front(j, sh, j->stream);
}

static void fill_fstage_args(fxg_job* j, FStageArgs* fa) {
  *fa = FStageArgs{};
  fa->st = j->d_st.as<DevStation>();
  fa->bands = j->d_bands.as<DevBand>();
  fa->ss = j->d_ss.as<DevStationSub>();
  fa->blocks = j->d_blocks.as<DevBlock>();
  fa->ftab = j->d_ftab.as<int2>();
  fa->lut8 = j->ctx->lut8.as<float>();
  fa->tw = j->d_tw.as<float2>();
  fa->ct = j->d_ct.as<unsigned char>();
  fa->spectra = j->d_spec.p;
  fa->prec = j->desc.xmac_precision;
  fa->scale = j->scale;
  fa->nblocks_subint = j->desc.blockspersend;
  fa->zbuf = j->d_z.p ? j->d_z.as<float2>() : nullptr;
  fa->rtab = j->d_rtab.p ? j->d_rtab.as<float2>() : nullptr;
  fa->zcap = j->plan.zcap;
  fa->rtabn = j->plan.rtabn;
  fa->zcb0 = 0;
  fa->cexp = j->d_cexp.p ? j->d_cexp.as<signed char>() : nullptr;
  fa->gstride = j->gstride;
  fa->bstride = j->bstride;
}

// identity of a subintegration's inputs: the prepared state is reused while it matches
static std::vector<unsigned char> subint_key(const fxg_job* j, const fxg_subint_desc* sub,
                                             const fxg_station_input* st, int data_on_device) {
  std::vector<unsigned char> k;
  auto put = [&](const void* p, size_t n) { const unsigned char* b = (const unsigned char*)p; k.insert(k.end(), b, b + n); };
  put(sub, sizeof(*sub));
  put(&data_on_device, sizeof(int));
  for (int d = 0; d < j->plan.nds; d++) {
    put(&st[d].data, sizeof(st[d].data)); put(&st[d].nbytes, sizeof(st[d].nbytes));
    put(&st[d].scan, 3 * sizeof(int32_t)); put(st[d].interpolator, sizeof(st[d].interpolator));
    put(&st[d].nvalidflags, sizeof(int32_t));
    if (st[d].validflags) put(st[d].validflags, sizeof(int32_t) * (size_t)std::max(0, st[d].nvalidflags));
  }
  return k;
}

// F and X stages chunk by chunk, weights, the cleared result, scatter and
// phase cal, launched on stream s into dres (a device result buffer).
static int launch_back(fxg_job* j, float* dres, hipStream_t s) {
  Plan& P = j->plan;
  const fxg_job_desc& D = j->desc;
  int rc;
  const int nblocks = D.blockspersend;
  const int prec = D.xmac_precision;
  const int chunk = j->chunk_blocks > 0 ? std::min(j->chunk_blocks, P.chunk_blocks) : P.chunk_blocks;
  FStageArgs fa;
  fill_fstage_args(j, &fa);
  KurtArgs ka{};
  if (j->kurt_on) {
    ka.spectra = j->d_spec.p; ka.prec = prec; ka.invscale = 1.0f / j->scale;
    ka.cexp = j->d_cexp.p ? j->d_cexp.as<signed char>() : nullptr;
    ka.nbands = (int)P.bands.size(); ka.C = P.C; ka.nout = j->kurt_nout;
    ka.band_stride = (long long)P.chunk_blocks * P.C;
    ka.gstride = j->gstride; ka.bstride = j->bstride;
    ka.nblocks = nblocks; ka.nrec = j->kurt_nrec; ka.s = j->d_kurt.as<float>();
    // Mode::zeroKurtosis at the start of every subintegration (core.cpp:712-714)
    HIPCHK(launch_zero(ka.s, 2LL * ka.nrec * ka.nbands * ka.C, s));
  }
  for (int c0 = 0; c0 < nblocks; c0 += chunk) {
    int cn = std::min(chunk, nblocks - c0);
    fa.chunk0 = c0;
    if ((rc = j->ev_begin(FXG_STAGE_FSTAGE))) return rc;
    if (j->nb2) {
      fa.bandlist = j->d_bandlist.as<int>();
      HIPCHK(launch_fstage(P.L, 2, P.st[0].cplx, fa, j->nb2, cn, s));
    }
    if (j->nb8) {
      fa.bandlist = j->d_bandlist.as<int>() + j->nb2;
      HIPCHK(launch_fstage(P.L, 8, P.st[0].cplx, fa, j->nb8, cn, s));
    }
    if ((rc = j->ev_end())) return rc;
    if (j->kurt_on) {
      ka.chunk0 = c0; ka.nchunk = cn;
      if ((rc = j->ev_begin(FXG_STAGE_KURTOSIS))) return rc;
      HIPCHK(launch_kurt_accum(ka, s));
      if ((rc = j->ev_end())) return rc;
    }
    for (size_t x = 0; x < P.xf.size(); x++) {
      const XmacFreq& xf = P.xf[x];
      XmacArgs xa{};
      xa.spec = j->d_spec.p;
      xa.in_off = j->d_inoff.as<long long>() + j->inoff_base[x];
      xa.block_stride = j->bstride;
      xa.gstride = j->gstride;
      xa.nin = (int)xf.in_station.size();
      xa.nchan = xf.nchan;
      xa.nblocks = cn;
      xmac_shape(xa.nin, xa.nchan, &xa, j->xmac_ct, j->xmac_tb);
      const bool fp8 = prec == FXG_XMAC_FP8;
      // above 64 inputs only the streamed kernel runs on the matrix cores
      const bool big = xa.nin > 64;
      const bool mfma = prec != FXG_XMAC_FP32 &&
                        (big ? !fp8 && xmac_mfma_streaming_supported(xa.nin, xa.nchan) &&
                                   xmac_mfma_streaming_ok(xa, xf.in_off.data())
                             : xmac_mfma_supported(xa.nin, xa.nchan));
      // FP32 spectra: the FP32 matrix cores under the streamed schedule, the
      // VALU register tiles otherwise
      const bool mfma32 = prec == FXG_XMAC_FP32 && j->xmac_streaming && xmac_mfma_f32_ok(xa, xf.in_off.data());
      if (fp8) { xa.cexp = j->d_cexp.as<signed char>(); xa.eoff = j->d_eoff.as<int>() + j->inoff_base[x]; }
      float2* V = j->d_V.as<float2>() + xf.voff;
      // few channel groups: split the blocks over grid rows too, into partial
      // matrices summed afterwards in row order
      const long long vsz = (long long)xa.nin * xa.nin * xa.nchan;
      const int nsplit =
          xmac_splits(mfma || mfma32 ? xa.nchan / 16 : (xa.nchan + xa.CT - 1) / xa.CT, cn, vsz, P.vpart_elems);
      xa.split_blocks = (cn + nsplit - 1) / nsplit;
      const int rows = (cn + xa.split_blocks - 1) / xa.split_blocks;
      if (rows > 1) { xa.V = j->d_vpart.as<float2>(); xa.vsplit = vsz; xa.accumulate = 0; }
      else { xa.V = V; xa.vsplit = 0; xa.accumulate = c0 > 0; }
      if ((rc = j->ev_begin(FXG_STAGE_XMAC))) return rc;
      if (fp8) {
        HIPCHK(launch_xmac_fp8(xa, s));
      } else if (mfma32) {
        HIPCHK(launch_xmac_mfma_f32(xa, j->d_zero.p, s));
      } else if (mfma) {
        if ((j->xmac_streaming || big) && xmac_mfma_streaming_ok(xa, xf.in_off.data()))
          HIPCHK(launch_xmac_mfma_streaming(xa, prec, j->d_zero.p, s));
        else
          HIPCHK(launch_xmac_mfma(xa, prec, s));
      } else {
        HIPCHK(launch_xmac_f32(xa, prec, s));
      }
      if (rows > 1) HIPCHK(launch_xmac_reduce(j->d_vpart.as<float2>(), rows, vsz, V, vsz, c0 > 0, s));
      if ((rc = j->ev_end())) return rc;
    }
    j->tap_chunk0 = c0;
    j->tap_chunkn = cn;
  }
  if (j->kurt_on) {
    if ((rc = j->ev_begin(FXG_STAGE_KURTOSIS))) return rc;
    HIPCHK(launch_kurt_final(ka, j->d_kurtout.as<float>(), j->d_kurtnz.as<int>(), s));
    if ((rc = j->ev_end())) return rc;
  }
  if ((rc = j->ev_begin(FXG_STAGE_WEIGHTS))) return rc;
  HIPCHK(launch_weights(j->d_st.as<DevStation>(), j->d_ss.as<DevStationSub>(), j->d_blocks.as<DevBlock>(), nblocks,
                        j->d_wterms.as<DevWTerm>(), (int)P.wterms.size(), j->d_weights.as<float>(), s));
  if ((rc = j->ev_end())) return rc;
  // the slot is zeroed before accumulation (core.cpp:319) by a kernel, so
  // that the sequence can live in a graph
  HIPCHK(launch_zero(dres, 2 * P.info.result_length, s));
  float vscale = 1.0f / (j->scale * j->scale);
  if ((rc = j->ev_begin(FXG_STAGE_SCATTER))) return rc;
  HIPCHK(launch_scatter(j->d_segs.as<DevSeg>(), (int)P.segs.size(), P.maxseglen, j->d_V.as<float2>(),
                        j->d_weights.as<float>(), vscale, dres, s));
  if ((rc = j->ev_end())) return rc;
  if (!P.pcal.empty()) {
    PcalArgs pa{};
    pa.st = j->d_st.as<DevStation>(); pa.bands = j->d_bands.as<DevBand>(); pa.ss = j->d_ss.as<DevStationSub>();
    pa.blocks = j->d_blocks.as<DevBlock>(); pa.ftab = j->d_ftab.as<int2>(); pa.lut8 = j->ctx->lut8.as<float>();
    pa.ct = j->d_ct.as<unsigned char>(); pa.pc = j->d_pcal.as<DevPcal>(); pa.aux = j->d_pcalaux.as<float2>();
    pa.tonebin = j->d_pcaltone.as<int>(); pa.part = j->d_pcalpart.as<float>(); pa.res = dres;
    pa.nblocks = nblocks; pa.ngroups = P.pcal_groups;
    if ((rc = j->ev_begin(FXG_STAGE_PCAL))) return rc;
    HIPCHK(launch_pcal(pa, (int)P.pcal.size(), s));
    if ((rc = j->ev_end())) return rc;
  }
  return 0;
}

// identity of a captured launch sequence: the shape, the result buffer and
// the per-subint buffers the kernels take as arguments
static std::vector<unsigned char> graph_key(const fxg_job* j, const LaunchShape& sh, const float* dres) {
  std::vector<unsigned char> k;
  auto put = [&](const void* p, size_t n) { const unsigned char* b = (const unsigned char*)p; k.insert(k.end(), b, b + n); };
  put(&sh, sizeof(sh));
  put(&dres, sizeof(dres));
  const void* ptrs[4] = {j->d_ss.p, j->d_flags.p, j->d_ftab.p, j->d_ct.p};
  put(ptrs, sizeof(ptrs));
  const int sched[8] = {j->xmac_streaming, j->xmac_ct, j->xmac_tb, j->chunk_blocks,
                        j->kurt_on, j->kurt_maxch, j->kurt_nrec, 0};
  put(sched, sizeof(sched));
  const void* kptrs[3] = {j->d_kurt.p, j->d_kurtout.p, j->d_kurtnz.p};
  put(kptrs, sizeof(kptrs));
  return k;
}

// Record fn's launches into a graph on stream s (fn enqueues on s and on
// streams that fork from it), instantiate it into *exec.
static int capture(hipStream_t s, hipGraphExec_t* exec, const std::function<int()>& fn) {
  HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  int rc = fn();
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(s, &g);
  if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
  if (e != hipSuccess) return set_err(FXG_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) { *exec = nullptr; return set_err(FXG_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e)); }
  return 0;
}

// Everything of one subintegration, enqueued on the job's stream: eagerly,
// or as one graph launch replayed while the launch shape stays the same.
static int enqueue_subint(fxg_job* j, const fxg_subint_desc* sub, const fxg_station_input* st,
                          int data_on_device, float* results, int results_on_device) {
  if (!j || !sub || !st || !results) return set_err(FXG_EINVAL, "null argument");
  Plan& P = j->plan;
  hipStream_t s = j->stream;
  HIPCHK(hipSetDevice(j->ctx->device));
  int rc;
  j->prepared_key.clear();
  LaunchShape sh;
  // host inputs without stage profiling: pipelined copies (the profiled
  // path keeps them on the job's stream, where the H2D stage is timed)
  int raw_set = -1;
  if ((rc = upload_subint(j, sub, st, data_on_device, s, &sh, j->profiling ? nullptr : &raw_set))) return rc;
  float* dres = results_on_device ? results : j->d_results.as<float>();
  if (j->use_graph && !j->profiling) {
    std::vector<unsigned char> key = graph_key(j, sh, dres);
    if (!j->graph_exec || key != j->graph_key) {
      if (j->graph_exec) { (void)hipGraphExecDestroy(j->graph_exec); j->graph_exec = nullptr; }
      j->graph_key.clear();
      rc = capture(s, &j->graph_exec, [&]() {
        int r = launch_front(j, sh, s);
        return r ? r : launch_back(j, dres, s);
      });
      if (rc) return rc;
      j->graph_key = key;
      j->graph_captures++;
    }
    HIPCHK(hipGraphLaunch(j->graph_exec, s));
  } else {
    if ((rc = launch_front(j, sh, s))) return rc;
    if ((rc = launch_back(j, dres, s))) return rc;
  }
  if (raw_set >= 0) {
    HIPCHK(hipEventRecord(j->raw_free[raw_set], s));
    j->raw_free_rec[raw_set] = true;
  }
  if (!results_on_device) {
    const size_t resbytes = sizeof(float) * 2 * (size_t)P.info.result_length;
    if ((rc = j->ev_begin(FXG_STAGE_D2H))) return rc;
    HIPCHK(hipMemcpyAsync(results, dres, resbytes, hipMemcpyDeviceToHost, s));
    if ((rc = j->ev_end())) return rc;
  }
  j->have_subint = 1;
  j->kurt_have = j->kurt_on;
  return 0;
}

int fxg_job_set_schedule(fxg_job* j, const fxg_schedule* sc) {
  if (!j || !sc) return set_err(FXG_EINVAL, "null argument");
  if (sc->xmac_ct < 0 || sc->xmac_tb < 0 || sc->chunk_blocks < 0) return set_err(FXG_EINVAL, "negative schedule entry");
  j->xmac_streaming = sc->xmac_streaming ? 1 : 0;
  j->xmac_ct = sc->xmac_ct;
  j->xmac_tb = sc->xmac_tb;
  j->chunk_blocks = sc->chunk_blocks;
  return 0;
}

int fxg_job_get_schedule(const fxg_job* j, fxg_schedule* sc) {
  if (!j || !sc) return set_err(FXG_EINVAL, "null argument");
  sc->xmac_streaming = j->xmac_streaming;
  sc->xmac_ct = j->xmac_ct;
  sc->xmac_tb = j->xmac_tb;
  sc->chunk_blocks = j->chunk_blocks;
  return 0;
}

// Candidate schedules one dimension at a time around `base`: the XMAC
// staging or tile first, then the pass length.
static std::vector<fxg_schedule> xmac_candidates(const fxg_job* j, const fxg_schedule& base) {
  const Plan& P = j->plan;
  std::vector<fxg_schedule> out;
  bool mfma = false, f32 = false;
  for (const XmacFreq& xf : P.xf) {
    const int nin = (int)xf.in_station.size();
    const bool reduced = j->desc.xmac_precision != FXG_XMAC_FP32;
    if (reduced && xmac_mfma_supported(nin, xf.nchan)) mfma = true;
    else if (!(reduced && nin > 64 && xmac_mfma_streaming_supported(nin, xf.nchan))) f32 = true;   // (streamed only)
  }
  if (mfma) {
    for (int st = 0; st < 2; st++) { fxg_schedule c = base; c.xmac_streaming = st; out.push_back(c); }
  }
  if (f32) {
    // the FP32 matrix cores (streamed schedule), then the VALU register tiles
    if (j->desc.xmac_precision == FXG_XMAC_FP32) { fxg_schedule c = base; c.xmac_streaming = 1; out.push_back(c); }
    for (int ct : {0, 4, 8, 16})
      for (int tb : {0, 4, 8, 16}) {
        if ((ct == 0) != (tb == 0)) continue;
        fxg_schedule c = base; c.xmac_streaming = 0; c.xmac_ct = ct; c.xmac_tb = tb; out.push_back(c);
      }
  }
  if (out.empty()) out.push_back(base);
  return out;
}

static std::vector<fxg_schedule> chunk_candidates(const fxg_job* j, const fxg_schedule& base) {
  std::vector<fxg_schedule> out;
  const int full = j->plan.chunk_blocks;
  for (int cb : {0, 1024, 256, 64, 16}) {
    if (cb > 0 && cb >= full) continue;
    fxg_schedule c = base; c.chunk_blocks = cb; out.push_back(c);
  }
  return out;
}

int fxg_job_autotune(fxg_job* j, const fxg_subint_desc* sub, const fxg_station_input* st, int data_on_device,
                     int reps, fxg_tune_result* results, int max_results, int* nresults) {
  if (!j || !sub || !st) return set_err(FXG_EINVAL, "null argument");
  if (reps < 1) reps = 1;
  HIPCHK(hipSetDevice(j->ctx->device));
  hipStream_t s = j->stream;
  int rc;
  const int saved_prof = j->profiling;
  j->profiling = 0;
  j->prepared_key.clear();
  LaunchShape sh;
  if ((rc = upload_subint(j, sub, st, data_on_device, s, &sh))) { j->profiling = saved_prof; return rc; }
  const size_t nres = 2 * (size_t)j->plan.info.result_length;
  std::vector<float> ref, cur(nres);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return set_err(FXG_EHIP, "hipEventCreate"); }
  std::vector<fxg_tune_result> tried;
  fxg_schedule best{};
  fxg_job_get_schedule(j, &best);
  double best_ms = 1e300;
  auto run = [&](const fxg_schedule& c, fxg_tune_result* r) -> int {
    fxg_job_set_schedule(j, &c);
    float* dres = j->d_results.as<float>();
    int q;
    if ((q = launch_front(j, sh, s)) || (q = launch_back(j, dres, s))) return q;     // untimed
    HIPCHK(hipEventRecord(e0, s));
    for (int k = 0; k < reps; k++)
      if ((q = launch_front(j, sh, s)) || (q = launch_back(j, dres, s))) return q;
    HIPCHK(hipEventRecord(e1, s));
    HIPCHK(hipMemcpyAsync(cur.data(), dres, sizeof(float) * nres, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    r->schedule = c;
    r->ms = ms / reps;
    if (ref.empty()) ref = cur;
    double num = 0, den = 0;
    for (size_t i = 0; i < nres; i++) { double d = (double)cur[i] - ref[i]; num += d * d; den += (double)ref[i] * ref[i]; }
    r->rel_err = den > 0 ? std::sqrt(num / den) : std::sqrt(num);
    r->accepted = r->rel_err <= 1e-5;
    r->pad = 0;
    return 0;
  };
  for (int phase = 0; phase < 2 && !rc; phase++) {
    std::vector<fxg_schedule> cand = phase == 0 ? xmac_candidates(j, best) : chunk_candidates(j, best);
    for (const fxg_schedule& c : cand) {
      fxg_tune_result r{};
      if ((rc = run(c, &r))) break;
      tried.push_back(r);
      if (r.accepted && r.ms < best_ms) { best_ms = r.ms; best = c; }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  j->profiling = saved_prof;
  fxg_job_set_schedule(j, &best);
  if (rc) return rc;
  if (nresults) *nresults = (int)tried.size();
  for (int i = 0; results && i < (int)tried.size() && i < max_results; i++) results[i] = tried[i];
  return 0;
}

int fxg_subint_process(fxg_job* j, const fxg_subint_desc* sub, const fxg_station_input* st,
                       int data_on_device, float* results, int results_on_device) {
  int rc = enqueue_subint(j, sub, st, data_on_device, results, results_on_device);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(j->stream));
  j->ev_collect();
  return 0;
}

int fxg_subint_submit(fxg_job* j, const fxg_subint_desc* sub, const fxg_station_input* st,
                      int data_on_device, float* results, int results_on_device, int64_t* ticket) {
  if (!ticket) return set_err(FXG_EINVAL, "null ticket");
  int rc = enqueue_subint(j, sub, st, data_on_device, results, results_on_device);
  if (rc) return rc;
  hipEvent_t e;
  if (!j->free_events.empty()) {
    e = j->free_events.back();
    j->free_events.pop_back();
  } else {
    HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  HIPCHK(hipEventRecord(e, j->stream));
  *ticket = j->next_ticket++;
  j->pending.push_back({*ticket, e});
  return 0;
}

int fxg_subint_query(fxg_job* j, int64_t ticket, int* done) {
  if (!j || !done) return set_err(FXG_EINVAL, "null argument");
  if (ticket <= 0 || ticket >= j->next_ticket) return set_err(FXG_EINVAL, "unknown ticket");
  *done = 1;
  for (auto& p : j->pending)
    if (p.first == ticket) {
      hipError_t q = hipEventQuery(p.second);
      if (q == hipErrorNotReady) *done = 0;
      else if (q != hipSuccess) return set_err(FXG_EHIP, std::string("hipEventQuery: ") + hipGetErrorString(q));
      break;
    }
  return 0;
}

int fxg_subint_wait(fxg_job* j, int64_t ticket) {
  if (!j) return set_err(FXG_EINVAL, "null job");
  if (ticket <= 0 || ticket >= j->next_ticket) return set_err(FXG_EINVAL, "unknown ticket");
  HIPCHK(hipSetDevice(j->ctx->device));
  // tickets complete in submission order on the job's stream
  while (!j->pending.empty() && j->pending.front().first <= ticket) {
    HIPCHK(hipEventSynchronize(j->pending.front().second));
    j->free_events.push_back(j->pending.front().second);
    j->pending.pop_front();
  }
  j->ev_collect();
  return 0;
}

// ---------------------------------------------------------------------------
// batches: one subintegration of each of several jobs in one graph launch;
// every job's sequence is a branch forked from the batch stream
// ---------------------------------------------------------------------------
struct fxg_batch {
  fxg_ctx* ctx = nullptr;
  std::vector<fxg_job*> jobs;
  hipStream_t stream = nullptr;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> join;
  hipGraphExec_t exec = nullptr;
  std::vector<unsigned char> key;
  long long captures = 0;
};

static void batch_free(fxg_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  for (hipEvent_t e : b->join) if (e) (void)hipEventDestroy(e);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int fxg_batch_create(fxg_ctx* c, fxg_job* const* jobs, int njobs, fxg_batch** out) {
  if (!c || !jobs || njobs < 1 || !out) return set_err(FXG_EINVAL, "null argument");
  for (int k = 0; k < njobs; k++) {
    if (!jobs[k] || jobs[k]->ctx != c) return set_err(FXG_EINVAL, "every job must belong to the context");
    for (int m = 0; m < k; m++)
      if (jobs[m] == jobs[k]) return set_err(FXG_EINVAL, "a job appears twice in the batch");
  }
  HIPCHK(hipSetDevice(c->device));
  fxg_batch* b = new fxg_batch();
  b->ctx = c;
  b->jobs.assign(jobs, jobs + njobs);
  b->join.assign(njobs, nullptr);
  bool ok = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) == hipSuccess;
  for (int k = 0; ok && k < njobs; k++) ok = hipEventCreateWithFlags(&b->join[k], hipEventDisableTiming) == hipSuccess;
  if (!ok) { batch_free(b); return set_err(FXG_EHIP, "batch stream / events"); }
  *out = b;
  return 0;
}

int fxg_batch_destroy(fxg_batch* b) {
  if (!b) return 0;
  (void)hipSetDevice(b->ctx->device);
  (void)hipStreamSynchronize(b->stream);
  batch_free(b);
  return 0;
}

int fxg_batch_graph_captures(const fxg_batch* b, int64_t* captures) {
  if (!b || !captures) return set_err(FXG_EINVAL, "null argument");
  *captures = b->captures;
  return 0;
}

int fxg_batch_process(fxg_batch* b, const fxg_subint_desc* subs, const fxg_station_input* const* stations,
                      int data_on_device, float* const* results, int results_on_device) {
  if (!b || !subs || !stations || !results) return set_err(FXG_EINVAL, "null argument");
  const int n = (int)b->jobs.size();
  for (int k = 0; k < n; k++) {
    if (!stations[k] || !results[k]) return set_err(FXG_EINVAL, "null station or result array");
    if (b->jobs[k]->profiling) return set_err(FXG_ESTATE, "stage profiling is not available inside a batch");
  }
  HIPCHK(hipSetDevice(b->ctx->device));
  hipStream_t s = b->stream;
  int rc;
  // work already queued on the jobs' own streams comes first
  for (int k = 0; k < n; k++) {
    HIPCHK(hipEventRecord(b->join[k], b->jobs[k]->stream));
    HIPCHK(hipStreamWaitEvent(s, b->join[k], 0));
  }
  std::vector<LaunchShape> sh(n);
  std::vector<float*> dres(n);
  std::vector<unsigned char> key;
  for (int k = 0; k < n; k++) {
    fxg_job* j = b->jobs[k];
    j->prepared_key.clear();
    if ((rc = upload_subint(j, &subs[k], stations[k], data_on_device, s, &sh[k]))) return rc;
    dres[k] = results_on_device ? results[k] : j->d_results.as<float>();
    std::vector<unsigned char> kk = graph_key(j, sh[k], dres[k]);
    key.insert(key.end(), kk.begin(), kk.end());
  }
  if (!b->exec || key != b->key) {
    if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
    b->key.clear();
    rc = capture(s, &b->exec, [&]() -> int {
      HIPCHK(hipEventRecord(b->fork, s));
      for (int k = 0; k < n; k++) {
        fxg_job* j = b->jobs[k];
        HIPCHK(hipStreamWaitEvent(j->stream, b->fork, 0));
        int r = launch_front(j, sh[k], j->stream);
        if (!r) r = launch_back(j, dres[k], j->stream);
        // the branch is joined even after a failed launch, so the capture ends cleanly
        HIPCHK(hipEventRecord(b->join[k], j->stream));
        HIPCHK(hipStreamWaitEvent(s, b->join[k], 0));
        if (r) return r;
      }
      return 0;
    });
    if (rc) return rc;
    b->key = key;
    b->captures++;
  }
  HIPCHK(hipGraphLaunch(b->exec, s));
  for (int k = 0; k < n; k++) {
    fxg_job* j = b->jobs[k];
    if (!results_on_device)
      HIPCHK(hipMemcpyAsync(results[k], dres[k], sizeof(float) * 2 * (size_t)j->plan.info.result_length,
                            hipMemcpyDeviceToHost, s));
  }
  HIPCHK(hipStreamSynchronize(s));
  for (fxg_job* j : b->jobs) { j->have_subint = 1; j->kurt_have = j->kurt_on; }
  return 0;
}

// digests of frames [f0, f0 + nf) of one datastream's input as the caller
// holds it now (host bytes, or device bytes through k_frame_digest)
static int frame_digests(fxg_job* j, const fxg_station_input& in, int data_on_device, int framebytes,
                         long long f0, long long nf, unsigned long long* out) {
  if (nf <= 0) return 0;
  if (!data_on_device) {
    for (long long f = 0; f < nf; f++) out[f] = host_frame_digest(in.data + (f0 + f) * framebytes, framebytes);
    return 0;
  }
  int rc;
  if ((rc = j->d_digest.ensure(sizeof(unsigned long long) * (size_t)nf))) return rc;
  HIPCHK(launch_frame_digest(in.data, f0, nf, framebytes, j->d_digest.as<unsigned long long>(), j->stream));
  HIPCHK(hipMemcpyAsync(out, j->d_digest.p, sizeof(unsigned long long) * (size_t)nf, hipMemcpyDeviceToHost,
                        j->stream));
  HIPCHK(hipStreamSynchronize(j->stream));
  return 0;
}

// Whether the frames that block `block` of datastream ds reads (one frame of
// margin either side of its sample window) still hold the bytes they held
// when the subintegration was prepared.
static int block_frames_unchanged(fxg_job* j, const fxg_station_input* st, int data_on_device, int ds, int block,
                                  bool* same) {
  const Plan& P = j->plan;
  const DevStation& S = P.st[ds];
  const DevBlock& g = j->prepared_blocks[(size_t)ds * j->desc.blockspersend + block];
  const long long nfr = st[ds].nbytes / S.framebytes;
  const long long t0 = g.nearestsample;
  const long long fa = std::max(0LL, (t0 >= 0 ? t0 / S.framesamples : -1) - 1);
  const long long fb = std::min(nfr - 1, (t0 + S.L) / std::max(1, S.framesamples) + 1);
  *same = true;
  if (fb < fa) return 0;
  std::vector<unsigned long long> now((size_t)(fb - fa + 1));
  int rc;
  if ((rc = frame_digests(j, st[ds], data_on_device, S.framebytes, fa, fb - fa + 1, now.data()))) return rc;
  for (long long f = fa; f <= fb; f++)
    if (now[(size_t)(f - fa)] != j->prepared_digest[(size_t)(j->prepared_fbase[ds] + f)]) { *same = false; break; }
  return 0;
}

int fxg_station_process_block(fxg_job* j, const fxg_subint_desc* sub, const fxg_station_input* st,
                              int data_on_device, int ds, int block, float* spectra, fxg_block_info* info) {
  if (!j || !sub || !st || !spectra) return set_err(FXG_EINVAL, "null argument");
  const Plan& P = j->plan;
  if (ds < 0 || ds >= P.nds || block < 0 || block >= j->desc.blockspersend) return set_err(FXG_EINVAL, "index");
  HIPCHK(hipSetDevice(j->ctx->device));
  int rc;
  std::vector<unsigned char> key = subint_key(j, sub, st, data_on_device);
  bool same = key == j->prepared_key;
  // the same description can hold new bytes (a refilled buffer): compare
  // the digests of the frames this block reads with their prepared ones
  if (same && (rc = block_frames_unchanged(j, st, data_on_device, ds, block, &same))) return rc;
  if (!same) {
    j->prepared_key.clear();
    if ((rc = prepare_subint(j, sub, st, data_on_device))) return rc;
    const size_t nblk = (size_t)P.nds * j->desc.blockspersend;
    j->prepared_blocks.resize(nblk);
    HIPCHK(hipMemcpyAsync(j->prepared_blocks.data(), j->d_blocks.p, sizeof(DevBlock) * nblk, hipMemcpyDeviceToHost,
                          j->stream));
    HIPCHK(hipStreamSynchronize(j->stream));
    j->prepared_fbase.assign(P.nds, 0);
    long long nf = 0;
    for (int d = 0; d < P.nds; d++) { j->prepared_fbase[d] = nf; nf += st[d].nbytes / P.st[d].framebytes; }
    j->prepared_digest.assign((size_t)nf, 0ull);
    for (int d = 0; d < P.nds; d++)
      if ((rc = frame_digests(j, st[d], data_on_device, P.st[d].framebytes, 0, st[d].nbytes / P.st[d].framebytes,
                              j->prepared_digest.data() + j->prepared_fbase[d])))
        return rc;
    j->prepared_key = key;
  }
  FStageArgs fa;
  fill_fstage_args(j, &fa);
  fa.chunk0 = block;
  fa.bandlist = j->d_iota.as<int>() + P.st[ds].band0;
  HIPCHK(launch_fstage(P.L, P.st[ds].nbits, P.st[ds].cplx, fa, P.st[ds].nchan, 1, j->stream));
  HIPCHK(hipStreamSynchronize(j->stream));
  j->ev_collect();
  j->have_subint = 1;
  j->tap_chunk0 = block;
  j->tap_chunkn = 1;
  for (int b = 0; b < P.st[ds].nchan; b++)
    if ((rc = fxg_tap_spectrum(j, ds, b, block, spectra + 2 * (size_t)b * P.C))) return rc;
  if (info) if ((rc = fxg_tap_block_info(j, ds, block, info))) return rc;
  return 0;
}

int fxg_job_set_profiling(fxg_job* j, int on) {
  if (!j) return set_err(FXG_EINVAL, "null job");
  j->profiling = on ? 1 : 0;
  return 0;
}

int fxg_job_stage_times(fxg_job* j, double* ms, int64_t* launches, int reset) {
  if (!j) return set_err(FXG_EINVAL, "null job");
  for (int i = 0; i < FXG_NSTAGES; i++) {
    if (ms) ms[i] = j->stage_ms[i];
    if (launches) launches[i] = j->stage_n[i];
    if (reset) { j->stage_ms[i] = 0; j->stage_n[i] = 0; }
  }
  return 0;
}

int fxg_tap_block_info(fxg_job* j, int ds, int block, fxg_block_info* out) {
  if (!j || !out) return set_err(FXG_EINVAL, "null argument");
  if (!j->have_subint) return set_err(FXG_ESTATE, "no subintegration processed");
  if (ds < 0 || ds >= j->plan.nds || block < 0 || block >= j->desc.blockspersend) return set_err(FXG_EINVAL, "index");
  HIPCHK(hipSetDevice(j->ctx->device));
  DevBlock g;
  HIPCHK(hipMemcpy(&g, j->d_blocks.as<DevBlock>() + (size_t)ds * j->desc.blockspersend + block, sizeof(g),
                   hipMemcpyDeviceToHost));
  std::memset(out, 0, sizeof(*out));
  out->status = g.status; out->nearestsample = g.nearestsample; out->integerdelay = g.integerdelay;
  out->weight = g.status == 0 ? g.weight : 0.0f; out->fracsampleerror = g.fse; out->intwalltime = g.intwall;
  out->a = g.a; out->b = g.b; out->fracwalltime = g.fracwall; out->c = g.c;
  return 0;
}

int fxg_tap_band_weights(fxg_job* j, int ds, int block, float* out) {
  if (!j || !out) return set_err(FXG_EINVAL, "null argument");
  if (!j->have_subint) return set_err(FXG_ESTATE, "no subintegration processed");
  const Plan& P = j->plan;
  if (ds < 0 || ds >= P.nds || block < 0 || block >= j->desc.blockspersend) return set_err(FXG_EINVAL, "index");
  HIPCHK(hipSetDevice(j->ctx->device));
  const int n = j->dss[ds].nrecbands;
  if (P.st[ds].edv4) {
    HIPCHK(hipMemcpy(out, j->d_pbw.as<float>() + j->pbw_off[ds] + (long long)block * n, sizeof(float) * n,
                     hipMemcpyDeviceToHost));
    DevBlock g;
    HIPCHK(hipMemcpy(&g, j->d_blocks.as<DevBlock>() + (size_t)ds * j->desc.blockspersend + block, sizeof(g),
                     hipMemcpyDeviceToHost));
    if (g.status != 0) std::fill(out, out + n, 0.0f);
    return 0;
  }
  fxg_block_info bi;
  int rc = fxg_tap_block_info(j, ds, block, &bi);
  if (rc) return rc;
  std::fill(out, out + n, bi.status == 0 ? bi.weight : 0.0f);
  return 0;
}

int fxg_tap_spectrum(fxg_job* j, int ds, int band, int block, float* out) {
  if (!j || !out) return set_err(FXG_EINVAL, "null argument");
  if (!j->have_subint) return set_err(FXG_ESTATE, "no subintegration processed");
  const Plan& P = j->plan;
  if (ds < 0 || ds >= P.nds || band < 0 || band >= j->dss[ds].nrecbands) return set_err(FXG_EINVAL, "index");
  if (block < j->tap_chunk0 || block >= j->tap_chunk0 + j->tap_chunkn)
    return set_err(FXG_ESTATE, "block is not in the resident chunk (the last chunk of the subintegration)");
  HIPCHK(hipSetDevice(j->ctx->device));
  const DevBand& B = P.bands[P.st[ds].band0 + band];
  long long off = B.specoff + (long long)(block - j->tap_chunk0) * j->bstride;
  // the C channels of the spectrum in channel order (gathered from the
  // 16-channel groups of the grouped layout)
  auto fetch = [&](void* dst, size_t es) -> int {
    const char* src = (const char*)j->d_spec.p + off * (long long)es;
    if (j->gstride == 16) HIPCHK(hipMemcpy(dst, src, es * (size_t)P.C, hipMemcpyDeviceToHost));
    else HIPCHK(hipMemcpy2D(dst, 16 * es, src, (size_t)j->gstride * es, 16 * es, (size_t)P.C / 16, hipMemcpyDeviceToHost));
    return 0;
  };
  int rc;
  if (j->desc.xmac_precision == FXG_XMAC_FP32) {
    if ((rc = fetch(out, sizeof(float2)))) return rc;
  } else if (j->desc.xmac_precision == FXG_XMAC_FP8) {
    std::vector<uint8_t> h(2 * (size_t)P.C);
    std::vector<signed char> ex((size_t)P.C);
    if ((rc = fetch(h.data(), 2))) return rc;
    HIPCHK(hipMemcpy(ex.data(), j->d_cexp.as<signed char>() + (size_t)(P.st[ds].band0 + band) * P.C, (size_t)P.C,
                     hipMemcpyDeviceToHost));
    for (size_t i = 0; i < h.size(); i++) {
      // e4m3: sign, 4 exponent bits (bias 7), 3 mantissa bits; 0x7f/0xff NaN
      const int sign = h[i] >> 7, e = (h[i] >> 3) & 15, m = h[i] & 7;
      float v = e == 0 ? std::ldexp((float)m, -9) : std::ldexp((float)(8 + m), e - 10);
      if (e == 15 && m == 7) v = std::nanf("");
      out[i] = std::ldexp(sign ? -v : v, -ex[i / 2]);
    }
  } else {
    std::vector<uint16_t> h(2 * (size_t)P.C);
    if ((rc = fetch(h.data(), 4))) return rc;
    for (size_t i = 0; i < h.size(); i++) {
      float v;
      if (j->desc.xmac_precision == FXG_XMAC_BF16) {
        uint32_t u = (uint32_t)h[i] << 16; std::memcpy(&v, &u, 4);
      } else {
        // IEEE half -> float
        uint32_t hv = h[i], sign = (hv >> 15) & 1, e = (hv >> 10) & 0x1f, m = hv & 0x3ff, u;
        if (e == 0) { v = std::ldexp((float)m, -24); if (sign) v = -v; out[i] = v / j->scale; continue; }
        if (e == 31) u = (sign << 31) | 0x7f800000u | (m << 13);
        else u = (sign << 31) | ((e - 15 + 127) << 23) | (m << 13);
        std::memcpy(&v, &u, 4);
      }
      out[i] = v / j->scale;
    }
  }
  return 0;
}

}  // extern "C"
