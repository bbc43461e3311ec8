// gpumode.cpp — Mode / Mk5Mode / Core / Model interfaces over libfxcorr_gpu.
#include "gpumode.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace fxgpu {

// ---------------------------------------------------------------------------
// DelayModel: model.cpp:229-304
// ---------------------------------------------------------------------------
f64 DelayModel::eval(const std::vector<f64>& c, f64 deltat) {
  // powers of deltat by repeated multiplication, dot product in index order
  f64 tp = 1.0, out = 0.0;
  for (size_t i = 0; i < c.size(); i++) {
    out += tp * c[i];
    tp *= deltat;
  }
  return out;
}

bool DelayModel::calculateDelayInterpolator(f64 offsettime, f64 timespan, int numincrements, int order,
                                            f64* delaycoeffs) const {
  if (order < 0 || order > 2 || polys_.empty() || !(inc_ > 0)) return false;
  const f64 t = offsettime + polyoffset_;
  int k = (int)(t / inc_);
  const int n = (int)polys_.size();
  if (k == n && (t - k * inc_) < 1e-6) k--;   // exactly at the end of the last interval
  if (k < 0 || k >= n) return false;
  const std::vector<f64>& c = polys_[k];
  const f64 mid = eval(c, t + timespan / 2.0 - k * inc_);
  if (order == 0) {
    delaycoeffs[0] = mid;
    return true;
  }
  const f64 s0 = eval(c, t - k * inc_);
  const f64 s2 = eval(c, t + timespan - k * inc_);
  if (order == 1) {
    delaycoeffs[0] = (s2 - s0) / numincrements;
    delaycoeffs[1] = s0 + (mid - (delaycoeffs[0] * numincrements / 2.0 + s0)) / 3.0;
    return true;
  }
  delaycoeffs[0] = (2.0 * s0 - 4.0 * mid + 2.0 * s2) / ((f64)numincrements * numincrements);
  delaycoeffs[1] = (-3.0 * s0 + 4.0 * mid - s2) / numincrements;
  delaycoeffs[2] = s0;
  return true;
}

// ---------------------------------------------------------------------------
// GpuMode
// ---------------------------------------------------------------------------
GpuMode::GpuMode(GpuCore* core, int dsindex, int numBufferedFFTs)
    : core_(core), ds_(dsindex), nbuf_(std::max(1, numBufferedFFTs)) {
  const fxg_datastream& d = core->dss_[dsindex];
  nrec_ = d.nrecbands;
  nzoom_ = d.nzoombands;
  nchan_ = core->freqs_[d.recfreq_index[0]].nchan;
  chanavg_ = std::max(1, core->freqs_[d.recfreq_index[0]].chans_to_avg);
  freqs_.assign(nbuf_, std::vector<cf32>((size_t)nrec_ * nchan_, cf32{0, 0}));
  conj_ = freqs_;
  dataweight_.assign(nbuf_, 0.0f);
  bandweight_.assign(nbuf_, std::vector<f32>((size_t)nrec_, 0.0f));
  for (int p = 0; p < 2; p++) {
    ac_[p].assign((size_t)nrec_ * nchan_, cf32{0, 0});
    acw_[p].assign(nrec_, 0.0f);
  }
  in_.sec = FXG_INVALID_SUBINT;
  flags_.assign((core->desc_.blockspersend + FXG_FLAGS_PER_INT - 1) / FXG_FLAGS_PER_INT, 0);
}

void GpuMode::setData(const u8* d, int dbytes, int dscan, int dsec, int dns) {
  in_.data = d;
  in_.nbytes = dbytes;
  in_.scan = dscan;
  in_.sec = dsec;
  in_.ns = dns;
}

void GpuMode::setValidFlags(const s32* v) { std::copy(v, v + flags_.size(), flags_.begin()); }

void GpuMode::setDelayModel(const DelayModel* model, int intclockseconds) {
  model_ = model;
  intclock_ = intclockseconds;
}

void GpuMode::setOffsets(int scan, int seconds, int ns) {
  offscan_ = scan;
  offsec_ = seconds;
  offns_ = ns;
  modelok_ = true;
  if (in_.sec <= FXG_INVALID_SUBINT) return;   // no valid data: the subint is ignored
  if (scan != in_.scan) {                      // data from another scan: torch it
    in_.nbytes = 0;
    return;
  }
  const fxg_datastream& d = core_->dss_[ds_];
  const fxg_freq& f = core_->freqs_[d.recfreq_index[0]];
  const f64 sampletime_us = 1.0 / (2.0 * f.bw_mhz);
  const int bps = core_->desc_.blockspersend;
  const f64 timespan = bps * 2.0 * f.nchan * sampletime_us / 1e6;
  f64 c[3] = {0, 0, 0};
  modelok_ = model_ != nullptr &&
             model_->calculateDelayInterpolator((f64)seconds + (f64)ns / 1e9, timespan, bps, 2, c);
  c[2] -= 1000000.0 * intclock_;
  for (int k = 0; k < 3; k++) in_.interpolator[k] = c[k];
}

void GpuMode::setOffsets(int scan, int seconds, int ns, const f64 interpolator[3], bool foundok) {
  offscan_ = scan;
  offsec_ = seconds;
  offns_ = ns;
  for (int k = 0; k < 3; k++) in_.interpolator[k] = interpolator[k];
  modelok_ = foundok;
}

void GpuMode::zeroAutocorrelations() {
  for (int p = 0; p < 2; p++) {
    std::fill(ac_[p].begin(), ac_[p].end(), cf32{0, 0});
    std::fill(acw_[p].begin(), acw_[p].end(), 0.0f);
  }
}

void GpuMode::averageFrequency() {
  if (chanavg_ == 1) return;
  const int outputchans = nchan_ / chanavg_;
  for (int p = 0; p < 2; p++)
    for (int j = 0; j < nrec_; j++) {
      cf32* a = &ac_[p][(size_t)j * nchan_];
      for (int k = 0; k < outputchans; k++) {
        cf32 m{0, 0};                                // genericMean_32fc
        for (int i = 0; i < chanavg_; i++) { m.re += a[k * chanavg_ + i].re; m.im += a[k * chanavg_ + i].im; }
        m.re /= chanavg_;
        m.im /= chanavg_;
        a[k] = m;
      }
    }
}

// acc += a * conj(b) in the operation order of vectorAddProduct_cf32 applied
// to (a, conj(b)) (architecture.h.in:580-586)
static inline void addprod_conj(cf32& acc, const cf32& a, const cf32& b) {
  f32 cre = b.re, cim = -b.im;
  acc.re += a.re * cre - a.im * cim;
  acc.im += a.re * cim + a.im * cre;
}

void GpuMode::process(int index, int subloopindex) {
  std::vector<cf32>& X = freqs_[subloopindex];
  fxg_block_info info{};
  dataweight_[subloopindex] = 0.0f;
  std::vector<f32>& bw = bandweight_[subloopindex];
  std::fill(bw.begin(), bw.end(), 0.0f);
  if (core_->station_block(ds_, index, (float*)X.data(), &info) != 0 || info.status != 0) {
    std::fill(X.begin(), X.end(), cf32{0, 0});
    std::fill(conj_[subloopindex].begin(), conj_[subloopindex].end(), cf32{0, 0});
    return;
  }
  dataweight_[subloopindex] = info.weight;
  if (core_->band_weights(ds_, index, bw.data()) != 0) std::fill(bw.begin(), bw.end(), info.weight);
  std::vector<cf32>& Xc = conj_[subloopindex];
  for (size_t i = 0; i < X.size(); i++) Xc[i] = cf32{X[i].re, -X[i].im};
  // autocorrelations (mode.cpp:1292-1375)
  const fxg_datastream& d = core_->dss_[ds_];
  const bool cross = core_->desc_.crosspol_autocorrs != 0;
  for (int i = 0; i < d.nrecfreqs; i++) {
    int idx[FXG_MAX_BANDS];
    int count = 0;
    for (int j = 0; j < nrec_; j++) {
      if (d.recband_freq[j] != i) continue;
      idx[count++] = j;
      const cf32* x = &X[(size_t)j * nchan_];
      cf32* a = &ac_[0][(size_t)j * nchan_];
      for (int c = 0; c < nchan_; c++) addprod_conj(a[c], x[c], x[c]);
      acw_[0][j] += bw[(size_t)j];                           // mode.cpp:1300-1307
    }
    if (count > 1 && cross) {
      const cf32* x0 = &X[(size_t)idx[0] * nchan_];
      const cf32* x1 = &X[(size_t)idx[1] * nchan_];
      cf32* a0 = &ac_[1][(size_t)idx[0] * nchan_];
      cf32* a1 = &ac_[1][(size_t)idx[1] * nchan_];
      for (int c = 0; c < nchan_; c++) {
        addprod_conj(a0[c], x0[c], x1[c]);
        addprod_conj(a1[c], x1[c], x0[c]);
      }
      // mode.cpp:1365-1373: the product of the two bands' weights for
      // INTERLACEDVDIF, else the block's weight
      const f32 wx = d.format == FXG_FORMAT_INTERLACEDVDIF ? bw[(size_t)idx[0]] * bw[(size_t)idx[1]] : info.weight;
      acw_[1][idx[0]] += wx;
      acw_[1][idx[1]] += wx;
    }
  }
}

const cf32* GpuMode::getFreqs(int band, int sub) const {
  const fxg_datastream& d = core_->dss_[ds_];
  if (band < nrec_) return &freqs_[sub][(size_t)band * nchan_];
  int z = band - nrec_;
  return &freqs_[sub][(size_t)d.zoomband_parent[z] * nchan_ + d.zoomband_choffset[z]];
}

const cf32* GpuMode::getConjugatedFreqs(int band, int sub) const {
  const fxg_datastream& d = core_->dss_[ds_];
  if (band < nrec_) return &conj_[sub][(size_t)band * nchan_];
  int z = band - nrec_;
  return &conj_[sub][(size_t)d.zoomband_parent[z] * nchan_ + d.zoomband_choffset[z]];
}

// mode.h:168: the record band's own weight (INTERLACEDVDIF), else the block's
f32 GpuMode::getDataWeight(int outputband, int sub) const {
  if (outputband >= 0 && outputband < nrec_) return bandweight_[sub][(size_t)outputband];
  return dataweight_[sub];
}

const cf32* GpuMode::getAutocorrelation(bool crosspol, int band) const {
  const fxg_datastream& d = core_->dss_[ds_];
  const std::vector<cf32>& a = ac_[crosspol ? 1 : 0];
  if (band < nrec_) return &a[(size_t)band * nchan_];
  int z = band - nrec_;   // zoom bands alias the (averaged) parent, mode.cpp:396
  return &a[(size_t)d.zoomband_parent[z] * nchan_ + d.zoomband_choffset[z] / chanavg_];
}

f32 GpuMode::getWeight(bool crosspol, int band) const {
  const fxg_datastream& d = core_->dss_[ds_];
  if (band >= nrec_) band = d.zoomband_parent[band - nrec_];
  return acw_[crosspol ? 1 : 0][band];
}

bool GpuMode::calculateAndAverageKurtosis(int thread) {
  kurtchan_ = std::min(core_->kurtchan_, nchan_);
  kurt_.assign((size_t)nrec_ * kurtchan_, 0.0f);
  int32_t n = 0, valid = 0;
  for (int b = 0; b < nrec_; b++) {
    if (fxg_job_get_kurtosis(core_->job_, thread, ds_, b, &kurt_[(size_t)b * kurtchan_], kurtchan_, &n, &valid) != 0)
      return false;
  }
  return valid != 0;
}

long long GpuMode::getEstimatedBytes() const {
  return (long long)(sizeof(cf32) * (2 * freqs_.size() * freqs_[0].size() + 2 * ac_[0].size()));
}

bool GpuMode::writeCrossAutoCorrs() const { return core_->desc_.crosspol_autocorrs != 0; }

int GpuMk5Mode::format() const { return core_->dss_[ds_].format; }

// ---------------------------------------------------------------------------
// GpuCore
// ---------------------------------------------------------------------------
GpuCore::GpuCore(int device, const fxg_job_desc& job, int numBufferedFFTs) {
  desc_ = job;
  freqs_.assign(job.freqs, job.freqs + job.nfreqs);
  dss_.assign(job.datastreams, job.datastreams + job.ndatastreams);
  if (job.nbaselines) bls_.assign(job.baselines, job.baselines + job.nbaselines);
  desc_.freqs = freqs_.data();
  desc_.datastreams = dss_.data();
  desc_.baselines = bls_.empty() ? nullptr : bls_.data();
  int rc = fxg_plan(&desc_, &info_);
  if (rc == 0) rc = fxg_ctx_create(device, &ctx_);
  if (rc == 0) rc = fxg_job_create(ctx_, &desc_, &job_);
  if (rc != 0) {
    err_ = fxg_last_error();
    return;
  }
  for (int d = 0; d < job.ndatastreams; d++) {
    if (dss_[d].format == FXG_FORMAT_MARK5B) modes_.emplace_back(new GpuMark5BMode(this, d, numBufferedFFTs));
    else modes_.emplace_back(new GpuVDIFMode(this, d, numBufferedFFTs));
  }
  ok_ = true;
}

GpuCore::~GpuCore() {
  if (job_) fxg_job_destroy(job_);
  if (ctx_) fxg_ctx_destroy(ctx_);
}

void GpuCore::gather(std::vector<fxg_station_input>& in, fxg_subint_desc& sub) const {
  in.resize(modes_.size());
  for (size_t d = 0; d < modes_.size(); d++) {
    const GpuMode& m = *modes_[d];
    in[d] = m.in_;
    in[d].nvalidflags = (int32_t)m.flags_.size();
    in[d].validflags = m.flags_.data();
    if (!m.modelok_) in[d].sec = FXG_INVALID_SUBINT;   // setOffsets could not find a model
  }
  sub = fxg_subint_desc{};
  const GpuMode& m0 = *modes_[0];
  sub.scan = m0.offscan_;
  sub.sec = m0.offsec_;
  sub.ns = m0.offns_;
  sub.scan_start_sec = scanstart_;
  sub.model_ok = 1;
}

int GpuCore::band_weights(int ds, int block, float* out) {
  if (!ok_) return FXG_ESTATE;
  int rc = fxg_tap_band_weights(job_, ds, block, out);
  if (rc) err_ = fxg_last_error();
  return rc;
}

int GpuCore::station_block(int ds, int block, float* spectra, fxg_block_info* info) {
  if (!ok_) return FXG_ESTATE;
  std::vector<fxg_station_input> in;
  fxg_subint_desc sub;
  gather(in, sub);
  int rc = fxg_station_process_block(job_, &sub, in.data(), 0, ds, block, spectra, info);
  if (rc) err_ = fxg_last_error();
  return rc;
}

int GpuCore::processdata(cf32* results) {
  if (!ok_) return FXG_ESTATE;
  std::vector<fxg_station_input> in;
  fxg_subint_desc sub;
  gather(in, sub);
  int rc = fxg_subint_process(job_, &sub, in.data(), 0, (float*)results, 0);
  if (rc) err_ = fxg_last_error();
  return rc;
}

int GpuCore::setDumpKurtosis(bool dump, int stadumpchannels, int numprocessthreads) {
  if (!ok_) return FXG_ESTATE;
  int rc = fxg_job_set_kurtosis(job_, dump ? 1 : 0, stadumpchannels, numprocessthreads);
  if (rc) err_ = fxg_last_error();
  else kurtchan_ = stadumpchannels;
  return rc;
}

long long GpuCore::getCoreResultBaselineOffset(int f, int b) const {
  int64_t o = -1;
  fxg_plan_baseline_offset(&desc_, f, b, &o);
  return o;
}
long long GpuCore::getCoreResultBWeightOffset(int f, int b) const {
  int64_t o = -1;
  fxg_plan_bweight_offset(&desc_, f, b, &o);
  return o;
}
long long GpuCore::getCoreResultAutocorrOffset(int d) const {
  int64_t o = -1;
  fxg_plan_autocorr_offset(&desc_, d, &o);
  return o;
}
long long GpuCore::getCoreResultACWeightOffset(int d) const {
  int64_t o = -1;
  fxg_plan_acweight_offset(&desc_, d, &o);
  return o;
}

}  // namespace fxgpu
