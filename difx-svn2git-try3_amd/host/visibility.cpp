// visibility.cpp — Visibility::addData / writedata / writedifx for one
// configuration, one phase centre, no pulsar binning, no phase-cal.
#include "visibility.h"

#include <cmath>
#include <cstdio>
#include <cstring>

namespace fxgpu {

namespace {
// Mode::decorrelationpercentage (mode.cpp:1594), indexed by nbits - 1
const float kDecorrelation[] = {0.63662f, 0.88f, 0.94f, 0.96f, 0.98f, 0.99f, 0.996f, 0.998f};
const unsigned int kSyncWord = 0xFF00FF00u;     // visibility.h:135
const int kBinaryHeaderVersion = 1;             // visibility.h:138
constexpr double kTiny = 0.000000001;           // mode.cpp:35

template <typename T>
void put(std::vector<char>& b, const T& v) {
  const char* p = reinterpret_cast<const char*>(&v);
  b.insert(b.end(), p, p + sizeof(T));
}
}  // namespace

GpuVisibility::GpuVisibility(const fxg_job_desc& job, const VisConfig& cfg) : cfg_(cfg) {
  desc_ = job;
  freqs_.assign(job.freqs, job.freqs + job.nfreqs);
  dss_.assign(job.datastreams, job.datastreams + job.ndatastreams);
  if (job.nbaselines) bls_.assign(job.baselines, job.baselines + job.nbaselines);
  desc_.freqs = freqs_.data();
  desc_.datastreams = dss_.data();
  desc_.baselines = bls_.empty() ? nullptr : bls_.data();
  if (fxg_plan(&desc_, &info_) != 0) {
    err_ = fxg_last_error();
    return;
  }
  if ((int)cfg_.telescope.size() < job.ndatastreams)
    for (int d = (int)cfg_.telescope.size(); d < job.ndatastreams; d++) cfg_.telescope.push_back(d);
  cfg_.tsys.resize(job.ndatastreams, 0.0);
  // isFrequencyUsed / isEquivalentFrequencyUsed and the largest product count
  used_.assign(job.nfreqs, 0);
  equiv_.assign(job.nfreqs, 0);
  int maxprod = 0;
  for (const fxg_baseline& b : bls_)
    for (int f = 0; f < b.nfreqs; f++) {
      used_[b.freq_index[f]] = 1;
      maxprod = std::max(maxprod, (int)b.npolproducts[f]);
    }
  for (int f = 0; f < job.nfreqs; f++) {
    if (used_[f]) continue;
    const fxg_freq& x = freqs_[f];
    const double xlo = x.freq_mhz - (x.lsb ? x.bw_mhz : 0.0);
    for (int g = 0; g < job.nfreqs && !equiv_[f]; g++) {
      const fxg_freq& y = freqs_[g];
      const double ylo = y.freq_mhz - (y.lsb ? y.bw_mhz : 0.0);
      if (used_[g] && x.bw_mhz - y.bw_mhz < kTiny && xlo - ylo < kTiny && x.nchan == y.nchan &&
          x.chans_to_avg == y.chans_to_avg)
        equiv_[f] = 1;
    }
  }
  autocorrwidth_ = (maxprod > 2 && cfg_.writeautocorrs) ? 2 : 1;          // visibility.cpp:55-58
  // visibility.cpp:1198-1200, 93
  if (cfg_.subintns <= 0) {
    err_ = "subintns must be positive";
    return;
  }
  meansubints_ = cfg_.inttime_s / ((double)cfg_.subintns / 1000000000.0);
  fftsperint_ = meansubints_ * job.blockspersend;
  subintsthisintegration_ = (int)(((long long)(cfg_.inttime_s * 1000000000.0)) / cfg_.subintns);
  results_.assign((size_t)info_.result_length, cf32{0, 0});
}

int GpuVisibility::ds_band_freq(int d, int k) const {
  const fxg_datastream& ds = dss_[d];
  return k < ds.nrecbands ? ds.recfreq_index[ds.recband_freq[k]] : ds.zoomband_freqindex[k - ds.nrecbands];
}

char GpuVisibility::ds_band_pol(int d, int k) const {
  const fxg_datastream& ds = dss_[d];
  return k < ds.nrecbands ? ds.recband_pol[k] : ds.zoomband_pol[k - ds.nrecbands];
}

bool GpuVisibility::freq_wanted(int f) const { return used_[f] || equiv_[f]; }

bool GpuVisibility::addData(const cf32* sub) {
  for (size_t i = 0; i < results_.size(); i++) {
    results_[i].re += sub[i].re;
    results_[i].im += sub[i].im;
  }
  currentsubints_++;
  return currentsubints_ >= subintsthisintegration_;
}

void GpuVisibility::setTime(int scanstartsec, int startseconds, int startns) {
  scanstart_ = scanstartsec;
  startsec_ = startseconds;
  startns_ = startns;
}

void GpuVisibility::zero() {
  std::fill(results_.begin(), results_.end(), cf32{0, 0});
  currentsubints_ = 0;
  recs_.clear();
}

int GpuVisibility::writedata(const std::vector<std::array<double, 3>>* uvw) {
  if (!err_.empty()) return FXG_ESTATE;
  recs_.clear();
  if (currentsubints_ == 0) return 0;                                     // nothing to write out
  float* fr = reinterpret_cast<float*>(results_.data());
  const int nds = (int)dss_.size(), nbl = (int)bls_.size();
  const int bps = desc_.blockspersend;
  auto offset = [&](int (*fn)(const fxg_job_desc*, int, int, int64_t*), int a, int b) {
    int64_t o = -1;
    fn(&desc_, a, b, &o);
    return (long long)o;
  };
  auto dsoffset = [&](int (*fn)(const fxg_job_desc*, int, int64_t*), int d) {
    int64_t o = -1;
    fn(&desc_, d, &o);
    return (long long)o;
  };
  // baseline weights (visibility.cpp:384-401)
  std::vector<std::vector<std::vector<float>>> bweight(nbl);
  for (int i = 0; i < nbl; i++) {
    bweight[i].resize(bls_[i].nfreqs);
    for (int j = 0; j < bls_[i].nfreqs; j++) {
      long long ri = offset(fxg_plan_bweight_offset, bls_[i].freq_index[j], i) * 2;
      for (int k = 0; k < bls_[i].npolproducts[j]; k++) bweight[i][j].push_back((float)(fr[ri++] / fftsperint_));
    }
  }
  // autocorrelation weights (visibility.cpp:418-436)
  std::vector<std::vector<std::vector<float>>> acw(nds);
  for (int d = 0; d < nds; d++) {
    const int nb = dss_[d].nrecbands + dss_[d].nzoombands;
    acw[d].assign(autocorrwidth_, std::vector<float>(nb, 0.0f));
    long long ri = dsoffset(fxg_plan_acweight_offset, d) * 2;
    for (int j = 0; j < autocorrwidth_; j++)
      for (int k = 0; k < nb; k++)
        if (freq_wanted(ds_band_freq(d, k))) acw[d][j][k] = (float)(fr[ri++] / fftsperint_);
  }
  // band averages of the autocorrelations for Tsys calibration (visibility.cpp:438-457)
  std::vector<std::vector<cf32>> calib(nds);
  for (int d = 0; d < nds; d++) {
    const int nb = dss_[d].nrecbands + dss_[d].nzoombands;
    calib[d].assign(nb, cf32{0, 0});
    if (!(cfg_.tsys[d] > 0.0)) continue;
    long long ri = dsoffset(fxg_plan_autocorr_offset, d);
    for (int k = 0; k < nb; k++) {
      const int f = ds_band_freq(d, k);
      if (!freq_wanted(f)) continue;
      const int fc = freqs_[f].nchan / freqs_[f].chans_to_avg;
      cf32 m{0, 0};
      for (int c = 0; c < fc; c++) { m.re += results_[ri + c].re; m.im += results_[ri + c].im; }
      m.re /= fc;
      m.im /= fc;
      calib[d][k] = m;
      ri += fc;
    }
  }
  auto zoom_modifier = [&](int d, int k) {           // zoom bandwidth / parent bandwidth
    const fxg_datastream& ds = dss_[d];
    const int z = k - ds.nrecbands;
    const int parent = ds.zoomband_parent[z];
    const double pbw = freqs_[ds.recfreq_index[ds.recband_freq[parent]]].bw_mhz;
    return (float)(freqs_[ds.zoomband_freqindex[z]].bw_mhz / pbw);
  };
  // calibrate each baseline (visibility.cpp:459-519)
  for (int i = 0; i < nbl; i++) {
    const int ds1 = bls_[i].ds1, ds2 = bls_[i].ds2;
    for (int j = 0; j < bls_[i].nfreqs; j++) {
      const int f = bls_[i].freq_index[j];
      long long ri = offset(fxg_plan_baseline_offset, f, i);
      const int fc = freqs_[f].nchan / freqs_[f].chans_to_avg;
      for (int k = 0; k < bls_[i].npolproducts[j]; k++) {
        const int b1 = bls_[i].band1[j][k], b2 = bls_[i].band2[j][k];
        float scale;
        if (cfg_.tsys[ds1] > 0.0 && cfg_.tsys[ds2] > 0.0) {
          const float divisor = kDecorrelation[dss_[ds1].nbits - 1] * kDecorrelation[dss_[ds2].nbits - 1] *
                                calib[ds1][b1].re * calib[ds2][b2].re / (acw[ds1][0][b1] * acw[ds2][0][b2]);
          if (divisor > 0.0 && bweight[i][j][k] > 0)
            scale = (float)(std::sqrt(cfg_.tsys[ds1] * cfg_.tsys[ds2] / divisor) / bweight[i][j][k]);
          else
            scale = 0.0f;
        } else if (bweight[i][j][k] > 0.0) {
          scale = (float)(1.0 / (bweight[i][j][k] * meansubints_ *
                                 ((float)(bps * 2 * fc * freqs_[f].chans_to_avg))));
          if (b1 >= dss_[ds1].nrecbands) scale *= std::sqrt(zoom_modifier(ds1, b1));
          if (b2 >= dss_[ds2].nrecbands) scale *= std::sqrt(zoom_modifier(ds2, b2));
        } else {
          scale = 0.0f;
        }
        if (scale > 0.0f)
          for (int c = 0; c < 2 * fc; c++) fr[2 * ri + c] *= scale;
        ri += fc;
      }
    }
  }
  // calibrate the autocorrelations (visibility.cpp:521-560)
  if (cfg_.writeautocorrs) {
    for (int d = 0; d < nds; d++) {
      long long ri = dsoffset(fxg_plan_autocorr_offset, d) * 2;
      const int nb = dss_[d].nrecbands + dss_[d].nzoombands;
      for (int j = 0; j < autocorrwidth_; j++)
        for (int k = 0; k < nb; k++) {
          const int f = ds_band_freq(d, k);
          if (!freq_wanted(f)) continue;
          const int fc = freqs_[f].nchan / freqs_[f].chans_to_avg;
          float scale = 0.0f;
          if (cfg_.tsys[d] > 0.0) {
            int other = k;
            if (j != 0) {                       // Configuration::getDMatchingBand: same freq, other pol
              for (int q = 0; q < nb; q++)
                if (q != k && ds_band_freq(d, q) == f && ds_band_pol(d, q) != ds_band_pol(d, k)) { other = q; break; }
            }
            const float divisor = std::sqrt(calib[d][k].re * calib[d][other].re);
            if (divisor > 0.0f) scale = (float)(cfg_.tsys[d] / divisor);
          } else if (acw[d][j][k] > 0.0f) {
            scale = (float)(1.0 / (acw[d][j][k] * meansubints_ * ((float)(bps * 2 * fc * freqs_[f].chans_to_avg))));
            if (k >= dss_[d].nrecbands) scale *= zoom_modifier(d, k);
          }
          if (scale > 0.0f)
            for (int c = 0; c < 2 * fc; c++) fr[ri + c] *= scale;
          ri += 2 * fc;
        }
    }
  }
  // writedifx (visibility.cpp:716-900): cross records, then autocorrelations
  const int total = cfg_.experseconds + scanstart_ + startsec_;
  const int dumpmjd = cfg_.expermjd + total / 86400;
  const double dumpseconds = double(total % 86400) + ((double)startns_) / 1000000000.0 + cfg_.inttime_s / 2.0;
  for (int i = 0; i < nbl; i++) {
    const int ds1 = bls_[i].ds1, ds2 = bls_[i].ds2;
    const int blnum = (cfg_.telescope[ds1] + 1) * 256 + (cfg_.telescope[ds2] + 1);   // Configuration::getBNumber
    for (int j = 0; j < bls_[i].nfreqs; j++) {
      const int f = bls_[i].freq_index[j];
      long long ri = offset(fxg_plan_baseline_offset, f, i);
      const int fc = freqs_[f].nchan / freqs_[f].chans_to_avg;
      for (int k = 0; k < bls_[i].npolproducts[j]; k++) {
        if (bweight[i][j][k] > 0.0f) {
          VisRecord r{};
          r.baseline = blnum; r.mjd = dumpmjd; r.seconds = dumpseconds;
          r.configindex = cfg_.configindex; r.sourceindex = cfg_.sourceindex; r.freqindex = f;
          r.polpair[0] = ds_band_pol(ds1, bls_[i].band1[j][k]);
          r.polpair[1] = ds_band_pol(ds2, bls_[i].band2[j][k]);
          r.pulsarbin = 0;
          r.weight = bweight[i][j][k];
          for (int q = 0; q < 3; q++) r.uvw[q] = uvw && (int)uvw->size() > i ? (*uvw)[i][q] : 0.0;
          r.data.assign(results_.begin() + ri, results_.begin() + ri + fc);
          recs_.push_back(std::move(r));
        }
        ri += fc;
      }
    }
  }
  if (cfg_.writeautocorrs) {
    for (int d = 0; d < nds; d++) {
      const int blnum = 257 * (cfg_.telescope[d] + 1);
      long long ri = dsoffset(fxg_plan_autocorr_offset, d);
      const int nb = dss_[d].nrecbands + dss_[d].nzoombands;
      for (int j = 0; j < autocorrwidth_; j++)
        for (int k = 0; k < nb; k++) {
          const int f = ds_band_freq(d, k);
          if (!freq_wanted(f)) continue;
          const int fc = freqs_[f].nchan / freqs_[f].chans_to_avg;
          if (acw[d][j][k] > 0.0f) {
            VisRecord r{};
            r.baseline = blnum; r.mjd = dumpmjd; r.seconds = dumpseconds;
            r.configindex = cfg_.configindex; r.sourceindex = cfg_.sourceindex; r.freqindex = f;
            r.polpair[0] = ds_band_pol(d, k);
            r.polpair[1] = j == 0 ? r.polpair[0] : (r.polpair[0] == 'R' ? 'L' : r.polpair[0] == 'L' ? 'R'
                                                      : r.polpair[0] == 'X' ? 'Y' : 'X');   // getOppositePol
            r.pulsarbin = 0;
            r.weight = acw[d][j][k];
            r.data.assign(results_.begin() + ri, results_.begin() + ri + fc);
            recs_.push_back(std::move(r));
          }
          ri += fc;
        }
    }
  }
  if (!cfg_.outfile.empty()) {
    std::vector<char> buf;
    for (const VisRecord& r : recs_) {
      put(buf, kSyncWord);
      put(buf, kBinaryHeaderVersion);
      put(buf, r.baseline);
      put(buf, r.mjd);
      put(buf, r.seconds);
      put(buf, r.configindex);
      put(buf, r.sourceindex);
      put(buf, r.freqindex);
      buf.push_back(r.polpair[0]);
      buf.push_back(r.polpair[1]);
      put(buf, r.pulsarbin);
      put(buf, r.weight);
      for (int q = 0; q < 3; q++) put(buf, r.uvw[q]);
      const char* p = reinterpret_cast<const char*>(r.data.data());
      buf.insert(buf.end(), p, p + sizeof(cf32) * r.data.size());
    }
    FILE* fo = std::fopen(cfg_.outfile.c_str(), "ab");
    if (!fo) {
      err_ = "cannot open " + cfg_.outfile;
      return FXG_EINVAL;
    }
    const size_t n = std::fwrite(buf.data(), 1, buf.size(), fo);
    std::fclose(fo);
    if (n != buf.size()) {
      err_ = "short write to " + cfg_.outfile;
      return FXG_EINVAL;
    }
  }
  return 0;
}

}  // namespace fxgpu
