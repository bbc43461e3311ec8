// visibility.h — the reference's Visibility (mpifxcorr/src/visibility.h:40-210)
// for one configuration with one phase centre and no pulsar binning: sums
// the slot results of one integration (addData), normalises them
// (writedata: visibility.cpp:348-560) and writes DiFX binary records
// (writedifx / writeDiFXHeader: visibility.cpp:716-900, 1050-1131).
#pragma once

#include <array>
#include <string>
#include <vector>

#include "fxcorr_gpu.h"
#include "gpumode.h"

namespace fxgpu {

struct VisConfig {
  int expermjd = 0;                 // experiment start (Configuration::getStartMJD)
  int experseconds = 0;             // and seconds (getStartSeconds)
  double inttime_s = 1.0;           // integration time (getIntTime)
  long long subintns = 0;           // subintegration length in ns (getSubintNS)
  int configindex = 0;
  int sourceindex = 0;              // the single phase centre's source index
  bool writeautocorrs = true;       // writeAutoCorrs
  std::vector<int> telescope;       // telescope index per datastream (default: datastream index)
  std::vector<double> tsys;         // a-priori Tsys per datastream; 0 = correlation coefficients
  std::string outfile;              // DIFX_<mjd>_<sec>.s0000.b0000 path; empty = do not write
};

// One DiFX binary record as written by writeDiFXHeader + the data.
struct VisRecord {
  int baseline, mjd;
  double seconds;
  int configindex, sourceindex, freqindex;
  char polpair[2];
  int pulsarbin;
  double weight;
  double uvw[3];
  std::vector<cf32> data;
};

class GpuVisibility {
 public:
  GpuVisibility(const fxg_job_desc& job, const VisConfig& cfg);
  const std::string& error() const { return err_; }
  // visibility.cpp:181-193: results += subint; true when the integration is complete
  bool addData(const cf32* subintresults);
  // the integration's time (Visibility::currentstartseconds / currentstartns
  // relative to the start of scan `scanstartsec` seconds after the experiment start)
  void setTime(int scanstartsec, int startseconds, int startns);
  // visibility.cpp:348-560 then writedifx (:716-900): normalises the summed
  // results in place and appends the records to cfg.outfile (when set).
  // uvw[baseline] are the baselines' (u,v,w) in metres (zeros when null).
  int writedata(const std::vector<std::array<double, 3>>* uvw = nullptr);
  // the records writedata produced (also what was written to the file)
  const std::vector<VisRecord>& records() const { return recs_; }
  const std::vector<cf32>& results() const { return results_; }
  int subintsThisIntegration() const { return subintsthisintegration_; }
  void zero();                      // start the next integration

 private:
  int ds_band_freq(int ds, int band) const;
  char ds_band_pol(int ds, int band) const;
  bool freq_wanted(int f) const;
  fxg_job_desc desc_{};
  std::vector<fxg_freq> freqs_;
  std::vector<fxg_datastream> dss_;
  std::vector<fxg_baseline> bls_;
  VisConfig cfg_;
  std::string err_;
  fxg_plan_info info_{};
  std::vector<char> used_, equiv_;
  int autocorrwidth_ = 1;
  std::vector<cf32> results_;
  int currentsubints_ = 0, subintsthisintegration_ = 1;
  double meansubints_ = 1.0, fftsperint_ = 1.0;
  int scanstart_ = 0, startsec_ = 0, startns_ = 0;
  std::vector<VisRecord> recs_;
};

}  // namespace fxgpu
