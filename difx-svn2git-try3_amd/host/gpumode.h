// gpumode.h — the reference's Mode / Mk5Mode / Core / Model interfaces for
// mpifxcorr's station and baseline processing, backed by libfxcorr_gpu.so on
// one MI355X.
//
//   GpuMode      Mode        (mpifxcorr/src/mode.h:50-355)
//   GpuMk5Mode   Mk5Mode     (mpifxcorr/src/mk5mode.h:40-87): mark5access
//                formats; GpuVDIFMode and GpuMark5BMode fix the format
//   GpuCore      Core::processdata for one slot (mpifxcorr/src/core.cpp:667)
//   DelayModel   Model::calculateDelayInterpolator (mpifxcorr/src/model.cpp:229)
//
// Method names, argument meaning and the zero-output behaviour on bad data
// follow the reference.  GpuCore::processdata runs a whole subintegration in
// one call; GpuMode::process(index, subloop) runs one FFT block for callers
// that drive Mode block by block and read getFreqs().
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "fxcorr_gpu.h"

namespace fxgpu {

typedef uint8_t u8;
typedef int32_t s32;
typedef float f32;
typedef double f64;
struct cf32 { f32 re, im; };

// Model::calculateDelayInterpolator for one antenna and one phase centre: the
// delay polynomials (microseconds; time in seconds since the start of each
// model interval) of consecutive model intervals of incsecs seconds each.
// polyoffset_s is the time of the scan's offset origin after the start of the
// first interval (model.cpp:242's polyoffset).
class DelayModel {
 public:
  DelayModel(std::vector<std::vector<f64>> polys_us, f64 incsecs, f64 polyoffset_s = 0.0)
      : polys_(std::move(polys_us)), inc_(incsecs), polyoffset_(polyoffset_s) {}
  // model.cpp:229-304: delaycoeffs[0..order] for numincrements equal steps over
  // timespan seconds from offsettime; false if order is not 0..2 or the time
  // lies outside the model's intervals.
  bool calculateDelayInterpolator(f64 offsettime_s, f64 timespan_s, int numincrements, int order,
                                  f64* delaycoeffs) const;
 private:
  static f64 eval(const std::vector<f64>& c, f64 deltat);
  std::vector<std::vector<f64>> polys_;
  f64 inc_, polyoffset_;
};

class GpuCore;

class GpuMode {
 public:
  virtual ~GpuMode() = default;
  // mode.cpp:1564-1573
  void setData(const u8* d, int dbytes, int dscan, int dsec, int dns);
  // mode.cpp:1557-1562
  void setValidFlags(const s32* v);
  // The datastream's delay model and integer clock seconds (what Mode reads
  // from Model and Configuration).
  void setDelayModel(const DelayModel* model, int intclockseconds = 0);
  // mode.cpp:1525-1555: records the subint time and fills the order-2 delay
  // interpolator from the delay model; a scan other than the data's torches the
  // data, a time the model does not cover torches the subintegration.
  void setOffsets(int scan, int seconds, int ns);
  // The same with the interpolator already computed (integer clock seconds
  // already removed); foundok = false torches the subintegration.
  void setOffsets(int scan, int seconds, int ns, const f64 interpolator[3], bool foundok = true);
  // mode.cpp:1480-1493
  void zeroAutocorrelations();
  // mode.cpp:1399-1424: autocorrelations averaged in place over the
  // datastream's channelstoaverage (its first recorded frequency's)
  void averageFrequency();
  // mode.cpp:649-1397 for one FFT block
  void process(int index, int subloopindex);
  // mode.h:199, 207 (band index counts zoom bands after the recorded bands)
  const cf32* getFreqs(int outputband, int subloopindex) const;
  const cf32* getConjugatedFreqs(int outputband, int subloopindex) const;
  // mode.h:168
  f32 getDataWeight(int outputband, int subloopindex) const;
  // mode.h:148, 161
  const cf32* getAutocorrelation(bool crosspol, int outputband) const;
  f32 getWeight(bool crosspol, int outputband) const;
  bool initialisedOK() const { return initok_; }
  // Spectral kurtosis of Core process thread `thread` over its blocks of the
  // last processdata (GpuCore::setDumpKurtosis): fetches every recorded
  // band's record and returns whether it is valid (mode.cpp:1426-1478);
  // getKurtosis(band) then points at that band's channels (mode.h:154).
  bool calculateAndAverageKurtosis(int thread = 0);
  const f32* getKurtosis(int outputband) const { return &kurt_[(size_t)outputband * kurtchan_]; }
  int getKurtosisChannels() const { return kurtchan_; }
  long long getEstimatedBytes() const;
  bool writeCrossAutoCorrs() const;
  int datastream() const { return ds_; }

 protected:
  GpuMode(GpuCore* core, int dsindex, int numBufferedFFTs);
  friend class GpuCore;
  GpuCore* core_;
  int ds_;
  bool initok_ = true;
  int nbuf_;
  int nrec_, nzoom_, nchan_, chanavg_;
  // inputs of the current subintegration (what Core hands Mode, core.cpp:705-730)
  fxg_station_input in_{};
  std::vector<s32> flags_;
  int offscan_ = 0, offsec_ = 0, offns_ = 0;
  bool modelok_ = true;
  const DelayModel* model_ = nullptr;
  int intclock_ = 0;
  // per subloop: spectra and conjugates of every recorded band, weights
  std::vector<std::vector<cf32>> freqs_, conj_;
  std::vector<f32> dataweight_;
  // per subloop and recorded band: Mode::getDataWeight (INTERLACEDVDIF
  // per-band weights, else the block's weight)
  std::vector<std::vector<f32>> bandweight_;
  // autocorrelations[2][recorded band][chan] and weights[2][recorded band]
  std::vector<cf32> ac_[2];
  std::vector<f32> acw_[2];
  // the fetched kurtosis records [recorded band][kurtchan_]
  std::vector<f32> kurt_;
  int kurtchan_ = 0;
};

// Mk5Mode: the mark5access-decoded formats (VDIF, Mark5B)
class GpuMk5Mode : public GpuMode {
 public:
  int format() const;
 protected:
  GpuMk5Mode(GpuCore* core, int dsindex, int nbuf) : GpuMode(core, dsindex, nbuf) {}
  friend class GpuCore;
};

class GpuVDIFMode : public GpuMk5Mode {
 protected:
  GpuVDIFMode(GpuCore* core, int dsindex, int nbuf) : GpuMk5Mode(core, dsindex, nbuf) {}
  friend class GpuCore;
};

class GpuMark5BMode : public GpuMk5Mode {
 protected:
  GpuMark5BMode(GpuCore* core, int dsindex, int nbuf) : GpuMk5Mode(core, dsindex, nbuf) {}
  friend class GpuCore;
};

// One correlator configuration on one GPU: the Modes of every datastream and
// the per-slot baseline processing.
class GpuCore {
 public:
  // job: the configuration tables; the descriptor's arrays are copied.
  GpuCore(int device, const fxg_job_desc& job, int numBufferedFFTs = 1);
  ~GpuCore();
  bool initialisedOK() const { return ok_; }
  const std::string& error() const { return err_; }
  // Configuration::getMode (configuration.cpp:954-1016): the datastream's Mode
  GpuMode* getMode(int dsindex) { return modes_[dsindex].get(); }
  int numDatastreams() const { return (int)modes_.size(); }
  // populateResultLengths: length of the slot's cf32 result buffer
  long long getCoreResultLength() const { return info_.result_length; }
  // Core::processdata for the whole slot: every block of every datastream
  // whose Mode has been given setData/setValidFlags/setOffsets; results are
  // zeroed and accumulated in the reference layout.  Returns 0 or an FXG_E* code.
  int processdata(cf32* results);
  // Offsets of the reference layout (configuration.h getCoreResult*Offset)
  long long getCoreResultBaselineOffset(int freq, int baseline) const;
  long long getCoreResultBWeightOffset(int freq, int baseline) const;
  long long getCoreResultAutocorrOffset(int ds) const;
  long long getCoreResultACWeightOffset(int ds) const;
  const fxg_job_desc& desc() const { return desc_; }
  // Configuration::dumpKurtosis, getSTADumpChannels and Core's number of
  // process threads (one record each, core.cpp:1054-1056): kurtosis is
  // accumulated from the next processdata on.  Returns 0 or an FXG_E* code.
  int setDumpKurtosis(bool dump, int stadumpchannels = 32, int numprocessthreads = 1);
  int scanStartSec() const { return scanstart_; }
  void setScanStartSec(int s) { scanstart_ = s; }

 private:
  friend class GpuMode;
  friend class GpuMk5Mode;
  int station_block(int ds, int block, float* spectra, fxg_block_info* info);
  int band_weights(int ds, int block, float* out);
  void gather(std::vector<fxg_station_input>& in, fxg_subint_desc& sub) const;
  bool ok_ = false;
  std::string err_;
  fxg_ctx* ctx_ = nullptr;
  fxg_job* job_ = nullptr;
  fxg_plan_info info_{};
  fxg_job_desc desc_{};
  std::vector<fxg_freq> freqs_;
  std::vector<fxg_datastream> dss_;
  std::vector<fxg_baseline> bls_;
  std::vector<std::unique_ptr<GpuMode>> modes_;
  int scanstart_ = 0;
  int kurtchan_ = 32;
};

}  // namespace fxgpu
