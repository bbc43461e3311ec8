// Drives the Mode/Core interfaces (difx-svn2git-try3_amd/host/gpumode.h) the
// way mpifxcorr's Core does, on a job and data written by the Python test:
//   (a) GpuCore::processdata for the whole slot;
//   (b) GpuMode::process block by block, with the baseline loop of
//       Core::processdata (core.cpp:826-1045) and averageAndSendAutocorrs
//       (core.cpp:1299-1369) done here on the host from getFreqs /
//       getConjugatedFreqs / getDataWeight / getAutocorrelation / getWeight.
// Writes both result buffers and exits non-zero if they differ by more than
// 1e-5 relative.  Usage: test_host_shim <dir>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "gpumode.h"

using namespace fxgpu;

static std::vector<char> slurp(const std::string& p) {
  FILE* f = fopen(p.c_str(), "rb");
  if (!f) { fprintf(stderr, "cannot open %s\n", p.c_str()); exit(2); }
  fseek(f, 0, SEEK_END);
  long n = ftell(f);
  fseek(f, 0, SEEK_SET);
  std::vector<char> b(n);
  if (n && fread(b.data(), 1, n, f) != (size_t)n) exit(2);
  fclose(f);
  return b;
}

static void spit(const std::string& p, const void* d, size_t n) {
  FILE* f = fopen(p.c_str(), "wb");
  fwrite(d, 1, n, f);
  fclose(f);
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  std::string dir = argv[1];
  std::vector<char> jb = slurp(dir + "/job.bin");
  const int32_t* h = (const int32_t*)jb.data();
  int nfreqs = h[0], nds = h[1], nbl = h[2], bps = h[3], prec = h[4], cross = h[5], scanstart = h[6];
  int subscan = h[7], subsec = h[8], subns = h[9];
  size_t off = 16 * sizeof(int32_t);
  std::vector<fxg_freq> freqs(nfreqs);
  std::vector<fxg_datastream> dss(nds);
  std::vector<fxg_baseline> bls(nbl);
  memcpy(freqs.data(), jb.data() + off, sizeof(fxg_freq) * nfreqs); off += sizeof(fxg_freq) * nfreqs;
  memcpy(dss.data(), jb.data() + off, sizeof(fxg_datastream) * nds); off += sizeof(fxg_datastream) * nds;
  memcpy(bls.data(), jb.data() + off, sizeof(fxg_baseline) * nbl); off += sizeof(fxg_baseline) * nbl;
  if (off != jb.size()) { fprintf(stderr, "job.bin size mismatch (%zu vs %zu)\n", off, jb.size()); return 2; }
  fxg_job_desc job{};
  job.nfreqs = nfreqs; job.ndatastreams = nds; job.nbaselines = nbl; job.blockspersend = bps;
  job.fringe_order = 1; job.crosspol_autocorrs = cross; job.xmac_precision = prec;
  job.freqs = freqs.data(); job.datastreams = dss.data(); job.baselines = bls.data();

  const int nbuf = 4;   // numBufferedFFTs
  GpuCore core(0, job, nbuf);
  if (!core.initialisedOK()) { fprintf(stderr, "GpuCore: %s\n", core.error().c_str()); return 3; }
  core.setScanStartSec(scanstart);
  std::vector<std::vector<char>> raws(nds);
  for (int d = 0; d < nds; d++) {
    raws[d] = slurp(dir + "/st" + std::to_string(d) + ".raw");
    std::vector<char> m = slurp(dir + "/st" + std::to_string(d) + ".meta");
    const int32_t* mi = (const int32_t*)m.data();
    int scan = mi[0], sec = mi[1], ns = mi[2], nflags = mi[3];
    const s32* flags = mi + 4;
    const double* interp = (const double*)(mi + 4 + nflags + (nflags & 1));
    GpuMode* mode = core.getMode(d);
    mode->zeroAutocorrelations();
    mode->setValidFlags(flags);
    mode->setData((const u8*)raws[d].data(), (int)raws[d].size(), scan, sec, ns);
    mode->setOffsets(subscan, subsec, subns, interp);
  }
  const long long rl = core.getCoreResultLength();
  std::vector<cf32> ra(rl), rb(rl, cf32{0, 0});
  // spectral kurtosis: 32 channels per record, two Core threads
  const int kthreads = 2, kchan = 32;
  if (core.setDumpKurtosis(true, kchan, kthreads)) { fprintf(stderr, "setDumpKurtosis: %s\n", core.error().c_str()); return 4; }
  int rc = core.processdata(ra.data());
  if (rc) { fprintf(stderr, "processdata: %d %s\n", rc, core.error().c_str()); return 4; }
  {
    // records [thread][datastream][recorded band][channel], then the valid flags [thread][datastream]
    std::vector<float> krec;
    std::vector<float> kvalid;
    for (int t = 0; t < kthreads; t++)
      for (int d = 0; d < nds; d++) {
        GpuMode* m = core.getMode(d);
        kvalid.push_back(m->calculateAndAverageKurtosis(t) ? 1.0f : 0.0f);
        for (int b = 0; b < dss[d].nrecbands; b++) {
          const f32* k = m->getKurtosis(b);
          krec.insert(krec.end(), k, k + m->getKurtosisChannels());
        }
      }
    krec.insert(krec.end(), kvalid.begin(), kvalid.end());
    spit(dir + "/kurtosis.bin", krec.data(), sizeof(float) * krec.size());
  }

  // (b) Mode::process per block + Core's baseline loop on the host
  std::vector<float> bw(2 * rl, 0.0f);
  for (int b0 = 0; b0 < bps; b0 += nbuf) {
    int nb = std::min(nbuf, bps - b0);
    for (int d = 0; d < nds; d++)
      for (int s = 0; s < nb; s++) core.getMode(d)->process(b0 + s, s);
    for (int f = 0; f < nfreqs; f++) {
      int nchan = freqs[f].nchan;
      for (int b = 0; b < nbl; b++) {
        long long bo = core.getCoreResultBaselineOffset(f, b);
        if (bo < 0) continue;
        int lf = -1;
        for (int k = 0; k < bls[b].nfreqs; k++) if (bls[b].freq_index[k] == f) lf = k;
        GpuMode* m1 = core.getMode(bls[b].ds1);
        GpuMode* m2 = core.getMode(bls[b].ds2);
        for (int s = 0; s < nb; s++)
          for (int p = 0; p < bls[b].npolproducts[lf]; p++) {
            const cf32* v1 = m1->getFreqs(bls[b].band1[lf][p], s);
            const cf32* v2 = m2->getConjugatedFreqs(bls[b].band2[lf][p], s);
            cf32* acc = &rb[bo + (long long)p * nchan];
            for (int c = 0; c < nchan; c++) {
              acc[c].re += v1[c].re * v2[c].re - v1[c].im * v2[c].im;
              acc[c].im += v1[c].re * v2[c].im + v1[c].im * v2[c].re;
            }
            float* w = (float*)&rb[core.getCoreResultBWeightOffset(f, b)] + p;
            *w += m1->getDataWeight(0, s) * m2->getDataWeight(0, s);
          }
      }
    }
  }
  // autocorrelations and their weights (core.cpp:1253-1369), one pass per slot
  for (int d = 0; d < nds; d++) {
    GpuMode* m = core.getMode(d);
    m->averageFrequency();
    const fxg_datastream& ds = dss[d];
    long long at = core.getCoreResultAutocorrOffset(d);
    float* wat = (float*)&rb[core.getCoreResultACWeightOffset(d)];
    int passes = 1;
    for (int b = 0; b < nbl; b++) for (int k = 0; k < bls[b].nfreqs; k++) if (bls[b].npolproducts[k] > 2) passes = 2;
    for (int pass = 0; pass < passes; pass++) {
      for (int k = 0; k < ds.nrecbands + ds.nzoombands; k++) {
        int fk = k < ds.nrecbands ? ds.recfreq_index[ds.recband_freq[k]] : ds.zoomband_freqindex[k - ds.nrecbands];
        bool used = false;
        for (int b = 0; b < nbl; b++) for (int q = 0; q < bls[b].nfreqs; q++) used |= bls[b].freq_index[q] == fk;
        if (!used) continue;
        int nchan = freqs[fk].nchan / freqs[fk].chans_to_avg;
        if (pass == 0 || m->writeCrossAutoCorrs()) {
          const cf32* a = m->getAutocorrelation(pass == 1, k);
          for (int c = 0; c < nchan; c++) { rb[at + c].re += a[c].re; rb[at + c].im += a[c].im; }
          *wat += m->getWeight(false, k);
        }
        wat++;
        at += nchan;
      }
    }
  }
  spit(dir + "/results_core.bin", ra.data(), sizeof(cf32) * rl);
  spit(dir + "/results_blocks.bin", rb.data(), sizeof(cf32) * rl);
  double num = 0, den = 0;
  for (long long i = 0; i < rl; i++) {
    double dr = (double)ra[i].re - rb[i].re, di = (double)ra[i].im - rb[i].im;
    num += dr * dr + di * di;
    den += (double)rb[i].re * rb[i].re + (double)rb[i].im * rb[i].im;
  }
  double rel = std::sqrt(num / den);
  printf("processdata vs per-block Mode::process: rel %.3e over %lld cf32\n", rel, rl);
  return rel < 1e-5 ? 0 : 1;
}
