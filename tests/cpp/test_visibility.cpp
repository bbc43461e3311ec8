// Drives fxgpu::GpuVisibility (difx-svn2git-try3_amd/host/visibility.h) on a
// slot result buffer written by the Python test: addData for nsub subints,
// setTime, writedata with a-priori Tsys (0 = correlation coefficients), the
// DiFX records appended to <dir>/out.difx.
// Usage: test_visibility <dir> <nsub> <tsys> <subintns> <inttime_s>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "visibility.h"

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

int main(int argc, char** argv) {
  if (argc < 6) return 2;
  std::string dir = argv[1];
  const int nsub = atoi(argv[2]);
  const double tsys = atof(argv[3]);
  const long long subintns = atoll(argv[4]);
  const double inttime = atof(argv[5]);
  std::vector<char> jb = slurp(dir + "/job.bin");
  const int32_t* h = (const int32_t*)jb.data();
  int nfreqs = h[0], nds = h[1], nbl = h[2], bps = h[3], cross = h[5];
  size_t off = 16 * sizeof(int32_t);
  std::vector<fxg_freq> freqs(nfreqs);
  std::vector<fxg_datastream> dss(nds);
  std::vector<fxg_baseline> bls(nbl);
  memcpy(freqs.data(), jb.data() + off, sizeof(fxg_freq) * nfreqs); off += sizeof(fxg_freq) * nfreqs;
  memcpy(dss.data(), jb.data() + off, sizeof(fxg_datastream) * nds); off += sizeof(fxg_datastream) * nds;
  memcpy(bls.data(), jb.data() + off, sizeof(fxg_baseline) * nbl);
  fxg_job_desc job{};
  job.nfreqs = nfreqs; job.ndatastreams = nds; job.nbaselines = nbl; job.blockspersend = bps;
  job.fringe_order = 1; job.crosspol_autocorrs = cross;
  job.freqs = freqs.data(); job.datastreams = dss.data(); job.baselines = bls.data();
  std::vector<char> rb = slurp(dir + "/results.bin");
  const cf32* res = (const cf32*)rb.data();
  VisConfig cfg;
  cfg.expermjd = 57000;
  cfg.experseconds = 86000;
  cfg.inttime_s = inttime;
  cfg.subintns = subintns;
  cfg.sourceindex = 3;
  cfg.tsys.assign(nds, tsys);
  for (int d = 0; d < nds; d++) cfg.telescope.push_back(2 * d + 1);
  cfg.outfile = dir + "/out.difx";
  GpuVisibility vis(job, cfg);
  if (!vis.error().empty()) { fprintf(stderr, "%s\n", vis.error().c_str()); return 3; }
  bool done = false;
  for (int s = 0; s < nsub; s++) done = vis.addData(res);
  vis.setTime(500, 7, 250000000);
  int rc = vis.writedata();
  if (rc) { fprintf(stderr, "writedata %d %s\n", rc, vis.error().c_str()); return 4; }
  printf("done=%d subintsthisintegration=%d records=%zu\n", done ? 1 : 0, vis.subintsThisIntegration(),
         vis.records().size());
  return 0;
}
