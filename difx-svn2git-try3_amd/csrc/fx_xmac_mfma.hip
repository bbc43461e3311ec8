// fx_xmac_mfma.hip — reduced-precision cross-multiply-accumulate on the
// gfx950 matrix cores (v_mfma_f32_32x32x16_{f16,bf16}, FP32 accumulation).
//
// For one frequency and one channel c, with X_a(t) the spectrum of input a
// (datastream band) at block t, the visibility matrix
//     V[a][b] = sum_t X_a(t) conj(X_b(t))                 (core.cpp:960)
// is two real products over K = 2T, k = 2t + {0 re, 1 im}:
//     Re V = A A^T,  A[a][2t] = re X_a(t),  A[a][2t+1] = im X_a(t)
//     Im V = Ah A^T, Ah[a][2t] = im X_a(t), Ah[a][2t+1] = -re X_a(t)
// The interleaved (re, im) half2 spectrum is the A operand as stored, and Ah
// is a half swap with one sign flip.  Re V is symmetric and Im V
// antisymmetric, so only the tiles (I <= J) of the 32x32 tile grid are
// accumulated.  One wavefront owns one channel; a workgroup of CG wavefronts
// owns CG consecutive channels and stages 32 blocks of every input through
// LDS, transposed to [channel][input][block] so that each lane's operand is
// one 16-byte LDS read.
#include <hip/hip_runtime.h>
#include "fx_kernels.h"

namespace fxg {

typedef _Float16 h8_t __attribute__((ext_vector_type(8)));
typedef __bf16 b8_t __attribute__((ext_vector_type(8)));
typedef float f16_t __attribute__((ext_vector_type(16)));

template <bool BF>
__device__ __forceinline__ f16_t mfma(uint4 a, uint4 b, f16_t c) {
  if constexpr (BF) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(b8_t, a), __builtin_bit_cast(b8_t, b), c, 0, 0, 0);
  } else {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(h8_t, a), __builtin_bit_cast(h8_t, b), c, 0, 0, 0);
  }
}

// (re, im) -> (im, -re) on each packed 16-bit pair
__device__ __forceinline__ unsigned hat(unsigned v) { return (v >> 16) | ((v ^ 0x8000u) << 16); }

template <int T, bool BF, int CG>
__global__ __launch_bounds__(64 * CG) void k_xmac_mfma(XmacArgs A) {
  // CG channels (wavefronts) per workgroup
  constexpr int NTH = 64 * CG;
  constexpr int TT = 32;            // blocks staged per pass
  constexpr int TP = TT + 4;        // padded row length in dwords
  constexpr int NR = 32 * T;        // inputs, padded to the tile size
  constexpr int NTILE = T * (T + 1) / 2;
  static_assert(CG % 4 == 0 && (NR * TT * (CG / 4)) % NTH == 0, "staging items must tile the workgroup");
  // per-channel stride: a multiple of 4 dwords that is 4 mod 16, so the four
  // channel rows one staging lane writes fall in different banks
  constexpr int CHS = NR * TP + 4;
  __shared__ unsigned lds[CG * CHS];

  const unsigned* spec = (const unsigned*)A.spec;
  const int tid = threadIdx.x;
  const int w = tid >> 6, l = tid & 63;
  // XCD-aware channel groups: workgroups are dealt to the 8 XCDs round-robin,
  // so group g = (bid mod 8) * (n/8) + bid/8 puts the two groups that share
  // each 128-byte line of a spectrum row on the same XCD (same L2)
  const int nwg = gridDim.x;
  const int bid = blockIdx.x;
  const int grp = (nwg % 8 == 0) ? (bid % 8) * (nwg / 8) + bid / 8 : bid;
  const int c0 = grp * CG;
  const int half = l >> 5, row = l & 31;
  const int sb0 = blockIdx.y * A.split_blocks;               // this row's blocks
  const int nb = min(A.nblocks - sb0, A.split_blocks);
  float2* const Vout = A.V + (long long)blockIdx.y * A.vsplit;

  f16_t accR[NTILE], accI[NTILE];
#pragma unroll
  for (int q = 0; q < NTILE; q++)
#pragma unroll
    for (int e = 0; e < 16; e++) { accR[q][e] = 0.0f; accI[q][e] = 0.0f; }

  for (int t0 = 0; t0 < nb; t0 += TT) {
    // ---- stage: NR inputs x TT blocks x CG channels, 16 bytes (4 channels) per item
    constexpr int NITEM = NR * TT * (CG / 4);
    for (int it = tid; it < NITEM; it += NTH) {
      constexpr int QN = CG / 4;      // 4-channel quads per staged row
      const int q = it % QN;
      const int t = (it / QN) % TT;
      const int a = it / (QN * TT);
      uint4 v = make_uint4(0u, 0u, 0u, 0u);
      if (a < A.nin && t0 + t < nb) {
        const long long e = A.in_off[a] + (long long)(sb0 + t0 + t) * A.block_stride + cidx(c0, A.gstride) + 4 * q;
        if ((e & 3) == 0) {
          v = *(const uint4*)(spec + e);
        } else {
          v.x = spec[e]; v.y = spec[e + 1]; v.z = spec[e + 2]; v.w = spec[e + 3];
        }
      }
      unsigned* d = lds + (4 * q) * CHS + a * TP + t;
      d[0] = v.x; d[CHS] = v.y; d[2 * CHS] = v.z; d[3 * CHS] = v.w;
    }
    __syncthreads();
    // ---- K steps of 8 blocks (16 real k values) each
    const unsigned* base = lds + w * CHS;
#pragma unroll
    for (int ks = 0; ks < TT / 8; ks++) {
      uint4 f[T], fh[T];
#pragma unroll
      for (int I = 0; I < T; I++) {
        f[I] = *(const uint4*)(base + (32 * I + row) * TP + 8 * ks + 4 * half);
        fh[I] = make_uint4(hat(f[I].x), hat(f[I].y), hat(f[I].z), hat(f[I].w));
      }
      int q = 0;
#pragma unroll
      for (int I = 0; I < T; I++)
#pragma unroll
        for (int J = I; J < T; J++, q++) {
          accR[q] = mfma<BF>(f[I], f[J], accR[q]);
          accI[q] = mfma<BF>(fh[I], f[J], accI[q]);
        }
    }
    __syncthreads();
  }

  // ---- write V[a][b][c] for this wavefront's channel
  const int c = c0 + w;
  if (c >= A.nchan) return;
  int q = 0;
#pragma unroll
  for (int I = 0; I < T; I++)
#pragma unroll
    for (int J = I; J < T; J++, q++) {
      const int b = 32 * J + row;
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int a = 32 * I + 8 * (r >> 2) + 4 * half + (r & 3);
        if (a >= A.nin || b >= A.nin) continue;
        float2* p = Vout + ((long long)a * A.nin + b) * A.nchan + c;
        float2 v = make_float2(accR[q][r], accI[q][r]);
        if (A.accumulate) { float2 o = *p; v.x += o.x; v.y += o.y; }
        *p = v;
      }
    }
}

// ---------------------------------------------------------------------------
// Streaming variant: the spectra go straight from HBM into LDS with gfx950's
// global_load_lds_dwordx4 (no registers) through a ring of NBUF stages of
// 8 blocks, NBUF - 1 of them in flight while the matrix cores work on the
// current one: one workgroup per CU keeps ~100 KB of loads outstanding, what
// HBM's latency needs for full rate (one stage in flight held it near half).
// LDS per stage: [row][block][CG channels], one 16 CG-byte row per (input,
// block), rows padded by 4 dwords.  One load instruction (2 CG lanes) fills
// the 8 blocks of one row; rows past the last block or past the last input
// read a 16-byte zero buffer.
//
// More than 64 inputs: the inputs form chunks of 64, and grid z runs over
// chunk pairs.  A diagonal workgroup (CROSS false, T = 2) loads chunk
// p = blockIdx.z as its 64 rows and accumulates that chunk's tiles I <= J; a
// cross workgroup (CROSS true, T = 4) loads chunk p as rows 0-63 and chunk
// q > p as rows 64-127, the z-th pair in (p, q) order, and accumulates the
// four 32x32 tiles between them.
// ---------------------------------------------------------------------------
// 16-byte load lanes per row and stage: 8 blocks of CG channels
constexpr int TT_LANES(int cg) { return 8 * cg / 4; }

template <int T, bool CROSS>
struct DlTiles {
  static constexpr int N = CROSS ? 4 : T * (T + 1) / 2;
  // tile pair q: rows of tile I, columns of tile J
  static constexpr __host__ __device__ int I(int q) {
    if (CROSS) return q >> 1;
    int i = 0;
    while (q >= T - i) { q -= T - i; i++; }
    return i;
  }
  static constexpr __host__ __device__ int J(int q) {
    if (CROSS) return 2 + (q & 1);
    int i = 0;
    while (q >= T - i) { q -= T - i; i++; }
    return i + q;
  }
};

template <int T, bool BF, bool CROSS>
__global__ __launch_bounds__(CROSS ? 512 : 1024) void k_xmac_mfma_dl(XmacArgs A, const uint4* __restrict__ zero16) {
  static_assert(!CROSS || T == 4, "a cross workgroup holds two 64-input chunks");
  // channels (wavefronts) per workgroup: a cross wavefront's 4 tiles of
  // accumulators need the registers of 2 wavefronts per SIMD
  constexpr int CG = CROSS ? 8 : 16;
  constexpr int LLANES = TT_LANES(CG);   // load lanes: 16-byte pieces of one row-stage
  constexpr int TT = 8;                  // blocks per stage: one K step of the matrix cores
  constexpr int NR = 32 * T;             // rows, padded to the tile size
  static_assert(LLANES == TT * CG / 4, "one load lane per 16-byte piece");
  using TL = DlTiles<T, CROSS>;
  constexpr int NTILE = TL::N;
  constexpr int SDW = TT * CG + 4;       // dwords per row per stage
  constexpr int STAGE = NR * SDW;        // dwords per stage
  constexpr int NBUF = T == 1 ? 8 : 4;   // ring of stages: 132-136 KB of LDS
  constexpr int LPW = NR / CG;           // load instructions per wave per stage
  // s_waitcnt with vmcnt = (NBUF - 2) LPW (the later stages may stay in
  // flight), expcnt and lgkmcnt left at their maxima
  constexpr int VMW = (NBUF - 2) * LPW;
  constexpr int WAITIMM = (VMW & 15) | (((VMW >> 4) & 3) << 14) | (7 << 4) | (15 << 8);
  __shared__ __align__(16) unsigned lds[NBUF * STAGE];

  const char* spec = (const char*)A.spec;
  const int tid = threadIdx.x;
  const int w = tid >> 6, l = tid & 63;
  const int nwg = gridDim.x;
  const int bid = blockIdx.x;
  const int grp = (nwg % 8 == 0) ? (bid % 8) * (nwg / 8) + bid / 8 : bid;   // XCD-aware groups
  const int c0 = grp * CG;
  const int half = l >> 5, row = l & 31;
  const int lrow = l / (CG / 4), lq = l % (CG / 4);   // load lane: block row and 4-channel quad
  const int sb0 = blockIdx.y * A.split_blocks;               // this row's blocks
  const int nb = min(A.nblocks - sb0, A.split_blocks);
  float2* const Vout = A.V + (long long)blockIdx.y * A.vsplit;
  // the two input chunks of this workgroup's rows (rows 0-63: chunk p,
  // rows 64-127: chunk q)
  int p = blockIdx.z, q = blockIdx.z;
  if (CROSS) {
    const int nch = (A.nin + 63) >> 6;
    int rem = blockIdx.z;
    p = 0;
    while (rem >= nch - 1 - p) { rem -= nch - 1 - p; p++; }
    q = p + 1 + rem;
  }
  auto input_of = [&](int r) { return r < 64 ? 64 * p + r : 64 * q + r - 64; };
  // the rows this wavefront loads, with their offsets read once into scalar
  // registers (a vector load here would make every later wait for it also
  // wait for the loads in flight)
  const int wu = __builtin_amdgcn_readfirstlane(w);
  long long ioff[LPW];
#pragma unroll
  for (int u = 0; u < LPW; u++) {
    const int a = input_of(wu * LPW + u);
    ioff[u] = a < A.nin ? A.in_off[a] + cidx(c0, A.gstride) : -1;
  }

  // every stage issues LPW loads per wavefront, also past the last block
  // (zeros), so that the wait counts below are the same for every stage
  auto issue = [&](int stage, int buf) {
    if (l < LLANES) {
      const int t = stage * TT + lrow;
#pragma unroll
      for (int u = 0; u < LPW; u++) {
        const void* src = zero16;
        if (ioff[u] >= 0 && t < nb)
          src = spec + 4 * (ioff[u] + (long long)(sb0 + t) * A.block_stride + 4 * lq);
        __builtin_amdgcn_global_load_lds(
            src, (__attribute__((address_space(3))) void*)(lds + buf * STAGE + (wu * LPW + u) * SDW), 16, 0, 0);
      }
    }
  };
  // workgroup barrier without __syncthreads' memory fence, which would wait
  // for every load in flight; the empty asm statements keep the compiler
  // from moving LDS accesses across it
  auto barrier = [] {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
  };

  f16_t accR[NTILE], accI[NTILE];
#pragma unroll
  for (int k = 0; k < NTILE; k++)
#pragma unroll
    for (int e = 0; e < 16; e++) { accR[k][e] = 0.0f; accI[k][e] = 0.0f; }

  const int nstage = (nb + TT - 1) / TT;
#pragma unroll
  for (int st = 0; st < NBUF - 1; st++) issue(st, st);
  __builtin_amdgcn_s_waitcnt(WAITIMM);   // stage 0 has landed
  barrier();
  for (int sg = 0; sg < nstage; sg++) {
    // the buffer the previous iteration read (everyone passed its barrier)
    issue(sg + NBUF - 1, (sg + NBUF - 1) % NBUF);
    const unsigned* base = lds + (sg % NBUF) * STAGE + w;      // this wavefront's channel
    uint4 f[T], fh[T];
#pragma unroll
    for (int I = 0; I < T; I++) {
      const unsigned* ptr = base + (32 * I + row) * SDW + 4 * half * CG;
      f[I] = make_uint4(ptr[0], ptr[CG], ptr[2 * CG], ptr[3 * CG]);
      fh[I] = make_uint4(hat(f[I].x), hat(f[I].y), hat(f[I].z), hat(f[I].w));
    }
#pragma unroll
    for (int k = 0; k < NTILE; k++) {
      accR[k] = mfma<BF>(f[TL::I(k)], f[TL::J(k)], accR[k]);
      accI[k] = mfma<BF>(fh[TL::I(k)], f[TL::J(k)], accI[k]);
    }
    __builtin_amdgcn_s_waitcnt(WAITIMM);   // this wavefront's loads of stage sg + 1 have landed
    barrier();                             // everyone's have, and nobody still reads stage sg
  }
  __builtin_amdgcn_s_waitcnt(0);           // the zero-filled stages past the end, before exit

  const int c = c0 + w;
  if (c >= A.nchan) return;
#pragma unroll
  for (int k = 0; k < NTILE; k++) {
    const int b = input_of(32 * TL::J(k) + row);
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int a = input_of(32 * TL::I(k) + 8 * (r >> 2) + 4 * half + (r & 3));
      if (a >= A.nin || b >= A.nin) continue;
      float2* ptr = Vout + ((long long)a * A.nin + b) * A.nchan + c;
      float2 v = make_float2(accR[k][r], accI[k][r]);
      if (A.accumulate) { float2 o = *ptr; v.x += o.x; v.y += o.y; }
      *ptr = v;
    }
  }
}

// ---------------------------------------------------------------------------
// FP32 on the matrix cores: v_mfma_f32_32x32x2_f32 takes one block per
// instruction (K = 2: lanes 0-31 give the real part of their row's input,
// lanes 32-63 the imaginary part), so with the same tiles as above
//     Re V += A A^T, Im V += Ah A^T,  Ah = (im, -re)
// in FP32 products and FP32 accumulation.  The spectra are row-major
// [band][block][channel] float2; a stage holds 8 blocks of CG channels of
// every row in LDS ([row][block][CG float2], rows padded by 4 dwords),
// filled by global_load_lds (one 16-byte piece, two channels, per lane) and
// double-buffered: the matrix cores need ~5 us per stage, the loads of the
// next one far less.  Rows and chunk pairs as in k_xmac_mfma_dl.
// ---------------------------------------------------------------------------
typedef float f32x16_t __attribute__((ext_vector_type(16)));

template <int T, bool CROSS>
__global__ __launch_bounds__(CROSS ? 512 : 1024) void k_xmac_mfma_f32(XmacArgs A, const uint4* __restrict__ zero16) {
  static_assert(!CROSS || T == 4, "a cross workgroup holds two 64-input chunks");
  constexpr int CG = CROSS ? 8 : 16;     // channels (wavefronts) per workgroup
  constexpr int TT = 8;                  // blocks per stage
  constexpr int NR = 32 * T;
  using TL = DlTiles<T, CROSS>;
  constexpr int NTILE = TL::N;
  constexpr int RBW = 2 * CG;            // dwords per (row, block)
  constexpr int SDW = TT * RBW + 4;      // dwords per row per stage
  constexpr int STAGE = NR * SDW;
  constexpr int LLANES = TT * CG / 2;    // 16-byte pieces per row-stage
  constexpr int LPW = NR / CG;           // rows each wavefront loads
  __shared__ __align__(16) unsigned lds[2 * STAGE];

  const char* spec = (const char*)A.spec;
  const int tid = threadIdx.x;
  const int w = tid >> 6, l = tid & 63;
  const int nwg = gridDim.x;
  const int bid = blockIdx.x;
  const int grp = (nwg % 8 == 0) ? (bid % 8) * (nwg / 8) + bid / 8 : bid;   // XCD-aware groups
  const int c0 = grp * CG;
  const int half = l >> 5, row = l & 31;
  const int lt = l / (CG / 2), lp = l % (CG / 2);   // load lane: block and 2-channel piece
  const int sb0 = blockIdx.y * A.split_blocks;
  const int nb = min(A.nblocks - sb0, A.split_blocks);
  float2* const Vout = A.V + (long long)blockIdx.y * A.vsplit;
  int p = blockIdx.z, q = blockIdx.z;
  if (CROSS) {
    const int nch = (A.nin + 63) >> 6;
    int rem = blockIdx.z;
    p = 0;
    while (rem >= nch - 1 - p) { rem -= nch - 1 - p; p++; }
    q = p + 1 + rem;
  }
  auto input_of = [&](int r) { return r < 64 ? 64 * p + r : 64 * q + r - 64; };
  const int wu = __builtin_amdgcn_readfirstlane(w);
  long long ioff[LPW];
#pragma unroll
  for (int u = 0; u < LPW; u++) {
    const int a = input_of(wu * LPW + u);
    ioff[u] = a < A.nin ? A.in_off[a] + c0 : -1;
  }
  auto issue = [&](int stage, int buf) {
    if (l < LLANES) {
      const int t = stage * TT + lt;
#pragma unroll
      for (int u = 0; u < LPW; u++) {
        const void* src = zero16;
        if (ioff[u] >= 0 && t < nb) src = spec + 8 * (ioff[u] + (long long)(sb0 + t) * A.block_stride + 2 * lp);
        __builtin_amdgcn_global_load_lds(
            src, (__attribute__((address_space(3))) void*)(lds + buf * STAGE + (wu * LPW + u) * SDW), 16, 0, 0);
      }
    }
  };
  auto barrier = [] {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
  };

  f32x16_t accR[NTILE], accI[NTILE];
#pragma unroll
  for (int k = 0; k < NTILE; k++)
#pragma unroll
    for (int e = 0; e < 16; e++) { accR[k][e] = 0.0f; accI[k][e] = 0.0f; }

  const int nstage = (nb + TT - 1) / TT;
  issue(0, 0);
  __builtin_amdgcn_s_waitcnt(0);
  barrier();
  for (int sg = 0; sg < nstage; sg++) {
    const int buf = sg & 1;
    if (sg + 1 < nstage) issue(sg + 1, buf ^ 1);
    const float* base = (const float*)(lds + buf * STAGE) + 2 * w;   // this wavefront's channel
    const int nt = min(TT, nb - sg * TT);
    for (int t = 0; t < nt; t++) {
      float f[T], fh[T];
#pragma unroll
      for (int I = 0; I < T; I++) {
        const float* ptr = base + (32 * I + row) * SDW + t * RBW;
        f[I] = ptr[half];                               // re (lanes 0-31), im (32-63)
        const float o = ptr[1 - half];
        fh[I] = half ? -o : o;                          // im, -re
      }
#pragma unroll
      for (int k = 0; k < NTILE; k++) {
        accR[k] = __builtin_amdgcn_mfma_f32_32x32x2f32(f[TL::I(k)], f[TL::J(k)], accR[k], 0, 0, 0);
        accI[k] = __builtin_amdgcn_mfma_f32_32x32x2f32(fh[TL::I(k)], f[TL::J(k)], accI[k], 0, 0, 0);
      }
    }
    __builtin_amdgcn_s_waitcnt(0);       // this wavefront's loads of the next stage have landed
    barrier();                           // everyone's have, and nobody still reads this stage
  }

  const int c = c0 + w;
  if (c >= A.nchan) return;
#pragma unroll
  for (int k = 0; k < NTILE; k++) {
    const int b = input_of(32 * TL::J(k) + row);
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int a = input_of(32 * TL::I(k) + 8 * (r >> 2) + 4 * half + (r & 3));
      if (a >= A.nin || b >= A.nin) continue;
      float2* ptr = Vout + ((long long)a * A.nin + b) * A.nchan + c;
      float2 v = make_float2(accR[k][r], accI[k][r]);
      if (A.accumulate) { float2 o = *ptr; v.x += o.x; v.y += o.y; }
      *ptr = v;
    }
  }
}

// ---------------------------------------------------------------------------
// FP8 variant: spectra stored as e4m3 (re, im) byte pairs, each (band,
// channel) scaled by 2^e in the F stage; v_mfma_scale_f32_32x32x64_f8f6f4
// takes 32 blocks per step (K = 64) and applies 2^-e per input row and
// column through its E8M0 block scales (one per lane: the lane's 32 values
// are one input at one channel), so the accumulators hold the unscaled
// products.  Staging goes through registers: two blocks of eight channels
// per item become one dword (two pairs) per channel in LDS, [channel][input]
// [block pair], so each lane's operand is two 16-byte LDS reads.
// ---------------------------------------------------------------------------
typedef int i8v_t __attribute__((ext_vector_type(8)));

// (re, im) -> (im, -re) on each packed fp8 pair (sign = bit 7 of the byte)
__device__ __forceinline__ unsigned hat8(unsigned v) {
  return (((v >> 8) & 0x00FF00FFu) | ((v << 8) & 0xFF00FF00u)) ^ 0x80008000u;
}

__device__ __forceinline__ f16_t mfma_fp8(const uint4 (&a)[2], const uint4 (&b)[2], f16_t c, int sa, int sb) {
  const i8v_t av = {(int)a[0].x, (int)a[0].y, (int)a[0].z, (int)a[0].w, (int)a[1].x, (int)a[1].y, (int)a[1].z, (int)a[1].w};
  const i8v_t bv = {(int)b[0].x, (int)b[0].y, (int)b[0].z, (int)b[0].w, (int)b[1].x, (int)b[1].y, (int)b[1].z, (int)b[1].w};
  return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(av, bv, c, 0, 0, 0, sa, 0, sb);
}

// eight consecutive (re, im) pairs from element e
__device__ __forceinline__ uint4 load8pairs(const unsigned short* p, long long e) {
  if ((e & 7) == 0) return *(const uint4*)(p + e);
  uint4 v;
  v.x = p[e] | ((unsigned)p[e + 1] << 16);
  v.y = p[e + 2] | ((unsigned)p[e + 3] << 16);
  v.z = p[e + 4] | ((unsigned)p[e + 5] << 16);
  v.w = p[e + 6] | ((unsigned)p[e + 7] << 16);
  return v;
}

template <int T, int CG>
__global__ __launch_bounds__(64 * CG) void k_xmac_fp8(XmacArgs A) {
  // CG channels (wavefronts) per workgroup: 16 for one 32-input tile, 8 for
  // two (the accumulators and operands of a wavefront need 128+ registers)
  constexpr int TT = 32;                 // blocks per K step
  constexpr int NR = 32 * T;             // inputs, padded to the tile size
  constexpr int NTILE = T * (T + 1) / 2;
  constexpr int TP = TT / 2 + 4;         // dwords per input row (16 block pairs + pad)
  constexpr int CHS = NR * TP + 4;       // dwords per channel
  __shared__ __align__(16) unsigned lds[CG * CHS];

  const unsigned short* spec = (const unsigned short*)A.spec;
  const int tid = threadIdx.x;
  const int w = tid >> 6, l = tid & 63;
  const int nwg = gridDim.x;
  const int bid = blockIdx.x;
  const int grp = (nwg % 8 == 0) ? (bid % 8) * (nwg / 8) + bid / 8 : bid;   // XCD-aware groups
  const int c0 = grp * CG;
  const int c = c0 + w;                  // this wavefront's channel
  const int half = l >> 5, row = l & 31;
  const int sb0 = blockIdx.y * A.split_blocks;
  const int nb = min(A.nblocks - sb0, A.split_blocks);
  float2* const Vout = A.V + (long long)blockIdx.y * A.vsplit;

  // E8M0 scale of this lane's row (input 32 I + row) at channel c: 2^-e
  int sc[T];
#pragma unroll
  for (int I = 0; I < T; I++) {
    const int a = 32 * I + row;
    sc[I] = a < A.nin ? 127 - (int)A.cexp[A.eoff[a] + c] : 127;
  }
  f16_t accR[NTILE], accI[NTILE];
#pragma unroll
  for (int q = 0; q < NTILE; q++)
#pragma unroll
    for (int e = 0; e < 16; e++) { accR[q][e] = 0.0f; accI[q][e] = 0.0f; }

  for (int t0 = 0; t0 < nb; t0 += TT) {
    // ---- stage: items (input, block pair, channel octet)
    constexpr int QN = CG / 8;           // channel octets
    constexpr int NITEM = NR * (TT / 2) * QN;
    for (int it = tid; it < NITEM; it += 64 * CG) {
      const int q = it % QN;
      const int p = (it / QN) % (TT / 2);
      const int a = it / (QN * (TT / 2));
      uint4 v0 = make_uint4(0u, 0u, 0u, 0u), v1 = v0;
      if (a < A.nin) {
        const long long e0 = A.in_off[a] + (long long)(sb0 + t0 + 2 * p) * A.block_stride + cidx(c0 + 8 * q, A.gstride);
        if (t0 + 2 * p < nb) v0 = load8pairs(spec, e0);
        if (t0 + 2 * p + 1 < nb) v1 = load8pairs(spec, e0 + A.block_stride);
      }
      const unsigned u0[4] = {v0.x, v0.y, v0.z, v0.w}, u1[4] = {v1.x, v1.y, v1.z, v1.w};
      unsigned* d = lds + (8 * q) * CHS + a * TP + p;
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const unsigned lo = (u0[k >> 1] >> (16 * (k & 1))) & 0xFFFFu;
        const unsigned hi = (u1[k >> 1] >> (16 * (k & 1))) & 0xFFFFu;
        d[k * CHS] = lo | (hi << 16);
      }
    }
    __syncthreads();
    // ---- one K step: lane (row, half) holds blocks 16 half .. 16 half + 15
    const unsigned* base = lds + w * CHS;
    uint4 f[T][2];
#pragma unroll
    for (int I = 0; I < T; I++) {
      const unsigned* rp = base + (32 * I + row) * TP + 8 * half;
      f[I][0] = *(const uint4*)rp;
      f[I][1] = *(const uint4*)(rp + 4);
    }
    int q = 0;
#pragma unroll
    for (int I = 0; I < T; I++) {
      uint4 fh[2];
#pragma unroll
      for (int h = 0; h < 2; h++)
        fh[h] = make_uint4(hat8(f[I][h].x), hat8(f[I][h].y), hat8(f[I][h].z), hat8(f[I][h].w));
#pragma unroll
      for (int J = I; J < T; J++, q++) {
        accR[q] = mfma_fp8(f[I], f[J], accR[q], sc[I], sc[J]);
        accI[q] = mfma_fp8(fh, f[J], accI[q], sc[I], sc[J]);
      }
    }
    __syncthreads();
  }

  if (c >= A.nchan) return;
  int q = 0;
#pragma unroll
  for (int I = 0; I < T; I++)
#pragma unroll
    for (int J = I; J < T; J++, q++) {
      const int b = 32 * J + row;
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int a = 32 * I + 8 * (r >> 2) + 4 * half + (r & 3);
        if (a >= A.nin || b >= A.nin) continue;
        float2* p = Vout + ((long long)a * A.nin + b) * A.nchan + c;
        float2 v = make_float2(accR[q][r], accI[q][r]);
        if (A.accumulate) { float2 o = *p; v.x += o.x; v.y += o.y; }
        *p = v;
      }
    }
}

hipError_t launch_xmac_fp8(const XmacArgs& a, hipStream_t stream) {
  if (!(a.nin <= 64 && a.nchan % 16 == 0) || !a.cexp || !a.eoff) return hipErrorInvalidValue;
  const unsigned rows = (unsigned)((a.nblocks + a.split_blocks - 1) / a.split_blocks);
  if (a.nin <= 32)
    hipLaunchKernelGGL((k_xmac_fp8<1, 16>), dim3((unsigned)(a.nchan / 16), rows), dim3(1024), 0, stream, a);
  else
    hipLaunchKernelGGL((k_xmac_fp8<2, 8>), dim3((unsigned)(a.nchan / 8), rows), dim3(512), 0, stream, a);
  return hipGetLastError();
}

bool xmac_mfma_supported(int nin, int nchan) { return nin <= 64 && nchan % 16 == 0; }
bool xmac_mfma_streaming_supported(int nin, int nchan) { return nin <= kXmacMfmaMaxInputs && nchan % 16 == 0; }

bool xmac_mfma_streaming_ok(const XmacArgs& a, const long long* in_off_host) {
  if (a.block_stride % 4) return false;
  for (int i = 0; i < a.nin; i++)
    if (in_off_host[i] % 4) return false;      // 16-byte aligned rows
  return true;
}

hipError_t launch_xmac_mfma_streaming(const XmacArgs& a, int prec, const void* zero16, hipStream_t stream) {
  if (!xmac_mfma_streaming_supported(a.nin, a.nchan)) return hipErrorInvalidValue;
  const unsigned gx = (unsigned)(a.nchan / 16), gy = (unsigned)((a.nblocks + a.split_blocks - 1) / a.split_blocks);
  const bool bf = prec == FXG_XMAC_BF16;
  const uint4* z = (const uint4*)zero16;
#define FX_DL(T, NZ)                                                                                           \
  do {                                                                                                         \
    if (bf) hipLaunchKernelGGL((k_xmac_mfma_dl<T, true, false>), dim3(gx, gy, NZ), dim3(1024), 0, stream, a, z); \
    else hipLaunchKernelGGL((k_xmac_mfma_dl<T, false, false>), dim3(gx, gy, NZ), dim3(1024), 0, stream, a, z);   \
  } while (0)
  if (a.nin <= 32) {
    FX_DL(1, 1);
  } else if (a.nin <= 64) {
    FX_DL(2, 1);
  } else {
    // every chunk of 64 inputs with itself, then every pair of chunks (8
    // channels per workgroup)
    const unsigned nch = (unsigned)((a.nin + 63) / 64);
    FX_DL(2, nch);
    const dim3 gc(2 * gx, gy, nch * (nch - 1) / 2);
    if (bf) hipLaunchKernelGGL((k_xmac_mfma_dl<4, true, true>), gc, dim3(512), 0, stream, a, z);
    else hipLaunchKernelGGL((k_xmac_mfma_dl<4, false, true>), gc, dim3(512), 0, stream, a, z);
  }
#undef FX_DL
  return hipGetLastError();
}

bool xmac_mfma_f32_ok(const XmacArgs& a, const long long* in_off_host) {
  if (a.nin > kXmacMfmaMaxInputs || a.nchan % 16 || a.block_stride % 2) return false;
  for (int i = 0; i < a.nin; i++)
    if (in_off_host[i] % 2) return false;      // 16-byte aligned channel pairs
  return true;
}

hipError_t launch_xmac_mfma_f32(const XmacArgs& a, const void* zero16, hipStream_t stream) {
  const unsigned gx = (unsigned)(a.nchan / 16), gy = (unsigned)((a.nblocks + a.split_blocks - 1) / a.split_blocks);
  const uint4* z = (const uint4*)zero16;
  if (a.nin <= 32) {
    hipLaunchKernelGGL((k_xmac_mfma_f32<1, false>), dim3(gx, gy), dim3(1024), 0, stream, a, z);
  } else if (a.nin <= 64) {
    hipLaunchKernelGGL((k_xmac_mfma_f32<2, false>), dim3(gx, gy), dim3(1024), 0, stream, a, z);
  } else {
    const unsigned nch = (unsigned)((a.nin + 63) / 64);
    hipLaunchKernelGGL((k_xmac_mfma_f32<2, false>), dim3(gx, gy, nch), dim3(1024), 0, stream, a, z);
    hipLaunchKernelGGL((k_xmac_mfma_f32<4, true>), dim3(2 * gx, gy, nch * (nch - 1) / 2), dim3(512), 0, stream, a,
                       z);
  }
  return hipGetLastError();
}

hipError_t launch_xmac_mfma(const XmacArgs& a, int prec, hipStream_t stream) {
  if (!xmac_mfma_supported(a.nin, a.nchan)) return hipErrorInvalidValue;
  constexpr int CG = 16;      // 64-byte row segments; one workgroup (148 KB of LDS) per CU
  dim3 grid((unsigned)(a.nchan / CG), (unsigned)((a.nblocks + a.split_blocks - 1) / a.split_blocks));
  const bool bf = prec == FXG_XMAC_BF16;
  if (a.nin <= 32) {
    if (bf) hipLaunchKernelGGL((k_xmac_mfma<1, true, CG>), grid, dim3(64 * CG), 0, stream, a);
    else hipLaunchKernelGGL((k_xmac_mfma<1, false, CG>), grid, dim3(64 * CG), 0, stream, a);
  } else {
    if (bf) hipLaunchKernelGGL((k_xmac_mfma<2, true, CG>), grid, dim3(64 * CG), 0, stream, a);
    else hipLaunchKernelGGL((k_xmac_mfma<2, false, CG>), grid, dim3(64 * CG), 0, stream, a);
  }
  return hipGetLastError();
}

}  // namespace fxg
