// Bulk-vertex sweep pipeline for bond dimension 32 on MFMA.
//
// For a site tensor psi[s, b1, b2, b3, b4] (d = 2, every chi = 32) the four
// outgoing BP messages (updated_message, src/caches/abstractbeliefpropagationcache.jl:225-239)
// are computed in two passes that share the absorbed pair:
//
//   pass A:  P[s, b1', b2', b3, b4] = sum_{b1,b2} M1[b1,b1'] M2[b2,b2'] psi[s,b1,b2,b3,b4]
//            T3 = sum_r P_r M4 psi_r^T,   T4 = sum_r P_r^T M3 psi_r      (r = (s, b1', b2'))
//   pass B:  the same with (3,4) absorbed and (1,2) open  ->  T1, T2
//
// where P_r and psi_r are the 32x32 (b3, b4) matrices of block r.  That is 12
// 32^3-GEMM units per vertex instead of the 16 of four independent messages.
// psi is kept in two block layouts (every 32x32 block contiguous), built once
// per site-tensor upload:
//   layout A: block (s + 2 b1 + 64 b2) holds [b3][b4];  layout B: block (s + 2 b3 + 64 b4) holds [b1][b2].
// "absorb" reads layout B (A) blocks, forms M1^T X M2 for each, and scatters them
// into P in layout A (B) through an LDS transpose; "fused" streams P and psi
// blocks of one layout and accumulates both messages in MFMA accumulators,
// feeding each intermediate accumulator straight back as the next B operand
// (its row order is the K order the other operand is loaded with).
//
// BF16 flavour: v_mfma_f32_32x32x16_bf16 with FP32 accumulation; psi is exact
// BF16, messages and intermediates are split into hi+lo BF16 terms
// (hi*hi + hi*lo + lo*hi).  FP64 flavour: v_mfma_f64_16x16x4_f64 on LDS tiles.
#include <algorithm>
#include <type_traits>
#include <cstdlib>
#include <vector>

#include "itnbp_ops.hpp"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int FC = 32;            // bond dimension of the pipeline
constexpr int FBLK = FC * FC;     // elements per block
constexpr int FNR = 2 * FC * FC;  // blocks per vertex (d = 2)
constexpr int FUSED_R_PER_WG = 128;
constexpr int FUSED_CHUNKS = FNR / FUSED_R_PER_WG;  // 16

struct FastItem {
  long long psi_in;   // absorb input blocks (element offset into the layout buffer)
  long long psi_out;  // fused blocks
  long long P;        // intermediate
  long long Ma, Mb;   // absorbed messages (rows, cols of the input blocks)
  long long Mz, Mw;   // open-axis messages (rows, cols of the fused blocks)
  long long out_z, out_w;
  int slot_z, slot_w;
  long long part;     // partial sums offset
  int site;           // index of the site's power-of-two scale (BF16/FP16 layouts)
};

__device__ inline void split_bf16(float x, __bf16& hi, __bf16& lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
}

// K slot (t, g, p) of a 32x32x16 operand <-> row of a 32x32 accumulator
__device__ inline int kperm(int t, int g, int p) { return 16 * t + 8 * (p >> 2) + 4 * g + (p & 3); }

#define MFMA_BF16(a, b, c) __builtin_amdgcn_mfma_f32_32x32x16_bf16((a), (b), (c), 0, 0, 0)

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

// 16-bit operand types of the 32x32x16 MFMAs (identical lane layouts): BF16
// (8 exponent bits) and FP16 (11-bit significand, 8x finer rounding)
template <typename T>
struct Mma;
template <>
struct Mma<__bf16> {
  typedef bf16x8 v8;
  static __device__ inline f32x16 mfma(v8 a, v8 b, f32x16 c) { return MFMA_BF16(a, b, c); }
};
template <>
struct Mma<_Float16> {
  typedef f16x8 v8;
  static __device__ inline f32x16 mfma(v8 a, v8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
  }
};

template <typename T>
__device__ inline void split16(float x, T& hi, T& lo) {
  hi = (T)x;
  lo = (T)(x - (float)hi);
}

// ---------------------------------------------------------------------------
// layout build: natural column-major psi -> layouts A and B
// ---------------------------------------------------------------------------
// Per-site power-of-two scales of the 16-bit layouts.  Site j's largest |psi|
// becomes m in [1/2, 1) after multiplying by 2^-e[j], so that the FP16
// intermediates (largest normal 65504, subnormals below 6.1e-5) hold the
// rotated blocks at every scale of psi; a Gram built from the scaled site is
// 2^-2e[j] times the true one and the finalisers multiply by 2^2e[j].
// k_site_absmax: grid (64, sites), partial maxima pmax[64 j + x];
// k_site_exp: e[j] from the 64 partials (0 for an all-zero or non-finite site).
template <typename TS>
__global__ __launch_bounds__(256) void k_site_absmax(const TS* __restrict__ psi, const long long* __restrict__ src_off,
                                                     long long nelem, double* __restrict__ pmax) {
  __shared__ double red[4];
  const TS* src = psi + src_off[blockIdx.y];
  double m = 0;
  for (long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x; n < nelem; n += (long long)gridDim.x * blockDim.x) {
    const double a = fabs(load_as<double>(src, n));
    m = (a > m || a != a) ? a : m;  // a NaN sticks
  }
  for (int k = 32; k > 0; k >>= 1) {
    const double o = __shfl_xor(m, k);
    m = (o > m || o != o) ? o : m;
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = (red[w] > m || red[w] != red[w]) ? red[w] : m;
    pmax[(long long)blockIdx.y * 64 + blockIdx.x] = m;
  }
}

__global__ void k_site_exp(const double* __restrict__ pmax, int n, int* __restrict__ sexp) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double m = 0;
  for (int x = 0; x < 64; ++x) {
    const double a = pmax[(long long)j * 64 + x];
    m = (a > m || a != a) ? a : m;
  }
  int e = 0;
  if (m > 0 && isfinite(m)) frexp(m, &e);
  sexp[j] = e;
}

__device__ inline double site_scale(const int* sexp, int j) { return sexp ? ldexp(1.0, -sexp[j]) : 1.0; }
__device__ inline double site_unscale2(const int* sexp, int j) { return sexp ? ldexp(1.0, 2 * sexp[j]) : 1.0; }

// Layout builds, coalesced through LDS.  A site of degree 4 (3) is the
// column-major matrix [2048 rows r = s + 2 b1 + 64 b2] x [1024 (32) columns,
// col = b3 + 32 b4 (b3)].
// k_lay_cols: block (s + 2 col) = [b1][b2] of every column (layout B of a
//   degree-4 site, layout C of a degree-3 site); grid (columns, sites).
// k_lay_rows: layout A, A[r][b3 32 + b4] = psi[r][b3 + 32 b4]; WG -> (b3, 64 rows),
//   grid (32 x 32, sites).
// Reads run down columns (64 or 2048 consecutive elements), writes along
// rows of 32 or 1024 consecutive elements.
template <typename TO>
__device__ inline TO to_out(double v) {
  if constexpr (std::is_same<TO, double>::value) return v;
  else return (TO)(float)v;
}

template <typename TS, typename TO>
__global__ __launch_bounds__(256) void k_lay_cols(const TS* __restrict__ psi, const long long* __restrict__ src_off,
                                                  const long long* __restrict__ dst_off, TO* __restrict__ out,
                                                  const int* __restrict__ sexp) {
  __shared__ double T[2][FC][FC + 1];  // [s][b1][b2]
  const int it = blockIdx.y, col = blockIdx.x;
  const TS* src = psi + src_off[it] + (long long)col * (2 * FBLK);
  TO* dst = out + dst_off[it] + (long long)(2 * col) * FBLK;
  const double sc = site_scale(sexp, it);
  for (int r = threadIdx.x; r < 2 * FBLK; r += blockDim.x)
    T[r & 1][(r >> 1) & 31][r >> 6] = load_as<double>(src, r) * sc;
  __syncthreads();
  for (int o = threadIdx.x; o < 2 * FBLK; o += blockDim.x)
    dst[o] = to_out<TO>(T[o >> 10][(o >> 5) & 31][o & 31]);
}

template <typename TS, typename TO>
__global__ __launch_bounds__(256) void k_lay_rows(const TS* __restrict__ psi, const long long* __restrict__ src_off,
                                                  const long long* __restrict__ dst_off, TO* __restrict__ out,
                                                  const int* __restrict__ sexp) {
  __shared__ double T[64][FC + 1];  // [row in tile][b4]
  const int it = blockIdx.y;
  const int b3 = blockIdx.x & 31, r0 = (blockIdx.x >> 5) * 64;
  const TS* src = psi + src_off[it];
  TO* dst = out + dst_off[it];
  const double sc = site_scale(sexp, it);
  for (int k = threadIdx.x; k < 64 * FC; k += blockDim.x) {
    const int rl = k & 63, b4 = k >> 6;
    T[rl][b4] = load_as<double>(src, (long long)(r0 + rl) + (long long)(2 * FBLK) * (b3 + FC * b4)) * sc;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < 64 * FC; k += blockDim.x) {
    const int b4 = k & 31, rl = k >> 5;
    dst[(long long)(r0 + rl) * FBLK + b3 * FC + b4] = to_out<TO>(T[rl][b4]);
  }
}

// layouts A and B of bulk sites (offsets src_off, dst_off; B follows A at +bstride)
template <typename TS, typename TO>
static void launch_layout_ab(hipStream_t stream, const TS* psi, const long long* src_off, const long long* dst_off,
                             TO* A, TO* B, int n, const int* sexp) {
  hipLaunchKernelGGL((k_lay_rows<TS, TO>), dim3(32 * 32, n), dim3(256), 0, stream, psi, src_off, dst_off, A, sexp);
  hipLaunchKernelGGL((k_lay_cols<TS, TO>), dim3(FC * FC, n), dim3(256), 0, stream, psi, src_off, dst_off, B, sexp);
}

// ---------------------------------------------------------------------------
// BF16 absorb: P[(s, x', y')][z][w] = (Ma^T X_{s,z,w} Mb)[x'][y']
// grid (128, items): WG q -> s = q&1, z = (q>>1)&31, w in [16 (q>>6), +16)
// ---------------------------------------------------------------------------
template <int NT>
__global__ __launch_bounds__(256) void k_absorb_bf16(const FastItem* __restrict__ items,
                                                     const __bf16* __restrict__ lay,
                                                     const float* __restrict__ msg,
                                                     float* __restrict__ P) {
  __shared__ float S[16 * 1056];
  const FastItem it = items[blockIdx.y];
  const int q = blockIdx.x;
  const int s = q & 1, z = (q >> 1) & 31, w0 = (q >> 6) * 16;
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6, g = l >> 5, c = l & 31;
  bf16x8 bMbh[2], bMbl[2], aMah[2], aMal[2];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      __bf16 h, lo;
      split_bf16(msg[it.Mb + (16 * t + 8 * g + p) + FC * c], h, lo);  // Mb[y][y']
      bMbh[t][p] = h;
      bMbl[t][p] = lo;
      split_bf16(msg[it.Ma + kperm(t, g, p) + FC * c], h, lo);        // Ma[x][x'], K permuted
      aMah[t][p] = h;
      aMal[t][p] = lo;
    }
#pragma unroll 1
  for (int i = 0; i < 4; ++i) {
    const int wl = 4 * wv + i, w = w0 + wl;
    const __bf16* X = lay + it.psi_in + (long long)(s + 2 * (z + FC * w)) * FBLK;
    f32x16 V = {};
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const bf16x8 aX = *reinterpret_cast<const bf16x8*>(X + c * FC + 16 * t + 8 * g);
      V = MFMA_BF16(aX, bMbh[t], V);
      if (NT > 1) V = MFMA_BF16(aX, bMbl[t], V);
    }
    bf16x8 bVh[2], bVl[2];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int p = 0; p < 8; ++p) {
        __bf16 h, lo;
        split_bf16(V[8 * t + p], h, lo);
        bVh[t][p] = h;
        bVl[t][p] = lo;
      }
    f32x16 Y = {};
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      Y = MFMA_BF16(aMah[t], bVh[t], Y);
      if (NT > 1) {
        Y = MFMA_BF16(aMah[t], bVl[t], Y);
        Y = MFMA_BF16(aMal[t], bVh[t], Y);
      }
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int xb = 8 * (v >> 2) + 4 * g + (v & 3);
      S[wl * 1056 + xb + 33 * c] = Y[v];
    }
  }
  __syncthreads();
  for (int q2 = threadIdx.x; q2 < FBLK; q2 += blockDim.x) {
    const int xb = q2 & 31, yb = q2 >> 5;
    const long long r = s + 2 * (xb + FC * yb);
    float* dst = P + it.P + r * FBLK + z * FC + w0;
    const int o = xb + 33 * yb;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      float4 v4 = make_float4(S[(4 * j) * 1056 + o], S[(4 * j + 1) * 1056 + o], S[(4 * j + 2) * 1056 + o],
                              S[(4 * j + 3) * 1056 + o]);
      *reinterpret_cast<float4*>(dst + 4 * j) = v4;
    }
  }
}

// ---------------------------------------------------------------------------
// BF16 fused: Tz += P_r Mw psi_r^T, Tw += P_r^T Mz psi_r over 128 blocks per WG
// grid (16, items); partial[item][chunk][2][32*32], [x*32 + y]
// ---------------------------------------------------------------------------
template <int NT>
__global__ __launch_bounds__(256) void k_fused_bf16(const FastItem* __restrict__ items,
                                                    const __bf16* __restrict__ lay,
                                                    const float* __restrict__ msg,
                                                    const float* __restrict__ P,
                                                    float* __restrict__ part) {
  __shared__ float Pb[4][32 * 33];
  __shared__ __bf16 Sb[4][32 * 34];
  __shared__ float Red[2 * FBLK];
  const FastItem it = items[blockIdx.y];
  const int q = blockIdx.x;
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6, g = l >> 5, c = l & 31;
  bf16x8 aMwh[2], aMwl[2], aMzh[2], aMzl[2];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      const int k = 16 * t + 8 * g + p;
      __bf16 h, lo;
      split_bf16(msg[it.Mw + k + FC * c], h, lo);  // Mw[w][w'] with w = k, w' = c
      aMwh[t][p] = h;
      aMwl[t][p] = lo;
      split_bf16(msg[it.Mz + c + FC * k], h, lo);  // Mz[z][z'] with z = c, z' = k
      aMzh[t][p] = h;
      aMzl[t][p] = lo;
    }
  f32x16 accTzT = {}, accTw = {};
  float* pb = Pb[wv];
  __bf16* sb = Sb[wv];
#pragma unroll 1
  for (int i = 0; i < FUSED_R_PER_WG / 4; ++i) {
    const long long r = (long long)q * FUSED_R_PER_WG + wv * (FUSED_R_PER_WG / 4) + i;
    const float* Pr = P + it.P + r * FBLK;
    const __bf16* Yr = lay + it.psi_out + r * FBLK;
    {
      const int row = l >> 1, col0 = (l & 1) * 16;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float4 v4 = *reinterpret_cast<const float4*>(Pr + l * 16 + 4 * j);
        float* d = pb + row * 33 + col0 + 4 * j;
        d[0] = v4.x;
        d[1] = v4.y;
        d[2] = v4.z;
        d[3] = v4.w;
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const bf16x8 v8 = *reinterpret_cast<const bf16x8*>(Yr + l * 16 + 8 * j);
        __bf16* d = sb + row * 34 + col0 + 8 * j;
#pragma unroll
        for (int p = 0; p < 8; ++p) d[p] = v8[p];
      }
    }
    __syncthreads();
    // step 1: X^T = Mw^T P_r^T
    f32x16 XT = {};
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      bf16x8 bh, bl;
#pragma unroll
      for (int p = 0; p < 8; ++p) {
        __bf16 h, lo;
        split_bf16(pb[c * 33 + 16 * t + 8 * g + p], h, lo);
        bh[p] = h;
        bl[p] = lo;
      }
      XT = MFMA_BF16(aMwh[t], bh, XT);
      if (NT > 1) {
        XT = MFMA_BF16(aMwh[t], bl, XT);
        XT = MFMA_BF16(aMwl[t], bh, XT);
      }
    }
    // step 2: Tz^T += psi_r X^T
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      bf16x8 a, bh, bl;
#pragma unroll
      for (int p = 0; p < 8; ++p) {
        a[p] = sb[c * 34 + kperm(t, g, p)];
        __bf16 h, lo;
        split_bf16(XT[8 * t + p], h, lo);
        bh[p] = h;
        bl[p] = lo;
      }
      accTzT = MFMA_BF16(a, bh, accTzT);
      if (NT > 1) accTzT = MFMA_BF16(a, bl, accTzT);
    }
    // step 3: Zm = Mz psi_r
    f32x16 Zm = {};
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      bf16x8 b;
#pragma unroll
      for (int p = 0; p < 8; ++p) b[p] = sb[(16 * t + 8 * g + p) * 34 + c];
      Zm = MFMA_BF16(aMzh[t], b, Zm);
      if (NT > 1) Zm = MFMA_BF16(aMzl[t], b, Zm);
    }
    // step 4: Tw += P_r^T Zm
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      bf16x8 ah, al, bh, bl;
#pragma unroll
      for (int p = 0; p < 8; ++p) {
        __bf16 h, lo;
        split_bf16(pb[kperm(t, g, p) * 33 + c], h, lo);
        ah[p] = h;
        al[p] = lo;
        split_bf16(Zm[8 * t + p], h, lo);
        bh[p] = h;
        bl[p] = lo;
      }
      accTw = MFMA_BF16(ah, bh, accTw);
      if (NT > 1) {
        accTw = MFMA_BF16(ah, bl, accTw);
        accTw = MFMA_BF16(al, bh, accTw);
      }
    }
    __syncthreads();
  }
  // reduce the four waves
  for (int w = 0; w < 4; ++w) {
    if (wv == w) {
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int row = 8 * (v >> 2) + 4 * g + (v & 3);
        if (w == 0) {
          Red[c * FC + row] = accTzT[v];            // Tz[x = c][y = row]
          Red[FBLK + row * FC + c] = accTw[v];      // Tw[x = row][y = c]
        } else {
          Red[c * FC + row] += accTzT[v];
          Red[FBLK + row * FC + c] += accTw[v];
        }
      }
    }
    __syncthreads();
  }
  float* dst = part + it.part + (long long)q * 2 * FBLK;
  for (int o = threadIdx.x; o < 2 * FBLK; o += blockDim.x) dst[o] = Red[o];
}

// ---------------------------------------------------------------------------
// BF16 absorb, compact flavour: P stored as one BF16 term.  WG (512 threads)
// -> (s, z), all 32 w; the 32 outputs per r are one 64-byte run written by 4
// lanes cooperatively from an LDS stage [r][w].
// grid (64, items)
// ---------------------------------------------------------------------------
constexpr int ST_LD = 40;  // LDS row stride (bf16) of the staging/tiles: 80 bytes, 16-byte aligned

// Work index of workgroup b of a 64-wide grid row.  Workgroups are dispatched
// round-robin over the 8 XCDs (chiplets, each with its own L2), so XCD x
// receives b = x, x + 8, ..., x + 56; they take the 8 consecutive work
// indices 8x .. 8x + 7.  Work indices 2z + s and 2(z + 1) + s write the two
// 64-byte halves of the same 128-byte lines, which then meet in one L2.
__device__ inline int xcd_local64(int b) { return ((b & 7) << 3) | (b >> 3); }

// TO: type of P and of the second product's operands (BF16 for the
// absorbed-pair passes, FP16 for rot12 of the eigenbasis pipeline); the first
// product multiplies the BF16 site tensor and stays in BF16.
template <typename TO>
__global__ __launch_bounds__(512) void k_absorb_bf16c(const FastItem* __restrict__ items,
                                                      const __bf16* __restrict__ lay,
                                                      const float* __restrict__ msg,
                                                      TO* __restrict__ P) {
  typedef typename Mma<TO>::v8 o8;
  __shared__ __attribute__((aligned(16))) TO S[FBLK * ST_LD];
  const FastItem it = items[blockIdx.y];
  const int q = xcd_local64(blockIdx.x);
  const int s = q & 1, z = q >> 1;
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6, g = l >> 5, c = l & 31;
  bf16x8 bMbh[2], bMbl[2];
  o8 aMah[2], aMal[2];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      __bf16 h, lo;
      split_bf16(msg[it.Mb + (16 * t + 8 * g + p) + FC * c], h, lo);
      bMbh[t][p] = h;
      bMbl[t][p] = lo;
      TO h2, lo2;
      split16(msg[it.Ma + kperm(t, g, p) + FC * c], h2, lo2);
      aMah[t][p] = h2;
      aMal[t][p] = lo2;
    }
  // the wave's four input blocks (8 KiB) are all requested before the first product
  bf16x8 aXs[4][2];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const __bf16* X = lay + it.psi_in + (long long)(s + 2 * (z + FC * (4 * wv + i))) * FBLK;
#pragma unroll
    for (int t = 0; t < 2; ++t) aXs[i][t] = *reinterpret_cast<const bf16x8*>(X + c * FC + 16 * t + 8 * g);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int w = 4 * wv + i;
    f32x16 V = {};
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      V = MFMA_BF16(aXs[i][t], bMbh[t], V);
      V = MFMA_BF16(aXs[i][t], bMbl[t], V);
    }
    o8 bV[2];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int p = 0; p < 8; ++p) bV[t][p] = (TO)V[8 * t + p];
    f32x16 Y = {};
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      Y = Mma<TO>::mfma(aMah[t], bV[t], Y);
      Y = Mma<TO>::mfma(aMal[t], bV[t], Y);
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int xb = 8 * (v >> 2) + 4 * g + (v & 3);
      S[(c + FC * xb) * ST_LD + w] = (TO)Y[v];  // row (y' + 32 x'), column w
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < FBLK * 4; k += blockDim.x) {
    const int o = k >> 2, part = k & 3;
    const int yb = o & 31, xb = o >> 5;
    const long long r = s + 2 * (xb + FC * yb);
    const o8 val = *reinterpret_cast<const o8*>(S + o * ST_LD + 8 * part);
    *reinterpret_cast<o8*>(P + it.P + r * FBLK + z * FC + 8 * part) = val;
  }
}

// ---------------------------------------------------------------------------
// BF16 fused, compact flavour: P and psi blocks in BF16, intermediates as one
// BF16 term, blocks double-buffered in LDS (prefetch of block i+1 overlaps the
// MFMAs of block i).  grid (16, items), 4 waves x 32 blocks
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_fused_bf16c(const FastItem* __restrict__ items,
                                                     const __bf16* __restrict__ lay,
                                                     const float* __restrict__ msg,
                                                     const __bf16* __restrict__ P,
                                                     float* __restrict__ part) {
  __shared__ __attribute__((aligned(16))) __bf16 Pb[2][4][FC * ST_LD];
  __shared__ __attribute__((aligned(16))) __bf16 Sb[2][4][FC * ST_LD];
  __shared__ float Red[2 * FBLK];
  const FastItem it = items[blockIdx.y];
  const int q = blockIdx.x;
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6, g = l >> 5, c = l & 31;
  bf16x8 aMwh[2], aMwl[2], aMzh[2], aMzl[2];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      const int k = 16 * t + 8 * g + p;
      __bf16 h, lo;
      split_bf16(msg[it.Mw + k + FC * c], h, lo);
      aMwh[t][p] = h;
      aMwl[t][p] = lo;
      split_bf16(msg[it.Mz + c + FC * k], h, lo);
      aMzh[t][p] = h;
      aMzl[t][p] = lo;
    }
  f32x16 accTzT = {}, accTw = {};
  const long long r0 = (long long)q * FUSED_R_PER_WG + wv * (FUSED_R_PER_WG / 4);
  const int row = l >> 1, col0 = (l & 1) * 16;
  bf16x8 rp[2], ry[2];
  {
    const __bf16* Pr = P + it.P + r0 * FBLK;
    const __bf16* Yr = lay + it.psi_out + r0 * FBLK;
    rp[0] = *reinterpret_cast<const bf16x8*>(Pr + l * 16);
    rp[1] = *reinterpret_cast<const bf16x8*>(Pr + l * 16 + 8);
    ry[0] = *reinterpret_cast<const bf16x8*>(Yr + l * 16);
    ry[1] = *reinterpret_cast<const bf16x8*>(Yr + l * 16 + 8);
    *reinterpret_cast<bf16x8*>(&Pb[0][wv][row * ST_LD + col0]) = rp[0];
    *reinterpret_cast<bf16x8*>(&Pb[0][wv][row * ST_LD + col0 + 8]) = rp[1];
    *reinterpret_cast<bf16x8*>(&Sb[0][wv][row * ST_LD + col0]) = ry[0];
    *reinterpret_cast<bf16x8*>(&Sb[0][wv][row * ST_LD + col0 + 8]) = ry[1];
  }
  __syncthreads();
  constexpr int NB = FUSED_R_PER_WG / 4;
#pragma unroll 1
  for (int i = 0; i < NB; ++i) {
    const int cur = i & 1;
    if (i + 1 < NB) {
      const __bf16* Pr = P + it.P + (r0 + i + 1) * FBLK;
      const __bf16* Yr = lay + it.psi_out + (r0 + i + 1) * FBLK;
      rp[0] = *reinterpret_cast<const bf16x8*>(Pr + l * 16);
      rp[1] = *reinterpret_cast<const bf16x8*>(Pr + l * 16 + 8);
      ry[0] = *reinterpret_cast<const bf16x8*>(Yr + l * 16);
      ry[1] = *reinterpret_cast<const bf16x8*>(Yr + l * 16 + 8);
    }
    const __bf16* pb = Pb[cur][wv];
    const __bf16* sb = Sb[cur][wv];
    // step 1: X^T = Mw^T P_r^T
    f32x16 XT = {};
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const bf16x8 b = *reinterpret_cast<const bf16x8*>(pb + c * ST_LD + 16 * t + 8 * g);
      XT = MFMA_BF16(aMwh[t], b, XT);
      XT = MFMA_BF16(aMwl[t], b, XT);
    }
    // step 2: Tz^T += psi_r X^T
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      bf16x8 a, b;
#pragma unroll
      for (int p = 0; p < 8; ++p) {
        a[p] = sb[c * ST_LD + kperm(t, g, p)];
        b[p] = (__bf16)XT[8 * t + p];
      }
      accTzT = MFMA_BF16(a, b, accTzT);
    }
    // step 3: Zm = Mz psi_r
    f32x16 Zm = {};
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      bf16x8 b;
#pragma unroll
      for (int p = 0; p < 8; ++p) b[p] = sb[(16 * t + 8 * g + p) * ST_LD + c];
      Zm = MFMA_BF16(aMzh[t], b, Zm);
      Zm = MFMA_BF16(aMzl[t], b, Zm);
    }
    // step 4: Tw += P_r^T Zm
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      bf16x8 a, b;
#pragma unroll
      for (int p = 0; p < 8; ++p) {
        a[p] = pb[kperm(t, g, p) * ST_LD + c];
        b[p] = (__bf16)Zm[8 * t + p];
      }
      accTw = MFMA_BF16(a, b, accTw);
    }
    if (i + 1 < NB) {
      *reinterpret_cast<bf16x8*>(&Pb[1 - cur][wv][row * ST_LD + col0]) = rp[0];
      *reinterpret_cast<bf16x8*>(&Pb[1 - cur][wv][row * ST_LD + col0 + 8]) = rp[1];
      *reinterpret_cast<bf16x8*>(&Sb[1 - cur][wv][row * ST_LD + col0]) = ry[0];
      *reinterpret_cast<bf16x8*>(&Sb[1 - cur][wv][row * ST_LD + col0 + 8]) = ry[1];
    }
    __syncthreads();
  }
  for (int w = 0; w < 4; ++w) {
    if (wv == w) {
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int rw = 8 * (v >> 2) + 4 * g + (v & 3);
        if (w == 0) {
          Red[c * FC + rw] = accTzT[v];
          Red[FBLK + rw * FC + c] = accTw[v];
        } else {
          Red[c * FC + rw] += accTzT[v];
          Red[FBLK + rw * FC + c] += accTw[v];
        }
      }
    }
    __syncthreads();
  }
  float* dst = part + it.part + (long long)q * 2 * FBLK;
  for (int o = threadIdx.x; o < 2 * FBLK; o += blockDim.x) dst[o] = Red[o];
}

// ---------------------------------------------------------------------------
// FP64 absorb / fused on v_mfma_f64_16x16x4_f64, operands from LDS
// ---------------------------------------------------------------------------
// 16x16x4 f64 operand layout: A[i][k] in lane (i + 16 k), B[k][j] in lane (j + 16 k);
// the 16x16 accumulator holds D[4 v + l / 16][l % 16] in lane l, value v (measured).
#define F64_DROW(l, v) (4 * (v) + ((l) >> 4))
#define F64_DCOL(l, v) ((l) & 15)

// C[32x32] (+)= A[32x32] * B[32x32]; A, B, C in LDS with row strides; one wave
__device__ inline void wave_gemm32_f64(const double* A, int lda, const double* B, int ldb, double* C,
                                       int ldc, bool accumulate) {
  const int l = threadIdx.x & 63;
  const int li = l & 15, lk = l >> 4;
  f64x4 acc[2][2];
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < 2; ++ni) acc[mi][ni] = f64x4{0, 0, 0, 0};
#pragma unroll
  for (int k0 = 0; k0 < 32; k0 += 4) {
    const double a0 = A[(li)*lda + k0 + lk];
    const double a1 = A[(16 + li) * lda + k0 + lk];
    const double b0 = B[(k0 + lk) * ldb + li];
    const double b1 = B[(k0 + lk) * ldb + 16 + li];
    acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
  }
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < 2; ++ni)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = 16 * mi + F64_DROW(l, v), j = 16 * ni + F64_DCOL(l, v);
        if (accumulate)
          C[i * ldc + j] += acc[mi][ni][v];
        else
          C[i * ldc + j] = acc[mi][ni][v];
      }
}

// one wave: C = A^T-free products use explicit transposed copies in LDS
constexpr int LD64 = 33;

// FP64 absorb: each wave handles 4 input blocks; WG q -> (s, z, 16 w) like the BF16 kernel
__global__ __launch_bounds__(256) void k_absorb_f64(const FastItem* __restrict__ items,
                                                    const double* __restrict__ lay,
                                                    const double* __restrict__ msg,
                                                    double* __restrict__ P) {
  __shared__ double MaT[32 * LD64], Mb[32 * LD64];
  __shared__ double X[4][32 * LD64], V[4][32 * LD64];
  __shared__ double S[8][FBLK];  // staged outputs of 8 inputs (64 KiB)
  const FastItem it = items[blockIdx.y];
  const int q = blockIdx.x;
  const int s = q & 1, z = (q >> 1) & 31, w0 = (q >> 6) * 16;
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int o = threadIdx.x; o < FBLK; o += blockDim.x) {
    const int a = o & 31, b = o >> 5;  // column-major message: M[a + 32 b]
    MaT[b * LD64 + a] = msg[it.Ma + o];  // Ma^T[x'][x] = Ma[x][x']
    Mb[a * LD64 + b] = msg[it.Mb + o];   // Mb[y][y']
  }
  __syncthreads();
  for (int half = 0; half < 2; ++half) {
    for (int i = 0; i < 2; ++i) {
      const int wl = 8 * half + 2 * wv + i;  // 0..15
      const int w = w0 + wl;
      const double* Xg = lay + it.psi_in + (long long)(s + 2 * (z + FC * w)) * FBLK;
      double* x = X[wv];
      for (int o = l; o < FBLK; o += 64) x[(o >> 5) * LD64 + (o & 31)] = Xg[o];
      __builtin_amdgcn_s_waitcnt(0);
      __builtin_amdgcn_wave_barrier();
      wave_gemm32_f64(x, LD64, Mb, LD64, V[wv], LD64, false);      // V = X Mb
      __builtin_amdgcn_s_waitcnt(0);
      __builtin_amdgcn_wave_barrier();
      wave_gemm32_f64(MaT, LD64, V[wv], LD64, x, LD64, false);     // Y = Ma^T V (into x)
      __builtin_amdgcn_s_waitcnt(0);
      __builtin_amdgcn_wave_barrier();
      double* st = S[2 * wv + i];
      for (int o = l; o < FBLK; o += 64) st[o] = x[(o >> 5) * LD64 + (o & 31)];
    }
    __syncthreads();
    for (int q2 = threadIdx.x; q2 < FBLK; q2 += blockDim.x) {
      const int xb = q2 & 31, yb = q2 >> 5;
      const long long r = s + 2 * (xb + FC * yb);
      double* dst = P + it.P + r * FBLK + z * FC + w0 + 8 * half;
#pragma unroll
      for (int j = 0; j < 8; ++j) dst[j] = S[j][xb * FC + yb];
    }
    __syncthreads();
  }
}

// FP64 fused: per block r, Tz += (P_r Mw) psi_r^T, Tw += P_r^T (Mz psi_r); one wave per block
__global__ __launch_bounds__(256) void k_fused_f64(const FastItem* __restrict__ items,
                                                   const double* __restrict__ lay,
                                                   const double* __restrict__ msg,
                                                   const double* __restrict__ P,
                                                   double* __restrict__ part) {
  __shared__ double Mw[32 * LD64], Mz[32 * LD64];
  __shared__ double Pr[4][32 * LD64], PrT[4][32 * LD64], Yt[4][32 * LD64], Yr[4][32 * LD64];
  const FastItem it = items[blockIdx.y];
  const int q = blockIdx.x;
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int o = threadIdx.x; o < FBLK; o += blockDim.x) {
    const int a = o & 31, b = o >> 5;
    Mw[a * LD64 + b] = msg[it.Mw + o];
    Mz[a * LD64 + b] = msg[it.Mz + o];
  }
  __syncthreads();
  const int li = l & 15, lk = l >> 4;
  f64x4 tz[2][2], tw[2][2];
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < 2; ++ni) {
      tz[mi][ni] = f64x4{0, 0, 0, 0};
      tw[mi][ni] = f64x4{0, 0, 0, 0};
    }
  double* pr = Pr[wv];
  double* prt = PrT[wv];
  double* yt = Yt[wv];
  double* yr = Yr[wv];
#pragma unroll 1
  for (int i = 0; i < FUSED_R_PER_WG / 4; ++i) {
    const long long r = (long long)q * FUSED_R_PER_WG + wv * (FUSED_R_PER_WG / 4) + i;
    const double* Pg = P + it.P + r * FBLK;
    const double* Yg = lay + it.psi_out + r * FBLK;
    for (int o = l; o < FBLK; o += 64) {
      const int a = o >> 5, b = o & 31;
      const double pv = Pg[o], yv = Yg[o];
      pr[a * LD64 + b] = pv;
      prt[b * LD64 + a] = pv;
      yr[a * LD64 + b] = yv;
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    wave_gemm32_f64(pr, LD64, Mw, LD64, yt, LD64, false);  // X = P_r Mw  (into yt)
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    // Tz += X psi_r^T : B[k][j] = psi_r[j][k]
#pragma unroll
    for (int k0 = 0; k0 < 32; k0 += 4) {
      const double a0 = yt[li * LD64 + k0 + lk], a1 = yt[(16 + li) * LD64 + k0 + lk];
      const double b0 = yr[li * LD64 + k0 + lk], b1 = yr[(16 + li) * LD64 + k0 + lk];
      tz[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, tz[0][0], 0, 0, 0);
      tz[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, tz[0][1], 0, 0, 0);
      tz[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, tz[1][0], 0, 0, 0);
      tz[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, tz[1][1], 0, 0, 0);
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    wave_gemm32_f64(Mz, LD64, yr, LD64, yt, LD64, false);  // Zm = Mz psi_r (into yt)
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    // Tw += P_r^T Zm
#pragma unroll
    for (int k0 = 0; k0 < 32; k0 += 4) {
      const double a0 = prt[li * LD64 + k0 + lk], a1 = prt[(16 + li) * LD64 + k0 + lk];
      const double b0 = yt[(k0 + lk) * LD64 + li], b1 = yt[(k0 + lk) * LD64 + 16 + li];
      tw[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, tw[0][0], 0, 0, 0);
      tw[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, tw[0][1], 0, 0, 0);
      tw[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, tw[1][0], 0, 0, 0);
      tw[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, tw[1][1], 0, 0, 0);
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
  }
  // reduce the four waves through LDS (reuse Pr as 2 x 1024 doubles per wave? use PrT)
  __syncthreads();
  double* red = &Pr[0][0];  // 4*32*33 doubles >= 2*1024
  for (int w = 0; w < 4; ++w) {
    if (wv == w) {
#pragma unroll
      for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int i = 16 * mi + F64_DROW(l, v), j = 16 * ni + F64_DCOL(l, v);
            if (w == 0) {
              red[i * FC + j] = tz[mi][ni][v];
              red[FBLK + i * FC + j] = tw[mi][ni][v];
            } else {
              red[i * FC + j] += tz[mi][ni][v];
              red[FBLK + i * FC + j] += tw[mi][ni][v];
            }
          }
    }
    __syncthreads();
  }
  double* dst = part + it.part + (long long)q * 2 * FBLK;
  for (int o = threadIdx.x; o < 2 * FBLK; o += blockDim.x) dst[o] = red[o];
}

// ---------------------------------------------------------------------------
// FP64, register-resident flavour.  v_mfma_f64_16x16x4_f64 lane layout (measured):
// A[i][k] in lane (i + 16 k), B[k][j] in lane (j + 16 k), D[4 v + l/16][l % 16]
// in lane l, value v.  A 32x32 product is 2x2 output tiles x 8 K-steps; an
// accumulator tile column is exactly the B fragment of the next product:
// K-step t of a 32-row operand held as acc[mi][ni][v] is acc[t / 4][ni][t % 4].
// ---------------------------------------------------------------------------
#define MFMA_F64(a, b, c) __builtin_amdgcn_mfma_f64_16x16x4f64((a), (b), (c), 0, 0, 0)
constexpr int LD8 = 33;   // padded LDS row stride (doubles) of a 32x32 block

struct Acc32 {
  f64x4 t[2][2];
};

__device__ inline void acc_zero(Acc32& a) {
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < 2; ++ni) a.t[mi][ni] = f64x4{0, 0, 0, 0};
}

// C += A_frag * B_frag, with A given as 16 fragments af[mi][t] and B as
// fragments bf[t][ni]
__device__ inline void mm32(Acc32& C, const double (&af)[2][8], const double (&bf)[8][2]) {
#pragma unroll
  for (int t = 0; t < 8; ++t)
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
      for (int ni = 0; ni < 2; ++ni) C.t[mi][ni] = MFMA_F64(af[mi][t], bf[t][ni], C.t[mi][ni]);
}

// C += A_frag * (accumulator B)
__device__ inline void mm32_accb(Acc32& C, const double (&af)[2][8], const Acc32& B) {
#pragma unroll
  for (int t = 0; t < 8; ++t)
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
      for (int ni = 0; ni < 2; ++ni) C.t[mi][ni] = MFMA_F64(af[mi][t], B.t[t >> 2][ni][t & 3], C.t[mi][ni]);
}

// FP64 absorb: WG (256 threads) -> (s, z, 8 consecutive w); each wave 2 inputs;
// Y = Ma^T (X Mb) in registers, staged [y' + 32 x'][w] for 64-byte output runs.
// grid (256, items)
__global__ __launch_bounds__(256) void k_absorb_f64r(const FastItem* __restrict__ items,
                                                     const double* __restrict__ lay,
                                                     const double* __restrict__ msg,
                                                     double* __restrict__ P) {
  constexpr int SROW = 10;  // staging row stride (doubles): 80 bytes
  __shared__ __attribute__((aligned(16))) double S[FBLK * SROW];
  const FastItem it = items[blockIdx.y];
  const int q = blockIdx.x;
  const int s = q & 1, z = (q >> 1) & 31, w0 = (q >> 6) * 8;
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int li = l & 15, lg = l >> 4;
  double bMb[8][2], aMaT[2][8];
#pragma unroll
  for (int t = 0; t < 8; ++t)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      // Mb[y][y'] = msg[Mb + y + 32 y'];  B fragment (k = 4t + lg, j = 16 h + li)
      bMb[t][h] = msg[it.Mb + (4 * t + lg) + FC * (16 * h + li)];
      // Ma^T[x'][x] = Ma[x][x'];  A fragment (i = 16 h + li, k = 4t + lg)
      aMaT[h][t] = msg[it.Ma + (4 * t + lg) + FC * (16 * h + li)];
    }
  // both input blocks of this wave are requested before the first product
  double aXs[2][2][8];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const double* X = lay + it.psi_in + (long long)(s + 2 * (z + FC * (w0 + 2 * wv + i))) * FBLK;
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int t = 0; t < 8; ++t) aXs[i][h][t] = X[(16 * h + li) * FC + 4 * t + lg];
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int wl = 2 * wv + i;
    const double(&aX)[2][8] = aXs[i];
    Acc32 V, Y;
    acc_zero(V);
    acc_zero(Y);
    mm32(V, aX, bMb);
    mm32_accb(Y, aMaT, V);
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
      for (int ni = 0; ni < 2; ++ni)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int xb = 16 * mi + 4 * v + lg, yb = 16 * ni + li;
          S[(yb + FC * xb) * SROW + wl] = Y.t[mi][ni][v];
        }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < FBLK * 4; k += blockDim.x) {
    const int o = k >> 2, part = k & 3;
    const int yb = o & 31, xb = o >> 5;
    const long long r = s + 2 * (xb + FC * yb);
    const double2 v2 = *reinterpret_cast<const double2*>(S + o * SROW + 2 * part);
    *reinterpret_cast<double2*>(P + it.P + r * FBLK + z * FC + w0 + 2 * part) = v2;
  }
}

// FP64 fused: per block r (one wave), Tz^T += psi_r (Mw^T P_r^T), Tw += P_r^T (Mz psi_r);
// the P_r and psi_r blocks are staged in LDS; grid (16, items)
__global__ __launch_bounds__(256) void k_fused_f64r(const FastItem* __restrict__ items,
                                                    const double* __restrict__ lay,
                                                    const double* __restrict__ msg,
                                                    const double* __restrict__ P,
                                                    double* __restrict__ part) {
  __shared__ double Pb[4][FC * LD8];
  __shared__ double Sb[4][FC * LD8];
  const FastItem it = items[blockIdx.y];
  const int q = blockIdx.x;
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int li = l & 15, lg = l >> 4;
  double aMwT[2][8], aMz[2][8];
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      // Mw^T[w'][w] = Mw[w][w'] : A fragment i = w' = 16h + li, k = w = 4t + lg
      aMwT[h][t] = msg[it.Mw + (4 * t + lg) + FC * (16 * h + li)];
      // Mz[z][z'] : A fragment i = z = 16h + li, k = z' = 4t + lg
      aMz[h][t] = msg[it.Mz + (16 * h + li) + FC * (4 * t + lg)];
    }
  Acc32 tzT, tw;
  acc_zero(tzT);
  acc_zero(tw);
  double* pb = Pb[wv];
  double* sb = Sb[wv];
  constexpr int NB = FUSED_R_PER_WG / 4;
  const long long r0 = (long long)q * FUSED_R_PER_WG + wv * NB;
  // software pipeline: the global loads of block i+1 are in flight (registers
  // nP, nY) while block i is multiplied out of LDS
  double2 nP[8], nY[8];
  auto fetch = [&](long long r) {
    const double* Pg = P + it.P + r * FBLK;
    const double* Yg = lay + it.psi_out + r * FBLK;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      nP[j] = *reinterpret_cast<const double2*>(Pg + 2 * l + 128 * j);
      nY[j] = *reinterpret_cast<const double2*>(Yg + 2 * l + 128 * j);
    }
  };
  auto stage = [&]() {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int o = 2 * l + 128 * j;  // two consecutive doubles per lane per step
      const int a = o >> 5, b = o & 31;
      pb[a * LD8 + b] = nP[j].x;
      pb[a * LD8 + b + 1] = nP[j].y;
      sb[a * LD8 + b] = nY[j].x;
      sb[a * LD8 + b + 1] = nY[j].y;
    }
  };
  fetch(r0);
  stage();
#pragma unroll 1
  for (int i = 0; i < NB; ++i) {
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    if (i + 1 < NB) fetch(r0 + i + 1);
    // X^T = Mw^T P_r^T : B fragment (k = w = 4t + lg, j = z = 16h + li) = P_r[z][w]
    {
      double bPT[8][2];
#pragma unroll
      for (int t = 0; t < 8; ++t)
#pragma unroll
        for (int h = 0; h < 2; ++h) bPT[t][h] = pb[(16 * h + li) * LD8 + 4 * t + lg];
      Acc32 XT;
      acc_zero(XT);
      mm32(XT, aMwT, bPT);
      // Tz^T += psi_r X^T : A fragment (i = y = 16h + li, k = w' = 4t + lg) = psi_r[y][w']
      double aY[2][8];
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int t = 0; t < 8; ++t) aY[h][t] = sb[(16 * h + li) * LD8 + 4 * t + lg];
      mm32_accb(tzT, aY, XT);
    }
    // Zm = Mz psi_r : B fragment (k = z' = 4t + lg, j = y = 16h + li) = psi_r[z'][y]
    {
      double bY[8][2];
#pragma unroll
      for (int t = 0; t < 8; ++t)
#pragma unroll
        for (int h = 0; h < 2; ++h) bY[t][h] = sb[(4 * t + lg) * LD8 + 16 * h + li];
      Acc32 Zm;
      acc_zero(Zm);
      mm32(Zm, aMz, bY);
      // Tw += P_r^T Zm : A fragment (i = x = 16h + li, k = z = 4t + lg) = P_r[z][x]
      double aPT[2][8];
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int t = 0; t < 8; ++t) aPT[h][t] = pb[(4 * t + lg) * LD8 + 16 * h + li];
      mm32_accb(tw, aPT, Zm);
    }
    // every lane's LDS reads of block i are issued before any lane overwrites it
    __builtin_amdgcn_wave_barrier();
    if (i + 1 < NB) stage();
  }
  __builtin_amdgcn_s_waitcnt(0);
  // reduce the four waves: Tz[x][y] from tz^T[y][x], Tw[x][y]
  __syncthreads();
  double* red = &Pb[0][0];  // 4 * 32 * 33 >= 2 * 1024
  for (int w = 0; w < 4; ++w) {
    if (wv == w) {
#pragma unroll
      for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int i = 16 * mi + 4 * v + lg, j = 16 * ni + li;
            if (w == 0) {
              red[j * FC + i] = tzT.t[mi][ni][v];
              red[FBLK + i * FC + j] = tw.t[mi][ni][v];
            } else {
              red[j * FC + i] += tzT.t[mi][ni][v];
              red[FBLK + i * FC + j] += tw.t[mi][ni][v];
            }
          }
    }
    __syncthreads();
  }
  double* dst = part + it.part + (long long)q * 2 * FBLK;
  for (int o = threadIdx.x; o < 2 * FBLK; o += blockDim.x) dst[o] = red[o];
}

// ---------------------------------------------------------------------------
// finaliser: sum partials (FP64), normalise, message_diff / residual, write
// ---------------------------------------------------------------------------
template <typename TP, typename TM>
__global__ __launch_bounds__(256) void k_fast_finalize(const FastItem* __restrict__ items,
                                                       const TP* __restrict__ part,
                                                       const TM* msg_in, TM* msg_out,
                                                       MsgStat* __restrict__ stats, int normalize,
                                                       const int* __restrict__ sexp) {
  __shared__ double T[FBLK];
  __shared__ double red[8];
  const FastItem it = items[blockIdx.y];
  const int mat = blockIdx.x;
  const TP* pp = part + it.part + (long long)mat * FBLK;
  const double us = site_unscale2(sexp, it.site);
  double n2 = 0;
  for (int o = threadIdx.x; o < FBLK; o += blockDim.x) {
    double sum = 0;
    for (int ch = 0; ch < FUSED_CHUNKS; ++ch) sum += (double)pp[(long long)ch * 2 * FBLK + o];
    sum *= us;
    const int x = o >> 5, y = o & 31;
    T[x + FC * y] = sum;  // message layout M[b_ket + chi b_bra]
    n2 += sum * sum;
  }
  auto bsum = [&](double v) {
    for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += red[i];
    return s;
  };
  n2 = bsum(n2);
  const double nrm = sqrt(n2);
  const double sc = (normalize && nrm != 0.0) ? 1.0 / nrm : 1.0;
  const long long off = mat == 0 ? it.out_z : it.out_w;
  const int slot = mat == 0 ? it.slot_z : it.slot_w;
  double dot = 0, o2 = 0, d2 = 0, bad = 0;
  for (int o = threadIdx.x; o < FBLK; o += blockDim.x) {
    const double nv = T[o] * sc;
    const double ov = (double)msg_in[off + o];
    if (!isfinite(nv)) bad = 1;
    dot += ov * nv;
    o2 += ov * ov;
    d2 += (nv - ov) * (nv - ov);
  }
  __syncthreads();
  for (int o = threadIdx.x; o < FBLK; o += blockDim.x) msg_out[off + o] = (TM)(T[o] * sc);
  dot = bsum(dot);
  o2 = bsum(o2);
  d2 = bsum(d2);
  bad = bsum(bad);
  if (threadIdx.x == 0 && slot >= 0) {
    const double nn2 = n2 * sc * sc;
    MsgStat st;
    st.diff = 1.0 - (dot * dot) / (o2 * nn2);
    st.res2 = nn2 > 0 ? d2 / nn2 : 0.0;
    st.norm = nrm;
    st.bad = (bad > 0 || !isfinite(nrm)) ? 1.0 : 0.0;
    stats[slot] = st;
  }
}

// ---------------------------------------------------------------------------
// Eigenbasis pipeline (symmetric messages).  With M_c = U_c diag(l_c) U_c^T for
// the four incoming messages, rotate every bond of psi into the eigenbasis of
// its message,
//   Y[s, r1, r2, r3, r4] = sum_b psi[s, b1, b2, b3, b4] prod_c U_c[b_c, r_c],
// and every outgoing message becomes a weighted Gram of one mode of Y rotated
// back (exact for any orthogonal U_c; eigenvalues of either sign):
//   T_c = U_c T'_c U_c^T,  T'_c[r, r'] = sum_{s, r_e (e != c)} prod_{e != c} l_e[r_e] Y[.., r, ..] Y[.., r', ..].
// Three streaming kernels per sweep:
//   rot12  = the absorb kernel with (U1, U2): psi layout-B block X[b1][b2] at (s, b3, b4)
//            -> R block (s + 2 r1 + 64 r2) = [b3][b4] of U1^T X U2
//   rot34  : R block Z -> Y block = U3^T Z U4 ([r3][r4]); T'3 += l1 l2 Y L4 Y^T and
//            T'4 += l1 l2 Y^T L3 Y in MFMA accumulators; Y scattered into the
//            Q block (s + 2 r3 + 64 r4) = [r1][r2]
//   gram12 : Q block D -> T'1 += l3 l4 D L2 D^T, T'2 += l3 l4 D^T L1 D
// then a finaliser rotates back, symmetrises, normalises and measures.  Per
// vertex and sweep that is 20 MiB of BF16 traffic (psi in, R out and in, Q
// out and in) instead of the 32 MiB of the absorbed-pair passes, and 9 GEMM
// units of 32^3 per block instead of 12.
// ---------------------------------------------------------------------------
constexpr int EIG_NPART = 16;  // partial Grams per (vertex, bond)
constexpr int ROT_NZ = 4;      // r1 values per rot34 workgroup (grid.x = 2 * 32 / ROT_NZ = EIG_NPART)

struct EigItem {
  long long R, Q;      // block buffers (element offsets)
  long long X;         // degree-3 sites: layout-C blocks of psi
  long long U[4];      // eigenvectors of the incoming message on bond c (column-major U[b + 32 r])
  long long L[4];      // its eigenvalues
  long long out[4];    // outgoing message offsets
  int slot[4];
  int site;            // index of the site's power-of-two scale (BF16/FP16 layouts)
  long long part;      // partial Grams [4 bonds][EIG_NPART][32 x 32]
};

__device__ inline void wave_sync() {
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
}

// Symmetric eigendecomposition of one 32x32 message per wave (cyclic parallel
// Jacobi, 16 disjoint rotations per step, round-robin pairing), input
// symmetrised as S = (M + M^T)/2.  With warm = 1 the previous sweep's
// eigenvectors U0 (in U64) seed it: Jacobi then diagonalises U0^T S U0, which
// is nearly diagonal once BP settles, and the rotations accumulate onto U0.
// Stops when the off-diagonal mass is below tol2 of the diagonal's (1e-28 for
// the FP64 pipeline, 1e-14 for the BF16 one, whose operands carry ~16 bits) or
// stops decreasing.  grid (ceil(n/4)), 256 threads.
template <typename TM>
__global__ __launch_bounds__(256) void k_msg_eigh(const long long* __restrict__ moff, int n,
                                                 const TM* __restrict__ msg, double* __restrict__ U64,
                                                 double* __restrict__ L64, float* __restrict__ U32,
                                                 float* __restrict__ L32, int* __restrict__ Lexp, int warm,
                                                 double tol2) {
  constexpr int LD = 33;
  __shared__ double As[4][FC * LD];
  __shared__ double Vs[4][FC * LD];
  __shared__ double rc[4][16], rs[4][16];
  __shared__ int rp[4][16], rq[4][16];
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int e = blockIdx.x * 4 + wv;
  if (e >= n) return;  // whole waves only; the kernel has no workgroup barrier
  double* A = As[wv];
  double* V = Vs[wv];
  const TM* M = msg + moff[e];
  double* u64 = U64 + (long long)e * FBLK;
  for (int o = l; o < FBLK; o += 64) {
    const int i = o & 31, j = o >> 5;
    A[i * LD + j] = 0.5 * ((double)M[i + FC * j] + (double)M[j + FC * i]);
    V[i * LD + j] = warm ? u64[i + FC * j] : ((i == j) ? 1.0 : 0.0);  // V[b][r] = U0[b + 32 r]
  }
  wave_sync();
  if (warm) {
    // A <- U0^T S U0.  Lane l owns column j = l & 31 and rows i = (l >> 5) + 2 m
    // (m = 0..15) of each product; the 16 accumulators are independent, so the
    // loads of one k-step are all in flight together.
    const int jc = l & 31, i0 = l >> 5;
    double t[16];
#pragma unroll
    for (int m = 0; m < 16; ++m) t[m] = 0;
#pragma unroll 2
    for (int q = 0; q < FC; ++q) {
      const double u = V[q * LD + jc];
#pragma unroll
      for (int m = 0; m < 16; ++m) t[m] += A[(i0 + 2 * m) * LD + q] * u;  // (S U0)[i][j]
    }
    wave_sync();
#pragma unroll
    for (int m = 0; m < 16; ++m) A[(i0 + 2 * m) * LD + jc] = t[m];
    wave_sync();
#pragma unroll
    for (int m = 0; m < 16; ++m) t[m] = 0;
#pragma unroll 2
    for (int q = 0; q < FC; ++q) {
      const double x = A[q * LD + jc];
#pragma unroll
      for (int m = 0; m < 16; ++m) t[m] += V[q * LD + i0 + 2 * m] * x;  // (U0^T S U0)[i][j]
    }
    wave_sync();
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      const int i = i0 + 2 * m;
      if (i <= jc) A[i * LD + jc] = t[m];
    }
    wave_sync();
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      const int i = i0 + 2 * m;
      if (i > jc) A[i * LD + jc] = A[jc * LD + i];  // exact symmetry from the upper triangle
    }
    wave_sync();
  }
  const int col = l & 31, half = l >> 5;
  double prev = INFINITY;
  for (int sweep = 0; sweep < 30; ++sweep) {
    double off = 0, dg = 0;
    for (int o = l; o < FBLK; o += 64) {
      const int i = o >> 5, j = o & 31;
      const double a = A[i * LD + j];
      if (i != j) off += a * a;
      else dg += a * a;
    }
    for (int m = 32; m > 0; m >>= 1) {
      off += __shfl_xor(off, m);
      dg += __shfl_xor(dg, m);
    }
    if (off <= tol2 * dg || off == 0.0 || off >= prev) break;
    prev = off;
    for (int st = 0; st < FC - 1; ++st) {
      if (l < 16) {
        auto at = [&](int pos) { return pos == 0 ? 0 : 1 + (pos - 1 + st) % (FC - 1); };
        int p = at(l), q = at(FC - 1 - l);
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
        double c = 1, s = 0;
        const double apq = A[p * LD + q];
        if (apq != 0.0) {
          const double zeta = (A[q * LD + q] - A[p * LD + p]) / (2 * apq);
          const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1 + zeta * zeta));
          c = 1 / sqrt(1 + t * t);
          s = c * t;
        }
        rc[wv][l] = c;
        rs[wv][l] = s;
        rp[wv][l] = p;
        rq[wv][l] = q;
      }
      wave_sync();
      // the 8 (disjoint) pairs of this lane: every operand is loaded before any
      // store, so the LDS latency is paid once per phase rather than per pair
      int pp[8], qq[8];
      double cc[8], ss[8];
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) {
        const int i = half + 2 * jj;
        pp[jj] = rp[wv][i];
        qq[jj] = rq[wv][i];
        cc[jj] = rc[wv][i];
        ss[jj] = rs[wv][i];
      }
      // rows: A <- J^T A
      double x0[8], x1[8];
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) {
        x0[jj] = A[pp[jj] * LD + col];
        x1[jj] = A[qq[jj] * LD + col];
      }
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) {
        A[pp[jj] * LD + col] = cc[jj] * x0[jj] - ss[jj] * x1[jj];
        A[qq[jj] * LD + col] = ss[jj] * x0[jj] + cc[jj] * x1[jj];
      }
      wave_sync();
      // columns: A <- A J, V <- V J
      double y0[8], y1[8];
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) {
        x0[jj] = A[col * LD + pp[jj]];
        x1[jj] = A[col * LD + qq[jj]];
        y0[jj] = V[col * LD + pp[jj]];
        y1[jj] = V[col * LD + qq[jj]];
      }
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) {
        A[col * LD + pp[jj]] = cc[jj] * x0[jj] - ss[jj] * x1[jj];
        A[col * LD + qq[jj]] = ss[jj] * x0[jj] + cc[jj] * x1[jj];
        V[col * LD + pp[jj]] = cc[jj] * y0[jj] - ss[jj] * y1[jj];
        V[col * LD + qq[jj]] = ss[jj] * y0[jj] + cc[jj] * y1[jj];
      }
      wave_sync();
    }
  }
  float* u32 = U32 ? U32 + (long long)e * FBLK : nullptr;
  for (int o = l; o < FBLK; o += 64) {
    const int b = o & 31, r = o >> 5;
    const double v = V[b * LD + r];
    u64[o] = v;  // U[b + 32 r]: column r is the eigenvector of eigenvalue r
    if (u32) u32[o] = (float)v;
  }
  // the FP32 copy is scaled by 2^-x so that its largest |lam| lies in [1/2, 1):
  // products of eigenvalues (the Gram weights of the 16-bit kernels) then stay
  // of order one whatever the messages' scale, and the finaliser multiplies the
  // Gram back by 2^(sum of the x of the bonds it weighs)
  const double lam = l < FC ? A[l * LD + l] : 0.0;
  if (l < FC) L64[(long long)e * FC + l] = lam;
  if (L32) {
    double m = fabs(lam);
    for (int k = 32; k > 0; k >>= 1) m = fmax(m, __shfl_xor(m, k));
    int x = 0;
    if (m > 0 && isfinite(m)) frexp(m, &x);
    if (l < FC) L32[(long long)e * FC + l] = (float)ldexp(lam, -x);
    if (l == 0) Lexp[e] = x;
  }
}

// rot34, BF16: WG (512 threads) -> (s, ROT_NZ consecutive r1, all 32 r2); wave
// wv -> r2 in [4 wv, 4 wv + 4).  The four R blocks of the next r1 are in flight
// while the current ones are multiplied.  grid (EIG_NPART, items)
// With GRAM = false it is rot12: the same block geometry (input block
// s + 2 (z + 32 w), output block s + 2 (x' + 32 y') at [z][w]) with U1, U2 in the
// U[2], U[3] slots of its items, psi layout B as R and the R buffer as Q.
// Staging tile row stride (16-bit elements): 72 bytes, so that the tile
// (73.7 KB) lets two workgroups share a CU, and the row stride is 18 banks
// (2-way conflicts on the column writes instead of 4-way at 80 bytes)
constexpr int ROT_LD = 36;

template <bool GRAM, typename TI>
__global__ __launch_bounds__(512) void k_rot34_bf16(const EigItem* __restrict__ items,
                                                    const TI* __restrict__ R,
                                                    const float* __restrict__ U,
                                                    const float* __restrict__ lam,
                                                    TI* __restrict__ Q, float* __restrict__ part) {
  typedef typename Mma<TI>::v8 v8;
  __shared__ __attribute__((aligned(16))) TI S[FBLK * ROT_LD];
  __shared__ __attribute__((aligned(16))) float lam34[2][FC];  // l3, l4
  const EigItem it = items[blockIdx.y];
  const int q = blockIdx.x;
  const int s = q & 1, z0 = (q >> 1) * ROT_NZ;
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6, g = l >> 5, c = l & 31;
  if (GRAM && threadIdx.x < 2 * FC)
    lam34[threadIdx.x >> 5][threadIdx.x & 31] = lam[((threadIdx.x >> 5) ? it.L[3] : it.L[2]) + (threadIdx.x & 31)];
  v8 aU3h[2], aU3l[2], bU4h[2], bU4l[2];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      TI h, lo;
      // U3 as A operand of U3^T V (lane r3, K = b3 in accumulator-row order) and as
      // B operand of V^T U3 (lane r3, same K order): the same values
      split16(U[it.U[2] + kperm(t, g, p) + FC * c], h, lo);
      aU3h[t][p] = h;
      aU3l[t][p] = lo;
      // U4 as B operand of Z U4 (lane r4, K = b4 natural)
      split16(U[it.U[3] + (16 * t + 8 * g + p) + FC * c], h, lo);
      bU4h[t][p] = h;
      bU4l[t][p] = lo;
    }
  __syncthreads();
  f32x16 T3 = {}, T4 = {};
  // the wave's four R blocks of the current and the next r1 (plain unrolled
  // code, no array references, so that the blocks stay in registers)
  v8 cur[4][2], nxt[4][2];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const TI* Z = R + it.R + (long long)(s + 2 * (z0 + FC * (4 * wv + i))) * FBLK;
#pragma unroll
    for (int t = 0; t < 2; ++t) cur[i][t] = *reinterpret_cast<const v8*>(Z + c * FC + 16 * t + 8 * g);
  }
#pragma unroll 1
  for (int zi = 0; zi < ROT_NZ; ++zi) {
    const int z = z0 + zi;
    if (zi + 1 < ROT_NZ) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const TI* Z = R + it.R + (long long)(s + 2 * (z + 1 + FC * (4 * wv + i))) * FBLK;
#pragma unroll
        for (int t = 0; t < 2; ++t) nxt[i][t] = *reinterpret_cast<const v8*>(Z + c * FC + 16 * t + 8 * g);
      }
    }
    const float l1 = GRAM ? lam[it.L[0] + z] : 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int w = 4 * wv + i;
      const float w12 = GRAM ? l1 * lam[it.L[1] + w] : 0.f;
      f32x16 V = {};
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        V = Mma<TI>::mfma(cur[i][t], bU4h[t], V);
        V = Mma<TI>::mfma(cur[i][t], bU4l[t], V);
      }
      v8 bV[2];
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int p = 0; p < 8; ++p) bV[t][p] = (TI)V[8 * t + p];
      f32x16 Y = {}, YT = {};
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        Y = Mma<TI>::mfma(aU3h[t], bV[t], Y);
        Y = Mma<TI>::mfma(aU3l[t], bV[t], Y);
        if (GRAM) {
          YT = Mma<TI>::mfma(bV[t], aU3h[t], YT);
          YT = Mma<TI>::mfma(bV[t], aU3l[t], YT);
        }
      }
#pragma unroll
      for (int t = 0; t < 2 && GRAM; ++t) {
        v8 a3, b3, a4, b4;
        // eigenvalue weights of accumulator rows kperm(t, g, 0..7): two runs of 4
        const float4 l3a = *reinterpret_cast<const float4*>(&lam34[0][16 * t + 4 * g]);
        const float4 l3b = *reinterpret_cast<const float4*>(&lam34[0][16 * t + 8 + 4 * g]);
        const float4 l4a = *reinterpret_cast<const float4*>(&lam34[1][16 * t + 4 * g]);
        const float4 l4b = *reinterpret_cast<const float4*>(&lam34[1][16 * t + 8 + 4 * g]);
        const float l3[8] = {l3a.x, l3a.y, l3a.z, l3a.w, l3b.x, l3b.y, l3b.z, l3b.w};
        const float l4[8] = {l4a.x, l4a.y, l4a.z, l4a.w, l4b.x, l4b.y, l4b.z, l4b.w};
#pragma unroll
        for (int p = 0; p < 8; ++p) {
          const float yt = YT[8 * t + p], y = Y[8 * t + p];
          a3[p] = (TI)yt;
          b3[p] = (TI)(w12 * l4[p] * yt);
          a4[p] = (TI)y;
          b4[p] = (TI)(w12 * l3[p] * y);
        }
        T3 = Mma<TI>::mfma(a3, b3, T3);
        T4 = Mma<TI>::mfma(a4, b4, T4);
      }
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int xb = 8 * (v >> 2) + 4 * g + (v & 3);
        S[(c + FC * xb) * ROT_LD + w] = (TI)Y[v];  // row (r4 + 32 r3), column r2
      }
      __builtin_amdgcn_sched_barrier(0);  // keep the four blocks' chains apart (register pressure)
    }
    __syncthreads();
    for (int k = threadIdx.x; k < FBLK * 4; k += blockDim.x) {
      const int o = k >> 2, pt = k & 3;
      const int yb = o & 31, xb = o >> 5;
      const long long qb = s + 2 * (xb + FC * yb);
      // 8-byte aligned tile rows: two 8-byte reads, one 16-byte store
      const uint2* src = reinterpret_cast<const uint2*>(S + o * ROT_LD + 8 * pt);
      const uint2 lo = src[0], hi = src[1];
      *reinterpret_cast<uint4*>(Q + it.Q + qb * FBLK + z * FC + 8 * pt) = make_uint4(lo.x, lo.y, hi.x, hi.y);
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      cur[i][0] = nxt[i][0];
      cur[i][1] = nxt[i][1];
    }
  }
  if (!GRAM) return;
  float* Red = reinterpret_cast<float*>(S);
  for (int w8 = 0; w8 < 8; ++w8) {
    if (wv == w8) {
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int rw = 8 * (v >> 2) + 4 * g + (v & 3);
        if (w8 == 0) {
          Red[rw * FC + c] = T3[v];
          Red[FBLK + rw * FC + c] = T4[v];
        } else {
          Red[rw * FC + c] += T3[v];
          Red[FBLK + rw * FC + c] += T4[v];
        }
      }
    }
    __syncthreads();
  }
  float* d3 = part + it.part + (long long)(2 * EIG_NPART + q) * FBLK;
  float* d4 = part + it.part + (long long)(3 * EIG_NPART + q) * FBLK;
  for (int o = threadIdx.x; o < FBLK; o += blockDim.x) {
    d3[o] = Red[o];
    d4[o] = Red[FBLK + o];
  }
}

// gram12, BF16: WG (256 threads) -> 4 NB consecutive Q blocks, wave -> NB (32
// for degree-4 sites, 1 for degree 3); the next 4 blocks of a wave are in
// flight while the current 4 are multiplied.  D^T comes from D I on the MFMA
// (exact).  Block qb has weight l3[(qb >> 1) & 31] l4[qb >> 6].
// grid (EIG_NPART, items)
template <int NB, typename TI>
__global__ __launch_bounds__(256) void k_gram12_bf16(const EigItem* __restrict__ items,
                                                     const TI* __restrict__ Q,
                                                     const float* __restrict__ lam,
                                                     float* __restrict__ part) {
  typedef typename Mma<TI>::v8 v8;
  __shared__ float Red[2 * FBLK];
  const EigItem it = items[blockIdx.y];
  const int q = blockIdx.x;
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6, g = l >> 5, c = l & 31;
  float lam1v[16], lam2v[16];
  v8 bI[2];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      lam1v[8 * t + p] = lam[it.L[0] + kperm(t, g, p)];   // K = r1 in accumulator-row order
      lam2v[8 * t + p] = lam[it.L[1] + 16 * t + 8 * g + p];  // K = r2 natural
      bI[t][p] = (TI)((16 * t + 8 * g + p) == c ? 1.0f : 0.0f);
    }
  f32x16 T1 = {}, T2 = {};
  constexpr int PF = NB < 4 ? NB : 4;
  const long long r0 = (long long)q * (4 * NB) + wv * NB;
  v8 cur[PF][2], nxt[PF][2];
  auto fetch = [&](long long r, v8 (&dst)[PF][2]) {
#pragma unroll
    for (int i = 0; i < PF; ++i) {
      const TI* D = Q + it.Q + (r + i) * FBLK;
#pragma unroll
      for (int t = 0; t < 2; ++t) dst[i][t] = *reinterpret_cast<const v8*>(D + c * FC + 16 * t + 8 * g);
    }
  };
  fetch(r0, cur);
#pragma unroll 1
  for (int i0 = 0; i0 < NB; i0 += PF) {
    if (i0 + PF < NB) fetch(r0 + i0 + PF, nxt);
#pragma unroll
    for (int i = 0; i < PF; ++i) {
      const long long qb = r0 + i0 + i;
      const int r3 = (int)((qb >> 1) & 31), r4 = (int)(qb >> 6);
      const float w34 = lam[it.L[2] + r3] * lam[it.L[3] + r4];
      f32x16 Dt = {};
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        v8 b1;
#pragma unroll
        for (int p = 0; p < 8; ++p) b1[p] = (TI)(w34 * lam2v[8 * t + p] * (float)cur[i][t][p]);
        T1 = Mma<TI>::mfma(cur[i][t], b1, T1);
        Dt = Mma<TI>::mfma(cur[i][t], bI[t], Dt);
      }
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        v8 a2, b2;
#pragma unroll
        for (int p = 0; p < 8; ++p) {
          const float d = Dt[8 * t + p];
          a2[p] = (TI)d;
          b2[p] = (TI)(w34 * lam1v[8 * t + p] * d);
        }
        T2 = Mma<TI>::mfma(a2, b2, T2);
      }
    }
#pragma unroll
    for (int i = 0; i < PF; ++i) {
      cur[i][0] = nxt[i][0];
      cur[i][1] = nxt[i][1];
    }
  }
  for (int w4 = 0; w4 < 4; ++w4) {
    if (wv == w4) {
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int rw = 8 * (v >> 2) + 4 * g + (v & 3);
        if (w4 == 0) {
          Red[rw * FC + c] = T1[v];
          Red[FBLK + rw * FC + c] = T2[v];
        } else {
          Red[rw * FC + c] += T1[v];
          Red[FBLK + rw * FC + c] += T2[v];
        }
      }
    }
    __syncthreads();
  }
  float* d1 = part + it.part + (long long)(0 * EIG_NPART + q) * FBLK;
  float* d2 = part + it.part + (long long)(1 * EIG_NPART + q) * FBLK;
  for (int o = threadIdx.x; o < FBLK; o += blockDim.x) {
    d1[o] = Red[o];
    d2[o] = Red[FBLK + o];
  }
}

// rot34, FP64: WG (256 threads) -> (s, 8 consecutive r2, ROT_NZ64 consecutive r1);
// wave wv -> r2 = w0 + 2 wv + {0, 1}.  One block of look-ahead: the loads of the
// next block are in flight during the 160 MFMAs of the current one.
// grid (EIG_NPART, items) = (2 s x 4 r2 groups x 2 r1 groups)
constexpr int ROT_NZ64 = 16;

// GRAM = false: rot12, as for the BF16 kernel
template <bool GRAM>
__global__ __launch_bounds__(256) void k_rot34_f64(const EigItem* __restrict__ items,
                                                   const double* __restrict__ R,
                                                   const double* __restrict__ U,
                                                   const double* __restrict__ lam,
                                                   double* __restrict__ Q, double* __restrict__ part) {
  constexpr int SROW = 10;
  __shared__ __attribute__((aligned(16))) double S[FBLK * SROW];
  const EigItem it = items[blockIdx.y];
  const int q = blockIdx.x;
  const int s = q & 1, w0 = ((q >> 1) & 3) * 8, z0 = (q >> 3) * ROT_NZ64;
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int li = l & 15, lg = l >> 4;
  double aU3T[2][8], bU4[8][2], lam3v[8], lam4v[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      aU3T[h][t] = U[it.U[2] + (4 * t + lg) + FC * (16 * h + li)];  // U3^T[16h+li][4t+lg]
      bU4[t][h] = U[it.U[3] + (4 * t + lg) + FC * (16 * h + li)];   // U4[4t+lg][16h+li]
    }
    lam3v[t] = GRAM ? lam[it.L[2] + 4 * t + lg] : 0.0;
    lam4v[t] = GRAM ? lam[it.L[3] + 4 * t + lg] : 0.0;
  }
  Acc32 T3, T4;
  acc_zero(T3);
  acc_zero(T4);
  double cur[2][8], nxt[2][8];
  auto fetch = [&](int j, double (&dst)[2][8]) {  // block j of this wave: z = z0 + j / 2, w = w0 + 2 wv + j % 2
    const int z = z0 + (j >> 1), w = w0 + 2 * wv + (j & 1);
    const double* Z = R + it.R + (long long)(s + 2 * (z + FC * w)) * FBLK;
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int t = 0; t < 8; ++t) dst[h][t] = Z[(16 * h + li) * FC + 4 * t + lg];
  };
  fetch(0, cur);
#pragma unroll 1
  for (int zi = 0; zi < ROT_NZ64; ++zi) {
    const int z = z0 + zi;
    const double l1 = GRAM ? lam[it.L[0] + z] : 0.0;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int j = 2 * zi + i;
      if (j + 1 < 2 * ROT_NZ64) fetch(j + 1, nxt);
      const int wl = 2 * wv + i;
      const double w12 = GRAM ? l1 * lam[it.L[1] + w0 + wl] : 0.0;
      Acc32 V, Y, YT;
      acc_zero(V);
      acc_zero(Y);
      acc_zero(YT);
      mm32(V, cur, bU4);       // V = Z U4
      mm32_accb(Y, aU3T, V);   // Y = U3^T V
#pragma unroll
      for (int t = 0; t < 8 && GRAM; ++t)  // Y^T = V^T U3
#pragma unroll
        for (int mi = 0; mi < 2; ++mi)
#pragma unroll
          for (int ni = 0; ni < 2; ++ni)
            YT.t[mi][ni] = MFMA_F64(V.t[t >> 2][mi][t & 3], aU3T[ni][t], YT.t[mi][ni]);
#pragma unroll
      for (int t = 0; t < 8 && GRAM; ++t)
#pragma unroll
        for (int mi = 0; mi < 2; ++mi)
#pragma unroll
          for (int ni = 0; ni < 2; ++ni) {
            // T3 += Y L4 Y^T (K = r4), T4 += Y^T L3 Y (K = r3)
            T3.t[mi][ni] = MFMA_F64(YT.t[t >> 2][mi][t & 3], w12 * lam4v[t] * YT.t[t >> 2][ni][t & 3], T3.t[mi][ni]);
            T4.t[mi][ni] = MFMA_F64(Y.t[t >> 2][mi][t & 3], w12 * lam3v[t] * Y.t[t >> 2][ni][t & 3], T4.t[mi][ni]);
          }
#pragma unroll
      for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int xb = 16 * mi + 4 * v + lg, yb = 16 * ni + li;
            S[(yb + FC * xb) * SROW + wl] = Y.t[mi][ni][v];
          }
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int t = 0; t < 8; ++t) cur[h][t] = nxt[h][t];
    }
    __syncthreads();
    for (int k = threadIdx.x; k < FBLK * 4; k += blockDim.x) {
      const int o = k >> 2, pt = k & 3;
      const int yb = o & 31, xb = o >> 5;
      const long long qb = s + 2 * (xb + FC * yb);
      *reinterpret_cast<double2*>(Q + it.Q + qb * FBLK + z * FC + w0 + 2 * pt) =
          *reinterpret_cast<const double2*>(S + o * SROW + 2 * pt);
    }
    __syncthreads();
  }
  if (!GRAM) return;
  double* Red = S;
  for (int w4 = 0; w4 < 4; ++w4) {
    if (wv == w4) {
#pragma unroll
      for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int rw = 16 * mi + 4 * v + lg, cl = 16 * ni + li;
            if (w4 == 0) {
              Red[rw * FC + cl] = T3.t[mi][ni][v];
              Red[FBLK + rw * FC + cl] = T4.t[mi][ni][v];
            } else {
              Red[rw * FC + cl] += T3.t[mi][ni][v];
              Red[FBLK + rw * FC + cl] += T4.t[mi][ni][v];
            }
          }
    }
    __syncthreads();
  }
  double* d3 = part + it.part + (long long)(2 * EIG_NPART + q) * FBLK;
  double* d4 = part + it.part + (long long)(3 * EIG_NPART + q) * FBLK;
  for (int o = threadIdx.x; o < FBLK; o += blockDim.x) {
    d3[o] = Red[o];
    d4[o] = Red[FBLK + o];
  }
}

// gram12, FP64: WG (256 threads) -> 4 NB consecutive Q blocks, wave -> NB.  Each
// block is loaded once (coalesced, one block of look-ahead in registers) into a
// per-wave LDS tile, from which both operand orientations (D for T'1, D^T for
// T'2) are read.  grid (EIG_NPART, items)
template <int NB>
__global__ __launch_bounds__(256) void k_gram12_f64(const EigItem* __restrict__ items,
                                                    const double* __restrict__ Q,
                                                    const double* __restrict__ lam,
                                                    double* __restrict__ part) {
  __shared__ double Db[4][FC * LD8];
  const EigItem it = items[blockIdx.y];
  const int q = blockIdx.x;
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int li = l & 15, lg = l >> 4;
  double lam1v[8], lam2v[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    lam1v[t] = lam[it.L[0] + 4 * t + lg];
    lam2v[t] = lam[it.L[1] + 4 * t + lg];
  }
  Acc32 T1, T2;
  acc_zero(T1);
  acc_zero(T2);
  const long long r0 = (long long)q * (4 * NB) + wv * NB;
  double* db = Db[wv];
  double2 nD[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) nD[j] = *reinterpret_cast<const double2*>(Q + it.Q + r0 * FBLK + 2 * l + 128 * j);
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int o = 2 * l + 128 * j, a = o >> 5, b = o & 31;
    db[a * LD8 + b] = nD[j].x;
    db[a * LD8 + b + 1] = nD[j].y;
  }
#pragma unroll 1
  for (int i = 0; i < NB; ++i) {
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    if (i + 1 < NB) {
#pragma unroll
      for (int j = 0; j < 8; ++j)
        nD[j] = *reinterpret_cast<const double2*>(Q + it.Q + (r0 + i + 1) * FBLK + 2 * l + 128 * j);
    }
    const long long qb = r0 + i;
    const int r3 = (int)((qb >> 1) & 31), r4 = (int)(qb >> 6);
    const double w34 = lam[it.L[2] + r3] * lam[it.L[3] + r4];
    double cr[2][8], cc[2][8];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        cr[h][t] = db[(16 * h + li) * LD8 + 4 * t + lg];  // D[16h+li][4t+lg]
        cc[h][t] = db[(4 * t + lg) * LD8 + 16 * h + li];  // D[4t+lg][16h+li]
      }
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
      for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni) {
          // T1 += D L2 D^T (K = r2), T2 += D^T L1 D (K = r1)
          T1.t[mi][ni] = MFMA_F64(cr[mi][t], w34 * lam2v[t] * cr[ni][t], T1.t[mi][ni]);
          T2.t[mi][ni] = MFMA_F64(cc[mi][t], w34 * lam1v[t] * cc[ni][t], T2.t[mi][ni]);
        }
    // every lane's reads of this block are issued before the tile is overwritten
    __builtin_amdgcn_wave_barrier();
    if (i + 1 < NB) {
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int o = 2 * l + 128 * j, a = o >> 5, b = o & 31;
        db[a * LD8 + b] = nD[j].x;
        db[a * LD8 + b + 1] = nD[j].y;
      }
    }
  }
  __syncthreads();
  double* Red = &Db[0][0];  // 4 * 32 * 33 >= 2 * 1024
  for (int w4 = 0; w4 < 4; ++w4) {
    if (wv == w4) {
#pragma unroll
      for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int rw = 16 * mi + 4 * v + lg, cl = 16 * ni + li;
            if (w4 == 0) {
              Red[rw * FC + cl] = T1.t[mi][ni][v];
              Red[FBLK + rw * FC + cl] = T2.t[mi][ni][v];
            } else {
              Red[rw * FC + cl] += T1.t[mi][ni][v];
              Red[FBLK + rw * FC + cl] += T2.t[mi][ni][v];
            }
          }
    }
    __syncthreads();
  }
  double* d1 = part + it.part + (long long)(0 * EIG_NPART + q) * FBLK;
  double* d2 = part + it.part + (long long)(1 * EIG_NPART + q) * FBLK;
  for (int o = threadIdx.x; o < FBLK; o += blockDim.x) {
    d1[o] = Red[o];
    d2[o] = Red[FBLK + o];
  }
}

// ---------------------------------------------------------------------------
// Degree-3 sites (lattice boundary), every chi = 32.  psi[s, b1, b2, b3] in
// layout C: block (s + 2 b3) = [b1][b2], 64 blocks.
//   rot12_d3: X -> U1^T X U2, in place of the block index (R3)
//   rot3_d3 : tile (s, r1): Z[b3][r2] = row r1 of R3 block (s + 2 b3); Y = U3^T Z
//             ([r3][r2]); T'3 += Y diag(l1[r1] l2[r2]) Y^T; row r1 of the Q3 block
//             (s + 2 r3) = Y[r3][:]
//   gram12<1> with l4 = 1 (a table of ones) gives T'1 and T'2.
// ---------------------------------------------------------------------------
constexpr int D3_NBLK = 2 * FC;  // blocks per degree-3 site

// grid (4, items), 256 threads: wave -> blocks 16 x + 4 wv + {0..3}
template <typename TI>
__global__ __launch_bounds__(256) void k_rot12_d3_bf16(const EigItem* __restrict__ items,
                                                       const __bf16* __restrict__ lay,
                                                       const float* __restrict__ U,
                                                       TI* __restrict__ R) {
  typedef typename Mma<TI>::v8 v8;
  const EigItem it = items[blockIdx.y];
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6, g = l >> 5, c = l & 31;
  v8 aU1h[2], aU1l[2];
  bf16x8 bU2h[2], bU2l[2];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      TI h1, l1;
      split16(U[it.U[0] + kperm(t, g, p) + FC * c], h1, l1);  // U1^T, K = b1 (accumulator rows)
      aU1h[t][p] = h1;
      aU1l[t][p] = l1;
      __bf16 h, lo;
      split_bf16(U[it.U[1] + (16 * t + 8 * g + p) + FC * c], h, lo);  // U2, K = b2
      bU2h[t][p] = h;
      bU2l[t][p] = lo;
    }
  bf16x8 aX[4][2];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const __bf16* X = lay + it.X + (long long)(16 * blockIdx.x + 4 * wv + i) * FBLK;
#pragma unroll
    for (int t = 0; t < 2; ++t) aX[i][t] = *reinterpret_cast<const bf16x8*>(X + c * FC + 16 * t + 8 * g);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    f32x16 V = {};
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      V = MFMA_BF16(aX[i][t], bU2h[t], V);
      V = MFMA_BF16(aX[i][t], bU2l[t], V);
    }
    v8 bV[2];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int p = 0; p < 8; ++p) bV[t][p] = (TI)V[8 * t + p];
    f32x16 Y = {};
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      Y = Mma<TI>::mfma(aU1h[t], bV[t], Y);
      Y = Mma<TI>::mfma(aU1l[t], bV[t], Y);
    }
    TI* dst = R + it.R + (long long)(16 * blockIdx.x + 4 * wv + i) * FBLK;
#pragma unroll
    for (int v = 0; v < 16; ++v) dst[(8 * (v >> 2) + 4 * g + (v & 3)) * FC + c] = (TI)Y[v];
  }
}

// grid (EIG_NPART, items), 256 threads: wave -> tile (s, r1) = (x & 1, x >> 1), x = 4 q + wv
template <typename TI>
__global__ __launch_bounds__(256) void k_rot3_d3_bf16(const EigItem* __restrict__ items,
                                                      const TI* __restrict__ R,
                                                      const float* __restrict__ U,
                                                      const float* __restrict__ lam,
                                                      TI* __restrict__ Q, float* __restrict__ part) {
  typedef typename Mma<TI>::v8 v8;
  __shared__ float Red[FBLK];
  const EigItem it = items[blockIdx.y];
  const int q = blockIdx.x;
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6, g = l >> 5, c = l & 31;
  const int x = 4 * q + wv, s = x & 1, r1 = x >> 1;
  v8 aU3h[2], aU3l[2], bZ[2];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      TI h, lo;
      split16(U[it.U[2] + (16 * t + 8 * g + p) + FC * c], h, lo);  // U3^T, K = b3 natural
      aU3h[t][p] = h;
      aU3l[t][p] = lo;
      // Z as B operand (lane r2, K = b3 natural)
      bZ[t][p] = R[it.R + (long long)(s + 2 * (16 * t + 8 * g + p)) * FBLK + r1 * FC + c];
    }
  f32x16 Y = {}, YT = {};
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    Y = Mma<TI>::mfma(aU3h[t], bZ[t], Y);
    Y = Mma<TI>::mfma(aU3l[t], bZ[t], Y);
    YT = Mma<TI>::mfma(bZ[t], aU3h[t], YT);
    YT = Mma<TI>::mfma(bZ[t], aU3l[t], YT);
  }
  const float l1 = lam[it.L[0] + r1];
  f32x16 T3 = {};
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    v8 a, b;
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      const float yt = YT[8 * t + p];
      a[p] = (TI)yt;
      b[p] = (TI)(l1 * lam[it.L[1] + kperm(t, g, p)] * yt);
    }
    T3 = Mma<TI>::mfma(a, b, T3);
  }
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const int r3 = 8 * (v >> 2) + 4 * g + (v & 3);
    Q[it.Q + (long long)(s + 2 * r3) * FBLK + r1 * FC + c] = (TI)Y[v];
  }
  for (int w4 = 0; w4 < 4; ++w4) {
    if (wv == w4) {
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int rw = 8 * (v >> 2) + 4 * g + (v & 3);
        if (w4 == 0) Red[rw * FC + c] = T3[v];
        else Red[rw * FC + c] += T3[v];
      }
    }
    __syncthreads();
  }
  float* d3 = part + it.part + (long long)(2 * EIG_NPART + q) * FBLK;
  for (int o = threadIdx.x; o < FBLK; o += blockDim.x) d3[o] = Red[o];
}

// ---------------------------------------------------------------------------
// Compensated flavours of the eigenbasis pipeline: R and Q in FP32, and every
// MFMA operand that is not exact in 16 bits carried as N terms of a 16-bit
// type (x = t0 + t1 + ... to FP32 rounding: t0 the rounding of x, every later
// term the rounding of what the earlier ones leave; the differences are exact
// in FP32).  A product a b takes the term products a_i b_j with i + j <= N - 1,
// summed in the FP32 MFMA accumulator; the site tensors enter unsplit.
//   PrecX3: three BF16 terms (about 24 significant bits); six products, three
//           with an exact operand.
//   PrecH2: two FP16 terms (about 22 significant bits); three products, two
//           with an exact operand.  FP16's 5-bit exponent suffices because the
//           16-bit layouts hold every site scaled by a power of two that puts
//           its largest entry in [1/2, 1) and the rotations are orthogonal;
//           a remainder below the smallest normal FP16 keeps an absolute error
//           of 2^-25 against operands of order one.  The BF16 site tensor
//           entries convert to FP16 exactly down to 2^-17.
// ---------------------------------------------------------------------------
constexpr int SX_LD = 33;  // FP32 staging row stride (odd: conflict-free column writes)

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));

struct PrecX3 {
  typedef __bf16 T;
  typedef bf16x8 v8;
  typedef bf16x2 v2;
  static constexpr int N = 3;
  static __device__ inline f32x16 mma(v8 a, v8 b, f32x16 c) { return MFMA_BF16(a, b, c); }
};
struct PrecH2 {
  typedef _Float16 T;
  typedef f16x8 v8;
  typedef f16x2 v2;
  static constexpr int N = 2;
  static __device__ inline f32x16 mma(v8 a, v8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
  }
};

template <class P>
struct Terms {
  typename P::v8 t[P::N];
};

// the two BF16 halves of a packed pair, widened to FP32 (exact)
__device__ inline f32x2 widen_bf16x2(bf16x2 v) {
  const unsigned u = __builtin_bit_cast(unsigned, v);
  f32x2 r;
  r.x = __builtin_bit_cast(float, u << 16);
  r.y = __builtin_bit_cast(float, u & 0xffff0000u);
  return r;
}

// N-term split of eight values, two at a time: every rounding is one packed
// conversion (round to nearest even) whose result is already the operand's
// register pair
template <class P>
__device__ inline Terms<P> split8(const float (&x)[8]) {
  Terms<P> o;
#pragma unroll
  for (int p = 0; p < 8; p += 2) {
    f32x2 r;
    r.x = x[p];
    r.y = x[p + 1];
#pragma unroll
    for (int i = 0; i < P::N; ++i) {
      const typename P::v2 h = __builtin_convertvector(r, typename P::v2);
      o.t[i][p] = h.x;
      o.t[i][p + 1] = h.y;
      if (i + 1 < P::N) {
        if constexpr (P::N == 3)
          r = r - widen_bf16x2(h);
        else
          r = r - __builtin_convertvector(h, f32x2);
      }
    }
  }
  return o;
}

// eight BF16 site tensor entries as an exact operand of the policy's MFMA
template <class P>
__device__ inline typename P::v8 exact8(bf16x8 v) {
  if constexpr (P::N == 3) {
    return v;
  } else {
    typename P::v8 o;
#pragma unroll
    for (int p = 0; p < 8; p += 2) {
      bf16x2 b;
      b.x = v[p];
      b.y = v[p + 1];
      const f16x2 h = __builtin_convertvector(widen_bf16x2(b), f16x2);
      o[p] = h.x;
      o[p + 1] = h.y;
    }
    return o;
  }
}

// every term product of order <= N - 1, smallest first
template <class P>
__device__ inline f32x16 mmN(const Terms<P>& a, const Terms<P>& b, f32x16 acc) {
  if constexpr (P::N == 3) {
    acc = P::mma(a.t[2], b.t[0], acc);
    acc = P::mma(a.t[1], b.t[1], acc);
    acc = P::mma(a.t[0], b.t[2], acc);
    acc = P::mma(a.t[1], b.t[0], acc);
    acc = P::mma(a.t[0], b.t[1], acc);
  } else {
    acc = P::mma(a.t[1], b.t[0], acc);
    acc = P::mma(a.t[0], b.t[1], acc);
  }
  return P::mma(a.t[0], b.t[0], acc);
}

// exact a times split b
template <class P>
__device__ inline f32x16 mm_exactN(typename P::v8 a, const Terms<P>& b, f32x16 acc) {
#pragma unroll
  for (int i = P::N - 1; i >= 0; --i) acc = P::mma(a, b.t[i], acc);
  return acc;
}

// split a times exact b
template <class P>
__device__ inline f32x16 mm_exactN_r(const Terms<P>& a, typename P::v8 b, f32x16 acc) {
#pragma unroll
  for (int i = P::N - 1; i >= 0; --i) acc = P::mma(a.t[i], b, acc);
  return acc;
}

// v with the sign bits of mask flipped (exact negation of selected elements)
template <typename V>
__device__ inline V flip_signs(V v, uint4 m) {
  uint4 u = __builtin_bit_cast(uint4, v);
  u.x ^= m.x;
  u.y ^= m.y;
  u.z ^= m.z;
  u.w ^= m.w;
  return __builtin_bit_cast(V, u);
}

// Weighted Gram term acc += sum_k w_k a[m][k] a[n][k] with one split: the
// A operand is sqrt|w_k| a[.][k] (split into N terms) and the B operand the
// same terms with the sign of w_k (mask) applied, so b = sign(w) sqrt|w| a exactly
template <class P>
__device__ inline f32x16 wgramN(const float (&aw)[8], uint4 mask, f32x16 acc) {
  const Terms<P> a = split8<P>(aw);
  Terms<P> b;
#pragma unroll
  for (int i = 0; i < P::N; ++i) b.t[i] = flip_signs(a.t[i], mask);
  return mmN<P>(a, b, acc);
}

// rot12: R block (s + 2 (x' + 32 y')) row z, columns w = (U1^T X U2)[x'][y'] of
// the 32 layout-B blocks X = (s + 2 (z + 32 w)).  Items: the absorb kernel's
// (psi_in = layout B, P = R offset, Ma = U1, Mb = U2).
// grid (64, items), 512 threads: WG -> (s, z) = (q & 1, q >> 1), wave -> w = 4 wv + {0..3}
template <class P>
__global__ __launch_bounds__(512) void k_rot12_s(const FastItem* __restrict__ items,
                                                  const __bf16* __restrict__ lay,
                                                  const float* __restrict__ U, float* __restrict__ R) {
  __shared__ float S[FBLK * SX_LD];  // row (y' + 32 x'), column w
  const FastItem it = items[blockIdx.y];
  const int q = blockIdx.x;
  const int s = q & 1, z = q >> 1;
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6, g = l >> 5, c = l & 31;
  Terms<P> bU2[2], aU1[2];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    float u2[8], u1[8];
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      u2[p] = U[it.Mb + (16 * t + 8 * g + p) + FC * c];  // U2, K = b2 natural
      u1[p] = U[it.Ma + kperm(t, g, p) + FC * c];        // U1^T, K = b1 in accumulator-row order
    }
    bU2[t] = split8<P>(u2);
    aU1[t] = split8<P>(u1);
  }
  typename P::v8 aXs[4][2];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const __bf16* X = lay + it.psi_in + (long long)(s + 2 * (z + FC * (4 * wv + i))) * FBLK;
#pragma unroll
    for (int t = 0; t < 2; ++t)
      aXs[i][t] = exact8<P>(*reinterpret_cast<const bf16x8*>(X + c * FC + 16 * t + 8 * g));
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int w = 4 * wv + i;
    f32x16 V = {};
#pragma unroll
    for (int t = 0; t < 2; ++t) V = mm_exactN<P>(aXs[i][t], bU2[t], V);
    f32x16 Y = {};
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      float v8[8];
#pragma unroll
      for (int p = 0; p < 8; ++p) v8[p] = V[8 * t + p];
      Y = mmN<P>(aU1[t], split8<P>(v8), Y);
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int xb = 8 * (v >> 2) + 4 * g + (v & 3);
      S[(c + FC * xb) * SX_LD + w] = Y[v];
    }
    __builtin_amdgcn_sched_barrier(0);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < FBLK * 8; k += blockDim.x) {
    const int o = k >> 3, pt = k & 7;
    const int yb = o & 31, xb = o >> 5;
    const long long r = s + 2 * (xb + FC * yb);
    const float* src = S + o * SX_LD + 4 * pt;
    *reinterpret_cast<float4*>(R + it.P + r * FBLK + z * FC + 4 * pt) = make_float4(src[0], src[1], src[2], src[3]);
  }
}

// rot34 with the Grams T'3, T'4 (the compensated twin of k_rot34_bf16<true>).
// grid (EIG_NPART, items), 512 threads: WG -> s = q & 1, r1 = z in
// [ROT_NZ (q >> 1), +ROT_NZ); per z the 32 R blocks (s + 2 (z + 32 w)) in two
// halves, wave wv -> w = 16 h + 2 wv + {0, 1}; the next half's two blocks are in
// flight while the current two are multiplied.  Y goes to the FP32 staging
// tile at column w ^ r3 (a swizzle that keeps the column writes, the row-wise
// copy-out and the transposed read-back conflict-free); the wave reads its
// block back transposed for T'3 instead of forming Y^T on the MFMA.
template <class P>
__global__ __launch_bounds__(512) void k_rot34_s(const EigItem* __restrict__ items, const float* __restrict__ R,
                                                  const float* __restrict__ U, const float* __restrict__ lam,
                                                  float* __restrict__ Q, float* __restrict__ part) {
  __shared__ float S[FBLK * SX_LD];  // row (r4 + 32 r3), column r2 ^ r3
  // sqrt|l3|, sqrt|l4| and the sign masks of l3, l4 for the K order of the Gram
  // operands (rows kperm(t, g, 0..7)), per (bond, t, g)
  __shared__ __attribute__((aligned(16))) float sq34[2][FC];
  __shared__ __attribute__((aligned(16))) uint4 msk34[2][2][2];
  const EigItem* itp = items + blockIdx.y;
  const int q = blockIdx.x;
  const int s = q & 1, z0 = (q >> 1) * ROT_NZ;
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6, g = l >> 5, c = l & 31;
  const long long Ro = itp->R, Qo = itp->Q, L0 = itp->L[0], L1 = itp->L[1];
  if (threadIdx.x < 2 * FC) {
    const float lv = lam[((threadIdx.x >> 5) ? itp->L[3] : itp->L[2]) + (threadIdx.x & 31)];
    sq34[threadIdx.x >> 5][threadIdx.x & 31] = sqrtf(fabsf(lv));
  }
  if (threadIdx.x < 8) {
    const int bd = threadIdx.x >> 2, t = (threadIdx.x >> 1) & 1, gg = threadIdx.x & 1;
    const long long lb = bd ? itp->L[3] : itp->L[2];
    unsigned mw[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float lo = lam[lb + kperm(t, gg, 2 * j)], hi = lam[lb + kperm(t, gg, 2 * j + 1)];
      mw[j] = (lo < 0.f ? 0x8000u : 0u) | (hi < 0.f ? 0x80000000u : 0u);
    }
    msk34[bd][t][gg] = make_uint4(mw[0], mw[1], mw[2], mw[3]);
  }
  Terms<P> aU3[2], bU4[2];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    float u3[8], u4[8];
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      u3[p] = U[itp->U[2] + kperm(t, g, p) + FC * c];
      u4[p] = U[itp->U[3] + (16 * t + 8 * g + p) + FC * c];
    }
    aU3[t] = split8<P>(u3);
    bU4[t] = split8<P>(u4);
  }
  __syncthreads();
  f32x16 T3 = {}, T4 = {};
  float4 cur[2][4], nxt[2][4];
  auto fetch = [&](int step, float4 (&dst)[2][4]) {
    const int z = z0 + (step >> 1), h = step & 1;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int w = 16 * h + 2 * wv + i;
      const float* Z = R + Ro + (long long)(s + 2 * (z + FC * w)) * FBLK + c * FC + 8 * g;
      dst[i][0] = *reinterpret_cast<const float4*>(Z);
      dst[i][1] = *reinterpret_cast<const float4*>(Z + 4);
      dst[i][2] = *reinterpret_cast<const float4*>(Z + 16);
      dst[i][3] = *reinterpret_cast<const float4*>(Z + 20);
    }
  };
  constexpr int NSTEP = 2 * ROT_NZ;
  fetch(0, cur);
#pragma unroll 1
  for (int step = 0; step < NSTEP; ++step) {
    const int z = z0 + (step >> 1), h = step & 1;
    if (step + 1 < NSTEP) fetch(step + 1, nxt);
    const float l1 = lam[L0 + z];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int w = 16 * h + 2 * wv + i;
      const float w12 = l1 * lam[L1 + w];
      const float sw12 = sqrtf(fabsf(w12));
      const unsigned nw = w12 < 0.f ? 0x80008000u : 0u;
      f32x16 V = {};
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const float x8[8] = {cur[i][2 * t].x, cur[i][2 * t].y, cur[i][2 * t].z, cur[i][2 * t].w,
                             cur[i][2 * t + 1].x, cur[i][2 * t + 1].y, cur[i][2 * t + 1].z, cur[i][2 * t + 1].w};
        V = mmN<P>(split8<P>(x8), bU4[t], V);
      }
      f32x16 Y = {};
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        float v8[8];
#pragma unroll
        for (int p = 0; p < 8; ++p) v8[p] = V[8 * t + p];
        Y = mmN<P>(aU3[t], split8<P>(v8), Y);
      }
      // Y[r3][r4] (lane r4, rows r3 = xb) to the tile
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int xb = 8 * (v >> 2) + 4 * g + (v & 3);
        S[(c + FC * xb) * SX_LD + (w ^ xb)] = Y[v];
      }
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const float4 s3a = *reinterpret_cast<const float4*>(&sq34[0][16 * t + 4 * g]);
        const float4 s3b = *reinterpret_cast<const float4*>(&sq34[0][16 * t + 8 + 4 * g]);
        const float4 s4a = *reinterpret_cast<const float4*>(&sq34[1][16 * t + 4 * g]);
        const float4 s4b = *reinterpret_cast<const float4*>(&sq34[1][16 * t + 8 + 4 * g]);
        const float s3[8] = {s3a.x, s3a.y, s3a.z, s3a.w, s3b.x, s3b.y, s3b.z, s3b.w};
        const float s4[8] = {s4a.x, s4a.y, s4a.z, s4a.w, s4b.x, s4b.y, s4b.z, s4b.w};
        uint4 m3 = msk34[0][t][g], m4 = msk34[1][t][g];
        m3.x ^= nw; m3.y ^= nw; m3.z ^= nw; m3.w ^= nw;
        m4.x ^= nw; m4.y ^= nw; m4.z ^= nw; m4.w ^= nw;
        float a8[8];
        // T'3: weights w12 l4 over the K index r4 = kperm(t, g, p) of the transposed
        // block, read back from the tile (lane r3 = c)
#pragma unroll
        for (int p = 0; p < 8; ++p)
          a8[p] = sw12 * s4[p] * S[(kperm(t, g, p) + FC * c) * SX_LD + (w ^ c)];
        T3 = wgramN<P>(a8, m4, T3);
        // T'4: weights w12 l3 over the K index r3 of Y
#pragma unroll
        for (int p = 0; p < 8; ++p) a8[p] = sw12 * s3[p] * Y[8 * t + p];
        T4 = wgramN<P>(a8, m3, T4);
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    if (h == 1) {
      __syncthreads();
      for (int k = threadIdx.x; k < FBLK * 8; k += blockDim.x) {
        const int o = k >> 3, pt = k & 7;
        const int yb = o & 31, xb = o >> 5;
        const long long qb = s + 2 * (xb + FC * yb);
        const float* src = S + o * SX_LD;
        const int c0 = 4 * pt;
        *reinterpret_cast<float4*>(Q + Qo + qb * FBLK + z * FC + c0) =
            make_float4(src[c0 ^ xb], src[(c0 + 1) ^ xb], src[(c0 + 2) ^ xb], src[(c0 + 3) ^ xb]);
      }
      __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) cur[i][j] = nxt[i][j];
  }
  float* Red = S;
  for (int w8 = 0; w8 < 8; ++w8) {
    if (wv == w8) {
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int rw = 8 * (v >> 2) + 4 * g + (v & 3);
        if (w8 == 0) {
          Red[rw * FC + c] = T3[v];
          Red[FBLK + rw * FC + c] = T4[v];
        } else {
          Red[rw * FC + c] += T3[v];
          Red[FBLK + rw * FC + c] += T4[v];
        }
      }
    }
    __syncthreads();
  }
  float* d3 = part + itp->part + (long long)(2 * EIG_NPART + q) * FBLK;
  float* d4 = part + itp->part + (long long)(3 * EIG_NPART + q) * FBLK;
  for (int o = threadIdx.x; o < FBLK; o += blockDim.x) {
    d3[o] = Red[o];
    d4[o] = Red[FBLK + o];
  }
}

// gram12 (T'1, T'2) of the FP32 Q blocks; the layout of k_gram12_bf16, two
// blocks of look-ahead per wave
template <class P, int NB>
__global__ __launch_bounds__(256) void k_gram12_s(const EigItem* __restrict__ items, const float* __restrict__ Q,
                                                   const float* __restrict__ lam, float* __restrict__ part) {
  __shared__ float Red[2 * FBLK];
  const EigItem* itp = items + blockIdx.y;
  const int q = blockIdx.x;
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6, g = l >> 5, c = l & 31;
  const long long Qo = itp->Q, L2 = itp->L[2], L3 = itp->L[3];
  // T'2 weights l1 over the K index r1 (rows kperm(t, g, p) of D^T): sqrt|l1| and
  // the sign masks of the single-split Gram
  float sl1v[16], lam2v[16];
  uint4 m1[2];
  typename P::v8 bI[2];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    unsigned mw[4] = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      const float l1 = lam[itp->L[0] + kperm(t, g, p)];
      sl1v[8 * t + p] = sqrtf(fabsf(l1));
      if (l1 < 0.f) mw[p >> 1] |= (p & 1) ? 0x80000000u : 0x8000u;
      lam2v[8 * t + p] = lam[itp->L[1] + 16 * t + 8 * g + p];
      bI[t][p] = (typename P::T)((16 * t + 8 * g + p) == c ? 1.0f : 0.0f);
    }
    m1[t] = make_uint4(mw[0], mw[1], mw[2], mw[3]);
  }
  f32x16 T1 = {}, T2 = {};
  constexpr int PF = NB < 2 ? NB : 2;
  const long long r0 = (long long)q * (4 * NB) + wv * NB;
  float4 cur[PF][4], nxt[PF][4];
  auto fetch = [&](long long r, float4 (&dst)[PF][4]) {
#pragma unroll
    for (int i = 0; i < PF; ++i) {
      const float* D = Q + Qo + (r + i) * FBLK + c * FC + 8 * g;
      dst[i][0] = *reinterpret_cast<const float4*>(D);
      dst[i][1] = *reinterpret_cast<const float4*>(D + 4);
      dst[i][2] = *reinterpret_cast<const float4*>(D + 16);
      dst[i][3] = *reinterpret_cast<const float4*>(D + 20);
    }
  };
  fetch(r0, cur);
#pragma unroll 1
  for (int i0 = 0; i0 < NB; i0 += PF) {
    if (i0 + PF < NB) fetch(r0 + i0 + PF, nxt);
#pragma unroll
    for (int i = 0; i < PF; ++i) {
      const long long qb = r0 + i0 + i;
      const int r3 = (int)((qb >> 1) & 31), r4 = (int)(qb >> 6);
      const float w34 = lam[L2 + r3] * lam[L3 + r4];
      f32x16 Dt = {};
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const float d8[8] = {cur[i][2 * t].x, cur[i][2 * t].y, cur[i][2 * t].z, cur[i][2 * t].w,
                             cur[i][2 * t + 1].x, cur[i][2 * t + 1].y, cur[i][2 * t + 1].z, cur[i][2 * t + 1].w};
        float b8[8];
#pragma unroll
        for (int p = 0; p < 8; ++p) b8[p] = w34 * lam2v[8 * t + p] * d8[p];
        const Terms<P> a = split8<P>(d8);
        T1 = mmN<P>(a, split8<P>(b8), T1);
        Dt = mm_exactN_r<P>(a, bI[t], Dt);  // D I: the transpose of the split D
      }
      const float sw34 = sqrtf(fabsf(w34));
      const unsigned nw = w34 < 0.f ? 0x80008000u : 0u;
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        float a8[8];
#pragma unroll
        for (int p = 0; p < 8; ++p) a8[p] = sw34 * sl1v[8 * t + p] * Dt[8 * t + p];
        const uint4 mk = make_uint4(m1[t].x ^ nw, m1[t].y ^ nw, m1[t].z ^ nw, m1[t].w ^ nw);
        T2 = wgramN<P>(a8, mk, T2);
      }
      __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int i = 0; i < PF; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) cur[i][j] = nxt[i][j];
  }
  for (int w4 = 0; w4 < 4; ++w4) {
    if (wv == w4) {
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int rw = 8 * (v >> 2) + 4 * g + (v & 3);
        if (w4 == 0) {
          Red[rw * FC + c] = T1[v];
          Red[FBLK + rw * FC + c] = T2[v];
        } else {
          Red[rw * FC + c] += T1[v];
          Red[FBLK + rw * FC + c] += T2[v];
        }
      }
    }
    __syncthreads();
  }
  float* d1 = part + itp->part + (long long)(0 * EIG_NPART + q) * FBLK;
  float* d2 = part + itp->part + (long long)(1 * EIG_NPART + q) * FBLK;
  for (int o = threadIdx.x; o < FBLK; o += blockDim.x) {
    d1[o] = Red[o];
    d2[o] = Red[FBLK + o];
  }
}

// degree-3 sites, compensated: rot12_d3 and rot3_d3 of the 16-bit kernels on
// FP32 R3 / Q3 blocks
template <class P>
__global__ __launch_bounds__(256) void k_rot12_d3_s(const EigItem* __restrict__ items,
                                                     const __bf16* __restrict__ lay,
                                                     const float* __restrict__ U, float* __restrict__ R) {
  const EigItem* itp = items + blockIdx.y;
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6, g = l >> 5, c = l & 31;
  Terms<P> aU1[2], bU2[2];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    float u1[8], u2[8];
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      u1[p] = U[itp->U[0] + kperm(t, g, p) + FC * c];
      u2[p] = U[itp->U[1] + (16 * t + 8 * g + p) + FC * c];
    }
    aU1[t] = split8<P>(u1);
    bU2[t] = split8<P>(u2);
  }
  typename P::v8 aX[4][2];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const __bf16* X = lay + itp->X + (long long)(16 * blockIdx.x + 4 * wv + i) * FBLK;
#pragma unroll
    for (int t = 0; t < 2; ++t)
      aX[i][t] = exact8<P>(*reinterpret_cast<const bf16x8*>(X + c * FC + 16 * t + 8 * g));
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    f32x16 V = {};
#pragma unroll
    for (int t = 0; t < 2; ++t) V = mm_exactN<P>(aX[i][t], bU2[t], V);
    f32x16 Y = {};
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      float v8[8];
#pragma unroll
      for (int p = 0; p < 8; ++p) v8[p] = V[8 * t + p];
      Y = mmN<P>(aU1[t], split8<P>(v8), Y);
    }
    float* dst = R + itp->R + (long long)(16 * blockIdx.x + 4 * wv + i) * FBLK;
#pragma unroll
    for (int v = 0; v < 16; ++v) dst[(8 * (v >> 2) + 4 * g + (v & 3)) * FC + c] = Y[v];
  }
}

template <class P>
__global__ __launch_bounds__(256) void k_rot3_d3_s(const EigItem* __restrict__ items, const float* __restrict__ R,
                                                    const float* __restrict__ U, const float* __restrict__ lam,
                                                    float* __restrict__ Q, float* __restrict__ part) {
  __shared__ float Red[FBLK];
  const EigItem* itp = items + blockIdx.y;
  const int q = blockIdx.x;
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6, g = l >> 5, c = l & 31;
  const int x = 4 * q + wv, s = x & 1, r1 = x >> 1;
  Terms<P> aU3[2], bZ[2];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    float u3[8], z8[8];
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      u3[p] = U[itp->U[2] + (16 * t + 8 * g + p) + FC * c];  // U3^T, K = b3 natural
      z8[p] = R[itp->R + (long long)(s + 2 * (16 * t + 8 * g + p)) * FBLK + r1 * FC + c];  // lane r2, K = b3
    }
    aU3[t] = split8<P>(u3);
    bZ[t] = split8<P>(z8);
  }
  f32x16 Y = {}, YT = {};
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    Y = mmN<P>(aU3[t], bZ[t], Y);
    YT = mmN<P>(bZ[t], aU3[t], YT);
  }
  const float l1 = lam[itp->L[0] + r1];
  f32x16 T3 = {};
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    float a8[8], b8[8];
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      a8[p] = YT[8 * t + p];
      b8[p] = l1 * lam[itp->L[1] + kperm(t, g, p)] * a8[p];
    }
    T3 = mmN<P>(split8<P>(a8), split8<P>(b8), T3);
  }
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const int r3 = 8 * (v >> 2) + 4 * g + (v & 3);
    Q[itp->Q + (long long)(s + 2 * r3) * FBLK + r1 * FC + c] = Y[v];
  }
  for (int w4 = 0; w4 < 4; ++w4) {
    if (wv == w4) {
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int rw = 8 * (v >> 2) + 4 * g + (v & 3);
        if (w4 == 0) Red[rw * FC + c] = T3[v];
        else Red[rw * FC + c] += T3[v];
      }
    }
    __syncthreads();
  }
  float* d3 = part + itp->part + (long long)(2 * EIG_NPART + q) * FBLK;
  for (int o = threadIdx.x; o < FBLK; o += blockDim.x) d3[o] = Red[o];
}

// FP64 degree-3 kernels, same decomposition on v_mfma_f64_16x16x4_f64
__global__ __launch_bounds__(256) void k_rot12_d3_f64(const EigItem* __restrict__ items,
                                                      const double* __restrict__ lay,
                                                      const double* __restrict__ U,
                                                      double* __restrict__ R) {
  const EigItem it = items[blockIdx.y];
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int li = l & 15, lg = l >> 4;
  double bU2[8][2], aU1T[2][8];
#pragma unroll
  for (int t = 0; t < 8; ++t)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      bU2[t][h] = U[it.U[1] + (4 * t + lg) + FC * (16 * h + li)];
      aU1T[h][t] = U[it.U[0] + (4 * t + lg) + FC * (16 * h + li)];
    }
  double cur[2][8], nxt[2][8];
  auto fetch = [&](int i, double (&dst)[2][8]) {
    const double* X = lay + it.X + (long long)(16 * blockIdx.x + 4 * wv + i) * FBLK;
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int t = 0; t < 8; ++t) dst[h][t] = X[(16 * h + li) * FC + 4 * t + lg];
  };
  fetch(0, cur);
#pragma unroll 1
  for (int i = 0; i < 4; ++i) {
    if (i + 1 < 4) fetch(i + 1, nxt);
    Acc32 V, Y;
    acc_zero(V);
    acc_zero(Y);
    mm32(V, cur, bU2);
    mm32_accb(Y, aU1T, V);
    double* dst = R + it.R + (long long)(16 * blockIdx.x + 4 * wv + i) * FBLK;
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
      for (int ni = 0; ni < 2; ++ni)
#pragma unroll
        for (int v = 0; v < 4; ++v) dst[(16 * mi + 4 * v + lg) * FC + 16 * ni + li] = Y.t[mi][ni][v];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int t = 0; t < 8; ++t) cur[h][t] = nxt[h][t];
  }
}

__global__ __launch_bounds__(256) void k_rot3_d3_f64(const EigItem* __restrict__ items,
                                                     const double* __restrict__ R,
                                                     const double* __restrict__ U,
                                                     const double* __restrict__ lam,
                                                     double* __restrict__ Q, double* __restrict__ part) {
  __shared__ double Red[FBLK];
  const EigItem it = items[blockIdx.y];
  const int q = blockIdx.x;
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int li = l & 15, lg = l >> 4;
  const int x = 4 * q + wv, s = x & 1, r1 = x >> 1;
  double aU3T[2][8], bZ[8][2], wv2[8];
  const double l1 = lam[it.L[0] + r1];
#pragma unroll
  for (int t = 0; t < 8; ++t) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      aU3T[h][t] = U[it.U[2] + (4 * t + lg) + FC * (16 * h + li)];                        // U3[4t+lg][16h+li]
      bZ[t][h] = R[it.R + (long long)(s + 2 * (4 * t + lg)) * FBLK + r1 * FC + 16 * h + li];  // Z[4t+lg][16h+li]
    }
    wv2[t] = l1 * lam[it.L[1] + 4 * t + lg];
  }
  Acc32 Y, YT, T3;
  acc_zero(Y);
  acc_zero(YT);
  acc_zero(T3);
#pragma unroll
  for (int t = 0; t < 8; ++t)
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
      for (int ni = 0; ni < 2; ++ni) {
        Y.t[mi][ni] = MFMA_F64(aU3T[mi][t], bZ[t][ni], Y.t[mi][ni]);   // Y = U3^T Z
        YT.t[mi][ni] = MFMA_F64(bZ[t][mi], aU3T[ni][t], YT.t[mi][ni]);  // Y^T = Z^T U3
      }
#pragma unroll
  for (int t = 0; t < 8; ++t)
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
      for (int ni = 0; ni < 2; ++ni)
        T3.t[mi][ni] = MFMA_F64(YT.t[t >> 2][mi][t & 3], wv2[t] * YT.t[t >> 2][ni][t & 3], T3.t[mi][ni]);
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < 2; ++ni)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r3 = 16 * mi + 4 * v + lg;
        Q[it.Q + (long long)(s + 2 * r3) * FBLK + r1 * FC + 16 * ni + li] = Y.t[mi][ni][v];
      }
  for (int w4 = 0; w4 < 4; ++w4) {
    if (wv == w4) {
#pragma unroll
      for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int rw = 16 * mi + 4 * v + lg, cl = 16 * ni + li;
            if (w4 == 0) Red[rw * FC + cl] = T3.t[mi][ni][v];
            else Red[rw * FC + cl] += T3.t[mi][ni][v];
          }
    }
    __syncthreads();
  }
  double* d3 = part + it.part + (long long)(2 * EIG_NPART + q) * FBLK;
  for (int o = threadIdx.x; o < FBLK; o += blockDim.x) d3[o] = Red[o];
}

// Finaliser of the eigenbasis pipeline: T'_c = sum of the partial Grams (FP64),
// T_c = U_c T'_c U_c^T, symmetrised, normalised, message_diff / residual
// against the previous message.  grid (4, items), 256 threads.
template <typename TP, typename TM>
__global__ __launch_bounds__(256) void k_eig_finalize(const EigItem* __restrict__ items,
                                                      const TP* __restrict__ part,
                                                      const double* __restrict__ U64, const TM* msg_in,
                                                      TM* msg_out, MsgStat* __restrict__ stats,
                                                      int normalize, const int* __restrict__ sexp,
                                                      const int* __restrict__ lexp) {
  constexpr int LD = 33;
  __shared__ double Tp[FC * LD], Us[FC * LD], X[FC * LD];
  __shared__ double red[8];
  const EigItem* itp = items + blockIdx.y;  // fields indexed by the bond: read in place
  const int cb = blockIdx.x;
  const TP* pp = part + itp->part + (long long)cb * EIG_NPART * FBLK;
  const long long uoff = itp->U[cb];
  for (int o = threadIdx.x; o < FBLK; o += blockDim.x) {
    double sum = 0;
    for (int ch = 0; ch < EIG_NPART; ++ch) sum += (double)pp[(long long)ch * FBLK + o];
    Tp[(o >> 5) * LD + (o & 31)] = sum;  // T'[r][r']
    const double u = U64[uoff + o];   // U[b + 32 r]
    Us[(o & 31) * LD + (o >> 5)] = u;    // Us[b][r]
  }
  __syncthreads();
  for (int o = threadIdx.x; o < FBLK; o += blockDim.x) {
    const int b = o >> 5, rr = o & 31;
    double acc = 0;
    for (int r = 0; r < FC; ++r) acc += Us[b * LD + r] * Tp[r * LD + rr];
    X[b * LD + rr] = acc;  // (U T')[b][r']
  }
  __syncthreads();
  for (int o = threadIdx.x; o < FBLK; o += blockDim.x) {
    const int b = o >> 5, bb = o & 31;
    double acc = 0;
    for (int r = 0; r < FC; ++r) acc += X[b * LD + r] * Us[bb * LD + r];
    Tp[b * LD + bb] = acc;  // T[b][b']
  }
  __syncthreads();
  double us = 0.5 * site_unscale2(sexp, itp->site);
  if (lexp) {
    // the eigenvalue scales of the bonds that weigh this Gram (L[3] of a
    // degree-3 site is the table of ones, scale 0)
    int x = 0;
    for (int b = 0; b < 4; ++b)
      if (b != cb) x += lexp[itp->L[b] / FC];
    us = ldexp(us, x);
  }
  double n2 = 0;
  for (int o = threadIdx.x; o < FBLK; o += blockDim.x) {
    const int i = o & 31, j = o >> 5;  // message layout M[i + 32 j]
    const double tv = us * (Tp[i * LD + j] + Tp[j * LD + i]);
    X[i * LD + j] = tv;
    n2 += tv * tv;
  }
  auto bsum = [&](double v) {
    for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += red[i];
    return s;
  };
  n2 = bsum(n2);
  const double nrm = sqrt(n2);
  const double sc = (normalize && nrm != 0.0) ? 1.0 / nrm : 1.0;
  const long long off = itp->out[cb];
  const int slot = itp->slot[cb];
  double dot = 0, o2 = 0, d2 = 0, bad = 0;
  for (int o = threadIdx.x; o < FBLK; o += blockDim.x) {
    const double nv = X[(o & 31) * LD + (o >> 5)] * sc;
    const double ov = (double)msg_in[off + o];
    if (!isfinite(nv)) bad = 1;
    dot += ov * nv;
    o2 += ov * ov;
    d2 += (nv - ov) * (nv - ov);
    msg_out[off + o] = (TM)nv;
  }
  dot = bsum(dot);
  o2 = bsum(o2);
  d2 = bsum(d2);
  bad = bsum(bad);
  if (threadIdx.x == 0 && slot >= 0) {
    const double nn2 = n2 * sc * sc;
    MsgStat st;
    st.diff = 1.0 - (dot * dot) / (o2 * nn2);
    st.res2 = nn2 > 0 ? d2 / nn2 : 0.0;
    st.norm = nrm;
    st.bad = (bad > 0 || !isfinite(nrm)) ? 1.0 : 0.0;
    stats[slot] = st;
  }
}

// ---------------------------------------------------------------------------
// One-site reduced density matrices (environment of a single vertex contracted
// with its ket and bra, src/caches/beliefpropagationcache.jl:100-105 followed by
// the open-index contraction of src/expect.jl:5-19).  With every incoming
// message symmetric, M_c = U_c diag(lam_c) U_c^T and
//   rho[s, s'] = sum_r prod_c lam_c(r_c) Y[s, r] Y[s', r],
// Y = psi rotated into the eigenbasis of every incoming message: the Q buffer of
// the eigenbasis pipeline.  Pair p of Q blocks holds s = 0, 1 of the same
// (r3, r4) = (p & 31, p >> 5) (bulk sites, 1024 pairs) or r3 = p (degree-3 sites,
// 32 pairs, lam_4 = a table of ones).  FP64 throughout.
// ---------------------------------------------------------------------------
constexpr int RHO_NPART = 16;  // partial sums per site (workgroups)

// grid (RHO_NPART, items), 256 threads: wave gw = 4 q + wv takes pairs gw,
// gw + 64, ...; lane l holds elements o = 2 l + 128 j (j < 8) of a block:
// r1 = o >> 5, r2 = (o & 31) + {0, 1}.  rpart[item][q][3] = (rho00, rho01, rho11).
__global__ __launch_bounds__(256) void k_rho_sgram_f64(const EigItem* __restrict__ items,
                                                       const double* __restrict__ Q,
                                                       const double* __restrict__ lam, int npair,
                                                       double* __restrict__ rpart) {
  __shared__ double red[4][3];
  const EigItem it = items[blockIdx.y];
  const int q = blockIdx.x;
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double w12[8][2];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int o = 2 * l + 128 * j, r1 = o >> 5, r2 = o & 31;
    const double l1 = lam[it.L[0] + r1];
    w12[j][0] = l1 * lam[it.L[1] + r2];
    w12[j][1] = l1 * lam[it.L[1] + r2 + 1];
  }
  double a00 = 0, a01 = 0, a11 = 0;
#pragma unroll 1
  for (int p = 4 * q + wv; p < npair; p += 4 * RHO_NPART) {
    const double w34 = lam[it.L[2] + (p & 31)] * lam[it.L[3] + (p >> 5)];
    const double* B0 = Q + it.Q + (long long)(2 * p) * FBLK;
    const double* B1 = B0 + FBLK;
    double2 x[8], y[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      x[j] = *reinterpret_cast<const double2*>(B0 + 2 * l + 128 * j);
      y[j] = *reinterpret_cast<const double2*>(B1 + 2 * l + 128 * j);
    }
    double c00 = 0, c01 = 0, c11 = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const double xa = w12[j][0] * x[j].x, xb = w12[j][1] * x[j].y;
      c00 += xa * x[j].x + xb * x[j].y;
      c01 += xa * y[j].x + xb * y[j].y;
      c11 += w12[j][0] * y[j].x * y[j].x + w12[j][1] * y[j].y * y[j].y;
    }
    a00 += w34 * c00;
    a01 += w34 * c01;
    a11 += w34 * c11;
  }
#pragma unroll
  for (int sh = 32; sh >= 1; sh >>= 1) {
    a00 += __shfl_xor(a00, sh);
    a01 += __shfl_xor(a01, sh);
    a11 += __shfl_xor(a11, sh);
  }
  if (l == 0) {
    red[wv][0] = a00;
    red[wv][1] = a01;
    red[wv][2] = a11;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int k = threadIdx.x;
    rpart[((long long)blockIdx.y * RHO_NPART + q) * 3 + k] = (red[0][k] + red[1][k]) + (red[2][k] + red[3][k]);
  }
}

// request i: rho (2 x 2, column-major rho[s + 2 s']) of site item[i] at out + off[i]
__global__ void k_rho_finalize(const double* __restrict__ rpart, const int* __restrict__ item,
                               const long long* __restrict__ off, int n, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* p = rpart + (long long)item[i] * RHO_NPART * 3;
  double a[3] = {0, 0, 0};
  for (int q = 0; q < RHO_NPART; ++q)
    for (int k = 0; k < 3; ++k) a[k] += p[3 * q + k];
  double* r = out + off[i];
  r[0] = a[0];
  r[1] = a[1];
  r[2] = a[1];
  r[3] = a[2];
}

// Half-pair environments for two-site RDMs (the pair environment of
// test/test_belief_propagation.jl:64-91 with one bond open): for open bond c,
//   G[(s, r_c), (s', r_c')] = sum over the other indices of prod lam * Y Y',
// a 64 x 64 Gram of the rotated site with K running over every other index, then
// A = (I x U_c) G (I x U_c)^T.  The K stream is cut into 32 x 32 tiles, two per
// step (s = 0 and s = 1 of the same rows and K range):
//   bulk, c = 0: tile (r3, r4) = Q block [r1][r2], rows r1, K = r2
//   bulk, c = 1: the same block read transposed, rows r2, K = r1
//   bulk, c = 2: tile (r4, m): row m of the blocks (r3, r4), rows r3, K = r2
//   bulk, c = 3: tile (r3, m): row m of the blocks (r3, r4), rows r4, K = r2
//   degree 3 (blocks s + 2 r3): c = 0, 1 as bulk with tile r3; c = 2 tile m, rows r3
struct PairTask {
  long long Q;     // rotated site: Q (bulk) or Q3 (degree-3) element offset
  long long L[4];  // eigenvalue offsets of the site's bonds (L[3] = ones for degree 3)
  long long U;     // eigenvectors of the incoming message on the open bond
  long long out;   // output offset of the 64 x 64 half-pair
  int c;           // open bond, 0-based
  int deg3;
};
constexpr int PG_NPART = 16;  // workgroups (partial Grams) per task

// grid (PG_NPART, tasks), 256 threads: wave gw = 4 q + wv takes tiles gw,
// gw + 64, ...; each tile pair is staged in the wave's two LDS tiles (one tile
// pair of look-ahead in registers).  part[task][q] = (G00, G01, G11), 32 x 32 each
// (row-major [r][r']).
__global__ __launch_bounds__(256) void k_pair_gram_f64(const PairTask* __restrict__ tasks,
                                                       const double* __restrict__ Qbulk,
                                                       const double* __restrict__ Q3,
                                                       const double* __restrict__ lam,
                                                       double* __restrict__ part) {
  __shared__ double Tb[4][2][FC * LD8];
  const PairTask tk = tasks[blockIdx.y];
  const double* Q = (tk.deg3 ? Q3 : Qbulk) + tk.Q;
  const int q = blockIdx.x;
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int li = l & 15, lg = l >> 4;
  const int c = tk.c;
  const bool tr = c == 1;
  const int ntile = tk.deg3 ? FC : FC * FC;
  const long long kw = c == 1 ? tk.L[0] : tk.L[1];
  double lamK[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) lamK[t] = lam[kw + 4 * t + lg];
  // s = 0 tile of step t: element [a][b] (loaded orientation) at base + a rs + b
  auto geom = [&](int t, long long& base, long long& rs, double& ws) {
    const int lo = t & 31, hi = t >> 5;
    if (tk.deg3) {
      if (c <= 1) {
        base = 2LL * t * FBLK;
        rs = FC;
        ws = lam[tk.L[2] + t];
      } else {
        base = (long long)FC * t;
        rs = 2LL * FBLK;
        ws = lam[tk.L[0] + t];
      }
    } else if (c <= 1) {
      base = 2LL * (lo + FC * hi) * FBLK;
      rs = FC;
      ws = lam[tk.L[2] + lo] * lam[tk.L[3] + hi];
    } else if (c == 2) {
      base = 2LL * FC * lo * FBLK + FC * hi;
      rs = 2LL * FBLK;
      ws = lam[tk.L[0] + hi] * lam[tk.L[3] + lo];
    } else {
      base = 2LL * lo * FBLK + FC * hi;
      rs = 2LL * FC * FBLK;
      ws = lam[tk.L[0] + hi] * lam[tk.L[2] + lo];
    }
  };
  Acc32 C00, C01, C11;
  acc_zero(C00);
  acc_zero(C01);
  acc_zero(C11);
  double* t0 = Tb[wv][0];
  double* t1 = Tb[wv][1];
  double2 n0[8], n1[8];
  auto fetch = [&](int t) {
    long long base, rs;
    double ws;
    geom(t, base, rs, ws);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int o = 2 * l + 128 * j, a = o >> 5, b = o & 31;
      n0[j] = *reinterpret_cast<const double2*>(Q + base + a * rs + b);
      n1[j] = *reinterpret_cast<const double2*>(Q + base + FBLK + a * rs + b);
    }
  };
  auto stage = [&]() {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int o = 2 * l + 128 * j, a = o >> 5, b = o & 31;
      t0[a * LD8 + b] = n0[j].x;
      t0[a * LD8 + b + 1] = n0[j].y;
      t1[a * LD8 + b] = n1[j].x;
      t1[a * LD8 + b + 1] = n1[j].y;
    }
  };
  const int t_first = 4 * q + wv, t_step = 4 * PG_NPART;
  if (t_first < ntile) {
    fetch(t_first);
    stage();
  }
#pragma unroll 1
  for (int t = t_first; t < ntile; t += t_step) {
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    const bool more = t + t_step < ntile;
    if (more) fetch(t + t_step);
    long long base, rs;
    double ws;
    geom(t, base, rs, ws);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      double a0[2], a1[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        // operand (row 16h + li, K 4k + lg) of the tile in row/K orientation
        const int ia = tr ? (4 * k + lg) * LD8 + 16 * h + li : (16 * h + li) * LD8 + 4 * k + lg;
        a0[h] = t0[ia];
        a1[h] = t1[ia];
      }
      const double w = ws * lamK[k];
#pragma unroll
      for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni) {
          C00.t[mi][ni] = MFMA_F64(a0[mi], w * a0[ni], C00.t[mi][ni]);
          C01.t[mi][ni] = MFMA_F64(a0[mi], w * a1[ni], C01.t[mi][ni]);
          C11.t[mi][ni] = MFMA_F64(a1[mi], w * a1[ni], C11.t[mi][ni]);
        }
    }
    // every lane's reads of this pair are issued before the tiles are overwritten
    __builtin_amdgcn_wave_barrier();
    if (more) stage();
  }
  __syncthreads();
  double* Red = &Tb[0][0][0];  // 4 * 2 * 32 * 33 >= 3 * 1024
  for (int w4 = 0; w4 < 4; ++w4) {
    if (wv == w4) {
#pragma unroll
      for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int o = (16 * mi + 4 * v + lg) * FC + 16 * ni + li;
            if (w4 == 0) {
              Red[o] = C00.t[mi][ni][v];
              Red[FBLK + o] = C01.t[mi][ni][v];
              Red[2 * FBLK + o] = C11.t[mi][ni][v];
            } else {
              Red[o] += C00.t[mi][ni][v];
              Red[FBLK + o] += C01.t[mi][ni][v];
              Red[2 * FBLK + o] += C11.t[mi][ni][v];
            }
          }
    }
    __syncthreads();
  }
  double* dst = part + ((long long)blockIdx.y * PG_NPART + q) * 3 * FBLK;
  for (int o = threadIdx.x; o < 3 * FBLK; o += blockDim.x) dst[o] = Red[o];
}

// Sum the partial Grams (fixed order), rotate back: X = G (I x U)^T, then
// A = (I x U) X, written as A[(s + 2 b) + 64 (s' + 2 b')].  grid (tasks), 256 threads.
__global__ __launch_bounds__(256) void k_pair_finalize(const PairTask* __restrict__ tasks,
                                                       const double* __restrict__ part,
                                                       const double* __restrict__ U64, double* __restrict__ out) {
  __shared__ double G[3][FBLK];   // G00, G01, G11 as [r][r']
  __shared__ double Us[FBLK];     // U[b + 32 r]
  __shared__ double X[4][FBLK];   // X_{ss'}[r][b'] = sum_r' G_{ss'}[r][r'] U[b' + 32 r']
  const PairTask tk = tasks[blockIdx.x];
  const double* p = part + (long long)blockIdx.x * PG_NPART * 3 * FBLK;
  for (int o = threadIdx.x; o < 3 * FBLK; o += blockDim.x) {
    double a = 0;
    for (int q = 0; q < PG_NPART; ++q) a += p[(long long)q * 3 * FBLK + o];
    G[o / FBLK][o % FBLK] = a;
  }
  for (int o = threadIdx.x; o < FBLK; o += blockDim.x) Us[o] = U64[tk.U + o];
  __syncthreads();
  for (int o = threadIdx.x; o < 4 * FBLK; o += blockDim.x) {
    const int ss = o >> 10, r = (o >> 5) & 31, b2 = o & 31;
    const int sa = ss & 1, sb = ss >> 1;
    double a = 0;
    for (int r2 = 0; r2 < FC; ++r2) {
      // G_{sa sb}[r][r2]; G10 = G01^T
      const double g = (sa == sb) ? G[sa ? 2 : 0][r * FC + r2] : (sa == 0 ? G[1][r * FC + r2] : G[1][r2 * FC + r]);
      a += g * Us[b2 + FC * r2];
    }
    X[ss][r * FC + b2] = a;
  }
  __syncthreads();
  double* A = out + tk.out;
  for (int o = threadIdx.x; o < 4 * FBLK; o += blockDim.x) {
    // A row x = s + 2 b, column y = s' + 2 b'
    const int x = o & 63, y = o >> 6;
    const int sa = x & 1, b = x >> 1, sb = y & 1, b2 = y >> 1;
    const double* Xs = X[sa + 2 * sb];
    double a = 0;
    for (int r = 0; r < FC; ++r) a += Us[b + FC * r] * Xs[r * FC + b2];
    A[o] = a;
  }
}

// ---------------------------------------------------------------------------
// host state
// ---------------------------------------------------------------------------
struct FastState {
  std::vector<int> bulk;            // eligible vertices
  std::vector<FastItem> items;      // 2 per bulk vertex (passes A and B)
  std::vector<int> out_msgs;        // message ids produced
  DevBuf d_items, lay16, lay64, P32, P16, P64, part32, part64, offs;
  bool lay16_ok = false, lay64_ok = false;
  int chunk[2] = {0, 0};            // items per absorb/fused launch pair, per precision (0: untuned)
  std::vector<float> tune_ms[2];    // autotuner timings of the candidates
  // eigenbasis pipeline
  bool eig = false;                 // chosen for the current phase
  std::vector<EigItem> eitems;      // 1 per bulk vertex
  std::vector<FastItem> k1items;    // rot12 descriptors (the absorb kernel's item type)
  std::vector<EigItem> k1eitems;    // rot12 descriptors for the rot34 kernels without Grams
  std::vector<long long> eig_moff;  // message offset of every eigendecomposed message
  DevBuf d_eitems, d_k1items, d_k1eitems, d_eig_moff, eigU64, eigL64, eigU32, eigL32;
  DevBuf eigLexp;  // power-of-two exponents of the FP32 eigenvalues, one per message (+ 0 for the ones)
  bool eig_warm = false;            // eigU64 holds an orthogonal matrix per slot (seeds the next eigh)
  long long eig_count = 0;          // eigh launches since the last allocation; every 16th starts cold
  // degree-3 sites (eigenbasis pipeline only)
  std::vector<int> deg3;            // eligible degree-3 vertices
  std::vector<int> deg3_used;       // those whose outgoing messages are all in the sequence
  std::vector<EigItem> e3items;
  DevBuf d_e3items, lay3_16, lay3_64, P3_16, P3_32, P3_64, offs3;
  bool lay3_16_ok = false, lay3_64_ok = false;
  // compensated 16-bit arithmetic (split operands, FP32 intermediates) in the
  // current reduced-precision phase; h2: two FP16 terms instead of three BF16
  // terms on the eigenbasis kernels
  bool x3 = false, h2 = false;
  // power-of-two exponents of the 16-bit layouts: bulk site bi at bi, degree-3
  // site k (index into deg3) at bulk.size() + k
  DevBuf sexp, pmax;
  // one-site density matrices: item lists, partial sums and request tables
  DevBuf rdm;
};

FastState* state_of(itnbp_net* net) { return reinterpret_cast<FastState*>(net->fast); }

// BF16 flavour: compact (P and the per-block intermediates as one BF16 term,
// default) or precise (P in FP32, every operand split hi+lo); ITNBP_BF16_PRECISE=1
// FP64 flavour: register-resident MFMA operands (default) or the LDS-tile
// GEMMs (ITNBP_F64_LDS=1)
bool fast_f64_lds() {
  const char* e = getenv("ITNBP_F64_LDS");
  return e && e[0] == '1';
}

bool fast_precise() {
  const char* e = getenv("ITNBP_BF16_PRECISE");
  return e && e[0] == '1';
}

// the eigenbasis pipeline is used when every message is symmetric, unless
// ITNBP_FAST_LEGACY=1 asks for the absorbed-pair passes
bool use_eig(const itnbp_net* net) {
  const char* e = getenv("ITNBP_FAST_LEGACY");
  return net->msgs_sym && !(e && e[0] == '1');
}

bool eligible(const itnbp_net* net, int v) {
  const VDesc& V = net->vdesc[v];
  if (V.k != 4 || V.dims[0] != 2) return false;
  for (int a = 1; a <= 4; ++a)
    if (V.dims[a] != FC) return false;
  return true;
}

bool eligible3(const itnbp_net* net, int v) {
  const VDesc& V = net->vdesc[v];
  if (V.k != 3 || V.dims[0] != 2) return false;
  for (int a = 1; a <= 3; ++a)
    if (V.dims[a] != FC) return false;
  return true;
}

int out_msg_id(const itnbp_net* net, int v, int a) {
  const int e = net->inc_edge[net->inc_ptr[v] + a - 1];
  return msg_id_into(net, e, net->edges[2 * e] == v ? net->edges[2 * e + 1] : net->edges[2 * e]);
}

}  // namespace

bool fast_path_available(itnbp_net* net, int acc_f64, int schedule) {
  if (schedule != ITNBP_SCHED_PARALLEL) return false;
  // the BF16 kernels need psi exactly representable in BF16
  if (!acc_f64 && net->dtype != ITNBP_BF16) return false;
  if (getenv("ITNBP_DISABLE_FAST")) return false;
  const bool eig = use_eig(net);
  for (int v = 0; v < net->nv; ++v)
    if (eligible(net, v) || (eig && eligible3(net, v))) return true;
  return false;
}

// compensated reduced-precision arithmetic: requested by the update
// (ITNBP_PREC_BF16X3*, ITNBP_PREC_*F16X2) or by ITNBP_BF16_PRECISE=1
static bool precise_of(const itnbp_net* net) { return net->low_x3 != 0 || fast_precise(); }

const char* fast_path_name(itnbp_net* net, int acc_f64) {
  if (use_eig(net))
    return acc_f64 ? "mfma-f64-eig-chi32"
                   : (precise_of(net) ? (net->low_x3 == 2 ? "mfma-f16x2-eig-chi32" : "mfma-bf16x3-eig-chi32")
                                      : "mfma-bf16-eig-chi32");
  return acc_f64 ? "mfma-f64-chi32" : (precise_of(net) ? "mfma-bf16x2-chi32" : "mfma-bf16c-chi32");
}

void release_fast_state(itnbp_net* net) {
  FastState* st = state_of(net);
  if (!st) return;
  free_buf(st->d_items);
  free_buf(st->lay16);
  free_buf(st->lay64);
  free_buf(st->P32);
  free_buf(st->P16);
  free_buf(st->P64);
  free_buf(st->part32);
  free_buf(st->part64);
  free_buf(st->offs);
  free_buf(st->d_eitems);
  free_buf(st->d_k1items);
  free_buf(st->d_k1eitems);
  free_buf(st->d_eig_moff);
  free_buf(st->eigU64);
  free_buf(st->eigL64);
  free_buf(st->eigU32);
  free_buf(st->eigL32);
  free_buf(st->eigLexp);
  free_buf(st->d_e3items);
  free_buf(st->lay3_16);
  free_buf(st->lay3_64);
  free_buf(st->P3_16);
  free_buf(st->P3_32);
  free_buf(st->P3_64);
  free_buf(st->offs3);
  free_buf(st->sexp);
  free_buf(st->pmax);
  free_buf(st->rdm);
  delete st;
  net->fast = nullptr;
}

void invalidate_fast_state(itnbp_net* net) {
  FastState* st = state_of(net);
  if (!st) return;
  st->lay16_ok = false;
  st->lay64_ok = false;
  st->lay3_16_ok = false;
  st->lay3_64_ok = false;
}

// Exponents of sites first .. first + n - 1 (psi offsets src, nelem elements each).
int site_exponents(itnbp_net* net, FastState* st, const long long* src, int n, long long nelem, int first) {
  itnbp_ctx* ctx = net->ctx;
  const size_t ns = st->bulk.size() + st->deg3.size();
  ITNBP_RETCHK(ensure_buf(ctx, st->sexp, ns * sizeof(int)));
  ITNBP_RETCHK(ensure_buf(ctx, st->pmax, ns * 64 * sizeof(double)));
  double* pm = (double*)st->pmax.p + (size_t)first * 64;
  dim3 g(64, n);
  switch (net->dtype) {
    case ITNBP_F64: hipLaunchKernelGGL(k_site_absmax<double>, g, dim3(256), 0, ctx->stream, (const double*)net->d_psi, src, nelem, pm); break;
    case ITNBP_F32: hipLaunchKernelGGL(k_site_absmax<float>, g, dim3(256), 0, ctx->stream, (const float*)net->d_psi, src, nelem, pm); break;
    default: hipLaunchKernelGGL(k_site_absmax<bf16_t>, g, dim3(256), 0, ctx->stream, (const bf16_t*)net->d_psi, src, nelem, pm); break;
  }
  ITNBP_HIPCHK(ctx, hipGetLastError());
  hipLaunchKernelGGL(k_site_exp, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, (const double*)pm, n,
                     (int*)st->sexp.p + first);
  ITNBP_HIPCHK(ctx, hipGetLastError());
  return ITNBP_OK;
}

// Mark the sequence entries the pipeline produces and prepare its buffers.
int fast_prepare(itnbp_net* net, const std::vector<int>& seq, int acc_f64, std::vector<char>& handled) {
  itnbp_ctx* ctx = net->ctx;
  handled.assign(seq.size(), 0);
  FastState* st = state_of(net);
  if (!st) {
    st = new FastState();
    net->fast = st;
    for (int v = 0; v < net->nv; ++v) {
      if (eligible(net, v)) st->bulk.push_back(v);
      if (eligible3(net, v)) st->deg3.push_back(v);
    }
  }
  const long long per = (long long)FNR * FBLK;  // elements per vertex layout
  // slots
  std::vector<int> pos(2 * net->ne, -1);
  for (size_t i = 0; i < seq.size(); ++i) pos[seq[i]] = (int)i;
  st->items.clear();
  const int nb = (int)st->bulk.size();
  for (int bi = 0; bi < nb; ++bi) {
    const int v = st->bulk[bi];
    const VDesc& V = net->vdesc[v];
    const int o1 = out_msg_id(net, v, 1), o2 = out_msg_id(net, v, 2), o3 = out_msg_id(net, v, 3),
              o4 = out_msg_id(net, v, 4);
    // every outgoing message of a bulk vertex must be in the sequence
    if (pos[o1] < 0 || pos[o2] < 0 || pos[o3] < 0 || pos[o4] < 0) return ITNBP_ERR_UNSUPPORTED;
    const long long lay = (long long)bi * per;
    FastItem a{}, b{};
    a.psi_in = nb * per + lay;  // layout B buffer follows layout A
    a.psi_out = lay;
    a.P = (long long)(2 * bi) * per;
    a.Ma = V.in_msg[1];
    a.Mb = V.in_msg[2];
    a.Mz = V.in_msg[3];
    a.Mw = V.in_msg[4];
    a.out_z = net->msg_off[o3];
    a.out_w = net->msg_off[o4];
    a.slot_z = pos[o3];
    a.slot_w = pos[o4];
    a.part = (long long)(2 * bi) * FUSED_CHUNKS * 2 * FBLK;
    b.psi_in = lay;
    b.psi_out = nb * per + lay;
    b.P = (long long)(2 * bi + 1) * per;
    b.Ma = V.in_msg[3];
    b.Mb = V.in_msg[4];
    b.Mz = V.in_msg[1];
    b.Mw = V.in_msg[2];
    b.out_z = net->msg_off[o1];
    b.out_w = net->msg_off[o2];
    b.slot_z = pos[o1];
    b.slot_w = pos[o2];
    b.part = (long long)(2 * bi + 1) * FUSED_CHUNKS * 2 * FBLK;
    a.site = b.site = bi;
    st->items.push_back(a);
    st->items.push_back(b);
    for (int o : {o1, o2, o3, o4}) handled[pos[o]] = 1;
  }
  const size_t ni = st->items.size();
  ITNBP_RETCHK(ensure_buf(ctx, st->d_items, ni * sizeof(FastItem)));
  if (ni > 0)
    ITNBP_HIPCHK(ctx, hipMemcpyAsync(st->d_items.p, st->items.data(), ni * sizeof(FastItem), hipMemcpyHostToDevice,
                                     ctx->stream));
  // eigenbasis pipeline: R = first half of the P buffer, Q = second half;
  // eigen slot 4 bi + (c - 1) holds the message into bulk vertex bi on bond c
  st->eig = use_eig(net);
  st->x3 = !acc_f64 && precise_of(net);
  st->h2 = st->x3 && net->low_x3 == 2;
  st->eitems.clear();
  st->k1items.clear();
  st->k1eitems.clear();
  st->eig_moff.clear();
  st->e3items.clear();
  st->deg3_used.clear();
  if (st->eig) {
    for (int bi = 0; bi < nb; ++bi) {
      const int v = st->bulk[bi];
      const VDesc& V = net->vdesc[v];
      EigItem e{};
      e.R = (long long)bi * per;
      e.Q = (long long)(nb + bi) * per;
      for (int c = 0; c < 4; ++c) {
        const long long slot = 4LL * bi + c;
        e.U[c] = slot * FBLK;
        e.L[c] = slot * FC;
        const int o = out_msg_id(net, v, c + 1);
        e.out[c] = net->msg_off[o];
        e.slot[c] = pos[o];
        st->eig_moff.push_back(V.in_msg[c + 1]);
      }
      e.part = (long long)bi * 4 * EIG_NPART * FBLK;
      e.site = bi;
      st->eitems.push_back(e);
      FastItem k{};
      k.psi_in = (long long)(nb + bi) * per;  // layout B
      k.P = e.R;
      k.Ma = e.U[0];
      k.Mb = e.U[1];
      st->k1items.push_back(k);
      EigItem k1 = e;
      k1.R = (long long)(nb + bi) * per;  // psi layout B
      k1.Q = e.R;
      k1.U[2] = e.U[0];
      k1.U[3] = e.U[1];
      st->k1eitems.push_back(k1);
    }
    ITNBP_RETCHK(ensure_buf(ctx, st->d_eitems, nb * sizeof(EigItem)));
    ITNBP_RETCHK(ensure_buf(ctx, st->d_k1items, nb * sizeof(FastItem)));
    ITNBP_RETCHK(ensure_buf(ctx, st->d_k1eitems, nb * sizeof(EigItem)));
    if (nb > 0) {
      ITNBP_HIPCHK(ctx, hipMemcpyAsync(st->d_eitems.p, st->eitems.data(), nb * sizeof(EigItem),
                                       hipMemcpyHostToDevice, ctx->stream));
      ITNBP_HIPCHK(ctx, hipMemcpyAsync(st->d_k1items.p, st->k1items.data(), nb * sizeof(FastItem),
                                       hipMemcpyHostToDevice, ctx->stream));
      ITNBP_HIPCHK(ctx, hipMemcpyAsync(st->d_k1eitems.p, st->k1eitems.data(), nb * sizeof(EigItem),
                                       hipMemcpyHostToDevice, ctx->stream));
    }
    // degree-3 sites: slots after the bulk ones; l4 points at a table of ones
    // stored after the last eigenvalue slot
    st->deg3_used.clear();
    st->e3items.clear();
    const long long per3 = (long long)D3_NBLK * FBLK;
    for (size_t k = 0; k < st->deg3.size(); ++k) {
      const int v = st->deg3[k];
      const VDesc& V = net->vdesc[v];
      const int o1 = out_msg_id(net, v, 1), o2 = out_msg_id(net, v, 2), o3 = out_msg_id(net, v, 3);
      if (pos[o1] < 0 || pos[o2] < 0 || pos[o3] < 0) continue;
      const int j = (int)st->deg3_used.size();
      st->deg3_used.push_back(v);
      EigItem e{};
      e.X = (long long)k * per3;  // layout C holds every degree-3 site
      e.site = nb + (int)k;
      e.R = (long long)j * per3;
      e.Q = (long long)(st->deg3.size() + j) * per3;
      const int oo[3] = {o1, o2, o3};
      for (int c = 0; c < 3; ++c) {
        const long long slot = (long long)st->eig_moff.size();
        e.U[c] = slot * FBLK;
        e.L[c] = slot * FC;
        e.out[c] = net->msg_off[oo[c]];
        e.slot[c] = pos[oo[c]];
        st->eig_moff.push_back(V.in_msg[c + 1]);
        handled[pos[oo[c]]] = 1;
      }
      e.part = (long long)(nb + j) * 4 * EIG_NPART * FBLK;
      st->e3items.push_back(e);
    }
    const size_t ne4 = st->eig_moff.size();
    for (EigItem& e : st->e3items) e.L[3] = (long long)ne4 * FC;  // ones
    const int n3 = (int)st->e3items.size();
    if (n3 > 0) {
      ITNBP_RETCHK(ensure_buf(ctx, st->d_e3items, n3 * sizeof(EigItem)));
      ITNBP_HIPCHK(ctx, hipMemcpyAsync(st->d_e3items.p, st->e3items.data(), n3 * sizeof(EigItem),
                                       hipMemcpyHostToDevice, ctx->stream));
    }
    if (ne4 * FBLK * 8 > st->eigU64.bytes) st->eig_warm = false;  // reallocated below: no seed
    ITNBP_RETCHK(ensure_buf(ctx, st->d_eig_moff, ne4 * 8));
    ITNBP_HIPCHK(ctx, hipMemcpyAsync(st->d_eig_moff.p, st->eig_moff.data(), ne4 * 8, hipMemcpyHostToDevice,
                                     ctx->stream));
    ITNBP_RETCHK(ensure_buf(ctx, st->eigU64, ne4 * FBLK * 8));
    ITNBP_RETCHK(ensure_buf(ctx, st->eigL64, (ne4 + 1) * FC * 8));
    static const double ones64[FC] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1,
                                      1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
    static const float ones32[FC] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1,
                                     1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
    ITNBP_HIPCHK(ctx, hipMemcpyAsync((double*)st->eigL64.p + ne4 * FC, ones64, sizeof(ones64),
                                     hipMemcpyHostToDevice, ctx->stream));
    if (!acc_f64) {
      ITNBP_RETCHK(ensure_buf(ctx, st->eigU32, ne4 * FBLK * 4));
      ITNBP_RETCHK(ensure_buf(ctx, st->eigL32, (ne4 + 1) * FC * 4));
      ITNBP_HIPCHK(ctx, hipMemcpyAsync((float*)st->eigL32.p + ne4 * FC, ones32, sizeof(ones32),
                                       hipMemcpyHostToDevice, ctx->stream));
      static const int zero_exp = 0;
      ITNBP_RETCHK(ensure_buf(ctx, st->eigLexp, (ne4 + 1) * 4));
      ITNBP_HIPCHK(ctx, hipMemcpyAsync((int*)st->eigLexp.p + ne4, &zero_exp, 4, hipMemcpyHostToDevice, ctx->stream));
    }
    // layout C of every degree-3 site and the block buffers of those in use
    const int n3all = (int)st->deg3.size();
    if (n3 > 0) {
      std::vector<long long> o3(2 * n3all);
      for (int k = 0; k < n3all; ++k) {
        o3[k] = net->psi_off[st->deg3[k]];
        o3[n3all + k] = (long long)k * per3;
      }
      ITNBP_RETCHK(ensure_buf(ctx, st->offs3, o3.size() * 8));
      ITNBP_HIPCHK(ctx, hipMemcpyAsync(st->offs3.p, o3.data(), o3.size() * 8, hipMemcpyHostToDevice, ctx->stream));
      ITNBP_HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
      const long long* s3 = (const long long*)st->offs3.p;
      const long long* d3 = s3 + n3all;
      dim3 g3(FC, n3all);
      if (!acc_f64) {
        ITNBP_RETCHK(ensure_buf(ctx, st->lay3_16, (size_t)n3all * per3 * 2));
        if (st->x3) ITNBP_RETCHK(ensure_buf(ctx, st->P3_32, (size_t)2 * n3all * per3 * 4));
        else ITNBP_RETCHK(ensure_buf(ctx, st->P3_16, (size_t)2 * n3all * per3 * 2));
        if (!st->lay3_16_ok) {
          ITNBP_RETCHK(site_exponents(net, st, s3, n3all, 2LL * FC * FC * FC, nb));
          __bf16* C = (__bf16*)st->lay3_16.p;
          const int* e3 = (const int*)st->sexp.p + nb;
          switch (net->dtype) {
            case ITNBP_F64: hipLaunchKernelGGL((k_lay_cols<double, __bf16>), g3, dim3(256), 0, ctx->stream, (const double*)net->d_psi, s3, d3, C, e3); break;
            case ITNBP_F32: hipLaunchKernelGGL((k_lay_cols<float, __bf16>), g3, dim3(256), 0, ctx->stream, (const float*)net->d_psi, s3, d3, C, e3); break;
            default: hipLaunchKernelGGL((k_lay_cols<bf16_t, __bf16>), g3, dim3(256), 0, ctx->stream, (const bf16_t*)net->d_psi, s3, d3, C, e3); break;
          }
          ITNBP_HIPCHK(ctx, hipGetLastError());
          st->lay3_16_ok = true;
        }
      } else {
        ITNBP_RETCHK(ensure_buf(ctx, st->lay3_64, (size_t)n3all * per3 * 8));
        ITNBP_RETCHK(ensure_buf(ctx, st->P3_64, (size_t)2 * n3all * per3 * 8));
        if (!st->lay3_64_ok) {
          double* C = (double*)st->lay3_64.p;
          switch (net->dtype) {
            case ITNBP_F64: hipLaunchKernelGGL((k_lay_cols<double, double>), g3, dim3(256), 0, ctx->stream, (const double*)net->d_psi, s3, d3, C, (const int*)nullptr); break;
            case ITNBP_F32: hipLaunchKernelGGL((k_lay_cols<float, double>), g3, dim3(256), 0, ctx->stream, (const float*)net->d_psi, s3, d3, C, (const int*)nullptr); break;
            default: hipLaunchKernelGGL((k_lay_cols<bf16_t, double>), g3, dim3(256), 0, ctx->stream, (const bf16_t*)net->d_psi, s3, d3, C, (const int*)nullptr); break;
          }
          ITNBP_HIPCHK(ctx, hipGetLastError());
          st->lay3_64_ok = true;
        }
      }
    }
  }
  // layouts
  std::vector<long long> offs(2 * nb);
  for (int bi = 0; bi < nb; ++bi) {
    offs[bi] = net->psi_off[st->bulk[bi]];
    offs[nb + bi] = (long long)bi * per;
  }
  ITNBP_RETCHK(ensure_buf(ctx, st->offs, offs.size() * 8));
  ITNBP_HIPCHK(ctx, hipMemcpyAsync(st->offs.p, offs.data(), offs.size() * 8, hipMemcpyHostToDevice, ctx->stream));
  ITNBP_HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  const long long* d_src = (const long long*)st->offs.p;
  const long long* d_dst = d_src + nb;
  if (!acc_f64 && !st->lay16_ok && nb > 0) {
    ITNBP_RETCHK(ensure_buf(ctx, st->lay16, (size_t)2 * nb * per * 2));
    ITNBP_RETCHK(site_exponents(net, st, d_src, nb, per, 0));
    __bf16* A = (__bf16*)st->lay16.p;
    __bf16* B = A + nb * per;
    const int* e = (const int*)st->sexp.p;
    switch (net->dtype) {
      case ITNBP_F64: launch_layout_ab(ctx->stream, (const double*)net->d_psi, d_src, d_dst, A, B, nb, e); break;
      case ITNBP_F32: launch_layout_ab(ctx->stream, (const float*)net->d_psi, d_src, d_dst, A, B, nb, e); break;
      default: launch_layout_ab(ctx->stream, (const bf16_t*)net->d_psi, d_src, d_dst, A, B, nb, e); break;
    }
    ITNBP_HIPCHK(ctx, hipGetLastError());
    st->lay16_ok = true;
  }
  if (acc_f64 && !st->lay64_ok && nb > 0) {
    ITNBP_RETCHK(ensure_buf(ctx, st->lay64, (size_t)2 * nb * per * 8));
    double* A = (double*)st->lay64.p;
    double* B = A + nb * per;
    switch (net->dtype) {
      case ITNBP_F64: launch_layout_ab(ctx->stream, (const double*)net->d_psi, d_src, d_dst, A, B, nb, (const int*)nullptr); break;
      case ITNBP_F32: launch_layout_ab(ctx->stream, (const float*)net->d_psi, d_src, d_dst, A, B, nb, (const int*)nullptr); break;
      default: launch_layout_ab(ctx->stream, (const bf16_t*)net->d_psi, d_src, d_dst, A, B, nb, (const int*)nullptr); break;
    }
    ITNBP_HIPCHK(ctx, hipGetLastError());
    st->lay64_ok = true;
  }
  // partial sums: the absorbed-pair passes need 2 items x FUSED_CHUNKS x 2 matrices
  // per bulk vertex, the eigenbasis pipeline 4 bonds x EIG_NPART per (bulk or degree-3) site
  const size_t nparts = std::max(ni * FUSED_CHUNKS * 2, (size_t)(nb + st->e3items.size()) * 4 * EIG_NPART);
  if (acc_f64) {
    ITNBP_RETCHK(ensure_buf(ctx, st->P64, (size_t)2 * nb * per * 8));
    ITNBP_RETCHK(ensure_buf(ctx, st->part64, nparts * FBLK * 8));
  } else {
    if (st->x3)
      ITNBP_RETCHK(ensure_buf(ctx, st->P32, (size_t)2 * nb * per * 4));
    else
      ITNBP_RETCHK(ensure_buf(ctx, st->P16, (size_t)2 * nb * per * 2));
    ITNBP_RETCHK(ensure_buf(ctx, st->part32, nparts * FBLK * 4));
  }
  ITNBP_HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return ITNBP_OK;
}

// absorb + fused launches for items [i0, i1)
static void launch_pair(itnbp_ctx* ctx, FastState* st, int acc_f64, const void* msg_in, int i0, int i1) {
  const int ni = i1 - i0;
  const FastItem* items = (const FastItem*)st->d_items.p + i0;
  const int e0 = prof_mark(ctx);
  if (!acc_f64) {
    if (st->x3) {
      hipLaunchKernelGGL((k_absorb_bf16<2>), dim3(128, ni), dim3(256), 0, ctx->stream, items,
                         (const __bf16*)st->lay16.p, (const float*)msg_in, (float*)st->P32.p);
    } else {
      hipLaunchKernelGGL(k_absorb_bf16c<__bf16>, dim3(64, ni), dim3(512), 0, ctx->stream, items,
                         (const __bf16*)st->lay16.p, (const float*)msg_in, (__bf16*)st->P16.p);
    }
  } else if (fast_f64_lds()) {
    hipLaunchKernelGGL(k_absorb_f64, dim3(128, ni), dim3(256), 0, ctx->stream, items,
                       (const double*)st->lay64.p, (const double*)msg_in, (double*)st->P64.p);
  } else {
    hipLaunchKernelGGL(k_absorb_f64r, dim3(256, ni), dim3(256), 0, ctx->stream, items,
                       (const double*)st->lay64.p, (const double*)msg_in, (double*)st->P64.p);
  }
  const int e1 = prof_mark(ctx);
  if (!acc_f64) {
    if (st->x3) {
      hipLaunchKernelGGL((k_fused_bf16<2>), dim3(FUSED_CHUNKS, ni), dim3(256), 0, ctx->stream, items,
                         (const __bf16*)st->lay16.p, (const float*)msg_in, (const float*)st->P32.p,
                         (float*)st->part32.p);
    } else {
      hipLaunchKernelGGL(k_fused_bf16c, dim3(FUSED_CHUNKS, ni), dim3(256), 0, ctx->stream, items,
                         (const __bf16*)st->lay16.p, (const float*)msg_in, (const __bf16*)st->P16.p,
                         (float*)st->part32.p);
    }
  } else if (fast_f64_lds()) {
    hipLaunchKernelGGL(k_fused_f64, dim3(FUSED_CHUNKS, ni), dim3(256), 0, ctx->stream, items,
                       (const double*)st->lay64.p, (const double*)msg_in, (const double*)st->P64.p,
                       (double*)st->part64.p);
  } else {
    hipLaunchKernelGGL(k_fused_f64r, dim3(FUSED_CHUNKS, ni), dim3(256), 0, ctx->stream, items,
                       (const double*)st->lay64.p, (const double*)msg_in, (const double*)st->P64.p,
                       (double*)st->part64.p);
  }
  const int e2 = prof_mark(ctx);
  prof_add(ctx, acc_f64 ? "absorb_f64" : "absorb_bf16", e0, e1);
  prof_add(ctx, acc_f64 ? "fused_f64" : "fused_bf16", e1, e2);
}

static void launch_all_pairs(itnbp_ctx* ctx, FastState* st, int acc_f64, const void* msg_in, int chunk) {
  const int ni = (int)st->items.size();
  for (int i0 = 0; i0 < ni; i0 += chunk) launch_pair(ctx, st, acc_f64, msg_in, i0, std::min(ni, i0 + chunk));
}

// Autotuner: the number of items per absorb/fused launch pair trades launch
// count against reuse of the intermediate P in the last-level cache (a chunk
// whose P fits in the 256 MiB Infinity Cache is read back from it).  Every
// candidate is timed on one real sweep (same inputs; the outputs are
// recomputed by the sweep that follows) and the fastest is kept per network
// and precision.  ITNBP_FAST_CHUNK=<items> forces a value.
static int tune_chunk(itnbp_net* net, FastState* st, int acc_f64, const void* msg_in) {
  itnbp_ctx* ctx = net->ctx;
  const int ni = (int)st->items.size();
  if (const char* e = getenv("ITNBP_FAST_CHUNK")) {
    const int c = atoi(e);
    if (c > 0) return std::min(c, ni);
  }
  if (st->chunk[acc_f64] > 0) return st->chunk[acc_f64];
  std::vector<int> cand{ni};
  for (int c : {512, 192, 96, 48, 24})
    if (c < ni) cand.push_back(c);
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess) return ni;
  if (hipEventCreate(&b) != hipSuccess) {
    (void)hipEventDestroy(a);
    return ni;
  }
  const bool prof = ctx->prof;
  ctx->prof = false;
  int best = ni;
  float best_ms = 1e30f;
  st->tune_ms[acc_f64].clear();
  launch_all_pairs(ctx, st, acc_f64, msg_in, ni);  // warm-up
  for (int c : cand) {
    (void)hipEventRecord(a, ctx->stream);
    launch_all_pairs(ctx, st, acc_f64, msg_in, c);
    (void)hipEventRecord(b, ctx->stream);
    (void)hipEventSynchronize(b);
    float ms = 1e30f;
    (void)hipEventElapsedTime(&ms, a, b);
    st->tune_ms[acc_f64].push_back(ms);
    if (ms < best_ms) {
      best_ms = ms;
      best = c;
    }
  }
  ctx->prof = prof;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  st->chunk[acc_f64] = best;
  return best;
}

int fast_chunk(itnbp_net* net, int acc_f64) {
  FastState* st = state_of(net);
  return st ? st->chunk[acc_f64 ? 1 : 0] : 0;
}

// Intermediates of the 16-bit eigenbasis pipeline: FP16 (default; 8x finer
// rounding than BF16, range kept by the per-site scales) or BF16
// (ITNBP_EIG_BF16=1)
static bool eig_bf16_intermediates() {
  const char* e = getenv("ITNBP_EIG_BF16");
  return e && e[0] == '1';
}

// stage 1 rot12, 2 rot34 (+ T'3, T'4), 3 gram12 (T'1, T'2), 4 degree-3 sites
template <typename TI>
static void eig16_stage(itnbp_ctx* ctx, FastState* st, int stage, int nb, int n3) {
  const EigItem* ei = (const EigItem*)st->d_eitems.p;
  const EigItem* e3 = (const EigItem*)st->d_e3items.p;
  const float* U = (const float*)st->eigU32.p;
  const float* L = (const float*)st->eigL32.p;
  TI* P = (TI*)st->P16.p;
  TI* P3 = (TI*)st->P3_16.p;
  float* part = (float*)st->part32.p;
  if (stage == 1 && nb > 0)
    hipLaunchKernelGGL(k_absorb_bf16c<TI>, dim3(64, nb), dim3(512), 0, ctx->stream,
                       (const FastItem*)st->d_k1items.p, (const __bf16*)st->lay16.p, U, P);
  if (stage == 2 && nb > 0)
    hipLaunchKernelGGL((k_rot34_bf16<true, TI>), dim3(EIG_NPART, nb), dim3(512), 0, ctx->stream, ei,
                       (const TI*)P, U, L, P, part);
  if (stage == 3 && nb > 0)
    hipLaunchKernelGGL((k_gram12_bf16<32, TI>), dim3(EIG_NPART, nb), dim3(256), 0, ctx->stream, ei, (const TI*)P,
                       L, part);
  if (stage == 4 && n3 > 0) {
    hipLaunchKernelGGL(k_rot12_d3_bf16<TI>, dim3(4, n3), dim3(256), 0, ctx->stream, e3,
                       (const __bf16*)st->lay3_16.p, U, P3);
    hipLaunchKernelGGL(k_rot3_d3_bf16<TI>, dim3(EIG_NPART, n3), dim3(256), 0, ctx->stream, e3, (const TI*)P3, U, L,
                       P3, part);
    hipLaunchKernelGGL((k_gram12_bf16<1, TI>), dim3(EIG_NPART, n3), dim3(256), 0, ctx->stream, e3, (const TI*)P3,
                       L, part);
  }
}

// the compensated flavours (FP32 R and Q, split 16-bit operands)
template <class Prec>
static void eigs_stage(itnbp_ctx* ctx, FastState* st, int stage, int nb, int n3) {
  const EigItem* ei = (const EigItem*)st->d_eitems.p;
  const EigItem* e3 = (const EigItem*)st->d_e3items.p;
  const float* U = (const float*)st->eigU32.p;
  const float* L = (const float*)st->eigL32.p;
  float* P = (float*)st->P32.p;
  float* P3 = (float*)st->P3_32.p;
  float* part = (float*)st->part32.p;
  if (stage == 1 && nb > 0)
    hipLaunchKernelGGL(k_rot12_s<Prec>, dim3(64, nb), dim3(512), 0, ctx->stream, (const FastItem*)st->d_k1items.p,
                       (const __bf16*)st->lay16.p, U, P);
  if (stage == 2 && nb > 0)
    hipLaunchKernelGGL(k_rot34_s<Prec>, dim3(EIG_NPART, nb), dim3(512), 0, ctx->stream, ei, (const float*)P, U, L, P,
                       part);
  if (stage == 3 && nb > 0)
    hipLaunchKernelGGL((k_gram12_s<Prec, 32>), dim3(EIG_NPART, nb), dim3(256), 0, ctx->stream, ei, (const float*)P, L, part);
  if (stage == 4 && n3 > 0) {
    hipLaunchKernelGGL(k_rot12_d3_s<Prec>, dim3(4, n3), dim3(256), 0, ctx->stream, e3, (const __bf16*)st->lay3_16.p, U,
                       P3);
    hipLaunchKernelGGL(k_rot3_d3_s<Prec>, dim3(EIG_NPART, n3), dim3(256), 0, ctx->stream, e3, (const float*)P3, U, L, P3,
                       part);
    hipLaunchKernelGGL((k_gram12_s<Prec, 1>), dim3(EIG_NPART, n3), dim3(256), 0, ctx->stream, e3, (const float*)P3, L, part);
  }
}

static void eig16(itnbp_ctx* ctx, FastState* st, int stage, int nb, int n3) {
  if (st->x3 && st->h2) eigs_stage<PrecH2>(ctx, st, stage, nb, n3);
  else if (st->x3) eigs_stage<PrecX3>(ctx, st, stage, nb, n3);
  else if (eig_bf16_intermediates()) eig16_stage<__bf16>(ctx, st, stage, nb, n3);
  else eig16_stage<_Float16>(ctx, st, stage, nb, n3);
}

// one sweep of the eigenbasis pipeline over every bulk and degree-3 site
static int eig_sweep(itnbp_ctx* ctx, FastState* st, int acc_f64, const void* msg_in, void* msg_out,
                     MsgStat* stats, int normalize) {
  const int nb = (int)st->eitems.size();
  const int n3 = (int)st->e3items.size();
  const int ne4 = (int)st->eig_moff.size();
  if (ne4 == 0) return ITNBP_OK;
  const EigItem* ei = (const EigItem*)st->d_eitems.p;
  const EigItem* e3 = (const EigItem*)st->d_e3items.p;
  // seeded by the previous eigenvectors, except every 16th launch, which starts
  // from the identity so that rounding in the accumulated rotations cannot build up
  const int warm = (st->eig_warm && st->eig_count % 16 != 0) ? 1 : 0;
  const int e0 = prof_mark(ctx);
  if (acc_f64)
    hipLaunchKernelGGL((k_msg_eigh<double>), dim3((ne4 + 3) / 4), dim3(256), 0, ctx->stream,
                       (const long long*)st->d_eig_moff.p, ne4, (const double*)msg_in, (double*)st->eigU64.p,
                       (double*)st->eigL64.p, (float*)nullptr, (float*)nullptr, (int*)nullptr, warm, 1e-28);
  else
    hipLaunchKernelGGL((k_msg_eigh<float>), dim3((ne4 + 3) / 4), dim3(256), 0, ctx->stream,
                       (const long long*)st->d_eig_moff.p, ne4, (const float*)msg_in, (double*)st->eigU64.p,
                       (double*)st->eigL64.p, (float*)st->eigU32.p, (float*)st->eigL32.p, (int*)st->eigLexp.p,
                       warm, 1e-14);
  st->eig_warm = true;
  ++st->eig_count;
  const int e1 = prof_mark(ctx);
  if (nb > 0) {
    // rot12.  FP64: the rot34 kernel without Grams (items k1e map U1, U2 to the
    // U[2], U[3] slots, psi layout B to R and the R buffer to Q), whose persistent
    // prefetching loop hides the FP64 products.  BF16: the absorb kernel, whose
    // two workgroups per CU stream faster than the rot34 loop's one.
    if (acc_f64) {
      const EigItem* k1e = (const EigItem*)st->d_k1eitems.p;
      hipLaunchKernelGGL(k_rot34_f64<false>, dim3(EIG_NPART, nb), dim3(256), 0, ctx->stream, k1e,
                         (const double*)st->lay64.p, (const double*)st->eigU64.p, (const double*)st->eigL64.p,
                         (double*)st->P64.p, (double*)nullptr);
    } else {
      eig16(ctx, st, 1, nb, n3);
    }
  }
  const int e2 = prof_mark(ctx);
  if (nb > 0) {
    if (acc_f64)
      hipLaunchKernelGGL(k_rot34_f64<true>, dim3(EIG_NPART, nb), dim3(256), 0, ctx->stream, ei, (const double*)st->P64.p,
                         (const double*)st->eigU64.p, (const double*)st->eigL64.p, (double*)st->P64.p,
                         (double*)st->part64.p);
    else
      eig16(ctx, st, 2, nb, n3);
  }
  const int e3m = prof_mark(ctx);
  if (nb > 0) {
    if (acc_f64)
      hipLaunchKernelGGL(k_gram12_f64<32>, dim3(EIG_NPART, nb), dim3(256), 0, ctx->stream, ei,
                         (const double*)st->P64.p, (const double*)st->eigL64.p, (double*)st->part64.p);
    else
      eig16(ctx, st, 3, nb, n3);
  }
  const int e4 = prof_mark(ctx);
  if (n3 > 0) {
    if (acc_f64) {
      hipLaunchKernelGGL(k_rot12_d3_f64, dim3(4, n3), dim3(256), 0, ctx->stream, e3, (const double*)st->lay3_64.p,
                         (const double*)st->eigU64.p, (double*)st->P3_64.p);
      hipLaunchKernelGGL(k_rot3_d3_f64, dim3(EIG_NPART, n3), dim3(256), 0, ctx->stream, e3,
                         (const double*)st->P3_64.p, (const double*)st->eigU64.p, (const double*)st->eigL64.p,
                         (double*)st->P3_64.p, (double*)st->part64.p);
      hipLaunchKernelGGL(k_gram12_f64<1>, dim3(EIG_NPART, n3), dim3(256), 0, ctx->stream, e3,
                         (const double*)st->P3_64.p, (const double*)st->eigL64.p, (double*)st->part64.p);
    } else {
      eig16(ctx, st, 4, nb, n3);
    }
  }
  const int e5 = prof_mark(ctx);
  if (acc_f64) {
    if (nb > 0)
      hipLaunchKernelGGL((k_eig_finalize<double, double>), dim3(4, nb), dim3(256), 0, ctx->stream, ei,
                         (const double*)st->part64.p, (const double*)st->eigU64.p, (const double*)msg_in,
                         (double*)msg_out, stats, normalize, (const int*)nullptr, (const int*)nullptr);
    if (n3 > 0)
      hipLaunchKernelGGL((k_eig_finalize<double, double>), dim3(3, n3), dim3(256), 0, ctx->stream, e3,
                         (const double*)st->part64.p, (const double*)st->eigU64.p, (const double*)msg_in,
                         (double*)msg_out, stats, normalize, (const int*)nullptr, (const int*)nullptr);
  } else {
    if (nb > 0)
      hipLaunchKernelGGL((k_eig_finalize<float, float>), dim3(4, nb), dim3(256), 0, ctx->stream, ei,
                         (const float*)st->part32.p, (const double*)st->eigU64.p, (const float*)msg_in,
                         (float*)msg_out, stats, normalize, (const int*)st->sexp.p, (const int*)st->eigLexp.p);
    if (n3 > 0)
      hipLaunchKernelGGL((k_eig_finalize<float, float>), dim3(3, n3), dim3(256), 0, ctx->stream, e3,
                         (const float*)st->part32.p, (const double*)st->eigU64.p, (const float*)msg_in,
                         (float*)msg_out, stats, normalize, (const int*)st->sexp.p, (const int*)st->eigLexp.p);
  }
  const int e6 = prof_mark(ctx);
  prof_add(ctx, "eigh", e0, e1);
  const char* fl = acc_f64 ? "f64" : (st->x3 ? (st->h2 ? "h2" : "x3") : "bf16");
  static const char* const names[4][4] = {{"rot12_f64", "rot34_f64", "gram12_f64", "deg3_f64"},
                                          {"rot12_x3", "rot34_x3", "gram12_x3", "deg3_x3"},
                                          {"rot12_h2", "rot34_h2", "gram12_h2", "deg3_h2"},
                                          {"rot12_bf16", "rot34_bf16", "gram12_bf16", "deg3_bf16"}};
  const int fi = fl[0] == 'f' ? 0 : fl[0] == 'x' ? 1 : fl[0] == 'h' ? 2 : 3;
  prof_add(ctx, names[fi][0], e1, e2);
  prof_add(ctx, names[fi][1], e2, e3m);
  prof_add(ctx, names[fi][2], e3m, e4);
  prof_add(ctx, names[fi][3], e4, e5);
  prof_add(ctx, "finalize", e5, e6);
  ITNBP_HIPCHK(ctx, hipGetLastError());
  return ITNBP_OK;
}

int fast_sweep(itnbp_net* net, int acc_f64, const void* msg_in, void* msg_out, MsgStat* stats,
               int normalize) {
  itnbp_ctx* ctx = net->ctx;
  FastState* st = state_of(net);
  if (!st) return ITNBP_OK;
  if (st->eig) return eig_sweep(ctx, st, acc_f64, msg_in, msg_out, stats, normalize);
  const int ni = (int)st->items.size();
  if (ni == 0) return ITNBP_OK;
  const FastItem* items = (const FastItem*)st->d_items.p;
  const int chunk = tune_chunk(net, st, acc_f64, msg_in);
  launch_all_pairs(ctx, st, acc_f64, msg_in, chunk);
  const int e2 = prof_mark(ctx);
  if (!acc_f64) {
    hipLaunchKernelGGL((k_fast_finalize<float, float>), dim3(2, ni), dim3(256), 0, ctx->stream, items,
                       (const float*)st->part32.p, (const float*)msg_in, (float*)msg_out, stats, normalize,
                       (const int*)st->sexp.p);
  } else {
    hipLaunchKernelGGL((k_fast_finalize<double, double>), dim3(2, ni), dim3(256), 0, ctx->stream, items,
                       (const double*)st->part64.p, (const double*)msg_in, (double*)msg_out, stats, normalize,
                       (const int*)nullptr);
  }
  prof_add(ctx, "finalize", e2, prof_mark(ctx));
  ITNBP_HIPCHK(ctx, hipGetLastError());
  return ITNBP_OK;
}

// ---- density matrices on the FP64 eigenbasis rotations ----------------------
// Every message symmetric and set, at least ITNBP_RDM_FAST_MIN (default 16)
// eligible sites among the requests: the eigendecomposition of every message
// into an eligible site is a fixed cost.
struct RotSites {
  std::vector<int> sb, s3;    // bulk and degree-3 sites rotated
  std::vector<int> item;      // vertex -> index into [bulk items, degree-3 items], -1 if none
  std::vector<EigItem> eitems;  // the rotated sites' items, bulk then degree-3
  char* tail = nullptr;       // device scratch after the item lists
};

static int rotated_sites_setup(itnbp_net* net, const std::vector<int>& want, size_t tail_bytes, RotSites& rs,
                               bool& ok) {
  itnbp_ctx* ctx = net->ctx;
  ok = false;
  if (want.empty() || !use_eig(net) || getenv("ITNBP_DISABLE_FAST")) return ITNBP_OK;
  for (int m = 0; m < 2 * net->ne; ++m)
    if (!net->msg_set[m]) return ITNBP_OK;
  int min_sites = 16;
  if (const char* e = getenv("ITNBP_RDM_FAST_MIN")) min_sites = std::max(1, atoi(e));
  int cand = 0;
  for (int v : want)
    if (v >= 0 && v < net->nv && (eligible(net, v) || eligible3(net, v))) ++cand;
  if (cand < min_sites) return ITNBP_OK;
  std::vector<int> seq(2 * net->ne);
  for (int m = 0; m < 2 * net->ne; ++m) seq[m] = m;
  std::vector<char> handled;
  const int fr = fast_prepare(net, seq, 1, handled);
  if (fr == ITNBP_ERR_UNSUPPORTED) return ITNBP_OK;
  ITNBP_RETCHK(fr);
  FastState* st = state_of(net);
  if (!st->eig || st->eig_moff.empty()) return ITNBP_OK;
  std::vector<int> bidx(net->nv, -1), tidx(net->nv, -1);
  for (size_t bi = 0; bi < st->bulk.size(); ++bi) bidx[st->bulk[bi]] = (int)bi;
  for (size_t j = 0; j < st->deg3_used.size(); ++j) tidx[st->deg3_used[j]] = (int)j;
  rs.sb.clear();
  rs.s3.clear();
  rs.item.assign(net->nv, -1);
  std::vector<char> seen(net->nv, 0);
  for (int v : want) {
    if (v < 0 || v >= net->nv || seen[v]) continue;
    seen[v] = 1;
    if (bidx[v] >= 0) rs.sb.push_back(v);
    else if (tidx[v] >= 0) rs.s3.push_back(v);
  }
  const int nbs = (int)rs.sb.size(), n3s = (int)rs.s3.size();
  if (nbs + n3s == 0) return ITNBP_OK;
  // device lists: [rot34 items nbs][degree-3 items n3s][rot12 items nbs]
  rs.eitems.clear();
  for (int k = 0; k < nbs; ++k) {
    rs.item[rs.sb[k]] = k;
    rs.eitems.push_back(st->eitems[bidx[rs.sb[k]]]);
  }
  for (int k = 0; k < n3s; ++k) {
    rs.item[rs.s3[k]] = nbs + k;
    rs.eitems.push_back(st->e3items[tidx[rs.s3[k]]]);
  }
  std::vector<EigItem> all = rs.eitems;
  for (int k = 0; k < nbs; ++k) all.push_back(st->k1eitems[bidx[rs.sb[k]]]);
  const size_t b_items = all.size() * sizeof(EigItem);
  const size_t at_tail = (b_items + 255) / 256 * 256;
  ITNBP_RETCHK(ensure_buf(ctx, st->rdm, at_tail + tail_bytes));
  char* base = (char*)st->rdm.p;
  ITNBP_HIPCHK(ctx, hipMemcpyAsync(base, all.data(), b_items, hipMemcpyHostToDevice, ctx->stream));
  rs.tail = base + at_tail;
  const EigItem* d_e = (const EigItem*)base;
  const EigItem* d_3 = d_e + nbs;
  const EigItem* d_k1 = d_e + nbs + n3s;
  const int ne4 = (int)st->eig_moff.size();
  const double* U = (const double*)st->eigU64.p;
  const double* L = (const double*)st->eigL64.p;
  // cold start: the rotations must be orthogonal to rounding
  hipLaunchKernelGGL((k_msg_eigh<double>), dim3((ne4 + 3) / 4), dim3(256), 0, ctx->stream,
                     (const long long*)st->d_eig_moff.p, ne4, (const double*)net->d_msg[net->cur],
                     (double*)st->eigU64.p, (double*)st->eigL64.p, (float*)nullptr, (float*)nullptr, (int*)nullptr, 0,
                     1e-28);
  st->eig_warm = true;
  if (nbs > 0) {
    double* P = (double*)st->P64.p;
    hipLaunchKernelGGL(k_rot34_f64<false>, dim3(EIG_NPART, nbs), dim3(256), 0, ctx->stream, d_k1,
                       (const double*)st->lay64.p, U, L, P, (double*)nullptr);
    hipLaunchKernelGGL(k_rot34_f64<false>, dim3(EIG_NPART, nbs), dim3(256), 0, ctx->stream, d_e, (const double*)P, U,
                       L, P, (double*)nullptr);
  }
  if (n3s > 0) {
    double* P3 = (double*)st->P3_64.p;
    hipLaunchKernelGGL(k_rot12_d3_f64, dim3(4, n3s), dim3(256), 0, ctx->stream, d_3, (const double*)st->lay3_64.p, U,
                       P3);
    hipLaunchKernelGGL(k_rot3_d3_f64, dim3(EIG_NPART, n3s), dim3(256), 0, ctx->stream, d_3, (const double*)P3, U, L,
                       P3, (double*)st->part64.p);
  }
  ITNBP_HIPCHK(ctx, hipGetLastError());
  ok = true;
  return ITNBP_OK;
}

// One-site density matrices of the listed vertices: rho of request i
// (unnormalised, column-major 2 x 2) goes to out + offs[i]; done[i] = 1 for the
// requests served here, the others are left to the generic kernels.
int fast_site_rdm(itnbp_net* net, int n, const int* verts, const long long* offs, double* out,
                  std::vector<char>& done) {
  itnbp_ctx* ctx = net->ctx;
  done.assign(n, 0);
  if (n <= 0) return ITNBP_OK;
  std::vector<int> want(verts, verts + n);
  // requests served: eligible sites (item >= 0 after the setup)
  const size_t tail = (size_t)n * (RHO_NPART * 3 * 8 + 8 + 4) + 256;
  RotSites rs;
  bool ok = false;
  const int e0 = prof_mark(ctx);
  ITNBP_RETCHK(rotated_sites_setup(net, want, tail, rs, ok));
  if (!ok) return ITNBP_OK;
  FastState* st = state_of(net);
  const int nbs = (int)rs.sb.size(), n3s = (int)rs.s3.size();
  std::vector<int> ritem;
  std::vector<long long> roff;
  for (int i = 0; i < n; ++i) {
    const int v = verts[i];
    if (v < 0 || v >= net->nv || rs.item[v] < 0) continue;
    done[i] = 1;
    ritem.push_back(rs.item[v]);
    roff.push_back(offs[i]);
  }
  const int nr = (int)ritem.size();
  double* rpart = (double*)rs.tail;
  char* tb = rs.tail + (size_t)(nbs + n3s) * RHO_NPART * 3 * 8;
  long long* d_off = (long long*)tb;
  int* d_item = (int*)(tb + (size_t)nr * 8);
  ITNBP_HIPCHK(ctx, hipMemcpyAsync(d_off, roff.data(), nr * 8, hipMemcpyHostToDevice, ctx->stream));
  ITNBP_HIPCHK(ctx, hipMemcpyAsync(d_item, ritem.data(), nr * 4, hipMemcpyHostToDevice, ctx->stream));
  const EigItem* d_e = (const EigItem*)st->rdm.p;
  const double* L = (const double*)st->eigL64.p;
  if (nbs > 0)
    hipLaunchKernelGGL(k_rho_sgram_f64, dim3(RHO_NPART, nbs), dim3(256), 0, ctx->stream, d_e,
                       (const double*)st->P64.p, L, FNR / 2, rpart);
  if (n3s > 0)
    hipLaunchKernelGGL(k_rho_sgram_f64, dim3(RHO_NPART, n3s), dim3(256), 0, ctx->stream, d_e + nbs,
                       (const double*)st->P3_64.p, L, D3_NBLK / 2, rpart + (size_t)nbs * RHO_NPART * 3);
  hipLaunchKernelGGL(k_rho_finalize, dim3((nr + 255) / 256), dim3(256), 0, ctx->stream, (const double*)rpart,
                     (const int*)d_item, (const long long*)d_off, nr, out);
  ITNBP_HIPCHK(ctx, hipGetLastError());
  prof_add(ctx, "rdm1_eig", e0, prof_mark(ctx));
  return ITNBP_OK;
}

// Half-pair environments of (site, axis) requests for two-site RDMs:
// A[(s + 2 b) + 64 (s' + 2 b')] = sum over the other bonds of psi psi' times their
// messages, the layout of the generic open contraction with one open bond,
// written at out + offs[i]; done[i] = 1 for the requests served here.
int fast_pair_halves(itnbp_net* net, int n, const int* sites, const int* axes, const long long* offs, double* out,
                     std::vector<char>& done) {
  itnbp_ctx* ctx = net->ctx;
  done.assign(n, 0);
  if (n <= 0) return ITNBP_OK;
  std::vector<int> want(sites, sites + n);
  const size_t tail = (size_t)n * (sizeof(PairTask) + (size_t)PG_NPART * 3 * FBLK * 8) + 256;
  RotSites rs;
  bool ok = false;
  const int e0 = prof_mark(ctx);
  ITNBP_RETCHK(rotated_sites_setup(net, want, tail, rs, ok));
  if (!ok) return ITNBP_OK;
  FastState* st = state_of(net);
  std::vector<PairTask> tasks;
  for (int i = 0; i < n; ++i) {
    const int v = sites[i];
    if (v < 0 || v >= net->nv || rs.item[v] < 0) continue;
    const EigItem& e = rs.eitems[rs.item[v]];
    const int deg3 = rs.item[v] >= (int)rs.sb.size() ? 1 : 0;
    const int c = axes[i] - 1;
    if (c < 0 || c >= (deg3 ? 3 : 4)) continue;
    PairTask t{};
    t.Q = e.Q;
    for (int k = 0; k < 4; ++k) t.L[k] = e.L[k];
    t.U = e.U[c];
    t.out = offs[i];
    t.c = c;
    t.deg3 = deg3;
    tasks.push_back(t);
    done[i] = 1;
  }
  const int nt = (int)tasks.size();
  if (nt == 0) return ITNBP_OK;
  PairTask* d_t = (PairTask*)rs.tail;
  double* part = (double*)(rs.tail + ((size_t)nt * sizeof(PairTask) + 255) / 256 * 256);
  ITNBP_HIPCHK(ctx, hipMemcpyAsync(d_t, tasks.data(), nt * sizeof(PairTask), hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_pair_gram_f64, dim3(PG_NPART, nt), dim3(256), 0, ctx->stream, (const PairTask*)d_t,
                     (const double*)st->P64.p, (const double*)st->P3_64.p, (const double*)st->eigL64.p, part);
  hipLaunchKernelGGL(k_pair_finalize, dim3(nt), dim3(256), 0, ctx->stream, (const PairTask*)d_t, (const double*)part,
                     (const double*)st->eigU64.p, out);
  ITNBP_HIPCHK(ctx, hipGetLastError());
  prof_add(ctx, "rdm2_eig", e0, prof_mark(ctx));
  return ITNBP_OK;
}
