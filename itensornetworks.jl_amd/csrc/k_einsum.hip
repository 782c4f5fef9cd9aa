// Device executor of small real tensor-network contractions (itnbp_einsum):
// the plan of itnbp_einsum.cpp run as unary steps (permute + trace + sum, one
// thread per output element) and batched FP64 GEMMs on v_mfma_f64_16x16x4_f64.
//
// Its user is the contraction of a partition's site tensors into the
// partition's tensor, the factor that BP over generalised partitions absorbs
// its messages into (src/caches/beliefpropagationcache.jl:20-35 builds the
// partitioned cache; abstractbeliefpropagationcache.jl:225-239 contracts a
// partition's factors with its incoming messages), and the partial traces that
// take a member site's density matrix out of its partition's.
#include <algorithm>
#include <string>
#include <vector>

#include "itnbp_einsum.hpp"
#include "itnbp_internal.hpp"
#include "itnbp_ops.hpp"

namespace {

using einsum_plan::kMaxRank;

struct UnaryArgs {
  int nout, nred;
  long long out_total, red_total;
  long long out_dims[kMaxRank], out_strides[kMaxRank];
  long long red_dims[kMaxRank], red_strides[kMaxRank];
};

// dst[o] = sum_r src[offset(o) + offset(r)]; grid-stride over the outputs
__global__ __launch_bounds__(256) void k_es_unary(const double* __restrict__ src, double* __restrict__ dst,
                                                  UnaryArgs a) {
  const long long step = (long long)gridDim.x * blockDim.x;
  for (long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x; o < a.out_total; o += step) {
    long long rem = o, base = 0;
    for (int i = 0; i < a.nout; ++i) {
      const long long q = rem / a.out_dims[i];
      base += (rem - q * a.out_dims[i]) * a.out_strides[i];
      rem = q;
    }
    double s = 0.0;
    for (long long r = 0; r < a.red_total; ++r) {
      long long rr = r, off = base;
      for (int i = 0; i < a.nred; ++i) {
        const long long q = rr / a.red_dims[i];
        off += (rr - q * a.red_dims[i]) * a.red_strides[i];
        rr = q;
      }
      s += src[off];
    }
    dst[o] = s;
  }
}

typedef double f64x4 __attribute__((ext_vector_type(4)));

// C[m + M (n + N b)] = sum_k A[m + M (k + K b)] Bt[n + N (k + K b)].
// 64 x 64 output tile per workgroup, 32 x 32 per wave (2 x 2 MFMA tiles of
// 16 x 16, K in steps of 4); operands straight from global memory (both have
// the output index fastest, so the 16 lanes of a k-row read 128 contiguous
// bytes).  v_mfma_f64_16x16x4_f64 lane layout: A[i][k] in lane i + 16 k,
// B[k][j] in lane j + 16 k, D[4 v + l / 16][l % 16] in lane l, value v.
// grid (ceil(M / 64), ceil(N / 64), min(batch, 65535)); batches beyond
// gridDim.z loop.
__global__ __launch_bounds__(256) void k_es_gemm(const double* __restrict__ A, const double* __restrict__ Bt,
                                                 double* __restrict__ C, long long M, long long N, long long K,
                                                 long long batch) {
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int li = l & 15, lk = l >> 4;
  const long long m0 = (long long)blockIdx.x * 64 + 32 * (wv & 1);
  const long long n0 = (long long)blockIdx.y * 64 + 32 * (wv >> 1);
  for (long long b = blockIdx.z; b < batch; b += gridDim.z) {
    const double* Ab = A + b * M * K;
    const double* Bb = Bt + b * N * K;
    f64x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = f64x4{0, 0, 0, 0};
    for (long long k0 = 0; k0 < K; k0 += 4) {
      const long long k = k0 + lk;
      const bool kin = k < K;
      double a[2], bv[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const long long m = m0 + 16 * i + li, n = n0 + 16 * i + li;
        a[i] = (kin && m < M) ? Ab[m + M * k] : 0.0;
        bv[i] = (kin && n < N) ? Bb[n + N * k] : 0.0;
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], bv[j], acc[i][j], 0, 0, 0);
    }
    double* Cb = C + b * M * N;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const long long m = m0 + 16 * i + 4 * v + lk, n = n0 + 16 * j + li;
          if (m < M && n < N) Cb[m + M * n] = acc[i][j][v];
        }
  }
}

// every slot of one contraction, back to back in one device buffer
constexpr long long kMaxSlotElems = 1LL << 28;   // 2 GiB per tensor
constexpr long long kMaxTotalElems = 1LL << 30;  // 8 GiB per call

int run_plan(itnbp_ctx* ctx, const einsum_plan::Plan& p, double* base, const std::vector<long long>& off) {
  for (const auto& s : p.steps) {
    if (s.gemm) {
      const auto& g = s.g;
      const long long gx = (g.m + 63) / 64, gy = (g.n + 63) / 64;
      if (gx > 0x7fffffffLL || gy > 65535)
        return itnbp_fail(ctx, ITNBP_ERR_UNSUPPORTED, "a product of the contraction is too wide for one launch");
      const unsigned gz = (unsigned)std::min<long long>(g.batch, 65535);
      hipLaunchKernelGGL(k_es_gemm, dim3((unsigned)gx, (unsigned)gy, gz), dim3(256), 0, ctx->stream,
                         base + off[g.a], base + off[g.bt], base + off[g.dst], (long long)g.m, (long long)g.n,
                         (long long)g.k, (long long)g.batch);
    } else {
      const auto& u = s.u;
      UnaryArgs a{};
      a.nout = (int)u.out_dims.size();
      a.nred = (int)u.red_dims.size();
      a.out_total = 1;
      a.red_total = 1;
      for (int i = 0; i < a.nout; ++i) {
        a.out_dims[i] = u.out_dims[i];
        a.out_strides[i] = u.out_strides[i];
        a.out_total *= u.out_dims[i];
      }
      for (int i = 0; i < a.nred; ++i) {
        a.red_dims[i] = u.red_dims[i];
        a.red_strides[i] = u.red_strides[i];
        a.red_total *= u.red_dims[i];
      }
      const long long blocks = std::max(1LL, std::min((a.out_total + 255) / 256, 8192LL));
      hipLaunchKernelGGL(k_es_unary, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, base + off[u.src],
                         base + off[u.dst], a);
    }
    ITNBP_HIPCHK(ctx, hipGetLastError());
  }
  return ITNBP_OK;
}

}  // namespace

extern "C" {

int itnbp_einsum(itnbp_ctx* ctx, int n, const int* ranks, const int* labels, const int* label_dims, int nlabels,
                 const double* data, int out_rank, const int* out_labels, double* out) {
  if (!ctx) return ITNBP_ERR_ARG;
  if (n < 1 || !ranks || !data || !out || nlabels < 0 || (nlabels > 0 && !label_dims) || out_rank < 0 ||
      (out_rank > 0 && !out_labels))
    return itnbp_fail(ctx, ITNBP_ERR_ARG, "itnbp_einsum: malformed arguments");
  return itnbp_guarded(ctx, [&]() -> int {
    int nax = 0;
    for (int i = 0; i < n; ++i) {
      if (ranks[i] < 0) return itnbp_fail(ctx, ITNBP_ERR_ARG, "itnbp_einsum: negative rank");
      nax += ranks[i];
    }
    if (nax > 0 && !labels) return itnbp_fail(ctx, ITNBP_ERR_ARG, "itnbp_einsum: labels missing");
    std::vector<int64_t> dims(nlabels);
    for (int x = 0; x < nlabels; ++x) dims[x] = label_dims[x];
    einsum_plan::Plan p;
    std::string err;
    if (einsum_plan::plan(n, ranks, labels, dims.data(), nlabels, out_rank, out_labels, kMaxSlotElems, p, err) != 0)
      return itnbp_fail(ctx, ITNBP_ERR_ARG, "itnbp_einsum: " + err);
    std::vector<long long> off(p.sizes.size());
    long long tot = 0;
    for (size_t i = 0; i < p.sizes.size(); ++i) {
      off[i] = tot;
      tot += p.sizes[i];
    }
    if (tot > kMaxTotalElems)
      return itnbp_fail(ctx, ITNBP_ERR_UNSUPPORTED, "itnbp_einsum: the contraction needs more than 8 GiB");
    DevBuf buf;
    std::unique_ptr<DevBuf, void (*)(DevBuf*)> guard(&buf, [](DevBuf* b) { free_buf(*b); });
    ITNBP_RETCHK(ensure_buf(ctx, buf, (size_t)std::max(tot, 1LL) * sizeof(double)));
    double* base = static_cast<double*>(buf.p);
    const long long nin = off.size() > (size_t)p.ninputs ? off[p.ninputs] : tot;
    ITNBP_HIPCHK(ctx, hipMemcpyAsync(base, data, (size_t)nin * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    ITNBP_RETCHK(run_plan(ctx, p, base, off));
    ITNBP_HIPCHK(ctx, hipMemcpyAsync(out, base + off[p.result], (size_t)p.sizes[p.result] * sizeof(double),
                                     hipMemcpyDeviceToHost, ctx->stream));
    ITNBP_HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return ITNBP_OK;
  });
}

}  // extern "C"
