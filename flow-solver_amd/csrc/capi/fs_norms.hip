// Sums of squares of a residual and an update held in device memory: the two
// norm loops of solve() (solve.C:785-793 for R, solve.C:885-893 for dQ), as
// device-pointer entry points for callers that keep their own arrays
// resident.
//
// Both entry points reduce in the same fixed order as the solver's own norms
// (chunks of kNormChunkNodes nodes, a 256-thread tree per chunk, then one
// wave over the chunk partials), so for the same arrays they return the same
// bits as each other and as fs_solve's r2.
//
// fs_state_norms allocates its chunk partials per call and waits for the
// stream before freeing them.
#include <hip/hip_runtime.h>

#include <string>

#include "../kernels/fs_kernels.hpp"
#include "fs_capi.h"
#include "fs_engine.hpp"

namespace {

constexpr int kThreads = 256;

// chunk ch: sum of R^2 and dQ^2 over the 4 components of nodes
// [ch*C, min(nn, (ch+1)*C)), C = kNormChunkNodes
__global__ void __launch_bounds__(kThreads) k_state_norm_chunk(int64_t nn, const double* __restrict__ R,
                                                               const double* __restrict__ dQ,
                                                               double* __restrict__ part) {
  __shared__ double sr[kThreads], sd[kThreads];
  const int64_t ch = blockIdx.x;
  const int64_t b = 4 * ch * fsk::kNormChunkNodes;
  const int64_t nend = (ch + 1) * fsk::kNormChunkNodes < nn ? (ch + 1) * fsk::kNormChunkNodes : nn;
  const int64_t e = 4 * nend;
  double a = 0.0, d = 0.0;
  for (int64_t i = b + threadIdx.x; i < e; i += kThreads) {
    a += R[i] * R[i];
    d += dQ[i] * dQ[i];
  }
  sr[threadIdx.x] = a;
  sd[threadIdx.x] = d;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sr[threadIdx.x] += sr[threadIdx.x + w];
      sd[threadIdx.x] += sd[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * ch] = sr[0];
    part[2 * ch + 1] = sd[0];
  }
}

// one wave: lane l sums chunks l, l+64, ...; then a fixed butterfly
__global__ void __launch_bounds__(64) k_state_norm_final(int64_t nchunk, const double* __restrict__ part,
                                                         double* __restrict__ out) {
  const int l = threadIdx.x;
  double a = 0.0, d = 0.0;
  for (int64_t ch = l; ch < nchunk; ch += 64) {
    a += part[2 * ch];
    d += part[2 * ch + 1];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off, 64);
    d += __shfl_xor(d, off, 64);
  }
  if (l == 0) {
    out[0] = a;
    out[1] = d;
  }
}

void enqueue(int64_t nn, const double* R, const double* dQ, double* part, double* out, hipStream_t st) {
  const int64_t nchunk = (nn + fsk::kNormChunkNodes - 1) / fsk::kNormChunkNodes;
  if (nchunk > 0)
    hipLaunchKernelGGL(k_state_norm_chunk, dim3((unsigned)nchunk), dim3(kThreads), 0, st, nn, R, dQ, part);
  hipLaunchKernelGGL(k_state_norm_final, dim3(1), dim3(64), 0, st, nchunk, part, out);
}

int check_args(int64_t nn, const double* R, const double* dQ, double* out) {
  if (nn < 0 || !out || (nn > 0 && (!R || !dQ))) return fse::set_error(FS_ERR_ARG, "bad arguments");
  // 32-bit grid: at most 2^31 - 1 chunks
  if ((nn + fsk::kNormChunkNodes - 1) / fsk::kNormChunkNodes > 0x7fffffffLL)
    return fse::set_error(FS_ERR_ARG, "node count too large");
  return FS_OK;
}

}  // namespace

extern "C" {

int fs_state_norms(int64_t nn, const double* d_R, const double* d_dQ, double* d_out, void* stream) {
  int rc = check_args(nn, d_R, d_dQ, d_out);
  if (rc) return rc;
  const int64_t nchunk = (nn + fsk::kNormChunkNodes - 1) / fsk::kNormChunkNodes;
  double* part = nullptr;
  FSE_HIP(hipMalloc((void**)&part, (size_t)(nchunk > 0 ? 2 * nchunk : 2) * sizeof(double)));
  hipStream_t st = (hipStream_t)stream;
  enqueue(nn, d_R, d_dQ, part, d_out, st);
  hipError_t e = hipGetLastError();
  hipError_t s = hipStreamSynchronize(st);
  (void)hipFree(part);
  if (e != hipSuccess) return fse::set_error(FS_ERR_HIP, std::string("norm launch: ") + hipGetErrorString(e));
  if (s != hipSuccess) return fse::set_error(FS_ERR_HIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(s));
  return FS_OK;
}

}  // extern "C"
