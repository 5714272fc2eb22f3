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
// stream before freeing them.  fs_state_norms_noalloc keeps the partials in
// the file-scope array `scratch` below (kNoAllocMaxChunks chunks), allocates
// nothing and only enqueues work, so it can run under stream capture or in a
// loop that must not allocate.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../kernels/fs_kernels.hpp"
#include "fs_capi.h"
#include "fs_engine.hpp"

namespace {

constexpr int kThreads = 256;
// 65,536 chunks of 1,024 nodes: up to 67,108,864 nodes, 1 MiB of partials
constexpr int64_t kNoAllocMaxChunks = 65536;
constexpr int kMaxDevices = 64;

// Chunk partials of fs_state_norms_noalloc, one copy per device for the life
// of the process.  Host code never passes `scratch` itself to a kernel: the
// name is a host-side stand-in, and the device address comes from
// hipGetSymbolAddress (scratch_on_current_device).
static __device__ __align__(256) double scratch[2 * kNoAllocMaxChunks];

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

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

int scratch_on_current_device(double** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fse::set_error(FS_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return fse::set_error(FS_ERR_HIP, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fse::set_error(FS_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
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

int fs_state_norms_noalloc(int64_t nn, const double* d_R, const double* d_dQ, double* d_out, void* stream) {
  int rc = check_args(nn, d_R, d_dQ, d_out);
  if (rc) return rc;
  if ((nn + fsk::kNormChunkNodes - 1) / fsk::kNormChunkNodes > kNoAllocMaxChunks)
    return fse::set_error(FS_ERR_ARG, "fs_state_norms_noalloc: at most " +
                                          std::to_string(kNoAllocMaxChunks * fsk::kNormChunkNodes) +
                                          " nodes (the fixed partials buffer); use fs_state_norms");
  double* part = nullptr;
  rc = scratch_on_current_device(&part);
  if (rc) return rc;
  enqueue(nn, d_R, d_dQ, part, d_out, (hipStream_t)stream);
  FSE_HIP(hipGetLastError());
  return FS_OK;
}

}  // extern "C"
