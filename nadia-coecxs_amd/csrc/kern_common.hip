// kern_common.hip -- single-workgroup final reductions (fixed order, so the
// results are deterministic for a given launch geometry).
#include "fft_device.h"
#include "kernels.h"

namespace nadia {

__global__ void __launch_bounds__(256) final_sum2_kernel(const double* p, int n, double* ratio, double* pair) {
  __shared__ double red[16];
  double a = 0, b = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    a += p[2 * i];
    b += p[2 * i + 1];
  }
  block_sum2(a, b, red);
  if (threadIdx.x == 0) {
    if (ratio) ratio[0] = a / b;
    if (pair) {
      pair[0] = a;
      pair[1] = b;
    }
  }
}

__global__ void __launch_bounds__(256) final_max_kernel(const double* p, int n, double* out) {
  __shared__ double red[8];
  double m = -INFINITY;
  for (int i = threadIdx.x; i < n; i += blockDim.x) m = p[i] > m ? p[i] : m;
  m = block_max(m, red);
  if (threadIdx.x == 0) out[0] = m;
}

__global__ void __launch_bounds__(256) final_sum1_kernel(const double* p, int n, double* out) {
  __shared__ double red[16];
  double a = 0, b = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) a += p[i];
  block_sum2(a, b, red);
  if (threadIdx.x == 0) out[0] = a;
}

// one workgroup per job: its row-pass partial sums -> errs[iter] (the same
// order as the reduction folded into the next column pass)
__global__ void __launch_bounds__(256) final_jobs_kernel(const JobDesc* jobs, int n, int iter) {
  __shared__ double red[16];
  const JobDesc* j = jobs + blockIdx.x;
  finalize_err256(j->partials, n, j->errs + iter, red);
}

// one workgroup per region: c_mean = sum |log mag| / sum |phase| over the
// region (ComplexConstraint::set_c_mean leaves a fixed value alone)
__global__ void __launch_bounds__(256) tc_final_kernel(TcDev d, const double* p, int nblocks) {
  __shared__ double red[16];
  const int k = blockIdx.x;
  double a = 0, b = 0;
  for (int i = threadIdx.x; i < nblocks; i += blockDim.x) {
    a += p[2 * ((long long)i * d.nregions + k)];
    b += p[2 * ((long long)i * d.nregions + k) + 1];
  }
  block_sum2(a, b, red);
  if (threadIdx.x == 0 && !d.fixed[k]) d.c_mean[k] = a / b;
}

void k_tc_final(const TcDev& d, const double* partials, int nblocks, hipStream_t st) {
  if (d.nregions > 0) hipLaunchKernelGGL(tc_final_kernel, dim3(d.nregions), dim3(256), 0, st, d, partials, nblocks);
}

void k_final_jobs(const JobDesc* jobs, int njobs, int n, int iter, hipStream_t st) {
  if (njobs > 0) hipLaunchKernelGGL(final_jobs_kernel, dim3(njobs), dim3(256), 0, st, jobs, n, iter);
}

// BaseCDI::update_n_best (NADIA/src/BaseCDI.c++:390-413) on the device: the
// error of the iteration just run (err) ranks against the kept errors
// (nbest_rank); the fused update pass ranks in its own error workgroup
// instead and only the copy is launched.
__global__ void nbest_rank_kernel(const double* err, double* best_err, int* perm, int n, int* target) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  nbest_rank(*err, best_err, perm, n, target);
}

// the snapshot copy of the estimate into buffer *target (when >= 0)
__global__ void nbest_copy_kernel(const unsigned long long* src, unsigned long long* const* bufs, const int* target,
                                  long long n8) {
  const int t = *target;
  if (t < 0) return;
  unsigned long long* dst = bufs[t];
  for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < n8; p += (long long)gridDim.x * blockDim.x)
    dst[p] = src[p];
}

void k_nbest(const double* err, double* best_err, int* perm, int n, int* target, const void* src, void* const* bufs,
             long long bytes, hipStream_t st) {
  hipLaunchKernelGGL(nbest_rank_kernel, dim3(1), dim3(64), 0, st, err, best_err, perm, n, target);
  k_nbest_copy(target, src, bufs, bytes, st);
}

void k_nbest_copy(const int* target, const void* src, void* const* bufs, long long bytes, hipStream_t st) {
  const long long n8 = bytes / 8;
  long long blocks = (n8 + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  hipLaunchKernelGGL(nbest_copy_kernel, dim3((unsigned)blocks), dim3(256), 0, st, (const unsigned long long*)src,
                     (unsigned long long* const*)bufs, (const int*)target, n8);
}

// best of the per-block XcBest partials: largest value, smallest index on ties
__global__ void xc_final_kernel(const XcBest* partials, int n, XcBest* out) {
  if (threadIdx.x != 0) return;
  XcBest b{0.0, -1};
  for (int k = 0; k < n; ++k) {
    const XcBest c = partials[k];
    if (c.index < 0) continue;
    if (b.index < 0 || c.value > b.value || (c.value == b.value && c.index < b.index)) b = c;
  }
  out[0] = b;
}

void k_xc_final(const XcBest* partials, int n, XcBest* out, hipStream_t st) {
  hipLaunchKernelGGL(xc_final_kernel, dim3(1), dim3(64), 0, st, partials, n, out);
}

__global__ void copy_doubles_kernel(CopyList l) {
  const int k = threadIdx.x;
  if (k < l.n) *l.dst[k] = *l.src[k];
}

void k_copy_doubles(const CopyList& l, hipStream_t st) {
  if (l.n > 0) hipLaunchKernelGGL(copy_doubles_kernel, dim3(1), dim3(kCopyMax), 0, st, l);
}

void k_final_sum2(const double* partials, int n, double* out_ratio, double* out_pair, hipStream_t st) {
  hipLaunchKernelGGL(final_sum2_kernel, dim3(1), dim3(256), 0, st, partials, n, out_ratio, out_pair);
}
void k_final_max(const double* partials, int n, double* out, hipStream_t st) {
  hipLaunchKernelGGL(final_max_kernel, dim3(1), dim3(256), 0, st, partials, n, out);
}
void k_final_sum1(const double* partials, int n, double* out, hipStream_t st) {
  hipLaunchKernelGGL(final_sum1_kernel, dim3(1), dim3(256), 0, st, partials, n, out);
}

}  // namespace nadia
