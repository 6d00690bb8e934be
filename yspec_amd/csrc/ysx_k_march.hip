// ysx_k_march.hip -- one march kernel and its launch function per object
// file: the build compiles this file twelve times, once per kernel
// (YSX_MARCH = 0 toroidal, 1 minors from the surface, 2 minors from the
// start level, 3 b-vector columns, 4 b-vector rows, 5 determinant rows) and
// work-distribution mode (YSX_PERSIST = 0 one
// block per work item, 1 persistent blocks on atomic queues), so that the
// large kernels compile in parallel.
#include "ysx_march_kernels.hpp"

#if !defined(YSX_MARCH) || !defined(YSX_PERSIST)
#error "compile with -DYSX_MARCH=0..5 -DYSX_PERSIST=0|1"
#endif

#if YSX_PERSIST
#define YSX_LAUNCH(name) launch_##name##_pers
#else
#define YSX_LAUNCH(name) launch_##name##_item
#endif

namespace ysx {

#if YSX_MARCH == 0
void YSX_LAUNCH(tor_march)(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                           const RunK* runs, Queue Q, StartData SD, cd* rec, int* cnt,
                           const StepStore* fs) {
  if (fs && fs->ys)
    hipLaunchKernelGGL((k_tor_march<YSX_PERSIST, true>), g, b, lds, s, gtab, runs, Q, SD, rec,
                       cnt, *fs);
  else
    hipLaunchKernelGGL((k_tor_march<YSX_PERSIST, false>), g, b, lds, s, gtab, runs, Q, SD, rec,
                       cnt, StepStore{});
}
#elif YSX_MARCH == 1
void YSX_LAUNCH(minor_top)(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                           const RunK* runs, Queue Q, SphState S, int* cnt) {
  hipLaunchKernelGGL(k_minor_top<YSX_PERSIST>, g, b, lds, s, gtab, runs, Q, S, cnt);
}
#elif YSX_MARCH == 2
void YSX_LAUNCH(minor_march)(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                             const RunK* runs, Queue Q, StartData SD, SphState S, int* cnt,
                             int drow_on, const StepStore* fs) {
  if (drow_on)
    hipLaunchKernelGGL((k_minor_march<YSX_PERSIST, true>), g, b, lds, s, gtab, runs, Q, SD, S, cnt,
                       StepStore{});
  else if (fs && fs->ys)
    hipLaunchKernelGGL((k_minor_march<YSX_PERSIST, false, true>), g, b, lds, s, gtab, runs, Q, SD,
                       S, cnt, *fs);
  else
    hipLaunchKernelGGL((k_minor_march<YSX_PERSIST, false>), g, b, lds, s, gtab, runs, Q, SD, S,
                       cnt, StepStore{});
}
#elif YSX_MARCH == 3
void YSX_LAUNCH(bvec_march)(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                            const RunK* runs, Queue Q, SphState S, cd* rec, int* cnt) {
  hipLaunchKernelGGL(k_bvec_march<YSX_PERSIST>, g, b, lds, s, gtab, runs, Q, S, rec, cnt);
}
#elif YSX_MARCH == 4
void YSX_LAUNCH(brow_march)(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                            const RunK* runs, Queue Q, SphState S, ShareK SH, cd* rec, int* cnt) {
  hipLaunchKernelGGL(k_brow_march<YSX_PERSIST>, g, b, lds, s, gtab, runs, Q, S, SH, rec, cnt);
}
#elif YSX_MARCH == 5
void YSX_LAUNCH(drow_march)(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                            const RunK* runs, Queue Q, SphState S, ShareK SH, int* cnt) {
  hipLaunchKernelGGL(k_drow_march<YSX_PERSIST>, g, b, lds, s, gtab, runs, Q, S, SH, cnt);
}
#else
#error "YSX_MARCH must be 0..5"
#endif

}  // namespace ysx
