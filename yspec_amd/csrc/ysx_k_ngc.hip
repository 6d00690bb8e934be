// ysx_k_ngc.hip -- the lockstep spheroidal march of gravity switches 0 and 1
// (both instantiations) and its launch function.
#include "ysx_march_ngc.hpp"

namespace ysx {

void launch_ngc_march(bool cow, dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                      const RunK* runs, Queue Q, cd* rec, int* cnt) {
  if (cow)
    hipLaunchKernelGGL(k_ngc_march<true>, g, b, lds, s, gtab, runs, Q, rec, cnt);
  else
    hipLaunchKernelGGL(k_ngc_march<false>, g, b, lds, s, gtab, runs, Q, rec, cnt);
}

}  // namespace ysx
