// gray_kernel.h -- launch arguments of the fused gray M1 kernels.
#ifndef BNS_NURATES_AMD_GRAY_KERNEL_H_
#define BNS_NURATES_AMD_GRAY_KERNEL_H_

#include <hip/hip_runtime.h>

#include <cstdint>

#include "gp19.h"
#include "gray_tables.h"

namespace bns {

struct PointFlags {  // OpacityFlags (bns_nurates.hpp:418-426)
  int use_abs_em, use_pair, use_brem, use_inelastic_scatt, use_iso;
};
struct PointParams {  // OpacityParams (bns_nurates.hpp:338-352)
  int use_dU, use_dm_eff, use_WM_ab, use_WM_sc, use_decay, brem_impl, use_NN_medium_corr,
      neglect_blocking;
};

struct GrayKernelArgs {
  int64_t npts;               // points in this launch
  int64_t ld;                 // elements between consecutive fields
  const double* pts;          // [21][ld] point block (device)
  const double* distr;        // [28][ld] distribution block or nullptr
  double* out;                // [20][ld] outputs (device)
  const double* tab;          // resident quadrature tables
  TableLayout L;
  PointFlags flags;
  PointParams params;
  int stim;                   // 1: ComputeM1Opacities, 0: ...NotStimulated
  double gamma36;             // GammaStirling(3.6) (functions.hpp:4460-4470)
  const Gp19Table* gp19;      // GP19 bremsstrahlung table in device memory, or nullptr
};

hipError_t launch_gray_m1_f64(const GrayKernelArgs& a, hipStream_t stream);
hipError_t launch_gray_m1_mixed(const GrayKernelArgs& a, hipStream_t stream);

}  // namespace bns

#endif  // BNS_NURATES_AMD_GRAY_KERNEL_H_
