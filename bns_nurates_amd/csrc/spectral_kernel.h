// spectral_kernel.h -- launch arguments of the spectral opacity kernel.
#ifndef BNS_NURATES_AMD_SPECTRAL_KERNEL_H_
#define BNS_NURATES_AMD_SPECTRAL_KERNEL_H_

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "gray_kernel.h"

namespace bns {

struct SpectralKernelArgs {
  int64_t npts;          // points in this launch
  int64_t ld;            // elements between consecutive fields of the point blocks
  const double* pts;     // [21][ld] point block (device)
  const double* distr;   // [28][ld] distribution block, or nullptr: reconstructed from M1
  int64_t nnu;           // energies per point
  const double* nu;      // energies [MeV]: nu[e] (nu_ld == 0) or nu[e * nu_ld + i]
  int64_t nu_ld;
  double* out;           // aos == 0: [16][nnu][ld], out[(f * nnu + e) * ld + i];
                         // aos == 1: SpectralOpacities[npts][nnu], out[(i * nnu + e) * 16 + f]
  int aos;
  const double* q1;      // 1D quadrature records (x, w) on [0, 1] (device)
  int n;                 // 1D quadrature order, 1..32
  int ech;               // energies per chunk (the LDS energy block)
  int64_t nchunk;        // chunks per point, ceil(nnu / ech)
  int64_t cpw;           // chunks per workgroup; ceil(nchunk / cpw) workgroups per point
  PointFlags flags;
  PointParams params;
  int stim;              // 1: ComputeSpectralOpacitiesStimulatedAbs, 0: ...NotStimulatedAbs
  double gamma36;        // GammaStirling(3.6)
  const Gp19Table* gp19; // GP19 bremsstrahlung table in device memory, or nullptr
};

// Energies per chunk for quadrature order n (a multiple of the energies one
// round of cells covers, floor(64 / 2n)).
int spectral_energies_per_chunk(int n);
// Chunks per workgroup: all chunks of a point in one workgroup (its point
// tables are built once) unless that leaves too few workgroups to fill the
// device.
int64_t spectral_chunks_per_group(int64_t npts, int64_t nchunk);
// Dynamic LDS bytes of one workgroup.
size_t spectral_lds_bytes(int n, int ech);

hipError_t launch_spectral_f64(const SpectralKernelArgs& a, hipStream_t stream);

}  // namespace bns

#endif  // BNS_NURATES_AMD_SPECTRAL_KERNEL_H_
