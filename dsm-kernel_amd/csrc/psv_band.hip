// psv_band.hip -- per-frequency P-SV band assembly on the device: the band
// pieces P0, P1, Q0, Q1 of every column (cala0 / cala1 / cala2 / calhml /
// calb0 / calb2 + overlapa / overlapb + cala's zone merge + calbc,
// calmat.f90:452-809, SGTpsv-MPI.f90:572-604), as psv_omega_band
// (fe_setup_psv.cpp) forms them on the host: the same expressions in the
// same order, built without FMA contraction, so the bits agree.  The
// element integrals are resident; a batch uploads only the per-zone
// moduli factors and (omega - i omegai)^2 of its frequencies.
//
// One lane per (frequency, band column).  A column takes the entries of
// at most two zones (the node two zones share), the first in zone order
// as P0/P1, the second as Q0/Q1.  A solid zone's column i (1-based in the
// zone) reads the element entries k of overlapa that land on it, each a
// short expression of the integrals of layer block n = ceil(k / 4).
#include "kernels.h"
#include "psv_band_lane.h"

namespace dsm {

__global__ __launch_bounds__(64) void psv_band_kernel(PsvBandArgs a)
{
    const int j = blockIdx.x * 64 + threadIdx.x;  // column, 0-based
    if (j < a.nn) psv_band_column(a, blockIdx.y, j);
}

void launch_psv_band(const PsvBandArgs &a, int nslots, hipStream_t s)
{
    hipLaunchKernelGGL(psv_band_kernel, dim3((a.nn + 63) / 64, nslots), dim3(64), 0, s, a);
}

}  // namespace dsm
