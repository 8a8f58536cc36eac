// bh_device.h -- device helpers shared by the kernel translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "bh_kernels.h"

namespace bh {

constexpr int kBlock = 256;

template <int D>
__device__ __forceinline__ double invert(const double (&J)[D][D], double (&Ji)[D][D]) {
  if constexpr (D == 2) {
    double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    double r = 1.0 / det;
    Ji[0][0] = J[1][1] * r;
    Ji[0][1] = -J[0][1] * r;
    Ji[1][0] = -J[1][0] * r;
    Ji[1][1] = J[0][0] * r;
    return det;
  } else {
    double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    double r = 1.0 / det;
    Ji[0][0] = c00 * r;
    Ji[1][0] = c01 * r;
    Ji[2][0] = c02 * r;
    Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * r;
    Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * r;
    Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * r;
    Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * r;
    Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * r;
    Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * r;
    return det;
  }
}

// Packed table layout (doubles), built by get_tables() in bh_api.hip:
//   [0, NQ)                         w_q
//   [NQ, NQ + NQ*NL*4)              per (q, l): ∇̂φ (D values) then φ, padded to 4
//   [.., + NQ*NG*D)                 geometric basis gradients ∂N_k/∂ξ
//   [.., + NQ*NG)                   geometric basis values N_k
template <int D, int NL, int NQ, int NG>
struct TabLayout {
  static constexpr int W = 0;
  static constexpr int R = NQ;
  static constexpr int DG = R + NQ * NL * 4;
  static constexpr int G = DG + NQ * NG * D;
  static constexpr int SIZE = G + NQ * NG;
};

// Row stride of the element-row workspace for a space of NL local dofs
// (ke_stride in bh_kernels.h).
template <int NL>
struct KeStride {
  static constexpr int value = ke_stride(NL);
};

template <int D>
struct SymN {
  static constexpr int N = D * (D + 1) / 2;
};

}  // namespace bh
