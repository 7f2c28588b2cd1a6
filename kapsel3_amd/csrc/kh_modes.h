// kh_modes.h -- per-mode spectral updates shared by the pointwise kernels
// (kh_spectral.hip) and the fused column passes (kh_fused.hip).
#pragma once

#include "kh_common.h"

namespace kh {

// One kept mode of Omega_k2zeta_k (fft_wrapper.cxx:227-263) followed by the
// slaved-Euler update of NS_solver_slavedEuler (fluid_solver.cxx:355-373):
// a[0..4] are the transformed products u2u3, u3u1, u1u2, u2^2-u3^2,
// u3^2-u1^2 of U2advection_k (operate_omega.cxx:31-47); zeta (x0, x1) is
// updated in place with dmy0 = -nu dt, dmy1 = 1/nu.
template <typename T>
__device__ __forceinline__ void slaved_euler_mode(const Grid &g, int ikx, int iky, int kz, const cpx<T> *a,
                                                  double dmy0, double dmy1, cpx<T> &x0, cpx<T> &x1) {
    const double k1 = ikx * g.wave_x, k2 = iky * g.wave_y, k3 = kz * g.wave_z;
    const double k1k2 = k1 * k2, k2k3 = k2 * k3, k3k1 = k3 * k1;
    const double K2   = (g.wave_x * ikx) * (g.wave_x * ikx) + (g.wave_y * iky) * (g.wave_y * iky) +
                      (g.wave_z * kz) * (g.wave_z * kz);
    const double IK2 = (K2 > 0.0) ? 1. / K2 : 0.0;
    double       om[3][2];
#pragma unroll
    for (int c = 0; c < 2; c++) {
        const double u2u3 = c ? a[0].y : a[0].x, u3u1 = c ? a[1].y : a[1].x, u1u2 = c ? a[2].y : a[2].x;
        const double u22_u32 = c ? a[3].y : a[3].x, u32_u12 = c ? a[4].y : a[4].x;
        const double u12_u22 = -u32_u12 - u22_u32;
        om[0][c] = -(k2k3 * u22_u32 + (k3 - k2) * (k3 + k2) * u2u3 + k3k1 * u1u2 - k1k2 * u3u1);
        om[1][c] = -(k3k1 * u32_u12 + (k1 - k3) * (k1 + k3) * u3u1 + k1k2 * u2u3 - k2k3 * u1u2);
        om[2][c] = -(k1k2 * u12_u22 + (k2 - k1) * (k2 + k1) * u1u2 + k2k3 * u3u1 - k3k1 * u2u3);
    }
    int c0, c1;
    if (ikx != 0) {
        c0 = 1;
        c1 = 2;
    } else if (iky != 0) {
        c0 = 2;
        c1 = 0;
    } else {
        c0 = 0;
        c1 = 1;
    }
    const double dmy = exp(dmy0 * K2) - 1.;
    const double nx0 = (double)x0.x + dmy * ((double)x0.x - dmy1 * IK2 * om[c0][0]);
    const double ny0 = (double)x0.y + dmy * ((double)x0.y - dmy1 * IK2 * om[c0][1]);
    const double nx1 = (double)x1.x + dmy * ((double)x1.x - dmy1 * IK2 * om[c1][0]);
    const double ny1 = (double)x1.y + dmy * ((double)x1.y - dmy1 * IK2 * om[c1][1]);
    x0               = cmk((T)nx0, (T)ny0);
    x1               = cmk((T)nx1, (T)ny1);
}

}  // namespace kh
