// kh_fused.h -- host launchers of the dealiasing-pruned fused passes
// (kh_fused.hip).
#pragma once

#include "kh_common.h"

namespace kh {

// up to five spectral components transformed by one launch (grid row = component)
template <typename T>
struct Comps {
    cpx<T> *p[5];
};

// y passes on columns kz < TRN_Z; x passes on kept (ky,kz) columns (a
// non-null skip: nothing is done once *skip is nonzero)
template <typename T, bool FWD>
hipError_t cols_y(const Grid &g, const Comps<T> &data, int ncomp, int in_mask, int out_mask, hipStream_t s,
                  const int *skip = nullptr);
template <typename T, bool FWD>
hipError_t cols_x(const Grid &g, const Comps<T> &data, int ncomp, int in_mask, int out_mask, hipStream_t s);
// zeta -> u_hat (truncated) -> inverse x pass, 3 components
template <typename T>
hipError_t x_zeta2u_inv(const Grid &g, const cpx<T> *z0, const cpx<T> *z1, cpx<T> *const u[3], double *ukdc,
                        int fixmask, hipStream_t s);
// inverse z (3) -> products -> forward z (5); for power-of-two NZ, p5[2] may
// be s3[0], p5[1] s3[2] and p5[0] s3[1] (and no other aliasing)
template <typename T>
hipError_t z_adv(const Grid &g, cpx<T> *const s3[3], cpx<T> *const p5[5], hipStream_t s);
// f = up - min(phi_sum,1) u* -> forward z (3); with `clear` (possible when
// z_force_clears(g)) up and phi_sum are left zero for the next step, and a
// non-null `tick` (same condition) is incremented once
// z_force_clears(g) also gates us2 (u* = us + us2), a null us (f = up) and
// skip (see zw_force); clear: bit 0 up, bit 1 phi_sum
bool z_force_clears(const Grid &g);
template <typename T>
hipError_t z_force(const Grid &g, T *const up[3], T *phis, T *const us[3], cpx<T> *const f[3], int clear,
                   int *tick, hipStream_t s, T *const us2[3] = nullptr, const int *skip = nullptr, Occ occ = {});
// the slip iteration's projection on the pruned transform: forward x (3) of
// the slip force field -> truncation + Solenoidal_uk, its curl into (a0, a1)
// at the kept modes and its k = 0 mode into dc -> inverse x (3), in place
template <typename T>
hipError_t x_slip(const Grid &g, cpx<T> *const f[3], cpx<T> *a0, cpx<T> *a1, double *dc, const int *skip,
                  hipStream_t s);
// forward x (5 products) -> Omega + slaved-Euler update of zeta at the kept modes
template <typename T>
hipError_t x_adv_ns(const Grid &g, const Comps<T> &prod, cpx<T> *z0, cpx<T> *z1, double dmy0, double dmy1,
                    hipStream_t s);
// forward x (3) -> projection + u* + curl -> zeta, uk_dc
template <typename T>
hipError_t x_f_zeta(const Grid &g, cpx<T> *const f[3], cpx<T> *z0, cpx<T> *z1, double *ukdc, hipStream_t s);
// zeta += (a0, a1), uk_dc += dc; kept_only: (a0, a1) read at the kept modes only
template <typename T>
hipError_t zeta_add(const Grid &g, cpx<T> *z0, cpx<T> *z1, const cpx<T> *a0, const cpx<T> *a1, double *ukdc,
                    const double *dc, hipStream_t s, int kept_only = 0);
template <typename T>
hipError_t zeta_clean(const Grid &g, cpx<T> *z0, cpx<T> *z1, hipStream_t s);

}  // namespace kh
