// kh_zwave.h -- z passes on the wavefront-scope FFT engine (kh_zwave.hip),
// used for power-of-two NZ from 8 to 512.
#pragma once

#include "kh_common.h"
#include "kh_fft.h"

namespace kh {

bool zw_supported(int L);
// rows of L reals -> half spectra (all L/2+1 modes); spectral rows `pitch` complex apart
template <typename T>
hipError_t zw_r2c(int L, const T *in, cpx<T> *out, long long nrows, int pitch, hipStream_t s, int kzv = -1);
// half spectra (modes kz < kzv, rest zero) -> rows of L reals * scale; grid y = component.
// mask (a real field, L >= 32): only 4-point vectors where it is nonzero are written
// (a non-null skip, masked passes only: nothing is done once *skip is nonzero;
// a non-null batch, masked passes only: its batch fields make one launch
// cover the batch's runs)
template <typename T>
hipError_t zw_c2r(int L, const ZIO<T> &io, int ncomp, long long nrows, T scale, int kzv, int pitch,
                  hipStream_t s, const T *mask = nullptr, const int *skip = nullptr, const Grid *batch = nullptr,
                  Occ occ = {});
// inverse z (3) -> advection products -> forward z (5), modes kz < TRN_Z
// inverse z pass of omega fused with u x omega (lengths >= 32)
bool zw_cross_supported(int L);
template <typename T>
hipError_t zw_c2r_cross(int L, const ZIO<T> &io, const T *const u[3], long long nrows, T scale, int kzv, int pitch,
                        hipStream_t s);
template <typename T>
hipError_t zw_adv(const Grid &g, cpx<T> *const s3[3], cpx<T> *const p5[5], T scale, hipStream_t s);
// f = up - min(phi_sum,1) u* -> forward z (3), modes kz < TRN_Z; u* = us +
// us2 when us2 is given, f = up when us is null (then phis may be null: up
// is read at every point instead of where phi_sum != 0); `clear` bit 0 leaves up
// zero, bit 1 phi_sum; a non-null `tick` is incremented once; nothing is
// done once a non-null *skip is nonzero
template <typename T>
struct ZForceIO {
    T       *up[3];
    T       *phis;
    const T *us[3];
    const T *us2[3];
    cpx<T>  *f[3];
    Occ      occ;  // with u* absent: where f may be nonzero (cleared with phi_sum)
};
template <typename T>
hipError_t zw_force(const Grid &g, T *const up[3], T *phis, T *const us[3], cpx<T> *const f[3], int clear,
                    int *tick, hipStream_t s, T *const us2[3] = nullptr, const int *skip = nullptr, Occ occ = {});

}  // namespace kh
