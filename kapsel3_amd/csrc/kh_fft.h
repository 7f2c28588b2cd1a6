// kh_fft.h -- host entry points of the 3D real FFT (see kh_fft.hip).
#pragma once

#include "kh_common.h"

namespace kh {

bool fft_length_supported(int L);

// Single passes (each one kernel launch).
template <typename T>
// kzvalid >= 0: only kz < kzvalid is written (the wavefront engine; other
// lengths write every kz)
hipError_t fft_z_r2c(const Grid &g, const T *in, cpx<T> *out, hipStream_t s, int kzvalid = -1);
// kzvalid >= 0: modes kz >= kzvalid are taken as zero (not read)
template <typename T>
hipError_t fft_z_c2r(const Grid &g, const cpx<T> *in, T *out, hipStream_t s, int kzvalid = -1);
// up to five components in one launch
template <typename T>
struct ZIO {
    const cpx<T> *in[5];
    T            *out[5];
};
template <typename T>
// mask: see zw_c2r (ignored by the LDS engine, which writes every point)
hipError_t fft_z_c2r_multi(const Grid &g, const ZIO<T> &io, int ncomp, hipStream_t s, int kzvalid = -1,
                           const T *mask = nullptr, Occ occ = {});
template <typename T, bool FWD>
hipError_t fft_y(const Grid &g, cpx<T> *data, hipStream_t s);
template <typename T, bool FWD>
hipError_t fft_x(const Grid &g, cpx<T> *data, hipStream_t s);

// Forward: real R[x][y][z] -> spectral S[x][y][kz] (unnormalised, exp(-i k.x)).
template <typename T>
hipError_t fft3d_r2c(const Grid &g, const T *real_in, cpx<T> *spec_out, hipStream_t s);

// Inverse: S -> R, normalised by 1/(NX NY NZ). S is overwritten (in-place
// x and y passes).
template <typename T>
hipError_t fft3d_c2r(const Grid &g, cpx<T> *spec_inout, T *real_out, hipStream_t s);

}  // namespace kh
