// Host-side 3D FFT drivers shared by fft.hip and ldl_sm.hip.
#pragma once

#include "fft_any.h"

namespace ldl {

// spec (N0,N1,N2/2+1) = scale * r2c(in); unnormalised for scale = 1.
template <typename T>
int r2c3d(const T* in, c2<T>* spec, const Mesh& m, T scale, const c2<T>* tw, hipStream_t st);

// out (N0,N1,N2) = scale * c2r(spec); tmp is one half-spectrum of scratch.
template <typename T>
int c2r3d(const c2<T>* spec, c2<T>* tmp, T* out, const Mesh& m, T scale, const c2<T>* tw, hipStream_t st);

}  // namespace ldl
