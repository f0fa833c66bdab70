// 3D r2c / c2r entry points (FFTW conventions, lib/parfft.h:33-138) built
// from the three LDS pencil passes of fft_core.h (power-of-two axes) or
// fft_any.h (axes with factors 2, 3, 5).
#include "fft.h"

namespace ldl {

template <typename T>
int r2c3d(const T* in, c2<T>* spec, const Mesh& m, T scale, const c2<T>* tw, hipStream_t st) {
    const int64_t nrows = m.n0 * m.n1, nh = m.nh();
    int rc;
    if ((rc = zpass_r2c<T>(nrows, m.n2, tw, RealRowLoad<T>{in, m.n2}, HalfRowStore<T>{spec, nh, T(1)}, st))) return rc;
    if ((rc = colpass<T, -1>(m.n1, m.n0, nh, m.n1 * nh, nh, tw, PlainColOp<T>{spec, spec, T(1)}, st))) return rc;
    if ((rc = colpass<T, -1>(m.n0, m.n1, nh, nh, m.n1 * nh, tw, PlainColOp<T>{spec, spec, scale}, st))) return rc;
    LDL_LAUNCH_CHECK();
    return LDL_OK;
}

template <typename T>
int c2r3d(const c2<T>* spec, c2<T>* tmp, T* out, const Mesh& m, T scale, const c2<T>* tw, hipStream_t st) {
    const int64_t nrows = m.n0 * m.n1, nh = m.nh();
    int rc;
    if ((rc = colpass<T, +1>(m.n0, m.n1, nh, nh, m.n1 * nh, tw, PlainColOp<T>{spec, tmp, T(1)}, st))) return rc;
    if ((rc = colpass<T, +1>(m.n1, m.n0, nh, m.n1 * nh, nh, tw, PlainColOp<T>{tmp, tmp, T(1)}, st))) return rc;
    if ((rc = zpass_c2r<T>(nrows, m.n2, tw, HalfRowLoad<T>{tmp, nh}, RealRowStore<T>{out, m.n2, scale}, st, nullptr))) return rc;
    LDL_LAUNCH_CHECK();
    return LDL_OK;
}

template <typename T>
static int pfft_impl(const T* in, T* out, const int64_t nmesh[3], double scale, hipStream_t st) {
    Mesh m;
    int rc = check_fft_mesh(nmesh, &m);
    if (rc) return rc;
    if (!in || !out) return fail(LDL_EINVAL, "NULL array argument");
    const c2<T>* tw;
    if ((rc = twiddles<T>(&tw))) return rc;
    return r2c3d<T>(in, reinterpret_cast<c2<T>*>(out), m, (T)scale, tw, st);
}

template <typename T>
static int pifft_impl(const T* in, T* out, const int64_t nmesh[3], double scale, void* ws, size_t wsb, hipStream_t st) {
    Mesh m;
    int rc = check_fft_mesh(nmesh, &m);
    if (rc) return rc;
    if (!in || !out) return fail(LDL_EINVAL, "NULL array argument");
    if (!ws || wsb < (size_t)m.hcells() * sizeof(c2<T>)) return fail(LDL_EINVAL, "pifft workspace too small (see ldl_fft_workspace_bytes)");
    const c2<T>* tw;
    if ((rc = twiddles<T>(&tw))) return rc;
    return c2r3d<T>(reinterpret_cast<const c2<T>*>(in), reinterpret_cast<c2<T>*>(ws), out, m, (T)scale, tw, st);
}

// out = w(k_z) scale in over an (N0, N1, N2/2+1) half spectrum (out may
// alias in); w by ``mode``: 0 -> 1; 1 -> c; 2 -> 1/c, with c = 1 on the
// k_z = 0, N2/2 planes and 2 elsewhere; 3 -> c2f's backward (model.py:23-29):
// v + conj(v) on z bins [1, min(N1/2, N2/2+1)), v elsewhere -- 2v for real
// bins, (2 Re v, 0) for complex ones.
// With ``E`` = T the bins are real numbers (a real kernel's half spectrum).
template <typename T, typename E>
__global__ void __launch_bounds__(256) k_zweight(const E* in, E* out, int64_t hcells, int64_t nh, int64_t n2, int64_t zend, int mode,
                                                 T scale) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < hcells; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t kz = i % nh;
        const bool edge = kz == 0 || 2 * kz == n2;
        T w = T(1);
        if (mode == 1) w = edge ? T(1) : T(2);
        else if (mode == 2) w = edge ? T(1) : T(0.5);
        else if (mode == 3) w = (kz >= 1 && kz < zend) ? T(2) : T(1);
        w *= scale;
        if constexpr (sizeof(E) == sizeof(T)) {
            out[i] = in[i] * w;
        } else {
            const E v = in[i];
            // c2f on complex bins: v + conj(v) = (2 Re v, 0) on the doubled bins
            const bool c2f = mode == 3 && w != scale;
            out[i] = c2f ? E{v.x * w, T(0)} : E{v.x * w, v.y * w};
        }
    }
}

template <typename T>
static int zweight_impl(const T* in, T* out, const int64_t nmesh[3], int32_t mode, double scale, hipStream_t st) {
    Mesh m;
    int rc = check_mesh(nmesh, &m);
    if (rc) return rc;
    if (!in || !out) return fail(LDL_EINVAL, "NULL array argument");
    const bool real = (mode & 4) != 0;
    mode &= ~4;
    if (mode < 0 || mode > 3) return fail(LDL_EINVAL, "spectrum weight mode must be 0..3, plus 4 for real bins");
    const int64_t nh = m.nh(), hc = m.hcells();
    const int64_t zend = std::min<int64_t>(m.n1 / 2, nh);
    const unsigned g = grid_for(hc, 256, 4096);
    if (real)
        hipLaunchKernelGGL((k_zweight<T, T>), dim3(g), dim3(256), 0, st, in, out, hc, nh, m.n2, zend, (int)mode, (T)scale);
    else
        hipLaunchKernelGGL((k_zweight<T, c2<T>>), dim3(g), dim3(256), 0, st, reinterpret_cast<const c2<T>*>(in),
                           reinterpret_cast<c2<T>*>(out), hc, nh, m.n2, zend, (int)mode, (T)scale);
    LDL_LAUNCH_CHECK();
    return LDL_OK;
}

template int r2c3d<float>(const float*, c2<float>*, const Mesh&, float, const c2<float>*, hipStream_t);
template int r2c3d<double>(const double*, c2<double>*, const Mesh&, double, const c2<double>*, hipStream_t);
template int c2r3d<float>(const c2<float>*, c2<float>*, float*, const Mesh&, float, const c2<float>*, hipStream_t);
template int c2r3d<double>(const c2<double>*, c2<double>*, double*, const Mesh&, double, const c2<double>*, hipStream_t);

}  // namespace ldl

using namespace ldl;

extern "C" {

size_t ldl_fft_workspace_bytes(const int64_t nmesh[3], int32_t is_f64) {
    Mesh m;
    if (check_mesh(nmesh, &m)) return 0;
    return (size_t)m.hcells() * (is_f64 ? 16 : 8);
}

int ldl_pfft_f32(const float* in, float* out, const int64_t nmesh[3], double scale, void* stream) {
    return pfft_impl<float>(in, out, nmesh, scale, as_stream(stream));
}
int ldl_pfft_f64(const double* in, double* out, const int64_t nmesh[3], double scale, void* stream) {
    return pfft_impl<double>(in, out, nmesh, scale, as_stream(stream));
}
int ldl_pifft_f32(const float* in, float* out, const int64_t nmesh[3], double scale, void* ws, size_t wsb, void* stream) {
    return pifft_impl<float>(in, out, nmesh, scale, ws, wsb, as_stream(stream));
}
int ldl_pifft_f64(const double* in, double* out, const int64_t nmesh[3], double scale, void* ws, size_t wsb, void* stream) {
    return pifft_impl<double>(in, out, nmesh, scale, ws, wsb, as_stream(stream));
}

int ldl_spectrum_zweight_f32(const float* in, float* out, const int64_t nmesh[3], int32_t mode, double scale, void* stream) {
    return zweight_impl<float>(in, out, nmesh, mode, scale, as_stream(stream));
}
int ldl_spectrum_zweight_f64(const double* in, double* out, const int64_t nmesh[3], int32_t mode, double scale, void* stream) {
    return zweight_impl<double>(in, out, nmesh, mode, scale, as_stream(stream));
}

}  // extern "C"
