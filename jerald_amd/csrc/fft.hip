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

}  // extern "C"
