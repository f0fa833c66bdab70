// libldl runtime: error channel, version, twiddle-table cache, buildplan.
#include <mutex>
#include <string>
#include <vector>

#include "ldl_common.h"

namespace ldl {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

// One twiddle table per (device, precision), built once with FP64 cos/sin and
// kept until the process ends.  tw[L + m] = exp(-2 pi i m / L), L = 1..2048.
static constexpr int kTwMax = 2048;
static std::mutex g_tw_mutex;
static std::vector<void*> g_tw_f32, g_tw_f64;

template <typename T>
static int build_table(void** dst) {
    std::vector<c2<T>> h(2 * kTwMax);
    h[0] = c2<T>{T(1), T(0)};
    for (int L = 1; L <= kTwMax; L <<= 1)
        for (int m = 0; m < L; m++) {
            // angle 2 pi m / L; cos/sin of the reduced fraction keep FP64 accuracy
            double a = -2.0 * M_PI * (double)m / (double)L;
            h[L + m] = c2<T>{(T)std::cos(a), (T)std::sin(a)};
        }
    void* d = nullptr;
    LDL_HIP_TRY(hipMalloc(&d, h.size() * sizeof(c2<T>)));
    LDL_HIP_TRY(hipMemcpy(d, h.data(), h.size() * sizeof(c2<T>), hipMemcpyHostToDevice));
    *dst = d;
    return LDL_OK;
}

template <typename T>
int twiddles(const c2<T>** out) {
    int dev = 0;
    LDL_HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_tw_mutex);
    std::vector<void*>& tab = sizeof(T) == 4 ? g_tw_f32 : g_tw_f64;
    if ((int)tab.size() <= dev) tab.resize(dev + 1, nullptr);
    if (!tab[dev]) {
        int rc = build_table<T>(&tab[dev]);
        if (rc) return rc;
    }
    *out = reinterpret_cast<const c2<T>*>(tab[dev]);
    return LDL_OK;
}

template int twiddles<float>(const c2<float>**);
template int twiddles<double>(const c2<double>**);

}  // namespace ldl

extern "C" {

const char* ldl_last_error(void) { return ldl::g_last_error.c_str(); }

int ldl_version(void) { return 1; }

int ldl_buildplan(const int64_t nmesh[3], int32_t out[2]) {
    ldl::Mesh m;
    int rc = ldl::check_fft_mesh(nmesh, &m);
    if (rc) return rc;
    const ldl::c2<float>* t32;
    const ldl::c2<double>* t64;
    if ((rc = ldl::twiddles<float>(&t32))) return rc;
    if ((rc = ldl::twiddles<double>(&t64))) return rc;
    if (out) {
        out[0] = (int32_t)m.n0;
        out[1] = 0;
    }
    return LDL_OK;
}

}  // extern "C"
