// libldl runtime: error channel, version, twiddle-table cache, buildplan.
#include <map>
#include <tuple>
#include <mutex>
#include <string>
#include <utility>
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

// Length-n tables tw[m] = exp(-2 pi i m / n), m < n, for transform lengths
// that are not powers of two; one per (device, precision, n).
static std::map<std::tuple<int, int, int>, void*> g_tw_n;

template <typename T>
int twiddles_n(int n, const c2<T>** out) {
    int dev = 0;
    LDL_HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_tw_mutex);
    const auto key = std::make_tuple(dev, (int)sizeof(T), n);
    auto it = g_tw_n.find(key);
    if (it == g_tw_n.end()) {
        std::vector<c2<T>> h(n);
        for (int m = 0; m < n; m++) {
            const double a = -2.0 * M_PI * (double)m / (double)n;
            h[m] = c2<T>{(T)std::cos(a), (T)std::sin(a)};
        }
        void* d = nullptr;
        LDL_HIP_TRY(hipMalloc(&d, h.size() * sizeof(c2<T>)));
        LDL_HIP_TRY(hipMemcpy(d, h.data(), h.size() * sizeof(c2<T>), hipMemcpyHostToDevice));
        it = g_tw_n.emplace(key, d).first;
    }
    *out = reinterpret_cast<const c2<T>*>(it->second);
    return LDL_OK;
}

template int twiddles_n<float>(int, const c2<float>**);
template int twiddles_n<double>(int, const c2<double>**);

// ------------------------------------------------- kernel timing (bench)
// When enabled, instrumented launches are bracketed by HIP events recorded on
// the launch stream; totals are read back (and the events released) later.
static std::mutex g_timer_mutex;
static bool g_timing = false;
static std::map<std::string, std::vector<std::pair<hipEvent_t, hipEvent_t>>> g_events;

bool timing_enabled() { return g_timing; }

void timer_begin(const char* name, hipStream_t st, void** token) {
    *token = nullptr;
    if (!g_timing) return;
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return;
    if (hipEventCreate(&b) != hipSuccess) { (void)hipEventDestroy(a); return; }
    (void)hipEventRecord(a, st);
    std::lock_guard<std::mutex> lock(g_timer_mutex);
    auto& v = g_events[name];
    v.emplace_back(a, b);
    *token = reinterpret_cast<void*>(v.size());
}

void timer_end(const char* name, hipStream_t st, void* token) {
    if (!token) return;
    std::lock_guard<std::mutex> lock(g_timer_mutex);
    auto& v = g_events[name];
    const size_t i = reinterpret_cast<size_t>(token) - 1;
    if (i < v.size()) (void)hipEventRecord(v[i].second, st);
}

static void timer_clear_locked() {
    for (auto& kv : g_events)
        for (auto& e : kv.second) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
    g_events.clear();
}

}  // namespace ldl

extern "C" {

const char* ldl_last_error(void) { return ldl::g_last_error.c_str(); }

int ldl_version(void) { return 1; }

int ldl_timer_enable(int32_t on) {
    std::lock_guard<std::mutex> lock(ldl::g_timer_mutex);
    ldl::timer_clear_locked();
    ldl::g_timing = on != 0;
    return LDL_OK;
}

int ldl_timer_read(const char* name, double* total_ms, int64_t* count) {
    if (!name || !total_ms || !count) return ldl::fail(LDL_EINVAL, "NULL argument");
    std::lock_guard<std::mutex> lock(ldl::g_timer_mutex);
    *total_ms = 0.0;
    *count = 0;
    auto it = ldl::g_events.find(name);
    if (it == ldl::g_events.end()) return LDL_OK;
    for (auto& e : it->second) {
        LDL_HIP_TRY(hipEventSynchronize(e.second));
        float ms = 0.f;
        LDL_HIP_TRY(hipEventElapsedTime(&ms, e.first, e.second));
        *total_ms += ms;
        *count += 1;
    }
    return LDL_OK;
}

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
