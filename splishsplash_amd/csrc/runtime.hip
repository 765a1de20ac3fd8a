// runtime.hip -- library-wide state (last error, device check), kernel
// timing, kernel constants and the free-standing C-ABI entry points.
#include <cmath>
#include <mutex>

#include "common.hpp"
#include "kernel_math.hpp"

namespace sph {

static thread_local std::string t_last_error;
void set_last_error(const std::string &m) { t_last_error = m; }

void require_device() {
    static std::once_flag once;
    static int status = SPH_OK;
    static std::string msg;
    std::call_once(once, [] {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
            status = SPH_ERR_DEVICE;
            msg = "no HIP device visible";
            return;
        }
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, 0) != hipSuccess) {
            status = SPH_ERR_DEVICE;
            msg = "hipGetDeviceProperties failed";
            return;
        }
        if (std::string(prop.gcnArchName).find("gfx950") == std::string::npos) {
            status = SPH_ERR_DEVICE;
            msg = std::string("device is ") + prop.gcnArchName + ", library built for gfx950";
        }
    });
    if (status != SPH_OK) throw Error(status, msg);
}

KernelConst make_kernel_const(double h) {
    KernelConst K{};
    const double pi = 3.14159265358979323846;
    const double h3 = h * h * h;
    K.h = (float)h;
    K.h2 = (float)h * (float)h;
    const double stepSize = h / (double)(KRES - 1);
    K.invStep_d = 1.0 / stepSize;
    K.invStep_f = (float)K.invStep_d;
    K.invStep2_d = K.invStep_d * K.invStep_d;
    K.k = (float)(8.0 / (pi * h3));
    K.l = (float)(48.0 / (pi * h3));
    K.W0 = K.k;
    K.invH2 = (float)(1.0 / (h * h));
    K.lh2 = (float)(48.0 / (pi * h3) / (h * h));
    K.stepq = (float)(1.0 / (double)(KRES - 1));
    K.h_d = h;
    return K;
}

// ------------------------------------------------------------ KernelTimer
KernelTimer::~KernelTimer() {
    for (auto &p : pending_) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto e : pool_) (void)hipEventDestroy(e);
}
int KernelTimer::stat_index(const char *name) {
    for (size_t i = 0; i < stats.size(); i++)
        if (stats[i].name == name) return (int)i;
    stats.push_back(Stat{name, 0, 0.0});
    return (int)stats.size() - 1;
}
hipEvent_t KernelTimer::get_event() {
    if (!pool_.empty()) {
        hipEvent_t e = pool_.back();
        pool_.pop_back();
        return e;
    }
    hipEvent_t e;
    SPH_HIP(hipEventCreate(&e));
    return e;
}
int KernelTimer::begin(hipStream_t s, const char *name) {
    Pending p;
    p.a = get_event();
    p.b = get_event();
    p.stat = stat_index(name);
    SPH_HIP(hipEventRecord(p.a, s));
    pending_.push_back(p);
    return (int)pending_.size() - 1;
}
void KernelTimer::end(hipStream_t s, int slot) {
    if (slot < 0 || slot >= (int)pending_.size()) return;
    (void)hipEventRecord(pending_[slot].b, s);
}
void KernelTimer::collect() {
    for (auto &p : pending_) {
        (void)hipEventSynchronize(p.b);
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            stats[p.stat].ms += ms;
            stats[p.stat].launches++;
        }
        pool_.push_back(p.a);
        pool_.push_back(p.b);
    }
    pending_.clear();
}
void KernelTimer::reset() {
    collect();
    stats.clear();
}

// ---------------------------------------------------------- kernel check
__global__ void k_kernel_eval(KernelConst K, const float *r, unsigned n, float *w, float *g) {
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // xi = r, xj = 0 so that xi - xj = r exactly
    const float4 xi = make_float4(r[3 * i], r[3 * i + 1], r[3 * i + 2], 0.f);
    const float4 xj = make_float4(0.f, 0.f, 0.f, 0.f);
    float3 rr;
    float ww, gg;
    kernel_WG(xi, xj, K, rr, ww, gg);
    w[i] = ww;
    g[3 * i] = gg * rr.x;
    g[3 * i + 1] = gg * rr.y;
    g[3 * i + 2] = gg * rr.z;
}

}  // namespace sph

using namespace sph;

extern "C" {

const char *sph_last_error(void) { return t_last_error.c_str(); }

int sph_version(void) { return 100; }

int sph_device_count(int *out) {
    if (!out) return SPH_ERR_INVALID;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *out = n;
    return SPH_OK;
}

int sph_device_synchronize(void) {
    return guarded([] {
        require_device();
        SPH_HIP(hipDeviceSynchronize());
    });
}

int sph_kernel_eval(double support_radius, const float *r, size_t n, float *w, float *g) {
    return guarded([&] {
        SPH_REQUIRE(r && w && g && support_radius > 0, SPH_ERR_INVALID, "sph_kernel_eval: bad arguments");
        require_device();
        if (n == 0) return;
        const KernelConst K = make_kernel_const(support_radius);
        DBuf<float> dr, dw, dg;
        dr.reserve(3 * n);
        dw.reserve(n);
        dg.reserve(3 * n);
        SPH_HIP(hipMemcpy(dr.get(), r, 3 * n * sizeof(float), hipMemcpyHostToDevice));
        k_kernel_eval<<<nblocks(n), BLOCK>>>(K, dr.get(), (unsigned)n, dw.get(), dg.get());
        SPH_HIP(hipGetLastError());
        SPH_HIP(hipMemcpy(w, dw.get(), n * sizeof(float), hipMemcpyDeviceToHost));
        SPH_HIP(hipMemcpy(g, dg.get(), 3 * n * sizeof(float), hipMemcpyDeviceToHost));
    });
}

}  // extern "C"
