// common.h -- host-side helpers shared by saship.hip and batch.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/saship.h"
#include "models.h"

namespace saship_host {

// thread-local error message and current device (defined in saship.hip)
std::string &last_error();
int &current_device();

inline int fail(int code, const std::string &msg)
{
    last_error() = msg;
    return code;
}

#define HIP_TRY(expr)                                                                         \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess)                                                                 \
            return saship_host::fail(SASHIP_EHIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

// Gauss-Legendre nodes and weights on [-1, 1], ascending, by Newton's method
// on the three-term Legendre recurrence (long double).
inline void gauss_legendre(int n, std::vector<double> &z, std::vector<double> &w)
{
    z.assign(n, 0.0);
    w.assign(n, 0.0);
    for (int i = 0; i < (n + 1) / 2; i++) {
        long double x = cosl(3.14159265358979323846264338327950288L * (i + 0.75L) / (n + 0.5L));
        for (int it = 0; it < 100; it++) {
            long double p0 = 1.0L, p1 = x;
            for (int k = 2; k <= n; k++) {
                const long double p2 = ((2.0L * k - 1.0L) * x * p1 - (k - 1.0L) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            const long double dp = n * (x * p1 - p0) / (x * x - 1.0L);
            const long double dx = p1 / dp;
            x -= dx;
            if (fabsl(dx) < 1e-19L) break;
        }
        long double p0 = 1.0L, p1 = x;
        for (int k = 2; k <= n; k++) {
            const long double p2 = ((2.0L * k - 1.0L) * x * p1 - (k - 1.0L) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        const long double dp = n * (x * p1 - p0) / (x * x - 1.0L);
        const long double wt = 2.0L / ((1.0L - x * x) * dp * dp);
        z[n - 1 - i] = (double)x;
        z[i] = (double)-x;
        w[n - 1 - i] = w[i] = (double)wt;
    }
    if (n % 2) z[n / 2] = 0.0;
}

inline int cu_count(int dev)
{
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
    return n;
}

// Model ids 0-4, 6-13 and 17-18 are form factors; 5, 14, 15 and 16 the
// hard-sphere, sticky hard-sphere, square-well and Hayter-Penfold MSA
// structure factors (radius_effective, volfraction and up to four more
// parameters, nothing polydisperse, no sld: no magnetic values).  17
// (core_multi_shell) and 18 (onion) have vector parameters, ten slots each;
// 19-21 are core_shell_ellipsoid, elliptical_cylinder, core_shell_bicelle;
// 22-32 models defined by their intensity alone (no <F>, FP64): guinier,
// gaussian_peak, lorentz, dab, gel_fit, mass_fractal, surface_fractal,
// mass_surface_fractal, star_polymer, mono_gauss_coil, fractal, and 33-40
// fractal_core_shell, lamellar_hg, lamellar_stack_caille,
// lamellar_hg_stack_caille, lamellar_stack_paracrystal, linear_pearls,
// polymer_micelle, raspberry.
constexpr int MODEL_HARDSPHERE = 5;
constexpr int MODEL_LAST = 40;
inline int model_npars(int id)
{
    static const int np[41] = {3, 6, 6, 8, 5, 2, 8, 8, 8, 8, 8, 7, 5, 4, 4, 4, 6, 24, 44, 9, 8, 10,
                               1, 2, 1, 1, 5, 3, 3, 4, 2, 2, 6, 8, 5, 6, 8, 6, 5, 10, 9};
    return np[id];
}
inline int model_maxpd(int id)
{
    static const int mp[41] = {1, 4, 4, 5, 2, 0, 5, 5, 4, 4, 4, 5, 2, 2, 0, 0, 0, 5, 5, 5, 5, 5,
                               0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 2, 2, 2, 3, 1, 0, 0, 3};
    return mp[id];
}
inline int model_nvalues(int id)
{
    static const int nv[41] = {15, 18, 18, 23, 20, 4, 20, 20, 20, 20, 20, 19, 17, 16, 6, 6, 8, 66, 116, 24, 20, 28,
                               3, 4, 3, 3, 7, 5, 5, 6, 4, 4, 18, 23, 20, 18, 23, 18, 17, 25, 24};
    return nv[id];
}
inline bool model_oriented(int id) { return id == 1 || id == 2 || id == 3 || (id >= 6 && id <= 11) || (id >= 19 && id <= 21); }
// orientation (theta, phi, psi): view and jitter about three axes
inline bool model_abc(int id) { return (id >= 6 && id <= 10) || id == 20; }
// bcc, fcc, sc paracrystals: FP64 product rule (Gauss-150 by default)
inline bool model_paracrystal(int id) { return id >= 8 && id <= 10; }
inline bool model_nested(int id) { return id == 6 || id == 7 || id == 20; }
inline bool model_structure(int id) { return id == MODEL_HARDSPHERE || id == 14 || id == 15 || id == 16; }
// models defined by I(q) alone (FP64 direct evaluation)
inline bool model_direct(int id) { return id >= 22 && id <= 40; }
// the 1D kernel returns <F> beside <F^2> (two outputs per q)
inline bool model_has_fq(int id) { return !model_structure(id) && !model_paracrystal(id) && !model_direct(id); }
// models with sld parameters have a magnetic kernel
inline bool model_magnetic(int id) { return model_nvalues(id) > 2 + model_npars(id); }
// floats of one set's spin constants (sas::spin_floats): 24 + 8 per sld slot
inline int model_spin_floats(int id)
{
    const int nmag = (model_nvalues(id) - 2 - model_npars(id) - 4) / 3;
    return 24 + 8 * (nmag > 0 ? nmag : 1);
}
// bytes of a kernel handle's meta block: npts at 0, the four sums at 64, spin constants at 128
constexpr int META_SPIN = 128, META_BYTES = 128 + 4 * (24 + 8 * 24);

// Separable cylinder 1D meshes: only the radius (parameter 2) and length
// (parameter 3) levels have more than one point, each at most SEP_MAX.
// Returns the length count rounded up to a multiple of 4 (the column
// count of the bilinear tables), or 0 when the mesh is not separable.
constexpr int SEP_MAX = 64;
inline int cylinder_separable(const int32_t *details)
{
    const int mp = model_maxpd(1);
    int nr = 0, nl = 0;
    for (int k = 0; k < mp; k++) {
        const int par = details[k], len = details[mp + k];
        if (par == 2) nr = len;
        else if (par == 3) nl = len;
        else if (len != 1) return 0;
    }
    if (nr < 1 || nr > SEP_MAX || nl < 1 || nl > SEP_MAX) return 0;
    return (nl + 3) / 4 * 4;
}

// "set s: " for messages about set s of a batch ("" for s < 0)
inline std::string set_prefix(int set) { return set < 0 ? std::string() : "set " + std::to_string(set) + ": "; }

// A separable cylinder set whose mesh the adaptive rule can split by
// rectangles: every point valid (radius, length >= 0) and accepted at this
// cutoff (weights >= 1e-150, so products stay positive; for cutoff > 0 the
// product of the smallest weights clears it by a margin), both grids
// ascending and equally spaced (the rotation mode of the expansion, which
// orders its tables by grid).  Out: the smallest and largest radius and
// half length.  Call after validate_set.
inline bool adaptive_separable(const int32_t *details, const double *values, double cutoff, double ext[4])
{
    const int mp = model_maxpd(1), nv = model_nvalues(1);
    const int32_t *par = details, *len = details + mp, *off = details + 2 * mp;
    const int num_weights = details[4 * mp + 1];
    const double *pv = values + nv, *pw = pv + num_weights;
    constexpr double WMIN = 1e-150;
    double wprod = 1.0;
    int kr = -1, kl = -1;
    for (int k = 0; k < mp; k++) {
        if (par[k] == 2) kr = k;
        else if (par[k] == 3) kl = k;
        if (par[k] == 2 || par[k] == 3) {
            double wmin = HUGE_VAL;
            for (int i = 0; i < len[k]; i++) wmin = std::min(wmin, pw[off[k] + i]);
            if (!(wmin >= WMIN)) return false;
            wprod *= wmin;
        } else {
            if (!(pw[off[k]] >= WMIN)) return false;
            wprod *= pw[off[k]];
        }
    }
    if (kr < 0 || kl < 0) return false;
    // every mesh weight above the cutoff (and above 0, which a cutoff of 0
    // also asks of the weight products), with a margin for their rounding
    if (!(wprod > std::max(cutoff, 0.0) * (1.0 + 1e-9)) || !(wprod >= 1e-290)) return false;
    for (int g = 0; g < 2; g++) {
        const int k = g ? kl : kr, n = len[k];
        const double *v = pv + off[k];
        const double d = n > 1 ? (v[n - 1] - v[0]) / (n - 1) : 0.0;
        for (int i = 0; i < n; i++) {
            if (!(v[i] >= 0.0)) return false;
            if (i > 0 && !(v[i] > v[i - 1])) return false;
            if (!(std::fabs(v[i] - (v[0] + i * d)) <= 1e-9 * std::fabs(v[n - 1]))) return false;
        }
        ext[2 * g] = g ? 0.5 * v[0] : v[0];
        ext[2 * g + 1] = g ? 0.5 * v[n - 1] : v[n - 1];
    }
    return true;
}

// The adaptive rule's size index for q * r (adaptive.c:11290-11297, FP64).
inline int adaptive_rule(double qr) { return qr < 10.0 ? 0 : qr < 100.0 ? 1 : qr < 750.0 ? 2 : 3; }
inline int adaptive_nodes(int k) { return k == 0 ? 20 : k == 1 ? 76 : k == 2 ? 500 : 5000; }

// Checks one set's details/values against the model before any device work:
// every index the kernels derive from them stays inside the buffers (set:
// the set's index in a batch, -1 for a single call; only named in messages).
inline int validate_set(int id, int kind, const int32_t *details, int64_t n_details,
                        int64_t n_values, int pd_start, int pd_stop, int set)
{
    const int mp = model_maxpd(id);
    if (n_details != 4 * mp + 4)
        return fail(SASHIP_EINVAL, set_prefix(set) + "details must have 4*max_pd+4 = " + std::to_string(4 * mp + 4) + " entries");
    const int32_t num_eval = details[4 * mp], num_weights = details[4 * mp + 1];
    if (num_eval < 0) return fail(SASHIP_EINVAL, set_prefix(set) + "num_eval < 0");
    if (pd_start < 0 || (pd_stop >= 0 && pd_stop < pd_start))
        return fail(SASHIP_EINVAL, set_prefix(set) + "need 0 <= pd_start <= pd_stop");
    if (pd_stop > num_eval) return fail(SASHIP_EINVAL, set_prefix(set) + "pd_stop exceeds num_eval");
    if (num_weights < 0 || n_values < model_nvalues(id) + 2 * (int64_t)num_weights)
        return fail(SASHIP_EINVAL, set_prefix(set) + "values shorter than nvalues + 2*num_weights");
    for (int i = 0; i < mp; i++) {
        const int par = details[i], len = details[mp + i], off = details[2 * mp + i], stride = details[3 * mp + i];
        if (par < 0 || par >= model_npars(id) || len < 1 || off < 0 || off + len > num_weights || stride < 1)
            return fail(SASHIP_EINVAL, set_prefix(set) + "inconsistent details buffer at pd level " + std::to_string(i));
    }
    if (kind == 2 && model_structure(id)) return fail(SASHIP_EINVAL, set_prefix(set) + "structure factors have no magnetic kernel");
    if (kind == 2 && !model_magnetic(id)) return fail(SASHIP_EINVAL, set_prefix(set) + "the model has no sld: no magnetic kernel");
    if (kind != 0 && model_oriented(id)) {
        const int tp = details[4 * mp + 3];
        const int last = tp + (model_abc(id) ? 2 : 1);
        if (tp < 0 || last >= model_npars(id)) return fail(SASHIP_EINVAL, set_prefix(set) + "theta_par out of range");
    }
    return 0;
}

// Grow a device buffer to at least `bytes` (contents are not kept).
template <class T>
int ensure(T *&ptr, size_t &cap, size_t bytes)
{
    if (bytes <= cap && ptr) return 0;
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
    const size_t want = std::max<size_t>(bytes, 256);
    HIP_TRY(hipMalloc((void **)&ptr, want));
    cap = want;
    return 0;
}

// Grow a pinned host buffer to at least `bytes` (contents are not kept).
inline int ensure_host(char *&ptr, size_t &cap, size_t bytes)
{
    if (bytes <= cap && ptr) return 0;
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    cap = 0;
    const size_t want = std::max<size_t>(bytes, 4096);
    HIP_TRY(hipHostMalloc((void **)&ptr, want, hipHostMallocDefault));
    cap = want;
    return 0;
}

}  // namespace saship_host

struct saship_model {
    int id = 0, quad_n = 0, dev = 0;
    float *tab = nullptr;     // 5 * ntab floats: a | b | w | c | d
    double *tab64 = nullptr;  // FP64 models: 5 * n doubles (Quad::g64)
    int ntab = 0;
    sas::Quad quad{};
    // kernels and batches made from the model hold a reference; the tables
    // are freed once the model is released and the last holder is gone
    std::atomic<int> refs{0};
    std::atomic<bool> released{false};
};

namespace saship_host {
inline void model_free(saship_model *m)
{
    (void)hipSetDevice(m->dev);
    if (m->tab) (void)hipFree(m->tab);
    if (m->tab64) (void)hipFree(m->tab64);
    delete m;
}
inline void model_retain(saship_model *m) { m->refs.fetch_add(1); }
inline void model_drop(saship_model *m)
{
    if (m->refs.fetch_sub(1) == 1 && m->released.load()) model_free(m);
}
// Host copy of a downloaded result into the caller's buffer: large copies
// (a batch of 2D images is tens of MB) are split over a small pool of
// worker threads that lives as long as the library, which also spreads the
// page faults of a freshly allocated destination.  One copy at a time goes
// through the pool; smaller ones are a plain memcpy.
class CopyPool {
public:
    static CopyPool &get()
    {
        static CopyPool pool;
        return pool;
    }
    void copy(void *dst, const void *src, size_t bytes)
    {
        std::lock_guard<std::mutex> one(submit_);
        {
            std::lock_guard<std::mutex> lk(mu_);
            dst_ = (char *)dst;
            src_ = (const char *)src;
            bytes_ = bytes;
            part_ = ((bytes / (workers_.size() + 1)) + 4095) & ~(size_t)4095;
            pending_ = (int)workers_.size();
            ++generation_;
        }
        cv_.notify_all();
        memcpy(dst, src, std::min(part_, bytes));        // the caller's share: piece 0
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [&] { return pending_ == 0; });
    }
    ~CopyPool()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : workers_) t.join();
    }

private:
    CopyPool()
    {
        const unsigned hw = std::thread::hardware_concurrency();
        const unsigned n = std::min(7u, hw > 1 ? hw - 1 : 1u);
        for (unsigned i = 0; i < n; i++) workers_.emplace_back([this, i] { run(i + 1); });
    }
    void run(unsigned piece)
    {
        uint64_t seen = 0;
        for (;;) {
            char *dst;
            const char *src;
            size_t bytes, part;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || generation_ != seen; });
                if (stop_) return;
                seen = generation_;
                dst = dst_;
                src = src_;
                bytes = bytes_;
                part = part_;
            }
            const size_t off = part * piece;
            if (off < bytes) memcpy(dst + off, src + off, std::min(part, bytes - off));
            {
                std::lock_guard<std::mutex> lk(mu_);
                if (--pending_ == 0) done_.notify_one();
            }
        }
    }
    std::vector<std::thread> workers_;
    std::mutex mu_, submit_;
    std::condition_variable cv_, done_;
    char *dst_ = nullptr;
    const char *src_ = nullptr;
    size_t bytes_ = 0, part_ = 0;
    int pending_ = 0;
    uint64_t generation_ = 0;
    bool stop_ = false;
};

inline void copy_out(void *dst, const void *src, size_t bytes)
{
    constexpr size_t PARALLEL_FROM = (size_t)2 << 20;
    if (bytes < PARALLEL_FROM) memcpy(dst, src, bytes);
    else CopyPool::get().copy(dst, src, bytes);
}

// Wait for the work enqueued on a stream (or up to an event) by polling it.
// The runtime's blocking waits return 10-20 us after the work has ended,
// which is most of a small call's overhead; a call waits for its own short
// device sequence, so it polls for up to SPIN_US and only then blocks.
constexpr double SPIN_US = 2000.0;

template <class Query, class Block>
hipError_t spin_wait(Query query, Block block)
{
    const auto t0 = std::chrono::steady_clock::now();
    for (;;) {
        const hipError_t e = query();
        if (e != hipErrorNotReady) return e;
        if (std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count() > SPIN_US)
            return block();
    }
}

// The device's address of p when p is device memory or host memory the
// device can address (pinned), else NULL.
inline void *device_alias(const void *p)
{
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return attr.devicePointer;
}

inline bool device_accessible(const void *p) { return device_alias(p) != nullptr; }

inline hipError_t stream_wait(hipStream_t st)
{
    return spin_wait([st] { return hipStreamQuery(st); }, [st] { return hipStreamSynchronize(st); });
}

inline hipError_t event_wait(hipEvent_t ev)
{
    return spin_wait([ev] { return hipEventQuery(ev); }, [ev] { return hipEventSynchronize(ev); });
}

}  // namespace saship_host
