// saship.hip -- model and single-kernel entry points of libsaship.so.
//
// One evaluation (saship_kernel_call) runs on the handle's stream:
//   H2D (values, details) -> k_expand -> k_fq1d | k_iq2d -> k_finalize -> D2H
// k_finalize reduces the per-split partial sums of every output with one
// wave64 per output, in FP64, through a fixed shuffle tree, and writes the
// reference result layout (kernel_iq.c:392-423, :826-847), adding to it when
// pd_start > 0.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "common.h"
#include "kernels.h"
#include "structure.h"

using namespace saship_host;

namespace saship_host {
std::string &last_error()
{
    static thread_local std::string err;
    return err;
}
int &current_device()
{
    static thread_local int dev = 0;
    return dev;
}
}  // namespace saship_host

namespace sas {

__global__ __launch_bounds__(256) void k_finalize(
    int nq, int nout, int nsplit, const double *__restrict__ partial,
    const double *__restrict__ sums, int accumulate, double *__restrict__ result)
{
    const int wave = (blockIdx.x * 256 + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (wave < nout * nq) {
        const int o = wave / nq, iq = wave - o * nq;
        const double *src = partial + ((size_t)o * nq + iq) * nsplit;
        double s = 0.0;
        for (int k = lane; k < nsplit; k += 64) s += src[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
        if (lane == 0) {
            const size_t idx = (size_t)iq * nout + o;
            result[idx] = accumulate ? result[idx] + s : s;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x < 4) {
        const size_t idx = (size_t)nout * nq + threadIdx.x;
        const double v = sums[threadIdx.x];
        result[idx] = accumulate ? result[idx] + v : v;
    }
}

// The separable single call's sums: one thread per output value (result
// index iq*2 + o, so that consecutive threads write consecutive doubles of
// the mapped host result), adding the node blocks' partial[(blk*2 + o)*nq
// + iq] in block order; the four weight sums after them.  Adds to the
// result when accumulate is set.
__global__ __launch_bounds__(256) void k_finalize_sep(int nq, int nblk, const double *__restrict__ partial,
                                                      const double *__restrict__ sums, int accumulate,
                                                      double *__restrict__ result)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t nv = 2 * (int64_t)nq;
    if (i < nv) {
        const int64_t iq = i >> 1, o = i & 1;
        const double *p = partial + o * nq + iq;
        double s = 0.0;
        for (int b = 0; b < nblk; b++) s += p[(size_t)b * 2 * nq];
        result[i] = accumulate ? result[i] + s : s;
    } else if (i < nv + 4) {
        const double v = sums[i - nv];
        result[i] = accumulate ? result[i] + v : v;
    }
}

// Structure-factor evaluation (FP64): nothing is polydisperse, so the mesh
// is one point of weight 1 whatever pd_start/pd_stop are (kernel_iq.c with
// MAX_PD 0), counted when 1 > cutoff.  One thread per q, q = |(qx, qy)| for
// 2D input; the weight sums are (1, 1, 1, 0): volume 1, no effective
// radius.  Adds to the result when accumulate is set.
__global__ __launch_bounds__(256) void k_structure(int model, int npar, int nq, int is2d,
                                                   const double *__restrict__ q64,
                                                   const double *__restrict__ values, int include,
                                                   int accumulate, double *__restrict__ result)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < nq) {
        const double q = is2d ? sqrt(q64[2 * i] * q64[2 * i] + q64[2 * i + 1] * q64[2 * i + 1]) : q64[i];
        double p[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int k = 0; k < npar && k < 6; k++) p[k] = values[2 + k];
        const double s = include ? structure_sq(model, q, p) : 0.0;
        result[i] = accumulate ? result[i] + s : s;
    }
    if (blockIdx.x == 0 && threadIdx.x < 4) {
        const double v = include && threadIdx.x < 3 ? 1.0 : 0.0;
        result[nq + threadIdx.x] = accumulate ? result[nq + threadIdx.x] + v : v;
    }
}

}  // namespace sas

struct saship_kernel {
    saship_model *m = nullptr;
    int nq = 0, is2d = 0, dev = 0, cus = 256;
    hipStream_t st = nullptr;
    float *q32 = nullptr;
    double *q64 = nullptr;
    char *d_args = nullptr;
    size_t args_cap = 0;
    float4 *d_recs = nullptr;
    size_t recs_cap = 0;
    char *d_meta = nullptr;          // npts (int) at 0, sums (4 doubles) at 64
    double *d_partial = nullptr;
    size_t partial_cap = 0;
    char *d_sep = nullptr;          // separable cylinder tables (R, H, A, B, dims)
    // cylinder 1D under the adaptive rule: whole separable meshes are
    // evaluated by a one-set batch (its rule-rectangle separable path)
    std::vector<double> q_host;
    saship_batch *adapt = nullptr;
    char *h_stage = nullptr;
    size_t stage_cap = 0;
    double *h_result = nullptr;      // pinned, mapped: the final kernel writes the result here
    double *hd_result = nullptr;     // device address of h_result
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timing = false;
    bool separable = true;           // use the separable cylinder path when it applies
    double t_ms = 0.0;
    int ncalls = 0;
    // graph replay of repeated calls of one shape (upload, launches, download)
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<int64_t> gkey;
    std::mutex mu;
};

namespace sas {
// Small uploads (values and details of one call) as a kernel argument: a
// one-workgroup copy into device memory instead of a DMA transfer, which
// costs more latency than the few hundred bytes take to move.
constexpr int STAGE_ARG_BYTES = 2048;
struct StageIn {
    int nbytes;
    alignas(16) unsigned char data[STAGE_ARG_BYTES];
};
__global__ __launch_bounds__(128) void k_stage_in(StageIn s, unsigned char *__restrict__ dst)
{
    for (int i = threadIdx.x * 16; i < s.nbytes; i += 128 * 16) {
        const int n = min(16, s.nbytes - i);
        if (n == 16) *(uint4 *)(dst + i) = *(const uint4 *)(s.data + i);
        else for (int k = 0; k < n; k++) dst[i + k] = s.data[i + k];
    }
}

// Separable expansion of a single call whose values (at byte 0) and details
// (at byte doff) come as the kernel argument: copied into LDS, read there.
__global__ __launch_bounds__(EXPAND_THREADS) void k_sep_expand_in(StageIn s, int doff, ExpandArgs a, sep::Tables t)
{
    __shared__ __align__(16) unsigned char buf[STAGE_ARG_BYTES];
    for (int i = threadIdx.x * 16; i < s.nbytes; i += EXPAND_THREADS * 16) {
        const int n = min(16, s.nbytes - i);
        if (n == 16) *(uint4 *)(buf + i) = *(const uint4 *)(s.data + i);
        else for (int k = 0; k < n; k++) buf[i + k] = s.data[i + k];
    }
    __syncthreads();
    a.values = (const double *)buf;
    a.det = (const int32_t *)(buf + doff);
    a.voff = nullptr;
    sep::sep_expand<0>(a, t);
}
}  // namespace sas

namespace {

// Where one evaluation reads q and leaves its intermediate and final sums:
// the kernel handle's own buffers, or the no-allocation scratch.
struct Target {
    hipStream_t st;
    const saship_model *m;
    int nq, is2d, cus;
    const float *q32;
    const double *q64;
    float4 *recs;        // compacted mesh records
    char *meta;          // npts (int) at 0, sums (4 doubles) at 64, spin constants at META_SPIN
    double *partial;     // per-split partial sums
    double *result;      // nout*nq + 4 doubles, reference layout
    char *sep;           // separable tables, sas::sep::tables_bytes(1, SEP_MAX)
    hipEvent_t ev0, ev1; // NULL: untimed
};

// Calls f((M *)nullptr) with M the device functor of model id.
template <class F>
int with_model(int id, F &&f)
{
    switch (id) {
    case 0: return f((sas::Sphere *)nullptr);
    case 1: return f((sas::Cylinder *)nullptr);
    case 2: return f((sas::Ellipsoid *)nullptr);
    case 3: return f((sas::CoreShellCylinder *)nullptr);
    case 4: return f((sas::CoreShellSphere *)nullptr);
    case 6: return f((sas::Parallelepiped *)nullptr);
    case 7: return f((sas::TriaxialEllipsoid *)nullptr);
    case 8: return f((sas::BccParacrystal *)nullptr);
    case 9: return f((sas::FccParacrystal *)nullptr);
    case 10: return f((sas::ScParacrystal *)nullptr);
    case 11: return f((sas::HollowCylinder *)nullptr);
    case 12: return f((sas::Vesicle *)nullptr);
    case 13: return f((sas::FuzzySphere *)nullptr);
    case 17: return f((sas::CoreMultiShell *)nullptr);
    case 18: return f((sas::Onion *)nullptr);
    case 19: return f((sas::CoreShellEllipsoid *)nullptr);
    case 20: return f((sas::EllipticalCylinder *)nullptr);
    case 21: return f((sas::CoreShellBicelle *)nullptr);
    case 22: return f((sas::Guinier *)nullptr);
    case 23: return f((sas::GaussianPeak *)nullptr);
    case 24: return f((sas::Lorentz *)nullptr);
    case 25: return f((sas::Dab *)nullptr);
    case 26: return f((sas::GelFit *)nullptr);
    case 27: return f((sas::MassFractal *)nullptr);
    case 28: return f((sas::SurfaceFractal *)nullptr);
    case 29: return f((sas::MassSurfaceFractal *)nullptr);
    case 30: return f((sas::StarPolymer *)nullptr);
    case 31: return f((sas::MonoGaussCoil *)nullptr);
    case 32: return f((sas::Fractal *)nullptr);
    case 33: return f((sas::FractalCoreShell *)nullptr);
    case 34: return f((sas::LamellarHg *)nullptr);
    case 35: return f((sas::LamellarStackCaille *)nullptr);
    case 36: return f((sas::LamellarHgStackCaille *)nullptr);
    case 37: return f((sas::LamellarStackParacrystal *)nullptr);
    case 38: return f((sas::LinearPearls *)nullptr);
    case 39: return f((sas::PolymerMicelle *)nullptr);
    case 40: return f((sas::Raspberry *)nullptr);
    default: return fail(SASHIP_EINVAL, "model " + std::to_string(id) + " has no mesh kernel");
    }
}

// Record width (float4 units) of model id for kernel kind.
int record_width(int id, int kind)
{
    return with_model(id, [&](auto *tag) {
        using M = std::remove_pointer_t<decltype(tag)>;
        return kind == 0 ? sas::rv1<M>() : sas::rv2<M>();
    });
}

// Mesh splits of the generic path: about 8 workgroups on every CU.
int split_count(int cus, int nq, int npts_max)
{
    const int ntiles = std::max(1, (nq + sas::EVAL_THREADS - 1) / sas::EVAL_THREADS);
    const int target = std::max(1, (cus * 8) / ntiles);
    const int ppb = std::max(1, (npts_max + target - 1) / target);
    return std::max(1, std::min(65535, (npts_max + ppb - 1) / ppb));
}

// Launch the kernels of one evaluation for model M with nsplit mesh splits
// (t.recs holds pd_stop - pd_start records, t.partial nout*nq*nsplit doubles).
template <class M>
int launch(const Target &t, int kind, const int32_t *d_det, const double *d_val, int n_details, int pd_start,
           int pd_stop, double cutoff, int reff_mode, int accumulate, int nsplit)
{
    using namespace sas;
    const int nq = t.nq;
    const int nout = kind == 0 ? 2 : 1;
    int *npts = (int *)t.meta;
    double *sums = (double *)(t.meta + 64);

    static_assert(META_SPIN + 4 * sas::spin_floats<M>() <= META_BYTES, "spin constants exceed the meta block");
    float *spin = (float *)(t.meta + META_SPIN);
    ExpandArgs ea{d_det, d_val, nullptr, nullptr, n_details, pd_start, pd_stop, cutoff, reff_mode,
                  t.recs, npts, sums, nullptr, kind == 2 ? spin : nullptr};
    switch (kind) {
    case 0: k_expand<M, 0><<<1, EXPAND_THREADS, 0, t.st>>>(ea); break;
    case 1: k_expand<M, 1><<<1, EXPAND_THREADS, 0, t.st>>>(ea); break;
    default: k_expand<M, 2><<<1, EXPAND_THREADS, 0, t.st>>>(ea); break;
    }
    HIP_TRY(hipGetLastError());

    const int ntiles = std::max(1, (nq + EVAL_THREADS - 1) / EVAL_THREADS);
    const dim3 grid(ntiles, nsplit);
    if (t.ev0) HIP_TRY(hipEventRecord(t.ev0, t.st));
    if (kind == 0) {
        if (t.m->quad_n > 0)
            k_fq1d<M, true><<<grid, EVAL_THREADS, 0, t.st>>>(nq, t.q32, t.q64, t.recs, npts, nsplit, t.m->quad,
                                                             t.partial);
        else
            k_fq1d<M, false><<<grid, EVAL_THREADS, 0, t.st>>>(nq, t.q32, t.q64, t.recs, npts, nsplit, t.m->quad,
                                                              t.partial);
    } else if (kind == 1) {
        k_iq2d<M, false><<<grid, EVAL_THREADS, 0, t.st>>>(nq, t.q64, d_val, nullptr, t.recs, npts, nsplit,
                                                          t.partial);
    } else {
        k_iq2d<M, true><<<grid, EVAL_THREADS, 0, t.st>>>(nq, t.q64, d_val, spin, t.recs, npts, nsplit, t.partial);
    }
    HIP_TRY(hipGetLastError());
    if (t.ev1) HIP_TRY(hipEventRecord(t.ev1, t.st));

    // models without <F> return only the <F^2> sums (nout 1 in 1D too)
    const int nres = M::FQ ? nout : 1;
    const int fblocks = std::max(1, (nres * nq + 3) / 4);
    k_finalize<<<fblocks, 256, 0, t.st>>>(nq, nres, nsplit, t.partial, sums, accumulate, t.result);
    HIP_TRY(hipGetLastError());
    return 0;
}

int launch_model(const Target &t, int kind, const int32_t *d_det, const double *d_val, int n_details, int pd_start,
                 int pd_stop, double cutoff, int reff_mode, int accumulate, int nsplit)
{
    return with_model(t.m->id, [&](auto *tag) {
        using M = std::remove_pointer_t<decltype(tag)>;
        return launch<M>(t, kind, d_det, d_val, n_details, pd_start, pd_stop, cutoff, reff_mode, accumulate, nsplit);
    });
}

// Separable cylinder path: tables, one wave per (q tile, node), FP64 node sums
// (t.sep holds the tables, t.partial 2*nq*nodes doubles).
// With `in` the values and details come as a kernel argument (in->data, the
// details at byte doff) straight into the expansion, with no staging copy.
int launch_sep(const Target &t, int nlp, const int32_t *d_det, const double *d_val, int n_details, int pd_start,
               int pd_stop, double cutoff, int reff_mode, int accumulate, const sas::StageIn *in = nullptr,
               int doff = 0)
{
    using namespace sas;
    const int nq = t.nq, nnode = t.m->quad.n;
    const sep::Tables tb = sep::carve(t.sep, 1, nlp);
    double *sums = (double *)(t.meta + 64);
    ExpandArgs ea{d_det, d_val, nullptr, nullptr, n_details, pd_start, pd_stop, cutoff, reff_mode,
                  nullptr, (int *)t.meta, sums, nullptr};
    if (in) k_sep_expand_in<<<1, EXPAND_THREADS, 0, t.st>>>(*in, doff, ea, tb);
    else sep::k_sep_expand<0><<<1, EXPAND_THREADS, 0, t.st>>>(ea, tb);
    HIP_TRY(hipGetLastError());
    const dim3 grid(std::max(1, (nq + 63) / 64), (nnode + 3) / 4);
    if (t.ev0) HIP_TRY(hipEventRecord(t.ev0, t.st));
    sep::k_fq1d_sep<0><<<grid, 256, 0, t.st>>>(nq, t.q32, tb, t.m->quad, t.partial);
    HIP_TRY(hipGetLastError());
    if (t.ev1) HIP_TRY(hipEventRecord(t.ev1, t.st));
    const int fblocks = (2 * nq + 4 + 255) / 256;
    k_finalize_sep<<<fblocks, 256, 0, t.st>>>(nq, (int)grid.y, t.partial, sums, accumulate, t.result);
    HIP_TRY(hipGetLastError());
    return 0;
}

// Structure-factor evaluation (nothing polydisperse).
int launch_structure(const Target &t, const double *d_val, double cutoff, int accumulate)
{
    const int id = t.m->id;
    if (t.ev0) HIP_TRY(hipEventRecord(t.ev0, t.st));
    const int include = 1.0 > cutoff ? 1 : 0;
    sas::k_structure<<<std::max(1, (t.nq + 255) / 256), 256, 0, t.st>>>(id, model_npars(id), t.nq, t.is2d, t.q64,
                                                                       d_val, include, accumulate, t.result);
    HIP_TRY(hipGetLastError());
    if (t.ev1) HIP_TRY(hipEventRecord(t.ev1, t.st));
    return 0;
}

// Which path evaluates a call: 0 structure factor, 1 separable cylinder
// (*nlp set), 2 generic mesh kernels.
int choose_path(const saship_model *m, int kind, const int32_t *details, bool separable, int *nlp)
{
    *nlp = 0;
    if (model_structure(m->id)) return 0;
    if (m->id == 1 && kind == 0 && m->quad_n > 0 && separable) *nlp = cylinder_separable(details);
    return *nlp ? 1 : 2;
}

static_assert(SEP_MAX == sas::sep::NRMAX, "separable table size");

// FP32 FMA issue-rate probe: eight independent chains per lane, no memory
// traffic besides one store per thread (the ceiling the evaluation kernels
// are compared with; the compiler flags forbid packed FP32 instructions).
__global__ __launch_bounds__(256) void k_fma_probe(float *out, int iters, float a, float b)
{
    float x[8];
#pragma unroll
    for (int k = 0; k < 8; k++) x[k] = (float)(threadIdx.x + k);
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int r = 0; r < 8; r++)
#pragma unroll
            for (int k = 0; k < 8; k++) x[k] = fmaf(x[k], a, b);
    }
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; k++) s += x[k];
    out[(size_t)blockIdx.x * 256 + threadIdx.x] = s;
}

}  // namespace

// ------------------------------------------------------------------ no-allocation scratch
// Device memory of saship_kernel_call_noalloc: staged values and details, FP32
// q, mesh records, the weight sums and the per-split partial sums of one call,
// carved per call.  Each device holds its own copy from the library's load
// until the process ends.  Every buffer in it is written by the call before it
// is read (nothing has to start at zero); calls that overlap on one device
// share it and race.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)32 << 20;
static __device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];

namespace sas {
__global__ __launch_bounds__(256) void k_q_to_f32(int64_t n, const double *__restrict__ q64, float *__restrict__ q32)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) q32[i] = (float)q64[i];
}
}  // namespace sas

namespace {
std::mutex scratch_mu;
std::vector<char *> scratch_dev;   // device address of `scratch` per device (NULL: not looked up)

// Device address of `scratch` on device dev, which must be current.  The
// symbol's host-side name is not a device address: kernels and copies get
// the address looked up here.
int scratch_address(int dev, char **out)
{
    std::lock_guard<std::mutex> lock(scratch_mu);
    if ((int)scratch_dev.size() <= dev) scratch_dev.resize(dev + 1, nullptr);
    if (!scratch_dev[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(SASHIP_EHIP, std::string("scratch address: ") + hipGetErrorString(e));
        scratch_dev[dev] = (char *)ptr;
    }
    *out = scratch_dev[dev];
    return 0;
}

}  // namespace


namespace sas {
// Diagnostic: one structure factor's coefficients (thread 0) and S at every q.
__global__ __launch_bounds__(256) void k_structure_eval(int model, const double *__restrict__ pars, int nq,
                                                        const double *__restrict__ q, double *__restrict__ prep,
                                                        double *__restrict__ sq)
{
    __shared__ double pr[S_PREP];
    if (threadIdx.x == 0) {
        for (int k = 0; k < S_PREP; k++) pr[k] = 0.0;
        structure_prepare(model, pars, pr);
        for (int k = 0; k < S_PREP; k++) prep[k] = pr[k];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nq; i += 256) sq[i] = structure_sq_prepared(model, q[i], pars, pr);
}
}  // namespace sas

namespace {
void drop_graph(saship_kernel *k)
{
    if (k->gexec) (void)hipGraphExecDestroy(k->gexec);
    if (k->graph) (void)hipGraphDestroy(k->graph);
    k->gexec = nullptr;
    k->graph = nullptr;
    k->gkey.clear();
}

// Wait for the call enqueued on the handle's stream (its download included),
// hand the result over and account its time.
int finish_call(saship_kernel *k, size_t rbytes, double *result)
{
    HIP_TRY(stream_wait(k->st));
    memcpy(result, k->h_result, rbytes);
    if (k->timing) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, k->ev0, k->ev1));
        k->t_ms += ms;
        k->ncalls++;
    }
    return 0;
}
}  // namespace

extern "C" {

const char *saship_last_error(void) { return last_error().c_str(); }

int saship_device_count(int32_t *count)
{
    if (!count) return fail(SASHIP_EINVAL, "count is NULL");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return 0;
}

int saship_init(int32_t device)
{
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n)
        return fail(SASHIP_EINVAL, "device " + std::to_string(device) + " out of range (" + std::to_string(n) + " devices)");
    HIP_TRY(hipSetDevice(device));
    current_device() = device;
    char *sp = nullptr;
    return scratch_address(device, &sp);
}

int saship_host_alloc(int64_t bytes, void **ptr)
{
    if (!ptr) return fail(SASHIP_EINVAL, "ptr is NULL");
    *ptr = nullptr;
    if (bytes < 1) return fail(SASHIP_EINVAL, "bytes must be >= 1");
    HIP_TRY(hipSetDevice(current_device()));
    void *p = nullptr;
    const hipError_t e =
        hipHostMalloc(&p, (size_t)bytes, hipHostMallocMapped | hipHostMallocCoherent | hipHostMallocPortable);
    if (e != hipSuccess || !p) return fail(SASHIP_ENOMEM, std::string("pinned host memory: ") + hipGetErrorString(e));
    *ptr = p;
    return 0;
}

int saship_host_free(void *ptr)
{
    if (!ptr) return 0;
    HIP_TRY(hipHostFree(ptr));
    return 0;
}

int saship_probe_fma(int32_t iters, double *tflops)
{
    if (!tflops) return fail(SASHIP_EINVAL, "tflops is NULL");
    if (iters < 1 || iters > (1 << 20)) return fail(SASHIP_EINVAL, "iters must be 1..2^20");
    HIP_TRY(hipSetDevice(current_device()));
    const int blocks = cu_count(current_device()) * 16;
    float *out = nullptr;
    HIP_TRY(hipMalloc((void **)&out, sizeof(float) * 256 * (size_t)blocks));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    float ms = 0.0f;
    if (e == hipSuccess) {
        k_fma_probe<<<blocks, 256>>>(out, 16, 0.999f, 0.001f);     // warm-up
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
    if (e == hipSuccess) {
        k_fma_probe<<<blocks, 256>>>(out, iters, 0.999f, 0.001f);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(out);
    if (e != hipSuccess) return fail(SASHIP_EHIP, std::string("fma probe: ") + hipGetErrorString(e));
    const double flops = 2.0 * 64.0 * iters * 256.0 * blocks;
    *tflops = ms > 0.0f ? flops / (ms * 1e-3) / 1e12 : 0.0;
    return 0;
}

int saship_model_create(int32_t model_id, int32_t quad_n, saship_model **out)
{
    if (!out) return fail(SASHIP_EINVAL, "out is NULL");
    *out = nullptr;
    if (model_id < 0 || model_id > MODEL_LAST)
        return fail(SASHIP_EINVAL, "unknown model id " + std::to_string(model_id));
    if (quad_n < 0 || quad_n > 100000) return fail(SASHIP_EINVAL, "quadrature size must be 0 (adaptive) or 1..100000");
    HIP_TRY(hipSetDevice(current_device()));
    saship_model *m = new (std::nothrow) saship_model();
    if (!m) return fail(SASHIP_ENOMEM, "out of host memory");
    m->id = model_id;
    m->quad_n = quad_n;
    m->dev = current_device();
    // orientation-quadrature tables (the models whose 1D kernel averages orientation)
    const bool cyl_family = model_id == 1 || model_id == 3 || model_id == 21;
    const bool ell = model_id == 2 || model_id == 19;
    const bool ppd = model_id == 6, tri = model_id == 7, hollow = model_id == 11, ecy = model_id == 20;
    std::vector<int> sizes;
    if (quad_n > 0) sizes = {quad_n};
    else sizes = {20, 76, 500, 5000};
    std::vector<float> a, b, w, c, d;
    std::vector<int> offs;
    if (cyl_family || ell || ppd || tri || hollow || ecy) {
        for (int n : sizes) {
            std::vector<double> z, wt;
            gauss_legendre(n, z, wt);
            offs.push_back((int)a.size());
            for (int i = 0; i < n; i++) {
                const double th = z[i] * sas::PI_4_D + sas::PI_4_D;   // [0, pi/2]
                const double u = z[i] * 0.5 + 0.5;                       // [0, 1]
                float ta = 0.0f, tb = 0.0f, tw = (float)wt[i], tc = 0.0f, td = 0.0f;
                if (cyl_family) {
                    ta = (float)sin(th);
                    tb = (float)cos(th);
                    tw = (float)(wt[i] * sin(th));
                } else if (ell) {
                    ta = (float)(u * u);
                } else if (hollow) {
                    // cos(theta) = u on [0, 1]
                    ta = (float)sqrt(1.0 - u * u);
                    tb = (float)u;
                } else if (ecy) {
                    // inner angle theta = 2 th on [0, pi]: cos^2 and sin^2 of th;
                    // outer cos(alpha) = u on [0, 1]
                    ta = (float)(cos(th) * cos(th));
                    tb = (float)(sin(th) * sin(th));
                    tc = (float)u;
                    td = (float)sqrt(1.0 - u * u);
                } else if (ppd) {
                    ta = (float)sin(th);
                    tb = (float)cos(th);
                    tc = (float)u;
                    td = (float)sqrt(1.0 - u * u);
                } else {
                    ta = (float)(u * u);
                    tc = (float)(sin(th) * sin(th));
                }
                a.push_back(ta);
                b.push_back(tb);
                w.push_back(tw);
                c.push_back(tc);
                d.push_back(td);
            }
        }
    } else {
        offs.push_back(0);
        a.push_back(0.0f);
        b.push_back(0.0f);
        w.push_back(0.0f);
        c.push_back(0.0f);
        d.push_back(0.0f);
    }
    m->ntab = (int)a.size();
    hipError_t e = hipMalloc((void **)&m->tab, sizeof(float) * 5 * m->ntab);
    const std::vector<float> *parts[5] = {&a, &b, &w, &c, &d};
    for (int k = 0; k < 5 && e == hipSuccess; k++)
        e = hipMemcpy(m->tab + k * m->ntab, parts[k]->data(), sizeof(float) * m->ntab, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (m->tab) (void)hipFree(m->tab);
        delete m;
        return fail(SASHIP_EHIP, std::string("quadrature upload: ") + hipGetErrorString(e));
    }
    m->quad.a = m->tab;
    m->quad.b = m->tab + m->ntab;
    m->quad.w = m->tab + 2 * m->ntab;
    m->quad.c = m->tab + 3 * m->ntab;
    m->quad.d = m->tab + 4 * m->ntab;
    m->quad.g64 = nullptr;
    if (model_paracrystal(model_id)) {
        // FP64 product rule over (theta, phi): sin/cos of both angles and w;
        // bcc, fcc: theta = pi/2 (z + 1), phi = pi (z + 1); sc: both pi/4 (z + 1)
        const int n = quad_n > 0 ? quad_n : 150;
        std::vector<double> z, wt, g(5 * (size_t)n);
        gauss_legendre(n, z, wt);
        const bool sc = model_id == 10;
        for (int i = 0; i < n; i++) {
            const double th = sc ? z[i] * sas::PI_4_D + sas::PI_4_D : (z[i] + 1.0) * (0.5 * sas::PI_D);
            const double ph = sc ? z[i] * sas::PI_4_D + sas::PI_4_D : (z[i] + 1.0) * sas::PI_D;
            g[i] = sin(th);
            g[n + i] = cos(th);
            g[2 * n + i] = sin(ph);
            g[3 * n + i] = cos(ph);
            g[4 * n + i] = wt[i];
        }
        e = hipMalloc((void **)&m->tab64, sizeof(double) * g.size());
        if (e == hipSuccess) e = hipMemcpy(m->tab64, g.data(), sizeof(double) * g.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(m->tab);
            if (m->tab64) (void)hipFree(m->tab64);
            delete m;
            return fail(SASHIP_EHIP, std::string("quadrature upload: ") + hipGetErrorString(e));
        }
        m->quad.g64 = m->tab64;
        m->quad_n = n;
    }
    m->quad.n = m->quad_n;
    for (int i = 0; i < 4; i++) m->quad.off[i] = i < (int)offs.size() ? offs[i] : 0;
    *out = m;
    return 0;
}

int saship_model_release(saship_model *m)
{
    if (!m) return 0;
    if (m->released.exchange(true)) return fail(SASHIP_EINVAL, "model released twice");
    if (m->refs.load() == 0) model_free(m);
    return 0;
}

int saship_kernel_create(saship_model *m, int32_t nq, const double *q, int32_t is2d, saship_kernel **out)
{
    if (!out) return fail(SASHIP_EINVAL, "out is NULL");
    *out = nullptr;
    if (!m) return fail(SASHIP_EINVAL, "model is NULL");
    if (nq < 0) return fail(SASHIP_EINVAL, "nq must be >= 0");
    if (nq > 0 && !q) return fail(SASHIP_EINVAL, "q is NULL");
    HIP_TRY(hipSetDevice(m->dev));
    saship_kernel *k = new (std::nothrow) saship_kernel();
    if (!k) return fail(SASHIP_ENOMEM, "out of host memory");
    k->m = m;
    model_retain(m);
    k->nq = nq;
    k->is2d = is2d ? 1 : 0;
    k->dev = m->dev;
    k->cus = cu_count(m->dev);
    const size_t nqv = (size_t)nq * (k->is2d ? 2 : 1);
    std::vector<float> q32(std::max<size_t>(nqv, 1));
    for (size_t i = 0; i < nqv; i++) q32[i] = (float)q[i];
    if (m->id == 1 && !k->is2d && m->quad_n == 0 && nq > 0) k->q_host.assign(q, q + nq);
    hipError_t e = hipStreamCreateWithFlags(&k->st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc((void **)&k->q32, sizeof(float) * std::max<size_t>(nqv, 1));
    if (e == hipSuccess) e = hipMalloc((void **)&k->q64, sizeof(double) * std::max<size_t>(nqv, 1));
    if (e == hipSuccess) e = hipMalloc((void **)&k->d_meta, META_BYTES);
    if (e == hipSuccess) e = hipHostMalloc((void **)&k->h_result, sizeof(double) * (2 * (size_t)nq + 4),
                                             hipHostMallocMapped | hipHostMallocCoherent);
    if (e == hipSuccess) e = hipHostGetDevicePointer((void **)&k->hd_result, k->h_result, 0);
    if (e == hipSuccess && nqv) e = hipMemcpy(k->q32, q32.data(), sizeof(float) * nqv, hipMemcpyHostToDevice);
    if (e == hipSuccess && nqv) e = hipMemcpy(k->q64, q, sizeof(double) * nqv, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipEventCreate(&k->ev0);
    if (e == hipSuccess) e = hipEventCreate(&k->ev1);
    if (e != hipSuccess) {
        std::string msg = std::string("kernel create: ") + hipGetErrorString(e);
        saship_kernel_release(k);
        return fail(SASHIP_EHIP, msg);
    }
    *out = k;
    return 0;
}

int saship_kernel_call(saship_kernel *k, int32_t kind, const int32_t *details, int32_t n_details,
                       const double *values, int32_t n_values, double cutoff, int32_t reff_mode,
                       int32_t pd_start, int32_t pd_stop, double *result)
{
    if (!k) return fail(SASHIP_EINVAL, "kernel is NULL");
    if (!details || !values || !result) return fail(SASHIP_EINVAL, "details, values and result must not be NULL");
    if (kind < 0 || kind > 2) return fail(SASHIP_EINVAL, "kind must be 0 (Iq), 1 (Iqxy) or 2 (Imagnetic)");
    if ((kind == 0) == (k->is2d != 0))
        return fail(SASHIP_EINVAL, kind == 0 ? "Iq kernel needs 1D q" : "2D kernels need (qx, qy) input");
    const int id = k->m->id;
    if (pd_stop < pd_start) return fail(SASHIP_EINVAL, "need 0 <= pd_start <= pd_stop");
    if (int e = validate_set(id, kind, details, n_details, n_values, pd_start, pd_stop, -1)) return e;
    std::lock_guard<std::mutex> lock(k->mu);
    HIP_TRY(hipSetDevice(k->dev));
    double ext[4];
    if (!k->q_host.empty() && kind == 0 && k->separable && pd_start == 0 &&
        pd_stop == details[4 * model_maxpd(id)] && cylinder_separable(details) &&
        adaptive_separable(details, values, cutoff, ext)) {
        // the whole mesh of a cylinder under the adaptive rule: the one-set
        // batch's separable path (the same raw sums, in another order)
        if (!k->adapt) {
            const int32_t nq1 = k->nq;
            if (int e = saship_batch_create(k->m, 1, &nq1, k->q_host.data(), 0, &k->adapt)) return e;
        }
        const int64_t voff[2] = {0, n_values};
        return saship_batch_call(k->adapt, 0, details, n_details, values, voff, cutoff, reff_mode, result, nullptr);
    }
    const int nq = k->nq, nout = kind == 0 && model_has_fq(id) ? 2 : 1;
    const size_t vbytes = sizeof(double) * (size_t)n_values;
    const size_t dbytes = sizeof(int32_t) * (size_t)n_details;
    const size_t doff = (vbytes + 255) & ~(size_t)255;
    if (int e = ensure_host(k->h_stage, k->stage_cap, doff + dbytes)) return e;
    if (int e = ensure(k->d_args, k->args_cap, doff + dbytes)) return e;
    memcpy(k->h_stage, values, vbytes);
    memcpy(k->h_stage + doff, details, dbytes);
    const size_t rbytes = sizeof(double) * ((size_t)nout * nq + 4);
    const int accumulate = pd_start > 0 ? 1 : 0;
    if (accumulate) memcpy(k->h_result, result, rbytes);
    // device buffers before any capture (allocations are never captured)
    int nlp = 0;
    const int path = choose_path(k->m, kind, details, k->separable, &nlp);
    const int npts_max = std::max(0, pd_stop - pd_start);
    int nsplit = 0;
    if (path == 1) {
        if (!k->d_sep) HIP_TRY(hipMalloc((void **)&k->d_sep, sas::sep::tables_bytes(1, SEP_MAX)));
        if (int e = ensure(k->d_partial, k->partial_cap, sizeof(double) * 2 * (size_t)nq * k->m->quad.n)) return e;
    } else if (path == 2) {
        nsplit = split_count(k->cus, nq, npts_max);
        if (int e = ensure(k->d_recs, k->recs_cap, sizeof(float4) * (size_t)std::max(1, npts_max) * record_width(id, kind)))
            return e;
        if (int e = ensure(k->d_partial, k->partial_cap, sizeof(double) * (kind == 0 ? 2 : 1) * (size_t)nq * nsplit))
            return e;
    }
    const int32_t *d_det = (const int32_t *)(k->d_args + doff);
    const double *d_val = (const double *)k->d_args;
    // the final kernel reads (accumulate) and writes the result in mapped
    // host memory: no transfer of the result in either direction
    Target t{k->st, k->m, nq, k->is2d, k->cus, k->q32, k->q64, k->d_recs, k->d_meta, k->d_partial, k->hd_result, k->d_sep,
             k->timing ? k->ev0 : nullptr, k->timing ? k->ev1 : nullptr};
    const size_t staged = doff + dbytes;
    const bool graphed = k->use_graph && !k->timing;
    // upload and evaluation: one sequence on the handle's stream (a captured
    // graph copies from the staging buffer: kernel arguments would be frozen)
    auto enqueue = [&]() -> int {
        if (!graphed && staged <= (size_t)sas::STAGE_ARG_BYTES) {
            sas::StageIn in;
            in.nbytes = (int)staged;
            memcpy(in.data, k->h_stage, staged);
            // the separable expansion reads the argument itself (nothing
            // later reads the values or details)
            if (path == 1)
                return launch_sep(t, nlp, d_det, d_val, n_details, pd_start, pd_stop, cutoff, reff_mode, accumulate,
                                  &in, (int)doff);
            sas::k_stage_in<<<1, 128, 0, k->st>>>(in, (unsigned char *)k->d_args);
            HIP_TRY(hipGetLastError());
        } else {
            HIP_TRY(hipMemcpyAsync(k->d_args, k->h_stage, staged, hipMemcpyHostToDevice, k->st));
        }
        int rc;
        if (path == 0) rc = launch_structure(t, d_val, cutoff, accumulate);
        else if (path == 1) rc = launch_sep(t, nlp, d_det, d_val, n_details, pd_start, pd_stop, cutoff, reff_mode, accumulate);
        else rc = launch_model(t, kind, d_det, d_val, n_details, pd_start, pd_stop, cutoff, reff_mode, accumulate, nsplit);
        return rc;
    };
    if (graphed) {
        int64_t cbits;
        memcpy(&cbits, &cutoff, sizeof cbits);
        const std::vector<int64_t> key = {kind, path, nlp, nsplit, (int64_t)(doff + dbytes), (int64_t)doff, accumulate, pd_start,
                                          pd_stop, reff_mode, cbits, (int64_t)rbytes, (int64_t)(intptr_t)k->d_args,
                                          (int64_t)(intptr_t)k->d_recs, (int64_t)(intptr_t)k->d_partial,
                                          (int64_t)(intptr_t)k->d_sep, (int64_t)(intptr_t)k->h_stage};
        if (!k->gexec || key != k->gkey) {
            drop_graph(k);
            HIP_TRY(hipStreamBeginCapture(k->st, hipStreamCaptureModeThreadLocal));
            const int rc = enqueue();
            hipGraph_t g = nullptr;
            const hipError_t ce = hipStreamEndCapture(k->st, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (ce != hipSuccess) return fail(SASHIP_EHIP, std::string("graph capture: ") + hipGetErrorString(ce));
            k->graph = g;
            HIP_TRY(hipGraphInstantiate(&k->gexec, g, nullptr, nullptr, 0));
            k->gkey = key;
        }
        HIP_TRY(hipGraphLaunch(k->gexec, k->st));
    } else {
        if (int e = enqueue()) return e;
    }
    return finish_call(k, rbytes, result);
}

int saship_kernel_call_noalloc(saship_model *m, int32_t kind, int32_t nq, const double *q,
                               const int32_t *details, int32_t n_details, const double *values, int32_t n_values,
                               double cutoff, int32_t reff_mode, int32_t pd_start, int32_t pd_stop,
                               double *result, void *stream)
{
    if (!m) return fail(SASHIP_EINVAL, "model is NULL");
    if (!details || !values || !result) return fail(SASHIP_EINVAL, "details, values and result must not be NULL");
    if (kind < 0 || kind > 2) return fail(SASHIP_EINVAL, "kind must be 0 (Iq), 1 (Iqxy) or 2 (Imagnetic)");
    if (nq < 0) return fail(SASHIP_EINVAL, "nq must be >= 0");
    if (nq > 0 && !q) return fail(SASHIP_EINVAL, "q is NULL");
    const int id = m->id;
    if (pd_stop < pd_start) return fail(SASHIP_EINVAL, "need 0 <= pd_start <= pd_stop");
    if (int e = validate_set(id, kind, details, n_details, n_values, pd_start, pd_stop, -1)) return e;
    HIP_TRY(hipSetDevice(m->dev));
    if ((nq > 0 && !device_accessible(q)) || !device_accessible(result))
        return fail(SASHIP_EINVAL, "q and result must be device memory");
    char *base = nullptr;
    if (int e = scratch_address(m->dev, &base)) return e;
    hipStream_t st = (hipStream_t)stream;
    const int is2d = kind != 0;
    const int cus = cu_count(m->dev);

    // carve: values | details | q32 | meta | records | separable tables | partials
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off = (off + bytes + 255) & ~(size_t)255;
        return at;
    };
    const size_t o_val = take(sizeof(double) * (size_t)n_values);
    const size_t o_det = take(sizeof(int32_t) * (size_t)n_details);
    const int64_t nqv = (int64_t)nq * (is2d ? 2 : 1);
    const size_t o_q32 = take(sizeof(float) * (size_t)std::max<int64_t>(1, nqv));
    const size_t o_meta = take(META_BYTES);
    int nlp = 0;
    const int path = choose_path(m, kind, details, true, &nlp);
    const int npts_max = std::max(0, pd_stop - pd_start);
    const size_t o_recs = take(path == 2 ? sizeof(float4) * (size_t)std::max(1, npts_max) * record_width(id, kind) : 0);
    const size_t o_sep = take(path == 1 ? sas::sep::tables_bytes(1, SEP_MAX) : 0);
    const size_t o_part = off;
    if (o_part > NOALLOC_SCRATCH_BYTES)
        return fail(SASHIP_EINVAL, "call needs more than the " + std::to_string(NOALLOC_SCRATCH_BYTES >> 20) +
                                       " MiB no-allocation scratch (mesh or values too large)");
    const size_t room = NOALLOC_SCRATCH_BYTES - o_part;
    int nsplit = 1;
    size_t part_bytes = 0;
    if (path == 1) {
        part_bytes = sizeof(double) * 2 * (size_t)nq * m->quad.n;
    } else if (path == 2) {
        const size_t per_split = sizeof(double) * (kind == 0 ? 2 : 1) * (size_t)std::max(1, nq);
        nsplit = split_count(cus, nq, npts_max);
        nsplit = (int)std::min<size_t>((size_t)nsplit, room / per_split);
        part_bytes = per_split * (size_t)nsplit;
    }
    if (nsplit < 1 || part_bytes > room)
        return fail(SASHIP_EINVAL, "call needs more than the " + std::to_string(NOALLOC_SCRATCH_BYTES >> 20) +
                                       " MiB no-allocation scratch (too many q points)");

    double *d_val = (double *)(base + o_val);
    int32_t *d_det = (int32_t *)(base + o_det);
    float *q32 = (float *)(base + o_q32);
    HIP_TRY(hipMemcpyAsync(d_val, values, sizeof(double) * (size_t)n_values, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(d_det, details, sizeof(int32_t) * (size_t)n_details, hipMemcpyHostToDevice, st));
    if (nqv > 0) {
        sas::k_q_to_f32<<<(unsigned)((nqv + 255) / 256), 256, 0, st>>>(nqv, q, q32);
        HIP_TRY(hipGetLastError());
    }
    Target t{st, m, nq, is2d, cus, q32, q, (float4 *)(base + o_recs), base + o_meta, (double *)(base + o_part),
             result, base + o_sep, nullptr, nullptr};
    const int accumulate = pd_start > 0 ? 1 : 0;
    if (path == 0) return launch_structure(t, d_val, cutoff, accumulate);
    if (path == 1) return launch_sep(t, nlp, d_det, d_val, n_details, pd_start, pd_stop, cutoff, reff_mode, accumulate);
    return launch_model(t, kind, d_det, d_val, n_details, pd_start, pd_stop, cutoff, reff_mode, accumulate, nsplit);
}

int saship_structure_eval(int32_t s_model, const double *pars, int32_t nq, const double *q, int32_t on_device,
                          double *prep, double *sq)
{
    if (!pars || !prep || (nq > 0 && (!q || !sq))) return fail(SASHIP_EINVAL, "NULL argument");
    if (!model_structure(s_model)) return fail(SASHIP_EINVAL, "model " + std::to_string(s_model) + " is not a structure factor");
    if (nq < 0) return fail(SASHIP_EINVAL, "nq must be >= 0");
    double p[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < model_npars(s_model); k++) p[k] = pars[k];
    const int id = s_model;     // the structure-factor ids are the model ids
    if (!on_device) {
        for (int k = 0; k < sas::S_PREP; k++) prep[k] = 0.0;
        sas::structure_prepare(id, p, prep);
        for (int i = 0; i < nq; i++) sq[i] = sas::structure_sq_prepared(id, q[i], p, prep);
        return 0;
    }
    HIP_TRY(hipSetDevice(current_device()));
    double *d = nullptr;
    const size_t n = 6 + (size_t)sas::S_PREP + 2 * (size_t)nq;
    HIP_TRY(hipMalloc((void **)&d, sizeof(double) * n));
    hipError_t e = hipMemcpy(d, p, sizeof(double) * 6, hipMemcpyHostToDevice);
    if (e == hipSuccess && nq > 0) e = hipMemcpy(d + 6 + sas::S_PREP, q, sizeof(double) * nq, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        sas::k_structure_eval<<<1, 256>>>(id, d, nq, d + 6 + sas::S_PREP, d + 6, d + 6 + sas::S_PREP + nq);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(prep, d + 6, sizeof(double) * sas::S_PREP, hipMemcpyDeviceToHost);
    if (e == hipSuccess && nq > 0) e = hipMemcpy(sq, d + 6 + sas::S_PREP + nq, sizeof(double) * nq, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(SASHIP_EHIP, std::string("structure eval: ") + hipGetErrorString(e));
    return 0;
}

int saship_kernel_set_path(saship_kernel *k, int32_t separable)
{
    if (!k) return fail(SASHIP_EINVAL, "kernel is NULL");
    std::lock_guard<std::mutex> lock(k->mu);
    k->separable = separable != 0;
    return 0;
}

int saship_kernel_set_graph(saship_kernel *k, int32_t enable)
{
    if (!k) return fail(SASHIP_EINVAL, "kernel is NULL");
    std::lock_guard<std::mutex> lock(k->mu);
    k->use_graph = enable != 0;
    drop_graph(k);
    return 0;
}

int saship_kernel_set_timing(saship_kernel *k, int32_t enable)
{
    if (!k) return fail(SASHIP_EINVAL, "kernel is NULL");
    std::lock_guard<std::mutex> lock(k->mu);
    k->timing = enable != 0;
    k->t_ms = 0.0;
    k->ncalls = 0;
    return 0;
}

int saship_kernel_get_timing(saship_kernel *k, double *eval_ms, int32_t *ncalls)
{
    if (!k || !eval_ms || !ncalls) return fail(SASHIP_EINVAL, "NULL argument");
    std::lock_guard<std::mutex> lock(k->mu);
    *eval_ms = k->t_ms;
    *ncalls = k->ncalls;
    return 0;
}

int saship_kernel_release(saship_kernel *k)
{
    if (!k) return 0;
    (void)hipSetDevice(k->dev);
    if (k->st) (void)hipStreamSynchronize(k->st);
    drop_graph(k);
    if (k->q32) (void)hipFree(k->q32);
    if (k->q64) (void)hipFree(k->q64);
    if (k->d_args) (void)hipFree(k->d_args);
    if (k->d_recs) (void)hipFree(k->d_recs);
    if (k->d_meta) (void)hipFree(k->d_meta);
    if (k->d_partial) (void)hipFree(k->d_partial);
    if (k->d_sep) (void)hipFree(k->d_sep);
    if (k->adapt) (void)saship_batch_release(k->adapt);
    if (k->h_stage) (void)hipHostFree(k->h_stage);
    if (k->h_result) (void)hipHostFree(k->h_result);
    if (k->ev0) (void)hipEventDestroy(k->ev0);
    if (k->ev1) (void)hipEventDestroy(k->ev1);
    if (k->st) (void)hipStreamDestroy(k->st);
    if (k->m) model_drop(k->m);
    delete k;
    return 0;
}

}  // extern "C"
