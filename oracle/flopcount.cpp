// flopcount.cpp -- counts the floating-point operations of the reference
// formulas.  TEST AND MEASUREMENT INFRASTRUCTURE ONLY, like the rest of
// oracle/: it gives the algorithmic FLOP figures of the benchmark workloads.
//
// sasoracle.c restates the reference's kernel (kernel_iq.c:317-858) and model
// formulas in a scalar type `oreal`.  Here that file is compiled as C++ with
// oreal = Counted, a double that tallies every operation it performs, so a
// kernel call counts exactly the operations the reference formulas take on
// the given inputs, on the branches actually taken (the J1 |x| <= 5 split,
// the 3j1(x)/x Taylor branch, the adaptive rule sizes).  Convention:
//   add, sub, mul, div, sqrt: 1 each; sin, cos, exp, log, pow, cbrt, asin,
//   expm1, atan, tgamma: 1 each, also tallied as transcendental; an FMA would be 2 (the
//   restatement has none: it is built without contraction).  Sign changes,
//   fabs, fmin/fmax and comparisons are not counted.
// Per-thread tallies (OpenMP over q) are summed after each call.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

namespace flopcount {

enum { ADD, MUL, DIV, SQRT, TRANS, NCAT };

struct Tally {
    uint64_t n[NCAT] = {0, 0, 0, 0, 0};
};

std::mutex registry_mu;
std::vector<Tally *> registry;

// one tally per thread, registered on first use and kept for the process
struct ThreadTally {
    Tally t;
    ThreadTally()
    {
        std::lock_guard<std::mutex> lock(registry_mu);
        registry.push_back(&t);
    }
};
inline Tally &tally()
{
    static thread_local ThreadTally tt;
    return tt.t;
}

}  // namespace flopcount

struct Counted {
    double v;
    Counted() = default;
    constexpr Counted(double x) : v(x) {}
    explicit operator double() const { return v; }
    explicit operator int() const { return (int)v; }
    Counted &operator+=(Counted o) { flopcount::tally().n[flopcount::ADD]++; v += o.v; return *this; }
    Counted &operator-=(Counted o) { flopcount::tally().n[flopcount::ADD]++; v -= o.v; return *this; }
    Counted &operator*=(Counted o) { flopcount::tally().n[flopcount::MUL]++; v *= o.v; return *this; }
    Counted &operator/=(Counted o) { flopcount::tally().n[flopcount::DIV]++; v /= o.v; return *this; }
};

inline Counted operator+(Counted a, Counted b) { flopcount::tally().n[flopcount::ADD]++; return Counted(a.v + b.v); }
inline Counted operator-(Counted a, Counted b) { flopcount::tally().n[flopcount::ADD]++; return Counted(a.v - b.v); }
inline Counted operator*(Counted a, Counted b) { flopcount::tally().n[flopcount::MUL]++; return Counted(a.v * b.v); }
inline Counted operator/(Counted a, Counted b) { flopcount::tally().n[flopcount::DIV]++; return Counted(a.v / b.v); }
inline Counted operator-(Counted a) { return Counted(-a.v); }
inline Counted operator+(Counted a) { return a; }
inline bool operator<(Counted a, Counted b) { return a.v < b.v; }
inline bool operator>(Counted a, Counted b) { return a.v > b.v; }
inline bool operator<=(Counted a, Counted b) { return a.v <= b.v; }
inline bool operator>=(Counted a, Counted b) { return a.v >= b.v; }
inline bool operator==(Counted a, Counted b) { return a.v == b.v; }
inline bool operator!=(Counted a, Counted b) { return a.v != b.v; }
inline bool operator!(Counted a) { return !a.v; }

// The wrappers are always inlined so that the compiler treats sin and cos of
// one argument as it does in the C build of the oracle (fused into one
// sincos call, whose last bit can differ from separate calls): the counted
// results then equal liboracle's bit for bit whatever else the file holds.
#define FLOPCOUNT_TRANS(fn)                                                               \
    __attribute__((always_inline)) inline Counted fn(Counted a)                          \
    {                                                                                     \
        flopcount::tally().n[flopcount::TRANS]++;                                         \
        return Counted(std::fn(a.v));                                                     \
    }
FLOPCOUNT_TRANS(sin)
FLOPCOUNT_TRANS(cos)
FLOPCOUNT_TRANS(exp)
FLOPCOUNT_TRANS(log)
FLOPCOUNT_TRANS(cbrt)
FLOPCOUNT_TRANS(asin)
FLOPCOUNT_TRANS(expm1)
FLOPCOUNT_TRANS(atan)
FLOPCOUNT_TRANS(tgamma)
#undef FLOPCOUNT_TRANS
inline Counted pow(Counted a, Counted b)
{
    flopcount::tally().n[flopcount::TRANS]++;
    return Counted(std::pow(a.v, b.v));
}
inline Counted sqrt(Counted a)
{
    flopcount::tally().n[flopcount::SQRT]++;
    return Counted(std::sqrt(a.v));
}
inline Counted fabs(Counted a) { return Counted(std::fabs(a.v)); }
inline Counted fmax(Counted a, Counted b) { return Counted(std::fmax(a.v, b.v)); }
inline Counted fmin(Counted a, Counted b) { return Counted(std::fmin(a.v, b.v)); }

#define SASORACLE_COUNTING 1
typedef Counted oreal;
#include "sasoracle.c"

extern "C" {

// The reference-shaped kernel call of sasoracle.c on counted arithmetic.
// values, q and result are doubles as for liboracle.so; counts receives
// (add+sub, mul, div, sqrt, transcendental) of this call.  Returns the
// status of the kernel call.
int flopcount_kernel(int model_id, int kind, int quad_n, int32_t nq, int32_t pd_start, int32_t pd_stop,
                     const int32_t *details, const double *values, int32_t n_values, const double *q,
                     double *result, int32_t n_result, double cutoff, int32_t reff_mode, uint64_t *counts)
{
    const int32_t nqv = kind == 0 ? nq : 2 * nq;
    std::vector<Counted> cv(values, values + n_values), cq(q, q + nqv), cr(result, result + n_result);
    {
        std::lock_guard<std::mutex> lock(flopcount::registry_mu);
        for (flopcount::Tally *t : flopcount::registry) *t = flopcount::Tally();
    }
    const int rc = sasoracle_kernel(model_id, kind, quad_n, nq, pd_start, pd_stop, details, cv.data(), cq.data(),
                                    cr.data(), Counted(cutoff), reff_mode);
    uint64_t sum[flopcount::NCAT] = {0, 0, 0, 0, 0};
    {
        std::lock_guard<std::mutex> lock(flopcount::registry_mu);
        for (flopcount::Tally *t : flopcount::registry)
            for (int k = 0; k < flopcount::NCAT; k++) sum[k] += t->n[k];
    }
    for (int k = 0; k < flopcount::NCAT; k++) counts[k] = sum[k];
    for (int32_t i = 0; i < n_result; i++) result[i] = cr[i].v;
    return rc;
}

void flopcount_set_threads(int n) { sasoracle_set_threads(n); }

}  // extern "C"
