// batch.hip -- many parameter sets, each with its own q vector and mesh, in
// one persistent launch.  No reference counterpart: the reference evaluates
// a fit population one Kernel call at a time (bumps_model.py:250-285).
//
// One call runs on the batch's stream:
//   H2D staging (values, details, work list, per-set tables, all sets)
//   k_expand        one workgroup per set (kernels.h); resets the work queue
//   k_batch_eval_*  persistent: CUs * waves_per_cu / 4 workgroups; every
//                   wave64 takes (set, 64-q tile, mesh chunk) items from an
//                   atomic counter -- longest items first -- and reads the
//                   item's records with wave-uniform (scalar) loads
//   k_batch_final   one workgroup per set: per output, the chunk partials in
//                   chunk order (FP64), the weight sums, and I(q) with the
//                   normalisation of Kernel.Iq (sasmodels/kernel.py:80-188),
//                   or of a P@S product with the hard-sphere S(q) evaluated
//                   there in FP64 (saship_batch_set_structure)
//   D2H             raw result and/or I(q)
// The whole sequence can be captured once in a hipGraph and replayed for
// every later call with the same shapes (a fit step).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "common.h"
#include "kernels.h"
#include "structure.h"

using namespace saship_host;

namespace sas {

constexpr int QTILE = 64;   // q points per work item: one per lane

struct BatchArgs {
    const float *q32;            // 1D: q; 2D: interleaved (qx, qy)
    const double *q64;           // 2D: interleaved (qx, qy) in FP64
    const int64_t *qoff;         // [nsets+1] q offsets (outputs)
    const int64_t *qdoff;        // [nsets] offsets of each set's q values (0 for all when the sets share q)
    const double *values;
    const int64_t *voff;         // [nsets]
    const float4 *recs;
    const int64_t *roff;         // [nsets] record offsets
    const int64_t *poff;         // [nsets] partial offsets (doubles)
    const int *npts;             // [nsets] accepted points (from k_expand)
    const float *spin;           // magnetic: per-set spin constants (from k_expand)
    const int4 *items;           // (set, tile, chunk, 0)
    const int *n_items;
    int pts_per_item;
    Quad quad;
    int grab;                    // items per queue atomic
    double *iq;                  // every set one chunk, no final kernel: I(q) out (NULL: partial sums)
    const double *sums;          // [nsets][4] weight sums (from k_expand), for iq
};

// I(q) of set s from its <F^2> sum when the set is one chunk: the final
// kernel's normalisation (k_batch_final without a structure factor), in its
// operation order, so the result is the same to the bit.
__device__ __forceinline__ double direct_iq(const BatchArgs &a, int s, double f2)
{
    const double *sm = a.sums + 4 * s;
    double total = sm[0];
    if (total == 0.0) total = 1.0;
    double shell = sm[2] / total;
    if (shell == 0.0) shell = 1.0;
    const double *v = a.values + a.voff[s];
    return v[0] / shell * (f2 / total) + v[1];
}

// The queue counter and the partial sums are separate __restrict__ kernel
// arguments: with every store of a kernel going through them, the compiler
// can read the read-only tables through the scalar cache.

// A wave's position in the work queue: it starts on its own counter, moves
// on when that one runs dry and stops after finding every counter dry in a
// row (counters only grow, so every item is handed out exactly once).
struct QueueCursor {
    int c, dry;
    int next = 0, end = 0;       // items taken by the last atomic and not yet handed out
};

__device__ __forceinline__ QueueCursor queue_start()
{
    // wave-uniform (the compiler cannot tell from threadIdx): keeps the
    // counter, the item and every table address derived from it in scalars
    const int c = (int)((blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) % WORK_QUEUES);
    return QueueCursor{__builtin_amdgcn_readfirstlane(c), 0, 0, 0};
}

// Next item of the wave, or -1 when the queue is empty.  One atomic takes
// `grab` consecutive items of a counter's range: small items (a few mesh
// points each) would otherwise spend their time on the counters' atomics.
__device__ __forceinline__ int queue_next(int *queue, int n_items, QueueCursor &cur, int grab = 1)
{
    if (cur.next < cur.end) return cur.next++;
    for (;;) {
        const int first = (int)((int64_t)cur.c * n_items / WORK_QUEUES);
        const int last = (int)((int64_t)(cur.c + 1) * n_items / WORK_QUEUES);
        int got = last;
        if ((threadIdx.x & 63) == 0) {
            int *ctr = queue + cur.c * WORK_QUEUE_STRIDE;
            // a stale read can only be low: then the atomic tells
            if (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < last - first)
                got = first + atomicAdd(ctr, grab);
        }
        const int item = __builtin_amdgcn_readfirstlane(__shfl(got, 0));
        if (item < last) {
            cur.dry = 0;
            cur.next = item + 1;
            cur.end = min(item + grab, last);
            return item;
        }
        if (++cur.dry == WORK_QUEUES) return -1;
        cur.c = cur.c + 1 == WORK_QUEUES ? 0 : cur.c + 1;
    }
}

template <class M, bool FIXED>
__global__ __launch_bounds__(256) void k_batch_eval_1d(BatchArgs a, int *__restrict__ queue,
                                                       double *__restrict__ partial)
{
    constexpr int RV = rv1<M>();
    const int lane = threadIdx.x & 63;
    const int n_items = *a.n_items;
    QueueCursor cur = queue_start();
    for (;;) {
        const int item = queue_next(queue, n_items, cur, a.grab);
        if (item < 0) break;
        const int4 it = a.items[item];
        const int s = it.x, tile = it.y, chunk = it.z;
        const int64_t q0 = a.qoff[s], qd0 = a.qdoff[s];
        const int nq = (int)(a.qoff[s + 1] - q0);
        const int iq = tile * QTILE + lane;
        const float q = iq < nq ? a.q32[qd0 + iq] : 0.0f;
        const int npts = a.npts[s];
        const int c0 = min(npts, chunk * a.pts_per_item);
        const int c1 = min(npts, c0 + a.pts_per_item);
        const float4 *rec = a.recs + a.roff[s] * RV;
        KSum a1, a2;
        double d1 = 0.0, d2 = 0.0;    // FP64 models
        bool shared = false;
        if constexpr (FIXED && M::SHARED_1D) {
            // every point of the chunk has the same lengths (a radius-only
            // mesh): length factors once per node for the whole chunk
            const float *rf = (const float *)rec;
            shared = c1 - c0 > 1 && M::shares_lengths(rf, 4 * RV, c0, c1);
            if (shared) M::eval1d_shared(q, rf, 4 * RV, c0, c1, a.quad, a1, a2);
        }
        if constexpr (!M::FP64 && RV == 1 && !M::SHARED_1D && !M::ORIENTED) {
            // one-float4 records of unoriented models: four records loaded
            // together per step, so their scalar loads and address
            // arithmetic are shared; the same sums in the same order
            int j = c0;
            for (; j + 4 <= c1; j += 4) {
                float4 r4[4];
#pragma unroll
                for (int t = 0; t < 4; t++) r4[t] = rec[j + t];
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    const float r[4] = {r4[t].x, r4[t].y, r4[t].z, r4[t].w};
                    float f1, f2;
                    M::template eval1d<FIXED>(q, r, a.quad, f1, f2);
                    a1.add(f1);
                    a2.add(f2);
                }
            }
            for (; j < c1; j++) {
                const float4 x = rec[j];
                const float r[4] = {x.x, x.y, x.z, x.w};
                float f1, f2;
                M::template eval1d<FIXED>(q, r, a.quad, f1, f2);
                a1.add(f1);
                a2.add(f2);
            }
            shared = true;     // summed
        }
        const double qd = M::FP64 && iq < nq ? a.q64[qd0 + iq] : 0.0;
        for (int j = c0; j < c1 && !shared; j++) {
            if constexpr (M::FP64 && RV > BIG_RECORD) {
                // a record too large for registers is read where it lies
                double f1, f2;
                M::eval1d_d(qd, (const float *)(rec + (size_t)j * RV), a.quad, f1, f2);
                d1 += f1;
                d2 += f2;
                continue;
            }
            float r[4 * RV];
            load_rec(rec + (size_t)j * RV, r, RV);
            if constexpr (M::FP64) {
                double f1, f2;
                M::eval1d_d(qd, r, a.quad, f1, f2);
                d1 += f1;
                d2 += f2;
            } else {
                float f1, f2;
                M::template eval1d<FIXED>(q, r, a.quad, f1, f2);
                a1.add(f1);
                a2.add(f2);
            }
        }
        if (iq < nq && a.iq) {
            a.iq[q0 + iq] = direct_iq(a, s, M::FP64 ? d2 : a2.value());
        } else if (iq < nq) {
            double *dst = partial + a.poff[s] + (size_t)chunk * 2 * nq;
            dst[iq] = M::FP64 ? d2 : a2.value();
            dst[nq + iq] = M::FP64 ? d1 : a1.value();
        }
    }
}

template <class M, bool MAG>
__global__ __launch_bounds__(256) void k_batch_eval_2d(BatchArgs a, int *__restrict__ queue,
                                                       double *__restrict__ partial)
{
    constexpr int RV = rv2<M>();
    const int lane = threadIdx.x & 63;
    const int n_items = *a.n_items;
    QueueCursor cur = queue_start();
    for (;;) {
        const int item = queue_next(queue, n_items, cur, a.grab);
        if (item < 0) break;
        const int4 it = a.items[item];
        const int s = it.x, tile = it.y, chunk = it.z;
        const int64_t q0 = a.qoff[s], qd0 = a.qdoff[s];
        const int nq = (int)(a.qoff[s + 1] - q0);
        const int iq = tile * QTILE + lane;
        const bool ok = iq < nq;
        Lane2D L;
        L.set(ok, ok ? a.q64[2 * (qd0 + iq)] : 0.0, ok ? a.q64[2 * (qd0 + iq) + 1] : 0.0);
        Contrast<M, MAG> con;
        con.setup(a.values + a.voff[s], MAG ? a.spin + (size_t)s * spin_floats<M>() : nullptr, L.qx, L.qy);
        const int npts = a.npts[s];
        const int c0 = min(npts, chunk * a.pts_per_item);
        const int c1 = min(npts, c0 + a.pts_per_item);
        const SAS_CONST float4 *rec = as_const(a.recs) + a.roff[s] * RV;
        KSum acc;
        for (int b0 = c0; b0 < c1; b0 += SUM_BLOCK) {
            const int b1 = min(c1, b0 + SUM_BLOCK);
            float blk = 0.0f;
            for (int j = b0; j < b1; j++) blk += point2d<M, MAG>(rec + (size_t)j * RV, L, con);
            acc.add(blk);
        }
        if (ok && a.iq) a.iq[q0 + iq] = direct_iq(a, s, acc.value());
        else if (ok) partial[a.poff[s] + (size_t)chunk * nq + iq] = acc.value();
    }
}

// Separable cylinder sets (radius x length product meshes, fixed Gauss
// rule): every wave takes (set, 64-q tile, block of Gauss nodes) items from
// the queue, sums its nodes (compensated FP32) from the set's tables and
// writes partial[poff[set] + (block*2 + o)*nq + iq], o = 0 for F^2, 1 for F.
// Items are numbered set by set in `order` (costliest first), tiles, then
// node blocks; ipre[p] is the first item of the p-th set in that order.
struct SepArgs {
    const float *q32;
    const int64_t *qoff;
    const int64_t *qdoff;
    sep::Tables t;
    Quad quad;
    const int *order;            // [nsets] set at each position
    const int *ipre;             // [nsets+1] item prefix counts
    int nsets;
    int nblk;                    // node blocks per (set, tile)
    int nodes_per_item;
    const int64_t *poff;
};

// One workgroup of four waves per item (set, 64-q tile, block of Gauss
// nodes): the block's nodes are dealt round-robin to the waves, each wave
// sums its nodes with compensation, and the four sums are combined in a
// fixed order through LDS, so one partial per output and q leaves the
// workgroup for the whole block (larger blocks: fewer partials in HBM, no
// loss of balance inside the workgroup).  ROT: every set's tables are in
// rotation mode (the host checked the meshes as for the adaptive rule),
// so the kernel carries the rank-one sums alone and fewer scalars; a set
// found otherwise gets NaN rather than a wrong sum.
template <bool ROT>
__global__ __launch_bounds__(256) void k_batch_sep(SepArgs a, int *__restrict__ queue, double *__restrict__ partial)
{
    __shared__ double red[2][4][64];
    __shared__ int next_item;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const SAS_CONST int *ipre = as_const(a.ipre);
    const int n_items = ipre[a.nsets];
    // the workgroup's cursor lives in wave 0; it starts on counter blockIdx % WORK_QUEUES
    QueueCursor cur{(int)(blockIdx.x % WORK_QUEUES), 0};
    int pos = 0;
    for (;;) {
        if (wave == 0) {
            const int it = queue_next(queue, n_items, cur);
            if (lane == 0) next_item = it;
        }
        __syncthreads();
        const int item = __builtin_amdgcn_readfirstlane(next_item);
        if (item < 0) break;
        // last position p with ipre[p] <= item: the previous item's set
        // when it still holds (items of one counter are consecutive); kept
        // wave-uniform so that every table read of the set is a scalar load
        pos = __builtin_amdgcn_readfirstlane(pos);
        if (item < ipre[pos] || item >= ipre[pos + 1]) {
            int lo = 0, hi = a.nsets;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (ipre[mid] <= item) lo = mid;
                else hi = mid;
            }
            pos = __builtin_amdgcn_readfirstlane(lo);
        }
        const int lo = pos;
        const int s = __builtin_amdgcn_readfirstlane(as_const(a.order)[lo]);
        const int rem = item - ipre[lo];
        const int tile = rem / a.nblk, blk = rem - tile * a.nblk;
        const int64_t q0 = a.qoff[s];
        const int nq = (int)(a.qoff[s + 1] - q0);
        const int iq = tile * QTILE + lane;
        const float q = iq < nq ? a.q32[a.qdoff[s] + iq] : 0.0f;
        const int n0 = blk * a.nodes_per_item, n1 = min(a.quad.n, n0 + a.nodes_per_item);
        KSum a1, a2;
        for (int node = n0 + wave; node < n1; node += 4) {
            float f1, f2;
            if constexpr (ROT) sep::node_sums_rot(a.t, s, q, a.quad.a[node], a.quad.b[node], f1, f2);
            else sep::node_sums(a.t, s, q, a.quad.a[node], a.quad.b[node], f1, f2);
            const float w = a.quad.w[node];
            a1.add(w * f1);
            a2.add(w * f2);
        }
        red[0][wave][lane] = a2.value();
        red[1][wave][lane] = a1.value();
        __syncthreads();
        if (wave < 2 && iq < nq) {
            // wave 0: <F^2>, wave 1: <F>, the four waves' sums in wave order
            double v = ((red[wave][0][lane] + red[wave][1][lane]) + red[wave][2][lane]) + red[wave][3][lane];
            if (ROT && as_const(a.t.dims)[4 * s + 2] != 2) v = __builtin_nan("");
            partial[a.poff[s] + (size_t)(blk * 2 + wave) * nq + iq] = v;
        }
        __syncthreads();
    }
}

// Separable cylinder sets under the adaptive rule (20/76/500/5000 nodes by
// q * max(radius, length/2), adaptive.c:11276-11298): items are (set, 64-q
// tile, rule, block of that rule's nodes), numbered tile by tile; titem[p]
// = (set, tile, k_lo | k_hi << 8, first item) lists the tiles in item order
// with the rules their q range can need.  Each lane counts, in FP64 as the
// reference chooses its rule, the grid prefixes below the rule's threshold
// and the previous one; a wave whose lanes all see the whole mesh under
// this rule (or none of it) takes the rotation-mode sums, one where a grid
// enters whole on every lane the band sums of the other grid, else the
// prefix sums.  One partial per (item, output, lane):
// partial[item*128 + o*64 + lane].
struct SepAdaptArgs {
    const float *q32;
    const double *q64;
    const int64_t *qoff;
    const int64_t *qdoff;
    sep::Tables t;
    Quad quad;
    const int4 *titem;
    int n_titem;
    int n_items;
    int nodes_per_item;
};

__device__ __forceinline__ int adaptive_n(int k) { return k == 0 ? 20 : k == 1 ? 76 : k == 2 ? 500 : 5000; }

__global__ __launch_bounds__(256) void k_batch_sep_adapt(SepAdaptArgs a, int *__restrict__ queue,
                                                         double *__restrict__ partial)
{
    __shared__ double red[2][4][64];
    __shared__ int next_item;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int4 *titem = a.titem;
    QueueCursor cur{(int)(blockIdx.x % WORK_QUEUES), 0};
    int pos = 0;
    for (;;) {
        if (wave == 0) {
            const int it = queue_next(queue, a.n_items, cur);
            if (lane == 0) next_item = it;
        }
        __syncthreads();
        const int item = __builtin_amdgcn_readfirstlane(next_item);
        if (item < 0) break;
        // the tile entry: last p with titem[p].w <= item
        pos = __builtin_amdgcn_readfirstlane(pos);
        if (!(titem[pos].w <= item && (pos + 1 == a.n_titem || item < titem[pos + 1].w))) {
            int lo = 0, hi = a.n_titem;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (titem[mid].w <= item) lo = mid;
                else hi = mid;
            }
            pos = __builtin_amdgcn_readfirstlane(lo);
        }
        const int4 te = titem[pos];
        const int s = te.x, tile = te.y, klo = te.z & 255, khi = te.z >> 8;
        const int P = a.nodes_per_item;
        int rem = item - te.w, k = klo;
        for (; k < khi; k++) {
            const int nb = (adaptive_n(k) + P - 1) / P;
            if (rem < nb) break;
            rem -= nb;
        }
        const int64_t q0 = a.qoff[s], qd0 = a.qdoff[s];
        const int nq = (int)(a.qoff[s + 1] - q0);
        const int iq = tile * QTILE + lane;
        const bool ok = iq < nq;
        const float q = ok ? a.q32[qd0 + iq] : 0.0f;
        const double qd = ok ? a.q64[qd0 + iq] : 0.0;
        const SAS_CONST int *dims = as_const(a.t.dims) + 4 * s;
        const int NR = dims[0], NL = dims[1];
        const double *G = a.t.G64 + (size_t)s * 2 * sep::NRMAX;
        // prefixes of the grids below this rule's threshold (ra, rb) and
        // below the previous rule's (pa, pb)
        int ra = NR, rb = NL, pa = 0, pb = 0;
        if (k < 3) {
            const double T = k == 0 ? 10.0 : k == 1 ? 100.0 : 750.0;
            ra = sep::rule_count(G, NR, qd, 1.0, T);
            rb = sep::rule_count(G + sep::NRMAX, NL, qd, 0.5, T);
        }
        if (k > 0) {
            const double T = k == 1 ? 10.0 : k == 2 ? 100.0 : 750.0;
            pa = sep::rule_count(G, NR, qd, 1.0, T);
            pb = sep::rule_count(G + sep::NRMAX, NL, qd, 0.5, T);
        }
        const bool none = !ok || (ra == pa && rb == pb);
        // the smaller rules' rectangle is empty where either prefix is
        const bool whole = ra == NR && rb == NL && (pa == 0 || pb == 0);
        const bool fast = sep::wave_all(whole || none);
        // one grid whole on every lane: the other grid's band alone is walked
        // (lanes with nothing to add get an empty band)
        const bool lband = !fast && sep::wave_all(none || (ra == NR && (pa == 0 || pa == NR || pb == 0)));
        const bool rband = !fast && !lband && sep::wave_all(none || (rb == NL && (pb == 0 || pb == NL || pa == 0)));
        const int blo = none ? sep::NRMAX : lband ? (pa == NR ? pb : 0) : (pb == NL ? pa : 0);
        const int bhi = none ? 0 : lband ? rb : ra;
        const int off = a.quad.off[k], n0 = rem * P, n1 = min(adaptive_n(k), n0 + P);
        KSum a1, a2;
        for (int node = n0 + wave; node < n1; node += 4) {
            float f1, f2;
            const float sn = a.quad.a[off + node], cs = a.quad.b[off + node];
            if (fast) sep::node_sums_rot(a.t, s, q, sn, cs, f1, f2);
            else if (lband) sep::band_sums<false>(a.t, s, q, sn, cs, blo, bhi, f1, f2);
            else if (rband) sep::band_sums<true>(a.t, s, q, sn, cs, blo, bhi, f1, f2);
            else sep::prefix_sums(a.t, s, q, sn, cs, ra, pa, rb, pb, f1, f2);
            if (none) f1 = f2 = 0.0f;
            const float w = a.quad.w[off + node];
            a1.add(w * f1);
            a2.add(w * f2);
        }
        red[0][wave][lane] = a2.value();
        red[1][wave][lane] = a1.value();
        __syncthreads();
        if (wave < 2) {
            // the tables of a mesh that is not whole on ascending grids cannot
            // be split by rule: NaN rather than a wrong sum
            double v = ((red[wave][0][lane] + red[wave][1][lane]) + red[wave][2][lane]) + red[wave][3][lane];
            if (dims[2] != 2) v = __builtin_nan("");
            partial[(size_t)item * 2 * QTILE + wave * QTILE + lane] = v;
        }
        __syncthreads();
    }
}

// P@S (sasmodels/product.py:425-528): S's parameters for set s, at the
// effective radius of the form (mode > 0) or the given one and at
// volfraction * form:shell volume ratio, from the staged structure
// parameters sp and the set's weight sums sm; total and shell with the
// guards of Kernel.Fq (total weight 0 -> 1, shell volume 0 -> 1).
__device__ __forceinline__ void structure_inputs(const double *sp, const double *sm, double *sq_pars, double &total,
                                                 double &shell)
{
    total = sm[0];
    if (total == 0.0) total = 1.0;
    shell = sm[2] / total;
    if (shell == 0.0) shell = 1.0;
    sq_pars[0] = sp[3] > 0.0 ? sm[3] / total : sp[0];
    sq_pars[1] = sp[1] * ((sm[1] / total) / shell);
    for (int k = 2; k < 6; k++) sq_pars[k] = sp[3 + k];
}

// One thread per set: the q-independent part of S (the Hayter MSA's
// coefficient iteration), before k_batch_final.
__global__ __launch_bounds__(64) void k_structure_prep(int nsets, int s_model, const double *__restrict__ spars,
                                                       const double *__restrict__ sums, double *__restrict__ sprep)
{
    const int s = blockIdx.x * 64 + threadIdx.x;
    if (s >= nsets) return;
    double sq_pars[6], total, shell;
    structure_inputs(spars + SASHIP_SPARS * s, sums + 4 * s, sq_pars, total, shell);
    structure_prepare(s_model, sq_pars, sprep + S_PREP * s);
}

// Per set: sum the chunk partials of every output in chunk order, add the
// weight sums, and form I(q) = scale * <F^2>/<V> + background with the
// guards of Kernel.Fq (total weight 0 -> 1, shell volume 0 -> 1); for P@S
// with structure factor SM (0: none), one instance per structure factor so
// that a plain batch carries none of their code and registers.
template <int SM>
__global__ __launch_bounds__(256) void k_batch_final(
    int nout, int nres, const int64_t *__restrict__ qoff, const int64_t *__restrict__ qdoff,
    const double *__restrict__ values,
    const int64_t *__restrict__ voff, const int64_t *__restrict__ poff,
    const int64_t *__restrict__ resoff, const int *__restrict__ nchunk,
    const int2 *__restrict__ tinfo, const int64_t *__restrict__ toff,
    const double *__restrict__ sums, const double *__restrict__ partial,
    const double *__restrict__ spars, const double *__restrict__ sprep,
    const double *__restrict__ q64, int is2d, double *__restrict__ result, double *__restrict__ iq_out)
{
    const int s = blockIdx.x;
    const int64_t q0 = qoff[s];
    const int nq = (int)(qoff[s + 1] - q0);
    const int nc = nchunk[s];
    const double *part = partial + poff[s];
    const double *sm = sums + 4 * s;
    double total = sm[0];
    if (total == 0.0) total = 1.0;
    double shell = sm[2] / total;
    if (shell == 0.0) shell = 1.0;
    const double scale = values[voff[s]], background = values[voff[s] + 1];
    double cscale = scale / shell;
    // P@S: S's parameters (structure_inputs); the number density is
    // volfraction / shell volume
    const double *sp = SM != 0 ? spars + SASHIP_SPARS * s : nullptr;
    double sq_pars[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};     // radius, volume fraction, further S parameters
    bool beta = false;
    if (sp) {
        structure_inputs(sp, sm, sq_pars, total, shell);
        beta = sp[2] > 0.0;
        if (sp[4] == 0.0) cscale *= sp[1];      // else the form carries its own volfraction
    }
    for (int e = blockIdx.y * 256 + threadIdx.x; e < nq; e += 256 * gridDim.y) {
        // chunk order is kept; the loads of 8 chunks are issued before their
        // adds so that their latencies overlap
        double f2 = 0.0, f1 = 0.0;
        if (tinfo) {
            // adaptive separable items: the tile's items in order, 2 x 64 partials each
            const int2 ti = tinfo[toff[s] + e / QTILE];
            const double *pt = partial + (size_t)ti.x * 2 * QTILE + e % QTILE;
            for (int c = 0; c < ti.y; c++) {
                f2 += pt[(size_t)c * 2 * QTILE];
                f1 += pt[(size_t)c * 2 * QTILE + QTILE];
            }
        } else {
        const size_t cstride = (size_t)nout * nq;
        int c = 0;
        for (; c + 8 <= nc; c += 8) {
            double a[8], b[8];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                a[k] = part[(size_t)(c + k) * cstride + e];
                b[k] = nout == 2 ? part[(size_t)(c + k) * cstride + nq + e] : 0.0;
            }
#pragma unroll
            for (int k = 0; k < 8; k++) {
                f2 += a[k];
                if (nout == 2) f1 += b[k];
            }
        }
        for (; c < nc; c++) {
            f2 += part[(size_t)c * cstride + e];
            if (nout == 2) f1 += part[(size_t)c * cstride + nq + e];
        }
        }
        if (result) {
            result[resoff[s] + (size_t)e * nres] = f2;
            if (nres == 2) result[resoff[s] + (size_t)e * nres + 1] = f1;
        }
        if (!iq_out) continue;
        double v = f2 / total;
        if (sp) {
            const int64_t g = qdoff[s] + e;
            const double q = is2d ? sqrt(q64[2 * g] * q64[2 * g] + q64[2 * g + 1] * q64[2 * g + 1]) : q64[g];
            const double S = structure_sq_prepared(SM, q, sq_pars, sprep + S_PREP * s);
            const double F1 = f1 / total;
            v = beta ? v + F1 * F1 * (S - 1.0) : v * S;
        }
        iq_out[q0 + e] = cscale * v + background;
    }
    if (result && blockIdx.y == 0 && threadIdx.x < 4)
        result[resoff[s] + (size_t)nres * nq + threadIdx.x] = sm[threadIdx.x];
}

// Resolution smearing of the batch's I(q): one thread per smeared point,
// out[o] = sum_k w[off[o] + k] * iq[from[o] + k*step[o]] for k < len[o]
// (FP64): contiguous bands of a 1D weight matrix (step 1) or the strided
// polar bins of a 2D pinhole (step = pixels of the set).
// Repeated bands (rep != NULL): the band is taken rep[o] times, each
// rstep[o] points further on, and the band sums combined with the weights
// rw[roff[o] ..] (the qy rows of a 2D slit).
struct SmearRepeat {
    const int32_t *rep;
    const int64_t *rstep, *roff;
    const double *rw;
};

__global__ __launch_bounds__(256) void k_batch_smear(int64_t n_out, const int64_t *__restrict__ from,
                                                     const int32_t *__restrict__ len,
                                                     const int64_t *__restrict__ step,
                                                     const int64_t *__restrict__ off,
                                                     const double *__restrict__ w, SmearRepeat r,
                                                     const double *__restrict__ iq, double *__restrict__ out)
{
    const int64_t o = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (o >= n_out) return;
    const double *src = iq + from[o];
    const double *wt = w + off[o];
    const int64_t st = step[o];
    const int n = len[o];
    if (!r.rep) {
        double acc = 0.0;
        for (int k = 0; k < n; k++) acc = fma(wt[k], src[k * st], acc);
        out[o] = acc;
        return;
    }
    const int64_t rs = r.rstep[o];
    const double *rw = r.rw + r.roff[o];
    double total = 0.0;
    for (int j = 0; j < r.rep[o]; j++) {
        const double *row = src + j * rs;
        double acc = 0.0;
        for (int k = 0; k < n; k++) acc = fma(wt[k], row[k * st], acc);
        total = fma(rw[j], acc, total);
    }
    out[o] = total;
}

}  // namespace sas

namespace batch_host {

// Host-side plan of one call: the staging layout and the work list.
struct Plan {
    int kind = 0, nout = 2, nres = 2, ndet = 0, rv = 1;   // partial outputs, result outputs
    int64_t n_values = 0, rec_total = 0, partial_total = 0, result_total = 0, nq_total = 0;
    int n_items = 0;
    int sep_nlp = 0;                  // separable path: bilinear table columns (0: generic path)
    int nodes_per_item = 0;           // separable path: Gauss nodes per work item
    size_t off_values = 0, off_voff = 0, off_det = 0, off_roff = 0, off_poff = 0, off_resoff = 0;
    size_t off_nchunk = 0, off_items = 0, off_nitems = 0, off_order = 0, off_ipre = 0, off_spars = 0, bytes = 0;
    bool structure = false;           // P@S: per-set structure-factor parameters staged at off_spars
    int nblk = 0;                     // separable path: node blocks per (set, tile)
    int max_nq = 0;                   // largest q count of a set
    double mean_pts = 0.0;            // generic path: mesh points per work item (upper bound)
    bool single_chunk = false;        // generic path: every set's mesh fits one work item
    bool sep_adapt = false;           // separable path under the adaptive rule (tiled partials)
    bool sep_rot = false;             // fixed-rule separable path, every set's mesh in rotation mode
    int n_titem = 0;                  // adaptive: tile entries (set, tile, rules, first item)
    size_t off_tinfo = 0, off_toff = 0;
};

}  // namespace batch_host

struct saship_batch {
    saship_model *m = nullptr;
    int nsets = 0, is2d = 0, dev = 0, cus = 256;
    hipStream_t st = nullptr;
    std::vector<int64_t> qoff;
    // per set, its first 64-q tile in tile_qmin/tile_qmax: the smallest and
    // largest q of each tile (the adaptive rule's range over a tile)
    std::vector<int64_t> tile_off;
    std::vector<double> tile_qmin, tile_qmax;
    float *q32 = nullptr;
    double *q64 = nullptr;
    int64_t *d_qoff = nullptr;
    int64_t *d_qdoff = nullptr;      // q value offset per set (all 0 when the sets share one q vector)
    // per-call buffers
    char *d_args = nullptr;
    size_t args_cap = 0;
    char *h_stage = nullptr;
    size_t stage_cap = 0;
    float4 *d_recs = nullptr;
    size_t recs_cap = 0;
    int *d_npts = nullptr;
    double *d_sums = nullptr;
    float *d_spin = nullptr;          // per-set spin constants of the magnetic kernel
    double *d_sprep = nullptr;        // P@S: the q-independent part of S per set (S_PREP doubles)
    int *d_queue = nullptr;
    double *d_partial = nullptr;
    size_t partial_cap = 0;
    double *d_out = nullptr;          // raw results then I(q)
    size_t out_cap = 0;
    char *d_sep = nullptr;            // separable cylinder tables
    size_t sep_cap = 0;
    // resolution smearing of I(q) (saship_batch_set_resolution)
    int64_t n_smear = 0;              // smeared points, all sets (0: no smearing)
    int64_t *d_sfrom = nullptr, *d_soff = nullptr, *d_sstep = nullptr;
    int32_t *d_slen = nullptr;
    double *d_sw = nullptr;
    int32_t *d_srep = nullptr;        // repeated bands (NULL: none)
    int64_t *d_srstep = nullptr, *d_sroff = nullptr;
    double *d_srw = nullptr;
    bool separable = true;            // use the separable path when every set allows it
    // P@S (saship_batch_set_structure): structure-factor model (0: none) and
    // per set (radius_effective, volfraction, structure_factor_mode, radius_effective_mode,
    // volfraction in the form factor, the structure factor's third and fourth parameter)
    int s_model = 0;
    std::vector<double> spars;
    char *h_out = nullptr;
    size_t hout_cap = 0;
    // launch shape
    int pts_per_item = 16;
    int waves_per_cu = 32;
    int grab = 0;                     // items per queue atomic (0: from the plan's mean item size)
    // graph replay
    bool use_graph = false;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    std::vector<int64_t> gkey;
    // the last successful call, for device-resident replays
    bool have_last = false;
    double last_cutoff = 0.0;
    int last_reff = 0;
    bool last_result = false, last_iq = false;
    batch_host::Plan last;
    // timing of the evaluation kernel
    hipEvent_t ev0 = nullptr, ev1 = nullptr;     // around the evaluation kernel
    hipEvent_t ev2 = nullptr, ev3 = nullptr;     // around the whole device sequence
    // large I(q) downloads go in chunks after the launch, each copied out
    // to the caller while the next one is in flight
    bool defer_iq = false;
    // the caller's I(q) array when it is pinned host memory: the last kernel
    // of the call writes I(q) there directly (no download, no host copy)
    double *iq_direct = nullptr;
    std::vector<hipEvent_t> chunk_ev;
    bool timing = false;
    double t_ms = 0.0, t_all_ms = 0.0;
    int ncalls = 0;
    std::mutex mu;
};

namespace {

using batch_host::Plan;

void drop_graph(saship_batch *b)
{
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
    b->gkey.clear();
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }


template <class M>
int rv_of(int kind) { return kind == 0 ? sas::rv1<M>() : sas::rv2<M>(); }

int record_width(int id, int kind)
{
    switch (id) {
    case 0: return rv_of<sas::Sphere>(kind);
    case 1: return rv_of<sas::Cylinder>(kind);
    case 2: return rv_of<sas::Ellipsoid>(kind);
    case 3: return rv_of<sas::CoreShellCylinder>(kind);
    case 4: return rv_of<sas::CoreShellSphere>(kind);
    case 6: return rv_of<sas::Parallelepiped>(kind);
    case 7: return rv_of<sas::TriaxialEllipsoid>(kind);
    case 8: return rv_of<sas::BccParacrystal>(kind);
    case 9: return rv_of<sas::FccParacrystal>(kind);
    case 10: return rv_of<sas::ScParacrystal>(kind);
    case 11: return rv_of<sas::HollowCylinder>(kind);
    case 12: return rv_of<sas::Vesicle>(kind);
    case 17: return rv_of<sas::CoreMultiShell>(kind);
    case 18: return rv_of<sas::Onion>(kind);
    case 19: return rv_of<sas::CoreShellEllipsoid>(kind);
    case 20: return rv_of<sas::EllipticalCylinder>(kind);
    case 21: return rv_of<sas::CoreShellBicelle>(kind);
    case 22: return rv_of<sas::Guinier>(kind);
    case 23: return rv_of<sas::GaussianPeak>(kind);
    case 24: return rv_of<sas::Lorentz>(kind);
    case 25: return rv_of<sas::Dab>(kind);
    case 26: return rv_of<sas::GelFit>(kind);
    case 27: return rv_of<sas::MassFractal>(kind);
    case 28: return rv_of<sas::SurfaceFractal>(kind);
    case 29: return rv_of<sas::MassSurfaceFractal>(kind);
    case 30: return rv_of<sas::StarPolymer>(kind);
    case 31: return rv_of<sas::MonoGaussCoil>(kind);
    case 32: return rv_of<sas::Fractal>(kind);
    case 33: return rv_of<sas::FractalCoreShell>(kind);
    case 34: return rv_of<sas::LamellarHg>(kind);
    case 35: return rv_of<sas::LamellarStackCaille>(kind);
    case 36: return rv_of<sas::LamellarHgStackCaille>(kind);
    case 37: return rv_of<sas::LamellarStackParacrystal>(kind);
    case 38: return rv_of<sas::LinearPearls>(kind);
    case 39: return rv_of<sas::PolymerMicelle>(kind);
    case 40: return rv_of<sas::Raspberry>(kind);
    default: return rv_of<sas::FuzzySphere>(kind);
    }
}

// Validate every set, lay out the staging buffer and fill it (host pinned).
int plan_and_stage(saship_batch *b, int kind, const int32_t *details, int32_t n_details,
                   const double *values, const int64_t *voff_in, double cutoff, Plan &p)
{
    const int id = b->m->id, ns = b->nsets;
    p.kind = kind;
    p.nout = kind == 0 ? 2 : 1;
    p.nres = kind == 0 && model_has_fq(id) ? 2 : 1;    // models without <F>: <F^2> only
    p.ndet = n_details;
    p.rv = record_width(id, kind);
    p.n_values = voff_in[ns];
    if (voff_in[0] != 0) return fail(SASHIP_EINVAL, "values_offset[0] must be 0");
    std::vector<int64_t> roff(ns), poff(ns), resoff(ns);
    std::vector<int32_t> nchunk(ns);
    int64_t items = 0;
    double pts_total = 0.0;
    const int P = b->pts_per_item;
    // separable cylinder path when every set's mesh allows it
    if (id == 1 && kind == 0 && b->m->quad_n > 0 && b->separable) {
        int nlp = 0;
        for (int s = 0; s < ns && (s == 0 || nlp); s++) {
            const int here = (int)n_details == model_maxpd(1) * 4 + 4 ? cylinder_separable(details + (size_t)s * n_details) : 0;
            nlp = here ? std::max(nlp, here) : 0;
        }
        p.sep_nlp = nlp;
    }
    // the adaptive rule: separable when every set's mesh is whole on ascending,
    // equally spaced grids (adaptive_separable); ext: per set the smallest and
    // largest radius and half length
    std::vector<double> ext;
    if (id == 1 && kind == 0 && b->m->quad_n == 0 && b->separable && n_details == model_maxpd(1) * 4 + 4) {
        int nlp = 0;
        ext.resize(4 * (size_t)ns);
        for (int s = 0; s < ns; s++) {
            const int32_t *det = details + (size_t)s * n_details;
            const int64_t nv = voff_in[s + 1] - voff_in[s];
            const int here = cylinder_separable(det);
            if (!here || nv < 0 || validate_set(id, kind, det, n_details, nv, 0, -1, s) != 0 ||
                !adaptive_separable(det, values + voff_in[s], cutoff, &ext[4 * (size_t)s])) {
                nlp = 0;
                last_error().clear();
                break;
            }
            nlp = std::max(nlp, here);
        }
        if (nlp) {
            p.sep_nlp = nlp;
            p.sep_adapt = true;
        }
    }
    const int nnode = p.sep_adapt ? 0 : b->m->quad.n;
    bool rot = p.sep_nlp && !p.sep_adapt;
    if (p.sep_adapt) p.nodes_per_item = P;
    else if (p.sep_nlp) p.nodes_per_item = std::min(P, nnode);
    const int nblk = p.sep_nlp ? (nnode + p.nodes_per_item - 1) / p.nodes_per_item : 0;
    p.nblk = nblk;
    for (int s = 0; s < ns; s++) {
        const int64_t nv = voff_in[s + 1] - voff_in[s];
        const int32_t *det = details + (size_t)s * n_details;
        if (nv < 0) return fail(SASHIP_EINVAL, set_prefix(s) + "values_offset must be non-decreasing");
        if (int e = validate_set(id, kind, det, n_details, nv, 0, -1, s)) return e;
        // fixed rule: every mesh whole on ascending, equally spaced grids (as
        // for the adaptive rule) puts every set's tables in rotation mode
        if (rot) {
            double ext[4];
            rot = adaptive_separable(det, values + voff_in[s], cutoff, ext);
        }
        const int64_t num_eval = det[n_details - 4];
        const int nq = (int)(b->qoff[s + 1] - b->qoff[s]);
        p.max_nq = std::max(p.max_nq, nq);
        roff[s] = p.rec_total;
        if (!p.sep_nlp) p.rec_total += num_eval;
        nchunk[s] = p.sep_nlp ? nblk : (int32_t)std::max<int64_t>(1, (num_eval + P - 1) / P);
        poff[s] = p.partial_total;
        p.partial_total += (int64_t)nchunk[s] * p.nout * nq;
        resoff[s] = p.result_total;
        p.result_total += (int64_t)p.nres * nq + 4;
        items += (int64_t)nchunk[s] * ((nq + sas::QTILE - 1) / sas::QTILE);
        pts_total += (double)std::max<int64_t>(num_eval, 1) * ((nq + sas::QTILE - 1) / sas::QTILE);
    }
    std::vector<int4> titem;
    std::vector<int2> tinfo;
    if (p.sep_adapt) {
        // tiles in set cost order; per tile the rules its q range can need
        // and their node blocks, numbered tile by tile
        std::vector<int> ord(ns);
        for (int s = 0; s < ns; s++) ord[s] = s;
        auto cost = [&](int s) { return details[(size_t)s * n_details + n_details - 4]; };
        std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return cost(x) > cost(y); });
        tinfo.assign((size_t)b->tile_off[ns], make_int2(0, 0));
        int64_t n = 0;
        for (int s : ord) {
            const double *e = &ext[4 * (size_t)s];
            for (int64_t t = b->tile_off[s]; t < b->tile_off[s + 1]; t++) {
                const int klo = adaptive_rule(b->tile_qmin[t] * std::max(e[0], e[2]));
                const int khi = adaptive_rule(b->tile_qmax[t] * std::max(e[1], e[3]));
                int count = 0;
                for (int k = klo; k <= khi; k++) count += (adaptive_nodes(k) + P - 1) / P;
                titem.push_back(make_int4(s, (int)(t - b->tile_off[s]), klo | (khi << 8), (int)n));
                tinfo[t] = make_int2((int)n, count);
                n += count;
                if (n > 0x7fffffff / (2 * sas::QTILE)) return fail(SASHIP_EINVAL, "too many work items in one batch");
            }
        }
        items = n;
        p.partial_total = n * 2 * sas::QTILE;
        p.n_titem = (int)titem.size();
        for (int s = 0; s < ns; s++) poff[s] = 0;
    }
    p.mean_pts = items > 0 ? pts_total / (double)items : 0.0;
    p.single_chunk = !p.sep_nlp && std::all_of(nchunk.begin(), nchunk.end(), [](int32_t c) { return c == 1; });
    p.sep_rot = rot;
    if (items > 0x7fffffff) return fail(SASHIP_EINVAL, "too many work items in one batch");
    p.n_items = (int)items;
    p.nq_total = b->qoff[ns];
    size_t off = 0;
    p.off_values = off; off = align256(off + sizeof(double) * p.n_values);
    p.off_voff = off; off = align256(off + sizeof(int64_t) * ns);
    p.off_det = off; off = align256(off + sizeof(int32_t) * (size_t)ns * n_details);
    p.off_roff = off; off = align256(off + sizeof(int64_t) * ns);
    p.off_poff = off; off = align256(off + sizeof(int64_t) * ns);
    p.off_resoff = off; off = align256(off + sizeof(int64_t) * ns);
    p.off_nchunk = off; off = align256(off + sizeof(int32_t) * ns);
    p.off_items = off;
    off = align256(off + sizeof(int4) * (size_t)(p.sep_adapt ? std::max(1, p.n_titem) : p.sep_nlp ? 1 : std::max(1, p.n_items)));
    if (p.sep_adapt) {
        p.off_tinfo = off; off = align256(off + sizeof(int2) * (size_t)std::max<int64_t>(1, b->tile_off[ns]));
        p.off_toff = off; off = align256(off + sizeof(int64_t) * ns);
    }
    p.off_nitems = off; off = align256(off + sizeof(int32_t));
    p.off_order = off; off = align256(off + sizeof(int32_t) * ns);
    p.off_ipre = off; off = align256(off + sizeof(int32_t) * (ns + 1));
    p.structure = b->s_model != 0;
    if (p.structure) {
        for (int s = 0; s < ns; s++)
            if (b->spars[SASHIP_SPARS * s + 2] > 0.0 && kind != 0)
                return fail(SASHIP_EINVAL, "set " + std::to_string(s) + ": the beta approximation needs 1D q");
        p.off_spars = off;
        off = align256(off + sizeof(double) * SASHIP_SPARS * ns);
    }
    p.bytes = off;
    if (int e = ensure_host(b->h_stage, b->stage_cap, p.bytes)) return e;
    char *h = b->h_stage;
    memcpy(h + p.off_values, values, sizeof(double) * p.n_values);
    memcpy(h + p.off_voff, voff_in, sizeof(int64_t) * ns);
    memcpy(h + p.off_det, details, sizeof(int32_t) * (size_t)ns * n_details);
    memcpy(h + p.off_roff, roff.data(), sizeof(int64_t) * ns);
    memcpy(h + p.off_poff, poff.data(), sizeof(int64_t) * ns);
    memcpy(h + p.off_resoff, resoff.data(), sizeof(int64_t) * ns);
    memcpy(h + p.off_nchunk, nchunk.data(), sizeof(int32_t) * ns);
    if (p.structure) memcpy(h + p.off_spars, b->spars.data(), sizeof(double) * SASHIP_SPARS * ns);
    if (p.sep_adapt) {
        memcpy(h + p.off_items, titem.data(), sizeof(int4) * titem.size());
        memcpy(h + p.off_tinfo, tinfo.data(), sizeof(int2) * tinfo.size());
        memcpy(h + p.off_toff, b->tile_off.data(), sizeof(int64_t) * ns);
    }
    // work list, longest items first: items are bucketed by the number of
    // mesh points their chunk can hold (an upper bound before pruning)
    int4 *it = (int4 *)(h + p.off_items);
    int64_t k = 0;
    if (p.sep_nlp) {
        // costliest sets first (mesh size as the proxy); the kernel numbers
        // the items of each set (tiles, then node blocks) from ipre
        int32_t *order = (int32_t *)(h + p.off_order), *ipre = (int32_t *)(h + p.off_ipre);
        for (int s = 0; s < ns; s++) order[s] = s;
        auto cost = [&](int s) { return details[(size_t)s * n_details + n_details - 4]; };
        std::stable_sort(order, order + ns, [&](int x, int y) { return cost(x) > cost(y); });
        ipre[0] = 0;
        for (int i = 0; i < ns; i++) {
            const int s = order[i];
            const int nq = (int)(b->qoff[s + 1] - b->qoff[s]);
            ipre[i + 1] = ipre[i] + ((nq + sas::QTILE - 1) / sas::QTILE) * nblk;
        }
    }
    // Counter c of the device queue hands out the slice
    // [c n / WORK_QUEUES, (c + 1) n / WORK_QUEUES) of the list in order, so the
    // longest-first order is dealt round-robin over the slices: every counter
    // starts on long items and ends on short ones, and waves that move to
    // another counter after theirs runs dry find short items there.
    int64_t slot_next[sas::WORK_QUEUES], slot_end[sas::WORK_QUEUES];
    for (int c = 0; c < sas::WORK_QUEUES; c++) {
        slot_next[c] = (int64_t)c * p.n_items / sas::WORK_QUEUES;
        slot_end[c] = (int64_t)(c + 1) * p.n_items / sas::WORK_QUEUES;
    }
    int deal = 0;
    for (int fill = p.sep_nlp ? 0 : P; fill >= 1; fill--) {
        for (int s = 0; s < ns; s++) {
            const int64_t num_eval = details[(size_t)s * n_details + n_details - 4];
            const int nq = (int)(b->qoff[s + 1] - b->qoff[s]);
            const int ntile = (nq + sas::QTILE - 1) / sas::QTILE;
            for (int c = 0; c < nchunk[s]; c++) {
                const int64_t here = std::min<int64_t>(P, num_eval - (int64_t)c * P);
                if (std::max<int64_t>(1, here) != fill) continue;
                for (int t = 0; t < ntile; t++) {
                    while (slot_next[deal] == slot_end[deal]) deal = (deal + 1) % sas::WORK_QUEUES;
                    it[slot_next[deal]++] = make_int4(s, t, c, 0);
                    deal = (deal + 1) % sas::WORK_QUEUES;
                    k++;
                }
            }
        }
    }
    if (k != p.n_items && !p.sep_nlp) return fail(SASHIP_EINVAL, "internal: work list size mismatch");
    *(int32_t *)(h + p.off_nitems) = p.n_items;
    return 0;
}

// Separable cylinder path: tables per set, then the node-block evaluation.
// Items per queue atomic: the batch's setting, else enough items to reach
// about 8 mesh points per atomic (at most 8 items) while every wave of the
// grid still gets one grab of its own.
int queue_grab(const saship_batch *b, const Plan &p, int nwaves)
{
    if (b->grab > 0) return b->grab;
    if (p.mean_pts >= 8.0 || p.mean_pts <= 0.0) return 1;
    const int want = std::min(8, (int)std::ceil(8.0 / p.mean_pts));
    const int room = std::max(1, p.n_items / std::max(1, nwaves));
    return std::max(1, std::min(want, room));
}

int enqueue_sep(saship_batch *b, const Plan &p, const sas::ExpandArgs &ea, const int64_t *d_poff)
{
    using namespace sas;
    const sep::Tables t = sep::carve(b->d_sep, b->nsets, p.sep_nlp);
    sep::k_sep_expand<1><<<b->nsets, EXPAND_THREADS, 0, b->st>>>(ea, t);
    HIP_TRY(hipGetLastError());
    const char *d = b->d_args;
    if (p.sep_adapt) {
        const SepAdaptArgs sa{b->q32, b->q64, b->d_qoff, b->d_qdoff, t, b->m->quad, (const int4 *)(d + p.off_items),
                              p.n_titem, p.n_items, p.nodes_per_item};
        const int grid = std::max(1, b->cus * b->waves_per_cu / 4);
        if (b->timing) HIP_TRY(hipEventRecord(b->ev0, b->st));
        k_batch_sep_adapt<<<grid, 256, 0, b->st>>>(sa, b->d_queue, b->d_partial);
        HIP_TRY(hipGetLastError());
        if (b->timing) HIP_TRY(hipEventRecord(b->ev1, b->st));
        return 0;
    }
    const SepArgs sa{b->q32, b->d_qoff, b->d_qdoff, t, b->m->quad, (const int *)(d + p.off_order), (const int *)(d + p.off_ipre),
                     b->nsets, p.nblk, p.nodes_per_item, d_poff};
    const int grid = std::max(1, b->cus * b->waves_per_cu / 4);
    if (b->timing) HIP_TRY(hipEventRecord(b->ev0, b->st));
    if (p.sep_rot) k_batch_sep<true><<<grid, 256, 0, b->st>>>(sa, b->d_queue, b->d_partial);
    else k_batch_sep<false><<<grid, 256, 0, b->st>>>(sa, b->d_queue, b->d_partial);
    HIP_TRY(hipGetLastError());
    if (b->timing) HIP_TRY(hipEventRecord(b->ev1, b->st));
    return 0;
}

// The device work of one call; with `copies` the staging upload before it
// and the result downloads after it (without: inputs and outputs stay in HBM).
// parts: the head (upload, expansion, evaluation) and/or the tail (structure
// factor, final sums, smearing, downloads).
constexpr int PART_HEAD = 1, PART_TAIL = 2;
template <class M>
int enqueue(saship_batch *b, const Plan &p, double cutoff, int reff_mode, bool want_result, bool want_iq,
            bool copies, int parts)
{
    using namespace sas;
    const char *d = b->d_args;
    const double *d_values = (const double *)(d + p.off_values);
    const int64_t *d_voff = (const int64_t *)(d + p.off_voff);
    const int32_t *d_det = (const int32_t *)(d + p.off_det);
    const int64_t *d_roff = (const int64_t *)(d + p.off_roff);
    const int64_t *d_poff = (const int64_t *)(d + p.off_poff);
    const int64_t *d_resoff = (const int64_t *)(d + p.off_resoff);
    const int32_t *d_nchunk = (const int32_t *)(d + p.off_nchunk);
    const int4 *d_items = (const int4 *)(d + p.off_items);
    const int *d_nitems = (const int *)(d + p.off_nitems);
    double *d_result = b->d_out;
    double *d_iq = b->d_out + p.result_total;

    // every set one chunk and plain I(q) asked: the evaluation kernel writes
    // I(q) itself (no partial sums, no final kernel); it then belongs to the
    // tail, whose output address may change from call to call
    const bool deval = p.single_chunk && !p.structure && want_iq && !want_result;
    double *direct = copies ? b->iq_direct : nullptr;
    auto launch_eval = [&](double *iq_out) -> int {
        const int grid = std::max(1, b->cus * b->waves_per_cu / 4);
        BatchArgs ba{b->q32, b->q64, b->d_qoff, b->d_qdoff, d_values, d_voff, b->d_recs, d_roff, d_poff, b->d_npts, b->d_spin,
                     d_items, d_nitems, b->pts_per_item, b->m->quad, queue_grab(b, p, grid * 4), iq_out, b->d_sums};
        if (b->timing) HIP_TRY(hipEventRecord(b->ev0, b->st));
        if (p.kind == 0) {
            if (b->m->quad_n > 0) k_batch_eval_1d<M, true><<<grid, 256, 0, b->st>>>(ba, b->d_queue, b->d_partial);
            else k_batch_eval_1d<M, false><<<grid, 256, 0, b->st>>>(ba, b->d_queue, b->d_partial);
        } else if (p.kind == 1) {
            k_batch_eval_2d<M, false><<<grid, 256, 0, b->st>>>(ba, b->d_queue, b->d_partial);
        } else {
            k_batch_eval_2d<M, true><<<grid, 256, 0, b->st>>>(ba, b->d_queue, b->d_partial);
        }
        HIP_TRY(hipGetLastError());
        if (b->timing) HIP_TRY(hipEventRecord(b->ev1, b->st));
        return 0;
    };

    if (parts & PART_HEAD) {
        if (copies) HIP_TRY(hipMemcpyAsync(b->d_args, b->h_stage, p.bytes, hipMemcpyHostToDevice, b->st));
        if (b->timing) HIP_TRY(hipEventRecord(b->ev2, b->st));
        ExpandArgs ea{d_det, d_values, d_voff, d_roff, p.ndet, 0, -1, cutoff, reff_mode,
                      b->d_recs, b->d_npts, b->d_sums, b->d_queue, b->d_spin};
        if (p.sep_nlp) {
            if (int e = enqueue_sep(b, p, ea, d_poff)) return e;
        } else {
            switch (p.kind) {
            case 0: k_expand<M, 0><<<b->nsets, EXPAND_THREADS, 0, b->st>>>(ea); break;
            case 1: k_expand<M, 1><<<b->nsets, EXPAND_THREADS, 0, b->st>>>(ea); break;
            default: k_expand<M, 2><<<b->nsets, EXPAND_THREADS, 0, b->st>>>(ea); break;
            }
            HIP_TRY(hipGetLastError());
            if (!deval)
                if (int e = launch_eval(nullptr)) return e;
        }
    }
    if (!(parts & PART_TAIL)) return 0;
    double *d_siq = d_iq + p.nq_total;
    if (deval) {
        if (int e = launch_eval(direct && b->n_smear == 0 ? direct : d_iq)) return e;
    } else {
    const double *d_spars = p.structure ? (const double *)(d + p.off_spars) : nullptr;
    if (p.structure) {
        k_structure_prep<<<(b->nsets + 63) / 64, 64, 0, b->st>>>(b->nsets, b->s_model, d_spars, b->d_sums, b->d_sprep);
        HIP_TRY(hipGetLastError());
    }
    const dim3 fgrid(b->nsets, std::max(1, (p.max_nq + 255) / 256));
    const int2 *d_tinfo = p.sep_adapt ? (const int2 *)(d + p.off_tinfo) : nullptr;
    const int64_t *d_toff = p.sep_adapt ? (const int64_t *)(d + p.off_toff) : nullptr;
    double *f_result = want_result ? d_result : nullptr;
    double *f_iq = !want_iq ? nullptr : direct && b->n_smear == 0 ? direct : d_iq;
#define SAS_FINAL(SM)                                                                                        \
    k_batch_final<SM><<<fgrid, 256, 0, b->st>>>(p.nout, p.nres, b->d_qoff, b->d_qdoff, d_values, d_voff, d_poff, d_resoff, \
                                                d_nchunk, d_tinfo, d_toff, b->d_sums, b->d_partial, d_spars,      \
                                                b->d_sprep, b->q64, b->is2d, f_result, f_iq)
    switch (p.structure ? b->s_model : 0) {
    case 0: SAS_FINAL(0); break;
    case S_HARDSPHERE: SAS_FINAL(S_HARDSPHERE); break;
    case S_STICKY: SAS_FINAL(S_STICKY); break;
    case S_SQUAREWELL: SAS_FINAL(S_SQUAREWELL); break;
    case S_HAYTER: SAS_FINAL(S_HAYTER); break;
    default: return fail(SASHIP_EINVAL, "unknown structure factor " + std::to_string(b->s_model));
    }
#undef SAS_FINAL
    HIP_TRY(hipGetLastError());
    }
    if (want_iq && b->n_smear > 0) {
        const int64_t blocks = (b->n_smear + 255) / 256;
        const SmearRepeat rep{b->d_srep, b->d_srstep, b->d_sroff, b->d_srw};
        k_batch_smear<<<(unsigned)blocks, 256, 0, b->st>>>(b->n_smear, b->d_sfrom, b->d_slen, b->d_sstep, b->d_soff,
                                                          b->d_sw, rep, d_iq, direct ? direct : d_siq);
        HIP_TRY(hipGetLastError());
    }
    if (b->timing) HIP_TRY(hipEventRecord(b->ev3, b->st));
    if (!copies) return 0;
    if (want_result)
        HIP_TRY(hipMemcpyAsync(b->h_out, d_result, sizeof(double) * p.result_total, hipMemcpyDeviceToHost, b->st));
    if (want_iq && !b->defer_iq && !direct)
        HIP_TRY(hipMemcpyAsync(b->h_out + sizeof(double) * p.result_total, b->n_smear > 0 ? d_siq : d_iq,
                               sizeof(double) * (b->n_smear > 0 ? b->n_smear : p.nq_total),
                               hipMemcpyDeviceToHost, b->st));
    return 0;
}

int enqueue_model(saship_batch *b, const Plan &p, double cutoff, int reff_mode, bool want_result, bool want_iq,
                  bool copies = true, int parts = PART_HEAD | PART_TAIL)
{
    switch (b->m->id) {
    case 0: return enqueue<sas::Sphere>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    case 1: return enqueue<sas::Cylinder>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    case 2: return enqueue<sas::Ellipsoid>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    case 3: return enqueue<sas::CoreShellCylinder>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    case 4: return enqueue<sas::CoreShellSphere>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    case 6: return enqueue<sas::Parallelepiped>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    case 7: return enqueue<sas::TriaxialEllipsoid>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    case 8: return enqueue<sas::BccParacrystal>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    case 9: return enqueue<sas::FccParacrystal>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    case 10: return enqueue<sas::ScParacrystal>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    case 11: return enqueue<sas::HollowCylinder>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    case 12: return enqueue<sas::Vesicle>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    case 17: return enqueue<sas::CoreMultiShell>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    case 18: return enqueue<sas::Onion>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    case 19: return enqueue<sas::CoreShellEllipsoid>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    case 20: return enqueue<sas::EllipticalCylinder>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    case 21: return enqueue<sas::CoreShellBicelle>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    case 22: return enqueue<sas::Guinier>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    case 23: return enqueue<sas::GaussianPeak>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    case 24: return enqueue<sas::Lorentz>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    case 25: return enqueue<sas::Dab>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    case 26: return enqueue<sas::GelFit>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    case 27: return enqueue<sas::MassFractal>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    case 28: return enqueue<sas::SurfaceFractal>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    case 29: return enqueue<sas::MassSurfaceFractal>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    case 30: return enqueue<sas::StarPolymer>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    case 31: return enqueue<sas::MonoGaussCoil>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    case 32: return enqueue<sas::Fractal>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    case 33: return enqueue<sas::FractalCoreShell>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    case 34: return enqueue<sas::LamellarHg>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    case 35: return enqueue<sas::LamellarStackCaille>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    case 36: return enqueue<sas::LamellarHgStackCaille>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    case 37: return enqueue<sas::LamellarStackParacrystal>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    case 38: return enqueue<sas::LinearPearls>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    case 39: return enqueue<sas::PolymerMicelle>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    case 40: return enqueue<sas::Raspberry>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    default: return enqueue<sas::FuzzySphere>(b, p, cutoff, reff_mode, want_result, want_iq, copies, parts);
    }
}

// Make every device buffer large enough for plan p (drops a stale graph).
int size_buffers(saship_batch *b, const Plan &p)
{
    bool grew = false;
    auto grow = [&](auto &ptr, size_t &cap, size_t bytes) -> int {
        if (bytes <= cap && ptr) return 0;
        grew = true;
        return ensure(ptr, cap, bytes);
    };
    if (int e = grow(b->d_args, b->args_cap, p.bytes)) return e;
    if (int e = grow(b->d_recs, b->recs_cap, sizeof(float4) * (size_t)std::max<int64_t>(1, p.rec_total) * p.rv)) return e;
    if (int e = grow(b->d_partial, b->partial_cap, sizeof(double) * (size_t)std::max<int64_t>(1, p.partial_total))) return e;
    if (int e = grow(b->d_out, b->out_cap, sizeof(double) * (size_t)(p.result_total + p.nq_total + b->n_smear))) return e;
    if (p.sep_nlp)
        if (int e = grow(b->d_sep, b->sep_cap, sas::sep::tables_bytes(b->nsets, p.sep_nlp))) return e;
    const size_t hb = sizeof(double) * (size_t)(p.result_total + std::max<int64_t>(p.nq_total, b->n_smear));
    if (hb > b->hout_cap || !b->h_out) {
        grew = true;
        if (int e = ensure_host(b->h_out, b->hout_cap, hb)) return e;
    }
    if (grew) drop_graph(b);
    return 0;
}

int run(saship_batch *b, int kind, const int32_t *details, int32_t n_details, const double *values,
        const int64_t *voff, double cutoff, int reff_mode, double *result, double *iq)
{
    Plan p;
    b->have_last = false;
    if (int e = plan_and_stage(b, kind, details, n_details, values, voff, cutoff, p)) return e;
    if (int e = size_buffers(b, p)) return e;
    const bool want_result = result != nullptr, want_iq = iq != nullptr;
    const size_t iq_bytes = want_iq ? sizeof(double) * (size_t)(b->n_smear > 0 ? b->n_smear : p.nq_total) : 0;
    constexpr size_t DEFER_IQ_FROM = (size_t)8 << 20, IQ_CHUNK = (size_t)4 << 20;
    // pinned output: the final (or smearing) kernel writes it; that tail is
    // launched eagerly (after the graph of the head) since the output's
    // address changes from call to call
    b->iq_direct = want_iq ? (double *)device_alias(iq) : nullptr;
    b->defer_iq = !b->iq_direct && iq_bytes >= DEFER_IQ_FROM;
    const int gparts = b->iq_direct ? PART_HEAD : PART_HEAD | PART_TAIL;
    if (b->use_graph && !b->timing) {
        std::vector<int64_t> key = {kind, (int64_t)p.bytes, p.rec_total, p.partial_total, p.result_total,
                                    reff_mode, want_result, want_iq, b->pts_per_item, b->waves_per_cu,
                                    p.sep_nlp, b->n_smear, b->s_model, b->defer_iq, gparts, p.single_chunk, p.sep_adapt, p.n_titem,
                                    p.sep_rot};
        int64_t cbits;
        memcpy(&cbits, &cutoff, sizeof cbits);
        key.push_back(cbits);
        if (!b->gexec || key != b->gkey) {
            drop_graph(b);
            HIP_TRY(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
            const int rc = enqueue_model(b, p, cutoff, reff_mode, want_result, want_iq, true, gparts);
            hipGraph_t g = nullptr;
            const hipError_t ce = hipStreamEndCapture(b->st, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (ce != hipSuccess) return fail(SASHIP_EHIP, std::string("graph capture: ") + hipGetErrorString(ce));
            b->graph = g;
            HIP_TRY(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
            b->gkey = key;
        }
        HIP_TRY(hipGraphLaunch(b->gexec, b->st));
        if (gparts == PART_HEAD)
            if (int e = enqueue_model(b, p, cutoff, reff_mode, want_result, want_iq, true, PART_TAIL)) return e;
    } else {
        if (int e = enqueue_model(b, p, cutoff, reff_mode, want_result, want_iq)) return e;
    }
    const size_t nchunk = b->defer_iq ? (iq_bytes + IQ_CHUNK - 1) / IQ_CHUNK : 0;
    if (nchunk) {
        const char *src = (const char *)(b->d_out + p.result_total + (b->n_smear > 0 ? p.nq_total : 0));
        char *stage = b->h_out + sizeof(double) * p.result_total;
        while (b->chunk_ev.size() < nchunk) {
            hipEvent_t ev = nullptr;
            HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
            b->chunk_ev.push_back(ev);
        }
        for (size_t c = 0; c < nchunk; c++) {
            const size_t off = c * IQ_CHUNK, len = std::min(IQ_CHUNK, iq_bytes - off);
            HIP_TRY(hipMemcpyAsync(stage + off, src + off, len, hipMemcpyDeviceToHost, b->st));
            HIP_TRY(hipEventRecord(b->chunk_ev[c], b->st));
        }
        for (size_t c = 0; c < nchunk; c++) {
            const size_t off = c * IQ_CHUNK, len = std::min(IQ_CHUNK, iq_bytes - off);
            HIP_TRY(event_wait(b->chunk_ev[c]));
            copy_out((char *)iq + off, stage + off, len);
        }
    }
    HIP_TRY(stream_wait(b->st));
    b->last = p;
    b->last_cutoff = cutoff;
    b->last_reff = reff_mode;
    b->last_result = want_result;
    b->last_iq = want_iq;
    b->have_last = true;
    if (want_result) copy_out(result, b->h_out, sizeof(double) * p.result_total);
    if (want_iq && !nchunk && !b->iq_direct) copy_out(iq, b->h_out + sizeof(double) * p.result_total, iq_bytes);
    b->iq_direct = nullptr;
    if (b->timing) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
        b->t_ms += ms;
        HIP_TRY(hipEventElapsedTime(&ms, b->ev2, b->ev3));
        b->t_all_ms += ms;
        b->ncalls++;
    }
    return 0;
}

}  // namespace

extern "C" {

int saship_batch_create(saship_model *m, int32_t nsets, const int32_t *nq, const double *q, int32_t is2d,
                        saship_batch **out)
{
    if (!out) return fail(SASHIP_EINVAL, "out is NULL");
    *out = nullptr;
    if (!m) return fail(SASHIP_EINVAL, "model is NULL");
    if (nsets < 1 || !nq) return fail(SASHIP_EINVAL, "need nsets >= 1 and a q-count array");
    if (model_structure(m->id))
        return fail(SASHIP_EINVAL, "a structure factor is batched through saship_batch_set_structure of its form factor");
    std::vector<int64_t> qoff(nsets + 1, 0);
    for (int s = 0; s < nsets; s++) {
        if (nq[s] < 0) return fail(SASHIP_EINVAL, "set " + std::to_string(s) + ": nq < 0");
        qoff[s + 1] = qoff[s] + nq[s];
    }
    const int64_t total = qoff[nsets];
    if (total > 0 && !q) return fail(SASHIP_EINVAL, "q is NULL");
    HIP_TRY(hipSetDevice(m->dev));
    saship_batch *b = new (std::nothrow) saship_batch();
    if (!b) return fail(SASHIP_ENOMEM, "out of host memory");
    b->m = m;
    model_retain(m);
    b->nsets = nsets;
    b->is2d = is2d ? 1 : 0;
    b->dev = m->dev;
    b->cus = cu_count(m->dev);
    b->qoff = qoff;
    b->tile_off.assign(nsets + 1, 0);
    for (int s = 0; s < nsets; s++) {
        const int nt = (nq[s] + sas::QTILE - 1) / sas::QTILE;
        b->tile_off[s + 1] = b->tile_off[s] + nt;
        for (int t = 0; t < nt; t++) {
            double lo = HUGE_VAL, hi = 0.0;
            for (int i = t * sas::QTILE; i < std::min(nq[s], (t + 1) * sas::QTILE); i++) {
                const size_t g = (size_t)qoff[s] + i;
                const double qv = is2d ? std::hypot(q[2 * g], q[2 * g + 1]) : std::fabs(q[g]);
                lo = std::min(lo, qv);
                hi = std::max(hi, qv);
            }
            b->tile_qmin.push_back(lo);
            b->tile_qmax.push_back(hi);
        }
    }
    // sets with the same q vector (a fit population on one data set) share
    // one copy on the device: the kernels read it through qdoff
    const size_t dims = b->is2d ? 2 : 1;
    bool shared = nsets > 1 && nq[0] > 0;
    for (int s = 1; s < nsets && shared; s++)
        shared = nq[s] == nq[0] && memcmp(q + (size_t)qoff[s] * dims, q, sizeof(double) * (size_t)nq[0] * dims) == 0;
    std::vector<int64_t> qdoff(nsets);
    for (int s = 0; s < nsets; s++) qdoff[s] = shared ? 0 : qoff[s];
    const size_t nqv = (shared ? (size_t)nq[0] : (size_t)total) * dims;
    std::vector<float> q32(std::max<size_t>(nqv, 1));
    for (size_t i = 0; i < nqv; i++) q32[i] = (float)q[i];
    hipError_t e = hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc((void **)&b->q32, sizeof(float) * std::max<size_t>(nqv, 1));
    if (e == hipSuccess) e = hipMalloc((void **)&b->q64, sizeof(double) * std::max<size_t>(nqv, 1));
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_qoff, sizeof(int64_t) * (nsets + 1));
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_qdoff, sizeof(int64_t) * nsets);
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_npts, sizeof(int) * nsets);
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_sums, sizeof(double) * 4 * nsets);
    if (e == hipSuccess && model_magnetic(m->id))
        e = hipMalloc((void **)&b->d_spin, sizeof(float) * model_spin_floats(m->id) * nsets);
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_sprep, sizeof(double) * sas::S_PREP * nsets);
    if (e == hipSuccess)
        e = hipMalloc((void **)&b->d_queue, sizeof(int) * sas::WORK_QUEUES * sas::WORK_QUEUE_STRIDE);
    if (e == hipSuccess && nqv) e = hipMemcpy(b->q32, q32.data(), sizeof(float) * nqv, hipMemcpyHostToDevice);
    if (e == hipSuccess && nqv) e = hipMemcpy(b->q64, q, sizeof(double) * nqv, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b->d_qoff, qoff.data(), sizeof(int64_t) * (nsets + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b->d_qdoff, qdoff.data(), sizeof(int64_t) * nsets, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    if (e == hipSuccess) e = hipEventCreate(&b->ev2);
    if (e == hipSuccess) e = hipEventCreate(&b->ev3);
    if (e != hipSuccess) {
        std::string msg = std::string("batch create: ") + hipGetErrorString(e);
        saship_batch_release(b);
        return fail(SASHIP_EHIP, msg);
    }
    *out = b;
    return 0;
}

int saship_batch_set_options(saship_batch *b, int32_t use_graph, int32_t pts_per_item, int32_t waves_per_cu)
{
    if (!b) return fail(SASHIP_EINVAL, "batch is NULL");
    if (pts_per_item < 1 || pts_per_item > 4096) return fail(SASHIP_EINVAL, "pts_per_item must be 1..4096");
    if (waves_per_cu < 4 || waves_per_cu > 128) return fail(SASHIP_EINVAL, "waves_per_cu must be 4..128");
    std::lock_guard<std::mutex> lock(b->mu);
    b->use_graph = use_graph != 0;
    b->pts_per_item = pts_per_item;
    b->waves_per_cu = waves_per_cu;
    drop_graph(b);
    b->have_last = false;      // the staged work list was cut for the old item size
    return 0;
}

int saship_batch_set_path(saship_batch *b, int32_t separable)
{
    if (!b) return fail(SASHIP_EINVAL, "batch is NULL");
    std::lock_guard<std::mutex> lock(b->mu);
    b->separable = separable != 0;
    drop_graph(b);
    b->have_last = false;      // the staged plan belongs to the other path
    return 0;
}

namespace batch_host {

// Shared body of the two resolution entry points; repeat == NULL: plain bands.
int set_resolution(saship_batch *b, const int32_t *n_out, const int32_t *start, const int32_t *length,
                   const int64_t *stride, const int64_t *offset, const double *weights, int64_t n_weights,
                   const int32_t *repeat, const int64_t *repeat_step, const int64_t *repeat_offset,
                   const double *repeat_weights, int64_t n_repeat_weights)
{
    if (!b) return fail(SASHIP_EINVAL, "batch is NULL");
    std::lock_guard<std::mutex> lock(b->mu);
    HIP_TRY(hipSetDevice(b->dev));
    drop_graph(b);
    b->have_last = false;      // the output buffer was sized for the old smearing
    void *old[] = {b->d_sfrom, b->d_soff, b->d_sstep, b->d_slen, b->d_sw, b->d_srep, b->d_srstep, b->d_sroff, b->d_srw};
    for (void *ptr : old)
        if (ptr) (void)hipFree(ptr);
    b->d_sfrom = b->d_soff = b->d_sstep = b->d_srstep = b->d_sroff = nullptr;
    b->d_slen = b->d_srep = nullptr;
    b->d_sw = b->d_srw = nullptr;
    b->n_smear = 0;
    if (!n_out) return 0;                          // smearing off
    if (!start || !length || !offset || (n_weights > 0 && !weights) || n_weights < 0)
        return fail(SASHIP_EINVAL, "start, length, offset and weights must not be NULL");
    if (repeat && (!repeat_step || !repeat_offset || (n_repeat_weights > 0 && !repeat_weights) || n_repeat_weights < 0))
        return fail(SASHIP_EINVAL, "repeat_step, repeat_offset and repeat_weights must not be NULL");
    int64_t total = 0;
    for (int s = 0; s < b->nsets; s++) {
        if (n_out[s] < 0) return fail(SASHIP_EINVAL, "set " + std::to_string(s) + ": n_out < 0");
        total += n_out[s];
    }
    std::vector<int64_t> from(std::max<int64_t>(total, 1)), steps(std::max<int64_t>(total, 1), 1);
    for (int s = 0, o = 0; s < b->nsets; s++) {
        const int64_t nq = b->qoff[s + 1] - b->qoff[s];
        for (int k = 0; k < n_out[s]; k++, o++) {
            const int64_t st = stride ? stride[o] : 1;
            const int64_t nr = repeat ? repeat[o] : 1, rs = repeat ? repeat_step[o] : 0;
            steps[o] = st;
            // the last point read: start + (nr - 1) rs + (length - 1) st
            if (start[o] < 0 || length[o] < 0 || st < 1 || nr < 0 || rs < 0 ||
                (length[o] > 0 && nr > 0 && start[o] + (nr - 1) * rs + (int64_t)(length[o] - 1) * st >= nq))
                return fail(SASHIP_EINVAL, "set " + std::to_string(s) + ": band of point " + std::to_string(k) +
                                               " outside the set's q points");
            if (offset[o] < 0 || offset[o] + length[o] > n_weights)
                return fail(SASHIP_EINVAL, "set " + std::to_string(s) + ": weights of point " + std::to_string(k) +
                                               " outside the weight array");
            if (repeat && (repeat_offset[o] < 0 || repeat_offset[o] + nr > n_repeat_weights))
                return fail(SASHIP_EINVAL, "set " + std::to_string(s) + ": repeat weights of point " +
                                               std::to_string(k) + " outside the repeat weight array");
            from[o] = b->qoff[s] + start[o];
        }
    }
    if (total == 0) return 0;
    hipError_t e = hipMalloc((void **)&b->d_sfrom, sizeof(int64_t) * total);
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_soff, sizeof(int64_t) * total);
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_sstep, sizeof(int64_t) * total);
    if (e == hipSuccess) e = hipMemcpy(b->d_sstep, steps.data(), sizeof(int64_t) * total, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_slen, sizeof(int32_t) * total);
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_sw, sizeof(double) * std::max<int64_t>(n_weights, 1));
    if (e == hipSuccess) e = hipMemcpy(b->d_sfrom, from.data(), sizeof(int64_t) * total, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b->d_soff, offset, sizeof(int64_t) * total, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b->d_slen, length, sizeof(int32_t) * total, hipMemcpyHostToDevice);
    if (e == hipSuccess && n_weights > 0) e = hipMemcpy(b->d_sw, weights, sizeof(double) * n_weights, hipMemcpyHostToDevice);
    if (repeat) {
        if (e == hipSuccess) e = hipMalloc((void **)&b->d_srep, sizeof(int32_t) * total);
        if (e == hipSuccess) e = hipMalloc((void **)&b->d_srstep, sizeof(int64_t) * total);
        if (e == hipSuccess) e = hipMalloc((void **)&b->d_sroff, sizeof(int64_t) * total);
        if (e == hipSuccess) e = hipMalloc((void **)&b->d_srw, sizeof(double) * std::max<int64_t>(n_repeat_weights, 1));
        if (e == hipSuccess) e = hipMemcpy(b->d_srep, repeat, sizeof(int32_t) * total, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(b->d_srstep, repeat_step, sizeof(int64_t) * total, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(b->d_sroff, repeat_offset, sizeof(int64_t) * total, hipMemcpyHostToDevice);
        if (e == hipSuccess && n_repeat_weights > 0)
            e = hipMemcpy(b->d_srw, repeat_weights, sizeof(double) * n_repeat_weights, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) return fail(SASHIP_EHIP, std::string("set resolution: ") + hipGetErrorString(e));
    b->n_smear = total;
    return 0;
}

}  // namespace batch_host

int saship_batch_set_resolution(saship_batch *b, const int32_t *n_out, const int32_t *start, const int32_t *length,
                                const int64_t *stride, const int64_t *offset, const double *weights, int64_t n_weights)
{
    return batch_host::set_resolution(b, n_out, start, length, stride, offset, weights, n_weights, nullptr, nullptr,
                                      nullptr, nullptr, 0);
}

int saship_batch_set_resolution_repeat(saship_batch *b, const int32_t *n_out, const int32_t *start,
                                       const int32_t *length, const int64_t *stride, const int64_t *offset,
                                       const double *weights, int64_t n_weights, const int32_t *repeat,
                                       const int64_t *repeat_step, const int64_t *repeat_offset,
                                       const double *repeat_weights, int64_t n_repeat_weights)
{
    if (!repeat) return fail(SASHIP_EINVAL, "repeat is NULL (use saship_batch_set_resolution)");
    return batch_host::set_resolution(b, n_out, start, length, stride, offset, weights, n_weights, repeat,
                                      repeat_step, repeat_offset, repeat_weights, n_repeat_weights);
}

int saship_batch_get_options(saship_batch *b, int32_t *use_graph, int32_t *pts_per_item, int32_t *waves_per_cu)
{
    if (!b || !use_graph || !pts_per_item || !waves_per_cu) return fail(SASHIP_EINVAL, "NULL argument");
    std::lock_guard<std::mutex> lock(b->mu);
    *use_graph = b->use_graph;
    *pts_per_item = b->pts_per_item;
    *waves_per_cu = b->waves_per_cu;
    return 0;
}

int saship_batch_call(saship_batch *b, int32_t kind, const int32_t *details, int32_t n_details,
                      const double *values, const int64_t *values_offset, double cutoff,
                      int32_t radius_effective_mode, double *result, double *iq)
{
    if (!b) return fail(SASHIP_EINVAL, "batch is NULL");
    if (!details || !values || !values_offset) return fail(SASHIP_EINVAL, "details, values and values_offset must not be NULL");
    if (kind < 0 || kind > 2) return fail(SASHIP_EINVAL, "kind must be 0 (Iq), 1 (Iqxy) or 2 (Imagnetic)");
    if ((kind == 0) == (b->is2d != 0))
        return fail(SASHIP_EINVAL, kind == 0 ? "Iq kernel needs 1D q" : "2D kernels need (qx, qy) input");
    std::lock_guard<std::mutex> lock(b->mu);
    if (b->s_model)
        for (int s = 0; s < b->nsets; s++) {
            const double er = b->spars[SASHIP_SPARS * s + 3];
            if (er != 0.0 && er != (double)radius_effective_mode)
                return fail(SASHIP_EINVAL, "set " + std::to_string(s) + ": structure radius_effective_mode " +
                                               std::to_string((int)er) + " differs from the call's " +
                                               std::to_string(radius_effective_mode));
        }
    HIP_TRY(hipSetDevice(b->dev));
    return run(b, kind, details, n_details, values, values_offset, cutoff, radius_effective_mode, result, iq);
}

int saship_batch_set_structure(saship_batch *b, int32_t s_model, const double *pars, int32_t nsets)
{
    if (!b) return fail(SASHIP_EINVAL, "batch is NULL");
    std::lock_guard<std::mutex> lock(b->mu);
    if (s_model == 0) {
        b->s_model = 0;
        b->spars.clear();
        return 0;
    }
    if (!model_structure(s_model)) return fail(SASHIP_EINVAL, "model " + std::to_string(s_model) + " is not a structure factor");
    if (model_structure(b->m->id)) return fail(SASHIP_EINVAL, "the batch model is itself a structure factor");
    if (!pars || nsets != b->nsets)
        return fail(SASHIP_EINVAL, "need " + std::to_string(SASHIP_SPARS) + " structure parameters for each of the batch's sets");
    for (int s = 0; s < nsets; s++) {
        const double *sp = pars + SASHIP_SPARS * s;
        const std::string where = "set " + std::to_string(s) + ": ";
        for (int k = 0; k < SASHIP_SPARS; k++)
            if (!std::isfinite(sp[k])) return fail(SASHIP_EINVAL, where + "structure parameters must be finite");
        if (sp[3] < 0.0 || sp[3] != std::floor(sp[3]) || sp[3] > 100.0)
            return fail(SASHIP_EINVAL, where + "radius_effective_mode must be a small non-negative integer");
    }
    b->spars.assign(pars, pars + SASHIP_SPARS * (size_t)nsets);
    b->s_model = s_model;
    b->have_last = false;
    return 0;
}

int saship_batch_replay(saship_batch *b, int32_t nsteps, double *device_ms)
{
    if (!b || !device_ms) return fail(SASHIP_EINVAL, "NULL argument");
    if (nsteps < 1) return fail(SASHIP_EINVAL, "nsteps must be >= 1");
    std::lock_guard<std::mutex> lock(b->mu);
    if (!b->have_last) return fail(SASHIP_EINVAL, "no successful saship_batch_call to replay");
    HIP_TRY(hipSetDevice(b->dev));
    const bool timing = b->timing;
    b->timing = false;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventRecord(e0, b->st);
    int rc = 0;
    for (int i = 0; i < nsteps && e == hipSuccess && rc == 0; i++)
        rc = enqueue_model(b, b->last, b->last_cutoff, b->last_reff, b->last_result, b->last_iq, false);
    if (e == hipSuccess && rc == 0) e = hipEventRecord(e1, b->st);
    if (e == hipSuccess && rc == 0) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess && rc == 0) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    b->timing = timing;
    if (rc) return rc;
    if (e != hipSuccess) return fail(SASHIP_EHIP, std::string("replay: ") + hipGetErrorString(e));
    *device_ms = ms;
    return 0;
}

int saship_batch_autotune(saship_batch *b, int32_t kind, const int32_t *details, int32_t n_details,
                          const double *values, const int64_t *values_offset, double cutoff,
                          int32_t *best_pts_per_item, int32_t *best_waves_per_cu)
{
    if (!b || !best_pts_per_item || !best_waves_per_cu) return fail(SASHIP_EINVAL, "NULL argument");
    if (!details || !values || !values_offset) return fail(SASHIP_EINVAL, "details, values and values_offset must not be NULL");
    if ((kind == 0) == (b->is2d != 0) || kind < 0 || kind > 2) return fail(SASHIP_EINVAL, "kind does not match the q input");
    std::lock_guard<std::mutex> lock(b->mu);
    HIP_TRY(hipSetDevice(b->dev));
    const bool graph = b->use_graph, timing = b->timing;
    const double t_keep = b->t_ms, t_all_keep = b->t_all_ms;
    const int n_keep = b->ncalls;
    b->use_graph = false;
    b->timing = true;
    int best_p = b->pts_per_item, best_w = b->waves_per_cu;
    double best_t = 1e300;
    const int pts[] = {4, 8, 16, 32, 64, 128};
    const int waves[] = {8, 16, 24, 32};
    constexpr int NP = sizeof(pts) / sizeof(pts[0]), NW = sizeof(waves) / sizeof(waves[0]);
    double tbest[NP][NW];
    int rc = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = fail(SASHIP_EHIP, "autotune events");
    // a shape's time: the call staged once, then the device time of three
    // back-to-back replays of its work (what the resident rate measures)
    auto measure = [&](double &t) -> int {
        b->timing = false;
        if (int e = run(b, kind, details, n_details, values, values_offset, cutoff, 0, nullptr, nullptr)) return e;
        HIP_TRY(hipEventRecord(e0, b->st));
        for (int i = 0; i < 3; i++)
            if (int e = enqueue_model(b, b->last, cutoff, 0, false, false, false)) return e;
        HIP_TRY(hipEventRecord(e1, b->st));
        HIP_TRY(hipEventSynchronize(e1));
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        t = std::min(t, (double)ms / 3.0);
        return 0;
    };
    // two passes over the shapes, each shape's fastest time over both: a
    // device still raising its clocks would otherwise favour the shapes
    // timed last
    for (int pass = 0; pass < 2 && rc == 0; pass++) {
        for (int ip = 0; ip < NP && rc == 0; ip++) {
            for (int iw = 0; iw < NW && rc == 0; iw++) {
                b->pts_per_item = pts[ip];
                b->waves_per_cu = waves[iw];
                double t = pass ? tbest[ip][iw] : 1e300;
                rc = measure(t);
                tbest[ip][iw] = t;
            }
        }
    }
    for (int ip = 0; ip < NP && rc == 0; ip++)
        for (int iw = 0; iw < NW; iw++)
            if (tbest[ip][iw] < best_t) {
                best_t = tbest[ip][iw];
                best_p = pts[ip];
                best_w = waves[iw];
            }
    b->pts_per_item = best_p;
    b->waves_per_cu = best_w;
    // then the items per queue atomic at that shape (0: the plan's choice)
    int best_g = b->grab;
    const int grabs[] = {0, 1, 2, 4, 8, 16};
    for (int G : grabs) {
        if (rc) break;
        b->grab = G;
        double t = 1e300;
        for (int rep = 0; rep < 2 && rc == 0; rep++) rc = measure(t);
        if (rc == 0 && t < best_t) {
            best_t = t;
            best_g = G;
        }
    }
    b->grab = best_g;
    b->use_graph = graph;
    b->timing = timing;
    b->t_ms = t_keep;
    b->t_all_ms = t_all_keep;
    b->ncalls = n_keep;
    drop_graph(b);
    b->have_last = false;      // the last staged plan is the sweep's last point
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (rc) return rc;
    *best_pts_per_item = best_p;
    *best_waves_per_cu = best_w;
    return 0;
}

int saship_batch_set_timing(saship_batch *b, int32_t enable)
{
    if (!b) return fail(SASHIP_EINVAL, "batch is NULL");
    std::lock_guard<std::mutex> lock(b->mu);
    b->timing = enable != 0;
    b->t_ms = 0.0;
    b->t_all_ms = 0.0;
    b->ncalls = 0;
    return 0;
}

int saship_batch_get_timing(saship_batch *b, double *eval_ms, int32_t *ncalls)
{
    if (!b || !eval_ms || !ncalls) return fail(SASHIP_EINVAL, "NULL argument");
    std::lock_guard<std::mutex> lock(b->mu);
    *eval_ms = b->t_ms;
    *ncalls = b->ncalls;
    return 0;
}

int saship_batch_release(saship_batch *b)
{
    if (!b) return 0;
    (void)hipSetDevice(b->dev);
    if (b->st) (void)hipStreamSynchronize(b->st);
    drop_graph(b);
    void *dev_ptrs[] = {b->q32, b->q64, b->d_qoff, b->d_qdoff, b->d_args, b->d_recs, b->d_npts, b->d_sums, b->d_spin, b->d_sprep,
                        b->d_queue, b->d_partial, b->d_out, b->d_sep, b->d_sfrom, b->d_soff, b->d_sstep, b->d_slen, b->d_sw,
                        b->d_srep, b->d_srstep, b->d_sroff, b->d_srw};
    for (void *ptr : dev_ptrs)
        if (ptr) (void)hipFree(ptr);
    if (b->h_stage) (void)hipHostFree(b->h_stage);
    if (b->h_out) (void)hipHostFree(b->h_out);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->ev2) (void)hipEventDestroy(b->ev2);
    if (b->ev3) (void)hipEventDestroy(b->ev3);
    for (hipEvent_t ev : b->chunk_ev) (void)hipEventDestroy(ev);
    if (b->st) (void)hipStreamDestroy(b->st);
    if (b->m) model_drop(b->m);
    delete b;
    return 0;
}

}  // extern "C"
