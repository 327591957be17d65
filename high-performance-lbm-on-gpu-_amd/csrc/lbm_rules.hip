// Sparse affine work on the population buffer: the link-rule pre-pass, affine
// readouts, and the momentum-exchange force reduction.
//
// A compiled affine set stores, term-major, slot[t][r] (int64 flat index into
// the (q, N) buffer), coeff[t][r] and constant[r], each row padded to m terms
// with zero coefficients (src/lbm/link_rules.py:54-103, 203-240). A row's
// value is constant + (coefficient * population) summed over the m terms in
// NumPy's pairwise order (8 interleaved partial sums for m >= 8), which is how
// the reference evaluates it (link_rules.py:269-272), so a rule writes the
// same bits the NumPy oracle computes.
//
// Rules are evaluated before any is written: rule values read pre-rule
// populations only, whatever the geometry (src/lbm/backends/kernels/link_rules.cu:8-11).
// Up to 4096 rules this is one workgroup with a barrier between the phases;
// above that, an evaluate launch into scratch followed by a write launch. When no
// rule reads a slot another rule writes (the host checks the compiled set), one
// launch of one rule per lane gives the same bits at any size.

#include "lbm_affine.h"
#include "lbm_force.h"

#include <algorithm>
#include <mutex>

#include "lbm_common.h"

namespace lbm {

template <class T>
__global__ __launch_bounds__(256) void k_affine_values(const T* __restrict__ F, const int64_t* __restrict__ slot,
                                                       const T* __restrict__ coeff, const T* __restrict__ constant,
                                                       int64_t n, int m, T* __restrict__ out)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) out[r] = affine_row(F, slot, coeff, constant, n, m, r);
}

// Rows of a concatenation of rule sets: row r sums only its own set's padded width
// width[r] (the term arrays are padded to the widest set), so each row rounds exactly as
// it does when its set is evaluated alone.
template <class T>
__global__ __launch_bounds__(256) void k_affine_values_w(const T* __restrict__ F, const int64_t* __restrict__ slot,
                                                         const T* __restrict__ coeff, const T* __restrict__ constant,
                                                         const int* __restrict__ width, int64_t n,
                                                         T* __restrict__ out)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) out[r] = affine_row(F, slot, coeff, constant, n, width[r], r);
}

template <class T>
__global__ __launch_bounds__(256) void k_rules_write(T* __restrict__ F, const int64_t* __restrict__ target,
                                                     const T* __restrict__ value, int64_t n)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) F[target[r]] = value[r];
}

// The whole pre-pass in one launch for a rule set in which no rule reads a slot that
// another rule writes (a rule may read its own target: its lane reads before it stores).
// Every value then comes from pre-rule populations whatever order the lanes run in, so the
// result equals the two-phase pass bit for bit. `width` (optional) gives each row its own
// padded term count, as k_affine_values_w.
template <class T>
__global__ __launch_bounds__(256) void k_rules_direct(T* __restrict__ F, const int64_t* __restrict__ slot,
                                                      const T* __restrict__ coeff, const T* __restrict__ constant,
                                                      const int* __restrict__ width,
                                                      const int64_t* __restrict__ target, int64_t n, int m)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int64_t t = target[r];
    const T v = affine_row(F, slot, coeff, constant, n, width ? width[r] : m, r);
    F[t] = v;
}

constexpr int SMALL_THREADS = 1024;
constexpr int SMALL_PER_THREAD = 4;

// One workgroup: evaluate up to 4096 rules into registers, barrier, write.
template <class T>
__global__ __launch_bounds__(SMALL_THREADS) void k_rules_small(T* __restrict__ F, const int64_t* __restrict__ slot,
                                                               const T* __restrict__ coeff,
                                                               const T* __restrict__ constant,
                                                               const int64_t* __restrict__ target, int64_t n, int m)
{
    rules_in_one_group<T, SMALL_PER_THREAD>(F, slot, coeff, constant, target, n, m, threadIdx.x, SMALL_THREADS);
}

// Force = sum_r weight[r] * value[r] (d components), the groups of the reduction as
// workgroups; with one group the sum goes to out, else to partial[b][3].
template <class T>
__global__ __launch_bounds__(FORCE_THREADS) void k_affine_force(const T* __restrict__ F,
                                                                const int64_t* __restrict__ slot,
                                                                const T* __restrict__ coeff,
                                                                const T* __restrict__ constant,
                                                                const double* __restrict__ weight, int64_t n, int m,
                                                                int d, double* __restrict__ out,
                                                                double* __restrict__ partial)
{
    __shared__ double part[3][FORCE_WAVES];
    double acc[3] = {0.0, 0.0, 0.0};
    const ForceRows<T> fr{slot, coeff, constant, weight, n, m, d};
    force_lane(F, fr, (int)gridDim.x, (int)blockIdx.x, (int)threadIdx.x, acc);
    force_tree(acc, part);
    if (threadIdx.x == 0) {
        if (gridDim.x == 1) {
            for (int a = 0; a < d; ++a) out[a] = part[a][0];
        } else {
            for (int a = 0; a < 3; ++a) partial[3 * blockIdx.x + a] = part[a][0];
        }
    }
}

// out[a] = sum over the groups' partial sums.
__global__ __launch_bounds__(FORCE_THREADS) void k_force_total(const double* __restrict__ partial, int groups, int d,
                                                               double* __restrict__ out)
{
    __shared__ double part[3][FORCE_WAVES];
    force_partials_total(partial, groups, d, out, part);
}

// The totals of a history of partial sums: row s of out (d values) from partial row s
// (groups x 3 values), one workgroup per row.
__global__ __launch_bounds__(FORCE_THREADS) void k_force_totals(const double* __restrict__ partial, int groups, int d,
                                                                double* __restrict__ out)
{
    __shared__ double part[3][FORCE_WAVES];
    const size_t row = blockIdx.x;
    force_partials_total(partial + row * 3 * (size_t)groups, groups, d, out + row * (size_t)d, part);
}

// The force of one step and the link-rule pre-pass of the next, in one launch of one
// workgroup (they are consecutive and read the same state). Lanes below FORCE_THREADS take
// the force rows exactly as k_affine_force does with one workgroup (rows tid + k *
// FORCE_THREADS, the same accumulation and tree), all lanes take the rules as
// k_rules_small does; every read of the state precedes the barrier, the rule writes follow
// it, so both give the bits of the two separate launches.
template <class T>
__global__ __launch_bounds__(SMALL_THREADS) void k_force_rules(
    T* __restrict__ F, const int64_t* __restrict__ fslot, const T* __restrict__ fcoeff,
    const T* __restrict__ fconstant, const double* __restrict__ weight, int64_t nf, int fm, int d,
    double* __restrict__ out, const int64_t* __restrict__ rslot, const T* __restrict__ rcoeff,
    const T* __restrict__ rconstant, const int64_t* __restrict__ rtarget, int64_t nr, int rm)
{
    static_assert(SMALL_THREADS == FORCE_THREADS, "one lane per force row and per rule");
    __shared__ double part[3][FORCE_WAVES];
    const int tid = threadIdx.x;
    double acc[3] = {0.0, 0.0, 0.0};
    force_lane(F, ForceRows<T>{fslot, fcoeff, fconstant, weight, nf, fm, d}, 1, 0, tid, acc);
    force_wave_sums(acc, part);
    T v[SMALL_PER_THREAD];
    int64_t t[SMALL_PER_THREAD];
    rules_values<T, SMALL_PER_THREAD>(F, rslot, rcoeff, rconstant, rtarget, nr, rm, tid, SMALL_THREADS, v, t);
    __syncthreads();  // also the force tree's barrier
    rules_store<T, SMALL_PER_THREAD>(F, nr, tid, SMALL_THREADS, v, t);
    force_waves_total(part);
    if (tid == 0)
        for (int a = 0; a < d; ++a) out[a] = part[a][0];
}

// Rule values of lbm_rules_apply_noalloc: the library's own device buffer, one per GPU
// (each device holds its own copy of this array until the process ends). Host code never
// passes the array's name to a kernel: on the host that name is only a stand-in for the
// symbol, so kernels get the device address that hipGetSymbolAddress returns for the
// current device (noalloc_scratch_ptr).
constexpr int64_t NOALLOC_MAX_RULES = int64_t(1) << 20;
static __device__ __align__(256) double rules_noalloc_scratch[NOALLOC_MAX_RULES];

constexpr int NOALLOC_MAX_DEVICES = 64;

// Device address of rules_noalloc_scratch on the current device, looked up on the first
// call on that device and kept. hipGetSymbolAddress allocates nothing.
inline void* noalloc_scratch_ptr()
{
    static std::mutex lock;
    static void* address[NOALLOC_MAX_DEVICES] = {};
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    require(dev >= 0 && dev < NOALLOC_MAX_DEVICES, "device index out of range");
    std::lock_guard<std::mutex> hold(lock);
    if (address[dev] == nullptr) {
        void* ptr = nullptr;
        hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(rules_noalloc_scratch)), "hipGetSymbolAddress(scratch)");
        require(ptr != nullptr, "hipGetSymbolAddress returned no address", LBM_ERR_HIP);
        address[dev] = ptr;
    }
    return address[dev];
}

// The launches of lbm_rules_apply on a given scratch buffer (used when n_rules > 4096).
template <class T>
void rules_apply_launch(T* f, const int64_t* slot, const T* coeff, const T* constant, const int64_t* target,
                        int64_t n_rules, int m, T* scratch, hipStream_t s)
{
    if (n_rules <= (int64_t)SMALL_THREADS * SMALL_PER_THREAD) {
        k_rules_small<T><<<1, SMALL_THREADS, 0, s>>>(f, slot, coeff, constant, target, n_rules, m);
    } else {
        require(scratch != nullptr, "more than 4096 rules need a scratch buffer of n_rules values");
        const unsigned blocks = (unsigned)((n_rules + 255) / 256);
        k_affine_values<T><<<blocks, 256, 0, s>>>(f, slot, coeff, constant, n_rules, m, scratch);
        k_rules_write<T><<<blocks, 256, 0, s>>>(f, target, scratch, n_rules);
    }
}

}  // namespace lbm

using namespace lbm;

extern "C" {

LBM_API int lbm_rules_apply(int dtype, void* f, const int64_t* slot, const void* coeff, const void* constant,
                            const int64_t* target, int64_t n_rules, int m, void* scratch, void* stream)
{
    return guarded("lbm_rules_apply", [&] {
        require(n_rules >= 0 && m >= 1 && m <= 128, "rules need 1 <= m <= 128 terms");
        if (n_rules == 0) return;
        require(f && slot && coeff && constant && target, "null pointer");
        dispatch_real(dtype, [&](auto Tt) {
            using T = decltype(Tt);
            rules_apply_launch<T>(static_cast<T*>(f), slot, static_cast<const T*>(coeff),
                                  static_cast<const T*>(constant), target, n_rules, m, static_cast<T*>(scratch),
                                  as_stream(stream));
            check_launch("rule pre-pass");
        });
    });
}

LBM_API int lbm_rules_apply_noalloc(int dtype, void* f, const int64_t* slot, const void* coeff, const void* constant,
                                    const int64_t* target, int64_t n_rules, int m, void* stream)
{
    return guarded("lbm_rules_apply_noalloc", [&] {
        require(n_rules >= 0 && m >= 1 && m <= 128, "rules need 1 <= m <= 128 terms");
        require(n_rules <= NOALLOC_MAX_RULES,
                "at most 1048576 rules: the library's scratch holds 2^20 rule values; use lbm_rules_apply with a "
                "caller scratch buffer for larger sets");
        require(dtype == LBM_F32 || dtype == LBM_F64, "unknown dtype code");
        if (n_rules == 0) return;
        require(f && slot && coeff && constant && target, "null pointer");
        void* scratch = n_rules > (int64_t)SMALL_THREADS * SMALL_PER_THREAD ? noalloc_scratch_ptr() : nullptr;
        dispatch_real(dtype, [&](auto Tt) {
            using T = decltype(Tt);
            rules_apply_launch<T>(static_cast<T*>(f), slot, static_cast<const T*>(coeff),
                                  static_cast<const T*>(constant), target, n_rules, m, static_cast<T*>(scratch),
                                  as_stream(stream));
            check_launch("rule pre-pass (library scratch)");
        });
    });
}

LBM_API int lbm_rules_apply_widths(int dtype, void* f, const int64_t* slot, const void* coeff, const void* constant,
                                   const int64_t* target, const int* width, int64_t n_rules, int m, void* scratch,
                                   void* stream)
{
    return guarded("lbm_rules_apply_widths", [&] {
        require(n_rules >= 0 && m >= 1 && m <= 128, "rules need 1 <= m <= 128 terms");
        if (n_rules == 0) return;
        require(f && slot && coeff && constant && target && width && scratch, "null pointer");
        dispatch_real(dtype, [&](auto Tt) {
            using T = decltype(Tt);
            const hipStream_t s = as_stream(stream);
            const unsigned blocks = (unsigned)((n_rules + 255) / 256);
            k_affine_values_w<T><<<blocks, 256, 0, s>>>(static_cast<const T*>(f), slot, static_cast<const T*>(coeff),
                                                        static_cast<const T*>(constant), width, n_rules,
                                                        static_cast<T*>(scratch));
            k_rules_write<T><<<blocks, 256, 0, s>>>(static_cast<T*>(f), target, static_cast<const T*>(scratch), n_rules);
            check_launch("batched rule pre-pass");
        });
    });
}

LBM_API int lbm_rules_apply_direct(int dtype, void* f, const int64_t* slot, const void* coeff, const void* constant,
                                   const int* width, const int64_t* target, int64_t n_rules, int m, void* stream)
{
    return guarded("lbm_rules_apply_direct", [&] {
        require(n_rules >= 0 && m >= 1 && m <= 128, "rules need 1 <= m <= 128 terms");
        if (n_rules == 0) return;
        require(f && slot && coeff && constant && target, "null pointer");
        dispatch_real(dtype, [&](auto Tt) {
            using T = decltype(Tt);
            k_rules_direct<T><<<(unsigned)((n_rules + 255) / 256), 256, 0, as_stream(stream)>>>(
                static_cast<T*>(f), slot, static_cast<const T*>(coeff), static_cast<const T*>(constant), width, target,
                n_rules, m);
            check_launch("one-launch rule pre-pass");
        });
    });
}

LBM_API int lbm_affine_values(int dtype, const void* f, const int64_t* slot, const void* coeff, const void* constant,
                              int64_t n_rows, int m, void* out, void* stream)
{
    return guarded("lbm_affine_values", [&] {
        require(n_rows >= 0 && m >= 1 && m <= 128, "affine rows need 1 <= m <= 128 terms");
        if (n_rows == 0) return;
        require(f && slot && coeff && constant && out, "null pointer");
        dispatch_real(dtype, [&](auto Tt) {
            using T = decltype(Tt);
            k_affine_values<T><<<(unsigned)((n_rows + 255) / 256), 256, 0, as_stream(stream)>>>(
                static_cast<const T*>(f), slot, static_cast<const T*>(coeff), static_cast<const T*>(constant), n_rows,
                m, static_cast<T*>(out));
            check_launch("affine values");
        });
    });
}

LBM_API int lbm_affine_force_rules(int dtype, void* f, const int64_t* force_slot, const void* force_coeff,
                                   const void* force_constant, const double* weight, int64_t n_rows, int force_m,
                                   int d, double* out, const int64_t* rule_slot, const void* rule_coeff,
                                   const void* rule_constant, const int64_t* rule_target, int64_t n_rules,
                                   int rule_m, void* stream)
{
    return guarded("lbm_affine_force_rules", [&] {
        require(n_rows >= 0 && n_rows <= FORCE_ONE_GROUP_ROWS, "the fused force takes at most 2048 rows");
        require(n_rules >= 0 && n_rules <= (int64_t)SMALL_THREADS * SMALL_PER_THREAD,
                "the fused pre-pass takes at most 4096 rules");
        require(force_m >= 1 && force_m <= 128 && rule_m >= 1 && rule_m <= 128, "affine rows need 1 <= m <= 128 terms");
        require(d >= 1 && d <= 3, "d must be 1, 2 or 3");
        require(f && out, "null pointer");
        require(n_rows == 0 || (force_slot && force_coeff && force_constant && weight), "null force rows");
        require(n_rules == 0 || (rule_slot && rule_coeff && rule_constant && rule_target), "null rules");
        dispatch_real(dtype, [&](auto Tt) {
            using T = decltype(Tt);
            k_force_rules<T><<<1, SMALL_THREADS, 0, as_stream(stream)>>>(
                static_cast<T*>(f), force_slot, static_cast<const T*>(force_coeff),
                static_cast<const T*>(force_constant), weight, n_rows, force_m, d, out, rule_slot,
                static_cast<const T*>(rule_coeff), static_cast<const T*>(rule_constant), rule_target, n_rules, rule_m);
            check_launch("fused force and pre-pass");
        });
    });
}

LBM_API int lbm_force_totals(const double* partial, int64_t rows, int groups, int d, double* out, void* stream)
{
    return guarded("lbm_force_totals", [&] {
        require(rows >= 0 && rows < (1ll << 31), "row count out of range");
        require(groups >= 1 && groups <= FORCE_MAX_GROUPS, "1 <= groups <= 1024");
        require(d >= 1 && d <= 3, "d must be 1, 2 or 3");
        if (rows == 0) return;
        require(partial && out, "null pointer");
        k_force_totals<<<(unsigned)rows, FORCE_THREADS, 0, as_stream(stream)>>>(partial, groups, d, out);
        check_launch("force totals");
    });
}

LBM_API int lbm_affine_force_groups(int64_t n_rows, int* groups)
{
    return guarded("lbm_affine_force_groups", [&] {
        require(n_rows >= 0 && groups != nullptr, "bad arguments");
        *groups = force_groups(n_rows);
    });
}

LBM_API int lbm_affine_force_scratch_size(int64_t n_rows, int64_t* bytes)
{
    return guarded("lbm_affine_force_scratch_size", [&] {
        require(n_rows >= 0 && bytes != nullptr, "bad arguments");
        const int groups = force_groups(n_rows);
        *bytes = groups == 1 ? 0 : (int64_t)sizeof(double) * 3 * groups;
    });
}

LBM_API int lbm_affine_force(int dtype, const void* f, const int64_t* slot, const void* coeff, const void* constant,
                             const double* weight, int64_t n_rows, int m, int d, double* out, void* scratch,
                             void* stream)
{
    return guarded("lbm_affine_force", [&] {
        require(n_rows >= 0 && m >= 1 && m <= 128, "affine rows need 1 <= m <= 128 terms");
        require(d >= 1 && d <= 3, "d must be 1, 2 or 3");
        require(out != nullptr, "null output");
        require(n_rows == 0 || (f && slot && coeff && constant && weight), "null pointer");
        const int groups = force_groups(n_rows);
        require(groups == 1 || scratch != nullptr,
                "more than 2048 rows need a scratch buffer of lbm_affine_force_scratch_size bytes");
        double* partial = static_cast<double*>(scratch);
        dispatch_real(dtype, [&](auto Tt) {
            using T = decltype(Tt);
            const hipStream_t s = as_stream(stream);
            k_affine_force<T><<<groups, FORCE_THREADS, 0, s>>>(static_cast<const T*>(f), slot,
                                                               static_cast<const T*>(coeff),
                                                               static_cast<const T*>(constant), weight, n_rows, m, d,
                                                               out, partial);
            if (groups > 1) k_force_total<<<1, FORCE_THREADS, 0, s>>>(partial, groups, d, out);
            check_launch("affine force");
        });
    });
}

}  // extern "C"
