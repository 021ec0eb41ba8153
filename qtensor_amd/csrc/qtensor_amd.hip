// qtensor_amd.hip -- CDNA4 (gfx950) kernels and the C ABI of the bucket-elimination engine.
//
// Two execution forms of the same bucket step:
//
//  * qb_step     -- one launch per bucket (the plain backend behind process_bucket).
//                   Grid-stride over output elements; every input offset is a bit
//                   gather ("pext") of the union offset, evaluated from precomputed
//                   runs of contiguous mask bits with one bit-field extract per run.
//  * qb_execute  -- persistent batched executor.  One work item = one compiled
//                   lightcone plan (one edge or one slice of an edge).  Items are
//                   pulled from an atomic queue; each item builds its gate tensors
//                   in LDS from FP64 angles, then runs its steps with intermediate
//                   tensors in LDS (small) or in a private slot of global scratch
//                   (large), synchronising only where the plan marks a hazard.
//
// Index convention (shared with the host plan compiler): a tensor over variables
// a_0 < a_1 < ... < a_{r-1} (all of size 2) is stored in C order, so a_0 is the most
// significant bit of the element offset.  In a step over union U, input j's offset
// for union offset u is pext(u, mask_j).

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>
#include <string>

#include "../../include/qtensor_amd.h"
#include "internal.h"

namespace {

thread_local std::string g_last_error;
int g_device = -1;
int g_num_cus = 0;
std::mutex g_mutex;

int set_error(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define QB_HIP(call)                                                                 \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess)                                                        \
            return set_error(QB_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

// Complex values as clang extended vectors (re, im): builtin loads and stores in
// every address space.
typedef float cf32 __attribute__((ext_vector_type(2)));
typedef double cf64 __attribute__((ext_vector_type(2)));
typedef int32_t i32x4 __attribute__((ext_vector_type(4)));
template <typename R> struct Cx;
template <> struct Cx<float> { using V = cf32; };
template <> struct Cx<double> { using V = cf64; };

// Pointers into LDS keep their address space so that LDS and global accesses are
// never merged into generic (flat) accesses.
#define QB_LDS __attribute__((address_space(3)))
#define QB_GLB __attribute__((address_space(1)))

template <typename V>
__device__ __forceinline__ V cmul(V a, V b) {
    V r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}

template <typename V>
__device__ __forceinline__ V cadd(V a, V b) {
    V r;
    r.x = a.x + b.x;
    r.y = a.y + b.y;
    return r;
}

// A group of M work items of one plan runs in lockstep in one executor block:
// their LDS (and scratch) elements are interleaved, element e of item i at
// e * M + i, so one group element holds element e of every item of the group
// and the group shares one pass over the plan's op records and lane tables.
// Groups of two complex64 items: (re0, im0, re1, im1), one 16-byte access.
typedef float cf32x2 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ cf32x2 cmul(cf32x2 a, cf32x2 b) {
    cf32x2 r;
    r[0] = a[0] * b[0] - a[1] * b[1];
    r[1] = a[0] * b[1] + a[1] * b[0];
    r[2] = a[2] * b[2] - a[3] * b[3];
    r[3] = a[2] * b[3] + a[3] * b[2];
    return r;
}

__device__ __forceinline__ cf32x2 cadd(cf32x2 a, cf32x2 b) {
    cf32x2 r;
    r[0] = a[0] + b[0];
    r[1] = a[1] + b[1];
    r[2] = a[2] + b[2];
    r[3] = a[3] + b[3];
    return r;
}

// acc + a b in four fused multiply-adds (instead of a product and an add)
template <typename V>
__device__ __forceinline__ V cfma(V acc, V a, V b) {
    V r;
    r.x = __builtin_fmaf(-a.y, b.y, __builtin_fmaf(a.x, b.x, acc.x));
    r.y = __builtin_fmaf(a.y, b.x, __builtin_fmaf(a.x, b.y, acc.y));
    return r;
}
template <>
__device__ __forceinline__ cf64 cfma(cf64 acc, cf64 a, cf64 b) {
    cf64 r;
    r.x = __builtin_fma(-a.y, b.y, __builtin_fma(a.x, b.x, acc.x));
    r.y = __builtin_fma(a.y, b.x, __builtin_fma(a.x, b.y, acc.y));
    return r;
}
__device__ __forceinline__ cf32x2 cfma(cf32x2 acc, cf32x2 a, cf32x2 b) {
    cf32x2 r;
    r[0] = __builtin_fmaf(-a[1], b[1], __builtin_fmaf(a[0], b[0], acc[0]));
    r[1] = __builtin_fmaf(a[1], b[0], __builtin_fmaf(a[0], b[1], acc[1]));
    r[2] = __builtin_fmaf(-a[3], b[3], __builtin_fmaf(a[2], b[2], acc[2]));
    r[3] = __builtin_fmaf(a[3], b[2], __builtin_fmaf(a[2], b[3], acc[3]));
    return r;
}

// the executor's element type for items of precision R run in groups of M
template <typename R, int M> struct PkOf;
template <typename R> struct PkOf<R, 1> { using T = typename Cx<R>::V; };
template <> struct PkOf<float, 2> { using T = cf32x2; };

template <typename V>
__device__ __forceinline__ V pk_fill(V x, V*) { return x; }
__device__ __forceinline__ cf32x2 pk_fill(cf32 x, cf32x2*) { return cf32x2{x.x, x.y, x.x, x.y}; }

// s + (s of the lane whose index differs in bit b), within a wave
template <typename V>
__device__ __forceinline__ V pk_xor_add(V s, int b) {
    s.x += __shfl_xor(s.x, b, 64);
    s.y += __shfl_xor(s.y, b, 64);
    return s;
}
__device__ __forceinline__ cf32x2 pk_xor_add(cf32x2 s, int b) {
    s[0] += __shfl_xor(s[0], b, 64);
    s[1] += __shfl_xor(s[1], b, 64);
    s[2] += __shfl_xor(s[2], b, 64);
    s[3] += __shfl_xor(s[3], b, 64);
    return s;
}

__device__ __forceinline__ uint32_t pext_loop(uint32_t x, uint32_t m) {
    uint32_t r = 0, bit = 1;
    while (m) {
        uint32_t low = m & (0u - m);
        if (x & low) r |= bit;
        bit <<= 1;
        m &= m - 1;
    }
    return r;
}

__device__ __forceinline__ uint32_t bfe(uint32_t x, uint32_t off, uint32_t w) {
    return __builtin_amdgcn_ubfe(x, off, w);
}

// Host-side pext-run decomposition: runs of contiguous set bits of `mask`,
// listed from the least significant; run r copies bits [src, src+w) of the
// union offset to bits [dst, dst+w) of the input offset.
int mask_runs(uint32_t mask, uint8_t* src, uint8_t* w, uint8_t* dst, int max_runs) {
    int n = 0;
    uint32_t d = 0;
    for (uint32_t b = 0; b < 32;) {
        if (!((mask >> b) & 1u)) { ++b; continue; }
        uint32_t s = b;
        while (b < 32 && ((mask >> b) & 1u)) ++b;
        if (n < max_runs) { src[n] = (uint8_t)s; w[n] = (uint8_t)(b - s); dst[n] = (uint8_t)d; }
        d += b - s;
        ++n;
    }
    return n;
}

// ---------------------------------------------------------------------------
// Plain path: one bucket step per launch
// ---------------------------------------------------------------------------

constexpr int kStepMaxRuns = 6;

struct StepInput {
    const void* ptr;
    uint32_t top;      // offset contributed by the summed variable = 1
    uint32_t mask;     // mask over the output bits (summed bit removed)
    uint32_t n_runs;   // > kStepMaxRuns: evaluate with pext_loop(mask)
    uint8_t src[kStepMaxRuns], w[kStepMaxRuns], dst[kStepMaxRuns];
};

struct StepArgs {
    StepInput in[QB_MAX_STEP_INPUTS];
    void* out;
    uint32_t n_in;
    uint32_t rank_out;
    uint32_t n_sum;    // 1 or 2
};

template <typename R>
__device__ __forceinline__ uint32_t step_offset(const StepInput& d, uint32_t o) {
    if (d.n_runs > (uint32_t)kStepMaxRuns) return pext_loop(o, d.mask);
    uint32_t off = 0;
#pragma unroll
    for (int r = 0; r < kStepMaxRuns; ++r)
        if ((uint32_t)r < d.n_runs) off |= bfe(o, d.src[r], d.w[r]) << d.dst[r];
    return off;
}

template <typename R>
__global__ __launch_bounds__(256) void step_kernel(const StepArgs a) {
    using V = typename Cx<R>::V;
    const uint64_t n_out = 1ull << a.rank_out;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t oo = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; oo < n_out; oo += stride) {
        const uint32_t o = (uint32_t)oo;
        V acc;
        acc.x = 0;
        acc.y = 0;
        for (uint32_t sv = 0; sv < a.n_sum; ++sv) {
            V prod;
            prod.x = 1;
            prod.y = 0;
#pragma unroll
            for (int j = 0; j < QB_MAX_STEP_INPUTS; ++j) {
                if ((uint32_t)j < a.n_in) {
                    const uint32_t off = step_offset<R>(a.in[j], o) + (sv ? a.in[j].top : 0u);
                    const V v = reinterpret_cast<const V*>(a.in[j].ptr)[off];
                    prod = j == 0 ? v : cmul(prod, v);
                }
            }
            acc = sv ? cadd(acc, prod) : prod;
        }
        reinterpret_cast<V*>(a.out)[o] = acc;
    }
}

// Specialised plain step for K <= 4 inputs.  Thread t of every block handles
// elements base + t; its low offsets pext(t, mask) are computed once, and the
// high part pext(base, mask) is block-uniform (scalar).  All input reads of an
// element are issued before the first product.
template <int K>
struct StepArgsK {
    const void* ptr[K];
    uint32_t mask[K];   // output bits the input carries (summed bit removed)
    uint32_t top[K];    // offset of the summed variable = 1
    uint32_t runs[K];   // pext runs of mask (n_runs << 30 | run0 | run1 << 15)
    uint32_t roff[K][8];// offset of register-bit pattern r in input j
    uint32_t rdep[8];   // output offset of register-bit pattern r
    uint32_t hi_runs[K];// pext runs of mask >> 8
    uint32_t out_hi;    // output bits >= 8 that are not register bits, shifted down by 8
    uint32_t dep_runs[2];// out_hi as deposit runs: count << 28 | (dst | width << 5) << 10 k
    uint32_t n_reg;     // register bits (0..3)
    void* out;
    uint32_t rank_out;  // output rank without the register bits (loop extent)
    uint32_t sum;
    int64_t zbytes;     // bytes between the arenas of consecutive slices (grid z)
};

// Deposit of the low bits of x into the bit runs listed in dep (at most 5 runs,
// lowest first; dep[0] bits 28..30 hold the count).
__device__ __forceinline__ uint32_t pdep_runs(uint32_t x, const uint32_t* dep) {
    const uint32_t n = dep[0] >> 28;
    uint32_t r = 0, src = 0;
#pragma unroll
    for (uint32_t k = 0; k < 5; ++k) {
        if (k >= n) break;
        const uint32_t f = k < 2 ? (dep[0] >> (10 * k)) : (dep[1] >> (10 * (k - 2)));
        const uint32_t dst = f & 31u, w = (f >> 5) & 31u;
        r |= __builtin_amdgcn_ubfe(x, src, w) << dst;
        src += w;
    }
    return r;
}

__device__ __forceinline__ uint32_t pdep_loop(uint32_t x, uint32_t m) {
    uint32_t r = 0;
    while (m) {
        const uint32_t low = m & (0u - m);
        if (x & 1u) r |= low;
        x >>= 1;
        m &= m - 1;
    }
    return r;
}

__device__ __forceinline__ uint32_t in_offset(uint32_t o, uint32_t mask, uint32_t runs);

template <typename R, int K>
__global__ __launch_bounds__(256) void step_kernel_k(const StepArgsK<K> a) {
    using V = typename Cx<R>::V;
    constexpr int kU = 4;   // grid-stride iterations in flight per thread
    const uint32_t tid = threadIdx.x;
    uint32_t off0[K], lo_bits[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        off0[j] = in_offset(tid, a.mask[j], a.runs[j]);
        lo_bits[j] = (uint32_t)__builtin_popcount(a.mask[j] & 255u);
    }
    const uint32_t n_reg = 1u << a.n_reg;
    const uint64_t n_out = 1ull << a.rank_out;
    const uint64_t stride = (uint64_t)gridDim.x * 256u;
    const int64_t zo = (int64_t)blockIdx.z * a.zbytes;       // this block's slice
    V* out = reinterpret_cast<V*>(reinterpret_cast<char*>(a.out) + zo);
    const V* p0 = reinterpret_cast<const V*>(reinterpret_cast<const char*>(a.ptr[0]) + zo);
    for (uint64_t base0 = (uint64_t)blockIdx.x * 256u; base0 < n_out; base0 += kU * stride) {
        uint32_t off[kU][K];
        uint64_t o0[kU];
        bool live[kU];
        V x0[kU], y0[kU];
        // offsets and the largest input's reads for kU block iterations
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const uint64_t base = base0 + (uint64_t)u * stride;
            live[u] = base + tid < n_out;
            const uint32_t hb = base < n_out ? pdep_runs((uint32_t)(base >> 8), a.dep_runs) : 0u;
#pragma unroll
            for (int j = 0; j < K; ++j)
                off[u][j] = off0[j] + (in_offset(hb, a.mask[j] >> 8, a.hi_runs[j]) << lo_bits[j]);
            o0[u] = ((uint64_t)hb << 8) + tid;
            if (live[u]) {
                x0[u] = p0[off[u][0]];
                if (a.sum) y0[u] = p0[off[u][0] + a.top[0]];
            }
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            if (!live[u]) continue;
            // input 0 is the largest: it carries no register bit, so it is read once
            // for all register patterns
#pragma unroll
            for (uint32_t r = 0; r < 8; ++r) {
                if (r >= n_reg) break;
                V x[K], y[K];
#pragma unroll
                for (int j = 1; j < K; ++j) {
                    const V* p = reinterpret_cast<const V*>(reinterpret_cast<const char*>(a.ptr[j]) + zo);
                    const uint32_t oj = off[u][j] + a.roff[j][r];
                    x[j] = p[oj];
                    if (a.sum) y[j] = p[oj + a.top[j]];
                }
                V v = x0[u];
#pragma unroll
                for (int j = 1; j < K; ++j) v = cmul(v, x[j]);
                if (a.sum) {
                    V q = y0[u];
#pragma unroll
                    for (int j = 1; j < K; ++j) q = cmul(q, y[j]);
                    v = cadd(v, q);
                }
                out[o0[u] + a.rdep[r]] = v;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Fused steps (wide plans): several bucket steps contracted in one pass
// ---------------------------------------------------------------------------

// Fused op record (int32 words, written by plan.PlanCompiler._encode_fused):
//   w0 = kOpFused | chain flag (1 << 12, kFusedChain below) | n_in | sync << 11 | len << 16
//   w1 = rank_u | rank_out << 8 | k << 16 | lo_bits << 24, w2 = out_loc, w3 = 0,
//   w4 = omask, w5 = smask (output and summed variables as bits of the union
//   offset), w6 = w7 = 0; per input j: loc, mask, 0, 0; then LO[o][j] for
//   o < 2^lo_bits and X[x][j] for x < 2^k, the offsets pext(pdep(o, omask), mask_j)
//   and pext(pdep(x, smask), mask_j), which the host checks against the masks.
// out[o] = sum over x of prod_j in_j[LO/high part of o + X[x][j]].  A chain op
// (three inputs) has the same value; the flag selects the factorised kernel.
constexpr int kFusedMaxIn = 8;
constexpr int kFusedHdr = 8;
constexpr uint32_t kOpFused = 1u << 14;
constexpr uint32_t kOpReduce = 1u << 13;

constexpr int kFusedHiRuns = 6;

struct FusedArgs {
    const char* in[kFusedMaxIn];  // input base addresses in the arena
    uint32_t mask[kFusedMaxIn];   // input variables as bits of the union offset
    // Offset of input j contributed by the high output bits hi = o >> lo_bits:
    // hi_n[j] <= kFusedHiRuns runs, run r copying bits [src, src + w) of hi to
    // bits [dst, dst + w) of the offset, packed src | w << 5 | dst << 10 in
    // hi_run[j][r]; more runs: pext(pdep(hi << lo_bits, omask), mask[j]).
    uint32_t hi_run[kFusedMaxIn][kFusedHiRuns];
    uint32_t hi_n[kFusedMaxIn];
    char* out;
    uint32_t omask;               // output variables as bits of the union offset
    uint32_t rank_out, k, lo_bits;
    int64_t zbytes;               // bytes between the arenas of consecutive slices (grid z)
};

template <int NI>
__device__ __forceinline__ uint32_t fused_hi_offset(const FusedArgs& a, int j, uint32_t hi) {
    const uint32_t n = a.hi_n[j];
    if (n > (uint32_t)kFusedHiRuns) return pext_loop(pdep_loop(hi << a.lo_bits, a.omask), a.mask[j]);
    uint32_t off = 0;
#pragma unroll
    for (int r = 0; r < kFusedHiRuns; ++r) {
        if ((uint32_t)r >= n) break;
        const uint32_t f = a.hi_run[j][r];
        off |= bfe(hi, f & 31u, (f >> 5) & 31u) << (f >> 10);
    }
    return off;
}

// out[o] = sum_x prod_j in_j[off_j(o) + x_tab[x][j]], o = hi << lo_bits | lo.
// Thread t of a block owns lo = t for every hi the block visits: its low
// offsets lo_tab[t][j] are read once, and the high part of every offset is
// uniform over the block (scalar arithmetic on the hi runs).  Blocks have
// 256 threads and lo_bits = min(8, rank_out); with rank_out < 8 only the
// first 2^rank_out threads of the one block work.  The tables come as
// separate restrict arguments so that their uniform reads are scalar loads;
// two summed assignments are in flight per iteration when k >= 1.
template <typename R, int NI>
__global__ __launch_bounds__(256) void fused_kernel(const FusedArgs a, const int32_t* __restrict__ lo_tab,
                                                    const int32_t* __restrict__ x_tab) {
    using V = typename Cx<R>::V;
    const uint32_t tid = threadIdx.x;
    if (tid >= (1u << a.lo_bits)) return;
    uint32_t lo_off[NI];
#pragma unroll
    for (int j = 0; j < NI; ++j) lo_off[j] = (uint32_t)lo_tab[tid * NI + j];
    const uint32_t n_x = 1u << a.k;
    const uint32_t n_hi = 1u << (a.rank_out - a.lo_bits);
    const int64_t zo = (int64_t)blockIdx.z * a.zbytes;       // this block's slice
    for (uint32_t hi = blockIdx.x; hi < n_hi; hi += gridDim.x) {
        const V* in[NI];
#pragma unroll
        for (int j = 0; j < NI; ++j)
            in[j] = reinterpret_cast<const V*>(a.in[j] + zo) + lo_off[j] + fused_hi_offset<NI>(a, j, hi);
        V acc = {0, 0};
        if (n_x == 1u) {
            V p = in[0][x_tab[0]];
#pragma unroll
            for (int j = 1; j < NI; ++j) p = cmul(p, in[j][x_tab[j]]);
            acc = p;
        } else {
            for (uint32_t x = 0; x < n_x; x += 2) {
                const int32_t* xt = x_tab + x * NI;
                V p = in[0][xt[0]], q = in[0][xt[NI]];
#pragma unroll
                for (int j = 1; j < NI; ++j) {
                    p = cmul(p, in[j][xt[j]]);
                    q = cmul(q, in[j][xt[NI + j]]);
                }
                acc = cadd(acc, cadd(p, q));
            }
        }
        reinterpret_cast<V*>(a.out + zo)[((uint64_t)hi << a.lo_bits) | tid] = acc;
    }
}

// Two-input fused op with register blocking: out[o] = sum_x A[a(o, x)] B[b(o, x)].
// Up to three (complex64; two for complex128) high output bits that A carries
// and B does not (regA) and as many that B carries and A does not (regB) are
// handled in registers: per summed
// assignment a thread loads 2^RA values of A and 2^RB of B and forms their
// 2^(RA+RB) products, instead of two loads per output.  Blocks visit the
// remaining high bits hk; hi = deposit(hk) over the non-register high bits
// (at most kFusedDepRuns runs, src | w << 5 | dst << 10 each).
constexpr int kFusedDepRuns = 5;

struct Fused2Args {
    uint32_t roff_a[8], roff_b[8];   // offset in A (B) of register pattern p of regA (regB)
    uint32_t rdep_a[8], rdep_b[8];   // high-bit pattern of register pattern p
    uint32_t dep_run[kFusedDepRuns];
    uint32_t n_dep, hk_bits;
};

template <typename R, int RA, int RB>
__global__ __launch_bounds__(256) void fused2_kernel(const FusedArgs a, const int32_t* __restrict__ lo_tab,
                                                     const int32_t* __restrict__ x_tab, const Fused2Args e) {
    using V = typename Cx<R>::V;
    constexpr int NA = 1 << RA, NB = 1 << RB;
    const uint32_t tid = threadIdx.x;
    if (tid >= (1u << a.lo_bits)) return;
    const uint32_t lo_a = (uint32_t)lo_tab[tid * 2], lo_b = (uint32_t)lo_tab[tid * 2 + 1];
    const int64_t zo = (int64_t)blockIdx.z * a.zbytes;       // this block's slice
    const V* __restrict__ in_a = reinterpret_cast<const V*>(a.in[0] + zo);
    const V* __restrict__ in_b = reinterpret_cast<const V*>(a.in[1] + zo);
    const uint32_t n_x = 1u << a.k;
    const uint32_t n_hk = 1u << e.hk_bits;
    for (uint32_t hk = blockIdx.x; hk < n_hk; hk += gridDim.x) {
        uint32_t hi = 0;
#pragma unroll
        for (int r = 0; r < kFusedDepRuns; ++r) {
            if ((uint32_t)r >= e.n_dep) break;
            const uint32_t f = e.dep_run[r];
            hi |= bfe(hk, f & 31u, (f >> 5) & 31u) << (f >> 10);
        }
        const V* pa = in_a + lo_a + fused_hi_offset<2>(a, 0, hi);
        const V* pb = in_b + lo_b + fused_hi_offset<2>(a, 1, hi);
        V acc[NA][NB];
#pragma unroll
        for (int i = 0; i < NA; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j) acc[i][j] = V{0, 0};
        for (uint32_t x = 0; x < n_x; ++x) {
            const uint32_t xa = (uint32_t)x_tab[2 * x], xb = (uint32_t)x_tab[2 * x + 1];
            V va[NA], vb[NB];
#pragma unroll
            for (int i = 0; i < NA; ++i) va[i] = pa[xa + e.roff_a[i]];
#pragma unroll
            for (int j = 0; j < NB; ++j) vb[j] = pb[xb + e.roff_b[j]];
#pragma unroll
            for (int i = 0; i < NA; ++i)
#pragma unroll
                for (int j = 0; j < NB; ++j) acc[i][j] = cfma(acc[i][j], va[i], vb[j]);
        }
        V* out = reinterpret_cast<V*>(a.out + zo);
#pragma unroll
        for (int i = 0; i < NA; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j)
                out[((uint64_t)(hi | e.rdep_a[i] | e.rdep_b[j]) << a.lo_bits) | tid] = acc[i][j];
    }
}

// Chain op: a fused three-input op out[o] = sum_{y, x} A[a] B[b] D[d] (the
// compiler's merge of a contraction T = sum A B with the contraction
// out = sum T D consuming it).  With x = the (at most kChainMaxX lowest)
// summed variables that D does not carry and y = the other summed variables,
// the sum factorises as out[o] = sum_y D[d(o, y)] (sum_x A[a(o, y, x)] B[b(o, y, x)]),
// so each inner sum is formed in registers and T is never stored.
//
// Output bits are split three ways by the host.  Register bits: up to two
// output bits that A carries and B does not (RR), as many of B's, and one that
// only D carries (RD); per (y, x) a thread loads 2^RR values of A and of B and
// forms their 2^2RR products, each multiplied into 2^RD outputs after the
// inner sum.  With PAIR (complex64), A's and B's lowest variables are such
// register bits, so each pair of register values is adjacent and one 16-byte
// load fetches both: the address units take four lanes per cycle whatever the
// width up to 16 bytes, so pairs halve their work, which bounds this kernel.
// Lane bits: the lowest 8 other output bits (thread t's offsets come from
// pdep(t, lane_mask)).  Block bits: the rest, visited through bit runs per
// target.  D (at most 2^kChainMaxD elements) and the offsets of every (y, x)
// are staged in LDS.  Without D (has_d = 0) the kernel runs a plain two-input
// op out = sum_x A B (D is the scalar 1 and y is empty): the same register
// tiles and pair loads for fused two-input ops.
constexpr uint32_t kFusedChain = 1u << 12;
constexpr int kChainMaxX = 4, kChainMaxY = 5, kChainMaxXY = 9, kChainMaxD = 10, kChainRuns = 8;

struct ChainArgs {
    uint32_t roff_a[4], roff_b[4];       // byte offset in A (B) of register pattern p
    uint32_t droff_a[4], droff_b[4], droff_d[2];  // element offset in D of each register pattern
    uint32_t oroff_a[4], oroff_b[4], oroff_d[2];  // element offset in the output of each register pattern
    uint32_t hk_run[4][kChainRuns];      // block bits -> offset bits in A, B, D, out (src | w << 5 | dst << 10)
    uint32_t hk_nrun[4];
    uint32_t lane_mask;                  // union bits of the 8 lane bits
    uint32_t xbits, ybits;               // union bits of the inner (x) and outer (y) summed variables
    uint32_t hk_bits, kx, ky, d_elems, has_d;
};

__device__ __forceinline__ uint32_t chain_runs(const uint32_t* run, uint32_t n, uint32_t hk) {
    uint32_t off = 0;
#pragma unroll
    for (int r = 0; r < kChainRuns; ++r) {
        if ((uint32_t)r >= n) break;
        const uint32_t f = run[r];
        off |= bfe(hk, f & 31u, (f >> 5) & 31u) << (f >> 10);
    }
    return off;
}

template <typename R, int RR, int RD, bool PAIR>
__global__ __launch_bounds__(256) void chain_kernel(const FusedArgs a, const ChainArgs e) {
    using V = typename Cx<R>::V;
    typedef float f32x4 __attribute__((ext_vector_type(4)));
    static_assert(!PAIR || (sizeof(V) == 8 && RR >= 1), "pairs: complex64 with a register bit per input");
    constexpr int NR = 1 << RR, ND = 1 << RD;
    __shared__ V dsh[1 << kChainMaxD];
    // byte offsets in A and B of summed assignment it = y << kx | x, and D's
    // element offset of its outer assignment y
    __shared__ uint4 xy_off[1 << kChainMaxXY];
    constexpr uint32_t EB = sizeof(V);
    const uint32_t tid = threadIdx.x;
    const int64_t zo = (int64_t)blockIdx.z * a.zbytes;       // this block's slice
    if (e.has_d) {
        const V* __restrict__ dsrc = reinterpret_cast<const V*>(a.in[2] + zo);
        for (uint32_t i = tid; i < e.d_elems; i += 256u) dsh[i] = dsrc[i];
    } else if (tid == 0) {
        dsh[0] = V{1, 0};
    }
    const uint32_t n_x = 1u << e.kx, n_y = 1u << e.ky, n_hk = 1u << e.hk_bits;
    const uint32_t n_it = n_y << e.kx;
    for (uint32_t it = tid; it < n_it; it += 256u) {
        const uint32_t uy = pdep_loop(it >> e.kx, e.ybits);
        const uint32_t u = pdep_loop(it & (n_x - 1u), e.xbits) | uy;
        xy_off[it] = make_uint4(pext_loop(u, a.mask[0]) * EB, pext_loop(u, a.mask[1]) * EB,
                                e.has_d ? pext_loop(uy, a.mask[2]) : 0u, 0u);
    }
    __syncthreads();
    // this lane's part of every offset
    const uint32_t ul = pdep_loop(tid, e.lane_mask);
    const uint32_t la = pext_loop(ul, a.mask[0]) * EB, lb = pext_loop(ul, a.mask[1]) * EB;
    const uint32_t ld = pext_loop(ul, a.mask[2]), lo = pext_loop(ul, a.omask);
    // A and B (each below 2 GiB) are read through 32-bit byte offsets from a
    // block-uniform base: one vector add per load
    const char* __restrict__ in_a = a.in[0] + zo;
    const char* __restrict__ in_b = a.in[1] + zo;
    V* out = reinterpret_cast<V*>(a.out + zo);
    uint32_t ra[NR], rb[NR];            // register patterns' byte offsets (uniform)
#pragma unroll
    for (int i = 0; i < NR; ++i) {
        ra[i] = e.roff_a[i];
        rb[i] = e.roff_b[i];
    }
    for (uint32_t hk = blockIdx.x; hk < n_hk; hk += gridDim.x) {
        const uint32_t oa = la + chain_runs(e.hk_run[0], e.hk_nrun[0], hk) * EB;
        const uint32_t ob = lb + chain_runs(e.hk_run[1], e.hk_nrun[1], hk) * EB;
        const uint32_t od = ld + chain_runs(e.hk_run[2], e.hk_nrun[2], hk);
        const uint32_t oo = lo + chain_runs(e.hk_run[3], e.hk_nrun[3], hk);
        V acc[NR][NR][ND];
#pragma unroll
        for (int i = 0; i < NR; ++i)
#pragma unroll
            for (int j = 0; j < NR; ++j)
#pragma unroll
                for (int d = 0; d < ND; ++d) acc[i][j][d] = V{0, 0};
        V t[NR][NR];
#pragma unroll
        for (int i = 0; i < NR; ++i)
#pragma unroll
            for (int j = 0; j < NR; ++j) t[i][j] = V{0, 0};
        // the tiles of summed assignment xo: 2^RR values of A and of B
        auto load = [&](const uint4& xo, V* va, V* vb) {
            const uint32_t xa = oa + xo.x, xb = ob + xo.y;
            if constexpr (PAIR) {
                // register pattern bit 0 is the input's lowest variable: p and
                // p + 1 are adjacent elements
#pragma unroll
                for (int i = 0; i < NR; i += 2) {
                    const f32x4 q = *reinterpret_cast<const f32x4*>(in_a + (xa + ra[i]));
                    va[i] = V{q[0], q[1]};
                    va[i + 1] = V{q[2], q[3]};
                }
#pragma unroll
                for (int j = 0; j < NR; j += 2) {
                    const f32x4 q = *reinterpret_cast<const f32x4*>(in_b + (xb + rb[j]));
                    vb[j] = V{q[0], q[1]};
                    vb[j + 1] = V{q[2], q[3]};
                }
            } else {
#pragma unroll
                for (int i = 0; i < NR; ++i) va[i] = *reinterpret_cast<const V*>(in_a + (xa + ra[i]));
#pragma unroll
                for (int j = 0; j < NR; ++j) vb[j] = *reinterpret_cast<const V*>(in_b + (xb + rb[j]));
            }
        };
        // products into T; after the last inner assignment of an outer one,
        // T times D into the outputs
        auto step = [&](uint32_t it, const uint4& xo, const V* va, const V* vb) {
#pragma unroll
            for (int i = 0; i < NR; ++i)
#pragma unroll
                for (int j = 0; j < NR; ++j) t[i][j] = cfma(t[i][j], va[i], vb[j]);
            if (((it + 1u) & (n_x - 1u)) == 0u) {
                const uint32_t yd = od + xo.z;
#pragma unroll
                for (int i = 0; i < NR; ++i)
#pragma unroll
                    for (int j = 0; j < NR; ++j) {
#pragma unroll
                        for (int d = 0; d < ND; ++d)
                            acc[i][j][d] = cfma(acc[i][j][d], t[i][j],
                                                dsh[yd + e.droff_a[i] + e.droff_b[j] + e.droff_d[d]]);
                        t[i][j] = V{0, 0};
                    }
            }
        };
        // two summed assignments per iteration: both tiles' loads are issued
        // before the first products
        uint32_t it = 0;
        for (; it + 1u < n_it; it += 2u) {
            const uint4 xo0 = xy_off[it], xo1 = xy_off[it + 1u];
            V a0[NR], b0[NR], a1[NR], b1[NR];
            load(xo0, a0, b0);
            load(xo1, a1, b1);
            step(it, xo0, a0, b0);
            step(it + 1u, xo1, a1, b1);
        }
        if (it < n_it) {
            const uint4 xo0 = xy_off[it];
            V a0[NR], b0[NR];
            load(xo0, a0, b0);
            step(it, xo0, a0, b0);
        }
#pragma unroll
        for (int i = 0; i < NR; ++i)
#pragma unroll
            for (int j = 0; j < NR; ++j)
#pragma unroll
                for (int d = 0; d < ND; ++d)
                    out[oo + e.oroff_a[i] + e.oroff_b[j] + e.oroff_d[d]] = acc[i][j][d];
    }
}

// Chain op on the matrix cores (complex64): per wave a 16 x 16 tile of the
// inner sum T[m][n] = sum_x A[m, x] B[x, n] over 4 output bits only A carries
// (m) and 4 only B carries (n), formed by v_mfma_f32_16x16x4f32 in chunks of 4
// inner assignments (four real products per complex chunk: Re += Ar Br - Ai Bi,
// Im += Ar Bi + Ai Br), then multiplied by D on the vector units into 2^RD
// outputs per T value (RD: an output bit only D carries).  Operand layout of
// the instruction (checked on the device): lane k 16 + i supplies A[i][k] and
// B[k][i]; lane (i / 4) 16 + j holds result element i % 4 of row i, column j.
// Each wave of a block takes its own block-bit value hk; D and the offsets of
// every (y, x) are staged in LDS as for chain_kernel.
struct ChainMfmaArgs {
    uint32_t m_mask, n_mask, d_mask;     // union bits of m, n and the D register bit
    uint32_t xbits, ybits;               // inner and outer summed variables
    uint32_t hk_run[4][kChainRuns];      // block bits -> offset bits in A, B, D, out
    uint32_t hk_nrun[4];
    uint32_t hk_bits, kx, ky, d_elems, has_d;
};

template <int RD>
__global__ __launch_bounds__(256) void chain_mfma_kernel(const FusedArgs a, const ChainMfmaArgs e) {
    using V = typename Cx<float>::V;
    typedef float f32x4 __attribute__((ext_vector_type(4)));
    constexpr int ND = 1 << RD;
    constexpr uint32_t EB = sizeof(V);
    __shared__ V dsh[1 << kChainMaxD];
    __shared__ uint4 xy_off[1 << kChainMaxXY];
    const uint32_t tid = threadIdx.x;
    const int64_t zo = (int64_t)blockIdx.z * a.zbytes;       // this block's slice
    if (e.has_d) {
        const V* __restrict__ dsrc = reinterpret_cast<const V*>(a.in[2] + zo);
        for (uint32_t i = tid; i < e.d_elems; i += 256u) dsh[i] = dsrc[i];
    } else if (tid == 0) {
        dsh[0] = V{1, 0};
    }
    const uint32_t n_x = 1u << e.kx, n_y = 1u << e.ky, n_hk = 1u << e.hk_bits;
    const uint32_t n_it = n_y << e.kx;
    for (uint32_t it = tid; it < n_it; it += 256u) {
        const uint32_t uy = pdep_loop(it >> e.kx, e.ybits);
        const uint32_t u = pdep_loop(it & (n_x - 1u), e.xbits) | uy;
        xy_off[it] = make_uint4(pext_loop(u, a.mask[0]) * EB, pext_loop(u, a.mask[1]) * EB,
                                e.has_d ? pext_loop(uy, a.mask[2]) : 0u, 0u);
    }
    __syncthreads();
    const uint32_t lane = tid & 63u, wave = tid >> 6;
    const uint32_t r = lane & 15u, kk = lane >> 4;
    // this lane's operand rows (A: m = r, B: n = r) and result rows m = 4 kk + i
    const uint32_t un = pdep_loop(r, e.n_mask);
    const uint32_t la = pext_loop(pdep_loop(r, e.m_mask), a.mask[0]) * EB;
    const uint32_t lb = pext_loop(un, a.mask[1]) * EB;
    const uint32_t dn = pext_loop(un, a.mask[2]), on = pext_loop(un, a.omask);
    uint32_t dm[4], om[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint32_t um = pdep_loop(kk * 4u + (uint32_t)i, e.m_mask);
        dm[i] = pext_loop(um, a.mask[2]);
        om[i] = pext_loop(um, a.omask);
    }
    uint32_t dmd[4][ND], od_[ND];        // D offsets of (result row i, D bit d)
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        const uint32_t ud = d ? e.d_mask : 0u;
        const uint32_t dd = pext_loop(ud, a.mask[2]);
        od_[d] = pext_loop(ud, a.omask);
#pragma unroll
        for (int i = 0; i < 4; ++i) dmd[i][d] = dm[i] + dd;
    }
    const char* __restrict__ in_a = a.in[0] + zo;
    const char* __restrict__ in_b = a.in[1] + zo;
    V* out = reinterpret_cast<V*>(a.out + zo);
    const uint32_t n_ch = n_x >> 2;         // chunks of 4 inner assignments (kx >= 2)
    for (uint32_t hk = blockIdx.x * 4u + wave; hk < n_hk; hk += gridDim.x * 4u) {
        const uint32_t oa = la + chain_runs(e.hk_run[0], e.hk_nrun[0], hk) * EB;
        const uint32_t ob = lb + chain_runs(e.hk_run[1], e.hk_nrun[1], hk) * EB;
        const uint32_t od = dn + chain_runs(e.hk_run[2], e.hk_nrun[2], hk);
        const uint32_t oo = on + chain_runs(e.hk_run[3], e.hk_nrun[3], hk);
        V acc[4][ND];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int d = 0; d < ND; ++d) acc[i][d] = V{0, 0};
        for (uint32_t y = 0; y < n_y; ++y) {
            V av[4], bv[4];
            uint32_t yd = 0;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                if ((uint32_t)c >= n_ch) break;
                const uint4 xo = xy_off[(y << e.kx) | ((uint32_t)c * 4u + kk)];
                yd = xo.z;
                av[c] = *reinterpret_cast<const V*>(in_a + (oa + xo.x));
                bv[c] = *reinterpret_cast<const V*>(in_b + (ob + xo.y));
            }
            f32x4 tr = {0.f, 0.f, 0.f, 0.f}, ti = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                if ((uint32_t)c >= n_ch) break;
                tr = __builtin_amdgcn_mfma_f32_16x16x4f32(av[c].x, bv[c].x, tr, 0, 0, 0);
                ti = __builtin_amdgcn_mfma_f32_16x16x4f32(av[c].x, bv[c].y, ti, 0, 0, 0);
                tr = __builtin_amdgcn_mfma_f32_16x16x4f32(-av[c].y, bv[c].y, tr, 0, 0, 0);
                ti = __builtin_amdgcn_mfma_f32_16x16x4f32(av[c].y, bv[c].x, ti, 0, 0, 0);
            }
            const V* dy = dsh + (od + yd);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int d = 0; d < ND; ++d) acc[i][d] = cfma(acc[i][d], V{tr[i], ti[i]}, dy[dmd[i][d]]);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int d = 0; d < ND; ++d) out[oo + om[i] + od_[d]] = acc[i][d];
    }
}

// Runs of gate-only ops (wide plans): fused ops that read gate tensors only
// are independent of each other (their outputs lie past the gate area), and
// the plan compiler puts them first.  One launch runs a run of them, one
// block per op (grid x) and slice (grid z): each op is decoded from its record
// in the device copy of the plan (checked on the host before the launch) and
// evaluated from its offset tables, thread t taking outputs t, t + 256, ...;
// the part of an offset above the table's low output bits is a pext of the
// output's high bits.
constexpr int kRunMaxOps = 64;
constexpr uint32_t kRunMaxTerms = 1u << 14;    // outputs x summed assignments x inputs of one op
constexpr uint32_t kLocBits = 0x7fffffffu;     // a location without its global-memory flag

struct OpRun {
    int32_t pos[kRunMaxOps];    // record positions in the plan
    int32_t n;
};

template <typename R>
__global__ __launch_bounds__(256) void gate_ops_kernel(const int32_t* __restrict__ plan, const OpRun run,
                                                       char* __restrict__ arena, int64_t zbytes) {
    using V = typename Cx<R>::V;
    char* base = arena + (int64_t)blockIdx.z * zbytes;       // this block's slice
    const int32_t* r = plan + run.pos[blockIdx.x];
    const uint32_t n_in = (uint32_t)r[0] & 15u, meta = (uint32_t)r[1];
    const uint32_t rank_out = (meta >> 8) & 255u, k = (meta >> 16) & 255u, lo_bits = meta >> 24;
    const uint32_t omask = (uint32_t)r[4];
    const int32_t* lo_tab = r + kFusedHdr + 4 * n_in;
    const int32_t* x_tab = lo_tab + (n_in << lo_bits);
    V* out = reinterpret_cast<V*>(base) + ((uint32_t)r[2] & kLocBits);
    for (uint32_t o = threadIdx.x; o < (1u << rank_out); o += 256u) {
        const uint32_t olo = o & ((1u << lo_bits) - 1u), ohi = o >> lo_bits;
        const uint32_t uhi = ohi ? pdep_loop(ohi << lo_bits, omask) : 0u;
        const V* in[kFusedMaxIn];
        for (uint32_t j = 0; j < n_in; ++j) {
            const uint32_t loc = (uint32_t)r[kFusedHdr + 4 * j] & kLocBits;
            const uint32_t mask = (uint32_t)r[kFusedHdr + 4 * j + 1];
            in[j] = reinterpret_cast<const V*>(base) + loc + (uint32_t)lo_tab[olo * n_in + j] +
                    (ohi ? pext_loop(uhi, mask) : 0u);
        }
        V acc = {0, 0};
        for (uint32_t x = 0; x < (1u << k); ++x) {
            const int32_t* xt = x_tab + x * n_in;
            V p = in[0][xt[0]];
            for (uint32_t j = 1; j < n_in; ++j) p = cmul(p, in[j][xt[j]]);
            acc = cadd(acc, p);
        }
        out[o] = acc;
    }
}

// ---------------------------------------------------------------------------
// Batched executor
// ---------------------------------------------------------------------------

// Plan layout (int32 words; offsets relative to the plan start):
//   header[kPlanHdr]: n_gates, n_ops, gates_off, ops_off, result_loc,
//                     lds_need, glb_need (low word), n_sliced, ops_words,
//                     glb_need (high word), 0, 0
//   gates: n_gates * kGateWords (distinct gate tensors only; equal leaves share one)
//   ops:   one record per op, read from global memory one op ahead of use
// LDS element 0 holds 1+0i, element 1 holds 0+0i, element 2 is a write sink;
// gate tensors start at element 4.
//
// Step op (one bucket step, all lanes of the item):
//   w0 = n_in | rank_out << 4 | sum << 9 | sync_glb << 11 | all_lds << 12 | len << 16
//   w1 = 0, w2 = out_loc
//   per input j, 4 words at 3 + 4 j: loc, top, mask, runs
//     runs = n_runs << 30 | run0 | run1 << 15, run = src | width << 5 | dst << 10;
//     n_runs == 3 means "more than two runs": the offset is pext(o, mask).
// Gather op (several small steps, one output element per lane, all in LDS):
//   w0 = kOpGather | n_steps | len << 16, w1 = active lanes n, w2 = lane table
//   offset, w3 = NI, the largest input count of its steps
//   lane table: n entries of NI + 1 words, entry l at word l * (NI + 1):
//     word 0: the LDS element lane l writes; word 1 + j: the LDS elements of
//     input j at the summed variable's 0 (bits 0-15) and 1 (bits 16-31).  A step
//     without a sum names element 1 (0+0i) as its first input's second element;
//     unused inputs name element 0 (1+0i).
// Reduction op (a chain of single-input sums over an LDS tensor's top k
// variables, one wave):
//   w0 = kOpReduce | rank_out << 4 | sync_glb << 11 | 4 << 16, w1 = k,
//   w2 = out_loc, w3 = in_loc; out[o] = sum over x < 2^k of in[o + x 2^rank_out]
constexpr int kPlanHdr = 12;
constexpr int kGateWords = 16;
constexpr int kInWords = 4;
constexpr int kStepHdr = 3;
constexpr int kExecMaxIn = 4;
constexpr uint32_t kGlobal = 0x80000000u;
constexpr uint32_t kOpGather = 1u << 15;
// LDS elements every item holds (written before its ops; LOC_ONE / LOC_ZERO in
// qtensor_amd/plan.py): 1 + 0i and 0 + 0i
constexpr uint32_t kLocOne = 0u, kLocZero = 1u;
constexpr int kGatherHdr = 4;
constexpr int kReduceHdr = 4;
// gate kinds (shared with qtensor_amd/gates.py and qtensor_amd/plan.py)
enum GateKind : int32_t {
    kConst = 0, kZPhase = 1, kXPhase = 2, kZZ = 3, kYPhase = 4,
    kH = 5, kCX = 6, kZ = 7, kM = 8,
};

// Gate entry words:
//  0 kind, 1 dst (LDS element offset), 2 n_elems, 3 src_map (4 bits per element),
//  4 j0, 5 c0 (float bits), 6 j1, 7 c1 (float bits), 8-9 alpha0 (double bits),
//  10 const_off, 11 n_slice | var0 << 8 | var1 << 16, 12 stride0, 13 stride1
// One sincospi evaluation serves every kind (the lanes of a wave build gates
// of different kinds; per-kind calls would run one after another).
__device__ void natural_gate(int32_t kind, double alpha, double* re, double* im) {
    for (int e = 0; e < 8; ++e) { re[e] = 0.0; im[e] = 0.0; }
    const double h = 0.70710678118654752440;
    const bool half = kind == kXPhase || kind == kYPhase || kind == kZZ;
    double s, c;
    sincospi(half ? alpha * 0.5 : alpha, &s, &c);
    switch (kind) {
    case kZPhase:
        re[0] = 1.0; re[1] = c; im[1] = s;
        break;
    case kXPhase:
    case kYPhase: {
        // e^{i pi a/2} [[c, -i s], [-i s, c]]  /  e^{i pi a/2} [[c, -s], [s, c]]
        const double gs = s, gc = c;  // global phase e^{i pi a/2} = gc + i gs
        re[0] = gc * c; im[0] = gs * c;
        re[3] = gc * c; im[3] = gs * c;
        if (kind == kXPhase) {
            // (gc + i gs)(-i s) = gs s - i gc s
            re[1] = gs * s; im[1] = -gc * s;
            re[2] = gs * s; im[2] = -gc * s;
        } else {
            re[1] = -gc * s; im[1] = -gs * s;
            re[2] = gc * s; im[2] = gs * s;
        }
        break;
    }
    case kZZ:
        re[0] = c; im[0] = -s;  // m
        re[1] = c; im[1] = s;   // p
        re[2] = c; im[2] = s;
        re[3] = c; im[3] = -s;
        break;
    case kH:
        re[0] = h; re[1] = h; re[2] = h; re[3] = -h;
        break;
    case kCX:  // [c][t_out][t_in] = (t_out == t_in ^ c)
        re[0] = 1.0; re[3] = 1.0; re[5] = 1.0; re[6] = 1.0;
        break;
    case kZ:
        re[0] = 1.0; re[1] = -1.0;
        break;
    case kM:
        re[0] = 1.0; re[3] = 1.0;
        break;
    default:
        break;
    }
}

// Angles of a work item: the executor loads an item's first kParamRegs angles
// with one uniform (scalar) load, so the lanes of the grid do not all read one
// cache line, and each wave keeps a copy in LDS (the params buffer is padded so
// the load stays inside it).  Higher indices are read from global memory.
constexpr int kParamRegs = 8;

__device__ __forceinline__ double angle_at(const QB_LDS double* lds_angles, int32_t n_lds, const double* params,
                                           int32_t j) {
    if (j < n_lds) return lds_angles[j];
    return params[j];
}

template <typename R, typename Dst>
__device__ __forceinline__ void build_gate(const int32_t* g, const double* params,
                                           const QB_LDS double* lds_angles, int32_t n_lds, uint32_t slice_bits,
                                           Dst lds, const typename Cx<R>::V* pool) {
    using V = typename Cx<R>::V;
    const int32_t kind = g[0];
    const int32_t dst = g[1];
    const int32_t n = g[2];
    const uint32_t src_map = (uint32_t)g[3];
    uint32_t shift = 0;
    const int32_t sl = g[11];
    const int32_t n_slice = sl & 0xff;
    if (n_slice > 0) shift += ((slice_bits >> ((sl >> 8) & 0xff)) & 1u) * (uint32_t)g[12];
    if (n_slice > 1) shift += ((slice_bits >> ((sl >> 16) & 0xff)) & 1u) * (uint32_t)g[13];
    if (kind == kConst) {
        for (int e = 0; e < n; ++e) lds[dst + e] = pool[g[10] + ((src_map >> (4 * e)) & 15u) + shift];
        return;
    }
    double alpha = __hiloint2double(g[9], g[8]);
    if (g[4] >= 0) alpha += (double)__int_as_float(g[5]) * angle_at(lds_angles, n_lds, params, g[4]);
    if (g[6] >= 0) alpha += (double)__int_as_float(g[7]) * angle_at(lds_angles, n_lds, params, g[6]);
    double re[8], im[8];
    natural_gate(kind, alpha, re, im);
    for (int e = 0; e < n; ++e) {
        const uint32_t s = ((src_map >> (4 * e)) & 15u) + shift;
        V v;
        v.x = (R)re[s & 7];
        v.y = (R)im[s & 7];
        lds[dst + e] = v;
    }
}

// Gate g for every item of a group whose items share one angle vector: the
// FP64 values are computed once, and item i's copy (its slice shift) goes to
// element e * M + i of the interleaved LDS data.
template <typename R, int M>
__device__ __forceinline__ void build_gate_group(const int32_t* g, const double* params,
                                                 const QB_LDS double* lds_angles, int32_t n_lds,
                                                 const uint32_t* slice_bits, QB_LDS typename Cx<R>::V* lds_v,
                                                 const typename Cx<R>::V* pool) {
    using V = typename Cx<R>::V;
    const int32_t kind = g[0];
    const int32_t dst = g[1];
    const int32_t n = g[2];
    const uint32_t src_map = (uint32_t)g[3];
    const int32_t sl = g[11];
    const int32_t n_slice = sl & 0xff;
    uint32_t shift[M];
#pragma unroll
    for (int i = 0; i < M; ++i) {
        shift[i] = 0;
        if (n_slice > 0) shift[i] += ((slice_bits[i] >> ((sl >> 8) & 0xff)) & 1u) * (uint32_t)g[12];
        if (n_slice > 1) shift[i] += ((slice_bits[i] >> ((sl >> 16) & 0xff)) & 1u) * (uint32_t)g[13];
    }
    if (kind == kConst) {
        for (int e = 0; e < n; ++e)
#pragma unroll
            for (int i = 0; i < M; ++i)
                lds_v[(uint32_t)(dst + e) * M + i] = pool[g[10] + ((src_map >> (4 * e)) & 15u) + shift[i]];
        return;
    }
    double alpha = __hiloint2double(g[9], g[8]);
    if (g[4] >= 0) alpha += (double)__int_as_float(g[5]) * angle_at(lds_angles, n_lds, params, g[4]);
    if (g[6] >= 0) alpha += (double)__int_as_float(g[7]) * angle_at(lds_angles, n_lds, params, g[6]);
    double re[8], im[8];
    natural_gate(kind, alpha, re, im);
    for (int e = 0; e < n; ++e) {
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const uint32_t s = ((src_map >> (4 * e)) & 15u) + shift[i];
            V v;
            v.x = (R)re[s & 7];
            v.y = (R)im[s & 7];
            lds_v[(uint32_t)(dst + e) * M + i] = v;
        }
    }
}

struct ExecArgs {
    const int32_t* program;
    uint32_t program_bytes;
    const int32_t* items;
    int32_t n_items;
    int32_t lds_elems;
    const double* params;
    const void* pool;
    void* scratch;
    int64_t scratch_elems;
    double* item_out;
    uint32_t* queue;
};

template <int ITEM_THREADS>
__device__ __forceinline__ void group_sync() {
    if constexpr (ITEM_THREADS == 64) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    } else {
        __syncthreads();
    }
}

// Global scratch of an executor item through a buffer resource (base = the
// item's slot, bounds = its size in bytes): out-of-range accesses read zero
// and drop writes instead of faulting, and buffer instructions stay distinct
// from LDS accesses, so both kinds of load can sit in one straight-line pass.
constexpr uint32_t kBufferDword3 = 0x00020000u;
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ cf32 buf_load(__amdgpu_buffer_rsrc_t r, uint32_t elem, cf32*) {
    return __builtin_bit_cast(cf32, __builtin_amdgcn_raw_buffer_load_b64(r, elem * 8u, 0, 0));
}
__device__ __forceinline__ cf64 buf_load(__amdgpu_buffer_rsrc_t r, uint32_t elem, cf64*) {
    return __builtin_bit_cast(cf64, __builtin_amdgcn_raw_buffer_load_b128(r, elem * 16u, 0, 0));
}
__device__ __forceinline__ void buf_store(__amdgpu_buffer_rsrc_t r, uint32_t elem, cf32 v) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, elem * 8u, 0, 0);
}
__device__ __forceinline__ void buf_store(__amdgpu_buffer_rsrc_t r, uint32_t elem, cf64 v) {
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, v), r, elem * 16u, 0, 0);
}
// a group's element (two complex64 values): one 16-byte access
__device__ __forceinline__ cf32x2 buf_load(__amdgpu_buffer_rsrc_t r, uint32_t elem, cf32x2*) {
    return __builtin_bit_cast(cf32x2, __builtin_amdgcn_raw_buffer_load_b128(r, elem * 16u, 0, 0));
}
__device__ __forceinline__ void buf_store(__amdgpu_buffer_rsrc_t r, uint32_t elem, cf32x2 v) {
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, v), r, elem * 16u, 0, 0);
}

// Item i's view of a group's interleaved LDS data: element e at e * M + i
// (the destination of its gate tensors).
template <typename V, int M>
struct LdsLane {
    QB_LDS V* p;
    __device__ __forceinline__ QB_LDS V& operator[](int32_t e) const { return p[(uint32_t)e * (uint32_t)M]; }
};

// Offset of union element o in an input whose mask has at most two runs: one
// bit-field extract and one shift per run (an absent run has width 0 and adds 0).
__device__ __forceinline__ uint32_t run_offset(uint32_t o, uint32_t runs) {
    const uint32_t a = bfe(o, runs & 31u, (runs >> 5) & 31u) << ((runs >> 10) & 31u);
    const uint32_t b = bfe(o, (runs >> 15) & 31u, (runs >> 20) & 31u) << ((runs >> 25) & 31u);
    return a | b;
}

__device__ __forceinline__ uint32_t in_offset(uint32_t o, uint32_t mask, uint32_t runs) {
    if ((runs >> 30) == 3u) return pext_loop(o, mask);
    return run_offset(o, runs);
}

#define QB_RL(r, k) ((uint32_t)__builtin_amdgcn_readlane((r), (k)))

// Op records are read from global memory (they stay cached) ahead of use,
// through a buffer resource over the program (`prog`): offsets are 32-bit and
// reads past the program return 0.  `st` is an op's word offset in the program.
// Lane l fetches word (l & 63) of the record.
__device__ __forceinline__ int32_t fetch_rec(__amdgpu_buffer_rsrc_t prog, uint32_t st, int lane) {
    return __builtin_amdgcn_raw_buffer_load_b32(prog, (uint32_t)(lane & 63) * 4u, st * 4u, 0);
}

// A lane's words of a gather op's lane table: entry l holds NI + 1 words.
static_assert(kExecMaxIn == 4, "fetch_lane reads four input words per lane");
struct GatherLane {
    uint32_t w[kExecMaxIn + 1];
};

// Lane l reads five words from its entry's start whatever the op's input
// count (one 16-byte and one 4-byte load; words past the entry belong to later
// entries or records, or to the zero words after the last plan, and are not
// used); lanes past the table read later records or, past the program, zeros,
// and skip the op.
__device__ __forceinline__ void fetch_lane(__amdgpu_buffer_rsrc_t prog, uint32_t st, int32_t r, int lane,
                                           uint32_t wave_lane0, uint32_t n, uint32_t ni, GatherLane& g) {
    // a wave none of whose lanes is active loads nothing (the op skips it)
    if (wave_lane0 >= n) return;
    const uint32_t so = (st + QB_RL(r, 2)) * 4u;
    const uint32_t vo = __umul24((uint32_t)lane, (ni + 1u) * 4u);   // lane < 2^10
    const auto a = __builtin_amdgcn_raw_buffer_load_b128(prog, vo, so, 0);
#pragma unroll
    for (int j = 0; j < 4; ++j) g.w[j] = (uint32_t)a[j];
    g.w[4] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(prog, vo + 16u, so, 0);
}

// One gather op: lane l computes its table entry's element,
// x_0 ... x_{NI-1} + y_0 ... y_{NI-1} (y_0 = 0 for a step without a sum).
template <typename P, int NI>
__device__ __forceinline__ void gather_op_n(const GatherLane& g, QB_LDS P* lds) {
    P x[NI], y[NI];
#pragma unroll
    for (int j = 0; j < NI; ++j) {
        x[j] = lds[g.w[1 + j] & 0xffffu];
        y[j] = lds[g.w[1 + j] >> 16];
    }
    // the summed term's last product is fused into the sum
    P out;
    if (NI == 1) {
        out = cadd(x[0], y[0]);
    } else if (NI == 2) {
        out = cfma(cmul(x[0], x[1]), y[0], y[1]);
    } else if (NI == 3) {
        out = cfma(cmul(cmul(x[0], x[1]), x[2]), cmul(y[0], y[1]), y[2]);
    } else {
        out = cfma(cmul(cmul(x[0], x[1]), cmul(x[2], x[3])), cmul(y[0], y[1]), cmul(y[2], y[3]));
    }
    lds[g.w[0]] = out;
}

// n: the op's active lanes, ni: its input count (both decoded from the record
// when its lane table was fetched)
template <typename P>
__device__ __forceinline__ void gather_op(uint32_t n, uint32_t ni, const GatherLane& g, QB_LDS P* lds, int lane) {
    if ((uint32_t)lane >= n) return;
    switch (ni) {
    case 1: gather_op_n<P, 1>(g, lds); break;
    case 2: gather_op_n<P, 2>(g, lds); break;
    case 3: gather_op_n<P, 3>(g, lds); break;
    default: gather_op_n<P, 4>(g, lds); break;
    }
}

// Input offsets of a step split at the item's lane bits: output element
// o = base + lane with base a multiple of ITEM_THREADS (a power of two) and
// lane < ITEM_THREADS share no bit, so pext(o, mask) = pext(lane, mask) |
// pext(base, mask).  The lane part is computed once per step; the base part is
// uniform and is computed by the scalar unit once per block of outputs.

// One reduction op, run by the item's first wave (n_out = 2^rank_out <= 64):
// lane l takes output o = l mod n_out and every (64 / n_out)-th term from
// l / n_out on; the lanes sharing an output then combine in a butterfly.
template <typename P>
__device__ __forceinline__ void reduce_op(int32_t r, uint32_t w0, QB_LDS P* lds, int lane) {
    if (lane >= 64) return;
    const uint32_t ro = (w0 >> 4) & 31u;
    const uint32_t terms = 1u << (QB_RL(r, 1) & 31u);
    const uint32_t out_loc = QB_RL(r, 2);
    const uint32_t in_loc = QB_RL(r, 3);
    const uint32_t n_out = 1u << ro;
    const uint32_t o = (uint32_t)lane & (n_out - 1u);
    const uint32_t g = (uint32_t)lane >> ro;
    const uint32_t step = ro < 6u ? 64u >> ro : 1u;   // plans keep rank_out <= 6
    // terms and step are powers of two: every lane makes terms / step trips,
    // or one when terms < step, where lanes past the terms add element
    // LOC_ZERO (0 + 0i) -- a uniform loop with no per-lane exit
    const uint32_t trips = terms >= step ? terms / step : 1u;
    P s = {};
#pragma unroll 4
    for (uint32_t t = 0; t < trips; ++t) {
        const uint32_t x = g + t * step;
        s = cadd(s, lds[x < terms ? in_loc + o + x * n_out : kLocZero]);
    }
    for (uint32_t b = n_out; b < 64u; b <<= 1) s = pk_xor_add(s, (int)b);
    if (g == 0) lds[out_loc + o] = s;
}

// Output blocks [base, base + U*ITEM_THREADS) of an all-LDS step with NI inputs,
// for every base below n_out in steps of U*ITEM_THREADS (n_out is a multiple of
// that when U > 1).
template <typename P, int ITEM_THREADS, int U, int NI>
__device__ __forceinline__ void step_lds_pass(uint32_t n_out, uint32_t n_sum, const uint32_t* loc,
                                              const uint32_t* top, const uint32_t* mask, const uint32_t* runs,
                                              uint32_t out_loc, QB_LDS P* lds, int lane) {
    uint32_t la[NI];
#pragma unroll
    for (int j = 0; j < NI; ++j) la[j] = loc[j] + in_offset((uint32_t)lane, mask[j], runs[j]);
    for (uint32_t base = 0; base < n_out; base += U * ITEM_THREADS) {
        P p[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t hb = base + u * ITEM_THREADS;
            P x[NI], y[NI];
#pragma unroll
            for (int j = 0; j < NI; ++j) {
                const uint32_t a = la[j] + in_offset(hb, mask[j], runs[j]);
                x[j] = lds[a];
                if (n_sum == 2) y[j] = lds[a + top[j]];
            }
            P q;
            if (NI == 1) {
                p[u] = x[0];
                if (n_sum == 2) q = y[0];
            } else if (NI == 2) {
                p[u] = cmul(x[0], x[1]);
                if (n_sum == 2) q = cmul(y[0], y[1]);
            } else if (NI == 3) {
                p[u] = cmul(cmul(x[0], x[1]), x[2]);
                if (n_sum == 2) q = cmul(cmul(y[0], y[1]), y[2]);
            } else {
                p[u] = cmul(cmul(x[0], x[1]), cmul(x[2], x[3]));
                if (n_sum == 2) q = cmul(cmul(y[0], y[1]), cmul(y[2], y[3]));
            }
            if (n_sum == 2) p[u] = cadd(p[u], q);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t o = base + u * ITEM_THREADS + (uint32_t)lane;
            if (U > 1 || o < n_out) lds[out_loc + o] = p[u];
        }
    }
}

template <typename P, int ITEM_THREADS, int NI>
__device__ __forceinline__ void step_lds_n(uint32_t n_out, uint32_t n_sum, const uint32_t* loc,
                                           const uint32_t* top, const uint32_t* mask, const uint32_t* runs,
                                           uint32_t out_loc, QB_LDS P* lds, int lane) {
    // a group of items unrolls less (its values are wider)
    constexpr bool wide = sizeof(P) > 16;
    if (NI < 4 && !wide && n_out >= 4u * ITEM_THREADS)
        step_lds_pass<P, ITEM_THREADS, 4, NI>(n_out, n_sum, loc, top, mask, runs, out_loc, lds, lane);
    else if (n_out >= 2u * ITEM_THREADS)
        step_lds_pass<P, ITEM_THREADS, 2, NI>(n_out, n_sum, loc, top, mask, runs, out_loc, lds, lane);
    else if (((uint32_t)__builtin_amdgcn_readfirstlane(lane) & ~63u) < n_out)   // waves past the outputs idle
        step_lds_pass<P, ITEM_THREADS, 1, NI>(n_out, n_sum, loc, top, mask, runs, out_loc, lds, lane);
}

// Output blocks of a step whose inputs or output live in global scratch
// (accessed through the slot's buffer resource `srd`).
template <typename P, int ITEM_THREADS, int U, int NI>
__device__ __forceinline__ void step_mixed_pass(uint32_t n_out, uint32_t n_sum,
                                                const uint32_t* loc, const uint32_t* top, const uint32_t* mask,
                                                const uint32_t* runs, uint32_t out_loc,
                                                QB_LDS P* lds, __amdgpu_buffer_rsrc_t srd, int lane) {
    uint32_t la[NI];
#pragma unroll
    for (int j = 0; j < NI; ++j) la[j] = (loc[j] & ~kGlobal) + in_offset((uint32_t)lane, mask[j], runs[j]);
    for (uint32_t base = 0; base < n_out; base += U * ITEM_THREADS) {
        P p[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t hb = base + u * ITEM_THREADS;
            P x[NI], y[NI];
#pragma unroll
            for (int j = 0; j < NI; ++j) {
                const uint32_t a = la[j] + in_offset(hb, mask[j], runs[j]);
                if (loc[j] & kGlobal) {
                    x[j] = buf_load(srd, a, (P*)nullptr);
                    if (n_sum == 2) y[j] = buf_load(srd, a + top[j], (P*)nullptr);
                } else {
                    x[j] = lds[a];
                    if (n_sum == 2) y[j] = lds[a + top[j]];
                }
            }
            P q;
            if (NI == 1) {
                p[u] = x[0];
                if (n_sum == 2) q = y[0];
            } else if (NI == 2) {
                p[u] = cmul(x[0], x[1]);
                if (n_sum == 2) q = cmul(y[0], y[1]);
            } else if (NI == 3) {
                p[u] = cmul(cmul(x[0], x[1]), x[2]);
                if (n_sum == 2) q = cmul(cmul(y[0], y[1]), y[2]);
            } else {
                p[u] = cmul(cmul(x[0], x[1]), cmul(x[2], x[3]));
                if (n_sum == 2) q = cmul(cmul(y[0], y[1]), cmul(y[2], y[3]));
            }
            if (n_sum == 2) p[u] = cadd(p[u], q);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t o = base + u * ITEM_THREADS + (uint32_t)lane;
            if (U > 1 || o < n_out) {
                if (out_loc & kGlobal)
                    buf_store(srd, (out_loc & ~kGlobal) + o, p[u]);
                else
                    lds[out_loc + o] = p[u];
            }
        }
    }
}

template <typename P, int ITEM_THREADS, int NI>
__device__ __forceinline__ void step_mixed_n(uint32_t n_out, uint32_t n_sum, const uint32_t* loc,
                                             const uint32_t* top, const uint32_t* mask, const uint32_t* runs,
                                             uint32_t out_loc, QB_LDS P* lds, __amdgpu_buffer_rsrc_t srd, int lane) {
    constexpr bool wide = sizeof(P) > 16;
    if (!wide && n_out >= 4u * ITEM_THREADS)
        step_mixed_pass<P, ITEM_THREADS, 4, NI>(n_out, n_sum, loc, top, mask, runs, out_loc, lds, srd, lane);
    else
        step_mixed_pass<P, ITEM_THREADS, 1, NI>(n_out, n_sum, loc, top, mask, runs, out_loc, lds, srd, lane);
}

// A step op with global scratch: every input is uniformly in LDS or in the
// item's scratch slot; reads of a pass are issued before its products.  Only
// the executor variant for batches with global scratch contains it (inline: a
// call would save and restore the callee-saved registers it uses every time).
template <typename P, int ITEM_THREADS>
__device__ __forceinline__ void step_mixed(int32_t r, uint32_t w0, QB_LDS P* lds, __amdgpu_buffer_rsrc_t srd,
                                           int lane) {
    const uint32_t out_loc = QB_RL(r, 2);
    const uint32_t n_in = w0 & 15u;
    const uint32_t n_out = 1u << ((w0 >> 4) & 31u);
    const uint32_t n_sum = ((w0 >> 9) & 1u) + 1u;
    uint32_t loc[kExecMaxIn], top[kExecMaxIn], mask[kExecMaxIn], runs[kExecMaxIn];
#pragma unroll
    for (int j = 0; j < kExecMaxIn; ++j) {
        if ((uint32_t)j < n_in) {
            loc[j] = QB_RL(r, 3 + 4 * j);
            top[j] = QB_RL(r, 4 + 4 * j);
            mask[j] = QB_RL(r, 5 + 4 * j);
            runs[j] = QB_RL(r, 6 + 4 * j);
        } else {
            loc[j] = 0u; top[j] = 0u; mask[j] = 0u; runs[j] = 0u;
        }
    }
    switch (n_in) {
    case 1: step_mixed_n<P, ITEM_THREADS, 1>(n_out, n_sum, loc, top, mask, runs, out_loc, lds, srd, lane); break;
    case 2: step_mixed_n<P, ITEM_THREADS, 2>(n_out, n_sum, loc, top, mask, runs, out_loc, lds, srd, lane); break;
    case 3: step_mixed_n<P, ITEM_THREADS, 3>(n_out, n_sum, loc, top, mask, runs, out_loc, lds, srd, lane); break;
    default: step_mixed_n<P, ITEM_THREADS, 4>(n_out, n_sum, loc, top, mask, runs, out_loc, lds, srd, lane); break;
    }
}

// One step op: all lanes of the item cover the step's output elements.  Lane l
// of each wave holds word l of the record; fields reach scalar registers by
// readlane.
template <typename P, int ITEM_THREADS>
__device__ __forceinline__ void step_op(int32_t r, uint32_t w0, QB_LDS P* lds, int lane) {
    const uint32_t out_loc = QB_RL(r, 2);
    const uint32_t n_in = w0 & 15u;
    const uint32_t rank_out = (w0 >> 4) & 31u;
    const uint32_t n_sum = ((w0 >> 9) & 1u) + 1u;
    uint32_t loc[kExecMaxIn], top[kExecMaxIn], mask[kExecMaxIn], runs[kExecMaxIn];
#pragma unroll
    for (int j = 0; j < kExecMaxIn; ++j) {
        if ((uint32_t)j < n_in) {
            loc[j] = QB_RL(r, 3 + 4 * j);
            top[j] = QB_RL(r, 4 + 4 * j);
            mask[j] = QB_RL(r, 5 + 4 * j);
            runs[j] = QB_RL(r, 6 + 4 * j);
        } else {
            loc[j] = 0u; top[j] = 0u; mask[j] = 0u; runs[j] = 0u;
        }
    }
    const uint32_t n_out = 1u << rank_out;
    // all-LDS: up to four blocks of ITEM_THREADS outputs per pass, every read
    // of a pass issued before its first product and its stores (outputs never
    // alias inputs); specialised by input count
    switch (n_in) {
    case 1: step_lds_n<P, ITEM_THREADS, 1>(n_out, n_sum, loc, top, mask, runs, out_loc, lds, lane); break;
    case 2: step_lds_n<P, ITEM_THREADS, 2>(n_out, n_sum, loc, top, mask, runs, out_loc, lds, lane); break;
    case 3: step_lds_n<P, ITEM_THREADS, 3>(n_out, n_sum, loc, top, mask, runs, out_loc, lds, lane); break;
    default: step_lds_n<P, ITEM_THREADS, 4>(n_out, n_sum, loc, top, mask, runs, out_loc, lds, lane); break;
    }
}

// Synchronisation before an op.  Only LDS data flows between ops unless the op
// is marked (bit 11) as touching global scratch written since the last marked op.
template <int ITEM_THREADS>
__device__ __forceinline__ void op_sync(uint32_t w0) {
    const bool glb = (w0 & (1u << 11)) != 0;
    if constexpr (ITEM_THREADS == 64) {
        // one wave: LDS operations complete in issue order, so the previous op's
        // LDS results are visible once the compiler keeps memory operations on
        // their side of this point; global data needs a real fence
        if (glb) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (glb) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    } else {
        if (glb) {
            __syncthreads();
        } else {
            // LDS-only fence: record prefetches in flight are not waited for
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
            __builtin_amdgcn_s_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
        }
    }
}

// All ops of one plan.  Software pipeline over the records: while op k runs,
// op k+1's lane table words (gather ops) and op k+2's record word are in
// flight.  r0 is op 0's record word, fetched by the caller.  One copy of the
// op bodies keeps the kernel inside the instruction cache.  GLB = false is the
// variant for batches without global scratch (every op all-LDS): no
// global-scratch step and no global fences.
template <typename P, int ITEM_THREADS, bool GLB>
__device__ __forceinline__ void run_ops(__amdgpu_buffer_rsrc_t prog, uint32_t st0, int32_t n_ops, int32_t r0,
                                        QB_LDS P* lds, __amdgpu_buffer_rsrc_t srd, int lane) {
    if (n_ops <= 0) return;
    const uint32_t wave_lane0 = (uint32_t)__builtin_amdgcn_readfirstlane(lane) & ~63u;
    uint32_t w0 = (uint32_t)__builtin_amdgcn_readfirstlane(r0);
    GatherLane g0;   // loaded and read only for gather ops
    uint32_t n0 = 0, ni0 = 0;
    if (w0 & kOpGather) {
        n0 = QB_RL(r0, 1);
        ni0 = QB_RL(r0, 3);
        fetch_lane(prog, st0, r0, lane, wave_lane0, n0, ni0, g0);
    }
    uint32_t st1 = st0 + (w0 >> 16);
    int32_t r1 = 0;
    if (n_ops > 1) r1 = fetch_rec(prog, st1, lane);
    for (int32_t k = 0; k < n_ops; ++k) {
        uint32_t w1 = 0, n1 = 0, ni1 = 0;
        GatherLane g1;
        uint32_t st2 = st1;
        int32_t r2 = 0;
        if (k + 1 < n_ops) {
            w1 = (uint32_t)__builtin_amdgcn_readfirstlane(r1);
            if (w1 & kOpGather) {
                n1 = QB_RL(r1, 1);
                ni1 = QB_RL(r1, 3);
                fetch_lane(prog, st1, r1, lane, wave_lane0, n1, ni1, g1);
            }
            st2 = st1 + (w1 >> 16);
            if (k + 2 < n_ops) r2 = fetch_rec(prog, st2, lane);
        }
        op_sync<ITEM_THREADS>(GLB ? w0 : 0u);
        if (w0 & kOpGather)
            gather_op<P>(n0, ni0, g0, lds, lane);
        else if (w0 & kOpReduce)
            reduce_op<P>(r0, w0, lds, lane);
        else if (!GLB || (w0 & (1u << 12)))
            step_op<P, ITEM_THREADS>(r0, w0, lds, lane);
        else
            step_mixed<P, ITEM_THREADS>(r0, w0, lds, srd, lane);
        w0 = w1;
        r0 = r1;
        g0 = g1;
        n0 = n1;
        ni0 = ni1;
        st1 = st2;
        r1 = r2;
    }
}

// complex64 items keep at most 96 VGPRs, so five waves fit on each SIMD (20
// per CU): as many items per CU as the LDS of the usual plans allows (for
// example ten 128-thread items of 16 KiB, or five 256-thread items of 26 KiB).
// A group of items holds wider values and needs more LDS per block: 128 VGPRs
// (four waves per SIMD).
template <typename R, int M>
constexpr int kExecWavesPerEU = sizeof(R) == 4 ? (M == 1 ? 5 : 4) : 1;

// LDS of an executor block, all dynamic so that the data region starts at
// address 0 (an element's byte address is then its index times the element
// size, one shift): lds_elems data elements per item of the group
// (interleaved), then each wave's copy of the first kParamRegs angles of every
// item of the group, then the queue ticket.
__host__ __device__ constexpr int exec_aux_bytes(int item_threads, int group) {
    return ((item_threads + 63) / 64) * group * kParamRegs * 8 + 16;
}

// The executor.  Block b runs group b first: items [b M, b M + M), which share
// one plan (the host checks that); later groups come from the queue.
template <typename R, int ITEM_THREADS, bool GLB, int M>
__global__ __launch_bounds__(ITEM_THREADS, (kExecWavesPerEU<R, M>)) void exec_kernel(
    const int32_t* __restrict__ program, uint32_t program_bytes, const int32_t* __restrict__ items,
    int32_t n_items, int32_t lds_elems, const double* __restrict__ params, const void* __restrict__ pool_v,
    void* __restrict__ scratch, int64_t scratch_elems, double* __restrict__ item_out,
    uint32_t* __restrict__ queue) {
    using V = typename Cx<R>::V;
    using P = typename PkOf<R, M>::T;
    static_assert(sizeof(P) == M * sizeof(V), "a group element is M item elements");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    // The kernel declares no static LDS, so its dynamic LDS starts at address
    // 0 (the launch checks that); element e of the data region is at byte
    // e * sizeof(P), which the compiler then folds into one shift.
    QB_LDS P* lds = (QB_LDS P*)(uintptr_t)0u;
    QB_LDS V* lds_v = (QB_LDS V*)(uintptr_t)0u;   // item i's element e at e * M + i
    QB_LDS double* angles_sh = (QB_LDS double*)(smem_raw + (size_t)lds_elems * sizeof(P));
    QB_LDS int32_t* item_sh = (QB_LDS int32_t*)(angles_sh + ((ITEM_THREADS + 63) / 64) * M * kParamRegs);
    const int lane = threadIdx.x;
    QB_GLB P* glb = (QB_GLB P*)(scratch) + (int64_t)blockIdx.x * scratch_elems;
    // the group's scratch slots as one bounds-checked buffer (qb_execute keeps
    // it below 4 GiB)
    const __amdgpu_buffer_rsrc_t srd = __builtin_amdgcn_make_buffer_rsrc(
        (void*)glb, (short)0, (int)(uint32_t)(scratch_elems * (int64_t)sizeof(P)), (int)kBufferDword3);
    const __amdgpu_buffer_rsrc_t prog = __builtin_amdgcn_make_buffer_rsrc(
        (void*)program, (short)0, (int)program_bytes, (int)kBufferDword3);
    const V* pool = reinterpret_cast<const V*>(pool_v);
    const int32_t n_groups = n_items / M;
    // the first group of every block is static (group = block index), so the
    // launch does not start with a burst of atomics on one address; later
    // groups come from the queue.  A ticket is taken only when a group is done
    // (groups finish at different times, so the atomics do not pile up), and
    // not at all when every group had a block of its own.  (Vector memory
    // operations complete in order, so an atomic issued early would hold up
    // every later load of the group until it returns.)
    int32_t grp = (int32_t)blockIdx.x;
    while (grp < n_groups) {
        const int32_t* __restrict__ it = items + (size_t)grp * M * QB_ITEM_WORDS;
        const int32_t* __restrict__ plan = program + it[0];
        const int32_t n_gates = plan[0];
        const int32_t* __restrict__ gates = plan + plan[2];
        // op 0's record is in flight while the gates are built
        const uint32_t st0 = (uint32_t)(it[0] + plan[3]);
        const int32_t r0 = fetch_rec(prog, st0, lane);
        if (lane == 0) {
            // materialised per group: a copy held across the group loop would
            // be spilled under the register cap
            R one = 1, zero = 0;
            asm volatile("" : "+v"(one), "+v"(zero));
            lds[kLocOne] = pk_fill(V{one, zero}, (P*)nullptr);
            lds[kLocZero] = pk_fill(V{zero, zero}, (P*)nullptr);
        }
        QB_LDS double* my_angles = angles_sh + (lane >> 6) * M * kParamRegs;
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const double* __restrict__ prm = params + it[i * QB_ITEM_WORDS + 1];
            double th[kParamRegs];
#pragma unroll
            for (int q = 0; q < kParamRegs; ++q) th[q] = prm[q];
            if ((lane & 63) == 0) {
#pragma unroll
                for (int q = 0; q < kParamRegs; ++q) my_angles[i * kParamRegs + q] = th[q];
            }
        }
        // the wave's own LDS writes are visible to its later LDS reads
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        bool same_angles = true;
#pragma unroll
        for (int i = 1; i < M; ++i) same_angles = same_angles && it[i * QB_ITEM_WORDS + 1] == it[1];
        if constexpr (M == 1) {
            const double* __restrict__ prm = params + it[1];
            const uint32_t slice_bits = (uint32_t)it[2];
            for (int32_t g = lane; g < n_gates; g += ITEM_THREADS)
                build_gate<R>(gates + (size_t)g * kGateWords, prm, my_angles, kParamRegs, slice_bits, lds, pool);
        } else if (same_angles) {
            // items of one request: one FP64 evaluation per gate for the group
            uint32_t sb[M];
#pragma unroll
            for (int i = 0; i < M; ++i) sb[i] = (uint32_t)it[i * QB_ITEM_WORDS + 2];
            for (int32_t g = lane; g < n_gates; g += ITEM_THREADS)
                build_gate_group<R, M>(gates + (size_t)g * kGateWords, params + it[1], my_angles, kParamRegs, sb,
                                       lds_v, pool);
        } else {
            // lane t builds gate t / M of item t % M
            for (int32_t t = lane; t < n_gates * M; t += ITEM_THREADS) {
                const int32_t g = t / M, i = t % M;
                int32_t aoff = it[1];
                uint32_t slice_bits = (uint32_t)it[2];
#pragma unroll
                for (int k = 1; k < M; ++k) {
                    if (i == k) {
                        aoff = it[k * QB_ITEM_WORDS + 1];
                        slice_bits = (uint32_t)it[k * QB_ITEM_WORDS + 2];
                    }
                }
                build_gate<R>(gates + (size_t)g * kGateWords, params + aoff, my_angles + i * kParamRegs, kParamRegs,
                              slice_bits, LdsLane<V, M>{lds_v + i}, pool);
            }
        }
        group_sync<ITEM_THREADS>();
        run_ops<P, ITEM_THREADS, GLB>(prog, st0, plan[1], r0, lds, srd, lane);
        group_sync<ITEM_THREADS>();
        if (lane < M) {
            // item `lane`'s result: LDS, or the group's scratch (through its buffer)
            const uint32_t rl = (uint32_t)plan[4];
            int32_t slot = it[3];
#pragma unroll
            for (int k = 1; k < M; ++k)
                if (lane == k) slot = it[k * QB_ITEM_WORDS + 3];
            const uint32_t e = (rl & ~kGlobal) * (uint32_t)M + (uint32_t)lane;
            const V r = (rl & kGlobal) ? buf_load(srd, e, (V*)nullptr) : lds_v[e];
            item_out[2 * (size_t)slot] = (double)r.x;
            item_out[2 * (size_t)slot + 1] = (double)r.y;
        }
        if ((int64_t)gridDim.x >= (int64_t)n_groups) break;
        if (lane == 0) *item_sh = (int32_t)(gridDim.x + atomicAdd(queue, 1u));
        group_sync<ITEM_THREADS>();
        grp = __builtin_amdgcn_readfirstlane(*item_sh);
        // item_sh is rewritten only after the next group's barriers
    }
}

// Queue reset before a launch whose items outnumber its blocks (only such
// launches take tickets).
__global__ void zero_u32(uint32_t* p) { *p = 0u; }

// Neumaier-compensated FP64 complex sums in a fixed order (deterministic):
// one wave per segment, lanes take strided items, then a shuffle tree combines
// (sum, compensation) pairs; the total is one block over the segment sums.
__device__ __forceinline__ void neumaier(double& s, double& c, double x) {
    const double t = s + x;
    c += (fabs(s) >= fabs(x)) ? ((s - t) + x) : ((x - t) + s);
    s = t;
}

__device__ __forceinline__ void wave_combine(double& s, double& c) {
    for (int off = 32; off > 0; off >>= 1) {
        const double os = __shfl_down(s, off, 64);
        const double oc = __shfl_down(c, off, 64);
        neumaier(s, c, os);
        c += oc;
    }
}

__global__ __launch_bounds__(64) void reduce_segments(const double* __restrict__ item_out,
                                                      const int32_t* __restrict__ seg, int32_t n_seg,
                                                      double* __restrict__ term_out) {
    const int32_t s = blockIdx.x;
    if (s >= n_seg) return;
    double sr = 0, cr = 0, si = 0, ci = 0;
    for (int32_t i = seg[s] + (int32_t)threadIdx.x; i < seg[s + 1]; i += 64) {
        neumaier(sr, cr, item_out[2 * (size_t)i]);
        neumaier(si, ci, item_out[2 * (size_t)i + 1]);
    }
    wave_combine(sr, cr);
    wave_combine(si, ci);
    if (threadIdx.x == 0) {
        term_out[2 * (size_t)s] = sr + cr;
        term_out[2 * (size_t)s + 1] = si + ci;
    }
}

// copy_to (optional): the terms are also written there (one item per term).
__global__ __launch_bounds__(256) void reduce_total(const double* __restrict__ term_out, int32_t n_seg,
                                                    double* __restrict__ total, double* __restrict__ copy_to) {
    __shared__ double sh[4][4];
    double sr = 0, cr = 0, si = 0, ci = 0;
    for (int32_t s = threadIdx.x; s < n_seg; s += blockDim.x) {
        const double re = term_out[2 * (size_t)s], im = term_out[2 * (size_t)s + 1];
        if (copy_to) {
            copy_to[2 * (size_t)s] = re;
            copy_to[2 * (size_t)s + 1] = im;
        }
        neumaier(sr, cr, re);
        neumaier(si, ci, im);
    }
    wave_combine(sr, cr);
    wave_combine(si, ci);
    const int w = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0) {
        sh[w][0] = sr;
        sh[w][1] = cr;
        sh[w][2] = si;
        sh[w][3] = ci;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double ar = 0, acr = 0, ai = 0, aci = 0;
        for (int k = 0; k < 4; ++k) {
            neumaier(ar, acr, sh[k][0]);
            acr += sh[k][1];
            neumaier(ai, aci, sh[k][2]);
            aci += sh[k][3];
        }
        total[0] = ar + acr;
        total[1] = ai + aci;
    }
}

// One executor instantiation: precision R, ITEM_THREADS T, global-scratch
// variant GLB, group size M.
template <typename R, int T, bool GLB, int M>
struct ExecCfg {
    static const void* kernel() { return reinterpret_cast<const void*>(&exec_kernel<R, T, GLB, M>); }
};

// Calls f(ExecCfg<...>{}) for a launch shape.  Groups of 2 items run complex64
// items of 64 to 512 threads; every other shape runs single items.
template <typename F>
int exec_dispatch(int32_t dtype, int item_threads, int group, bool glb, F&& f) {
#define QB_EXEC_T(R, G, M)                                             \
    switch (item_threads) {                                            \
    case 64: return f(ExecCfg<R, 64, G, M>{});                         \
    case 128: return f(ExecCfg<R, 128, G, M>{});                       \
    case 256: return f(ExecCfg<R, 256, G, M>{});                       \
    case 512: return f(ExecCfg<R, 512, G, M>{});                       \
    default: break;                                                    \
    }
#define QB_EXEC_T1(R, G)                                               \
    QB_EXEC_T(R, G, 1)                                                 \
    if (item_threads == 1024) return f(ExecCfg<R, 1024, G, 1>{});
    if (group == 1) {
        if (dtype == QB_C128) {
            if (glb) { QB_EXEC_T1(double, true) } else { QB_EXEC_T1(double, false) }
        } else {
            if (glb) { QB_EXEC_T1(float, true) } else { QB_EXEC_T1(float, false) }
        }
        return set_error(QB_ERR_INVALID, "item_threads %d not in {64,128,256,512,1024}", item_threads);
    }
    if (group == 2 && dtype == QB_C64) {
        if (glb) { QB_EXEC_T(float, true, 2) } else { QB_EXEC_T(float, false, 2) }
        return set_error(QB_ERR_INVALID, "groups of 2 items run item_threads 64, 128, 256 or 512, not %d", item_threads);
    }
#undef QB_EXEC_T1
#undef QB_EXEC_T
    return set_error(QB_ERR_INVALID, "group %d: 1, or 2 for complex64", group);
}

template <typename R, int T, bool GLB, int M>
int exec_launch_cfg(ExecCfg<R, T, GLB, M> cfg, const ExecArgs& a, int grid, int lds_bytes, hipStream_t st) {
    // the kernel addresses its LDS data from address 0: it must declare no
    // static LDS (checked once per instantiation)
    static int static_lds = -1;
    if (static_lds < 0) {
        hipFuncAttributes attr;
        QB_HIP(hipFuncGetAttributes(&attr, cfg.kernel()));
        static_lds = (int)attr.sharedSizeBytes;
    }
    if (static_lds != 0)
        return set_error(QB_ERR_UNSUPPORTED, "executor kernel declares %d bytes of static LDS", static_lds);
    hipLaunchKernelGGL((exec_kernel<R, T, GLB, M>), dim3(grid), dim3(T), lds_bytes, st, a.program, a.program_bytes,
                       a.items, a.n_items, a.lds_elems, a.params, a.pool, a.scratch, a.scratch_elems, a.item_out,
                       a.queue);
    QB_HIP(hipGetLastError());
    return QB_OK;
}

// Blocks resident per CU (*grid_per_cu) and items resident on the device
// (*n_slots: blocks x CUs x group) for a launch shape.  Batches whose plans need
// no global scratch run the variant without the global-scratch path (glb).
int slots_for(int32_t item_threads, int32_t lds_elems, int32_t dtype, int32_t group, bool glb, int32_t* n_slots,
              int* grid_per_cu) {
    if (g_device < 0) return set_error(QB_ERR_NOT_INIT, "qb_init was not called");
    if (lds_elems < 0) return set_error(QB_ERR_INVALID, "lds_elems < 0");
    if (group < 1) return set_error(QB_ERR_INVALID, "group %d < 1", group);
    const int elem = dtype == QB_C128 ? 16 : 8;
    const int64_t lds_bytes = (int64_t)lds_elems * elem * group + exec_aux_bytes(item_threads, group);
    if (lds_bytes > 160 * 1024)
        return set_error(QB_ERR_UNSUPPORTED, "LDS request %lld bytes exceeds 160 KiB", (long long)lds_bytes);
    int blocks = 0;
    const int rc = exec_dispatch(dtype, item_threads, group, glb, [&](auto cfg) {
        QB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, cfg.kernel(), item_threads, (size_t)lds_bytes));
        return QB_OK;
    });
    if (rc) return rc;
    if (blocks < 1) return set_error(QB_ERR_UNSUPPORTED, "executor shape does not fit on a CU");
    *grid_per_cu = blocks;
    *n_slots = blocks * g_num_cus * group;
    return QB_OK;
}

// Wide plans (every tensor in global memory): gate build and result copy.
// Slice z (grid z) uses slice assignment slice_bits + z and the arena at
// arena + z zbytes.
template <typename R>
__global__ void build_gates_global(const int32_t* __restrict__ gates, int32_t n_gates,
                                   const double* __restrict__ params, uint32_t slice_bits,
                                   void* __restrict__ arena, int64_t zbytes) {
    using V = typename Cx<R>::V;
    const int32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n_gates)
        build_gate<R>(gates + (size_t)g * kGateWords, params, (const QB_LDS double*)nullptr, 0,
                      slice_bits + blockIdx.z,
                      reinterpret_cast<V*>(reinterpret_cast<char*>(arena) + (int64_t)blockIdx.z * zbytes),
                      (const V*)nullptr);
}

// Result of a wide plan: n elements from arena[loc], element e of slice z
// (grid z; its arena at arena + z zbytes) as complex128 to out[2 (e stride + z)].
template <typename R>
__global__ void store_result(const void* __restrict__ arena, uint32_t loc, uint32_t n, int64_t stride,
                             double* __restrict__ out, int64_t zbytes) {
    using V = typename Cx<R>::V;
    const V* src = reinterpret_cast<const V*>(reinterpret_cast<const char*>(arena) + (int64_t)blockIdx.z * zbytes);
    double* dst = out + 2 * (int64_t)blockIdx.z;
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
        const V r = src[loc + e];
        dst[2 * (int64_t)e * stride] = (double)r.x;
        dst[2 * (int64_t)e * stride + 1] = (double)r.y;
    }
}

// Host: pext runs word of a mask (n_runs << 30 | run0 | run1 << 15); 3 << 30 when
// the mask has more than two runs.
uint32_t runs_word(uint32_t mask) {
    uint8_t src[2], w[2], dst[2];
    const int n = mask_runs(mask, src, w, dst, 2);
    if (n > 2) return 3u << 30;
    uint32_t r = (uint32_t)n << 30;
    for (int k = 0; k < n; ++k) r |= ((uint32_t)src[k] | ((uint32_t)w[k] << 5) | ((uint32_t)dst[k] << 10)) << (15 * k);
    return r;
}

uint32_t pext_host(uint32_t x, uint32_t m) {
    uint32_t r = 0, bit = 1;
    for (; m; m &= m - 1, bit <<= 1)
        if (x & m & (0u - m)) r |= bit;
    return r;
}

uint32_t pdep_host(uint32_t x, uint32_t m) {
    uint32_t r = 0;
    for (; m; m &= m - 1, x >>= 1)
        if (x & 1u) r |= m & (0u - m);
    return r;
}

// Launch the plain step for K <= 4 inputs.  Inputs are reordered so the largest
// comes first; up to three output bits >= 8 that the largest input does not
// carry become register bits (each thread reads the large input once and
// writes the 2^n outputs that share it).  The loop runs over the remaining
// ("kept") output bits; an element's offset in input j is
// pext(tid, m_j & 255) + pext(pdep(base >> 8, kept >> 8), m_j >> 8) << popc(m_j & 255)
// plus the register pattern's pext(pdep(r, reg), m_j).
template <int K>
int launch_step_k(const StepArgs& a, const uint32_t* runs_unused, int32_t dtype, int grid_unused, hipStream_t st,
                  int nz = 1, int64_t zbytes = 0) {
    (void)runs_unused;
    (void)grid_unused;
    int order[K];
    for (int j = 0; j < K; ++j) order[j] = j;
    int big = 0;
    for (int j = 1; j < K; ++j) {
        const int sj = __builtin_popcount(a.in[j].mask) + (a.in[j].top ? 1 : 0);
        const int sb = __builtin_popcount(a.in[big].mask) + (a.in[big].top ? 1 : 0);
        if (sj > sb) big = j;
    }
    order[0] = big;
    order[big] = 0;
    const uint32_t full = a.rank_out >= 32 ? 0xffffffffu : ((1u << a.rank_out) - 1u);
    uint32_t missing = full & ~a.in[big].mask & ~255u;
    uint32_t reg = 0;
    int n_reg = 0;
    if (a.rank_out >= 11) {
        for (int b = 31; b >= 8 && n_reg < 3; --b)
            if ((missing >> b) & 1u) {
                reg |= 1u << b;
                ++n_reg;
            }
    }
    const uint32_t keep = full & ~reg;
    StepArgsK<K> k;
    memset(&k, 0, sizeof(k));
    for (int j = 0; j < K; ++j) {
        const StepInput& in = a.in[order[j]];
        k.ptr[j] = in.ptr;
        k.top[j] = in.top;
        k.mask[j] = in.mask;
        k.runs[j] = runs_word(in.mask & 255u);
        for (uint32_t r = 0; r < (1u << n_reg); ++r) k.roff[j][r] = pext_host(pdep_host(r, reg), in.mask);
    }
    for (uint32_t r = 0; r < (1u << n_reg); ++r) k.rdep[r] = pdep_host(r, reg);
    for (int j = 0; j < K; ++j) k.hi_runs[j] = runs_word(k.mask[j] >> 8);
    k.out_hi = keep >> 8;
    {
        // keep >> 8 has at most n_reg + 1 runs
        uint32_t m = k.out_hi, n = 0, w0 = 0, w1 = 0;
        for (uint32_t b = 0; b < 32 && m;) {
            if (!((m >> b) & 1u)) { ++b; continue; }
            uint32_t s0 = b;
            while (b < 32 && ((m >> b) & 1u)) ++b;
            const uint32_t f = s0 | ((b - s0) << 5);
            if (n < 2) w0 |= f << (10 * n); else w1 |= f << (10 * (n - 2));
            ++n;
            m &= ~(((b < 32) ? (1u << b) : 0u) - (1u << s0));
        }
        k.dep_runs[0] = w0 | (n << 28);
        k.dep_runs[1] = w1;
    }
    k.n_reg = (uint32_t)n_reg;
    k.out = a.out;
    k.rank_out = a.rank_out - (uint32_t)n_reg;
    k.sum = a.n_sum == 2 ? 1u : 0u;
    k.zbytes = zbytes;
    const uint64_t n_out = 1ull << k.rank_out;
    const uint64_t want = (n_out + 255) / 256;
    const uint64_t cap = (uint64_t)g_num_cus * 8;
    const int grid = (int)(want < cap ? (want > 0 ? want : 1) : cap);
    if (dtype == QB_C128)
        hipLaunchKernelGGL((step_kernel_k<double, K>), dim3(grid, 1, nz), dim3(256), 0, st, k);
    else
        hipLaunchKernelGGL((step_kernel_k<float, K>), dim3(grid, 1, nz), dim3(256), 0, st, k);
    QB_HIP(hipGetLastError());
    return QB_OK;
}

}  // namespace

// Deposit runs of the set bits of `keep` (below hi_bits) for the kernels'
// block loop over hk: run r copies bits [src, src + w) of hk to [dst, dst + w)
// of hi (src | w << 5 | dst << 10).  Checked against pdep for every bit and for
// all of them.  Returns false when more than kFusedDepRuns runs are needed.
static bool deposit_runs(uint32_t keep, uint32_t hi_bits, uint32_t* dep_run, uint32_t* n_dep, uint32_t* hk_bits) {
    uint32_t n = 0, src = 0;
    for (uint32_t b = 0; b < hi_bits;) {
        if (!((keep >> b) & 1u)) { ++b; continue; }
        const uint32_t b0 = b;
        while (b < hi_bits && ((keep >> b) & 1u)) ++b;
        if (n >= (uint32_t)kFusedDepRuns) return false;
        dep_run[n++] = src | ((b - b0) << 5) | (b0 << 10);
        src += b - b0;
    }
    for (uint32_t b = 0; b <= src; ++b) {
        const uint32_t hk = b < src ? (1u << b) : ((1u << src) - 1u);
        uint32_t hi = 0;
        for (uint32_t r = 0; r < n; ++r) {
            const uint32_t f = dep_run[r], wd = (f >> 5) & 31u;
            hi |= ((hk >> (f & 31u)) & ((1u << wd) - 1u)) << (f >> 10);
        }
        if (hi != pdep_host(hk, keep)) return false;
    }
    *n_dep = n;
    *hk_bits = src;
    return true;
}

// Bit runs mapping the block bits (the set bits of hmask, in order, as the
// bits of hk) to offset bits of a tensor over `mask`: the kernel's chain_runs
// must equal pext(pdep(hk, hmask), mask); checked for every bit and all of them.
static bool chain_hk_runs(uint32_t hmask, uint32_t mask, uint32_t* run, uint32_t* n_run) {
    uint32_t n = 0, k = 0, prev_k = 0, prev_t = 0;
    for (uint32_t m = hmask; m; m &= m - 1, ++k) {
        const uint32_t u = (uint32_t)__builtin_ctz(m);
        if (!((mask >> u) & 1u)) continue;
        const uint32_t t = (uint32_t)__builtin_popcount(mask & ((1u << u) - 1u));
        if (n > 0 && k == prev_k + 1 && t == prev_t + 1) {
            run[n - 1] += 1u << 5;
        } else {
            if (n >= (uint32_t)kChainRuns) return false;
            run[n++] = k | (1u << 5) | (t << 10);
        }
        prev_k = k;
        prev_t = t;
    }
    const uint32_t bits = (uint32_t)__builtin_popcount(hmask);
    for (uint32_t b = 0; b <= bits; ++b) {
        const uint32_t hk = b < bits ? (1u << b) : (bits >= 32 ? 0xffffffffu : (1u << bits) - 1u);
        uint32_t got = 0;
        for (uint32_t r = 0; r < n; ++r) {
            const uint32_t f = run[r], w = (f >> 5) & 31u;
            got |= ((hk >> (f & 31u)) & ((1u << w) - 1u)) << (f >> 10);
        }
        if (got != pext_host(pdep_host(hk, hmask), mask)) return false;
    }
    *n_run = n;
    return true;
}

// Register, lane and block bits of a chain kernel launch (see chain_kernel)
// for inputs A, B (and D when has_d) of `a`, inner summed variables xbits and
// outer ybits.  Returns false when no split fits.
struct ChainPlan {
    ChainArgs e;
    uint32_t rr, rdb;
    bool pair;
};

static bool chain_config(const FusedArgs& a, uint32_t xbits, uint32_t ybits, bool has_d, int32_t dtype,
                         ChainPlan* cp) {
    const uint32_t omask = a.omask;
    const uint32_t ma = a.mask[0], mb = a.mask[1], md = has_d ? a.mask[2] : 0u;
    const uint32_t kx = (uint32_t)__builtin_popcount(xbits), ky = (uint32_t)__builtin_popcount(ybits);
    const uint32_t eb = dtype == QB_C128 ? 16u : 8u;
    const int lg = eb == 16u ? 4 : 3;
    if (a.rank_out < 8 || ky > (uint32_t)kChainMaxY || kx + ky > (uint32_t)kChainMaxXY ||
        __builtin_popcount(md) > kChainMaxD || __builtin_popcount(ma) + lg > 31 || __builtin_popcount(mb) + lg > 31)
        return false;
    ChainArgs& e = cp->e;
    memset(&e, 0, sizeof(e));
    e.kx = kx;
    e.ky = ky;
    e.xbits = xbits;
    e.ybits = ybits;
    e.has_d = has_d ? 1u : 0u;
    e.d_elems = 1u << __builtin_popcount(md);
    // output bits that only A (B, D) carries, highest first
    uint32_t ca[32], cb[32], cd[32], na = 0, nb = 0, nd = 0;
    for (int u = 31; u >= 0; --u) {
        if (!((omask >> u) & 1u)) continue;
        const bool in_a = (ma >> u) & 1u, in_b = (mb >> u) & 1u, in_d = (md >> u) & 1u;
        if (in_a && !in_b) ca[na++] = (uint32_t)u;
        else if (in_b && !in_a) cb[nb++] = (uint32_t)u;
        else if (in_d && !in_a && !in_b) cd[nd++] = (uint32_t)u;
    }
    // pairs: A's and B's lowest variables are output bits only they carry
    // (16-byte loads: A, B and every slice's region 16-byte aligned)
    const uint32_t pa = ma ? (uint32_t)__builtin_ctz(ma) : 32u, pb = mb ? (uint32_t)__builtin_ctz(mb) : 32u;
    const bool aligned = ((reinterpret_cast<uintptr_t>(a.in[0]) | reinterpret_cast<uintptr_t>(a.in[1]) |
                           (uintptr_t)a.zbytes) & 15u) == 0;
    const bool can_pair = dtype == QB_C64 && aligned && pa < 32u && pb < 32u && ((omask >> pa) & 1u) &&
                          !((mb >> pa) & 1u) && ((omask >> pb) & 1u) && !((ma >> pb) & 1u);
    uint32_t rr = 0, rdb = 0, ramask = 0, rbmask = 0, rdmask = 0;
    bool pair = false, ok = false;
    for (int cfg = 0; cfg < 10 && !ok; ++cfg) {
        // by preference: pairs with 2 or 1 register bits per input, then
        // without pairs 2, 1, 0; each with and without a D bit
        const bool want_pair = cfg < 4;
        const uint32_t want_r = cfg < 2 ? 2u : cfg < 4 ? 1u : cfg < 6 ? 2u : cfg < 8 ? 1u : 0u;
        const uint32_t want_d = (cfg & 1) ? 0u : 1u;
        if (want_pair && !can_pair) continue;
        pair = want_pair;
        ramask = rbmask = 0;
        uint32_t ga = 0, gb = 0;
        if (pair) {
            ramask = 1u << pa;
            rbmask = 1u << pb;
            ga = gb = 1;
        }
        for (uint32_t i = 0; i < na && ga < want_r; ++i)
            if (!((ramask >> ca[i]) & 1u)) { ramask |= 1u << ca[i]; ++ga; }
        for (uint32_t i = 0; i < nb && gb < want_r; ++i)
            if (!((rbmask >> cb[i]) & 1u)) { rbmask |= 1u << cb[i]; ++gb; }
        rr = ga < gb ? ga : gb;
        if (pair && rr < 1) continue;
        // equal counts: drop the highest extra bits (never a pair bit, the lowest)
        while ((uint32_t)__builtin_popcount(ramask) > rr) ramask &= ~(1u << (31u - (uint32_t)__builtin_clz(ramask)));
        while ((uint32_t)__builtin_popcount(rbmask) > rr) rbmask &= ~(1u << (31u - (uint32_t)__builtin_clz(rbmask)));
        rdb = (want_d && nd > 0) ? 1u : 0u;
        rdmask = rdb ? (1u << cd[0]) : 0u;
        const uint32_t regs = ramask | rbmask | rdmask;
        if ((uint32_t)__builtin_popcount(omask & ~regs) < 8u) continue;
        // lane bits: the lowest 8 other output bits; block bits: the rest
        uint32_t lane = 0;
        for (uint32_t m = omask & ~regs, n = 0; n < 8u; m &= m - 1, ++n) lane |= m & (0u - m);
        const uint32_t hmask = omask & ~regs & ~lane;
        const uint32_t tm[4] = {ma, mb, md, omask};
        ok = true;
        for (int j = 0; j < 4 && ok; ++j) ok = chain_hk_runs(hmask, tm[j], e.hk_run[j], &e.hk_nrun[j]);
        if (!ok) continue;
        e.lane_mask = lane;
        e.hk_bits = (uint32_t)__builtin_popcount(hmask);
    }
    if (!ok) return false;
    if (pair && ((ramask & (0u - ramask)) != (1u << pa) || (rbmask & (0u - rbmask)) != (1u << pb))) return false;
    for (uint32_t p = 0; p < (1u << rr); ++p) {
        const uint32_t ua = pdep_host(p, ramask), ub = pdep_host(p, rbmask);
        e.roff_a[p] = pext_host(ua, ma) * eb;
        e.roff_b[p] = pext_host(ub, mb) * eb;
        e.droff_a[p] = pext_host(ua, md);
        e.droff_b[p] = pext_host(ub, md);
        e.oroff_a[p] = pext_host(ua, omask);
        e.oroff_b[p] = pext_host(ub, omask);
    }
    if (pair)
        for (uint32_t p = 0; p < (1u << rr); p += 2)
            if (e.roff_a[p + 1] != e.roff_a[p] + eb || e.roff_b[p + 1] != e.roff_b[p] + eb) return false;
    for (uint32_t q = 0; q < (1u << rdb); ++q) {
        const uint32_t ud = pdep_host(q, rdmask);
        e.droff_d[q] = pext_host(ud, md);
        e.oroff_d[q] = pext_host(ud, omask);
    }
    cp->rr = rr;
    cp->rdb = rdb;
    cp->pair = pair;
    return true;
}

static int launch_chain_plan(const FusedArgs& a, const ChainPlan& cp, int32_t dtype, hipStream_t st, int nz) {
    const uint64_t gcap = (uint64_t)g_num_cus * 16;
    const uint64_t want = 1ull << cp.e.hk_bits;
    const int grid = (int)(want < gcap ? want : gcap);
    const uint32_t rr = cp.rr, rdb = cp.rdb;
    const bool pair = cp.pair;
    const ChainArgs& e = cp.e;
#define QB_CHAIN_CASE(RR_, RD_)                                                                           \
    if (!pair && rr == RR_ && rdb == RD_) {                                                               \
        if (dtype == QB_C128)                                                                             \
            hipLaunchKernelGGL((chain_kernel<double, RR_, RD_, false>), dim3(grid, 1, nz), dim3(256), 0, st, a, e); \
        else                                                                                              \
            hipLaunchKernelGGL((chain_kernel<float, RR_, RD_, false>), dim3(grid, 1, nz), dim3(256), 0, st, a, e);  \
    }
#define QB_CHAIN_PAIR_CASE(RR_, RD_)                                                                      \
    if (pair && rr == RR_ && rdb == RD_)                                                                  \
        hipLaunchKernelGGL((chain_kernel<float, RR_, RD_, true>), dim3(grid, 1, nz), dim3(256), 0, st, a, e);
    QB_CHAIN_CASE(0, 0) QB_CHAIN_CASE(0, 1) QB_CHAIN_CASE(1, 0) QB_CHAIN_CASE(1, 1) QB_CHAIN_CASE(2, 0)
    QB_CHAIN_CASE(2, 1)
    QB_CHAIN_PAIR_CASE(1, 0) QB_CHAIN_PAIR_CASE(1, 1) QB_CHAIN_PAIR_CASE(2, 0) QB_CHAIN_PAIR_CASE(2, 1)
#undef QB_CHAIN_PAIR_CASE
#undef QB_CHAIN_CASE
    QB_HIP(hipGetLastError());
    return QB_OK;
}

// The matrix-core form of a chain launch (complex64): 4 output bits only A
// carries (the lowest: consecutive lanes read consecutive elements), 4 only B
// carries, one only D carries when there is one, 4 to 16 inner assignments
// (kx 2..4); false when the op does not have them.
static bool chain_mfma_config(const FusedArgs& a, uint32_t xbits, uint32_t ybits, bool has_d, int32_t dtype,
                              ChainMfmaArgs* e, uint32_t* rd_out) {
    const uint32_t omask = a.omask;
    const uint32_t ma = a.mask[0], mb = a.mask[1], md = has_d ? a.mask[2] : 0u;
    const uint32_t kx = (uint32_t)__builtin_popcount(xbits), ky = (uint32_t)__builtin_popcount(ybits);
    if (dtype != QB_C64 || kx < 2 || kx > (uint32_t)kChainMaxX || ky > (uint32_t)kChainMaxY ||
        kx + ky > (uint32_t)kChainMaxXY || __builtin_popcount(md) > kChainMaxD ||
        __builtin_popcount(ma) + 3 > 31 || __builtin_popcount(mb) + 3 > 31)
        return false;
    uint32_t mm = 0, nm = 0, dmask = 0, na = 0, nb = 0;
    for (uint32_t u = 0; u < 32; ++u) {
        if (!((omask >> u) & 1u)) continue;
        const bool in_a = (ma >> u) & 1u, in_b = (mb >> u) & 1u, in_d = (md >> u) & 1u;
        if (in_a && !in_b && na < 4) { mm |= 1u << u; ++na; }
        else if (in_b && !in_a && nb < 4) { nm |= 1u << u; ++nb; }
        else if (in_d && !in_a && !in_b && !dmask) dmask = 1u << u;
    }
    if (na < 4 || nb < 4) return false;
    memset(e, 0, sizeof(*e));
    const uint32_t hmask = omask & ~(mm | nm | dmask);
    const uint32_t tm[4] = {ma, mb, md, omask};
    for (int j = 0; j < 4; ++j)
        if (!chain_hk_runs(hmask, tm[j], e->hk_run[j], &e->hk_nrun[j])) return false;
    e->m_mask = mm;
    e->n_mask = nm;
    e->d_mask = dmask;
    e->xbits = xbits;
    e->ybits = ybits;
    e->hk_bits = (uint32_t)__builtin_popcount(hmask);
    e->kx = kx;
    e->ky = ky;
    e->d_elems = 1u << __builtin_popcount(md);
    e->has_d = has_d ? 1u : 0u;
    *rd_out = dmask ? 1u : 0u;
    return true;
}

static int launch_chain_mfma(const FusedArgs& a, const ChainMfmaArgs& e, uint32_t rd, hipStream_t st, int nz) {
    const uint64_t waves = 1ull << e.hk_bits;
    const uint64_t want = (waves + 3) / 4;
    const uint64_t gcap = (uint64_t)g_num_cus * 16;
    const int grid = (int)(want < gcap ? want : gcap);
    if (rd)
        hipLaunchKernelGGL(chain_mfma_kernel<1>, dim3(grid, 1, nz), dim3(256), 0, st, a, e);
    else
        hipLaunchKernelGGL(chain_mfma_kernel<0>, dim3(grid, 1, nz), dim3(256), 0, st, a, e);
    QB_HIP(hipGetLastError());
    return QB_OK;
}

// Launch a chain op (kFusedChain): inputs A, B, D; the inner summed variables
// are the (at most kChainMaxX lowest) summed variables D lacks.  `a` holds the
// checked inputs and masks.
static int launch_chain(const FusedArgs& a, uint32_t n_in, uint32_t smask, int32_t k_op, int32_t dtype,
                        hipStream_t st, int nz) {
    const uint32_t md = a.mask[2];
    uint32_t xbits = 0;
    for (uint32_t m = smask & ~md, n = 0; m && n < (uint32_t)kChainMaxX; m &= m - 1, ++n) xbits |= m & (0u - m);
    const uint32_t ybits = smask & ~xbits;
    const uint32_t kx = (uint32_t)__builtin_popcount(xbits), ky = (uint32_t)__builtin_popcount(ybits);
    const uint32_t rd = (uint32_t)__builtin_popcount(md);
    if (n_in != 3 || a.rank_out < 8 || kx < 1 || ky > (uint32_t)kChainMaxY || rd > (uint32_t)kChainMaxD)
        return set_error(QB_ERR_INVALID, "chain op %d: %u inputs, rank %u, %u inner and %u outer summed variables, "
                         "third input of rank %u", k_op, n_in, a.rank_out, kx, ky, rd);
    ChainMfmaArgs em;
    uint32_t rdm = 0;
    if (chain_mfma_config(a, xbits, ybits, true, dtype, &em, &rdm)) return launch_chain_mfma(a, em, rdm, st, nz);
    ChainPlan cp;
    if (!chain_config(a, xbits, ybits, true, dtype, &cp))
        return set_error(QB_ERR_UNSUPPORTED, "chain op %d: no register and lane split fits (or an input of 2 GiB "
                         "or more)", k_op);
    return launch_chain_plan(a, cp, dtype, st, nz);
}

// Parse, check and launch one fused op of a wide plan (host copy `w`, device
// copy `plan_dev`).  Every offset the kernel forms is a pext of a union offset
// by an input's mask: the tables are recomputed here from the masks and must
// match, and each input and the output must lie inside the arena.
int launch_fused(const int32_t* w, int64_t plan_words, const int32_t* plan_dev, int64_t pos, int32_t k_op,
                 char* arena, int64_t elem, int64_t cap, int32_t dtype, hipStream_t st, int nz, int64_t zbytes,
                 bool check_only) {
    if (pos + kFusedHdr > plan_words) return set_error(QB_ERR_INVALID, "fused op %d outside the plan", k_op);
    const uint32_t w0 = (uint32_t)w[pos];
    const uint32_t n_in = w0 & 15u;
    const uint32_t meta = (uint32_t)w[pos + 1];
    const uint32_t rank_u = meta & 255u, rank_out = (meta >> 8) & 255u, k = (meta >> 16) & 255u,
                   lo_bits = meta >> 24;
    const uint32_t out_loc = (uint32_t)w[pos + 2];
    const uint32_t omask = (uint32_t)w[pos + 4], smask = (uint32_t)w[pos + 5];
    if (n_in < 1 || n_in > (uint32_t)kFusedMaxIn) return set_error(QB_ERR_INVALID, "fused op %d: %u inputs", k_op, n_in);
    const bool chain = (w0 & kFusedChain) != 0;
    if (rank_u > (chain ? 31u : 30u) || rank_out + k != rank_u || k > 8 || lo_bits != (rank_out < 8 ? rank_out : 8u) ||
        (omask & smask) || (uint32_t)__builtin_popcount(omask) != rank_out ||
        (uint32_t)__builtin_popcount(smask) != k || ((omask | smask) >> rank_u))
        return set_error(QB_ERR_INVALID, "fused op %d: inconsistent shape", k_op);
    const int64_t tab = pos + kFusedHdr + 4 * (int64_t)n_in;
    const int64_t len = tab + ((int64_t)n_in << lo_bits) + ((int64_t)n_in << k) - pos;
    if ((int64_t)(w0 >> 16) != len || pos + len > plan_words)
        return set_error(QB_ERR_INVALID, "fused op %d: record length", k_op);
    if (!(out_loc & kGlobal) || (int64_t)(out_loc & ~kGlobal) + (1ll << rank_out) > cap)
        return set_error(QB_ERR_INVALID, "fused op %d output outside the arena", k_op);
    FusedArgs a;
    memset(&a, 0, sizeof(a));
    for (uint32_t j = 0; j < n_in; ++j) {
        const uint32_t loc = (uint32_t)w[pos + kFusedHdr + 4 * j], mask = (uint32_t)w[pos + kFusedHdr + 4 * j + 1];
        if (mask >> rank_u) return set_error(QB_ERR_INVALID, "fused op %d input %u mask", k_op, j);
        if (!(loc & kGlobal) || (int64_t)(loc & ~kGlobal) + (1ll << __builtin_popcount(mask)) > cap)
            return set_error(QB_ERR_INVALID, "fused op %d input %u outside the arena", k_op, j);
        a.in[j] = arena + (int64_t)(loc & ~kGlobal) * elem;
        a.mask[j] = mask;
        for (uint32_t o = 0; o < (1u << lo_bits); ++o)
            if ((uint32_t)w[tab + o * n_in + j] != pext_host(pdep_host(o, omask), mask))
                return set_error(QB_ERR_INVALID, "fused op %d: output table", k_op);
        for (uint32_t x = 0; x < (1u << k); ++x)
            if ((uint32_t)w[tab + ((int64_t)n_in << lo_bits) + x * n_in + j] != pext_host(pdep_host(x, smask), mask))
                return set_error(QB_ERR_INVALID, "fused op %d: summed-variable table", k_op);
    }
    // high-bit runs: hi bit b is output bit lo_bits + b, i.e. union bit u_b (the
    // (lo_bits + b)-th set bit of omask); it lands on bit popcount(mask & (2^u_b - 1))
    // of input j's offset when the input carries u_b
    {
        uint32_t ubit[32];
        uint32_t n_ob = 0;
        for (uint32_t u = 0; u < 32; ++u)
            if ((omask >> u) & 1u) ubit[n_ob++] = u;
        for (uint32_t j = 0; j < n_in; ++j) {
            const uint32_t mask = a.mask[j];
            uint32_t n = 0, prev_b = 0, prev_d = 0;
            for (uint32_t b = 0; b + lo_bits < n_ob; ++b) {
                const uint32_t u = ubit[b + lo_bits];
                if (!((mask >> u) & 1u)) continue;
                const uint32_t d = (uint32_t)__builtin_popcount(mask & ((1u << u) - 1u));
                if (n > 0 && n <= (uint32_t)kFusedHiRuns && b == prev_b + 1 && d == prev_d + 1) {
                    a.hi_run[j][n - 1] += 1u << 5;        // extend the run by one bit
                } else {
                    if (n < (uint32_t)kFusedHiRuns) a.hi_run[j][n] = b | (1u << 5) | (d << 10);
                    ++n;
                }
                prev_b = b;
                prev_d = d;
            }
            a.hi_n[j] = n;
            // the kernel's evaluation must equal the pext of the deposit for every
            // high bit alone and for all of them (each bit lies in one run)
            const uint32_t hi_bits = rank_out - lo_bits;
            for (uint32_t b = 0; b <= hi_bits; ++b) {
                const uint32_t hi = b < hi_bits ? (1u << b) : ((1u << hi_bits) - 1u);
                uint32_t got = pext_host(pdep_host(hi << lo_bits, omask), mask);
                if (n <= (uint32_t)kFusedHiRuns) {
                    got = 0;
                    for (uint32_t r = 0; r < n; ++r) {
                        const uint32_t f = a.hi_run[j][r], src = f & 31u, wd = (f >> 5) & 31u;
                        got |= ((hi >> src) & ((1u << wd) - 1u)) << (f >> 10);
                    }
                }
                if (got != pext_host(pdep_host(hi << lo_bits, omask), mask))
                    return set_error(QB_ERR_UNSUPPORTED, "fused op %d input %u: high-bit runs", k_op, j);
            }
        }
    }
    if (check_only) return QB_OK;
    const int32_t* lo_tab = plan_dev + tab;                        // (1 << lo_bits) x n_in
    const int32_t* x_tab = plan_dev + tab + ((int64_t)n_in << lo_bits);  // (1 << k) x n_in
    a.out = arena + (int64_t)(out_loc & ~kGlobal) * elem;
    a.omask = omask;
    a.rank_out = rank_out;
    a.k = k;
    a.lo_bits = lo_bits;
    a.zbytes = zbytes;
    if (chain) return launch_chain(a, n_in, smask, k_op, dtype, st, nz);
    const uint64_t gcap = (uint64_t)g_num_cus * 16;
    if (n_in == 2 && rank_out >= lo_bits + 1) {
        // register bits: high output bits carried by one input only
        const uint32_t hi_bits = rank_out - lo_bits;
        uint32_t ubit[32], n_ob = 0;
        for (uint32_t u = 0; u < 32; ++u)
            if ((omask >> u) & 1u) ubit[n_ob++] = u;
        // complex64: up to three register bits per input (an 8 x 8 tile of
        // outputs per thread: a quarter of a load per output and summed
        // assignment); complex128 holds at most 4 x 4 in registers.  Fewer
        // register bits when the remaining high bits would split into more
        // deposit runs than the kernel takes.
        uint32_t reg_a[3], reg_b[3], ra = 0, rb = 0;
        for (uint32_t max_reg = dtype == QB_C128 ? 2u : 3u; max_reg > 0; --max_reg) {
            ra = rb = 0;
            uint32_t regmask = 0;
            for (uint32_t b = hi_bits; b-- > 0;) {          // the highest bits first
                const uint32_t u = ubit[b + lo_bits];
                const bool in_a = (a.mask[0] >> u) & 1u, in_b = (a.mask[1] >> u) & 1u;
                if (in_a && !in_b && ra < max_reg) {
                    reg_a[ra++] = b;
                    regmask |= 1u << b;
                } else if (in_b && !in_a && rb < max_reg) {
                    reg_b[rb++] = b;
                    regmask |= 1u << b;
                }
            }
            const uint32_t keep = ((1u << hi_bits) - 1u) & ~regmask;
            uint32_t runs = 0;
            for (uint32_t b = 0; b < hi_bits; ++b)
                if (((keep >> b) & 1u) && (b == 0 || !((keep >> (b - 1)) & 1u))) ++runs;
            if (runs <= (uint32_t)kFusedDepRuns) break;
        }
        {
            // the chain kernel without D: lane bits of its own choosing, so
            // register bits may come from the low output bits too, and 16-byte
            // pair loads; taken when it loads less per product
            ChainPlan cp;
            if (chain_config(a, smask, 0u, false, dtype, &cp)) {
                const double nr = (double)(1u << cp.rr);
                const double cost_c = 2.0 * nr / (cp.pair ? 2.0 : 1.0) / (nr * nr);
                const double cost_f = (double)((1u << ra) + (1u << rb)) / (double)(1u << (ra + rb));
                if (cost_c < cost_f) return launch_chain_plan(a, cp, dtype, st, nz);
            }
        }
        if (ra + rb > 0) {
            Fused2Args e;
            memset(&e, 0, sizeof(e));
            uint32_t regmask = 0;
            for (uint32_t i = 0; i < ra; ++i) regmask |= 1u << reg_a[i];
            for (uint32_t i = 0; i < rb; ++i) regmask |= 1u << reg_b[i];
            const uint32_t keep = ((1u << hi_bits) - 1u) & ~regmask;
            for (uint32_t p = 0; p < (1u << ra); ++p) {
                uint32_t hp = 0;
                for (uint32_t i = 0; i < ra; ++i) if ((p >> i) & 1u) hp |= 1u << reg_a[i];
                e.rdep_a[p] = hp;
                e.roff_a[p] = pext_host(pdep_host(hp << lo_bits, omask), a.mask[0]);
            }
            for (uint32_t p = 0; p < (1u << rb); ++p) {
                uint32_t hp = 0;
                for (uint32_t i = 0; i < rb; ++i) if ((p >> i) & 1u) hp |= 1u << reg_b[i];
                e.rdep_b[p] = hp;
                e.roff_b[p] = pext_host(pdep_host(hp << lo_bits, omask), a.mask[1]);
            }
            // deposit runs of hk into keep; checked against pdep for every bit and all
            uint32_t n = 0, src = 0;
            for (uint32_t b = 0; b < hi_bits;) {
                if (!((keep >> b) & 1u)) { ++b; continue; }
                uint32_t b0 = b;
                while (b < hi_bits && ((keep >> b) & 1u)) ++b;
                if (n >= (uint32_t)kFusedDepRuns) return set_error(QB_ERR_UNSUPPORTED, "fused op %d: deposit runs", k_op);
                e.dep_run[n++] = src | ((b - b0) << 5) | (b0 << 10);
                src += b - b0;
            }
            e.n_dep = n;
            e.hk_bits = src;
            for (uint32_t b = 0; b <= e.hk_bits; ++b) {
                const uint32_t hk = b < e.hk_bits ? (1u << b) : ((1u << e.hk_bits) - 1u);
                uint32_t hi = 0;
                for (uint32_t r = 0; r < n; ++r) {
                    const uint32_t f = e.dep_run[r], wd = (f >> 5) & 31u;
                    hi |= ((hk >> (f & 31u)) & ((1u << wd) - 1u)) << (f >> 10);
                }
                if (hi != pdep_host(hk, keep))
                    return set_error(QB_ERR_UNSUPPORTED, "fused op %d: deposit runs", k_op);
            }
            const uint64_t want2 = 1ull << e.hk_bits;
            const int grid2 = (int)(want2 < gcap ? want2 : gcap);
#define QB_FUSED2_CASE(RA_, RB_)                                                                         \
            if (ra == RA_ && rb == RB_) {                                                                \
                if (dtype == QB_C128)                                                                    \
                    hipLaunchKernelGGL((fused2_kernel<double, RA_, RB_>), dim3(grid2, 1, nz), dim3(256), 0, st, a, lo_tab, \
                                       x_tab, e);                                                        \
                else                                                                                     \
                    hipLaunchKernelGGL((fused2_kernel<float, RA_, RB_>), dim3(grid2, 1, nz), dim3(256), 0, st, a, lo_tab,  \
                                       x_tab, e);                                                        \
            }
#define QB_FUSED2_CASE_F(RA_, RB_)                                                                       \
            if (ra == RA_ && rb == RB_)                                                                  \
                hipLaunchKernelGGL((fused2_kernel<float, RA_, RB_>), dim3(grid2, 1, nz), dim3(256), 0, st, a, lo_tab, x_tab, e);
            QB_FUSED2_CASE(0, 1) QB_FUSED2_CASE(0, 2) QB_FUSED2_CASE(1, 0) QB_FUSED2_CASE(1, 1)
            QB_FUSED2_CASE(1, 2) QB_FUSED2_CASE(2, 0) QB_FUSED2_CASE(2, 1) QB_FUSED2_CASE(2, 2)
            QB_FUSED2_CASE_F(0, 3) QB_FUSED2_CASE_F(1, 3) QB_FUSED2_CASE_F(2, 3) QB_FUSED2_CASE_F(3, 0)
            QB_FUSED2_CASE_F(3, 1) QB_FUSED2_CASE_F(3, 2) QB_FUSED2_CASE_F(3, 3)
#undef QB_FUSED2_CASE_F
#undef QB_FUSED2_CASE
            QB_HIP(hipGetLastError());
            return QB_OK;
        }
    }
    const uint64_t want = 1ull << (rank_out - lo_bits);     // one block per value of the high bits
    const int grid = (int)(want < gcap ? want : gcap);
#define QB_FUSED_CASE(N)                                                                      \
    case N:                                                                                   \
        if (dtype == QB_C128)                                                                 \
            hipLaunchKernelGGL((fused_kernel<double, N>), dim3(grid, 1, nz), dim3(256), 0, st, a, lo_tab, x_tab); \
        else                                                                                  \
            hipLaunchKernelGGL((fused_kernel<float, N>), dim3(grid, 1, nz), dim3(256), 0, st, a, lo_tab, x_tab);  \
        break;
    switch (n_in) {
        QB_FUSED_CASE(1)
        QB_FUSED_CASE(2)
        QB_FUSED_CASE(3)
        QB_FUSED_CASE(4)
        QB_FUSED_CASE(5)
        QB_FUSED_CASE(6)
        QB_FUSED_CASE(7)
        QB_FUSED_CASE(8)
    }
#undef QB_FUSED_CASE
    QB_HIP(hipGetLastError());
    return QB_OK;
}

namespace qb_internal {
int fail(int code, const char* msg) { return set_error(code, "%s", msg); }
}  // namespace qb_internal

// Host check of one executor plan at word `off` of a program: every record and
// lane table lies inside the program, and every LDS element an op, a gate or the
// result touches lies inside the item's data region (lds_elems elements: the
// per-wave angle copies and the queue ticket sit right after it), every global
// location inside the item's scratch slot (scratch_elems elements), every angle
// index inside the item's angle vector (n_angles doubles from the item's
// offset, plus the kParamRegs words of padding the executor may read), and
// every offset the kernel derives from a mask agrees with that mask.
// `const_ok`: whether gates of kind kConst (copied from the constant pool) are
// allowed.
static int check_item_plan(const int32_t* prog, int64_t words, int64_t off, int32_t item_threads, int32_t lds_elems,
                           int64_t scratch_elems, int64_t n_angles, bool const_ok) {
    if (off < 0 || off + kPlanHdr > words) return set_error(QB_ERR_INVALID, "plan at %lld: header outside the program", (long long)off);
    const int32_t* p = prog + off;
    const int32_t n_gates = p[0], n_ops = p[1], gates_off = p[2], ops_off = p[3];
    const int64_t lds_need = (uint32_t)p[5];
    const int64_t glb_need = (int64_t)(uint32_t)p[6] | ((int64_t)(uint32_t)p[9] << 32);
    if (n_gates < 0 || n_ops < 0 || gates_off < kPlanHdr || ops_off < kPlanHdr)
        return set_error(QB_ERR_INVALID, "plan at %lld: bad header", (long long)off);
    if (p[11] != 0)
        return set_error(QB_ERR_INVALID, "plan at %lld has %d open variables; the executor returns scalars",
                         (long long)off, p[11]);
    if (lds_need > lds_elems || lds_need > 65536)
        return set_error(QB_ERR_INVALID, "plan at %lld needs %lld LDS elements, the launch reserves %d",
                         (long long)off, (long long)lds_need, lds_elems);
    if (glb_need > scratch_elems)
        return set_error(QB_ERR_INVALID, "plan at %lld needs %lld scratch elements, the slot holds %lld",
                         (long long)off, (long long)glb_need, (long long)scratch_elems);
    if (off + gates_off + (int64_t)n_gates * kGateWords > words)
        return set_error(QB_ERR_INVALID, "plan at %lld: gate table outside the program", (long long)off);
    auto lds_ok = [&](int64_t first, int64_t last) { return first >= 0 && first <= last && last < lds_need; };
    auto loc_ok = [&](uint32_t loc, int64_t n) {
        if (loc & kGlobal) return (int64_t)(loc & ~kGlobal) + n <= glb_need;
        return lds_ok((int64_t)loc, (int64_t)loc + n - 1);
    };
    for (int32_t g = 0; g < n_gates; ++g) {
        const int32_t* e = p + gates_off + (int64_t)g * kGateWords;
        const int32_t kind = e[0], dst = e[1], n = e[2];
        if (kind < kConst || kind > kM || (kind == kConst && !const_ok))
            return set_error(QB_ERR_INVALID, "plan at %lld gate %d: kind %d has no device builder", (long long)off, g, kind);
        if (n < 1 || n > 8 || dst < 4 || !lds_ok(dst, (int64_t)dst + n - 1))
            return set_error(QB_ERR_INVALID, "plan at %lld gate %d outside the LDS data", (long long)off, g);
        for (int k = 0; k < 2; ++k) {
            const int32_t j = e[4 + 2 * k];
            if (j < -1 || (int64_t)j >= n_angles)
                return set_error(QB_ERR_INVALID, "plan at %lld gate %d: angle index %d", (long long)off, g, j);
        }
        if ((e[11] & 0xff) > 2) return set_error(QB_ERR_INVALID, "plan at %lld gate %d: slice word", (long long)off, g);
    }
    int64_t pos = off + ops_off;
    for (int32_t k = 0; k < n_ops; ++k) {
        if (pos + 4 > words) return set_error(QB_ERR_INVALID, "plan at %lld op %d outside the program", (long long)off, k);
        const uint32_t w0 = (uint32_t)prog[pos];
        const int64_t len = w0 >> 16;
        if (pos + len > words || len < 4)
            return set_error(QB_ERR_INVALID, "plan at %lld op %d: record length %lld", (long long)off, k, (long long)len);
        const int32_t* r = prog + pos;
        if (w0 & kOpGather) {
            const uint32_t n_act = (uint32_t)r[1], tab = (uint32_t)r[2], ni = (uint32_t)r[3];
            if (ni < 1 || ni > (uint32_t)kExecMaxIn || n_act < 1 || n_act > (uint32_t)item_threads ||
                (int64_t)tab + (int64_t)n_act * (ni + 1) > len)
                return set_error(QB_ERR_INVALID, "plan at %lld op %d: gather table", (long long)off, k);
            for (uint32_t l = 0; l < n_act; ++l) {
                const int32_t* t = r + tab + (int64_t)l * (ni + 1);
                if (!lds_ok((uint32_t)t[0], (uint32_t)t[0]) || (uint32_t)t[0] < 2)
                    return set_error(QB_ERR_INVALID, "plan at %lld op %d lane %u writes outside the LDS data", (long long)off, k, l);
                for (uint32_t j = 0; j < ni; ++j) {
                    const uint32_t w = (uint32_t)t[1 + j];
                    if (!lds_ok(w & 0xffffu, w & 0xffffu) || !lds_ok(w >> 16, w >> 16))
                        return set_error(QB_ERR_INVALID, "plan at %lld op %d lane %u reads outside the LDS data", (long long)off, k, l);
                }
            }
        } else if (w0 & kOpReduce) {
            const uint32_t ro = (w0 >> 4) & 31u, kk = (uint32_t)r[1];
            if (ro > 6 || kk > 24 || len != kReduceHdr)
                return set_error(QB_ERR_INVALID, "plan at %lld op %d: reduction shape", (long long)off, k);
            const uint32_t out_loc = (uint32_t)r[2], in_loc = (uint32_t)r[3];
            if ((out_loc & kGlobal) || (in_loc & kGlobal) || !lds_ok(out_loc, (int64_t)out_loc + (1ll << ro) - 1) ||
                !lds_ok(in_loc, (int64_t)in_loc + (1ll << (ro + kk)) - 1) || out_loc < 2)
                return set_error(QB_ERR_INVALID, "plan at %lld op %d: reduction outside the LDS data", (long long)off, k);
        } else if (w0 & kOpFused) {
            return set_error(QB_ERR_INVALID, "plan at %lld op %d: fused ops run on the wide path only", (long long)off, k);
        } else {
            const uint32_t n_in = w0 & 15u, rank_out = (w0 >> 4) & 31u, sum = (w0 >> 9) & 1u;
            const bool all_lds = (w0 >> 12) & 1u;
            if (n_in < 1 || n_in > (uint32_t)kExecMaxIn || len != kStepHdr + kInWords * (int64_t)n_in || rank_out > 24)
                return set_error(QB_ERR_INVALID, "plan at %lld op %d: step shape", (long long)off, k);
            const uint32_t out_loc = (uint32_t)r[2];
            bool any_glb = (out_loc & kGlobal) != 0;
            if (!loc_ok(out_loc, 1ll << rank_out) || (!(out_loc & kGlobal) && out_loc < 2))
                return set_error(QB_ERR_INVALID, "plan at %lld op %d: output outside its memory", (long long)off, k);
            for (uint32_t j = 0; j < n_in; ++j) {
                const uint32_t loc = (uint32_t)r[3 + 4 * j], top = (uint32_t)r[4 + 4 * j];
                const uint32_t mask = (uint32_t)r[5 + 4 * j], runs = (uint32_t)r[6 + 4 * j];
                const int64_t n_mask = 1ll << __builtin_popcount(mask);
                if ((mask >> rank_out) != 0u || runs != runs_word(mask) ||
                    !(top == 0u || (sum && top == (uint32_t)n_mask)))
                    return set_error(QB_ERR_INVALID, "plan at %lld op %d input %u: mask, runs or top", (long long)off, k, j);
                if (!loc_ok(loc, n_mask * (top ? 2 : 1)))
                    return set_error(QB_ERR_INVALID, "plan at %lld op %d input %u outside its memory", (long long)off, k, j);
                any_glb = any_glb || (loc & kGlobal);
            }
            if (all_lds && any_glb)
                return set_error(QB_ERR_INVALID, "plan at %lld op %d: marked all-LDS but touches scratch", (long long)off, k);
            if (any_glb && scratch_elems <= 0)
                return set_error(QB_ERR_INVALID, "plan at %lld op %d touches scratch, the launch has none", (long long)off, k);
        }
        pos += len;
    }
    const uint32_t rl = (uint32_t)p[4];
    if (!loc_ok(rl, 1)) return set_error(QB_ERR_INVALID, "plan at %lld: result outside its memory", (long long)off);
    return QB_OK;
}

namespace qb_internal {
// The executor launch behind qb_execute: `scratch` holds `n_slots_avail` slots
// of d->scratch_elems elements, `queue` one uint32.
int execute_impl(const qb_exec_desc* d, void* scratch, int32_t n_slots_avail, uint32_t* queue, void* stream) {
    if (d->dtype != QB_C64 && d->dtype != QB_C128) return set_error(QB_ERR_INVALID, "dtype %d", d->dtype);
    if (d->n_items < 0) return set_error(QB_ERR_INVALID, "n_items < 0");
    if (!d->program || !d->items || !d->params || !d->item_out || !queue)
        return set_error(QB_ERR_INVALID, "null device pointer in descriptor");
    const int32_t group = d->group > 1 ? d->group : 1;
    if (d->n_items % group != 0)
        return set_error(QB_ERR_INVALID, "n_items %d is not a multiple of group %d", d->n_items, group);
    const bool glb = d->scratch_elems > 0;
    int32_t n_slots = 0;
    int per_cu = 0;
    int rc = slots_for(d->item_threads, d->lds_elems, d->dtype, group, glb, &n_slots, &per_cu);
    if (rc) return rc;
    const int32_t n_groups = d->n_items / group;
    int grid = per_cu * g_num_cus;
    if (n_groups < grid) grid = n_groups;
    if (grid < 1) grid = 1;
    if ((int64_t)grid * group > n_slots_avail && glb)
        return set_error(QB_ERR_INVALID, "scratch holds %d slots, launch needs %lld", n_slots_avail,
                         (long long)grid * group);
    if (d->scratch_elems > 0 && !scratch) return set_error(QB_ERR_INVALID, "scratch is null");
    if (d->scratch_elems < 0 || d->scratch_elems * (d->dtype == QB_C128 ? 16 : 8) * group > (int64_t)0xFFFFFFFF)
        return set_error(QB_ERR_UNSUPPORTED, "scratch of a group (%d x %lld elements) exceeds 4 GiB", group,
                         (long long)d->scratch_elems);
    if (d->program_words < 0 || d->program_words > (1 << 30))
        return set_error(QB_ERR_INVALID, "program_words %d outside 0..2^30", d->program_words);
    ExecArgs a;
    a.program = d->program;
    a.program_bytes = d->program_words ? 4u * (uint32_t)d->program_words : 0xFFFFFFFFu;
    a.items = d->items;
    a.n_items = d->n_items;
    a.lds_elems = d->lds_elems;
    a.params = d->params;
    a.pool = d->const_pool;
    a.scratch = scratch;
    a.scratch_elems = d->scratch_elems;
    a.item_out = d->item_out;
    a.queue = queue;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (grid < n_groups) {
        // a fill kernel, not a memset: the launch may be captured in a graph
        hipLaunchKernelGGL(zero_u32, dim3(1), dim3(1), 0, st, queue);
        QB_HIP(hipGetLastError());
    }
    const int elem = d->dtype == QB_C128 ? 16 : 8;
    const int lds_bytes = d->lds_elems * elem * group + exec_aux_bytes(d->item_threads, group);
    return exec_dispatch(d->dtype, d->item_threads, group, glb,
                         [&](auto cfg) { return exec_launch_cfg(cfg, a, grid, lds_bytes, st); });
}

int resident_items(const qb_exec_desc* d, int32_t* n_slots) {
    int per_cu = 0;
    return slots_for(d->item_threads, d->lds_elems, d->dtype, d->group > 1 ? d->group : 1, true, n_slots, &per_cu);
}

int current_device() { return g_device; }
}  // namespace qb_internal

extern "C" {

int qb_version(void) { return 100; }

const char* qb_last_error(void) { return g_last_error.c_str(); }

int qb_init(int device) {
    std::lock_guard<std::mutex> lock(g_mutex);
    int n = 0;
    QB_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return set_error(QB_ERR_INVALID, "device %d of %d", device, n);
    QB_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    QB_HIP(hipGetDeviceProperties(&prop, device));
    g_num_cus = prop.multiProcessorCount;
    // this device's copy of qb_execute_noalloc's static workspace
    const int rc = qb_internal::noalloc_init(device);
    if (rc) return rc;
    g_device = device;
    return QB_OK;
}

int qb_shutdown(void) {
    std::lock_guard<std::mutex> lock(g_mutex);
    g_device = -1;
    return QB_OK;
}

int qb_step(const qb_step_desc* d, const void* const* inputs, void* out, void* stream) {
    if (g_device < 0) return set_error(QB_ERR_NOT_INIT, "qb_init was not called");
    if (!d || !inputs || !out) return set_error(QB_ERR_INVALID, "null descriptor or pointer");
    if (d->n_inputs < 1 || d->n_inputs > QB_MAX_STEP_INPUTS)
        return set_error(QB_ERR_INVALID, "n_inputs=%d outside [1,%d]", d->n_inputs, QB_MAX_STEP_INPUTS);
    if (d->rank_u < 0 || d->rank_u > QB_MAX_RANK)
        return set_error(QB_ERR_UNSUPPORTED, "rank_u=%d outside [0,%d]", d->rank_u, QB_MAX_RANK);
    if (d->sum && d->rank_u < 1) return set_error(QB_ERR_INVALID, "sum over an empty union");
    if (d->dtype != QB_C64 && d->dtype != QB_C128) return set_error(QB_ERR_INVALID, "dtype %d", d->dtype);
    const uint32_t full = d->rank_u >= 32 ? 0xffffffffu : ((1u << d->rank_u) - 1u);
    StepArgs a;
    memset(&a, 0, sizeof(a));
    a.n_in = (uint32_t)d->n_inputs;
    a.n_sum = d->sum ? 2u : 1u;
    a.rank_out = (uint32_t)(d->rank_u - (d->sum ? 1 : 0));
    a.out = out;
    const uint32_t top_bit = d->sum ? (1u << (d->rank_u - 1)) : 0u;
    for (int j = 0; j < d->n_inputs; ++j) {
        const uint32_t m = d->masks[j];
        if (m & ~full) return set_error(QB_ERR_INVALID, "mask %d has bits outside rank %d", j, d->rank_u);
        if (!inputs[j]) return set_error(QB_ERR_INVALID, "input %d is null", j);
        StepInput& in = a.in[j];
        in.ptr = inputs[j];
        const uint32_t low = m & ~top_bit;
        in.mask = low;
        in.top = (m & top_bit) ? (1u << __builtin_popcount(low)) : 0u;
        in.n_runs = (uint32_t)mask_runs(low, in.src, in.w, in.dst, kStepMaxRuns);
    }
    const uint64_t n_out = 1ull << a.rank_out;
    const uint64_t want = (n_out + 255) / 256;
    const uint64_t cap = (uint64_t)g_num_cus * 8;
    const int grid = (int)(want < cap ? (want > 0 ? want : 1) : cap);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (d->n_inputs <= 4) {
        uint32_t runs[QB_MAX_STEP_INPUTS];
        for (int j = 0; j < d->n_inputs; ++j) runs[j] = runs_word(a.in[j].mask);
        switch (d->n_inputs) {
        case 1: return launch_step_k<1>(a, runs, d->dtype, grid, st);
        case 2: return launch_step_k<2>(a, runs, d->dtype, grid, st);
        case 3: return launch_step_k<3>(a, runs, d->dtype, grid, st);
        default: return launch_step_k<4>(a, runs, d->dtype, grid, st);
        }
    }
    if (d->dtype == QB_C128)
        hipLaunchKernelGGL(step_kernel<double>, dim3(grid), dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL(step_kernel<float>, dim3(grid), dim3(256), 0, st, a);
    QB_HIP(hipGetLastError());
    return QB_OK;
}

int qb_exec_slots(int32_t item_threads, int32_t lds_elems, int32_t dtype, int32_t group, int32_t* n_slots) {
    if (!n_slots) return set_error(QB_ERR_INVALID, "n_slots is null");
    int per_cu = 0;
    return slots_for(item_threads, lds_elems, dtype, group > 1 ? group : 1, true, n_slots, &per_cu);
}

int qb_execute(const qb_exec_desc* d, void* stream) {
    if (!d) return set_error(QB_ERR_INVALID, "null descriptor");
    if (!d->queue) return set_error(QB_ERR_INVALID, "null device pointer in descriptor");
    return qb_internal::execute_impl(d, d->scratch, d->n_slots, d->queue, stream);
}

int qb_wide_execute(const qb_wide_desc* d, void* stream) {
    if (g_device < 0) return set_error(QB_ERR_NOT_INIT, "qb_init was not called");
    if (!d || !d->plan_host || !d->plan_dev || !d->params || !d->arena || !d->out)
        return set_error(QB_ERR_INVALID, "null pointer in wide descriptor");
    if (d->dtype != QB_C64 && d->dtype != QB_C128) return set_error(QB_ERR_INVALID, "dtype %d", d->dtype);
    const int32_t* w = d->plan_host;
    if (d->plan_words < kPlanHdr) return set_error(QB_ERR_INVALID, "plan shorter than its header");
    const int32_t n_gates = w[0], n_ops = w[1], gates_off = w[2], ops_off = w[3];
    const uint32_t result_loc = (uint32_t)w[4];
    const int64_t elem = d->dtype == QB_C128 ? 16 : 8;
    const int64_t cap = d->arena_elems;
    // slices z = 0 .. nz-1 (assignments slice_bits + z) in lockstep: every
    // launch covers all of them (grid z), slice z in arena region z
    const int nz = d->n_slices > 0 ? d->n_slices : 1;
    if (nz > 1024) return set_error(QB_ERR_INVALID, "n_slices %d > 1024", d->n_slices);
    if ((uint64_t)d->slice_bits + (uint64_t)(nz - 1) > 0xffffffffull)
        return set_error(QB_ERR_INVALID, "slice assignments beyond 32 bits");
    if (cap < 0 || cap > ((int64_t)1 << 40)) return set_error(QB_ERR_INVALID, "arena_elems %lld", (long long)cap);
    const int64_t zbytes = cap * elem;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if ((int64_t)gates_off + (int64_t)n_gates * kGateWords > d->plan_words)
        return set_error(QB_ERR_INVALID, "gate table outside the plan");
    for (int32_t g = 0; g < n_gates; ++g) {
        const int32_t* e = w + gates_off + (size_t)g * kGateWords;
        if (e[0] <= kConst || e[0] > kM)   // no constant pool on the wide path
            return set_error(QB_ERR_INVALID, "gate %d: kind %d has no device builder", g, e[0]);
        if (e[2] < 1 || e[2] > 8 || e[1] < 0 || (int64_t)e[1] + e[2] > cap)
            return set_error(QB_ERR_INVALID, "gate %d outside the arena", g);
    }
    if (n_gates > 0) {
        const int blocks = (n_gates + 255) / 256;
        if (d->dtype == QB_C128)
            hipLaunchKernelGGL(build_gates_global<double>, dim3(blocks, 1, nz), dim3(256), 0, st,
                               d->plan_dev + gates_off, n_gates, d->params, d->slice_bits, d->arena, zbytes);
        else
            hipLaunchKernelGGL(build_gates_global<float>, dim3(blocks, 1, nz), dim3(256), 0, st,
                               d->plan_dev + gates_off, n_gates, d->params, d->slice_bits, d->arena, zbytes);
        QB_HIP(hipGetLastError());
    }
    char* arena = reinterpret_cast<char*>(d->arena);
    // the gate area ends where the last gate tensor does; consecutive fused ops
    // that read it only (each checked as any fused op) run in one launch
    int64_t gate_end = 0;
    for (int32_t g = 0; g < n_gates; ++g) {
        const int32_t* e = w + gates_off + (size_t)g * kGateWords;
        if ((int64_t)e[1] + e[2] > gate_end) gate_end = (int64_t)e[1] + e[2];
    }
    OpRun run;
    run.n = 0;
    auto flush_run = [&]() -> int {
        if (run.n == 0) return QB_OK;
        if (d->dtype == QB_C128)
            hipLaunchKernelGGL(gate_ops_kernel<double>, dim3(run.n, 1, nz), dim3(256), 0, st, d->plan_dev, run,
                               arena, zbytes);
        else
            hipLaunchKernelGGL(gate_ops_kernel<float>, dim3(run.n, 1, nz), dim3(256), 0, st, d->plan_dev, run,
                               arena, zbytes);
        run.n = 0;
        QB_HIP(hipGetLastError());
        return QB_OK;
    };
    int64_t pos = ops_off;
    for (int32_t k = 0; k < n_ops; ++k) {
        if (pos + kStepHdr > d->plan_words) return set_error(QB_ERR_INVALID, "op %d outside the plan", k);
        const uint32_t w0 = (uint32_t)w[pos];
        if (w0 & (kOpGather | kOpReduce)) return set_error(QB_ERR_INVALID, "wide plans take step ops only");
        if (w0 & kOpFused) {
            // a candidate for a run of gate-only ops, from the record as it
            // stands (launch_fused then checks it like any fused op)
            const uint32_t n_in = w0 & 15u;
            bool gate_only = !(w0 & kFusedChain) && n_in >= 1 && n_in <= (uint32_t)kFusedMaxIn &&
                             pos + kFusedHdr + 4 * (int64_t)n_in <= d->plan_words && pos <= INT32_MAX;
            if (gate_only) {
                const uint32_t meta = (uint32_t)w[pos + 1];
                const uint32_t ro = (meta >> 8) & 255u, kk = (meta >> 16) & 255u;
                gate_only = ro + kk < 32u && ((uint64_t)n_in << (ro + kk)) <= (uint64_t)kRunMaxTerms &&
                            (int64_t)((uint32_t)w[pos + 2] & kLocBits) >= gate_end;
            }
            for (uint32_t j = 0; j < n_in && gate_only; ++j) {
                const uint32_t loc = (uint32_t)w[pos + kFusedHdr + 4 * j] & kLocBits;
                const uint32_t mask = (uint32_t)w[pos + kFusedHdr + 4 * j + 1];
                gate_only = (int64_t)loc + (1ll << __builtin_popcount(mask)) <= gate_end;
            }
            int rc;
            if (gate_only) {
                if ((rc = launch_fused(w, d->plan_words, d->plan_dev, pos, k, arena, elem, cap, d->dtype, st, nz,
                                       zbytes, true)))            // checks only
                    return rc;
                run.pos[run.n++] = (int32_t)pos;
                if (run.n == kRunMaxOps && (rc = flush_run())) return rc;
            } else {
                if ((rc = flush_run())) return rc;
                if ((rc = launch_fused(w, d->plan_words, d->plan_dev, pos, k, arena, elem, cap, d->dtype, st, nz,
                                       zbytes, false)))
                    return rc;
            }
            pos += w0 >> 16;
            continue;
        }
        {
            const int rc = flush_run();
            if (rc) return rc;
        }
        const uint32_t n_in = w0 & 15u, rank_out = (w0 >> 4) & 31u, sum = (w0 >> 9) & 1u;
        const uint32_t out_loc = (uint32_t)w[pos + 2];
        if (n_in < 1 || n_in > (uint32_t)kExecMaxIn || pos + kStepHdr + kInWords * (int64_t)n_in > d->plan_words)
            return set_error(QB_ERR_INVALID, "op %d has %u inputs", k, n_in);
        if (!(out_loc & kGlobal) || (int64_t)(out_loc & ~kGlobal) + (1ll << rank_out) > cap)
            return set_error(QB_ERR_INVALID, "op %d output outside the arena", k);
        StepArgs a;
        memset(&a, 0, sizeof(a));
        a.n_in = n_in;
        a.n_sum = sum ? 2u : 1u;
        a.rank_out = rank_out;
        a.out = arena + (int64_t)(out_loc & ~kGlobal) * elem;
        uint32_t runs[kExecMaxIn];
        for (uint32_t j = 0; j < n_in; ++j) {
            const int32_t* in = w + pos + kStepHdr + kInWords * j;
            const uint32_t loc = (uint32_t)in[0];
            const uint32_t top = (uint32_t)in[1], mask = (uint32_t)in[2];
            const int64_t n_elems = (1ll << __builtin_popcount(mask)) * (top ? 2 : 1);
            // the kernel reads the summed variable's 1 at offset + top
            if (top != 0u && !(sum && (int64_t)top == (1ll << __builtin_popcount(mask))))
                return set_error(QB_ERR_INVALID, "op %d input %u: top %u is not 2^popcount(mask)", k, j, top);
            if (!(loc & kGlobal) || (int64_t)(loc & ~kGlobal) + n_elems > cap)
                return set_error(QB_ERR_INVALID, "op %d input %u outside the arena", k, j);
            if (mask >> rank_out) return set_error(QB_ERR_INVALID, "op %d input %u mask beyond rank", k, j);
            a.in[j].ptr = arena + (int64_t)(loc & ~kGlobal) * elem;
            a.in[j].mask = mask;
            a.in[j].top = top;
            runs[j] = runs_word(mask);
        }
        const uint64_t n_out = 1ull << rank_out;
        const uint64_t want = (n_out + 255) / 256;
        const uint64_t gcap = (uint64_t)g_num_cus * 8;
        const int grid = (int)(want < gcap ? (want > 0 ? want : 1) : gcap);
        int rc;
        switch (n_in) {
        case 1: rc = launch_step_k<1>(a, runs, d->dtype, grid, st, nz, zbytes); break;
        case 2: rc = launch_step_k<2>(a, runs, d->dtype, grid, st, nz, zbytes); break;
        case 3: rc = launch_step_k<3>(a, runs, d->dtype, grid, st, nz, zbytes); break;
        default: rc = launch_step_k<4>(a, runs, d->dtype, grid, st, nz, zbytes); break;
        }
        if (rc) return rc;
        pos += w0 >> 16;
    }
    {
        const int rc = flush_run();
        if (rc) return rc;
    }
    const int32_t rank = w[11];
    if (rank < 0 || rank > 24) return set_error(QB_ERR_INVALID, "result rank %d", rank);
    const uint32_t n_res = 1u << rank;
    if (!(result_loc & kGlobal) || (int64_t)(result_loc & ~kGlobal) + n_res > cap)
        return set_error(QB_ERR_INVALID, "result outside the arena");
    const int64_t stride = d->out_stride > 0 ? d->out_stride : 1;
    const int rblocks = (int)((n_res + 255) / 256 < 1024 ? (n_res + 255) / 256 : 1024);
    if (nz > 1 && n_res > 1 && stride < nz)
        return set_error(QB_ERR_INVALID, "out_stride %lld < n_slices %d: slice results would overlap",
                         (long long)stride, nz);
    if (d->dtype == QB_C128)
        hipLaunchKernelGGL(store_result<double>, dim3(rblocks, 1, nz), dim3(256), 0, st, d->arena,
                           result_loc & ~kGlobal, n_res, stride, d->out, zbytes);
    else
        hipLaunchKernelGGL(store_result<float>, dim3(rblocks, 1, nz), dim3(256), 0, st, d->arena,
                           result_loc & ~kGlobal, n_res, stride, d->out, zbytes);
    QB_HIP(hipGetLastError());
    return QB_OK;
}

int qb_reduce_terms(const double* item_out, const int32_t* seg, int32_t n_segments,
                    double* term_out, double* total, void* stream) {
    if (g_device < 0) return set_error(QB_ERR_NOT_INIT, "qb_init was not called");
    if (!item_out || !term_out || !total) return set_error(QB_ERR_INVALID, "null pointer");
    if (n_segments < 1 || n_segments > (1 << 20))
        return set_error(QB_ERR_INVALID, "n_segments=%d outside [1, 2^20]", n_segments);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (!seg) {
        // one item per term: the terms are the item results (one launch)
        hipLaunchKernelGGL(reduce_total, dim3(1), dim3(256), 0, st, item_out, n_segments, total, term_out);
        QB_HIP(hipGetLastError());
        return QB_OK;
    }
    hipLaunchKernelGGL(reduce_segments, dim3(n_segments), dim3(64), 0, st, item_out, seg, n_segments,
                       term_out);
    QB_HIP(hipGetLastError());
    hipLaunchKernelGGL(reduce_total, dim3(1), dim3(256), 0, st, term_out, n_segments, total,
                       (double*)nullptr);
    QB_HIP(hipGetLastError());
    return QB_OK;
}

// ---------------------------------------------------------------------------
// Plan handles: a batched program, its work items and every buffer one
// energy evaluation needs, owned by the library.
// ---------------------------------------------------------------------------

struct qb_plan {
    int device = -1;
    int32_t dtype = QB_C64;
    int32_t item_threads = 0;
    int32_t lds_elems = 0;
    int32_t n_items = 0;
    int32_t n_terms = 0;
    int32_t n_params = 0;
    int32_t n_slots = 0;
    int32_t group = 1;
    int64_t scratch_elems = 0;
    int64_t program_words = 0;
    bool unit_seg = false;       // every term has exactly one item
    int32_t* program = nullptr;
    int32_t* items = nullptr;
    int32_t* seg = nullptr;
    double* item_out = nullptr;
    void* scratch = nullptr;
    uint32_t* queue = nullptr;
    double* pool = nullptr;
};

static void plan_free(qb_plan* p) {
    if (!p) return;
    if (p->device >= 0) (void)hipSetDevice(p->device);
    (void)hipFree(p->program);
    (void)hipFree(p->items);
    (void)hipFree(p->seg);
    (void)hipFree(p->item_out);
    (void)hipFree(p->scratch);
    (void)hipFree(p->queue);
    (void)hipFree(p->pool);
    delete p;
}

int qb_check_program(const int32_t* program, int64_t program_words, const int32_t* items, int32_t n_items,
                     int32_t n_params, int32_t item_threads, int32_t lds_elems, int64_t scratch_elems,
                     int32_t group) {
    if (!program || !items || n_items < 0 || program_words < 1 || n_params < 0 || lds_elems < 0 || scratch_elems < 0 ||
        group < 0)
        return set_error(QB_ERR_INVALID, "null or negative program, items or sizes");
    if (group > 1) {
        if (group != 2) return set_error(QB_ERR_INVALID, "group %d: 1 or 2", group);
        if (n_items % group != 0)
            return set_error(QB_ERR_INVALID, "n_items %d is not a multiple of group %d", n_items, group);
        for (int32_t i = 0; i < n_items; ++i)
            if (items[(size_t)i * QB_ITEM_WORDS] != items[(size_t)(i - i % group) * QB_ITEM_WORDS])
                return set_error(QB_ERR_INVALID, "item %d: its group's items do not share one plan", i);
    }
    if (program_words > (1 << 30))
        return set_error(QB_ERR_INVALID, "program_words=%lld > 2^30", (long long)program_words);
    if (item_threads != 64 && item_threads != 128 && item_threads != 256 && item_threads != 512 && item_threads != 1024)
        return set_error(QB_ERR_INVALID, "item_threads %d not in {64,128,256,512,1024}", item_threads);
    // output slots must be a permutation of 0..n_items-1; angle availability per plan
    std::vector<uint8_t> slot_seen((size_t)n_items, 0);
    std::map<int32_t, int64_t> plan_angles;   // plan offset -> fewest angles any of its items has
    for (int32_t i = 0; i < n_items; ++i) {
        const int32_t* it = items + (size_t)i * QB_ITEM_WORDS;
        if (it[0] < 0 || (int64_t)it[0] + kPlanHdr > program_words)
            return set_error(QB_ERR_INVALID, "item %d: plan offset %d outside the program", i, it[0]);
        if (it[1] < 0 || it[1] > n_params) return set_error(QB_ERR_INVALID, "item %d: angle offset %d", i, it[1]);
        if (it[3] < 0 || it[3] >= n_items || slot_seen[(size_t)it[3]])
            return set_error(QB_ERR_INVALID, "item %d: output slot %d (slots must be a permutation)", i, it[3]);
        slot_seen[(size_t)it[3]] = 1;
        const int64_t avail = (int64_t)n_params - it[1];
        auto f = plan_angles.find(it[0]);
        if (f == plan_angles.end()) plan_angles.emplace(it[0], avail);
        else if (avail < f->second) f->second = avail;
    }
    for (const auto& kv : plan_angles) {
        const int rc = check_item_plan(program, program_words, kv.first, item_threads, lds_elems, scratch_elems,
                                       kv.second, false);
        if (rc) return rc;
        const int32_t n_sliced = program[kv.first + 7];
        if (n_sliced < 0 || n_sliced > 24)
            return set_error(QB_ERR_INVALID, "plan at %d: %d sliced variables", kv.first, n_sliced);
    }
    return QB_OK;
}

int qb_plan_create(const qb_plan_desc* d, qb_plan** out) {
    if (g_device < 0) return set_error(QB_ERR_NOT_INIT, "qb_init was not called");
    if (!d || !out) return set_error(QB_ERR_INVALID, "null descriptor or output");
    *out = nullptr;
    if (d->dtype != QB_C64 && d->dtype != QB_C128) return set_error(QB_ERR_INVALID, "dtype %d", d->dtype);
    if (!d->program || d->program_words < 1 || !d->items || d->n_items < 1 || !d->seg || d->n_terms < 1 ||
        d->n_params < 0 || d->scratch_elems < 0)
        return set_error(QB_ERR_INVALID, "empty or negative program, items, segments or sizes");
    if (d->n_terms > (1 << 20)) return set_error(QB_ERR_INVALID, "n_terms=%d > 2^20", d->n_terms);
    if (d->program_words > (1 << 30))
        return set_error(QB_ERR_INVALID, "program_words=%lld > 2^30", (long long)d->program_words);
    // host-side checks of what the kernels will index
    for (int32_t t = 0; t < d->n_terms; ++t)
        if (d->seg[t] > d->seg[t + 1]) return set_error(QB_ERR_INVALID, "segments must not decrease");
    const int32_t group = d->group > 1 ? d->group : 1;
    // every item of a term in some segment; padding items (completing a group)
    // may follow the last segment
    if (d->seg[0] != 0 || d->seg[d->n_terms] > d->n_items || d->seg[d->n_terms] + group <= d->n_items)
        return set_error(QB_ERR_INVALID, "segments must cover items 0..n_items (up to one group's padding)");
    bool unit_seg = true;
    for (int32_t t = 0; t < d->n_terms; ++t) unit_seg = unit_seg && d->seg[t + 1] - d->seg[t] == 1;
    int rc0 = qb_check_program(d->program, d->program_words, d->items, d->n_items, d->n_params, d->item_threads,
                               d->lds_elems, d->scratch_elems, group);
    if (rc0) return rc0;
    int32_t n_slots = 0;
    int per_cu = 0;
    int rc = slots_for(d->item_threads, d->lds_elems, d->dtype, group, true, &n_slots, &per_cu);
    if (rc) return rc;
    qb_plan* p = new qb_plan();
    p->device = g_device;
    p->dtype = d->dtype;
    p->item_threads = d->item_threads;
    p->lds_elems = d->lds_elems;
    p->n_items = d->n_items;
    p->n_terms = d->n_terms;
    p->n_params = d->n_params;
    p->group = group;
    p->n_slots = n_slots < d->n_items ? n_slots : d->n_items;
    p->scratch_elems = d->scratch_elems;
    p->program_words = d->program_words;
    p->unit_seg = unit_seg;
    const size_t elem = d->dtype == QB_C128 ? 16 : 8;
    hipError_t e = hipSuccess;
#define QB_PLAN_TRY(call)                                                          \
    do {                                                                           \
        e = (call);                                                                \
        if (e != hipSuccess) {                                                     \
            plan_free(p);                                                          \
            return set_error(QB_ERR_HIP, "%s: %s", #call, hipGetErrorString(e));   \
        }                                                                          \
    } while (0)
    QB_PLAN_TRY(hipMalloc(&p->program, sizeof(int32_t) * (size_t)d->program_words));
    QB_PLAN_TRY(hipMalloc(&p->items, sizeof(int32_t) * QB_ITEM_WORDS * (size_t)d->n_items));
    QB_PLAN_TRY(hipMalloc(&p->seg, sizeof(int32_t) * ((size_t)d->n_terms + 1)));
    QB_PLAN_TRY(hipMalloc(&p->item_out, sizeof(double) * 2 * (size_t)d->n_items));
    QB_PLAN_TRY(hipMalloc(&p->queue, sizeof(uint32_t)));
    QB_PLAN_TRY(hipMalloc(&p->pool, sizeof(double) * 4));
    if (d->scratch_elems > 0)
        QB_PLAN_TRY(hipMalloc(&p->scratch, elem * (size_t)d->scratch_elems * (size_t)p->n_slots));
    QB_PLAN_TRY(hipMemcpy(p->program, d->program, sizeof(int32_t) * (size_t)d->program_words,
                          hipMemcpyHostToDevice));
    QB_PLAN_TRY(hipMemcpy(p->items, d->items, sizeof(int32_t) * QB_ITEM_WORDS * (size_t)d->n_items,
                          hipMemcpyHostToDevice));
    QB_PLAN_TRY(hipMemcpy(p->seg, d->seg, sizeof(int32_t) * ((size_t)d->n_terms + 1), hipMemcpyHostToDevice));
    QB_PLAN_TRY(hipMemset(p->queue, 0, sizeof(uint32_t)));
    QB_PLAN_TRY(hipMemset(p->pool, 0, sizeof(double) * 4));
#undef QB_PLAN_TRY
    *out = p;
    return QB_OK;
}

int qb_plan_execute(qb_plan* p, const double* params, double* term_out, double* total, void* stream) {
    if (!p) return set_error(QB_ERR_INVALID, "null plan");
    if (!params || !term_out || !total) return set_error(QB_ERR_INVALID, "null device pointer");
    qb_exec_desc d;
    d.program = p->program;
    d.items = p->items;
    d.n_items = p->n_items;
    d.item_threads = p->item_threads;
    d.lds_elems = p->lds_elems;
    d.dtype = p->dtype;
    d.params = params;
    d.const_pool = p->pool;
    d.scratch = p->scratch;
    d.scratch_elems = p->scratch_elems;
    d.n_slots = p->n_slots;
    d.program_words = (int32_t)p->program_words;
    d.item_out = p->item_out;
    d.queue = p->queue;
    d.group = p->group;
    int rc = qb_execute(&d, stream);
    if (rc) return rc;
    return qb_reduce_terms(p->item_out, p->unit_seg ? nullptr : p->seg, p->n_terms, term_out, total, stream);
}

int qb_plan_destroy(qb_plan* p) {
    plan_free(p);
    return QB_OK;
}

}  // extern "C"
