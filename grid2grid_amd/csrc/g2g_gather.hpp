// Device code of the fused gather (gfx950, wave64): element operations, the
// copy queues, the LDS transposes, the scaled epilogue and the k_gather
// kernel template. Included by g2g_gather_variants.hip, which instantiates
// one (workgroup size, feature level) family of variants per object, and by
// g2g_kernels.hip for the shared element helpers.
//
// Data movement only: elements move as bit patterns through 4/8/16-byte
// vector registers; the one arithmetic path (beta*dst + alpha*op(src), used
// only when the caller scales) reproduces the reference build's x86
// arithmetic on bit patterns.
#pragma once

#include "g2g_index.hpp"

// no FMA contraction anywhere (the build also passes -ffp-contract=off)
#pragma clang fp contract(off)

#include <algorithm>

namespace g2g {

typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// A 16-byte vector at an address that is only 4-byte aligned. HSA code
// objects run in unaligned-access mode, where global vector accesses need
// dword alignment only.
typedef uint32_t u32x4a4 __attribute__((ext_vector_type(4), aligned(4)));

// Pointers into global memory (address space 1): the copy loops use global_*
// instructions rather than flat_* ones, which would also hold the LDS/scalar
// wait counter.
#define GLOBAL __attribute__((address_space(1)))
typedef GLOBAL unsigned char gbyte;
// Pointers into LDS (address space 3): 32-bit addresses whose constant parts
// fold into the ds_* instructions' offset fields.
#define LOCAL __attribute__((address_space(3)))

template <int B> struct bits_t;
template <> struct bits_t<4> { using type = uint32_t; };
template <> struct bits_t<8> { using type = u32x2; };
template <> struct bits_t<16> { using type = u32x4; };

enum { DT_F32 = 0, DT_F64 = 1, DT_C64 = 2, DT_C128 = 3 };

// ------------------------------------------------------------ element ops
//
// The scaled epilogue dst = beta*dst + alpha*op(src) (reference
// memory_utils.hpp:37, 237, 251) on bit patterns, with the arithmetic of the
// reference's x86-64 build: IEEE round-to-nearest products and sums (no FMA;
// denormals kept: the code object preserves f32 and f64 denormals), an
// invalid operation gives the x86 default NaN (sign and quiet bit set), an
// operation with a NaN operand returns the first source operand if it is a
// NaN, else the other one, quieted (Intel SDM vol. 1, table 4-7, SSE). The
// hardware's own NaN choice and default NaN (positive) are never used: NaN
// operands are resolved before the operation. Complex products are GCC's:
// inline (ac - bd, ad + bc), recomputed as libgcc's __mulsc3 / __muldc3 (C11
// Annex G.5.1, example _Cmultd) when both parts are NaN. Which operand is the
// first source depends on the reference's call site and the element's
// position in it (ref_order).

__device__ __forceinline__ double u2d(uint32_t lo, uint32_t hi) {
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
__device__ __forceinline__ u32x2 d2u(double v) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    return u32x2{(uint32_t)b, (uint32_t)(b >> 32)};
}

enum : uint32_t {
    ORD_M1S = 1u,  // beta*dst: beta first (complex: first __mul?c3 argument)
    ORD_M2S = 2u,  // alpha*src: alpha first
    ORD_ARE = 4u,  // real part of the sum: alpha*src first
    ORD_AIM = 8u   // imaginary part of the sum: alpha*src first
};

// IEEE single / double bit patterns with the x86 NaN rules; a is the first source.
struct xf32 {
    using U = uint32_t;
    static constexpr U SIGN = 0x80000000u, QUIET = 0x00400000u, DEFNAN = 0xffc00000u;
    static constexpr U INF = 0x7f800000u, ONE = 0x3f800000u;
    __device__ static bool nan(U x) { return (x & ~SIGN) > INF; }
    __device__ static bool inf(U x) { return (x & ~SIGN) == INF; }
    __device__ static U fix(U r) { return nan(r) ? DEFNAN : r; }
    __device__ static U mul(U a, U b) {
        if (nan(a)) return a | QUIET;
        if (nan(b)) return b | QUIET;
        return fix(__float_as_uint(__uint_as_float(a) * __uint_as_float(b)));
    }
    __device__ static U add(U a, U b) {
        if (nan(a)) return a | QUIET;
        if (nan(b)) return b | QUIET;
        return fix(__float_as_uint(__uint_as_float(a) + __uint_as_float(b)));
    }
    __device__ static U sub(U a, U b) {
        if (nan(a)) return a | QUIET;
        if (nan(b)) return b | QUIET;
        return fix(__float_as_uint(__uint_as_float(a) - __uint_as_float(b)));
    }
};
struct xf64 {
    using U = uint64_t;
    static constexpr U SIGN = 0x8000000000000000ull, QUIET = 0x0008000000000000ull;
    static constexpr U DEFNAN = 0xfff8000000000000ull, INF = 0x7ff0000000000000ull;
    static constexpr U ONE = 0x3ff0000000000000ull;
    __device__ static bool nan(U x) { return (x & ~SIGN) > INF; }
    __device__ static bool inf(U x) { return (x & ~SIGN) == INF; }
    __device__ static U fix(U r) { return nan(r) ? DEFNAN : r; }
    __device__ static double f(U x) { return __longlong_as_double((long long)x); }
    __device__ static U b(double v) { return (U)__double_as_longlong(v); }
    __device__ static U mul(U x, U y) {
        if (nan(x)) return x | QUIET;
        if (nan(y)) return y | QUIET;
        return fix(b(f(x) * f(y)));
    }
    __device__ static U add(U x, U y) {
        if (nan(x)) return x | QUIET;
        if (nan(y)) return y | QUIET;
        return fix(b(f(x) + f(y)));
    }
    __device__ static U sub(U x, U y) {
        if (nan(x)) return x | QUIET;
        if (nan(y)) return y | QUIET;
        return fix(b(f(x) - f(y)));
    }
};

// copysign(isinf(x) ? 1 : 0, x)
template <typename F>
__device__ __forceinline__ typename F::U box(typename F::U x) {
    return (x & F::SIGN) | (F::inf(x) ? F::ONE : (typename F::U)0);
}

// libgcc __mul?c3(a, b, c, d) = (a + bi)(c + di), C11 Annex G.5.1: products
// a*c, b*d, a*d, c*b; when both parts are NaN, infinities are boxed, NaNs
// of the other factor (or, after an overflow, all NaNs) become signed zeros,
// and the parts are recomputed times infinity.
template <typename F>
__device__ __attribute__((noinline)) void mulxc3(typename F::U a, typename F::U b, typename F::U c,
                                                 typename F::U d, typename F::U &re, typename F::U &im) {
    using U = typename F::U;
    const U ac = F::mul(a, c), bd = F::mul(b, d), ad = F::mul(a, d), bc = F::mul(c, b);
    U x = F::sub(ac, bd), y = F::add(ad, bc);
    if (F::nan(x) && F::nan(y)) {
        bool recalc = false;
        if (F::inf(a) || F::inf(b)) {
            a = box<F>(a);
            b = box<F>(b);
            if (F::nan(c)) c &= F::SIGN;
            if (F::nan(d)) d &= F::SIGN;
            recalc = true;
        }
        if (F::inf(c) || F::inf(d)) {
            c = box<F>(c);
            d = box<F>(d);
            if (F::nan(a)) a &= F::SIGN;
            if (F::nan(b)) b &= F::SIGN;
            recalc = true;
        }
        if (!recalc && (F::inf(ac) || F::inf(bd) || F::inf(ad) || F::inf(bc))) {
            if (F::nan(a)) a &= F::SIGN;
            if (F::nan(b)) b &= F::SIGN;
            if (F::nan(c)) c &= F::SIGN;
            if (F::nan(d)) d &= F::SIGN;
            recalc = true;
        }
        if (recalc) {
            x = F::mul(F::INF, F::sub(F::mul(a, c), F::mul(b, d)));
            y = F::mul(F::INF, F::add(F::mul(a, d), F::mul(b, c)));
        }
    }
    re = x;
    im = y;
}

// l*r as GCC emits std::complex operator*: inline, then __mul?c3 with the
// call site's argument order (swap: r first) when both parts are NaN.
template <typename F>
__device__ __forceinline__ void cmul(typename F::U lr, typename F::U li, typename F::U rr, typename F::U ri,
                                     bool swap, typename F::U &x, typename F::U &y) {
    x = F::sub(F::mul(lr, rr), F::mul(li, ri));
    y = F::add(F::mul(lr, ri), F::mul(li, rr));
    if (F::nan(x) && F::nan(y)) {
        if (swap)
            mulxc3<F>(rr, ri, lr, li, x, y);
        else
            mulxc3<F>(lr, li, rr, ri, x, y);
    }
}

// beta*d + alpha*s of one element in operand order `ord`.
__device__ __forceinline__ uint32_t scale4(uint32_t s, uint32_t d, const request &rq, uint32_t ord) {
    const uint32_t a = (uint32_t)rq.alpha[0], b = (uint32_t)rq.beta[0];
    const uint32_t bd = (ord & ORD_M1S) ? xf32::mul(b, d) : xf32::mul(d, b);
    const uint32_t as = (ord & ORD_M2S) ? xf32::mul(a, s) : xf32::mul(s, a);
    return (ord & ORD_ARE) ? xf32::add(as, bd) : xf32::add(bd, as);
}

__device__ __forceinline__ u32x2 scale8(u32x2 s, u32x2 d, const request &rq, uint32_t ord) {
    if (rq.dtype == DT_F64) {
        const uint64_t a = rq.alpha[0], b = rq.beta[0];
        const uint64_t dv = ((uint64_t)d.y << 32) | d.x, sv = ((uint64_t)s.y << 32) | s.x;
        const uint64_t bd = (ord & ORD_M1S) ? xf64::mul(b, dv) : xf64::mul(dv, b);
        const uint64_t as = (ord & ORD_M2S) ? xf64::mul(a, sv) : xf64::mul(sv, a);
        const uint64_t r = (ord & ORD_ARE) ? xf64::add(as, bd) : xf64::add(bd, as);
        return u32x2{(uint32_t)r, (uint32_t)(r >> 32)};
    }
    // complex<float>
    const uint32_t ar = (uint32_t)rq.alpha[0], ai = (uint32_t)(rq.alpha[0] >> 32);
    const uint32_t br = (uint32_t)rq.beta[0], bi = (uint32_t)(rq.beta[0] >> 32);
    uint32_t bdr, bdi, asr, asi;
    cmul<xf32>(br, bi, d.x, d.y, !(ord & ORD_M1S), bdr, bdi);
    cmul<xf32>(ar, ai, s.x, s.y, !(ord & ORD_M2S), asr, asi);
    return u32x2{(ord & ORD_ARE) ? xf32::add(asr, bdr) : xf32::add(bdr, asr),
                 (ord & ORD_AIM) ? xf32::add(asi, bdi) : xf32::add(bdi, asi)};
}

__device__ __forceinline__ u32x4 scale16(u32x4 s, u32x4 d, const request &rq, uint32_t ord) {
    const uint64_t dr = ((uint64_t)d.y << 32) | d.x, di = ((uint64_t)d.w << 32) | d.z;
    const uint64_t sr = ((uint64_t)s.y << 32) | s.x, si = ((uint64_t)s.w << 32) | s.z;
    uint64_t bdr, bdi, asr, asi;
    cmul<xf64>(rq.beta[0], rq.beta[1], dr, di, !(ord & ORD_M1S), bdr, bdi);
    cmul<xf64>(rq.alpha[0], rq.alpha[1], sr, si, !(ord & ORD_M2S), asr, asi);
    const uint64_t re = (ord & ORD_ARE) ? xf64::add(asr, bdr) : xf64::add(bdr, asr);
    const uint64_t im = (ord & ORD_AIM) ? xf64::add(asi, bdi) : xf64::add(bdi, asi);
    return u32x4{(uint32_t)re, (uint32_t)(re >> 32), (uint32_t)im, (uint32_t)(im >> 32)};
}

__device__ __forceinline__ uint32_t scale_bits(uint32_t s, uint32_t d, const request &rq, uint32_t ord) {
    return scale4(s, d, rq, ord);
}
__device__ __forceinline__ u32x2 scale_bits(u32x2 s, u32x2 d, const request &rq, uint32_t ord) {
    return scale8(s, d, rq, ord);
}
__device__ __forceinline__ u32x4 scale_bits(u32x4 s, u32x4 d, const request &rq, uint32_t ord) {
    return scale16(s, d, rq, ord);
}

// Operand order of the reference's scaled loops for final element (r, c) of
// a piece of rows x cols (what its compiled loops do, pinned by the oracle
// against the reference's own build):
//   PIECE_SITE_COL   copy_and_scale over one column (copy2D_and_scale's
//     strided branch, memory_utils.hpp:113-120): real types put the element
//     first in both products, beta*dst first in the sum in the vector body
//     (r < rows - rows % (16/sizeof(T))) and for one-row columns, alpha*src
//     first in the remainder;
//   PIECE_SITE_FLAT  copy_and_scale over the whole contiguous piece
//     (:108-110): real types take element pairs with the element first in
//     both products and the last element of an odd count with the scalar
//     first; alpha*src first in the sum;
//   PIECE_SITE_TRANS copy_transpose_and_scale of a local transposed piece
//     (:186-256, tiles of 512/sizeof(T)): real types element first, alpha*src
//     first in the sum; complex types pass (alpha, el) to __mul?c3 in the
//     diagonal tiles and (el, alpha) in the others (complex<float> also puts
//     alpha*src first in the real part of the sum there).
// Complex types otherwise pass (beta, dst) and (alpha, src) to __mul?c3 and
// put beta*dst first in both parts of the sum.
__device__ __forceinline__ uint32_t ref_order(uint32_t dtype, uint32_t site, uint32_t r, uint32_t c,
                                              uint32_t rows, uint32_t cols) {
    const bool real = dtype == DT_F32 || dtype == DT_F64;
    if (site == PIECE_SITE_COL) {
        if (!real) return ORD_M1S | ORD_M2S;
        const uint32_t vf = dtype == DT_F32 ? 4u : 2u;
        return (rows == 1 || r < rows - rows % vf) ? 0u : ORD_ARE;
    }
    if (site == PIECE_SITE_FLAT) {
        if (!real) return ORD_M1S | ORD_M2S;
        const uint64_t n = (uint64_t)rows * cols, i = (uint64_t)c * rows + r;
        return ((n & 1) && i == n - 1) ? (ORD_M1S | ORD_M2S | ORD_ARE) : ORD_ARE;
    }
    if (real) return ORD_ARE;
    const uint32_t tile = dtype == DT_C64 ? 64u : 32u;  // 512 / sizeof(T)
    if (c / tile == r / tile) return ORD_M1S | ORD_M2S;
    return dtype == DT_C64 ? (ORD_M1S | ORD_ARE) : ORD_M1S;
}

// std::conj: sign bit of the imaginary part (the host sets REQ_CONJ only for
// complex types; conjugate() of real types is the identity).
__device__ __forceinline__ uint32_t conj_bits(uint32_t v) { return v; }
__device__ __forceinline__ u32x2 conj_bits(u32x2 v) {
    v.y ^= 0x80000000u;
    return v;
}
__device__ __forceinline__ u32x4 conj_bits(u32x4 v) {
    v.w ^= 0x80000000u;
    return v;
}

// ---------------------------------------------------------- copy paths

// Global loads/stores, optionally with the non-temporal hint (streaming data
// that is touched exactly once).
template <bool NT, typename T>
__device__ __forceinline__ T ld(const GLOBAL T *p) {
    if constexpr (NT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}
template <bool NT, typename T>
__device__ __forceinline__ void st(GLOBAL T *p, T v) {
    if constexpr (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

// Unsigned 32-bit division by a divisor fixed for a whole piece, as a
// multiply-high and shifts (Granlund & Montgomery, "Division by invariant
// integers using multiplication", 1994, fig. 4.1).
struct fastdiv {
    uint32_t mul, s1, s2;
    __device__ __forceinline__ explicit fastdiv(uint32_t d) {
        const uint32_t l = ceil_log2(d);
        mul = (uint32_t)((((uint64_t)1 << 32) * (((uint64_t)1 << l) - d)) / d + 1);
        s1 = l > 0 ? 1u : 0u;
        s2 = l > 0 ? l - 1 : 0u;
    }
    __device__ __forceinline__ uint32_t div(uint32_t n) const {
        const uint32_t t = __umulhi(mul, n);
        return (t + ((n - t) >> s1)) >> s2;
    }
};

// 16-byte accesses at dword-aligned addresses.
template <bool NT>
__device__ __forceinline__ u32x4 ld_a4(const gbyte *p) {
    const GLOBAL u32x4a4 *q = (const GLOBAL u32x4a4 *)p;
    if constexpr (NT)
        return __builtin_nontemporal_load(q);
    else
        return *q;
}
template <bool NT>
__device__ __forceinline__ void st_a4(gbyte *p, u32x4 v) {
    GLOBAL u32x4a4 *q = (GLOBAL u32x4a4 *)p;
    const u32x4a4 w = v;
    if constexpr (NT)
        __builtin_nontemporal_store(w, q);
    else
        *q = w;
}

// Element e of a 16-byte vector holding 16/ES elements of ES bytes.
template <int ES>
__device__ __forceinline__ typename bits_t<ES>::type vec_elem(const u32x4 &v, int e) {
    if constexpr (ES == 4)
        return e == 0 ? v.x : e == 1 ? v.y : e == 2 ? v.z : v.w;
    else if constexpr (ES == 8)
        return e == 0 ? u32x2{v.x, v.y} : u32x2{v.z, v.w};
    else
        return v;
}
template <int ES>
__device__ __forceinline__ void vec_set(u32x4 &v, int e, typename bits_t<ES>::type x) {
    if constexpr (ES == 4) {
        if (e == 0) v.x = x;
        else if (e == 1) v.y = x;
        else if (e == 2) v.z = x;
        else v.w = x;
    } else if constexpr (ES == 8) {
        if (e == 0) {
            v.x = x.x;
            v.y = x.y;
        } else {
            v.z = x.x;
            v.w = x.y;
        }
    } else {
        v = x;
    }
}

// Copy every slot of the queued SEG_VEC segments. Slots are dealt to the
// threads in order, so consecutive lanes move consecutive 16-byte vectors of a
// column (a wave spans several short columns, or one stretch of a long one,
// and several pieces), and every thread keeps UB bytes in flight whatever the
// piece shapes. Addresses need only be dword-aligned. Every slot of a column
// of 16 bytes or more is one 16-byte load and one 16-byte store: the last slot
// of a ragged column rewrites up to 12 bytes of the slot before it with the
// same values (source and destination do not overlap), so no wave takes a
// dword-by-dword path for the columns' tails.
template <int THREADS, int UB, bool NT>
__device__ __forceinline__ void flat_copy(const seg_entry *segs, uint32_t total) {
    constexpr int U = UB / 16;
    // the segment of this thread's current slot, kept in registers (a
    // thread's slots only increase, so it moves forward and rarely)
    uint32_t j = 0;
    seg_entry e = segs[0];
    uint32_t next0 = segs[1].slot0;
    for (uint32_t base = threadIdx.x; base < total; base += (uint32_t)THREADS * U) {
        u32x4 r[U];
        gbyte *dp[U];
        uint32_t tail[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t k = base + (uint32_t)u * THREADS;
            tail[u] = 4;
            if (k < total) {
                if (next0 <= k) {  // walk the slot starts, then load the one entry needed
                    do {
                        ++j;
                        next0 = segs[j + 1].slot0;
                    } while (next0 <= k);
                    e = segs[j];
                }
                uint64_t so, dof;
                vec_slot(e, k - e.slot0, so, dof, tail[u]);
                const gbyte *sp = (const gbyte *)e.s + so;
                dp[u] = (gbyte *)e.d + dof;
                if (tail[u] == 4) {
                    r[u] = ld_a4<NT>(sp);
                } else {
                    const GLOBAL uint32_t *q = (const GLOBAL uint32_t *)sp;
                    r[u].x = q[0];
                    r[u].y = tail[u] > 1 ? q[1] : 0u;
                    r[u].z = tail[u] > 2 ? q[2] : 0u;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (base + (uint32_t)u * THREADS >= total) continue;
            if (tail[u] == 4) {
                st_a4<NT>(dp[u], r[u]);
            } else {
                GLOBAL uint32_t *q = (GLOBAL uint32_t *)dp[u];
                q[0] = r[u].x;
                if (tail[u] > 1) q[1] = r[u].y;
                if (tail[u] > 2) q[2] = r[u].z;
            }
        }
    }
}

// Copy every slot of the queued SEG_TRANS / SEG_CONJ segments of 4- or
// 8-byte elements as VE x VE register tiles, VE = 16/ES. A slot is a group of
// VE final rows [h*VE, +VE) by a group of VE final columns [g*VE, +VE)
// (seg_entry.vpc = the piece's row groups): VE 16-byte reads, one per final
// row (VE consecutive elements of source storage column i), a transpose in
// registers, and VE 16-byte writes, one per final column (VE consecutive
// rows). Consecutive lanes take consecutive row groups of one column group, so
// each write instruction stores one contiguous run per destination column.
// A partial row group (rows not a multiple of VE) or column group (the
// piece's last, seg_entry.ncols) moves element by element.
template <int THREADS, bool NT, int ES>
__device__ __forceinline__ void flat_copy_tiles(const seg_entry *segs, uint32_t total) {
    static_assert(ES == 4 || ES == 8, "register tiles of 4- and 8-byte elements");
    using T = typename bits_t<ES>::type;
    constexpr uint32_t VE = 16 / ES;
    uint32_t j = 0;
    seg_entry e = segs[0];
    uint32_t next0 = segs[1].slot0;
    for (uint32_t k = threadIdx.x; k < total; k += (uint32_t)THREADS) {
        if (next0 <= k) {  // walk the slot starts, then load the one entry needed
            do {
                ++j;
                next0 = segs[j + 1].slot0;
            } while (next0 <= k);
            e = segs[j];
        }
        uint32_t c, i0, nc, nr;  // nc columns from c, nr rows from i0
        tile_slot<ES>(e, k - e.slot0, c, i0, nc, nr);
        const bool conj = ES > 4 && e.kind == SEG_CONJ;
        const gbyte *sp = (const gbyte *)e.s + (uint64_t)i0 * e.lds_b + (uint64_t)c * ES;
        u32x4 in[VE];
#pragma unroll
        for (uint32_t y = 0; y < VE; ++y) {
            in[y] = u32x4{0u, 0u, 0u, 0u};
            if (y < nr) {
                const gbyte *q = sp + (uint64_t)y * e.lds_b;
                if (nc == VE) {
                    in[y] = ld_a4<NT>(q);
                } else {
#pragma unroll
                    for (uint32_t x = 0; x < VE; ++x)
                        if (x < nc) vec_set<ES>(in[y], x, *(const GLOBAL T *)(q + x * ES));
                }
            }
        }
        gbyte *dp = (gbyte *)e.d + (uint64_t)c * e.ldd_b + (uint64_t)i0 * ES;
#pragma unroll
        for (uint32_t x = 0; x < VE; ++x) {
            if (x >= nc) continue;
            u32x4 out;
#pragma unroll
            for (uint32_t y = 0; y < VE; ++y) {
                T v = vec_elem<ES>(in[y], x);
                if (conj) v = conj_bits(v);
                vec_set<ES>(out, y, v);
            }
            gbyte *q = dp + (uint64_t)x * e.ldd_b;
            if (nr == VE) {
                st_a4<NT>(q, out);
            } else {
#pragma unroll
                for (uint32_t y = 0; y < VE; ++y)
                    if (y < nr) st<NT>((GLOBAL T *)(q + y * ES), vec_elem<ES>(out, y));
            }
        }
    }
}

// Copy every slot of the queued SEG_TRANS / SEG_CONJ segments of ES-byte
// elements. A slot is one final row i of a group of VE = 16/ES final columns:
// one 16-byte read of VE consecutive elements of source storage column i
// (dword-aligned), written as one element into each of VE destination
// columns. Consecutive lanes take consecutive rows, so every one of those VE
// stores is contiguous across the lanes. The last group of a piece may hold
// fewer than VE columns (seg_entry.ncols = the piece's column count).
template <int THREADS, bool NT, int ES>
__device__ __forceinline__ void flat_copy_elems(const seg_entry *segs, uint32_t total) {
    using T = typename bits_t<ES>::type;
    constexpr uint32_t VE = 16 / ES;
    constexpr int U = 4;
    uint32_t j = 0;
    seg_entry e = segs[0];
    uint32_t next0 = segs[1].slot0;
    for (uint32_t base = threadIdx.x; base < total; base += (uint32_t)THREADS * U) {
        u32x4 r[U];
        gbyte *dp[U];
        uint32_t cnt[U];
        uint64_t ldd[U];
        bool conj[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t k = base + (uint32_t)u * THREADS;
            cnt[u] = 0;
            if (k < total) {
                if (next0 <= k) {  // walk the slot starts, then load the one entry needed
                    do {
                        ++j;
                        next0 = segs[j + 1].slot0;
                    } while (next0 <= k);
                    e = segs[j];
                }
                uint32_t c, v;  // final row v of cnt final columns from c
                elem_slot<ES>(e, k - e.slot0, c, v, cnt[u]);
                dp[u] = (gbyte *)e.d + c * e.ldd_b + (uint64_t)v * ES;
                ldd[u] = e.ldd_b;
                conj[u] = ES > 4 && e.kind == SEG_CONJ;
                const gbyte *sp = (const gbyte *)e.s + (uint64_t)v * e.lds_b + (uint64_t)c * ES;
                if (cnt[u] == VE) {
                    r[u] = ld_a4<NT>(sp);
                } else {  // ragged last group: element by element
#pragma unroll
                    for (uint32_t x = 0; x < VE; ++x)
                        if (x < cnt[u]) vec_set<ES>(r[u], x, *(const GLOBAL T *)(sp + x * ES));
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (uint32_t x = 0; x < VE; ++x) {
                if (x < cnt[u]) {
                    T val = vec_elem<ES>(r[u], x);
                    if (conj[u]) val = conj_bits(val);
                    st<NT>((GLOBAL T *)(dp[u] + x * ldd[u]), val);
                }
            }
        }
    }
}

// A scaled piece (transposed or not), one element per lane slot, slots in
// column-major order of the final columns [c0, c1): final element (r, c) =
// beta*dst + alpha*op(source element), op = transpose (source (r, c) at
// src + (r*lds + c)*ES) and conjugation as the request says, in the operand
// order of the reference call site that scales the piece (ref_order, from
// the site in piece.req and the element's position in the whole piece).
template <int THREADS, int ES>
__device__ __forceinline__ void copy_scaled(const piece &pc, const request &rq, uint32_t c0, uint32_t c1) {
    using T = typename bits_t<ES>::type;
    constexpr int U = 4;
    const uint32_t rows = pc.rows;
    const uint32_t slots = (c1 - c0) * rows;
    const bool trans = (rq.flags & REQ_TRANS) != 0, conj = (rq.flags & REQ_CONJ) != 0;
    const uint32_t site = (pc.req & PIECE_SITE_MASK) >> PIECE_SITE_SHIFT;
    const uint64_t lds_b = (uint64_t)pc.lds * ES, ldd_b = (uint64_t)pc.ldd * ES;
    const gbyte *s = (const gbyte *)pc.src;
    gbyte *d = (gbyte *)pc.dst;
    const fastdiv per_col(rows);
    for (uint32_t base = threadIdx.x; base < slots; base += (uint32_t)THREADS * U) {
        T a[U], b[U];
        uint32_t rr[U], cc[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t k = base + (uint32_t)u * THREADS;
            const uint32_t c = per_col.div(k);
            rr[u] = k - c * rows;
            cc[u] = c0 + c;
            if (k < slots) {
                const uint64_t so = trans ? (uint64_t)rr[u] * lds_b + (uint64_t)cc[u] * ES
                                          : (uint64_t)cc[u] * lds_b + (uint64_t)rr[u] * ES;
                a[u] = *(const GLOBAL T *)(s + so);
                b[u] = *(const GLOBAL T *)(d + (uint64_t)cc[u] * ldd_b + (uint64_t)rr[u] * ES);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (base + (uint32_t)u * THREADS >= slots) continue;
            const T x = conj ? conj_bits(a[u]) : a[u];
            const uint32_t ord = ref_order(rq.dtype, site, rr[u], cc[u], rows, pc.cols);
            *(GLOBAL T *)(d + (uint64_t)cc[u] * ldd_b + (uint64_t)rr[u] * ES) = scale_bits(x, b[u], rq, ord);
        }
    }
}

// Transposing copy: destination(i, j) = op(source(j, i)) for final rows
// i < rows and final columns j < ncols. Source element (i, j) lies at
// s + (i*lds + j)*ES: storage rows run along final columns. A tile is TC final
// rows (TC = 64, or 32 / 16 for pieces with fewer rows) by TR = 16384/(TC*ES)
// final columns. Phase 1 reads TR*ES contiguous bytes (256-1024) of each of
// the TC source storage columns; phase 2 writes TC*ES contiguous bytes of
// each of the TR destination columns. Both phases move 16-byte vectors
// (dword-aligned addresses); LDS holds the tile in storage order, rows padded
// by one element against bank conflicts. Vectors that would run past the
// piece move element by element.
template <int THREADS, bool NT, int ES, uint32_t TC>
__device__ __forceinline__ void copy_transposed(const gbyte *__restrict__ s,
                                                gbyte *__restrict__ d, uint64_t lds_b,
                                                uint64_t ldd_b, uint32_t rows, uint32_t ncols,
                                                uint32_t flags, LOCAL unsigned char *smem) {
    using T = typename bits_t<ES>::type;
    using TG = tile_geom<ES, TC, THREADS>;
    constexpr uint32_t VE = TG::VE, TR = TG::TR, LDW = TG::LDW, PER = TG::PER;
    LOCAL T *tile = (LOCAL T *)smem;               // tile[c * LDW + r]
    const bool conj = (flags & REQ_CONJ) != 0;
    const uint32_t n_ic = (rows + TC - 1) / TC;
    const uint32_t n_tiles = n_ic * ((ncols + TR - 1) / TR);
    // Phase 1 of tile t+1 (global loads into registers) is issued before
    // phase 2 of tile t (LDS to global), so loads and stores overlap.
    u32x4 v1[PER];
    uint32_t cnt1[PER];
    auto load_tile = [&](uint32_t t) {
        const uint32_t jc = (t / n_ic) * TR, ic = (t % n_ic) * TC;
#pragma unroll
        for (uint32_t k = 0; k < PER; ++k) {
            uint32_t c, r0;
            TG::load_index(threadIdx.x + k * THREADS, c, r0);
            const uint32_t i = ic + c, j0 = jc + r0;
            cnt1[k] = (i < rows && j0 < ncols) ? min(VE, ncols - j0) : 0u;
            const gbyte *p = s + i * lds_b + (uint64_t)j0 * ES;
            if (cnt1[k] == VE) {
                v1[k] = ld_a4<NT>(p);
            } else {
#pragma unroll
                for (uint32_t e = 0; e < VE; ++e)
                    if (e < cnt1[k]) vec_set<ES>(v1[k], (int)e, *(const GLOBAL T *)(p + e * ES));
            }
        }
    };
    if (n_tiles) load_tile(0);
    for (uint32_t t = 0; t < n_tiles; ++t) {
        const uint32_t jc = (t / n_ic) * TR, ic = (t % n_ic) * TC;
#pragma unroll
        for (uint32_t k = 0; k < PER; ++k) {
            uint32_t c, r0;
            TG::load_index(threadIdx.x + k * THREADS, c, r0);
            LOCAL T *tp = tile + c * LDW + r0;
#pragma unroll
            for (uint32_t e = 0; e < VE; ++e)
                if (e < cnt1[k]) tp[e] = vec_elem<ES>(v1[k], (int)e);
        }
        __syncthreads();
        if (t + 1 < n_tiles) load_tile(t + 1);
#pragma unroll
        for (uint32_t k = 0; k < PER; ++k) {
            uint32_t r, w;
            TG::store_index(threadIdx.x + k * THREADS, r, w);
            const uint32_t j = jc + r, i0 = ic + w * VE;
            if (j < ncols && i0 < rows) {
                gbyte *p = d + j * ldd_b + (uint64_t)i0 * ES;
                const LOCAL T *tp = tile + (w * VE) * LDW + r;
                if (i0 + VE <= rows) {
                    u32x4 v;
#pragma unroll
                    for (uint32_t e = 0; e < VE; ++e) {
                        T x = tp[e * LDW];
                        if (conj) x = conj_bits(x);
                        vec_set<ES>(v, e, x);
                    }
                    st_a4<NT>(p, v);
                } else {
                    for (uint32_t e = 0; e < rows - i0; ++e) {
                        T x = tp[e * LDW];
                        if (conj) x = conj_bits(x);
                        *(GLOBAL T *)(p + e * ES) = x;
                    }
                }
            }
        }
        __syncthreads();
    }
}

// A value the caller passes identically to every lane, kept in scalar
// registers (arguments of a called function arrive in vector registers).
// readfirstlane returns int: each half goes back through uint32_t, so the low
// half is zero-extended, never sign-extended into the high one.
__device__ __forceinline__ uint32_t uniform32(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)x);
}
__device__ __forceinline__ uint64_t uniform64(uint64_t x) {
    return ((uint64_t)uniform32((uint32_t)(x >> 32)) << 32) | (uint64_t)uniform32((uint32_t)x);
}

// The same transposing copy over whole tiles only: n_ic tiles of TC final
// rows down, n_tiles in all (n_tiles / n_ic column strips of TR final
// columns), every element of every tile inside the piece, so no vector needs
// a bounds check. A pass moves G = trans_tiles<THREADS>() consecutive tiles
// (row tiles first), each through its own LDS area; the global loads of the
// next pass are issued before the stores of the current one. Addresses are a
// tile base in scalar registers plus a 32-bit offset per thread and vector,
// fixed for the whole call: the caller guarantees TC*lds_b and TR*ldd_b
// below 2^32.
template <int THREADS, bool NT, int ES, uint32_t TC>
__device__ __forceinline__ void copy_transposed_full(const gbyte *__restrict__ s_in,
                                                     gbyte *__restrict__ d_in, uint64_t lds_in,
                                                     uint64_t ldd_in, uint32_t n_ic_in, uint32_t n_tiles_in,
                                                     uint32_t flags, LOCAL unsigned char *smem) {
    using T = typename bits_t<ES>::type;
    using TG = tile_geom<ES, TC, THREADS>;
    constexpr uint32_t G = trans_tiles<THREADS>();
    constexpr uint32_t VE = TG::VE, TR = TG::TR, LDW = TG::LDW, PER = TG::PER;
    constexpr uint32_t TILE_ELEMS = TG::TILE_ELEMS;  // LDS area of one tile, in elements
    const gbyte *s = (const gbyte *)uniform64((uint64_t)s_in);
    gbyte *d = (gbyte *)uniform64((uint64_t)d_in);
    const uint64_t lds_b = uniform64(lds_in), ldd_b = uniform64(ldd_in);
    const uint32_t n_ic = uniform32(n_ic_in), n_tiles = uniform32(n_tiles_in);
    const bool conj = uniform32(flags & REQ_CONJ) != 0;
    LOCAL T *tile = (LOCAL T *)smem;               // area g: tile[g*TILE_ELEMS + c*LDW + r]
    // per thread and vector k: source offset inside a tile (storage column,
    // element), destination offset (final column, row), LDS positions
    uint32_t so[PER], dof[PER], lw[PER], lr[PER];
#pragma unroll
    for (uint32_t k = 0; k < PER; ++k) {
        uint32_t c, r0, r, w;
        TG::load_index(threadIdx.x + k * THREADS, c, r0);
        so[k] = c * (uint32_t)lds_b + r0 * ES;
        lw[k] = c * LDW + r0;
        TG::store_index(threadIdx.x + k * THREADS, r, w);
        dof[k] = r * (uint32_t)ldd_b + w * VE * ES;
        lr[k] = (w * VE) * LDW + r;
    }
    const uint32_t n_pass = (n_tiles + G - 1) / G;
    u32x4 v[G][PER];
    auto load_pass = [&](uint32_t pass) {
#pragma unroll
        for (uint32_t g = 0; g < G; ++g) {
            const uint32_t t = pass * G + g;
            if (t < n_tiles) {
                const uint32_t jc = (t / n_ic) * TR, ic = (t % n_ic) * TC;
                const gbyte *base = s + ic * lds_b + (uint64_t)jc * ES;
#pragma unroll
                for (uint32_t k = 0; k < PER; ++k) v[g][k] = ld_a4<NT>(base + so[k]);
            }
        }
    };
    if (n_pass) load_pass(0);
    for (uint32_t pass = 0; pass < n_pass; ++pass) {
#pragma unroll
        for (uint32_t g = 0; g < G; ++g) {
            if (pass * G + g < n_tiles) {
#pragma unroll
                for (uint32_t k = 0; k < PER; ++k) {
                    LOCAL T *tp = tile + g * TILE_ELEMS + lw[k];
#pragma unroll
                    for (uint32_t e = 0; e < VE; ++e) tp[e] = vec_elem<ES>(v[g][k], (int)e);
                }
            }
        }
        __syncthreads();
        if (pass + 1 < n_pass) load_pass(pass + 1);
        // one tile at a time: its LDS and global addresses are computed in
        // the loop rather than held for every tile of the pass
#pragma unroll 1
        for (uint32_t g = 0; g < G; ++g) {
            const uint32_t t = pass * G + g;
            if (t < n_tiles) {
                const uint32_t jc = (t / n_ic) * TR, ic = (t % n_ic) * TC;
                gbyte *base = d + jc * ldd_b + (uint64_t)ic * ES;
                const LOCAL T *tg = tile + g * TILE_ELEMS;
#pragma unroll
                for (uint32_t k = 0; k < PER; ++k) {
                    gbyte *p = base + dof[k];
                    const LOCAL T *tp = tg + lr[k];
                    u32x4 o;
#pragma unroll
                    for (uint32_t e = 0; e < VE; ++e) {
                        T x = tp[e * LDW];
                        if (conj) x = conj_bits(x);
                        vec_set<ES>(o, e, x);
                    }
                    st_a4<NT>(p, o);
                }
            }
        }
        __syncthreads();
    }
}

// One part of a transposed piece: FULL = its whole tiles (copy_transposed_full,
// with n_ic = x, n_tiles = y), else a strip through the bounds-checked path
// (rows = x, ncols = y). Not inlined: a called function keeps the registers of
// the transpose out of the queue loops of the gather, and the caller keeps
// its own live values in registers the callee leaves alone, so every callee
// is kept small: one per part, element size and tile height.
template <int THREADS, bool NT, int ES, uint32_t TC, bool FULL>
__device__ __attribute__((noinline)) void transpose_part(const gbyte *__restrict__ s, gbyte *__restrict__ d,
                                                         uint64_t lds_b, uint64_t ldd_b, uint32_t x, uint32_t y,
                                                         uint32_t flags, LOCAL unsigned char *smem) {
    if constexpr (FULL)
        copy_transposed_full<THREADS, NT, ES, TC>(s, d, lds_b, ldd_b, x, y, flags, smem);
    else
        copy_transposed<THREADS, NT, ES, TC>(s, d, lds_b, ldd_b, x, y, flags, smem);
}

// A transposed piece with tiles of TC final rows: whole tiles through the
// full-tile path, then the bottom strip (the last rows % TC final rows, every
// column) and the right strip (the last ncols % TR final columns of the other
// rows) through the bounds-checked one.
template <int THREADS, bool NT, int ES, uint32_t TC>
__device__ __forceinline__ void transpose_tc(const gbyte *s, gbyte *d, uint64_t lds_b, uint64_t ldd_b,
                                             uint32_t rows, uint32_t ncols, uint32_t flags,
                                             unsigned char *smem_g) {
    constexpr uint32_t TR = kTileBytes / (TC * ES);
    LOCAL unsigned char *smem = (LOCAL unsigned char *)smem_g;
    // whole tiles: unless a tile spans 4 GiB or more of either side (32-bit
    // offsets inside a tile)
    const bool small_ld = (uint64_t)TC * lds_b < (1ull << 32) && (uint64_t)TR * ldd_b < (1ull << 32);
    const uint32_t rf = small_ld ? rows / TC * TC : 0u, cf = ncols / TR * TR;
    if (rf && cf)
        transpose_part<THREADS, NT, ES, TC, true>(s, d, lds_b, ldd_b, rf / TC, (rf / TC) * (cf / TR), flags, smem);
    if (rf < rows)
        transpose_part<THREADS, NT, ES, TC, false>(s + rf * lds_b, d + (uint64_t)rf * ES, lds_b, ldd_b, rows - rf,
                                                  ncols, flags, smem);
    if (rf && cf < ncols)
        transpose_part<THREADS, NT, ES, TC, false>(s + (uint64_t)cf * ES, d + cf * ldd_b, lds_b, ldd_b, rf,
                                                  ncols - cf, flags, smem);
}

// A transposed piece of `rows` final rows: tiles of 64 final rows, or 32 / 16
// for shorter pieces.
template <int THREADS, bool NT, int ES>
__device__ __forceinline__ void transpose_call(const gbyte *s, gbyte *d, uint64_t lds_b, uint64_t ldd_b,
                                               uint32_t rows, uint32_t ncols, uint32_t flags,
                                               unsigned char *smem) {
    if (rows > 32)
        transpose_tc<THREADS, NT, ES, 64>(s, d, lds_b, ldd_b, rows, ncols, flags, smem);
    else if (rows > 16)
        transpose_tc<THREADS, NT, ES, 32>(s, d, lds_b, ldd_b, rows, ncols, flags, smem);
    else
        transpose_tc<THREADS, NT, ES, 16>(s, d, lds_b, ldd_b, rows, ncols, flags, smem);
}

// Large transposed pieces and scaled pieces: moved by the whole workgroup
// right away.
template <int THREADS, bool NT, int FEAT>
__device__ __forceinline__ void process_piece(const piece &pc, const request &rq, uint32_t c0,
                                              uint32_t c1, unsigned char *smem) {
    const uint32_t es = rq.esize;
    if (rq.flags & REQ_SCALE) {
        if constexpr (FEAT == FEAT_ALL) {  // the host sends scaling to this variant only
            if (es == 4)
                copy_scaled<THREADS, 4>(pc, rq, c0, c1);
            else if (es == 8)
                copy_scaled<THREADS, 8>(pc, rq, c0, c1);
            else
                copy_scaled<THREADS, 16>(pc, rq, c0, c1);
        }
        return;
    }
    const uint32_t ncols = c1 - c0;
    const uint64_t lds_b = (uint64_t)pc.lds * es;
    const uint64_t ldd_b = (uint64_t)pc.ldd * es;
    gbyte *d = (gbyte *)pc.dst + (uint64_t)c0 * ldd_b;
    const gbyte *s = (const gbyte *)pc.src + (uint64_t)c0 * es;
    if (es == 4)
        transpose_call<THREADS, NT, 4>(s, d, lds_b, ldd_b, pc.rows, ncols, rq.flags, smem);
    else if (es == 8)
        transpose_call<THREADS, NT, 8>(s, d, lds_b, ldd_b, pc.rows, ncols, rq.flags, smem);
    else
        transpose_call<THREADS, NT, 16>(s, d, lds_b, ldd_b, pc.rows, ncols, rq.flags, smem);
}

__device__ __forceinline__ uint32_t wave_scan_u32(uint32_t x, uint32_t lane) {
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    return x;
}

// Work item t covers bytes [lo, hi) of the piece stream (pieces concatenated
// in table order, each column-major; item_range: tile_bytes-long items, the
// last ones of a split stream shorter). It moves
// every column (every group of 256/sizeof(T) columns for transposed pieces)
// whose first byte lies in that range, so work items partition the columns
// exactly. Pieces are examined 64 at a time, one per lane (every wave does the
// same arithmetic; wave 0 writes the queues): untransposed, unscaled column
// runs are queued for flat_copy and small transposed pieces for
// flat_copy_elems; once the queues are drained, the pieces that are moved on
// their own (large transposed pieces, scaled pieces) follow one by one, each
// by the whole workgroup. Pieces write disjoint destinations, so their order
// does not matter.
template <int THREADS, int UB, bool NT, int FEAT>
__device__ __forceinline__ void process_tile(const work_rec &job, const kreqs &kr, unsigned char *smem,
                                             seg_entry *segs, seg_entry *tsegs) {
    const uint64_t lo = job.lo, hi = job.hi;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t es = kr.r[0].esize;  // both requests of a desc share the element type
    // pieces past the work item's last possible one are not even loaded
    const uint32_t n = job.last_piece + 1;
    const piece *pieces = job.pieces;
    const uint64_t *pend = job.pend;
    // p: the work item's first piece; a lane-0 piece past the work item ends
    // the loop below without a separate check
    uint32_t p = job.first_piece;
    for (;;) {
        if (p >= n) break;
        const uint32_t q = p + lane;
        uint32_t kind = PK_NONE, c0 = 0, c1 = 0, rows = 0, flags = 0, vpc = 0;
        piece pc;
        bool in = q < n;
        if (in) {
            pc = pieces[q];
            const uint64_t ps = q ? pend[q - 1] : 0;
            const uint64_t pe = pend[q];
            in = ps < hi;
            flags = (pc.req & PIECE_REQ1) ? kr.r[1].flags : kr.r[0].flags;
            rows = pc.rows;
            piece_cols(rows, pc.cols, flags, es, ps, pe, lo, hi, c0, c1);
            if (in) {
                kind = piece_kind(FEAT, flags, rows, c0, c1, es);
                vpc = piece_vpc(kind, rows, es);
            }
        }
        // pieces past the work item come after every piece inside it
        const uint64_t out_mask = __ballot(!in);
        const uint32_t end = out_mask ? (uint32_t)__ffsll((unsigned long long)out_mask) - 1 : 64u;
        const bool is_v = kind == PK_VEC, is_e = kind == PK_ELEM;
        const uint64_t mv = __ballot(is_v), me = __ballot(is_e);
        uint64_t md = __ballot(kind == PK_DIRECT);
        const uint32_t nseg = (uint32_t)__popcll(mv), ntseg = (uint32_t)__popcll(me);
        if (nseg | ntseg) {
            const uint32_t slots = (is_v || is_e) ? piece_groups(kind, c1 - c0, es) * vpc : 0u;
            const uint32_t sv = wave_scan_u32(is_v ? slots : 0u, lane);
            const uint32_t se = wave_scan_u32(is_e ? slots : 0u, lane);
            const uint32_t total = __shfl(sv, 63, 64), ttotal = __shfl(se, 63, 64);
            if (threadIdx.x < 64) {
                const uint64_t below = (1ull << lane) - 1;
                if (is_v || is_e) {
                    seg_entry &e = is_v ? segs[__popcll(mv & below)] : tsegs[__popcll(me & below)];
                    e = make_seg(pc, kind, flags, es, c0, c1 - c0, vpc, is_v ? sv - slots : se - slots);
                }
                if (lane == 0) {  // sentinels
                    segs[nseg].slot0 = total;
                    tsegs[ntseg].slot0 = ttotal;
                }
            }
            __syncthreads();
            if (nseg) flat_copy<THREADS, UB, NT>(segs, total);
            if constexpr (FEAT != FEAT_COPY) {
                if (ntseg) {
                    if (es == 4) flat_copy_tiles<THREADS, NT, 4>(tsegs, ttotal);
                    else if (es == 8) flat_copy_tiles<THREADS, NT, 8>(tsegs, ttotal);
                    else flat_copy_elems<THREADS, NT, 16>(tsegs, ttotal);
                }
            }
            __syncthreads();  // the queues are rewritten by the next examination
        }
        if constexpr (FEAT != FEAT_COPY) {
            while (md) {
                const uint32_t d = p + (uint32_t)__ffsll((unsigned long long)md) - 1;
                md &= md - 1;
                const piece dp = pieces[d];
                // field by field: selecting one of the two structs by address
                // would keep both in private memory
                const bool r1 = (dp.req & PIECE_REQ1) != 0;
                request rq;
                rq.dtype = r1 ? kr.r[1].dtype : kr.r[0].dtype;
                rq.flags = r1 ? kr.r[1].flags : kr.r[0].flags;
                rq.esize = es;
                rq.pad = 0;
                rq.alpha[0] = r1 ? kr.r[1].alpha[0] : kr.r[0].alpha[0];
                rq.alpha[1] = r1 ? kr.r[1].alpha[1] : kr.r[0].alpha[1];
                rq.beta[0] = r1 ? kr.r[1].beta[0] : kr.r[0].beta[0];
                rq.beta[1] = r1 ? kr.r[1].beta[1] : kr.r[0].beta[1];
                uint32_t d0, d1;
                piece_cols(dp.rows, dp.cols, rq.flags, es, d ? pend[d - 1] : 0, pend[d], lo, hi, d0, d1);
                process_piece<THREADS, NT, FEAT>(dp, rq, d0, d1, smem);
            }
        }
        if (end < 64) break;
        p += 64;
    }
}

// Global work index w -> (transform, local work item) -> columns to move.
template <int THREADS, int UB, bool NT, int FEAT>
__device__ __forceinline__ void process_work(const gather_args &a, uint32_t w, uint32_t *claim, uint32_t &next,
                                             unsigned char *smem, seg_entry *segs, seg_entry *tsegs) {
    work_rec job;
    if (!a.descs) {
        job.pieces = a.one.pieces;
        job.pend = a.one.pend;
        item_range(a.one.tile_bytes, a.one.split_from, a.one.split_bytes, a.one.total_bytes, w, job.lo, job.hi);
        job.first_piece = a.one.first_piece[w];
        job.n_pieces = a.one.n_pieces;
        job.desc = 0;
        job.last_piece = w + 1 < a.one.n_tiles ? a.one.first_piece[w + 1] : a.one.n_pieces - 1;
    } else {
        job = a.work[w];  // all the pieces' loads wait on
    }
    // Persistent launches claim the next work index here: after the loads
    // above, so that waiting for them does not wait for the claim too (memory
    // operations retire in order), and before the work, whose latency then
    // hides the claim's.
    asm volatile("" ::: "memory");
    if (claim && threadIdx.x == 0) next = atomicAdd(claim, 1u);
    if (!a.descs) {
        process_tile<THREADS, UB, NT, FEAT>(job, a.one.kr, smem, segs, tsegs);
    } else {
        const kreqs kr = a.descs[job.desc].kr;
        process_tile<THREADS, UB, NT, FEAT>(job, kr, smem, segs, tsegs);
    }
}

// Direct mode: one workgroup per work index. Persistent mode: a grid sized to
// the chip pulls work indices from a device counter until all are taken.
template <bool PERSIST, int THREADS, int UB, bool NT, int FEAT>
__global__ __launch_bounds__(THREADS) void k_gather(gather_args a) {
    // One LDS block, the queues first: their fields then sit at addresses that
    // fit the ds_* immediate offsets (16 bits), ahead of the transpose tiles
    // of one pass (unused by the copy-only variant).
    struct lds_t {
        seg_entry segs[kMaxSeg + 1];
        seg_entry tsegs[FEAT == FEAT_COPY ? 1 : kMaxSeg + 1];
        uint32_t t;  // persistent launches: the next work index
        __attribute__((aligned(16))) unsigned char smem[FEAT == FEAT_COPY ? 16 : trans_tiles<THREADS>() * kSmemBytes];
    };
    __shared__ lds_t lds;
    seg_entry *segs = lds.segs, *tsegs = lds.tsegs;
    unsigned char *smem = lds.smem;
    uint32_t next = 0;
    if (!PERSIST) {
        uint32_t w = blockIdx.x;
        if (a.xcd_ranges) {
            // workgroups are dealt to the 8 XCDs round-robin: give each XCD
            // (and its L2) one contiguous range of work items instead
            const uint32_t n = gridDim.x, x = w & 7u, q = n >> 3, r = n & 7u;
            w = x * q + min(x, r) + (w >> 3);
        }
        process_work<THREADS, UB, NT, FEAT>(a, w, nullptr, next, smem, segs, tsegs);
        return;
    }
    uint32_t &s_t = lds.t;
    if (!a.xcd_ranges) {
        // the next work index is claimed while the current one is moved
        if (threadIdx.x == 0) s_t = atomicAdd(a.queue, 1u);
        __syncthreads();
        uint32_t t = __builtin_amdgcn_readfirstlane(s_t);
        __syncthreads();
        while (t < a.total_tiles) {
            process_work<THREADS, UB, NT, FEAT>(a, t, a.queue, next, smem, segs, tsegs);
            if (threadIdx.x == 0) s_t = next;
            __syncthreads();
            t = __builtin_amdgcn_readfirstlane(s_t);
            __syncthreads();
        }
        return;
    }
    // One contiguous range of work items and one counter per XCD (workgroups
    // are dealt to the 8 XCDs round-robin); a workgroup drains its own XCD's
    // range, then helps the others in turn, so every range is drained and
    // every workgroup exits after visiting all eight.
    const uint32_t n = a.total_tiles, q = n >> 3, r = n & 7u, x0 = blockIdx.x & 7u;
    for (uint32_t i = 0; i < 8; ++i) {
        const uint32_t x = (x0 + i) & 7u;
        const uint32_t lo = x * q + min(x, r), len = q + (x < r ? 1u : 0u);
        if (threadIdx.x == 0) s_t = atomicAdd(a.queue + x, 1u);
        __syncthreads();
        uint32_t c = __builtin_amdgcn_readfirstlane(s_t);
        __syncthreads();
        while (c < len) {
            process_work<THREADS, UB, NT, FEAT>(a, lo + c, a.queue + x, next, smem, segs, tsegs);
            if (threadIdx.x == 0) s_t = next;
            __syncthreads();
            c = __builtin_amdgcn_readfirstlane(s_t);
            __syncthreads();
        }
    }
}

// Resets the work counters; a kernel (not a memset) so that it can sit in a
// captured hipGraph.
template <int WORDS>
__global__ void k_reset_queue(uint32_t *q) {
    if (threadIdx.x < (unsigned)WORDS) q[threadIdx.x] = 0;
}

template <bool P, int TH, int UB, bool NT, int F>
hipError_t launch_variant(const gather_args &a, int grid_limit, hipStream_t stream) {
    if (!P) {
        hipLaunchKernelGGL((k_gather<false, TH, UB, NT, F>), dim3(a.total_tiles), dim3(TH), 0, stream, a);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_reset_queue<kQueueWords>, dim3(1), dim3(64), 0, stream, a.queue);
    const uint32_t grid = std::min<uint32_t>(a.total_tiles, (uint32_t)std::max(1, grid_limit));
    hipLaunchKernelGGL((k_gather<true, TH, UB, NT, F>), dim3(grid), dim3(TH), 0, stream, a);
    return hipGetLastError();
}

// Every variant of one workgroup size and feature level: launch mode, bytes
// in flight, non-temporal hint.
template <int TH, int F>
hipError_t launch_family(const gather_args &a, const launch_cfg &c, int gl, hipStream_t s) {
    if (c.persistent) {
        if (c.unroll_bytes >= 128)
            return c.nontemporal ? launch_variant<true, TH, 128, true, F>(a, gl, s)
                                 : launch_variant<true, TH, 128, false, F>(a, gl, s);
        return c.nontemporal ? launch_variant<true, TH, 64, true, F>(a, gl, s)
                             : launch_variant<true, TH, 64, false, F>(a, gl, s);
    }
    if (c.unroll_bytes >= 128)
        return c.nontemporal ? launch_variant<false, TH, 128, true, F>(a, gl, s)
                             : launch_variant<false, TH, 128, false, F>(a, gl, s);
    return c.nontemporal ? launch_variant<false, TH, 64, true, F>(a, gl, s)
                         : launch_variant<false, TH, 64, false, F>(a, gl, s);
}

// One object per (workgroup size, feature level) family (g2g_gather_variants.hip).
#define G2G_FAMILY_DECL(TH, F) \
    hipError_t launch_gather_##TH##_##F(const gather_args &a, const launch_cfg &c, int gl, hipStream_t s);
G2G_FAMILY_DECL(256, 0)
G2G_FAMILY_DECL(256, 1)
G2G_FAMILY_DECL(256, 2)
G2G_FAMILY_DECL(512, 0)
G2G_FAMILY_DECL(512, 1)
G2G_FAMILY_DECL(512, 2)
#undef G2G_FAMILY_DECL

}  // namespace g2g
