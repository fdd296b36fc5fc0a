// HIP kernels of the MI355X grid2grid transform (gfx950, wave64).
//
// Data movement only: elements move as bit patterns through 4/8/16-byte
// vector registers; the one arithmetic path (beta*dst + alpha*op(src), used
// only when the caller scales) is compiled without FMA contraction.
#include "g2g_internal.hpp"

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

__device__ __forceinline__ double u2d(uint32_t lo, uint32_t hi) {
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
__device__ __forceinline__ u32x2 d2u(double v) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    return u32x2{(uint32_t)b, (uint32_t)(b >> 32)};
}

// beta*dst + alpha*src, evaluated as round(round(beta*dst) + round(alpha*src))
// (memory_utils.hpp:37 in the reference, whose x86-64 build has no FMA);
// complex products are (ac - bd, ad + bc).
#pragma clang fp contract(off)
__device__ __forceinline__ uint32_t scale4(uint32_t s, uint32_t d, const request &rq) {
    const float a = __uint_as_float((uint32_t)rq.alpha[0]);
    const float b = __uint_as_float((uint32_t)rq.beta[0]);
    const float bd = b * __uint_as_float(d);
    const float as = a * __uint_as_float(s);
    return __float_as_uint(bd + as);
}

__device__ __forceinline__ u32x2 scale8(u32x2 s, u32x2 d, const request &rq) {
    u32x2 r;
    if (rq.dtype == DT_F64) {
        const double a = __longlong_as_double((long long)rq.alpha[0]);
        const double b = __longlong_as_double((long long)rq.beta[0]);
        const double bd = b * u2d(d.x, d.y);
        const double as = a * u2d(s.x, s.y);
        r = d2u(bd + as);
    } else {  // complex<float>
        const float ar = __uint_as_float((uint32_t)rq.alpha[0]);
        const float ai = __uint_as_float((uint32_t)(rq.alpha[0] >> 32));
        const float br = __uint_as_float((uint32_t)rq.beta[0]);
        const float bi = __uint_as_float((uint32_t)(rq.beta[0] >> 32));
        const float dr = __uint_as_float(d.x), di = __uint_as_float(d.y);
        const float sr = __uint_as_float(s.x), si = __uint_as_float(s.y);
        const float bdr0 = br * dr, bdr1 = bi * di;
        const float bdi0 = br * di, bdi1 = bi * dr;
        const float asr0 = ar * sr, asr1 = ai * si;
        const float asi0 = ar * si, asi1 = ai * sr;
        const float bd_re = bdr0 - bdr1, bd_im = bdi0 + bdi1;
        const float as_re = asr0 - asr1, as_im = asi0 + asi1;
        r.x = __float_as_uint(bd_re + as_re);
        r.y = __float_as_uint(bd_im + as_im);
    }
    return r;
}

__device__ __forceinline__ u32x4 scale16(u32x4 s, u32x4 d, const request &rq) {
    const double ar = __longlong_as_double((long long)rq.alpha[0]);
    const double ai = __longlong_as_double((long long)rq.alpha[1]);
    const double br = __longlong_as_double((long long)rq.beta[0]);
    const double bi = __longlong_as_double((long long)rq.beta[1]);
    const double dr = u2d(d.x, d.y), di = u2d(d.z, d.w);
    const double sr = u2d(s.x, s.y), si = u2d(s.z, s.w);
    const double bdr0 = br * dr, bdr1 = bi * di;
    const double bdi0 = br * di, bdi1 = bi * dr;
    const double asr0 = ar * sr, asr1 = ai * si;
    const double asi0 = ar * si, asi1 = ai * sr;
    const double bd_re = bdr0 - bdr1, bd_im = bdi0 + bdi1;
    const double as_re = asr0 - asr1, as_im = asi0 + asi1;
    const u32x2 re = d2u(bd_re + as_re), im = d2u(bd_im + as_im);
    return u32x4{re.x, re.y, im.x, im.y};
}
#pragma clang fp contract(on)

__device__ __forceinline__ uint32_t scale_bits(uint32_t s, uint32_t d, const request &rq) {
    return scale4(s, d, rq);
}
__device__ __forceinline__ u32x2 scale_bits(u32x2 s, u32x2 d, const request &rq) {
    return scale8(s, d, rq);
}
__device__ __forceinline__ u32x4 scale_bits(u32x4 s, u32x4 d, const request &rq) {
    return scale16(s, d, rq);
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

__device__ __forceinline__ uint32_t ceil_log2(uint32_t v) {
    return v <= 1 ? 0u : 32u - (uint32_t)__clz(v - 1);
}

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

// Runs queued for one flat copy.
// SEG_VEC: columns of colb bytes starting at s (source, leading dimension
//   lds_b bytes) and d (destination, ldd_b), each cut into vpc = ceil(colb/16)
//   16-byte slots; when colb is not a multiple of 16 the last slot is the
//   column's last 16 bytes (overlapping the slot before it), or, in a column
//   shorter than 16 bytes, its 1-3 dwords.
// SEG_TRANS / SEG_CONJ (a queue of their own): a small transposed piece of
//   ncols final columns: destination element (i, c) at d + c*ldd_b + i*es
//   comes from s + i*lds_b + c*es, its imaginary sign bit flipped for
//   SEG_CONJ. One slot per group of 16/es final columns and, for 4- and 8-byte
//   elements, group of 16/es final rows (vpc = ceil(rows*es/16)), or final
//   row (16-byte elements, vpc = final rows).
// slot0 numbers the segment's first slot in its queue.
enum : uint32_t { SEG_VEC = 0, SEG_TRANS = 1, SEG_CONJ = 2 };
struct seg_entry {
    uint64_t s, d;
    uint64_t lds_b, ldd_b;
    uint32_t colb;
    uint32_t vpc;
    uint32_t slot0;
    uint32_t vmul;  // division magic for vpc
    uint32_t kind;
    uint32_t ncols;  // element segments: final columns of the piece
};
constexpr int kMaxSeg = 64;
// Transposed pieces whose final columns hold fewer bytes than this (counting
// at most 8 bytes an element: f32 below 128 rows, the other types below 64),
// or whose final column ranges are narrower (bytes), go to the element queue
// instead of LDS tiles.
#ifndef G2G_TRANS_ELEM_ROWB
#define G2G_TRANS_ELEM_ROWB 512
#endif
#ifndef G2G_TRANS_ELEM_COLB
#define G2G_TRANS_ELEM_COLB 256
#endif
constexpr uint32_t kTransElemRowBytes = G2G_TRANS_ELEM_ROWB;
constexpr uint32_t kTransElemColBytes = G2G_TRANS_ELEM_COLB;

__device__ __forceinline__ uint32_t magic_div(uint32_t n, uint32_t d, uint32_t mul) {
    const uint32_t l = ceil_log2(d);
    const uint32_t t = __umulhi(mul, n);
    return l ? (t + ((n - t) >> 1)) >> (l - 1) : n;
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
                const uint32_t rel = k - e.slot0;
                const uint32_t c = magic_div(rel, e.vpc, e.vmul);
                const uint32_t v = rel - c * e.vpc;
                // column byte offset of the slot (v*16 < colb: no wrap)
                uint32_t off = v * 16;
                if (e.colb >= 16 && off > e.colb - 16) off = e.colb - 16;
                const gbyte *sp = (const gbyte *)e.s + c * e.lds_b + off;
                dp[u] = (gbyte *)e.d + c * e.ldd_b + off;
                // dwords in this slot: 4, or 1-3 in a column shorter than 16 bytes
                tail[u] = e.colb < 16 ? e.colb >> 2 : 4;
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
        const uint32_t rel = k - e.slot0;
        const uint32_t g = magic_div(rel, e.vpc, e.vmul);
        const uint32_t h = rel - g * e.vpc;
        const uint32_t c = g * VE, i0 = h * VE;
        const uint32_t nc = min(VE, e.ncols - c);               // columns of the group
        const uint32_t nr = min(VE, e.colb / (uint32_t)ES - i0);  // rows of the group
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
                const uint32_t rel = k - e.slot0;
                const uint32_t g = magic_div(rel, e.vpc, e.vmul);
                const uint32_t v = rel - g * e.vpc;
                const uint32_t c = g * VE;
                cnt[u] = min(VE, e.ncols - c);
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

// Scaled column copy, one element per lane slot.
template <int THREADS, int ES>
__device__ __forceinline__ void copy_cols_scaled(const gbyte *__restrict__ s,
                                                 gbyte *__restrict__ d, uint64_t lds_b,
                                                 uint64_t ldd_b, uint32_t rows, uint32_t ncols,
                                                 const request &rq) {
    using T = typename bits_t<ES>::type;
    constexpr int U = 4;
    const uint32_t tid = threadIdx.x;
    const uint32_t slots = ncols * rows;
    const fastdiv per_col(rows);
    for (uint32_t base = tid; base < slots; base += (uint32_t)THREADS * U) {
        T a[U], b[U];
        uint64_t dof[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t k = base + (uint32_t)u * THREADS;
            const uint32_t c = per_col.div(k);
            const uint32_t v = k - c * rows;
            dof[u] = c * ldd_b + (uint64_t)v * ES;
            if (k < slots) {
                a[u] = *(const GLOBAL T *)(s + c * lds_b + (uint64_t)v * ES);
                b[u] = *(const GLOBAL T *)(d + dof[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (base + (uint32_t)u * THREADS < slots) *(GLOBAL T *)(d + dof[u]) = scale_bits(a[u], b[u], rq);
    }
}

constexpr uint32_t kTileBytes = 16384;  // data bytes of one transpose tile
constexpr uint32_t kSmemBytes = kTileBytes + 64 * 16;  // plus one padding element per tile row

// Transpose tiles moved per pass of the full-tile path: enough that every
// thread keeps 128 bytes of loads in flight (a transposing workgroup of 512
// threads shares a CU with only one other, so one 16 KiB tile per pass left a
// CU with 32 KiB of loads in flight). The pass's LDS (4 x 17 KiB for 512
// threads, 2 x 17 KiB for 256) still lets 2 (512) or 3 (256) workgroups share
// a CU.
template <int THREADS>
__host__ __device__ constexpr uint32_t trans_tiles() {
    return THREADS >= 512 ? 4u : 2u;
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
template <int THREADS, bool NT, int ES, uint32_t TC, bool SCALE>
__device__ __forceinline__ void copy_transposed(const gbyte *__restrict__ s,
                                                gbyte *__restrict__ d, uint64_t lds_b,
                                                uint64_t ldd_b, uint32_t rows, uint32_t ncols,
                                                const request &rq, LOCAL unsigned char *smem) {
    using T = typename bits_t<ES>::type;
    constexpr uint32_t VE = 16 / ES;                     // elements per vector
    constexpr uint32_t TR = kTileBytes / (TC * ES);      // final columns per tile
    constexpr uint32_t LDW = TR + 1;                     // LDS row length in elements
    constexpr uint32_t RV = TR / VE;                     // vectors per source run
    constexpr uint32_t RVL = __builtin_ctz(RV);
    constexpr uint32_t W = TC * ES / 16;                 // vectors per final column of a tile
    constexpr uint32_t WL = __builtin_ctz(W);
    constexpr uint32_t PER = kTileBytes / 16 / THREADS;  // vectors per thread per phase
    static_assert(TC * LDW * ES <= kSmemBytes, "transpose tile exceeds its LDS");
    LOCAL T *tile = (LOCAL T *)smem;               // tile[c * LDW + r]
    const bool conj = (rq.flags & REQ_CONJ) != 0;
    const bool scale = SCALE && (rq.flags & REQ_SCALE) != 0;
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
            const uint32_t q = threadIdx.x + k * THREADS;
            const uint32_t c = q >> RVL, r0 = (q & (RV - 1)) * VE;
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
            const uint32_t q = threadIdx.x + k * THREADS;
            const uint32_t c = q >> RVL, r0 = (q & (RV - 1)) * VE;
            LOCAL T *tp = tile + c * LDW + r0;
#pragma unroll
            for (uint32_t e = 0; e < VE; ++e)
                if (e < cnt1[k]) tp[e] = vec_elem<ES>(v1[k], (int)e);
        }
        __syncthreads();
        if (t + 1 < n_tiles) load_tile(t + 1);
#pragma unroll
        for (uint32_t k = 0; k < PER; ++k) {
            const uint32_t q = threadIdx.x + k * THREADS;
            const uint32_t r = q >> WL, w = q & (W - 1);
            const uint32_t j = jc + r, i0 = ic + w * VE;
            if (j < ncols && i0 < rows) {
                gbyte *p = d + j * ldd_b + (uint64_t)i0 * ES;
                const LOCAL T *tp = tile + (w * VE) * LDW + r;
                if (i0 + VE <= rows) {
                    u32x4 v, old;
                    if (scale) old = ld_a4<false>(p);
#pragma unroll
                    for (uint32_t e = 0; e < VE; ++e) {
                        T x = tp[e * LDW];
                        if (conj) x = conj_bits(x);
                        if (scale) x = scale_bits(x, vec_elem<ES>(old, e), rq);
                        vec_set<ES>(v, e, x);
                    }
                    st_a4<NT>(p, v);
                } else {
                    for (uint32_t e = 0; e < rows - i0; ++e) {
                        T x = tp[e * LDW];
                        if (conj) x = conj_bits(x);
                        GLOBAL T *dp = (GLOBAL T *)(p + e * ES);
                        if (scale) x = scale_bits(x, *dp, rq);
                        *dp = x;
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
template <int THREADS, bool NT, int ES, uint32_t TC, bool SCALE>
__device__ __forceinline__ void copy_transposed_full(const gbyte *__restrict__ s_in,
                                                     gbyte *__restrict__ d_in, uint64_t lds_in,
                                                     uint64_t ldd_in, uint32_t n_ic_in, uint32_t n_tiles_in,
                                                     const request &rq, LOCAL unsigned char *smem) {
    using T = typename bits_t<ES>::type;
    constexpr uint32_t G = trans_tiles<THREADS>();
    constexpr uint32_t VE = 16 / ES;
    constexpr uint32_t TR = kTileBytes / (TC * ES);
    constexpr uint32_t LDW = TR + 1;
    constexpr uint32_t RV = TR / VE;
    constexpr uint32_t RVL = __builtin_ctz(RV);
    constexpr uint32_t W = TC * ES / 16;
    constexpr uint32_t WL = __builtin_ctz(W);
    constexpr uint32_t PER = kTileBytes / 16 / THREADS;
    constexpr uint32_t TILE_ELEMS = kSmemBytes / ES;     // LDS area of one tile, in elements
    static_assert(TC * LDW * ES <= kSmemBytes, "transpose tile exceeds its LDS");
    const gbyte *s = (const gbyte *)uniform64((uint64_t)s_in);
    gbyte *d = (gbyte *)uniform64((uint64_t)d_in);
    const uint64_t lds_b = uniform64(lds_in), ldd_b = uniform64(ldd_in);
    const uint32_t n_ic = uniform32(n_ic_in), n_tiles = uniform32(n_tiles_in);
    const bool conj = uniform32(rq.flags & REQ_CONJ) != 0;
    const bool scale = SCALE && uniform32(rq.flags & REQ_SCALE) != 0;
    LOCAL T *tile = (LOCAL T *)smem;               // area g: tile[g*TILE_ELEMS + c*LDW + r]
    // per thread and vector k: source offset inside a tile (storage column,
    // element), destination offset (final column, row), LDS positions
    uint32_t so[PER], dof[PER], lw[PER], lr[PER];
#pragma unroll
    for (uint32_t k = 0; k < PER; ++k) {
        const uint32_t q = threadIdx.x + k * THREADS;
        const uint32_t c = q >> RVL, r0 = (q & (RV - 1)) * VE;
        so[k] = c * (uint32_t)lds_b + r0 * ES;
        lw[k] = c * LDW + r0;
        const uint32_t r = q >> WL, w = q & (W - 1);
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
                    u32x4 o, old;
                    if (scale) old = ld_a4<false>(p);
#pragma unroll
                    for (uint32_t e = 0; e < VE; ++e) {
                        T x = tp[e * LDW];
                        if (conj) x = conj_bits(x);
                        if (scale) x = scale_bits(x, vec_elem<ES>(old, e), rq);
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
// is kept small: one per part, element size and tile height. The request
// arrives as scalars: a struct argument would be copied to the caller's
// private memory.
template <int THREADS, bool NT, int ES, uint32_t TC, bool SCALE, bool FULL>
__device__ __attribute__((noinline)) void transpose_part(const gbyte *__restrict__ s, gbyte *__restrict__ d,
                                                         uint64_t lds_b, uint64_t ldd_b, uint32_t x, uint32_t y,
                                                         uint32_t flags, uint32_t dtype, uint64_t a0, uint64_t a1,
                                                         uint64_t b0, uint64_t b1, LOCAL unsigned char *smem) {
    request rq;
    rq.dtype = dtype;
    rq.flags = flags;
    rq.esize = ES;
    rq.pad = 0;
    rq.alpha[0] = a0;
    rq.alpha[1] = a1;
    rq.beta[0] = b0;
    rq.beta[1] = b1;
    if constexpr (FULL)
        copy_transposed_full<THREADS, NT, ES, TC, SCALE>(s, d, lds_b, ldd_b, x, y, rq, smem);
    else
        copy_transposed<THREADS, NT, ES, TC, SCALE>(s, d, lds_b, ldd_b, x, y, rq, smem);
}

// A transposed piece with tiles of TC final rows: whole tiles through the
// full-tile path, then the bottom strip (the last rows % TC final rows, every
// column) and the right strip (the last ncols % TR final columns of the other
// rows) through the bounds-checked one.
template <int THREADS, bool NT, int ES, uint32_t TC, int FEAT>
__device__ __forceinline__ void transpose_tc(const gbyte *s, gbyte *d, uint64_t lds_b, uint64_t ldd_b,
                                             uint32_t rows, uint32_t ncols, const request &rq,
                                             unsigned char *smem_g) {
    constexpr bool SC = FEAT == FEAT_ALL;
    constexpr uint32_t TR = kTileBytes / (TC * ES);
    LOCAL unsigned char *smem = (LOCAL unsigned char *)smem_g;
    const uint32_t dt = SC ? rq.dtype : 0u;
    const uint64_t a0 = SC ? rq.alpha[0] : 0u, a1 = SC ? rq.alpha[1] : 0u;
    const uint64_t b0 = SC ? rq.beta[0] : 0u, b1 = SC ? rq.beta[1] : 0u;
    // whole tiles: unless a tile spans 4 GiB or more of either side (32-bit
    // offsets inside a tile)
    const bool small_ld = (uint64_t)TC * lds_b < (1ull << 32) && (uint64_t)TR * ldd_b < (1ull << 32);
    const uint32_t rf = small_ld ? rows / TC * TC : 0u, cf = ncols / TR * TR;
    if (rf && cf)
        transpose_part<THREADS, NT, ES, TC, SC, true>(s, d, lds_b, ldd_b, rf / TC, (rf / TC) * (cf / TR), rq.flags,
                                                     dt, a0, a1, b0, b1, smem);
    if (rf < rows)
        transpose_part<THREADS, NT, ES, TC, SC, false>(s + rf * lds_b, d + (uint64_t)rf * ES, lds_b, ldd_b,
                                                      rows - rf, ncols, rq.flags, dt, a0, a1, b0, b1, smem);
    if (rf && cf < ncols)
        transpose_part<THREADS, NT, ES, TC, SC, false>(s + (uint64_t)cf * ES, d + cf * ldd_b, lds_b, ldd_b, rf,
                                                      ncols - cf, rq.flags, dt, a0, a1, b0, b1, smem);
}

// A transposed piece of `rows` final rows: tiles of 64 final rows, or 32 / 16
// for shorter pieces.
template <int THREADS, bool NT, int ES, int FEAT>
__device__ __forceinline__ void transpose_call(const gbyte *s, gbyte *d, uint64_t lds_b, uint64_t ldd_b,
                                               uint32_t rows, uint32_t ncols, const request &rq,
                                               unsigned char *smem) {
    if (rows > 32)
        transpose_tc<THREADS, NT, ES, 64, FEAT>(s, d, lds_b, ldd_b, rows, ncols, rq, smem);
    else if (rows > 16)
        transpose_tc<THREADS, NT, ES, 32, FEAT>(s, d, lds_b, ldd_b, rows, ncols, rq, smem);
    else
        transpose_tc<THREADS, NT, ES, 16, FEAT>(s, d, lds_b, ldd_b, rows, ncols, rq, smem);
}

// Transposed or scaled pieces: moved by the whole workgroup right away.
template <int THREADS, bool NT, int FEAT>
__device__ __forceinline__ void process_piece(const piece &pc, const request &rq, uint32_t c0,
                                              uint32_t c1, unsigned char *smem) {
    const uint32_t es = rq.esize;
    const uint32_t ncols = c1 - c0;
    const uint64_t lds_b = (uint64_t)pc.lds * es;
    const uint64_t ldd_b = (uint64_t)pc.ldd * es;
    gbyte *d = (gbyte *)pc.dst + (uint64_t)c0 * ldd_b;
    if (rq.flags & REQ_TRANS) {
        const gbyte *s = (const gbyte *)pc.src + (uint64_t)c0 * es;
        if (es == 4)
            transpose_call<THREADS, NT, 4, FEAT>(s, d, lds_b, ldd_b, pc.rows, ncols, rq, smem);
        else if (es == 8)
            transpose_call<THREADS, NT, 8, FEAT>(s, d, lds_b, ldd_b, pc.rows, ncols, rq, smem);
        else
            transpose_call<THREADS, NT, 16, FEAT>(s, d, lds_b, ldd_b, pc.rows, ncols, rq, smem);
        return;
    }
    if constexpr (FEAT != FEAT_ALL) return;  // the host never sends scaling here
    const gbyte *s = (const gbyte *)pc.src + (uint64_t)c0 * lds_b;
    if (es == 4)
        copy_cols_scaled<THREADS, 4>(s, d, lds_b, ldd_b, pc.rows, ncols, rq);
    else if (es == 8)
        copy_cols_scaled<THREADS, 8>(s, d, lds_b, ldd_b, pc.rows, ncols, rq);
    else
        copy_cols_scaled<THREADS, 16>(s, d, lds_b, ldd_b, pc.rows, ncols, rq);
}

// Columns [c0, c1) of a piece that belong to the work item covering stream
// bytes [lo, hi), for a piece occupying stream bytes [ps, pe): the columns
// (groups of 256/es columns for transposed pieces) whose first byte lies in
// [lo, hi).
__device__ __forceinline__ void piece_cols(uint32_t rows, uint32_t cols, uint32_t flags, uint32_t es,
                                           uint64_t ps, uint64_t pe, uint64_t lo, uint64_t hi,
                                           uint32_t &c0, uint32_t &c1) {
    c0 = 0;
    c1 = cols;
    if (ps < lo || pe > hi) {  // piece shared with neighbouring work items
        const uint64_t q = (flags & REQ_TRANS) ? 256u / es : 1u;
        const uint64_t qb = (uint64_t)rows * es * q;
        const uint64_t a0 = lo > ps ? ((lo - ps + qb - 1) / qb) * q : 0;
        const uint64_t a1 = ((hi - ps + qb - 1) / qb) * q;
        c0 = (uint32_t)min(a0, (uint64_t)cols);
        c1 = (uint32_t)min(a1, (uint64_t)cols);
    }
}

__device__ __forceinline__ uint32_t wave_scan_u32(uint32_t x, uint32_t lane) {
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    return x;
}

enum : uint32_t { PK_NONE = 0, PK_VEC = 1, PK_ELEM = 2, PK_DIRECT = 3 };

// Work item t covers bytes [t*tile_bytes, (t+1)*tile_bytes) of the piece
// stream (pieces concatenated in table order, each column-major). It moves
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
            flags = pc.req ? kr.r[1].flags : kr.r[0].flags;
            rows = pc.rows;
            piece_cols(rows, pc.cols, flags, es, ps, pe, lo, hi, c0, c1);
            if (in && c0 < c1) {
                const bool trans = (flags & REQ_TRANS) != 0;
                if (FEAT == FEAT_COPY) {
                    // the copy-only variant is launched for untransposed,
                    // unscaled requests only; anything else is left alone
                    kind = (flags & (REQ_TRANS | REQ_SCALE)) ? PK_NONE : PK_VEC;
                } else if (flags & REQ_SCALE) {
                    kind = PK_DIRECT;
                } else if (trans) {
                    // transposed pieces that would leave an LDS tile mostly
                    // empty go to the element queue (register tiles)
                    kind = ((uint64_t)rows * min(es, 8u) < kTransElemRowBytes ||
                            (uint64_t)(c1 - c0) * es < kTransElemColBytes)
                               ? PK_ELEM : PK_DIRECT;
                } else {
                    kind = PK_VEC;
                }
                // copy slots: 16-byte vectors per column; element slots: groups
                // of 16/es final rows (register tiles), or rows for 16-byte types
                vpc = kind == PK_VEC ? (rows * es + 15) >> 4 : es < 16 ? (rows * es + 15) >> 4 : rows;
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
            // element segments: one slot per final row and group of 16/es columns
            const uint32_t groups = is_e ? (c1 - c0 + (16u / es) - 1) / (16u / es) : c1 - c0;
            const uint32_t slots = (is_v || is_e) ? groups * vpc : 0u;
            const uint32_t sv = wave_scan_u32(is_v ? slots : 0u, lane);
            const uint32_t se = wave_scan_u32(is_e ? slots : 0u, lane);
            const uint32_t total = __shfl(sv, 63, 64), ttotal = __shfl(se, 63, 64);
            if (threadIdx.x < 64) {
                const uint64_t below = (1ull << lane) - 1;
                if (is_v || is_e) {
                    const uint64_t ldd_b = (uint64_t)pc.ldd * es, lds_b = (uint64_t)pc.lds * es;
                    seg_entry &e = is_v ? segs[__popcll(mv & below)] : tsegs[__popcll(me & below)];
                    e.s = pc.src + (uint64_t)c0 * (is_v ? lds_b : es);
                    e.d = pc.dst + (uint64_t)c0 * ldd_b;
                    e.lds_b = lds_b;
                    e.ldd_b = ldd_b;
                    e.colb = rows * es;
                    e.vpc = vpc;
                    e.slot0 = is_v ? sv - slots : se - slots;
                    e.vmul = is_v ? pc.vmul : div_magic(vpc);
                    e.kind = is_v ? SEG_VEC : (flags & REQ_CONJ) ? SEG_CONJ : SEG_TRANS;
                    e.ncols = c1 - c0;
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
                const bool r1 = dp.req != 0;
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
        job.lo = (uint64_t)w * a.one.tile_bytes;
        job.hi = min(job.lo + a.one.tile_bytes, a.one.total_bytes);
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
__global__ void k_reset_queue(uint32_t *q) {
    if (threadIdx.x < (unsigned)kQueueWords) q[threadIdx.x] = 0;
}

template <bool P, int TH, int UB, bool NT, int F>
static hipError_t launch_variant(const gather_args &a, int grid_limit, hipStream_t stream) {
    if (!P) {
        hipLaunchKernelGGL((k_gather<false, TH, UB, NT, F>), dim3(a.total_tiles), dim3(TH), 0, stream, a);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(64), 0, stream, a.queue);
    const uint32_t grid = std::min<uint32_t>(a.total_tiles, (uint32_t)std::max(1, grid_limit));
    hipLaunchKernelGGL((k_gather<true, TH, UB, NT, F>), dim3(grid), dim3(TH), 0, stream, a);
    return hipGetLastError();
}

template <bool P, int TH, int UB, bool NT>
static hipError_t launch_feat(const gather_args &a, const launch_cfg &c, int gl, hipStream_t s) {
    if (c.features == FEAT_COPY) return launch_variant<P, TH, UB, NT, FEAT_COPY>(a, gl, s);
    if (c.features == FEAT_TRANS) return launch_variant<P, TH, UB, NT, FEAT_TRANS>(a, gl, s);
    return launch_variant<P, TH, UB, NT, FEAT_ALL>(a, gl, s);
}

template <bool P, int TH, int UB>
static hipError_t launch_nt(const gather_args &a, const launch_cfg &c, int gl, hipStream_t s) {
    return c.nontemporal ? launch_feat<P, TH, UB, true>(a, c, gl, s) : launch_feat<P, TH, UB, false>(a, c, gl, s);
}

template <bool P, int TH>
static hipError_t launch_ub(const gather_args &a, const launch_cfg &c, int gl, hipStream_t s) {
    return c.unroll_bytes >= 128 ? launch_nt<P, TH, 128>(a, c, gl, s) : launch_nt<P, TH, 64>(a, c, gl, s);
}

template <bool P>
static hipError_t launch_th(const gather_args &a, const launch_cfg &c, int gl, hipStream_t s) {
    return c.threads >= 512 ? launch_ub<P, 512>(a, c, gl, s) : launch_ub<P, 256>(a, c, gl, s);
}

hipError_t launch_gather(const gather_args &a_in, const launch_cfg &c, int cus, hipStream_t stream) {
    if (a_in.total_tiles == 0) return hipSuccess;
    gather_args a = a_in;
    a.xcd_ranges = c.xcd_ranges ? 1u : 0u;
    // persistent grids: 16 waves per CU (two 512-thread or four 256-thread
    // workgroups); a third resident 512-thread workgroup, which the copy-only
    // kernels' register counts allow, made the 32768^2 f64 block-cyclic
    // transform ~5 % slower
    const int per_cu = c.threads >= 512 ? 2 : 4;
    const int gl = cus * per_cu;
    return c.persistent ? launch_th<true>(a, c, gl, stream) : launch_th<false>(a, c, gl, stream);
}

__global__ void k_work_records(const work_desc *descs, work_rec *work) {
    const uint32_t k = blockIdx.x;
    const work_desc &d = descs[k];
    for (uint32_t t = threadIdx.x; t < d.n_tiles; t += blockDim.x) {
        work_rec r;
        r.pieces = d.pieces;
        r.pend = d.pend;
        r.lo = (uint64_t)t * d.tile_bytes;
        r.hi = min(r.lo + d.tile_bytes, d.total_bytes);
        r.first_piece = d.first_piece[t];
        r.n_pieces = d.n_pieces;
        r.desc = k;
        r.last_piece = t + 1 < d.n_tiles ? d.first_piece[t + 1] : d.n_pieces - 1;
        work[d.tile_base + t] = r;
    }
}

hipError_t launch_work_records(const work_desc *descs, int32_t n, work_rec *work, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_work_records, dim3((unsigned)n), dim3(256), 0, stream, descs, work);
    return hipGetLastError();
}

// -------------------------------------------------------------- planning

// One thread per union cell p = uj*n_ur + ui: the piece where initial cell
// (urows[ui].iblk, ucols[uj].iblk) meets final cell (fblk, fblk). The source
// pointer follows block::subblock (reference block.cpp:165-196): for a
// transposed initial layout the storage rows run along final columns.
__global__ void k_plan_pieces(const span *urows, int32_t n_ur, const span *ucols, int32_t n_uc,
                              const int32_t *irs, const int32_t *ics, const int32_t *frs,
                              const int32_t *fcs, const cell *iblk, int32_t n_ic,
                              const cell *fblk, int32_t n_fc, uint32_t es, int trans,
                              piece *pieces, uint64_t *bytes, uint32_t *ranks) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (int64_t)n_ur * n_uc) return;
    const int32_t ui = (int32_t)(p % n_ur), uj = (int32_t)(p / n_ur);
    const span R = urows[ui], C = ucols[uj];
    const cell F = fblk[(int64_t)R.fblk * n_fc + C.fblk];
    const cell I = iblk[(int64_t)R.iblk * n_ic + C.iblk];
    const uint32_t rows = (uint32_t)(R.end - R.start), cols = (uint32_t)(C.end - C.start);
    const uint64_t dst =
        F.ptr + ((uint64_t)(C.start - fcs[C.fblk]) * F.ld + (uint64_t)(R.start - frs[R.fblk])) * es;
    uint64_t src;
    if (!trans)
        src = I.ptr + ((uint64_t)(C.start - ics[C.iblk]) * I.ld + (uint64_t)(R.start - irs[R.iblk])) * es;
    else
        src = I.ptr + ((uint64_t)(R.start - irs[R.iblk]) * I.ld + (uint64_t)(C.start - ics[C.iblk])) * es;
    piece pc;
    pc.src = src;
    pc.dst = dst;
    pc.lds = (uint32_t)I.ld;
    pc.ldd = (uint32_t)F.ld;
    pc.rows = rows;
    pc.cols = cols;
    pc.req = 0;
    pc.vmul = div_magic((rows * es + 15) / 16);
    pieces[p] = pc;
    bytes[p] = (uint64_t)rows * cols * es;
    ranks[p] = ((uint32_t)I.owner << 16) | (uint32_t)F.owner;
}

hipError_t launch_plan_pieces(const span *urows, int32_t n_ur, const span *ucols, int32_t n_uc,
                              const int32_t *irs, const int32_t *ics, const int32_t *frs,
                              const int32_t *fcs, const cell *iblk, int32_t n_ic,
                              const cell *fblk, int32_t n_fc, uint32_t esize, int trans,
                              piece *pieces, uint64_t *bytes, uint32_t *ranks,
                              hipStream_t stream) {
    const int64_t n = (int64_t)n_ur * n_uc;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_plan_pieces, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, urows,
                       n_ur, ucols, n_uc, irs, ics, frs, fcs, iblk, n_ic, fblk, n_fc, esize, trans,
                       pieces, bytes, ranks);
    return hipGetLastError();
}

// Sub-piece k = g * n_ur + ui: final columns [c0, c0 + w) of piece (ui, uj),
// (uj, c0) = chunk g. The destination advances by whole columns; so does the
// source of an untransposed piece, while a transposed piece's source storage
// rows run along final columns (element (i, j) at src + (i*lds + j)*es).
__global__ void k_chunk_pieces(const piece *pieces, int32_t n_ur, const uint32_t *chunk_uj,
                               const uint32_t *chunk_c0, int64_t n_chunks, uint32_t w, uint32_t es,
                               int trans, piece *out, uint64_t *bytes) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_chunks * n_ur) return;
    const int64_t g = k / n_ur;
    const int32_t ui = (int32_t)(k - g * n_ur);
    const uint32_t uj = chunk_uj[g], c0 = chunk_c0[g];
    piece pc = pieces[(int64_t)uj * n_ur + ui];
    const uint32_t nc = min(w, pc.cols - c0);
    pc.src += trans ? (uint64_t)c0 * es : (uint64_t)c0 * pc.lds * es;
    pc.dst += (uint64_t)c0 * pc.ldd * es;
    pc.cols = nc;
    out[k] = pc;
    bytes[k] = (uint64_t)pc.rows * nc * es;
}

hipError_t launch_chunk_pieces(const piece *pieces, int32_t n_ur, const uint32_t *chunk_uj,
                               const uint32_t *chunk_c0, int64_t n_chunks, uint32_t w, uint32_t esize,
                               int trans, piece *out, uint64_t *bytes, hipStream_t stream) {
    const int64_t n = n_chunks * n_ur;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_chunk_pieces, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, pieces, n_ur,
                       chunk_uj, chunk_c0, n_chunks, w, esize, trans, out, bytes);
    return hipGetLastError();
}

__global__ void k_piece_bytes(const piece *pieces, int64_t n, kreqs kr, uint64_t *bytes) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const piece pc = pieces[p];
    const uint32_t es = pc.req ? kr.r[1].esize : kr.r[0].esize;
    bytes[p] = (uint64_t)pc.rows * pc.cols * es;
}

hipError_t launch_piece_bytes(const piece *pieces, int64_t n, const kreqs &kr, uint64_t *bytes,
                              hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_piece_bytes, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream,
                       pieces, n, kr, bytes);
    return hipGetLastError();
}

// ------------------------------------------------------- prefix scan

constexpr int kScanItems = 8;
constexpr int kScanBlock = kThreads * kScanItems;  // 2048 values per workgroup

__device__ __forceinline__ uint64_t wave_inclusive_scan(uint64_t x, uint32_t lane) {
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint64_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    return x;
}

// Inclusive scan of one 2048-value block; writes the block total.
__global__ __launch_bounds__(kThreads) void k_scan_block(uint64_t *data, int64_t n,
                                                         uint64_t *block_sums) {
    __shared__ uint64_t wsum[kThreads / 64];
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int64_t base = (int64_t)blockIdx.x * kScanBlock + (int64_t)tid * kScanItems;
    uint64_t v[kScanItems];
    uint64_t run = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint64_t x = (base + k < n) ? data[base + k] : 0;
        run += x;
        v[k] = run;
    }
    const uint64_t incl = wave_inclusive_scan(run, lane);
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    uint64_t wave_base = 0;
    for (uint32_t k = 0; k < w; ++k) wave_base += wsum[k];
    const uint64_t excl = wave_base + incl - run;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) data[base + k] = v[k] + excl;
    if (tid == kThreads - 1 && block_sums) block_sums[blockIdx.x] = excl + run;
}

__global__ __launch_bounds__(kThreads) void k_scan_add(uint64_t *data, int64_t n,
                                                       const uint64_t *block_sums) {
    if (blockIdx.x == 0) return;
    const uint64_t add = block_sums[blockIdx.x - 1];
    const int64_t base = (int64_t)blockIdx.x * kScanBlock;
    for (int k = threadIdx.x; k < kScanBlock; k += kThreads)
        if (base + k < n) data[base + k] += add;
}

size_t scan_scratch_elems(int64_t n) {
    size_t total = 0;
    int64_t m = n;
    while (m > kScanBlock) {
        m = (m + kScanBlock - 1) / kScanBlock;
        total += (size_t)m;
    }
    return total + 1;
}

hipError_t inclusive_scan_u64(uint64_t *data, int64_t n, uint64_t *scratch, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int64_t nb = (n + kScanBlock - 1) / kScanBlock;
    if (nb == 1) {
        hipLaunchKernelGGL(k_scan_block, dim3(1), dim3(kThreads), 0, stream, data, n, (uint64_t *)nullptr);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_scan_block, dim3((unsigned)nb), dim3(kThreads), 0, stream, data, n, scratch);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = inclusive_scan_u64(scratch, nb, scratch + nb, stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(kThreads), 0, stream, data, n,
                       (const uint64_t *)scratch);
    return hipGetLastError();
}

__global__ void k_first_piece(const uint64_t *pend, int64_t n, uint64_t tb, uint32_t *fp) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint64_t s = p ? pend[p - 1] : 0, e = pend[p];
    if (e <= s) return;
    for (uint64_t t = (s + tb - 1) / tb; t * tb < e; ++t) fp[t] = (uint32_t)p;
}

hipError_t launch_first_piece(const uint64_t *pend, int64_t n, uint64_t tile_bytes,
                              uint32_t *first_piece, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_first_piece, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, pend, n,
                       tile_bytes, first_piece);
    return hipGetLastError();
}

// ---------------------------------------------- wire order of the plain form

constexpr int kPairChunk = 2048;  // pieces per chunk of the multi-split
constexpr uint32_t kNoKey = 0xffffffffu;

__device__ __forceinline__ uint32_t pair_key(uint32_t r, int32_t P) {
    const uint32_t s = r >> 16, d = r & 0xffffu;
    return s == d ? kNoKey : s * (uint32_t)P + d;
}

// Per chunk, elements sent per (source, destination) pair -> hist[key][chunk].
__global__ __launch_bounds__(kThreads) void k_pair_hist(const uint32_t *ranks, const piece *pieces,
                                                        int64_t n, int32_t P, uint64_t *hist,
                                                        int64_t n_chunks) {
    extern __shared__ unsigned long long h[];
    const uint32_t K = (uint32_t)P * (uint32_t)P;
    for (uint32_t k = threadIdx.x; k < K; k += kThreads) h[k] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kPairChunk;
    for (int i = threadIdx.x; i < kPairChunk; i += kThreads) {
        const int64_t p = base + i;
        if (p >= n) break;
        const uint32_t key = pair_key(ranks[p], P);
        if (key != kNoKey) atomicAdd(&h[key], (unsigned long long)pieces[p].rows * pieces[p].cols);
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < K; k += kThreads) hist[(int64_t)k * n_chunks + blockIdx.x] = h[k];
}

// Exclusive scan over chunks per key; pair totals.
__global__ void k_pair_scan(uint64_t *hist, int64_t n_chunks, int32_t P, uint64_t *pair_counts) {
    const int64_t key = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (key >= (int64_t)P * P) return;
    uint64_t run = 0;
    for (int64_t c = 0; c < n_chunks; ++c) {
        const uint64_t v = hist[key * n_chunks + c];
        hist[key * n_chunks + c] = run;
        run += v;
    }
    pair_counts[key] = run;
}

// One wave per chunk walks its pieces 64 at a time; for every distinct pair
// key among the 64 lanes a wavefront prefix scan of the piece sizes gives each
// piece its offset inside the pair's segment.
__global__ __launch_bounds__(64) void k_pair_within(const uint32_t *ranks, const piece *pieces,
                                                    int64_t n, int32_t P, const uint64_t *hist,
                                                    int64_t n_chunks, uint64_t *within) {
    extern __shared__ unsigned long long run[];
    const uint32_t K = (uint32_t)P * (uint32_t)P;
    const uint32_t lane = threadIdx.x;
    for (uint32_t k = lane; k < K; k += 64) run[k] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kPairChunk;
    for (int g = 0; g < kPairChunk; g += 64) {
        const int64_t p = base + g + lane;
        const bool valid = p < n;
        const uint32_t key = valid ? pair_key(ranks[p], P) : kNoKey;
        const uint64_t el = valid ? (uint64_t)pieces[p].rows * pieces[p].cols : 0;
        if (valid && key == kNoKey) within[p] = ~0ull;
        unsigned long long pending = __ballot(key != kNoKey);
        while (pending) {
            const int leader = __ffsll((long long)pending) - 1;
            const uint32_t k = __shfl(key, leader, 64);
            const bool mine = key == k;
            const uint64_t x = mine ? el : 0;
            const uint64_t incl = wave_inclusive_scan(x, lane);
            const uint64_t tot = __shfl(incl, 63, 64);
            const uint64_t b = hist[(int64_t)k * n_chunks + blockIdx.x] + run[k];
            if (mine) within[p] = b + incl - x;
            __syncthreads();
            if (lane == 0) run[k] += tot;
            __syncthreads();
            pending &= ~__ballot(mine);
        }
    }
}

hipError_t launch_pair_offsets(const uint32_t *ranks, const piece *pieces, int64_t n,
                               int32_t n_ranks, uint64_t *hist, int64_t n_chunks,
                               uint64_t *within, uint64_t *pair_counts, hipStream_t stream) {
    const size_t K = (size_t)n_ranks * n_ranks;
    const size_t shm = K * sizeof(unsigned long long);
    if (n == 0) {
        return hipMemsetAsync(pair_counts, 0, K * sizeof(uint64_t), stream);
    }
    hipLaunchKernelGGL(k_pair_hist, dim3((unsigned)n_chunks), dim3(kThreads), shm, stream, ranks, pieces,
                       n, n_ranks, hist, n_chunks);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_pair_scan, dim3((unsigned)((K + 255) / 256)), dim3(256), 0, stream, hist, n_chunks,
                       n_ranks, pair_counts);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_pair_within, dim3((unsigned)n_chunks), dim3(64), shm, stream, ranks, pieces, n,
                       n_ranks, (const uint64_t *)hist, n_chunks, within);
    return hipGetLastError();
}

// Stage tables of the plain form. Stage A: pack remote pieces into the
// source rank's send buffer (packed column-major, ld = piece rows, as
// copy_block_to_buffer, reference communication_data.cpp:101-117) and copy
// local pieces straight to the destination (copy_local_blocks, :224-252).
// Stage C: unpack remote pieces from the destination rank's receive buffer
// (copy_block_from_buffer, :119-130); local pieces become empty.
__global__ void k_stage_pieces(const piece *pieces, const uint32_t *ranks, const uint64_t *within,
                               int64_t n, int32_t P, const uint64_t *sbase, const uint64_t *sdspl,
                               const uint64_t *rbase, const uint64_t *rdspl, uint64_t send_arena,
                               uint64_t recv_arena, uint32_t es, piece *sa, piece *sc) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const piece pc = pieces[p];
    const uint32_t r = ranks[p];
    const uint32_t s = r >> 16, d = r & 0xffffu;
    piece a = pc, c = pc;
    if (s != d) {
        const uint64_t w = within[p];
        const uint64_t soff = sbase[s] + sdspl[(int64_t)s * P + d] + w;
        const uint64_t roff = rbase[d] + rdspl[(int64_t)d * P + s] + w;
        a.dst = send_arena + soff * es;
        a.ldd = pc.rows;
        a.req = 0;
        c.src = recv_arena + roff * es;
        c.lds = pc.rows;
        c.req = 0;
    } else {
        a.req = 1;
        c.src = 0;
        c.dst = 0;
        c.cols = 0;
        c.req = 0;
    }
    sa[p] = a;
    sc[p] = c;
}

hipError_t launch_stage_pieces(const piece *pieces, const uint32_t *ranks, const uint64_t *within,
                               int64_t n, int32_t n_ranks, const uint64_t *sbase,
                               const uint64_t *sdspl, const uint64_t *rbase,
                               const uint64_t *rdspl, uint64_t send_arena, uint64_t recv_arena,
                               uint32_t esize, piece *stage_a, piece *stage_c,
                               hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_stage_pieces, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, pieces,
                       ranks, within, n, n_ranks, sbase, sdspl, rbase, rdspl, send_arena, recv_arena,
                       esize, stage_a, stage_c);
    return hipGetLastError();
}

// --------------------------------------------------- seeded test values

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    uint64_t z = x + 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// v(seed, i, j) as raw bits in a u32x4 (unused lanes zero).
__device__ __forceinline__ u32x4 seeded_value(int dtype, uint64_t seed, int64_t i, int64_t j,
                                              int64_t n_cols) {
    const uint64_t key = (uint64_t)i * (uint64_t)n_cols + (uint64_t)j;
    u32x4 r = {0u, 0u, 0u, 0u};
    if (dtype == DT_F32) {
        const uint64_t x = splitmix64(seed ^ key);
        r.x = __float_as_uint((float)(x >> 40) * 0x1p-23f - 1.0f);
    } else if (dtype == DT_F64) {
        const uint64_t x = splitmix64(seed ^ key);
        const u32x2 v = d2u((double)(x >> 11) * 0x1p-52 - 1.0);
        r.x = v.x;
        r.y = v.y;
    } else if (dtype == DT_C64) {
        const uint64_t x = splitmix64(seed ^ key);
        r.x = __float_as_uint((float)(x >> 40) * 0x1p-23f - 1.0f);
        r.y = __float_as_uint((float)((x >> 16) & 0xffffffull) * 0x1p-23f - 1.0f);
    } else {
        const uint64_t x0 = splitmix64(seed ^ (2 * key));
        const uint64_t x1 = splitmix64(seed ^ (2 * key + 1));
        const u32x2 v0 = d2u((double)(x0 >> 11) * 0x1p-52 - 1.0);
        const u32x2 v1 = d2u((double)(x1 >> 11) * 0x1p-52 - 1.0);
        r = u32x4{v0.x, v0.y, v1.x, v1.y};
    }
    return r;
}

__device__ __forceinline__ uint32_t dtype_size(int dtype) {
    return dtype == DT_F32 ? 4u : (dtype == DT_C128 ? 16u : 8u);
}

__global__ __launch_bounds__(kThreads) void k_fill(const fill_block *blocks, int dtype,
                                                   uint64_t seed, int64_t n_cols) {
    const fill_block b = blocks[blockIdx.x];
    const uint32_t es = dtype_size(dtype);
    const int32_t rows = b.r1 - b.r0, cols = b.c1 - b.c0;
    for (int32_t j = blockIdx.y; j < cols; j += gridDim.y) {
        unsigned char *col = (unsigned char *)b.ptr + (uint64_t)j * b.ld * es;
        for (int32_t i = threadIdx.x; i < rows; i += kThreads) {
            const u32x4 v = seeded_value(dtype, seed, b.r0 + i, b.c0 + j, n_cols);
            unsigned char *e = col + (uint64_t)i * es;
            if (es == 4)
                *(uint32_t *)e = v.x;
            else if (es == 8)
                *(u32x2 *)e = u32x2{v.x, v.y};
            else
                *(u32x4 *)e = v;
        }
    }
}

__global__ __launch_bounds__(kThreads) void k_verify(const fill_block *blocks, int dtype,
                                                     uint64_t seed, int trans, int conj,
                                                     int64_t a_n_cols, unsigned long long *bad) {
    const fill_block b = blocks[blockIdx.x];
    const uint32_t es = dtype_size(dtype);
    const int32_t rows = b.r1 - b.r0, cols = b.c1 - b.c0;
    unsigned long long cnt = 0;
    for (int32_t j = blockIdx.y; j < cols; j += gridDim.y) {
        const unsigned char *col = (const unsigned char *)b.ptr + (uint64_t)j * b.ld * es;
        for (int32_t i = threadIdx.x; i < rows; i += kThreads) {
            const int64_t gi = b.r0 + i, gj = b.c0 + j;
            u32x4 w = trans ? seeded_value(dtype, seed, gj, gi, a_n_cols)
                            : seeded_value(dtype, seed, gi, gj, a_n_cols);
            if (conj) {
                if (dtype == DT_C64) w.y ^= 0x80000000u;
                if (dtype == DT_C128) w.w ^= 0x80000000u;
            }
            const unsigned char *e = col + (uint64_t)i * es;
            bool ok;
            if (es == 4)
                ok = *(const uint32_t *)e == w.x;
            else if (es == 8) {
                const u32x2 g = *(const u32x2 *)e;
                ok = g.x == w.x && g.y == w.y;
            } else {
                const u32x4 g = *(const u32x4 *)e;
                ok = g.x == w.x && g.y == w.y && g.z == w.z && g.w == w.w;
            }
            cnt += ok ? 0ull : 1ull;
        }
    }
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_down(cnt, off, 64);
    if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(bad, cnt);
}

hipError_t launch_fill(const fill_block *blocks, int32_t n_blocks, int dtype, uint64_t seed,
                       int64_t n_cols, hipStream_t stream) {
    if (n_blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)n_blocks, 32), dim3(kThreads), 0, stream, blocks, dtype,
                       seed, n_cols);
    return hipGetLastError();
}

// Streaming reference rates of the device (test support): each thread moves
// 8 16-byte vectors per grid stride with non-temporal hints; copy reads
// src and writes dst, read only reads (folding into one value that is
// written only if it matches a constant), write only writes. No scratch,
// every access below n vectors.
template <int KIND>
__global__ __launch_bounds__(kThreads) void k_stream(const u32x4 *__restrict__ src, u32x4 *__restrict__ dst,
                                                     uint64_t n) {
    constexpr int U = 8;
    const uint64_t stride = (uint64_t)gridDim.x * kThreads * U;
    u32x4 acc = {0u, 0u, 0u, 0u};
    const u32x4 fill = {0x3ff00000u, 0u, 0x3ff00000u, threadIdx.x};
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads * U + threadIdx.x; i < n; i += stride) {
        u32x4 r[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (KIND != 2 && i + (uint64_t)u * kThreads < n) r[u] = __builtin_nontemporal_load(src + i + u * kThreads);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (i + (uint64_t)u * kThreads >= n) continue;
            if (KIND == 0) __builtin_nontemporal_store(r[u], dst + i + u * kThreads);
            if (KIND == 1) acc ^= r[u];
            if (KIND == 2) __builtin_nontemporal_store(fill, dst + i + u * kThreads);
        }
    }
    if (KIND == 1 && (acc.x ^ acc.y ^ acc.z ^ acc.w) == 0x9e3779b9u && n) dst[0] = acc;
}

hipError_t launch_stream(int kind, const void *src, void *dst, uint64_t bytes, int grid, hipStream_t stream) {
    const uint64_t n = bytes / 16;
    const u32x4 *s = (const u32x4 *)src;
    u32x4 *d = (u32x4 *)dst;
    if (kind == 0) hipLaunchKernelGGL(k_stream<0>, dim3((unsigned)grid), dim3(kThreads), 0, stream, s, d, n);
    else if (kind == 1) hipLaunchKernelGGL(k_stream<1>, dim3((unsigned)grid), dim3(kThreads), 0, stream, s, d, n);
    else hipLaunchKernelGGL(k_stream<2>, dim3((unsigned)grid), dim3(kThreads), 0, stream, s, d, n);
    return hipGetLastError();
}

hipError_t launch_verify(const fill_block *blocks, int32_t n_blocks, int dtype, uint64_t seed,
                         int trans, int conj, int64_t a_n_cols, unsigned long long *bad,
                         hipStream_t stream) {
    if (n_blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_verify, dim3((unsigned)n_blocks, 32), dim3(kThreads), 0, stream, blocks, dtype,
                       seed, trans, conj, a_n_cols, bad);
    return hipGetLastError();
}

}  // namespace g2g
