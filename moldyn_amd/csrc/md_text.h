// md_text.h -- decimal text of doubles and integers as the reference's files
// hold them.
//
// The reference writes every f64 of its CSV files through the `csv` 1.3 crate
// (core/src/save_data.rs:165-213, 291-300) and of potentials.json through
// `serde_json` 1.0 (solver/src/solver/potential.rs:104-122).  Both print
// an f64 with the `ryu` crate: the shortest decimal that reads back to the same
// double (closest to it when several are as short, ties to an even last
// digit), laid out as
//   "0.0", "-0.0", "NaN", "inf", "-inf",
//   digits d_1..d_n times 10^k with kk = n + k:
//     0 <= k and kk <= 16   ->  d_1..d_n, (kk - n) zeros, ".0"      1234e7  -> 12340000000.0
//     0 < kk <= 16          ->  d_1..d_kk "." d_kk+1..d_n            1234e-2 -> 12.34
//     -5 < kk <= 0          ->  "0." (-kk zeros) d_1..d_n            1234e-6 -> 0.001234
//     n == 1                ->  d_1 "e" (kk - 1)                     1e30
//     otherwise             ->  d_1 "." d_2..d_n "e" (kk - 1)        1.234e33
// with the exponent printed without "+" or leading zeros.
//
// shortest_decimal() restates Ryu (U. Adams, PLDI 2018) as the ryu crate
// runs it: the interval of decimals that round to the double, scaled by a
// 125-bit power of 5 (md_text_tables.h), then digits removed while the
// interval still holds a shorter candidate.  The functions are usable on the
// device and the host; the frame kernels (md_frames.hip) call them on the GPU.
// The digit-removal loop and the log10/log2 approximations follow the
// structure of Ryu's reference implementation in C (d2s.c of
// github.com/ulfjack/ryu, by Ulf Adams, licensed Apache-2.0 or BSL-1.0),
// written here for this code's own types and tables.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "md_text_tables.h"

namespace md {

#define MD_TEXT_FN __host__ __device__ __forceinline__

// longest text of one double ("-1.2345678901234567e-308")
constexpr int kF64Chars = 24;

struct Pow5Tables {
    const uint64_t (*inv)[2];  // MD_POW5_INV_COUNT entries {lo, hi}
    const uint64_t (*pos)[2];  // MD_POW5_COUNT entries
};

// value = digits * 10^exp10; sign, special kinds apart
struct Decimal64 {
    uint64_t digits;
    int32_t exp10;
    int32_t kind;  // 0 finite nonzero, 1 zero, 2 NaN, 3 infinity
    bool neg;
};

MD_TEXT_FN uint32_t txt_pow5bits(int32_t e) { return (((uint32_t)e * 1217359u) >> 19) + 1u; }
MD_TEXT_FN uint32_t txt_log10_pow2(int32_t e) { return ((uint32_t)e * 78913u) >> 18; }
MD_TEXT_FN uint32_t txt_log10_pow5(int32_t e) { return ((uint32_t)e * 732923u) >> 20; }

MD_TEXT_FN bool txt_pow5_divides(uint64_t v, uint32_t p) {
    // 5^p divides v (v > 0)
    uint32_t c = 0;
    while (c < p && v % 5u == 0) {
        v /= 5u;
        ++c;
    }
    return c >= p;
}

MD_TEXT_FN uint64_t txt_mul_shift(uint64_t m, const uint64_t* mul, int32_t j) {
    const unsigned __int128 b0 = (unsigned __int128)m * mul[0];
    const unsigned __int128 b2 = (unsigned __int128)m * mul[1];
    return (uint64_t)(((b0 >> 64) + b2) >> (j - 64));
}

MD_TEXT_FN Decimal64 shortest_decimal(double x, const Pow5Tables& tb) {
    Decimal64 d{0, 0, 0, false};
    uint64_t bits;
    __builtin_memcpy(&bits, &x, 8);
    d.neg = (bits >> 63) != 0;
    const uint64_t frac = bits & ((1ull << 52) - 1);
    const uint32_t bexp = (uint32_t)(bits >> 52) & 0x7ffu;
    if (bexp == 0x7ffu) {
        d.kind = frac ? 2 : 3;
        return d;
    }
    if (bexp == 0 && frac == 0) {
        d.kind = 1;
        return d;
    }
    // x = mv 2^e2 with mv = 4 m2; the neighbours' midpoints are mp 2^e2, mm 2^e2
    int32_t e2;
    uint64_t m2;
    if (bexp == 0) {
        e2 = 1 - 1075 - 2;
        m2 = frac;
    } else {
        e2 = (int32_t)bexp - 1075 - 2;
        m2 = (1ull << 52) | frac;
    }
    const bool accept = (m2 & 1) == 0;  // round-half-even readers take the bounds of an even mantissa
    const uint64_t mv = 4 * m2;
    const uint32_t mm_shift = (frac != 0 || bexp <= 1) ? 1u : 0u;  // below a power of two the gap halves
    const uint64_t mp = mv + 2, mm = mv - 1 - mm_shift;
    // vr, vp, vm = floor of mv, mp, mm times 2^e2 / 10^e10; *_tz: the floor was exact
    uint64_t vr, vp, vm;
    int32_t e10;
    bool vm_tz = false, vr_tz = false;
    if (e2 >= 0) {
        const uint32_t q = txt_log10_pow2(e2) - (e2 > 3 ? 1u : 0u);
        e10 = (int32_t)q;
        const int32_t k = 125 + (int32_t)txt_pow5bits((int32_t)q) - 1;
        const int32_t j = -e2 + (int32_t)q + k;
        const uint64_t* mul = tb.inv[q];
        vr = txt_mul_shift(mv, mul, j);
        vp = txt_mul_shift(mp, mul, j);
        vm = txt_mul_shift(mm, mul, j);
        if (q <= 21) {
            // at most one of mm, mv, mp is a multiple of 5
            if (mv % 5u == 0)
                vr_tz = txt_pow5_divides(mv, q);
            else if (accept)
                vm_tz = txt_pow5_divides(mm, q);
            else if (txt_pow5_divides(mp, q))
                --vp;
        }
    } else {
        const uint32_t q = txt_log10_pow5(-e2) - (-e2 > 1 ? 1u : 0u);
        e10 = (int32_t)q + e2;
        const int32_t i = -e2 - (int32_t)q;
        const int32_t k = (int32_t)txt_pow5bits(i) - 125;
        const int32_t j = (int32_t)q - k;
        const uint64_t* mul = tb.pos[i];
        vr = txt_mul_shift(mv, mul, j);
        vp = txt_mul_shift(mp, mul, j);
        vm = txt_mul_shift(mm, mul, j);
        if (q <= 1) {
            // mv has two trailing zero bits, mp one, mm one when mm_shift is 1
            vr_tz = true;
            if (accept)
                vm_tz = mm_shift == 1;
            else
                --vp;
        } else if (q < 63) {
            vr_tz = (mv & ((1ull << q) - 1)) == 0;
        }
    }
    // remove digits while a shorter decimal stays inside (vm, vp]
    int32_t removed = 0;
    uint32_t last = 0;  // last digit removed from vr
    uint64_t out;
    if (vm_tz || vr_tz) {
        for (;;) {
            const uint64_t vp10 = vp / 10u, vm10 = vm / 10u;
            if (vp10 <= vm10) break;
            const uint64_t vr10 = vr / 10u;
            vm_tz = vm_tz && (vm - 10u * vm10) == 0;
            vr_tz = vr_tz && last == 0;
            last = (uint32_t)(vr - 10u * vr10);
            vr = vr10;
            vp = vp10;
            vm = vm10;
            ++removed;
        }
        if (vm_tz) {
            // the lower bound itself is a candidate: keep removing its zeros
            for (;;) {
                const uint64_t vm10 = vm / 10u;
                if (vm - 10u * vm10 != 0) break;
                const uint64_t vr10 = vr / 10u;
                vr_tz = vr_tz && last == 0;
                last = (uint32_t)(vr - 10u * vr10);
                vr = vr10;
                vp = vp / 10u;
                vm = vm10;
                ++removed;
            }
        }
        if (vr_tz && last == 5 && vr % 2 == 0) last = 4;  // exactly halfway: keep the even digit
        out = vr + (((vr == vm && (!accept || !vm_tz)) || last >= 5) ? 1u : 0u);
    } else {
        bool up = false;
        for (;;) {
            const uint64_t vp10 = vp / 10u, vm10 = vm / 10u;
            if (vp10 <= vm10) break;
            const uint64_t vr10 = vr / 10u;
            up = (vr - 10u * vr10) >= 5;
            vr = vr10;
            vp = vp10;
            vm = vm10;
            ++removed;
        }
        out = vr + ((vr == vm || up) ? 1u : 0u);
    }
    d.digits = out;
    d.exp10 = e10 + removed;
    return d;
}

MD_TEXT_FN int txt_digit_count(uint64_t v) {
    int n = 1;
    while (v >= 10u) {
        v /= 10u;
        ++n;
    }
    return n;
}

// length of the text of d (layout in the header comment)
MD_TEXT_FN int decimal_text_len(const Decimal64& d) {
    if (d.kind == 2) return 3;
    if (d.kind == 3) return d.neg ? 4 : 3;
    const int s = d.neg ? 1 : 0;
    if (d.kind == 1) return s + 3;
    const int n = txt_digit_count(d.digits);
    const int k = d.exp10, kk = n + k;
    if (k >= 0 && kk <= 16) return s + kk + 2;
    if (kk > 0 && kk <= 16) return s + n + 1;
    if (kk > -5 && kk <= 0) return s + 2 - kk + n;
    const int e = kk - 1, ae = e < 0 ? -e : e;
    const int elen = (e < 0 ? 1 : 0) + (ae >= 100 ? 3 : ae >= 10 ? 2 : 1);
    return s + (n == 1 ? 1 : n + 1) + 1 + elen;
}

// writes the text of d at out (decimal_text_len(d) bytes); returns the length
template <typename Out>
MD_TEXT_FN int decimal_text(const Decimal64& d, Out out) {
    int p = 0;
    if (d.kind == 2) {
        out[0] = 'N';
        out[1] = 'a';
        out[2] = 'N';
        return 3;
    }
    if (d.neg) out[p++] = '-';
    if (d.kind == 3) {
        out[p] = 'i';
        out[p + 1] = 'n';
        out[p + 2] = 'f';
        return p + 3;
    }
    if (d.kind == 1) {
        out[p] = '0';
        out[p + 1] = '.';
        out[p + 2] = '0';
        return p + 3;
    }
    char dig[20];
    const int n = txt_digit_count(d.digits);
    uint64_t v = d.digits;
    for (int i = n - 1; i >= 0; --i) {
        dig[i] = (char)('0' + (int)(v % 10u));
        v /= 10u;
    }
    const int k = d.exp10, kk = n + k;
    if (k >= 0 && kk <= 16) {
        for (int i = 0; i < n; ++i) out[p++] = dig[i];
        for (int i = n; i < kk; ++i) out[p++] = '0';
        out[p++] = '.';
        out[p++] = '0';
    } else if (kk > 0 && kk <= 16) {
        for (int i = 0; i < kk; ++i) out[p++] = dig[i];
        out[p++] = '.';
        for (int i = kk; i < n; ++i) out[p++] = dig[i];
    } else if (kk > -5 && kk <= 0) {
        out[p++] = '0';
        out[p++] = '.';
        for (int i = 0; i < -kk; ++i) out[p++] = '0';
        for (int i = 0; i < n; ++i) out[p++] = dig[i];
    } else {
        out[p++] = dig[0];
        if (n > 1) {
            out[p++] = '.';
            for (int i = 1; i < n; ++i) out[p++] = dig[i];
        }
        out[p++] = 'e';
        int e = kk - 1;
        if (e < 0) {
            out[p++] = '-';
            e = -e;
        }
        if (e >= 100) {
            out[p++] = (char)('0' + e / 100);
            out[p++] = (char)('0' + (e / 10) % 10);
            out[p++] = (char)('0' + e % 10);
        } else if (e >= 10) {
            out[p++] = (char)('0' + e / 10);
            out[p++] = (char)('0' + e % 10);
        } else {
            out[p++] = (char)('0' + e);
        }
    }
    return p;
}

MD_TEXT_FN int uint_text_len(uint64_t v) { return txt_digit_count(v); }

template <typename Out>
MD_TEXT_FN int uint_text(uint64_t v, Out out) {
    const int n = txt_digit_count(v);
    for (int i = n - 1; i >= 0; --i) {
        out[i] = (char)('0' + (int)(v % 10u));
        v /= 10u;
    }
    return n;
}

}  // namespace md
