// Peak picking of one raw spectrum, as MSQuery::pickpeaks does it
// (msquery.cpp:338-423): sort the peaks by intensity with the reference's
// key/value quicksort (keyval.h:49-103), set the most intense peak to
// base_int, scale the next qalen-1 peaks by base_int / max and keep them
// while they stay >= min_int. The kept peaks are the last `n_kept` of the
// sorted arrays, in ascending-intensity order.
#pragma once
#include <cstdint>
#include <utility>

namespace gx {

// (int) of a double as x86-64 computes it (cvttsd2si): truncation toward
// zero, and INT32_MIN for NaN and values outside the int32 range, which is
// what the reference's `int *= double` stores for them.
inline int32_t to_int32_x86(double x) {
    return (x > -2147483649.0 && x < 2147483648.0) ? (int32_t)x : INT32_MIN;
}

// Ascending sort of key[0..n) carrying val along, with the partition steps of
// KeyVal_Parallel_Internal (keyval.h:49-103): middle pivot, Hoare scan, swap
// and step both cursors. Equal keys end where those steps leave them, so the
// order of ties is the reference's. Subranges are independent, so the order
// they are visited in does not change the result; the smaller one is taken
// first, which bounds the stack by log2(n).
inline void keyval_sort(int32_t *key, int32_t *val, int64_t n) {
    if (n < 2) return;
    int64_t stk[2 * 64];
    int sp = 0;
    stk[sp++] = 0;
    stk[sp++] = n - 1;
    while (sp > 0) {
        const int64_t right = stk[--sp], left = stk[--sp];
        int64_t i = left, j = right;
        const int32_t pivot = key[(left + right) / 2];
        while (i <= j) {
            while (key[i] < pivot) i++;
            while (key[j] > pivot) j--;
            if (i <= j) {
                std::swap(key[i], key[j]);
                std::swap(val[i], val[j]);
                i++;
                j--;
            }
        }
        const bool lo = left < j, hi = i < right;
        if (lo && hi) {
            // push the larger range first so the smaller is popped next
            if (j - left > right - i) {
                stk[sp++] = left, stk[sp++] = j;
                stk[sp++] = i, stk[sp++] = right;
            } else {
                stk[sp++] = i, stk[sp++] = right;
                stk[sp++] = left, stk[sp++] = j;
            }
        } else if (lo) {
            stk[sp++] = left, stk[sp++] = j;
        } else if (hi) {
            stk[sp++] = i, stk[sp++] = right;
        }
    }
}

// Normalisation and cut of an intensity-sorted spectrum (msquery.cpp:350-376,
// the non-STL branch); returns the number of kept peaks.
inline int64_t normalize_sorted(int32_t *in, int64_t n, int32_t base_int, int32_t min_int, int32_t qalen) {
    if (n <= 0) return 0;
    const double factor = (double)base_int / (double)in[n - 1];
    in[n - 1] = base_int;
    int64_t kept = 1;
    for (int64_t j = n - 2; j >= n - qalen && j >= 0; j--) {
        in[j] = to_int32_x86((double)in[j] * factor);
        if (in[j] >= min_int)
            kept++;
        else
            break;
    }
    return kept >= qalen ? qalen : kept;
}

// Both steps on the intensity/m-z arrays of one spectrum, in place.
inline int64_t pick_peaks(int32_t *in, int32_t *mz, int64_t n, int32_t base_int, int32_t min_int, int32_t qalen) {
    keyval_sort(in, mz, n);
    return normalize_sorted(in, n, base_int, min_int, qalen);
}

}  // namespace gx
