// Device code every walk shares: a walk's moves packed 2 bits each, the CIGAR text they make, and the placing
// of every lane's text in one output buffer a wave at a time (reference cigar.mojo:157-216 for the text).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dinara {

// A walk's moves, 2 bits each, 32 to a 64-bit word, in the order the walk makes them (from the pair's end):
// 0 two equal letters ('=', or 'M' for both kinds without eqx), 1 two unequal ('X'), 2 'D', 3 'I'.
struct OpTape {
    uint64_t *words;
    uint64_t held;
    int64_t count;

    __device__ void put(uint32_t code) {
        held |= (uint64_t)code << (2 * (count & 31));
        if ((++count & 31) == 0) {
            words[(count >> 5) - 1] = held;
            held = 0;
        }
    }
    __device__ void close() {
        if (count & 31) words[count >> 5] = held;
    }
};

__device__ __forceinline__ int decimal_digits(uint32_t v) {
    return v < 10u ? 1 : v < 100u ? 2 : v < 1000u ? 3 : v < 10000u ? 4 : v < 100000u ? 5 : v < 1000000u ? 6
         : v < 10000000u ? 7 : v < 100000000u ? 8 : v < 1000000000u ? 9 : 10;
}

// The CIGAR text of a walk's moves: runs of one move, each its length then its letter, in the alignment's order
// (the moves' own order when the sequences were walked reversed, else back to front). Returns the text's length;
// writes it to `out` unless null. Runs are found a word at a time: the moves that differ from the run's are the
// nonzero 2-bit fields of the word xor the run's code repeated.
__device__ __forceinline__ int64_t cigar_text(int eqx, int forward, const uint64_t *words, int64_t count, char *out) {
    const char letter_of[4] = {eqx ? '=' : 'M', 'X', 'D', 'I'};
    int64_t length = 0;
    auto emit = [&](uint32_t code, int64_t run) {
        const uint32_t v = (uint32_t)run;
        const int digits = decimal_digits(v);
        if (out) {
            uint32_t rest = v;
            for (int i = digits - 1; i >= 0; --i) {
                out[length + i] = (char)('0' + rest % 10u);
                rest /= 10u;
            }
            out[length + digits] = letter_of[code];
        }
        length += digits + 1;
    };
    auto code_at = [&](int64_t p) { return (uint32_t)(words[p >> 5] >> (2 * (p & 31))) & 3u; };
    if (forward) {
        int64_t p = 0;
        while (p < count) {
            const uint32_t code = code_at(p);
            const uint64_t repeated = (uint64_t)code * 0x5555555555555555ull;
            int64_t q = p;
            while (q < count) {
                const int off = (int)(q & 31);
                const uint64_t x = (words[q >> 5] ^ repeated) >> (2 * off);
                if (x == 0) {
                    q += 32 - off;
                    continue;
                }
                q += __builtin_ctzll(x) >> 1;
                break;
            }
            q = q < count ? q : count;
            emit(code, q - p);
            p = q;
        }
    } else {
        int64_t p = count - 1;
        while (p >= 0) {
            const uint32_t code = code_at(p);
            const uint64_t repeated = (uint64_t)code * 0x5555555555555555ull;
            int64_t q = p;
            while (q >= 0) {
                const int off = (int)(q & 31);
                const uint64_t x = (words[q >> 5] ^ repeated) << (2 * (31 - off));
                if (x == 0) {
                    q -= off + 1;
                    continue;
                }
                q -= __builtin_clzll(x) >> 1;
                break;
            }
            emit(code, p - q);
            p = q;
        }
    }
    return length;
}

// Claims room for every lane's CIGAR in the shared output with one atomic a wave: an exclusive prefix sum of the
// lengths across the wave, its total added to the counter by the last lane. Every lane of the wave must call it
// (a lane with nothing to publish brings length 0).
__device__ __forceinline__ unsigned long long wave_claim(unsigned long long *counter, unsigned long long length) {
    const int lane = threadIdx.x & 63;
    unsigned long long inclusive = length;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned long long other = __shfl_up(inclusive, d);
        if (lane >= d) inclusive += other;
    }
    unsigned long long base = 0;
    if (lane == 63 && inclusive > 0) base = atomicAdd(counter, inclusive);
    base = __shfl(base, 63);
    return base + inclusive - length;
}

}  // namespace dinara
