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
    // `moves` (1 to 24) moves at once, the first in bits 0-1 of `codes`.
    __device__ void put_many(uint64_t codes, int moves) {
        const int fill = (int)(count & 31);
        held |= codes << (2 * fill);
        if (fill + moves >= 32) {
            words[count >> 5] = held;
            held = fill + moves > 32 ? codes >> (2 * (32 - fill)) : 0;
        }
        count += moves;
    }
    __device__ void close() {
        if (count & 31) words[count >> 5] = held;
    }
};

__device__ __forceinline__ int decimal_digits(uint32_t v) {
    return v < 10u ? 1 : v < 100u ? 2 : v < 1000u ? 3 : v < 10000u ? 4 : v < 100000u ? 5 : v < 1000000u ? 6
         : v < 10000000u ? 7 : v < 100000000u ? 8 : v < 1000000000u ? 9 : 10;
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

// The CIGAR text of one lane's tape, by that lane alone (the form for short tapes, where a whole wave on each
// pair costs more than it saves): runs of one move, each its length then its letter, in the alignment's order (the
// moves' own order when the sequences were walked reversed, else back to front). Returns the text's length;
// writes it to `out` unless null. Runs are found a word at a time: the moves that differ from the run's are the
// nonzero 2-bit fields of the word xor the run's code repeated.
__device__ __forceinline__ int64_t lane_cigar_text(int eqx, int forward, const uint64_t *words, int64_t count,
                                                   char *out) {
    const uint32_t letters = (uint32_t)(eqx ? '=' : 'M') | ((uint32_t)'X' << 8) | ((uint32_t)'D' << 16) |
                             ((uint32_t)'I' << 24);
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
            out[length + digits] = (char)(letters >> (8 * code));
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

// Moves at alignment positions p0 .. p0 + 31 of a tape of `count` moves (p0 a multiple of 32 below count), field k
// (bits 2k, 2k + 1) the move at p0 + k, zero past the end. The alignment takes the tape's order when `forward` (the
// sequences were walked reversed), else its reverse: position p is then tape move count - 1 - p, so the 32 moves
// are a window of the tape ending at count - 1 - p0, its fields put back to front.
__device__ __forceinline__ uint64_t moves_at(const uint64_t *words, int64_t count, int forward, int64_t p0) {
    if (forward) return words[p0 >> 5];  // the last word's fields past `count` are zero
    const int64_t low = count - 32 - p0;  // tape index of position p0 + 31
    uint64_t window;
    if (low < 0) {
        window = words[0] << (2 * -low);  // tape moves below 0 come in as zero fields
    } else {
        const int shift = (int)(low & 31);
        window = words[low >> 5] >> (2 * shift);
        if (shift) window |= words[(low >> 5) + 1] << (64 - 2 * shift);
    }
    const uint64_t reversed = __builtin_bitreverse64(window);
    return ((reversed >> 1) & 0x5555555555555555ull) | ((reversed & 0x5555555555555555ull) << 1);
}

// Bits 0, 2, 4, ... 62 of x packed into 32 bits.
__device__ __forceinline__ uint32_t field_bits(uint64_t x) {
    x &= 0x5555555555555555ull;
    x = (x | (x >> 1)) & 0x3333333333333333ull;
    x = (x | (x >> 2)) & 0x0F0F0F0F0F0F0F0Full;
    x = (x | (x >> 4)) & 0x00FF00FF00FF00FFull;
    x = (x | (x >> 8)) & 0x0000FFFF0000FFFFull;
    return (uint32_t)(x | (x >> 16));
}

// The CIGAR text of one tape of `count` moves, made by the whole wave: runs of one move, each its length then its
// letter, in the alignment's order. Returns the text's length (in every lane); writes it to `out` unless null.
// Every lane of the wave must call it with the same arguments.
//
// The tape goes by in chunks of 64 words (2048 moves), lane l holding word l of the chunk. A run starts where a
// move differs from the one before it (a bit of `starts`); it ends at the next start in the lane, else at the first
// start of the next lane that has one. The chunk's last run is carried into the next chunk, where it either goes on
// (the next chunk opens with the same move: lane 0's first run counts the carried moves too) or ends (lane 0 writes
// it before its own runs). Each lane writes its runs' text after the text of the lanes below it.
__device__ int64_t wave_cigar_text(int eqx, int forward, const uint64_t *words, int64_t count, char *out) {
    const int lane = threadIdx.x & 63;
    const uint32_t letters = (uint32_t)(eqx ? '=' : 'M') | ((uint32_t)'X' << 8) | ((uint32_t)'D' << 16) |
                             ((uint32_t)'I' << 24);
    const int64_t total_words = (count + 31) >> 5;
    int64_t written = 0;         // text of the chunks before this one
    int64_t carried = 0;         // moves of the run carried in (0: none)
    uint32_t carried_move = 0;
    uint64_t upcoming = lane < total_words ? moves_at(words, count, forward, (int64_t)lane * 32) : 0;
    for (int64_t base = 0; base < total_words; base += 64) {
        const int64_t p0 = (base + lane) * 32;
        const int64_t left = count - p0;
        const int valid = left <= 0 ? 0 : left >= 32 ? 32 : (int)left;
        const uint64_t v = upcoming;
        if (base + 64 + lane < total_words) upcoming = moves_at(words, count, forward, p0 + 64 * 32);
        const uint64_t before = __shfl_up(v, 1) >> 62;
        const uint64_t differ = v ^ ((v << 2) | before);
        uint32_t starts = field_bits(differ | (differ >> 1));
        starts &= valid >= 32 ? ~0u : (1u << valid) - 1u;
        if (lane == 0) starts |= 1u;
        const int64_t chunk_end = count < (base + 64) * 32 ? count : (base + 64) * 32;
        const bool last_chunk = chunk_end == count;
        const unsigned long long with = __ballot(starts != 0);
        const int last_lane = 63 - __builtin_clzll(with);
        const unsigned long long above = lane == 63 ? 0ull : (with >> (lane + 1)) << (lane + 1);
        const int64_t first = p0 + (starts ? __builtin_ctz(starts) : 0);
        const int64_t next_first = __shfl(first, above ? __builtin_ctzll(above) : lane);
        // The carried run goes on into lane 0's first run when the chunk opens with its move, else ends here.
        const bool goes_on = base > 0 && (uint32_t)(v & 3u) == carried_move;
        const bool ends_here = base > 0 && !goes_on;
        auto runs = [&](char *to) -> int64_t {
            int64_t text = 0;
            auto put = [&](uint32_t move, int64_t moves) {
                const uint32_t run = (uint32_t)moves;
                const int digits = decimal_digits(run);
                if (to) {
                    uint32_t rest = run;
                    for (int i = digits - 1; i >= 0; --i) {
                        to[text + i] = (char)('0' + rest % 10u);
                        rest /= 10u;
                    }
                    to[text + digits] = (char)(letters >> (8 * move));
                }
                text += digits + 1;
            };
            if (lane == 0 && ends_here) put(carried_move, carried);
            uint32_t rest = starts;
            while (rest) {
                const int k = __builtin_ctz(rest);
                rest &= rest - 1;
                int64_t end;
                if (rest) {
                    end = p0 + __builtin_ctz(rest);
                } else if (lane != last_lane) {
                    end = next_first;
                } else if (!last_chunk) {
                    break;  // the chunk's last run: carried
                } else {
                    end = chunk_end;
                }
                put((uint32_t)(v >> (2 * k)) & 3u, end - (p0 + k) + (lane == 0 && k == 0 && goes_on ? carried : 0));
            }
            return text;
        };
        const int64_t text = runs(nullptr);
        int64_t inclusive = text;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int64_t other = __shfl_up(inclusive, d);
            if (lane >= d) inclusive += other;
        }
        if (out) runs(out + written + inclusive - text);
        written += __shfl(inclusive, 63);
        if (!last_chunk) {
            // The chunk's last run: from the last lane's last start to the chunk's end, with the moves carried
            // into it when it is lane 0's first run.
            const int k = starts ? 31 - __builtin_clz(starts) : 0;
            const int64_t from = __shfl(p0 + k, last_lane);
            const uint32_t move = (uint32_t)__shfl((int)((v >> (2 * k)) & 3u), last_lane);
            carried = chunk_end - from + (from == base * 32 && goes_on ? carried : 0);
            carried_move = move;
        }
    }
    return written;
}

// Writes the CIGAR of each lane's tape (`moves` moves at `words`; `moves` below zero for a lane with none) into the
// shared output: the texts' lengths first, then one atomic claims the wave's room (an exclusive prefix sum across
// the lanes) and the texts are written. Long tapes are made one at a time by the whole wave, short ones by their
// own lanes. Every lane of the wave must call it.
__device__ __forceinline__ void publish_wave(int eqx, int forward, const uint64_t *words, int64_t moves,
                                             char *cigars, unsigned long long *used, int64_t *offset,
                                             int32_t *length) {
    const int lane = threadIdx.x & 63;
    // Short tapes (none past one 2048-move step): each lane makes its own text.
    int64_t longest = moves;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int64_t other = __shfl_xor(longest, d);
        longest = other > longest ? other : longest;
    }
    if (longest <= 2048) {
        const int64_t text = moves >= 0 ? lane_cigar_text(eqx, forward, words, moves, nullptr) : 0;
        const unsigned long long at = wave_claim(used, (unsigned long long)text);
        if (moves >= 0) {
            lane_cigar_text(eqx, forward, words, moves, cigars + at);
            *offset = (int64_t)at;
            *length = (int32_t)text;
        }
        return;
    }
    const unsigned long long walked = __ballot(moves >= 0);
    const unsigned long long tape = (unsigned long long)(uintptr_t)words;
    int64_t mine = 0;
    for (unsigned long long rest = walked; rest; rest &= rest - 1) {
        const int from = __builtin_ctzll(rest);
        const int64_t text = wave_cigar_text(eqx, forward, reinterpret_cast<const uint64_t *>(__shfl(tape, from)),
                                             __shfl(moves, from), nullptr);
        if (lane == from) mine = text;
    }
    const unsigned long long at = wave_claim(used, (unsigned long long)mine);
    for (unsigned long long rest = walked; rest; rest &= rest - 1) {
        const int from = __builtin_ctzll(rest);
        wave_cigar_text(eqx, forward, reinterpret_cast<const uint64_t *>(__shfl(tape, from)), __shfl(moves, from),
                        cigars + __shfl(at, from));
    }
    if (moves >= 0) {
        *offset = (int64_t)at;
        *length = (int32_t)mine;
    }
}

}  // namespace dinara
