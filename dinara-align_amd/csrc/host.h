// Host-side helpers the batch calls share.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <functional>
#include <thread>
#include <vector>

namespace dinara {

// Runs body(lo, hi) over [0, count) cut into `threads` contiguous shares, each on a thread of its own; one share
// on the caller's thread when `threads` is one or less.
inline void parallel_for(int64_t count, int64_t threads, const std::function<void(int64_t, int64_t)> &body) {
    const int64_t workers = std::max<int64_t>(1, std::min<int64_t>(threads, count));
    if (workers <= 1) {
        body(0, count);
        return;
    }
    std::vector<std::thread> pool;
    for (int64_t w = 0; w < workers; ++w)
        pool.emplace_back([&, w]() { body(count * w / workers, count * (w + 1) / workers); });
    for (auto &t : pool) t.join();
}

// Byte classes for packing letters: A, C, G and T carry their 2-bit codes (0, 1, 2, 3) in bits 0-1 and a bit
// each of their own, every other byte the OTHER bit.
struct ByteClasses {
    static constexpr uint8_t OTHER = 64;
    uint8_t cls[256];
    ByteClasses() {
        for (int i = 0; i < 256; ++i) cls[i] = OTHER;
        cls['A'] = 0 | 4;
        cls['C'] = 1 | 8;
        cls['G'] = 2 | 16;
        cls['T'] = 3 | 32;
    }
    // The union of the classes of `len` letters: OTHER set when any is not A, C, G or T.
    uint32_t seen(const char *text, int64_t len) const {
        const unsigned char *u = reinterpret_cast<const unsigned char *>(text);
        uint32_t acc = 0;
        for (int64_t i = 0; i < len; ++i) acc |= cls[u[i]];
        return acc;
    }
    // Packs `len` letters of A, C, G and T, 16 codes to a 32-bit word (letter i at bits 2i), word w at out[64 w],
    // `words` words in all (zero past the letters); back to front when `reversed`.
    void pack2(const char *text, int64_t len, bool reversed, uint32_t *out, int32_t words) const {
        const unsigned char *u = reinterpret_cast<const unsigned char *>(text);
        const int64_t whole = len / 16;
        for (int64_t w = 0; w < whole; ++w) {
            uint32_t word = 0;
            if (reversed) {
                const unsigned char *at = u + len - 1 - 16 * w;
                for (int i = 0; i < 16; ++i) word |= (uint32_t)(cls[at[-i]] & 3) << (2 * i);
            } else {
                const unsigned char *at = u + 16 * w;
                for (int i = 0; i < 16; ++i) word |= (uint32_t)(cls[at[i]] & 3) << (2 * i);
            }
            out[w * 64] = word;
        }
        int64_t w = whole;
        if (w < words && 16 * w < len) {
            uint32_t word = 0;
            for (int64_t i = 0; 16 * w + i < len; ++i) {
                const unsigned char byte = reversed ? u[len - 1 - (16 * w + i)] : u[16 * w + i];
                word |= (uint32_t)(cls[byte] & 3) << (2 * i);
            }
            out[w * 64] = word;
            ++w;
        }
        for (; w < words; ++w) out[w * 64] = 0;
    }
};

// The decisions of one sweep launch may take at most this much device memory; a larger launch is cut in pieces.
constexpr size_t ARENA_BUDGET = size_t(16) << 30;

}  // namespace dinara
