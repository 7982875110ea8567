// Host-only pieces with no device code: the serving loop's first-fit free
// list and the grouped autotuner's table format.  Free of HIP headers, so a
// CPU unit test builds them with AddressSanitizer and UBSan.
#pragma once

#include <cstddef>
#include <cstdint>
#include <istream>
#include <iterator>
#include <list>
#include <map>
#include <sstream>
#include <string>
#include <tuple>

namespace omb {

// First-fit allocator over the offsets [0, capacity) of one buffer, in
// 256-byte granules.  Free blocks are kept sorted by offset and adjacent
// free blocks are merged on release.
class FirstFit {
   public:
    static constexpr size_t kGranule = 256;
    static size_t granules(size_t bytes) {
        const size_t b = bytes == 0 ? 1 : bytes;
        return (b + kGranule - 1) / kGranule * kGranule;
    }
    void reset(size_t capacity) {
        cap_ = granules(capacity);
        free_.clear();
        free_.push_back({0, cap_});
    }
    size_t capacity() const { return cap_; }
    // The lowest free offset with room for `bytes` (rounded to granules).
    bool alloc(size_t bytes, size_t* off) {
        bytes = granules(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->size >= bytes) {
                *off = it->off;
                it->off += bytes;
                it->size -= bytes;
                if (it->size == 0) free_.erase(it);
                return true;
            }
        }
        return false;
    }
    // Return a block obtained from alloc (same offset and byte count).
    void release(size_t off, size_t bytes) {
        bytes = granules(bytes);
        auto it = free_.begin();
        while (it != free_.end() && it->off < off) ++it;
        it = free_.insert(it, {off, bytes});
        auto next = std::next(it);
        if (next != free_.end() && it->off + it->size == next->off) {
            it->size += next->size;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->off + prev->size == it->off) {
                prev->size += it->size;
                free_.erase(it);
            }
        }
    }
    size_t free_bytes() const {
        size_t s = 0;
        for (const auto& b : free_) s += b.size;
        return s;
    }
    size_t free_blocks() const { return free_.size(); }

   private:
    struct Block {
        size_t off, size;
    };
    size_t cap_ = 0;
    std::list<Block> free_;
};

// Tile classes of grouped plans: plain FP32 tiles (f32_*), the split tile
// shared by every split mode, and the GEMV row chunk (never in a table).
enum Cls : int { C32 = 0, C64 = 1, C128x64 = 2, CSPLIT = 3, CGEMV = 4, NCLS = 5 };
constexpr const char* kClsName[NCLS] = {"f32_32x32", "f32_64x64", "f32_128x64",
                                        "bf16x3_64x64", "gemv_64rows"};

// Whether a GEMM tile class computes the arithmetic of `mode` (modes 0-3:
// plain FP32, split-BF16 with 6 or 3 products, scaled split-FP16).
inline bool class_serves_mode(int cls, int mode) {
    if (mode == 0) return cls == C32 || cls == C64 || cls == C128x64;
    if (mode >= 1 && mode <= 3) return cls == CSPLIT;
    return false;
}

// Autotuner table key: mode and the (log2 M, log2 N, log2 K) bucket.
struct TuneKey {
    int mode, bm, bn, bk;
    bool operator<(const TuneKey& o) const {
        return std::tie(mode, bm, bn, bk) < std::tie(o.mode, o.bm, o.bn, o.bk);
    }
};

// Parse the autotuner's text table: one "mode log2M log2N log2K class" line
// per bucket; blank lines and lines starting with '#' are skipped.  A line
// with missing or extra fields, a mode outside 0-3, a bucket outside
// [0, 62], an unknown class or a class that does not compute the line's
// mode is an error: false, with the offending line in *err, and *out
// unchanged.
inline bool parse_tune_table(std::istream& in, std::map<TuneKey, int>* out, std::string* err) {
    std::map<TuneKey, int> loaded;
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream is(line);
        TuneKey k{};
        std::string name, extra;
        if (!(is >> k.mode >> k.bm >> k.bn >> k.bk >> name) || (is >> extra)) {
            *err = "malformed line: " + line;
            return false;
        }
        if (k.mode < 0 || k.mode > 3) {
            *err = "unknown mode in line: " + line;
            return false;
        }
        if (k.bm < 0 || k.bm > 62 || k.bn < 0 || k.bn > 62 || k.bk < 0 || k.bk > 62) {
            *err = "bucket out of range in line: " + line;
            return false;
        }
        int cls = -1;
        for (int c = 0; c < NCLS; ++c)
            if (name == kClsName[c]) cls = c;
        if (cls < 0 || cls == CGEMV) {
            *err = "unknown class " + name;
            return false;
        }
        if (!class_serves_mode(cls, k.mode)) {
            *err = "class " + name + " does not compute mode " + std::to_string(k.mode);
            return false;
        }
        loaded[k] = cls;
    }
    for (auto& kv : loaded) (*out)[kv.first] = kv.second;
    return true;
}

}  // namespace omb
