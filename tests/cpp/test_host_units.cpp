// CPU unit tests of the library's host-only pieces (csrc/host_util.hpp):
// the serving loop's first-fit arena and the autotuner's table parser.
// Built by tests/test_host_units.py with g++ -fsanitize=address,undefined;
// prints "ok <checks>" and exits 0, or names the failed check and exits 1.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <sstream>
#include <vector>

#include "host_util.hpp"

static int g_checks = 0;
#define CHECK(cond)                                                        \
    do {                                                                   \
        ++g_checks;                                                        \
        if (!(cond)) {                                                     \
            std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond); \
            std::exit(1);                                                  \
        }                                                                  \
    } while (0)

using omb::FirstFit;

static void first_fit_basics() {
    FirstFit f;
    f.reset(1000);  // rounded up to whole granules
    CHECK(f.capacity() == 1024);
    size_t a = 1, b = 1, c = 1, d = 1;
    CHECK(f.alloc(1, &a) && a == 0);          // one granule
    CHECK(f.alloc(300, &b) && b == 256);      // two granules
    CHECK(f.alloc(256, &c) && c == 768);
    CHECK(!f.alloc(1, &d));                   // full
    CHECK(f.free_bytes() == 0 && f.free_blocks() == 0);
    f.release(b, 300);                        // hole [256, 768)
    CHECK(f.free_bytes() == 512 && f.free_blocks() == 1);
    CHECK(!f.alloc(513, &d));                 // needs three granules
    CHECK(f.alloc(200, &d) && d == 256);      // first fit: the hole's start
    f.release(a, 1);                          // [0, 256) free, [512, 768) free
    CHECK(f.free_blocks() == 2 && f.free_bytes() == 512);
    f.release(d, 200);                        // [0, 768) coalesces three ways
    CHECK(f.free_blocks() == 1 && f.free_bytes() == 768);
    CHECK(f.alloc(768, &d) && d == 0);
    f.release(d, 768);
    f.release(c, 256);
    CHECK(f.free_blocks() == 1 && f.free_bytes() == 1024);
    // zero bytes still take one granule (distinct offsets per holder)
    size_t z0 = 9, z1 = 9;
    CHECK(f.alloc(0, &z0) && f.alloc(0, &z1) && z0 != z1);
}

// Budget refusal: a request larger than the capacity never fits, and one that
// fits only after a release waits for it.
static void first_fit_budget() {
    FirstFit f;
    f.reset(4096);
    size_t o1 = 0, o2 = 0, o3 = 0;
    CHECK(!f.alloc(4097, &o1));
    CHECK(f.alloc(3000, &o1));
    CHECK(!f.alloc(2000, &o2));
    CHECK(f.alloc(1000, &o3));                // the remaining 1024 bytes
    f.release(o1, 3000);
    CHECK(f.alloc(2000, &o2) && o2 == 0);
}

// Random alloc/release against a byte map of owners: no two live blocks
// overlap, every block lies inside the capacity, and after releasing all of
// them the free list is one block again.
static void first_fit_random() {
    std::mt19937_64 rng(12345);
    FirstFit f;
    const size_t cap = 1 << 20;
    f.reset(cap);
    std::vector<int> owner(cap / FirstFit::kGranule, -1);
    struct Live { size_t off, bytes; };
    std::vector<Live> live;
    for (int step = 0; step < 20000; ++step) {
        const bool do_alloc = live.empty() || (rng() % 3 != 0);
        if (do_alloc) {
            const size_t bytes = 1 + rng() % 40000;
            size_t off = 0;
            if (!f.alloc(bytes, &off)) continue;
            const size_t g0 = off / FirstFit::kGranule;
            const size_t g1 = (off + FirstFit::granules(bytes)) / FirstFit::kGranule;
            CHECK(off % FirstFit::kGranule == 0 && off + bytes <= cap);
            for (size_t g = g0; g < g1; ++g) {
                CHECK(owner[g] < 0);
                owner[g] = step;
            }
            live.push_back({off, bytes});
        } else {
            const size_t i = rng() % live.size();
            const Live l = live[i];
            const size_t g0 = l.off / FirstFit::kGranule;
            const size_t g1 = (l.off + FirstFit::granules(l.bytes)) / FirstFit::kGranule;
            for (size_t g = g0; g < g1; ++g) owner[g] = -1;
            f.release(l.off, l.bytes);
            live[i] = live.back();
            live.pop_back();
        }
        size_t used = 0;
        for (const Live& l : live) used += FirstFit::granules(l.bytes);
        CHECK(used + f.free_bytes() == f.capacity());
    }
    for (const Live& l : live) f.release(l.off, l.bytes);
    CHECK(f.free_blocks() == 1 && f.free_bytes() == cap);
}

static bool parse(const char* text, std::map<omb::TuneKey, int>* out, std::string* err) {
    std::istringstream in(text);
    return omb::parse_tune_table(in, out, err);
}

static void tune_table() {
    std::map<omb::TuneKey, int> t;
    std::string err;
    CHECK(parse("# mode log2M log2N log2K class\n0 4 5 6 f32_32x32\n\n0 7 7 7 f32_128x64\n"
                "1 5 5 5 bf16x3_64x64\n3 6 6 6 bf16x3_64x64\n",
                &t, &err));
    CHECK(t.size() == 4);
    CHECK((t[omb::TuneKey{0, 4, 5, 6}] == omb::C32));
    CHECK((t[omb::TuneKey{0, 7, 7, 7}] == omb::C128x64));
    CHECK((t[omb::TuneKey{3, 6, 6, 6}] == omb::CSPLIT));
    // every malformed table is refused whole: the map is unchanged
    const char* bad[] = {
        "0 4 5 6\n",                      // missing class
        "0 4 5 6 f32_32x32 extra\n",      // extra field
        "x 4 5 6 f32_32x32\n",            // not a number
        "7 4 5 6 f32_32x32\n",            // unknown mode
        "-1 4 5 6 f32_32x32\n",
        "0 4 99 6 f32_32x32\n",           // bucket out of range
        "0 4 5 -2 f32_32x32\n",
        "0 4 5 6 f32_16x16\n",            // unknown class
        "0 4 5 6 gemv_64rows\n",          // GEMV chunks are not GEMM tiles
        "0 4 5 6 bf16x3_64x64\n",         // split tile for plain FP32
        "1 4 5 6 f32_64x64\n",            // plain tile for split-BF16
        "3 4 5 6 f32_32x32\n",            // plain tile for scaled FP16
        "0 1 1 1 f32_64x64\n2 4 5 6 f32_128x64\n",  // one bad line spoils the table
    };
    for (const char* b : bad) {
        std::map<omb::TuneKey, int> before = t;
        err.clear();
        CHECK(!parse(b, &t, &err));
        CHECK(!err.empty());
        CHECK(t.size() == before.size());
    }
    for (int mode = 0; mode < 4; ++mode)
        for (int c = 0; c < omb::NCLS; ++c)
            CHECK(omb::class_serves_mode(c, mode) ==
                  (mode == 0 ? (c == omb::C32 || c == omb::C64 || c == omb::C128x64)
                             : c == omb::CSPLIT));
}

int main() {
    first_fit_basics();
    first_fit_budget();
    first_fit_random();
    tune_table();
    std::printf("ok %d\n", g_checks);
    return 0;
}
