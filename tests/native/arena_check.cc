// Host-only checks of the patch server's arena allocator and admission
// order (dengo_amd/csrc/arena.h), built with g++ by tests/test_server_host.py.
// Prints "ok" and exits 0 when every check holds.
#include <stdio.h>
#include <stdlib.h>

#include <random>
#include <vector>

#include "arena.h"

#define CHECK(c)                                                         \
    do {                                                                 \
        if (!(c)) {                                                      \
            fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            exit(1);                                                     \
        }                                                                \
    } while (0)

struct Req {
    size_t bytes;
    int bypassed;
};

int main()
{
    using dengo::Arena;
    // alignment, exhaustion, coalescing
    Arena a(4096);
    CHECK(a.size() == 4096);
    const long long x = a.alloc(1), y = a.alloc(300), z = a.alloc(256);
    CHECK(x == 0 && y == 256 && z == 768);
    CHECK(a.used() == 1024 && a.peak() == 1024);
    CHECK(a.alloc(4096) == -1 && !a.fits(3073) && a.fits(3072));
    a.release(256, 300);           // hole [256, 768)
    CHECK(a.free_runs() == 2 && a.largest_free() == 3072);
    CHECK(a.alloc(512) == 256);    // first fit takes the hole
    a.release(256, 512);
    a.release(0, 1);
    a.release(768, 256);           // everything coalesces back to one run
    CHECK(a.free_runs() == 1 && a.largest_free() == 4096 && a.used() == 0);
    CHECK(a.peak() == 1024);

    // random alloc/free against a byte map: blocks never overlap, stay in
    // bounds, and the free space always adds up
    std::mt19937_64 rng(5);
    Arena b(1 << 20);
    std::vector<char> owner(1 << 20, 0);
    struct Blk { long long off; size_t n; };
    std::vector<Blk> live;
    for (int it = 0; it < 20000; ++it) {
        if (live.empty() || rng() % 3 != 0) {
            const size_t n = 1 + rng() % 20000;
            const long long off = b.alloc(n);
            if (off < 0) {
                CHECK(!b.fits(n));
                continue;
            }
            const size_t r = Arena::round_up(n);
            CHECK(off % Arena::kAlign == 0 && (size_t)off + r <= b.size());
            for (size_t i = 0; i < r; ++i) {
                CHECK(owner[off + i] == 0);
                owner[off + i] = 1;
            }
            live.push_back({off, n});
        } else {
            const size_t k = rng() % live.size();
            const Blk bl = live[k];
            live.erase(live.begin() + k);
            const size_t r = Arena::round_up(bl.n);
            for (size_t i = 0; i < r; ++i) owner[bl.off + i] = 0;
            b.release(bl.off, bl.n);
        }
        size_t used = 0;
        for (const Blk &bl : live) used += Arena::round_up(bl.n);
        CHECK(b.used() == used);
    }
    for (const Blk &bl : live) b.release(bl.off, bl.n);
    CHECK(b.free_runs() == 1 && b.used() == 0);

    // admission: first request that fits, bounded bypass of a large one
    std::deque<Req> w = {{900, 0}, {100, 0}, {100, 0}, {100, 0}};
    size_t avail = 250;
    auto fits = [&](const Req &r) { return r.bytes <= avail; };
    CHECK(dengo::pick_admission(w, fits, 2) == 1);
    CHECK(w[0].bypassed == 1);
    w.erase(w.begin() + 1);
    CHECK(dengo::pick_admission(w, fits, 2) == 1);
    CHECK(w[0].bypassed == 2);
    w.erase(w.begin() + 1);
    CHECK(dengo::pick_admission(w, fits, 2) == -1);   // the large one now blocks
    avail = 1000;
    CHECK(dengo::pick_admission(w, fits, 2) == 0);
    std::deque<Req> e;
    CHECK(dengo::pick_admission(e, fits, 2) == -1);
    printf("ok\n");
    return 0;
}
