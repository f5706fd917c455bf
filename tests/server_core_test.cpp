// Host test of quicktex_amd/csrc/server_core.h: the admission arena against a
// byte-map model under random alloc/free, and the FIFO admission rule.
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <random>
#include <vector>

#include "server_core.h"

#define CHECK(c)                                                        \
    do {                                                                \
        if (!(c)) {                                                     \
            std::printf("FAIL line %d: %s\n", __LINE__, #c);            \
            return 1;                                                   \
        }                                                               \
    } while (0)

int main() {
    using qtx::srv::Arena;
    std::mt19937_64 rng(12345);
    for (int trial = 0; trial < 50; trial++) {
        const size_t cap = 256 * (64 + rng() % 4000);
        Arena a(cap, 256);
        std::vector<unsigned char> owner(cap / 256, 0);  // 0 free, else allocation id
        struct Live { size_t off, n; unsigned char id; };
        std::vector<Live> live;
        unsigned char next_id = 1;
        for (int step = 0; step < 4000; step++) {
            if (live.empty() || rng() % 3) {
                const size_t n = 1 + rng() % (cap / 8);
                size_t off = 0;
                // model: first fit over the byte map
                const size_t units = (n + 255) / 256;
                long first = -1;
                for (size_t u = 0, run = 0; u < owner.size(); u++) {
                    run = owner[u] ? 0 : run + 1;
                    if (run == units) { first = (long)(u + 1 - units); break; }
                }
                const bool ok = a.alloc(n, off);
                CHECK(ok == (first >= 0));
                if (ok) {
                    CHECK(off == (size_t)first * 256);
                    CHECK(off % 256 == 0 && off + n <= cap);
                    for (size_t u = off / 256; u < off / 256 + units; u++) { CHECK(owner[u] == 0); owner[u] = next_id; }
                    live.push_back({off, n, next_id});
                    next_id = (unsigned char)(next_id % 250 + 1);
                }
            } else {
                const size_t k = rng() % live.size();
                const Live l = live[k];
                a.release(l.off, l.n);
                for (size_t u = l.off / 256; u < l.off / 256 + (l.n + 255) / 256; u++) { CHECK(owner[u] == l.id); owner[u] = 0; }
                live.erase(live.begin() + (long)k);
            }
            size_t used = 0;
            for (auto &l : live) used += (l.n + 255) / 256 * 256;
            CHECK(a.used() == used);
            CHECK(a.peak() >= a.used());
        }
        for (auto &l : live) a.release(l.off, l.n);
        CHECK(a.used() == 0);
        CHECK(a.free_runs() == 1 && a.largest_free() == a.capacity());  // fully coalesced
    }
    // reach of a failed batch: device-side faults and a lost device end the server,
    // refused calls only their batch (hipError_t values)
    {
        using qtx::srv::FaultScope;
        using qtx::srv::fault_scope;
        for (int code : {3, 4, 53, 100, 101, 201, 214, 700, 702, 709, 710, 719, 999})
            CHECK(fault_scope(code) == FaultScope::Server);
        for (int code : {1, 2, 9, 13, 17, 98, 400, 401, 701, 720, 801, 900, 901})
            CHECK(fault_scope(code) == FaultScope::Batch);
        std::deque<int> w = {5, 7, 9, 7};
        CHECK(qtx::srv::remove_waiting(w, 7) && w.size() == 3 && w[0] == 5 && w[1] == 9 && w[2] == 7);
        CHECK(!qtx::srv::remove_waiting(w, 4) && w.size() == 3);
    }
    // FIFO admission: stops at the first request that does not fit
    {
        Arena a(256 * 10, 256);
        std::deque<size_t> waiting = {512, 1024, 2048, 256}, taken;
        std::vector<size_t> offs;
        auto reserve = [&](size_t &n) { size_t off; if (!a.alloc(n, off)) return false; offs.push_back(off); return true; };
        size_t k = qtx::srv::admit_fifo(waiting, 16, reserve, taken);
        CHECK(k == 2 && taken.size() == 2 && waiting.size() == 2 && waiting.front() == 2048);  // 256 does not overtake
        k = qtx::srv::admit_fifo(waiting, 16, reserve, taken);
        CHECK(k == 0);
        a.release(offs[0], 512);
        a.release(offs[1], 1024);
        k = qtx::srv::admit_fifo(waiting, 1, reserve, taken);
        CHECK(k == 1 && waiting.size() == 1);  // max_take honoured
        k = qtx::srv::admit_fifo(waiting, 16, reserve, taken);
        CHECK(k == 1 && waiting.empty());
    }
    std::printf("server_core ok\n");
    return 0;
}
