// Host-only check of the graph cache's policy (giant-impact_amd/csrc/gi_graph_cache.h)
// against a simulated device: graphs are launched asynchronously and stay
// "in flight" until the next device synchronisation.  Asserted throughout:
//   - no graph is destroyed while it has launches in flight;
//   - no graph the current call holds is destroyed;
//   - every graph returned to a caller is live, and stays live while held;
//   - the cache never exceeds its capacity unless every entry is held.
// The sequences are those of the steppers: gi_impact_steps asks for its
// one-step graph, then its eight-step graph holding the first, then (with a
// recenter) the recenter graph holding both, and launches them without a
// synchronisation; the sequential ensemble baseline does that for 64
// simulations in turn, each with its own arrays and plan generation.
#include "gi_graph_cache.h"

#include <cstdio>
#include <cstdlib>
#include <initializer_list>
#include <set>
#include <vector>

namespace {

struct Device {
    std::set<int> live, in_flight, destroyed;
    int next_id = 1, syncs = 0, captures = 0;
};

int g_fail = 0;
#define CHECK(c)                                                          \
    do {                                                                  \
        if (!(c)) {                                                       \
            std::fprintf(stderr, "FAILED %s at line %d\n", #c, __LINE__); \
            ++g_fail;                                                     \
        }                                                                 \
    } while (0)

int get(gi::GraphCache<int>& cache, Device& dev, const std::vector<unsigned char>& key, unsigned long long gen,
        std::initializer_list<int> keep) {
    int out = -1;
    auto capture = [&](int* p) {
        *p = dev.next_id++;
        dev.live.insert(*p);
        ++dev.captures;
        return 0;
    };
    auto sync = [&]() {
        dev.in_flight.clear();
        ++dev.syncs;
        return 0;
    };
    auto destroy = [&](int& p) {
        CHECK(dev.in_flight.count(p) == 0);  // nothing of it still runs
        for (int h : keep) CHECK(h != p);    // the caller does not hold it
        CHECK(dev.live.count(p) == 1);
        dev.live.erase(p);
        dev.destroyed.insert(p);
    };
    auto held = [&](const int& p) {
        for (int h : keep)
            if (h == p) return true;
        return false;
    };
    int rc = cache.get(key, gen, capture, sync, destroy, held, &out);
    CHECK(rc == 0);
    CHECK(dev.live.count(out) == 1);
    size_t nheld = 0;
    for (auto& e : cache.entries) nheld += held(e.payload) ? 1 : 0;
    CHECK(cache.entries.size() <= cache.capacity || nheld + 1 >= cache.entries.size());
    return out;
}

std::vector<unsigned char> key_of(int kind, int sim, int steps) {
    std::vector<unsigned char> k;
    for (int v : {kind, sim, steps})
        for (int b = 0; b < 4; ++b) k.push_back((unsigned char)(v >> (8 * b)));
    return k;
}

// One gi_impact_steps call: the one-step graph, the eight-step graph holding
// it, optionally the recenter graph holding both; all launched, no sync.
void impact_steps_call(gi::GraphCache<int>& cache, Device& dev, int sim, unsigned long long gen, bool recenter) {
    const int one = get(cache, dev, key_of(1, sim, 1), gen, {});
    const int many = get(cache, dev, key_of(1, sim, 8), gen, {one});
    int rec = -1;
    if (recenter) rec = get(cache, dev, key_of(2, sim, 0), gen, {one, many});
    for (int g : {one, many, rec})
        if (g >= 0) {
            CHECK(dev.live.count(g) == 1);  // still live when launched
            dev.in_flight.insert(g);
        }
}

}  // namespace

int main() {
    {
        // the sequential baseline of the ensemble leg: 64 simulations, each
        // re-planned (a new generation) and run through the stepper twice
        gi::GraphCache<int> cache;
        cache.capacity = 8;
        Device dev;
        unsigned long long gen = 0;
        for (int rep = 0; rep < 3; ++rep)
            for (int s = 0; s < 64; ++s) {
                ++gen;
                impact_steps_call(cache, dev, s, gen, s % 3 == 0);
                impact_steps_call(cache, dev, s, gen, s % 3 == 0);  // hits
            }
        CHECK(cache.entries.size() <= 8);
        CHECK(dev.syncs > 0);
        // every capture beyond the capacity was paired with one destruction
        CHECK((int)dev.live.size() == (int)cache.entries.size());
        CHECK(dev.captures == (int)(dev.live.size() + dev.destroyed.size()));
    }
    {
        // a hit moves to the back: the least recently used is evicted first
        gi::GraphCache<int> cache;
        cache.capacity = 2;
        Device dev;
        const int a = get(cache, dev, key_of(9, 0, 0), 1, {});
        const int b = get(cache, dev, key_of(9, 1, 0), 1, {});
        CHECK(get(cache, dev, key_of(9, 0, 0), 1, {}) == a);  // hit, a is now the most recent
        const int c = get(cache, dev, key_of(9, 2, 0), 1, {});
        CHECK(dev.destroyed.count(b) == 1 && dev.live.count(a) == 1 && dev.live.count(c) == 1);
        // a different generation of the same key is a different graph
        const int a2 = get(cache, dev, key_of(9, 0, 0), 2, {c});
        CHECK(a2 != a && dev.destroyed.count(a) == 1);
    }
    {
        // every entry held: the cache grows instead of evicting
        gi::GraphCache<int> cache;
        cache.capacity = 2;
        Device dev;
        const int a = get(cache, dev, key_of(7, 0, 0), 1, {});
        const int b = get(cache, dev, key_of(7, 1, 0), 1, {a});
        const int c = get(cache, dev, key_of(7, 2, 0), 1, {a, b});
        CHECK(cache.entries.size() == 3 && dev.destroyed.empty());
        dev.in_flight = {a, b, c};
        get(cache, dev, key_of(7, 3, 0), 1, {c});  // evicts a after a sync, never c
        CHECK(dev.destroyed.count(a) == 1 && dev.live.count(c) == 1);
    }
    if (g_fail) {
        std::fprintf(stderr, "%d check(s) failed\n", g_fail);
        return 1;
    }
    std::printf("graph cache policy: all checks passed\n");
    return 0;
}
