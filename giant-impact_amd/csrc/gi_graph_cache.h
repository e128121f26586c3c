// gi_graph_cache.h -- the graph cache's lookup and eviction policy, free of
// device types so that the host alone can exercise it.
//
// Entries are (key, plan generation, payload) with the least recently used
// first.  A lookup hit becomes the most recently used.  A miss captures a new
// payload; when the cache is full the least recently used entry the caller
// does not hold is evicted -- after a device synchronisation, since launches
// of that graph from an earlier, unsynchronised call may still be in flight
// -- and the new entry is appended.  When every entry is held the cache grows
// past its capacity instead (a caller holding several graphs at once asks for
// them one by one and passes those it already has as held), and shrinks back
// at the next miss that holds fewer.
#pragma once

#include <cstddef>
#include <utility>
#include <vector>

namespace gi {

template <typename Payload>
struct GraphCache {
    struct Entry {
        std::vector<unsigned char> key;
        unsigned long long generation = 0;
        Payload payload{};
    };
    std::vector<Entry> entries;  // least recently used first
    size_t capacity = 8;

    // capture(Payload*) -> int, sync() -> int, destroy(Payload&), held(const
    // Payload&) -> bool.  On success *out is a payload that stays valid until
    // the next call of this function that does not hold it.
    template <typename Capture, typename Sync, typename Destroy, typename Held>
    int get(const std::vector<unsigned char>& key, unsigned long long generation, Capture capture, Sync sync,
            Destroy destroy, Held held, Payload* out) {
        auto no_update = [](Payload&, bool* done) {
            *done = false;
            return 0;
        };
        return get(key, generation, capture, no_update, sync, destroy, held, out);
    }

    // As above, and on a miss an unheld entry with the same key but another
    // plan generation is first offered to update(Payload& existing, bool*
    // done) -> int, which may bring it up to date in place (the executable
    // graph kept, its parameters replaced) after a device synchronisation;
    // only when that is not possible is a new payload captured.
    template <typename Capture, typename Update, typename Sync, typename Destroy, typename Held>
    int get(const std::vector<unsigned char>& key, unsigned long long generation, Capture capture, Update update,
            Sync sync, Destroy destroy, Held held, Payload* out) {
        for (size_t k = 0; k < entries.size(); ++k)
            if (entries[k].generation == generation && entries[k].key == key) {
                Entry hit = std::move(entries[k]);
                entries.erase(entries.begin() + (long)k);
                entries.push_back(std::move(hit));
                *out = entries.back().payload;
                return 0;
            }
        int rc;
        for (size_t k = 0; k < entries.size(); ++k)
            if (entries[k].key == key && !held(entries[k].payload)) {
                if ((rc = sync())) return rc;
                bool done = false;
                if ((rc = update(entries[k].payload, &done))) return rc;
                if (!done) break;
                Entry hit = std::move(entries[k]);
                entries.erase(entries.begin() + (long)k);
                hit.generation = generation;
                entries.push_back(std::move(hit));
                *out = entries.back().payload;
                return 0;
            }
        Entry e;
        e.key = key;
        e.generation = generation;
        rc = capture(&e.payload);
        if (rc) return rc;
        // Evict least recently used unheld entries until the new one fits.
        bool synced = false;
        for (size_t k = 0; entries.size() >= capacity && k < entries.size();) {
            if (held(entries[k].payload)) {
                ++k;
                continue;
            }
            if (!synced && (rc = sync())) {
                destroy(e.payload);
                return rc;
            }
            synced = true;
            destroy(entries[k].payload);
            entries.erase(entries.begin() + (long)k);
        }
        entries.push_back(std::move(e));
        *out = entries.back().payload;
        return 0;
    }

    template <typename Destroy>
    void clear(Destroy destroy) {
        for (auto& e : entries) destroy(e.payload);
        entries.clear();
    }
};

}  // namespace gi
