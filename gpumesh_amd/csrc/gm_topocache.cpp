// gm_topocache.cpp -- shared, size-bounded cache of built topologies.
//
// Meshes that differ only in vertex positions (the same connectivity and the
// same fixed flags) have the same patch lists, neighbour lists, colouring and
// renumbering: everything NodeGroups::build and Mesh::compileTopology derive
// depends on connectivity alone.  A serving stream of jittered cube meshes
// repeats a few hundred connectivities, so the host-side build (the costliest
// host step of admitting a mesh) runs once per connectivity.  Lookups compare
// the full input arrays, never a hash alone, so a hit is always exact.
#include <algorithm>
#include <cstring>
#include <list>
#include <memory>
#include <mutex>

#include "gm_topology.h"

namespace gmt
{

namespace
{

uint64_t mix(uint64_t h, uint64_t w)
{
    h ^= w * 0x9E3779B97F4A7C15ull;
    h = (h << 27) | (h >> 37);
    return h * 0xC2B2AE3D27D4EB4Full + 0x165667B19E3779F9ull;
}

// Four independent lanes over 32-byte strides, so the multiply chains
// overlap (a mesh of 6M tets hashes 96 MB per lookup).
uint64_t hashBytes(uint64_t h, const void* data, size_t n)
{
    const unsigned char* p = (const unsigned char*)data;
    uint64_t l[4] = {h, h ^ 0x9E3779B97F4A7C15ull, h ^ 0xC2B2AE3D27D4EB4Full, h ^ 0x165667B19E3779F9ull};
    size_t i = 0;
    for(; i + 32 <= n; i += 32)
        for(int k = 0; k < 4; ++k)
        {
            uint64_t w;
            std::memcpy(&w, p + i + 8 * k, 8);
            l[k] = mix(l[k], w);
        }
    for(; i + 8 <= n; i += 8)
    {
        uint64_t w;
        std::memcpy(&w, p + i, 8);
        l[0] = mix(l[0], w);
    }
    uint64_t tail = 0;
    std::memcpy(&tail, p + i, n - i);
    return mix(mix(mix(mix(l[0], l[1]), l[2]), l[3]), tail ^ (uint64_t)n);
}

// The inputs a cached topology was built from (compared in full on lookup,
// outside the cache's lock), and the topology.
struct Key
{
    uint64_t key;
    int64_t nverts, ntets, nhexs, npris;
    std::vector<uint8_t> fixed;
    std::vector<int32_t> cons;      // empty without sliding constraints
    std::vector<uint32_t> tets, hexs, pris;
    std::shared_ptr<const Topology> topo;
};

struct Entry
{
    std::shared_ptr<const Key> k;
    size_t bytes;
};

std::mutex g_mutex;
std::list<Entry> g_entries;     // most recently used first
size_t g_capacity = 0, g_bytes = 0;
int64_t g_hits = 0, g_misses = 0;

size_t topologyBytes(const Topology& t)
{
    // The device layout built from it on the first upload (slots twice,
    // patch, neighbour and element arrays, fixed flags, corner slots) is
    // counted with it.
    const size_t layout = 2 * t.fixed.size() + 4 * (2 * t.fixed.size() + t.eoff.size() + 2 * t.eref.size() +
                                                   t.noff.size() + t.nbr.size() + t.tets.size() + t.hexs.size());
    return 2 * t.fixed.size() + 4 * (t.tets.size() + t.hexs.size() + t.eoff.size() + t.eref.size() + t.noff.size() +
                                     t.nbr.size() + t.groupNodes.size() + t.serialGroup.size() + t.cons.size()) +
           8 * (t.permutation.size() + t.groupOff.size() + t.runOff.size()) +
           4 * (t.groupMaxPatch.size() + t.runMaxPatch.size()) + layout;
}

void evictTo(size_t cap)
{
    while(g_bytes > cap && !g_entries.empty())
    {
        g_bytes -= g_entries.back().bytes;
        g_entries.pop_back();
    }
}

} // namespace

void setCacheCapacity(size_t bytes)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    g_capacity = bytes;
    evictTo(bytes);
}

size_t cacheCapacity()
{
    std::lock_guard<std::mutex> lock(g_mutex);
    return g_capacity;
}

bool isCached(const Topology* t)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    for(const Entry& e : g_entries)
        if(e.k->topo.get() == t) return true;
    return false;
}

void cacheStats(int64_t* out4)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    out4[0] = g_hits;
    out4[1] = g_misses;
    out4[2] = (int64_t)g_entries.size();
    out4[3] = (int64_t)g_bytes;
}

std::shared_ptr<const Topology> buildShared(int64_t nverts, const double* pos, const int32_t* fixed,
                                            const int32_t* cons, int64_t ntets, const uint32_t* tets, int64_t nhexs,
                                            const uint32_t* hexs, int64_t npris, const uint32_t* pris,
                                            std::vector<double>& renumbered, std::string& err)
{
    std::vector<uint8_t> fx((size_t)std::max<int64_t>(nverts, 0));
    for(int64_t v = 0; v < nverts; ++v) fx[v] = fixed[v] != 0;
    // The constraint ids take part in the key (they decide vertex types and
    // hence the numbering); a mesh without sliding vertices keys on none.
    std::vector<int32_t> cs;
    if(cons)
        for(int64_t v = 0; v < nverts; ++v)
            if(cons[v] >= 0 && !fixed[v])
            {
                cs.assign(cons, cons + nverts);
                break;
            }
    size_t cap;
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        cap = g_capacity;
    }
    uint64_t key = 0;
    if(cap > 0 && nverts > 0)
    {
        key = hashBytes(0x243F6A8885A308D3ull, &nverts, sizeof(nverts));
        key = hashBytes(key, fx.data(), fx.size());
        key = hashBytes(key, cs.data(), 4 * cs.size());
        key = hashBytes(key, tets, 16 * (size_t)ntets);
        key = hashBytes(key, hexs, 32 * (size_t)nhexs);
        key = hashBytes(key, pris, 24 * (size_t)npris);
        // Candidates with this hash, taken under the lock; their inputs are
        // immutable and kept alive by the shared pointers, so the full
        // comparison runs without it.
        std::vector<std::shared_ptr<const Key>> cand;
        {
            std::lock_guard<std::mutex> lock(g_mutex);
            for(const Entry& e : g_entries)
                if(e.k->key == key && e.k->nverts == nverts && e.k->ntets == ntets && e.k->nhexs == nhexs &&
                   e.k->npris == npris)
                    cand.push_back(e.k);
        }
        std::shared_ptr<const Key> match;
        for(const auto& k : cand)
        {
            if(k->fixed != fx || k->cons != cs) continue;
            if(ntets && std::memcmp(k->tets.data(), tets, 16 * (size_t)ntets) != 0) continue;
            if(nhexs && std::memcmp(k->hexs.data(), hexs, 32 * (size_t)nhexs) != 0) continue;
            if(npris && std::memcmp(k->pris.data(), pris, 24 * (size_t)npris) != 0) continue;
            match = k;
            break;
        }
        std::shared_ptr<const Topology> hit;
        if(match)
        {
            hit = match->topo;
            std::lock_guard<std::mutex> lock(g_mutex);
            ++g_hits;
            for(auto it = g_entries.begin(); it != g_entries.end(); ++it)
                if(it->k == match)
                {
                    g_entries.splice(g_entries.begin(), g_entries, it);   // most recently used
                    break;
                }
        }
        if(hit)
        {
            renumbered.resize(3 * (size_t)nverts);
            for(int64_t v = 0; v < nverts; ++v)
            {
                const int64_t p = hit->permutation[v];
                renumbered[3 * p] = pos[3 * v];
                renumbered[3 * p + 1] = pos[3 * v + 1];
                renumbered[3 * p + 2] = pos[3 * v + 2];
            }
            return hit;
        }
    }
    auto t = std::make_shared<Topology>();
    if(!build(nverts, pos, fixed, cs.empty() ? nullptr : cs.data(), ntets, tets, nhexs, hexs, npris, pris, *t, err))
        return nullptr;
    renumbered = std::move(t->pos);
    t->pos.clear();
    t->pos.shrink_to_fit();
    std::shared_ptr<const Topology> ct = t;
    if(cap > 0 && nverts > 0)
    {
        auto k = std::make_shared<Key>();
        k->key = key;
        k->nverts = nverts;
        k->ntets = ntets;
        k->nhexs = nhexs;
        k->npris = npris;
        k->fixed = std::move(fx);
        k->cons = std::move(cs);
        k->tets.assign(tets, tets + 4 * (size_t)ntets);
        k->hexs.assign(hexs, hexs + 8 * (size_t)nhexs);
        k->pris.assign(pris, pris + 6 * (size_t)npris);
        k->topo = ct;
        Entry e;
        e.bytes = topologyBytes(*ct) + k->fixed.size() +
                  4 * (k->cons.size() + k->tets.size() + k->hexs.size() + k->pris.size());
        e.k = std::move(k);
        std::lock_guard<std::mutex> lock(g_mutex);
        ++g_misses;
        if(e.bytes <= g_capacity)
        {
            g_bytes += e.bytes;
            g_entries.push_front(std::move(e));
            evictTo(g_capacity);
        }
    }
    return ct;
}

} // namespace gmt
