// Per-kernel timing: a start and a stop event on the launch stream around each
// instrumented launch; durations are summed per kernel name when read. Events come
// from a pool, so steady-state timing creates none.
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"

namespace hasty_hip {

namespace {

struct Pending {
    const char* name;
    hipEvent_t a, b;
};
struct Stat {
    double ms = 0.0;
    int64_t n = 0;
};

std::mutex g_mu;
bool g_on = false;
std::vector<hipEvent_t> g_pool;
std::vector<Pending> g_pending;
std::map<std::string, Stat> g_stats;

hipEvent_t take_event()
{
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    HASTY_HIP_CHECK(hipEventCreate(&e));
    return e;
}

// Folds finished pairs into the statistics; with wait, waits for every pair.
void collect(bool wait)
{
    size_t keep = 0;
    for (size_t i = 0; i < g_pending.size(); ++i) {
        Pending& p = g_pending[i];
        if (!wait && hipEventQuery(p.b) != hipSuccess) {
            g_pending[keep++] = p;
            continue;
        }
        HASTY_HIP_CHECK(hipEventSynchronize(p.b));
        float ms = 0.f;
        HASTY_HIP_CHECK(hipEventElapsedTime(&ms, p.a, p.b));
        Stat& st = g_stats[p.name];
        st.ms += ms;
        st.n += 1;
        g_pool.push_back(p.a);
        g_pool.push_back(p.b);
    }
    g_pending.resize(keep);
}

}  // namespace

bool timing_on() { return g_on; }

hipEvent_t timing_begin(hipStream_t s)
{
    std::lock_guard<std::mutex> lk(g_mu);
    hipEvent_t e = take_event();
    HASTY_HIP_CHECK(hipEventRecord(e, s));
    return e;
}

void timing_end(const char* name, hipEvent_t start, hipStream_t s)
{
    std::lock_guard<std::mutex> lk(g_mu);
    hipEvent_t e = take_event();
    HASTY_HIP_CHECK(hipEventRecord(e, s));
    g_pending.push_back({name, start, e});
    if (g_pending.size() >= 2048) collect(false);
}

void timing_enable(bool on)
{
    std::lock_guard<std::mutex> lk(g_mu);
    collect(true);
    g_stats.clear();
    g_on = on;
}

int timing_entries()
{
    std::lock_guard<std::mutex> lk(g_mu);
    collect(true);
    return (int)g_stats.size();
}

bool timing_read(int index, std::string* name, double* ms, int64_t* count)
{
    std::lock_guard<std::mutex> lk(g_mu);
    collect(true);
    if (index < 0 || index >= (int)g_stats.size()) return false;
    auto it = g_stats.begin();
    std::advance(it, index);
    *name = it->first;
    *ms = it->second.ms;
    *count = it->second.n;
    return true;
}

}  // namespace hasty_hip
