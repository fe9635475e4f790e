// CPU unit test of the serving loop's admission policy (ag_admit.h) with a
// mock byte counter standing in for device allocations: no GPU, no HIP.
// Prints one line per case and exits non-zero on the first failure.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "ag_admit.h"

using ag::AdmitHolder;
using ag::plan_admission;

static int failures = 0;

static void expect(bool ok, const char *what)
{
    std::printf("%s %s\n", ok ? "PASS" : "FAIL", what);
    if (!ok) failures++;
}

int main()
{
    // mock: a batch of `cap` columns reserves 1000 bytes per column + 500
    auto bytes = [](uint32_t cap) { return (uint64_t)cap * 1000 + 500; };

    {   // room for the full capacity: nothing evicted
        auto p = plan_admission(100000, 0, 64, {}, bytes);
        expect(p.cap == 64 && p.evict.empty(), "empty server admits the full capacity");
    }
    {   // budget for 16 columns only: capacity lowered in steps of 8
        auto p = plan_admission(16500, 0, 64, {}, bytes);
        expect(p.cap == 16 && p.evict.empty(), "capacity lowered to what fits");
    }
    {   // below 8 columns: no room, nothing evicted
        auto p = plan_admission(8499, 0, 64, {}, bytes);
        expect(p.cap == 0 && p.evict.empty(), "no room below 8 columns");
    }
    {   // idle holders evicted largest first, only while the full capacity does not fit
        std::vector<AdmitHolder> h = {{10000, true}, {30000, true}, {20000, true}};
        // reserved 60000; need 64500 for 64 columns; budget 100000 -> evict 30000 (held 30000)
        auto p = plan_admission(100000, 60000, 64, h, bytes);
        expect(p.cap == 64 && p.evict.size() == 1 && p.evict[0] == 1,
               "largest idle holder evicted first, just enough");
    }
    {   // busy holders are never evicted; the capacity shrinks instead
        std::vector<AdmitHolder> h = {{50000, false}, {5000, true}};
        // reserved 55000, budget 80000: evict the idle 5000 -> 50000 held, 30000 left -> 24 columns
        auto p = plan_admission(80000, 55000, 64, h, bytes);
        expect(p.cap == 24 && p.evict.size() == 1 && p.evict[0] == 1,
               "busy holder kept, idle one evicted, capacity lowered");
    }
    {   // even evicting every idle holder leaves no room: evict nothing
        std::vector<AdmitHolder> h = {{70000, false}, {2000, true}};
        // after the eviction 78000 - 70000 = 8000 < 8500 (8 columns)
        auto p = plan_admission(78000, 72000, 64, h, bytes);
        expect(p.cap == 0 && p.evict.empty(), "no pointless eviction when nothing fits");
        // with 80000 the same eviction makes room for 8 columns
        auto q = plan_admission(80000, 72000, 64, h, bytes);
        expect(q.cap == 8 && q.evict.size() == 1 && q.evict[0] == 1, "eviction that makes room");
    }
    {   // equal sizes: eviction order follows holder order (stable)
        std::vector<AdmitHolder> h = {{30000, true}, {30000, true}, {30000, true}};
        // held 90000 -> 60000 -> 30000 -> 0: only then do 64 columns (64500) fit in 70000
        auto p = plan_admission(70000, 90000, 64, h, bytes);
        expect(p.cap == 64 && p.evict == std::vector<size_t>({0, 1, 2}),
               "equal holders evicted in holder order");
    }
    {   // want 8 with an exact fit
        auto p = plan_admission(8500, 0, 8, {}, bytes);
        expect(p.cap == 8, "exact fit of 8 columns");
    }
    return failures ? 1 : 0;
}
