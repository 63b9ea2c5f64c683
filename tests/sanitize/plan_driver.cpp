// plan_driver.cpp -- drives the host range planner (flowvpm.jl_amd/csrc/
// range_plan.cpp) on its own, with the device copies replaced by host
// memcpy, so that it can run under AddressSanitizer + UndefinedBehavior-
// Sanitizer and under ThreadSanitizer on the CPU.  Each case checks the plan
// against a brute-force reading of the list; exit status 0 and a last line
// "plan_driver OK" mean every case passed.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "range_plan.h"
#include "flowvpm_hip.h"

namespace vpm {
thread_local std::string g_error;
int fail(int status, const std::string& msg)
{
    g_error = msg;
    return status;
}
}  // namespace vpm

using namespace vpm;

// Host stand-in of the field's pinned buffer and device plan buffer.
struct HostSink final : PlanSink {
    std::vector<char> host, dev;
    int64_t nitems = 0, nsrc = 0;
    int64_t fail_at = -1;     // copy_piece of the piece holding this pair fails
    size_t items_bytes() const { return sizeof(RangeItem) * (size_t)nitems; }
    size_t bytes() const { return items_bytes() + sizeof(SrcRange) * (size_t)nsrc; }
    int stage(int64_t ni, int64_t ns, RangeItem*& items, SrcRange*& src) override
    {
        nitems = ni;
        nsrc = ns;
        host.assign(bytes() + 64, 0);
        items = (RangeItem*)host.data();
        src = (SrcRange*)(host.data() + items_bytes());
        return VPM_OK;
    }
    int prepare_copies(int64_t, int64_t) override
    {
        dev.assign(bytes() + 64, 0);
        return VPM_OK;
    }
    int copy_piece(int64_t a, int64_t b) override
    {
        if (a <= fail_at && fail_at < b) return 7;
        const size_t off = items_bytes() + sizeof(SrcRange) * (size_t)a;
        std::memcpy(dev.data() + off, host.data() + off, sizeof(SrcRange) * (size_t)(b - a));
        return 0;
    }
    int report(int code) override { return fail(VPM_EDEVICE, "copy failed with " + std::to_string(code)); }
    const RangeItem* items() const { return (const RangeItem*)host.data(); }
    const SrcRange* src(bool copied) const
    {
        return (const SrcRange*)((copied ? dev.data() : host.data()) + items_bytes());
    }
};

static int g_failures = 0;
#define CHECK(c, ...)                                                   \
    do {                                                                \
        if (!(c)) {                                                     \
            std::printf("FAIL %s:%d: ", __FILE__, __LINE__);            \
            std::printf(__VA_ARGS__);                                   \
            std::printf("\n");                                          \
            ++g_failures;                                               \
            return;                                                     \
        }                                                               \
    } while (0)

struct List { std::vector<int64_t> tb, te, sb, se; };

// For every target, the source ranges of the pairs covering it, in pair
// order, from the plan and from the list itself; they must be equal.
static void check_plan(const char* name, const List& L, int64_t np, const HostSink& s, const PlanShape& sh)
{
    const int64_t n = (int64_t)L.tb.size();
    std::vector<std::vector<std::pair<int32_t, int32_t>>> want(np), got(np);
    for (int64_t k = 0; k < n; ++k)
        if (L.sb[k] < L.se[k])
            for (int64_t t = L.tb[k]; t < L.te[k]; ++t) want[t].push_back({(int32_t)L.sb[k], (int32_t)L.se[k]});
    const RangeItem* it = s.items();
    const SrcRange* src = s.src(sh.src_copied);
    const int64_t chunk = sh.wave ? 64 / sh.R : 256;
    std::vector<int> owner(np, 0);
    for (int64_t m = 0; m < sh.nitems; ++m) {
        CHECK(it[m].t_begin >= 0 && it[m].t_begin < it[m].t_end && it[m].t_end <= np, "%s item %ld targets", name,
              (long)m);
        CHECK(it[m].t_end - it[m].t_begin <= chunk, "%s item %ld longer than %ld", name, (long)m, (long)chunk);
        CHECK(it[m].l_begin >= 0 && it[m].l_begin <= it[m].l_end && it[m].l_end <= sh.nsrc, "%s item %ld list",
              name, (long)m);
        for (int64_t t = it[m].t_begin; t < it[m].t_end; ++t) {
            ++owner[t];
            for (int64_t l = it[m].l_begin; l < it[m].l_end; ++l)
                if (src[l].b < src[l].e) got[t].push_back({src[l].b, src[l].e});
        }
    }
    for (int64_t t = 0; t < np; ++t) {
        CHECK(owner[t] <= 1, "%s target %ld in %d items", name, (long)t, owner[t]);
        CHECK(got[t] == want[t], "%s target %ld: %zu source ranges planned, %zu listed", name, (long)t,
              got[t].size(), want[t].size());
    }
}

static List leaf_list(int64_t np, int64_t leaf, int rep, bool shuffle, unsigned seed)
{
    List L;
    std::vector<int64_t> starts;
    for (int64_t a = 0; a < np; a += leaf) starts.push_back(a);
    std::vector<int64_t> order(starts.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (int64_t)i;
    if (shuffle) std::shuffle(order.begin(), order.end(), std::mt19937(seed));
    for (int64_t i : order)
        for (int r = 0; r < rep; ++r)
            for (int64_t a : starts) {
                L.tb.push_back(starts[i]);
                L.te.push_back(std::min(starts[i] + leaf, np));
                L.sb.push_back(a);
                L.se.push_back(std::min(a + leaf, np));
            }
    return L;
}

static int run(const List& L, int64_t np, bool wave, HostSink& s, PlanShape& sh)
{
    return plan_ranges(np, L.tb.data(), L.te.data(), L.sb.data(), L.se.data(), (int64_t)L.tb.size(), wave, 8, s,
                       sh);
}

static void case_long_leaf_list()
{
    // 1024 particles in 4-particle leaves, every leaf against every leaf 40
    // times over: 2.6M pairs, eight planner threads
    const int64_t np = 1024;
    List L = leaf_list(np, 4, 40, false, 0);
    HostSink s;
    PlanShape sh;
    CHECK(run(L, np, true, s, sh) == VPM_OK, "long leaf list: %s", g_error.c_str());
    CHECK(sh.wave && sh.R == 8 && sh.src_copied, "long leaf list shape wave=%d R=%d", sh.wave, sh.R);
    CHECK(sh.nsrc == (int64_t)L.tb.size() && sh.nitems == np / 4, "long leaf list counts");
    check_plan("long leaf list", L, np, s, sh);
}

static void case_errors_lowest_index()
{
    const int64_t np = 1024;
    List base = leaf_list(np, 4, 40, false, 0);
    struct Bad { int col; int64_t k, v; };
    const std::vector<std::vector<Bad>> cases = {
        {{3, 2300000, -1}, {3, 1100000, np + 1}},                    // source ranges, two chunks
        {{1, 2000000, np + 3}, {0, 1500000, -2}, {2, 500000, -1}},   // targets are checked first
    };
    const int64_t first[] = {1100000, 1500000};
    for (size_t c = 0; c < cases.size(); ++c) {
        List L = base;
        for (const Bad& b : cases[c]) {
            std::vector<int64_t>* col[] = {&L.tb, &L.te, &L.sb, &L.se};
            (*col[b.col])[b.k] = b.v;
        }
        HostSink s;
        PlanShape sh;
        CHECK(run(L, np, true, s, sh) == VPM_EINVAL, "bad list %zu accepted", c);
        const std::string want = "range pair " + std::to_string(first[c]) + " ";
        CHECK(g_error.find(want) != std::string::npos, "bad list %zu: %s", c, g_error.c_str());
    }
    HostSink s;
    s.fail_at = int64_t(1) << 20;
    PlanShape sh;
    CHECK(run(base, np, true, s, sh) == VPM_EDEVICE, "failed copy not reported");
}

static void case_shuffled_and_general(unsigned seed)
{
    const int64_t np = 3000;
    std::mt19937 rng(seed);
    // shuffled leaves of 16 with empty target ranges: still the leaf form
    List L = leaf_list(np, 16, 2, true, seed);
    for (int k = 0; k < 50; ++k) {
        const int64_t a = (int64_t)(rng() % np);
        L.tb.push_back(a); L.te.push_back(a); L.sb.push_back(0); L.se.push_back(np);
    }
    HostSink s;
    PlanShape sh;
    CHECK(run(L, np, true, s, sh) == VPM_OK, "shuffled leaves: %s", g_error.c_str());
    CHECK(sh.wave && sh.R == 4 && sh.src_copied, "shuffled leaves shape");
    check_plan("shuffled leaves", L, np, s, sh);
    // overlapping target ranges, repeated and empty source ranges: the
    // general interval plan
    List G;
    for (int k = 0; k < 20000; ++k) {
        const int64_t a = (int64_t)(rng() % np), b = std::min<int64_t>(np, a + (int64_t)(rng() % 300));
        const int64_t c = (int64_t)(rng() % np), d = std::min<int64_t>(np, c + (int64_t)(rng() % 200));
        G.tb.push_back(a); G.te.push_back(b); G.sb.push_back(c); G.se.push_back(d);
    }
    for (bool wave : {false, true}) {
        HostSink g;
        PlanShape gs;
        CHECK(run(G, np, wave, g, gs) == VPM_OK, "general list: %s", g_error.c_str());
        CHECK(!gs.src_copied, "general list copied early");
        check_plan(wave ? "general list (wave)" : "general list", G, np, g, gs);
    }
}

int main()
{
    case_long_leaf_list();
    case_errors_lowest_index();
    case_shuffled_and_general(11);
    case_shuffled_and_general(12);
    if (g_failures) {
        std::printf("plan_driver: %d failure(s)\n", g_failures);
        return 1;
    }
    std::printf("plan_driver OK\n");
    return 0;
}
