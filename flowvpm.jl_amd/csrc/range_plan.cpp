// range_plan.cpp -- host planner of (target range, source range) pair lists
// (FastMultipole's direct list, fmm.direct! per pair, src/FLOWVPM_fmm.jl:
// 102-168).  Host-only C++: built with g++ into the library, and on its own
// under ASan/UBSan and TSan by the CPU tests.
//
// Target ranges of different pairs may overlap, so in general the target
// index space is cut at every range end into disjoint elementary intervals,
// and each interval gets the source ranges of every pair covering it, in
// pair order.  The usual FMM list is simpler: every target range is a leaf,
// ranges are identical or disjoint, and the pairs of one target range are
// listed together; then the list itself is the per-interval source list and
// the plan is one linear pass, run on up to kPlanThreads host threads.
#include "range_plan.h"

#include <algorithm>
#include <exception>
#include <new>
#include <thread>
#include <vector>

#include "../../include/flowvpm_hip.h"

namespace vpm {

namespace {

int range_error(int64_t k)
{
    return fail(VPM_EINVAL, "range pair " + std::to_string(k) + " outside [0, np)");
}

// Runs fn(k0, k1, t) over [0, n) cut into T <= tmax contiguous chunks of at
// least kPlanPiece pairs, chunk t on thread t (the caller's thread takes
// chunk 0; a chunk whose thread cannot start runs on the caller's).  fn must
// not throw.  Returns T.
template <class Fn>
int plan_parallel(int64_t n, int tmax, Fn&& fn)
{
    int T = (int)std::min<int64_t>(n / kPlanPiece + 1, tmax);
    const unsigned hc = std::thread::hardware_concurrency();
    if (hc > 0) T = std::min(T, (int)hc);
    if (T <= 1) {
        fn((int64_t)0, n, 0);
        return 1;
    }
    std::vector<std::thread> th;
    std::vector<int> inline_chunks;
    th.reserve(T);
    inline_chunks.reserve(T);
    for (int t = 1; t < T; ++t) {
        try {
            th.emplace_back(fn, n * t / T, n * (t + 1) / T, t);
        } catch (...) {
            inline_chunks.push_back(t);
        }
    }
    fn((int64_t)0, n / T, 0);
    for (int t : inline_chunks) fn(n * t / T, n * (t + 1) / T, t);
    for (std::thread& x : th) x.join();
    return T;
}

// Lowest failing pair index over the chunks' first failures (-1 if none).
int64_t first_bad(const int64_t* bad, int T)
{
    int64_t k = -1;
    for (int t = 0; t < T; ++t)
        if (bad[t] >= 0 && (k < 0 || bad[t] < k)) k = bad[t];
    return k;
}

bool bad_range(int64_t b, int64_t e, int64_t np) { return b < 0 || e > np || b > e; }

int plan_impl(int64_t np, const int64_t* tb, const int64_t* te, const int64_t* sb, const int64_t* se,
              int64_t npairs, bool allow_wave, int rmax, PlanSink& sink, PlanShape& shape, int tmax)
{
    shape = PlanShape();
    // runs of consecutive pairs with one target range, found per chunk of the
    // list (a run belongs to the chunk it starts in) and joined in list order
    struct Run { int64_t tb, te, l0, l1; };
    struct Part { std::vector<Run> runs; int64_t tot = 0, cnt = 0; bool oom = false; };
    Part parts[kPlanThreads];
    int64_t bad[kPlanThreads];
    for (int t = 0; t < kPlanThreads; ++t) bad[t] = -1;
    auto same = [tb, te](int64_t a, int64_t b) { return tb[a] == tb[b] && te[a] == te[b]; };
    const int T = plan_parallel(npairs, tmax, [&](int64_t k0, int64_t k1, int t) {
        Part& pt = parts[t];
        int64_t k = k0;
        while (k < k1 && k > 0 && same(k, k - 1)) ++k;
        while (k < k1) {
            if (bad_range(tb[k], te[k], np)) { bad[t] = k; return; }
            int64_t m = k + 1;
            while (m < npairs && same(m, k)) ++m;
            try {
                pt.runs.push_back({tb[k], te[k], k, m});
            } catch (...) {   // no exception leaves a worker thread
                pt.oom = true;
                return;
            }
            if (te[k] > tb[k]) { pt.tot += te[k] - tb[k]; ++pt.cnt; }
            k = m;
        }
    });
    for (int t = 0; t < T; ++t)
        if (parts[t].oom) throw std::bad_alloc();
    if (int64_t k = first_bad(bad, T); k >= 0) return range_error(k);
    std::vector<Run> runs;
    int64_t tot = 0, cnt = 0;
    {
        size_t nr = 0;
        for (int t = 0; t < T; ++t) nr += parts[t].runs.size();
        runs.reserve(nr);
        for (int t = 0; t < T; ++t) {
            runs.insert(runs.end(), parts[t].runs.begin(), parts[t].runs.end());
            tot += parts[t].tot;
            cnt += parts[t].cnt;
        }
    }
    shape.wave = allow_wave && cnt > 0 && tot <= 128 * cnt;
    // leaves of <= 8, 16 or 32 targets on average share a wave's lanes 8, 4
    // or 2 ways
    int R = !shape.wave ? 1 : tot <= 8 * cnt ? 8 : tot <= 16 * cnt ? 4 : tot <= 32 * cnt ? 2 : 1;
    shape.R = std::min(R, std::max(rmax, 1));
    const int64_t chunk = shape.wave ? 64 / shape.R : 256;
    auto pieces = [chunk](int64_t a, int64_t b) { return b > a ? (b - a + chunk - 1) / chunk : 0; };

    // Leaf lists: target ranges identical or disjoint, the pairs of one range
    // together.  Empty target ranges hold no targets and are left out of the
    // check.  Lists whose runs come in target order, as FMM lists do, need no
    // sort.
    bool leaves = true;
    {
        const Run* prev = nullptr;
        bool sorted = true;
        for (const Run& r : runs) {
            if (r.te == r.tb) continue;
            if (prev && r.tb < prev->tb) { sorted = false; break; }
            if (prev && r.tb < prev->te) { leaves = false; break; }   // overlap, or one range in two runs
            prev = &r;
        }
        if (!sorted) {
            std::vector<const Run*> order;
            for (const Run& r : runs)
                if (r.te > r.tb) order.push_back(&r);
            std::sort(order.begin(), order.end(), [](const Run* a, const Run* b) { return a->tb < b->tb; });
            for (size_t r = 1; r < order.size() && leaves; ++r)
                if (order[r]->tb < order[r - 1]->te) leaves = false;
        }
    }
    if (leaves) {
        int64_t nitems = 0;
        for (const Run& r : runs) nitems += pieces(r.tb, r.te);
        RangeItem* items;
        SrcRange* src;
        if (int rc = sink.stage(nitems, npairs, items, src)) return rc;
        shape.nitems = nitems;
        shape.nsrc = npairs;
        // each piece of packed source ranges goes to the device as soon as it
        // is written, behind the packing of the next
        if (int rc = sink.prepare_copies(nitems, npairs)) return rc;
        int codes[kPlanThreads];
        for (int t = 0; t < kPlanThreads; ++t) { bad[t] = -1; codes[t] = 0; }
        const int T2 = plan_parallel(npairs, tmax, [&](int64_t k0, int64_t k1, int t) {
            for (int64_t a = k0; a < k1; a += kPlanPiece) {
                const int64_t b = std::min(a + kPlanPiece, k1);
                for (int64_t k = a; k < b; ++k) {
                    if (bad_range(sb[k], se[k], np)) { bad[t] = k; return; }
                    src[k] = SrcRange{(int32_t)sb[k], (int32_t)se[k]};
                }
                if (int c = sink.copy_piece(a, b)) { codes[t] = c; return; }
            }
        });
        if (int64_t k = first_bad(bad, T2); k >= 0) return range_error(k);
        for (int t = 0; t < T2; ++t)
            if (codes[t]) return sink.report(codes[t]);
        shape.src_copied = true;
        int64_t m = 0;
        for (const Run& r : runs)
            for (int64_t t = r.tb; t < r.te; t += chunk)
                items[m++] = {t, std::min<int64_t>(t + chunk, r.te), r.l0, r.l1};
        return VPM_OK;
    }

    // general lists: elementary intervals between all range ends
    std::vector<int64_t> cuts;
    cuts.reserve(2 * (size_t)npairs);
    for (int64_t k = 0; k < npairs; ++k) {
        if (bad_range(sb[k], se[k], np)) return range_error(k);
        if (tb[k] < te[k] && sb[k] < se[k]) { cuts.push_back(tb[k]); cuts.push_back(te[k]); }
    }
    if (cuts.empty()) return VPM_OK;
    std::sort(cuts.begin(), cuts.end());
    cuts.erase(std::unique(cuts.begin(), cuts.end()), cuts.end());
    const size_t nint = cuts.size() - 1;
    std::vector<int64_t> count(nint + 1, 0);
    for (int64_t k = 0; k < npairs; ++k) {
        if (!(tb[k] < te[k] && sb[k] < se[k])) continue;
        const size_t lo = std::lower_bound(cuts.begin(), cuts.end(), tb[k]) - cuts.begin();
        const size_t hi = std::lower_bound(cuts.begin(), cuts.end(), te[k]) - cuts.begin();
        for (size_t m = lo; m < hi; ++m) count[m + 1]++;
    }
    for (size_t m = 0; m < nint; ++m) count[m + 1] += count[m];
    int64_t nitems = 0;
    for (size_t m = 0; m < nint; ++m)
        if (count[m + 1] != count[m]) nitems += pieces(cuts[m], cuts[m + 1]);
    RangeItem* items;
    SrcRange* src;
    if (int rc = sink.stage(nitems, count[nint], items, src)) return rc;
    shape.nitems = nitems;
    shape.nsrc = count[nint];
    std::vector<int64_t> fill(count.begin(), count.end() - 1);
    for (int64_t k = 0; k < npairs; ++k) {
        if (!(tb[k] < te[k] && sb[k] < se[k])) continue;
        const size_t lo = std::lower_bound(cuts.begin(), cuts.end(), tb[k]) - cuts.begin();
        const size_t hi = std::lower_bound(cuts.begin(), cuts.end(), te[k]) - cuts.begin();
        for (size_t m = lo; m < hi; ++m) src[fill[m]++] = SrcRange{(int32_t)sb[k], (int32_t)se[k]};
    }
    int64_t q = 0;
    for (size_t m = 0; m < nint; ++m) {
        if (count[m + 1] == count[m]) continue;
        for (int64_t t = cuts[m]; t < cuts[m + 1]; t += chunk)
            items[q++] = {t, std::min<int64_t>(t + chunk, cuts[m + 1]), count[m], count[m + 1]};
    }
    return VPM_OK;
}

}  // namespace

int plan_ranges(int64_t np, const int64_t* tb, const int64_t* te, const int64_t* sb, const int64_t* se,
                int64_t npairs, bool allow_wave, int rmax, PlanSink& sink, PlanShape& shape, int threads)
{
    const int tmax = threads > 0 ? std::min(threads, kPlanThreads) : kPlanThreads;
    try {
        return plan_impl(np, tb, te, sb, se, npairs, allow_wave, rmax, sink, shape, tmax);
    } catch (const std::exception&) {
        shape = PlanShape();
        return fail(VPM_ENOMEM, "host allocation failed while planning a range list");
    }
}

}  // namespace vpm
