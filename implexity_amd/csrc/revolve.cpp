// Binomial checkpoint schedule; see revolve.h.
#include "revolve.h"

#include <algorithm>
#include <string>

namespace lbm {
namespace {

using u128 = unsigned __int128;
constexpr u128 kSaturated = u128(1) << 100;

// C(s + r, r), saturated (revolve.rs:60-71).
u128 binom(uint64_t s, int64_t r) {
    if (r < 0) return 0;
    u128 v = 1;
    for (int64_t i = 1; i <= r; i++) {
        v = v * u128(s + uint64_t(i)) / u128(i);
        if (v >= kSaturated) return kSaturated;
    }
    return v;
}

// r * n - C(s+r, r) * r / (s + 1) with the smallest r such that C(s+r, r) >= n
// (revolve.rs:95-109): total advances of the optimal reversal of a chain of n steps.
u128 chain_advances(uint64_t length, uint64_t snapshots) {
    if (length <= 1) return 0;
    if (snapshots == 0) return kSaturated;
    u128 r = 0, v = 1;
    while (v < u128(length)) {
        r += 1;
        v = v * (u128(snapshots) + r) / r;
    }
    return r * u128(length) - v * r / (u128(snapshots) + 1);
}

// Largest optimal first advance of a chain (revolve.rs:237-253).
uint64_t split(uint64_t length, uint64_t slots) {
    if (slots == 0) throw std::invalid_argument("a chain of more than one step needs a snapshot slot");
    if (slots == 1) return length - 1;
    const int64_t r = int64_t(revolve_repetitions(length, slots));
    const u128 l = length;
    const u128 tb = binom(slots - 1, r);
    const u128 tb1 = binom(slots - 1, r - 1);
    u128 lo = std::max<u128>(std::max<u128>(1, binom(slots, r - 2)), l > tb ? l - tb : 0);
    u128 hi = std::min<u128>(std::min<u128>(l - 1, binom(slots, r - 1)), l > tb1 ? l - tb1 : 0);
    if (lo > hi)
        throw std::invalid_argument("no optimal split for a chain of " + std::to_string(length) + " steps with " +
                                    std::to_string(slots) + " slots");
    return uint64_t(hi);
}

struct Builder {
    uint64_t steps;
    std::vector<Action> actions;
    std::vector<uint64_t> free;  // kept sorted descending; back() is the smallest free slot
    bool has_cursor = false;
    uint64_t cursor = 0;
    uint64_t in_use = 0, peak = 0, advanced = 0;
    bool reversing = false;
    uint64_t reverse_start = 0;

    Builder(uint64_t n, uint64_t slots) : steps(n) {
        for (uint64_t s = slots; s-- > 0;) free.push_back(s);
    }
    uint64_t take_slot() {
        if (free.empty()) throw std::invalid_argument("checkpoint schedule ran out of snapshot slots");
        const uint64_t s = free.back();
        free.pop_back();
        in_use++;
        peak = std::max(peak, in_use);
        return s;
    }
    void release(uint64_t slot) {
        actions.push_back({ActKind::Release, slot, 0});
        in_use = in_use ? in_use - 1 : 0;
        free.push_back(slot);
        std::sort(free.begin(), free.end(), std::greater<uint64_t>());
    }
    void advance(uint64_t from, uint64_t to) {
        if (to > from) {
            actions.push_back({ActKind::Advance, from, to});
            advanced += to - from;
            has_cursor = true;
            cursor = to;
        }
    }
    void ensure_cursor(uint64_t position, uint64_t slot) {
        if (!has_cursor || cursor != position) {
            actions.push_back({ActKind::Restore, position, slot});
            has_cursor = true;
            cursor = position;
        }
    }
    void reverse_step(uint64_t step) {
        if (!reversing) {
            actions.push_back({ActKind::Advance, steps - 1, steps});
            advanced += 1;
            reversing = true;
            reverse_start = actions.size();
        }
        actions.push_back({ActKind::Reverse, step, 0});
        has_cursor = false;
    }
    void chain(uint64_t a, uint64_t length, uint64_t own, uint64_t slots) {
        for (;;) {
            if (length == 0) return;
            if (length == 1) {
                ensure_cursor(a, own);
                reverse_step(a + 1);
                return;
            }
            const uint64_t m = split(length, slots);
            ensure_cursor(a, own);
            advance(a, a + m);
            if (length - m > 1) {
                const uint64_t slot = take_slot();
                actions.push_back({ActKind::Snapshot, a + m, slot});
                chain(a + m, length - m, slot, slots - 1);
                release(slot);
            } else {
                reverse_step(a + length);
            }
            length = m;
        }
    }
};

}  // namespace

uint64_t revolve_repetitions(uint64_t steps, uint64_t snapshots) {
    if (steps <= 1) return 0;
    if (snapshots == 0) return UINT64_MAX;
    uint64_t r = 0;
    u128 v = 1;
    while (v < u128(steps)) {
        r++;
        v = v * u128(snapshots + r) / u128(r);
    }
    return r;
}

uint64_t revolve_optimal_forward_steps(uint64_t steps, uint64_t snapshots) {
    if (steps == 0) return 0;
    if (steps == 1) return 1;
    if (snapshots == 0) return UINT64_MAX;
    const u128 v = chain_advances(steps, snapshots) + 1;
    return v > u128(UINT64_MAX) ? UINT64_MAX : uint64_t(v);
}

Schedule revolve_schedule(uint64_t steps, uint64_t snapshots) {
    if (steps == 0) throw std::invalid_argument("a checkpoint schedule needs at least one step");
    if (snapshots == 0) throw std::invalid_argument("a checkpoint schedule needs at least one snapshot slot");
    Builder b(steps, snapshots);
    const uint64_t own = b.take_slot();
    b.actions.push_back({ActKind::Snapshot, 0, own});
    b.has_cursor = true;
    b.cursor = 0;
    b.chain(0, steps, own, snapshots);
    Schedule s;
    s.steps = steps;
    s.actions = std::move(b.actions);
    s.reverse_start = b.reverse_start;
    s.forward_steps = b.advanced;
    s.peak = b.peak;
    return s;
}

Schedule every_schedule(uint64_t steps, uint64_t every) {
    if (steps == 0) throw std::invalid_argument("a checkpoint schedule needs at least one step");
    if (every == 0) throw std::invalid_argument("checkpoint interval must be at least one step");
    Schedule s;
    s.steps = steps;
    const uint64_t nseg = (steps + every - 1) / every;
    // slots 0..nseg-1 hold the segment starts, nseg.. the states inside a segment
    auto& A = s.actions;
    A.push_back({ActKind::Snapshot, 0, 0});
    uint64_t fwd = 0;
    for (uint64_t g = 0; g < nseg; g++) {
        const uint64_t a = g * every, e = std::min(a + every, steps);
        A.push_back({ActKind::Advance, a, e});
        fwd += e - a;
        if (e < steps) A.push_back({ActKind::Snapshot, e, g + 1});
    }
    s.reverse_start = A.size();
    for (uint64_t g = nseg; g-- > 0;) {
        const uint64_t a = g * every, e = std::min(a + every, steps);
        // recompute a+1 .. e-1 into the segment slots (the state e-1 is the cursor)
        A.push_back({ActKind::Restore, a, g});
        for (uint64_t t = a; t + 1 < e; t++) {
            A.push_back({ActKind::Advance, t, t + 1});
            fwd++;
            if (t + 2 < e) A.push_back({ActKind::Snapshot, t + 1, nseg + (t - a)});
        }
        for (uint64_t k = e; k > a; k--) {
            if (k < e) {
                // cursor must hold state k-1: a segment slot (k-1 > a) or the segment start
                const uint64_t slot = (k - 1 == a) ? g : nseg + (k - 2 - a);
                A.push_back({ActKind::Restore, k - 1, slot});
            }
            A.push_back({ActKind::Reverse, k, 0});
        }
        for (uint64_t t = a + 1; t + 1 < e; t++) A.push_back({ActKind::Release, nseg + (t - a - 1), 0});
        A.push_back({ActKind::Release, g, 0});
    }
    s.forward_steps = fwd;
    s.peak = nseg + (every > 2 ? every - 2 : 0);
    return s;
}

}  // namespace lbm
