// Binomial (revolve) checkpoint schedule.
//
// Same algorithm and action list as implexity-ad's BinomialSchedule
// (crates/implexity-ad/src/revolve.rs:38-329): an initial snapshot of state 0, a
// recursive chain that advances by the optimal split, snapshots, recurses with one
// slot less and reverses; the first reverse is preceded by the advance to the final
// state.  Kept alike so that a schedule built here and one built by the reference
// drive identical recomputation (forward_steps == optimal_forward_steps(n, s)).
#pragma once
#include <cstdint>
#include <stdexcept>
#include <vector>

namespace lbm {

enum class ActKind : int { Advance = 0, Snapshot = 1, Restore = 2, Reverse = 3, Release = 4 };

struct Action {
    ActKind kind;
    uint64_t a;  // from / step
    uint64_t b;  // to / slot
};

struct Schedule {
    uint64_t steps = 0;
    std::vector<Action> actions;
    uint64_t reverse_start = 0;
    uint64_t forward_steps = 0;
    uint64_t peak = 0;
};

uint64_t revolve_repetitions(uint64_t steps, uint64_t snapshots);
uint64_t revolve_optimal_forward_steps(uint64_t steps, uint64_t snapshots);
// Throws std::invalid_argument on an impossible request (no steps or no slots).
Schedule revolve_schedule(uint64_t steps, uint64_t snapshots);

// Checkpointing::Every(k) of scan_adjoint (crates/implexity-ad/src/scan.rs:140-246)
// as an action list for the same driver: a snapshot every `every` steps in the first
// sweep; in reverse, each segment [s, e) is recomputed from its snapshot with all of its
// states kept (e - s - 1 steps) and reversed.  forward_steps = n + sum (len - 1);
// slots = ceil(n / every) + max(every - 2, 0).
Schedule every_schedule(uint64_t steps, uint64_t every);

}  // namespace lbm
