// Host-only test of the node-counter lending rules (gpuboscia_amd/csrc/gb_slots.h):
// simulated launches on several streams, with completion markers that
// signal only when the simulated launch ends, check that no two launches
// that can run at the same time ever hold the same slot, that captured
// slots are never lent again, and that captures never need a marker query.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <set>
#include <vector>

#include "../../gpuboscia_amd/csrc/gb_slots.h"

namespace {

int failures = 0;
#define CHECK(cond, msg)                                            \
    do {                                                            \
        if (!(cond)) {                                              \
            std::printf("FAIL %s (line %d)\n", msg, __LINE__);      \
            ++failures;                                             \
        }                                                           \
    } while (0)

// A marker is the id of the simulated launch recorded into it; it has
// signalled once that launch has finished.
struct Sim {
    std::set<int> running;     // launch ids still running
    std::vector<int> holder;   // slot -> launch id using it (-1: none)
};

}  // namespace

int main() {
    using Table = gb::SlotTable<int>;
    std::mt19937 rng(12345);
    // 1. Random eager traffic on 7 streams, launches finishing out of order.
    {
        Table t;
        Sim sim;
        int next_launch = 0;
        auto done = [&](int mark) { return sim.running.count(mark) == 0; };
        for (int step = 0; step < 20000; ++step) {
            if (t.fresh() < 8) t.grow(16);
            sim.holder.resize(t.size(), -1);
            const int slot = t.take(false, done);
            CHECK(slot >= 0, "an eager take finds a slot after growth");
            if (slot < 0) break;
            const int prev = sim.holder[(size_t)slot];
            CHECK(prev < 0 || sim.running.count(prev) == 0, "a lent slot is never in use by a running launch");
            const int id = next_launch++;
            sim.running.insert(id);
            sim.holder[(size_t)slot] = id;
            t.give_back(slot, false, [&](int& m) { m = id; return true; });
            // finish a random subset of running launches
            std::vector<int> fin;
            for (int r : sim.running)
                if (rng() % 3 == 0) fin.push_back(r);
            for (int r : fin) sim.running.erase(r);
        }
        CHECK(t.size() < 20000, "slots are reused");
    }
    // 2. Captured slots stay with their graph; captures take only never-used slots.
    {
        Table t;
        t.grow(4);
        bool queried = false;
        auto spy = [&](int) { queried = true; return true; };
        const int a = t.take(false, spy);
        t.give_back(a, false, [](int& m) { m = 1; return true; });
        const int c1 = t.take(true, spy);
        CHECK(c1 >= 0 && c1 != a, "a capture takes a never-used slot");
        CHECK(!queried, "a capture queries no marker");
        t.give_back(c1, true, [](int&) { return true; });
        for (int k = 0; k < 10; ++k) {
            const int e = t.take(false, [](int) { return true; });
            CHECK(e != c1, "a captured slot is never lent again");
            if (e >= 0) t.give_back(e, false, [](int& m) { m = 2; return true; });
        }
        // exhaust never-used slots under capture
        int taken = 0;
        for (;;) {
            const int c = t.take(true, spy);
            if (c < 0) break;
            t.give_back(c, true, [](int&) { return true; });
            ++taken;
        }
        CHECK(taken == 2, "a capture can take exactly the remaining never-used slots");
        CHECK(!queried, "captures never query markers");
    }
    // 3. A failed marker record retires the slot.
    {
        Table t;
        t.grow(2);
        const int a = t.take(false, [](int) { return true; });
        t.give_back(a, false, [](int&) { return false; });
        const int b = t.take(false, [](int) { return true; });
        t.give_back(b, false, [](int& m) { m = 0; return true; });
        const int c = t.take(false, [](int) { return true; });
        CHECK(b != a && c != a, "a slot whose marker could not be recorded is retired");
    }
    // 4. Busy slots are never lent twice even when their old markers signalled.
    {
        Table t;
        t.grow(3);
        std::set<int> held;
        for (int k = 0; k < 3; ++k) {
            const int s = t.take(false, [](int) { return true; });
            CHECK(s >= 0 && held.count(s) == 0, "distinct slots while all are lent");
            held.insert(s);
        }
        CHECK(t.take(false, [](int) { return true; }) < 0, "no slot while all are lent");
    }
    if (failures) {
        std::printf("%d failure(s)\n", failures);
        return 1;
    }
    std::printf("slot table: all checks passed\n");
    return 0;
}
