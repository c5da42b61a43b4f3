// Reuse rules of the library's per-launch device slots (the node counters of
// the persistent kernels).  Plain C++ with no HIP dependency, so the rules
// are unit-tested on the host (tests/cpp/slot_table_test.cpp).
//
// A slot is lent to one call at a time.  When the call has queued its
// launches it gives the slot back with a completion marker recorded on its
// stream (a HIP event); the slot is lent again only once that marker has
// signalled, so two launches that may run at the same time never share a
// slot.  A slot lent while the call's stream is being captured into a graph
// belongs to that graph for good (its replays may run at any later time)
// and is never lent again.  Captured calls take only slots that were never
// used, so a capture needs no marker query; eager calls keep `reserve` such
// slots available.
#pragma once
#include <cstddef>
#include <vector>

namespace gb {

template <class Marker>
class SlotTable {
  public:
    struct Slot {
        Marker mark{};
        bool marked = false;    // a marker was recorded at the last give-back
        bool busy = false;      // lent out
        bool captured = false;  // owned by a captured graph
    };

    // Lends a slot: an eager call takes a marked slot whose marker has
    // signalled (done(mark) true), else a never-used one; a capturing call
    // takes only never-used slots.  Returns -1 when none is available
    // (the caller grows the table and asks again).
    template <class Done>
    int take(bool capturing, Done&& done) {
        const size_t n = slots_.size();
        if (!capturing) {
            for (size_t k = 0; k < n; ++k) {
                const size_t i = (next_ + k) % n;
                Slot& s = slots_[i];
                if (s.busy || s.captured || !s.marked) continue;
                if (!done(s.mark)) continue;
                s.busy = true;
                next_ = i + 1;
                return static_cast<int>(i);
            }
        }
        for (size_t i = 0; i < n; ++i) {
            Slot& s = slots_[i];
            if (!s.busy && !s.captured && !s.marked) {
                s.busy = true;
                --fresh_;
                return static_cast<int>(i);
            }
        }
        return -1;
    }

    // Gives slot i back.  capturing: the slot stays with the graph.
    // Otherwise record(mark) records the completion marker on the call's
    // stream; if it fails the slot is retired (never lent again).
    template <class Record>
    void give_back(int i, bool capturing, Record&& record) {
        Slot& s = slots_[static_cast<size_t>(i)];
        s.busy = false;
        if (capturing) {
            s.captured = true;
            return;
        }
        if (record(s.mark)) {
            s.marked = true;
        } else {
            s.captured = true;
        }
    }

    void grow(size_t count) {
        slots_.resize(slots_.size() + count);
        fresh_ += count;
    }
    size_t size() const { return slots_.size(); }
    size_t fresh() const { return fresh_; }
    const Slot& slot(size_t i) const { return slots_[i]; }
    Marker& marker(size_t i) { return slots_[i].mark; }
    void clear() {
        slots_.clear();
        fresh_ = 0;
        next_ = 0;
    }

  private:
    std::vector<Slot> slots_;
    size_t fresh_ = 0;
    size_t next_ = 0;
};

}  // namespace gb
