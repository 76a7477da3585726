// Host-side bookkeeping of the resident engine: buffer handles and the
// transaction journal (no HIP dependency; tests build it with g++).
//
// Journal.  Inside a transaction the first write to an element saves its old
// value into the journal arena (device memory the engine owns); later writes
// to that element save nothing, and a rollback copies every saved run back.
// The saved elements of a buffer are kept as disjoint sorted runs [lo, hi)
// in an ordered map, so bookkeeping costs grow with the number of runs an
// operation meets, not with its length, and only buffers the transaction
// touched carry any state.  An operation hands in every range it is about to
// write; the journal plans the saves for all of them at once and refuses the
// operation as a whole when the arena cannot hold them, before anything is
// copied or recorded.
//
// Semantics follow the reference's resident transactions
// (native/src/icicle_resident.cu:142-250: first-write saves, whole-operation
// preflight, rollback and commit; crates/gpu/tests/transform_packet.rs:131,
// incremental_session.rs:63 for the refusal behaviour).
#pragma once
#include <stdint.h>

#include <algorithm>
#include <iterator>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "errors.hpp"

namespace dvgpu {
namespace resident {

// Element kinds of resident buffers and their stored / boundary widths.
enum Kind : uint32_t { KIND_FR = 0, KIND_G1 = 1, KIND_BASES = 2 };
constexpr uint32_t stored_width(uint32_t kind) { return kind == KIND_FR ? 32 : (kind == KIND_G1 ? 128 : 64); }
constexpr uint32_t boundary_width(uint32_t kind) { return kind == KIND_FR ? 32 : 64; }

// Handles: the owner's tag in the high 32 bits and slot + 1 in the low 32,
// so 0 is never a handle and two owners' handles never collide.
inline uint64_t handle_of(uint32_t tag, uint32_t slot) { return (uint64_t(tag) << 32) | (uint64_t(slot) + 1); }

inline uint32_t slot_of(uint64_t handle, uint32_t tag, size_t slots) {
    require(uint32_t(handle >> 32) == tag, "resident handle belongs to another owner");
    const uint64_t low = handle & 0xffffffffull;
    require(low >= 1 && low <= slots, "unknown resident handle");
    return uint32_t(low - 1);
}

inline uint64_t mul_or_throw(uint64_t a, uint64_t b) {
    require(!a || b <= UINT64_MAX / a, "resident size overflows 64 bits");
    return a * b;
}

struct Run {
    uint64_t lo, hi;  // elements [lo, hi)
};

// Disjoint, sorted, non-touching runs of element indices.
class RunSet {
  public:
    bool empty() const { return runs_.empty(); }
    size_t size() const { return runs_.size(); }
    void clear() { runs_.clear(); }

    // Append the parts of [lo, hi) that no run covers, in ascending order.
    void holes(uint64_t lo, uint64_t hi, std::vector<Run>& out) const {
        if (lo >= hi) return;
        auto it = runs_.upper_bound(lo);
        if (it != runs_.begin()) {
            const auto before = std::prev(it);
            if (before->second > lo) lo = before->second;
        }
        for (; lo < hi && it != runs_.end() && it->first < hi; ++it) {
            if (it->first > lo) out.push_back({lo, it->first});
            lo = std::max(lo, it->second);
        }
        if (lo < hi) out.push_back({lo, hi});
    }

    // Add [lo, hi), merged with every run it overlaps or touches.
    void add(uint64_t lo, uint64_t hi) {
        if (lo >= hi) return;
        auto it = runs_.upper_bound(lo);
        if (it != runs_.begin()) {
            const auto before = std::prev(it);
            if (before->second >= lo) {
                lo = before->first;
                hi = std::max(hi, before->second);
                it = runs_.erase(before);
            }
        }
        while (it != runs_.end() && it->first <= hi) {
            hi = std::max(hi, it->second);
            it = runs_.erase(it);
        }
        runs_.emplace(lo, hi);
    }

    template <class F>
    void each(F&& f) const {
        for (const auto& r : runs_) f(r.first, r.second);
    }

  private:
    std::map<uint64_t, uint64_t> runs_;  // lo -> hi
};

// A range an operation is about to write: elements [lo, hi) of buffer
// `slot`, each `width` bytes.
struct WriteRange {
    uint32_t slot;
    uint64_t lo, hi;
    uint32_t width;
};

// A saved run: `count` elements from element `lo` of buffer `slot`, copied
// to byte `arena` of the journal arena.
struct SavedRun {
    uint32_t slot;
    uint64_t lo, count;
    uint32_t width;
    uint64_t arena;
    uint64_t bytes() const { return count * width; }
};

class Journal {
  public:
    explicit Journal(uint64_t arena_bytes = 0) : capacity_(arena_bytes) {}

    uint64_t capacity() const { return capacity_; }
    uint64_t used() const { return used_; }
    bool open() const { return open_; }
    // Saved runs in the order they were saved (a rollback walks them backwards).
    const std::vector<SavedRun>& log() const { return log_; }
    // Upper bound of the run records one transaction can hold: every run
    // saves at least one element of at least 32 bytes.
    uint64_t record_bound() const { return std::max<uint64_t>(1, capacity_ / 32); }

    void begin() {
        require(!open_, "a resident transaction is already open");
        reset();
        open_ = true;
    }

    // The saves `writes` need: none outside a transaction; inside one, the
    // parts of every range that are not saved yet (ranges of one call may
    // overlap each other: each element is saved once), placed after the used
    // part of the arena.  Throws, recording nothing, when they do not fit.
    std::vector<SavedRun> plan(const std::vector<WriteRange>& writes) const {
        std::vector<SavedRun> saves;
        if (!open_) return saves;
        std::unordered_map<uint32_t, RunSet> planned;
        std::vector<Run> unsaved, fresh;
        uint64_t need = 0;
        for (const WriteRange& w : writes) {
            unsaved.clear();
            const auto s = saved_.find(w.slot);
            if (s == saved_.end()) {
                if (w.lo < w.hi) unsaved.push_back({w.lo, w.hi});
            } else {
                s->second.holes(w.lo, w.hi, unsaved);
            }
            RunSet& mine = planned[w.slot];
            for (const Run& r : unsaved) {
                fresh.clear();
                mine.holes(r.lo, r.hi, fresh);
                for (const Run& f : fresh) {
                    const uint64_t bytes = mul_or_throw(f.hi - f.lo, w.width);
                    require(bytes <= UINT64_MAX - need, "resident size overflows 64 bits");
                    saves.push_back({w.slot, f.lo, f.hi - f.lo, w.width, used_ + need});
                    need += bytes;
                }
                mine.add(r.lo, r.hi);
            }
        }
        if (need > capacity_ - used_)
            throw InvalidArgument("resident transaction journal is full: the operation must save " +
                                  std::to_string(need) + " bytes and " + std::to_string(capacity_ - used_) +
                                  " of " + std::to_string(capacity_) + " are free");
        return saves;
    }

    // Record planned saves once their copies are enqueued.
    void record(const std::vector<SavedRun>& saves) {
        for (const SavedRun& s : saves) {
            saved_[s.slot].add(s.lo, s.lo + s.count);
            log_.push_back(s);
            used_ = std::max(used_, s.arena + s.bytes());
        }
    }

    // End the transaction: after a commit, or once a rollback has restored
    // every saved run.
    void close() {
        reset();
        open_ = false;
    }

  private:
    void reset() {
        saved_.clear();
        log_.clear();
        used_ = 0;
    }

    uint64_t capacity_ = 0, used_ = 0;
    bool open_ = false;
    std::unordered_map<uint32_t, RunSet> saved_;
    std::vector<SavedRun> log_;
};

// Elements of the destination that a transposing copy writes: elements
// [start, start + n) of a rows x cols row-major matrix, element r * cols + c
// landing at c * rows + r.  The elements of one source column c inside the
// window form one run of destination rows, so the result is at most
// min(cols, n) runs.
inline void transpose_destinations(uint64_t rows, uint64_t cols, uint64_t start, uint64_t n, std::vector<Run>& out) {
    if (!n || !cols) return;
    const uint64_t last = start + n - 1;
    const uint64_t columns = std::min(cols, n);
    for (uint64_t t = 0; t < columns; ++t) {
        const uint64_t first = start + t;  // first element of its column in the window
        const uint64_t c = first % cols;
        const uint64_t r0 = first / cols, r1 = (last - c) / cols;
        out.push_back({c * rows + r0, c * rows + r1 + 1});
    }
}

}  // namespace resident
}  // namespace dvgpu
