// Ring allocator of the serving loop's device ring (and its pinned host
// mirror at the same offsets).  Plain C++, so the host-side checker in
// tests/native/ring_check.cpp builds it with g++.
//
// The serving loop allocates one block per batch at admission and releases it
// when the batch's results are delivered, in completion order, not admission
// order.  Its admission rule relies on one property of this allocator: with
// no block live, any size up to the capacity is granted.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <deque>

namespace spl {

// Ring allocator over [0, cap): blocks are allocated at the head and may be
// released in any order; space is reclaimed from the oldest block forward.
class Ring {
  public:
    explicit Ring(uint64_t cap = 0) : cap_(cap) {}

    bool alloc(uint64_t size, uint64_t *off) {
        if (size == 0 || size > cap_) return false;
        uint64_t at;
        if (blocks_.empty()) {
            at = 0;
        } else {
            const uint64_t tail = blocks_.front().off;
            const bool wrapped = blocks_.back().off < tail;
            if (!wrapped) {
                if (head_ + size <= cap_) at = head_;
                else if (size <= tail) at = 0;
                else return false;
            } else {
                if (head_ + size <= tail) at = head_;
                else return false;
            }
        }
        blocks_.push_back({at, size, false});
        head_ = at + size;
        used_ += size;
        peak_ = std::max(peak_, used_);
        *off = at;
        return true;
    }

    void release(uint64_t off) {
        for (auto &b : blocks_)
            if (b.off == off && !b.freed) {
                b.freed = true;
                used_ -= b.size;
                break;
            }
        while (!blocks_.empty() && blocks_.front().freed) blocks_.pop_front();
        if (blocks_.empty()) head_ = 0;
    }

    uint64_t peak() const { return peak_; }
    uint64_t cap() const { return cap_; }

  private:
    struct Block {
        uint64_t off, size;
        bool freed;
    };
    uint64_t cap_, head_ = 0, used_ = 0, peak_ = 0;
    std::deque<Block> blocks_;
};

}  // namespace spl
