// pzl_candlist.h -- the candidate list with the semantics of
// insert_new_accelcand (accel_utils.c:382-480).
//
// The reference keeps a singly linked GSList sorted by r and, for each hit
// above threshold (in scan order), walks from the head to the first node
// with r >= rr (or the last node).  A hit within ACCEL_CLOSEST_R (15) bins of
// the previous or that node overwrites it when its sigma is higher (and, if it
// overwrote the previous node, removes the next one when that is also close and
// weaker); otherwise it is inserted as a new node.
//
// The list is held in a vector.  The head walk is a binary search over the
// running maximum of r (equal to r itself while the list is sorted, which
// every operation preserves except one rare overwrite at the tail), and
// consecutive hits of one plane row move the search position by at most a
// node, so a cursor makes the typical hit O(1).
#pragma once

#include <algorithm>
#include <cmath>
#include <vector>

#include "pzl_accel.h"

namespace pzl {

class CandList {
  public:
    static constexpr double kClosest = 15.0;

    // decode(i) -> PzlCand for the hit being inserted; called only when the
    // list changes (rare), so callers can defer z/w/numharm decoding.
    template <typename Decode>
    void insert(double rr, float sigma, Decode &&decode)
    {
        const size_t n = v_.size();
        if (n == 0) {
            v_.push_back(decode());
            touch(0);
            return;
        }
        size_t tmp = find(rr);
        if (tmp == n) tmp = n - 1;
        const bool has_prev = tmp > 0;
        const double next_diff = std::fabs(rr - v_[tmp].r);
        const double prev_diff = has_prev ? std::fabs(rr - v_[tmp - 1].r) : kClosest + 1.0;
        if (prev_diff < kClosest) {
            const size_t prev = tmp - 1;
            bool added = false;
            if (v_[prev].sigma < sigma) {
                set(prev, decode());
                added = true;
            }
            if (next_diff < kClosest && v_[tmp].sigma < sigma) {
                if (added) erase(tmp);
                else set(tmp, decode());
            }
        } else if (next_diff < kClosest) {
            if (v_[tmp].sigma < sigma) set(tmp, decode());
        } else if (tmp + 1 == n && v_[tmp].r < rr - kClosest) {
            insert_at(n, decode());
        } else {
            insert_at(tmp, decode());  // before tmp (at the head when tmp == 0)
        }
    }

    const std::vector<PzlCand> &items() const { return v_; }

  private:
    // First index with r >= rr (the reference's walk from the head), n if
    // none.  The walk stops at the first node whose r reaches rr, which is the
    // first index where the running maximum of r reaches rr; that running
    // maximum is non-decreasing even when an overwrite left the list out of
    // order, so a binary search (or the cursor) finds it.
    size_t find(double rr)
    {
        const size_t n = v_.size();
        refresh();
        const double *m = pm_.data();
        if (pos_ <= n && (pos_ == 0 || m[pos_ - 1] < rr) && (pos_ == n || m[pos_] >= rr)) return pos_;
        if (pos_ < n && m[pos_] < rr && (pos_ + 1 == n || m[pos_ + 1] >= rr)) return ++pos_;
        pos_ = std::lower_bound(m, m + n, rr) - m;
        return pos_;
    }
    void refresh()
    {
        const size_t n = v_.size();
        if (pm_.size() != n) pm_.resize(n);
        for (size_t t = dirty_; t < n; t++) pm_[t] = t ? std::max(pm_[t - 1], v_[t].r) : v_[t].r;
        dirty_ = n;
    }
    void touch(size_t i) { dirty_ = std::min(dirty_, i); }
    void set(size_t i, const PzlCand &c)
    {
        v_[i] = c;
        touch(i);
    }
    void erase(size_t i)
    {
        v_.erase(v_.begin() + i);
        touch(i);
    }
    void insert_at(size_t i, const PzlCand &c)
    {
        v_.insert(v_.begin() + i, c);
        touch(i);
    }

    std::vector<PzlCand> v_;
    std::vector<double> pm_;  // running maximum of v_[..].r
    size_t dirty_ = 0;        // pm_ is valid below this index
    size_t pos_ = 0;          // last find() result (a hint)
};

}  // namespace pzl
