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
// The list is held in a vector.  While it is non-decreasing in r (every
// operation preserves that except one rare overwrite), the head walk is a
// lower_bound, and consecutive hits of one plane row move the search position
// by at most a node, so a cursor makes the typical hit O(1).  When an
// overwrite leaves the vector out of order the walk is done literally.
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
            pos_valid_ = false;
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
    // first index with r >= rr (walking as the reference does), n if none
    size_t find(double rr)
    {
        const size_t n = v_.size();
        if (inversions_ == 0) {
            if (pos_valid_ && pos_ <= n && (pos_ == 0 || v_[pos_ - 1].r < rr) && (pos_ == n || v_[pos_].r >= rr))
                return pos_;
            if (pos_valid_ && pos_ < n && v_[pos_].r < rr && (pos_ + 1 == n || v_[pos_ + 1].r >= rr)) {
                pos_ = pos_ + 1;
                return pos_;
            }
            pos_ = std::lower_bound(v_.begin(), v_.end(), rr, [](const PzlCand &a, double x) { return a.r < x; }) -
                   v_.begin();
            pos_valid_ = true;
            return pos_;
        }
        size_t t = 0;
        while (t + 1 < n && v_[t].r < rr) t++;
        if (t + 1 == n && v_[t].r < rr) return n;
        return t;
    }
    int inv(size_t i) const { return (i + 1 < v_.size() && v_[i].r > v_[i + 1].r) ? 1 : 0; }
    void set(size_t i, const PzlCand &c)
    {
        if (i > 0) inversions_ -= inv(i - 1);
        inversions_ -= inv(i);
        v_[i] = c;
        if (i > 0) inversions_ += inv(i - 1);
        inversions_ += inv(i);
        pos_valid_ = false;
    }
    void erase(size_t i)
    {
        if (i > 0) inversions_ -= inv(i - 1);
        inversions_ -= inv(i);
        v_.erase(v_.begin() + i);
        if (i > 0) inversions_ += inv(i - 1);
        pos_valid_ = false;
    }
    void insert_at(size_t i, const PzlCand &c)
    {
        if (i > 0) inversions_ -= inv(i - 1);
        v_.insert(v_.begin() + i, c);
        if (i > 0) inversions_ += inv(i - 1);
        inversions_ += inv(i);
        pos_valid_ = false;
    }

    std::vector<PzlCand> v_;
    long long inversions_ = 0;
    size_t pos_ = 0;
    bool pos_valid_ = false;
};

}  // namespace pzl
