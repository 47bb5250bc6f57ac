/*
 * mk_esys.h -- event-based pacing for the host library.
 *
 * Same rules as the ESys pacing system of myokit/_sim/pacing.h:
 *   - events are kept in a queue ordered by start time; scheduling an event
 *     at the same start as a queued one is an error (ESys_ScheduleEvent,
 *     pacing.h:191-213);
 *   - AdvanceTime(t) repeatedly processes the next start/stop while
 *     t >= tnext (with a DBL_EPSILON-relative tolerance, ESys_eq/ESys_geq,
 *     pacing.h:69-80): an active event ends at tdown, a queued event starts
 *     and sets the level, recurring events are re-queued at start + period
 *     until their multiplier runs out (pacing.h:482-548);
 *   - tdown snaps to the next event's start when the two are
 *     indistinguishable.
 */
#ifndef MK_ESYS_H
#define MK_ESYS_H

#include <cfloat>
#include <cmath>
#include <vector>

struct MkEvent {
    double level, duration, start, period, multiplier;
    int next;   /* index of the next event in the queue, or -1 */
};

class MkESys {
public:
    enum { OK = 0, NEGATIVE_TIME = -40, SIMULTANEOUS = -50, INVALID = -20 };

    /* rows: n x (level, start, duration, period, multiplier) */
    int init(double t0, const double* rows, long long n)
    {
        ev_.clear();
        head_ = fire_ = -1;
        time_ = tnext_ = tdown_ = t0;
        level_ = 0;
        for (long long k = 0; k < n; ++k) {
            const double* r = rows + 5 * k;
            MkEvent e;
            e.level = r[0];
            e.start = r[1];
            e.duration = r[2];
            e.period = r[3];
            e.multiplier = r[4];
            e.next = -1;
            if (e.period == 0 && e.multiplier != 0) return INVALID;
            if (e.period < 0 || e.multiplier < 0) return INVALID;
            ev_.push_back(e);
        }
        for (int k = 0; k < (int)ev_.size(); ++k) {
            int f = schedule(head_, k);
            if (f != OK) return f;
        }
        return advance(t0);
    }

    int advance(double t)
    {
        if (t < time_) return NEGATIVE_TIME;
        time_ = t;
        while (geq(time_, tnext_)) {
            if (fire_ >= 0 && geq(tnext_, tdown_)) {
                fire_ = -1;
                level_ = 0;
            }
            if (head_ >= 0 && geq(tnext_, ev_[head_].start)) {
                fire_ = head_;
                MkEvent& f = ev_[fire_];
                head_ = f.next;
                tdown_ = f.start + f.duration;
                level_ = f.level;
                if (f.period > 0) {
                    if (f.multiplier != 1) {
                        if (f.multiplier > 1) f.multiplier -= 1;
                        f.start += f.period;
                        int r = schedule(head_, fire_);
                        if (r != OK) return r;
                    } else {
                        f.period = 0;
                    }
                }
                if (head_ >= 0 && eq(ev_[head_].start, tdown_))
                    tdown_ = ev_[head_].start;
            }
            tnext_ = HUGE_VAL;
            if (fire_ >= 0 && tnext_ > tdown_) tnext_ = tdown_;
            if (head_ >= 0 && tnext_ > ev_[head_].start)
                tnext_ = ev_[head_].start;
        }
        return OK;
    }

    double next_time() const { return tnext_; }
    double level() const { return level_; }
    double time() const { return time_; }

private:
    static bool eq(double a, double b)
    {
        if (a == b) return true;
        double s = std::fabs(a) > std::fabs(b) ? std::fabs(a) : std::fabs(b);
        return std::fabs(a - b) / s < DBL_EPSILON;
    }
    static bool geq(double a, double b) { return a >= b || eq(a, b); }

    /* Inserts event k into the queue starting at `head` (updates head_). */
    int schedule(int head, int k)
    {
        MkEvent& add = ev_[k];
        add.next = -1;
        int flag = OK;
        if (head < 0) {
            head_ = k;
            return OK;
        }
        if (add.start < ev_[head].start) {
            add.next = head;
            head_ = k;
            return OK;
        }
        int e = head;
        while (ev_[e].next >= 0 && add.start >= ev_[ev_[e].next].start)
            e = ev_[e].next;
        if (add.start == ev_[e].start) flag = SIMULTANEOUS;
        add.next = ev_[e].next;
        ev_[e].next = k;
        head_ = head;
        return flag;
    }

    std::vector<MkEvent> ev_;
    int head_ = -1, fire_ = -1;
    double time_ = 0, tnext_ = 0, tdown_ = 0, level_ = 0;
};

#endif
