// TimeGovernor of the reference restated for the equation handles (Darcy and
// transport): marks, constraints, fixed step until a mark, steady form.
#include <algorithm>
#include <cmath>
#include <limits>

#include "time_governor.hh"

namespace fl {

// ---------------------------------------------------------------------------
// TimeGovernor restated (time_governor.cc:75-117, 215-237, 261-355,
// time_governor.hh:205-211, 347-355; TimeMarks::next time_marks.cc:125-132)
// ---------------------------------------------------------------------------
namespace {
const double kInf = std::numeric_limits<double>::infinity();
const double kEps = std::numeric_limits<double>::epsilon();
}  // namespace

void TimeGovernorR::init(double end, double save_step) {
    steady = false;
    level = 0;
    time = 0.0;
    end_time = end;
    end_fixed = 0.0;
    max_ts = end;
    time_step = end;
    last_time_step = 0.0;
    upper = kInf;
    lower = kEps;
    min_ts = kEps;
    marks.clear();
    add_mark(0.0);
    add_mark(end);
    for (double t = 0.0; t <= end * (1 + 1e-12); t += save_step) add_mark(t);
}
bool TimeGovernorR::lt(double o) const {
    double b = o - 16 * kEps * std::max(std::fabs(time), std::fabs(o));
    return !(time >= b);
}
bool TimeGovernorR::ge(double o) const {
    double b = o - 16 * kEps * std::max(std::fabs(time), std::fabs(o));
    return time >= b;
}
bool TimeGovernorR::is_end() const { return ge(end_time) || time == kInf; }
void TimeGovernorR::add_mark(double t) {
    auto it = std::lower_bound(marks.begin(), marks.end(), t);
    if (it != marks.end() && *it == t) return;
    marks.insert(it, t);
}
double TimeGovernorR::next_mark() const {
    auto it = std::lower_bound(marks.begin(), marks.end(), time);
    while (it != marks.end() && !lt(*it)) ++it;
    return it == marks.end() ? kInf : *it;
}
// fixed step (the `init_dt` key, time_governor.cc:100-107): min = max =
// both constraints = dt; marks at start, end and the equation's output marks
// every save_step (none for save_step <= 0)
void TimeGovernorR::init_fixed(double end, double dt, double save_step) {
    steady = false;
    level = 0;
    time = 0.0;
    end_time = end;
    end_fixed = 0.0;
    time_step = max_ts = min_ts = upper = lower = dt;
    last_time_step = 0.0;
    marks.clear();
    add_mark(0.0);
    add_mark(end);
    if (save_step > 0.0)
        for (double t = 0.0; t <= end * (1 + 1e-12); t += save_step) add_mark(t);
}
// steady governor (time_governor.cc:151-168): t = 0, dt = inf until next_time
// moves it to t = inf (end), dt = 0
void TimeGovernorR::init_steady() {
    steady = true;
    level = 0;
    time = 0.0;
    end_time = kInf;
    end_fixed = 0.0;
    time_step = kInf;
    last_time_step = 0.0;
    upper = max_ts = kInf;
    lower = min_ts = kEps;
    marks.assign(1, 0.0);
}
double TimeGovernorR::estimate_dt() const {
    if (time == kInf || is_end()) return 0.0;
    if (steady) return level == 0 ? kInf : 0.0;
    if (lt(end_fixed)) return fixed_dt;
    double full = next_mark() - time;
    double est = std::min(full, upper);
    est = std::max(est, lower);
    int n = (int)std::ceil(full / est);
    int nf = (int)std::floor(full / est);
    if (std::fabs(full / est - nf) < 1e-14) n = nf;
    return full / n;
}
int TimeGovernorR::set_upper_constraint(double u) {
    if (upper < u) return -1;
    if (lower <= u) { upper = u; return 0; }
    return 1;
}
double TimeGovernorR::fix_dt_until_mark() {
    if (steady) return 0.0;
    end_fixed = -kInf;
    fixed_dt = estimate_dt();
    dt_fixed_now = true;
    return end_fixed = next_mark();
}
void TimeGovernorR::next_time() {
    if (time == kInf || is_end()) return;
    if (steady) {
        last_time_step = end_time;
        level = 1;
        time_step = 0.0;
        time = end_time;
        dt_changed = false;
        return;
    }
    if (lt(end_fixed)) {
        fixed_dt = (end_fixed - time) / std::round((end_fixed - time) / fixed_dt);
        last_time_step = time_step;
        time_step = fixed_dt;
        if (dt_fixed_now) { dt_fixed_now = false; dt_changed = (last_time_step != time_step); }
        else dt_changed = false;
    } else {
        last_time_step = time_step;
        time_step = estimate_dt();
        dt_changed = (last_time_step != time_step);
    }
    time += time_step;
    level++;
    upper = std::min(end_time - time, max_ts);
    lower = min_ts;
}

}  // namespace fl
