// TimeGovernor of the reference (time_governor.hh / time_governor.cc)
// restated on the host: the time, step, marks and step constraints of one
// equation.  Steady form (steady Darcy: t = 0 and dt = inf until the solve,
// then t = inf and dt = 0), fixed-step form (the `init_dt` key) and the
// output-mark form of the transport (marks every save_step).
#pragma once
#include <vector>

namespace fl {

struct TimeGovernorR {
    double time = 0, end_time = 0, time_step = 0, last_time_step = 0, fixed_dt = 0, end_fixed = 0;
    double upper = 0, lower = 0, max_ts = 0, min_ts = 0;
    bool dt_fixed_now = false, dt_changed = true, steady = false;
    int level = 0;
    std::vector<double> marks;
    void init(double end, double save_step);  // marks at 0, end and every save_step
    void init_fixed(double end, double dt, double save_step);  // fixed step dt, marks at 0, end, every save_step
    void init_steady();
    bool lt(double o) const;
    bool ge(double o) const;
    bool is_end() const;
    void add_mark(double t);
    double next_mark() const;
    double estimate_dt() const;
    double estimate_time() const { return time + estimate_dt(); }
    int set_upper_constraint(double u);
    double fix_dt_until_mark();
    void next_time();
};

}  // namespace fl
