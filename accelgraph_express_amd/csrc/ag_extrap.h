// Extrapolation rule of personalized PageRank, shared by the batched sweep
// (k_ppr_reduce, per column) and the single-vector sweep (k_pr_finalize).
//
// Aitken-type extrapolation (Kamvar, Haveliwala, Manning, Golub, "Extrapolation
// methods for accelerating PageRank computations", WWW 2003): once the L1
// residual of an iterate falls by the same ratio rho on consecutive sweeps,
// its error is dominated by one mode e_k = rho^k v, and x + rho / (1 - rho)
// (T(x) - x) removes that mode.  On R-MAT graphs the ratio settles at 0.3-0.7
// after a few sweeps.  A sweep that took the step cannot certify a stop (its
// residual belongs to the iterate it read), so callers stop only on a plain
// sweep's residual and the contraction bound ||x - x*||_1 <= d / (1 - d) tol
// holds as for plain power iteration.
//
// constexpr functions are host and device code alike under hipcc, so the rule
// is unit-tested on the CPU (tests/c/test_extrapolation.cpp).
#pragma once

#include <cstdint>

namespace ag {

struct ExtrapRule {
    static constexpr uint32_t kMinPlain = 3;   // plain sweeps before (and between) steps
    static constexpr double kAgree = 0.02;     // consecutive ratios agree within 2%
    static constexpr double kRhoMin = 0.2, kRhoMax = 0.95;
};

// One sweep's update of the rule.  R: this sweep's residual; prev: the last
// sweep's (0 or not finite before the first); stepped: this sweep stored an
// extrapolated iterate.  since / rho: the rule's state (plain sweeps since the
// start or the last step, last ratio).  Returns the weight w = rho / (1 - rho)
// for the next sweep's step, 0 for a plain sweep.
constexpr float extrapolation_weight(double R, double prev, bool stepped, uint32_t &since, double &rho)
{
    if (stepped) {
        since = 0;
        rho = 0.0;
        return 0.0f;
    }
    const double r = (prev > 0.0 && prev < 1e300) ? R / prev : 0.0;
    since += 1;
    const double diff = r > rho ? r - rho : rho - r;
    const bool fire = since >= ExtrapRule::kMinPlain && r >= ExtrapRule::kRhoMin &&
                      r <= ExtrapRule::kRhoMax && diff <= ExtrapRule::kAgree * r;
    rho = r;
    return fire ? (float)(r / (1.0 - r)) : 0.0f;
}

} // namespace ag
