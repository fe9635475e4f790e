// CPU unit test of the extrapolation rule shared by the batched and the
// single-vector personalized PageRank sweeps (ag_extrap.h): when it fires, with
// what weight, and that the steps it takes speed up a linear fixed-point
// iteration while every stop is still certified by a plain sweep's residual.
// No GPU, no HIP.  Prints one line per case and exits non-zero on a failure.
#include <cmath>
#include <cstdio>
#include <vector>

#include "ag_extrap.h"

using ag::extrapolation_weight;

static int failures = 0;

static void expect(bool ok, const char *what)
{
    std::printf("%s %s\n", ok ? "PASS" : "FAIL", what);
    if (!ok) failures++;
}

// Feeds residuals R_k = r0 rho^k; returns the sweep index (1-based) of the
// first nonzero weight, 0 if none within n sweeps, and that weight.
static int first_fire(double r0, double rho, int n, float &w)
{
    uint32_t since = 0;
    double last = 0.0, prev = INFINITY;
    for (int k = 1; k <= n; k++) {
        const double R = r0 * std::pow(rho, k);
        w = extrapolation_weight(R, prev, false, since, last);
        if (w != 0.0f) return k;
        prev = R;
    }
    return 0;
}

// x_{k+1} = A x_k + b with A = diag(lam) (the slowest mode first) and a fixed
// point x* = b / (1 - lam); returns sweeps until the plain-sweep residual
// ||x_{k+1} - x_k||_1 <= tol, and the final error ||x - x*||_1.
static int iterate(bool extrap, double tol, double &err)
{
    const double lam[3] = {0.6, 0.25, -0.1};
    const double b[3] = {0.4, 0.75, 1.1};
    double x[3] = {0.0, 0.0, 0.0};
    uint32_t since = 0;
    double rho = 0.0, prev = INFINITY;
    float w = 0.0f;
    for (int k = 1; k <= 200; k++) {
        double R = 0.0, nx[3];
        for (int i = 0; i < 3; i++) {
            const double t = lam[i] * x[i] + b[i];
            R += std::fabs(t - x[i]);
            nx[i] = w != 0.0f ? t + w * (t - x[i]) : t;
        }
        const bool stepped = w != 0.0f;
        for (int i = 0; i < 3; i++) x[i] = nx[i];
        if (R <= tol && !stepped) {
            err = 0.0;
            for (int i = 0; i < 3; i++) err += std::fabs(x[i] - b[i] / (1.0 - lam[i]));
            return k;
        }
        w = extrap ? extrapolation_weight(R, prev, stepped, since, rho) : 0.0f;
        prev = R;
    }
    err = INFINITY;
    return 0;
}

int main()
{
    float w = 0.0f;
    expect(first_fire(1.0, 0.61, 20, w) == 3, "a steady ratio fires on the third plain sweep");
    expect(std::fabs(w - 0.61 / 0.39) < 1e-5, "weight rho / (1 - rho)");
    expect(first_fire(1.0, 0.1, 20, w) == 0, "ratios below 0.2 never fire (fast enough)");
    expect(first_fire(1.0, 0.97, 20, w) == 0, "ratios above 0.95 never fire");

    {   // ratios that wander by more than 2% never fire
        uint32_t since = 0;
        double rho = 0.0, prev = INFINITY, R = 1.0;
        bool fired = false;
        for (int k = 0; k < 20; k++) {
            R *= (k % 2) ? 0.5 : 0.6;
            fired |= extrapolation_weight(R, prev, false, since, rho) != 0.0f;
            prev = R;
        }
        expect(!fired, "alternating ratios 0.5 / 0.6 never fire");
    }
    {   // after a step: the state resets and three plain sweeps pass first
        uint32_t since = 5;
        double rho = 0.61;
        expect(extrapolation_weight(0.01, 0.02, true, since, rho) == 0.0f && since == 0 &&
                   rho == 0.0,
               "a stepped sweep resets the state and takes no step");
        double prev = 0.01, R = 0.01;
        int fired_at = 0;
        for (int k = 1; k <= 6 && !fired_at; k++) {
            R *= 0.5;
            if (extrapolation_weight(R, prev, false, since, rho) != 0.0f) fired_at = k;
            prev = R;
        }
        expect(fired_at == 3, "next step after three plain sweeps");
    }
    {   // the first sweep has no previous residual
        uint32_t since = 0;
        double rho = 0.0;
        expect(extrapolation_weight(1.0, INFINITY, false, since, rho) == 0.0f && rho == 0.0,
               "no ratio before the second sweep");
    }
    {   // a linear iteration: fewer sweeps, and the stop stays certified
        double e_plain = 0.0, e_ext = 0.0;
        const double tol = 1e-10;
        const int n_plain = iterate(false, tol, e_plain);
        const int n_ext = iterate(true, tol, e_ext);
        std::printf("linear iteration: %d sweeps plain, %d extrapolated; errors %.2e %.2e\n",
                    n_plain, n_ext, e_plain, e_ext);
        expect(n_plain > 0 && n_ext > 0 && n_ext < n_plain, "extrapolation takes fewer sweeps");
        // ||x - x*||_1 <= max|lam| / (1 - max|lam|) * tol for a plain stop
        expect(e_ext <= 0.6 / 0.4 * tol && e_plain <= 0.6 / 0.4 * tol,
               "both stops within the contraction bound");
    }
    std::printf("%s\n", failures ? "FAILED" : "OK");
    return failures ? 1 : 0;
}
