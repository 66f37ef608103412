// chkn_bl.h -- inflow functions of the inflow_function boundary condition (host side).
//
// The ghost-cell states of an inflow_function side depend only on the cell positions, so the
// host evaluates them once at set-up (bcs.cu:321-352 per ghost cell, and per boundary face
// for the viscous boundary condition, bcs.cu:790-792).  The laminar boundary-layer profile is
// a natural cubic spline in y through 12 samples of T and velx (spline.cu:20-84, with the
// second derivatives from Gauss-Jordan elimination with partial pivoting, rsla.cu:196-227),
// clamped to the end samples outside their range (config.cu:73-93).  All arithmetic keeps
// the reference's operation order.
#pragma once

#include <cmath>
#include <stdexcept>
#include <utility>

#include "chkn_physics.h"

namespace chkn {
namespace bl {

constexpr int kSamples = 12;  // config.cu:31

// Natural cubic spline through n+1 = kSamples knots (spline.cu:20-109).
struct Spline {
    static constexpr int n = kSamples - 1;
    double xs[n + 1], ys[n + 1], a[n], b[n], c[n];

    void set(const double* x, const double* y)
    {
        for (int i = 0; i <= n; ++i) { xs[i] = x[i]; ys[i] = y[i]; }
        double h[n];
        for (int i = 0; i < n; ++i) h[i] = xs[i + 1] - xs[i];
        constexpr int m = n - 1;  // interior knots: the unknown second derivatives
        double A[m][m + 1];
        for (int r = 0; r < m; ++r)
            for (int q = 0; q <= m; ++q) A[r][q] = 0.0;
        for (int k = 0; k < m; ++k) {
            const int i = k + 1;
            if (k > 0) A[k][k - 1] = h[i - 1];
            A[k][k] = 2.0 * (h[i - 1] + h[i]);
            if (k < m - 1) A[k][k + 1] = h[i];
            A[k][m] = 6.0 * ((ys[i + 1] - ys[i]) / h[i] - (ys[i] - ys[i - 1]) / h[i - 1]);
        }
        // Gauss-Jordan elimination to [I | sigma], pivot = largest |entry| at or below the
        // diagonal, first one on ties.
        for (int j = 0; j < m; ++j) {
            int p = j;
            for (int i = j + 1; i < m; ++i)
                if (std::fabs(A[i][j]) > std::fabs(A[p][j])) p = i;
            if (std::fabs(A[p][j]) < 1.0e-16)
                throw std::runtime_error("Gauss-Jordan elimination failed while solving for sigma.");
            if (p != j)
                for (int q = 0; q <= m; ++q) std::swap(A[p][q], A[j][q]);
            const double ajj = A[j][j];
            for (int q = 0; q <= m; ++q) A[j][q] /= ajj;
            for (int i = 0; i < m; ++i) {
                if (i == j) continue;
                const double aij = A[i][j];
                for (int q = 0; q <= m; ++q) A[i][q] -= aij * A[j][q];
            }
        }
        double sigma[n + 1];
        sigma[0] = 0.0;
        sigma[n] = 0.0;
        for (int i = 0; i < m; ++i) sigma[i + 1] = A[i][m];
        for (int i = 0; i < n; ++i) {
            a[i] = (sigma[i + 1] - sigma[i]) / (6.0 * h[i]);
            b[i] = sigma[i] / 2.0;
            c[i] = (ys[i + 1] - ys[i]) / h[i] - (2.0 * sigma[i] + sigma[i + 1]) * h[i] / 6.0;
        }
    }

    double eval(double x) const
    {
        int i = 0;
        if (!(x < xs[0])) {
            i = n - 1;
            while (x < xs[i] && i > 0) i -= 1;
        }
        const double dx = x - xs[i];
        return ((a[i] * dx + b[i]) * dx + c[i]) * dx + ys[i];
    }
};

// Gas and velocity of an inflow ghost cell or face: the primitive six plus p, T, a.
struct State {
    double rho, e, YB, vx, vy, vz, p, T, a;
};

// gas.cu:67-73 (update_from_pT) with the velocity set.
inline State from_pT(double p, double T, double YB, double vx, double vy, double vz)
{
    State s;
    s.p = p;
    s.T = T;
    s.YB = YB;
    s.e = T * kCv - YB * kQ;
    s.rho = p / (T * kR);
    s.a = std::sqrt(kGamma * kR * T);
    s.vx = vx;
    s.vy = vy;
    s.vz = vz;
    return s;
}

struct Profile {
    double p_e = 0.0;
    Spline T, velx;

    // numbers: p_e, then y, T, velx of each sample (the file's order).
    void set(const double* numbers)
    {
        p_e = numbers[0];
        double ys[kSamples], Ts[kSamples], us[kSamples];
        for (int i = 0; i < kSamples; ++i) {
            ys[i] = numbers[1 + 3 * i];
            Ts[i] = numbers[2 + 3 * i];
            us[i] = numbers[3 + 3 * i];
        }
        T.set(ys, Ts);
        velx.set(ys, us);
    }

    // config.cu:73-93
    State eval(double y) const
    {
        double t, u;
        if (y <= T.xs[0]) {
            t = T.ys[0];
            u = velx.ys[0];
        } else if (y >= T.xs[Spline::n]) {
            t = T.ys[Spline::n];
            u = velx.ys[Spline::n];
        } else {
            t = T.eval(y);
            u = velx.eval(y);
        }
        return from_pT(p_e, t, 0.0, u, 0.0, 0.0);
    }
};

// bcs.cu:321-352: the state of inflow function fn (1 supersonic_vortex and 3
// manufactured_solution are fixed placeholder states in the reference; 2 is the profile)
// at position (x, y, z).
inline State inflow_function(int fn, const Profile* prof, double y)
{
    if (fn == 2) return prof->eval(y);
    return from_pT(100.0e3, 300.0, 0.0, 0.0, 0.0, 0.0);
}

}  // namespace bl
}  // namespace chkn
