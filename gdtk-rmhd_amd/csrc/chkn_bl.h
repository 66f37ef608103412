// chkn_bl.h -- inflow functions of the inflow_function boundary condition (host side).
//
// The ghost-cell states of an inflow_function side depend only on the cell positions, so the
// host evaluates them once at set-up (bcs.cu:321-352 per ghost cell, and per boundary face
// for the viscous boundary condition, bcs.cu:790-792).  The laminar boundary-layer profile is
// a natural cubic spline in y through 12 samples of T and velx (spline.cu:20-84, with the
// second derivatives from Gauss-Jordan elimination with partial pivoting, rsla.cu:196-227),
// clamped to the end samples outside their range (config.cu:73-93).  All arithmetic keeps
// the reference's operation order and, in the FP32 flavour, its types: the two matrix
// entries written with double literals, the pivot test against a double threshold and the
// segment polynomial are evaluated in double as the reference's FP32 build evaluates them.
#pragma once

#include <cmath>
#include <stdexcept>
#include <utility>

#include "chkn_physics.h"

namespace chkn {
namespace bl {

constexpr int kSamples = 12;  // config.cu:31
constexpr int kProfileValues = 1 + 3 * kSamples;  // p_e, then y, T, velx of each sample

// Natural cubic spline through n+1 = kSamples knots (spline.cu:20-109).
struct Spline {
    static constexpr int n = kSamples - 1;
    real xs[n + 1], ys[n + 1], a[n], b[n], c[n];

    void set(const real* x, const real* y)
    {
        for (int i = 0; i <= n; ++i) { xs[i] = x[i]; ys[i] = y[i]; }
        real h[n];
        for (int i = 0; i < n; ++i) h[i] = xs[i + 1] - xs[i];
        constexpr int m = n - 1;  // interior knots: the unknown second derivatives
        real A[m][m + 1];
        for (int r = 0; r < m; ++r)
            for (int q = 0; q <= m; ++q) A[r][q] = real(0);
        for (int k = 0; k < m; ++k) {
            const int i = k + 1;
            if (k > 0) A[k][k - 1] = h[i - 1];
            A[k][k] = (real)(2.0 * (double)(h[i - 1] + h[i]));
            if (k < m - 1) A[k][k + 1] = h[i];
            A[k][m] = (real)(6.0 * (double)((ys[i + 1] - ys[i]) / h[i] - (ys[i] - ys[i - 1]) / h[i - 1]));
        }
        // Gauss-Jordan elimination to [I | sigma], pivot = largest |entry| at or below the
        // diagonal, first one on ties.
        for (int j = 0; j < m; ++j) {
            int p = j;
            for (int i = j + 1; i < m; ++i)
                if (std::fabs(A[i][j]) > std::fabs(A[p][j])) p = i;
            if ((double)std::fabs(A[p][j]) < 1.0e-16)
                throw std::runtime_error("Gauss-Jordan elimination failed while solving for sigma.");
            if (p != j)
                for (int q = 0; q <= m; ++q) std::swap(A[p][q], A[j][q]);
            const real ajj = A[j][j];
            for (int q = 0; q <= m; ++q) A[j][q] /= ajj;
            for (int i = 0; i < m; ++i) {
                if (i == j) continue;
                const real aij = A[i][j];
                for (int q = 0; q <= m; ++q) A[i][q] -= aij * A[j][q];
            }
        }
        real sigma[n + 1];
        sigma[0] = real(0);
        sigma[n] = real(0);
        for (int i = 0; i < m; ++i) sigma[i + 1] = A[i][m];
        for (int i = 0; i < n; ++i) {
            a[i] = (sigma[i + 1] - sigma[i]) / (real(6) * h[i]);
            b[i] = sigma[i] / real(2);
            c[i] = (ys[i + 1] - ys[i]) / h[i] - (real(2) * sigma[i] + sigma[i + 1]) * h[i] / real(6);
        }
    }

    real eval(real x) const
    {
        int i = 0;
        if (!(x < xs[0])) {
            i = n - 1;
            while (x < xs[i] && i > 0) i -= 1;
        }
        const double dx = (double)(x - xs[i]);  // spline.cu:94-95 evaluates the piece in double
        return (real)((((double)a[i] * dx + (double)b[i]) * dx + (double)c[i]) * dx + (double)ys[i]);
    }
};

// Gas and velocity of an inflow ghost cell or face: the primitive six plus p, T, a.
struct State {
    real rho, e, YB, vx, vy, vz, p, T, a;
};

// gas.cu:67-73 (update_from_pT) with the velocity set.
inline State from_pT(real p, real T, real YB, real vx, real vy, real vz)
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
    real p_e = real(0);
    Spline T, velx;

    // numbers: p_e, then y, T, velx of each sample (the file's order).
    void set(const real* numbers)
    {
        p_e = numbers[0];
        real ys[kSamples], Ts[kSamples], us[kSamples];
        for (int i = 0; i < kSamples; ++i) {
            ys[i] = numbers[1 + 3 * i];
            Ts[i] = numbers[2 + 3 * i];
            us[i] = numbers[3 + 3 * i];
        }
        T.set(ys, Ts);
        velx.set(ys, us);
    }

    // config.cu:73-93
    State eval(real y) const
    {
        real t, u;
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
        return from_pT(p_e, t, real(0), u, real(0), real(0));
    }
};

// bcs.cu:321-352: the state of inflow function fn (1 supersonic_vortex and 3
// manufactured_solution are fixed placeholder states in the reference; 2 is the profile)
// at position (x, y, z).
inline State inflow_function(int fn, const Profile* prof, real y)
{
    if (fn == 2) return prof->eval(y);
    return from_pT(real(100.0e3), real(300.0), real(0), real(0), real(0), real(0));
}

}  // namespace bl
}  // namespace chkn
