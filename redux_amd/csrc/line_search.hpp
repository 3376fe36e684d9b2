// The conjugate-gradient line search as a resumable state machine, compiled for the device
// (one thread per patch in a round of probes) and, for its tests, for the host.  The
// includer provides __host__ / __device__ (hip_runtime.h) and <cmath>.
#pragma once

// One patch's math::bracket + math::brent (helpers.cpp:58-216) as conjugate_rdx_iterate
// calls them (conjugate_gradient.cpp:196-231: the first probe at 1e-21 / lambda, bracket
// with limit 100, brent with tol 0.1 and limit 100), written as a state machine that stops
// at every function value it needs: the probe x - stp lambda p, i.e. the line's phases at
// s = -stp lambda, is evaluated between two steps and its metric read on resumption.  The
// arithmetic is the host driver's, operation for operation (no contraction), so a device
// line search takes the same probes and ends at the same point.
struct LsState {
    int pc;                 // resume point; 0 = a new line (fa and lambda set), -1 = idle / done
    int iter;               // brent's iteration
    double lambda, fa0;
    double a, b, mid, fa, fb, fmid, u, fu, ulim;     // bracket
    double ba, bb, x, w, v, fx, fw, fv, d, e;        // brent
    double stepOut, fOut;                            // the line's end: step and metric
    int nProbe;                                      // probes (metric evaluations) so far
};

__host__ __device__ inline double ls_sign(double a, double b) { return b >= 0.0 ? fabs(a) : -fabs(a); }

// Advance the state machine with f (the metric of the last probe); returns 1 and sets *stp
// when the next probe is wanted, 0 when the line search has ended (stepOut, fOut).
__host__ __device__ inline int ls_advance(LsState& S, double f, double* stp) {
#pragma clang fp contract(off)
    constexpr double GOLD = 1.6180339887498949025, TINYV = 1.0E-32, CGOLD = 2.0 - 1.6180339887498949025;
    constexpr int LIMIT = 100;
    constexpr double TOL = 0.1;
    switch (S.pc) {
    case 0:                                 // conjugate_gradient.cpp:196-198: fc at stepc
        S.fa = S.fa0;
        S.a = 0.0;
        S.b = 1E-21 / S.lambda;
        S.mid = 0.0;
        S.fmid = 0.0;
        S.pc = 1;
        *stp = S.b;
        return 1;
    case 1:                                 // bracket(a = 0, b = stepc)
        S.fb = f;
        if (S.fb > S.fa) {
            double t = S.a; S.a = S.b; S.b = t;
            t = S.fa; S.fa = S.fb; S.fb = t;
        }
        S.mid = S.b + GOLD * (S.b - S.a);
        S.pc = 2;
        *stp = S.mid;
        return 1;
    case 2:
        S.fmid = f;
        goto loop;
    case 3:
        S.fu = f;
        if (S.fu < S.fmid) {
            S.a = S.b; S.b = S.u; S.fa = S.fb; S.fb = S.fu;
            goto bracket_end;
        } else if (S.fu > S.fb) {
            S.mid = S.u; S.fmid = S.fu;
            goto bracket_end;
        }
        S.u = S.mid + GOLD * (S.mid - S.b);
        S.pc = 7;
        *stp = S.u;
        return 1;
    case 5:
        S.fu = f;
        if (S.fu < S.fmid) {                // SHFT(b, mid, u, mid + G (mid - b)), then f(u)
            S.b = S.mid;
            S.mid = S.u;
            S.u = S.mid + GOLD * (S.mid - S.b);
            S.pc = 6;
            *stp = S.u;
            return 1;
        }
        goto shift;
    case 6:                                 // SHFT(fb, fmid, fu, f(u))
        S.fb = S.fmid;
        S.fmid = S.fu;
        S.fu = f;
        goto shift;
    case 7:
        S.fu = f;
        goto shift;
    case 10:                                // brent's first value
        S.fx = f;
        S.fu = S.fv = S.fw = S.fx;
        S.d = 0.0;
        S.e = 0.0;
        S.iter = 1;
        goto brent_loop;
    case 11:
        S.fu = f;
        if (S.fu <= S.fx) {
            if (S.u >= S.x) S.ba = S.x; else S.bb = S.x;
            S.v = S.w; S.w = S.x; S.x = S.u;
            S.fv = S.fw; S.fw = S.fx; S.fx = S.fu;
        } else {
            if (S.u < S.x) S.ba = S.u; else S.bb = S.u;
            if (S.fu <= S.fw || S.w == S.x) { S.v = S.w; S.w = S.u; S.fv = S.fw; S.fw = S.fu; }
            else if (S.fu <= S.fv || S.v == S.x || S.v == S.w) { S.v = S.u; S.fv = S.fu; }
        }
        ++S.iter;
        goto brent_loop;
    default:
        return 0;
    }
shift:                                      // SHFT(a, b, mid, u), SHFT(fa, fb, fmid, fu)
    S.a = S.b; S.b = S.mid; S.mid = S.u;
    S.fa = S.fb; S.fb = S.fmid; S.fmid = S.fu;
loop:
    if (S.fb > S.fmid) {
        const double rr = (S.b - S.a) * (S.fb - S.fmid);
        const double q = (S.b - S.mid) * (S.fb - S.fa);
        const double m1 = fabs(q - rr);
        S.u = S.b - ((S.b - S.mid) * q - (S.b - S.a) * rr) / (2.0 * ls_sign(m1 > TINYV ? m1 : TINYV, q - rr));
        S.ulim = S.b + LIMIT * (S.mid - S.b);
        if ((S.b - S.u) * (S.u - S.mid) > 0.0) {
            S.pc = 3;
        } else if ((S.mid - S.u) * (S.u - S.ulim) > 0.0) {
            S.pc = 5;
        } else if ((S.mid - S.ulim) * (S.ulim - S.u) >= 0.0) {
            S.u = S.ulim;
            S.pc = 7;
        } else {
            S.u = S.mid + GOLD * (S.mid - S.b);
            S.pc = 7;
        }
        *stp = S.u;
        return 1;
    }
bracket_end:
    if (S.a > S.b) { double t = S.a; S.a = S.b; S.b = t; t = S.fa; S.fa = S.fb; S.fb = t; }
    if (S.mid > S.b) { double t = S.mid; S.mid = S.b; S.b = t; t = S.fmid; S.fmid = S.fb; S.fb = t; }
    if (S.mid < S.a) { double t = S.mid; S.mid = S.a; S.a = t; t = S.fmid; S.fmid = S.fa; S.fa = t; }
    // brent(a, b, mid, tol 0.1, limit 100)
    S.ba = S.a < S.b ? S.a : S.b;
    S.bb = S.a > S.b ? S.a : S.b;
    S.x = S.mid;
    S.u = S.v = S.w = S.x;
    S.pc = 10;
    *stp = S.x;
    return 1;
brent_loop:
    if (S.iter > LIMIT) goto finished;
    {
        const double xm = 0.5 * (S.ba + S.bb);
        const double tol1 = TOL * fabs(S.x) + TINYV, tol2 = 2.0 * tol1;
        if (fabs(S.x - xm) <= (tol2 - 0.5 * (S.bb - S.ba))) goto finished;
        if (fabs(S.e) > tol1) {
            const double rr = (S.x - S.w) * (S.fx - S.fv);
            double q = (S.x - S.v) * (S.fx - S.fw);
            double pp = (S.x - S.v) * q - (S.x - S.w) * rr;
            q = 2.0 * (q - rr);
            if (q > 0.0) pp = -pp;
            q = fabs(q);
            const double etemp = S.e;
            S.e = S.d;
            if (fabs(pp) >= fabs(0.5 * q * etemp) || pp <= q * (S.ba - S.x) || pp >= q * (S.bb - S.x)) {
                S.e = (S.x >= xm ? S.ba - S.x : S.bb - S.x);
                S.d = CGOLD * S.e;
            } else {
                S.d = pp / q;
                S.u = S.x + S.d;
                if (S.u - S.ba < tol2 || S.bb - S.u < tol2) S.d = ls_sign(tol1, xm - S.x);
            }
        } else {
            S.e = (S.x >= xm ? S.ba - S.x : S.bb - S.x);
            S.d = CGOLD * S.e;
        }
        if (fabs(S.d) < tol1) S.d *= (tol1 / fabs(S.d));
        S.u = ((fabs(S.d) >= tol1) ? S.x + S.d : S.x + ls_sign(tol1, S.d));
        S.pc = 11;
        *stp = S.u;
        return 1;
    }
finished:
    S.stepOut = S.x;
    S.fOut = S.fx;
    S.pc = -1;
    return 0;
}
