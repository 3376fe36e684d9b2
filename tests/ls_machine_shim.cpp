// Host build of the device line search's state machine (redux_amd/csrc/line_search.hpp) for
// its CPU tests: one line driven to its end on a function given as a callback, recording
// every probe the machine asks for.
#include <cmath>

#include "line_search.hpp"

extern "C" int ls_trace(double fa, double lambda, double (*f)(double), double* steps, int cap, double* out) {
    LsState S{};
    S.pc = 0;
    S.fa0 = fa;
    S.lambda = lambda;
    double fv = 0.0, stp = 0.0;
    int n = 0;
    while (ls_advance(S, fv, &stp)) {
        if (n == cap) return -1;
        steps[n++] = stp;
        fv = f(stp);
    }
    out[0] = S.stepOut;
    out[1] = S.fOut;
    return n;
}
