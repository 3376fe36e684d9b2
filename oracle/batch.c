/*
 * Oracle (test infrastructure only): run o_solve over several independent
 * patches on host threads -- the CPU baseline decomposition "patches in
 * parallel across cores" (the reference farms one patch per worker,
 * momfbdjob.cpp:363-377).
 */
#include "mfbd_oracle.h"

#include <omp.h>

int o_solve_many(const OProblem* prs, OResult* res, int n, int nThreads)
{
    int rc = 0;
    if (nThreads > 0) omp_set_num_threads(nThreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(| : rc)
    for (int i = 0; i < n; ++i) rc |= o_solve(&prs[i], &res[i]);
    return rc;
}
