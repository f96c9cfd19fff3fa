// Application driver: the time loop of the reference application
// (applications/solvers/compressible/CUDA/cudaEuler2DmixO3RK3solver/
// cudaEuler2DmixO3RK3solver.cpp:18-88): setDeltaT when adjustTimeStep,
// advance the time, evolve, write at write times, log every 10 iterations.
//
//   cudaEuler2DmixO3RK3solver -case <dir>
//
// <dir>/system/{controlDict, fvSchemes, afarMeshDict}; fields are written to
// <dir>/<time>/{p, U, T, shockIndicator}.
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>

#include "basicCUDAsolver.h"

int main(int argc, char* argv[]) {
    std::string caseDir = ".";
    for (int i = 1; i < argc; ++i)
        if (!strcmp(argv[i], "-case") && i + 1 < argc) caseDir = argv[++i];
    try {
        afarhost::runTime runTime(caseDir);
        afarhost::caseMesh mesh(caseDir);
        const bool adjustTimeStep = runTime.controlDict().lookupBool("adjustTimeStep");
        printf("creating cudaSolver...\n");
        afarhost::basicCUDAsolver cudaSolver(runTime, mesh);
        printf("create cudaSolver completed.\n\n");
        unsigned ITER = 0;
        const auto t0 = std::chrono::steady_clock::now();
        auto elapsed = [&]() {
            return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        };
        printf("\nITER    Time[s]    deltaT[s]    ExecutionTime[s]\n");
        while (runTime.run()) {
            if (adjustTimeStep) cudaSolver.setDeltaT(runTime);
            ++runTime;
            ++ITER;
            cudaSolver.evolve(runTime.deltaT());
            if (runTime.write()) cudaSolver.writeField(runTime);
            if (ITER % 200 == 0) printf("\nITER    Time[s]    deltaT[s]    ExecutionTime[s]\n");
            if (ITER % 10 == 0)
                printf("%u    %s    %g    %g\n", ITER, runTime.timeName().c_str(), runTime.deltaT(), elapsed());
        }
        printf("\nevolve operation completed.\n");
        printf("ITER    Time[s]    deltaT[s]    ExecutionTime[s]\n");
        printf("%u    %s    %g    %g\n", ITER, runTime.timeName().c_str(), runTime.deltaT(), elapsed());
    } catch (const std::exception& e) {
        fprintf(stderr, "Error: %s\n", e.what());
        return 1;
    }
    return 0;
}
