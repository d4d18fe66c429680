// BatchLayout command line, a drop-in for the reference's Test/BatchLayout.cpp:26-188.
//
// Same flags, defaults and quirks (BatchLayout.cpp:57-129): -init is clamped to 1 when out
// of [0,1]; -initf forces file initialisation; -engt outside [0,1] resets to 0.01 and any
// value other than 0.01 selects the converged algo 2; -algo outside [0,8] becomes 2; the
// thread count is capped like the reference's.  Same stdout lines (a wrapper reads the
// wall-time line and the output-path line by position), same output file names, same
// Results.txt row.  Layouts run on the GPU through libbatchlayout.so.
//
// Build-only flags use a prefix the reference never parses:
//   --gpu-precision mixed|fp32|fp64   pair arithmetic (default mixed: FP64 with an FP32-seeded
//                               reciprocal; fp32: FP32 |d|^2 and reciprocal; fp64: the
//                               reference's FP64 operation order)
// Algorithms 0, 1, 5 and 7 (sequential / naive / greedy / algebraic baselines of the
// reference) are not provided by this build; they exit with status 1 and a message.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <fstream>
#include <iostream>
#include <string>
#include <thread>
#include <vector>

#include "algorithms.hpp"

using bl::algorithms;

static void helpmessage() {
    printf("\n");
    printf("Usage of BatchLayout tool:\n");
    printf("-input <string>, full path of input file (required).\n");
    printf("-output <string>, directory where output file will be stored.\n");
    printf("-batch <int>, size of minibatch.\n");
    printf("-init <int>, any of 0 or 1, 1 - random initialization, 0 - greedy initialization.\n");
    printf("-initf <string>, this will overwrite -init and initialize coordinates from a given file.\n");
    printf("-iter <int>, number of iteration.\n");
    printf("-threads <int>, number of threads, default value is maximum available threads in the machine.\n");
    printf("-algo <int>, an integer among 0, 1, 2, 3, 4, 5 and 6.\n");
    printf("        2 - for parallel layout generation using cache blocking stochastic minibatch update.\n");
    printf("        3 - for parallel layout generation using linlog mode, (0,-1)-energy model.\n");
    printf("        4 - for parallel layout generation using barnes-hut repulsive force approximation.\n");
    printf("        6 - for parallel layout generation, 80%% time in -algo 4 and last 20%% time in -algo 2.\n");
    printf("        8 - for parallel layout generation, (1,-1)-energy model.\n");
    printf("        (0, 1, 5 and 7 are not provided by this GPU build)\n");
    printf("-bht <float>, barnes-hut threshold parameter.\n");
    printf("-engt <float>, convergenece criteria. It should be a floating point number between 0 to 1 exclusive.\n");
    printf("It indicates if energy value is improved less than this percenrages then optimization will stop and "
           "return layout. If this value is set, then number of iteration will be overwritten.\n");
    printf("-weight <float>, a real number to include edge weight.\n");
    printf("--gpu-precision <mixed|fp32|fp64>, pair arithmetic on the GPU (build extension).\n");
    printf("-h, show help message.\n");
    printf("default: -weight 1.0 -bht 1.2 -engt 0.01 -batch 256 -iter 600 -threads MAX -algo 2 -init 0\n\n");
}

int main(int argc, char *argv[]) {
    double energyThreshold = 0.01, bhThreshold = 1.2, weight = 1.0;
    const int maxThreads = (int)std::max(1u, std::thread::hardware_concurrency());
    int init = 0, batchsize = 256, iterations = 600, numberOfThreads = maxThreads, algoOption = 2;
    int precision = BL_PREC_MIXED;
    std::string inputfile, initfile, outputfile, algoname = "CACHE", initname = "GREEDY";
    for (int p = 0; p < argc; p++) {
        const bool has = p + 1 < argc;
        if (strcmp(argv[p], "-input") == 0 && has) inputfile = argv[p + 1];
        if (strcmp(argv[p], "-output") == 0 && has) outputfile = argv[p + 1];
        if (strcmp(argv[p], "-batch") == 0 && has) batchsize = atoi(argv[p + 1]);
        if (strcmp(argv[p], "-init") == 0 && init != 2 && has) {
            init = atoi(argv[p + 1]);
            if (init < 0 || init > 1) init = 1;
        }
        if (strcmp(argv[p], "-iter") == 0 && has) iterations = atoi(argv[p + 1]);
        if (strcmp(argv[p], "-threads") == 0 && has) numberOfThreads = atoi(argv[p + 1]);
        if (strcmp(argv[p], "-algo") == 0 && has) algoOption = atoi(argv[p + 1]);
        if (strcmp(argv[p], "-bht") == 0 && has) bhThreshold = atof(argv[p + 1]);
        if (strcmp(argv[p], "-initf") == 0 && has) {
            initfile = argv[p + 1];
            init = 2;
        }
        if (strcmp(argv[p], "-engt") == 0 && has) {
            energyThreshold = atof(argv[p + 1]);
            if (energyThreshold < 0.0 || energyThreshold > 1.0) energyThreshold = 0.01;
        }
        if (strcmp(argv[p], "-weight") == 0 && has) weight = atof(argv[p + 1]);
        if (strcmp(argv[p], "--gpu-precision") == 0 && has) {
            if (strcmp(argv[p + 1], "fp64") == 0) precision = BL_PREC_FP64;
            else if (strcmp(argv[p + 1], "fp32") == 0) precision = BL_PREC_FP32;
            else precision = BL_PREC_MIXED;
        }
        if (strcmp(argv[p], "-h") == 0) {
            helpmessage();
            exit(1);
        }
    }
    if (inputfile.size() < 1 || argc < 2) {
        helpmessage();
        exit(1);
    }
    if (init == 1) initname = "RANDOM";
    if (init == 2) initname = "FILE";
    printf("\n");
    printf("@@@@@                   @      @                                     \n");
    printf("@    @                  @      @                                     \n");
    printf("@  @   @@@@@@ @  @@@@@@ @      @      @@@@@@ @    @ @@@@@@ @    @ @  \n");
    printf("@@@         @ @@ @    @ @@@@@@ @           @ @    @ @    @ @    @ @@ \n");
    printf("@  @   @@@@@@ @  @      @    @ @      @@@@@@ @@@@@@ @    @ @    @ @  \n");
    printf("@    @ @    @ @  @      @    @ @      @    @      @ @    @ @    @ @  \n");
    printf("@@@@@  @@@@@@ @@ @@@@@@ @    @ @@@@@@ @@@@@@ @@@@@@ @@@@@@ @@@@@@ @@ \n\n");
    fflush(stdout);
    bl::CSRGraph A_csr;
    const int rc = bl::SetInputMatricesAsCSR(A_csr, inputfile);
    if (rc != BL_OK) {
        fprintf(stderr, "batchlayout: cannot read %s (status %d)\n", inputfile.c_str(), rc);
        return 1;
    }
    if (algoOption < 0 || algoOption > 8) algoOption = 2;
    if (maxThreads < numberOfThreads) {
        numberOfThreads = maxThreads;
        if (A_csr.rows < 1000000 && numberOfThreads > 32) numberOfThreads = 32;
    }
    if (algoOption == 0 || algoOption == 1 || algoOption == 5 || algoOption == 7) {
        fprintf(stderr, "batchlayout: -algo %d is not provided by this GPU build (use 2, 3, 4, 6 or 8)\n",
                algoOption);
        return 1;
    }
    std::vector<double> outputvec;
    algorithms algo(A_csr, inputfile, outputfile, init, weight, energyThreshold, initfile);
    algo.precision = precision;
    if (algo.status != BL_OK) return 1;
    if (algoOption == 2) {
        algoname = "CACHE";
        if (fabs(energyThreshold - 0.01) > 0)
            outputvec = algo.cacheBlockingminiBatchForceDirectedAlgorithmConverged(iterations, numberOfThreads,
                                                                                   batchsize, 0);
        else
            outputvec = algo.cacheBlockingminiBatchForceDirectedAlgorithm(iterations, numberOfThreads, batchsize, 0);
    } else if (algoOption == 3) {
        algoname = "LINLOG";
        outputvec = algo.LinLogcacheBlockingminiBatchForceDirectedAlgorithm(iterations, numberOfThreads, batchsize);
    } else if (algoOption == 4) {
        algoname = "BHCACHE";
        outputvec = algo.BarnesHutApproximation(iterations, numberOfThreads, batchsize, bhThreshold, 0);
    } else if (algoOption == 6) {
        algoname = "BCGRDY";
        outputvec = algo.approxCacheBlockBH(iterations, numberOfThreads, batchsize);
    } else if (algoOption == 8) {
        algoname = "FA";
        outputvec = algo.FAcacheBlockingminiBatchForceDirectedAlgorithm(iterations, numberOfThreads, batchsize);
    }
    if (algo.status != BL_OK) return 1;
    std::ofstream output("Results.txt", std::ofstream::app);
    output << algoname << "\t" << initname << "\t";
    output << iterations << "\t" << numberOfThreads << "\t" << batchsize << "\t";
    output << outputvec[0] << "\t" << outputvec[1];
    output << std::endl;
    return 0;
}
