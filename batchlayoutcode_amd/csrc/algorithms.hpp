// C++ host mirror of the reference's `algorithms` class (sample/algorithms.h:50-85),
// implemented on the C ABI of libbatchlayout.so.  Same constructor arguments, method
// names, return values ({ENERGY, seconds}), stdout lines and output-file names as the
// reference, so code written against the reference class ports by changing the include.
// The layouts run on the GPU; NUMOFTHREADS is accepted and ignored.
#pragma once

#include <string>
#include <vector>

#include "batchlayout.h"

namespace bl {

// CSR as built by the reference's SetInputMatricesAsCSR (sample/commonutility.h:44-53).
struct CSRGraph {
    int32_t rows = 0;
    int64_t nnz = 0;
    std::vector<int32_t> rowptr, colids;
};

// Reads the .mtx exactly like SetInputMatricesAsCSR and prints the reference's messages
// (IO.h:60-155, commonutility.h:25-41).  Returns a BL_E_* status.
int SetInputMatricesAsCSR(CSRGraph &g, const std::string &inputfile);

class algorithms {
   public:
    CSRGraph graph;
    std::vector<double> nCoordinates;  // n x 2, FP64, as Coordinate<double>[n]
    double W, threshold;
    int init;
    std::string filename, outputdir, initfile;
    int precision = BL_PREC_MIXED;     // GPU pair arithmetic (build extension)
    int status = BL_OK;                // last GPU status (the reference has none)

    // algorithms.cpp:3-15
    algorithms(CSRGraph &A_csr, std::string input, std::string outputd, int init, double weight, double th,
               std::string ifile);
    ~algorithms();
    algorithms(const algorithms &) = delete;
    algorithms &operator=(const algorithms &) = delete;

    void randInit();            // algorithms.cpp:16-28 (sequential: the -threads 1 sequence)
    void initDFS();             // algorithms.cpp:29-97
    void fileInitialization();  // algorithms.cpp:98-119

    std::vector<double> cacheBlockingminiBatchForceDirectedAlgorithm(int ITERATIONS, int NUMOFTHREADS, int BATCHSIZE,
                                                                     int flag = 0);
    std::vector<double> cacheBlockingminiBatchForceDirectedAlgorithmConverged(int ITERATIONS, int NUMOFTHREADS,
                                                                              int BATCHSIZE, int flag = 0);
    std::vector<double> LinLogcacheBlockingminiBatchForceDirectedAlgorithm(int ITERATIONS, int NUMOFTHREADS,
                                                                           int BATCHSIZE);
    std::vector<double> FAcacheBlockingminiBatchForceDirectedAlgorithm(int ITERATIONS, int NUMOFTHREADS,
                                                                       int BATCHSIZE);
    std::vector<double> BarnesHutApproximation(int ITERATIONS, int NUMOFTHREADS, int BATCHSIZE, double TH,
                                               int flag = 0);
    std::vector<double> approxCacheBlockBH(int ITERATIONS, int NUMOFTHREADS, int BATCHSIZE);
    void writeToFile(const std::string &f);  // algorithms.cpp:1244-1256

   private:
    bl_graph *g_ = nullptr;
    void initialize();
    int gpu_run(int algo, int iters, int batch, double bht, int converged, double step0, bl_stats *st);
};

}  // namespace bl
