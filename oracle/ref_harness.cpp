// Fixture harness over the reference's own sources (test infrastructure only).
//
// Compiled by oracle/Makefile together with sample/algorithms.cpp from the
// reference checkout into oracle/_ref/ (never committed, never shipped).  It
// drives the reference's `algorithms` class and `BarnesHut`/`MortonCode`
// exactly as Test/BatchLayout.cpp:120-174 does, and prints the state at full
// precision (%.17g) instead of the 6-digit output files, so that the oracle
// under oracle/ can be pinned against the real reference.
//
// Modes:
//   run  <mtx> <initf> <algo> <iters> <batch> <bht> <weight> <engt> <threads>
//        -> "E <energy>" then n lines "x y" (%.17g) of nCoordinates
//   bh   <pointsfile> <theta>
//        -> Morton scales, sorted permutation, tree (pre-order), forces
//   csr  <mtx>
//        -> rows nnz, rowptr, colids of the reference's CSR
#include <omp.h>
#include <stdio.h>
#include <stdint.h>
#include <iostream>
#include <algorithm>
#include <numeric>
#include <functional>
#include <fstream>
#include <iterator>
#include <ctime>
#include <cstring>
#include <cmath>
#include <string>
#include <sstream>
#include <random>
#include <map>
#include <unordered_map>
#include <vector>
#include <stack>
#include <array>
#include <limits>
#include <cassert>
#include <tuple>
#include <ostream>
#include <parallel/algorithm>
#include <parallel/numeric>
#include <parallel/tags.h>

// The quadtree root is a private member of BarnesHut (BarnesHut.h:112); the
// harness needs it to dump the topology at full precision.
#define private public
#include "sample/algorithms.h"
#undef private

using namespace std;

static void dump_csr(const char *mtx) {
    CSR<int, double> A;
    SetInputMatricesAsCSR(A, string(mtx));
    printf("CSR %d %d\n", A.rows, A.nnz);
    for (int i = 0; i <= A.rows; i++) printf("%d\n", A.rowptr[i]);
    for (int k = 0; k < A.nnz; k++) printf("%d\n", A.colids[k]);
}

static void run_algo(int argc, char **argv) {
    string mtx = argv[2], initf = argv[3];
    int algo = atoi(argv[4]), iters = atoi(argv[5]), batch = atoi(argv[6]);
    double bht = atof(argv[7]), weight = atof(argv[8]), engt = atof(argv[9]);
    int threads = atoi(argv[10]);
    CSR<int, double> A;
    SetInputMatricesAsCSR(A, mtx);
    A.Sorted();
    algorithms al(A, mtx, string("/tmp/"), 2, weight, engt, initf);
    vector<double> out;
    if (algo == 2) out = al.cacheBlockingminiBatchForceDirectedAlgorithm(iters, threads, batch, 0);
    else if (algo == 21) out = al.cacheBlockingminiBatchForceDirectedAlgorithmConverged(iters, threads, batch, 0);
    else if (algo == 3) out = al.LinLogcacheBlockingminiBatchForceDirectedAlgorithm(iters, threads, batch);
    else if (algo == 4) out = al.BarnesHutApproximation(iters, threads, batch, bht, 0);
    else if (algo == 6) out = al.approxCacheBlockBH(iters, threads, batch);
    else if (algo == 8) out = al.FAcacheBlockingminiBatchForceDirectedAlgorithm(iters, threads, batch);
    else { fprintf(stderr, "unknown algo\n"); exit(2); }
    printf("E %.17g\n", out[0]);
    for (int i = 0; i < A.rows; i++)
        printf("%.17g %.17g\n", al.nCoordinates[i].getX(), al.nCoordinates[i].getY());
}

struct PtKey { uint64_t a, b; bool operator<(const PtKey &o) const { return a < o.a || (a == o.a && b < o.b); } };
static PtKey key_of(double x, double y) { PtKey k; memcpy(&k.a, &x, 8); memcpy(&k.b, &y, 8); return k; }

static map<PtKey, int> *g_index;
static int g_next_node;

// Pre-order dump of the reference's (uncompressed) quadtree.
//   N <id> <parent> <quadrant> <diameter> <cx> <cy>
//   L <parent> <quadrant> <point index>
static void dump_node(I_Node_2 *nd, int parent, int quad) {
    int id = g_next_node++;
    printf("N %d %d %d %.17g %.17g %.17g\n", id, parent, quad, nd->diameter, nd->getX(), nd->getY());
    for (int q = 0; q < 4; q++) {
        if (nd->_i_Coordinates[q] != NULL) dump_node(nd->_i_Coordinates[q], id, q);
        else if (nd->_Coordinates[q] != NULL) {
            Coordinate<double> *p = nd->_Coordinates[q];
            printf("L %d %d %d\n", id, q, (*g_index)[key_of(p->getX(), p->getY())]);
        }
    }
}

static void run_bh(int argc, char **argv) {
    FILE *f = fopen(argv[2], "r");
    double theta = atof(argv[3]);
    vector<Coordinate<double> > pts;
    double x, y;
    while (fscanf(f, "%lf %lf", &x, &y) == 2) pts.push_back(Coordinate<double>(x, y));
    fclose(f);
    int n = (int)pts.size();
    map<PtKey, int> index;
    for (int i = 0; i < n; i++) index[key_of(pts[i].getX(), pts[i].getY())] = i;
    g_index = &index;
    vector<Coordinate<double> > work(pts);
    {
        MortonCode mc(work.data(), n);
        printf("M %.17g\n", mc.gridLength(1));
        for (int i = 0; i < n; i++)
            printf("S %u %u\n", mc.xscale(pts[i].getX()), mc.yscale(pts[i].getY()));
    }
    BarnesHut bh(work.data(), n, theta);
    for (int i = 0; i < n; i++) printf("P %d\n", index[key_of(work[i].getX(), work[i].getY())]);
    g_next_node = 0;
    dump_node(bh.quadtree, -1, -1);
    for (int i = 0; i < n; i++) {
        Coordinate<double> fr = bh.calcRepForce(pts[i]);
        printf("F %.17g %.17g\n", fr.getX(), fr.getY());
    }
}

int main(int argc, char **argv) {
    if (argc >= 3 && strcmp(argv[1], "csr") == 0) dump_csr(argv[2]);
    else if (argc >= 11 && strcmp(argv[1], "run") == 0) run_algo(argc, argv);
    else if (argc >= 4 && strcmp(argv[1], "bh") == 0) run_bh(argc, argv);
    else { fprintf(stderr, "usage: see header\n"); return 2; }
    return 0;
}
