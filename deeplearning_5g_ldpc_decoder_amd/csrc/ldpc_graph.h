// Compile-time base-graph row layout shared by the HIP kernels: row r of base
// graph BG owns edges [start(r), start(r+1)) of the edge tables in
// nr_bg_tables.h, so loops over a row's edges unroll with constant bounds.
// Include after the HIP runtime header (or the CPU emulation header).
#pragma once
#include "nr_bg_tables.h"

namespace nrldpc {

template <int BG> struct Graph;
template <> struct Graph<1> {
    static constexpr int kRows = NR_BG1_ROWS;
    static constexpr int kEdges = NR_BG1_EDGES;
    __host__ __device__ static constexpr int start(int r) { return kBG1RowStart[r]; }
};
template <> struct Graph<2> {
    static constexpr int kRows = NR_BG2_ROWS;
    static constexpr int kEdges = NR_BG2_EDGES;
    __host__ __device__ static constexpr int start(int r) { return kBG2RowStart[r]; }
};

}  // namespace nrldpc
