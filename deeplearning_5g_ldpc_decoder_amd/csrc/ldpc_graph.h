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
    static constexpr int kInfoCols = 22;   // systematic columns (kb)
    static constexpr int kCols = 68;
    __host__ __device__ static constexpr int start(int r) { return kBG1RowStart[r]; }
    // base-graph column of edge e (compile-time use only)
    __host__ __device__ static constexpr int col(int e) { return kBG1EdgeCol[e]; }
};
template <> struct Graph<2> {
    static constexpr int kRows = NR_BG2_ROWS;
    static constexpr int kEdges = NR_BG2_EDGES;
    static constexpr int kInfoCols = 10;
    static constexpr int kCols = 52;
    __host__ __device__ static constexpr int start(int r) { return kBG2RowStart[r]; }
    __host__ __device__ static constexpr int col(int e) { return kBG2EdgeCol[e]; }
};

// Barrier groups of the layered schedule.  Two rows with disjoint column
// sets touch disjoint beliefs, so running them in either order (or
// interleaved) gives the same beliefs as the reference's row order.
// Consecutive rows are grouped greedily while each new row is column-disjoint
// from every row already in its group; the workgroup barrier is needed only
// between groups.  barrier_after[r]: a barrier follows row r.
template <int BG>
struct RowGroups {
    bool barrier_after[Graph<BG>::kRows];
    int groups;
    static constexpr bool share(int a, int b) {
        for (int e = Graph<BG>::start(a); e < Graph<BG>::start(a + 1); ++e)
            for (int f = Graph<BG>::start(b); f < Graph<BG>::start(b + 1); ++f)
                if (Graph<BG>::col(e) == Graph<BG>::col(f)) return true;
        return false;
    }
    constexpr RowGroups() : barrier_after(), groups(1) {
        int g0 = 0;
        for (int x = 1; x < Graph<BG>::kRows; ++x) {
            bool conflict = false;
            for (int y = g0; y < x && !conflict; ++y) conflict = share(x, y);
            barrier_after[x - 1] = conflict;
            if (conflict) {
                g0 = x;
                ++groups;
            }
        }
        barrier_after[Graph<BG>::kRows - 1] = true;
    }
};
template <int BG> inline constexpr RowGroups<BG> kRowGroups{};
static_assert(kRowGroups<1>.groups == 32 && kRowGroups<2>.groups == 28, "row groups");

// Edge of row r on base-graph column c, or -1 (compile-time).
template <int BG>
__host__ __device__ constexpr int find_edge_ct(int r, int c) {
    for (int e = Graph<BG>::start(r); e < Graph<BG>::start(r + 1); ++e)
        if (Graph<BG>::col(e) == c) return e;
    return -1;
}

}  // namespace nrldpc
