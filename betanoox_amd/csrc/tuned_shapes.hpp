// tuned_shapes.hpp -- default tile shapes of the FP32 SELL SpMV on gfx950.
// Written by `python -m betanoox_amd.autotune --emit-header` from a sweep of
// bhip_tune_spmv over the seeded workload generators; edit by re-running it.
// A row applies to matrices with B DOFs per node and at most max_slices 64-node
// slices; the first matching row wins.  Shape ids are those of bhip_tune_spmv.
#pragma once

struct TunedSellRow {
    int b;
    long long max_slices;
    int shape;
};

constexpr TunedSellRow kTunedSell[] = {
    // B = 3, measured at n = 16, 24, 32, 40, 48, 56, 64, 72, 80, 96, 112 (elastic_cube)
    {3, 777, 2},
    {3, 2306, 3},
    {3, 3524, 1},
    {3, 5107, 2},
    {3, 1LL << 40, 9},
    // B = 2, measured at n = 128, 256, 512, 768, 1024, 1536, 2048 (plane_stress_quad)
    {2, 2061, 0},
    {2, 6165, 9},
    {2, 24617, 3},
    {2, 1LL << 40, 9},
    // B = 1, measured at n = 32, 48, 64, 80, 96, 112, 128, 160 (diffusion_cube)
    {1, 27500, 0},
    {1, 1LL << 40, 9},
};
