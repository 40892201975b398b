// pp_xcd.h -- XCD-aware block numbering for gfx950 (8 XCDs, one L2 each).
#pragma once
#include <hip/hip_runtime.h>

namespace pp {

// Blocks are dealt to the 8 XCDs round-robin by linear id, and each XCD has its own L2.
// Renumbering so that each XCD's blocks take consecutive logical indices puts neighbouring
// tiles -- which share cache lines when a tile row is shorter than 128 bytes -- on one XCD.
__device__ __forceinline__ int xcd_logical(int lin, int total)
{
    constexpr int X = 8;
    const int q = total / X, r = total % X;
    const int xcd = lin % X, k = lin / X;
    return (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + k;
}

}  // namespace pp
