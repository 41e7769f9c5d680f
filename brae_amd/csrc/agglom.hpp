// Host-side pairwise agglomeration: the coarse hierarchy's ADDRESSING, built once per mesh on the CPU
// (it is a greedy, order-dependent walk over the faces) and uploaded. Two numberings, each bit-exact to
// its reference:
//   Brae      brae's device-path hierarchy (device_amg.cu:533-685 agglomerate, :444-522
//             coarseFacesByBucket, :1346-1431 buildAMG): forward greedy pairing, coarse faces as sorted
//             unique (lo, hi) pairs, coarsen until n <= target cells.
//   OpenFOAM  faceAreaPair (pair_gamg_agglomeration_cpp.cu:28-345): perturbed weights supplied by the
//             caller, direction flag alternating every pass, coarse faces owner-major in first-appearance
//             order, stop when a pass would go under nCellsInCoarsestLevel.
//   Pair      the host CPU path's GAMG (src/OpenFOAM/matrices/gamg.cu:41-172 agglomeratePair + coarsen):
//             each cell's upper-side faces then its lower-side ones, direction alternating every pass from
//             forward, singletons numbered last, a reverse pass renumbered n - 1 - id; coarse faces as sorted
//             unique (lo, hi) pairs; stop when a pass would reach nCellsInCoarsest cells or fewer.
#pragma once

#include <cstdint>
#include <vector>

namespace mfv {

struct AggLevel
{
    int nFine = 0, nFineFaces = 0, nCoarse = 0, nCoarseFaces = 0;
    std::vector<int> map;           // [nFine] fine cell -> coarse cell
    std::vector<int> cOwn, cNei;    // [nCoarseFaces] coarse addressing, owner < neighbour, upper-triangular
    std::vector<int> faceRestrict;  // [nFineFaces] >= 0 coarse face, < 0: -(coarse cell + 1)
    std::vector<int8_t> faceFlip;   // [nFineFaces]
};

enum class AggKind : int { Brae = 0, OpenFOAM = 1, Pair = 2 };

// fineOwner/fineNeighbour: the mesh's internal faces; weights per internal face.
// target: brae's TARGET (coarsen while n > target) or OpenFOAM's nCellsInCoarsestLevel.
// forward: OpenFOAM's static direction flag (in/out); ignored by Brae.
std::vector<AggLevel> agglomerate(AggKind kind, int nCells, const std::vector<int>& fineOwner,
                                  const std::vector<int>& fineNeighbour, const std::vector<double>& weights,
                                  int target, bool& forward);

// Composes each run of `merge` consecutive levels into one (OpenFOAM's mergeLevels); 1 leaves them.
std::vector<AggLevel> mergeLevels(std::vector<AggLevel> levels, int merge);

} // namespace mfv
