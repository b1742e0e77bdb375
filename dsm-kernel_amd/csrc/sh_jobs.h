// sh_jobs.h -- rows of the SH synthesis contraction: for every catalog
// component, the first coefficient row q0 (sh_lane.h), the number of theta
// functions nb, the azimuthal order |m| = mu and the theta functions b
// (0 Y0, 1 Y1, 2 Y2, 3 Y3, 4 Z23, 5 Z33; legendre_lane.h).
// kind 0: TSGT_SH (10 components = P-SV slots 7,12..20, udertotsgtSH),
// kind 1: RSGT_SH (5 = slots 6..10, udertorsgtSH), kind 2: SYNN_SH (5).
#pragma once

namespace dsm {

struct ShJobDef {
    int q0, nb, mu, b[4], kind, comp;
};

constexpr ShJobDef SH_JOBS[] = {
    {0, 2, 1, {2, 5, 0, 0}, 0, 0},  {2, 2, 2, {2, 5, 0, 0}, 0, 1},  {4, 2, 1, {0, 1, 0, 0}, 0, 2},
    {6, 2, 1, {0, 1, 0, 0}, 0, 3},  {8, 4, 1, {2, 5, 4, 3}, 0, 4},  {12, 4, 1, {2, 5, 4, 3}, 0, 5},
    {16, 2, 2, {0, 1, 0, 0}, 0, 6}, {18, 2, 2, {0, 1, 0, 0}, 0, 7}, {20, 4, 2, {2, 5, 4, 3}, 0, 8},
    {24, 4, 2, {2, 5, 4, 3}, 0, 9}, {28, 2, 1, {2, 5, 0, 0}, 1, 0}, {30, 2, 1, {2, 5, 0, 0}, 1, 1},
    {32, 2, 1, {4, 3, 0, 0}, 1, 2}, {34, 1, 1, {0, 0, 0, 0}, 1, 3}, {35, 1, 1, {1, 0, 0, 0}, 1, 4},
    {0, 1, 2, {0, 0, 0, 0}, 2, 0},  {1, 1, 2, {0, 0, 0, 0}, 2, 1},  {2, 1, 2, {1, 0, 0, 0}, 2, 2},
    {3, 1, 1, {0, 0, 0, 0}, 2, 3},  {4, 1, 1, {1, 0, 0, 0}, 2, 4},
};
constexpr int SH_NJOBS = sizeof(SH_JOBS) / sizeof(SH_JOBS[0]);
constexpr int SH_NQ = 36, SH_NQS = 5;

}  // namespace dsm
