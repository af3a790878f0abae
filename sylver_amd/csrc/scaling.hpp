// Matching-based scalings and ordering of a symmetric matrix (host code).
//
// The reference computes these in SPRAL's spral_scaling and spral_match_order
// modules (called from src/spldlt_factorize_mod.F90:734-835 and
// src/spldlt_analyse_mod.F90:789-820); SPRAL is not part of the reference's
// tree, so the methods here are stated from their published descriptions:
//  - hungarian_scale_sym: the maximum-product matching of Duff & Koster
//    (MC64, SIAM J. Matrix Anal. Appl. 22, 2001) found by shortest augmenting
//    paths (Dijkstra with dual potentials), made symmetric as in Duff &
//    Pralet (SIAM J. Matrix Anal. Appl. 27, 2005): s_i = exp((u_i + v_i -
//    log cmax_i) / 2), so that every |s_i a_ij s_j| <= 1;
//  - auction_scale_sym: the same assignment problem by Bertsekas' auction
//    with epsilon scaling (Ann. Oper. Res. 14, 1988); its prices give duals
//    that satisfy the dual constraints to within epsilon;
//  - match_order: the matching split into 1x1 and 2x2 candidate pivots
//    (cycles of the matching permutation cut into pairs), the pairs
//    compressed to single vertices, the compressed graph ordered by nested
//    dissection and expanded with each pair kept adjacent (Hogg & Scott,
//    RAL-TR-2008-020 / Duff & Pralet).
// The norm equilibration (MC77, Ruiz 2001) runs on the device with the
// factorization (numeric.hip).
#pragma once

#include <cstdint>
#include <vector>

namespace sylver_amd {

// Inputs: lower-triangle CSC, 1-based (ptr[n+1], row), values val.
// Outputs in the caller's variable order (0-based arrays of n):
//  scaling[i] > 0; match[i] = variable matched to i, or -1 (unmatched).
// Return: 0; 1 structurally singular (scaling still computed when
// scale_if_singular, unmatched variables scaled so their largest scaled
// entry is 1); -2 structurally singular and not scale_if_singular.
int hungarian_scale_sym(int n, const int64_t* ptr, const int* row, const double* val, bool scale_if_singular,
                        double* scaling, int* match);

// Auction variant (epsilon-optimal): same outputs; entries of the scaled
// matrix are bounded by exp(eps_final) instead of 1. Return 0 or 1 (some
// variables left unmatched within the iteration limit).
int auction_scale_sym(int n, const int64_t* ptr, const int* row, const double* val, double* scaling,
                      int* match);

// Matching-based ordering: order[v] = 1-based pivot position of variable v;
// scaling[v] the Hungarian scaling (kept by the analysis for scaling = 3).
// Return 0, or 1 when the matrix is structurally singular (warning).
int match_order(int n, const int64_t* ptr, const int* row, const double* val, int* order, double* scaling);

}  // namespace sylver_amd
