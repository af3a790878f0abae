// Symbolic analysis on the CPU: the part of spldlt_analyse that feeds the
// device numeric phase.
//
// Outputs use the reference's conventions (all 1-based, as produced by SPRAL
// basic_analyse and build_map and consumed by the SymbolicTree constructor,
// src/SymbolicTree.cxx:7-102):
//   sptr[nnodes+1]    first pivot of each supernode
//   sparent[nnodes]   parent supernode, nnodes+1 for a root
//   rptr[nnodes+1]    (int64) start of each node's row list in rlist
//   rlist[...]        row indices (pivot order), node's own pivots first, ascending
//   nptr[nnodes+1]    (int64) start of each node's A-map in nlist
//   nlist[2*nnz]      (int64) pairs (source index in val, 1-based position
//                     (col-sptr)*nrow + row) -- src/spldlt_analyse_mod.F90:130-232
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace sylver_amd {

struct SymbolicFront {
   int nrow = 0;           // rows of the front (rptr difference)
   int ncol = 0;           // pivots of the front (sptr difference)
   int parent = -1;        // 0-based parent, nnodes for the virtual root
   int first_child = -1;   // child linked list in prepend order
   int next_child = -1;
   int least_desc = 0;
   int64_t rlist_off = 0;  // 0-based offset of the node's rows in rlist
   int64_t amap_off = 0;   // 0-based offset (in pairs) of the node's A-map in nlist
   int64_t num_a = 0;
};

struct Analysis {
   int n = 0;
   // Cleaned input pattern (1-based CSC lower) used for build_map, and the map
   // from entries of the caller's val to it (only when cleaning changed it).
   std::vector<int64_t> ptr;
   std::vector<int> row;
   bool has_conv_map = false;
   std::vector<int64_t> conv_src;   // caller val index (0-based) per contribution
   std::vector<int64_t> conv_dst;   // clean index (0-based) it adds into
   int64_t user_nnz = 0;
   // Pivot order: order[v] = 1-based position of variable v+1; invp inverse.
   std::vector<int> order;
   std::vector<int> invp;
   // Supernodal structure
   int nnodes = 0;
   std::vector<int> sptr;
   std::vector<int> sparent;
   std::vector<int64_t> rptr;
   std::vector<int> rlist;
   std::vector<int64_t> nptr;
   std::vector<int64_t> nlist;
   // Symbolic tree (src/SymbolicTree.cxx:7-102); fronts[nnodes] is the root.
   std::vector<SymbolicFront> fronts;
   int64_t nfactor = 0;     // sum nrow*ncol
   // Per-front kernel variant of the posdef factorization chosen by the cost
   // model or the autotuner: -1 the default rule, 0 the tiled large-front
   // path, 1 the one-workgroup kernel; variant_gen changes with every update.
   std::vector<signed char> variant;
   int variant_gen = 0;
   // Caller-given front layout (batched fronts in the caller's memory):
   // offset and leading dimension per front; empty = the plan's own arena.
   std::vector<int64_t> fixed_off;
   std::vector<int> fixed_ld;
   // Scaling from the matching-based ordering (ordering = 2), caller's
   // variable order; used by spldlt_factorize with scaling = 3.
   std::vector<double> scaling;
   // Statistics reported through sylver_inform_t
   int64_t num_factor = 0;  // entries in L (lower)
   int64_t num_flops = 0;   // sum over nodes of sum_{r=m-n+1}^{m} r^2
   int maxfront = 0;
   int maxdepth = 0;
   int matrix_dup = 0;
   int matrix_outrange = 0;
   int matrix_missing_diag = 0;
   int nemin = 32;
};

// Returns the inform flag (0 on success, <0 error, >0 warning). `val` is
// only read for the matching-based ordering (ordering = 2).
int analyse(int n, int* order, const int64_t* ptr, const int* row, const double* val, bool check,
            int ordering, int nemin, Analysis& out);

// Values of the cleaned pattern (duplicates summed, out-of-range entries
// dropped) from the caller's val.
std::vector<double> clean_values(const Analysis& a, const double* val);

// General nested dissection ordering of the graph of a symmetric matrix given
// by its lower CSC pattern (1-based). order[v] = 1-based position.
void nested_dissection(int n, const int64_t* ptr, const int* row, int* order);

// build_map restated for this library (src/spldlt_analyse_mod.F90:130-232).
void build_map(int n, const int64_t* ptr, const int* row, const int* perm,
               const int* invp, int nnodes, const int* sptr, const int64_t* rptr,
               const int* rlist, int64_t* nptr, int64_t* nlist);

// Symbolic tree construction (src/SymbolicTree.cxx:7-102).
void build_symbolic_tree(Analysis& a);

// Reference flop model for one node (src/spldlt_analyse_mod.F90:1744-1764).
int64_t node_flops(int64_t m, int64_t n);

}  // namespace sylver_amd
