// Symbolic analysis of the mole-reactor preconditioner: the sparsity pattern
// of J = d(ydot)/dy that IdealGasConstPressureMoleReactor::getJacobianElements
// fills (src/zeroD/IdealGasConstPressureMoleReactor.cpp:108-174), a
// minimum-degree elimination order, the fill of an LU factorisation without
// pivoting in that order, and the device schedules of the numeric
// factorisation and the two triangular solves.
//
// The reference factors M = I - gamma J with Eigen's IncompleteLUT (drop
// tolerance 1e-12 of the row norm, fill factor N/2 per row;
// src/numerics/AdaptivePreconditioner.cpp:24-81); here M is factored
// completely on the fixed fill pattern, no entries dropped, so the
// preconditioner is exact for the given J.
#pragma once
#include <vector>

namespace cg {

struct SparseLUHost {
    int N = 0;                 // state size, internal order [n_0..n_{K-1}, T]
    int nnz = 0;               // entries of L + U (fill included)
    std::vector<int> perm;     // elimination position -> internal index
    std::vector<int> iperm;    // internal index -> elimination position
    std::vector<int> pos;      // N x N internal (row, col) -> value slot, -1 outside the fill pattern
    std::vector<int> row, col; // value slot -> internal (row, col)
    std::vector<int> diag;     // elimination position k -> slot of (k, k)
    // factorisation step k: L entries of column k (slots), then the updates
    // a_ij -= l_ik u_kj as (slot ij, slot ik, slot kj) triples
    std::vector<int> lptr, lslot, lrow;   // lrow: elimination position of each L entry's row
    std::vector<int> optr, ops;           // ops: 3 ints per update
    // backward solve: U column k above the diagonal (slots, row positions)
    std::vector<int> uptr, uslot, urow;
    long updates() const { return static_cast<long>(ops.size() / 3); }
};

// pattern: N x N booleans (row-major, internal order); the diagonal is added.
SparseLUHost sparse_lu_analyse(int N, const std::vector<char>& pattern);

}  // namespace cg
