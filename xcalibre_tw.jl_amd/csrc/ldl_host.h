// Host LDL^T factorization for the LDL preconditioner (ldl_host.cpp).
#pragma once

#include <vector>

namespace fvh {

// L D L^T = P A P^T with P the ordering perm (position k holds row perm[k]).
// L is unit lower triangular, stored by columns without its diagonal (Lp, Li,
// Lx; rows ascending within a column), in the permuted numbering.
struct LdlHost {
    int n = 0;
    std::vector<int> perm, Lp, Li;
    std::vector<double> Lx, D;
};

// Nested-dissection ordering of a symmetric pattern from the cells' centres;
// with `bounds`, also the ranges of positions that each leaf and each
// separator occupy (range g is [bounds[g], bounds[g + 1]), in order).
void ldl_order(int n, const int *rowptr, const int *col, const double *x, const double *y, std::vector<int> &perm,
               std::vector<int> *bounds = nullptr);

// Numeric factorization; A symmetric (values read from its upper triangle).
int ldl_factor(int n, const int *rowptr, const int *col, const double *val, const std::vector<int> &perm, LdlHost &F);

}  // namespace fvh
