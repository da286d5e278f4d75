// Sparse LU plugin interface; same virtual signatures as the reference
// (include/gap/solver/lu_solver_interface.h:12-48 of the reference).
#pragma once

#include "gap/core/types.h"

namespace gap::solver {

class ILUSolver {
  public:
    virtual ~ILUSolver() = default;
    virtual bool factorize(SparseMatrix const& matrix) = 0;
    virtual ComplexVector solve(ComplexVector const& rhs) = 0;
    virtual bool update_factorization(SparseMatrix const& matrix) = 0;
    virtual BackendType get_backend_type() const noexcept = 0;
    virtual bool is_factorized() const noexcept = 0;
};

}  // namespace gap::solver
