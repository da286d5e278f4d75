// Y-bus builder plugin interface; same virtual signatures as the reference
// (include/gap/admittance/admittance_interface.h:12-38 of the reference).
#pragma once

#include <memory>
#include <vector>

#include "gap/core/types.h"

namespace gap::admittance {

class IAdmittanceMatrix {
  public:
    virtual ~IAdmittanceMatrix() = default;
    virtual std::unique_ptr<SparseMatrix> build_admittance_matrix(NetworkData const& network_data) = 0;
    virtual std::unique_ptr<SparseMatrix> update_admittance_matrix(
        SparseMatrix const& matrix, std::vector<BranchData> const& branch_changes) = 0;
    virtual BackendType get_backend_type() const noexcept = 0;
};

}  // namespace gap::admittance
