// Input/output module interface: the same class, methods and semantics as the
// reference's include/gap/io/io_interface.h:13-40, so callers written against
// it (src/main/main.cpp) compile unchanged against this tree.
#pragma once

#include <memory>
#include <string>

#include "gap/core/types.h"

namespace gap::io {

class IIOModule {
  public:
    virtual ~IIOModule() = default;

    // Network data of an input file.
    virtual NetworkData read_network_data(std::string const& filename) = 0;

    // Bus voltage phasors (per unit) and the convergence outcome, to a file.
    virtual void write_results(std::string const& filename, ComplexVector const& bus_voltages, bool converged,
                               int iterations) = 0;

    // true when the file parses as an input dataset this module reads.
    virtual bool validate_input_format(std::string const& filename) = 0;
};

}  // namespace gap::io
