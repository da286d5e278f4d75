// Data model of the power-flow plugin boundary.  Names, field meanings and
// defaults match the reference model (include/gap/core/types.h:15-357 of the
// reference) so that a caller written against it compiles unchanged; the GPU
// mirror structs of the reference (GPUSparseMatrix, GPUPowerFlowData) are not
// part of this boundary because device data lives behind the C ABI (gap_hip.h).
#pragma once

#include <cmath>
#include <complex>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace gap {

using Float = double;
using Complex = std::complex<Float>;
using ComplexVector = std::vector<Complex>;

enum class BusType { PQ = 0, PV = 1, SLACK = 2 };
enum class ApplianceType { SOURCE = 0, LOADGEN = 1, SHUNT = 2 };
enum class BranchType { LINE = 0, TRAFO = 1, GENERIC = 2 };

// CSR with int32 indices; duplicates allowed (parallel branches).
struct SparseMatrix {
    std::vector<int> row_ptr;
    std::vector<int> col_idx;
    std::vector<Complex> values;
    int num_rows = 0;
    int num_cols = 0;
    int nnz = 0;
};

struct BusData {
    int id = 0;
    Float u_rated = 0.0;  // V
    BusType bus_type{BusType::PQ};
    int8_t energized{0};
    Float u{0.0};
    Float u_pu{0.0};
    Float u_angle{0.0};
    Float active_power{0.0};    // W, generator sign
    Float reactive_power{0.0};  // VAr, generator sign
};

struct BranchData {
    int id = 0;
    int from_bus = 0;
    int8_t from_status = 1;
    int to_bus = 0;
    int8_t to_status = 1;
    int8_t status = 1;
    BranchType branch_type{BranchType::LINE};
    Float p_from{0.0}, p_to{0.0}, q_from{0.0}, q_to{0.0}, s_from{0.0}, s_to{0.0};
    Float i_from{0.0}, i_from_angle{0.0}, i_to{0.0}, i_to_angle{0.0}, loading{0.0};
    Float r1{1e-6};  // ohm
    Float x1{1e-6};  // ohm
    Float g1{0.0};   // S
    Float b1{0.0};   // S
    Float k{1.0};
    Float theta{0.0};
    Float sn{0.0};
    Float i_n{1000.0};
    void set_from_pgm_capacitive_params(Float c1, Float tan1, Float frequency = 50.0) {
        if (c1 > 0) {
            b1 = 2.0 * M_PI * frequency * c1;
            g1 = b1 * tan1;
        }
    }
};

struct ApplianceData {
    int id = 0;
    int node = 0;
    int8_t status = 1;
    ApplianceType type{ApplianceType::LOADGEN};
    Float p{0.0}, q{0.0}, i{0.0}, s{0.0};
    Float u_ref{1.0}, u_ref_angle{0.0}, sk{1e10}, rx_ratio{0.1};
    Float p_specified{0.0};  // W
    Float q_specified{0.0};  // VAr
    Float g1{0.0};           // S (shunt)
    Float b1{0.0};           // S (shunt)
};

struct NetworkData {
    std::vector<BusData> buses;
    std::vector<BranchData> branches;
    std::vector<ApplianceData> appliances;
    int num_buses = 0;
    int num_branches = 0;
    int num_appliances = 0;
    std::string version{"1.0"};
    std::string type{"input"};
    bool is_batch{false};

    // The reference's consistency check (types.h:337-351 of the reference,
    // called by its JSON reader and bound in Python): every bus has a positive
    // rated voltage and every branch r1 >= 0, x1 != 0 and a positive rated
    // current.  Appliances are not checked.
    bool validate_pgm_compliance() const {
        for (BusData const& b : buses)
            if (b.u_rated <= 0.0) return false;
        for (BranchData const& br : branches)
            if (br.r1 < 0.0 || br.x1 == 0.0 || br.i_n <= 0.0) return false;
        return true;
    }
};

// GPU_HIP fills the slot the reference left commented out (types.h:357).
enum class BackendType { CPU, GPU_CUDA, GPU_HIP };

inline std::string backend_type_to_string(BackendType t) {
    switch (t) {
        case BackendType::CPU: return "CPU";
        case BackendType::GPU_CUDA: return "GPU_CUDA";
        case BackendType::GPU_HIP: return "GPU_HIP";
    }
    return "Unknown";
}

inline std::string appliance_type_to_string(ApplianceType t) {
    switch (t) {
        case ApplianceType::SOURCE: return "SOURCE";
        case ApplianceType::LOADGEN: return "LOADGEN";
        case ApplianceType::SHUNT: return "SHUNT";
    }
    return "Unknown";
}

inline std::string bus_type_to_string(BusType t) {
    switch (t) {
        case BusType::PQ: return "PQ";
        case BusType::PV: return "PV";
        case BusType::SLACK: return "SLACK";
    }
    return "Unknown";
}

}  // namespace gap
