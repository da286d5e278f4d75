// Power Grid Model (PGM) JSON input / output for the MI355X backend: the
// replacement of the reference's JSON module (src/io/json_io.cpp), which does
// not build with libstdc++ 11, reads loads with the wrong sign, drops shunts
// and has a stub writer (json_io.cpp:195-211, 426-486, 507-517).
//
// Reading follows the conversion of the reference's validated path (the PGM
// harness feeding solve_simple_power_flow, tests/pgm_validation/
// test_pgm_validation.py:160-221 and json_io/gap_json_parser.py:96-350):
// nodes sorted by id; line r1, x1 to per unit on Z_max = max(u_rated)^2 / S_base
// and b = 2 pi 50 c1 Z_max, back to ohm / siemens on Z_0 = u_rated[0]^2 / S_base;
// from / to as node indices; a node with a source is a slack bus at u_ref, every
// other node PQ at 1.0 pu with its in-service loads summed with a negative sign.
// The float operations are the harness's, in its order, so the rows are
// bit-identical to the ones it builds.  Only node / line / sym_load / source
// are read, as on that path.
//
// Writing produces a PGM-style output dataset: per node u_pu, u_angle (rad),
// u (V) and the injection p, q (W, VAr) from the branch flows; per line the
// flows p, q, i, s at both ends.
#pragma once

#include <string>
#include <string_view>
#include <utility>
#include <vector>

#include "gap/io/io_interface.h"

namespace gap::pgm {

// A parsed JSON value (RFC 8259; numbers as double).
struct Json {
    enum Kind { Null, Bool, Number, String, Array, Object };
    Kind kind = Null;
    bool boolean = false;
    double number = 0.0;
    std::string string;
    std::vector<Json> array;
    std::vector<std::pair<std::string, Json>> object;

    Json const* find(std::string_view key) const;  // object member or nullptr
};

// Parses a whole document; throws std::runtime_error naming the byte offset.
Json parse_json(std::string const& text);

// Bus rows [id, u_rated, type, u_pu, u_angle, p, q] and branch rows
// [id, from, to, r, x, b] as the harness passes them to solve_simple_power_flow.
struct Rows {
    std::vector<std::vector<double>> bus, branch;
    std::vector<double> branch_i_n;  // rated current of each branch row's line (PGM i_n, A; 0 if absent)
};
Rows rows_from_pgm(Json const& doc, double base_power);

// NetworkData from rows, as solve_simple_power_flow converts them
// (bindings/solver_bindings.cpp:64-118); a line's rated current, when the rows
// carry it, goes to BranchData::i_n (it does not enter the solve).
NetworkData network_from_rows(Rows const& rows);

// Number formatting that round-trips a double.
std::string fmt(double x);

class PgmJsonIO final : public io::IIOModule {
  public:
    explicit PgmJsonIO(double base_power = 1e6) : base_power_(base_power) {}

    NetworkData read_network_data(std::string const& filename) override;
    // Nodes only (ids and ratings from the last network read); see write_output.
    void write_results(std::string const& filename, ComplexVector const& bus_voltages, bool converged,
                       int iterations) override;
    bool validate_input_format(std::string const& filename) override;

    // Full output dataset: nodes with injections and lines with the flows of
    // `branches` (BranchData result fields filled by the solver).
    void write_output(std::string const& filename, NetworkData const& net, ComplexVector const& bus_voltages,
                      std::vector<BranchData> const& branches, bool converged, int iterations,
                      double final_mismatch) const;

    Rows const& last_rows() const { return rows_; }
    double base_power() const { return base_power_; }

  private:
    double base_power_;
    Rows rows_;
    NetworkData last_;
};

}  // namespace gap::pgm
