// PGM JSON input / output (see pgm_io.hpp).
#include "pgm_io.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <map>
#include <numeric>
#include <sstream>
#include <stdexcept>

namespace gap::pgm {

// ------------------------------------------------------------------ JSON
Json const* Json::find(std::string_view key) const {
    if (kind != Object) return nullptr;
    for (auto const& kv : object)
        if (kv.first == key) return &kv.second;
    return nullptr;
}

namespace {

class Parser {
  public:
    explicit Parser(std::string const& t) : s_(t) {}

    Json document() {
        Json v = value(0);
        ws();
        if (i_ != s_.size()) fail("trailing characters");
        return v;
    }

  private:
    std::string const& s_;
    size_t i_ = 0;
    static constexpr int kMaxDepth = 256;

    [[noreturn]] void fail(char const* what) const {
        throw std::runtime_error(std::string("JSON: ") + what + " at byte " + std::to_string(i_));
    }
    void ws() {
        while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\t' || s_[i_] == '\n' || s_[i_] == '\r')) ++i_;
    }
    bool eat(char c) {
        ws();
        if (i_ < s_.size() && s_[i_] == c) {
            ++i_;
            return true;
        }
        return false;
    }
    void expect(char c) {
        if (!eat(c)) fail((std::string("expected '") + c + "'").c_str());
    }
    bool literal(char const* word) {
        size_t const n = std::char_traits<char>::length(word);
        if (s_.compare(i_, n, word) == 0) {
            i_ += n;
            return true;
        }
        return false;
    }
    static void put_utf8(std::string& out, unsigned cp) {
        if (cp < 0x80) {
            out += static_cast<char>(cp);
        } else if (cp < 0x800) {
            out += static_cast<char>(0xc0 | (cp >> 6));
            out += static_cast<char>(0x80 | (cp & 0x3f));
        } else if (cp < 0x10000) {
            out += static_cast<char>(0xe0 | (cp >> 12));
            out += static_cast<char>(0x80 | ((cp >> 6) & 0x3f));
            out += static_cast<char>(0x80 | (cp & 0x3f));
        } else {
            out += static_cast<char>(0xf0 | (cp >> 18));
            out += static_cast<char>(0x80 | ((cp >> 12) & 0x3f));
            out += static_cast<char>(0x80 | ((cp >> 6) & 0x3f));
            out += static_cast<char>(0x80 | (cp & 0x3f));
        }
    }
    unsigned hex4() {
        if (i_ + 4 > s_.size()) fail("short \\u escape");
        unsigned v = 0;
        for (int k = 0; k < 4; ++k) {
            char const c = s_[i_++];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= static_cast<unsigned>(c - '0');
            else if (c >= 'a' && c <= 'f') v |= static_cast<unsigned>(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= static_cast<unsigned>(c - 'A' + 10);
            else fail("bad \\u escape");
        }
        return v;
    }
    std::string str() {
        // at the opening quote
        ++i_;
        std::string out;
        while (true) {
            if (i_ >= s_.size()) fail("unterminated string");
            char const c = s_[i_++];
            if (c == '"') return out;
            if (static_cast<unsigned char>(c) < 0x20) fail("control character in string");
            if (c != '\\') {
                out += c;
                continue;
            }
            if (i_ >= s_.size()) fail("unterminated escape");
            char const e = s_[i_++];
            switch (e) {
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case '/': out += '/'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'n': out += '\n'; break;
                case 'r': out += '\r'; break;
                case 't': out += '\t'; break;
                case 'u': {
                    unsigned cp = hex4();
                    if (cp >= 0xd800 && cp < 0xdc00) {
                        if (!(literal("\\u"))) fail("lone high surrogate");
                        unsigned const lo = hex4();
                        if (lo < 0xdc00 || lo >= 0xe000) fail("bad low surrogate");
                        cp = 0x10000 + ((cp - 0xd800) << 10) + (lo - 0xdc00);
                    } else if (cp >= 0xdc00 && cp < 0xe000) {
                        fail("lone low surrogate");
                    }
                    put_utf8(out, cp);
                    break;
                }
                default: fail("bad escape");
            }
        }
    }
    double num() {
        size_t const b = i_;
        if (s_[i_] == '-') ++i_;
        auto digits = [&] {
            size_t const d = i_;
            while (i_ < s_.size() && s_[i_] >= '0' && s_[i_] <= '9') ++i_;
            return i_ > d;
        };
        if (i_ < s_.size() && s_[i_] == '0') {
            ++i_;
        } else if (!digits()) {
            fail("bad number");
        }
        if (i_ < s_.size() && s_[i_] == '.') {
            ++i_;
            if (!digits()) fail("bad fraction");
        }
        if (i_ < s_.size() && (s_[i_] == 'e' || s_[i_] == 'E')) {
            ++i_;
            if (i_ < s_.size() && (s_[i_] == '+' || s_[i_] == '-')) ++i_;
            if (!digits()) fail("bad exponent");
        }
        std::string const tok = s_.substr(b, i_ - b);
        return std::strtod(tok.c_str(), nullptr);  // correctly rounded (glibc)
    }
    Json value(int depth) {
        if (depth > kMaxDepth) fail("nesting too deep");
        ws();
        if (i_ >= s_.size()) fail("unexpected end");
        Json v;
        char const c = s_[i_];
        if (c == '{') {
            ++i_;
            v.kind = Json::Object;
            if (eat('}')) return v;
            do {
                ws();
                if (i_ >= s_.size() || s_[i_] != '"') fail("expected member name");
                std::string key = str();
                expect(':');
                v.object.emplace_back(std::move(key), value(depth + 1));
            } while (eat(','));
            expect('}');
        } else if (c == '[') {
            ++i_;
            v.kind = Json::Array;
            if (eat(']')) return v;
            do {
                v.array.push_back(value(depth + 1));
            } while (eat(','));
            expect(']');
        } else if (c == '"') {
            v.kind = Json::String;
            v.string = str();
        } else if (c == '-' || (c >= '0' && c <= '9')) {
            v.kind = Json::Number;
            v.number = num();
        } else if (literal("true")) {
            v.kind = Json::Bool;
            v.boolean = true;
        } else if (literal("false")) {
            v.kind = Json::Bool;
        } else if (literal("null")) {
            v.kind = Json::Null;
        } else {
            fail("unexpected character");
        }
        return v;
    }
};

std::string read_file(std::string const& filename) {
    std::ifstream f(filename, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + filename);
    std::ostringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

Json const& member(Json const& obj, char const* key, char const* what) {
    Json const* v = obj.find(key);
    if (!v) throw std::runtime_error(std::string(what) + " without '" + key + "'");
    return *v;
}
double number(Json const& obj, char const* key, char const* what) {
    Json const& v = member(obj, key, what);
    if (v.kind == Json::Bool) return v.boolean ? 1.0 : 0.0;
    if (v.kind != Json::Number) throw std::runtime_error(std::string(what) + " '" + key + "' is not a number");
    return v.number;
}
double number_or(Json const& obj, char const* key, double dflt, char const* what) {
    return obj.find(key) ? number(obj, key, what) : dflt;
}
std::vector<Json> const& component(Json const& data, char const* name) {
    static std::vector<Json> const empty;
    Json const* c = data.find(name);
    if (!c) return empty;
    if (c->kind != Json::Array) throw std::runtime_error(std::string("'") + name + "' is not an array");
    return c->array;
}

}  // namespace

Json parse_json(std::string const& text) { return Parser(text).document(); }

std::string fmt(double x) {
    if (!std::isfinite(x)) return "null";
    char buf[40];
    std::snprintf(buf, sizeof buf, "%.17g", x);
    return buf;
}

// ------------------------------------------------------------------ rows
Rows rows_from_pgm(Json const& doc, double base_power) {
    Json const& data = doc.find("data") ? *doc.find("data") : doc;
    auto const& nodes = component(data, "node");
    if (nodes.empty()) throw std::runtime_error("No nodes found in input data");
    auto const& lines = component(data, "line");
    if (lines.empty()) throw std::runtime_error("No lines found in input data");
    auto const& sources = component(data, "source");
    if (sources.empty()) throw std::runtime_error("No sources found in input data");
    auto const& loads = component(data, "sym_load");

    size_t const n = nodes.size();
    std::vector<int> ids(n);
    std::vector<double> ur(n);
    for (size_t k = 0; k < n; ++k) {
        ids[k] = static_cast<int>(number(nodes[k], "id", "node"));
        ur[k] = number(nodes[k], "u_rated", "node");
    }
    std::vector<size_t> order(n);
    std::iota(order.begin(), order.end(), size_t{0});
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return ids[a] < ids[b]; });
    std::vector<int> node_id(n);
    std::vector<double> u_rated(n);
    std::map<int, int> idx;
    for (size_t k = 0; k < n; ++k) {
        node_id[k] = ids[order[k]];
        u_rated[k] = ur[order[k]];
        idx[node_id[k]] = static_cast<int>(k);
    }
    auto index_of = [&](Json const& obj, char const* key, char const* what) {
        int const id = static_cast<int>(number(obj, key, what));
        auto it = idx.find(id);
        if (it == idx.end()) throw std::runtime_error(std::string(what) + " refers to unknown node " + std::to_string(id));
        return it->second;
    };

    double const base_voltage = *std::max_element(u_rated.begin(), u_rated.end());
    double const z_max = base_voltage * base_voltage / base_power;
    double const omega = 2.0 * M_PI * 50.0;
    double const z0 = (u_rated[0] * u_rated[0]) / base_power;

    // sources: slack at u_ref (the first source on a node)
    std::vector<int> src_type(n, 0);
    std::vector<double> src_u(n, 1.0);
    for (auto const& s : sources) {
        int const i = index_of(s, "node", "source");
        double const u_ref = number(s, "u_ref", "source");
        if (src_type[i] != 2) {
            src_type[i] = 2;
            src_u[i] = u_ref;
        }
    }
    // loads: -(p / S_base) * S_base summed in load order over in-service loads
    std::vector<double> p_load(n, 0.0), q_load(n, 0.0);
    for (auto const& l : loads) {
        int const i = index_of(l, "node", "sym_load");
        int const status = static_cast<int>(number_or(l, "status", 1.0, "sym_load"));
        double const p_pu = number(l, "p_specified", "sym_load") / base_power;
        double const q_pu = number(l, "q_specified", "sym_load") / base_power;
        if (status) {
            p_load[i] -= p_pu * base_power;
            q_load[i] -= q_pu * base_power;
        }
    }
    Rows r;
    for (size_t i = 0; i < n; ++i)
        r.bus.push_back({static_cast<double>(node_id[i]), u_rated[i], static_cast<double>(src_type[i]), src_u[i], 0.0,
                         p_load[i], q_load[i]});
    for (auto const& l : lines) {
        int const id = static_cast<int>(number(l, "id", "line"));
        int const f = index_of(l, "from_node", "line"), t = index_of(l, "to_node", "line");
        int const status = static_cast<int>(number_or(l, "from_status", 1.0, "line")) *
                           static_cast<int>(number_or(l, "to_status", 1.0, "line"));
        double const r_pu = number(l, "r1", "line") / z_max;
        double const x_pu = number(l, "x1", "line") / z_max;
        double const b_pu = omega * number_or(l, "c1", 0.0, "line") * z_max;
        if (status) {
            r.branch.push_back({static_cast<double>(id), static_cast<double>(f), static_cast<double>(t), r_pu * z0,
                                x_pu * z0, b_pu / z0});
            r.branch_i_n.push_back(number_or(l, "i_n", 0.0, "line"));
        }
    }
    return r;
}

NetworkData network_from_rows(Rows const& rows) {
    NetworkData net;
    for (auto const& row : rows.bus) {
        if (row.size() < 7)
            throw std::runtime_error("Bus data must have at least 7 elements: [id, u_rated, bus_type, u_pu, u_angle, p, q]");
        BusData b;
        b.id = static_cast<int>(row[0]);
        b.u_rated = row[1];
        b.bus_type = static_cast<BusType>(static_cast<int>(row[2]));
        b.energized = true;
        b.u_pu = row[3];
        b.u_angle = row[4];
        b.active_power = row[5];
        b.reactive_power = row[6];
        net.buses.push_back(b);
    }
    for (auto const& row : rows.branch) {
        if (row.size() < 5) throw std::runtime_error("Branch data must have at least 5 elements: [id, from_bus, to_bus, r, x]");
        BranchData br;
        br.id = static_cast<int>(row[0]);
        br.from_bus = static_cast<int>(row[1]);
        br.to_bus = static_cast<int>(row[2]);
        br.status = 1;
        br.r1 = row[3];
        br.x1 = row[4];
        br.b1 = row.size() > 5 ? row[5] : 0.0;
        if (net.branches.size() < rows.branch_i_n.size()) br.i_n = rows.branch_i_n[net.branches.size()];
        net.branches.push_back(br);
    }
    net.num_buses = static_cast<int>(net.buses.size());
    net.num_branches = static_cast<int>(net.branches.size());
    net.num_appliances = 0;
    return net;
}

// ------------------------------------------------------------------ module
NetworkData PgmJsonIO::read_network_data(std::string const& filename) {
    rows_ = rows_from_pgm(parse_json(read_file(filename)), base_power_);
    last_ = network_from_rows(rows_);
    // as the reference reader (json_io.cpp:496-499): a warning, not an error
    if (!last_.validate_pgm_compliance())
        std::cerr << "warning: " << filename << ": network data does not meet PGM compliance standards\n";
    return last_;
}

bool PgmJsonIO::validate_input_format(std::string const& filename) {
    try {
        (void)rows_from_pgm(parse_json(read_file(filename)), base_power_);
        return true;
    } catch (std::exception const&) {
        return false;
    }
}

namespace {

void write_text(std::string const& filename, std::string const& text) {
    std::ofstream f(filename, std::ios::binary);
    if (!f) throw std::runtime_error("cannot write " + filename);
    f << text;
    if (!f) throw std::runtime_error("write failed: " + filename);
}

std::string node_entry(BusData const& b, Complex v, double p, double q, bool with_power) {
    std::string e = "{\"id\":" + std::to_string(b.id) + ",\"energized\":1,\"u_pu\":" + fmt(std::abs(v)) +
                    ",\"u_angle\":" + fmt(std::arg(v)) + ",\"u\":" + fmt(std::abs(v) * b.u_rated);
    if (with_power) e += ",\"p\":" + fmt(p) + ",\"q\":" + fmt(q);
    return e + "}";
}

std::string meta(bool converged, int iterations, double final_mismatch, double base) {
    return "\"meta\":{\"solver\":\"gap_amd (MI355X HIP)\",\"calculation_method\":\"newton_raphson\",\"converged\":" +
           std::string(converged ? "true" : "false") + ",\"iterations\":" + std::to_string(iterations) +
           ",\"final_mismatch\":" + fmt(final_mismatch) + ",\"base_power_va\":" + fmt(base) + "}";
}

}  // namespace

void PgmJsonIO::write_results(std::string const& filename, ComplexVector const& bus_voltages, bool converged,
                              int iterations) {
    if (bus_voltages.size() != last_.buses.size())
        throw std::runtime_error("write_results: voltage count does not match the network read last");
    std::string out = "{\"version\":\"1.0\",\"type\":\"sym_output\",\"is_batch\":false,\"attributes\":{},\"data\":{\"node\":[";
    for (size_t i = 0; i < bus_voltages.size(); ++i)
        out += (i ? "," : "") + node_entry(last_.buses[i], bus_voltages[i], 0.0, 0.0, false);
    out += "]}," + meta(converged, iterations, 0.0, base_power_) + "}\n";
    write_text(filename, out);
}

void PgmJsonIO::write_output(std::string const& filename, NetworkData const& net, ComplexVector const& v,
                             std::vector<BranchData> const& branches, bool converged, int iterations,
                             double final_mismatch) const {
    size_t const n = net.buses.size();
    if (v.size() != n) throw std::runtime_error("write_output: voltage count does not match the network");
    // node injections: what leaves the node into its branches (and shunts)
    std::map<int, size_t> idx;
    for (size_t i = 0; i < n; ++i) idx[net.buses[i].id] = i;
    std::vector<double> p(n, 0.0), q(n, 0.0);
    for (auto const& b : branches) {
        auto f = idx.find(b.from_bus), t = idx.find(b.to_bus);
        if (f != idx.end()) {
            p[f->second] += b.p_from;
            q[f->second] += b.q_from;
        }
        if (t != idx.end()) {
            p[t->second] += b.p_to;
            q[t->second] += b.q_to;
        }
    }
    double const zb = n ? net.buses[0].u_rated * net.buses[0].u_rated / base_power_ : 1.0;
    for (auto const& a : net.appliances) {
        if (a.type != ApplianceType::SHUNT) continue;
        auto it = idx.find(a.node);
        if (it == idx.end()) continue;
        double const m2 = std::norm(v[it->second]);
        p[it->second] += m2 * a.g1 * zb * base_power_;
        q[it->second] -= m2 * a.b1 * zb * base_power_;
    }
    std::string out = "{\"version\":\"1.0\",\"type\":\"sym_output\",\"is_batch\":false,\"attributes\":{},\"data\":{\"node\":[";
    for (size_t i = 0; i < n; ++i) out += (i ? "," : "") + node_entry(net.buses[i], v[i], p[i], q[i], true);
    out += "],\"line\":[";
    for (size_t k = 0; k < branches.size(); ++k) {
        BranchData const& b = branches[k];
        out += (k ? "," : "") + std::string("{\"id\":") + std::to_string(b.id) + ",\"energized\":" +
               std::to_string(b.status ? 1 : 0) + ",\"p_from\":" + fmt(b.p_from) + ",\"q_from\":" + fmt(b.q_from) +
               ",\"i_from\":" + fmt(b.i_from) + ",\"s_from\":" + fmt(b.s_from) + ",\"p_to\":" + fmt(b.p_to) +
               ",\"q_to\":" + fmt(b.q_to) + ",\"i_to\":" + fmt(b.i_to) + ",\"s_to\":" + fmt(b.s_to) + ",\"loading\":" + fmt(b.loading) + "}";
    }
    out += "]}," + meta(converged, iterations, final_mismatch, base_power_) + "}\n";
    write_text(filename, out);
}

}  // namespace gap::pgm
