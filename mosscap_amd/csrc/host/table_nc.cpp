// Reader for the tabulated LTE hydrogen EOS table (eos.table_path), the file
// TabulatedLteH::init opens (TabulatedLteH.hpp:30-55): scalars
// log_eint_rho_start/step, n_eint_rho, log_rho_start/step, n_rho and the 2-D
// variables y_table and log_T_table [n_rho][n_eint_rho]. The file is read in
// the netCDF classic formats (CDF-1, CDF-2 with 64-bit offsets, CDF-5 with
// 64-bit sizes and types), parsed here without a netCDF library; numeric
// variables of type byte/short/int/float/double/int64 are converted to
// double. netCDF-4 (HDF5) files are refused with a hint to convert them.
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>

#include "simulation.hpp"

namespace mh::host {

namespace {

constexpr uint32_t NC_DIMENSION = 10, NC_VARIABLE = 11, NC_ATTRIBUTE = 12;

struct Reader {
    std::vector<unsigned char> d;
    size_t p = 0;
    int version = 1;
    std::string path;

    [[noreturn]] void bad(const std::string& what) const {
        throw std::runtime_error("eos.table_path " + path + ": " + what);
    }
    uint64_t be(int n) {
        if (p + n > d.size()) bad("truncated netCDF header");
        uint64_t v = 0;
        for (int i = 0; i < n; ++i) v = (v << 8) | d[p++];
        return v;
    }
    uint64_t count() { return be(version == 5 ? 8 : 4); }        // NON_NEG
    uint64_t offset() { return be(version == 1 ? 4 : 8); }       // OFFSET
    std::string name() {
        const uint64_t n = count();
        if (p + n > d.size()) bad("truncated netCDF header");
        std::string s(d.begin() + p, d.begin() + p + n);
        p += n;
        p = (p + 3) & ~(size_t)3;
        return s;
    }
    static int type_size(uint32_t t) {
        switch (t) {
            case 1: case 2: case 7: return 1;  // byte, char, ubyte
            case 3: case 8: return 2;          // short, ushort
            case 4: case 5: case 9: return 4;  // int, float, uint
            case 6: case 10: case 11: return 8;  // double, int64, uint64
            default: return 0;
        }
    }
    void skip_atts() {
        const uint32_t tag = (uint32_t)be(4);
        const uint64_t n = count();
        if (tag == 0 && n == 0) return;
        if (tag != NC_ATTRIBUTE) bad("malformed attribute list");
        for (uint64_t a = 0; a < n; ++a) {
            name();
            const uint32_t t = (uint32_t)be(4);
            const uint64_t m = count();
            const int sz = type_size(t);
            if (sz == 0) bad("unknown attribute type");
            p += (size_t)((m * sz + 3) & ~(uint64_t)3);
        }
    }
    double value(uint32_t t, size_t at) const {
        uint64_t u = 0;
        const int n = type_size(t);
        for (int i = 0; i < n; ++i) u = (u << 8) | d[at + i];
        switch (t) {
            case 1: return (double)(int8_t)u;
            case 3: return (double)(int16_t)u;
            case 4: return (double)(int32_t)u;
            case 5: { float f; uint32_t w = (uint32_t)u; std::memcpy(&f, &w, 4); return f; }
            case 6: { double v; std::memcpy(&v, &u, 8); return v; }
            case 7: return (double)(uint8_t)u;
            case 8: return (double)(uint16_t)u;
            case 9: return (double)(uint32_t)u;
            case 10: return (double)(int64_t)u;
            case 11: return (double)u;
            default: return NAN;
        }
    }
};

struct NcVar {
    std::vector<uint64_t> shape;
    uint32_t type = 0;
    uint64_t begin = 0;
};

}  // namespace

LteTable read_lte_table(const std::string& path) {
    Reader r;
    r.path = path;
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("eos.table_path " + path + ": cannot open the file");
    r.d.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    if (r.d.size() >= 4 && std::memcmp(r.d.data(), "\x89HDF", 4) == 0)
        r.bad("netCDF-4 (HDF5) files are not read here; convert the table with `nccopy -k cdf5`");
    if (r.d.size() < 8 || std::memcmp(r.d.data(), "CDF", 3) != 0) r.bad("not a netCDF classic file");
    r.version = r.d[3];
    if (r.version != 1 && r.version != 2 && r.version != 5) r.bad("unknown netCDF classic version");
    r.p = 4;
    r.count();  // numrecs (the table has no record variables)
    std::vector<uint64_t> dims;
    {
        const uint32_t tag = (uint32_t)r.be(4);
        const uint64_t n = r.count();
        if (!(tag == 0 && n == 0)) {
            if (tag != NC_DIMENSION) r.bad("malformed dimension list");
            for (uint64_t i = 0; i < n; ++i) {
                r.name();
                dims.push_back(r.count());
            }
        }
    }
    r.skip_atts();
    std::map<std::string, NcVar> vars;
    {
        const uint32_t tag = (uint32_t)r.be(4);
        const uint64_t n = r.count();
        if (!(tag == 0 && n == 0)) {
            if (tag != NC_VARIABLE) r.bad("malformed variable list");
            for (uint64_t i = 0; i < n; ++i) {
                const std::string nm = r.name();
                NcVar v;
                const uint64_t nd = r.count();
                for (uint64_t k = 0; k < nd; ++k) {
                    const uint64_t id = r.count();
                    if (id >= dims.size()) r.bad("variable " + nm + " names an unknown dimension");
                    v.shape.push_back(dims[id]);
                }
                r.skip_atts();
                v.type = (uint32_t)r.be(4);
                r.count();  // vsize
                v.begin = r.offset();
                if (Reader::type_size(v.type) == 0 || v.type == 2) r.bad("variable " + nm + " is not numeric");
                vars[nm] = v;
            }
        }
    }
    auto get = [&](const std::string& nm, size_t expect) {
        const auto it = vars.find(nm);
        if (it == vars.end()) r.bad("no variable " + nm);
        size_t n = 1;
        for (uint64_t s : it->second.shape) n *= (size_t)s;
        if (n != expect) r.bad("variable " + nm + " has " + std::to_string(n) + " values, expected " +
                               std::to_string(expect));
        const int sz = Reader::type_size(it->second.type);
        if (it->second.begin + (uint64_t)n * sz > r.d.size()) r.bad("variable " + nm + " runs past the end");
        std::vector<double> out(n);
        for (size_t i = 0; i < n; ++i) out[i] = r.value(it->second.type, (size_t)it->second.begin + i * sz);
        return out;
    };
    LteTable t;
    t.log_eint_rho_start = get("log_eint_rho_start", 1)[0];
    t.log_eint_rho_step = get("log_eint_rho_step", 1)[0];
    t.n_eint_rho = (int)get("n_eint_rho", 1)[0];
    t.log_rho_start = get("log_rho_start", 1)[0];
    t.log_rho_step = get("log_rho_step", 1)[0];
    t.n_rho = (int)get("n_rho", 1)[0];
    if (t.n_eint_rho < 1 || t.n_rho < 1) r.bad("empty table axes");
    const size_t m = (size_t)t.n_eint_rho * (size_t)t.n_rho;
    t.y_table = get("y_table", m);
    t.log_T_table = get("log_T_table", m);
    return t;
}

}  // namespace mh::host
