// Output in the netCDF classic format with 64-bit offsets (CDF-2), written
// directly (no netCDF library): the file layout of write_output
// (Output.cpp:88-278) — global attributes, x/y/z cell-centre axes, then Q
// (and W) as [var][z][y][x] per output time. With output.single_file the
// records go along an unlimited "time" dimension of one file; otherwise each
// output is its own file "{name}_{count:05d}.nc" holding one time value.
//
// Format (netCDF classic file format specification, Unidata): big-endian;
// header = magic "CDF\x02", numrecs, dimension list, global attribute list,
// variable list (name, dimids, attributes, type, vsize, begin offset); data
// of fixed-size variables follows at their offsets, then the records, each
// holding every record variable's slab in variable order.
#include <cstdio>
#include <cstring>
#include <fstream>

#include "simulation.hpp"

namespace mh::host {

namespace {

constexpr uint32_t NC_DIMENSION = 10, NC_VARIABLE = 11, NC_ATTRIBUTE = 12;
constexpr uint32_t NC_CHAR = 2, NC_INT = 4, NC_DOUBLE = 6;

struct Buf {
    std::vector<unsigned char> b;
    void u32(uint32_t v) {
        for (int s = 24; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s));
    }
    void u64(uint64_t v) {
        for (int s = 56; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s));
    }
    void pad() {
        while (b.size() % 4) b.push_back(0);
    }
    void name(const std::string& s) {
        u32((uint32_t)s.size());
        b.insert(b.end(), s.begin(), s.end());
        pad();
    }
};

struct Attr {
    std::string name;
    uint32_t type;
    std::string text;
    int32_t ival = 0;
};

struct Dim {
    std::string name;
    uint32_t len;  // 0 = unlimited
};

struct Var {
    std::string name;
    std::vector<int> dims;
    uint64_t elems_per_record;  // elements of one record (all elements for fixed variables)
    bool record;
    uint64_t begin = 0;
};

void put_be_doubles(std::vector<unsigned char>& out, const double* v, size_t n) {
    out.resize(n * 8);
    for (size_t k = 0; k < n; ++k) {
        uint64_t u;
        std::memcpy(&u, &v[k], 8);
        for (int s = 0; s < 8; ++s) out[k * 8 + s] = (unsigned char)(u >> (56 - 8 * s));
    }
}

// A file layout: header bytes plus each variable's offset.
struct Layout {
    std::vector<Dim> dims;
    std::vector<Attr> attrs;
    std::vector<Var> vars;
    uint64_t header_size = 0;
    uint64_t recsize = 0;

    std::vector<unsigned char> header(uint32_t numrecs) const {
        Buf h;
        h.b = {'C', 'D', 'F', 2};
        h.u32(numrecs);
        h.u32(NC_DIMENSION);
        h.u32((uint32_t)dims.size());
        for (const Dim& d : dims) {
            h.name(d.name);
            h.u32(d.len);
        }
        if (attrs.empty()) {
            h.u32(0);
            h.u32(0);
        } else {
            h.u32(NC_ATTRIBUTE);
            h.u32((uint32_t)attrs.size());
            for (const Attr& a : attrs) {
                h.name(a.name);
                h.u32(a.type);
                if (a.type == NC_CHAR) {
                    h.u32((uint32_t)a.text.size());
                    h.b.insert(h.b.end(), a.text.begin(), a.text.end());
                    h.pad();
                } else {
                    h.u32(1);
                    h.u32((uint32_t)a.ival);
                }
            }
        }
        h.u32(NC_VARIABLE);
        h.u32((uint32_t)vars.size());
        for (const Var& v : vars) {
            h.name(v.name);
            h.u32((uint32_t)v.dims.size());
            for (int d : v.dims) h.u32((uint32_t)d);
            h.u32(0);  // no variable attributes
            h.u32(0);
            h.u32(NC_DOUBLE);
            const uint64_t vsize = v.elems_per_record * 8;
            h.u32(vsize > 0xFFFFFFFCull ? 0xFFFFFFFFu : (uint32_t)vsize);
            h.u64(v.begin);
        }
        return h.b;
    }

    void finalize() {
        header_size = header(0).size();
        uint64_t off = header_size;
        for (Var& v : vars)
            if (!v.record) {
                v.begin = off;
                off += v.elems_per_record * 8;
            }
        recsize = 0;
        for (Var& v : vars)
            if (v.record) {
                v.begin = off + recsize;
                recsize += v.elems_per_record * 8;
            }
    }
};

std::string path_of(const Simulation& sim, const std::string& file) {
    const std::string& d = sim.out_cfg.directory;
    return (d.empty() || d == ".") ? file : d + "/" + file;
}

Layout make_layout(const Simulation& sim, bool single_file) {
    const auto& cfg = sim.out_cfg;
    const auto& sz = sim.state.sz;
    const int nd = sim.num_dim;
    Layout L;
    auto txt = [&](const std::string& n, const std::string& v) { L.attrs.push_back(Attr{n, NC_CHAR, v, 0}); };
    auto num = [&](const std::string& n, int v) {
        for (Attr& a : L.attrs)
            if (a.name == n) {
                a.ival = v;
                return;
            }
        L.attrs.push_back(Attr{n, NC_INT, "", v});
    };
    // write_header, Output.cpp:88-159
    txt("hydro_program", "mosscap");
#ifdef MOSSCAP_GIT_HASH
    txt("hydro_git_hash", MOSSCAP_GIT_HASH);
#else
    txt("hydro_git_hash", "unknown");
#endif
    txt("problem", cfg.problem_name);
    num("num_ghost", sz.ng);
    num("single_file", cfg.single_file ? 1 : 0);
    num("write_conserved", cfg.variables.conserved ? 1 : 0);
    num("write_primitive", cfg.variables.primitive ? 1 : 0);
    num("write_fluxes", cfg.variables.fluxes ? 1 : 0);
    num("write_source_term", cfg.variables.source ? 1 : 0);
    if (cfg.variables.conserved) {  // write_cons_header, Output.cpp:18-50
        num("irho", 0);
        num("imx", 1);
        if (nd > 1) num("imy", 2);
        if (nd > 2) num("imz", 3);
        num("iene", 1 + nd);
    }
    if (cfg.variables.primitive) {  // write_prim_header, Output.cpp:52-86
        num("irho", 0);
        num("ivx", 1);
        if (nd > 1) num("ivy", 2);
        if (nd > 2) num("ivz", 3);
        num("ipre", 1 + nd);
    }
    // write_axes, Output.cpp:161-199: dimensions x, y, z, time, then var
    L.dims = {{"x", (uint32_t)sz.xc}, {"y", (uint32_t)sz.yc}, {"z", (uint32_t)sz.zc}};
    L.vars.push_back(Var{"x", {0}, (uint64_t)sz.xc, false});
    if (nd > 1) L.vars.push_back(Var{"y", {1}, (uint64_t)sz.yc, false});
    if (nd > 2) L.vars.push_back(Var{"z", {2}, (uint64_t)sz.zc, false});
    L.dims.push_back({"time", single_file ? 0u : 1u});
    L.dims.push_back({"var", (uint32_t)sim.state.nvar});
    const uint64_t field = (uint64_t)sim.state.nvar * sz.zc * sz.yc * sz.xc;
    const uint64_t cells = (uint64_t)sz.zc * sz.yc * sz.xc;
    const bool eos_fields = !sim.eos.is_constant;  // ion_frac and T (Output.cpp:237-242, 261-266)
    if (single_file) {
        L.vars.push_back(Var{"time", {3}, 1, true});
        if (cfg.variables.conserved) L.vars.push_back(Var{"Q", {3, 4, 2, 1, 0}, field, true});
        if (cfg.variables.primitive) L.vars.push_back(Var{"W", {3, 4, 2, 1, 0}, field, true});
        if (eos_fields) {
            L.vars.push_back(Var{"ion_frac", {3, 2, 1, 0}, cells, true});
            L.vars.push_back(Var{"T", {3, 2, 1, 0}, cells, true});
        }
    } else {
        L.vars.push_back(Var{"time", {3}, 1, false});
        if (cfg.variables.conserved) L.vars.push_back(Var{"Q", {4, 2, 1, 0}, field, false});
        if (cfg.variables.primitive) L.vars.push_back(Var{"W", {4, 2, 1, 0}, field, false});
        if (eos_fields) {
            L.vars.push_back(Var{"ion_frac", {2, 1, 0}, cells, false});
            L.vars.push_back(Var{"T", {2, 1, 0}, cells, false});
        }
    }
    L.finalize();
    return L;
}

// EOS fields of the last update_eos (zeros before any, e.g. without a device)
std::vector<double> eos_y(const Simulation& sim) {
    return sim.eos.y_space.empty() ? std::vector<double>((size_t)sim.state.cells(), 0.0) : sim.eos.y_space;
}
std::vector<double> eos_T(const Simulation& sim) {
    return sim.eos.T_space.empty() ? std::vector<double>((size_t)sim.state.cells(), 0.0) : sim.eos.T_space;
}

void write_at(std::fstream& f, uint64_t off, const std::vector<unsigned char>& bytes) {
    f.seekp((std::streamoff)off);
    f.write(reinterpret_cast<const char*>(bytes.data()), (std::streamsize)bytes.size());
}

}  // namespace

// write_output, Output.cpp:201-278.
bool write_output(Simulation& sim) {
    auto& cfg = sim.out_cfg;
    if (sim.ctx) download_state(sim);
    const bool single = cfg.single_file;
    char count[16];
    std::snprintf(count, sizeof count, "%05d", cfg.output_count);
    const std::string file = path_of(sim, single ? cfg.filename + ".nc" : cfg.filename + "_" + count + ".nc");
    const Layout L = make_layout(sim, single);
    const bool create = cfg.prev_output_time < 0.0 || !single;

    std::vector<double> W;
    if (cfg.variables.primitive) W = primitive_state(sim);
    std::vector<unsigned char> bytes;

    std::fstream f;
    uint32_t rec = 0;
    if (create) {
        std::ofstream(file, std::ios::binary | std::ios::trunc).close();
        f.open(file, std::ios::binary | std::ios::in | std::ios::out);
        if (!f) throw std::runtime_error("cannot create output file " + file);
        write_at(f, 0, L.header(single ? 0 : 0));
        // axes (fixed variables)
        for (const Var& v : L.vars) {
            if (v.record) continue;
            std::vector<double> vals;
            if (v.name == "x")
                for (int i = 0; i < sim.state.sz.xc; ++i) vals.push_back(sim.state.pos_x(i));
            else if (v.name == "y")
                for (int j = 0; j < sim.state.sz.yc; ++j) vals.push_back(sim.state.pos_y(j));
            else if (v.name == "z")
                for (int k = 0; k < sim.state.sz.zc; ++k) vals.push_back(sim.state.pos_z(k));
            else if (v.name == "time")
                vals.push_back(sim.time);
            else if (v.name == "Q")
                vals = sim.state.Q;
            else if (v.name == "W")
                vals = W;
            else if (v.name == "ion_frac")
                vals = eos_y(sim);
            else if (v.name == "T")
                vals = eos_T(sim);
            put_be_doubles(bytes, vals.data(), vals.size());
            write_at(f, v.begin, bytes);
        }
    } else {
        f.open(file, std::ios::binary | std::ios::in | std::ios::out);
        if (!f) throw std::runtime_error("cannot open output file " + file);
        unsigned char n[4];
        f.seekg(4);
        f.read(reinterpret_cast<char*>(n), 4);
        rec = ((uint32_t)n[0] << 24) | ((uint32_t)n[1] << 16) | ((uint32_t)n[2] << 8) | n[3];
    }
    if (single) {
        for (const Var& v : L.vars) {
            if (!v.record) continue;
            const uint64_t off = v.begin + (uint64_t)rec * L.recsize;
            if (v.name == "time") {
                put_be_doubles(bytes, &sim.time, 1);
            } else {
                const std::vector<double> src = v.name == "Q"          ? sim.state.Q
                                                : v.name == "W"        ? W
                                                : v.name == "ion_frac" ? eos_y(sim)
                                                                       : eos_T(sim);
                put_be_doubles(bytes, src.data(), src.size());
            }
            write_at(f, off, bytes);
        }
        write_at(f, 0, L.header(rec + 1));
    }
    f.close();
    if (!f) throw std::runtime_error("writing " + file + " failed");
    cfg.prev_output_time = sim.time;
    cfg.output_count += 1;
    return true;
}

}  // namespace mh::host
