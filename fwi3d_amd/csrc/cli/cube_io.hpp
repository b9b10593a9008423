// Hypercube files of the drop-in executables: a text header of "key=value" words
// (n1..n6, o1..o6, d1..d6, label1..label6, esize, in=<binary path>, data_format) and a
// binary file of the samples, first axis fastest. The binary format (format=1) is the
// reference's binRead / binWrite (src/include/IO.hpp:181-430): native byte order,
// esize 4 or 8 on input, the binary written as <datapath>/<output>@ next to the header.
// The SEP format (format=0) uses the same header text as SEPlib's .H files, written
// here without the SEPlib library: data_format="native_float" on output, and on input
// "xdr_float" (SEPlib's default, big-endian) or native samples.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "params.hpp"

namespace fwi3d_cli {

struct Axis {
    int n = 1;
    float o = 0.f, d = 1.f;
    std::string label;
};

struct Cube {
    std::vector<Axis> axes;
    std::vector<float> v;
    size_t count() const {
        size_t c = 1;
        for (const Axis& a : axes) c *= (size_t)a.n;
        return c;
    }
    void alloc() { v.assign(count(), 0.f); }
};

inline std::string strip_quotes(std::string s) {
    if (s.size() >= 2 && s.front() == '"' && s.back() == '"') s = s.substr(1, s.size() - 2);
    return s;
}

// Read a header and its samples. ndim0 keeps at least that many axes (trailing axes of
// length 1 beyond it are dropped, as the reference does).
inline Cube read_cube(const std::string& header, int ndim0 = 1) {
    std::ifstream f(header);
    if (!f.is_open()) throw std::runtime_error("The input file " + header + " could not be opened. Check the path or the file name");
    Axis ax[6];
    int esize = 4;
    std::string in, fmt;
    std::string line;
    while (std::getline(f, line)) {
        if (!line.empty() && line[0] == '#') continue;
        std::istringstream ss(line);
        std::string w;
        while (ss >> w) {
            const size_t eq = w.find('=');
            if (eq == std::string::npos) continue;
            const std::string k = w.substr(0, eq), val = w.substr(eq + 1);
            if (k.size() == 2 && (k[0] == 'n' || k[0] == 'o' || k[0] == 'd') && k[1] >= '1' && k[1] <= '6') {
                Axis& a = ax[k[1] - '1'];
                if (k[0] == 'n') a.n = from_token<int>(val);
                else if (k[0] == 'o') a.o = from_token<float>(val);
                else a.d = from_token<float>(val);
            } else if (k.size() == 6 && k.compare(0, 5, "label") == 0 && k[5] >= '1' && k[5] <= '6') {
                ax[k[5] - '1'].label = val;
            } else if (k == "esize") {
                esize = from_token<int>(val);
            } else if (k == "in") {
                in = strip_quotes(val);
            } else if (k == "data_format") {
                fmt = strip_quotes(val);
            }
        }
    }
    int ndim = 6;
    while (ndim > 1 && ax[ndim - 1].n == 1 && ndim != ndim0) ndim--;
    Cube c;
    c.axes.assign(ax, ax + ndim);
    c.alloc();
    const size_t n = c.count();
    FILE* b = std::fopen(in.c_str(), "rb");
    if (!b) throw std::runtime_error("The input file " + in + " could not be read. Check the path or the file name");
    std::fseek(b, 0, SEEK_END);
    const long bytes = std::ftell(b);
    std::rewind(b);
    if (bytes != (long)(n * esize)) {
        std::fclose(b);
        throw std::runtime_error("The binary file contains " + std::to_string(bytes) +
                                 " bytes whereas the description file suggests " + std::to_string(n * esize) + " bytes");
    }
    const bool big = fmt.find("xdr") != std::string::npos;
    size_t got = 0;
    if (esize == 4) {
        got = std::fread(c.v.data(), 4, n, b);
        if (big)
            for (float& x : c.v) {
                uint32_t u;
                std::memcpy(&u, &x, 4);
                u = __builtin_bswap32(u);
                std::memcpy(&x, &u, 4);
            }
    } else if (esize == 8) {
        std::vector<double> tmp(n);
        got = std::fread(tmp.data(), 8, n, b);
        for (size_t i = 0; i < n; i++) {
            double x = tmp[i];
            if (big) {
                uint64_t u;
                std::memcpy(&u, &x, 8);
                u = __builtin_bswap64(u);
                std::memcpy(&x, &u, 8);
            }
            c.v[i] = (float)x;
        }
    } else {
        std::fclose(b);
        throw std::runtime_error("unsupported esize " + std::to_string(esize));
    }
    std::fclose(b);
    if (got != n) throw std::runtime_error("The binary file was not read properly");
    return c;
}

// Write the header and the samples (FP32). The binary goes to <dir>/<output>@ with dir
// the datapath, else $DATAPATH, else $PWD, else ".".
inline void write_cube(const Cube& c, const std::string& output, bool binary_format, const std::string& datapath) {
    std::string dir = datapath;
    if (dir == "none") {
        const char* p = std::getenv("DATAPATH");
        if (!p) p = std::getenv("PWD");
        dir = p ? p : ".";
    }
    const std::string bin = dir + "/" + output + "@";
    std::ofstream h(output);
    if (!h.is_open()) throw std::runtime_error("The output file " + output + " could not be created. Check for the path");
    for (size_t i = 0; i < c.axes.size(); i++) {
        h << "n" << i + 1 << "=" << c.axes[i].n << "\n";
        h << "o" << i + 1 << "=" << c.axes[i].o << "\n";
        h << "d" << i + 1 << "=" << c.axes[i].d << "\n";
        if (!binary_format && !c.axes[i].label.empty()) h << "label" << i + 1 << "=" << c.axes[i].label << "\n";
    }
    h << "esize=4\n";
    if (!binary_format) h << "data_format=\"native_float\"\n";
    h << "in=" << bin << "\n";
    h.close();
    FILE* b = std::fopen(bin.c_str(), "wb");
    if (!b) throw std::runtime_error("The output file " + bin + " could not be created. Check the path");
    const size_t n = c.count();
    const size_t put = std::fwrite(c.v.data(), 4, n, b);
    std::fclose(b);
    if (put != n) throw std::runtime_error("short write to " + bin);
}

}  // namespace fwi3d_cli
