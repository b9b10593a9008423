// C ABI of the drop-in executables' host layer (parameter parsing and hypercube files),
// built into fwi3d_amd/libfwi3d_cli.so so that the parsing and the file formats can be
// checked without a GPU.
#include <cstring>

#include "cube_io.hpp"
#include "params_json.hpp"

extern "C" {

// readParameters on argv (argv[0] is the program name), dumped as JSON into out[cap].
// Returns 0, or -(needed bytes) when cap is too small.
int fwi3d_cli_params_json(int argc, char** argv, char* out, int cap) {
    fwi3d_cli::RunParam p;
    fwi3d_cli::read_parameters(std::vector<std::string>(argv, argv + argc), p);
    const std::string j = fwi3d_cli::params_json(p);
    if ((int)j.size() + 1 > cap) return -(int)j.size() - 1;
    std::memcpy(out, j.c_str(), j.size() + 1);
    return 0;
}

// Read a hypercube file: axes (n, o, d) into n[6], o[6], d[6], *ndim; samples into
// data[cap] (NULL: header only). Returns 0, -1 on error (message in err[256]).
int fwi3d_cli_read_cube(const char* header, int ndim0, int* ndim, int* n, float* o, float* d, float* data,
                        long long cap, char* err) {
    try {
        const fwi3d_cli::Cube c = fwi3d_cli::read_cube(header, ndim0);
        *ndim = (int)c.axes.size();
        for (int i = 0; i < 6; i++) {
            n[i] = i < *ndim ? c.axes[i].n : 1;
            o[i] = i < *ndim ? c.axes[i].o : 0.f;
            d[i] = i < *ndim ? c.axes[i].d : 1.f;
        }
        if (data) {
            if ((long long)c.v.size() > cap) throw std::runtime_error("buffer too small");
            std::memcpy(data, c.v.data(), c.v.size() * sizeof(float));
        }
        return 0;
    } catch (const std::exception& e) {
        if (err) std::snprintf(err, 256, "%s", e.what());
        return -1;
    }
}

// Write a hypercube file (format 1: binary + header, 0: SEP header).
int fwi3d_cli_write_cube(const char* header, int ndim, const int* n, const float* o, const float* d, const float* data,
                         int format, const char* datapath, char* err) {
    try {
        fwi3d_cli::Cube c;
        for (int i = 0; i < ndim; i++) c.axes.push_back(fwi3d_cli::Axis{n[i], o[i], d[i], ""});
        c.v.assign(data, data + c.count());
        fwi3d_cli::write_cube(c, header, format != 0, datapath ? datapath : "none");
        return 0;
    } catch (const std::exception& e) {
        if (err) std::snprintf(err, 256, "%s", e.what());
        return -1;
    }
}

}  // extern "C"
