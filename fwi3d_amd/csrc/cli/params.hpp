// Run parameters of the drop-in executables: the keys, defaults and parsing rules of the
// reference's param struct and readParameters (src/include/param.hpp:8-71,
// src/cpp/param.cpp:4-253), so that existing parfiles, command lines and srcoord files
// mean the same thing here.
//
//   * a value is given as key=value; the first argument that starts with "key=" wins and
//     the value is the first whitespace-separated token after "=", converted the way a
//     C++ stream converts it (failed conversions give 0 / false / "");
//   * parfile=<file>: every line not starting with '#' is one such argument; arguments
//     on the command line are applied after the parfile, so they override it;
//   * list values (B-spline controls and multiplicities) are comma-separated;
//   * srcoord=<file>: lines "shot_id sx sy sz rx ry rz rdip raz"; consecutive lines with
//     the same shot id add receivers to that shot (a line with a new id starts a shot;
//     a blank line repeats the previous id, with zero coordinates);
//     without the file, ns shots and nr receivers are laid out from sx0/sxinc, ...
#pragma once

#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace fwi3d_cli {

struct RunParam {
    // time
    int nt = 0, sub = 0;
    float courant = 0.5f, dt = -1.f, tmax = 0.f, fmax = 10.f;
    std::string resampling = "sinc";
    int sinc_half_length = 11;
    // geometry
    int ns = 1, nr = 1, nscomp = 3, nrcomp = 3;
    float sx0 = 0, sy0 = 0, sz0 = 0, rx0 = 0, ry0 = 0, rz0 = 0, rdip0 = 0, raz0 = 0;
    float sxinc = 0, syinc = 0, szinc = 0, rxinc = 0, ryinc = 0, rzinc = 0, rdipinc = 0, razinc = 0;
    std::string srcoord = "none";
    bool srcoord_from_file = false;
    std::vector<std::vector<float>> sxyz;                // [ns][3]
    std::vector<std::vector<std::vector<float>>> rxyz;   // [ns][nr_s][5] x, y, z, dip, azimuth
    float gl = 0;  // DAS gauge length (0: point receivers)
    // source mechanism, receiver type
    bool mt = false;
    float fangle = 0, fazimuth = 0;
    float mxx = 1, myy = 1, mzz = 1, mxy = 1, mxz = 1, myz = 1;
    int seismotype = 0;
    // boundaries
    int bc_top = 1, bc_bottom = 1, bc_left = 1, bc_right = 1, bc_front = 1, bc_back = 1;
    int taper_top = 0, taper_bottom = 0, taper_left = 0, taper_right = 0, taper_front = 0, taper_back = 0;
    float free_surface_stiffness = 1.05f, taper_strength = 0.05f;
    // model bounds
    float vpmin = 0.2f, vpmax = 8, vsmin = 0.1f, vsmax = 5, rhomin = 0.1f, rhomax = 8;
    float deltamin = -0.5f, deltamax = 1, epsilonmin = 0, epsilonmax = 1, gammamin = 0, gammamax = 1;
    float vmax = 8, vmin = 0.1f;
    // model parameterization
    int nmodels = 3, model_parameterization = 1;
    bool bsplines = false, soft_clip = false, inversion1d = false;
    int bs_nx = 3, bs_ny = 3, bs_nz = 3;
    std::vector<int> bs_mx, bs_my, bs_mz;
    std::vector<float> bs_controlx, bs_controly, bs_controlz;
    std::string horizon_file = "none";
    // inversion
    std::string nlsolver = "lbfgs", lsearch = "regular_wolfe", mask_file = "none", weights_file = "none",
                filter_file = "none", filter_phase = "default", inverse_diagonal_hessian_file = "none",
                prior_file = "none";
    float threshold = 0, lambda = 0, reg_xweight = 1, reg_yweight = 1, reg_zweight = 1, scale_source_log_clip = 1;
    int lbfgs_m = 5, niter = 0, max_trial = 10, isave = 10, envelop = 0, regularization = -1, scale_source_times = 0;
    bool normalize = false, normalize_obj_func = false;
    float ls_a0 = 1, ls_a1 = 0, ls_c1 = 1e-4f, ls_c2 = 0.9f, ls_max_step = 1e6f;
    bool ls_version = false;
    // miscellaneous
    int nthreads = 0, verbose = 1, compression_rate = 0;
    bool format = false;  // 0: SEP header (.H), 1: binary + header
    std::string datapath = "none";
};

// Stream conversion of one token (a failed conversion yields the zero value).
template <class T>
T from_token(const std::string& s) {
    std::istringstream ss(s);
    T v{};
    ss >> v;
    return v;
}

// Value of the first argument of the form key=value (args[0] is the program name).
inline bool arg_value(const std::vector<std::string>& args, const std::string& key, std::string& out) {
    const std::string pre = key + "=";
    for (size_t i = 1; i < args.size(); i++)
        if (args[i].compare(0, pre.size(), pre) == 0) {
            out = args[i].substr(pre.size());
            return true;
        }
    return false;
}

template <class T>
void get(const std::vector<std::string>& args, const std::string& key, T& v) {
    std::string s;
    if (arg_value(args, key, s)) v = from_token<T>(s);
}

template <class T>
void get(const std::vector<std::string>& args, const std::string& key, std::vector<T>& v) {
    std::string s;
    if (!arg_value(args, key, s)) return;
    v.clear();
    size_t start = 0;
    for (;;) {
        const size_t comma = s.find(',', start);
        v.push_back(from_token<T>(s.substr(start, comma == std::string::npos ? std::string::npos : comma - start)));
        if (comma == std::string::npos) break;
        start = comma + 1;
    }
}

// Every key the drivers read from the command line or a parfile.
inline void apply_args(const std::vector<std::string>& a, RunParam& p) {
    get(a, "resampling", p.resampling); get(a, "srcoord", p.srcoord); get(a, "horizon", p.horizon_file);
    get(a, "nlsolver", p.nlsolver); get(a, "lsearch", p.lsearch); get(a, "mask", p.mask_file);
    get(a, "weights", p.weights_file); get(a, "filter", p.filter_file);
    get(a, "inverse_diagonal_hessian", p.inverse_diagonal_hessian_file); get(a, "prior", p.prior_file);
    get(a, "datapath", p.datapath);
    get(a, "courant", p.courant); get(a, "dt", p.dt); get(a, "fmax", p.fmax);
    get(a, "sx0", p.sx0); get(a, "sy0", p.sy0); get(a, "sz0", p.sz0);
    get(a, "rx0", p.rx0); get(a, "ry0", p.ry0); get(a, "rz0", p.rz0); get(a, "rdip0", p.rdip0); get(a, "raz0", p.raz0);
    get(a, "sxinc", p.sxinc); get(a, "syinc", p.syinc); get(a, "szinc", p.szinc);
    get(a, "rxinc", p.rxinc); get(a, "ryinc", p.ryinc); get(a, "rzinc", p.rzinc);
    get(a, "rdipinc", p.rdipinc); get(a, "razinc", p.razinc);
    get(a, "free_surface_stiffness", p.free_surface_stiffness); get(a, "taper_strength", p.taper_strength);
    get(a, "gl", p.gl);
    get(a, "mxx", p.mxx); get(a, "myy", p.myy); get(a, "mzz", p.mzz); get(a, "mxy", p.mxy); get(a, "mxz", p.mxz);
    get(a, "myz", p.myz); get(a, "fangle", p.fangle); get(a, "fazimuth", p.fazimuth);
    get(a, "vpmin", p.vpmin); get(a, "vpmax", p.vpmax); get(a, "vsmin", p.vsmin); get(a, "vsmax", p.vsmax);
    get(a, "rhomin", p.rhomin); get(a, "rhomax", p.rhomax); get(a, "deltamin", p.deltamin);
    get(a, "deltamax", p.deltamax); get(a, "epsilonmin", p.epsilonmin); get(a, "epsilonmax", p.epsilonmax);
    get(a, "gammamin", p.gammamin); get(a, "gammamax", p.gammamax);
    get(a, "bs_controlx", p.bs_controlx); get(a, "bs_controly", p.bs_controly); get(a, "bs_controlz", p.bs_controlz);
    get(a, "threshold", p.threshold); get(a, "ls_a0", p.ls_a0); get(a, "ls_a1", p.ls_a1); get(a, "ls_c1", p.ls_c1);
    get(a, "ls_c2", p.ls_c2); get(a, "ls_max_step", p.ls_max_step); get(a, "lambda", p.lambda);
    get(a, "reg_xweight", p.reg_xweight); get(a, "reg_yweight", p.reg_yweight); get(a, "reg_zweight", p.reg_zweight);
    get(a, "scale_source_log_clip", p.scale_source_log_clip);
    get(a, "ns", p.ns); get(a, "nr", p.nr); get(a, "seismotype", p.seismotype);
    get(a, "bc_top", p.bc_top); get(a, "bc_bottom", p.bc_bottom); get(a, "bc_left", p.bc_left);
    get(a, "bc_right", p.bc_right); get(a, "bc_front", p.bc_front); get(a, "bc_back", p.bc_back);
    get(a, "taper_top", p.taper_top); get(a, "taper_bottom", p.taper_bottom); get(a, "taper_left", p.taper_left);
    get(a, "taper_right", p.taper_right); get(a, "taper_front", p.taper_front); get(a, "taper_back", p.taper_back);
    get(a, "sinc_half_length", p.sinc_half_length); get(a, "sub", p.sub);
    get(a, "model_parameterization", p.model_parameterization);
    get(a, "bs_nx", p.bs_nx); get(a, "bs_ny", p.bs_ny); get(a, "bs_nz", p.bs_nz);
    get(a, "lbfgs_m", p.lbfgs_m); get(a, "niter", p.niter); get(a, "max_trial", p.max_trial); get(a, "isave", p.isave);
    get(a, "envelop", p.envelop); get(a, "regularization", p.regularization);
    get(a, "scale_source_times", p.scale_source_times); get(a, "verbose", p.verbose); get(a, "nthreads", p.nthreads);
    get(a, "compression_rate", p.compression_rate);
    get(a, "mt", p.mt); get(a, "bsplines", p.bsplines); get(a, "soft_clip", p.soft_clip);
    get(a, "inversion1d", p.inversion1d); get(a, "normalize", p.normalize);
    get(a, "normalize_obj_func", p.normalize_obj_func); get(a, "ls_version", p.ls_version); get(a, "format", p.format);
    get(a, "bs_mx", p.bs_mx); get(a, "bs_my", p.bs_my); get(a, "bs_mz", p.bs_mz);
}

// The lines of a parfile as arguments (a missing file adds nothing).
inline std::vector<std::string> parfile_args(const std::string& path) {
    std::vector<std::string> a = {"parfile"};
    std::ifstream f(path);
    std::string line;
    while (f && std::getline(f, line))
        if (line.empty() || line[0] != '#') a.push_back(line + " ");
    return a;
}

// srcoord file: shots and their receivers (a missing file leaves the layout to the keys).
inline void read_srcoord(const std::string& path, RunParam& p) {
    std::ifstream f(path);
    if (!f.is_open()) return;
    p.sxyz.clear();
    p.rxyz.clear();
    p.srcoord_from_file = true;
    std::string line;
    bool first = true;
    int prev_id = 0;
    float sid = 0;  // kept across lines: a blank line leaves the previous shot id
    while (std::getline(f, line)) {
        if (!line.empty() && line[0] == '#') continue;
        std::istringstream ss(line);
        ss >> sid;
        const int id = (int)sid;
        std::vector<float> s(3, 0.f), r(5, 0.f);
        for (float& v : s) ss >> v;
        for (float& v : r) ss >> v;
        if (first || id != prev_id) {
            p.sxyz.push_back(s);
            p.rxyz.push_back({r});
        } else {
            p.rxyz.back().push_back(r);
        }
        first = false;
        prev_id = id;
    }
}

// readParameters: parfile, then the command line, then the geometry.
inline void read_parameters(const std::vector<std::string>& args, RunParam& p) {
    std::string parfile = "none";
    get(args, "parfile", parfile);
    apply_args(parfile_args(parfile), p);
    apply_args(args, p);
    read_srcoord(p.srcoord, p);
    if (p.srcoord_from_file) {
        p.ns = (int)p.sxyz.size();
        return;
    }
    p.sxyz.clear();
    p.rxyz.clear();
    for (int s = 0; s < p.ns; s++) {
        p.sxyz.push_back({p.sx0 + s * p.sxinc, p.sy0 + s * p.syinc, p.sz0 + s * p.szinc});
        std::vector<std::vector<float>> rs;
        for (int r = 0; r < p.nr; r++)
            rs.push_back({p.rx0 + r * p.rxinc, p.ry0 + r * p.ryinc, p.rz0 + r * p.rzinc, p.rdip0 + r * p.rdipinc,
                          p.raz0 + r * p.razinc});
        p.rxyz.push_back(rs);
    }
}

}  // namespace fwi3d_cli
