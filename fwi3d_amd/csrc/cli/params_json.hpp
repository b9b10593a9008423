// Text dump of the parsed run parameters (every key the drivers read, plus the shot
// geometry), used to compare this parser with the reference's readParameters.
#pragma once

#include <cstdio>
#include <string>
#include <vector>

#include "params.hpp"

namespace fwi3d_cli {

inline std::string jnum(double x) {
    char b[64];
    std::snprintf(b, sizeof b, "%.9g", x);
    return b;
}
inline std::string jstr(const std::string& s) {
    std::string o = "\"";
    for (char c : s) {
        if (c == '"' || c == '\\') o += '\\';
        o += c;
    }
    return o + "\"";
}
template <class T>
std::string jlist(const std::vector<T>& v) {
    std::string o = "[";
    for (size_t i = 0; i < v.size(); i++) o += (i ? "," : "") + jnum((double)v[i]);
    return o + "]";
}

inline std::string params_json(const RunParam& p) {
    std::string o = "{";
    auto kv = [&](const char* k, const std::string& v) { o += (o.size() > 1 ? "," : "") + jstr(k) + ":" + v; };
    auto f = [&](const char* k, double v) { kv(k, jnum(v)); };
    kv("resampling", jstr(p.resampling)); kv("srcoord", jstr(p.srcoord)); kv("horizon", jstr(p.horizon_file));
    kv("nlsolver", jstr(p.nlsolver)); kv("lsearch", jstr(p.lsearch)); kv("mask", jstr(p.mask_file));
    kv("weights", jstr(p.weights_file)); kv("filter", jstr(p.filter_file));
    kv("inverse_diagonal_hessian", jstr(p.inverse_diagonal_hessian_file)); kv("prior", jstr(p.prior_file));
    kv("datapath", jstr(p.datapath));
    f("courant", p.courant); f("dt", p.dt); f("fmax", p.fmax);
    f("sx0", p.sx0); f("sy0", p.sy0); f("sz0", p.sz0); f("rx0", p.rx0); f("ry0", p.ry0); f("rz0", p.rz0);
    f("rdip0", p.rdip0); f("raz0", p.raz0); f("sxinc", p.sxinc); f("syinc", p.syinc); f("szinc", p.szinc);
    f("rxinc", p.rxinc); f("ryinc", p.ryinc); f("rzinc", p.rzinc); f("rdipinc", p.rdipinc); f("razinc", p.razinc);
    f("free_surface_stiffness", p.free_surface_stiffness); f("taper_strength", p.taper_strength); f("gl", p.gl);
    f("mxx", p.mxx); f("myy", p.myy); f("mzz", p.mzz); f("mxy", p.mxy); f("mxz", p.mxz); f("myz", p.myz);
    f("fangle", p.fangle); f("fazimuth", p.fazimuth);
    f("vpmin", p.vpmin); f("vpmax", p.vpmax); f("vsmin", p.vsmin); f("vsmax", p.vsmax); f("rhomin", p.rhomin);
    f("rhomax", p.rhomax); f("deltamin", p.deltamin); f("deltamax", p.deltamax); f("epsilonmin", p.epsilonmin);
    f("epsilonmax", p.epsilonmax); f("gammamin", p.gammamin); f("gammamax", p.gammamax);
    kv("bs_controlx", jlist(p.bs_controlx)); kv("bs_controly", jlist(p.bs_controly));
    kv("bs_controlz", jlist(p.bs_controlz));
    f("threshold", p.threshold); f("ls_a0", p.ls_a0); f("ls_a1", p.ls_a1); f("ls_c1", p.ls_c1); f("ls_c2", p.ls_c2);
    f("ls_max_step", p.ls_max_step); f("lambda", p.lambda); f("reg_xweight", p.reg_xweight);
    f("reg_yweight", p.reg_yweight); f("reg_zweight", p.reg_zweight); f("scale_source_log_clip", p.scale_source_log_clip);
    f("ns", p.ns); f("nr", p.nr); f("seismotype", p.seismotype);
    f("bc_top", p.bc_top); f("bc_bottom", p.bc_bottom); f("bc_left", p.bc_left); f("bc_right", p.bc_right);
    f("bc_front", p.bc_front); f("bc_back", p.bc_back); f("taper_top", p.taper_top); f("taper_bottom", p.taper_bottom);
    f("taper_left", p.taper_left); f("taper_right", p.taper_right); f("taper_front", p.taper_front);
    f("taper_back", p.taper_back); f("sinc_half_length", p.sinc_half_length); f("sub", p.sub);
    f("model_parameterization", p.model_parameterization); f("bs_nx", p.bs_nx); f("bs_ny", p.bs_ny);
    f("bs_nz", p.bs_nz); f("lbfgs_m", p.lbfgs_m); f("niter", p.niter); f("max_trial", p.max_trial);
    f("isave", p.isave); f("envelop", p.envelop); f("regularization", p.regularization);
    f("scale_source_times", p.scale_source_times); f("verbose", p.verbose); f("nthreads", p.nthreads);
    f("compression_rate", p.compression_rate);
    f("mt", p.mt); f("bsplines", p.bsplines); f("soft_clip", p.soft_clip); f("inversion1d", p.inversion1d);
    f("normalize", p.normalize); f("normalize_obj_func", p.normalize_obj_func); f("ls_version", p.ls_version);
    f("format", p.format);
    kv("bs_mx", jlist(p.bs_mx)); kv("bs_my", jlist(p.bs_my)); kv("bs_mz", jlist(p.bs_mz));
    f("srcoord_from_file", p.srcoord_from_file);
    std::string s = "[";
    for (size_t i = 0; i < p.sxyz.size(); i++) s += (i ? "," : "") + jlist(p.sxyz[i]);
    kv("sxyz", s + "]");
    s = "[";
    for (size_t i = 0; i < p.rxyz.size(); i++) {
        s += i ? ",[" : "[";
        for (size_t r = 0; r < p.rxyz[i].size(); r++) s += (r ? "," : "") + jlist(p.rxyz[i][r]);
        s += "]";
    }
    kv("rxyz", s + "]");
    return o + "}";
}

}  // namespace fwi3d_cli
