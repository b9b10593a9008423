"""Host layer of the drop-in executables (WE_MODELING_3D.x, FWI3D.x): the run-parameter
parser against the reference's own readParameters / readCoord (src/cpp/param.cpp:4-253,
compiled from the reference checkout into oracle/_ref/libref_param.so), field by field,
and the hypercube files (src/include/IO.hpp:181-430) against their format.
No GPU is used."""
import ctypes as C
import json
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "fwi3d_amd", "libfwi3d_cli.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libref_param.so")


@pytest.fixture(scope="module")
def cli():
    if not os.path.exists(CLI):
        pytest.fail("fwi3d_amd/libfwi3d_cli.so is not built (python -m fwi3d_amd.build)")
    return C.CDLL(CLI)


def _argv(args):
    arr = (C.c_char_p * len(args))(*[a.encode() for a in args])
    return len(args), arr


def parse(lib, fn, args):
    argc, argv = _argv(["prog"] + list(args))
    buf = C.create_string_buffer(1 << 20)
    rc = getattr(lib, fn)(argc, argv, buf, len(buf))
    assert rc == 0
    return json.loads(buf.value.decode())


CASES = [
    [],
    ["ns=3", "nr=4", "sx0=0.1", "sxinc=0.05", "sy0=0.2", "sz0=0.03", "rx0=0.01", "rxinc=0.02", "ry0=0.5",
     "rz0=0.02", "rdip0=0.3", "rdipinc=0.1", "raz0=-0.2", "razinc=0.05"],
    ["bc_top=2", "bc_bottom=0", "taper_left=20", "taper_right=15", "taper_front=3", "taper_back=7",
     "taper_strength=0.08", "free_surface_stiffness=1.2", "courant=0.45", "dt=-1", "sub=-1", "fmax=25",
     "resampling=linear", "sinc_half_length=7", "seismotype=1", "mt=1", "mxx=2.5", "myz=-0.5", "fangle=0.3",
     "fazimuth=1.1", "gl=0.01"],
    ["vpmin=1.5", "vpmax=4.5", "vsmin=0.6", "vsmax=2.9", "rhomin=1.9", "rhomax=2.8", "deltamin=-0.1",
     "deltamax=0.3", "epsilonmin=0.01", "epsilonmax=0.4", "gammamin=0.02", "gammamax=0.5",
     "model_parameterization=3", "soft_clip=1", "bsplines=1", "bs_controlx=0,0.5,1.5", "bs_mx=2,1,2",
     "bs_controlz=0.1,0.2", "bs_mz=2,2", "inversion1d=0", "horizon=h.H"],
    ["nlsolver=nlcg", "lsearch=strong_wolfe", "niter=15", "max_trial=7", "threshold=1e-3", "lbfgs_m=8",
     "ls_a0=0.5", "ls_a1=0.1", "ls_c1=1e-3", "ls_c2=0.8", "ls_max_step=100", "ls_version=1", "isave=3",
     "envelop=2", "normalize=1", "normalize_obj_func=1", "regularization=1", "lambda=0.25", "reg_xweight=2",
     "reg_yweight=3", "reg_zweight=0.5", "scale_source_times=2", "scale_source_log_clip=0.3",
     "mask=m.H", "weights=w.H", "filter=f.H", "inverse_diagonal_hessian=ih.H", "prior=p.H", "datapath=/tmp/x",
     "verbose=3", "nthreads=4", "compression_rate=8", "format=1"],
    # the first occurrence of a key wins; values are the first token; bool/int from streams
    ["ns=2", "ns=5", "niter=12abc", "mt=true", "format=2", "sx0=1e-2", "resampling=sinc extra"],
]


@pytest.mark.skipif(not os.path.exists(REF), reason="reference parser not built (needs the reference checkout)")
@pytest.mark.parametrize("k", range(len(CASES)))
def test_command_line_matches_reference(cli, k):
    ref = C.CDLL(REF)
    a = parse(cli, "fwi3d_cli_params_json", CASES[k])
    b = parse(ref, "ref_read_parameters_json", CASES[k])
    assert a == b


@pytest.mark.skipif(not os.path.exists(REF), reason="reference parser not built (needs the reference checkout)")
def test_parfile_srcoord_match_reference(cli, tmp_path):
    ref = C.CDLL(REF)
    par = tmp_path / "run.p"
    par.write_text("# comment line\nbc_top=1\ntaper_bottom=25\nniter=4 trailing words\n  verbose=2\n\n"
                   "mt=1\nsub=4\nsrcoord=" + str(tmp_path / "geom.txt") + "\ncourant=0.3\n")
    (tmp_path / "geom.txt").write_text(
        "# shot x y z rx ry rz dip az\n"
        "0 0.10 0.20 0.03 0.01 0.02 0.02 0 0\n"
        "0 0.10 0.20 0.03 0.05 0.02 0.02 0.1 0.2\n"
        "0 0.10 0.20 0.03 0.09 0.02 0.02\n"        # short line: missing values read as 0
        "1 0.30 0.25 0.03 0.01 0.04 0.02 0 0\n"
        "1 0.30 0.25 0.03 0.03 0.04 0.02 0 0\n"
        "\n"                                       # a blank line: previous shot id, zero coordinates
        "2 0.50 0.30 0.04 0.11 0.12 0.13 0.5 1.5\n")
    for extra in ([], ["courant=0.45", "taper_bottom=10"], ["srcoord=none", "ns=2", "nr=3", "rxinc=0.01"]):
        args = ["parfile=" + str(par)] + extra
        a = parse(cli, "fwi3d_cli_params_json", args)
        b = parse(ref, "ref_read_parameters_json", args)
        assert a == b
    a = parse(cli, "fwi3d_cli_params_json", ["parfile=" + str(par)])
    assert a["srcoord_from_file"] == 1 and a["ns"] == 3 and len(a["rxyz"][0]) == 3 and len(a["rxyz"][1]) == 3
    assert a["courant"] == pytest.approx(0.3) and a["niter"] == 4 and a["verbose"] == 1  # "  verbose" is no key


def test_defaults_are_the_reference_param_defaults(cli):
    """src/include/param.hpp:8-71 (independent of the reference build)."""
    a = parse(cli, "fwi3d_cli_params_json", [])
    want = dict(courant=0.5, dt=-1, fmax=10, resampling="sinc", sinc_half_length=11, ns=1, nr=1, mt=0, mxx=1,
                bc_top=1, bc_back=1, taper_top=0, free_surface_stiffness=1.05, taper_strength=0.05, vpmin=0.2,
                vpmax=8, vsmin=0.1, vsmax=5, rhomin=0.1, rhomax=8, deltamin=-0.5, deltamax=1, epsilonmin=0,
                epsilonmax=1, gammamin=0, gammamax=1, model_parameterization=1, nlsolver="lbfgs",
                lsearch="regular_wolfe", lbfgs_m=5, niter=0, max_trial=10, isave=10, regularization=-1, ls_a0=1,
                ls_c1=1e-4, ls_c2=0.9, ls_max_step=1e6, verbose=1, format=0, datapath="none", srcoord="none")
    for k, v in want.items():
        assert a[k] == (pytest.approx(v) if isinstance(v, float) else v), k
    assert a["sxyz"] == [[0, 0, 0]] and a["rxyz"] == [[[0, 0, 0, 0, 0]]]


def _write(cli, path, data, axes, fmt, datapath):
    n = (C.c_int * 6)(*[a[0] for a in axes])
    o = (C.c_float * 6)(*[a[1] for a in axes])
    d = (C.c_float * 6)(*[a[2] for a in axes])
    err = C.create_string_buffer(256)
    data = np.ascontiguousarray(data, np.float32)
    rc = cli.fwi3d_cli_write_cube(str(path).encode(), len(axes), n, o, d, C.c_void_p(data.ctypes.data), fmt,
                                  str(datapath).encode(), err)
    assert rc == 0, err.value


def _read(cli, path, ndim0=1, cap=1 << 22):
    nd = C.c_int()
    n, o, d = (C.c_int * 6)(), (C.c_float * 6)(), (C.c_float * 6)()
    out = np.zeros(cap, np.float32)
    err = C.create_string_buffer(256)
    rc = cli.fwi3d_cli_read_cube(str(path).encode(), ndim0, C.byref(nd), n, o, d, C.c_void_p(out.ctypes.data),
                                 cap, err)
    return rc, err.value.decode(), nd.value, list(n)[: nd.value], list(o)[: nd.value], list(d)[: nd.value], out


@pytest.mark.parametrize("fmt", [1, 0])
def test_cube_write_format_and_round_trip(cli, tmp_path, monkeypatch, fmt):
    rng = np.random.default_rng(3)
    data = rng.standard_normal((2, 5, 4, 7)).astype(np.float32)  # [n4][n3][n2][n1]
    axes = [(7, 0.0, 0.01), (4, 0.5, 0.02), (5, -1.0, 0.03), (2, 0.0, 1.0)]
    monkeypatch.chdir(tmp_path)
    hdr = "m.H"  # the binary goes to <datapath>/<output>@
    _write(cli, hdr, data, axes, fmt, tmp_path)
    text = open(hdr).read()
    kv = dict(line.split("=", 1) for line in text.split())
    assert kv["n1"] == "7" and kv["n2"] == "4" and kv["n3"] == "5" and kv["n4"] == "2"
    assert float(kv["d2"]) == pytest.approx(0.02) and float(kv["o3"]) == -1 and kv["esize"] == "4"
    assert kv["in"] == str(tmp_path) + "/m.H@"
    assert ("data_format" in kv) == (fmt == 0)
    raw = np.fromfile(kv["in"], dtype="<f4")
    assert np.array_equal(raw, data.ravel())
    rc, err, nd, n, o, d, out = _read(cli, hdr, 4)
    assert rc == 0, err
    assert nd == 4 and n == [7, 4, 5, 2]
    assert np.array_equal(out[: data.size], data.ravel())


def test_cube_read_header_variants(cli, tmp_path):
    """Several words per line, comments, a quoted in=, trailing unit axes dropped unless
    kept by ndim0, esize=8 samples and SEPlib's big-endian xdr_float samples."""
    v = np.arange(24, dtype=np.float64).reshape(2, 3, 4) * 0.5
    (tmp_path / "a.bin").write_bytes(v.astype("<f8").tobytes())
    (tmp_path / "a.H").write_text(f'# a comment n1=99\nn1=4 o1=0.1 d1=0.25 n2=3\nd2=2 n3=2 n4=1\nesize=8\n'
                                  f'in="{tmp_path}/a.bin"\n')
    rc, err, nd, n, o, d, out = _read(cli, tmp_path / "a.H")
    assert rc == 0, err
    assert nd == 3 and n == [4, 3, 2] and o[0] == pytest.approx(0.1) and d[1] == 2
    assert np.array_equal(out[:24], v.ravel().astype(np.float32))
    rc, _, nd, n, *_ = _read(cli, tmp_path / "a.H", ndim0=4)
    assert nd == 4 and n == [4, 3, 2, 1]
    (tmp_path / "b.bin").write_bytes(v.astype(">f4").tobytes())
    (tmp_path / "b.H").write_text(f'n1=4 n2=6\ndata_format="xdr_float"\nin={tmp_path}/b.bin\n')
    rc, err, nd, n, *_, out = _read(cli, tmp_path / "b.H")
    assert rc == 0 and n == [4, 6] and np.array_equal(out[:24], v.ravel().astype(np.float32))
    # size mismatch is an error, as in the reference
    (tmp_path / "c.H").write_text(f"n1=5 n2=6\nin={tmp_path}/b.bin\n")
    rc, err, *_ = _read(cli, tmp_path / "c.H")
    assert rc != 0 and "bytes" in err
