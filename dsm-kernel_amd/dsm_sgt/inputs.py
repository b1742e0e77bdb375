"""Readers of the reference's input files into dsm_sgt_input_t.

Grammar of src/SGTsh/others.f90:1-110 (pinput, readDSMconf, readpsvmodel):
lines whose first column is '#' are dropped, reading stops at a line that
starts with 'end'; the parameter file's first four records are a120
strings (DSM.config path, output directory, model path, model name), the
rest is list-directed: tlen / rmin rmax rdelta / r0 / thetamin thetamax
thetadelta / imin imax / 3 or 4 switches (rsgt tsgt synn [psgt]).
The model file is nzone followed by 28 numbers per zone, each zone read
starting on a fresh record (SGTpsv-MPI.f90:100-105).
"""
import ctypes as C
import os

MAXZONE = 64


class DsmSgtInput(C.Structure):
    """Layout of dsm_sgt_input_t (include/dsm_sgt.h)."""
    _fields_ = [
        ("nzone", C.c_int32), ("maxlmax", C.c_int32),
        ("imin", C.c_int32), ("imax", C.c_int32),
        ("rsgtswitch", C.c_int32), ("tsgtswitch", C.c_int32),
        ("synnswitch", C.c_int32), ("psgtswitch", C.c_int32),
        ("vrmin", C.c_double * MAXZONE), ("vrmax", C.c_double * MAXZONE),
        ("rho", (C.c_double * 4) * MAXZONE), ("vpv", (C.c_double * 4) * MAXZONE),
        ("vph", (C.c_double * 4) * MAXZONE), ("vsv", (C.c_double * 4) * MAXZONE),
        ("vsh", (C.c_double * 4) * MAXZONE), ("eta", (C.c_double * 4) * MAXZONE),
        ("qmu", C.c_double * MAXZONE), ("qkappa", C.c_double * MAXZONE),
        ("re", C.c_double), ("ratc", C.c_double), ("ratl", C.c_double),
        ("omegai_seed", C.c_double), ("tlen", C.c_double),
        ("rmin_", C.c_double), ("rmax_", C.c_double), ("rdelta_", C.c_double),
        ("r0min", C.c_double),
        ("thetamin", C.c_double), ("thetamax", C.c_double), ("thetadelta", C.c_double),
        ("r_list_n", C.c_int32), ("theta_list_n", C.c_int32),
        ("r_list", C.c_void_p), ("theta_list", C.c_void_p),
    ]


def _records(text):
    recs = []
    for line in text.splitlines():
        if line[:1] == "#":
            continue
        if line[:3] == "end":
            break
        recs.append(line)
    return recs


def _f(tok):
    return float(tok.replace("d", "e").replace("D", "e"))


def _fields(rec):
    return rec.replace(",", " ").split()


def read_model(text):
    recs = _records(text)
    nzone = int(_fields(recs[0])[0])
    if not 1 <= nzone <= MAXZONE:
        raise ValueError(f"nzone={nzone} outside [1,{MAXZONE}]")
    zones, k = [], 1
    for _ in range(nzone):
        vals = []
        while len(vals) < 28:
            if k >= len(recs):
                raise ValueError("model file ends inside a zone")
            vals += [_f(t) for t in _fields(recs[k])]
            k += 1
        zones.append(vals[:28])
    return zones


def read_dsm_config(text):
    recs = [r for r in _records(text) if r.strip()]
    v = [_fields(recs[i])[0] for i in range(5)]
    return dict(re=_f(v[0]), ratc=_f(v[1]), ratl=_f(v[2]), omegai_seed=_f(v[3]), maxlmax=int(v[4]))


def read_params(text):
    recs = _records(text)
    p = dict(dsmconf=recs[0].strip(), outputdir=recs[1].strip(), model=recs[2].strip(),
             modelname=recs[3].strip())
    rest = [_fields(r) for r in recs[4:] if r.strip()]
    p["tlen"] = _f(rest[0][0])
    p["rmin_"], p["rmax_"], p["rdelta_"] = (_f(t) for t in rest[1][:3])
    p["r0min"] = _f(rest[2][0])
    p["thetamin"], p["thetamax"], p["thetadelta"] = (_f(t) for t in rest[3][:3])
    p["imin"], p["imax"] = int(rest[4][0]), int(rest[4][1])
    sw = [int(t) for t in rest[5][:4]]
    sw += [0] * (4 - len(sw))
    p["rsgtswitch"], p["tsgtswitch"], p["synnswitch"], p["psgtswitch"] = sw
    return p


def make_input(zones, conf, params):
    s = DsmSgtInput()
    s.nzone = len(zones)
    for i, z in enumerate(zones):
        s.vrmin[i], s.vrmax[i] = z[0], z[1]
        for j in range(4):
            s.rho[i][j], s.vpv[i][j], s.vph[i][j] = z[2 + j], z[6 + j], z[10 + j]
            s.vsv[i][j], s.vsh[i][j], s.eta[i][j] = z[14 + j], z[18 + j], z[22 + j]
        s.qmu[i], s.qkappa[i] = z[26], z[27]
    for k in ("re", "ratc", "ratl", "omegai_seed", "maxlmax"):
        setattr(s, k, conf[k])
    for k in ("tlen", "rmin_", "rmax_", "rdelta_", "r0min", "thetamin", "thetamax", "thetadelta",
              "imin", "imax", "rsgtswitch", "tsgtswitch", "synnswitch", "psgtswitch"):
        setattr(s, k, params[k])
    return s


def read_inputs(param_text, base_dir="."):
    """Parse a parameter file and the DSM.config / model files it names
    (relative paths resolved against base_dir).  Returns (input, params)."""
    p = read_params(param_text)

    def rd(path):
        with open(path if os.path.isabs(path) else os.path.join(base_dir, path)) as f:
            return f.read()

    return make_input(read_model(rd(p["model"])), read_dsm_config(rd(p["dsmconf"])), p), p
