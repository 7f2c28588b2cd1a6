"""Output_udf records of a run (output.cxx:268-375) written as a UDF file.

The reference appends one record per frame to its output UDF (every GTS
steps, before the step of that ts, sp_3d_ns.cxx:863-950): E (1.0), t (the
step count) and, per particle, the members of class outParticle in the
order of its definition -- R, R_raw, v, q, omega, then the previous step's
f_hydro, torque_hydro, f_r, torque_r (0), f_slip, torque_slip. Frames come
from the device (Run.get_output, or Server.frames for served runs) as
FRAME_PARTICLE_DTYPE records already in that order; this module turns them
into that record form and writes a self-describing UDF file: the `def`
section of the record, then one data block per frame.
"""
import numpy as np

from . import udf

# the record's definition (the output schema's def section: outParticle, E,
# t, Particles[]); class and member names are those the reference writes
RECORD_DEF = """\\begin{def}
class Vector3d:{x: float, y: float, z: float}
class Quaternion:{q0: float, q1: float, q2: float, q3: float}
class outParticle:{R:Vector3d, R_raw:Vector3d, v:Vector3d, q:Quaternion, omega:Vector3d,
                   f_hydro:Vector3d, torque_hydro:Vector3d, f_r:Vector3d, torque_r:Vector3d,
                   f_slip:Vector3d, torque_slip:Vector3d}
E: float "total kinetic energy of the system"
t: float "total time"
Particles[]: outParticle
\\end{def}
"""

_VEC = ("R", "R_raw", "v", "omega", "f_hydro", "torque_hydro", "f_r", "torque_r", "f_slip", "torque_slip")


def frame_records(t, frames):
    """Output_udf's records for frames (t: step counts, frames: (nframes, np)
    FRAME_PARTICLE_DTYPE) as nested dicts keyed by the record's names."""
    out = []
    for k in range(len(t)):
        parts = []
        for p in frames[k]:
            rec = {m: {"x": float(p[m][0]), "y": float(p[m][1]), "z": float(p[m][2])} for m in _VEC}
            rec["q"] = {f"q{i}": float(p["q"][i]) for i in range(4)}
            parts.append(rec)
        out.append({"E": 1.0, "t": float(t[k]), "Particles": parts})
    return out


def write_output_udf(path, t, frames):
    """Write the frames as Output_udf records into a UDF file at path."""
    _, schema, classes, _ = _schema()
    out = [RECORD_DEF]
    for rec in frame_records(t, frames):
        out.append("\\begin{data}\n")
        for name in ("E", "t", "Particles"):
            dims, typ = schema[name]
            out.append(name + ":")
            udf._dump(rec[name], dims, typ, classes, out)
            out.append("\n")
        out.append("\\end{data}\n")
    with open(path, "w") as f:
        f.write("".join(out))


def read_output_udf(path):
    """(t, frames) back from an output UDF file (FRAME_PARTICLE_DTYPE)."""
    from . import FRAME_PARTICLE_DTYPE
    recs = udf.read_udf_records(path)
    t = np.array([int(round(r["t"])) for r in recs], dtype=np.int64)
    np_ = len(recs[0]["Particles"]) if recs else 0
    frames = np.zeros((len(recs), np_), dtype=FRAME_PARTICLE_DTYPE)
    for k, r in enumerate(recs):
        for j, p in enumerate(r["Particles"]):
            for m in _VEC:
                frames[k, j][m] = [p[m]["x"], p[m]["y"], p[m]["z"]]
            frames[k, j]["q"] = [p["q"][f"q{i}"] for i in range(4)]
    return t, frames


def _schema():
    return udf.parse_udf_text(RECORD_DEF)
