"""Reader for KAPSEL's UDF input decks and restart records.

A deck (e.g. Examples/02/gravity.udf) is a `\\begin{data}` section of
positional values whose meaning comes from the schema of its included
`define.udf` (`\\begin{global_def}` / `\\begin{def}`: classes, nested
structs, `select` fields, `name[]` arrays). `read_udf` joins the two into a
nested dict keyed by the schema's names, so `deck["mesh"]["NPX"]` reads the
value input.cxx gets from `io_parser("mesh.NPX")`. `deck_params` then maps
the keys input.cxx reads for the SPM step (Navier_Stokes or
Shear_Navier_Stokes_Lees_Edwards fluid, spherical particles; input.cxx:742-833,
1235-1470, 1700-1905, 2305-2310) onto `Params` (and the deck's shear), and `deck_particles` / `restore_deck_state` give the initial
particles and, for `resume.Calculation = CONTINUE`, the saved state
(Save_Restart_udf's record, resume.cxx:10-74, 381-506).

The grammar follows the decks and schemas shipped with the reference: `//`
comments, strings in double or single quotes (a backslash before a newline
continues a string), values separated by optional commas, `{...}` for a
struct's fields in schema order, `[...]` for an array. A struct whose data
group stops early leaves its trailing fields absent; `io_parser` then takes
its default (0 for the optional janus members).
"""
import math
import os
import re

import numpy as np

from ._abi import (FP64, JAX_NONE, JAX_X, JAX_Y, JAX_Z, JP_MOTOR, JP_NONE, JP_OBSTACLE, JP_SLIP, CTime,
                   Params, make_species, new_particles)

_TOKEN = re.compile(r"""
    (?P<ws>\s+|//[^\n]*)
  | (?P<str>"(?:[^"\\]|\\.|\\\n)*"|'(?:[^'\\]|\\.|\\\n)*')
  | (?P<cmd>\\[A-Za-z_]+)
  | (?P<unit>\[(?:[A-Za-z]|\d+/)[A-Za-z0-9_/^*.\-() ]*\])
  | (?P<num>[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?)
  | (?P<name>[A-Za-z_][A-Za-z_0-9]*)
  | (?P<punct>\[\]|[{}\[\](),:])
""", re.VERBOSE)

BASIC = {"int", "float", "double", "string", "bool"}


class UdfError(ValueError):
    pass


def _tokens(text):
    pos, out = 0, []
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m:
            raise UdfError(f"unexpected text near {text[pos:pos + 40]!r}")
        pos = m.end()
        kind = m.lastgroup
        if kind in ("ws", "unit"):  # unit annotations of the schema carry no data
            continue
        v = m.group()
        if kind == "str":
            v = v[1:-1].replace("\\\n", "")
        out.append((kind, v))
    return out


class _Stream:
    def __init__(self, toks):
        self.t, self.i = toks, 0

    def peek(self, k=0):
        j = self.i + k
        return self.t[j] if j < len(self.t) else (None, None)

    def next(self):
        tok = self.peek()
        self.i += 1
        return tok

    def expect(self, v):
        tok = self.next()
        if tok[1] != v:
            raise UdfError(f"expected {v!r}, found {tok[1]!r}")

    def skip(self, v):
        if self.peek()[1] == v:
            self.i += 1
            return True
        return False


# ---------------------------------------------------------------- schema
# A type is ("basic", name) | ("select", choices) | ("class", name) |
# ("struct", [(field, dims, type), ...]); dims = number of [] after a name.

def _parse_type(s):
    kind, v = s.peek()
    if v == "{":
        return ("struct", _parse_fields(s))
    if kind == "name" and v == "select":
        s.next()
        s.expect("{")
        choices = []
        while not s.skip("}"):
            k, c = s.next()
            if k == "str":
                choices.append(c)
        return ("select", choices)
    if kind == "name":
        s.next()
        return ("basic", v) if v in BASIC else ("class", v)
    raise UdfError(f"bad type at {v!r}")


def _parse_field(s):
    kind, name = s.next()
    if kind != "name":
        raise UdfError(f"field name expected, found {name!r}")
    dims = 0
    while s.skip("[]"):
        dims += 1
    s.expect(":")
    typ = _parse_type(s)
    if s.peek()[0] == "str":  # the field's comment
        s.next()
    return name, dims, typ


def _parse_fields(s):
    s.expect("{")
    fields = []
    while not s.skip("}"):
        if s.skip(","):
            continue
        fields.append(_parse_field(s))
    return fields


def _parse_def(s, schema, classes, end):
    while True:
        kind, v = s.peek()
        if kind is None:
            raise UdfError(f"missing \\end{{{end}}}")
        if kind == "cmd" and v == "\\end":
            return
        if kind == "name" and v == "class":
            s.next()
            _, cname = s.next()
            s.expect(":")
            classes[cname] = ("struct", _parse_fields(s))
            continue
        name, dims, typ = _parse_field(s)
        schema[name] = (dims, typ)


# ------------------------------------------------------------------ data

def _parse_value(s):
    kind, v = s.peek()
    if v == "[]":
        s.next()
        return ("array", [])
    if v in ("{", "["):
        close = "}" if v == "{" else "]"
        s.next()
        items = []
        while not s.skip(close):
            if s.skip(","):
                continue
            items.append(_parse_value(s))
        return ("group" if v == "{" else "array", items)
    s.next()
    if kind == "num":
        return ("num", v)
    if kind in ("str", "name"):
        return ("str", v)
    raise UdfError(f"bad value {v!r}")


def _bind(value, dims, typ, classes):
    """Positional data -> Python values for a schema type."""
    if dims:
        if value[0] != "array":
            raise UdfError("array expected")
        return [_bind(v, dims - 1, typ, classes) for v in value[1]]
    t, arg = typ
    if t == "class":
        if arg not in classes:
            raise UdfError(f"unknown class {arg}")
        return _bind(value, 0, classes[arg], classes)
    if t == "struct":
        if value[0] != "group":
            raise UdfError("struct value expected")
        out = {}
        for (name, fd, ft), v in zip(arg, value[1]):
            out[name] = _bind(v, fd, ft, classes)
        return out
    if t == "select" or (t == "basic" and arg == "string"):
        return value[1]
    if t == "basic":
        if value[0] != "num":
            raise UdfError(f"number expected, found {value[1]!r}")
        return int(float(value[1])) if arg == "int" else float(value[1])
    raise UdfError(f"unknown type {typ}")


def _sections(text, base_dir, seen):
    """Expand \\include{"file"} relative to the including file; drop unit
    definition blocks and a byte-order mark."""
    text = re.sub(r"\\begin\{unit\}.*?\\end\{unit\}", "", text.lstrip("\ufeff"), flags=re.S)
    out = []
    pos = 0
    for m in re.finditer(r'\\include\{\s*"([^"]+)"\s*\}', text):
        out.append(text[pos:m.start()])
        path = os.path.join(base_dir, m.group(1))
        if path in seen:
            raise UdfError(f"include cycle at {path}")
        with open(path) as f:
            out.append(_sections(f.read(), os.path.dirname(path), seen | {path}))
        pos = m.end()
    out.append(text[pos:])
    return "".join(out)


def read_udf(path):
    """Parse a UDF file (with its includes) -> dict of the data section's
    values keyed by the schema's names. The header block is skipped."""
    return read_udf_schema(path)[0]


def read_udf_schema(path):
    """read_udf, also returning the schema: (data, schema, classes)."""
    data, schema, classes, _ = _read(path)
    return data, schema, classes


def read_udf_records(path):
    """Every `\\begin{data}` block of a UDF file as its own dict, in file
    order: the records an output file holds (Output_udf appends one per
    frame with newRecord, output.cxx:268-375)."""
    return _read(path)[3]


def _read(path):
    with open(path) as f:
        return parse_udf_text(_sections(f.read(), os.path.dirname(os.path.abspath(path)), {os.path.abspath(path)}))


def parse_udf_text(text):
    """UDF text with its includes already expanded -> (data, schema,
    classes, records)."""
    s = _Stream(_tokens(text))
    schema, classes, data, records = {}, {}, {}, []
    depth_header = 0
    while s.peek()[0] is not None:
        kind, v = s.next()
        if kind != "cmd":
            if depth_header:
                continue
            raise UdfError(f"unexpected {v!r} outside a section")
        s.expect("{")
        _, sec = s.next()
        s.expect("}")
        if v == "\\begin" and sec == "header":
            depth_header += 1
        elif v == "\\end" and sec == "header":
            depth_header -= 1
        elif depth_header:
            continue
        elif v == "\\begin" and sec in ("global_def", "def"):
            _parse_def(s, schema, classes, sec)
        elif v == "\\begin" and sec == "data":
            rec = {}
            records.append(rec)
            while not (s.peek()[0] == "cmd" and s.peek()[1] == "\\end"):
                _, name = s.next()
                while s.skip("[]"):
                    pass
                s.expect(":")
                value = _parse_value(s)
                if name not in schema:
                    raise UdfError(f"{name} has no definition")
                dims, typ = schema[name]
                data[name] = rec[name] = _bind(value, dims, typ, classes)
        elif v == "\\end":
            continue
        else:
            raise UdfError(f"unknown section {v}{{{sec}}}")
    return data, schema, classes, records


# ------------------------------------------------------------------ write

def _default(dims, typ, classes):
    if dims:
        return []
    t, arg = typ
    if t == "class":
        return _default(0, classes[arg], classes)
    if t == "struct":
        return {name: _default(fd, ft, classes) for name, fd, ft in arg}
    if t == "select":
        return arg[0] if arg else ""
    if t == "basic":
        return "" if arg == "string" else (0 if arg == "int" else 0.0)
    raise UdfError(f"unknown type {typ}")


def _dump(value, dims, typ, classes, out):
    """Python value -> positional UDF text in schema order (missing members
    take the schema's zero value)."""
    if dims:
        out.append("[")
        for k, v in enumerate(value):
            if k:
                out.append(",")
            _dump(v, dims - 1, typ, classes, out)
        out.append("]")
        return
    t, arg = typ
    if t == "class":
        _dump(value, 0, classes[arg], classes, out)
    elif t == "struct":
        out.append("{")
        for k, (name, fd, ft) in enumerate(arg):
            if k:
                out.append(",")
            v = value.get(name) if isinstance(value, dict) else None
            _dump(_default(fd, ft, classes) if v is None else v, fd, ft, classes, out)
        out.append("}")
    elif t == "select" or (t == "basic" and arg == "string"):
        out.append('"' + str(value) + '"')
    elif t == "basic" and arg == "int":
        out.append(str(int(value)))
    else:
        out.append(repr(float(value)))


def write_udf(path, data, schema, classes, include="define.udf"):
    """Write `data` (read_udf's form) as a deck that includes `include`."""
    out = [f'\\include{{"{include}"}}\n\n\\begin{{data}}\n']
    for name, v in data.items():
        if name not in schema:
            raise UdfError(f"{name} has no definition")
        dims, typ = schema[name]
        out.append(name + ":")
        _dump(v, dims, typ, classes, out)
        out.append("\n")
    out.append("\\end{data}\n")
    with open(path, "w") as f:
        f.write("".join(out))


# --------------------------------------------------------- deck -> run

_JAX = {"X": JAX_X, "Y": JAX_Y, "Z": JAX_Z, "NONE": JAX_NONE}                 # input.cxx:73
_JP = {"TUMBLER": JP_MOTOR, "SQUIRMER": JP_SLIP, "OBSTACLE": JP_OBSTACLE, "OFF": JP_NONE}  # input.cxx:74
_LJP = {"12:6": 0, "24:12": 1, "36:18": 2}                                      # input.cxx:1729-1741
_GDIR = {"-X": 0, "-Y": 1, "-Z": 2}                                             # input.cxx:1713-1722


class Deck:
    """What the SPM step needs from a deck: params (Params), counts,
    snapshots (GTS, Num_snap; MSTEP = GTS*Num_snap, the loop running
    MSTEP + 1 steps, sp_3d_ns.cxx:857), placement and orientation rules."""

    def __init__(self, data, params, counts, gts, num_snap, distribution, orientation, shear=None):
        self.data, self.params, self.counts = data, params, counts
        self.gts, self.num_snap = gts, num_snap
        self.mstep = gts * num_snap
        self.distribution, self.orientation = distribution, orientation
        # (rate, ac, frequency) of a Shear_Navier_Stokes_Lees_Edwards deck
        self.shear = shear

    @property
    def nparticles(self):
        return int(sum(self.counts))


def _vec(v):
    return tuple(float(v[k]) for k in ("x", "y", "z")) if isinstance(v, dict) else (0.0, 0.0, 0.0)


def deck_params(data, precision=FP64, seed=12345):
    """input.cxx's reading of a Navier_Stokes or Shear_Navier_Stokes_Lees_Edwards
    (input.cxx:796-833) / spherical_particle deck."""
    ce = data["constitutive_eq"]
    if ce["type"] not in ("Navier_Stokes", "Shear_Navier_Stokes_Lees_Edwards"):
        raise UdfError(f"constitutive_eq {ce['type']!r}: this step implements Navier_Stokes and "
                       "Shear_Navier_Stokes_Lees_Edwards")
    ns = ce[ce["type"]]
    shear = None
    if ce["type"] == "Shear_Navier_Stokes_Lees_Edwards":
        ef = ns["External_field"]
        if ef["type"] == "DC":
            shear = (float(ef["DC"]["Shear_rate"]), False, 0.0)
        elif ef["type"] == "AC":
            shear = (float(ef["AC"]["Shear_rate"]), True, float(ef["AC"]["Frequency"]))
        else:
            raise UdfError(f"External_field {ef['type']!r}")
    ot = data["object_type"]
    if ot["type"] != "spherical_particle":
        raise UdfError(f"object_type {ot['type']!r}: this step implements spherical_particle")
    species, counts = [], []
    for sp in ot["spherical_particle"]["Particle_spec"]:
        jp = _JP[sp.get("janus_propulsion", "OFF")]
        motor, slip = jp == JP_MOTOR, jp == JP_SLIP
        species.append(make_species(
            mass_ratio=sp["MASS_RATIO"], propulsion=jp, janus_axis=_JAX[sp.get("janus_axis", "NONE")],
            janus_force=_vec(sp.get("janus_force")) if motor else (0, 0, 0),
            janus_torque=_vec(sp.get("janus_torque")) if motor else (0, 0, 0),
            slip_vel=sp.get("janus_slip_vel", 0.0) if slip else 0.0,
            slip_mode=sp.get("janus_slip_mode", 0.0) if slip else 0.0,
            rotlet_c1=sp.get("janus_rotlet_C1", 0.0) if slip else 0.0,
            rotlet_c2=sp.get("janus_rotlet_dipole_C2", 0.0) if slip else 0.0))
        counts.append(int(sp["Particle_number"]))
    mesh = data["mesh"]
    n = (1 << mesh["NPX"], 1 << mesh["NPY"], 1 << mesh["NPZ"])
    dx, rho, eta = ns["DX"], ns["RHO"], ns["ETA"]
    ti = data["time_increment"]
    if ti["type"] == "auto":  # DT = factor / (NU KMAX2), input.cxx:469-476, 518
        kmax2 = sum((2 * math.pi / (nd * dx) * ((nd + 2) // 3)) ** 2 for nd in n)
        dt = ti["auto"]["factor"] / ((eta / rho) * kmax2)
    elif ti["type"] == "manual":
        dt = ti["manual"]["delta_t"]
    else:
        raise UdfError(f"invalid time_increment {ti['type']!r}")
    sw = data["switch"]
    lj_trunc = {"ON": 1, "OFF": 0, "NONE": -1}[sw["LJ_truncate"]]
    fix = sw["FIX_CELL"]
    gr = data["gravity"]
    P = Params(n, dx=dx, rho=rho, eta=eta, kbt=ns["kBT"], alpha_v=ns["alpha_v"], alpha_o=ns["alpha_o"], dt=dt,
               a=data["A"], a_xi=data["A_XI"], g=gr["G"], g_dir=_GDIR[gr["G_direction"]],
               rotation=1 if sw["ROTATION"] == "ON" else 0,
               fix_cell=tuple(1 if fix[k] == "ON" else 0 for k in ("x", "y", "z")),
               lj_truncate=lj_trunc, lj_powers=_LJP[data["LJ_powers"]], epsilon=data["EPSILON"],
               species=species, slip_iter_max=int(sw.get("SLIP_iter", 4)), slip_tol=float(sw.get("SLIP_tol", 0.005)),
               precision=precision, seed=seed)
    out = data["output"]
    return Deck(data, P, counts, int(out["GTS"]), int(out["Num_snap"]), sw["INIT_distribution"],
                sw["INIT_orientation"], shear)


def read_deck(path, precision=FP64, seed=12345):
    return deck_params(read_udf(path), precision=precision, seed=seed)


def deck_particles(deck, seed=10):
    """Initial particles: species in Particle_spec order; positions from
    user_specify or, for uniform_random, random sequential addition with the
    reference's minimum centre distance 1.05 * 2a (init_particle.cxx:65;
    glibc rand() itself is not reproduced); orientation space_align (q = 1),
    random (seeded) or user_specify."""
    from .workloads import place_spheres, random_quaternions
    n = deck.nparticles
    c = deck.params.c
    spec = np.repeat(np.arange(len(deck.counts)), deck.counts)
    dist = deck.distribution
    user = dist.get("user_specify", {}).get("Particles", [])
    if dist["type"] == "user_specify":
        if len(user) < n:
            raise UdfError(f"user_specify holds {len(user)} particles, the deck asks for {n}")
        pos = np.array([_vec(u["R"]) for u in user[:n]])
    elif dist["type"] == "uniform_random":
        box = np.array([c.nx, c.ny, c.nz]) * c.dx
        pos = place_spheres(n, box, c.a * c.dx, gap=0.05 * 2 * c.a * c.dx, seed=seed)
    else:
        raise UdfError(f"INIT_distribution {dist['type']!r} is not supported")
    p = new_particles(n)
    p["spec"] = spec
    p["x"] = pos
    p["x_nopbc"] = pos
    p["x_previous"] = pos
    if dist["type"] == "user_specify":
        p["v"] = np.array([_vec(u["v"]) for u in user[:n]])
        p["omega"] = np.array([_vec(u["omega"]) for u in user[:n]])
    if deck.orientation == "random":
        q = random_quaternions(n, seed)
    elif deck.orientation == "user_specify":
        q = np.array([[u["q"][k] for k in ("q0", "q1", "q2", "q3")] for u in user[:n]])
    else:
        q = np.tile([1.0, 0.0, 0.0, 0.0], (n, 1))
    p["q"] = q
    p["q_old"] = q
    return p


_SAVED = [("R", "x"), ("v", "v"), ("v_old", "v_old"), ("f_hydro", "f_hydro"),
          ("f_hydro_previous", "f_hydro_previous"), ("f_hydro1", "f_hydro1"), ("f_slip", "f_slip"),
          ("f_slip_previous", "f_slip_previous"), ("fr", "fr"), ("fr_previous", "fr_previous"),
          ("omega", "omega"), ("omega_old", "omega_old"), ("torque_hydro", "torque_hydro"),
          ("torque_hydro_previous", "torque_hydro_previous"), ("torque_hydro1", "torque_hydro1"),
          ("torque_slip", "torque_slip"), ("torque_slip_previous", "torque_slip_previous")]


def restore_deck_state(run, deck):
    """resume.Calculation = CONTINUE: load the saved jikan, particles, zeta
    and uk_dc (Save_Restart_udf, resume.cxx:10-74, 381-506) into `run`.
    Returns False for a NEW deck."""
    res = deck.data.get("resume", {})
    if res.get("Calculation", "NEW") != "CONTINUE":
        return False
    sd = res["CONTINUE"]["Saved_Data"]
    if len(sd["Particles"]) != run.np:
        raise UdfError(f"restart record holds {len(sd['Particles'])} particles, the run has {run.np}")
    # the record carries no species: it follows Particle_spec order, as at
    # initialisation
    p = new_particles(run.np)
    p["spec"] = np.repeat(np.arange(len(deck.counts)), deck.counts)[:run.np]
    for k, sp in enumerate(sd["Particles"]):
        for a, b in _SAVED:
            if a in sp:
                p[k][b] = _vec(sp[a])
        if "R_raw" in sp:
            p[k]["x_nopbc"] = _vec(sp["R_raw"])
        for a in ("torque_r", "torque_r_previous"):
            if a in sp:
                p[k][a] = _vec(sp[a])
        for a, b in (("q", "q"), ("q_old", "q_old")):
            if a in sp:
                p[k][b] = [sp[a][s] for s in ("q0", "q1", "q2", "q3")]
    run.set_particles(p)
    z = sd["Zeta"]
    z0 = np.array([[[c["zeta0"] for c in row] for row in plane] for plane in z])
    z1 = np.array([[[c["zeta1"] for c in row] for row in plane] for plane in z])
    run.set_fluid(z0, z1, np.array(_vec(sd["uk_dc"])))
    t = run.get_time()
    run.set_time(CTime(int(sd["jikan"]["ts"]), float(sd["jikan"]["time"]), t.dt_fluid, t.hdt_fluid, t.dt_md,
                       t.hdt_md))
    if deck.shear is not None:
        # the strain is restored (resume.cxx:562-567); Shear_rate_eff is not
        # saved, so the resumed process starts it from its initial value 0
        # until the first step sets it
        run.set_shear(*deck.shear)
        run.set_shear_state(float(sd.get("oblique", {}).get("degree_oblique", 0.0)), 0.0)
    return True


def restart_record(run):
    """The resume entry Save_Restart_udf writes (resume.cxx:10-74, particle
    members 381-506) for the state of `run`, in read_udf's form."""
    z0, z1, uk = run.get_fluid()
    t = run.get_time()
    p = run.get_particles()

    def v3(a):
        return {"x": float(a[0]), "y": float(a[1]), "z": float(a[2])}

    def q4(a):
        return {"q0": float(a[0]), "q1": float(a[1]), "q2": float(a[2]), "q3": float(a[3])}

    parts = []
    for k in range(len(p)):
        rec = {a: v3(p[k][b]) for a, b in _SAVED}
        rec["R_raw"] = v3(p[k]["x_nopbc"])
        rec["torque_r"] = v3(p[k]["torque_r"])
        rec["torque_r_previous"] = v3(p[k]["torque_r_previous"])
        rec["q"], rec["q_old"] = q4(p[k]["q"]), q4(p[k]["q_old"])
        parts.append(rec)
    zeta = [[[{"zeta0": float(a), "zeta1": float(b)} for a, b in zip(r0, r1)] for r0, r1 in zip(p0, p1)]
            for p0, p1 in zip(z0.tolist(), z1.tolist())]
    return {"Calculation": "CONTINUE",
            "CONTINUE": {"Saved_Data": {"jikan": {"ts": int(t.ts), "time": float(t.time)}, "Particles": parts,
                                        "Zeta": zeta, "uk_dc": v3(uk),
                                        "oblique": {"degree_oblique": float(run.shear_state()[0])}}}}


def save_restart_udf(run, deck_path, out_path, include=None):
    """Write deck_path's data with resume = CONTINUE and the run's state,
    the restart deck the reference writes (resume.cxx:10-74); `include`
    names the schema file the new deck includes (default: the one
    deck_path includes, resolved relative to out_path's folder by the
    reader)."""
    data, schema, classes = read_udf_schema(deck_path)
    if "resume" not in schema:
        raise UdfError(f"{deck_path}: its schema has no resume entry")
    data["resume"] = restart_record(run)
    if include is None:
        with open(deck_path) as f:
            m = re.search(r'\\include\{\s*"([^"]+)"\s*\}', f.read())
        include = m.group(1) if m else "define.udf"
    write_udf(out_path, data, schema, classes, include=include)
