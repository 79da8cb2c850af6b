"""Read a reference input deck (msg123.main and the files it names) into a
Problem, for the decks this solver covers: a transient run (sim_type 1) on a
type-7 uniform grid with 3D float32 soil files, a water-table initial state
and a constant, uniform 2D recharge.  Anything else raises ValueError naming
the setting, instead of being approximated.

Formats (the reference's readers):
  - namelists in the order of read_main_file / read_inp_set / read_out_set
    (Read_input.f90:23-60), group syntax of Fortran namelist input;
  - grid type 7: a header line, then "corner width" for x, y and z
    (Read_input.f90:548-566), node k at corner + width*(k-1);
  - retention and parameter indirection files (&inretn_type/&inretn_path,
    &inparm_type/&inparm_path, Read_input.f90:632-680) with type 6 = a 3D
    float32 stream file, x fastest (read_3d_r4bin, Read_module.f90:446-479,
    opened access='stream', Utility_module.f90:111-112);
  - the time-input file (&tinp_type/&tinp_path/&tinp_unit,
    Read_input.f90:732-736), rech_type 4 = a 2D float32 stream file in mm/h;
    with rech_unit set (SEC, MIN, HOU or DAY; conv_unit,
    Utility_module.f90:322-372) the file holds records, each its end time in
    that unit and then the 2D values (skip_file / read_next,
    Read_module.f90:516-583,984-1065), read into rech_etime / rech_mmh;
  - the calculation time from sdate/edate (calc_etime, Read_input.f90:1158-1218)
    in float32 seconds.

    from msg123_amd import deck, gpu
    p, opts = deck.read_deck("path/to/deck")
    ctx = gpu.Context(gpu.BoxMesh(p), precon=opts["precon_type"])
"""
from __future__ import annotations

import datetime
import os
import re

import numpy as np

from msg123_amd.configs import Problem

_TOKEN = re.compile(r"""'(?:[^']|'')*'|"(?:[^"]|"")*"|[^\s,=/&!'"]+|[=,/&!]""")


def _value(tok: str):
    if tok[0] in "'\"":
        return tok[1:-1].replace(tok[0] * 2, tok[0])
    t = tok.lower()
    if t in (".true.", "t", ".t."):
        return True
    if t in (".false.", "f", ".f."):
        return False
    try:
        return int(t)
    except ValueError:
        pass
    try:
        return float(t.replace("d", "e"))
    except ValueError:
        raise ValueError(f"unreadable namelist value {tok!r}") from None


def parse_namelists(text: str) -> dict:
    """{group: {name: value or [values]}} of every &group ... / in `text`
    (names lower-cased; `!` starts a comment outside strings)."""
    lines = []
    for line in text.splitlines():
        out, q = [], None
        for ch in line:
            if q:
                q = None if ch == q else q
            elif ch in "'\"":
                q = ch
            elif ch == "!":
                break
            out.append(ch)
        lines.append("".join(out))
    toks = _TOKEN.findall("\n".join(lines))
    groups, i = {}, 0
    while i < len(toks):
        if toks[i] != "&":
            i += 1
            continue
        name = toks[i + 1].lower()
        i += 2
        grp, key, vals = {}, None, []
        while i < len(toks) and toks[i] not in ("/", "&"):
            if i + 1 < len(toks) and toks[i + 1] == "=":
                if key is not None:
                    grp[key] = vals[0] if len(vals) == 1 else vals
                key, vals = toks[i].lower(), []
                i += 2
                continue
            if toks[i] != ",":
                vals.append(_value(toks[i]))
            i += 1
        if key is not None:
            grp[key] = vals[0] if len(vals) == 1 else vals
        groups[name] = grp
        i += 1
    return groups


def _need(cond, what):
    if not cond:
        raise ValueError(f"deck not supported by this solver: {what}")


def _etime(sdate, edate) -> float:
    s = datetime.datetime(*[int(v) for v in sdate])
    e = datetime.datetime(*[int(v) for v in edate])
    return float(np.float32((e - s).total_seconds()))


def read_deck(d: str, name: str = "deck"):
    """(Problem, options) of the deck in directory `d`; options holds
    precon_type, the set_amg values and sim_name."""
    def path(f):
        return os.path.join(d, f)

    def nml(f):
        with open(path(f)) as fh:
            return parse_namelists(fh.read())

    main = nml("msg123.main")
    sim = main.get("set_simulation", {})
    _need(sim.get("sim_type") == 1, f"sim_type {sim.get('sim_type')} (only 1, transient)")
    ct = main["set_calc_time"]
    _need(str(ct.get("calc_unit", "")).strip().upper() == "SEC",
          f"calc_unit {ct.get('calc_unit')!r} (only 'SEC')")
    end_time = _etime(ct["sdate"], ct["edate"])
    sol = main["set_solution"]
    _need(float(sol.get("incr_multi", 1.2)) == 1.2, "incr_multi other than 1.2")
    g = main["set_grid"]
    nx, ny, nz = int(g["gridx"]), int(g["gridy"]), int(g["gridz"])
    _need(g.get("grid_type") == 7, f"grid_type {g.get('grid_type')} (only 7)")
    with open(path(g["grid_file"])) as fh:
        rows = [ln.split() for ln in fh.read().splitlines()[1:4]]
    (_, dx), (_, dy), (z0, dzw) = [(float(a), float(b)) for a, b in rows]
    _need(z0 == 0.0 and dzw < 0.0 and dx > 0.0 and dy > 0.0,
          "grid corner z other than 0 or a z width that is not negative")
    n = nx * ny * nz

    def f32(f, count):
        a = np.fromfile(path(f), dtype="<f4")
        _need(a.size >= count, f"{f} holds {a.size} values, {count} needed")
        return a[:count].astype(np.float32)

    rp = main["set_retn_parm"]
    _need(rp.get("retn_type") == 0 and rp.get("parm_type") == 0,
          "retn_type/parm_type other than 0 (per-cell files)")
    rt, pt = nml(rp["retn_file"]), nml(rp["parm_file"])
    soil = {}
    for grp, path_grp, items in (
            (rt["inretn_type"], rt["inretn_path"],
             (("vana", "in_vapath"), ("vann", "in_vnpath"), ("resi", "in_repath"))),
            (pt["inparm_type"], pt["inparm_path"],
             (("pakx", "in_kxpath"), ("paky", "in_kypath"), ("pakz", "in_kzpath"),
              ("pass", "in_sspath"), ("pats", "in_tspath")))):
        for key, pkey in items:
            _need(grp.get(key + "_type") == 6, f"{key}_type {grp.get(key + '_type')} (only 6)")
            soil[key] = f32(path_grp[pkey], n)
    _need(np.array_equal(soil["pakx"], soil["paky"]) and np.array_equal(soil["pakx"], soil["pakz"]),
          "anisotropic conductivity (pakx, paky, pakz differ)")
    ini = main["set_init"]
    _need(ini.get("init_type") == 7, f"init_type {ini.get('init_type')} (only 7, a water table)")
    ti = nml(main["set_time_input"]["time_input_file"])
    _need(ti["tinp_type"].get("rech_type") == 4, "rech_type other than 4 (a 2D binary file)")
    units = {k: str(v).strip().upper() for k, v in ti.get("tinp_unit", {}).items()}
    _need(not any(v for k, v in units.items() if k != "rech_unit"),
          "a time-varying input other than recharge (tinp_unit)")
    rech_etime = rech_mmh = None
    runit = units.get("rech_unit", "")
    if runit:
        multi = {"SEC": np.float32(1.0), "MIN": np.float32(60.0), "HOU": np.float32(3600.0),
                 "DAY": np.float32(86400.0)}.get(runit)
        _need(multi is not None, f"rech_unit {runit} (only SEC, MIN, HOU, DAY)")
        raw = np.fromfile(path(ti["tinp_path"]["rech_path"]), dtype="<f4")
        rec = 1 + nx * ny
        _need(raw.size >= rec and raw.size % rec == 0,
              f"a recharge file of {raw.size} values (records of 1 + {nx * ny})")
        recs = raw.reshape(-1, rec)
        _need(np.all(recs[:, 1:] == recs[:, 1:2]), "a recharge that varies over the top surface")
        rech_etime = (recs[:, 0] * multi).astype(np.float32)
        rech_mmh = recs[:, 1].astype(np.float32)
        _need(np.all(np.diff(rech_etime) > 0), "recharge record end times that do not increase")
        rech = rech_mmh[:1]
    else:
        rech = f32(ti["tinp_path"]["rech_path"], nx * ny)
        _need(np.all(rech == rech[0]), "a recharge that varies over the top surface")
    p = Problem(
        name=name, nx=nx, ny=ny, nz=nz, dx=dx, dy=dy, dz=-dzw,
        alpha=soil["vana"], vg_n=soil["vann"], theta_r=soil["resi"], theta_s=soil["pats"],
        ks=soil["pakx"], ss=soil["pass"], wt_depth=float(ini["init_dept"]),
        recharge_mmh=float(rech[0]), init_step=float(sol["init_step"]),
        max_step=float(sol["max_step"]), end_time=end_time,
        maxout_iter=int(sol["maxout_iter"]), maxinn_iter=int(sol["maxinn_iter"]),
        criteria=float(sol["criteria"]), rech_etime=rech_etime, rech_mmh=rech_mmh)
    opts = dict(precon_type=int(sol.get("precon_type", 0)), amg=main.get("set_amg", {}),
                sim_name=sim.get("sim_name", ""))
    return p, opts
