"""Reader for the subset of Smoldyn's configuration language used by the
reference's particle-path regression models (examples/S95_regression).

Statement semantics follow simreadstring (source/Smoldyn/smolsim.cpp:626-2263)
for the statements listed in HANDLERS; display-only statements are ignored and
anything else raises Unsupported.  Statements are applied in file order, which
matters because molecule placement draws from the random stream seeded by
random_seed.
"""
import re

IGNORED = {"graphics", "graphic_iter", "graphic_delay", "color", "display_size", "text_display",
           "output_files", "cmd", "frame_thickness", "frame_color", "grid_thickness", "background_color",
           "tiff_name", "tiff_iter", "light", "quit_at_end", "end_file"}
# display-only statements inside a surface block (surfreadstring, smolsurface.c)
SURF_IGNORED = {"color", "thickness", "polygon", "shininess", "stipple", "max_panels", "end_surface"}
# panel parameter counts by shape and dimension (surfreadstring "panel")
PANEL_NPARAM = {("rect", 1): 1, ("rect", 2): 3, ("rect", 3): 5, ("tri", 1): 1, ("tri", 2): 4, ("tri", 3): 9,
                ("sph", 1): 2, ("sph", 2): 4, ("sph", 3): 6, ("cyl", 3): 9, ("hemi", 3): 9, ("disk", 2): 5,
                ("disk", 3): 8}
PANEL_SHAPE = {"rect": "rect", "r": "rect", "tri": "tri", "t": "tri", "sph": "sph", "sphere": "sph", "s": "sph",
               "disk": "disk", "d": "disk", "cyl": "cyl", "hemi": "hemi"}
STATES = {"soln": 0, "fsoln": 0, "front": 1, "back": 2, "up": 3, "down": 4, "bsoln": 5, "all": 6}


class Unsupported(Exception):
    pass


_FLOAT = re.compile(r"[-+]?(\d+\.?\d*|\.\d+)([eE][-+]?\d+)?")


def _range(tok):
    """Parse "lo-hi" the way sscanf("%lg-%lg") does in the mol statement."""
    m1 = _FLOAT.match(tok)
    if not m1 or m1.end() >= len(tok) or tok[m1.end()] != "-":
        return None
    m2 = _FLOAT.match(tok, m1.end() + 1)
    if not m2:
        return None
    return float(m1.group(0)), float(m2.group(0))


def _num(tok, defines):
    tok = defines.get(tok, tok)
    return float(tok)


def read_config(text, sim_factory, defines=None):
    """Apply a config to a simulation created by sim_factory(dim, low, high).

    Walls are only known after the boundaries statements, so statements are
    buffered until the first statement that needs the simulation object."""
    defines = dict(defines or {})
    lines = []
    skip = [False]
    for raw in text.splitlines():
        line = raw.split("#", 1)[0].strip()
        if not line:
            continue
        w = line.split()
        if w[0] == "define":
            defines[w[1]] = " ".join(w[2:])
            continue
        if w[0] == "ifdefine":
            skip.append(skip[-1] or w[1] not in defines)
            continue
        if w[0] == "ifundefine":
            skip.append(skip[-1] or w[1] in defines)
            continue
        if w[0] == "else":
            top = skip.pop()
            skip.append((not top) if not skip[-1] else True)
            continue
        if w[0] == "endif":
            skip.pop()
            continue
        if skip[-1]:
            continue
        words = [defines.get(t, t) for t in w]
        lines.append(words)
        if w[0] == "end_file":
            break

    st = dict(dim=None, bounds={}, walls={}, seed=None)
    pending = []  # statements applied once the simulation exists
    for words in lines:
        k = words[0]
        if k == "dim":
            st["dim"] = int(words[1])
        elif k == "boundaries":
            d = {"x": 0, "y": 1, "z": 2}.get(words[1], None)
            d = int(words[1]) if d is None else d
            st["bounds"][d] = (float(words[2]), float(words[3]))
            t = words[4] if len(words) > 4 else "t"
            st["walls"][(d, 0)] = t
            st["walls"][(d, 1)] = t
        elif k in ("low_wall", "high_wall"):
            d = {"x": 0, "y": 1, "z": 2}.get(words[1], None)
            d = int(words[1]) if d is None else d
            lo, hi = st["bounds"].get(d, (None, None))
            if k == "low_wall":
                lo = float(words[2])
                st["walls"][(d, 0)] = words[3]
            else:
                hi = float(words[2])
                st["walls"][(d, 1)] = words[3]
            st["bounds"][d] = (lo, hi)
        elif k in IGNORED:
            continue
        else:
            pending.append(words)

    dim = st["dim"]
    low = [st["bounds"][d][0] for d in range(dim)]
    high = [st["bounds"][d][1] for d in range(dim)]
    sim = sim_factory(dim, low, high)
    for (d, side), t in st["walls"].items():
        sim.wall(d, side, t)
    times = [0.0, None, None]
    rxns = {}
    surfaces = {}
    panels = {}  # (surface, panel name) -> panel id
    cur = [None, {}]  # current surface index, shape counters
    for words in pending:
        k = words[0]
        if cur[0] is not None and k != "end_surface" and k not in ("start_surface",):
            _surface_statement(sim, dim, cur, words, panels)
            continue
        if k == "start_surface":
            name = words[1]
            if name not in surfaces:
                surfaces[name] = sim.add_surface(name)
            cur[0] = surfaces[name]
            cur[1] = cur[1] if cur[0] in panels.get("__counters__", {}) else {}
            continue
        if k == "end_surface":
            cur[0] = None
            continue
        if k == "random_seed":
            sim.seed(int(words[1]))
        elif k == "time_start":
            times[0] = float(words[1])
        elif k == "time_stop":
            times[1] = float(words[1])
        elif k == "time_step":
            times[2] = float(words[1])
            sim.times(times[0], times[1], times[2])
        elif k in ("species", "names", "name"):
            for nm in words[1:]:
                sim.species(nm)
        elif k == "difc":
            nm, val = words[1], float(words[2])
            if nm == "all":
                for i in range(1, 64):
                    if sim.index_exists(i):
                        sim.difc(i, val, 0)
            else:
                sim.difc(sim.index(nm), val, 0)
        elif k == "molecule_lists":
            for nm in words[1:]:
                sim.add_list(nm)
        elif k == "mol_list":
            sim.set_list(sim.index(words[1]), sim.list_index(words[2]), 6)
        elif k == "molperbox":
            sim.partitions("molperbox", float(words[1]))
        elif k == "boxsize":
            sim.partitions("boxsize", float(words[1]))
        elif k == "reaction":
            name = words[1]
            rest = words[2:]
            rate = float(rest[-1])
            arrow = rest.index("->")
            rct = [t for t in rest[:arrow] if t not in ("+", "0")]
            prd = [t for t in rest[arrow + 1:-1] if t not in ("+", "0")]
            rxns[name] = sim.reaction(name, [sim.index(t) for t in rct], [sim.index(t) for t in prd], rate)
        elif k == "confspread_radius":
            sim.confspread(rxns[words[1]], float(words[2]))
        elif k == "product_placement":
            sim.revparam(rxns[words[1]], words[2], float(words[3]) if len(words) > 3 else 0.0)
        elif k == "mol":
            n = int(words[1])
            i = sim.index(words[2])
            lo, hi = list(low), list(high)
            for d in range(dim):
                tok = words[3 + d]
                if tok == "u":
                    continue
                rng = _range(tok)
                if rng:
                    lo[d], hi[d] = rng
                else:
                    lo[d] = hi[d] = float(tok)
            sim.add_mols(n, i, lo, hi)
        elif k == "accuracy":
            sim.accuracy(float(words[1]))
        else:
            raise Unsupported(k)
    return sim


def _species_state(sim, tok):
    """species[(state)] -> (ident, state); 'all' is ident -5 (molstring2index1)."""
    ms = 0
    if "(" in tok:
        tok, st = tok[:-1].split("(")
        ms = STATES[st]
    return (-5 if tok == "all" else sim.index(tok)), ms


def _surface_statement(sim, dim, cur, words, panels):
    from oracle_bind import FACE, PS, SA
    s = cur[0]
    k = words[0]
    if k in SURF_IGNORED:
        return
    if k == "action":
        # new format: species face action; old format: face species action
        a, b = words[1], words[2]
        if b in FACE and b not in ("none",):
            sp, face = a, b
        else:
            sp, face = b, a
        ident, ms = _species_state(sim, sp)
        newident = sim.index(words[4]) if len(words) > 4 else ident
        sim.surface_action(s, ident, ms, FACE[face], SA[words[3]], newident)
    elif k in ("rate", "rate_internal"):
        ident, ms = _species_state(sim, words[1])
        ms1, ms2 = STATES[words[2]], STATES[words[3]]
        newident = sim.index(words[5]) if len(words) > 5 else -5
        sim.surface_rate(s, ident, ms, ms1, ms2, float(words[4]), 1 if k == "rate" else 2, newident)
    elif k == "panel":
        shape = PANEL_SHAPE[words[1]]
        toks = words[2:]
        axis = None
        if shape == "rect":
            axis, toks = toks[0], toks[1:]
        npar = PANEL_NPARAM.get((shape, dim))
        if npar is None:
            raise Unsupported("panel %s in %dD" % (shape, dim))
        nums, name = [], None
        for t in toks:
            try:
                nums.append(float(t))
            except ValueError:
                name = t
        if shape in ("sph", "disk") and len(nums) < npar:  # drawing slices/stacks default
            nums += [20.0] * (npar - len(nums))
        cnt = cur[1].get(shape, 0)
        cur[1][shape] = cnt + 1
        pid = sim.add_panel(s, PS[shape], nums[:npar], axis)
        panels[(s, name or "%s%d" % (shape, cnt))] = pid
    elif k == "jump":
        p1, f1, arrow, p2, f2 = words[1:6]
        sim.panel_jump(panels[(s, p1)], FACE[f1], panels[(s, p2)], FACE[f2], arrow == "<->")
    else:
        raise Unsupported("surface statement " + k)
